"""The drop-in boundary on the CPU: the Julia shim (dielectricbie.jl_amd/julia/DielectricBIEHip.jl), the C
header (include/dbie.h) and the ctypes binding (dbie/_lib.py) agree on every struct layout and on every
ccall signature the shim makes; geometry built from the caller's own panel quadrature keeps it bit for
bit (FlatPanel.points/.weights/.gl_xs/.gl_ws, src/core/panels.jl:5-22).  When `julia` is installed the
shim's own test script runs; otherwise that test is skipped with the reason."""
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "dbie.h")
SHIM = os.path.join(ROOT, "dielectricbie.jl_amd", "julia", "DielectricBIEHip.jl")
SHIM_TEST = os.path.join(ROOT, "dielectricbie.jl_amd", "julia", "test_shim.jl")

# C type -> canonical, Julia type -> canonical
C_CANON = {"double": "f64", "int": "i32", "dbie_precision": "i32", "dbie_kernel": "i32", "dbie_status": "i32",
           "int64_t": "i64", "size_t": "usize", "void": "void"}
JL_CANON = {"Float64": "f64", "Cint": "i32", "Int64": "i64", "Csize_t": "usize", "Cvoid": "void", "UInt8": "u8"}


def strip_comments(text):
    return re.sub(r"/\*.*?\*/", " ", text, flags=re.S)


def header_structs():
    """{struct name: [(field, canonical type), ...]} of the header's typedef structs."""
    text = strip_comments(open(HEADER).read())
    out = {}
    for body, name in re.findall(r"typedef struct\s*\{(.*?)\}\s*(\w+)\s*;", text, flags=re.S):
        fields = []
        for decl in body.split(";"):
            decl = decl.strip()
            if not decl:
                continue
            m = re.match(r"([\w\s]+?)\s+([\w\s,\[\]0-9]+)$", decl)
            assert m, decl
            ctype = m.group(1).strip()
            for f in m.group(2).split(","):
                f = f.strip()
                fields.append((re.sub(r"\[.*\]", "", f), C_CANON.get(ctype, ctype)))
        out[name] = fields
    return out


def header_functions():
    """{name: (return canonical, [param canonical types])}."""
    text = strip_comments(open(HEADER).read())
    out = {}
    for ret, name, params in re.findall(r"([\w\s\*]+?)\s*\b(dbie_\w+)\s*\(([^;{]*?)\)\s*;", text, flags=re.S):
        ret = ret.split()[-1] if ret.strip() else ret
        plist = []
        params = " ".join(params.split())
        if params and params != "void":
            for p in params.split(","):
                p = p.strip().replace("const ", "")
                m = re.match(r"([\w\s]+?)\s*(\**)\s*(\w+)$", p)
                assert m, (name, p)
                base, stars = m.group(1).strip(), m.group(2)
                plist.append(canon_c(base, stars))
        out[name] = (canon_c(ret.replace("*", "").strip(), "*" * ret.count("*")), plist)
    return out


def canon_c(base, stars):
    if stars:
        inner = C_CANON.get(base, "handle" if base.startswith("dbie_") else base)
        if base == "char":
            inner = "u8"
        if base in ("dbie_near_opts", "dbie_solve_opts", "dbie_solve_stats", "dbie_server_opts",
                    "dbie_request_info", "dbie_server_summary"):
            return f"ref<{base}>"
        if base == "void" or inner == "handle":
            inner = "void"
        return "ptr<" * len(stars) + inner + ">" * len(stars)
    return C_CANON.get(base, base)


def jl_structs():
    text = open(SHIM).read()
    out = {}
    for kind, name, body in re.findall(r"(mutable struct|struct)\s+(\w+)\s*\n(.*?)\nend", text, flags=re.S):
        fields = []
        for line in body.splitlines():
            m = re.match(r"\s*(\w+)::(\w+)\s*$", line)
            if m:
                fields.append((m.group(1), JL_CANON.get(m.group(2), m.group(2))))
        out[name] = fields
    return out


def canon_jl(t):
    t = t.strip()
    m = re.match(r"(Ptr|Ref)\{(.*)\}$", t)
    if m:
        inner = m.group(2)
        if m.group(1) == "Ref" and inner in JL_STRUCT_TO_C:
            return f"ref<{JL_STRUCT_TO_C[inner]}>"
        return f"ptr<{canon_jl(inner)}>"
    return JL_CANON.get(t, t)


JL_STRUCT_TO_C = {"NearOpts": "dbie_near_opts", "SolveOpts": "dbie_solve_opts", "SolveStats": "dbie_solve_stats"}


def split_top(s):
    """Split at commas not nested in braces or parentheses."""
    out, depth, cur = [], 0, ""
    for ch in s:
        if ch in "({":
            depth += 1
        elif ch in ")}":
            depth -= 1
        if ch == "," and depth == 0:
            out.append(cur)
            cur = ""
        else:
            cur += ch
    if cur.strip():
        out.append(cur)
    return [x.strip() for x in out]


def jl_ccalls():
    """[(symbol, return type, [arg types])] of every ccall in the shim."""
    text = open(SHIM).read()
    calls = []
    for m in re.finditer(r"ccall\(\(:(\w+),\s*lib\),\s*(\w+),\s*\(", text):
        i = m.end()
        depth, j = 1, i
        while depth:
            depth += {"(": 1, ")": -1}.get(text[j], 0)
            j += 1
        args = text[i:j - 1]
        types = [canon_jl(t) for t in split_top(args) if t]
        calls.append((m.group(1), JL_CANON.get(m.group(2), m.group(2)), types))
    return calls


def test_struct_layouts_match_header_and_ctypes(dbie):
    from dbie import _lib as L
    import ctypes as C
    hs = header_structs()
    js = jl_structs()
    ct = {"dbie_near_opts": L.NearOpts, "dbie_solve_opts": L.SolveOpts, "dbie_solve_stats": L.SolveStats}
    ccanon = {C.c_double: "f64", C.c_int: "i32", C.c_int64: "i64"}
    for jname, cname in JL_STRUCT_TO_C.items():
        assert js[jname] == hs[cname], (jname, js[jname], hs[cname])
        assert [(f, ccanon[t]) for f, t in ct[cname]._fields_] == hs[cname], cname


def test_shim_ccalls_match_header():
    funcs = header_functions()
    calls = jl_ccalls()
    assert len(calls) >= 10
    for sym, ret, args in calls:
        assert sym in funcs, f"{sym} is not declared in include/dbie.h"
        cret, cargs = funcs[sym]
        assert ret == cret, (sym, ret, cret)
        assert len(args) == len(cargs), (sym, args, cargs)
        for a, c in zip(args, cargs):
            # Julia passes Vector{Cint}/Vector{Float64} etc. as Ptr{T}; handles as Ptr{Cvoid}
            assert a == c, (sym, args, cargs)


def test_julia_shim_runs_where_julia_exists():
    julia = shutil.which("julia")
    if julia is None:
        pytest.skip("julia is not installed in this image: the shim's layouts and ccall signatures are "
                    "checked against include/dbie.h by the other tests of this file instead")
    r = subprocess.run([julia, SHIM_TEST], capture_output=True, text=True, timeout=1200,
                       env=dict(os.environ, DBIE_LIB=os.path.join(ROOT, "dielectricbie.jl_amd", "dbie", "libdbie.so")))
    assert r.returncode == 0, r.stdout + r.stderr


def test_caller_nodes_kept_bit_for_bit(dbie):
    """Nodes, weights and rule perturbed by one ulp come back unchanged from dbie_geom_nodes."""
    g = dbie.single_dielectric_box3d(1.2, 0.8, 0.6, 4, 0.2, 4.0, 1.0)
    corners, normals, is_edge = g.panels()
    nd = g.nodes
    x, w = dbie.gausslegendre(4)
    pts = np.nextafter(nd[0:3].T, np.inf)
    wts = np.nextafter(nd[6], 0.0)
    g2 = dbie.interface_from_panels(corners, normals, 4, 4.0, 1.0, is_edge, points=pts, weights=wts,
                                    gl_xs=np.nextafter(x, np.inf), gl_ws=w)
    nd2 = g2.nodes
    assert np.array_equal(nd2[0:3].T, pts)
    assert np.array_equal(nd2[6], wts)
    assert np.array_equal(nd2[3:6], nd[3:6])
    # the unperturbed nodes reproduce the mesher's geometry exactly
    g3 = dbie.interface_from_panels(corners, normals, 4, 4.0, 1.0, is_edge, points=nd[0:3].T, weights=nd[6],
                                    gl_xs=x, gl_ws=w)
    assert np.array_equal(g3.nodes, nd)
    with pytest.raises(ValueError):
        dbie.interface_from_panels(corners, normals, 4, 4.0, 1.0, is_edge, points=pts[:10], weights=wts,
                                   gl_xs=x, gl_ws=w)
