"""CPU FP64 oracle for the boundary-operator path (test infrastructure only).

This module is the checker.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` import it; the product package never does.

It wraps ``oracle/liboracle.so`` (a C restatement of the reference, see
``dbie_oracle.c`` for the per-function ``file:line`` citations) with numpy, and
adds the pieces that are plain linear algebra:

* ``solve_lu``  -- restates ``src/utils/linear_algebra.jl:1`` with scipy's LAPACK LU;
* ``gmres``     -- restates ``Krylov.gmres`` as used at ``src/utils/linear_algebra.jl:3-6``
  (full, unrestarted GMRES, stop when ``||r|| <= atol + rtol*||b||``).  Krylov.jl's
  iterate trajectory is not pinned by any reference test; only the converged
  solution is compared.
"""
from __future__ import annotations

import ctypes as C
import math
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


def build(force: bool = False) -> str:
    """Compile liboracle.so with the committed Makefile."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-B" if force else "-s", "-C", _HERE, "liboracle.so"], check=True,
                       stdout=subprocess.DEVNULL)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        vp = C.c_void_p
        L.orc_gauss_legendre.argtypes = [C.c_int, _dp, _dp]
        L.orc_best_grid.argtypes = [C.c_double, C.c_double, C.c_double, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.orc_iface_box.restype = vp
        L.orc_iface_box.argtypes = [C.c_double] * 3 + [C.c_int] + [C.c_double] * 4
        L.orc_iface_from_panels.restype = vp
        L.orc_iface_from_panels.argtypes = [C.c_int, C.c_int, _dp, _dp, _ip, _dp, _dp]
        L.orc_iface_free.argtypes = [vp]
        L.orc_rect_adaptive_panels.restype = C.c_int
        L.orc_rect_adaptive_panels.argtypes = [_dp, _dp, _ip, _ip, C.c_double, C.c_double, C.c_int, _dp, _ip]
        L.orc_iface_sizes.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.orc_iface_nodes.argtypes = [vp, _dp]
        L.orc_iface_panels.argtypes = [vp, _dp, _dp, _ip]
        L.orc_iface_set_eps.argtypes = [vp, _dp, _dp]
        L.orc_diag.argtypes = [vp, _dp]
        for n in ("orc_dense_DT", "orc_dense_S", "orc_dense_D", "orc_lhs_dense"):
            getattr(L, n).argtypes = [vp, _dp]
        L.orc_apply_direct.argtypes = [vp, C.c_int, C.c_int, _dp, _dp]
        L.orc_apply_DT_rows.argtypes = [vp, C.c_int, _ip, _dp, _dp]
        L.orc_pottrg.argtypes = [vp, C.c_int, _dp, C.c_int, _dp, _dp]
        L.orc_rhs_points.argtypes = [vp, C.c_int, _dp, _dp, C.c_double, _dp]
        L.orc_rhs_volume.argtypes = [vp, C.c_int, _dp, _dp, _dp, C.c_double, _dp]
        L.orc_screen_sharp.argtypes = [vp, C.c_int, _dp, _dp, _dp]
        L.orc_bernstein_rho_panel.restype = C.c_double
        L.orc_bernstein_rho_panel.argtypes = [vp, C.c_int, _dp]
        L.orc_neighbor_list.restype = vp
        L.orc_neighbor_list.argtypes = [vp, C.c_int, C.c_double, C.c_int]
        L.orc_nlist_sizes.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.orc_nlist_get.argtypes = [vp, _ip, _ip, _ip, _ip, _ip]
        L.orc_nlist_free.argtypes = [vp]
        L.orc_corrections.restype = C.c_int
        L.orc_corrections.argtypes = [vp, vp, C.c_double, C.c_double, C.c_int, C.c_int, _dp]
        L.orc_corrections_mode.restype = C.c_int
        L.orc_corrections_mode.argtypes = [vp, vp, C.c_int, C.c_double, C.c_double, C.c_int, C.c_int, _dp]
        L.orc_corrections_apply.argtypes = [vp, vp, _dp, C.c_int, _dp, _dp]
        L.orc_laplace3d_grad.restype = C.c_double
        L.orc_laplace3d_grad.argtypes = [_dp, _dp, _dp]
        L.orc_laplace3d_pot.restype = C.c_double
        L.orc_laplace3d_pot.argtypes = [_dp, _dp]
        L.orc_bary_row.argtypes = [C.c_int, _dp, _dp, C.c_double, _dp]
        L.orc_rho_from_pole.restype = C.c_double
        L.orc_rho_from_pole.argtypes = [C.c_double, C.c_double]
        L.orc_num_threads.restype = C.c_int
        _lp = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
        L.orc_tmap_build.restype = vp
        L.orc_tmap_build.argtypes = [vp, C.c_int, _dp, C.c_double, C.c_double]
        L.orc_tmap_sizes.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_long),
                                     C.POINTER(C.c_long), C.POINTER(C.c_long)]
        L.orc_tmap_get.argtypes = [vp, _dp, _ip, _ip, _ip, _ip, _dp, _lp, _ip, _dp]
        L.orc_tmap_apply.argtypes = [vp, C.c_int, _dp, _dp]
        L.orc_reaction.argtypes = [vp, C.c_int, _lp, _dp, C.c_int, _dp, _dp]
        L.orc_tmap_free.argtypes = [vp]
        L.orc_pot_panel_hcub.restype = C.c_int
        L.orc_pot_panel_hcub.argtypes = [vp, C.c_int, _dp, C.c_double, _dp]
        _lib = L
    return _lib


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def gauss_legendre(n: int):
    x = np.zeros(n); w = np.zeros(n)
    if lib().orc_gauss_legendre(n, x, w) != 0:
        raise ValueError("n must be >= 1")
    return x, w


def best_grid(W: float, H: float, alpha: float):
    m = C.c_int(); n = C.c_int()
    rc = lib().orc_best_grid(W, H, alpha, C.byref(m), C.byref(n))
    if rc != 0:
        raise ValueError(f"best_grid: invalid arguments (code {rc})")
    return m.value, n.value


def laplace3d_grad(s, t, n):
    return lib().orc_laplace3d_grad(_f64(s), _f64(t), _f64(n))


def laplace3d_pot(s, t):
    return lib().orc_laplace3d_pot(_f64(s), _f64(t))


def bary_row(x, w, xq):
    r = np.zeros(len(x))
    lib().orc_bary_row(len(x), _f64(x), _f64(w), float(xq), r)
    return r


# --------------------------------------------------------------------------- #
# Multi-box geometry (src/shape/box3d_multi.jl:4-386), restated in plain Python.
# Boxes are (center (x, y, z), Lx, Ly, Lz); points are 3-tuples of floats.
# --------------------------------------------------------------------------- #
def _norm3(v):
    return math.sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2])


def _dot3(a, b):
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]


def _jmin2(x, y):
    """Julia's min on floats: -0.0 counts as below 0.0."""
    return y if (y < x or (y == x and math.copysign(1.0, y) < 0)) else x


def _jmax2(x, y):
    return y if (y > x or (y == x and math.copysign(1.0, y) > 0)) else x


def _jmin(it):
    it = list(it)
    m = it[0]
    for v in it[1:]:
        m = _jmin2(m, v)
    return m


def _jmax(it):
    it = list(it)
    m = it[0]
    for v in it[1:]:
        m = _jmax2(m, v)
    return m


def box_faces_at_center(center, Lx, Ly, Lz):
    """_box3d_faces_at_center (src/shape/box3d_multi.jl:4-30): six (a, b, c, d, normal)."""
    cx, cy, cz = center
    hx, hy, hz = Lx / 2, Ly / 2, Lz / 2
    v = [None,
         (cx + hx, cy + hy, cz + hz), (cx - hx, cy + hy, cz + hz), (cx - hx, cy - hy, cz + hz),
         (cx + hx, cy - hy, cz + hz), (cx + hx, cy + hy, cz - hz), (cx - hx, cy + hy, cz - hz),
         (cx - hx, cy - hy, cz - hz), (cx + hx, cy - hy, cz - hz)]
    return [(v[1], v[2], v[3], v[4], (0.0, 0.0, 1.0)), (v[5], v[8], v[7], v[6], (0.0, 0.0, -1.0)),
            (v[8], v[5], v[1], v[4], (1.0, 0.0, 0.0)), (v[7], v[3], v[2], v[6], (-1.0, 0.0, 0.0)),
            (v[6], v[2], v[1], v[5], (0.0, 1.0, 0.0)), (v[7], v[8], v[4], v[3], (0.0, -1.0, 0.0))]


def _plane_axes(n):
    an = (abs(n[0]), abs(n[1]), abs(n[2]))
    if an[0] > 0.5:
        ax = (1, 2)
    elif an[1] > 0.5:
        ax = (0, 2)
    else:
        ax = (0, 1)
    nax = next(k for k in range(3) if an[k] > 0.5)
    return ax[0], ax[1], nax


def _point(ax1, ax2, nax, u, v, w):
    p = [0.0, 0.0, 0.0]
    p[ax1] = u
    p[ax2] = v
    p[nax] = w
    return tuple(p)


def rect_overlap_3d(a1, b1, c1, d1, n1, a2, b2, c2, d2, n2, tol=1e-10):
    """_rect_overlap_3d (src/shape/box3d_multi.jl:36-132) -> (bool, (a, b, c, d) or None)."""
    if _norm3((n1[0] + n2[0], n1[1] + n2[1], n1[2] + n2[2])) > tol:
        return False, None
    for q in (a2, b2, c2, d2):
        if abs(_dot3(n1, (q[0] - a1[0], q[1] - a1[1], q[2] - a1[2]))) > tol:
            return False, None
    ax1, ax2, nax = _plane_axes(n1)
    c1s, c2s = (a1, b1, c1, d1), (a2, b2, c2, d2)
    lo_u = _jmax2(_jmin(c[ax1] for c in c1s), _jmin(c[ax1] for c in c2s))
    hi_u = _jmin2(_jmax(c[ax1] for c in c1s), _jmax(c[ax1] for c in c2s))
    lo_v = _jmax2(_jmin(c[ax2] for c in c1s), _jmin(c[ax2] for c in c2s))
    hi_v = _jmin2(_jmax(c[ax2] for c in c1s), _jmax(c[ax2] for c in c2s))
    if hi_u - lo_u < tol or hi_v - lo_v < tol:
        return False, None
    w = a1[nax]
    e_ab = (b1[ax1] - a1[ax1], b1[ax2] - a1[ax2])
    e_ad = (d1[ax1] - a1[ax1], d1[ax2] - a1[ax2])
    mk = lambda u, v: _point(ax1, ax2, nax, u, v, w)  # noqa: E731
    if abs(e_ab[0]) > tol:
        us, ue = (lo_u, hi_u) if e_ab[0] >= 0 else (hi_u, lo_u)
        vs, ve = (lo_v, hi_v) if e_ad[1] >= 0 else (hi_v, lo_v)
        return True, (mk(us, vs), mk(ue, vs), mk(ue, ve), mk(us, ve))
    us, ue = (lo_u, hi_u) if e_ad[0] >= 0 else (hi_u, lo_u)
    vs, ve = (lo_v, hi_v) if e_ab[1] >= 0 else (hi_v, lo_v)
    return True, (mk(us, vs), mk(us, ve), mk(ue, ve), mk(ue, vs))


def detect_shared_faces_3d(boxes):
    """_detect_shared_faces_3d (src/shape/box3d_multi.jl:139-162) -> [(region, i, j, normal)]
    with 0-based i < j and the normal pointing from box j to box i."""
    out = []
    for i in range(len(boxes)):
        fi = box_faces_at_center(*boxes[i])
        for j in range(i + 1, len(boxes)):
            fj = box_faces_at_center(*boxes[j])
            for (a1, b1, c1, d1, n1) in fi:
                for (a2, b2, c2, d2, n2) in fj:
                    ok, reg = rect_overlap_3d(a1, b1, c1, d1, n1, a2, b2, c2, d2, n2)
                    if ok:
                        out.append((reg, i, j, (-n1[0], -n1[1], -n1[2])))
    return out


def _sorted_unique(xs):
    """sort!(unique!(xs)) with Julia's isequal/isless on floats (-0.0 before 0.0, kept apart)."""
    seen, out = set(), []
    for x in xs:
        key = (x, math.copysign(1.0, x))
        if key not in seen:
            seen.add(key)
            out.append(x)
    return sorted(out, key=lambda x: (x, math.copysign(1.0, x)))


def subtract_rects_from_face_3d(a, b, c, d, normal, shared, tol=1e-10):
    """_subtract_rects_from_face_3d (src/shape/box3d_multi.jl:169-308) -> [(a, b, c, d, edges, corners)]."""
    if not shared:
        return [(a, b, c, d, (True,) * 4, (True,) * 4)]
    ax1, ax2, nax = _plane_axes(normal)
    w = a[nax]
    fc = (a, b, c, d)
    fu0, fu1 = _jmin(p[ax1] for p in fc), _jmax(p[ax1] for p in fc)
    fv0, fv1 = _jmin(p[ax2] for p in fc), _jmax(p[ax2] for p in fc)
    us, vs = [fu0, fu1], [fv0, fv1]
    for r in shared:
        for q in r:
            us.append(q[ax1])
            vs.append(q[ax2])
    us = [u for u in _sorted_unique(us) if fu0 - tol <= u <= fu1 + tol]
    vs = [v for v in _sorted_unique(vs) if fv0 - tol <= v <= fv1 + tol]

    def is_shared(um, vm):
        for r in shared:
            if (_jmin(q[ax1] for q in r) - tol <= um <= _jmax(q[ax1] for q in r) + tol and
                    _jmin(q[ax2] for q in r) - tol <= vm <= _jmax(q[ax2] for q in r) + tol):
                return True
        return False

    nu, nv = len(us) - 1, len(vs) - 1
    sh = [[False] * nv for _ in range(nu)]
    for i in range(nu):
        for j in range(nv):
            if us[i + 1] - us[i] < tol or vs[j + 1] - vs[j] < tol:
                sh[i][j] = True
            else:
                sh[i][j] = is_shared((us[i] + us[i + 1]) / 2, (vs[j] + vs[j + 1]) / 2)

    def blocked(ci, cj):
        return ci < 0 or ci >= nu or cj < 0 or cj >= nv or sh[ci][cj]

    pc = [[any(blocked(gi + di, gj + dj) for di, dj in ((0, 0), (-1, 0), (0, -1), (-1, -1)))
           for gj in range(nv + 1)] for gi in range(nu + 1)]
    out = []
    for i in range(nu):
        for j in range(nv):
            if sh[i][j]:
                continue
            mk = lambda u, v: _point(ax1, ax2, nax, u, v, w)  # noqa: E731
            edges = (j == 0 or sh[i][j - 1], i == nu - 1 or sh[i + 1][j],
                     j == nv - 1 or sh[i][j + 1], i == 0 or sh[i - 1][j])
            corners = (pc[i][j], pc[i + 1][j], pc[i + 1][j + 1], pc[i][j + 1])
            out.append((mk(us[i], vs[j]), mk(us[i + 1], vs[j]), mk(us[i + 1], vs[j + 1]), mk(us[i], vs[j + 1]),
                        edges, corners))
    return out


def multi_box3d_face_regions(boxes, epses, eps_out):
    """_multi_box3d_face_regions (src/shape/box3d_multi.jl:313-356) ->
    [(a, b, c, d, normal, eps_in, eps_out, edges, corners)]: external pieces (box order, face
    order) then the shared faces."""
    shared = detect_shared_faces_3d(boxes)
    out = []
    for bid, box in enumerate(boxes):
        for (fa, fb, fc, fd, fn) in box_faces_at_center(*box):
            mine = []
            for (reg, lo, hi, _n) in shared:
                if bid in (lo, hi):
                    ok, ov = rect_overlap_3d(fa, fb, fc, fd, fn, *reg, (-fn[0], -fn[1], -fn[2]))
                    if ok:
                        mine.append(ov)
            for (ra, rb, rc, rd, ie, ic) in subtract_rects_from_face_3d(fa, fb, fc, fd, fn, mine):
                out.append((ra, rb, rc, rd, fn, epses[bid], eps_out, ie, ic))
    for (reg, lo, hi, n) in shared:
        out.append((*reg, n, epses[hi], epses[lo], (True,) * 4, (True,) * 4))
    return out


def rect_adaptive_panels(a, b, c, d, normal, edges, corners, alpha, l_ec):
    """rect_panel3d_adaptive_panels (src/shape/box3d.jl:99-133) -> (corners P x 4 x 3, is_edge P)."""
    abcd = np.array([*a, *b, *c, *d], np.float64)
    e = np.array([int(x) for x in edges], np.int32)
    k = np.array([int(x) for x in corners], np.int32)
    cap = 1024
    while True:
        cs = np.zeros(12 * cap)
        ie = np.zeros(cap, np.int32)
        n = lib().orc_rect_adaptive_panels(abcd, _f64(normal), e, k, float(l_ec), float(alpha), cap, cs, ie)
        if n < 0:
            raise ValueError("best_grid found no tiling")
        if n <= cap:
            return cs[:12 * n].reshape(n, 4, 3), ie[:n]
        cap = n


# --------------------------------------------------------------------------- #
# Screened volume sources (src/core/sources.jl:315-349, 408-515), plain Python loops.
# --------------------------------------------------------------------------- #
def _smooth(d, bw):
    """_smooth_screen_factor (src/core/sources.jl:321-323)."""
    return (1.0 + math.erf(d / bw)) / 2.0


def _box_screen(p, lo, hi, bw):
    """_box_screen (src/core/sources.jl:325-332)."""
    return (_smooth(p[0] - lo[0], bw) * _smooth(hi[0] - p[0], bw) * _smooth(p[1] - lo[1], bw) *
            _smooth(hi[1] - p[1], bw) * _smooth(p[2] - lo[2], bw) * _smooth(hi[2] - p[2], bw))


def _in_box(p, lo, hi, tol):
    """_point_in_box (src/core/sources.jl:315-319)."""
    return all(lo[k] - tol <= p[k] <= hi[k] + tol for k in range(3))


def screen_one_region(pos, rho, lo, hi, eps_in, eps_out, mode, bandwidth=0.0, tol=0.0):
    """_screened_volume_density (src/core/sources.jl:334-366): mode 'sharp', 'mix' or 'inv'."""
    out = np.zeros(len(rho))
    for s in range(len(rho)):
        p = tuple(pos[s])
        if mode == "sharp":
            e = eps_in if _in_box(p, lo, hi, tol) else eps_out
        else:
            sc = _box_screen(p, lo, hi, bandwidth)
            e = eps_in * sc + eps_out * (1.0 - sc) if mode == "mix" else 1.0 / (sc / eps_in + (1.0 - sc) / eps_out)
        out[s] = rho[s] / e
    return out


def screen_boxes(pos, rho, boxes, epses, eps_out, mode, bandwidth=0.0, tol=None):
    """_screened_volume_density_multibox (src/core/sources.jl:408-493) with the tol default of
    screened_volume_source(boxes, ...) (:495-515); boxes = [(center, Lx, Ly, Lz)]."""
    if tol is None:
        tol = math.sqrt(np.finfo(float).eps) * max(max(abs(b[1]), abs(b[2]), abs(b[3])) for b in boxes)
    los = [tuple(b[0][k] - b[1 + k] / 2 for k in range(3)) for b in boxes]
    his = [tuple(b[0][k] + b[1 + k] / 2 for k in range(3)) for b in boxes]
    out = np.zeros(len(rho))
    for s in range(len(rho)):
        p = tuple(pos[s])
        if mode == "sharp":
            e = eps_out
            for b in range(len(boxes)):
                if _in_box(p, los[b], his[b], tol):
                    e = epses[b]
                    break
            out[s] = rho[s] / e
            continue
        acc = 0.0
        sum_p = 0.0
        for b in range(len(boxes)):
            P = _box_screen(p, los[b], his[b], bandwidth)
            acc += epses[b] * P if mode == "mix" else P / epses[b]
            sum_p += P
        out[s] = rho[s] / (acc + eps_out * (1.0 - sum_p)) if mode == "mix" else rho[s] * (acc + (1.0 - sum_p) / eps_out)
    return out


class Interface:
    """A DielectricInterface restated on the CPU (uniform n_quad)."""

    @classmethod
    def multi_box(cls, n_quad, l_ec, boxes, epses, eps_out=1.0, alpha=float(np.sqrt(2.0))):
        """multi_dielectric_box3d (src/shape/box3d_multi.jl:358-386); boxes = [(center, Lx, Ly, Lz)]."""
        if len(boxes) != len(epses) or not boxes:
            raise ValueError("need one permittivity per box and at least one box")
        cs, ns, es, ei, eo = [], [], [], [], []
        for (ra, rb, rc, rd, fn, e_in, e_out, ie, ic) in multi_box3d_face_regions(boxes, epses, eps_out):
            c, e = rect_adaptive_panels(ra, rb, rc, rd, fn, ie, ic, alpha, l_ec)
            cs.append(c)
            es.append(e)
            ns.append(np.tile(np.asarray(fn, np.float64), (len(e), 1)))
            ei.append(np.full(len(e), float(e_in)))
            eo.append(np.full(len(e), float(e_out)))
        return cls.from_panels(np.concatenate(cs), np.concatenate(ns), n_quad, np.concatenate(ei),
                               np.concatenate(eo), np.concatenate(es))

    def __init__(self, handle):
        if not handle:
            raise ValueError("oracle interface construction failed")
        self.h = C.c_void_p(handle)
        P = C.c_int(); N = C.c_int(); nq = C.c_int()
        lib().orc_iface_sizes(self.h, C.byref(P), C.byref(N), C.byref(nq))
        self.P, self.N, self.nq = P.value, N.value, nq.value

    @classmethod
    def box(cls, Lx, Ly, Lz, n_quad, l_ec, eps_in, eps_out, alpha=float(np.sqrt(2.0))):
        """single_dielectric_box3d (src/shape/box3d_single.jl:49-70)."""
        return cls(lib().orc_iface_box(Lx, Ly, Lz, n_quad, l_ec, eps_in, eps_out, alpha))

    @classmethod
    def from_panels(cls, corners, normals, n_quad, eps_in, eps_out, is_edge=None):
        corners = _f64(corners).reshape(-1, 4, 3)
        P = corners.shape[0]
        ie = np.zeros(P, np.int32) if is_edge is None else np.ascontiguousarray(is_edge, dtype=np.int32)
        return cls(lib().orc_iface_from_panels(P, n_quad, corners.ravel(), _f64(normals).ravel(), ie,
                                               _f64(np.broadcast_to(eps_in, P)), _f64(np.broadcast_to(eps_out, P))))

    def __del__(self):
        try:
            if getattr(self, "h", None) and _lib is not None:
                _lib.orc_iface_free(self.h)
                self.h = None
        except Exception:
            pass

    def nodes(self):
        out = np.zeros(7 * self.N)
        lib().orc_iface_nodes(self.h, out)
        return out.reshape(7, self.N)

    def panels(self):
        c = np.zeros(12 * self.P); n = np.zeros(3 * self.P); e = np.zeros(self.P, np.int32)
        lib().orc_iface_panels(self.h, c, n, e)
        return c.reshape(self.P, 4, 3), n.reshape(self.P, 3), e

    def set_eps(self, eps_in, eps_out):
        lib().orc_iface_set_eps(self.h, _f64(np.broadcast_to(eps_in, self.P)), _f64(np.broadcast_to(eps_out, self.P)))

    def diag(self):
        t = np.zeros(self.N)
        lib().orc_diag(self.h, t)
        return t

    def _dense(self, name):
        M = np.zeros(self.N * self.N)
        getattr(lib(), name)(self.h, M)
        return M.reshape(self.N, self.N).T  # stored column-major

    def dense_DT(self):
        return self._dense("orc_dense_DT")

    def dense_S(self):
        return self._dense("orc_dense_S")

    def dense_D(self):
        return self._dense("orc_dense_D")

    def lhs_dense(self):
        return self._dense("orc_lhs_dense")

    def apply_direct(self, X, mode=0):
        X = np.asarray(X, np.float64)
        vec = X.ndim == 1
        Xm = X.reshape(self.N, -1)
        K = Xm.shape[1]
        Y = np.zeros(self.N * K)
        lib().orc_apply_direct(self.h, mode, K, _f64(Xm.T).ravel(), Y)
        Y = Y.reshape(K, self.N).T
        return Y[:, 0] if vec else Y

    def apply_DT_rows(self, rows, x):
        rows = np.ascontiguousarray(rows, dtype=np.int32)
        y = np.zeros(len(rows))
        lib().orc_apply_DT_rows(self.h, len(rows), rows, _f64(x), y)
        return y

    def pottrg(self, targets, X):
        targets = _f64(targets).reshape(-1, 3)
        X = np.asarray(X, np.float64)
        vec = X.ndim == 1
        Xm = X.reshape(self.N, -1)
        M, K = targets.shape[0], Xm.shape[1]
        out = np.zeros(M * K)
        lib().orc_pottrg(self.h, M, targets.ravel(), K, _f64(Xm.T).ravel(), out)
        out = out.reshape(K, M).T
        return out[:, 0] if vec else out

    def rhs_points(self, src, q, eps_src):
        src = _f64(src).reshape(-1, 3)
        q = _f64(np.atleast_1d(q))
        K = src.shape[0]
        F = np.zeros(self.N * K)
        lib().orc_rhs_points(self.h, K, src.ravel(), q, float(eps_src), F)
        return F.reshape(K, self.N).T

    def rhs_volume(self, pos, w, rho, eps_src):
        pos = _f64(pos).reshape(-1, 3)
        F = np.zeros(self.N)
        lib().orc_rhs_volume(self.h, pos.shape[0], pos.ravel(), _f64(w), _f64(rho), float(eps_src), F)
        return F

    def screen_sharp(self, pos, rho):
        pos = _f64(pos).reshape(-1, 3)
        out = np.zeros(pos.shape[0])
        if lib().orc_screen_sharp(self.h, pos.shape[0], pos.ravel(), _f64(rho), out) != 0:
            raise ValueError("screened_volume_source requires a uniform eps across the interface")
        return out

    def bernstein_rho_panel(self, p, trg):
        return lib().orc_bernstein_rho_panel(self.h, int(p), _f64(trg))

    def neighbor_list(self, max_order=8, atol=1e-9, correct_edges=True):
        return NeighborList(self, lib().orc_neighbor_list(self.h, max_order, atol, int(bool(correct_edges))))


class NeighborList:
    """build_neighbor_list output (src/kernel/laplace3d_near.jl:240-392), 0-based panels."""

    def __init__(self, iface, handle):
        self.iface = iface
        self.h = C.c_void_p(handle)
        nu = C.c_int(); na = C.c_int()
        lib().orc_nlist_sizes(self.h, C.byref(nu), C.byref(na))
        self.n_up, self.n_adp = nu.value, na.value
        self.up_i = np.zeros(self.n_up, np.int32); self.up_j = np.zeros(self.n_up, np.int32)
        self.up_p = np.zeros(self.n_up, np.int32)
        self.adp_i = np.zeros(self.n_adp, np.int32); self.adp_j = np.zeros(self.n_adp, np.int32)
        lib().orc_nlist_get(self.h, self.up_i, self.up_j, self.up_p, self.adp_i, self.adp_j)
        self.blocks = None

    def __del__(self):
        try:
            if getattr(self, "h", None) and _lib is not None:
                _lib.orc_nlist_free(self.h)
                self.h = None
        except Exception:
            pass

    @property
    def upsample(self):
        return {(int(i), int(j)): int(p) for i, j, p in zip(self.up_i, self.up_j, self.up_p)}

    @property
    def adaptive(self):
        return {(int(i), int(j)) for i, j in zip(self.adp_i, self.adp_j)}

    def corrections(self, adp_atol=1e-9, adp_rtol=float(np.sqrt(np.finfo(float).eps)), adp_nGL=0, adp_max_depth=20,
                    mode="DT"):
        """laplace3d_DT_corrections (mode "DT") or laplace3d_D_corrections (mode "D")
        (src/kernel/laplace3d_near_upsampling.jl:127-282); the blocks are kept for apply()."""
        npp = self.iface.nq ** 2
        nb = self.n_up + self.n_adp
        out = np.zeros(max(nb, 1) * npp * npp)
        m = {"DT": 0, "D": 1}[mode]
        self.hits = lib().orc_corrections_mode(self.iface.h, self.h, m, adp_atol, adp_rtol, adp_nGL, adp_max_depth,
                                               out)
        self.blocks = out[: nb * npp * npp].reshape(nb, npp, npp).transpose(0, 2, 1)  # [b, t, c]
        return self.blocks

    def pairs(self):
        """(src panel, trg panel) for each block, upsampled first then adaptive."""
        return np.concatenate([self.up_i, self.adp_i]), np.concatenate([self.up_j, self.adp_j])

    def apply(self, X):
        X = np.asarray(X, np.float64)
        vec = X.ndim == 1
        Xm = X.reshape(self.iface.N, -1)
        K = Xm.shape[1]
        blk = _f64(self.blocks.transpose(0, 2, 1)).ravel()
        Y = np.zeros(self.iface.N * K)
        lib().orc_corrections_apply(self.iface.h, self.h, blk, K, _f64(Xm.T).ravel(), Y)
        Y = Y.reshape(K, self.iface.N).T
        return Y[:, 0] if vec else Y

    def sparse(self):
        """The correction matrix as scipy CSR (rows = targets, cols = sources)."""
        import scipy.sparse as sp
        npp = self.iface.nq ** 2
        src, trg = self.pairs()
        r = (trg[:, None, None] * npp + np.arange(npp)[None, :, None]) + 0 * np.arange(npp)[None, None, :]
        c = (src[:, None, None] * npp + np.arange(npp)[None, None, :]) + 0 * np.arange(npp)[None, :, None]
        return sp.csr_matrix((self.blocks.ravel(), (r.ravel(), c.ravel())), shape=(self.iface.N, self.iface.N))


class CorrectedOperator:
    """y = Dᵀ x (direct sum) + C x + t∘x : the FP64 operator of
    lhs_dielectric_box3d_fmm3d_corrected / BatchedDielectricOperator
    (src/solver/dielectric_box3d.jl:48-92, src/solver/multi_rhs.jl:60-130) with the
    FMM replaced by its exact direct sum."""

    def __init__(self, iface, max_order=8, up_tol=1e-9, correct_edges=True, adaptive_atol=None,
                 adaptive_rtol=float(np.sqrt(np.finfo(float).eps)), adaptive_n_GL=0, adaptive_max_depth=20):
        self.iface = iface
        self.nl = iface.neighbor_list(max_order, up_tol, correct_edges)
        self.nl.corrections(up_tol if adaptive_atol is None else adaptive_atol, adaptive_rtol, adaptive_n_GL,
                            adaptive_max_depth)
        self.Cs = self.nl.sparse()
        self.t = iface.diag()
        self.N = iface.N

    def __matmul__(self, X):
        X = np.asarray(X, np.float64)
        Y = self.iface.apply_direct(X, 0)
        Y = Y + self.Cs @ X + (self.t[:, None] * X if X.ndim == 2 else self.t * X)
        return Y


def solve_lu(A, b):
    """solve_lu (src/utils/linear_algebra.jl:1): LAPACK LU with partial pivoting."""
    import scipy.linalg as sl
    return sl.lu_solve(sl.lu_factor(A), b)


def gmres(apply, b, rtol=1e-10, atol=0.0, itmax=500, M=None):
    """Full (unrestarted) GMRES with modified Gram-Schmidt and Givens rotations.
    ``M`` is an optional right preconditioner (vector of diagonal entries)."""
    b = np.asarray(b, np.float64)
    n = b.shape[0]
    beta = np.linalg.norm(b)
    if beta == 0.0:
        return np.zeros(n), 0, 0.0
    tol = atol + rtol * beta
    V = np.zeros((itmax + 1, n))
    H = np.zeros((itmax + 1, itmax))
    cs = np.zeros(itmax); sn = np.zeros(itmax)
    g = np.zeros(itmax + 1); g[0] = beta
    V[0] = b / beta
    k = 0
    res = beta
    for k in range(itmax):
        z = V[k] * M if M is not None else V[k]
        w = apply(z)
        for i in range(k + 1):
            H[i, k] = V[i] @ w
            w = w - H[i, k] * V[i]
        H[k + 1, k] = np.linalg.norm(w)
        if H[k + 1, k] != 0.0:
            V[k + 1] = w / H[k + 1, k]
        for i in range(k):
            tmp = cs[i] * H[i, k] + sn[i] * H[i + 1, k]
            H[i + 1, k] = -sn[i] * H[i, k] + cs[i] * H[i + 1, k]
            H[i, k] = tmp
        d = np.hypot(H[k, k], H[k + 1, k])
        cs[k] = H[k, k] / d; sn[k] = H[k + 1, k] / d
        H[k, k] = d; H[k + 1, k] = 0.0
        g[k + 1] = -sn[k] * g[k]; g[k] = cs[k] * g[k]
        res = abs(g[k + 1])
        if res <= tol:
            k += 1
            break
    else:
        k = itmax
    y = np.linalg.solve(np.triu(H[:k, :k]), g[:k])
    x = V[:k].T @ y
    if M is not None:
        x = x * M
    return x, k, res / beta


class TargetMap:
    """laplace3d_pottrg_fmm3d_corrected_hcubature restated on the CPU
    (src/kernel/laplace3d_near_hcubature.jl:337-375), direct sum for the FMM."""

    def __init__(self, iface: "Interface", targets, atol: float, range_factor: float = 5.0):
        t = _f64(np.asarray(targets, np.float64).reshape(-1, 3))
        self.M = t.shape[0]
        self.N = iface.N
        self._iface = iface
        self.h = C.c_void_p(lib().orc_tmap_build(iface.h, self.M, t.ravel(), atol, range_factor))
        Nr = C.c_int(); Pr = C.c_int(); nnz = C.c_long(); npairs = C.c_long(); ncap = C.c_long()
        lib().orc_tmap_sizes(self.h, C.byref(Nr), C.byref(Pr), C.byref(nnz), C.byref(npairs), C.byref(ncap))
        self.N_refined, self.P_refined = Nr.value, Pr.value
        self.nnz, self.n_pairs, self.n_capped = nnz.value, npairs.value, ncap.value

    def __del__(self):
        try:
            if getattr(self, "h", None) and _lib is not None:
                _lib.orc_tmap_free(self.h)
                self.h = None
        except Exception:
            pass

    def parts(self):
        """(refined nodes 4 x Nr, parent, from_split, P as (ptr, col, val), C as (ptr, col, val))."""
        Nr, Pr, M = self.N_refined, self.P_refined, self.M
        npp = Nr // max(Pr, 1)
        xyzw = np.zeros(4 * Nr); par = np.zeros(Pr, np.int32); spl = np.zeros(Pr, np.int32)
        pp = np.zeros(Nr + 1, np.int32); pc = np.zeros(Nr * npp, np.int32); pv = np.zeros(Nr * npp)
        cp = np.zeros(M + 1, np.int64); cc = np.zeros(max(self.nnz, 1), np.int32); cv = np.zeros(max(self.nnz, 1))
        lib().orc_tmap_get(self.h, xyzw, par, spl, pp, pc, pv, cp, cc, cv)
        n = int(pp[-1])
        return (xyzw.reshape(4, Nr), par, spl, (pp, pc[:n], pv[:n]), (cp, cc[:self.nnz], cv[:self.nnz]))

    def apply(self, S):
        S = np.asarray(S, np.float64).reshape(self.N, -1)
        K = S.shape[1]
        U = np.zeros(self.M * K)
        lib().orc_tmap_apply(self.h, K, _f64(S.T).ravel(), U)
        return U.reshape(K, self.M).T

    def reaction(self, grp_ptr, q, S):
        S = np.asarray(S, np.float64).reshape(self.N, -1)
        K = S.shape[1]
        grp = np.ascontiguousarray(grp_ptr, dtype=np.int64)
        G = len(grp) - 1
        R = np.zeros(G * K)
        lib().orc_reaction(self.h, G, grp, _f64(q), K, _f64(S.T).ravel(), R)
        return R.reshape(K, G).T


def pot_panel_hcub(iface: "Interface", p: int, target, atol: float):
    """Adaptive-cubature potential of panel p against its n_quad^2 Lagrange basis at one target
    (laplace3d_pot_panel_hcubature, src/kernel/laplace3d_near_hcubature.jl:86-150)."""
    I = np.zeros(iface.nq ** 2)
    ok = lib().orc_pot_panel_hcub(iface.h, p, _f64(target), atol, I)
    return I, bool(ok)
