"""CPU-side checks of the product library: it loads, exports every symbol of include/dbie.h,
and its host setup (mesher, near-pair list) equals the oracle bit for bit.  No GPU needed."""
import ctypes as C
import math
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "dbie.h")

BOXES = [
    (1.2, 0.8, 0.6, 4, 0.2),      # README quickstart (config 1)
    (2.4, 1.6, 1.2, 4, 0.05),     # configs 2/3
    (1.2, 0.8, 0.6, 5, 0.03),     # config 5
    (1.0, 1.0, 1.0, 2, 0.2),
    (0.5, 0.5, 0.5, 4, 0.1),      # smallest config-4 geometry
    (3.0, 3.0, 1.0, 6, 0.2),
]


def header_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(dbie_[a-z0-9_]+)\(", text)))


def test_library_exports_every_header_symbol(dbie):
    lib = C.CDLL(dbie.LIB_PATH)
    syms = header_symbols()
    assert len(syms) >= 25
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing


def test_python_binding_covers_header(dbie):
    from dbie import _lib
    assert set(header_symbols()) <= set(_lib.SIGNATURES)


@pytest.mark.parametrize("box", BOXES)
def test_geometry_bit_exact_vs_oracle(dbie, O, box):
    g = dbie.single_dielectric_box3d(*box, 4.0, 1.0)
    o = O.Interface.box(*box, 4.0, 1.0)
    assert (g.n_panels, g.n_points) == (o.P, o.N)
    assert np.array_equal(g.nodes, o.nodes())
    c1, n1, e1 = g.panels()
    c2, n2, e2 = o.panels()
    assert np.array_equal(c1, c2) and np.array_equal(n1, n2) and np.array_equal(e1, e2.astype(bool))


@pytest.mark.parametrize("box", BOXES[:4] + BOXES[5:])
def test_neighbor_list_equals_oracle(dbie, O, box):
    g = dbie.single_dielectric_box3d(*box, 4.0, 1.0)
    o = O.Interface.box(*box, 4.0, 1.0)
    for correct_edges in (True, False):
        up, ad = dbie.build_neighbor_list(g, 8, 1e-9, correct_edges=correct_edges)
        nl = o.neighbor_list(8, 1e-9, correct_edges)
        assert up == nl.upsample
        assert ad == nl.adaptive


def test_gauss_legendre_and_best_grid_match_oracle(dbie, O):
    for n in range(1, 9):
        x, w = dbie.gausslegendre(n)
        xo, wo = O.gauss_legendre(n)
        assert np.array_equal(x, xo) and np.array_equal(w, wo)
    assert dbie.best_grid_mn(20.0, 1.0, math.sqrt(2)) == (15, 1)
    assert dbie.best_grid_mn(3.0, 2.0, 1.0) == (3, 2)
    # extreme aspect ratios, whose continued fractions run past the denominator bound of 10^6
    rng = np.random.default_rng(3)
    cases = [(1.0, 1e-9, 1.1), (1.0, 3e-7, 1.05), (math.pi, 1.0, 1.0001), (1e6, 1.0, 1.5), (1.0, 1e-15, 2.0)]
    cases += [tuple(rng.uniform(0.01, 100, 2)) + (float(rng.uniform(1, 1.01)),) for _ in range(200)]
    for W, H, a in cases:
        try:
            got = dbie.best_grid_mn(W, H, a)
        except ValueError:
            got = None
        try:
            want = O.best_grid(W, H, a)
        except ValueError:
            want = None
        assert got == want, (W, H, a)
    with pytest.raises(ValueError):
        dbie.best_grid_mn(0.0, 1.0, 1.0)
    with pytest.raises(ValueError):
        dbie.best_grid_mn(1.0, 1.0, 0.99)


def test_from_panels_matches_oracle(dbie, O):
    corners = np.array([[(-0.5, -0.5, 0.0), (0.5, -0.5, 0.0), (0.5, 0.5, 0.0), (-0.5, 0.5, 0.0)],
                        [(-0.5, -0.5, 0.2), (0.5, -0.5, 0.2), (0.5, 0.5, 0.2), (-0.5, 0.5, 0.2)]])
    normals = np.array([[0, 0, 1.0], [0, 0, 1.0]])
    g = dbie.interface_from_panels(corners, normals, 2, 2.0, 1.0)
    o = O.Interface.from_panels(corners, normals, 2, 2.0, 1.0)
    assert np.array_equal(g.nodes, o.nodes())
    up, ad = dbie.build_neighbor_list(g, 12, 1e-3, correct_edges=True)
    nl = o.neighbor_list(12, 1e-3, True)
    assert up == nl.upsample and ad == nl.adaptive


def test_bad_arguments_raise(dbie):
    with pytest.raises(ValueError):
        dbie.single_dielectric_box3d(-1.0, 1.0, 1.0, 4, 0.2, 4.0, 1.0)
    with pytest.raises(ValueError):
        dbie.single_dielectric_box3d(1.0, 1.0, 1.0, 0, 0.2, 4.0, 1.0)
    with pytest.raises(ValueError):
        dbie.single_dielectric_box3d(1.0, 1.0, 1.0, 4, 0.0, 4.0, 1.0)
    g = dbie.single_dielectric_box3d(1.0, 1.0, 1.0, 2, 0.5, 4.0, 1.0)
    with pytest.raises(ValueError):
        dbie.build_neighbor_list(g, 8, 2.0)
    from dbie import _lib
    assert _lib.lib().dbie_autotune(5, 0) == _lib.BAD_ARG
    assert _lib.lib().dbie_autotune(-1, 1) == 0


def test_no_cpu_fallback_without_gpu(dbie):
    """On a host without a gfx950 GPU the device entry points fail loudly."""
    import subprocess
    import sys
    code = ("import sys; sys.path.insert(0, %r); import dbie; from dbie import _lib;"
            "g = dbie.single_dielectric_box3d(1.0,1.0,1.0,2,0.5,4.0,1.0)\n"
            "try:\n  dbie.BatchedDielectricOperator(g)\n  print('HAS_GPU')\n"
            "except dbie.DbieError as e:\n  print('NO_DEVICE' if e.code == _lib.NO_DEVICE else 'OTHER', e)\n"
            % os.path.join(ROOT, "dielectricbie.jl_amd"))
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120).stdout
    assert ("NO_DEVICE" in out) or ("HAS_GPU" in out), out
