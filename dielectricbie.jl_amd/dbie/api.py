"""Host-side mirror of DielectricBIE.jl's boundary-operator API, backed by libdbie on the GPU.

Names, argument meaning and error behaviour follow the reference (Julia) functions
cited in each docstring; Julia's 1-based panel/node indices become 0-based here.
Arrays are numpy float64.  Every compute call goes through the C ABI of
libdbie.so; nothing here computes kernel sums on the CPU.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

from . import _lib as L

SQRT_EPS = math.sqrt(np.finfo(np.float64).eps)


# --------------------------------------------------------------------------- #
# Utilities
# --------------------------------------------------------------------------- #
def gausslegendre(n: int):
    """FastGaussQuadrature.gausslegendre(n): ascending nodes and weights on [-1, 1]."""
    x = np.zeros(n); w = np.zeros(n)
    L.check(L.lib().dbie_gauss_legendre(int(n), L.dptr(x), L.dptr(w)))
    return x, w


def best_grid_mn(W: float, H: float, alpha: float):
    """best_grid_mn (src/utils/best_grid.jl:123-126); raises ValueError (ArgumentError) on bad input."""
    m = C.c_int(); n = C.c_int()
    L.check(L.lib().dbie_best_grid(float(W), float(H), float(alpha), C.byref(m), C.byref(n)))
    return m.value, n.value


best_grid = best_grid_mn


# --------------------------------------------------------------------------- #
# Sources (src/core/sources.jl)
# --------------------------------------------------------------------------- #
@dataclass
class PointSource:
    """PointSource{T,3}(point, charge) (src/core/sources.jl:5-8)."""
    point: tuple
    charge: float


@dataclass
class VolumeSource:
    """VolumeSource (src/core/sources.jl:20-32): positions 3 x S, weights S, density S."""
    positions: np.ndarray
    weights: np.ndarray
    density: np.ndarray

    def __post_init__(self):
        self.positions = np.asarray(self.positions, np.float64)
        if self.positions.ndim == 2 and self.positions.shape[0] != 3 and self.positions.shape[1] == 3:
            self.positions = self.positions.T
        self.positions = np.ascontiguousarray(self.positions.reshape(3, -1))
        self.weights = np.ascontiguousarray(self.weights, np.float64).ravel()
        self.density = np.ascontiguousarray(self.density, np.float64).ravel()
        S = self.positions.shape[1]
        if len(self.weights) != S or len(self.density) != S:
            raise ValueError("points, weights and density must have the same length")

    @classmethod
    def from_points(cls, points, weights, density, tol: float = 0.0):
        """VolumeSource(points::Vector{NTuple{3}}, weights, density; tol) (src/core/sources.jl:249-261)
        with the |density| >= tol truncation of _truncate_volume_source (:100-112)."""
        pos = np.asarray(points, np.float64).reshape(-1, 3).T
        w = np.asarray(weights, np.float64); rho = np.asarray(density, np.float64)
        if tol < 0:
            raise ValueError("tol must be >= 0")
        if tol > 0:
            keep = np.abs(rho) >= tol
            if not keep.any():
                raise ValueError(f"all source points are truncated by tol={tol}")
            pos, w, rho = pos[:, keep], w[keep], rho[keep]
        return cls(pos, w, rho)

    @classmethod
    def point_charge(cls, point, charge: float):
        """A point charge as a one-point VolumeSource (src/core/sources.jl:249)."""
        return cls(np.asarray(point, np.float64).reshape(3, 1), np.ones(1), np.array([float(charge)]))


class SharpScreening:
    """SharpScreening (src/core/sources.jl:265): rho / eps_in inside the box, rho / eps_out outside."""
    code = 0
    bandwidth = 0.0


class SoftMixPermittivity:
    """SoftMixPermittivity(bandwidth) (src/core/sources.jl:267-273): eps mixed by the box's smooth
    indicator (products of (1 + erf(d / bandwidth)) / 2 over the six faces)."""
    code = 1

    def __init__(self, bandwidth: float):
        if not bandwidth > 0:
            raise ValueError("bandwidth must be positive")
        self.bandwidth = float(bandwidth)


class SoftMixInversePermittivity(SoftMixPermittivity):
    """SoftMixInversePermittivity(bandwidth) (src/core/sources.jl:275-281): 1/eps mixed instead."""
    code = 2


def _screen_mode(mode):
    if mode is None:
        raise TypeError("screened_volume_source requires a screening mode (MethodError in the reference)")
    if not isinstance(mode, (SharpScreening, SoftMixPermittivity)):
        raise TypeError(f"unknown screening mode {type(mode).__name__}")
    return mode.code, mode.bandwidth


def box_array(boxes) -> np.ndarray:
    """n x 6 (centre x, y, z, Lx, Ly, Lz) from the reference's (center = (x, y, z), Lx, Ly, Lz)
    boxes, given as dicts, objects with those attributes or (center, Lx, Ly, Lz) tuples."""
    rows = []
    for b in boxes:
        if isinstance(b, dict):
            c, lx, ly, lz = b["center"], b["Lx"], b["Ly"], b["Lz"]
        elif hasattr(b, "center"):
            c, lx, ly, lz = b.center, b.Lx, b.Ly, b.Lz
        else:
            c, lx, ly, lz = b
        rows.append([float(c[0]), float(c[1]), float(c[2]), float(lx), float(ly), float(lz)])
    return np.ascontiguousarray(np.array(rows, np.float64).reshape(-1, 6))


# --------------------------------------------------------------------------- #
# Geometry (src/core/panels.jl, src/shape/box3d_single.jl)
# --------------------------------------------------------------------------- #
class DielectricInterface:
    """DielectricInterface of flat tensor-Gauss-Legendre panels (src/core/panels.jl:64-68).

    Node order: panel order, then (i outer along b-a, j inner along d-a)."""

    def __init__(self, handle, eps_in, eps_out):
        self._h = C.c_void_p(handle)
        P = C.c_int64(); N = C.c_int64(); nq = C.c_int()
        L.check(L.lib().dbie_geom_sizes(self._h, C.byref(P), C.byref(N), C.byref(nq)))
        self.n_panels, self.n_points, self.n_quad = P.value, N.value, nq.value
        self.eps_in = np.ascontiguousarray(np.broadcast_to(np.asarray(eps_in, np.float64), (self.n_panels,)))
        self.eps_out = np.ascontiguousarray(np.broadcast_to(np.asarray(eps_out, np.float64), (self.n_panels,)))
        self._nodes = None
        self._panels = None
        self._ops = {}

    def __del__(self):
        try:
            if getattr(self, "_h", None) and L._lib is not None:
                L._lib.dbie_geom_destroy(self._h)
                self._h = None
        except Exception:
            pass

    def __repr__(self):
        return f"DielectricInterface with {self.n_panels} panels in Float64 ({self.n_points} points)"

    @property
    def nodes(self) -> np.ndarray:
        """7 x N SoA: x, y, z, nx, ny, nz, w."""
        if self._nodes is None:
            out = np.zeros(7 * self.n_points)
            L.check(L.lib().dbie_geom_nodes(self._h, L.dptr(out)))
            self._nodes = out.reshape(7, self.n_points)
        return self._nodes

    @property
    def points(self) -> np.ndarray:
        return self.nodes[0:3]

    @property
    def normals(self) -> np.ndarray:
        return self.nodes[3:6]

    @property
    def weights(self) -> np.ndarray:
        return self.nodes[6]

    def panels(self):
        """(corners P x 4 x 3, normals P x 3, is_edge P)."""
        if self._panels is None:
            c = np.zeros(12 * self.n_panels); n = np.zeros(3 * self.n_panels); e = np.zeros(self.n_panels, np.int32)
            L.check(L.lib().dbie_geom_panels(self._h, L.dptr(c), L.dptr(n), L.iptr(e)))
            self._panels = (c.reshape(-1, 4, 3), n.reshape(-1, 3), e.astype(bool))
        return self._panels

    def diag(self) -> np.ndarray:
        t = 0.5 * (self.eps_out + self.eps_in) / (self.eps_out - self.eps_in)
        return np.repeat(t, self.n_quad ** 2)


def num_points(interface: DielectricInterface) -> int:
    """num_points (src/core/panels.jl:117)."""
    return interface.n_points


def all_weights(interface: DielectricInterface) -> np.ndarray:
    """all_weights (src/core/panels.jl:118)."""
    return interface.weights.copy()


def single_dielectric_box3d(Lx, Ly, Lz, n_quad: int, l_ec, eps_in, eps_out, T=np.float64, *, alpha=math.sqrt(2.0)):
    """single_dielectric_box3d (src/shape/box3d_single.jl:49-70)."""
    if T not in (np.float64, float):
        raise TypeError("only Float64 geometry is supported")
    h = C.c_void_p()
    L.check(L.lib().dbie_geometry_box(float(Lx), float(Ly), float(Lz), int(n_quad), float(l_ec), float(alpha),
                                      C.byref(h)))
    return DielectricInterface(h.value, eps_in, eps_out)


def multi_dielectric_box3d(n_quad: int, l_ec, boxes, epses, eps_out=1.0, *, alpha=math.sqrt(2.0)):
    """multi_dielectric_box3d (src/shape/box3d_multi.jl:358-386): touching axis-aligned boxes with
    permittivity epses[b] inside box b and eps_out outside; shared faces carry
    (eps_in, eps_out) = (eps of the higher-numbered box, eps of the lower)."""
    B = box_array(boxes)
    e = np.ascontiguousarray(epses, np.float64).ravel()
    if len(e) != B.shape[0]:
        raise ValueError("Number of boxes must match number of permittivities")
    h = C.c_void_p()
    L.check(L.lib().dbie_geometry_multi_box(int(n_quad), float(l_ec), B.shape[0], L.dptr(B.ravel()), L.dptr(e),
                                            float(eps_out), float(alpha), C.byref(h)))
    P = C.c_int64()
    L.check(L.lib().dbie_geom_sizes(h, C.byref(P), None, None))
    ei = np.zeros(P.value); eo = np.zeros(P.value)
    L.check(L.lib().dbie_geom_eps(h, L.dptr(ei), L.dptr(eo)))
    return DielectricInterface(h.value, ei, eo)


def interface_from_panels(corners, normals, n_quad: int, eps_in, eps_out, is_edge=None, *, points=None,
                          weights=None, gl_xs=None, gl_ws=None, bary_weights=None):
    """DielectricInterface(panels, eps_in, eps_out) of rect_panel3d_discretize panels
    (src/shape/box3d.jl:22-64); corners P x 4 x 3 ordered (a, b, c, d).  With points (N x 3),
    weights (N), gl_xs and gl_ws (n_quad; bary_weights optional) the panels' own quadrature
    (FlatPanel fields, src/core/panels.jl:5-22) is used as given instead of being regenerated."""
    c = np.ascontiguousarray(corners, np.float64).reshape(-1, 12)
    P = c.shape[0]
    n = np.ascontiguousarray(normals, np.float64).reshape(P, 3)
    e = np.zeros(P, np.int32) if is_edge is None else np.ascontiguousarray(is_edge, np.int32)
    h = C.c_void_p()
    if points is None:
        L.check(L.lib().dbie_geometry_from_panels(P, int(n_quad), L.dptr(c.ravel()), L.dptr(n.ravel()), L.iptr(e),
                                                  C.byref(h)))
    else:
        Npts = P * int(n_quad) ** 2
        pts = np.ascontiguousarray(points, np.float64).reshape(Npts, 3)
        w = np.ascontiguousarray(weights, np.float64).reshape(Npts)
        gx = np.ascontiguousarray(gl_xs, np.float64).reshape(int(n_quad))
        gw = np.ascontiguousarray(gl_ws, np.float64).reshape(int(n_quad))
        bw = None if bary_weights is None else L.dptr(np.ascontiguousarray(bary_weights, np.float64).reshape(int(n_quad)))
        L.check(L.lib().dbie_geometry_from_panel_nodes(P, int(n_quad), L.dptr(c.ravel()), L.dptr(n.ravel()),
                                                       L.iptr(e), L.dptr(pts.ravel()), L.dptr(w), L.dptr(gx),
                                                       L.dptr(gw), bw, C.byref(h)))
    return DielectricInterface(h.value, eps_in, eps_out)


def build_neighbor_list(interface: DielectricInterface, max_order: int, atol: float, *, correct_edges: bool = False):
    """build_neighbor_list (src/kernel/laplace3d_near.jl:240-392) -> (upsample, adaptive) with 0-based
    (source panel, target panel) keys; upsample maps to p_up."""
    nu = C.c_int64(); na = C.c_int64()
    lib = L.lib()
    L.check(lib.dbie_neighbor_list(interface._h, int(max_order), float(atol), int(bool(correct_edges)),
                                   C.byref(nu), C.byref(na), None, None, None, None, None))
    us = np.zeros(nu.value, np.int32); ut = np.zeros(nu.value, np.int32); uo = np.zeros(nu.value, np.int32)
    asrc = np.zeros(na.value, np.int32); atrg = np.zeros(na.value, np.int32)
    L.check(lib.dbie_neighbor_list(interface._h, int(max_order), float(atol), int(bool(correct_edges)),
                                   C.byref(nu), C.byref(na), L.iptr(us), L.iptr(ut), L.iptr(uo), L.iptr(asrc),
                                   L.iptr(atrg)))
    upsample = {(int(i), int(j)): int(p) for i, j, p in zip(us, ut, uo)}
    adaptive = {(int(i), int(j)) for i, j in zip(asrc, atrg)}
    return upsample, adaptive


# --------------------------------------------------------------------------- #
# Operators
# --------------------------------------------------------------------------- #
class BatchedDielectricOperator:
    """BatchedDielectricOperator (src/solver/multi_rhs.jl:46-141), resident in HBM.

    ``op @ X`` / ``mul(Y, X)`` apply A = Dᵀ + C + diag(t) to a vector or an N x K block.
    The far field is the exact direct sum (the reference's FMM at fmm_tol approximates it)."""

    def __init__(self, interface: DielectricInterface, *, corrected=True, up_tol=1e-9, max_order=8,
                 correct_edges=True, adaptive_atol=None, adaptive_rtol=SQRT_EPS, adaptive_n_GL=0,
                 adaptive_max_depth=20, n_threads=0, device_build=True):
        o = L.NearOpts()
        L.lib().dbie_near_opts_default(C.byref(o))
        o.up_tol = float(up_tol)
        o.max_order = int(max_order)
        o.correct_edges = int(bool(correct_edges))
        o.adaptive_atol = float(up_tol if adaptive_atol is None else adaptive_atol)
        o.adaptive_rtol = float(adaptive_rtol)
        o.adaptive_n_GL = int(adaptive_n_GL)
        o.adaptive_max_depth = int(adaptive_max_depth)
        o.build_corrections = int(bool(corrected))
        o.n_threads = int(n_threads)
        o.device_build = int(bool(device_build))
        h = C.c_void_p()
        L.check(L.lib().dbie_operator_create(interface._h, L.dptr(interface.eps_in), L.dptr(interface.eps_out),
                                             C.byref(o), C.byref(h)))
        self._h = C.c_void_p(h.value)
        self.interface = interface
        self.n = interface.n_points
        N = C.c_int64(); nb = C.c_int64(); bd = C.c_int()
        L.check(L.lib().dbie_op_sizes(self._h, C.byref(N), C.byref(nb), C.byref(bd)))
        self.n_blocks, self.block_dim = nb.value, bd.value

    def __del__(self):
        try:
            if getattr(self, "_h", None) and L._lib is not None:
                L._lib.dbie_operator_destroy(self._h)
                self._h = None
        except Exception:
            pass

    @property
    def shape(self):
        return (self.n, self.n)

    def size(self, d=None):
        return self.shape if d is None else (self.n if d <= 2 else 1)

    dtype = np.float64

    def set_eps(self, eps_in, eps_out):
        ei = np.ascontiguousarray(np.broadcast_to(np.asarray(eps_in, np.float64), (self.interface.n_panels,)))
        eo = np.ascontiguousarray(np.broadcast_to(np.asarray(eps_out, np.float64), (self.interface.n_panels,)))
        L.check(L.lib().dbie_op_set_eps(self._h, L.dptr(ei), L.dptr(eo)))

    def corrections(self, kernel="DT"):
        """Correction blocks of laplace3d_DT_corrections (kernel "DT") or laplace3d_D_corrections ("D"),
        src/kernel/laplace3d_near_upsampling.jl:272-282: (src panels, trg panels, blocks[b, t, c]) with
        C[trg*bd+t, src*bd+c] = blocks[b,t,c]."""
        nb, bd = self.n_blocks, self.block_dim
        src = np.zeros(nb, np.int32); trg = np.zeros(nb, np.int32); blk = np.zeros(max(nb, 1) * bd * bd)
        kid = {"DT": L.OP_DT, "D": L.OP_D}[kernel]
        L.check(L.lib().dbie_op_corrections_kernel(self._h, kid, L.iptr(src), L.iptr(trg), L.dptr(blk)))
        return src, trg, blk[: nb * bd * bd].reshape(nb, bd, bd)

    def device_bytes(self) -> int:
        b = C.c_int64()
        L.check(L.lib().dbie_op_device_bytes(self._h, C.byref(b)))
        return b.value

    def setup_ms(self) -> dict:
        """Host time of the operator build: near pairs, correction blocks, derived data, uploads (ms)."""
        t = np.zeros(4)
        L.check(L.lib().dbie_op_setup_ms(self._h, L.dptr(t)))
        return dict(zip(("pairs", "corrections", "derived", "upload"), map(float, t)))

    def tuned_shape(self, K, precision="fp64"):
        """The autotuner's cached far-field launch shape for K columns: (splits, variant), or None."""
        sp = C.c_int(); var = C.c_int()
        L.check(L.lib().dbie_autotune_choice(self._h, int(K), L.PRECISIONS[precision] if isinstance(precision, str)
                                             else int(precision), C.byref(sp), C.byref(var)))
        return (sp.value, var.value) if sp.value > 0 else None

    def set_tuned_shape(self, K, precision, splits, variant=-1):
        """Pin the far-field launch shape for this operator's class at K columns."""
        L.check(L.lib().dbie_autotune_set(self._h, int(K), L.PRECISIONS[precision] if isinstance(precision, str)
                                          else int(precision), int(splits), int(variant)))

    def graphs(self, enable=None):
        """Switch HIP-graph replay of repeated applies on/off (None: unchanged); returns the
        operator's (captures, replays) so far."""
        c = C.c_int64(); r = C.c_int64()
        L.check(L.lib().dbie_op_graphs(self._h, -1 if enable is None else int(bool(enable)), C.byref(c), C.byref(r)))
        return c.value, r.value

    def stored_far(self, mode=None):
        """Allow (True) or forbid (False) the stored split-FP16 kernel matrix for applies of more than
        128 columns (None: unchanged); returns the device bytes it holds now (0 until built)."""
        b = C.c_int64()
        L.check(L.lib().dbie_op_stored_far(self._h, -1 if mode is None else int(bool(mode)), C.byref(b)))
        return b.value

    def profile(self, enable=True):
        """Record HIP events around the far-field and combine launches (kernel timing)."""
        L.check(L.lib().dbie_op_profile(self._h, int(bool(enable))))

    def profile_read(self):
        """{kernel name: (launches, total ms)} since the last read; clears the records."""
        out = {}
        for kid, name in ((0, "far_fp64"), (1, "far_lowprec"), (2, "near_combine")):
            n = C.c_int64(); ms = C.c_double()
            L.check(L.lib().dbie_op_profile_read(self._h, kid, C.byref(n), C.byref(ms)))
            out[name] = (n.value, ms.value)
        return out

    def apply(self, X, precision="fp64"):
        X = np.asarray(X, np.float64)
        vec = X.ndim == 1
        Xf = L.fortran(X, self.n)
        K = Xf.shape[1]
        Y = np.zeros((self.n, K), order="F")
        L.check(L.lib().dbie_apply(self._h, K, L.dptr(Xf), self.n, L.dptr(Y), self.n,
                                   L.PRECISIONS[precision] if isinstance(precision, str) else int(precision)))
        return Y[:, 0] if vec else Y

    def apply_kernel(self, X, kernel="DT", corrected=False):
        """Plain laplace3d_DT / _S / _D matvec (no corrections, no diagonal); corrected=True adds the near
        corrections of that kernel (laplace3d_DT_fmm3d_corrected / laplace3d_D_fmm3d_corrected,
        src/kernel/laplace3d_near.jl:395-445)."""
        X = np.asarray(X, np.float64)
        vec = X.ndim == 1
        Xf = L.fortran(X, self.n)
        K = Xf.shape[1]
        Y = np.zeros((self.n, K), order="F")
        kid = {"DT": L.OP_DT, "S": L.OP_S, "D": L.OP_D}[kernel]
        fn = L.lib().dbie_apply_kernel_corrected if corrected else L.lib().dbie_apply_kernel
        L.check(fn(self._h, kid, K, L.dptr(Xf), self.n, L.dptr(Y), self.n))
        return Y[:, 0] if vec else Y

    def __matmul__(self, X):
        return self.apply(X)

    def __mul__(self, X):
        return self.apply(X)

    def mul(self, Y, X):
        """mul!(Y, op, X) (src/solver/multi_rhs.jl:103-136)."""
        Y[...] = self.apply(X).reshape(Y.shape)
        return Y

    def dense(self) -> np.ndarray:
        """lhs_dielectric_box3d: dense Dᵀ + diag(t) assembled on the GPU (N <= 65535)."""
        A = np.zeros((self.n, self.n), order="F")
        L.check(L.lib().dbie_lhs_dense(self._h, L.dptr(A), self.n))
        return A

    def solve(self, F, *, rtol=1e-10, atol=0.0, itmax=500, restart=0, precondition=True, precision="fp64",
              inner_rtol=1e-4, max_refine=10, check_residual=True, raise_on_fail=True, krylov="columns",
              recycle=True, verify=True):
        """GMRES solve of A Σ = F for all columns of F on the GPU; returns (Σ, stats dict).
        ``krylov``: "columns" (one Krylov space per column, shared block matvecs) or "block"
        (one block Krylov space for all columns, as Krylov.block_gmres).  ``recycle`` (mixed
        precision, per-column Krylov): refinements after the first deflate the first inner solve's
        Krylov space instead of starting a new one.  ``verify`` (mixed precision): end only on an
        FP64 residual (True) or also on the refined residual estimate, as the FP64 GMRES ends on its
        Krylov estimate (False; ``check_residual`` then reports the true residual)."""
        F = np.asarray(F, np.float64)
        vec = F.ndim == 1
        Ff = L.fortran(F, self.n)
        K = Ff.shape[1]
        S = np.zeros((self.n, K), order="F")
        o = L.SolveOpts()
        L.lib().dbie_solve_opts_default(C.byref(o))
        o.rtol, o.atol, o.itmax, o.restart = float(rtol), float(atol), int(itmax), int(restart)
        o.precondition = _precond_code(precondition)
        o.precision = L.PRECISIONS[precision] if isinstance(precision, str) else int(precision)
        o.inner_rtol, o.max_refine, o.check_residual = float(inner_rtol), int(max_refine), int(bool(check_residual))
        o.krylov = _krylov_code(krylov)
        o.recycle = int(bool(recycle))
        o.verify = int(bool(verify))
        st = L.SolveStats()
        status = L.lib().dbie_solve_block(self._h, K, L.dptr(Ff), self.n, L.dptr(S), self.n, C.byref(o),
                                          C.byref(st))
        stats = st.as_dict()
        stats["residual_history"] = self.residual_history()
        if status == L.NOT_CONVERGED and not raise_on_fail:
            pass
        else:
            L.check(status)
        return (S[:, 0] if vec else S), stats

    def apply_device(self, dX: "DeviceMatrix", dY: "DeviceMatrix", precision="fp64"):
        """Y = A X with X and Y resident in device memory (stream-ordered; returns once queued)."""
        if dX.shape[0] != self.n or dY.shape[0] != self.n or dX.shape[1] != dY.shape[1]:
            raise ValueError("DimensionMismatch: X and Y must be N x K")
        L.check(L.lib().dbie_apply_device(self._h, dX.shape[1], dX.ptr, self.n, dY.ptr, self.n,
                                          L.PRECISIONS[precision] if isinstance(precision, str) else int(precision)))

    def synchronize(self):
        """Wait for the work queued on the operator's stream (apply_device returns once queued)."""
        L.check(L.lib().dbie_op_synchronize(self._h))

    def solve_device(self, dF: "DeviceMatrix", dSigma: "DeviceMatrix", *, rtol=1e-10, atol=0.0, itmax=500,
                     restart=0, precondition=True, precision="fp64", inner_rtol=1e-4, max_refine=10,
                     check_residual=True, raise_on_fail=True, krylov="columns", recycle=True, verify=True):
        """Solve with F and Σ resident in device memory (dbie_solve_block_device); returns stats."""
        if dF.shape[0] != self.n or dSigma.shape != dF.shape:
            raise ValueError("DimensionMismatch: device blocks must be N x K and equal in shape")
        o = L.SolveOpts()
        L.lib().dbie_solve_opts_default(C.byref(o))
        o.rtol, o.atol, o.itmax, o.restart = float(rtol), float(atol), int(itmax), int(restart)
        o.precondition = _precond_code(precondition)
        o.precision = L.PRECISIONS[precision] if isinstance(precision, str) else int(precision)
        o.inner_rtol, o.max_refine, o.check_residual = float(inner_rtol), int(max_refine), int(bool(check_residual))
        o.krylov = _krylov_code(krylov)
        o.recycle = int(bool(recycle))
        o.verify = int(bool(verify))
        st = L.SolveStats()
        status = L.lib().dbie_solve_block_device(self._h, dF.shape[1], dF.ptr, self.n, dSigma.ptr, self.n,
                                                 C.byref(o), C.byref(st))
        if not (status == L.NOT_CONVERGED and not raise_on_fail):
            L.check(status)
        return dict(st.as_dict(), residual_history=self.residual_history())

    def residual_history(self):
        """[(kind, max_k ||r_k|| / ||F_k||)] of the last solve: kind "krylov" (an iteration's residual
        estimate), "fp64" (a true FP64 residual), "updated" (low-precision update / inner estimate)."""
        n = C.c_int()
        L.check(L.lib().dbie_op_residual_history(self._h, 0, C.byref(n), None))
        buf = np.zeros(2 * max(n.value, 1))
        L.check(L.lib().dbie_op_residual_history(self._h, n.value, C.byref(n), L.dptr(buf)))
        kinds = ("krylov", "fp64", "updated")
        return [(kinds[int(buf[2 * i])], float(buf[2 * i + 1])) for i in range(n.value)]

    def pottrg(self, targets, Sigma):
        """Induced potential u[σ](t) = Σ_i w_i σ_i /(4π|t - x_i|) at targets (3 x M or M x 3)."""
        t = np.asarray(targets, np.float64)
        if t.ndim == 2 and t.shape[0] == 3 and t.shape[1] != 3:
            t = t.T
        t = np.ascontiguousarray(t.reshape(-1, 3))
        M = t.shape[0]
        Sg = np.asarray(Sigma, np.float64)
        vec = Sg.ndim == 1
        Sf = L.fortran(Sg, self.n)
        K = Sf.shape[1]
        U = np.zeros((M, K), order="F")
        L.check(L.lib().dbie_pottrg(self._h, M, L.dptr(t.ravel()), K, L.dptr(Sf), self.n, L.dptr(U), M))
        return U[:, 0] if vec else U


def solve_fused(ops, Fs, *, rtol=1e-10, atol=0.0, itmax=500, restart=0, precision="fp64", inner_rtol=1e-5,
                max_refine=10, check_residual=True, raise_on_fail=True, krylov="columns", recycle=True,
                verify=True, precondition=True):
    """Solve the block systems A_b Σ_b = F_b of several operators (different geometries) as one Krylov
    process (dbie_solve_fused): every iteration applies all operators at once on the GPU.  Returns
    ([Σ_b], stats)."""
    B = len(ops)
    if B != len(Fs) or B < 1:
        raise ValueError("one right-hand side block per operator")
    Ff = [L.fortran(F, op.n) for op, F in zip(ops, Fs)]
    Ss = [np.zeros(F.shape, order="F") for F in Ff]
    ha = (C.c_void_p * B)(*[op._h.value for op in ops])
    Ka = np.array([F.shape[1] for F in Ff], np.int32)
    Fa = (L._dp * B)(*[L.dptr(F) for F in Ff])
    Sa = (L._dp * B)(*[L.dptr(S) for S in Ss])
    ld = np.array([op.n for op in ops], np.int64)
    o = L.SolveOpts()
    L.lib().dbie_solve_opts_default(C.byref(o))
    o.rtol, o.atol, o.itmax, o.restart = float(rtol), float(atol), int(itmax), int(restart)
    o.precision = L.PRECISIONS[precision] if isinstance(precision, str) else int(precision)
    o.inner_rtol, o.max_refine, o.check_residual = float(inner_rtol), int(max_refine), int(bool(check_residual))
    o.krylov = _krylov_code(krylov)
    o.recycle = int(bool(recycle))
    o.verify = int(bool(verify))
    o.precondition = _precond_code(precondition)  # 2 / 3: each member's near-field preconditioner
    st = L.SolveStats()
    status = L.lib().dbie_solve_fused(B, ha, L.iptr(Ka), Fa, L.i64ptr(ld), Sa, L.i64ptr(ld), C.byref(o), C.byref(st))
    if not (status == L.NOT_CONVERGED and not raise_on_fail):
        L.check(status)
    return Ss, st.as_dict()


def _precond_code(precondition) -> int:
    """True / 1 / "diagonal": diagonal scaling (the reference's); 2 / "near": the near-field Neumann term;
    3 / "near2": two Jacobi sweeps of the near field (two Neumann terms); False / 0 / "none": no
    preconditioner."""
    if isinstance(precondition, str):
        return {"none": 0, "diagonal": 1, "near": 2, "near2": 3}[precondition]
    return int(precondition)


def _krylov_code(krylov) -> int:
    codes = {"columns": 0, "block": 1}
    if krylov not in codes:
        raise ValueError('krylov must be "columns" or "block"')
    return codes[krylov]


class DeviceMatrix:
    """An N x K column-major FP64 block in device memory (HBM), owned by this object."""

    def __init__(self, rows: int, cols: int):
        self.shape = (int(rows), int(cols))
        self.nbytes = 8 * self.shape[0] * self.shape[1]
        p = C.c_void_p()
        L.check(L.lib().dbie_device_alloc(self.nbytes, C.byref(p)))
        self.ptr = p

    @classmethod
    def from_host(cls, A):
        A = L.fortran(A)
        d = cls(*A.shape)
        L.check(L.lib().dbie_memcpy_h2d(d.ptr, A.ctypes.data_as(C.c_void_p), d.nbytes))
        return d

    def copy_from(self, other: "DeviceMatrix"):
        """Device-to-device copy (a copy kernel) of an equally shaped block."""
        if other.shape != self.shape:
            raise ValueError("DimensionMismatch: shapes differ")
        L.check(L.lib().dbie_copy_block_device(self.shape[0], self.shape[1], other.ptr, self.shape[0], self.ptr,
                                               self.shape[0]))

    def to_host(self) -> np.ndarray:
        out = np.zeros(self.shape, order="F")
        L.check(L.lib().dbie_memcpy_d2h(out.ctypes.data_as(C.c_void_p), self.ptr, self.nbytes))
        return out

    def __del__(self):
        try:
            if getattr(self, "ptr", None) and L._lib is not None:
                L._lib.dbie_device_free(self.ptr)
                self.ptr = None
        except Exception:
            pass


def autotune(enable=None, clear=False):
    """Switch the far-field launch-shape autotuner on/off (None: unchanged); clear=True forgets
    every cached choice.  The tuner times candidate source split counts (and, for "fp16x3" at
    K > 128, the kernel variant) on the first apply of each shape class."""
    L.check(L.lib().dbie_autotune(-1 if enable is None else int(bool(enable)), int(bool(clear))))


def _op_for(interface, key, **kw):
    op = interface._ops.get(key)
    if op is None:
        op = BatchedDielectricOperator(interface, **kw)
        interface._ops[key] = op
    return op


def lhs_dielectric_box3d(interface: DielectricInterface) -> np.ndarray:
    """lhs_dielectric_box3d (src/solver/dielectric_box3d.jl:1-17): dense Dᵀ + diag(t)."""
    return _op_for(interface, ("plain",), corrected=False).dense()


def lhs_dielectric_box3d_fmm3d(interface: DielectricInterface, tol: float = 1e-6):
    """lhs_dielectric_box3d_fmm3d (src/solver/dielectric_box3d.jl:20-44): matrix-free Dᵀ + diag(t).
    ``tol`` is accepted for API compatibility; the GPU sum is exact."""
    return _op_for(interface, ("plain",), corrected=False)


def lhs_dielectric_box3d_fmm3d_corrected(interface: DielectricInterface, fmm_tol: float, up_tol: float,
                                         max_order: int, *, correct_edges=True, adaptive_atol=None,
                                         adaptive_rtol=SQRT_EPS, adaptive_n_GL=0, adaptive_max_depth=20):
    """lhs_dielectric_box3d_fmm3d_corrected (src/solver/dielectric_box3d.jl:48-92)."""
    kw = dict(up_tol=up_tol, max_order=max_order, correct_edges=correct_edges,
              adaptive_atol=up_tol if adaptive_atol is None else adaptive_atol, adaptive_rtol=adaptive_rtol,
              adaptive_n_GL=adaptive_n_GL, adaptive_max_depth=adaptive_max_depth)
    return _op_for(interface, ("corr",) + tuple(sorted(kw.items())), **kw)


def batched_lhs_dielectric_box3d_fmm3d_corrected(interface: DielectricInterface, fmm_tol: float, up_tol: float,
                                                 max_order: int, **kw):
    """batched_lhs_dielectric_box3d_fmm3d_corrected (src/solver/multi_rhs.jl:60-100)."""
    return lhs_dielectric_box3d_fmm3d_corrected(interface, fmm_tol, up_tol, max_order, **kw)


class _KernelMap:
    """LinearMap{Float64} of a corrected layer potential (K + C_K, no diagonal) on a resident operator."""

    def __init__(self, op: BatchedDielectricOperator, kernel: str):
        self.op, self.kernel = op, kernel
        self.shape = op.shape

    def __matmul__(self, X):
        return self.op.apply_kernel(X, self.kernel, corrected=True)

    def __mul__(self, X):
        return self @ X


def laplace3d_DT_fmm3d_corrected(interface: DielectricInterface, fmm_tol: float, up_tol: float, max_order: int, *,
                                 correct_edges=True, adaptive_atol=None, adaptive_rtol=SQRT_EPS, adaptive_n_GL=0,
                                 adaptive_max_depth=20):
    """laplace3d_DT_fmm3d_corrected (src/kernel/laplace3d_near.jl:395-419): x -> Dt x + C x (the far
    field is the exact direct sum; fmm_tol is accepted for the reference's signature)."""
    op = BatchedDielectricOperator(interface, up_tol=up_tol, max_order=max_order, correct_edges=correct_edges,
                                   adaptive_atol=adaptive_atol, adaptive_rtol=adaptive_rtol,
                                   adaptive_n_GL=adaptive_n_GL, adaptive_max_depth=adaptive_max_depth)
    return _KernelMap(op, "DT")


def laplace3d_D_fmm3d_corrected(interface: DielectricInterface, fmm_tol: float, up_tol: float, max_order: int, *,
                                correct_edges=True, adaptive_atol=None, adaptive_rtol=SQRT_EPS, adaptive_n_GL=0,
                                adaptive_max_depth=20):
    """laplace3d_D_fmm3d_corrected (src/kernel/laplace3d_near.jl:421-445): x -> D x + C_D x with the
    double-layer corrections laplace3d_D_corrections (source normal,
    src/kernel/laplace3d_near_upsampling.jl:278-282) on the same near-pair list."""
    op = BatchedDielectricOperator(interface, up_tol=up_tol, max_order=max_order, correct_edges=correct_edges,
                                   adaptive_atol=adaptive_atol, adaptive_rtol=adaptive_rtol,
                                   adaptive_n_GL=adaptive_n_GL, adaptive_max_depth=adaptive_max_depth)
    return _KernelMap(op, "D")


def dielectric_diagonal_scaling(interface: DielectricInterface) -> np.ndarray:
    """dielectric_diagonal_scaling (src/solver/multi_rhs.jl:234-247): 1 / t per node."""
    return 1.0 / interface.diag()


def _rhs_op(interface):
    ops = interface._ops
    if ops:
        return next(iter(ops.values()))
    return _op_for(interface, ("plain",), corrected=False)


def screened_volume_source(*args, tol=None) -> VolumeSource:
    """screened_volume_source (src/core/sources.jl:368-515), computed on the GPU.  Forms:

    * ``(interface, vs, mode)``: the interface's bounding box and uniform permittivities;
    * ``(boxes, epses, eps_out, vs, mode)``: several boxes (first box holding a point wins for
      SharpScreening; partition-of-unity mixing for the soft modes);
    * ``(Lx, Ly, Lz, vs, eps_in, eps_out, mode)``: one box centred at the origin.

    ``tol`` (SharpScreening's widening of the boxes) defaults to sqrt(eps) times the largest side."""
    t = -1.0 if tol is None else float(tol)
    if len(args) in (2, 3) and isinstance(args[0], DielectricInterface):
        interface, vs = args[0], args[1]
        code, bw = _screen_mode(args[2] if len(args) == 3 else None)
        op = _rhs_op(interface)
        rho = np.zeros(vs.density.size)
        L.check(L.lib().dbie_screen_interface(op._h, vs.density.size, L.dptr(np.ascontiguousarray(vs.positions.T)),
                                              L.dptr(vs.density), code, float(bw), t, L.dptr(rho)))
    elif len(args) == 5:
        boxes, epses, eps_out, vs, mode = args
        B = box_array(boxes)
        e = np.ascontiguousarray(epses, np.float64).ravel()
        if len(e) != B.shape[0]:
            raise ValueError("Number of boxes must match number of permittivities")
        rho = _screen_boxes(vs, B, e, float(eps_out), mode, t, 0)
    elif len(args) == 7:
        Lx, Ly, Lz, vs, eps_in, eps_out, mode = args
        B = np.array([[0.0, 0.0, 0.0, float(Lx), float(Ly), float(Lz)]])
        rho = _screen_boxes(vs, B, np.array([float(eps_in)]), float(eps_out), mode, t, 1)
    else:
        raise TypeError("screened_volume_source(interface, vs, mode), (boxes, epses, eps_out, vs, mode) or "
                        "(Lx, Ly, Lz, vs, eps_in, eps_out, mode)")
    return VolumeSource(vs.positions.copy(), vs.weights.copy(), rho)


def _screen_boxes(vs, B, e, eps_out, mode, tol, form):
    code, bw = _screen_mode(mode)
    S = vs.density.size
    rho = np.zeros(S)
    pos = np.ascontiguousarray(vs.positions.T) if S else np.zeros((1, 3))
    dens = vs.density if S else np.zeros(1)
    L.check(L.lib().dbie_screen_boxes(S, L.dptr(pos.ravel()), L.dptr(dens), B.shape[0], L.dptr(B.ravel()), L.dptr(e),
                                      eps_out, code, float(bw), tol, form, L.dptr(rho) if S else L.dptr(np.zeros(1))))
    return rho


def rhs_dielectric_box3d(interface: DielectricInterface, source, eps_src=None, *rest) -> np.ndarray:
    """rhs_dielectric_box3d for a PointSource (src/solver/dielectric_box3d.jl:95-104), a VolumeSource
    with eps_src (:106-120), a VolumeSource screened by the interface (:122-124), or
    ``(interface, boxes, epses, eps_out, vs)``: screened box by box (:224-232)."""
    if rest:
        if len(rest) != 2:
            raise TypeError("rhs_dielectric_box3d(interface, boxes, epses, eps_out, vs)")
        return rhs_volume_boxes(interface, [rest[1]], source, eps_src, rest[0])[:, 0]
    op = _rhs_op(interface)
    N = interface.n_points
    if isinstance(source, PointSource):
        if eps_src is None:
            raise TypeError("eps_src is required for a PointSource")
        src = np.ascontiguousarray(source.point, np.float64).reshape(3)
        q = np.array([float(source.charge)])
        F = np.zeros((N, 1), order="F")
        L.check(L.lib().dbie_rhs_points(op._h, 1, L.dptr(src), L.dptr(q), float(eps_src), L.dptr(F), N))
        return F[:, 0]
    if isinstance(source, VolumeSource):
        return _rhs_volume(op, [source], screening=eps_src is None, eps_src=1.0 if eps_src is None else eps_src)[:, 0]
    raise TypeError(f"unsupported source type {type(source).__name__}")


def rhs_point_charges(interface: DielectricInterface, points, charges, eps_src: float) -> np.ndarray:
    """N x K RHS with one column per point charge (the PointSource rule applied column-wise)."""
    op = _rhs_op(interface)
    pts = np.ascontiguousarray(points, np.float64).reshape(-1, 3)
    q = np.ascontiguousarray(charges, np.float64).ravel()
    K = pts.shape[0]
    if len(q) != K:
        raise ValueError("points and charges must have the same length")
    F = np.zeros((interface.n_points, K), order="F")
    L.check(L.lib().dbie_rhs_points(op._h, K, L.dptr(pts.ravel()), L.dptr(q), float(eps_src), L.dptr(F),
                                    interface.n_points))
    return F


def rhs_volume_boxes(interface: DielectricInterface, vss, boxes, epses, eps_out=1.0, mode=None) -> np.ndarray:
    """N x K right-hand side of K volume sources screened box by box (default SharpScreening), eps_src = 1:
    rhs_dielectric_box3d_fmm3d(interface, vss, thresh; screen_boxes, screen_epses, screen_eps_out)
    (src/solver/multi_rhs.jl:168-211) with the direct sum for the FMM."""
    code, bw = _screen_mode(SharpScreening() if mode is None else mode)
    B = box_array(boxes)
    e = np.ascontiguousarray(epses, np.float64).ravel()
    if len(e) != B.shape[0]:
        raise ValueError("Number of boxes must match number of permittivities")
    op = _rhs_op(interface)
    col_ptr, xyz, w, rho = _pack_sources(vss)
    K, N = len(vss), op.n
    if K == 0:
        return np.zeros((N, 0))
    F = np.zeros((N, K), order="F")
    L.check(L.lib().dbie_rhs_volume_boxes(op._h, K, L.i64ptr(col_ptr), L.dptr(xyz.ravel()), L.dptr(w), L.dptr(rho),
                                          B.shape[0], L.dptr(B.ravel()), L.dptr(e), float(eps_out), code, float(bw),
                                          L.dptr(F), N))
    return F


def _pack_sources(vss):
    K = len(vss)
    col_ptr = np.zeros(K + 1, np.int64)
    col_ptr[1:] = np.cumsum([v.positions.shape[1] for v in vss])
    xyz = np.ascontiguousarray(np.concatenate([v.positions.T for v in vss], axis=0)) if K else np.zeros((0, 3))
    w = np.ascontiguousarray(np.concatenate([v.weights for v in vss])) if K else np.zeros(0)
    rho = np.ascontiguousarray(np.concatenate([v.density for v in vss])) if K else np.zeros(0)
    if xyz.size == 0:
        xyz = np.zeros((1, 3)); w = np.zeros(1); rho = np.zeros(1)
    return col_ptr, xyz, w, rho


def _rhs_volume(op, vss, screening: bool, eps_src: float = 1.0) -> np.ndarray:
    K, N = len(vss), op.n
    if K == 0:
        return np.zeros((N, 0))
    col_ptr, xyz, w, rho = _pack_sources(vss)
    F = np.zeros((N, K), order="F")
    L.check(L.lib().dbie_rhs_volume(op._h, K, L.i64ptr(col_ptr), L.dptr(xyz.ravel()), L.dptr(w), L.dptr(rho),
                                    int(bool(screening)), float(eps_src), L.dptr(F), N))
    return F


def rhs_dielectric_box3d_fmm3d(interface: DielectricInterface, sources, *args, screen_boxes=None,
                               screen_epses=None, screen_eps_out=1.0):
    """rhs_dielectric_box3d_fmm3d: single VolumeSource (src/solver/dielectric_box3d.jl:127-159) or
    Vector{VolumeSource} -> N x K (src/solver/multi_rhs.jl:168-211).  Forms:
    (interface, vs, eps_src, thresh), (interface, vs, thresh), (interface, vss, thresh), the latter with
    optional box-by-box screening (screen_boxes, screen_epses, screen_eps_out), and
    (interface, boxes, epses, eps_out, vs, thresh) (src/solver/dielectric_box3d.jl:234-243)."""
    if len(args) == 4 and isinstance(args[2], VolumeSource):
        return rhs_volume_boxes(interface, [args[2]], sources, args[0], args[1])[:, 0]
    if screen_boxes is not None:
        vss = [sources] if isinstance(sources, VolumeSource) else list(sources)
        return rhs_volume_boxes(interface, vss, screen_boxes, screen_epses, screen_eps_out)
    op = _rhs_op(interface)
    if isinstance(sources, VolumeSource):
        if len(args) == 2:
            return _rhs_volume(op, [sources], screening=False, eps_src=float(args[0]))[:, 0]
        return _rhs_volume(op, [sources], screening=True)[:, 0]
    return _rhs_volume(op, list(sources), screening=True)


# --------------------------------------------------------------------------- #
# Solvers
# --------------------------------------------------------------------------- #
def solve_lu(A, b):
    """solve_lu (src/utils/linear_algebra.jl:1): LU with partial pivoting, on the GPU (rocSOLVER)."""
    A = L.fortran(A)
    n = A.shape[0]
    if A.shape[1] != n:
        raise ValueError("DimensionMismatch: A must be square")
    b = np.asarray(b, np.float64)
    vec = b.ndim == 1
    B = L.fortran(b, n).copy(order="F")
    L.check(L.lib().dbie_solve_lu(n, L.dptr(A), n, B.shape[1], L.dptr(B), n))
    return B[:, 0] if vec else B


def solve_gmres(A, b, atol: float = 1e-6, rtol: float = 1e-6, itmax: int = 1000):
    """solve_gmres (src/utils/linear_algebra.jl:3-6): unpreconditioned GMRES on the GPU."""
    if not isinstance(A, BatchedDielectricOperator):
        raise TypeError("solve_gmres needs a GPU operator (lhs_dielectric_box3d_fmm3d[_corrected])")
    x, _ = A.solve(b, rtol=rtol, atol=atol, itmax=itmax, precondition=False, raise_on_fail=False)
    return x


def solve_dielectric_box3d_block(interface: DielectricInterface, vss, *, fmm_tol=1e-9, up_tol=1e-9, max_order=8,
                                 rtol=1e-10, atol=0.0, itmax=500, precondition=True, correct_edges=True,
                                 screen_boxes=None, screen_epses=None, screen_eps_out=1.0,
                                 precision="fp64", inner_rtol=1e-4, max_refine=10):
    """solve_dielectric_box3d_block (src/solver/multi_rhs.jl:270-286) -> (Σ, stats).
    With ``screen_boxes`` the sources are screened box by box (several permittivity regions);
    ``precision`` other than "fp64" selects the mixed-precision refinement solve."""
    op = batched_lhs_dielectric_box3d_fmm3d_corrected(interface, fmm_tol, up_tol, max_order,
                                                      correct_edges=correct_edges)
    F = rhs_dielectric_box3d_fmm3d(interface, list(vss), fmm_tol, screen_boxes=screen_boxes,
                                   screen_epses=screen_epses, screen_eps_out=screen_eps_out)
    return op.solve(F, rtol=rtol, atol=atol, itmax=itmax, precondition=precondition, precision=precision,
                    inner_rtol=inner_rtol, max_refine=max_refine)


# ---------------------------------------------------------------------------------------------
# Off-surface targets, Gaussian charge blobs and the reaction matrix
# ---------------------------------------------------------------------------------------------
def gausshermite(n: int):
    """Physicists' Gauss-Hermite rule (weight exp(-x^2)), ascending nodes."""
    x = np.zeros(n); w = np.zeros(n)
    L.check(L.lib().dbie_gauss_hermite(int(n), L.dptr(x), L.dptr(w)))
    return x, w


def GaussianVolumeSource(center, sigma: float, n: int, tol: float) -> VolumeSource:
    """GaussianVolumeSource (src/core/sources.jl:517-548): unit-charge Gaussian of width sigma on an
    n^3 midpoint grid spanning +-sqrt(2 sigma^2 log(1/tol)) about the centre."""
    if n < 1 or not sigma > 0 or not (0 < tol < 1):
        raise ValueError("need n >= 1, sigma > 0 and tol in (0, 1)")
    two_s2 = 2.0 * sigma * sigma
    support = math.sqrt(two_s2 * math.log(1.0 / tol))
    norm = 1.0 / (math.sqrt(2.0 * math.pi) * sigma) ** 3
    h = 2.0 * support / n
    off0 = -support + h / 2.0
    axes = [np.array([c + off0 + i * h for i in range(n)]) for c in center]
    X, Y, Z = np.meshgrid(*axes, indexing="ij")
    r2 = (X - center[0]) ** 2 + (Y - center[1]) ** 2 + (Z - center[2]) ** 2
    # point order ix outer, iz inner (_volume_source_positions, src/core/sources.jl:129-151)
    pts = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
    w = np.full(n ** 3, h * h * h)
    rho = (norm * np.exp(-r2 / two_s2)).ravel()
    return VolumeSource.from_points(pts, w, rho)


def gaussian_blob_gh(center, sigma: float, n: int = 6) -> VolumeSource:
    """Unit-charge isotropic Gaussian blob exp(-|y-c|^2 / (2 sigma^2)) / (2 pi sigma^2)^(3/2) as an
    n x n x n tensor Gauss-Hermite cloud: nodes c + sqrt(2) sigma (x_i, x_j, x_k), charges
    q = pi^(-3/2) w_i w_j w_k normalised to sum 1, stored as weights q / rho(y) and density rho(y)
    (so that sum(weights * density) = 1).  Point order: i (x) outer, then j, then k."""
    x, w = gausshermite(n)
    s = math.sqrt(2.0) * sigma
    I, J, K = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    I, J, K = I.ravel(), J.ravel(), K.ravel()
    pts = np.stack([center[0] + s * x[I], center[1] + s * x[J], center[2] + s * x[K]], axis=1)
    q = w[I] * w[J] * w[K]
    q = q / q.sum()
    r2 = (s * x[I]) ** 2 + (s * x[J]) ** 2 + (s * x[K]) ** 2
    rho = np.exp(-r2 / (2.0 * sigma * sigma)) / (2.0 * math.pi * sigma * sigma) ** 1.5
    return VolumeSource.from_points(pts, q / rho, rho)


def blob_lattice_centers(Lx, Ly, Lz, nx=16, ny=8, nz=8):
    """Centres of an nx x ny x nz lattice interior to the box: x_c = -Lx/2 + (i - 1/2) Lx/nx, ...;
    order x outer, then y, then z."""
    xs = [-Lx / 2 + (i + 0.5) * Lx / nx for i in range(nx)]
    ys = [-Ly / 2 + (j + 0.5) * Ly / ny for j in range(ny)]
    zs = [-Lz / 2 + (k + 0.5) * Lz / nz for k in range(nz)]
    return np.array([(x, y, z) for x in xs for y in ys for z in zs])


class TargetPotentialMap:
    """The M x N map sigma -> u[sigma](targets) of laplace3d_pottrg_fmm3d_corrected_hcubature
    (src/kernel/laplace3d_near_hcubature.jl:337-375): direct sum over the interface refined around
    the targets plus the adaptive-cubature near-target corrections.  device_cubature=True integrates
    the near-target cubatures on the GPU (bit-identical to the host cubature) instead of on n_threads
    host threads; device_refine=True (with it) also builds the refinement and the prolongation on the
    GPU (bit-identical to the host build), keeping the prolongation there.  Evaluation runs on the GPU."""

    def __init__(self, interface: DielectricInterface, targets, hcubature_atol: float, range_factor: float = 5.0,
                 n_threads: int = 0, device_cubature: bool = True, device_refine: bool = True):
        t = np.asarray(targets, np.float64)
        if t.ndim == 2 and t.shape[0] == 3 and t.shape[1] != 3:
            t = t.T
        t = np.ascontiguousarray(t.reshape(-1, 3))
        self.M, self.n = t.shape[0], interface.n_points
        h = C.c_void_p()
        L.check(L.lib().dbie_tmap_create_ex(interface._h, self.M, L.dptr(t.ravel()), float(hcubature_atol),
                                            float(range_factor), int(n_threads),
                                            (2 if device_refine else 1) if device_cubature else 0,
                                            C.byref(h)))
        self._h = h
        self._iface = interface
        v = [C.c_int64() for _ in range(6)]
        L.check(L.lib().dbie_tmap_sizes(self._h, *[C.byref(x) for x in v]))
        _, _, self.n_refined, self.n_pairs, self.nnz, self.n_capped = [x.value for x in v]

    def __del__(self):
        try:
            if getattr(self, "_h", None) and L._lib is not None:
                L._lib.dbie_tmap_destroy(self._h)
                self._h = None
        except Exception:
            pass

    @property
    def shape(self):
        return (self.M, self.n)

    def setup_ms(self):
        """Build phases in ms: refinement, prolongation, pairs + cubatures, assembly."""
        t = np.zeros(4)
        L.check(L.lib().dbie_tmap_setup_ms(self._h, L.dptr(t)))
        return dict(zip(("refine", "prolong", "cubatures", "assembly"), t.tolist()))

    def refined(self):
        """(nodes 4 x N_refined: x, y, z, w; parent panel and from_split per refined panel)."""
        npp = self._iface.n_quad ** 2
        Pr = self.n_refined // npp
        xyzw = np.zeros(4 * self.n_refined); par = np.zeros(Pr, np.int32); spl = np.zeros(Pr, np.int32)
        L.check(L.lib().dbie_tmap_refined(self._h, L.dptr(xyzw), L.iptr(par), L.iptr(spl)))
        return xyzw.reshape(4, -1), par, spl.astype(bool)

    def matrices(self):
        """(prolongation, corrections) as CSR triples (ptr, col, val)."""
        Nr, M = self.n_refined, self.M
        pp = np.zeros(Nr + 1, np.int64)
        L.check(L.lib().dbie_tmap_matrices(self._h, L.i64ptr(pp), None, None, None, None, None))
        npn = int(pp[-1])
        pc = np.zeros(max(npn, 1), np.int32); pv = np.zeros(max(npn, 1))
        cp = np.zeros(M + 1, np.int64); cc = np.zeros(max(self.nnz, 1), np.int32); cv = np.zeros(max(self.nnz, 1))
        L.check(L.lib().dbie_tmap_matrices(self._h, L.i64ptr(pp), L.iptr(pc), L.dptr(pv), L.i64ptr(cp), L.iptr(cc),
                                           L.dptr(cv)))
        return (pp, pc[:npn], pv[:npn]), (cp, cc[:self.nnz], cv[:self.nnz])

    def apply(self, Sigma):
        Sg = np.asarray(Sigma, np.float64)
        vec = Sg.ndim == 1
        Sf = L.fortran(Sg, self.n)
        K = Sf.shape[1]
        U = np.zeros((self.M, K), order="F")
        L.check(L.lib().dbie_tmap_apply(self._h, K, L.dptr(Sf), self.n, L.dptr(U), max(self.M, 1)))
        return U[:, 0] if vec else U

    def __matmul__(self, Sigma):
        return self.apply(Sigma)

    def reaction(self, grp_ptr, q, Sigma):
        """R[a,b] = sum_{s in group a} q_s u[Sigma_b](t_s) (G x K); Sigma on the host or a DeviceMatrix."""
        grp = np.ascontiguousarray(grp_ptr, dtype=np.int64)
        G = len(grp) - 1
        qv = np.ascontiguousarray(q, dtype=np.float64)
        if qv.shape != (self.M,):
            raise ValueError(f"DimensionMismatch: q must have {self.M} entries")
        if isinstance(Sigma, DeviceMatrix):
            if Sigma.shape[0] != self.n:
                raise ValueError("DimensionMismatch: Sigma rows")
            K = Sigma.shape[1]
            R = np.zeros((G, K), order="F")
            L.check(L.lib().dbie_reaction_matrix_device(self._h, G, L.i64ptr(grp), L.dptr(qv), K, Sigma.ptr, self.n,
                                                        L.dptr(R), G))
            return R
        Sf = L.fortran(Sigma, self.n)
        K = Sf.shape[1]
        R = np.zeros((G, K), order="F")
        L.check(L.lib().dbie_reaction_matrix(self._h, G, L.i64ptr(grp), L.dptr(qv), K, L.dptr(Sf), self.n, L.dptr(R),
                                             G))
        return R


def laplace3d_pottrg_fmm3d_corrected_hcubature(interface: DielectricInterface, targets, fmm_tol: float,
                                               hcubature_atol: float, range_factor: float = 5.0):
    """laplace3d_pottrg_fmm3d_corrected_hcubature (src/kernel/laplace3d_near_hcubature.jl:337-375).
    fmm_tol is accepted and ignored: the base potential is the exact direct sum."""
    return TargetPotentialMap(interface, targets, hcubature_atol, range_factor)


def _sources_targets(sources):
    pts = np.concatenate([np.asarray(vs.positions, np.float64).T for vs in sources])
    q = np.concatenate([np.asarray(vs.weights, np.float64) * np.asarray(vs.density, np.float64) for vs in sources])
    grp = np.zeros(len(sources) + 1, np.int64)
    grp[1:] = np.cumsum([len(vs.weights) for vs in sources])
    return pts, q, grp


def reaction_matrix(interface: DielectricInterface, sources, Sigma, *, lhs_tol: float = 1e-9,
                    range_factor: float = 5.0, tmap: TargetPotentialMap | None = None):
    """Induced term of four_index_matrix (src/solver/multi_rhs.jl:299-321):
    R[a,b] = sum_{s in source a} w_s rho_s u[Sigma_b](y_s), with u from the near-corrected
    target map at tolerance lhs_tol.  Unlike the reference, the sources need not share one grid:
    the targets are the union of all source points.  Returns (R, tmap)."""
    pts, q, grp = _sources_targets(sources)
    if tmap is None:
        tmap = TargetPotentialMap(interface, pts, lhs_tol, range_factor)
    elif tmap.M != len(q):
        raise ValueError("DimensionMismatch: target map does not match the sources")
    return tmap.reaction(grp, q, Sigma), tmap


# ---------------------------------------------------------------------------------------------
# Request server (continuous batching of many block solves)
# ---------------------------------------------------------------------------------------------
class Server:
    """Continuous-batching request loop on the GPU (libdbie's dbie_server_*).

    submit_box() queues a single dielectric box with K point charges (eps_src = eps_in) and
    returns an id at once; wait() blocks until it is done (or a timeout) and returns its info;
    result() returns the N x K density and forgets the request.  Requests are admitted by device
    memory footprint, set up on host threads while others run on the GPU, and ready requests on
    one cached operator with equal options are fused into one block solve."""

    def __init__(self, *, hbm_budget: float = 0.0, setup_workers: int = 2, setup_threads: int = 0,
                 solve_slots: int = 2, max_batch_rhs: int = 256, cache_ops: bool = True, inject_oom_every: int = 0):
        o = L.ServerOpts()
        L.lib().dbie_server_opts_default(C.byref(o))
        o.hbm_budget, o.setup_workers, o.setup_threads = float(hbm_budget), int(setup_workers), int(setup_threads)
        o.solve_slots, o.max_batch_rhs, o.cache_ops = int(solve_slots), int(max_batch_rhs), int(bool(cache_ops))
        o.inject_oom_every = int(inject_oom_every)
        h = C.c_void_p()
        L.check(L.lib().dbie_server_create(C.byref(o), C.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None) and L._lib is not None:
            L._lib.dbie_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def submit_box(self, Lx, Ly, Lz, n_quad, l_ec, eps_in, eps_out, points, charges, *, precision="fp16x3",
                   rtol=1e-10, itmax=500, inner_rtol=1e-5, max_refine=10, verify=True, precondition=True) -> int:
        pts = np.ascontiguousarray(points, np.float64).reshape(-1, 3)
        q = np.ascontiguousarray(charges, np.float64).ravel()
        if len(q) != pts.shape[0]:
            raise ValueError("points and charges must have the same length")
        o = L.SolveOpts()
        L.lib().dbie_solve_opts_default(C.byref(o))
        o.rtol, o.itmax = float(rtol), int(itmax)
        o.precision = L.PRECISIONS[precision] if isinstance(precision, str) else int(precision)
        o.inner_rtol, o.max_refine = float(inner_rtol), int(max_refine)
        o.verify = int(bool(verify))
        o.precondition = _precond_code(precondition)
        rid = C.c_int64()
        L.check(L.lib().dbie_server_submit_box(self._h, Lx, Ly, Lz, int(n_quad), l_ec, eps_in, eps_out, len(q),
                                               L.dptr(pts.ravel()), L.dptr(q), C.byref(o), C.byref(rid)))
        return rid.value

    def wait(self, rid: int, timeout: float | None = None):
        """Info dict of the request; None if it is still running after `timeout` seconds."""
        info = L.RequestInfo()
        st = L.lib().dbie_server_wait(self._h, int(rid), -1.0 if timeout is None else float(timeout), C.byref(info))
        if st == L.PENDING:
            return None
        L.check(st)
        return info.as_dict()

    def info(self, rid: int) -> dict:
        """Info dict of the request in whatever state it is (no waiting)."""
        info = L.RequestInfo()
        st = L.lib().dbie_server_wait(self._h, int(rid), 0.0, C.byref(info))
        if st not in (L.OK, L.PENDING):
            L.check(st)
        return info.as_dict()

    def pause(self, paused: bool = True):
        """Stop (or resume) starting new block solves; setup goes on, ready requests accumulate."""
        L.check(L.lib().dbie_server_pause(self._h, int(bool(paused))))

    def result(self, rid: int) -> np.ndarray:
        """N x K density of a finished request (raises if it failed); the request is forgotten."""
        info = self.wait(rid)
        if info["status"] not in (L.OK,):
            L.lib().dbie_server_take(self._h, int(rid), None, 0)
            raise L.DbieError(info["status"], f"request {rid} failed: {info['message']}")
        S = np.zeros((info["N"], info["K"]), order="F")
        L.check(L.lib().dbie_server_take(self._h, int(rid), L.dptr(S), info["N"]))
        return S

    def forget(self, rid: int):
        L.check(L.lib().dbie_server_take(self._h, int(rid), None, 0))

    def cancel(self, rid: int):
        L.check(L.lib().dbie_server_cancel(self._h, int(rid)))

    def stats(self) -> dict:
        s = L.ServerSummary()
        L.check(L.lib().dbie_server_stats(self._h, C.byref(s)))
        return s.as_dict()


def serving_requests(n: int, seed: int = 0, rate: float = 8.0, l_ecs=(0.1, 0.07, 0.05), kmax: int = 64):
    """Seeded request stream of the serving workload: Poisson arrivals at `rate` per second, box sides
    ~U(0.5, 2.0), l_ec uniform over l_ecs, n_quad 4, eps_in ~U(2, 80), eps_out 1, 1..kmax charges
    uniform in the box at least 0.05 from every face, q ~U(-1, 1).  Returns a list of dicts."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for _ in range(n):
        t += rng.exponential(1.0 / rate)
        L3 = rng.uniform(0.5, 2.0, 3)
        l_ec = float(l_ecs[rng.integers(len(l_ecs))])
        eps_in = float(rng.uniform(2.0, 80.0))
        K = int(rng.integers(1, kmax + 1))
        half = L3 / 2 - 0.05
        pts = rng.uniform(-half, half, (K, 3))
        q = rng.uniform(-1.0, 1.0, K)
        out.append({"t": t, "L": tuple(float(v) for v in L3), "l_ec": l_ec, "eps_in": eps_in, "eps_out": 1.0,
                    "n_quad": 4, "points": pts, "charges": q})
    return out

