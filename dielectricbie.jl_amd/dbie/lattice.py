"""Lattice-batch caller of the block solver and the on-disk batch result.

``solve_dielectric_lattice_batch`` mirrors src/solver/lattice_batch.jl:285-304: one shared
interface for a batch of pair densities, the sources screened box by box, one block solve on the
GPU.  ``BatchResult`` with ``save_batch_result`` / ``load_batch_result`` / ``is_complete_batch``
mirrors src/solver/batch_io.jl:6-68: the same fields, a format-version guard, shape checks on load
and atomic writes (temporary file on the same filesystem, then rename), so that a batch file
either exists complete or not at all.  The container is a NumPy ``.npz`` archive read with
``allow_pickle=False`` (the reference serializes Julia objects, which no other language reads);
the interface is stored as its panels (corners, normals, edge flags, permittivities).
"""
from __future__ import annotations

import json
import os
import secrets
import zipfile
from dataclasses import dataclass, field

import numpy as np

from . import api as A

BATCH_FORMAT_VERSION = 1


@dataclass
class LatticeBatch:
    """LatticeBatch (src/solver/lattice_batch.jl:49-60): K pair densities sharing n grid points."""
    pair_ids: np.ndarray        # K x 2 (orbital i, orbital j) per column
    gidx: np.ndarray            # n x 3 global grid indices
    positions: np.ndarray       # 3 x n
    weights: np.ndarray         # n
    densities: np.ndarray       # n x K raw (unscreened) pair densities
    lattice_basis: object = None

    def __post_init__(self):
        self.pair_ids = np.asarray(self.pair_ids, np.int64).reshape(-1, 2)
        self.gidx = np.asarray(self.gidx, np.int64).reshape(-1, 3)
        self.positions = np.ascontiguousarray(self.positions, np.float64).reshape(3, -1)
        self.weights = np.ascontiguousarray(self.weights, np.float64).ravel()
        self.densities = np.asarray(self.densities, np.float64).reshape(self.positions.shape[1], -1)
        n, K = self.densities.shape
        if len(self.weights) != n or len(self.gidx) != n or len(self.pair_ids) != K:
            raise ValueError("LatticeBatch: gidx, positions, weights, densities and pair_ids disagree")


def num_pairs(b: LatticeBatch) -> int:
    return b.densities.shape[1]


def envelope_volume_source(b: LatticeBatch) -> A.VolumeSource:
    """envelope_volume_source (src/solver/lattice_batch.jl:154-171): per-point root sum of squares
    over the batch's columns (accumulated column by column, then the square root)."""
    acc = np.zeros(b.densities.shape[0])
    for k in range(b.densities.shape[1]):
        acc += b.densities[:, k] ** 2
    return A.VolumeSource(b.positions.copy(), b.weights.copy(), np.sqrt(acc))


def batch_volume_sources(b: LatticeBatch):
    """batch_volume_sources (src/solver/lattice_batch.jl:173-184): one VolumeSource per column."""
    return [A.VolumeSource(b.positions.copy(), b.weights.copy(), b.densities[:, k].copy())
            for k in range(num_pairs(b))]


def solve_dielectric_lattice_batch(boxes, epses, eps_out, b: LatticeBatch, *, n_quad: int, l_ec: float,
                                   fmm_tol: float, gmres_rtol: float, interface=None, up_tol=None,
                                   max_order: int = 8, itmax: int = 500, precondition: bool = True,
                                   correct_edges: bool = True, precision: str = "fp64", inner_rtol: float = 1e-5,
                                   alpha: float = float(np.sqrt(2.0))):
    """solve_dielectric_lattice_batch (src/solver/lattice_batch.jl:285-304) -> dict(sigma, interface,
    sources, stats).

    The reference refines the shared interface on the batch envelope with
    multi_dielectric_box3d_rhs_adaptive, whose near-target RHS values come from the TKM volume
    solver; that refinement is not part of this library.  Pass the refined ``interface`` (for
    instance the reference's own, through ``interface_from_panels``); without one the boxes are
    meshed by ``multi_dielectric_box3d`` (edge/corner refinement to ``l_ec`` only).  The solve
    screens every source box by box and runs one block solve on the GPU (FP64, or mixed precision
    with ``precision``)."""
    if interface is None:
        interface = A.multi_dielectric_box3d(n_quad, l_ec, boxes, epses, eps_out, alpha=alpha)
    sources = batch_volume_sources(b)
    sigma, stats = A.solve_dielectric_box3d_block(
        interface, sources, fmm_tol=fmm_tol, up_tol=fmm_tol if up_tol is None else up_tol, max_order=max_order,
        rtol=gmres_rtol, itmax=itmax, precondition=precondition, correct_edges=correct_edges,
        screen_boxes=boxes, screen_epses=epses, screen_eps_out=eps_out, precision=precision,
        inner_rtol=inner_rtol)
    return {"sigma": sigma, "interface": interface, "sources": sources, "stats": stats}


@dataclass
class BatchResult:
    """BatchResult (src/solver/batch_io.jl:6-29): what the post-evaluation phase needs from one
    solved batch (positions are recomputed from ``gidx``)."""
    version: int
    batch_id: int
    pair_ids: np.ndarray      # K x 2
    gidx: np.ndarray          # n x 3
    weights: np.ndarray       # n
    densities: np.ndarray     # n x K
    interface: object         # DielectricInterface (stored as its panels)
    sigma: np.ndarray         # N x K
    stats: dict = field(default_factory=dict)


def batch_result(batch_id: int, b: LatticeBatch, solved: dict, extra_stats=None) -> BatchResult:
    """Pack a solved batch (solve_batch_core, src/campaign/tasks.jl:181-201)."""
    st = {"niter": int(solved["stats"].get("niter", 0)), "dof": int(solved["sigma"].shape[0]),
          "n_support": int(len(b.gidx)), "K": int(num_pairs(b))}
    st.update(extra_stats or {})
    return BatchResult(BATCH_FORMAT_VERSION, int(batch_id), b.pair_ids.copy(), b.gidx.copy(), b.weights.copy(),
                       b.densities.copy(), solved["interface"], np.asarray(solved["sigma"]), st)


def save_batch_result(path: str, br: BatchResult) -> str:
    """Atomic write (src/solver/batch_io.jl:31-41): a temporary file next to ``path``, then rename."""
    d = os.path.dirname(path)
    if d:
        os.makedirs(d, exist_ok=True)
    itf = br.interface
    corners, normals, is_edge = itf.panels()
    tmp = f"{path}.tmp.{os.getpid()}_{secrets.randbits(32)}"
    try:
        with open(tmp, "wb") as fh:
            np.savez(fh, version=np.int64(br.version), batch_id=np.int64(br.batch_id),
                     pair_ids=np.asarray(br.pair_ids, np.int64), gidx=np.asarray(br.gidx, np.int64),
                     weights=np.asarray(br.weights, np.float64), densities=np.asarray(br.densities, np.float64),
                     sigma=np.asarray(br.sigma, np.float64),
                     stats=np.frombuffer(json.dumps(br.stats, sort_keys=True).encode(), np.uint8),
                     itf_corners=corners, itf_normals=normals, itf_is_edge=is_edge.astype(np.int32),
                     itf_eps_in=np.asarray(itf.eps_in, np.float64), itf_eps_out=np.asarray(itf.eps_out, np.float64),
                     itf_n_quad=np.int64(itf.n_quad))
            fh.flush()
            os.fsync(fh.fileno())
        os.replace(tmp, path)
    finally:
        if os.path.exists(tmp):
            os.remove(tmp)
    return path


def load_batch_result(path: str) -> BatchResult:
    """Read and validate (src/solver/batch_io.jl:43-58): format version and the shapes of every
    field against the densities.  Raises ValueError on any problem."""
    try:
        with np.load(path, allow_pickle=False) as z:
            f = {k: z[k] for k in z.files}
    except (OSError, ValueError, EOFError, zipfile.BadZipFile) as e:
        raise ValueError(f"{path}: not a batch result ({e})") from None
    need = {"version", "batch_id", "pair_ids", "gidx", "weights", "densities", "sigma", "stats", "itf_corners",
            "itf_normals", "itf_is_edge", "itf_eps_in", "itf_eps_out", "itf_n_quad"}
    if not need <= set(f):
        raise ValueError(f"{path}: not a batch result (missing {sorted(need - set(f))})")
    if int(f["version"]) != BATCH_FORMAT_VERSION:
        raise ValueError(f"{path}: format version {int(f['version'])} != {BATCH_FORMAT_VERSION}")
    dens = f["densities"]
    if dens.ndim != 2:
        raise ValueError(f"{path}: densities must be n x K")
    n, K = dens.shape
    if len(f["gidx"]) != n:
        raise ValueError(f"{path}: gidx/densities mismatch")
    if len(f["weights"]) != n:
        raise ValueError(f"{path}: weights/densities mismatch")
    if len(f["pair_ids"]) != K:
        raise ValueError(f"{path}: pair_ids/densities mismatch")
    if f["sigma"].ndim != 2 or f["sigma"].shape[1] != K:
        raise ValueError(f"{path}: sigma/densities mismatch")
    itf = A.interface_from_panels(f["itf_corners"], f["itf_normals"], int(f["itf_n_quad"]), f["itf_eps_in"],
                                  f["itf_eps_out"], f["itf_is_edge"])
    if itf.n_points != f["sigma"].shape[0]:
        raise ValueError(f"{path}: sigma/interface mismatch")
    return BatchResult(int(f["version"]), int(f["batch_id"]), f["pair_ids"], f["gidx"], f["weights"], dens, itf,
                       f["sigma"], json.loads(bytes(f["stats"]).decode()))


def is_complete_batch(path: str) -> bool:
    """True iff ``path`` exists and loads as a valid current-version batch result
    (src/solver/batch_io.jl:60-68)."""
    if not os.path.isfile(path):
        return False
    try:
        load_batch_result(path)
        return True
    except ValueError:
        return False
