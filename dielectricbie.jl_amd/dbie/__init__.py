"""dbie -- MI355X (gfx950) boundary-operator engine with DielectricBIE.jl's API.

Host-side mirror of the reference's operator/solver interface; all arithmetic
runs in libdbie.so (HIP kernels for gfx950 behind a C ABI, include/dbie.h).
"""
from ._lib import DbieError, NotConverged, lib, LIB_PATH  # noqa: F401
from .api import (  # noqa: F401
    BatchedDielectricOperator,
    DeviceMatrix,
    GaussianVolumeSource,
    TargetPotentialMap,
    DielectricInterface,
    PointSource,
    Server,
    SharpScreening,
    SoftMixInversePermittivity,
    SoftMixPermittivity,
    VolumeSource,
    all_weights,
    autotune,
    batched_lhs_dielectric_box3d_fmm3d_corrected,
    best_grid,
    best_grid_mn,
    blob_lattice_centers,
    box_array,
    build_neighbor_list,
    dielectric_diagonal_scaling,
    gaussian_blob_gh,
    gausshermite,
    gausslegendre,
    interface_from_panels,
    lhs_dielectric_box3d,
    multi_dielectric_box3d,
    lhs_dielectric_box3d_fmm3d,
    lhs_dielectric_box3d_fmm3d_corrected,
    laplace3d_D_fmm3d_corrected,
    laplace3d_DT_fmm3d_corrected,
    laplace3d_pottrg_fmm3d_corrected_hcubature,
    num_points,
    rhs_dielectric_box3d,
    rhs_dielectric_box3d_fmm3d,
    reaction_matrix,
    rhs_point_charges,
    rhs_volume_boxes,
    screened_volume_source,
    serving_requests,
    single_dielectric_box3d,
    solve_dielectric_box3d_block,
    solve_fused,
    solve_gmres,
    solve_lu,
)

from .lattice import (  # noqa: F401
    BATCH_FORMAT_VERSION,
    BatchResult,
    LatticeBatch,
    batch_result,
    batch_volume_sources,
    envelope_volume_source,
    is_complete_batch,
    load_batch_result,
    save_batch_result,
    solve_dielectric_lattice_batch,
)

__version__ = "0.1.0"
