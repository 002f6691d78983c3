/*
 * dbie.h -- C ABI of libdbie, an MI355X (gfx950) implementation of
 * DielectricBIE.jl's boundary-operator hot path.
 *
 * Conventions
 *   - All functions return dbie_status (0 = OK).  On error a thread-local
 *     message is available from dbie_last_error().
 *   - Matrices are column-major FP64 with an explicit leading dimension,
 *     exactly like Julia's Matrix{Float64} (the reference's caller type).
 *   - Indices of panels/nodes are 0-based.  Node order is panel order, then
 *     (i outer along b-a, j inner along d-a), matching
 *     src/shape/box3d.jl:44-59 of the reference.
 *   - Host-pointer entry points are synchronous: they copy in, run on the
 *     library's HIP stream and copy out before returning.  The *_device
 *     variants take device pointers and are stream-ordered on the handle's
 *     stream (dbie_op_stream); they return before the GPU finishes.
 *   - A handle is not thread-safe; distinct handles are independent.
 *
 * Each entry point names the reference interface (file:line) it replaces.
 */
#ifndef DBIE_H
#define DBIE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int dbie_status;
enum {
    DBIE_OK = 0,
    DBIE_BAD_ARG = 1,        /* ArgumentError in the reference */
    DBIE_DIM_MISMATCH = 2,   /* DimensionMismatch (src/solver/multi_rhs.jl:105) */
    DBIE_OOM = 3,            /* device or host allocation failed */
    DBIE_HIP_ERROR = 4,      /* a HIP runtime call failed */
    DBIE_NOT_CONVERGED = 5,  /* iterative solve hit itmax (Krylov stats.solved == false) */
    DBIE_NO_DEVICE = 6,      /* no usable gfx950 device */
    DBIE_PENDING = 7         /* request server: the request has not finished yet */
};

/* Precision of the far-field kernel x RHS-block contraction. */
typedef enum {
    DBIE_FP64 = 0,   /* FP64 tile generation and FP64 accumulation (the plain path) */
    DBIE_FP32 = 1,   /* FP32 tile generation, FP32 accumulation */
    DBIE_BF16 = 2,   /* FP32 tile generation, BF16 MFMA contraction, FP32 accumulation */
    DBIE_BF16X3 = 3, /* FP32 generation, split-BF16 (hi*hi+hi*lo+lo*hi) MFMA, FP32 accumulation */
    DBIE_FP16X3 = 4  /* FP64 differences, FP32 generation, scaled split-FP16 (3 products) MFMA,
                        FP32 accumulation: entry error about 2e-7 */
} dbie_precision;

/* Which boundary operator a plain apply evaluates (src/kernel/laplace3d.jl:18-52). */
typedef enum {
    DBIE_OP_DT = 0,  /* adjoint double layer  DT[i,j] = w_j n_i.(x_i-y_j)/(4 pi r^3) */
    DBIE_OP_S = 1,   /* single layer          S[i,j]  = w_j /(4 pi r)                  */
    DBIE_OP_D = 2    /* double layer          D[i,j]  = w_j n_j.(x_j-x_i)/(4 pi r^3)  */
} dbie_kernel;

typedef struct dbie_geom dbie_geom;  /* a DielectricInterface of flat tensor-GL panels */
typedef struct dbie_op dbie_op;      /* a BatchedDielectricOperator resident on the GPU */

/* Near-field options: keyword arguments of
   lhs_dielectric_box3d_fmm3d_corrected (src/solver/dielectric_box3d.jl:48-58) and
   batched_lhs_dielectric_box3d_fmm3d_corrected (src/solver/multi_rhs.jl:60-66). */
typedef struct {
    double up_tol;           /* Bernstein threshold atol (default 1e-9) */
    int max_order;           /* p_up ceiling (default 8) */
    int correct_edges;       /* build touching-pair (adaptive) blocks (default 1) */
    double adaptive_atol;    /* default = up_tol */
    double adaptive_rtol;    /* default sqrt(eps) */
    int adaptive_n_GL;       /* 0 = source panel's n_quad */
    int adaptive_max_depth;  /* default 20 */
    int build_corrections;   /* 0 = plain (uncorrected) operator, as lhs_dielectric_box3d_fmm3d */
    int n_threads;           /* host threads for the near-field build (0 = all cores) */
    int device_build;        /* build the correction blocks on the GPU (default 1; 0 = host threads) */
} dbie_near_opts;

/* Solve options: keywords of solve_dielectric_box3d_block (src/solver/multi_rhs.jl:270-276)
   and solve_gmres (src/utils/linear_algebra.jl:3). */
typedef struct {
    double rtol;             /* default 1e-10 */
    double atol;             /* default 0 */
    int itmax;               /* default 500 */
    int restart;             /* Krylov basis size before restart (0 = min(itmax, 100)) */
    int precondition;        /* right preconditioner: 1 = diagonal scaling 1/t
                                (dielectric_diagonal_scaling, default), 2 = near-field: one Neumann
                                term of (diag(t) + N)^-1, N the near panel pairs' full interaction (one
                                near-field product per iteration; GMRES needs ~40% fewer iterations;
                                no counterpart in the reference), 3 = near-field with two Jacobi
                                sweeps z <- D^-1 (x - N z) (both Neumann terms, two near-field
                                products), 0 = none */
    dbie_precision precision;/* DBIE_FP64 = plain GMRES; otherwise mixed-precision IR */
    double inner_rtol;       /* mixed: inner-solve relative tolerance (default 1e-4) */
    int max_refine;          /* mixed: maximum FP64 refinement steps (default 10) */
    int check_residual;      /* FP64 path (and mixed with verify = 0): evaluate the true FP64
                                residual after the solve (one extra matvec, outside time_ms; fills
                                max_rel_residual; default 1) */
    int krylov;              /* 0 = GMRES per column on a shared block matvec (default),
                                1 = block GMRES: one block Krylov space for all K columns
                                (Krylov.block_gmres, src/solver/multi_rhs.jl:144-148) */
    int recycle;             /* mixed precision with krylov = 0: the inner solves after the first
                                deflate the first one's Krylov space (GCRO: minimise over it and
                                keep new directions orthogonal to its image), so the refinement
                                continues that space instead of restarting (default 1) */
    int verify;              /* mixed precision: 1 = the solve ends only on an FP64 residual that
                                meets the tolerance (default); 0 = it may also end, as the FP64
                                GMRES does, on its residual estimate: after the first FP64
                                residual the inner solve runs to half the remaining tolerance, and
                                when every column's inner residual norm meets half the tolerance
                                and the low-precision operator's bound on its part of the residual
                                (2 x its relative error x 10 x ||correction||) a quarter of it, the
                                solve ends (split FP16: corrections up to 2.5e-4 ||F|| at rtol
                                1e-10; the forms without near-pair exclusion in practice always
                                end on an FP64 residual); check_residual then
                                evaluates the true FP64 residual after the solve, as for FP64 */
} dbie_solve_opts;

typedef struct {
    int niter;               /* total Krylov iterations (inner iterations for mixed) */
    int nrefine;             /* FP64 residual evaluations (mixed) */
    int solved;              /* all columns reached the tolerance */
    double max_rel_residual; /* max_k ||F_k - A Sigma_k|| / ||F_k|| in FP64 */
    double time_ms;          /* device time of the solve */
} dbie_solve_stats;

/* ---- library ---------------------------------------------------------- */
const char *dbie_version(void);
/* Copies the calling thread's last error message into buf (NUL-terminated). */
dbie_status dbie_last_error(char *buf, size_t len);
/* Initialise the HIP device (returns DBIE_NO_DEVICE if there is no GPU). */
dbie_status dbie_init(int device);
/* Default-initialise option structs. */
void dbie_near_opts_default(dbie_near_opts *o);
void dbie_solve_opts_default(dbie_solve_opts *o);

/* ---- geometry (host) -------------------------------------------------- */
/* gausslegendre(n) (FastGaussQuadrature, ascending nodes) */
dbie_status dbie_gauss_legendre(int n, double *x, double *w);
/* best_grid_mn (src/utils/best_grid.jl:8-126) */
dbie_status dbie_best_grid(double W, double H, double alpha, int *m, int *n);
/* single_dielectric_box3d (src/shape/box3d_single.jl:49-70) */
dbie_status dbie_geometry_box(double Lx, double Ly, double Lz, int n_quad, double l_ec, double alpha,
                              dbie_geom **out);
/* DielectricInterface(panels, ...) of rect_panel3d_discretize panels
   (src/shape/box3d.jl:22-64, src/core/panels.jl:64-68); corners P x 4 x 3
   (lb, rb, rt, lt), normals P x 3, is_edge P (may be NULL). */
dbie_status dbie_geometry_from_panels(int64_t P, int n_quad, const double *corners, const double *normals,
                                      const int *is_edge, dbie_geom **out);
/* DielectricInterface(panels, ...) with the panels' own quadrature (FlatPanel.points, .weights, .gl_xs,
   .gl_ws, .bary_weights; src/core/panels.jl:5-22,64-68), taken as given rather than regenerated:
   points 3 x N (x, y, z of node n at 3n .. 3n+2, panel order, node order of the panels), weights N,
   gl_xs / gl_ws / bary_weights n_quad (one rule for every panel; bary_weights may be NULL: computed from
   gl_xs, gl_ws).  Corners, normals and is_edge as for dbie_geometry_from_panels. */
dbie_status dbie_geometry_from_panel_nodes(int64_t P, int n_quad, const double *corners, const double *normals,
                                           const int *is_edge, const double *points, const double *weights,
                                           const double *gl_xs, const double *gl_ws, const double *bary_weights,
                                           dbie_geom **out);
/* multi_dielectric_box3d (src/shape/box3d_multi.jl:358-386): n_boxes axis-aligned boxes, boxes
   n_boxes x 6 (centre x, y, z, Lx, Ly, Lz; the reference's (center, Lx, Ly, Lz) NamedTuples),
   permittivity epses[b] inside box b and eps_out outside.  Panels: the pieces of every box face not
   shared with another box (box order, face order), then the shared faces; each region meshed by
   rect_panel3d_adaptive_panels.  dbie_geom_eps returns the panels' permittivities. */
dbie_status dbie_geometry_multi_box(int n_quad, double l_ec, int n_boxes, const double *boxes, const double *epses,
                                    double eps_out, double alpha, dbie_geom **out);
/* per-panel eps_in / eps_out (interface.eps_in / eps_out, src/core/panels.jl:64-68) of a geometry
   whose mesher assigned them (dbie_geometry_multi_box); DBIE_BAD_ARG for other geometries */
dbie_status dbie_geom_eps(const dbie_geom *g, double *eps_in, double *eps_out);
dbie_status dbie_geom_sizes(const dbie_geom *g, int64_t *P, int64_t *N, int *n_quad);
/* node SoA, 7 x N rows x,y,z,nx,ny,nz,w (eachpoint + all_weights, src/core/panels.jl:85-118) */
dbie_status dbie_geom_nodes(const dbie_geom *g, double *xyz_nrm_w);
dbie_status dbie_geom_panels(const dbie_geom *g, double *corners, double *normals, int *is_edge);
void dbie_geom_destroy(dbie_geom *g);

/* ---- near-field list (host, exposed for checking) --------------------- */
/* build_neighbor_list (src/kernel/laplace3d_near.jl:240-392).  Call with NULL
   arrays to get the counts, then again with arrays of those sizes. Pairs are
   sorted by (source panel, target panel). */
dbie_status dbie_neighbor_list(const dbie_geom *g, int max_order, double atol, int correct_edges,
                               int64_t *n_up, int64_t *n_adp, int *up_src, int *up_trg, int *up_order,
                               int *adp_src, int *adp_trg);

/* ---- operator --------------------------------------------------------- */
/* batched_lhs_dielectric_box3d_fmm3d_corrected (src/solver/multi_rhs.jl:60-100):
   geometry SoA, correction blocks (laplace3d_DT_corrections,
   src/kernel/laplace3d_near_upsampling.jl:272) and diag t are built and kept
   resident in HBM.  eps_in/eps_out are per panel (length P). */
dbie_status dbie_operator_create(const dbie_geom *g, const double *eps_in, const double *eps_out,
                                 const dbie_near_opts *opts, dbie_op **out);
void dbie_operator_destroy(dbie_op *op);
dbie_status dbie_op_sizes(const dbie_op *op, int64_t *N, int64_t *n_blocks, int *block_dim);
/* Host time of the operator build in ms (no counterpart in the reference): t[0] near-pair list,
   t[1] correction blocks, t[2] derived near-field data (full near blocks, pair mask), t[3] layout
   permutation + uploads. */
dbie_status dbie_op_setup_ms(const dbie_op *op, double *t);
/* Update eps (diag only; geometry and corrections depend on geometry alone). */
dbie_status dbie_op_set_eps(dbie_op *op, const double *eps_in, const double *eps_out);
/* Export the correction blocks: pairs (src, trg) and blocks[b][t + c*bd] (FP64). */
dbie_status dbie_op_corrections(const dbie_op *op, int *src, int *trg, double *blocks);
/* As dbie_op_corrections for the corrections of kernel DBIE_OP_DT or DBIE_OP_D (laplace3d_DT_corrections /
   laplace3d_D_corrections, src/kernel/laplace3d_near_upsampling.jl:272-282); the D blocks are built on the
   first call. */
dbie_status dbie_op_corrections_kernel(dbie_op *op, dbie_kernel kernel, int *src, int *trg, double *blocks);
/* HBM bytes the operator holds. */
dbie_status dbie_op_device_bytes(const dbie_op *op, int64_t *bytes);

/* mul!(Y, op, X) (src/solver/multi_rhs.jl:103-136): Y = DT X + C X + t.*X.
   precision selects the far-field contraction (DBIE_FP64 = plain path). */
dbie_status dbie_apply(dbie_op *op, int K, const double *X, int64_t ldx, double *Y, int64_t ldy,
                       dbie_precision precision);
/* Plain boundary operators without corrections/diag: laplace3d_DT / _S / _D applied
   matrix-free (src/kernel/laplace3d.jl:18-52 and the *_fmm3d maps :84-137). */
dbie_status dbie_apply_kernel(dbie_op *op, dbie_kernel kernel, int K, const double *X, int64_t ldx,
                              double *Y, int64_t ldy);
/* laplace3d_DT_fmm3d_corrected / laplace3d_D_fmm3d_corrected (src/kernel/laplace3d_near.jl:395-445):
   Y = K X + C_K X for kernel DBIE_OP_DT or DBIE_OP_D (no diagonal), C_K the near corrections of that
   kernel on the operator's near-pair list (laplace3d_DT_corrections / laplace3d_D_corrections,
   src/kernel/laplace3d_near_upsampling.jl:272-282; the D blocks use the source normal and are built on
   the first call).  DBIE_BAD_ARG for DBIE_OP_S or an operator created without corrections. */
dbie_status dbie_apply_kernel_corrected(dbie_op *op, dbie_kernel kernel, int K, const double *X, int64_t ldx,
                                        double *Y, int64_t ldy);
/* Dense lhs_dielectric_box3d (src/solver/dielectric_box3d.jl:1-17): A = DT + diag(t), N x N. */
dbie_status dbie_lhs_dense(dbie_op *op, double *A, int64_t lda);
/* solve_lu (src/utils/linear_algebra.jl:1): LU with partial pivoting on the GPU. */
dbie_status dbie_solve_lu(int64_t n, const double *A, int64_t lda, int nrhs, double *B, int64_t ldb);

/* ---- right-hand sides --------------------------------------------------- */
/* rhs_dielectric_box3d(interface, PointSource, eps_src) (src/solver/dielectric_box3d.jl:95-104),
   one column per charge: F[i,k] = -q_k n_i.(x_i-s_k)/(4 pi r^3 eps_src). */
dbie_status dbie_rhs_points(dbie_op *op, int K, const double *src_xyz, const double *q, double eps_src,
                            double *F, int64_t ldf);
/* rhs_dielectric_box3d(interface, VolumeSource[, eps_src]) (src/solver/dielectric_box3d.jl:106-124) and
   the Vector{VolumeSource} form (src/solver/multi_rhs.jl:168-211).  Column k uses source points
   col_ptr[k] .. col_ptr[k+1]-1 of xyz (3 x S), w, rho.  screening = 1 applies SharpScreening
   through the interface bounds (src/core/sources.jl:390-406) and eps_src is then 1. */
dbie_status dbie_rhs_volume(dbie_op *op, int K, const int64_t *col_ptr, const double *xyz, const double *w,
                            const double *rho, int screening, double eps_src, double *F, int64_t ldf);
/* rhs_dielectric_box3d(interface, boxes, epses, eps_out, vs) (src/solver/dielectric_box3d.jl:224-232)
   and rhs_dielectric_box3d_fmm3d(interface, vss; screen_boxes, screen_epses, screen_eps_out)
   (src/solver/multi_rhs.jl:168-211): the sources screened box by box
   (screened_volume_source(boxes, epses, eps_out, vs, mode), src/core/sources.jl:408-515; boxes
   n_boxes x 6 = centre x, y, z, Lx, Ly, Lz; mode 0 SharpScreening, 1 SoftMixPermittivity(bandwidth),
   2 SoftMixInversePermittivity(bandwidth); default tolerance), then eps_src = 1. */
dbie_status dbie_rhs_volume_boxes(dbie_op *op, int K, const int64_t *col_ptr, const double *xyz, const double *w,
                                  const double *rho, int n_boxes, const double *boxes, const double *epses,
                                  double eps_out, int mode, double bandwidth, double *F, int64_t ldf);
/* screened_volume_source on the GPU: rho_out[s] = the density of point s (xyz 3 x S) screened by
   the permittivity at its position.  form 0: the several-box overload
   (src/core/sources.jl:408-515); form 1: the one-region overload screened_volume_source(Lx, Ly, Lz,
   vs, eps_in, eps_out, mode) (:315-388) with one box (epses[0] = eps_in).  tol < 0 takes the
   reference's default (sqrt(eps) x the largest side). */
dbie_status dbie_screen_boxes(int64_t S, const double *xyz, const double *rho, int n_boxes, const double *boxes,
                              const double *epses, double eps_out, int mode, double bandwidth, double tol, int form,
                              double *rho_out);
/* screened_volume_source(interface, vs, mode; tol) (src/core/sources.jl:390-406) through the
   operator's interface bounds and uniform permittivities (DBIE_BAD_ARG when they vary). */
dbie_status dbie_screen_interface(dbie_op *op, int64_t S, const double *xyz, const double *rho, int mode,
                                  double bandwidth, double tol, double *rho_out);

/* ---- solves ------------------------------------------------------------ */
/* solve_dielectric_box3d_block / Krylov.block_gmres (src/solver/multi_rhs.jl:270-286) and
   solve_gmres (src/utils/linear_algebra.jl:3-6): A Sigma = F for K columns on the GPU.
   precision = DBIE_FP64 runs FP64 GMRES; any other value runs mixed-precision iterative
   refinement (FP64 residuals, low-precision inner GMRES). */
dbie_status dbie_solve_block(dbie_op *op, int K, const double *F, int64_t ldf, double *Sigma, int64_t lds,
                             const dbie_solve_opts *opts, dbie_solve_stats *stats);

/* Residual history of the last solve on this operator (the Krylov stats the reference returns,
   src/solver/multi_rhs.jl:285): *n entries of (kind, max_k ||r_k|| / ||F_k||) into kind_value (2 doubles
   each, at most cap entries written); kind 0 = a Krylov iteration's residual estimate (outer GMRES, or the
   low-precision inner GMRES of a mixed solve), 1 = a true FP64 residual, 2 = a residual updated with the
   low-precision operator or taken from the inner solve's estimate. */
dbie_status dbie_op_residual_history(const dbie_op *op, int cap, int *n, double *kind_value);

/* No counterpart in the reference (its callers run solve_dielectric_box3d_block once per interface,
   src/solver/multi_rhs.jl:270-286, src/solver/lattice_batch.jl:285-304): the block systems of B operators
   of different geometries solved as ONE Krylov process.  Member b brings K[b] columns F[b] (N_b x K[b],
   leading dimension ldf[b]) and receives Sigma[b]; the columns are held in N_max x sum(K) blocks (rows
   past a member's N are zero), and every iteration applies all members at once, each on its own stream
   with its own tuned launch shape and HIP graphs, so that small systems fill the GPU together and share
   one host synchronisation per iteration.  Each operator may appear once (merge its columns).  With
   opts->precondition = 1 the right preconditioner (a scalar per member for single-contrast interfaces)
   is left out, GMRES being invariant to it, and multi-contrast members are solved unpreconditioned;
   with 2 or 3 each member's own near-field preconditioner acts on the member's rows and columns, on the
   member's stream.  opts and stats as for
   dbie_solve_block (one stats record for the whole solve), except that opts->krylov must be 0: block
   GMRES orthogonalises the columns against each other, which would mix columns of different operators
   (DBIE_BAD_ARG otherwise). */
dbie_status dbie_solve_fused(int B, dbie_op *const *ops, const int *K, const double *const *F, const int64_t *ldf,
                             double *const *Sigma, const int64_t *lds, const dbie_solve_opts *opts,
                             dbie_solve_stats *stats);

/* ---- induced potential off the surface --------------------------------- */
/* laplace3d_pottrg / laplace3d_pottrg_fmm3d (src/kernel/laplace3d.jl:69-82,169-191):
   U[t,k] = sum_i w_i Sigma[i,k] / (4 pi |trg_t - x_i|) for M targets (3 x M). */
dbie_status dbie_pottrg(dbie_op *op, int64_t M, const double *trg, int K, const double *Sigma, int64_t lds,
                        double *U, int64_t ldu);

/* ---- near-corrected potential at off-surface targets -------------------- */
typedef struct dbie_tmap dbie_tmap;  /* the M x N map Sigma -> u[Sigma](targets) */

/* laplace3d_pottrg_fmm3d_corrected_hcubature(interface, targets, fmm_tol, hcubature_atol,
   range_factor) (src/kernel/laplace3d_near_hcubature.jl:337-375): refines the interface around
   the targets (src/kernel/laplace3d_near.jl:99-150), builds the prolongation of densities onto
   it (:152-225) and the near-target corrections by adaptive cubature to hcub_atol for every
   non-edge refined panel within range_factor * l / n_quad of a target (:47-70,
   src/kernel/laplace3d_near_hcubature.jl:193-251).  Host-side build on n_threads threads
   (0 = all cores); no device is needed until the first evaluation.  trg: 3 x M. */
dbie_status dbie_tmap_create(const dbie_geom *g, int64_t M, const double *trg, double hcub_atol,
                             double range_factor, int n_threads, dbie_tmap **out);
/* As dbie_tmap_create (same reference function); device_cubature = 1 integrates the near-target
   cubatures on the GPU (one wave per panel-target pair, the host cubature's operation order, so
   the map is bit-identical); device_cubature = 2 also builds the refinement around the targets
   (one thread per coarse panel walking its subtree in the host's order) and the prolongation on
   the GPU, bit-identical to the host build, and keeps the prolongation there (for n_quad <= 5;
   otherwise, or when a subtree is deeper than 21 levels, the host builds both). */
dbie_status dbie_tmap_create_ex(const dbie_geom *g, int64_t M, const double *trg, double hcub_atol,
                                double range_factor, int n_threads, int device_cubature, dbie_tmap **out);
void dbie_tmap_destroy(dbie_tmap *tm);
/* sizes: targets, coarse nodes, refined nodes, cubature pairs, correction nonzeros, and
   cubatures that stopped at the evaluation cap (2e6 integrand calls) before reaching hcub_atol */
/* No counterpart in the reference: wall-clock ms of the build phases (refinement, prolongation,
   near-target pairs + cubatures, correction assembly), for setup timing. */
dbie_status dbie_tmap_setup_ms(const dbie_tmap *tm, double *t /* [4] */);
dbie_status dbie_tmap_sizes(const dbie_tmap *tm, int64_t *M, int64_t *N, int64_t *N_refined,
                            int64_t *n_pairs, int64_t *nnz_corr, int64_t *n_capped);
/* refined nodes as 4 x N_refined rows (x; y; z; w), parent panel and split flag per refined
   panel (any pointer may be NULL) */
dbie_status dbie_tmap_refined(const dbie_tmap *tm, double *xyzw, int *parent, int *from_split);
/* prolongation (CSR, N_refined rows, coarse-node columns) and corrections (CSR, M rows,
   refined-node columns); sizes from dbie_tmap_sizes (prolongation nnz = p_ptr[N_refined]) */
dbie_status dbie_tmap_matrices(const dbie_tmap *tm, int64_t *p_ptr, int *p_col, double *p_val,
                               int64_t *c_ptr, int *c_col, double *c_val);
/* U (M x K) = u[Sigma] at the targets: the returned LinearMap applied to N x K densities */
dbie_status dbie_tmap_apply(dbie_tmap *tm, int K, const double *Sigma, int64_t lds, double *U, int64_t ldu);
dbie_status dbie_tmap_apply_device(dbie_tmap *tm, int K, const double *dSigma, int64_t lds, double *dU,
                                   int64_t ldu);
/* Reaction matrix, the induced term of four_index_matrix (src/solver/multi_rhs.jl:299-321):
   R[a,b] = sum_{s in group a} q_s u[Sigma_b](trg_s), groups = target ranges
   [grp_ptr[a], grp_ptr[a+1]) covering 0..M-1, q_s = w_s rho_s.  R: G x K (host, ldr). */
dbie_status dbie_reaction_matrix(dbie_tmap *tm, int G, const int64_t *grp_ptr, const double *q, int K,
                                 const double *Sigma, int64_t lds, double *R, int64_t ldr);
dbie_status dbie_reaction_matrix_device(dbie_tmap *tm, int G, const int64_t *grp_ptr, const double *q, int K,
                                        const double *dSigma, int64_t lds, double *R, int64_t ldr);
/* physicists' Gauss-Hermite rule (weight exp(-x^2)), ascending nodes; used for Gaussian charge
   blobs (no counterpart in the reference, whose GaussianVolumeSource uses a midpoint grid,
   src/core/sources.jl:517-548) */
dbie_status dbie_gauss_hermite(int n, double *x, double *w);

/* ---- request server: continuous batching of many block solves ----------- */
/* No counterpart in the reference, whose callers loop over solve_dielectric_box3d_block
   (src/solver/multi_rhs.jl:270-286).  Requests (a box and its point charges) are admitted by
   device-memory footprint, set up on host threads (mesh, near field, RHS) while other requests
   run, and solved by a few concurrent slots; ready requests with the same options are fused into one
   solve: those on one cached operator (keyed by geometry; the permittivities set its diagonal) as one
   column block, those on other operators as members of a fused Krylov process (dbie_solve_fused,
   per-column Krylov only).  Results are the same alone and in a mix, within the solve tolerance. */
typedef struct dbie_server dbie_server;
typedef struct {
    double hbm_budget;     /* device bytes for operators + workspaces (0 = 80% of free memory) */
    int setup_workers;     /* concurrent request setups (default 2) */
    int setup_threads;     /* host threads per near-field build (0 = cores / setup_workers) */
    int solve_slots;       /* concurrent block solves, each on its operator's stream (default 2) */
    int max_batch_rhs;     /* columns fused into one block solve (default 256) */
    int cache_ops;         /* keep idle operators for later requests (default 1) */
    int inject_oom_every;  /* test hook: every n-th operator build fails with DBIE_OOM (0 = off) */
} dbie_server_opts;
typedef struct {
    int64_t id;
    dbie_status status;    /* the request's own outcome (DBIE_NOT_CONVERGED, DBIE_OOM, ...) */
    int state;             /* 0 queued, 1 setup, 2 ready, 3 solving, 4 done, 5 failed, 6 cancelled */
    int64_t N;
    int K, batch_K;        /* own columns; columns of the block solve it was part of */
    int solved, niter, nrefine;
    double max_rel_residual;
    int op_cache_hit;
    double queue_ms, setup_ms, solve_ms, total_ms;  /* arrival -> setup, setup, solve, arrival -> result */
    char message[256];
} dbie_request_info;
typedef struct {
    int64_t completed, failed, operator_builds, cache_hits, batches, fused_requests, cached_operators;
    int64_t bytes_in_use, peak_bytes, budget_bytes;
    int64_t multi_geometry_batches;  /* batches that fused requests of several operators (dbie_solve_fused) */
} dbie_server_summary;

void dbie_server_opts_default(dbie_server_opts *o);
dbie_status dbie_server_create(const dbie_server_opts *o, dbie_server **out);
/* stops the workers after the solves in progress and frees every operator */
void dbie_server_destroy(dbie_server *s);
/* single_dielectric_box3d(Lx, Ly, Lz, n_quad, l_ec, eps_in, eps_out) with K point charges (pts 3 x K,
   q) at eps_src = eps_in; opts NULL = defaults with DBIE_FP16X3.  Returns at once with an id. */
dbie_status dbie_server_submit_box(dbie_server *s, double Lx, double Ly, double Lz, int n_quad, double l_ec,
                                   double eps_in, double eps_out, int K, const double *pts, const double *q,
                                   const dbie_solve_opts *opts, int64_t *id);
/* waits up to timeout_s seconds (< 0: no limit); DBIE_PENDING if unfinished.  info may be NULL. */
dbie_status dbie_server_wait(dbie_server *s, int64_t id, double timeout_s, dbie_request_info *info);
/* copies the N x K density of a finished request (Sigma may be NULL) and forgets the request */
dbie_status dbie_server_take(dbie_server *s, int64_t id, double *Sigma, int64_t lds);
dbie_status dbie_server_cancel(dbie_server *s, int64_t id);
/* paused != 0: solve slots start no new block solve (setup continues), e.g. to drain the device or
   to let ready requests accumulate; 0 resumes */
dbie_status dbie_server_pause(dbie_server *s, int paused);
dbie_status dbie_server_stats(dbie_server *s, dbie_server_summary *st);

/* ---- kernel timing (HIP events on the operator's stream) ----------------- */
/* No counterpart in the reference (it has no profiler hooks); used by the benchmark to
   time the dominant kernels inside a timed region.  kernel ids: 0 = far field FP64,
   1 = far field low precision (FP32/BF16/BF16X3/FP16X3), 2 = near-field combine.  Reading
   synchronizes the stream and clears the records. */
dbie_status dbie_op_profile(dbie_op *op, int enable);
dbie_status dbie_op_profile_read(dbie_op *op, int kernel, int64_t *launches, double *total_ms);

/* ---- launch-shape autotuner ---------------------------------------------- */
/* No counterpart in the reference (its far field is a CPU loop with no launch shape).  The
   far-field launch shape (source split count; for the split-FP16 form at K > 128 also the kernel
   variant) is tuned online per shape class (precision, near exclusion, nodes per panel, K in
   32-column groups, N to a quarter octave): the first applies of a class each run one untried
   candidate under HIP events, without extra launches or waits, and once all are timed the fastest
   is kept for the process.  enable: 1 on (the default; environment DBIE_AUTOTUNE=0 starts it off),
   0 off (default shapes), -1 unchanged; clear != 0 forgets the cached choices. */
dbie_status dbie_autotune(int enable, int clear);
/* the settled choice for this operator's class at K columns: *splits = 0 while the class is still
   being tuned or was never applied (either output may be NULL) */
dbie_status dbie_autotune_choice(const dbie_op *op, int K, dbie_precision precision, int *splits, int *variant);
/* pins the choice for this operator's class at K columns (e.g. one saved from an earlier run):
   splits >= 1 (clamped to what the launch accepts), variant -1 = default kernel, 0 = the general
   split-FP16 kernel, 1 = the warp-specialised split-FP16 kernel with one consumer wave per SIMD,
   2 / 3 / 4 = the GEMM over the stored split-FP16 kernel matrix (2: four waves per workgroup, the
   stored entries loaded into registers; 3: eight waves, through the LDS; 4: eight waves, into
   registers; see dbie_op_stored_far; as 1 where the operator keeps none) */
dbie_status dbie_autotune_set(const dbie_op *op, int K, dbie_precision precision, int splits, int variant);

/* ---- HIP graphs ------------------------------------------------------------ */
/* No counterpart in the reference.  An operator apply whose arguments (buffers, K, precision,
   launch shape) repeat, as in every GMRES iteration, is captured as a HIP graph on its second call
   and replayed from then on (on by default; environment DBIE_GRAPHS=0 starts operators with it
   off).  enable: 1 on, 0 off (eager launches; captured graphs stay cached), -1 unchanged.
   captures / replays (may be NULL) receive the operator's counts so far. */
dbie_status dbie_op_graphs(dbie_op *op, int enable, int64_t *captures, int64_t *replays);

/* No counterpart in the reference (its far field is FMM3D).  The split-FP16 far field of applies
   with more than 128 columns can stream a stored copy of its kernel matrix (hi/lo fp16 entries of
   every cross-plane node pair, 4 bytes per pair: 13.6 GB at N = 64,384) instead of generating the
   entries on the fly; the results are bit-identical at the same launch shape.  The operator builds
   it on the first apply whose tuned launch shape uses it, when free device memory holds it with
   max(8 GiB, 1/16 of the device) to spare.  mode: 1 allowed (the default; environment
   DBIE_STORED_FAR=0 starts operators with 0), 0 never (frees a stored matrix), -1 unchanged.
   bytes (may be NULL) receives the device bytes the stored matrix holds (0 if none). */
dbie_status dbie_op_stored_far(dbie_op *op, int mode, int64_t *bytes);

/* ---- device-pointer variants (stream-ordered on the operator's stream) ---- */
void *dbie_op_stream(dbie_op *op);
/* waits until the work queued on the operator's stream (device-pointer applies) has finished */
dbie_status dbie_op_synchronize(dbie_op *op);
dbie_status dbie_apply_device(dbie_op *op, int K, const double *dX, int64_t ldx, double *dY, int64_t ldy,
                              dbie_precision precision);
/* No counterpart in the reference: dbie_apply_device for callers that cannot allocate device memory
   (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  Every
   workspace (operand staging, partial sums, X copy) comes from one fixed 256 MiB device array the
   library holds per GPU until the process ends; the call allocates nothing and launches on the
   operator's stream (stream-ordered; dbie_op_synchronize waits).  DBIE_BAD_ARG when the apply needs
   more scratch (one source split of N x K FP64 partial sums plus the operand staging; N = 64,384,
   K = 256 fits).  The launch shape is the autotuner's settled one (no tuning here), graphs are not
   used.  Every call on one GPU shares that GPU's array: calls that overlap on one GPU (two streams,
   two threads) race.  The first split-FP16 call on an operator also computes its fp16 scale and
   synchronizes the stream. */
dbie_status dbie_apply_device_noalloc(dbie_op *op, int K, const double *dX, int64_t ldx, double *dY, int64_t ldy,
                                      dbie_precision precision);
/* dbie_solve_block with F and Sigma resident in device memory (solve_dielectric_box3d_block,
   src/solver/multi_rhs.jl:270-286); returns after the solve completes. */
dbie_status dbie_solve_block_device(dbie_op *op, int K, const double *dF, int64_t ldf, double *dSigma, int64_t lds,
                                    const dbie_solve_opts *opts, dbie_solve_stats *stats);
/* Device memory helpers for callers without a GPU array library (no counterpart in the
   reference, whose arrays live in host memory). */
dbie_status dbie_device_alloc(size_t bytes, void **dptr);
dbie_status dbie_device_free(void *dptr);
dbie_status dbie_memcpy_h2d(void *dst, const void *src, size_t bytes);
dbie_status dbie_memcpy_d2h(void *dst, const void *src, size_t bytes);
/* Device-to-device copy of an N x K column-major block by a kernel (synchronous). */
dbie_status dbie_copy_block_device(int64_t N, int K, const double *dsrc, int64_t lds, double *ddst, int64_t ldd);

#ifdef __cplusplus
}
#endif
#endif /* DBIE_H */
