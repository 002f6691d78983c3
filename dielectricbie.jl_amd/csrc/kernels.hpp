// kernels.hpp -- launchers of the gfx950 kernels (implemented in *.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

namespace dbie {

// Device-resident node streams of one interface (structure of arrays).
struct DevGeom {
    int64_t N = 0;
    const double *x = nullptr, *y = nullptr, *z = nullptr;      // node coordinates
    const double *nx = nullptr, *ny = nullptr, *nz = nullptr;   // node normals
    const double *w = nullptr;                                  // quadrature weights
    const float *xf = nullptr, *yf = nullptr, *zf = nullptr;    // FP32 copies (mixed path)
    const float *nxf = nullptr, *nyf = nullptr, *nzf = nullptr;
    const float *wf = nullptr;
    // face structure for skipping same-plane tiles: plane id per 64-node tile (-1 = mixed)
    const int *tile_plane = nullptr;
    // near panel pairs: bit (p, q) at near_mask[p * mask_words + q / 32] (target p, source q);
    // rows padded with two zero words.  npp = nodes per panel.
    const uint32_t *near_mask = nullptr;
    int npp = 0;
    int mask_words = 0;
};

// Block-sparse near correction on the device.
struct DevBSR {
    int P = 0, bd = 0;
    int64_t nb = 0;
    const int *row_ptr = nullptr, *col = nullptr;
    const double *val64 = nullptr;  // nb x bd x bd blocks, element (t, c) at bsr_storage_index(bd, mfma_order, t, c)
    const float *val32 = nullptr;
    const float *valup32 = nullptr; // nb blocks of the full near interaction (correction + direct), FP32
    // the full near blocks for the split-FP16 operator (bd = 16): per block the hi and then the lo
    // halves of 2^rexp[p] C[t, c] / w_c, row-major (t, c), as _Float16; rexp per block row p
    const void *valh = nullptr;
    const int *rexp = nullptr;
    // super-rows of kNearSuperRows consecutive block rows (the split-FP16 near kernel): the sorted union
    // of their block columns, sr_col[sr_ptr[q] .. sr_ptr[q + 1]), and per union entry u the block of each
    // row, sr_blk[u * kNearSuperRows + r] (-1: none)
    const int *sr_ptr = nullptr, *sr_col = nullptr, *sr_blk = nullptr;
    int sr_q = 0;
    bool mfma_order = false;        // bd = 16: blocks stored in MFMA operand order
};
// Storage position of block element (t, c): MFMA operand order for bd = 16
// (k-slice c/4, then lane (c%4)*16 + t), row-major t*bd + c otherwise.
int64_t bsr_storage_index(int bd, bool mfma_order, int t, int c);

enum FarMode { FAR_DT = 0, FAR_S = 1, FAR_D = 2 };

// Launch shape of a far-field contraction, as picked by the autotuner: splits = number of
// source splits (0: the built-in heuristic), variant = -1: the warp-specialised split-FP16
// kernel where it applies (two consumer waves per SIMD), 1: the same with one consumer wave per
// SIMD, 0: never, 2 / 3 / 4: the streamed GEMM over the stored kernel matrix (2: four waves per
// workgroup, the stored entries straight into registers; 3: eight waves, through the LDS; 4: eight
// waves, into registers) where the operator keeps one, else as 1.
struct FarTune {
    int splits = 0;
    int variant = -1;
};

// The split-FP16 far field's kernel matrix kept in HBM (launch variants 2, 3 and 4; K > 128): per 256-target super-tile and per 32-source stage of each source
// tile of 256 nodes that is not on the super-tile's plane, a 32 KiB block of hi/lo fp16 entries in
// MFMA operand order, as the generating kernels compute them.  tile_off (ntile x (ntile + 1) ints,
// device): stored source tiles before (super-tile T, source tile st) in row-major order.
struct FarStored {
    const void *g = nullptr;
    const int *tile_off = nullptr;
    int ntile = 0;
};
// host: fills tile_off from the plane ids of the 256-node tiles; returns the number of blocks
int64_t far_stored_layout(const std::vector<int> &tile_plane, int64_t N, std::vector<int> &tile_off);
size_t far_stored_block_bytes();
// writes every stored block (gscale: the operator's fp16 scale; exclude_near as in far_field_mfma)
void far_stored_build(hipStream_t st, const DevGeom &g, float gscale, bool exclude_near, const FarStored &fs);

// Far-field partial sums: part[s][k][i] = sum over sources in split s of G(i,j) w_j X[j,k]
// (G of the chosen mode, self pair excluded, 1/(4 pi) applied).  Returns the split count S.
int far_field_f64(hipStream_t st, const DevGeom &g, FarMode mode, int K, const double *X, int64_t ldx,
                  double *part, int64_t part_cap, const FarTune &tune = FarTune());
// exclude_near (16-node panels with a near mask): leave the near panel pairs out, as the split-FP16
// far field does
int far_field_f32(hipStream_t st, const DevGeom &g, int K, const double *X, int64_t ldx, double *part,
                  int64_t part_cap, const FarTune &tune = FarTune(), bool exclude_near = false);
// MFMA contraction of the DT far field: tiles generated on the fly and contracted with the
// RHS block.  prec: 0 = FP64 MFMA, 2 = BF16 MFMA, 3 = split BF16 (3 products), 4 = split FP16.
// scratch: far_field_mfma_scratch_bytes(N, prec, K) bytes of device memory (the
// pre-scaled RHS block w.*X in FP64, or its tile-major bf16 split).
// gscale: power-of-two scale of the kernel entries for the fp16 form (prec 4).  exclude_near
// (prec 4 only): leave out the panel pairs of g.near_mask, whose full interaction the caller adds.
// ev0/ev1 (optional) are recorded immediately before and after the contraction kernel.
int far_field_mfma(hipStream_t st, const DevGeom &g, int prec, int K, const double *X, int64_t ldx, double *part,
                   int64_t part_cap, void *scratch, float gscale, bool exclude_near, hipEvent_t ev0 = nullptr,
                   hipEvent_t ev1 = nullptr, const FarTune &tune = FarTune(), const FarStored *fs = nullptr);
// largest useful split count of a far-field launch (splits hold >= 256 sources)
int far_field_max_splits(int64_t N);
size_t far_field_mfma_scratch_bytes(int64_t N, int prec, int K);
// Power-of-two scale for the fp16 kernel entries: 2^(14 - e) with max |G_ij| < 2^e over all
// node pairs.  tmp4: 4 bytes of device memory.  Synchronises the stream; 0 on a HIP error.
float far_field_f16_scale(hipStream_t st, const DevGeom &g, void *tmp4);
bool far_field_mfma_supported(int prec, int K);
int64_t far_field_part_elems(int64_t N, int K);  // workspace needed for any far-field launch

// The split-FP16 far field's operands of an apply (in its scratch): hi and lo halves of
// 2^col_exponent(cmax[k]) w_j X[j, k], per tile of `tile` sources and 16-column block one chunk
// in MFMA B-operand order (zsplit_idx in devmath.hpp).
struct NearH3 {
    const void *zh = nullptr, *zl = nullptr;
    const unsigned *cmax = nullptr;
    int Kp = 0, tile = 0;
};
void far_f16_operands(void *scratch, int64_t N, int K, NearH3 *out);

// Y[:,k] = sum_s part[s][k] + dscale*diag.*X[:,k] + C X[:,k] (C in FP64 or FP32 values).
enum NearVals { NEAR_C64 = 0, NEAR_C32 = 1, NEAR_UP32 = 2 };
// NEAR_UP32 with 16 x 16 blocks and K >= 16: with h3 (the far field's split-FP16 operands of the
// same X) and C.valh, the full near blocks multiply them on the FP16 MFMA (three products of the
// hi/lo splits); otherwise xt, near_xt_bytes(N, K) bytes of scratch, takes a tile-major FP32 copy
// of X for the FP32 blocks, stored lane-major (bsr_lane_major_index).
// yscale (split-FP16 path only): Y[i, k] is multiplied by yscale[i]; then Y may be X itself
void near_combine(hipStream_t st, const DevBSR &C, NearVals vals, int S, const double *part, int64_t N, int K,
                  const double *X, int64_t ldx, const double *diag, double *Y, int64_t ldy, float *xt = nullptr,
                  const NearH3 *h3 = nullptr, const double *yscale = nullptr);
// The split-FP16 operands of X (zh, zl, cmax as far_f16_operands lays them out in `scratch`, which
// holds far_field_mfma_scratch_bytes(N, DBIE_FP16X3, K) bytes): what the far field computes first.
void f16_split_operands(hipStream_t st, int64_t N, int K, const double *X, int64_t ldx, const double *w, void *scratch);
constexpr int kNearSuperRows = 4;
// C.valh and C.rexp (device arrays of nb * 512 halves and P ints) from C.valup32 and the node weights w
void near_h3_pack(hipStream_t st, const DevBSR &C, const double *w, void *valh, int *rexp);
size_t near_xt_bytes(int64_t N, int K);
int64_t bsr_lane_major_index(int t, int c);

// Right-hand sides.
void rhs_points(hipStream_t st, const DevGeom &g, int K, const double *src, const double *q, double eps_src,
                double *F, int64_t ldf);
void rhs_volume(hipStream_t st, const DevGeom &g, int K, const int64_t *col_ptr, const double *pos,
                const double *wrho, double *F, int64_t ldf, double scale);
void screen_sharp(hipStream_t st, int64_t S, const double *pos, const double *rho, const double *w,
                  const double *lo, const double *hi, double tol, double eps_in, double eps_out, double *wrho);

// Screening of volume source densities by boxes of permittivity (device arrays lo, hi: nb x 3,
// eps: nb).  mode: SCREEN_SHARP, SCREEN_SOFT_MIX (eps mixed), SCREEN_SOFT_MIX_INV (1/eps mixed);
// single = the one-region formulas of the interface overloads (nb = 1).
enum ScreenMode { SCREEN_SHARP = 0, SCREEN_SOFT_MIX = 1, SCREEN_SOFT_MIX_INV = 2 };
struct ScreenBoxes {
    int nb = 0;
    const double *lo = nullptr, *hi = nullptr, *eps = nullptr;
    double eps_out = 1.0, tol = 0.0, bandwidth = 1.0;
    int mode = SCREEN_SHARP;
    bool single = false;
};
// out[s] = w[s] * rho'_s (w != NULL) or rho'_s, rho' the screened density
void screen_boxes(hipStream_t st, int64_t S, const double *pos, const double *rho, const double *w,
                  const ScreenBoxes &sb, double *out);

// Off-surface single-layer potential U[t,k] = sum_i w_i X[i,k]/(4 pi |t - x_i|).
void pottrg(hipStream_t st, const DevGeom &g, int64_t M, const double *trg, int K, const double *X, int64_t ldx,
            double *U, int64_t ldu);

// Dense A = DT + diag(t) (column-major, lda).
void dense_lhs(hipStream_t st, const DevGeom &g, const double *diag, double *A, int64_t lda);

// ---- Krylov primitives on column blocks (N x K, column-major, ld = N) ----
// A Krylov basis V is stored in FP64 or FP32 (the low-precision inner solves); the vectors it
// multiplies and the coefficients are FP64, and every sum is accumulated in FP64.
// h[i*K + k] = <V_i[:,k], W[:,k]> for i < nv   (V_i = V + i*N*K); partial buffer of
// dots_partial_elems(N, K, nv) doubles.
int64_t dots_partial_elems(int64_t N, int K, int nv);
void block_dots(hipStream_t st, const double *V, int nv, const double *W, int64_t N, int K, double *partial,
                double *h);
void block_dots(hipStream_t st, const float *V, int nv, const double *W, int64_t N, int K, double *partial,
                double *h);
// W[:,k] -= sum_i V_i[:,k] h[i*K+k]
void block_subtract(hipStream_t st, const double *V, int nv, const double *h, double *W, int64_t N, int K);
void block_subtract(hipStream_t st, const float *V, int nv, const double *h, double *W, int64_t N, int K);
// out[:,k] = in[:,k] * s[k] (s may be inverted: out = in / s when invert)
void col_scale(hipStream_t st, const double *in, const double *s, bool invert, double *out, int64_t N, int K);
void col_scale(hipStream_t st, const double *in, const double *s, bool invert, float *out, int64_t N, int K);
// out[:, k] = in[:, k] / sqrt(sq[k]) (zero where sq[k] <= 0)
void col_normalize(hipStream_t st, const double *in, const double *sq, double *out, int64_t N, int K);
void col_normalize(hipStream_t st, const double *in, const double *sq, float *out, int64_t N, int K);
// coef[i, k] = a[i, k] - om[i, k] sum_{l < n1[k]} om[l, k] a[l, k] for i < n1[k], zero for n1[k] <= i < nb
// (a, om, coef: nb x K, element (i, k) at i K + k)
void deflate_coef(hipStream_t st, const double *a, const double *om, const int *n1, int nb, int K, double *coef);
// out[:,k] = d .* in[:,k]  (d per row; d null: out = in)
// leading dimensions ldi / ldo (0: N)
void row_scale(hipStream_t st, const double *d, const double *in, double *out, int64_t N, int K, int64_t ldi = 0,
               int64_t ldo = 0);
void row_scale(hipStream_t st, const double *d, const float *in, double *out, int64_t N, int K, int64_t ldi = 0,
               int64_t ldo = 0);
// X[:,k] += sum_i V_i[:,k] y[i*K+k]
void block_combine(hipStream_t st, const double *V, int nv, const double *y, double *X, int64_t N, int K);
void block_combine(hipStream_t st, const float *V, int nv, const double *y, double *X, int64_t N, int K);
// R = F - R (residual from AX stored in R), column-wise
// out = out - z + s .* py (row scale s; py and z N x K contiguous, out with leading dimension ldo, 0: N)
void jacobi_update(hipStream_t st, const double *s, const double *py, const double *z, double *out, int64_t N, int K,
                   int64_t ldo = 0);
void sub_from(hipStream_t st, const double *F, double *R, int64_t N, int K);
// X += D
void add_to(hipStream_t st, const double *D, double *X, int64_t N, int K);
// X -= D
void sub_to(hipStream_t st, const double *D, double *X, int64_t N, int K);
// copy with leading dimensions
void copy_cols(hipStream_t st, const double *src, int64_t lds, double *dst, int64_t ldd, int64_t N, int K);
// C = beta C + sum of nb partial p x q blocks stored back to back in P (C has leading dimension ldc)
void sum_blocks(hipStream_t st, int nb, int p, int q, const double *P, double beta, double *C, int64_t ldc);
// dst (rows x cols) = unit lower trapezoid of src (Householder vectors below a geqrf R)
void unit_lower_copy(hipStream_t st, int rows, int cols, const double *src, int64_t lds, double *dst, int64_t ldd);
// dst = upper triangle of the n x n block src (zeros below the diagonal)
void triu_copy(hipStream_t st, int n, const double *src, int64_t lds, double *dst, int64_t ldd);

// ---- off-surface targets (targets.hip) ----
// Y (R x K) = A X (+ Y if accumulate) for CSR A with R rows whose columns index rows of X.
void csr_spmm(hipStream_t st, int64_t R, const int64_t *ptr, const int *col, const double *val, int K, const double *X,
              int64_t ldx, double *Y, int64_t ldy, bool accumulate);
// Qt (g.N x G, column-major) = w_r / (4 pi) sum_{s in group a} q_s / |t_s - y_r|; groups by grp_ptr (G + 1).
// Qc (N x G, ld N) = P^T Qr (Qr: Nr x G, ld Nr) for a prolongation whose refined nodes of coarse
// panel c are first[c] .. first[c] + count[c] - 1 and interpolate from that panel's npp nodes only.
// Returns false (nothing launched) for npp > 32.
bool prolong_t(hipStream_t st, int P, int npp, const int64_t *first, const int *count, const int64_t *p_ptr,
               const int *p_col, const double *p_val, int G, const double *Qr, int64_t Nr, double *Qc, int64_t N);
void reaction_q(hipStream_t st, const DevGeom &g, int G, const int64_t *grp_ptr, const double *trg, const double *q,
                double *Qt);
// Refinement of an interface around targets and the prolongation onto it (targets.hip), the host
// builder's arithmetic restated.  RefineGrid: the host TargetGrid on the device (targets t, 3 x M;
// cell lists start / idx).  refine_count: leaves per coarse panel (-1: subtree deeper than the
// device stack); refine_emit: the leaves of panel p from offset[p] on.  rules: glx[nq], bary[nq] of
// the coarse interface; p_ptr: rows of the prolongation (1 or nq^2 entries per refined node).
struct RefineGrid {
    const double *t = nullptr;
    double lo[3] = {0, 0, 0}, h = 1;
    int64_t dim[3] = {1, 1, 1};
    const int64_t *start = nullptr;
    const int *idx = nullptr;
};
void refine_count(hipStream_t st, int P, const double *corners, const RefineGrid &G, int64_t M, double limit,
                  double range_factor, int nq, int64_t *count);
void refine_emit(hipStream_t st, int P, const double *corners, const RefineGrid &G, int64_t M, double limit,
                 double range_factor, int nq, const int64_t *offset, double *out_corners, int *parent, int *from_split);
void refined_nodes(hipStream_t st, int64_t Pr, int nq, const double *corners, const double *glx, const double *glw,
                   double *x, double *y, double *z, double *w);
void prolongation(hipStream_t st, int64_t Pr, int nq, const double *coarse_corners, const double *rules,
                  const int *parent, const int *from_split, const double *x, const double *y, const double *z,
                  const int64_t *p_ptr, int *p_col, double *p_val);
// Y[idx[e]] += val[e] for n entries with distinct idx.
void scatter_add(hipStream_t st, int64_t n, const int64_t *idx, const double *val, double *Y);

// ---- near-field correction blocks on the device (nearbuild.hip) ----
// Upsampled pairs from per-source-panel caches (upos then Mt, at cache_off, order cache_nu) and
// touching pairs from host K_up blocks (bd x bd row-major): writes the correction (FP64, FP32) and
// the full near interaction (FP32) of each pair's BSR slot.  max_nu: largest cache order (LDS size).
// Returns false (nothing launched) when bd x max_nu^2 doubles exceed kNearBuildLds.
constexpr size_t kNearBuildLds = 64 * 1024;
// Near-pair candidate scan on the device (the middle step of the pair list).  pdata: kPairData
// doubles per panel (pair_prep); corners: 12 per panel; is_edge: per panel.  With ptr == nullptr it
// writes count[i] (candidates of source panel i); with ptr (P + 1 offsets) it writes the candidates:
// target panel, smallest Bernstein radius over its nodes in the ball (inf if none), touch flag.
constexpr int kPairData = 20;
void pair_scan(hipStream_t st, const DevGeom &g, int P, int npp, const double *pdata, const double *corners,
               const int *is_edge, double tc2, bool correct_edges, int64_t *count, const int64_t *ptr, int *ctrg,
               double *crho, unsigned char *ctouch);
// mode (both builders): 0 = the Dt correction blocks, 1 = the D blocks (source normal, FP64 values
// only: val32 and vup32 may be null).
// Touching pairs on the device: K_up blocks (n_adp x bd x bd, row-major) and max-depth flags of the
// quadtree rule, one wave per pair.  adp_map: 10 doubles per pair (cc, bma, dma, scale of the source
// panel); rules: glx[nq], bary[nq], ns[nGL], ws[nGL].  Returns false (nothing launched) when the
// walk's LDS, adaptive_blocks_lds(), exceeds kNearBuildLds.  hit[k]: 1 = a cell reached max_depth
// unconverged, 2 = the walk stopped after max_nodes nodes (block incomplete; build it on the host).
size_t adaptive_blocks_lds(int nq, int nGL, int max_depth);
bool adaptive_blocks(hipStream_t st, const DevGeom &g, int nq, int nGL, int n_adp, const int *adp_src,
                     const int *adp_trg, const int *adp_rep, const double *adp_map, const double *rules, double atol,
                     double rtol, int max_depth, int max_nodes, double *kup, int *hit, int mode = 0);
// The near pairs' full interaction (correction + direct Dt entries) in FP64, C.nb x bd x bd values in
// the storage order of C.val64 (the near-field preconditioner of FP64 solves).
void near_full_blocks(hipStream_t st, const DevGeom &g, const DevBSR &C, double *full64);
bool build_near_blocks(hipStream_t st, const DevGeom &g, int bd, bool mfma_order, int n_up, const int *up_src,
                       const int *up_trg, const int *up_slot, const int *cache_off, const int *cache_nu,
                       const double *caches, int max_nu, int n_adp, const int *adp_src, const int *adp_trg,
                       const int *adp_slot, const double *adp_kup, double *val64, float *val32, float *vup32,
                       int mode = 0);

// Near-target cubatures on the device (laplace3d_pottrg_corrections_hcubature): for each pair,
// the adaptive Genz-Malik cubature over [-1, 1]^2 of the panel potential against the n_quad^2
// Lagrange basis functions, summed error <= atol, in the host cubature's operation order.
// pdat: 13 doubles per pair (panel centre, b - a, d - a, area scale, target); rules: glx[nq],
// bary[nq].  Iout: n_quad^2 per pair; status: 0 converged, 1 stopped at maxevals, 2 more than
// bcap boxes (not computed).  boxes: target_cubature_scratch(nwg, nq, bcap) bytes.
constexpr int kCubatureMaxNpp = 64;
struct CubatureRule {
    double l2, l3, l4, l5;
    double w[5], wp[4];
};
size_t target_cubature_scratch(int nwg, int nq, int bcap);
void target_cubatures(hipStream_t st, int npairs, int nq, const double *pdat, const double *rules,
                      const CubatureRule &rule, double atol, long long maxevals, int bcap, int nwg, double *boxes,
                      double *Iout, int *status);

}  // namespace dbie
