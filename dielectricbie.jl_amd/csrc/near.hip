// near.hip -- block-sparse near-field correction fused with the far-field split
// reduction and the diagonal contrast term (mul! of BatchedDielectricOperator,
// src/solver/multi_rhs.jl:103-130):
//     Y[i,k] = sum_s part[s][k][i] + t_i X[i,k] + sum_b C_b[i - bd*p, :] . X[bd*src_b + :, k]
// Block rows are target panels; each block is a dense bd x bd (bd = n_quad^2) tile.  C_b is the
// correction block (FP64 or FP32), or, for the split-FP16 operator whose far field leaves the
// near panel pairs out, the full near interaction (correction + direct term, FP32).
//
// Two kernels:
//  * K >= 16 and bd = 16: FP64 MFMA (v_mfma_f64_16x16x4_f64).  One workgroup per block
//    row and 256 columns; the A operand is the correction block, stored in MFMA operand
//    order so that each k-slice is one contiguous 512-byte load; the B operand is the
//    matching 16-row slice of X.  Block rows are remapped so that consecutive rows (which
//    share most source panels) run on the same XCD and reuse its L2.
//  * K >= 16 and bd = 4, 9 or 25 (n_quad = 2, 3, 5; row-major blocks): FP64 MFMA on the block
//    padded to RT x 16 rows and KS x 4 source columns, one workgroup per block row and 256
//    columns, operands loaded straight from the block and X (no LDS, no barrier).
//  * otherwise: one thread per (row, column), VALU.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <stdexcept>

#include "devmath.hpp"
#include "kernels.hpp"

namespace dbie {

// Position of block element (t, c) in storage: MFMA order for bd = 16
// (k-slice c/4, then lane (c%4)*16 + t), row-major otherwise.
__host__ __device__ inline int blk_idx(int bd, bool mfma_order, int t, int c) {
    return mfma_order ? (c >> 2) * 64 + (c & 3) * 16 + t : t * bd + c;
}
// Lane-major order of a 16 x 16 block (full near blocks of the split-FP16 operator): lane
// (c/4)*16 + t holds columns 4(c/4) .. 4(c/4)+3 contiguously, one 16-byte load per lane.
__host__ __device__ inline int lane_major_idx(int t, int c) { return ((c >> 2) * 16 + t) * 4 + (c & 3); }

int64_t bsr_storage_index(int bd, bool mfma_order, int t, int c) { return blk_idx(bd, mfma_order, t, c); }
int64_t bsr_lane_major_index(int t, int c) { return lane_major_idx(t, c); }

namespace {
typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef _Float16 v8h __attribute__((ext_vector_type(8)));

// One output (row, column) per CS threads: thread cs of the group takes the block columns
// c = cs, cs + CS, ... of every block of the row, and the group's partial sums meet in LDS.  Thread
// tid = row + bd (cs + CS (kc + KTe pp)).
template <typename VT>
__global__ __launch_bounds__(256) void near_combine_kernel(int P, int bd, int mfma_order, int CS,
                                                           const int *__restrict__ row_ptr,
                                                           const int *__restrict__ col, const VT *__restrict__ val,
                                                           int S, const double *__restrict__ part, int64_t N, int K,
                                                           int KTe, int RP, const double *__restrict__ X, int64_t ldx,
                                                           const double *__restrict__ diag, double *__restrict__ Y,
                                                           int64_t ldy) {
    __shared__ double s_red[256];
    const int tid = threadIdx.x;
    const int row = tid % bd;
    int rest = tid / bd;
    const int cs = rest % CS;
    rest /= CS;
    const int kc = rest % KTe;
    const int pp = rest / KTe;
    const int p = blockIdx.x * RP + pp;
    const int k = blockIdx.y * KTe + kc;
    const bool valid = pp < RP && p < P && k < K;
    const int64_t i = (int64_t)(valid ? p : 0) * bd + row;
    const double *xk = X + (int64_t)(valid ? k : 0) * ldx;
    double corr = 0.0;
    if (valid) {
        const int b1 = row_ptr[p + 1];
        for (int b = row_ptr[p]; b < b1; ++b) {
            const VT *v = val + (int64_t)b * bd * bd;
            const double *xs = xk + (int64_t)col[b] * bd;
            for (int c = cs; c < bd; c += CS)
                corr += (double)v[mfma_order == 2 ? lane_major_idx(row, c) : blk_idx(bd, mfma_order != 0, row, c)] * xs[c];
        }
    }
    if (CS > 1) {
        s_red[tid] = corr;
        __syncthreads();
    }
    if (!valid || cs != 0) return;
    for (int j = 1; j < CS; ++j) corr += s_red[tid + j * bd];
    double acc = diag[i] * xk[i];
    for (int s2 = 0; s2 < S; ++s2) acc += part[((int64_t)s2 * K + k) * N + i];
    Y[i + (int64_t)k * ldy] = acc + corr;
}

// XCD-aware block-row order: workgroups are dispatched round-robin over 8 XCDs, so
// workgroup w runs on XCD w % 8; give each XCD a contiguous range of block rows.
__device__ inline int xcd_row(int w, int P) {
    const int per = (P + 7) / 8;
    const int r = (w % 8) * per + w / 8;
    return r;
}

// NQ column blocks of 16 per wave: a workgroup covers 64 NQ columns (NQ = 1, 2, 4 by K).
template <typename VT, int NQ>
__global__ __launch_bounds__(256) void near_mfma_kernel(int P, const int *__restrict__ row_ptr,
                                                        const int *__restrict__ col, const VT *__restrict__ val,
                                                        int S, const double *__restrict__ part, int64_t N, int K,
                                                        const double *__restrict__ X, int64_t ldx,
                                                        const double *__restrict__ diag, double *__restrict__ Y,
                                                        int64_t ldy) {
    constexpr int CW = 64 * NQ;         // columns per workgroup
    __shared__ double s_out[CW][17];  // [column within the column group][row], padded
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int p = xcd_row(blockIdx.x, P);
    const int k0 = blockIdx.y * CW;
    if (p >= P) return;  // uniform per workgroup
    v4d acc[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) acc[q] = v4d{0.0, 0.0, 0.0, 0.0};
    // column of this lane for each of its NQ column blocks (clamped; masked by a zero multiplier)
    int64_t coff[NQ];
    double cmask[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        const int c = k0 + (wave + 4 * q) * 16 + (lane & 15);
        cmask[q] = c < K ? 1.0 : 0.0;
        coff[q] = (int64_t)(c < K ? c : K - 1) * ldx;
    }
    const int b0 = row_ptr[p], b1 = row_ptr[p + 1];
    // operands of one block: A (4 k-slices) and X (4 k-slices x 4 column blocks), prefetched one block ahead
    double an[4], xn[4][NQ];
    auto load = [&](int b, double (&a)[4], double (&xv)[4][NQ]) {
        const VT *v = val + (int64_t)b * 256;
        const int64_t src_row = (int64_t)col[b] * 16 + (lane >> 4);
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) {
            a[kk] = (double)v[kk * 64 + lane];  // C_b[t = lane%16][c = kk*4 + lane/16]
#pragma unroll
            for (int q = 0; q < NQ; ++q) xv[kk][q] = X[src_row + kk * 4 + coff[q]];
        }
    };
    if (b0 < b1) load(b0, an, xn);
    for (int b = b0; b < b1; ++b) {
        double a[4], xv[4][NQ];
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) {
            a[kk] = an[kk];
#pragma unroll
            for (int q = 0; q < NQ; ++q) xv[kk][q] = xn[kk][q];
        }
        if (b + 1 < b1) load(b + 1, an, xn);
#pragma unroll
        for (int kk = 0; kk < 4; ++kk)
#pragma unroll
            for (int q = 0; q < NQ; ++q) acc[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[kk], xv[kk][q], acc[q], 0, 0, 0);
    }
    // D(row = 4r + lane/16, col = lane%16) -> LDS [col][row]
#pragma unroll
    for (int q = 0; q < NQ; ++q)
#pragma unroll
        for (int r = 0; r < 4; ++r) s_out[(wave + 4 * q) * 16 + (lane & 15)][4 * r + (lane >> 4)] = acc[q][r] * cmask[q];
    __syncthreads();
    // each thread finishes CW / 16 outputs: row = tid % 16, columns tid/16 + 16*m
    const int row = tid & 15;
    const int64_t i = (int64_t)p * 16 + row;
    const double t = diag[i];
    for (int m = 0; m < CW / 16; ++m) {
        const int cl = (tid >> 4) + 16 * m;
        const int c = k0 + cl;
        if (c >= K) continue;
        double y = s_out[cl][row] + t * X[i + (int64_t)c * ldx];
        for (int s = 0; s < S; ++s) y += part[((int64_t)s * K + c) * N + i];
        Y[i + (int64_t)c * ldy] = y;
    }
}

// Row-major bd x bd blocks (bd <= 16 RT, bd <= 4 KS) on v_mfma_f64_16x16x4_f64, padded to RT x 16
// rows and 4 KS source columns: each lane loads its A operands straight from the block (L1-resident:
// the four waves read the same block) with zeros outside it, and its B operands from the source
// panel's rows of X, so the loop has no barrier and no LDS; each wave multiplies 4 column blocks of
// 16 columns.  The epilogue adds the partial sums and the diagonal term from the registers.
template <typename VT, int RT, int KS>
__global__ __launch_bounds__(256) void near_mfma_rm_kernel(int P, int bd, const int *__restrict__ row_ptr,
                                                           const int *__restrict__ col, const VT *__restrict__ val,
                                                           int S, const double *__restrict__ part, int64_t N, int K,
                                                           const double *__restrict__ X, int64_t ldx,
                                                           const double *__restrict__ diag, double *__restrict__ Y,
                                                           int64_t ldy) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int p = xcd_row(blockIdx.x, P);
    const int k0 = blockIdx.y * 256;
    if (p >= P) return;  // uniform per workgroup
    v4d acc[RT][4];
#pragma unroll
    for (int rt = 0; rt < RT; ++rt)
#pragma unroll
        for (int q = 0; q < 4; ++q) acc[rt][q] = v4d{0.0, 0.0, 0.0, 0.0};
    int64_t coff[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int c = k0 + (wave + 4 * q) * 16 + (lane & 15);
        coff[q] = (int64_t)(c < K ? c : K - 1) * ldx;
    }
    // this lane's A element (t, c) of row tile rt and k-slice kk: t = 16 rt + lane % 16,
    // c = 4 kk + lane / 16; offsets into the block, or -1 outside it
    int aoff[RT][KS];
#pragma unroll
    for (int rt = 0; rt < RT; ++rt)
#pragma unroll
        for (int kk = 0; kk < KS; ++kk) {
            const int t = 16 * rt + (lane & 15), c = 4 * kk + (lane >> 4);
            aoff[rt][kk] = t < bd && c < bd ? t * bd + c : -1;
        }
    const int b0 = row_ptr[p], b1 = row_ptr[p + 1];
    const int nv = bd * bd;
    for (int b = b0; b < b1; ++b) {
        const VT *v = val + (int64_t)b * nv;
        const int64_t src = (int64_t)col[b] * bd;
#pragma unroll
        for (int kk = 0; kk < KS; ++kk) {
            const int c = kk * 4 + (lane >> 4);
            double xv[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) xv[q] = c < bd ? X[src + c + coff[q]] : 0.0;
#pragma unroll
            for (int rt = 0; rt < RT; ++rt) {
                const double a = aoff[rt][kk] >= 0 ? (double)v[aoff[rt][kk]] : 0.0;
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    acc[rt][q] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, xv[q], acc[rt][q], 0, 0, 0);
            }
        }
    }
    // D(row = 4r + lane/16, col = lane%16) of row tile rt: row 16 rt + 4 r + lane / 16 of the block row
    const int64_t ib = (int64_t)p * bd;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int c = k0 + (wave + 4 * q) * 16 + (lane & 15);
        if (c >= K) continue;
#pragma unroll
        for (int rt = 0; rt < RT; ++rt)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = 16 * rt + 4 * r + (lane >> 4);
                if (row >= bd) continue;
                const int64_t i = ib + row;
                double y = acc[rt][q][r] + diag[i] * X[i + (int64_t)c * ldx];
                for (int s2 = 0; s2 < S; ++s2) y += part[((int64_t)s2 * K + c) * N + i];
                Y[i + (int64_t)c * ldy] = y;
            }
    }
}

// X (N x K, FP64) -> tile-major FP32 copy Xt[((j / 16) * Kpad + k) * 16 + j % 16], zero beyond K
__global__ void x_tiles_f32_kernel(const double *__restrict__ X, int64_t ldx, int64_t N, int K, int Kpad,
                                   float *__restrict__ Xt) {
    const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int k = blockIdx.y;
    if (j >= N) return;
    Xt[((j >> 4) * Kpad + k) * 16 + (j & 15)] = k < K ? (float)X[j + (int64_t)k * ldx] : 0.0f;
}

// Full near blocks (FP32, lane-major) times the tile-major FP32 copy of X on the FP64 MFMA: per
// block one 16-byte load of A and one of B per column block and lane.  MFMA kk contracts source
// columns c = 4 (lane / 16) + kk, a permutation of the block's 16 columns.
template <int NQ>
__global__ __launch_bounds__(256) void near_up32_kernel(int P, const int *__restrict__ row_ptr,
                                                        const int *__restrict__ col, const float *__restrict__ val,
                                                        int S, const double *__restrict__ part, int64_t N, int K,
                                                        const float *__restrict__ Xt, int Kpad,
                                                        const double *__restrict__ X, int64_t ldx,
                                                        const double *__restrict__ diag, double *__restrict__ Y,
                                                        int64_t ldy) {
    constexpr int CW = 64 * NQ;  // columns per workgroup
    __shared__ double s_out[CW][17];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int p = xcd_row(blockIdx.x, P);
    const int k0 = blockIdx.y * CW;
    if (p >= P) return;  // uniform per workgroup
    v4d acc[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) acc[q] = v4d{0.0, 0.0, 0.0, 0.0};
    const int b0 = row_ptr[p], b1 = row_ptr[p + 1];
    float4 an, xn[NQ];
    auto load = [&](int b, float4 &a, float4 (&xv)[NQ]) {
        a = *reinterpret_cast<const float4 *>(val + (int64_t)b * 256 + lane * 4);
        const float *xs = Xt + (int64_t)col[b] * Kpad * 16 + 4 * (lane >> 4);
#pragma unroll
        for (int q = 0; q < NQ; ++q)
            xv[q] = *reinterpret_cast<const float4 *>(xs + (int64_t)(k0 + (wave + 4 * q) * 16 + (lane & 15)) * 16);
    };
    if (b0 < b1) load(b0, an, xn);
    for (int b = b0; b < b1; ++b) {
        const float4 a = an;
        float4 xv[NQ];
#pragma unroll
        for (int q = 0; q < NQ; ++q) xv[q] = xn[q];
        if (b + 1 < b1) load(b + 1, an, xn);
        const double av[4] = {a.x, a.y, a.z, a.w};
#pragma unroll
        for (int kk = 0; kk < 4; ++kk)
#pragma unroll
            for (int q = 0; q < NQ; ++q) {
                const double bv = kk == 0 ? xv[q].x : kk == 1 ? xv[q].y : kk == 2 ? xv[q].z : xv[q].w;
                acc[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[kk], bv, acc[q], 0, 0, 0);
            }
    }
#pragma unroll
    for (int q = 0; q < NQ; ++q)
#pragma unroll
        for (int r = 0; r < 4; ++r) s_out[(wave + 4 * q) * 16 + (lane & 15)][4 * r + (lane >> 4)] = acc[q][r];
    __syncthreads();
    const int row = tid & 15;
    const int64_t i = (int64_t)p * 16 + row;
    const double t = diag[i];
    for (int m = 0; m < CW / 16; ++m) {
        const int cl = (tid >> 4) + 16 * m;
        const int c = k0 + cl;
        if (c >= K) continue;
        double y = s_out[cl][row] + t * X[i + (int64_t)c * ldx];
        for (int s = 0; s < S; ++s) y += part[((int64_t)s * K + c) * N + i];
        Y[i + (int64_t)c * ldy] = y;
    }
}
// One workgroup per block row p: the row's full near blocks (FP32, lane-major) as 2^rexp[p] C[t, c] / w_c
// split into hi and lo halves (row-major per block: hi at b * 512 + 16 t + c, lo 256 further), with
// rexp[p] chosen so that the row's largest |C[t, c] / w_c| scales below 2^14.  Dividing by the source
// weights lets the blocks multiply the far field's split operands w .* X directly.
__global__ __launch_bounds__(256) void near_h3_pack_kernel(const int *__restrict__ row_ptr, const int *__restrict__ col,
                                                           const float *__restrict__ vup, const double *__restrict__ w,
                                                           _Float16 *__restrict__ vh, int *__restrict__ rexp) {
    __shared__ double s_m[4];
    const int p = blockIdx.x;
    const int b0 = row_ptr[p], b1 = row_ptr[p + 1];
    const int tid = threadIdx.x, t = tid >> 4, c = tid & 15;
    double m = 0.0;
    for (int b = b0; b < b1; ++b)
        m = fmax(m, fabs((double)vup[(int64_t)b * 256 + lane_major_idx(t, c)] / w[(int64_t)col[b] * 16 + c]));
    for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_down(m, off, 64));
    if ((tid & 63) == 0) s_m[tid >> 6] = m;
    __syncthreads();
    m = fmax(fmax(s_m[0], s_m[1]), fmax(s_m[2], s_m[3]));
    int e = 0;
    if (m > 0.0) (void)frexp(m, &e);  // m < 2^e
    const int r = 14 - e;
    if (tid == 0) rexp[p] = r;
    for (int b = b0; b < b1; ++b) {
        const double v = ldexp((double)vup[(int64_t)b * 256 + lane_major_idx(t, c)] / w[(int64_t)col[b] * 16 + c], r);
        const _Float16 hi = (_Float16)(float)v;
        vh[(int64_t)b * 512 + 16 * t + c] = hi;
        vh[(int64_t)b * 512 + 256 + 16 * t + c] = (_Float16)(float)(v - (double)(float)hi);
    }
}

// The far field's split-FP16 operands are in the layout of zsplit_idx (tiles of kZt sources); the
// host checks the far field's tile height against it at run time.  Column part of the index of a
// 16-byte unit (8 sources of one column): zsplit_idx = zsplit_row(j) + zsplit_col(k) for 8-aligned j.
__device__ inline int64_t zsplit_row(int64_t j, int Kp) { return (j >> 5) * (int64_t)Kp * kZt + ((j >> 3) & 3) * 128; }
__device__ inline int zsplit_col(int k) { return (k & ~15) * kZt + (k & 15) * 8; }

// Full near blocks of the split-FP16 operator on the FP16 MFMA (v_mfma_f32_16x16x32_f16): the 32-deep
// contraction takes two blocks of the row at a time, lanes of k-groups 0-1 the first (sources 0-7,
// 8-15), 2-3 the second, and each block pair costs the three products hi.hi + hi.lo + lo.hi of the
// blocks' and the far field's splits (the operands the far field of the same apply left in its
// scratch: no conversion of X here).  One workgroup per block row and 64 NQ columns, NQ column
// blocks of 16 per wave; the next pair's operands are loaded while the current one multiplies.
// The epilogue undoes the row and column scalings and adds the partial sums and the diagonal term.
template <int NQ>
__global__ __launch_bounds__(256) void near_h3_kernel(int P, const int *__restrict__ row_ptr, const int *__restrict__ col,
                                                      const _Float16 *__restrict__ vh, const int *__restrict__ rexp,
                                                      int S, const double *__restrict__ part, int64_t N, int K,
                                                      const _Float16 *__restrict__ zh, const _Float16 *__restrict__ zl,
                                                      const unsigned *__restrict__ cmax, int Kp,
                                                      const double *X, int64_t ldx, const double *__restrict__ diag,
                                                      const double *__restrict__ yscale, double *Y, int64_t ldy) {
    constexpr int CW = 64 * NQ;  // columns per workgroup
    __shared__ double s_out[CW][17];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int p = xcd_row(blockIdx.x, P);
    const int k0 = blockIdx.y * CW;
    if (p >= P) return;  // uniform per workgroup
    // lane: row / column t, sources 8 h .. 8 h + 7 of the pair's block `second`
    const int t = lane & 15, h = (lane >> 4) & 1, second = lane >> 5;
    int64_t zoff[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        const int c = k0 + (wave + 4 * q) * 16 + t;
        zoff[q] = zsplit_col(c < K ? c : K - 1);  // columns past K: clamped, never stored
    }
    v4f acc[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) acc[q] = v4f{0.f, 0.f, 0.f, 0.f};
    const int b0 = row_ptr[p], b1 = row_ptr[p + 1];
    auto load = [&](int b, uint4 &ah, uint4 &al, uint4 (&bh)[NQ], uint4 (&bl)[NQ]) {
        const int bb = b + second;
        const int bv = bb < b1 ? bb : b;  // an odd row's last pair: the second block's A is zero
        const _Float16 *a = vh + (int64_t)bv * 512 + 16 * t + 8 * h;
        ah = *reinterpret_cast<const uint4 *>(a);
        al = *reinterpret_cast<const uint4 *>(a + 256);
        if (bb >= b1) ah = al = uint4{0u, 0u, 0u, 0u};
        const int64_t tb = zsplit_row((int64_t)col[bv] * 16 + 8 * h, Kp);  // the panel's sources 8 h .. 8 h + 7
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            bh[q] = *reinterpret_cast<const uint4 *>(zh + tb + zoff[q]);
            bl[q] = *reinterpret_cast<const uint4 *>(zl + tb + zoff[q]);
        }
    };
    uint4 nah, nal, nbh[NQ], nbl[NQ];
    if (b0 < b1) load(b0, nah, nal, nbh, nbl);
    for (int b = b0; b < b1; b += 2) {
        const v8h ah = __builtin_bit_cast(v8h, nah), al = __builtin_bit_cast(v8h, nal);
        v8h bh[NQ], bl[NQ];
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            bh[q] = __builtin_bit_cast(v8h, nbh[q]);
            bl[q] = __builtin_bit_cast(v8h, nbl[q]);
        }
        if (b + 2 < b1) load(b + 2, nah, nal, nbh, nbl);
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            acc[q] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bh[q], acc[q], 0, 0, 0);
            acc[q] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bl[q], acc[q], 0, 0, 0);
            acc[q] = __builtin_amdgcn_mfma_f32_16x16x32_f16(al, bh[q], acc[q], 0, 0, 0);
        }
    }
    // D(row = 4 (lane / 16) + r, col = lane % 16) -> LDS [col][row], scaled back
    const int rp = rexp[p];
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        const int c = k0 + (wave + 4 * q) * 16 + (lane & 15);
        const double sc = ldexp(1.0, -(rp + col_exponent(cmax[c < K ? c : K - 1])));
#pragma unroll
        for (int r = 0; r < 4; ++r) s_out[(wave + 4 * q) * 16 + (lane & 15)][4 * (lane >> 4) + r] = (double)acc[q][r] * sc;
    }
    __syncthreads();
    const int row = tid & 15;
    const int64_t i = (int64_t)p * 16 + row;
    const double dg = diag[i];
    const double ys = yscale ? yscale[i] : 1.0;
    for (int m = 0; m < CW / 16; ++m) {
        const int cl = (tid >> 4) + 16 * m;
        const int c = k0 + cl;
        if (c >= K) continue;
        double y = s_out[cl][row] + dg * X[i + (int64_t)c * ldx];
        for (int s2 = 0; s2 < S; ++s2) y += part[((int64_t)s2 * K + c) * N + i];
        Y[i + (int64_t)c * ldy] = yscale ? ys * y : y;
    }
}
// The full near blocks of RT consecutive block rows against the far field's split operands: one
// workgroup per super-row and 64 NQ columns.  Adjacent panels share most of their near panels, so the
// workgroup walks the union of the rows' block columns and loads each source panel's split operands
// once for all RT rows (a row without a block at that column multiplies zeros); the 32-deep contraction
// takes two union entries at a time, as near_h3_kernel takes two blocks.  The epilogue writes from the
// registers (no LDS): lane (row 4 (lane / 16) + r, column lane % 16) of each 16 x 16 tile.
template <int NQ, int RT>
__global__ __launch_bounds__(256) void near_h3_sr_kernel(int P, int Q, const int *__restrict__ sr_ptr,
                                                         const int *__restrict__ sr_col, const int *__restrict__ sr_blk,
                                                         const _Float16 *__restrict__ vh, const int *__restrict__ rexp,
                                                         int S, const double *__restrict__ part, int64_t N, int K,
                                                         const _Float16 *__restrict__ zh, const _Float16 *__restrict__ zl,
                                                         const unsigned *__restrict__ cmax, int Kp,
                                                         const double *X, int64_t ldx,
                                                         const double *__restrict__ diag,
                                                         const double *__restrict__ yscale, double *Y, int64_t ldy) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int q = xcd_row(blockIdx.x, Q);
    const int k0 = blockIdx.y * 64 * NQ;
    if (q >= Q) return;  // uniform per workgroup
    const int t = lane & 15, h = (lane >> 4) & 1, second = lane >> 5;
    int64_t zoff[NQ];
#pragma unroll
    for (int c = 0; c < NQ; ++c) {
        const int col = k0 + (wave + 4 * c) * 16 + t;
        zoff[c] = zsplit_col(col < K ? col : K - 1);  // columns past K: clamped, never stored
    }
    v4f acc[RT][NQ];
#pragma unroll
    for (int r = 0; r < RT; ++r)
#pragma unroll
        for (int c = 0; c < NQ; ++c) acc[r][c] = v4f{0.f, 0.f, 0.f, 0.f};
    const int u0 = sr_ptr[q], u1 = sr_ptr[q + 1];
    auto load = [&](int u, uint4 (&ah)[RT], uint4 (&al)[RT], uint4 (&bh)[NQ], uint4 (&bl)[NQ]) {
        const int uu = u + second;
        const int uv = uu < u1 ? uu : u;  // an odd count's last pair: the second entry's A is zero
#pragma unroll
        for (int r = 0; r < RT; ++r) {
            const int b = uu < u1 ? sr_blk[(int64_t)uv * RT + r] : -1;
            const _Float16 *a = vh + (int64_t)(b >= 0 ? b : 0) * 512 + 16 * t + 8 * h;
            ah[r] = *reinterpret_cast<const uint4 *>(a);
            al[r] = *reinterpret_cast<const uint4 *>(a + 256);
            if (b < 0) ah[r] = al[r] = uint4{0u, 0u, 0u, 0u};
        }
        const int64_t tb = zsplit_row((int64_t)sr_col[uv] * 16 + 8 * h, Kp);
#pragma unroll
        for (int c = 0; c < NQ; ++c) {
            bh[c] = *reinterpret_cast<const uint4 *>(zh + tb + zoff[c]);
            bl[c] = *reinterpret_cast<const uint4 *>(zl + tb + zoff[c]);
        }
    };
    uint4 nah[RT], nal[RT], nbh[NQ], nbl[NQ];
    if (u0 < u1) load(u0, nah, nal, nbh, nbl);
    for (int u = u0; u < u1; u += 2) {
        v8h ah[RT], al[RT], bh[NQ], bl[NQ];
#pragma unroll
        for (int r = 0; r < RT; ++r) {
            ah[r] = __builtin_bit_cast(v8h, nah[r]);
            al[r] = __builtin_bit_cast(v8h, nal[r]);
        }
#pragma unroll
        for (int c = 0; c < NQ; ++c) {
            bh[c] = __builtin_bit_cast(v8h, nbh[c]);
            bl[c] = __builtin_bit_cast(v8h, nbl[c]);
        }
        if (u + 2 < u1) load(u + 2, nah, nal, nbh, nbl);
#pragma unroll
        for (int r = 0; r < RT; ++r)
#pragma unroll
            for (int c = 0; c < NQ; ++c) {
                acc[r][c] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[r], bh[c], acc[r][c], 0, 0, 0);
                acc[r][c] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[r], bl[c], acc[r][c], 0, 0, 0);
                acc[r][c] = __builtin_amdgcn_mfma_f32_16x16x32_f16(al[r], bh[c], acc[r][c], 0, 0, 0);
            }
    }
#pragma unroll
    for (int c = 0; c < NQ; ++c) {
        const int col = k0 + (wave + 4 * c) * 16 + (lane & 15);
        if (col >= K) continue;
        const int ce = col_exponent(cmax[col]);
#pragma unroll
        for (int r = 0; r < RT; ++r) {
            const int p = q * RT + r;
            if (p >= P) continue;
            const double sc = ldexp(1.0, -(rexp[p] + ce));
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int64_t i = (int64_t)p * 16 + 4 * (lane >> 4) + e;
                double y = (double)acc[r][c][e] * sc + diag[i] * X[i + (int64_t)col * ldx];
                for (int s2 = 0; s2 < S; ++s2) y += part[((int64_t)s2 * K + col) * N + i];
                Y[i + (int64_t)col * ldy] = yscale ? yscale[i] * y : y;
            }
        }
    }
}
}  // namespace

// column blocks of 16 per wave of the near kernels: a workgroup covers 64, 128 or 256 columns
static int near_nq(int K) { return K <= 64 ? 1 : (K <= 128 ? 2 : 4); }
static int near_kpad(int K) {
    const int cw = 64 * near_nq(K);
    return (K + cw - 1) / cw * cw;
}

size_t near_xt_bytes(int64_t N, int K) { return sizeof(float) * (size_t)N * (size_t)near_kpad(K); }

void near_h3_pack(hipStream_t st, const DevBSR &C, const double *w, void *valh, int *rexp) {
    if (C.P < 1) return;
    near_h3_pack_kernel<<<(unsigned)C.P, 256, 0, st>>>(C.row_ptr, C.col, C.valup32, w, static_cast<_Float16 *>(valh),
                                                        rexp);
}

void near_combine(hipStream_t st, const DevBSR &C, NearVals vals, int S, const double *part, int64_t N, int K,
                  const double *X, int64_t ldx, const double *diag, double *Y, int64_t ldy, float *xt,
                  const NearH3 *h3, const double *yscale) {
    const int bd = C.bd;
    const float *v32 = vals == NEAR_UP32 ? C.valup32 : (vals == NEAR_C32 ? C.val32 : nullptr);
    const int nq = near_nq(K), Kpad = near_kpad(K), cw = 64 * nq;
    // (the split-FP16 operands cover Kp >= 16 columns, so any K >= 1 runs here when they are given)
    if (vals == NEAR_UP32 && C.mfma_order && bd == 16 && K >= 1 && h3 && C.valh) {
        if (h3->tile != kZt) throw std::logic_error("near_combine: split-FP16 operand tiles of an unexpected height");
        const int per = (C.P + 7) / 8;
        dim3 grid((unsigned)(8 * per), (unsigned)(Kpad / cw));
        const _Float16 *vh = static_cast<const _Float16 *>(C.valh);
        const _Float16 *zh = static_cast<const _Float16 *>(h3->zh), *zl = static_cast<const _Float16 *>(h3->zl);
        if (C.sr_ptr) {
            const int perq = (C.sr_q + 7) / 8;
            dim3 gq((unsigned)(8 * perq), (unsigned)(Kpad / cw));
#define DBIE_SR(Q) near_h3_sr_kernel<Q, kNearSuperRows><<<gq, 256, 0, st>>>(C.P, C.sr_q, C.sr_ptr, C.sr_col, C.sr_blk, \
                                                                              vh, C.rexp, S, part, N, K, zh, zl, h3->cmax, \
                                                                              h3->Kp, X, ldx, diag, yscale, Y, ldy)
            if (nq == 1) DBIE_SR(1);
            else if (nq == 2) DBIE_SR(2);
            else DBIE_SR(4);
#undef DBIE_SR
            return;
        }
#define DBIE_H3(Q) near_h3_kernel<Q><<<grid, 256, 0, st>>>(C.P, C.row_ptr, C.col, vh, C.rexp, S, part, N, K, zh, zl, \
                                                           h3->cmax, h3->Kp, X, ldx, diag, yscale, Y, ldy)
        if (nq == 1) DBIE_H3(1);
        else if (nq == 2) DBIE_H3(2);
        else DBIE_H3(4);
#undef DBIE_H3
        return;
    }
    if (yscale) throw std::logic_error("near_combine: an output row scale needs the split-FP16 near blocks");
    if (vals == NEAR_UP32 && C.mfma_order && bd == 16 && K >= 16) {  // xt: near_xt_bytes(N, K)
        x_tiles_f32_kernel<<<dim3((unsigned)((N + 255) / 256), (unsigned)Kpad), 256, 0, st>>>(X, ldx, N, K, Kpad, xt);
        const int per = (C.P + 7) / 8;
        dim3 grid((unsigned)(8 * per), (unsigned)(Kpad / cw));
#define DBIE_UP32(Q) near_up32_kernel<Q><<<grid, 256, 0, st>>>(C.P, C.row_ptr, C.col, C.valup32, S, part, N, K, xt, \
                                                               Kpad, X, ldx, diag, Y, ldy)
        if (nq == 1) DBIE_UP32(1);
        else if (nq == 2) DBIE_UP32(2);
        else DBIE_UP32(4);
#undef DBIE_UP32
        return;
    }
    if (C.mfma_order && bd == 16 && K >= 16) {
        const int per = (C.P + 7) / 8;
        dim3 grid((unsigned)(8 * per), (unsigned)((K + cw - 1) / cw));
#define DBIE_NM(Q) near_mfma_kernel<double, Q><<<grid, 256, 0, st>>>(C.P, C.row_ptr, C.col, C.val64, S, part, N, K, X, \
                                                                      ldx, diag, Y, ldy)
        if (nq == 1) DBIE_NM(1);
        else if (nq == 2) DBIE_NM(2);
        else DBIE_NM(4);
#undef DBIE_NM
        return;
    }
    if (!C.mfma_order && K >= 16 && (bd == 4 || bd == 9 || bd == 25)) {
        const int per = (C.P + 7) / 8;
        dim3 grid((unsigned)(8 * per), (unsigned)((K + 255) / 256));
#define DBIE_NEAR_RM(RT, KS)                                                                                      \
    do {                                                                                                          \
        if (v32)                                                                                                  \
            near_mfma_rm_kernel<float, RT, KS><<<grid, 256, 0, st>>>(C.P, bd, C.row_ptr, C.col, v32, S, part, N, K, \
                                                                     X, ldx, diag, Y, ldy);                       \
        else                                                                                                      \
            near_mfma_rm_kernel<double, RT, KS><<<grid, 256, 0, st>>>(C.P, bd, C.row_ptr, C.col, C.val64, S, part, \
                                                                      N, K, X, ldx, diag, Y, ldy);                \
    } while (0)
        if (bd == 4) DBIE_NEAR_RM(1, 1);
        else if (bd == 9) DBIE_NEAR_RM(1, 3);
        else DBIE_NEAR_RM(2, 7);
#undef DBIE_NEAR_RM
        return;
    }
    // CS threads per (row, column) output (up to 4, within a 64-thread workgroup for blocks of at
    // most 64 rows), 64 / (bd CS) block rows times columns per workgroup, so that a few columns
    // still spread over every CU with short per-thread loops
    const int nt = bd <= 64 ? 64 : 256;
    int CS = 1;
    while (CS < 4 && bd * CS * 2 <= nt) CS *= 2;
    const int KT = std::max(1, nt / (bd * CS));
    const int KTe = std::min(K, KT);
    const int RP = std::max(1, KT / KTe);
    dim3 grid((unsigned)((C.P + RP - 1) / RP), (unsigned)((K + KTe - 1) / KTe));
    const int mo = (vals == NEAR_UP32 && C.mfma_order) ? 2 : (C.mfma_order ? 1 : 0);
    if (v32)
        near_combine_kernel<float><<<grid, nt, 0, st>>>(C.P, bd, mo, CS, C.row_ptr, C.col, v32, S, part, N, K, KTe,
                                                         RP, X, ldx, diag, Y, ldy);
    else
        near_combine_kernel<double><<<grid, nt, 0, st>>>(C.P, bd, mo, CS, C.row_ptr, C.col, C.val64, S, part, N, K,
                                                          KTe, RP, X, ldx, diag, Y, ldy);
}

}  // namespace dbie
