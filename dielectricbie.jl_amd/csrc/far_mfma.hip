// far_mfma.hip -- far-field Dᵀ x RHS-block contraction on the CDNA4 matrix cores.
//
// Y[i,k] = sum_j G(i,j) Z[j,k],  G(i,j) = n_i.(x_i - y_j) / r^3 (self pair 0),  Z = w .* X.
// The kernel matrix G is never stored in HBM.  A workgroup of 8 waves owns 128 targets
// (eight 16-target slices) and all columns of its column group.  Per source stage:
//   * every wave contracts the target slices it owns with its own 16-column blocks, reading
//     the generated tile of stage s from LDS (laid out exactly as the MFMA A operand) and the
//     B operand (w.*X) from a tile-major copy in L2, prefetched one stage ahead;
//   * interleaved with those MFMAs, wave w generates on VALU the entries of stage s+1 for
//     target slice w and stores them into the next LDS tile;
//   * one barrier per stage.  Tiles are triple-buffered so that a wave writing stage s+1
//     never overwrites a tile a slower wave still reads.
//
//   FP64: v_mfma_f64_16x16x4_f64, stage = 16 sources (4 k-slices of 4)
//   BF16: v_mfma_f32_16x16x32_bf16, stage = 64 sources (2 k-slices of 32), FP32 generation,
//         FP32 accumulation; the BF16X3 form adds hi*lo + lo*hi of a two-term bf16 split.
// Operand layouts on gfx950 (measured with a one-hot probe): A lane l = (row l%16,
// k-group l/16); B lane l = (k-group l/16, col l%16); D col = l%16, row = 4*r + l/16 for
// the FP64 op and 4*(l/16) + r for the BF16 op.
//
// Source tiles whose nodes all share the target block's plane are skipped (exact zeros).
// Each source split (gridDim.z) writes its own FP64 partial sum.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <type_traits>

#include <algorithm>

#include "devmath.hpp"

#include "kernels.hpp"

namespace dbie {

namespace {
typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef __bf16 v2bf __attribute__((ext_vector_type(2)));
typedef float v2f __attribute__((ext_vector_type(2)));

constexpr double kInv4Pi = 1.0 / (4.0 * 3.14159265358979311599796346854);
constexpr int kWaves = 8;
constexpr int kThreads = kWaves * 64;
constexpr int kTargets = kWaves * 16;  // 128 targets per workgroup

// Work items per stage are (target slice, 16-column block).  With NCB >= 8 blocks each
// wave owns NCB/8 column blocks and all 8 slices; with NCB < 8 each column block is
// shared by 8/NCB waves, each owning NCB of the slices.
template <int NCB>
struct Items {
    static constexpr int kCbPerWave = NCB >= kWaves ? NCB / kWaves : 1;
    static constexpr int kSlicesPerWave = NCB >= kWaves ? kWaves : NCB;
    static constexpr int kN = kCbPerWave * kSlicesPerWave;
    __device__ static int cb(int wave, int q) { return NCB >= kWaves ? wave * kCbPerWave + q : wave % NCB; }
    __device__ static int slice(int wave, int s) { return NCB >= kWaves ? s : (wave / NCB) + s * (kWaves / NCB); }
};

// ---------------------------------------------------------------- FP64
constexpr int kS64 = 16;  // sources per stage

// Tile-major w.*X: Zt[((j/16)*Kp + col)*16 + j%16]
__global__ void scale_tiles_f64(const double *__restrict__ X, int64_t ldx, const double *__restrict__ w, int64_t N,
                                int K, int Kp, double *__restrict__ Zt) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y;
    const int64_t ntile = (N + kS64 - 1) / kS64;
    if (j >= ntile * kS64) return;
    const double z = (j < N && k < K) ? w[j] * X[j + (int64_t)k * ldx] : 0.0;
    Zt[((j / kS64) * Kp + k) * kS64 + (j % kS64)] = z;
}

// Stage sequence of a source split: stages start every kS sources, same-plane 256-node
// tiles are skipped (their entries are exact zeros).
struct StageWalk {
    int64_t jend;
    int tplane;
    const int *tile_plane;
    __device__ int64_t next(int64_t jb) const {
        while (jb < jend && tplane >= 0 && tile_plane[jb / 256] == tplane) jb += 256 - (jb % 256);
        return jb;
    }
};

// Padding sources (past the split) are parked far away: their kernel entries underflow
// to zero, so no per-entry validity select is needed.  The self pair (r2 = 0) has
// n.(x - y) = 0 and r2 is clamped before the reciprocal square root.
constexpr double kFar64 = 1.0e150;
constexpr float kFar32 = 1.0e18f;

template <int NCB>
__global__ __launch_bounds__(kThreads) void far_mfma_f64_kernel(DevGeom g, int K, int Kp, const double *__restrict__ Zt,
                                                                double *__restrict__ part, int src_per_split) {
    using I = Items<NCB>;
    constexpr int KS = kS64 / 4;
    __shared__ double s_g[3][kWaves][KS][64];  // [stage % 3][slice][k-slice][lane]
    __shared__ double4 s_pos[2][kS64];         // [stage % 2][source]
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int64_t N = g.N;
    const int64_t i0 = (int64_t)blockIdx.x * kTargets;
    const int k0 = blockIdx.y * NCB * 16;
    const int64_t jb0 = (int64_t)blockIdx.z * src_per_split;
    const StageWalk walk{min(N, jb0 + src_per_split), g.tile_plane ? g.tile_plane[i0 / 256] : -1, g.tile_plane};

    int64_t tg = i0 + wave * 16 + (lane & 15);
    tg = tg < N ? tg : N - 1;
    const double tx = g.x[tg], ty = g.y[tg], tz = g.z[tg];
    const double tn0 = g.nx[tg], tn1 = g.ny[tg], tn2 = g.nz[tg];

    v4d acc[I::kN];
#pragma unroll
    for (int q = 0; q < I::kN; ++q) acc[q] = v4d{0.0, 0.0, 0.0, 0.0};
    double bcur[I::kCbPerWave][KS], bnxt[I::kCbPerWave][KS];
    auto load_b = [&](int64_t jb, double (&b)[I::kCbPerWave][KS]) {
        const double *zt = Zt + ((jb / kS64) * Kp + k0) * kS64;
#pragma unroll
        for (int q = 0; q < I::kCbPerWave; ++q)
#pragma unroll
            for (int ks = 0; ks < KS; ++ks)
                b[q][ks] = zt[(I::cb(wave, q) * 16 + (lane & 15)) * kS64 + ks * 4 + (lane >> 4)];
    };
    auto load_pos = [&](int64_t jb, int pb) {
        if (tid < kS64) {
            const int64_t j = jb + tid;
            s_pos[pb][tid] = j < walk.jend ? make_double4(g.x[j], g.y[j], g.z[j], 0.0)
                                           : make_double4(kFar64, kFar64, kFar64, 0.0);
        }
    };
    auto gen = [&](int pb, int ks) {  // entry (slice target, source ks*4 + lane/16)
        const double4 sp = s_pos[pb][ks * 4 + (lane >> 4)];
        const double dx = tx - sp.x, dy = ty - sp.y, dz = tz - sp.z;
        const double r2 = fmax(dx * dx + dy * dy + dz * dz, 1e-200);
        const double rinv = rsqrt(r2);
        return (tn0 * dx + tn1 * dy + tn2 * dz) * (rinv * rinv * rinv);
    };

    // prologue: stage 0 generated, stage 1 positions staged, B(0) loaded (skipped, uniformly
    // per workgroup, when the whole split is same-plane; the zero partial is still written)
    int64_t jc = walk.next(jb0);
    int64_t j1 = walk.jend;
    if (jc < walk.jend) {
        j1 = walk.next(jc + kS64);
        load_pos(jc, 0);
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) s_g[0][wave][ks][lane] = gen(0, ks);
        if (j1 < walk.jend) load_pos(j1, 1);
        load_b(jc, bcur);
        __syncthreads();
    }
    int sg = 0, pb = 1;
    while (jc < walk.jend) {
        const bool more = j1 < walk.jend;
        const int64_t j2 = more ? walk.next(j1 + kS64) : walk.jend;
        // branch-free body: without a next stage the generation below writes a tile that is
        // never read and the B prefetch re-reads the current (valid) stage
        load_b(more ? j1 : jc, bnxt);
        double gnext[KS];
        // MFMAs of stage jc interleaved with generation of stage j1
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) {
            gnext[ks] = gen(pb, ks);
#pragma unroll
            for (int s = 0; s < I::kSlicesPerWave; ++s) {
                const double a = s_g[sg][I::slice(wave, s)][ks][lane];
#pragma unroll
                for (int q = 0; q < I::kCbPerWave; ++q)
                    acc[q * I::kSlicesPerWave + s] =
                        __builtin_amdgcn_mfma_f64_16x16x4f64(a, bcur[q][ks], acc[q * I::kSlicesPerWave + s], 0, 0, 0);
            }
        }
        // ask the scheduler to co-issue: one MFMA, then two VALU, repeated
#pragma unroll
        for (int m = 0; m < KS * I::kN; ++m) {
            __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
            __builtin_amdgcn_sched_group_barrier(0x002, 2, 0);
        }
        const int sn = sg == 2 ? 0 : sg + 1;
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) s_g[sn][wave][ks][lane] = gnext[ks];
        if (j2 < walk.jend) load_pos(j2, pb ^ 1);
        __syncthreads();
#pragma unroll
        for (int q = 0; q < I::kCbPerWave; ++q)
#pragma unroll
            for (int ks = 0; ks < KS; ++ks) bcur[q][ks] = bnxt[q][ks];
        sg = sn;
        pb ^= 1;
        jc = j1;
        j1 = j2;
    }
    // D: col = lane%16, row = 4*r + lane/16
#pragma unroll
    for (int q = 0; q < I::kCbPerWave; ++q)
#pragma unroll
        for (int s = 0; s < I::kSlicesPerWave; ++s) {
            const int col = k0 + I::cb(wave, q) * 16 + (lane & 15);
            if (col >= K) continue;
            double *dst = part + ((int64_t)blockIdx.z * K + col) * N;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int64_t i = i0 + I::slice(wave, s) * 16 + 4 * r + (lane >> 4);
                if (i < N) dst[i] = acc[q * I::kSlicesPerWave + s][r] * kInv4Pi;
            }
        }
}

// ---------------------------------------------------------------- BF16
constexpr int kSb = 64;  // sources per stage (two k-slices of 32)

// w.*X split into bf16 hi + lo, tile-major: [j/64][col][j%64], Kp columns.
__global__ void split_tiles_bf16(const double *__restrict__ X, int64_t ldx, const double *__restrict__ w, int64_t N,
                                 int K, int Kp, __bf16 *__restrict__ zh, __bf16 *__restrict__ zl) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y;
    const int64_t ntile = (N + kSb - 1) / kSb;
    if (j >= ntile * kSb) return;
    __bf16 h = (__bf16)0.0f, l = (__bf16)0.0f;
    if (j < N && k < K) {
        const double zd = w[j] * X[j + (int64_t)k * ldx];
        h = (__bf16)(float)zd;
        l = (__bf16)(float)(zd - (double)(float)h);
    }
    const int64_t idx = ((j / kSb) * Kp + k) * kSb + (j % kSb);
    zh[idx] = h;
    if (zl) zl[idx] = l;
}

template <int NCB, bool X3>
__global__ __launch_bounds__(kThreads) void far_mfma_bf16_kernel(DevGeom g, int K, int Kp, const __bf16 *__restrict__ zh,
                                                                 const __bf16 *__restrict__ zl, double *__restrict__ part,
                                                                 int src_per_split) {
    using I = Items<NCB>;
    constexpr int KS = kSb / 32;
    constexpr int NG = X3 ? 2 : 1;                 // hi (and lo) planes of the generated tile
    __shared__ uint4 s_g[3][NG][kWaves][KS][64];    // [stage % 3][plane][slice][k-slice][lane]
    __shared__ float4 s_pos[2][kSb];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int64_t N = g.N;
    const int64_t i0 = (int64_t)blockIdx.x * kTargets;
    const int k0 = blockIdx.y * NCB * 16;
    const int64_t jb0 = (int64_t)blockIdx.z * src_per_split;
    const StageWalk walk{min(N, jb0 + src_per_split), g.tile_plane ? g.tile_plane[i0 / 256] : -1, g.tile_plane};

    int64_t tg = i0 + wave * 16 + (lane & 15);
    tg = tg < N ? tg : N - 1;
    const float tx = g.xf[tg], ty = g.yf[tg], tz = g.zf[tg];
    const float tn0 = g.nxf[tg], tn1 = g.nyf[tg], tn2 = g.nzf[tg];

    v4f acc[I::kN];
#pragma unroll
    for (int q = 0; q < I::kN; ++q) acc[q] = v4f{0.f, 0.f, 0.f, 0.f};
    constexpr int NL = X3 ? I::kCbPerWave : 1;
    uint4 bh_cur[I::kCbPerWave][KS], bh_nxt[I::kCbPerWave][KS];
    uint4 bl_cur[NL][KS], bl_nxt[NL][KS];
    auto load_b = [&](int64_t jb, uint4 (&bh)[I::kCbPerWave][KS], uint4 (&bl)[NL][KS]) {
        const int64_t base = ((jb / kSb) * Kp + k0) * kSb;
#pragma unroll
        for (int q = 0; q < I::kCbPerWave; ++q)
#pragma unroll
            for (int ks = 0; ks < KS; ++ks) {
                const int64_t off = base + (int64_t)(I::cb(wave, q) * 16 + (lane & 15)) * kSb + ks * 32 + 8 * (lane >> 4);
                bh[q][ks] = *reinterpret_cast<const uint4 *>(zh + off);
                if (X3) bl[q < NL ? q : 0][ks] = *reinterpret_cast<const uint4 *>(zl + off);
            }
    };
    auto load_pos = [&](int64_t jb, int pb) {
        if (tid < kSb) {
            const int64_t j = jb + tid;
            s_pos[pb][tid] = j < walk.jend ? make_float4(g.xf[j], g.yf[j], g.zf[j], 0.0f)
                                           : make_float4(kFar32, kFar32, kFar32, 0.0f);
        }
    };
    // one pair of entries (sources e, e+1 of k-slice ks) -> packed bf16 hi (and lo)
    auto gen2 = [&](int pb, int ks, int e, unsigned &hi, unsigned &lo) {
        float v[2];
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            const float4 sp = s_pos[pb][ks * 32 + 8 * (lane >> 4) + e + u];
            const float dx = tx - sp.x, dy = ty - sp.y, dz = tz - sp.z;
            const float r2 = fmaxf(dx * dx + dy * dy + dz * dz, 1e-20f);
            const float rinv = __builtin_amdgcn_rsqf(r2);
            v[u] = (tn0 * dx + tn1 * dy + tn2 * dz) * (rinv * rinv * rinv);
        }
        const v2bf h = __builtin_convertvector(v2f{v[0], v[1]}, v2bf);
        hi = __builtin_bit_cast(unsigned, h);
        if (X3) {
            const v2f hf = __builtin_convertvector(h, v2f);
            const v2bf l = __builtin_convertvector(v2f{v[0] - hf[0], v[1] - hf[1]}, v2bf);
            lo = __builtin_bit_cast(unsigned, l);
        }
    };
    auto gen_slice = [&](int pb, int ks, uint4 &h4, uint4 &l4) {
        unsigned h[4], l[4] = {0u, 0u, 0u, 0u};
#pragma unroll
        for (int e = 0; e < 4; ++e) gen2(pb, ks, 2 * e, h[e], l[e]);
        h4 = uint4{h[0], h[1], h[2], h[3]};
        l4 = uint4{l[0], l[1], l[2], l[3]};
    };

    int64_t jc = walk.next(jb0);
    int64_t j1 = walk.jend;
    if (jc < walk.jend) {
        j1 = walk.next(jc + kSb);
        load_pos(jc, 0);
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) {
            uint4 h4, l4;
            gen_slice(0, ks, h4, l4);
            s_g[0][0][wave][ks][lane] = h4;
            if (X3) s_g[0][NG - 1][wave][ks][lane] = l4;
        }
        if (j1 < walk.jend) load_pos(j1, 1);
        load_b(jc, bh_cur, bl_cur);
        __syncthreads();
    }
    int sg = 0, pb = 1;
    while (jc < walk.jend) {
        const bool more = j1 < walk.jend;
        const int64_t j2 = more ? walk.next(j1 + kSb) : walk.jend;
        // branch-free body: without a next stage the generation below writes a tile that is
        // never read and the B prefetch re-reads the current (valid) stage
        load_b(more ? j1 : jc, bh_nxt, bl_nxt);
        uint4 gh[KS], gl[KS];
        // MFMAs of stage jc interleaved with generation of stage j1
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) {
            gen_slice(pb, ks, gh[ks], gl[ks]);
#pragma unroll
            for (int s = 0; s < I::kSlicesPerWave; ++s) {
                const int sl = I::slice(wave, s);
                const v8bf ah = __builtin_bit_cast(v8bf, s_g[sg][0][sl][ks][lane]);
                const v8bf al = __builtin_bit_cast(v8bf, s_g[sg][NG - 1][sl][ks][lane]);
#pragma unroll
                for (int q = 0; q < I::kCbPerWave; ++q) {
                    v4f &c = acc[q * I::kSlicesPerWave + s];
                    const v8bf bh = __builtin_bit_cast(v8bf, bh_cur[q][ks]);
                    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh, c, 0, 0, 0);
                    if (X3) {
                        const v8bf bl = __builtin_bit_cast(v8bf, bl_cur[q < NL ? q : 0][ks]);
                        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl, c, 0, 0, 0);
                        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh, c, 0, 0, 0);
                    }
                }
            }
        }
        // ask the scheduler to co-issue: one MFMA, then a few VALU, repeated
#pragma unroll
        for (int m = 0; m < KS * I::kN * (X3 ? 3 : 1); ++m) {
            __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
            __builtin_amdgcn_sched_group_barrier(0x002, X3 ? 3 : 8, 0);
        }
        const int sn = sg == 2 ? 0 : sg + 1;
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) {
            s_g[sn][0][wave][ks][lane] = gh[ks];
            if (X3) s_g[sn][NG - 1][wave][ks][lane] = gl[ks];
        }
        if (j2 < walk.jend) load_pos(j2, pb ^ 1);
        __syncthreads();
#pragma unroll
        for (int q = 0; q < I::kCbPerWave; ++q)
#pragma unroll
            for (int ks = 0; ks < KS; ++ks) {
                bh_cur[q][ks] = bh_nxt[q][ks];
                if (X3) bl_cur[q < NL ? q : 0][ks] = bl_nxt[q < NL ? q : 0][ks];
            }
        sg = sn;
        pb ^= 1;
        jc = j1;
        j1 = j2;
    }
    // D: col = lane%16, row = 4*(lane/16) + r
#pragma unroll
    for (int q = 0; q < I::kCbPerWave; ++q)
#pragma unroll
        for (int s = 0; s < I::kSlicesPerWave; ++s) {
            const int col = k0 + I::cb(wave, q) * 16 + (lane & 15);
            if (col >= K) continue;
            double *dst = part + ((int64_t)blockIdx.z * K + col) * N;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int64_t i = i0 + I::slice(wave, s) * 16 + 4 * (lane >> 4) + r;
                if (i < N) dst[i] = (double)acc[q * I::kSlicesPerWave + s][r] * kInv4Pi;
            }
        }
}

// ---------------------------------------------------------------- FP16 x3
// Two-term fp16 split of both operands (11-bit mantissas, ~22 bits together) with
// hi*hi + hi*lo + lo*hi products accumulated in FP32: about 2e-7 per entry.  Coordinates are
// taken relative to the tile's first target in FP64 and then rounded to FP32.  Range: entries are scaled by 2^-eG (|G| <= 1/r_min^2 <= 2^14 after scaling)
// and each column of w.*X by 2^e_k (max |.| in [2^13, 2^14)); both are undone on output.
typedef _Float16 v8h __attribute__((ext_vector_type(8)));
typedef unsigned v4u __attribute__((ext_vector_type(4)));
typedef _Float16 v2h __attribute__((ext_vector_type(2)));

__global__ void zero_u32_kernel(unsigned *__restrict__ p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

// per-column max |w_j X_jk| as float bits (non-negative floats order like unsigned ints): a
// 2-D grid of 4096-row chunks x columns, each workgroup folding its chunk into cmax[k] with one
// atomic max (cmax cleared beforehand)
constexpr int kColmaxRows = 4096;
__global__ __launch_bounds__(256) void colmax_kernel(const double *__restrict__ X, int64_t ldx,
                                                     const double *__restrict__ w, int64_t N,
                                                     unsigned *__restrict__ cmax) {
    __shared__ float s_m[4];
    const int k = blockIdx.y;
    const double *xk = X + (int64_t)k * ldx;
    const int64_t j0 = (int64_t)blockIdx.x * kColmaxRows;
    const int64_t j1 = min(N, j0 + kColmaxRows);
    float m = 0.0f;
    for (int64_t j = j0 + threadIdx.x; j < j1; j += 256) m = fmaxf(m, fabsf((float)(w[j] * xk[j])));
    for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_down(m, off, 64));
    if ((threadIdx.x & 63) == 0) s_m[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) atomicMax(cmax + k, __float_as_uint(fmaxf(fmaxf(s_m[0], s_m[1]), fmaxf(s_m[2], s_m[3]))));
}

__global__ void split_tiles_f16(const double *__restrict__ X, int64_t ldx, const double *__restrict__ w, int64_t N,
                                int K, int Kp, const unsigned *__restrict__ cmax, _Float16 *__restrict__ zh,
                                _Float16 *__restrict__ zl) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y;
    if (j >= (N + kSb - 1) / kSb * kSb) return;  // the operands cover whole 64-source stages
    _Float16 h = (_Float16)0.0f, l = (_Float16)0.0f;
    if (j < N && k < K) {
        const double zd = ldexp(w[j] * X[j + (int64_t)k * ldx], col_exponent(cmax[k]));
        h = (_Float16)(float)zd;
        l = (_Float16)(float)(zd - (double)(float)h);
    }
    const int64_t idx = zsplit_idx(j, k, Kp);
    zh[idx] = h;
    zl[idx] = l;
}

// Near-pair exclusion (EXCL > 0): pairs of panels (target panel, source panel) flagged in
// nmask (one bit per pair, rows of mwords 32-bit words) are left out of the sum; the caller adds
// their accurate interaction (upsampled/adaptive blocks) in the near-field kernel.  The low
// precision then never meets the large direct entries of close pairs, whose cancellation
// against the corrections would amplify its rounding.  EXCL = 1: npp = 16 (a lane's 8 sources in
// a k-slice lie in one panel); EXCL = 2: any npp (per-source test).
struct NearMask {
    const uint32_t *bits = nullptr;
    int npp = 0;
    int mwords = 0;
};

// Lane n of each 16-lane row broadcast to the whole row (DPP row_newbcast).
template <int N>
__device__ inline float row_bcast_c(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), 0x150 + N, 0xF, 0xF, true));
}
__device__ inline float row_bcast(float v, int n) {  // n is a constant after unrolling
    switch (n) {
    case 0: return row_bcast_c<0>(v);
    case 1: return row_bcast_c<1>(v);
    case 2: return row_bcast_c<2>(v);
    case 3: return row_bcast_c<3>(v);
    case 4: return row_bcast_c<4>(v);
    case 5: return row_bcast_c<5>(v);
    case 6: return row_bcast_c<6>(v);
    case 7: return row_bcast_c<7>(v);
    case 8: return row_bcast_c<8>(v);
    case 9: return row_bcast_c<9>(v);
    case 10: return row_bcast_c<10>(v);
    case 11: return row_bcast_c<11>(v);
    case 12: return row_bcast_c<12>(v);
    case 13: return row_bcast_c<13>(v);
    case 14: return row_bcast_c<14>(v);
    default: return row_bcast_c<15>(v);
    }
}

template <int NCB, int EXCL>
__global__ __launch_bounds__(kThreads) void far_mfma_f16x3_kernel(DevGeom g, int K, int Kp, const _Float16 *__restrict__ zh,
                                                                  const _Float16 *__restrict__ zl,
                                                                  const unsigned *__restrict__ cmax, float gscale,
                                                                  double *__restrict__ part, int src_per_split,
                                                                  NearMask nm) {
    using I = Items<NCB>;
    constexpr int KS = kSb / 32;
    __shared__ uint4 s_g[2][2][kWaves][KS][64];  // [stage parity][hi/lo][slice][k-slice][lane]
    // source coordinates relative to the tile origin, staged in LDS.  Unless EXCL = 2, lane
    // 16 g + 8 ks + e of each wave reads only source 32 ks + 8 g + e of the stage and the lanes of
    // row g take the others by a row broadcast (one LDS read per lane and stage)
    constexpr bool kRegPos = EXCL != 2;
    __shared__ float4 s_pos[2][kSb];
    __shared__ int s_rel[EXCL == 2 ? 2 : 1][kSb];  // source panel - first panel of the stage
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int64_t N = g.N;
    const int64_t i0 = (int64_t)blockIdx.x * kTargets;
    const int k0 = blockIdx.y * NCB * 16;
    const int64_t jb0 = (int64_t)blockIdx.z * src_per_split;
    const StageWalk walk{min(N, jb0 + src_per_split), g.tile_plane ? g.tile_plane[i0 / 256] : -1, g.tile_plane};

    int64_t tg = i0 + wave * 16 + (lane & 15);
    tg = tg < N ? tg : N - 1;
    // coordinates relative to the first target of the tile, rounded to FP32 after the FP64
    // subtraction: the difference of two such values carries an error of about
    // eps_32 (|d| + 2 * tile extent), small against |d| for pairs outside the near list
    const int64_t io = i0 < N ? i0 : N - 1;
    const double ox = g.x[io], oy = g.y[io], oz = g.z[io];
    const float tx = (float)(g.x[tg] - ox), ty = (float)(g.y[tg] - oy), tz = (float)(g.z[tg] - oz);
    const float tn0 = g.nxf[tg] * gscale, tn1 = g.nyf[tg] * gscale, tn2 = g.nzf[tg] * gscale;
    const uint32_t *mrow = EXCL ? nm.bits + (tg / nm.npp) * (int64_t)nm.mwords : nullptr;
    // 64 mask bits starting at the first source panel of the stage at js
    auto load_win = [&](int64_t js) -> uint64_t {
        if (!EXCL || js >= walk.jend) return 0;
        const int64_t sp0 = js / nm.npp;
        const int64_t w0 = sp0 >> 5;
        const int sh = (int)(sp0 & 31);
        const uint64_t lo = (uint64_t)mrow[w0] | ((uint64_t)mrow[w0 + 1] << 32);
        const uint64_t hi = mrow[w0 + 2];
        return sh ? ((lo >> sh) | (hi << (64 - sh))) : lo;
    };

    v4f acc[I::kN];
#pragma unroll
    for (int q = 0; q < I::kN; ++q) acc[q] = v4f{0.f, 0.f, 0.f, 0.f};
    // B operands of the current and the next stage; the loop is unrolled by two so that the two
    // halves swap roles without register copies (a copy would wait for the loads in flight)
    uint4 bh[2][I::kCbPerWave][KS], bl[2][I::kCbPerWave][KS];
    auto load_b = [&](int64_t jb, uint4 (&bh)[I::kCbPerWave][KS], uint4 (&bl)[I::kCbPerWave][KS]) {
#pragma unroll
        for (int q = 0; q < I::kCbPerWave; ++q)
#pragma unroll
            for (int ks = 0; ks < KS; ++ks) {
                // k-slice ks of the stage is one operand tile: its chunk of column block cb, lane l at 16 l
                const int64_t off = ((jb / kZt + ks) * Kp + k0 + I::cb(wave, q) * 16) * kZt + lane * 8;
                bh[q][ks] = *reinterpret_cast<const uint4 *>(zh + off);
                bl[q][ks] = *reinterpret_cast<const uint4 *>(zl + off);
            }
    };
    const int my_src = 32 * ((lane >> 3) & 1) + 8 * (lane >> 4) + (lane & 7);
    float px = 0.f, py = 0.f, pz = 0.f;  // relative FP32 coordinates of my_src in the generated stage
    auto take_pos = [&](int pb) {
        const float4 sp = s_pos[pb][my_src];
        px = sp.x;
        py = sp.y;
        pz = sp.z;
    };
    auto load_pos = [&](int64_t jb, int pb) {
        if (tid < kSb) {
            const int64_t j = jb + tid;
            const bool in = j < walk.jend;
            s_pos[pb][tid] = in ? make_float4((float)(g.x[j] - ox), (float)(g.y[j] - oy), (float)(g.z[j] - oz), 0.f)
                                : make_float4(1.0e18f, 1.0e18f, 1.0e18f, 0.f);
            if (EXCL == 2) s_rel[pb][tid] = in ? (int)(j / nm.npp - jb / nm.npp) : 0;
        }
    };
    // excl: bit mask of this lane's excluded sources in k-slice ks of the stage in buffer pb
    auto gen2 = [&](int pb, int ks, int e, uint64_t win, float n0, float n1, float n2, unsigned &hi, unsigned &lo) {
        float v[2];
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            const int sidx = ks * 32 + 8 * (lane >> 4) + e + u;
            float sx, sy, sz;
            if (kRegPos) {
                sx = row_bcast(px, 8 * ks + e + u);
                sy = row_bcast(py, 8 * ks + e + u);
                sz = row_bcast(pz, 8 * ks + e + u);
            } else {
                const float4 sp = s_pos[pb][sidx];
                sx = sp.x;
                sy = sp.y;
                sz = sp.z;
            }
            const float dx = tx - sx, dy = ty - sy, dz = tz - sz;
            // + 1e-20 keeps the self pair finite (d = 0: rinv^3 = 1e30, entry 0 * 1e30 = 0) at the
            // cost of no instruction: it folds into the first fused multiply-add
            float r2 = fmaf(dz, dz, fmaf(dy, dy, fmaf(dx, dx, 1e-20f)));
            if (EXCL == 2) r2 = ((win >> (s_rel[pb][sidx] & 63)) & 1) ? 1.0e30f : r2;
            const float rinv = __builtin_amdgcn_rsqf(r2);
            v[u] = (n0 * dx + n1 * dy + n2 * dz) * (rinv * rinv * rinv);
        }
        const v2h h = __builtin_convertvector(v2f{v[0], v[1]}, v2h);
        const v2f hf = __builtin_convertvector(h, v2f);
        const v2h l = __builtin_convertvector(v2f{v[0] - hf[0], v[1] - hf[1]}, v2h);
        hi = __builtin_bit_cast(unsigned, h);
        lo = __builtin_bit_cast(unsigned, l);
    };
    auto gen_slice = [&](int pb, int ks, uint64_t win, uint4 &h4, uint4 &l4) {
        unsigned h[4], l[4];
        float n0 = tn0, n1 = tn1, n2 = tn2;
        if (EXCL == 1) {  // npp = 16: the 8 sources lie in panel 2 ks + (lane >> 4) / 2 of the stage;
                          // a zero normal makes their entries exactly zero
            const bool ex = (win >> (2 * ks + ((lane >> 4) >> 1))) & 1;
            n0 = ex ? 0.f : n0;
            n1 = ex ? 0.f : n1;
            n2 = ex ? 0.f : n2;
        }
#pragma unroll
        for (int e = 0; e < 4; ++e) gen2(pb, ks, 2 * e, win, n0, n1, n2, h[e], l[e]);
        h4 = uint4{h[0], h[1], h[2], h[3]};
        l4 = uint4{l[0], l[1], l[2], l[3]};
    };

    int64_t jc = walk.next(jb0);
    int64_t j1 = walk.jend;
    uint64_t win[2] = {0, 0};  // exclusion window of the stage generated in an iteration of parity C
    if (jc < walk.jend) {
        j1 = walk.next(jc + kSb);
        load_pos(jc, 0);
        const uint64_t win0 = load_win(jc);
        __syncthreads();
        if (kRegPos) take_pos(0);
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) {
            uint4 h4, l4;
            gen_slice(0, ks, win0, h4, l4);
            s_g[0][0][wave][ks][lane] = h4;
            s_g[0][1][wave][ks][lane] = l4;
        }
        if (j1 < walk.jend) load_pos(j1, 1);
        win[0] = load_win(j1);
        load_b(jc, bh[0], bl[0]);
        __syncthreads();
    }
    // one stage: MFMAs on stage jc (LDS tile s_g[C], B operands bh/bl[C]) while stage j1 is generated
    // into s_g[1 - C]; loads for the stage after are issued into the other parity
    auto step = [&](auto cc) {
        constexpr int C = decltype(cc)::value;
        const bool more = j1 < walk.jend;
        const int64_t j2 = more ? walk.next(j1 + kSb) : walk.jend;
        load_b(more ? j1 : jc, bh[1 - C], bl[1 - C]);
        win[1 - C] = load_win(j2);
        if (kRegPos) take_pos(1 - C);
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) {
            uint4 gh, gl;
            gen_slice(1 - C, ks, win[C], gh, gl);
            // s_g[1 - C] is not read in this stage: store the generated slice at once
            s_g[1 - C][0][wave][ks][lane] = gh;
            s_g[1 - C][1][wave][ks][lane] = gl;
#pragma unroll
            for (int s = 0; s < I::kSlicesPerWave; ++s) {
                const int sl = I::slice(wave, s);
                const v8h ah = __builtin_bit_cast(v8h, s_g[C][0][sl][ks][lane]);
                const v8h al = __builtin_bit_cast(v8h, s_g[C][1][sl][ks][lane]);
                // the three products of one accumulator are issued kCbPerWave apart, so that
                // consecutive MFMAs never depend on each other
#pragma unroll
                for (int q = 0; q < I::kCbPerWave; ++q)
                    acc[q * I::kSlicesPerWave + s] = __builtin_amdgcn_mfma_f32_16x16x32_f16(
                        ah, __builtin_bit_cast(v8h, bh[C][q][ks]), acc[q * I::kSlicesPerWave + s], 0, 0, 0);
#pragma unroll
                for (int q = 0; q < I::kCbPerWave; ++q)
                    acc[q * I::kSlicesPerWave + s] = __builtin_amdgcn_mfma_f32_16x16x32_f16(
                        ah, __builtin_bit_cast(v8h, bl[C][q][ks]), acc[q * I::kSlicesPerWave + s], 0, 0, 0);
#pragma unroll
                for (int q = 0; q < I::kCbPerWave; ++q)
                    acc[q * I::kSlicesPerWave + s] = __builtin_amdgcn_mfma_f32_16x16x32_f16(
                        al, __builtin_bit_cast(v8h, bh[C][q][ks]), acc[q * I::kSlicesPerWave + s], 0, 0, 0);
            }
        }
#pragma unroll
        for (int m = 0; m < KS * I::kN * 3; ++m) {
            __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
            __builtin_amdgcn_sched_group_barrier(0x002, 3, 0);
        }
        if (j2 < walk.jend) load_pos(j2, C);
        __syncthreads();
        jc = j1;
        j1 = j2;
    };
    while (jc < walk.jend) {
        step(std::integral_constant<int, 0>{});
        if (!(jc < walk.jend)) break;
        step(std::integral_constant<int, 1>{});
    }
    const double out_scale = kInv4Pi / (double)gscale;
#pragma unroll
    for (int q = 0; q < I::kCbPerWave; ++q)
#pragma unroll
        for (int s = 0; s < I::kSlicesPerWave; ++s) {
            const int col = k0 + I::cb(wave, q) * 16 + (lane & 15);
            if (col >= K) continue;
            const double cs = ldexp(out_scale, -col_exponent(cmax[col]));
            double *dst = part + ((int64_t)blockIdx.z * K + col) * N;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int64_t i = i0 + I::slice(wave, s) * 16 + 4 * (lane >> 4) + r;
                if (i < N) dst[i] = (double)acc[q * I::kSlicesPerWave + s][r] * cs;
            }
        }
}

// Warp-specialised split-FP16 far field for 256-column blocks: 128 targets x 256 columns per
// workgroup, 32 sources per stage, 12 waves.  Waves 0-3 (one per SIMD) generate the hi/lo A tiles of
// the next stage into LDS (VALU: 2 target slices x 8 sources per lane; each lane loads its own
// source coordinates two stages ahead into registers) and copy the B operands of the stage after
// next global -> LDS (three LDS buffers); waves 4-11 (two per SIMD) run the MFMAs of the current
// stage, each on 4 target slices x 4 column blocks, reading A and B from LDS only.  One barrier per
// stage.  Generation, scaling, splitting and near-pair exclusion are those of
// far_mfma_f16x3_kernel.
constexpr int kWsSb = 32;        // sources per stage
// CW consumer waves: 8 (two per SIMD, each 4 target slices x 4 column blocks) or 4 (one per SIMD,
// each 8 target slices x 4 column blocks: a quarter less LDS traffic per MFMA, twice the
// accumulator registers); the autotuner picks between them
template <int CW>
constexpr int ws_threads() { return (4 + CW) * 64; }
constexpr int kWsMaskWords = 512;  // near-mask words per target panel staged in LDS (16384 source panels)

// Workgroup barrier for the producers: their LDS writes are complete (lgkmcnt(0)) and all but
// their newest VM operations (the copies and loads issued for the stage after next) have landed,
// so that the B operands copied a stage earlier are in LDS for every wave after the barrier.
// s_waitcnt immediates in the gfx9 encoding (vmcnt [3:0] and [15:14], expcnt [6:4], lgkmcnt [11:8]).
// __syncthreads would wait for every VM operation in flight.
template <int VM>  // VM = 63: no wait on VM operations
__device__ inline void ws_producer_barrier() {
    static_assert(VM >= 0 && VM <= 63, "vmcnt field");
    __builtin_amdgcn_s_waitcnt(0x0070 | (VM & 15) | ((VM >> 4) << 14));
    asm volatile("s_barrier" ::: "memory");
}

template <int EXCL, int CW>
__global__ __launch_bounds__(ws_threads<CW>()) void far_ws_f16x3_kernel(DevGeom g, int K, int Kp,
                                                                  const _Float16 *__restrict__ zh,
                                                                  const _Float16 *__restrict__ zl,
                                                                  const unsigned *__restrict__ cmax, float gscale,
                                                                  double *__restrict__ part, int src_per_split,
                                                                  NearMask nm, const int *__restrict__ tplane) {
    static_assert(EXCL >= 0 && EXCL <= 2, "exclusion mode");
    static_assert(CW == 4 || CW == 8, "4 or 8 consumer waves");
    constexpr int RS = 32 / CW;  // target slices per consumer wave
    __shared__ uint4 s_a[2][8][2][64];   // [stage parity][target slice][hi/lo][lane]
    __shared__ uint4 s_b[3][16][2][64];  // [stage % 3][column block][hi/lo][lane]
    __shared__ double s_pos[3][3][kWsSb];  // [stage % 3][x, y, z][source]
    // near-mask rows of the tile's target panels: the 8 panels of the 8 target slices (EXCL = 1,
    // npp = 16), or the at most 8 panels the 128 targets touch (EXCL = 2, npp >= 19)
    __shared__ uint32_t s_mask[EXCL ? 8 : 1][EXCL ? kWsMaskWords : 1];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const bool producer = wave < 4;
    const int64_t N = g.N;
    const int64_t i0 = (int64_t)blockIdx.x * kTargets;
    const int k0 = blockIdx.y * 256;
    const int64_t jb0 = (int64_t)blockIdx.z * src_per_split;
    // tplane (the plane ids of the 256-node tiles) is a restrict kernel argument, so that its
    // wave-uniform reads go through the scalar cache and never wait on vector loads in flight
    const StageWalk walk{min(N, jb0 + src_per_split), tplane ? tplane[i0 / 256] : -1, tplane};
    // Both roles walk the same stages (jc, then j1, j2) and pass one barrier per stage; their
    // loops are separate so that each keeps only its own state in registers.
    auto after = [&](int64_t j) { return j < walk.jend ? walk.next(j + kWsSb) : walk.jend; };
    int64_t jc = walk.next(jb0);
    int64_t j1 = after(jc);

    if (producer) {
        const int pw = wave;  // slices 2 pw, 2 pw + 1
        const int64_t io = i0 < N ? i0 : N - 1;
        const double ox = g.x[io], oy = g.y[io], oz = g.z[io];
        // two targets per lane, FP32 coordinates relative to the tile's first target, normals
        // scaled by gscale
        float tx[2], ty[2], tz[2], tn0[2], tn1[2], tn2[2];
        int tpr[2] = {0, 0};  // EXCL = 2: mask row (target panel - first panel of the tile)
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            int64_t tg = i0 + (2 * pw + u) * 16 + (lane & 15);
            tg = tg < N ? tg : N - 1;
            if (EXCL == 2) tpr[u] = (int)(tg / nm.npp - i0 / nm.npp);
            tx[u] = (float)(g.x[tg] - ox);
            ty[u] = (float)(g.y[tg] - oy);
            tz[u] = (float)(g.z[tg] - oz);
            tn0[u] = g.nxf[tg] * gscale;
            tn1[u] = g.nyf[tg] * gscale;
            tn2[u] = g.nzf[tg] * gscale;
        }
        // near-mask rows of the 8 target panels over the split's source panels, into LDS (the
        // host keeps them within kWsMaskWords); read after the first barrier
        const int64_t wbase = EXCL ? (jb0 / nm.npp) >> 5 : 0;
        if (EXCL && walk.jend > jb0) {
            const int nw = (int)((((walk.jend - 1) / nm.npp) >> 5) - wbase + 1);
            const int64_t tp_last = (N - 1) / nm.npp;
            for (int idx = tid; idx < 8 * nw; idx += 256) {
                const int sl = idx / nw, w = idx - sl * nw;
                int64_t tp;
                if (EXCL == 1) {
                    const int64_t tg = i0 + sl * 16;
                    tp = (tg < N ? tg : N - 1) / nm.npp;
                } else {
                    tp = i0 / nm.npp + sl;
                    tp = tp < tp_last ? tp : tp_last;
                }
                s_mask[sl][w] = nm.bits[tp * (int64_t)nm.mwords + wbase + w];
            }
        }
        // Every VM operation of the producers is a global -> LDS copy of the stage after next (B
        // operands and source coordinates), so the compiler never waits on one for a register; the
        // barriers wait for the previous stage's copies.  This wave copies column blocks
        // 4 pw .. 4 pw + 3 (eight 1 KiB copies per stage) and, for pw < 3, coordinate pw of the
        // stage's 32 sources (one 256-byte copy).
        typedef __attribute__((address_space(3))) void *lptr_t;
        const int zbytes = (int)(((N + kSb - 1) / kSb) * Kp * kSb * 2);  // the host keeps this below 2^31
        const __amdgpu_buffer_rsrc_t rh = __builtin_amdgcn_make_buffer_rsrc((void *)zh, (short)0, zbytes, 0x00020000);
        const __amdgpu_buffer_rsrc_t rl = __builtin_amdgcn_make_buffer_rsrc((void *)zl, (short)0, zbytes, 0x00020000);
        const double *coord = pw == 0 ? g.x : (pw == 1 ? g.y : g.z);
        const __amdgpu_buffer_rsrc_t rc =
            __builtin_amdgcn_make_buffer_rsrc((void *)coord, (short)0, (int)(N * (int64_t)sizeof(double)), 0x00020000);
        const int loff = 4 * pw * 1024 + lane * 16;  // column block 4 pw: lane l at 16 l of its chunk
        // The coordinate copy goes first: a barrier preceded by vmcnt(8) then has every wave's
        // coordinates of the stage landed (the other producer waves read them in generate) while
        // the 8 B-operand copies, read by the consumers a stage later, may still fly.
        auto copy_stage = [&](int64_t js, int pb) {
            const int soff = (int)(((js / kZt) * Kp + k0) * kZt * 2);
            // 64 lanes x 4 bytes = the stage's 32 doubles (reads past the end return zero)
            if (pw < 3)
                __builtin_amdgcn_raw_ptr_buffer_load_lds(rc, (lptr_t)&s_pos[pb][pw][0], 4, lane * 4,
                                                         (int)(js * (int64_t)sizeof(double)), 0, 0);
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int cb = 4 * pw + q;
                __builtin_amdgcn_raw_ptr_buffer_load_lds(rh, (lptr_t)&s_b[pb][cb][0][0], 16, loff + q * 1024, soff, 0, 0);
                __builtin_amdgcn_raw_ptr_buffer_load_lds(rl, (lptr_t)&s_b[pb][cb][1][0], 16, loff + q * 1024, soff, 0, 0);
            }
        };
        // A tiles of stage js (coordinates in s_pos[cb]) into s_a[pb].  Lane 16 r + e (and
        // 16 r + e + 8) takes source 8 r + e; the row broadcast hands it to the whole row.
        auto generate = [&](int pb, int cb, int64_t js) {
            const int sl = 8 * (lane >> 4) + (lane & 7);
            const bool in = js + sl < walk.jend;
            const float px = in ? (float)(s_pos[cb][0][sl] - ox) : 1.0e18f;
            const float py = in ? (float)(s_pos[cb][1][sl] - oy) : 1.0e18f;
            const float pz = in ? (float)(s_pos[cb][2][sl] - oz) : 1.0e18f;
            float n0[2], n1[2], n2[2];
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                n0[u] = tn0[u];
                n1[u] = tn1[u];
                n2[u] = tn2[u];
                if (EXCL == 1) {  // the lane's 8 sources lie in panel js / 16 + (lane >> 4) / 2
                    const int64_t sp0 = js / nm.npp;  // even: both panels' bits lie in one word
                    const bool ex = (s_mask[2 * pw + u][(sp0 >> 5) - wbase] >> ((sp0 & 31) + ((lane >> 4) >> 1))) & 1u;
                    n0[u] = ex ? 0.f : n0[u];
                    n1[u] = ex ? 0.f : n1[u];
                    n2[u] = ex ? 0.f : n2[u];
                }
            }
            // EXCL = 2 (npp >= 19): the lane's 8 consecutive sources js + 8 (lane >> 4) + e lie in
            // panel pa for e < bnd and in panel pa + 1 otherwise; exA / exB are those panels' bits
            int bnd = 8;
            bool exA[2] = {false, false}, exB[2] = {false, false};
            if (EXCL == 2) {
                const int64_t s0 = js + 8 * (lane >> 4);
                const int64_t pa = s0 / nm.npp;
                bnd = (int)((pa + 1) * nm.npp - s0);
                auto bit = [&](int row, int64_t sp) -> bool {
                    int w = (int)((sp >> 5) - wbase);  // padding sources past the split may point
                    w = w < 0 ? 0 : (w > kWsMaskWords - 1 ? kWsMaskWords - 1 : w);  // one word beyond
                    return (s_mask[row][w] >> (sp & 31)) & 1u;
                };
#pragma unroll
                for (int u = 0; u < 2; ++u) {
                    exA[u] = bit(tpr[u], pa);
                    exB[u] = bit(tpr[u], pa + 1);
                }
            }
            unsigned h[2][4], l[2][4];
#pragma unroll
            for (int e = 0; e < 8; e += 2) {
                float v[2][2];
#pragma unroll
                for (int d = 0; d < 2; ++d) {
                    const float sx = row_bcast(px, e + d), sy = row_bcast(py, e + d), sz = row_bcast(pz, e + d);
#pragma unroll
                    for (int u = 0; u < 2; ++u) {
                        const float dx = tx[u] - sx, dy = ty[u] - sy, dz = tz[u] - sz;
                        const float r2 = fmaf(dz, dz, fmaf(dy, dy, fmaf(dx, dx, 1e-20f)));
                        const float rinv = __builtin_amdgcn_rsqf(r2);
                        v[u][d] = (n0[u] * dx + n1[u] * dy + n2[u] * dz) * (rinv * rinv * rinv);
                        if (EXCL == 2) v[u][d] = (e + d < bnd ? exA[u] : exB[u]) ? 0.f : v[u][d];
                    }
                }
#pragma unroll
                for (int u = 0; u < 2; ++u) {
                    const v2h hh = __builtin_convertvector(v2f{v[u][0], v[u][1]}, v2h);
                    const v2f hf = __builtin_convertvector(hh, v2f);
                    const v2h ll = __builtin_convertvector(v2f{v[u][0] - hf[0], v[u][1] - hf[1]}, v2h);
                    h[u][e / 2] = __builtin_bit_cast(unsigned, hh);
                    l[u][e / 2] = __builtin_bit_cast(unsigned, ll);
                }
            }
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                s_a[pb][2 * pw + u][0][lane] = uint4{h[u][0], h[u][1], h[u][2], h[u][3]};
                s_a[pb][2 * pw + u][1][lane] = uint4{l[u][0], l[u][1], l[u][2], l[u][3]};
            }
        };
        int bc = 0;  // LDS buffer of stage jc (B operands, coordinates)
        if (jc < walk.jend) {
            copy_stage(jc, 0);
            if (j1 < walk.jend) {
                copy_stage(j1, 1);
                ws_producer_barrier<8>();  // stage jc's copies, stage j1's coordinates (and the mask
                                           // rows) are in LDS; stage j1's B operands may fly
            } else {
                ws_producer_barrier<0>();
            }
            generate(0, 0, jc);
            ws_producer_barrier<63>();
        }
        // stage jc (parity C, buffer bc): generate stage j1 into s_a[1 - C] from s_pos[bc + 1] while
        // the consumers multiply stage jc; copy stage j2 into buffer bc + 2 (mod 3).  Stage j1's
        // coordinates landed before the previous barrier (in every wave); before this barrier
        // stage j1's B operands and stage j2's coordinates land, and only stage j2's B operands
        // (this wave's newest 8 copies) may fly.
        auto step = [&](auto cc) {
            constexpr int C = decltype(cc)::value;
            const int64_t j2 = after(j1);
            const bool more2 = j2 < walk.jend;
            if (more2) copy_stage(j2, bc == 0 ? 2 : bc - 1);
            generate(1 - C, bc == 2 ? 0 : bc + 1, j1);
            if (!more2) ws_producer_barrier<0>();
            else ws_producer_barrier<8>();
            jc = j1;
            j1 = j2;
            bc = bc == 2 ? 0 : bc + 1;
        };
        while (jc < walk.jend) {
            step(std::integral_constant<int, 0>{});
            if (!(jc < walk.jend)) break;
            step(std::integral_constant<int, 1>{});
        }
        return;
    }

    // An (empty) statement that clobbers an accumulation register makes the compiler keep the
    // MFMA accumulators in AGPRs, which leaves the VGPR file to the producers' VALU work on the
    // same SIMD (with the accumulators in VGPRs the two hardly overlap).
    asm volatile("" ::: "a0");
    const int cw = wave - 4;
    const int sg = cw >> 2, cg = cw & 3;  // target slices RS sg .. RS sg + RS - 1, column blocks 4 cg .. 4 cg + 3
    v4f acc[RS][4];
#pragma unroll
    for (int s = 0; s < RS; ++s)
#pragma unroll
        for (int q = 0; q < 4; ++q) acc[s][q] = v4f{0.f, 0.f, 0.f, 0.f};
    if (jc < walk.jend) {
        __syncthreads();
        __syncthreads();
    }
    // stage jc: A in s_a[c], B in s_b[bc] (the producers have made both complete before the barrier)
    for (int c = 0, bc = 0; jc < walk.jend; c ^= 1, bc = bc == 2 ? 0 : bc + 1) {
        uint4 bh[4], bl[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            bh[q] = s_b[bc][4 * cg + q][0][lane];
            bl[q] = s_b[bc][4 * cg + q][1][lane];
        }
#pragma unroll
        for (int s = 0; s < RS; ++s) {
            const v8h ah = __builtin_bit_cast(v8h, s_a[c][RS * sg + s][0][lane]);
            const v8h al = __builtin_bit_cast(v8h, s_a[c][RS * sg + s][1][lane]);
            // the three products of an accumulator written as one chain: the compiler then keeps
            // every accumulator in its own AGPRs (it interleaves the chains by itself)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                v4f t = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, __builtin_bit_cast(v8h, bh[q]), acc[s][q], 0, 0, 0);
                t = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, __builtin_bit_cast(v8h, bl[q]), t, 0, 0, 0);
                acc[s][q] = __builtin_amdgcn_mfma_f32_16x16x32_f16(al, __builtin_bit_cast(v8h, bh[q]), t, 0, 0, 0);
            }
        }
        // keep this stage's MFMAs (and the registers they read) on this side of the barrier
        __builtin_amdgcn_sched_barrier(0);
        __syncthreads();
        jc = j1;
        j1 = after(j1);
    }
    const double out_scale = kInv4Pi / (double)gscale;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int col = k0 + (4 * cg + q) * 16 + (lane & 15);
        if (col >= K) continue;
        const double cs = ldexp(out_scale, -col_exponent(cmax[col]));
        double *dst = part + ((int64_t)blockIdx.z * K + col) * N;
#pragma unroll
        for (int s = 0; s < RS; ++s)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int64_t i = i0 + (RS * sg + s) * 16 + 4 * (lane >> 4) + r;
                if (i < N) dst[i] = (double)acc[s][q][r] * cs;
            }
    }
}

// ---------------------------------------------------------------- stored split-FP16 far field
// The split-FP16 far field of a 256-column block moves five times more MFMA operand bytes per stage
// through the LDS than it generates entries, and the generation (about 18 VALU instructions per
// entry) competes with the MFMAs for every SIMD.  With 288 GB of HBM the kernel matrix itself fits:
// at N = 64,384 its cross-plane entries are 3.4e9 hi/lo fp16 pairs, 13.6 GB.  far_g_build_kernel
// writes them once per operator, exactly as the generating kernels compute them, and
// far_gs_f16x3_kernel streams them as the A operand of an MFMA GEMM: no VALU work, the HBM stream
// of the stored entries (4 bytes per pair and apply) against the three products of the split.
//
// Storage: per 256-target super-tile T (two 128-target halves, each generated relative to its own
// first target as the 128-target kernels do) and per 32-source stage of each source tile of 256
// nodes that the stage walk does not skip for T (same-plane tiles hold exact zeros), one block of
// 16 target slices x hi/lo x 64 lanes x 8 halves (32 KiB) in MFMA A-operand order.  tile_off[T *
// (ntile + 1) + st] counts the stored source tiles of T before source tile st; block b of tile
// st's stages is at tile_off[...] * 8 + (stage within the tile).
constexpr int kGsT = 256;                // targets per super-tile
constexpr int kGsChunks = 32;            // 1 KiB chunks per block: [slice][hi/lo]
constexpr int kGsBlockU4 = kGsChunks * 64;
constexpr int kGsBlockBytes = kGsBlockU4 * 16;

// grid (super-tiles, source tiles), 8 waves: wave w generates target slices 2 w, 2 w + 1 of the
// super-tile for the 8 stages of source tile blockIdx.y.  EXCL as in far_ws_f16x3_kernel (1: npp =
// 16, a zero normal per 8-source group; 2: per source), so that every stored entry equals the one the
// generating kernel computes for the same pair, bit for bit.
template <int EXCL>
__global__ __launch_bounds__(512) void far_g_build_kernel(DevGeom g, float gscale, NearMask nm, uint4 *__restrict__ gst,
                                                          const int *__restrict__ tile_off, int ntile) {
    const int T = blockIdx.x, st = blockIdx.y;
    const int64_t N = g.N;
    const int tpid = g.tile_plane ? g.tile_plane[T] : -1;
    if (tpid >= 0 && g.tile_plane[st] == tpid) return;
    const int64_t blk0 = (int64_t)tile_off[(int64_t)T * (ntile + 1) + st] * 8;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int h = w >> 2, pw = w & 3;  // 128-target half, producer slot within it (slices 2 pw, 2 pw + 1)
    const int64_t i0 = (int64_t)T * kGsT + 128 * h;
    const int64_t io = i0 < N ? i0 : N - 1;
    const double ox = g.x[io], oy = g.y[io], oz = g.z[io];
    float tx[2], ty[2], tz[2], tn0[2], tn1[2], tn2[2];
    int64_t tpan[2] = {0, 0};  // target panel of the mask row
#pragma unroll
    for (int u = 0; u < 2; ++u) {
        int64_t tg = i0 + (2 * pw + u) * 16 + (lane & 15);
        tg = tg < N ? tg : N - 1;
        tx[u] = (float)(g.x[tg] - ox);
        ty[u] = (float)(g.y[tg] - oy);
        tz[u] = (float)(g.z[tg] - oz);
        tn0[u] = g.nxf[tg] * gscale;
        tn1[u] = g.nyf[tg] * gscale;
        tn2[u] = g.nzf[tg] * gscale;
        if (EXCL == 1) {
            const int64_t ts = i0 + (2 * pw + u) * 16;
            tpan[u] = (ts < N ? ts : N - 1) / nm.npp;
        } else if (EXCL == 2) {
            tpan[u] = tg / nm.npp;
        }
    }
    auto bit = [&](int64_t tp, int64_t sp) -> bool {
        return (nm.bits[tp * (int64_t)nm.mwords + (sp >> 5)] >> (sp & 31)) & 1u;
    };
    for (int q = 0; q < 8; ++q) {
        const int64_t js = (int64_t)st * 256 + 32 * q;
        if (js >= N) break;
        const int64_t s0 = js + 8 * (lane >> 4);  // the lane's 8 sources s0 .. s0 + 7
        float n0[2], n1[2], n2[2];
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            n0[u] = tn0[u];
            n1[u] = tn1[u];
            n2[u] = tn2[u];
            if (EXCL == 1) {
                const bool ex = bit(tpan[u], js / nm.npp + ((lane >> 4) >> 1));
                n0[u] = ex ? 0.f : n0[u];
                n1[u] = ex ? 0.f : n1[u];
                n2[u] = ex ? 0.f : n2[u];
            }
        }
        unsigned hw[2][4], lw[2][4];
#pragma unroll
        for (int e = 0; e < 8; e += 2) {
            float v[2][2];
#pragma unroll
            for (int d = 0; d < 2; ++d) {
                const int64_t j = s0 + e + d;
                const bool in = j < N;
                const float sx = in ? (float)(g.x[j] - ox) : 1.0e18f;
                const float sy = in ? (float)(g.y[j] - oy) : 1.0e18f;
                const float sz = in ? (float)(g.z[j] - oz) : 1.0e18f;
#pragma unroll
                for (int u = 0; u < 2; ++u) {
                    const float dx = tx[u] - sx, dy = ty[u] - sy, dz = tz[u] - sz;
                    const float r2 = fmaf(dz, dz, fmaf(dy, dy, fmaf(dx, dx, 1e-20f)));
                    const float rinv = __builtin_amdgcn_rsqf(r2);
                    v[u][d] = (n0[u] * dx + n1[u] * dy + n2[u] * dz) * (rinv * rinv * rinv);
                    if (EXCL == 2) v[u][d] = (in && bit(tpan[u], j / nm.npp)) ? 0.f : v[u][d];
                }
            }
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                const v2h hh = __builtin_convertvector(v2f{v[u][0], v[u][1]}, v2h);
                const v2f hf = __builtin_convertvector(hh, v2f);
                const v2h ll = __builtin_convertvector(v2f{v[u][0] - hf[0], v[u][1] - hf[1]}, v2h);
                hw[u][e / 2] = __builtin_bit_cast(unsigned, hh);
                lw[u][e / 2] = __builtin_bit_cast(unsigned, ll);
            }
        }
        uint4 *blk = gst + (blk0 + q) * kGsBlockU4;
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            const int slice = 8 * h + 2 * pw + u;
            blk[(slice * 2 + 0) * 64 + lane] = uint4{hw[u][0], hw[u][1], hw[u][2], hw[u][3]};
            blk[(slice * 2 + 1) * 64 + lane] = uint4{lw[u][0], lw[u][1], lw[u][2], lw[u][3]};
        }
    }
}

// Streamed GEMM over the stored entries: 256 targets x 256 columns per workgroup, 32 sources per
// stage.  CW = 4 waves (one per SIMD, 2 x 2: 128 targets x 128 columns each, 256 accumulator
// registers) or CW = 8 (two per SIMD, 2 x 4: 128 targets x 64 columns).  Every wave copies its share
// of the stage's stored block (32 KiB, global -> LDS, three-deep ring: issued two stages ahead) and
// of the B operands (the far field's fp16 hi/lo split of w.*X, 32 KiB, two-deep: one stage ahead),
// then multiplies from LDS.  Each ring slot is its own LDS variable and the stage loop is unrolled
// over the six slot combinations, so that the compiler sees that a copy in flight never writes
// what the wave reads.  One barrier per stage waits for the copies issued a stage earlier.  The
// accumulation order per output (stage by stage: hi*hi, hi*lo, lo*hi) is that of
// far_ws_f16x3_kernel, so with the same source split the partial sums are bit-identical.
template <int CW>
__global__ __launch_bounds__(CW * 64) void far_gs_f16x3_kernel(const int *__restrict__ tplane, int64_t N, int K, int Kp,
                                                               const _Float16 *__restrict__ zh,
                                                               const _Float16 *__restrict__ zl,
                                                               const unsigned *__restrict__ cmax, float gscale,
                                                               double *__restrict__ part, int src_per_split,
                                                               const uint4 *__restrict__ gst,
                                                               const int *__restrict__ tile_off, int ntile) {
    static_assert(CW == 4 || CW == 8, "4 or 8 waves");
    constexpr int NCB = CW == 4 ? 8 : 4;  // 16-column blocks per wave
    constexpr int NC = kGsChunks / CW;    // chunks of A and of B each wave copies per stage
    __shared__ uint4 s_a0[kGsChunks][64], s_a1[kGsChunks][64], s_a2[kGsChunks][64];
    __shared__ uint4 s_b0[kGsChunks][64], s_b1[kGsChunks][64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int th = CW == 4 ? wave >> 1 : wave >> 2;  // target half: slices 8 th .. 8 th + 7
    const int cg = CW == 4 ? wave & 1 : wave & 3;    // column blocks NCB cg .. NCB cg + NCB - 1
    const int T = blockIdx.x;
    const int k0 = blockIdx.y * 256;
    const int64_t jb0 = (int64_t)blockIdx.z * src_per_split;
    const StageWalk walk{min(N, jb0 + src_per_split), tplane ? tplane[T] : -1, tplane};
    auto after = [&](int64_t j) { return j < walk.jend ? walk.next(j + kWsSb) : walk.jend; };
    const int *toff = tile_off + (int64_t)T * (ntile + 1);
    const int t0 = toff[0];
    typedef __attribute__((address_space(3))) void *lptr_t;
    const __amdgpu_buffer_rsrc_t ra = __builtin_amdgcn_make_buffer_rsrc(
        (void *)(gst + (int64_t)t0 * 8 * kGsBlockU4), (short)0, (toff[ntile] - t0) * 8 * kGsBlockBytes, 0x00020000);
    const int zbytes = (int)(((N + kSb - 1) / kSb) * Kp * kSb * 2);  // the host keeps this below 2^31
    const __amdgpu_buffer_rsrc_t rh = __builtin_amdgcn_make_buffer_rsrc((void *)zh, (short)0, zbytes, 0x00020000);
    const __amdgpu_buffer_rsrc_t rl = __builtin_amdgcn_make_buffer_rsrc((void *)zl, (short)0, zbytes, 0x00020000);
    auto copy_a = [&](int64_t js, uint4 (&sa)[kGsChunks][64]) {
        const int soff = ((toff[js / 256] - t0) * 8 + (int)((js % 256) / 32)) * kGsBlockBytes;
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            const int ch = NC * wave + c;
            __builtin_amdgcn_raw_ptr_buffer_load_lds(ra, (lptr_t)&sa[ch][0], 16, ch * 1024 + lane * 16, soff, 0, 0);
        }
    };
    auto copy_b = [&](int64_t js, uint4 (&sb)[kGsChunks][64]) {
        const int soff = (int)(((js / kZt) * Kp + k0) * kZt * 2);
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            const int ch = NC * wave + c;  // column block ch / 2, hi (even) or lo (odd)
            const int loff = (ch >> 1) * 1024 + lane * 16;
            __builtin_amdgcn_raw_ptr_buffer_load_lds((c & 1) ? rl : rh, (lptr_t)&sb[ch][0], 16, loff, soff, 0, 0);
        }
    };
    v4f acc[8][NCB];
#pragma unroll
    for (int r = 0; r < 8; ++r)
#pragma unroll
        for (int q = 0; q < NCB; ++q) acc[r][q] = v4f{0.f, 0.f, 0.f, 0.f};
    // The A operands of target slice r + 1 are read into the other register pair before the MFMAs of
    // slice r are issued (scheduling barriers keep that order), so no slice waits on its LDS read.
    auto compute = [&](const uint4 (&sa)[kGsChunks][64], const uint4 (&sb)[kGsChunks][64]) {
        uint4 bh[NCB], bl[NCB], ah[2], al[2];
        ah[0] = sa[(8 * th) * 2 + 0][lane];
        al[0] = sa[(8 * th) * 2 + 1][lane];
#pragma unroll
        for (int q = 0; q < NCB; ++q) {
            bh[q] = sb[(NCB * cg + q) * 2 + 0][lane];
            bl[q] = sb[(NCB * cg + q) * 2 + 1][lane];
        }
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            if (r < 7) {
                ah[(r + 1) & 1] = sa[(8 * th + r + 1) * 2 + 0][lane];
                al[(r + 1) & 1] = sa[(8 * th + r + 1) * 2 + 1][lane];
            }
            const v8h xh = __builtin_bit_cast(v8h, ah[r & 1]);
            const v8h xl = __builtin_bit_cast(v8h, al[r & 1]);
#pragma unroll
            for (int q = 0; q < NCB; ++q) {
                v4f t = __builtin_amdgcn_mfma_f32_16x16x32_f16(xh, __builtin_bit_cast(v8h, bh[q]), acc[r][q], 0, 0, 0);
                t = __builtin_amdgcn_mfma_f32_16x16x32_f16(xh, __builtin_bit_cast(v8h, bl[q]), t, 0, 0, 0);
                acc[r][q] = __builtin_amdgcn_mfma_f32_16x16x32_f16(xl, __builtin_bit_cast(v8h, bh[q]), t, 0, 0, 0);
            }
            // this region: the LDS reads (slice 0: A(0), B, A(1); later slices: A(r + 1)) ahead of
            // the slice's MFMAs
            if (r == 0) __builtin_amdgcn_sched_group_barrier(0x100, 4 + 2 * NCB, 0);
            else if (r < 7) __builtin_amdgcn_sched_group_barrier(0x100, 2, 0);
            __builtin_amdgcn_sched_group_barrier(0x008, 3 * NCB, 0);
            __builtin_amdgcn_sched_barrier(0);
        }
    };
    int64_t jc = walk.next(jb0);
    int64_t j1 = after(jc);
    int64_t j2 = after(j1);
    // The copies are issued without branches (a stage past the walk copies the current stage again
    // into a slot no later stage reads): the compiler's wait analysis then keeps the copies in
    // flight apart from the LDS reads of other slots across every join of the control flow.
    if (jc < walk.jend) {
        copy_a(jc, s_a0);
        copy_b(jc, s_b0);
        copy_a(j1 < walk.jend ? j1 : jc, s_a1);
        ws_producer_barrier<NC>();  // stage jc in LDS; the next stage's block may fly
    }
    // one stage: A in slot a, B in slot b; copies of B(j1) into the other B slot and of A(j2) into
    // slot (a + 2) % 3, both read last in the previous stage (every wave has passed its barrier)
    auto stage = [&](uint4 (&sa)[kGsChunks][64], uint4 (&sb)[kGsChunks][64], uint4 (&na)[kGsChunks][64],
                     uint4 (&nb)[kGsChunks][64]) {
        copy_b(j1 < walk.jend ? j1 : jc, nb);
        copy_a(j2 < walk.jend ? j2 : jc, na);
        // the copies go out first: the barrier at the end of the stage waits for them
        __builtin_amdgcn_sched_barrier(0);
        compute(sa, sb);
        // keep this stage's LDS reads on this side of the barrier
        __builtin_amdgcn_sched_barrier(0);
        ws_producer_barrier<NC>();  // A(j1), B(j1) landed; A(j2) may fly
        jc = j1;
        j1 = j2;
        j2 = after(j2);
        return jc < walk.jend;
    };
    while (jc < walk.jend) {
        if (!stage(s_a0, s_b0, s_a2, s_b1)) break;
        if (!stage(s_a1, s_b1, s_a0, s_b0)) break;
        if (!stage(s_a2, s_b0, s_a1, s_b1)) break;
        if (!stage(s_a0, s_b1, s_a2, s_b0)) break;
        if (!stage(s_a1, s_b0, s_a0, s_b1)) break;
        if (!stage(s_a2, s_b1, s_a1, s_b0)) break;
    }
    __builtin_amdgcn_s_waitcnt(0xC07F & ~0xC00F);  // vmcnt(0): no copy outlives the loop
    const double out_scale = kInv4Pi / (double)gscale;
#pragma unroll
    for (int q = 0; q < NCB; ++q) {
        const int col = k0 + (NCB * cg + q) * 16 + (lane & 15);
        if (col >= K) continue;
        const double cs = ldexp(out_scale, -col_exponent(cmax[col]));
        double *dst = part + ((int64_t)blockIdx.z * K + col) * N;
#pragma unroll
        for (int r = 0; r < 8; ++r)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int64_t i = (int64_t)T * kGsT + (8 * th + r) * 16 + 4 * (lane >> 4) + e;
                if (i < N) dst[i] = (double)acc[r][q][e] * cs;
            }
    }
}

// Streamed GEMM with the stored entries loaded straight into registers: 256 targets x 256 columns
// per workgroup, W = 4 waves (one per SIMD) or 8 (two per SIMD); wave w owns target slices RS w ..
// RS w + RS - 1 (RS = 16 / W: 64 or 32 targets) and all 16 column blocks (16 RS accumulator
// registers per lane).  A wave's part of a stage's block (its RS slices, hi and lo: 2 RS KiB,
// contiguous) is used by that wave alone, so it goes from HBM into registers with non-temporal loads
// three stages ahead (four register sets), without passing through the LDS; only the B operands,
// which all waves read, are copied global -> LDS, into a four-deep ring, also three stages ahead.
// The copies of the LDS path then carry 32 KiB per stage instead of 64, and each CU keeps about
// 190 KiB of loads in flight.  With W = 8 a SIMD has a second wave to issue MFMAs while one waits.
// Accumulation order per output as in far_ws_f16x3_kernel (bit-identical partial sums at the same
// source split).
template <int W>
__global__ __launch_bounds__(W * 64) void far_gr_f16x3_kernel(
    const int *__restrict__ tplane, int64_t N, int K, int Kp, const _Float16 *__restrict__ zh,
    const _Float16 *__restrict__ zl, const unsigned *__restrict__ cmax, float gscale, double *__restrict__ part,
    int src_per_split, const uint4 *__restrict__ gst, const int *__restrict__ tile_off, int ntile) {
    static_assert(W == 4 || W == 8, "4 or 8 waves");
    constexpr int NC = kGsChunks / W;  // B chunks each wave copies per stage (8 or 4)
    constexpr int RS = 16 / W;         // target slices per wave
    constexpr int NA = 2 * RS;         // A chunks per wave and stage (hi and lo per slice)
    __shared__ uint4 s_b0[kGsChunks][64], s_b1[kGsChunks][64], s_b2[kGsChunks][64], s_b3[kGsChunks][64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int T = blockIdx.x;
    const int k0 = blockIdx.y * 256;
    const int64_t jb0 = (int64_t)blockIdx.z * src_per_split;
    const StageWalk walk{min(N, jb0 + src_per_split), tplane ? tplane[T] : -1, tplane};
    auto after = [&](int64_t j) { return j < walk.jend ? walk.next(j + kWsSb) : walk.jend; };
    const int *toff = tile_off + (int64_t)T * (ntile + 1);
    const int t0 = toff[0];
    // A through a buffer resource over this super-tile's blocks: the per-stage offset is a scalar,
    // the per-lane offset fixed, so no vector address is computed (or waited for) per stage
    const __amdgpu_buffer_rsrc_t ra = __builtin_amdgcn_make_buffer_rsrc(
        (void *)(gst + (int64_t)t0 * 8 * kGsBlockU4), (short)0, (toff[ntile] - t0) * 8 * kGsBlockBytes, 0x00020000);
    const int aoff = ((NA * wave) * 64 + lane) * 16;
    typedef __attribute__((address_space(3))) void *lptr_t;
    const int zbytes = (int)(((N + kSb - 1) / kSb) * Kp * kSb * 2);  // the host keeps this below 2^31
    const __amdgpu_buffer_rsrc_t rh = __builtin_amdgcn_make_buffer_rsrc((void *)zh, (short)0, zbytes, 0x00020000);
    const __amdgpu_buffer_rsrc_t rl = __builtin_amdgcn_make_buffer_rsrc((void *)zl, (short)0, zbytes, 0x00020000);
    auto copy_b = [&](int64_t js, uint4 (&sb)[kGsChunks][64]) {
        const int soff = (int)(((js / kZt) * Kp + k0) * kZt * 2);
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            const int ch = NC * wave + c;  // column block ch / 2, hi (even) or lo (odd)
            const int loff = (ch >> 1) * 1024 + lane * 16;
            __builtin_amdgcn_raw_ptr_buffer_load_lds((c & 1) ? rl : rh, (lptr_t)&sb[ch][0], 16, loff, soff, 0, 0);
        }
    };
    // chunk c of this wave: target slice RS wave + c / 2, hi (c even) or lo (c odd)
    auto load_a = [&](int64_t js, uint4 (&a)[NA]) {
        const int soff = ((toff[js / 256] - t0) * 8 + (int)((js % 256) / 32)) * kGsBlockBytes;
#pragma unroll
        for (int c = 0; c < NA; ++c)  // non-temporal (aux 2): the stream is read once
            a[c] = __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(ra, aoff + c * 1024, soff, 2));
    };
    v4f acc[RS][16];
#pragma unroll
    for (int r = 0; r < RS; ++r)
#pragma unroll
        for (int q = 0; q < 16; ++q) acc[r][q] = v4f{0.f, 0.f, 0.f, 0.f};
    // column blocks in pairs: the B operands of the next pair are read before this pair's 6 RS MFMAs
    auto compute = [&](const uint4 (&sb)[kGsChunks][64], const uint4 (&a)[NA]) {
        uint4 bb[2][4];
#pragma unroll
        for (int u = 0; u < 4; ++u) bb[0][u] = sb[u][lane];
#pragma unroll
        for (int p = 0; p < 8; ++p) {
            if (p < 7) {
#pragma unroll
                for (int u = 0; u < 4; ++u) bb[(p + 1) & 1][u] = sb[4 * (p + 1) + u][lane];
            }
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int q = 2 * p + h;
                const v8h bh = __builtin_bit_cast(v8h, bb[p & 1][2 * h]);
                const v8h bl = __builtin_bit_cast(v8h, bb[p & 1][2 * h + 1]);
#pragma unroll
                for (int r = 0; r < RS; ++r) {
                    const v8h xh = __builtin_bit_cast(v8h, a[2 * r]);
                    const v8h xl = __builtin_bit_cast(v8h, a[2 * r + 1]);
                    v4f t = __builtin_amdgcn_mfma_f32_16x16x32_f16(xh, bh, acc[r][q], 0, 0, 0);
                    t = __builtin_amdgcn_mfma_f32_16x16x32_f16(xh, bl, t, 0, 0, 0);
                    acc[r][q] = __builtin_amdgcn_mfma_f32_16x16x32_f16(xl, bh, t, 0, 0, 0);
                }
            }
            if (p == 0) __builtin_amdgcn_sched_group_barrier(0x100, 8, 0);
            else if (p < 7) __builtin_amdgcn_sched_group_barrier(0x100, 4, 0);
            __builtin_amdgcn_sched_group_barrier(0x008, 6 * RS, 0);
            __builtin_amdgcn_sched_barrier(0);
        }
    };
    int64_t jc = walk.next(jb0);
    int64_t j1 = after(jc);
    int64_t j2 = after(j1);
    int64_t j3 = after(j2);
    uint4 a0[NA], a1[NA], a2[NA], a3[NA];
    // Loads are issued without branches (a stage past the walk loads the current stage again into a
    // slot no later stage reads), so that the compiler's wait analysis stays exact across the loop.
    // Per stage a wave issues NC copies and NA = NC loads.
    if (jc < walk.jend) {
        copy_b(jc, s_b0);
        load_a(jc, a0);
        copy_b(j1 < walk.jend ? j1 : jc, s_b1);
        load_a(j1 < walk.jend ? j1 : jc, a1);
        copy_b(j2 < walk.jend ? j2 : jc, s_b2);
        load_a(j2 < walk.jend ? j2 : jc, a2);
        // a copy into the fourth slot too (overwritten, in issue order, by the first stage's copy):
        // the loop is then entered with copies pending into the four slots in the order a loop
        // iteration leaves them, and the compiler's wait analysis keeps them apart from the reads
        copy_b(jc, s_b3);
        ws_producer_barrier<5 * NC>();  // B(jc), A(jc) in LDS / registers; later loads may fly
    }
    // one stage: B in ring slot b, A in register set a; B(j3) into slot (b + 3) % 4 (read last in the
    // previous stage, before its barrier) and A(j3) into register set (a + 3) % 4 (consumed last in
    // the previous stage)
    auto stage = [&](const uint4 (&sb)[kGsChunks][64], const uint4 (&a)[NA], uint4 (&nb)[kGsChunks][64],
                     uint4 (&na)[NA]) {
        const int64_t jn = j3 < walk.jend ? j3 : jc;
        copy_b(jn, nb);
        load_a(jn, na);
        __builtin_amdgcn_sched_barrier(0);
        compute(sb, a);
        __builtin_amdgcn_sched_barrier(0);
        ws_producer_barrier<5 * NC>();  // B(j1) landed; A(j1), B(j2), A(j2), B(j3), A(j3) may fly
        jc = j1;
        j1 = j2;
        j2 = j3;
        j3 = after(j3);
        return jc < walk.jend;
    };
    while (jc < walk.jend) {
        if (!stage(s_b0, a0, s_b3, a3)) break;
        if (!stage(s_b1, a1, s_b0, a0)) break;
        if (!stage(s_b2, a2, s_b1, a1)) break;
        if (!stage(s_b3, a3, s_b2, a2)) break;
    }
    __builtin_amdgcn_s_waitcnt(0x0070);  // vmcnt(0) lgkmcnt(0): no load outlives the loop
    const double out_scale = kInv4Pi / (double)gscale;
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int col = k0 + q * 16 + (lane & 15);
        if (col >= K) continue;
        const double cs = ldexp(out_scale, -col_exponent(cmax[col]));
        double *dst = part + ((int64_t)blockIdx.z * K + col) * N;
#pragma unroll
        for (int r = 0; r < RS; ++r)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int64_t i = (int64_t)T * kGsT + (RS * wave + r) * 16 + 4 * (lane >> 4) + e;
                if (i < N) dst[i] = (double)acc[r][q][e] * cs;
            }
    }
}

// 16-column blocks per workgroup for K columns (power of two, at most 16)
int ncb_for(int K) {
    const int nb = (K + 15) / 16;
    int ncb = 1;
    while (ncb < nb && ncb < 16) ncb *= 2;
    return ncb;
}

// Source split count (splits start on 256-node tiles).
int mfma_splits(int64_t N, int K, int64_t groups, int64_t part_cap, int forced, int64_t *per_out) {
    int64_t s = std::max<int64_t>(1, (768 + groups - 1) / groups);
    s = std::min<int64_t>(s, std::max<int64_t>(1, N / 2048));
    if (forced > 0) s = std::min<int64_t>(forced, std::max<int64_t>(1, N / 256));
    while (s > 1 && s * K * N > part_cap) --s;
    int64_t per = (N + s - 1) / s;
    per = (per + 255) / 256 * 256;
    *per_out = per;
    return (int)((N + per - 1) / per);
}

// max over all node pairs of |n_i . (x_i - y_j)| / r^3, in the arithmetic of the fp16 tile
// generator (FP64 differences rounded to FP32), as float bits.  grid: (target tiles, source splits)
__global__ __launch_bounds__(256) void gmax_kernel(DevGeom g, unsigned *__restrict__ out) {
    __shared__ double4 s_pos[256];
    const int64_t N = g.N;
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t ic = i < N ? i : N - 1;
    const double tx = g.x[ic], ty = g.y[ic], tz = g.z[ic];
    const float n0 = g.nxf[ic], n1 = g.nyf[ic], n2 = g.nzf[ic];
    const int64_t per = ((N + gridDim.y - 1) / gridDim.y + 255) / 256 * 256;
    const int64_t j0 = (int64_t)blockIdx.y * per, j1 = min(N, j0 + per);
    float m = 0.0f;
    for (int64_t jb = j0; jb < j1; jb += 256) {
        __syncthreads();
        const int64_t j = jb + threadIdx.x;
        s_pos[threadIdx.x] = j < j1 ? make_double4(g.x[j], g.y[j], g.z[j], 0.0) : make_double4(tx, ty, tz, 0.0);
        __syncthreads();
        for (int u = 0; u < 256; ++u) {
            const double4 sp = s_pos[u];
            const float dx = (float)(tx - sp.x), dy = (float)(ty - sp.y), dz = (float)(tz - sp.z);
            const float r2 = dx * dx + dy * dy + dz * dz;
            if (r2 > 0.0f) {
                const float rinv = __builtin_amdgcn_rsqf(r2);
                m = fmaxf(m, fabsf(n0 * dx + n1 * dy + n2 * dz) * (rinv * rinv * rinv));
            }
        }
    }
    if (i >= N) m = 0.0f;
    for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_down(m, off, 64));
    if ((threadIdx.x & 63) == 0) atomicMax(out, __float_as_uint(m));
}
}  // namespace

float far_field_f16_scale(hipStream_t st, const DevGeom &g, void *tmp4) {
    unsigned *d = static_cast<unsigned *>(tmp4);
    unsigned h = 0;
    zero_u32_kernel<<<1, 64, 0, st>>>(d, 1);
    gmax_kernel<<<dim3((unsigned)((g.N + 255) / 256), 8), 256, 0, st>>>(g, d);
    if (hipMemcpyAsync(&h, d, sizeof(unsigned), hipMemcpyDeviceToHost, st) != hipSuccess) return 0.0f;
    if (hipStreamSynchronize(st) != hipSuccess) return 0.0f;
    float gm;
    std::memcpy(&gm, &h, sizeof gm);
    if (!(gm > 0.0f) || !std::isfinite(gm)) return 1.0f;
    int e;
    (void)std::frexp(gm, &e);  // gm < 2^e
    return std::ldexp(1.0f, 14 - e);  // scaled entries stay below 2^14
}

int64_t far_stored_layout(const std::vector<int> &tile_plane, int64_t N, std::vector<int> &tile_off) {
    const int nt = (int)((N + 255) / 256);
    tile_off.assign((size_t)nt * (nt + 1), 0);
    int64_t c = 0;
    for (int T = 0; T < nt; ++T) {
        const int tp = (size_t)T < tile_plane.size() ? tile_plane[T] : -1;
        int *row = &tile_off[(size_t)T * (nt + 1)];
        for (int st = 0; st < nt; ++st) {
            row[st] = (int)c;
            if (!(tp >= 0 && (size_t)st < tile_plane.size() && tile_plane[st] == tp)) ++c;
        }
        row[nt] = (int)c;
    }
    return c * 8;
}

size_t far_stored_block_bytes() { return (size_t)kGsBlockBytes; }

void far_stored_build(hipStream_t st, const DevGeom &g, float gscale, bool exclude_near, const FarStored &fs) {
    NearMask nm;
    nm.bits = g.near_mask;
    nm.npp = g.npp;
    nm.mwords = g.mask_words;
    const int ex = (!exclude_near || !g.near_mask) ? 0 : (g.npp == 16 ? 1 : 2);
    const dim3 grid((unsigned)fs.ntile, (unsigned)fs.ntile);
    uint4 *gst = static_cast<uint4 *>(const_cast<void *>(fs.g));
    if (ex == 1) far_g_build_kernel<1><<<grid, 512, 0, st>>>(g, gscale, nm, gst, fs.tile_off, fs.ntile);
    else if (ex == 2) far_g_build_kernel<2><<<grid, 512, 0, st>>>(g, gscale, nm, gst, fs.tile_off, fs.ntile);
    else far_g_build_kernel<0><<<grid, 512, 0, st>>>(g, gscale, nm, gst, fs.tile_off, fs.ntile);
}

bool far_field_mfma_supported(int prec, int K) {
    if (K < 16) return false;
    return prec == 0 || prec == 2 || prec == 3 || prec == 4;
}

size_t far_field_mfma_scratch_bytes(int64_t N, int prec, int K) {
    const int ncb = ncb_for(K);
    const int Kp = ((K + ncb * 16 - 1) / (ncb * 16)) * ncb * 16;
    if (prec == 0) return sizeof(double) * (size_t)((N + kS64 - 1) / kS64) * kS64 * Kp;
    const size_t zelems = (size_t)((N + kSb - 1) / kSb) * Kp * kSb;
    if (prec == 4) return sizeof(_Float16) * zelems * 2 + sizeof(unsigned) * (size_t)Kp;
    return sizeof(__bf16) * zelems * (prec == 3 ? 2 : 1);
}

void far_f16_operands(void *scratch, int64_t N, int K, NearH3 *out) {
    const int ncb = ncb_for(K);
    const int Kp = ((K + ncb * 16 - 1) / (ncb * 16)) * ncb * 16;
    const size_t zelems = (size_t)((N + kSb - 1) / kSb) * Kp * kSb;
    const _Float16 *zh = static_cast<const _Float16 *>(scratch);
    out->zh = zh;
    out->zl = zh + zelems;
    out->cmax = reinterpret_cast<const unsigned *>(zh + 2 * zelems);
    out->Kp = Kp;
    out->tile = kZt;
}

void f16_split_operands(hipStream_t st, int64_t N, int K, const double *X, int64_t ldx, const double *w, void *scratch) {
    const int ncb = ncb_for(K);
    const int Kp = ((K + ncb * 16 - 1) / (ncb * 16)) * ncb * 16;
    const int64_t ntile = (N + kSb - 1) / kSb;
    const size_t zelems = (size_t)ntile * Kp * kSb;
    _Float16 *zh = static_cast<_Float16 *>(scratch);
    _Float16 *zl = zh + zelems;
    unsigned *cmax = reinterpret_cast<unsigned *>(zl + zelems);
    zero_u32_kernel<<<(unsigned)((K + 255) / 256), 256, 0, st>>>(cmax, K);
    colmax_kernel<<<dim3((unsigned)((N + kColmaxRows - 1) / kColmaxRows), (unsigned)K), 256, 0, st>>>(X, ldx, w, N,
                                                                                                     cmax);
    split_tiles_f16<<<dim3((unsigned)((ntile * kSb + 255) / 256), (unsigned)Kp), 256, 0, st>>>(X, ldx, w, N, K, Kp, cmax,
                                                                                                zh, zl);
}

int far_field_max_splits(int64_t N) { return (int)std::max<int64_t>(1, N / 256); }

int far_field_mfma(hipStream_t st, const DevGeom &g, int prec, int K, const double *X, int64_t ldx, double *part,
                   int64_t part_cap, void *scratch, float gscale, bool exclude_near, hipEvent_t ev0, hipEvent_t ev1,
                   const FarTune &tune, const FarStored *fs) {
    auto mark = [&](hipEvent_t e) {
        if (e) (void)hipEventRecord(e, st);
    };
    const int64_t N = g.N;
    const int64_t bx = (N + kTargets - 1) / kTargets;
    const int ncb = ncb_for(K);
    const int CB = ncb * 16;
    const int by = (K + CB - 1) / CB;
    const int Kp = by * CB;
    int64_t per;
    const int S = mfma_splits(N, K, bx * by, part_cap, tune.splits, &per);
    dim3 grid((unsigned)bx, (unsigned)by, (unsigned)S);
    if (prec == 0) {
        double *Zt = static_cast<double *>(scratch);
        const int64_t ntile = (N + kS64 - 1) / kS64;
        scale_tiles_f64<<<dim3((unsigned)((ntile * kS64 + 255) / 256), (unsigned)Kp), 256, 0, st>>>(X, ldx, g.w, N, K,
                                                                                                     Kp, Zt);
        mark(ev0);
        switch (ncb) {
#define DBIE_F64(n) case n: far_mfma_f64_kernel<n><<<grid, kThreads, 0, st>>>(g, K, Kp, Zt, part, (int)per); break;
            DBIE_F64(1) DBIE_F64(2) DBIE_F64(4) DBIE_F64(8) DBIE_F64(16)
#undef DBIE_F64
        }
        mark(ev1);
        return S;
    }
    const int64_t ntile = (N + kSb - 1) / kSb;
    const size_t zelems = (size_t)ntile * Kp * kSb;
    if (prec == 4) {
        _Float16 *zh = static_cast<_Float16 *>(scratch);
        _Float16 *zl = zh + zelems;
        unsigned *cmax = reinterpret_cast<unsigned *>(zl + zelems);
        f16_split_operands(st, N, K, X, ldx, g.w, scratch);
        NearMask nm;
        nm.bits = g.near_mask;
        nm.npp = g.npp;
        nm.mwords = g.mask_words;
        const int ex = (!exclude_near || !g.near_mask) ? 0 : (g.npp == 16 ? 1 : 2);
        mark(ev0);
        const int64_t mask_words_split = (per / std::max(g.npp, 1)) / 32 + 2;
        // the stored kernel matrix (variants 2 and 3), where the operator keeps one
        if (tune.variant >= 2 && fs && fs->g && ncb == 16 && zelems * sizeof(_Float16) < ((size_t)1 << 31) &&
            fs->ntile == (int)((N + kGsT - 1) / kGsT)) {
            const dim3 gg((unsigned)fs->ntile, (unsigned)by, (unsigned)S);
            const uint4 *gst = static_cast<const uint4 *>(fs->g);
            if (tune.variant == 2)
                far_gr_f16x3_kernel<4><<<gg, 4 * 64, 0, st>>>(g.tile_plane, N, K, Kp, zh, zl, cmax, gscale, part,
                                                              (int)per, gst, fs->tile_off, fs->ntile);
            else if (tune.variant == 4)
                far_gr_f16x3_kernel<8><<<gg, 8 * 64, 0, st>>>(g.tile_plane, N, K, Kp, zh, zl, cmax, gscale, part,
                                                              (int)per, gst, fs->tile_off, fs->ntile);
            else
                far_gs_f16x3_kernel<8><<<gg, 512, 0, st>>>(g.tile_plane, N, K, Kp, zh, zl, cmax, gscale, part, (int)per,
                                                           gst, fs->tile_off, fs->ntile);
            mark(ev1);
            return S;
        }
        // the warp-specialised kernel: exclusion per lane (npp = 16) or per source for panels of at
        // least 19 nodes (the 128 targets of a tile then touch at most 8 target panels)
        if (tune.variant != 0 && ncb == 16 && (ex != 2 || g.npp >= 19) && zelems * sizeof(_Float16) < ((size_t)1 << 31) &&
            (ex == 0 || mask_words_split <= kWsMaskWords)) {
#define DBIE_WS(E, C)                                                                                         \
    far_ws_f16x3_kernel<E, C><<<grid, ws_threads<C>(), 0, st>>>(g, K, Kp, zh, zl, cmax, gscale, part, (int)per, nm, \
                                                               g.tile_plane)
            if (tune.variant >= 1) {  // 2, 3 without a stored kernel matrix: one consumer wave per SIMD
                if (ex == 2) DBIE_WS(2, 4);
                else if (ex == 1) DBIE_WS(1, 4);
                else DBIE_WS(0, 4);
            } else {
                if (ex == 2) DBIE_WS(2, 8);
                else if (ex == 1) DBIE_WS(1, 8);
                else DBIE_WS(0, 8);
            }
#undef DBIE_WS
            mark(ev1);
            return S;
        }
        switch (ncb * 4 + ex) {
#define DBIE_H3(n, e) \
    case n * 4 + e: far_mfma_f16x3_kernel<n, e><<<grid, kThreads, 0, st>>>(g, K, Kp, zh, zl, cmax, gscale, part, (int)per, nm); break;
#define DBIE_H3E(n) DBIE_H3(n, 0) DBIE_H3(n, 1) DBIE_H3(n, 2)
            DBIE_H3E(1) DBIE_H3E(2) DBIE_H3E(4) DBIE_H3E(8) DBIE_H3E(16)
#undef DBIE_H3E
#undef DBIE_H3
        }
        mark(ev1);
        return S;
    }
    const bool x3 = prec == 3;
    __bf16 *zh = static_cast<__bf16 *>(scratch);
    __bf16 *zl = x3 ? zh + zelems : nullptr;
    split_tiles_bf16<<<dim3((unsigned)((ntile * kSb + 255) / 256), (unsigned)Kp), 256, 0, st>>>(X, ldx, g.w, N, K, Kp,
                                                                                                 zh, zl);
    mark(ev0);
    if (x3) {
        switch (ncb) {
#define DBIE_B3(n) case n: far_mfma_bf16_kernel<n, true><<<grid, kThreads, 0, st>>>(g, K, Kp, zh, zl, part, (int)per); break;
            DBIE_B3(1) DBIE_B3(2) DBIE_B3(4) DBIE_B3(8) DBIE_B3(16)
#undef DBIE_B3
        }
    } else {
        switch (ncb) {
#define DBIE_B1(n) case n: far_mfma_bf16_kernel<n, false><<<grid, kThreads, 0, st>>>(g, K, Kp, zh, nullptr, part, (int)per); break;
            DBIE_B1(1) DBIE_B1(2) DBIE_B1(4) DBIE_B1(8) DBIE_B1(16)
#undef DBIE_B1
        }
    }
    mark(ev1);
    return S;
}

}  // namespace dbie
