// far.hip -- far-field layer-potential sums on VALU (FP64 and FP32), gfx950.
//
// One thread owns one target node and KC right-hand-side columns; source nodes
// are streamed through LDS in tiles of 256.  The source range is split over
// gridDim.z so that a launch has enough workgroups to fill 256 CUs; each split
// writes its own partial sum (no atomics, bitwise reproducible) and the
// near-field combine kernel reduces the splits.
//
// Kernels (src/kernel/laplace3d.jl:1-52, the FMM maps at :84-137 replaced by the
// exact direct sum):
//   DT: sum_j w_j X_j n_i.(x_i - y_j) / (4 pi r^3)
//   S : sum_j w_j X_j / (4 pi r)
//   D : sum_j w_j X_j n_j.(y_j - x_i) / (4 pi r^3)
// Tiles whose targets and sources all lie on one plane are skipped for DT/D:
// their entries are exactly zero in FP64 (n.(x - y) = 0 bit-exactly on a face).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <stdexcept>

#include "kernels.hpp"

namespace dbie {

namespace {
constexpr int kTile = 256;
constexpr double kInv4Pi = 1.0 / (4.0 * 3.14159265358979311599796346854);

template <int MODE, int KC>
__global__ __launch_bounds__(kTile) void far_f64_kernel(DevGeom g, int K, const double *__restrict__ X, int64_t ldx,
                                                        double *__restrict__ part, int src_per_split) {
    __shared__ double4 s_pos[kTile];
    __shared__ double4 s_nrm[MODE == FAR_D ? kTile : 1];
    __shared__ double s_q[KC][kTile];
    const int64_t N = g.N;
    const int tid = threadIdx.x;
    const int64_t i = (int64_t)blockIdx.x * kTile + tid;
    const int k0 = blockIdx.y * KC;
    const int64_t jb0 = (int64_t)blockIdx.z * src_per_split;
    const int64_t jend = min(N, jb0 + src_per_split);
    const bool valid = i < N;
    const int64_t ii = valid ? i : N - 1;
    const double tx = g.x[ii], ty = g.y[ii], tz = g.z[ii];
    const double tnx = g.nx[ii], tny = g.ny[ii], tnz = g.nz[ii];
    const int tplane = g.tile_plane ? g.tile_plane[blockIdx.x] : -1;
    double acc[KC];
#pragma unroll
    for (int c = 0; c < KC; ++c) acc[c] = 0.0;

    for (int64_t jb = jb0; jb < jend; jb += kTile) {
        if (MODE != FAR_S && tplane >= 0 && g.tile_plane[jb / kTile] == tplane) continue;  // exact zeros
        const int64_t j = jb + tid;
        __syncthreads();
        if (j < jend) {
            s_pos[tid] = make_double4(g.x[j], g.y[j], g.z[j], 0.0);
            if (MODE == FAR_D) s_nrm[tid] = make_double4(g.nx[j], g.ny[j], g.nz[j], 0.0);
            const double wj = g.w[j];
#pragma unroll
            for (int c = 0; c < KC; ++c) s_q[c][tid] = (k0 + c < K) ? wj * X[j + (int64_t)(k0 + c) * ldx] : 0.0;
        }
        __syncthreads();
        const int nt = (int)min<int64_t>(kTile, jend - jb);
        for (int jj = 0; jj < nt; ++jj) {
            const double4 sp = s_pos[jj];
            const double dx = tx - sp.x, dy = ty - sp.y, dz = tz - sp.z;
            const double r2 = dx * dx + dy * dy + dz * dz;
            // 1/sqrt(r2): the FP32 hardware estimate y (relative error below 2e-7 for r2 in the normal
            // FP32 range, which node distances are) and one third-order (Halley) step in FP64,
            // y (1 + e / 2 + 3 e^2 / 8) with e = 1 - r2 y^2, which cubes the relative error
            double rinv = (double)__builtin_amdgcn_rsqf((float)r2);
            const double e = fma(-r2, rinv * rinv, 1.0);
            rinv = fma(rinv, e * fma(0.375, e, 0.5), rinv);
            rinv = r2 > 0.0 ? rinv : 0.0;
            double gk;
            if (MODE == FAR_DT) {
                gk = (tnx * dx + tny * dy + tnz * dz) * (rinv * rinv * rinv);
            } else if (MODE == FAR_S) {
                gk = rinv;
            } else {
                const double4 sn = s_nrm[jj];
                gk = -(sn.x * dx + sn.y * dy + sn.z * dz) * (rinv * rinv * rinv);
            }
#pragma unroll
            for (int c = 0; c < KC; ++c) acc[c] += gk * s_q[c][jj];
        }
    }
    if (!valid) return;
#pragma unroll
    for (int c = 0; c < KC; ++c)
        if (k0 + c < K) part[((int64_t)blockIdx.z * K + k0 + c) * N + i] = acc[c] * kInv4Pi;
}

// FP32 tile generation and accumulation (mixed-precision inner operator, DT only).  A source's KC
// charges are contiguous in LDS (KC / 4 vector reads per pair), and 1/r is the hardware reciprocal
// square root (about 1 ulp) of r^2 clamped at 1e-20: the self pair has n.(x - y) = 0 and stays 0.
// EXCL (16-node panels): the near panel pairs of g.near_mask are left out, as the split-FP16 far
// field leaves them out, and the caller adds their full interaction; a 256-source tile covers 16
// source panels, 16 bits of the target panel's mask row.
template <int KC, bool EXCL>
__global__ __launch_bounds__(kTile) void far_f32_kernel(DevGeom g, int K, const double *__restrict__ X, int64_t ldx,
                                                        double *__restrict__ part, int src_per_split) {
    static_assert(KC == 1 || KC == 2 || KC % 4 == 0, "charges per source: 1, 2 or a multiple of 4");
    __shared__ float4 s_pos[kTile];
    __shared__ __align__(16) float s_q[kTile][KC];
    const int64_t N = g.N;
    const int tid = threadIdx.x;
    const int64_t i = (int64_t)blockIdx.x * kTile + tid;
    const int k0 = blockIdx.y * KC;
    const int64_t jb0 = (int64_t)blockIdx.z * src_per_split;
    const int64_t jend = min(N, jb0 + src_per_split);
    const bool valid = i < N;
    const int64_t ii = valid ? i : N - 1;
    const float tx = g.xf[ii], ty = g.yf[ii], tz = g.zf[ii];
    const float tnx = g.nxf[ii], tny = g.nyf[ii], tnz = g.nzf[ii];
    const int tplane = g.tile_plane ? g.tile_plane[blockIdx.x] : -1;
    float acc[KC];
#pragma unroll
    for (int c = 0; c < KC; ++c) acc[c] = 0.0f;
    for (int64_t jb = jb0; jb < jend; jb += kTile) {
        if (tplane >= 0 && g.tile_plane[jb / kTile] == tplane) continue;
        const int64_t j = jb + tid;
        __syncthreads();
        if (j < jend) {
            s_pos[tid] = make_float4(g.xf[j], g.yf[j], g.zf[j], 0.0f);
            const double wj = g.w[j];
#pragma unroll
            for (int c = 0; c < KC; ++c) s_q[tid][c] = (k0 + c < K) ? (float)(wj * X[j + (int64_t)(k0 + c) * ldx]) : 0.0f;
        }
        __syncthreads();
        const int nt = (int)min<int64_t>(kTile, jend - jb);
        // source panels jb / 16 .. jb / 16 + 15 (jb is a multiple of 256): bits 0-15 or 16-31 of one word
        const uint32_t mbits =
            EXCL ? g.near_mask[(ii >> 4) * (int64_t)g.mask_words + (jb >> 9)] >> ((jb >> 4) & 31) : 0u;
        for (int jj = 0; jj < nt; ++jj) {
            const float4 sp = s_pos[jj];
            const float dx = tx - sp.x, dy = ty - sp.y, dz = tz - sp.z;
            const float r2 = fmaf(dz, dz, fmaf(dy, dy, fmaf(dx, dx, 1e-20f)));
            const float rinv = __builtin_amdgcn_rsqf(r2);
            float gk = (tnx * dx + tny * dy + tnz * dz) * (rinv * rinv * rinv);
            if (EXCL) gk = ((mbits >> (jj >> 4)) & 1u) ? 0.0f : gk;
            if constexpr (KC % 4 == 0) {
#pragma unroll
                for (int c = 0; c < KC; c += 4) {
                    const float4 q = *reinterpret_cast<const float4 *>(&s_q[jj][c]);
                    acc[c] += gk * q.x;
                    acc[c + 1] += gk * q.y;
                    acc[c + 2] += gk * q.z;
                    acc[c + 3] += gk * q.w;
                }
            } else {
#pragma unroll
                for (int c = 0; c < KC; ++c) acc[c] += gk * s_q[jj][c];
            }
        }
    }
    if (!valid) return;
#pragma unroll
    for (int c = 0; c < KC; ++c)
        if (k0 + c < K) part[((int64_t)blockIdx.z * K + k0 + c) * N + i] = (double)acc[c] * kInv4Pi;
}

// Split count: enough workgroups to cover 256 CUs several times, >= 4 source tiles per split.
int choose_splits(int64_t N, int64_t groups) {
    const int64_t target = 2048;
    int64_t s = (target + groups - 1) / groups;
    const int64_t max_s = std::max<int64_t>(1, N / (4 * kTile));
    s = std::min(std::max<int64_t>(s, 1), std::min<int64_t>(max_s, 64));
    return (int)s;
}

template <class Launch>
int far_launch(int64_t N, int K, int KC, int64_t part_cap, int forced, Launch &&launch) {
    const int64_t bx = (N + kTile - 1) / kTile;
    const int by = (K + KC - 1) / KC;
    int S = forced > 0 ? (int)std::min<int64_t>(forced, std::max<int64_t>(1, N / kTile)) : choose_splits(N, bx * by);
    while (S > 1 && (int64_t)S * K * N > part_cap) --S;
    // splits start on tile boundaries so that the plane table applies
    int64_t per = (N + S - 1) / S;
    per = (per + kTile - 1) / kTile * kTile;
    S = (int)((N + per - 1) / per);
    launch(dim3((unsigned)bx, (unsigned)by, (unsigned)S), (int)per);
    return S;
}
}  // namespace

int64_t far_field_part_elems(int64_t N, int K) {
    const int64_t bx = (N + kTile - 1) / kTile;
    return (int64_t)choose_splits(N, bx) * std::max(K, 1) * N + 64 * N;
}

int far_field_f64(hipStream_t st, const DevGeom &g, FarMode mode, int K, const double *X, int64_t ldx,
                  double *part, int64_t part_cap, const FarTune &tune) {
    const int KC = K >= 8 ? 8 : (K >= 4 ? 4 : (K >= 2 ? 2 : 1));
    return far_launch(g.N, K, KC, part_cap, tune.splits, [&](dim3 grid, int per) {
#define DBIE_FAR(M, C) far_f64_kernel<M, C><<<grid, kTile, 0, st>>>(g, K, X, ldx, part, per)
#define DBIE_FAR_KC(M)           \
    switch (KC) {                \
    case 1: DBIE_FAR(M, 1); break; \
    case 2: DBIE_FAR(M, 2); break; \
    case 4: DBIE_FAR(M, 4); break; \
    default: DBIE_FAR(M, 8); break; \
    }
        if (mode == FAR_DT) { DBIE_FAR_KC(FAR_DT) }
        else if (mode == FAR_S) { DBIE_FAR_KC(FAR_S) }
        else { DBIE_FAR_KC(FAR_D) }
#undef DBIE_FAR_KC
#undef DBIE_FAR
    });
}

int far_field_f32(hipStream_t st, const DevGeom &g, int K, const double *X, int64_t ldx, double *part,
                  int64_t part_cap, const FarTune &tune, bool exclude_near) {
    if (exclude_near && !(g.near_mask && g.npp == 16))
        throw std::logic_error("far_field_f32: near-pair exclusion needs 16-node panels and the near mask");
    // charges per source and workgroup: up to 16, so that K <= 16 computes each pair's kernel once
    const int KC = K > 8 ? 16 : (K > 4 ? 8 : (K >= 3 ? 4 : (K == 2 ? 2 : 1)));
    return far_launch(g.N, K, KC, part_cap, tune.splits, [&](dim3 grid, int per) {
#define DBIE_F32(C)                                                                          \
    case C:                                                                                  \
        if (exclude_near) far_f32_kernel<C, true><<<grid, kTile, 0, st>>>(g, K, X, ldx, part, per);  \
        else far_f32_kernel<C, false><<<grid, kTile, 0, st>>>(g, K, X, ldx, part, per);              \
        break;
        switch (KC) {
            DBIE_F32(1) DBIE_F32(2) DBIE_F32(4) DBIE_F32(8) DBIE_F32(16)
        }
#undef DBIE_F32
    });
}

}  // namespace dbie
