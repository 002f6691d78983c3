// far_mfma.hip -- far-field Dᵀ x RHS-block contraction on the CDNA4 matrix cores.
//
// Y[i,k] = sum_j G(i,j) Z[j,k],  G(i,j) = n_i.(x_i - y_j) / r^3 (self pair 0),  Z = w .* X.
// The kernel matrix G is never stored: every wave generates its A-operand tile
// on VALU, directly in the MFMA register layout, and contracts it with the
// RHS block staged in LDS.  One lane owns one (target, source) entry per MFMA
// k-slice, so generation needs no LDS round trip.
//
//   FP64: v_mfma_f64_16x16x4_f64   (A: 16 targets x 4 sources, lane l -> (l%16, l/16))
//   BF16: v_mfma_f32_16x16x32_bf16 (A: 16 targets x 32 sources, lane l -> (l%16, 8*(l/16)+0..7))
//         with FP32 generation and FP32 accumulation; the BF16X3 form adds the
//         hi*lo + lo*hi cross products of a two-term bf16 split of G and Z.
// D (16 x 16) lands in lane l, register r at col l%16 and row 4*r + l/16 for the FP64
// op but row 4*(l/16) + r for the BF16 op (both measured with a one-hot probe on gfx950).
//
// Workgroup: 4 waves x 32 targets = 128 targets; a 16-column block per MFMA, NCB blocks
// per workgroup.  Source tiles whose nodes all share the target block's plane are
// skipped (exact zeros).  Each source split writes its own partial sum (transposed
// through LDS so that the stores are coalesced along targets).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "kernels.hpp"

namespace dbie {

namespace {
typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

constexpr double kInv4Pi = 1.0 / (4.0 * 3.14159265358979311599796346854);
constexpr int kTargets = 128;  // targets per workgroup
constexpr int kOutStride = kTargets + 4;

// ---------------------------------------------------------------- FP64
constexpr int kTS64 = 32;         // sources per LDS stage
constexpr int kZS64 = kTS64 + 4;  // row stride (doubles) of the [col][source] tile

template <int NCB>
__global__ __launch_bounds__(256) void far_mfma_f64_kernel(DevGeom g, int K, const double *__restrict__ Z,
                                                           double *__restrict__ part, int src_per_split) {
    constexpr int CB = NCB * 16;
    constexpr int kSmem = (CB * kZS64 > 16 * kOutStride) ? CB * kZS64 : 16 * kOutStride;
    __shared__ double4 s_pos[kTS64];
    __shared__ double s_mem[kSmem];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int64_t N = g.N;
    const int64_t i0 = (int64_t)blockIdx.x * kTargets;
    const int k0 = blockIdx.y * CB;
    const int64_t jb0 = (int64_t)blockIdx.z * src_per_split;
    const int64_t jend = min(N, jb0 + src_per_split);

    double tx[2], ty[2], tz[2], tn0[2], tn1[2], tn2[2];
#pragma unroll
    for (int a = 0; a < 2; ++a) {
        int64_t t = i0 + wave * 32 + a * 16 + (lane & 15);
        t = t < N ? t : N - 1;
        tx[a] = g.x[t]; ty[a] = g.y[t]; tz[a] = g.z[t];
        tn0[a] = g.nx[t]; tn1[a] = g.ny[t]; tn2[a] = g.nz[t];
    }
    v4d acc[2][NCB];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int cb = 0; cb < NCB; ++cb) acc[a][cb] = v4d{0.0, 0.0, 0.0, 0.0};
    const int tplane = g.tile_plane ? g.tile_plane[i0 / 256] : -1;

    for (int64_t jb = jb0; jb < jend; jb += kTS64) {
        if (tplane >= 0 && g.tile_plane[jb / 256] == tplane) continue;  // uniform per workgroup
        __syncthreads();
        if (tid < kTS64) {
            const int64_t j = jb + tid;
            s_pos[tid] = j < jend ? make_double4(g.x[j], g.y[j], g.z[j], 1.0) : make_double4(0.0, 0.0, 0.0, 0.0);
        }
        {
            const int s = tid & (kTS64 - 1);
            const int64_t j = jb + s;
            for (int c = tid / kTS64; c < CB; c += 256 / kTS64) {
                const int col = k0 + c;
                s_mem[c * kZS64 + s] = (j < jend && col < K) ? Z[j + (int64_t)col * N] : 0.0;
            }
        }
        __syncthreads();
#pragma unroll 2
        for (int ks = 0; ks < kTS64 / 4; ++ks) {
            const int sidx = ks * 4 + (lane >> 4);
            const double4 sp = s_pos[sidx];
            double gv[2];
#pragma unroll
            for (int a = 0; a < 2; ++a) {
                const double dx = tx[a] - sp.x, dy = ty[a] - sp.y, dz = tz[a] - sp.z;
                const double r2 = dx * dx + dy * dy + dz * dz;
                const double rinv = (r2 > 0.0 && sp.w != 0.0) ? rsqrt(r2) : 0.0;
                gv[a] = (tn0[a] * dx + tn1[a] * dy + tn2[a] * dz) * (rinv * rinv * rinv);
            }
#pragma unroll
            for (int cb = 0; cb < NCB; ++cb) {
                const double b = s_mem[(cb * 16 + (lane & 15)) * kZS64 + sidx];
                acc[0][cb] = __builtin_amdgcn_mfma_f64_16x16x4f64(gv[0], b, acc[0][cb], 0, 0, 0);
                acc[1][cb] = __builtin_amdgcn_mfma_f64_16x16x4f64(gv[1], b, acc[1][cb], 0, 0, 0);
            }
        }
    }
    // transpose each 16-column block through LDS; coalesced stores along targets
#pragma unroll
    for (int cb = 0; cb < NCB; ++cb) {
        __syncthreads();
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int r = 0; r < 4; ++r)
                s_mem[(lane & 15) * kOutStride + wave * 32 + a * 16 + 4 * r + (lane >> 4)] = acc[a][cb][r];
        __syncthreads();
        const int t = tid & (kTargets - 1);
        const int64_t i = i0 + t;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int c = (tid >> 7) + 2 * q;
            const int col = k0 + cb * 16 + c;
            if (col < K && i < N) part[((int64_t)blockIdx.z * K + col) * N + i] = s_mem[c * kOutStride + t] * kInv4Pi;
        }
    }
}

// ---------------------------------------------------------------- BF16
constexpr int kTSb = 64;         // sources per LDS stage (two k-slices of 32)
constexpr int kZSb = kTSb + 8;   // row stride (bf16) of the [col][source] tile

// Z = w .* X split into bf16 hi + lo, stored tile-major: [j / 64][col][j % 64], Kp columns.
__global__ void split_bf16_kernel(const double *__restrict__ X, int64_t ldx, const double *__restrict__ w, int64_t N,
                                  int K, int Kp, __bf16 *__restrict__ zh, __bf16 *__restrict__ zl) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y;
    const int64_t ntile = (N + kTSb - 1) / kTSb;
    if (j >= ntile * kTSb) return;
    float z = 0.0f, lo = 0.0f;
    __bf16 h = (__bf16)0.0f;
    if (j < N && k < K) {
        const double zd = w[j] * X[j + (int64_t)k * ldx];
        z = (float)zd;
        h = (__bf16)z;
        lo = (float)(zd - (double)(float)h);
    }
    const int64_t idx = ((j / kTSb) * Kp + k) * kTSb + (j % kTSb);
    zh[idx] = h;
    if (zl) zl[idx] = (__bf16)lo;
}

template <int NCB, bool X3>
__global__ __launch_bounds__(256) void far_mfma_bf16_kernel(DevGeom g, int K, int Kp, const __bf16 *__restrict__ zh,
                                                            const __bf16 *__restrict__ zl, double *__restrict__ part,
                                                            int src_per_split) {
    constexpr int CB = NCB * 16;
    constexpr int kZ = CB * kZSb;  // bf16 elements per LDS tile
    constexpr int kOutF = 16 * kOutStride;
    constexpr int kSmemBytes = ((X3 ? 2 : 1) * kZ * 2 > kOutF * 4) ? (X3 ? 2 : 1) * kZ * 2 : kOutF * 4;
    __shared__ float4 s_pos[kTSb];
    __shared__ __attribute__((aligned(16))) unsigned char s_raw[kSmemBytes];
    __bf16 *s_zh = reinterpret_cast<__bf16 *>(s_raw);
    __bf16 *s_zl = s_zh + kZ;
    float *s_out = reinterpret_cast<float *>(s_raw);

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int64_t N = g.N;
    const int64_t i0 = (int64_t)blockIdx.x * kTargets;
    const int k0 = blockIdx.y * CB;
    const int64_t jb0 = (int64_t)blockIdx.z * src_per_split;
    const int64_t jend = min(N, jb0 + src_per_split);

    float tx[2], ty[2], tz[2], tn0[2], tn1[2], tn2[2];
#pragma unroll
    for (int a = 0; a < 2; ++a) {
        int64_t t = i0 + wave * 32 + a * 16 + (lane & 15);
        t = t < N ? t : N - 1;
        tx[a] = g.xf[t]; ty[a] = g.yf[t]; tz[a] = g.zf[t];
        tn0[a] = g.nxf[t]; tn1[a] = g.nyf[t]; tn2[a] = g.nzf[t];
    }
    v4f acc[2][NCB];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int cb = 0; cb < NCB; ++cb) acc[a][cb] = v4f{0.f, 0.f, 0.f, 0.f};
    const int tplane = g.tile_plane ? g.tile_plane[i0 / 256] : -1;

    for (int64_t jb = jb0; jb < jend; jb += kTSb) {
        if (tplane >= 0 && g.tile_plane[jb / 256] == tplane) continue;  // uniform per workgroup
        __syncthreads();
        if (tid < kTSb) {
            const int64_t j = jb + tid;
            s_pos[tid] = j < jend ? make_float4(g.xf[j], g.yf[j], g.zf[j], 1.0f) : make_float4(0.f, 0.f, 0.f, 0.f);
        }
        {
            // tile-major global layout: the (CB x 64) block of this tile is contiguous per column
            const int64_t tile = jb / kTSb;
            const uint4 *gh = reinterpret_cast<const uint4 *>(zh + (tile * Kp + k0) * kTSb);
            const uint4 *gl = X3 ? reinterpret_cast<const uint4 *>(zl + (tile * Kp + k0) * kTSb) : nullptr;
            for (int e = tid; e < CB * (kTSb / 8); e += 256) {
                const int c = e / (kTSb / 8), q = e % (kTSb / 8);
                *reinterpret_cast<uint4 *>(s_zh + c * kZSb + 8 * q) = gh[e];
                if (X3) *reinterpret_cast<uint4 *>(s_zl + c * kZSb + 8 * q) = gl[e];
            }
        }
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < kTSb / 32; ++ks) {
            const int sbase = ks * 32 + 8 * (lane >> 4);
            v8bf ah[2], al[2];
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const float4 sp = s_pos[sbase + e];
#pragma unroll
                for (int a = 0; a < 2; ++a) {
                    const float dx = tx[a] - sp.x, dy = ty[a] - sp.y, dz = tz[a] - sp.z;
                    const float r2 = dx * dx + dy * dy + dz * dz;
                    const float rinv = (r2 > 0.0f && sp.w != 0.0f) ? __frsqrt_rn(r2) : 0.0f;
                    const float gk = (tn0[a] * dx + tn1[a] * dy + tn2[a] * dz) * (rinv * rinv * rinv);
                    const __bf16 h = (__bf16)gk;
                    ah[a][e] = h;
                    if (X3) al[a][e] = (__bf16)(gk - (float)h);
                }
            }
#pragma unroll
            for (int cb = 0; cb < NCB; ++cb) {
                const int off = (cb * 16 + (lane & 15)) * kZSb + sbase;
                const v8bf bh = *reinterpret_cast<const v8bf *>(s_zh + off);
                acc[0][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[0], bh, acc[0][cb], 0, 0, 0);
                acc[1][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[1], bh, acc[1][cb], 0, 0, 0);
                if (X3) {
                    const v8bf bl = *reinterpret_cast<const v8bf *>(s_zl + off);
                    acc[0][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[0], bl, acc[0][cb], 0, 0, 0);
                    acc[1][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[1], bl, acc[1][cb], 0, 0, 0);
                    acc[0][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[0], bh, acc[0][cb], 0, 0, 0);
                    acc[1][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[1], bh, acc[1][cb], 0, 0, 0);
                }
            }
        }
    }
#pragma unroll
    for (int cb = 0; cb < NCB; ++cb) {
        __syncthreads();
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int r = 0; r < 4; ++r)
                s_out[(lane & 15) * kOutStride + wave * 32 + a * 16 + 4 * (lane >> 4) + r] = acc[a][cb][r];
        __syncthreads();
        const int t = tid & (kTargets - 1);
        const int64_t i = i0 + t;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int c = (tid >> 7) + 2 * q;
            const int col = k0 + cb * 16 + c;
            if (col < K && i < N) part[((int64_t)blockIdx.z * K + col) * N + i] = (double)s_out[c * kOutStride + t] * kInv4Pi;
        }
    }
}

// Source split count for a launch of `groups` workgroups (splits start on 256-node tiles).
int mfma_splits(int64_t N, int K, int64_t groups, int64_t part_cap, int64_t *per_out) {
    int64_t s = std::max<int64_t>(1, (1536 + groups - 1) / groups);
    s = std::min<int64_t>(s, std::max<int64_t>(1, N / 2048));
    while (s > 1 && s * K * N > part_cap) --s;
    int64_t per = (N + s - 1) / s;
    per = (per + 255) / 256 * 256;
    *per_out = per;
    return (int)((N + per - 1) / per);
}

__global__ void scale_rows_kernel(const double *__restrict__ X, int64_t ldx, const double *__restrict__ w, int64_t N,
                                  double *__restrict__ Z) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y;
    if (j < N) Z[j + (int64_t)k * N] = w[j] * X[j + (int64_t)k * ldx];
}
}  // namespace

bool far_field_mfma_supported(int prec, int K) {
    if (K < 16) return false;
    return prec == 0 || prec == 2 || prec == 3;
}

size_t far_field_mfma_scratch_bytes(int64_t N, int prec, int K) {
    if (prec == 0) return sizeof(double) * (size_t)N * K;
    const int ncb = std::min(16, (K + 15) / 16);
    const int CB = ncb * 16;
    const int Kp = ((K + CB - 1) / CB) * CB;
    const size_t zelems = (size_t)((N + kTSb - 1) / kTSb) * Kp * kTSb;
    return sizeof(__bf16) * zelems * (prec == 3 ? 2 : 1);
}

int far_field_mfma(hipStream_t st, const DevGeom &g, int prec, int K, const double *X, int64_t ldx, double *part,
                   int64_t part_cap, void *scratch) {
    const int64_t N = g.N;
    const int64_t bx = (N + kTargets - 1) / kTargets;
    if (prec == 0) {
        const int ncb = std::min(8, (K + 15) / 16);
        const int CB = ncb * 16;
        const int by = (K + CB - 1) / CB;
        int64_t per;
        const int S = mfma_splits(N, K, bx * by, part_cap, &per);
        double *Z = static_cast<double *>(scratch);
        scale_rows_kernel<<<dim3((unsigned)((N + 255) / 256), (unsigned)K), 256, 0, st>>>(X, ldx, g.w, N, Z);
        dim3 grid((unsigned)bx, (unsigned)by, (unsigned)S);
        switch (ncb) {
#define DBIE_F64(n) case n: far_mfma_f64_kernel<n><<<grid, 256, 0, st>>>(g, K, Z, part, (int)per); break;
            DBIE_F64(1) DBIE_F64(2) DBIE_F64(3) DBIE_F64(4) DBIE_F64(5) DBIE_F64(6) DBIE_F64(7) DBIE_F64(8)
#undef DBIE_F64
        }
        return S;
    }
    const bool x3 = prec == 3;
    const int ncb = std::min(16, (K + 15) / 16);
    const int CB = ncb * 16;
    const int by = (K + CB - 1) / CB;
    const int Kp = by * CB;
    const int64_t ntile = (N + kTSb - 1) / kTSb;
    const size_t zelems = (size_t)ntile * Kp * kTSb;
    __bf16 *zh = static_cast<__bf16 *>(scratch);
    __bf16 *zl = x3 ? zh + zelems : nullptr;
    split_bf16_kernel<<<dim3((unsigned)((ntile * kTSb + 255) / 256), (unsigned)Kp), 256, 0, st>>>(X, ldx, g.w, N, K, Kp,
                                                                                                   zh, zl);
    int64_t per;
    const int S = mfma_splits(N, K, bx * by, part_cap, &per);
    dim3 grid((unsigned)bx, (unsigned)by, (unsigned)S);
    if (x3) {
        switch (ncb) {
#define DBIE_B3(n) case n: far_mfma_bf16_kernel<n, true><<<grid, 256, 0, st>>>(g, K, Kp, zh, zl, part, (int)per); break;
            DBIE_B3(1) DBIE_B3(2) DBIE_B3(3) DBIE_B3(4) DBIE_B3(5) DBIE_B3(6) DBIE_B3(7) DBIE_B3(8)
            DBIE_B3(9) DBIE_B3(10) DBIE_B3(11) DBIE_B3(12) DBIE_B3(13) DBIE_B3(14) DBIE_B3(15) DBIE_B3(16)
#undef DBIE_B3
        }
    } else {
        switch (ncb) {
#define DBIE_B1(n) case n: far_mfma_bf16_kernel<n, false><<<grid, 256, 0, st>>>(g, K, Kp, zh, nullptr, part, (int)per); break;
            DBIE_B1(1) DBIE_B1(2) DBIE_B1(3) DBIE_B1(4) DBIE_B1(5) DBIE_B1(6) DBIE_B1(7) DBIE_B1(8)
            DBIE_B1(9) DBIE_B1(10) DBIE_B1(11) DBIE_B1(12) DBIE_B1(13) DBIE_B1(14) DBIE_B1(15) DBIE_B1(16)
#undef DBIE_B1
        }
    }
    return S;
}

}  // namespace dbie
