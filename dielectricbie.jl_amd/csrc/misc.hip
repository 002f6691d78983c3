// misc.hip -- right-hand sides, off-surface potential, dense assembly and the
// Krylov primitives of the GMRES loop (all FP64, gfx950).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "kernels.hpp"

namespace dbie {

namespace {
constexpr double kInv4Pi = 1.0 / (4.0 * 3.14159265358979311599796346854);
constexpr int kTile = 256;

// rhs_dielectric_box3d(interface, PointSource, eps_src) (src/solver/dielectric_box3d.jl:95-104)
__global__ __launch_bounds__(256) void rhs_points_kernel(DevGeom g, int K, const double *__restrict__ src,
                                                         const double *__restrict__ q, double inv_eps,
                                                         double *__restrict__ F, int64_t ldf) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y;
    if (i >= g.N || k >= K) return;
    const double dx = g.x[i] - src[3 * k], dy = g.y[i] - src[3 * k + 1], dz = g.z[i] - src[3 * k + 2];
    const double r2 = dx * dx + dy * dy + dz * dz;
    const double rinv = rsqrt(r2);
    const double nd = g.nx[i] * dx + g.ny[i] * dy + g.nz[i] * dz;
    F[i + (int64_t)k * ldf] = -q[k] * (nd * rinv * rinv * rinv) * kInv4Pi * inv_eps;
}

// rhs_dielectric_box3d(interface, VolumeSource, eps_src) (src/solver/dielectric_box3d.jl:106-120):
// column k sums the source points col_ptr[k] .. col_ptr[k+1]-1 with charges wrho = w.*rho.
__global__ __launch_bounds__(kTile) void rhs_volume_kernel(DevGeom g, const int64_t *__restrict__ col_ptr,
                                                           const double *__restrict__ pos,
                                                           const double *__restrict__ wrho, double scale,
                                                           double *__restrict__ F, int64_t ldf) {
    __shared__ double4 s_src[kTile];
    const int tid = threadIdx.x;
    const int64_t i = (int64_t)blockIdx.x * kTile + tid;
    const int k = blockIdx.y;
    const int64_t ii = i < g.N ? i : g.N - 1;
    const double tx = g.x[ii], ty = g.y[ii], tz = g.z[ii], nx = g.nx[ii], ny = g.ny[ii], nz = g.nz[ii];
    const int64_t s0 = col_ptr[k], s1 = col_ptr[k + 1];
    double acc = 0.0;
    for (int64_t sb = s0; sb < s1; sb += kTile) {
        __syncthreads();
        const int64_t s = sb + tid;
        if (s < s1) s_src[tid] = make_double4(pos[3 * s], pos[3 * s + 1], pos[3 * s + 2], wrho[s]);
        __syncthreads();
        const int nt = (int)min<int64_t>(kTile, s1 - sb);
        for (int m = 0; m < nt; ++m) {
            const double4 sp = s_src[m];
            const double dx = tx - sp.x, dy = ty - sp.y, dz = tz - sp.z;
            const double r2 = dx * dx + dy * dy + dz * dz;
            const double rinv = rsqrt(r2);
            acc += sp.w * ((nx * dx + ny * dy + nz * dz) * (rinv * rinv * rinv));
        }
    }
    if (i < g.N) F[i + (int64_t)k * ldf] = -acc * kInv4Pi * scale;
}

// SharpScreening through the interface box (src/core/sources.jl:315-336,350-366):
// wrho_s = w_s * (rho_s / eps(y_s)), eps = eps_in inside the tolerance-padded box.
__global__ void screen_sharp_kernel(int64_t S, const double *__restrict__ pos, const double *__restrict__ rho,
                                    const double *__restrict__ w, double lo0, double lo1, double lo2, double hi0,
                                    double hi1, double hi2, double tol, double eps_in, double eps_out,
                                    double *__restrict__ wrho) {
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= S) return;
    const double x = pos[3 * s], y = pos[3 * s + 1], z = pos[3 * s + 2];
    const bool in = (lo0 - tol <= x && x <= hi0 + tol) && (lo1 - tol <= y && y <= hi1 + tol) &&
                    (lo2 - tol <= z && z <= hi2 + tol);
    wrho[s] = w[s] * (rho[s] / (in ? eps_in : eps_out));
}

// (1 + erf(d / bandwidth)) / 2 (_smooth_screen_factor, src/core/sources.jl:321-323)
__device__ inline double smooth_step(double d, double bw) { return (1.0 + erf(d / bw)) / 2.0; }

// Screened density rho / eps(y) of source points (src/core/sources.jl:315-349 for one region,
// :408-493 for several boxes).  Sharp: eps of the first box whose tol-widened closure holds the
// point, else eps_out.  Soft: the boxes' smooth indicators P_b (products of six error-function
// steps) mix eps (or 1/eps) with eps_out; each sum runs in box order from zero like the reference.
__global__ void screen_boxes_kernel(int64_t S, const double *__restrict__ pos, const double *__restrict__ rho,
                                    const double *__restrict__ w, ScreenBoxes sb, double *__restrict__ out) {
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= S) return;
    const double p[3] = {pos[3 * s], pos[3 * s + 1], pos[3 * s + 2]};
    double r;
    if (sb.mode == SCREEN_SHARP) {
        double eps = sb.eps_out;
        for (int b = 0; b < sb.nb; ++b) {
            const double *lo = sb.lo + 3 * b, *hi = sb.hi + 3 * b;
            if ((lo[0] - sb.tol <= p[0] && p[0] <= hi[0] + sb.tol) && (lo[1] - sb.tol <= p[1] && p[1] <= hi[1] + sb.tol) &&
                (lo[2] - sb.tol <= p[2] && p[2] <= hi[2] + sb.tol)) {
                eps = sb.eps[b];
                break;
            }
        }
        r = rho[s] / eps;
    } else {
        double mix = 0.0, sum_p = 0.0;
        for (int b = 0; b < sb.nb; ++b) {
            const double *lo = sb.lo + 3 * b, *hi = sb.hi + 3 * b;
            const double P = smooth_step(p[0] - lo[0], sb.bandwidth) * smooth_step(hi[0] - p[0], sb.bandwidth) *
                             smooth_step(p[1] - lo[1], sb.bandwidth) * smooth_step(hi[1] - p[1], sb.bandwidth) *
                             smooth_step(p[2] - lo[2], sb.bandwidth) * smooth_step(hi[2] - p[2], sb.bandwidth);
            if (sb.single) {
                mix = P;  // one region: the smooth indicator s itself
            } else {
                mix += sb.mode == SCREEN_SOFT_MIX ? sb.eps[b] * P : P / sb.eps[b];
                sum_p += P;
            }
        }
        if (sb.single) {
            const double eps = sb.mode == SCREEN_SOFT_MIX ? sb.eps[0] * mix + sb.eps_out * (1.0 - mix)
                                                          : 1.0 / (mix / sb.eps[0] + (1.0 - mix) / sb.eps_out);
            r = rho[s] / eps;
        } else if (sb.mode == SCREEN_SOFT_MIX) {
            r = rho[s] / (mix + sb.eps_out * (1.0 - sum_p));
        } else {
            r = rho[s] * (mix + (1.0 - sum_p) / sb.eps_out);
        }
    }
    out[s] = w ? w[s] * r : r;
}

// laplace3d_pottrg (src/kernel/laplace3d.jl:69-82): off-surface targets, KC columns per thread.
template <int KC>
__global__ __launch_bounds__(kTile) void pottrg_kernel(DevGeom g, int64_t M, const double *__restrict__ trg, int K,
                                                       const double *__restrict__ X, int64_t ldx,
                                                       double *__restrict__ U, int64_t ldu) {
    __shared__ double4 s_pos[kTile];
    __shared__ double s_q[KC][kTile];
    const int tid = threadIdx.x;
    const int64_t t = (int64_t)blockIdx.x * kTile + tid;
    const int k0 = blockIdx.y * KC;
    const int64_t tt = t < M ? t : M - 1;
    const double tx = trg[3 * tt], ty = trg[3 * tt + 1], tz = trg[3 * tt + 2];
    double acc[KC];
#pragma unroll
    for (int c = 0; c < KC; ++c) acc[c] = 0.0;
    for (int64_t jb = 0; jb < g.N; jb += kTile) {
        __syncthreads();
        const int64_t j = jb + tid;
        if (j < g.N) {
            s_pos[tid] = make_double4(g.x[j], g.y[j], g.z[j], 0.0);
#pragma unroll
            for (int c = 0; c < KC; ++c) s_q[c][tid] = (k0 + c < K) ? g.w[j] * X[j + (int64_t)(k0 + c) * ldx] : 0.0;
        }
        __syncthreads();
        const int nt = (int)min<int64_t>(kTile, g.N - jb);
        for (int jj = 0; jj < nt; ++jj) {
            const double4 sp = s_pos[jj];
            const double dx = tx - sp.x, dy = ty - sp.y, dz = tz - sp.z;
            const double rinv = rsqrt(dx * dx + dy * dy + dz * dz);
#pragma unroll
            for (int c = 0; c < KC; ++c) acc[c] += rinv * s_q[c][jj];
        }
    }
    if (t >= M) return;
#pragma unroll
    for (int c = 0; c < KC; ++c)
        if (k0 + c < K) U[t + (int64_t)(k0 + c) * ldu] = acc[c] * kInv4Pi;
}

// lhs_dielectric_box3d (src/solver/dielectric_box3d.jl:1-17): A = DT + diag(t)
__global__ void dense_lhs_kernel(DevGeom g, const double *__restrict__ diag, double *__restrict__ A, int64_t lda) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t j = blockIdx.y;
    if (i >= g.N) return;
    double v;
    if (i == j) {
        v = diag[i];
    } else {
        const double dx = g.x[i] - g.x[j], dy = g.y[i] - g.y[j], dz = g.z[i] - g.z[j];
        const double r = sqrt(dx * dx + dy * dy + dz * dz);
        const double ir = 1.0 / r;
        const double ir3 = ir * ir * ir;
        v = ((g.nx[i] * (ir3 * dx) + g.ny[i] * (ir3 * dy) + g.nz[i] * (ir3 * dz)) * kInv4Pi) * g.w[j];
    }
    A[i + j * lda] = v;
}

// ---- Krylov primitives ----
constexpr int kChunk = 4096;  // rows per dot-product workgroup (16 per thread)

__device__ inline double wave_sum(double v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

template <typename TV>
__global__ __launch_bounds__(256) void dots_kernel(const TV *__restrict__ V, int nv, const double *__restrict__ W,
                                                   int64_t N, int K, double *__restrict__ partial) {
    __shared__ double red[4];
    const int64_t c = blockIdx.x;
    const int k = blockIdx.y;
    const int64_t n0 = c * kChunk;
    const int tid = threadIdx.x;
    double wv[kChunk / 256];
    const double *wk = W + (int64_t)k * N;
#pragma unroll
    for (int e = 0; e < kChunk / 256; ++e) {
        const int64_t n = n0 + e * 256 + tid;
        wv[e] = n < N ? wk[n] : 0.0;
    }
    const int64_t nchunks = gridDim.x;
    for (int i = 0; i < nv; ++i) {
        const TV *vk = V + ((int64_t)i * K + k) * N;
        double s = 0.0;
#pragma unroll
        for (int e = 0; e < kChunk / 256; ++e) {
            const int64_t n = n0 + e * 256 + tid;
            if (n < N) s += (double)vk[n] * wv[e];
        }
        s = wave_sum(s);
        if ((tid & 63) == 0) red[tid >> 6] = s;
        __syncthreads();
        if (tid == 0) partial[((int64_t)i * K + k) * nchunks + c] = (red[0] + red[1]) + (red[2] + red[3]);
        __syncthreads();
    }
}

__global__ void dots_reduce_kernel(const double *__restrict__ partial, int64_t nchunks, int64_t count,
                                   double *__restrict__ h) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= count) return;
    double s = 0.0;
    for (int64_t c = 0; c < nchunks; ++c) s += partial[e * nchunks + c];
    h[e] = s;
}

template <int SIGN, typename TV>
__global__ __launch_bounds__(256) void combine_kernel(const TV *__restrict__ V, int nv, const double *__restrict__ y,
                                                      double *__restrict__ W, int64_t N, int K) {
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y;
    if (n >= N) return;
    double s = 0.0;
    for (int i = 0; i < nv; ++i) s += (double)V[((int64_t)i * K + k) * N + n] * y[(int64_t)i * K + k];
    if (SIGN < 0) W[(int64_t)k * N + n] -= s;
    else W[(int64_t)k * N + n] += s;
}

template <typename TO>
__global__ void col_scale_kernel(const double *__restrict__ in, const double *__restrict__ s, int invert,
                                 TO *__restrict__ out, int64_t N, int K) {
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y;
    if (n >= N) return;
    const double v = in[(int64_t)k * N + n];
    double r;
    if (invert) r = s[k] != 0.0 ? v / s[k] : 0.0;
    else r = v * s[k];
    out[(int64_t)k * N + n] = (TO)r;
}

// out[:, k] = in[:, k] / sqrt(sq[k]) (zero where sq[k] <= 0): a new Arnoldi vector from its squared norm
// without a host round trip
template <typename TO>
__global__ void col_normalize_kernel(const double *__restrict__ in, const double *__restrict__ sq,
                                     TO *__restrict__ out, int64_t N, int K) {
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y;
    if (n >= N) return;
    const double s = sq[k];
    out[(int64_t)k * N + n] = (TO)(s > 0.0 ? in[(int64_t)k * N + n] / sqrt(s) : 0.0);
}

// coef[i, k] = a[i, k] - om[i, k] sum_l om[l, k] a[l, k] for i < n1[k], zero otherwise (the projection
// onto a subspace of the span of nb basis blocks given by its unit normal om, column by column)
__global__ void deflate_coef_kernel(const double *__restrict__ a, const double *__restrict__ om,
                                    const int *__restrict__ n1, int nb, int K, double *__restrict__ coef) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= K) return;
    const int n = n1[k];
    double wa = 0.0;
    for (int i = 0; i < n; ++i) wa += om[(int64_t)i * K + k] * a[(int64_t)i * K + k];
    for (int i = 0; i < nb; ++i)
        coef[(int64_t)i * K + k] = i < n ? a[(int64_t)i * K + k] - om[(int64_t)i * K + k] * wa : 0.0;
}

// d may be null: out = in (a conversion to FP64)
template <typename TI>
__global__ void row_scale_kernel(const double *__restrict__ d, const TI *__restrict__ in, int64_t ldi,
                                 double *__restrict__ out, int64_t ldo, int64_t N, int K) {
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y;
    if (n < N) out[(int64_t)k * ldo + n] = (d ? d[n] : 1.0) * (double)in[(int64_t)k * ldi + n];
}

// out = out - z + s .* py (the second Jacobi sweep of the near-field preconditioner)
__global__ void jacobi_update_kernel(const double *__restrict__ s, const double *__restrict__ py,
                                     const double *__restrict__ z, double *__restrict__ out, int64_t ldo, int64_t N,
                                     int K) {
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t idx = (int64_t)blockIdx.y * N + n, io = (int64_t)blockIdx.y * ldo + n;
    if (n < N) out[io] = out[io] - z[idx] + s[n] * py[idx];
}

__global__ void sub_from_kernel(const double *__restrict__ F, double *__restrict__ R, int64_t N, int K) {
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y;
    if (n < N) R[(int64_t)k * N + n] = F[(int64_t)k * N + n] - R[(int64_t)k * N + n];
}

__global__ void add_to_kernel(const double *__restrict__ D, double *__restrict__ X, int64_t N, int K) {
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y;
    if (n < N) X[(int64_t)k * N + n] += D[(int64_t)k * N + n];
}

__global__ void sub_to_kernel(const double *__restrict__ D, double *__restrict__ X, int64_t N, int K) {
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y;
    if (n < N) X[(int64_t)k * N + n] -= D[(int64_t)k * N + n];
}

__global__ void copy_cols_kernel(const double *__restrict__ src, int64_t lds, double *__restrict__ dst, int64_t ldd,
                                 int64_t N) {
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y;
    if (n < N) dst[n + (int64_t)k * ldd] = src[n + (int64_t)k * lds];
}

// dst(0:n, 0:n) = upper triangle of src(0:n, 0:n), zeros below the diagonal
__global__ void triu_copy_kernel(int n, const double *__restrict__ src, int64_t lds, double *__restrict__ dst,
                                 int64_t ldd) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    const int c = blockIdx.y;
    if (r < n) dst[r + (int64_t)c * ldd] = r <= c ? src[r + (int64_t)c * lds] : 0.0;
}

// dst(0:rows, 0:cols) = the unit lower trapezoid of src (ones on the diagonal, zeros above): the
// Householder vectors that geqrf leaves below R
__global__ void unit_lower_copy_kernel(int rows, const double *__restrict__ src, int64_t lds, double *__restrict__ dst,
                                       int64_t ldd) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    const int c = blockIdx.y;
    if (r < rows) dst[r + (int64_t)c * ldd] = r > c ? src[r + (int64_t)c * lds] : (r == c ? 1.0 : 0.0);
}

// C(i, j) = beta C(i, j) + sum_b P_b(i, j) over nb partial p x q blocks stored back to back
__global__ void sum_blocks_kernel(int nb, int p, int q, const double *__restrict__ P, double beta,
                                  double *__restrict__ C, int64_t ldc) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t pq = (int64_t)p * q;
    if (e >= pq) return;
    double s = 0.0;
    for (int b = 0; b < nb; ++b) s += P[(int64_t)b * pq + e];
    const int64_t i = e % p, j = e / p;
    double *c = C + i + j * ldc;
    *c = beta == 0.0 ? s : beta * *c + s;
}

inline dim3 grid2(int64_t N, int K) { return dim3((unsigned)((N + 255) / 256), (unsigned)K); }
}  // namespace

void rhs_points(hipStream_t st, const DevGeom &g, int K, const double *src, const double *q, double eps_src,
                double *F, int64_t ldf) {
    rhs_points_kernel<<<grid2(g.N, K), 256, 0, st>>>(g, K, src, q, 1.0 / eps_src, F, ldf);
}

void rhs_volume(hipStream_t st, const DevGeom &g, int K, const int64_t *col_ptr, const double *pos,
                const double *wrho, double *F, int64_t ldf, double scale) {
    rhs_volume_kernel<<<grid2(g.N, K), kTile, 0, st>>>(g, col_ptr, pos, wrho, scale, F, ldf);
}

void screen_sharp(hipStream_t st, int64_t S, const double *pos, const double *rho, const double *w, const double *lo,
                  const double *hi, double tol, double eps_in, double eps_out, double *wrho) {
    if (S <= 0) return;
    screen_sharp_kernel<<<(unsigned)((S + 255) / 256), 256, 0, st>>>(S, pos, rho, w, lo[0], lo[1], lo[2], hi[0],
                                                                      hi[1], hi[2], tol, eps_in, eps_out, wrho);
}

void screen_boxes(hipStream_t st, int64_t S, const double *pos, const double *rho, const double *w,
                  const ScreenBoxes &sb, double *out) {
    if (S <= 0) return;
    screen_boxes_kernel<<<(unsigned)((S + 255) / 256), 256, 0, st>>>(S, pos, rho, w, sb, out);
}

void pottrg(hipStream_t st, const DevGeom &g, int64_t M, const double *trg, int K, const double *X, int64_t ldx,
            double *U, int64_t ldu) {
    const int KC = K >= 8 ? 8 : (K >= 4 ? 4 : 1);
    dim3 grid((unsigned)((M + kTile - 1) / kTile), (unsigned)((K + KC - 1) / KC));
    if (KC == 8) pottrg_kernel<8><<<grid, kTile, 0, st>>>(g, M, trg, K, X, ldx, U, ldu);
    else if (KC == 4) pottrg_kernel<4><<<grid, kTile, 0, st>>>(g, M, trg, K, X, ldx, U, ldu);
    else pottrg_kernel<1><<<grid, kTile, 0, st>>>(g, M, trg, K, X, ldx, U, ldu);
}

void dense_lhs(hipStream_t st, const DevGeom &g, const double *diag, double *A, int64_t lda) {
    dense_lhs_kernel<<<dim3((unsigned)((g.N + 255) / 256), (unsigned)g.N), 256, 0, st>>>(g, diag, A, lda);
}

int64_t dots_partial_elems(int64_t N, int K, int nv) { return ((N + kChunk - 1) / kChunk) * (int64_t)K * nv; }

template <typename TV>
void block_dots_t(hipStream_t st, const TV *V, int nv, const double *W, int64_t N, int K, double *partial, double *h) {
    const int64_t nchunks = (N + kChunk - 1) / kChunk;
    dots_kernel<TV><<<dim3((unsigned)nchunks, (unsigned)K), 256, 0, st>>>(V, nv, W, N, K, partial);
    const int64_t count = (int64_t)nv * K;
    dots_reduce_kernel<<<(unsigned)((count + 255) / 256), 256, 0, st>>>(partial, nchunks, count, h);
}
void block_dots(hipStream_t st, const double *V, int nv, const double *W, int64_t N, int K, double *partial,
                double *h) {
    block_dots_t(st, V, nv, W, N, K, partial, h);
}
void block_dots(hipStream_t st, const float *V, int nv, const double *W, int64_t N, int K, double *partial,
                double *h) {
    block_dots_t(st, V, nv, W, N, K, partial, h);
}

void block_subtract(hipStream_t st, const double *V, int nv, const double *h, double *W, int64_t N, int K) {
    combine_kernel<-1, double><<<grid2(N, K), 256, 0, st>>>(V, nv, h, W, N, K);
}
void block_subtract(hipStream_t st, const float *V, int nv, const double *h, double *W, int64_t N, int K) {
    combine_kernel<-1, float><<<grid2(N, K), 256, 0, st>>>(V, nv, h, W, N, K);
}

void block_combine(hipStream_t st, const double *V, int nv, const double *y, double *X, int64_t N, int K) {
    combine_kernel<1, double><<<grid2(N, K), 256, 0, st>>>(V, nv, y, X, N, K);
}
void block_combine(hipStream_t st, const float *V, int nv, const double *y, double *X, int64_t N, int K) {
    combine_kernel<1, float><<<grid2(N, K), 256, 0, st>>>(V, nv, y, X, N, K);
}

void col_scale(hipStream_t st, const double *in, const double *s, bool invert, double *out, int64_t N, int K) {
    col_scale_kernel<double><<<grid2(N, K), 256, 0, st>>>(in, s, invert ? 1 : 0, out, N, K);
}
void col_scale(hipStream_t st, const double *in, const double *s, bool invert, float *out, int64_t N, int K) {
    col_scale_kernel<float><<<grid2(N, K), 256, 0, st>>>(in, s, invert ? 1 : 0, out, N, K);
}

void col_normalize(hipStream_t st, const double *in, const double *sq, double *out, int64_t N, int K) {
    col_normalize_kernel<double><<<grid2(N, K), 256, 0, st>>>(in, sq, out, N, K);
}
void col_normalize(hipStream_t st, const double *in, const double *sq, float *out, int64_t N, int K) {
    col_normalize_kernel<float><<<grid2(N, K), 256, 0, st>>>(in, sq, out, N, K);
}

void deflate_coef(hipStream_t st, const double *a, const double *om, const int *n1, int nb, int K, double *coef) {
    deflate_coef_kernel<<<(unsigned)((K + 63) / 64), 64, 0, st>>>(a, om, n1, nb, K, coef);
}

void row_scale(hipStream_t st, const double *d, const double *in, double *out, int64_t N, int K, int64_t ldi,
               int64_t ldo) {
    row_scale_kernel<double><<<grid2(N, K), 256, 0, st>>>(d, in, ldi > 0 ? ldi : N, out, ldo > 0 ? ldo : N, N, K);
}
void row_scale(hipStream_t st, const double *d, const float *in, double *out, int64_t N, int K, int64_t ldi,
               int64_t ldo) {
    row_scale_kernel<float><<<grid2(N, K), 256, 0, st>>>(d, in, ldi > 0 ? ldi : N, out, ldo > 0 ? ldo : N, N, K);
}

void jacobi_update(hipStream_t st, const double *s, const double *py, const double *z, double *out, int64_t N, int K,
                   int64_t ldo) {
    jacobi_update_kernel<<<grid2(N, K), 256, 0, st>>>(s, py, z, out, ldo > 0 ? ldo : N, N, K);
}

void sub_from(hipStream_t st, const double *F, double *R, int64_t N, int K) {
    sub_from_kernel<<<grid2(N, K), 256, 0, st>>>(F, R, N, K);
}

void add_to(hipStream_t st, const double *D, double *X, int64_t N, int K) {
    add_to_kernel<<<grid2(N, K), 256, 0, st>>>(D, X, N, K);
}

void sub_to(hipStream_t st, const double *D, double *X, int64_t N, int K) {
    sub_to_kernel<<<grid2(N, K), 256, 0, st>>>(D, X, N, K);
}

void copy_cols(hipStream_t st, const double *src, int64_t lds, double *dst, int64_t ldd, int64_t N, int K) {
    copy_cols_kernel<<<grid2(N, K), 256, 0, st>>>(src, lds, dst, ldd, N);
}

void sum_blocks(hipStream_t st, int nb, int p, int q, const double *P, double beta, double *C, int64_t ldc) {
    const int64_t pq = (int64_t)p * q;
    if (pq > 0) sum_blocks_kernel<<<(unsigned)((pq + 255) / 256), 256, 0, st>>>(nb, p, q, P, beta, C, ldc);
}

void unit_lower_copy(hipStream_t st, int rows, int cols, const double *src, int64_t lds, double *dst, int64_t ldd) {
    if (rows > 0 && cols > 0) unit_lower_copy_kernel<<<grid2(rows, cols), 256, 0, st>>>(rows, src, lds, dst, ldd);
}

void triu_copy(hipStream_t st, int n, const double *src, int64_t lds, double *dst, int64_t ldd) {
    if (n > 0) triu_copy_kernel<<<grid2(n, n), 256, 0, st>>>(n, src, lds, dst, ldd);
}

}  // namespace dbie
