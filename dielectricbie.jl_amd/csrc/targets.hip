// targets.hip -- device side of the off-surface target evaluation and the reaction matrix.
//
//  * csr_spmm: Y = A X (+ Y) for a CSR matrix A and a column-major block X; used for the
//    prolongation onto the refined interface (_refined_interface_prolongation,
//    src/kernel/laplace3d_near.jl:152-225), the near-target corrections
//    (laplace3d_pottrg_corrections_hcubature, src/kernel/laplace3d_near_hcubature.jl:193-251)
//    and the transposed prolongation of the reaction weights.  One thread per (row, column):
//    rows are short (1..n_quad^2 entries for the prolongation, a few panels' nodes for the
//    corrections), so the kernel is bound by the gathers of X from HBM/L2.
//  * prolong_t: Qc = P^T Qr for the reaction weights, per coarse panel one dense
//    n_quad^2 x (refined nodes of the panel) block of P^T times those rows of Qr on FP64 MFMA
//    (the refined panels of one coarse panel are consecutive, and their nodes interpolate from
//    that panel's nodes only).
//  * reaction_q: Qt[r, a] = w_r / (4 pi) * sum_{s in group a} q_s / |t_s - y_r|, the weight of
//    refined node r in group a's potential integral.  With it the reaction matrix of
//    four_index_matrix (src/solver/multi_rhs.jl:299-321, induced term)
//        R[a, b] = sum_{s in a} q_s u[Sigma_b](t_s)
//    becomes R = (P^T (Qt + Qcorr))^T Sigma: one N x G weight block and one DGEMM instead of the
//    potential at every target (2 M N K flops).  FP64 VALU, LDS-staged source points.
//  * target_cubatures: the adaptive two-dimensional Genz-Malik cubatures of the near-target
//    corrections (one workgroup of one wave per (refined panel, target) pair), with every sum in
//    the order of the host cubature (targets.cpp, hcubature2 / gm_apply), so the integrals and
//    the evaluation counts are those of the host build.
#include <hip/hip_runtime.h>

#include "devmath.hpp"
#include "kernels.hpp"

namespace dbie {
namespace {
constexpr double kFourPiT = 4.0 * 3.14159265358979311599796346854;
typedef double v4d __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(256) void csr_spmm_kernel(int64_t R, const int64_t *__restrict__ ptr,
                                                       const int *__restrict__ col, const double *__restrict__ val,
                                                       const double *__restrict__ X, int64_t ldx, double *__restrict__ Y,
                                                       int64_t ldy, int accumulate) {
    const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (r >= R) return;
    const int64_t k = blockIdx.y;
    const double *xk = X + k * ldx;
    double acc = 0.0;
    for (int64_t e = ptr[r]; e < ptr[r + 1]; ++e) acc += val[e] * xk[col[e]];
    double *y = Y + r + k * ldy;
    *y = accumulate ? *y + acc : acc;
}

// 1 / sqrt(d2) for d2 in the normal FP32 range: the FP32 hardware estimate of the FP32-rounded d2
// (relative error e0 below 2e-7) refined by one Newton step y <- y (3 - d2 y^2) / 2 in FP64, whose
// relative error is -1.5 e0^2 - 0.5 e0^3, below 7e-14 (one-signed: the step never overshoots)
__device__ inline double rsqrt_f64(double d2) {
    const double y = (double)__builtin_amdgcn_rsqf((float)d2);
    return y * fma(-0.5 * d2, y * y, 1.5);
}

__global__ __launch_bounds__(256) void reaction_q_kernel(DevGeom g, const int64_t *__restrict__ grp_ptr,
                                                         const double *__restrict__ trg, const double *__restrict__ q,
                                                         double *__restrict__ Qt) {
    __shared__ double4 s_src[256];
    const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int a = blockIdx.y;
    const int64_t jc = j < g.N ? j : g.N - 1;
    const double x = g.x[jc], y = g.y[jc], z = g.z[jc];
    const int64_t s0 = grp_ptr[a], s1 = grp_ptr[a + 1];
    double acc = 0.0;
    for (int64_t sb = s0; sb < s1; sb += 256) {
        __syncthreads();
        const int64_t s = sb + threadIdx.x;
        if (s < s1) s_src[threadIdx.x] = make_double4(trg[3 * s], trg[3 * s + 1], trg[3 * s + 2], q[s]);
        __syncthreads();
        const int n = (int)min<int64_t>(256, s1 - sb);
        for (int u = 0; u < n; ++u) {
            const double4 p = s_src[u];
            const double dx = p.x - x, dy = p.y - y, dz = p.z - z;
            acc = fma(p.w, rsqrt_f64(dx * dx + dy * dy + dz * dz), acc);
        }
    }
    if (j < g.N) Qt[j + (int64_t)a * g.N] = g.w[j] * acc / kFourPiT;
}

// XCD-aware order of coarse panels: workgroup w runs on XCD w % 8; each XCD takes a contiguous
// range of panels.
__device__ inline int xcd_panel(int w, int P) { return (w % 8) * ((P + 7) / 8) + w / 8; }

// Qc[c npp + k, a] = sum_{r refined node of coarse panel c} P[r, c npp + k] Qr[r, a]: one workgroup
// per coarse panel and 256 columns.  Chunks of 32 refined nodes: the P^T chunk (npp x 32, zero
// beyond npp rows and the panel's nodes) goes to LDS, the matching rows of Qr are the B operands of
// v_mfma_f64_16x16x4_f64 (RT row tiles of 16 coarse nodes, 4 column blocks of 16 per wave).
template <int RT>
__global__ __launch_bounds__(256) void prolong_t_kernel(int P, int npp, const int64_t *__restrict__ first,
                                                        const int *__restrict__ count, const int64_t *__restrict__ p_ptr,
                                                        const int *__restrict__ p_col, const double *__restrict__ p_val,
                                                        int G, const double *__restrict__ Qr, int64_t Nr,
                                                        double *__restrict__ Qc, int64_t N) {
    constexpr int CH = 32;
    constexpr int LP = CH + 1;
    constexpr int LO = 16 * RT + 1;
    __shared__ double s_p[16 * RT][LP];  // [coarse node k][refined node within the chunk]
    __shared__ double s_out[256][LO];    // [column within the 256-column group][coarse node]
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int c = xcd_panel(blockIdx.x, P);
    const int a0 = blockIdx.y * 256;
    if (c >= P) return;  // uniform per workgroup
    v4d acc[RT][4];
#pragma unroll
    for (int rt = 0; rt < RT; ++rt)
#pragma unroll
        for (int q = 0; q < 4; ++q) acc[rt][q] = v4d{0.0, 0.0, 0.0, 0.0};
    int64_t coff[4];
    double cmask[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int a = a0 + (wave + 4 * q) * 16 + (lane & 15);
        cmask[q] = a < G ? 1.0 : 0.0;
        coff[q] = (int64_t)(a < G ? a : G - 1) * Nr;
    }
    const int64_t r0 = first[c], col0 = (int64_t)c * npp;
    const int n = count[c];
    for (int base = 0; base < n; base += CH) {
        __syncthreads();  // the previous chunk has been read
        for (int idx = tid; idx < 16 * RT * LP; idx += 256) (&s_p[0][0])[idx] = 0.0;
        __syncthreads();
        const int nc = min(CH, n - base);
        for (int idx = tid; idx < nc * npp; idx += 256) {
            const int u = idx / npp, j = idx - u * npp;
            const int64_t r = r0 + base + u;
            const int64_t e = p_ptr[r] + j;
            if (e < p_ptr[r + 1]) {
                const int64_t k = p_col[e] - col0;
                if (k >= 0 && k < npp) s_p[k][u] = p_val[e];
            }
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < CH / 4; ++kk) {
            const int u = kk * 4 + (lane >> 4);
            double xv[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) xv[q] = u < nc ? Qr[r0 + base + u + coff[q]] : 0.0;
#pragma unroll
            for (int rt = 0; rt < RT; ++rt) {
                const double a = s_p[rt * 16 + (lane & 15)][u];
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    acc[rt][q] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, xv[q], acc[rt][q], 0, 0, 0);
            }
        }
    }
    // D(row = 4r + lane/16, col = lane%16) of row tile rt -> LDS [col][16 rt + row]
#pragma unroll
    for (int rt = 0; rt < RT; ++rt)
#pragma unroll
        for (int q = 0; q < 4; ++q)
#pragma unroll
            for (int r = 0; r < 4; ++r)
                s_out[(wave + 4 * q) * 16 + (lane & 15)][16 * rt + 4 * r + (lane >> 4)] = acc[rt][q][r] * cmask[q];
    __syncthreads();
    for (int idx = tid; idx < 256 * npp; idx += 256) {
        const int cl = idx / npp, k = idx - cl * npp;
        const int a = a0 + cl;
        if (a >= G) break;  // idx only grows
        Qc[col0 + k + (int64_t)a * N] = s_out[cl][k];
    }
}

__global__ void scatter_add_kernel(int64_t n, const int64_t *__restrict__ idx, const double *__restrict__ val,
                                   double *__restrict__ Y) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e < n) Y[idx[e]] += val[e];
}

// ---- near-target cubatures ------------------------------------------------------------------
// Box storage per slot: a0, a1, b0, b1, E (-1 once split), kdiv, I[npp].  Slots are appended in
// creation order, so slot order is the host's sequence order.
constexpr int kBoxHead = 6;

// One 64-lane workgroup per pair (grid-stride over the pairs).  pdat per pair: panel centre,
// b - a, d - a, area scale, target (13 doubles).  status: 0 converged, 1 stopped at maxevals,
// 2 box storage full (the host redoes that pair).
__global__ __launch_bounds__(64) void target_cubature_kernel(int npairs, int nq, const double *__restrict__ pdat,
                                                             const double *__restrict__ rules, CubatureRule gm,
                                                             double atol, long long maxevals, int bcap,
                                                             double *__restrict__ boxes, double *__restrict__ Iout,
                                                             int *__restrict__ status) {
#pragma clang fp contract(off)
    __shared__ double s_v[34][kCubatureMaxNpp];  // integrand at the 17 points of two boxes
    __shared__ double s_sq[2][3][kCubatureMaxNpp];  // per-component squares of the two boxes
    __shared__ double s_ab[2][4];                   // the boxes being evaluated
    __shared__ double s_red[2][3];                  // E, kdiv, per box
    __shared__ double s_sel[2];                     // argmax: E, slot
    const int lane = threadIdx.x;
    const int npp = nq * nq, stride = kBoxHead + npp;
    const double *glx = rules, *bary = rules + nq;
    double *bx = boxes + (size_t)blockIdx.x * bcap * stride;

    for (int pr = blockIdx.x; pr < npairs; pr += gridDim.x) {
        const double *pd = pdat + (size_t)pr * 13;
        const double cc0 = pd[0], cc1 = pd[1], cc2 = pd[2], bm0 = pd[3], bm1 = pd[4], bm2 = pd[5];
        const double dm0 = pd[6], dm1 = pd[7], dm2 = pd[8], scale = pd[9], t0 = pd[10], t1 = pd[11], t2 = pd[12];

        // the integrand at point (u, v): kernel times the tensor Lagrange basis (targets.cpp, f)
        auto integrand = [&](double u, double v, double *vals) {
            double rx[16], ry[16];
            bary_row(nq, glx, bary, u, rx);
            bary_row(nq, glx, bary, v, ry);
            const double p0 = cc0 + bm0 * (u / 2) + dm0 * (v / 2);
            const double p1 = cc1 + bm1 * (u / 2) + dm1 * (v / 2);
            const double p2 = cc2 + bm2 * (u / 2) + dm2 * (v / 2);
            const double dx = p0 - t0, dy = p1 - t1, dz = p2 - t2;
            const double kv = (1.0 / sqrt(dx * dx + dy * dy + dz * dz)) / kFourPiT * scale;
            for (int ii = 0; ii < nq; ++ii)
                for (int jj = 0; jj < nq; ++jj) vals[ii * nq + jj] = kv * rx[ii] * ry[jj];
        };
        // rule applications on the nb boxes in s_ab: I into slots dst[k], E and kdiv into s_red
        auto apply = [&](int nb, const int *dst) {
            if (lane < 17 * nb) {
                const int k = lane / 17, j = lane - 17 * k;
                const double *ab = s_ab[k];
                const double c0 = 0.5 * (ab[0] + ab[2]), c1 = 0.5 * (ab[1] + ab[3]);
                const double D0 = 0.5 * fabs(ab[2] - ab[0]), D1 = 0.5 * fabs(ab[3] - ab[1]);
                double u = c0, v = c1;
                if (j >= 1 && j < 9) {  // axis points: +-D l2, +-D l3 along axis (j - 1) / 4
                    const int ax = (j - 1) / 4, r = (j - 1) % 4;
                    const double p = (ax == 0 ? D0 : D1) * (r < 2 ? gm.l2 : gm.l3);
                    const double pu = ax == 0 ? p : 0.0, pv = ax == 0 ? 0.0 : p;
                    if (r % 2 == 0) {
                        u = c0 + pu;
                        v = c1 + pv;
                    } else {
                        u = c0 - pu;
                        v = c1 - pv;
                    }
                } else if (j >= 9) {  // corner points, s0 outer, s1 inner; l4 then l5
                    const int q = (j - 9) / 2, l5 = (j - 9) % 2;
                    const double s0 = q < 2 ? -1.0 : 1.0, s1 = q % 2 == 0 ? -1.0 : 1.0;
                    const double lam = l5 ? gm.l5 : gm.l4;
                    u = c0 + s0 * D0 * lam;
                    v = c1 + s1 * D1 * lam;
                }
                integrand(u, v, s_v[lane]);
            }
            __syncthreads();
            if (lane < npp * nb) {
                const int k = lane / npp, i = lane - k * npp;
                const double *ab = s_ab[k];
                const double D0 = 0.5 * fabs(ab[2] - ab[0]), D1 = 0.5 * fabs(ab[3] - ab[1]);
                const double V = D0 * D1;
                const double(*fv)[kCubatureMaxNpp] = s_v + 17 * k;
                const double f1 = fv[0][i];
                double f2 = 0.0, f3 = 0.0, f4 = 0.0, f5 = 0.0;
                for (int ax = 0; ax < 2; ++ax) {
                    const double f2i = fv[1 + 4 * ax][i] + fv[2 + 4 * ax][i];
                    const double f3i = fv[3 + 4 * ax][i] + fv[4 + 4 * ax][i];
                    f2 += f2i;
                    f3 += f3i;
                    const double e = f3i + 12 * f1 - 7 * f2i;  // fourth divided difference
                    s_sq[k][ax][i] = e * e;
                }
                for (int q = 0; q < 4; ++q) {
                    f4 += fv[9 + 2 * q][i];
                    f5 += fv[10 + 2 * q][i];
                }
                const double I = V * (gm.w[0] * f1 + gm.w[1] * f2 + gm.w[2] * f3 + gm.w[3] * f4 + gm.w[4] * f5);
                const double Ip = V * (gm.wp[0] * f1 + gm.wp[1] * f2 + gm.wp[2] * f3 + gm.wp[3] * f4);
                s_sq[k][2][i] = (I - Ip) * (I - Ip);
                bx[(size_t)dst[k] * stride + kBoxHead + i] = I;
            }
            __syncthreads();
            if (lane < nb) {  // serial sums in component order, as on the host
                const int k = lane;
                double d0 = 0.0, d1 = 0.0, e2 = 0.0;
                for (int i = 0; i < npp; ++i) d0 += s_sq[k][0][i];
                for (int i = 0; i < npp; ++i) d1 += s_sq[k][1][i];
                for (int i = 0; i < npp; ++i) e2 += s_sq[k][2][i];
                d0 = sqrt(d0);
                d1 = sqrt(d1);
                double maxdd = 0.0;
                int kdiv = 0;
                if (d0 > maxdd) {
                    maxdd = d0;
                    kdiv = 0;
                }
                if (d1 > maxdd) {
                    maxdd = d1;
                    kdiv = 1;
                }
                const double E = sqrt(e2);
                s_red[k][0] = E;
                s_red[k][1] = (double)kdiv;
                double *h = bx + (size_t)dst[k] * stride;
                h[0] = s_ab[k][0];
                h[1] = s_ab[k][1];
                h[2] = s_ab[k][2];
                h[3] = s_ab[k][3];
                h[4] = E;
                h[5] = (double)kdiv;
            }
            __syncthreads();
        };

        if (lane == 0) {
            s_ab[0][0] = -1.0;
            s_ab[0][1] = -1.0;
            s_ab[0][2] = 1.0;
            s_ab[0][3] = 1.0;
        }
        __syncthreads();
        const int d0[2] = {0, 0};
        apply(1, d0);
        long long evals = 17;
        int n = 1, st = 0;
        double Etot = s_red[0][0];
        if (!(Etot <= atol || evals >= maxevals)) {
            for (;;) {
                // the box with the largest error, the oldest among equal errors
                double be = -2.0;
                int bs = 0x7fffffff;
                for (int sl = lane; sl < n; sl += 64) {
                    const double e = bx[(size_t)sl * stride + 4];
                    if (e > be) {
                        be = e;
                        bs = sl;
                    }
                }
                for (int off = 32; off > 0; off >>= 1) {
                    const double oe = __shfl_xor(be, off, 64);
                    const int os = __shfl_xor(bs, off, 64);
                    if (oe > be || (oe == be && os < bs)) {
                        be = oe;
                        bs = os;
                    }
                }
                if (n + 2 > bcap) {
                    st = 2;
                    break;
                }
                const double *old = bx + (size_t)bs * stride;
                const double oE = old[4];
                const int kd = (int)old[5];
                if (lane == 0) {
                    const double a0 = old[0], a1 = old[1], b0 = old[2], b1 = old[3];
                    const double ak = kd == 0 ? a0 : a1, bk = kd == 0 ? b0 : b1;
                    const double mid = ak + (bk - ak) / 2;
                    s_ab[0][0] = a0;
                    s_ab[0][1] = a1;
                    s_ab[0][2] = kd == 0 ? mid : b0;
                    s_ab[0][3] = kd == 0 ? b1 : mid;
                    s_ab[1][0] = kd == 0 ? mid : a0;
                    s_ab[1][1] = kd == 0 ? a1 : mid;
                    s_ab[1][2] = b0;
                    s_ab[1][3] = b1;
                }
                __syncthreads();
                if (lane == 0) bx[(size_t)bs * stride + 4] = -1.0;  // split: no longer a leaf
                const int dst[2] = {n, n + 1};
                apply(2, dst);
                n += 2;
                evals += 34;
                Etot += s_red[0][0] + s_red[1][0] - oE;
                if (Etot <= atol || evals >= maxevals) break;
            }
        }
        if (st == 0) st = Etot <= atol ? 0 : 1;
        // sum over the leaves in creation order
        if (lane < npp) {
            double acc = 0.0;
            for (int sl = 0; sl < n; ++sl)
                if (bx[(size_t)sl * stride + 4] >= 0.0) acc += bx[(size_t)sl * stride + kBoxHead + lane];
            Iout[(size_t)pr * npp + lane] = acc;
        }
        if (lane == 0) status[pr] = st;
        __syncthreads();
    }
}
// ---- refinement of the interface around the targets (_refine_interface_for_targets,
// src/kernel/laplace3d_near.jl:99-150) and the prolongation onto it (_refined_interface_prolongation,
// :152-225), restating the host builder (targets.cpp) operation for operation without contraction, so
// that every decision (split or keep) and every value are those of the host.

__device__ inline double len3_k(double a, double b, double c) {
#pragma clang fp contract(off)
    return sqrt(a * a + b * b + c * c);
}
__device__ inline double panel_max_length_k(const double *c) {
#pragma clang fp contract(off)
    const double l1 = len3_k(c[0] - c[3], c[1] - c[4], c[2] - c[5]);
    const double l2 = len3_k(c[3] - c[6], c[4] - c[7], c[5] - c[8]);
    return l1 < l2 ? l2 : l1;
}
// clamped cell coordinate of the host TargetGrid
__device__ inline int64_t grid_coord(const RefineGrid &G, int k, double v) {
#pragma clang fp contract(off)
    int64_t c = (int64_t)floor((v - G.lo[k]) / G.h);
    c = c < 0 ? 0 : c;
    return c > G.dim[k] - 1 ? G.dim[k] - 1 : c;
}
// some target with |t - c|^2 <= r^2 (TargetGrid::ball)
__device__ bool any_in_ball(const RefineGrid &G, const double *c, double r) {
#pragma clang fp contract(off)
    int64_t a[3], b[3];
    for (int k = 0; k < 3; ++k) {
        a[k] = grid_coord(G, k, c[k] - r);
        b[k] = grid_coord(G, k, c[k] + r);
    }
    const double r2 = r * r;
    for (int64_t i = a[0]; i <= b[0]; ++i)
        for (int64_t j = a[1]; j <= b[1]; ++j)
            for (int64_t k = a[2]; k <= b[2]; ++k) {
                const int64_t cc = (i * G.dim[1] + j) * G.dim[2] + k;
                for (int64_t q = G.start[cc]; q < G.start[cc + 1]; ++q) {
                    const double *p = G.t + 3 * (int64_t)G.idx[q];
                    const double dx = p[0] - c[0], dy = p[1] - c[1], dz = p[2] - c[2];
                    if (dx * dx + dy * dy + dz * dz <= r2) return true;
                }
            }
    return false;
}

constexpr int kRefineStack = 64;  // items per thread; deeper refinements fall back to the host

// One thread per coarse panel, its subtree walked as the host stack walks it (children pushed in
// order 0..3, so visited 3..0).  EMIT = false: count[p] = leaves of panel p (-1: stack overflow);
// EMIT = true: the leaves from offset[p] on (corners, parent, from_split).
template <bool EMIT>
__global__ __launch_bounds__(64) void refine_kernel(int P, const double *__restrict__ corners, RefineGrid G, int64_t M,
                                                    double limit, double range_factor, int nq, int64_t *__restrict__ count,
                                                    const int64_t *__restrict__ offset, double *__restrict__ out_c,
                                                    int *__restrict__ parent, int *__restrict__ from_split) {
#pragma clang fp contract(off)
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    double stk[kRefineStack][12];
    unsigned char spl[kRefineStack];
    int sp = 1;
    for (int k = 0; k < 12; ++k) stk[0][k] = corners[12 * (int64_t)p + k];
    spl[0] = 0;
    int64_t n = 0;
    while (sp > 0) {
        --sp;
        double c[12];
        for (int k = 0; k < 12; ++k) c[k] = stk[sp][k];
        const unsigned char was_split = spl[sp];
        const double l = panel_max_length_k(c);
        double cc[3];
        for (int k = 0; k < 3; ++k) cc[k] = (c[k] + c[3 + k] + c[6 + k] + c[9 + k]) / 4;
        const double r = range_factor * l / nq;
        const bool near = M > 0 && any_in_ball(G, cc, r);
        if (near && l > limit) {
            if (sp + 4 > kRefineStack) {
                if (!EMIT) count[p] = -1;
                return;
            }
            // _split_panel4 (src/kernel/laplace3d_near.jl:78-97): (a,ab,m,da), (ab,b,bc,m), (m,bc,c,cd), (da,m,cd,d)
            double ab[3], bc[3], cd[3], da[3], m[3];
            for (int k = 0; k < 3; ++k) {
                ab[k] = (c[k] + c[3 + k]) / 2;
                bc[k] = (c[3 + k] + c[6 + k]) / 2;
                cd[k] = (c[6 + k] + c[9 + k]) / 2;
                da[k] = (c[9 + k] + c[k]) / 2;
                m[k] = (c[k] + c[3 + k] + c[6 + k] + c[9 + k]) / 4;
            }
            const double *quads[4][4] = {{c, ab, m, da}, {ab, c + 3, bc, m}, {m, bc, c + 6, cd}, {da, m, cd, c + 9}};
            for (int q = 0; q < 4; ++q) {
                for (int v = 0; v < 4; ++v)
                    for (int k = 0; k < 3; ++k) stk[sp][3 * v + k] = quads[q][v][k];
                spl[sp] = 1;
                ++sp;
            }
        } else {
            if (EMIT) {
                const int64_t o = offset[p] + n;
                for (int k = 0; k < 12; ++k) out_c[12 * o + k] = c[k];
                parent[o] = p;
                from_split[o] = was_split;
            }
            ++n;
        }
    }
    if (!EMIT) count[p] = n;
}

// rect_panel3d_discretize nodes and weights of the refined panels (fill_nodes, geometry.cpp)
__global__ __launch_bounds__(256) void refined_nodes_kernel(int64_t Pr, int nq, const double *__restrict__ corners,
                                                            const double *__restrict__ glx,
                                                            const double *__restrict__ glw, double *__restrict__ x,
                                                            double *__restrict__ y, double *__restrict__ z,
                                                            double *__restrict__ w) {
#pragma clang fp contract(off)
    const int npp = nq * nq;
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= Pr * npp) return;
    const int64_t p = n / npp;
    const int ij = (int)(n - p * npp), i = ij / nq, j = ij - i * nq;
    const double *A = corners + 12 * p, *B = A + 3, *C = A + 6, *D = A + 9;
    const double Lab = len3_k(B[0] - A[0], B[1] - A[1], B[2] - A[2]);
    const double Lbc = len3_k(C[0] - B[0], C[1] - B[1], C[2] - B[2]);
    double v[3];
    for (int k = 0; k < 3; ++k) {
        const double cc = (A[k] + B[k] + C[k] + D[k]) / 4;
        v[k] = cc + (B[k] - A[k]) * (glx[i] / 2) + (D[k] - A[k]) * (glx[j] / 2);
    }
    x[n] = v[0];
    y[n] = v[1];
    z[n] = v[2];
    w[n] = glw[i] * glw[j] * Lab * Lbc / 4;
}

// Prolongation rows: a node of an unsplit panel copies its coarse node; one of a split panel takes
// the tensor Lagrange interpolant of its parent (coarse rule glx, bary), at the node's local
// coordinates from the 2 x 2 normal equations of the parent's edge vectors.
__global__ __launch_bounds__(256) void prolong_kernel(int64_t Pr, int nq, const double *__restrict__ coarse_corners,
                                                      const double *__restrict__ rules, const int *__restrict__ parent,
                                                      const int *__restrict__ from_split, const double *__restrict__ x,
                                                      const double *__restrict__ y, const double *__restrict__ z,
                                                      const int64_t *__restrict__ p_ptr, int *__restrict__ p_col,
                                                      double *__restrict__ p_val) {
#pragma clang fp contract(off)
    const int npp = nq * nq;
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= Pr * npp) return;
    const int64_t q = n / npp;
    const int k = (int)(n - q * npp);
    const int par = parent[q];
    const int64_t col0 = (int64_t)par * npp;
    int64_t e = p_ptr[n];
    if (!from_split[q]) {
        p_col[e] = (int)(col0 + k);
        p_val[e] = 1.0;
        return;
    }
    const double *glx = rules, *bary = rules + nq;
    const double *c = coarse_corners + 12 * (int64_t)par;
    const double *a = c, *b = c + 3, *d = c + 9;
    double cc[3], bma[3], dma[3];
    for (int t = 0; t < 3; ++t) {
        cc[t] = (c[t] + c[3 + t] + c[6 + t] + c[9 + t]) / 4;
        bma[t] = b[t] - a[t];
        dma[t] = d[t] - a[t];
    }
    const double bb = bma[0] * bma[0] + bma[1] * bma[1] + bma[2] * bma[2];
    const double bd = bma[0] * dma[0] + bma[1] * dma[1] + bma[2] * dma[2];
    const double dd = dma[0] * dma[0] + dma[1] * dma[1] + dma[2] * dma[2];
    const double det = bb * dd - bd * bd;
    const double inv11 = dd / det, inv12 = -bd / det, inv22 = bb / det;
    const double xv[3] = {x[n] - cc[0], y[n] - cc[1], z[n] - cc[2]};
    const double rhs1 = bma[0] * xv[0] + bma[1] * xv[1] + bma[2] * xv[2];
    const double rhs2 = dma[0] * xv[0] + dma[1] * xv[1] + dma[2] * xv[2];
    const double s = inv11 * rhs1 + inv12 * rhs2;
    const double t = inv12 * rhs1 + inv22 * rhs2;
    double rx[16], ry[16];
    bary_row(nq, glx, bary, 2 * s, rx);
    bary_row(nq, glx, bary, 2 * t, ry);
    for (int ii = 0; ii < nq; ++ii)
        for (int jj = 0; jj < nq; ++jj, ++e) {
            p_col[e] = (int)(col0 + ii * nq + jj);
            p_val[e] = rx[ii] * ry[jj];
        }
}
}  // namespace

void refine_count(hipStream_t st, int P, const double *corners, const RefineGrid &G, int64_t M, double limit,
                  double range_factor, int nq, int64_t *count) {
    if (P > 0)
        refine_kernel<false><<<(P + 63) / 64, 64, 0, st>>>(P, corners, G, M, limit, range_factor, nq, count, nullptr,
                                                            nullptr, nullptr, nullptr);
}
void refine_emit(hipStream_t st, int P, const double *corners, const RefineGrid &G, int64_t M, double limit,
                 double range_factor, int nq, const int64_t *offset, double *out_corners, int *parent, int *from_split) {
    if (P > 0)
        refine_kernel<true><<<(P + 63) / 64, 64, 0, st>>>(P, corners, G, M, limit, range_factor, nq, nullptr, offset,
                                                           out_corners, parent, from_split);
}
void refined_nodes(hipStream_t st, int64_t Pr, int nq, const double *corners, const double *glx, const double *glw,
                   double *x, double *y, double *z, double *w) {
    const int64_t n = Pr * nq * nq;
    if (n > 0)
        refined_nodes_kernel<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(Pr, nq, corners, glx, glw, x, y, z, w);
}
void prolongation(hipStream_t st, int64_t Pr, int nq, const double *coarse_corners, const double *rules,
                  const int *parent, const int *from_split, const double *x, const double *y, const double *z,
                  const int64_t *p_ptr, int *p_col, double *p_val) {
    const int64_t n = Pr * nq * nq;
    if (n > 0)
        prolong_kernel<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(Pr, nq, coarse_corners, rules, parent, from_split,
                                                                     x, y, z, p_ptr, p_col, p_val);
}

void csr_spmm(hipStream_t st, int64_t R, const int64_t *ptr, const int *col, const double *val, int K, const double *X,
              int64_t ldx, double *Y, int64_t ldy, bool accumulate) {
    if (R <= 0 || K <= 0) return;
    csr_spmm_kernel<<<dim3((unsigned)((R + 255) / 256), (unsigned)K), 256, 0, st>>>(R, ptr, col, val, X, ldx, Y, ldy,
                                                                                     accumulate ? 1 : 0);
}

void reaction_q(hipStream_t st, const DevGeom &g, int G, const int64_t *grp_ptr, const double *trg, const double *q,
                double *Qt) {
    if (g.N <= 0 || G <= 0) return;
    reaction_q_kernel<<<dim3((unsigned)((g.N + 255) / 256), (unsigned)G), 256, 0, st>>>(g, grp_ptr, trg, q, Qt);
}

bool prolong_t(hipStream_t st, int P, int npp, const int64_t *first, const int *count, const int64_t *p_ptr,
               const int *p_col, const double *p_val, int G, const double *Qr, int64_t Nr, double *Qc, int64_t N) {
    if (npp > 32) return false;
    if (P <= 0 || G <= 0) return true;
    const int per = (P + 7) / 8;
    dim3 grid((unsigned)(8 * per), (unsigned)((G + 255) / 256));
    if (npp <= 16)
        prolong_t_kernel<1><<<grid, 256, 0, st>>>(P, npp, first, count, p_ptr, p_col, p_val, G, Qr, Nr, Qc, N);
    else
        prolong_t_kernel<2><<<grid, 256, 0, st>>>(P, npp, first, count, p_ptr, p_col, p_val, G, Qr, Nr, Qc, N);
    return true;
}

void scatter_add(hipStream_t st, int64_t n, const int64_t *idx, const double *val, double *Y) {
    if (n <= 0) return;
    scatter_add_kernel<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(n, idx, val, Y);
}

size_t target_cubature_scratch(int nwg, int nq, int bcap) {
    return sizeof(double) * (size_t)nwg * bcap * (kBoxHead + nq * nq);
}

void target_cubatures(hipStream_t st, int npairs, int nq, const double *pdat, const double *rules,
                      const CubatureRule &rule, double atol, long long maxevals, int bcap, int nwg, double *boxes,
                      double *Iout, int *status) {
    if (npairs <= 0) return;
    target_cubature_kernel<<<(unsigned)nwg, 64, 0, st>>>(npairs, nq, pdat, rules, rule, atol, maxevals, bcap, boxes,
                                                         Iout, status);
}

}  // namespace dbie
