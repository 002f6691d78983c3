// nearbuild.hip -- correction blocks of the upsampled near pairs built on the device.
//
// For a near pair (source panel i, target panel j) with upsampling order n_up
// (_laplace3d_corrections, src/kernel/laplace3d_near_upsampling.jl:127-276):
//     K_up[t, c] = sum_a Mt_i[c, a] * G(u_a, x_t, n_j)        (u_a: n_up^2 upsampled nodes of panel i)
//     C[t, c]    = K_up[t, c] - w_c G(y_c, x_t, n_j)           (direct quadrature term)
// with G = laplace3d_grad (src/kernel/laplace3d.jl:9-15).  Mt_i and u_a come from the host
// (per-source-panel caches); one workgroup builds one block: the kernel values for all
// (target, upsampled node) pairs go to LDS, then each thread forms its (t, c) sums in the
// order of the host builder.  Outputs: the correction C in FP64 and FP32 and the full near
// interaction K_up in FP32, in the storage orders of the near-field kernels.  Touching pairs
// arrive as K_up blocks (from the quadtree kernel below, or from the host) and only get their
// direct term here.
#include <hip/hip_runtime.h>

#include <cfloat>

#include "devmath.hpp"
#include "kernels.hpp"

namespace dbie {
namespace {
constexpr double kFourPiB = 4.0 * 3.14159265358979311599796346854;

// laplace3d_grad(s, t, n), with the operation order of the host builder
__device__ inline double grad_k(double s0, double s1, double s2, double t0, double t1, double t2, double n0,
                                double n1, double n2) {
#pragma clang fp contract(off)
    const double d0 = s0 - t0, d1 = s1 - t1, d2 = s2 - t2;
    const double r = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
    const double ir = 1.0 / r;
    const double ir3 = ir * ir * ir;
    return (n0 * (ir3 * (t0 - s0)) + n1 * (ir3 * (t1 - s1)) + n2 * (ir3 * (t2 - s2))) / kFourPiB;
}

// Kernel of a correction block between a point y of the source panel and a target x: MODE 0 (:DT)
// laplace3d_grad(y, x, n_trg), MODE 1 (:D) laplace3d_grad(x, y, n_src)
// (src/kernel/laplace3d_near_upsampling.jl:183-197, src/kernel/laplace3d_near_adaptive.jl:169-175)
template <int MODE>
__device__ inline double corr_k(double y0, double y1, double y2, double x0, double x1, double x2, double tn0,
                                double tn1, double tn2, double sn0, double sn1, double sn2) {
    return MODE == 1 ? grad_k(x0, x1, x2, y0, y1, y2, sn0, sn1, sn2) : grad_k(y0, y1, y2, x0, x1, x2, tn0, tn1, tn2);
}

__device__ inline int store_idx(int bd, bool mfma, int t, int c) {
    return mfma ? (c >> 2) * 64 + (c & 3) * 16 + t : t * bd + c;
}
__device__ inline int store_up_idx(int bd, bool mfma, int t, int c) {
    return mfma ? ((c >> 2) * 16 + t) * 4 + (c & 3) : t * bd + c;
}

// One workgroup per item.  up = 1: upsampled pair (K_up from the cache); up = 0: K_up given (kup).
// val32 and vup32 may be null (MODE 1 builds the FP64 blocks only).
template <bool UP, int MODE>
__global__ __launch_bounds__(256) void near_blocks_kernel(DevGeom g, int bd, int mfma, const int *__restrict__ src,
                                                          const int *__restrict__ trg, const int *__restrict__ slot,
                                                          const int *__restrict__ cache_off,
                                                          const int *__restrict__ cache_nu,
                                                          const double *__restrict__ caches,
                                                          const double *__restrict__ kup, int mt_lds,
                                                          double *__restrict__ val64, float *__restrict__ val32,
                                                          float *__restrict__ vup32) {
#pragma clang fp contract(off)
    extern __shared__ double sm[];  // UP: kv (bd x nu2), then Mt (bd x nu2) when mt_lds
    const int it = blockIdx.x;
    const int i = src[it], j = trg[it];
    const int64_t sb = (int64_t)i * bd, tb = (int64_t)j * bd;
    const double n0 = g.nx[tb], n1 = g.ny[tb], n2 = g.nz[tb];
    const double s0 = g.nx[sb], s1 = g.ny[sb], s2 = g.nz[sb];
    const size_t out0 = (size_t)slot[it] * bd * bd;
    int nu2 = 0;
    const double *Mt = nullptr;
    double *kv = sm;
    if (UP) {
        const int nu = cache_nu[i];
        nu2 = nu * nu;
        const double *upos = caches + cache_off[i];
        Mt = upos + 3 * nu2;
        double *sMt = sm + bd * nu2;
        for (int q = threadIdx.x; q < bd * nu2; q += blockDim.x) {
            const int t = q / nu2, a = q - t * nu2;
            kv[q] = corr_k<MODE>(upos[3 * a], upos[3 * a + 1], upos[3 * a + 2], g.x[tb + t], g.y[tb + t], g.z[tb + t],
                                 n0, n1, n2, s0, s1, s2);
            if (mt_lds) sMt[q] = Mt[q];
        }
        if (mt_lds) Mt = sMt;
        __syncthreads();
    }
    for (int o = threadIdx.x; o < bd * bd; o += blockDim.x) {
        const int t = o / bd, c = o - t * bd;
        double s = 0.0;
        if (UP) {
            const double *mr = Mt + (size_t)c * nu2;
            const double *kr = kv + (size_t)t * nu2;
            for (int a = 0; a < nu2; ++a) s += mr[a] * kr[a];
        } else {
            s = kup[(size_t)it * bd * bd + o];
        }
        const double direct = corr_k<MODE>(g.x[sb + c], g.y[sb + c], g.z[sb + c], g.x[tb + t], g.y[tb + t],
                                           g.z[tb + t], n0, n1, n2, s0, s1, s2) *
                              g.w[sb + c];
        const double v = s - direct;
        val64[out0 + store_idx(bd, mfma != 0, t, c)] = v;
        if (val32) val32[out0 + store_idx(bd, mfma != 0, t, c)] = (float)v;
        if (vup32) vup32[out0 + store_up_idx(bd, mfma != 0, t, c)] = (float)s;
    }
}
// Touching pairs: the quadtree of walk_adaptive_tree (src/kernel/laplace3d_near_adaptive.jl:119-224)
// driven by the representative target, depth first with an explicit stack in LDS, and the block
// moments of adaptive_block_moments_inplace! (:308-387) accumulated leaf by leaf in walk order.
// Every sum runs in the order of the host quadtree (Quadtree::walk / adaptive_block), so the
// split decisions and the blocks are those of the host.  One wave per pair; kup gets K_up
// (bd x bd row-major, no direct term), hit the max-depth flag (1), or 2 when the walk stopped at
// max_nodes nodes (the host then builds that block).  MODE as in near_blocks_kernel.
template <int MODE>
__global__ __launch_bounds__(64) void adaptive_blocks_kernel(DevGeom g, int nq, int nGL, const int *__restrict__ src,
                                                             const int *__restrict__ trg, const int *__restrict__ rep,
                                                             const double *__restrict__ pmap,
                                                             const double *__restrict__ rules, double atol,
                                                             double rtol, int max_depth, int max_nodes,
                                                             double *__restrict__ kup, int *__restrict__ hit) {
#pragma clang fp contract(off)
    extern __shared__ double sm[];
    __shared__ int s_top, s_leaf;
    const int it = blockIdx.x, tid = threadIdx.x, nt = blockDim.x;
    const int npp = nq * nq, bd = npp, G2 = nGL * nGL;
    const int ent = 5 + npp, cap = 3 * max_depth + 4;
    double *stk = sm;                   // cap x (ulo, uhi, vlo, vhi, depth, moments[npp])
    double *cm = stk + cap * ent;       // 4 x npp: moments of the four cells in cb
    double *wk = cm + 4 * npp;          // 4 x G2: weight x kernel at the representative target
    double *wuv = wk + 4 * G2;          // 4 x G2: weights
    double *rxs = wuv + 4 * G2;         // 4 x nGL x nq: interpolation rows in u
    double *rys = rxs + 4 * nGL * nq;   // 4 x nGL x nq: interpolation rows in v
    double *pts = rys + 4 * nGL * nq;   // 4 x G2 x 3: quadrature points
    double *kt = pts + 12 * G2;         // G2 x bd: kernel at the panel's targets (one leaf)
    double *blk = kt + G2 * bd;         // bd x bd
    double *cb = blk + bd * bd;         // 4 x (ulo, uhi, vlo, vhi)

    const double *mp = pmap + (size_t)it * 10;
    const double cc0 = mp[0], cc1 = mp[1], cc2 = mp[2], bm0 = mp[3], bm1 = mp[4], bm2 = mp[5];
    const double dm0 = mp[6], dm1 = mp[7], dm2 = mp[8], scale = mp[9];
    const double *glx = rules, *bary = rules + nq, *ns = rules + 2 * nq, *ws = ns + nGL;
    const int64_t tb = (int64_t)trg[it] * bd;
    const int64_t rn = tb + rep[it];
    const double tx = g.x[rn], ty = g.y[rn], tz = g.z[rn];
    const double n0 = g.nx[tb], n1 = g.ny[tb], n2 = g.nz[tb];
    const int64_t sbn = (int64_t)src[it] * bd;
    const double s0 = g.nx[sbn], s1 = g.ny[sbn], s2 = g.nz[sbn];

    // moments of the first ncell cells of cb into cm (Quadtree::moments)
    auto moments = [&](int ncell) {
        for (int idx = tid; idx < ncell * G2; idx += nt) {
            const int k = idx / G2, ab = idx - k * G2, a = ab / nGL, b = ab - a * nGL;
            const double ulo = cb[4 * k], uhi = cb[4 * k + 1], vlo = cb[4 * k + 2], vhi = cb[4 * k + 3];
            const double hu = (uhi - ulo) / 2, hv = (vhi - vlo) / 2, mu = (uhi + ulo) / 2, mv = (vhi + vlo) / 2;
            const double sc = hu * hv * scale;
            const double u = mu + hu * ns[a], v = mv + hv * ns[b];
            if (b == 0) bary_row(nq, glx, bary, u, rxs + (k * nGL + a) * nq);
            if (a == 0) bary_row(nq, glx, bary, v, rys + (k * nGL + b) * nq);
            const double y0 = cc0 + bm0 * (u / 2) + dm0 * (v / 2);
            const double y1 = cc1 + bm1 * (u / 2) + dm1 * (v / 2);
            const double y2 = cc2 + bm2 * (u / 2) + dm2 * (v / 2);
            const double w = ws[a] * ws[b] * sc;
            pts[3 * idx] = y0;
            pts[3 * idx + 1] = y1;
            pts[3 * idx + 2] = y2;
            wuv[idx] = w;
            wk[idx] = w * corr_k<MODE>(y0, y1, y2, tx, ty, tz, n0, n1, n2, s0, s1, s2);
        }
        __syncthreads();
        for (int idx = tid; idx < ncell * npp; idx += nt) {
            const int k = idx / npp, q = idx - k * npp, mx = q / nq, my = q - mx * nq;
            double s = 0.0;
            for (int a = 0; a < nGL; ++a) {
                const double rx = rxs[(k * nGL + a) * nq + mx];
                for (int b = 0; b < nGL; ++b) s += wk[k * G2 + a * nGL + b] * rx * rys[(k * nGL + b) * nq + my];
            }
            cm[idx] = s;
        }
        __syncthreads();
    };

    for (int e = tid; e < bd * bd; e += nt) blk[e] = 0.0;
    if (tid == 0) {
        cb[0] = -1.0; cb[1] = 1.0; cb[2] = -1.0; cb[3] = 1.0;
    }
    __syncthreads();
    moments(1);
    for (int q = tid; q < ent; q += nt) stk[q] = q < 4 ? cb[q] : (q == 4 ? 0.0 : cm[q - 5]);
    if (tid == 0) s_top = 1;
    int hitflag = 0, nodes = 0;
    __syncthreads();
    while (s_top > 0 && nodes < max_nodes) {
        ++nodes;
        const int top = s_top;
        const double *nd = stk + (size_t)(top - 1) * ent;
        const double ulo = nd[0], uhi = nd[1], vlo = nd[2], vhi = nd[3];
        const int depth = (int)nd[4];
        const double mu = (uhi + ulo) / 2, mv = (vhi + vlo) / 2;
        if (tid == 0) {
            const double c[16] = {ulo, mu, vlo, mv, mu, uhi, vlo, mv, ulo, mu, mv, vhi, mu, uhi, mv, vhi};
            for (int q = 0; q < 16; ++q) cb[q] = c[q];
        }
        __syncthreads();
        moments(4);
        if (tid == 0) {
            double err = 0.0, pmax = 0.0;
            for (int q = 0; q < npp; ++q) {
                const double e = fabs(nd[5 + q] - (cm[q] + cm[npp + q] + cm[2 * npp + q] + cm[3 * npp + q]));
                if (q == 0 || e > err) err = e;
                const double a = fabs(nd[5 + q]);
                if (q == 0 || a > pmax) pmax = a;
            }
            const double rp = rtol * pmax;
            const double tl = atol < rp ? rp : atol;
            if (depth == max_depth && err >= tl) hitflag = 1;
            s_leaf = (err < tl || depth == max_depth) ? 1 : 0;
            s_top = top - 1;
        }
        __syncthreads();
        if (s_leaf) {
            // the four cells are leaves: block moments at every target of the panel, cell by cell
            for (int k = 0; k < 4; ++k) {
                for (int idx = tid; idx < G2 * bd; idx += nt) {
                    const int ab = idx / bd, t = idx - ab * bd;
                    const double *y = pts + 3 * (k * G2 + ab);
                    kt[idx] = corr_k<MODE>(y[0], y[1], y[2], g.x[tb + t], g.y[tb + t], g.z[tb + t], n0, n1, n2, s0,
                                           s1, s2);
                }
                __syncthreads();
                for (int e = tid; e < bd * bd; e += nt) {
                    const int t = e / bd, c = e - t * bd, mx = c / nq, my = c - mx * nq;
                    double s = blk[e];
                    for (int a = 0; a < nGL; ++a) {
                        const double rx = rxs[(k * nGL + a) * nq + mx];
                        for (int b = 0; b < nGL; ++b) {
                            const double wm = wuv[k * G2 + a * nGL + b] * rx * rys[(k * nGL + b) * nq + my];
                            s += wm * kt[(a * nGL + b) * bd + t];
                        }
                    }
                    blk[e] = s;
                }
                __syncthreads();
            }
        } else {
            // push the children last-to-first so that the first child is walked next
            const int base = top - 1;
            for (int idx = tid; idx < 4 * ent; idx += nt) {
                const int sl = idx / ent, q = idx - sl * ent, k = 3 - sl;
                stk[(size_t)(base + sl) * ent + q] =
                    q < 4 ? cb[4 * k + q] : (q == 4 ? (double)(depth + 1) : cm[k * npp + q - 5]);
            }
            __syncthreads();
            if (tid == 0) s_top = base + 4;
        }
        __syncthreads();
    }
    for (int e = tid; e < bd * bd; e += nt) kup[(size_t)it * bd * bd + e] = blk[e];
    if (tid == 0) hit[it] = s_top > 0 ? 2 : hitflag;  // 2: node budget spent, block incomplete
}
// ---- near-pair scan (build_neighbor_list, src/kernel/laplace3d_near.jl:240-392) ----------------
// The scalar helpers below repeat the host's (nearfield.cpp) operations one for one, so that the
// Bernstein radii, and with them every pair decision, are bit-identical.

// Julia Base.Math.hypot, FMA-corrected branch
__device__ double hypot_cr_d(double x, double y) {
#pragma clang fp contract(off)
    if (isinf(x) || isinf(y)) return INFINITY;
    double ax = fabs(x), ay = fabs(y);
    if (ay > ax) { const double t = ax; ax = ay; ay = t; }
    if (ay <= ax * sqrt(DBL_EPSILON / 2.0)) return ax;
    double scale = DBL_EPSILON * sqrt(DBL_MIN);
    if (ax > sqrt(DBL_MAX / 2.0)) {
        ax *= scale; ay *= scale; scale = 1.0 / scale;
    } else if (ay < sqrt(DBL_MIN)) {
        ax /= scale; ay /= scale;
    } else {
        scale = 1.0;
    }
    double h = sqrt(fma(ax, ax, ay * ay));
    const double h2 = h * h, ax2 = ax * ax;
    h -= (fma(-ay, ay, h2 - ax2) + fma(h, h, -h2) - fma(ax, ax, -ax2)) / (2.0 * h);
    return h * scale;
}

// Julia Base sqrt(::Complex{Float64})
__device__ void complex_sqrt_d(double x, double y, double &re, double &im) {
#pragma clang fp contract(off)
    if (x == 0.0 && y == 0.0) { re = 0.0; im = y; return; }
    int k = 0;
    double rho = x * x + y * y;
    if (!isfinite(rho) && (isinf(x) || isinf(y))) {
        rho = INFINITY;
    } else if (isinf(rho) || (rho == 0.0 && (x != 0.0 || y != 0.0)) ||
               rho < 4.9406564584124654e-324 / (2.0 * DBL_EPSILON * DBL_EPSILON)) {
        const double fx = fabs(x), fy = fabs(y);
        const double m = fx < fy ? fy : fx;
        k = m == 0.0 ? 0 : ilogb(m);
        const double xk = ldexp(x, -k), yk = ldexp(y, -k);
        rho = xk * xk + yk * yk;
    }
    if (isfinite(x)) rho = ldexp(fabs(x), -k) + sqrt(rho);
    if (k % 2 != 0) {
        k = (k - 1) / 2;
    } else {
        k = k / 2 - 1;
        rho += rho;
    }
    rho = ldexp(sqrt(rho), k);
    double xi = rho, eta = y;
    if (rho != 0.0) {
        if (isfinite(eta)) eta = (eta / rho) / 2.0;
        if (x < 0.0) { xi = fabs(eta); eta = copysign(rho, y); }
    }
    re = xi; im = eta;
}

// bernstein_rho_from_pole (src/utils/bernstein.jl:12-17)
__device__ double rho_pole_d(double zr, double zi) {
#pragma clang fp contract(off)
    double sr, si;
    complex_sqrt_d(zr * zr - zi * zi - 1.0, zr * zi + zi * zr, sr, si);
    const double a = hypot_cr_d(zr + sr, zi + si), b = hypot_cr_d(zr - sr, zi - si);
    return a < b ? b : a;
}

// bernstein_rho_panel (src/utils/bernstein.jl:33-62); f: panel data (frame at offset 5)
__device__ double bernstein_rho_d(const double *f, double t0, double t1, double t2) {
#pragma clang fp contract(off)
    const double r0 = t0 - f[5], r1 = t1 - f[6], r2 = t2 - f[7];
    const double xt = r0 * f[8] + r1 * f[9] + r2 * f[10];
    const double yt = r0 * f[11] + r1 * f[12] + r2 * f[13];
    const double zt = r0 * f[14] + r1 * f[15] + r2 * f[16];
    const double L1 = f[17], L2 = f[18];
    const double cy0 = yt < -L2 ? -L2 : yt;
    const double cy = L2 < cy0 ? L2 : cy0;
    const double dy = yt - cy;
    const double rho1 = rho_pole_d(xt / L1, sqrt(dy * dy + zt * zt) / L1);
    const double cx0 = xt < -L1 ? -L1 : xt;
    const double cx = L1 < cx0 ? L1 : cx0;
    const double dx = xt - cx;
    const double rho2 = rho_pole_d(yt / L2, sqrt(dx * dx + zt * zt) / L2);
    return rho2 < rho1 ? rho2 : rho1;
}

// One workgroup per source panel i, one thread per target panel j in chunks of 256: candidates are
// target panels with a node in the ball around i (smallest Bernstein radius over those nodes, in
// node order) or, with correct_edges, a corner within tol_corner of one of i's.  Pass WRITE = false
// counts them; WRITE = true writes them, ascending j, at ptr[i].
template <bool WRITE>
__global__ __launch_bounds__(256) void pair_scan_kernel(DevGeom g, int P, int npp, const double *__restrict__ pdata,
                                                        const double *__restrict__ corners,
                                                        const int *__restrict__ is_edge, double tc2, int correct_edges,
                                                        int64_t *__restrict__ count, const int64_t *__restrict__ ptr,
                                                        int *__restrict__ ctrg, double *__restrict__ crho,
                                                        unsigned char *__restrict__ ctouch) {
#pragma clang fp contract(off)
    __shared__ int s_wave[4];
    const int i = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    if (!correct_edges && is_edge[i]) {
        if (!WRITE && tid == 0) count[i] = 0;
        return;
    }
    double fi[kPairData];
    for (int k = 0; k < kPairData; ++k) fi[k] = pdata[(size_t)i * kPairData + k];
    const double cx = fi[0], cy = fi[1], cz = fi[2], ri2 = fi[3], ri = fi[19];
    int64_t running = 0;
    for (int base = 0; base < P; base += 256) {
        const int j = base + tid;
        bool seen = false, touch = false;
        double rmin = INFINITY;
        if (j < P && j != i) {
            const double *fj = pdata + (size_t)j * kPairData;
            const double e0 = fj[0] - cx, e1 = fj[1] - cy, e2 = fj[2] - cz;
            const double dc = sqrt(e0 * e0 + e1 * e1 + e2 * e2);
            // no node of j can be in the ball when dc - rad_j > ri (with slack for rounding)
            if (dc - fj[4] <= ri * (1.0 + 1e-9) + 1e-300) {
                for (int t = 0; t < npp; ++t) {
                    const int64_t n = (int64_t)j * npp + t;
                    const double x = g.x[n], y = g.y[n], z = g.z[n];
                    const double d0 = x - cx, d1 = y - cy, d2 = z - cz;
                    if (d0 * d0 + d1 * d1 + d2 * d2 > ri2) continue;
                    const double rho = bernstein_rho_d(fi, x, y, z);
                    if (!seen) { seen = true; rmin = rho; }
                    else if (rho < rmin) rmin = rho;
                }
            }
            if (correct_edges) {
                const double *ci = corners + (size_t)i * 12, *cj = corners + (size_t)j * 12;
                for (int a = 0; a < 4 && !touch; ++a)
                    for (int b = 0; b < 4 && !touch; ++b) {
                        const double d0 = cj[3 * b] - ci[3 * a], d1 = cj[3 * b + 1] - ci[3 * a + 1],
                                     d2 = cj[3 * b + 2] - ci[3 * a + 2];
                        touch = d0 * d0 + d1 * d1 + d2 * d2 <= tc2;
                    }
            }
        }
        const bool flag = seen || touch;
        const uint64_t mask = __ballot(flag);
        const int before = __popcll(mask & ((1ull << lane) - 1ull));
        if (lane == 0) s_wave[wv] = __popcll(mask);
        __syncthreads();
        int woff = 0, total = 0;
        for (int w = 0; w < 4; ++w) {
            if (w < wv) woff += s_wave[w];
            total += s_wave[w];
        }
        if (WRITE && flag) {
            const int64_t o = ptr[i] + running + woff + before;
            if (o < ptr[i + 1]) {
                ctrg[o] = j;
                crho[o] = seen ? rmin : INFINITY;
                ctouch[o] = touch ? 1 : 0;
            }
        }
        running += total;
        __syncthreads();
    }
    if (!WRITE && tid == 0) count[i] = running;
}

// The near pairs' full interaction in FP64, correction block + direct Dt term (the upsampled value the
// builder split into the two), in the storage order of the FP64 blocks: one workgroup per block row
// (target panel), the direct term computed as near_blocks_kernel computes it.
__global__ __launch_bounds__(256) void near_full64_kernel(DevGeom g, int bd, int mfma, const int *__restrict__ row_ptr,
                                                          const int *__restrict__ col, const double *__restrict__ val64,
                                                          double *__restrict__ full64) {
    const int p = blockIdx.x;
    const int64_t tb = (int64_t)p * bd;
    const double n0 = g.nx[tb], n1 = g.ny[tb], n2 = g.nz[tb];
    for (int b = row_ptr[p]; b < row_ptr[p + 1]; ++b) {
        const int64_t sb = (int64_t)col[b] * bd;
        const double s0 = g.nx[sb], s1 = g.ny[sb], s2 = g.nz[sb];
        const size_t out0 = (size_t)b * bd * bd;
        for (int o = threadIdx.x; o < bd * bd; o += blockDim.x) {
            const int t = o / bd, c = o - t * bd;
            const size_t idx = out0 + store_idx(bd, mfma != 0, t, c);
            const double direct = corr_k<0>(g.x[sb + c], g.y[sb + c], g.z[sb + c], g.x[tb + t], g.y[tb + t],
                                            g.z[tb + t], n0, n1, n2, s0, s1, s2) *
                                  g.w[sb + c];
            full64[idx] = val64[idx] + direct;
        }
    }
}
}  // namespace

void near_full_blocks(hipStream_t st, const DevGeom &g, const DevBSR &C, double *full64) {
    if (C.P < 1 || C.nb < 1) return;
    near_full64_kernel<<<(unsigned)C.P, 256, 0, st>>>(g, C.bd, C.mfma_order ? 1 : 0, C.row_ptr, C.col, C.val64, full64);
}

void pair_scan(hipStream_t st, const DevGeom &g, int P, int npp, const double *pdata, const double *corners,
               const int *is_edge, double tc2, bool correct_edges, int64_t *count, const int64_t *ptr, int *ctrg,
               double *crho, unsigned char *ctouch) {
    if (P <= 0) return;
    if (ptr == nullptr)
        pair_scan_kernel<false><<<P, 256, 0, st>>>(g, P, npp, pdata, corners, is_edge, tc2, correct_edges ? 1 : 0,
                                                   count, nullptr, nullptr, nullptr, nullptr);
    else
        pair_scan_kernel<true><<<P, 256, 0, st>>>(g, P, npp, pdata, corners, is_edge, tc2, correct_edges ? 1 : 0,
                                                  nullptr, ptr, ctrg, crho, ctouch);
}

size_t adaptive_blocks_lds(int nq, int nGL, int max_depth) {
    const size_t npp = (size_t)nq * nq, G2 = (size_t)nGL * nGL;
    const size_t cap = 3 * (size_t)max_depth + 4;
    return sizeof(double) * (cap * (5 + npp) + 4 * npp + 8 * G2 + 8 * (size_t)nGL * nq + 12 * G2 + G2 * npp +
                             npp * npp + 16);
}

bool adaptive_blocks(hipStream_t st, const DevGeom &g, int nq, int nGL, int n_adp, const int *adp_src,
                     const int *adp_trg, const int *adp_rep, const double *adp_map, const double *rules, double atol,
                     double rtol, int max_depth, int max_nodes, double *kup, int *hit, int mode) {
    const size_t lds = adaptive_blocks_lds(nq, nGL, max_depth);
    if (lds > kNearBuildLds || max_depth < 0) return false;
    if (n_adp > 0) {
        if (mode == 1)
            adaptive_blocks_kernel<1><<<n_adp, 64, lds, st>>>(g, nq, nGL, adp_src, adp_trg, adp_rep, adp_map, rules,
                                                              atol, rtol, max_depth, max_nodes, kup, hit);
        else
            adaptive_blocks_kernel<0><<<n_adp, 64, lds, st>>>(g, nq, nGL, adp_src, adp_trg, adp_rep, adp_map, rules,
                                                              atol, rtol, max_depth, max_nodes, kup, hit);
    }
    return true;
}

bool build_near_blocks(hipStream_t st, const DevGeom &g, int bd, bool mfma_order, int n_up, const int *up_src,
                       const int *up_trg, const int *up_slot, const int *cache_off, const int *cache_nu,
                       const double *caches, int max_nu, int n_adp, const int *adp_src, const int *adp_trg,
                       const int *adp_slot, const double *adp_kup, double *val64, float *val32, float *vup32,
                       int mode) {
    const int mo = mfma_order ? 1 : 0;
    if (n_up > 0) {
        // kernel values in LDS; the quadrature matrix too when both fit in 64 KiB
        const size_t kv = sizeof(double) * (size_t)bd * max_nu * max_nu;
        if (kv > kNearBuildLds) return false;
        const int mt_lds = 2 * kv <= kNearBuildLds ? 1 : 0;
        const size_t lds = (1 + mt_lds) * kv;
        if (mode == 1)
            near_blocks_kernel<true, 1><<<n_up, 256, lds, st>>>(g, bd, mo, up_src, up_trg, up_slot, cache_off, cache_nu,
                                                                caches, nullptr, mt_lds, val64, val32, vup32);
        else
            near_blocks_kernel<true, 0><<<n_up, 256, lds, st>>>(g, bd, mo, up_src, up_trg, up_slot, cache_off, cache_nu,
                                                                caches, nullptr, mt_lds, val64, val32, vup32);
    }
    if (n_adp > 0) {
        if (mode == 1)
            near_blocks_kernel<false, 1><<<n_adp, 256, 0, st>>>(g, bd, mo, adp_src, adp_trg, adp_slot, nullptr, nullptr,
                                                                nullptr, adp_kup, 0, val64, val32, vup32);
        else
            near_blocks_kernel<false, 0><<<n_adp, 256, 0, st>>>(g, bd, mo, adp_src, adp_trg, adp_slot, nullptr, nullptr,
                                                                nullptr, adp_kup, 0, val64, val32, vup32);
    }
    return true;
}

}  // namespace dbie
