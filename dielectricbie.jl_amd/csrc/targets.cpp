// targets.cpp -- induced potential at off-surface targets with near-target corrections:
// laplace3d_pottrg_fmm3d_corrected_hcubature (src/kernel/laplace3d_near_hcubature.jl:337-375).
//
//   u(t) = sum_r w_r (P sigma)_r / (4 pi |t - y_r|) + sum_r C[t, r] (P sigma)_r
//
// y_r are the nodes of the interface refined around the targets
// (_refine_interface_for_targets, src/kernel/laplace3d_near.jl:99-150), P interpolates the
// coarse density onto them (_refined_interface_prolongation, :152-225), and C holds, for every
// non-edge refined panel and every target within range_factor * l / n_quad of its centre
// (build_target_neighbor_list, :47-70), the adaptively integrated panel potential against each
// Lagrange basis function minus the direct quadrature term
// (laplace3d_pottrg_corrections_hcubature, src/kernel/laplace3d_near_hcubature.jl:193-251).
//
// The adaptive cubature is the two-dimensional Genz-Malik rule (degree 7 with an embedded
// degree-5 error estimate, 17 points; A. C. Genz and A. A. Malik, J. Comput. Appl. Math. 6
// (1980) 295-302) driven by bisection of the box with the largest error along the axis of the
// largest fourth divided difference until the summed error is <= atol -- the scheme of
// HCubature.jl's hcubature(f, a, b; atol), which the reference calls.  Range queries use a
// uniform cell grid over the targets with the test |t - c|^2 <= r^2.
#include <algorithm>
#include <array>
#include <chrono>
#include <cfloat>
#include <cmath>
#include <memory>
#include <numeric>

#include "internal.hpp"

namespace dbie {
namespace {

inline double len3(double a, double b, double c) { return std::sqrt(a * a + b * b + c * c); }

// ---- uniform grid over the targets -------------------------------------------------------
struct TargetGrid {
    const double *t = nullptr;
    double lo[3] = {0, 0, 0}, h = 1;
    int64_t dim[3] = {1, 1, 1};
    std::vector<int64_t> start;
    std::vector<int> idx;

    TargetGrid(const double *trg, int64_t M, double cell) : t(trg) {
        double hi[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
        for (int k = 0; k < 3; ++k) lo[k] = DBL_MAX;
        for (int64_t j = 0; j < M; ++j)
            for (int k = 0; k < 3; ++k) {
                lo[k] = std::min(lo[k], trg[3 * j + k]);
                hi[k] = std::max(hi[k], trg[3 * j + k]);
            }
        if (M == 0) {
            for (int k = 0; k < 3; ++k) lo[k] = hi[k] = 0;
        }
        double ext = 0;
        for (int k = 0; k < 3; ++k) ext = std::max(ext, hi[k] - lo[k]);
        h = std::max(cell, ext / 256 + 1e-300);
        for (int k = 0; k < 3; ++k) dim[k] = std::max<int64_t>(1, (int64_t)((hi[k] - lo[k]) / h) + 1);
        const int64_t nc = dim[0] * dim[1] * dim[2];
        start.assign(nc + 1, 0);
        std::vector<int64_t> cell_of(M);
        for (int64_t j = 0; j < M; ++j) {
            cell_of[j] = cell_index(&trg[3 * j]);
            ++start[cell_of[j] + 1];
        }
        for (int64_t c = 0; c < nc; ++c) start[c + 1] += start[c];
        idx.resize(M);
        std::vector<int64_t> fill(start.begin(), start.end() - 1);
        for (int64_t j = 0; j < M; ++j) idx[fill[cell_of[j]]++] = (int)j;
    }
    int64_t coord(int k, double v) const {
        const int64_t c = (int64_t)std::floor((v - lo[k]) / h);
        return std::min(std::max<int64_t>(c, 0), dim[k] - 1);
    }
    int64_t cell_index(const double *p) const { return (coord(0, p[0]) * dim[1] + coord(1, p[1])) * dim[2] + coord(2, p[2]); }
    // visit(j) for every target with |t_j - c|^2 <= r^2; stops when visit returns false
    template <class F>
    void ball(const double *c, double r, F &&visit) const {
        int64_t a[3], b[3];
        for (int k = 0; k < 3; ++k) {
            a[k] = coord(k, c[k] - r);
            b[k] = coord(k, c[k] + r);
        }
        const double r2 = r * r;
        for (int64_t i = a[0]; i <= b[0]; ++i)
            for (int64_t j = a[1]; j <= b[1]; ++j)
                for (int64_t k = a[2]; k <= b[2]; ++k) {
                    const int64_t cc = (i * dim[1] + j) * dim[2] + k;
                    for (int64_t q = start[cc]; q < start[cc + 1]; ++q) {
                        const int id = idx[q];
                        const double *p = &t[3 * (int64_t)id];
                        const double dx = p[0] - c[0], dy = p[1] - c[1], dz = p[2] - c[2];
                        if (dx * dx + dy * dy + dz * dz <= r2)
                            if (!visit(id)) return;
                    }
                }
    }
};

void panel_center(const double *c, double *cc) {
    for (int k = 0; k < 3; ++k) cc[k] = (c[k] + c[3 + k] + c[6 + k] + c[9 + k]) / 4;
}
double panel_max_length(const double *c) {
    return std::max(len3(c[0] - c[3], c[1] - c[4], c[2] - c[5]), len3(c[3] - c[6], c[4] - c[7], c[5] - c[8]));
}

// _split_panel4 (src/kernel/laplace3d_near.jl:78-97): children (a,ab,cc,da), (ab,b,bc,cc),
// (cc,bc,c,cd), (da,cc,cd,d)
void split4(const double *c, double out[4][12]) {
    const double *a = c, *b = c + 3, *cc3 = c + 6, *d = c + 9;
    double ab[3], bc[3], cd[3], da[3], m[3];
    for (int k = 0; k < 3; ++k) {
        ab[k] = (a[k] + b[k]) / 2;
        bc[k] = (b[k] + cc3[k]) / 2;
        cd[k] = (cc3[k] + d[k]) / 2;
        da[k] = (d[k] + a[k]) / 2;
        m[k] = (a[k] + b[k] + cc3[k] + d[k]) / 4;
    }
    const double *quads[4][4] = {{a, ab, m, da}, {ab, b, bc, m}, {m, bc, cc3, cd}, {da, m, cd, d}};
    for (int q = 0; q < 4; ++q)
        for (int v = 0; v < 4; ++v)
            for (int k = 0; k < 3; ++k) out[q][3 * v + k] = quads[q][v][k];
}

// ---- Genz-Malik rule in two dimensions ---------------------------------------------------
constexpr int kMaxVals = 64;
struct Box {
    double a[2], b[2];
    double I[kMaxVals];
    double E;
    int kdiv;
    int64_t seq;
};

struct GMRule {
    double l2, l3, l4, l5;
    double w[5], wp[4];
    GMRule() {
        l2 = std::sqrt(9.0 / 70.0);
        l3 = std::sqrt(9.0 / 10.0);
        l4 = std::sqrt(9.0 / 10.0);
        l5 = std::sqrt(9.0 / 19.0);
        const double n = 2, p2 = 4;  // 2^n
        w[0] = p2 * (12824 - 9120 * n + 400 * n * n) / 19683;
        w[1] = p2 * 980 / 6561;
        w[2] = p2 * (1820 - 400 * n) / 19683;
        w[3] = p2 * 200 / 19683;
        w[4] = 6859.0 / 19683;
        wp[0] = p2 * (729 - 950 * n + 50 * n * n) / 729;
        wp[1] = p2 * 245 / 486;
        wp[2] = p2 * (265 - 100 * n) / 1458;
        wp[3] = p2 * 25 / 729;
    }
};

inline double vnorm(const double *v, int m) {
    double s = 0;
    for (int i = 0; i < m; ++i) s += v[i] * v[i];
    return std::sqrt(s);
}

// One rule application on [a, b]: fills box.I, box.E, box.kdiv.  f(u, v, out[m]).
template <class F>
void gm_apply(const GMRule &g, F &f, int m, Box &bx) {
    double c[2], D[2];
    for (int k = 0; k < 2; ++k) {
        c[k] = 0.5 * (bx.a[k] + bx.b[k]);
        D[k] = 0.5 * std::fabs(bx.b[k] - bx.a[k]);
    }
    const double V = D[0] * D[1];
    double f1[kMaxVals], f2[kMaxVals], f3[kMaxVals], f4[kMaxVals], f5[kMaxVals], tp[kMaxVals], tm[kMaxVals];
    f(c[0], c[1], f1);
    for (int i = 0; i < m; ++i) f2[i] = f3[i] = f4[i] = f5[i] = 0;
    bx.kdiv = 0;
    double maxdd = 0;
    for (int ax = 0; ax < 2; ++ax) {
        double f2i[kMaxVals], f3i[kMaxVals];
        double p[2] = {0, 0};
        p[ax] = D[ax] * g.l2;
        f(c[0] + p[0], c[1] + p[1], tp);
        f(c[0] - p[0], c[1] - p[1], tm);
        for (int i = 0; i < m; ++i) f2i[i] = tp[i] + tm[i];
        p[ax] = D[ax] * g.l3;
        f(c[0] + p[0], c[1] + p[1], tp);
        f(c[0] - p[0], c[1] - p[1], tm);
        for (int i = 0; i < m; ++i) f3i[i] = tp[i] + tm[i];
        double dd = 0;
        for (int i = 0; i < m; ++i) {
            f2[i] += f2i[i];
            f3[i] += f3i[i];
            const double e = f3i[i] + 12 * f1[i] - 7 * f2i[i];  // fourth divided difference
            dd += e * e;
        }
        dd = std::sqrt(dd);
        if (dd > maxdd) {
            maxdd = dd;
            bx.kdiv = ax;
        }
    }
    for (int s0 = -1; s0 <= 1; s0 += 2)
        for (int s1 = -1; s1 <= 1; s1 += 2) {
            f(c[0] + s0 * D[0] * g.l4, c[1] + s1 * D[1] * g.l4, tp);
            f(c[0] + s0 * D[0] * g.l5, c[1] + s1 * D[1] * g.l5, tm);
            for (int i = 0; i < m; ++i) {
                f4[i] += tp[i];
                f5[i] += tm[i];
            }
        }
    double e2 = 0;
    for (int i = 0; i < m; ++i) {
        bx.I[i] = V * (g.w[0] * f1[i] + g.w[1] * f2[i] + g.w[2] * f3[i] + g.w[3] * f4[i] + g.w[4] * f5[i]);
        const double Ip = V * (g.wp[0] * f1[i] + g.wp[1] * f2[i] + g.wp[2] * f3[i] + g.wp[3] * f4[i]);
        e2 += (bx.I[i] - Ip) * (bx.I[i] - Ip);
    }
    bx.E = std::sqrt(e2);
}

struct BoxLess {  // max-heap on E; among equal errors the older box first
    bool operator()(const Box &x, const Box &y) const { return x.E < y.E || (x.E == y.E && x.seq > y.seq); }
};

// Adaptive cubature of f over [-1,1]^2 to summed error <= atol.  Returns false if the
// evaluation cap was reached first.
template <class F>
bool hcubature2(const GMRule &g, F &f, int m, double atol, int64_t maxevals, double *I, std::vector<Box> &heap) {
    heap.clear();
    Box b0;
    b0.a[0] = b0.a[1] = -1;
    b0.b[0] = b0.b[1] = 1;
    b0.seq = 0;
    gm_apply(g, f, m, b0);
    int64_t evals = 17;
    if (b0.E <= atol || evals >= maxevals) {
        std::copy(b0.I, b0.I + m, I);
        return b0.E <= atol;
    }
    double Itot[kMaxVals];
    std::copy(b0.I, b0.I + m, Itot);
    double Etot = b0.E;
    heap.push_back(b0);
    int64_t seq = 1;
    for (;;) {
        std::pop_heap(heap.begin(), heap.end(), BoxLess());
        const Box old = heap.back();
        heap.pop_back();
        const int kd = old.kdiv;
        const double mid = old.a[kd] + (old.b[kd] - old.a[kd]) / 2;
        Box c1 = old, c2 = old;
        c1.b[kd] = mid;
        c2.a[kd] = mid;
        c1.seq = seq++;
        c2.seq = seq++;
        gm_apply(g, f, m, c1);
        gm_apply(g, f, m, c2);
        evals += 34;
        for (int i = 0; i < m; ++i) Itot[i] += c1.I[i] + c2.I[i] - old.I[i];
        Etot += c1.E + c2.E - old.E;
        heap.push_back(c1);
        std::push_heap(heap.begin(), heap.end(), BoxLess());
        heap.push_back(c2);
        std::push_heap(heap.begin(), heap.end(), BoxLess());
        if (Etot <= atol || evals >= maxevals) break;
    }
    // re-sum over the boxes in sequence order (independent of heap layout)
    std::sort(heap.begin(), heap.end(), [](const Box &x, const Box &y) { return x.seq < y.seq; });
    double Es = 0;
    for (int i = 0; i < m; ++i) I[i] = 0;
    for (const Box &bx : heap) {
        for (int i = 0; i < m; ++i) I[i] += bx.I[i];
        Es += bx.E;
    }
    return Etot <= atol;
}

}  // namespace

std::array<double, 13> genz_malik_rule() {
    const GMRule g;
    return {g.l2, g.l3, g.l4, g.l5, g.w[0], g.w[1], g.w[2], g.w[3], g.w[4], g.wp[0], g.wp[1], g.wp[2], g.wp[3]};
}

TargetMap *build_target_map(const Geometry &g, int64_t M, const double *trg, double atol, double range_factor,
                            int n_threads, const TargetCubatureFn &device, const TargetRefineFn &refine) {
    if (M < 0 || (M > 0 && !trg)) fail(DBIE_BAD_ARG, "targets: need M >= 0 and a 3 x M array");
    if (!(atol > 0)) fail(DBIE_BAD_ARG, "hcubature atol must be positive");
    if (!(range_factor > 0)) fail(DBIE_BAD_ARG, "range_factor must be positive");
    if (M > INT32_MAX) fail(DBIE_BAD_ARG, "too many targets");
    const int nq = g.nq, npp = g.npp();
    if (npp > kMaxVals) fail(DBIE_BAD_ARG, "n_quad too large for the target corrections");
    std::unique_ptr<TargetMap> tm(new TargetMap);
    tm->M = M;
    tm->trg.assign(trg, trg + 3 * M);

    // panel size limit = smallest panel (max side)
    double limit = DBL_MAX;
    for (int64_t p = 0; p < g.P; ++p) limit = std::min(limit, panel_max_length(&g.corners[12 * p]));
    using Clk = std::chrono::steady_clock;
    auto ms_since = [](Clk::time_point a) { return std::chrono::duration<double, std::milli>(Clk::now() - a).count(); };
    auto tp = Clk::now();
    const TargetGrid grid(tm->trg.data(), M, range_factor * limit / nq);

    if (refine) {  // both phases on the device
        RefineInput in;
        in.g = &g;
        in.trg = tm->trg.data();
        in.M = M;
        for (int k = 0; k < 3; ++k) {
            in.lo[k] = grid.lo[k];
            in.dim[k] = grid.dim[k];
        }
        in.h = grid.h;
        in.start = &grid.start;
        in.idx = &grid.idx;
        in.limit = limit;
        in.range_factor = range_factor;
        RefineOutput out;
        if (refine(in, out)) {
            const int64_t Pr = (int64_t)out.parent.size();
            std::vector<double> rn((size_t)Pr * 3), pts((size_t)Pr * npp * 3);
            std::vector<int> re(Pr);
            for (int64_t q = 0; q < Pr; ++q) {
                for (int k = 0; k < 3; ++k) rn[3 * q + k] = g.normal[3 * (int64_t)out.parent[q] + k];
                re[q] = g.is_edge[out.parent[q]];
            }
            for (int64_t n = 0; n < Pr * npp; ++n) {
                pts[3 * n] = out.x[n];
                pts[3 * n + 1] = out.y[n];
                pts[3 * n + 2] = out.z[n];
            }
            double gx[16], gw[16];
            gauss_legendre(nq, gx, gw);  // the rule make_from_panels gives the refined interface
            std::unique_ptr<Geometry> gr(make_from_panel_nodes(Pr, nq, out.corners.data(), rn.data(), re.data(),
                                                               pts.data(), out.w.data(), gx, gw, nullptr));
            tm->ref = std::move(*gr);
            tm->parent = std::move(out.parent);
            tm->from_split = std::move(out.from_split);
            tm->p_ptr = std::move(out.p_ptr);
            tm->prolong_device = true;
            tm->setup_ms[0] = out.ms_refine;
            tm->setup_ms[1] = ms_since(tp) - out.ms_refine;
            tp = Clk::now();
        }
    }
    if (!tm->prolong_device) {
    // ---- refinement around the targets (stack, last in first out) ----
    struct Item {
        int parent;
        bool split;
        double c[12];
    };
    std::vector<Item> stack;
    stack.reserve(g.P * 2);
    for (int64_t p = 0; p < g.P; ++p) {
        Item it;
        it.parent = (int)p;
        it.split = false;
        std::copy(&g.corners[12 * p], &g.corners[12 * p] + 12, it.c);
        stack.push_back(it);
    }
    std::vector<double> rc, rn;
    std::vector<int> re;
    while (!stack.empty()) {
        const Item it = stack.back();
        stack.pop_back();
        const double l = panel_max_length(it.c);
        double cc[3];
        panel_center(it.c, cc);
        const double r = range_factor * l / nq;
        bool near = false;
        if (M > 0) grid.ball(cc, r, [&](int) { near = true; return false; });
        if (near && l > limit) {
            double ch[4][12];
            split4(it.c, ch);
            for (int q = 0; q < 4; ++q) {
                Item c;
                c.parent = it.parent;
                c.split = true;
                std::copy(ch[q], ch[q] + 12, c.c);
                stack.push_back(c);
            }
        } else {
            rc.insert(rc.end(), it.c, it.c + 12);
            rn.insert(rn.end(), &g.normal[3 * (int64_t)it.parent], &g.normal[3 * (int64_t)it.parent] + 3);
            re.push_back(g.is_edge[it.parent]);
            tm->parent.push_back(it.parent);
            tm->from_split.push_back(it.split ? 1 : 0);
        }
    }
    const int64_t Pr = (int64_t)tm->parent.size();
    {
        std::unique_ptr<Geometry> gr(make_from_panels(Pr, nq, rc.data(), rn.data(), re.data()));
        tm->ref = std::move(*gr);
    }
    const Geometry &R = tm->ref;

    tm->setup_ms[0] = ms_since(tp);
    tp = Clk::now();
    // ---- prolongation (coarse density -> refined nodes) ----
    // a refined node of an unsplit panel copies its coarse node; one of a split panel takes the
    // tensor Lagrange interpolant of its parent panel (n_quad^2 entries).  Rows are counted first,
    // then filled in parallel.
    tm->p_ptr.assign((size_t)Pr * npp + 1, 0);
    for (int64_t q = 0; q < Pr; ++q)
        for (int k = 0; k < npp; ++k) tm->p_ptr[q * npp + k + 1] = tm->from_split[q] ? npp : 1;
    for (int64_t r = 0; r < Pr * npp; ++r) tm->p_ptr[r + 1] += tm->p_ptr[r];
    tm->p_col.resize(tm->p_ptr.back());
    tm->p_val.resize(tm->p_ptr.back());
    {
        const int nt = n_threads > 0 ? n_threads : (int)std::max(1u, std::thread::hardware_concurrency());
        parallel_for(Pr, nt, [&](int64_t q, int) {
            const int par = tm->parent[q];
            const int64_t col0 = (int64_t)par * npp;
            if (!tm->from_split[q]) {
                for (int k = 0; k < npp; ++k) {
                    const int64_t e = tm->p_ptr[q * npp + k];
                    tm->p_col[e] = (int)(col0 + k);
                    tm->p_val[e] = 1.0;
                }
                return;
            }
            double rx[16], ry[16];
            const double *c = &g.corners[12 * (int64_t)par];
            const double *a = c, *b = c + 3, *d = c + 9;
            double cc[3], bma[3], dma[3];
            panel_center(c, cc);
            for (int k = 0; k < 3; ++k) {
                bma[k] = b[k] - a[k];
                dma[k] = d[k] - a[k];
            }
            const double bb = bma[0] * bma[0] + bma[1] * bma[1] + bma[2] * bma[2];
            const double bd = bma[0] * dma[0] + bma[1] * dma[1] + bma[2] * dma[2];
            const double dd = dma[0] * dma[0] + dma[1] * dma[1] + dma[2] * dma[2];
            const double det = bb * dd - bd * bd;
            const double inv11 = dd / det, inv12 = -bd / det, inv22 = bb / det;
            for (int k = 0; k < npp; ++k) {
                const int64_t n = q * npp + k;
                const double x[3] = {R.x[n] - cc[0], R.y[n] - cc[1], R.z[n] - cc[2]};
                const double rhs1 = bma[0] * x[0] + bma[1] * x[1] + bma[2] * x[2];
                const double rhs2 = dma[0] * x[0] + dma[1] * x[1] + dma[2] * x[2];
                const double s = inv11 * rhs1 + inv12 * rhs2;
                const double t = inv12 * rhs1 + inv22 * rhs2;
                barycentric_row(nq, g.glx, g.bary, 2 * s, rx);
                barycentric_row(nq, g.glx, g.bary, 2 * t, ry);
                int64_t e = tm->p_ptr[n];
                for (int ii = 0; ii < nq; ++ii)
                    for (int jj = 0; jj < nq; ++jj, ++e) {
                        tm->p_col[e] = (int)(col0 + ii * nq + jj);
                        tm->p_val[e] = rx[ii] * ry[jj];
                    }
            }
        });
    }
    tm->setup_ms[1] = ms_since(tp);
    tp = Clk::now();
    }  // host refinement and prolongation
    const Geometry &R = tm->ref;
    const int64_t Pr = (int64_t)tm->parent.size();
    // ---- near-target corrections on the refined interface ----
    // pairs (refined panel, target) in panel order, targets ascending per panel
    const int nt = n_threads > 0 ? n_threads : (int)std::max(1u, std::thread::hardware_concurrency());
    std::vector<std::vector<int>> near_of(Pr);
    parallel_for(Pr, nt, [&](int64_t q, int) {
        if (R.is_edge[q] || M == 0) return;
        const double *c = &R.corners[12 * q];
        double cc[3];
        panel_center(c, cc);
        const double l = panel_max_length(c);
        std::vector<int> &near = near_of[q];
        grid.ball(cc, range_factor * l / nq, [&](int id) { near.push_back(id); return true; });
        std::sort(near.begin(), near.end());
    });
    TargetCubatureJob job;
    job.nq = nq;
    job.atol = atol;
    job.maxevals = 2000000;
    for (int64_t q = 0; q < Pr; ++q)
        for (const int tg : near_of[q]) {
            job.panel.push_back((int)q);
            job.target.push_back(tg);
        }
    std::vector<std::vector<int>>().swap(near_of);
    const int64_t npairs = (int64_t)job.panel.size();
    if (npairs > INT32_MAX) fail(DBIE_BAD_ARG, "too many near-target pairs");
    job.pdat.resize((size_t)npairs * 13);
    parallel_for(npairs, nt, [&](int64_t k, int) {
        const double *c = &R.corners[12 * (int64_t)job.panel[k]];
        double *d = &job.pdat[(size_t)k * 13];
        panel_center(c, d);
        for (int i = 0; i < 3; ++i) {
            d[3 + i] = c[3 + i] - c[i];  // b - a
            d[6 + i] = c[9 + i] - c[i];  // d - a
        }
        d[9] = len3(d[3], d[4], d[5]) * len3(d[6], d[7], d[8]) / 4;
        for (int i = 0; i < 3; ++i) d[10 + i] = tm->trg[3 * (int64_t)job.target[k] + i];
    });
    std::vector<double> I((size_t)npairs * npp);
    std::vector<int> st(npairs, 2);
    if (device && npairs > 0) device(job, g, I.data(), st.data());
    // the host integrates every pair the device did not (all of them without a device)
    const GMRule rule;
    parallel_for(npairs, nt, [&](int64_t k, int) {
        if (st[k] != 2) return;
        const double *d = &job.pdat[(size_t)k * 13];
        auto f = [&](double u, double v, double *vals) {
            double rx[16], ry[16];
            barycentric_row(nq, g.glx, g.bary, u, rx);
            barycentric_row(nq, g.glx, g.bary, v, ry);
            double p[3];
            for (int i = 0; i < 3; ++i) p[i] = d[i] + d[3 + i] * (u / 2) + d[6 + i] * (v / 2);
            const double dx = p[0] - d[10], dy = p[1] - d[11], dz = p[2] - d[12];
            const double kv = (1.0 / std::sqrt(dx * dx + dy * dy + dz * dz)) / kFourPi * d[9];
            for (int ii = 0; ii < nq; ++ii)
                for (int jj = 0; jj < nq; ++jj) vals[ii * nq + jj] = kv * rx[ii] * ry[jj];
        };
        std::vector<Box> heap;
        st[k] = hcubature2(rule, f, npp, atol, job.maxevals, &I[(size_t)k * npp], heap) ? 0 : 1;
    });
    tm->setup_ms[2] = ms_since(tp);
    tp = Clk::now();
    // corrections: integral minus the direct quadrature term
    struct Trip {
        int t, c;
        double v;
    };
    std::vector<std::vector<Trip>> out(nt);
    parallel_for(npairs, nt, [&](int64_t k, int tid) {
        const int64_t q = job.panel[k];
        const int tg = job.target[k];
        const double *tp = &tm->trg[3 * (int64_t)tg];
        for (int i = 0; i < npp; ++i) {
            const int64_t n = q * npp + i;
            const double dx = R.x[n] - tp[0], dy = R.y[n] - tp[1], dz = R.z[n] - tp[2];
            const double direct = R.w[n] * ((1.0 / std::sqrt(dx * dx + dy * dy + dz * dz)) / kFourPi);
            const double v = I[(size_t)k * npp + i] - direct;
            if (v == 0.0) continue;
            out[tid].push_back(Trip{tg, (int)n, v});
        }
    });
    size_t total = 0;
    for (auto &o : out) total += o.size();
    std::vector<Trip> all;
    all.reserve(total);
    for (auto &o : out) {
        all.insert(all.end(), o.begin(), o.end());
        std::vector<Trip>().swap(o);
    }
    std::sort(all.begin(), all.end(), [](const Trip &x, const Trip &y) { return x.t < y.t || (x.t == y.t && x.c < y.c); });
    tm->c_ptr.assign(M + 1, 0);
    tm->c_col.resize(all.size());
    tm->c_val.resize(all.size());
    for (size_t e = 0; e < all.size(); ++e) {
        ++tm->c_ptr[all[e].t + 1];
        tm->c_col[e] = all[e].c;
        tm->c_val[e] = all[e].v;
    }
    for (int64_t t = 0; t < M; ++t) tm->c_ptr[t + 1] += tm->c_ptr[t];
    tm->setup_ms[3] = ms_since(tp);
    tm->n_pairs = npairs;
    tm->n_capped = std::count(st.begin(), st.end(), 1);
    tm->n_device = device ? npairs - std::count(st.begin(), st.end(), 2) : 0;
    return tm.release();
}

}  // namespace dbie
