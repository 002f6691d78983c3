// api_targets.cpp -- C ABI of the off-surface target map and the reaction matrix.
//
//   dbie_tmap_create      laplace3d_pottrg_fmm3d_corrected_hcubature (src/kernel/laplace3d_near_hcubature.jl:337-375)
//   dbie_tmap_apply       the returned LinearMap applied to densities: U = u[Sigma] at the targets
//   dbie_reaction_matrix  four_index_matrix (src/solver/multi_rhs.jl:299-321), induced term
//
// dbie_tmap_create builds the refined interface, the prolongation and the near-target
// corrections on the host (targets.cpp, no device needed); dbie_tmap_create_ex can integrate the
// near-target cubatures on the device (bit-identical results).  The first evaluation uploads the
// map and the device evaluates U or R.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

#include <array>
#include <chrono>
#include <cstring>
#include <memory>

#include "internal.hpp"
#include "kernels.hpp"
#include "runtime.hpp"

using namespace dbie;

struct dbie_tmap {
    std::unique_ptr<TargetMap> h;
    int64_t N = 0;  // coarse nodes
    hipStream_t stream = nullptr;
    rocblas_handle blas = nullptr;
    DBuf x, y, z, w, trg;                      // refined nodes, targets
    DBuf p_ptr, p_col, p_val;                  // prolongation (N_ref rows)
    DBuf pt_ptr, pt_col, pt_val;               // transposed prolongation (N rows)
    DBuf par_first, par_count;                 // refined nodes of each coarse panel (consecutive)
    bool par_blocks = false;                   // the prolongation is block-structured by coarse panel
    DBuf c_ptr, c_col, c_val;                  // corrections (M rows)
    DBuf xr, xin, uout;                        // workspaces
    bool resident = false;                     // everything above is on the device
    DevGeom dg;
    ~dbie_tmap() {
        if (blas) (void)rocblas_destroy_handle(blas);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

void check_tm(const dbie_tmap *tm) {
    if (!tm) fail(DBIE_BAD_ARG, "target map handle is NULL");
}

void blas_check(rocblas_status s, const char *what) {
    if (s != rocblas_status_success) fail(DBIE_HIP_ERROR, std::string(what) + ": rocBLAS status " + std::to_string((int)s));
}

// Uploads the refined nodes, targets and sparse matrices on first device use.
void make_resident(dbie_tmap *tm) {
    if (tm->resident) return;
    ensure_device();
    const TargetMap &h = *tm->h;
    if (!tm->stream) DBIE_HIP(hipStreamCreateWithFlags(&tm->stream, hipStreamNonBlocking));
    hipStream_t st = tm->stream;
    const Geometry &R = h.ref;
    if (!h.prolong_device) {  // the device builder left the refined nodes and the prolongation there
        upload(tm->x, R.x.data(), R.N, st);
        upload(tm->y, R.y.data(), R.N, st);
        upload(tm->z, R.z.data(), R.N, st);
        upload(tm->w, R.w.data(), R.N, st);
        upload(tm->p_ptr, h.p_ptr.data(), h.p_ptr.size(), st);
        upload(tm->p_col, h.p_col.data(), h.p_col.size(), st);
        upload(tm->p_val, h.p_val.data(), h.p_val.size(), st);
    }
    upload(tm->trg, h.trg.data(), h.trg.size(), st);
    upload(tm->c_ptr, h.c_ptr.data(), h.c_ptr.size(), st);
    upload(tm->c_col, h.c_col.data(), h.c_col.size(), st);
    upload(tm->c_val, h.c_val.data(), h.c_val.size(), st);
    if (h.prolong_device) {
        // by construction the refined nodes of coarse panel c are consecutive (its subtree is walked
        // before the next panel's) and interpolate from c's nodes only
        const int npp = h.ref.nq * h.ref.nq;
        const int64_t P = tm->N / std::max(1, npp);
        std::vector<int64_t> first(P, -1);
        std::vector<int> count(P, 0);
        for (size_t qp = 0; qp < h.parent.size(); ++qp) {
            const int c = h.parent[qp];
            if (first[c] < 0) first[c] = (int64_t)qp * npp;
            count[c] += npp;
        }
        upload(tm->par_first, first.data(), first.size(), st);
        upload(tm->par_count, count.data(), count.size(), st);
        tm->par_blocks = true;
    } else {  // transpose of the prolongation: rows = coarse nodes
        const int64_t N = tm->N, Nr = R.N;
        std::vector<int64_t> tp(N + 1, 0);
        for (int c : h.p_col) ++tp[c + 1];
        for (int64_t c = 0; c < N; ++c) tp[c + 1] += tp[c];
        std::vector<int> tc(h.p_col.size());
        std::vector<double> tv(h.p_col.size());
        std::vector<int64_t> fill(tp.begin(), tp.end() - 1);
        for (int64_t r = 0; r < Nr; ++r)
            for (int64_t e = h.p_ptr[r]; e < h.p_ptr[r + 1]; ++e) {
                const int64_t at = fill[h.p_col[e]]++;
                tc[at] = (int)r;
                tv[at] = h.p_val[e];
            }
        upload(tm->pt_ptr, tp.data(), tp.size(), st);
        upload(tm->pt_col, tc.data(), tc.size(), st);
        upload(tm->pt_val, tv.data(), tv.size(), st);
    }
    if (!h.prolong_device) {  // refined nodes per coarse panel: consecutive (the refinement finishes a panel's subtree
       // before the next panel), interpolating from that panel's nodes only
        const int64_t P = tm->N / std::max(1, h.ref.nq * h.ref.nq);
        const int npp = h.ref.nq * h.ref.nq;
        std::vector<int64_t> first(P, -1);
        std::vector<int> count(P, 0);
        bool ok = P * npp == tm->N;
        int prev = -1;
        for (size_t qp = 0; qp < h.parent.size() && ok; ++qp) {
            const int c = h.parent[qp];
            if (c < 0 || c >= P) { ok = false; break; }
            if (c != prev) {
                if (first[c] >= 0) { ok = false; break; }  // a second run of the same parent
                first[c] = (int64_t)qp * npp;
                prev = c;
            }
            count[c] += npp;
            for (int64_t r = (int64_t)qp * npp; r < (int64_t)(qp + 1) * npp && ok; ++r)
                for (int64_t e = h.p_ptr[r]; e < h.p_ptr[r + 1]; ++e)
                    if (h.p_col[e] / npp != c) { ok = false; break; }
        }
        for (int64_t c = 0; c < P && ok; ++c) ok = first[c] >= 0;
        tm->par_blocks = ok && npp <= 32;
        if (tm->par_blocks) {
            upload(tm->par_first, first.data(), first.size(), st);
            upload(tm->par_count, count.data(), count.size(), st);
        }
    }
    DevGeom &d = tm->dg;
    d.N = R.N;
    d.x = tm->x.as<double>();
    d.y = tm->y.as<double>();
    d.z = tm->z.as<double>();
    d.w = tm->w.as<double>();
    DBIE_HIP(hipStreamSynchronize(st));
    tm->resident = true;
}

// U (M x K) = base(P Sigma) + C (P Sigma), everything on the device.
void tmap_apply_dev(dbie_tmap *tm, int K, const double *dS, int64_t lds, double *dU, int64_t ldu) {
    make_resident(tm);
    const TargetMap &h = *tm->h;
    const int64_t Nr = h.ref.N, M = h.M;
    hipStream_t st = tm->stream;
    tm->xr.ensure(sizeof(double) * (size_t)Nr * K);
    double *Xr = tm->xr.as<double>();
    csr_spmm(st, Nr, tm->p_ptr.as<int64_t>(), tm->p_col.as<int>(), tm->p_val.as<double>(), K, dS, lds, Xr, Nr, false);
    if (M > 0) {
        pottrg(st, tm->dg, M, tm->trg.as<double>(), K, Xr, Nr, dU, ldu);
        csr_spmm(st, M, tm->c_ptr.as<int64_t>(), tm->c_col.as<int>(), tm->c_val.as<double>(), K, Xr, Nr, dU, ldu, true);
    }
    DBIE_HIP(hipGetLastError());
}

// R (G x K, host, ldr) for densities on the device.
void reaction_dev(dbie_tmap *tm, int G, const int64_t *grp_ptr, const double *q, int K, const double *dS, int64_t lds,
                  double *R, int64_t ldr) {
    make_resident(tm);
    const TargetMap &h = *tm->h;
    const int64_t Nr = h.ref.N, M = h.M, N = tm->N;
    if (G < 1) fail(DBIE_BAD_ARG, "need at least one group");
    if (!grp_ptr || !q || !R) fail(DBIE_BAD_ARG, "NULL group pointer, weights or output");
    if (grp_ptr[0] != 0 || grp_ptr[G] != M) fail(DBIE_DIM_MISMATCH, "groups must cover targets 0..M-1");
    for (int a = 0; a < G; ++a)
        if (grp_ptr[a + 1] < grp_ptr[a]) fail(DBIE_BAD_ARG, "group pointers must be non-decreasing");
    if (ldr < G) fail(DBIE_DIM_MISMATCH, "ldr smaller than the number of groups");
    hipStream_t st = tm->stream;
    // near-target part of the weights: Qcorr[r, a] = sum_{s in a} q_s C[s, r]  (host, sparse)
    std::vector<std::vector<std::pair<int64_t, double>>> per(G);
    parallel_for(G, 0, [&](int64_t a, int) {
        std::vector<double> acc;
        std::vector<int> touched;
        std::vector<char> seen;
        for (int64_t s = grp_ptr[a]; s < grp_ptr[a + 1]; ++s)
            for (int64_t e = h.c_ptr[s]; e < h.c_ptr[s + 1]; ++e) {
                if (acc.empty()) {
                    acc.assign(Nr, 0.0);
                    seen.assign(Nr, 0);
                }
                const int r = h.c_col[e];
                if (!seen[r]) {
                    seen[r] = 1;
                    touched.push_back(r);
                }
                acc[r] += q[s] * h.c_val[e];
            }
        std::sort(touched.begin(), touched.end());
        auto &o = per[a];
        o.reserve(touched.size());
        for (int r : touched) o.emplace_back((int64_t)r + a * Nr, acc[r]);
    });
    std::vector<int64_t> cidx;
    std::vector<double> cval;
    for (auto &o : per)
        for (auto &e : o) {
            cidx.push_back(e.first);
            cval.push_back(e.second);
        }
    DBuf dgrp, dq, dQt, dQct, dR, dci, dcv;
    upload(dgrp, grp_ptr, (size_t)G + 1, st);
    upload(dq, q, (size_t)M, st);
    upload(dci, cidx.data(), cidx.size(), st);
    upload(dcv, cval.data(), cval.size(), st);
    dQt.ensure(sizeof(double) * (size_t)Nr * G);
    dQct.ensure(sizeof(double) * (size_t)N * G);
    dR.ensure(sizeof(double) * (size_t)G * K);
    reaction_q(st, tm->dg, G, dgrp.as<int64_t>(), tm->trg.as<double>(), dq.as<double>(), dQt.as<double>());
    scatter_add(st, (int64_t)cidx.size(), dci.as<int64_t>(), dcv.as<double>(), dQt.as<double>());
    const int npp = h.ref.nq * h.ref.nq;
    if (!(tm->par_blocks && prolong_t(st, (int)(N / npp), npp, tm->par_first.as<int64_t>(), tm->par_count.as<int>(),
                                      tm->p_ptr.as<int64_t>(), tm->p_col.as<int>(), tm->p_val.as<double>(), G,
                                      dQt.as<double>(), Nr, dQct.as<double>(), N)))
        csr_spmm(st, N, tm->pt_ptr.as<int64_t>(), tm->pt_col.as<int>(), tm->pt_val.as<double>(), G, dQt.as<double>(),
                 Nr, dQct.as<double>(), N, false);
    DBIE_HIP(hipGetLastError());
    if (!tm->blas) {
        blas_check(rocblas_create_handle(&tm->blas), "rocblas_create_handle");
        blas_check(rocblas_set_stream(tm->blas, st), "rocblas_set_stream");
    }
    const double one = 1.0, zero = 0.0;
    blas_check(rocblas_dgemm(tm->blas, rocblas_operation_transpose, rocblas_operation_none, G, K, (rocblas_int)N, &one,
                             dQct.as<double>(), N, dS, lds, &zero, dR.as<double>(), G),
               "rocblas_dgemm");
    DBIE_HIP(hipMemcpy2DAsync(R, sizeof(double) * ldr, dR.p, sizeof(double) * G, sizeof(double) * G, K,
                              hipMemcpyDeviceToHost, st));
    DBIE_HIP(hipStreamSynchronize(st));
}

}  // namespace

extern "C" {

dbie_status dbie_tmap_create(const dbie_geom *geom, int64_t M, const double *trg, double hcub_atol, double range_factor,
                             int n_threads, dbie_tmap **out) {
    return dbie_tmap_create_ex(geom, M, trg, hcub_atol, range_factor, n_threads, 0, out);
}

namespace {
// The near-target cubatures on the device (target_cubatures); pairs whose boxes outgrow the
// per-workgroup storage come back with status 2 and are integrated on the host.
void device_cubatures(const TargetCubatureJob &job, const Geometry &g, double *I, int *status) {
    ensure_device();
    const int npairs = (int)job.panel.size();
    const int nq = job.nq, npp = nq * nq;
    if (npp > kCubatureMaxNpp) fail(DBIE_BAD_ARG, "n_quad too large for the device cubature");
    hipStream_t st = nullptr;
    DBIE_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    struct StreamGuard {
        hipStream_t s;
        ~StreamGuard() { (void)hipStreamDestroy(s); }
    } sg{st};
    std::vector<double> rules(2 * nq);
    std::copy(g.glx, g.glx + nq, rules.begin());
    std::copy(g.bary, g.bary + nq, rules.begin() + nq);
    const std::array<double, 13> r = genz_malik_rule();
    CubatureRule rule;
    rule.l2 = r[0];
    rule.l3 = r[1];
    rule.l4 = r[2];
    rule.l5 = r[3];
    for (int i = 0; i < 5; ++i) rule.w[i] = r[4 + i];
    for (int i = 0; i < 4; ++i) rule.wp[i] = r[9 + i];
    constexpr int kBoxCap = 512, kMaxWg = 2048;
    const int nwg = std::min(npairs, kMaxWg);
    DBuf pd, ru, bx, io, so;
    upload(pd, job.pdat.data(), job.pdat.size(), st);
    upload(ru, rules.data(), rules.size(), st);
    bx.ensure(target_cubature_scratch(nwg, nq, kBoxCap));
    io.ensure(sizeof(double) * (size_t)npairs * npp);
    so.ensure(sizeof(int) * (size_t)npairs);
    target_cubatures(st, npairs, nq, pd.as<double>(), ru.as<double>(), rule, job.atol, (long long)job.maxevals, kBoxCap,
                     nwg, bx.as<double>(), io.as<double>(), so.as<int>());
    DBIE_HIP(hipGetLastError());
    DBIE_HIP(hipMemcpyAsync(I, io.p, sizeof(double) * (size_t)npairs * npp, hipMemcpyDeviceToHost, st));
    DBIE_HIP(hipMemcpyAsync(status, so.p, sizeof(int) * (size_t)npairs, hipMemcpyDeviceToHost, st));
    DBIE_HIP(hipStreamSynchronize(st));
}
}  // namespace

namespace {
// The refinement around the targets and the prolongation on the device (targets.hip): the leaves
// counted per coarse panel, placed in the host's order (panels from the last to the first, each
// subtree depth first), their nodes computed, the prolongation filled; the refined nodes and the
// prolongation stay in the map's device buffers.  Declines for n_quad > 5 (the reaction fold's
// block kernel takes at most 32 nodes per panel) and when a subtree is deeper than the device stack.
bool device_refine(dbie_tmap *tm, const RefineInput &in, RefineOutput &out) {
    using Clk = std::chrono::steady_clock;
    const auto t0 = Clk::now();
    const Geometry &g = *in.g;
    const int nq = g.nq, npp = nq * nq;
    if (npp > 32 || g.P > INT32_MAX) return false;
    ensure_device();
    if (!tm->stream) DBIE_HIP(hipStreamCreateWithFlags(&tm->stream, hipStreamNonBlocking));
    hipStream_t st = tm->stream;
    const int P = (int)g.P;
    DBuf dco(true), dtrg(true), dstart(true), didx(true), dcnt(true), doff(true), drc(true), dpar(true), dspl(true),
        drule(true), dgl(true);
    upload(dco, g.corners.data(), g.corners.size(), st);
    upload(dtrg, in.trg, 3 * (size_t)in.M, st);
    upload(dstart, in.start->data(), in.start->size(), st);
    upload(didx, in.idx->data(), in.idx->size(), st);
    RefineGrid G;
    G.t = dtrg.as<double>();
    for (int k = 0; k < 3; ++k) {
        G.lo[k] = in.lo[k];
        G.dim[k] = in.dim[k];
    }
    G.h = in.h;
    G.start = dstart.as<int64_t>();
    G.idx = didx.as<int>();
    dcnt.ensure(sizeof(int64_t) * std::max(P, 1));
    refine_count(st, P, dco.as<double>(), G, in.M, in.limit, in.range_factor, nq, dcnt.as<int64_t>());
    DBIE_HIP(hipGetLastError());
    std::vector<int64_t> cnt(P), off(P);
    if (P > 0) DBIE_HIP(hipMemcpyAsync(cnt.data(), dcnt.p, sizeof(int64_t) * P, hipMemcpyDeviceToHost, st));
    DBIE_HIP(hipStreamSynchronize(st));
    int64_t Pr = 0;
    for (int p = P - 1; p >= 0; --p) {  // the host pops the coarse panels from the last one
        if (cnt[p] < 0) return false;
        off[p] = Pr;
        Pr += cnt[p];
    }
    if (Pr * npp > INT32_MAX) return false;
    upload(doff, off.data(), off.size(), st);
    drc.ensure(sizeof(double) * 12 * std::max<int64_t>(Pr, 1));
    dpar.ensure(sizeof(int) * std::max<int64_t>(Pr, 1));
    dspl.ensure(sizeof(int) * std::max<int64_t>(Pr, 1));
    refine_emit(st, P, dco.as<double>(), G, in.M, in.limit, in.range_factor, nq, doff.as<int64_t>(), drc.as<double>(),
                dpar.as<int>(), dspl.as<int>());
    DBIE_HIP(hipGetLastError());
    // refined nodes with the rule make_from_panels gives a new interface
    double gx[16], gw[16];
    gauss_legendre(nq, gx, gw);
    std::vector<double> gl(gx, gx + nq);
    gl.insert(gl.end(), gw, gw + nq);
    upload(dgl, gl.data(), gl.size(), st);
    const int64_t Nr = Pr * npp;
    tm->x.ensure(sizeof(double) * std::max<int64_t>(Nr, 1));
    tm->y.ensure(sizeof(double) * std::max<int64_t>(Nr, 1));
    tm->z.ensure(sizeof(double) * std::max<int64_t>(Nr, 1));
    tm->w.ensure(sizeof(double) * std::max<int64_t>(Nr, 1));
    refined_nodes(st, Pr, nq, drc.as<double>(), dgl.as<double>(), dgl.as<double>() + nq, tm->x.as<double>(),
                  tm->y.as<double>(), tm->z.as<double>(), tm->w.as<double>());
    DBIE_HIP(hipGetLastError());
    out.parent.resize(Pr);
    out.from_split.resize(Pr);
    out.corners.resize((size_t)Pr * 12);
    out.x.resize(Nr);
    out.y.resize(Nr);
    out.z.resize(Nr);
    out.w.resize(Nr);
    if (Pr > 0) {
        DBIE_HIP(hipMemcpyAsync(out.parent.data(), dpar.p, sizeof(int) * Pr, hipMemcpyDeviceToHost, st));
        DBIE_HIP(hipMemcpyAsync(out.from_split.data(), dspl.p, sizeof(int) * Pr, hipMemcpyDeviceToHost, st));
        DBIE_HIP(hipMemcpyAsync(out.corners.data(), drc.p, sizeof(double) * 12 * Pr, hipMemcpyDeviceToHost, st));
        DBIE_HIP(hipMemcpyAsync(out.x.data(), tm->x.p, sizeof(double) * Nr, hipMemcpyDeviceToHost, st));
        DBIE_HIP(hipMemcpyAsync(out.y.data(), tm->y.p, sizeof(double) * Nr, hipMemcpyDeviceToHost, st));
        DBIE_HIP(hipMemcpyAsync(out.z.data(), tm->z.p, sizeof(double) * Nr, hipMemcpyDeviceToHost, st));
        DBIE_HIP(hipMemcpyAsync(out.w.data(), tm->w.p, sizeof(double) * Nr, hipMemcpyDeviceToHost, st));
    }
    DBIE_HIP(hipStreamSynchronize(st));
    const auto t1 = Clk::now();
    out.ms_refine = std::chrono::duration<double, std::milli>(t1 - t0).count();
    // prolongation: 1 entry per node of an unsplit panel, n_quad^2 per node of a split one
    out.p_ptr.assign(Nr + 1, 0);
    int64_t e = 0;
    for (int64_t q = 0; q < Pr; ++q) {
        const int64_t per = out.from_split[q] ? npp : 1;
        for (int k = 0; k < npp; ++k) {
            out.p_ptr[q * npp + k] = e;
            e += per;
        }
    }
    out.p_ptr[Nr] = e;
    upload(tm->p_ptr, out.p_ptr.data(), out.p_ptr.size(), st);
    tm->p_col.ensure(sizeof(int) * std::max<int64_t>(e, 1));
    tm->p_val.ensure(sizeof(double) * std::max<int64_t>(e, 1));
    std::vector<double> rules(g.glx, g.glx + nq);
    rules.insert(rules.end(), g.bary, g.bary + nq);
    upload(drule, rules.data(), rules.size(), st);
    prolongation(st, Pr, nq, dco.as<double>(), drule.as<double>(), dpar.as<int>(), dspl.as<int>(), tm->x.as<double>(),
                 tm->y.as<double>(), tm->z.as<double>(), tm->p_ptr.as<int64_t>(), tm->p_col.as<int>(),
                 tm->p_val.as<double>());
    DBIE_HIP(hipGetLastError());
    DBIE_HIP(hipStreamSynchronize(st));
    out.ms_prolong = std::chrono::duration<double, std::milli>(Clk::now() - t1).count();
    return true;
}
}  // namespace

dbie_status dbie_tmap_create_ex(const dbie_geom *geom, int64_t M, const double *trg, double hcub_atol,
                                double range_factor, int n_threads, int device_cubature, dbie_tmap **out) {
    return guard([&] {
        if (!geom || !out) fail(DBIE_BAD_ARG, "NULL geometry or output");
        if (device_cubature < 0 || device_cubature > 2) fail(DBIE_BAD_ARG, "device must be 0, 1 or 2");
        *out = nullptr;
        std::unique_ptr<dbie_tmap> tm(new dbie_tmap);
        TargetCubatureFn dev;
        if (device_cubature >= 1) dev = device_cubatures;
        TargetRefineFn ref;
        dbie_tmap *tp = tm.get();
        if (device_cubature == 2) ref = [tp](const RefineInput &in, RefineOutput &o) { return device_refine(tp, in, o); };
        tm->h.reset(build_target_map(geom->g, M, trg, hcub_atol, range_factor, n_threads, dev, ref));
        tm->N = geom->g.N;
        *out = tm.release();
    });
}

void dbie_tmap_destroy(dbie_tmap *tm) { delete tm; }

dbie_status dbie_tmap_sizes(const dbie_tmap *tm, int64_t *M, int64_t *N, int64_t *N_refined, int64_t *n_pairs,
                            int64_t *nnz_corr, int64_t *n_capped) {
    return guard([&] {
        check_tm(tm);
        const TargetMap &h = *tm->h;
        if (M) *M = h.M;
        if (N) *N = tm->N;
        if (N_refined) *N_refined = h.ref.N;
        if (n_pairs) *n_pairs = h.n_pairs;
        if (nnz_corr) *nnz_corr = (int64_t)h.c_val.size();
        if (n_capped) *n_capped = h.n_capped;
    });
}

dbie_status dbie_tmap_setup_ms(const dbie_tmap *tm, double *t) {
    return guard([&] {
        check_tm(tm);
        if (!t) fail(DBIE_BAD_ARG, "NULL output");
        for (int i = 0; i < 4; ++i) t[i] = tm->h->setup_ms[i];
    });
}

dbie_status dbie_tmap_refined(const dbie_tmap *tm, double *xyzw, int *parent, int *from_split) {
    return guard([&] {
        check_tm(tm);
        const TargetMap &h = *tm->h;
        const Geometry &R = h.ref;
        if (xyzw) {
            std::memcpy(xyzw, R.x.data(), sizeof(double) * R.N);
            std::memcpy(xyzw + R.N, R.y.data(), sizeof(double) * R.N);
            std::memcpy(xyzw + 2 * R.N, R.z.data(), sizeof(double) * R.N);
            std::memcpy(xyzw + 3 * R.N, R.w.data(), sizeof(double) * R.N);
        }
        if (parent) std::copy(h.parent.begin(), h.parent.end(), parent);
        if (from_split) std::copy(h.from_split.begin(), h.from_split.end(), from_split);
    });
}

dbie_status dbie_tmap_matrices(const dbie_tmap *tm, int64_t *p_ptr, int *p_col, double *p_val, int64_t *c_ptr,
                               int *c_col, double *c_val) {
    return guard([&] {
        check_tm(tm);
        const TargetMap &h = *tm->h;
        if (p_ptr) std::copy(h.p_ptr.begin(), h.p_ptr.end(), p_ptr);
        if (h.prolong_device) {  // the entries live on the device
            const int64_t nnz = h.p_ptr.empty() ? 0 : h.p_ptr.back();
            if (p_col && nnz) DBIE_HIP(hipMemcpy(p_col, tm->p_col.p, sizeof(int) * nnz, hipMemcpyDeviceToHost));
            if (p_val && nnz) DBIE_HIP(hipMemcpy(p_val, tm->p_val.p, sizeof(double) * nnz, hipMemcpyDeviceToHost));
        } else {
            if (p_col) std::copy(h.p_col.begin(), h.p_col.end(), p_col);
            if (p_val) std::copy(h.p_val.begin(), h.p_val.end(), p_val);
        }
        if (c_ptr) std::copy(h.c_ptr.begin(), h.c_ptr.end(), c_ptr);
        if (c_col) std::copy(h.c_col.begin(), h.c_col.end(), c_col);
        if (c_val) std::copy(h.c_val.begin(), h.c_val.end(), c_val);
    });
}

dbie_status dbie_tmap_apply(dbie_tmap *tm, int K, const double *Sigma, int64_t lds, double *U, int64_t ldu) {
    return guard([&] {
        check_tm(tm);
        const int64_t N = tm->N, M = tm->h->M;
        if (K < 1 || !Sigma || (!U && M > 0)) fail(DBIE_BAD_ARG, "K must be >= 1 and buffers non-NULL");
        if (lds < N || ldu < M) fail(DBIE_DIM_MISMATCH, "leading dimension too small");
        make_resident(tm);
        hipStream_t st = tm->stream;
        tm->xin.ensure(sizeof(double) * (size_t)N * K);
        tm->uout.ensure(sizeof(double) * (size_t)std::max<int64_t>(M, 1) * K);
        DBIE_HIP(hipMemcpy2DAsync(tm->xin.p, sizeof(double) * N, Sigma, sizeof(double) * lds, sizeof(double) * N, K,
                                  hipMemcpyHostToDevice, st));
        tmap_apply_dev(tm, K, tm->xin.as<double>(), N, tm->uout.as<double>(), M);
        if (M > 0)
            DBIE_HIP(hipMemcpy2DAsync(U, sizeof(double) * ldu, tm->uout.p, sizeof(double) * M, sizeof(double) * M, K,
                                      hipMemcpyDeviceToHost, st));
        DBIE_HIP(hipStreamSynchronize(st));
    });
}

dbie_status dbie_tmap_apply_device(dbie_tmap *tm, int K, const double *dSigma, int64_t lds, double *dU, int64_t ldu) {
    return guard([&] {
        check_tm(tm);
        if (K < 1 || !dSigma || !dU) fail(DBIE_BAD_ARG, "K must be >= 1 and buffers non-NULL");
        if (lds < tm->N || ldu < tm->h->M) fail(DBIE_DIM_MISMATCH, "leading dimension too small");
        tmap_apply_dev(tm, K, dSigma, lds, dU, ldu);
        DBIE_HIP(hipStreamSynchronize(tm->stream));
    });
}

dbie_status dbie_reaction_matrix(dbie_tmap *tm, int G, const int64_t *grp_ptr, const double *q, int K,
                                 const double *Sigma, int64_t lds, double *R, int64_t ldr) {
    return guard([&] {
        check_tm(tm);
        const int64_t N = tm->N;
        if (K < 1 || !Sigma) fail(DBIE_BAD_ARG, "K must be >= 1 and Sigma non-NULL");
        if (lds < N) fail(DBIE_DIM_MISMATCH, "lds smaller than N");
        make_resident(tm);
        hipStream_t st = tm->stream;
        tm->xin.ensure(sizeof(double) * (size_t)N * K);
        DBIE_HIP(hipMemcpy2DAsync(tm->xin.p, sizeof(double) * N, Sigma, sizeof(double) * lds, sizeof(double) * N, K,
                                  hipMemcpyHostToDevice, st));
        reaction_dev(tm, G, grp_ptr, q, K, tm->xin.as<double>(), N, R, ldr);
    });
}

dbie_status dbie_reaction_matrix_device(dbie_tmap *tm, int G, const int64_t *grp_ptr, const double *q, int K,
                                        const double *dSigma, int64_t lds, double *R, int64_t ldr) {
    return guard([&] {
        check_tm(tm);
        if (K < 1 || !dSigma) fail(DBIE_BAD_ARG, "K must be >= 1 and Sigma non-NULL");
        if (lds < tm->N) fail(DBIE_DIM_MISMATCH, "lds smaller than N");
        reaction_dev(tm, G, grp_ptr, q, K, dSigma, lds, R, ldr);
    });
}

dbie_status dbie_gauss_hermite(int n, double *x, double *w) {
    return guard([&] {
        if (!x || !w) fail(DBIE_BAD_ARG, "NULL output");
        gauss_hermite(n, x, w);
    });
}

}  // extern "C"
