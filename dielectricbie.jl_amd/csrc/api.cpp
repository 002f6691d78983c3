// api.cpp -- the C ABI of libdbie: handles, device residency, operator apply and
// the GMRES / mixed-precision refinement drivers.
//
// Host code owns control flow only; every O(N) or larger step runs on the GPU.
// Per Krylov iteration the host reads back the (j+2) x K Hessenberg column to
// update the per-column Givens rotations (a few KB), as Krylov.jl does on the CPU.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <atomic>
#include <tuple>
#include <type_traits>
#include <map>
#include <mutex>

#include "internal.hpp"
#include "kernels.hpp"
#include "runtime.hpp"

namespace dbie {

static thread_local std::string g_last_error;
void set_error(const std::string &msg) { g_last_error = msg; }
[[noreturn]] void fail(dbie_status code, const std::string &msg) { throw Error{code, msg}; }
void hip_check(hipError_t e, const char *what) {
    if (e != hipSuccess) fail(DBIE_HIP_ERROR, std::string(what) + ": " + hipGetErrorString(e));
}

namespace {
// size class of a pooled block: 256-byte granules below 1 MiB, then eight classes per octave
size_t pool_class(size_t b) {
    if (b <= ((size_t)1 << 20)) return (b + 255) & ~(size_t)255;
    int e = 63 - __builtin_clzll((unsigned long long)b);  // 2^e <= b < 2^(e+1)
    const size_t step = (size_t)1 << (e - 3);
    return (b + step - 1) / step * step;
}
constexpr size_t kPoolMaxBlock = (size_t)8 << 30;
constexpr size_t kPoolMaxCached = (size_t)16 << 30;
struct ThreadPool {
    std::multimap<size_t, void *> free_blocks;
    size_t cached = 0;
    void trim() {
        for (auto &kv : free_blocks) (void)hipFree(kv.second);
        free_blocks.clear();
        cached = 0;
    }
    ~ThreadPool() { trim(); }
};
ThreadPool &thread_pool() {
    static thread_local ThreadPool tp;
    return tp;
}
}  // namespace

void *pool_take(size_t bytes, size_t *got) {
    const size_t c = pool_class(std::max<size_t>(bytes, 1));
    ThreadPool &tp = thread_pool();
    auto it = tp.free_blocks.find(c);
    if (it != tp.free_blocks.end()) {
        void *p = it->second;
        tp.free_blocks.erase(it);
        tp.cached -= c;
        *got = c;
        return p;
    }
    void *p = nullptr;
    if (hipMalloc(&p, c) != hipSuccess) {
        (void)hipGetLastError();
        tp.trim();  // give the cached blocks back and try once more
        if (hipMalloc(&p, c) != hipSuccess) {
            (void)hipGetLastError();
            return nullptr;
        }
    }
    *got = c;
    return p;
}

void pool_give(void *p, size_t bytes) {
    ThreadPool &tp = thread_pool();
    if (bytes >= kPoolMaxBlock || tp.cached + bytes > kPoolMaxCached) {
        (void)hipFree(p);
        return;
    }
    tp.free_blocks.emplace(bytes, p);
    tp.cached += bytes;
}

namespace {
std::once_flag g_init_flag;
dbie_status g_init_status = DBIE_OK;
std::string g_init_msg;

}  // namespace

void ensure_device() {
    std::call_once(g_init_flag, [] {
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
            g_init_status = DBIE_NO_DEVICE;
            g_init_msg = "no HIP device visible";
            return;
        }
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, 0) != hipSuccess) {
            g_init_status = DBIE_NO_DEVICE;
            g_init_msg = "hipGetDeviceProperties failed";
            return;
        }
        if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
            g_init_status = DBIE_NO_DEVICE;
            g_init_msg = std::string("libdbie is built for gfx950 only; device is ") + prop.gcnArchName;
        }
    });
    if (g_init_status != DBIE_OK) fail(g_init_status, g_init_msg);
}

}  // namespace dbie

using namespace dbie;

// The Krylov space of one low-precision inner solve, kept for the refinements after it: the
// basis V (nblk N x K blocks, in GmresWork::V1) and per column the Arnoldi steps k1, the Givens
// rotations cs/sn that reduced the (k1+1) x k1 Hessenberg H to R1 (stored in R, layout of the
// solve's Hessenberg), and omega = Omega^T e_k1 with Omega = G_(k1-1) ... G_0, the unit vector
// orthogonal to range(H).  (A M) V_k1 = V_(k1+1) H, so C = V_(k1+1) Omega^T[:, :k1] is an orthonormal
// basis of the image of U = V_k1 R1^-1 (A M U = C), the recycled pair of GCRO.
struct KrylovRecycle {
    bool valid = false;
    int nblk = 0, m = 0;
    int elem = 8;  // bytes per element of the stored basis (V1): 8 (FP64) or 4 (FP32)
    std::vector<int> k1;
    std::vector<double> cs, sn, R, omega;
};

struct GmresWork {
    DBuf V, W, Z, U, partial, hbuf, nbuf;
    DBuf V1, abuf, cbuf;  // recycled Krylov basis of an earlier inner solve, its projection coefficients
    DBuf om, n1;          // the recycled space's unit normals omega and block counts on the device
    PinnedBuf pin;        // host staging of the per-iteration coefficients, norms and projections
    KrylovRecycle rec;
    DBuf H, G, tau, Y;  // block GMRES: block Hessenberg, projected right-hand side, reflectors, coefficients
    DBuf gpart, R, info;  // block GMRES: split-K partial products, K x K triangle, rocSOLVER info
    DBuf PV, PT;          // block GMRES: Householder vectors (2K x K) and T factors (K x K) per Hessenberg step
    DBuf PY, PX, PZ;      // near-field preconditioner: its near product (N x K FP64), staging of its input (FP32
                          // copy, or the split-FP16 operands), the first sweep's result (two sweeps)
    // take the workspaces from (and return them to) the calling thread's block cache
    void set_pooled() {
        for (DBuf *d : {&V, &W, &Z, &U, &partial, &hbuf, &nbuf, &V1, &abuf, &cbuf, &om, &n1, &H, &G, &tau, &Y, &gpart,
                        &R, &info, &PV, &PT, &PY, &PX, &PZ})
            d->pooled = true;
    }
};

struct dbie_op {
    Geometry g;                  // host copy of the interface
    std::vector<double> diag_h;  // t per node
    std::vector<double> eps_in, eps_out;
    hipStream_t stream = nullptr;
    // resident device data
    DBuf x, y, z, nx, ny, nz, w, xf, yf, zf, nxf, nyf, nzf, wf, plane, diag;
    DBuf bsr_row, bsr_col, bsr_v64, bsr_v32, bsr_vup32, nmask;
    DBuf bsr_vh, bsr_rexp;  // split-FP16 full near blocks (2^rexp C / w, hi and lo halves) and row exponents
    DBuf sr_ptr, sr_col, sr_blk;  // super-rows of the split-FP16 near kernel
    dbie_near_opts nopts{};       // near options the operator was built with
    DBuf bsr_d64;                 // correction blocks of the double layer D (built on first use)
    bool d_built = false;
    DBuf bsr_full64;              // the near pairs' full interaction in FP64 (the near-field preconditioner
    bool full64_built = false;    // of FP64 solves; built on first use)
    // residual history of the last solve: (kind, max_k ||r_k|| / ||F_k||) per entry; kind 0 = a Krylov
    // iteration's residual estimate, 1 = the true FP64 residual, 2 = a residual updated with the
    // low-precision operator or taken from the inner solve's estimate.  hist_scale: 1 / ||F_k||
    std::vector<double> hist, hist_scale;
    void note(int kind, const std::vector<double> &r) {
        if (hist_scale.size() != r.size()) return;
        double m = 0.0;
        for (size_t k = 0; k < r.size(); ++k) m = std::max(m, r[k] * hist_scale[k]);
        hist.push_back((double)kind);
        hist.push_back(m);
    }
    DevGeom dg;
    DevBSR bsr;
    int adaptive_hits = 0;
    // workspaces (grown on demand)
    DBuf part, xin, yout, tmp, mscratch, gmax, xt32;
    GmresWork gw;                 // GMRES workspaces, reused across solves
    double setup_ms[4] = {0, 0, 0, 0};  // near pairs, correction blocks (host part), derived data, uploads + device block build
    float f16_gscale = 0.0f;  // power-of-two scale of the fp16 kernel entries (0 = not yet computed)
    // the split-FP16 far field's kernel matrix kept in HBM (launch variants 2 - 4, K > 128): built on
    // the first apply whose launch shape uses it.  stored_mode 1: when device memory allows (default;
    // DBIE_STORED_FAR=0 starts operators with 0), 0: never.  gst_state 0: not built, 1: built, -1:
    // declined for this operator (not enough free device memory)
    DBuf gst, gst_off;
    int gst_ntile = 0;
    int stored_mode = 1;
    int gst_state = 0;
    bool gst_excl = false;        // built with the near panel pairs left out
    std::vector<int> tile_plane_h;  // plane ids of the 256-node tiles (host copy)
    // optional per-kernel timing: event pairs recorded around launches
    bool profiling = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> prof[3];
    // operator applies captured as HIP graphs: an apply whose arguments (buffers, K, precision,
    // launch shape) repeat is captured on its second call and replayed from then on
    struct ApplyGraph {
        const double *X;
        int64_t ldx;
        double *Y;
        int64_t ldy;
        int K, prec, splits, variant;
        hipGraphExec_t exec;    // null: seen once, not captured yet
        void *part, *scratch, *xt;  // workspaces the graph writes; a reallocation retires it
        void *gst;              // the stored kernel matrix the graph reads (or null); a change retires it
        uint64_t last_use;
    };
    rocblas_handle blas = nullptr;  // on the operator's stream (block GMRES), created on first use
    bool graphs = true;
    bool graphs_failed = false;  // a capture failed: applies stay eager
    std::vector<ApplyGraph> agraphs;
    uint64_t graph_tick = 0;
    int64_t graph_captures = 0, graph_replays = 0;
    void drop_graphs() {
        for (auto &a : agraphs)
            if (a.exec) (void)hipGraphExecDestroy(a.exec);
        agraphs.clear();
    }
    // A fused operator (dbie_solve_fused): no geometry of its own; its N x K blocks hold the columns of
    // several member operators (member b: columns c0 .. c0 + K - 1, rows 0 .. its N - 1, the rows past
    // that zero), and an apply runs every member's apply on the member's stream, forked from and joined
    // back to this operator's stream.
    struct Member {
        dbie_op *op;
        int c0, K;
        hipEvent_t done;
    };
    std::vector<Member> members;
    hipEvent_t fork_ev = nullptr;
    ~dbie_op() {
        if (stream) (void)hipStreamSynchronize(stream);
        for (auto &m : members)
            if (m.done) (void)hipEventDestroy(m.done);
        if (fork_ev) (void)hipEventDestroy(fork_ev);
        drop_graphs();
        if (blas) (void)rocblas_destroy_handle(blas);
        for (auto &v : prof)
            for (auto &e : v) {
                (void)hipEventDestroy(e.first);
                (void)hipEventDestroy(e.second);
            }
        if (stream) (void)hipStreamDestroy(stream);
    }
    int64_t N() const { return g.N; }
};

namespace dbie {
namespace {

// quadtree nodes one device walk of a touching pair may visit before that pair goes to the host
constexpr int kQuadtreeNodeBudget = 1 << 16;

// Near-pair list with the candidate scan on the device (pair_prep and finish_pair_list on the host).
PairList device_pair_list(const Geometry &g, const DevGeom &dg, int max_order, double atol, bool correct_edges,
                          hipStream_t st) {
    const int64_t P = g.P;
    if (P > INT32_MAX) fail(DBIE_BAD_ARG, "too many panels");
    const PairPrep pp = pair_prep(g, max_order, atol, correct_edges);
    DBuf pd(true), co(true), ie(true), cnt(true), dptr(true), dtrg(true), drho(true), dtouch(true);
    upload(pd, pp.pdata.data(), pp.pdata.size(), st);
    upload(co, g.corners.data(), g.corners.size(), st);
    upload(ie, g.is_edge.data(), g.is_edge.size(), st);
    cnt.ensure(sizeof(int64_t) * P);
    pair_scan(st, dg, (int)P, g.npp(), pd.as<double>(), co.as<double>(), ie.as<int>(), pp.tc2, correct_edges,
              cnt.as<int64_t>(), nullptr, nullptr, nullptr, nullptr);
    DBIE_HIP(hipGetLastError());
    PairCandidates c;
    c.ptr.assign(P + 1, 0);
    DBIE_HIP(hipMemcpyAsync(c.ptr.data() + 1, cnt.p, sizeof(int64_t) * P, hipMemcpyDeviceToHost, st));
    DBIE_HIP(hipStreamSynchronize(st));
    for (int64_t i = 0; i < P; ++i) c.ptr[i + 1] += c.ptr[i];
    const int64_t total = c.ptr[P];
    upload(dptr, c.ptr.data(), c.ptr.size(), st);
    dtrg.ensure(sizeof(int) * std::max<int64_t>(total, 1));
    drho.ensure(sizeof(double) * std::max<int64_t>(total, 1));
    dtouch.ensure(std::max<int64_t>(total, 1));
    pair_scan(st, dg, (int)P, g.npp(), pd.as<double>(), co.as<double>(), ie.as<int>(), pp.tc2, correct_edges,
              nullptr, dptr.as<int64_t>(), dtrg.as<int>(), drho.as<double>(), dtouch.as<unsigned char>());
    DBIE_HIP(hipGetLastError());
    c.trg.resize(total);
    c.rho.resize(total);
    c.touch.resize(total);
    if (total) {
        DBIE_HIP(hipMemcpyAsync(c.trg.data(), dtrg.p, sizeof(int) * total, hipMemcpyDeviceToHost, st));
        DBIE_HIP(hipMemcpyAsync(c.rho.data(), drho.p, sizeof(double) * total, hipMemcpyDeviceToHost, st));
        DBIE_HIP(hipMemcpyAsync(c.touch.data(), dtouch.p, total, hipMemcpyDeviceToHost, st));
    }
    DBIE_HIP(hipStreamSynchronize(st));
    return finish_pair_list(g, pp, c);
}

void check_op(const dbie_op *op) {
    if (!op) fail(DBIE_BAD_ARG, "operator handle is NULL");
}

// Plane id per 256-node tile: nodes with the same panel normal and the same n.x (computed as the
// kernels' entries are) share a plane id; a tile gets that id when all its nodes share it, else -1.
// For axis-aligned normals two nodes of one plane id have n.(x - y) = 0 exactly, so the far-field
// kernels skip same-plane tile pairs at no cost to accuracy.  The key is taken per node, not per panel,
// so that nodes a caller supplies off their panel's plane never have their entries skipped.
std::vector<int> tile_planes(const Geometry &g) {
    const int npp = g.npp();
    std::map<std::array<double, 4>, int> ids;
    std::vector<int> nid(g.N);
    for (int64_t i = 0; i < g.N; ++i) {
        const double *n = &g.normal[3 * (i / npp)];
        const std::array<double, 4> key{n[0], n[1], n[2], n[0] * g.x[i] + n[1] * g.y[i] + n[2] * g.z[i]};
        auto it = ids.find(key);
        if (it == ids.end()) {
            const int id = ids.size() < 4096 ? (int)ids.size() : -1;
            it = ids.emplace(key, id).first;
        }
        nid[i] = it->second;
    }
    const int64_t nt = (g.N + 255) / 256;
    std::vector<int> tp(nt);
    for (int64_t t = 0; t < nt; ++t) {
        const int64_t n0 = t * 256, n1 = std::min<int64_t>(g.N, n0 + 256);
        int id = nid[n0];
        for (int64_t n = n0; n < n1 && id >= 0; ++n)
            if (nid[n] != id) id = -1;
        tp[t] = id;
    }
    return tp;
}

void compute_diag(dbie_op *op, const double *eps_in, const double *eps_out) {
    const Geometry &g = op->g;
    const int npp = g.npp();
    op->eps_in.assign(eps_in, eps_in + g.P);
    op->eps_out.assign(eps_out, eps_out + g.P);
    op->diag_h.resize(g.N);
    for (int64_t p = 0; p < g.P; ++p) {
        const double t = 0.5 * (eps_out[p] + eps_in[p]) / (eps_out[p] - eps_in[p]);
        for (int j = 0; j < npp; ++j) op->diag_h[p * npp + j] = t;
    }
    upload(op->diag, op->diag_h.data(), g.N, op->stream);
}

// Correction blocks of one kernel mode (0: laplace3d_DT_corrections, 1: laplace3d_D_corrections,
// src/kernel/laplace3d_near_upsampling.jl:127-282) for the operator's geometry with the near options o:
// the near-pair list (build_neighbor_list, src/kernel/laplace3d_near.jl:240-392; device candidate scan
// when o.device_build), then the blocks.  On the device path the values land in v64 (and v32, vup32
// when given) in the near kernels' storage order and true is returned; otherwise bs.val holds them
// (row-major blocks) and false is returned.  bs gets the BSR structure either way.  t: ms of the pair
// list, the host part of the block build, and the device block build.
bool build_correction_blocks(dbie_op *op, const dbie_near_opts &o, int mode, BlockSparse &bs, DBuf *v64, DBuf *v32,
                             DBuf *vup32, double *t) {
    using Clk = std::chrono::steady_clock;
    auto ms = [](Clk::time_point a, Clk::time_point b) {
        return std::chrono::duration<double, std::milli>(b - a).count();
    };
    const Geometry &g = op->g;
    const int64_t N = g.N;
    hipStream_t st = op->stream;
    const int bd = g.npp();
    const bool mfma_order = bd == 16;
    bs.bd = bd;
    NearOpts no;
    no.up_tol = o.up_tol;
    no.max_order = o.max_order;
    no.correct_edges = o.correct_edges != 0;
    no.adaptive_atol = o.adaptive_atol;
    no.adaptive_rtol = o.adaptive_rtol;
    no.adaptive_n_GL = o.adaptive_n_GL;
    no.adaptive_max_depth = o.adaptive_max_depth;
    no.n_threads = o.n_threads;
    no.mode = mode;
    // the device builder keeps a block's kernel values in LDS
    const bool on_device =
        o.device_build && (double)bd * o.max_order * o.max_order * sizeof(double) <= (double)kNearBuildLds;
    const auto t0 = Clk::now();
    DevGeom dgp;
    dgp.N = N;
    dgp.x = op->x.as<double>(); dgp.y = op->y.as<double>(); dgp.z = op->z.as<double>();
    const PairList pl = o.device_build ? device_pair_list(g, dgp, o.max_order, o.up_tol, o.correct_edges != 0, st)
                                       : build_pair_list(g, o.max_order, o.up_tol, o.correct_edges != 0, o.n_threads);
    const auto t1 = Clk::now();
    t[0] = ms(t0, t1);
    if (!on_device) {
        bs = build_corrections(g, pl, no);
        t[1] = ms(t1, Clk::now());
        t[2] = 0.0;
        return false;
    }
    // host: pair order, source caches, touching-pair quadtrees; device: every block
    const int nGL = o.adaptive_n_GL == 0 ? g.nq : o.adaptive_n_GL;
    const bool dev_quad = adaptive_blocks_lds(g.nq, nGL, o.adaptive_max_depth) <= kNearBuildLds &&
                          o.adaptive_max_depth >= 0;
    const CorrectionPlan cp = plan_corrections(g, pl, no, !dev_quad);
    const auto t2 = Clk::now();
    t[1] = ms(t1, t2);
    bs.row_ptr = cp.row_ptr;
    bs.col = cp.col;
    bs.adaptive_hits = cp.adaptive_hits;
    const size_t nv = (size_t)cp.nblocks() * bd * bd;
    v64->ensure(sizeof(double) * std::max<size_t>(nv, 1));
    if (v32) v32->ensure(sizeof(float) * std::max<size_t>(nv, 1));
    if (vup32) vup32->ensure(sizeof(float) * std::max<size_t>(nv, 1));
    DBuf us(true), ut(true), ul(true), co(true), cn(true), ca(true), as(true), at(true), al(true), ak(true);
    upload(us, cp.up_src.data(), cp.up_src.size(), st);
    upload(ut, cp.up_trg.data(), cp.up_trg.size(), st);
    upload(ul, cp.up_slot.data(), cp.up_slot.size(), st);
    upload(co, cp.cache_off.data(), cp.cache_off.size(), st);
    upload(cn, cp.cache_nu.data(), cp.cache_nu.size(), st);
    upload(ca, cp.caches.data(), cp.caches.size(), st);
    upload(as, cp.adp_src.data(), cp.adp_src.size(), st);
    upload(at, cp.adp_trg.data(), cp.adp_trg.size(), st);
    upload(al, cp.adp_slot.data(), cp.adp_slot.size(), st);
    int max_nu = 0;
    for (int v : cp.cache_nu) max_nu = std::max(max_nu, v);
    DevGeom dgb;
    dgb.N = N;
    dgb.x = op->x.as<double>(); dgb.y = op->y.as<double>(); dgb.z = op->z.as<double>();
    dgb.nx = op->nx.as<double>(); dgb.ny = op->ny.as<double>(); dgb.nz = op->nz.as<double>();
    dgb.w = op->w.as<double>();
    const int n_adp = (int)cp.adp_src.size();
    DBuf ar(true), am(true), ru(true), ah(true);
    std::vector<int> hits;
    if (dev_quad) {
        // touching pairs: quadtrees on the device, K_up blocks into ak
        const std::vector<double> rules = quadtree_rules(g, nGL);
        upload(ar, cp.adp_rep.data(), cp.adp_rep.size(), st);
        upload(am, cp.adp_map.data(), cp.adp_map.size(), st);
        upload(ru, rules.data(), rules.size(), st);
        ak.ensure(sizeof(double) * std::max<size_t>((size_t)n_adp * bd * bd, 1));
        ah.ensure(sizeof(int) * std::max(n_adp, 1));
        if (!adaptive_blocks(st, dgb, g.nq, nGL, n_adp, as.as<int>(), at.as<int>(), ar.as<int>(), am.as<double>(),
                             ru.as<double>(), o.adaptive_atol, o.adaptive_rtol, o.adaptive_max_depth,
                             kQuadtreeNodeBudget, ak.as<double>(), ah.as<int>(), mode))
            fail(DBIE_BAD_ARG, "near block build: quadtree too deep for LDS");
        DBIE_HIP(hipGetLastError());
        hits.resize(n_adp);
        if (n_adp > 0) {
            DBIE_HIP(hipMemcpyAsync(hits.data(), ah.p, sizeof(int) * n_adp, hipMemcpyDeviceToHost, st));
            DBIE_HIP(hipStreamSynchronize(st));
        }
        bs.adaptive_hits = 0;
        std::vector<double> blk((size_t)bd * bd);
        for (int k = 0; k < n_adp; ++k) {
            if (hits[k] == 2) {  // walk over the node budget: this block on the host
                hits[k] = touching_block(g, cp.adp_src[k], cp.adp_trg[k], no, blk.data()) ? 1 : 0;
                h2d(ak.as<double>() + (size_t)k * bd * bd, blk.data(), sizeof(double) * blk.size(), st);
            }
            bs.adaptive_hits += hits[k];
        }
    } else {
        upload(ak, cp.adp_kup.data(), cp.adp_kup.size(), st);
    }
    const bool ok = build_near_blocks(st, dgb, bd, mfma_order, (int)cp.up_src.size(), us.as<int>(), ut.as<int>(),
                                      ul.as<int>(), co.as<int>(), cn.as<int>(), ca.as<double>(), max_nu, n_adp,
                                      as.as<int>(), at.as<int>(), al.as<int>(), ak.as<double>(), v64->as<double>(),
                                      v32 ? v32->as<float>() : nullptr, vup32 ? vup32->as<float>() : nullptr, mode);
    if (!ok) fail(DBIE_BAD_ARG, "near block build: upsampling order too high for LDS");
    DBIE_HIP(hipGetLastError());
    DBIE_HIP(hipStreamSynchronize(st));
    t[2] = ms(t2, Clk::now());
    return true;
}

// The near corrections of the double layer D (laplace3d_D_corrections,
// src/kernel/laplace3d_near_upsampling.jl:278-282) on the operator's near-pair list, built on first use
// into bsr_d64 in the storage order of the FP64 near kernels.  The pair list depends on the geometry
// alone, so the blocks sit in the slots of the Dt corrections.
void ensure_d_corrections(dbie_op *op) {
    if (op->d_built) return;
    if (!op->nopts.build_corrections) fail(DBIE_BAD_ARG, "the operator was created without near corrections");
    BlockSparse bs;
    double t[3];
    const bool dev = build_correction_blocks(op, op->nopts, 1, bs, &op->bsr_d64, nullptr, nullptr, t);
    if (bs.nblocks() != op->bsr.nb) fail(DBIE_HIP_ERROR, "D corrections: pair list differs from the operator's");
    if (!dev) {
        const int bd = bs.bd;
        if (bd == 16) {
            parallel_for(bs.nblocks(), op->nopts.n_threads, [&](int64_t b, int) {
                double tmpb[256];
                double *blk = &bs.val[(size_t)b * 256];
                for (int tt = 0; tt < 16; ++tt)
                    for (int c = 0; c < 16; ++c) tmpb[bsr_storage_index(16, true, tt, c)] = blk[tt * 16 + c];
                std::copy(tmpb, tmpb + 256, blk);
            });
        }
        upload(op->bsr_d64, bs.val.data(), std::max<size_t>(bs.val.size(), 1), op->stream);
    }
    op->d_built = true;
}

// Event pair around a launch when profiling is on (kernel id: 0 far FP64, 1 far low precision, 2 combine).
struct ProfScope {
    dbie_op *op;
    int id;
    hipEvent_t e1 = nullptr;
    ProfScope(dbie_op *o, int k) : op(o), id(k) {
        if (!op->profiling) return;
        hipEvent_t e0;
        DBIE_HIP(hipEventCreate(&e0));
        DBIE_HIP(hipEventCreate(&e1));
        DBIE_HIP(hipEventRecord(e0, op->stream));
        op->prof[id].push_back({e0, e1});
    }
    ~ProfScope() {
        if (e1) (void)hipEventRecord(e1, op->stream);
    }
};

void check_precision(int p) {
    if (p < DBIE_FP64 || p > DBIE_FP16X3) fail(DBIE_BAD_ARG, "unknown precision");
}

// Y = A X on the device (X, Y: N x K with leading dims).
// ---- launch-shape autotuner ------------------------------------------------------------------
// The far-field contraction's source split count (and, for the split-FP16 form at K > 128, the
// kernel variant) is picked per shape class: (precision, near exclusion, nodes per panel, K in
// 32-column groups, N to a quarter octave).  The tuner is online: the first applies of a class
// each run the next untried candidate with HIP events around the whole apply (far field and
// combine), nothing is run twice and nothing waits for a timing; timings are collected when their
// events have completed, and once every candidate has one the fastest is kept for the process.
// Different split counts sum the partial sums in a different order, so results differ at the
// rounding level between candidates (well inside every solve tolerance).
struct TuneKey {
    int prec, excl, npp, kgroups, nbucket;
    bool operator<(const TuneKey &o) const {
        return std::tie(prec, excl, npp, kgroups, nbucket) <
               std::tie(o.prec, o.excl, o.npp, o.kgroups, o.nbucket);
    }
};
// Every candidate is timed kTuneRounds times (round robin) and keeps its fastest time: a single
// timing of a launch can be off by several per cent (clock ramps, the first touch of a buffer).
constexpr size_t kTuneRounds = 2;
struct TuneState {
    std::vector<FarTune> cand;
    std::vector<float> ms;  // per candidate: fastest time so far; < 0 until timed
    std::vector<int> nt;    // per candidate: timings received
    size_t next = 0;        // next trial (candidate next % cand.size())
    bool done = false;
    FarTune best;
};
struct PendingTiming {
    TuneKey key;
    size_t idx;
    hipEvent_t e0, e1;
};
std::mutex g_tune_mu;
std::map<TuneKey, TuneState> g_tune;
std::vector<PendingTiming> g_pending;
std::atomic<int> g_autotune{-1};  // -1: not read yet (DBIE_AUTOTUNE=0 switches it off)

bool autotune_on() {
    int v = g_autotune.load();
    if (v < 0) {
        const char *e = std::getenv("DBIE_AUTOTUNE");
        v = (e && e[0] == '0') ? 0 : 1;
        g_autotune.store(v);
    }
    return v != 0;
}

// whether the far field of this apply leaves the near panel pairs out (split-FP16 MFMA form)
// whether a split-FP16 apply leaves the near panel pairs out of its far field: the MFMA kernels, and for
// K < 16 the FP32 far field on 16-node panels (without the exclusion the near pairs' large direct
// entries would meet their corrections in low precision: at N = 62,208 and K = 4, 28 inner iterations
// and two refinements instead of 15 and one)
bool excludes_near(const dbie_op *op, int K, dbie_precision prec) {
    return prec == DBIE_FP16X3 && op->dg.near_mask && op->bsr.valup32 &&
           (far_field_mfma_supported((int)prec, K) || (K < 16 && op->dg.npp == 16));
}

// whether an apply that leaves the near panel pairs out of its far field takes their full interaction
// from the split-FP16 blocks (no FP32 copy of X needed)
bool near_uses_h3(const dbie_op *op, int K) { return op->bsr.valh && op->bsr.mfma_order && op->bsr.bd == 16 && K >= 16; }

// split-FP16 applies at K < 16 (FP32 far field without the near pairs): the near pairs' full blocks
// multiply split-FP16 operands made for them (padded to 16 columns) where the apply has the scratch
bool small_k_h3(const dbie_op *op, int K, dbie_precision prec) {
    return K < 16 && excludes_near(op, K, prec) && op->bsr.valh && op->bsr.mfma_order && op->bsr.bd == 16;
}

TuneKey tune_key(const dbie_op *op, int K, dbie_precision prec, bool exclude) {
    return TuneKey{(int)prec, exclude ? 1 : 0, op->g.npp(), (K + 31) / 32,
                   (int)std::lround(4.0 * std::log2((double)std::max<int64_t>(op->N(), 1)))};
}

// Candidate launch shapes for one apply: split counts the launch accepts whose partial sums fit
// in 2 GB, each with the default kernel and (split-FP16, K > 128) the general kernel, the
// warp-specialised kernel with one consumer wave per SIMD and, at one or two splits, the GEMM over
// the stored kernel matrix in its three forms (an operator that keeps no
// stored matrix runs those as the warp-specialised kernel).
std::vector<FarTune> tune_candidates(int64_t N, int K, dbie_precision prec) {
    std::vector<FarTune> cand;
    const int smax = far_field_max_splits(N);
    const bool ws = prec == DBIE_FP16X3 && K > 128;
    for (int sp : {1, 2, 3, 4, 6, 8, 12, 16, 24, 32}) {
        if (sp > smax || (double)sp * K * N * sizeof(double) > 2.0e9) break;
        cand.push_back(FarTune{sp, -1});
        if (ws) {
            cand.push_back(FarTune{sp, 0});
            cand.push_back(FarTune{sp, 1});
            // the stored kernel matrix: 256-target workgroups fill the CUs at one or two splits
            if (sp <= 2) {
                cand.push_back(FarTune{sp, 2});
                cand.push_back(FarTune{sp, 3});
                cand.push_back(FarTune{sp, 4});
            }
        }
    }
    return cand;
}

// the fastest timed candidate (the first one while none is timed)
void settle(TuneState &ts) {
    size_t bi = 0;
    float bm = -1.f;
    bool all = ts.next >= kTuneRounds * ts.cand.size();
    for (size_t i = 0; i < ts.cand.size(); ++i) {
        if (i < ts.nt.size() && ts.nt[i] < (int)kTuneRounds) all = false;
        if (ts.ms[i] < 0) {
            all = false;
            continue;
        }
        if (bm < 0 || ts.ms[i] < bm) {
            bm = ts.ms[i];
            bi = i;
        }
    }
    ts.best = ts.cand.empty() ? FarTune() : ts.cand[bi];
    ts.done = all;
}

// Collects the timings whose events have completed (caller holds g_tune_mu).
void poll_timings_locked() {
    for (size_t i = 0; i < g_pending.size();) {
        PendingTiming &p = g_pending[i];
        const hipError_t q = hipEventQuery(p.e1);
        if (q == hipErrorNotReady) {
            ++i;
            continue;
        }
        float ms = 1e30f;  // a failed query rules the candidate out
        if (q != hipSuccess || hipEventElapsedTime(&ms, p.e0, p.e1) != hipSuccess) ms = 1e30f;
        auto it = g_tune.find(p.key);
        if (it != g_tune.end() && p.idx < it->second.ms.size() && !it->second.done) {
            TuneState &ts = it->second;
            ts.ms[p.idx] = ts.ms[p.idx] < 0 ? ms : std::min(ts.ms[p.idx], ms);
            if (p.idx < ts.nt.size()) ++ts.nt[p.idx];
            settle(ts);
        }
        (void)hipEventDestroy(p.e0);
        (void)hipEventDestroy(p.e1);
        p = g_pending.back();
        g_pending.pop_back();
    }
}

void apply_dev_cfg(dbie_op *op, int K, const double *X, int64_t ldx, double *Y, int64_t ldy, dbie_precision prec,
                   const FarTune &tune);

// HIP-graph form of apply_dev_cfg for repeated applies (GMRES calls the operator on the same
// buffers every iteration).  The first call with a set of arguments runs eagerly (it also sizes
// the workspaces and computes the fp16 scale); the second is captured on the operator's stream and
// instantiated; later ones replay the graph.  The kernels need no memsets or host synchronisation,
// so the graph holds kernel nodes only.  A graph is retired when a workspace it writes has been
// reallocated by another apply; at most 8 are kept per operator (least recently used goes).
void apply_graphed(dbie_op *op, int K, const double *X, int64_t ldx, double *Y, int64_t ldy, dbie_precision prec,
                   const FarTune &tune) {
    if (!op->graphs || op->graphs_failed || op->profiling) {
        apply_dev_cfg(op, K, X, ldx, Y, ldy, prec, tune);
        return;
    }
    using AG = dbie_op::ApplyGraph;
    auto live = [&](const AG &a) {
        return a.part == op->part.p && a.scratch == op->mscratch.p && a.xt == op->xt32.p && a.gst == op->gst.p;
    };
    AG *hit = nullptr;
    for (size_t i = 0; i < op->agraphs.size();) {
        AG &a = op->agraphs[i];
        if (a.exec && !live(a)) {  // a workspace moved: retire
            (void)hipGraphExecDestroy(a.exec);
            a = op->agraphs.back();
            op->agraphs.pop_back();
            continue;
        }
        if (a.X == X && a.ldx == ldx && a.Y == Y && a.ldy == ldy && a.K == K && a.prec == (int)prec &&
            a.splits == tune.splits && a.variant == tune.variant)
            hit = &a;
        ++i;
    }
    if (hit && hit->exec) {
        hit->last_use = ++op->graph_tick;
        DBIE_HIP(hipGraphLaunch(hit->exec, op->stream));
        ++op->graph_replays;
        return;
    }
    if (!hit) {  // first sighting: run eagerly and remember the arguments
        apply_dev_cfg(op, K, X, ldx, Y, ldy, prec, tune);
        if (op->agraphs.size() >= 8) {
            auto lru = std::min_element(op->agraphs.begin(), op->agraphs.end(),
                                        [](const AG &a, const AG &b) { return a.last_use < b.last_use; });
            if (lru->exec) (void)hipGraphExecDestroy(lru->exec);
            *lru = op->agraphs.back();
            op->agraphs.pop_back();
        }
        op->agraphs.push_back(AG{X, ldx, Y, ldy, K, (int)prec, tune.splits, tune.variant, nullptr, nullptr, nullptr,
                                 nullptr, nullptr, ++op->graph_tick});
        return;
    }
    // second sighting: capture, instantiate, launch
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool ok = hipStreamBeginCapture(op->stream, hipStreamCaptureModeThreadLocal) == hipSuccess;
    if (ok) {
        try {
            apply_dev_cfg(op, K, X, ldx, Y, ldy, prec, tune);
        } catch (...) {
            ok = false;
        }
        ok = hipStreamEndCapture(op->stream, &graph) == hipSuccess && ok && graph;
        ok = ok && hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) == hipSuccess;
        if (graph) (void)hipGraphDestroy(graph);
    }
    if (!ok) {  // capture is not available here: stay eager for this operator
        (void)hipGetLastError();
        if (exec) (void)hipGraphExecDestroy(exec);
        op->graphs_failed = true;
        op->drop_graphs();
        apply_dev_cfg(op, K, X, ldx, Y, ldy, prec, tune);
        return;
    }
    hit->exec = exec;
    hit->part = op->part.p;
    hit->scratch = op->mscratch.p;
    hit->xt = op->xt32.p;
    hit->gst = op->gst.p;
    hit->last_use = ++op->graph_tick;
    ++op->graph_captures;
    DBIE_HIP(hipGraphLaunch(exec, op->stream));
    ++op->graph_replays;
}

void apply_dev(dbie_op *op, int K, const double *X, int64_t ldx, double *Y, int64_t ldy, dbie_precision prec);

// Apply of a fused operator: each member's columns through the member's own apply (tuned shape, HIP
// graph replay) on the member's stream; the rows past a member's N are set to zero.  The member
// streams start after the work queued on the fused stream and the fused stream continues after all
// of them, so the members' applies overlap on the device.
void fused_apply(dbie_op *op, int K, const double *X, int64_t ldx, double *Y, int64_t ldy, dbie_precision prec) {
    int total = 0;
    for (const auto &m : op->members) total = std::max(total, m.c0 + m.K);
    if (K != total) fail(DBIE_DIM_MISMATCH, "fused apply: K differs from the members' columns");
    const int64_t N = op->N();
    DBIE_HIP(hipEventRecord(op->fork_ev, op->stream));
    for (auto &m : op->members) {
        DBIE_HIP(hipStreamWaitEvent(m.op->stream, op->fork_ev, 0));
        apply_dev(m.op, m.K, X + (size_t)m.c0 * ldx, ldx, Y + (size_t)m.c0 * ldy, ldy, prec);
        const int64_t Nm = m.op->N();
        if (Nm < N)
            DBIE_HIP(hipMemset2DAsync(Y + (size_t)m.c0 * ldy + Nm, sizeof(double) * ldy, 0, sizeof(double) * (N - Nm),
                                      m.K, m.op->stream));
        DBIE_HIP(hipEventRecord(m.done, m.op->stream));
        DBIE_HIP(hipStreamWaitEvent(op->stream, m.done, 0));
    }
}

void apply_dev(dbie_op *op, int K, const double *X, int64_t ldx, double *Y, int64_t ldy, dbie_precision prec) {
    if (!op->members.empty()) {
        fused_apply(op, K, X, ldx, Y, ldy, prec);
        return;
    }
    if (!autotune_on() || K < 1) {
        apply_graphed(op, K, X, ldx, Y, ldy, prec, FarTune());
        return;
    }
    const TuneKey key = tune_key(op, K, prec, excludes_near(op, K, prec));
    FarTune use;
    bool timed = false;
    size_t idx = 0;
    int max_splits = 0;
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        poll_timings_locked();
        auto it = g_tune.find(key);
        if (it == g_tune.end()) {
            TuneState ts;
            ts.cand = tune_candidates(op->N(), K, prec);
            ts.ms.assign(ts.cand.size(), -1.f);
            ts.nt.assign(ts.cand.size(), 0);
            if (ts.cand.size() <= 1) {
                ts.next = kTuneRounds * ts.cand.size();
                ts.nt.assign(ts.cand.size(), (int)kTuneRounds);
            }
            settle(ts);
            it = g_tune.emplace(key, ts).first;
        }
        TuneState &ts = it->second;
        if (!ts.done && ts.next < kTuneRounds * ts.cand.size()) {
            idx = ts.next++ % ts.cand.size();
            use = ts.cand[idx];
            timed = true;
            for (const FarTune &c : ts.cand) max_splits = std::max(max_splits, c.splits);
        } else {
            use = ts.best;
        }
    }
    if (!timed) {
        apply_graphed(op, K, X, ldx, Y, ldy, prec, use);
        return;
    }
    // partial sums at their largest first, so that no timed apply allocates
    op->part.ensure(sizeof(double) * ((int64_t)max_splits * K * op->N() + 64 * op->N()));
    hipEvent_t e0, e1;
    DBIE_HIP(hipEventCreate(&e0));
    DBIE_HIP(hipEventCreate(&e1));
    DBIE_HIP(hipEventRecord(e0, op->stream));
    apply_dev_cfg(op, K, X, ldx, Y, ldy, prec, use);
    DBIE_HIP(hipEventRecord(e1, op->stream));
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_pending.push_back(PendingTiming{key, idx, e0, e1});
}

// Workspaces of one apply: the far-field partial sums (part_cap doubles), the far-field operand
// staging (w.*X tile-major, or its fp16/bf16 split), the tile-major FP32 copy of X for the full
// near blocks, and 4 bytes for the fp16 scale reduction.
struct ApplyWs {
    double *part = nullptr;
    int64_t part_cap = 0;
    void *mscratch = nullptr;
    float *xt32 = nullptr;
    void *tmp4 = nullptr;
};

// Partial-sum elements an apply with this launch shape needs (the launchers reduce the split
// count to what part_cap holds, never below one split of K x N).
int64_t apply_part_need(int64_t N, int K, const FarTune &tune) {
    return std::max<int64_t>(far_field_part_elems(N, K), (int64_t)tune.splits * std::max(K, 1) * N + 64 * N);
}

// a forced split count stays within what the launch accepts and 2 GB of partial sums
FarTune clamp_tune(int64_t N, int K, const FarTune &tune_in) {
    FarTune tune = tune_in;
    if (tune.splits > 0) {
        const int64_t cap = std::max<int64_t>(1, (int64_t)(2.0e9 / (8.0 * std::max(K, 1) * N)));
        tune.splits = (int)std::min<int64_t>({(int64_t)tune.splits, (int64_t)far_field_max_splits(N), cap});
    }
    return tune;
}

// Y = A X with the given workspaces (no allocation here).
void apply_core(dbie_op *op, int K, const double *X, int64_t ldx, double *Y, int64_t ldy, dbie_precision prec,
                const FarTune &tune, const ApplyWs &ws) {
    const int64_t N = op->N();
    int S;
    bool exclude = false;  // far field leaves the near panel pairs out (their full interaction is added below)
    if (prec != DBIE_FP32 && far_field_mfma_supported((int)prec, K)) {
        // profiling events bracket the contraction kernel alone (not the operand preparation)
        hipEvent_t e0 = nullptr, e1 = nullptr;
        if (op->profiling) {
            DBIE_HIP(hipEventCreate(&e0));
            DBIE_HIP(hipEventCreate(&e1));
            op->prof[prec == DBIE_FP64 ? 0 : 1].push_back({e0, e1});
        }
        if (prec == DBIE_FP16X3 && op->f16_gscale == 0.0f) {
            op->f16_gscale = far_field_f16_scale(op->stream, op->dg, ws.tmp4);
            if (op->f16_gscale == 0.0f) DBIE_HIP(hipGetLastError());
        }
        exclude = excludes_near(op, K, prec);
        FarStored fs;
        if (op->gst_state == 1 && op->gst_excl == exclude) {
            fs.g = op->gst.p;
            fs.tile_off = op->gst_off.as<int>();
            fs.ntile = op->gst_ntile;
        }
        S = far_field_mfma(op->stream, op->dg, (int)prec, K, X, ldx, ws.part, ws.part_cap, ws.mscratch,
                           op->f16_gscale, exclude, e0, e1, tune, fs.g ? &fs : nullptr);
    } else {
        ProfScope ps(op, prec == DBIE_FP64 ? 0 : 1);
        if (prec == DBIE_FP64) {
            S = far_field_f64(op->stream, op->dg, FAR_DT, K, X, ldx, ws.part, ws.part_cap, tune);
        } else {
            exclude = excludes_near(op, K, prec);
            S = far_field_f32(op->stream, op->dg, K, X, ldx, ws.part, ws.part_cap, tune, exclude);
        }
    }
    {
        ProfScope ps(op, 2);
        const NearVals nv = exclude ? NEAR_UP32 : (prec != DBIE_FP64 ? NEAR_C32 : NEAR_C64);
        // the full near blocks multiply the far field's split-FP16 operands of this X where they
        // were packed for it, else a tile-major FP32 copy of X
        NearH3 h3;
        const bool small_h3 = ws.mscratch && small_k_h3(op, K, prec);
        const bool use_h3 = exclude && (near_uses_h3(op, K) || small_h3);
        if (small_h3) f16_split_operands(op->stream, N, K, X, ldx, op->dg.w, ws.mscratch);
        if (use_h3) far_f16_operands(ws.mscratch, N, K, &h3);
        near_combine(op->stream, op->bsr, nv, S, ws.part, N, K, X, ldx, op->diag.as<double>(), Y, ldy,
                     nv == NEAR_UP32 && !use_h3 ? ws.xt32 : nullptr, use_h3 ? &h3 : nullptr);
    }
    DBIE_HIP(hipGetLastError());
}

// Builds the operator's stored kernel matrix (split-FP16 far field) when free device memory holds it
// with a reserve of max(8 GiB, 1/16 of the device) to spare; otherwise the operator declines it for
// good.  Never under stream capture (the first, eager apply of a launch shape builds it).
void build_stored(dbie_op *op, bool exclude) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(op->stream, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) {
        (void)hipGetLastError();
        return;
    }
    std::vector<int> off;
    const int64_t nblk = far_stored_layout(op->tile_plane_h, op->N(), off);
    const size_t bytes = (size_t)nblk * far_stored_block_bytes();
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) {
        (void)hipGetLastError();
        op->gst_state = -1;
        return;
    }
    const size_t reserve = std::max<size_t>((size_t)8 << 30, total_b / 16);
    if (nblk <= 0 || bytes + reserve > free_b) {
        op->gst_state = -1;
        return;
    }
    try {
        op->gst.ensure(bytes);
        upload(op->gst_off, off.data(), off.size(), op->stream);
    } catch (const Error &) {
        op->gst.release();
        op->gst_off.release();
        op->gst_state = -1;
        return;
    }
    op->gst_ntile = (int)((op->N() + 255) / 256);
    FarStored fs;
    fs.g = op->gst.p;
    fs.tile_off = op->gst_off.as<int>();
    fs.ntile = op->gst_ntile;
    far_stored_build(op->stream, op->dg, op->f16_gscale, exclude, fs);
    DBIE_HIP(hipGetLastError());
    op->gst_excl = exclude;
    op->gst_state = 1;
}

void apply_dev_cfg(dbie_op *op, int K, const double *X, int64_t ldx, double *Y, int64_t ldy, dbie_precision prec,
                   const FarTune &tune_in) {
    const int64_t N = op->N();
    const FarTune tune = clamp_tune(N, K, tune_in);
    if (prec == DBIE_FP16X3 && tune.variant >= 2 && K > 128 && op->stored_mode == 1 && op->gst_state == 0) {
        if (op->f16_gscale == 0.0f) {
            op->gmax.ensure(sizeof(unsigned));
            op->f16_gscale = far_field_f16_scale(op->stream, op->dg, op->gmax.p);
            if (op->f16_gscale == 0.0f) DBIE_HIP(hipGetLastError());
        }
        build_stored(op, excludes_near(op, K, prec));
    }
    const int64_t need = apply_part_need(N, K, tune);
    op->part.ensure(need * sizeof(double));
    ApplyWs ws;
    ws.part = op->part.as<double>();
    ws.part_cap = need;
    if (prec != DBIE_FP32 && far_field_mfma_supported((int)prec, K)) {
        if (prec == DBIE_FP16X3 && op->f16_gscale == 0.0f) {
            op->gmax.ensure(sizeof(unsigned));
            ws.tmp4 = op->gmax.p;
        }
        op->mscratch.ensure(far_field_mfma_scratch_bytes(N, (int)prec, K));
        ws.mscratch = op->mscratch.p;
        if (excludes_near(op, K, prec) && !near_uses_h3(op, K)) {
            op->xt32.ensure(near_xt_bytes(N, K));
            ws.xt32 = op->xt32.as<float>();
        }
    } else if (small_k_h3(op, K, prec)) {
        op->mscratch.ensure(far_field_mfma_scratch_bytes(N, (int)DBIE_FP16X3, K));
        ws.mscratch = op->mscratch.p;
    }
    apply_core(op, K, X, ldx, Y, ldy, prec, tune, ws);
}

// ---- screening boxes ---------------------------------------------------------------------------
// Boxes of permittivity for the screened volume sources, on the host and (after upload) on the device.
struct HostBoxes {
    std::vector<double> lo, hi, eps;
    ScreenBoxes sb;  // device pointers filled by upload_boxes
};

void check_screen_mode(int mode, double bandwidth) {
    if (mode < SCREEN_SHARP || mode > SCREEN_SOFT_MIX_INV) fail(DBIE_BAD_ARG, "screening mode must be 0, 1 or 2");
    if (mode != SCREEN_SHARP && !(bandwidth > 0)) fail(DBIE_BAD_ARG, "bandwidth must be positive");
}

// screened_volume_source(boxes, epses, eps_out, vs, mode; tol) (src/core/sources.jl:495-515): box b spans
// centre -/+ sides/2; tol < 0 takes the default sqrt(eps) * the largest side of any box.  form 1 selects
// the one-region formulas (screened_volume_source(Lx, Ly, Lz, vs, eps_in, eps_out, mode), :368-388).
HostBoxes box_list(int nb, const double *boxes, const double *epses, double eps_out, int mode, double bandwidth,
                   double tol, int form) {
    if (nb < 1 || !boxes || !epses) fail(DBIE_BAD_ARG, "need at least one box and one permittivity per box");
    if (form != 0 && form != 1) fail(DBIE_BAD_ARG, "form must be 0 (several boxes) or 1 (one region)");
    if (form == 1 && nb != 1) fail(DBIE_BAD_ARG, "the one-region form takes one box");
    check_screen_mode(mode, bandwidth);
    HostBoxes hb;
    double max_dim = 0.0;
    for (int b = 0; b < nb; ++b) {
        const double *bx = boxes + 6 * b;
        for (int k = 0; k < 3; ++k) {
            hb.lo.push_back(bx[k] - bx[3 + k] / 2);
            hb.hi.push_back(bx[k] + bx[3 + k] / 2);
        }
        max_dim = std::max(max_dim, std::max(std::max(std::fabs(bx[3]), std::fabs(bx[4])), std::fabs(bx[5])));
        hb.eps.push_back(epses[b]);
    }
    hb.sb.nb = nb;
    hb.sb.eps_out = eps_out;
    hb.sb.tol = tol >= 0 ? tol : std::sqrt(2.220446049250313e-16) * max_dim;
    hb.sb.mode = mode;
    hb.sb.bandwidth = bandwidth;
    hb.sb.single = form == 1;
    return hb;
}

// the interface's bounding box with its uniform permittivities (_interface_box_bounds,
// _uniform_interface_eps, src/core/sources.jl:283-313) and the default tolerance (:390-400)
HostBoxes interface_box(const dbie_op *op) {
    const Geometry &g = op->g;
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int64_t p = 0; p < g.P; ++p)
        for (int c = 0; c < 4; ++c)
            for (int k = 0; k < 3; ++k) {
                lo[k] = std::min(lo[k], g.corners[12 * p + 3 * c + k]);
                hi[k] = std::max(hi[k], g.corners[12 * p + 3 * c + k]);
            }
    for (int64_t p = 1; p < g.P; ++p)
        if (op->eps_in[p] != op->eps_in[0] || op->eps_out[p] != op->eps_out[0])
            fail(DBIE_BAD_ARG, "screened_volume_source requires a uniform eps across the interface");
    HostBoxes hb;
    hb.lo.assign(lo, lo + 3);
    hb.hi.assign(hi, hi + 3);
    hb.eps.assign(1, op->eps_in[0]);
    hb.sb.nb = 1;
    hb.sb.eps_out = op->eps_out[0];
    hb.sb.tol = std::sqrt(2.220446049250313e-16) *
                std::max(std::max(std::fabs(hi[0] - lo[0]), std::fabs(hi[1] - lo[1])), std::fabs(hi[2] - lo[2]));
    hb.sb.mode = SCREEN_SHARP;
    hb.sb.single = true;
    return hb;
}

// copies the box arrays into dbuf (lo, hi, eps back to back) and points sb at them
ScreenBoxes upload_boxes(const HostBoxes &hb, DBuf &dbuf, hipStream_t st) {
    std::vector<double> all(hb.lo);
    all.insert(all.end(), hb.hi.begin(), hb.hi.end());
    all.insert(all.end(), hb.eps.begin(), hb.eps.end());
    upload(dbuf, all.data(), all.size(), st);
    ScreenBoxes sb = hb.sb;
    sb.lo = dbuf.as<double>();
    sb.hi = sb.lo + 3 * sb.nb;
    sb.eps = sb.hi + 3 * sb.nb;
    return sb;
}

// screened densities of S host points (device arithmetic)
void screen_host(hipStream_t st, int64_t S, const double *xyz, const double *rho, const HostBoxes &hb,
                 double *rho_out) {
    if (S == 0) return;
    DBuf dpos(true), drho(true), dout(true), dbox(true);
    upload(dpos, xyz, 3 * (size_t)S, st);
    upload(drho, rho, (size_t)S, st);
    dout.ensure(sizeof(double) * S);
    screen_boxes(st, S, dpos.as<double>(), drho.as<double>(), nullptr, upload_boxes(hb, dbox, st), dout.as<double>());
    DBIE_HIP(hipGetLastError());
    DBIE_HIP(hipMemcpyAsync(rho_out, dout.p, sizeof(double) * S, hipMemcpyDeviceToHost, st));
    DBIE_HIP(hipStreamSynchronize(st));
}

// rhs_dielectric_box3d for K volume sources (src/solver/dielectric_box3d.jl:106-124,
// src/solver/multi_rhs.jl:151-211): column k sums the points col_ptr[k] .. col_ptr[k+1]-1, each
// with charge w * rho' (rho' screened by hb, or rho itself when hb is NULL), times 1 / eps_src.
void rhs_volume_host(dbie_op *op, int K, const int64_t *col_ptr, const double *xyz, const double *w,
                     const double *rho, const HostBoxes *hb, double eps_src, double *F, int64_t ldf) {
    const int64_t N = op->N();
    if (K < 1 || !col_ptr || !F) fail(DBIE_BAD_ARG, "K >= 1, col_ptr and F required");
    if (ldf < N) fail(DBIE_DIM_MISMATCH, "ldf < N");
    const int64_t S = col_ptr[K];
    if (col_ptr[0] != 0 || S < 0) fail(DBIE_BAD_ARG, "col_ptr must start at 0");
    for (int k = 0; k < K; ++k)
        if (col_ptr[k + 1] < col_ptr[k]) fail(DBIE_BAD_ARG, "col_ptr must be non-decreasing");
    if (S > 0 && (!xyz || !w || !rho)) fail(DBIE_BAD_ARG, "source arrays required");
    hipStream_t st = op->stream;
    DBuf dcp(true), dpos(true), dw(true), drho(true), dwr(true), dbox(true);
    upload(dcp, col_ptr, (size_t)K + 1, st);
    upload(dpos, xyz, 3 * (size_t)S, st);
    upload(dw, w, (size_t)S, st);
    upload(drho, rho, (size_t)S, st);
    dwr.ensure(sizeof(double) * std::max<int64_t>(S, 1));
    if (hb) {
        screen_boxes(st, S, dpos.as<double>(), drho.as<double>(), dw.as<double>(), upload_boxes(*hb, dbox, st),
                     dwr.as<double>());
    } else {
        // charges w .* rho, 1 / eps_src applied at the end
        const double lo[3] = {-INFINITY, -INFINITY, -INFINITY}, hi[3] = {INFINITY, INFINITY, INFINITY};
        screen_sharp(st, S, dpos.as<double>(), drho.as<double>(), dw.as<double>(), lo, hi, 0.0, 1.0, 1.0,
                     dwr.as<double>());
    }
    op->yout.ensure(sizeof(double) * N * K);
    rhs_volume(st, op->dg, K, dcp.as<int64_t>(), dpos.as<double>(), dwr.as<double>(), op->yout.as<double>(), N,
               1.0 / eps_src);
    DBIE_HIP(hipGetLastError());
    DBIE_HIP(hipMemcpy2DAsync(F, sizeof(double) * ldf, op->yout.p, sizeof(double) * N, sizeof(double) * N, K,
                              hipMemcpyDeviceToHost, st));
    DBIE_HIP(hipStreamSynchronize(st));
}

// Column 2-norms of an N x K block (contiguous), returned on the host.
void col_norms(dbie_op *op, const double *V, int64_t N, int K, DBuf &partial, DBuf &h, std::vector<double> &out) {
    partial.ensure(dots_partial_elems(N, K, 1) * sizeof(double));
    h.ensure(sizeof(double) * K);
    block_dots(op->stream, V, 1, V, N, K, partial.as<double>(), h.as<double>());
    out.resize(K);
    DBIE_HIP(hipMemcpyAsync(out.data(), h.p, sizeof(double) * K, hipMemcpyDeviceToHost, op->stream));
    DBIE_HIP(hipStreamSynchronize(op->stream));
    for (auto &v : out) v = std::sqrt(std::max(v, 0.0));
}

struct GmresOut {
    int iters = 0;
    bool solved = false;
    std::vector<double> resid;  // per column: the residual norm the Krylov process ended with (from the
                                // Givens recurrence per column, the explicit block residual for block GMRES)
};

// Right preconditioner of the Krylov solves (solve option precondition): 1 = the diagonal scaling
// M = D^-1, D = diag(t) (dielectric_diagonal_scaling, src/solver/multi_rhs.jl:234-247); 2 = one Neumann
// term of the near-field inverse, M = D^-1 - D^-1 N D^-1 with N the near panel pairs' full interaction
// (correction + direct Dt entries), i.e. (D + N)^-1 to first order.  For the box it takes GMRES from 15
// to 9 iterations (eps_in 4) and from 23 to 13 (eps_in 80) at the cost of one near-field product per
// iteration.  Applied as z = D^-1 v, Y = N z - t z (a near product whose diagonal is -t), M v = -D^-1 Y;
// FP64 solves multiply the FP64 full blocks, the low-precision inner solves the FP32 ones.
struct Precond {
    const double *dinv = nullptr;   // 1 / t
    const double *ndinv = nullptr;  // -1 / t (near form)
    const double *negt = nullptr;   // -t (near form)
    DevBSR near;                    // the blocks of N (near form)
    NearVals vals = NEAR_C64;
    bool use_near = false;
    const double *w = nullptr;      // quadrature weights: the split-FP16 form multiplies w .* x
    bool h3 = false;                // near form on the split-FP16 blocks of the low-precision operator
    int sweeps = 1;                 // near form: Jacobi sweeps z <- z0 - D^-1 N z from z = z0 = D^-1 in
    // fused operator: member b's own preconditioner on its rows and columns (the fused dinv unused)
    std::vector<Precond> mem;
};

template <typename TV>
void apply_precond(dbie_op *op, const Precond &pc, const TV *in, double *out, int64_t N, int K, int64_t ldi = 0,
                   int64_t ldo = 0);

// Fused operator: each member's preconditioner on its columns and first N_m rows (leading dimension the
// fused N) on the member's stream, the rows past N_m set to zero; forked from and joined to the fused
// stream as the fused apply is.
template <typename TV>
void apply_precond_fused(dbie_op *op, const Precond &pc, const TV *in, double *out, int64_t N) {
    DBIE_HIP(hipEventRecord(op->fork_ev, op->stream));
    for (size_t b = 0; b < op->members.size(); ++b) {
        const dbie_op::Member &m = op->members[b];
        DBIE_HIP(hipStreamWaitEvent(m.op->stream, op->fork_ev, 0));
        const int64_t Nm = m.op->N();
        apply_precond(m.op, pc.mem[b], in + (size_t)m.c0 * N, out + (size_t)m.c0 * N, Nm, m.K, N, N);
        if (Nm < N)
            DBIE_HIP(hipMemset2DAsync(out + (size_t)m.c0 * N + Nm, sizeof(double) * N, 0, sizeof(double) * (N - Nm),
                                      m.K, m.op->stream));
        DBIE_HIP(hipEventRecord(m.done, m.op->stream));
        DBIE_HIP(hipStreamWaitEvent(op->stream, m.done, 0));
    }
}

// out = M in (in and out may alias; leading dimensions ldi, ldo, 0: N).  Near form: z0 = D^-1 in, then
// each sweep z <- z0 - D^-1 N z (one sweep: the first Neumann term of (D + N)^-1; two: the second as well).
template <typename TV>
void apply_precond(dbie_op *op, const Precond &pc, const TV *in, double *out, int64_t N, int K, int64_t ldi,
                   int64_t ldo) {
    if (!pc.mem.empty()) {
        apply_precond_fused(op, pc, in, out, N);
        return;
    }
    if (ldi <= 0) ldi = N;
    if (ldo <= 0) ldo = N;
    hipStream_t st = op->stream;
    row_scale(st, pc.dinv, in, out, N, K, ldi, ldo);
    if (!pc.use_near) return;
    GmresWork &ws = op->gw;
    if (pc.h3) {  // any K: the split pads the operands to 16 columns
        // z <- -(1/t) .* (N z - t .* z0) with z0 = out: N z from z's split-FP16 operands (the three
        // products of the operator's near field), the scaling in the near kernel's epilogue; the last
        // sweep writes out in place (each element read and written by one thread)
        ws.PX.ensure(far_field_mfma_scratch_bytes(N, (int)DBIE_FP16X3, K));
        if (pc.sweeps > 1) ws.PY.ensure(sizeof(double) * (size_t)N * K);
        NearH3 h3;
        far_f16_operands(ws.PX.p, N, K, &h3);
        const double *z = out;
        int64_t ldz = ldo;
        for (int s = 0; s < pc.sweeps; ++s) {
            f16_split_operands(st, N, K, z, ldz, pc.w, ws.PX.p);
            const bool last = s + 1 == pc.sweeps;
            double *y = last ? out : ws.PY.as<double>();
            const int64_t ldy = last ? ldo : N;
            near_combine(st, pc.near, pc.vals, 0, nullptr, N, K, out, ldo, pc.negt, y, ldy, nullptr, &h3, pc.ndinv);
            z = y;
            ldz = ldy;
        }
        return;
    }
    ws.PY.ensure(sizeof(double) * (size_t)N * K);
    float *xt = nullptr;
    if (pc.vals == NEAR_UP32) {
        ws.PX.ensure(near_xt_bytes(N, K));
        xt = ws.PX.as<float>();
    }
    if (pc.sweeps == 1) {
        near_combine(st, pc.near, pc.vals, 0, nullptr, N, K, out, ldo, pc.negt, ws.PY.as<double>(), N, xt, nullptr);
        row_scale(st, pc.ndinv, ws.PY.as<double>(), out, N, K, N, ldo);
        return;
    }
    // two sweeps: z1 = z0 - D^-1 N z0 into PZ, then out = z0 - D^-1 N z1 = z0 - z1 - D^-1 (N z1 - t .* z1)
    ws.PZ.ensure(sizeof(double) * (size_t)N * K);
    double *z1 = ws.PZ.as<double>(), *py = ws.PY.as<double>();
    near_combine(st, pc.near, pc.vals, 0, nullptr, N, K, out, ldo, pc.negt, py, N, xt, nullptr);
    row_scale(st, pc.ndinv, py, z1, N, K);
    near_combine(st, pc.near, pc.vals, 0, nullptr, N, K, z1, N, pc.negt, py, N, xt, nullptr);
    jacobi_update(st, pc.ndinv, py, z1, out, N, K, ldo);
}

// GMRES(m) for K independent columns sharing each block matvec.  Right
// preconditioner pc (null: none).  X receives the solution (initial guess zero).
// tol_k = atol + rtol * ||B_k||.
//
// keep: store this solve's last Krylov cycle for later solves with the same operator.
// use: GCRO deflation with a stored space (C, U), A M U = C: every cycle starts from
// r0 = (I - C C^T) r with c = C^T r, each new Arnoldi vector is made orthogonal to C as well
// (b_j = C^T A M v_j), and the cycle's correction is M (U (c - B y) + V y), which minimises the
// residual over span(U) + the new Krylov space: the refinement continues the earlier space.
template <typename TV>
GmresOut gmres_t(dbie_op *op, int K, const double *B, double *X, dbie_precision prec, double rtol, double atol,
                  int itmax, int restart, const Precond *pc, KrylovRecycle *keep, const KrylovRecycle *use) {
    const int64_t N = op->N();
    const size_t NK = (size_t)N * K;
    const int m = std::max(1, std::min(restart > 0 ? restart : 100, std::max(itmax, 1)));
    // workspaces live in the operator and are reused by later solves; the Krylov basis grows
    // on demand (most solves stop long before the restart length)
    GmresWork &ws = op->gw;
    DBuf &W = ws.W, &Z = ws.Z, &U = ws.U, &partial = ws.partial, &hbuf = ws.hbuf, &nbuf = ws.nbuf;
    const bool defl = use && use->valid && use->nblk > 0;
    if (defl && use->elem != (int)sizeof(double)) fail(DBIE_BAD_ARG, "recycled Krylov space not stored in FP64");
    const int nb1 = defl ? use->nblk : 0, um = defl ? use->m : 0;
    const int cgs_passes = prec == DBIE_FP64 ? 2 : 1;
    W.ensure(sizeof(double) * NK);
    U.ensure(sizeof(double) * NK);
    constexpr bool kF64 = std::is_same<TV, double>::value;
    if (pc || !kF64) Z.ensure(sizeof(double) * NK);
    partial.ensure(sizeof(double) * std::max(dots_partial_elems(N, K, std::max(m + 1, nb1)), dots_partial_elems(N, K, 1)));
    hbuf.ensure(sizeof(double) * (size_t)(cgs_passes == 2 ? 2 * (m + 1) : m + 1) * K);
    nbuf.ensure(sizeof(double) * (size_t)(m + 1) * K);
    // pinned host staging: V1^T W of the last deflation (nb1 x K), then an Arnoldi step's CGS
    // coefficients of both passes ((m + 1) x K each) and the squared norms (K)
    ws.pin.ensure(sizeof(double) * ((size_t)nb1 + 2 * (size_t)(m + 1) + 1) * K);
    double *const a_h = ws.pin.as<double>();
    double *const ph1 = a_h + (size_t)nb1 * K, *const ph2 = ph1 + (size_t)(m + 1) * K, *const phn = ph2 + (size_t)(m + 1) * K;
    if (defl) {
        ws.abuf.ensure(sizeof(double) * (size_t)nb1 * K);
        ws.cbuf.ensure(sizeof(double) * (size_t)nb1 * K);
        std::vector<double> om_h((size_t)nb1 * K, 0.0);
        std::vector<int> n1_h(K, 0);
        for (int k = 0; k < K; ++k) {
            const int k1 = use->k1[k];
            if (k1 <= 0) continue;
            n1_h[k] = k1 + 1;
            for (int i = 0; i <= k1; ++i) om_h[(size_t)i * K + k] = use->omega[(size_t)k * (um + 1) + i];
        }
        upload(ws.om, om_h.data(), om_h.size(), op->stream);
        upload(ws.n1, n1_h.data(), n1_h.size(), op->stream);
    }
    hipStream_t st = op->stream;
    double *Wp = W.as<double>();
    const double *V1p = defl ? ws.V1.as<double>() : nullptr;  // kept by an FP64-basis solve
    // basis capacity in N x K blocks; growing copies the blocks already built
    auto basis = [&](int blocks) -> TV * {
        blocks = std::min(blocks, m + 1);
        if (ws.V.bytes >= sizeof(TV) * NK * blocks) return ws.V.as<TV>();
        const int cap = std::min(m + 1, std::max(blocks, std::max(16, 2 * (int)(ws.V.bytes / (sizeof(TV) * NK)))));
        DBuf nv(ws.V.pooled);
        nv.ensure(sizeof(TV) * NK * cap);
        if (ws.V.p) {
            DBIE_HIP(hipMemcpyAsync(nv.p, ws.V.p, ws.V.bytes, hipMemcpyDeviceToDevice, st));
            DBIE_HIP(hipStreamSynchronize(st));
        }
        std::swap(ws.V.p, nv.p);
        std::swap(ws.V.bytes, nv.bytes);
        return ws.V.as<TV>();
    };
    TV *Vp = basis(2);

    // Wd <- (I - C C^T) Wd with C C^T = V1 (I - omega omega^T) V1^T on the span of V1_(k1+1), on the
    // device; V1^T Wd is copied back for the projections, which project() forms after the next sync:
    // tproj (K x (um+1)) = Omega V1^T Wd per column, whose first k1 entries are C^T Wd
    std::vector<double> tproj;
    auto deflate = [&](double *Wd) {
        block_dots(st, V1p, nb1, Wd, N, K, partial.as<double>(), ws.abuf.as<double>());
        deflate_coef(st, ws.abuf.as<double>(), ws.om.as<double>(), ws.n1.as<int>(), nb1, K, ws.cbuf.as<double>());
        block_subtract(st, V1p, nb1, ws.cbuf.as<double>(), Wd, N, K);
        DBIE_HIP(hipMemcpyAsync(a_h, ws.abuf.p, sizeof(double) * (size_t)nb1 * K, hipMemcpyDeviceToHost, st));
    };
    auto project = [&]() {
        tproj.assign((size_t)(um + 1) * K, 0.0);
        for (int k = 0; k < K; ++k) {
            const int k1 = use->k1[k];
            if (k1 <= 0) continue;
            double *tk = &tproj[(size_t)k * (um + 1)];
            for (int i = 0; i <= k1; ++i) tk[i] = a_h[(size_t)i * K + k];
            const double *c = &use->cs[(size_t)k * um], *s = &use->sn[(size_t)k * um];
            for (int i = 0; i < k1; ++i) {
                const double u = c[i] * tk[i] + s[i] * tk[i + 1];
                tk[i + 1] = -s[i] * tk[i] + c[i] * tk[i + 1];
                tk[i] = u;
            }
        }
    };

    DBIE_HIP(hipMemsetAsync(X, 0, sizeof(double) * NK, st));
    // r0 = B (x0 = 0) lives in W
    DBIE_HIP(hipMemcpyAsync(Wp, B, sizeof(double) * NK, hipMemcpyDeviceToDevice, st));
    std::vector<double> beta, tol(K), cs((size_t)m * K), sn((size_t)m * K), gv((size_t)(m + 1) * K),
        hh((size_t)(m + 1) * K), yv((size_t)m * K);
    // The Hessenberg columns of the cycle, grown one Arnoldi step at a time (the restart length can be
    // far above the steps a solve takes; storage for all of it would be K (m + 1) m doubles per call):
    // Hc[j][k (j + 2) + i] = H_k(i, j), i <= j + 1.  Deflation: Bc[j][k um + i] = B_k(i, j).
    std::vector<std::vector<double>> Hc, Bc;
    auto Hat = [&](int k, int i, int l) -> double & { return Hc[l][(size_t)k * (l + 2) + i]; };
    std::vector<double> cvec;  // deflation: C^T r of the cycle's start (um per column)
    std::vector<int> nconv(K);
    std::vector<char> active(K), conv(K);
    GmresOut res;
    if (defl) {  // tolerances from ||B||, before the first projection
        col_norms(op, Wp, N, K, partial, nbuf, beta);
        for (int k = 0; k < K; ++k) {
            if (!std::isfinite(beta[k])) fail(DBIE_BAD_ARG, "right-hand side contains NaN or Inf");
            tol[k] = atol + rtol * beta[k];
        }
    }
    // X += M (V y + U (c - B y)) for the cycle just ended (y = 0 for inactive columns)
    auto update = [&](int nmax) {
        DBIE_HIP(hipMemsetAsync(U.p, 0, sizeof(double) * NK, st));
        if (nmax > 0) {
            h2d(hbuf.p, yv.data(), sizeof(double) * (size_t)nmax * K, st);
            block_combine(st, Vp, nmax, hbuf.as<double>(), U.as<double>(), N, K);
        }
        if (defl) {
            std::vector<double> z((size_t)nb1 * K, 0.0);
            for (int k = 0; k < K; ++k) {
                const int k1 = use->k1[k];
                if (k1 <= 0) continue;
                std::vector<double> rhs(k1);
                for (int i = 0; i < k1; ++i) {
                    double v = cvec[(size_t)k * um + i];
                    for (int j = 0; j < nconv[k]; ++j)
                        v -= Bc[j][(size_t)k * um + i] * yv[(size_t)j * K + k];
                    rhs[i] = v;
                }
                const double *Rk = &use->R[(size_t)k * (um + 1) * um];
                for (int i = k1 - 1; i >= 0; --i) {
                    double v = rhs[i];
                    for (int l = i + 1; l < k1; ++l) v -= Rk[(size_t)l * (um + 1) + i] * z[(size_t)l * K + k];
                    z[(size_t)i * K + k] = v / Rk[(size_t)i * (um + 1) + i];
                }
            }
            h2d(ws.abuf.p, z.data(), sizeof(double) * z.size(), st);
            block_combine(st, V1p, nb1, ws.abuf.as<double>(), U.as<double>(), N, K);
        }
        if (pc) apply_precond(op, *pc, U.as<double>(), U.as<double>(), N, K);
        add_to(st, U.as<double>(), X, N, K);
    };
    bool first = true;
    for (;;) {
        if (defl) deflate(Wp);
        col_norms(op, Wp, N, K, partial, nbuf, beta);  // synchronizes the stream
        if (defl) {
            project();
            cvec.assign((size_t)um * K, 0.0);
            for (int k = 0; k < K; ++k)
                for (int i = 0; i < use->k1[k]; ++i) cvec[(size_t)k * um + i] = tproj[(size_t)k * (um + 1) + i];
            Bc.clear();
        }
        bool any = false;
        for (int k = 0; k < K; ++k) {
            if (first && !defl) {
                if (!std::isfinite(beta[k])) fail(DBIE_BAD_ARG, "right-hand side contains NaN or Inf");
                tol[k] = atol + rtol * beta[k];
            }
            active[k] = beta[k] > tol[k] && beta[k] > 0.0;
            conv[k] = !active[k];
            nconv[k] = 0;
            any = any || active[k];
        }
        first = false;
        if (!any || res.iters >= itmax) {
            if (defl) {  // the projection alone: X += M U c
                std::fill(yv.begin(), yv.end(), 0.0);
                update(0);
            }
            res.resid = beta;
            res.solved = !any;
            break;
        }
        // V0 = r / beta
        Hc.clear();
        h2d(nbuf.p, beta.data(), sizeof(double) * K, st);
        col_scale(st, Wp, nbuf.as<double>(), true, Vp, N, K);
        std::fill(gv.begin(), gv.end(), 0.0);
        for (int k = 0; k < K; ++k) gv[(size_t)k * (m + 1)] = active[k] ? beta[k] : 0.0;
        std::vector<double> cur(beta);  // per column: residual estimate of the cycle so far
        int j = 0;
        for (; j < m && res.iters < itmax; ++j) {
            Vp = basis(j + 2);
            const TV *Vj = Vp + (size_t)j * NK;
            const double *Zin;
            if (pc || !kF64) {  // M v_j in FP64 (the apply's input)
                if (pc) apply_precond(op, *pc, Vj, Z.as<double>(), N, K);
                else row_scale(st, nullptr, Vj, Z.as<double>(), N, K);
                Zin = Z.as<double>();
            } else {
                Zin = reinterpret_cast<const double *>(Vj);
            }
            apply_dev(op, K, Zin, N, Wp, N, prec);
            if (defl) deflate(Wp);
            // classical Gram-Schmidt, twice (CGS2) for FP64 solves; once for the low-precision inner
            // solves of the refinement, whose operator is accurate to ~1e-7 and whose tolerance is
            // ~1e-5: the orthogonality CGS1 keeps is far below both
            block_dots(st, Vp, j + 1, Wp, N, K, partial.as<double>(), hbuf.as<double>());
            block_subtract(st, Vp, j + 1, hbuf.as<double>(), Wp, N, K);
            const double *h1 = ph1, *h2 = ph2;
            if (cgs_passes == 2) {
                // the second pass's coefficients in a buffer of their own (nbuf is reused below)
                block_dots(st, Vp, j + 1, Wp, N, K, partial.as<double>(), hbuf.as<double>() + (size_t)(j + 1) * K);
                block_subtract(st, Vp, j + 1, hbuf.as<double>() + (size_t)(j + 1) * K, Wp, N, K);
                DBIE_HIP(hipMemcpyAsync(ph2, hbuf.as<double>() + (size_t)(j + 1) * K, sizeof(double) * (j + 1) * K,
                                        hipMemcpyDeviceToHost, st));
            } else {
                std::fill(ph2, ph2 + (size_t)(j + 1) * K, 0.0);
            }
            DBIE_HIP(hipMemcpyAsync(ph1, hbuf.p, sizeof(double) * (j + 1) * K, hipMemcpyDeviceToHost, st));
            // the next basis vector from the squared norms on the device; one sync brings the
            // coefficients, the norms and the deflation projections to the host
            block_dots(st, Wp, 1, Wp, N, K, partial.as<double>(), nbuf.as<double>());
            col_normalize(st, Wp, nbuf.as<double>(), Vp + (size_t)(j + 1) * NK, N, K);
            DBIE_HIP(hipMemcpyAsync(phn, nbuf.p, sizeof(double) * K, hipMemcpyDeviceToHost, st));
            DBIE_HIP(hipStreamSynchronize(st));
            std::vector<double> hn(phn, phn + K);
            for (auto &v : hn) v = std::sqrt(std::max(v, 0.0));
            Hc.emplace_back((size_t)K * (j + 2), 0.0);
            if (defl) Bc.emplace_back((size_t)K * um, 0.0);
            if (defl) {
                project();
                for (int k = 0; k < K; ++k)
                    for (int i = 0; i < use->k1[k]; ++i)
                        Bc[j][(size_t)k * um + i] = tproj[(size_t)k * (um + 1) + i];
            }
            ++res.iters;
            bool all = true;
            for (int k = 0; k < K; ++k) {
                if (conv[k]) continue;
                double *col = &Hat(k, 0, j);
                for (int i = 0; i <= j; ++i) col[i] = h1[(size_t)i * K + k] + h2[(size_t)i * K + k];
                col[j + 1] = hn[k];
                double *c = &cs[(size_t)k * m], *s = &sn[(size_t)k * m], *g = &gv[(size_t)k * (m + 1)];
                for (int i = 0; i < j; ++i) {
                    const double t = c[i] * col[i] + s[i] * col[i + 1];
                    col[i + 1] = -s[i] * col[i] + c[i] * col[i + 1];
                    col[i] = t;
                }
                const double d = std::hypot(col[j], col[j + 1]);
                c[j] = d != 0.0 ? col[j] / d : 1.0;
                s[j] = d != 0.0 ? col[j + 1] / d : 0.0;
                col[j] = d;
                col[j + 1] = 0.0;
                g[j + 1] = -s[j] * g[j];
                g[j] = c[j] * g[j];
                nconv[k] = j + 1;
                cur[k] = std::fabs(g[j + 1]);
                if (std::fabs(g[j + 1]) <= tol[k] || hn[k] == 0.0) conv[k] = 1;
                else all = false;
            }
            op->note(0, cur);
            if (all) { ++j; break; }
        }
        // y_k = H_k^{-1} g_k (upper triangular, size nconv[k])
        std::fill(yv.begin(), yv.end(), 0.0);
        for (int k = 0; k < K; ++k) {
            if (!active[k]) continue;
            const int nk = nconv[k];
            const double *g = &gv[(size_t)k * (m + 1)];
            std::vector<double> y(nk);
            for (int i = nk - 1; i >= 0; --i) {
                double s = g[i];
                for (int l = i + 1; l < nk; ++l) s -= Hat(k, i, l) * y[l];
                y[i] = s / Hat(k, i, i);
            }
            for (int i = 0; i < nk; ++i) yv[(size_t)i * K + k] = y[i];
        }
        int nmax = 0;
        for (int k = 0; k < K; ++k) nmax = std::max(nmax, nconv[k]);
        update(nmax);
        res.resid.resize(K);
        for (int k = 0; k < K; ++k)
            res.resid[k] = active[k] ? std::fabs(gv[(size_t)k * (m + 1) + nconv[k]]) : beta[k];
        bool all_conv = true;
        for (int k = 0; k < K; ++k) all_conv = all_conv && conv[k];
        if (all_conv) { res.solved = true; break; }
        if (res.iters >= itmax) break;
        // restart: r = B - A X
        apply_dev(op, K, X, N, Wp, N, DBIE_FP64 == prec ? DBIE_FP64 : prec);
        sub_from(st, B, Wp, N, K);
    }
    DBIE_HIP(hipStreamSynchronize(st));
    if (keep) {
        keep->valid = false;
        int nmax = 0;
        for (int k = 0; k < K; ++k) nmax = std::max(nmax, nconv[k]);
        if (nmax > 0 && !defl) {
            // the last cycle's basis moves to V1 (its buffer is reused by the next solve's basis)
            std::swap(ws.V.p, ws.V1.p);
            std::swap(ws.V.bytes, ws.V1.bytes);
            keep->nblk = nmax + 1;
            keep->elem = (int)sizeof(TV);
            // stored for the steps the cycle took (mk), not the restart length
            const int mk = nmax;
            keep->m = mk;
            keep->k1 = nconv;
            keep->omega.assign((size_t)K * (mk + 1), 0.0);
            keep->cs.assign((size_t)K * mk, 0.0);
            keep->sn.assign((size_t)K * mk, 0.0);
            keep->R.assign((size_t)K * (mk + 1) * mk, 0.0);
            for (int k = 0; k < K; ++k) {
                const int k1 = nconv[k];
                double *om = &keep->omega[(size_t)k * (mk + 1)];
                om[k1] = 1.0;
                const double *c = &cs[(size_t)k * m], *s = &sn[(size_t)k * m];
                for (int i = k1 - 1; i >= 0; --i) {  // omega = G_0^T ... G_(k1-1)^T e_k1
                    const double x0 = om[i], x1 = om[i + 1];
                    om[i] = c[i] * x0 - s[i] * x1;
                    om[i + 1] = s[i] * x0 + c[i] * x1;
                }
                for (int i = 0; i < mk; ++i) {
                    keep->cs[(size_t)k * mk + i] = c[i];
                    keep->sn[(size_t)k * mk + i] = s[i];
                }
                double *Rk = &keep->R[(size_t)k * (mk + 1) * mk];
                for (int l = 0; l < k1; ++l)
                    for (int i = 0; i <= l + 1; ++i) Rk[(size_t)l * (mk + 1) + i] = Hat(k, i, l);
            }
            keep->valid = true;
        }
    }
    return res;
}

GmresOut gmres(dbie_op *op, int K, const double *B, double *X, dbie_precision prec, double rtol, double atol, int itmax,
               int restart, const Precond *pc, KrylovRecycle *keep = nullptr, const KrylovRecycle *use = nullptr) {
    // The low-precision inner solves of the refinement keep their Krylov basis in FP32: their operator
    // is accurate to ~1e-7 and their tolerance ~1e-5, and the orthogonalisation's passes over the
    // basis (the solve's largest HBM traffic after the operator) halve.  Not a space kept for a later
    // refinement: that one works on a residual ~1e-5 of the first, and a kept space rounded to FP32
    // deflates it so poorly that it takes about twice the iterations (measured at N = 64,384,
    // K = 256: 39 inner iterations instead of 19).  The deflating solve's own basis may be FP32.
    const bool f32 = prec != DBIE_FP64 && !keep;
    return f32 ? gmres_t<float>(op, K, B, X, prec, rtol, atol, itmax, restart, pc, keep, use)
               : gmres_t<double>(op, K, B, X, prec, rtol, atol, itmax, restart, pc, keep, use);
}

void blas_ok(rocblas_status st, const char *what) {
    if (st != rocblas_status_success) fail(DBIE_HIP_ERROR, std::string(what) + ": rocBLAS status " + std::to_string((int)st));
}

rocblas_handle op_blas(dbie_op *op) {
    if (!op->blas) {
        blas_ok(rocblas_create_handle(&op->blas), "rocblas_create_handle");
        blas_ok(rocblas_set_stream(op->blas, op->stream), "rocblas_set_stream");
    }
    return op->blas;
}

// C (p x q, leading dimension ldc) = beta C + A^T B for A (N x p) and B (N x q), both with leading
// dimension N: a GEMM whose reduction runs over the long dimension, split into row chunks computed as
// one strided-batched GEMM and summed by a kernel (one GEMM call over all N rows runs at a small
// fraction of the FP64 matrix peak for these shapes).
void gram(dbie_op *op, GmresWork &ws, const double *A, int p, const double *B, int q, int64_t N, double *C,
          int64_t ldc, double beta) {
    rocblas_handle h = op_blas(op);
    const double one = 1.0, zero = 0.0;
    const int64_t nc = std::max<int64_t>(1, std::min<int64_t>(64, N / 2048));
    const int64_t rc = N / nc;  // rows per chunk; the last N - nc rc rows go in one more GEMM
    ws.gpart.ensure(sizeof(double) * (size_t)nc * p * q);
    blas_ok(rocblas_dgemm_strided_batched(h, rocblas_operation_transpose, rocblas_operation_none, p, q, (rocblas_int)rc,
                                          &one, A, (rocblas_int)N, rc, B, (rocblas_int)N, rc, &zero,
                                          ws.gpart.as<double>(), p, (rocblas_stride)p * q, (rocblas_int)nc),
            "rocblas_dgemm_strided_batched");
    sum_blocks(op->stream, (int)nc, p, q, ws.gpart.as<double>(), beta, C, ldc);
    const int64_t tail = N - nc * rc;
    if (tail > 0)
        blas_ok(rocblas_dgemm(h, rocblas_operation_transpose, rocblas_operation_none, p, q, (rocblas_int)tail, &one,
                              A + nc * rc, (rocblas_int)N, B + nc * rc, (rocblas_int)N, &one, C, (rocblas_int)ldc),
                "rocblas_dgemm");
}

// W (N x K) = Q R with Q overwriting W, R upper triangular (K x K, leading dimension ldr): Cholesky QR
// twice (G = W^T W = R1^T R1, W <- W R1^-1, again, R = R2 R1), or once with npass = 1.  Returns false,
// with W unchanged in meaning, when a Gram matrix is not numerically positive definite.
bool chol_qr2(dbie_op *op, GmresWork &ws, double *W, int64_t N, int K, double *R, int64_t ldr, int npass = 2) {
    rocblas_handle h = op_blas(op);
    hipStream_t st = op->stream;
    const double one = 1.0;
    ws.R.ensure(sizeof(double) * (size_t)K * K);
    ws.info.ensure(sizeof(rocblas_int));
    double *G = ws.R.as<double>();
    for (int pass = 0; pass < npass; ++pass) {
        gram(op, ws, W, K, W, K, N, G, K, 0.0);
        blas_ok(rocsolver_dpotrf(h, rocblas_fill_upper, K, G, K, ws.info.as<rocblas_int>()), "rocsolver_dpotrf");
        rocblas_int info = 0;
        DBIE_HIP(hipMemcpyAsync(&info, ws.info.p, sizeof(info), hipMemcpyDeviceToHost, st));
        DBIE_HIP(hipStreamSynchronize(st));
        if (info != 0) return false;
        blas_ok(rocblas_dtrsm(h, rocblas_side_right, rocblas_fill_upper, rocblas_operation_none,
                              rocblas_diagonal_non_unit, (rocblas_int)N, K, &one, G, K, W, (rocblas_int)N),
                "rocblas_dtrsm");
        if (pass == 0) {
            triu_copy(st, K, G, K, R, ldr);
        } else {  // R = R2 R1
            blas_ok(rocblas_dtrmm(h, rocblas_side_left, rocblas_fill_upper, rocblas_operation_none,
                                  rocblas_diagonal_non_unit, K, K, &one, G, K, R, (rocblas_int)ldr, R, (rocblas_int)ldr),
                    "rocblas_dtrmm");
        }
    }
    return true;
}

// Block GMRES(m) (Krylov.block_gmres as called by _block_gmres_solve, src/solver/multi_rhs.jl:144-148):
// one block Krylov space span{R, A M R, (A M)^2 R, ...} for all K columns, right preconditioner M =
// diag(dinv).  Block Arnoldi with classical Gram-Schmidt twice as two pairs of GEMMs over the whole
// basis; each new block is orthonormalised by Cholesky QR twice (Householder QR, rocSOLVER
// geqrf/orgqr, when a Gram matrix is not positive definite), whose R is the sub-diagonal block of
// the block Hessenberg Hbar.  The small least-squares problem min ||E1 S0 - Hbar Y|| is solved through
// the normal equations with an incrementally extended Cholesky factor of Hbar^T Hbar (Hbar = V^T A M V
// has the conditioning of the operator on the Krylov space, a second-kind operator here, so squaring
// it costs no accuracy the 1e-10 stopping test can see), and the residual E1 S0 - Hbar Y is formed
// explicitly: its column norms are the residual norms of the K columns.  Stops when every column meets
// atol + rtol ||B_k|| (the reference's Krylov.jl tests the Frobenius norm of the block residual; the
// per-column test is the stricter of the two).  X receives the solution (initial guess zero).
GmresOut block_gmres(dbie_op *op, int K, const double *B, double *X, dbie_precision prec, double rtol, double atol,
                     int itmax, int restart, const Precond *pc) {
    const int64_t N = op->N();
    const size_t NK = (size_t)N * K;
    const int m = std::max(1, std::min(restart > 0 ? restart : 40, std::max(itmax, 1)));
    GmresWork &ws = op->gw;
    hipStream_t st = op->stream;
    rocblas_handle h = op_blas(op);
    const rocblas_int KK = K;
    const int64_t ldh = (int64_t)(m + 1) * K;  // Hbar: (m+1)K x mK; residual (m+1)K x K
    const int64_t ldr = (int64_t)m * K;        // Cholesky factor of Hbar^T Hbar: mK x mK (upper)
    ws.W.ensure(sizeof(double) * NK);
    ws.U.ensure(sizeof(double) * NK);
    if (pc) ws.Z.ensure(sizeof(double) * NK);
    ws.V.ensure(sizeof(double) * NK * (m + 1));
    ws.H.ensure(sizeof(double) * (size_t)ldh * m * K);
    ws.G.ensure(sizeof(double) * (size_t)ldh * K);   // residual E1 S0 - Hbar Y
    ws.Y.ensure(sizeof(double) * (size_t)ldh * K);   // CGS coefficients, then Y
    ws.tau.ensure(sizeof(double) * (size_t)K);
    ws.PV.ensure(sizeof(double) * (size_t)ldr * ldr);  // Cholesky factor
    ws.PT.ensure(sizeof(double) * (size_t)(ldr + 2 * K) * K);  // right-hand side z, S0, K x K scratch
    ws.partial.ensure(sizeof(double) * std::max(dots_partial_elems(N, K, 1), dots_partial_elems(ldh, K, 1)));
    ws.nbuf.ensure(sizeof(double) * K);
    ws.info.ensure(sizeof(rocblas_int));
    double *Vp = ws.V.as<double>(), *Wp = ws.W.as<double>(), *Hp = ws.H.as<double>(), *Res = ws.G.as<double>();
    double *Yp = ws.Y.as<double>(), *taup = ws.tau.as<double>(), *Rp = ws.PV.as<double>();
    double *zp = ws.PT.as<double>(), *S0 = zp + (size_t)ldr * K, *Dp = S0 + (size_t)K * K;
    const double one = 1.0, zero = 0.0, mone = -1.0;
    std::vector<double> bnorm, tol(K), rn;
    col_norms(op, B, N, K, ws.partial, ws.nbuf, bnorm);
    for (int k = 0; k < K; ++k) {
        if (!std::isfinite(bnorm[k])) fail(DBIE_BAD_ARG, "right-hand side contains NaN or Inf");
        tol[k] = atol + rtol * bnorm[k];
    }
    DBIE_HIP(hipMemsetAsync(X, 0, sizeof(double) * NK, st));
    DBIE_HIP(hipMemcpyAsync(Wp, B, sizeof(double) * NK, hipMemcpyDeviceToDevice, st));
    GmresOut res;
    // V (N x K) = Q R in place (V holds a copy of W), R (upper triangular) into Rdst
    // FP64 solves orthogonalise twice (block CGS2, Cholesky QR twice); the low-precision inner solves
    // of the refinement once each: their operator is accurate to ~1e-7 and their tolerance ~1e-5
    const int npass = prec == DBIE_FP64 ? 2 : 1;
    auto orthonormalise = [&](double *V, double *Rdst, int64_t ld) {
        if (chol_qr2(op, ws, V, N, K, Rdst, ld, npass)) return;
        // the Gram matrix lost definiteness (nearly dependent directions): Householder QR
        DBIE_HIP(hipMemcpyAsync(V, Wp, sizeof(double) * NK, hipMemcpyDeviceToDevice, st));
        blas_ok(rocsolver_dgeqrf(h, (rocblas_int)N, KK, V, (rocblas_int)N, taup), "rocsolver_dgeqrf");
        triu_copy(st, K, V, N, Rdst, ld);
        blas_ok(rocsolver_dorgqr(h, (rocblas_int)N, KK, KK, V, (rocblas_int)N, taup), "rocsolver_dorgqr");
    };
    for (;;) {
        // residual check of the current iterate (W holds R = B - A X)
        col_norms(op, Wp, N, K, ws.partial, ws.nbuf, rn);
        res.resid = rn;
        bool all = true;
        for (int k = 0; k < K; ++k) all = all && rn[k] <= tol[k];
        if (all) { res.solved = true; break; }
        if (res.iters >= itmax) break;
        // R = V0 S0
        DBIE_HIP(hipMemcpyAsync(Vp, Wp, sizeof(double) * NK, hipMemcpyDeviceToDevice, st));
        orthonormalise(Vp, S0, K);
        int j = 0;
        bool conv = false, breakdown = false;
        for (; j < m && res.iters < itmax; ++j) {
            double *Vj = Vp + (size_t)j * NK, *Vn = Vp + (size_t)(j + 1) * NK;
            double *Hc = Hp + (size_t)j * K * ldh;  // block column j of Hbar
            const rocblas_int nb = (rocblas_int)((j + 1) * K), jk = (rocblas_int)(j * K);
            const double *Zin = Vj;
            if (pc) {
                apply_precond(op, *pc, Vj, ws.Z.as<double>(), N, K);
                Zin = ws.Z.as<double>();
            }
            apply_dev(op, K, Zin, N, Wp, N, prec);
            // block CGS2: Hc = V^T W, W -= V Hc, twice (the second correction accumulated into Hc)
            DBIE_HIP(hipMemsetAsync(Hc, 0, sizeof(double) * (size_t)ldh * K, st));
            for (int pass = 0; pass < npass; ++pass) {
                gram(op, ws, Vp, nb, Wp, K, N, Yp, ldh, 0.0);
                blas_ok(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, (rocblas_int)N, KK, nb, &mone,
                                      Vp, (rocblas_int)N, Yp, (rocblas_int)ldh, &one, Wp, (rocblas_int)N),
                        "rocblas_dgemm");
                blas_ok(rocblas_dgeam(h, rocblas_operation_none, rocblas_operation_none, nb, KK, &one, Hc,
                                      (rocblas_int)ldh, &one, Yp, (rocblas_int)ldh, Hc, (rocblas_int)ldh),
                        "rocblas_dgeam");
            }
            // W = V_{j+1} R: R is the sub-diagonal block of block column j
            DBIE_HIP(hipMemcpyAsync(Vn, Wp, sizeof(double) * NK, hipMemcpyDeviceToDevice, st));
            orthonormalise(Vn, Hc + (size_t)(j + 1) * K, ldh);
            // extend the Cholesky factor R of Hbar^T Hbar by block column j:
            //   Rc = R^-T (Hbar_old^T h),  Rd = chol(h^T h - Rc^T Rc)
            double *Rc = Rp + (size_t)jk * ldr, *Rd = Rc + jk;
            if (j > 0) {
                blas_ok(rocblas_dgemm(h, rocblas_operation_transpose, rocblas_operation_none, jk, KK, nb, &one, Hp,
                                      (rocblas_int)ldh, Hc, (rocblas_int)ldh, &zero, Rc, (rocblas_int)ldr),
                        "rocblas_dgemm");
                blas_ok(rocblas_dtrsm(h, rocblas_side_left, rocblas_fill_upper, rocblas_operation_transpose,
                                      rocblas_diagonal_non_unit, jk, KK, &one, Rp, (rocblas_int)ldr, Rc, (rocblas_int)ldr),
                        "rocblas_dtrsm");
            }
            blas_ok(rocblas_dgemm(h, rocblas_operation_transpose, rocblas_operation_none, KK, KK, nb + KK, &one, Hc,
                                  (rocblas_int)ldh, Hc, (rocblas_int)ldh, &zero, Rd, (rocblas_int)ldr),
                    "rocblas_dgemm");
            if (j > 0)
                blas_ok(rocblas_dgemm(h, rocblas_operation_transpose, rocblas_operation_none, KK, KK, jk, &mone, Rc,
                                      (rocblas_int)ldr, Rc, (rocblas_int)ldr, &one, Rd, (rocblas_int)ldr),
                        "rocblas_dgemm");
            blas_ok(rocsolver_dpotrf(h, rocblas_fill_upper, KK, Rd, (rocblas_int)ldr, ws.info.as<rocblas_int>()),
                    "rocsolver_dpotrf");
            // right-hand side z = Hbar^T E1 S0: block j is Hbar(0:K, block j)^T S0
            blas_ok(rocblas_dgemm(h, rocblas_operation_transpose, rocblas_operation_none, KK, KK, KK, &one, Hc,
                                  (rocblas_int)ldh, S0, KK, &zero, zp + jk, (rocblas_int)ldr),
                    "rocblas_dgemm");
            ++res.iters;
            rocblas_int info = 0;
            DBIE_HIP(hipMemcpyAsync(&info, ws.info.p, sizeof(info), hipMemcpyDeviceToHost, st));
            DBIE_HIP(hipStreamSynchronize(st));
            if (info != 0) {  // Hbar lost rank: the Krylov space is exhausted; solve with the previous blocks
                breakdown = true;
                break;
            }
            // Y = (R^T R)^-1 z, residual E1 S0 - Hbar Y and its column norms
            blas_ok(rocblas_dgeam(h, rocblas_operation_none, rocblas_operation_none, nb, KK, &one, zp, (rocblas_int)ldr,
                                  &zero, zp, (rocblas_int)ldr, Yp, (rocblas_int)ldh),
                    "rocblas_dgeam");
            blas_ok(rocblas_dtrsm(h, rocblas_side_left, rocblas_fill_upper, rocblas_operation_transpose,
                                  rocblas_diagonal_non_unit, nb, KK, &one, Rp, (rocblas_int)ldr, Yp, (rocblas_int)ldh),
                    "rocblas_dtrsm");
            blas_ok(rocblas_dtrsm(h, rocblas_side_left, rocblas_fill_upper, rocblas_operation_none,
                                  rocblas_diagonal_non_unit, nb, KK, &one, Rp, (rocblas_int)ldr, Yp, (rocblas_int)ldh),
                    "rocblas_dtrsm");
            const rocblas_int nr = nb + KK;  // rows of Hbar's first j + 1 block columns
            blas_ok(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, nr, KK, nb, &mone, Hp,
                                  (rocblas_int)ldh, Yp, (rocblas_int)ldh, &zero, Res, nr),
                    "rocblas_dgemm");
            blas_ok(rocblas_dgeam(h, rocblas_operation_none, rocblas_operation_none, KK, KK, &one, Res, nr, &one, S0, KK,
                                  Res, nr),
                    "rocblas_dgeam");
            std::vector<double> resn;
            col_norms(op, Res, nr, K, ws.partial, ws.nbuf, resn);
            res.resid = resn;
            op->note(0, resn);
            conv = true;
            for (int k = 0; k < K && conv; ++k) conv = resn[k] <= tol[k];
            if (conv) { ++j; break; }
        }
        if (breakdown && j > 0) {  // Y for the first j blocks (the factor of those is intact)
            const rocblas_int nb = (rocblas_int)(j * K);
            blas_ok(rocblas_dgeam(h, rocblas_operation_none, rocblas_operation_none, nb, KK, &one, zp, (rocblas_int)ldr,
                                  &zero, zp, (rocblas_int)ldr, Yp, (rocblas_int)ldh),
                    "rocblas_dgeam");
            blas_ok(rocblas_dtrsm(h, rocblas_side_left, rocblas_fill_upper, rocblas_operation_transpose,
                                  rocblas_diagonal_non_unit, nb, KK, &one, Rp, (rocblas_int)ldr, Yp, (rocblas_int)ldh),
                    "rocblas_dtrsm");
            blas_ok(rocblas_dtrsm(h, rocblas_side_left, rocblas_fill_upper, rocblas_operation_none,
                                  rocblas_diagonal_non_unit, nb, KK, &one, Rp, (rocblas_int)ldr, Yp, (rocblas_int)ldh),
                    "rocblas_dtrsm");
        }
        // X += M V(:, 0:jK) Y
        const rocblas_int nj = (rocblas_int)(j * K);
        if (nj > 0) {
            blas_ok(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, (rocblas_int)N, KK, nj, &one, Vp,
                                  (rocblas_int)N, Yp, (rocblas_int)ldh, &zero, ws.U.as<double>(), (rocblas_int)N),
                    "rocblas_dgemm");
            if (pc) apply_precond(op, *pc, ws.U.as<double>(), ws.U.as<double>(), N, K);
            add_to(st, ws.U.as<double>(), X, N, K);
        }
        if (conv) { res.solved = true; break; }
        if (res.iters >= itmax || (breakdown && j == 0)) break;
        // restart (or continue after a breakdown): W = B - A X
        apply_dev(op, K, X, N, Wp, N, prec);
        sub_from(st, B, Wp, N, K);
    }
    DBIE_HIP(hipStreamSynchronize(st));
    return res;
}

// Solve A X = F for K columns, F and X on the device (N x K, ld = N).  Returns DBIE_OK or
// DBIE_NOT_CONVERGED (X and stats filled either way); throws on argument or runtime errors.
dbie_status solve_dev(dbie_op *op, int K, const double *dF, double *dX, const dbie_solve_opts *opts,
                      dbie_solve_stats *stats) {
    const int64_t N = op->N();
    dbie_solve_opts o;
    if (opts) o = *opts;
    else dbie_solve_opts_default(&o);
    if (o.itmax < 1) fail(DBIE_BAD_ARG, "itmax must be >= 1");
    if (!(o.rtol >= 0) || !(o.atol >= 0)) fail(DBIE_BAD_ARG, "tolerances must be >= 0");
    check_precision(o.precision);
    if (o.krylov != 0 && o.krylov != 1) fail(DBIE_BAD_ARG, "krylov must be 0 (per column) or 1 (block)");
    if (o.precondition < 0 || o.precondition > 3)
        fail(DBIE_BAD_ARG, "precondition must be 0 (none), 1 (diagonal scaling), 2 (near-field) or 3 (near-field, "
                           "two sweeps)");
    hipStream_t st = op->stream;
    const size_t NK = (size_t)N * K;
    DBuf dR(true), dD(true), dT(true), dinv(true), partial(true), hb(true);
    // preconditioners: pc64 for FP64 solves, pcl for the low-precision inner solves (the near form
    // multiplies FP64 or FP32 / split-FP16 full blocks).  A fused operator takes each member's own
    // near-field preconditioner on the member's rows and columns, or (diagonal scaling) its unit diagonal
    Precond pc64, pcl;
    const Precond *p64 = nullptr, *plo = nullptr;
    std::vector<DBuf> mdinv(op->members.size());
    // the preconditioners of one (non-fused) operator o2, its inverse diagonals in buf
    auto build = [&](dbie_op *o2, DBuf &buf, Precond &q64, Precond &ql) {
        const int64_t n2 = o2->N();
        const bool near = o.precondition >= 2 && o2->members.empty() && o2->bsr.nb > 0;
        std::vector<double> inv((size_t)n2 * (near ? 3 : 1));
        for (int64_t i = 0; i < n2; ++i) {
            inv[i] = 1.0 / o2->diag_h[i];
            if (near) {
                inv[n2 + i] = -inv[i];
                inv[2 * n2 + i] = -o2->diag_h[i];
            }
        }
        upload(buf, inv.data(), inv.size(), st);
        q64.dinv = buf.as<double>();
        if (near) {
            q64.ndinv = buf.as<double>() + n2;
            q64.negt = buf.as<double>() + 2 * n2;
        }
        q64.sweeps = o.precondition == 3 ? 2 : 1;
        ql = q64;
        if (near && o.precision == DBIE_FP64) {
            if (!o2->full64_built) {
                o2->bsr_full64.ensure(sizeof(double) * (size_t)o2->bsr.nb * o2->bsr.bd * o2->bsr.bd);
                near_full_blocks(st, o2->dg, o2->bsr, o2->bsr_full64.as<double>());
                DBIE_HIP(hipGetLastError());
                o2->full64_built = true;
            }
            q64.near = o2->bsr;
            q64.near.val64 = o2->bsr_full64.as<double>();
            q64.near.val32 = q64.near.valup32 = nullptr;
            q64.near.valh = nullptr;
            q64.near.sr_ptr = nullptr;
            q64.vals = NEAR_C64;
            q64.use_near = true;
        }
        if (near && o.precision != DBIE_FP64 && o2->bsr.valup32) {
            ql.near = o2->bsr;
            ql.h3 = o2->bsr.valh && o2->bsr.mfma_order && o2->bsr.bd == 16;
            ql.w = o2->w.as<double>();
            if (!ql.h3) {
                ql.near.valh = nullptr;
                ql.near.sr_ptr = nullptr;
            }
            ql.vals = NEAR_UP32;
            ql.use_near = true;
        }
    };
    if (o.precondition) {
        build(op, dinv, pc64, pcl);
        if (o.precondition >= 2 && !op->members.empty()) {
            pc64.mem.resize(op->members.size());
            pcl.mem.resize(op->members.size());
            for (size_t b = 0; b < op->members.size(); ++b)
                build(op->members[b].op, mdinv[b], pc64.mem[b], pcl.mem[b]);
        }
        p64 = &pc64;
        plo = &pcl;
    }
    dbie_solve_stats s{};
    std::vector<double> bnorm, rn;
    // residual history of this solve, relative to ||F_k||
    col_norms(op, dF, N, K, partial, hb, bnorm);
    op->hist.clear();
    op->hist_scale.resize(K);
    for (int k = 0; k < K; ++k) op->hist_scale[k] = bnorm[k] > 0 ? 1.0 / bnorm[k] : 0.0;
    struct ScaleReset {
        dbie_op *op;
        ~ScaleReset() { op->hist_scale.clear(); }
    } scale_reset{op};
    hipEvent_t e0, e1;
    DBIE_HIP(hipEventCreate(&e0));
    DBIE_HIP(hipEventCreate(&e1));
    DBIE_HIP(hipEventRecord(e0, st));
    if (o.precision == DBIE_FP64) {
        const GmresOut r = o.krylov == 1 ? block_gmres(op, K, dF, dX, DBIE_FP64, o.rtol, o.atol, o.itmax, o.restart, p64)
                                         : gmres(op, K, dF, dX, DBIE_FP64, o.rtol, o.atol, o.itmax, o.restart, p64);
        s.niter = r.iters;
        s.solved = r.solved;
    } else {
        // Mixed precision: FP64 iterative refinement around low-precision inner GMRES.
        // While corrections are large the residual is recomputed with the FP64 operator;
        // once they are small (||D_k|| <= 1e-3 ||X_k||) it is updated with the low-precision
        // operator, R -= A_low D, whose error is ~delta*||A||*||D||, and a final FP64
        // residual verifies convergence before the solve returns.
        dR.ensure(sizeof(double) * NK);
        dD.ensure(sizeof(double) * NK);
        col_norms(op, dF, N, K, partial, hb, bnorm);
        DBIE_HIP(hipMemsetAsync(dX, 0, sizeof(double) * NK, st));
        DBIE_HIP(hipMemcpyAsync(dR.p, dF, sizeof(double) * NK, hipMemcpyDeviceToDevice, st));
        rn = bnorm;
        bool exact = true;
        int n_exact = 0;
        std::vector<double> dn, xn;
        // per column: bound on |true residual - rn| from the low-precision residual updates since the
        // last FP64 residual
        std::vector<double> ubound(K, 0.0);
        const bool trace = std::getenv("DBIE_TRACE") != nullptr;
        auto worst = [&](const std::vector<double> &v) {
            double m = 0.0;
            for (int k = 0; k < K; ++k) m = std::max(m, bnorm[k] > 0 ? v[k] / bnorm[k] : v[k]);
            return m;
        };
        for (int it = 0;; ++it) {
            bool done = true;
            std::vector<double> mask(K);
            for (int k = 0; k < K; ++k) {
                const bool ok = rn[k] <= o.atol + o.rtol * bnorm[k];
                mask[k] = ok ? 0.0 : 1.0;
                done = done && ok;
            }
            if (done && exact) { s.solved = 1; break; }
            if (done) {  // verify the updated residual with the FP64 operator
                if (trace) std::fprintf(stderr, "[dbie] verify: updated residual %.3e\n", worst(rn));
                apply_dev(op, K, dX, N, dR.as<double>(), N, DBIE_FP64);
                sub_from(st, dF, dR.as<double>(), N, K);
                ++s.nrefine;
                exact = true;
                ++n_exact;
                std::fill(ubound.begin(), ubound.end(), 0.0);
                col_norms(op, dR.as<double>(), N, K, partial, hb, rn);
                op->note(1, rn);
                continue;
            }
            if (it >= o.max_refine || s.niter >= o.itmax) break;
            // without the FP64 verification the solve may end on this inner solve's residual norms:
            // once a true residual is known, run it to half of what the tolerance still allows
            double irtol = o.inner_rtol;
            if (!o.verify && n_exact >= 1) {
                double r_need = o.inner_rtol;
                for (int k = 0; k < K; ++k)
                    if (mask[k] != 0.0 && rn[k] > 0.0) r_need = std::min(r_need, 0.5 * (o.atol + o.rtol * bnorm[k]) / rn[k]);
                irtol = std::max(r_need, 1e-7);
            }
            hb.ensure(sizeof(double) * K);
            h2d(hb.p, mask.data(), sizeof(double) * K, st);
            col_scale(st, dR.as<double>(), hb.as<double>(), false, dR.as<double>(), N, K);
            // the first inner solve keeps its Krylov space, the later ones deflate it
            KrylovRecycle &rec = op->gw.rec;
            if (it == 0) rec.valid = false;
            const bool recyc = o.recycle != 0 && o.krylov == 0;
            const GmresOut r =
                o.krylov == 1 ? block_gmres(op, K, dR.as<double>(), dD.as<double>(), o.precision, irtol, 0.0,
                                            o.itmax - s.niter, o.restart, plo)
                              : gmres(op, K, dR.as<double>(), dD.as<double>(), o.precision, irtol, 0.0,
                                      o.itmax - s.niter, o.restart, plo, recyc && it == 0 ? &rec : nullptr,
                                      recyc && it > 0 ? &rec : nullptr);
            s.niter += r.iters;
            if (trace) std::fprintf(stderr, "[dbie] refine %d: inner iters %d, residual %.3e (%s)\n", it, r.iters,
                                    worst(rn), exact ? "fp64" : "updated");
            add_to(st, dD.as<double>(), dX, N, K);
            bool small = n_exact >= 1;
            if (small) {
                col_norms(op, dD.as<double>(), N, K, partial, hb, dn);
                col_norms(op, dX, N, K, partial, hb, xn);
                for (int k = 0; k < K && small; ++k)
                    if (mask[k] != 0.0 && dn[k] > 1e-3 * xn[k]) small = false;
            }
            // With small corrections the updated residual R - A_low D differs from the true one by at
            // most delta ||A|| ||D|| (delta: twice the low-precision operator's measured relative
            // error, ||A|| taken as 10), and these bounds add up over the updates since the last FP64
            // residual.  Split FP16 has delta 1e-8 also on the edge-localised vectors corrections are
            // (its far field leaves the near pairs out); the other forms keep the near pairs' direct
            // entries in their far field, whose rounding cancellation against the corrections
            // amplifies about a hundredfold on such vectors (delta 1e-4).
            std::vector<double> ub_next;
            if (small) {
                const double delta = o.precision == DBIE_FP16X3 ? 1e-8 : o.precision == DBIE_BF16 ? 1e-1 : 1e-4;
                ub_next = ubound;
                for (int k = 0; k < K; ++k)
                    if (mask[k] != 0.0) ub_next[k] += delta * 10.0 * dn[k];
            }
            // Without verification (verify = 0) the solve ends here when every column's inner residual
            // norm meets half the tolerance and the accumulated bound a quarter of it (split FP16 at
            // rtol 1e-10: corrections up to 2.5e-4 ||F_k|| after an FP64 residual; the other forms in
            // practice end on an FP64 residual).
            if (!o.verify && small && (int)r.resid.size() == K) {
                bool end = true;
                for (int k = 0; k < K && end; ++k) {
                    const double tk = o.atol + o.rtol * bnorm[k];
                    if (mask[k] != 0.0 && !(r.resid[k] <= 0.5 * tk && ub_next[k] <= 0.25 * tk)) end = false;
                }
                if (end) {
                    for (int k = 0; k < K; ++k)
                        if (mask[k] != 0.0) rn[k] = r.resid[k];
                    op->note(2, rn);
                    s.solved = 1;
                    break;
                }
            }
            // The updated residual's norm is the one the inner GMRES ended with, up to the loss of
            // orthogonality of the Krylov basis.  When those norms already meet the tolerance ten
            // times over, the update would only confirm it: go to the FP64 verification at once (a
            // failed verification costs an FP64 apply and another inner solve, hence the margin).
            bool est_ok = small && (int)r.resid.size() == K;
            for (int k = 0; k < K && est_ok; ++k)
                if (mask[k] != 0.0 && !(r.resid[k] <= 0.1 * (o.atol + o.rtol * bnorm[k]))) est_ok = false;
            if (est_ok) {
                for (int k = 0; k < K; ++k)
                    if (mask[k] != 0.0) rn[k] = r.resid[k];
                exact = false;
                ubound = ub_next;
                op->note(2, rn);
                continue;
            }
            if (small) {
                dT.ensure(sizeof(double) * NK);
                apply_dev(op, K, dD.as<double>(), N, dT.as<double>(), N, o.precision);
                sub_to(st, dT.as<double>(), dR.as<double>(), N, K);
                exact = false;
                ubound = ub_next;
            } else {
                apply_dev(op, K, dX, N, dR.as<double>(), N, DBIE_FP64);
                sub_from(st, dF, dR.as<double>(), N, K);
                ++s.nrefine;
                exact = true;
                ++n_exact;
                std::fill(ubound.begin(), ubound.end(), 0.0);
            }
            col_norms(op, dR.as<double>(), N, K, partial, hb, rn);
            op->note(exact ? 1 : 2, rn);
        }
    }
    DBIE_HIP(hipEventRecord(e1, st));
    DBIE_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    DBIE_HIP(hipEventElapsedTime(&ms, e0, e1));
    s.time_ms = ms;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if ((o.precision == DBIE_FP64 || !o.verify) && o.check_residual) {
        dR.ensure(sizeof(double) * NK);
        col_norms(op, dF, N, K, partial, hb, bnorm);
        apply_dev(op, K, dX, N, dR.as<double>(), N, DBIE_FP64);
        sub_from(st, dF, dR.as<double>(), N, K);
        col_norms(op, dR.as<double>(), N, K, partial, hb, rn);
        op->note(1, rn);
    }
    double mx = 0.0;
    for (size_t k = 0; k < rn.size(); ++k) mx = std::max(mx, bnorm[k] > 0 ? rn[k] / bnorm[k] : rn[k]);
    s.max_rel_residual = rn.empty() ? -1.0 : mx;
    DBIE_HIP(hipStreamSynchronize(st));
    if (stats) *stats = s;
    return s.solved ? DBIE_OK : DBIE_NOT_CONVERGED;
}

}  // namespace

dbie_status solve_fused_host(int B, dbie_op *const *ops, const int *Ks, const double *const *F, const int64_t *ldf,
                             double *const *S, const int64_t *lds, const dbie_solve_opts *opts,
                             dbie_solve_stats *stats) {
    if (B < 1 || !ops || !Ks || !F || !ldf || !S || !lds) fail(DBIE_BAD_ARG, "fused solve: B >= 1 and arrays required");
    // block Arnoldi mixes columns (block orthogonalisation), and the columns of different members belong
    // to different operators: the fused solve runs GMRES per column
    if (opts && opts->krylov != 0)
        fail(DBIE_BAD_ARG, "fused solve: krylov must be 0 (GMRES per column; block GMRES would mix the columns "
                           "of different operators)");
    ensure_device();
    int64_t Nmax = 0;
    int Ktot = 0;
    for (int b = 0; b < B; ++b) {
        check_op(ops[b]);
        if (!ops[b]->members.empty()) fail(DBIE_BAD_ARG, "fused solve: a member is itself a fused operator");
        for (int c = 0; c < b; ++c)
            if (ops[c] == ops[b]) fail(DBIE_BAD_ARG, "fused solve: an operator appears twice (merge its columns)");
        if (Ks[b] < 1 || !F[b] || !S[b]) fail(DBIE_BAD_ARG, "fused solve: K >= 1 and buffers per member");
        if (ldf[b] < ops[b]->N() || lds[b] < ops[b]->N()) fail(DBIE_DIM_MISMATCH, "leading dimension smaller than N");
        Nmax = std::max(Nmax, ops[b]->N());
        Ktot += Ks[b];
    }
    std::unique_ptr<dbie_op> f(new dbie_op);
    f->gw.set_pooled();  // the next fused solve of this thread reuses the Krylov workspaces
    f->g.N = Nmax;
    f->graphs = false;  // the members replay their own graphs
    f->diag_h.assign(Nmax, 1.0);  // right preconditioner: each member is a single contrast (a scalar
                                  // multiple of the identity, which GMRES is invariant to)
    DBIE_HIP(hipStreamCreateWithFlags(&f->stream, hipStreamNonBlocking));
    DBIE_HIP(hipEventCreateWithFlags(&f->fork_ev, hipEventDisableTiming));
    int c0 = 0;
    for (int b = 0; b < B; ++b) {
        hipEvent_t e = nullptr;
        DBIE_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        f->members.push_back(dbie_op::Member{ops[b], c0, Ks[b], e});
        c0 += Ks[b];
    }
    hipStream_t st = f->stream;
    const size_t NK = (size_t)Nmax * Ktot;
    DBuf dF(true), dX(true);
    dF.ensure(sizeof(double) * NK);
    dX.ensure(sizeof(double) * NK);
    DBIE_HIP(hipMemsetAsync(dF.p, 0, sizeof(double) * NK, st));
    for (const auto &m : f->members) {
        const int b = (int)(&m - f->members.data());
        DBIE_HIP(hipMemcpy2DAsync(dF.as<double>() + (size_t)m.c0 * Nmax, sizeof(double) * Nmax, F[b],
                                  sizeof(double) * ldf[b], sizeof(double) * m.op->N(), m.K, hipMemcpyHostToDevice, st));
    }
    const dbie_status rc = solve_dev(f.get(), Ktot, dF.as<double>(), dX.as<double>(), opts, stats);
    for (const auto &m : f->members) {
        const int b = (int)(&m - f->members.data());
        DBIE_HIP(hipMemcpy2DAsync(S[b], sizeof(double) * lds[b], dX.as<double>() + (size_t)m.c0 * Nmax,
                                  sizeof(double) * Nmax, sizeof(double) * m.op->N(), m.K, hipMemcpyDeviceToHost, st));
    }
    DBIE_HIP(hipStreamSynchronize(st));
    return rc;
}

size_t apply_noalloc_bytes(const dbie_op *op, int K, dbie_precision prec) {
    const int64_t N = op->N();
    auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
    size_t b = 256;  // fp16 scale reduction
    if (prec != DBIE_FP32 && far_field_mfma_supported((int)prec, K)) {
        b += up(far_field_mfma_scratch_bytes(N, (int)prec, K));
        if (excludes_near(op, K, prec) && !near_uses_h3(op, K)) b += up(near_xt_bytes(N, K));
    } else if (small_k_h3(op, K, prec)) {
        b += up(far_field_mfma_scratch_bytes(N, (int)DBIE_FP16X3, K));
    }
    return b + sizeof(double) * (size_t)std::max(K, 1) * (size_t)N;  // one split of partial sums
}

void apply_noalloc(dbie_op *op, int K, const double *X, int64_t ldx, double *Y, int64_t ldy, dbie_precision prec,
                   void *scratch, size_t bytes) {
    const int64_t N = op->N();
    const size_t need = apply_noalloc_bytes(op, K, prec);
    if (need > bytes)
        fail(DBIE_BAD_ARG, "apply without allocation: N = " + std::to_string(N) + ", K = " + std::to_string(K) +
                               " needs " + std::to_string(need) + " bytes of scratch, the library holds " +
                               std::to_string(bytes));
    // the launch shape the tuner settled on for this class (no exploration: it would time the
    // launch with events), else the built-in default
    FarTune tune;
    if (autotune_on()) {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto it = g_tune.find(tune_key(op, K, prec, excludes_near(op, K, prec)));
        if (it != g_tune.end() && it->second.done) tune = it->second.best;
    }
    tune = clamp_tune(N, K, tune);
    auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
    char *base = static_cast<char *>(scratch);
    size_t off = 256;
    ApplyWs ws;
    ws.tmp4 = base;
    if (prec != DBIE_FP32 && far_field_mfma_supported((int)prec, K)) {
        ws.mscratch = base + off;
        off += up(far_field_mfma_scratch_bytes(N, (int)prec, K));
        if (excludes_near(op, K, prec) && !near_uses_h3(op, K)) {
            ws.xt32 = reinterpret_cast<float *>(base + off);
            off += up(near_xt_bytes(N, K));
        }
    } else if (small_k_h3(op, K, prec)) {  // as dbie_apply_device: the near blocks' split operands
        ws.mscratch = base + off;
        off += up(far_field_mfma_scratch_bytes(N, (int)DBIE_FP16X3, K));
    }
    ws.part = reinterpret_cast<double *>(base + off);
    ws.part_cap = (int64_t)((bytes - off) / sizeof(double));  // >= K x N: at least one split fits
    // no kernel timing here: event creation is a library allocation
    struct NoProf {
        dbie_op *op;
        bool was;
        ~NoProf() { op->profiling = was; }
    } np{op, op->profiling};
    op->profiling = false;
    apply_core(op, K, X, ldx, Y, ldy, prec, tune, ws);
}

}  // namespace dbie

// ============================== C ABI ========================================
extern "C" {

const char *dbie_version(void) { return "dbie-mi355x 0.1 (gfx950)"; }

dbie_status dbie_last_error(char *buf, size_t len) {
    if (!buf || len == 0) return DBIE_BAD_ARG;
    std::strncpy(buf, g_last_error.c_str(), len - 1);
    buf[len - 1] = '\0';
    return DBIE_OK;
}

dbie_status dbie_init(int device) {
    return guard([&] {
        ensure_device();
        DBIE_HIP(hipSetDevice(device));
    });
}

void dbie_near_opts_default(dbie_near_opts *o) {
    if (!o) return;
    o->up_tol = 1e-9;
    o->max_order = 8;
    o->correct_edges = 1;
    o->adaptive_atol = 1e-9;
    o->adaptive_rtol = 1.4901161193847656e-8;
    o->adaptive_n_GL = 0;
    o->adaptive_max_depth = 20;
    o->build_corrections = 1;
    o->n_threads = 0;
    o->device_build = 1;
}

void dbie_solve_opts_default(dbie_solve_opts *o) {
    if (!o) return;
    o->rtol = 1e-10;
    o->atol = 0.0;
    o->itmax = 500;
    o->restart = 0;
    o->precondition = 1;
    o->precision = DBIE_FP64;
    o->inner_rtol = 1e-4;
    o->max_refine = 10;
    o->check_residual = 1;
    o->krylov = 0;
    o->recycle = 1;
    o->verify = 1;
}

dbie_status dbie_gauss_legendre(int n, double *x, double *w) {
    return guard([&] {
        if (!x || !w) fail(DBIE_BAD_ARG, "NULL output");
        gauss_legendre(n, x, w);
    });
}

dbie_status dbie_best_grid(double W, double H, double alpha, int *m, int *n) {
    return guard([&] {
        if (!m || !n) fail(DBIE_BAD_ARG, "NULL output");
        if (!(W > 0 && H > 0)) fail(DBIE_BAD_ARG, "W and H must be positive");
        if (alpha < 1) fail(DBIE_BAD_ARG, "alpha must be >= 1");
        if (!best_grid(W, H, alpha, m, n)) fail(DBIE_BAD_ARG, "no (m,n) found within search limits");
    });
}

dbie_status dbie_geometry_box(double Lx, double Ly, double Lz, int n_quad, double l_ec, double alpha,
                              dbie_geom **out) {
    return guard([&] {
        if (!out) fail(DBIE_BAD_ARG, "NULL output");
        std::unique_ptr<Geometry> g(make_box(Lx, Ly, Lz, n_quad, l_ec, alpha));
        auto *h = new dbie_geom;
        h->g = std::move(*g);
        *out = h;
    });
}

dbie_status dbie_geometry_from_panels(int64_t P, int n_quad, const double *corners, const double *normals,
                                      const int *is_edge, dbie_geom **out) {
    return guard([&] {
        if (!out) fail(DBIE_BAD_ARG, "NULL output");
        std::unique_ptr<Geometry> g(make_from_panels(P, n_quad, corners, normals, is_edge));
        auto *h = new dbie_geom;
        h->g = std::move(*g);
        *out = h;
    });
}

dbie_status dbie_geometry_from_panel_nodes(int64_t P, int n_quad, const double *corners, const double *normals,
                                           const int *is_edge, const double *points, const double *weights,
                                           const double *gl_xs, const double *gl_ws, const double *bary_weights,
                                           dbie_geom **out) {
    return guard([&] {
        if (!out) fail(DBIE_BAD_ARG, "NULL output");
        std::unique_ptr<Geometry> g(
            make_from_panel_nodes(P, n_quad, corners, normals, is_edge, points, weights, gl_xs, gl_ws, bary_weights));
        auto *h = new dbie_geom;
        h->g = std::move(*g);
        *out = h;
    });
}

dbie_status dbie_geometry_multi_box(int n_quad, double l_ec, int n_boxes, const double *boxes, const double *epses,
                                    double eps_out, double alpha, dbie_geom **out) {
    return guard([&] {
        if (!out) fail(DBIE_BAD_ARG, "NULL output");
        std::unique_ptr<dbie_geom> h(new dbie_geom);
        std::unique_ptr<Geometry> g(
            make_multi_box(n_quad, l_ec, n_boxes, boxes, epses, eps_out, alpha, h->eps_in, h->eps_out));
        h->g = std::move(*g);
        *out = h.release();
    });
}

dbie_status dbie_geom_eps(const dbie_geom *gh, double *eps_in, double *eps_out) {
    return guard([&] {
        if (!gh) fail(DBIE_BAD_ARG, "geometry handle is NULL");
        if (gh->eps_in.size() != (size_t)gh->g.P) fail(DBIE_BAD_ARG, "this geometry carries no permittivities");
        if (eps_in) std::memcpy(eps_in, gh->eps_in.data(), sizeof(double) * gh->g.P);
        if (eps_out) std::memcpy(eps_out, gh->eps_out.data(), sizeof(double) * gh->g.P);
    });
}

dbie_status dbie_geom_sizes(const dbie_geom *g, int64_t *P, int64_t *N, int *n_quad) {
    return guard([&] {
        if (!g) fail(DBIE_BAD_ARG, "geometry handle is NULL");
        if (P) *P = g->g.P;
        if (N) *N = g->g.N;
        if (n_quad) *n_quad = g->g.nq;
    });
}

dbie_status dbie_geom_nodes(const dbie_geom *gh, double *out) {
    return guard([&] {
        if (!gh || !out) fail(DBIE_BAD_ARG, "NULL argument");
        const Geometry &g = gh->g;
        const int64_t N = g.N;
        const int npp = g.npp();
        std::memcpy(out, g.x.data(), sizeof(double) * N);
        std::memcpy(out + N, g.y.data(), sizeof(double) * N);
        std::memcpy(out + 2 * N, g.z.data(), sizeof(double) * N);
        for (int64_t i = 0; i < N; ++i)
            for (int k = 0; k < 3; ++k) out[(3 + k) * N + i] = g.normal[3 * (i / npp) + k];
        std::memcpy(out + 6 * N, g.w.data(), sizeof(double) * N);
    });
}

dbie_status dbie_geom_panels(const dbie_geom *gh, double *corners, double *normals, int *is_edge) {
    return guard([&] {
        if (!gh) fail(DBIE_BAD_ARG, "NULL argument");
        const Geometry &g = gh->g;
        if (corners) std::memcpy(corners, g.corners.data(), sizeof(double) * 12 * g.P);
        if (normals) std::memcpy(normals, g.normal.data(), sizeof(double) * 3 * g.P);
        if (is_edge) std::memcpy(is_edge, g.is_edge.data(), sizeof(int) * g.P);
    });
}

void dbie_geom_destroy(dbie_geom *g) { delete g; }

dbie_status dbie_neighbor_list(const dbie_geom *gh, int max_order, double atol, int correct_edges, int64_t *n_up,
                               int64_t *n_adp, int *up_src, int *up_trg, int *up_order, int *adp_src, int *adp_trg) {
    return guard([&] {
        if (!gh || !n_up || !n_adp) fail(DBIE_BAD_ARG, "NULL argument");
        if (!(atol > 0 && atol < 1)) fail(DBIE_BAD_ARG, "atol must be in (0, 1)");
        if (max_order < 1 || max_order > 16) fail(DBIE_BAD_ARG, "max_order must be in 1..16");
        const PairList pl = build_pair_list(gh->g, max_order, atol, correct_edges != 0, 0);
        *n_up = (int64_t)pl.up_src.size();
        *n_adp = (int64_t)pl.adp_src.size();
        if (up_src) std::copy(pl.up_src.begin(), pl.up_src.end(), up_src);
        if (up_trg) std::copy(pl.up_trg.begin(), pl.up_trg.end(), up_trg);
        if (up_order) std::copy(pl.up_order.begin(), pl.up_order.end(), up_order);
        if (adp_src) std::copy(pl.adp_src.begin(), pl.adp_src.end(), adp_src);
        if (adp_trg) std::copy(pl.adp_trg.begin(), pl.adp_trg.end(), adp_trg);
    });
}

dbie_status dbie_operator_create(const dbie_geom *gh, const double *eps_in, const double *eps_out,
                                 const dbie_near_opts *opts, dbie_op **out) {
    return guard([&] {
        if (!gh || !eps_in || !eps_out || !out) fail(DBIE_BAD_ARG, "NULL argument");
        ensure_device();
        dbie_near_opts o;
        if (opts) o = *opts;
        else dbie_near_opts_default(&o);
        if (o.max_order < 1 || o.max_order > 16) fail(DBIE_BAD_ARG, "max_order must be in 1..16");
        if (!(o.up_tol > 0 && o.up_tol < 1)) fail(DBIE_BAD_ARG, "up_tol must be in (0, 1)");
        std::unique_ptr<dbie_op> op(new dbie_op);
        op->g = gh->g;
        const Geometry &g = op->g;
        DBIE_HIP(hipStreamCreateWithFlags(&op->stream, hipStreamNonBlocking));
        {
            const char *e = std::getenv("DBIE_GRAPHS");
            op->graphs = !(e && e[0] == '0');
            const char *f = std::getenv("DBIE_STORED_FAR");
            op->stored_mode = (f && f[0] == '0') ? 0 : 1;
        }
        hipStream_t st = op->stream;
        const int64_t N = g.N;
        // node streams
        std::vector<double> nrm[3];
        for (int k = 0; k < 3; ++k) nrm[k].resize(N);
        for (int64_t i = 0; i < N; ++i)
            for (int k = 0; k < 3; ++k) nrm[k][i] = g.normal[3 * (i / g.npp()) + k];
        upload(op->x, g.x.data(), N, st);
        upload(op->y, g.y.data(), N, st);
        upload(op->z, g.z.data(), N, st);
        upload(op->nx, nrm[0].data(), N, st);
        upload(op->ny, nrm[1].data(), N, st);
        upload(op->nz, nrm[2].data(), N, st);
        upload(op->w, g.w.data(), N, st);
        auto f32 = [&](const std::vector<double> &v) {
            std::vector<float> r(v.size());
            for (size_t i = 0; i < v.size(); ++i) r[i] = (float)v[i];
            return r;
        };
        const std::vector<float> fx = f32(g.x), fy = f32(g.y), fz = f32(g.z), fnx = f32(nrm[0]), fny = f32(nrm[1]),
                                 fnz = f32(nrm[2]), fw = f32(g.w);
        upload(op->xf, fx.data(), N, st);
        upload(op->yf, fy.data(), N, st);
        upload(op->zf, fz.data(), N, st);
        upload(op->nxf, fnx.data(), N, st);
        upload(op->nyf, fny.data(), N, st);
        upload(op->nzf, fnz.data(), N, st);
        upload(op->wf, fw.data(), N, st);
        const std::vector<int> tp = tile_planes(g);
        upload(op->plane, tp.data(), tp.size(), st);
        op->tile_plane_h = tp;
        // near-field corrections
        BlockSparse bs;
        bs.bd = g.npp();
        using Clk = std::chrono::steady_clock;
        auto ms = [](Clk::time_point a, Clk::time_point b) {
            return std::chrono::duration<double, std::milli>(b - a).count();
        };
        const auto t0 = Clk::now();
        auto t2 = t0;
        const int npp = g.npp(), bd = bs.bd;
        const bool mfma_order = bd == 16;
        bool device_built = false;
        if (o.build_corrections) {
            double tb[3] = {0, 0, 0};
            device_built = build_correction_blocks(op.get(), o, 0, bs, &op->bsr_v64, &op->bsr_v32, &op->bsr_vup32, tb);
            op->setup_ms[0] = tb[0];
            op->setup_ms[1] = tb[1];
            op->setup_ms[3] = tb[2];
            t2 = Clk::now();
        } else {
            bs.row_ptr.assign(g.P + 1, 0);
        }
        op->nopts = o;
        op->adaptive_hits = bs.adaptive_hits;
        // panel-pair mask of the near list (the split-FP16 far field leaves these pairs out)
        const int mwords = (int)((g.P + 31) / 32) + 2;
        std::vector<uint32_t> mask((size_t)g.P * mwords, 0u);
        for (int64_t p = 0; p < g.P; ++p)
            for (int b = bs.row_ptr[p]; b < bs.row_ptr[p + 1]; ++b)
                mask[(size_t)p * mwords + bs.col[b] / 32] |= 1u << (bs.col[b] % 32);
        const auto t3 = Clk::now();
        op->setup_ms[2] = ms(t2, t3);
        upload(op->bsr_row, bs.row_ptr.data(), bs.row_ptr.size(), st);
        upload(op->bsr_col, bs.col.data(), bs.col.size(), st);
        upload(op->nmask, mask.data(), mask.size(), st);
        if (!device_built) {
            // host-built blocks: full near interaction (correction + direct Dt term), storage orders
            std::vector<double> vup(bs.val.size());
            parallel_for(g.P, o.n_threads, [&](int64_t p, int) {
                for (int b = bs.row_ptr[p]; b < bs.row_ptr[p + 1]; ++b) {
                    const int64_t q = bs.col[b];
                    for (int t = 0; t < bd; ++t) {
                        const int64_t i = p * npp + t;
                        const double *n = &g.normal[3 * p];
                        for (int c = 0; c < bd; ++c) {
                            const int64_t j = q * npp + c;
                            const double dx = g.x[i] - g.x[j], dy = g.y[i] - g.y[j], dz = g.z[i] - g.z[j];
                            const double r2 = dx * dx + dy * dy + dz * dz;
                            double kd = 0.0;
                            if (r2 > 0.0) {
                                const double ir = 1.0 / std::sqrt(r2);
                                kd = (n[0] * dx + n[1] * dy + n[2] * dz) * (ir * ir * ir) / kFourPi * g.w[j];
                            }
                            const size_t at = (size_t)b * bd * bd + (size_t)t * bd + c;
                            vup[at] = bs.val[at] + kd;
                        }
                    }
                }
            });
            if (mfma_order) {  // permute each block into the kernels' operand orders
                parallel_for(bs.nblocks(), o.n_threads, [&](int64_t b, int) {
                    double tmpb[256];
                    double *blk = &bs.val[(size_t)b * 256];
                    for (int t = 0; t < 16; ++t)
                        for (int c = 0; c < 16; ++c) tmpb[bsr_storage_index(16, true, t, c)] = blk[t * 16 + c];
                    std::copy(tmpb, tmpb + 256, blk);
                    double *bu = &vup[(size_t)b * 256];
                    for (int t = 0; t < 16; ++t)
                        for (int c = 0; c < 16; ++c) tmpb[bsr_lane_major_index(t, c)] = bu[t * 16 + c];
                    std::copy(tmpb, tmpb + 256, bu);
                });
            }
            upload(op->bsr_v64, bs.val.data(), bs.val.size(), st);
            const std::vector<float> v32 = f32(bs.val);
            upload(op->bsr_v32, v32.data(), v32.size(), st);
            const std::vector<float> u32 = f32(vup);
            upload(op->bsr_vup32, u32.data(), u32.size(), st);
        }
        DBIE_HIP(hipStreamSynchronize(st));
        compute_diag(op.get(), eps_in, eps_out);
        DevGeom &d = op->dg;
        d.N = N;
        d.x = op->x.as<double>(); d.y = op->y.as<double>(); d.z = op->z.as<double>();
        d.nx = op->nx.as<double>(); d.ny = op->ny.as<double>(); d.nz = op->nz.as<double>();
        d.w = op->w.as<double>();
        d.xf = op->xf.as<float>(); d.yf = op->yf.as<float>(); d.zf = op->zf.as<float>();
        d.nxf = op->nxf.as<float>(); d.nyf = op->nyf.as<float>(); d.nzf = op->nzf.as<float>();
        d.wf = op->wf.as<float>();
        d.tile_plane = op->plane.as<int>();
        d.npp = npp;
        d.mask_words = mwords;
        d.near_mask = bs.nblocks() > 0 ? op->nmask.as<uint32_t>() : nullptr;
        DevBSR &c = op->bsr;
        c.P = (int)g.P;
        c.bd = bs.bd;
        c.nb = bs.nblocks();
        c.row_ptr = op->bsr_row.as<int>();
        c.col = op->bsr_col.as<int>();
        c.val64 = op->bsr_v64.as<double>();
        c.val32 = op->bsr_v32.as<float>();
        c.valup32 = bs.nblocks() > 0 ? op->bsr_vup32.as<float>() : nullptr;
        c.mfma_order = mfma_order;
        if (mfma_order && c.valup32) {
            op->bsr_vh.ensure(sizeof(uint16_t) * 512 * (size_t)c.nb);
            op->bsr_rexp.ensure(sizeof(int) * (size_t)c.P);
            near_h3_pack(st, c, op->w.as<double>(), op->bsr_vh.p, op->bsr_rexp.as<int>());
            DBIE_HIP(hipGetLastError());
            c.valh = op->bsr_vh.p;
            c.rexp = op->bsr_rexp.as<int>();
            // super-rows: kNearSuperRows consecutive block rows (neighbouring panels of a face) with the
            // sorted union of their block columns
            constexpr int RT = kNearSuperRows;
            const int Q = (int)((g.P + RT - 1) / RT);
            std::vector<int> sp(Q + 1, 0), scol, sblk;
            std::vector<std::pair<int, int>> ent;  // (block column, block index)
            for (int q = 0; q < Q; ++q) {
                ent.clear();
                for (int r = 0; r < RT && q * RT + r < g.P; ++r) {
                    const int p = q * RT + r;
                    for (int b = bs.row_ptr[p]; b < bs.row_ptr[p + 1]; ++b) ent.push_back({bs.col[b], b});
                }
                std::sort(ent.begin(), ent.end());
                for (size_t e = 0; e < ent.size();) {
                    const int cl = ent[e].first;
                    scol.push_back(cl);
                    const size_t base = sblk.size();
                    sblk.resize(base + RT, -1);
                    for (; e < ent.size() && ent[e].first == cl; ++e) {
                        const int b = ent[e].second;
                        const int r = (int)(std::upper_bound(bs.row_ptr.begin(), bs.row_ptr.end(), b) - bs.row_ptr.begin()) - 1 - q * RT;
                        sblk[base + r] = b;
                    }
                }
                sp[q + 1] = (int)scol.size();
            }
            upload(op->sr_ptr, sp.data(), sp.size(), st);
            upload(op->sr_col, scol.data(), scol.size(), st);
            upload(op->sr_blk, sblk.data(), sblk.size(), st);
            c.sr_ptr = op->sr_ptr.as<int>();
            c.sr_col = op->sr_col.as<int>();
            c.sr_blk = op->sr_blk.as<int>();
            c.sr_q = Q;
        }
        DBIE_HIP(hipStreamSynchronize(st));
        op->setup_ms[3] += ms(t3, Clk::now());
        *out = op.release();
    });
}

void dbie_operator_destroy(dbie_op *op) { delete op; }

dbie_status dbie_op_sizes(const dbie_op *op, int64_t *N, int64_t *n_blocks, int *block_dim) {
    return guard([&] {
        check_op(op);
        if (N) *N = op->N();
        if (n_blocks) *n_blocks = op->bsr.nb;
        if (block_dim) *block_dim = op->bsr.bd;
    });
}

dbie_status dbie_op_set_eps(dbie_op *op, const double *eps_in, const double *eps_out) {
    return guard([&] {
        check_op(op);
        if (!eps_in || !eps_out) fail(DBIE_BAD_ARG, "NULL eps");
        compute_diag(op, eps_in, eps_out);
        DBIE_HIP(hipStreamSynchronize(op->stream));
    });
}

static void export_corrections(dbie_op *op, dbie_kernel kernel, int *src, int *trg, double *blocks) {
    check_op(op);
    if (kernel != DBIE_OP_DT && kernel != DBIE_OP_D) fail(DBIE_BAD_ARG, "corrections exist for DBIE_OP_DT and DBIE_OP_D");
    const DevBSR &c = op->bsr;
    const double *v64 = c.val64;
    if (kernel == DBIE_OP_D) {
        ensure_d_corrections(op);
        v64 = op->bsr_d64.as<double>();
    }
    std::vector<int> rp(c.P + 1), col(c.nb);
    DBIE_HIP(hipMemcpy(rp.data(), c.row_ptr, sizeof(int) * (c.P + 1), hipMemcpyDeviceToHost));
    if (c.nb) DBIE_HIP(hipMemcpy(col.data(), c.col, sizeof(int) * c.nb, hipMemcpyDeviceToHost));
    for (int p = 0; p < c.P; ++p)
        for (int b = rp[p]; b < rp[p + 1]; ++b) {
            if (src) src[b] = col[b];
            if (trg) trg[b] = p;
        }
    if (blocks && c.nb) {
        const size_t bb = (size_t)c.bd * c.bd;
        std::vector<double> raw(c.nb * bb);
        DBIE_HIP(hipMemcpy(raw.data(), v64, sizeof(double) * raw.size(), hipMemcpyDeviceToHost));
        for (int64_t b = 0; b < c.nb; ++b)
            for (int t = 0; t < c.bd; ++t)
                for (int cc = 0; cc < c.bd; ++cc)
                    blocks[b * bb + (size_t)t * c.bd + cc] = raw[b * bb + bsr_storage_index(c.bd, c.mfma_order, t, cc)];
    }
}

dbie_status dbie_op_corrections(const dbie_op *op, int *src, int *trg, double *blocks) {
    return guard([&] { export_corrections(const_cast<dbie_op *>(op), DBIE_OP_DT, src, trg, blocks); });
}

dbie_status dbie_op_corrections_kernel(dbie_op *op, dbie_kernel kernel, int *src, int *trg, double *blocks) {
    return guard([&] { export_corrections(op, kernel, src, trg, blocks); });
}

dbie_status dbie_op_setup_ms(const dbie_op *op, double *t) {
    return guard([&] {
        check_op(op);
        if (!t) fail(DBIE_BAD_ARG, "NULL output");
        for (int i = 0; i < 4; ++i) t[i] = op->setup_ms[i];
    });
}

dbie_status dbie_op_device_bytes(const dbie_op *op, int64_t *bytes) {
    return guard([&] {
        check_op(op);
        if (!bytes) fail(DBIE_BAD_ARG, "NULL output");
        int64_t b = 0;
        for (const DBuf *d : {&op->x, &op->y, &op->z, &op->nx, &op->ny, &op->nz, &op->w, &op->xf, &op->yf, &op->zf,
                              &op->nxf, &op->nyf, &op->nzf, &op->wf, &op->plane, &op->diag, &op->bsr_row,
                              &op->bsr_col, &op->bsr_v64, &op->bsr_v32, &op->bsr_vup32, &op->bsr_vh, &op->bsr_rexp, &op->sr_ptr,
                              &op->sr_col, &op->sr_blk, &op->bsr_d64, &op->bsr_full64,
                              &op->nmask, &op->part,
                              &op->xin, &op->yout, &op->tmp, &op->mscratch, &op->xt32, &op->gst, &op->gst_off})
            b += (int64_t)d->bytes;
        // the Krylov workspaces the operator keeps for later solves
        const GmresWork &w = op->gw;
        for (const DBuf *d : {&w.V, &w.W, &w.Z, &w.U, &w.partial, &w.hbuf, &w.nbuf, &w.V1, &w.abuf, &w.cbuf, &w.om,
                              &w.n1, &w.H, &w.G, &w.tau, &w.Y, &w.gpart, &w.R, &w.info, &w.PV, &w.PT})
            b += (int64_t)d->bytes;
        *bytes = b;
    });
}

void *dbie_op_stream(dbie_op *op) { return op ? (void *)op->stream : nullptr; }

dbie_status dbie_op_residual_history(const dbie_op *op, int cap, int *n, double *kind_value) {
    return guard([&] {
        check_op(op);
        if (!n) fail(DBIE_BAD_ARG, "NULL output");
        const int total = (int)(op->hist.size() / 2);
        *n = total;
        if (kind_value && cap > 0) std::copy(op->hist.begin(), op->hist.begin() + 2 * std::min(cap, total), kind_value);
    });
}

dbie_status dbie_op_synchronize(dbie_op *op) {
    return guard([&] {
        check_op(op);
        DBIE_HIP(hipStreamSynchronize(op->stream));
    });
}

dbie_status dbie_autotune(int enable, int clear) {
    return guard([&] {
        if (enable < -1 || enable > 1) fail(DBIE_BAD_ARG, "enable must be -1, 0 or 1");
        if (enable >= 0) {
            autotune_on();  // read the environment first so that it cannot override this call
            g_autotune.store(enable);
        }
        if (clear) {
            std::lock_guard<std::mutex> lk(g_tune_mu);
            for (PendingTiming &p : g_pending) {  // events of launches still in flight are released
                (void)hipEventDestroy(p.e0);      // when those complete
                (void)hipEventDestroy(p.e1);
            }
            g_pending.clear();
            g_tune.clear();
        }
    });
}

dbie_status dbie_autotune_choice(const dbie_op *op, int K, dbie_precision precision, int *splits, int *variant) {
    return guard([&] {
        check_op(op);
        if (K < 1) fail(DBIE_BAD_ARG, "K must be >= 1");
        check_precision(precision);
        FarTune t;
        t.splits = 0;
        {
            std::lock_guard<std::mutex> lk(g_tune_mu);
            poll_timings_locked();
            auto it = g_tune.find(tune_key(op, K, precision, excludes_near(op, K, precision)));
            if (it != g_tune.end() && it->second.done) t = it->second.best;
        }
        if (splits) *splits = t.splits;
        if (variant) *variant = t.variant;
    });
}

dbie_status dbie_autotune_set(const dbie_op *op, int K, dbie_precision precision, int splits, int variant) {
    return guard([&] {
        check_op(op);
        if (K < 1 || splits < 1) fail(DBIE_BAD_ARG, "K and splits must be >= 1");
        if (variant < -1 || variant > 4) fail(DBIE_BAD_ARG, "variant must be -1 ... 4");
        check_precision(precision);
        TuneState ts;
        ts.cand = {FarTune{splits, variant}};
        ts.ms = {0.f};
        ts.nt = {(int)kTuneRounds};
        ts.next = kTuneRounds;
        ts.done = true;
        ts.best = ts.cand[0];
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_tune[tune_key(op, K, precision, excludes_near(op, K, precision))] = ts;
    });
}

dbie_status dbie_op_graphs(dbie_op *op, int enable, int64_t *captures, int64_t *replays) {
    return guard([&] {
        check_op(op);
        if (enable >= 0) {
            DBIE_HIP(hipStreamSynchronize(op->stream));
            op->graphs = enable != 0;  // captured graphs stay cached while switched off
        }
        if (captures) *captures = op->graph_captures;
        if (replays) *replays = op->graph_replays;
    });
}

dbie_status dbie_op_stored_far(dbie_op *op, int mode, int64_t *bytes) {
    return guard([&] {
        check_op(op);
        if (mode < -1 || mode > 1) fail(DBIE_BAD_ARG, "mode must be -1, 0 or 1");
        if (mode >= 0 && mode != op->stored_mode) {
            DBIE_HIP(hipStreamSynchronize(op->stream));
            op->stored_mode = mode;
            if (mode == 0) {  // graphs that read the matrix retire when it goes (their liveness check)
                op->gst.release();
                op->gst_off.release();
            }
            op->gst_state = 0;
        }
        if (bytes) *bytes = op->gst_state == 1 ? (int64_t)op->gst.bytes : 0;
    });
}

dbie_status dbie_op_profile(dbie_op *op, int enable) {
    return guard([&] {
        check_op(op);
        op->profiling = enable != 0;
    });
}

dbie_status dbie_op_profile_read(dbie_op *op, int kernel, int64_t *launches, double *total_ms) {
    return guard([&] {
        check_op(op);
        if (kernel < 0 || kernel > 2) fail(DBIE_BAD_ARG, "kernel id must be 0, 1 or 2");
        DBIE_HIP(hipStreamSynchronize(op->stream));
        double ms = 0.0;
        for (auto &e : op->prof[kernel]) {
            float t = 0.f;
            DBIE_HIP(hipEventElapsedTime(&t, e.first, e.second));
            ms += t;
            (void)hipEventDestroy(e.first);
            (void)hipEventDestroy(e.second);
        }
        if (launches) *launches = (int64_t)op->prof[kernel].size();
        if (total_ms) *total_ms = ms;
        op->prof[kernel].clear();
    });
}

dbie_status dbie_apply_device(dbie_op *op, int K, const double *dX, int64_t ldx, double *dY, int64_t ldy,
                              dbie_precision precision) {
    return guard([&] {
        check_op(op);
        if (K < 1 || !dX || !dY) fail(DBIE_BAD_ARG, "K must be >= 1 and buffers non-NULL");
        if (ldx < op->N() || ldy < op->N()) fail(DBIE_DIM_MISMATCH, "leading dimension smaller than N");
        check_precision(precision);
        apply_dev(op, K, dX, ldx, dY, ldy, precision);
    });
}

dbie_status dbie_apply(dbie_op *op, int K, const double *X, int64_t ldx, double *Y, int64_t ldy,
                       dbie_precision precision) {
    return guard([&] {
        check_op(op);
        const int64_t N = op->N();
        if (K < 1 || !X || !Y) fail(DBIE_BAD_ARG, "K must be >= 1 and buffers non-NULL");
        if (ldx < N || ldy < N) fail(DBIE_DIM_MISMATCH, "leading dimension smaller than N");
        check_precision(precision);
        hipStream_t st = op->stream;
        op->xin.ensure(sizeof(double) * N * K);
        op->yout.ensure(sizeof(double) * N * K);
        DBIE_HIP(hipMemcpy2DAsync(op->xin.p, sizeof(double) * N, X, sizeof(double) * ldx, sizeof(double) * N, K,
                                  hipMemcpyHostToDevice, st));
        apply_dev(op, K, op->xin.as<double>(), N, op->yout.as<double>(), N, precision);
        DBIE_HIP(hipMemcpy2DAsync(Y, sizeof(double) * ldy, op->yout.p, sizeof(double) * N, sizeof(double) * N, K,
                                  hipMemcpyDeviceToHost, st));
        DBIE_HIP(hipStreamSynchronize(st));
    });
}

// Plain boundary operator (corrected = false) or operator with its near corrections (corrected = true),
// no diagonal: far-field partial sums reduced by the combine with zero diag and the given blocks.
static void apply_kernel_host(dbie_op *op, dbie_kernel kernel, bool corrected, int K, const double *X, int64_t ldx,
                              double *Y, int64_t ldy) {
    check_op(op);
    const int64_t N = op->N();
    if (K < 1 || !X || !Y) fail(DBIE_BAD_ARG, "K must be >= 1 and buffers non-NULL");
    if (ldx < N || ldy < N) fail(DBIE_DIM_MISMATCH, "leading dimension smaller than N");
    if (kernel < DBIE_OP_DT || kernel > DBIE_OP_D) fail(DBIE_BAD_ARG, "unknown kernel");
    if (corrected && kernel == DBIE_OP_S) fail(DBIE_BAD_ARG, "near corrections exist for DBIE_OP_DT and DBIE_OP_D");
    hipStream_t st = op->stream;
    DevBSR C = op->bsr;
    DBuf zrow(true);
    if (corrected) {
        if (!op->nopts.build_corrections) fail(DBIE_BAD_ARG, "the operator was created without near corrections");
        if (kernel == DBIE_OP_D) {
            ensure_d_corrections(op);
            C.val64 = op->bsr_d64.as<double>();
        }
    } else {
        std::vector<int> zeros(op->g.P + 1, 0);
        upload(zrow, zeros.data(), zeros.size(), st);
        C.row_ptr = zrow.as<int>();
    }
    op->xin.ensure(sizeof(double) * N * K);
    op->yout.ensure(sizeof(double) * N * K);
    const int64_t need = far_field_part_elems(N, K);
    op->part.ensure(need * sizeof(double));
    DBIE_HIP(hipMemcpy2DAsync(op->xin.p, sizeof(double) * N, X, sizeof(double) * ldx, sizeof(double) * N, K,
                              hipMemcpyHostToDevice, st));
    const int S = far_field_f64(st, op->dg, (FarMode)kernel, K, op->xin.as<double>(), N, op->part.as<double>(), need);
    op->tmp.ensure(sizeof(double) * N);
    DBIE_HIP(hipMemsetAsync(op->tmp.p, 0, sizeof(double) * N, st));
    near_combine(st, C, NEAR_C64, S, op->part.as<double>(), N, K, op->xin.as<double>(), N, op->tmp.as<double>(),
                 op->yout.as<double>(), N);
    DBIE_HIP(hipGetLastError());
    DBIE_HIP(hipMemcpy2DAsync(Y, sizeof(double) * ldy, op->yout.p, sizeof(double) * N, sizeof(double) * N, K,
                              hipMemcpyDeviceToHost, st));
    DBIE_HIP(hipStreamSynchronize(st));
}

dbie_status dbie_apply_kernel(dbie_op *op, dbie_kernel kernel, int K, const double *X, int64_t ldx, double *Y,
                              int64_t ldy) {
    return guard([&] { apply_kernel_host(op, kernel, false, K, X, ldx, Y, ldy); });
}

dbie_status dbie_apply_kernel_corrected(dbie_op *op, dbie_kernel kernel, int K, const double *X, int64_t ldx,
                                        double *Y, int64_t ldy) {
    return guard([&] { apply_kernel_host(op, kernel, true, K, X, ldx, Y, ldy); });
}

dbie_status dbie_lhs_dense(dbie_op *op, double *A, int64_t lda) {
    return guard([&] {
        check_op(op);
        const int64_t N = op->N();
        if (!A) fail(DBIE_BAD_ARG, "NULL output");
        if (lda < N) fail(DBIE_DIM_MISMATCH, "lda < N");
        if (N > 65535) fail(DBIE_BAD_ARG, "dense assembly is limited to N <= 65535");
        DBuf dA;
        dA.ensure(sizeof(double) * N * N);
        dense_lhs(op->stream, op->dg, op->diag.as<double>(), dA.as<double>(), N);
        DBIE_HIP(hipGetLastError());
        DBIE_HIP(hipMemcpy2DAsync(A, sizeof(double) * lda, dA.p, sizeof(double) * N, sizeof(double) * N, N,
                                  hipMemcpyDeviceToHost, op->stream));
        DBIE_HIP(hipStreamSynchronize(op->stream));
    });
}

dbie_status dbie_solve_lu(int64_t n, const double *A, int64_t lda, int nrhs, double *B, int64_t ldb) {
    return guard([&] {
        ensure_device();
        if (n < 1 || nrhs < 1 || !A || !B) fail(DBIE_BAD_ARG, "n, nrhs >= 1 and non-NULL buffers required");
        if (lda < n || ldb < n) fail(DBIE_DIM_MISMATCH, "leading dimension smaller than n");
        if (n > 2147483647) fail(DBIE_BAD_ARG, "n too large");
        rocblas_handle h;
        if (rocblas_create_handle(&h) != rocblas_status_success) fail(DBIE_HIP_ERROR, "rocblas_create_handle failed");
        std::unique_ptr<_rocblas_handle, decltype(&rocblas_destroy_handle)> hg(h, rocblas_destroy_handle);
        DBuf dA, dB, dP, dI;
        dA.ensure(sizeof(double) * n * n);
        dB.ensure(sizeof(double) * n * nrhs);
        dP.ensure(sizeof(rocblas_int) * n);
        dI.ensure(sizeof(rocblas_int));
        DBIE_HIP(hipMemcpy2D(dA.p, sizeof(double) * n, A, sizeof(double) * lda, sizeof(double) * n, n,
                             hipMemcpyHostToDevice));
        DBIE_HIP(hipMemcpy2D(dB.p, sizeof(double) * n, B, sizeof(double) * ldb, sizeof(double) * n, nrhs,
                             hipMemcpyHostToDevice));
        if (rocsolver_dgetrf(h, (rocblas_int)n, (rocblas_int)n, dA.as<double>(), (rocblas_int)n, dP.as<rocblas_int>(),
                             dI.as<rocblas_int>()) != rocblas_status_success)
            fail(DBIE_HIP_ERROR, "rocsolver_dgetrf failed");
        rocblas_int info = 0;
        DBIE_HIP(hipMemcpy(&info, dI.p, sizeof(info), hipMemcpyDeviceToHost));
        if (info != 0) fail(DBIE_BAD_ARG, "matrix is singular (zero pivot at " + std::to_string(info) + ")");
        if (rocsolver_dgetrs(h, rocblas_operation_none, (rocblas_int)n, nrhs, dA.as<double>(), (rocblas_int)n,
                             dP.as<rocblas_int>(), dB.as<double>(), (rocblas_int)n) != rocblas_status_success)
            fail(DBIE_HIP_ERROR, "rocsolver_dgetrs failed");
        DBIE_HIP(hipMemcpy2D(B, sizeof(double) * ldb, dB.p, sizeof(double) * n, sizeof(double) * n, nrhs,
                             hipMemcpyDeviceToHost));
    });
}

dbie_status dbie_rhs_points(dbie_op *op, int K, const double *src_xyz, const double *q, double eps_src, double *F,
                            int64_t ldf) {
    return guard([&] {
        check_op(op);
        const int64_t N = op->N();
        if (K < 1 || !src_xyz || !q || !F) fail(DBIE_BAD_ARG, "K >= 1 and non-NULL buffers required");
        if (ldf < N) fail(DBIE_DIM_MISMATCH, "ldf < N");
        if (eps_src == 0.0) fail(DBIE_BAD_ARG, "eps_src must be nonzero");
        hipStream_t st = op->stream;
        DBuf ds(true), dq(true);
        upload(ds, src_xyz, 3 * (size_t)K, st);
        upload(dq, q, K, st);
        op->yout.ensure(sizeof(double) * N * K);
        rhs_points(st, op->dg, K, ds.as<double>(), dq.as<double>(), eps_src, op->yout.as<double>(), N);
        DBIE_HIP(hipGetLastError());
        DBIE_HIP(hipMemcpy2DAsync(F, sizeof(double) * ldf, op->yout.p, sizeof(double) * N, sizeof(double) * N, K,
                                  hipMemcpyDeviceToHost, st));
        DBIE_HIP(hipStreamSynchronize(st));
    });
}

dbie_status dbie_rhs_volume(dbie_op *op, int K, const int64_t *col_ptr, const double *xyz, const double *w,
                            const double *rho, int screening, double eps_src, double *F, int64_t ldf) {
    return guard([&] {
        check_op(op);
        if (!screening && eps_src == 0.0) fail(DBIE_BAD_ARG, "eps_src must be nonzero");
        HostBoxes hb;
        if (screening) hb = interface_box(op);
        rhs_volume_host(op, K, col_ptr, xyz, w, rho, screening ? &hb : nullptr, screening ? 1.0 : eps_src, F, ldf);
    });
}

dbie_status dbie_rhs_volume_boxes(dbie_op *op, int K, const int64_t *col_ptr, const double *xyz, const double *w,
                                  const double *rho, int n_boxes, const double *boxes, const double *epses,
                                  double eps_out, int mode, double bandwidth, double *F, int64_t ldf) {
    return guard([&] {
        check_op(op);
        const HostBoxes hb = box_list(n_boxes, boxes, epses, eps_out, mode, bandwidth, -1.0, 0);
        rhs_volume_host(op, K, col_ptr, xyz, w, rho, &hb, 1.0, F, ldf);
    });
}

dbie_status dbie_screen_boxes(int64_t S, const double *xyz, const double *rho, int n_boxes, const double *boxes,
                              const double *epses, double eps_out, int mode, double bandwidth, double tol, int form,
                              double *rho_out) {
    return guard([&] {
        ensure_device();
        if (S < 0 || (S > 0 && (!xyz || !rho || !rho_out))) fail(DBIE_BAD_ARG, "S >= 0 and source arrays required");
        const HostBoxes hb = box_list(n_boxes, boxes, epses, eps_out, mode, bandwidth, tol, form);
        screen_host(nullptr, S, xyz, rho, hb, rho_out);
    });
}

dbie_status dbie_screen_interface(dbie_op *op, int64_t S, const double *xyz, const double *rho, int mode,
                                  double bandwidth, double tol, double *rho_out) {
    return guard([&] {
        check_op(op);
        if (S < 0 || (S > 0 && (!xyz || !rho || !rho_out))) fail(DBIE_BAD_ARG, "S >= 0 and source arrays required");
        HostBoxes hb = interface_box(op);
        check_screen_mode(mode, bandwidth);
        hb.sb.mode = mode;
        hb.sb.bandwidth = bandwidth;
        if (tol >= 0) hb.sb.tol = tol;
        screen_host(op->stream, S, xyz, rho, hb, rho_out);
    });
}

dbie_status dbie_pottrg(dbie_op *op, int64_t M, const double *trg, int K, const double *Sigma, int64_t lds, double *U,
                        int64_t ldu) {
    return guard([&] {
        check_op(op);
        const int64_t N = op->N();
        if (M < 1 || K < 1 || !trg || !Sigma || !U) fail(DBIE_BAD_ARG, "M, K >= 1 and non-NULL buffers required");
        if (lds < N || ldu < M) fail(DBIE_DIM_MISMATCH, "leading dimension too small");
        hipStream_t st = op->stream;
        DBuf dt, dS, dU;
        upload(dt, trg, 3 * (size_t)M, st);
        dS.ensure(sizeof(double) * N * K);
        dU.ensure(sizeof(double) * M * K);
        DBIE_HIP(hipMemcpy2DAsync(dS.p, sizeof(double) * N, Sigma, sizeof(double) * lds, sizeof(double) * N, K,
                                  hipMemcpyHostToDevice, st));
        pottrg(st, op->dg, M, dt.as<double>(), K, dS.as<double>(), N, dU.as<double>(), M);
        DBIE_HIP(hipGetLastError());
        DBIE_HIP(hipMemcpy2DAsync(U, sizeof(double) * ldu, dU.p, sizeof(double) * M, sizeof(double) * M, K,
                                  hipMemcpyDeviceToHost, st));
        DBIE_HIP(hipStreamSynchronize(st));
    });
}

dbie_status dbie_solve_block(dbie_op *op, int K, const double *F, int64_t ldf, double *Sigma, int64_t lds,
                             const dbie_solve_opts *opts, dbie_solve_stats *stats) {
    dbie_status rc = DBIE_OK;
    const dbie_status g = guard([&] {
        check_op(op);
        const int64_t N = op->N();
        if (K < 1 || !F || !Sigma) fail(DBIE_BAD_ARG, "K >= 1 and non-NULL buffers required");
        if (ldf < N || lds < N) fail(DBIE_DIM_MISMATCH, "leading dimension smaller than N");
        hipStream_t st = op->stream;
        DBuf dF(true), dX(true);
        dF.ensure(sizeof(double) * N * K);
        dX.ensure(sizeof(double) * N * K);
        DBIE_HIP(hipMemcpy2DAsync(dF.p, sizeof(double) * N, F, sizeof(double) * ldf, sizeof(double) * N, K,
                                  hipMemcpyHostToDevice, st));
        rc = solve_dev(op, K, dF.as<double>(), dX.as<double>(), opts, stats);
        DBIE_HIP(hipMemcpy2DAsync(Sigma, sizeof(double) * lds, dX.p, sizeof(double) * N, sizeof(double) * N, K,
                                  hipMemcpyDeviceToHost, st));
        DBIE_HIP(hipStreamSynchronize(st));
    });
    if (g != DBIE_OK) return g;
    if (rc == DBIE_NOT_CONVERGED) set_error("solve did not reach the tolerance within itmax / max_refine");
    return rc;
}

dbie_status dbie_solve_fused(int B, dbie_op *const *ops, const int *K, const double *const *F, const int64_t *ldf,
                             double *const *Sigma, const int64_t *lds, const dbie_solve_opts *opts,
                             dbie_solve_stats *stats) {
    dbie_status rc = DBIE_OK;
    const dbie_status g = guard([&] { rc = solve_fused_host(B, ops, K, F, ldf, Sigma, lds, opts, stats); });
    if (g != DBIE_OK) return g;
    if (rc == DBIE_NOT_CONVERGED) set_error("solve did not reach the tolerance within itmax / max_refine");
    return rc;
}

dbie_status dbie_solve_block_device(dbie_op *op, int K, const double *dF, int64_t ldf, double *dSigma, int64_t lds,
                                    const dbie_solve_opts *opts, dbie_solve_stats *stats) {
    dbie_status rc = DBIE_OK;
    const dbie_status g = guard([&] {
        check_op(op);
        const int64_t N = op->N();
        if (K < 1 || !dF || !dSigma) fail(DBIE_BAD_ARG, "K >= 1 and non-NULL buffers required");
        if (ldf < N || lds < N) fail(DBIE_DIM_MISMATCH, "leading dimension smaller than N");
        hipStream_t st = op->stream;
        DBuf cF(true), cX(true);
        const double *F = dF;
        double *X = dSigma;
        if (ldf != N) {
            cF.ensure(sizeof(double) * N * K);
            copy_cols(st, dF, ldf, cF.as<double>(), N, N, K);
            F = cF.as<double>();
        }
        if (lds != N) {
            cX.ensure(sizeof(double) * N * K);
            X = cX.as<double>();
        }
        rc = solve_dev(op, K, F, X, opts, stats);
        if (lds != N) copy_cols(st, X, N, dSigma, lds, N, K);
        DBIE_HIP(hipStreamSynchronize(st));
    });
    if (g != DBIE_OK) return g;
    if (rc == DBIE_NOT_CONVERGED) set_error("solve did not reach the tolerance within itmax / max_refine");
    return rc;
}

dbie_status dbie_device_alloc(size_t bytes, void **dptr) {
    return guard([&] {
        ensure_device();
        if (!dptr) fail(DBIE_BAD_ARG, "NULL output");
        if (hipMalloc(dptr, std::max<size_t>(bytes, 1)) != hipSuccess) {
            (void)hipGetLastError();
            *dptr = nullptr;
            fail(DBIE_OOM, "hipMalloc of " + std::to_string(bytes) + " bytes failed");
        }
    });
}

dbie_status dbie_device_free(void *dptr) {
    return guard([&] {
        if (dptr) DBIE_HIP(hipFree(dptr));
    });
}

dbie_status dbie_memcpy_h2d(void *dst, const void *src, size_t bytes) {
    return guard([&] {
        if (bytes && (!dst || !src)) fail(DBIE_BAD_ARG, "NULL pointer");
        if (bytes) DBIE_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    });
}

dbie_status dbie_copy_block_device(int64_t N, int K, const double *dsrc, int64_t lds, double *ddst, int64_t ldd) {
    return guard([&] {
        ensure_device();
        if (N < 0 || K < 0 || ((N > 0 && K > 0) && (!dsrc || !ddst))) fail(DBIE_BAD_ARG, "bad copy arguments");
        if (lds < N || ldd < N) fail(DBIE_DIM_MISMATCH, "leading dimension smaller than N");
        if (N > 0 && K > 0) {
            copy_cols(nullptr, dsrc, lds, ddst, ldd, N, K);
            DBIE_HIP(hipGetLastError());
            DBIE_HIP(hipStreamSynchronize(nullptr));
        }
    });
}

dbie_status dbie_memcpy_d2h(void *dst, const void *src, size_t bytes) {
    return guard([&] {
        if (bytes && (!dst || !src)) fail(DBIE_BAD_ARG, "NULL pointer");
        if (bytes) DBIE_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    });
}

}  // extern "C"
