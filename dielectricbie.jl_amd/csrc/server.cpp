// server.cpp -- continuous-batching request loop for many block solves on one GPU.
//
// No counterpart in the reference (a Julia library whose callers loop over
// solve_dielectric_box3d_block, src/solver/multi_rhs.jl:270-286).  A request is a single
// dielectric box (single_dielectric_box3d, src/shape/box3d_single.jl:49-70) with K point charges
// (rhs_dielectric_box3d, src/solver/dielectric_box3d.jl:95-104); its answer is the N x K density.
//
//   arrivals --> [queue] --admission (HBM budget)--> setup workers --> [ready] --> solve slots
//                                                  (mesh, near field,   (fuse ready requests of one
//                                                   RHS, on host cores)   operator, block solve on the
//                                                                         operator's own stream)
//
// * Admission: a request enters setup only if the device bytes of the cached operators, the
//   in-flight reservations and its own estimate (operator + Krylov workspace) fit the budget;
//   idle cached operators are evicted (least recently used) to make room.  A request that does
//   not fit even alone fails with DBIE_OOM; the rest keep running.  Admission is first in,
//   first out: the head of the queue waits for memory rather than being overtaken.
// * Operator cache: keyed by geometry (box sides, n_quad, l_ec); the permittivities only set the
//   operator's diagonal, which is updated before a solve with other permittivities.  A hit skips the
//   build.
// * Continuous batching: a free slot takes the oldest ready request and every other ready request with
//   the same solve options, up to max_batch_rhs columns: those on the same operator (and permittivities)
//   become one column block, and small requests on other free operators join the same solve as
//   further members of a fused Krylov process (dbie_solve_fused: every iteration applies all members
//   at once, each on its own stream, with one host synchronisation) as long as padding to the largest
//   member wastes at most half the Krylov work.  Requests that arrive while others are on the GPU are set up on the host meanwhile and start
//   as slots free up; the slots run on separate streams, so their device work overlaps.
// * Failures (allocation, non-convergence, bad arguments) are per request and reported in its
//   info; a cancelled request is dropped from the queue or its result discarded.
#include <hip/hip_runtime.h>

#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <tuple>
#include <unordered_map>

#include "internal.hpp"
#include "runtime.hpp"

namespace {

using Clock = std::chrono::steady_clock;
double ms_since(Clock::time_point t0, Clock::time_point t1) {
    return std::chrono::duration<double, std::milli>(t1 - t0).count();
}

std::string last_error_text() {
    char buf[1024];
    dbie_last_error(buf, sizeof buf);
    return buf;
}

using GeoKey = std::tuple<double, double, double, int, double>;

struct OpEntry {
    GeoKey key;
    dbie_op *op = nullptr;
    int64_t N = 0, P = 0;
    double eps_in = 0, eps_out = 0;  // permittivities the operator's diagonal holds
    size_t bytes = 0;         // device bytes (estimate until built, then measured)
    bool building = true;
    bool failed = false;
    bool busy = false;        // a slot is solving on it
    int users = 0;            // requests bound to it (setup .. done)
    uint64_t last_use = 0;
    ~OpEntry() {
        if (op) dbie_operator_destroy(op);
    }
};

enum class St { Queued, Setup, Ready, Solving, Done, Failed, Cancelled };

struct Request {
    int64_t id = 0;
    double L[3] = {0, 0, 0};
    int nq = 4;
    double l_ec = 0.1, eps_in = 2, eps_out = 1;
    int K = 0;
    std::vector<double> pts, q;
    dbie_solve_opts sopts;
    St st = St::Queued;
    dbie_status status = DBIE_OK;
    std::string msg;
    std::shared_ptr<OpEntry> ent;
    std::vector<double> F, sigma;
    int64_t N = 0;
    dbie_solve_stats stats{};
    size_t pending_ws = 0;     // admission estimates: Krylov workspace, operator
    size_t pending_op = 0;
    size_t reserved = 0;       // workspace bytes held while in flight
    size_t op_reserved = 0;    // operator estimate held while it is being built
    bool cache_hit = false;
    bool cancel = false;
    int batch_K = 0;
    Clock::time_point t_arrive, t_setup0, t_setup1, t_solve0, t_solve1;
};

// Device bytes of an operator before it exists: node streams, plane tiles, near blocks (FP64 +
// two FP32 copies + indices, about 45 blocks of n_quad^4 per panel), and the far-field scratch.
size_t estimate_op_bytes(int64_t N, int nq) {
    const int64_t npp = (int64_t)nq * nq;
    const int64_t P = N / std::max<int64_t>(npp, 1);
    const double blocks = 50.0 * P;
    return (size_t)(N * 8 * 24 + blocks * npp * npp * 16.0 + blocks * 8 + 64.0 * 1024 * 1024);
}
// Krylov basis and the recycled basis of mixed-precision refinements (up to about 2 x 25 blocks for
// inner solves of up to 24 iterations), work vectors and the far-field partial sums for K columns; block
// GMRES (krylov = 1) also keeps its block Hessenberg ((m + 1) K x m K), the Cholesky factor of its
// normal equations ((m K)^2) and three (m + 1) K x K blocks, m the restart length (40 by default).
size_t estimate_ws_bytes(int64_t N, int K, const dbie_solve_opts &o) {
    size_t b = (size_t)N * K * 8 * 56 + (size_t)N * K * 8 * 6;
    if (o.krylov == 1) {
        const double m = o.restart > 0 ? o.restart : 40;
        const double mk = m * K;
        b += (size_t)(8.0 * ((m + 1) * K * mk + mk * mk + 3.0 * (m + 1) * K * K));
    }
    return b;
}

}  // namespace

struct dbie_server {
    dbie_server_opts o;
    std::mutex mu;
    std::condition_variable cv;
    bool stop = false;
    bool paused = false;   // slots start no new batch while set
    int64_t next_id = 1;
    uint64_t tick = 0;
    std::map<int64_t, std::shared_ptr<Request>> reqs;  // every request not yet taken
    std::deque<std::shared_ptr<Request>> queue;        // waiting for admission (FIFO)
    std::deque<std::shared_ptr<Request>> ready;        // set up, waiting for a slot
    std::map<GeoKey, std::shared_ptr<OpEntry>> cache;
    size_t budget = 0, inflight = 0, peak = 0;
    int64_t builds = 0, hits = 0, done = 0, failed = 0, batches = 0, fused = 0, multi = 0;
    std::vector<std::thread> threads;

    size_t cached_bytes() const {
        size_t b = 0;
        for (auto &kv : cache) b += kv.second->bytes;
        return b;
    }
    size_t in_use() const { return cached_bytes() + inflight; }
    void note_peak() { peak = std::max(peak, in_use()); }
    // drop idle cached operators (oldest first) until `need` more bytes fit
    bool make_room(size_t need) {
        while (in_use() + need > budget) {
            std::shared_ptr<OpEntry> victim;
            for (auto &kv : cache) {
                auto &e = kv.second;
                if (e->users == 0 && !e->building && !e->busy && (!victim || e->last_use < victim->last_use)) victim = e;
            }
            if (!victim) return false;
            cache.erase(victim->key);
        }
        return true;
    }
    void finish(const std::shared_ptr<Request> &r, St st, dbie_status status, const std::string &msg) {
        r->st = st;
        r->status = status;
        r->msg = msg;
        r->t_solve1 = Clock::now();
        inflight -= r->reserved + r->op_reserved;
        r->reserved = r->op_reserved = 0;
        if (r->ent) {
            r->ent->users--;
            if (r->ent->failed && r->ent->users == 0) cache.erase(r->ent->key);
            else if (!o.cache_ops && r->ent->users == 0 && !r->ent->busy) cache.erase(r->ent->key);
            r->ent.reset();
        }
        if (st == St::Done) ++done;
        else ++failed;
        r->F.clear();
        r->F.shrink_to_fit();
        cv.notify_all();
    }

    // ---------------------------------------------------------------- setup worker
    void setup_loop() {
        std::unique_lock<std::mutex> lk(mu);
        for (;;) {
            cv.wait(lk, [&] { return stop || pick_head(); });
            if (stop) return;
            const std::shared_ptr<Request> r = queue.front();
            queue.pop_front();
            // admission: bind or create the operator entry and reserve memory
            const GeoKey key{r->L[0], r->L[1], r->L[2], r->nq, r->l_ec};
            auto it = cache.find(key);
            std::shared_ptr<OpEntry> ent;
            bool build = false;
            if (it != cache.end()) {
                ent = it->second;
                r->cache_hit = true;
                ++hits;
            } else {
                ent = std::make_shared<OpEntry>();
                ent->key = key;
                ent->bytes = 0;
                cache[key] = ent;
                build = true;
            }
            ent->users++;
            ent->last_use = ++tick;
            r->ent = ent;
            r->reserved = r->pending_ws;
            r->op_reserved = build ? r->pending_op : 0;
            inflight += r->reserved + r->op_reserved;
            note_peak();
            r->st = St::Setup;
            r->t_setup0 = Clock::now();
            lk.unlock();
            dbie_status s = DBIE_OK;
            std::string msg;
            if (build) {
                s = build_op(*r, *ent, msg);
            }
            lk.lock();
            if (!build) {  // wait for a concurrent build of the same operator
                cv.wait(lk, [&] { return stop || !ent->building; });
                if (stop) return;
                if (ent->failed) {
                    s = DBIE_OOM;
                    msg = "operator build failed for this geometry";
                }
            } else {
                ent->building = false;
                inflight -= r->op_reserved;
                r->op_reserved = 0;
                if (s != DBIE_OK) ent->failed = true;
                cv.notify_all();
            }
            if (s == DBIE_OK) {
                r->N = ent->N;
                lk.unlock();
                s = make_rhs(*r, *ent, msg);
                lk.lock();
            }
            r->t_setup1 = Clock::now();
            if (s != DBIE_OK) {
                finish(r, St::Failed, s, msg);
                continue;
            }
            if (r->cancel) {
                finish(r, St::Cancelled, DBIE_OK, "cancelled");
                continue;
            }
            r->st = St::Ready;
            ready.push_back(r);
            cv.notify_all();
        }
    }
    // True when the head of the queue can be admitted now (called with the lock held).  Cancelled
    // heads and heads that can never fit the budget are finished here.
    bool pick_head() {
        while (!queue.empty()) {
            const std::shared_ptr<Request> r = queue.front();
            if (r->cancel) {
                queue.pop_front();
                finish(r, St::Cancelled, DBIE_OK, "cancelled");
                continue;
            }
            const GeoKey key{r->L[0], r->L[1], r->L[2], r->nq, r->l_ec};
            const size_t need = r->pending_ws + (cache.count(key) ? 0 : r->pending_op);
            if (need > budget) {
                queue.pop_front();
                finish(r, St::Failed, DBIE_OOM,
                       "request needs about " + std::to_string(need) + " device bytes; budget " + std::to_string(budget));
                continue;
            }
            return make_room(need);
        }
        return false;
    }
    dbie_status build_op(Request &r, OpEntry &ent, std::string &msg) {
        if (o.inject_oom_every > 0 && (++build_count % o.inject_oom_every) == 0) {
            msg = "injected allocation failure (test hook)";
            return DBIE_OOM;
        }
        dbie_geom *g = nullptr;
        dbie_status s = dbie_geometry_box(r.L[0], r.L[1], r.L[2], r.nq, r.l_ec, std::sqrt(2.0), &g);
        if (s != DBIE_OK) {
            msg = last_error_text();
            return s;
        }
        int64_t P = 0, N = 0;
        int nq = 0;
        dbie_geom_sizes(g, &P, &N, &nq);
        std::vector<double> ei(P, r.eps_in), eo(P, r.eps_out);
        dbie_near_opts no;
        dbie_near_opts_default(&no);
        no.n_threads = o.setup_threads;
        dbie_op *op = nullptr;
        s = dbie_operator_create(g, ei.data(), eo.data(), &no, &op);
        if (s != DBIE_OK) msg = last_error_text();
        dbie_geom_destroy(g);
        if (s != DBIE_OK) return s;
        // admission counts the operator's resident bytes and the Krylov workspaces only: no stored
        // far-field matrix for served operators
        dbie_op_stored_far(op, 0, nullptr);
        int64_t b = 0;
        dbie_op_device_bytes(op, &b);
        std::lock_guard<std::mutex> lk(mu);
        ent.op = op;
        ent.N = N;
        ent.P = P;
        ent.eps_in = r.eps_in;
        ent.eps_out = r.eps_out;
        ent.bytes = (size_t)b;
        ++builds;
        return DBIE_OK;
    }
    dbie_status make_rhs(Request &r, OpEntry &ent, std::string &msg) {
        r.F.assign((size_t)ent.N * r.K, 0.0);
        // the RHS kernel runs on the operator's stream; a slot may be solving on it, so take the
        // entry's turn like a solve would
        std::unique_lock<std::mutex> lk(mu);
        cv.wait(lk, [&] { return stop || !ent.busy; });
        if (stop) return DBIE_HIP_ERROR;
        ent.busy = true;
        lk.unlock();
        const dbie_status s = dbie_rhs_points(ent.op, r.K, r.pts.data(), r.q.data(), r.eps_in, r.F.data(), ent.N);
        if (s != DBIE_OK) msg = last_error_text();
        lk.lock();
        ent.busy = false;
        cv.notify_all();
        return s;
    }

    // ---------------------------------------------------------------- solve slot
    static bool same_opts(const dbie_solve_opts &a, const dbie_solve_opts &b) {
        return a.rtol == b.rtol && a.atol == b.atol && a.itmax == b.itmax && a.restart == b.restart &&
               a.precondition == b.precondition && a.precision == b.precision && a.inner_rtol == b.inner_rtol &&
               a.max_refine == b.max_refine && a.krylov == b.krylov &&
               a.recycle == b.recycle && a.verify == b.verify && a.check_residual == b.check_residual;
    }
    // One member of a batch: requests on one operator with the same permittivities, one column block.
    struct Group {
        std::shared_ptr<OpEntry> ent;
        double eps_in = 0, eps_out = 0;
        std::vector<std::shared_ptr<Request>> reqs;
        int K = 0;
    };
    static constexpr int kMaxMembers = 8;  // operators fused into one Krylov process
    // Requests join a fused Krylov process of several operators only when their own apply is small
    // (N <= 32,768 and N^2 K <= 2e10 pair-column products; DBIE_FUSE_MAX_N / DBIE_FUSE_MAX_WORK
    // override): there the per-iteration launch and synchronisation costs dominate and one fused
    // process of 8 such systems takes half the time of solving them one after another (N = 4,992,
    // 1-14 columns, split-FP16 and FP64); on the serving stream of 32 requests (N 7.5k-54k, 1-64
    // columns) these bounds fuse a quarter of the requests and raise served throughput by 6% over
    // N <= 16,384 and 4e9.  Larger applies fill the GPU on their own; fusing them runs every member
    // for the slowest member's iteration count.
    static bool fusable(const Request &r) {
        static const double max_n = env_double("DBIE_FUSE_MAX_N", 32768.0);
        static const double max_work = env_double("DBIE_FUSE_MAX_WORK", 2.0e10);
        return r.N <= max_n && (double)r.N * (double)r.N * r.K <= max_work;
    }
    static double env_double(const char *name, double dflt) {
        const char *e = std::getenv(name);
        if (!e || !*e) return dflt;
        char *end = nullptr;
        const double v = std::strtod(e, &end);
        return end && *end == '\0' && v >= 0 ? v : dflt;
    }

    void slot_loop() {
        std::unique_lock<std::mutex> lk(mu);
        for (;;) {
            cv.wait(lk, [&] {
                if (stop) return true;
                if (paused) return false;
                for (auto &r : ready)
                    if (!r->ent->busy) return true;
                return false;
            });
            if (stop) return;
            // the oldest ready request whose operator is free, its companions on that operator, and
            // ready requests on other free operators (fused members)
            std::shared_ptr<Request> h;
            for (auto &r : ready)
                if (!r->ent->busy) {
                    h = r;
                    break;
                }
            std::vector<Group> groups;
            int Ktot = 0;
            int64_t Nmax = 0;
            double useful = 0.0;  // sum of N x K over the batch
            size_t reserved = 0;  // workspace the batch's requests hold
            size_t extra = 0;     // fused solve: padded workspace beyond those reservations
            for (auto it = ready.begin(); it != ready.end();) {
                const std::shared_ptr<Request> r = *it;
                bool take = false;
                if (!r->ent->busy && same_opts(r->sopts, h->sopts) &&
                    (groups.empty() || Ktot + r->K <= o.max_batch_rhs)) {
                    Group *g = nullptr;
                    for (auto &gg : groups)
                        if (gg.ent == r->ent) g = &gg;
                    const int64_t nN = std::max(Nmax, r->N);
                    const int nK = Ktot + r->K;
                    const double nu = useful + (double)r->N * r->K;
                    // padding to the largest member may waste at most half the Krylov work
                    const bool pad_ok = groups.empty() || (double)nN * nK <= 2.0 * nu;
                    const size_t nres = reserved + r->reserved;
                    const bool multi_next = groups.size() + (g ? 0 : 1) > 1;
                    size_t nextra = 0;
                    if (multi_next) {
                        const size_t need = estimate_ws_bytes(nN, nK, h->sopts) + 16 * (size_t)nN * nK;
                        nextra = need > nres ? need - nres : 0;
                    }
                    const bool mem_ok = !multi_next || in_use() + nextra <= budget;
                    if (g) {
                        take = g->eps_in == r->eps_in && g->eps_out == r->eps_out && pad_ok && mem_ok;
                    } else {
                        // other operators join only the per-column Krylov solve (block GMRES would mix
                        // the columns of different operators)
                        take = groups.empty() || (h->sopts.krylov == 0 && (int)groups.size() < kMaxMembers &&
                                                  fusable(*h) && fusable(*r) && pad_ok && mem_ok);
                    }
                    if (take) {
                        if (!g) {
                            groups.push_back(Group{r->ent, r->eps_in, r->eps_out, {}, 0});
                            g = &groups.back();
                        }
                        g->reqs.push_back(r);
                        g->K += r->K;
                        Ktot = nK;
                        Nmax = nN;
                        useful = nu;
                        reserved = nres;
                        extra = nextra;
                    }
                }
                if (take) it = ready.erase(it);
                else ++it;
            }
            int nreq = 0;
            for (auto &g : groups) {
                g.ent->busy = true;
                g.ent->last_use = ++tick;
                nreq += (int)g.reqs.size();
            }
            ++batches;
            fused += nreq - 1;
            if (groups.size() > 1) ++multi;
            inflight += extra;
            note_peak();
            const auto t0 = Clock::now();
            for (auto &g : groups)
                for (auto &r : g.reqs) {
                    r->st = St::Solving;
                    r->t_solve0 = t0;
                    r->batch_K = Ktot;
                }
            lk.unlock();
            dbie_status s = DBIE_OK;
            std::string msg;
            // the operators' diagonals for the batch's permittivities
            for (auto &g : groups) {
                if (g.ent->eps_in == g.eps_in && g.ent->eps_out == g.eps_out) continue;
                std::vector<double> ei(g.ent->P, g.eps_in), eo(g.ent->P, g.eps_out);
                s = dbie_op_set_eps(g.ent->op, ei.data(), eo.data());
                if (s != DBIE_OK) {
                    msg = last_error_text();
                    break;
                }
                g.ent->eps_in = g.eps_in;
                g.ent->eps_out = g.eps_out;
            }
            std::vector<std::vector<double>> F(groups.size()), S(groups.size());
            for (size_t b = 0; b < groups.size(); ++b) {
                const int64_t N = groups[b].ent->N;
                F[b].resize((size_t)N * groups[b].K);
                S[b].resize((size_t)N * groups[b].K);
                int64_t col = 0;
                for (auto &r : groups[b].reqs) {
                    std::memcpy(&F[b][(size_t)col * N], r->F.data(), sizeof(double) * N * r->K);
                    col += r->K;
                }
            }
            dbie_solve_stats st{};
            if (s == DBIE_OK) {
                if (groups.size() == 1) {
                    const int64_t N = groups[0].ent->N;
                    s = dbie_solve_block(groups[0].ent->op, Ktot, F[0].data(), N, S[0].data(), N, &h->sopts, &st);
                } else {
                    std::vector<dbie_op *> ops;
                    std::vector<int> Ks;
                    std::vector<const double *> Fp;
                    std::vector<double *> Sp;
                    std::vector<int64_t> ld;
                    for (size_t b = 0; b < groups.size(); ++b) {
                        ops.push_back(groups[b].ent->op);
                        Ks.push_back(groups[b].K);
                        Fp.push_back(F[b].data());
                        Sp.push_back(S[b].data());
                        ld.push_back(groups[b].ent->N);
                    }
                    dbie_status rc = DBIE_OK;
                    s = dbie::guard([&] {
                        rc = dbie::solve_fused_host((int)groups.size(), ops.data(), Ks.data(), Fp.data(), ld.data(),
                                                    Sp.data(), ld.data(), &h->sopts, &st);
                    });
                    if (s == DBIE_OK) s = rc;
                }
                if (s != DBIE_OK) msg = last_error_text();
            }
            lk.lock();
            inflight -= extra;
            for (size_t b = 0; b < groups.size(); ++b) {
                const int64_t N = groups[b].ent->N;
                int64_t col = 0;
                for (auto &r : groups[b].reqs) {
                    r->stats = st;
                    if (s == DBIE_OK || s == DBIE_NOT_CONVERGED)
                        r->sigma.assign(S[b].begin() + (size_t)col * N, S[b].begin() + (size_t)(col + r->K) * N);
                    col += r->K;
                    if (r->cancel) finish(r, St::Cancelled, DBIE_OK, "cancelled");
                    else if (s == DBIE_OK) finish(r, St::Done, DBIE_OK, "");
                    else finish(r, St::Failed, s, msg);
                }
                auto &ent = groups[b].ent;
                ent->busy = false;
                int64_t bytes = 0;
                if (ent->op && dbie_op_device_bytes(ent->op, &bytes) == DBIE_OK) ent->bytes = (size_t)bytes;
                if (!o.cache_ops && ent->users == 0) cache.erase(ent->key);
            }
            note_peak();
            cv.notify_all();
        }
    }
    int64_t build_count = 0;
};

extern "C" {

void dbie_server_opts_default(dbie_server_opts *o) {
    if (!o) return;
    o->hbm_budget = 0.0;
    o->setup_workers = 2;
    o->setup_threads = 0;
    o->solve_slots = 2;
    o->max_batch_rhs = 256;
    o->cache_ops = 1;
    o->inject_oom_every = 0;
}

dbie_status dbie_server_create(const dbie_server_opts *opts, dbie_server **out) {
    return dbie::guard([&] {
        if (!out) dbie::fail(DBIE_BAD_ARG, "NULL output");
        *out = nullptr;
        dbie::ensure_device();
        std::unique_ptr<dbie_server> s(new dbie_server);
        if (opts) s->o = *opts;
        else dbie_server_opts_default(&s->o);
        dbie_server_opts &o = s->o;
        if (o.setup_workers < 1 || o.solve_slots < 1 || o.max_batch_rhs < 1)
            dbie::fail(DBIE_BAD_ARG, "setup_workers, solve_slots and max_batch_rhs must be >= 1");
        if (o.setup_threads <= 0)
            o.setup_threads = std::max(1, (int)std::thread::hardware_concurrency() / o.setup_workers);
        size_t freeb = 0, total = 0;
        DBIE_HIP(hipMemGetInfo(&freeb, &total));
        s->budget = o.hbm_budget > 0 ? (size_t)o.hbm_budget : (size_t)(0.8 * (double)freeb);
        for (int i = 0; i < o.setup_workers; ++i) s->threads.emplace_back([p = s.get()] { p->setup_loop(); });
        for (int i = 0; i < o.solve_slots; ++i) s->threads.emplace_back([p = s.get()] { p->slot_loop(); });
        *out = s.release();
    });
}

void dbie_server_destroy(dbie_server *s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
        s->cv.notify_all();
    }
    for (auto &t : s->threads) t.join();
    s->reqs.clear();
    s->queue.clear();
    s->ready.clear();
    s->cache.clear();
    delete s;
}

dbie_status dbie_server_submit_box(dbie_server *s, double Lx, double Ly, double Lz, int n_quad, double l_ec,
                                   double eps_in, double eps_out, int K, const double *pts, const double *q,
                                   const dbie_solve_opts *opts, int64_t *id) {
    return dbie::guard([&] {
        if (!s || !id) dbie::fail(DBIE_BAD_ARG, "NULL server or id");
        if (K < 1 || !pts || !q) dbie::fail(DBIE_BAD_ARG, "need K >= 1 charges with positions and values");
        if (!(Lx > 0 && Ly > 0 && Lz > 0 && l_ec > 0) || n_quad < 1 || n_quad > 8)
            dbie::fail(DBIE_BAD_ARG, "invalid box parameters");
        if (!(eps_in > 0 && eps_out > 0) || eps_in == eps_out)
            dbie::fail(DBIE_BAD_ARG, "permittivities must be positive and differ (t = (eo+ei)/(2(eo-ei)))");
        auto r = std::make_shared<Request>();
        r->L[0] = Lx;
        r->L[1] = Ly;
        r->L[2] = Lz;
        r->nq = n_quad;
        r->l_ec = l_ec;
        r->eps_in = eps_in;
        r->eps_out = eps_out;
        r->K = K;
        r->pts.assign(pts, pts + 3 * (size_t)K);
        r->q.assign(q, q + K);
        if (opts) r->sopts = *opts;
        else {
            dbie_solve_opts_default(&r->sopts);
            r->sopts.precision = DBIE_FP16X3;
        }
        r->sopts.check_residual = 0;
        // node count of the mesh, for the admission estimate (host mesher, no device work)
        dbie_geom *g = nullptr;
        if (dbie_geometry_box(Lx, Ly, Lz, n_quad, l_ec, std::sqrt(2.0), &g) != DBIE_OK)
            dbie::fail(DBIE_BAD_ARG, last_error_text());
        int64_t P = 0, N = 0;
        int nq = 0;
        dbie_geom_sizes(g, &P, &N, &nq);
        dbie_geom_destroy(g);
        r->pending_op = estimate_op_bytes(N, n_quad);
        r->pending_ws = estimate_ws_bytes(N, K, r->sopts);  // a request is solved with all its columns
        r->t_arrive = Clock::now();
        std::lock_guard<std::mutex> lk(s->mu);
        r->id = s->next_id++;
        s->reqs[r->id] = r;
        s->queue.push_back(r);
        *id = r->id;
        s->cv.notify_all();
    });
}

dbie_status dbie_server_wait(dbie_server *s, int64_t id, double timeout_s, dbie_request_info *info) {
    return dbie::guard([&] {
        if (!s) dbie::fail(DBIE_BAD_ARG, "NULL server");
        std::unique_lock<std::mutex> lk(s->mu);
        auto it = s->reqs.find(id);
        if (it == s->reqs.end()) dbie::fail(DBIE_BAD_ARG, "unknown request id " + std::to_string(id));
        const std::shared_ptr<Request> r = it->second;
        auto finished = [&] { return r->st == St::Done || r->st == St::Failed || r->st == St::Cancelled; };
        if (timeout_s < 0) s->cv.wait(lk, finished);
        else s->cv.wait_for(lk, std::chrono::duration<double>(timeout_s), finished);
        if (info) {
            std::memset(info, 0, sizeof *info);
            info->id = r->id;
            info->status = r->status;
            info->state = (int)r->st;
            info->N = r->N;
            info->K = r->K;
            info->batch_K = r->batch_K;
            info->solved = r->stats.solved;
            info->niter = r->stats.niter;
            info->nrefine = r->stats.nrefine;
            info->max_rel_residual = r->stats.max_rel_residual;
            info->op_cache_hit = r->cache_hit ? 1 : 0;
            if (finished()) {
                const bool set = r->t_setup0 != Clock::time_point{};
                info->queue_ms = set ? ms_since(r->t_arrive, r->t_setup0) : 0.0;
                info->setup_ms = set ? ms_since(r->t_setup0, r->t_setup1) : 0.0;
                info->solve_ms = r->t_solve0 != Clock::time_point{} ? ms_since(r->t_solve0, r->t_solve1) : 0.0;
                info->total_ms = ms_since(r->t_arrive, r->t_solve1);
            }
            std::strncpy(info->message, r->msg.c_str(), sizeof info->message - 1);
        }
        if (!finished()) dbie::fail(DBIE_PENDING, "request " + std::to_string(id) + " still running");
    });
}

dbie_status dbie_server_take(dbie_server *s, int64_t id, double *Sigma, int64_t lds) {
    return dbie::guard([&] {
        if (!s) dbie::fail(DBIE_BAD_ARG, "NULL server");
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->reqs.find(id);
        if (it == s->reqs.end()) dbie::fail(DBIE_BAD_ARG, "unknown request id " + std::to_string(id));
        const std::shared_ptr<Request> r = it->second;
        if (r->st != St::Done && r->st != St::Failed && r->st != St::Cancelled)
            dbie::fail(DBIE_PENDING, "request still running");
        if (Sigma) {
            if (r->sigma.empty()) dbie::fail(DBIE_BAD_ARG, "request has no result: " + r->msg);
            if (lds < r->N) dbie::fail(DBIE_DIM_MISMATCH, "lds smaller than N");
            for (int k = 0; k < r->K; ++k)
                std::memcpy(Sigma + (size_t)k * lds, &r->sigma[(size_t)k * r->N], sizeof(double) * r->N);
        }
        s->reqs.erase(it);
    });
}

dbie_status dbie_server_cancel(dbie_server *s, int64_t id) {
    return dbie::guard([&] {
        if (!s) dbie::fail(DBIE_BAD_ARG, "NULL server");
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->reqs.find(id);
        if (it == s->reqs.end()) dbie::fail(DBIE_BAD_ARG, "unknown request id " + std::to_string(id));
        const std::shared_ptr<Request> r = it->second;
        r->cancel = true;
        if (r->st == St::Queued) {
            for (auto q = s->queue.begin(); q != s->queue.end(); ++q)
                if (*q == r) {
                    s->queue.erase(q);
                    break;
                }
            s->finish(r, St::Cancelled, DBIE_OK, "cancelled");
        } else if (r->st == St::Ready) {
            for (auto q = s->ready.begin(); q != s->ready.end(); ++q)
                if (*q == r) {
                    s->ready.erase(q);
                    break;
                }
            s->finish(r, St::Cancelled, DBIE_OK, "cancelled");
        }
        s->cv.notify_all();
    });
}

dbie_status dbie_server_pause(dbie_server *s, int paused) {
    return dbie::guard([&] {
        if (!s) dbie::fail(DBIE_BAD_ARG, "NULL server");
        std::lock_guard<std::mutex> lk(s->mu);
        s->paused = paused != 0;
        s->cv.notify_all();
    });
}

dbie_status dbie_server_stats(dbie_server *s, dbie_server_summary *out) {
    return dbie::guard([&] {
        if (!s || !out) dbie::fail(DBIE_BAD_ARG, "NULL argument");
        std::lock_guard<std::mutex> lk(s->mu);
        out->completed = s->done;
        out->failed = s->failed;
        out->operator_builds = s->builds;
        out->cache_hits = s->hits;
        out->batches = s->batches;
        out->fused_requests = s->fused;
        out->multi_geometry_batches = s->multi;
        out->cached_operators = (int64_t)s->cache.size();
        out->bytes_in_use = (int64_t)s->in_use();
        out->peak_bytes = (int64_t)s->peak;
        out->budget_bytes = (int64_t)s->budget;
    });
}

}  // extern "C"
