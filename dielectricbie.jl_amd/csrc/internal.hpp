// internal.hpp -- shared host-side types of libdbie (not part of the C ABI).
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <functional>
#include <cstdint>
#include <thread>
#include <string>
#include <vector>

#include "../../include/dbie.h"

namespace dbie {

// ---- error handling -------------------------------------------------------
void set_error(const std::string &msg);
struct Error {
    dbie_status code;
    std::string msg;
};
[[noreturn]] void fail(dbie_status code, const std::string &msg);
void hip_check(hipError_t e, const char *what);
#define DBIE_HIP(call) ::dbie::hip_check((call), #call)

// Dynamic-schedule loop over [0, n) on n_threads host threads (0 = all cores);
// body(i, thread_index).
template <class F>
void parallel_for(int64_t n, int n_threads, F &&body) {
    int nt = n_threads > 0 ? n_threads : (int)std::max(1u, std::thread::hardware_concurrency());
    nt = (int)std::min<int64_t>(nt, std::max<int64_t>(n, 1));
    std::atomic<int64_t> next{0};
    auto worker = [&](int tid) {
        for (;;) {
            const int64_t i = next.fetch_add(1);
            if (i >= n) break;
            body(i, tid);
        }
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < nt; ++t) pool.emplace_back(worker, t);
    worker(0);
    for (auto &th : pool) th.join();
}

// ---- geometry --------------------------------------------------------------
constexpr double kFourPi = 4.0 * 3.14159265358979311599796346854;

struct Geometry {
    int64_t P = 0, N = 0;
    int nq = 0;
    std::vector<double> corners;  // P x 4 x 3 (a, b, c, d)
    std::vector<double> normal;   // P x 3
    std::vector<int> is_edge;     // P
    double glx[16] = {}, glw[16] = {}, bary[16] = {};
    // node streams (structure of arrays)
    std::vector<double> x, y, z, w;
    int npp() const { return nq * nq; }
};

void gauss_legendre(int n, double *x, double *w);  // throws on n < 1
bool best_grid(double W, double H, double alpha, int *m, int *n);
Geometry *make_box(double Lx, double Ly, double Lz, int nq, double l_ec, double alpha);
Geometry *make_from_panels(int64_t P, int nq, const double *corners, const double *normals, const int *is_edge);
// as make_from_panels with the caller's nodes (points 3 x N interleaved, weights N) and one rule for every
// panel (gl_xs, gl_ws; bary may be null: computed from them) taken as given
Geometry *make_from_panel_nodes(int64_t P, int nq, const double *corners, const double *normals, const int *is_edge,
                                const double *points, const double *weights, const double *gl_xs, const double *gl_ws,
                                const double *bary);
// multi_dielectric_box3d: boxes nb x 6 (centre x, y, z, Lx, Ly, Lz); per-panel permittivities out
Geometry *make_multi_box(int nq, double l_ec, int nb, const double *boxes, const double *epses, double eps_out,
                         double alpha, std::vector<double> &eps_in_p, std::vector<double> &eps_out_p);
// physicists' Gauss-Hermite rule (weight exp(-x^2)), ascending nodes; throws on n < 1
void gauss_hermite(int n, double *x, double *w);
void barycentric_row(int n, const double *x, const double *lam, double xq, double *r);

// ---- near field --------------------------------------------------------------
struct NearOpts {
    double up_tol = 1e-9;
    int max_order = 8;
    bool correct_edges = true;
    double adaptive_atol = 1e-9;
    double adaptive_rtol = 1.4901161193847656e-8;
    int adaptive_n_GL = 0;
    int adaptive_max_depth = 20;
    int n_threads = 0;
    int mode = 0;  // correction kernel: 0 = laplace3d_DT_corrections, 1 = laplace3d_D_corrections
};

struct PairList {
    // sorted by (src, trg); upsampled pairs carry order >= nq, touching pairs order = 0
    std::vector<int> up_src, up_trg, up_order;
    std::vector<int> adp_src, adp_trg;
};

PairList build_pair_list(const Geometry &g, int max_order, double atol, bool correct_edges, int n_threads);

// The pair list in three steps (build_pair_list runs all three on the host; the device scan
// replaces the middle one).  Per-panel data, kPairData doubles per panel: centre[3], ri^2,
// node radius, Bernstein frame cc[3] e1[3] e2[3] n[3] L1 L2, ri (kPairData = 20).
struct PairPrep {
    int max_order = 8;
    double atol = 1e-9, rho_star = 0, tc2 = 0;
    bool correct_edges = true;
    std::vector<double> len, off, pdata;
};
PairPrep pair_prep(const Geometry &g, int max_order, double atol, bool correct_edges);
// candidates per source panel (CSR over source panels, ascending target panel): the smallest
// Bernstein radius over the target's nodes in the ball (inf if none) and the corner-touch flag
struct PairCandidates {
    std::vector<int64_t> ptr;
    std::vector<int> trg;
    std::vector<double> rho;
    std::vector<unsigned char> touch;
};
PairList finish_pair_list(const Geometry &g, const PairPrep &pp, const PairCandidates &c);

// Block-sparse (BSR) correction matrix: block rows = target panels.
struct BlockSparse {
    int bd = 0;                    // block dimension = nq^2
    std::vector<int> row_ptr;      // P + 1
    std::vector<int> col;          // source panel per block
    std::vector<double> val;       // nb x bd x bd, element (t, c) at t*bd + c (row-major)
    int64_t nblocks() const { return (int64_t)col.size(); }
    int adaptive_hits = 0;         // touching pairs that reached max_depth
};

BlockSparse build_corrections(const Geometry &g, const PairList &pl, const NearOpts &o);

// Host part of the device correction build: BSR structure, the pair items with their BSR slots,
// per-source-panel caches (upsampled nodes, interpolation-quadrature matrix Mt) for the upsampled
// pairs, and the adaptive (touching-pair) K_up blocks, row-major, without the direct term.
struct CorrectionPlan {
    int bd = 0;
    std::vector<int> row_ptr, col;
    std::vector<int> up_src, up_trg, up_slot;
    std::vector<int> cache_off, cache_nu;   // per source panel (-1 / 0: no upsampled pairs)
    std::vector<double> caches;             // per panel: upos (3 n_up^2), then Mt (bd x n_up^2)
    std::vector<int> adp_src, adp_trg, adp_slot;
    std::vector<double> adp_kup;            // per touching pair: bd x bd (host quadtrees only)
    std::vector<int> adp_rep;               // per touching pair: representative target node in its panel
    std::vector<double> adp_map;            // per touching pair: source panel map cc[3], bma[3], dma[3], scale
    int adaptive_hits = 0;
    int64_t nblocks() const { return (int64_t)col.size(); }
};
// host_quadtrees = false leaves the touching pairs' quadtrees to the device (adp_rep, adp_map filled,
// adp_kup empty).
CorrectionPlan plan_corrections(const Geometry &g, const PairList &pl, const NearOpts &o, bool host_quadtrees);
// Gauss-Legendre rule and barycentric data the device quadtree walk needs: glx[nq], bary[nq],
// ns[nGL], ws[nGL] (nGL = adaptive_n_GL, or nq when 0).
std::vector<double> quadtree_rules(const Geometry &g, int nGL);
// K_up block of one touching pair on the host (source panel i, target panel j; bd x bd row-major,
// no direct term); returns the max-depth flag.
bool touching_block(const Geometry &g, int64_t i, int64_t j, const NearOpts &o, double *blk);

// ---- off-surface targets ------------------------------------------------------
// laplace3d_pottrg_fmm3d_corrected_hcubature (src/kernel/laplace3d_near_hcubature.jl:337-375):
// the interface refined around the targets, the prolongation of densities onto it, and the
// near-target corrections (adaptive cubature minus the direct quadrature term).
struct TargetMap {
    int64_t M = 0;
    std::vector<double> trg;          // 3 x M
    Geometry ref;                     // refined interface
    std::vector<int> parent;          // refined panel -> coarse panel
    std::vector<int> from_split;      // refined panel came from a split
    std::vector<int64_t> p_ptr;       // prolongation, CSR: N_ref rows x N columns
    std::vector<int> p_col;
    std::vector<double> p_val;
    std::vector<int64_t> c_ptr;       // corrections, CSR: M rows x N_ref columns (sorted columns)
    std::vector<int> c_col;
    std::vector<double> c_val;
    int64_t n_pairs = 0;              // (panel, target) cubatures
    int64_t n_capped = 0;             // cubatures that hit the evaluation cap
    int64_t n_device = 0;             // cubatures integrated on the device
    bool prolong_device = false;      // refinement and prolongation built on the device: p_col / p_val
                                      // live there only (p_ptr is kept on the host too)
    double setup_ms[4] = {0, 0, 0, 0};  // refinement, prolongation, pairs + cubatures, assembly
};
// The near-target cubatures of a target map: pair k = (refined panel panel[k], target target[k]);
// pdat: panel centre, b - a, d - a, area scale, target (13 doubles per pair).
struct TargetCubatureJob {
    int nq = 0;
    double atol = 0;
    int64_t maxevals = 0;
    std::vector<int> panel, target;
    std::vector<double> pdat;
};
// Optional device integrator: fills I (n_quad^2 per pair) and status (0 converged, 1 stopped at
// maxevals, 2 not computed: the host integrates those pairs).  g: the coarse geometry (rules).
// Genz-Malik two-dimensional rule: l2, l3, l4, l5, w[5], wp[4] (the host cubature's constants)
std::array<double, 13> genz_malik_rule();
using TargetCubatureFn = std::function<void(const TargetCubatureJob &, const Geometry &g, double *I, int *status)>;
// Optional device builder of the refinement and the prolongation (the first two phases of the target
// map).  In: the coarse interface, the targets and the host's uniform grid over them (cells of size
// h from lo, dim cells per axis, target lists start / idx), the panel size limit and range factor.
// Out: the refined panels (corners, parent, from_split) in the host's order, their nodes, and the
// prolongation's row pointers; the prolongation's entries stay on the device (the hook's owner keeps
// them).  Returns false (nothing written) when it declines; the host then builds both phases.
struct RefineInput {
    const Geometry *g = nullptr;
    const double *trg = nullptr;
    int64_t M = 0;
    double lo[3] = {0, 0, 0}, h = 1;
    int64_t dim[3] = {1, 1, 1};
    const std::vector<int64_t> *start = nullptr;
    const std::vector<int> *idx = nullptr;
    double limit = 0, range_factor = 0;
};
struct RefineOutput {
    std::vector<double> corners, x, y, z, w;
    std::vector<int> parent, from_split;
    std::vector<int64_t> p_ptr;
    double ms_refine = 0, ms_prolong = 0;
};
using TargetRefineFn = std::function<bool(const RefineInput &, RefineOutput &)>;
TargetMap *build_target_map(const Geometry &g, int64_t M, const double *trg, double atol, double range_factor,
                            int n_threads, const TargetCubatureFn &device = nullptr,
                            const TargetRefineFn &refine = nullptr);

// ---- operator apply without allocation (noalloc.hip) ----------------------------
// Scratch bytes dbie_apply_device needs when every workspace comes from one caller buffer: the
// far-field operand staging, the tile-major X copy of the full near blocks, one source split of
// partial sums and 256 bytes for the fp16 scale reduction.
size_t apply_noalloc_bytes(const dbie_op *op, int K, dbie_precision prec);
// Y = A X on the operator's stream with its workspaces carved out of scratch (bytes long, device
// memory, 256-byte aligned); throws DBIE_BAD_ARG when apply_noalloc_bytes exceeds bytes.  Uses the
// tuner's settled launch shape (never explores), no HIP graphs, no kernel timing.
void apply_noalloc(dbie_op *op, int K, const double *X, int64_t ldx, double *Y, int64_t ldy, dbie_precision prec,
                   void *scratch, size_t bytes);

// dbie_solve_fused: one Krylov process for the block systems of B operators (host blocks).
dbie_status solve_fused_host(int B, dbie_op *const *ops, const int *Ks, const double *const *F, const int64_t *ldf,
                             double *const *S, const int64_t *lds, const dbie_solve_opts *opts,
                             dbie_solve_stats *stats);

}  // namespace dbie

// Opaque handles of the C ABI.
struct dbie_geom {
    dbie::Geometry g;
    std::vector<double> eps_in, eps_out;  // per panel, when the mesher assigns permittivities
};
