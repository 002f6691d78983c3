// targets.hip -- device side of the off-surface target evaluation and the reaction matrix.
//
//  * csr_spmm: Y = A X (+ Y) for a CSR matrix A and a column-major block X; used for the
//    prolongation onto the refined interface (_refined_interface_prolongation,
//    src/kernel/laplace3d_near.jl:152-225), the near-target corrections
//    (laplace3d_pottrg_corrections_hcubature, src/kernel/laplace3d_near_hcubature.jl:193-251)
//    and the transposed prolongation of the reaction weights.  One thread per (row, column):
//    rows are short (1..n_quad^2 entries for the prolongation, a few panels' nodes for the
//    corrections), so the kernel is bound by the gathers of X from HBM/L2.
//  * reaction_q: Qt[r, a] = w_r / (4 pi) * sum_{s in group a} q_s / |t_s - y_r|, the weight of
//    refined node r in group a's potential integral.  With it the reaction matrix of
//    four_index_matrix (src/solver/multi_rhs.jl:299-321, induced term)
//        R[a, b] = sum_{s in a} q_s u[Sigma_b](t_s)
//    becomes R = (P^T (Qt + Qcorr))^T Sigma: one N x G weight block and one DGEMM instead of the
//    potential at every target (2 M N K flops).  FP64 VALU, LDS-staged source points.
#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace dbie {
namespace {
constexpr double kFourPiT = 4.0 * 3.14159265358979311599796346854;

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
            acc += p.w / sqrt(dx * dx + dy * dy + dz * dz);
        }
    }
    if (j < g.N) Qt[j + (int64_t)a * g.N] = g.w[j] * acc / kFourPiT;
}

__global__ void scatter_add_kernel(int64_t n, const int64_t *__restrict__ idx, const double *__restrict__ val,
                                   double *__restrict__ Y) {
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e < n) Y[idx[e]] += val[e];
}
}  // namespace

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

void scatter_add(hipStream_t st, int64_t n, const int64_t *idx, const double *val, double *Y) {
    if (n <= 0) return;
    scatter_add_kernel<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(n, idx, val, Y);
}

}  // namespace dbie
