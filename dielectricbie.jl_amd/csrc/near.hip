// near.hip -- block-sparse near-field correction fused with the far-field split
// reduction and the diagonal contrast term (mul! of BatchedDielectricOperator,
// src/solver/multi_rhs.jl:103-130):
//     Y[i,k] = sum_s part[s][k][i] + t_i X[i,k] + sum_b C_b[i - bd*p, :] . X[bd*src_b + :, k]
// Block rows are target panels; each block is a dense bd x bd (bd = n_quad^2)
// row-major tile.  Values are read once per launch (HBM-bound for K = 1).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "kernels.hpp"

namespace dbie {

namespace {
template <typename VT>
__global__ __launch_bounds__(256) void near_combine_kernel(int P, int bd, const int *__restrict__ row_ptr,
                                                           const int *__restrict__ col, const VT *__restrict__ val,
                                                           int S, const double *__restrict__ part, int64_t N, int K,
                                                           int KTe, int RP, const double *__restrict__ X, int64_t ldx,
                                                           const double *__restrict__ diag, double *__restrict__ Y,
                                                           int64_t ldy) {
    const int tid = threadIdx.x;
    const int row = tid % bd;
    const int rest = tid / bd;
    const int kc = rest % KTe;
    const int pp = rest / KTe;
    if (pp >= RP) return;
    const int p = blockIdx.x * RP + pp;
    const int k = blockIdx.y * KTe + kc;
    if (p >= P || k >= K) return;
    const int64_t i = (int64_t)p * bd + row;
    const double *xk = X + (int64_t)k * ldx;
    double acc = diag[i] * xk[i];
    for (int s = 0; s < S; ++s) acc += part[((int64_t)s * K + k) * N + i];
    double corr = 0.0;
    const int b1 = row_ptr[p + 1];
    for (int b = row_ptr[p]; b < b1; ++b) {
        const VT *v = val + ((int64_t)b * bd + row) * bd;
        const double *xs = xk + (int64_t)col[b] * bd;
        for (int c = 0; c < bd; ++c) corr += (double)v[c] * xs[c];
    }
    Y[i + (int64_t)k * ldy] = acc + corr;
}
}  // namespace

void near_combine(hipStream_t st, const DevBSR &C, bool use_f32, int S, const double *part, int64_t N, int K,
                  const double *X, int64_t ldx, const double *diag, double *Y, int64_t ldy) {
    const int bd = C.bd;
    const int KT = std::max(1, 256 / bd);
    const int KTe = std::min(K, KT);
    const int RP = std::max(1, KT / KTe);
    dim3 grid((unsigned)((C.P + RP - 1) / RP), (unsigned)((K + KTe - 1) / KTe));
    if (use_f32 && C.val32)
        near_combine_kernel<float><<<grid, 256, 0, st>>>(C.P, bd, C.row_ptr, C.col, C.val32, S, part, N, K, KTe, RP, X,
                                                          ldx, diag, Y, ldy);
    else
        near_combine_kernel<double><<<grid, 256, 0, st>>>(C.P, bd, C.row_ptr, C.col, C.val64, S, part, N, K, KTe, RP,
                                                           X, ldx, diag, Y, ldy);
}

}  // namespace dbie
