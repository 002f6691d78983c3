// devmath.hpp -- device restatements of host helpers whose floating-point decisions must match
// the host bit for bit (same operations, same order, no contraction).
#pragma once
#include <hip/hip_runtime.h>

namespace dbie {

// barycentric_row (geometry.cpp): row of Lagrange basis values at xq for the nodes x with
// barycentric weights lam (src/utils/barycentric.jl:6-95)
__device__ inline void bary_row(int n, const double *x, const double *lam, double xq, double *r) {
#pragma clang fp contract(off)
    const double rtol = 0x1p-26;  // sqrt(DBL_EPSILON)
    for (int i = 0; i < n; ++i) {
        const double ax = fabs(xq), bx = fabs(x[i]);
        if (xq == x[i] || fabs(xq - x[i]) <= rtol * (ax < bx ? bx : ax)) {
            for (int k = 0; k < n; ++k) r[k] = 0.0;
            r[i] = 1.0;
            return;
        }
    }
    double den = 0.0;
    for (int i = 0; i < n; ++i) {
        r[i] = lam[i] / (xq - x[i]);
        den += r[i];
    }
    const double inv = 1.0 / den;
    for (int i = 0; i < n; ++i) r[i] *= inv;
}

// Exponent shift of a split-FP16 operand column: with m = max |w_j X_jk| (float bits, as the
// column-max kernel leaves them), column k is scaled by 2^col_exponent so that its entries stay
// below 2^14 (0 for an all-zero column).
__device__ inline int col_exponent(unsigned bits) {
    const float m = __uint_as_float(bits);
    if (!(m > 0.0f)) return 0;
    int e;
    (void)frexpf(m, &e);  // m < 2^e
    return 14 - e;
}

// Split-FP16 operands (hi and lo halves of the scaled w .* X, Kp columns): per tile of kZt = 32
// sources and block of 16 columns one 1 KiB chunk in MFMA B-operand order -- group of 8 sources,
// column, source within the group -- so that lane l of a wave reading 16 bytes at chunk + 16 l holds
// column l % 16, sources 8 (l / 16) .. 8 (l / 16) + 7, and a stage of 32 sources reads whole cache
// lines.  zsplit_idx is the element index of source j, column k.
constexpr int kZt = 32;
__device__ inline int64_t zsplit_idx(int64_t j, int k, int Kp) {
    return ((j >> 5) * Kp + (k & ~15)) * kZt + ((j >> 3) & 3) * 128 + (k & 15) * 8 + (j & 7);
}

}  // namespace dbie
