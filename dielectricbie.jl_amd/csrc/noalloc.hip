// noalloc.hip -- the operator apply for callers that cannot allocate device memory (code under
// HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
//
// Every workspace of the apply (far-field operand staging, partial sums, the FP32 copy of X for
// the full near blocks, the fp16 scale reduction) is carved out of one file-scope device array,
// `scratch`, of kScratchBytes.  Each GPU the library runs on holds its own copy of the array (at
// its own address) from the first use on that device until the process ends.  Host code never
// passes the array's name to a kernel or a copy: the name is a host-side stand-in, not a device
// address.  The address is looked up once per device with hipGetSymbolAddress, with that device
// current, and only that pointer is used.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "internal.hpp"
#include "runtime.hpp"

namespace dbie {
namespace {

// 256 MiB: the split-FP16 apply at N = 64,384, K = 256 (operand split 66 MB, X copy 66 MB, one
// split of partial sums 132 MB) fits; so does the FP64 MFMA apply of the same size.
constexpr size_t kScratchBytes = (size_t)256 << 20;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

// The device address of this GPU's scratch array (looked up on the first call on each device).
void *device_scratch() {
    int dev = 0;
    DBIE_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) fail(DBIE_BAD_ARG, "device ordinal out of range for the scratch table");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            fail(DBIE_HIP_ERROR, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[dev] = p;
    }
    return g_scratch_ptr[dev];
}

}  // namespace
}  // namespace dbie

using namespace dbie;

extern "C" {

dbie_status dbie_apply_device_noalloc(dbie_op *op, int K, const double *dX, int64_t ldx, double *dY, int64_t ldy,
                                      dbie_precision precision) {
    return guard([&] {
        if (!op) fail(DBIE_BAD_ARG, "operator handle is NULL");
        if (K < 1 || !dX || !dY) fail(DBIE_BAD_ARG, "K must be >= 1 and buffers non-NULL");
        int64_t N = 0;
        if (dbie_op_sizes(op, &N, nullptr, nullptr) != DBIE_OK) fail(DBIE_BAD_ARG, "bad operator handle");
        if (ldx < N || ldy < N) fail(DBIE_DIM_MISMATCH, "leading dimension smaller than N");
        if (precision < DBIE_FP64 || precision > DBIE_FP16X3) fail(DBIE_BAD_ARG, "unknown precision");
        ensure_device();
        apply_noalloc(op, K, dX, ldx, dY, ldy, precision, device_scratch(), kScratchBytes);
    });
}

}  // extern "C"
