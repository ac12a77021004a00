// Self-check of the LU block kernel's reciprocal (lu.hip, block_step): for
// every normal float d with |d| < 2^126 (the pivots that kernel accepts),
// v_rcp_f32 followed by one FMA correction must equal the correctly rounded
// division 1.0f / d bit for bit.  Built beside libpiro_hip.so as
// libpiro_checks.so and run by the LU tests; not part of the library's ABI.
#include <hip/hip_runtime.h>
#include <cstdint>

namespace {

__global__ __launch_bounds__(256) void k_rcp_check(unsigned long long *counts)
{
    unsigned long long bad_in = 0, bad_out = 0;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t u = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; u < (1ull << 32); u += stride) {
        const uint32_t bits = (uint32_t)u, ab = bits & 0x7fffffffu;
        if (ab < 0x00800000u || ab >= 0x7f800000u) continue;  // zero, subnormal, infinite, NaN
        const float d = __uint_as_float(bits);
        const float ieee = 1.0f / d;
        const float r0 = __builtin_amdgcn_rcpf(d);
        const float r1 = __builtin_fmaf(__builtin_fmaf(-d, r0, 1.0f), r0, r0);
        if (__float_as_uint(r1) != __float_as_uint(ieee)) {
            if (ab < 0x7e800000u) ++bad_in;
            else ++bad_out;
        }
    }
    if (bad_in) atomicAdd(&counts[0], bad_in);
    if (bad_out) atomicAdd(&counts[1], bad_out);
}

}  // namespace

// counts[0]: mismatches among the accepted pivots (must be 0); counts[1]:
// among normal pivots of magnitude 2^126 or more (flagged by the kernel).
// Returns 0, or a nonzero HIP error code.
extern "C" int piro_check_reciprocal(unsigned long long *counts)
{
    unsigned long long *d = nullptr;
    hipError_t e = hipMalloc(&d, 2 * sizeof(unsigned long long));
    if (e != hipSuccess) return (int)e;
    e = hipMemset(d, 0, 2 * sizeof(unsigned long long));
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_rcp_check, dim3(4096), dim3(256), 0, 0, d);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(counts, d, 2 * sizeof(unsigned long long), hipMemcpyDeviceToHost);
    (void)hipFree(d);
    return (int)e;
}
