// piro_internal.hpp -- shared plumbing of libpiro_hip (gfx950 only).
//
// Every extern "C" entry point goes through PIRO_API_BEGIN/END: errors are
// turned into a status code plus a thread-local message (piro_last_error),
// never an exception or exit across the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>

#include "../../include/piro_hip.h"

namespace piro {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

void set_last_error(const std::string &msg);

#define PIRO_CHECK_HIP(expr)                                                   \
    do {                                                                       \
        hipError_t _e = (expr);                                                \
        if (_e != hipSuccess)                                                  \
            throw ::piro::Error(PIRO_EHIP, std::string(#expr) + ": " +         \
                                               hipGetErrorString(_e));         \
    } while (0)

#define PIRO_REQUIRE(cond, msg)                                                \
    do {                                                                       \
        if (!(cond)) throw ::piro::Error(PIRO_EINVAL, (msg));                  \
    } while (0)

#define PIRO_API_BEGIN try {
#define PIRO_API_END                                                           \
    return PIRO_OK;                                                            \
    }                                                                          \
    catch (const ::piro::Error &e) {                                           \
        ::piro::set_last_error(e.what());                                      \
        return e.code;                                                         \
    }                                                                          \
    catch (const std::exception &e) {                                          \
        ::piro::set_last_error(e.what());                                      \
        return PIRO_EINTERNAL;                                                 \
    }                                                                          \
    catch (...) {                                                              \
        ::piro::set_last_error("unknown error");                               \
        return PIRO_EINTERNAL;                                                 \
    }

// Growable device scratch owned by a context (reduction partials, geam
// counts, LU panels ...).  Reallocation only grows.
struct Scratch {
    void *ptr = nullptr;
    size_t bytes = 0;
    void *get(size_t need) {
        if (need > bytes) {
            if (ptr) PIRO_CHECK_HIP(hipFree(ptr));
            ptr = nullptr;
            bytes = 0;
            size_t n = need < (1u << 20) ? (1u << 20) : need;
            PIRO_CHECK_HIP(hipMalloc(&ptr, n));
            bytes = n;
        }
        return ptr;
    }
    void release() {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        bytes = 0;
    }
};

}  // namespace piro

struct piro_ctx {
    int device = 0;
    hipStream_t stream = nullptr;   // null = legacy default stream
    bool own_stream = false;
    int num_cus = 256;
    int num_xcds = 8;
    piro::Scratch scratch[4];
    piro::Scratch lu_keep;  // the LU's copy of the matrix while it speculates
    hipEvent_t ev0 = nullptr, ev1 = nullptr;  // piro_timer_start/stop
    // autotuned FE-step choice keyed by the last matrix shape
    int fe_variant = -1;
    int64_t fe_tuned_rows = -1, fe_tuned_nnz = -1;
    // live piro_batch / piro_surrogate objects made on this context: it
    // refuses to be destroyed under them
    int live_children = 0;
    // LU panels try the exchange-free columns first (same factors either way)
    int lu_speculate = 1;
    // the LU solves' pinned host buffers (pivots, then the row permutation;
    // 2 lu_host_n ints) and the events after the pivots' copy to the host
    // and after the permutation's copy back
    int32_t *lu_host = nullptr;
    int64_t lu_host_n = 0;
    // pinned host memory for the small read-backs of the synchronous calls
    // (a factorisation's info and flag, the refinement's norms and error
    // word, the JLI state), one call at a time; made at first use
    void *host_read = nullptr;
    hipEvent_t lu_piv_ev = nullptr, lu_perm_ev = nullptr;
};

namespace piro {

constexpr size_t kHostReadBytes = 1024;

// the context's pinned read-back buffer (kHostReadBytes)
inline void *host_read(piro_ctx *c)
{
    if (!c->host_read) PIRO_CHECK_HIP(hipHostMalloc(&c->host_read, kHostReadBytes, hipHostMallocDefault));
    return c->host_read;
}

inline piro_ctx *ctx_of(piro_ctx *c) {
    if (!c) throw Error(PIRO_EINVAL, "null context");
    PIRO_CHECK_HIP(hipSetDevice(c->device));
    return c;
}

inline unsigned grid_for(int64_t n, int block) {
    int64_t g = (n + block - 1) / block;
    if (g < 1) g = 1;
    if (g > 0x7fffffff) throw Error(PIRO_EINVAL, "grid too large");
    return (unsigned)g;
}

inline void check_launch() { PIRO_CHECK_HIP(hipGetLastError()); }

}  // namespace piro
