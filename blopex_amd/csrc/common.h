// Internal runtime shared by the HIP sources of libblopex_amd.so:
// one library stream, error reporting, device scratch and pinned staging.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstdlib>

#include "hip_multivector.h"

namespace bx {

// Records msg as the library's last error (blopex_amd_last_error).
void set_error(const char *fmt, ...) __attribute__((format(printf, 1, 2)));
// Records the error, prints it and aborts: for void callbacks of the C API,
// whose contract violations the reference turns into assert aborts too
// (fortran_options.h:18).
[[noreturn]] void fatal(const char *fmt, ...) __attribute__((format(printf, 1, 2)));

#define BX_HIP(expr)                                                                      \
    do {                                                                                  \
        hipError_t e_ = (expr);                                                           \
        if (e_ != hipSuccess)                                                             \
            ::bx::fatal("%s:%d: %s failed: %s", __FILE__, __LINE__, #expr,               \
                        hipGetErrorString(e_));                                           \
    } while (0)

#define BX_REQUIRE(cond, ...)                                                             \
    do {                                                                                  \
        if (!(cond)) ::bx::fatal(__VA_ARGS__);                                            \
    } while (0)

// The stream every kernel and copy of the library is ordered on.
hipStream_t stream();
int device();
int num_cus();

// Library-owned device scratch, grown on demand and kept between calls
// (blopex_amd_release_workspace frees it); slot s is independent of the
// other slots.  Contents are undefined on return.
void *scratch(int slot, size_t bytes);
// Device address of the fixed array of hip_Multi_VectorInnerProdNoAlloc on
// the library's device, looked up once per device (nullptr on failure).
double *noalloc_setup();
// Pinned host staging buffer of at least `bytes`.
void *pinned(int slot, size_t bytes);
// Small host->device copy ordered on the stream; the host array may be reused
// on return (it goes through a pinned staging slot whose previous copy is
// waited for before reuse).
void upload_async(void *dst, const void *src, size_t bytes);
// Device->host copy on the stream, then waits for it.
void download_sync(void *dst, const void *src, size_t bytes);
// The same copy in two halves: download_start enqueues it, download_finish
// waits for the copy alone (work enqueued in between keeps running) and
// copies the bytes out.  One such copy at a time; download_sync may run in
// between (it stages through another pinned buffer).
void download_start(const void *src, size_t bytes);
void download_finish(void *dst, size_t bytes);

inline size_t elem_size(int precision) {
    return precision == HIPMV_FP32 ? 4 : precision == HIPMV_C64 ? 16 : 8;
}

// Device-side status of one solve; kernels after a fatal failure are no-ops.
struct DevStatus {
    int abort;      // set on a failure that ends the solve
    int x_info;     // Cholesky info of X'BX (initial orthonormalisation)
    int r_info;     // Cholesky info of R'BR
    int p_info;     // Cholesky info of P'BP (nonzero: P dropped this iteration)
    int gevp_info;  // Rayleigh-Ritz failure (dsygv-style info)
    int pad[3];
    double condest; // log10 of the pencil's B-part condition estimate + 1 (cond_limit > 0)
};

// Column maps passed to kernels by value (kernel arguments).
constexpr int kMaxCols = 128;
struct ColMap {
    int n;
    int idx[kMaxCols];
};

ColMap colmap_of(const hip_Multi_Vector *x);
ColMap colmap_identity(int n);

inline unsigned grid_for(size_t work, unsigned per_block, unsigned cap = 1u << 20) {
    size_t g = (work + per_block - 1) / per_block;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (unsigned)g;
}

}  // namespace bx
