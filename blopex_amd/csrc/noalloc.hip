// hip_Multi_VectorInnerProdNoAlloc: the MultiInnerProd slot of the HIP
// interpreter (v = X(active)' Y(active), host h x w, leading dimension gh)
// for callers that cannot allocate device memory.  Its Gram kernels write
// per-workgroup partial tiles and the reduced h x w result into `scratch`,
// a file-scope device array of fixed size, instead of library workspace that
// grows on demand; the result reaches v through one 2D copy.
//
// Each device holds its own copy of `scratch` for the life of the process.
// Its device address is looked up once per device (noalloc_setup, from
// blopex_amd_init or the entry point's first call on that device); host code
// passes that address, never the array's name, to the kernels.  Calls that
// overlap on one device share the array and race: the partial tiles and the
// result are written whole by every call, so nothing needs clearing, but
// concurrent calls would overwrite each other's.
#include <cstring>

#include "kernels.h"

namespace {
constexpr int kNoAllocCols = 64;                         // h, w at most
constexpr size_t kNoAllocPartWords = (size_t)1040 * 256;  // partial tiles (FP64 words)
constexpr size_t kNoAllocWords = kNoAllocPartWords + (size_t)kNoAllocCols * kNoAllocCols;
}  // namespace

// file scope, outside the library namespace (whose scratch() is the growing
// workspace): the fixed array of this entry point
static __device__ __align__(256) double scratch[kNoAllocWords];

namespace bx {

namespace {

constexpr int kMaxDevices = 64;
double *g_scratch_ptr[kMaxDevices] = {};

}  // namespace

// Looks up the device address of `scratch` on the library's device (made
// current for the call); returns nullptr with the error recorded on failure.
double *noalloc_setup() {
    const int dev = device();
    if (dev < 0 || dev >= kMaxDevices) {
        set_error("hip_Multi_VectorInnerProdNoAlloc: device %d out of range", dev);
        return nullptr;
    }
    if (g_scratch_ptr[dev]) return g_scratch_ptr[dev];
    int cur = -1;
    if (hipGetDevice(&cur) != hipSuccess) {
        set_error("hip_Multi_VectorInnerProdNoAlloc: hipGetDevice failed");
        return nullptr;
    }
    if (cur != dev && hipSetDevice(dev) != hipSuccess) {
        set_error("hip_Multi_VectorInnerProdNoAlloc: hipSetDevice(%d) failed", dev);
        return nullptr;
    }
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(::scratch));
    if (cur != dev) (void)hipSetDevice(cur);
    if (e != hipSuccess || !p) {
        set_error("hip_Multi_VectorInnerProdNoAlloc: hipGetSymbolAddress failed: %s",
                  hipGetErrorString(e));
        return nullptr;
    }
    g_scratch_ptr[dev] = (double *)p;
    return g_scratch_ptr[dev];
}

}  // namespace bx

using namespace bx;

extern "C" int hip_Multi_VectorInnerProdNoAlloc(const hip_Multi_Vector *x,
                                                const hip_Multi_Vector *y, BlopexInt gh,
                                                BlopexInt h, BlopexInt w, double *v) {
    if (!x || !y || !v) {
        set_error("hip_Multi_VectorInnerProdNoAlloc: null argument");
        return 1;
    }
    const bool real = (x->precision == HIPMV_FP64 || x->precision == HIPMV_FP32);
    if (!real || y->precision != x->precision || y->size != x->size) {
        set_error("hip_Multi_VectorInnerProdNoAlloc: X and Y must be real blocks of one "
                  "precision and height");
        return 1;
    }
    if (h != x->num_active_vectors || w != y->num_active_vectors) {
        set_error("hip_Multi_VectorInnerProdNoAlloc: %d x %d result for %d x %d active columns",
                  (int)h, (int)w, (int)x->num_active_vectors, (int)y->num_active_vectors);
        return 1;
    }
    if (h > kNoAllocCols || w > kNoAllocCols || gh < h) {
        set_error("hip_Multi_VectorInnerProdNoAlloc: %d x %d result (leading dimension %d) "
                  "above the %d x %d maximum",
                  (int)h, (int)w, (int)gh, kNoAllocCols, kNoAllocCols);
        return 1;
    }
    if (h == 0 || w == 0) return 0;
    double *ws = noalloc_setup();
    if (!ws) return 1;
    double *part = ws, *g = ws + kNoAllocPartWords;
    gram_cols_into(x->precision, x->size, x->data, x->ld, colmap_of(x), y->data, y->ld,
                   colmap_of(y), g, h, part, kNoAllocPartWords);
    hipError_t e = hipMemcpy2DAsync(v, (size_t)gh * sizeof(double), g, (size_t)h * sizeof(double),
                                    (size_t)h * sizeof(double), (size_t)w, hipMemcpyDeviceToHost,
                                    stream());
    if (e == hipSuccess) e = hipStreamSynchronize(stream());
    if (e != hipSuccess) {
        set_error("hip_Multi_VectorInnerProdNoAlloc: %s", hipGetErrorString(e));
        return 1;
    }
    return 0;
}
