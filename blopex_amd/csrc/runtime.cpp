// Library runtime: stream, errors, scratch.
#include <cstdarg>
#include <cstring>
#include <mutex>

#include "common.h"

namespace bx {

namespace {
char g_err[1024] = "";
std::mutex g_mu;
hipStream_t g_stream = nullptr;
int g_device = -1;
int g_cus = 0;
constexpr int kSlots = 16;
void *g_scratch[kSlots] = {};
size_t g_scratch_bytes[kSlots] = {};
void *g_pinned[kSlots] = {};
size_t g_pinned_bytes[kSlots] = {};

void vset(const char *fmt, va_list ap) { vsnprintf(g_err, sizeof(g_err), fmt, ap); }

void init_locked() {
    if (g_stream) return;
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) {
        snprintf(g_err, sizeof(g_err), "no HIP device: %s", hipGetErrorString(e));
        fprintf(stderr, "blopex_amd: %s\n", g_err);
        abort();
    }
    g_device = dev;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) == hipSuccess) g_cus = prop.multiProcessorCount;
    if (g_cus <= 0) g_cus = 256;
    // default priority, non-blocking w.r.t. the null stream
    e = hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        snprintf(g_err, sizeof(g_err), "hipStreamCreate failed: %s", hipGetErrorString(e));
        fprintf(stderr, "blopex_amd: %s\n", g_err);
        abort();
    }
}
}  // namespace

void set_error(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vset(fmt, ap);
    va_end(ap);
}

void fatal(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vset(fmt, ap);
    va_end(ap);
    fprintf(stderr, "blopex_amd: %s\n", g_err);
    fflush(stderr);
    abort();
}

hipStream_t stream() {
    if (!g_stream) {
        std::lock_guard<std::mutex> lk(g_mu);
        init_locked();
    }
    return g_stream;
}

int device() {
    stream();
    return g_device;
}

int num_cus() {
    stream();
    return g_cus;
}

void *scratch(int slot, size_t bytes) {
    BX_REQUIRE(slot >= 0 && slot < kSlots, "scratch slot %d out of range", slot);
    if (bytes == 0) bytes = 256;
    if (g_scratch_bytes[slot] < bytes) {
        hipStream_t s = stream();
        if (g_scratch[slot]) {
            BX_HIP(hipStreamSynchronize(s));
            BX_HIP(hipFree(g_scratch[slot]));
        }
        size_t grow = bytes + bytes / 4;
        BX_HIP(hipMalloc(&g_scratch[slot], grow));
        g_scratch_bytes[slot] = grow;
    }
    return g_scratch[slot];
}

void *pinned(int slot, size_t bytes) {
    BX_REQUIRE(slot >= 0 && slot < kSlots, "pinned slot %d out of range", slot);
    if (bytes == 0) bytes = 256;
    if (g_pinned_bytes[slot] < bytes) {
        if (g_pinned[slot]) {
            BX_HIP(hipStreamSynchronize(stream()));
            BX_HIP(hipHostFree(g_pinned[slot]));
        }
        size_t grow = bytes + bytes / 4;
        BX_HIP(hipHostMalloc(&g_pinned[slot], grow, hipHostMallocDefault));
        g_pinned_bytes[slot] = grow;
    }
    return g_pinned[slot];
}

namespace {
constexpr int kStage = 4;
hipEvent_t g_stage_ev[kStage] = {};
bool g_stage_pending[kStage] = {};
int g_stage_next = 0;
}  // namespace

void upload_async(void *dst, const void *src, size_t bytes) {
    if (bytes == 0) return;
    hipStream_t s = stream();
    int slot = g_stage_next;
    g_stage_next = (g_stage_next + 1) % kStage;
    if (!g_stage_ev[slot]) BX_HIP(hipEventCreateWithFlags(&g_stage_ev[slot], hipEventDisableTiming));
    if (g_stage_pending[slot]) BX_HIP(hipEventSynchronize(g_stage_ev[slot]));
    void *h = pinned(kSlots - 1 - slot, bytes);
    memcpy(h, src, bytes);
    BX_HIP(hipMemcpyAsync(dst, h, bytes, hipMemcpyHostToDevice, s));
    BX_HIP(hipEventRecord(g_stage_ev[slot], s));
    g_stage_pending[slot] = true;
}

void download_sync(void *dst, const void *src, size_t bytes) {
    if (bytes == 0) return;
    hipStream_t s = stream();
    void *h = pinned(kSlots - 1 - kStage, bytes);
    BX_HIP(hipMemcpyAsync(h, src, bytes, hipMemcpyDeviceToHost, s));
    BX_HIP(hipStreamSynchronize(s));
    memcpy(dst, h, bytes);
}

namespace {
hipEvent_t g_download_ev = nullptr;
// a pinned slot of its own, so that work enqueued between download_start and
// download_finish (a caller's operator among it) may use download_sync
constexpr int kAheadSlot = kSlots - 2 - kStage;
}  // namespace

void download_start(const void *src, size_t bytes) {
    hipStream_t s = stream();
    void *h = pinned(kAheadSlot, bytes ? bytes : 1);
    if (!g_download_ev) BX_HIP(hipEventCreateWithFlags(&g_download_ev, hipEventDisableTiming));
    if (bytes) BX_HIP(hipMemcpyAsync(h, src, bytes, hipMemcpyDeviceToHost, s));
    BX_HIP(hipEventRecord(g_download_ev, s));
}

void download_finish(void *dst, size_t bytes) {
    BX_HIP(hipEventSynchronize(g_download_ev));
    if (bytes) memcpy(dst, pinned(kAheadSlot, bytes), bytes);
}

ColMap colmap_of(const hip_Multi_Vector *x) {
    ColMap m;
    BX_REQUIRE(x->num_active_vectors <= kMaxCols,
               "block operations support at most %d active columns (got %d)", kMaxCols,
               (int)x->num_active_vectors);
    m.n = x->num_active_vectors;
    for (int i = 0; i < m.n; i++) m.idx[i] = x->active_indices[i];
    return m;
}

ColMap colmap_identity(int n) {
    ColMap m;
    BX_REQUIRE(n <= kMaxCols, "at most %d columns (got %d)", kMaxCols, n);
    m.n = n;
    for (int i = 0; i < n; i++) m.idx[i] = i;
    return m;
}

}  // namespace bx

extern "C" {

int blopex_amd_init(void) {
    bx::stream();
    // the device address of the no-allocation entry point's array, looked up
    // once per device here rather than at its first call
    return bx::noalloc_setup() ? 0 : 1;
}

void *blopex_amd_stream(void) { return (void *)bx::stream(); }

int blopex_amd_synchronize(void) {
    hipError_t e = hipStreamSynchronize(bx::stream());
    if (e != hipSuccess) {
        bx::set_error("hipStreamSynchronize: %s", hipGetErrorString(e));
        return 1;
    }
    return 0;
}

const char *blopex_amd_last_error(void) { return bx::g_err; }

int blopex_amd_release_workspace(void) {
    if (hipStreamSynchronize(bx::stream()) != hipSuccess) return 1;
    for (int i = 0; i < bx::kSlots; i++) {
        if (bx::g_scratch[i]) hipFree(bx::g_scratch[i]);
        bx::g_scratch[i] = nullptr;
        bx::g_scratch_bytes[i] = 0;
    }
    return 0;
}
}
