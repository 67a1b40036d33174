// psa_runtime.hip - device initialisation, carrier table, per-thread streams.
#include <math.h>
#include <atomic>
#include "psa_internal.h"

namespace psa {

static uint32_t h_carr[256];
static uint32_t *d_carr = nullptr;
static std::once_flag g_init_once;
static std::atomic<bool> g_ready{false};

// Carrier table of the mixing LUT, src/sdr_func.c:121-124:
// carr = (int8)floor(cos/sin(-2 pi i / 256) / (1/11.2f) + 0.5).  The LUT of
// the reference, mix_tbl[(j << 8) + i] = cpx8(j) * carr(i) stored as int8
// (src/sdr_func.c:125-130), is evaluated from this table in the kernels.
static void build_carrier()
{
    const double PI_D = 3.1415926535897932;
    for (int i = 0; i < 256; i++) {
        int8_t cI = (int8_t)floor(cos(-2.0 * PI_D * i / 256) / PSA_CSCALE + 0.5);
        int8_t cQ = (int8_t)floor(sin(-2.0 * PI_D * i / 256) / PSA_CSCALE + 0.5);
        h_carr[i] = (uint32_t)(uint8_t)cI | ((uint32_t)(uint8_t)cQ << 8);
    }
}

void ensure_init()
{
    std::call_once(g_init_once, [] {
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess || n < 1) {
            PSA_FAIL("no HIP device available (libpocketsdr_amd requires an MI355X / gfx950 GPU)");
        }
        build_carrier();
        PSA_CHECK(hipMalloc(&d_carr, sizeof(h_carr)));
        PSA_CHECK(hipMemcpy(d_carr, h_carr, sizeof(h_carr), hipMemcpyHostToDevice));
        g_ready.store(true);
    });
}

bool init_done()
{
    return g_ready.load();
}

const uint32_t *host_carrier()
{
    ensure_init();
    return h_carr;
}

const uint32_t *dev_carrier()
{
    ensure_init();
    return d_carr;
}

void *Scratch::get(size_t bytes)
{
    if (bytes == 0) bytes = 1;
    if (bytes > size) {
        if (ptr) PSA_CHECK(hipFree(ptr));
        size_t n = bytes + bytes / 4;
        PSA_CHECK(hipMalloc(&ptr, n));
        size = n;
    }
    return ptr;
}

void *ThreadCtx::host_pinned(size_t bytes)
{
    if (bytes > pinned_size) {
        if (pinned) PSA_CHECK(hipHostFree(pinned));
        size_t n = bytes + bytes / 4;
        PSA_CHECK(hipHostMalloc(&pinned, n, hipHostMallocDefault));
        pinned_size = n;
    }
    return pinned;
}

// Thread-local context: created on first use by each calling thread.  The
// stream has the default priority.  Buffers are released with the thread.
struct CtxHolder {
    ThreadCtx c;
    ~CtxHolder()
    {
        for (auto &s : c.s) if (s.ptr) (void)hipFree(s.ptr);
        if (c.pinned) (void)hipHostFree(c.pinned);
        if (c.stream) (void)hipStreamDestroy(c.stream);
    }
};

ThreadCtx &tctx()
{
    thread_local CtxHolder h;
    if (!h.c.stream) {
        ensure_init();
        PSA_CHECK(hipStreamCreateWithFlags(&h.c.stream, hipStreamNonBlocking));
    }
    return h.c;
}

}  // namespace psa

extern "C" const char *sdr_amd_version(void)
{
    return "pocketsdr_amd 0.1 (gfx950)";
}
