// lx_noalloc.hip -- field readback that allocates nothing.
//
// lx_sim_read_field gathers a component into an FP64 staging buffer of the
// component's size (allocated on first use) and copies that to the host.
// lx_sim_read_field_noalloc gives the same values to callers that cannot
// allocate device memory (tight serving loops with a fixed footprint,
// harnesses that forbid allocation; it synchronizes the simulation's stream,
// so it is not itself recorded into a graph):
// it gathers the component in slices of whole x indices through `scratch`, a
// fixed device array that every GPU the library runs on holds from module
// load until the process ends, and copies each slice to its place in the
// caller's buffer.  The gather kernel is the one lx_sim_read_field uses, so
// the two entry points return identical values.
//
// `scratch` is a device symbol: host code never passes its name to a kernel
// or a copy.  Its address on each device is looked up once, with that device
// current, by hipGetSymbolAddress (noalloc_init, called by lx_ctx_create) and
// kept per device; kernels and copies take that address.  Every call on one
// GPU shares that GPU's array: calls that overlap on one GPU race.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>

#include "../../include/luxum.h"
#include "lx_sim.hpp"
#include "lx_util.hpp"

namespace lx {

constexpr size_t kScratchDoubles = size_t(4) << 20;   // 32 MiB of FP64
static __device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
static double *g_scratch[kMaxDevices];
static std::mutex g_scratch_mu;

// Look up this device's copy of `scratch` (the device must be current).
lx_status noalloc_init(int device) {
    if (device < 0 || device >= kMaxDevices) return fail(LX_EINVAL, "device %d out of range", device);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch[device]) return LX_OK;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return fail(LX_EDEVICE, "hipGetSymbolAddress(scratch) failed on device %d: %s", device, hipGetErrorString(e));
    g_scratch[device] = static_cast<double *>(p);
    return LX_OK;
}

}  // namespace lx

using namespace lx;

#define NA_TRY(x)                                                                                          \
    do {                                                                                                   \
        hipError_t _e = (x);                                                                               \
        if (_e != hipSuccess) return fail(LX_EDEVICE, "%s failed: %s", #x, hipGetErrorString(_e));          \
    } while (0)

extern "C" lx_status lx_sim_read_field_noalloc(lx_sim *s, int32_t comp, double *dst, size_t n) {
    if (!s || !dst) return fail(LX_EINVAL, "lx_sim_read_field_noalloc: null pointer");
    int ext[3];
    if (comp < 0 || comp > 5 || !comp_extent(s->N, s->planar, comp, ext)) return fail(LX_EINVAL, "bad component %d", comp);
    const size_t m = (size_t)ext[0] * ext[1] * ext[2];
    if (n != m) return fail(LX_EINVAL, "buffer holds %zu values, component has %zu", n, m);
    if (m == 0) return LX_OK;
    const size_t slice = (size_t)ext[1] * ext[2];   // one x index: contiguous in the host layout
    if (slice > kScratchDoubles)
        return fail(LX_EINVAL,
                    "lx_sim_read_field_noalloc: one x slice of component %d holds %zu values, more than the %zu "
                    "of the allocation-free scratch (use lx_sim_read_field)",
                    comp, slice, kScratchDoubles);
    const int dev = s->ctx->device;
    NA_TRY(hipSetDevice(dev));
    double *buf = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_scratch_mu);
        buf = (dev >= 0 && dev < kMaxDevices) ? g_scratch[dev] : nullptr;
    }
    if (!buf) return fail(LX_EDEVICE, "lx_sim_read_field_noalloc: scratch of device %d not set up", dev);
    member_enter(s);
    const double scale = (comp >= 3 ? 1.0 / s->eta0 : 1.0) / s->sscale;
    const int rows = (int)std::min<size_t>(kScratchDoubles / slice, (size_t)ext[0]);
    for (int i0 = 0; i0 < ext[0]; i0 += rows) {
        const int nr = std::min(rows, ext[0] - i0);
        // the gather of x indices [i0, i0+nr): the field shifted by i0 (x is
        // the fastest device axis), written from the start of the scratch
        if (s->fp16) {
            Grid<__half> g = make_grid<__half>(s);
            g.F += i0;
            launch_gather(g, comp, nr, ext[1], ext[2], scale, buf, s->cs);
        } else {
            Grid<float> g = make_grid<float>(s);
            g.F += i0;
            launch_gather(g, comp, nr, ext[1], ext[2], scale, buf, s->cs);
        }
        lx_status st = check_launch();
        if (st) return st;
        // the next slice's gather follows this copy on the same stream
        NA_TRY(hipMemcpyAsync(dst + (size_t)i0 * slice, buf, sizeof(double) * nr * slice, hipMemcpyDeviceToHost,
                              s->cs));
    }
    NA_TRY(hipStreamSynchronize(s->cs));
    return LX_OK;
}
