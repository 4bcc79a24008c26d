// lx_dma.cpp -- device-to-host copies through hsa_amd_memory_async_copy
// (the reasons are in lx_dma.hpp).
#include "lx_dma.hpp"

#include <hsa/hsa_ext_amd.h>

#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <deque>
#include <mutex>
#include <thread>

#include "lx_util.hpp"

namespace lx {
namespace {

struct Hsa {
    bool ok = false;
    hsa_agent_t cpu{0};
};

const Hsa &hsa() {
    static Hsa h;
    static std::once_flag once;
    std::call_once(once, [] {
        if (hsa_init() != HSA_STATUS_SUCCESS) return;   // reference-counted: HIP holds the runtime open already
        (void)hsa_iterate_agents(
            [](hsa_agent_t a, void *d) {
                hsa_device_type_t t;
                if (hsa_agent_get_info(a, HSA_AGENT_INFO_DEVICE, &t) == HSA_STATUS_SUCCESS && t == HSA_DEVICE_TYPE_CPU) {
                    static_cast<Hsa *>(d)->cpu = a;
                    return HSA_STATUS_INFO_BREAK;
                }
                return HSA_STATUS_SUCCESS;
            },
            &h);
        h.ok = h.cpu.handle != 0;
    });
    return h;
}

bool agent_is(hsa_agent_t a, hsa_device_type_t want) {
    hsa_device_type_t t;
    return a.handle && hsa_agent_get_info(a, HSA_AGENT_INFO_DEVICE, &t) == HSA_STATUS_SUCCESS && t == want;
}

// The allocation holding [p, p + bytes) and the agent owning it, when the
// HSA runtime allocated it and its owner is of type `want`.
bool owned_range(const void *p, size_t bytes, hsa_device_type_t want, hsa_agent_t *owner) {
    hsa_amd_pointer_info_t info{};
    info.size = sizeof(info);
    if (hsa_amd_pointer_info(p, &info, nullptr, nullptr, nullptr) != HSA_STATUS_SUCCESS) return false;
    if (info.type != HSA_EXT_POINTER_TYPE_HSA && info.type != HSA_EXT_POINTER_TYPE_HSA_VMEM) return false;
    if (!agent_is(info.agentOwner, want)) return false;
    // hipHostMalloc memory is mapped at one address for the host and the GPU
    const char *base = static_cast<const char *>(want == HSA_DEVICE_TYPE_CPU ? info.hostBaseAddress : info.agentBaseAddress);
    const char *q = static_cast<const char *>(p);
    if (!base || q < base || q + bytes > base + info.sizeInBytes) return false;
    *owner = info.agentOwner;
    return true;
}

bool agents(const void *dst, const void *src, size_t bytes, hsa_agent_t *cpu, hsa_agent_t *gpu) {
    if (!hsa().ok || !dst || !src || !bytes) return false;
    return owned_range(dst, bytes, HSA_DEVICE_TYPE_CPU, cpu) && owned_range(src, bytes, HSA_DEVICE_TYPE_GPU, gpu);
}

lx_status ensure_signal(DmaSlot *slot) {
    if (slot->sig.handle) return LX_OK;
    if (hsa_signal_create(0, 0, nullptr, &slot->sig) != HSA_STATUS_SUCCESS) {
        slot->sig.handle = 0;
        return fail(LX_EDEVICE, "hsa_signal_create failed");
    }
    return LX_OK;
}

// Issue the copy (completion: the slot's signal drops to 0); a copy the
// runtime refuses is made with hipMemcpy instead.
void issue(DmaSlot *slot, void *dst, const void *src, size_t bytes) {
    hsa_agent_t cpu, gpu;
    if (agents(dst, src, bytes, &cpu, &gpu) &&
        hsa_amd_memory_async_copy(dst, cpu, src, gpu, bytes, 0, nullptr, slot->sig) == HSA_STATUS_SUCCESS)
        return;
    if (hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost) != hipSuccess) {
        (void)hipGetLastError();
        slot->failed = true;
    }
    hsa_signal_store_screlease(slot->sig, 0);
}

// One lane stores the mark at system scope (the host polls it).
__global__ void k_dma_mark(uint64_t *flag, uint64_t seq) {
    if (threadIdx.x == 0) __hip_atomic_store(flag, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

void wait_mark(const uint64_t *flag, uint64_t seq) {
    for (int i = 0; __atomic_load_n(flag, __ATOMIC_ACQUIRE) < seq; ++i)
        std::this_thread::sleep_for(std::chrono::microseconds(i < 200 ? 5 : 50));
}

struct Job {
    DmaSlot *slot;
    uint64_t seq;
    void *dst;
    const void *src;
    size_t bytes;
};

// One worker per process; never destroyed (it blocks on its queue at exit).
struct Worker {
    std::mutex mu;
    std::condition_variable cv;
    std::deque<Job> q;
    std::thread th;
};

void work(Worker *w) {
    for (;;) {
        Job j;
        {
            std::unique_lock<std::mutex> lk(w->mu);
            w->cv.wait(lk, [&] { return !w->q.empty(); });
            j = w->q.front();
            w->q.pop_front();
        }
        wait_mark(j.slot->flag, j.seq);
        issue(j.slot, j.dst, j.src, j.bytes);
    }
}

Worker *worker() {
    static Worker *w = [] {
        Worker *x = new Worker;
        x->th = std::thread(work, x);
        x->th.detach();
        return x;
    }();
    return w;
}

}  // namespace

bool dma_eligible(const void *dst, const void *src, size_t bytes) {
    hsa_agent_t cpu, gpu;
    return agents(dst, src, bytes, &cpu, &gpu);
}

lx_status dma_mark(DmaSlot *slot, hipStream_t stream) {
    lx_status st = ensure_signal(slot);
    if (st) return st;
    if (!dma_idle(slot)) return fail(LX_EDEVICE, "copy slot still busy");
    if (!slot->flag) {
        void *p = nullptr;
        if (hipHostMalloc(&p, sizeof(uint64_t), hipHostMallocCoherent) != hipSuccess) {
            (void)hipGetLastError();
            return fail(LX_ENOMEM, "pinned flag allocation failed");
        }
        slot->flag = static_cast<uint64_t *>(p);
        __atomic_store_n(slot->flag, slot->seq, __ATOMIC_RELEASE);
    }
    ++slot->seq;
    hipLaunchKernelGGL(k_dma_mark, dim3(1), dim3(64), 0, stream, slot->flag, slot->seq);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(LX_EDEVICE, "mark launch failed: %s", hipGetErrorString(e));
    return LX_OK;
}

lx_status dma_enqueue(DmaSlot *slot, void *dst, const void *src, size_t bytes) {
    if (!slot->flag || !dma_idle(slot)) return fail(LX_EDEVICE, "copy slot not marked or still busy");
    slot->failed = false;
    hsa_signal_store_screlease(slot->sig, 1);
    Worker *w = worker();
    {
        std::lock_guard<std::mutex> lk(w->mu);
        w->q.push_back(Job{slot, slot->seq, dst, src, bytes});
    }
    w->cv.notify_one();
    return LX_OK;
}

lx_status dma_copy_marked(DmaSlot *slot, void *dst, const void *src, size_t bytes) {
    if (!slot->flag || !dma_idle(slot)) return fail(LX_EDEVICE, "copy slot not marked or still busy");
    slot->failed = false;
    hsa_signal_store_screlease(slot->sig, 1);
    wait_mark(slot->flag, slot->seq);
    issue(slot, dst, src, bytes);
    return dma_wait(slot);
}

bool dma_idle(const DmaSlot *slot) { return !slot->sig.handle || hsa_signal_load_scacquire(slot->sig) == 0; }

lx_status dma_wait(DmaSlot *slot) {
    if (!slot->sig.handle) return LX_OK;
    while (hsa_signal_wait_scacquire(slot->sig, HSA_SIGNAL_CONDITION_LT, 1, UINT64_MAX, HSA_WAIT_STATE_BLOCKED) >= 1) {
    }
    if (slot->failed) {
        slot->failed = false;
        return fail(LX_EDEVICE, "device-to-host copy failed");
    }
    return LX_OK;
}

void dma_release(DmaSlot *slot) {
    if (slot->sig.handle) {
        (void)dma_wait(slot);
        (void)hsa_signal_destroy(slot->sig);
        slot->sig.handle = 0;
    }
    if (slot->flag) (void)hipHostFree(slot->flag);
    slot->flag = nullptr;
}

}  // namespace lx
