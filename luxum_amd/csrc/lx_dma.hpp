// lx_dma.hpp -- device-to-host copies on the GPU's DMA engine, issued through
// the HSA runtime instead of a HIP stream.
//
// A step graph with parallel branches (the CPML shell launch beside the
// interior launch) does not start while a hipMemcpyAsync from device to host
// on another stream is running, so a 134 MB snapshot copy (2.4 ms at the
// host link's 57 GB/s) used to cost its whole duration on the step stream.
// The same copy issued with hsa_amd_memory_async_copy runs beside the graphs.
//
// Only host memory that the HSA runtime allocated (hipHostMalloc) and device
// memory of the GPU take this path (dma_eligible); pageable or registered
// host memory keeps the HIP stream copy.
#pragma once
#include <hip/hip_runtime.h>
#include <hsa/hsa.h>

#include <atomic>
#include <cstddef>
#include <cstdint>

#include "../../include/luxum.h"

namespace lx {

// One copy in flight at a time.  The signal holds 1 from dma_enqueue until
// the copy has landed (or failed), then 0.  The source is ready once *flag
// (pinned host memory) reaches seq: dma_mark enqueues that store behind the
// producing kernel, so no thread waits on an event of a stream that a graph
// capture may be using (that would invalidate the capture).
struct DmaSlot {
    hsa_signal_t sig{0};
    uint64_t *flag = nullptr;
    uint64_t seq = 0;
    std::atomic<bool> failed{false};   // set by the copying thread before the signal drops
};

// dst is host memory allocated by the HSA runtime and src device memory, and
// both ranges lie inside their allocations.
bool dma_eligible(const void *dst, const void *src, size_t bytes);

// Enqueue on `stream` (behind the kernel writing the source) the store that
// marks the slot's source ready.  The slot must be idle.
lx_status dma_mark(DmaSlot *slot, hipStream_t stream);

// Copy bytes from src (device) to dst (host) once the last mark has landed:
// a worker thread waits for it and issues the copy.  dst and src must stay
// untouched until dma_wait(slot) returns.
lx_status dma_enqueue(DmaSlot *slot, void *dst, const void *src, size_t bytes);

// The same copy, issued from the calling thread once the mark has landed;
// returns when the copy has.
lx_status dma_copy_marked(DmaSlot *slot, void *dst, const void *src, size_t bytes);

// No copy of the slot is queued or running.
bool dma_idle(const DmaSlot *slot);

// Wait until the slot's copy has landed; LX_EDEVICE if it failed.
lx_status dma_wait(DmaSlot *slot);

// Wait, then free the slot's signal and flag.
void dma_release(DmaSlot *slot);

}  // namespace lx
