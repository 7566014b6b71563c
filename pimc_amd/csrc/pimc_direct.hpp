// pimc_direct.hpp -- kernel dispatch straight into a user-mode AQL queue, for the
// launches whose latency a Markov chain waits on (short Delta V requests of a
// context, the serving form's Delta V batches).
//
// hipLaunchKernel spends ~3 us of host time per launch here and its tail reaches
// 7-11 us; a dispatch packet written into an HSA queue of our own, with the
// kernel arguments in device memory written through the PCIe BAR, costs ~1.9 us
// of host time (most of it the read-back that orders the argument writes before
// the doorbell) with a flat tail.  A context's call is synchronous: it waits for
// the kernel's published results before it returns, and the context's HIP
// stream is quiet when the packet is written, so the two queues never run work
// of one context at the same time.  The server sends a batch's two kernels as
// packets on a queue of the slot's own, in order, and retires the batch from
// its published records.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace pimc {

// A kernel as the packet processor sees it: its code object entry and the
// segment sizes the packet carries.
struct DirectKernel {
    uint64_t object = 0;
    uint32_t kernarg_bytes = 0, group_bytes = 0;
    bool ok() const { return object != 0; }
};

// One caller's ring of kernel-argument blocks in device memory.  Slot k is
// rewritten kDirectSlots dispatches later; the caller waits for each dispatch's
// results (published after every argument read) before its next one, so a
// slot is never rewritten while a kernel can still read it.
constexpr int kDirectSlots = 64;
constexpr size_t kDirectSlotBytes = 4096;
struct DirectRing {
    char* base = nullptr;
    int device = -1;
    unsigned next = 0;
    void* own_queue = nullptr;   // a queue of the ring's own (in order), or null: the device's shared queue
};

// A piece of the argument block: bytes [offset, offset + bytes) come from src.
struct ArgPiece {
    const void* src;
    uint32_t offset, bytes;
};

// Sets up the device's queue on first use and allocates the ring; false when
// direct dispatch is unavailable (the caller launches through HIP instead).
// own_queue: the ring gets an in-order queue of its own, for one thread whose
// packets depend on each other; otherwise it shares the device's queue, whose
// packets the caller orders by waiting for each result.
bool direct_ring_create(int device, DirectRing* r, bool own_queue = false);
// Waits for every packet queued so far on the ring's queue, then frees the ring
// (and its own queue).
void direct_ring_release(DirectRing* r);
// Device memory the host writes through the BAR (made visible by
// direct_flush), for data a packet's kernel reads; null when unavailable.
void* direct_alloc_visible(int device, size_t bytes);
void direct_free_visible(void* p);
// Resolves a __global__ function (its host stub) in the code object this
// library registered; false unless the code object's argument block is exactly
// 'bytes' long and the kernel needs no scratch.
bool direct_kernel(int device, const void* host_stub, hipStream_t stream, uint32_t bytes, DirectKernel* k);
// Writes the argument pieces into the ring's next slot, makes them visible to
// the device, and queues one dispatch of grid_x workgroups of block_x lanes.
void direct_dispatch(DirectRing* r, const DirectKernel& k, const ArgPiece* pieces, int n_pieces, uint32_t grid_x,
                     uint32_t block_x);
// The same in three steps, for several packets behind one flush: copy the
// pieces into the next slot (returns its address), make every BAR write so far
// visible to the device, queue the packet.
char* direct_stage(DirectRing* r, const ArgPiece* pieces, int n_pieces);
void direct_flush(int device);
void direct_submit(DirectRing* r, const DirectKernel& k, const char* kernargs, uint32_t grid_x, uint32_t block_x);

}  // namespace pimc
