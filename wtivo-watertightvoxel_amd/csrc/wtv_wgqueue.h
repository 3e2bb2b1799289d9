// wtv_wgqueue.h -- device helpers shared by the list kernels of libwtv_gpupr:
// device-scope adds, LDS-staged workgroup queues and contiguous shares.
// Device code only: include from .hip sources.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace wtv {

__device__ __forceinline__ uint32_t atom_add(uint32_t* p, uint32_t v) {
    return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ long long atom_add(long long* p, long long v) {
    return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned long long atom_add(unsigned long long* p, unsigned long long v) {
    return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// wave64: rank of this lane among the set lanes of m below it
__device__ __forceinline__ uint32_t lane_rank(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(m >> 32),
                                     __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(m), 0u));
}

// ------------------------------------------- LDS-staged workgroup queue
// Items are appended to an LDS buffer (one LDS atomic per wave, ranks from
// ballot + mbcnt), then the workgroup reserves space in the global list with
// a single device atomic and copies the items out coalesced.  The queue is
// flushed lazily -- when the next iteration could overflow it, and at kernel
// end -- because device-scope atomics on one address serialise at ~12 ns each
// (~83M/s) on MI355X: one flush per iteration of a 78M-vertex sweep would
// cost milliseconds.  At least 16 KiB of LDS per workgroup; the capacity holds
// five appends per thread (20 KiB at 1024 threads).
template <int TPB>
struct WgQueue {
    static constexpr int kCap = 5 * TPB > 4096 ? 5 * TPB : 4096;
    uint32_t cnt;
    uint32_t base;
    uint32_t buf[kCap];
};

// All active lanes of the wave must call this (pred may be false).
template <class Q>
__device__ __forceinline__ void q_push(Q& q, bool pred, uint32_t val) {
    const uint64_t m = __ballot(pred);
    if (m == 0) return;
    const uint32_t rank = lane_rank(m);
    const int leader = __ffsll(static_cast<unsigned long long>(m)) - 1;
    uint32_t off = 0;
    if (pred && rank == 0) off = atomicAdd(&q.cnt, static_cast<uint32_t>(__popcll(m)));
    off = __shfl(off, leader);
    if (pred) q.buf[off + rank] = val;
}

// Every thread of the workgroup must call this the same number of times.
template <class Q>
__device__ __forceinline__ void q_flush(Q& q, uint32_t* list, uint32_t* gcnt) {
    __syncthreads();
    const uint32_t c = q.cnt;
    __syncthreads();
    if (threadIdx.x == 0) {
        q.base = c ? atom_add(gcnt, c) : 0u;
        q.cnt = 0u;
    }
    __syncthreads();
    const uint32_t b = q.base;
    for (uint32_t i = threadIdx.x; i < c; i += blockDim.x) list[b + i] = q.buf[i];
    __syncthreads();
}

// Flushes only when fewer than `room` slots are free, i.e. when the next
// iteration (at most `room` appends) could overflow.  Uniform: every thread
// reads the count between two barriers.  Every thread of the workgroup must
// call this the same number of times, and call q_flush once at the end.
template <class Q>
__device__ __forceinline__ void q_flush_if(Q& q, uint32_t* list, uint32_t* gcnt, uint32_t room) {
    __syncthreads();
    const uint32_t c = q.cnt;
    __syncthreads();
    if (c + room > static_cast<uint32_t>(Q::kCap)) {
        if (threadIdx.x == 0) {
            q.base = atom_add(gcnt, c);
            q.cnt = 0u;
        }
        __syncthreads();
        const uint32_t b = q.base;
        for (uint32_t i = threadIdx.x; i < c; i += blockDim.x) list[b + i] = q.buf[i];
        __syncthreads();
    }
}

template <class Q>
__device__ __forceinline__ void q_init(Q& q) {
    if (threadIdx.x == 0) q.cnt = 0u;
    __syncthreads();
}

// A workgroup's contiguous share of m items (list entries or node ids) in
// steps of `step`: the items [b, e).  The appends of a workgroup, flushed
// lazily every few steps, then come from one region of its input, so the
// lists the next kernel reads keep neighbouring vertices in neighbouring
// entries and a wave's rows share cache lines.  A grid-stride loop would mix
// into each flush steps gridDim.x * step entries apart, and over many
// sub-rounds the lists drift towards a random order (on the 77.5M-node
// Delaunay graph, push visits from a random-order list take 3 to 6 times
// as long as from an id-ordered one).  With m <= gridDim.x * step the share
// is the single step at blockIdx.x * step, as in a grid-stride loop.
struct Span {
    uint64_t b, e;
};
__device__ __forceinline__ Span wg_span(uint64_t m, uint32_t step) {
    const uint64_t steps = (m + step - 1u) / step;
    const uint64_t per = (steps + gridDim.x - 1u) / gridDim.x;
    const uint64_t b = static_cast<uint64_t>(blockIdx.x) * per * step;
    const uint64_t e = b + per * step;
    return Span{b < m ? b : m, e < m ? e : m};
}

}  // namespace wtv
