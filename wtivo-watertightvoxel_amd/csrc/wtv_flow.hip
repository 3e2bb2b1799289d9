// wtv_flow.hip -- maximum flow with conservation: the solve plus an excess
// return (the second phase of push-relabel).
//
// The solve (as the reference's, gpu_push_relabel_fast.cu:406-571) stops at a
// maximum preflow: every vertex that cannot reach the sink keeps whatever
// excess it holds.  That already fixes the sink flow and the minimum cut (the
// minimal sink set is unique), but not a flow: the source side does not
// conserve.  The excess goes back to the source here by solving a second
// problem with the same solver on the residual graph:
//
//   excess  e(v) = max(term0, 0) - sink flow(v) - sum_k (res0[v][k] - res[v][k])
//   source side (partition 0): term''(v) = e(v) - max(term(v), 0)
//       the flow the source still sends into v (term > 0 is never changed by
//       the solve) can take back up to term(v) at once; what is left of e(v)
//       is supply (term'' > 0), and unused return capacity is a sink
//       (term'' < 0) -- the reference's own terminal netting
//       (wtivo_gpupr_bindings.cpp:286-290)
//   sink side: term'' = 0 (no vertex there holds excess, and no residual arc
//       leads into it from the source side, so it takes no part)
//
// Every unit of excess came from the source, so it has a residual path back
// and the second solve delivers all supply.  Afterwards the source side's
// terminal is the flow the source still sends, max(0, -term''), and the sink
// side's terminal is restored.  The sink flow, the labels and the cut are the
// first solve's.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <vector>

#include "wtv_internal.h"

namespace wtv {
namespace {

constexpr int kB = 256;
constexpr int kSlots = 64;

inline void ckf(hipError_t e, const char* where) {
    if (e == hipSuccess) return;
    fail(WTV_EDEVICE, std::string(kPrefix) + "HIP at " + where + ": " + hipGetErrorString(e));
}

struct FlowCtl {
    unsigned long long excess[kSlots];  // sum of e(v) over the source side
    unsigned long long supply[kSlots];  // sum of max(term'', 0)
    int32_t bad;                        // 1 negative excess, 2 term'' outside int32
};

__device__ __forceinline__ unsigned long long wave_sum(unsigned long long x) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
    return x;
}

__global__ void __launch_bounds__(kB) k_excess_terms(uint32_t n, const uint4* __restrict__ res0,
                                                     const int32_t* __restrict__ term0,
                                                     const uint4* __restrict__ res, int32_t* __restrict__ term,
                                                     const uint8_t* __restrict__ part, int32_t* __restrict__ saved,
                                                     FlowCtl* ctl) {
    unsigned long long ex = 0, sup = 0;
    const uint32_t stride = gridDim.x * kB;
    for (uint32_t v = blockIdx.x * kB + threadIdx.x; v < n; v += stride) {
        const int32_t t = term[v];
        saved[v] = t;
        if (part[v] != 0) {  // sink side
            term[v] = 0;
            continue;
        }
        const int32_t t0 = term0[v];
        const uint4 a = res0[v], b = res[v];
        const long long out = (static_cast<long long>(a.x) - b.x) + (static_cast<long long>(a.y) - b.y) +
                              (static_cast<long long>(a.z) - b.z) + (static_cast<long long>(a.w) - b.w);
        const long long sinkf = t0 < 0 ? static_cast<long long>(t) - t0 : 0ll;
        const long long e = (t0 > 0 ? static_cast<long long>(t0) : 0ll) - sinkf - out;
        if (e < 0) {
            __hip_atomic_exchange(&ctl->bad, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            continue;
        }
        const long long tt = e - (t > 0 ? static_cast<long long>(t) : 0ll);
        if (tt > 2147483647ll || tt < -2147483647ll) {
            __hip_atomic_exchange(&ctl->bad, 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            continue;
        }
        term[v] = static_cast<int32_t>(tt);
        ex += static_cast<unsigned long long>(e);
        sup += tt > 0 ? static_cast<unsigned long long>(tt) : 0ull;
    }
    ex = wave_sum(ex);
    sup = wave_sum(sup);
    if ((threadIdx.x & 63) == 0) {
        const uint32_t sl = blockIdx.x % kSlots;
        if (ex) __hip_atomic_fetch_add(&ctl->excess[sl], ex, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (sup) __hip_atomic_fetch_add(&ctl->supply[sl], sup, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

__global__ void __launch_bounds__(kB) k_flow_finish(uint32_t n, const uint8_t* __restrict__ part,
                                                    int32_t* __restrict__ term, const int32_t* __restrict__ saved) {
    const uint32_t stride = gridDim.x * kB;
    for (uint32_t v = blockIdx.x * kB + threadIdx.x; v < n; v += stride)
        term[v] = part[v] != 0 ? saved[v] : (term[v] < 0 ? -term[v] : 0);
}

struct DevF {
    void* p = nullptr;
    ~DevF() { if (p) (void)hipFree(p); }
};

}  // namespace

void pr_solve_flow(uint32_t n, int32_t* d_nbr, uint32_t* d_res, int32_t* d_term, uint8_t* partition,
                   bool partition_on_device, const wtv_params& p, wtv_stats& st, int64_t& returned,
                   void* stream_v) {
    returned = 0;
    if (n == 0) {
        pr_solve(n, d_nbr, d_res, d_term, partition, partition_on_device, p, st, stream_v);
        return;
    }
    if (!d_nbr || !d_res || !d_term || !partition) fail(WTV_EINVAL, std::string(kPrefix) + "null graph pointer");
    auto s = static_cast<hipStream_t>(stream_v);
    DevF res0, term0, saved, part, part2, ctl;
    ckf(hipMalloc(&res0.p, static_cast<size_t>(n) * 16u), "flow res0");
    ckf(hipMalloc(&term0.p, static_cast<size_t>(n) * 4u), "flow term0");
    ckf(hipMalloc(&saved.p, static_cast<size_t>(n) * 4u), "flow saved");
    ckf(hipMalloc(&part2.p, n), "flow partition 2");
    ckf(hipMalloc(&ctl.p, sizeof(FlowCtl)), "flow ctl");
    uint8_t* d_part = partition;
    if (!partition_on_device) {
        ckf(hipMalloc(&part.p, n), "flow partition");
        d_part = static_cast<uint8_t*>(part.p);
    }
    ckf(hipMemcpyAsync(res0.p, d_res, static_cast<size_t>(n) * 16u, hipMemcpyDeviceToDevice, s), "res0 copy");
    ckf(hipMemcpyAsync(term0.p, d_term, static_cast<size_t>(n) * 4u, hipMemcpyDeviceToDevice, s), "term0 copy");
    ckf(hipMemsetAsync(ctl.p, 0, sizeof(FlowCtl), s), "flow ctl clear");

    // phase 1: the maximum preflow, its sink flow and labels
    pr_solve(n, d_nbr, d_res, d_term, d_part, true, p, st, stream_v);

    // phase 2: return the excess
    const int grid = static_cast<int>(std::min<uint64_t>((static_cast<uint64_t>(n) + kB - 1) / kB, 8192));
    auto* fc = static_cast<FlowCtl*>(ctl.p);
    k_excess_terms<<<grid, kB, 0, s>>>(n, static_cast<const uint4*>(res0.p), static_cast<const int32_t*>(term0.p),
                                       reinterpret_cast<const uint4*>(d_res), d_term, d_part,
                                       static_cast<int32_t*>(saved.p), fc);
    ckf(hipGetLastError(), "excess terms");
    FlowCtl h{};
    ckf(hipMemcpyAsync(&h, fc, sizeof(FlowCtl), hipMemcpyDeviceToHost, s), "flow ctl read");
    ckf(hipStreamSynchronize(s), "flow ctl sync");
    if (h.bad == 1) fail(WTV_EINVARIANT, std::string(kPrefix) + "excess return: negative excess (code=6)");
    if (h.bad == 2) fail(WTV_EOVERFLOW, std::string(kPrefix) + "excess return: a vertex's excess exceeds int32");
    unsigned long long excess = 0, supply = 0;
    for (int i = 0; i < kSlots; ++i) {
        excess += h.excess[i];
        supply += h.supply[i];
    }
    wtv_stats st2{};
    if (supply) pr_solve(n, d_nbr, d_res, d_term, static_cast<uint8_t*>(part2.p), true, p, st2, stream_v);
    if (static_cast<unsigned long long>(st2.raw_flow) != supply)
        fail(WTV_EINVARIANT, std::string(kPrefix) + "excess return left excess behind (code=7)");
    k_flow_finish<<<grid, kB, 0, s>>>(n, d_part, d_term, static_cast<const int32_t*>(saved.p));
    ckf(hipGetLastError(), "flow finish");
    if (!partition_on_device)
        ckf(hipMemcpyAsync(partition, d_part, n, hipMemcpyDeviceToHost, s), "partition D2H");
    ckf(hipStreamSynchronize(s), "flow sync");
    returned = static_cast<int64_t>(excess);
    st.solve_seconds += st2.solve_seconds;
    st.final_relabel_seconds += st2.final_relabel_seconds;
}

}  // namespace wtv
