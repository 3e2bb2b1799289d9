// wtv_order.hip -- locality renumbering of a single-graph solve.
//
// Every kernel of the solve gathers the rows of a vertex's neighbours, so it
// runs near HBM speed only when neighbours sit near each other in memory.  A
// mesher that exports its cells in container order (the reference exports
// CGAL's finite cells as they are stored, wtivo_core.cpp:130-153) can hand
// over a graph whose arcs jump across the whole id range: the 77.5M-tet
// Delaunay graph of 12M genus-3 points solves in 574 ms with its tets
// shuffled against 333 ms in the tetrahedraliser's spatial order, the same
// flow and labels.  pr_solve therefore samples the arcs first and, when
// fewer than half of them join ids less than 4096 apart, solves a renumbered
// copy of the graph:
//   1. clusters: every M-th node (M >= 1024) seeds a cluster, and a
//      multi-source BFS over the adjacency (every arc, whatever its residual)
//      gives each node its distance to the nearest seeds and the smallest
//      seed index among them -- an atomic minimum, the same result on every
//      run;
//   2. order: a stable radix sort (rocPRIM) by (cluster, distance), ties by
//      input id, so a cluster's nodes are contiguous, nearest-first;
//   3. the copy: rows gathered in the new order, neighbour ids mapped;
//   4. the solve of the copy, then its residuals, terminals and partition
//      scattered back to the caller's ids.
// The maximum flow value and the partition (the minimal sink set) do not
// depend on the numbering, so the results equal a solve in the input order;
// the final residuals are a valid maximum preflow of the caller's graph
// (which preflow the discharge ends in depends on the order it visits).
// A renumbering whose workspace cannot be allocated is skipped: the solve
// then runs in the input order.
#include <hip/hip_runtime.h>

#include <rocprim/device/device_radix_sort.hpp>

#include <chrono>
#include <string>

#include "wtv_internal.h"

namespace wtv {
namespace {

constexpr int kOrdTPB = 256;
constexpr uint32_t kUnreached = 0xFFFFFFFFu;
constexpr uint32_t kLocalSpan = 4096u;     // an arc is local below this id distance
constexpr uint32_t kSampleRows = 65536u;   // rows the disorder test reads
constexpr uint32_t kMinAutoNodes = 1u << 20;  // smaller graphs stay in the caches
constexpr int kMaxLevels = 254;            // distances fit the key's low byte

inline void cko(hipError_t e, const char* where) {
    if (e == hipSuccess) return;
    fail(WTV_EDEVICE, std::string(kPrefix) + "HIP at " + where + ": " + hipGetErrorString(e));
}

inline int grid_for(uint64_t m) {
    return static_cast<int>(std::max<uint64_t>(1, std::min<uint64_t>((m + kOrdTPB - 1) / kOrdTPB, 16384)));
}

__device__ __forceinline__ bool valid_id(int u, uint32_t n) { return u >= 0 && static_cast<uint32_t>(u) < n; }

// Sampled rows v = i * stride: cnt[0] += valid arcs, cnt[1] += arcs to ids
// less than kLocalSpan away.
__global__ void __launch_bounds__(kOrdTPB) k_locality_sample(uint32_t n, uint32_t rows, uint32_t stride,
                                                             const int4* __restrict__ nbr4,
                                                             unsigned long long* __restrict__ cnt) {
    __shared__ unsigned int s_valid, s_local;
    if (threadIdx.x == 0) s_valid = s_local = 0u;
    __syncthreads();
    unsigned int valid = 0, local = 0;
    for (uint32_t i = blockIdx.x * kOrdTPB + threadIdx.x; i < rows; i += gridDim.x * kOrdTPB) {
        const uint32_t v = i * stride;
        const int4 nb = nbr4[v];
        const int ids[4] = {nb.x, nb.y, nb.z, nb.w};
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if (!valid_id(ids[k], n)) continue;
            ++valid;
            const uint32_t u = static_cast<uint32_t>(ids[k]);
            if ((u > v ? u - v : v - u) < kLocalSpan) ++local;
        }
    }
    atomicAdd(&s_valid, valid);
    atomicAdd(&s_local, local);
    __syncthreads();
    if (threadIdx.x == 0) {
        atomicAdd(&cnt[0], static_cast<unsigned long long>(s_valid));
        atomicAdd(&cnt[1], static_cast<unsigned long long>(s_local));
    }
}

// w[v] = distance << 24 | cluster; seeds (v % M == 0) are at distance 0.
__global__ void __launch_bounds__(kOrdTPB) k_cluster_seed(uint32_t n, uint32_t M, uint32_t* __restrict__ w) {
    for (uint32_t v = blockIdx.x * kOrdTPB + threadIdx.x; v < n; v += gridDim.x * kOrdTPB)
        w[v] = v % M == 0u ? v / M : kUnreached;
}

// One BFS level L, pushed from its frontier (the nodes at distance L): each
// offers distance L + 1 and its cluster to its neighbours with an atomic
// minimum on the packed word, so a node reached at this level ends with the
// smallest cluster among its frontier neighbours, whatever the timing.  A
// node already reached has a smaller word and keeps it; a node reached in
// this pass is at distance L + 1 and not part of this pass's frontier.  A
// plain load first skips the atomic where it cannot lower the word.
// *grown counts the nodes reached (the atomic that found them unreached).
__global__ void __launch_bounds__(kOrdTPB) k_cluster_level(uint32_t n, uint32_t L, const int4* __restrict__ nbr4,
                                                           uint32_t* __restrict__ w, unsigned int* grown) {
    unsigned int mine = 0;
    for (uint32_t v = blockIdx.x * kOrdTPB + threadIdx.x; v < n; v += gridDim.x * kOrdTPB) {
        const uint32_t x = w[v];
        if (x == kUnreached || (x >> 24) != L) continue;
        const uint32_t offer = ((L + 1u) << 24) | (x & 0xFFFFFFu);
        const int4 nb = nbr4[v];
        const int ids[4] = {nb.x, nb.y, nb.z, nb.w};
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if (!valid_id(ids[k], n)) continue;
            uint32_t* t = &w[ids[k]];
            if (__hip_atomic_load(t, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) <= offer) continue;
            if (atomicMin(t, offer) == kUnreached) ++mine;
        }
    }
    // one atomic per wave
    const unsigned long long lanes = __ballot(mine != 0u);
    unsigned int sum = mine;
    for (int off = 32; off > 0; off >>= 1) sum += __shfl_xor(sum, off);
    if (lanes && (threadIdx.x & 63u) == 0u) atomicAdd(grown, sum);
}

// In place: w[v] becomes the sort key (cluster << 8 | distance; unreached
// nodes after every cluster), ids[v] = v the sort's values.
__global__ void __launch_bounds__(kOrdTPB) k_order_keys(uint32_t n, uint32_t clusters, uint32_t* __restrict__ w,
                                                        uint32_t* __restrict__ ids) {
    for (uint32_t v = blockIdx.x * kOrdTPB + threadIdx.x; v < n; v += gridDim.x * kOrdTPB) {
        const uint32_t x = w[v];
        w[v] = x == kUnreached ? clusters << 8 : ((x & 0xFFFFFFu) << 8) | (x >> 24);
        ids[v] = v;
    }
}

__global__ void __launch_bounds__(kOrdTPB) k_invert(uint32_t n, const uint32_t* __restrict__ perm,
                                                    uint32_t* __restrict__ inv) {
    for (uint32_t j = blockIdx.x * kOrdTPB + threadIdx.x; j < n; j += gridDim.x * kOrdTPB) inv[perm[j]] = j;
}

// Row j of the copy is input row perm[j], its neighbour ids mapped; an id
// out of range stays as it is, so the solve reports it as it would have.
__global__ void __launch_bounds__(kOrdTPB) k_gather_rows(uint32_t n, const uint32_t* __restrict__ perm,
                                                         const uint32_t* __restrict__ inv,
                                                         const int4* __restrict__ nbr4, const uint4* __restrict__ res4,
                                                         const int32_t* __restrict__ term, int4* __restrict__ nbr_o,
                                                         uint4* __restrict__ res_o, int32_t* __restrict__ term_o) {
    for (uint32_t j = blockIdx.x * kOrdTPB + threadIdx.x; j < n; j += gridDim.x * kOrdTPB) {
        const uint32_t o = perm[j];
        int4 nb = nbr4[o];
        auto map = [&](int u) { return valid_id(u, n) ? static_cast<int>(inv[u]) : u; };
        nb.x = map(nb.x);
        nb.y = map(nb.y);
        nb.z = map(nb.z);
        nb.w = map(nb.w);
        nbr_o[j] = nb;
        res_o[j] = res4[o];
        term_o[j] = term[o];
    }
}

__global__ void __launch_bounds__(kOrdTPB) k_scatter_rows(uint32_t n, const uint32_t* __restrict__ perm,
                                                          const uint4* __restrict__ res_o,
                                                          const int32_t* __restrict__ term_o,
                                                          const uint8_t* __restrict__ part_o, uint4* __restrict__ res4,
                                                          int32_t* __restrict__ term, uint8_t* __restrict__ part) {
    for (uint32_t j = blockIdx.x * kOrdTPB + threadIdx.x; j < n; j += gridDim.x * kOrdTPB) {
        const uint32_t o = perm[j];
        res4[o] = res_o[j];
        term[o] = term_o[j];
        part[o] = part_o[j];
    }
}

// Device memory of the renumbering; freed only after the work queued on
// the stream that uses it has finished, failures included.
struct OrderBuf {
    hipStream_t s;
    void* p = nullptr;
    explicit OrderBuf(hipStream_t st) : s(st) {}
    ~OrderBuf() {
        if (!p) return;
        (void)hipStreamSynchronize(s);
        (void)hipFree(p);
    }
};

inline size_t up256(size_t x) { return (x + 255u) / 256u * 256u; }

// Whether the sampled arcs are mostly long (the disorder test).
bool disordered(uint32_t n, const int32_t* d_nbr, hipStream_t s) {
    OrderBuf b(s);
    if (hipMalloc(&b.p, 16) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    auto* cnt = static_cast<unsigned long long*>(b.p);
    cko(hipMemsetAsync(cnt, 0, 16, s), "order sample clear");
    const uint32_t rows = std::min(n, kSampleRows);
    const uint32_t stride = n / rows;
    k_locality_sample<<<grid_for(rows), kOrdTPB, 0, s>>>(n, rows, stride, reinterpret_cast<const int4*>(d_nbr), cnt);
    cko(hipGetLastError(), "order sample");
    unsigned long long h[2] = {0, 0};
    cko(hipMemcpyAsync(h, cnt, 16, hipMemcpyDeviceToHost, s), "order sample D2H");
    cko(hipStreamSynchronize(s), "order sample");
    return h[0] > 0 && 2u * h[1] < h[0];
}

}  // namespace

void pr_solve(uint32_t n, int32_t* d_nbr, uint32_t* d_res, int32_t* d_term, uint8_t* partition,
              bool partition_on_device, const wtv_params& prm, wtv_stats& st, void* stream) {
    using Clock = std::chrono::steady_clock;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const bool want = n > 1u && n <= 0x7FFFFFFFu && d_nbr && d_res && d_term && partition &&
                      (prm.renumber > 0 || (prm.renumber == 0 && n >= kMinAutoNodes));
    const auto t0 = Clock::now();
    if (!want) {
        pr_solve_in_order(n, d_nbr, d_res, d_term, partition, partition_on_device, prm, st, stream);
        return;
    }
    if (prm.renumber == 0 && !disordered(n, d_nbr, s)) {
        const double test_ms = std::chrono::duration<double, std::milli>(Clock::now() - t0).count();
        pr_solve_in_order(n, d_nbr, d_res, d_term, partition, partition_on_device, prm, st, stream);
        st.t_order_ms = test_ms;
        return;
    }
    // cluster size: at least 1024 nodes, and few enough clusters for 24 bits
    const uint32_t M = std::max<uint32_t>(1024u, (n + 0xFFFFFEu) / 0xFFFFFFu);
    const uint32_t clusters = (n + M - 1u) / M;
    unsigned end_bit = 8u;
    while ((1ull << (end_bit - 8u)) <= clusters) ++end_bit;  // keys < (clusters + 1) << 8
    // workspace: keys 4n (then the inverse), sorted keys 4n (then a host
    // partition's staging), ids 4n, perm 4n, rows 36n, partition n, counter
    const size_t nz = n;
    size_t sort_bytes = 0;
    cko(rocprim::radix_sort_pairs(nullptr, sort_bytes, static_cast<const uint32_t*>(nullptr),
                                  static_cast<uint32_t*>(nullptr), static_cast<const uint32_t*>(nullptr),
                                  static_cast<uint32_t*>(nullptr), nz, 0u, end_bit, s),
        "radix sort size");
    const size_t o_w = 0, o_keys = o_w + up256(4 * nz), o_ids = o_keys + up256(4 * nz),
                 o_perm = o_ids + up256(4 * nz), o_nbr = o_perm + up256(4 * nz), o_res = o_nbr + up256(16 * nz),
                 o_term = o_res + up256(16 * nz), o_part = o_term + up256(4 * nz), o_cnt = o_part + up256(nz),
                 o_sort = o_cnt + 256, total = o_sort + up256(sort_bytes);
    OrderBuf b(s);
    if (hipMalloc(&b.p, total) != hipSuccess) {
        (void)hipGetLastError();  // no room for the copy: solve in the input order
        pr_solve_in_order(n, d_nbr, d_res, d_term, partition, partition_on_device, prm, st, stream);
        return;
    }
    char* base = static_cast<char*>(b.p);
    auto* w = reinterpret_cast<uint32_t*>(base + o_w);
    auto* keys = reinterpret_cast<uint32_t*>(base + o_keys);
    auto* ids = reinterpret_cast<uint32_t*>(base + o_ids);
    auto* perm = reinterpret_cast<uint32_t*>(base + o_perm);
    auto* nbr_o = reinterpret_cast<int4*>(base + o_nbr);
    auto* res_o = reinterpret_cast<uint4*>(base + o_res);
    auto* term_o = reinterpret_cast<int32_t*>(base + o_term);
    auto* part_o = reinterpret_cast<uint8_t*>(base + o_part);
    auto* grown = reinterpret_cast<unsigned int*>(base + o_cnt);
    const int4* nbr4 = reinterpret_cast<const int4*>(d_nbr);
    const int g = grid_for(n);

    // 1. clusters: levels until none grows (a check every 4 levels)
    k_cluster_seed<<<g, kOrdTPB, 0, s>>>(n, M, w);
    cko(hipGetLastError(), "cluster seeds");
    for (int L = 0; L < kMaxLevels;) {
        cko(hipMemsetAsync(grown, 0, 4, s), "cluster count clear");
        const int batch = std::min(4, kMaxLevels - L);
        for (int i = 0; i < batch; ++i, ++L)
            k_cluster_level<<<g, kOrdTPB, 0, s>>>(n, static_cast<uint32_t>(L), nbr4, w, grown);
        cko(hipGetLastError(), "cluster level");
        unsigned int h = 0;
        cko(hipMemcpyAsync(&h, grown, 4, hipMemcpyDeviceToHost, s), "cluster count D2H");
        cko(hipStreamSynchronize(s), "cluster levels");
        if (h == 0u) break;
    }
    // 2. order
    k_order_keys<<<g, kOrdTPB, 0, s>>>(n, clusters, w, ids);
    cko(hipGetLastError(), "order keys");
    cko(rocprim::radix_sort_pairs(base + o_sort, sort_bytes, static_cast<const uint32_t*>(w), keys,
                                  static_cast<const uint32_t*>(ids), perm, nz, 0u, end_bit, s),
        "radix sort");
    // 3. the copy (the key arrays are free again: w holds the inverse)
    k_invert<<<g, kOrdTPB, 0, s>>>(n, perm, w);
    k_gather_rows<<<g, kOrdTPB, 0, s>>>(n, perm, w, nbr4, reinterpret_cast<const uint4*>(d_res), d_term, nbr_o, res_o,
                                        term_o);
    cko(hipGetLastError(), "order copy");
    const double order_ms = std::chrono::duration<double, std::milli>(Clock::now() - t0).count();
    // 4. solve the copy and scatter back
    pr_solve_in_order(n, reinterpret_cast<int32_t*>(nbr_o), reinterpret_cast<uint32_t*>(res_o), term_o, part_o, true,
                      prm, st, stream);
    const auto t1 = Clock::now();
    uint8_t* part_dst = partition_on_device ? partition : reinterpret_cast<uint8_t*>(keys);
    k_scatter_rows<<<g, kOrdTPB, 0, s>>>(n, perm, res_o, term_o, part_o, reinterpret_cast<uint4*>(d_res), d_term,
                                         part_dst);
    cko(hipGetLastError(), "order scatter");
    if (!partition_on_device) cko(hipMemcpyAsync(partition, part_dst, nz, hipMemcpyDeviceToHost, s), "partition D2H");
    cko(hipStreamSynchronize(s), "order scatter");
    st.t_order_ms = order_ms + std::chrono::duration<double, std::milli>(Clock::now() - t1).count();
    st.renumbered = 1u;
    st.solver_workspace_bytes += total;
}

}  // namespace wtv
