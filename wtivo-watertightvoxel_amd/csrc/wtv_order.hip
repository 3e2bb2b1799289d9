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
#include "wtv_wgqueue.h"

namespace wtv {
namespace {

constexpr int kOrdTPB = 256;
constexpr uint32_t kUnreached = 0xFFFFFFFFu;
constexpr uint32_t kLocalSpan = 4096u;     // an arc is local below this id distance
constexpr uint32_t kSampleRows = 65536u;   // rows the disorder test reads
constexpr uint32_t kMinAutoNodes = 1u << 20;  // smaller graphs stay in the caches
constexpr int kMaxLevels = 254;            // distances fit the key's low byte
constexpr int kListGrid = 2048;            // workgroups of a frontier-list level

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

// w[v] = distance << 24 | cluster; seeds (v % M == 0) are at distance 0 and
// form the first frontier, f0[c] = c * M; cnt[0] = their number.
__global__ void __launch_bounds__(kOrdTPB) k_cluster_seed(uint32_t n, uint32_t M, uint32_t clusters,
                                                          uint32_t* __restrict__ w, uint32_t* __restrict__ f0,
                                                          uint32_t* __restrict__ cnt) {
    for (uint32_t v = blockIdx.x * kOrdTPB + threadIdx.x; v < n; v += gridDim.x * kOrdTPB) {
        const bool seed = v % M == 0u;
        w[v] = seed ? v / M : kUnreached;
        if (seed) f0[v / M] = v;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        cnt[0] = clusters;
        cnt[1] = 0u;
        cnt[2] = 0u;
    }
}

// One BFS level L, pushed from its frontier list `fin` (the nodes at
// distance L, cnt[L % 3] of them): each offers distance L + 1 and its
// cluster to its neighbours with an atomic minimum on the packed word, so a
// node reached at this level ends with the smallest cluster among its
// frontier neighbours, whatever the timing.  A node already reached has a
// smaller word and keeps it.  The atomic that finds a node unreached is
// unique, and appends it to `fout` (count in cnt[(L + 1) % 3], zeroed by the
// level before; this level zeroes cnt[(L + 2) % 3] for the next).  A plain
// load first skips the atomic where it cannot lower the word.  The lists
// replace a sweep of all n words per level (28 levels on the 77.5M-node
// Delaunay graph).
__global__ void __launch_bounds__(kOrdTPB) k_cluster_level(uint32_t n, uint32_t L, const int4* __restrict__ nbr4,
                                                           uint32_t* __restrict__ w, const uint32_t* __restrict__ fin,
                                                           uint32_t* __restrict__ fout, uint32_t* cnt) {
    __shared__ WgQueue<kOrdTPB> q;
    q_init(q);
    const uint32_t c = cnt[L % 3u];
    uint32_t* out_cnt = &cnt[(L + 1u) % 3u];
    if (blockIdx.x == 0 && threadIdx.x == 0) cnt[(L + 2u) % 3u] = 0u;
    const Span sp = wg_span(c, kOrdTPB);
    for (uint64_t base = sp.b; base < sp.e; base += kOrdTPB) {
        const uint64_t i = base + threadIdx.x;
        bool got[4] = {false, false, false, false};
        int ids[4] = {-1, -1, -1, -1};
        if (i < sp.e) {
            const uint32_t u = fin[i];
            const uint32_t offer = ((L + 1u) << 24) | (w[u] & 0xFFFFFFu);
            const int4 nb = nbr4[u];
            ids[0] = nb.x;
            ids[1] = nb.y;
            ids[2] = nb.z;
            ids[3] = nb.w;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                if (!valid_id(ids[k], n)) continue;
                uint32_t* t = &w[ids[k]];
                if (__hip_atomic_load(t, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) <= offer) continue;
                got[k] = atomicMin(t, offer) == kUnreached;
            }
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) q_push(q, got[k], static_cast<uint32_t>(ids[k]));
        q_flush_if(q, fout, out_cnt, 4 * kOrdTPB);
    }
    q_flush(q, fout, out_cnt);
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

// The results back in the caller's order: node o takes row inv[o] of the
// copy.  Reads are random and writes coalesced, the cheaper way round: a
// scatter's random 16-, 4- and 1-byte writes each cost a read-modify-write
// of a whole line at the memory.
__global__ void __launch_bounds__(kOrdTPB) k_scatter_rows(uint32_t n, const uint32_t* __restrict__ inv,
                                                          const uint4* __restrict__ res_o,
                                                          const int32_t* __restrict__ term_o,
                                                          const uint8_t* __restrict__ part_o, uint4* __restrict__ res4,
                                                          int32_t* __restrict__ term, uint8_t* __restrict__ part) {
    for (uint32_t o = blockIdx.x * kOrdTPB + threadIdx.x; o < n; o += gridDim.x * kOrdTPB) {
        const uint32_t j = inv[o];
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
    auto* lcnt = reinterpret_cast<uint32_t*>(base + o_cnt);
    const int4* nbr4 = reinterpret_cast<const int4*>(d_nbr);
    const int g = grid_for(n);

    // 1. clusters: levels until the frontier empties (a check every 4 levels);
    // the frontiers ping-pong between the sort's key and permutation arrays,
    // free until the sort
    k_cluster_seed<<<g, kOrdTPB, 0, s>>>(n, M, clusters, w, keys, lcnt);
    cko(hipGetLastError(), "cluster seeds");
    uint32_t* fr[2] = {keys, perm};
    for (int L = 0; L < kMaxLevels;) {
        const int batch = std::min(4, kMaxLevels - L);
        for (int i = 0; i < batch; ++i, ++L)
            k_cluster_level<<<kListGrid, kOrdTPB, 0, s>>>(n, static_cast<uint32_t>(L), nbr4, w, fr[L & 1],
                                                          fr[(L + 1) & 1], lcnt);
        cko(hipGetLastError(), "cluster level");
        uint32_t h[3] = {0u, 0u, 0u};
        cko(hipMemcpyAsync(h, lcnt, sizeof(h), hipMemcpyDeviceToHost, s), "cluster count D2H");
        cko(hipStreamSynchronize(s), "cluster levels");
        if (h[L % 3] == 0u) break;  // level L's frontier is empty
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
    // the copy finished, so that t_order_ms holds the renumbering's whole time
    // (the solve waits for it anyway)
    cko(hipStreamSynchronize(s), "order copy");
    const double order_ms = std::chrono::duration<double, std::milli>(Clock::now() - t0).count();
    // 4. solve the copy and scatter back
    pr_solve_in_order(n, reinterpret_cast<int32_t*>(nbr_o), reinterpret_cast<uint32_t*>(res_o), term_o, part_o, true,
                      prm, st, stream);
    const auto t1 = Clock::now();
    uint8_t* part_dst = partition_on_device ? partition : reinterpret_cast<uint8_t*>(keys);
    k_scatter_rows<<<g, kOrdTPB, 0, s>>>(n, w, res_o, term_o, part_o, reinterpret_cast<uint4*>(d_res), d_term,
                                         part_dst);
    cko(hipGetLastError(), "order scatter");
    if (!partition_on_device) cko(hipMemcpyAsync(partition, part_dst, nz, hipMemcpyDeviceToHost, s), "partition D2H");
    cko(hipStreamSynchronize(s), "order scatter");
    st.t_order_ms = order_ms + std::chrono::duration<double, std::milli>(Clock::now() - t1).count();
    st.renumbered = 1u;
    st.solver_workspace_bytes += total;
}

}  // namespace wtv
