// wtv_graph_build.hip -- graph_cut_fast on the device: classify, reduced ids,
// exact capacities, terminal folding, solve, label scatter.
//
// Restates graph_cut_gpupr_fast_v630 (native/gpupr/wtivo_gpupr_bindings.cpp:100-397)
// with the topology never leaving HBM (the reference copies it to the host in
// 262,144-row chunks twice and uploads the reduced graph chunk by chunk,
// :139-148,227-232,305-313):
//   k_classify  label 0 -> fixed source (-2); label != 0 with a hull face ->
//               fixed sink (-3); otherwise free (:153-171); neighbour range
//               check; per-tile free counts and reference counts
//   k_tile_scan exclusive scan of the per-tile free counts
//   k_assign    reduced id = rank among free tets in increasing tet order
//               (:179-188), free_old[id] = tet
//   k_build     capacity_exact_formula (:65-98) for every slot of every free
//               tet: FP64 cross product, one rounding of 1e7*0.5*|c| to fp32,
//               pen = area or area*(1+fill_T), truncation to int64, range
//               [0, 2^30); free neighbour -> arc, fixed neighbour -> folded into
//               term = src - snk, base += min(src, snk) (:253-294)
//   k_scatter   out = labels; out[free_old[q]] = partition[q] (:357-364)
// Bit-exactness of the FP64 part needs no FMA contraction (pragma below) and a
// correctly rounded sqrt; both are checked against the host formula in the
// GPU tests.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <sstream>
#include <vector>

#include "wtv_internal.h"

namespace wtv {
namespace {

constexpr int kB = 256;           // threads per workgroup
constexpr uint32_t kTile = 4096;  // tets per scan tile (16 per thread)

inline void ckg(hipError_t e, const char* where) {
    if (e == hipSuccess) return;
    fail(WTV_EDEVICE, std::string(kPrefix) + "HIP at " + where + ": " + hipGetErrorString(e));
}

// Counters are spread over kSlots addresses (slot = workgroup index mod
// kSlots): device-scope atomics on one address serialise at ~12 ns each on
// MI355X, and a 78M-tet build has ~19k workgroups.  The host sums the slots.
constexpr int kSlots = 64;
struct BuildCtl {
    unsigned long long free_n[kSlots], fixed_source[kSlots], fixed_sink[kSlots], full_refs[kSlots],
        free_refs[kSlots], terminal_edges[kSlots];
    long long base_flow[kSlots];
    int32_t bad_range;  // neighbour out of range (classify)
    int32_t bad_build;  // 1 hull face on a free tet, 3 bad class, 4 term range, 5 bad vertex, 6 overflow
};
struct BuildSums {
    unsigned long long free_n = 0, fixed_source = 0, fixed_sink = 0, full_refs = 0, free_refs = 0,
                       terminal_edges = 0;
    long long base_flow = 0;
    int32_t bad_range = 0, bad_build = 0;
};
__device__ __forceinline__ uint32_t cslot() { return blockIdx.x % kSlots; }

__device__ __forceinline__ unsigned long long wsum(unsigned long long x) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
    return x;
}
__device__ __forceinline__ long long wsum(long long x) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
    return x;
}
__device__ __forceinline__ void add_u64(unsigned long long* p, unsigned long long v) {
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void add_i64(long long* p, long long v) {
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void flag(int32_t* p, int v) {
    __hip_atomic_exchange(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ int slot(const int4& v, int k) { return k == 0 ? v.x : k == 1 ? v.y : k == 2 ? v.z : v.w; }

// Workgroup sum, added once per workgroup.
template <class T>
__device__ __forceinline__ void wg_add(T* dst, T x, T* lds) {
    x = wsum(x);
    if (__lane_id() == 0) lds[threadIdx.x >> 6] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        T t = 0;
        for (int i = 0; i < kB / 64; ++i) t += lds[i];
        if (t) {
            if constexpr (sizeof(T) == 8 && static_cast<T>(-1) < 0) add_i64(reinterpret_cast<long long*>(dst), t);
            else add_u64(reinterpret_cast<unsigned long long*>(dst), static_cast<unsigned long long>(t));
        }
    }
    __syncthreads();
}

// One tile of kTile consecutive tets per workgroup.
__global__ void __launch_bounds__(kB) k_classify(uint32_t N, const int4* __restrict__ nbr4,
                                                 const uint8_t* __restrict__ labels, int32_t* __restrict__ cls,
                                                 uint32_t* __restrict__ tile_free, BuildCtl* ctl) {
    __shared__ unsigned long long lds[kB / 64];
    const uint32_t t0 = blockIdx.x * kTile;
    unsigned long long nf = 0, ns = 0, nk = 0, refs = 0;
    int bad = 0;
    for (uint32_t j = threadIdx.x; j < kTile; j += kB) {
        const uint32_t i = t0 + j;
        if (i >= N) break;
        const int4 nb = nbr4[i];
        bool exterior = false;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int x = slot(nb, k);
            if (x < 0) exterior = true;
            else {
                ++refs;
                if (static_cast<uint32_t>(x) >= N) bad = 1;
            }
        }
        int32_t c;
        if (labels[i] == 0u) { c = -2; ++ns; }
        else if (exterior) { c = -3; ++nk; }
        else { c = -1; ++nf; }
        cls[i] = c;
    }
    if (bad) flag(&ctl->bad_range, 1);
    wg_add(&ctl->fixed_source[cslot()], ns, lds);
    wg_add(&ctl->fixed_sink[cslot()], nk, lds);
    wg_add(&ctl->full_refs[cslot()], refs, lds);
    // per-tile free count (also summed into free_n)
    unsigned long long x = wsum(nf);
    if (__lane_id() == 0) lds[threadIdx.x >> 6] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t = 0;
        for (int i = 0; i < kB / 64; ++i) t += lds[i];
        tile_free[blockIdx.x] = static_cast<uint32_t>(t);
        if (t) add_u64(&ctl->free_n[cslot()], t);
    }
}

// Exclusive scan of the tile counts (one workgroup; tiles <= 2^20 per call).
__global__ void __launch_bounds__(kB) k_tile_scan(uint32_t tiles, uint32_t* __restrict__ tile_free) {
    __shared__ uint32_t wsums[kB / 64];
    __shared__ uint32_t carry;
    if (threadIdx.x == 0) carry = 0u;
    __syncthreads();
    for (uint32_t base = 0; base < tiles; base += kB) {
        const uint32_t i = base + threadIdx.x;
        const uint32_t v = i < tiles ? tile_free[i] : 0u;
        // inclusive scan within the wave
        uint32_t x = v;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t y = __shfl_up(x, o);
            if (static_cast<int>(__lane_id()) >= o) x += y;
        }
        if (__lane_id() == 63) wsums[threadIdx.x >> 6] = x;
        __syncthreads();
        uint32_t wpre = 0;
        for (int w = 0; w < static_cast<int>(threadIdx.x >> 6); ++w) wpre += wsums[w];
        const uint32_t c = carry;
        if (i < tiles) tile_free[i] = c + wpre + x - v;
        __syncthreads();
        if (threadIdx.x == kB - 1) carry = c + wpre + x;
        __syncthreads();
    }
}

// Reduced ids in increasing tet order: rank of each free tet inside its tile,
// processed as 16 rounds of 256 consecutive tets.
__global__ void __launch_bounds__(kB) k_assign(uint32_t N, int32_t* __restrict__ cls,
                                               const uint32_t* __restrict__ tile_base,
                                               uint32_t* __restrict__ free_old) {
    __shared__ uint32_t wcnt[kB / 64];
    __shared__ uint32_t carry;
    if (threadIdx.x == 0) carry = tile_base[blockIdx.x];
    __syncthreads();
    const uint32_t t0 = blockIdx.x * kTile;
    for (uint32_t j0 = 0; j0 < kTile; j0 += kB) {
        const uint32_t i = t0 + j0 + threadIdx.x;
        const bool fr = i < N && cls[i] == -1;
        const uint64_t m = __ballot(fr);
        const uint32_t rank = __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(m >> 32),
                                                        __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(m), 0u));
        if (__lane_id() == 0) wcnt[threadIdx.x >> 6] = static_cast<uint32_t>(__popcll(m));
        __syncthreads();
        uint32_t wpre = 0;
        for (int w = 0; w < static_cast<int>(threadIdx.x >> 6); ++w) wpre += wcnt[w];
        const uint32_t c = carry;
        if (fr) {
            const uint32_t id = c + wpre + rank;
            cls[i] = static_cast<int32_t>(id);
            free_old[id] = i;
        }
        __syncthreads();
        if (threadIdx.x == 0) carry = c + wcnt[0] + wcnt[1] + wcnt[2] + wcnt[3];
        __syncthreads();
    }
}

// capacity_exact_formula (wtivo_gpupr_bindings.cpp:65-98); 0 ok, 1 bad vertex, 2 overflow.
__device__ __forceinline__ int cap_exact(const double* __restrict__ V, size_t nv, int a, int b, int c, bool differ,
                                         double fill_T, int32_t* out) {
#pragma clang fp contract(off)
    if (a < 0 || b < 0 || c < 0 || static_cast<size_t>(a) >= nv || static_cast<size_t>(b) >= nv ||
        static_cast<size_t>(c) >= nv)
        return 1;
    const double* p0 = V + static_cast<size_t>(a) * 3u;
    const double* p1 = V + static_cast<size_t>(b) * 3u;
    const double* p2 = V + static_cast<size_t>(c) * 3u;
    const double x1 = p1[0] - p0[0], y1 = p1[1] - p0[1], z1 = p1[2] - p0[2];
    const double x2 = p2[0] - p0[0], y2 = p2[1] - p0[1], z2 = p2[2] - p0[2];
    const double cx = y1 * z2 - z1 * y2;
    const double cy = z1 * x2 - x1 * z2;
    const double cz = x1 * y2 - y1 * x2;
    const float area_f = static_cast<float>(1.0E7 * 0.5 * __builtin_sqrt(cx * cx + cy * cy + cz * cz));
    const double area = static_cast<double>(area_f);
    const double pen = differ ? area : area * (1.0 + fill_T);
    if (!(pen > -9.2e18 && pen < 9.2e18)) return 2;
    const long long cap64 = static_cast<long long>(pen);
    if (cap64 < 0 || cap64 >= 1073741824LL) return 2;
    *out = static_cast<int32_t>(cap64);
    return 0;
}

__global__ void __launch_bounds__(kB) k_build(uint32_t N, const double* __restrict__ V, size_t nv,
                                              const int4* __restrict__ tets4, const int4* __restrict__ nbr4,
                                              const uint8_t* __restrict__ labels, const int32_t* __restrict__ cls,
                                              double fill_T, int4* __restrict__ nbr_r, uint4* __restrict__ res_r,
                                              int32_t* __restrict__ term_r, BuildCtl* ctl) {
    __shared__ unsigned long long lds_u[kB / 64];
    __shared__ long long lds_i[kB / 64];
    constexpr int lut[4][3] = {{0, 1, 2}, {0, 1, 3}, {0, 2, 3}, {1, 2, 3}};
    unsigned long long refs = 0, terms = 0;
    long long base = 0;
    int bad = 0;
    const uint32_t stride = gridDim.x * kB;
    for (uint32_t i = blockIdx.x * kB + threadIdx.x; i < N; i += stride) {
        const int32_t rr = cls[i];
        if (rr < 0) continue;
        const int4 tr = tets4[i];
        const int4 nb = nbr4[i];
        const int tv[4] = {tr.x, tr.y, tr.z, tr.w};
        const uint8_t li = labels[i];
        int nr[4] = {-1, -1, -1, -1};
        uint32_t rs[4] = {0u, 0u, 0u, 0u};
        long long src = 0, snk = 0;
#pragma unroll
        for (int fi = 0; fi < 4; ++fi) {
            const int j = slot(nb, fi);
            if (j < 0) { bad = 1; continue; }
            if (static_cast<uint32_t>(j) >= N) { bad = 3; continue; }
            const int32_t m = cls[j];
            int32_t cap = 0;
            const int rc = cap_exact(V, nv, tv[lut[fi][0]], tv[lut[fi][1]], tv[lut[fi][2]], li != labels[j], fill_T,
                                     &cap);
            if (rc == 1) { bad = 5; continue; }
            if (rc == 2) { bad = 6; continue; }
            if (m >= 0) {
                nr[fi] = m;
                rs[fi] = static_cast<uint32_t>(cap);
                ++refs;
            } else if (m == -2) { src += cap; ++terms; }
            else if (m == -3) { snk += cap; ++terms; }
            else bad = 3;
        }
        base += src < snk ? src : snk;
        const long long diff = src - snk;
        if (diff < -2147483648LL || diff > 2147483647LL) bad = 4;
        nbr_r[rr] = make_int4(nr[0], nr[1], nr[2], nr[3]);
        res_r[rr] = make_uint4(rs[0], rs[1], rs[2], rs[3]);
        term_r[rr] = static_cast<int32_t>(diff);
    }
    if (bad) flag(&ctl->bad_build, bad);
    wg_add(&ctl->free_refs[cslot()], refs, lds_u);
    wg_add(&ctl->terminal_edges[cslot()], terms, lds_u);
    wg_add(&ctl->base_flow[cslot()], base, lds_i);
}

__global__ void __launch_bounds__(kB) k_scatter(uint32_t n, const uint32_t* __restrict__ free_old,
                                                const uint8_t* __restrict__ part, uint8_t* __restrict__ out) {
    const uint32_t stride = gridDim.x * kB;
    for (uint32_t q = blockIdx.x * kB + threadIdx.x; q < n; q += stride) out[free_old[q]] = part[q] ? 1u : 0u;
}

// ------------------------------------------------------------------------
// Surface extraction (wtivo_gpupr_bindings.cpp:400-577): oriented boundary
// triangles of label-0 tets, tet-major / slot-minor, then the used vertices
// compacted in ascending original id.
// ------------------------------------------------------------------------
__device__ __forceinline__ int surf_faces(uint8_t li, const int4& nb, uint32_t N, const uint8_t* __restrict__ labels,
                                          int* take4, int* bad) {
    int c = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int j = slot(nb, k);
        int t = 0;
        if (li == 0u) {
            if (j < 0) t = 1;
            else if (static_cast<uint32_t>(j) >= N) *bad = 1;
            else t = labels[j] != 0u;
        }
        take4[k] = t;
        c += t;
    }
    return c;
}

__global__ void __launch_bounds__(kB) k_surf_count(uint32_t N, const int4* __restrict__ nbr4,
                                                   const uint8_t* __restrict__ labels, uint32_t* __restrict__ tile_cnt,
                                                   BuildCtl* ctl) {
    __shared__ unsigned long long lds[kB / 64];
    const uint32_t t0 = blockIdx.x * kTile;
    unsigned long long c = 0;
    int bad = 0;
    for (uint32_t j = threadIdx.x; j < kTile; j += kB) {
        const uint32_t i = t0 + j;
        if (i >= N) break;
        const uint8_t li = labels[i];
        if (li != 0u) continue;
        int take4[4];
        c += surf_faces(li, nbr4[i], N, labels, take4, &bad);
    }
    if (bad) flag(&ctl->bad_range, 1);
    unsigned long long x = wsum(c);
    if (__lane_id() == 0) lds[threadIdx.x >> 6] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t = 0;
        for (int i = 0; i < kB / 64; ++i) t += lds[i];
        tile_cnt[blockIdx.x] = static_cast<uint32_t>(t);
        if (t) add_u64(&ctl->free_n[cslot()], t);  // total faces
    }
}

__global__ void __launch_bounds__(kB) k_surf_emit(uint32_t N, const int4* __restrict__ tets4,
                                                  const int4* __restrict__ nbr4, const uint8_t* __restrict__ labels,
                                                  const double* __restrict__ V, size_t nv,
                                                  const uint32_t* __restrict__ tile_off, int32_t* __restrict__ raw,
                                                  uint8_t* __restrict__ used, BuildCtl* ctl) {
#pragma clang fp contract(off)
    constexpr int lut[4][3] = {{0, 1, 2}, {0, 1, 3}, {0, 2, 3}, {1, 2, 3}};
    constexpr int opp[4] = {3, 2, 1, 0};
    __shared__ uint32_t wsums[kB / 64];
    __shared__ uint32_t carry;
    if (threadIdx.x == 0) carry = tile_off[blockIdx.x];
    __syncthreads();
    const uint32_t t0 = blockIdx.x * kTile;
    int bad = 0;
    for (uint32_t j0 = 0; j0 < kTile; j0 += kB) {
        const uint32_t i = t0 + j0 + threadIdx.x;
        int take4[4] = {0, 0, 0, 0};
        uint32_t c = 0;
        int4 nb = make_int4(-1, -1, -1, -1);
        uint8_t li = 1u;
        if (i < N) {
            li = labels[i];
            if (li == 0u) {
                nb = nbr4[i];
                int dummy = 0;
                c = static_cast<uint32_t>(surf_faces(li, nb, N, labels, take4, &dummy));
            }
        }
        // exclusive scan of c over the workgroup round
        uint32_t x = c;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t y = __shfl_up(x, o);
            if (static_cast<int>(__lane_id()) >= o) x += y;
        }
        if (__lane_id() == 63) wsums[threadIdx.x >> 6] = x;
        __syncthreads();
        uint32_t wpre = 0;
        for (int w = 0; w < static_cast<int>(threadIdx.x >> 6); ++w) wpre += wsums[w];
        const uint32_t base = carry;
        uint32_t pos = base + wpre + x - c;
        if (c) {
            const int4 tr = tets4[i];
            const int tv[4] = {tr.x, tr.y, tr.z, tr.w};
            for (int k = 0; k < 4; ++k) {
                if (!take4[k]) continue;
                int a0 = tv[lut[k][0]], b0 = tv[lut[k][1]], c0 = tv[lut[k][2]];
                const int fourth = tv[opp[k]];
                if (a0 < 0 || b0 < 0 || c0 < 0 || fourth < 0 || static_cast<size_t>(a0) >= nv ||
                    static_cast<size_t>(b0) >= nv || static_cast<size_t>(c0) >= nv ||
                    static_cast<size_t>(fourth) >= nv) {
                    bad = 2;
                    raw[3u * pos] = raw[3u * pos + 1u] = raw[3u * pos + 2u] = 0;
                    ++pos;
                    continue;
                }
                int t;
                if (a0 > b0) { t = a0; a0 = b0; b0 = t; }
                if (b0 > c0) { t = b0; b0 = c0; c0 = t; }
                if (a0 > b0) { t = a0; a0 = b0; b0 = t; }
                const double* v0 = V + static_cast<size_t>(a0) * 3u;
                const double* v1 = V + static_cast<size_t>(b0) * 3u;
                const double* v2 = V + static_cast<size_t>(c0) * 3u;
                const double* vp = V + static_cast<size_t>(fourth) * 3u;
                const double x1 = v1[0] - v0[0], y1 = v1[1] - v0[1], z1 = v1[2] - v0[2];
                const double x2 = v2[0] - v0[0], y2 = v2[1] - v0[1], z2 = v2[2] - v0[2];
                const double cx = y1 * z2 - z1 * y2;
                const double cy = z1 * x2 - x1 * z2;
                const double cz = x1 * y2 - y1 * x2;
                const double px = vp[0] - v0[0], py = vp[1] - v0[1], pz = vp[2] - v0[2];
                const double vol = cx * px + cy * py + cz * pz;
                raw[3u * pos] = a0;
                raw[3u * pos + 1u] = vol < 0.0 ? b0 : c0;
                raw[3u * pos + 2u] = vol < 0.0 ? c0 : b0;
                used[a0] = 1u;
                used[b0] = 1u;
                used[c0] = 1u;
                ++pos;
            }
        }
        __syncthreads();
        if (threadIdx.x == kB - 1) carry = base + wpre + x;
        __syncthreads();
    }
    if (bad) flag(&ctl->bad_build, bad);
}

// Compaction of a 0/1 flag array in index order: per-tile counts, then ranks.
__global__ void __launch_bounds__(kB) k_flag_count(uint32_t n, const uint8_t* __restrict__ flags,
                                                   uint32_t* __restrict__ tile_cnt) {
    __shared__ uint32_t lds[kB / 64];
    const uint32_t t0 = blockIdx.x * kTile;
    uint32_t c = 0;
    for (uint32_t j = threadIdx.x; j < kTile; j += kB) {
        const uint32_t i = t0 + j;
        if (i < n) c += flags[i] != 0u;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o);
    if (__lane_id() == 0) lds[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) tile_cnt[blockIdx.x] = lds[0] + lds[1] + lds[2] + lds[3];
}

__global__ void __launch_bounds__(kB) k_flag_rank(uint32_t n, const uint8_t* __restrict__ flags,
                                                  const uint32_t* __restrict__ tile_base, int32_t* __restrict__ rank_out) {
    __shared__ uint32_t wcnt[kB / 64];
    __shared__ uint32_t carry;
    if (threadIdx.x == 0) carry = tile_base[blockIdx.x];
    __syncthreads();
    const uint32_t t0 = blockIdx.x * kTile;
    for (uint32_t j0 = 0; j0 < kTile; j0 += kB) {
        const uint32_t i = t0 + j0 + threadIdx.x;
        const bool f = i < n && flags[i] != 0u;
        const uint64_t m = __ballot(f);
        const uint32_t r = __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(m >> 32),
                                                     __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(m), 0u));
        if (__lane_id() == 0) wcnt[threadIdx.x >> 6] = static_cast<uint32_t>(__popcll(m));
        __syncthreads();
        uint32_t wpre = 0;
        for (int w = 0; w < static_cast<int>(threadIdx.x >> 6); ++w) wpre += wcnt[w];
        const uint32_t c = carry;
        if (i < n) rank_out[i] = f ? static_cast<int32_t>(c + wpre + r) : -1;
        __syncthreads();
        if (threadIdx.x == 0) carry = c + wcnt[0] + wcnt[1] + wcnt[2] + wcnt[3];
        __syncthreads();
    }
}

__global__ void __launch_bounds__(kB) k_surf_finish(uint32_t nv, uint64_t nidx, const double* __restrict__ V,
                                                    const int32_t* __restrict__ remap, double* __restrict__ ov,
                                                    int32_t* __restrict__ raw) {
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kB;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * kB + threadIdx.x; i < nv; i += stride) {
        const int32_t r = remap[i];
        if (r >= 0) {
            ov[3u * static_cast<size_t>(r)] = V[3u * i];
            ov[3u * static_cast<size_t>(r) + 1u] = V[3u * i + 1u];
            ov[3u * static_cast<size_t>(r) + 2u] = V[3u * i + 2u];
        }
    }
    for (uint64_t q = static_cast<uint64_t>(blockIdx.x) * kB + threadIdx.x; q < nidx; q += stride)
        raw[q] = remap[raw[q]];
}

struct Dev {
    void* p = nullptr;
    ~Dev() { if (p) (void)hipFree(p); }
    template <class T> T* as() const { return static_cast<T*>(p); }
};

// Copies the control block to the host and sums the counter slots.
BuildSums read_build_ctl(const BuildCtl* d, hipStream_t s, const char* where) {
    std::vector<BuildCtl> raw(1);
    ckg(hipMemcpyAsync(raw.data(), d, sizeof(BuildCtl), hipMemcpyDeviceToHost, s), where);
    ckg(hipStreamSynchronize(s), where);
    const BuildCtl& r = raw[0];
    BuildSums h;
    for (int i = 0; i < kSlots; ++i) {
        h.free_n += r.free_n[i];
        h.fixed_source += r.fixed_source[i];
        h.fixed_sink += r.fixed_sink[i];
        h.full_refs += r.full_refs[i];
        h.free_refs += r.free_refs[i];
        h.terminal_edges += r.terminal_edges[i];
        h.base_flow += r.base_flow[i];
    }
    h.bad_range = r.bad_range;
    h.bad_build = r.bad_build;
    return h;
}

}  // namespace

// Scratch of one reduced-graph build: the class of every tet, free-tet counts
// per tile and the control block.
struct BuildScratch {
    Dev dCls, dTiles, dCtl;
    uint32_t tiles = 0;
};

// Classification and its counters (wtivo_gpupr_bindings.cpp:139-188): label 0
// -> fixed source, label != 0 with a hull face -> fixed sink, else free.
// Fills the classify counters of info and returns the free tet count.
uint64_t classify_tets(const int32_t* dN, size_t N, const uint8_t* d_lab, BuildScratch& bs, BuildInfo& info,
                       hipStream_t s) {
    std::memset(&info, 0, sizeof(info));
    const uint32_t n32 = static_cast<uint32_t>(N);
    bs.tiles = static_cast<uint32_t>((N + kTile - 1) / kTile);
    if (bs.tiles > (1u << 20)) fail(WTV_EINVAL, ">2^32 full tets unsupported");
    ckg(hipMalloc(&bs.dCls.p, N * 4u), "malloc classes");
    ckg(hipMalloc(&bs.dTiles.p, static_cast<size_t>(bs.tiles) * 4u), "malloc tiles");
    ckg(hipMalloc(&bs.dCtl.p, sizeof(BuildCtl)), "malloc build ctl");
    ckg(hipMemsetAsync(bs.dCtl.p, 0, sizeof(BuildCtl), s), "build ctl clear");
    const int4* nbr4 = reinterpret_cast<const int4*>(dN);
    BuildCtl* ctl = bs.dCtl.as<BuildCtl>();
    k_classify<<<bs.tiles, kB, 0, s>>>(n32, nbr4, d_lab, bs.dCls.as<int32_t>(), bs.dTiles.as<uint32_t>(), ctl);
    ckg(hipGetLastError(), "classify");
    const BuildSums h = read_build_ctl(ctl, s, "classify D2H");
    if (h.bad_range) fail(WTV_EINVAL, "neighbor out of range");
    if (h.full_refs & 1ull) fail(WTV_EINVAL, "neighbor reference count not even");
    info.free_n = h.free_n;
    info.fixed_source = h.fixed_source;
    info.fixed_sink = h.fixed_sink;
    info.full_refs = h.full_refs;
    if (h.free_n > static_cast<uint64_t>(std::numeric_limits<int32_t>::max()))
        fail(WTV_EINVAL, ">INT32_MAX reduced nodes unsupported");
    return h.free_n;
}

// The rest of the build after classify_tets found free tets: reduced ids in
// increasing tet order, capacity_exact_formula and terminal folding
// (wtivo_gpupr_bindings.cpp:179-318).  Outputs hold at least free_n entries;
// their contents match the reference's host build exactly.
void build_graph(const double* V, bool v_on_device, size_t nv, const int32_t* dT, const int32_t* dN, size_t N,
                 const uint8_t* d_lab, double fill_T, int32_t* gN, uint32_t* gR, int32_t* gT, uint32_t* free_old,
                 BuildScratch& bs, BuildInfo& info, hipStream_t s) {
    const uint32_t n32 = static_cast<uint32_t>(N);
    const int4* nbr4 = reinterpret_cast<const int4*>(dN);
    const int4* tets4 = reinterpret_cast<const int4*>(dT);
    BuildCtl* ctl = bs.dCtl.as<BuildCtl>();
    Dev dV;
    const double* vdev = V;
    if (!v_on_device) {
        ckg(hipMalloc(&dV.p, std::max<size_t>(nv, 1) * 24u), "malloc vertices");
        pinned_to_device(V, nv * 24u, dV.p, 0, s);
        vdev = dV.as<double>();
    }
    k_tile_scan<<<1, kB, 0, s>>>(bs.tiles, bs.dTiles.as<uint32_t>());
    k_assign<<<bs.tiles, kB, 0, s>>>(n32, bs.dCls.as<int32_t>(), bs.dTiles.as<uint32_t>(), free_old);
    const int grid = static_cast<int>(std::min<uint64_t>((N + kB - 1) / kB, 8192));
    k_build<<<grid, kB, 0, s>>>(n32, vdev, nv, tets4, nbr4, d_lab, bs.dCls.as<int32_t>(), fill_T,
                                reinterpret_cast<int4*>(gN), reinterpret_cast<uint4*>(gR), gT, ctl);
    ckg(hipGetLastError(), "build");
    const BuildSums h = read_build_ctl(ctl, s, "build D2H");
    // the reference's build lambda turns every failure, the capacity
    // overflow_error included, into one runtime_error (:292,296-297)
    if (h.bad_build == 6) fail(WTV_EINVAL, "invalid graph reduction/capacity state (capacity outside safe 30-bit range)");
    if (h.bad_build == 5) fail(WTV_EINVAL, "invalid graph reduction/capacity state (invalid tet vertex)");
    if (h.bad_build) fail(WTV_EINVAL, "invalid graph reduction/capacity state");
    if (h.free_refs & 1ull) fail(WTV_EINVAL, "free reference count not even");
    info.free_refs = h.free_refs;
    info.terminal_edges = h.terminal_edges;
    info.base_flow = h.base_flow;
}

// Builds the reduced graph on the device into caller buffers holding at least
// free_n entries.  Returns false when there are no free nodes (the caller then
// returns the input labels).
bool build_reduced(const double* V, bool v_on_device, size_t nv, const int32_t* dT, const int32_t* dN, size_t N,
                   const uint8_t* d_lab, double fill_T, int32_t* gN, uint32_t* gR, int32_t* gT, uint32_t* free_old,
                   BuildInfo& info, hipStream_t s) {
    BuildScratch bs;
    if (classify_tets(dN, N, d_lab, bs, info, s) == 0) return false;
    build_graph(V, v_on_device, nv, dT, dN, N, d_lab, fill_T, gN, gR, gT, free_old, bs, info, s);
    return true;
}

void graph_cut_fast(const double* V, size_t nv, const int32_t* dT, const int32_t* dN, size_t N, const uint8_t* lp,
                    double fill_T, int requested_threads, int period, uint64_t max_rounds, int local_steps,
                    uint8_t* out_labels, wtv_cut_result& R, wtv_stats& st, void* stream_v, bool resident) {
    using Clock = std::chrono::steady_clock;
    const auto all0 = Clock::now();
    std::memset(&R, 0, sizeof(R));
    std::memset(&st, 0, sizeof(st));
    if ((!V && nv) || (!dT && N) || (!dN && N) || (!lp && N) || (!out_labels && N))
        fail(WTV_EINVAL, "null pointer argument");
    if (N > static_cast<size_t>(std::numeric_limits<uint32_t>::max()))
        fail(WTV_EINVAL, ">2^32 full tets unsupported");
    R.threads = std::max(1, requested_threads);
    R.local_steps = local_steps;
    if (N == 0) {
        R.native_total_seconds = std::chrono::duration<double>(Clock::now() - all0).count();
        return;
    }
    hipStream_t s = static_cast<hipStream_t>(stream_v);
    // WTV_TRACE=1: one stderr line per phase with its wall time (diagnostics)
    const char* trace_env = std::getenv("WTV_TRACE");
    const bool trace = trace_env && *trace_env && *trace_env != '0';
    auto mark = all0;
    auto phase = [&](const char* what) {
        if (!trace) return;
        (void)hipStreamSynchronize(s);
        const auto now = Clock::now();
        std::fprintf(stderr, "[wtv] graph_cut_fast %-18s %8.3f ms\n", what,
                     std::chrono::duration<double, std::milli>(now - mark).count());
        mark = now;
    };
    // resident: vertices, labels and output labels are device buffers
    Dev dLab;
    const uint8_t* d_lab = lp;
    if (!resident) {
        ckg(hipMalloc(&dLab.p, N), "malloc labels");
        pinned_to_device(lp, N, dLab.p, 0, s);
        d_lab = dLab.as<uint8_t>();
    }
    phase("labels upload");

    const auto build0 = Clock::now();
    BuildInfo bi{};
    BuildScratch bs;
    const uint64_t free_N = classify_tets(dN, N, d_lab, bs, bi, s);
    phase("classify");
    R.full_edges = bi.full_refs / 2u;
    R.fixed_source = bi.fixed_source;
    R.fixed_sink = bi.fixed_sink;
    if (free_N == 0) {
        if (resident) {
            ckg(hipMemcpyAsync(out_labels, lp, N, hipMemcpyDeviceToDevice, s), "labels copy");
            ckg(hipStreamSynchronize(s), "labels copy");
        } else {
            std::memcpy(out_labels, lp, N);
        }
        R.build_upload_seconds = std::chrono::duration<double>(Clock::now() - build0).count();
        R.native_total_seconds = std::chrono::duration<double>(Clock::now() - all0).count();
        return;
    }

    // VRAM precheck after classification, as the reference (:200-212): the
    // reduced graph (36 B) plus this solver's workspace (26 B; its default
    // knobs run no pull-mode sub-rounds), the reduced id
    // map (4 B) and the partition (1 B) per free node, the vertices and output
    // labels still to be placed, and 256 MiB of headroom
    size_t gpu_free = 0, gpu_total = 0;
    ckg(hipMemGetInfo(&gpu_free, &gpu_total), "hipMemGetInfo");
    const uint64_t need = free_N * (36ull + 26ull + 4ull + 1ull) +
                          (resident ? 0ull : static_cast<uint64_t>(nv) * 24u + static_cast<uint64_t>(N)) +
                          (256ull << 20);
    if (need > static_cast<uint64_t>(gpu_free)) fail(WTV_ENOMEM, "insufficient VRAM for reduced graph + solver");
    phase("vram precheck");
    Dev gN, gR, gT, dFreeOld;
    ckg(hipMalloc(&gN.p, free_N * 16u), "malloc graph nbr");
    ckg(hipMalloc(&gR.p, free_N * 16u), "malloc graph res");
    ckg(hipMalloc(&gT.p, free_N * 4u), "malloc graph term");
    ckg(hipMalloc(&dFreeOld.p, free_N * 4u), "malloc free_old");
    phase("graph malloc");
    build_graph(V, resident, nv, dT, dN, N, d_lab, fill_T, gN.as<int32_t>(), gR.as<uint32_t>(), gT.as<int32_t>(),
                dFreeOld.as<uint32_t>(), bs, bi, s);
    const auto build1 = Clock::now();
    phase("build");

    wtv_params prm;
    wtv_params_default(&prm);
    prm.global_relabel_period = period;
    prm.max_rounds = max_rounds;
    prm.local_steps = local_steps;
    Dev dPart, dOut;
    ckg(hipMalloc(&dPart.p, free_N), "malloc partition");
    pr_solve(static_cast<uint32_t>(free_N), gN.as<int32_t>(), gR.as<uint32_t>(), gT.as<int32_t>(), dPart.as<uint8_t>(),
             true, prm, st, stream_v);
    phase("solve");

    // labels (:357-364): out = labels; out[free_old[q]] = partition[q]
    uint8_t* d_out = out_labels;
    if (!resident) {
        ckg(hipMalloc(&dOut.p, N), "malloc out labels");
        d_out = dOut.as<uint8_t>();
    }
    if (d_out != d_lab) ckg(hipMemcpyAsync(d_out, d_lab, N, hipMemcpyDeviceToDevice, s), "labels copy");
    k_scatter<<<static_cast<int>(std::min<uint64_t>((free_N + kB - 1) / kB, 8192)), kB, 0, s>>>(
        static_cast<uint32_t>(free_N), dFreeOld.as<uint32_t>(), dPart.as<uint8_t>(), d_out);
    ckg(hipGetLastError(), "scatter");
    phase("scatter");
    if (!resident) pinned_to_host(d_out, N, out_labels, 0, s);
    ckg(hipStreamSynchronize(s), "labels sync");
    phase("labels download");

    R.adjusted_flow = st.raw_flow + bi.base_flow;
    R.raw_flow = st.raw_flow;
    R.graph_bytes = free_N * 36ull;
    R.workspace_bytes = st.solver_workspace_bytes;
    R.free_edges = bi.free_refs / 2u;
    R.terminal_edges = bi.terminal_edges;
    R.macro_rounds = st.rounds;
    R.global_relabels = st.global_relabels;
    R.map_bytes = static_cast<uint64_t>(N) * sizeof(int32_t) + free_N * sizeof(uint32_t);
    R.build_upload_seconds = std::chrono::duration<double>(build1 - build0).count();
    R.solve_seconds = st.solve_seconds;
    R.native_total_seconds = std::chrono::duration<double>(Clock::now() - all0).count();
}

void build_reduced_graph(const double* V, size_t nv, const int32_t* dT, const int32_t* dN, size_t N,
                         const uint8_t* labels, double fill_T, int32_t* gN, uint32_t* gR, int32_t* gT,
                         uint32_t* free_old, BuildInfo& info, void* stream_v) {
    std::memset(&info, 0, sizeof(info));
    if (!N) return;
    if (!V || !dT || !dN || !labels || !gN || !gR || !gT || !free_old) fail(WTV_EINVAL, "null pointer argument");
    if (N > static_cast<size_t>(std::numeric_limits<uint32_t>::max()))
        fail(WTV_EINVAL, ">2^32 full tets unsupported");
    hipStream_t s = static_cast<hipStream_t>(stream_v);
    Dev dLab;
    ckg(hipMalloc(&dLab.p, N), "malloc labels");
    pinned_to_device(labels, N, dLab.p, 0, s);
    build_reduced(V, false, nv, dT, dN, N, dLab.as<uint8_t>(), fill_T, gN, gR, gT, free_old, info, s);
    ckg(hipStreamSynchronize(s), "build");
}

void surface_extract(const uint8_t* lp, const double* V, size_t nv, const int32_t* dT, const int32_t* dN, size_t N,
                     std::vector<double>& OV, std::vector<int32_t>& OF, void* stream_v) {
    OV.clear();
    OF.clear();
    if ((!lp && N) || (!V && nv) || (!dT && N) || (!dN && N)) fail(WTV_EINVAL, "null pointer argument");
    if (N > static_cast<size_t>(std::numeric_limits<uint32_t>::max()) ||
        nv > static_cast<size_t>(std::numeric_limits<int32_t>::max()))
        fail(WTV_EINVAL, ">2^32 full tets unsupported");
    if (N == 0) return;
    hipStream_t s = static_cast<hipStream_t>(stream_v);
    const uint32_t n32 = static_cast<uint32_t>(N);
    const uint32_t tiles = static_cast<uint32_t>((N + kTile - 1) / kTile);
    const uint32_t vtiles = static_cast<uint32_t>((std::max<size_t>(nv, 1) + kTile - 1) / kTile);
    Dev dLab, dTiles, dCtl, dV, dUsed, dRemap, dVt;
    ckg(hipMalloc(&dLab.p, N), "malloc labels");
    ckg(hipMalloc(&dTiles.p, static_cast<size_t>(tiles) * 4u), "malloc tiles");
    ckg(hipMalloc(&dCtl.p, sizeof(BuildCtl)), "malloc ctl");
    pinned_to_device(lp, N, dLab.p, 0, s);
    ckg(hipMemsetAsync(dCtl.p, 0, sizeof(BuildCtl), s), "ctl clear");
    BuildCtl* ctl = dCtl.as<BuildCtl>();
    BuildSums h;
    const int4* nbr4 = reinterpret_cast<const int4*>(dN);
    const int4* tets4 = reinterpret_cast<const int4*>(dT);
    k_surf_count<<<tiles, kB, 0, s>>>(n32, nbr4, dLab.as<uint8_t>(), dTiles.as<uint32_t>(), ctl);
    ckg(hipGetLastError(), "surface count");
    h = read_build_ctl(ctl, s, "surface D2H");
    if (h.bad_range) fail(WTV_EINVAL, "surface neighbor out of range");
    const uint64_t F = h.free_n;
    if (F * 3u > 0x7FFFFFFFull) fail(WTV_EINVAL, "surface too large");
    Dev dRaw;
    ckg(hipMalloc(&dRaw.p, std::max<uint64_t>(F, 1) * 12u), "malloc faces");
    ckg(hipMalloc(&dV.p, std::max<size_t>(nv, 1) * 24u), "malloc vertices");
    ckg(hipMalloc(&dUsed.p, std::max<size_t>(nv, 1)), "malloc used");
    ckg(hipMalloc(&dRemap.p, std::max<size_t>(nv, 1) * 4u), "malloc remap");
    ckg(hipMalloc(&dVt.p, static_cast<size_t>(vtiles) * 4u), "malloc vertex tiles");
    pinned_to_device(V, nv * 24u, dV.p, 0, s);
    ckg(hipMemsetAsync(dUsed.p, 0, std::max<size_t>(nv, 1), s), "used clear");
    k_tile_scan<<<1, kB, 0, s>>>(tiles, dTiles.as<uint32_t>());
    k_surf_emit<<<tiles, kB, 0, s>>>(n32, tets4, nbr4, dLab.as<uint8_t>(), dV.as<double>(), nv, dTiles.as<uint32_t>(),
                                     dRaw.as<int32_t>(), dUsed.as<uint8_t>(), ctl);
    ckg(hipGetLastError(), "surface emit");
    h = read_build_ctl(ctl, s, "surface D2H");
    if (h.bad_build) fail(WTV_EINVAL, "surface invalid tet vertex index");
    const uint32_t nv32 = static_cast<uint32_t>(nv);
    k_flag_count<<<vtiles, kB, 0, s>>>(nv32, dUsed.as<uint8_t>(), dVt.as<uint32_t>());
    k_tile_scan<<<1, kB, 0, s>>>(vtiles, dVt.as<uint32_t>());
    k_flag_rank<<<vtiles, kB, 0, s>>>(nv32, dUsed.as<uint8_t>(), dVt.as<uint32_t>(), dRemap.as<int32_t>());
    uint32_t last_base = 0, last_cnt = 0;
    // used count = last tile base + flags in the last tile
    ckg(hipMemcpyAsync(&last_base, dVt.as<uint32_t>() + (vtiles - 1), 4, hipMemcpyDeviceToHost, s), "used D2H");
    ckg(hipStreamSynchronize(s), "used sync");
    std::vector<uint8_t> tail(std::min<size_t>(nv, kTile));
    const size_t tail0 = static_cast<size_t>(vtiles - 1) * kTile;
    if (nv > tail0) {
        ckg(hipMemcpyAsync(tail.data(), dUsed.as<uint8_t>() + tail0, nv - tail0, hipMemcpyDeviceToHost, s), "tail");
        ckg(hipStreamSynchronize(s), "tail sync");
        for (size_t i = 0; i < nv - tail0; ++i) last_cnt += tail[i] != 0u;
    }
    const size_t used = static_cast<size_t>(last_base) + last_cnt;
    Dev dOV;
    ckg(hipMalloc(&dOV.p, std::max<size_t>(used, 1) * 24u), "malloc out vertices");
    const int grid = static_cast<int>(std::min<uint64_t>((std::max<uint64_t>(nv, F * 3u) + kB - 1) / kB, 8192));
    k_surf_finish<<<std::max(grid, 1), kB, 0, s>>>(nv32, F * 3u, dV.as<double>(), dRemap.as<int32_t>(),
                                                   dOV.as<double>(), dRaw.as<int32_t>());
    ckg(hipGetLastError(), "surface finish");
    OV.resize(used * 3u);
    OF.resize(F * 3u);
    if (used) ckg(hipMemcpyAsync(OV.data(), dOV.p, used * 24u, hipMemcpyDeviceToHost, s), "OV D2H");
    if (F) ckg(hipMemcpyAsync(OF.data(), dRaw.p, F * 12u, hipMemcpyDeviceToHost, s), "OF D2H");
    ckg(hipStreamSynchronize(s), "surface sync");
}

}  // namespace wtv
