// wtv_topology.hip -- 4-slot tet neighbours built on the device.
//
// Restates the neighbour export of tetrahedralize_neighbors
// (native/core/wtivo_core.cpp:136-154) for meshes that arrive with tets only
// (generated lattices, meshes from other tools): slot fi of tet i holds the
// tet across face {012,013,023,123}[fi], i.e. the face opposite local vertex
// 3-fi (CGAL's neighbor(3-fi), the slot map {3,2,1,0} of :136), and -1 on the
// hull.  In a conforming mesh that tet is the unique other tet holding the
// same three vertices, so matching faces by their vertex sets reproduces the
// export exactly.
//
//   k_fill32       neighbours := -1, table := empty
//   k_span_sample  mean vertex-id span of 64k sampled tets (decides the
//                  probe start below; the result does not depend on it)
//   k_face_match   one thread per face slot (4N): sort the face's three
//                  vertex ids, hash them, and probe an open-addressing table
//                  of face references (tet*4 + slot).  An empty entry is
//                  claimed with a CAS; an occupied one is compared exactly by
//                  re-reading the occupant's vertices from the immutable tets
//                  (the 96-bit key is never stored, so no entry is ever seen
//                  half-written).  The second holder of a face marks the
//                  entry paired with an atomic OR, which lands on a line the
//                  probe has just touched, and writes both neighbour entries;
//                  a holder that finds the entry already paired is a third
//                  one (non-manifold input -> EINVAL).
//
// Probe start.  With spatially coherent vertex ids (small spans) a probe
// starts in a 32-entry "home" line picked by the face's smallest vertex id,
// so the faces of a wave's tets, which share few vertices, claim and pair on
// few lines (1.4x faster than hashed starts on a lattice-ordered mesh); after
// 128 probes a crowded home run continues from a hashed start.  With random
// vertex ids home lines only crowd the table, and every probe starts hashed.
//
// Table entries are 32-bit for N < 2^29 and 64-bit beyond (the top bit is the
// paired mark); the table has a power of two >= 6N entries, so it is at most
// two thirds full even when every face is a hull face (4N distinct faces), and
// every probe sequence ends at an empty entry or a match.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <vector>

#include "wtv_internal.h"

namespace wtv {
namespace {

constexpr int kB = 256;
constexpr uint32_t kMaxWgs = 1u << 20;  // grid-stride beyond 2^28 threads
constexpr uint64_t kHomeLine = 32;      // probes start in a 32-entry line picked by the smallest vertex id
constexpr uint64_t kHomeProbes = 128;   // after that many probes a face continues from a hashed start
constexpr uint64_t kSpanSamples = 65536;  // tets sampled to decide between home and hashed starts
constexpr double kSpanDiv = 16.0;         // home starts when the mean tet span is below nv / kSpanDiv

// error codes written to the device flag
constexpr int32_t kBadTet = 1;       // vertex id outside [0, nv) or repeated in a tet
constexpr int32_t kNonManifold = 2;  // a face held by three or more tets
constexpr int32_t kTableFull = 3;    // probe sequence exhausted (cannot happen for T > 4N)

inline void ckt(hipError_t e, const char* where) {
    if (e == hipSuccess) return;
    fail(WTV_EDEVICE, std::string(kPrefix) + "HIP at " + where + ": " + hipGetErrorString(e));
}

__global__ void __launch_bounds__(kB) k_fill32(uint32_t* __restrict__ p, uint64_t n, uint32_t v) {
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kB;
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(kB) + threadIdx.x; i < n; i += stride) p[i] = v;
}

__device__ __forceinline__ void flag_err(int32_t* p, int32_t v) {
    __hip_atomic_exchange(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// The three vertex ids of face fi of tet v, ascending: slot fi leaves out
// local vertex 3-fi.
__device__ __forceinline__ void face_sorted(const int4& v, int fi, uint32_t& a, uint32_t& b, uint32_t& c) {
    uint32_t x = static_cast<uint32_t>(fi == 3 ? v.y : v.x);
    uint32_t y = static_cast<uint32_t>(fi >= 2 ? v.z : v.y);
    uint32_t z = static_cast<uint32_t>(fi == 0 ? v.z : v.w);
    const uint32_t lo = min(x, y), hi = max(x, y);
    x = lo;
    y = hi;
    const uint32_t m = min(y, z);
    z = max(y, z);
    y = m;
    a = min(x, y);
    b = max(x, y);
    c = z;
}

__device__ __forceinline__ uint64_t face_hash(uint32_t a, uint32_t b, uint32_t c) {
    uint64_t k = ((static_cast<uint64_t>(a) << 32) | b) * 0x9E3779B97F4A7C15ull;
    k ^= static_cast<uint64_t>(c) * 0xC2B2AE3D27D4EB4Full;
    k ^= k >> 31;
    k *= 0xBF58476D1CE4E5B9ull;
    k ^= k >> 29;
    k *= 0x94D049BB133111EBull;
    k ^= k >> 32;
    return k;
}

// Vertex-id span (largest id - smallest id) of sampled tets, summed per
// workgroup.  Small spans mean spatially coherent vertex numbering: then the
// faces of a tet and of its neighbours share home lines of nearby ids.
__global__ void __launch_bounds__(kB) k_span_sample(uint64_t N, uint64_t step, const int4* __restrict__ tets4,
                                                    unsigned long long* __restrict__ part) {
    __shared__ unsigned long long red[kB / 64];
    const uint64_t q = (blockIdx.x * static_cast<uint64_t>(kB) + threadIdx.x) * step;
    unsigned long long sp = 0;
    if (q < N) {
        const int4 v = tets4[q];
        const uint32_t x = static_cast<uint32_t>(v.x), y = static_cast<uint32_t>(v.y), z = static_cast<uint32_t>(v.z),
                       w = static_cast<uint32_t>(v.w);
        sp = max(max(x, y), max(z, w)) - min(min(x, y), min(z, w));
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) sp += __shfl_xor(sp, o);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = sp;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t = 0;
        for (int k = 0; k < kB / 64; ++k) t += red[k];
        part[blockIdx.x] = t;
    }
}

// Table entry: a face reference (tet*4 + slot) whose top bit marks a face
// that has found its second holder.
template <class E>
__global__ void __launch_bounds__(kB) k_face_match(uint64_t slots, uint32_t nv, const int4* __restrict__ tets4,
                                                   E* __restrict__ table, uint64_t mask, double home_scale,
                                                   int home_mode, int32_t* __restrict__ nbr,
                                                   int32_t* __restrict__ err) {
    constexpr E kEmpty = static_cast<E>(~static_cast<E>(0));
    constexpr E kPaired = static_cast<E>(static_cast<E>(1) << (sizeof(E) * 8 - 1));
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kB;
    for (uint64_t t = blockIdx.x * static_cast<uint64_t>(kB) + threadIdx.x; t < slots; t += stride) {
        const uint64_t i = t >> 2;
        const int fi = static_cast<int>(t & 3);
        const int4 v = tets4[i];
        const bool in_range = static_cast<uint32_t>(v.x) < nv && static_cast<uint32_t>(v.y) < nv &&
                              static_cast<uint32_t>(v.z) < nv && static_cast<uint32_t>(v.w) < nv;
        const bool distinct = v.x != v.y && v.x != v.z && v.x != v.w && v.y != v.z && v.y != v.w && v.z != v.w;
        if (!(in_range && distinct)) {
            if (fi == 0) flag_err(err, kBadTet);
            continue;
        }
        uint32_t a, b, c;
        face_sorted(v, fi, a, b, c);
        const E me = static_cast<E>(t);
        const uint64_t hf = face_hash(a, b, c);
        // home mode: start in the line of the smallest vertex id, so the faces
        // of nearby tets (few, nearby vertex ids) probe and CAS few lines
        uint64_t h = hf & mask;
        if (home_mode)
            h = ((static_cast<uint64_t>(static_cast<double>(a) * home_scale) & ~uint64_t{kHomeLine - 1}) +
                 (hf & (kHomeLine - 1))) & mask;
        bool done = false;
        for (uint64_t p = 0; p <= mask + kHomeProbes && !done; ++p, h = (h + 1) & mask) {
            if (home_mode && p == kHomeProbes) h = (hf >> 20) & mask;  // crowded home run: hashed restart
            E cur = table[h];  // entries change only from empty: a stale empty costs a CAS
            if (cur == kEmpty) {
                cur = atomicCAS(&table[h], kEmpty, me);
                if (cur == kEmpty) {
                    done = true;  // first holder of this face: the partner writes both entries
                    break;
                }
            }
            const E ref = static_cast<E>(cur & ~kPaired);
            const uint64_t j = static_cast<uint64_t>(ref) >> 2;
            if (j == i) continue;  // another face of this tet
            const int fj = static_cast<int>(ref & 3);
            uint32_t a2, b2, c2;
            face_sorted(tets4[j], fj, a2, b2, c2);
            if (a2 == a && b2 == b && c2 == c) {
                const E old = atomicOr(&table[h], kPaired);
                if (old & kPaired) {
                    flag_err(err, kNonManifold);  // the face already has two holders
                } else {
                    nbr[t] = static_cast<int32_t>(j);
                    nbr[j * 4 + fj] = static_cast<int32_t>(i);
                }
                done = true;
            }
        }
        if (!done) flag_err(err, kTableFull);
    }
}

struct DevBuf {
    void* p = nullptr;
    ~DevBuf() { if (p) (void)hipFree(p); }
};

uint32_t grid_for(uint64_t threads) {
    return static_cast<uint32_t>(std::min<uint64_t>(std::max<uint64_t>((threads + kB - 1) / kB, 1), kMaxWgs));
}

// Samples up to kSpanSamples tets evenly over the mesh; true when their mean
// vertex-id span is below nv / kSpanDiv.  The choice changes only where
// probes start, never the result.
bool locality(const int32_t* d_tets, size_t N, size_t nv, hipStream_t s) {
    const uint64_t S = std::min<uint64_t>(N, kSpanSamples);
    const uint64_t step = N / S;
    const uint32_t blocks = static_cast<uint32_t>((S + kB - 1) / kB);
    DevBuf sp;
    ckt(hipMalloc(&sp.p, blocks * sizeof(unsigned long long)), "span alloc");
    hipLaunchKernelGGL(k_span_sample, dim3(blocks), dim3(kB), 0, s, static_cast<uint64_t>(S) * step, step,
                       reinterpret_cast<const int4*>(d_tets), static_cast<unsigned long long*>(sp.p));
    ckt(hipGetLastError(), "k_span_sample launch");
    std::vector<unsigned long long> h(blocks);
    ckt(hipMemcpyAsync(h.data(), sp.p, blocks * sizeof(unsigned long long), hipMemcpyDeviceToHost, s), "span read");
    ckt(hipStreamSynchronize(s), "span sync");
    double sum = 0;
    for (unsigned long long x : h) sum += static_cast<double>(x);
    return sum / static_cast<double>(S) < static_cast<double>(nv) / kSpanDiv;
}

}  // namespace

void tet_neighbors(const int32_t* d_tets, size_t N, size_t nv, int32_t* d_nbr, void* stream_v) {
    if (N == 0) return;
    if (!d_tets || !d_nbr) fail(WTV_EINVAL, std::string(kPrefix) + "null tets/neighbors pointer");
    if (N > 0x7FFFFFFFull) fail(WTV_EINVAL, std::string(kPrefix) + "tet ids must fit int32 (N < 2^31)");
    if (nv == 0 || nv > 0x80000000ull) fail(WTV_EINVAL, std::string(kPrefix) + "vertex count must be in [1, 2^31]");
    auto s = static_cast<hipStream_t>(stream_v);
    const uint64_t slots = 4ull * N;
    uint64_t T = 1024;
    while (T < 6ull * N) T <<= 1;
    const bool wide = slots > 0x7FFFFFFFull;  // references need the top bit free for the paired mark
    const size_t esz = wide ? 8u : 4u;
    DevBuf table, err;
    const hipError_t e = hipMalloc(&table.p, T * esz);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        fail(WTV_ENOMEM, std::string(kPrefix) + "face table of " + std::to_string(T * esz) + " bytes does not fit");
    }
    ckt(hipMalloc(&err.p, sizeof(int32_t)), "err alloc");
    int32_t* d_err = static_cast<int32_t*>(err.p);
    hipLaunchKernelGGL(k_fill32, dim3(1), dim3(kB), 0, s, reinterpret_cast<uint32_t*>(d_err), 1ull, 0u);
    hipLaunchKernelGGL(k_fill32, dim3(grid_for(slots)), dim3(kB), 0, s, reinterpret_cast<uint32_t*>(d_nbr), slots,
                       0xFFFFFFFFu);
    const uint64_t twords = T * esz / 4;
    hipLaunchKernelGGL(k_fill32, dim3(grid_for(twords)), dim3(kB), 0, s, static_cast<uint32_t*>(table.p), twords,
                       0xFFFFFFFFu);
    const double home_scale = static_cast<double>(T) / static_cast<double>(nv);
    const auto* t4 = reinterpret_cast<const int4*>(d_tets);
    const int home_mode = locality(d_tets, N, nv, s) ? 1 : 0;
    if (wide)
        hipLaunchKernelGGL(k_face_match<unsigned long long>, dim3(grid_for(slots)), dim3(kB), 0, s, slots,
                           static_cast<uint32_t>(nv), t4, static_cast<unsigned long long*>(table.p), T - 1,
                           home_scale, home_mode, d_nbr, d_err);
    else
        hipLaunchKernelGGL(k_face_match<uint32_t>, dim3(grid_for(slots)), dim3(kB), 0, s, slots,
                           static_cast<uint32_t>(nv), t4, static_cast<uint32_t*>(table.p), T - 1, home_scale,
                           home_mode, d_nbr, d_err);
    ckt(hipGetLastError(), "k_face_match launch");
    int32_t h_err = 0;
    ckt(hipMemcpyAsync(&h_err, d_err, sizeof(h_err), hipMemcpyDeviceToHost, s), "err read");
    ckt(hipStreamSynchronize(s), "tet_neighbors sync");
    if (h_err == kBadTet)
        fail(WTV_EINVAL, std::string(kPrefix) + "tets hold a vertex id outside [0, nv) or a repeated vertex");
    if (h_err == kNonManifold) fail(WTV_EINVAL, std::string(kPrefix) + "a face is shared by more than two tets");
    if (h_err != 0) fail(WTV_EINVARIANT, std::string(kPrefix) + "face table exhausted (code 3)");
}

}  // namespace wtv
