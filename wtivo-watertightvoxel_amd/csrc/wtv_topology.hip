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
//   k_fill32      neighbours := -1, table := empty
//   k_face_match  one thread per face slot (4N): sort the face's three vertex
//                 ids, hash them, and probe an open-addressing table of face
//                 references (tet*4 + slot).  An empty entry is claimed with a
//                 CAS; an occupied one is compared exactly by re-reading the
//                 occupant's vertices from the immutable tets (the 96-bit key
//                 is never stored, so no entry is ever seen half-written).  On
//                 a match both neighbour entries are written; the partner's
//                 with a CAS from -1, which fails only when a third tet holds
//                 the same face (non-manifold input -> EINVAL).
//
// Table entries are 32-bit for N < 2^30 and 64-bit beyond; the table has a
// power of two >= 6N entries, so a closed mesh (about 2N distinct faces) fills
// a third of it at most and every probe sequence ends at an empty entry or a
// match.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>

#include "wtv_internal.h"

namespace wtv {
namespace {

constexpr int kB = 256;
constexpr uint32_t kMaxWgs = 1u << 20;  // grid-stride beyond 2^28 threads

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

template <class E>
__global__ void __launch_bounds__(kB) k_face_match(uint64_t slots, uint32_t nv, const int4* __restrict__ tets4,
                                                   E* __restrict__ table, uint64_t mask, int32_t* __restrict__ nbr,
                                                   int32_t* __restrict__ err) {
    constexpr E kEmpty = static_cast<E>(~static_cast<E>(0));
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
        uint64_t h = face_hash(a, b, c) & mask;
        bool done = false;
        for (uint64_t p = 0; p <= mask && !done; ++p, h = (h + 1) & mask) {
            E cur = table[h];  // entries change once, empty -> reference: a stale empty only costs a CAS
            if (cur == kEmpty) {
                cur = atomicCAS(&table[h], kEmpty, me);
                if (cur == kEmpty) {
                    done = true;  // first holder of this face: the partner writes both entries
                    break;
                }
            }
            const uint64_t j = static_cast<uint64_t>(cur) >> 2;
            if (j == i) continue;  // another face of this tet
            const int fj = static_cast<int>(cur & 3);
            uint32_t a2, b2, c2;
            face_sorted(tets4[j], fj, a2, b2, c2);
            if (a2 == a && b2 == b && c2 == c) {
                nbr[t] = static_cast<int32_t>(j);
                const int32_t old = atomicCAS(&nbr[j * 4 + fj], -1, static_cast<int32_t>(i));
                if (old != -1) flag_err(err, kNonManifold);
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
    const bool wide = slots > 0xFFFFFFFEull;
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
    const auto* t4 = reinterpret_cast<const int4*>(d_tets);
    if (wide)
        hipLaunchKernelGGL(k_face_match<unsigned long long>, dim3(grid_for(slots)), dim3(kB), 0, s, slots,
                           static_cast<uint32_t>(nv), t4, static_cast<unsigned long long*>(table.p), T - 1, d_nbr,
                           d_err);
    else
        hipLaunchKernelGGL(k_face_match<uint32_t>, dim3(grid_for(slots)), dim3(kB), 0, s, slots,
                           static_cast<uint32_t>(nv), t4, static_cast<uint32_t*>(table.p), T - 1, d_nbr, d_err);
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
