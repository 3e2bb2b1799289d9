// wtv_udf.hip -- tet labels from a sparse distance grid, on the device.
//
// Restates SparseUDF.sample_tet_labels (native/vdb/wtivo_vdb.cpp:1519-1584):
// for every tet, the FP64 centroid (the four vertices summed in order, times
// 0.25) is mapped into the padded unit box, ((c - bbox_min) / size) *
// (1 - 2 padding) + padding, and the grid is sampled there with OpenVDB's
// BoxSampler in world space; label = 0 when the sample is <= threshold, else 1.
//
// The grid is the one the reference samples: a float grid with a uniform
// linear transform (Transform::createLinearTransform(voxel_size), so index =
// world * (1 / voxel_size)), stored sparsely as 8^3 leaves.  The caller hands
// over the leaves (index-space origins and 512 values each, OpenVDB's leaf
// order offset = (x&7)<<6 | (y&7)<<3 | (z&7)) and the background value that
// every voxel outside a leaf holds.  In HBM the grid is a leaf pool plus a
// dense int32 index over the leaf-cell bounding box (-1 = background).
//
// Sampling follows the published BoxSampler (openvdb/tools/Interpolation.h):
// i = floor(p) per axis, uvw = p - i, the 2x2x2 values v[dx][dy][dz] at
// i + (dx,dy,dz), and three rounds of lerp(a, b, w) = a + float((b - a) * w),
// with (b - a) in float, the product in double and one rounding back to float:
// first along z, then y, then x.  No FMA contraction anywhere (pragma below),
// so every label equals the host restatement's bit for bit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <memory>
#include <string>
#include <vector>

#include "wtv_internal.h"

#pragma clang fp contract(off)

namespace wtv {
namespace {

constexpr int kB = 256;
constexpr uint32_t kMaxWgs = 1u << 20;

inline void cku(hipError_t e, const char* where) {
    if (e == hipSuccess) return;
    fail(WTV_EDEVICE, std::string(kPrefix) + "HIP at " + where + ": " + hipGetErrorString(e));
}

uint32_t grid_for(uint64_t threads) {
    return static_cast<uint32_t>(std::min<uint64_t>(std::max<uint64_t>((threads + kB - 1) / kB, 1), kMaxWgs));
}

struct DevGrid {
    const float* values;   // [leaves][512]
    const int32_t* index;  // [dz][dy][dx] leaf cells -> leaf id or -1
    int32_t ox, oy, oz;    // first leaf cell (origin / 8)
    int32_t dx, dy, dz;    // leaf-cell box extent
    float background;
    double inv_voxel;      // 1 / voxel_size
};

__global__ void __launch_bounds__(kB) k_fill_i32(int32_t* __restrict__ p, uint64_t n, int32_t v) {
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kB;
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(kB) + threadIdx.x; i < n; i += stride) p[i] = v;
}

// index[cell of leaf l] = l; a second leaf on the same cell sets *dup.
__global__ void __launch_bounds__(kB) k_index_leaves(uint64_t L, const int32_t* __restrict__ origin, int32_t ox,
                                                     int32_t oy, int32_t oz, int32_t dx, int32_t dy,
                                                     int32_t* __restrict__ index, int32_t* __restrict__ bad) {
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kB;
    for (uint64_t l = blockIdx.x * static_cast<uint64_t>(kB) + threadIdx.x; l < L; l += stride) {
        const int32_t cx = (origin[3 * l + 0] >> 3) - ox;
        const int32_t cy = (origin[3 * l + 1] >> 3) - oy;
        const int32_t cz = (origin[3 * l + 2] >> 3) - oz;
        const uint64_t cell = (static_cast<uint64_t>(cz) * dy + cy) * dx + cx;
        if (atomicCAS(&index[cell], -1, static_cast<int32_t>(l)) != -1)
            __hip_atomic_exchange(bad, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

__device__ __forceinline__ float grid_value(const DevGrid& g, int64_t x, int64_t y, int64_t z) {
    const int64_t cx = (x >> 3) - g.ox, cy = (y >> 3) - g.oy, cz = (z >> 3) - g.oz;
    if (cx < 0 || cy < 0 || cz < 0 || cx >= g.dx || cy >= g.dy || cz >= g.dz) return g.background;
    const int32_t leaf = g.index[(cz * g.dy + cy) * g.dx + cx];
    if (leaf < 0) return g.background;
    const uint32_t off = (static_cast<uint32_t>(x & 7) << 6) | (static_cast<uint32_t>(y & 7) << 3) |
                         static_cast<uint32_t>(z & 7);
    return g.values[static_cast<uint64_t>(leaf) * 512u + off];
}

__device__ __forceinline__ float lerp_vdb(float a, float b, double w) {
    const double t = static_cast<double>(b - a) * w;
    return a + static_cast<float>(t);
}

// Index coordinate -> (floor, fraction).  A coordinate beyond +-2^40 lies
// outside every grid; it is clamped so the cell lookup sees background.
__device__ __forceinline__ int64_t split(double p, double& frac) {
    const double f = floor(p);
    frac = p - f;
    return static_cast<int64_t>(fmin(fmax(f, -1099511627776.0), 1099511627776.0));
}

__global__ void __launch_bounds__(kB) k_sample_labels(uint64_t N, const int4* __restrict__ tets4,
                                                      const double* __restrict__ V, uint64_t nv, DevGrid g,
                                                      double bx, double by, double bz, double sx, double sy,
                                                      double sz, double scale, double padding, double threshold,
                                                      uint8_t* __restrict__ labels, int32_t* __restrict__ bad) {
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kB;
    for (uint64_t i = blockIdx.x * static_cast<uint64_t>(kB) + threadIdx.x; i < N; i += stride) {
        const int4 t = tets4[i];
        const int32_t q[4] = {t.x, t.y, t.z, t.w};
        double cx = 0.0, cy = 0.0, cz = 0.0;
        bool ok = true;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if (q[k] < 0 || static_cast<uint64_t>(q[k]) >= nv) {
                ok = false;
                continue;
            }
            const double* p = V + static_cast<uint64_t>(q[k]) * 3u;
            cx += p[0];
            cy += p[1];
            cz += p[2];
        }
        if (!ok) {
            __hip_atomic_exchange(bad, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            labels[i] = 1;
            continue;
        }
        cx *= 0.25;
        cy *= 0.25;
        cz *= 0.25;
        const double nx = ((cx - bx) / sx) * scale + padding;
        const double ny = ((cy - by) / sy) * scale + padding;
        const double nz = ((cz - bz) / sz) * scale + padding;
        double u, v, w;
        const int64_t ix = split(nx * g.inv_voxel, u);
        const int64_t iy = split(ny * g.inv_voxel, v);
        const int64_t iz = split(nz * g.inv_voxel, w);
        const float v000 = grid_value(g, ix, iy, iz), v001 = grid_value(g, ix, iy, iz + 1);
        const float v010 = grid_value(g, ix, iy + 1, iz), v011 = grid_value(g, ix, iy + 1, iz + 1);
        const float v100 = grid_value(g, ix + 1, iy, iz), v101 = grid_value(g, ix + 1, iy, iz + 1);
        const float v110 = grid_value(g, ix + 1, iy + 1, iz), v111 = grid_value(g, ix + 1, iy + 1, iz + 1);
        const float d = lerp_vdb(lerp_vdb(lerp_vdb(v000, v001, w), lerp_vdb(v010, v011, w), v),
                                 lerp_vdb(lerp_vdb(v100, v101, w), lerp_vdb(v110, v111, w), v), u);
        labels[i] = static_cast<double>(d) <= threshold ? 0 : 1;
    }
}

struct Udf {
    int device = 0;
    void* values = nullptr;
    void* index = nullptr;
    DevGrid g{};
    uint64_t leaves = 0;
    uint64_t bytes = 0;
    ~Udf() {
        if (values) (void)hipFree(values);
        if (index) (void)hipFree(index);
    }
};

struct DevBuf {
    void* p = nullptr;
    ~DevBuf() { if (p) (void)hipFree(p); }
};

}  // namespace

void* udf_create(const int32_t* origin, const float* values, size_t L, float background, double voxel_size,
                 void* stream_v) {
    if (L == 0) fail(WTV_EINVAL, std::string(kPrefix) + "SparseUDF: grid is empty");
    if (!origin || !values) fail(WTV_EINVAL, std::string(kPrefix) + "SparseUDF: null leaf arrays");
    if (!(voxel_size > 0.0) || !std::isfinite(voxel_size))
        fail(WTV_EINVAL, std::string(kPrefix) + "SparseUDF: voxel_size must be > 0");
    if (L > 0x7FFFFFFFull) fail(WTV_EINVAL, std::string(kPrefix) + "SparseUDF: more than 2^31-1 leaves");
    int64_t lo[3] = {INT64_MAX, INT64_MAX, INT64_MAX}, hi[3] = {INT64_MIN, INT64_MIN, INT64_MIN};
    for (size_t l = 0; l < L; ++l)
        for (int a = 0; a < 3; ++a) {
            const int32_t o = origin[3 * l + a];
            if (o & 7) fail(WTV_EINVAL, std::string(kPrefix) + "SparseUDF: leaf origins must be multiples of 8");
            lo[a] = std::min<int64_t>(lo[a], o >> 3);
            hi[a] = std::max<int64_t>(hi[a], o >> 3);
        }
    const int64_t ext[3] = {hi[0] - lo[0] + 1, hi[1] - lo[1] + 1, hi[2] - lo[2] + 1};
    const double cells = static_cast<double>(ext[0]) * static_cast<double>(ext[1]) * static_cast<double>(ext[2]);
    if (cells > 2147483647.0)
        fail(WTV_ENOMEM, std::string(kPrefix) + "SparseUDF: leaf-cell box of " + std::to_string(cells) +
                             " cells exceeds 2^31");
    auto s = static_cast<hipStream_t>(stream_v);
    auto u = std::make_unique<Udf>();
    cku(hipGetDevice(&u->device), "hipGetDevice");
    const uint64_t ncell = static_cast<uint64_t>(cells);
    hipError_t e = hipMalloc(&u->values, L * 512u * sizeof(float));
    if (e == hipSuccess) e = hipMalloc(&u->index, ncell * sizeof(int32_t));
    if (e != hipSuccess) {
        (void)hipGetLastError();
        fail(WTV_ENOMEM, std::string(kPrefix) + "SparseUDF: device grid does not fit");
    }
    DevBuf d_origin, d_bad;
    cku(hipMalloc(&d_origin.p, L * 3u * sizeof(int32_t)), "origin alloc");
    cku(hipMalloc(&d_bad.p, sizeof(int32_t)), "flag alloc");
    pinned_to_device(values, L * 512u * sizeof(float), u->values, 0, s);
    cku(hipMemcpyAsync(d_origin.p, origin, L * 3u * sizeof(int32_t), hipMemcpyHostToDevice, s), "origin upload");
    auto* idx = static_cast<int32_t*>(u->index);
    auto* bad = static_cast<int32_t*>(d_bad.p);
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(kB), 0, s, bad, 1ull, 0);
    hipLaunchKernelGGL(k_fill_i32, dim3(grid_for(ncell)), dim3(kB), 0, s, idx, ncell, -1);
    hipLaunchKernelGGL(k_index_leaves, dim3(grid_for(L)), dim3(kB), 0, s, static_cast<uint64_t>(L),
                       static_cast<const int32_t*>(d_origin.p), static_cast<int32_t>(lo[0]),
                       static_cast<int32_t>(lo[1]), static_cast<int32_t>(lo[2]), static_cast<int32_t>(ext[0]),
                       static_cast<int32_t>(ext[1]), idx, bad);
    cku(hipGetLastError(), "k_index_leaves launch");
    int32_t h_bad = 0;
    cku(hipMemcpyAsync(&h_bad, bad, sizeof(h_bad), hipMemcpyDeviceToHost, s), "flag read");
    cku(hipStreamSynchronize(s), "udf_create sync");
    if (h_bad) fail(WTV_EINVAL, std::string(kPrefix) + "SparseUDF: two leaves share an origin");
    u->g.values = static_cast<const float*>(u->values);
    u->g.index = idx;
    u->g.ox = static_cast<int32_t>(lo[0]);
    u->g.oy = static_cast<int32_t>(lo[1]);
    u->g.oz = static_cast<int32_t>(lo[2]);
    u->g.dx = static_cast<int32_t>(ext[0]);
    u->g.dy = static_cast<int32_t>(ext[1]);
    u->g.dz = static_cast<int32_t>(ext[2]);
    u->g.background = background;
    u->g.inv_voxel = 1.0 / voxel_size;
    u->leaves = L;
    u->bytes = L * 512u * sizeof(float) + ncell * sizeof(int32_t);
    return u.release();
}

void udf_destroy(void* h) { delete static_cast<Udf*>(h); }

uint64_t udf_bytes(const void* h) { return static_cast<const Udf*>(h)->bytes; }

double udf_sample_tet_labels(const void* h, const double* V, size_t nv, bool vertices_on_device,
                             const int32_t* d_tets, size_t N, const double* bmin, const double* bmax, double padding,
                             double threshold, uint8_t* labels, bool labels_on_device, void* stream_v) {
    if (!h) fail(WTV_EINVAL, std::string(kPrefix) + "SparseUDF.sample_tet_labels: grid is empty");
    if (!bmin || !bmax) fail(WTV_EINVAL, std::string(kPrefix) + "SparseUDF.sample_tet_labels: bbox must have length 3");
    const double sx = bmax[0] - bmin[0], sy = bmax[1] - bmin[1], sz = bmax[2] - bmin[2];
    if (!(sx > 0 && sy > 0 && sz > 0))
        fail(WTV_EINVAL, std::string(kPrefix) + "SparseUDF.sample_tet_labels: degenerate bbox");
    if (N == 0) return 0.0;
    if (!V || !d_tets || !labels) fail(WTV_EINVAL, std::string(kPrefix) + "SparseUDF.sample_tet_labels: null array");
    const auto* u = static_cast<const Udf*>(h);
    auto s = static_cast<hipStream_t>(stream_v);
    const auto t0 = std::chrono::steady_clock::now();
    DevBuf dV, dL, d_bad;
    const double* vdev = V;
    if (!vertices_on_device) {
        cku(hipMalloc(&dV.p, std::max<size_t>(nv, 1) * 24u), "vertices alloc");
        vdev = static_cast<const double*>(dV.p);
    }
    cku(hipMalloc(&d_bad.p, sizeof(int32_t)), "flag alloc");
    uint8_t* out = labels;
    if (!labels_on_device) {
        cku(hipMalloc(&dL.p, N), "labels alloc");
        out = static_cast<uint8_t*>(dL.p);
    }
    if (!vertices_on_device) pinned_to_device(V, nv * 24u, dV.p, 0, s);
    auto* bad = static_cast<int32_t*>(d_bad.p);
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(kB), 0, s, bad, 1ull, 0);
    hipLaunchKernelGGL(k_sample_labels, dim3(grid_for(N)), dim3(kB), 0, s, static_cast<uint64_t>(N),
                       reinterpret_cast<const int4*>(d_tets), vdev,
                       static_cast<uint64_t>(nv), u->g, bmin[0], bmin[1], bmin[2], sx, sy, sz, 1.0 - 2.0 * padding,
                       padding, threshold, out, bad);
    cku(hipGetLastError(), "k_sample_labels launch");
    int32_t h_bad = 0;
    cku(hipMemcpyAsync(&h_bad, bad, sizeof(h_bad), hipMemcpyDeviceToHost, s), "flag read");
    cku(hipStreamSynchronize(s), "sample sync");
    if (!labels_on_device) pinned_to_host(out, N, labels, 0, s);
    if (h_bad) fail(WTV_EINVAL, std::string(kPrefix) + "SparseUDF.sample_tet_labels: tet index out of range");
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace wtv
