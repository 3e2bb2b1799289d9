// wtv_upload.cpp -- pinned, multi-stream copies between host memory and HBM.
//
// topology_upload: an [rows, 4] int32 topology matrix into HBM.
// Replaces topology_to_cuda_chunked (wtivo.py:284-309), which repacks
// 262,144-row slices of a possibly Fortran-ordered array (Eigen's MatrixXi is
// column-major) into a pageable buffer and copies each slice synchronously.
// Here W worker threads each own a HIP stream and two pinned staging buffers
// of kChunkRows rows: a worker packs chunk c (a memcpy for a C-ordered source,
// a 4-stream interleave for a column-contiguous one, a strided gather
// otherwise) into one buffer while the DMA of its previous chunk runs from the
// other.  Chunks are dealt round-robin, so W copies are in flight at once.
//
// pinned_to_device / pinned_to_host: the same pipeline for a plain byte range
// (vertex arrays and labels of the cut): pageable memory is copied through
// the pinned buffers by several threads instead of one runtime-staged copy.
//
// The pinned buffers are allocated with hipHostMalloc on first use and kept
// for the life of the process (at most kMaxWorkers * 2 * 4 MiB); copies on
// one process are serialised by a mutex.  Every call is synchronous: it
// returns once the data has arrived, after the work already queued on the
// caller's stream (which may still be writing or releasing that memory) has
// finished.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "wtv_internal.h"

namespace wtv {
namespace {

constexpr size_t kChunkRows = 262144;  // 4 MiB of int32x4 rows per staging buffer
constexpr size_t kChunkBytes = kChunkRows * 16u;
constexpr int kMaxWorkers = 16;

inline void cku(hipError_t e, const char* where) {
    if (e == hipSuccess) return;
    fail(WTV_EDEVICE, std::string(kPrefix) + "HIP at " + where + ": " + hipGetErrorString(e));
}

struct Worker {
    int device = -1;
    hipStream_t stream = nullptr;
    int32_t* pinned[2] = {nullptr, nullptr};
    hipEvent_t done[2] = {nullptr, nullptr};
};

std::mutex g_mu;
Worker g_workers[kMaxWorkers];

// Creates the worker's stream, events and pinned buffers on first use (and
// its stream and events again if the current device changed).
void ready(Worker& w, int device) {
    if (w.device != device) {
        if (w.stream) (void)hipStreamDestroy(w.stream);
        for (auto& e : w.done)
            if (e) (void)hipEventDestroy(e);
        w.stream = nullptr;
        w.done[0] = w.done[1] = nullptr;
        cku(hipStreamCreateWithFlags(&w.stream, hipStreamNonBlocking), "upload stream");
        for (auto& e : w.done) cku(hipEventCreateWithFlags(&e, hipEventDisableTiming), "upload event");
        w.device = device;
    }
    for (auto& p : w.pinned)
        if (!p) cku(hipHostMalloc(reinterpret_cast<void**>(&p), kChunkRows * 16u, hipHostMallocPortable),
                    "upload pinned buffer");
}

// Packs rows [r0, r1) of the strided source into dst as row-major int32x4.
void pack(const int32_t* src, int64_t rs, int64_t cs, size_t r0, size_t r1, int32_t* dst) {
    const size_t n = r1 - r0;
    if (rs == 4 && cs == 1) {
        std::memcpy(dst, src + r0 * 4, n * 16u);
        return;
    }
    if (rs == 1) {
        const int32_t* c0 = src + r0;
        const int32_t* c1 = c0 + cs;
        const int32_t* c2 = c1 + cs;
        const int32_t* c3 = c2 + cs;
        for (size_t r = 0; r < n; ++r) {
            dst[4 * r + 0] = c0[r];
            dst[4 * r + 1] = c1[r];
            dst[4 * r + 2] = c2[r];
            dst[4 * r + 3] = c3[r];
        }
        return;
    }
    for (size_t r = 0; r < n; ++r) {
        const int32_t* row = src + static_cast<int64_t>(r0 + r) * rs;
        for (int k = 0; k < 4; ++k) dst[4 * r + k] = row[k * cs];
    }
}

// Runs `chunks` chunks of at most kChunkBytes over W workers, each with its
// own stream and two pinned buffers, after the work queued on `caller`.
// Host to device: fill(c, buf) packs chunk c into the pinned buffer, which is
// copied to dev + c * kChunkBytes while the worker packs its next chunk.  Device to host: the chunk at dev + c *
// kChunkBytes is copied into a pinned buffer and handed to drain(c, buf,
// bytes) once it has arrived, while the next chunk's copy runs.
template <class Fill, class Drain>
void pipeline(bool to_device, size_t chunks, size_t total_bytes, int threads, char* dev, hipStream_t caller,
              Fill fill, Drain drain) {
    using Clock = std::chrono::steady_clock;
    const auto t0 = Clock::now();
    std::lock_guard<std::mutex> lk(g_mu);
    int device = 0;
    cku(hipGetDevice(&device), "hipGetDevice");
    int W = threads > 0 ? threads : static_cast<int>(std::thread::hardware_concurrency());
    W = static_cast<int>(std::min<size_t>(std::max(1, std::min(W, kMaxWorkers)), chunks));
    for (int w = 0; w < W; ++w) ready(g_workers[w], device);

    // the caller's queued work first: dev may be memory it just released or
    // data it is still writing
    hipEvent_t start = nullptr;
    cku(hipEventCreateWithFlags(&start, hipEventDisableTiming), "copy start event");
    struct EvGuard {
        hipEvent_t e;
        ~EvGuard() { (void)hipEventDestroy(e); }
    } guard{start};
    cku(hipEventRecord(start, caller), "copy start record");
    for (int w = 0; w < W; ++w) cku(hipStreamWaitEvent(g_workers[w].stream, start, 0), "copy wait");

    std::vector<int> status(W, WTV_OK);
    std::vector<std::string> msg(W);
    auto run = [&](int w) {
        Worker& wk = g_workers[w];
        try {
            cku(hipSetDevice(device), "hipSetDevice");
            bool used[2] = {false, false};
            size_t held[2] = {0, 0}, held_bytes[2] = {0, 0};
            int b = 0;
            for (size_t c = static_cast<size_t>(w); c < chunks; c += static_cast<size_t>(W), b ^= 1) {
                const size_t bytes = std::min(kChunkBytes, total_bytes - c * kChunkBytes);
                if (used[b]) {
                    cku(hipEventSynchronize(wk.done[b]), "copy buffer wait");
                    if (!to_device) drain(held[b], reinterpret_cast<const char*>(wk.pinned[b]), held_bytes[b]);
                }
                if (to_device) {
                    fill(c, reinterpret_cast<char*>(wk.pinned[b]));
                    cku(hipMemcpyAsync(dev + c * kChunkBytes, wk.pinned[b], bytes, hipMemcpyHostToDevice,
                                       wk.stream), "copy H2D");
                } else {
                    cku(hipMemcpyAsync(wk.pinned[b], dev + c * kChunkBytes, bytes, hipMemcpyDeviceToHost,
                                       wk.stream), "copy D2H");
                }
                cku(hipEventRecord(wk.done[b], wk.stream), "copy record");
                used[b] = true;
                held[b] = c;
                held_bytes[b] = bytes;
            }
            cku(hipStreamSynchronize(wk.stream), "copy sync");
            if (!to_device)
                for (int q = 0; q < 2; ++q)
                    if (used[q]) drain(held[q], reinterpret_cast<const char*>(wk.pinned[q]), held_bytes[q]);
        } catch (const Error& e) {
            status[w] = e.code;
            msg[w] = e.what();
            (void)hipStreamSynchronize(wk.stream);
        }
    };
    const auto t1 = Clock::now();
    std::vector<std::thread> pool;
    pool.reserve(W - 1);
    for (int w = 1; w < W; ++w) pool.emplace_back(run, w);
    run(0);
    const auto t2 = Clock::now();
    for (auto& t : pool) t.join();
    const auto t3 = Clock::now();
    static const bool trace = [] {
        const char* e = std::getenv("WTV_TRACE");
        return e && *e && *e != '0';
    }();
    if (trace)
        std::fprintf(stderr, "[wtv] copy %s %zu B, %d workers: setup %.3f ms, worker 0 %.3f ms, join %.3f ms\n",
                     to_device ? "H2D" : "D2H", total_bytes, W,
                     std::chrono::duration<double, std::milli>(t1 - t0).count(),
                     std::chrono::duration<double, std::milli>(t2 - t1).count(),
                     std::chrono::duration<double, std::milli>(t3 - t2).count());
    for (int w = 0; w < W; ++w)
        if (status[w] != WTV_OK) fail(status[w], msg[w]);
}

}  // namespace

void topology_upload(const int32_t* host, size_t rows, int64_t row_stride, int64_t col_stride, int32_t* d_out,
                     int threads, void* stream_v) {
    if (rows == 0) return;
    if (!host || !d_out) fail(WTV_EINVAL, std::string(kPrefix) + "null topology pointer");
    if (row_stride == 0 && rows > 1) fail(WTV_EINVAL, std::string(kPrefix) + "row stride 0 with more than one row");
    const size_t chunks = (rows + kChunkRows - 1) / kChunkRows;
    pipeline(true, chunks, rows * 16u, threads, reinterpret_cast<char*>(d_out), static_cast<hipStream_t>(stream_v),
             [&](size_t c, char* buf) {
                 const size_t r0 = c * kChunkRows, r1 = std::min(rows, r0 + kChunkRows);
                 pack(host, row_stride, col_stride, r0, r1, reinterpret_cast<int32_t*>(buf));
             },
             [](size_t, const char*, size_t) {});
}

void pinned_to_device(const void* host, size_t bytes, void* d_out, int threads, void* stream_v) {
    if (bytes == 0) return;
    if (!host || !d_out) fail(WTV_EINVAL, std::string(kPrefix) + "null copy pointer");
    const char* src = static_cast<const char*>(host);
    pipeline(true, (bytes + kChunkBytes - 1) / kChunkBytes, bytes, threads, static_cast<char*>(d_out),
             static_cast<hipStream_t>(stream_v),
             [&](size_t c, char* buf) {
                 std::memcpy(buf, src + c * kChunkBytes, std::min(kChunkBytes, bytes - c * kChunkBytes));
             },
             [](size_t, const char*, size_t) {});
}

void pinned_to_host(const void* d_src, size_t bytes, void* host, int threads, void* stream_v) {
    if (bytes == 0) return;
    if (!host || !d_src) fail(WTV_EINVAL, std::string(kPrefix) + "null copy pointer");
    char* dst = static_cast<char*>(host);
    pipeline(false, (bytes + kChunkBytes - 1) / kChunkBytes, bytes, threads,
             const_cast<char*>(static_cast<const char*>(d_src)), static_cast<hipStream_t>(stream_v),
             [](size_t, char*) {},
             [&](size_t c, const char* buf, size_t n) { std::memcpy(dst + c * kChunkBytes, buf, n); });
}

}  // namespace wtv
