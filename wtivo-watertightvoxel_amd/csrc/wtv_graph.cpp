// wtv_graph.cpp -- reduced-graph build + full cut (host side of graph_cut_fast).
//
// Restates graph_cut_gpupr_fast_v630 (native/gpupr/wtivo_gpupr_bindings.cpp:100-397):
//   classify      label 0 -> fixed source; label != 0 with a hull face -> fixed
//                 sink; otherwise free (:153-171); reduced id = rank among free
//                 tets in increasing tet order (:179-188)
//   capacity      capacity_exact_formula (:65-98): FP64 cross product, one
//                 rounding of 1e7*0.5*|c| to fp32, pen = area or
//                 area*(1+fill_T), truncation to int64, [0,2^30) range check
//   fold          arcs to fixed tets become terminal capacity:
//                 term = src - snk, base += min(src, snk) (:265-290)
//   scatter       out = labels; out[free_old[q]] = partition[q] (:357-364)
// Differences by design: the whole topology comes down in one transfer (the
// reference streams 262,144-row chunks through a 4 MiB pageable buffer,
// :139-148,227-232), the reduced graph goes up in one transfer per array, and
// the early-out for free_N == 0 reports the full 17 scalars (the reference
// returns a shorter tuple there, :191-196).
//
// Must be compiled with -ffp-contract=off: the capacity formula is bit-exact
// only without FMA contraction.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <sstream>
#include <thread>
#include <vector>

#include "wtv_internal.h"

namespace wtv {
namespace {

inline void ckh(hipError_t e, const char* where) {
    if (e == hipSuccess) return;
    fail(WTV_EDEVICE, std::string(kPrefix) + "HIP at " + where + ": " + hipGetErrorString(e));
}

// Contiguous ranges over `threads` std::threads; serial below 32,768 items
// (the reference's par_chunks policy, wtivo_gpupr_bindings.cpp:37-52).
template <class Fn>
void par_chunks(size_t n, int threads, Fn fn) {
    if (n == 0) return;
    threads = std::max(1, std::min<int>(threads, static_cast<int>(std::min<size_t>(n, 1u << 30))));
    if (threads == 1 || n < 32768) { fn(size_t{0}, n, 0); return; }
    std::vector<std::thread> ts;
    ts.reserve(static_cast<size_t>(threads));
    const size_t step = (n + static_cast<size_t>(threads) - 1) / static_cast<size_t>(threads);
    for (int t = 0; t < threads; ++t) {
        const size_t a = static_cast<size_t>(t) * step;
        const size_t b = std::min(n, a + step);
        if (a >= b) break;
        ts.emplace_back([=, &fn]() { fn(a, b, t); });
    }
    for (auto& th : ts) th.join();
}

constexpr int kFaceLut[4][3] = {{0, 1, 2}, {0, 1, 3}, {0, 2, 3}, {1, 2, 3}};

// 0 ok; 1 invalid vertex; 2 overflow.
inline int capacity_exact(const double* V, size_t nv, const int32_t* tr, int fi, uint8_t li, uint8_t lj,
                          double fill_T, int32_t* cap_out) {
    const int a = tr[kFaceLut[fi][0]], b = tr[kFaceLut[fi][1]], c = tr[kFaceLut[fi][2]];
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
    const float area_f = static_cast<float>(1.0E7 * 0.5 * std::sqrt(cx * cx + cy * cy + cz * cz));
    const double area = static_cast<double>(area_f);
    const double pen = (li != lj) ? area : area * (1.0 + fill_T);
    if (!(pen > -9.2e18 && pen < 9.2e18)) return 2;
    const int64_t cap64 = static_cast<int64_t>(pen);
    if (cap64 < 0 || cap64 >= 1073741824LL) return 2;
    *cap_out = static_cast<int32_t>(cap64);
    return 0;
}

struct DevArr {
    void* p = nullptr;
    ~DevArr() { if (p) (void)hipFree(p); }
};

}  // namespace

void graph_cut_fast(const double* V, size_t nv, const int32_t* dT, const int32_t* dN, size_t N,
                    const uint8_t* lp, double fill_T, int requested_threads, int period, uint64_t max_rounds,
                    int local_steps, uint8_t* out_labels, wtv_cut_result& R, wtv_stats& st, void* stream_v) {
    using Clock = std::chrono::steady_clock;
    const auto all0 = Clock::now();
    std::memset(&R, 0, sizeof(R));
    std::memset(&st, 0, sizeof(st));
    if (!V || (!dT && N) || (!dN && N) || (!lp && N) || (!out_labels && N))
        fail(WTV_EINVAL, "null pointer argument");
    if (N > static_cast<size_t>(std::numeric_limits<uint32_t>::max()))
        fail(WTV_EINVAL, ">2^32 full tets unsupported");
    const int threads = std::max(1, requested_threads);
    hipStream_t s = static_cast<hipStream_t>(stream_v);

    // Topology comes down in one transfer each (borrowed device buffers).
    std::vector<int32_t> ht(N * 4u), hn(N * 4u);
    if (N) {
        ckh(hipMemcpyAsync(hn.data(), dN, N * 4u * sizeof(int32_t), hipMemcpyDeviceToHost, s), "nbr D2H");
        ckh(hipMemcpyAsync(ht.data(), dT, N * 4u * sizeof(int32_t), hipMemcpyDeviceToHost, s), "tets D2H");
        ckh(hipStreamSynchronize(s), "topology D2H");
    }

    // ---- classify (wtivo_gpupr_bindings.cpp:153-171)
    std::vector<int32_t> full_to_free(N, -1);
    std::vector<size_t> lf(threads, 0), ls(threads, 0), lk(threads, 0), lr(threads, 0);
    std::atomic<int> bad{0};
    par_chunks(N, threads, [&](size_t a, size_t b, int tid) {
        size_t qf = 0, qs = 0, qk = 0, qr = 0;
        for (size_t i = a; i < b; ++i) {
            bool exterior = false;
            for (int fi = 0; fi < 4; ++fi) {
                const int32_t j = hn[i * 4u + static_cast<size_t>(fi)];
                if (j < 0) exterior = true;
                else {
                    ++qr;
                    if (static_cast<size_t>(j) >= N) bad.store(1);
                }
            }
            if (lp[i] == 0u) { full_to_free[i] = -2; ++qs; }
            else if (exterior) { full_to_free[i] = -3; ++qk; }
            else { full_to_free[i] = -1; ++qf; }
        }
        lf[tid] = qf; ls[tid] = qs; lk[tid] = qk; lr[tid] = qr;
    });
    if (bad.load()) fail(WTV_EINVAL, "neighbor out of range");
    size_t free_N = 0, fixed_source = 0, fixed_sink = 0, full_refs = 0;
    for (int t = 0; t < threads; ++t) { free_N += lf[t]; fixed_source += ls[t]; fixed_sink += lk[t]; full_refs += lr[t]; }
    if (full_refs & 1u) fail(WTV_EINVAL, "neighbor reference count not even");

    // ---- reduced ids in increasing tet order (:179-188), prefix over chunks
    std::vector<uint32_t> free_old(free_N);
    {
        const int T = (N < 32768) ? 1 : threads;
        const size_t step = (N + static_cast<size_t>(T) - 1) / static_cast<size_t>(T);
        std::vector<size_t> cnt(static_cast<size_t>(T) + 1, 0);
        par_chunks(static_cast<size_t>(T), T, [&](size_t a, size_t b, int) {
            for (size_t t = a; t < b; ++t) {
                size_t c = 0;
                const size_t i0 = t * step, i1 = std::min(N, i0 + step);
                for (size_t i = i0; i < i1; ++i) c += full_to_free[i] == -1;
                cnt[t + 1] = c;
            }
        });
        for (int t = 0; t < T; ++t) cnt[t + 1] += cnt[t];
        if (cnt[T] != free_N) fail(WTV_EINVAL, "reduced id mismatch");
        par_chunks(static_cast<size_t>(T), T, [&](size_t a, size_t b, int) {
            for (size_t t = a; t < b; ++t) {
                size_t rid = cnt[t];
                const size_t i0 = t * step, i1 = std::min(N, i0 + step);
                for (size_t i = i0; i < i1; ++i)
                    if (full_to_free[i] == -1) {
                        full_to_free[i] = static_cast<int32_t>(rid);
                        free_old[rid] = static_cast<uint32_t>(i);
                        ++rid;
                    }
            }
        });
    }
    R.full_edges = full_refs / 2u;
    R.fixed_source = fixed_source;
    R.fixed_sink = fixed_sink;
    R.threads = threads;
    R.local_steps = local_steps;

    if (free_N == 0) {
        if (N) std::memcpy(out_labels, lp, N);
        R.native_total_seconds = std::chrono::duration<double>(Clock::now() - all0).count();
        return;
    }
    if (free_N > static_cast<size_t>(std::numeric_limits<int32_t>::max()))
        fail(WTV_EINVAL, ">INT32_MAX reduced nodes unsupported");

    const uint64_t graph_bytes = static_cast<uint64_t>(free_N) * (4ull * 4u + 4ull * 4u + 4u);
    // this solver's workspace: excess 8 + height 4 + mark 4 + 3 lists 12 + rslot 1 per node
    const uint64_t workspace_est = static_cast<uint64_t>(free_N) * 29ull + 8192ull;
    size_t gpu_free = 0, gpu_total = 0;
    ckh(hipMemGetInfo(&gpu_free, &gpu_total), "hipMemGetInfo");
    const uint64_t safety = 256ull * 1024ull * 1024ull;
    if (graph_bytes + workspace_est + safety > static_cast<uint64_t>(gpu_free))
        fail(WTV_ENOMEM, "insufficient VRAM for reduced graph + solver");

    // ---- build (:253-294)
    const auto build0 = Clock::now();
    std::vector<int32_t> h_nbr(free_N * 4u, -1);
    std::vector<uint32_t> h_res(free_N * 4u, 0u);
    std::vector<int32_t> h_term(free_N, 0);
    std::vector<uint64_t> lrefs(threads, 0), lterm(threads, 0);
    std::vector<int64_t> lbase(threads, 0);
    std::atomic<int> bad2{0};
    par_chunks(N, threads, [&](size_t a, size_t b, int tid) {
        uint64_t refs = 0, terms = 0;
        int64_t base = 0;
        for (size_t i = a; i < b; ++i) {
            const int32_t rr = full_to_free[i];
            if (rr < 0) continue;
            const size_t local = static_cast<size_t>(rr);
            int64_t src = 0, snk = 0;
            const int32_t* tr = &ht[i * 4u];
            for (int fi = 0; fi < 4; ++fi) {
                const int32_t jj = hn[i * 4u + static_cast<size_t>(fi)];
                if (jj < 0) { bad2.store(1); continue; }
                const size_t j = static_cast<size_t>(jj);
                if (j >= N) { bad2.store(2); continue; }
                const int32_t m = full_to_free[j];
                int32_t cap = 0;
                const int rc = capacity_exact(V, nv, tr, fi, lp[i], lp[j], fill_T, &cap);
                if (rc == 1) { bad2.store(5); continue; }
                if (rc == 2) { bad2.store(6); continue; }
                if (m >= 0) {
                    h_nbr[local * 4u + static_cast<size_t>(fi)] = m;
                    h_res[local * 4u + static_cast<size_t>(fi)] = static_cast<uint32_t>(cap);
                    ++refs;
                } else if (m == -2) { src += cap; ++terms; }
                else if (m == -3) { snk += cap; ++terms; }
                else bad2.store(3);
            }
            base += std::min(src, snk);
            const int64_t diff = src - snk;
            if (diff < std::numeric_limits<int32_t>::min() || diff > std::numeric_limits<int32_t>::max())
                bad2.store(4);
            else h_term[local] = static_cast<int32_t>(diff);
        }
        lrefs[tid] = refs; lterm[tid] = terms; lbase[tid] = base;
    });
    // The reference's build lambda catches every exception, including the
    // capacity overflow_error, and reports one runtime_error
    // (wtivo_gpupr_bindings.cpp:292,296-297); the detail is appended here.
    const int b2 = bad2.load();
    if (b2 == 6) fail(WTV_EINVAL, "invalid graph reduction/capacity state (capacity outside safe 30-bit range)");
    if (b2 == 5) fail(WTV_EINVAL, "invalid graph reduction/capacity state (invalid tet vertex)");
    if (b2) fail(WTV_EINVAL, "invalid graph reduction/capacity state");
    uint64_t free_refs = 0, terminal_edges = 0;
    int64_t base_flow = 0;
    for (int t = 0; t < threads; ++t) { free_refs += lrefs[t]; terminal_edges += lterm[t]; base_flow += lbase[t]; }

    DevArr gN, gR, gT;
    ckh(hipMalloc(&gN.p, free_N * 4u * sizeof(int32_t)), "malloc graph nbr");
    ckh(hipMalloc(&gR.p, free_N * 4u * sizeof(uint32_t)), "malloc graph res");
    ckh(hipMalloc(&gT.p, free_N * sizeof(int32_t)), "malloc graph term");
    ckh(hipMemcpyAsync(gN.p, h_nbr.data(), free_N * 4u * sizeof(int32_t), hipMemcpyHostToDevice, s), "upload nbr");
    ckh(hipMemcpyAsync(gR.p, h_res.data(), free_N * 4u * sizeof(uint32_t), hipMemcpyHostToDevice, s), "upload res");
    ckh(hipMemcpyAsync(gT.p, h_term.data(), free_N * sizeof(int32_t), hipMemcpyHostToDevice, s), "upload term");
    ckh(hipStreamSynchronize(s), "graph upload sync");
    const auto build1 = Clock::now();
    if (free_refs & 1ull) fail(WTV_EINVAL, "free reference count not even");

    std::vector<int32_t>().swap(full_to_free);
    std::vector<int32_t>().swap(ht);
    std::vector<int32_t>().swap(hn);
    std::vector<int32_t>().swap(h_nbr);
    std::vector<uint32_t>().swap(h_res);
    std::vector<int32_t>().swap(h_term);

    wtv_params prm;
    wtv_params_default(&prm);
    prm.global_relabel_period = period;
    prm.max_rounds = max_rounds;
    prm.local_steps = local_steps;
    std::vector<uint8_t> part(free_N);
    pr_solve(static_cast<uint32_t>(free_N), static_cast<int32_t*>(gN.p), static_cast<uint32_t*>(gR.p),
             static_cast<int32_t*>(gT.p), part.data(), false, prm, st, stream_v);

    std::memcpy(out_labels, lp, N);
    par_chunks(free_N, threads, [&](size_t a, size_t b, int) {
        for (size_t q = a; q < b; ++q) out_labels[free_old[q]] = part[q] ? 1u : 0u;
    });

    R.adjusted_flow = st.raw_flow + base_flow;
    R.raw_flow = st.raw_flow;
    R.graph_bytes = graph_bytes;
    R.workspace_bytes = st.solver_workspace_bytes;
    R.free_edges = free_refs / 2u;
    R.terminal_edges = terminal_edges;
    R.macro_rounds = st.rounds;
    R.global_relabels = st.global_relabels;
    R.map_bytes = static_cast<uint64_t>(N) * sizeof(int32_t) + static_cast<uint64_t>(free_N) * sizeof(uint32_t);
    R.build_upload_seconds = std::chrono::duration<double>(build1 - build0).count();
    R.solve_seconds = st.solve_seconds;
    R.native_total_seconds = std::chrono::duration<double>(Clock::now() - all0).count();
}

}  // namespace wtv
