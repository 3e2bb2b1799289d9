// wtv_solver.hip -- push-relabel max-flow + residual BFS partition for gfx950.
//
// Replaces cellocut_gpu_push_relabel_fast (native/gpupr/gpu_push_relabel_fast.cu:406-571).
// Same arithmetic as the reference: uint32 residuals, int32 terminals
// (>0 source supply, <0 remaining sink capacity), int64 excess, uint64 sink
// flow; source height n, unreached height n+1; active = excess > 0 and
// height < n (gpu_push_relabel_fast.cu:149-166); relabel to 1 + min height over
// residual arcs, the sink counting as height 0 when term < 0 and the source as
// height n when term > 0 (gpu_push_relabel_fast.cu:274-302).
//
// Schedule (this file's own design, not the reference's): every sub-round is
// two kernels over device-resident worklists.
//   push    : heights are read-only.  Each active vertex pushes to the sink
//             (h == 1) and along every admissible arc (h(v) == h(u) + 1) as
//             long as it has excess.  Its own residual slots and terminal are
//             written with plain stores (no other thread writes them in this
//             phase: admissibility is antisymmetric); the receiver's reverse
//             residual and excess take device-scope atomic adds.
//   relabel : every vertex that was active or received flow recomputes its
//             height from its residual arcs and re-enters the worklist.
// Because heights never change during a push phase and only grow in the
// relabel phase, the height function stays a valid labelling
// (h(v) <= h(u) + 1 on every residual arc) with any number of threads, which
// is what makes the final BFS partition the unique minimal sink set.
//
// List appends are staged in LDS (ballot + mbcnt ranks, one LDS atomic per
// wave) and reach the global list with one device atomic per workgroup flush.
// Counters live in device memory and are zeroed by the kernel before the one
// that appends to them, so the host never waits per round: it reads the
// active count once per macro round (local_steps sub-rounds).
//
// Global relabel = level-synchronous BFS from the sink over residual arcs
// (gpu_push_relabel_fast.cu:85-147,360-402), with reverse slots precomputed
// once per solve into one byte per vertex (2 bits per slot) instead of the
// reference's per-arc 4-slot search (gpu_push_relabel_fast.cu:56-67).  Levels
// 1 and 2 come from one sweep over the terminals; a global relabel never
// lowers a height, so the active list after it is the previous list filtered
// by the new heights rather than a sweep over all vertices.  Vertices the BFS
// does not reach keep a sentinel height that every reader takes as n+1.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <optional>
#include <sstream>
#include <type_traits>
#include <vector>

#include "wtv_internal.h"
#include "wtv_wgqueue.h"

namespace wtv {
namespace {

constexpr uint32_t kSent = 0xFFFFFFFFu;
// bit 4 of a vertex's residual-mask byte during a BFS: not reached yet
constexpr uint32_t kUnreached = 16u;
// bit 5, once reached at level >= 2: the parity of that level
constexpr uint32_t kParityShift = 5u;
__device__ __forceinline__ uint8_t reached_byte(uint32_t mask_byte, uint32_t level) {
    return static_cast<uint8_t>((mask_byte & (kUnreached - 1u)) | ((level & 1u) << kParityShift));
}
constexpr int kTPB = 256;

// Device control block: worklist counters, flow, work counters, error word.
enum : int {
    kStVisits,
    kStPushes,
    kStRelabels,
    kStBfsPops,
    kStSinkFlow,
    kStBuChecks,
    kStSeedChecks,  // vertices with term >= 0 examined by k_bfs_seed
    kNumStats
};
constexpr int kStatSlots = 64;

// Phase of a list position: position % 6 fixes both the active-list buffer
// (phase & 1) and the count slot (phase % 3).
__host__ __device__ inline uint32_t lbuf(uint32_t ph) { return ph & 1u; }
__host__ __device__ inline uint32_t lslot(uint32_t ph, uint32_t ahead = 0u) { return (ph + ahead) % 3u; }

struct Ctl {
    // Active-list counts, a ring of three indexed by position % 3 (the
    // lists themselves ping-pong between two buffers by position % 2).  A
    // sub-round at position t reads lc[t % 3], appends to lc[(t + 1) % 3]
    // (zero when it starts) and zeroes lc[(t + 2) % 3] for the next one, so
    // a single kernel can read one count, append to the next and clear the
    // one after without a separate reset launch.
    uint32_t lc[3];
    uint32_t rl_cnt[2];   // relabel list counts (ping-pong)
    uint32_t bfs_cnt[3];  // BFS frontier counts, ring indexed by level % 3
    uint32_t bfs_levels;
    uint32_t tail_d;      // level reached by the one-workgroup BFS tail kernel
    uint32_t tail_sub;    // sub-rounds run by the one-workgroup discharge kernel
    unsigned long long bu_levels;
    int32_t error;
    uint32_t violations;
    // Work counters and the sink flow, each spread over kStatSlots addresses:
    // device-scope atomics on one address serialise at ~12 ns each on MI355X
    // (about 83M/s, returning or not), so one counter per wave of a large
    // launch would cost tens of microseconds.  Readers sum the slots.
    unsigned long long stat[kNumStats][kStatSlots];
};

inline unsigned long long stat_sum(const Ctl& c, int f) {
    unsigned long long t = 0;
    for (int i = 0; i < kStatSlots; ++i) t += c.stat[f][i];
    return t;
}

// ---------------------------------------------------------------- atomics
__device__ __forceinline__ uint32_t atom_xchg(uint32_t* p, uint32_t v) {
    return __hip_atomic_exchange(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ bool atom_cas(uint32_t* p, uint32_t expected, uint32_t desired) {
    return __hip_atomic_compare_exchange_strong(p, &expected, desired, __ATOMIC_RELAXED,
                                                __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ long long atom_load(const long long* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void set_error(Ctl* c, int code) {
    __hip_atomic_exchange(&c->error, code, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ------------------------------------------------------- wave64 primitives
__device__ __forceinline__ unsigned long long wave_sum(unsigned long long x) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
    return x;
}

// Sum per-lane counters over the wave and add them once per wave, to one of
// kStatSlots addresses picked by the wave's index in the grid.
__device__ __forceinline__ void flush_stat(Ctl* ctl, int f, unsigned long long x) {
    x = wave_sum(x);
    if (__lane_id() == 0 && x) {
        const uint32_t w = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
        atom_add(&ctl->stat[f][w % kStatSlots], x);
    }
}

__device__ __forceinline__ int slot_of(const int4& v, int k) {
    return k == 0 ? v.x : k == 1 ? v.y : k == 2 ? v.z : v.w;
}
__device__ __forceinline__ uint32_t slot_of(const uint4& v, int k) {
    return k == 0 ? v.x : k == 1 ? v.y : k == 2 ? v.z : v.w;
}

// Source height of v: n for a single graph; in batched mode the node count
// of v's own graph, so every graph keeps its own height conventions.
template <bool kB>
__device__ __forceinline__ uint32_t src_h(uint32_t v, uint32_t n, const uint32_t* __restrict__ gid,
                                          const uint32_t* __restrict__ gn) {
    if constexpr (kB) return gn[gid[v]];
    else return n;
}

// ---------------------------------------------------------------- kernels



// Batched mode: graph g occupies nodes [off[g], off[g+1]) of the union and
// its neighbour ids are local to g.  Writes union-wide ids and the graph id.
__global__ void __launch_bounds__(kTPB) k_globalize(uint32_t n, uint32_t G, const unsigned long long* __restrict__ off,
                                                    const int4* __restrict__ nbr_local, int4* __restrict__ nbr_global,
                                                    uint32_t* __restrict__ gid, Ctl* ctl) {
    const uint32_t stride = gridDim.x * kTPB;
    // largest g with off[g] <= x
    auto graph_of = [&](uint32_t x) {
        uint32_t lo = 0, hi = G;
        while (hi - lo > 1u) {
            const uint32_t mid = (lo + hi) >> 1;
            if (off[mid] <= x) lo = mid; else hi = mid;
        }
        return lo;
    };
    for (uint32_t v = blockIdx.x * kTPB + threadIdx.x; v < n; v += stride) {
        // one search per wave, for its first vertex; a lane past that graph's
        // end (the wave straddles graphs) searches for its own
        const uint32_t v0 = __shfl(v, 0);
        uint32_t lo = 0;
        if (__lane_id() == 0) lo = graph_of(v0);
        lo = __shfl(lo, 0);
        if (off[lo + 1] <= v) lo = graph_of(v);
        const uint32_t base = static_cast<uint32_t>(off[lo]);
        const uint32_t ng = static_cast<uint32_t>(off[lo + 1] - off[lo]);
        const int4 a = nbr_local[v];
        int4 b;
        int err = 0;
        auto map = [&](int x) -> int {
            if (x < 0) return -1;
            if (static_cast<uint32_t>(x) >= ng) { err = 3; return -1; }
            return static_cast<int>(base + static_cast<uint32_t>(x));
        };
        b.x = map(a.x); b.y = map(a.y); b.z = map(a.z); b.w = map(a.w);
        if (err) set_error(ctl, err);
        nbr_global[v] = b;
        gid[v] = lo;
    }
}

// Reverse slots and preflow init (gpu_push_relabel_fast.cu:69-83) of the
// nodes [lo, hi), whose neighbours lie in [lo, hi) too (a graph of the
// arena, or the whole solve).  With `wl` the vertices holding source supply
// are appended to it (counter *cnt); without, the pass only validates.
// Errors go to *err_word when given (a graph of the arena), else to ctl.
__global__ void __launch_bounds__(kTPB) k_prepare(uint32_t lo, uint32_t hi, const int4* __restrict__ nbr4,
                                                  const int32_t* __restrict__ term, uint8_t* __restrict__ rslot,
                                                  long long* __restrict__ excess, uint32_t* __restrict__ wl,
                                                  uint32_t* cnt, Ctl* ctl, int32_t* err_word) {
    __shared__ WgQueue<kTPB> q;
    q_init(q);
    // uniform trip count per workgroup: every thread reaches each queue flush
    const Span sp = wg_span(hi - lo, kTPB);
    for (uint64_t base = sp.b; base < sp.e; base += kTPB) {
        const uint32_t v = lo + static_cast<uint32_t>(base) + threadIdx.x;
        bool cand = false;
        if (base + threadIdx.x < sp.e) {
            const int4 nb = nbr4[v];
            uint32_t code = 0;
            int err = 0;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const int u = slot_of(nb, k);
                if (u < 0) continue;
                if (static_cast<uint32_t>(u) < lo || static_cast<uint32_t>(u) >= hi) { err = 3; continue; }
                const int4 nu = nbr4[u];
                const int iv = static_cast<int>(v);
                const int j = nu.x == iv ? 0 : nu.y == iv ? 1 : nu.z == iv ? 2 : nu.w == iv ? 3 : -1;
                if (j < 0) { err = 2; continue; }
                code |= static_cast<uint32_t>(j) << (2 * k);
            }
            if (err) {
                if (err_word) __hip_atomic_exchange(err_word, err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                else set_error(ctl, err);
            }
            rslot[v] = static_cast<uint8_t>(code);
            const int32_t t = term[v];
            excess[v] = t > 0 ? static_cast<long long>(t) : 0LL;
            cand = wl && t > 0;  // every vertex with excess joins the candidate list
        }
        if (wl) {
            q_push(q, cand, v);
            q_flush_if(q, wl, cnt, kTPB);
        }
    }
    if (wl) q_flush(q, wl, cnt);
}

// Batched solves: the graphs that still hold an active vertex.  Graphs share
// no arcs, so a graph without one stays finished; the global relabels before
// the final one skip it (its heights no longer matter, and the final BFS
// recomputes them all).
__global__ void __launch_bounds__(kTPB) k_clear_alive(uint32_t G, uint8_t* __restrict__ galive) {
    for (uint32_t g = blockIdx.x * kTPB + threadIdx.x; g < G; g += gridDim.x * kTPB) galive[g] = 0u;
}
__global__ void __launch_bounds__(kTPB) k_mark_alive(uint32_t ph, const uint32_t* __restrict__ wl,
                                                     const uint32_t* __restrict__ gid, uint8_t* __restrict__ galive,
                                                     const Ctl* ctl) {
    const uint32_t cnt = ctl->lc[lslot(ph)];
    for (uint32_t i = blockIdx.x * kTPB + threadIdx.x; i < cnt; i += gridDim.x * kTPB) galive[gid[wl[i]]] = 1u;
}

// BFS seeds: vertices with remaining sink capacity get height 1
// (gpu_push_relabel_fast.cu:85-101); every other vertex starts unreached.
// No seed list is built: level 1 always expands bottom-up.
__global__ void k_bfs_begin(Ctl* ctl, uint32_t ph) {
    ctl->bfs_cnt[0] = 0u;
    ctl->bfs_cnt[1] = 0u;
    ctl->bfs_cnt[2] = 0u;
    ctl->lc[lslot(ph, 1u)] = 0u;  // receives the filtered active list
    ctl->lc[lslot(ph, 2u)] = 0u;  // receives the first sub-round's list
    ctl->rl_cnt[0] = 0u;       // the next push phase starts an empty relabel list
    ctl->rl_cnt[1] = 0u;
    ctl->bfs_levels += 1u;     // level 1 is expanded by k_bfs_seed
}

// BFS start (gpu_push_relabel_fast.cu:90-139, levels 1 and 2 in one sweep).
// Level 1 is every vertex with residual sink capacity (term < 0).  Level 2 is
// found bottom-up in the same pass: a vertex with term >= 0 and a residual arc
// into a term < 0 neighbour.  Testing the neighbour's terminal instead of its
// height removes the dependency on the level-1 writes, so one sweep replaces a
// seed sweep plus a bottom-up level.  Each thread takes kSeedUnroll vertices
// per iteration (coalesced, one stride of kTPB apart), so the workgroup queue
// flushes once per kSeedUnroll * kTPB vertices.  Level 2 goes to `fout` with
// its count in bfs_cnt[2] (zeroed by k_bfs_begin).
//
// The sweep also writes nzm[v], bit k set when res[v][k] != 0, for every
// vertex with term >= 0 -- every vertex the levels can still claim.  The
// residuals do not change during a BFS, so the levels test an arc into the
// frontier with one byte from an n-byte array instead of a word from the
// 16n-byte residual rows: a random top-down probe then hits a line that
// holds 64 vertices' masks (and stays in cache far better) instead of 4
// vertices' residual rows, and a bottom-up level reads 1 byte per vertex
// where it read 16.  Bit 4 (kUnreached) marks a vertex the BFS has not
// reached: a top-down probe reads only the neighbour's byte, and touches its
// height only for the claim itself.  Bit 5 holds the parity of the level that
// reached the vertex (0 for level 2, which this sweep writes), which is all a
// bottom-up level needs to know of a reached neighbour.
constexpr int kSeedUnroll = 4;
static_assert(kSeedUnroll * kTPB <= WgQueue<kTPB>::kCap, "queue holds one append per vertex per iteration");
__global__ void __launch_bounds__(kTPB) k_bfs_seed(uint32_t n, const int32_t* __restrict__ term,
                                                   const int4* __restrict__ nbr4, const uint32_t* __restrict__ res,
                                                   uint32_t* __restrict__ height, uint8_t* __restrict__ nzm,
                                                   uint32_t* __restrict__ fout, uint32_t first,
                                                   const uint32_t* __restrict__ gid,
                                                   const uint8_t* __restrict__ galive,
                                                   const uint8_t* __restrict__ gfresh, Ctl* ctl) {
    __shared__ WgQueue<kTPB> q;
    q_init(q);
    const uint32_t span = kTPB * kSeedUnroll;
    const Span sp = wg_span(n, span);
    unsigned long long work = 0;
    for (uint64_t base = sp.b; base < sp.e; base += span) {
        if (galive) {
            // a tile whose graphs are all finished is skipped by the whole
            // workgroup (graph ids ascend with the node id)
            const uint64_t last = base + span < n ? base + span - 1u : static_cast<uint64_t>(n) - 1u;
            bool any = false;
            for (uint32_t g = gid[base]; g <= gid[last] && !any; ++g) any = galive[g] != 0u;
            if (!any) continue;
        }
        // arena solves: a tile lies in one graph; a graph admitted since the
        // last relabel has no heights yet, so all of them are written
        const uint32_t tfirst = gfresh ? static_cast<uint32_t>(gfresh[gid[base]]) : first;
        uint32_t v[kSeedUnroll];
        int32_t t[kSeedUnroll];
#pragma unroll
        for (int j = 0; j < kSeedUnroll; ++j) {
            const uint64_t x = base + static_cast<uint64_t>(j) * kTPB + threadIdx.x;
            v[j] = static_cast<uint32_t>(x);
            t[j] = x < n ? term[x] : -1;
        }
        bool found[kSeedUnroll];
#pragma unroll
        for (int j = 0; j < kSeedUnroll; ++j) {
            found[j] = false;
            if (static_cast<uint64_t>(base) + static_cast<uint64_t>(j) * kTPB + threadIdx.x < n && t[j] >= 0) {
                ++work;
                const int4 nb = nbr4[v[j]];
                const uint4 r = reinterpret_cast<const uint4*>(res)[v[j]];
                const uint32_t mask = (r.x != 0u) | (r.y != 0u) << 1 | (r.z != 0u) << 2 | (r.w != 0u) << 3;
                int32_t tu[4];
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const int u = slot_of(nb, k);
                    tu[k] = (u >= 0 && slot_of(r, k) != 0u) ? term[u] : 0;
                }
#pragma unroll
                for (int k = 0; k < 4; ++k) found[j] |= tu[k] < 0;
                nzm[v[j]] = static_cast<uint8_t>(mask | (found[j] ? 0u : kUnreached));
            }
        }
#pragma unroll
        for (int j = 0; j < kSeedUnroll; ++j) {
            // a vertex with residual sink capacity keeps height 1 once set:
            // valid labels give it h <= 1, and term < 0 never returns once
            // the sink arc is saturated, so later relabels skip those writes
            if (static_cast<uint64_t>(base) + static_cast<uint64_t>(j) * kTPB + threadIdx.x < n &&
                (t[j] >= 0 || tfirst)) {
                height[v[j]] = t[j] < 0 ? 1u : (found[j] ? 2u : kSent);
                if (t[j] < 0) nzm[v[j]] = 0u;  // height 1 for good: never unreached again
            }
            q_push(q, found[j], v[j]);
        }
        q_flush_if(q, fout, &ctl->bfs_cnt[2], kSeedUnroll * kTPB);
    }
    q_flush(q, fout, &ctl->bfs_cnt[2]);
    flush_stat(ctl, kStSeedChecks, work);
}

// Top-down claims of frontier vertex u (valid lanes only): each unreached
// neighbour v with residual v->u > 0 (bit rslot of nzm[v]) gets height d+1 by
// CAS.  The height of v and its mask are loaded together (both addresses are
// known once nbr[u] and rslot[u] are in) to keep the dependent-latency chain
// short.
__device__ __forceinline__ void td_claims(bool valid, uint32_t u, uint32_t d, const int4* __restrict__ nbr4,
                                          const uint8_t* __restrict__ rslot, uint8_t* nzm,
                                          uint32_t* height, int4& nb, bool claim[4]) {
    nb = make_int4(-1, -1, -1, -1);
    uint32_t rs = 0;
    if (valid) {
        nb = nbr4[u];
        rs = rslot[u];
    }
    uint32_t mv[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int v = slot_of(nb, k);
        mv[k] = v >= 0 ? static_cast<uint32_t>(nzm[v]) : 0u;
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        claim[k] = false;
        // unreached with a residual arc into u; a stale bit only costs a failed CAS
        if ((mv[k] & kUnreached) != 0u && ((mv[k] >> ((rs >> (2 * k)) & 3u)) & 1u) != 0u) {
            claim[k] = atom_cas(&height[slot_of(nb, k)], kSent, d + 1u);
            if (claim[k]) nzm[slot_of(nb, k)] = reached_byte(mv[k], d + 1u);
        }
    }
}

// One BFS level from height d to d+1 (gpu_push_relabel_fast.cu:103-139),
// d >= 2.  Top-down over the frontier list when it is small; bottom-up over
// all vertices when the frontier is large: each unreached v checks its own
// arcs v->u for a residual arc into level d, needing neither the reverse slot
// nor a CAS.
template <int TPB>
__global__ void __launch_bounds__(TPB) k_bfs_level(uint32_t n, uint32_t d, uint32_t bu_threshold,
                                                    const uint32_t* __restrict__ fin, uint32_t* __restrict__ fout,
                                                    const int4* __restrict__ nbr4, const uint8_t* __restrict__ rslot,
                                                    uint8_t* nzm, uint32_t* height, Ctl* ctl) {
    __shared__ WgQueue<TPB> q;
    // The first frontier entry of this thread is loaded before the count is
    // known (the buffer holds n entries; entries past the count are ignored),
    // so the two loads overlap instead of forming two round trips.
    const uint32_t i0 = blockIdx.x * TPB + threadIdx.x;
    const uint32_t u0 = i0 < n ? fin[i0] : 0u;
    q_init(q);
    const uint32_t cnt = ctl->bfs_cnt[d % 3u];
    uint32_t* out_cnt = &ctl->bfs_cnt[(d + 1u) % 3u];
    const bool bottom_up = cnt > bu_threshold;  // bu_threshold >= 4096, so cnt > 0
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        ctl->bfs_cnt[(d + 2u) % 3u] = 0u;
        if (cnt) ctl->bfs_levels += 1u;
        if (bottom_up) ctl->bu_levels += 1u;
        // every frontier entry is expanded once: one add per level
        else if (cnt) atom_add(&ctl->stat[kStBfsPops][d % kStatSlots], static_cast<unsigned long long>(cnt));
    }
    unsigned long long work = 0;
    if (bottom_up) {
        const Span sp = wg_span(n, TPB);
        for (uint64_t base = sp.b; base < sp.e; base += TPB) {
            const uint32_t v = static_cast<uint32_t>(base) + threadIdx.x;
            bool found = false;
            // A reached vertex u with a residual arc v -> u from an unreached
            // v sits at level d or d + 1: at a level below d, v would have
            // been reached at the level after it (level 1, the sink seeds,
            // makes its neighbours level 2 in the seed sweep).  So the parity
            // bit of u's byte tells level d from a vertex reached in this
            // same pass, and the level reads bytes (64 per line) where it
            // read heights (16 per line).  Bytes of free arena ranges, which
            // the seed sweep does not write, have no arcs and find nothing.
            const uint32_t m = v < n ? static_cast<uint32_t>(nzm[v]) : 0u;
            if ((m & kUnreached) != 0u) {
                ++work;
                const int4 nb = nbr4[v];
                uint32_t mu[4];
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const int u = slot_of(nb, k);
                    mu[k] = (u >= 0 && ((m >> k) & 1u)) ? static_cast<uint32_t>(nzm[u]) : kUnreached;
                }
#pragma unroll
                for (int k = 0; k < 4; ++k)
                    found |= (mu[k] & kUnreached) == 0u && ((mu[k] >> kParityShift) & 1u) == (d & 1u);
                if (found) {
                    height[v] = d + 1u;
                    nzm[v] = reached_byte(m, d + 1u);
                }
            }
            q_push(q, found, v);
            q_flush_if(q, fout, out_cnt, TPB);
        }
        q_flush(q, fout, out_cnt);
        flush_stat(ctl, kStBuChecks, work);
    } else {
        // the prefetched entry is this thread's first when the share is the
        // single step at blockIdx.x * TPB (and for workgroup 0 always)
        const Span sp = wg_span(cnt, TPB);
        for (uint64_t base = sp.b; base < sp.e; base += TPB) {
            const uint32_t i = static_cast<uint32_t>(base) + threadIdx.x;
            const bool valid = i < sp.e;
            const uint32_t u = !valid ? 0u : (base == blockIdx.x * TPB ? u0 : fin[i]);
            int4 nb;
            bool claim[4];
            td_claims(valid, u, d, nbr4, rslot, nzm, height, nb, claim);
#pragma unroll
            for (int k = 0; k < 4; ++k) q_push(q, claim[k], static_cast<uint32_t>(slot_of(nb, k)));
            q_flush_if(q, fout, out_cnt, 4 * TPB);
        }
        q_flush(q, fout, out_cnt);
    }
}

// BFS tail: one workgroup expands level after level top-down, with
// __syncthreads between levels instead of a launch per level, while the
// frontier holds at most kTailMax vertices (up to max_levels levels).  Leaves
// the frontier of the level it stops at in f[d & 1] with its count in the
// counter ring, exactly as a sequence of k_bfs_level launches would.
constexpr uint32_t kTailMax = 1024;
constexpr double kRelabelCost = 1.0;
// With fused sub-rounds off, active lists of more than this many vertices
// run pull-mode sub-rounds by default.  On the 77.5M-node Delaunay graph a
// pull-mode sub-round costs 4.6 ms against 7.9 ms for a push/relabel pair
// while the list holds all ~37M source vertices, and the two meet near 2-3M
// vertices: the apply sweep reads two words per neighbour where the relabel
// reads one.  A fused sub-round (4.5 ms there, and less after it as the list
// shrinks faster) beats both, so with fused sub-rounds pull mode runs only on
// request (wtv_params.pull_min > 0).
constexpr uint64_t kPullMinList = 4194304u;
// The one-workgroup kernels run 1024 threads: a list of up to 1024 items
// takes one pass per level or phase.
constexpr int kTailTPB = 1024;
__global__ void __launch_bounds__(kTailTPB) k_bfs_tail(uint32_t d0, uint32_t max_levels, uint32_t* f0, uint32_t* f1,
                                                   const int4* __restrict__ nbr4, const uint8_t* __restrict__ rslot,
                                                   uint8_t* nzm, uint32_t* height, Ctl* ctl) {
    __shared__ uint32_t s_cnt;
    uint32_t* f[2] = {f0, f1};
    uint32_t cur = ctl->bfs_cnt[d0 % 3u];
    uint32_t d = d0, levels = 0;
    unsigned long long pops = 0;
    while (cur > 0u && cur <= kTailMax && levels < max_levels) {
        const uint32_t* fin = f[d & 1u];
        uint32_t* fout = f[(d + 1u) & 1u];
        if (threadIdx.x == 0) s_cnt = 0u;
        __syncthreads();
        for (uint32_t base = 0; base < cur; base += kTailTPB) {
            const uint32_t i = base + threadIdx.x;
            const bool valid = i < cur;
            const uint32_t u = valid ? fin[i] : 0u;
            pops += valid;
            int4 nb;
            bool claim[4];
            td_claims(valid, u, d, nbr4, rslot, nzm, height, nb, claim);
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const uint64_t m = __ballot(claim[k]);
                if (m == 0) continue;
                const uint32_t rank = lane_rank(m);
                const int leader = __ffsll(static_cast<unsigned long long>(m)) - 1;
                uint32_t off = 0;
                if (claim[k] && rank == 0) off = atomicAdd(&s_cnt, static_cast<uint32_t>(__popcll(m)));
                off = __shfl(off, leader);
                if (claim[k]) fout[off + rank] = static_cast<uint32_t>(slot_of(nb, k));
            }
        }
        __syncthreads();
        cur = s_cnt;
        ++d;
        ++levels;
        __syncthreads();
    }
    flush_stat(ctl, kStBfsPops, pops);
    if (threadIdx.x == 0) {
        ctl->bfs_cnt[d % 3u] = cur;
        ctl->bfs_cnt[(d + 1u) % 3u] = 0u;
        ctl->tail_d = d;
        ctl->bfs_levels += levels;
    }
}

// Active list after a global relabel (gpu_push_relabel_fast.cu:149-166:
// {excess > 0, height < n}).  A global relabel only raises heights (BFS
// distances bound every valid labelling from above), so no vertex becomes
// active: the new list is the old one -- which holds every active vertex --
// filtered by the new heights.  Entries are distinct, so survivors are too.
// Unreached vertices keep the sentinel height, which every reader treats as
// the reference's n+1 (gpu_push_relabel_fast.cu:141-147).
template <bool kB, int TPB>
__global__ void __launch_bounds__(TPB) k_filter_active(uint32_t n, uint32_t ph, const uint32_t* __restrict__ src,
                                                        uint32_t* __restrict__ dst,
                                                        const long long* __restrict__ excess,
                                                        const uint32_t* __restrict__ height,
                                                        const uint32_t* __restrict__ gid,
                                                        const uint32_t* __restrict__ gn, Ctl* ctl) {
    __shared__ WgQueue<TPB> q;
    q_init(q);
    const uint32_t cnt = ctl->lc[lslot(ph)];
    uint32_t* out_cnt = &ctl->lc[lslot(ph, 1u)];
    const Span sp = wg_span(cnt, TPB);
    for (uint64_t base = sp.b; base < sp.e; base += TPB) {
        const uint32_t i = static_cast<uint32_t>(base) + threadIdx.x;
        bool a = false;
        uint32_t v = 0;
        if (i < sp.e) {
            v = src[i];
            a = height[v] < src_h<kB>(v, n, gid, gn) && excess[v] > 0;
        }
        q_push(q, a, v);
        q_flush_if(q, dst, out_cnt, TPB);
    }
    q_flush(q, dst, out_cnt);
}

// Batched mode: adds each lane's sink flow to its graph's total.  The lanes
// of a wave nearly always belong to one graph, so the wave adds one sum per
// distinct graph id instead of one atomic per sink push (atomics on one
// address serialise).  All lanes of the wave must call this.
template <bool kB>
__device__ __forceinline__ void flush_graph_flow(bool valid, uint32_t v, unsigned long long f,
                                                 const uint32_t* __restrict__ gid,
                                                 unsigned long long* __restrict__ gflow) {
    if constexpr (kB) {
        const uint32_t g = valid && f ? gid[v] : 0u;
        uint64_t pending = __ballot(f != 0ull);
        while (pending) {
            const int leader = __ffsll(static_cast<unsigned long long>(pending)) - 1;
            const uint32_t gl = __shfl(g, leader);
            const bool mine = f != 0ull && g == gl;
            const unsigned long long sum = wave_sum(mine ? f : 0ull);
            if (__lane_id() == static_cast<unsigned>(leader)) atom_add(&gflow[gl], sum);
            if (mine) f = 0ull;
            pending = __ballot(f != 0ull);
        }
    }
}

// Push step of one worklist vertex v (heights read-only during the phase).
// keep: v still holds excess after its pushes; cand[k]: neighbour k received
// its first excess.  Both come from the excess atomics' returned values, so
// each vertex with excess at the end of the phase enters the relabel list
// exactly once, with no membership flag:
//  - excess never goes negative and only v's own thread lowers excess[v];
//  - the add that finds a receiver at exactly 0 is unique (adds are positive),
//    and that receiver is then not in the worklist or has just pushed all of
//    its excess out, in which case its own subtraction returned exactly what
//    it pushed and it did not keep itself;
//  - a receiver whose excess was positive already is kept by its own visit
//    (it is in the worklist, as is every vertex with excess and a height
//    below the source height), or lies at or above the source height and
//    stays inactive.
template <bool kB>
__device__ __forceinline__ void push_vertex(bool valid, uint32_t v, uint32_t n, const int4* __restrict__ nbr4,
                                            const uint8_t* __restrict__ rslot, uint32_t* res,
                                            int32_t* __restrict__ term, long long* excess,
                                            const uint32_t* __restrict__ height,
                                            const uint32_t* __restrict__ gid, const uint32_t* __restrict__ gn,
                                            bool& keep, bool cand[4], int4& nb, unsigned long long& visits,
                                            unsigned long long& pushes, unsigned long long& sinkf,
                                            unsigned long long& sdlt) {
    sdlt = 0ull;
    // every load that depends on v alone is issued at once (one round trip),
    // before the activity test that needs two of them
    long long e = 0;
    uint32_t h = kSent, sh = 0u;
    nb = make_int4(-1, -1, -1, -1);
    uint4 r = make_uint4(0u, 0u, 0u, 0u);
    uint32_t rs = 0;
    int32_t t = 0;
    if (valid) {
        e = atom_load(&excess[v]);
        h = height[v];
        nb = nbr4[v];
        r = reinterpret_cast<const uint4*>(res)[v];
        rs = rslot[v];
        t = term[v];
        sh = src_h<kB>(v, n, gid, gn);
    }
    const bool act = (e > 0) & (h < sh);
    uint32_t hu[4] = {kSent, kSent, kSent, kSent};
    long long pushed = 0;
    if (act) {
        ++visits;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int u = slot_of(nb, k);
            if (u >= 0) hu[k] = height[u];
        }
        // v -> sink (gpu_push_relabel_fast.cu:220-236)
        if (t < 0 && h == 1u) {
            const long long cap = -static_cast<long long>(t);
            const long long dlt = e < cap ? e : cap;
            term[v] = static_cast<int32_t>(static_cast<long long>(t) + dlt);
            e -= dlt;
            pushed += dlt;
            sinkf += static_cast<unsigned long long>(dlt);
            sdlt = static_cast<unsigned long long>(dlt);
            ++pushes;
        }
    } else {
        nb = make_int4(-1, -1, -1, -1);
    }
    // the amounts first (registers only), then the stores and every atomic
    // back to back: one memory round trip for all of them instead of one per
    // pushed arc.  v's slots are stored one by one: a neighbour pushing into
    // v adds to one of v's other slots at the same time (never to a slot v
    // pushes along: that needs the opposite height order).
    uint32_t dl[4] = {0u, 0u, 0u, 0u};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int u = slot_of(nb, k);
        const uint32_t rk = slot_of(r, k);
        if (e > 0 && u >= 0 && rk != 0u && h == hu[k] + 1u) {
            dl[k] = static_cast<long long>(rk) < e ? rk : static_cast<uint32_t>(e);
            e -= dl[k];
            pushed += dl[k];
            ++pushes;
        }
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (dl[k]) {
            res[static_cast<size_t>(v) * 4u + k] = slot_of(r, k) - dl[k];
            const uint32_t back = (rs >> (2 * k)) & 3u;
            atom_add(&res[static_cast<size_t>(slot_of(nb, k)) * 4u + back], dl[k]);
        }
    }
    // old[k] is read only where dl[k] != 0, where the atomic set it; no
    // default value, so no merge at the branch joins forces a wait there
    long long old[4];
#pragma unroll
    for (int k = 0; k < 4; ++k)
        if (dl[k]) old[k] = atom_add(&excess[slot_of(nb, k)], static_cast<long long>(dl[k]));
    long long own = 0ll;
    if (pushed) own = atom_add(&excess[v], -pushed);
#pragma unroll
    for (int k = 0; k < 4; ++k) cand[k] = dl[k] != 0u && old[k] == 0ll;
    keep = act && (pushed == 0 || own != pushed);
}

// Relabel step of one relabel-list vertex: height = 1 + min over residual
// arcs (sink 0 when term < 0, source = own source height when term > 0);
// returns whether v stays active.
template <bool kB>
__device__ __forceinline__ bool relabel_vertex(uint32_t v, uint32_t n, const int4* __restrict__ nbr4,
                                               const uint32_t* __restrict__ res, const int32_t* __restrict__ term,
                                               const long long* __restrict__ excess, uint32_t* height,
                                               const uint32_t* __restrict__ gid,
                                               const uint32_t* __restrict__ gn, Ctl* ctl,
                                               unsigned long long& relabels) {
    // every load that depends on v alone at once (one round trip), then the
    // residual neighbours' heights at once (a second one)
    long long e = 0;
    uint32_t h = 0u, sh = 0u;
    int32_t t = 0;
    int4 nb = make_int4(-1, -1, -1, -1);
    uint4 r = make_uint4(0u, 0u, 0u, 0u);
    if (v < n + 1u) {  // list entries are vertex ids: always
        e = excess[v];
        h = height[v];
        sh = src_h<kB>(v, n, gid, gn);
        t = term[v];
        nb = nbr4[v];
        r = reinterpret_cast<const uint4*>(res)[v];
    }
    if (!((e > 0) & (h < sh))) return false;
    uint32_t x[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int u = slot_of(nb, k);
        x[k] = u >= 0 && slot_of(r, k) != 0u ? height[u] : kSent;
    }
    uint32_t mh = kSent;
    if (t < 0) mh = 0u;
    else if (t > 0) mh = sh;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int u = slot_of(nb, k);
        if (u >= 0 && slot_of(r, k) != 0u) {
            // a vertex the last BFS did not reach keeps the sentinel; it
            // stands for the reference's normalised height n+1
            // (gpu_push_relabel_fast.cu:141-147)
            const uint32_t y = x[k] == kSent ? sh + 1u : x[k];
            mh = y < mh ? y : mh;
        }
    }
    if (mh == kSent) {
        set_error(ctl, 5);
    } else if (mh + 1u > h) {
        h = mh + 1u;
        height[v] = h;
        ++relabels;
    }
    return h < sh;
}

// Push phase of the sub-round with phase ph (heights read-only).
template <bool kB, int TPB>
__global__ void __launch_bounds__(TPB) k_push(uint32_t n, uint32_t ph, const uint32_t* __restrict__ wl,
                                               uint32_t* __restrict__ rl, const int4* __restrict__ nbr4,
                                               const uint8_t* __restrict__ rslot, uint32_t* res,
                                               int32_t* __restrict__ term, long long* excess,
                                               const uint32_t* __restrict__ height,
                                               const uint32_t* __restrict__ gid, const uint32_t* __restrict__ gn,
                                               unsigned long long* __restrict__ gflow, Ctl* ctl) {
    __shared__ WgQueue<TPB> q;
    q_init(q);
    const uint32_t cnt = ctl->lc[lslot(ph)];
    if (blockIdx.x == 0 && threadIdx.x == 0) ctl->lc[lslot(ph, 2u)] = 0u;
    uint32_t* rl_cnt = &ctl->rl_cnt[lbuf(ph)];
    unsigned long long visits = 0, pushes = 0, sinkf = 0;
    const Span sp = wg_span(cnt, TPB);
    for (uint64_t base = sp.b; base < sp.e; base += TPB) {
        const uint32_t i = static_cast<uint32_t>(base) + threadIdx.x;
        const bool valid = i < sp.e;
        const uint32_t v = valid ? wl[i] : 0u;
        bool keep, cand[4];
        int4 nb;
        unsigned long long sdlt;
        push_vertex<kB>(valid, v, n, nbr4, rslot, res, term, excess, height, gid, gn, keep, cand, nb, visits,
                        pushes, sinkf, sdlt);
        flush_graph_flow<kB>(valid, v, sdlt, gid, gflow);
        // every vertex left with excess is relabel-checked exactly once
        q_push(q, keep, v);
#pragma unroll
        for (int k = 0; k < 4; ++k) q_push(q, cand[k], static_cast<uint32_t>(slot_of(nb, k)));
        q_flush_if(q, rl, rl_cnt, 5 * TPB);
    }
    q_flush(q, rl, rl_cnt);
    flush_stat(ctl, kStVisits, visits);
    flush_stat(ctl, kStPushes, pushes);
    flush_stat(ctl, kStSinkFlow, sinkf);
}

// Relabel phase of the sub-round with phase ph: recompute heights of every
// vertex that was active or received flow, and rebuild the active list.
template <bool kB, int TPB>
__global__ void __launch_bounds__(TPB) k_relabel(uint32_t n, uint32_t ph, const uint32_t* __restrict__ rl,
                                                  uint32_t* __restrict__ wl_next, const int4* __restrict__ nbr4,
                                                  const uint32_t* __restrict__ res, const int32_t* __restrict__ term,
                                                  const long long* __restrict__ excess, uint32_t* height,
                                                  const uint32_t* __restrict__ gid,
                                                  const uint32_t* __restrict__ gn, Ctl* ctl) {
    __shared__ WgQueue<TPB> q;
    q_init(q);
    const uint32_t cnt = ctl->rl_cnt[lbuf(ph)];
    if (blockIdx.x == 0 && threadIdx.x == 0) ctl->rl_cnt[lbuf(ph) ^ 1u] = 0u;
    uint32_t* wl_cnt = &ctl->lc[lslot(ph, 1u)];
    unsigned long long relabels = 0;
    const Span sp = wg_span(cnt, TPB);
    for (uint64_t base = sp.b; base < sp.e; base += TPB) {
        const uint32_t i = static_cast<uint32_t>(base) + threadIdx.x;
        bool act = false;
        uint32_t v = 0;
        if (i < sp.e) {
            v = rl[i];
            act = relabel_vertex<kB>(v, n, nbr4, res, term, excess, height, gid, gn, ctl, relabels);
        }
        q_push(q, act, v);
        q_flush_if(q, wl_next, wl_cnt, TPB);
    }
    q_flush(q, wl_next, wl_cnt);
    flush_stat(ctl, kStRelabels, relabels);
}

// ------------------------------------------------ fused push-relabel sub-rounds
// One kernel per sub-round: a visit pushes along admissible arcs, relabels
// itself when it keeps excess -- by at most 2 levels -- and pushes again
// along the arcs its new height made admissible.  Heights stay a valid
// labelling with every visit running concurrently: a neighbour u that
// pushes into v (creating v -> u) read h(u) = h_old(v) + 1 and so
// h_new(v) <= h_old(v) + 2 = h(u) + 1; a neighbour u that v pushes into after
// raising (creating u -> v) was read at h_new(v) - 1 and rises by at most 2
// after that read, to at most h_new(v) + 1; a vertex visits once per
// sub-round, so it rises once.  Appends go straight to the next active list
// (keep and first-excess receivers, as k_push).

// One fused visit of v (all lanes call it; `valid` false for padding lanes).
// keep: v stays active and joins the next list; cand[k]: the receiver on
// slot k had no excess before this visit's push (it joins the next list).
template <bool kB>
__device__ __forceinline__ void fused_vertex(bool valid, uint32_t v, uint32_t n, const int4* __restrict__ nbr4,
                                             const uint8_t* __restrict__ rslot, uint32_t* res, int32_t* term,
                                             long long* excess, uint32_t* height, const uint32_t* __restrict__ gid,
                                             const uint32_t* __restrict__ gn, bool& keep, bool cand[4],
                                             int4& nb, unsigned long long& visits, unsigned long long& pushes,
                                             unsigned long long& sinkf, unsigned long long& relabels,
                                             unsigned long long& sdlt) {
    long long e = 0;
    uint32_t h = kSent, sh = 0u, rs = 0u;
    nb = make_int4(-1, -1, -1, -1);
    uint4 r = make_uint4(0u, 0u, 0u, 0u);
    int32_t t = 0;
    if (valid) {
        e = atom_load(&excess[v]);
        h = height[v];
        nb = nbr4[v];
        r = reinterpret_cast<const uint4*>(res)[v];
        rs = rslot[v];
        t = term[v];
        sh = src_h<kB>(v, n, gid, gn);
    }
    const bool act = (e > 0) & (h < sh);
    uint32_t hu[4] = {kSent, kSent, kSent, kSent};
    long long pushed = 0;
    sdlt = 0ull;
    if (act) {
        ++visits;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int u = slot_of(nb, k);
            if (u >= 0) hu[k] = height[u];
        }
        if (t < 0 && h == 1u) {
            const long long cap = -static_cast<long long>(t);
            const long long dlt = e < cap ? e : cap;
            t = static_cast<int32_t>(static_cast<long long>(t) + dlt);
            term[v] = t;
            e -= dlt;
            pushed += dlt;
            sinkf += static_cast<unsigned long long>(dlt);
            sdlt = static_cast<unsigned long long>(dlt);
            ++pushes;
        }
    }
    uint32_t dl[4] = {0u, 0u, 0u, 0u};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int u = slot_of(nb, k);
        const uint32_t rk = slot_of(r, k);
        if (act && e > 0 && u >= 0 && rk != 0u && h == hu[k] + 1u) {
            dl[k] = static_cast<long long>(rk) < e ? rk : static_cast<uint32_t>(e);
            e -= dl[k];
            pushed += dl[k];
            ++pushes;
        }
    }
    uint32_t hf = h;
    if (act && e > 0) {
        // relabel on what v read: 1 + the lowest residual neighbour (the
        // sink when term < 0, the source height when term > 0).  A vertex
        // that sees no residual arc keeps its height and stays listed: a
        // concurrent push into v adds the reverse residual and the excess
        // with two atomics, and v's loads may see the excess add before the
        // residual one; the next visit sees both.
        uint32_t mh = kSent;
        if (t < 0) mh = 0u;
        else if (t > 0) mh = sh;
#pragma unroll
        for (int k = 0; k < 4; ++k)
            if (slot_of(nb, k) >= 0 && slot_of(r, k) - dl[k] != 0u) {
                const uint32_t y = hu[k] == kSent ? sh + 1u : hu[k];
                mh = y < mh ? y : mh;
            }
        if (mh != kSent) {
            hf = mh + 1u < h + 2u ? mh + 1u : h + 2u;
            if (hf > h) {
                height[v] = hf;
                ++relabels;
            }
            // the arcs the new height made admissible
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const int u = slot_of(nb, k);
                const uint32_t rk = slot_of(r, k) - dl[k];
                if (e > 0 && u >= 0 && rk != 0u && hf == hu[k] + 1u) {
                    const uint32_t d2 = static_cast<long long>(rk) < e ? rk : static_cast<uint32_t>(e);
                    dl[k] += d2;
                    e -= d2;
                    pushed += d2;
                    ++pushes;
                }
            }
        }
    }
    // a neighbour may push into the very slot v pushes along in this
    // sub-round (one of the two raised), so v's own slots take atomic
    // subtractions too; v's loaded residual is never above the slot's
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (dl[k]) {
            atom_add(&res[static_cast<size_t>(v) * 4u + k], 0u - dl[k]);
            const uint32_t back = (rs >> (2 * k)) & 3u;
            atom_add(&res[static_cast<size_t>(slot_of(nb, k)) * 4u + back], dl[k]);
        }
    }
    long long old[4];
#pragma unroll
    for (int k = 0; k < 4; ++k)
        if (dl[k]) old[k] = atom_add(&excess[slot_of(nb, k)], static_cast<long long>(dl[k]));
    long long own = 0ll;
    if (pushed) own = atom_add(&excess[v], -pushed);
#pragma unroll
    for (int k = 0; k < 4; ++k) cand[k] = dl[k] != 0u && old[k] == 0ll;
    keep = act && (pushed == 0 || own != pushed) && hf < sh;
}

template <bool kB, int TPB>
__global__ void __launch_bounds__(TPB) k_push_fused(uint32_t n, uint32_t ph, const uint32_t* __restrict__ wl,
                                                     uint32_t* __restrict__ wl_next, const int4* __restrict__ nbr4,
                                                     const uint8_t* __restrict__ rslot, uint32_t* res, int32_t* term,
                                                     long long* excess, uint32_t* height,
                                                     const uint32_t* __restrict__ gid, const uint32_t* __restrict__ gn,
                                                     unsigned long long* __restrict__ gflow, Ctl* ctl) {
    __shared__ WgQueue<TPB> q;
    q_init(q);
    const uint32_t cnt = ctl->lc[lslot(ph)];
    uint32_t* out_cnt = &ctl->lc[lslot(ph, 1u)];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        ctl->lc[lslot(ph, 2u)] = 0u;
        ctl->rl_cnt[lbuf(ph) ^ 1u] = 0u;  // a two-phase sub-round may follow
    }
    unsigned long long visits = 0, pushes = 0, sinkf = 0, relabels = 0;
    const Span sp = wg_span(cnt, TPB);
    for (uint64_t base = sp.b; base < sp.e; base += TPB) {
        const uint32_t i = static_cast<uint32_t>(base) + threadIdx.x;
        const bool valid = i < sp.e;
        const uint32_t v = valid ? wl[i] : 0u;
        bool keep, cand[4];
        int4 nb;
        unsigned long long sdlt;
        fused_vertex<kB>(valid, v, n, nbr4, rslot, res, term, excess, height, gid, gn, keep, cand, nb, visits,
                         pushes, sinkf, relabels, sdlt);
        flush_graph_flow<kB>(valid, v, sdlt, gid, gflow);
        q_push(q, keep, v);
#pragma unroll
        for (int k = 0; k < 4; ++k) q_push(q, cand[k], static_cast<uint32_t>(slot_of(nb, k)));
        q_flush_if(q, wl_next, out_cnt, 5 * TPB);
    }
    q_flush(q, wl_next, out_cnt);
    flush_stat(ctl, kStVisits, visits);
    flush_stat(ctl, kStPushes, pushes);
    flush_stat(ctl, kStSinkFlow, sinkf);
    flush_stat(ctl, kStRelabels, relabels);
}

// ------------------------------------------------------ pull-mode sub-rounds
// Large active lists run sub-rounds without atomics.  In the push phase a
// vertex writes what it pushes along each slot into its own row of `out` and
// marks itself and its receivers in the byte map `touched`; it updates only
// its own residuals, terminal and excess, with plain stores.  The apply
// phase then sweeps the byte map in id order: each marked vertex reads what
// each neighbour pushed into it -- the neighbour's `out` entry for the slot
// that points back to it -- adds that to its excess and reverse residuals,
// clears the entry, relabels itself, and joins the next active list when
// still active.  Every entry has exactly one reader (the vertex on that
// slot), which is marked whenever the entry is non-zero, so every entry is
// zero again when the phase ends; entries of neighbours that did not push
// read as zero.  A sub-round computes what a k_push/k_relabel pair computes:
// heights are read-only while pushing, and the relabel sees every residual
// change of the sub-round.  The atomic sub-round pays three memory-side
// atomics per pushing vertex (receiver excess, reverse residual, own
// excess), each a read-modify-write of a 64-byte line at the memory, and
// those bound its launches while the list holds millions of vertices; here
// every access is a plain load or store, and the map sweep costs n bytes.
constexpr int kPullTPB = 256;
constexpr uint32_t kPullStep = 16u * kPullTPB;  // map bytes per sweep step (16 per lane)

template <bool kB>
__global__ void __launch_bounds__(kPullTPB) k_push_pull(uint32_t n, uint32_t ph, const uint32_t* __restrict__ wl,
                                                        const int4* __restrict__ nbr4, uint32_t* __restrict__ res,
                                                        int32_t* __restrict__ term, long long* __restrict__ excess,
                                                        const uint32_t* __restrict__ height,
                                                        uint32_t* __restrict__ out, uint8_t* __restrict__ touched,
                                                        const uint32_t* __restrict__ gid,
                                                        const uint32_t* __restrict__ gn,
                                                        unsigned long long* __restrict__ gflow, Ctl* ctl) {
    const uint32_t cnt = ctl->lc[lslot(ph)];
    if (blockIdx.x == 0 && threadIdx.x == 0) ctl->lc[lslot(ph, 2u)] = 0u;
    unsigned long long visits = 0, pushes = 0, sinkf = 0;
    const Span sp = wg_span(cnt, kPullTPB);
    for (uint64_t base = sp.b; base < sp.e; base += kPullTPB) {
        const uint32_t i = static_cast<uint32_t>(base) + threadIdx.x;
        const bool valid = i < sp.e;
        const uint32_t v = valid ? wl[i] : 0u;
        // nothing else writes v's row, terminal or excess in this phase
        long long e = 0;
        uint32_t h = kSent, sh = 0u;
        int4 nb = make_int4(-1, -1, -1, -1);
        uint4 r = make_uint4(0u, 0u, 0u, 0u);
        int32_t t = 0;
        if (valid) {
            e = excess[v];
            h = height[v];
            nb = nbr4[v];
            r = reinterpret_cast<const uint4*>(res)[v];
            t = term[v];
            sh = src_h<kB>(v, n, gid, gn);
        }
        const bool act = (e > 0) & (h < sh);
        uint32_t hu[4] = {kSent, kSent, kSent, kSent};
        unsigned long long sdlt = 0ull;
        long long pushed = 0;
        if (act) {
            ++visits;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const int u = slot_of(nb, k);
                if (u >= 0) hu[k] = height[u];
            }
            // v -> sink (gpu_push_relabel_fast.cu:220-236)
            if (t < 0 && h == 1u) {
                const long long cap = -static_cast<long long>(t);
                const long long dlt = e < cap ? e : cap;
                term[v] = static_cast<int32_t>(static_cast<long long>(t) + dlt);
                e -= dlt;
                pushed += dlt;
                sinkf += static_cast<unsigned long long>(dlt);
                sdlt = static_cast<unsigned long long>(dlt);
                ++pushes;
            }
        }
        uint32_t dl[4] = {0u, 0u, 0u, 0u};
        bool arcs = false;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int u = slot_of(nb, k);
            const uint32_t rk = slot_of(r, k);
            if (act && e > 0 && u >= 0 && rk != 0u && h == hu[k] + 1u) {
                dl[k] = static_cast<long long>(rk) < e ? rk : static_cast<uint32_t>(e);
                e -= dl[k];
                pushed += dl[k];
                arcs = true;
                ++pushes;
            }
        }
        if (arcs) {
#pragma unroll
            for (int k = 0; k < 4; ++k)
                if (dl[k]) {
                    res[static_cast<size_t>(v) * 4u + k] = slot_of(r, k) - dl[k];
                    touched[slot_of(nb, k)] = 1u;
                }
            // every entry of the row is zero before this store
            reinterpret_cast<uint4*>(out)[v] = make_uint4(dl[0], dl[1], dl[2], dl[3]);
        }
        if (pushed) excess[v] = e;
        // still holding excess: relabel-checked in the apply phase
        if (act && e > 0) touched[v] = 1u;
        flush_graph_flow<kB>(valid, v, sdlt, gid, gflow);
    }
    flush_stat(ctl, kStVisits, visits);
    flush_stat(ctl, kStPushes, pushes);
    flush_stat(ctl, kStSinkFlow, sinkf);
}

// Apply phase of a pull-mode sub-round with phase ph: a sweep of the first m
// bytes of the touched map (clearing it) in id order; every marked vertex
// takes in what its neighbours pushed into it, relabels, and joins the next
// active list when active.  Every vertex that holds excess after the push
// phase is marked: the active ones that kept some and every receiver.
template <bool kB>
__global__ void __launch_bounds__(kPullTPB) k_apply_relabel(uint32_t n, uint32_t m, uint32_t ph,
                                                            uint32_t* __restrict__ wl_next,
                                                            const int4* __restrict__ nbr4,
                                                            const uint8_t* __restrict__ rslot,
                                                            uint32_t* __restrict__ res,
                                                            const int32_t* __restrict__ term,
                                                            long long* __restrict__ excess, uint32_t* height,
                                                            uint32_t* __restrict__ out,
                                                            uint8_t* __restrict__ touched,
                                                            const uint32_t* __restrict__ gid,
                                                            const uint32_t* __restrict__ gn, Ctl* ctl) {
    __shared__ WgQueue<kPullTPB> q;
    __shared__ uint32_t s_ids[kPullStep];
    __shared__ uint32_t s_wsum[kPullTPB / 64];
    __shared__ uint32_t s_total;
    q_init(q);
    if (blockIdx.x == 0 && threadIdx.x == 0) ctl->rl_cnt[lbuf(ph) ^ 1u] = 0u;  // a two-phase sub-round may follow
    uint32_t* wl_cnt = &ctl->lc[lslot(ph, 1u)];
    const uint32_t lane = __lane_id(), wave = threadIdx.x >> 6;
    unsigned long long relabels = 0;
    const Span sp = wg_span(m, kPullStep);
    for (uint64_t base = sp.b; base < sp.e; base += kPullStep) {
        // 16 map bytes per lane, each 0 or 1
        const uint64_t o = base + 16u * threadIdx.x;
        uint32_t w[4] = {0u, 0u, 0u, 0u};
        if (o + 16u <= m) {
            const uint4 x = *reinterpret_cast<const uint4*>(touched + o);
            w[0] = x.x;
            w[1] = x.y;
            w[2] = x.z;
            w[3] = x.w;
        } else {
            for (uint32_t b = 0; b < 16u; ++b)
                if (o + b < m && touched[o + b]) w[b >> 2] |= 1u << (8u * (b & 3u));
        }
        const uint32_t c = __popc(w[0]) + __popc(w[1]) + __popc(w[2]) + __popc(w[3]);
        // exclusive prefix of c over the workgroup: lanes, then waves
        uint32_t incl = c;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const uint32_t y = __shfl_up(incl, d);
            if (static_cast<int>(lane) >= d) incl += y;
        }
        if (lane == 63u) s_wsum[wave] = incl;
        __syncthreads();
        uint32_t pos = incl - c;
        for (uint32_t x = 0; x < wave; ++x) pos += s_wsum[x];
        if (threadIdx.x == kPullTPB - 1) s_total = pos + c;
        if (c) {
#pragma unroll
            for (int b = 0; b < 16; ++b)
                if ((w[b >> 2] >> (8 * (b & 3))) & 1u) s_ids[pos++] = static_cast<uint32_t>(o) + static_cast<uint32_t>(b);
            if (o + 16u <= m) *reinterpret_cast<uint4*>(touched + o) = make_uint4(0u, 0u, 0u, 0u);
            else
                for (uint32_t b = 0; b < 16u; ++b)
                    if (o + b < m) touched[o + b] = 0u;
        }
        __syncthreads();
        const uint32_t total = s_total;
        for (uint32_t j = 0; j < total; j += kPullTPB) {
            const uint32_t k = j + threadIdx.x;
            bool act = false;
            uint32_t x = 0;
            if (k < total) {
                x = s_ids[k];
                long long e = excess[x];
                uint32_t h = height[x];
                const uint32_t sh = src_h<kB>(x, n, gid, gn);
                const int32_t t = term[x];
                const int4 nb = nbr4[x];
                const uint4 r = reinterpret_cast<const uint4*>(res)[x];
                const uint32_t rs = rslot[x];
                // each neighbour's entry for the slot back to x and its
                // height, all at once
                uint32_t in[4], hu[4];
#pragma unroll
                for (int kk = 0; kk < 4; ++kk) {
                    const int u = slot_of(nb, kk);
                    in[kk] = 0u;
                    hu[kk] = kSent;
                    if (u >= 0) {
                        in[kk] = out[static_cast<size_t>(u) * 4u + ((rs >> (2 * kk)) & 3u)];
                        hu[kk] = height[u];
                    }
                }
                long long gain = 0;
                uint32_t rr[4] = {r.x, r.y, r.z, r.w};
#pragma unroll
                for (int kk = 0; kk < 4; ++kk)
                    if (in[kk] != 0u) {
                        rr[kk] += in[kk];
                        gain += in[kk];
                        res[static_cast<size_t>(x) * 4u + kk] = rr[kk];
                        out[static_cast<size_t>(slot_of(nb, kk)) * 4u + ((rs >> (2 * kk)) & 3u)] = 0u;
                    }
                if (gain) {
                    e += gain;
                    excess[x] = e;
                }
                // relabel (as relabel_vertex)
                if ((e > 0) & (h < sh)) {
                    uint32_t mh = kSent;
                    if (t < 0) mh = 0u;
                    else if (t > 0) mh = sh;
#pragma unroll
                    for (int kk = 0; kk < 4; ++kk) {
                        if (slot_of(nb, kk) >= 0 && rr[kk] != 0u) {
                            const uint32_t y = hu[kk] == kSent ? sh + 1u : hu[kk];
                            mh = y < mh ? y : mh;
                        }
                    }
                    if (mh == kSent) {
                        set_error(ctl, 5);
                    } else if (mh + 1u > h) {
                        h = mh + 1u;
                        height[x] = h;
                        ++relabels;
                    }
                    act = h < sh;
                }
            }
            q_push(q, act, x);
            q_flush_if(q, wl_next, wl_cnt, kPullTPB);
        }
        __syncthreads();  // s_ids and s_wsum are rewritten by the next step
    }
    q_flush(q, wl_next, wl_cnt);
    flush_stat(ctl, kStRelabels, relabels);
}

// LDS-counter append for a single-workgroup kernel: ranks from ballot + mbcnt,
// one LDS atomic per wave, the item goes straight to the global list.
__device__ __forceinline__ void lds_append(bool pred, uint32_t val, uint32_t* list, uint32_t* s_cnt) {
    const uint64_t m = __ballot(pred);
    if (m == 0) return;
    const uint32_t rank = lane_rank(m);
    const int leader = __ffsll(static_cast<unsigned long long>(m)) - 1;
    uint32_t off = 0;
    if (pred && rank == 0) off = atomicAdd(s_cnt, static_cast<uint32_t>(__popcll(m)));
    off = __shfl(off, leader);
    if (pred) list[off + rank] = val;
}

// Discharge tail: when the active list is small, one workgroup runs the same
// sub-rounds back to back -- fused visits (kF) or a push and a relabel phase
// -- with __syncthreads instead of kernel boundaries.  Stops when the list
// empties, grows beyond `cap`, after max_sub sub-rounds, or once rel_budget
// height increases were made (the host then runs a global relabel).  Leaves
// the counters exactly as the same number of multi-workgroup sub-rounds would.
template <bool kB, bool kF>
__global__ void __launch_bounds__(kTailTPB) k_discharge_tail(uint32_t n, uint32_t ph0, uint32_t cap, uint32_t max_sub,
                                                         unsigned long long rel_budget, uint32_t* wl0, uint32_t* wl1,
                                                         uint32_t* rl, const int4* __restrict__ nbr4,
                                                         const uint8_t* __restrict__ rslot, uint32_t* res,
                                                         int32_t* __restrict__ term, long long* excess,
                                                         uint32_t* height,
                                                         const uint32_t* __restrict__ gid,
                                                         const uint32_t* __restrict__ gn,
                                                         unsigned long long* __restrict__ gflow, Ctl* ctl) {
    __shared__ uint32_t s_cnt;
    __shared__ uint32_t s_cur;
    __shared__ unsigned long long s_rel;
    __shared__ unsigned long long s_wrel[kTailTPB / 64];
    uint32_t* wl[2] = {wl0, wl1};
    if (threadIdx.x == 0) {
        s_cur = ctl->lc[lslot(ph0)];
        s_rel = 0ull;
    }
    __syncthreads();
    unsigned long long visits = 0, pushes = 0, sinkf = 0, relabels_all = 0;
    uint32_t sub = 0;
    for (;;) {
        const uint32_t cur = s_cur;
        if (cur == 0u || cur > cap || sub >= max_sub || s_rel >= rel_budget) break;
        const uint32_t p = lbuf(ph0 + sub);
        unsigned long long rel = 0;
        if constexpr (kF) {
            if (threadIdx.x == 0) s_cnt = 0u;
            __syncthreads();
            for (uint32_t base = 0; base < cur; base += kTailTPB) {
                const uint32_t i = base + threadIdx.x;
                const bool valid = i < cur;
                const uint32_t v = valid ? wl[p][i] : 0u;
                bool keep, cand[4];
                int4 nb;
                unsigned long long sdlt;
                fused_vertex<kB>(valid, v, n, nbr4, rslot, res, term, excess, height, gid, gn, keep, cand, nb,
                                 visits, pushes, sinkf, rel, sdlt);
                flush_graph_flow<kB>(valid, v, sdlt, gid, gflow);
                lds_append(keep, v, wl[p ^ 1u], &s_cnt);
#pragma unroll
                for (int k = 0; k < 4; ++k)
                    lds_append(cand[k], static_cast<uint32_t>(slot_of(nb, k)), wl[p ^ 1u], &s_cnt);
            }
        } else {
            // push phase
            if (threadIdx.x == 0) s_cnt = 0u;
            __syncthreads();
            for (uint32_t base = 0; base < cur; base += kTailTPB) {
                const uint32_t i = base + threadIdx.x;
                const bool valid = i < cur;
                const uint32_t v = valid ? wl[p][i] : 0u;
                bool keep, cand[4];
                int4 nb;
                unsigned long long sdlt;
                push_vertex<kB>(valid, v, n, nbr4, rslot, res, term, excess, height, gid, gn, keep, cand, nb,
                                visits, pushes, sinkf, sdlt);
                flush_graph_flow<kB>(valid, v, sdlt, gid, gflow);
                lds_append(keep, v, rl, &s_cnt);
#pragma unroll
                for (int k = 0; k < 4; ++k) lds_append(cand[k], static_cast<uint32_t>(slot_of(nb, k)), rl, &s_cnt);
            }
            __syncthreads();
            const uint32_t rcnt = s_cnt;
            __syncthreads();
            // relabel phase
            if (threadIdx.x == 0) s_cnt = 0u;
            __syncthreads();
            for (uint32_t base = 0; base < rcnt; base += kTailTPB) {
                const uint32_t i = base + threadIdx.x;
                bool act = false;
                uint32_t v = 0;
                if (i < rcnt) {
                    v = rl[i];
                    act = relabel_vertex<kB>(v, n, nbr4, res, term, excess, height, gid, gn, ctl, rel);
                }
                lds_append(act, v, wl[p ^ 1u], &s_cnt);
            }
        }
        relabels_all += rel;
        rel = wave_sum(rel);
        if (__lane_id() == 0) s_wrel[threadIdx.x >> 6] = rel;
        __syncthreads();
        if (threadIdx.x == 0) {
            unsigned long long t = 0;
            for (int w = 0; w < kTailTPB / 64; ++w) t += s_wrel[w];
            s_rel += t;
            s_cur = s_cnt;
        }
        __syncthreads();
        ++sub;
    }
    flush_stat(ctl, kStVisits, visits);
    flush_stat(ctl, kStPushes, pushes);
    flush_stat(ctl, kStSinkFlow, sinkf);
    flush_stat(ctl, kStRelabels, relabels_all);
    if (threadIdx.x == 0) {
        ctl->lc[lslot(ph0 + sub)] = s_cur;
        ctl->lc[lslot(ph0 + sub, 1u)] = 0u;
        ctl->rl_cnt[0] = 0u;
        ctl->rl_cnt[1] = 0u;
        ctl->tail_sub = sub;
    }
}

// ---------------------------------------------------- arena (serving) kernels
// A serving session keeps many graphs in one node arena.  Graph slot g owns
// nodes [lo, lo + ag), ag = its node count ng rounded up to a multiple of
// 1024 (the seed sweep's tile), so every tile belongs to one graph; nodes
// past ng are padding without arcs or terminals.  Freed ranges are inert:
// no arcs, no terminals, no excess, height 0 (never taken for unreached)
// and the sentinel graph slot, whose tiles the sweeps skip.

// Admission of graph slot g: union-wide neighbour ids (ids outside [0, ng)
// are device invariant 3, written to the slot's error word), graph ids,
// inert padding.
__global__ void __launch_bounds__(kTPB) k_admit(uint32_t lo, uint32_t ng, uint32_t ag, uint32_t g,
                                                const int4* __restrict__ nbr_local, int4* __restrict__ nbrg,
                                                uint4* __restrict__ res, int32_t* __restrict__ term,
                                                uint32_t* __restrict__ gid, int32_t* gerr) {
    int err = 0;
    for (uint32_t i = blockIdx.x * kTPB + threadIdx.x; i < ag; i += gridDim.x * kTPB) {
        const uint32_t v = lo + i;
        gid[v] = g;
        if (i < ng) {
            const int4 a = nbr_local[i];
            auto map = [&](int x) -> int {
                if (x < 0) return -1;
                if (static_cast<uint32_t>(x) >= ng) { err = 3; return -1; }
                return static_cast<int>(lo + static_cast<uint32_t>(x));
            };
            nbrg[v] = make_int4(map(a.x), map(a.y), map(a.z), map(a.w));
        } else {
            nbrg[v] = make_int4(-1, -1, -1, -1);
            res[v] = make_uint4(0u, 0u, 0u, 0u);
            term[v] = 0;
        }
    }
    if (err) __hip_atomic_exchange(&gerr[g], err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Per-slot tables of graph slot g: node count (its source height), flow,
// residency, the fresh flag (no heights yet) and a clear error word.
__global__ void k_set_slot(uint32_t g, uint32_t ng, uint32_t* gn, unsigned long long* gflow, uint8_t* gres,
                           uint8_t* gfresh, int32_t* gerr, uint8_t resident) {
    gerr[g] = 0;
    gn[g] = ng;
    gflow[g] = 0ull;
    gres[g] = resident;
    gfresh[g] = resident;
}

__global__ void k_fill_u8(uint8_t* p, uint32_t n, uint8_t x) {
    for (uint32_t i = blockIdx.x * kTPB + threadIdx.x; i < n; i += gridDim.x * kTPB) p[i] = x;
}

// partition of graph slot g (gpu_push_relabel_fast.cu:315-324) into its
// caller's buffer; vertices still active are invariant violations
__global__ void __launch_bounds__(kTPB) k_partition_range(uint32_t lo, uint32_t ng, uint32_t sh,
                                                          const uint32_t* __restrict__ height,
                                                          const long long* __restrict__ excess,
                                                          uint8_t* __restrict__ out, Ctl* ctl) {
    unsigned long long bad = 0;
    for (uint32_t i = blockIdx.x * kTPB + threadIdx.x; i < ng; i += gridDim.x * kTPB) {
        const uint32_t h = height[lo + i];
        out[i] = h < sh + 1u ? 1u : 0u;
        if (excess[lo + i] > 0 && h < sh) ++bad;
    }
    bad = wave_sum(bad);
    if (__lane_id() == 0 && bad) atom_add(&ctl->violations, static_cast<uint32_t>(bad));
}

// Freed nodes [lo, hi) become inert.
__global__ void __launch_bounds__(kTPB) k_inert(uint32_t lo, uint32_t hi, int4* __restrict__ nbrg,
                                                int32_t* __restrict__ term, long long* __restrict__ excess,
                                                uint32_t* __restrict__ height, uint32_t* __restrict__ gid,
                                                uint32_t sentinel) {
    for (uint32_t v = lo + blockIdx.x * kTPB + threadIdx.x; v < hi; v += gridDim.x * kTPB) {
        nbrg[v] = make_int4(-1, -1, -1, -1);
        term[v] = 0;
        excess[v] = 0;
        height[v] = 0u;
        gid[v] = sentinel;
    }
}

// partition = reaches the sink (gpu_push_relabel_fast.cu:315-324); counts any
// vertex still active after the final BFS (an invariant violation).
template <bool kB>
__global__ void __launch_bounds__(kTPB) k_partition(uint32_t n, const uint32_t* __restrict__ height,
                                                    const long long* __restrict__ excess,
                                                    uint8_t* __restrict__ part, const uint32_t* __restrict__ gid,
                                                    const uint32_t* __restrict__ gn, Ctl* ctl) {
    const uint32_t stride = gridDim.x * kTPB;
    unsigned long long bad = 0;
    for (uint32_t v = blockIdx.x * kTPB + threadIdx.x; v < n; v += stride) {
        const uint32_t h = height[v];
        const uint32_t sh = src_h<kB>(v, n, gid, gn);
        part[v] = h < sh + 1u ? 1u : 0u;
        if (excess[v] > 0 && h < sh) ++bad;
    }
    bad = wave_sum(bad);
    if (__lane_id() == 0 && bad) atom_add(&ctl->violations, static_cast<uint32_t>(bad));
}

// --------------------------------------------------------------- host side
inline void ck(hipError_t e, const char* where) {
    if (e == hipSuccess) return;
    std::ostringstream oss;
    oss << kPrefix << "HIP failure at " << where << ": " << hipGetErrorString(e);
    fail(WTV_EDEVICE, oss.str());
}

struct DeviceBuf {
    void* p = nullptr;
    ~DeviceBuf() { if (p) (void)hipFree(p); }
};

// Per-launch HIP event pairs on the solver's stream (profile mode only).
// Completed pairs are harvested in launch order; events are not reused within
// a solve, so pairs of launches still in flight can stay pending.
struct KernelTimer {
    enum Kind { kPush = 0, kRelabel = 1, kLevel = 2, kTail = 3 };
    bool on = false;
    hipStream_t s = nullptr;
    std::vector<hipEvent_t> pool;
    std::vector<std::pair<int, size_t>> pending;  // (kind, index of start event)
    size_t head = 0;                              // first pending entry not yet harvested
    double ms[4] = {0, 0, 0, 0};
    uint64_t count[4] = {0, 0, 0, 0};
    ~KernelTimer() { for (auto e : pool) (void)hipEventDestroy(e); }
    hipEvent_t next() {
        hipEvent_t e;
        ck(hipEventCreate(&e), "event create");
        pool.push_back(e);
        return e;
    }
    void begin(int kind) {
        if (!on) return;
        pending.emplace_back(kind, pool.size());
        ck(hipEventRecord(next(), s), "event");
    }
    void end() {
        if (!on) return;
        ck(hipEventRecord(next(), s), "event");
    }
    void harvest() {
        if (!on) return;
        while (head < pending.size()) {
            const auto& pr = pending[head];
            if (hipEventQuery(pool[pr.second + 1]) != hipSuccess) break;
            float t = 0.f;
            ck(hipEventElapsedTime(&t, pool[pr.second], pool[pr.second + 1]), "event time");
            ms[pr.first] += t;
            count[pr.first] += 1;
            ++head;
        }
    }
};

inline size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

__global__ void k_noop() {}

// Cost of the event pair the profile mode puts around every launch, on
// stream s: an empty kernel of the push launch shape timed kCal times with a
// pair of its own, against kCal back-to-back launches in one pair.  The pair
// adds the dispatch and completion latency that back-to-back launches hide,
// which a kernel trace does not count as the kernel's duration.
double event_pair_overhead_us(int grid, int wg, hipStream_t s) {
    constexpr int kCal = 32;
    std::vector<hipEvent_t> ev(2 * kCal + 2, nullptr);
    struct Guard {
        std::vector<hipEvent_t>& e;
        ~Guard() { for (auto x : e) if (x) (void)hipEventDestroy(x); }
    } guard{ev};
    for (auto& e : ev) ck(hipEventCreate(&e), "event create");
    for (int i = 0; i < kCal; ++i) {
        ck(hipEventRecord(ev[2 * i], s), "event");
        k_noop<<<grid, wg, 0, s>>>();
        ck(hipEventRecord(ev[2 * i + 1], s), "event");
    }
    ck(hipEventRecord(ev[2 * kCal], s), "event");
    for (int i = 0; i < kCal; ++i) k_noop<<<grid, wg, 0, s>>>();
    ck(hipEventRecord(ev[2 * kCal + 1], s), "event");
    ck(hipGetLastError(), "noop launch");
    ck(hipEventSynchronize(ev[2 * kCal + 1]), "event sync");
    double pairs = 0.0;
    for (int i = 0; i < kCal; ++i) {
        float t = 0.f;
        ck(hipEventElapsedTime(&t, ev[2 * i], ev[2 * i + 1]), "event time");
        pairs += t;
    }
    float batch = 0.f;
    ck(hipEventElapsedTime(&batch, ev[2 * kCal], ev[2 * kCal + 1]), "event time");
    return std::max(0.0, (pairs - static_cast<double>(batch)) / kCal * 1e3);
}

// Resources of one solve: the solver's own non-blocking stream (default
// priority; graph capture needs a stream other than the null stream, and the
// caller's stream is ordered before and after through events), a pinned block
// for control-block snapshots and the events of the host loop.  They come
// from a process-wide pool, per device, and go back to it when the solve
// ends, so a process holds as many sets as it ever ran solves at once: server
// worker threads that come and go create none of their own, and no set is
// ever released (nothing is torn down at process exit).
constexpr int kSolveEvents = 10;
constexpr int kCtlSnapshots = 5;
struct SolveRes {
    int dev = 0;
    hipStream_t s = nullptr;
    void* hctl = nullptr;  // kCtlSnapshots * sizeof(Ctl), pinned
    hipEvent_t ev[kSolveEvents] = {};
};

class ResPool {
  public:
    SolveRes acquire() {
        int dev = 0;
        ck(hipGetDevice(&dev), "hipGetDevice");
        {
            std::lock_guard<std::mutex> lk(mu_);
            for (size_t i = 0; i < free_.size(); ++i)
                if (free_[i].dev == dev) {
                    SolveRes r = free_[i];
                    free_.erase(free_.begin() + static_cast<std::ptrdiff_t>(i));
                    return r;
                }
        }
        SolveRes r;
        r.dev = dev;
        struct Partial {  // releases what was created if a later step fails
            SolveRes& r;
            bool done = false;
            ~Partial() {
                if (done) return;
                for (auto e : r.ev) if (e) (void)hipEventDestroy(e);
                if (r.hctl) (void)hipHostFree(r.hctl);
                if (r.s) (void)hipStreamDestroy(r.s);
            }
        } partial{r};
        ck(hipStreamCreateWithFlags(&r.s, hipStreamNonBlocking), "stream create");
        ck(hipHostMalloc(&r.hctl, kCtlSnapshots * sizeof(Ctl), hipHostMallocDefault), "pinned ctl");
        for (auto& e : r.ev) ck(hipEventCreate(&e), "event create");
        partial.done = true;
        return r;
    }
    void release(const SolveRes& r) {
        std::lock_guard<std::mutex> lk(mu_);
        free_.push_back(r);
    }

  private:
    std::mutex mu_;
    std::vector<SolveRes> free_;
};

ResPool& res_pool() {
    static ResPool* pool = new ResPool;  // never destroyed: no HIP calls at exit
    return *pool;
}

// A set of solve resources for the lifetime of one solve.
struct ResLease {
    SolveRes r;
    ResLease() : r(res_pool().acquire()) {}
    ~ResLease() { res_pool().release(r); }
    ResLease(const ResLease&) = delete;
    ResLease& operator=(const ResLease&) = delete;
};

// Waits for a stream when the scope ends, also when an exception unwinds, so
// that the buffers declared before it outlive every kernel queued on it.
struct DrainOnExit {
    hipStream_t s;
    ~DrainOnExit() { (void)hipStreamSynchronize(s); }
};

struct GraphExec {
    hipGraphExec_t exec = nullptr;
    ~GraphExec() { if (exec) (void)hipGraphExecDestroy(exec); }
};

struct Layout {
    size_t height, excess, rslot, nzm, wl0, wl1, rl, ctl, gid, nbrg, gn, goff, gflow, galive, gres, gfresh, gerr,
        ares, aterm, pout, ptouched, total;
};

// One allocation carved into 256-byte aligned arrays.  Batched mode adds the
// graph id per node, a union-wide copy of the neighbour ids and three
// per-graph tables (node count, offsets, sink flow).  An arena (serving) also
// holds the residual rows and terminals of its graphs, two more per-slot
// byte tables (resident, fresh) and a per-slot error word (admission checks).
// Pull-mode sub-rounds add 17 B per node: the push rows and the touched map.
Layout layout_for(uint32_t n, uint32_t G, bool arena = false, bool pull = false) {
    Layout L{};
    size_t o = 0;
    auto take = [&](size_t bytes) { size_t at = o; o = align_up(o + bytes, 256); return at; };
    const size_t N = n;
    L.ctl = take(sizeof(Ctl));
    L.excess = take(N * 8u);
    L.height = take(N * 4u);
    L.wl0 = take(N * 4u);
    L.wl1 = take(N * 4u);
    L.rl = take(N * 4u);
    L.rslot = take(N);
    L.nzm = take(N);
    if (pull) {
        L.pout = take(N * 16u);
        L.ptouched = take(N);
    }
    if (G) {
        L.gid = take(N * 4u);
        L.nbrg = take(N * 16u);
        L.gn = take(static_cast<size_t>(G) * 4u);
        L.goff = take((static_cast<size_t>(G) + 1u) * 8u);
        L.gflow = take(static_cast<size_t>(G) * 8u);
        L.galive = take(G);
        if (arena) {
            L.gres = take(G);
            L.gfresh = take(G);
            L.gerr = take(static_cast<size_t>(G) * 4u);
            L.ares = take(N * 16u);
            L.aterm = take(N * 4u);
        }
    }
    L.total = o;
    return L;
}

// Calls f with std::integral_constant<int, wg> for the supported list-kernel
// workgroup sizes (checked by the caller).
template <class F>
void launch_wg(int wg, F&& f) {
    if (wg == 1024) f(std::integral_constant<int, 1024>{});
    else if (wg == 512) f(std::integral_constant<int, 512>{});
    else f(std::integral_constant<int, 256>{});
}

struct BatchArgs {
    uint32_t G;
    const uint64_t* h_off;  // G+1 node offsets (host)
    int64_t* h_flows;       // G raw flows (host, out)
};

}  // namespace

// The device arrays of an arena solve, as the feeder that fills it sees them.
struct ArenaView {
    hipStream_t s;
    Ctl* ctl;
    int4* nbrg;
    uint4* res;
    int32_t* term;
    long long* excess;
    uint32_t* height;
    uint8_t* rslot;
    uint32_t* gid;
    uint32_t* gn;
    unsigned long long* gflow;
    uint8_t* galive;  // graphs with an active vertex (after a relabel's filter)
    uint8_t* gres;    // resident graphs
    uint8_t* gfresh;  // graphs admitted since the last relabel
    int32_t* gerr;    // per slot: the first failed admission check of its graph
    uint32_t* wl[2];
    uint32_t capacity;  // nodes
    uint32_t slots;     // graph slots; slot `slots` is the sentinel of free nodes
    uint32_t hwm;       // node bound of the sweeps: the end of the highest used range
};

// A source of graphs for an arena solve (the serving loop).  solve_t calls it
// at every global relabel, and when the active list drains, on the solve's
// own host thread and stream.
struct Feeder {
    virtual ~Feeder() = default;
    // Before a global relabel: admit waiting graphs (their vertices holding
    // source supply are appended to the active list wl[p]) and set hwm.
    virtual void admit(ArenaView& a, uint32_t ph) = 0;
    // After the relabel and the active-list filter: finish every resident
    // graph without an active vertex (its heights are final).
    virtual void finish(ArenaView& a, uint32_t ph) = 0;
    // Requests are waiting that the arena could take now.
    virtual bool pending() = 0;
    // The active list is empty and nothing is resident: block until a graph
    // can be admitted (true) or the session stops (false).
    virtual bool wait_for_work() = 0;
    // Graphs are resident (admitted, not yet finished).
    virtual bool resident() = 0;
};

// Optional arguments of solve_t.
struct SolveExtras {
    void* ws = nullptr;         // workspace provided by the caller (else allocated)
    bool noalloc = false;       // allocate nothing: caller's stream, no events or graphs
    Feeder* feeder = nullptr;   // arena solve fed at relabels (ws: layout_for(n, slots + 1, true, pull))
    uint32_t slots = 0;         // graph slots of the arena
};

// X.noalloc: the solve allocates nothing (no device or pinned memory,
// streams, events or graphs).  X.ws is then a workspace of at least
// layout_for(n, 0, false, false).total bytes (no pull-mode arrays: its sub-rounds
// are all atomic ones), the kernels run on the caller's stream, the
// host reads the control block into thread-local pageable memory, and every
// wait is a stream synchronisation.
// X.feeder: an arena solve (kB): n is the arena's node capacity, the graphs
// arrive through the feeder at global relabels and leave at the relabel that
// finds them without an active vertex; d_nbr, d_res, d_term and partition are
// unused (the arena holds the rows, the feeder writes each graph's results).
template <bool kB>
void solve_t(uint32_t n, const int32_t* d_nbr, uint32_t* d_res, int32_t* d_term, uint8_t* partition,
             bool partition_on_device, const wtv_params& prm, wtv_stats& st, void* stream_v, const BatchArgs* B,
             const SolveExtras& X = SolveExtras{}) {
    using Clock = std::chrono::steady_clock;
    std::memset(&st, 0, sizeof(st));
    if (!n) return;
    const bool noalloc = X.noalloc;
    Feeder* const feeder = kB ? X.feeder : nullptr;
    if (noalloc && prm.profile != 0)
        fail(WTV_EINVAL, std::string(kPrefix) + "profile mode needs events: not available without allocation");
    if (!feeder && (!d_nbr || !d_res || !d_term || !partition))
        fail(WTV_EINVAL, std::string(kPrefix) + "null graph pointer");
    if (n > 0x7FFFFFFFu) fail(WTV_EINVAL, std::string(kPrefix) + ">INT32_MAX reduced nodes unsupported");
    int local_steps = prm.local_steps;
    if (local_steps < 1) local_steps = 1;
    if (local_steps > 32) local_steps = 32;
    int period = prm.global_relabel_period < 0 ? 0 : prm.global_relabel_period;
    uint64_t max_rounds = prm.max_rounds ? prm.max_rounds : 2000000ull;
    const double time_limit = prm.time_limit_s > 0 ? prm.time_limit_s : 240.0;
    const int tpb = kTPB;  // sweeps over all vertices
    // 2048 workgroups per list launch: the BFS levels take ~6% less per
    // relabel than with 1024 on the 77.5M-tet Delaunay graph, the discharge
    // launches no longer (4096: the same)
    int grid = prm.grid_wgs > 0 ? prm.grid_wgs : 2048;
    // workgroup size of the list kernels (push, relabel, BFS level, filter)
    const int wg = prm.wg_size == 0 ? 256 : prm.wg_size;
    if (wg != 256 && wg != 512 && wg != 1024)
        fail(WTV_EINVAL, std::string(kPrefix) + "wg_size must be 0 (256), 256, 512 or 1024");
    // node bound of the sweeps: n, or an arena's high-water mark
    uint32_t nn = feeder ? 0u : n;
    auto sweep_grid_of = [&](uint32_t m) {
        return std::max(1, static_cast<int>(std::min<uint64_t>((static_cast<uint64_t>(m) + tpb - 1) / tpb, 8192)));
    };
    int sweep_grid = sweep_grid_of(nn);
    // a level whose frontier holds more than n/32 vertices expands bottom-up
    uint32_t bu_threshold = std::max<uint32_t>(nn / 32u, 4096u);
    hipStream_t cs = static_cast<hipStream_t>(stream_v);
    std::optional<ResLease> lease;
    if (!noalloc) lease.emplace();
    hipStream_t s = noalloc ? cs : lease->r.s;
    struct EvRef {
        hipEvent_t e;  // null without allocation: waits become stream synchronisations
    };
    auto lease_ev = [&](int i) -> EvRef { return EvRef{noalloc ? nullptr : lease->r.ev[i]}; };
    auto ev_record = [&](const EvRef& e) {
        if (e.e) ck(hipEventRecord(e.e, s), "event");
    };
    auto ev_wait = [&](const EvRef& e, const char* where) {
        if (e.e) ck(hipEventSynchronize(e.e), where);
        else ck(hipStreamSynchronize(s), where);
    };
    EvRef ev_caller = lease_ev(0), ev_done = lease_ev(1);
    if (!noalloc) {
        ck(hipEventRecord(ev_caller.e, cs), "event");
        ck(hipStreamWaitEvent(s, ev_caller.e, 0), "stream wait");
    }
    const bool use_graph = prm.profile == 0 && prm.no_graph == 0 && !noalloc;
    // WTV_TRACE=1: one stderr line per macro round / global relabel (diagnostics)
    const char* trace_env = std::getenv("WTV_TRACE");
    const bool trace = trace_env && *trace_env && *trace_env != '0';

    const uint32_t G = kB ? (feeder ? X.slots + 1u : B->G) : 0u;
    // pull-mode sub-rounds (never in the allocation-free solve): lists of more
    // than pull_min vertices
    const bool pull_on = !noalloc && pull_enabled(prm);
    const Layout L = layout_for(n, G, feeder != nullptr, pull_on);
    DeviceBuf ws;
    if (!X.ws) ck(hipMalloc(&ws.p, L.total), "workspace malloc");
    // every exit from here on, failures included, first waits for the kernels
    // still queued on s (a speculative round, a BFS batch), then frees ws
    DrainOnExit drain{s};
    st.solver_workspace_bytes = L.total;
    char* base = static_cast<char*>(X.ws ? X.ws : ws.p);
    Ctl* ctl = reinterpret_cast<Ctl*>(base + L.ctl);
    long long* excess = reinterpret_cast<long long*>(base + L.excess);
    uint32_t* height = reinterpret_cast<uint32_t*>(base + L.height);
    uint32_t* wl[2] = {reinterpret_cast<uint32_t*>(base + L.wl0), reinterpret_cast<uint32_t*>(base + L.wl1)};
    uint32_t* rl = reinterpret_cast<uint32_t*>(base + L.rl);
    uint8_t* rslot = reinterpret_cast<uint8_t*>(base + L.rslot);
    uint8_t* nzm = reinterpret_cast<uint8_t*>(base + L.nzm);
    const uint64_t pull_min = prm.pull_min > 0 ? static_cast<uint64_t>(prm.pull_min) : kPullMinList;
    uint32_t* pout = pull_on ? reinterpret_cast<uint32_t*>(base + L.pout) : nullptr;
    uint8_t* ptouched = pull_on ? reinterpret_cast<uint8_t*>(base + L.ptouched) : nullptr;
    const int4* nbr4 = reinterpret_cast<const int4*>(d_nbr);
    uint32_t* gid = kB ? reinterpret_cast<uint32_t*>(base + L.gid) : nullptr;
    uint32_t* gn = kB ? reinterpret_cast<uint32_t*>(base + L.gn) : nullptr;
    unsigned long long* goff = kB ? reinterpret_cast<unsigned long long*>(base + L.goff) : nullptr;
    unsigned long long* gflow = kB ? reinterpret_cast<unsigned long long*>(base + L.gflow) : nullptr;
    uint8_t* galive = kB ? reinterpret_cast<uint8_t*>(base + L.galive) : nullptr;
    ArenaView av{};
    if (feeder) {
        // the arena's own rows stand for the caller's arrays
        d_res = reinterpret_cast<uint32_t*>(base + L.ares);
        d_term = reinterpret_cast<int32_t*>(base + L.aterm);
        nbr4 = reinterpret_cast<const int4*>(base + L.nbrg);
        av = ArenaView{s, ctl, reinterpret_cast<int4*>(base + L.nbrg), reinterpret_cast<uint4*>(d_res), d_term,
                       excess, height, rslot, gid, gn, gflow, galive, reinterpret_cast<uint8_t*>(base + L.gres),
                       reinterpret_cast<uint8_t*>(base + L.gfresh), reinterpret_cast<int32_t*>(base + L.gerr),
                       {wl[0], wl[1]}, n, X.slots, 0u};
    }

    // pinned snapshots, or thread-local pageable ones without allocation
    thread_local Ctl tl_ctl[kCtlSnapshots];
    Ctl* hctl = noalloc ? tl_ctl : static_cast<Ctl*>(lease->r.hctl);
    Ctl* snap = hctl + 1;   // two pipelined snapshots of the discharge loop
    Ctl* bsnap = hctl + 3;  // two of the BFS level batches
    EvRef bfs_ev[2] = {lease_ev(2), lease_ev(3)};

    KernelTimer* ktp = nullptr;
    auto read_ctl = [&](const char* where) {
        ck(hipMemcpyAsync(hctl, ctl, sizeof(Ctl), hipMemcpyDeviceToHost, s), where);
        ck(hipStreamSynchronize(s), where);
        if (ktp) ktp->harvest();
    };
    auto check_err = [&](const char* stage) {
        read_ctl(stage);
        if (hctl->error) {
            std::ostringstream oss;
            oss << kPrefix << "device invariant failed at " << stage << " code=" << hctl->error;
            fail(WTV_EINVARIANT, oss.str());
        }
    };

    EvRef ev_t0 = lease_ev(4), ev_t1 = lease_ev(5), ev_b0 = lease_ev(6), ev_b1 = lease_ev(7);
    const auto host_t0 = Clock::now();
    double bfs_ms = 0.0;
    // Cost guard of the adaptive relabel trigger: a global relabel may run
    // once the time since the last one reaches kRelabelCost times that one's
    // duration, so relabels take at most about 1/(1+kRelabelCost) of a solve
    // (a relabel costs ~1 ms on a soft-seed 78M graph, but ~15 ms where the
    // BFS sweeps most of the graph, as in the reference capacity model).
    double last_gr_ms = 0.0;
    auto last_gr_end = Clock::now();
    KernelTimer kt;
    kt.on = prm.profile != 0;
    kt.s = s;
    ktp = &kt;

    ck(hipMemsetAsync(ctl, 0, sizeof(Ctl), s), "ctl clear");
    if (pull_on) {
        // every push row entry zero, no node marked
        ck(hipMemsetAsync(pout, 0, static_cast<size_t>(n) * 16u, s), "pull rows clear");
        ck(hipMemsetAsync(ptouched, 0, static_cast<size_t>(n), s), "pull map clear");
    }
    std::vector<uint32_t> h_gn;
    if (feeder) {
        // no graph resident: every slot free, the sentinel never resident
        k_fill_u8<<<1, kTPB, 0, s>>>(av.gres, G, 0u);
        k_fill_u8<<<1, kTPB, 0, s>>>(av.gfresh, G, 0u);
        k_fill_u8<<<1, kTPB, 0, s>>>(galive, G, 0u);
        ck(hipGetLastError(), "arena tables");
    } else if constexpr (kB) {
        h_gn.resize(G);
        for (uint32_t g = 0; g < G; ++g) h_gn[g] = static_cast<uint32_t>(B->h_off[g + 1] - B->h_off[g]);
        ck(hipMemcpyAsync(gn, h_gn.data(), G * 4u, hipMemcpyHostToDevice, s), "batch sizes");
        ck(hipMemcpyAsync(goff, B->h_off, (G + 1u) * 8u, hipMemcpyHostToDevice, s), "batch offsets");
        ck(hipMemsetAsync(gflow, 0, G * 8u, s), "batch flows");
    }
    ev_record(ev_t0);
    if (!feeder) {
        if constexpr (kB) {
            int4* nbrg = reinterpret_cast<int4*>(base + L.nbrg);
            k_globalize<<<sweep_grid, tpb, 0, s>>>(n, G, goff, reinterpret_cast<const int4*>(d_nbr), nbrg, gid,
                                                   ctl);
            ck(hipGetLastError(), "globalize");
            nbr4 = nbrg;
        }
        k_prepare<<<sweep_grid, tpb, 0, s>>>(0u, n, nbr4, d_term, rslot, excess, wl[0], &ctl->lc[0], ctl,
                                             nullptr);
        ck(hipGetLastError(), "prepare");
    }
    check_err("prepare");

    uint64_t sub = 0;  // sub-rounds run (statistics)
    // List position: the active list is wl[pos & 1], its count lc[pos % 3].
    // It advances with every sub-round and once per active-list filter after
    // a global relabel; kernels take its phase pos % 6.
    uint64_t pos = 0;
    auto phase = [&](uint64_t at) { return static_cast<uint32_t>(at % 6u); };
    // BFS from the sink.  Its frontiers ping-pong between the relabel list
    // (dead between sub-rounds) and the idle worklist, so the active list
    // wl[pos & 1] survives for the filter that follows.
    // skip_done (batched solves, relabels between the first and the final
    // one): the seed sweep passes over graphs without an active vertex
    auto global_relabel = [&](bool skip_done) {
        ev_record(ev_b0);
        const auto host_b0 = Clock::now();
        const uint32_t p = static_cast<uint32_t>(pos & 1u);
        uint32_t* bf[2] = {rl, wl[p ^ 1u]};
        k_bfs_begin<<<1, 1, 0, s>>>(ctl, phase(pos));
        ck(hipGetLastError(), "bfs begin");
        // an arena relabels every resident graph: its seed sweep skips the
        // tiles of free nodes and writes every height of fresh graphs
        const bool skip = kB && !feeder && skip_done && st.global_relabels > 0;
        if (skip) {
            k_clear_alive<<<std::max(1, static_cast<int>(std::min<uint64_t>((G + tpb - 1) / tpb, 1024))), tpb, 0,
                            s>>>(G, galive);
            k_mark_alive<<<grid, tpb, 0, s>>>(phase(pos), wl[p], gid, galive, ctl);
            ck(hipGetLastError(), "mark alive");
        }
        k_bfs_seed<<<sweep_grid, tpb, 0, s>>>(nn, d_term, nbr4, d_res, height, nzm, bf[0],
                                              st.global_relabels == 0 ? 1u : 0u, gid,
                                              feeder ? av.gres : (skip ? galive : nullptr),
                                              feeder ? av.gfresh : nullptr, ctl);
        ck(hipGetLastError(), "bfs seed");
        // Levels run in batches of 8 launches.  Batches are pipelined: batch
        // b+1 is enqueued before the frontier count after batch b is read, and
        // a level whose frontier is empty returns at once, so running past the
        // last level costs a few microseconds instead of a host round trip
        // per batch.
        uint32_t d = 2;
        const uint32_t batch = 8;
        int next = 0;
        struct Mark {
            int slot;
            uint32_t d;  // level whose frontier count the snapshot holds
        };
        auto enqueue_batch = [&]() -> Mark {
            for (uint32_t b = 0; b < batch; ++b, ++d) {
                kt.begin(KernelTimer::kLevel);
                launch_wg(wg, [&](auto t) {
                    constexpr int T = decltype(t)::value;
                    k_bfs_level<T><<<grid, T, 0, s>>>(nn, d, bu_threshold, bf[d & 1u], bf[(d + 1u) & 1u], nbr4,
                                                      rslot, nzm, height, ctl);
                });
                kt.end();
            }
            ck(hipGetLastError(), "bfs expand");
            const int sl = next;
            next ^= 1;
            ck(hipMemcpyAsync(&bsnap[sl], ctl, sizeof(Ctl), hipMemcpyDeviceToHost, s), "bfs snapshot");
            ev_record(bfs_ev[sl]);
            return Mark{sl, d};
        };
        Mark prev = enqueue_batch();
        for (;;) {
            const Mark cur = enqueue_batch();  // speculative
            ev_wait(bfs_ev[prev.slot], "bfs snapshot wait");
            if (ktp) ktp->harvest();
            const uint32_t last = bsnap[prev.slot].bfs_cnt[prev.d % 3u];
            if (last == 0u) break;  // the speculative batch ran on empty frontiers
            if (static_cast<uint64_t>(prev.d) > static_cast<uint64_t>(n) + 1ull)
                fail(WTV_EINVARIANT, std::string(kPrefix) + "BFS exceeded n+1 levels");
            if (last <= kTailMax) {
                // small frontier: one workgroup runs the levels back to back,
                // from the level the speculative batch reached (it takes the
                // frontier only while it holds at most kTailMax vertices)
                k_bfs_tail<<<1, kTailTPB, 0, s>>>(cur.d, 4096u, bf[0], bf[1], nbr4, rslot, nzm, height, ctl);
                ck(hipGetLastError(), "bfs tail");
                read_ctl("bfs tail");
                d = hctl->tail_d;
                if (hctl->bfs_cnt[d % 3u] == 0u) break;
                if (static_cast<uint64_t>(d) > static_cast<uint64_t>(n) + 1ull)
                    fail(WTV_EINVARIANT, std::string(kPrefix) + "BFS exceeded n+1 levels");
                prev = enqueue_batch();
                continue;
            }
            prev = cur;
        }
        ev_record(ev_b1);
        ev_wait(ev_b1, "event sync");
        float ms = 0.f;
        if (noalloc) ms = std::chrono::duration<float, std::milli>(Clock::now() - host_b0).count();
        else ck(hipEventElapsedTime(&ms, ev_b0.e, ev_b1.e), "event time");
        bfs_ms += ms;
        last_gr_ms = ms;
        last_gr_end = Clock::now();
        ++st.global_relabels;
        if (trace) read_ctl("trace");  // the level count, current
        if (trace)
            std::fprintf(stderr, "[wtv] global relabel %llu: levels=%u %.3f ms\n",
                         static_cast<unsigned long long>(st.global_relabels), hctl->bfs_levels, ms);
    };
    auto filter_active = [&]() -> uint32_t {
        const uint32_t p = static_cast<uint32_t>(pos & 1u);
        launch_wg(wg, [&](auto t) {
            constexpr int T = decltype(t)::value;
            k_filter_active<kB, T><<<grid, T, 0, s>>>(n, phase(pos), wl[p], wl[p ^ 1u], excess, height, gid, gn,
                                                      ctl);
        });
        ck(hipGetLastError(), "filter active");
        ++pos;
        read_ctl("active count");
        return hctl->lc[lslot(phase(pos))];
    };

    // An arena admits graphs before a relabel (which sets their heights) and
    // finishes, after the relabel and the filter, every graph left without
    // an active vertex: that relabel's BFS ran on its final residual graph.
    auto relabel_and_filter = [&](bool skip_done) -> uint32_t {
        if (feeder) {
            feeder->admit(av, phase(pos));
            nn = av.hwm;
            sweep_grid = sweep_grid_of(nn);
            bu_threshold = std::max<uint32_t>(nn / 32u, 4096u);
        }
        global_relabel(skip_done);
        const uint32_t act = filter_active();
        if (feeder) feeder->finish(av, phase(pos));
        return act;
    };

    uint32_t active = 0;
    if (!feeder) {
        global_relabel(false);
        check_err("global relabel");
        active = filter_active();
    }

    int macro_period = 0;
    if (period > 0) macro_period = std::max(1, (period + local_steps - 1) / local_steps);
    // Adaptive cadence: a global relabel also runs once the relabel work since
    // the last one exceeds alpha * (active vertices right after it) + 1024.
    const double alpha = prm.relabel_alpha == 0.0 ? 2.0 : prm.relabel_alpha;
    uint64_t relabels_at_gr = stat_sum(*hctl, kStRelabels);
    uint64_t active_at_gr = active;
    // share of the vertices the last global relabel's BFS expanded
    uint64_t pops_at_gr = stat_sum(*hctl, kStBfsPops);
    double last_gr_reach = nn ? static_cast<double>(pops_at_gr) / static_cast<double>(nn) : 0.0;

    // One macro round = local_steps sub-rounds.  Rounds are replayed from a
    // captured graph (one per mode and starting phase) and pipelined: round
    // r+1 is enqueued before the counters of round r are read, so the device
    // never idles on the host's decision.
    // [mode][starting phase]; mode 0 = fused push-relabel, 1 = pull-mode
    // sub-rounds, 2 = push and relabel kernels
    GraphExec gexec[3][6];
    const int small_mode = prm.fused < 0 ? 2 : 0;
    // the latest active-list size the host has read; it picks the mode of
    // the rounds launched next (either mode gives the same results)
    uint64_t known_active = active;
    auto launch_round = [&]() {
        const int mode = pull_on && known_active > pull_min ? 1 : small_mode;
        auto enqueue = [&](uint64_t pos0) {
            for (int j = 0; j < local_steps; ++j) {
                const uint32_t ph = phase(pos0 + static_cast<uint64_t>(j));
                const uint32_t p = lbuf(ph);
                if (mode == 0) {
                    kt.begin(KernelTimer::kPush);
                    launch_wg(wg, [&](auto t) {
                        constexpr int T = decltype(t)::value;
                        k_push_fused<kB, T><<<grid, T, 0, s>>>(n, ph, wl[p], wl[p ^ 1u], nbr4, rslot, d_res, d_term,
                                                                excess, height, gid, gn, gflow, ctl);
                    });
                    kt.end();
                    continue;
                }
                if (mode == 1) {
                    kt.begin(KernelTimer::kPush);
                    k_push_pull<kB><<<grid, kPullTPB, 0, s>>>(n, ph, wl[p], nbr4, d_res, d_term, excess, height, pout,
                                                              ptouched, gid, gn, gflow, ctl);
                    kt.end();
                    kt.begin(KernelTimer::kRelabel);
                    k_apply_relabel<kB><<<grid, kPullTPB, 0, s>>>(n, n, ph, wl[p ^ 1u], nbr4, rslot, d_res, d_term,
                                                                  excess, height, pout, ptouched, gid, gn, ctl);
                    kt.end();
                    continue;
                }
                kt.begin(KernelTimer::kPush);
                launch_wg(wg, [&](auto t) {
                    constexpr int T = decltype(t)::value;
                    k_push<kB, T><<<grid, T, 0, s>>>(n, ph, wl[p], rl, nbr4, rslot, d_res, d_term, excess, height,
                                                     gid, gn, gflow, ctl);
                });
                kt.end();
                kt.begin(KernelTimer::kRelabel);
                launch_wg(wg, [&](auto t) {
                    constexpr int T = decltype(t)::value;
                    k_relabel<kB, T><<<grid, T, 0, s>>>(n, ph, rl, wl[p ^ 1u], nbr4, d_res, d_term, excess, height,
                                                        gid, gn, ctl);
                });
                kt.end();
            }
        };
        if (use_graph) {
            GraphExec& ge = gexec[mode][phase(pos)];
            if (!ge.exec) {
                hipGraph_t g = nullptr;
                ck(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "begin capture");
                enqueue(pos);
                ck(hipStreamEndCapture(s, &g), "end capture");
                const hipError_t e = hipGraphInstantiate(&ge.exec, g, nullptr, nullptr, 0);
                (void)hipGraphDestroy(g);
                ck(e, "graph instantiate");
            }
            ck(hipGraphLaunch(ge.exec, s), "graph launch");
        } else {
            enqueue(pos);
        }
        if (mode == 1) ++st.pull_rounds;
        if (mode == 0) ++st.fused_rounds;
        ck(hipGetLastError(), "discharge");
        sub += static_cast<uint64_t>(local_steps);
        pos += static_cast<uint64_t>(local_steps);
    };
    struct Pending {
        int slot;
        uint64_t pos_after;
        EvRef* ev;
    };
    EvRef snap_ev[2] = {lease_ev(8), lease_ev(9)};
    int next_slot = 0;
    auto snapshot = [&]() -> Pending {
        const int sl = next_slot;
        next_slot ^= 1;
        ck(hipMemcpyAsync(&snap[sl], ctl, sizeof(Ctl), hipMemcpyDeviceToHost, s), "snapshot");
        ev_record(snap_ev[sl]);
        return Pending{sl, pos, &snap_ev[sl]};
    };

    // Small active lists (<= tail_cap) go to the one-workgroup discharge kernel.
    const uint32_t tail_cap = prm.tail_cap == 0 ? 512u : (prm.tail_cap < 0 ? 0u : static_cast<uint32_t>(prm.tail_cap));
    uint64_t relabels_now = relabels_at_gr;
    auto do_global_relabel = [&]() {
        active = relabel_and_filter(prm.skip_finished >= 0);
        relabels_now = relabels_at_gr = stat_sum(*hctl, kStRelabels);
        const uint64_t pops = stat_sum(*hctl, kStBfsPops);
        last_gr_reach = nn ? static_cast<double>(pops - pops_at_gr) / static_cast<double>(nn) : 0.0;
        pops_at_gr = pops;
        active_at_gr = active;
    };
    // A relabel whose BFS expands most of the graph costs a sweep of it
    // however little of the heights the discharge has changed, so it must be
    // paid for by a longer discharge interval: the cost factor rises from 1
    // when the last BFS expanded three quarters of the vertices or fewer to 4
    // when it expanded all of them.
    auto affordable = [&]() -> bool {
        const double reach = std::min(1.0, last_gr_reach);
        const double factor = kRelabelCost * (1.0 + 12.0 * std::max(0.0, reach - 0.75));
        return std::chrono::duration<double, std::milli>(Clock::now() - last_gr_end).count() >= factor * last_gr_ms;
    };
    auto work_trigger = [&]() -> bool {
        return (alpha > 0.0 &&
                static_cast<double>(relabels_now - relabels_at_gr) > alpha * static_cast<double>(active_at_gr) + 1024.0 &&
                affordable()) ||
               // requests waiting for an arena join at the next affordable relabel
               (feeder && feeder->pending() && affordable());
    };
    auto check_device_error = [&](const Ctl& c) {
        if (c.error) {
            std::ostringstream oss;
            oss << kPrefix << "device invariant failed at discharge code=" << c.error;
            fail(WTV_EINVARIANT, oss.str());
        }
    };

    auto solve0 = Clock::now();
    // the wall-clock safety limit (gpu_push_relabel_fast.cu:538-544), checked
    // before every macro round the host decides on; the kernels still queued
    // finish before the failure unwinds (DrainOnExit)
    auto check_time = [&]() {
        const double el = std::chrono::duration<double>(Clock::now() - solve0).count();
        if (el > time_limit) {
            std::ostringstream oss;
            oss << kPrefix << "solve exceeded " << time_limit << "s safety limit";
            fail(WTV_ELIMIT, oss.str());
        }
    };
    uint64_t rounds0 = 0;  // an arena counts its limits per busy period
    for (;;) {
        if (!active) {
            if (!feeder) break;
            // an empty list: finish the resident graphs (all done) and admit
            // what waits, or wait for requests when nothing is resident
            if (!feeder->resident() && !feeder->wait_for_work()) break;
            if (!feeder->resident()) {
                solve0 = Clock::now();
                rounds0 = st.rounds;
            }
            do_global_relabel();
            continue;
        }
        if (st.rounds - rounds0 >= max_rounds) fail(WTV_ELIMIT, std::string(kPrefix) + "max macro rounds exceeded");
        check_time();
        if (active <= tail_cap) {
            // ---- one workgroup, sub-rounds back to back
            unsigned long long budget = ~0ull;
            if (alpha > 0.0) {
                double left = alpha * static_cast<double>(active_at_gr) + 1024.0 -
                              static_cast<double>(relabels_now - relabels_at_gr);
                if (left <= 0.0) {
                    if (affordable()) {
                        do_global_relabel();
                        continue;
                    }
                    left = alpha * static_cast<double>(active_at_gr) + 1024.0;  // one more interval first
                }
                budget = static_cast<unsigned long long>(left) + 1ull;
            }
            const uint64_t rounds_left = max_rounds - (st.rounds - rounds0);
            const uint32_t max_sub = static_cast<uint32_t>(
                std::min<uint64_t>(4096ull, rounds_left * static_cast<uint64_t>(local_steps)));
            const auto tail0 = Clock::now();
            const uint32_t tail_in = active;
            const uint64_t tail_rel0 = relabels_now;
            kt.begin(KernelTimer::kTail);
            if (small_mode == 0)
                k_discharge_tail<kB, true><<<1, kTailTPB, 0, s>>>(n, phase(pos), tail_cap, max_sub, budget, wl[0],
                                                                  wl[1], rl, nbr4, rslot, d_res, d_term, excess,
                                                                  height, gid, gn, gflow, ctl);
            else
                k_discharge_tail<kB, false><<<1, kTailTPB, 0, s>>>(n, phase(pos), tail_cap, max_sub, budget, wl[0],
                                                                   wl[1], rl, nbr4, rslot, d_res, d_term, excess,
                                                                   height, gid, gn, gflow, ctl);
            kt.end();
            ck(hipGetLastError(), "discharge tail");
            read_ctl("discharge tail");
            check_device_error(*hctl);
            const uint32_t ts = hctl->tail_sub;
            st.tail_sub_rounds += ts;
            const uint64_t before = st.rounds;
            sub += ts;
            pos += ts;
            st.rounds += std::max<uint64_t>(1ull, (ts + static_cast<uint64_t>(local_steps) - 1u) /
                                                      static_cast<uint64_t>(local_steps));
            relabels_now = stat_sum(*hctl, kStRelabels);
            active = hctl->lc[lslot(phase(pos))];
            if (trace)
                std::fprintf(stderr, "[wtv] tail: active %u -> %u, %u sub-rounds, %llu height increases, %.3f ms\n",
                             tail_in, active, ts, static_cast<unsigned long long>(relabels_now - tail_rel0),
                             std::chrono::duration<double, std::milli>(Clock::now() - tail0).count());
            const bool by_period = macro_period > 0 && st.rounds / static_cast<uint64_t>(macro_period) !=
                                                           before / static_cast<uint64_t>(macro_period);
            if (active && (by_period || work_trigger())) do_global_relabel();
            continue;
        }
        // ---- large active list: pipelined macro rounds
        known_active = active;
        launch_round();
        ++st.rounds;
        Pending prev = snapshot();
        for (;;) {
            // speculative next round (a no-op on the device if prev drained the list)
            const bool spec = st.rounds - rounds0 < max_rounds;
            Pending cur{};
            if (spec) {
                launch_round();
                cur = snapshot();
            }
            ev_wait(*prev.ev, "snapshot wait");
            if (ktp) ktp->harvest();
            const Ctl& S = snap[prev.slot];
            check_device_error(S);
            const uint32_t act_prev = S.lc[lslot(phase(prev.pos_after))];
            known_active = act_prev;
            relabels_now = stat_sum(S, kStRelabels);
            if (trace)
                std::fprintf(stderr, "[wtv] round %llu: active=%u relabels=%llu\n",
                             static_cast<unsigned long long>(st.rounds), act_prev,
                             stat_sum(S, kStRelabels));
            if (act_prev == 0u) {
                if (spec) ev_wait(*cur.ev, "drain");
                active = 0;
                break;
            }
            if (!spec) fail(WTV_ELIMIT, std::string(kPrefix) + "max macro rounds exceeded");
            check_time();
            ++st.rounds;
            const bool by_period = macro_period > 0 && (st.rounds % static_cast<uint64_t>(macro_period)) == 0ull;
            if (by_period || work_trigger()) {
                do_global_relabel();  // runs after the speculative round
                break;
            }
            if (act_prev <= tail_cap) {
                ev_wait(*cur.ev, "switch");
                const Ctl& C2 = snap[cur.slot];
                check_device_error(C2);
                active = C2.lc[lslot(phase(cur.pos_after))];
                relabels_now = stat_sum(C2, kStRelabels);
                break;
            }
            prev = cur;
        }
    }
    check_err("solve completion");
    st.solve_seconds = std::chrono::duration<double>(Clock::now() - solve0).count();
    if (feeder) return;  // the session stops: every graph was finished at a relabel

    const auto fr0 = Clock::now();
    global_relabel(false);  // every graph: the labels come from this BFS
    st.final_relabel_seconds = std::chrono::duration<double>(Clock::now() - fr0).count();

    uint8_t* d_part = partition_on_device ? partition : reinterpret_cast<uint8_t*>(rl);
    struct CallerOrder {  // the caller's stream waits for the solver's work, also on failure
        hipStream_t cs, s;
        hipEvent_t e;
        ~CallerOrder() {
            if (e && hipEventRecord(e, s) == hipSuccess) (void)hipStreamWaitEvent(cs, e, 0);
        }
    } order{cs, s, ev_done.e};  // without allocation s is the caller's stream
    k_partition<kB><<<sweep_grid, tpb, 0, s>>>(n, height, excess, d_part, gid, gn, ctl);
    ck(hipGetLastError(), "partition");
    ev_record(ev_t1);
    if (!partition_on_device)
        ck(hipMemcpyAsync(partition, d_part, static_cast<size_t>(n), hipMemcpyDeviceToHost, s), "partition D2H");
    read_ctl("final");
    if (hctl->error) {
        std::ostringstream oss;
        oss << kPrefix << "device invariant failed at final code=" << hctl->error;
        fail(WTV_EINVARIANT, oss.str());
    }
    if (hctl->violations) {
        std::ostringstream oss;
        oss << kPrefix << "device invariant failed at partition: " << hctl->violations
            << " vertices still active after the final BFS";
        fail(WTV_EINVARIANT, oss.str());
    }
    const unsigned long long sink_flow = stat_sum(*hctl, kStSinkFlow);
    if (sink_flow > static_cast<unsigned long long>(INT64_MAX))
        fail(WTV_EINVARIANT, std::string(kPrefix) + "flow exceeds int64");
    st.raw_flow = static_cast<int64_t>(sink_flow);
    if constexpr (kB) {
        std::vector<unsigned long long> f(G);
        ck(hipMemcpyAsync(f.data(), gflow, G * 8u, hipMemcpyDeviceToHost, s), "batch flows D2H");
        ck(hipStreamSynchronize(s), "batch flows");
        unsigned long long sum = 0;
        for (uint32_t g = 0; g < G; ++g) {
            B->h_flows[g] = static_cast<int64_t>(f[g]);
            sum += f[g];
        }
        if (sum != sink_flow) fail(WTV_EINVARIANT, std::string(kPrefix) + "per-graph flows do not add up");
    }
    st.visits = stat_sum(*hctl, kStVisits);
    st.pushes = stat_sum(*hctl, kStPushes);
    st.relabels = stat_sum(*hctl, kStRelabels);
    st.bfs_pops = stat_sum(*hctl, kStBfsPops);
    st.bfs_levels = hctl->bfs_levels;
    st.bfs_bu_levels = hctl->bu_levels;
    st.bfs_bu_checks = stat_sum(*hctl, kStBuChecks);
    st.bfs_seed_checks = stat_sum(*hctl, kStSeedChecks);
    st.sub_rounds = sub;
    float ms = 0.f;
    if (noalloc) ms = std::chrono::duration<float, std::milli>(Clock::now() - host_t0).count();
    else ck(hipEventElapsedTime(&ms, ev_t0.e, ev_t1.e), "event time");
    st.t_solve_ms = ms;
    st.t_bfs_ms = bfs_ms;
    kt.harvest();
    st.t_push_ms = kt.ms[KernelTimer::kPush];
    st.t_relabel_ms = kt.ms[KernelTimer::kRelabel];
    st.t_level_ms = kt.ms[KernelTimer::kLevel];
    st.n_push = kt.count[KernelTimer::kPush];
    st.n_relabel = kt.count[KernelTimer::kRelabel];
    st.n_level = kt.count[KernelTimer::kLevel];
    st.t_tail_ms = kt.ms[KernelTimer::kTail];
    st.n_tail = kt.count[KernelTimer::kTail];
    st.t_event_overhead_us = kt.on ? event_pair_overhead_us(grid, wg, s) : 0.0;
}

void pr_solve_in_order(uint32_t n, int32_t* d_nbr, uint32_t* d_res, int32_t* d_term, uint8_t* partition,
                       bool partition_on_device, const wtv_params& prm, wtv_stats& st, void* stream) {
    solve_t<false>(n, d_nbr, d_res, d_term, partition, partition_on_device, prm, st, stream, nullptr);
}

// ------------------------------------------------ allocation-free solve
// Workspace of wtv_pr_solve_noalloc: a file-scope device array, so that the
// call allocates nothing.  It holds the layout of an n-node solve for every
// n up to WTV_NOALLOC_MAX_NODES (control block, then 26 B per node: excess 8,
// height 4, two active lists and the relabel list 4 each, reverse slots 1,
// residual masks 1, each array 256-byte aligned).  Every device the library
// runs on holds its own copy, 27.3 MB, until the process ends.
constexpr size_t kNoallocMaxNodes = WTV_NOALLOC_MAX_NODES;
constexpr size_t kScratchBytes = (sizeof(Ctl) + 255u) / 256u * 256u + 26u * kNoallocMaxNodes;
static_assert(kNoallocMaxNodes % 256u == 0u, "per-node arrays stay 256-byte aligned");
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
// Device address of `scratch`, looked up once per device with that device
// current.  In host code the array's name is a host-side stand-in, not a
// device address: kernels and copies get this address instead.
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

void* scratch_of_current_device() {
    int dev = 0;
    ck(hipGetDevice(&dev), "hipGetDevice");
    if (dev < 0 || dev >= kMaxDevices) fail(WTV_EINVAL, std::string(kPrefix) + "device ordinal beyond 63");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        ck(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "scratch symbol address");
        if (!p) fail(WTV_EDEVICE, std::string(kPrefix) + "scratch symbol has no device address");
        g_scratch_ptr[dev] = p;
    }
    return g_scratch_ptr[dev];
}
}  // namespace

void pr_solve_noalloc(uint32_t n, int32_t* d_nbr, uint32_t* d_res, int32_t* d_term, uint8_t* partition,
                      bool partition_on_device, const wtv_params& prm, wtv_stats& st, void* stream) {
    std::memset(&st, 0, sizeof(st));
    if (n > kNoallocMaxNodes) {
        std::ostringstream oss;
        oss << kPrefix << "allocation-free solve: " << n << " nodes exceed its fixed workspace of "
            << kNoallocMaxNodes << " nodes";
        fail(WTV_EINVAL, oss.str());
    }
    if (layout_for(n, 0, false, false).total > kScratchBytes)
        fail(WTV_EINVARIANT, std::string(kPrefix) + "scratch layout");
    SolveExtras X;
    X.ws = scratch_of_current_device();
    X.noalloc = true;
    solve_t<false>(n, d_nbr, d_res, d_term, partition, partition_on_device, prm, st, stream, nullptr, X);
}

void batch_solve(uint32_t G, const uint64_t* h_off, const int32_t* d_nbr, uint32_t* d_res, int32_t* d_term,
                 uint8_t* partition, bool partition_on_device, int64_t* h_flows, const wtv_params& prm,
                 wtv_stats& st, void* stream) {
    std::memset(&st, 0, sizeof(st));
    if (!G) return;
    if (!h_off || !h_flows) fail(WTV_EINVAL, std::string(kPrefix) + "null batch pointer");
    if (h_off[0] != 0) fail(WTV_EINVAL, std::string(kPrefix) + "batch offsets must start at 0");
    for (uint32_t g = 0; g < G; ++g) {
        if (h_off[g + 1] < h_off[g]) fail(WTV_EINVAL, std::string(kPrefix) + "batch offsets must not decrease");
        h_flows[g] = 0;
    }
    const uint64_t total = h_off[G];
    if (total > 0x7FFFFFFFull) fail(WTV_EINVAL, std::string(kPrefix) + ">INT32_MAX reduced nodes unsupported");
    BatchArgs B{G, h_off, h_flows};
    solve_t<true>(static_cast<uint32_t>(total), d_nbr, d_res, d_term, partition, partition_on_device, prm, st,
                  stream, &B);
}


// ------------------------------------------------------------ serving sessions
namespace {

constexpr uint32_t kArenaAlign = 1024u;  // k_bfs_seed's tile: kTPB * kSeedUnroll
static_assert(kArenaAlign == kTPB * kSeedUnroll, "arena ranges align to the seed sweep's tiles");

inline uint32_t align_nodes(uint32_t n) { return (n + kArenaAlign - 1u) / kArenaAlign * kArenaAlign; }

// Continuous batching on one arena: requests join at global relabels, when
// they fit the free node ranges and graph slots, and each graph leaves at
// the first relabel after which it has no active vertex (that relabel's BFS
// ran on its final residual graph, so its partition is final): its
// partition, residuals and terminals go back to the caller's buffers and its
// node range becomes inert.  Graphs share no arcs and keep their own source
// heights, so each result equals a lone solve of the same graph.
class ArenaSession final : public Feeder {
  public:
    ArenaSession(SessionQueue& q, uint32_t capacity, uint32_t slots, int32_t worker)
        : q_(q), capacity_(capacity), slots_(slots), worker_(worker) {
        free_[0] = capacity;
        for (uint32_t g = slots; g > 0; --g) free_slots_.push_back(g - 1u);
    }

    void admit(ArenaView& a, uint32_t ph) override {
        a.hwm = hwm_;
        if (alone_) return;  // a graph solved alone admits no other
        std::vector<SessionJob> jobs = q_.take(largest_free(), static_cast<uint32_t>(free_slots_.size()),
                                               !res_.empty());
        if (jobs.empty()) return;
        try {
            admit_round(a, ph, jobs);
        } catch (...) {
            // the requests taken and not yet resident or finished are not
            // lost: back to the queue, to be solved alone
            std::vector<SessionJob> rest;
            for (const SessionJob& j : jobs)
                if (j.ticket) rest.push_back(j);
            if (!rest.empty()) q_.requeue_alone(rest);
            throw;
        }
        a.hwm = hwm_;
    }

    void finish(ArenaView& a, uint32_t ph) override {
        if (res_.empty()) return;
        const uint32_t G = slots_ + 1u;
        const int gg = std::max(1, static_cast<int>(std::min<uint32_t>((G + kTPB - 1u) / kTPB, 1024u)));
        k_clear_alive<<<gg, kTPB, 0, a.s>>>(G, a.galive);
        k_mark_alive<<<1024, kTPB, 0, a.s>>>(ph, a.wl[lbuf(ph)], a.gid, a.galive, a.ctl);
        k_fill_u8<<<gg, kTPB, 0, a.s>>>(a.gfresh, G, 0u);
        ck(hipGetLastError(), "arena alive marks");
        alive_.resize(G);
        ck(hipMemcpyAsync(alive_.data(), a.galive, G, hipMemcpyDeviceToHost, a.s), "alive marks D2H");
        ck(hipStreamSynchronize(a.s), "alive marks");
        std::vector<size_t> done;
        for (size_t i = 0; i < res_.size(); ++i)
            if (!alive_[res_[i].slot]) done.push_back(i);
        if (done.empty()) return;
        clear_violations(a);
        for (size_t i : done) {
            const Res& r = res_[i];
            const uint32_t ng = r.job.req.n;
            k_partition_range<<<grid_of(ng), kTPB, 0, a.s>>>(r.lo, ng, ng, a.height, a.excess,
                                                             r.job.req.d_partition, a.ctl);
            ck(hipMemcpyAsync(r.job.req.d_res, a.res + r.lo, static_cast<size_t>(ng) * 16u,
                              hipMemcpyDeviceToDevice, a.s), "results D2D");
            ck(hipMemcpyAsync(r.job.req.d_term, a.term + r.lo, static_cast<size_t>(ng) * 4u,
                              hipMemcpyDeviceToDevice, a.s), "results D2D");
        }
        ck(hipGetLastError(), "arena partition");
        flows_.resize(G);
        ck(hipMemcpyAsync(flows_.data(), a.gflow, static_cast<size_t>(G) * 8u, hipMemcpyDeviceToHost, a.s),
           "flows D2H");
        uint32_t viol = 0;
        ck(hipMemcpyAsync(&viol, &a.ctl->violations, 4u, hipMemcpyDeviceToHost, a.s), "violations D2H");
        ck(hipStreamSynchronize(a.s), "arena results");
        const double t = q_.now();
        for (size_t k = done.size(); k-- > 0;) {
            const Res r = res_[done[k]];
            const int status = viol ? WTV_EINVARIANT : WTV_OK;
            q_.finish(r.job.ticket, status, static_cast<int64_t>(flows_[r.slot]), r.batch, worker_, r.start, t);
            release(a, r);
            res_.erase(res_.begin() + static_cast<std::ptrdiff_t>(done[k]));
        }
        if (res_.empty()) alone_ = false;
        a.hwm = hwm_;
    }

    bool pending() override {
        return !alone_ && q_.waiting(largest_free(), static_cast<uint32_t>(free_slots_.size()), !res_.empty());
    }
    bool wait_for_work() override { return q_.wait_nonempty(); }
    bool resident() override { return !res_.empty(); }

    // A shared solve failed: every resident request goes back to the queue to
    // be solved alone (a request already alone fails with the error).
    void abort(int code) {
        std::vector<SessionJob> again;
        const double t = q_.now();
        for (const Res& r : res_) {
            if (r.job.alone) q_.finish(r.job.ticket, code, 0, r.batch, worker_, r.start, t);
            else again.push_back(r.job);
        }
        for (SessionJob& j : again) j.alone = true;
        if (!again.empty()) q_.requeue_alone(again);
        res_.clear();
        free_.clear();
        free_[0] = capacity_;
        free_slots_.clear();
        for (uint32_t g = slots_; g > 0; --g) free_slots_.push_back(g - 1u);
        hwm_ = 0;
        alone_ = false;
    }

  private:
    struct Res {
        SessionJob job;
        uint32_t slot, lo, ag;
        int32_t batch;  // graphs resident once it joined, itself included
        double start;
    };

    static int grid_of(uint32_t m) {
        return std::max(1, static_cast<int>(std::min<uint64_t>((static_cast<uint64_t>(m) + kTPB - 1) / kTPB, 8192)));
    }

    uint64_t largest_free() const {
        uint64_t best = 0;
        for (const auto& f : free_) best = std::max<uint64_t>(best, f.second);
        return best;
    }

    // first fit of `ag` nodes; the caller checked it fits
    uint32_t carve(uint32_t ag) {
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < ag) continue;
            const uint32_t lo = it->first, len = it->second;
            free_.erase(it);
            if (len > ag) free_[lo + ag] = len - ag;
            return lo;
        }
        fail(WTV_EINVARIANT, std::string(kPrefix) + "arena range lost");
    }

    void clear_violations(ArenaView& a) { ck(hipMemsetAsync(&a.ctl->violations, 0, 4u, a.s), "violations clear"); }

    int32_t read_error(ArenaView& a, const char* where) {
        int32_t err = 0;
        ck(hipMemcpyAsync(&err, &a.ctl->error, 4u, hipMemcpyDeviceToHost, a.s), where);
        ck(hipStreamSynchronize(a.s), where);
        return err;
    }

    // One admission round: every taken request is copied in and validated
    // (ids in range, reverse arcs) into its slot's error word; one read of
    // the words then fails the bad requests alone, before any of their
    // vertices joins the active list, and the rest enter the solve.  A job
    // whose ticket is set to 0 has been finished or made resident.
    void admit_round(ArenaView& a, uint32_t ph, std::vector<SessionJob>& jobs) {
        // an error of the discharge so far belongs to the shared solve, not
        // to these requests: the solve fails and its graphs are solved alone
        if (const int32_t e0 = read_error(a, "admit check")) {
            std::ostringstream oss;
            oss << kPrefix << "device invariant failed at discharge code=" << e0;
            fail(WTV_EINVARIANT, oss.str());
        }
        struct Placed {
            size_t job;
            Res r;
        };
        std::vector<Placed> placed;
        const double t = q_.now();
        for (size_t i = 0; i < jobs.size(); ++i) {
            SessionJob& j = jobs[i];
            const uint32_t ng = j.req.n;
            if (ng == 0) {
                q_.finish(j.ticket, WTV_OK, 0, static_cast<int32_t>(res_.size() + placed.size()) + 1, worker_, t, t);
                j.ticket = 0;
                continue;
            }
            const uint32_t ag = align_nodes(ng);
            const uint32_t lo = carve(ag);
            const uint32_t g = free_slots_.back();
            free_slots_.pop_back();
            Res r{j, g, lo, ag, static_cast<int32_t>(res_.size() + placed.size()) + 1, t};
            placed.push_back(Placed{i, r});
            hwm_ = std::max(hwm_, lo + ag);
            if (j.ready) ck(hipStreamWaitEvent(a.s, static_cast<hipEvent_t>(j.ready), 0), "ready wait");
            k_set_slot<<<1, 1, 0, a.s>>>(g, ng, a.gn, a.gflow, a.gres, a.gfresh, a.gerr, 1u);
            k_admit<<<grid_of(ag), kTPB, 0, a.s>>>(lo, ng, ag, g, reinterpret_cast<const int4*>(j.req.d_nbr), a.nbrg,
                                                   a.res, a.term, a.gid, a.gerr);
            ck(hipMemcpyAsync(a.res + lo, j.req.d_res, static_cast<size_t>(ng) * 16u, hipMemcpyDeviceToDevice, a.s),
               "admit res");
            ck(hipMemcpyAsync(a.term + lo, j.req.d_term, static_cast<size_t>(ng) * 4u, hipMemcpyDeviceToDevice, a.s),
               "admit term");
            k_prepare<<<grid_of(ag), kTPB, 0, a.s>>>(lo, lo + ag, a.nbrg, a.term, a.rslot, a.excess, nullptr, nullptr,
                                                     a.ctl, a.gerr + g);
            ck(hipGetLastError(), "admit");
        }
        if (placed.empty()) return;
        const uint32_t G = slots_ + 1u;
        errs_.resize(G);
        ck(hipMemcpyAsync(errs_.data(), a.gerr, static_cast<size_t>(G) * 4u, hipMemcpyDeviceToHost, a.s),
           "admit errors D2H");
        ck(hipStreamSynchronize(a.s), "admit errors");
        for (const Placed& pl : placed) {
            const Res& r = pl.r;
            jobs[pl.job].ticket = 0;
            if (errs_[r.slot]) {
                q_.finish(r.job.ticket, WTV_EINVARIANT, 0, r.batch, worker_, r.start, q_.now());
                release(a, r);
                continue;
            }
            k_prepare<<<grid_of(r.ag), kTPB, 0, a.s>>>(r.lo, r.lo + r.ag, a.nbrg, a.term, a.rslot, a.excess,
                                                       a.wl[lbuf(ph)], &a.ctl->lc[lslot(ph)], a.ctl, a.gerr + r.slot);
            ck(hipGetLastError(), "admit supply");
            res_.push_back(r);
            if (r.job.alone) alone_ = true;
        }
        // a release above recomputed the bound before the later graphs joined
        hwm_ = 0;
        for (const Res& x : res_) hwm_ = std::max(hwm_, x.lo + x.ag);
    }

    // the graph's nodes become inert and its range and slot free
    void release(ArenaView& a, const Res& r) {
        k_inert<<<grid_of(r.ag), kTPB, 0, a.s>>>(r.lo, r.lo + r.ag, a.nbrg, a.term, a.excess, a.height, a.gid,
                                                 slots_);
        k_set_slot<<<1, 1, 0, a.s>>>(r.slot, 0u, a.gn, a.gflow, a.gres, a.gfresh, a.gerr, 0u);
        ck(hipGetLastError(), "release");
        free_slots_.push_back(r.slot);
        // merge the range with its free neighbours
        uint32_t lo = r.lo, len = r.ag;
        auto nx = free_.lower_bound(lo);
        if (nx != free_.end() && nx->first == lo + len) {
            len += nx->second;
            nx = free_.erase(nx);
        }
        if (nx != free_.begin()) {
            auto pv = std::prev(nx);
            if (pv->first + pv->second == lo) {
                lo = pv->first;
                len += pv->second;
                free_.erase(pv);
            }
        }
        free_[lo] = len;
        hwm_ = 0;
        for (const Res& x : res_)
            if (x.slot != r.slot) hwm_ = std::max(hwm_, x.lo + x.ag);
    }

    SessionQueue& q_;
    const uint32_t capacity_, slots_;
    const int32_t worker_;
    std::map<uint32_t, uint32_t> free_;  // free node ranges: start -> length
    std::vector<uint32_t> free_slots_;
    std::vector<Res> res_;
    std::vector<uint8_t> alive_;
    std::vector<int32_t> errs_;
    std::vector<unsigned long long> flows_;
    uint32_t hwm_ = 0;
    bool alone_ = false;
};

}  // namespace

bool pull_enabled(const wtv_params& p) { return p.pull_min > 0 || (p.pull_min == 0 && p.fused < 0); }

uint64_t arena_bytes_per_node(bool pull) {
    // rows 36 B, workspace 26 B (43 B with the pull-mode arrays), graph id 4 B,
    // union ids 16 B ... as layout_for
    return layout_for(1u << 20, 2u, true, pull).total / (1u << 20);
}

void session_serve(SessionQueue& q, const wtv_params& prm, uint32_t capacity, uint32_t slots, void* stream,
                   int32_t worker) {
    capacity = std::max(kArenaAlign, capacity / kArenaAlign * kArenaAlign);
    const Layout L = layout_for(capacity, slots + 1u, true, pull_enabled(prm));
    DeviceBuf ws;
    ck(hipMalloc(&ws.p, L.total), "arena malloc");
    ArenaSession sess(q, capacity, slots, worker);
    SolveExtras X;
    X.ws = ws.p;
    X.feeder = &sess;
    X.slots = slots;
    for (;;) {
        wtv_stats st{};
        try {
            solve_t<true>(capacity, nullptr, nullptr, nullptr, nullptr, true, prm, st, stream, nullptr, X);
            return;  // the queue stopped with nothing resident
        } catch (const Error& e) {
            sess.abort(e.code);
        }
    }
}
}  // namespace wtv
