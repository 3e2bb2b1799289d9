/*
 * wtv_oracle.c -- CPU restatement of the WTiVo graph-cut hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and
 * never as the thing that is measured or shipped.  The product path is the
 * HIP library under wtivo-watertightvoxel_amd/csrc and fails loudly when that
 * library is missing.
 *
 * Contents (each function cites the reference code it restates):
 *   wtvo_capacity      native/gpupr/wtivo_gpupr_bindings.cpp:65-98
 *   wtvo_reduce        native/gpupr/wtivo_gpupr_bindings.cpp:139-318
 *   wtvo_scatter       native/gpupr/wtivo_gpupr_bindings.cpp:357-364
 *   wtvo_partition     native/gpupr/gpu_push_relabel_fast.cu:85-147,315-324,360-402
 *   wtvo_pr_solve      native/gpupr/gpu_push_relabel_fast.cu:69-324,406-571
 *                      (the same queue-round schedule, executed sequentially)
 *   wtvo_bk_solve      Boykov & Kolmogorov, "An experimental comparison of
 *                      min-cut/max-flow algorithms for energy minimization in
 *                      vision", IEEE PAMI 26(9), 2004 -- the standard CPU
 *                      graph-cut solver, used as the CPU baseline.
 *   wtvo_certificate   max-flow/min-cut certificate (no reference counterpart;
 *                      the reference ships no checker).
 *   wtvo_flow_certificate  the same for a maximum flow (conservation at every
 *                      vertex), as left by the excess return; pinned by
 *                      Boykov-Kolmogorov's results, which are flows.
 *   wtvo_sample_labels native/vdb/wtivo_vdb.cpp:1519-1584 with OpenVDB's
 *                      BoxSampler (trilinear, openvdb/tools/Interpolation.h)
 *                      restated from its published source; OpenVDB is not in
 *                      this image, so parity with OpenVDB itself is pinned
 *                      only through the closed forms in tests/test_oracle.py
 *                      (constant and linear fields, voxel-centre samples,
 *                      background outside leaves).
 *
 * Pinning (what makes these functions trustworthy, see tests/test_oracle.py):
 *   capacity  - hand-computed known answers (right-triangle faces), the
 *               overflow throw at 2^30, degenerate faces.
 *   reduce    - invariants (free+src+sink = N, even refs, terminal_edges +
 *               free_refs = 4n, monotone reduced ids, base = sum min) and a
 *               pure-Python brute force on tiny meshes.
 *   max-flow  - scipy.sparse.csgraph.maximum_flow (independent implementation)
 *               on random graphs; pr_solve == bk_solve on every instance.
 *   partition - 2^n enumeration of all cuts on graphs with n <= 14 free nodes:
 *               the labels are the unique minimal sink set of the min cut.
 *   scatter   - literal label arrays worked out by hand (labels 0, 1, 7 mixed,
 *               free tets from the classification of a 3^3 Kuhn mesh); a scatter
 *               indexed by reduced id, unnormalised bytes or a zero start fail.
 *   surface   - a three-tet mesh whose faces and vertices are written out by
 *               hand: tet-major / slot-minor order, orientation, vertices in
 *               ascending original id; slot-major order, first-use compaction
 *               or unoriented faces fail.  Closed-surface invariants on Kuhn
 *               blobs (each directed edge once, enclosed volume).
 *
 * Graph representation (the reduced graph of the reference):
 *   n nodes; nbr[4v+k] = neighbour id or -1; res[4v+k] = residual capacity of
 *   arc v->nbr[4v+k] (uint32); term[v] > 0 = source supply, term[v] < 0 =
 *   remaining sink capacity (gpu_push_relabel_fast.cu:79-80,221-235).
 *
 * Build: gcc -O3 -march=x86-64-v3 -ffp-contract=off -fPIC -shared (oracle/Makefile).
 */
#define _POSIX_C_SOURCE 199309L
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

enum {
    WTVO_OK = 0,
    WTVO_EINVAL = 1,     /* invalid argument / graph (runtime_error)        */
    WTVO_EOVERFLOW = 2,  /* capacity outside [0,2^30) (overflow_error)      */
    WTVO_EINVARIANT = 3, /* internal invariant (missing reverse arc, ...)   */
    WTVO_ELIMIT = 5,     /* round limit                                     */
    WTVO_ENOMEM = 6
};

/* ------------------------------------------------------------------------ */
/* Capacity: wtivo_gpupr_bindings.cpp:65-98                                  */
/* ------------------------------------------------------------------------ */
static const int k_face_lut[4][3] = {{0, 1, 2}, {0, 1, 3}, {0, 2, 3}, {1, 2, 3}};

/* Returns the capacity; *status = 0 ok, 1 invalid vertex, 2 overflow. */
int32_t wtvo_capacity(const double* V, size_t nv, const int32_t* tet4, int fi,
                      uint8_t label_i, uint8_t label_j, double fill_T,
                      int* status)
{
    const int a = tet4[k_face_lut[fi][0]];
    const int b = tet4[k_face_lut[fi][1]];
    const int c = tet4[k_face_lut[fi][2]];
    *status = 0;
    if (a < 0 || b < 0 || c < 0 || (size_t)a >= nv || (size_t)b >= nv ||
        (size_t)c >= nv) {
        *status = 1;
        return 0;
    }
    const double* p0 = V + (size_t)a * 3u;
    const double* p1 = V + (size_t)b * 3u;
    const double* p2 = V + (size_t)c * 3u;
    const double x1 = p1[0] - p0[0], y1 = p1[1] - p0[1], z1 = p1[2] - p0[2];
    const double x2 = p2[0] - p0[0], y2 = p2[1] - p0[1], z2 = p2[2] - p0[2];
    const double cx = y1 * z2 - z1 * y2;
    const double cy = z1 * x2 - x1 * z2;
    const double cz = x1 * y2 - y1 * x2;
    /* FP64 product, correctly rounded sqrt, one rounding to FP32. */
    const float area_f = (float)(1.0E7 * 0.5 * sqrt(cx * cx + cy * cy + cz * cz));
    const double area = (double)area_f;
    const double pen = (label_i != label_j) ? area : area * (1.0 + fill_T);
    /* C cast truncates toward zero, like static_cast<int64_t>. Guard the
       conversion itself against values outside int64 (UB in C). */
    if (!(pen > -9.2e18 && pen < 9.2e18)) {
        *status = 2;
        return 0;
    }
    const int64_t cap64 = (int64_t)pen;
    if (cap64 < 0 || cap64 >= 1073741824LL) {
        *status = 2;
        return 0;
    }
    return (int32_t)cap64;
}

/* ------------------------------------------------------------------------ */
/* Reduction: classify, reduced ids, capacities, terminal folding            */
/* wtivo_gpupr_bindings.cpp:139-188 (classify, ids), :227-318 (build)        */
/* ------------------------------------------------------------------------ */
typedef struct wtvo_reduce_info {
    uint64_t free_n;
    uint64_t fixed_source;
    uint64_t fixed_sink;
    uint64_t full_refs;
    uint64_t free_refs;
    uint64_t terminal_edges;
    int64_t base_flow;
} wtvo_reduce_info;

/*
 * Caller provides full_to_free[N], free_old[N], nbr_r[4N], res[4N], term[N]
 * (only the first free_n / 4*free_n entries are written).
 *
 * Like the reference, classify and build run over contiguous ranges of tets
 * on `threads` threads with per-range counters summed afterwards, and the
 * reduced ids are assigned in one serial pass in tet order (bindings.cpp:
 * 147-188 classify and ids, :253-294 build), so every output is the same for
 * any thread count.  An error is reported for the first range (in tet
 * order) that hits one.
 */
typedef struct reduce_job {
    size_t lo, hi;
    const double* V;
    size_t nv;
    const int32_t* tets;
    const int32_t* nbrs;
    const uint8_t* labels;
    double fill_T;
    size_t N;
    int32_t* full_to_free;
    int32_t* nbr_r;
    uint32_t* res;
    int32_t* term;
    wtvo_reduce_info cnt; /* this range's counters */
    int rc;
} reduce_job;

/* classify (bindings.cpp:153-171) */
static void* reduce_classify(void* arg)
{
    reduce_job* J = (reduce_job*)arg;
    for (size_t i = J->lo; i < J->hi; ++i) {
        int exterior = 0;
        for (int fi = 0; fi < 4; ++fi) {
            const int32_t j = J->nbrs[i * 4u + (size_t)fi];
            if (j < 0)
                exterior = 1;
            else {
                ++J->cnt.full_refs;
                if ((size_t)j >= J->N) {
                    J->rc = WTVO_EINVAL; /* neighbor out of range */
                    return NULL;
                }
            }
        }
        if (J->labels[i] == 0u) {
            J->full_to_free[i] = -2;
            ++J->cnt.fixed_source;
        } else if (exterior) {
            J->full_to_free[i] = -3;
            ++J->cnt.fixed_sink;
        } else {
            J->full_to_free[i] = -1;
            ++J->cnt.free_n;
        }
    }
    return NULL;
}

/* build (bindings.cpp:253-294) */
static void* reduce_build(void* arg)
{
    reduce_job* J = (reduce_job*)arg;
    for (size_t i = J->lo; i < J->hi; ++i) {
        const int32_t rr = J->full_to_free[i];
        if (rr < 0) continue;
        const size_t local = (size_t)rr;
        int64_t src = 0, snk = 0;
        const int32_t* tr = J->tets + i * 4u;
        for (int fi = 0; fi < 4; ++fi) {
            J->nbr_r[local * 4u + (size_t)fi] = -1;
            J->res[local * 4u + (size_t)fi] = 0u;
        }
        for (int fi = 0; fi < 4; ++fi) {
            const int32_t jj = J->nbrs[i * 4u + (size_t)fi];
            if (jj < 0) { J->rc = WTVO_EINVAL; return NULL; } /* free tets have no hull face */
            const size_t j = (size_t)jj;
            const int32_t m = J->full_to_free[j];
            int st = 0;
            const int32_t cap = wtvo_capacity(J->V, J->nv, tr, fi, J->labels[i], J->labels[j], J->fill_T, &st);
            if (st == 1) { J->rc = WTVO_EINVAL; return NULL; }
            if (st == 2) { J->rc = WTVO_EOVERFLOW; return NULL; }
            if (m >= 0) {
                J->nbr_r[local * 4u + (size_t)fi] = m;
                J->res[local * 4u + (size_t)fi] = (uint32_t)cap;
                ++J->cnt.free_refs;
            } else if (m == -2) {
                src += cap;
                ++J->cnt.terminal_edges;
            } else if (m == -3) {
                snk += cap;
                ++J->cnt.terminal_edges;
            } else {
                J->rc = WTVO_EINVAL;
                return NULL;
            }
        }
        J->cnt.base_flow += (src < snk) ? src : snk;
        const int64_t diff = src - snk;
        if (diff < -2147483648LL || diff > 2147483647LL) { J->rc = WTVO_EINVAL; return NULL; }
        J->term[local] = (int32_t)diff;
    }
    return NULL;
}

/* Runs fn over the jobs, jobs[1..] on threads of their own; the first
 * nonzero rc in range order. */
static int run_ranges(reduce_job* jobs, int nj, void* (*fn)(void*))
{
    pthread_t th[256];
    int started = 0;
    for (int t = 1; t < nj; ++t) {
        if (pthread_create(&th[t], NULL, fn, &jobs[t]) != 0) break;
        started = t;
    }
    fn(&jobs[0]);
    for (int t = started + 1; t < nj; ++t) fn(&jobs[t]); /* any thread that did not start */
    for (int t = 1; t <= started; ++t) pthread_join(th[t], NULL);
    for (int t = 0; t < nj; ++t)
        if (jobs[t].rc) return jobs[t].rc;
    return WTVO_OK;
}

int wtvo_reduce_mt(size_t N, const double* V, size_t nv, const int32_t* tets,
                   const int32_t* nbrs, const uint8_t* labels, double fill_T,
                   int32_t* full_to_free, uint32_t* free_old, int32_t* nbr_r,
                   uint32_t* res, int32_t* term, wtvo_reduce_info* info, int threads)
{
    memset(info, 0, sizeof(*info));
    if (threads < 1) threads = 1;
    if (threads > 256) threads = 256;
    if ((size_t)threads > N / 4096u + 1u) threads = (int)(N / 4096u + 1u);
    reduce_job jobs[256];
    for (int t = 0; t < threads; ++t) {
        reduce_job* J = &jobs[t];
        memset(J, 0, sizeof(*J));
        J->lo = N * (size_t)t / (size_t)threads;
        J->hi = N * (size_t)(t + 1) / (size_t)threads;
        J->V = V; J->nv = nv; J->tets = tets; J->nbrs = nbrs; J->labels = labels; J->fill_T = fill_T;
        J->N = N; J->full_to_free = full_to_free; J->nbr_r = nbr_r; J->res = res; J->term = term;
    }
    int rc = run_ranges(jobs, threads, reduce_classify);
    if (rc) return rc;
    for (int t = 0; t < threads; ++t) {
        info->full_refs += jobs[t].cnt.full_refs;
        info->fixed_source += jobs[t].cnt.fixed_source;
        info->fixed_sink += jobs[t].cnt.fixed_sink;
        info->free_n += jobs[t].cnt.free_n;
    }
    if (info->full_refs & 1u) return WTVO_EINVAL; /* count not even (:177) */
    /* reduced ids in increasing tet order (:179-188) */
    uint64_t rid = 0;
    for (size_t i = 0; i < N; ++i) {
        if (full_to_free[i] == -1) {
            full_to_free[i] = (int32_t)rid;
            free_old[rid] = (uint32_t)i;
            ++rid;
        }
    }
    if (info->free_n > 2147483647ull) return WTVO_EINVAL;
    rc = run_ranges(jobs, threads, reduce_build);
    if (rc) return rc;
    for (int t = 0; t < threads; ++t) {
        info->free_refs += jobs[t].cnt.free_refs;
        info->terminal_edges += jobs[t].cnt.terminal_edges;
        info->base_flow += jobs[t].cnt.base_flow;
    }
    if (info->free_refs & 1ull) return WTVO_EINVAL;
    return WTVO_OK;
}

int wtvo_reduce(size_t N, const double* V, size_t nv, const int32_t* tets,
                const int32_t* nbrs, const uint8_t* labels, double fill_T,
                int32_t* full_to_free, uint32_t* free_old, int32_t* nbr_r,
                uint32_t* res, int32_t* term, wtvo_reduce_info* info)
{
    return wtvo_reduce_mt(N, V, nv, tets, nbrs, labels, fill_T, full_to_free, free_old, nbr_r, res, term, info, 1);
}

/* Label reconstruction (bindings.cpp:357-364): out = labels; out[free_old[q]] = part[q]. */
void wtvo_scatter(size_t N, const uint8_t* labels, uint64_t free_n,
                  const uint32_t* free_old, const uint8_t* partition, uint8_t* out)
{
    memcpy(out, labels, N);
    for (uint64_t q = 0; q < free_n; ++q) out[free_old[q]] = partition[q] ? 1u : 0u;
}

/* ------------------------------------------------------------------------ */
/* Residual BFS from the sink (global relabel + partition)                   */
/* gpu_push_relabel_fast.cu:85-147 (seed, expand, normalize), :360-402,      */
/* :315-324 (partition = dist < n+1).                                        */
/* ------------------------------------------------------------------------ */
static int reverse_slot(const int32_t* nbr, uint32_t u, uint32_t v)
{
    for (int k = 0; k < 4; ++k)
        if (nbr[(size_t)u * 4u + (size_t)k] == (int32_t)v) return k;
    return -1;
}

/* height[v] = BFS distance to the sink over residual arcs, n+1 if unreached.
   Returns number of vertex expansions (frontier pops) or -1 on a missing
   reverse slot. */
static int64_t bfs_heights(uint32_t n, const int32_t* nbr, const uint32_t* res,
                           const int32_t* term, uint32_t* height, uint32_t* queue)
{
    const uint32_t sent = 0xFFFFFFFFu;
    size_t head = 0, tail = 0;
    for (uint32_t v = 0; v < n; ++v) {
        height[v] = sent;
        if (term[v] < 0) {
            height[v] = 1u;
            queue[tail++] = v;
        }
    }
    int64_t pops = 0;
    while (head < tail) {
        const uint32_t u = queue[head++];
        ++pops;
        const uint32_t hu = height[u];
        for (int ku = 0; ku < 4; ++ku) {
            const int32_t vv = nbr[(size_t)u * 4u + (size_t)ku];
            if (vv < 0) continue;
            const uint32_t v = (uint32_t)vv;
            if (height[v] != sent) continue;
            const int kv = reverse_slot(nbr, v, u);
            if (kv < 0) return -1;
            if (res[(size_t)v * 4u + (size_t)kv] == 0u) continue;
            height[v] = hu + 1u;
            queue[tail++] = v;
        }
    }
    for (uint32_t v = 0; v < n; ++v)
        if (height[v] == sent) height[v] = n + 1u;
    return pops;
}

int wtvo_partition(uint32_t n, const int32_t* nbr, const uint32_t* res,
                   const int32_t* term, uint8_t* partition)
{
    uint32_t* height = (uint32_t*)malloc((size_t)n * 4u + 4u);
    uint32_t* queue = (uint32_t*)malloc((size_t)n * 4u + 4u);
    if (!height || !queue) {
        free(height);
        free(queue);
        return WTVO_ENOMEM;
    }
    const int64_t pops = bfs_heights(n, nbr, res, term, height, queue);
    for (uint32_t v = 0; v < n; ++v) partition[v] = (height[v] < n + 1u) ? 1u : 0u;
    free(height);
    free(queue);
    return pops < 0 ? WTVO_EINVARIANT : WTVO_OK;
}

/* ------------------------------------------------------------------------ */
/* Push-relabel Phase I with the reference's queue-round schedule, executed  */
/* sequentially: gpu_push_relabel_fast.cu:406-571 (driver), :69-83 (init),   */
/* :149-166 (active set), :185-313 (discharge), :360-402 (global relabel).   */
/* ------------------------------------------------------------------------ */
typedef struct wtvo_stats {
    double solve_seconds;
    double final_relabel_seconds;
    uint64_t rounds;
    uint64_t global_relabels;
    int64_t raw_flow;
    uint64_t visits;   /* discharge steps (vertex-visits) */
    uint64_t pushes;   /* successful pushes (arc or sink) */
    uint64_t relabels;
    uint64_t bfs_pops; /* BFS vertex expansions over all global relabels */
} wtvo_stats;

static double now_s(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

int wtvo_pr_solve(uint32_t n, const int32_t* nbr, uint32_t* res, int32_t* term,
                  uint8_t* partition, int period, uint64_t max_rounds,
                  int local_steps, wtvo_stats* st)
{
    memset(st, 0, sizeof(*st));
    if (!n) return WTVO_OK;
    /* clamps (.cu:421-424) */
    if (local_steps < 1) local_steps = 1;
    if (local_steps > 32) local_steps = 32;
    if (period < 0) period = 0;
    if (!max_rounds) max_rounds = 2000000ull;
    const uint32_t source_h = n, inf_h = n + 1u;
    const size_t N = n;
    uint32_t* height = (uint32_t*)malloc(N * 4u);
    int64_t* excess = (int64_t*)malloc(N * 8u);
    uint32_t* cur = (uint32_t*)malloc(N * 4u);
    uint32_t* nxt = (uint32_t*)malloc(N * 4u);
    uint8_t* queued = (uint8_t*)calloc(N, 1u);
    uint32_t* bfsq = (uint32_t*)malloc(N * 4u);
    int rc = WTVO_OK;
    if (!height || !excess || !cur || !nxt || !queued || !bfsq) {
        rc = WTVO_ENOMEM;
        goto done;
    }
    for (size_t v = 0; v < N; ++v) {
        excess[v] = term[v] > 0 ? (int64_t)term[v] : 0; /* init_preflow */
        height[v] = inf_h;
    }
    {
        int64_t pops = bfs_heights(n, nbr, res, term, height, bfsq);
        if (pops < 0) { rc = WTVO_EINVARIANT; goto done; }
        st->bfs_pops += (uint64_t)pops;
        ++st->global_relabels;
    }
    size_t ncur = 0;
    for (size_t v = 0; v < N; ++v) {
        queued[v] = 0;
        if (excess[v] > 0 && height[v] < source_h) {
            queued[v] = 1;
            cur[ncur++] = (uint32_t)v;
        }
    }
    int macro = 0;
    if (period > 0) {
        macro = (period + local_steps - 1) / local_steps;
        if (macro < 1) macro = 1;
    }
    uint64_t sink_flow = 0;
    const double t0 = now_s();
    while (ncur) {
        if (st->rounds >= max_rounds) { rc = WTVO_ELIMIT; goto done; }
        size_t nnext = 0;
        for (size_t qi = 0; qi < ncur; ++qi) {
            const uint32_t v = cur[qi];
            queued[v] = 0;
            const size_t vb = (size_t)v * 4u;
            for (int step = 0; step < local_steps; ++step) {
                int64_t ex = excess[v];
                if (ex <= 0) break;
                const uint32_t hv = height[v];
                if (hv >= source_h) break;
                ++st->visits;
                int32_t tv = term[v];
                if (tv < 0 && hv == 1u) { /* push to sink (.cu:222-236) */
                    const int64_t cap = -(int64_t)tv;
                    const int64_t d = ex < cap ? ex : cap;
                    if (d) {
                        term[v] = (int32_t)((int64_t)tv + d);
                        excess[v] -= d;
                        sink_flow += (uint64_t)d;
                        ex -= d;
                        ++st->pushes;
                    }
                }
                for (int k = 0; k < 4 && ex > 0; ++k) { /* arc pushes (.cu:240-269) */
                    const int32_t uu = nbr[vb + (size_t)k];
                    if (uu < 0) continue;
                    const uint32_t u = (uint32_t)uu;
                    if (hv != height[u] + 1u) continue;
                    const uint32_t r = res[vb + (size_t)k];
                    if (!r) continue;
                    const uint32_t d = (uint32_t)((uint64_t)r < (uint64_t)ex ? (uint64_t)r : (uint64_t)ex);
                    const int rk = reverse_slot(nbr, u, v);
                    if (rk < 0) { rc = WTVO_EINVARIANT; goto done; }
                    res[vb + (size_t)k] -= d;
                    res[(size_t)u * 4u + (size_t)rk] += d;
                    excess[v] -= d;
                    const int64_t old_u = excess[u];
                    excess[u] += d;
                    ex -= d;
                    ++st->pushes;
                    if (old_u == 0 && !queued[u]) {
                        queued[u] = 1;
                        nxt[nnext++] = u;
                    }
                }
                ex = excess[v];
                if (ex <= 0) break;
                /* relabel (.cu:275-302) */
                uint32_t min_h = 0xFFFFFFFFu;
                tv = term[v];
                if (tv < 0) min_h = 0u;
                if (tv > 0 && source_h < min_h) min_h = source_h;
                for (int k = 0; k < 4; ++k) {
                    const int32_t uu = nbr[vb + (size_t)k];
                    if (uu < 0) continue;
                    if (!res[vb + (size_t)k]) continue;
                    const uint32_t hu = height[(uint32_t)uu];
                    if (hu < min_h) min_h = hu;
                }
                if (min_h == 0xFFFFFFFFu) { rc = WTVO_EINVARIANT; goto done; }
                const uint32_t new_h = min_h + 1u;
                if (new_h > hv) {
                    height[v] = new_h;
                    ++st->relabels;
                    continue;
                }
                break;
            }
            if (excess[v] > 0 && height[v] < source_h && !queued[v]) {
                queued[v] = 1;
                nxt[nnext++] = v;
            }
        }
        ++st->rounds;
        { uint32_t* t = cur; cur = nxt; nxt = t; }
        ncur = nnext;
        if (macro > 0 && ncur && (st->rounds % (uint64_t)macro) == 0ull) {
            int64_t pops = bfs_heights(n, nbr, res, term, height, bfsq);
            if (pops < 0) { rc = WTVO_EINVARIANT; goto done; }
            st->bfs_pops += (uint64_t)pops;
            ++st->global_relabels;
            ncur = 0;
            for (size_t v = 0; v < N; ++v) {
                queued[v] = 0;
                if (excess[v] > 0 && height[v] < source_h) {
                    queued[v] = 1;
                    cur[ncur++] = (uint32_t)v;
                }
            }
        }
    }
    st->solve_seconds = now_s() - t0;
    {
        const double f0 = now_s();
        int64_t pops = bfs_heights(n, nbr, res, term, height, bfsq);
        if (pops < 0) { rc = WTVO_EINVARIANT; goto done; }
        st->bfs_pops += (uint64_t)pops;
        ++st->global_relabels;
        st->final_relabel_seconds = now_s() - f0;
    }
    for (size_t v = 0; v < N; ++v) partition[v] = height[v] < inf_h ? 1u : 0u;
    if (sink_flow > (uint64_t)INT64_MAX) { rc = WTVO_EINVAL; goto done; }
    st->raw_flow = (int64_t)sink_flow;
done:
    free(height);
    free(excess);
    free(cur);
    free(nxt);
    free(queued);
    free(bfsq);
    return rc;
}

/* ------------------------------------------------------------------------ */
/* Boykov-Kolmogorov max-flow (PAMI 2004), restated from the paper:          */
/* two search trees S (from the source) and T (to the sink), growth,         */
/* augmentation, adoption of orphans; with the distance/timestamp heuristic. */
/* Works on the same 4-slot residual representation; terminal arcs are the   */
/* node's term value (source supply > 0, sink capacity < 0).                 */
/* ------------------------------------------------------------------------ */
#define BK_NONE 0xFFu   /* free node                    */
#define BK_TERM 0xFEu   /* parent is the terminal       */
#define BK_ORPH 0xFDu   /* orphan                       */
#define BK_INF_D 0x7FFFFFFF

typedef struct bk_ctx {
    uint32_t n;
    const int32_t* nbr;
    uint32_t* res;
    int64_t* tcap;   /* >0: residual source->v ; <0: residual v->sink */
    uint8_t* rs;     /* reverse slot per arc: rs[4v+k] */
    uint8_t* par;    /* parent slot (0..3) or BK_NONE/BK_TERM/BK_ORPH */
    uint8_t* tree;   /* 0 = free, 1 = S, 2 = T */
    int32_t* ts;     /* timestamp */
    int32_t* dist;
    uint32_t* aq;    /* active FIFO ring */
    uint8_t* inq;
    size_t aq_head, aq_tail, aq_cap;
    uint32_t* oq;    /* orphan FIFO */
    size_t oq_head, oq_tail;
    int32_t time;
    int64_t flow;
} bk_ctx;

static void bk_push_active(bk_ctx* c, uint32_t v)
{
    if (c->inq[v]) return;
    c->inq[v] = 1;
    c->aq[c->aq_tail] = v;
    c->aq_tail = (c->aq_tail + 1) % c->aq_cap;
}

static void bk_push_orphan(bk_ctx* c, uint32_t v)
{
    c->par[v] = BK_ORPH;
    c->oq[c->oq_tail++] = v;
}

/* Residual capacity of the arc used to move flow from a to b, where b's
   parent is a (S tree): arc a->b. For the T tree, child b has parent a and the
   flow goes b->a. The parent slot of v is stored as v's slot k (nbr[v][k] is
   the parent). */
static inline uint32_t* bk_arc(bk_ctx* c, uint32_t v, int k) { return &c->res[(size_t)v * 4u + (size_t)k]; }

/* Walk from v to its root; returns 1 if it reaches a terminal, sets dist. */
static int bk_origin_dist(bk_ctx* c, uint32_t v, int32_t* d_out)
{
    int32_t d = 0;
    uint32_t x = v;
    for (;;) {
        if (c->ts[x] == c->time) { d += c->dist[x]; break; }
        const uint8_t p = c->par[x];
        ++d;
        if (p == BK_TERM) { c->ts[x] = c->time; c->dist[x] = 1; break; }
        if (p == BK_ORPH || p == BK_NONE) return 0;
        x = (uint32_t)c->nbr[(size_t)x * 4u + p];
    }
    /* mark the path with distances for later queries */
    uint32_t y = v;
    int32_t dd = d;
    while (c->ts[y] != c->time) {
        c->ts[y] = c->time;
        c->dist[y] = dd--;
        y = (uint32_t)c->nbr[(size_t)y * 4u + c->par[y]];
    }
    *d_out = d;
    return 1;
}

static void bk_augment(bk_ctx* c, uint32_t s_side, int k_mid, uint32_t t_side)
{
    /* path: source ... -> s_side -(k_mid)-> t_side -> ... sink */
    int64_t bott = (int64_t)*bk_arc(c, s_side, k_mid);
    /* S side: walk to the source; flow goes parent -> child, arc parent->child
       is the reverse slot of the child's parent slot. */
    for (uint32_t x = s_side;;) {
        const uint8_t p = c->par[x];
        if (p == BK_TERM) {
            if (c->tcap[x] < bott) bott = c->tcap[x];
            break;
        }
        const uint32_t y = (uint32_t)c->nbr[(size_t)x * 4u + p];
        const int rk = c->rs[(size_t)x * 4u + p];
        const int64_t r = (int64_t)*bk_arc(c, y, rk);
        if (r < bott) bott = r;
        x = y;
    }
    for (uint32_t x = t_side;;) {
        const uint8_t p = c->par[x];
        if (p == BK_TERM) {
            if (-c->tcap[x] < bott) bott = -c->tcap[x];
            break;
        }
        const int64_t r = (int64_t)*bk_arc(c, x, p);
        if (r < bott) bott = r;
        x = (uint32_t)c->nbr[(size_t)x * 4u + p];
    }
    /* apply */
    {
        const int rk = c->rs[(size_t)s_side * 4u + (size_t)k_mid];
        *bk_arc(c, s_side, k_mid) -= (uint32_t)bott;
        *bk_arc(c, t_side, rk) += (uint32_t)bott;
    }
    for (uint32_t x = s_side;;) {
        const uint8_t p = c->par[x];
        if (p == BK_TERM) {
            c->tcap[x] -= bott;
            if (c->tcap[x] == 0) bk_push_orphan(c, x);
            break;
        }
        const uint32_t y = (uint32_t)c->nbr[(size_t)x * 4u + p];
        const int rk = c->rs[(size_t)x * 4u + p];
        *bk_arc(c, y, rk) -= (uint32_t)bott;
        *bk_arc(c, x, p) += (uint32_t)bott;
        if (*bk_arc(c, y, rk) == 0u) bk_push_orphan(c, x);
        x = y;
    }
    for (uint32_t x = t_side;;) {
        const uint8_t p = c->par[x];
        if (p == BK_TERM) {
            c->tcap[x] += bott;
            if (c->tcap[x] == 0) bk_push_orphan(c, x);
            break;
        }
        const uint32_t y = (uint32_t)c->nbr[(size_t)x * 4u + p];
        const int rk = c->rs[(size_t)x * 4u + p];
        *bk_arc(c, x, p) -= (uint32_t)bott;
        *bk_arc(c, y, rk) += (uint32_t)bott;
        if (*bk_arc(c, x, p) == 0u) bk_push_orphan(c, x);
        x = y;
    }
    c->flow += bott;
}

static void bk_adopt(bk_ctx* c, uint32_t v)
{
    const int in_s = (c->tree[v] == 1);
    int best_k = -1;
    int32_t best_d = BK_INF_D;
    for (int k = 0; k < 4; ++k) {
        const int32_t ww = c->nbr[(size_t)v * 4u + (size_t)k];
        if (ww < 0) continue;
        const uint32_t w = (uint32_t)ww;
        if (c->tree[w] != c->tree[v]) continue;
        /* S: need residual w->v; T: need residual v->w */
        const uint32_t r = in_s ? *bk_arc(c, w, c->rs[(size_t)v * 4u + (size_t)k]) : *bk_arc(c, v, k);
        if (!r) continue;
        int32_t d;
        if (!bk_origin_dist(c, w, &d)) continue;
        if (d < best_d) { best_d = d; best_k = k; }
    }
    if (best_k >= 0) {
        c->par[v] = (uint8_t)best_k;
        c->ts[v] = c->time;
        c->dist[v] = best_d + 1;
        return;
    }
    /* no valid parent: v becomes free; its children become orphans, and its
       same-tree neighbours with residual towards/from v become active */
    for (int k = 0; k < 4; ++k) {
        const int32_t ww = c->nbr[(size_t)v * 4u + (size_t)k];
        if (ww < 0) continue;
        const uint32_t w = (uint32_t)ww;
        if (c->tree[w] != c->tree[v]) continue;
        const uint32_t r = in_s ? *bk_arc(c, w, c->rs[(size_t)v * 4u + (size_t)k]) : *bk_arc(c, v, k);
        if (r) bk_push_active(c, w);
        const uint8_t pw = c->par[w];
        if (pw < 4 && c->nbr[(size_t)w * 4u + pw] == (int32_t)v) bk_push_orphan(c, w);
    }
    c->tree[v] = 0;
    c->par[v] = BK_NONE;
}

int wtvo_bk_solve(uint32_t n, const int32_t* nbr, uint32_t* res, int32_t* term,
                  uint8_t* partition, int64_t* flow_out)
{
    *flow_out = 0;
    if (!n) return WTVO_OK;
    bk_ctx c;
    memset(&c, 0, sizeof(c));
    const size_t N = n;
    c.n = n;
    c.nbr = nbr;
    c.res = res;
    c.tcap = (int64_t*)malloc(N * 8u);
    c.rs = (uint8_t*)malloc(N * 4u);
    c.par = (uint8_t*)malloc(N);
    c.tree = (uint8_t*)calloc(N, 1u);
    c.ts = (int32_t*)calloc(N, 4u);
    c.dist = (int32_t*)calloc(N, 4u);
    c.aq_cap = N + 1u;
    c.aq = (uint32_t*)malloc(c.aq_cap * 4u);
    c.inq = (uint8_t*)calloc(N, 1u);
    c.oq = (uint32_t*)malloc((N * 8u + 16u) * 4u); /* orphan FIFO, reset per augment */
    int rc = WTVO_OK;
    if (!c.tcap || !c.rs || !c.par || !c.tree || !c.ts || !c.dist || !c.aq || !c.inq || !c.oq) {
        rc = WTVO_ENOMEM;
        goto done;
    }
    for (size_t v = 0; v < N; ++v) {
        for (int k = 0; k < 4; ++k) {
            const int32_t u = nbr[v * 4u + (size_t)k];
            c.rs[v * 4u + (size_t)k] = 0;
            if (u < 0) continue;
            const int rk = reverse_slot(nbr, (uint32_t)u, (uint32_t)v);
            if (rk < 0) { rc = WTVO_EINVARIANT; goto done; }
            c.rs[v * 4u + (size_t)k] = (uint8_t)rk;
        }
        c.tcap[v] = term[v];
        c.par[v] = BK_NONE;
        if (term[v] > 0) { c.tree[v] = 1; c.par[v] = BK_TERM; c.dist[v] = 1; bk_push_active(&c, (uint32_t)v); }
        else if (term[v] < 0) { c.tree[v] = 2; c.par[v] = BK_TERM; c.dist[v] = 1; bk_push_active(&c, (uint32_t)v); }
    }
    c.time = 0;
    while (c.aq_head != c.aq_tail) {
        const uint32_t v = c.aq[c.aq_head];
        /* keep v at the head while it still may find a path */
        if (c.tree[v] == 0) { c.inq[v] = 0; c.aq_head = (c.aq_head + 1) % c.aq_cap; continue; }
        int found = 0;
        uint32_t s_side = 0, t_side = 0;
        int k_mid = 0;
        for (int k = 0; k < 4 && !found; ++k) {
            const int32_t ww = nbr[(size_t)v * 4u + (size_t)k];
            if (ww < 0) continue;
            const uint32_t w = (uint32_t)ww;
            if (c.tree[v] == 1) {
                if (!res[(size_t)v * 4u + (size_t)k]) continue;
                if (c.tree[w] == 0) {
                    c.tree[w] = 1;
                    c.par[w] = c.rs[(size_t)v * 4u + (size_t)k];
                    c.ts[w] = c.ts[v];
                    c.dist[w] = c.dist[v] + 1;
                    bk_push_active(&c, w);
                } else if (c.tree[w] == 2) {
                    found = 1; s_side = v; t_side = w; k_mid = k;
                } else if (c.ts[w] <= c.ts[v] && c.dist[w] > c.dist[v]) {
                    c.par[w] = c.rs[(size_t)v * 4u + (size_t)k];
                    c.ts[w] = c.ts[v];
                    c.dist[w] = c.dist[v] + 1;
                }
            } else {
                const int rk = c.rs[(size_t)v * 4u + (size_t)k];
                if (!res[(size_t)w * 4u + (size_t)rk]) continue;
                if (c.tree[w] == 0) {
                    c.tree[w] = 2;
                    c.par[w] = (uint8_t)rk;
                    c.ts[w] = c.ts[v];
                    c.dist[w] = c.dist[v] + 1;
                    bk_push_active(&c, w);
                } else if (c.tree[w] == 1) {
                    found = 1; s_side = w; t_side = v; k_mid = rk;
                } else if (c.ts[w] <= c.ts[v] && c.dist[w] > c.dist[v]) {
                    c.par[w] = (uint8_t)rk;
                    c.ts[w] = c.ts[v];
                    c.dist[w] = c.dist[v] + 1;
                }
            }
        }
        if (!found) {
            c.inq[v] = 0;
            c.aq_head = (c.aq_head + 1) % c.aq_cap;
            continue;
        }
        ++c.time;
        c.oq_head = c.oq_tail = 0;
        bk_augment(&c, s_side, k_mid, t_side);
        while (c.oq_head < c.oq_tail) {
            const uint32_t o = c.oq[c.oq_head++];
            if (c.oq_head == c.oq_tail) c.oq_head = c.oq_tail = 0;
            bk_adopt(&c, o);
        }
        /* v stays at the head of the queue (it may still have paths) */
    }
    /* Report terminals in the Phase-I convention of the reference: source arcs
       count as saturated (term > 0 unchanged; any unused supply is excess left
       on the source side), term < 0 is the remaining sink capacity. */
    for (size_t v = 0; v < N; ++v) term[v] = term[v] > 0 ? term[v] : (int32_t)c.tcap[v];
    *flow_out = c.flow;
    rc = wtvo_partition(n, nbr, res, term, partition);
done:
    free(c.tcap); free(c.rs); free(c.par); free(c.tree); free(c.ts);
    free(c.dist); free(c.aq); free(c.inq); free(c.oq);
    return rc;
}

/* ------------------------------------------------------------------------ */
/* Certificate: preflow validity + cut capacity == flow value + the labels   */
/* are exactly the vertices that reach the sink in the residual graph.       */
/* Returns 0 if all checks pass, else a small code naming the first failure. */
/*  1 arc-pair capacity sum changed        2 negative excess                 */
/*  3 flow value mismatch with sink arcs   4 residual arc S->T               */
/*  5 S vertex with remaining sink cap     6 T vertex with excess            */
/*  7 cut capacity != flow                 8 labels != residual reachability  */
/*  9 missing reverse slot                                                    */
/* ------------------------------------------------------------------------ */
int wtvo_certificate(uint32_t n, const int32_t* nbr, const uint32_t* res0,
                     const int32_t* term0, const uint32_t* res,
                     const int32_t* term, const uint8_t* partition,
                     int64_t raw_flow)
{
    const size_t N = n;
    int64_t sink_total = 0;
    for (size_t v = 0; v < N; ++v) {
        int64_t inflow = term0[v] > 0 ? term0[v] : 0; /* source arcs saturated */
        if (term0[v] > 0 && term[v] != term0[v]) return 3;
        if (term0[v] <= 0) {
            const int64_t used = (int64_t)term[v] - (int64_t)term0[v];
            if (used < 0 || term[v] > 0) return 3;
            sink_total += used;
            inflow -= used;
        }
        for (int k = 0; k < 4; ++k) {
            const int32_t uu = nbr[v * 4u + (size_t)k];
            if (uu < 0) continue;
            const int rk = reverse_slot(nbr, (uint32_t)uu, (uint32_t)v);
            if (rk < 0) return 9;
            const size_t e = v * 4u + (size_t)k, r = (size_t)uu * 4u + (size_t)rk;
            if ((uint64_t)res[e] + res[r] != (uint64_t)res0[e] + res0[r]) return 1;
            /* net flow v->u = res0[e] - res[e] (as signed) */
            inflow -= (int64_t)res0[e] - (int64_t)res[e];
        }
        if (inflow < 0) return 2;
        if (partition[v] && inflow != 0) return 6;
    }
    if (sink_total != raw_flow) return 3;
    int64_t cut = 0;
    for (size_t v = 0; v < N; ++v) {
        const int s_side = !partition[v];
        if (s_side) {
            if (term[v] < 0) return 5;
            if (term0[v] < 0) cut += -(int64_t)term0[v];
            for (int k = 0; k < 4; ++k) {
                const int32_t uu = nbr[v * 4u + (size_t)k];
                if (uu < 0 || !partition[uu]) continue;
                if (res[v * 4u + (size_t)k]) return 4;
                cut += res0[v * 4u + (size_t)k];
            }
        } else if (term0[v] > 0) {
            cut += term0[v];
        }
    }
    if (cut != raw_flow) return 7;
    uint8_t* p2 = (uint8_t*)malloc(N + 1u);
    if (!p2) return WTVO_ENOMEM;
    int rc = wtvo_partition(n, nbr, res, term, p2);
    if (rc == WTVO_OK && memcmp(p2, partition, N) != 0) rc = 8;
    else if (rc != WTVO_OK) rc = 9;
    free(p2);
    return rc;
}

/* Flow certificate: 0 when (res, term) is a maximum FLOW -- every vertex
 * conserves -- with sink flow raw_flow and `partition` its minimal sink set.
 * term0 > 0: term is the source flow still sent into v, in [0, term0];
 * term0 <= 0: term - term0 is the sink flow out of v (term <= 0).
 * 1 broken residual pair, 2 a vertex that does not conserve, 3 terminal out
 * of range or sink total != raw_flow, 4 residual arc from the source side to
 * the sink side, 7 cut capacity != raw_flow, 8 partition not the minimal sink
 * set, 9 missing reverse arc. */
int wtvo_flow_certificate(uint32_t n, const int32_t* nbr, const uint32_t* res0,
                          const int32_t* term0, const uint32_t* res,
                          const int32_t* term, const uint8_t* partition,
                          int64_t raw_flow)
{
    const size_t N = n;
    int64_t sink_total = 0;
    for (size_t v = 0; v < N; ++v) {
        int64_t bal = 0;
        if (term0[v] > 0) {
            if (term[v] < 0 || term[v] > term0[v]) return 3;
            bal += term[v];
        } else {
            const int64_t used = (int64_t)term[v] - (int64_t)term0[v];
            if (used < 0 || term[v] > 0) return 3;
            sink_total += used;
            bal -= used;
        }
        for (int k = 0; k < 4; ++k) {
            const int32_t uu = nbr[v * 4u + (size_t)k];
            if (uu < 0) continue;
            const int rk = reverse_slot(nbr, (uint32_t)uu, (uint32_t)v);
            if (rk < 0) return 9;
            const size_t e = v * 4u + (size_t)k, r = (size_t)uu * 4u + (size_t)rk;
            if ((uint64_t)res[e] + res[r] != (uint64_t)res0[e] + res0[r]) return 1;
            bal -= (int64_t)res0[e] - (int64_t)res[e];
        }
        if (bal != 0) return 2;
    }
    if (sink_total != raw_flow) return 3;
    int64_t cut = 0;
    for (size_t v = 0; v < N; ++v) {
        if (!partition[v]) {
            if (term0[v] < 0) cut += -(int64_t)term0[v];
            for (int k = 0; k < 4; ++k) {
                const int32_t uu = nbr[v * 4u + (size_t)k];
                if (uu < 0 || !partition[uu]) continue;
                if (res[v * 4u + (size_t)k]) return 4;
                cut += res0[v * 4u + (size_t)k];
            }
        } else if (term0[v] > 0) {
            cut += term0[v];
        }
    }
    if (cut != raw_flow) return 7;
    uint8_t* p2 = (uint8_t*)malloc(N + 1u);
    if (!p2) return WTVO_ENOMEM;
    int rc = wtvo_partition(n, nbr, res, term, p2);
    if (rc == WTVO_OK && memcmp(p2, partition, N) != 0) rc = 8;
    else if (rc != WTVO_OK) rc = 9;
    free(p2);
    return rc;
}

/* ------------------------------------------------------------------------ */
/* Surface extraction: wtivo_gpupr_bindings.cpp:400-577.                      */
/* Boundary triangles of label-0 tets (hull faces and faces towards a tet     */
/* with label != 0), tet-major / slot-minor order; vertices sorted, then      */
/* oriented by the sign of the 4th vertex's volume; vertices compacted in     */
/* ascending original id.  Two calls: first with faces == NULL to get the     */
/* counts, then with buffers of those sizes.                                  */
/* Returns 0 ok, 1 neighbour out of range, 2 invalid tet vertex.              */
/* ------------------------------------------------------------------------ */
int wtvo_surface(size_t N, const uint8_t* labels, const double* V, size_t nv,
                 const int32_t* tets, const int32_t* nbrs, size_t* n_faces,
                 size_t* n_used, double* out_v, int32_t* out_f)
{
    static const int opp[4] = {3, 2, 1, 0};
    size_t F = 0;
    for (size_t i = 0; i < N; ++i) {
        if (labels[i] != 0u) continue;
        for (int fi = 0; fi < 4; ++fi) {
            const int32_t j = nbrs[i * 4u + (size_t)fi];
            if (j >= 0 && (size_t)j >= N) return 1;
            if (j < 0 || labels[j] != 0u) ++F;
        }
    }
    *n_faces = F;
    int32_t* raw = (int32_t*)malloc((F * 3u + 1u) * sizeof(int32_t));
    int32_t* remap = (int32_t*)malloc((nv + 1u) * sizeof(int32_t));
    if (!raw || !remap) { free(raw); free(remap); return WTVO_ENOMEM; }
    size_t p = 0;
    for (size_t i = 0; i < N; ++i) {
        if (labels[i] != 0u) continue;
        const int32_t* tr = tets + i * 4u;
        for (int fi = 0; fi < 4; ++fi) {
            const int32_t j = nbrs[i * 4u + (size_t)fi];
            if (!(j < 0 || labels[j] != 0u)) continue;
            int a0 = tr[k_face_lut[fi][0]], b0 = tr[k_face_lut[fi][1]], c0 = tr[k_face_lut[fi][2]];
            const int fourth = tr[opp[fi]];
            if (a0 < 0 || b0 < 0 || c0 < 0 || fourth < 0 || (size_t)a0 >= nv || (size_t)b0 >= nv ||
                (size_t)c0 >= nv || (size_t)fourth >= nv) {
                free(raw); free(remap);
                return 2;
            }
            int t;
            if (a0 > b0) { t = a0; a0 = b0; b0 = t; }
            if (b0 > c0) { t = b0; b0 = c0; c0 = t; }
            if (a0 > b0) { t = a0; a0 = b0; b0 = t; }
            const double* v0 = V + (size_t)a0 * 3u;
            const double* v1 = V + (size_t)b0 * 3u;
            const double* v2 = V + (size_t)c0 * 3u;
            const double* vp = V + (size_t)fourth * 3u;
            const double x1 = v1[0] - v0[0], y1 = v1[1] - v0[1], z1 = v1[2] - v0[2];
            const double x2 = v2[0] - v0[0], y2 = v2[1] - v0[1], z2 = v2[2] - v0[2];
            const double cx = y1 * z2 - z1 * y2;
            const double cy = z1 * x2 - x1 * z2;
            const double cz = x1 * y2 - y1 * x2;
            const double px = vp[0] - v0[0], py = vp[1] - v0[1], pz = vp[2] - v0[2];
            const double vol = cx * px + cy * py + cz * pz;
            raw[3u * p] = a0;
            raw[3u * p + 1u] = (vol < 0.0) ? b0 : c0;
            raw[3u * p + 2u] = (vol < 0.0) ? c0 : b0;
            ++p;
        }
    }
    for (size_t v = 0; v < nv; ++v) remap[v] = -1;
    for (size_t q = 0; q < 3u * F; ++q) remap[raw[q]] = -2;
    size_t used = 0;
    for (size_t v = 0; v < nv; ++v)
        if (remap[v] == -2) remap[v] = (int32_t)used++;
    *n_used = used;
    if (out_v && out_f) {
        for (size_t v = 0; v < nv; ++v)
            if (remap[v] >= 0) memcpy(out_v + (size_t)remap[v] * 3u, V + v * 3u, 3u * sizeof(double));
        for (size_t q = 0; q < 3u * F; ++q) out_f[q] = remap[raw[q]];
    }
    free(raw);
    free(remap);
    return WTVO_OK;
}

/* ------------------------------------------------------------------------ */
/* Tet labels from a sparse grid: SparseUDF.sample_tet_labels               */
/* (native/vdb/wtivo_vdb.cpp:1519-1584) with OpenVDB's BoxSampler as        */
/* published (openvdb/tools/Interpolation.h: floorVec3, uvw = p - floor(p), */
/* trilinearInterpolation's lerp a + ValueT((b - a) * w), z then y then x)  */
/* and the linear transform of Transform::createLinearTransform(voxel)      */
/* (index = world * (1 / voxel)).  Leaves are found by binary search over   */
/* their sorted cell keys.                                                   */
/* ------------------------------------------------------------------------ */
typedef struct {
    int64_t key;
    int64_t leaf;
} wtvo_leaf_key;

static int64_t wtvo_cell_key(int64_t cx, int64_t cy, int64_t cz)
{
    /* cells lie within +-2^20 on each axis for any grid this oracle is given */
    return ((cx + (1 << 20)) << 42) | ((cy + (1 << 20)) << 21) | (cz + (1 << 20));
}

static int wtvo_key_cmp(const void* a, const void* b)
{
    const int64_t x = ((const wtvo_leaf_key*)a)->key, y = ((const wtvo_leaf_key*)b)->key;
    return (x > y) - (x < y);
}

static float wtvo_grid_value(const wtvo_leaf_key* keys, size_t L, const float* values, float background,
                             int64_t x, int64_t y, int64_t z)
{
    const int64_t cx = x >> 3, cy = y >> 3, cz = z >> 3;
    const int64_t lim = (int64_t)1 << 20;
    if (cx < -lim || cx >= lim || cy < -lim || cy >= lim || cz < -lim || cz >= lim) return background;
    wtvo_leaf_key probe = {wtvo_cell_key(cx, cy, cz), 0};
    const wtvo_leaf_key* hit = (const wtvo_leaf_key*)bsearch(&probe, keys, L, sizeof(wtvo_leaf_key), wtvo_key_cmp);
    if (!hit) return background;
    const size_t off = ((size_t)(x & 7) << 6) | ((size_t)(y & 7) << 3) | (size_t)(z & 7);
    return values[(size_t)hit->leaf * 512u + off];
}

static float wtvo_lerp(float a, float b, double w)
{
    const double t = (double)(b - a) * w;
    return a + (float)t;
}

/* labels[i] (and, if values_out, the sampled value) for every tet.  Returns
 * WTVO_EINVAL for a degenerate bbox, a vertex id out of range, a leaf origin
 * that is not a multiple of 8 or outside +-2^23, or two leaves on one cell. */
int wtvo_sample_labels(size_t L, const int32_t* origins, const float* values, float background,
                       double voxel_size, size_t nv, const double* V, size_t N, const int32_t* tets,
                       const double* bmin, const double* bmax, double padding, double threshold,
                       uint8_t* labels, float* values_out)
{
    const double sx = bmax[0] - bmin[0], sy = bmax[1] - bmin[1], sz = bmax[2] - bmin[2];
    if (!(sx > 0 && sy > 0 && sz > 0)) return WTVO_EINVAL;
    wtvo_leaf_key* keys = (wtvo_leaf_key*)malloc((L ? L : 1) * sizeof(wtvo_leaf_key));
    if (!keys) return WTVO_ENOMEM;
    for (size_t l = 0; l < L; ++l) {
        const int32_t* o = origins + 3u * l;
        for (int a = 0; a < 3; ++a)
            if ((o[a] & 7) || o[a] < -(1 << 23) || o[a] >= (1 << 23)) {
                free(keys);
                return WTVO_EINVAL;
            }
        keys[l].key = wtvo_cell_key(o[0] >> 3, o[1] >> 3, o[2] >> 3);
        keys[l].leaf = (int64_t)l;
    }
    qsort(keys, L, sizeof(wtvo_leaf_key), wtvo_key_cmp);
    for (size_t l = 1; l < L; ++l)
        if (keys[l].key == keys[l - 1].key) {
            free(keys);
            return WTVO_EINVAL;
        }
    const double scale = 1.0 - 2.0 * padding;
    const double inv = 1.0 / voxel_size;
    int rc = WTVO_OK;
    for (size_t i = 0; i < N; ++i) {
        double cx = 0.0, cy = 0.0, cz = 0.0;
        for (int q = 0; q < 4; ++q) {
            const int32_t vi = tets[4u * i + (size_t)q];
            if (vi < 0 || (size_t)vi >= nv) {
                free(keys);
                return WTVO_EINVAL;
            }
            cx += V[3u * (size_t)vi + 0];
            cy += V[3u * (size_t)vi + 1];
            cz += V[3u * (size_t)vi + 2];
        }
        cx *= 0.25;
        cy *= 0.25;
        cz *= 0.25;
        const double p[3] = {(((cx - bmin[0]) / sx) * scale + padding) * inv,
                             (((cy - bmin[1]) / sy) * scale + padding) * inv,
                             (((cz - bmin[2]) / sz) * scale + padding) * inv};
        int64_t ix[3];
        double uvw[3];
        for (int a = 0; a < 3; ++a) {
            double f = floor(p[a]);
            uvw[a] = p[a] - f;
            if (f < -1099511627776.0) f = -1099511627776.0;
            if (f > 1099511627776.0) f = 1099511627776.0;
            ix[a] = (int64_t)f;
        }
        float v[2][2][2];
        for (int dx = 0; dx < 2; ++dx)
            for (int dy = 0; dy < 2; ++dy)
                for (int dz = 0; dz < 2; ++dz)
                    v[dx][dy][dz] = wtvo_grid_value(keys, L, values, background, ix[0] + dx, ix[1] + dy, ix[2] + dz);
        const float d = wtvo_lerp(wtvo_lerp(wtvo_lerp(v[0][0][0], v[0][0][1], uvw[2]),
                                            wtvo_lerp(v[0][1][0], v[0][1][1], uvw[2]), uvw[1]),
                                  wtvo_lerp(wtvo_lerp(v[1][0][0], v[1][0][1], uvw[2]),
                                            wtvo_lerp(v[1][1][0], v[1][1][1], uvw[2]), uvw[1]),
                                  uvw[0]);
        labels[i] = ((double)d <= threshold) ? 0u : 1u;
        if (values_out) values_out[i] = d;
    }
    free(keys);
    return rc;
}
