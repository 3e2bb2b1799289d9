// wtv_delaunay.cpp -- 3D Delaunay tetrahedralisation of a point set, for the
// benchmark and test meshes of the production mesh family (libwtv_delaunay.so,
// not part of the solver library).
//
// The reference meshes its proxy points with CGAL's Delaunay_3 and exports
// each cell's vertices and 4-slot neighbours (native/core/wtivo_core.cpp:72-165).
// CGAL is not in this image and scipy's Qhull needs ~84 s per million points
// here, so config-size meshes (12M points, ~78M tets) come from this file.
//
// Method: incremental Bowyer-Watson insertion in Morton order, with an
// infinite vertex closing the hull so that every tet has four neighbours.
// Each point is located by a visibility walk from the previous insertion, the
// tets whose open circumsphere contains it (the cavity) are collected by a
// search over neighbours, and the cavity is replaced by the cone from the
// point to its boundary faces.  Coordinates are integers below 2^21 per axis,
// so the two predicates are exact: a floating-point evaluation with a static
// error bound (Shewchuk's filters for orient3d and insphere), and __int128
// arithmetic when the bound does not decide.  Ties are broken by symbolic
// perturbation of the lifted coordinates with the newest point perturbed
// most, which makes a point on a circumsphere lie outside it; the result is
// then the Delaunay triangulation (the unique one for points in general
// position, the same tets Qhull and CGAL return).
//
// Orientation: finite tets are positive, orient(a,b,c,d) = det[b-a; c-a;
// d-a] > 0.  Infinite tets keep the orientation they would have if the
// infinite vertex were a point beyond their hull face.  tn[4t+i] = 4u+j: tet
// u is across the face of t opposite its vertex i, and t is across u's face
// opposite its vertex j.
#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

namespace {

using i64 = int64_t;
using i128 = __int128;

constexpr int32_t kInf = -1;

struct P3 {
    double x, y, z;  // integer valued, |.| < 2^21
};

// orient(a,b,c,d) = det[b-a; c-a; d-a]; +1, 0 or -1, exact.
int orient(const P3& a, const P3& b, const P3& c, const P3& d) {
    const double bx = b.x - a.x, by = b.y - a.y, bz = b.z - a.z;
    const double cx = c.x - a.x, cy = c.y - a.y, cz = c.z - a.z;
    const double dx = d.x - a.x, dy = d.y - a.y, dz = d.z - a.z;
    const double m1 = cy * dz - cz * dy, m2 = cz * dx - cx * dz, m3 = cx * dy - cy * dx;
    const double det = bx * m1 + by * m2 + bz * m3;
    const double perm = std::fabs(bx) * (std::fabs(cy * dz) + std::fabs(cz * dy)) +
                        std::fabs(by) * (std::fabs(cz * dx) + std::fabs(cx * dz)) +
                        std::fabs(bz) * (std::fabs(cx * dy) + std::fabs(cy * dx));
    const double bound = 7.771561172376103e-16 * perm;  // (7 + 56 eps) eps, eps = 2^-53
    if (det > bound) return 1;
    if (-det > bound) return -1;
    const i64 Bx = (i64)bx, By = (i64)by, Bz = (i64)bz, Cx = (i64)cx, Cy = (i64)cy, Cz = (i64)cz;
    const i64 Dx = (i64)dx, Dy = (i64)dy, Dz = (i64)dz;
    const i128 e = (i128)Bx * (Cy * Dz - Cz * Dy) + (i128)By * (Cz * Dx - Cx * Dz) + (i128)Bz * (Cx * Dy - Cy * Dx);
    return e > 0 ? 1 : (e < 0 ? -1 : 0);
}

// +1 when e lies inside the circumsphere of the positively oriented tet
// (a,b,c,d), -1 outside, 0 on it; exact.  The value is -det of the rows
// (q - e, |q - e|^2), q = a, b, c, d.
int insphere(const P3& a, const P3& b, const P3& c, const P3& d, const P3& e) {
    const double aex = a.x - e.x, aey = a.y - e.y, aez = a.z - e.z;
    const double bex = b.x - e.x, bey = b.y - e.y, bez = b.z - e.z;
    const double cex = c.x - e.x, cey = c.y - e.y, cez = c.z - e.z;
    const double dex = d.x - e.x, dey = d.y - e.y, dez = d.z - e.z;
    const double ab = aex * bey - bex * aey, bc = bex * cey - cex * bey, cd = cex * dey - dex * cey;
    const double da = dex * aey - aex * dey, ac = aex * cey - cex * aey, bd = bex * dey - dex * bey;
    const double abc = aez * bc - bez * ac + cez * ab;
    const double bcd = bez * cd - cez * bd + dez * bc;
    const double cda = cez * da + dez * ac + aez * cd;
    const double dab = dez * ab + aez * bd + bez * da;
    const double al = aex * aex + aey * aey + aez * aez, bl = bex * bex + bey * bey + bez * bez;
    const double cl = cex * cex + cey * cey + cez * cez, dl = dex * dex + dey * dey + dez * dez;
    // det of the rows (q - e, |q - e|^2)
    const double det = (dl * abc - cl * dab) + (bl * cda - al * bcd);
    const double fab = std::fabs(aex * bey) + std::fabs(bex * aey), fbc = std::fabs(bex * cey) + std::fabs(cex * bey);
    const double fcd = std::fabs(cex * dey) + std::fabs(dex * cey), fda = std::fabs(dex * aey) + std::fabs(aex * dey);
    const double fac = std::fabs(aex * cey) + std::fabs(cex * aey), fbd = std::fabs(bex * dey) + std::fabs(dex * bey);
    const double perm = (fcd * std::fabs(bez) + fbd * std::fabs(cez) + fbc * std::fabs(dez)) * al +
                        (fda * std::fabs(cez) + fac * std::fabs(dez) + fcd * std::fabs(aez)) * bl +
                        (fab * std::fabs(dez) + fbd * std::fabs(aez) + fda * std::fabs(bez)) * cl +
                        (fbc * std::fabs(aez) + fac * std::fabs(bez) + fab * std::fabs(cez)) * dl;
    const double bound = 1.7763568394002532e-15 * perm;  // (16 + 224 eps) eps
    if (det > bound) return -1;
    if (-det > bound) return 1;
    const i64 Ax = (i64)aex, Ay = (i64)aey, Az = (i64)aez, Bx = (i64)bex, By = (i64)bey, Bz = (i64)bez;
    const i64 Cx = (i64)cex, Cy = (i64)cey, Cz = (i64)cez, Dx = (i64)dex, Dy = (i64)dey, Dz = (i64)dez;
    const i64 iab = Ax * By - Bx * Ay, ibc = Bx * Cy - Cx * By, icd = Cx * Dy - Dx * Cy;
    const i64 ida = Dx * Ay - Ax * Dy, iac = Ax * Cy - Cx * Ay, ibd = Bx * Dy - Dx * By;
    const i128 iabc = (i128)Az * ibc - (i128)Bz * iac + (i128)Cz * iab;
    const i128 ibcd = (i128)Bz * icd - (i128)Cz * ibd + (i128)Dz * ibc;
    const i128 icda = (i128)Cz * ida + (i128)Dz * iac + (i128)Az * icd;
    const i128 idab = (i128)Dz * iab + (i128)Az * ibd + (i128)Bz * ida;
    const i128 ial = (i128)(Ax * Ax + Ay * Ay + Az * Az), ibl = (i128)(Bx * Bx + By * By + Bz * Bz);
    const i128 icl = (i128)(Cx * Cx + Cy * Cy + Cz * Cz), idl = (i128)(Dx * Dx + Dy * Dy + Dz * Dz);
    const i128 x = (idl * iabc - icl * idab) + (ibl * icda - ial * ibcd);
    return x > 0 ? -1 : (x < 0 ? 1 : 0);
}

inline uint64_t spread3(uint64_t v) {  // 21 bits -> every third bit
    v &= 0x1FFFFFull;
    v = (v | (v << 32)) & 0x1F00000000FFFFull;
    v = (v | (v << 16)) & 0x1F0000FF0000FFull;
    v = (v | (v << 8)) & 0x100F00F00F00F00Full;
    v = (v | (v << 4)) & 0x10C30C30C30C30C3ull;
    v = (v | (v << 2)) & 0x1249249249249249ull;
    return v;
}

class Delaunay {
  public:
    // threads <= 1: every point inserted by one thread in Morton order.
    // Otherwise a random sample of 65536 points first, then the rest in
    // Morton order split into one contiguous range per thread.
    int run(const int32_t* xyz, int64_t n, int threads) {
        n_ = n;
        if (n < 4) return 2;
        if (n > (int64_t(1) << 30)) return 1;
        // internal vertex k is the k-th point in Morton order, so that the
        // points a predicate reads sit close together in memory; repeated
        // points keep only their first copy (the others stay unused)
        std::vector<std::pair<uint64_t, int32_t>> key(static_cast<size_t>(n));
        for (int64_t i = 0; i < n; ++i) {
            const int32_t x = xyz[3 * i], y = xyz[3 * i + 1], z = xyz[3 * i + 2];
            if (x < 0 || y < 0 || z < 0 || x >= (1 << 21) || y >= (1 << 21) || z >= (1 << 21)) return 1;
            key[i] = {spread3(static_cast<uint64_t>(x)) | spread3(static_cast<uint64_t>(y)) << 1 |
                          spread3(static_cast<uint64_t>(z)) << 2,
                      static_cast<int32_t>(i)};
        }
        std::sort(key.begin(), key.end());
        orig_.reserve(static_cast<size_t>(n));
        P_.reserve(static_cast<size_t>(n));
        for (int64_t k = 0; k < n; ++k) {
            if (k && key[k].first == key[k - 1].first) {
                ++repeated_;
                continue;
            }
            const int32_t i = key[k].second;
            orig_.push_back(i);
            P_.push_back(P3{double(xyz[3 * i]), double(xyz[3 * i + 1]), double(xyz[3 * i + 2])});
        }
        key.clear();
        key.shrink_to_fit();
        nv_ = static_cast<int64_t>(P_.size());
        if (nv_ < 4) return 2;
        if (threads < 1) threads = 1;
        if (nv_ < 200000) threads = 1;
        // a 3D Delaunay triangulation has about 6.5 tets per point; room for
        // 7.5 (and the infinite tets).  The array is not initialised: slots
        // below top_ are written when handed out, the rest are never read, so
        // its pages are first touched by the inserting threads.
        cap_ = static_cast<size_t>(nv_) * 15u / 2u + 65536u;
        T_.reset(new Tet[cap_]);
        vt_.assign(static_cast<size_t>(nv_), -1);
        W_.resize(static_cast<size_t>(threads));
        for (int i = 0; i < threads; ++i) {
            W_[i].id = static_cast<uint32_t>(i);
            W_[i].rng = 2463534242u + 7919u * static_cast<uint32_t>(i);
        }
        Worker& w0 = W_[0];
        if (!first_tet(w0)) return 4;  // all points coplanar
        const int32_t* f = first_;
        auto is_first = [&](int32_t p) { return p == f[0] || p == f[1] || p == f[2] || p == f[3]; };
        std::vector<uint8_t> done(static_cast<size_t>(nv_), 0);
        for (int i = 0; i < 4; ++i) done[static_cast<size_t>(f[i])] = 1;
        if (threads > 1) {
            // a coarse triangulation of the whole domain: a random sample
            std::vector<int32_t> sample;
            uint64_t s = 0x2545F4914F6CDD1Dull;
            for (int k = 0; k < 65536; ++k) {
                s ^= s << 13;
                s ^= s >> 7;
                s ^= s << 17;
                const int32_t p = static_cast<int32_t>(s % static_cast<uint64_t>(nv_));
                if (!done[static_cast<size_t>(p)]) {
                    done[static_cast<size_t>(p)] = 1;
                    sample.push_back(p);
                }
            }
            std::sort(sample.begin(), sample.end());
            for (int32_t p : sample)
                if (insert(w0, p) != 0) return 5;
        }
        std::vector<int32_t> rest;
        rest.reserve(static_cast<size_t>(nv_));
        for (int32_t p = 0; p < nv_; ++p)
            if (!done[static_cast<size_t>(p)] && !is_first(p)) rest.push_back(p);
        done.clear();
        done.shrink_to_fit();
        std::atomic<int> status{0};
        auto body = [&](int ti) {
            Worker& w = W_[static_cast<size_t>(ti)];
            const size_t lo = rest.size() * static_cast<size_t>(ti) / static_cast<size_t>(threads);
            const size_t hi = rest.size() * static_cast<size_t>(ti + 1) / static_cast<size_t>(threads);
            for (size_t k = lo; k < hi && status.load(std::memory_order_relaxed) == 0; ++k) {
                const int rc = insert(w, rest[k]);
                if (rc == 1) w.retry.push_back(rest[k]);
                else if (rc != 0) status.store(rc);
            }
            // points that met another thread's locks, while progress is made
            for (size_t before = w.retry.size() + 1; !w.retry.empty() && w.retry.size() < before;) {
                before = w.retry.size();
                std::vector<int32_t> again;
                again.swap(w.retry);
                for (int32_t p : again) {
                    const int rc = insert(w, p);
                    if (rc == 1) w.retry.push_back(p);
                    else if (rc != 0) status.store(rc);
                }
            }
        };
        std::vector<std::thread> pool;
        for (int i = 1; i < threads; ++i) pool.emplace_back(body, i);
        body(0);
        for (auto& t : pool) t.join();
        if (status.load() != 0) return status.load();
        // what is left, alone: no other thread holds a lock now
        for (auto& w : W_)
            for (int32_t p : w.retry)
                if (insert(w0, p) != 0) return 5;
        return 0;
    }

    int64_t repeated() const { return repeated_; }
    int threads() const { return static_cast<int>(W_.size()); }

    int64_t finite_tets() const {
        int64_t c = 0;
        for (size_t t = 0; t < top_; ++t)
            if (alive(t) && !infinite(t)) ++c;
        return c;
    }

    // Finite tets in a canonical order: by their smallest vertex in Morton
    // order, then by the other three; each tet starts at its smallest vertex
    // and keeps its orientation.  Vertex ids are the caller's; slot fi holds
    // the neighbour across local face {012,013,023,123}[fi], the face
    // opposite vertex 3-fi, -1 on the hull.
    void export_mesh(int32_t* tets, int32_t* nbrs, int threads) const {
        struct Rec {
            int32_t v[4];  // rotated: v[0] smallest, orientation kept
            int32_t t;     // storage index
        };
        const size_t top = top_;
        const size_t T = static_cast<size_t>(std::max(1, threads));
        auto parallel = [&](auto&& fn) {  // fn(lo, hi) over [0, top) in T ranges
            std::vector<std::thread> pool;
            for (size_t j = 1; j < T; ++j) pool.emplace_back(fn, top * j / T, top * (j + 1) / T);
            fn(size_t(0), top / T);
            for (auto& th : pool) th.join();
        };
        auto rotated = [&](size_t t) {
            const int32_t* v = T_[t].v;
            int m = 0;
            for (int i = 1; i < 4; ++i)
                if (v[i] < v[m]) m = i;
            // an even permutation that brings position m to the front ...
            static const int rot[4][4] = {{0, 1, 2, 3}, {1, 0, 3, 2}, {2, 3, 0, 1}, {3, 2, 1, 0}};
            Rec r;
            for (int i = 0; i < 4; ++i) r.v[i] = v[rot[m][i]];
            // ... then an even rotation of the last three, smallest first
            while (!(r.v[1] < r.v[2] && r.v[1] < r.v[3])) {
                const int32_t x = r.v[1];
                r.v[1] = r.v[2];
                r.v[2] = r.v[3];
                r.v[3] = x;
            }
            r.t = static_cast<int32_t>(t);
            return r;
        };
        // counting sort by the smallest vertex, then each bucket by the others
        std::vector<uint32_t> start(static_cast<size_t>(nv_) + 1u, 0u);
        parallel([&](size_t lo, size_t hi) {
            for (size_t t = lo; t < hi; ++t) {
                if (!alive(t) || infinite(t)) continue;
                const int32_t* v = T_[t].v;
                const int32_t m = std::min(std::min(v[0], v[1]), std::min(v[2], v[3]));
                __atomic_fetch_add(&start[static_cast<size_t>(m) + 1u], 1u, __ATOMIC_RELAXED);
            }
        });
        for (size_t i = 1; i < start.size(); ++i) start[i] += start[i - 1];
        const size_t nt = start.back();
        std::vector<Rec> rec(nt);
        std::vector<uint32_t> cursor(start.begin(), start.end() - 1);
        parallel([&](size_t lo, size_t hi) {
            for (size_t t = lo; t < hi; ++t) {
                if (!alive(t) || infinite(t)) continue;
                const Rec r = rotated(t);
                rec[__atomic_fetch_add(&cursor[static_cast<size_t>(r.v[0])], 1u, __ATOMIC_RELAXED)] = r;
            }
        });
        cursor.clear();
        cursor.shrink_to_fit();
        auto less = [](const Rec& a, const Rec& b) {
            if (a.v[1] != b.v[1]) return a.v[1] < b.v[1];
            const int32_t a2 = std::min(a.v[2], a.v[3]), b2 = std::min(b.v[2], b.v[3]);
            if (a2 != b2) return a2 < b2;
            return std::max(a.v[2], a.v[3]) < std::max(b.v[2], b.v[3]);
        };
        std::vector<int32_t> id(top, -1);
        {
            std::vector<std::thread> pool;
            const size_t nv = static_cast<size_t>(nv_);
            auto fn = [&](size_t lo, size_t hi) {
                for (size_t b = lo; b < hi; ++b) {
                    std::sort(rec.begin() + start[b], rec.begin() + start[b + 1], less);
                    for (size_t o = start[b]; o < start[b + 1]; ++o)
                        id[static_cast<size_t>(rec[o].t)] = static_cast<int32_t>(o);
                }
            };
            for (size_t j = 1; j < T; ++j) pool.emplace_back(fn, nv * j / T, nv * (j + 1) / T);
            fn(0, nv / T);
            for (auto& th : pool) th.join();
        }
        std::vector<std::thread> pool;
        auto out = [&](size_t lo, size_t hi) {
            for (size_t o = lo; o < hi; ++o) {
                const Rec& r = rec[o];
                const Tet& tt = T_[static_cast<size_t>(r.t)];
                for (int i = 0; i < 4; ++i) {
                    tets[4 * o + i] = orig_[static_cast<size_t>(r.v[i])];
                    // the neighbour across the face opposite r.v[i]
                    int j = 0;
                    while (tt.v[j] != r.v[i]) ++j;
                    nbrs[4 * o + 3 - i] = id[static_cast<size_t>(tt.n[j] >> 2)];
                }
            }
        };
        for (size_t j = 1; j < T; ++j) pool.emplace_back(out, nt * j / T, nt * (j + 1) / T);
        out(0, nt / T);
        for (auto& th : pool) th.join();
    }

    // Every finite tet positive; neighbour links symmetric and on shared faces.
    int check() const {
        for (size_t t = 0; t < top_; ++t) {
            if (!alive(t)) continue;
            const int32_t* v = T_[t].v;
            if (!infinite(t) && orient(pt(v[0]), pt(v[1]), pt(v[2]), pt(v[3])) <= 0) return 1;
            for (int i = 0; i < 4; ++i) {
                const int32_t l = T_[t].n[i];
                const size_t u = static_cast<size_t>(l >> 2);
                const int j = l & 3;
                if (!alive(u) || T_[u].n[j] != static_cast<int32_t>(4 * t + i)) return 2;
                int shared = 0;
                for (int a = 0; a < 4; ++a)
                    for (int b = 0; b < 4; ++b)
                        if (a != i && b != j && v[a] == T_[u].v[b]) ++shared;
                if (shared != 3) return 3;
            }
        }
        return 0;
    }

  private:
    // One tet: vertices (kInf = the infinite vertex), links n[i] = 4u + j to
    // the tet across the face opposite v[i], and its lock word: 0 free,
    // kDead deleted, else the token of the insertion that holds it (tok:
    // locked, tok + 1: in that insertion's cavity, tok + 2: tested outside).
    struct Tet {
        int32_t v[4];
        int32_t n[4];
        uint32_t lk;
    };
    struct Bnd {
        int32_t v[4];
        int32_t out;  // link of the outside tet's face: 4u + j
        int i;        // position of the vertex p replaces
    };
    struct Slot {
        uint32_t gen = 0;
        uint64_t key = 0;
        int32_t val = 0;
    };
    // Open-addressing table with generation stamps: entries of earlier
    // generations count as empty, so clearing is one increment.
    struct Table {
        std::vector<Slot> s = std::vector<Slot>(1024);
        int bits = 10;
        uint32_t gen = 0;
        void reset(size_t items) {
            ++gen;
            while (s.size() < 4u * items + 16u) {
                s.resize(s.size() * 2u);
                ++bits;
            }
        }
        // slot of key, or the empty slot where it would go
        Slot& find(uint64_t key) {
            const uint32_t mask = static_cast<uint32_t>(s.size() - 1u);
            for (uint32_t h = static_cast<uint32_t>((key * 0x9E3779B97F4A7C15ull) >> (64 - bits));;
                 h = (h + 1u) & mask) {
                Slot& x = s[h];
                if (x.gen != gen || x.key == key) return x;
            }
        }
    };
    // state of one inserting thread
    struct Worker {
        uint32_t id = 0, attempts = 0, rng = 1;
        std::vector<int32_t> free, cavity, locked, retry;
        std::vector<Bnd> bnd;
        Table edges;
        int32_t last = -1;
    };

    bool alive(size_t t) const { return get_lk(static_cast<int32_t>(t)) != kDead; }
    uint32_t get_lk(int32_t t) const { return __atomic_load_n(&T_[static_cast<size_t>(t)].lk, __ATOMIC_RELAXED); }
    void set_lk(int32_t t, uint32_t x) { __atomic_store_n(&T_[static_cast<size_t>(t)].lk, x, __ATOMIC_RELAXED); }
    bool infinite(size_t t) const {
        const int32_t* v = T_[t].v;
        return v[0] < 0 || v[1] < 0 || v[2] < 0 || v[3] < 0;
    }
    const P3& pt(int32_t v) const { return P_[static_cast<size_t>(v)]; }

    // a fresh tet slot, locked with tok: this worker's freed slots first
    int32_t new_tet(Worker& w, uint32_t tok) {
        int32_t t;
        if (!w.free.empty()) {
            t = w.free.back();
            w.free.pop_back();
        } else {
            const size_t x = top_.fetch_add(1);
            if (x >= cap_) return -1;
            t = static_cast<int32_t>(x);
        }
        set_lk(t, tok);
        w.locked.push_back(t);
        return t;
    }

    bool first_tet(Worker& w) {
        const int32_t a = 0, b = 1;
        int64_t ic = -1, id = -1;
        for (int64_t k = 2; k < nv_ && ic < 0; ++k) {
            const P3& q = pt(static_cast<int32_t>(k));
            const double ux = pt(b).x - pt(a).x, uy = pt(b).y - pt(a).y, uz = pt(b).z - pt(a).z;
            const double vx = q.x - pt(a).x, vy = q.y - pt(a).y, vz = q.z - pt(a).z;
            if (uy * vz - uz * vy != 0 || uz * vx - ux * vz != 0 || ux * vy - uy * vx != 0) ic = k;
        }
        if (ic < 0) return false;
        int32_t c = static_cast<int32_t>(ic);
        for (int64_t k = 2; k < nv_ && id < 0; ++k) {
            if (k == ic) continue;
            if (orient(pt(a), pt(b), pt(c), pt(static_cast<int32_t>(k))) != 0) id = k;
        }
        if (id < 0) return false;
        int32_t d = static_cast<int32_t>(id);
        if (orient(pt(a), pt(b), pt(c), pt(d)) < 0) std::swap(c, d);
        first_[0] = a;
        first_[1] = b;
        first_[2] = c;
        first_[3] = d;
        // the finite tet and, for each of its faces i, an infinite tet: the
        // finite tet with vertex i made infinite and two others swapped, so
        // that it is positive with the infinite vertex beyond face i
        const int32_t t0 = new_tet(w, 0u);
        const int32_t v0[4] = {a, b, c, d};
        for (int i = 0; i < 4; ++i) {
            T_[t0].v[i] = v0[i];
            vt_[static_cast<size_t>(v0[i])] = t0;
        }
        int32_t ti[4];
        for (int i = 0; i < 4; ++i) {
            ti[i] = new_tet(w, 0u);
            int32_t x[4] = {a, b, c, d};
            x[i] = kInf;
            std::swap(x[(i + 1) & 3], x[(i + 2) & 3]);
            for (int j = 0; j < 4; ++j) T_[ti[i]].v[j] = x[j];
        }
        w.locked.clear();
        // links by matching faces among the five tets
        const int32_t all[5] = {t0, ti[0], ti[1], ti[2], ti[3]};
        for (int x = 0; x < 5; ++x)
            for (int i = 0; i < 4; ++i) {
                bool found = false;
                for (int y = 0; y < 5 && !found; ++y) {
                    if (y == x) continue;
                    for (int j = 0; j < 4 && !found; ++j)
                        if (same_face(all[x], i, all[y], j)) {
                            T_[all[x]].n[i] = 4 * all[y] + j;
                            found = true;
                        }
                }
                if (!found) return false;
            }
        w.last = t0;
        return true;
    }

    bool same_face(int32_t t, int i, int32_t u, int j) const {
        int32_t f[3], g[3];
        int a = 0, b = 0;
        for (int k = 0; k < 4; ++k) {
            if (k != i) f[a++] = T_[t].v[k];
            if (k != j) g[b++] = T_[u].v[k];
        }
        std::sort(f, f + 3);
        std::sort(g, g + 3);
        return f[0] == g[0] && f[1] == g[1] && f[2] == g[2];
    }

    // orientation of tet t with its vertex at position i replaced by point p
    int orient_with(int32_t t, int i, const P3& p) const {
        const int32_t* v = T_[t].v;
        const P3* q[4];
        for (int k = 0; k < 4; ++k) q[k] = k == i ? &p : &P_[static_cast<size_t>(v[k])];
        return orient(*q[0], *q[1], *q[2], *q[3]);
    }

    // A tet near point p, for a walk whose last tet is gone: one incident to
    // the nearest inserted vertex before p in Morton order (or after it).
    // Every vertex of a deleted tet is a vertex of a tet that replaces it,
    // and each insertion records one new tet for every vertex it touches.
    int32_t hint(int32_t p) const {
        for (int32_t u = p - 1; u >= 0; --u) {
            const int32_t t = __atomic_load_n(&vt_[static_cast<size_t>(u)], __ATOMIC_RELAXED);
            if (t >= 0) return t;
        }
        for (int64_t u = p + 1; u < nv_; ++u) {
            const int32_t t = __atomic_load_n(&vt_[static_cast<size_t>(u)], __ATOMIC_RELAXED);
            if (t >= 0) return t;
        }
        return 0;
    }

    static uint32_t rnd(Worker& w) {
        w.rng ^= w.rng << 13;
        w.rng ^= w.rng >> 17;
        w.rng ^= w.rng << 5;
        return w.rng;
    }

    // Inserts point pid.  Every tet is read only after this insertion has
    // locked it (an acquire CAS from 0 to its token) and written only while
    // locked; the locks are released with release stores when the insertion
    // ends.  A lock held by another thread ends the attempt: it releases
    // what it holds and returns 1 (retry later), so no thread ever waits.
    // 0: inserted; 5: the point could not be placed.
    int insert(Worker& w, int32_t pid) {
        const P3& p = pt(pid);
        const uint32_t tok = 4u * (1u + w.id + static_cast<uint32_t>(W_.size()) * w.attempts++);
        w.locked.clear();
        w.cavity.clear();
        w.bnd.clear();
        auto lock = [&](int32_t t) -> int {  // 1 held by us, 0 busy, -1 deleted
            uint32_t cur = 0;
            if (__atomic_compare_exchange_n(&T_[static_cast<size_t>(t)].lk, &cur, tok, false, __ATOMIC_ACQUIRE,
                                            __ATOMIC_RELAXED)) {
                w.locked.push_back(t);
                return 1;
            }
            if (cur >= tok && cur <= tok + 2u) return 1;
            return cur == kDead ? -1 : 0;
        };
        auto release = [&]() {
            for (int32_t t : w.locked) {
                uint32_t* lk = &T_[static_cast<size_t>(t)].lk;
                const uint32_t x = __atomic_load_n(lk, __ATOMIC_RELAXED);
                if (x >= tok && x <= tok + 2u) __atomic_store_n(lk, 0u, __ATOMIC_RELEASE);
            }
            w.locked.clear();
        };
        // ---- walk to a tet in conflict with p
        int32_t t = w.last;
        int r = t >= 0 ? lock(t) : -1;
        if (r < 0) {
            t = hint(pid);
            r = lock(t);
        }
        if (r <= 0) {
            release();
            return 1;
        }
        if (infinite(static_cast<size_t>(t))) {
            int k = 0;
            while (T_[t].v[k] != kInf) ++k;
            t = T_[t].n[k] >> 2;
            if (lock(t) <= 0) {
                release();
                return 1;
            }
        }
        int from = -1;
        for (size_t steps = 0;; ++steps) {
            if (infinite(static_cast<size_t>(t))) break;
            const uint32_t rr = rnd(w);
            int next = -1;
            for (int j = 0; j < 4; ++j) {
                const int i = static_cast<int>((rr + static_cast<uint32_t>(j)) & 3u);
                if (i == from) continue;
                if (orient_with(t, i, p) < 0) {
                    next = i;
                    break;
                }
            }
            if (next < 0) break;
            const int32_t l = T_[t].n[next];
            t = l >> 2;
            from = l & 3;
            if (lock(t) <= 0 || steps > 1000000) {
                release();
                return steps > 1000000 ? 5 : 1;
            }
        }
        // ---- the cavity: tets whose circumsphere contains p
        int c = conflict(t, p, lock);
        if (c < 0) {
            release();
            return 1;
        }
        if (c == 0) {
            release();
            return 5;
        }
        set_lk(t, tok + 1u);
        w.cavity.push_back(t);
        for (size_t q = 0; q < w.cavity.size(); ++q) {
            const int32_t a = w.cavity[q];
            for (int i = 0; i < 4; ++i) {
                const int32_t l = T_[a].n[i];
                const int32_t u = l >> 2;
                const uint32_t lu = get_lk(u);
                if (lu == tok + 1u) continue;
                if (lu != tok + 2u) {
                    if (lu != tok && lock(u) <= 0) {
                        release();
                        return 1;
                    }
                    c = conflict(u, p, lock);
                    if (c < 0) {
                        release();
                        return 1;
                    }
                    if (c > 0) {
                        set_lk(u, tok + 1u);
                        w.cavity.push_back(u);
                        continue;
                    }
                    set_lk(u, tok + 2u);
                }
                const Tet& ta = T_[a];
                w.bnd.push_back(Bnd{{ta.v[0], ta.v[1], ta.v[2], ta.v[3]}, l, i});
            }
        }
        // ---- the cone from p over the cavity's boundary; the three faces of
        // a new tet through p are matched by the boundary edge they contain
        for (int32_t a : w.cavity) {
            set_lk(a, kDead);
            w.free.push_back(a);
        }
        w.edges.reset(3u * w.bnd.size() / 2u + 1u);
        int32_t last = -1;
        for (const Bnd& b : w.bnd) {
            const int32_t nt = new_tet(w, tok);
            if (nt < 0) {
                release();
                return 6;
            }
            Tet& T = T_[static_cast<size_t>(nt)];
            for (int k = 0; k < 4; ++k) T.v[k] = b.v[k];
            T.v[b.i] = pid;
            T.n[b.i] = b.out;
            T_[static_cast<size_t>(b.out >> 2)].n[b.out & 3] = 4 * nt + b.i;
            for (int k = 0; k < 4; ++k)
                if (T.v[k] >= 0) __atomic_store_n(&vt_[static_cast<size_t>(T.v[k])], nt, __ATOMIC_RELAXED);
            for (int k = 0; k < 4; ++k) {
                if (k == b.i) continue;
                int32_t e0 = -2, e1 = -2;
                for (int m = 0; m < 4; ++m) {
                    if (m == b.i || m == k) continue;
                    if (e0 == -2) e0 = T.v[m]; else e1 = T.v[m];
                }
                if (e0 > e1) std::swap(e0, e1);
                const uint64_t key = (static_cast<uint64_t>(static_cast<uint32_t>(e0)) << 32) |
                                     static_cast<uint32_t>(e1);
                Slot& x = w.edges.find(key);
                if (x.gen != w.edges.gen) {
                    x.gen = w.edges.gen;
                    x.key = key;
                    x.val = 4 * nt + k;
                } else {
                    T.n[k] = x.val;
                    T_[static_cast<size_t>(x.val >> 2)].n[x.val & 3] = 4 * nt + k;
                    x.key = ~0ull;  // an edge is shared by exactly two new tets
                }
            }
            if (last < 0 || !infinite(static_cast<size_t>(nt))) last = nt;
        }
        w.last = last;
        release();
        return 0;
    }

    // Whether p lies in the open circumsphere of tet t (held by this
    // insertion): 1 yes, 0 no, -1 another thread holds a tet it must read.
    template <class Lock>
    int conflict(int32_t t, const P3& p, Lock& lock) const {
        const int32_t* v = T_[t].v;
        int k = -1;
        for (int i = 0; i < 4; ++i)
            if (v[i] == kInf) k = i;
        if (k < 0) return insphere(pt(v[0]), pt(v[1]), pt(v[2]), pt(v[3]), p) > 0 ? 1 : 0;
        const int o = orient_with(t, k, p);
        if (o != 0) return o > 0 ? 1 : 0;
        // p on the plane of the hull face: inside its circumcircle, i.e.
        // inside the circumsphere of the finite tet on the other side
        const int32_t u = T_[t].n[k] >> 2;
        if (lock(u) <= 0) return -1;
        const int32_t* w = T_[u].v;
        return insphere(pt(w[0]), pt(w[1]), pt(w[2]), pt(w[3]), p) > 0 ? 1 : 0;
    }

    static constexpr uint32_t kDead = 0xFFFFFFFFu;

    int64_t n_ = 0, nv_ = 0, repeated_ = 0;
    std::vector<P3> P_;          // points in Morton order
    std::vector<int32_t> orig_;  // caller's id of each point
    std::unique_ptr<Tet[]> T_;
    std::vector<int32_t> vt_;    // a tet incident to each inserted vertex, -1 before
    std::vector<Worker> W_;
    size_t cap_ = 0;
    std::atomic<size_t> top_{0};
    int32_t first_[4] = {-1, -1, -1, -1};
};

}  // namespace

extern "C" {

// Triangulates n points with integer coordinates in [0, 2^21) (xyz int32[n*3])
// on `threads` threads (<= 1: one).  Repeated points are used once: the
// first copy is a vertex, the others stay unused (*repeated counts them).
// Returns a handle, or NULL with *status set: 1 bad coordinate or n beyond
// 2^30, 2 fewer than 4 distinct points, 4 all coplanar, 5 a point could not
// be inserted, 6 more tets than the 8 per point provisioned.  *n_tets is the
// finite tet count.
void* wtvd_triangulate(const int32_t* xyz, int64_t n, int threads, int64_t* n_tets, int64_t* repeated,
                       int* status) {
    auto* d = new Delaunay;
    const int rc = d->run(xyz, n, threads);
    if (status) *status = rc;
    if (rc != 0) {
        delete d;
        return nullptr;
    }
    if (n_tets) *n_tets = d->finite_tets();
    if (repeated) *repeated = d->repeated();
    return d;
}

// Writes the finite tets (int32[n_tets*4], caller's vertex ids, canonical
// order) and their 4-slot neighbours (int32[n_tets*4], slot fi across face
// {012,013,023,123}[fi], -1 on the hull).
void wtvd_export(const void* h, int32_t* tets, int32_t* nbrs) {
    const auto* d = static_cast<const Delaunay*>(h);
    d->export_mesh(tets, nbrs, d->threads());
}

// 0 when every finite tet is positively oriented and every neighbour link is
// symmetric and joins tets on a shared face.
int wtvd_check(const void* h) { return static_cast<const Delaunay*>(h)->check(); }

void wtvd_free(void* h) { delete static_cast<Delaunay*>(h); }

}  // extern "C"
