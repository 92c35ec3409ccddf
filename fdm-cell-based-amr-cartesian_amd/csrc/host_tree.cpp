// host_tree.cpp -- quadtree topology, mesh operations, leaf lists and initial conditions.
// Each routine cites the reference routine whose behaviour it keeps. Neighbour finding must
// reproduce the reference bit for bit (it decides stencil shapes), so the finders follow
// the reference's case analysis exactly, including its asymmetric diagonal null rules.
#include "host_tree.hpp"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include <shared_mutex>
#include <stdexcept>
#include <exception>
#include <thread>

#include <sys/mman.h>

namespace fdm {

static inline double pow2i(int k) { return (double)(1 << k); }

uint64_t splitmix64(uint64_t& s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
// host threads of the regrid-time work: the hardware's, at most 16, or FDMAMR_HOST_THREADS
static unsigned host_threads() {
    unsigned nt = std::thread::hardware_concurrency();
    if (const char* e = std::getenv("FDMAMR_HOST_THREADS")) nt = (unsigned)std::max(1, std::atoi(e));
    return std::min<unsigned>(std::max(nt, 1u), 16u);
}

void parallel_for(size_t n, const std::function<void(size_t, size_t)>& f, size_t serial_below) {
    const unsigned nt = host_threads();
    if (n < serial_below || n < 2 || nt == 1) { f(0, n); return; }
    std::vector<std::thread> th;
    std::vector<std::exception_ptr> errs(nt);
    const size_t chunk = (n + nt - 1) / nt;
    for (unsigned k = 0; k < nt; k++) {
        const size_t b = k * chunk, e = std::min(n, b + chunk);
        if (b < e)
            th.emplace_back([&f, &errs, k, b, e] {
                try { f(b, e); } catch (...) { errs[k] = std::current_exception(); }
            });
    }
    for (auto& t : th) t.join();
    for (auto& x : errs)
        if (x) std::rethrow_exception(x);
}

void* big_alloc(size_t bytes) {
    void* p = nullptr;
    if (posix_memalign(&p, (size_t)2 << 20, bytes) != 0 || !p) throw std::bad_alloc();
    (void)madvise(p, bytes, MADV_HUGEPAGE);  // advice only: ignored where huge pages are off
    return p;
}
void big_free(void* p) { std::free(p); }

static double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}
PhaseTimer::PhaseTimer() : on(std::getenv("FDMAMR_TIMING") != nullptr), t0(now_ms()) {}
void PhaseTimer::mark(const char* phase) {
    if (!on) return;
    const double t = now_ms();
    std::fprintf(stderr, "[fdmamr] %-28s %9.1f ms\n", phase, t - t0);
    t0 = t;
}

static double u01(uint64_t& s) { return (double)(splitmix64(s) >> 11) * 0x1.0p-53; }
static double uab(uint64_t& s, double a, double b) { return a + (b - a) * u01(s); }

void Tree::configure(double xmin, double xmax, double ymin, double ymax, double h_,
                     double gamma_, double Ma_) {
    h = h_;
    xf = 1.0;
    gamma = gamma_;
    Ma = Ma_;
    pxmin = xmin; pxmax = xmax; pymin = ymin; pymax = ymax;
    Xmin = xmin - 3.0 * xf * h;                     // MAIN.F90:55-56
    Xmax = xmax + 3.0 * xf * h + 1.0e-10;
    Ymin = ymin - 3.0 * h;
    Ymax = ymax + 3.0 * h + 1.0e-10;
    cv = 1.0 / (gamma * (gamma - 1.0) * Ma * Ma);   // MAIN.F90:61
}

int Tree::gridcross(int a, int b) const {           // mesh.F90:89-142 (gridsort == 0)
    if (a < 4 && b < 4) return -5;
    if (a < 4 && b > m - 3) return -5;
    if (a > n - 3 && b > m - 3) return -5;
    if (a > n - 3 && b < 4) return -5;
    if (a < 4) return -1;
    if (a > n - 3) return -3;
    if (b > m - 3) return -2;
    if (b < 4) return -4;
    return 0;
}

void Tree::creat_mesh() {                           // mesh.F90:9-47
    m = (int)((Xmax - Xmin) / (xf * h));
    n = (int)((Ymax - Ymin) / h);
    total = m * n;
    father.assign(total, -1); son.assign(total, -1); lvl.assign(total, 0); sp.assign(total, 0);
    ni.resize(total); nj.resize(total); cross.resize(total);
    x.resize(total); y.resize(total);
    U.assign(total, {0, 0, 0, 0, 0, 0, 0});
    rot.assign(total, 0.0); div.assign(total, 0.0);
    fresh.assign(total, 0);
    free_blocks.clear();
    for (int i = 1; i <= total; i++) {
        int a = (i - 1) / m + 1;
        int b = i - m * (a - 1);
        int k = i - 1;
        ni[k] = a; nj[k] = b;
        x[k] = 0.5 * (2.0 * b - 1.0) * xf * h + Xmin;
        y[k] = (a - 0.5) * h + Ymin;
        cross[k] = gridcross(a, b);
    }
    idneighbor();
}

// Node allocation of a parallel pass. The node arrays extend past the nodes in use (`used`);
// son blocks come from the free list or the tail, under `mu`. Workers hold `grow` shared while
// they process cells. A worker that finds the tail exhausted releases its share and takes
// `grow` exclusively, which waits until no other worker is inside a cell, then resizes the
// arrays; no reference into them is held across an allocation.
struct PassPool {
    std::mutex mu;
    std::shared_mutex grow;
    size_t used = 0;
};
// Each worker takes son blocks from the pool a few at a time into its own cache; a worker
// returns the blocks it did not use to the free list, marked released, at the end of the pass.
static thread_local std::vector<int32_t>* tl_blocks = nullptr;
// ... and counts its refines and coarsens on its own, adding them to Tree::mod_count at the end
static thread_local long* tl_mods = nullptr;
constexpr int kBlockBatch = 64;

void Tree::resize_nodes(size_t nn) {
    if (nn > (size_t)INT_MAX) throw std::runtime_error("quadtree node count exceeds 2^31");
    if (nn > father.capacity()) {
        // with room to grow, so that further passes do not copy the arrays again; the copy
        // into the new storage runs on all host threads
        const size_t cap = nn + nn / 4;
        auto grow = [&](auto& v) {
            std::decay_t<decltype(v)> w;
            w.reserve(cap);
            w.resize(v.size());
            using T = typename std::decay_t<decltype(v)>::value_type;
            parallel_for(v.size(), [&](size_t b, size_t e) { std::memcpy(w.data() + b, v.data() + b, (e - b) * sizeof(T)); });
            v.swap(w);
        };
        grow(father); grow(son); grow(lvl); grow(sp); grow(ni); grow(nj); grow(cross);
        grow(x); grow(y); grow(U); grow(rot); grow(div); grow(fresh);
    }
    father.resize(nn); son.resize(nn); lvl.resize(nn); sp.resize(nn); ni.resize(nn);
    nj.resize(nn); cross.resize(nn); x.resize(nn); y.resize(nn); U.resize(nn); rot.resize(nn);
    div.resize(nn); fresh.resize(nn);
}

int Tree::alloc_block() {
    if (pool) {
        std::vector<int32_t>& cache = *tl_blocks;
        for (;;) {
            if (!cache.empty()) {
                int b = cache.back();
                cache.pop_back();
                return b;
            }
            {
                std::lock_guard<std::mutex> g(pool->mu);
                while (!free_blocks.empty() && cache.size() < kBlockBatch) {
                    cache.push_back(free_blocks.back());
                    free_blocks.pop_back();
                }
                while (cache.size() < kBlockBatch && pool->used + 4 <= father.size()) {
                    cache.push_back((int32_t)pool->used);
                    pool->used += 4;
                }
                if (!cache.empty()) {
                    // hand out the lowest id first, as a serial allocation would
                    std::reverse(cache.begin(), cache.end());
                    continue;
                }
            }
            pool->grow.unlock_shared();
            std::exception_ptr err;
            {
                std::unique_lock<std::shared_mutex> ex(pool->grow);
                try {
                    if (pool->used + 4 > father.size())
                        resize_nodes(father.size() + std::max<size_t>(father.size() / 8, (size_t)1 << 20));
                } catch (...) {
                    err = std::current_exception();
                }
            }
            pool->grow.lock_shared();
            if (err) std::rethrow_exception(err);
        }
    }
    if (!free_blocks.empty()) {
        int b = free_blocks.back();
        free_blocks.pop_back();
        return b;
    }
    int b = (int)father.size();
    resize_nodes(father.size() + 4);
    return b;
}

void Tree::release_subtree(int c) {
    int s = son[c];
    if (s < 0) return;
    for (int k = 0; k < 4; k++) release_subtree(s + k);
    for (int k = 0; k < 4; k++) { father[s + k] = -2; son[s + k] = -1; }
    son[c] = -1;
    // in a parallel pass the block goes to the worker's own cache, for its next allocations;
    // what the worker has not used at the end of the pass returns to the free list
    if (pool) tl_blocks->push_back(s);
    else free_blocks.push_back(s);
}

void Tree::compact() {
    PhaseTimer pt;
    const size_t NA = n_alloc();
    // nodes below each base cell, then their first new id
    std::vector<int64_t> first((size_t)total + 1, 0);
    parallel_for((size_t)total, [&](size_t b, size_t e) {
        std::vector<int32_t> st;
        for (size_t i = b; i < e; i++) {
            int64_t cnt = 0;
            st.assign(1, (int32_t)i);
            while (!st.empty()) {
                const int c = st.back();
                st.pop_back();
                if (son[c] < 0) continue;
                cnt += 4;
                for (int k = 0; k < 4; k++) st.push_back(son[c] + k);
            }
            first[i + 1] = cnt;
        }
    });
    pt.mark("compact: count");
    first[0] = total;
    for (int i = 0; i < total; i++) first[i + 1] += first[i];
    const size_t nn = (size_t)first[total];
    if (nn > (size_t)INT_MAX) throw std::runtime_error("quadtree node count exceeds 2^31");
    BigVec<int32_t> id(NA);  // -1 for released nodes
    parallel_for(NA, [&](size_t b, size_t e) { std::fill(id.begin() + b, id.begin() + e, -1); });
    // block pre-order below each base cell: a node's four sons, then each son's subtree
    parallel_for((size_t)total, [&](size_t b, size_t e) {
        std::vector<int32_t> st;
        for (size_t i = b; i < e; i++) {
            int32_t next = (int32_t)first[i];
            id[i] = (int32_t)i;
            st.assign(1, (int32_t)i);
            while (!st.empty()) {
                const int c = st.back();
                st.pop_back();
                if (son[c] < 0) continue;
                for (int k = 0; k < 4; k++) id[son[c] + k] = next++;
                for (int k = 3; k >= 0; k--) st.push_back(son[c] + k);
            }
        }
    });
    pt.mark("compact: ids");
    BigVec<int32_t> old(nn);  // every entry is written: each new id has one old node
    parallel_for(NA, [&](size_t b, size_t e) {
        for (size_t c = b; c < e; c++)
            if (id[c] >= 0) old[id[c]] = (int32_t)c;
    });
    // in place through a spare buffer kept for the next regrid (the arrays keep their capacity)
    auto permute = [&](auto& v, auto map) {
        using T = typename std::decay_t<decltype(v)>::value_type;
        if (spare_bytes < NA * sizeof(T)) {
            spare.reset();
            spare_bytes = std::max(NA * sizeof(T) + NA * sizeof(T) / 4, BIG_ALLOC_MIN);
            // not zeroed: first touched in parallel
            spare = std::shared_ptr<unsigned char[]>(static_cast<unsigned char*>(big_alloc(spare_bytes)), big_free);
        }
        T* src = reinterpret_cast<T*>(spare.get());
        parallel_for(NA, [&](size_t b, size_t e) { std::memcpy(src + b, v.data() + b, (e - b) * sizeof(T)); });
        parallel_for(nn, [&](size_t b, size_t e) {
            for (size_t k = b; k < e; k++) v[k] = map(src[old[k]]);
        });
        v.resize(nn);
    };
    auto same = [](auto a) { return a; };
    auto link = [&](int32_t a) { return a < 0 ? a : id[a]; };
    permute(father, link);
    permute(son, link);
    permute(lvl, same); permute(sp, same); permute(ni, same); permute(nj, same); permute(cross, same);
    pt.mark("compact: links");
    permute(x, same); permute(y, same); permute(U, same); permute(rot, same); permute(div, same);
    permute(fresh, same);
    pt.mark("compact: values");
    free_blocks.clear();
    nb.clear();
}

// Run f(cell) for every base cell on host threads, in chunks of consecutive cells taken from a
// shared counter, for passes whose cells are independent (each reads and writes only its own
// subtree), so any order gives the tree a serial loop gives. New son blocks come from a
// PassPool whose tail starts with room for `room` nodes.
template <class F>
static void run_cells(Tree& t, unsigned nt, size_t room, const F& f) {
    constexpr int kChunk = 32;
    PassPool pool;
    pool.used = t.n_alloc();
    PhaseTimer pt;
    t.resize_nodes(pool.used + std::max(room, (size_t)1 << 16));
    pt.mark("  pass: node arrays");
    t.pool = &pool;
    std::atomic<int> next{0};
    std::atomic<bool> abort{false};
    std::vector<std::exception_ptr> errs(nt);
    auto worker = [&](unsigned w) {
        std::vector<int32_t> cache;
        long mods = 0;
        tl_blocks = &cache;
        tl_mods = &mods;
        for (;;) {
            const int k0 = next.fetch_add(kChunk, std::memory_order_relaxed);
            if (k0 >= t.total || abort.load(std::memory_order_relaxed)) break;
            pool.grow.lock_shared();
            try {
                for (int k = k0; k < std::min(t.total, k0 + kChunk); k++) f(k);
            } catch (...) {
                errs[w] = std::current_exception();
                abort = true;
            }
            pool.grow.unlock_shared();
        }
        {
            std::lock_guard<std::mutex> g(pool.mu);
            for (int32_t b : cache) {
                for (int k = 0; k < 4; k++) { t.father[b + k] = -2; t.son[b + k] = -1; }
                t.free_blocks.push_back(b);
            }
        }
        tl_blocks = nullptr;
        tl_mods = nullptr;
        if (t.mod_count) t.mod_count->fetch_add(mods, std::memory_order_relaxed);
    };
    std::vector<std::thread> th;
    for (unsigned w = 1; w < nt; w++) th.emplace_back(worker, w);
    worker(0);
    for (auto& x : th) x.join();
    pt.mark("  pass: cells");
    t.pool = nullptr;
    t.resize_nodes(pool.used);
    for (auto& e : errs)
        if (e) std::rethrow_exception(e);
}

// Run visit(cell) for the base cells that want(cell) selects, on host threads, row by row: a
// worker takes the next row and walks it by increasing column, and starts cell (r, c) once row
// r - 1 has finished its columns up to c + lag. With lag = 0, every cell that shares an edge
// with an earlier cell of the row-major order starts after it; with lag = 2, every cell of an
// earlier row in the same column within four rows, or within two rows and two columns, does
// (row r - 1 at c + 2 itself waited for row r - 2 at c + 4, and so on). Cells that interact
// only in those ways then see the tree exactly as in a serial loop over the cells in index
// order. want() reads only the per-cell level bounds; visit() runs while the worker holds the
// pool's growth lock shared, and never waits for another worker. Returns the number of cells
// visited.
template <class W, class V>
static size_t run_rows(Tree& t, int lag, unsigned nt, size_t room, const W& want, const V& visit) {
    const int m = t.m, n = t.n;
    constexpr int kPublish = 16;  // skipped columns between progress stores (a visit always stores)
    PassPool pool;
    pool.used = t.n_alloc();
    t.resize_nodes(pool.used + std::max(room, (size_t)1 << 16));
    t.pool = &pool;
    std::unique_ptr<std::atomic<int>[]> done(new std::atomic<int>[(size_t)n]);
    for (int r = 0; r < n; r++) done[r].store(0, std::memory_order_relaxed);
    std::atomic<int> next_row{0};
    std::atomic<bool> abort{false};
    std::atomic<size_t> visited{0};
    std::vector<std::exception_ptr> errs(nt);
    // with FDMAMR_TIMING: each worker's time inside visits
    const bool timing = std::getenv("FDMAMR_TIMING") != nullptr;
    std::vector<double> busy(nt, 0.0);
    const double t_start = timing ? now_ms() : 0.0;
    auto worker = [&](unsigned w) {
        std::vector<int32_t> cache;
        long mods = 0;
        size_t visits = 0;
        tl_blocks = &cache;
        tl_mods = &mods;
        for (;;) {
            const int r = next_row.fetch_add(1, std::memory_order_relaxed);
            if (r >= n) break;
            int seen = r == 0 ? m : 0;  // columns of row r - 1 known to be finished
            for (int c = 0; c < m; c++) {
                const int need = std::min(m, c + lag + 1);
                for (int spin = 0; seen < need; spin++) {
                    seen = done[r - 1].load(std::memory_order_acquire);
                    if (seen < need && spin > 2000) std::this_thread::yield();
                }
                const int i = r * m + c;
                bool visited_cell = false;
                if (!abort.load(std::memory_order_relaxed) && want(i)) {
                    pool.grow.lock_shared();
                    visits++;
                    visited_cell = true;
                    const double v0 = timing ? now_ms() : 0.0;
                    try {
                        visit(i);
                    } catch (...) {
                        errs[w] = std::current_exception();
                        abort = true;
                    }
                    if (timing) busy[w] += now_ms() - v0;
                    pool.grow.unlock_shared();
                }
                // after a visit at once, so that the next row follows a slow stretch closely
                if (visited_cell || (c + 1) % kPublish == 0 || c + 1 == m)
                    done[r].store(c + 1, std::memory_order_release);
            }
        }
        {
            std::lock_guard<std::mutex> g(pool.mu);
            for (int32_t b : cache) {
                for (int k = 0; k < 4; k++) { t.father[b + k] = -2; t.son[b + k] = -1; }
                t.free_blocks.push_back(b);
            }
        }
        tl_blocks = nullptr;
        tl_mods = nullptr;
        if (t.mod_count) t.mod_count->fetch_add(mods, std::memory_order_relaxed);
        visited += visits;
    };
    std::vector<std::thread> th;
    for (unsigned w = 1; w < nt; w++) th.emplace_back(worker, w);
    worker(0);
    for (auto& x : th) x.join();
    if (timing) {
        double sum = 0.0, mx = 0.0;
        for (double b : busy) { sum += b; mx = std::max(mx, b); }
        std::fprintf(stderr, "[fdmamr]   rows: %.1f ms, workers inside visits %.1f ms in all, at most %.1f ms each\n",
                     now_ms() - t_start, sum, mx);
    }
    t.pool = nullptr;
    t.resize_nodes(pool.used);
    for (auto& e : errs)
        if (e) std::rethrow_exception(e);
    return visited.load();
}

// parallel passes only on meshes with enough base cells to share out
static unsigned pass_threads(const Tree& t) {
    const unsigned nt = host_threads();
    return t.total >= 4096 ? nt : 1;
}

void Tree::refine(int c) {                          // mesh.F90:199-260
    // Refining a node that already has sons replaces its subtree (the reference overwrites
    // the son pointers, mesh.F90:203); the old subtree is unreachable afterwards.
    if (son[c] >= 0) release_subtree(c);
    int s = alloc_block();
    son[c] = s;
    double ox = xf * h / pow2i(lvl[c] + 2), oy = h / pow2i(lvl[c] + 2);
    static const int sx[4] = {-1, 1, 1, -1}, sy[4] = {-1, -1, 1, 1};
    for (int k = 0; k < 4; k++) {
        int q = s + k;
        father[q] = c; son[q] = -1; ni[q] = ni[c]; nj[q] = nj[c];
        lvl[q] = lvl[c] + 1; sp[q] = k + 1;
        x[q] = sx[k] < 0 ? x[c] - ox : x[c] + ox;
        y[q] = sy[k] < 0 ? y[c] - oy : y[c] + oy;
        cross[q] = gridcross(ni[q], nj[q]);
        U[q] = U[c];
        rot[q] = rot[c]; div[q] = div[c];
        fresh[q] = 1;
    }
    if (mod_count) {
        if (tl_mods) ++*tl_mods;
        else mod_count->fetch_add(1, std::memory_order_relaxed);
    }
    if (track_levels) {
        // the new leaves are at lvl + 1; a replaced subtree may have held only deeper ones
        const size_t cell = (size_t)(ni[c] - 1) * m + nj[c] - 1;
        cell_lmax[cell] = std::max<int8_t>(cell_lmax[cell], (int8_t)(lvl[c] + 1));
        cell_lmin[cell] = std::min<int8_t>(cell_lmin[cell], (int8_t)(lvl[c] + 1));
        cell_dirty[cell] = 1;
        cell_mod[cell] = mod_phase;
    }
}

void Tree::restrict_mean(int c, double* var) const { // XuanGua.F90:359-398
    if (son[c] < 0) { for (int i = 0; i < 7; i++) var[i] = U[c][i]; return; }
    double sv[4][7];
    for (int k = 0; k < 4; k++) {
        int q = son[c] + k;
        if (son[q] >= 0) restrict_mean(q, sv[k]);
        else for (int i = 0; i < 7; i++) sv[k][i] = U[q][i];
    }
    for (int i = 0; i < 7; i++) var[i] = 0.25 * (((sv[0][i] + sv[1][i]) + sv[2][i]) + sv[3][i]);
}

void Tree::coarse(int c) {                          // mesh.F90:397-408
    double v[7];
    restrict_mean(c, v);
    for (int i = 0; i < 7; i++) U[c][i] = v[i];
    release_subtree(c);
    fresh[c] = 1;
    if (mod_count) {
        if (tl_mods) ++*tl_mods;
        else mod_count->fetch_add(1, std::memory_order_relaxed);
    }
    if (track_levels) {
        const size_t cell = (size_t)(ni[c] - 1) * m + nj[c] - 1;
        cell_lmin[cell] = std::min<int8_t>(cell_lmin[cell], (int8_t)lvl[c]);
        cell_dirty[cell] = 1;
        cell_mod[cell] = mod_phase;
    }
}

static void cell_levels(const Tree& t, int i, std::vector<int32_t>& st, int& lo, int& hi) {
    lo = 127; hi = 0;
    st.assign(1, (int32_t)i);
    while (!st.empty()) {
        const int c = st.back();
        st.pop_back();
        if (t.son[c] < 0) { lo = std::min(lo, t.lvl[c]); hi = std::max(hi, t.lvl[c]); continue; }
        for (int k = 0; k < 4; k++) st.push_back(t.son[c] + k);
    }
}
void Tree::leaf_levels() {
    cell_lmin.assign(total, 127);
    cell_lmax.assign(total, 0);
    cell_dirty.assign(total, 0);
    cell_mod.assign(total, INT32_MIN);
    parallel_for((size_t)total, [&](size_t b, size_t e) {
        std::vector<int32_t> st;
        for (size_t i = b; i < e; i++) {
            int lo, hi;
            cell_levels(*this, (int)i, st, lo, hi);
            cell_lmin[i] = (int8_t)lo;
            cell_lmax[i] = (int8_t)hi;
        }
    });
}
void Tree::leaf_levels_dirty() {
    parallel_for((size_t)total, [&](size_t b, size_t e) {
        std::vector<int32_t> st;
        for (size_t i = b; i < e; i++) {
            if (!cell_dirty[i]) continue;
            int lo, hi;
            cell_levels(*this, (int)i, st, lo, hi);
            cell_lmin[i] = (int8_t)lo;
            cell_lmax[i] = (int8_t)hi;
            cell_dirty[i] = 0;
        }
    });
}

// ---------------------------------------------------------------- neighbours
int Tree::init_up(int c) const { return ni[c] == n ? -1 : ni[c] * m + nj[c] - 1; }
int Tree::init_down(int c) const { return ni[c] == 1 ? -1 : ni[c] * m - 2 * m + nj[c] - 1; }
int Tree::init_left(int c) const { return nj[c] == 1 ? -1 : ni[c] * m - m + nj[c] - 2; }
int Tree::init_right(int c) const { return nj[c] == m ? -1 : ni[c] * m - m + nj[c]; }

int Tree::up(int c) const {                         // neighbor.F90:65-96
    int f = father[c];
    if (f < 0) return init_up(c);
    if (sp[c] == 1) return son[f] + 3;
    if (sp[c] == 2) return son[f] + 2;
    int fn = up(f);
    if (fn >= 0) {
        if (son[fn] < 0) return fn;
        if (sp[c] == 4) return son[fn] + 0;
        if (sp[c] == 3) return son[fn] + 1;
    }
    return -1;
}
int Tree::down(int c) const {                       // neighbor.F90:98-129
    int f = father[c];
    if (f < 0) return init_down(c);
    if (sp[c] == 4) return son[f] + 0;
    if (sp[c] == 3) return son[f] + 1;
    int fn = down(f);
    if (fn >= 0) {
        if (son[fn] < 0) return fn;
        if (sp[c] == 1) return son[fn] + 3;
        if (sp[c] == 2) return son[fn] + 2;
    }
    return -1;
}
int Tree::left(int c) const {                       // neighbor.F90:131-162
    int f = father[c];
    if (f < 0) return init_left(c);
    if (sp[c] == 2) return son[f] + 0;
    if (sp[c] == 3) return son[f] + 3;
    int fn = left(f);
    if (fn >= 0) {
        if (son[fn] < 0) return fn;
        if (sp[c] == 4) return son[fn] + 2;
        if (sp[c] == 1) return son[fn] + 1;
    }
    return -1;
}
int Tree::right(int c) const {                      // neighbor.F90:164-196
    int f = father[c];
    if (f < 0) return init_right(c);
    if (sp[c] == 1) return son[f] + 1;
    if (sp[c] == 4) return son[f] + 2;
    int fn = right(f);
    if (fn >= 0) {
        if (son[fn] < 0) return fn;
        if (sp[c] == 2) return son[fn] + 0;
        if (sp[c] == 3) return son[fn] + 3;
    }
    return -1;
}

// Diagonal finders: one table-driven routine for the four corners. For corner (V,H) with
// V in {up,down} and H in {left,right} the reference's cases are:
//   base cell                 -> null if H or V neighbour is null, else V(H(c))
//   sp == inner               -> a sibling
//   sp == "descend" position  -> V(H(c)) then descend through son `ds` while coarser
//   sp == "H-side" position   -> null if H(c) null or coarser, else V(H(c))
//   sp == "V-side" position   -> null if V(c) null or coarser, else H(V(c))
struct DiagRule { int sib_sp, sib_son, desc_sp, desc_son, hside_sp, vside_sp; };
// upleft (neighbor.F90:199-246), upright (:249-296), downright (:299-346), downleft (:349-396)
static const DiagRule kUL = {2, 3, 4, 1, 1, 3};
static const DiagRule kUR = {1, 2, 3, 0, 2, 4};
static const DiagRule kDR = {4, 1, 2, 3, 3, 1};
static const DiagRule kDL = {3, 0, 1, 2, 4, 2};

// One corner's finder given the node's own H and V neighbours hn, vn and the face finders H, V
// of other nodes (the recursive finders, or idnei's stored results)
template <class HF, class VF>
static int diag(const Tree& t, const DiagRule& R, int c, int hn, int vn, const HF& H, const VF& V) {
    const int f = t.father[c];
    if (f < 0) return (hn < 0 || vn < 0) ? -1 : V(hn);
    if (t.sp[c] == R.sib_sp) return t.son[f] + R.sib_son;
    if (t.sp[c] == R.desc_sp) {
        if (hn < 0 || vn < 0) return -1;
        int r = V(hn);
        while (r >= 0 && t.lvl[r] < t.lvl[c] && t.son[r] >= 0) r = t.son[r] + R.desc_son;
        return r;
    }
    if (t.sp[c] == R.hside_sp) return (hn < 0 || t.lvl[hn] < t.lvl[c]) ? -1 : V(hn);
    if (t.sp[c] == R.vside_sp) return (vn < 0 || t.lvl[vn] < t.lvl[c]) ? -1 : H(vn);
    return -1;
}
int Tree::upleft(int c) const {
    return diag(*this, kUL, c, left(c), up(c), [&](int q) { return left(q); }, [&](int q) { return up(q); });
}
int Tree::upright(int c) const {
    return diag(*this, kUR, c, right(c), up(c), [&](int q) { return right(q); }, [&](int q) { return up(q); });
}
int Tree::downright(int c) const {
    return diag(*this, kDR, c, right(c), down(c), [&](int q) { return right(q); }, [&](int q) { return down(q); });
}
int Tree::downleft(int c) const {
    return diag(*this, kDL, c, left(c), down(c), [&](int q) { return left(q); }, [&](int q) { return down(q); });
}

// The face finders' recursion (neighbor.F90:65-196) as a table: for a son in position sp of
// its father, the neighbour in direction dir is the sibling kSib[dir][sp] when that is >= 0;
// otherwise it follows from the father's neighbour fn in that direction: null if fn is null,
// fn itself if fn is a leaf, else fn's son kAcross[dir][sp]. Directions in Nb order (up, down,
// right, left), positions 1..4 (index 0 unused).
static const int8_t kSib[4][5] = {{-1, 3, 2, -1, -1}, {-1, -1, -1, 1, 0}, {-1, 1, -1, -1, 2}, {-1, -1, 0, 3, -1}};
static const int8_t kAcross[4][5] = {{-1, -1, -1, 1, 0}, {-1, 3, 2, -1, -1}, {-1, -1, 0, 3, -1}, {-1, 1, -1, -1, 2}};

// idnei (neighbor.F90:412-449) evaluated top-down: a node's face neighbours come from its
// father's stored ones (the same values the recursive finders return), base cells' from the
// base grid
void Tree::idnei_faces(int c) {
    int32_t* q = &nb[8 * (size_t)c];
    const int f = father[c];
    if (f < 0) {
        q[NB_UP] = init_up(c); q[NB_DOWN] = init_down(c); q[NB_RIGHT] = init_right(c); q[NB_LEFT] = init_left(c);
    } else {
        const int p = sp[c];
        const int32_t* qf = &nb[8 * (size_t)f];
        for (int dir = 0; dir < 4; dir++) {
            if (kSib[dir][p] >= 0) { q[dir] = son[f] + kSib[dir][p]; continue; }
            const int fn = qf[dir];
            q[dir] = fn < 0 ? -1 : son[fn] < 0 ? fn : son[fn] + kAcross[dir][p];
        }
    }
    if (son[c] >= 0) for (int k = 0; k < 4; k++) idnei_faces(son[c] + k);
}
// the corners from the stored face neighbours of the node and of its face neighbours
void Tree::idnei_corners(int c) {
    int32_t* q = &nb[8 * (size_t)c];
    if (cross[c] >= 0) {
        auto up_of = [&](int r) { return nb[8 * (size_t)r + NB_UP]; };
        auto down_of = [&](int r) { return nb[8 * (size_t)r + NB_DOWN]; };
        auto left_of = [&](int r) { return nb[8 * (size_t)r + NB_LEFT]; };
        auto right_of = [&](int r) { return nb[8 * (size_t)r + NB_RIGHT]; };
        q[NB_UPLEFT] = diag(*this, kUL, c, q[NB_LEFT], q[NB_UP], left_of, up_of);
        q[NB_UPRIGHT] = diag(*this, kUR, c, q[NB_RIGHT], q[NB_UP], right_of, up_of);
        q[NB_DOWNLEFT] = diag(*this, kDL, c, q[NB_LEFT], q[NB_DOWN], left_of, down_of);
        q[NB_DOWNRIGHT] = diag(*this, kDR, c, q[NB_RIGHT], q[NB_DOWN], right_of, down_of);
    } else {
        q[NB_UPLEFT] = q[NB_UPRIGHT] = q[NB_DOWNLEFT] = q[NB_DOWNRIGHT] = -1;
    }
    if (son[c] >= 0) for (int k = 0; k < 4; k++) idnei_corners(son[c] + k);
}
void Tree::idneighbor() {
    // idnei_faces / idnei_corners write all eight entries of every node below a base cell; only released nodes
    // (father -2), which no walk reaches, are cleared here
    nb.resize(8 * n_alloc());
    parallel_for(n_alloc(), [&](size_t b, size_t e) {
        for (size_t c = b; c < e; c++)
            if (father[c] == -2) std::fill(nb.begin() + 8 * c, nb.begin() + 8 * c + 8, -1);
    });
    // each base cell's subtree writes its own entries: first every node's face neighbours, then
    // (when all are stored) the corners, which read those of other nodes
    parallel_for((size_t)total, [&](size_t b, size_t e) {
        for (size_t i = b; i < e; i++) idnei_faces((int)i);
    });
    parallel_for((size_t)total, [&](size_t b, size_t e) {
        for (size_t i = b; i < e; i++) idnei_corners((int)i);
    });
}

// ---------------------------------------------------------------- leaf lists
void Tree::ns_walk(int c, bool fill, int& cnt) {    // solution.F90:618-670
    if (son[c] >= 0) {
        int s = son[c];
        ns_walk(s + 0, fill, cnt); ns_walk(s + 1, fill, cnt);
        ns_walk(s + 3, fill, cnt); ns_walk(s + 2, fill, cnt);
    } else {
        if (cross[c] < 0) return;
        if (fill) ns[cnt] = c;
        cnt++;
    }
}
void Tree::bc_walk(int c, bool fill, int& cnt) {    // solution.F90:634-693
    if (son[c] >= 0) {
        for (int k = 0; k < 4; k++) bc_walk(son[c] + k, fill, cnt);
    } else {
        if (cross[c] == -5) return;
        if (cross[c] < 0) {
            if (fill) bcl[cnt] = c;
            cnt++;
        }
    }
}
void Tree::idgrid() {
    // per base cell counts, their prefix sums, then each base cell's walk writes its own range
    std::vector<int> na((size_t)total + 1, 0), nbc((size_t)total + 1, 0);
    parallel_for((size_t)total, [&](size_t b0, size_t e0) {
        for (size_t i = b0; i < e0; i++) {
            int a = 0, b = 0;
            ns_walk((int)i, false, a);
            bc_walk((int)i, false, b);
            na[i + 1] = a;
            nbc[i + 1] = b;
        }
    });
    for (int i = 0; i < total; i++) { na[i + 1] += na[i]; nbc[i + 1] += nbc[i]; }
    ns_beg.assign(na.begin(), na.end());
    ns.assign(na[total], -1);
    bcl.assign(nbc[total], -1);
    parallel_for((size_t)total, [&](size_t b0, size_t e0) {
        for (size_t i = b0; i < e0; i++) {
            int a = na[i], b = nbc[i];
            ns_walk((int)i, true, a);
            bc_walk((int)i, true, b);
        }
    });
}

// ---------------------------------------------------------------- boundaries
int Tree::bc_kind_of(int cr) const {
    return cr == -4 ? bc[0] : cr == -2 ? bc[1] : cr == -1 ? bc[2] : cr == -3 ? bc[3] : -1;
}

// Partner node of a ghost leaf g (BC_cyc, solution.F90:109-168): the periodic partner base
// cell, descended by the ghost's son path. The reference walks the path leaf-first
// (sp(1) = the leaf's own position, solution.F90:142-159); that order is kept for the
// periodic kind. The transmissive kind (new) uses the clamped interior base cell and the
// root-first path. Returns -1 for the inflow kind.
int Tree::bc_partner(int g) const {
    int a = cross[g];
    int kind = bc_kind_of(a);
    int p;
    if (kind == 6) {
        int i;
        if (a == -4) i = ni[g] * m + nj[g] - 6;
        else if (a == -2) i = ni[g] * m + nj[g] - 2 * m + 6;
        else if (a == -3) i = ni[g] * m + nj[g] - (n - 5) * m;
        else i = ni[g] * m + nj[g] - (7 - n) * m;
        p = i - 1;
    } else if (kind == 1) {
        int a1 = ni[g], b1 = nj[g];
        if (a == -4) b1 = 4;
        else if (a == -2) b1 = m - 3;
        else if (a == -3) a1 = n - 3;
        else a1 = 4;
        p = (a1 - 1) * m + b1 - 1;
    } else {
        return -1;
    }
    int cl = lvl[g];
    if (cl > 0) {
        int path[32];
        int cf = g;
        for (int j = 0; j < cl; j++) { path[j] = sp[cf]; cf = father[cf]; }
        for (int j = 0; j < cl; j++) {
            if (son[p] < 0) break;
            int s = kind == 6 ? path[j] : path[cl - 1 - j];
            p = son[p] + (s - 1);
        }
    }
    return p;
}

void Tree::boundary_fill_host() {                   // initial_boundary, solution.F90:71-107
    for (int g : bcl) {
        int kind = bc_kind_of(cross[g]);
        if (kind == 2) {
            auto& u = U[g];
            for (int i = 0; i < 7; i++) u[i] = 0.0;
            u[0] = inflow[0]; u[1] = inflow[1]; u[2] = inflow[2]; u[4] = inflow[3];
            u[3] = (u[4] * gamma * Ma * Ma) / u[0];
            continue;
        }
        int p = bc_partner(g);
        if (p < 0) continue;
        double v[7];
        restrict_mean(p, v);
        for (int i = 0; i < 7; i++) U[g][i] = v[i];
    }
}

// ---------------------------------------------------------------- canonical ids
std::vector<int32_t> Tree::canonical_ids(int32_t* n_out) const {
    std::vector<int32_t> id(n_alloc(), -1);
    int32_t next = total;
    for (int i = 0; i < total; i++) id[i] = i;
    // block pre-order: assign a node's four sons, then recurse into each
    std::vector<int32_t> stack;
    for (int i = 0; i < total; i++) {
        stack.push_back(i);
        while (!stack.empty()) {
            int c = stack.back();
            stack.pop_back();
            if (son[c] < 0) continue;
            for (int k = 0; k < 4; k++) id[son[c] + k] = next++;
            for (int k = 3; k >= 0; k--) stack.push_back(son[c] + k);
        }
    }
    if (n_out) *n_out = next;
    return id;
}

// ---------------------------------------------------------------- regrid
void Tree::rot_div_walk(int c, std::vector<int32_t>& out) const { // solution.F90:442-466
    if (son[c] < 0 && (cross[c] == 0 || cross[c] == 1)) out.push_back(c);
    else if (son[c] >= 0) for (int k = 0; k < 4; k++) rot_div_walk(son[c] + k, out);
}
void Tree::rot_div_sums(double& srot, double& sdiv, int& ne) const { // solution.F90:424-440
    // the visiting order is collected on all host threads, in runs of base cells, and the
    // sums then run through the runs in order on one thread, so the additions are the
    // reference's, in its order
    const size_t nrun = std::min<size_t>(1024, std::max(1, total));
    const size_t per = ((size_t)total + nrun - 1) / nrun;
    std::vector<std::vector<int32_t>> runs(nrun);
    parallel_for(nrun, [&](size_t r0, size_t r1) {
        for (size_t r = r0; r < r1; r++)
            for (size_t i = r * per; i < std::min((size_t)total, (r + 1) * per); i++)
                if (cross[i] == 0 || cross[i] == 1 || cross[i] == 2) rot_div_walk((int)i, runs[r]);
    }, 1);
    double a = 0, b = 0;
    size_t cnt = 0;
    for (const auto& run : runs) {
        for (int32_t k : run) { a = a + rot[k] * rot[k]; b = b + div[k] * div[k]; }
        cnt += run.size();
    }
    srot = a;
    sdiv = b;
    ne = (int)cnt;
}

void Tree::coarse_sub(int c, double arot, double adiv, int lvl_amr, double cg_rot,
                      double cg_div) {              // solution.F90:490-517
    int s = son[c];
    if (s >= 0 && lvl[c] < lvl_amr && son[s] < 0 && son[s + 1] < 0 && son[s + 2] < 0 && son[s + 3] < 0) {
        if ((rot[s] + rot[s + 1] + rot[s + 2] + rot[s + 3]) / 4 < cg_rot * arot &&
            (div[s] + div[s + 1] + div[s + 2] + div[s + 3]) / 4 < cg_div * adiv)
            coarse(c);
    } else if (s >= 0) {
        for (int k = 0; k < 4; k++)
            if (son[s + k] >= 0) coarse_sub(s + k, arot, adiv, lvl_amr, cg_rot, cg_div);
    }
}
void Tree::ref_sub(int c, double arot, double adiv, int lvl_amr, double rg_rot,
                   double rg_div) {                 // solution.F90:519-535
    if (son[c] < 0 && lvl[c] < lvl_amr) {
        if (rot[c] > rg_rot * arot || div[c] > rg_div * adiv) refine(c);
    } else if (son[c] >= 0 && lvl[c] < lvl_amr) {
        int s = son[c];
        for (int k = 0; k < 4; k++) ref_sub(s + k, arot, adiv, lvl_amr, rg_rot, rg_div);
    }
}
void Tree::solute_ref(double arot, double adiv, int lvl_amr, double rg_rot, double rg_div,
                      double cg_rot, double cg_div) { // solution.F90:468-488
    fresh.resize(n_alloc());
    parallel_for(n_alloc(), [&](size_t b, size_t e) { std::fill(fresh.begin() + b, fresh.begin() + e, 0); });
    auto cell = [&](int i) {
        if (cross[i] == 0 || cross[i] == 1) ref_sub(i, arot, adiv, lvl_amr, rg_rot, rg_div);
        if (son[i] >= 0) coarse_sub(i, arot, adiv, lvl_amr, cg_rot, cg_div);
    };
    // each base cell reads and writes only its own subtree: any order gives the same tree
    const unsigned nt = pass_threads(*this);
    if (nt > 1) {
        // ref_sub refines each leaf below lvl_amr above the goals at most once (its new sons
        // are not visited), and coarse_sub only releases: that bounds the new nodes
        std::atomic<size_t> refines{0};
        parallel_for(n_alloc(), [&](size_t b, size_t e) {
            size_t k = 0;
            for (size_t c = b; c < e; c++)
                if (father[c] != -2 && son[c] < 0 && lvl[c] < lvl_amr &&
                    (rot[c] > rg_rot * arot || div[c] > rg_div * adiv))
                    k++;
            refines += k;
        });
        PhaseTimer pt;
        run_cells(*this, nt, 4 * refines.load(), cell);
        if (pt.on) std::fprintf(stderr, "[fdmamr]   refine candidates %zu\n", refines.load());
    } else {
        for (int i = 0; i < total; i++) cell(i);
    }
}

void Tree::gridmodify1(int c) {                     // mesh.F90:305-353
    if (son[c] >= 0) for (int k = 0; k < 4; k++) gridmodify1(son[c] + k);
    if (cross[c] != 0) return;
    // the finders only read the tree, so each is evaluated only when the outcome needs it
    auto has_sons = [&](int q) { return q >= 0 && son[q] >= 0; };
    if (son[c] < 0) {
        if ((has_sons(left(c)) && has_sons(right(c))) || (has_sons(up(c)) && has_sons(down(c)))) refine(c);
    } else {
        int s = son[c];
        if (son[s] < 0 && son[s + 1] < 0 && son[s + 2] < 0 && son[s + 3] < 0)
            if ((!has_sons(left(c)) && !has_sons(right(c))) || (!has_sons(up(c)) && !has_sons(down(c))))
                coarse(c);
    }
}
void Tree::gridmodify4(int c) {                     // mesh.F90:355-395
    if (son[c] >= 0) { for (int k = 0; k < 4; k++) gridmodify4(son[c] + k); return; }
    for (int dir = 1; dir <= 4; dir++) {
        int nbr = dir == 1 ? right(c) : dir == 2 ? left(c) : dir == 3 ? up(c) : down(c);
        if (nbr < 0) continue;
        if (lvl[nbr] <= lvl[c] - 2) refine(nbr);
        int nn = dir == 1 ? right(nbr) : dir == 2 ? left(nbr) : dir == 3 ? up(nbr) : down(nbr);
        if (nn >= 0 && lvl[nn] <= lvl[c] - 2) refine(nn);
    }
}
void Tree::solute_gridmodify(int lvl_amr) {         // solution.F90:537-556
    // The reference refreshes the cached neighbour pointers after each pass (idneighbor),
    // but gridmodify1 and gridmodify4 search neighbours afresh through father/son links and
    // never read the cache, so only the final refresh (by the caller) can be observed.
    // In parallel, the base cells run row by row (run_rows). gridmodify1 on a base cell writes
    // its own subtree and reads those of the four cells beside it: cells that differ in both
    // row and column do not interact (lag 0). gridmodify4 also refines the neighbour's
    // neighbour: it reads and writes the cells up to two away in a row or a column, and cells
    // (r, c), (r', c') interact when they share a row or column within four, or
    // |r - r'|, |c - c'| <= 2 (lag 2). The node numbering then depends on the threads' timing,
    // so the tree is renumbered.
    //
    // A base cell whose leaves and those of the four cells beside it are all on one level
    // gives gridmodify1 nothing to do (each leaf's neighbours are leaves of its own level, each
    // parent of leaves has parents of leaves beside it), and a base cell whose deepest leaf is
    // at level L, with no leaf below level L - 1 in it or beside it, gives gridmodify4 nothing
    // (a leaf at level L >= 2 finds its neighbour and the neighbour's neighbour within the cells
    // beside its own); the parallel passes skip such cells, by the level bounds of each cell.
    // The bounds are exact at the start; within a pass refine and coarse widen those of the
    // cells they change, and between passes those cells' bounds are made exact again.
    //
    // From the second pass on, a phase (gridmodify1 or gridmodify4 over the cells) also skips a
    // cell when nothing it reads has changed since the same phase of the previous pass visited
    // or skipped it: no cell of its read set (itself and the cells beside it for gridmodify1,
    // the cells up to two away in its row and column for gridmodify4) changed in that phase,
    // in the phase between, or so far in this one. Its visit then sees the tree it saw last
    // time, after a visit that changed nothing (a change would have marked a cell of its read
    // set), so it would change nothing again.
    const unsigned nt = pass_threads(*this);
    PhaseTimer pt;
    auto span = [&](int i, int& lo, int& hi) {
        const int r = i / m, c = i % m;
        lo = cell_lmin[i]; hi = cell_lmax[i];
        const int nbr[4] = {c > 0 ? i - 1 : -1, c < m - 1 ? i + 1 : -1, r > 0 ? i - m : -1, r < n - 1 ? i + m : -1};
        for (int j : nbr)
            if (j >= 0) { lo = std::min<int>(lo, cell_lmin[j]); hi = std::max<int>(hi, cell_lmax[j]); }
    };
    // A pass (gridmodify1 then gridmodify4 over all cells) that changes nothing leaves the tree
    // as it found it, so every later pass would change nothing either: those are skipped.
    std::atomic<long> mods{0};
    size_t visits = 0;
    mod_count = &mods;
    struct Reset {
        Tree& t;
        ~Reset() { t.mod_count = nullptr; t.track_levels = false; }
    } reset{*this};
    if (nt > 1) {
        leaf_levels();
        track_levels = true;
        pt.mark("leaf levels");
    }
    auto report = [&](const char* phase) {
        if (pt.on) std::fprintf(stderr, "[fdmamr]   %s: %zu cells visited, %ld changes\n", phase, visits, mods.load());
        pt.mark(phase);
    };
    // true when no cell of the cross of the given radius around cell i changed in this phase or
    // the two before it (never in the first pass)
    auto settled = [&](int i, int radius) {
        if (mod_phase < 2) return false;
        const int r = i / m, c = i % m, since = mod_phase - 2;
        if (cell_mod[i] >= since) return false;
        for (int d = 1; d <= radius; d++)
            if ((c - d >= 0 && cell_mod[i - d] >= since) || (c + d < m && cell_mod[i + d] >= since) ||
                (r - d >= 0 && cell_mod[i - d * m] >= since) || (r + d < n && cell_mod[i + d * m] >= since))
                return false;
        return true;
    };
    mod_phase = 0;
    for (int k = 1; k <= lvl_amr; k++) {
        mods = 0;
        if (nt > 1) {
            visits = run_rows(*this, 0, nt, n_alloc() / 16, [&](int i) {
                if (settled(i, 1)) return false;
                int lo, hi;
                span(i, lo, hi);
                return lo != hi;
            }, [&](int i) { gridmodify1(i); });
            leaf_levels_dirty();
            report("gridmodify1");
            mod_phase++;
            visits = run_rows(*this, 2, nt, n_alloc() / 16, [&](int i) {
                if (settled(i, 2)) return false;
                int lo, hi;
                span(i, lo, hi);
                return cell_lmax[i] >= 2 && lo < cell_lmax[i] - 1;
            }, [&](int i) { gridmodify4(i); });
            leaf_levels_dirty();
            report("gridmodify4");
            mod_phase++;
        } else {
            for (int i = 0; i < total; i++) gridmodify1(i);
            pt.mark("gridmodify1");
            for (int i = 0; i < total; i++) gridmodify4(i);
            pt.mark("gridmodify4");
        }
        if (mods.load() == 0) break;
    }
    track_levels = false;
    compact();
    pt.mark("compact");
}

// ---------------------------------------------------------------- initial conditions
struct Bubble { double x, y, r, ratio; };

void Tree::initial(int ic, uint64_t seed, const double* a) {
    static const double PI_REF = 3.1415926535897932;
    double q[4][4] = {{0}}, post[4] = {0};
    std::vector<Bubble> bub;
    uint64_t s = seed;
    if (ic == 1) {
        for (int k = 0; k < 4; k++) {
            q[k][0] = uab(s, 0.5, 2.0); q[k][1] = uab(s, -0.8, 0.8);
            q[k][2] = uab(s, -0.8, 0.8); q[k][3] = uab(s, 0.1, 1.5);
        }
    } else if (ic == 2) {
        // Rankine-Hugoniot state behind a Mach a[1] shock running into rho=1, p=1, u=v=0
        double g = gamma, M = a[1];
        double p2 = 1.0 + 2.0 * g / (g + 1.0) * (M * M - 1.0);
        double r2 = (g + 1.0) * M * M / ((g - 1.0) * M * M + 2.0);
        double sh = M * std::sqrt(g);
        post[0] = r2; post[1] = sh * (1.0 - 1.0 / r2); post[2] = 0.0; post[3] = p2;
        int nbub = (int)a[2];
        double xs = a[0];
        for (int k = 0; k < nbub && k < 16; k++) {
            for (int tries = 0; tries < 100000; tries++) {
                double r = uab(s, a[3], a[4]);
                double ratio = uab(s, a[5], a[6]);
                double bx = uab(s, xs + r + 0.05, pxmax - r);
                double by = uab(s, pymin + r, pymax - r);
                bool ok = true;
                for (auto& b : bub) {
                    double dx = bx - b.x, dy = by - b.y;
                    if (std::sqrt(dx * dx + dy * dy) < r + b.r) { ok = false; break; }
                }
                if (ok) { bub.push_back({bx, by, r, ratio}); break; }
            }
        }
    }
    auto setp = [&](int c, double r, double u, double v, double p) {
        auto& w = U[c];
        for (int i = 0; i < 7; i++) w[i] = 0.0;
        w[0] = r; w[1] = u; w[2] = v; w[4] = p;
        w[3] = (w[4] * gamma * Ma * Ma) / w[0];
    };
    // every leaf of the tree (vortex is applied to all cells, solution.F90:16-19)
    for (size_t c = 0; c < n_alloc(); c++) {
        if (father[c] == -2 || son[c] >= 0) continue;
        double X = x[c], Y = y[c];
        if (ic == 0) {                              // solution.F90:31-69
            const double cc = 5.0;
            double rr = (X - 5.0) * (X - 5.0) + (Y - 5.0) * (Y - 5.0);
            double dT = -(gamma - 1.0) * (cc * cc) * std::exp(1.0 - rr) / (8.0 * gamma * PI_REF * PI_REF);
            auto& w = U[c];
            for (int i = 0; i < 7; i++) w[i] = 0.0;
            w[0] = 1.0 + std::pow(1.0 + dT, 1.0 / (gamma - 1.0)) - 1.0;
            w[1] = 1.0 - (Y - 5.0) * cc / (2.0 * PI_REF) * std::exp(0.5 * (1.0 - rr));
            w[2] = 1.0 + (X - 5.0) * cc / (2.0 * PI_REF) * std::exp(0.5 * (1.0 - rr));
            w[4] = 1.0 + std::pow(1.0 + dT, gamma / (gamma - 1.0)) - 1.0;
            w[3] = (w[4] * gamma * Ma * Ma) / w[0];
        } else if (ic == 1) {
            int k = (X >= a[0]) ? (Y >= a[1] ? 0 : 3) : (Y >= a[1] ? 1 : 2);
            setp((int)c, q[k][0], q[k][1], q[k][2], q[k][3]);
        } else if (ic == 2) {
            if (X < a[0]) setp((int)c, post[0], post[1], post[2], post[3]);
            else {
                double r = 1.0;
                for (auto& b : bub) {
                    double dx = X - b.x, dy = Y - b.y;
                    if (dx * dx + dy * dy < b.r * b.r) r = b.ratio;
                }
                setp((int)c, r, 0.0, 0.0, 1.0);
            }
        } else {
            setp((int)c, 1.0 + 0.2 * std::sin(PI_REF * (X + Y)), 1.0, 1.0, 1.0);
        }
        rot[c] = 0.0; div[c] = 0.0;
    }
}

}  // namespace fdm
