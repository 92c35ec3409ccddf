// host_tables.cpp -- regrid-time flattening of the quadtree (see host_tables.hpp).
#include "host_tables.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <atomic>
#include <unordered_map>

namespace fdm {

static const double PI_REF = 3.1415926535897932;   // module.F90:9
static const double DIR_EPS = 0x1.b7cdfep-34;      // REAL(4) 1e-10, XuanGua.F90:40

static inline double pow2i(int k) { return (double)(1 << k); }

int brinv(int n, const double* a, double* r) {     // XuanGua.F90:593-660 (column-major)
#define R(i, j) r[((j) - 1) * n + ((i) - 1)]
    int is[16], js[16];
    for (int k = 0; k < n * n; k++) r[k] = a[k];
    for (int k = 1; k <= n; k++) {
        double d = 0.0;
        for (int i = k; i <= n; i++)
            for (int j = k; j <= n; j++)
                if (std::fabs(R(i, j)) > d) { d = std::fabs(R(i, j)); is[k - 1] = i; js[k - 1] = j; }
        if (d + 1.0 == 1.0) return 0;  // singular: the reference returns the partial matrix
        for (int j = 1; j <= n; j++) { double t = R(k, j); R(k, j) = R(is[k - 1], j); R(is[k - 1], j) = t; }
        for (int i = 1; i <= n; i++) { double t = R(i, k); R(i, k) = R(i, js[k - 1]); R(i, js[k - 1]) = t; }
        R(k, k) = 1.0 / R(k, k);
        for (int j = 1; j <= n; j++) if (j != k) R(k, j) = R(k, j) * R(k, k);
        for (int i = 1; i <= n; i++)
            if (i != k)
                for (int j = 1; j <= n; j++)
                    if (j != k) R(i, j) = R(i, j) - R(i, k) * R(k, j);
        for (int i = 1; i <= n; i++) if (i != k) R(i, k) = -R(i, k) * R(k, k);
    }
    for (int k = n; k >= 1; k--) {
        for (int j = 1; j <= n; j++) { double t = R(k, j); R(k, j) = R(js[k - 1], j); R(js[k - 1], j) = t; }
        for (int i = 1; i <= n; i++) { double t = R(i, k); R(i, k) = R(i, is[k - 1]); R(i, is[k - 1]) = t; }
    }
#undef R
    return 1;
}

int get_direction(double xx, double yy, double x, double y) {   // XuanGua.F90:33-43
    double dx = xx - x, dy = yy - y;
    // The octants are bounded by |dy| / |dx| = tan(22.5 deg) and its inverse. Away from those
    // lines (relative margin 1e-7, far above the arctangent's rounding and the 1e-10 shift) the
    // octant follows from comparisons; on a stencil the offsets are small multiples of the cell
    // sizes, whose ratios never come near the irrational bound, so the arctangent below is the
    // rare fallback.
    const double ax = std::fabs(dx), ay = std::fabs(dy), t = 0.41421356237309503;  // tan(22.5 deg)
    const double tol = 1e-7 * (ax + ay);
    if (ax + ay > 0 && std::fabs(ay - t * ax) > tol && std::fabs(ax - t * ay) > tol) {
        if (ay < t * ax) return dx > 0 ? 1 : 5;
        if (ax < t * ay) return dy > 0 ? 3 : 7;
        return dx > 0 ? (dy > 0 ? 2 : 8) : (dy > 0 ? 4 : 6);
    }
    double angle = (std::atan2(dy, dx) + DIR_EPS) * 180.0 / PI_REF;
    int fl = (int)std::floor((angle + 22.5) / 45.0 + 8);
    return fl % 8 + 1;
}

namespace {

struct List { int a[16]; int n = 0; };

void add_to(List& l, int c) {                       // XuanGua.F90:11-30
    for (int i = 0; i < l.n; i++) if (l.a[i] == c) return;
    l.a[l.n++] = c;
}

// XuanGua.F90:152-177 (one pass; the reference's 4x repeat adds nothing after the first)
void search5(const Tree& t, int c, double x, double y, List& l1, List& l2, bool* used) {
    if (t.son[c] < 0) {
        int dir = get_direction(t.x[c], t.y[c], x, y);
        if (!used[dir - 1]) {
            if (dir == 1 || dir == 5 || dir == 3 || dir == 7) add_to(l1, c);
            else add_to(l2, c);
            used[dir - 1] = true;
        }
    } else {
        for (int k = 0; k < 4; k++) search5(t, t.son[c] + k, x, y, l1, l2, used);
    }
}

void parent_lists(const Tree& t, int c, List& l1, List& l2) {  // XuanGua.F90:45-86
    bool used[8] = {false};
    l1.n = l2.n = 0;
    for (int k = 0; k < 4; k++) search5(t, t.son[c] + k, t.x[c], t.y[c], l1, l2, used);
    if (l1.n + l2.n < 8) {
        const int32_t* q = &t.nb[8 * (size_t)c];
        const int order[4] = {NB_LEFT, NB_RIGHT, NB_DOWN, NB_UP};
        for (int k = 0; k < 4; k++)
            if (q[order[k]] >= 0) search5(t, q[order[k]], t.x[c], t.y[c], l1, l2, used);
    }
}

void child_lists(const Tree& t, int c, List& l1, List& l2) {   // XuanGua.F90:96-141
    bool used[8] = {false};
    l1.n = l2.n = 0;
    add_to(l1, c);
    const int32_t* q = &t.nb[8 * (size_t)c];
    const int order[8] = {NB_DOWNLEFT, NB_LEFT, NB_UPLEFT, NB_DOWN, NB_UP, NB_DOWNRIGHT, NB_RIGHT, NB_UPRIGHT};
    for (int k = 0; k < 8; k++)
        if (q[order[k]] >= 0) search5(t, q[order[k]], t.x[c], t.y[c], l1, l2, used);
}

// Inputs of a slot's WLS-WENO geometry: the stencil coordinates, the evaluation point and the
// cell's box, in the frame the record uses (the reference's absolute frame, or the cell-local
// scaled one: x' = (x - xc)/hx, box [-1/2, 1/2]^2). Unused entries are zero, so two slots have
// the same geometry record exactly when their inputs are equal bit for bit.
struct WlsIn {
    int32_t m, m1;
    double xx, yy, px[9], py[9], box[4];  // box: x_min, x_max, y_min, y_max
};

void wls_inputs(const Tree& t, int cell, double xx, double yy, const int32_t* nodes, int m, int m1, WlsIn& in) {
    std::memset(&in, 0, sizeof(in));
    in.m = m;
    in.m1 = m1;
    const double hxc = t.xf * t.h / pow2i(t.lvl[cell]), hyc = t.h / pow2i(t.lvl[cell]);
    if (t.wls_local) {
        xx = (xx - t.x[cell]) / hxc;
        yy = (yy - t.y[cell]) / hyc;
    }
    in.xx = xx;
    in.yy = yy;
    for (int i = 0; i < m; i++) {
        const int q = nodes[i];
        double px = t.x[q], py = t.y[q];
        if (t.wls_local) { px = (px - t.x[cell]) / hxc; py = (py - t.y[cell]) / hyc; }
        in.px[i] = px;
        in.py[i] = py;
    }
    if (t.wls_local) {
        in.box[0] = -0.5; in.box[1] = 0.5; in.box[2] = -0.5; in.box[3] = 0.5;
    } else {
        const double hx = t.xf * t.h / std::ldexp(1.0, t.lvl[cell] + 1);
        const double hy = t.h / std::ldexp(1.0, t.lvl[cell] + 1);
        in.box[0] = t.x[cell] - hx; in.box[1] = t.x[cell] + hx;
        in.box[2] = t.y[cell] - hy; in.box[3] = t.y[cell] + hy;
    }
}

// WLS-WENO geometry record of one slot (XuanGua.F90:206-255 and the box of :318-323)
void wls_record(const WlsIn& in, double* d) {
    const int m1 = in.m1, m = in.m;
    const double xx = in.xx, yy = in.yy;
    double w[16], P[16][6];
    for (int i = 0; i < m; i++) {
        const double px = in.px[i], py = in.py[i];
        double dd = std::sqrt((px - xx) * (px - xx) + (py - yy) * (py - yy));
        w[i] = 1.0 / (dd + 1e-6);
        d[WLS_PX + i] = px; d[WLS_PY + i] = py; d[WLS_W + i] = w[i];
        P[i][0] = 1.0; P[i][1] = px; P[i][2] = py;
        P[i][3] = px * py; P[i][4] = px * px; P[i][5] = py * py;
    }
    double A6[36], A3a[9], A3b[9];
    for (int r = 0; r < 6; r++)
        for (int c = 0; c < 6; c++) {
            double s = 0.0;
            for (int i = 0; i < m; i++) s = s + P[i][r] * (w[i] * P[i][c]);
            A6[c * 6 + r] = s;
        }
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) {
            double s = 0.0;
            for (int i = 0; i < m1; i++) s = s + P[i][r] * (w[i] * P[i][c]);
            A3a[c * 3 + r] = s;
            s = 0.0;
            for (int i = m1; i < m; i++) s = s + P[i][r] * (w[i] * P[i][c]);
            A3b[c * 3 + r] = s;
        }
    brinv(6, A6, d + WLS_INV6);
    brinv(3, A3a, d + WLS_INV3A);
    brinv(3, A3b, d + WLS_INV3B);
    d[WLS_XX] = xx; d[WLS_YY] = yy;
    const double x_min = in.box[0], x_max = in.box[1], y_min = in.box[2], y_max = in.box[3];
    d[WLS_BOX + 0] = x_max - x_min;
    d[WLS_BOX + 1] = y_max - y_min;
    d[WLS_BOX + 2] = x_max * x_max - x_min * x_min;
    d[WLS_BOX + 3] = y_max * y_max - y_min * y_min;
    d[WLS_BOX + 4] = x_max * x_max * x_max - x_min * x_min * x_min;
    d[WLS_BOX + 5] = y_max * y_max * y_max - y_min * y_min * y_min;
}

// Fold a slot's normal-equation solves into FP32 coefficient matrices (all products in
// FP64, rounded once): arr = C phi with C = (P^T W P)^-1 P^T W, and the fitted value
// basis(xx,yy)^T arr = c^T phi.
void fast_record(const double* d, const int32_t* iv, float* f) {
    const int m = iv[WLSI_M], m1 = iv[WLSI_M1], m2 = m - m1;
    const double xx = d[WLS_XX], yy = d[WLS_YY];
    double P[9][6];
    for (int i = 0; i < m; i++) {
        const double px = d[WLS_PX + i], py = d[WLS_PY + i];
        P[i][0] = 1.0; P[i][1] = px; P[i][2] = py; P[i][3] = px * py; P[i][4] = px * px; P[i][5] = py * py;
    }
    const double bl[6] = {1.0, xx, yy, xx * yy, xx * xx, yy * yy};
    for (int i = 0; i < m; i++) {
        double cl = 0.0;
        for (int r = 0; r < 6; r++) {
            double s = 0.0;
            for (int c = 0; c < 6; c++) s += d[WLS_INV6 + c * 6 + r] * P[i][c];
            s *= d[WLS_W + i];
            if (r > 0) f[WF_C6 + (r - 1) * 9 + i] = (float)s;
            cl += bl[r] * s;
        }
        f[WF_CL + i] = (float)cl;
    }
    for (int j = 0; j < m1 && j < 5; j++) {
        double cs = 0.0;
        for (int r = 0; r < 3; r++) {
            double s = 0.0;
            for (int c = 0; c < 3; c++) s += d[WLS_INV3A + c * 3 + r] * P[j][c];
            s *= d[WLS_W + j];
            if (r > 0) f[WF_C3A + (r - 1) * 5 + j] = (float)s;
            cs += bl[r] * s;
        }
        f[WF_CS1 + j] = (float)cs;
    }
    for (int j = 0; j < m2 && j < 4; j++) {
        double cs = 0.0;
        for (int r = 0; r < 3; r++) {
            double s = 0.0;
            for (int c = 0; c < 3; c++) s += d[WLS_INV3B + c * 3 + r] * P[m1 + j][c];
            s *= d[WLS_W + m1 + j];
            if (r > 0) f[WF_C3B + (r - 1) * 9 + m1 + j] = (float)s;
            cs += bl[r] * s;
        }
        f[WF_CS2 + m1 + j] = (float)cs;
    }
    for (int k = 0; k < 6; k++) f[WF_BOX + k] = (float)d[WLS_BOX + k];
}

struct Key {
    int kind, node;
    uint64_t xb, yb;
    bool operator==(const Key& o) const { return kind == o.kind && node == o.node && xb == o.xb && yb == o.yb; }
};
struct KeyHash {
    size_t operator()(const Key& k) const {
        uint64_t h = (uint64_t)k.node * 0x9E3779B97F4A7C15ULL ^ (uint64_t)k.kind;
        h ^= k.xb + 0x9E3779B97F4A7C15ULL + (h << 6) + (h >> 2);
        h ^= k.yb + 0x9E3779B97F4A7C15ULL + (h << 6) + (h >> 2);
        return (size_t)h;
    }
};
uint64_t bits(double v) { uint64_t b; std::memcpy(&b, &v, 8); return b; }

}  // namespace

void stencil_point(const Tree& t, int c, int k, int kind, int tg, double* px, double* py) {
    if (kind == ST_PARENT_WLS || kind == ST_PARENT_RESTRICT) {
        *px = t.x[tg];
        *py = t.y[tg];
        return;
    }
    const int dir = k & 3;
    const bool two = k >= 4;
    const double off = two ? (dir < 2 ? 2.0 * t.xf * t.h : 2.0 * t.h) / pow2i(t.lvl[c])
                           : (dir < 2 ? t.xf * t.h : t.h) / pow2i(t.lvl[c]);
    double xx = t.x[c], yy = t.y[c];
    if (dir == 0) xx = t.x[c] + off;
    else if (dir == 1) xx = t.x[c] - off;
    else if (dir == 2) yy = t.y[c] + off;
    else yy = t.y[c] - off;
    *px = xx;
    *py = yy;
}

int stencil_lists(const Tree& t, int tg, int kind, int32_t* nb9, int32_t* m1, int32_t* m2) {
    List l1, l2;
    if (kind == ST_PARENT_WLS || kind == ST_PARENT_RESTRICT) parent_lists(t, tg, l1, l2);
    else if (kind == ST_CHILD_WLS || kind == ST_CHILD_COPY) child_lists(t, tg, l1, l2);
    *m1 = l1.n;
    *m2 = l2.n;
    for (int j = 0; j < 9; j++) nb9[j] = -1;
    for (int j = 0; j < l1.n && j < 9; j++) nb9[j] = l1.a[j];
    for (int j = 0; j < l2.n && l1.n + j < 9; j++) nb9[l1.n + j] = l2.a[j];
    return 0;
}

// v resized to n elements; beyond its capacity the old contents are dropped, not copied (the
// caller writes every element)
template <class V>
static void resize_fresh(V& v, size_t n) {
    if (n > v.capacity()) {
        v.clear();
        v.shrink_to_fit();
        v.reserve(n + n / 8);
    }
    v.resize(n);
}
// v resized to n with every element set to val, on all host threads
template <class V, class T>
static void fill_par(V& v, size_t n, T val) {
    resize_fresh(v, n);
    parallel_for(n, [&](size_t b, size_t e) { std::fill(v.begin() + b, v.begin() + e, val); });
}

// out = the i in [0, n) for which pred(i) holds, ascending, appended; on all host threads (a
// prefix count over chunks, then each chunk writes its own range)
template <class V, class P>
static void append_where(size_t n, const P& pred, V& out) {
    const size_t nch = 256, per = (n + nch - 1) / nch;
    std::vector<size_t> at(nch + 1, 0);
    parallel_for(nch, [&](size_t b, size_t e) {
        for (size_t ch = b; ch < e; ch++)
            for (size_t i = ch * per; i < std::min(n, (ch + 1) * per); i++) at[ch + 1] += pred(i) ? 1 : 0;
    }, 1);
    for (size_t ch = 0; ch < nch; ch++) at[ch + 1] += at[ch];
    const size_t o = out.size();
    out.resize(o + at[nch]);
    parallel_for(nch, [&](size_t b, size_t e) {
        for (size_t ch = b; ch < e; ch++) {
            size_t k = o + at[ch];
            for (size_t i = ch * per; i < std::min(n, (ch + 1) * per); i++)
                if (pred(i)) out[k++] = (int32_t)i;
        }
    }, 1);
}

void build_tables(const Tree& t, Tables& tb) {
    PhaseTimer pt;
    const size_t NA = t.n_alloc();
    {
        // The largest arrays keep their storage from the previous build (resized, not cleared:
        // every element is written below), which spares the time of mapping fresh memory.
        Tables old = std::move(tb);
        tb = Tables();
        tb.stencil.swap(old.stencil); tb.st_kind.swap(old.st_kind); tb.st_node.swap(old.st_node);
        tb.off16.swap(old.off16);
        tb.sel_loc.swap(old.sel_loc); tb.state_of_node.swap(old.state_of_node);
        tb.leaf_node.swap(old.leaf_node); tb.node_son.swap(old.node_son); tb.lvl_leaf.swap(old.lvl_leaf);
        tb.face_share.swap(old.face_share); tb.ns_pos.swap(old.ns_pos); tb.dev_of_ns.swap(old.dev_of_ns);
        tb.wls_i.swap(old.wls_i); tb.sel_nbr.swap(old.sel_nbr); tb.face_lo.swap(old.face_lo);
    }
    tb.N = (int)t.ns.size();
    tb.G = (int)t.bcl.size();
    fill_par(tb.state_of_node, NA, -1);
    // Device order of the interior leaves: the base cells along a Morton (Z-order) curve over
    // the base grid, the leaves of each base cell in the reference's cell_NS order (son1, son2,
    // son4, son3: Morton again), so leaves that are neighbours in either direction are close in
    // memory. ns_pos / dev_of_ns translate to and from cell_NS order at the C-ABI.
    {
        auto spread = [](uint32_t v) {
            uint64_t x = v;
            x = (x | (x << 16)) & 0x0000FFFF0000FFFFull;
            x = (x | (x << 8)) & 0x00FF00FF00FF00FFull;
            x = (x | (x << 4)) & 0x0F0F0F0F0F0F0F0Full;
            x = (x | (x << 2)) & 0x3333333333333333ull;
            x = (x | (x << 1)) & 0x5555555555555555ull;
            return x;
        };
        // the base cells that hold interior leaves (cell_NS walks the base cells row-major,
        // ns_beg marks where each starts), by Morton key
        if (t.ns_beg.size() != (size_t)t.total + 1 || t.ns_beg[t.total] != tb.N)
            throw std::runtime_error("leaf lists out of date");
        std::vector<std::pair<uint64_t, int>> runs;  // (Morton key, base cell)
        for (int b = 0; b < t.total; b++)
            if (t.ns_beg[b + 1] > t.ns_beg[b])
                runs.emplace_back(spread((uint32_t)(b % t.m)) | (spread((uint32_t)(b / t.m)) << 1), b);
        std::sort(runs.begin(), runs.end());
        std::vector<int> at(runs.size() + 1, 0);
        for (size_t r = 0; r < runs.size(); r++)
            at[r + 1] = at[r] + t.ns_beg[runs[r].second + 1] - t.ns_beg[runs[r].second];
        resize_fresh(tb.ns_pos, tb.N);
        resize_fresh(tb.dev_of_ns, tb.N);
        parallel_for(runs.size(), [&](size_t r0, size_t r1) {
            for (size_t r = r0; r < r1; r++) {
                int i = at[r];
                for (int e = t.ns_beg[runs[r].second]; e < t.ns_beg[runs[r].second + 1]; e++, i++) {
                    tb.ns_pos[i] = e;
                    tb.dev_of_ns[e] = i;
                }
            }
        });
    }
    pt.mark("tables: device order");
    resize_fresh(tb.leaf_node, (size_t)tb.N + tb.G);
    parallel_for((size_t)tb.N, [&](size_t b, size_t e) {
        for (size_t i = b; i < e; i++) {
            const int c = t.ns[tb.ns_pos[i]];
            tb.state_of_node[c] = (int32_t)i;
            tb.leaf_node[i] = c;
        }
    });
    for (int i = 0; i < tb.G; i++) { tb.state_of_node[t.bcl[i]] = tb.N + i; tb.leaf_node[(size_t)tb.N + i] = t.bcl[i]; }
    // passive leaves: corners (never updated, solution.F90:644,682-686), in node order
    {
        std::vector<std::vector<int32_t>> part(64);
        const size_t per = (NA + 63) / 64;
        parallel_for(64, [&](size_t b, size_t e) {
            for (size_t k = b; k < e; k++)
                for (size_t c = k * per; c < std::min(NA, (k + 1) * per); c++)
                    if (t.father[c] != -2 && t.son[c] < 0 && tb.state_of_node[c] < 0) part[k].push_back((int32_t)c);
        }, 1);
        for (auto& v : part)
            for (int32_t c : v) {
                tb.state_of_node[c] = (int)tb.leaf_node.size();
                tb.leaf_node.push_back(c);
            }
    }
    tb.L = (int)tb.leaf_node.size();
    tb.P = tb.L - tb.N - tb.G;
    resize_fresh(tb.node_son, NA);
    parallel_for(NA, [&](size_t b, size_t e) { std::copy(t.son.begin() + b, t.son.begin() + e, tb.node_son.begin() + b); });

    pt.mark("tables: leaf numbering");
    const int N = tb.N;
    // every entry of these is written by the classification below
    resize_fresh(tb.stencil, 8 * (size_t)N);
    resize_fresh(tb.lvl_leaf, N);
    resize_fresh(tb.st_kind, 8 * (size_t)N);
    resize_fresh(tb.st_node, 8 * (size_t)N);

    pt.mark("tables: allocation");
    // phase 1 (parallel): locate every stencil entry as the reference's gathers do
    // (Tip/Tim/Tjp/Tjm for k < 4, Tipp/Timm/Tjpp/Tjmm for k >= 4, XuanGua.F90:401-590). A
    // same-level leaf is a copy; a same-level parent and a coarser leaf are level jumps, whose
    // kind (WLS-WENO, or the fallbacks when the stencil has fewer than 7 points) depends only
    // on the jump's key: the parent node, or the coarse leaf and the point.
    constexpr int8_t CAND_PARENT = -1, CAND_CHILD = -2;
    std::atomic<int> lmax_at{0}, missing{0};
    parallel_for((size_t)N, [&](size_t b, size_t e) {
        int lm = 0;
        for (size_t ii = b; ii < e; ii++) {
            const int i = (int)ii;
            const int c = tb.leaf_node[i];
            tb.lvl_leaf[i] = (int8_t)t.lvl[c];
            lm = std::max(lm, t.lvl[c]);
            const int32_t* q = &t.nb[8 * (size_t)c];
            for (int k = 0; k < 8; k++) {
                const int dir = k & 3;
                const bool two = k >= 4;
                const int nbk = dir == 0 ? NB_RIGHT : dir == 1 ? NB_LEFT : dir == 2 ? NB_UP : NB_DOWN;
                const int cn = q[nbk];
                if (cn < 0) { missing = 1; continue; }
                int tg;
                bool same;
                if (!two) { tg = cn; same = t.lvl[cn] == t.lvl[c]; }
                else if (t.lvl[cn] == t.lvl[c]) {
                    int cnn = t.nb[8 * (size_t)cn + nbk];
                    if (cnn < 0) { missing = 1; continue; }
                    tg = cnn; same = t.lvl[cnn] == t.lvl[cn];
                } else { tg = cn; same = false; }
                int8_t kind;
                const size_t d = (size_t)i * 8 + k;
                if (same && t.son[tg] < 0) {
                    kind = ST_COPY;
                    tb.stencil[d] = tb.state_of_node[tg];
                } else {
                    kind = same ? CAND_PARENT : CAND_CHILD;
                }
                tb.st_kind[d] = kind;
                tb.st_node[d] = tg;
            }
        }
        int cur = lmax_at.load();
        while (lm > cur && !lmax_at.compare_exchange_weak(cur, lm)) {}
    });
    if (missing) throw std::runtime_error("interior leaf without its stencil neighbours");
    pt.mark("tables: classify");
    const int lmax = lmax_at.load();

    // phase 2: the distinct level-jump keys, numbered in order of first occurrence. The jump
    // entries are sharded by key hash over the threads; each shard finds the first entry of
    // each of its keys, and a prefix count over the first entries numbers the keys.
    BigVec<int64_t> jd;  // entries that are level jumps, ascending
    {
        const unsigned nch = 64;
        const size_t per = ((size_t)N + nch - 1) / nch;
        std::vector<std::vector<int64_t>> part(nch);
        parallel_for(nch, [&](size_t b, size_t e) {
            for (size_t ch = b; ch < e; ch++)
                for (size_t d = ch * per * 8; d < std::min((size_t)N, (ch + 1) * per) * 8; d++)
                    if (tb.st_kind[d] < 0) part[ch].push_back((int64_t)d);
        }, 1);
        std::vector<size_t> at(nch + 1, 0);
        for (unsigned ch = 0; ch < nch; ch++) at[ch + 1] = at[ch] + part[ch].size();
        jd.resize(at[nch]);  // every element is copied below
        parallel_for(nch, [&](size_t b, size_t e) {
            for (size_t ch = b; ch < e; ch++) std::copy(part[ch].begin(), part[ch].end(), jd.begin() + at[ch]);
        }, 1);
    }
    pt.mark("tables: jump entries");
    const size_t J = jd.size();
    auto key_of = [&](size_t j) {
        const size_t d = (size_t)jd[j];
        const bool parent = tb.st_kind[d] == CAND_PARENT;
        if (parent) return Key{1, tb.st_node[d], 0, 0};
        double px, py;
        stencil_point(t, tb.leaf_node[d / 8], (int)(d % 8), ST_CHILD_WLS, tb.st_node[d], &px, &py);
        return Key{3, tb.st_node[d], bits(px), bits(py)};
    };
    pt.mark("tables: jump keys");
    // (the large temporaries below are not zeroed: every element is written before it is read,
    // on all host threads)
    BigVec<int32_t> rep(J);  // first jump entry with the same key
    {
        auto mix = [](uint64_t z) {
            z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
            z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
            return z ^ (z >> 31);
        };
        BigVec<uint64_t> hj(J);
        BigVec<Key> kj(J);  // each entry's key, computed once
        parallel_for(J, [&](size_t b, size_t e) {
            for (size_t j = b; j < e; j++) {
                const Key k = key_of(j);
                kj[j] = k;
                hj[j] = mix(mix(mix((uint64_t)k.kind << 32 | (uint32_t)k.node) ^ k.xb) ^ k.yb);
            }
        });
        // 256 shards by the top hash bits, each an open-addressing table of first entries small
        // enough to stay in a core's cache. The entries are first sorted by shard (stably:
        // chunks of entries in order, each chunk's entries of a shard at that chunk's offset),
        // so each shard reads only its own.
        constexpr unsigned ns = 256, shard_shift = 56;
        const size_t nch = 64, per = (J + nch - 1) / nch;
        std::vector<size_t> off(ns * nch + 1, 0);
        parallel_for(nch, [&](size_t b, size_t e) {
            for (size_t ch = b; ch < e; ch++)
                for (size_t j = ch * per; j < std::min(J, (ch + 1) * per); j++) off[1 + (hj[j] >> shard_shift) * nch + ch]++;
        }, 1);
        for (size_t k = 0; k < ns * nch; k++) off[k + 1] += off[k];
        // entry numbers, hashes and keys in shard order
        BigVec<int32_t> by_shard(J);
        BigVec<uint64_t> hs(J);
        BigVec<Key> ks(J);
        parallel_for(nch, [&](size_t b, size_t e) {
            for (size_t ch = b; ch < e; ch++) {
                std::vector<size_t> at(ns);
                for (unsigned sh = 0; sh < ns; sh++) at[sh] = off[sh * nch + ch];
                for (size_t j = ch * per; j < std::min(J, (ch + 1) * per); j++) {
                    const size_t k = at[hj[j] >> shard_shift]++;
                    by_shard[k] = (int32_t)j;
                    hs[k] = hj[j];
                    ks[k] = kj[j];
                }
            }
        }, 1);
        parallel_for(ns, [&](size_t b, size_t e) {
            std::vector<int32_t> first;  // shard positions of the first entries
            for (size_t sh = b; sh < e; sh++) {
                const size_t k0 = off[sh * nch], k1 = off[(sh + 1) * nch];
                const size_t cnt = k1 - k0;
                size_t cap = 16;
                while (cap < 2 * cnt) cap <<= 1;
                const size_t mask = cap - 1;
                first.assign(cap, -1);
                for (size_t k = k0; k < k1; k++) {
                    size_t at = hs[k] & mask;
                    int32_t f;
                    while ((f = first[at]) >= 0) {
                        if (hs[(size_t)f] == hs[k] && ks[(size_t)f] == ks[k]) break;
                        at = (at + 1) & mask;
                    }
                    if (f < 0) first[at] = f = (int32_t)k;
                    rep[by_shard[k]] = by_shard[(size_t)f];
                }
            }
        }, 1);
    }
    pt.mark("tables: dedup");
    // per distinct key (in order): its first entry, then the kind its stencil lists give. The
    // keys are numbered by a prefix count of the first entries over chunks of entries; every
    // other entry then takes its first entry's number.
    BigVec<int32_t> uid(J), ufirst;
    {
        const size_t nch = 256, per = (J + nch - 1) / nch;
        std::vector<size_t> at(nch + 1, 0);
        parallel_for(nch, [&](size_t b, size_t e) {
            for (size_t ch = b; ch < e; ch++)
                for (size_t j = ch * per; j < std::min(J, (ch + 1) * per); j++) at[ch + 1] += rep[j] == (int32_t)j;
        }, 1);
        for (size_t ch = 0; ch < nch; ch++) at[ch + 1] += at[ch];
        ufirst.resize(at[nch]);
        parallel_for(nch, [&](size_t b, size_t e) {
            for (size_t ch = b; ch < e; ch++) {
                size_t u = at[ch];
                for (size_t j = ch * per; j < std::min(J, (ch + 1) * per); j++)
                    if (rep[j] == (int32_t)j) { uid[j] = (int32_t)u; ufirst[u++] = (int32_t)j; }
            }
        }, 1);
        parallel_for(J, [&](size_t b, size_t e) {
            for (size_t j = b; j < e; j++)
                if (rep[j] != (int32_t)j) uid[j] = uid[rep[j]];
        });
    }
    pt.mark("tables: slot numbering");
    const size_t U = ufirst.size();
    BigVec<int8_t> ukind(U), um(U), um1(U);
    BigVec<int32_t> unodes(U * 9);
    parallel_for(U, [&](size_t b, size_t e) {
        for (size_t u = b; u < e; u++) {
            std::fill(&unodes[u * 9], &unodes[u * 9] + 9, -1);
            const size_t d = (size_t)jd[ufirst[u]];
            const int tg = tb.st_node[d];
            const bool parent = tb.st_kind[d] == CAND_PARENT;
            List l1, l2;
            if (parent) parent_lists(t, tg, l1, l2);
            else child_lists(t, tg, l1, l2);
            const int m = l1.n + l2.n;
            ukind[u] = parent ? (m < 7 ? ST_PARENT_RESTRICT : ST_PARENT_WLS) : (m < 7 ? ST_CHILD_COPY : ST_CHILD_WLS);
            um[u] = (int8_t)m;
            um1[u] = (int8_t)l1.n;
            if (m >= 7) {
                if (m > 9) throw std::runtime_error("WLS-WENO stencil with more than 9 points");
                for (int i = 0; i < l1.n; i++) unodes[u * 9 + i] = l1.a[i];
                for (int i = 0; i < l2.n; i++) unodes[u * 9 + l1.n + i] = l2.a[i];
            }
        }
    });
    pt.mark("tables: lists");
    // WLS slots and restriction slots, each in order of first occurrence
    // (numbered by prefix counts over chunks of keys)
    BigVec<int32_t> uref(U);
    std::vector<int32_t> wls_u;  // per WLS slot: its key
    {
        const size_t nch = 256, per = (U + nch - 1) / nch;
        std::vector<size_t> nw(nch + 1, 0), nr(nch + 1, 0);
        parallel_for(nch, [&](size_t b, size_t e) {
            for (size_t ch = b; ch < e; ch++)
                for (size_t u = ch * per; u < std::min(U, (ch + 1) * per); u++) {
                    nw[ch + 1] += ukind[u] == ST_PARENT_WLS || ukind[u] == ST_CHILD_WLS;
                    nr[ch + 1] += ukind[u] == ST_PARENT_RESTRICT;
                }
        }, 1);
        for (size_t ch = 0; ch < nch; ch++) { nw[ch + 1] += nw[ch]; nr[ch + 1] += nr[ch]; }
        wls_u.resize(nw[nch]);
        tb.rst_node.resize(nr[nch]);
        parallel_for(nch, [&](size_t b, size_t e) {
            for (size_t ch = b; ch < e; ch++) {
                size_t w = nw[ch], r = nr[ch];
                for (size_t u = ch * per; u < std::min(U, (ch + 1) * per); u++) {
                    const int tg = tb.st_node[(size_t)jd[ufirst[u]]];
                    switch (ukind[u]) {
                        case ST_PARENT_WLS:
                        case ST_CHILD_WLS: uref[u] = 1000000000 + (int32_t)w; wls_u[w++] = (int32_t)u; break;
                        case ST_PARENT_RESTRICT: uref[u] = -2 - (int32_t)r; tb.rst_node[r++] = tg; break;
                        default: uref[u] = tb.state_of_node[tg];
                    }
                }
            }
        }, 1);
    }
    tb.S_w = (int)wls_u.size();
    tb.S_r = (int)tb.rst_node.size();
    parallel_for(J, [&](size_t b, size_t e) {
        for (size_t j = b; j < e; j++) {
            const size_t d = (size_t)jd[j];
            const int u = uid[j];
            int32_t r = uref[u];
            if (r >= 1000000000) r = tb.L + (r - 1000000000);
            else if (r <= -2) r = tb.L + tb.S_w + (-2 - r);
            tb.stencil[d] = r;
            tb.st_kind[d] = ukind[u];
        }
    });
    pt.mark("tables: slot lists");

    // Records by pattern: in the cell-local frame a slot's geometry (coordinates in units of
    // the cell size, weights, inverted normal matrices, smoothness box) depends only on the
    // shape of the stencil around the level jump, so slots with the same shape have
    // bit-identical records and share one, numbered in slot order; wls_i[.. + WLSI_PAT] names
    // it. The record is a function of the slot's inputs (WlsIn), so equal inputs mean equal
    // records; both the FP64 records and the FP32 ones derived from them are stored once per
    // pattern. In the reference's absolute frame every record differs and each slot keeps its
    // own. Each thread finds the distinct inputs of a run of slots; the runs are merged in slot
    // order and the records computed once per pattern.
    fill_par(tb.wls_i, (size_t)tb.S_w * WLSI_STRIDE, -1);
    {
        struct InHash {
            size_t operator()(const WlsIn& in) const {
                uint64_t h = 1469598103934665603ull;
                const uint64_t* w = reinterpret_cast<const uint64_t*>(&in);
                for (size_t k = 0; k < sizeof(WlsIn) / 8; k++) { h ^= w[k]; h *= 1099511628211ull; }
                return (size_t)(h ^ (h >> 29));
            }
        };
        struct InEq {
            bool operator()(const WlsIn& a, const WlsIn& b) const { return std::memcmp(&a, &b, sizeof(WlsIn)) == 0; }
        };
        static_assert(sizeof(WlsIn) % 8 == 0, "WlsIn hashed as 64-bit words");
        auto inputs = [&](int s, WlsIn& in) {
            const size_t u = (size_t)wls_u[s];
            const size_t d = (size_t)jd[ufirst[u]];
            double px, py;
            stencil_point(t, tb.leaf_node[d / 8], (int)(d % 8), ukind[u], tb.st_node[d], &px, &py);
            wls_inputs(t, tb.st_node[d], px, py, &unodes[u * 9], um[u], um1[u], in);
        };
        const size_t nrun = std::max<size_t>(1, std::min<size_t>(64, (size_t)tb.S_w / 4096));
        const size_t per = ((size_t)tb.S_w + nrun - 1) / nrun;
        std::vector<std::vector<int32_t>> run_first(nrun);  // first slot of each local pattern
        BigVec<int32_t> lpat(tb.S_w);
        parallel_for(nrun, [&](size_t b, size_t e) {
            for (size_t r = b; r < e; r++) {
                std::unordered_map<WlsIn, int32_t, InHash, InEq> loc;
                WlsIn in;
                for (size_t s = r * per; s < std::min((size_t)tb.S_w, (r + 1) * per); s++) {
                    inputs((int)s, in);
                    auto it = loc.emplace(in, (int32_t)run_first[r].size());
                    if (it.second) run_first[r].push_back((int32_t)s);
                    lpat[s] = it.first->second;
                }
            }
        }, 1);
        std::unordered_map<WlsIn, int32_t, InHash, InEq> glob;
        std::vector<int32_t> pat_slot;  // first slot of each pattern
        std::vector<std::vector<int32_t>> run_map(nrun);
        for (size_t r = 0; r < nrun; r++) {
            WlsIn in;
            for (int32_t s0 : run_first[r]) {
                inputs(s0, in);
                auto it = glob.emplace(in, (int32_t)pat_slot.size());
                if (it.second) pat_slot.push_back(s0);
                run_map[r].push_back(it.first->second);
            }
        }
        tb.n_pat = (int)pat_slot.size();
        tb.wls_d.assign((size_t)tb.n_pat * WLS_STRIDE, 0.0);
        tb.wls_f.assign((size_t)tb.n_pat * WF_STRIDE, 0.0f);
        parallel_for((size_t)tb.S_w, [&](size_t b, size_t e) {
            for (size_t s = b; s < e; s++) {
                int32_t* iv = &tb.wls_i[s * WLSI_STRIDE];
                const size_t u = (size_t)wls_u[s];
                iv[WLSI_M] = um[u];
                iv[WLSI_M1] = um1[u];
                for (int i = 0; i < um[u]; i++) iv[WLSI_IDS + i] = tb.state_of_node[unodes[u * 9 + i]];
                iv[WLSI_PAT] = run_map[s / per][lpat[s]];
            }
        });
        parallel_for((size_t)tb.n_pat, [&](size_t b, size_t e) {
            WlsIn in;
            for (size_t p = b; p < e; p++) {
                const int s0 = pat_slot[p];
                inputs(s0, in);
                wls_record(in, &tb.wls_d[p * WLS_STRIDE]);
                fast_record(&tb.wls_d[p * WLS_STRIDE], &tb.wls_i[(size_t)s0 * WLSI_STRIDE], &tb.wls_f[p * WF_STRIDE]);
            }
        });
    }
    pt.mark("tables: slot patterns");
    tb.bc_partner.resize(tb.G);
    for (int g = 0; g < tb.G; g++) tb.bc_partner[g] = t.bc_partner(t.bcl[g]);
    tb.l_max = lmax;

    // internal nodes whose restriction mean is needed, grouped by level, deepest first
    {
        std::unordered_map<int, int> idx_of;
        std::vector<std::vector<int>> by_level(64);
        std::vector<int> stack;
        auto need = [&](int root) {
            stack.push_back(root);
            while (!stack.empty()) {
                int c = stack.back();
                stack.pop_back();
                if (t.son[c] < 0 || idx_of.count(c)) continue;
                idx_of[c] = -1;
                by_level[t.lvl[c]].push_back(c);
                for (int k = 0; k < 4; k++) stack.push_back(t.son[c] + k);
            }
        };
        for (int32_t p : tb.bc_partner) if (p >= 0) need(p);
        for (int32_t r : tb.rst_node) need(r);
        std::vector<int> order;
        tb.int_level_start.clear();
        for (int l = 63; l >= 0; l--) {
            if (by_level[l].empty()) continue;
            tb.int_level_start.push_back((int)order.size());
            for (int c : by_level[l]) { idx_of[c] = (int)order.size(); order.push_back(c); }
        }
        tb.int_level_start.push_back((int)order.size());
        tb.n_int = (int)order.size();
        auto ref = [&](int c) { return t.son[c] < 0 ? tb.state_of_node[c] : -(1 + idx_of.at(c)); };
        tb.int_child.resize(4 * (size_t)tb.n_int);
        for (int i = 0; i < tb.n_int; i++)
            for (int k = 0; k < 4; k++) tb.int_child[4 * (size_t)i + k] = ref(t.son[order[i]] + k);
        tb.rst_ref.resize(tb.S_r);
        for (int s = 0; s < tb.S_r; s++) tb.rst_ref[s] = idx_of.at(tb.rst_node[s]);
        tb.bc_ref.resize(tb.G);
        for (int g = 0; g < tb.G; g++) tb.bc_ref[g] = tb.bc_partner[g] < 0 ? REF_INFLOW : ref(tb.bc_partner[g]);
        // flattened fill (fast path)
        const int items = tb.G + tb.S_r;
        tb.fill_beg.assign(items + 1, 0);
        tb.fill_leaf.clear();
        tb.fill_w.clear();
        std::vector<std::pair<int32_t, float>> work;
        for (int k = 0; k < items; k++) {
            tb.fill_beg[k] = (int)tb.fill_leaf.size();
            const int32_t r = k < tb.G ? tb.bc_ref[k] : -(1 + tb.rst_ref[k - tb.G]);
            if (r == REF_INFLOW) {
                tb.fill_leaf.push_back(-1);
                tb.fill_w.push_back(0.f);
                continue;
            }
            work.assign(1, {r, 1.0f});
            while (!work.empty()) {
                const auto [q, w] = work.back();
                work.pop_back();
                if (q >= 0) {
                    tb.fill_leaf.push_back(q);
                    tb.fill_w.push_back(w);
                } else {
                    const size_t i = (size_t)(-(q + 1));
                    for (int c = 3; c >= 0; c--) work.push_back({tb.int_child[4 * i + c], 0.25f * w});
                }
            }
        }
        tb.fill_beg[items] = (int)tb.fill_leaf.size();
        tb.fill_flat_ok = true;
        for (int32_t q : tb.fill_leaf)
            if (q >= tb.N) tb.fill_flat_ok = false;
    }

    pt.mark("tables: restriction + fill");
    // Ghosts whose boundary partner is an interior leaf hold an exact copy of that leaf's state
    // after every fill, and all inflow ghosts hold the same inflow state, so every table reads
    // the leaf itself, or the first inflow ghost, instead. The fills still write every ghost
    // (the exported state keeps them), but no stage reads the aliased ones, and the fast
    // path's per-stage fill covers only the rest: the first inflow ghost, restriction means
    // and restriction slots (fill_fast). A WLS slot that still reads a ghost marks the point
    // in the ghost mask of its size word, so that the fast path can form the value from the
    // ghost's fill list in the same launch that writes the ghost.
    tb.ghost_alias.assign(tb.G, -1);
    int inflow0 = -1;
    for (int g = 0; g < tb.G; g++) {
        if (tb.bc_ref[g] >= 0 && tb.bc_ref[g] < tb.N) tb.ghost_alias[g] = tb.bc_ref[g];
        else if (tb.bc_ref[g] == REF_INFLOW) {
            if (inflow0 < 0) inflow0 = g;
            else tb.ghost_alias[g] = tb.N + inflow0;
        }
    }
    auto redirect = [&](int32_t r) -> int32_t {
        return (r >= tb.N && r < tb.N + tb.G && tb.ghost_alias[r - tb.N] >= 0) ? tb.ghost_alias[r - tb.N] : r;
    };
    // off-axis points of the 5x5 selector stencil (state ids, ghost aliases resolved)
    resize_fresh(tb.off16, 16 * (size_t)N);
    std::atomic<size_t> descents{0};
    parallel_for((size_t)N, [&](size_t b0, size_t e0) {
        size_t nd = 0;
        for (size_t ii = b0; ii < e0; ii++) {
            const int i = (int)ii;
            const int c = tb.leaf_node[i];
            const double dy = t.h / pow2i(t.lvl[c]), dx = t.xf * dy;
            // The points are cell centres at the leaf's level, so each lies at least half a
            // leaf width inside every box of that level or coarser: the nearest ancestor whose
            // box holds it is on the path from its base cell, and the descent can start there.
            // The ancestors with their half widths, nearest first (listed when first needed):
            int anc[64], na = -1;
            double ax[64], ay[64], ahx[64], ahy[64];
            // A same-level leaf at a along x whose stencil has a same-level leaf at b along y
            // (or the other way round): that leaf has the point at its centre, so it is the
            // leaf the descent below would find, and its id is in that stencil row already
            // (a same-level leaf's stencil id is its state id; the stencil's ghost aliases are
            // resolved after this). The same-level interior leaves at a along x and at b along
            // y, once per leaf:
            static constexpr int kXe[4] = {5, 1, 0, 4}, kYe[4] = {7, 3, 2, 6};  // offsets -2, -1, 1, 2
            auto same_level_leaf = [&](size_t e) -> int32_t {
                const int32_t j = tb.stencil[e];
                return tb.st_kind[e] == ST_COPY && j < N ? j : -1;
            };
            int32_t jx[4], jy[4];
            for (int u = 0; u < 4; u++) {
                jx[u] = same_level_leaf((size_t)i * 8 + kXe[u]);
                jy[u] = same_level_leaf((size_t)i * 8 + kYe[u]);
            }
            int q = 0;
            for (int vb = 0; vb < 4; vb++)
                for (int ua = 0; ua < 4; ua++, q++) {
                    const int a = ua < 2 ? ua - 2 : ua - 1, b = vb < 2 ? vb - 2 : vb - 1;
                    size_t e2 = SIZE_MAX;
                    if (jx[ua] >= 0 && tb.st_kind[(size_t)jx[ua] * 8 + kYe[vb]] == ST_COPY)
                        e2 = (size_t)jx[ua] * 8 + kYe[vb];
                    else if (jy[vb] >= 0 && tb.st_kind[(size_t)jy[vb] * 8 + kXe[ua]] == ST_COPY)
                        e2 = (size_t)jy[vb] * 8 + kXe[ua];
                    if (e2 != SIZE_MAX) {
                        tb.off16[(size_t)q * N + i] = redirect(tb.stencil[e2]);
                        continue;
                    }
                    const double px = t.x[c] + a * dx, py = t.y[c] + b * dy;
                    // Otherwise, when the x-neighbour (or y-neighbour) is a same-level leaf, the
                    // point is the one its stencil entry along y (along x) names, and that
                    // entry's node (a coarser leaf, or a same-level node whose sons the descent
                    // enters) holds it: the descent starts there.
                    int node = -1;
                    if (jx[ua] >= 0) node = tb.st_node[(size_t)jx[ua] * 8 + kYe[vb]];
                    else if (jy[vb] >= 0) node = tb.st_node[(size_t)jy[vb] * 8 + kXe[ua]];
                    if (node >= 0) {
                        while (t.son[node] >= 0) {
                            const bool rt = px > t.x[node], upq = py > t.y[node];
                            node = t.son[node] + (rt && upq ? 2 : rt ? 1 : upq ? 3 : 0);
                        }
                        tb.off16[(size_t)q * N + i] = redirect(tb.state_of_node[node]);
                        continue;
                    }
                    if (na < 0) {
                        na = 0;
                        for (int f = t.father[c], k = 1; f >= 0 && na < 64; f = t.father[f], k++) {
                            anc[na] = f; ax[na] = t.x[f]; ay[na] = t.y[f];
                            ahx[na] = 0.5 * dx * (double)(1 << k); ahy[na] = 0.5 * dy * (double)(1 << k);
                            na++;
                        }
                    }
                    for (int k = 0; k < na; k++)
                        if (std::fabs(px - ax[k]) < ahx[k] && std::fabs(py - ay[k]) < ahy[k]) { node = anc[k]; break; }
                    if (node < 0) {
                        int nj1 = (int)std::floor((px - t.Xmin) / (t.xf * t.h)) + 1;
                        int ni1 = (int)std::floor((py - t.Ymin) / t.h) + 1;
                        nj1 = std::min(std::max(nj1, 1), t.m);
                        ni1 = std::min(std::max(ni1, 1), t.n);
                        node = (ni1 - 1) * t.m + nj1 - 1;
                    }
                    while (t.son[node] >= 0) {
                        const bool rt = px > t.x[node], upq = py > t.y[node];
                        node = t.son[node] + (rt && upq ? 2 : rt ? 1 : upq ? 3 : 0);
                    }
                    tb.off16[(size_t)q * N + i] = redirect(tb.state_of_node[node]);
                    nd++;
                }
        }
        descents += nd;
    });
    if (pt.on) std::fprintf(stderr, "[fdmamr]   off16: %zu of %zu points by descent from an ancestor\n", descents.load(), 16 * (size_t)N);
    pt.mark("tables: off16");
    parallel_for(tb.stencil.size(), [&](size_t b0, size_t e0) {
        for (size_t k = b0; k < e0; k++) tb.stencil[k] = redirect(tb.stencil[k]);
    });
    parallel_for((size_t)tb.S_w, [&](size_t b0, size_t e0) {
        for (size_t s = b0; s < e0; s++) {
            int32_t* iv = &tb.wls_i[s * WLSI_STRIDE];
            int gmask = 0;
            for (int j = 0; j < iv[WLSI_M]; j++) {
                int32_t& r = iv[WLSI_IDS + j];
                r = redirect(r);
                if (r >= tb.N && r < tb.N + tb.G) gmask |= 1 << j;
            }
            iv[WLSI_M] |= gmask << WLSI_GSHIFT;
        }
    });
    // shared faces (the FP64 stage computes such a face once): after every redirect, so the
    // comparison is on the ids the kernels read
    resize_fresh(tb.face_share, N);
    parallel_for((size_t)N, [&](size_t b0, size_t e0) {
        for (size_t ii = b0; ii < e0; ii++) {
            const int32_t* sv = &tb.stencil[ii * 8];
            uint8_t f = 0;
            const int32_t L = sv[1], B = sv[3];
            if (L >= 0 && L < N) {
                const int32_t* sl = &tb.stencil[(size_t)L * 8];
                if (sl[0] == (int32_t)ii && sl[1] == sv[5] && sl[4] == sv[0]) f |= 1;
            }
            if (B >= 0 && B < N) {
                const int32_t* sb = &tb.stencil[(size_t)B * 8];
                if (sb[2] == (int32_t)ii && sb[3] == sv[7] && sb[6] == sv[2]) f |= 2;
            }
            tb.face_share[ii] = f;
        }
    });
    tb.face_lo.clear();
    append_where((size_t)N, [&](size_t i) { return !(tb.face_share[i] & 1); }, tb.face_lo);
    tb.n_lo_x = (int)tb.face_lo.size();
    append_where((size_t)N, [&](size_t i) { return !(tb.face_share[i] & 2); }, tb.face_lo);
    tb.n_lo_y = (int)tb.face_lo.size() - tb.n_lo_x;
    // slots that read a ghost go last (slots [0, S_wp) read none), each part in its old order
    {
        BigVec<int32_t> plain, ghosted;
        auto reads_ghost = [&](size_t s) { return (tb.wls_i[s * WLSI_STRIDE + WLSI_M] >> WLSI_GSHIFT) != 0; };
        append_where((size_t)tb.S_w, [&](size_t s) { return !reads_ghost(s); }, plain);
        append_where((size_t)tb.S_w, reads_ghost, ghosted);
        tb.S_wp = (int)plain.size();
        BigVec<int32_t> newidx(tb.S_w);
        parallel_for(plain.size(), [&](size_t b0, size_t e0) {
            for (size_t k = b0; k < e0; k++) newidx[plain[k]] = (int32_t)k;
        });
        parallel_for(ghosted.size(), [&](size_t b0, size_t e0) {
            for (size_t k = b0; k < e0; k++) newidx[ghosted[k]] = (int32_t)(plain.size() + k);
        });
        if (tb.S_wp < tb.S_w) {
            BigVec<int32_t> wi(tb.wls_i.size());
            parallel_for((size_t)tb.S_w, [&](size_t b0, size_t e0) {
                for (size_t s = b0; s < e0; s++)
                    std::memcpy(&wi[(size_t)newidx[s] * WLSI_STRIDE], &tb.wls_i[s * WLSI_STRIDE],
                                WLSI_STRIDE * sizeof(int32_t));
            });
            tb.wls_i.swap(wi);
            parallel_for(tb.stencil.size(), [&](size_t b0, size_t e0) {
                for (size_t k = b0; k < e0; k++) {
                    int32_t& r = tb.stencil[k];
                    if (r >= tb.L && r < tb.L + tb.S_w) r = tb.L + newidx[r - tb.L];
                }
            });
        }
    }
    // the ghost points of the slots [S_wp, S_w) as fixed-size entry lists, so that the fast
    // path gathers them without walking the fill lists
    {
        const int D = tb.S_w - tb.S_wp;
        tb.dep_hdr.assign((size_t)D * 4, 0);
        tb.dep_leaf.assign((size_t)D * 2 * DEP_ENT, -1);
        tb.dep_w.assign((size_t)D * 2 * DEP_ENT, 0.f);
        for (int d = 0; d < D; d++) {
            const int32_t* iv = &tb.wls_i[(size_t)(tb.S_wp + d) * WLSI_STRIDE];
            const int gm = iv[WLSI_M] >> WLSI_GSHIFT;
            int cnt = 0;
            bool fits = true;
            for (int j = 0; j < 9 && fits; j++) {
                if (!((gm >> j) & 1)) continue;
                const int k = iv[WLSI_IDS + j] - tb.N;
                const int b0 = tb.fill_beg[k], e0 = tb.fill_beg[k + 1];
                if (cnt == 2 || e0 - b0 > DEP_ENT) { fits = false; break; }
                tb.dep_hdr[(size_t)d * 4 + 1 + cnt] = j;
                for (int u = 0; u < e0 - b0; u++) {
                    const size_t o = ((size_t)d * 2 + cnt) * DEP_ENT + u;
                    const int32_t q = tb.fill_leaf[b0 + u];
                    tb.dep_leaf[o] = q < 0 ? DEP_INFLOW : q;
                    tb.dep_w[o] = q < 0 ? 1.f : tb.fill_w[b0 + u];
                }
                cnt++;
            }
            tb.dep_hdr[(size_t)d * 4] = fits ? cnt : DEP_GENERIC;
        }
    }
    // each slot block's distinct records, in first-use order
    {
        const int nbs = (tb.S_w + SLOT_BLK - 1) / SLOT_BLK;
        tb.slot_blk_pat.assign((size_t)nbs * SLOT_BLK_PATS, -1);
        tb.slot_lpat.assign(tb.S_w, 255);
        tb.slot_uncov.clear();
        std::vector<uint8_t> blk_uncov(nbs, 0);
        parallel_for((size_t)nbs, [&](size_t b0, size_t e0) {
            for (size_t b = b0; b < e0; b++) {
                int32_t* bp = &tb.slot_blk_pat[b * SLOT_BLK_PATS];
                int np = 0;
                for (int s = (int)b * SLOT_BLK; s < std::min(tb.S_w, ((int)b + 1) * SLOT_BLK); s++) {
                    const int32_t pat = tb.wls_i[(size_t)s * WLSI_STRIDE + WLSI_PAT];
                    int q = 0;
                    while (q < np && bp[q] != pat) q++;
                    if (q == np && np < SLOT_BLK_PATS) bp[np++] = pat;
                    if (q < np) tb.slot_lpat[s] = (uint8_t)q;
                    else blk_uncov[b] = 1;
                }
            }
        }, 64);
        for (int b = 0; b < nbs; b++)
            if (blk_uncov[b])
                for (int s = b * SLOT_BLK; s < std::min(tb.S_w, (b + 1) * SLOT_BLK); s++)
                    if (tb.slot_lpat[s] == 255) tb.slot_uncov.push_back(s);
    }
    tb.merge_ok = tb.fill_flat_ok;
    for (size_t d = 0; d * 4 < tb.dep_hdr.size(); d++)
        if (tb.dep_hdr[d * 4] == DEP_GENERIC) tb.merge_ok = false;
    tb.fill_fast.clear();
    for (int k = 0; k < tb.G + tb.S_r; k++)
        if (k >= tb.G || tb.ghost_alias[k] < 0) tb.fill_fast.push_back(k);

    pt.mark("tables: aliases, faces, slots");
    // selector tile neighbourhoods (see Tables::sel_nbr)
    {
        const int ntiles = (N + 15) / 16;
        auto point_src = [&](int i, int p) -> int32_t {
            if (p >= 25 || p == 12) return i;
            const int a = p % 5 - 2, b = p / 5 - 2;
            if (b == 0) return tb.stencil[(size_t)i * 8 + (a == 1 ? 0 : a == -1 ? 1 : a == 2 ? 4 : 5)];
            if (a == 0) return tb.stencil[(size_t)i * 8 + (b == 1 ? 2 : b == -1 ? 3 : b == 2 ? 6 : 7)];
            const int q16 = (b + 2 - (b > 0 ? 1 : 0)) * 4 + (a + 2 - (a > 0 ? 1 : 0));
            return tb.off16[(size_t)q16 * N + i];
        };
        // Tiles in runs; each run keeps its tiles' lists in one buffer, and the buffers are
        // then placed at the prefix sums of the tile counts.
        const size_t nrun = std::min<size_t>(256, std::max(1, ntiles));
        const size_t per = ((size_t)ntiles + nrun - 1) / nrun;
        std::vector<std::vector<int32_t>> runbuf(nrun);
        std::vector<int32_t> tile_cnt(ntiles);
        resize_fresh(tb.sel_loc, (size_t)N * 32);
        parallel_for(nrun, [&](size_t r0, size_t r1) {
            // The tile's own leaves come first, at positions 0..15 (every leaf reads itself);
            // the other points in order of first use, found through an open-addressing set of
            // at most 16 x 28 keys (1024 slots, cleared after each tile).
            std::vector<int32_t> key(1024, -1), val(1024, 0), used, U;
            for (size_t r = r0; r < r1; r++)
                for (size_t T = r * per; T < std::min((size_t)ntiles, (r + 1) * per); T++) {
                    U.clear();
                    const int i0 = (int)T * 16, i1 = std::min(N, i0 + 16);
                    for (int i = i0; i < i1; i++) U.push_back(i);
                    for (int i = i0; i < i1; i++)
                        for (int g = 0; g < 4; g++)
                            for (int j = 0; j < 7; j++) {
                                const int32_t src = point_src(i, g + 4 * j);
                                if (src >= i0 && src < i1) {
                                    tb.sel_loc[(size_t)i * 32 + g * 8 + j] = (uint8_t)(src - i0);
                                    continue;
                                }
                                uint32_t h = ((uint32_t)src * 2654435761u) >> 22;
                                while (key[h] != -1 && key[h] != src) h = (h + 1) & 1023;
                                if (key[h] == -1) {
                                    key[h] = src;
                                    val[h] = (int)U.size();
                                    U.push_back(src);
                                    used.push_back((int)h);
                                }
                                tb.sel_loc[(size_t)i * 32 + g * 8 + j] = (uint8_t)std::min(val[h], 255);
                            }
                    // byte 7 of every lane: the centre (point 12 = lane 0, entry 3)
                    for (int i = i0; i < i1; i++)
                        for (int g = 0; g < 4; g++) tb.sel_loc[(size_t)i * 32 + g * 8 + 7] = tb.sel_loc[(size_t)i * 32 + 3];
                    for (int h : used) key[h] = -1;
                    used.clear();
                    if ((int)U.size() <= SEL_NBR_CAP) {
                        runbuf[r].insert(runbuf[r].end(), U.begin(), U.end());
                        tile_cnt[T] = (int32_t)U.size();
                    } else {
                        runbuf[r].resize(runbuf[r].size() + SEL_NBR_CAP + 1, 0);  // marker: gathered directly
                        tile_cnt[T] = SEL_NBR_CAP + 1;
                    }
                }
        }, 1);
        tb.sel_nbr_beg.assign(ntiles + 1, 0);
        for (int T = 0; T < ntiles; T++) tb.sel_nbr_beg[T + 1] = tb.sel_nbr_beg[T] + tile_cnt[T];
        resize_fresh(tb.sel_nbr, (size_t)tb.sel_nbr_beg[ntiles]);  // every element is copied below
        std::atomic<int> bad_nbr{0};
        parallel_for(nrun, [&](size_t r0, size_t r1) {
            for (size_t r = r0; r < r1; r++) {
                if (r * per >= (size_t)ntiles) continue;
                const size_t at = (size_t)tb.sel_nbr_beg[r * per];
                std::copy(runbuf[r].begin(), runbuf[r].end(), tb.sel_nbr.begin() + at);
                for (int32_t q : runbuf[r])
                    if (q < 0 || q >= tb.L + tb.S_w + tb.S_r) bad_nbr = 1;
            }
        }, 1);
        if (bad_nbr) throw std::runtime_error("selector neighbourhood index out of range");
    }
    pt.mark("tables: selector tiles");
    // the largest tables are checked on all host threads
    std::atomic<int> bad_loc{0}, bad_off{0}, bad_alias{0};
    parallel_for(tb.sel_nbr_beg.empty() ? 0 : tb.sel_nbr_beg.size() - 1, [&](size_t b0, size_t e0) {
        for (size_t T = b0; T < e0; T++) {
            const int nU = tb.sel_nbr_beg[T + 1] - tb.sel_nbr_beg[T];
            if (nU > SEL_NBR_CAP) continue;
            for (size_t k = T * 16 * 32; k < std::min((size_t)N, T * 16 + 16) * 32; k++)
                if (tb.sel_loc[k] >= nU) bad_loc = 1;
        }
    });
    if (bad_loc) throw std::runtime_error("selector neighbourhood position out of range");
    parallel_for(tb.off16.size(), [&](size_t b0, size_t e0) {
        for (size_t k = b0; k < e0; k++)
            if (tb.off16[k] < 0 || tb.off16[k] >= tb.L) bad_off = 1;
    });
    if (bad_off) throw std::runtime_error("selector stencil index out of range");
    for (size_t k = 0; k < tb.int_child.size(); k++) {
        const int32_t r = tb.int_child[k];
        const int self = (int)(k / 4);
        // a child is a leaf state or an internal node evaluated in an earlier (deeper) group
        if (r >= tb.L || (r < 0 && -(r + 1) >= self))
            throw std::runtime_error("restriction child out of range");
    }
    for (int32_t q : tb.fill_leaf)
        if (q < -1 || q >= tb.L) throw std::runtime_error("flattened fill leaf out of range");
    for (int32_t k : tb.fill_fast)
        if (k < 0 || k >= tb.G + tb.S_r) throw std::runtime_error("fast fill item out of range");
    for (int32_t p : tb.slot_blk_pat)
        if (p < -1 || p >= tb.n_pat) throw std::runtime_error("staged slot record out of range");
    parallel_for((size_t)tb.S_w, [&](size_t b0, size_t e0) {
        for (size_t s = b0; s < e0; s++) {
            const int lp = tb.slot_lpat[s];
            if (lp != 255 && (lp >= SLOT_BLK_PATS ||
                              tb.slot_blk_pat[(s / SLOT_BLK) * SLOT_BLK_PATS + lp] !=
                                  tb.wls_i[s * WLSI_STRIDE + WLSI_PAT]))
                throw std::runtime_error("staged slot record wrong");
        }
    });
    for (int32_t s : tb.slot_uncov)
        if (s < 0 || s >= tb.S_w || tb.slot_lpat[s] != 255) throw std::runtime_error("unstaged slot list wrong");
    for (size_t d = 0; d * 4 < tb.dep_hdr.size(); d++) {
        const int cnt = tb.dep_hdr[d * 4];
        if (cnt == DEP_GENERIC) continue;
        if (cnt < 1 || cnt > 2) throw std::runtime_error("ghost entry count out of range");
        for (int p = 0; p < cnt; p++) {
            if (tb.dep_hdr[d * 4 + 1 + p] < 0 || tb.dep_hdr[d * 4 + 1 + p] > 8) throw std::runtime_error("ghost point out of range");
            for (int u = 0; u < DEP_ENT; u++) {
                const int32_t q = tb.dep_leaf[(d * 2 + p) * DEP_ENT + u];
                if (q != -1 && q != DEP_INFLOW && (q < 0 || q >= tb.N)) throw std::runtime_error("ghost entry out of range");
            }
        }
    }
    parallel_for(tb.stencil.size(), [&](size_t b0, size_t e0) {
        for (size_t k = b0; k < e0; k++) {
            const int32_t r = tb.stencil[k];
            if (r >= tb.N && r < tb.N + tb.G && tb.ghost_alias[r - tb.N] >= 0) bad_alias = 1;
        }
    });
    if (bad_alias) throw std::runtime_error("stencil reads an aliased ghost");
    for (int32_t r : tb.rst_ref)
        if (r < 0 || r >= tb.n_int) throw std::runtime_error("restriction slot ref out of range");
    for (int32_t r : tb.bc_ref)
        if (r != REF_INFLOW && (r >= tb.L || (r < 0 && -(r + 1) >= tb.n_int)))
            throw std::runtime_error("ghost ref out of range");
    parallel_for((size_t)tb.S_w, [&](size_t b0, size_t e0) {
        for (size_t s = b0; s < e0; s++) {
            const int32_t* iv = &tb.wls_i[s * WLSI_STRIDE];
            const int m = iv[WLSI_M] & WLSI_MMASK, gmask = iv[WLSI_M] >> WLSI_GSHIFT;
            if (m < 1 || m > 9 || iv[WLSI_M1] < 0 || iv[WLSI_M1] > m || (gmask >> m) != 0)
                throw std::runtime_error("WLS stencil size out of range");
            for (int i = 0; i < m; i++) {
                const int32_t r = iv[WLSI_IDS + i];
                if (r < 0 || r >= tb.L) throw std::runtime_error("WLS id out of range");
                // the ghost mask names exactly the points that are ghost leaves
                if (((gmask >> i) & 1) != (r >= tb.N && r < tb.N + tb.G)) throw std::runtime_error("WLS ghost mask wrong");
            }
            if ((gmask != 0) != ((int)s >= tb.S_wp)) throw std::runtime_error("WLS slot order wrong");
            if (iv[WLSI_PAT] < 0 || iv[WLSI_PAT] >= tb.n_pat) throw std::runtime_error("WLS record index out of range");
        }
    });
    for (int32_t q : tb.rst_node)
        if (q < 0 || (size_t)q >= NA || t.son[q] < 0) throw std::runtime_error("restriction node invalid");
    for (int32_t q : tb.bc_partner)
        if (q >= 0 && ((size_t)q >= NA || t.father[q] == -2 || (t.son[q] < 0 && tb.state_of_node[q] < 0)))
            throw std::runtime_error("boundary partner invalid");
    parallel_for(NA, [&](size_t b0, size_t e0) {
        for (size_t c = b0; c < e0; c++) {
            if (t.father[c] == -2) continue;
            if (t.son[c] >= 0 && (size_t)t.son[c] + 3 >= NA) throw std::runtime_error("son index out of range");
            if (t.son[c] < 0 && tb.state_of_node[c] < 0) throw std::runtime_error("leaf without state slot");
        }
    });
    pt.mark("tables: checks");
    // 16-byte stencil records (Tables::st16): per block of ST16_BLK leaves the smallest id of
    // a ghost or slot its leaves read is the base of the far entries; then every record is
    // encoded and decoded back, and must give the 32-byte record's ids
    {
        const size_t nblk = ((size_t)N + ST16_BLK - 1) / ST16_BLK;
        tb.st16_base.assign(std::max<size_t>(nblk, 1), 0);
        resize_fresh(tb.st16, (size_t)N * 8);
        std::atomic<int> escaped{0}, bad{0};
        parallel_for(nblk, [&](size_t b0, size_t e0) {
            int esc = 0;
            for (size_t b = b0; b < e0; b++) {
                const size_t i0 = b * ST16_BLK, i1 = std::min((size_t)N, i0 + ST16_BLK);
                int32_t base = INT32_MAX;
                for (size_t k = i0 * 8; k < i1 * 8; k++)
                    if (tb.stencil[k] >= N) base = std::min(base, tb.stencil[k]);
                if (base == INT32_MAX) base = 0;
                tb.st16_base[b] = base;
                for (size_t i = i0; i < i1; i++) {
                    int16_t e[8];
                    bool ok = true;
                    for (int k = 0; k < 8; k++) {
                        const int64_t r = tb.stencil[i * 8 + k];
                        const int64_t d = r - (int64_t)i, f = r - base;
                        if (r < N && d > -ST16_FAR && d < ST16_FAR) e[k] = (int16_t)d;
                        else if (r >= N && f >= 0 && f < 32768 - ST16_FAR) e[k] = (int16_t)(ST16_FAR + f);
                        else ok = false;
                    }
                    for (int k = 0; k < 8; k++) {
                        tb.st16[i * 8 + k] = ok ? e[k] : ST16_ESC;
                        const int32_t back = e[k] >= ST16_FAR ? base + (e[k] - ST16_FAR) : (int32_t)i + e[k];
                        if (ok && back != tb.stencil[i * 8 + k]) bad = 1;
                    }
                    esc += !ok;
                }
            }
            escaped += esc;
        });
        if (bad) throw std::runtime_error("compressed stencil record wrong");
        tb.st16_escaped = escaped;
    }
    pt.mark("tables: 16-byte records");
}

}  // namespace fdm
