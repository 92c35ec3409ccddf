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

// WLS-WENO geometry of one slot (XuanGua.F90:206-255 and the box of :318-323)
void wls_geometry(const Tree& t, int cell, double xx, double yy, const List& l1, const List& l2,
                  const std::vector<int32_t>& state_of_node, double* d, int32_t* iv) {
    int m1 = l1.n, m = l1.n + l2.n;
    double px[16], py[16], w[16], P[16][6];
    // cell-local scaled frame (optional): x' = (x - xc)/hx, box [-1/2, 1/2]^2
    const double hxc = t.xf * t.h / pow2i(t.lvl[cell]), hyc = t.h / pow2i(t.lvl[cell]);
    if (t.wls_local) {
        xx = (xx - t.x[cell]) / hxc;
        yy = (yy - t.y[cell]) / hyc;
    }
    for (int i = 0; i < m; i++) {
        int q = i < m1 ? l1.a[i] : l2.a[i - m1];
        px[i] = t.x[q]; py[i] = t.y[q];
        if (t.wls_local) { px[i] = (px[i] - t.x[cell]) / hxc; py[i] = (py[i] - t.y[cell]) / hyc; }
        double dd = std::sqrt((px[i] - xx) * (px[i] - xx) + (py[i] - yy) * (py[i] - yy));
        w[i] = 1.0 / (dd + 1e-6);
        iv[WLSI_IDS + i] = state_of_node[q];
        d[WLS_PX + i] = px[i]; d[WLS_PY + i] = py[i]; d[WLS_W + i] = w[i];
        P[i][0] = 1.0; P[i][1] = px[i]; P[i][2] = py[i];
        P[i][3] = px[i] * py[i]; P[i][4] = px[i] * px[i]; P[i][5] = py[i] * py[i];
    }
    iv[WLSI_M] = m; iv[WLSI_M1] = m1;
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
    double hx = t.xf * t.h / std::ldexp(1.0, t.lvl[cell] + 1);
    double hy = t.h / std::ldexp(1.0, t.lvl[cell] + 1);
    double x_min = t.x[cell] - hx, x_max = t.x[cell] + hx;
    double y_min = t.y[cell] - hy, y_max = t.y[cell] + hy;
    if (t.wls_local) { x_min = -0.5; x_max = 0.5; y_min = -0.5; y_max = 0.5; }
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

void build_tables(const Tree& t, Tables& tb) {
    const size_t NA = t.n_alloc();
    tb = Tables();
    tb.N = (int)t.ns.size();
    tb.G = (int)t.bcl.size();
    tb.state_of_node.assign(NA, -1);
    tb.leaf_node.clear();
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
        // runs of cell_NS that share a base cell (cell_NS walks the base cells row-major)
        std::vector<std::pair<uint64_t, int>> runs;  // (Morton key of the base cell, first position)
        for (int e = 0; e < tb.N; e++) {
            const int c = t.ns[e];
            if (e == 0 || t.ni[c] != t.ni[t.ns[e - 1]] || t.nj[c] != t.nj[t.ns[e - 1]])
                runs.emplace_back(spread((uint32_t)(t.nj[c] - 1)) | (spread((uint32_t)(t.ni[c] - 1)) << 1), e);
        }
        std::vector<int> run_end(runs.size());
        for (size_t r = 0; r < runs.size(); r++) run_end[r] = r + 1 < runs.size() ? runs[r + 1].second : tb.N;
        std::vector<int> order(runs.size());
        for (size_t r = 0; r < runs.size(); r++) order[r] = (int)r;
        std::sort(order.begin(), order.end(), [&](int a, int b) { return runs[a].first < runs[b].first; });
        tb.ns_pos.resize(tb.N);
        tb.dev_of_ns.resize(tb.N);
        int i = 0;
        for (int r : order)
            for (int e = runs[r].second; e < run_end[r]; e++) { tb.ns_pos[i] = e; tb.dev_of_ns[e] = i; i++; }
    }
    for (int i = 0; i < tb.N; i++) {
        const int c = t.ns[tb.ns_pos[i]];
        tb.state_of_node[c] = i;
        tb.leaf_node.push_back(c);
    }
    for (int i = 0; i < tb.G; i++) { tb.state_of_node[t.bcl[i]] = tb.N + i; tb.leaf_node.push_back(t.bcl[i]); }
    // passive leaves: corners (never updated, solution.F90:644,682-686)
    for (size_t c = 0; c < NA; c++)
        if (t.father[c] != -2 && t.son[c] < 0 && tb.state_of_node[c] < 0) {
            tb.state_of_node[c] = (int)tb.leaf_node.size();
            tb.leaf_node.push_back((int)c);
        }
    tb.L = (int)tb.leaf_node.size();
    tb.P = tb.L - tb.N - tb.G;
    tb.node_son.assign(t.son.begin(), t.son.end());

    const int N = tb.N;
    tb.stencil.assign(8 * (size_t)N, -1);
    tb.lvl_leaf.assign(N, 0);
    tb.st_kind.assign(8 * (size_t)N, 0);
    tb.st_node.assign(8 * (size_t)N, -1);
    tb.st_px.assign(8 * (size_t)N, 0.0);
    tb.st_py.assign(8 * (size_t)N, 0.0);

    // phase 1 (parallel): classify every stencil entry as the reference's gathers do
    // (Tip/Tim/Tjp/Tjm for k < 4, Tipp/Timm/Tjpp/Tjmm for k >= 4, XuanGua.F90:401-590)
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
                double off = two ? (dir < 2 ? 2.0 * t.xf * t.h : 2.0 * t.h) / pow2i(t.lvl[c])
                                 : (dir < 2 ? t.xf * t.h : t.h) / pow2i(t.lvl[c]);
                double xx = t.x[c], yy = t.y[c];
                if (dir == 0) xx = t.x[c] + off;
                else if (dir == 1) xx = t.x[c] - off;
                else if (dir == 2) yy = t.y[c] + off;
                else yy = t.y[c] - off;
                int tg;
                bool same;
                if (!two) { tg = cn; same = t.lvl[cn] == t.lvl[c]; }
                else if (t.lvl[cn] == t.lvl[c]) {
                    int cnn = t.nb[8 * (size_t)cn + nbk];
                    if (cnn < 0) { missing = 1; continue; }
                    tg = cnn; same = t.lvl[cnn] == t.lvl[cn];
                } else { tg = cn; same = false; }
                int kind;
                if (same) {
                    if (t.son[tg] < 0) kind = ST_COPY;
                    else {
                        List l1, l2;
                        parent_lists(t, tg, l1, l2);
                        xx = t.x[tg]; yy = t.y[tg];
                        kind = l1.n + l2.n < 7 ? ST_PARENT_RESTRICT : ST_PARENT_WLS;
                    }
                } else {
                    List l1, l2;
                    child_lists(t, tg, l1, l2);
                    kind = l1.n + l2.n < 7 ? ST_CHILD_COPY : ST_CHILD_WLS;
                }
                const size_t d = (size_t)i * 8 + k;
                tb.st_kind[d] = (int8_t)kind;
                tb.st_node[d] = tg;
                tb.st_px[d] = xx; tb.st_py[d] = yy;
            }
        }
        int cur = lmax_at.load();
        while (lm > cur && !lmax_at.compare_exchange_weak(cur, lm)) {}
    });
    if (missing) throw std::runtime_error("interior leaf without its stencil neighbours");
    const int lmax = lmax_at.load();

    // phase 2 (serial, deterministic order): one slot per distinct level-jump value
    std::unordered_map<Key, int, KeyHash> wls_map, rst_map;
    {
        // size the slot map once (a level-jump value per stencil entry at most)
        size_t jumps = 0;
        for (size_t d = 0; d < tb.st_kind.size(); d++)
            jumps += tb.st_kind[d] == ST_PARENT_WLS || tb.st_kind[d] == ST_CHILD_WLS;
        wls_map.reserve(jumps);
    }
    std::vector<int32_t> wls_node;      // per WLS slot: the WLS cell
    std::vector<int8_t> wls_parent;     // per WLS slot: parent (1) or child (0) stencil
    std::vector<double> wls_xx, wls_yy;
    for (int i = 0; i < N; i++)
        for (int k = 0; k < 8; k++) {
            const size_t d = (size_t)i * 8 + k;
            const int kind = tb.st_kind[d], tg = tb.st_node[d];
            int ref;
            if (kind == ST_COPY || kind == ST_CHILD_COPY) ref = tb.state_of_node[tg];
            else if (kind == ST_PARENT_RESTRICT) {
                Key key{2, tg, 0, 0};
                auto it = rst_map.find(key);
                int s;
                if (it == rst_map.end()) { s = (int)tb.rst_node.size(); rst_map.emplace(key, s); tb.rst_node.push_back(tg); }
                else s = it->second;
                ref = -2 - s;  // patched to L + S_w + s below
            } else {
                const bool parent = kind == ST_PARENT_WLS;
                Key key{parent ? 1 : 3, tg, parent ? 0 : bits(tb.st_px[d]), parent ? 0 : bits(tb.st_py[d])};
                auto it = wls_map.find(key);
                int s;
                if (it == wls_map.end()) {
                    s = (int)wls_node.size();
                    wls_map.emplace(key, s);
                    wls_node.push_back(tg);
                    wls_parent.push_back(parent ? 1 : 0);
                    wls_xx.push_back(tb.st_px[d]);
                    wls_yy.push_back(tb.st_py[d]);
                } else s = it->second;
                ref = 1000000000 + s;  // patched to L + s below
            }
            tb.stencil[(size_t)i * 8 + k] = ref;
        }
    tb.S_w = (int)wls_node.size();
    tb.S_r = (int)tb.rst_node.size();
    for (auto& r : tb.stencil) {
        if (r >= 1000000000) r = tb.L + (r - 1000000000);
        else if (r <= -2) r = tb.L + tb.S_w + (-2 - r);
    }

    // phase 3 (parallel): stencil lists and FP64 geometry of every WLS slot
    std::vector<double> recd((size_t)tb.S_w * WLS_STRIDE, 0.0);
    tb.wls_i.assign((size_t)tb.S_w * WLSI_STRIDE, -1);
    parallel_for((size_t)tb.S_w, [&](size_t b, size_t e) {
        for (size_t s = b; s < e; s++) {
            List l1, l2;
            const int tg = wls_node[s];
            if (wls_parent[s]) parent_lists(t, tg, l1, l2);
            else child_lists(t, tg, l1, l2);
            wls_geometry(t, tg, wls_xx[s], wls_yy[s], l1, l2, tb.state_of_node, &recd[s * WLS_STRIDE],
                         &tb.wls_i[s * WLSI_STRIDE]);
        }
    });
    // Records by pattern: in the cell-local frame a slot's geometry (coordinates in units of
    // the cell size, weights, inverted normal matrices, smoothness box) depends only on the
    // shape of the stencil around the level jump, so slots with the same shape have
    // bit-identical records and share one, numbered in slot order; wls_i[.. + WLSI_PAT] names
    // it. Both the FP64 records and the FP32 ones derived from them are stored once per
    // pattern. In the reference's absolute frame every record differs and each slot keeps its
    // own.
    {
        auto same = [&](size_t a, size_t b2) {
            const int32_t* ia = &tb.wls_i[a * WLSI_STRIDE];
            const int32_t* ib = &tb.wls_i[b2 * WLSI_STRIDE];
            return ia[WLSI_M] == ib[WLSI_M] && ia[WLSI_M1] == ib[WLSI_M1] &&
                   std::memcmp(&recd[a * WLS_STRIDE], &recd[b2 * WLS_STRIDE], WLS_STRIDE * sizeof(double)) == 0;
        };
        auto hash_rec = [&](size_t s) {
            uint64_t h = 1469598103934665603ull;
            const uint32_t* w = reinterpret_cast<const uint32_t*>(&recd[s * WLS_STRIDE]);
            for (int k = 0; k < 2 * WLS_STRIDE; k++) { h ^= w[k]; h *= 1099511628211ull; }
            h ^= (uint64_t)tb.wls_i[s * WLSI_STRIDE + WLSI_M] * 131u + (uint64_t)tb.wls_i[s * WLSI_STRIDE + WLSI_M1];
            return h * 1099511628211ull;
        };
        std::vector<uint64_t> hs(tb.S_w);
        parallel_for((size_t)tb.S_w, [&](size_t b, size_t e) {
            for (size_t s = b; s < e; s++) hs[s] = hash_rec(s);
        });
        std::unordered_map<uint64_t, std::vector<int>> buckets;
        std::vector<int> pat_slot;  // first slot of each pattern
        for (int s = 0; s < tb.S_w; s++) {
            auto& bk = buckets[hs[s]];
            int pid = -1;
            for (int p : bk)
                if (same((size_t)pat_slot[p], (size_t)s)) { pid = p; break; }
            if (pid < 0) {
                pid = (int)pat_slot.size();
                pat_slot.push_back(s);
                bk.push_back(pid);
            }
            tb.wls_i[(size_t)s * WLSI_STRIDE + WLSI_PAT] = pid;
        }
        tb.n_pat = (int)pat_slot.size();
        tb.wls_d.assign((size_t)tb.n_pat * WLS_STRIDE, 0.0);
        tb.wls_f.assign((size_t)tb.n_pat * WF_STRIDE, 0.0f);
        for (int p = 0; p < tb.n_pat; p++) {
            const size_t s0 = (size_t)pat_slot[p];
            std::memcpy(&tb.wls_d[(size_t)p * WLS_STRIDE], &recd[s0 * WLS_STRIDE], WLS_STRIDE * sizeof(double));
            fast_record(&tb.wls_d[(size_t)p * WLS_STRIDE], &tb.wls_i[s0 * WLSI_STRIDE], &tb.wls_f[(size_t)p * WF_STRIDE]);
        }
    }
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
    parallel_for(tb.stencil.size(), [&](size_t b0, size_t e0) {
        for (size_t k = b0; k < e0; k++) tb.stencil[k] = redirect(tb.stencil[k]);
    });
    for (int s = 0; s < tb.S_w; s++) {
        int32_t* iv = &tb.wls_i[(size_t)s * WLSI_STRIDE];
        int gmask = 0;
        for (int j = 0; j < iv[WLSI_M]; j++) {
            int32_t& r = iv[WLSI_IDS + j];
            r = redirect(r);
            if (r >= tb.N && r < tb.N + tb.G) gmask |= 1 << j;
        }
        iv[WLSI_M] |= gmask << WLSI_GSHIFT;
    }
    // shared faces (the FP64 stage computes such a face once): after every redirect, so the
    // comparison is on the ids the kernels read
    tb.face_share.assign(N, 0);
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
    for (int i = 0; i < N; i++)
        if (!(tb.face_share[i] & 1)) tb.face_lo.push_back(i);
    tb.n_lo_x = (int)tb.face_lo.size();
    for (int i = 0; i < N; i++)
        if (!(tb.face_share[i] & 2)) tb.face_lo.push_back(i);
    tb.n_lo_y = (int)tb.face_lo.size() - tb.n_lo_x;
    // slots that read a ghost go last (slots [0, S_wp) read none), each part in its old order
    {
        std::vector<int32_t> newidx(tb.S_w);
        int np = 0;
        for (int s = 0; s < tb.S_w; s++)
            if ((tb.wls_i[(size_t)s * WLSI_STRIDE + WLSI_M] >> WLSI_GSHIFT) == 0) newidx[s] = np++;
        tb.S_wp = np;
        for (int s = 0; s < tb.S_w; s++)
            if ((tb.wls_i[(size_t)s * WLSI_STRIDE + WLSI_M] >> WLSI_GSHIFT) != 0) newidx[s] = np++;
        if (tb.S_wp < tb.S_w) {
            std::vector<int32_t> wi(tb.wls_i.size());
            for (int s = 0; s < tb.S_w; s++)
                std::memcpy(&wi[(size_t)newidx[s] * WLSI_STRIDE], &tb.wls_i[(size_t)s * WLSI_STRIDE],
                            WLSI_STRIDE * sizeof(int32_t));
            tb.wls_i.swap(wi);
            for (auto& r : tb.stencil)
                if (r >= tb.L && r < tb.L + tb.S_w) r = tb.L + newidx[r - tb.L];
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
        for (int b = 0; b < nbs; b++) {
            int32_t* bp = &tb.slot_blk_pat[(size_t)b * SLOT_BLK_PATS];
            int np = 0;
            for (int s = b * SLOT_BLK; s < std::min(tb.S_w, (b + 1) * SLOT_BLK); s++) {
                const int32_t pat = tb.wls_i[(size_t)s * WLSI_STRIDE + WLSI_PAT];
                int q = 0;
                while (q < np && bp[q] != pat) q++;
                if (q == np && np < SLOT_BLK_PATS) bp[np++] = pat;
                if (q < np) tb.slot_lpat[s] = (uint8_t)q;
                else tb.slot_uncov.push_back(s);
            }
        }
    }
    tb.merge_ok = tb.fill_flat_ok;
    for (size_t d = 0; d * 4 < tb.dep_hdr.size(); d++)
        if (tb.dep_hdr[d * 4] == DEP_GENERIC) tb.merge_ok = false;
    tb.fill_fast.clear();
    for (int k = 0; k < tb.G + tb.S_r; k++)
        if (k >= tb.G || tb.ghost_alias[k] < 0) tb.fill_fast.push_back(k);

    // off-axis points of the 5x5 selector stencil
    tb.off16.assign(16 * (size_t)N, 0);
    parallel_for((size_t)N, [&](size_t b0, size_t e0) {
        for (size_t ii = b0; ii < e0; ii++) {
            const int i = (int)ii;
            const int c = tb.leaf_node[i];
            const double dy = t.h / pow2i(t.lvl[c]), dx = t.xf * dy;
            int q = 0;
            for (int b = -2; b <= 2; b++)
                for (int a = -2; a <= 2; a++) {
                    if (a == 0 || b == 0) continue;
                    const double px = t.x[c] + a * dx, py = t.y[c] + b * dy;
                    int nj1 = (int)std::floor((px - t.Xmin) / (t.xf * t.h)) + 1;
                    int ni1 = (int)std::floor((py - t.Ymin) / t.h) + 1;
                    nj1 = std::min(std::max(nj1, 1), t.m);
                    ni1 = std::min(std::max(ni1, 1), t.n);
                    int node = (ni1 - 1) * t.m + nj1 - 1;
                    while (t.son[node] >= 0) {
                        const bool rt = px > t.x[node], upq = py > t.y[node];
                        node = t.son[node] + (rt && upq ? 2 : rt ? 1 : upq ? 3 : 0);
                    }
                    tb.off16[(size_t)q * N + i] = redirect(tb.state_of_node[node]);
                    q++;
                }
        }
    });
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
        std::vector<std::vector<int32_t>> nbr(ntiles);
        tb.sel_loc.assign((size_t)N * 32, 0);
        parallel_for((size_t)ntiles, [&](size_t b0, size_t e0) {
            // open-addressing set of at most 16 x 28 keys (1024 slots, cleared after each tile)
            std::vector<int32_t> key(1024, -1), val(1024, 0), used;
            for (size_t T = b0; T < e0; T++) {
                std::vector<int32_t>& U = nbr[T];
                U.clear();
                const int i0 = (int)T * 16, i1 = std::min(N, i0 + 16);
                for (int i = i0; i < i1; i++)
                    for (int g = 0; g < 4; g++)
                        for (int j = 0; j < 7; j++) {
                            const int32_t src = point_src(i, g + 4 * j);
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
            }
        });
        tb.sel_nbr_beg.assign(ntiles + 1, 0);
        tb.sel_nbr.clear();
        for (int T = 0; T < ntiles; T++) {
            tb.sel_nbr_beg[T] = (int)tb.sel_nbr.size();
            if ((int)nbr[T].size() <= SEL_NBR_CAP) tb.sel_nbr.insert(tb.sel_nbr.end(), nbr[T].begin(), nbr[T].end());
            else tb.sel_nbr.resize(tb.sel_nbr.size() + SEL_NBR_CAP + 1, 0);  // marker: gathered directly
        }
        tb.sel_nbr_beg[ntiles] = (int)tb.sel_nbr.size();
    }
    for (int32_t r : tb.sel_nbr)
        if (r < 0 || r >= tb.L + tb.S_w + tb.S_r) throw std::runtime_error("selector neighbourhood index out of range");
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
    for (int s = 0; s < tb.S_w; s++) {
        const int lp = tb.slot_lpat[s];
        if (lp != 255 && (lp >= SLOT_BLK_PATS ||
                          tb.slot_blk_pat[(size_t)(s / SLOT_BLK) * SLOT_BLK_PATS + lp] !=
                              tb.wls_i[(size_t)s * WLSI_STRIDE + WLSI_PAT]))
            throw std::runtime_error("staged slot record wrong");
    }
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
    for (int s = 0; s < tb.S_w; s++) {
        const int32_t* iv = &tb.wls_i[(size_t)s * WLSI_STRIDE];
        const int m = iv[WLSI_M] & WLSI_MMASK, gmask = iv[WLSI_M] >> WLSI_GSHIFT;
        if (m < 1 || m > 9 || iv[WLSI_M1] < 0 || iv[WLSI_M1] > m || (gmask >> m) != 0)
            throw std::runtime_error("WLS stencil size out of range");
        for (int i = 0; i < m; i++) {
            const int32_t r = iv[WLSI_IDS + i];
            if (r < 0 || r >= tb.L) throw std::runtime_error("WLS id out of range");
            // the ghost mask names exactly the points that are ghost leaves
            if (((gmask >> i) & 1) != (r >= tb.N && r < tb.N + tb.G)) throw std::runtime_error("WLS ghost mask wrong");
        }
        if ((gmask != 0) != (s >= tb.S_wp)) throw std::runtime_error("WLS slot order wrong");
        if (iv[WLSI_PAT] < 0 || iv[WLSI_PAT] >= tb.n_pat) throw std::runtime_error("WLS record index out of range");
    }
    for (int32_t q : tb.rst_node)
        if (q < 0 || (size_t)q >= NA || t.son[q] < 0) throw std::runtime_error("restriction node invalid");
    for (int32_t q : tb.bc_partner)
        if (q >= 0 && ((size_t)q >= NA || t.father[q] == -2 || (t.son[q] < 0 && tb.state_of_node[q] < 0)))
            throw std::runtime_error("boundary partner invalid");
    for (size_t c = 0; c < NA; c++) {
        if (t.father[c] == -2) continue;
        if (t.son[c] >= 0 && (size_t)t.son[c] + 3 >= NA) throw std::runtime_error("son index out of range");
        if (t.son[c] < 0 && tb.state_of_node[c] < 0) throw std::runtime_error("leaf without state slot");
    }
}

}  // namespace fdm
