// host_tree.cpp -- quadtree topology, mesh operations, leaf lists and initial conditions.
// Each routine cites the reference routine whose behaviour it keeps. Neighbour finding must
// reproduce the reference bit for bit (it decides stencil shapes), so the finders follow
// the reference's case analysis exactly, including its asymmetric diagonal null rules.
#include "host_tree.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <exception>
#include <thread>

namespace fdm {

static inline double pow2i(int k) { return (double)(1 << k); }

uint64_t splitmix64(uint64_t& s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
void parallel_for(size_t n, const std::function<void(size_t, size_t)>& f) {
    unsigned nt = std::thread::hardware_concurrency();
    if (const char* e = std::getenv("FDMAMR_HOST_THREADS")) nt = (unsigned)std::max(1, std::atoi(e));
    nt = std::min<unsigned>(std::max(nt, 1u), 16u);
    if (n < 4096 || nt == 1) { f(0, n); return; }
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

int Tree::alloc_block() {
    if (!free_blocks.empty()) {
        int b = free_blocks.back();
        free_blocks.pop_back();
        return b;
    }
    int b = (int)father.size();
    size_t nn = father.size() + 4;
    father.resize(nn); son.resize(nn); lvl.resize(nn); sp.resize(nn); ni.resize(nn);
    nj.resize(nn); cross.resize(nn); x.resize(nn); y.resize(nn); U.resize(nn); rot.resize(nn);
    div.resize(nn);
    return b;
}

void Tree::release_subtree(int c) {
    int s = son[c];
    if (s < 0) return;
    for (int k = 0; k < 4; k++) release_subtree(s + k);
    for (int k = 0; k < 4; k++) { father[s + k] = -2; son[s + k] = -1; }
    free_blocks.push_back(s);
    son[c] = -1;
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

#define DIAG_BODY(RULE, HFN, VFN)                                                      \
    int hn = HFN(c), vn = VFN(c), f = father[c];                                         \
    if (f < 0) return (hn < 0 || vn < 0) ? -1 : VFN(hn);                                 \
    if (sp[c] == RULE.sib_sp) return son[f] + RULE.sib_son;                              \
    if (sp[c] == RULE.desc_sp) {                                                         \
        if (hn < 0 || vn < 0) return -1;                                                 \
        int r = VFN(hn);                                                                 \
        while (r >= 0 && lvl[r] < lvl[c] && son[r] >= 0) r = son[r] + RULE.desc_son;     \
        return r;                                                                        \
    }                                                                                    \
    if (sp[c] == RULE.hside_sp) return (hn < 0 || lvl[hn] < lvl[c]) ? -1 : VFN(hn);      \
    if (sp[c] == RULE.vside_sp) return (vn < 0 || lvl[vn] < lvl[c]) ? -1 : HFN(vn);      \
    return -1;

int Tree::upleft(int c) const { DIAG_BODY(kUL, left, up) }
int Tree::upright(int c) const { DIAG_BODY(kUR, right, up) }
int Tree::downright(int c) const { DIAG_BODY(kDR, right, down) }
int Tree::downleft(int c) const { DIAG_BODY(kDL, left, down) }
#undef DIAG_BODY

void Tree::idnei(int c) {                           // neighbor.F90:412-449
    int32_t* q = &nb[8 * (size_t)c];
    q[NB_UP] = up(c); q[NB_DOWN] = down(c); q[NB_RIGHT] = right(c); q[NB_LEFT] = left(c);
    if (cross[c] >= 0) {
        q[NB_UPLEFT] = upleft(c); q[NB_UPRIGHT] = upright(c);
        q[NB_DOWNLEFT] = downleft(c); q[NB_DOWNRIGHT] = downright(c);
    } else {
        q[NB_UPLEFT] = q[NB_UPRIGHT] = q[NB_DOWNLEFT] = q[NB_DOWNRIGHT] = -1;
    }
    if (son[c] >= 0) for (int k = 0; k < 4; k++) idnei(son[c] + k);
}
void Tree::idneighbor() {
    nb.assign(8 * n_alloc(), -1);
    // the finders only read the tree; each base cell's subtree writes its own entries
    parallel_for((size_t)total, [&](size_t b, size_t e) {
        for (size_t i = b; i < e; i++) idnei((int)i);
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
    int a = 0, b = 0;
    for (int i = 0; i < total; i++) { ns_walk(i, false, a); bc_walk(i, false, b); }
    ns.assign(a, -1); bcl.assign(b, -1);
    a = b = 0;
    for (int i = 0; i < total; i++) { ns_walk(i, true, a); bc_walk(i, true, b); }
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
void Tree::rot_div_order(std::vector<int32_t>& out) const {       // solution.F90:424-440
    out.clear();
    for (int i = 0; i < total; i++)
        if (cross[i] == 0 || cross[i] == 1 || cross[i] == 2) rot_div_walk(i, out);
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
    for (int i = 0; i < total; i++) {
        if (cross[i] == 0 || cross[i] == 1) ref_sub(i, arot, adiv, lvl_amr, rg_rot, rg_div);
        if (son[i] >= 0) coarse_sub(i, arot, adiv, lvl_amr, cg_rot, cg_div);
    }
}

void Tree::gridmodify1(int c) {                     // mesh.F90:305-353
    if (son[c] >= 0) for (int k = 0; k < 4; k++) gridmodify1(son[c] + k);
    if (cross[c] != 0) return;
    int jp = up(c), jm = down(c), im = left(c), ip = right(c);
    bool upS = jp >= 0 && son[jp] >= 0, dnS = jm >= 0 && son[jm] >= 0;
    bool lfS = im >= 0 && son[im] >= 0, rtS = ip >= 0 && son[ip] >= 0;
    if (son[c] < 0) {
        if ((lfS && rtS) || (upS && dnS)) refine(c);
    } else {
        int s = son[c];
        if (son[s] < 0 && son[s + 1] < 0 && son[s + 2] < 0 && son[s + 3] < 0)
            if ((!lfS && !rtS) || (!upS && !dnS)) coarse(c);
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
    for (int k = 1; k <= lvl_amr; k++) {
        for (int i = 0; i < total; i++) gridmodify1(i);
        for (int i = 0; i < total; i++) gridmodify4(i);
    }
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
