// Host regrid and table build under the sanitizers (AddressSanitizer + UndefinedBehavior-
// Sanitizer, or ThreadSanitizer): a shock-bubble mesh with 134 x 70 base cells (enough for the
// parallel passes) regridded from synthetic sensors several times on several host threads,
// each time also on one thread; the trees and tables must agree. Exit code 0 on success.
// Built and run by tests/test_sanitizers.py (CPU only).
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <string>

#include "host_tables.hpp"

using namespace fdm;

static uint64_t fnv(uint64_t h, const void* p, size_t n) {
    const unsigned char* b = static_cast<const unsigned char*>(p);
    for (size_t k = 0; k < n; k++) { h ^= b[k]; h *= 1099511628211ull; }
    return h;
}

static uint64_t digest(const Tree& t, const Tables& tb) {
    uint64_t h = 1469598103934665603ull;
    h = fnv(h, t.father.data(), t.father.size() * 4);
    h = fnv(h, t.son.data(), t.son.size() * 4);
    h = fnv(h, t.x.data(), t.x.size() * 8);
    h = fnv(h, t.U.data(), t.U.size() * 56);
    h = fnv(h, t.nb.data(), t.nb.size() * 4);
    h = fnv(h, tb.stencil.data(), tb.stencil.size() * 4);
    h = fnv(h, tb.wls_i.data(), tb.wls_i.size() * 4);
    h = fnv(h, tb.wls_d.data(), tb.wls_d.size() * 8);
    h = fnv(h, tb.off16.data(), tb.off16.size() * 4);
    h = fnv(h, tb.sel_nbr.data(), tb.sel_nbr.size() * 4);
    h = fnv(h, tb.sel_loc.data(), tb.sel_loc.size());
    h = fnv(h, tb.face_lo.data(), tb.face_lo.size() * 4);
    return h;
}

// the cached neighbours (idneighbor, evaluated top-down) against the recursive finders
static bool neighbours_match_finders(const Tree& t) {
    for (size_t c = 0; c < t.n_alloc(); c++) {
        if (t.father[c] == -2) continue;
        const int ci = (int)c;
        int want[8] = {t.up(ci), t.down(ci), t.right(ci), t.left(ci), -1, -1, -1, -1};
        if (t.cross[c] >= 0) {
            want[NB_UPLEFT] = t.upleft(ci);
            want[NB_UPRIGHT] = t.upright(ci);
            want[NB_DOWNLEFT] = t.downleft(ci);
            want[NB_DOWNRIGHT] = t.downright(ci);
        }
        for (int k = 0; k < 8; k++)
            if (t.nb[8 * c + k] != want[k]) {
                std::printf("node %zu neighbour %d: cached %d, finder %d\n", c, k, t.nb[8 * c + k], want[k]);
                return false;
            }
    }
    return true;
}

static void mesh(Tree& t) {
    t.configure(0, 2, 0, 1, 1.0 / 64, 1.4, 1.0);
    t.bc[0] = 2; t.bc[1] = 1; t.bc[2] = 1; t.bc[3] = 1;
    t.inflow[0] = 2.0; t.inflow[1] = 1.0; t.inflow[2] = 0; t.inflow[3] = 4.0;
    t.wls_local = 1;
    t.creat_mesh();
    const double par[8] = {0.4, 2.0, 3, 0.1, 0.2, 0.1, 5.0, 0};
    t.initial(2, 3, par);
    t.idgrid();
    t.boundary_fill_host();
}

static void regrid(Tree& t, Tables& tb, int k, const char* threads) {
    setenv("FDMAMR_HOST_THREADS", threads, 1);
    for (int c : t.ns) {
        const double x = t.x[c], y = t.y[c];
        double s = 0;
        for (int r = 0; r < 4; r++) {
            const double cx = 0.3 + 0.45 * r + 0.05 * k, cy = 0.5 + 0.25 * std::sin(1.3 * r + k);
            const double d = std::fabs(std::hypot(x - cx, y - cy) - 0.15 - 0.02 * r);
            s += std::exp(-d * d / 2e-3);
        }
        const uint64_t hh = ((uint64_t)(x * 4096) * 73856093u) ^ ((uint64_t)(y * 4096) * 19349663u);
        s *= 0.3 + (double)((hh ^ (hh >> 13)) & 0xFFFF) / 65536.0;
        t.rot[c] = s;
        t.div[c] = 0.7 * s;
    }
    double sr = 0, sd = 0;
    int ne = 0;
    t.rot_div_sums(sr, sd, ne);
    const double ar = std::sqrt(sr / ne), ad = std::sqrt(sd / ne);
    t.solute_ref(ar, ad, 3, 1.0, 1.0, 0.4, 0.4);
    t.solute_gridmodify(3);
    t.idneighbor();
    t.idgrid();
    build_tables(t, tb);
}

int main() {
    Tree a, b;
    Tables ta, tbb;
    mesh(a);
    mesh(b);
    for (int k = 0; k < 4; k++) {
        regrid(a, ta, k, "4");
        regrid(b, tbb, k, "1");
        const uint64_t ha = digest(a, ta), hb = digest(b, tbb);
        std::printf("regrid %d: %zu leaves, %d WLS slots, digest %016llx %s\n", k, a.ns.size(), ta.S_w,
                    (unsigned long long)ha, ha == hb ? "(= one thread)" : "(DIFFERS from one thread)");
        if (ha != hb) return 1;
        if (!neighbours_match_finders(a) || !neighbours_match_finders(b)) return 1;
    }
    return 0;
}
