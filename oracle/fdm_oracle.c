/*
 * fdm_oracle.c -- TEST INFRASTRUCTURE ONLY (see fdm_oracle.h).
 *
 * Plain-C FP64 restatement of the reference solver, structured like the reference: a
 * pointer quadtree, per-leaf recursive stencil gathers, WLS-WENO solved afresh by every
 * requester. Each routine cites the reference file:line it follows. Build with
 * -ffp-contract=off (oracle/Makefile) so every multiply and add rounds separately, as the
 * reference's x86-64 build does.
 *
 * Parity status of each part (what pins it) is tabulated in tests/test_oracle_*.py; the
 * reference itself does not compile with the image's only Fortran compiler, so no part of
 * this file is checked against reference output.
 */
#include "fdm_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* REAL(4) literals widened to REAL(8) by the reference (exact values). */
static const double WCNS_D1 = 0x1.745d18p-4;   /* 1./11.  nnd_center_RK.F90:331 */
static const double WCNS_D2 = 0x1.d1745ep-1;   /* 10./11. nnd_center_RK.F90:331 */
static const double WCNS_EPS = 0x1.16c2p-133;  /* 1.e-40 (float subnormal) nnd_center_RK.F90:319 */
static const double DIR_EPS = 0x1.b7cdfep-34;  /* 1e-10  XuanGua.F90:40 */
static const double PI_REF = 3.1415926535897932; /* module.F90:9 */

typedef struct Node Node;
struct Node {                       /* type(grid), module.F90:24-35 */
    int num[2];
    int lvl, sort, cross;
    double x, y;
    Node *father, *son1, *son2, *son3, *son4;
    Node *upnei, *downnei, *leftnei, *rightnei;
    Node *upleftnei, *downrightnei, *downleftnei, *uprightnei;
    int sp;
    double U[7], Ut[7], Un[7];
    double rot, div;
    int canon;                      /* canonical id (export numbering) */
};

struct Orc {
    OrcParams p;
    double Xmin, Xmax, Ymin, Ymax, h, x_factor;
    int m, n, total;
    Node* cell;
    Node** cell_NS; Node** cell_BC;
    int nResgrid, nBCgrid;
    double gama, cv;
    double tstep, tt;
    int it;
    double lambda_x, lambda_y;
    int l_max;
    double Res_max[6], Res_rms[6];
    double Srot, Sdiv; int Ne;
    int Neq;
    /* every son block ever allocated (freed at orc_free; the reference leaks them) */
    Node** blocks; int nblocks, cap_blocks;
    int n_canon;
    /* per-leaf residual terms fu*dx*dy of the last stage (cell_NS order), reduced serially in
     * leaf order after the (OpenMP) leaf loop, so the sums do not depend on the thread count */
    double (*fu_leaf)[4];
    int fu_cap;
};

/* ------------------------------------------------------------------------------------ */
/* small helpers                                                                         */
/* ------------------------------------------------------------------------------------ */
static double pow2i(int k) { return (double)(1 << k); } /* 2**k integer power, then REAL */

uint64_t orc_splitmix64(uint64_t* s) {
    uint64_t z = (*s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
static double u01(uint64_t* s) { return (double)(orc_splitmix64(s) >> 11) * 0x1.0p-53; }
static double uab(uint64_t* s, double a, double b) { return a + (b - a) * u01(s); }

/* ------------------------------------------------------------------------------------ */
/* mesh.F90                                                                              */
/* ------------------------------------------------------------------------------------ */
static int gridsort(double xx, double yy) { (void)xx; (void)yy; return 0; } /* mesh.F90:81-87 */

static int gridcross(const Orc* o, const Node* c) {               /* mesh.F90:89-142 */
    int m = o->m, n = o->n;
    if (c->num[0] < 4 && c->num[1] < 4) return -5;
    if (c->num[0] < 4 && c->num[1] > m - 3) return -5;
    if (c->num[0] > n - 3 && c->num[1] > m - 3) return -5;
    if (c->num[0] > n - 3 && c->num[1] < 4) return -5;
    if (c->num[0] < 4) return -1;
    if (c->num[0] > n - 3) return -3;
    if (c->num[1] > m - 3) return -2;
    if (c->num[1] < 4) return -4;
    /* gridsort is identically 0, so t1..t4 = 0 and t0 = 0 */
    return 0;
}

static Node* alloc_sons(Orc* o) {
    Node* b = (Node*)calloc(4, sizeof(Node));
    if (o->nblocks == o->cap_blocks) {
        o->cap_blocks = o->cap_blocks ? 2 * o->cap_blocks : 1024;
        o->blocks = (Node**)realloc(o->blocks, sizeof(Node*) * o->cap_blocks);
    }
    o->blocks[o->nblocks++] = b;
    return b;
}

static void refine_cell(Orc* o, Node* c) {                         /* mesh.F90:199-260 */
    Node* b = alloc_sons(o);  /* re-refining an internal node overwrites its sons (mesh.F90:203) */
    c->son1 = &b[0]; c->son2 = &b[1]; c->son3 = &b[2]; c->son4 = &b[3];
    double ox = o->x_factor * o->h / pow2i(c->lvl + 2);
    double oy = o->h / pow2i(c->lvl + 2);
    Node* s[4] = {c->son1, c->son2, c->son3, c->son4};
    const double sx[4] = {-1, +1, +1, -1}, sy[4] = {-1, -1, +1, +1};
    for (int k = 0; k < 4; k++) {
        Node* q = s[k];
        q->father = c; q->num[0] = c->num[0]; q->num[1] = c->num[1];
        q->lvl = c->lvl + 1; q->sp = k + 1;
        q->x = sx[k] < 0 ? c->x - ox : c->x + ox;
        q->y = sy[k] < 0 ? c->y - oy : c->y + oy;
        q->sort = gridsort(q->x, q->y);
        q->cross = gridcross(o, q);
        memcpy(q->U, c->U, sizeof q->U);
        memcpy(q->Ut, c->U, sizeof q->Ut);
        memcpy(q->Un, c->U, sizeof q->Un);
        q->rot = c->rot; q->div = c->div;
        q->son1 = q->son2 = q->son3 = q->son4 = NULL;
    }
}

/* ------------------------------------------------------------------------------------ */
/* neighbor.F90                                                                          */
/* ------------------------------------------------------------------------------------ */
static Node* init_up(const Orc* o, const Node* c) {                 /* neighbor.F90:6-18 */
    if (c->num[0] == o->n) return NULL;
    return &o->cell[c->num[0] * o->m + c->num[1] - 1];
}
static Node* init_down(const Orc* o, const Node* c) {               /* neighbor.F90:21-33 */
    if (c->num[0] == 1) return NULL;
    return &o->cell[c->num[0] * o->m - 2 * o->m + c->num[1] - 1];
}
static Node* init_left(const Orc* o, const Node* c) {               /* neighbor.F90:36-48 */
    if (c->num[1] == 1) return NULL;
    return &o->cell[c->num[0] * o->m - o->m + c->num[1] - 1 - 1];
}
static Node* init_right(const Orc* o, const Node* c) {              /* neighbor.F90:51-63 */
    if (c->num[1] == o->m) return NULL;
    return &o->cell[c->num[0] * o->m - o->m + c->num[1] + 1 - 1];
}

static Node* up_nb(const Orc* o, Node* c) {                         /* neighbor.F90:65-96 */
    Node* f = c->father;
    if (!f) return init_up(o, c);
    if (c->sp == 1) return f->son4;
    if (c->sp == 2) return f->son3;
    Node* fn = up_nb(o, f);
    if (fn) {
        if (!fn->son1) return fn;
        if (c->sp == 4) return fn->son1;
        if (c->sp == 3) return fn->son2;
    }
    return NULL;
}
static Node* down_nb(const Orc* o, Node* c) {                       /* neighbor.F90:98-129 */
    Node* f = c->father;
    if (!f) return init_down(o, c);
    if (c->sp == 4) return f->son1;
    if (c->sp == 3) return f->son2;
    Node* fn = down_nb(o, f);
    if (fn) {
        if (!fn->son1) return fn;
        if (c->sp == 1) return fn->son4;
        if (c->sp == 2) return fn->son3;
    }
    return NULL;
}
static Node* left_nb(const Orc* o, Node* c) {                       /* neighbor.F90:131-162 */
    Node* f = c->father;
    if (!f) return init_left(o, c);
    if (c->sp == 2) return f->son1;
    if (c->sp == 3) return f->son4;
    Node* fn = left_nb(o, f);
    if (fn) {
        if (!fn->son1) return fn;
        if (c->sp == 4) return fn->son3;
        if (c->sp == 1) return fn->son2;
    }
    return NULL;
}
static Node* right_nb(const Orc* o, Node* c) {                      /* neighbor.F90:164-196 */
    Node* f = c->father;
    if (!f) return init_right(o, c);
    if (c->sp == 1) return f->son2;
    if (c->sp == 4) return f->son3;
    Node* fn = right_nb(o, f);
    if (fn) {
        if (!fn->son1) return fn;
        if (c->sp == 2) return fn->son1;
        if (c->sp == 3) return fn->son4;
    }
    return NULL;
}

static Node* upleft_nb(const Orc* o, Node* c) {                     /* neighbor.F90:199-246 */
    Node* ln = left_nb(o, c);
    Node* un = up_nb(o, c);
    Node* f = c->father;
    if (!f) return (!ln || !un) ? NULL : up_nb(o, ln);
    if (c->sp == 2) return f->son4;
    if (c->sp == 4) {
        if (!ln || !un) return NULL;
        Node* r = up_nb(o, ln);
        while (r && r->lvl < c->lvl && r->son2) r = r->son2;
        return r;
    }
    if (c->sp == 1) return (!ln || ln->lvl < c->lvl) ? NULL : up_nb(o, ln);
    if (c->sp == 3) return (!un || un->lvl < c->lvl) ? NULL : left_nb(o, un);
    return NULL;
}
static Node* upright_nb(const Orc* o, Node* c) {                    /* neighbor.F90:249-296 */
    Node* rn = right_nb(o, c);
    Node* un = up_nb(o, c);
    Node* f = c->father;
    if (!f) return (!rn || !un) ? NULL : up_nb(o, rn);
    if (c->sp == 1) return f->son3;
    if (c->sp == 3) {
        if (!rn || !un) return NULL;
        Node* r = up_nb(o, rn);
        while (r && r->lvl < c->lvl && r->son1) r = r->son1;
        return r;
    }
    if (c->sp == 2) return (!rn || rn->lvl < c->lvl) ? NULL : up_nb(o, rn);
    if (c->sp == 4) return (!un || un->lvl < c->lvl) ? NULL : right_nb(o, un);
    return NULL;
}
static Node* downright_nb(const Orc* o, Node* c) {                  /* neighbor.F90:299-346 */
    Node* rn = right_nb(o, c);
    Node* dn = down_nb(o, c);
    Node* f = c->father;
    if (!f) return (!rn || !dn) ? NULL : down_nb(o, rn);
    if (c->sp == 4) return f->son2;
    if (c->sp == 2) {
        if (!rn || !dn) return NULL;
        Node* r = down_nb(o, rn);
        while (r && r->lvl < c->lvl && r->son4) r = r->son4;
        return r;
    }
    if (c->sp == 3) return (!rn || rn->lvl < c->lvl) ? NULL : down_nb(o, rn);
    if (c->sp == 1) return (!dn || dn->lvl < c->lvl) ? NULL : right_nb(o, dn);
    return NULL;
}
static Node* downleft_nb(const Orc* o, Node* c) {                   /* neighbor.F90:349-396 */
    Node* ln = left_nb(o, c);
    Node* dn = down_nb(o, c);
    Node* f = c->father;
    if (!f) return (!ln || !dn) ? NULL : down_nb(o, ln);
    if (c->sp == 3) return f->son1;
    if (c->sp == 1) {
        if (!ln || !dn) return NULL;
        Node* r = down_nb(o, ln);
        while (r && r->lvl < c->lvl && r->son3) r = r->son3;
        return r;
    }
    if (c->sp == 4) return (!ln || ln->lvl < c->lvl) ? NULL : down_nb(o, ln);
    if (c->sp == 2) return (!dn || dn->lvl < c->lvl) ? NULL : left_nb(o, dn);
    return NULL;
}

static void idnei(const Orc* o, Node* c) {                          /* neighbor.F90:412-449 */
    c->upnei = up_nb(o, c);
    c->downnei = down_nb(o, c);
    c->rightnei = right_nb(o, c);
    c->leftnei = left_nb(o, c);
    if (c->cross >= 0) {
        c->upleftnei = upleft_nb(o, c);
        c->uprightnei = upright_nb(o, c);
        c->downleftnei = downleft_nb(o, c);
        c->downrightnei = downright_nb(o, c);
    } else {
        c->upleftnei = c->uprightnei = c->downleftnei = c->downrightnei = NULL;
    }
    if (c->son1) { idnei(o, c->son1); idnei(o, c->son2); idnei(o, c->son3); idnei(o, c->son4); }
}
static void idneighbor(Orc* o) {                                    /* neighbor.F90:399-409 */
    for (int i = 0; i < o->total; i++) idnei(o, &o->cell[i]);
}

/* ------------------------------------------------------------------------------------ */
/* XuanGua.F90: restriction, dense inverse, WLS-WENO, stencil gathers                    */
/* ------------------------------------------------------------------------------------ */
static void interpA(const Node* c, double* var) {                   /* XuanGua.F90:359-398 */
    if (!c->son1) { for (int i = 0; i < 7; i++) var[i] = c->U[i]; return; }
    double s[4][7];
    const Node* sons[4] = {c->son1, c->son2, c->son3, c->son4};
    for (int k = 0; k < 4; k++) {
        if (sons[k]->son1) interpA(sons[k], s[k]);
        else for (int i = 0; i < 7; i++) s[k][i] = sons[k]->U[i];
    }
    for (int i = 0; i < 7; i++) var[i] = 0.25 * (((s[0][i] + s[1][i]) + s[2][i]) + s[3][i]);
}

int32_t orc_brinv(int32_t n, const double* a, double* r) {         /* XuanGua.F90:593-660 */
    /* column-major n x n, as Fortran */
#define R(i, j) r[((j) - 1) * n + ((i) - 1)]
    int is[16], js[16];
    for (int k = 0; k < n * n; k++) r[k] = a[k];
    for (int k = 1; k <= n; k++) {
        double d = 0.0;
        for (int i = k; i <= n; i++)
            for (int j = k; j <= n; j++)
                if (fabs(R(i, j)) > d) { d = fabs(R(i, j)); is[k - 1] = i; js[k - 1] = j; }
        if (d + 1.0 == 1.0) return 0;  /* prints and returns the partial matrix */
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

int32_t orc_get_direction(double xx, double yy, double x, double y) { /* XuanGua.F90:33-43 */
    double dx = xx - x, dy = yy - y;
    double angle = (atan2(dy, dx) + DIR_EPS) * 180.0 / PI_REF;
    int fl = (int)floor((angle + 22.5) / 45.0 + 8);
    return fl % 8 + 1;
}

typedef struct { const Node* a[16]; int n; } NList;

static void add_to_list(NList* l, const Node* c) {                  /* XuanGua.F90:11-30 */
    for (int i = 0; i < l->n; i++) if (l->a[i] == c) return;
    l->a[l->n++] = c;
}

static void search_cell_5(const Node* c, double x, double y, NList* l1, NList* l2, int* used) {
    /* XuanGua.F90:152-177; the 4x repeat of the son loop adds nothing after the first pass
     * (every leaf's direction is marked on first visit), so one pass is equivalent. */
    if (!c->son1) {
        int dir = orc_get_direction(c->x, c->y, x, y);
        if (!used[dir - 1]) {
            if (dir == 1 || dir == 5 || dir == 3 || dir == 7) add_to_list(l1, c);
            else add_to_list(l2, c);
            used[dir - 1] = 1;
        }
    } else {
        search_cell_5(c->son1, x, y, l1, l2, used);
        search_cell_5(c->son2, x, y, l1, l2, used);
        search_cell_5(c->son3, x, y, l1, l2, used);
        search_cell_5(c->son4, x, y, l1, l2, used);
    }
}

/* smoothness integrals, XuanGua.F90:309-355. Box geometry terms are passed precomputed. */
typedef struct { double dxb, dyb, x2, y2, x3, y3; } Box;

static Box make_box(double cx, double cy, int lvl, double h, double x_factor) {
    double hx = x_factor * h / ldexp(1.0, lvl + 1);                 /* XuanGua.F90:318-323 */
    double hy = h / ldexp(1.0, lvl + 1);
    double x_min = cx - hx, x_max = cx + hx, y_min = cy - hy, y_max = cy + hy;
    Box b;
    b.dxb = x_max - x_min; b.dyb = y_max - y_min;
    b.x2 = x_max * x_max - x_min * x_min; b.y2 = y_max * y_max - y_min * y_min;
    b.x3 = x_max * x_max * x_max - x_min * x_min * x_min;
    b.y3 = y_max * y_max * y_max - y_min * y_min * y_min;
    return b;
}
static double isq_poly(double k1, double k2, double k3, const Box* b) { /* XuanGua.F90:342-355 */
    double term1 = k1 * k1 * b->dxb * b->dyb;
    double term2 = 2 * k1 * k2 * b->x2 / 2 * b->dyb + 2 * k1 * k3 * b->dxb * b->y2 / 2;
    double term3 = k2 * k2 * b->x3 / 3 * b->dyb + 2 * k2 * k3 * b->x2 / 2 * b->y2 / 2 +
                   k3 * k3 * b->dxb * b->y3 / 3;
    return term1 + term2 + term3;
}
static double smooth2(const double* a, const Box* b) {              /* XuanGua.F90:325-334 */
    double beta = 0.0;
    beta = beta + isq_poly(a[1], 2 * a[3], a[4], b);
    beta = beta + isq_poly(a[2], a[4], 2 * a[5], b);
    beta = beta + (2 * a[3]) * (2 * a[3]) * b->dxb * b->dyb;
    beta = beta + a[4] * a[4] * b->dxb * b->dyb;
    beta = beta + (2 * a[5]) * (2 * a[5]) * b->dxb * b->dyb;
    return beta;
}
static double smooth1(const double* a, const Box* b) {              /* XuanGua.F90:336-338 */
    return (a[1] * a[1] + a[2] * a[2]) * b->dxb * b->dyb;
}

/* WLSI_US_WENO core on explicit coordinates/values: XuanGua.F90:179-307.
 * px,py,val: m points (first m1 = list 1); var4 receives rho,u,v,p. */
static void wls_core(int m, int m1, const double* px, const double* py, const double (*val)[4],
                     double xx, double yy, const Box* box, double* var4) {
    int m2 = m - m1;
    double w[16], P[16][6];
    for (int i = 0; i < m; i++) {
        double d = sqrt((px[i] - xx) * (px[i] - xx) + (py[i] - yy) * (py[i] - yy));
        w[i] = 1.0 / (d + 1e-6);
    }
    for (int i = 0; i < m; i++) {
        P[i][0] = 1.0; P[i][1] = px[i]; P[i][2] = py[i];
        P[i][3] = px[i] * py[i]; P[i][4] = px[i] * px[i]; P[i][5] = py[i] * py[i];
    }
    /* PTWP = P^T (W P); (W P)(i,c) = w_i P(i,c) exactly (diagonal matmul) */
    double A6[36], A3a[9], A3b[9], I6[36], I3a[9], I3b[9];
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
    orc_brinv(6, A6, I6);
    orc_brinv(3, A3a, I3a);
    orc_brinv(3, A3b, I3b);
    const double lw[3] = {0.98, 0.01, 0.01};
    for (int k = 0; k < 4; k++) {
        double wphi[16];
        for (int i = 0; i < m; i++) wphi[i] = w[i] * val[i][k];
        double b6[6], al[6], b3a[3], b3b[3], as1[3], as2[3];
        for (int r = 0; r < 6; r++) {
            double s = 0.0;
            for (int i = 0; i < m; i++) s = s + P[i][r] * wphi[i];
            b6[r] = s;
        }
        for (int r = 0; r < 6; r++) {
            double s = 0.0;
            for (int c = 0; c < 6; c++) s = s + I6[c * 6 + r] * b6[c];
            al[r] = s;
        }
        double vl = al[0] + xx * al[1] + yy * al[2] + xx * yy * al[3] + xx * xx * al[4] + yy * yy * al[5];
        double sl = smooth2(al, box);
        for (int r = 0; r < 3; r++) {
            double s = 0.0;
            for (int i = 0; i < m1; i++) s = s + P[i][r] * wphi[i];
            b3a[r] = s;
            s = 0.0;
            for (int i = m1; i < m; i++) s = s + P[i][r] * wphi[i];
            b3b[r] = s;
        }
        (void)m2;
        for (int r = 0; r < 3; r++) {
            double s = 0.0, t = 0.0;
            for (int c = 0; c < 3; c++) { s = s + I3a[c * 3 + r] * b3a[c]; t = t + I3b[c * 3 + r] * b3b[c]; }
            as1[r] = s; as2[r] = t;
        }
        double vs1 = as1[0] + xx * as1[1] + yy * as1[2];
        double vs2 = as2[0] + xx * as2[1] + yy * as2[2];
        double beta[3] = {sl, smooth1(as1, box), smooth1(as2, box)};
        double tau = (fabs(beta[0] - beta[1]) + fabs(beta[0] - beta[2])) / 2.0;
        tau = tau * tau;
        double alpha[3], om[3];
        for (int j = 0; j < 3; j++) alpha[j] = lw[j] * (1.0 + tau / (1e-6 + beta[j]));
        double sa = alpha[0] + alpha[1] + alpha[2];
        for (int j = 0; j < 3; j++) om[j] = alpha[j] / sa;
        var4[k] = om[0] * (1.0 / lw[0] * vl - (lw[1] / lw[0] * vs1 + lw[2] / lw[0] * vs2)) +
                  (om[1] * vs1 + om[2] * vs2);
    }
}

/* var(1:7) from a WLS call: k=1,2,3,5 set; var(4) left undefined by the reference -> 0 */
static void wls_nodes(const Orc* o, const Node* c, double xx, double yy, const NList* l1,
                      const NList* l2, double* var) {               /* XuanGua.F90:179-222 */
    int m1 = l1->n, m = l1->n + l2->n;
    double px[16], py[16], val[16][4];
    for (int i = 0; i < m; i++) {
        const Node* q = i < m1 ? l1->a[i] : l2->a[i - m1];
        px[i] = q->x; py[i] = q->y;
        val[i][0] = q->U[0]; val[i][1] = q->U[1]; val[i][2] = q->U[2]; val[i][3] = q->U[4];
    }
    Box b;
    if (o->p.wls_local) {
        /* new frame: (x - xc)/hx, (y - yc)/hy about the WLS cell; its box is [-1/2, 1/2]^2 */
        double hx = o->x_factor * o->h / pow2i(c->lvl), hy = o->h / pow2i(c->lvl);
        for (int i = 0; i < m; i++) { px[i] = (px[i] - c->x) / hx; py[i] = (py[i] - c->y) / hy; }
        xx = (xx - c->x) / hx;
        yy = (yy - c->y) / hy;
        b = make_box(0.0, 0.0, 0, 1.0, 1.0);
    } else {
        b = make_box(c->x, c->y, c->lvl, o->h, o->x_factor);
    }
    double v4[4];
    wls_core(m, m1, px, py, (const double(*)[4])val, xx, yy, &b, v4);
    var[0] = v4[0]; var[1] = v4[1]; var[2] = v4[2]; var[3] = 0.0; var[4] = v4[3];
    var[5] = 0.0; var[6] = 0.0;
}

void orc_wls_weno(int32_t m, int32_t m1, const double* pts, double xx, double yy, double cx,
                  double cy, int32_t lvl, double h, double x_factor, double* out4) {
    double px[16], py[16], val[16][4];
    for (int i = 0; i < m; i++) {
        px[i] = pts[i * 6 + 0]; py[i] = pts[i * 6 + 1];
        for (int k = 0; k < 4; k++) val[i][k] = pts[i * 6 + 2 + k];
    }
    Box b = make_box(cx, cy, lvl, h, x_factor);
    wls_core(m, m1, px, py, (const double(*)[4])val, xx, yy, &b, out4);
}

/* gather lists as parent_value_weno / child_value_weno do */
static void parent_lists(const Node* c, NList* l1, NList* l2) {     /* XuanGua.F90:45-86 */
    int used[8] = {0};
    l1->n = l2->n = 0;
    double x = c->x, y = c->y;
    search_cell_5(c->son1, x, y, l1, l2, used);
    search_cell_5(c->son2, x, y, l1, l2, used);
    search_cell_5(c->son3, x, y, l1, l2, used);
    search_cell_5(c->son4, x, y, l1, l2, used);
    if (l1->n + l2->n < 8) {
        const Node* nb[4] = {c->leftnei, c->rightnei, c->downnei, c->upnei};
        for (int k = 0; k < 4; k++) if (nb[k]) search_cell_5(nb[k], x, y, l1, l2, used);
    }
}
static void child_lists(const Node* c, NList* l1, NList* l2) {      /* XuanGua.F90:96-141 */
    int used[8] = {0};
    l1->n = l2->n = 0;
    double x = c->x, y = c->y;
    add_to_list(l1, c);
    const Node* nb[8] = {c->downleftnei, c->leftnei, c->upleftnei, c->downnei,
                         c->upnei, c->downrightnei, c->rightnei, c->uprightnei};
    for (int k = 0; k < 8; k++) if (nb[k]) search_cell_5(nb[k], x, y, l1, l2, used);
}

static void parent_value_weno(const Orc* o, const Node* c, double* var) {
    NList l1, l2;
    parent_lists(c, &l1, &l2);
    if (l1.n + l2.n < 7) interpA(c, var);                           /* XuanGua.F90:87-88 */
    else wls_nodes(o, c, c->x, c->y, &l1, &l2, var);
}
static void child_value_weno(const Orc* o, double xx, double yy, const Node* c, double* var) {
    NList l1, l2;
    child_lists(c, &l1, &l2);
    if (l1.n + l2.n < 7) { for (int i = 0; i < 7; i++) var[i] = c->U[i]; } /* XuanGua.F90:143-144 */
    else wls_nodes(o, c, xx, yy, &l1, &l2, var);
}

/* The eight stencil gathers Tip..Tjmm (XuanGua.F90:401-590), as one routine:
 * k = 0:+1x 1:-1x 2:+1y 3:-1y 4:+2x 5:-2x 6:+2y 7:-2y.
 * Returns the classification (see orc_stencil_info) and fills var when var != NULL. */
static int stencil(const Orc* o, const Node* c, int k, double* var, const Node** node_out,
                   double* pxo, double* pyo, NList* l1o, NList* l2o) {
    const int dir = k & 3, two = k >= 4;
    const Node* cn;
    switch (dir) {
        case 0: cn = c->rightnei; break;
        case 1: cn = c->leftnei; break;
        case 2: cn = c->upnei; break;
        default: cn = c->downnei; break;
    }
    double xx = c->x, yy = c->y;
    /* virtual point: XuanGua.F90:415,438,442,463,486,490,511,534,538,559,582,586 */
    double off = two ? (dir < 2 ? 2.0 * o->x_factor * o->h : 2.0 * o->h) / pow2i(c->lvl)
                     : (dir < 2 ? o->x_factor * o->h : o->h) / pow2i(c->lvl);
    if (dir == 0) xx = c->x + off;
    else if (dir == 1) xx = c->x - off;
    else if (dir == 2) yy = c->y + off;
    else yy = c->y - off;

    const Node* t;    /* node the value comes from */
    int same;         /* 1: t is at the requester's level */
    if (!two) { t = cn; same = (cn->lvl == c->lvl); }
    else if (cn->lvl == c->lvl) {
        const Node* cnn = dir == 0 ? cn->rightnei : dir == 1 ? cn->leftnei : dir == 2 ? cn->upnei : cn->downnei;
        t = cnn; same = (cnn->lvl == cn->lvl);
    } else { t = cn; same = 0; }

    NList l1, l2; l1.n = l2.n = 0;
    int kind;
    if (same) {
        if (!t->son1) {
            kind = 0;
            if (var) for (int i = 0; i < 7; i++) var[i] = t->U[i];
        } else {
            parent_lists(t, &l1, &l2);
            if (l1.n + l2.n < 7) { kind = 2; if (var) interpA(t, var); }
            else { kind = 1; if (var) wls_nodes(o, t, t->x, t->y, &l1, &l2, var); }
            xx = t->x; yy = t->y;
        }
    } else {
        child_lists(t, &l1, &l2);
        if (l1.n + l2.n < 7) { kind = 4; if (var) for (int i = 0; i < 7; i++) var[i] = t->U[i]; }
        else { kind = 3; if (var) wls_nodes(o, t, xx, yy, &l1, &l2, var); }
    }
    if (node_out) *node_out = t;
    if (pxo) *pxo = xx;
    if (pyo) *pyo = yy;
    if (l1o) *l1o = l1;
    if (l2o) *l2o = l2;
    return kind;
}

/* ------------------------------------------------------------------------------------ */
/* nnd_center_RK.F90: WCNS3 flux derivative and the per-leaf Euler substep               */
/* ------------------------------------------------------------------------------------ */
void orc_wcns_z3(const double* u1, const double* u2, const double* u3, const double* R,
                 double* hp) {                                      /* nnd_center_RK.F90:310-353 */
    double uL[4];
    for (int i = 0; i < 4; i++) {
        double IS1 = (-u1[i] + u2[i]) * (-u1[i] + u2[i]);
        double IS2 = (-u2[i] + u3[i]) * (-u2[i] + u3[i]);
        double t3 = u1[i] - 2. * u2[i] + u3[i];
        double IS3 = t3 * t3;
        double p1 = u2[i];
        double p2 = ((-u1[i] + 6.0 * u2[i] + 3.0 * u3[i]) / 8.0 - WCNS_D1 * p1) / WCNS_D2;
        IS1 = IS1 < IS2 ? IS1 : IS2;  /* min(IS1, IS2) */
        IS2 = IS3 + u1[i] * u1[i] + u2[i] * u2[i] + u3[i] * u3[i] - u1[i] * u2[i] - u2[i] * u3[i] - u1[i] * u3[i];
        double tq = 10. * fabs(IS2 - 3. * IS1);
        double tau3 = tq * tq;
        double a1 = WCNS_D1 * (1.0 + tau3 / (WCNS_EPS + IS1));
        double a2 = WCNS_D2 * (1.0 + tau3 / (WCNS_EPS + IS2));
        double w1 = a1 / (a1 + a2), w2 = a2 / (a1 + a2);
        uL[i] = w1 * p1 + w2 * p2;
    }
    /* R column-major 4x4 */
    for (int r = 0; r < 4; r++)
        hp[r] = R[0 * 4 + r] * uL[0] + R[1 * 4 + r] * uL[1] + R[2 * 4 + r] * uL[2] + R[3 * 4 + r] * uL[3];
}

void orc_eigen(int32_t dir, const double* U, double gama, double cv, double* Rm, double* Ri) {
    /* REcalc (dir 0, nnd_center_RK.F90:114-169) / RFcalc (dir 1, :236-291); column-major */
#define RR(i, j) Rm[((j) - 1) * 4 + (i) - 1]
#define RI(i, j) Ri[((j) - 1) * 4 + (i) - 1]
    double r = U[0], u = U[1] / U[0], v = U[2] / U[0];
    (void)cv; /* U0(4) (temperature) is computed but unused */
    double p = (gama - 1.0) * r * (U[3] / r - 0.5 * (u * u + v * v));
    double a = sqrt(gama * p / r);
    double q2 = 0.5 * (u * u + v * v);
    double g1 = gama - 1.0, mg = 1.0 - gama;
    if (dir == 0) {
        RR(1, 1) = 1.0; RR(2, 1) = u - a; RR(3, 1) = v; RR(4, 1) = a * a / (gama - 1) + q2 - u * a;
        RR(1, 2) = 1.0; RR(2, 2) = u; RR(3, 2) = v; RR(4, 2) = q2;
        RR(1, 3) = 1.0; RR(2, 3) = u + a; RR(3, 3) = v; RR(4, 3) = a * a / (gama - 1) + q2 + u * a;
        RR(1, 4) = 0.0; RR(2, 4) = 0.0; RR(3, 4) = -1.0; RR(4, 4) = -v;
        RI(1, 1) = (g1 * q2 + a * u) / (2.0 * a * a);
        RI(2, 1) = (a * a - g1 * q2) / (a * a);
        RI(3, 1) = (g1 * q2 - a * u) / (2.0 * a * a);
        RI(4, 1) = v;
        RI(1, 2) = (mg * u - a) / (2.0 * a * a);
        RI(2, 2) = (g1 * u) / (a * a);
        RI(3, 2) = (mg * u + a) / (2.0 * a * a);
        RI(4, 2) = 0.0;
        RI(1, 3) = (mg * v) / (2.0 * a * a);
        RI(2, 3) = (g1 * v) / (a * a);
        RI(3, 3) = (mg * v) / (2.0 * a * a);
        RI(4, 3) = -1.0;
    } else {
        RR(1, 1) = 1.0; RR(2, 1) = u; RR(3, 1) = v - a; RR(4, 1) = a * a / (gama - 1) + q2 - v * a;
        RR(1, 2) = 1.0; RR(2, 2) = u; RR(3, 2) = v; RR(4, 2) = q2;
        RR(1, 3) = 1.0; RR(2, 3) = u; RR(3, 3) = v + a; RR(4, 3) = a * a / (gama - 1) + q2 + v * a;
        RR(1, 4) = 0.0; RR(2, 4) = 1.0; RR(3, 4) = 0.0; RR(4, 4) = u;
        RI(1, 1) = (g1 * q2 + a * v) / (2.0 * a * a);
        RI(2, 1) = (a * a - g1 * q2) / (a * a);
        RI(3, 1) = (g1 * q2 - a * v) / (2.0 * a * a);
        RI(4, 1) = -u;
        RI(1, 2) = (mg * u) / (2.0 * a * a);
        RI(2, 2) = (g1 * u) / (a * a);
        RI(3, 2) = (mg * u) / (2.0 * a * a);
        RI(4, 2) = 1.0;
        RI(1, 3) = (mg * v - a) / (2.0 * a * a);
        RI(2, 3) = (g1 * v) / (a * a);
        RI(3, 3) = (mg * v + a) / (2.0 * a * a);
        RI(4, 3) = 0.0;
    }
    RI(1, 4) = g1 / (2.0 * a * a);
    RI(2, 4) = mg / (a * a);
    RI(3, 4) = g1 / (2.0 * a * a);
    RI(4, 4) = 0.0;
#undef RR
#undef RI
}

/* compute_E / compute_F (nnd_center_RK.F90:171-185,293-307): prim (rho,u,v,T,p) at 5 points
 * -> conservative Uc[pt][4] and flux Fl[pt][4] */
static void cons_flux(int dir, const double* prim5x5, double gama, double Uc[5][4], double Fl[5][4]) {
    for (int j = 0; j < 5; j++) {
        const double* q = prim5x5 + 5 * j;
        double r = q[0], u = q[1], v = q[2], p = q[4];
        Uc[j][0] = r;
        Uc[j][1] = r * u;
        Uc[j][2] = r * v;
        Uc[j][3] = p / (gama - 1.0) + 0.5 * r * (u * u + v * v);
        if (dir == 0) {
            Fl[j][0] = r * u;
            Fl[j][1] = r * u * u + p;
            Fl[j][2] = r * u * v;
            Fl[j][3] = u * (p / (gama - 1.0) + 0.5 * r * (u * u + v * v) + p);
        } else {
            Fl[j][0] = r * v;
            Fl[j][1] = r * u * v;
            Fl[j][2] = r * v * v + p;
            Fl[j][3] = v * (p / (gama - 1.0) + 0.5 * r * (u * u + v * v) + p);
        }
    }
}

void orc_flux_derivative(int32_t dir, const double* prim5x5, double d, double lam, double gama,
                         double cv, double* dEF) {   /* nnd_center_RK.F90:66-112 / 187-234 */
    double Uc[5][4], Fl[5][4];
    cons_flux(dir, prim5x5, gama, Uc, Fl);
    double flux[2][4];
    for (int f = 0; f < 2; f++) {   /* f=0: face i-1/2 (points 1..4), f=1: i+1/2 (2..5) */
        double avg[4], Rm[16], Ri[16];
        for (int k = 0; k < 4; k++) avg[k] = (Uc[f + 1][k] + Uc[f + 2][k]) / 2.0;
        orc_eigen(dir, avg, gama, cv, Rm, Ri);
        double w[5][4], g[5][4];
        for (int j = 0; j < 5; j++)
            for (int r = 0; r < 4; r++) {
                double s = 0.0, t = 0.0;
                for (int c = 0; c < 4; c++) { s = s + Ri[c * 4 + r] * Uc[j][c]; t = t + Ri[c * 4 + r] * Fl[j][c]; }
                w[j][r] = s; g[j][r] = t;
            }
        double wL[4], wR[4], gL[4], gR[4];
        int a = f, b = f + 1, c = f + 2, e = f + 3;
        orc_wcns_z3(w[a], w[b], w[c], Rm, wL);
        orc_wcns_z3(w[e], w[c], w[b], Rm, wR);
        orc_wcns_z3(g[a], g[b], g[c], Rm, gL);
        orc_wcns_z3(g[e], g[c], g[b], Rm, gR);
        for (int k = 0; k < 4; k++) flux[f][k] = 0.5 * (gL[k] + gR[k] - lam * (wR[k] - wL[k]));
    }
    for (int n = 0; n < 4; n++) {
        if (dir == 0)
            dEF[n] = 4.0 / (3.0 * d) * (flux[1][n] - flux[0][n]) - 1.0 / (6.0 * d) * (Fl[3][n] - Fl[1][n]);
        else
            dEF[n] = 4.0 / 3.0 / d * (flux[1][n] - flux[0][n]) - 1.0 / 6.0 / d * (Fl[3][n] - Fl[1][n]);
    }
}

/* solution_cell + LXL + Res_LXL (solution.F90:302-312, nnd_center_RK.F90:15-64). The leaf's
 * residual terms fu*dx*dy (nnd_center_RK.F90:31-36) go to res4; the caller folds them into
 * Res_max / Res_rms in leaf order. */
static void solution_cell(const Orc* o, Node* c, double* res4) {
    double v[8][7];
    for (int k = 0; k < 8; k++) stencil(o, c, k, v[k], NULL, NULL, NULL, NULL, NULL);
    double dy = o->h / pow2i(c->lvl);
    double dx = o->x_factor * dy;
    double U[4], fu[4], Ut[4], dE[4], dF[4];
    U[0] = c->U[0];
    U[1] = c->U[0] * c->U[1];
    U[2] = c->U[0] * c->U[2];
    U[3] = c->U[4] / (o->gama - 1.0) + 0.5 * c->U[0] * (c->U[1] * c->U[1] + c->U[2] * c->U[2]);
    double colx[25], coly[25];
    const double* px[5] = {v[5], v[1], c->U, v[0], v[4]};  /* imm, im, c, ip, ipp */
    const double* py[5] = {v[7], v[3], c->U, v[2], v[6]};
    for (int j = 0; j < 5; j++)
        for (int q = 0; q < 5; q++) { colx[5 * j + q] = px[j][q]; coly[5 * j + q] = py[j][q]; }
    orc_flux_derivative(0, colx, dx, o->lambda_x, o->gama, o->cv, dE);
    orc_flux_derivative(1, coly, dy, o->lambda_y, o->gama, o->cv, dF);
    for (int i = 0; i < 4; i++) fu[i] = -dE[i] - dF[i];
    for (int i = 0; i < 4; i++) {
        Ut[i] = U[i] + fu[i] * o->tstep;
        fu[i] = fu[i] * dx * dy;
        res4[i] = fu[i];
    }
    c->Ut[0] = Ut[0];
    c->Ut[1] = Ut[1] / Ut[0];
    c->Ut[2] = Ut[2] / Ut[0];
    c->Ut[3] = (Ut[3] - 0.5 * Ut[0] * (c->Ut[1] * c->Ut[1] + c->Ut[2] * c->Ut[2])) / (o->cv * Ut[0]);
    c->Ut[4] = (o->gama - 1.0) * c->Ut[0] * (Ut[3] / c->Ut[0] - 0.5 * (c->Ut[1] * c->Ut[1] + c->Ut[2] * c->Ut[2]));
}

/* LXL for interior leaves [i0, i1): the OpenMP loop of Time_advance_RK (solution.F90:279-284,
 * PARALLEL DO SCHEDULE(DYNAMIC)) with the residual reduction (the reference's REDUCTION
 * clause) done afterwards in leaf order, so results do not depend on the thread count. */
static void fu_reserve(Orc* o) {
    if (o->fu_cap < o->nResgrid) {
        free(o->fu_leaf);
        o->fu_cap = o->nResgrid + 1;
        o->fu_leaf = (double(*)[4])malloc(sizeof(double[4]) * (size_t)o->fu_cap);
    }
}
static void leaves_lxl(Orc* o, int i0, int i1) {
    fu_reserve(o);
    double(*fu)[4] = o->fu_leaf;
#pragma omp parallel for schedule(dynamic, 16)
    for (int i = i0; i < i1; i++) solution_cell(o, o->cell_NS[i], fu[i]);
    for (int i = i0; i < i1; i++)
        for (int k = 0; k < 4; k++) {
            if (fabs(fu[i][k]) > o->Res_max[k]) o->Res_max[k] = fabs(fu[i][k]);
            o->Res_rms[k] = o->Res_rms[k] + fu[i][k] * fu[i][k];
        }
}

/* ------------------------------------------------------------------------------------ */
/* solution.F90: boundaries, RK3, leaf lists, AMR                                        */
/* ------------------------------------------------------------------------------------ */
static void set_prim(Orc* o, Node* c, double r, double u, double v, double p) {
    for (int i = 0; i < 7; i++) c->U[i] = 0.0;
    c->U[0] = r; c->U[1] = u; c->U[2] = v; c->U[4] = p;
    c->U[3] = (c->U[4] * o->gama * o->p.Ma * o->p.Ma) / c->U[0];   /* solution.F90:55 */
}

/* partner base cell of a ghost (BC_cyc index arithmetic, solution.F90:116-136, and the
 * clamped interior cell for the new transmissive kind) */
static Node* bc_partner(Orc* o, Node* c, int a, int kind) {
    int m = o->m, n = o->n, i;
    if (kind == ORC_BC_PERIODIC) {
        if (a == -4) i = c->num[0] * m + c->num[1] - 6;
        else if (a == -2) i = c->num[0] * m + c->num[1] - 2 * m + 6;
        else if (a == -3) i = c->num[0] * m + c->num[1] - (n - 5) * m;
        else i = c->num[0] * m + c->num[1] - (7 - n) * m;
        return &o->cell[i - 1];
    }
    int ni = c->num[0], nj = c->num[1];
    if (a == -4) nj = 4;
    else if (a == -2) nj = m - 3;
    else if (a == -3) ni = n - 3;
    else ni = 4;
    return &o->cell[(ni - 1) * m + nj - 1];
}

static void bc_fill(Orc* o, Node* c, int a) {                       /* solution.F90:109-168 */
    int kind = a == -4 ? o->p.bc_xm : a == -2 ? o->p.bc_xp : a == -3 ? o->p.bc_yp : o->p.bc_ym;
    if (kind == ORC_BC_INFLOW) {
        set_prim(o, c, o->p.inflow[0], o->p.inflow[1], o->p.inflow[2], o->p.inflow[3]);
        return;
    }
    if (kind != ORC_BC_PERIODIC && kind != ORC_BC_TRANSMISSIVE) return;
    Node* cn = bc_partner(o, c, a, kind);
    if (c->lvl != 0) {
        int sp[32], cl = c->lvl;
        Node* cf = c;
        for (int j = 0; j < cl; j++) { sp[j] = cf->sp; cf = cf->father; }
        /* the reference descends using sp(1) (the leaf's own position) first
         * (solution.F90:148-159); the new transmissive kind descends root-first */
        for (int j = 0; j < cl; j++) {
            if (!cn->son1) break;
            int s = kind == ORC_BC_PERIODIC ? sp[j] : sp[cl - 1 - j];
            cn = s == 1 ? cn->son1 : s == 2 ? cn->son2 : s == 3 ? cn->son3 : cn->son4;
        }
    }
    double v[7];
    interpA(cn, v);
    for (int i = 0; i < 7; i++) c->U[i] = v[i];
}

static void body_boundary(Orc* o, Node* c) {                        /* solution.F90:314-356 */
    if (c->sort != 0) return;
    if (c->cross == -4 || c->cross == -2 || c->cross == -3 || c->cross == -1) bc_fill(o, c, c->cross);
}

static void initial_boundary(Orc* o, Node* c) {                     /* solution.F90:71-107 */
    if (c->son1) {
        initial_boundary(o, c->son1); initial_boundary(o, c->son2);
        initial_boundary(o, c->son3); initial_boundary(o, c->son4);
    } else body_boundary(o, c);
}

/* ---- initial conditions ---- */
typedef struct { double x, y, r, ratio; } Bubble;
typedef struct {
    double q[4][4];      /* Riemann quadrant states rho,u,v,p (quadrants 1..4) */
    double post[4];      /* shock: post-shock state */
    Bubble b[16]; int nb;
} ICData;

static void ic_prepare(const Orc* o, ICData* d) {
    uint64_t s = o->p.seed;
    const double* a = o->p.ic_param;
    memset(d, 0, sizeof *d);
    if (o->p.ic_kind == ORC_IC_RIEMANN4) {
        for (int qd = 0; qd < 4; qd++) {
            d->q[qd][0] = uab(&s, 0.5, 2.0);
            d->q[qd][1] = uab(&s, -0.8, 0.8);
            d->q[qd][2] = uab(&s, -0.8, 0.8);
            d->q[qd][3] = uab(&s, 0.1, 1.5);
        }
    } else if (o->p.ic_kind == ORC_IC_SHOCK_BUBBLES) {
        double g = o->gama, M = a[1];
        double p2 = 1.0 + 2.0 * g / (g + 1.0) * (M * M - 1.0);
        double r2 = (g + 1.0) * M * M / ((g - 1.0) * M * M + 2.0);
        double sh = M * sqrt(g);
        d->post[0] = r2; d->post[1] = sh * (1.0 - 1.0 / r2); d->post[2] = 0.0; d->post[3] = p2;
        int nb = (int)a[2];
        double xs = a[0];
        for (int k = 0; k < nb && k < 16; k++) {
            for (int tries = 0; tries < 100000; tries++) {
                double r = uab(&s, a[3], a[4]);
                double ratio = uab(&s, a[5], a[6]);
                double x = uab(&s, xs + r + 0.05, o->p.xmax - r);
                double y = uab(&s, o->p.ymin + r, o->p.ymax - r);
                int ok = 1;
                for (int j = 0; j < d->nb; j++) {
                    double dx = x - d->b[j].x, dy = y - d->b[j].y;
                    if (sqrt(dx * dx + dy * dy) < r + d->b[j].r) { ok = 0; break; }
                }
                if (ok) { d->b[d->nb].x = x; d->b[d->nb].y = y; d->b[d->nb].r = r; d->b[d->nb].ratio = ratio; d->nb++; break; }
            }
        }
    }
}

static void ic_point(Orc* o, const ICData* d, Node* c) {
    const double* a = o->p.ic_param;
    double x = c->x, y = c->y;
    switch (o->p.ic_kind) {
        case ORC_IC_VORTEX: {                                       /* solution.F90:31-69 */
            const double cc = 5.0;
            double rr = (x - 5.0) * (x - 5.0) + (y - 5.0) * (y - 5.0);
            double dT = -(o->gama - 1.0) * (cc * cc) * exp(1.0 - rr) / (8.0 * o->gama * PI_REF * PI_REF);
            for (int i = 0; i < 7; i++) c->U[i] = 0.0;
            c->U[0] = 1.0 + pow(1.0 + dT, 1.0 / (o->gama - 1.0)) - 1.0;
            c->U[1] = 1.0 - (y - 5.0) * cc / (2.0 * PI_REF) * exp(0.5 * (1.0 - rr));
            c->U[2] = 1.0 + (x - 5.0) * cc / (2.0 * PI_REF) * exp(0.5 * (1.0 - rr));
            c->U[4] = 1.0 + pow(1.0 + dT, o->gama / (o->gama - 1.0)) - 1.0;
            c->U[3] = (c->U[4] * o->gama * o->p.Ma * o->p.Ma) / c->U[0];
            break;
        }
        case ORC_IC_RIEMANN4: {
            int qd = (x >= a[0]) ? (y >= a[1] ? 0 : 3) : (y >= a[1] ? 1 : 2);
            set_prim(o, c, d->q[qd][0], d->q[qd][1], d->q[qd][2], d->q[qd][3]);
            break;
        }
        case ORC_IC_SHOCK_BUBBLES: {
            if (x < a[0]) set_prim(o, c, d->post[0], d->post[1], d->post[2], d->post[3]);
            else {
                double r = 1.0;
                for (int j = 0; j < d->nb; j++) {
                    double dx = x - d->b[j].x, dy = y - d->b[j].y;
                    if (dx * dx + dy * dy < d->b[j].r * d->b[j].r) r = d->b[j].ratio;
                }
                set_prim(o, c, r, 0.0, 0.0, 1.0);
            }
            break;
        }
        default: {  /* density wave */
            set_prim(o, c, 1.0 + 0.2 * sin(PI_REF * (x + y)), 1.0, 1.0, 1.0);
            break;
        }
    }
    for (int i = 0; i < 7; i++) { c->Ut[i] = c->U[i]; c->Un[i] = c->U[i]; }
    c->rot = 0.0; c->div = 0.0;
}

static void ic_tree(Orc* o, const ICData* d, Node* c) {
    if (c->son1) { ic_tree(o, d, c->son1); ic_tree(o, d, c->son2); ic_tree(o, d, c->son3); ic_tree(o, d, c->son4); }
    else ic_point(o, d, c);
}

/* ---- leaf lists (solution.F90:593-693) ---- */
static void ns_walk(Orc* o, Node* c, int fill) {
    if (c->son1) {
        ns_walk(o, c->son1, fill); ns_walk(o, c->son2, fill);
        ns_walk(o, c->son4, fill); ns_walk(o, c->son3, fill);
    } else if (c->sort == 0) {
        if (c->cross < 0) return;
        if (fill) o->cell_NS[o->nResgrid] = c;
        o->nResgrid++;
    }
}
static void bc_walk(Orc* o, Node* c, int fill) {
    if (c->son1) {
        bc_walk(o, c->son1, fill); bc_walk(o, c->son2, fill);
        bc_walk(o, c->son3, fill); bc_walk(o, c->son4, fill);
    } else if (c->sort == 1 || c->sort == 0) {
        if (c->cross == -5) return;
        if (c->sort == 1 || c->cross < 0) {
            if (fill) o->cell_BC[o->nBCgrid] = c;
            o->nBCgrid++;
        }
    }
}
static void IDgrid(Orc* o) {
    free(o->cell_NS); free(o->cell_BC);
    o->nResgrid = o->nBCgrid = 0;
    for (int i = 0; i < o->total; i++) { ns_walk(o, &o->cell[i], 0); bc_walk(o, &o->cell[i], 0); }
    o->cell_NS = (Node**)malloc(sizeof(Node*) * (o->nResgrid + 1));
    o->cell_BC = (Node**)malloc(sizeof(Node*) * (o->nBCgrid + 1));
    o->nResgrid = o->nBCgrid = 0;
    for (int i = 0; i < o->total; i++) { ns_walk(o, &o->cell[i], 1); bc_walk(o, &o->cell[i], 1); }
}

/* ---- AMR (solution.F90:408-589, mesh.F90:305-408) ---- */
static void sensors_leaf(Orc* o, Node* c) {                         /* solution.F90:442-466 */
    double dy = o->h / pow2i(c->lvl);
    double dx = o->x_factor * dy;
    double vp[7], vm[7], wp[7], wm[7];
    stencil(o, c, 0, vp, NULL, NULL, NULL, NULL, NULL);
    stencil(o, c, 1, vm, NULL, NULL, NULL, NULL, NULL);
    stencil(o, c, 2, wp, NULL, NULL, NULL, NULL, NULL);
    stencil(o, c, 3, wm, NULL, NULL, NULL, NULL, NULL);
    double vx = (vp[2] - vm[2]) / (2.0 * dx), uy = (wp[1] - wm[1]) / (2.0 * dy);
    c->rot = fabs(vx - uy) * pow(dy, 1.5);                          /* solution.F90:558-567 */
    double ux = (vp[0] - vm[0]) / (2.0 * dx), vy = (wp[0] - wm[0]) / (2.0 * dy);
    c->div = (fabs(ux) + fabs(vy)) * pow(dy, 1.5);                  /* solution.F90:569-580 */
    if (o->p.amr_sensor == 1) c->rot = c->div;  /* density-gradient only: both tests on div */
    o->Srot = o->Srot + c->rot * c->rot;
    o->Sdiv = o->Sdiv + c->div * c->div;
    o->Ne = o->Ne + 1;
}
static void solute_div_rot(Orc* o, Node* c) {
    if (!c->son1 && (c->cross == 0 || c->cross == 1)) sensors_leaf(o, c);
    else if (c->son1) {
        solute_div_rot(o, c->son1); solute_div_rot(o, c->son2);
        solute_div_rot(o, c->son3); solute_div_rot(o, c->son4);
    }
}
static void rot_div(Orc* o) {                                       /* solution.F90:424-440 */
    o->Srot = 0; o->Sdiv = 0; o->Ne = 0;
    for (int i = 0; i < o->total; i++) {
        Node* c = &o->cell[i];
        if (c->cross == 0 || c->cross == 1 || c->cross == 2) solute_div_rot(o, c);
    }
}

static void coarse_cell(Node* c) {                                  /* mesh.F90:397-408 */
    double v[7];
    interpA(c, v);
    for (int i = 0; i < 7; i++) { c->U[i] = v[i]; c->Un[i] = v[i]; }
    c->son1 = c->son2 = c->son3 = c->son4 = NULL;  /* storage kept until orc_free */
}

static void coarse_sub(Orc* o, Node* c, double arot, double adiv) { /* solution.F90:490-517 */
    if (c->son1 && c->lvl < o->p.lvl_amr && !c->son1->son1 && !c->son2->son1 && !c->son3->son1 && !c->son4->son1) {
        if ((c->son1->rot + c->son2->rot + c->son3->rot + c->son4->rot) / 4 < o->p.coarse_goal_rot * arot &&
            (c->son1->div + c->son2->div + c->son3->div + c->son4->div) / 4 < o->p.coarse_goal_div * adiv)
            coarse_cell(c);
    } else if (c->son1) {
        if (c->son1->son1) coarse_sub(o, c->son1, arot, adiv);
        if (c->son2->son1) coarse_sub(o, c->son2, arot, adiv);
        if (c->son3->son1) coarse_sub(o, c->son3, arot, adiv);
        if (c->son4->son1) coarse_sub(o, c->son4, arot, adiv);
    }
}
static void ref_sub(Orc* o, Node* c, double arot, double adiv) {    /* solution.F90:519-535 */
    if (!c->son1 && c->lvl < o->p.lvl_amr) {
        if (c->rot > o->p.refine_goal_rot * arot || c->div > o->p.refine_goal_div * adiv) refine_cell(o, c);
    } else if (c->son1 && c->lvl < o->p.lvl_amr) {
        ref_sub(o, c->son1, arot, adiv); ref_sub(o, c->son2, arot, adiv);
        ref_sub(o, c->son3, arot, adiv); ref_sub(o, c->son4, arot, adiv);
    }
}
static void solute_ref(Orc* o) {                                    /* solution.F90:468-488 */
    double arot = sqrt(o->Srot / o->Ne), adiv = sqrt(o->Sdiv / o->Ne);
    for (int i = 0; i < o->total; i++) {
        Node* c = &o->cell[i];
        if (c->cross == 0 || c->cross == 1) ref_sub(o, c, arot, adiv);
        if (c->son1) coarse_sub(o, c, arot, adiv);
    }
}

static void gridmodify1(Orc* o, Node* c) {                          /* mesh.F90:305-353 */
    if (c->son1) { gridmodify1(o, c->son1); gridmodify1(o, c->son2); gridmodify1(o, c->son3); gridmodify1(o, c->son4); }
    if (c->cross != 0) return;
    Node *jp = up_nb(o, c), *jm = down_nb(o, c), *im = left_nb(o, c), *ip = right_nb(o, c);
    int up = jp && jp->son1, dn = jm && jm->son1, lf = im && im->son1, rt = ip && ip->son1;
    if (!c->son1) {
        if ((lf && rt) || (up && dn)) refine_cell(o, c);
    } else if (!c->son1->son1 && !c->son2->son1 && !c->son3->son1 && !c->son4->son1) {
        if ((!lf && !rt) || (!up && !dn)) coarse_cell(c);
    }
}
static void gridmodify4(Orc* o, Node* c) {                          /* mesh.F90:355-395 */
    if (c->son1) { gridmodify4(o, c->son1); gridmodify4(o, c->son2); gridmodify4(o, c->son3); gridmodify4(o, c->son4); return; }
    for (int dir = 1; dir <= 4; dir++) {
        Node* nb = dir == 1 ? right_nb(o, c) : dir == 2 ? left_nb(o, c) : dir == 3 ? up_nb(o, c) : down_nb(o, c);
        if (!nb) continue;
        if (nb->lvl <= c->lvl - 2) refine_cell(o, nb);
        Node* nn = dir == 1 ? right_nb(o, nb) : dir == 2 ? left_nb(o, nb) : dir == 3 ? up_nb(o, nb) : down_nb(o, nb);
        if (nn && nn->lvl <= c->lvl - 2) refine_cell(o, nn);
    }
}
static void gridmodify_all(Orc* o) {                                /* solution.F90:537-556 */
    for (int k = 1; k <= o->p.lvl_amr; k++) {
        for (int i = 0; i < o->total; i++) gridmodify1(o, &o->cell[i]);
        idneighbor(o);
        for (int i = 0; i < o->total; i++) gridmodify4(o, &o->cell[i]);
        idneighbor(o);
    }
}
static void amr_regrid(Orc* o) {                                    /* solution.F90:413-419 body */
    rot_div(o);
    solute_ref(o);
    gridmodify_all(o);
    idneighbor(o);
    IDgrid(o);
}

/* ------------------------------------------------------------------------------------ */
/* driver-level API                                                                      */
/* ------------------------------------------------------------------------------------ */
Orc* orc_new(const OrcParams* p) {
    Orc* o = (Orc*)calloc(1, sizeof(Orc));
    o->p = *p;
    o->x_factor = 1.0;
    o->h = p->h;
    o->gama = p->gamma;
    /* MAIN.F90:55-56 ghost extension; MAIN.F90:61 cv */
    o->Xmin = p->xmin - 3.0 * o->x_factor * o->h;
    o->Xmax = p->xmax + 3.0 * o->x_factor * o->h + 1.0e-10;
    o->Ymin = p->ymin - 3.0 * o->h;
    o->Ymax = p->ymax + 3.0 * o->h + 1.0e-10;
    o->cv = 1.0 / (o->gama * (o->gama - 1.0) * p->Ma * p->Ma);
    o->tstep = p->tstep;
    o->Neq = 4;
    return o;
}

void orc_free(Orc* o) {
    if (!o) return;
    for (int i = 0; i < o->nblocks; i++) free(o->blocks[i]);
    free(o->blocks); free(o->cell); free(o->cell_NS); free(o->cell_BC); free(o->fu_leaf);
    free(o);
}

static void creat_mesh(Orc* o) {                                    /* mesh.F90:9-47 */
    o->m = (int)((o->Xmax - o->Xmin) / (o->x_factor * o->h));
    o->n = (int)((o->Ymax - o->Ymin) / o->h);
    o->total = o->m * o->n;
    o->cell = (Node*)calloc(o->total, sizeof(Node));
    for (int i = 1; i <= o->total; i++) {
        int ni = (i - 1) / o->m + 1;
        int nj = i - o->m * (ni - 1);
        Node* c = &o->cell[i - 1];
        c->num[0] = ni; c->num[1] = nj; c->lvl = 0;
        c->x = 0.5 * (2.0 * nj - 1.0) * o->x_factor * o->h + o->Xmin;
        c->y = (ni - 0.5) * o->h + o->Ymin;
        c->sp = 0;
        c->sort = gridsort(c->x, c->y);
        c->father = NULL;
        c->cross = gridcross(o, c);
    }
}

int32_t orc_init(Orc* o) {
    creat_mesh(o);
    idneighbor(o);
    ICData d;
    ic_prepare(o, &d);
    for (int i = 0; i < o->total; i++) ic_tree(o, &d, &o->cell[i]);   /* solution.F90:16-19 */
    for (int i = 0; i < o->total; i++) initial_boundary(o, &o->cell[i]);
    IDgrid(o);
    for (int r = 0; r < o->p.pre_adapt; r++) {
        /* new: adapt the initial condition (sensor + refine + balance), then re-evaluate it */
        amr_regrid(o);
        for (int i = 0; i < o->total; i++) ic_tree(o, &d, &o->cell[i]);
        for (int i = 0; i < o->total; i++) initial_boundary(o, &o->cell[i]);
    }
    return 0;
}

static void time_advance_rk(Orc* o) {                               /* solution.F90:267-298 */
    for (int i = 0; i < o->nResgrid; i++) memcpy(o->cell_NS[i]->Un, o->cell_NS[i]->U, sizeof(double) * 7);
    static const double ra[3] = {0.0, 3.0 / 4.0, 1.0 / 3.0}, rb[3] = {1.0, 1.0 / 4.0, 2.0 / 3.0};
    for (int krk = 0; krk < 3; krk++) {
        for (int i = 0; i < 6; i++) { o->Res_max[i] = 0.0; o->Res_rms[i] = 0.0; }
        leaves_lxl(o, 0, o->nResgrid);
        for (int i = 0; i < o->nResgrid; i++) {                     /* updateRK, solution.F90:386-404 */
            Node* c = o->cell_NS[i];
            for (int j = 0; j < o->Neq + 1; j++) c->U[j] = ra[krk] * c->Un[j] + rb[krk] * c->Ut[j];
        }
        for (int i = 0; i < o->nBCgrid; i++) body_boundary(o, o->cell_BC[i]);
    }
}

int32_t orc_advance_rk3(Orc* o, double dt, double lam_x, double lam_y) {
    o->tstep = dt; o->lambda_x = lam_x; o->lambda_y = lam_y;
    time_advance_rk(o);
    return 0;
}

int32_t orc_step(Orc* o) {                                          /* solution.F90:183-235 */
    o->it = o->it + 1;
    o->l_max = 0;
    for (int i = 0; i < o->nResgrid; i++) {
        Node* c = o->cell_NS[i];
        double a = sqrt(o->gama * c->U[4] / c->U[0]);
        double lx = fabs(c->U[1]) + a, ly = fabs(c->U[2]) + a;
        if (lx > o->lambda_x) o->lambda_x = lx;
        if (ly > o->lambda_y) o->lambda_y = ly;
        if (c->lvl > o->l_max) o->l_max = c->lvl;
    }
    if (o->p.iflag_localdt)
        o->tstep = o->p.cfl / (o->lambda_x / (o->x_factor * o->h / pow2i(o->l_max)) + o->lambda_y / (o->h / pow2i(o->l_max)));
    if (o->tt + o->tstep > o->p.t_end) o->tstep = o->p.t_end - o->tt;
    time_advance_rk(o);
    o->tt = o->tt + o->tstep;
    int bad = 0;
    for (int i = 0; i < o->Neq; i++) {
        o->Res_rms[i] = sqrt(o->Res_rms[i] / (1.0 * o->nResgrid));
        if (isnan(o->Res_rms[i])) bad = 1;
    }
    if (bad) return -1;
    if (o->p.iflag_amr && o->it >= o->p.nstep_adapt && o->it % o->p.pstep_adapt == 1) {
        amr_regrid(o);
        for (int i = 0; i < o->nBCgrid; i++) body_boundary(o, o->cell_BC[i]);
    }
    return 0;
}

/* ---- accessors ---- */
int32_t orc_n_leaf(const Orc* o) { return o->nResgrid; }
int32_t orc_n_bc(const Orc* o) { return o->nBCgrid; }
int32_t orc_base_dims(const Orc* o, int32_t* m, int32_t* n) { *m = o->m; *n = o->n; return 0; }

static int count_nodes(const Node* c) {
    if (!c->son1) return 1;
    return 1 + count_nodes(c->son1) + count_nodes(c->son2) + count_nodes(c->son3) + count_nodes(c->son4);
}
int32_t orc_n_node(const Orc* o) {
    int s = 0;
    for (int i = 0; i < o->total; i++) s += count_nodes(&o->cell[i]);
    return s;
}

/* canonical numbering: base cells first, then for each refined node its 4 sons as a
 * consecutive block, blocks emitted in a pre-order walk of the base cells */
static void canon_walk(Orc* o, Node* c) {
    if (!c->son1) return;
    Node* s[4] = {c->son1, c->son2, c->son3, c->son4};
    for (int k = 0; k < 4; k++) s[k]->canon = o->n_canon++;
    for (int k = 0; k < 4; k++) canon_walk(o, s[k]);
}
static void canon_assign(Orc* o) {
    o->n_canon = o->total;
    for (int i = 0; i < o->total; i++) o->cell[i].canon = i;
    for (int i = 0; i < o->total; i++) canon_walk(o, &o->cell[i]);
}

static void export_node(const Node* c, int32_t* father, int32_t* son, int32_t* lvl, int32_t* sp,
                        int32_t* num2, int32_t* cross, double* x, double* y, double* prim7) {
    int k = c->canon;
    father[k] = c->father ? c->father->canon : -1;
    son[k] = c->son1 ? c->son1->canon : -1;
    lvl[k] = c->lvl; sp[k] = c->sp; num2[2 * k] = c->num[0]; num2[2 * k + 1] = c->num[1];
    cross[k] = c->cross; x[k] = c->x; y[k] = c->y;
    for (int i = 0; i < 7; i++) prim7[7 * k + i] = c->U[i];
    if (c->son1) {
        export_node(c->son1, father, son, lvl, sp, num2, cross, x, y, prim7);
        export_node(c->son2, father, son, lvl, sp, num2, cross, x, y, prim7);
        export_node(c->son3, father, son, lvl, sp, num2, cross, x, y, prim7);
        export_node(c->son4, father, son, lvl, sp, num2, cross, x, y, prim7);
    }
}

int32_t orc_tree_export(const Orc* o, int32_t* father, int32_t* son, int32_t* lvl, int32_t* sp,
                        int32_t* num2, int32_t* cross, double* x, double* y, double* prim7) {
    canon_assign((Orc*)o);
    for (int i = 0; i < o->total; i++) export_node(&o->cell[i], father, son, lvl, sp, num2, cross, x, y, prim7);
    return o->n_canon;
}

int32_t orc_get_leaves(const Orc* o, double* x, double* y, int32_t* lvl, int32_t* node, double* prim5) {
    canon_assign((Orc*)o);
    for (int i = 0; i < o->nResgrid; i++) {
        const Node* c = o->cell_NS[i];
        if (x) x[i] = c->x;
        if (y) y[i] = c->y;
        if (lvl) lvl[i] = c->lvl;
        if (node) node[i] = c->canon;
        if (prim5) for (int k = 0; k < 5; k++) prim5[5 * i + k] = c->U[k];
    }
    return o->nResgrid;
}
int32_t orc_get_bc_leaves(const Orc* o, int32_t* node, double* prim5) {
    canon_assign((Orc*)o);
    for (int i = 0; i < o->nBCgrid; i++) {
        const Node* c = o->cell_BC[i];
        if (node) node[i] = c->canon;
        if (prim5) for (int k = 0; k < 5; k++) prim5[5 * i + k] = c->U[k];
    }
    return o->nBCgrid;
}
int32_t orc_set_leaves(Orc* o, const double* prim5) {
    for (int i = 0; i < o->nResgrid; i++)
        for (int k = 0; k < 5; k++) o->cell_NS[i]->U[k] = prim5[5 * i + k];
    for (int i = 0; i < o->nBCgrid; i++) body_boundary(o, o->cell_BC[i]);
    return 0;
}
int32_t orc_get_scalars(const Orc* o, double* out8) {
    out8[0] = o->tt; out8[1] = o->tstep; out8[2] = o->lambda_x; out8[3] = o->lambda_y;
    out8[4] = o->l_max; out8[5] = o->it; out8[6] = 0; out8[7] = 0;
    return 0;
}
int32_t orc_get_residual(const Orc* o, double* rmax, double* rrms) {
    for (int i = 0; i < 4; i++) { rmax[i] = o->Res_max[i]; rrms[i] = o->Res_rms[i]; }
    return 0;
}

int32_t orc_stencil_info(const Orc* o, int32_t i, int32_t k, int32_t* kind, int32_t* node,
                         double* px, double* py, int32_t* m1, int32_t* m2, int32_t* nb) {
    if (i < 0 || i >= o->nResgrid || k < 0 || k > 7) return -1;
    canon_assign((Orc*)o);
    const Node* t; NList l1, l2;
    *kind = stencil(o, o->cell_NS[i], k, NULL, &t, px, py, &l1, &l2);
    *node = t->canon;
    *m1 = l1.n; *m2 = l2.n;
    for (int j = 0; j < l1.n; j++) nb[j] = l1.a[j]->canon;
    for (int j = 0; j < l2.n; j++) nb[l1.n + j] = l2.a[j]->canon;
    return 0;
}
int32_t orc_stencil_value(const Orc* o, int32_t i, int32_t k, double* var7) {
    if (i < 0 || i >= o->nResgrid || k < 0 || k > 7) return -1;
    stencil(o, o->cell_NS[i], k, var7, NULL, NULL, NULL, NULL, NULL);
    return 0;
}

/* amr_regrid with the leaves' sensors given (cell_NS order) instead of computed by rot_div:
 * the sums Srot, Sdiv, Ne in rot_div's visiting order (solution.F90:424-466), solute_ref,
 * gridmodify, idneighbor, IDgrid (solution.F90:413-419). amr_sensor 1 tests div only, as
 * sensors_leaf does. The ghosts are not refilled. */
static void sum_given(Orc* o, Node* c) {
    if (!c->son1 && (c->cross == 0 || c->cross == 1)) {
        if (o->p.amr_sensor == 1) c->rot = c->div;
        o->Srot = o->Srot + c->rot * c->rot;
        o->Sdiv = o->Sdiv + c->div * c->div;
        o->Ne = o->Ne + 1;
    } else if (c->son1) {
        sum_given(o, c->son1); sum_given(o, c->son2); sum_given(o, c->son3); sum_given(o, c->son4);
    }
}
int32_t orc_regrid_sensors(Orc* o, const double* rot, const double* div) {
    for (int i = 0; i < o->nResgrid; i++) {
        o->cell_NS[i]->rot = rot[i];
        o->cell_NS[i]->div = div[i];
    }
    o->Srot = 0; o->Sdiv = 0; o->Ne = 0;
    for (int i = 0; i < o->total; i++) {
        Node* c = &o->cell[i];
        if (c->cross == 0 || c->cross == 1 || c->cross == 2) sum_given(o, c);
    }
    solute_ref(o);
    gridmodify_all(o);
    idneighbor(o);
    IDgrid(o);
    return 0;
}

int32_t orc_sensors(Orc* o, double* rot, double* div, double* s2) {
    rot_div(o);
    for (int i = 0; i < o->nResgrid; i++) {
        if (rot) rot[i] = o->cell_NS[i]->rot;
        if (div) div[i] = o->cell_NS[i]->div;
    }
    if (s2) { s2[0] = o->Srot; s2[1] = o->Sdiv; }
    return o->Ne;
}

double orc_smoothness(int32_t order, const double* coeff, double cx, double cy, int32_t lvl,
                      double h, double x_factor) {
    Box b = make_box(cx, cy, lvl, h, x_factor);
    return order == 2 ? smooth2(coeff, &b) : smooth1(coeff, &b);
}

static const Node* find_canon(const Node* c, int id) {
    if (c->canon == id) return c;
    if (!c->son1) return NULL;
    const Node* s[4] = {c->son1, c->son2, c->son3, c->son4};
    for (int k = 0; k < 4; k++) {
        const Node* r = find_canon(s[k], id);
        if (r) return r;
    }
    return NULL;
}

int32_t orc_neighbors(const Orc* o, int32_t node, int32_t* out8) {
    canon_assign((Orc*)o);
    const Node* c = NULL;
    for (int i = 0; i < o->total && !c; i++) c = find_canon(&o->cell[i], node);
    if (!c) return -1;
    const Node* nb[8] = {c->upnei, c->downnei, c->rightnei, c->leftnei,
                         c->upleftnei, c->uprightnei, c->downleftnei, c->downrightnei};
    for (int k = 0; k < 8; k++) out8[k] = nb[k] ? nb[k]->canon : -1;
    return 0;
}

/* one forward-Euler substep (LXL for every interior leaf, nnd_center_RK.F90:15-44) from the
 * current state; writes Ut (rho,u,v,T,p) in cell_NS order and leaves U unchanged */
int32_t orc_substep(Orc* o, double dt, double lam_x, double lam_y, double* prim5) {
    o->tstep = dt; o->lambda_x = lam_x; o->lambda_y = lam_y;
    for (int i = 0; i < 6; i++) { o->Res_max[i] = 0.0; o->Res_rms[i] = 0.0; }
    leaves_lxl(o, 0, o->nResgrid);
    for (int i = 0; i < o->nResgrid; i++)
        for (int k = 0; k < 5; k++) prim5[5 * i + k] = o->cell_NS[i]->Ut[k];
    return 0;
}

/* ------------------------------------------------------------------------------------ */
/* test-harness entry points: hand-built trees, whole-tree state, threads, timing        */
/* ------------------------------------------------------------------------------------ */
int32_t orc_set_threads(int32_t n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}

static Node* node_by_canon(const Orc* o, int32_t id) {
    canon_assign((Orc*)o);
    const Node* c = NULL;
    for (int i = 0; i < o->total && !c; i++) c = find_canon(&o->cell[i], id);
    return (Node*)c;
}

/* Refine_cell (mesh.F90:199-260) on the node with canonical id `node` (sons copy its state).
 * Canonical ids change afterwards; call orc_rebuild before using the tree. */
int32_t orc_refine(Orc* o, int32_t node) {
    Node* c = node_by_canon(o, node);
    if (!c || c->son1) return -1;
    refine_cell(o, c);
    return 0;
}

/* idneighbor + IDgrid after hand edits; with reinit_ic the initial condition is evaluated
 * again on every leaf; then the ghosts are filled (initial_boundary) */
int32_t orc_rebuild(Orc* o, int32_t reinit_ic) {
    idneighbor(o);
    if (reinit_ic) {
        ICData d;
        ic_prepare(o, &d);
        for (int i = 0; i < o->total; i++) ic_tree(o, &d, &o->cell[i]);
    }
    for (int i = 0; i < o->total; i++) initial_boundary(o, &o->cell[i]);
    IDgrid(o);
    return 0;
}

static void set_node_walk(Node* c, const double* prim7) {
    for (int i = 0; i < 7; i++) { c->U[i] = prim7[7 * (size_t)c->canon + i]; c->Ut[i] = c->U[i]; c->Un[i] = c->U[i]; }
    if (c->son1) {
        set_node_walk(c->son1, prim7); set_node_walk(c->son2, prim7);
        set_node_walk(c->son3, prim7); set_node_walk(c->son4, prim7);
    }
}
/* U(1:7) of every node, canonical numbering (the ghosts are not refilled) */
int32_t orc_set_node_values(Orc* o, const double* prim7) {
    canon_assign(o);
    for (int i = 0; i < o->total; i++) set_node_walk(&o->cell[i], prim7);
    return 0;
}

/* Replace the tree by a flattened one in canonical numbering (as orc_tree_export writes it;
 * base cells 0..m*n-1 with this oracle's m, n). idneighbor and IDgrid follow; the ghosts keep
 * the given values. */
int32_t orc_tree_import(Orc* o, int32_t n_node, const int32_t* father, const int32_t* son,
                        const int32_t* lvl, const int32_t* sp, const int32_t* num2, const int32_t* cross,
                        const double* x, const double* y, const double* prim7) {
    if (n_node < o->total) return -1;
    for (int i = 0; i < o->nblocks; i++) free(o->blocks[i]);
    o->nblocks = 0;
    Node** at = (Node**)calloc((size_t)n_node, sizeof(Node*));
    for (int k = 0; k < o->total; k++) at[k] = &o->cell[k];
    for (int k = 0; k < n_node; k++) {
        if (son[k] < 0) continue;
        if (son[k] < o->total || son[k] + 3 >= n_node) { free(at); return -1; }
        Node* b = alloc_sons(o);
        for (int j = 0; j < 4; j++) at[son[k] + j] = &b[j];
    }
    for (int k = 0; k < n_node; k++) {
        Node* c = at[k];
        if (!c) { free(at); return -1; }
        c->num[0] = num2[2 * k]; c->num[1] = num2[2 * k + 1];
        c->lvl = lvl[k]; c->sp = sp[k]; c->cross = cross[k]; c->sort = 0;
        c->x = x[k]; c->y = y[k];
        c->father = father[k] >= 0 ? at[father[k]] : NULL;
        if (son[k] >= 0) {
            c->son1 = at[son[k]]; c->son2 = at[son[k] + 1]; c->son3 = at[son[k] + 2]; c->son4 = at[son[k] + 3];
        } else {
            c->son1 = c->son2 = c->son3 = c->son4 = NULL;
        }
        for (int i = 0; i < 7; i++) { c->U[i] = prim7[7 * (size_t)k + i]; c->Ut[i] = c->U[i]; c->Un[i] = c->U[i]; }
        c->rot = c->div = 0.0;
    }
    free(at);
    idneighbor(o);
    IDgrid(o);
    return 0;
}

/* Time_advance_RK restricted to interior leaves [i0, i1) (no ghost fill): the CPU baseline's
 * timed unit on a slice of a large mesh */
int32_t orc_rk3_range(Orc* o, int32_t i0, int32_t i1, double dt, double lam_x, double lam_y) {
    if (i0 < 0 || i1 > o->nResgrid || i0 > i1) return -1;
    o->tstep = dt; o->lambda_x = lam_x; o->lambda_y = lam_y;
    for (int i = i0; i < i1; i++) memcpy(o->cell_NS[i]->Un, o->cell_NS[i]->U, sizeof(double) * 7);
    static const double ra[3] = {0.0, 3.0 / 4.0, 1.0 / 3.0}, rb[3] = {1.0, 1.0 / 4.0, 2.0 / 3.0};
    for (int krk = 0; krk < 3; krk++) {
        for (int i = 0; i < 6; i++) { o->Res_max[i] = 0.0; o->Res_rms[i] = 0.0; }
        leaves_lxl(o, i0, i1);
        for (int i = i0; i < i1; i++) {
            Node* c = o->cell_NS[i];
            for (int j = 0; j < o->Neq + 1; j++) c->U[j] = ra[krk] * c->Un[j] + rb[krk] * c->Ut[j];
        }
    }
    return 0;
}

/* orc_neighbors for every node: out8[8 * canonical id + k] */
static void nb_walk(const Node* c, int32_t* out) {
    const Node* nb[8] = {c->upnei, c->downnei, c->rightnei, c->leftnei,
                         c->upleftnei, c->uprightnei, c->downleftnei, c->downrightnei};
    for (int k = 0; k < 8; k++) out[8 * (size_t)c->canon + k] = nb[k] ? nb[k]->canon : -1;
    if (c->son1) { nb_walk(c->son1, out); nb_walk(c->son2, out); nb_walk(c->son3, out); nb_walk(c->son4, out); }
}
int32_t orc_neighbors_all(const Orc* o, int32_t* out8) {
    canon_assign((Orc*)o);
    for (int i = 0; i < o->total; i++) nb_walk(&o->cell[i], out8);
    return 0;
}

/* orc_stencil_info and orc_stencil_value for every interior leaf and position: arrays of
 * N*8 entries (nb: N*8*9, var: N*8*7; either may be NULL) */
int32_t orc_stencil_all(const Orc* o, int32_t* kind, int32_t* node, double* px, double* py, int32_t* m1,
                        int32_t* m2, int32_t* nb, double* var) {
    canon_assign((Orc*)o);
#pragma omp parallel for schedule(dynamic, 16)
    for (int i = 0; i < o->nResgrid; i++)
        for (int k = 0; k < 8; k++) {
            const size_t d = (size_t)i * 8 + k;
            const Node* t; NList l1, l2;
            kind[d] = stencil(o, o->cell_NS[i], k, var ? var + 7 * d : NULL, &t, px + d, py + d, &l1, &l2);
            node[d] = t->canon;
            m1[d] = l1.n; m2[d] = l2.n;
            if (nb) {
                for (int j = 0; j < 9; j++) nb[9 * d + j] = -1;
                for (int j = 0; j < l1.n && j < 9; j++) nb[9 * d + j] = l1.a[j]->canon;
                for (int j = 0; j < l2.n && l1.n + j < 9; j++) nb[9 * d + l1.n + j] = l2.a[j]->canon;
            }
        }
    return 0;
}
