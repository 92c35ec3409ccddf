// host_tables.hpp -- flattening of the quadtree into the device tables of the hot path.
//
// The reference resolves every stencil value afresh in every stage by walking cached
// neighbour pointers (Tip..Tjmm, XuanGua.F90:401-590) and, at refinement-level jumps, by
// gathering a WLS-WENO stencil and inverting three normal-equation matrices
// (XuanGua.F90:45-356). All of that depends only on the mesh, so it is done once per regrid
// here: each interior leaf gets eight state indices; each distinct level-jump value becomes a
// "slot" whose geometry (weights, inverted matrices, smoothness box) is precomputed in FP64.
#pragma once
#include <cstdint>
#include <vector>

#include "host_tree.hpp"

namespace fdm {

// per-slot record layout (doubles), shared with the device code
enum WlsLayout {
    WLS_INV6 = 0,    // 36: inverse of P^T W P (6x6), column-major
    WLS_INV3A = 36,  // 9: inverse for list 1 (3x3)
    WLS_INV3B = 45,  // 9: inverse for list 2 (3x3)
    WLS_PX = 54,     // 9 stencil x
    WLS_PY = 63,     // 9 stencil y
    WLS_W = 72,      // 9 inverse-distance weights
    WLS_XX = 81, WLS_YY = 82,
    WLS_BOX = 83,    // 6: dxb, dyb, x2, y2, x3, y3
    WLS_STRIDE = 96
};
// WLSI_M: stencil size m in the low byte; bit WLSI_GSHIFT + j set when point j is a ghost leaf
// (read through its fill list by the fast path). WLSI_PAT: index of the slot's records in
// wls_d / wls_f (slots with bit-identical geometry share them).
enum WlsIntLayout { WLSI_M = 0, WLSI_M1 = 1, WLSI_IDS = 2, WLSI_PAT = 11, WLSI_STRIDE = 12 };
constexpr int WLSI_MMASK = 0xFF, WLSI_GSHIFT = 8;
// ghost entries of the slots that read ghost leaves (Tables::dep_hdr): up to 2 ghost points of
// up to DEP_ENT entries each; DEP_INFLOW marks the inflow state, DEP_GENERIC a slot that does
// not fit (the mesh then takes a fill launch per stage)
constexpr int DEP_ENT = 4, DEP_INFLOW = -2, DEP_GENERIC = 255;

// FP32 slot record of the fast path (floats): the least-squares solves folded into
// coefficient matrices on the host in FP64 (C6 = (P^T W P)^-1 P^T W etc.), plus the value
// weights c = basis(xx,yy)^T C, so a slot costs a few small mat-vecs per variable. Only the
// rows the smoothness indicators read are kept (the constant-term rows enter through the
// value weights alone). Every row is indexed by stencil point and zero outside its list
// (list 1: points [0, min(m1, 5)); list 2: points [m1, m1 + min(m - m1, 4))), so the kernel
// needs no data-dependent indexing.
enum WlsFastLayout {
    WF_C6 = 0,      // rows 1..5 of 6 x 9, row-major [(r-1)*9 + i]
    WF_C3A = 45,    // rows 1..2 of 3 x 5 (list 1, points 0..4)
    WF_C3B = 55,    // rows 1..2 of 3 x 9 (list 2, by point)
    WF_CL = 73,     // 9 value weights of the quadratic fit
    WF_CS1 = 82,    // 5 value weights of linear fit 1
    WF_CS2 = 87,    // 9 value weights of linear fit 2 (by point)
    WF_BOX = 96,    // 6 box terms (dxb, dyb, x2, y2, x3, y3)
    WF_STRIDE = 104
};

constexpr int32_t REF_INFLOW = -2147483647 - 1;
constexpr int SEL_NBR_CAP = 192;  // positions fit a byte
// compressed stencil records (Tables::st16)
constexpr int ST16_BLK = 256, ST16_FAR = 16384;
constexpr int16_t ST16_ESC = -32768;
// fast-path slot launch: slots per block (two lanes each) and the most distinct FP32 records a
// block stages in LDS
constexpr int SLOT_BLK = 128, SLOT_BLK_PATS = 48;

// stencil classification, identical to the reference's gathers
enum StKind { ST_COPY = 0, ST_PARENT_WLS = 1, ST_PARENT_RESTRICT = 2, ST_CHILD_WLS = 3, ST_CHILD_COPY = 4 };

struct Tables {
    int N = 0, G = 0, P = 0, L = 0;    // interior, ghost, passive (corner) leaves, L = N+G+P
    int S_w = 0, S_r = 0;              // WLS slots, restriction slots
    int S_wp = 0;                      // WLS slots [0, S_wp) read no ghost leaf (the rest do)
    int l_max = 0;
    BigVec<int32_t> state_of_node;       // node -> state index (leaves) or -1
    BigVec<int32_t> leaf_node;           // state index -> node
    // interior leaves are stored in device order (base cells along a Morton curve); ns_pos[i]
    // is the cell_NS position of device leaf i, dev_of_ns its inverse
    BigVec<int32_t> ns_pos, dev_of_ns;
    BigVec<int32_t> stencil;             // [N][8]: state index, or L + slot index
    // The same records in 16 bytes for the fast path's stage launch: entry e of leaf i is the
    // leaf i + e for |e| < ST16_FAR, or the state st16_base[i / ST16_BLK] + (e - ST16_FAR) for
    // e >= ST16_FAR (ghosts and level-jump slots, numbered in leaf order); a leaf with an entry
    // neither form holds has every entry ST16_ESC and is read from `stencil`.
    BigVec<int16_t> st16;                // [N][8]
    BigVec<int32_t> st16_base;           // [ceil(N / ST16_BLK)]
    int st16_escaped = 0;                // leaves read from `stencil`
    // bit 0: the leaf's x-1/2 face reads the same four state indices as its left neighbour's
    // x+1/2 face (so both compute the same flux, bit for bit); bit 1: the same in y
    BigVec<uint8_t> face_share;          // [N]
    // the leaves whose x-1/2 face is not shared, ascending, then those whose y-1/2 face is not
    BigVec<int32_t> face_lo;
    int n_lo_x = 0, n_lo_y = 0;
    BigVec<int8_t> lvl_leaf;             // [N]
    // diagnostics (host only)
    BigVec<int8_t> st_kind;              // [N*8]
    BigVec<int32_t> st_node;             // [N*8]
    // slots
    std::vector<double> wls_d;           // [n_pat * WLS_STRIDE], one per pattern
    BigVec<int32_t> wls_i;               // [S_w * WLSI_STRIDE]
    std::vector<float> wls_f;            // [n_pat * WF_STRIDE] fast-path records, one per pattern
    int n_pat = 0;
    std::vector<int32_t> rst_node;       // [S_r]
    // ghost fill: partner node per ghost (-1: inflow state)
    std::vector<int32_t> bc_partner;     // [G]
    // Restriction means (interpA, XuanGua.F90:359-398) of the internal nodes that ghosts and
    // restriction slots need, evaluated bottom-up one level per launch. Child / source refs:
    // >= 0 leaf state index, < 0 internal node -(1 + index); REF_INFLOW for the inflow kind.
    int n_int = 0;
    std::vector<int32_t> int_child;      // [n_int][4]
    std::vector<int32_t> int_level_start;  // group boundaries, deepest level first
    std::vector<int32_t> rst_ref;        // [S_r] internal index of each restriction slot
    std::vector<int32_t> bc_ref;         // [G]
    // The same fill flattened for the fast path: item k (ghost k, then restriction slot
    // k - G) is sum_j fill_w[j] * state[fill_leaf[j]] over j in [fill_beg[k], fill_beg[k+1]),
    // the weights 4^-depth of the leaves below the referenced node; fill_leaf = -1 marks the
    // inflow state. One pass instead of one per tree level; the order of the sum differs from
    // the nested means, so the FP64 path keeps those.
    std::vector<int32_t> fill_beg, fill_leaf;
    std::vector<float> fill_w;
    bool fill_flat_ok = true;  // every summed leaf is interior (no ghost read while written)
    // ghost g -> the state index the tables read instead: the interior leaf it copies, or the
    // first inflow ghost for the other inflow ghosts (-1: read the ghost itself). The fast path
    // refills only the items in fill_fast (ghosts without an alias, then restriction slots, as
    // item ids k of the flattened fill).
    std::vector<int32_t> ghost_alias;    // [G]
    std::vector<int32_t> fill_fast;
    // slot S_wp + d: dep_hdr[4d] = number of ghost points (or DEP_GENERIC), [4d+1], [4d+2]
    // their point indices; dep_leaf / dep_w [d][2][DEP_ENT]: leaf (-1: none) and weight of
    // each entry of the ghost's fill list
    std::vector<int32_t> dep_hdr, dep_leaf;
    std::vector<float> dep_w;
    // the distinct FP32 slot records of each block of SLOT_BLK slots (the first SLOT_BLK_PATS in
    // slot order; -1 pads), and per slot its record's position in that list (255: not listed,
    // read from wls_f directly)
    std::vector<int32_t> slot_blk_pat;   // [ceil(S_w / SLOT_BLK)][SLOT_BLK_PATS]
    std::vector<uint8_t> slot_lpat;      // [S_w]
    std::vector<int32_t> slot_uncov;     // the slots with slot_lpat 255
    // the fast path refills ghosts in its slot launch when the flattened fill is usable and
    // every ghost-reading slot fits its fixed-size entry lists (else a fill launch per stage)
    bool merge_ok = true;
    // selector input: state index of the leaf containing each of the 16 off-axis points of
    // the leaf's 5x5 stencil (x + a dx, y + b dy), a,b in -2..2, a != 0, b != 0, row-major
    // in b then a. New rule (the reference has no 5x5 stencil): descend from the base cell
    // into the son whose quadrant contains the point, ties to the west/south son.
    BigVec<int32_t> off16;               // [16][N]
    // Selector tiles (16 consecutive interior leaves): the distinct state indices their 5x5
    // stencils read, tile k's in sel_nbr[sel_nbr_beg[k], sel_nbr_beg[k+1]), and for every
    // leaf the positions of its points in that list: sel_loc[leaf][g][j] (uint8) is point
    // g + 4j of the 5x5 grid (p = (b+2)*5 + (a+2); the centre for p > 24), j < 7, for the four
    // lanes g of the leaf, and sel_loc[leaf][g][7] the centre's position. A tile with more than
    // SEL_NBR_CAP distinct indices stores none (count SEL_NBR_CAP + 1) and is gathered directly.
    std::vector<int32_t> sel_nbr_beg;
    BigVec<int32_t> sel_nbr;
    BigVec<uint8_t> sel_loc;             // [N][4][8]
    // device topology for restriction means
    BigVec<int32_t> node_son;            // [n_alloc]
};

void build_tables(const Tree& t, Tables& tb);
// WLS-WENO stencil lists of one classified entry (diagnostics): node ids, list sizes
int stencil_lists(const Tree& t, int tg, int kind, int32_t* nb9, int32_t* m1, int32_t* m2);
// the point stencil entry k of leaf node c evaluates (kind, target tg as classified): the
// parent's centre for a parent entry, else the point at one or two cell widths (Tip..Tjmm)
void stencil_point(const Tree& t, int c, int k, int kind, int tg, double* px, double* py);

// reference helpers reused by diagnostics
int brinv(int n, const double* a, double* r);                   // XuanGua.F90:593-660
int get_direction(double xx, double yy, double x, double y);    // XuanGua.F90:33-43

}  // namespace fdm
