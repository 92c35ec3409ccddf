// fdmamr_device.hpp -- device-side table views and step scalars shared by the kernels and the
// host-side context.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "host_tables.hpp"

namespace fdm {

// maximum quadtree depth handled by the on-device restriction walk
constexpr int MAX_TREE_DEPTH = 24;

// Workgroups are dispatched round-robin over the 8 XCDs (block b to XCD b % 8). This renumbers
// them so that each XCD processes one contiguous range of logical blocks: neighbouring leaves
// share stencil data, and the range then stays in that XCD's L2. A bijection on [0, gridDim.x).
// xcd_block_n applies the same renumbering to the first nb blocks of a larger grid (the
// active prefix of a launch whose size is only known on the device).
__device__ __forceinline__ int xcd_block_n(int b, int nb) {
    const int q = nb >> 3, r = nb & 7, x = b & 7;
    return x * q + (x < r ? x : r) + (b >> 3);
}
__device__ __forceinline__ int xcd_block() { return xcd_block_n(blockIdx.x, gridDim.x); }

// ---- wave64 reductions on DPP lane moves (VALU, no LDS round trips) ----
// Lanes whose DPP source is outside the pattern or whose row is masked off read 0, the
// identity of a sum and of a maximum of non-negative values; the result is valid in lane 63.
template <int CTRL, int ROWS>
__device__ __forceinline__ double dpp_move(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(b & 0xffffffffll), CTRL, ROWS, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, ROWS, 0xF, false);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}
// quad_perm [1,0,3,2] and [2,3,0,1] sum quads, row_half_mirror and row_mirror the 16-lane
// rows; row_bcast:15 (rows 1, 3) and row_bcast:31 (rows 2, 3) carry the rows into lane 63
template <bool MAX>
__device__ __forceinline__ double wave_reduce63(double v) {
    auto op = [](double a, double b) { return MAX ? fmax(a, b) : a + b; };
    v = op(v, dpp_move<0xB1, 0xF>(v));
    v = op(v, dpp_move<0x4E, 0xF>(v));
    v = op(v, dpp_move<0x141, 0xF>(v));
    v = op(v, dpp_move<0x140, 0xF>(v));
    v = op(v, dpp_move<0x142, 0xA>(v));
    v = op(v, dpp_move<0x143, 0xC>(v));
    return v;
}

// Per-block stage partials (256 threads, every thread active): max |r| and sum r^2 of the
// four residual components to partials[blk * 8 ...] and, with lampart, the block's largest
// |u|+c and |v|+c to lampart[blk * 2 ...].
__device__ __forceinline__ void block_partials(const double* rmax, const double* rsum, double* partials,
                                               int blk, double* lampart, double lx, double ly) {
    __shared__ double red[4][10];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const double a = wave_reduce63<true>(rmax[k]);
        const double b = wave_reduce63<false>(rsum[k]);
        if (lane == 63) { red[wv][k] = a; red[wv][4 + k] = b; }
    }
    if (lampart) {
        lx = wave_reduce63<true>(lx);
        ly = wave_reduce63<true>(ly);
        if (lane == 63) { red[wv][8] = lx; red[wv][9] = ly; }
    }
    __syncthreads();
    if (threadIdx.x < 8) {
        const int k = threadIdx.x;
        double acc = red[0][k];
        for (int w = 1; w < 4; w++) acc = k < 4 ? fmax(acc, red[w][k]) : acc + red[w][k];
        partials[(size_t)blk * 8 + k] = acc;
    } else if (lampart && threadIdx.x < 10) {
        const int k = threadIdx.x;
        double acc = red[0][k];
        for (int w = 1; w < 4; w++) acc = fmax(acc, red[w][k]);
        lampart[(size_t)blk * 2 + (k - 8)] = acc;
    }
}

// zero partials of a block with no leaves
__device__ __forceinline__ void block_partials_zero(double* partials, int blk, double* lampart) {
    if (threadIdx.x < 8) partials[(size_t)blk * 8 + threadIdx.x] = 0.0;
    else if (lampart && threadIdx.x < 10) lampart[(size_t)blk * 2 + (threadIdx.x - 8)] = 0.0;
}

// Device views of the flattened tables (all pointers are device memory).
struct DevTables {
    int N, G, P, L, S_w, S_r;
    int S_wp;                   // WLS slots [0, S_wp) read no ghost leaf
    double gamma, cv, h, xf, Ma;
    double inflow[4];
    const int32_t* stencil;     // [N][8]: one 32-byte record per leaf (+1x, -1x, +1y, -1y, +2x, -2x, +2y, -2y)
    const int16_t* st16;        // [N][8]: the same in 16 bytes (Tables::st16)
    const int32_t* st16_base;   // [ceil(N / ST16_BLK)]
    const int8_t* lvl;          // [N]
    const double* wls_d;        // [S_w][WLS_STRIDE]
    const int32_t* wls_i;       // [S_w][WLSI_STRIDE]
    const int32_t* rst_node;    // [S_r]
    const int32_t* bc_partner;  // [G]
    const int32_t* node_son;    // [n_node]
    const int32_t* node_state;  // [n_node]
    int n_int;
    const int32_t* int_child;   // [n_int][4]
    const int32_t* rst_ref;     // [S_r]
    const int32_t* bc_ref;      // [G]
    const int4* dep_hdr;        // [S_w - S_wp]: ghost points of the slots that read ghosts
    const int4* dep_leaf;       // [S_w - S_wp][2]: DEP_ENT leaves each
    const float4* dep_w;        // [S_w - S_wp][2]: their weights
    const int32_t* slot_blk_pat;  // [ceil(S_w / SLOT_BLK)][SLOT_BLK_PATS] records staged per block
    const uint8_t* slot_lpat;     // [S_w] position of the slot's record in its block's list
    const int32_t* slot_uncov;    // [S_unc] the slots whose record is not in their block's list
    int S_unc;
    const uint8_t* face_share;    // [N] bit 0: x-1/2 face = left neighbour's x+1/2; bit 1: same in y
    const int32_t* face_lo;       // leaves whose x-1/2 face is not shared ([0, n_lo_x)), then y-1/2
    int n_lo_x, n_lo_y;
};

// the eight stencil ids of interior leaf i (two 16-byte loads of its record)
__device__ __forceinline__ void load_stencil(const DevTables& t, int i, int s[8]) {
    const int4* r = reinterpret_cast<const int4*>(t.stencil + (size_t)i * 8);
    const int4 a = r[0], b = r[1];
    s[0] = a.x; s[1] = a.y; s[2] = a.z; s[3] = a.w;
    s[4] = b.x; s[5] = b.y; s[6] = b.z; s[7] = b.w;
}

// the same from the 16-byte record (one 16-byte load and the block's base); a leaf whose
// record is escaped reads the 32-byte one
__device__ __forceinline__ void load_stencil16(const DevTables& t, int i, int s[8]) {
    const int4 r = *reinterpret_cast<const int4*>(t.st16 + (size_t)i * 8);
    const int base = t.st16_base[i / ST16_BLK];
    const int w[4] = {r.x, r.y, r.z, r.w};
    int e[8];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        e[2 * k] = (int)(int16_t)(w[k] & 0xFFFF);
        e[2 * k + 1] = w[k] >> 16;
    }
    if (e[0] == ST16_ESC) {
        load_stencil(t, i, s);
        return;
    }
#pragma unroll
    for (int k = 0; k < 8; k++) s[k] = e[k] >= ST16_FAR ? base + (e[k] - ST16_FAR) : i + e[k];
}

// Scalars of the device-resident step loop (solution.F90:183-215).
struct DevScalars {
    double dt, lam_x, lam_y, tt, t_end, cfl, dt_fixed, h, xf;
    double res_max[4], res_sum[4], res_rms[4];
    // status: 0, -3 NaN residual (solution.F90:209-214), -4 non-positive rho or p in a stage;
    // pos_bad: set by a stage when any leaf's FP64 rho or p of its result is <= 0
    int32_t localdt, l_max, steps_done, status, pos_bad;
};

void launch_slots(const DevTables& t, double4* Q, const double* T, hipStream_t st);
// FP64 stage; with reduce, residual partials per block, and with lampart also the wave-speed
// partials of the output state (the next step's lambda, 2 doubles per block)
void launch_stage(const DevTables& t, const double4* Qin, const double4* Qn, const double* Tn,
                  double4* Qout, double* Tout, const DevScalars* sc, double alpha, double beta,
                  int reduce, double* partials, double* lampart, int* posflag, hipStream_t st);
// the same stage with shared faces: every leaf's x+1/2, y+1/2 faces into fxp, fyp and its
// unshared lower faces into fxm, fym ([N] each), one face per thread, then the update reading
// them (bit-identical to launch_stage)
void launch_stage_shared(const DevTables& t, const double4* Qin, const double4* Qn, const double* Tn,
                         double4* Qout, double* Tout, const DevScalars* sc, double alpha, double beta, int reduce,
                         double* partials, double* lampart, int* posflag, double4* fxp, double4* fyp, double4* fxm,
                         double4* fym, hipStream_t st);
// ghost fill and restriction slots after a stage (restriction means one level per launch)
void launch_restrict_fill(const DevTables& t, const std::vector<int32_t>& level_start, double4* Q,
                          double* T, float4* Qf, double4* Qi, double* Ti, hipStream_t st);
void launch_wavespeed(const DevTables& t, const double4* Q, double* partials, hipStream_t st);
void launch_step_begin(const double* partials, int nb, DevScalars* sc, hipStream_t st);
// End of a step: residual, time advance and (with lampart, the final stage's wave speeds) the
// next step's lambda and dt. Partials: nb blocks, or with meta the fast path's split (full
// blocks from 0, linear blocks from lin0). parts: STEP_PARTS * 10 doubles of scratch.
constexpr int STEP_PARTS = 64;
void launch_step_end(const double* partials, int nb, const int* meta, int lin0, int n_leaf, int advance_time,
                     DevScalars* sc, const double* lampart, double* parts, hipStream_t st);
void launch_sensors(const DevTables& t, const double4* Q, double* rot, double* div, hipStream_t st);
void launch_positivity(const DevTables& t, const double4* Q, int* nbad, hipStream_t st);
// fdmamr_max_wavespeed_noalloc: the wave-speed reduction through a static device scratch array
// (NOALLOC_BLOCKS blocks of 256 leaves at most). noalloc_scratch_init takes the array's address
// on `device`, which must be current; launch_wavespeed_noalloc writes the two maxima to the
// host array out2 (asynchronously on st).
constexpr int NOALLOC_BLOCKS = 131072;
constexpr int NOALLOC_DEVICES = 64;
hipError_t noalloc_scratch_init(int device);
hipError_t launch_wavespeed_noalloc(const DevTables& t, const double4* Q, int device, double* out2,
                                    hipStream_t st);
void launch_stencil_values(const DevTables& t, const double4* Q, int i, double* out, hipStream_t st);

// fast path (fdmamr_fast.hip)
void launch_slots_fast(const DevTables& t, const float* wf, const double4* Q, float4* Qf, hipStream_t st);
// the fill items items[0, nitems) of the flattened lists and the WLS slots of the same state in
// one launch (points of a slot that are ghost leaves are formed from their fill lists)
void launch_slots_fill(const DevTables& t, const float* wf, double4* Q, float4* Qf, const int32_t* items,
                       int nitems, const int32_t* beg, const int32_t* leaf, const float* w, hipStream_t st);
// Selector judge folded into the labelling stage: flags == nullptr for none. cnt: the selector
// counters (troubled, judged, rejected), bound: the label bound.
struct JudgeArgs {
    uint8_t* flags = nullptr;
    int* cnt = nullptr;
    float bound = 0.f;
};
// lists: cleared leaves then troubled leaves (meta[0] = number cleared; the troubled part feeds
// the full path, the cleared leaves are taken from flags bit 0), or nullptr for the full path
// on every leaf; residual partials occupy stage_fast_partial_blocks() blocks. With
// label, each troubled leaf's largest weight deviation goes to label, and with jd.flags (lists
// only) the judge runs on them in the same launch. posflag (or nullptr) is set to 1 when a
// leaf's FP64 result has rho <= 0 or p <= 0.
// The mixed launch (lists != nullptr) takes its troubled units from claim counters kept in meta
// after the list split: meta holds LIST_META_INTS ints, [CLAIM_OFF, LIST_META_INTS) zero before
// the first launch (each launch leaves them zero). STAGE_EARLY claimer blocks come before the
// linear blocks and STAGE_LATE after them (multiples of CLAIM_GROUPS, the XCD count).
constexpr int CLAIM_GROUPS = 8, CLAIM_STRIDE = 16, CLAIM_OFF = 16;
constexpr int LIST_META_INTS = CLAIM_OFF + CLAIM_STRIDE * CLAIM_GROUPS + CLAIM_STRIDE;
constexpr int STAGE_EARLY = 384, STAGE_LATE = 1536;
static_assert(STAGE_EARLY % CLAIM_GROUPS == 0 && STAGE_LATE % CLAIM_GROUPS == 0, "claimers per XCD group");
void launch_stage_fast(const DevTables& t, const int32_t* lists, int* meta, const uint8_t* flags,
                       const double4* Qin, const float4* Qf_in, const double4* Qn, double4* Qout, float4* Qf_out,
                       const DevScalars* sc, double alpha, double beta, int reduce, double* partials,
                       double* lampart, float* label, const JudgeArgs& jd, int* posflag, hipStream_t st);
int stage_fast_partial_blocks(int n, bool lists);
// leaf lists from selector flags (bit 0: troubled); blk_cnt holds list_blocks(n) ints, the
// troubled count of each block of LIST_BLK leaves (turned into the blocks' prefix offsets).
// launch_list_write is the prefix scan and second pass alone, for counts written by another
// kernel.
constexpr int LIST_BLK = 2048;
void launch_build_lists(int n, const uint8_t* flags, int* blk_cnt, int32_t* lists, int* meta, hipStream_t st);
void launch_list_write(int n, const uint8_t* flags, int* blk_cnt, int32_t* lists, int* meta,
                       hipStream_t st);
int list_blocks(int n);
void launch_shadow(const double4* Q, float4* Qf, int n, hipStream_t st);
void launch_stencil_values_f(const DevTables& t, const float4* Qf, int i, float* out, hipStream_t st);
// ghosts and restriction slots of the fast path in one pass (flattened weighted sums)
void launch_fill_flat(const DevTables& t, double4* Q, float4* Qf, const int32_t* beg, const int32_t* leaf,
                      const float* w, hipStream_t st);

}  // namespace fdm
