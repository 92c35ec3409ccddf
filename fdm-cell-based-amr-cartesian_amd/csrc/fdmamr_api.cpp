// fdmamr_api.cpp -- the C-ABI (include/fdmamr.h): context, device buffers, stage sequencing
// and the regrid crossing (device state -> host tree -> new tables -> device).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/fdmamr.h"
#include "fdmamr_device.hpp"
#include "fdmamr_selector.hpp"
#include "host_tables.hpp"
#include "host_tree.hpp"

using namespace fdm;

static_assert(FDMAMR_NOALLOC_MAX_LEAVES == NOALLOC_BLOCKS * 256, "no-allocation scratch size");


template <typename T>
struct DBuf {
    T* p = nullptr;
    size_t cap = 0;
};

struct fdmamr_ctx {
    fdmamr_params p{};
    Tree tree;
    Tables tb;
    hipStream_t stream = nullptr;
    bool host_only = false;  // mesh and table building only; no device calls
    bool tables_ready = false;
    std::string err;

    DBuf<double4> Q[3];
    DBuf<double> T[3];
    DBuf<float4> Qf[3];      // FP32 shadow of the stage buffers (fast path gathers)
    DBuf<double4> Qint;      // restriction means of internal nodes (scratch, per stage)
    DBuf<double> Tint;
    DBuf<int32_t> int_child, rst_ref, bc_ref;
    DBuf<int32_t> fill_beg, fill_leaf, fill_fast;  // flattened fill (fast path)
    DBuf<int32_t> dep_hdr, dep_leaf;                // ghost entries of ghost-reading slots
    DBuf<float> dep_w;
    DBuf<float> fill_w;
    DBuf<float> wls_f;       // FP32 slot records (fast path)
    DBuf<uint8_t> flags;     // selector flags per interior leaf (bit 0: troubled)
    DBuf<float> label;       // per-leaf WCNS weight deviation (labels / judge)
    bool use_flags = false;  // fast path consults the flags
    bool merge_off = false;  // keep the slot launch and the ghost fill apart even where they merge
    DBuf<int32_t> lists;     // cleared leaves, then troubled leaves (from the flags)
    DBuf<int32_t> slot_blk_pat;  // distinct slot records of each slot block
    DBuf<uint8_t> slot_lpat;     // each slot's position in its block's list
    DBuf<int32_t> slot_uncov;    // slots whose record is not staged
    int n_cu = 256;
    DBuf<int> list_cnt, list_meta;

    // learned selector
    DBuf<int32_t> off16;
    DBuf<int32_t> sel_nbr_beg, sel_nbr;  // selector tile neighbourhoods
    DBuf<uint8_t> sel_loc;
    DBuf<float> sel_w, sel_m, sel_v, sel_grad, sel_gpart, sel_lpart, sel_loss, sel_prob;
    DBuf<uint16_t> sel_pk;
    DBuf<int> sel_cnt;
    DBuf<uint16_t> ds_x;     // training rows (BF16 features)
    DBuf<uint8_t> ds_y;
    DBuf<int32_t> ds_idx;
    int ds_rows = 0;
    bool sel_ready = false;
    float sel_threshold = 0.5f, sel_judge = 0.02f, sel_bound = 0.05f;
    uint32_t sel_step = 0;
    int adam_t = 0;
    DBuf<int32_t> stencil, wls_i, rst, bcp, son, state;
    DBuf<int16_t> st16;        // 16-byte stencil records (fast-path stage)
    DBuf<int32_t> st16_base;
    DBuf<uint8_t> face_share;  // FP64 path: which lower faces equal the neighbour's upper face
    DBuf<double4> fxp, fyp;    // FP64 path: every leaf's x+1/2 and y+1/2 face fluxes of a stage
    DBuf<double4> fxm, fym;    // FP64 path: the lower faces that are not shared
    DBuf<int32_t> face_lo;
    bool share_off = false;    // FP64 path: compute every face in the update kernel
    DBuf<int8_t> lvl;
    DBuf<double> wls_d, partials, lampart, rot, div, tmp, parts;
    DBuf<int> nbad;
    DevScalars* sc = nullptr;
    DevTables dt{};

    // pinned staging for large host-to-device copies (two buffers, alternating)
    void* stage_buf[2] = {nullptr, nullptr};
    hipEvent_t stage_ev[2] = {nullptr, nullptr};
    bool stage_pending[2] = {false, false};

    // kernel timing with HIP events on the context's stream (bench / profiling)
    bool profiling = false;
    std::vector<hipEvent_t> ev_pool;
    std::vector<std::pair<int, int>> ev_stage;  // (start, stop) indices around k_stage launches
    std::vector<std::pair<int, int>> ev_select; // (start, stop) indices around k_select launches
    int ev_used = 0;
};

namespace {

int fail(fdmamr_ctx* c, int code, const std::string& msg) {
    if (c) c->err = msg;
    return code;
}

#define HIPCHK(ctx, expr)                                                                   \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess)                                                               \
            return fail(ctx, -2, std::string("HIP error in ") + #expr + ": " + hipGetErrorString(e_)); \
    } while (0)

template <typename T>
hipError_t ensure(DBuf<T>& b, size_t n) {
    if (n == 0) n = 1;
    if (b.cap >= n) return hipSuccess;
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.cap = 0;
    size_t want = n + n / 4 + 64;
    hipError_t e = hipMalloc((void**)&b.p, want * sizeof(T));
    if (e == hipSuccess) b.cap = want;
    return e;
}
template <typename T>
void release(DBuf<T>& b) {
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.cap = 0;
}

int nblocks(int n, int b) { return (n + b - 1) / b; }

int tree_depth(const Tree& t) {
    int d = 0;
    for (size_t c = 0; c < t.n_alloc(); c++)
        if (t.father[c] != -2) d = std::max(d, t.lvl[c]);
    return d;
}

// Host-to-device copy of a large host array through two pinned staging buffers: host threads
// copy one chunk into a buffer while the DMA engine reads the other, so the transfer runs at
// the pinned rate rather than the runtime's single-threaded pageable path. Enqueued on the
// context's stream; the source may be reused once the stream has passed the copy.
constexpr size_t STAGE_BYTES = (size_t)32 << 20;
hipError_t upload_staged(fdmamr_ctx* c, void* dst, const void* src, size_t bytes) {
    if (bytes == 0) return hipSuccess;
    if (bytes < ((size_t)4 << 20)) return hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream);
    hipError_t e = hipSuccess;
    for (int k = 0; k < 2 && e == hipSuccess; k++) {
        if (!c->stage_buf[k]) e = hipHostMalloc(&c->stage_buf[k], STAGE_BYTES, hipHostMallocDefault);
        if (e == hipSuccess && !c->stage_ev[k]) e = hipEventCreateWithFlags(&c->stage_ev[k], hipEventDisableTiming);
    }
    if (e != hipSuccess) return e;
    for (size_t off = 0, k = 0; off < bytes; off += STAGE_BYTES, k ^= 1) {
        const size_t n = std::min(STAGE_BYTES, bytes - off);
        if (c->stage_pending[k]) {
            e = hipEventSynchronize(c->stage_ev[k]);
            if (e != hipSuccess) return e;
            c->stage_pending[k] = false;
        }
        unsigned char* buf = static_cast<unsigned char*>(c->stage_buf[k]);
        const unsigned char* from = static_cast<const unsigned char*>(src) + off;
        parallel_for(n, [&](size_t b, size_t e2) { std::memcpy(buf + b, from + b, e2 - b); }, (size_t)1 << 20);
        e = hipMemcpyAsync(static_cast<unsigned char*>(dst) + off, buf, n, hipMemcpyHostToDevice, c->stream);
        if (e == hipSuccess) e = hipEventRecord(c->stage_ev[k], c->stream);
        if (e != hipSuccess) return e;
        c->stage_pending[k] = true;
    }
    return hipSuccess;
}

// Device-to-host copy of a large array through the same staging buffers: the DMA engine fills
// one buffer while host threads empty the other. Synchronises the context's stream.
hipError_t download_staged(fdmamr_ctx* c, void* dst, const void* src, size_t bytes) {
    if (bytes == 0) return hipStreamSynchronize(c->stream);
    if (bytes < ((size_t)4 << 20)) {
        hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream);
        return e == hipSuccess ? hipStreamSynchronize(c->stream) : e;
    }
    hipError_t e = hipSuccess;
    for (int k = 0; k < 2 && e == hipSuccess; k++) {
        if (!c->stage_buf[k]) e = hipHostMalloc(&c->stage_buf[k], STAGE_BYTES, hipHostMallocDefault);
        if (e == hipSuccess && !c->stage_ev[k]) e = hipEventCreateWithFlags(&c->stage_ev[k], hipEventDisableTiming);
        if (e == hipSuccess && c->stage_pending[k]) e = hipEventSynchronize(c->stage_ev[k]);
        c->stage_pending[k] = false;
    }
    if (e != hipSuccess) return e;
    const size_t nchunk = (bytes + STAGE_BYTES - 1) / STAGE_BYTES;
    auto issue = [&](size_t ch) -> hipError_t {
        const size_t off = ch * STAGE_BYTES, n = std::min(STAGE_BYTES, bytes - off);
        hipError_t r = hipMemcpyAsync(c->stage_buf[ch & 1], static_cast<const unsigned char*>(src) + off, n,
                                      hipMemcpyDeviceToHost, c->stream);
        return r == hipSuccess ? hipEventRecord(c->stage_ev[ch & 1], c->stream) : r;
    };
    e = issue(0);
    for (size_t ch = 0; ch < nchunk && e == hipSuccess; ch++) {
        if (ch + 1 < nchunk) e = issue(ch + 1);
        if (e == hipSuccess) e = hipEventSynchronize(c->stage_ev[ch & 1]);
        if (e != hipSuccess) break;
        const size_t off = ch * STAGE_BYTES, n = std::min(STAGE_BYTES, bytes - off);
        const unsigned char* buf = static_cast<const unsigned char*>(c->stage_buf[ch & 1]);
        unsigned char* to = static_cast<unsigned char*>(dst) + off;
        parallel_for(n, [&](size_t b, size_t e2) { std::memcpy(to + b, buf + b, e2 - b); }, (size_t)1 << 20);
    }
    return e == hipSuccess ? hipStreamSynchronize(c->stream) : e;
}

int upload_tables(fdmamr_ctx* c) {
    Tree& t = c->tree;
    Tables& tb = c->tb;
    const int slots = tb.S_w + tb.S_r;
    const size_t ext = (size_t)tb.L + slots;
    for (int k = 0; k < 3; k++) {
        HIPCHK(c, ensure(c->Q[k], ext));
        HIPCHK(c, ensure(c->T[k], ext));
        HIPCHK(c, ensure(c->Qf[k], ext));
    }
    HIPCHK(c, ensure(c->wls_f, tb.wls_f.size()));
    HIPCHK(c, ensure(c->flags, tb.N));
    HIPCHK(c, ensure(c->lists, tb.N));
    HIPCHK(c, ensure(c->list_cnt, list_blocks(tb.N)));
    HIPCHK(c, ensure(c->list_meta, LIST_META_INTS));
    // the mixed stage's claim counters start at zero (each launch leaves them so)
    HIPCHK(c, hipMemsetAsync(c->list_meta.p, 0, LIST_META_INTS * sizeof(int), c->stream));
    HIPCHK(c, ensure(c->label, tb.N));
    HIPCHK(c, ensure(c->stencil, tb.stencil.size()));
    HIPCHK(c, ensure(c->lvl, tb.lvl_leaf.size()));
    HIPCHK(c, ensure(c->wls_d, tb.wls_d.size()));
    HIPCHK(c, ensure(c->wls_i, tb.wls_i.size()));
    HIPCHK(c, ensure(c->rst, tb.rst_node.size()));
    HIPCHK(c, ensure(c->bcp, tb.bc_partner.size()));
    HIPCHK(c, ensure(c->son, tb.node_son.size()));
    HIPCHK(c, ensure(c->state, tb.state_of_node.size()));
    const int nb = std::max(1, nblocks(tb.N, 256));
    HIPCHK(c, ensure(c->partials, (size_t)stage_fast_partial_blocks(tb.N, true) * 8));
    HIPCHK(c, ensure(c->lampart, (size_t)stage_fast_partial_blocks(tb.N, true) * 2));
    HIPCHK(c, ensure(c->rot, tb.N));
    HIPCHK(c, ensure(c->div, tb.N));
    HIPCHK(c, ensure(c->tmp, 64));
    HIPCHK(c, ensure(c->parts, STEP_PARTS * 10));
    HIPCHK(c, ensure(c->nbad, 1));
    hipStream_t s = c->stream;
    auto up = [&](void* d, const void* h, size_t bytes) -> hipError_t { return upload_staged(c, d, h, bytes); };
    HIPCHK(c, up(c->stencil.p, tb.stencil.data(), tb.stencil.size() * 4));
    HIPCHK(c, ensure(c->st16, std::max<size_t>(tb.st16.size(), 8)));
    HIPCHK(c, ensure(c->st16_base, std::max<size_t>(tb.st16_base.size(), 1)));
    if (!tb.st16.empty()) HIPCHK(c, up(c->st16.p, tb.st16.data(), tb.st16.size() * 2));
    if (!tb.st16_base.empty()) HIPCHK(c, up(c->st16_base.p, tb.st16_base.data(), tb.st16_base.size() * 4));
    HIPCHK(c, ensure(c->face_share, tb.face_share.size()));
    HIPCHK(c, up(c->face_share.p, tb.face_share.data(), tb.face_share.size()));
    if (c->p.flux_precision == FDMAMR_FLUX_FP64) {
        HIPCHK(c, ensure(c->fxp, tb.N));
        HIPCHK(c, ensure(c->fyp, tb.N));
        HIPCHK(c, ensure(c->fxm, tb.N));
        HIPCHK(c, ensure(c->fym, tb.N));
        HIPCHK(c, ensure(c->face_lo, tb.face_lo.size()));
        HIPCHK(c, up(c->face_lo.p, tb.face_lo.data(), tb.face_lo.size() * 4));
    }
    HIPCHK(c, up(c->lvl.p, tb.lvl_leaf.data(), tb.lvl_leaf.size()));
    HIPCHK(c, up(c->wls_d.p, tb.wls_d.data(), tb.wls_d.size() * 8));
    HIPCHK(c, up(c->wls_i.p, tb.wls_i.data(), tb.wls_i.size() * 4));
    HIPCHK(c, up(c->rst.p, tb.rst_node.data(), tb.rst_node.size() * 4));
    HIPCHK(c, up(c->bcp.p, tb.bc_partner.data(), tb.bc_partner.size() * 4));
    HIPCHK(c, up(c->son.p, tb.node_son.data(), tb.node_son.size() * 4));
    HIPCHK(c, up(c->state.p, tb.state_of_node.data(), tb.state_of_node.size() * 4));
    // state of every leaf (interior, ghost, corner) into all three stage buffers
    BigVec<double4> q(tb.L);
    BigVec<double> tv(tb.L);
    parallel_for((size_t)tb.L, [&](size_t b, size_t e) {
        for (size_t i = b; i < e; i++) {
            const auto& u = t.U[tb.leaf_node[i]];
            q[i] = make_double4(u[0], u[1], u[2], u[4]);
            tv[i] = u[3];
        }
    });
    HIPCHK(c, up(c->Q[0].p, q.data(), q.size() * sizeof(double4)));
    HIPCHK(c, up(c->T[0].p, tv.data(), tv.size() * 8));
    for (int k = 1; k < 3; k++) {
        HIPCHK(c, hipMemcpyAsync(c->Q[k].p, c->Q[0].p, q.size() * sizeof(double4), hipMemcpyDeviceToDevice, s));
        HIPCHK(c, hipMemcpyAsync(c->T[k].p, c->T[0].p, tv.size() * 8, hipMemcpyDeviceToDevice, s));
    }
    HIPCHK(c, up(c->wls_f.p, tb.wls_f.data(), tb.wls_f.size() * 4));
    HIPCHK(c, ensure(c->Qint, tb.n_int));
    HIPCHK(c, ensure(c->Tint, tb.n_int));
    HIPCHK(c, ensure(c->int_child, tb.int_child.size()));
    HIPCHK(c, ensure(c->rst_ref, tb.rst_ref.size()));
    HIPCHK(c, ensure(c->bc_ref, tb.bc_ref.size()));
    HIPCHK(c, up(c->int_child.p, tb.int_child.data(), tb.int_child.size() * 4));
    HIPCHK(c, up(c->rst_ref.p, tb.rst_ref.data(), tb.rst_ref.size() * 4));
    HIPCHK(c, up(c->bc_ref.p, tb.bc_ref.data(), tb.bc_ref.size() * 4));
    HIPCHK(c, ensure(c->fill_beg, tb.fill_beg.size()));
    HIPCHK(c, ensure(c->fill_leaf, tb.fill_leaf.size()));
    HIPCHK(c, ensure(c->fill_w, tb.fill_w.size()));
    HIPCHK(c, up(c->fill_beg.p, tb.fill_beg.data(), tb.fill_beg.size() * 4));
    HIPCHK(c, up(c->fill_leaf.p, tb.fill_leaf.data(), tb.fill_leaf.size() * 4));
    HIPCHK(c, up(c->fill_w.p, tb.fill_w.data(), tb.fill_w.size() * 4));
    HIPCHK(c, ensure(c->dep_hdr, tb.dep_hdr.size()));
    HIPCHK(c, ensure(c->dep_leaf, tb.dep_leaf.size()));
    HIPCHK(c, ensure(c->dep_w, tb.dep_w.size()));
    if (!tb.dep_hdr.empty()) {
        HIPCHK(c, up(c->dep_hdr.p, tb.dep_hdr.data(), tb.dep_hdr.size() * 4));
        HIPCHK(c, up(c->dep_leaf.p, tb.dep_leaf.data(), tb.dep_leaf.size() * 4));
        HIPCHK(c, up(c->dep_w.p, tb.dep_w.data(), tb.dep_w.size() * 4));
    }
    HIPCHK(c, ensure(c->fill_fast, std::max<size_t>(tb.fill_fast.size(), 1)));
    if (!tb.fill_fast.empty()) HIPCHK(c, up(c->fill_fast.p, tb.fill_fast.data(), tb.fill_fast.size() * 4));
    HIPCHK(c, ensure(c->off16, tb.off16.size()));
    HIPCHK(c, ensure(c->sel_prob, tb.N));
    HIPCHK(c, up(c->off16.p, tb.off16.data(), tb.off16.size() * 4));
    HIPCHK(c, ensure(c->sel_nbr_beg, tb.sel_nbr_beg.size()));
    HIPCHK(c, ensure(c->sel_nbr, tb.sel_nbr.size()));
    HIPCHK(c, ensure(c->sel_loc, tb.sel_loc.size()));
    HIPCHK(c, up(c->sel_nbr_beg.p, tb.sel_nbr_beg.data(), tb.sel_nbr_beg.size() * 4));
    HIPCHK(c, up(c->sel_nbr.p, tb.sel_nbr.data(), tb.sel_nbr.size() * 4));
    HIPCHK(c, up(c->sel_loc.p, tb.sel_loc.data(), tb.sel_loc.size()));
    HIPCHK(c, ensure(c->slot_blk_pat, std::max<size_t>(tb.slot_blk_pat.size(), 1)));
    HIPCHK(c, ensure(c->slot_lpat, std::max<size_t>(tb.slot_lpat.size(), 1)));
    if (!tb.slot_blk_pat.empty()) HIPCHK(c, up(c->slot_blk_pat.p, tb.slot_blk_pat.data(), tb.slot_blk_pat.size() * 4));
    if (!tb.slot_lpat.empty()) HIPCHK(c, up(c->slot_lpat.p, tb.slot_lpat.data(), tb.slot_lpat.size()));
    HIPCHK(c, ensure(c->slot_uncov, std::max<size_t>(tb.slot_uncov.size(), 1)));
    if (!tb.slot_uncov.empty()) HIPCHK(c, up(c->slot_uncov.p, tb.slot_uncov.data(), tb.slot_uncov.size() * 4));
    // every leaf troubled until the selector says otherwise; the leaves the last regrid made
    // (copied or averaged values) and the leaves whose stencil reads one of them stay
    // troubled through the selector's first decision
    {
        std::vector<uint8_t> f0(tb.N);
        const bool have_fresh = t.fresh.size() == t.n_alloc();
        auto fresh_node = [&](int q) {
            if (q < 0) return false;
            if (t.fresh[q]) return true;
            const int sq = t.son[q];  // a parent's level-jump value reads its sons
            return sq >= 0 && (t.fresh[sq] || t.fresh[sq + 1] || t.fresh[sq + 2] || t.fresh[sq + 3]);
        };
        parallel_for((size_t)tb.N, [&](size_t b, size_t e) {
            for (size_t i = b; i < e; i++) {
                bool fr = false;
                if (have_fresh) {
                    fr = t.fresh[tb.leaf_node[i]] != 0;
                    for (int k = 0; k < 8 && !fr; k++) fr = fresh_node(tb.st_node[i * 8 + k]);
                }
                f0[i] = (uint8_t)(1 | (fr ? 8 : 0));
            }
        });
        if (tb.N > 0) HIPCHK(c, hipMemcpyAsync(c->flags.p, f0.data(), tb.N, hipMemcpyHostToDevice, s));
        HIPCHK(c, hipStreamSynchronize(s));
    }
    launch_build_lists(tb.N, c->flags.p, c->list_cnt.p, c->lists.p, c->list_meta.p, s);
    DevTables& d = c->dt;
    d.N = tb.N; d.G = tb.G; d.P = tb.P; d.L = tb.L; d.S_w = tb.S_w; d.S_r = tb.S_r; d.S_wp = tb.S_wp;
    d.gamma = t.gamma; d.cv = t.cv; d.h = t.h; d.xf = t.xf; d.Ma = t.Ma;
    for (int k = 0; k < 4; k++) d.inflow[k] = t.inflow[k];
    d.stencil = c->stencil.p; d.lvl = c->lvl.p;
    d.st16 = c->st16.p; d.st16_base = c->st16_base.p; d.wls_d = c->wls_d.p; d.wls_i = c->wls_i.p;
    d.dep_hdr = reinterpret_cast<const int4*>(c->dep_hdr.p);
    d.dep_leaf = reinterpret_cast<const int4*>(c->dep_leaf.p);
    d.dep_w = reinterpret_cast<const float4*>(c->dep_w.p);
    d.slot_blk_pat = c->slot_blk_pat.p;
    d.slot_lpat = c->slot_lpat.p;
    d.slot_uncov = c->slot_uncov.p;
    d.face_share = c->face_share.p;
    d.face_lo = c->face_lo.p;
    d.n_lo_x = tb.n_lo_x;
    d.n_lo_y = tb.n_lo_y;
    d.S_unc = (int)tb.slot_uncov.size();
    d.rst_node = c->rst.p; d.bc_partner = c->bcp.p; d.node_son = c->son.p; d.node_state = c->state.p;
    d.n_int = tb.n_int; d.int_child = c->int_child.p; d.rst_ref = c->rst_ref.p; d.bc_ref = c->bc_ref.p;
    // ghosts and restriction slots of the uploaded state, FP32 shadows of all three buffers
    launch_restrict_fill(d, tb.int_level_start, c->Q[0].p, c->T[0].p, nullptr, c->Qint.p, c->Tint.p, s);
    for (int k = 0; k < 3; k++) launch_shadow(c->Q[k].p, c->Qf[k].p, (int)ext, s);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(s));
    return 0;
}

void fill_after_stage(fdmamr_ctx* c, int buf) {
    const bool fast = c->p.flux_precision == FDMAMR_FLUX_FP32;
    if (fast && c->tb.fill_flat_ok) {
        launch_fill_flat(c->dt, c->Q[buf].p, c->Qf[buf].p, c->fill_beg.p, c->fill_leaf.p, c->fill_w.p, c->stream);
        return;
    }
    launch_restrict_fill(c->dt, c->tb.int_level_start, c->Q[buf].p, fast ? nullptr : c->T[buf].p,
                         fast ? c->Qf[buf].p : nullptr, c->Qint.p, c->Tint.p, c->stream);
}

// the fast path refills the ghosts inside its slot launch
bool slots_merged(const fdmamr_ctx* c) { return c->tb.merge_ok && !c->merge_off; }

// After a run of steps on the fast path with merged fills, every ghost and restriction slot
// of the final state (buffer 0), so that the state seen from outside is the same as after a
// fill per stage.
void finish_steps(fdmamr_ctx* c) {
    if (c->p.flux_precision == FDMAMR_FLUX_FP32 && slots_merged(c))
        launch_fill_flat(c->dt, c->Q[0].p, c->Qf[0].p, c->fill_beg.p, c->fill_leaf.p, c->fill_w.p, c->stream);
}

// device state (interior + ghost leaves of buffer 0) -> host tree
int pull_state(fdmamr_ctx* c) {
    Tables& tb = c->tb;
    const int n = tb.N + tb.G;
    BigVec<double4> q(n);
    BigVec<double> tv(n);
    const bool fast = c->p.flux_precision == FDMAMR_FLUX_FP32;
    HIPCHK(c, download_staged(c, q.data(), c->Q[0].p, (size_t)n * sizeof(double4)));
    if (!fast) HIPCHK(c, download_staged(c, tv.data(), c->T[0].p, (size_t)n * 8));
    const double gm = c->tree.gamma * c->tree.Ma * c->tree.Ma;
    parallel_for((size_t)n, [&](size_t b, size_t e) {
        for (size_t i = b; i < e; i++) {
            auto& u = c->tree.U[tb.leaf_node[i]];
            u[0] = q[i].x; u[1] = q[i].y; u[2] = q[i].z; u[4] = q[i].w;
            // the fast path does not carry the passive T: recompute it from the equation of state
            u[3] = fast ? q[i].w * gm / q[i].x : tv[i];
        }
    });
    return 0;
}

int check_ready(fdmamr_ctx* c) {
    if (!c) return -1;
    if (c->host_only) return fail(c, -5, "host-only context: no device");
    if (!c->tables_ready) return fail(c, -1, "tables not built: call fdmamr_build_tables first");
    return 0;
}

// selector counters summed over their slots (synchronises the stream)
hipError_t read_counters(fdmamr_ctx* c, int out[CNT_N]) {
    std::vector<int> all(CNT_N * CNT_SLOTS);
    hipError_t e = hipMemcpyAsync(all.data(), c->sel_cnt.p, all.size() * 4, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    for (int k = 0; k < CNT_N; k++) {
        out[k] = 0;
        for (int sl = 0; sl < CNT_SLOTS; sl++) out[k] += all[sl * CNT_N + k];
    }
    return e;
}

int next_event(fdmamr_ctx* c) {
    if (c->ev_used == (int)c->ev_pool.size()) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return -1;
        c->ev_pool.push_back(e);
    }
    return c->ev_used++;
}

// one SSP-RK3 step from buffer 0 back into buffer 0 (solution.F90:267-298); with prepare_next
// the final stage also reduces the wave speeds of its output and the step's end computes the
// next step's lambda and dt (what wavespeed + step_begin would do at the next step's start)
void enqueue_rk3(fdmamr_ctx* c, int advance_time, bool prepare_next) {
    static const double ra[3] = {0.0, 3.0 / 4.0, 1.0 / 3.0};   // solution.F90:392-393
    static const double rb[3] = {1.0, 1.0 / 4.0, 2.0 / 3.0};
    static const int in[3] = {0, 1, 2}, out[3] = {1, 2, 0};
    const DevTables& d = c->dt;
    hipStream_t s = c->stream;
    const bool fast = c->p.flux_precision == FDMAMR_FLUX_FP32;
    // the learned selector runs once per step, on the step's input state, before stage 1;
    // its judge sample takes the full path in stage 1 and is checked in the same launch
    const bool select = fast && c->p.selector && c->sel_ready;
    // flagged fast path: cleared and troubled leaves run as separate launches over the lists
    // built from the flags; the judge's flag changes after stage 1 are not folded back into
    // the lists, so judged leaves stay on the full path for the whole step
    const int32_t* lists = fast && (c->use_flags || select) ? c->lists.p : nullptr;
    // fast path with flattened fills: each stage's input gets its unaliased ghosts and
    // restriction slots in the launch that evaluates its WLS slots (finish_steps completes
    // the ghosts of the final state)
    const bool merged = fast && slots_merged(c);
    for (int k = 0; k < 3; k++) {
        if (merged)
            launch_slots_fill(d, c->wls_f.p, c->Q[in[k]].p, c->Qf[in[k]].p, c->fill_fast.p,
                              (int)c->tb.fill_fast.size(), c->fill_beg.p, c->fill_leaf.p, c->fill_w.p, s);
        else if (fast) launch_slots_fast(d, c->wls_f.p, c->Q[in[k]].p, c->Qf[in[k]].p, s);
        else launch_slots(d, c->Q[in[k]].p, c->T[in[k]].p, s);
        if (select && k == 0) {
            SelectParams sp{c->sel_threshold, (uint32_t)(c->sel_judge * 65536.0f), c->sel_step++};
            int f0 = -1, f1 = -1;
            if (c->profiling && (f0 = next_event(c)) >= 0) (void)hipEventRecord(c->ev_pool[f0], s);
            launch_select(d, c->Qf[0].p, c->off16.p, c->sel_pk.p, c->sel_w.p, sp, c->flags.p, nullptr, c->n_cu,
                          c->sel_nbr_beg.p, c->sel_nbr.p, c->sel_loc.p, c->sel_cnt.p, s);
            if (f0 >= 0 && (f1 = next_event(c)) >= 0) {
                (void)hipEventRecord(c->ev_pool[f1], s);
                c->ev_select.emplace_back(f0, f1);
            }
            launch_build_lists(d.N, c->flags.p, c->list_cnt.p, c->lists.p, c->list_meta.p, s);
        }
        int e0 = -1, e1 = -1;
        // every stage flags a non-positive FP64 density or pressure of its result (no host sync)
        int* posflag = &c->sc->pos_bad;
        if (c->profiling && (e0 = next_event(c)) >= 0) (void)hipEventRecord(c->ev_pool[e0], s);
        if (fast)
            launch_stage_fast(d, lists, c->list_meta.p, c->flags.p, c->Q[in[k]].p, c->Qf[in[k]].p, c->Q[0].p,
                              c->Q[out[k]].p, c->Qf[out[k]].p, c->sc, ra[k], rb[k], k == 2 ? 1 : 0,
                              c->partials.p, (prepare_next && k == 2) ? c->lampart.p : nullptr,
                              (select && k == 0) ? c->label.p : nullptr,
                              (select && k == 0) ? JudgeArgs{c->flags.p, c->sel_cnt.p, c->sel_bound} : JudgeArgs{},
                              posflag, s);
        else if (c->share_off)
            launch_stage(d, c->Q[in[k]].p, c->Q[0].p, c->T[0].p, c->Q[out[k]].p, c->T[out[k]].p, c->sc, ra[k],
                         rb[k], k == 2 ? 1 : 0, c->partials.p, (prepare_next && k == 2) ? c->lampart.p : nullptr,
                         posflag, s);
        else
            launch_stage_shared(d, c->Q[in[k]].p, c->Q[0].p, c->T[0].p, c->Q[out[k]].p, c->T[out[k]].p, c->sc,
                                ra[k], rb[k], k == 2 ? 1 : 0, c->partials.p,
                                (prepare_next && k == 2) ? c->lampart.p : nullptr, posflag, c->fxp.p, c->fyp.p,
                                c->fxm.p, c->fym.p, s);
        if (e0 >= 0 && (e1 = next_event(c)) >= 0) {
            (void)hipEventRecord(c->ev_pool[e1], s);
            c->ev_stage.emplace_back(e0, e1);
        }
        if (!merged) fill_after_stage(c, out[k]);
    }
    const int npart = fast ? stage_fast_partial_blocks(d.N, lists != nullptr) : nblocks(d.N, 256);
    launch_step_end(c->partials.p, npart, lists ? c->list_meta.p : nullptr, nblocks(4 * d.N, 256), d.N, advance_time,
                    c->sc, prepare_next ? c->lampart.p : nullptr, c->parts.p, s);
}

}  // namespace

extern "C" {

int32_t fdmamr_create(const fdmamr_params* p, fdmamr_ctx** out) {
    if (!p || !out) return -1;
    *out = nullptr;
    if (!(p->h > 0) || !(p->gamma > 1.0) || p->xmax <= p->xmin || p->ymax <= p->ymin) return -1;
    if (p->flux_precision != FDMAMR_FLUX_FP64 && p->flux_precision != FDMAMR_FLUX_FP32) return -1;
    const int kinds[4] = {p->bc_xm, p->bc_xp, p->bc_ym, p->bc_yp};
    for (int k : kinds)
        if (k != FDMAMR_BC_PERIODIC && k != FDMAMR_BC_TRANSMISSIVE && k != FDMAMR_BC_INFLOW) return -1;
    const bool host_only = p->device == FDMAMR_DEVICE_HOST_ONLY;
    if (!host_only) {
        int ndev = 0;
        if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return -5;
        if (p->device < 0 || p->device >= ndev) return -5;
    }
    fdmamr_ctx* c = new (std::nothrow) fdmamr_ctx();
    if (!c) return -1;
    c->p = *p;
    c->host_only = host_only;
    if (!host_only && (hipSetDevice(p->device) != hipSuccess || hipStreamCreate(&c->stream) != hipSuccess ||
                       hipMalloc((void**)&c->sc, sizeof(DevScalars)) != hipSuccess ||
                       noalloc_scratch_init(p->device) != hipSuccess)) {
        if (c->sc) (void)hipFree(c->sc);
        if (c->stream) (void)hipStreamDestroy(c->stream);
        delete c;
        return -2;
    }
    if (!host_only) {
        int n_cu = 0;
        if (hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, p->device) == hipSuccess && n_cu > 0)
            c->n_cu = n_cu;
    }
    c->tree.configure(p->xmin, p->xmax, p->ymin, p->ymax, p->h, p->gamma, p->Ma > 0 ? p->Ma : 1.0);
    for (int k = 0; k < 4; k++) { c->tree.bc[k] = kinds[k]; c->tree.inflow[k] = p->inflow[k]; }
    c->tree.wls_local = p->wls_local ? 1 : 0;
    *out = c;
    return 0;
}

void fdmamr_destroy(fdmamr_ctx* c) {
    if (!c) return;
    if (c->host_only) { delete c; return; }
    (void)hipSetDevice(c->p.device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (int k = 0; k < 3; k++) { release(c->Q[k]); release(c->T[k]); }
    release(c->stencil); release(c->st16); release(c->st16_base); release(c->face_share); release(c->fxp); release(c->fyp); release(c->fxm); release(c->fym); release(c->face_lo); release(c->wls_i); release(c->rst); release(c->bcp); release(c->son);
    release(c->state); release(c->lvl); release(c->wls_d); release(c->partials); release(c->lampart);
    release(c->rot); release(c->div); release(c->tmp); release(c->parts); release(c->nbad);
    release(c->Qf[0]); release(c->Qf[1]); release(c->Qf[2]); release(c->wls_f); release(c->flags);
    release(c->label); release(c->off16); release(c->sel_w); release(c->sel_m); release(c->sel_v);
    release(c->sel_grad); release(c->sel_gpart); release(c->sel_lpart); release(c->sel_loss);
    release(c->sel_prob); release(c->sel_pk); release(c->sel_cnt); release(c->ds_x); release(c->ds_y);
    release(c->ds_idx); release(c->lists); release(c->slot_blk_pat); release(c->slot_lpat); release(c->slot_uncov); release(c->list_cnt); release(c->list_meta);

    release(c->fill_beg); release(c->fill_leaf); release(c->fill_w); release(c->fill_fast);
    release(c->dep_hdr); release(c->dep_leaf); release(c->dep_w);
    release(c->sel_nbr_beg); release(c->sel_nbr); release(c->sel_loc);
    for (auto e : c->ev_pool) (void)hipEventDestroy(e);
    for (int k = 0; k < 2; k++) {
        if (c->stage_ev[k]) (void)hipEventDestroy(c->stage_ev[k]);
        if (c->stage_buf[k]) (void)hipHostFree(c->stage_buf[k]);
    }
    if (c->sc) (void)hipFree(c->sc);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

int32_t fdmamr_last_error(const fdmamr_ctx* c, char* buf, int32_t len) {
    if (!c || !buf || len <= 0) return -1;
    std::snprintf(buf, (size_t)len, "%s", c->err.c_str());
    return 0;
}

int32_t fdmamr_creat_mesh(fdmamr_ctx* c) {
    if (!c) return -1;
    try {
        c->tree.creat_mesh();
        c->tree.idgrid();
    } catch (const std::exception& e) {
        return fail(c, -1, e.what());
    }
    c->tables_ready = false;
    return 0;
}

int32_t fdmamr_initial(fdmamr_ctx* c, int32_t ic_kind, uint64_t seed, const double* ic_param8) {
    if (!c) return -1;
    if (c->tree.total == 0) return fail(c, -1, "no mesh: call fdmamr_creat_mesh first");
    if (ic_kind < 0 || ic_kind > 3) return fail(c, -1, "unknown initial condition");
    double z[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    c->tree.initial(ic_kind, seed, ic_param8 ? ic_param8 : z);
    c->tree.idgrid();
    c->tree.boundary_fill_host();
    c->tables_ready = false;
    return 0;
}

int32_t fdmamr_tree_import(fdmamr_ctx* c, int32_t m, int32_t n, int32_t n_node,
                           const int32_t* father, const int32_t* son, const int32_t* lvl,
                           const int32_t* sp, const int32_t* num2, const int32_t* cross,
                           const double* x, const double* y, const double* prim7) {
    if (!c) return -1;
    // the current tables describe the old tree: nothing may use them once the import starts
    c->tables_ready = false;
    if (m < 7 || n < 7 || (int64_t)m * n > n_node || !father || !son || !lvl || !sp || !num2 || !cross || !x ||
        !y || !prim7)
        return fail(c, -1, "fdmamr_tree_import: invalid argument");
    // Validate everything before the context's tree is touched: base cells in the reference's
    // cell(i) order (mesh.F90:19-26), every other node the son sp of its father, son blocks of
    // four consecutive ids that name their father, levels one below the father's.
    const int total = m * n;
    auto bad = [&](int k, const char* what) {
        return fail(c, -1, std::string("fdmamr_tree_import: node ") + std::to_string(k) + ": " + what);
    };
    for (int k = 0; k < n_node; k++) {
        if (cross[k] < -5 || cross[k] > 2) return bad(k, "cross out of range");
        if (!std::isfinite(x[k]) || !std::isfinite(y[k])) return bad(k, "centre not finite");
        if (k < total) {
            if (father[k] != -1 || lvl[k] != 0 || sp[k] != 0) return bad(k, "base cell with a father, level or position");
            if (num2[2 * k] != k / m + 1 || num2[2 * k + 1] != k % m + 1) return bad(k, "base cell index (num) wrong");
        } else {
            const int f = father[k];
            if (f < 0 || f >= n_node) return bad(k, "father out of range");
            if (son[f] < total || k < son[f] || k > son[f] + 3) return bad(k, "not a son of its father");
            if (sp[k] != k - son[f] + 1) return bad(k, "son position (sp) wrong");
            if (lvl[k] != lvl[f] + 1) return bad(k, "level is not the father's plus one");
            if (num2[2 * k] != num2[2 * f] || num2[2 * k + 1] != num2[2 * f + 1]) return bad(k, "base cell index differs from the father's");
        }
        if (lvl[k] < 0 || lvl[k] >= MAX_TREE_DEPTH - 1) return bad(k, "level out of range");
        if (son[k] != -1) {
            if (son[k] < total || son[k] > n_node - 4) return bad(k, "son index out of range");
            for (int j = 0; j < 4; j++)
                if (father[son[k] + j] != k) return bad(k, "son block does not name this node as father");
        }
    }
    Tree& t = c->tree;
    BigVec<int32_t> vni(n_node), vnj(n_node);
    for (int k = 0; k < n_node; k++) { vni[k] = num2[2 * k]; vnj[k] = num2[2 * k + 1]; }
    BigVec<std::array<double, 7>> U(n_node);
    for (int k = 0; k < n_node; k++)
        for (int i = 0; i < 7; i++) U[k][i] = prim7[7 * (size_t)k + i];
    t.m = m; t.n = n; t.total = total;
    t.father.assign(father, father + n_node);
    t.son.assign(son, son + n_node);
    t.lvl.assign(lvl, lvl + n_node);
    t.sp.assign(sp, sp + n_node);
    t.cross.assign(cross, cross + n_node);
    t.ni.swap(vni);
    t.nj.swap(vnj);
    t.x.assign(x, x + n_node);
    t.y.assign(y, y + n_node);
    t.U.swap(U);
    t.rot.assign(n_node, 0.0);
    t.div.assign(n_node, 0.0);
    t.fresh.assign(n_node, 0);
    t.free_blocks.clear();
    t.idneighbor();
    t.idgrid();
    return 0;
}

int32_t fdmamr_tree_size(fdmamr_ctx* c, int32_t* n_node) {
    if (!c || !n_node) return -1;
    c->tree.canonical_ids(n_node);
    return 0;
}

int32_t fdmamr_tree_export(fdmamr_ctx* c, int32_t* n_node, int32_t* father, int32_t* son,
                           int32_t* lvl, int32_t* sp, int32_t* num2, int32_t* cross, double* x,
                           double* y, double* prim7) {
    if (!c) return -1;
    if (c->tables_ready && !c->host_only) { int r = pull_state(c); if (r) return r; }
    const Tree& t = c->tree;
    int32_t nn = 0;
    std::vector<int32_t> id = t.canonical_ids(&nn);
    if (n_node) *n_node = nn;
    for (size_t k = 0; k < t.n_alloc(); k++) {
        int j = id[k];
        if (j < 0) continue;
        if (father) father[j] = t.father[k] >= 0 ? id[t.father[k]] : -1;
        if (son) son[j] = t.son[k] >= 0 ? id[t.son[k]] : -1;
        if (lvl) lvl[j] = t.lvl[k];
        if (sp) sp[j] = t.sp[k];
        if (num2) { num2[2 * j] = t.ni[k]; num2[2 * j + 1] = t.nj[k]; }
        if (cross) cross[j] = t.cross[k];
        if (x) x[j] = t.x[k];
        if (y) y[j] = t.y[k];
        if (prim7) for (int i = 0; i < 7; i++) prim7[7 * (size_t)j + i] = t.U[k][i];
    }
    return 0;
}

int32_t fdmamr_build_tables(fdmamr_ctx* c) {
    if (!c) return -1;
    if (c->tree.total == 0) return fail(c, -1, "no mesh");
    if (!c->host_only && hipSetDevice(c->p.device) != hipSuccess) return fail(c, -2, "hipSetDevice failed");
    PhaseTimer pt;
    try {
        c->tree.idgrid();
        pt.mark("IDgrid");
        if (tree_depth(c->tree) >= MAX_TREE_DEPTH - 1) return fail(c, -1, "quadtree deeper than supported");
        build_tables(c->tree, c->tb);
        pt.mark("build_tables");
    } catch (const std::exception& e) {
        return fail(c, -1, std::string("table build: ") + e.what());
    }
    if (c->host_only) {
        c->tables_ready = true;
        return 0;
    }
    int r = upload_tables(c);
    if (r) return r;
    pt.mark("upload");
    c->tables_ready = true;
    return 0;
}

int32_t fdmamr_counts(fdmamr_ctx* c, int32_t* n_leaf, int32_t* n_ghost, int32_t* n_slot_wls,
                      int32_t* n_slot_restrict, int32_t* l_max) {
    if (!c) return -1;
    if (n_leaf) *n_leaf = c->tb.N;
    if (n_ghost) *n_ghost = c->tb.G;
    if (n_slot_wls) *n_slot_wls = c->tb.S_w;
    if (n_slot_restrict) *n_slot_restrict = c->tb.S_r;
    if (l_max) *l_max = c->tb.l_max;
    return 0;
}

int32_t fdmamr_set_state(fdmamr_ctx* c, const double* prim5) {
    int r = check_ready(c);
    if (r) return r;
    if (!prim5) return fail(c, -1, "null state");
    const int N = c->tb.N;
    std::vector<double4> q(N);
    std::vector<double> tv(N);
    for (int i = 0; i < N; i++) {
        const double* u = prim5 + 5 * (size_t)c->tb.ns_pos[i];  // device order from cell_NS order
        q[i] = make_double4(u[0], u[1], u[2], u[4]);
        tv[i] = u[3];
    }
    HIPCHK(c, hipMemcpyAsync(c->Q[0].p, q.data(), N * sizeof(double4), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->T[0].p, tv.data(), N * 8, hipMemcpyHostToDevice, c->stream));
    // ghosts follow the new interior state (as after the reference's AMR, solution.F90:224-229)
    if (c->p.flux_precision == FDMAMR_FLUX_FP32) launch_shadow(c->Q[0].p, c->Qf[0].p, N, c->stream);
    fill_after_stage(c, 0);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

// state [off, off + n) of buffer 0 to host AoS (interior leaves: from device to cell_NS order)
static int get_range(fdmamr_ctx* c, int off, int n, double* prim5) {
    std::vector<double4> q(n);
    std::vector<double> tv(n);
    if (n > 0) {
        HIPCHK(c, hipMemcpyAsync(q.data(), c->Q[0].p + off, n * sizeof(double4), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipMemcpyAsync(tv.data(), c->T[0].p + off, n * 8, hipMemcpyDeviceToHost, c->stream));
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    const bool fast = c->p.flux_precision == FDMAMR_FLUX_FP32;
    const double gm = c->tree.gamma * c->tree.Ma * c->tree.Ma;
    const bool interior = off == 0;
    for (int i = 0; i < n; i++) {
        double* u = prim5 + 5 * (size_t)(interior ? c->tb.ns_pos[i] : i);
        u[0] = q[i].x; u[1] = q[i].y; u[2] = q[i].z; u[4] = q[i].w;
        u[3] = fast ? q[i].w * gm / q[i].x : tv[i];
    }
    return 0;
}

int32_t fdmamr_get_state(fdmamr_ctx* c, double* prim5) {
    int r = check_ready(c);
    if (r) return r;
    if (!prim5) return fail(c, -1, "null state");
    return get_range(c, 0, c->tb.N, prim5);
}

int32_t fdmamr_get_ghost_state(fdmamr_ctx* c, double* prim5) {
    int r = check_ready(c);
    if (r) return r;
    if (!prim5) return fail(c, -1, "null state");
    return get_range(c, c->tb.N, c->tb.G, prim5);
}

int32_t fdmamr_leaf_geometry(fdmamr_ctx* c, double* x, double* y, int32_t* lvl, int32_t* node) {
    if (!c) return -1;
    const Tree& t = c->tree;
    std::vector<int32_t> id;
    if (node) id = t.canonical_ids(nullptr);
    for (size_t i = 0; i < t.ns.size(); i++) {
        int k = t.ns[i];
        if (x) x[i] = t.x[k];
        if (y) y[i] = t.y[k];
        if (lvl) lvl[i] = t.lvl[k];
        if (node) node[i] = id[k];
    }
    return 0;
}

int32_t fdmamr_max_wavespeed(fdmamr_ctx* c, double* lam_x, double* lam_y, int32_t* l_max) {
    int r = check_ready(c);
    if (r) return r;
    const int nb = std::max(1, nblocks(c->tb.N, 256));
    if (c->tb.N > 0) launch_wavespeed(c->dt, c->Q[0].p, c->lampart.p, c->stream);
    HIPCHK(c, hipGetLastError());
    std::vector<double> part((size_t)nb * 2, 0.0);
    if (c->tb.N > 0)
        HIPCHK(c, hipMemcpyAsync(part.data(), c->lampart.p, part.size() * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    double lx = 0.0, ly = 0.0;
    for (int b = 0; b < nb; b++) { lx = std::max(lx, part[2 * b]); ly = std::max(ly, part[2 * b + 1]); }
    if (lam_x) *lam_x = lx;
    if (lam_y) *lam_y = ly;
    if (l_max) *l_max = c->tb.l_max;
    return 0;
}

int32_t fdmamr_max_wavespeed_noalloc(fdmamr_ctx* c, double* lam_x, double* lam_y, int32_t* l_max) {
    int r = check_ready(c);
    if (r) return r;
    if (c->tb.N > FDMAMR_NOALLOC_MAX_LEAVES)
        return fail(c, -1, "fdmamr_max_wavespeed_noalloc: " + std::to_string(c->tb.N) +
                               " leaves exceed its scratch (" + std::to_string(FDMAMR_NOALLOC_MAX_LEAVES) +
                               " at most); use fdmamr_max_wavespeed");
    double out[2] = {0.0, 0.0};
    if (c->tb.N > 0) {
        HIPCHK(c, launch_wavespeed_noalloc(c->dt, c->Q[0].p, c->p.device, out, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    if (lam_x) *lam_x = out[0];
    if (lam_y) *lam_y = out[1];
    if (l_max) *l_max = c->tb.l_max;
    return 0;
}

int32_t fdmamr_advance_rk3(fdmamr_ctx* c, double dt, double lam_x, double lam_y,
                           double* res_max4, double* res_rms_sum4) {
    int r = check_ready(c);
    if (r) return r;
    if (c->tb.N == 0) return 0;
    DevScalars s{};
    s.dt = dt; s.lam_x = lam_x; s.lam_y = lam_y; s.h = c->tree.h; s.xf = c->tree.xf;
    HIPCHK(c, hipMemcpyAsync(c->sc, &s, sizeof s, hipMemcpyHostToDevice, c->stream));
    enqueue_rk3(c, 0, false);
    finish_steps(c);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipMemcpyAsync(&s, c->sc, sizeof s, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (int k = 0; k < 4; k++) {
        if (res_max4) res_max4[k] = s.res_max[k];
        if (res_rms_sum4) res_rms_sum4[k] = s.res_sum[k];
    }
    if (s.pos_bad) return fail(c, -4, "non-positive density or pressure in a stage of the step");
    for (int k = 0; k < 4; k++)
        if (s.res_sum[k] != s.res_sum[k]) return fail(c, -3, "NaN residual");
    return 0;
}

int32_t fdmamr_run_steps(fdmamr_ctx* c, int32_t n_steps, int32_t localdt, double cfl,
                         double dt_fixed, double t_end, double* tt_inout, double* lam_inout2,
                         double* scalars8) {
    int r = check_ready(c);
    if (r) return r;
    if (n_steps < 0) return fail(c, -1, "negative step count");
    DevScalars s{};
    s.tt = tt_inout ? *tt_inout : 0.0;
    s.t_end = t_end; s.cfl = cfl; s.dt_fixed = dt_fixed; s.localdt = localdt;
    s.l_max = c->tb.l_max; s.h = c->tree.h; s.xf = c->tree.xf;
    s.lam_x = lam_inout2 ? lam_inout2[0] : 0.0;
    s.lam_y = lam_inout2 ? lam_inout2[1] : 0.0;
    HIPCHK(c, hipMemcpyAsync(c->sc, &s, sizeof s, hipMemcpyHostToDevice, c->stream));
    const int nb = std::max(1, nblocks(c->tb.N, 256));
    for (int it = 0; it < n_steps && c->tb.N > 0; it++) {
        if (it == 0) {
            launch_wavespeed(c->dt, c->Q[0].p, c->lampart.p, c->stream);
            launch_step_begin(c->lampart.p, nb, c->sc, c->stream);
        }
        enqueue_rk3(c, 1, it + 1 < n_steps);
    }
    if (n_steps > 0 && c->tb.N > 0) finish_steps(c);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipMemcpyAsync(&s, c->sc, sizeof s, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (tt_inout) *tt_inout = s.tt;
    if (lam_inout2) { lam_inout2[0] = s.lam_x; lam_inout2[1] = s.lam_y; }
    if (scalars8) {
        scalars8[0] = s.tt; scalars8[1] = s.dt; scalars8[2] = s.lam_x; scalars8[3] = s.lam_y;
        for (int k = 0; k < 4; k++) scalars8[4 + k] = s.res_rms[k];
    }
    if (s.status == -4) return fail(c, -4, "non-positive density or pressure in a stage of a step");
    if (s.status != 0) return fail(c, -3, "NaN residual");
    return 0;
}

int32_t fdmamr_check_positivity(fdmamr_ctx* c, int32_t* n_bad) {
    int r = check_ready(c);
    if (r) return r;
    HIPCHK(c, hipMemsetAsync(c->nbad.p, 0, sizeof(int), c->stream));
    if (c->tb.N > 0) launch_positivity(c->dt, c->Q[0].p, c->nbad.p, c->stream);
    int nb = 0;
    HIPCHK(c, hipMemcpyAsync(&nb, c->nbad.p, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (n_bad) *n_bad = nb;
    return nb ? fail(c, -4, "non-positive density or pressure") : 0;
}

int32_t fdmamr_sensors(fdmamr_ctx* c, double* rot, double* div) {
    int r = check_ready(c);
    if (r) return r;
    const int N = c->tb.N;
    launch_slots(c->dt, c->Q[0].p, c->T[0].p, c->stream);
    if (N > 0) launch_sensors(c->dt, c->Q[0].p, c->rot.p, c->div.p, c->stream);
    HIPCHK(c, hipGetLastError());
    BigVec<double> a(N), b(N);
    HIPCHK(c, download_staged(c, a.data(), c->rot.p, (size_t)N * 8));
    HIPCHK(c, download_staged(c, b.data(), c->div.p, (size_t)N * 8));
    Tree& t = c->tree;
    double f_of_lvl[64];
    for (int l = 0; l < 64; l++) f_of_lvl[l] = l < 31 ? std::pow(t.h / (double)(1 << l), 1.5) : 0.0;  // dy**1.5, solution.F90:565,578
    parallel_for((size_t)N, [&](size_t b0, size_t e0) {
        for (size_t i = b0; i < e0; i++) {
            const int k = c->tb.leaf_node[i];
            const double f = f_of_lvl[t.lvl[k]];
            t.rot[k] = a[i] * f;
            t.div[k] = b[i] * f;
            const int e = c->tb.ns_pos[i];                // outputs in cell_NS order
            if (rot) rot[e] = t.rot[k];
            if (div) div[e] = t.div[k];
        }
    });
    return 0;
}

// AMR body after rot_div (solution.F90:413-419): RMS of the sensors, solute_ref,
// solute_gridmodify, idneighbor, IDgrid + tables. The tree's rot/div hold the sensors.
static int regrid_from_sensors(fdmamr_ctx* c, int32_t lvl_amr, double refine_rot, double refine_div,
                               double coarse_rot, double coarse_div, int32_t sensor_mode) {
    PhaseTimer pt;
    Tree& t = c->tree;
    if (sensor_mode == 1)
        parallel_for(t.ns.size(), [&](size_t b, size_t e) {
            for (size_t i = b; i < e; i++) t.rot[t.ns[i]] = t.div[t.ns[i]];
        });
    // Srot, Sdiv, Ne in rot_div's visiting order (solution.F90:430-465)
    double srot = 0, sdiv = 0;
    int ne = 0;
    t.rot_div_sums(srot, sdiv, ne);
    double arot = std::sqrt(srot / ne), adiv = std::sqrt(sdiv / ne);
    try {
        t.solute_ref(arot, adiv, lvl_amr, refine_rot, refine_div, coarse_rot, coarse_div);
        pt.mark("solute_ref");
        t.solute_gridmodify(lvl_amr);
        pt.mark("solute_gridmodify");
        t.idneighbor();
        pt.mark("idneighbor");
    } catch (const std::exception& e) {
        return fail(c, -1, e.what());
    }
    // the upload refills the ghosts from the new interior state (solution.F90:224-229)
    return fdmamr_build_tables(c);
}

int32_t fdmamr_regrid(fdmamr_ctx* c, int32_t lvl_amr, double refine_rot, double refine_div,
                      double coarse_rot, double coarse_div, int32_t sensor_mode) {
    int r = check_ready(c);
    if (r) return r;
    PhaseTimer pt;
    r = pull_state(c);
    if (r) return r;
    pt.mark("state download");
    r = fdmamr_sensors(c, nullptr, nullptr);
    if (r) return r;
    pt.mark("sensors");
    return regrid_from_sensors(c, lvl_amr, refine_rot, refine_div, coarse_rot, coarse_div, sensor_mode);
}

int32_t fdmamr_regrid_sensors(fdmamr_ctx* c, int32_t lvl_amr, double refine_rot, double refine_div,
                              double coarse_rot, double coarse_div, int32_t sensor_mode, const double* rot,
                              const double* div) {
    if (!c) return -1;
    // a host-only context regrids its host tree (no state to download, no upload)
    if (!c->tables_ready) return fail(c, -1, "tables not built: call fdmamr_build_tables first");
    if (!rot || !div) return fail(c, -1, "fdmamr_regrid_sensors: null sensor array");
    if (!c->host_only) {
        int r = pull_state(c);
        if (r) return r;
    }
    Tree& t = c->tree;
    std::atomic<int> bad{0};
    parallel_for(t.ns.size(), [&](size_t b, size_t e) {
        for (size_t i = b; i < e; i++)
            if (!std::isfinite(rot[i]) || !std::isfinite(div[i])) bad = 1;
    });
    if (bad) return fail(c, -1, "fdmamr_regrid_sensors: sensor value not finite");
    parallel_for(t.ns.size(), [&](size_t b, size_t e) {
        for (size_t i = b; i < e; i++) {
            t.rot[t.ns[i]] = rot[i];
            t.div[t.ns[i]] = div[i];
        }
    });
    return regrid_from_sensors(c, lvl_amr, refine_rot, refine_div, coarse_rot, coarse_div, sensor_mode);
}

int32_t fdmamr_stencil_info(fdmamr_ctx* c, int32_t i, int32_t* kind8, int32_t* node8,
                            double* px8, double* py8, int32_t* m1_8, int32_t* m2_8,
                            int32_t* nb8x9) {
    if (!c || !c->tables_ready || i < 0 || i >= c->tb.N) return fail(c, -1, "fdmamr_stencil_info: bad leaf");
    std::vector<int32_t> id = c->tree.canonical_ids(nullptr);
    const Tables& tb = c->tb;
    const int dev = tb.dev_of_ns[i];
    for (int k = 0; k < 8; k++) {
        size_t d = (size_t)dev * 8 + k;
        if (kind8) kind8[k] = tb.st_kind[d];
        if (node8) node8[k] = id[tb.st_node[d]];
        double px, py;
        stencil_point(c->tree, tb.leaf_node[dev], k, tb.st_kind[d], tb.st_node[d], &px, &py);
        if (px8) px8[k] = px;
        if (py8) py8[k] = py;
        int32_t nb9[9], a = 0, b = 0;
        stencil_lists(c->tree, tb.st_node[d], tb.st_kind[d], nb9, &a, &b);
        if (m1_8) m1_8[k] = a;
        if (m2_8) m2_8[k] = b;
        if (nb8x9)
            for (int j = 0; j < 9; j++) nb8x9[k * 9 + j] = nb9[j] >= 0 ? id[nb9[j]] : -1;
    }
    return 0;
}

int32_t fdmamr_selector_set_flags(fdmamr_ctx* c, const uint8_t* flags, int32_t use) {
    int r = check_ready(c);
    if (r) return r;
    if (c->p.flux_precision != FDMAMR_FLUX_FP32) return fail(c, -1, "selector needs the FP32 flux mode");
    std::vector<uint8_t> fd(c->tb.N);
    if (flags && c->tb.N > 0) {
        for (int i = 0; i < c->tb.N; i++) fd[i] = flags[c->tb.ns_pos[i]];
        HIPCHK(c, hipMemcpyAsync(c->flags.p, fd.data(), c->tb.N, hipMemcpyHostToDevice, c->stream));
    }
    launch_build_lists(c->tb.N, c->flags.p, c->list_cnt.p, c->lists.p, c->list_meta.p, c->stream);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->use_flags = use != 0;
    return 0;
}

int32_t fdmamr_selector_get_flags(fdmamr_ctx* c, uint8_t* flags) {
    int r = check_ready(c);
    if (r) return r;
    if (!flags) return fail(c, -1, "null flags");
    std::vector<uint8_t> fd(c->tb.N);
    if (c->tb.N > 0)
        HIPCHK(c, hipMemcpyAsync(fd.data(), c->flags.p, c->tb.N, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (int i = 0; i < c->tb.N; i++) flags[c->tb.ns_pos[i]] = fd[i];
    return 0;
}

int32_t fdmamr_leaf_lists(fdmamr_ctx* c, int32_t* lists, int32_t* n_clear) {
    int r = check_ready(c);
    if (r) return r;
    if (!lists || !n_clear) return fail(c, -1, "null list buffer");
    // the lists as the last step built them, or from the current flags for a fresh context
    if (c->tb.N > 0) {
        HIPCHK(c, hipMemcpyAsync(lists, c->lists.p, (size_t)c->tb.N * 4, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipMemcpyAsync(n_clear, c->list_meta.p, 4, hipMemcpyDeviceToHost, c->stream));
    } else {
        *n_clear = 0;
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    // device leaf indices -> cell_NS positions, each list ascending
    const int nc = std::min(std::max(*n_clear, 0), c->tb.N);
    for (int k = 0; k < c->tb.N; k++) {
        if (lists[k] < 0 || lists[k] >= c->tb.N) return fail(c, -2, "fdmamr_leaf_lists: corrupt list entry");
        lists[k] = c->tb.ns_pos[lists[k]];
    }
    std::sort(lists, lists + nc);
    std::sort(lists + nc, lists + c->tb.N);
    return 0;
}

int32_t fdmamr_labels(fdmamr_ctx* c, float* dev) {
    int r = check_ready(c);
    if (r) return r;
    if (!dev) return fail(c, -1, "null label buffer");
    if (c->tb.N == 0) return 0;
    // one full-path FP32 substep into the scratch buffer 2 with label output; the state in
    // buffer 0 is untouched
    DevScalars s{};
    s.dt = 0.0; s.h = c->tree.h; s.xf = c->tree.xf;
    HIPCHK(c, hipMemcpyAsync(c->sc, &s, sizeof s, hipMemcpyHostToDevice, c->stream));
    launch_shadow(c->Q[0].p, c->Qf[0].p, c->tb.L + c->tb.S_w + c->tb.S_r, c->stream);
    launch_slots_fast(c->dt, c->wls_f.p, c->Q[0].p, c->Qf[0].p, c->stream);
    launch_stage_fast(c->dt, nullptr, nullptr, nullptr, c->Q[0].p, c->Qf[0].p, c->Q[0].p, c->Q[2].p, c->Qf[2].p,
                      c->sc, 0.0, 1.0, 0, c->partials.p, nullptr, c->label.p, JudgeArgs{}, nullptr, c->stream);
    HIPCHK(c, hipGetLastError());
    std::vector<float> lab(c->tb.N);
    HIPCHK(c, hipMemcpyAsync(lab.data(), c->label.p, c->tb.N * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (int i = 0; i < c->tb.N; i++) dev[c->tb.ns_pos[i]] = lab[i];
    return 0;
}

// ---------------------------------------------------------------- learned selector
static int sel_alloc(fdmamr_ctx* c) {
    HIPCHK(c, ensure(c->sel_w, P_TOTAL));
    HIPCHK(c, ensure(c->sel_m, P_TOTAL));
    HIPCHK(c, ensure(c->sel_v, P_TOTAL));
    HIPCHK(c, ensure(c->sel_grad, P_TOTAL));
    HIPCHK(c, ensure(c->sel_pk, KW_TOTAL));
    HIPCHK(c, ensure(c->sel_cnt, CNT_N * CNT_SLOTS));
    HIPCHK(c, ensure(c->sel_loss, 4));
    return 0;
}

static int sel_upload(fdmamr_ctx* c, const std::vector<float>& w) {
    int r = sel_alloc(c);
    if (r) return r;
    std::vector<float> z(P_TOTAL, 0.0f);
    HIPCHK(c, hipMemcpyAsync(c->sel_w.p, w.data(), P_TOTAL * 4, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->sel_m.p, z.data(), P_TOTAL * 4, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->sel_v.p, z.data(), P_TOTAL * 4, hipMemcpyHostToDevice, c->stream));
    const std::vector<int> zc(CNT_N * CNT_SLOTS, 0);
    HIPCHK(c, hipMemcpyAsync(c->sel_cnt.p, zc.data(), zc.size() * 4, hipMemcpyHostToDevice, c->stream));
    launch_pack(c->sel_w.p, c->sel_pk.p, c->stream);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->adam_t = 0;
    c->sel_ready = true;
    return 0;
}

int32_t fdmamr_selector_init(fdmamr_ctx* c, uint64_t seed) {
    if (!c || c->host_only) return fail(c, -5, "no device");
    std::vector<float> w(P_TOTAL, 0.0f);
    uint64_t s = seed;
    auto nrm = [&]() {  // Box-Muller from SplitMix64
        double u1 = ((splitmix64(s) >> 11) + 1.0) * 0x1.0p-53, u2 = (splitmix64(s) >> 11) * 0x1.0p-53;
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    };
    for (int i = 0; i < 75; i++)
        for (int o = 0; o < MLP_H; o++) w[P_W1 + i * MLP_H + o] = (float)(nrm() * std::sqrt(2.0 / 75));
    for (int k = 0; k < MLP_H * MLP_H; k++) {
        w[P_W2 + k] = (float)(nrm() * std::sqrt(2.0 / MLP_H));
        w[P_W3 + k] = (float)(nrm() * std::sqrt(2.0 / MLP_H));
    }
    for (int k = 0; k < MLP_H; k++) w[P_W4 + k] = (float)(nrm() * std::sqrt(1.0 / MLP_H));
    return sel_upload(c, w);
}

int32_t fdmamr_selector_set_weights(fdmamr_ctx* c, const float* w, int32_t n) {
    if (!c || c->host_only) return fail(c, -5, "no device");
    if (!w || n != P_TOTAL) return fail(c, -1, "selector weights: wrong size");
    std::vector<float> v(w, w + n);
    for (int i = 75; i < MLP_IN; i++)
        for (int o = 0; o < MLP_H; o++) v[P_W1 + i * MLP_H + o] = 0.0f;
    return sel_upload(c, v);
}

int32_t fdmamr_selector_get_weights(fdmamr_ctx* c, float* w, int32_t n) {
    if (!c || !c->sel_ready || !w || n != P_TOTAL) return fail(c, -1, "selector weights unavailable");
    HIPCHK(c, hipMemcpyAsync(w, c->sel_w.p, P_TOTAL * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

int32_t fdmamr_selector_config(fdmamr_ctx* c, float threshold, float judge_fraction, float label_bound) {
    if (!c || !(threshold > 0.f && threshold <= 1.f) || !(judge_fraction >= 0.f && judge_fraction <= 1.f) ||
        !(label_bound > 0.f))
        return fail(c, -1, "selector config out of range");
    c->sel_threshold = threshold;
    c->sel_judge = judge_fraction;
    c->sel_bound = label_bound;
    return 0;
}

int32_t fdmamr_selector_infer(fdmamr_ctx* c, float* prob, int32_t* n_troubled) {
    int r = check_ready(c);
    if (r) return r;
    if (!c->sel_ready) return fail(c, -1, "selector has no weights");
    const int N = c->tb.N;
    launch_shadow(c->Q[0].p, c->Qf[0].p, c->tb.L + c->tb.S_w + c->tb.S_r, c->stream);
    launch_slots_fast(c->dt, c->wls_f.p, c->Q[0].p, c->Qf[0].p, c->stream);
    std::vector<int> cnt0(CNT_N * CNT_SLOTS, 0);
    HIPCHK(c, hipMemcpyAsync(c->sel_cnt.p, cnt0.data(), cnt0.size() * 4, hipMemcpyHostToDevice, c->stream));
    SelectParams sp{c->sel_threshold, 0u, c->sel_step};  // no judge sample here
    launch_select(c->dt, c->Qf[0].p, c->off16.p, c->sel_pk.p, c->sel_w.p, sp, c->flags.p, c->sel_prob.p, c->n_cu,
                  c->sel_nbr_beg.p, c->sel_nbr.p, c->sel_loc.p, c->sel_cnt.p, c->stream);
    // no judge sample was drawn: the judge pass only counts the troubled leaves
    launch_judge(N, c->label.p, c->sel_bound, c->flags.p, c->sel_cnt.p, c->stream);
    launch_build_lists(N, c->flags.p, c->list_cnt.p, c->lists.p, c->list_meta.p, c->stream);
    HIPCHK(c, hipGetLastError());
    std::vector<float> pd(N);
    if (prob && N > 0) HIPCHK(c, hipMemcpyAsync(pd.data(), c->sel_prob.p, N * 4, hipMemcpyDeviceToHost, c->stream));
    int cnt[CNT_N];
    HIPCHK(c, read_counters(c, cnt));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (prob)
        for (int i = 0; i < N; i++) prob[c->tb.ns_pos[i]] = pd[i];
    if (n_troubled) *n_troubled = cnt[CNT_TROUBLED];
    c->use_flags = true;
    return 0;
}

int32_t fdmamr_selector_features(fdmamr_ctx* c, uint16_t* x) {
    int r = check_ready(c);
    if (r) return r;
    if (!x) return fail(c, -1, "null feature buffer");
    const int N = c->tb.N;
    DBuf<uint16_t> tmp;
    HIPCHK(c, ensure(tmp, (size_t)N * MLP_IN));
    launch_shadow(c->Q[0].p, c->Qf[0].p, c->tb.L + c->tb.S_w + c->tb.S_r, c->stream);
    launch_slots_fast(c->dt, c->wls_f.p, c->Q[0].p, c->Qf[0].p, c->stream);
    launch_features(c->dt, c->Qf[0].p, c->off16.p, tmp.p, 0, c->stream);
    hipError_t e = hipGetLastError();
    std::vector<uint16_t> rows((size_t)N * MLP_IN);
    if (e == hipSuccess && N > 0)
        e = hipMemcpyAsync(rows.data(), tmp.p, (size_t)N * MLP_IN * 2, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    release(tmp);
    if (e != hipSuccess) return fail(c, -2, hipGetErrorString(e));
    for (int i = 0; i < N; i++)
        std::memcpy(x + (size_t)c->tb.ns_pos[i] * MLP_IN, rows.data() + (size_t)i * MLP_IN, MLP_IN * 2);
    return 0;
}

int32_t fdmamr_selector_clear_dataset(fdmamr_ctx* c) {
    if (!c) return -1;
    c->ds_rows = 0;
    return 0;
}

int32_t fdmamr_selector_collect(fdmamr_ctx* c, int32_t* n_rows, int32_t* n_positive) {
    int r = check_ready(c);
    if (r) return r;
    const int N = c->tb.N;
    const size_t need = (size_t)c->ds_rows + N;
    // rows are numbered in int32 (element offsets are 64-bit)
    if (need > 0x7fffffff) return fail(c, -1, "training set too large");
    if (c->ds_x.cap < need * MLP_IN) {
        // grow by half again, keeping the rows collected so far
        DBuf<uint16_t> nx;
        DBuf<uint8_t> ny;
        HIPCHK(c, ensure(nx, need * MLP_IN + need * MLP_IN / 2));
        HIPCHK(c, ensure(ny, need + need / 2));
        if (c->ds_rows) {
            HIPCHK(c, hipMemcpyAsync(nx.p, c->ds_x.p, (size_t)c->ds_rows * MLP_IN * 2, hipMemcpyDeviceToDevice, c->stream));
            HIPCHK(c, hipMemcpyAsync(ny.p, c->ds_y.p, (size_t)c->ds_rows, hipMemcpyDeviceToDevice, c->stream));
        }
        HIPCHK(c, hipStreamSynchronize(c->stream));
        release(c->ds_x);
        release(c->ds_y);
        c->ds_x = nx;
        c->ds_y = ny;
    }
    // labels: one full-path FP32 substep with label output (state unchanged)
    std::vector<float> dummy;
    DevScalars s{};
    s.h = c->tree.h; s.xf = c->tree.xf;
    HIPCHK(c, hipMemcpyAsync(c->sc, &s, sizeof s, hipMemcpyHostToDevice, c->stream));
    launch_shadow(c->Q[0].p, c->Qf[0].p, c->tb.L + c->tb.S_w + c->tb.S_r, c->stream);
    launch_slots_fast(c->dt, c->wls_f.p, c->Q[0].p, c->Qf[0].p, c->stream);
    launch_stage_fast(c->dt, nullptr, nullptr, nullptr, c->Q[0].p, c->Qf[0].p, c->Q[0].p, c->Q[2].p, c->Qf[2].p,
                      c->sc, 0.0, 1.0, 0, c->partials.p, nullptr, c->label.p, JudgeArgs{}, nullptr, c->stream);
    launch_features(c->dt, c->Qf[0].p, c->off16.p, c->ds_x.p, c->ds_rows, c->stream);
    launch_labels_to_y(N, c->label.p, c->sel_bound, c->ds_y.p, c->ds_rows, c->stream);
    HIPCHK(c, hipGetLastError());
    std::vector<uint8_t> yy(N);
    if (N > 0) HIPCHK(c, hipMemcpyAsync(yy.data(), c->ds_y.p + c->ds_rows, N, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->ds_rows += N;
    int pos = 0;
    for (uint8_t v : yy) pos += v;
    if (n_rows) *n_rows = c->ds_rows;
    if (n_positive) *n_positive = pos;
    return 0;
}

static int sel_grad_rows(fdmamr_ctx* c, const uint16_t* X, const uint8_t* y, const int32_t* rows, int n,
                         float pos_weight) {
    const int ntiles = (n + 15) / 16;
    const int nb = std::max(1, std::min(ntiles, 512));
    HIPCHK(c, ensure(c->sel_gpart, (size_t)nb * P_TOTAL));
    HIPCHK(c, ensure(c->sel_lpart, nb));
    launch_mlp_grad(c->sel_pk.p, c->sel_w.p, X, y, rows, n, pos_weight, 1.0f / (float)std::max(n, 1),
                    c->sel_gpart.p, c->sel_lpart.p, nb, c->sel_grad.p, c->sel_loss.p, c->stream);
    HIPCHK(c, hipGetLastError());
    return 0;
}

int32_t fdmamr_selector_train(fdmamr_ctx* c, int32_t iters, int32_t batch, float lr, float pos_weight,
                              uint64_t seed, float* loss) {
    if (!c || c->host_only) return fail(c, -5, "no device");
    if (!c->sel_ready) return fail(c, -1, "selector has no weights");
    if (c->ds_rows == 0 || iters < 0 || batch <= 0) return fail(c, -1, "nothing to train on");
    HIPCHK(c, ensure(c->ds_idx, batch));
    std::vector<int32_t> idx(batch);
    uint64_t s = seed;
    float last = 0.f;
    for (int it = 0; it < iters; it++) {
        for (int b = 0; b < batch; b++) idx[b] = (int32_t)(splitmix64(s) % (uint64_t)c->ds_rows);
        HIPCHK(c, hipMemcpyAsync(c->ds_idx.p, idx.data(), batch * 4, hipMemcpyHostToDevice, c->stream));
        int r = sel_grad_rows(c, c->ds_x.p, c->ds_y.p, c->ds_idx.p, batch, pos_weight);
        if (r) return r;
        c->adam_t++;
        const float b1 = 0.9f, b2 = 0.999f;
        launch_adam(c->sel_w.p, c->sel_grad.p, c->sel_m.p, c->sel_v.p, lr, b1, b2, 1e-8f,
                    1.0f - std::pow(b1, (float)c->adam_t), 1.0f - std::pow(b2, (float)c->adam_t), c->stream);
        launch_pack(c->sel_w.p, c->sel_pk.p, c->stream);
        HIPCHK(c, hipGetLastError());
        // the host-side index buffer is rewritten next iteration: wait for this one's copy
        HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    if (iters > 0) {
        HIPCHK(c, hipMemcpyAsync(&last, c->sel_loss.p, 4, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    if (loss) *loss = last / (float)batch;
    return 0;
}

int32_t fdmamr_selector_grad(fdmamr_ctx* c, const uint16_t* x, const uint8_t* y, int32_t n, float pos_weight,
                             float* grad, float* loss) {
    if (!c || c->host_only) return fail(c, -5, "no device");
    if (!c->sel_ready || !x || !y || n <= 0) return fail(c, -1, "selector grad: bad arguments");
    DBuf<uint16_t> dx;
    DBuf<uint8_t> dy;
    HIPCHK(c, ensure(dx, (size_t)n * MLP_IN));
    HIPCHK(c, ensure(dy, n));
    HIPCHK(c, hipMemcpyAsync(dx.p, x, (size_t)n * MLP_IN * 2, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(dy.p, y, n, hipMemcpyHostToDevice, c->stream));
    int r = sel_grad_rows(c, dx.p, dy.p, nullptr, n, pos_weight);
    float l = 0.f;
    if (!r) {
        if (grad) HIPCHK(c, hipMemcpyAsync(grad, c->sel_grad.p, P_TOTAL * 4, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipMemcpyAsync(&l, c->sel_loss.p, 4, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    release(dx);
    release(dy);
    if (loss) *loss = l / (float)n;
    return r;
}

int32_t fdmamr_selector_stats(fdmamr_ctx* c, int32_t* troubled, int32_t* judged, int32_t* rejected,
                              int32_t reset) {
    if (!c || !c->sel_ready) return fail(c, -1, "selector has no weights");
    int cnt[CNT_N];
    HIPCHK(c, read_counters(c, cnt));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (troubled) *troubled = cnt[CNT_TROUBLED];
    if (judged) *judged = cnt[CNT_JUDGED];
    if (rejected) *rejected = cnt[CNT_REJECTED];
    if (reset) {
        std::vector<int> z(CNT_N * CNT_SLOTS, 0);
        HIPCHK(c, hipMemcpyAsync(c->sel_cnt.p, z.data(), z.size() * 4, hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    return 0;
}

int32_t fdmamr_selector_judge_counter(fdmamr_ctx* c, int64_t set_to, uint32_t* current) {
    if (!c) return -1;
    if (set_to > (int64_t)UINT32_MAX) return fail(c, -1, "fdmamr_selector_judge_counter: value out of range");
    if (current) *current = c->sel_step;
    if (set_to >= 0) c->sel_step = (uint32_t)set_to;
    return 0;
}

int32_t fdmamr_profile(fdmamr_ctx* c, int32_t enable) {
    if (!c || c->host_only) return -1;
    c->profiling = enable != 0;
    c->ev_stage.clear();
    c->ev_select.clear();
    c->ev_used = 0;
    return 0;
}

int32_t fdmamr_profile_read(fdmamr_ctx* c, double* stage_ms_total, int32_t* stage_launches) {
    if (!c || c->host_only) return -1;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    double tot = 0.0;
    for (auto& pr : c->ev_stage) {
        float ms = 0.f;
        HIPCHK(c, hipEventElapsedTime(&ms, c->ev_pool[pr.first], c->ev_pool[pr.second]));
        tot += ms;
    }
    if (stage_ms_total) *stage_ms_total = tot;
    if (stage_launches) *stage_launches = (int32_t)c->ev_stage.size();
    c->ev_stage.clear();
    if (c->ev_select.empty()) c->ev_used = 0;
    return 0;
}

int32_t fdmamr_profile_read_select(fdmamr_ctx* c, double* ms_total, int32_t* launches, int64_t* network_tiles) {
    if (!c || c->host_only) return -1;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    double tot = 0.0;
    for (auto& pr : c->ev_select) {
        float ms = 0.f;
        HIPCHK(c, hipEventElapsedTime(&ms, c->ev_pool[pr.first], c->ev_pool[pr.second]));
        tot += ms;
    }
    if (ms_total) *ms_total = tot;
    if (launches) *launches = (int32_t)c->ev_select.size();
    if (network_tiles) {
        *network_tiles = 0;
        if (c->sel_cnt.p) {
            int cnt[CNT_N];
            HIPCHK(c, read_counters(c, cnt));
            HIPCHK(c, hipStreamSynchronize(c->stream));
            *network_tiles = cnt[CNT_TILES];
        }
    }
    c->ev_select.clear();
    if (c->ev_stage.empty()) c->ev_used = 0;
    return 0;
}

int32_t fdmamr_stencil_values_fast(fdmamr_ctx* c, int32_t i, float* val8x4) {
    int r = check_ready(c);
    if (r) return r;
    if (c->p.flux_precision != FDMAMR_FLUX_FP32) return fail(c, -1, "fdmamr_stencil_values_fast: FP32 contexts only");
    if (i < 0 || i >= c->tb.N || !val8x4) return fail(c, -1, "fdmamr_stencil_values_fast: bad leaf");
    // the slot and ghost work that precedes a stage, on state 0 (its values are unchanged)
    if (slots_merged(c))
        launch_slots_fill(c->dt, c->wls_f.p, c->Q[0].p, c->Qf[0].p, c->fill_fast.p, (int)c->tb.fill_fast.size(),
                          c->fill_beg.p, c->fill_leaf.p, c->fill_w.p, c->stream);
    else
        launch_slots_fast(c->dt, c->wls_f.p, c->Q[0].p, c->Qf[0].p, c->stream);
    launch_stencil_values_f(c->dt, c->Qf[0].p, c->tb.dev_of_ns[i], reinterpret_cast<float*>(c->tmp.p), c->stream);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipMemcpyAsync(val8x4, c->tmp.p, 32 * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

int32_t fdmamr_set_face_share(fdmamr_ctx* c, int32_t enable) {
    if (!c) return -1;
    c->share_off = enable == 0;
    return 0;
}

int32_t fdmamr_set_slot_merge(fdmamr_ctx* c, int32_t enable) {
    if (!c) return -1;
    c->merge_off = enable == 0;
    return 0;
}

int32_t fdmamr_table_info(fdmamr_ctx* c, int32_t* info8) {  // 11 entries
    if (!c || !info8) return -1;
    if (!c->tables_ready) return fail(c, -2, "fdmamr_table_info: tables not built");
    const Tables& tb = c->tb;
    int aliased = 0, generic = 0;
    for (int32_t a : tb.ghost_alias) aliased += a >= 0;
    for (size_t d = 0; d * 4 < tb.dep_hdr.size(); d++) generic += tb.dep_hdr[d * 4] == DEP_GENERIC;
    // selector tiles gathered directly (too many distinct values), tiles with more than 64,
    // and the largest count of a tile staged in LDS
    int direct = 0, over64 = 0, max_nu = 0;
    for (size_t k = 0; k + 1 < tb.sel_nbr_beg.size(); k++) {
        const int nu = tb.sel_nbr_beg[k + 1] - tb.sel_nbr_beg[k];
        if (nu > SEL_NBR_CAP) direct++;
        else max_nu = std::max(max_nu, nu);
        over64 += nu > 64;
    }
    int shx = 0, shy = 0;
    for (uint8_t f : tb.face_share) { shx += f & 1; shy += (f >> 1) & 1; }
    const int32_t v[11] = {tb.S_wp, tb.S_w, aliased, (int32_t)tb.fill_fast.size(), generic, direct, over64, max_nu,
                           shx, shy, tb.st16_escaped};
    for (int k = 0; k < 11; k++) info8[k] = v[k];
    return 0;
}

int32_t fdmamr_stencil_values(fdmamr_ctx* c, int32_t i, double* val8x4) {
    int r = check_ready(c);
    if (r) return r;
    if (i < 0 || i >= c->tb.N || !val8x4) return fail(c, -1, "fdmamr_stencil_values: bad leaf");
    launch_slots(c->dt, c->Q[0].p, c->T[0].p, c->stream);
    launch_stencil_values(c->dt, c->Q[0].p, c->tb.dev_of_ns[i], c->tmp.p, c->stream);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipMemcpyAsync(val8x4, c->tmp.p, 32 * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

}  // extern "C"
