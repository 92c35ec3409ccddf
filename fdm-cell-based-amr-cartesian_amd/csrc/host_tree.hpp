// host_tree.hpp -- host-side quadtree of the MI355X solver library.
//
// The reference keeps a pointer-linked quadtree (type(grid), module.F90:24-35). Here the
// tree is a structure of arrays indexed by node id; the four sons of a node occupy four
// consecutive ids (son1..son4 = SW, SE, NE, NW, mesh.F90:204-248). The tree lives on the
// host and changes only at regrid time; the per-step hot path runs on flattened device
// tables built from it (host_tables.cpp).
#pragma once
#include <array>
#include <cstdint>
#include <functional>
#include <atomic>
#include <memory>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

namespace fdm {

struct PassPool;  // node allocation of a parallel pass over the base cells (host_tree.cpp)

// Allocator whose resize leaves new elements uninitialised, for the node, per-leaf and
// per-entry arrays whose every element is written before it is read: growing them costs
// nothing, and fresh memory is first touched by the host threads that fill it.
// Large blocks are 2 MiB aligned and advised for transparent huge pages, which cuts the page
// faults of first touch and the TLB misses of the regrid's scattered reads.
void* big_alloc(size_t bytes);
void big_free(void* p);
constexpr size_t BIG_ALLOC_MIN = (size_t)4 << 20;
template <class T>
struct NoInit : std::allocator<T> {
    template <class U>
    struct rebind { using other = NoInit<U>; };
    NoInit() = default;
    template <class U>
    NoInit(const NoInit<U>&) {}
    T* allocate(size_t n) {
        if (n * sizeof(T) < BIG_ALLOC_MIN) return std::allocator<T>::allocate(n);
        return static_cast<T*>(big_alloc(n * sizeof(T)));
    }
    void deallocate(T* p, size_t n) {
        if (n * sizeof(T) < BIG_ALLOC_MIN) std::allocator<T>::deallocate(p, n);
        else big_free(p);
    }
    template <class U>
    void construct(U* p) noexcept(std::is_nothrow_default_constructible<U>::value) { ::new ((void*)p) U; }
    template <class U, class... A>
    void construct(U* p, A&&... a) { ::new ((void*)p) U(std::forward<A>(a)...); }
};
template <class T>
using BigVec = std::vector<T, NoInit<T>>;


enum Nb { NB_UP = 0, NB_DOWN, NB_RIGHT, NB_LEFT, NB_UPLEFT, NB_UPRIGHT, NB_DOWNLEFT, NB_DOWNRIGHT };

struct Tree {
    // ---- configuration (set_parameter, MAIN.F90:29-71) ----
    double h = 0.5, xf = 1.0;
    double Xmin = 0, Xmax = 0, Ymin = 0, Ymax = 0;  // ghost-extended (MAIN.F90:55-56)
    double pxmin = 0, pxmax = 0, pymin = 0, pymax = 0;  // physical domain (MAIN.F90:35-36)
    double gamma = 1.4, Ma = 1.0, cv = 0;
    int bc[4] = {6, 6, 6, 6};                       // xm, xp, ym, yp
    int wls_local = 0;                              // WLS-WENO frame (see fdmamr_params)
    double inflow[4] = {0, 0, 0, 0};
    int m = 0, n = 0, total = 0;

    // ---- node arrays ----
    BigVec<int32_t> father, son, lvl, sp, ni, nj, cross;
    BigVec<double> x, y;
    BigVec<std::array<double, 7>> U;                // U(1:7) of every node
    BigVec<double> rot, div;
    // 1 for the nodes the current regrid made leaves (refined sons, coarsened parents): their
    // values are copies or means, not evolved states
    BigVec<uint8_t> fresh;
    BigVec<int32_t> nb;                             // 8 cached neighbours per node (idnei)
    std::vector<int32_t> free_blocks;               // first ids of released son blocks

    // ---- leaf lists (IDgrid) ----
    std::vector<int32_t> ns, bcl;                   // cell_NS, cell_BC
    std::vector<int32_t> ns_beg;                    // cell_NS position of each base cell's first leaf

    void configure(double xmin, double xmax, double ymin, double ymax, double h_, double gamma_,
                   double Ma_);
    void creat_mesh();                              // mesh.F90:9-47
    int gridcross(int ni_, int nj_) const;          // mesh.F90:89-142
    bool is_leaf(int c) const { return son[c] < 0; }
    size_t n_alloc() const { return father.size(); }

    void refine(int c);                             // mesh.F90:199-260
    void coarse(int c);                             // mesh.F90:397-408
    void release_subtree(int c);
    // Renumber the nodes canonically (base cells, then son blocks in block pre-order, as
    // canonical_ids) and drop released blocks; the cached neighbours must be rebuilt after.
    void compact();
    void resize_nodes(size_t nn);
    PassPool* pool = nullptr;                       // set while a parallel pass runs
    std::shared_ptr<unsigned char[]> spare;         // compact()'s staging buffer, kept
    size_t spare_bytes = 0;
    // per base cell: bounds on the levels of the leaves below it (a lower bound on the least,
    // an upper bound on the greatest), kept by refine and coarse while track_levels is set,
    // which also mark the cell in cell_dirty
    std::vector<int8_t> cell_lmin, cell_lmax;
    std::vector<uint8_t> cell_dirty;
    std::vector<int32_t> cell_mod;                  // the last balance phase that changed the cell
    int32_t mod_phase = 0;                          // the balance phase running
    bool track_levels = false;
    std::atomic<long>* mod_count = nullptr;         // refines + coarsens, counted when set
    void leaf_levels();                             // exact bounds of every base cell
    void leaf_levels_dirty();                       // exact bounds of the marked cells; clears the marks

    // neighbour finders (neighbor.F90:6-396); -1 = null
    int init_up(int c) const;
    int init_down(int c) const;
    int init_left(int c) const;
    int init_right(int c) const;
    int up(int c) const;
    int down(int c) const;
    int left(int c) const;
    int right(int c) const;
    int upleft(int c) const;
    int upright(int c) const;
    int downright(int c) const;
    int downleft(int c) const;
    void idneighbor();                              // neighbor.F90:399-449
    void idgrid();                                  // solution.F90:593-693

    void restrict_mean(int c, double* var7) const;  // interpA, XuanGua.F90:359-398
    int bc_partner(int g) const;                    // BC_cyc partner after the son-path descent
    int bc_kind_of(int cross_) const;

    // canonical numbering: base cells, then son blocks in a block pre-order walk
    std::vector<int32_t> canonical_ids(int32_t* n_out) const;

    // ---- regrid ----
    // solute_ref (solution.F90:468-535) on per-node rot/div; returns nothing
    void solute_ref(double arot, double adiv, int lvl_amr, double rg_rot, double rg_div,
                    double cg_rot, double cg_div);
    void gridmodify1(int c);                        // mesh.F90:305-353
    void gridmodify4(int c);                        // mesh.F90:355-395
    void solute_gridmodify(int lvl_amr);            // solution.F90:537-556
    // Srot, Sdiv and Ne of rot_div: the sums of the squared sensors over the interior leaves,
    // added in the order rot_div visits them (solution.F90:430-465)
    void rot_div_sums(double& srot, double& sdiv, int& ne) const;

    // ---- initial conditions (host, driver support) ----
    void initial(int ic_kind, uint64_t seed, const double* ic_param8);
    void boundary_fill_host();                      // initial_boundary (solution.F90:71-107)

   private:
    int alloc_block();
    void ns_walk(int c, bool fill, int& cnt);
    void bc_walk(int c, bool fill, int& cnt);
    void idnei_faces(int c);
    void idnei_corners(int c);
    void coarse_sub(int c, double arot, double adiv, int lvl_amr, double cg_rot, double cg_div);
    void ref_sub(int c, double arot, double adiv, int lvl_amr, double rg_rot, double rg_div);
    void rot_div_walk(int c, std::vector<int32_t>& out) const;
};

// SplitMix64 shared by the IC generators
uint64_t splitmix64(uint64_t& s);

// Run f(begin, end) over [0, n) split across host threads (regrid-time work only); ranges
// shorter than serial_below run on the calling thread.
void parallel_for(size_t n, const std::function<void(size_t, size_t)>& f, size_t serial_below = 4096);

// Wall time of the phases of a regrid / table build, printed to stderr when the environment
// variable FDMAMR_TIMING is set (diagnostics).
struct PhaseTimer {
    PhaseTimer();
    void mark(const char* phase);
    bool on;
    double t0;
};

}  // namespace fdm
