// Kernel launch interface shared by the host orchestration (ctx.hip) and the
// kernel translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pic_device.h"

namespace pic {

template <typename R>
struct PView {
    R *x, *y, *vx, *vy, *vz;
    uint32_t *id;
    int *cell_start;
    int cap;
};

template <typename R>
struct NodeConst {
    int Gsize, ngx, ngy, nsp, CondNUMR;
    R hdx, hdy, idx, idy, idxy, hxy, hyx;
    R q_density[5];
    const int *Boundary, *Face, *CondID, *A_idx;
    const R *Area, *eps;
    unsigned long long *sigma_hits;   // null when no cell border leads into a dielectric
};

// One 64-lane wavefront per cell: the wave's cell (named by its lower-left
// node), or -1 past the last cell; bx is the block's index in its grid.
__device__ inline int wave_cell_b(int ncy, int ngy, int Csize, int &lane, int bx) {
    lane = threadIdx.x & 63;
    // wave-uniform: computed in scalar registers
    const int w = __builtin_amdgcn_readfirstlane((int)((bx * blockDim.x + threadIdx.x) >> 6));
    if (w >= Csize) return -1;
    return (w / ncy) * ngy + (w % ncy);
}
__device__ inline int wave_cell(int ncy, int ngy, int Csize, int &lane) {
    return wave_cell_b(ncy, ngy, Csize, lane, (int)blockIdx.x);
}

// [k0, k1) of cell c, clamped to the array capacity: after a capacity
// overflow (reported through the error flag) the offsets may point past the
// arrays, and no kernel may follow them there
template <typename R>
__device__ inline void cell_range(const PView<R> &p, int c, int &k0, int &k1) {
    k0 = min(p.cell_start[c], p.cap);
    k1 = min(p.cell_start[c + 1], p.cap);
    if (k1 < k0) k1 = k0;
}

__device__ inline unsigned long long lanes_below(int lane) {
    return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

// particle views of all species of a case (kernel argument)
template <typename R>
struct PSet {
    PView<R> s[5];
};

// Device-side clock of the fused step: the step number and time of the step
// about to run.  Kept in device memory so that a captured step graph needs no
// per-step kernel arguments; the deposit at the end of each step advances it.
struct StepClock {
    uint32_t step;
    uint32_t pad;
    double t;
};

// fused-step extras of the node deposit kernel: FP64 profile sums, the
// particle id counters (advanced by the step's product count) and the clock
template <typename R>
struct DepositExtras {
    int on;
    double *sden, *sphi, *sex, *sey;
    const R *tot, *Ex, *Ey;
    uint32_t *next_id;
    const int *prod_total_end;
    StepClock *clk;
    double dt;
};

constexpr int MAX_SRC = 16;
constexpr int EEDF_BINS = 256;
struct VoltConst {
    int ncond, nsrc;
    int flag[MAX_SRC], mid[MAX_SRC];
    double dc[MAX_SRC], ac[MAX_SRC], w[MAX_SRC], ph[MAX_SRC];   // values already rounded to the path precision
};

template <typename R> void launch_push(const DevConst<R> &, const PView<R> &, int, uint8_t *, hipStream_t);
// every species of a staged step, for the all-species launches of a whole
// step (push + count, sort scatter)
template <typename R>
struct StagedSet {
    PView<R> p[5], q[5];
    uint8_t *tag[5];
    int *key[5], *cnt[5];
    const int *base[5];
};
template <typename R>
void launch_push_count(const DevConst<R> &, const StagedSet<R> &, int nsp, int *condcharge, hipStream_t);
template <typename R>
void launch_sort_scatter_all(const DevConst<R> &, const StagedSet<R> &, int nsp, int *err, hipStream_t);
template <typename R>
void launch_sort_count(const DevConst<R> &, const PView<R> &, int, const uint8_t *, int *, int *, int *, hipStream_t);
template <typename R>
void launch_sort_scatter(const DevConst<R> &, const PView<R> &, const PView<R> &, const uint8_t *, const int *,
                         const int *, int, int *, hipStream_t);
template <typename R>
void launch_mcc_collide(const DevConst<R> &, const PView<R> &, int, uint32_t, unsigned long long *, int *, int *, R *,
                        hipStream_t);
template <typename R>
void launch_mcc_insert(const DevConst<R> &, const PView<R> &, const PView<R> &, const int *, const int *,
                       const PView<R> &, const int *, const R *, int, const uint32_t *, int, int *, hipStream_t);
// fused step: count pass (push, walls, MCC; collided particles written back
// in place, flag byte per particle, ionisation products in stash, 6 arrays
// of stash_stride) and scatter pass (sorted write of survivors and products)
template <typename R>
void launch_fused_count(const DevConst<R> &, const PView<R> &, int isp, const StepClock *clk, int *cnt, int *pcnt,
                        unsigned long long *mccr, int *condcharge, uint8_t *flag, R *stash, size_t stash_stride,
                        int *err, hipStream_t);
template <typename R>
void launch_fused_scatter(const DevConst<R> &, const PView<R> &p, const PView<R> &q, const PView<R> &qion, int isp,
                          const uint8_t *flag, const R *stash, size_t stash_stride, const int *base,
                          const int *pbase_e, const int *pbase_i, const int *pid_off, const uint32_t *next_id, int cap,
                          int cap_ion, int *err, hipStream_t);
template <typename R>
void launch_deposit(const DevConst<R> &, const NodeConst<R> &, const PView<R> *, R *part, R *den, R *sigma_sp,
                    R *Sigma, R *Source, R *b, const DepositExtras<R> *ex, hipStream_t);
// fused step: phi from the FP64 solution, electrode voltages and total potential in one kernel
template <typename R>
void launch_phi_potential(const NodeConst<R> &n, const VoltConst &v, const StepClock *clk, const double *x,
                          const R *lap, R *V, R *phi, R *tot, hipStream_t s);
// electron energy histogram (EEDF_BINS bins over [0, emax) eV) added into hist
template <typename R>
void launch_eedf(const PView<R> &electrons, int Gsize, R escale, R emax, unsigned long long *hist, hipStream_t s);
template <typename R>
void launch_efield_only(const NodeConst<R> &n, const R *phi, const R *tot, const R *den, R *Sigma, R *Ex, R *Ey,
                        hipStream_t s);
template <typename R>
void launch_efield(const NodeConst<R> &, const VoltConst &, double, const R *, R *, const R *, R *, const R *, R *, R *,
                   R *, hipStream_t);
template <typename R>
void launch_accumulate(int, int, const R *, const R *, const R *, const R *, double *, double *, double *, double *,
                       hipStream_t);
void launch_totals(int, const int *, const int *, const int *, const int *, int *, int *, int *, hipStream_t);
void launch_assign_bases(int, const int *, const int *, const int *, const int *, const int *, const int *, const int *,
                         int *, int *, int *, hipStream_t);
void launch_scan(const int *in, int *out, int n, int *tmp, hipStream_t s);
int scan_tmp_ints(int n);
void launch_scan_multi(int k, const int *const *in, int *const *out, int n, int *tmp, hipStream_t s);
// the staged sort's per-species arrays, for its all-species launches
constexpr int SORT_MAXSP = 8;
struct SortSpecies {
    const int *cnt[SORT_MAXSP];
    int *total[SORT_MAXSP];
    const int *start[SORT_MAXSP];
    int *base[SORT_MAXSP];
};
void launch_totals_sp(int Gsize, int nsp, const int *inv_start, const int *inv_ent, const SortSpecies &j,
                      hipStream_t s);
void launch_assign_bases_sp(int Gsize, int nsp, const int *inv_start, const int *inv_ent, const SortSpecies &j,
                            hipStream_t s);
// fused step, between the passes: per-destination totals of both species,
// their scans (cell_start of both species, product order) and the base of
// every (source cell, slot) entry; tmp holds scan_tmp_ints(G) ints
void launch_bases_pair(int G, const int *inv_start, const int *inv_ent, const int *cnt_e, const int *cnt_i,
                       const int *prod_cnt, int *total_e, int *alive_e, int *total_i, int *alive_i, int *prod_total,
                       int *start_e, int *start_i, int *prod_start, int *base_e, int *base_i, int *pbase_e,
                       int *pbase_i, int *pid_off, int *tmp, hipStream_t s);
void launch_cell_counts(int Gsize, const int *cell_start, const int *pcnt, int *total, hipStream_t s);
void launch_fill_int(int *p, int v, int n, hipStream_t s);
void launch_bump_ids(uint32_t *next_id, const int *prod_total_end, int nsp, hipStream_t s);
template <typename R> void launch_rhs_to_fp64(int N, const R *b, double *b64, hipStream_t s);
template <typename R> void launch_x_to_phi(int Gsize, const int *A_idx, const double *x, R *phi, hipStream_t s);

// ---- field solver (pcg.hip)
struct PcgResult {
    int iters;          // total inner iterations
    int outer;          // refinement cycles (1 for the plain path)
    double rel_res;     // final ||b - A x|| / ||b|| in FP64
    int status;         // 0 converged, 1 hit the iteration limit, 3 abandoned (a grid barrier wait timed out)
};

// Operator in padded 5-entry row form (missing entries have col = row, val = 0).
struct PcgOperator {
    int N;
    const int *col;          // 5N
    const float *valf;       // 5N
    const double *vald;      // 5N
    const float *mdiagf;     // N   Jacobi preconditioner 1/diag
    const double *mdiagd;
};

struct PcgWork;   // device scratch owned by the context
PcgWork *pcg_work_create(int N, int *err, int want_blocks, hipStream_t stream);
void pcg_work_destroy(PcgWork *);
int pcg_grid_blocks(PcgWork *);
// per-step diagnostics (diag.hip)
struct DiagConst {
    int G, ngy, ncy, Csize, nsp, ncond;
    double dt;
    double mass[5], np2c[5];
};
int diag_record_len(int nsp, int ncond);
template <typename R>
void launch_diag_sample(const DiagConst &d, const PSet<R> &ps, double *part, const R *Sigma, const R *Area,
                        const int *CondID, const int *condcharge, int *cc_prev, const R *V, long long step,
                        double *rec, hipStream_t s);
// metastable fluid (fluid.hip): device state and arrays of one case
struct FluidState {
    int nave, conti;
};
template <typename R>
struct FluidDev {
    int ncx, ncy, ngx, ngy, G, nf, nx, ny, n_ave, TnRct, gas;
    R dtc, dx, dy, dt, denscale_e, RR7, RR9;
    R dsc[8];   // Denscale of every particle species
    double dtm, sigv1;
    const int *x_x1, *x_x2, *x_y, *y_x, *y_y1, *y_y2, *plasma;
    const R *x_g1, *x_g2, *y_g1, *y_g2, *ax, *bx, *ay, *by, *BackDen1, *BackDen2, *RR;
    R *den, *src, *flux_x, *flux_y, *xs, *gs, *win, *nodeN;
    unsigned long long *mark;
    FluidState *state;
};
template <typename R>
void launch_fluid_step(const FluidDev<R> &d, const R *den_e, const unsigned long long *mccr, const R *P1, R *P2,
                       hipStream_t s);
// V000 dump particle records of one species, in slot order (dump.hip)
template <typename R>
void launch_dump_pack(const PView<R> &p, int ncy, int ngy, int Csize, uint32_t *rec, hipStream_t s);

// clears the grid barrier after an abandoned solve (stream-ordered after it)
hipError_t pcg_work_reset(PcgWork *, hipStream_t);
// picks the persistent solver's threads per block and grid size for this
// operator by timing fixed-length trial solves (cached per size and path);
// synchronises the stream
int pcg_autotune(PcgWork *, const PcgOperator &A, int refined, hipStream_t st, double *best_us_per_iter);
int pcg_launch_shape(const PcgWork *, int *blocks, int *threads);

// x (FP64, warm start, in/out), b (FP64).  Writes the result to *res (device
// memory) so the solve can run inside a captured graph.
hipError_t pcg_solve_plain_fp64(const PcgOperator &A, const double *b, double *x, double tol, int maxit, PcgWork *w,
                                PcgResult *res, hipStream_t s);
hipError_t pcg_solve_refined(const PcgOperator &A, const double *b, double *x, double tol, double inner_tol,
                             int maxit, PcgWork *w, PcgResult *res, hipStream_t s);

// ---- separable fast solve (spectral.hip), FP64 throughout
struct SpecDev {
    int nx, ny;
    const double *Vxi, *VyiT, *Vx, *VyT, *dinv;   // dinv = 1/(cx lx_a + cy ly_b), 0 for a null mode
    double *T1, *T2, *r;                            // nx*ny work matrices, r = residual
};
struct SpecState {
    int done, outer, status, pad;
    double rel;
};
// b32 = right-hand side (FP32 or FP64), converted into b (FP64) by the first residual
template <typename RB>
void spectral_solve_refined(const PcgOperator &A, const SpecDev &sp, const RB *b32, double *b, double *x,
                            double tol, int max_cycles, SpecState *state, double *partials, PcgResult *res,
                            hipStream_t st);

}  // namespace pic
