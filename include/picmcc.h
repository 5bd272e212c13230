/*
 * picmcc.h - C ABI of the MI355X PIC-MCC step library (libpicmcc.so).
 *
 * Two layers:
 *  1. pic_* functions on an explicit context (one discharge case: deck,
 *     device-resident particles, fields and random-stream state).  Each
 *     returns 0 or a negative status; pic_last_error() gives the message.
 *  2. The reference's zero-argument stage functions (inc/cuda_main.cuh:36-61,
 *     inc/main.h:20,34-41) as shims acting on a default context opened with
 *     pic_shim_open(); they print the error and exit(1) like the reference's
 *     gpuErrchk (src/cuda_main.cu:2-10).
 * No torch types, plain pointers and sizes only.
 */
#ifndef PICMCC_H
#define PICMCC_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pic_ctx pic_ctx;

enum pic_precision {
    PIC_FP32_REFINED = 0, /* FP32 particles/fields, FP32 PCG inside FP64 refinement */
    PIC_FP64_PLAIN = 1    /* the reference algorithm promoted to FP64 */
};

typedef struct {
    int precision;        /* enum pic_precision */
    int device;           /* HIP device ordinal */
    double inner_tol;     /* refined path: relative residual drop per FP32 inner solve (0 = 1e-4) */
    int pcg_maxit;        /* iteration cap per field solve (0 = 100000, the reference's cap) */
    int load;             /* 1 = load the initial particles from the deck (default when opt is NULL) */
    int field_solver;     /* 0 = fast separable solve when the operator allows it (refined path), 1 = PCG,
                             2 = fast separable solve when the operator allows it, either precision */
} pic_options;

typedef struct {
    int ngx, ngy, nsp, A_size, CondNUMR, TnRct, MCCn, nvel;
    double dt, dx, dy, t;
    long long tstep;
    int precision;
    long long np[5];
    long long cap[5];
} pic_info;

/* ---------------- context ---------------- */
/* deck = JSON text in the reference format (src/start.c:19-1431) */
int pic_create(const char *deck_json, const pic_options *opt, pic_ctx **out);
/* same, reading the deck from a file (the reference's -i <case.json>, src/main.c:27-39) */
int pic_create_from_file(const char *path, const pic_options *opt, pic_ctx **out);
int pic_destroy(pic_ctx *ctx);
const char *pic_last_error(void);
int pic_info_get(pic_ctx *ctx, pic_info *info);
int pic_sync(pic_ctx *ctx);   /* waits for the stream; reports device-side overflow */
/* non-blocking: 1 when the case's enqueued work has finished, 0 while it
 * runs, <0 on error (no counterpart in the reference, which synchronises
 * after every stage; used by the ensemble server to overlap cases) */
int pic_query(pic_ctx *ctx);
/* free and total device memory of a device, in bytes (no counterpart in the
 * reference; the ensemble server sizes its admission budget with it) */
int pic_device_memory(int device, size_t *free_bytes, size_t *total_bytes);
/* waits for all work on a device and reports any device error (no
 * counterpart in the reference; test harnesses call it between cases) */
int pic_device_sync(int device);

/* ---------------- stage table (src/cuda_main.cu:112-235) ---------------- */
int pic_advance_time(pic_ctx *ctx);  /* t += dt, tstep++ (src/cuda_main.cu:114-116) */
int pic_field_solve(pic_ctx *ctx);   /* PCG_SOLVER, src/cuda_Field.cu:133-181 */
int pic_efield(pic_ctx *ctx);        /* Efield_cuda_Basic, src/cuda_Field.cu:2-42 */
int pic_move(pic_ctx *ctx);          /* Move_cuda, src/cuda_Move.cu:3-7 */
int pic_sort_boundary(pic_ctx *ctx); /* SortBounndary_cuda, src/cuda_Sortboundary.cu:4-13 */
int pic_mcc(pic_ctx *ctx);           /* MCC_Ar_cuda, src/cuda_mcc.cu:97-106; for gas type 1
                                       MCC_O2_cuda, src/cuda_mcc.cu:72-96; for gas type 2
                                       MCC_ArO2_cuda, src/cuda_mcc.cu:21-71 */
int pic_deposit(pic_ctx *ctx);       /* Deposit_cuda / Deposit_Basic, src/cuda_Deposit.cu:3-29 */
int pic_diagnostic(pic_ctx *ctx);    /* profile accumulators, src/cuda_Diagnostic.cu:505-547 */
/* electron energy distribution: adds a 256-bin histogram of the electrons'
 * energies (eV, bins over [0, emax_ev), the top bin also takes higher
 * energies) to the "eedf" array; a different emax_ev restarts it.  The
 * reference's EEDF diagnostic is commented out (src/start.c:934-967). */
int pic_eedf_sample(pic_ctx *ctx, double emax_ev);
/* Diagnostics of the last completed step, in FP64 (src/cuda_Diagnostic.cu:37-267):
 * appends one record to the case's history (enqueued, no host
 * synchronisation; the newest 4096 records are kept).  Record layout,
 * doubles: t, step, np[nsp], kinetic energy per species (sum of m v^2/2 x
 * np2c, J per metre of depth), charge of each conductor (sum of Sigma x
 * Area over its nodes, the reference's Surf_charge), particles of each
 * (species, conductor) absorbed since the previous record, conductor
 * voltages.  The reference's conduction current is q_density x absorbed /
 * (steps between records x dt); its displacement current is the backward
 * difference of the conductor charge. */
int pic_diag_sample(pic_ctx *ctx);
/* copies up to max_records of the history, oldest first, into out
 * (max_records x *rec_len doubles; out may be NULL to query the sizes) */
int pic_diag_history(pic_ctx *ctx, double *out, int max_records, int *nrec, int *rec_len);
/* Neutral fluid (the reference's continuity solver and source update of
 * its warm-up loop, src/cuda_main.cu:84, src/cuda_Fluid.cu:34-178,
 * src/cuda_Diagnostic.cu:37-110, 548-744): Ar* for argon; O2A, O2B, OP, OD
 * for O2; Ar*, O2A, O2B, OP, OD for Ar/O2.  With on != 0 every step, after
 * the deposit, adds the electron density to the averaging window; every
 * AverageIter+1-th step averages it and the collision rates, forms every
 * fluid species' source and switches the continuity solve on; from then on
 * each step solves the species' continuity equations (ADI,
 * Scharfetter-Gummel, time step dt x TimeStep_Conti) and hands the
 * densities to the nodes, where argon's electron null-collision
 * probability and the oxygen collision steps' target densities follow
 * them.  -75 when AverageIter is 0.  Arrays: fl_den, fl_src, fl_flux_x,
 * fl_flux_y (species-major), fl_win, FluidN, P2 (argon), fl_state (window
 * count, solve on). */
int pic_fluid_enable(pic_ctx *ctx, int on);
/* Checkpoint in the reference's dump format, version 0.0.0 (V000_DUMP,
 * src/cuda_SaveDump.cu:5-106, with the version header of SaveDumpFile,
 * src/cuda_SaveDump.cu:151-155): time, per species np2c, count and
 * particles (node index, absolute position in cell units, velocity),
 * fluid arrays, histories (from pic_diag_sample records), per-node
 * densities and surface charge, fields, averages and collision rates.
 * Synchronises the case's stream. */
int pic_dump_save(pic_ctx *ctx, const char *path);
/* Resume from such a file (LoadDumpFile / V000_LoadDUMP, src/load.c:5-120,
 * then the particle bucketing of Set_Particle_cuda, src/cuda_Particle.cu:3-58,
 * and the initial deposit, src/cuda_main.cu:28).  Restores the step and
 * time, the particles (ids renumbered 0..n-1 in cell order), the surface
 * charge per species, the Poisson and total potentials and E; counters and
 * averages start again.  -63: the dump's gas type differs from the deck's;
 * -64: np2c differs (the reference warns and adopts the dump's value);
 * -30: more particles than the capacity; -61: the file ends early;
 * -65: a particle outside the grid. */
int pic_dump_load(pic_ctx *ctx, const char *path);
/* ---------------- batches: several cases stepped by shared launches ----------------
 * No counterpart in the reference (which runs one case per process).  A
 * batch steps its cases together: each kernel of the fused step runs once
 * for all of them over a grid that concatenates the cases' grids, from job
 * records prepared at creation; one step of the batch is one graph launch.
 * The cases must share device and precision, be argon cases on the
 * separable field solve (pic_field_solver_kind 1) with equal solve cycles,
 * and have stage timers off (-70 otherwise).  Each case gives the same bits
 * as when stepped alone.  A batch step is ordered after the work already
 * enqueued on the cases' streams and before work enqueued on them later, so
 * pic_sync / pic_get_* on a case see the batch's steps.  A case stepped
 * outside its batch invalidates the batch (pic_batch_step returns -71); a
 * case may not be destroyed while in a batch (pic_destroy returns -72). */
typedef struct pic_batch pic_batch;
int pic_batch_create(pic_ctx *const *cases, int n, pic_batch **out);
int pic_batch_step(pic_batch *b, int nsteps);   /* enqueued, no host synchronisation */
int pic_batch_sync(pic_batch *b);               /* waits; reports each case's errors (pic_sync) */
int pic_batch_destroy(pic_batch *b);            /* the cases stay valid */
/* one full timestep through the stage functions above, nsteps times */
int pic_run_staged(pic_ctx *ctx, int nsteps);
/* one full timestep on the fused path (push+walls+MCC counted, scanned and
 * scattered in two passes), nsteps times; no host synchronisation inside */
int pic_step(pic_ctx *ctx, int nsteps);

/* ---------------- state exchange ---------------- */
/* particle arrays of species isp: element type double for PIC_FP64_PLAIN,
 * float otherwise; buffers must hold np (from pic_info_get) elements;
 * cell_start holds ngx*ngy+1 ints */
int pic_get_particles(pic_ctx *ctx, int isp, void *x, void *y, void *vx, void *vy, void *vz, uint32_t *id,
                      int32_t *cell_start);
int pic_set_particles(pic_ctx *ctx, int isp, long long n, const void *x, const void *y, const void *vx,
                      const void *vy, const void *vz, const uint32_t *id, const int32_t *cell_start,
                      uint32_t next_id);
/* named arrays (node fields, setup maps, tables, counters); see the
 * library's array list in pic_array_names() */
int pic_array_info(pic_ctx *ctx, const char *name, size_t *count, int *elsize, int *is_float);
int pic_get_array(pic_ctx *ctx, const char *name, void *host, size_t bytes);
int pic_set_array(pic_ctx *ctx, const char *name, const void *host, size_t bytes);
const char *pic_array_names(void);
int pic_set_step(pic_ctx *ctx, long long tstep);
int pic_last_field_solve(pic_ctx *ctx, int *iters, int *outer, double *rel_res);
/* per-stage GPU time of pic_step, from HIP events on the case's stream:
 * ms[0] field solve, [1] E field, [2] electron count pass (push+walls+MCC),
 * [3] ion count pass, [4] scan/bases, [5] electron scatter, [6] ion scatter,
 * [7] deposit, [8] accumulators; summed over *steps steps since enabling */
int pic_stage_timing(pic_ctx *ctx, int enable);
int pic_stage_times(pic_ctx *ctx, double *ms, int n, long long *steps);
/* 0 = Jacobi PCG, 1 = separable fast solve (FP64 fast-Poisson inverse in FP64 refinement) */
int pic_field_solver_kind(pic_ctx *ctx);
/* launch shape of the persistent PCG (blocks in the grid, threads per block),
 * chosen at pic_create by timing trial solves of the case's operator (kept
 * per operator size for the process; the environment variable
 * PIC_PCG_BLOCKS fixes the grid instead); 0, 0 with the separable solve.
 * No counterpart in the reference, whose PCG_float launch is fixed. */
int pic_field_solver_shape(pic_ctx *ctx, int *blocks, int *threads);

/* ---------------- batched Poisson solves ---------------- */
/* Many independent 5-point problems (unit spacing, operator = the reference
 * field operator's stencil with eps = 1, src/load.c:121-245) solved together.
 * Problem p has nx[p] x ny[p] unknowns, unknown (a, b) at a*ny + b of its
 * block (blocks concatenated in problem order); bc[4p..4p+3] = left (a = 0),
 * right, bottom (b = 0), top: 0 Neumann (mirror), 1 Dirichlet (0 V).
 * No counterpart in the reference, which solves one problem per case. */
typedef struct pic_pbatch pic_pbatch;
int pic_pbatch_create(int nprob, const int *nx, const int *ny, const int *bc, int device, pic_pbatch **out);
int pic_pbatch_destroy(pic_pbatch *b);
long long pic_pbatch_unknowns(pic_pbatch *b);
int pic_pbatch_set_rhs(pic_pbatch *b, const double *rhs);          /* host array of all unknowns */
int pic_pbatch_solve(pic_pbatch *b, double tol, int max_cycles);   /* enqueued, from x = 0 */
int pic_pbatch_sync(pic_pbatch *b);
/* waits; copies x and per-problem relative residuals (either may be NULL);
 * returns the number of problems above tol, or < 0 on error */
int pic_pbatch_result(pic_pbatch *b, double *x, double *rel_res);
const char *pic_pbatch_last_error(void);
/* pic_pbatch_create + pic_pbatch_solve + pic_pbatch_result in one call that
 * allocates no device memory.  No counterpart in the reference: it is for
 * callers that cannot allocate (code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc).  b_dev and x_dev are device arrays of
 * all unknowns, laid out as for pic_pbatch_create; rel_res (host, nprob
 * entries) may be NULL; stream is a hipStream_t (NULL = the device's null
 * stream).  Tables, eigenbases and work vectors live in a fixed 64 MiB
 * device array that every call on a device shares (a batch needs about
 * 24 bytes per unknown plus 16 n^2 bytes per distinct side length n, and 8
 * bytes per unknown for 1/(lx+ly)); a larger batch returns -3.  Calls that
 * overlap on one device race.  Waits for the solve; returns the number of
 * problems above tol, or < 0 on error. */
int pic_pbatch_solve_noalloc(int nprob, const int *nx, const int *ny, const int *bc, int device, const double *b_dev,
                             double *x_dev, double tol, int max_cycles, double *rel_res, void *stream);

/* ---------------- host-only setup (no device needed) ---------------- */
/* deck -> geometry maps, field operator, Laplace solutions, tables and the
 * initial particle load, computed on the host exactly as pic_create does */
typedef struct pic_setup_h pic_setup_h;
int pic_setup_open(const char *deck_json, int precision, int load, pic_setup_h **out);
int pic_setup_array_info(pic_setup_h *h, const char *name, size_t *count, int *elsize, int *is_float);
int pic_setup_get_array(pic_setup_h *h, const char *name, void *host, size_t bytes);
int pic_setup_scalar(pic_setup_h *h, const char *name, double *value);
void pic_setup_close(pic_setup_h *h);
/* Reaction switches of gas type 0 (Ar), 1 (O2) or 2 (Ar/O2): the name of
 * the deck's object under "Collision_On_Off", then every reaction key in
 * reaction order, newline-separated (src/start.c:1003-1420); NULL for an
 * unknown type.  For gas types 1 and 2 pic_mcc applies the O2 or Ar/O2
 * collision step, and pic_step runs the stages one by one. */
const char *pic_reaction_names(int gas);

/* ---------------- reference-named zero-argument shims ---------------- */
int pic_shim_open(const char *deck_path, int precision);
void pic_shim_close(void);
/* the context the shims act on (NULL before pic_shim_open), so that a C
 * driver that calls the stage shims one by one can advance the clock with
 * pic_advance_time, as the reference's loop does with t += dt; tstep++
 * (src/cuda_main.cu:114-116), and read results with the pic_* calls */
pic_ctx *pic_shim_ctx(void);
void main_cuda(void);           /* src/cuda_main.cu:11 (runs RFCycles cycles) */
void start_cuda(void);          /* src/cuda_start.cu:2 */
void info_Device(void);         /* src/cuda_start.cu:36 */
void Set_Device_Parameter(void);
void Set_Particle_cuda(void);
void Set_NullCollisionTime_cuda(void);
void Set_SortBoundary_cuda(void);
void Set_MatrixPCG_cuda(void);
void PCG_SOLVER_Laplace(void);
void PCG_SOLVER(void);          /* src/cuda_Field.cu:133 */
void Efield_cuda_Basic(void);   /* src/cuda_Field.cu:2 */
void Move_cuda(void);           /* src/cuda_Move.cu:3 */
void SortBounndary_cuda(void);  /* src/cuda_Sortboundary.cu:4 (sic) */
void MCC_Ar_cuda(void);         /* src/cuda_mcc.cu:97 dispatch */
void MCC_O2_cuda(void);         /* src/cuda_mcc.cu:7-10 dispatch (gas type 1) */
void MCC_ArO2_cuda(void);       /* src/cuda_mcc.cu:2-6 dispatch (gas type 2) */
void Deposit_cuda(void);        /* src/cuda_Deposit.cu:3 */
void Deposit_Basic(void);       /* src/cuda_Deposit.cu:21 */
void Diagnostic(void);          /* DIAG table entry (src/cuda_start.cu:33): the profile
                                   accumulators of src/cuda_Diagnostic.cu:505-547 */

#ifdef __cplusplus
}
#endif
#endif
