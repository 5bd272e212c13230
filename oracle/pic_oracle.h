/*
 * ORACLE - test infrastructure only.  A CPU restatement of the reference
 * ainch/Hybrid-Final per-timestep PIC-MCC loop (src/cuda_main.cu:112-235)
 * and of the host setup it depends on.  Only tests/, __graft_entry__.smoke()
 * and bench.py's CPU-baseline leg may load it, and only as the checker.
 *
 * Built twice from the same source: liborc_f32.so (REAL=float, the
 * reference's own precision) and liborc_f64.so (REAL=double, the reference
 * algorithm promoted to FP64).
 *
 * Two changes to the reference are deliberate and shared with the GPU path:
 *   1. random numbers come from Philox4x32-10 keyed on (seed, particle id,
 *      step, stage tag) instead of cuRAND XORWOW per cell;
 *   2. particle order inside a cell is fixed: stayers in their old order,
 *      then arrivals grouped by source cell (ascending) in old order, then
 *      collision products in parent order (parent species ascending).
 * Everything else follows the reference; each function cites the lines it
 * restates.  Reference defects that are undefined behaviour are resolved as
 * noted at the function.  Three more are fixed, as on the GPU path: an
 * offset left outside its cell by a move of more than one cell is clamped
 * into it, a mirrored offset of 1 is kept below 1, and particles deleted at
 * walls are counted (Lost).  The others are preserved.
 *
 * O2 and Ar/O2 decks: orc_create sets up the grid, the field operator, the
 * backgrounds, the fluid and the rate constants only (no particles;
 * orc_step returns -19); the oxygen collision step is orc_mcc_o2 on
 * host-set-up inputs, the oxygen fluid orc_fluid_step.
 *
 * How each part is pinned (reference-built fixtures, known answers, closed
 * forms, statistics of the collision channels, the reaction equations) is
 * listed with the tests in tests/test_oracle_pins.py,
 * tests/test_setup_parity.py and tests/test_fluid_pins.py.
 */
#ifndef PIC_ORACLE_H
#define PIC_ORACLE_H
#include <stdint.h>

#ifndef REAL
#define REAL double
#endif
typedef REAL real;

#define ORC_MAXL 32      /* max boundary / conductor / source / dielectric lines */
#define ORC_MAXSP 5
#define ORC_MAXRCT 100
#define ORC_MAXFSP 5   /* fluid species: Ar* | O2A O2B OP OD | Ar* O2A O2B OP OD */

/* Flat copy of the JSON case deck (src/start.c:19-1431).  Filled from the
 * deck by oracle/oracle.py with parson's "missing key reads as 0" rule. */
typedef struct {
    double xlength, ylength, zlength;
    int ngx, ngy;
    int nbnd;
    int bx0[ORC_MAXL], bx1[ORC_MAXL], by0[ORC_MAXL], by1[ORC_MAXL], bbc[ORC_MAXL];
    double btemp[ORC_MAXL];
    int ncond;
    int cmid[ORC_MAXL], cx0[ORC_MAXL], cx1[ORC_MAXL], cy0[ORC_MAXL], cy1[ORC_MAXL];
    double ctemp[ORC_MAXL], cR[ORC_MAXL], cL[ORC_MAXL], cC[ORC_MAXL];
    int nsrc;
    int smid[ORC_MAXL];
    double sdc[ORC_MAXL], spow[ORC_MAXL], sac[ORC_MAXL], sfreq[ORC_MAXL], sphase[ORC_MAXL];
    int ndiel;
    int dmid[ORC_MAXL], dx0[ORC_MAXL], dx1[ORC_MAXL], dy0[ORC_MAXL], dy1[ORC_MAXL];
    double deps[ORC_MAXL];
    int gas;                      /* 0 Ar, 1 O2, 2 Ar/O2 */
    double total_pres;
    double bg_ratio[2], bg_temp[2];   /* Argon, Oxygen */
    int fl_loadtype;
    double fl_x0, fl_x1, fl_y0, fl_y1;
    double fl_dens[5], fl_temp[5], fl_quench[5];   /* Ar*, O2A, O2B, OP, OD */
    int sp_loadtype;
    double sp_x0, sp_x1, sp_y0, sp_y1;
    double sp_dens[ORC_MAXSP], sp_temp[ORC_MAXSP], sp_np2c[ORC_MAXSP];
    int sp_ratio[ORC_MAXSP], sp_maxnp[ORC_MAXSP];
    int dt_pic;
    int dt_conti;                 /* continuity time step in PIC steps (TimeStep_Conti) */
    double pcg_tol;
    int n_ave, n_smt, np_limit;
    double coll_flag[ORC_MAXRCT];
    int64_t seed;                 /* Park-Miller loader seed (reference: getpid) */
    int64_t rng_seed;             /* Philox key for the time loop */
} orc_deck;

enum { ORC_DIRICHLET = 1, ORC_NEUMANN, ORC_CONDUCTOR, ORC_DIELECTRIC };
enum { ORC_LEFT, ORC_RIGHT, ORC_UP, ORC_DOWN, ORC_UL, ORC_UR, ORC_LL, ORC_LR, ORC_NOFACE };

typedef struct {
    int n, cap;
    real *x, *y, *vx, *vy, *vz;
    uint32_t *id;
    int *cell_start;              /* Gsize+1, bucket of node-indexed cell c */
    uint32_t next_id;
    /* second set of particle arrays that the OpenMP step rebuilds buckets
     * into and swaps with the first (NULL until used) */
    real *sx, *sy, *svx, *svy, *svz;
    uint32_t *sid;
} orc_species;

typedef struct {
    orc_deck deck;
    int ngx, ngy, ncx, ncy, Gsize, Csize, nsp;
    real dx, dy, zlength, dt, dt_dx, dt_dy, hdx, hdy, idx, idy, idxy;
    double dt_d;
    /* species constants (src/start.c:2005-2017) */
    real mass[ORC_MAXSP], q[ORC_MAXSP], np2c[ORC_MAXSP], q_density[ORC_MAXSP], vti[ORC_MAXSP];
    real qm[ORC_MAXSP], Escale[ORC_MAXSP], Ascale[ORC_MAXSP], Denscale[ORC_MAXSP];
    /* background + fluid */
    real bg_mass[2], bg_pres[2], bg_dens[2];
    real fl_mass[5], fl_dens[5];
    /* geometry (src/start.c:1432-1851) */
    int *Boundary, *CondID, *Face, *DensRegion;      /* Gsize */
    real *Area, *Temp;                               /* Gsize */
    int *PlasmaRegion; real *eps_r, *eps;            /* Csize */
    int *StructIdx;                                  /* (ncx+2)(ncy+2) */
    int CondNUMR, ncond;
    /* field matrix (src/load.c:121-245): CSR 0-based here */
    int A_size;
    int *A_idx;                                      /* Gsize, 1-based, 0 = eliminated */
    int *Ai, *Aj; real *Aval, *Atemp_val; real *Mdiag;
    real *cond_b;                                    /* CondNUMR x A_size */
    real *temp_b;
    /* Laplace solutions and gas fields */
    real *lap_phi;                                   /* CondNUMR x Gsize */
    real *BackDen1, *BackVel1, *BackDen2, *BackVel2; /* Gsize */
    real *FluidN;                                    /* nfsp x Gsize */
    int nfsp;
    /* sources */
    int *Efield_Flag;
    /* cross-section tables */
    int N_LOGX, TnRct;
    real idLOGX;
    real *xe, *xee;                                  /* N_LOGX */
    real *cx;                                        /* N_LOGX x TnRct (column r) */
    real Th_e[ORC_MAXRCT], mofM[ORC_MAXRCT], RR[ORC_MAXRCT], Flag[ORC_MAXRCT];
    real sigv[32];
    int MCCn; real dt_mcc;
    real *P1, *P2, *Pion;                            /* per-node collision probabilities */
    /* Maxwellian inverse-CDF table (src/Particle.c:176-210) */
    int nvel; real *vsave;
    /* particles */
    orc_species sp[ORC_MAXSP];
    int *tag;                                        /* per-particle move tag (MCID) */
    /* field state */
    real *phi;        /* Poisson part at nodes */
    real *TotPot;     /* total potential */
    real *Sigma, *Source, *Ex, *Ey, *den, *sigma_sp;  /* den, sigma_sp: nsp x Gsize */
    real *b;          /* PCG right-hand side, A_size */
    double *xsol;     /* last solution (warm start / diagnostics) */
    real *V_cond;     /* conductor voltages this step */
    real *CondCharge; /* nsp x CondNUMR */
    real *Lost;       /* nsp: particles deleted at a wall with no conductor or dielectric */
    int64_t *mccr;    /* Gsize x TnRct event counts, cumulative */
    /* profile accumulators (FP64, src/cuda_Diagnostic.cu:505-547) */
    double *sum_den, *sum_phi, *sum_ex, *sum_ey;
    int n_accum;
    int64_t tstep;
    double t;
    int err;
    /* banded LU of the field operator (FP64) */
    double *lu; int lu_bw;
    /* metastable fluid (src/Fluid.c, src/cuda_Fluid.cu): x lines (fixed y)
     * and y lines (fixed x) of plasma cells with their end factors per fluid
     * species; per species the cell diffusion coefficient and density
     * [ncx][ncy], the x-face coefficients and fluxes [ngx][ncy], the y-face
     * ones [ncx][ngy], the source [ncx][ncy]; the averaging window */
    real dtc;
    int nlx, nly;
    int *lx_x1, *lx_x2, *lx_y, *ly_x, *ly_y1, *ly_y2;
    real *lx_f1, *lx_f2, *ly_f1, *ly_f2;            /* nfsp x lines */
    real *fD, *fden, *fsrc, *fax, *fbx, *fflx, *fay, *fby, *ffly;
    int fluid_on, nave_count, conti_flag;
    real *win_den;                                  /* Gsize: electron density summed over the window */
    int64_t *mccr_mark;                             /* Gsize x TnRct: counts at the last average */
} orc_case;

/* Collision step of the oxygen chemistries, standalone: O2 (gas type 1,
 * chem 0; species e, O2+, O+, O-) or Ar/O2 (gas type 2, chem 1; species e,
 * Ar+, O2+, O+, O-).  The caller passes the setup arrays and the species'
 * buckets, sorted by node with cell_start[G+1] and room for cap particles.
 * The buckets are rebuilt in place; next_id, mccr (G x TnRct) and, when
 * not NULL, delcnt (G: O- removed by electron detachment) are advanced. */
typedef struct {
    int G, TnRct, nlogx, nvel, MCCn;
    real idLOGX, dt, dtm, idx, idy;
    real np2c[5], Escale[5], sigv[25];
    const real *Th_e, *xe, *cx, *vsave, *BackDen1, *BackVel1, *FluidN, *den;
    uint64_t seed;
    uint32_t step;
    int ions;   /* O2: 0 = only electrons collide, as the reference's MCC_O2_Basic (src/cuda_mcc.cu:76) */
    int chem;   /* 0 O2, 1 Ar/O2 */
    const real *BackDen2, *BackVel2;   /* Ar/O2: the O2 background */
    int64_t *delcnt;
} orc_o2_in;
typedef struct {
    real *x, *y, *vx, *vy, *vz;
    uint32_t *id;
    int *cell_start;
    int n, cap;
} orc_bucket;
int orc_mcc_o2(const orc_o2_in *k, orc_bucket *sp, uint32_t *next_id, int64_t *mccr);

#endif
