// Host-side case setup: the JSON deck and everything the time loop needs
// that is computed once (geometry maps, field operator, Laplace solutions,
// gas fields, cross-section tables, null-collision constants, Maxwellian
// table, initial particles).  Restates src/start.c:19-2052, src/load.c:121-245,
// src/Efield.c:3-14, src/PhysicsData.c:638-705,1579-3275,
// src/cuda_mcc.cu:202-501, src/Particle.c:3-230 and
// src/cuda_Field.cu:182-236,646-678.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace pic {

enum { DIRICHLET = 1, NEUMANN, CONDUCTOR, DIELECTRIC };
enum { F_LEFT, F_RIGHT, F_UP, F_DOWN, F_UL, F_UR, F_LL, F_LR, F_NOFACE };
constexpr int MAXSP = 5;
constexpr int AR_NRCT = 10;
constexpr int N_LOGX = 81;

struct Line { int x0 = 0, x1 = 0, y0 = 0, y1 = 0, id = 0, bc = 0; double temp = 0, a = 0, b = 0, c = 0; };
struct Src { int mid = 0; double dc = 0, pow = 0, ac = 0, freq = 0, phase = 0; };

// One case deck, with the reference's key names and "missing reads as 0".
struct Deck {
    double xlength = 0, ylength = 0, zlength = 0;
    int ngx = 0, ngy = 0;
    std::vector<Line> bnd, cond, diel;   // diel: a = epsilon; cond: a,b,c = R,L,C
    std::vector<Src> src;
    int gas = 0;
    double total_pres = 0, bg_ratio[2] = {0, 0}, bg_temp[2] = {0, 0};
    int fl_loadtype = 0;
    double fl_dens[5] = {0}, fl_temp[5] = {0}, fl_quench[5] = {0};
    int sp_loadtype = 0;
    double sp_x0 = 0, sp_x1 = 0, sp_y0 = 0, sp_y1 = 0;
    double sp_dens[MAXSP] = {0}, sp_temp[MAXSP] = {0}, sp_np2c[MAXSP] = {0};
    int sp_ratio[MAXSP] = {0}, sp_maxnp[MAXSP] = {0};
    int dt_pic = 0, dt_conti = 0;
    double pcg_tol = 0;
    int n_ave = 0, n_smt = 0, np_limit = 0;
    double coll_flag[100] = {0};
    int64_t seed = 1, rng_seed = 0;
    int o2_ion_coll = 0;   // O2: 1 = O2+, O+ and O- collide too (the reference's O2 step collides electrons only)
    int rf_cycles = 0;
};

// Parses deck text; returns 0 or a negative code with a message.
int parse_deck(const std::string &text, Deck &d, std::string &err);
// "Collision_On_Off" object name, then the reaction keys of gas type 0-2,
// newline-separated
std::string reaction_names(int gas);

template <typename R>
struct SpeciesInit {
    std::vector<R> x, y, vx, vy, vz;
    std::vector<uint32_t> id;
    std::vector<int> cell_start;   // Gsize + 1
    int n = 0, cap = 0;
};

template <typename R>
struct Setup {
    Deck deck;
    int ngx = 0, ngy = 0, ncx = 0, ncy = 0, Gsize = 0, Csize = 0, nsp = 0, nfsp = 0, nbg = 1, CondNUMR = 0;
    R dx = 0, dy = 0, zl = 0, dt = 0, dt_dx = 0, dt_dy = 0, hdx = 0, hdy = 0, idx = 0, idy = 0, idxy = 0;
    R mass[MAXSP] = {0}, q[MAXSP] = {0}, np2c[MAXSP] = {0}, q_density[MAXSP] = {0}, vti[MAXSP] = {0};
    R qm[MAXSP] = {0}, Escale[MAXSP] = {0}, Ascale[MAXSP] = {0}, Denscale[MAXSP] = {0};
    R bg_mass[2] = {0, 0}, bg_pres[2] = {0, 0}, bg_dens[2] = {0, 0}, fl_mass[5] = {0}, fl_dens[5] = {0};
    // geometry
    std::vector<int> Boundary, CondID, Face, DensRegion, PlasmaRegion, StructIdx, A_idx;
    std::vector<R> Area, Temp, eps_r, eps;
    // field operator, CSR 0-based (row = unknown)
    int A_size = 0;
    std::vector<int> Ai, Aj;
    std::vector<R> Aval, Atemp, Mdiag, cond_b, temp_b;
    // Laplace solutions (CondNUMR x Gsize) and gas fields; BackDen2 and
    // BackVel2 are the second background gas (O2 of an Ar/O2 mixture)
    std::vector<R> lap_phi, BackDen1, BackVel1, BackDen2, BackVel2, FluidN;
    std::vector<int> Efield_Flag;
    // collision data
    int TnRct = 0, MCCn = 1;
    R idLOGX = 0, dt_mcc = 0;
    std::vector<R> xe, xee, cx;      // cx: N_LOGX x TnRct
    R Th_e[100] = {0}, Flag[100] = {0}, mofM[100] = {0}, RR[100] = {0};
    R sigv[32] = {0};                // max sigma*v per (projectile, target) pair
    int nsigv = 0;                   // pairs: 3 (Ar), 20 (O2), 25 (Ar/O2)
    R colprob[MAXSP] = {0};          // collision probability per step, per projectile species
    std::vector<R> P1, P2, Pion;
    int nvel = 0;
    std::vector<R> vsave;
    SpeciesInit<R> sp[MAXSP];
    // metastable fluid (src/Fluid.c:3-256, src/cuda_Fluid.cu:3-33, 179-301):
    // lines of consecutive plasma cells along x (fixed y) and along y (fixed
    // x); per fluid species the wall flux factor at each line end, the
    // diffusion coefficient per cell, the Scharfetter-Gummel coefficients per
    // cell face, the initial density and face fluxes.  Layouts per species:
    // cell arrays [ncx][ncy], x faces [ngx][ncy], y faces [ncx][ngy].
    R dtc = 0;
    int fx_n = 0, fy_n = 0;
    std::vector<int> fx_x1, fx_x2, fx_y, fy_x, fy_y1, fy_y2;
    std::vector<R> fx_g1, fx_g2, fy_g1, fy_g2;           // nfsp x lines
    std::vector<R> fl_D, fl_den, fl_ax, fl_bx, fl_flux_x, fl_ay, fl_by, fl_flux_y;
};

template <typename R>
int build_setup(const Deck &d, Setup<R> &s, std::string &err, bool load = true);

// Separable field operator: when the unknowns form a rectangle with uniform
// coefficients, A = cx Tx (x) I + I (x) cy Ty with Tx, Ty the 1-D second
// difference operators (Neumann mirror or Dirichlet ends), so
// A^-1 = (Vx (x) Vy) diag(1/(cx lx_a + cy ly_b)) (Vx^-1 (x) Vy^-1).
// Row-major nx x nx / ny x ny FP64 matrices.
struct Spectral {
    bool ok = false;
    bool exact = false;   // every operator entry equals the separable model bit for bit
    int nx = 0, ny = 0;
    double cx = 0, cy = 0;
    std::vector<double> Vx, Vxi, Vy, Vyi, lx, ly;
};
template <typename R>
bool build_spectral(const Setup<R> &S, Spectral &sp, std::string &why);

// FP64 banded LU without pivoting (the field operator is an M-matrix).
struct BandLU {
    int N = 0, bw = 0;
    std::vector<double> a;
    template <typename R>
    int factor(int n, const std::vector<int> &Ai, const std::vector<int> &Aj, const std::vector<R> &val);
    void solve(const double *b, double *x) const;
};

}  // namespace pic
