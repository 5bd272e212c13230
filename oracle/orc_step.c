/*
 * ORACLE - test infrastructure only (see pic_oracle.h).
 * Per-timestep restatement of src/cuda_main.cu:112-235 for argon:
 * field solve, E-field, push, sort/boundary, null-collision MCC, deposit,
 * profile accumulation.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "pic_oracle.h"
#include "detmath.h"
#include "philox_ref.h"

#define EPS0 8.8542e-12
#define ORC_PI 3.14159265358979323846

int orc_setup(orc_case *c);
void orc_band_solve(int N, const double *lu, int bw, const double *b, double *x);

/* elementary functions of the collision physics in the path precision:
 * the FP32 build uses detmath.h's single-precision routines, the FP64 build
 * the double ones (the GPU path makes the same choice) */
#ifdef ORC_FLOAT
#define R_LOG10(x) dmf_log10(x)
#define R_POW(a, b) dmf_pow(a, b)
#define R_SINCOS(x, s, c) dmf_sincos(x, s, c)
#define R_TAN(x) dmf_tan(x)
#define R_ATAN(x) dmf_atan(x)
#else
#define R_LOG10(x) dm_log10(x)
#define R_POW(a, b) dm_pow(a, b)
#define R_SINCOS(x, s, c) (*(s) = dm_sin(x), *(c) = dm_cos(x))
#define R_TAN(x) dm_tan(x)
#define R_ATAN(x) dm_atan(x)
#endif

static void *zalloc(size_t n, size_t s) {
    void *p = calloc(n ? n : 1, s);
    if (!p) { fprintf(stderr, "oracle: out of memory\n"); exit(2); }
    return p;
}

/* ================= lifecycle ================= */
orc_case *orc_create(const orc_deck *d, int *err) {
    orc_case *c = zalloc(1, sizeof(orc_case));
    c->deck = *d;
    int e = orc_setup(c);
    if (err) *err = e;
    if (e) { return c; }
    int G = c->Gsize, s;
    c->phi = zalloc(G, sizeof(real)); c->TotPot = zalloc(G, sizeof(real));
    c->Sigma = zalloc(G, sizeof(real)); c->Source = zalloc(G, sizeof(real));
    c->Ex = zalloc(G, sizeof(real)); c->Ey = zalloc(G, sizeof(real));
    c->den = zalloc((size_t)c->nsp * G, sizeof(real));
    c->sigma_sp = zalloc((size_t)c->nsp * G, sizeof(real));
    c->b = zalloc(c->A_size, sizeof(real));
    c->xsol = zalloc(c->A_size, sizeof(double));
    c->V_cond = zalloc(c->CondNUMR, sizeof(real));
    c->CondCharge = zalloc((size_t)c->nsp * c->CondNUMR, sizeof(real));
    c->Lost = zalloc((size_t)c->nsp, sizeof(real));
    c->mccr = zalloc((size_t)G * c->TnRct, sizeof(int64_t));
    c->sum_den = zalloc((size_t)c->nsp * G, sizeof(double));
    c->sum_phi = zalloc(G, sizeof(double));
    c->sum_ex = zalloc(G, sizeof(double)); c->sum_ey = zalloc(G, sizeof(double));
    int maxcap = 0;
    for (s = 0; s < c->nsp; ++s) if (c->sp[s].cap > maxcap) maxcap = c->sp[s].cap;
    c->tag = zalloc((size_t)c->nsp * maxcap, sizeof(int));
    c->Efield_Flag = zalloc(c->CondNUMR, sizeof(int));
    {
        int i, j;
        for (i = 0; i < c->CondNUMR; ++i) {
            int cnt = 0;
            for (j = 0; j < c->deck.nsrc; ++j) if (c->deck.smid[j] == i + 1) cnt++;
            c->Efield_Flag[i] = (cnt > 0 && cnt < 4) ? 1 : 0;
        }
    }
    return c;
}

/* Frees a case and everything it owns (free(NULL) is a no-op, so a case
 * whose setup failed part-way is released too). */
void orc_destroy(orc_case *c) {
    int s;
    if (!c) return;
    void *own[] = {c->Boundary, c->CondID, c->Face, c->DensRegion, c->Area, c->Temp, c->PlasmaRegion, c->eps_r,
                   c->eps, c->StructIdx, c->A_idx, c->Ai, c->Aj, c->Aval, c->Atemp_val, c->Mdiag, c->cond_b,
                   c->temp_b, c->lap_phi, c->BackDen1, c->BackVel1, c->BackDen2, c->BackVel2, c->FluidN,
                   c->Efield_Flag, c->xe, c->xee, c->cx, c->P1, c->P2, c->Pion, c->vsave, c->tag, c->phi, c->TotPot,
                   c->Sigma, c->Source, c->Ex, c->Ey, c->den, c->sigma_sp, c->b, c->xsol, c->V_cond, c->CondCharge,
                   c->Lost, c->mccr, c->sum_den, c->sum_phi, c->sum_ex, c->sum_ey, c->lu,
                   c->lx_x1, c->lx_x2, c->lx_y, c->ly_x, c->ly_y1, c->ly_y2, c->lx_f1, c->lx_f2, c->ly_f1,
                   c->ly_f2, c->fD, c->fden, c->fsrc, c->fax, c->fbx, c->fflx, c->fay, c->fby, c->ffly,
                   c->win_den, c->mccr_mark};
    for (s = 0; s < (int)(sizeof own / sizeof own[0]); ++s) free(own[s]);
    for (s = 0; s < ORC_MAXSP; ++s) {
        orc_species *p = &c->sp[s];
        free(p->x); free(p->y); free(p->vx); free(p->vy); free(p->vz); free(p->id); free(p->cell_start);
        free(p->sx); free(p->sy); free(p->svx); free(p->svy); free(p->svz); free(p->sid);
    }
    free(c);
}

/* ================= field solve: src/cuda_Field.cu:133-181, 639-645 =========
 * The oracle solves A x = b exactly (FP64 banded LU); b comes from the last
 * deposit.  x is scattered to node phi with 0 at eliminated nodes. */
void orc_field(orc_case *c) {
    int N = c->A_size, g;
    double *bd = zalloc(N, sizeof(double));
    for (g = 0; g < N; ++g) bd[g] = (double)c->b[g];
    orc_band_solve(N, c->lu, c->lu_bw, bd, c->xsol);
    for (g = 0; g < c->Gsize; ++g) c->phi[g] = c->A_idx[g] ? (real)c->xsol[c->A_idx[g] - 1] : 0;
    free(bd);
}

/* ================= E-field: src/cuda_Field.cu:2-42, 500-558, 698-846 ======= */
static void eps_index(const orc_case *c, int TID, int *e_tid, int *e_left, int *e_down, int *e_cross) {
    int ngx = c->ngx, ngy = c->ngy;
    int x = TID / ngy, y = TID % ngy;
    int ET = x * (ngy - 1) + y;
    int EL = x ? ET - ngy + 1 : ET;
    int ED = y ? ET - 1 : ET;
    int EC = EL + ED - ET;
    int xo = (x == ngx - 1) ? ngy - 1 : 0;
    int yo = (y == ngy - 1) ? 1 : 0;
    ET -= (xo + yo); EL -= yo; ED -= xo;
    *e_tid = ET; *e_left = EL; *e_down = ED; *e_cross = EC;
}

static void cond_sigma(orc_case *c, const real *Phi, real *Sigma) {
    int ngx = c->ngx, ngy = c->ngy, TID, i;
    for (TID = 0; TID < c->Gsize; ++TID) {
        int x = TID / ngy, y = TID % ngy;
        int ET, EL, ED, EC;
        eps_index(c, TID, &ET, &EL, &ED, &EC);
        int pl = x ? TID - ngy : TID, pr = (x == ngx - 1) ? TID : TID + ngy;
        int pd = y ? TID - 1 : TID, pu = (y == ngy - 1) ? TID : TID + 1;
        real ss = 0;
        for (i = 0; i < c->nsp; ++i) ss += c->den[TID + i * c->Gsize] * c->q_density[i];
        int b = c->Boundary[TID], f = c->Face[TID];
        if (!((b == ORC_CONDUCTOR || b == ORC_DIRICHLET) && f != ORC_NOFACE)) continue;
        const real *e = c->eps;
        real hdx = c->hdx, hdy = c->hdy, idxy = c->idxy, P = Phi[TID];
        switch (f) {
        case ORC_UP: Sigma[TID] = hdy * (e[ET] + e[EL]) * (P - Phi[pu]) - hdx * ss; break;
        case ORC_DOWN: Sigma[TID] = hdy * (e[ED] + e[EC]) * (P - Phi[pd]) - hdx * ss; break;
        case ORC_LEFT: Sigma[TID] = hdx * (e[EL] + e[EC]) * (P - Phi[pl]) - hdy * ss; break;
        case ORC_RIGHT: Sigma[TID] = hdx * (e[ET] + e[ED]) * (P - Phi[pr]) - hdy * ss; break;
        case ORC_UL: Sigma[TID] = hdy * (e[ET] + e[EL]) * (P - Phi[pu]) + hdx * (e[EL] + e[EC]) * (P - Phi[pl]) - idxy * ss; break;
        case ORC_UR: Sigma[TID] = hdy * (e[ET] + e[EL]) * (P - Phi[pu]) + hdx * (e[ET] + e[ED]) * (P - Phi[pr]) - idxy * ss; break;
        case ORC_LL: Sigma[TID] = hdy * (e[ED] + e[EC]) * (P - Phi[pd]) + hdx * (e[EL] + e[EC]) * (P - Phi[pl]) - idxy * ss; break;
        case ORC_LR: Sigma[TID] = hdy * (e[ED] + e[EC]) * (P - Phi[pd]) + hdx * (e[ET] + e[ED]) * (P - Phi[pr]) - idxy * ss; break;
        default: break;
        }
    }
}

static void v_to_efield(orc_case *c, const real *Sigma, const real *Phi) {
    int ngx = c->ngx, ngy = c->ngy, TID;
    const real *e = c->eps;
    real dx = c->dx, dy = c->dy, hdx = c->hdx, hdy = c->hdy, idx = c->idx, idy = c->idy;
    for (TID = 0; TID < c->Gsize; ++TID) {
        int x = TID / ngy, y = TID % ngy;
        int pl = x ? TID - ngy : TID, pr = (x == ngx - 1) ? TID : TID + ngy;
        int pd = y ? TID - 1 : TID, pu = (y == ngy - 1) ? TID : TID + 1;
        int ET, EL, ED, EC;
        eps_index(c, TID, &ET, &EL, &ED, &EC);
        int b = c->Boundary[TID], f = c->Face[TID];
        real P = Phi[TID], Ex, Ey;
        real hxy = (real)(0.5 * (double)dx / (double)dy), hyx = (real)(0.5 * (double)dy / (double)dx);
        int wall = (b == ORC_CONDUCTOR || b == ORC_NEUMANN || b == ORC_DIRICHLET);
        int die = (b == ORC_DIELECTRIC || b >= 100);
        int hard = (b == ORC_CONDUCTOR || b == ORC_DIRICHLET || b == ORC_DIELECTRIC);
        if (f == ORC_LEFT) {
            if (wall) { Ex = -2 * Sigma[TID] / (e[EL] + e[EC]); Ey = 0; }
            else if (die) { Ex = idx * (Phi[pl] - P); Ey = hdy * (Phi[pd] - Phi[pu]); }
            else {
                Ex = (real)0.5 * (e[ET] + e[ED]) * (Phi[pr] - P) * idx;
                Ex -= hxy * e[ET] * (P - Phi[pu]) * idy;
                Ex -= hxy * e[ED] * (P - Phi[pd]) * idy;
                Ex /= (real)-0.5 * (e[EL] + e[EC]);
                Ey = hdy * (Phi[pd] - Phi[pu]);
            }
        } else if (f == ORC_RIGHT) {
            if (wall) { Ex = 2 * Sigma[TID] / (e[ET] + e[ED]); Ey = 0; }
            else if (die) { Ex = idx * (P - Phi[pr]); Ey = hdy * (Phi[pd] - Phi[pu]); }
            else {
                Ex = (real)0.5 * (e[EL] + e[EC]) * (Phi[pl] - P) * idx;
                Ex -= hxy * e[EL] * (P - Phi[pu]) * idy;
                Ex -= hxy * e[EC] * (P - Phi[pd]) * idy;
                Ex /= (real)0.5 * (e[ET] + e[ED]);
                Ey = hdy * (Phi[pd] - Phi[pu]);
            }
        } else if (f == ORC_UP) {
            if (wall) { Ex = 0; Ey = 2 * Sigma[TID] / (e[ET] + e[EL]); }
            else if (die) { Ex = hdx * (Phi[pl] - Phi[pr]); Ey = idy * (P - Phi[pu]); }
            else {
                Ex = hdx * (Phi[pl] - Phi[pr]);
                Ey = (real)0.5 * (e[ED] + e[EC]) * (Phi[pd] - P) * idy;
                Ey -= hyx * e[ED] * (P - Phi[pr]) * idx;
                Ey -= hyx * e[EC] * (P - Phi[pl]) * idx;
                Ey /= (real)0.5 * (e[ET] + e[EL]);
            }
        } else if (f == ORC_DOWN) {
            if (wall) { Ex = 0; Ey = -2 * Sigma[TID] / (e[ED] + e[EC]); }
            else if (die) { Ex = hdx * (Phi[pl] - Phi[pr]); Ey = idy * (Phi[pd] - P); }
            else {
                Ex = hdx * (Phi[pl] - Phi[pr]);
                Ey = (real)0.5 * (e[ET] + e[EL]) * (Phi[pu] - P) * idy;
                Ey -= hyx * e[ET] * (P - Phi[pr]) * idx;
                Ey -= hyx * e[EL] * (P - Phi[pl]) * idx;
                Ey /= (real)-0.5 * (e[ED] + e[EC]);
            }
        } else if (f == ORC_UL) {
            if (hard) { Ex = idx * (Phi[pl] - P); Ey = idy * (P - Phi[pu]); }
            else { Ex = hdx * (Phi[pl] - Phi[pr]); Ey = hdy * (Phi[pd] - Phi[pu]); }
        } else if (f == ORC_UR) {
            if (hard) { Ex = idx * (P - Phi[pr]); Ey = idy * (P - Phi[pu]); }
            else { Ex = hdx * (Phi[pl] - Phi[pr]); Ey = hdy * (Phi[pd] - Phi[pu]); }
        } else if (f == ORC_LL) {
            if (hard) { Ex = idx * (Phi[pl] - P); Ey = idy * (Phi[pd] - P); }
            else { Ex = hdx * (Phi[pl] - Phi[pr]); Ey = hdy * (Phi[pd] - Phi[pu]); }
        } else if (f == ORC_LR) {
            if (hard) { Ex = idx * (P - Phi[pr]); Ey = idy * (Phi[pd] - P); }
            else { Ex = hdx * (Phi[pl] - Phi[pr]); Ey = hdy * (Phi[pd] - Phi[pu]); }
        } else if (b == ORC_NEUMANN) {
            if (x == 0 || x == ngx - 1) { Ex = 0; Ey = hdy * (Phi[pd] - Phi[pu]); }
            else if (y == 0 || y == ngy - 1) { Ex = hdx * (Phi[pl] - Phi[pr]); Ey = 0; }
            else { Ex = c->Ex[TID]; Ey = c->Ey[TID]; }   /* reference leaves E unchanged */
        } else {
            Ex = hdx * (Phi[pl] - Phi[pr]); Ey = hdy * (Phi[pd] - Phi[pu]);
        }
        c->Ex[TID] = Ex; c->Ey[TID] = Ey;
    }
}

/* conductor voltages at time t (src/cuda_Field.cu:9-20): V = DC + AC sin(w t + phase) */
void orc_voltages(orc_case *c, double t) {
    int i, j;
    for (i = 0; i < c->CondNUMR; ++i) {
        real v = 0;
        if (c->Efield_Flag[i]) {
            for (j = 0; j < c->deck.nsrc; ++j)
                if (c->deck.smid[j] == i + 1) {
                    real w = (real)(2 * ORC_PI * (double)(real)c->deck.sfreq[j]);
                    real ph = (real)(ORC_PI / 180 * (double)(real)c->deck.sphase[j]);
                    v = (real)((double)v + ((double)(real)c->deck.sdc[j] +
                                            (double)(real)c->deck.sac[j] * dm_sin((double)w * t + (double)ph)));
                }
        }
        c->V_cond[i] = v;
    }
}

void orc_efield(orc_case *c) {
    int g, i;
    orc_voltages(c, c->t);
    for (g = 0; g < c->Gsize; ++g) {
        real tp = 0;
        for (i = 0; i < c->CondNUMR; ++i) tp += c->V_cond[i] * c->lap_phi[(size_t)i * c->Gsize + g];
        tp += (real)1 * c->phi[g];
        c->TotPot[g] = tp;
    }
    cond_sigma(c, c->phi, c->Sigma);     /* Poisson part only, as the reference */
    v_to_efield(c, c->Sigma, c->TotPot);
}

/* ================= push: src/cuda_Move.cu:181-270 =================
 * tag = relative cell step (0 = stays).  Positions are offsets in the cell.
 * As in the reference, a particle moves at most one cell per axis and step.
 * Deliberate change: an offset still outside [0, 1) after that single wrap
 * (a move of more than one cell) is clamped into the destination cell; the
 * reference leaves it outside, which corrupts the next CIC weights and can
 * grow without bound. */
#ifdef ORC_FLOAT
#define ORC_ONE_MINUS 0.99999994f           /* 1 - 2^-24, largest float below 1 */
#else
#define ORC_ONE_MINUS 0.99999999999999989   /* 1 - 2^-53, largest double below 1 */
#endif
void orc_move(orc_case *c) {
    int isp, cell, ngy = c->ngy;
    for (isp = 0; isp < c->nsp; ++isp) {
        orc_species *s = &c->sp[isp];
        int *tag = c->tag + (size_t)isp * s->cap;
        real As = c->Ascale[isp];
        for (cell = 0; cell < c->Gsize; ++cell) {
            int k0 = s->cell_start[cell], k1 = s->cell_start[cell + 1], k;
            if (k0 == k1) continue;
            real ews = c->Ex[cell], ewn = c->Ex[cell + 1], ees = c->Ex[cell + ngy], een = c->Ex[cell + ngy + 1];
            real yws = c->Ey[cell], ywn = c->Ey[cell + 1], yes = c->Ey[cell + ngy], yen = c->Ey[cell + ngy + 1];
            for (k = k0; k < k1; ++k) {
                real lx = s->x[k], ly = s->y[k];
                real dvx = ews * (1 - lx) * (1 - ly) + ewn * (1 - lx) * ly + ees * lx * (1 - ly) + een * lx * ly;
                real dvy = yws * (1 - lx) * (1 - ly) + ywn * (1 - lx) * ly + yes * lx * (1 - ly) + yen * lx * ly;
                real mx = s->vx[k] + dvx * As;
                real my = s->vy[k] + dvy * As;
                lx += mx * c->dt_dx;
                ly += my * c->dt_dy;
                int t = 0;
                if (ly >= 1 || ly < 0 || lx >= 1 || lx < 0) {
                    if (ly >= 1) { t += 1; ly -= 1; }
                    else if (ly < 0) { t -= 1; ly += 1; }
                    if (lx >= 1) { t += ngy; lx -= 1; }
                    else if (lx < 0) { t -= ngy; lx += 1; }
                    if (lx >= 1) lx = ORC_ONE_MINUS; else if (lx < 0) lx = 0;
                    if (ly >= 1) ly = ORC_ONE_MINUS; else if (ly < 0) ly = 0;
                }
                s->x[k] = lx; s->y[k] = ly; s->vx[k] = mx; s->vy[k] = my;
                tag[k] = t;
            }
        }
    }
}

/* ================= sort + boundary: src/cuda_Sortboundary.cu:41-258 =========
 * Returns 1 and sets *dest (species-local node-indexed cell) when the
 * particle survives; 0 when it is absorbed (conductor), deleted (Dirichlet)
 * or deposited as surface charge (dielectric).  Position/velocity are
 * transformed in place for Neumann reflections.  Structure codes the
 * reference leaves undefined (e.g. -3, -5 from a Dirichlet node next to a
 * Neumann wall) delete the particle. */
#define RABS(v) ((v) < 0 ? -(v) : (v))
static int resolve_crosser(orc_case *c, int isp, int ID, int MCID, real *px, real *py, real *pvx, real *pvy, int *dest) {
    int ngy = c->ngy;
    int x = ID / ngy;
    int SM;
    if (MCID > 1) SM = (ID + ngy + 2 + x) + (MCID + 1);
    else if (MCID < -1) SM = (ID + ngy + 2 + x) + (MCID - 1);
    else SM = (ID + ngy + 2 + x) + MCID;
    const int *S = c->StructIdx;
    int CID = ID + MCID;
    int SI = S[SM];
    real X = *px, Y = *py, VX = *pvx, VY = *pvy;
    if (SI == 0) { *dest = CID; return 1; }
    if (SI >= 1 && SI < 100) {
        int pa = 0, pb = 0; real a = 0, b = 0, delta;
        if (MCID == 1) { delta = X - VX / VY * Y; pa = CID; a = 1 - delta; pb = CID + ngy; b = delta; }
        else if (MCID == -1) { delta = X + VX / VY * (1 - Y); pa = ID; a = 1 - delta; pb = ID + ngy; b = delta; }
        else if (MCID == ngy) { delta = Y - VY / VX * X; pa = CID; a = 1 - delta; pb = CID + 1; b = delta; }
        else if (MCID == -ngy) { delta = Y + VY / VX * (1 - X); pa = ID; a = 1 - delta; pb = ID + 1; b = delta; }
        else if (MCID == ngy + 1) {
            delta = X - VX / VY * Y;
            if (delta >= 0) {
                if (S[SM - 1]) { delta = Y - VY / VX * X; pa = CID - 1; a = -delta; pb = CID; b = 1 + delta; }
                else { pa = CID; a = 1 - delta; pb = CID + ngy; b = delta; }
            } else {
                if (S[SM - ngy - 1]) { pa = ID + 1; a = -delta; pb = CID; b = 1 + delta; }
                else { delta = Y - VY / VX * X; pa = CID; a = 1 - delta; pb = CID + 1; b = delta; }
            }
        } else if (MCID == ngy - 1) {
            delta = X + VX / VY * (1 - Y);
            if (delta >= 0) {
                if (S[SM + 1]) { delta = Y - VY / VX * X; pa = CID + 1; a = 2 - delta; pb = CID + 2; b = delta - 1; }
                else { pa = CID + 1; a = 1 - delta; pb = CID + 1 + ngy; b = delta; }
            } else {
                if (S[SM - ngy - 1]) { pa = ID; a = -delta; pb = CID + 1; b = 1 + delta; }
                else { delta = Y - VY / VX * X; pa = CID; a = 1 - delta; pb = CID + 1; b = delta; }
            }
        } else if (MCID == -ngy + 1) {
            delta = X - VX / VY * Y;
            if (delta < 1) {
                if (S[SM - 1]) { delta = Y + VY / VX * (1 - X); pa = CID + ngy - 1; a = -delta; pb = CID + ngy; b = 1 + delta; }
                else { pa = CID; a = 1 - delta; pb = CID + ngy; b = delta; }
            } else {
                if (S[SM + ngy + 1]) { pa = ID + 1; a = 2 - delta; pb = ID + 1 + ngy; b = delta - 1; }
                else { delta = Y + VY / VX * (1 - X); pa = ID + 1; a = 1 - delta; pb = ID + 2; b = delta; }
            }
        } else if (MCID == -ngy - 1) {
            delta = X + VX / VY * (1 - Y);
            if (delta < 1) {
                if (S[SM + 1]) { delta = Y + VY / VX * (1 - X); pa = ID; a = 2 - delta; pb = ID + 1; b = delta - 1; }
                else { pa = CID; a = 1 - delta; pb = ID; b = delta; }
            } else {
                if (S[SM + ngy + 1]) { pa = ID; a = 2 - delta; pb = ID + ngy; b = delta - 1; }
                else { delta = Y + VY / VX * (1 - X); pa = ID - 1; a = 1 - delta; pb = ID; b = delta; }
            }
        }
        real *sg = c->sigma_sp + (size_t)isp * c->Gsize;
        sg[pa] += a;
        sg[pb] += b;
        return 0;
    }
    if (SI >= 100) { c->CondCharge[isp * c->CondNUMR + (SI - 101)] += 1; return 0; }
    int sum, fx, fy, fvx, fvy;
    switch (SI) {
    case -2: sum = ngy; fx = 0; fy = 1; fvx = -1; fvy = 1; break;
    case -6: sum = -ngy; fx = 0; fy = 1; fvx = -1; fvy = 1; break;
    case -4: sum = 1; fx = 1; fy = 0; fvx = 1; fvy = -1; break;
    case -8: sum = -1; fx = 1; fy = 0; fvx = 1; fvy = -1; break;
    case -12: sum = ngy + 1; fx = 0; fy = 0; fvx = -1; fvy = -1; break;
    case -16: sum = 1 - ngy; fx = 0; fy = 0; fvx = -1; fvy = -1; break;
    case -20: sum = ngy - 1; fx = 0; fy = 0; fvx = -1; fvy = -1; break;
    case -24: sum = -ngy - 1; fx = 0; fy = 0; fvx = -1; fvy = -1; break;
    default: c->Lost[isp] += 1; return 0;   /* -1 Dirichlet, and undefined codes: deleted */
    }
    *dest = CID + sum;
    /* a mirrored offset of 0 (clamped onto the wall) would be 1: keep it inside */
    if (!fx) { *px = 1 - X; if (*px >= 1) *px = ORC_ONE_MINUS; }
    if (!fy) { *py = 1 - Y; if (*py >= 1) *py = ORC_ONE_MINUS; }
    *pvx = fvx * VX;
    *pvy = fvy * VY;
    return 1;
}

/* Rebucket after the push.  Order in the new bucket of cell d: particles of
 * d that did not cross (old order), then survivors arriving in d, grouped by
 * source cell in ascending order, each group in old order. */
int orc_sort(orc_case *c) {
    int isp;
    for (isp = 0; isp < c->nsp; ++isp) {
        orc_species *s = &c->sp[isp];
        int *tag = c->tag + (size_t)isp * s->cap;
        int n = s->n, G = c->Gsize, cell, k;
        int *destof = zalloc(n, sizeof(int));
        char *cls = zalloc(n, 1);
        int *scnt = zalloc(G, sizeof(int)), *acnt = zalloc(G, sizeof(int));
        for (cell = 0; cell < G; ++cell)
            for (k = s->cell_start[cell]; k < s->cell_start[cell + 1]; ++k) {
                if (tag[k] == 0) { destof[k] = cell; cls[k] = 0; scnt[cell]++; continue; }
                int d;
                if (resolve_crosser(c, isp, cell, tag[k], &s->x[k], &s->y[k], &s->vx[k], &s->vy[k], &d)) {
                    destof[k] = d; cls[k] = 1; acnt[d]++;
                } else { destof[k] = -1; cls[k] = 2; }
            }
        int *start = zalloc(G + 1, sizeof(int));
        for (cell = 0; cell < G; ++cell) start[cell + 1] = start[cell] + scnt[cell] + acnt[cell];
        int total = start[G];
        real *nx = zalloc(s->cap, sizeof(real)), *ny = zalloc(s->cap, sizeof(real));
        real *nvx = zalloc(s->cap, sizeof(real)), *nvy = zalloc(s->cap, sizeof(real)), *nvz = zalloc(s->cap, sizeof(real));
        uint32_t *nid = zalloc(s->cap, sizeof(uint32_t));
        int *sfill = zalloc(G, sizeof(int)), *afill = zalloc(G, sizeof(int));
        for (k = 0; k < n; ++k) {
            if (cls[k] == 2) continue;
            int d = destof[k], pos;
            if (cls[k] == 0) pos = start[d] + sfill[d]++;
            else pos = start[d] + scnt[d] + afill[d]++;
            nx[pos] = s->x[k]; ny[pos] = s->y[k];
            nvx[pos] = s->vx[k]; nvy[pos] = s->vy[k]; nvz[pos] = s->vz[k]; nid[pos] = s->id[k];
        }
        free(s->x); free(s->y); free(s->vx); free(s->vy); free(s->vz); free(s->id);
        s->x = nx; s->y = ny; s->vx = nvx; s->vy = nvy; s->vz = nvz; s->id = nid;
        memcpy(s->cell_start, start, (G + 1) * sizeof(int));
        s->n = total;
        free(destof); free(cls); free(scnt); free(acnt); free(start); free(sfill); free(afill);
        memset(tag, 0, sizeof(int) * s->cap);
    }
    return 0;
}

/* ================= MCC (argon): src/cuda_mcc.cu:97-106, 139-201;
 *                   src/cuda_mccAr.cu:41-232, 707-777 ================= */
#define TWO_PI_R ((real)(2.0 * ORC_PI))
#if defined(ORC_FLOAT)
#define RSQRT(v) sqrtf(v)
#else
#define RSQRT(v) sqrt(v)
#endif

static real ar_sigma(const orc_case *c, int r, real engy) {
    if (engy == 0) return 0;
    int N = c->N_LOGX, T = c->TnRct;
    real lengy = R_LOG10(engy);
    real ee1 = c->idLOGX * (lengy - c->xe[0]);
    int ee2 = (int)ee1;
    real a1 = ee1 - (real)ee2;
    real a2 = 1 - a1;
    if (lengy < c->xe[0]) return c->cx[r];
    if (lengy > c->xe[N - 1])
        return (real)((double)c->cx[(N - 1) * T + r] * 0.1 * dm_exp(-(double)(lengy - c->xe[N - 1])));
    if (ee2 >= N - 1) return c->cx[(N - 1) * T + r];   /* reference reads one past the table */
    return a2 * c->cx[ee2 * T + r] + a1 * c->cx[(ee2 + 1) * T + r];
}

static void anewvel(real energy, real vel, real *nvx, real *nvy, real *nvz, real r1, real r2) {
    real coschi;
    if (energy < (real)1e-30) coschi = 1;
    else coschi = (energy + 2 - 2 * R_POW(energy + 1, r1)) / energy;
    real sinchi = RSQRT(RABS(1 - coschi * coschi));
    real phi1 = TWO_PI_R * r2;
    real cosphi, sinphi;
    R_SINCOS(phi1, &sinphi, &cosphi);
    real r13 = *nvx, r23 = *nvy, r33 = *nvz, up1, up2, up3;
    if (r33 == 1) { up1 = 0; up2 = 1; up3 = 0; } else { up1 = 0; up2 = 0; up3 = 1; }
    real r12 = r23 * up3 - r33 * up2;
    real r22 = r33 * up1 - r13 * up3;
    real r32 = r13 * up2 - r23 * up1;
    real mag = RSQRT(r12 * r12 + r22 * r22 + r32 * r32);
    r12 /= mag; r22 /= mag; r32 /= mag;
    real r11 = r22 * r33 - r32 * r23;
    real r21 = r32 * r13 - r12 * r33;
    real r31 = r12 * r23 - r22 * r13;
    *nvx = vel * (r11 * sinchi * cosphi + r12 * sinchi * sinphi + r13 * coschi);
    *nvy = vel * (r21 * sinchi * cosphi + r22 * sinchi * sinphi + r23 * coschi);
    *nvz = vel * (r31 * sinchi * cosphi + r32 * sinchi * sinphi + r33 * coschi);
}

static void newvel_ionsc(real *vx, real *vy, real *vz, real vel, real r1, real r2) {
    real coschi = RSQRT(r1);
    real sinchi = RSQRT(RABS(1 - coschi * coschi));
    real phi1 = TWO_PI_R * r2;
    real cosphi, sinphi;
    R_SINCOS(phi1, &sinphi, &cosphi);
    real r13 = *vx / vel, r23 = *vy / vel, r33 = *vz / vel, up1, up2, up3;
    if (r33 == 1) { up1 = 0; up2 = 1; up3 = 0; } else { up1 = 0; up2 = 0; up3 = 1; }
    real r12 = r23 * up3 - r33 * up2;
    real r22 = r33 * up1 - r13 * up3;
    real r32 = r13 * up2 - r23 * up1;
    real mag = RSQRT(r12 * r12 + r22 * r22 + r32 * r32);
    r12 /= mag; r22 /= mag; r32 /= mag;
    real r11 = r22 * r33 - r32 * r23;
    real r21 = r32 * r13 - r12 * r33;
    real r31 = r12 * r23 - r22 * r13;
    *vx = vel * coschi * (r11 * sinchi * cosphi + r12 * sinchi * sinphi + r13 * coschi);
    *vy = vel * coschi * (r21 * sinchi * cosphi + r22 * sinchi * sinphi + r23 * coschi);
    *vz = vel * coschi * (r31 * sinchi * cosphi + r32 * sinchi * sinphi + r33 * coschi);
}

static void dmaxwellv(real *vx, real *vy, real *vz, real vsaven, real vti, real Rphi, real Rthe) {
    real aphi = TWO_PI_R * Rphi;
    real costheta = 1 - 2 * Rthe;
    real sintheta = RSQRT(1 - costheta * costheta);
    real cosa, sina;
    R_SINCOS(aphi, &sina, &cosa);
    *vx = vti * vsaven * sintheta * cosa;
    *vy = vti * vsaven * sintheta * sina;
    *vz = vti * vsaven * costheta;
}

#define U(st) ((real)orc_uniform(st))
#define MCC_TAG(isp) ((uint32_t)(((isp) << 8) | 1u))

typedef struct { real x, y, vx, vy, vz; int cell; } orc_prod;

/* Growable product lists of one MCC pass (electrons and ions pair up). */
typedef struct { orc_prod *e, *i; int n, cap; } orc_prods;

static void prods_init(orc_prods *P, int cap) {
    P->cap = cap < 16 ? 16 : cap;
    P->n = 0;
    P->e = zalloc(P->cap, sizeof(orc_prod));
    P->i = zalloc(P->cap, sizeof(orc_prod));
}

/* Collisions of the ions and electrons of cell ID; products go to P in
 * parent order.  Touches only the cell's particles and its rows of mccr. */
static void mcc_cell(orc_case *c, int ID, orc_prods *P) {
    orc_species *se = &c->sp[0], *si = &c->sp[1];
    int k;
    uint64_t seed = (uint64_t)c->deck.rng_seed;
    uint32_t step = (uint32_t)c->tstep;
    int T = c->TnRct;
    {
        /* ions (Direct_Argon_ArIon) */
        real Prob = c->Pion[ID];
        for (k = si->cell_start[ID]; k < si->cell_start[ID + 1]; ++k) {
            orc_stream st; orc_stream_init(&st, seed, si->id[k], step, MCC_TAG(1));
            real R1 = U(&st);
            if (R1 > Prob) continue;
            int n = (int)((real)(c->nvel - 1) * U(&st));
            real vnx, vny, vnz;
            real rphi = U(&st), rthe = U(&st);
            dmaxwellv(&vnx, &vny, &vnz, c->vsave[n], c->BackVel1[ID], rphi, rthe);
            real VX = si->vx[k] - vnx, VY = si->vy[k] - vnz, VZ = si->vz[k] - vny;  /* y/z swap as reference */
            real dum = VX * VX + VY * VY + VZ * VZ;
            real engy = c->Escale[1] * dum;
            real vel = RSQRT(dum);
            R1 = U(&st) * c->sigv[2] / vel;
            real sum = ar_sigma(c, 5, engy);
            if (R1 <= sum) {
                si->vx[k] = vnx; si->vy[k] = vny; si->vz[k] = vnz;
                c->mccr[(size_t)ID * T + 5]++;
            } else if (R1 <= (sum += ar_sigma(c, 6, engy))) {
                real a = U(&st), b = U(&st);
                newvel_ionsc(&VX, &VY, &VZ, vel, a, b);
                si->vx[k] = VX + vnx; si->vy[k] = VY + vny; si->vz[k] = VZ + vnz;
                c->mccr[(size_t)ID * T + 6]++;
            }
        }
        /* electrons (Direct_Argon_Electron) */
        real P1 = c->P1[ID], P2 = c->P2[ID];
        for (k = se->cell_start[ID]; k < se->cell_start[ID + 1]; ++k) {
            orc_stream st; orc_stream_init(&st, seed, se->id[k], step, MCC_TAG(0));
            int coll = 0, j;
            real R1 = 0;
            for (j = 0; j < c->MCCn; ++j) { R1 = U(&st); if (R1 < P2) { coll = 1; break; } }
            if (!coll) continue;
            int target = (R1 <= P1) ? 0 : 1;
            real VX = se->vx[k], VY = se->vy[k], VZ = se->vz[k];
            real dum = VX * VX + VY * VY + VZ * VZ;
            real vel = RSQRT(dum), vel2 = 0, rengy = 0;
            VX /= vel; VY /= vel; VZ /= vel;
            real engy = c->Escale[0] * dum;
            real R2, sum;
            int type = 1;
            if (target == 0) {
                R2 = U(&st) * c->sigv[0] / vel;
                sum = ar_sigma(c, 0, engy);
                if (R2 <= sum) { type = 0; c->mccr[(size_t)ID * T + 0]++; }
                else if (engy > c->Th_e[1] && R2 <= (sum += ar_sigma(c, 1, engy))) {
                    type = 0; engy -= c->Th_e[1]; vel = RSQRT(RABS(engy) / c->Escale[0]); c->mccr[(size_t)ID * T + 1]++;
                } else if (engy > c->Th_e[2] && R2 <= (sum += ar_sigma(c, 2, engy))) {
                    type = 0; engy -= c->Th_e[2]; vel = RSQRT(RABS(engy) / c->Escale[0]); c->mccr[(size_t)ID * T + 2]++;
                } else if (engy > c->Th_e[3] && R2 <= (sum += ar_sigma(c, 3, engy))) {
                    type = 2; engy -= c->Th_e[3];
                    real u = U(&st);
                    rengy = (real)10 * R_TAN(u * R_ATAN(engy / (real)20));
                    engy -= rengy;
                    vel = RSQRT(RABS(rengy) / c->Escale[0]);
                    vel2 = RSQRT(RABS(engy) / c->Escale[0]);
                    c->mccr[(size_t)ID * T + 3]++;
                }
            } else {
                R2 = U(&st) * c->sigv[1] / vel;
                sum = ar_sigma(c, 4, engy);
                if (engy > c->Th_e[4] && R2 <= sum) {
                    type = 2; engy -= c->Th_e[4];
                    real u = U(&st);
                    rengy = (real)10 * R_TAN(u * R_ATAN(engy / (real)20));
                    engy -= rengy;
                    vel = RSQRT(RABS(rengy) / c->Escale[0]);
                    vel2 = RSQRT(RABS(engy) / c->Escale[0]);
                    c->mccr[(size_t)ID * T + 4]++;
                }
            }
            if (type == 0) {
                real a = U(&st), b = U(&st);
                anewvel(engy, vel, &VX, &VY, &VZ, a, b);
                se->vx[k] = VX; se->vy[k] = VY; se->vz[k] = VZ;
            } else if (type == 2) {
                orc_prod *e = &P->e[P->n], *ion = &P->i[P->n];
                P->n++;
                e->cell = ID; e->x = se->x[k]; e->y = se->y[k];
                e->vx = VX; e->vy = VY; e->vz = VZ;
                real a = U(&st), b = U(&st);
                anewvel(rengy, vel, &e->vx, &e->vy, &e->vz, a, b);
                ion->cell = ID; ion->x = se->x[k]; ion->y = se->y[k];
                int n = (int)((real)(c->nvel - 1) * U(&st));
                real rphi = U(&st), rthe = U(&st);
                dmaxwellv(&ion->vx, &ion->vy, &ion->vz, c->vsave[n], c->BackVel1[ID], rphi, rthe);
                a = U(&st); b = U(&st);
                anewvel(engy, vel2, &VX, &VY, &VZ, a, b);
                se->vx[k] = VX; se->vy[k] = VY; se->vz[k] = VZ;
            }
            if (P->n >= P->cap - 1) {
                P->cap *= 2;
                P->e = realloc(P->e, P->cap * sizeof(orc_prod)); P->i = realloc(P->i, P->cap * sizeof(orc_prod));
            }
        }
    }
}

/* The species' second array set (allocated on first use). */
static void spare_get(orc_species *s, real **x, real **y, real **vx, real **vy, real **vz, uint32_t **id) {
    if (!s->sx) {
        s->sx = malloc(s->cap * sizeof(real)); s->sy = malloc(s->cap * sizeof(real));
        s->svx = malloc(s->cap * sizeof(real)); s->svy = malloc(s->cap * sizeof(real));
        s->svz = malloc(s->cap * sizeof(real)); s->sid = malloc(s->cap * sizeof(uint32_t));
    }
    *x = s->sx; *y = s->sy; *vx = s->svx; *vy = s->svy; *vz = s->svz; *id = s->sid;
}

/* makes the second set current and the current one the second */
static void spare_swap(orc_species *s) {
    real *t;
    uint32_t *u;
    t = s->x; s->x = s->sx; s->sx = t;
    t = s->y; s->y = s->sy; s->sy = t;
    t = s->vx; s->vx = s->svx; s->svx = t;
    t = s->vy; s->vy = s->svy; s->svy = t;
    t = s->vz; s->vz = s->svz; s->svz = t;
    u = s->id; s->id = s->sid; s->sid = u;
}

/* Appends products to each bucket in parent order; returns 0 or -30 on
 * capacity overflow.  reuse: rebuild into the species' second array set
 * instead of fresh arrays. */
static int mcc_append(orc_case *c, orc_prod *pe, orc_prod *pi, int npe, int npi, int reuse) {
    orc_species *se = &c->sp[0], *si = &c->sp[1];
    int G = c->Gsize;
    /* append products: bucket d = old bucket ++ products of d (parent order);
     * new ids in global (cell, parent) order */
    orc_species *sps[2] = {se, si};
    orc_prod *lists[2] = {pe, pi};
    int nps[2] = {npe, npi}, s2, rc = 0;
    for (s2 = 0; s2 < 2; ++s2) {
        orc_species *s = sps[s2];
        int np = nps[s2];
        if (s->n + np > s->cap) { rc = -30; continue; }
        if (np == 0) continue;   /* buckets unchanged */
        int *pc = zalloc(G, sizeof(int)), d, q;
        for (q = 0; q < np; ++q) pc[lists[s2][q].cell]++;
        int *start = zalloc(G + 1, sizeof(int));
        for (d = 0; d < G; ++d) start[d + 1] = start[d] + (s->cell_start[d + 1] - s->cell_start[d]) + pc[d];
        int tot = start[G];
        real *nx, *ny, *nvx, *nvy, *nvz;
        uint32_t *nid;
        if (reuse) spare_get(s, &nx, &ny, &nvx, &nvy, &nvz, &nid);
        else {
            nx = zalloc(s->cap, sizeof(real)); ny = zalloc(s->cap, sizeof(real));
            nvx = zalloc(s->cap, sizeof(real)); nvy = zalloc(s->cap, sizeof(real)); nvz = zalloc(s->cap, sizeof(real));
            nid = zalloc(s->cap, sizeof(uint32_t));
        }
        int *fill = zalloc(G, sizeof(int));
#pragma omp parallel for schedule(dynamic, 64)
        for (d = 0; d < G; ++d) {
            int o = s->cell_start[d], m = s->cell_start[d + 1] - o, r;
            for (r = 0; r < m; ++r) {
                int pos = start[d] + r;
                nx[pos] = s->x[o + r]; ny[pos] = s->y[o + r];
                nvx[pos] = s->vx[o + r]; nvy[pos] = s->vy[o + r]; nvz[pos] = s->vz[o + r]; nid[pos] = s->id[o + r];
            }
            fill[d] = start[d] + m;
        }
        for (q = 0; q < np; ++q) {     /* lists are already in (cell, parent) order */
            orc_prod *p = &lists[s2][q];
            int pos = fill[p->cell]++;
            nx[pos] = p->x; ny[pos] = p->y; nvx[pos] = p->vx; nvy[pos] = p->vy; nvz[pos] = p->vz;
            nid[pos] = s->next_id++;
        }
        if (reuse) spare_swap(s);
        else {
            free(s->x); free(s->y); free(s->vx); free(s->vy); free(s->vz); free(s->id);
            s->x = nx; s->y = ny; s->vx = nvx; s->vy = nvy; s->vz = nvz; s->id = nid;
        }
        memcpy(s->cell_start, start, (G + 1) * sizeof(int));
        s->n = tot;
        free(pc); free(start); free(fill);
    }
    return rc;
}

int orc_mcc(orc_case *c) {
    orc_prods P;
    int ID, rc;
    prods_init(&P, c->sp[0].n + 16);
    for (ID = 0; ID < c->Gsize; ++ID) mcc_cell(c, ID, &P);
    rc = mcc_append(c, P.e, P.i, P.n, P.n, 0);
    free(P.e); free(P.i);
    return rc;
}

/* ================= deposit: src/cuda_Deposit.cu:3-104, 277-326 ============= */
void orc_deposit(orc_case *c) {
    int G = c->Gsize, ngy = c->ngy, ngx = c->ngx, isp, cell, k, g, i;
    for (g = 0; g < c->nsp * G; ++g) c->den[g] = 0;
    for (isp = 0; isp < c->nsp; ++isp) {
        orc_species *s = &c->sp[isp];
        real *den = c->den + (size_t)isp * G;
        for (cell = 0; cell < G; ++cell) {
            int k0 = s->cell_start[cell], k1 = s->cell_start[cell + 1];
            if (k0 == k1) continue;
            real WS = 0, WN = 0, ES = 0, EN = 0;
            for (k = k0; k < k1; ++k) {
                real lx = s->x[k], ly = s->y[k];
                WS += (1 - lx) * (1 - ly);
                WN += (1 - lx) * ly;
                ES += lx * (1 - ly);
                EN += lx * ly;
            }
            den[cell] += WS; den[cell + 1] += WN; den[cell + ngy] += ES; den[cell + ngy + 1] += EN;
        }
        for (g = 0; g < G; ++g) {     /* DepositBoundary */
            int f = c->Face[g];
            if (f == ORC_LEFT || f == ORC_RIGHT || f == ORC_UP || f == ORC_DOWN) den[g] *= 2;
            else if (f == ORC_UL || f == ORC_LL || f == ORC_UR || f == ORC_LR) den[g] *= 4;
            if (c->Boundary[g] == ORC_NEUMANN) den[g] *= 2;
            else if (c->Boundary[g] == ORC_DIRICHLET && f != ORC_NOFACE && c->CondID[g] != 0) den[g] *= 2;
        }
    }
    /* SumSource; the reference leaves SD uninitialised on wall nodes whose
     * inward neighbour is neither dielectric nor conductor - read as 0 here */
    for (g = 0; g < G; ++g) {
        int x = g / ngy, y = g % ngy, nb = -1;
        real SD = 0, ss = 0;
        if (x == 0) nb = g + ngy;
        else if (x == ngx - 1) nb = g - ngy;
        else if (y == 0) nb = g + 1;
        else if (y == ngy - 1) nb = g - 1;
        int bnd = nb >= 0 ? c->Boundary[nb] : c->Boundary[g];
        real fac = nb >= 0 ? 2 : 1;
        if (bnd == ORC_DIELECTRIC) {
            for (i = 0; i < c->nsp; ++i) ss += c->sigma_sp[(size_t)i * G + g] * c->q_density[i];
            c->Sigma[g] = fac * ss / c->Area[g];
            SD = ss;
        } else if (bnd == ORC_CONDUCTOR) {
            for (i = 0; i < c->nsp; ++i) c->sigma_sp[(size_t)i * G + g] = 0;
            SD = 0;
        }
        ss = 0;
        for (i = 0; i < c->nsp; ++i) ss += c->den[(size_t)i * G + g] * c->q_density[i];
        c->Source[g] = (real)((double)(ss + SD) / EPS0);
    }
    for (g = 0; g < G; ++g) if (c->A_idx[g]) c->b[c->A_idx[g] - 1] = c->Source[g];
}

/* ================= accumulators: src/cuda_Diagnostic.cu:505-547 (FP64) ===== */
void orc_accumulate(orc_case *c) {
    int g, G = c->Gsize;
    for (g = 0; g < c->nsp * G; ++g) c->sum_den[g] += (double)c->den[g];
    for (g = 0; g < G; ++g) {
        c->sum_phi[g] += (double)c->TotPot[g];
        c->sum_ex[g] += (double)c->Ex[g];
        c->sum_ey[g] += (double)c->Ey[g];
    }
    c->n_accum++;
}

void orc_fluid_step(orc_case *c);

/* ================= one timestep: src/cuda_main.cu:112-235 ================= */
int orc_step(orc_case *c) {
    if (c->deck.gas != 0) return -19;   /* oxygen gases: fluid only (their collision step is orc_mcc_o2) */
    c->tstep++;
    c->t = (double)c->tstep * (double)c->dt;
    orc_field(c);
    orc_efield(c);
    orc_move(c);
    int e = orc_sort(c);
    if (e) return e;
    e = orc_mcc(c);
    if (e) return e;
    orc_deposit(c);
    if (c->fluid_on) orc_fluid_step(c);
    orc_accumulate(c);
    return 0;
}

int orc_run(orc_case *c, int nsteps) {
    int i, e;
    for (i = 0; i < nsteps; ++i) { e = orc_step(c); if (e) return e; }
    return 0;
}

/* ================= metastable fluid step: src/cuda_Fluid.cu:34-178,
 *                   src/cuda_Diagnostic.cu:37-110, 534-560, 634-642 ====== */
/* Calculate_Flux_x / _y (src/cuda_Fluid.cu:51-90) */
void orc_fluid_flux_x(orc_case *c, int f) {
    int k, i, ncy = c->ncy;
    size_t NX = (size_t)c->ngx * ncy, NC = (size_t)c->ncx * ncy;
    const real *n = c->fden + f * NC, *ax = c->fax + f * NX, *bx = c->fbx + f * NX;
    real *fx = c->fflx + f * NX;
    for (k = 0; k < c->nlx; k++) {
        int x1 = c->lx_x1[k], x2 = c->lx_x2[k], yy = c->lx_y[k];
        fx[x1 * ncy + yy] = -c->lx_f1[f * c->nlx + k] * bx[x1 * ncy + yy] * n[x1 * ncy + yy];
        for (i = x1 + 1; i <= x2; i++) {
            real left = n[(i - 1) * ncy + yy], right = n[i * ncy + yy];
            fx[i * ncy + yy] = ax[i * ncy + yy] * left - bx[i * ncy + yy] * right;
        }
        fx[(x2 + 1) * ncy + yy] = c->lx_f2[f * c->nlx + k] * ax[(x2 + 1) * ncy + yy] * n[x2 * ncy + yy];
    }
}
void orc_fluid_flux_y(orc_case *c, int f) {
    int k, j, ncy = c->ncy, ngy = c->ngy;
    size_t NY = (size_t)c->ncx * ngy, NC = (size_t)c->ncx * ncy;
    const real *n = c->fden + f * NC, *ay = c->fay + f * NY, *by = c->fby + f * NY;
    real *fy = c->ffly + f * NY;
    for (k = 0; k < c->nly; k++) {
        int xx = c->ly_x[k], y1 = c->ly_y1[k], y2 = c->ly_y2[k];
        fy[xx * ngy + y1] = -c->ly_f1[f * c->nly + k] * by[xx * ngy + y1] * n[xx * ncy + y1];
        for (j = y1 + 1; j <= y2; j++) {
            real lower = n[xx * ncy + j - 1], upper = n[xx * ncy + j];
            fy[xx * ngy + j] = ay[xx * ngy + j] * lower - by[xx * ngy + j] * upper;
        }
        fy[xx * ngy + y2 + 1] = c->ly_f2[f * c->nly + k] * ay[xx * ngy + y2 + 1] * n[xx * ncy + y2];
    }
}

/* tridiag (src/cuda_Fluid.cu:164-178): Thomas algorithm; 0 when a pivot
 * vanishes (the reference then skips the back substitution) */
int orc_tridiag(const real *a, const real *b, const real *cc, const real *d, real *x, real *gam, int n) {
    int j;
    real bet;
    x[0] = d[0] / (bet = b[0]);
    for (j = 1; j < n; j++) {
        gam[j - 1] = cc[j - 1] / bet;
        bet = b[j] - a[j] * cc[j - 1] / bet;
        x[j] = (d[j] - a[j] * x[j - 1]) / bet;
        if (bet == 0) return 0;
    }
    for (j = n - 2; j >= 0; j--) x[j] -= gam[j] * x[j + 1];
    return 1;
}

/* Solve_Density_x / _y (src/cuda_Fluid.cu:91-163).  A line whose system
 * meets a zero pivot keeps its density (the reference reads uninitialised
 * values there). */
static void fluid_solve_x(orc_case *c, int f) {
    int k, i, ncy = c->ncy, ngy = c->ngy;
    size_t NX = (size_t)c->ngx * ncy, NY = (size_t)c->ncx * ngy, NC = (size_t)c->ncx * ncy;
    real *n = c->fden + f * NC;
    const real *ax = c->fax + f * NX, *bx = c->fbx + f * NX, *fy = c->ffly + f * NY, *src = c->fsrc + f * NC;
    real del_f = c->dtc / c->dx;
    int m = c->ngx + 1;
    real *sub = malloc(m * sizeof(real)), *dia = malloc(m * sizeof(real)), *sup = malloc(m * sizeof(real));
    real *rhs = malloc(m * sizeof(real)), *x = malloc(m * sizeof(real)), *gam = malloc(m * sizeof(real));
    for (k = 0; k < c->nlx; k++) {
        int x1 = c->lx_x1[k], x2 = c->lx_x2[k], yy = c->lx_y[k], num = x2 - x1 + 1;
        for (i = 0; i < num; i++) {
            int xx = i + x1;
            sub[i] = -ax[xx * ncy + yy] * del_f;
            dia[i] = 1 + (ax[(xx + 1) * ncy + yy] + bx[xx * ncy + yy]) * del_f;
            sup[i] = -bx[(xx + 1) * ncy + yy] * del_f;
            rhs[i] = n[xx * ncy + yy] + src[xx * ncy + yy] * c->dtc - (fy[xx * ngy + yy + 1] - fy[xx * ngy + yy]) * c->dtc / c->dy;
        }
        if (!orc_tridiag(sub, dia, sup, rhs, x, gam, num)) continue;
        for (i = 0; i < num; i++)
            n[(i + x1) * ncy + yy] = (c->PlasmaRegion[(i + x1) * ncy + yy] == 1 && x[i] > 1e9) ? x[i] : 0;
    }
    free(sub); free(dia); free(sup); free(rhs); free(x); free(gam);
}
static void fluid_solve_y(orc_case *c, int f) {
    int k, j, ncy = c->ncy, ngy = c->ngy;
    size_t NX = (size_t)c->ngx * ncy, NY = (size_t)c->ncx * ngy, NC = (size_t)c->ncx * ncy;
    real *n = c->fden + f * NC;
    const real *ay = c->fay + f * NY, *by = c->fby + f * NY, *fx = c->fflx + f * NX, *src = c->fsrc + f * NC;
    real del_f = c->dtc / c->dy;
    int m = c->ngy + 1;
    real *sub = malloc(m * sizeof(real)), *dia = malloc(m * sizeof(real)), *sup = malloc(m * sizeof(real));
    real *rhs = malloc(m * sizeof(real)), *x = malloc(m * sizeof(real)), *gam = malloc(m * sizeof(real));
    for (k = 0; k < c->nly; k++) {
        int xx = c->ly_x[k], y1 = c->ly_y1[k], y2 = c->ly_y2[k], num = y2 - y1 + 1;
        for (j = 0; j < num; j++) {
            int yy = j + y1;
            sub[j] = -ay[xx * ngy + yy] * del_f;
            dia[j] = 1 + (ay[xx * ngy + yy + 1] + by[xx * ngy + yy]) * del_f;
            sup[j] = -by[xx * ngy + yy + 1] * del_f;
            rhs[j] = n[xx * ncy + yy] + src[xx * ncy + yy] * c->dtc - (fx[(xx + 1) * ncy + yy] - fx[xx * ncy + yy]) * c->dtc / c->dx;
        }
        if (!orc_tridiag(sub, dia, sup, rhs, x, gam, num)) continue;
        for (j = 0; j < num; j++)
            n[xx * ncy + j + y1] = (c->PlasmaRegion[xx * ncy + j + y1] == 1 && x[j] > 1e9) ? x[j] : 0;
    }
    free(sub); free(dia); free(sup); free(rhs); free(x); free(gam);
}

/* Solve_Continuity_eqn (src/cuda_Fluid.cu:34-42) */
void orc_fluid_continuity(orc_case *c) {
    int f;
    for (f = 0; f < c->nfsp; f++) {
        fluid_solve_x(c, f);
        orc_fluid_flux_x(c, f);
        fluid_solve_y(c, f);
        orc_fluid_flux_y(c, f);
    }
}

/* Averaged event rates (src/cuda_Diagnostic.cu:548-633): reaction r's count
 * over the window, times Denscale of the species whose collisions count it
 * over N_ave dt.  rate_range covers reactions r0..r1-1 of one species
 * ((Denscale / N_ave) / dt, then times the count); rate_late is the form the
 * reference uses for its last few entries (Denscale times (1 / N_ave / dt),
 * then times the count). */
static void rate_range(orc_case *c, real *M, int g, int r0, int r1, int sp) {
    int r, T = c->TnRct;
    real t = c->Denscale[sp] / (real)c->deck.n_ave / c->dt;
    for (r = r0; r < r1; r++) M[r] = t * (real)(c->mccr[(size_t)g * T + r] - c->mccr_mark[(size_t)g * T + r]);
}
static void rate_late(orc_case *c, real *M, int g, int r, int sp) {
    int T = c->TnRct;
    real t = (real)1 / (real)c->deck.n_ave / c->dt;
    M[r] = c->Denscale[sp] * t * (real)(c->mccr[(size_t)g * T + r] - c->mccr_mark[(size_t)g * T + r]);
}

/* The fluid sources of one node, per fluid species (src/cuda_Diagnostic.cu:
 * 634-744).  M holds the averaged rates; the rate-constant reactions are
 * set here from the node's fluid (last sync) and background densities, then
 * each species' source is its production minus its loss as the reference
 * sums them.  Eden = 0.25 x the averaged electron density. */
static void fluid_sources(orc_case *c, int g, real *M, real Eden, real *S) {
    int G = c->Gsize;
    const real *F = c->FluidN;
    real B1 = c->BackDen1[g];
    if (c->deck.gas == 0) {
        /* Ar*: S = R2 - R4 - k7 n_e n - 2 R8 - k9 n_Ar n */
        real nm = F[g];
        real k7 = c->RR[7] * Eden * nm;
        real k9 = c->RR[9] * B1 * nm;
        S[0] = M[2] - M[4] - k7 - 2 * M[8] - k9;
    } else if (c->deck.gas == 1) {
        /* O2A, O2B, OP, OD (src/cuda_Diagnostic.cu:643-683) */
        real sum, diff;
        M[59] = c->RR[59] * F[2 * G + g] * F[3 * G + g];
        M[60] = c->RR[60] * F[3 * G + g] * B1;
        M[61] = c->RR[61] * F[3 * G + g] * B1;
        M[62] = c->RR[62] * F[3 * G + g] * B1;
        M[63] = c->RR[63] * F[g] * F[2 * G + g];
        M[64] = c->RR[64] * F[g] * B1;
        M[65] = c->RR[65] * F[g] * F[g];
        M[66] = c->RR[66] * F[G + g] * B1;
        sum = M[3] + M[61];
        diff = M[14] + M[15] + M[16] + M[17] + M[18] + M[19] + M[20] + M[21] + M[46] + M[51] + M[56] + M[63] + M[64]
             + 2 * M[65];
        S[0] = sum - diff;
        sum = M[4] + M[62];
        diff = M[22] + M[23] + M[24] + M[25] + M[26] + M[27] + M[28] + M[29] + M[52] + M[57] + M[58] + M[66];
        S[1] = sum - diff;
        sum = M[6] + 2 * M[7] + M[8] + M[11] + M[13] + M[15] + 2 * M[18] + M[19] + M[21] + M[23] + 2 * M[26] + M[27]
            + M[29] + M[30] + M[31] + M[40] + M[42] + M[44] + 2 * M[45] + M[46] + M[50] + M[53] + M[56] + M[57]
            + M[58] + M[59] + M[60] + M[61] + M[62];
        diff = M[33] + M[34] + M[35] + M[36] + M[37] + M[38] + M[43] + M[47];
        S[2] = sum - diff;
        sum = M[8] + 2 * M[9] + M[19] + 2 * M[20] + M[27] + 2 * M[28] + M[31] + M[33];
        diff = M[39] + M[40] + M[59] + M[60] + M[61] + M[62];
        S[3] = sum - diff;
    } else {
        /* Ar*, O2A, O2B, OP, OD (src/cuda_Diagnostic.cu:684-744) */
        real sum, diff, B2 = c->BackDen2[g];
        real N1 = F[g], N2 = F[G + g], N3 = F[2 * G + g], N4 = F[3 * G + g], N5 = F[4 * G + g];
        M[68] = c->RR[68] * Eden * N1;
        M[71] = c->RR[71] * N4 * N5;
        M[72] = c->RR[72] * N5 * B2;
        M[73] = c->RR[73] * N5 * B2;
        M[74] = c->RR[74] * N5 * B2;
        M[75] = c->RR[75] * N2 * N4;
        M[76] = c->RR[76] * N2 * B2;
        M[77] = c->RR[77] * N2 * N2;
        M[78] = c->RR[78] * N3 * B2;
        M[82] = c->RR[82] * N1 * B1;
        M[83] = c->RR[83] * N1 * N4;
        M[84] = c->RR[84] * N1 * N4;
        M[85] = c->RR[85] * N1 * B2;
        M[86] = c->RR[86] * N1 * B2;
        M[87] = c->RR[87] * N1 * B2;
        sum = M[2];
        diff = M[4] + M[68] + 2 * M[81] + M[82] + M[83] + M[84] + M[85] + M[86] + M[87];
        S[0] = sum - diff;
        sum = M[8] + M[73];
        diff = M[19] + M[20] + M[21] + M[22] + M[23] + M[24] + M[25] + M[26] + M[51] + M[56] + M[30] + M[75] + M[76]
             + 2 * M[77];
        S[1] = sum - diff;
        sum = M[9] + M[74];
        diff = M[27] + M[28] + M[29] + M[30] + M[31] + M[32] + M[33] + M[34] + M[57] + M[64] + M[70] + M[78];
        S[2] = sum - diff;
        sum = M[11] + 2 * M[12] + M[13] + M[16] + M[18] + M[20] + 2 * M[23] + M[24] + M[26] + M[28] + 2 * M[31]
            + M[32] + M[34] + M[35] + M[36] + M[45] + M[47] + M[49] + 2 * M[50] + M[51] + M[55] + M[60] + M[63]
            + M[64] + M[70] + M[71] + M[72] + M[73] + M[74] + M[69] + 2 * M[85] + M[86];
        diff = M[38] + M[39] + M[40] + M[41] + M[42] + M[43] + M[48] + M[52] + M[79] + M[83];
        S[3] = sum - diff;
        sum = M[13] + 2 * M[14] + M[24] + 2 * M[25] + M[32] + 2 * M[33] + M[36] + M[38] + M[83] + M[86];
        diff = M[44] + M[45] + M[71] + M[72] + M[73] + M[74];
        S[4] = sum - diff;
    }
}

/* The fluid part of the diagnostic (src/cuda_Diagnostic.cu:44-110,
 * 534-744): every step adds the electron node density to the window; the
 * (N_ave+1)-th call averages it and the collision rates since the last
 * average (per gas: argon 0-6 and 8 by the electron Denscale; O2 0-40 e,
 * 41-46 O-, 47-52 O2+, 53-57 O+, 58 O-; Ar/O2 0-45 e, 46-51 O-, 52-59 O2+,
 * 60-64 O+, 65-67 Ar+, 69-70 O+, 79-80 Ar+, 81 e), computes every fluid
 * species' source per node (fluid_sources), hands it to the cells and
 * switches the continuity solve on; once on, the cell densities of every
 * fluid species go to the nodes, and for argon the electron null-collision
 * probability P2 follows the Ar* density (the oxygen collision steps read
 * the node densities themselves). */
void orc_fluid_diag(orc_case *c) {
    int g, r, f, G = c->Gsize, T = c->TnRct, ngy = c->ngy, ncy = c->ncy, nf = c->nfsp, gas = c->deck.gas;
    int avg = c->nave_count == c->deck.n_ave;
    size_t NC = (size_t)c->ncx * ncy;
    for (g = 0; g < G; g++) {
        c->win_den[g] += c->den[g];
        if (!avg) continue;
        real ave_e = c->Denscale[0] / (real)c->deck.n_ave * c->win_den[g];
        real M[ORC_MAXRCT], S[ORC_MAXFSP];
        for (r = 0; r < T; r++) M[r] = 0;
        if (gas == 0) {
            rate_range(c, M, g, 0, 7, 0);
            rate_range(c, M, g, 8, 9, 0);
        } else if (gas == 1) {
            rate_range(c, M, g, 0, 41, 0);
            rate_range(c, M, g, 41, 47, 3);
            rate_range(c, M, g, 47, 53, 1);
            rate_range(c, M, g, 53, 58, 2);
            rate_late(c, M, g, 58, 3);
        } else {
            rate_range(c, M, g, 0, 46, 0);
            rate_range(c, M, g, 46, 52, 4);
            rate_range(c, M, g, 52, 60, 2);
            rate_range(c, M, g, 60, 65, 3);
            rate_range(c, M, g, 65, 68, 1);
            rate_late(c, M, g, 69, 3);
            rate_late(c, M, g, 70, 3);
            rate_late(c, M, g, 79, 1);
            rate_late(c, M, g, 80, 1);
            rate_late(c, M, g, 81, 0);
        }
        for (r = 0; r < T; r++) c->mccr_mark[(size_t)g * T + r] = c->mccr[(size_t)g * T + r];
        c->win_den[g] = 0;
        fluid_sources(c, g, M, (real)(0.25 * ave_e), S);
        int x = g / ngy, y = g % ngy;
        if (x < c->ncx && y < ncy)
            for (f = 0; f < nf; f++) c->fsrc[f * NC + x * ncy + y] = S[f];
    }
    if (avg) { c->nave_count = 0; c->conti_flag = 1; }
    else c->nave_count++;
    if (!c->conti_flag) return;
    /* Sync_Fluid_GFCtoGFG_forDen (src/cuda_Fluid.cu:302-320) and the P2 of
     * the argon electron null test (src/cuda_mccAr.cu:57-58) */
    {
        int i, j;
        for (f = 0; f < nf; f++)
            for (i = 0; i < c->ncx; i++)
                for (j = 0; j < ncy; j++) {
                    real v = c->fden[f * NC + i * ncy + j];
                    c->FluidN[(size_t)f * G + i * ngy + j] = v;
                    if (i == c->ncx - 1 || j == ncy - 1) c->FluidN[(size_t)f * G + (i + 1) * ngy + j + 1] = v;
                }
        if (gas == 0) {
            double dtm = (double)c->dt_mcc;
            for (g = 0; g < G; g++)
                c->P2[g] = c->P1[g] + (real)(1.0 - dm_exp(-dtm * (double)c->sigv[1] * (double)c->FluidN[g]));
        }
    }
}

/* one step of the fluid coupling after the deposit: the continuity solve
 * (once switched on), then its part of the diagnostic */
void orc_fluid_step(orc_case *c) {
    if (c->conti_flag) orc_fluid_continuity(c);
    orc_fluid_diag(c);
}

/* ================= OpenMP step (CPU baseline only) =========================
 * The same step with the particle loops split over cell tiles: the push of
 * every cell, the bucket rebuild of the particles that stay in their cell,
 * and the collisions of every cell run in parallel; the wall crossings (which
 * add to shared surface charge and counters) and the placement of crossers
 * and collision products run in particle order on one thread, so the result
 * is the sequential step's, bit for bit.  Built only with -fopenmp. */
#ifdef _OPENMP
#include <omp.h>

static void orc_move_par(orc_case *c) {
    int isp, ngy = c->ngy;
    for (isp = 0; isp < c->nsp; ++isp) {
        orc_species *s = &c->sp[isp];
        int *tag = c->tag + (size_t)isp * s->cap;
        real As = c->Ascale[isp];
        int cell;
#pragma omp parallel for schedule(dynamic, 64)
        for (cell = 0; cell < c->Gsize; ++cell) {
            int k0 = s->cell_start[cell], k1 = s->cell_start[cell + 1], k;
            if (k0 == k1) continue;
            real ews = c->Ex[cell], ewn = c->Ex[cell + 1], ees = c->Ex[cell + ngy], een = c->Ex[cell + ngy + 1];
            real yws = c->Ey[cell], ywn = c->Ey[cell + 1], yes = c->Ey[cell + ngy], yen = c->Ey[cell + ngy + 1];
            for (k = k0; k < k1; ++k) {
                real lx = s->x[k], ly = s->y[k];
                real dvx = ews * (1 - lx) * (1 - ly) + ewn * (1 - lx) * ly + ees * lx * (1 - ly) + een * lx * ly;
                real dvy = yws * (1 - lx) * (1 - ly) + ywn * (1 - lx) * ly + yes * lx * (1 - ly) + yen * lx * ly;
                real mx = s->vx[k] + dvx * As;
                real my = s->vy[k] + dvy * As;
                lx += mx * c->dt_dx;
                ly += my * c->dt_dy;
                int t = 0;
                if (ly >= 1 || ly < 0 || lx >= 1 || lx < 0) {
                    if (ly >= 1) { t += 1; ly -= 1; }
                    else if (ly < 0) { t -= 1; ly += 1; }
                    if (lx >= 1) { t += ngy; lx -= 1; }
                    else if (lx < 0) { t -= ngy; lx += 1; }
                    if (lx >= 1) lx = ORC_ONE_MINUS; else if (lx < 0) lx = 0;
                    if (ly >= 1) ly = ORC_ONE_MINUS; else if (ly < 0) ly = 0;
                }
                s->x[k] = lx; s->y[k] = ly; s->vx[k] = mx; s->vy[k] = my;
                tag[k] = t;
            }
        }
    }
}

static int orc_sort_par(orc_case *c) {
    int isp;
    for (isp = 0; isp < c->nsp; ++isp) {
        orc_species *s = &c->sp[isp];
        int *tag = c->tag + (size_t)isp * s->cap;
        int n = s->n, G = c->Gsize, cell, k;
        int *destof = zalloc(n, sizeof(int));
        char *cls = zalloc(n, 1);
        int *scnt = zalloc(G, sizeof(int)), *acnt = zalloc(G, sizeof(int));
        /* stayers in parallel; crossers (cls 3) resolved below in order */
#pragma omp parallel for schedule(dynamic, 64)
        for (cell = 0; cell < G; ++cell) {
            int m = 0, kk;
            for (kk = s->cell_start[cell]; kk < s->cell_start[cell + 1]; ++kk)
                if (tag[kk] == 0) { destof[kk] = cell; cls[kk] = 0; ++m; }
                else cls[kk] = 3;
            scnt[cell] = m;
        }
        for (cell = 0; cell < G; ++cell)
            for (k = s->cell_start[cell]; k < s->cell_start[cell + 1]; ++k) {
                if (cls[k] != 3) continue;
                int d;
                if (resolve_crosser(c, isp, cell, tag[k], &s->x[k], &s->y[k], &s->vx[k], &s->vy[k], &d)) {
                    destof[k] = d; cls[k] = 1; acnt[d]++;
                } else { destof[k] = -1; cls[k] = 2; }
            }
        int *start = zalloc(G + 1, sizeof(int));
        for (cell = 0; cell < G; ++cell) start[cell + 1] = start[cell] + scnt[cell] + acnt[cell];
        int total = start[G];
        real *nx, *ny, *nvx, *nvy, *nvz;
        uint32_t *nid;
        spare_get(s, &nx, &ny, &nvx, &nvy, &nvz, &nid);
#pragma omp parallel for schedule(dynamic, 64)
        for (cell = 0; cell < G; ++cell) {
            int pos = start[cell], kk;
            for (kk = s->cell_start[cell]; kk < s->cell_start[cell + 1]; ++kk) {
                if (cls[kk] != 0) continue;
                nx[pos] = s->x[kk]; ny[pos] = s->y[kk];
                nvx[pos] = s->vx[kk]; nvy[pos] = s->vy[kk]; nvz[pos] = s->vz[kk]; nid[pos] = s->id[kk];
                ++pos;
            }
        }
        int *afill = zalloc(G, sizeof(int));
        for (k = 0; k < n; ++k) {
            if (cls[k] != 1) continue;
            int d = destof[k], pos = start[d] + scnt[d] + afill[d]++;
            nx[pos] = s->x[k]; ny[pos] = s->y[k];
            nvx[pos] = s->vx[k]; nvy[pos] = s->vy[k]; nvz[pos] = s->vz[k]; nid[pos] = s->id[k];
        }
        spare_swap(s);
        memcpy(s->cell_start, start, (G + 1) * sizeof(int));
        s->n = total;
        free(destof); free(cls); free(scnt); free(acnt); free(start); free(afill);
#pragma omp parallel for schedule(static)
        for (k = 0; k < n; ++k) tag[k] = 0;   /* the push tags only [0, n) */
    }
    return 0;
}

/* collisions over contiguous cell tiles, one per thread; each tile's
 * products are in (cell, parent) order, so the tiles concatenated in order
 * give the sequential list */
static int orc_mcc_par(orc_case *c) {
    int nt = omp_get_max_threads(), t, G = c->Gsize, rc;
    orc_prods *P = zalloc(nt, sizeof(orc_prods));
#pragma omp parallel num_threads(nt)
    {
        int me = omp_get_thread_num(), ID;
        int a = (int)((long)G * me / nt), b = (int)((long)G * (me + 1) / nt);
        prods_init(&P[me], 64);
        for (ID = a; ID < b; ++ID) mcc_cell(c, ID, &P[me]);
    }
    int n = 0;
    for (t = 0; t < nt; ++t) n += P[t].n;
    orc_prod *pe = zalloc(n + 1, sizeof(orc_prod)), *pi = zalloc(n + 1, sizeof(orc_prod));
    for (t = 0, n = 0; t < nt; ++t) {
        memcpy(pe + n, P[t].e, P[t].n * sizeof(orc_prod));
        memcpy(pi + n, P[t].i, P[t].n * sizeof(orc_prod));
        n += P[t].n;
        free(P[t].e); free(P[t].i);
    }
    free(P);
    rc = mcc_append(c, pe, pi, n, n, 1);
    free(pe); free(pi);
    return rc;
}

/* nsteps of the argon step with `threads` OpenMP threads (0: the runtime's
 * default); returns 0, or the first step's error */
int orc_run_par(orc_case *c, int nsteps, int threads) {
    int i, e;
    if (threads > 0) omp_set_num_threads(threads);
    for (i = 0; i < nsteps; ++i) {
        c->tstep++;
        c->t = (double)c->tstep * (double)c->dt;
        orc_field(c);
        orc_efield(c);
        orc_move_par(c);
        if ((e = orc_sort_par(c))) return e;
        if ((e = orc_mcc_par(c))) return e;
        orc_deposit(c);
        orc_accumulate(c);
    }
    return 0;
}

int orc_threads(void) { return omp_get_max_threads(); }
#endif

/* ================= accessors for the Python test harness ================= */
typedef struct { const char *name; void *ptr; long count; int elsize; } orc_arr;

static int lookup(orc_case *c, const char *name, orc_arr *out) {
    long G = c->Gsize, C = c->Csize;
    int rs = (int)sizeof(real);
#define ENT(nm, p, cnt, es) if (!strcmp(name, nm)) { out->ptr = (void *)(p); out->count = (cnt); out->elsize = (es); return 0; }
    ENT("Boundary", c->Boundary, G, 4) ENT("CondID", c->CondID, G, 4) ENT("Face", c->Face, G, 4)
    ENT("DensRegion", c->DensRegion, G, 4) ENT("Area", c->Area, G, rs) ENT("Temp", c->Temp, G, rs)
    ENT("PlasmaRegion", c->PlasmaRegion, C, 4) ENT("eps_r", c->eps_r, C, rs) ENT("eps", c->eps, C, rs)
    ENT("StructIdx", c->StructIdx, (long)(c->ncx + 2) * (c->ncy + 2), 4)
    ENT("A_idx", c->A_idx, G, 4) ENT("Ai", c->Ai, c->A_size + 1, 4)
    ENT("Aj", c->Aj, c->A_size > 0 ? c->Ai[c->A_size] : 0, 4) ENT("Aval", c->Aval, c->A_size > 0 ? c->Ai[c->A_size] : 0, rs)
    ENT("Atemp", c->Atemp_val, c->A_size > 0 ? c->Ai[c->A_size] : 0, rs)
    ENT("Mdiag", c->Mdiag, c->A_size, rs) ENT("cond_b", c->cond_b, (long)c->CondNUMR * c->A_size, rs)
    ENT("temp_b", c->temp_b, c->A_size, rs) ENT("lap_phi", c->lap_phi, (long)c->CondNUMR * G, rs)
    ENT("BackDen1", c->BackDen1, G, rs) ENT("BackVel1", c->BackVel1, G, rs)
    ENT("BackDen2", c->BackDen2, c->BackDen2 ? G : 0, rs) ENT("BackVel2", c->BackVel2, c->BackVel2 ? G : 0, rs)
    ENT("RR", c->RR, c->TnRct, rs)
    ENT("Denscale", c->Denscale, c->nsp, rs) ENT("FluidN", c->FluidN, (long)c->nfsp * G, rs)
    ENT("xe", c->xe, c->N_LOGX, rs) ENT("xee", c->xee, c->N_LOGX, rs) ENT("cx", c->cx, (long)c->N_LOGX * c->TnRct, rs)
    ENT("sigv", c->sigv, 3, rs) ENT("P1", c->P1, G, rs) ENT("P2", c->P2, G, rs) ENT("Pion", c->Pion, G, rs)
    ENT("vsave", c->vsave, c->nvel, rs) ENT("phi", c->phi, G, rs) ENT("TotPot", c->TotPot, G, rs)
    ENT("Sigma", c->Sigma, G, rs) ENT("Source", c->Source, G, rs) ENT("Ex", c->Ex, G, rs) ENT("Ey", c->Ey, G, rs)
    ENT("den", c->den, (long)c->nsp * G, rs) ENT("sigma_sp", c->sigma_sp, (long)c->nsp * G, rs)
    ENT("b", c->b, c->A_size, rs) ENT("xsol", c->xsol, c->A_size, 8) ENT("V_cond", c->V_cond, c->CondNUMR, rs)
    ENT("CondCharge", c->CondCharge, (long)c->nsp * c->CondNUMR, rs) ENT("Lost", c->Lost, c->nsp, rs) ENT("mccr", c->mccr, G * c->TnRct, 8)
    ENT("sum_den", c->sum_den, (long)c->nsp * G, 8) ENT("sum_phi", c->sum_phi, G, 8)
    ENT("sum_ex", c->sum_ex, G, 8) ENT("sum_ey", c->sum_ey, G, 8) ENT("lu", c->lu, (long)c->A_size * (2 * c->lu_bw + 1), 8)
    {
        long NF = c->nfsp, NC = (long)c->ncx * c->ncy, NX = (long)c->ngx * c->ncy, NY = (long)c->ncx * c->ngy;
        ENT("fl_x1", c->lx_x1, c->nlx, 4) ENT("fl_x2", c->lx_x2, c->nlx, 4) ENT("fl_xy", c->lx_y, c->nlx, 4)
        ENT("fl_yx", c->ly_x, c->nly, 4) ENT("fl_y1", c->ly_y1, c->nly, 4) ENT("fl_y2", c->ly_y2, c->nly, 4)
        ENT("fl_xg1", c->lx_f1, NF * c->nlx, rs) ENT("fl_xg2", c->lx_f2, NF * c->nlx, rs)
        ENT("fl_yg1", c->ly_f1, NF * c->nly, rs) ENT("fl_yg2", c->ly_f2, NF * c->nly, rs)
        ENT("fl_D", c->fD, NF * NC, rs) ENT("fl_den", c->fden, NF * NC, rs) ENT("fl_src", c->fsrc, NF * NC, rs)
        ENT("fl_ax", c->fax, NF * NX, rs) ENT("fl_bx", c->fbx, NF * NX, rs) ENT("fl_flux_x", c->fflx, NF * NX, rs)
        ENT("fl_ay", c->fay, NF * NY, rs) ENT("fl_by", c->fby, NF * NY, rs) ENT("fl_flux_y", c->ffly, NF * NY, rs)
        ENT("fl_win", c->win_den, G, rs)
    }
    if (!strncmp(name, "sp", 2) && name[2] >= '0' && name[2] <= '9' && name[3] == '.') {
        int s = name[2] - '0';
        if (s >= c->nsp) return -1;
        orc_species *p = &c->sp[s];
        const char *f = name + 4;
        if (!strcmp(f, "x")) { out->ptr = p->x; out->count = p->cap; out->elsize = rs; return 0; }
        if (!strcmp(f, "y")) { out->ptr = p->y; out->count = p->cap; out->elsize = rs; return 0; }
        if (!strcmp(f, "vx")) { out->ptr = p->vx; out->count = p->cap; out->elsize = rs; return 0; }
        if (!strcmp(f, "vy")) { out->ptr = p->vy; out->count = p->cap; out->elsize = rs; return 0; }
        if (!strcmp(f, "vz")) { out->ptr = p->vz; out->count = p->cap; out->elsize = rs; return 0; }
        if (!strcmp(f, "id")) { out->ptr = p->id; out->count = p->cap; out->elsize = 4; return 0; }
        if (!strcmp(f, "cell_start")) { out->ptr = p->cell_start; out->count = G + 1; out->elsize = 4; return 0; }
        if (!strcmp(f, "tag")) { out->ptr = c->tag + (size_t)s * p->cap; out->count = p->cap; out->elsize = 4; return 0; }
    }
#undef ENT
    return -1;
}

void *orc_array(orc_case *c, const char *name, long *count, int *elsize) {
    orc_arr a;
    if (lookup(c, name, &a)) { *count = 0; *elsize = 0; return 0; }
    *count = a.count; *elsize = a.elsize;
    return a.ptr;
}

/* scalar getter/setter by name */
double orc_scalar(orc_case *c, const char *name) {
    if (!strcmp(name, "dt")) return (double)c->dt;
    if (!strcmp(name, "dx")) return (double)c->dx;
    if (!strcmp(name, "dy")) return (double)c->dy;
    if (!strcmp(name, "t")) return c->t;
    if (!strcmp(name, "tstep")) return (double)c->tstep;
    if (!strcmp(name, "ngx")) return c->ngx;
    if (!strcmp(name, "ngy")) return c->ngy;
    if (!strcmp(name, "nsp")) return c->nsp;
    if (!strcmp(name, "A_size")) return c->A_size;
    if (!strcmp(name, "CondNUMR")) return c->CondNUMR;
    if (!strcmp(name, "MCCn")) return c->MCCn;
    if (!strcmp(name, "dt_mcc")) return (double)c->dt_mcc;
    if (!strcmp(name, "nvel")) return c->nvel;
    if (!strcmp(name, "TnRct")) return c->TnRct;
    if (!strcmp(name, "n_accum")) return c->n_accum;
    if (!strcmp(name, "dtc")) return (double)c->dtc;
    if (!strcmp(name, "fl_nx")) return c->nlx;
    if (!strcmp(name, "fl_ny")) return c->nly;
    if (!strcmp(name, "nfsp")) return c->nfsp;
    if (!strcmp(name, "conti_flag")) return c->conti_flag;
    if (!strcmp(name, "nave_count")) return c->nave_count;
    if (!strncmp(name, "np", 2)) return c->sp[name[2] - '0'].n;
    if (!strncmp(name, "cap", 3)) return c->sp[name[3] - '0'].cap;
    if (!strncmp(name, "next_id", 7)) return c->sp[name[7] - '0'].next_id;
    if (!strncmp(name, "Ascale", 6)) return (double)c->Ascale[name[6] - '0'];
    if (!strncmp(name, "Escale", 6)) return (double)c->Escale[name[6] - '0'];
    if (!strncmp(name, "vti", 3)) return (double)c->vti[name[3] - '0'];
    if (!strncmp(name, "q_density", 9)) return (double)c->q_density[name[9] - '0'];
    if (!strncmp(name, "Denscale", 8)) return (double)c->Denscale[name[8] - '0'];
    return 0.0 / 0.0;
}

void orc_set_np(orc_case *c, int isp, int n) { c->sp[isp].n = n; }
void orc_set_fluid(orc_case *c, int on) { c->fluid_on = on; }
void orc_set_tstep(orc_case *c, long tstep) { c->tstep = tstep; c->t = (double)tstep * (double)c->dt; }
int orc_sizeof_deck(void) { return (int)sizeof(orc_deck); }
int orc_sizeof_real(void) { return (int)sizeof(real); }

/* ================= helpers exported for the pinning tests ================= */
void orc_philox(const uint32_t *ctr, const uint32_t *key, uint32_t *out) { orc_philox4x32_10(ctr, key, out); }

double orc_detmath(int fn, double x, double y) {
    switch (fn) {
    case 0: return dm_exp(x);
    case 1: return dm_log(x);
    case 2: return dm_log10(x);
    case 3: return dm_sin(x);
    case 4: return dm_cos(x);
    case 5: return dm_tan(x);
    case 6: return dm_atan(x);
    case 7: return dm_pow(x, y);
    default: return 0.0;
    }
}

/* the single-precision routines of detmath.h, element by element:
 * fn 0 exp, 1 log, 2 log10, 3 sin, 4 cos, 5 tan, 6 atan, 7 pow(x, y) */
void orc_detmathf(int fn, int n, const float *x, const float *y, float *out) {
    int i;
    for (i = 0; i < n; ++i) {
        float v = 0.0f, s, c;
        switch (fn) {
        case 0: v = dmf_exp(x[i]); break;
        case 1: v = dmf_log(x[i]); break;
        case 2: v = dmf_log10(x[i]); break;
        case 3: dmf_sincos(x[i], &s, &c); v = s; break;
        case 4: dmf_sincos(x[i], &s, &c); v = c; break;
        case 5: v = dmf_tan(x[i]); break;
        case 6: v = dmf_atan(x[i]); break;
        case 7: v = dmf_pow(x[i], y[i]); break;
        default: break;
        }
        out[i] = v;
    }
}

/* Test hook: one call of a collision kinematics routine of the MCC above,
 * with its random numbers given.  which = 0 anewvel {E, vel, ux, uy, uz, r1,
 * r2} -> {vx, vy, vz}; 1 newvel_ionsc {vx, vy, vz, vel, r1, r2} -> {vx, vy,
 * vz}; 2 dmaxwellv {vsave_n, vti, Rphi, Rthe} -> {vx, vy, vz}; 3 the
 * ionisation energy sharing {E, u} -> {secondary energy}. */
int orc_kinematics(int which, const double *in, double *out) {
    real a = 0, b = 0, cc = 0;
    switch (which) {
    case 0:
        a = (real)in[2]; b = (real)in[3]; cc = (real)in[4];
        anewvel((real)in[0], (real)in[1], &a, &b, &cc, (real)in[5], (real)in[6]);
        break;
    case 1:
        a = (real)in[0]; b = (real)in[1]; cc = (real)in[2];
        newvel_ionsc(&a, &b, &cc, (real)in[3], (real)in[4], (real)in[5]);
        break;
    case 2:
        dmaxwellv(&a, &b, &cc, (real)in[0], (real)in[1], (real)in[2], (real)in[3]);
        break;
    case 3: {
        real engy = (real)in[0], u = (real)in[1];
        a = (real)10 * R_TAN(u * R_ATAN(engy / (real)20));
        break;
    }
    default:
        return -1;
    }
    out[0] = (double)a; out[1] = (double)b; out[2] = (double)cc;
    return 0;
}

/* n draws of the Park-Miller generator from seed (src/Particle.c:211-226) */
void orc_lcg_seq(long seed, int n, long *out) {
    long a = 16807, m = 2147483647, q = 127773, r = 2836;
    int i;
    for (i = 0; i < n; ++i) {
        long hi = seed / q, lo = seed - q * hi;
        seed = a * lo - r * hi;
        if (seed <= 0) seed += m;
        out[i] = seed;
    }
}

/* argon cross section lookup, for the collision-statistics tests */
double orc_ar_sigma(orc_case *c, int r, double engy) { return (double)ar_sigma(c, r, (real)engy); }

/* ================= MCC (O2, gas type 1; Ar/O2, gas type 2):
 *   src/cuda_mcc.cu:21-96 (MCC_ArO2_Basic, MCC_O2_Basic);
 *   src/cuda_mccO2.cu:867-977 and src/cuda_mccArO2.cu:1130-1265 (target tests);
 *   src/cuda_mccO2.cu:2-866 and src/cuda_mccArO2.cu:51-1129 (handlers) ============
 * Each particle draws from its own stream (seed, id, step, (species << 8) | 1)
 * in the reference's order: the target draws (up to MCCn for the sub-cycled
 * species, one for the others), the neutral's velocity (ions that collide in
 * the neutral's frame: index, phi, theta), the reaction draw, then for an
 * ionisation with energy sharing the sharing draw, each product's draws (an
 * electron's two scattering angles, a heavy product's Maxwellian), and the two
 * angles of the scattered projectile.  Reactions are the rows of o2_chem.h and
 * aro2_chem.h. */
#include "aro2_chem.h"
#include "o2_chem.h"

typedef struct { int r, sum, type, emode; double gain; int p0, p1, th; } o2_row;
typedef struct { int sp, flag, sig, first, n, vsrc; } o2_grp;
typedef struct { int sp, flag, sig, kind, idx; } o2_tgt;
#define ORC_O2_ROW(r, sm, ty, em, g, a, b) {r, sm, ty, em, g, a, b, r},
#define ORC_ARO2_ROW(r, sm, ty, em, g, a, b, th) {r, sm, ty, em, g, a, b, th},
#define ORC_O2_GRP(s, f, sg, fi, n) {s, f, sg, fi, n, (s) == 1 || (s) == 2 ? 1 : 0},
#define ORC_ARO2_GRP(s, f, sg, fi, n, v) {s, f, sg, fi, n, v},
#define ORC_TGT(s, f, sg, kd, ix) {s, f, sg, kd, ix},
static const o2_row o2_rows[2][CHEM_MAXROWS] = {{O2_REACTIONS(ORC_O2_ROW)}, {ARO2_REACTIONS(ORC_ARO2_ROW)}};
static const o2_grp o2_grps[2][CHEM_MAXGROUPS] = {{O2_GROUPS(ORC_O2_GRP)}, {ARO2_GROUPS(ORC_ARO2_GRP)}};
static const o2_tgt o2_tgts[2][CHEM_MAXTARGETS] = {{O2_TARGETS(ORC_TGT)}, {ARO2_TARGETS(ORC_TGT)}};
static const int o2_group0[2][CHEM_MAXSP] = {O2_GROUP0, ARO2_GROUP0};
static const int o2_ngroup[2][CHEM_MAXSP] = {O2_NGROUP, ARO2_NGROUP};
static const int o2_nsp[2] = {O2_NSP, ARO2_NSP}, o2_ntgt[2] = {O2_NTARGETS, ARO2_NTARGETS};
static const int o2_neg[2] = {O2_NEG, ARO2_NEG}, o2_sub[2] = {O2_SUBCYCLED, ARO2_SUBCYCLED};
static const uint32_t o2_vtag[2] = {O2_VICTIM_TAG, ARO2_VICTIM_TAG};

/* O2_CrossSection / ArO2_CrossSection (src/cuda_mccO2.cu:978-...,
 * src/cuda_mccArO2.cu:2461-...): the argon lookup's form on the table */
static real o2_sigma(const orc_o2_in *k, int r, real engy) {
    if (engy == 0) return 0;
    int N = k->nlogx, T = k->TnRct;
    real lengy = R_LOG10(engy);
    real ee1 = k->idLOGX * (lengy - k->xe[0]);
    int ee2 = (int)ee1;
    real a1 = ee1 - (real)ee2;
    real a2 = 1 - a1;
    if (lengy < k->xe[0]) return k->cx[r];
    if (lengy > k->xe[N - 1])
        return (real)((double)k->cx[(N - 1) * T + r] * 0.1 * dm_exp(-(double)(lengy - k->xe[N - 1])));
    if (ee2 >= N - 1) return k->cx[(N - 1) * T + r];
    return a2 * k->cx[ee2 * T + r] + a1 * k->cx[(ee2 + 1) * T + r];
}

static void o2_maxw(const orc_o2_in *k, real vti, orc_stream *st, real *v) {
    int n = (int)((real)(k->nvel - 1) * U(st));
    real rphi = U(st), rthe = U(st);
    dmaxwellv(&v[0], &v[1], &v[2], k->vsave[n], vti, rphi, rthe);
}

typedef struct {
    int r, dead, np, detach;
    int ps[2];
    real pv[2][3];
} o2_out;

/* one collision of species s with the target of group grp at node g */
static void o2_collide(const orc_o2_in *k, int s, int grp, int g, orc_stream *st, real *v, o2_out *o) {
    const o2_grp *G = &o2_grps[k->chem][grp];
    const o2_row *rows = o2_rows[k->chem];
    real vn[3] = {0, 0, 0};
    int heavy = G->vsrc != 0, q;
    o->r = -1; o->dead = 0; o->np = 0; o->detach = 0;
    if (heavy) o2_maxw(k, G->vsrc == 2 ? k->BackVel2[g] : k->BackVel1[g], st, vn);
    real VX = v[0] - vn[0], VY = v[1] - vn[1], VZ = v[2] - vn[2];
    real dum = VX * VX + VY * VY + VZ * VZ;
    real vel = RSQRT(dum);
    if (!heavy) { VX /= vel; VY /= vel; VZ /= vel; }
    real engy = k->Escale[s] * dum;
    real R1 = U(st) * k->sigv[G->sig] / vel;
    real sum = 0;
    const o2_row *row = 0;
    for (q = G->first; q < G->first + G->n; ++q) {
        if (!(engy > k->Th_e[rows[q].r])) continue;
        real sg = o2_sigma(k, rows[q].r, engy);
        sum = rows[q].sum ? sum + sg : sg;
        if (R1 <= sum) { row = &rows[q]; break; }
    }
    if (!row) return;
    o->r = row->r;
    if (row->type == O2T_LOSS || row->type == O2T_ION || row->type == O2T_DETACH) {
        if (row->emode == O2E_LOSS) { engy -= k->Th_e[row->th]; vel = RSQRT(RABS(engy) / k->Escale[s]); }
        if (row->emode == O2E_GAIN) { engy += (real)row->gain; vel = RSQRT(RABS(engy) / k->Escale[s]); }
        int prods[2] = {row->p0, row->type == O2T_ION ? row->p1 : -1}, j;
        for (j = 0; j < 2; ++j) {
            if (prods[j] < 0) continue;
            o->ps[o->np] = prods[j];
            if (prods[j] == 0) { o->pv[o->np][0] = VX; o->pv[o->np][1] = VY; o->pv[o->np][2] = VZ; }
            else o2_maxw(k, k->BackVel1[g], st, o->pv[o->np]);
            o->np++;
        }
        if (row->type == O2T_DETACH) o->detach = 1;
        real a = U(st), b = U(st);
        anewvel(engy, vel, &VX, &VY, &VZ, a, b);
        v[0] = VX; v[1] = VY; v[2] = VZ;
    } else if (row->type == O2T_IONSHARE) {
        /* src/cuda_mccArO2.cu:108-114 (energy sharing), 470-510 (products) */
        int j;
        engy -= k->Th_e[row->th];
        real u = U(st);
        real rengy = (real)10 * R_TAN(u * R_ATAN(engy / (real)20));
        engy -= rengy;
        real velr = RSQRT(RABS(rengy) / k->Escale[s]);
        real vel2 = RSQRT(RABS(engy) / k->Escale[s]);
        int prods[2] = {row->p0, row->p1};
        for (j = 0; j < 2; ++j) {
            o->ps[o->np] = prods[j];
            if (prods[j] == 0) {
                real ex = VX, ey = VY, ez = VZ;
                real a = U(st), b = U(st);
                anewvel(rengy, velr, &ex, &ey, &ez, a, b);
                o->pv[o->np][0] = ex; o->pv[o->np][1] = ey; o->pv[o->np][2] = ez;
            } else {
                o2_maxw(k, k->BackVel1[g], st, o->pv[o->np]);
            }
            o->np++;
        }
        real a = U(st), b = U(st);
        anewvel(engy, vel2, &VX, &VY, &VZ, a, b);
        v[0] = VX; v[1] = VY; v[2] = VZ;
    } else if (row->type == O2T_ATTACH) {
        o->dead = 1;
        o->ps[0] = row->p0; o2_maxw(k, k->BackVel1[g], st, o->pv[0]); o->np = 1;
    } else if (row->type == O2T_DEL) {
        o->dead = 1;
    } else if (row->type == O2T_SCAT) {
        real a = U(st), b = U(st);
        newvel_ionsc(&VX, &VY, &VZ, vel, a, b);
        if (heavy) { v[0] = VX + vn[0]; v[1] = VY + vn[1]; v[2] = VZ + vn[2]; }
        else { v[0] = VX; v[1] = VY; v[2] = VZ; }
    } else if (row->type == O2T_CXKEEP) {
        v[0] = vn[0]; v[1] = vn[1]; v[2] = vn[2];
    } else if (row->type == O2T_CXTO) {
        o->dead = 1;
        o->ps[0] = row->p0; o->pv[0][0] = vn[0]; o->pv[0][1] = vn[1]; o->pv[0][2] = vn[2]; o->np = 1;
    } else { /* O2T_NEGDET */
        o->dead = 1;
        o->ps[0] = 0; o->pv[0][0] = VX; o->pv[0][1] = VY; o->pv[0][2] = VZ; o->np = 1;
    }
}

typedef struct { uint32_t key; int pos; } o2_key;
static int o2_key_cmp(const void *a, const void *b) {
    const o2_key *x = a, *y = b;
    if (x->key != y->key) return x->key < y->key ? -1 : 1;
    return x->pos - y->pos;
}

int orc_mcc_o2(const orc_o2_in *k, orc_bucket *B, uint32_t *next_id, int64_t *mccr) {
    const int ch = k->chem, nsp = o2_nsp[ch], ntgt = o2_ntgt[ch], neg = o2_neg[ch];
    int G = k->G, g, s, t, kk, j, rc = 0;
    real *prob = zalloc((size_t)G * ntgt, sizeof(real));
    for (g = 0; g < G; ++g) {
        real acc = 0;
        int sp = -1;
        for (t = 0; t < ntgt; ++t) {
            const o2_tgt *T = &o2_tgts[ch][t];
            if (T->sp != sp) { sp = T->sp; acc = 0; }
            real x = (((o2_sub[ch] >> T->sp) & 1) ? k->dtm : k->dt) * k->sigv[T->sig];
            if (T->kind == 0) x = x * (T->idx ? k->BackDen2[g] : k->BackDen1[g]);
            else if (T->kind == 1) x = x * k->FluidN[(size_t)T->idx * G + g];
            else x = x * k->den[(size_t)T->idx * G + g] * k->np2c[T->idx] * k->idx * k->idy;
            acc = (acc + (real)1) - (real)dm_exp(-(double)x);
            prob[(size_t)g * ntgt + t] = acc;
        }
    }
    /* collisions: outcome per particle */
    int *dead[CHEM_MAXSP];
    o2_out *out[CHEM_MAXSP];
    int *delreq = zalloc(G, sizeof(int));
    for (s = 0; s < nsp; ++s) {
        dead[s] = zalloc(B[s].n + 1, sizeof(int));
        out[s] = zalloc(B[s].n + 1, sizeof(o2_out));
    }
    for (s = 0; s < nsp; ++s) {
        if (s != 0 && !k->ions && ch == 0) continue;   /* O2: the reference's thread bound TID < Gsize */
        int ng = o2_ngroup[ch][s], g0 = o2_group0[ch][s];
        int draws = ((o2_sub[ch] >> s) & 1) ? k->MCCn : 1;
        for (g = 0; g < G; ++g) {
            const real *P = prob + (size_t)g * ntgt + g0;
            real Tprob = P[ng - 1];
            for (kk = B[s].cell_start[g]; kk < B[s].cell_start[g + 1]; ++kk) {
                orc_stream st;
                orc_stream_init(&st, k->seed, B[s].id[kk], k->step, (uint32_t)((s << 8) | 1));
                real R1 = 0;
                out[s][kk].r = -1;
                out[s][kk].np = 0;
                for (j = 0; j < draws; ++j) { R1 = U(&st); if (R1 < Tprob) break; }
                if (!(R1 < Tprob)) continue;
                int flag = 0;
                while (flag < ng - 1 && R1 > P[flag]) ++flag;
                real v[3] = {B[s].vx[kk], B[s].vy[kk], B[s].vz[kk]};
                o2_collide(k, s, g0 + flag, g, &st, v, &out[s][kk]);
                if (out[s][kk].r < 0) continue;
                mccr[(size_t)g * k->TnRct + out[s][kk].r]++;
                if (out[s][kk].dead) dead[s][kk] = 1;
                else { B[s].vx[kk] = v[0]; B[s].vy[kk] = v[1]; B[s].vz[kk] = v[2]; }
                if (out[s][kk].detach) delreq[g]++;
            }
        }
    }
    /* O- removed by electron detachment: per node, min(requests, O- that
     * survived their own collision), smallest Philox keys first */
    for (g = 0; g < G; ++g) {
        if (!delreq[g]) continue;
        int a = B[neg].cell_start[g], b = B[neg].cell_start[g + 1], m = 0;
        o2_key *keys = zalloc(b - a + 1, sizeof(o2_key));
        for (kk = a; kk < b; ++kk) {
            if (dead[neg][kk]) continue;
            uint32_t ctr[4] = {B[neg].id[kk], k->step, o2_vtag[ch], 0};
            uint32_t key[2] = {(uint32_t)k->seed, (uint32_t)(k->seed >> 32)}, o[4];
            orc_philox4x32_10(ctr, key, o);
            keys[m].key = o[0]; keys[m].pos = kk; ++m;
        }
        qsort(keys, m, sizeof(o2_key), o2_key_cmp);
        int kdel = delreq[g] < m ? delreq[g] : m;
        for (j = 0; j < kdel; ++j) dead[neg][keys[j].pos] = 2;
        if (k->delcnt) k->delcnt[g] += kdel;
        free(keys);
    }
    /* rebuild: survivors in order, then products by (source species, parent,
     * slot).  Every species' rebuild reads the old buckets of all species, so
     * the new ones are built aside and copied back at the end. */
    real *nx[CHEM_MAXSP], *ny[CHEM_MAXSP], *nvx[CHEM_MAXSP], *nvy[CHEM_MAXSP], *nvz[CHEM_MAXSP];
    uint32_t *nidv[CHEM_MAXSP];
    int *start[CHEM_MAXSP], nnew[CHEM_MAXSP];
    for (t = 0; t < nsp; ++t) {
        orc_bucket *D = &B[t];
        int cap = D->cap, pos = 0;
        uint32_t nid = next_id[t];
        nx[t] = zalloc(cap + 1, sizeof(real)); ny[t] = zalloc(cap + 1, sizeof(real));
        nvx[t] = zalloc(cap + 1, sizeof(real)); nvy[t] = zalloc(cap + 1, sizeof(real));
        nvz[t] = zalloc(cap + 1, sizeof(real));
        nidv[t] = zalloc(cap + 1, sizeof(uint32_t));
        start[t] = zalloc(G + 1, sizeof(int));
        for (g = 0; g < G; ++g) {
            start[t][g] = pos;
            for (kk = D->cell_start[g]; kk < D->cell_start[g + 1]; ++kk) {
                if (dead[t][kk]) continue;
                if (pos < cap) {
                    nx[t][pos] = D->x[kk]; ny[t][pos] = D->y[kk];
                    nvx[t][pos] = D->vx[kk]; nvy[t][pos] = D->vy[kk]; nvz[t][pos] = D->vz[kk];
                    nidv[t][pos] = D->id[kk];
                }
                ++pos;
            }
            int s2;
            for (s2 = 0; s2 < nsp; ++s2)
                for (kk = B[s2].cell_start[g]; kk < B[s2].cell_start[g + 1]; ++kk) {
                    const o2_out *o = &out[s2][kk];
                    for (j = 0; j < o->np; ++j) {
                        if (o->ps[j] != t) continue;
                        if (pos < cap) {
                            nx[t][pos] = B[s2].x[kk]; ny[t][pos] = B[s2].y[kk];
                            nvx[t][pos] = o->pv[j][0]; nvy[t][pos] = o->pv[j][1]; nvz[t][pos] = o->pv[j][2];
                            nidv[t][pos] = nid;
                        }
                        ++nid;
                        ++pos;
                    }
                }
        }
        start[t][G] = pos;
        if (pos > cap) rc = -30;
        nnew[t] = pos < cap ? pos : cap;
        next_id[t] = nid;
    }
    for (t = 0; t < nsp; ++t) {
        orc_bucket *D = &B[t];
        int n = nnew[t];
        memcpy(D->x, nx[t], n * sizeof(real)); memcpy(D->y, ny[t], n * sizeof(real));
        memcpy(D->vx, nvx[t], n * sizeof(real)); memcpy(D->vy, nvy[t], n * sizeof(real));
        memcpy(D->vz, nvz[t], n * sizeof(real));
        memcpy(D->id, nidv[t], n * sizeof(uint32_t));
        memcpy(D->cell_start, start[t], (G + 1) * sizeof(int));
        D->n = n;
        free(nx[t]); free(ny[t]); free(nvx[t]); free(nvy[t]); free(nvz[t]); free(nidv[t]); free(start[t]);
    }
    for (s = 0; s < nsp; ++s) { free(dead[s]); free(out[s]); }
    free(delreq); free(prob);
    return rc;
}
