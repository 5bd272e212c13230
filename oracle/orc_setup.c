/*
 * ORACLE - test infrastructure only (see pic_oracle.h).
 * Setup restatement: deck-derived constants, geometry maps, field operator,
 * direct FP64 solves, cross-section tables, null-collision constants,
 * Maxwellian table and the Park-Miller loader.
 */
#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "pic_oracle.h"
#include "detmath.h"
#include "xsec_data.inc"

#define EPS0 8.8542e-12
#define CQ 1.602e-19
#define AMU 1.66053e-27
#define NPERTORR 8.3221e20
#define DBLMIN 1e-200
#define NVTS 3.3
#define ORC_PI 3.14159265358979323846

static void *xcalloc(size_t n, size_t s) {
    void *p = calloc(n ? n : 1, s);
    if (!p) { fprintf(stderr, "oracle: out of memory\n"); exit(2); }
    return p;
}

/* ---------- deck-derived constants: src/start.c:19-760, 1432-1460, 1999-2023;
 *            src/Efield.c:3-14 ---------- */
static int derive_constants(orc_case *c) {
    const orc_deck *d = &c->deck;
    int i;
    c->ngx = d->ngx; c->ngy = d->ngy;
    if (d->nbnd < 4) return -10;                       /* start.c:56-59 */
    if (d->ncond < 2) return -11;                      /* start.c:95-98 */
    if (d->ncond > 100) return -11;
    for (i = 0; i < d->nbnd; ++i) {
        if (d->bx0[i] > d->bx1[i] || d->bx1[i] >= d->ngx) return -12;
        if (d->by0[i] > d->by1[i] || d->by1[i] >= d->ngy) return -12;
        if (d->bbc[i] > 4) return -12;
    }
    /* conductor M_ID must start at 1 and be non-decreasing (start.c:103-133) */
    {
        int ids[100], nid = 0;
        for (i = 0; i < d->ncond; ++i) {
            int m = d->cmid[i], k, dup = 0;
            if (m == 0) return -13;
            if (nid == 0) { if (m != 1) return -13; ids[nid++] = m; continue; }
            if (ids[nid - 1] > m) return -13;
            for (k = 0; k < nid; ++k) if (ids[k] == m) dup = 1;
            if (!dup) ids[nid++] = m;
        }
        c->CondNUMR = nid;
    }
    for (i = 0; i < d->ncond; ++i) {
        if (d->cx0[i] > d->cx1[i] || d->cx1[i] >= d->ngx) return -14;
        if (d->cy0[i] > d->cy1[i] || d->cy1[i] >= d->ngy) return -14;
    }
    c->ncond = d->ncond;
    /* sources: must match a conductor; max frequency sets dt (start.c:176-220) */
    float max_freq = 0.0f;
    for (i = 0; i < d->nsrc; ++i) {
        int k, ok = 0;
        for (k = 0; k < d->ncond; ++k) if (d->cmid[k] == d->smid[i]) ok = 1;
        if (!ok) return -15;
        if ((float)d->sfreq[i] != 0.0f && (float)d->sfreq[i] > max_freq) max_freq = (float)d->sfreq[i];
        if ((float)d->sdc[i] == 0.0f && (float)d->spow[i] == 0.0f && (float)d->sac[i] == 0.0f) return -15;
    }
    if (max_freq <= 0.0f) return -16;
    if (d->dt_pic == 0) return -17;
    if (d->gas < 0 || d->gas > 2) return -18;

    c->Gsize = c->ngx * c->ngy;
    c->ncx = c->ngx - 1; c->ncy = c->ngy - 1;
    c->Csize = c->ncx * c->ncy;
    real xl = (real)d->xlength, yl = (real)d->ylength;
    c->zlength = (real)d->zlength;
    c->dx = xl / c->ncx;                               /* start.c:1442-1443 */
    c->dy = yl / c->ncy;
    c->idx = (real)1 / c->dx;
    c->idy = (real)1 / c->dy;
    c->idxy = (real)(1.5 / (double)(c->dx + c->dy));
    c->hdx = (real)(0.5 / (double)c->dx);
    c->hdy = (real)(0.5 / (double)c->dy);
    /* dt = 1/f_max/TimeStep_PIC (src/Efield.c:7); Max_FREQ is a float */
    if (sizeof(real) == sizeof(float)) {
        c->dt = (real)(1.0 / (double)max_freq / (double)(float)d->dt_pic);
    } else {
        double mf = 0.0;
        for (i = 0; i < d->nsrc; ++i) if (d->sfreq[i] > mf) mf = d->sfreq[i];
        c->dt = (real)(1.0 / mf / (double)d->dt_pic);
    }
    c->dt_d = (double)c->dt;
    c->dt_dx = c->dt / c->dx;
    c->dt_dy = c->dt / c->dy;

    /* gas and species (start.c:304-725) */
    if (d->gas == 0) {
        c->nsp = 2; c->nfsp = 1;
        c->bg_mass[0] = (real)(39.950 * AMU);
        c->bg_pres[0] = (real)d->total_pres;
        c->bg_dens[0] = (real)(NPERTORR * (double)c->bg_pres[0] / ((double)(real)d->bg_temp[0] + DBLMIN));
        c->fl_mass[0] = (real)(39.950 * AMU);
        c->mass[0] = (real)(5.485799486e-4 * AMU); c->q[0] = (real)(-1.0 * CQ);
        c->mass[1] = (real)(39.9500 * AMU);        c->q[1] = (real)(1.0 * CQ);
    } else if (d->gas == 1) {
        /* O2 (start.c:420-550): O2 background; fluid O2A, O2B, OP, OD;
         * particles e, O2+, O+, O- */
        c->nsp = 4; c->nfsp = 4;
        c->bg_mass[0] = (real)(32.0 * AMU);
        c->bg_pres[0] = (real)d->total_pres;
        c->bg_dens[0] = (real)(NPERTORR * (double)c->bg_pres[0] / ((double)(real)d->bg_temp[1] + DBLMIN));
        c->fl_mass[0] = c->fl_mass[1] = (real)(32.0 * AMU);
        c->fl_mass[2] = c->fl_mass[3] = (real)(16.0 * AMU);
        c->mass[0] = (real)(5.485799486e-4 * AMU); c->q[0] = (real)(-1.0 * CQ);
        c->mass[1] = (real)(32.0 * AMU);           c->q[1] = (real)(1.0 * CQ);
        c->mass[2] = (real)(16.0 * AMU);           c->q[2] = (real)(1.0 * CQ);
        c->mass[3] = (real)(16.0 * AMU);           c->q[3] = (real)(-1.0 * CQ);
    } else {
        /* Ar/O2 (start.c:551-725): the O2 share of the pressure is
         * r_O2 / (r_Ar + r_O2) (float), Ar takes 1 - that; fluid Ar*, O2A,
         * O2B, OP, OD; particles e, Ar+, O2+, O+, O- */
        real r1 = (real)d->bg_ratio[0], r2 = (real)d->bg_ratio[1];
        real f2 = r2 / (r1 + r2);
        real f1 = (real)(1.0 - (double)f2);
        c->nsp = 5; c->nfsp = 5;
        c->bg_mass[0] = (real)(39.950 * AMU);
        c->bg_pres[0] = (real)d->total_pres * f1;
        c->bg_dens[0] = (real)(NPERTORR * (double)c->bg_pres[0] / ((double)(real)d->bg_temp[0] + DBLMIN));
        c->bg_mass[1] = (real)(32.0 * AMU);
        c->bg_pres[1] = (real)d->total_pres * f2;
        c->bg_dens[1] = (real)(NPERTORR * (double)c->bg_pres[1] / ((double)(real)d->bg_temp[1] + DBLMIN));
        c->fl_mass[0] = (real)(39.950 * AMU);
        c->fl_mass[1] = c->fl_mass[2] = (real)(32.0 * AMU);
        c->fl_mass[3] = c->fl_mass[4] = (real)(16.0 * AMU);
        c->mass[0] = (real)(5.485799486e-4 * AMU); c->q[0] = (real)(-1.0 * CQ);
        c->mass[1] = (real)(39.950 * AMU);         c->q[1] = (real)(1.0 * CQ);
        c->mass[2] = (real)(32.0 * AMU);           c->q[2] = (real)(1.0 * CQ);
        c->mass[3] = (real)(16.0 * AMU);           c->q[3] = (real)(1.0 * CQ);
        c->mass[4] = (real)(16.0 * AMU);           c->q[4] = (real)(-1.0 * CQ);
    }
    for (i = 0; i < c->nfsp; ++i) c->fl_dens[i] = (real)d->fl_dens[i];
    for (i = 0; i < c->nsp; ++i) {
        c->np2c[i] = (real)d->sp_np2c[i];
        real temp = (real)d->sp_temp[i];
        c->q_density[i] = c->q[i] * c->np2c[i];
        c->vti[i] = (real)sqrt(CQ * (double)temp / (double)c->mass[i]);
        c->qm[i] = c->q[i] / c->mass[i];
        c->Escale[i] = (real)(0.5 * (double)c->mass[i] / CQ);
        c->Ascale[i] = c->qm[i] * c->dt;
        c->Denscale[i] = c->np2c[i] / c->dx / c->dy;
    }
    return 0;
}

/* ---------- geometry: src/start.c:1464-1775 ---------- */
static void geometry(orc_case *c) {
    const orc_deck *d = &c->deck;
    int ngx = c->ngx, ngy = c->ngy, ncx = c->ncx, ncy = c->ncy;
    int i, j, k, ID, CID;
    real dx = c->dx, dy = c->dy, zl = c->zlength;
    c->Boundary = xcalloc(c->Gsize, sizeof(int));
    c->CondID = xcalloc(c->Gsize, sizeof(int));
    c->Face = xcalloc(c->Gsize, sizeof(int));
    c->DensRegion = xcalloc(c->Gsize, sizeof(int));
    c->Area = xcalloc(c->Gsize, sizeof(real));
    c->Temp = xcalloc(c->Gsize, sizeof(real));
    c->PlasmaRegion = xcalloc(c->Csize, sizeof(int));
    c->eps_r = xcalloc(c->Csize, sizeof(real));
    c->eps = xcalloc(c->Csize, sizeof(real));
    for (i = 0; i < c->Gsize; ++i) { c->Face[i] = ORC_NOFACE; c->Area[i] = dy * zl; }
    for (i = 0; i < c->Csize; ++i) { c->PlasmaRegion[i] = 1; c->eps_r[i] = 1; }
    for (j = 0; j < ngy; ++j)
        for (i = 0; i < ngx; ++i) {
            ID = i * ngy + j;
            for (k = 0; k < d->nbnd; ++k) {
                if (i >= d->bx0[k] && i <= d->bx1[k] && j >= d->by0[k] && j <= d->by1[k]) {
                    if (d->bbc[k] == ORC_DIRICHLET) {
                        c->Boundary[ID] = ORC_DIRICHLET; c->Temp[ID] = (real)d->btemp[k];
                    } else if (c->Boundary[ID] != ORC_DIRICHLET && d->bbc[k] == ORC_NEUMANN) {
                        c->Boundary[ID] = ORC_NEUMANN; c->Temp[ID] = (real)d->btemp[k];
                    }
                }
            }
        }
    for (k = 0; k < d->ndiel; ++k)
        for (i = d->dx0[k]; i <= d->dx1[k]; ++i)
            for (j = d->dy0[k]; j <= d->dy1[k]; ++j) {
                ID = i * ngy + j;
                if (c->Boundary[ID] == 0) c->Boundary[ID] = ORC_DIELECTRIC;
                if (i < d->dx1[k] && j < d->dy1[k]) {
                    CID = i * ncy + j;
                    c->PlasmaRegion[CID] = 0; c->eps_r[CID] = (real)d->deps[k];
                }
            }
    for (k = 0; k < d->ncond; ++k)
        for (i = d->cx0[k]; i <= d->cx1[k]; ++i)
            for (j = d->cy0[k]; j <= d->cy1[k]; ++j) {
                ID = i * ngy + j;
                c->CondID[ID] = d->cmid[k];
                if (c->Boundary[ID] == 0 || c->Boundary[ID] == ORC_DIELECTRIC) {
                    c->Boundary[ID] = ORC_CONDUCTOR; c->Temp[ID] = (real)d->ctemp[k];
                }
                if (i < d->cx1[k] && j < d->cy1[k]) {
                    CID = i * ncy + j;
                    c->PlasmaRegion[CID] = 0; c->eps_r[CID] = 0;
                }
            }
    for (i = 0; i < ncx; ++i)
        for (j = 0; j < ncy; ++j) {
            if (c->PlasmaRegion[i * ncy + j] == 1) {
                ID = i * ngy + j;
                c->DensRegion[ID] = 1; c->DensRegion[ID + 1] = 1;
                c->DensRegion[ID + ngy] = 1; c->DensRegion[ID + ngy + 1] = 1;
            }
        }
    /* interior dielectric / conductor faces (start.c:1550-1625).  The
     * reference compares eps_r through int variables P1..P4. */
    for (i = 1; i < ngx - 1; ++i)
        for (j = 1; j < ngy - 1; ++j) {
            ID = i * ngy + j; CID = i * ncy + j;
            int P1 = (int)c->eps_r[CID], P2 = (int)c->eps_r[CID - ncy];
            int P3 = (int)c->eps_r[CID - 1], P4 = (int)c->eps_r[CID - ncy - 1];
            int b = c->Boundary[ID];
            if (b != ORC_DIELECTRIC && b != ORC_CONDUCTOR) continue;
            int f = -1;
            if ((P2 == 1 && P1 == 1 && P4 == 1) || (P2 != 0 && P1 != 0 && P4 != 0 && P3 == 0)) f = ORC_UL;
            else if ((P2 == 1 && P1 == 1 && P3 == 1) || (P2 != 0 && P1 != 0 && P3 != 0 && P4 == 0)) f = ORC_UR;
            else if ((P2 == 1 && P1 == 1 && P4 != 1 && P3 != 1) || (P2 != 0 && P1 != 0 && P4 == 0 && P3 == 0)) f = ORC_UP;
            else if ((P4 == 1 && P3 == 1 && P2 == 1) || (P2 != 0 && P4 != 0 && P3 != 0 && P1 == 0)) f = ORC_LL;
            else if ((P4 == 1 && P3 == 1 && P1 == 1) || (P1 != 0 && P4 != 0 && P3 != 0 && P2 == 0)) f = ORC_LR;
            else if ((P4 == 1 && P3 == 1 && P1 != 1 && P2 != 1) || (P4 != 0 && P3 != 0 && P1 == 0 && P2 == 0)) f = ORC_DOWN;
            else if ((P4 == 1 && P2 == 1 && P1 != 1 && P3 != 1) || (P4 != 0 && P2 != 0 && P1 == 0 && P3 == 0)) f = ORC_LEFT;
            else if ((P4 != 1 && P2 != 1 && P1 == 1 && P3 == 1) || (P4 == 0 && P2 == 0 && P1 != 0 && P3 != 0)) f = ORC_RIGHT;
            if (f < 0) continue;
            c->Face[ID] = f;
            if (f == ORC_LEFT || f == ORC_RIGHT) c->Area[ID] = dy * zl;
            else if (f == ORC_UP || f == ORC_DOWN) c->Area[ID] = dx * zl;
            else c->Area[ID] = (real)(0.5 * (double)(dx + dy) * (double)zl);
        }
    /* wall nodes (start.c:1627-1757).  CondID lookups outside the node array
     * (the reference reads one past the end at the top-right corner) read 0. */
#define CID_AT(n) (((n) >= 0 && (n) < c->Gsize) ? c->CondID[(n)] : 0)
    for (j = 1; j < ncy; ++j) {
        int side;
        for (side = 0; side < 2; ++side) {
            ID = side ? ncx * ngy + j : j;
            int in = side ? ID - ngy : ngy + ID;     /* inward neighbour */
            int up = CID_AT(ID + 1), dn = CID_AT(ID - 1), inw = CID_AT(in);
            if (c->Boundary[ID] == ORC_DIRICHLET) {
                c->Face[ID] = side ? ORC_LEFT : ORC_RIGHT; c->Area[ID] = dy * zl;
                if (up == 0 && dn != 0 && inw != 0) { c->Face[ID] = ORC_UP; c->Area[ID] = dx * zl; }
                else if (up != 0 && dn == 0 && inw != 0) { c->Face[ID] = ORC_DOWN; c->Area[ID] = dx * zl; }
                else if (up != 0 && dn != 0 && inw != 0) { c->Face[ID] = ORC_NOFACE; c->Area[ID] = dy * zl; }
                else if (up == 0 && dn == 0 && inw != 0) { c->Face[ID] = ORC_NOFACE; c->Area[ID] = dy * zl; }
            } else if (c->Boundary[ID] == ORC_NEUMANN) {
                c->Face[ID] = ORC_NOFACE; c->Area[ID] = dy * zl;
                if (up == 0 && dn != 0 && inw != 0) { c->Face[ID] = ORC_UP; c->Area[ID] = dx * zl; }
                else if (up != 0 && dn == 0 && inw != 0) { c->Face[ID] = ORC_DOWN; c->Area[ID] = dx * zl; }
                else if (up != 0 && dn != 0 && inw == 0) { c->Face[ID] = side ? ORC_LEFT : ORC_RIGHT; c->Area[ID] = dy * zl; }
            }
        }
    }
    for (i = 1; i < ncx; ++i) {
        int side;
        for (side = 0; side < 2; ++side) {
            ID = side ? i * ngy + ncy : i * ngy;
            int in = side ? ID - 1 : ID + 1;
            int rt = CID_AT(ID + ngy), lf = CID_AT(ID - ngy), inw = CID_AT(in);
            if (c->Boundary[ID] == ORC_DIRICHLET) {
                c->Face[ID] = side ? ORC_DOWN : ORC_UP; c->Area[ID] = dx * zl;
                if (rt == 0 && lf != 0 && inw != 0) { c->Face[ID] = ORC_RIGHT; c->Area[ID] = dy * zl; }
                else if (rt != 0 && lf == 0 && inw != 0) { c->Face[ID] = ORC_LEFT; c->Area[ID] = dy * zl; }
                else if (rt != 0 && lf != 0 && inw != 0) { c->Face[ID] = ORC_NOFACE; c->Area[ID] = dy * zl; }
                else if (rt == 0 && lf == 0 && inw != 0) { c->Face[ID] = ORC_NOFACE; c->Area[ID] = dy * zl; }
            } else if (c->Boundary[ID] == ORC_NEUMANN) {
                c->Face[ID] = ORC_NOFACE; c->Area[ID] = dy * zl;
                if (rt == 0 && lf != 0 && inw != 0) { c->Face[ID] = ORC_RIGHT; c->Area[ID] = dy * zl; }
                else if (rt != 0 && lf == 0 && inw != 0) { c->Face[ID] = ORC_LEFT; c->Area[ID] = dy * zl; }
                else if (rt != 0 && lf != 0 && inw == 0) { c->Face[ID] = side ? ORC_DOWN : ORC_UP; c->Area[ID] = dx * zl; }
            }
        }
    }
    {
        int corners[4] = {0, ncx * ngy, ncy, ncx * ngy + ncy};
        int faces[4] = {ORC_UR, ORC_UL, ORC_LR, ORC_LL};
        for (k = 0; k < 4; ++k) {
            ID = corners[k];
            if (c->Boundary[ID] == ORC_DIRICHLET && CID_AT(ID + 1) == 0) {
                c->Face[ID] = faces[k];
                c->Area[ID] = (real)(0.5 * (double)(dx + dy) * (double)zl);
            }
        }
    }
#undef CID_AT
    /* padded structure map (start.c:1776-1851) */
    int W = ncy + 2;
    c->StructIdx = xcalloc((size_t)(ncx + 2) * W, sizeof(int));
#define SI(a, b) c->StructIdx[(a) * W + (b)]
    for (i = 0; i < ncx; ++i)
        for (j = 0; j < ncy; ++j) {
            SI(i + 1, j + 1) = 0;
            for (k = 0; k < d->ndiel; ++k)
                if (i >= d->dx0[k] && i < d->dx1[k] && j >= d->dy0[k] && j < d->dy1[k]) SI(i + 1, j + 1) = d->dmid[k];
            for (k = 0; k < d->ncond; ++k)
                if (i >= d->cx0[k] && i < d->cx1[k] && j >= d->cy0[k] && j < d->cy1[k]) SI(i + 1, j + 1) = 100 + d->cmid[k];
        }
    for (i = 1; i < ngx; ++i) {
        if ((c->Boundary[(i - 1) * ngy] == ORC_NEUMANN || c->Boundary[i * ngy] == ORC_NEUMANN) && SI(i, 1) == 0)
            SI(i, 0) = -c->Boundary[(i - 1) * ngy] - 2;
        else SI(i, 0) = SI(i, 1) ? SI(i, 1) : -1;
        if ((c->Boundary[(i - 1) * ngy + ncy] == ORC_NEUMANN || c->Boundary[i * ngy + ncy] == ORC_NEUMANN) && SI(i, ngy - 1) == 0)
            SI(i, ngy) = -c->Boundary[(i - 1) * ngy + ncy] - 6;
        else SI(i, ngy) = SI(i, ngy - 1) ? SI(i, ngy - 1) : -1;
    }
    for (j = 1; j < ngy; ++j) {
        if ((c->Boundary[j - 1] == ORC_NEUMANN || c->Boundary[j] == ORC_NEUMANN) && SI(1, j) == 0)
            SI(0, j) = -c->Boundary[j - 1];
        else SI(0, j) = SI(1, j) ? SI(1, j) : -1;
        if ((c->Boundary[ncx * ngy + j - 1] == ORC_NEUMANN || c->Boundary[ncx * ngy + j] == ORC_NEUMANN) && SI(ngx - 1, j) == 0)
            SI(ngx, j) = -c->Boundary[ncx * ngy + j] - 4;
        else SI(ngx, j) = SI(ngx - 1, j) ? SI(ngx - 1, j) : -1;
    }
    SI(0, 0) = -c->Boundary[0];
    SI(ncx + 1, 0) = -c->Boundary[(ncx - 1) * ngy];
    SI(0, ncy + 1) = -c->Boundary[ngy + ncy];
    /* src/start.c:1848 reads node (ncy-1, ncy), past the node array when
     * ncy > ncx; there the top-wall node (ncx-1, ncy) is used */
    SI(ncx + 1, ncy + 1) = -c->Boundary[(size_t)(ncy <= ncx ? ncy - 1 : ncx - 1) * ngy + ncy];
#undef SI
    for (i = 0; i < c->Csize; ++i) c->eps[i] = (real)(EPS0 * (double)c->eps_r[i]);
}

/* ---------- field operator: src/start.c:1911-1935, src/load.c:121-245 ---------- */
static void field_matrix(orc_case *c) {
    int ngx = c->ngx, ngy = c->ngy, ncx = c->ncx, ncy = c->ncy;
    int i, j;
    c->A_idx = xcalloc(c->Gsize, sizeof(int));
    int n = 1;
    for (i = 0; i < ngx; ++i)
        for (j = 0; j < ngy; ++j) {
            int g = i * ngy + j;
            if (!c->CondID[g] && c->Boundary[g] != ORC_DIRICHLET) c->A_idx[g] = n++;
        }
    int N = n - 1;
    c->A_size = N;
    c->Ai = xcalloc(N + 1, sizeof(int));
    c->Aj = xcalloc(5 * (size_t)N, sizeof(int));
    c->Aval = xcalloc(5 * (size_t)N, sizeof(real));
    c->Atemp_val = xcalloc(5 * (size_t)N, sizeof(real));
    c->Mdiag = xcalloc(N, sizeof(real));
    c->cond_b = xcalloc((size_t)c->CondNUMR * N, sizeof(real));
    c->temp_b = xcalloc(N, sizeof(real));
    real dydx = (real)(0.5 * (double)c->dy / (double)c->dx);
    real dxdy = (real)(0.5 * (double)c->dx / (double)c->dy);
    const real *er = c->eps_r;
    int m = 0, l = 0;
    c->Ai[0] = 0;
    for (i = 0; i < ngx; ++i)
        for (j = 0; j < ngy; ++j) {
            int g = i * ngy + j;
            if (!c->A_idx[g]) continue;
            int x1, y1, y2, x2;
            x1 = (i == 0) ? i : i - 1; y1 = (j == ncy) ? j - 1 : j; y2 = (j == 0) ? j : j - 1;
            real L = dydx * (er[x1 * ncy + y1] + er[x1 * ncy + y2]);
            x1 = (i == ncx) ? i - 1 : i;
            real R = dydx * (er[x1 * ncy + y1] + er[x1 * ncy + y2]);
            y1 = (j == ncy) ? j - 1 : j; x1 = (i == ncx) ? i - 1 : i; x2 = (i == 0) ? i : i - 1;
            real U = dxdy * (er[x1 * ncy + y1] + er[x2 * ncy + y1]);
            y1 = (j == 0) ? j : j - 1;
            real D = dxdy * (er[x1 * ncy + y1] + er[x2 * ncy + y1]);
            int LN = (i == 0) ? 2 : 1, RN = (i == ncx) ? 2 : 1, UN = (j == ncy) ? 2 : 1, DN = (j == 0) ? 2 : 1;
            int nb;
            if (i) {   /* left */
                nb = (i - 1) * ngy + j;
                if (c->CondID[nb]) { c->cond_b[(size_t)(c->CondID[nb] - 1) * N + l] += L; c->temp_b[l] += L * c->Temp[nb]; }
                else if (c->Boundary[nb] == ORC_DIRICHLET) c->temp_b[l] += L * c->Temp[nb];
                else { c->Aval[m] = (-L) * RN; c->Atemp_val[m] = (-L) * RN; c->Aj[m] = c->A_idx[nb] - 1; m++; }
            }
            if (j) {   /* down */
                nb = i * ngy + j - 1;
                if (c->CondID[nb]) { c->cond_b[(size_t)(c->CondID[nb] - 1) * N + l] += D; c->temp_b[l] += D * c->Temp[nb]; }
                else if (c->Boundary[nb] == ORC_DIRICHLET) c->temp_b[l] += D * c->Temp[nb];
                else { c->Aval[m] = (-D) * UN; c->Atemp_val[m] = (-D); c->Aj[m] = c->A_idx[nb] - 1; m++; }
            }
            c->Aval[m] = L + D + U + R;             /* middle */
            c->Mdiag[l] = (real)1 / c->Aval[m];
            c->Atemp_val[m] = L + D + U + R;
            c->Aj[m] = c->A_idx[g] - 1; m++;
            if (j != ncy) {   /* up */
                nb = i * ngy + j + 1;
                if (c->CondID[nb]) { c->cond_b[(size_t)(c->CondID[nb] - 1) * N + l] += U; c->temp_b[l] += U * c->Temp[nb]; }
                else if (c->Boundary[nb] == ORC_DIRICHLET) c->temp_b[l] += U * c->Temp[nb];
                else { c->Aval[m] = (-U) * DN; c->Atemp_val[m] = (-U); c->Aj[m] = c->A_idx[nb] - 1; m++; }
            }
            if (i != ncx) {   /* right */
                nb = (i + 1) * ngy + j;
                if (c->CondID[nb]) { c->cond_b[(size_t)(c->CondID[nb] - 1) * N + l] += R; c->temp_b[l] += R * c->Temp[nb]; }
                else if (c->Boundary[nb] == ORC_DIRICHLET) c->temp_b[l] += R * c->Temp[nb];
                else { c->Aval[m] = (-R) * LN; c->Atemp_val[m] = (-R); c->Aj[m] = c->A_idx[nb] - 1; m++; }
            }
            l++;
            c->Ai[l] = m;
        }
}

/* ---------- FP64 banded LU (no pivoting; the operator is an M-matrix) ---------- */
int orc_band_factor(int N, const int *Ai, const int *Aj, const real *val, double **lu_out, int *bw_out) {
    int bw = 0, r, k;
    for (r = 0; r < N; ++r)
        for (k = Ai[r]; k < Ai[r + 1]; ++k) {
            int dcol = Aj[k] - r;
            if (dcol < 0) dcol = -dcol;
            if (dcol > bw) bw = dcol;
        }
    int W = 2 * bw + 1;
    double *lu = calloc((size_t)N * W, sizeof(double));
    if (!lu) return -1;
#define LU(rr, cc) lu[(size_t)(rr) * W + ((cc) - (rr) + bw)]
    for (r = 0; r < N; ++r)
        for (k = Ai[r]; k < Ai[r + 1]; ++k) LU(r, Aj[k]) += (double)val[k];
    for (k = 0; k < N; ++k) {
        double piv = LU(k, k);
        int rmax = k + bw < N - 1 ? k + bw : N - 1;
        int i2, j2;
        for (i2 = k + 1; i2 <= rmax; ++i2) {
            double f = LU(i2, k);
            if (f == 0.0) continue;
            f /= piv;
            LU(i2, k) = f;
            for (j2 = k + 1; j2 <= rmax; ++j2) LU(i2, j2) -= f * LU(k, j2);
        }
    }
#undef LU
    *lu_out = lu; *bw_out = bw;
    return 0;
}

void orc_band_solve(int N, const double *lu, int bw, const double *b, double *x) {
    int W = 2 * bw + 1, r, k;
#define LU(rr, cc) lu[(size_t)(rr) * W + ((cc) - (rr) + bw)]
    for (r = 0; r < N; ++r) {
        double s = b[r];
        int k0 = r - bw > 0 ? r - bw : 0;
        for (k = k0; k < r; ++k) s -= LU(r, k) * x[k];
        x[r] = s;
    }
    for (r = N - 1; r >= 0; --r) {
        double s = x[r];
        int k1 = r + bw < N - 1 ? r + bw : N - 1;
        for (k = r + 1; k <= k1; ++k) s -= LU(r, k) * x[k];
        x[r] = s / LU(r, r);
    }
#undef LU
}

/* ---------- Laplace pre-solves and gas fields: src/cuda_Field.cu:182-236,
 *            646-678.  The reference runs these with its FP32 PCG; the
 *            oracle solves them exactly (FP64 banded LU). ---------- */
static int laplace_presolves(orc_case *c) {
    int N = c->A_size, k, g;
    if (orc_band_factor(N, c->Ai, c->Aj, c->Aval, &c->lu, &c->lu_bw)) return -20;
    double *b = xcalloc(N, sizeof(double)), *x = xcalloc(N, sizeof(double));
    c->lap_phi = xcalloc((size_t)c->CondNUMR * c->Gsize, sizeof(real));
    for (k = 0; k < c->CondNUMR; ++k) {
        int r;
        for (r = 0; r < N; ++r) b[r] = (double)c->cond_b[(size_t)k * N + r];
        orc_band_solve(N, c->lu, c->lu_bw, b, x);
        real *phi = c->lap_phi + (size_t)k * c->Gsize;
        for (g = 0; g < c->Gsize; ++g) {
            if (c->A_idx[g]) phi[g] = (real)x[c->A_idx[g] - 1];
            else if (c->CondID[g] == k + 1) phi[g] = 1;
        }
    }
    /* temperature field with the temperature operator */
    {
        double *lut; int bwt, r;
        if (orc_band_factor(N, c->Ai, c->Aj, c->Atemp_val, &lut, &bwt)) return -20;
        for (r = 0; r < N; ++r) b[r] = (double)c->temp_b[r];
        orc_band_solve(N, lut, bwt, b, x);
        free(lut);
        for (g = 0; g < c->Gsize; ++g) if (c->A_idx[g]) c->Temp[g] = (real)x[c->A_idx[g] - 1];
    }
    c->BackDen1 = xcalloc(c->Gsize, sizeof(real)); c->BackVel1 = xcalloc(c->Gsize, sizeof(real));
    c->BackDen2 = xcalloc(c->Gsize, sizeof(real)); c->BackVel2 = xcalloc(c->Gsize, sizeof(real));
    for (g = 0; g < c->Gsize; ++g) {
        if (!c->DensRegion[g]) continue;
        double T = (double)c->Temp[g];
        c->BackDen1[g] = (real)(NPERTORR * (double)c->bg_pres[0] / (T * 8.6142e-5 + DBLMIN));
        c->BackVel1[g] = (real)sqrt(T * 1.38e-23 / (double)c->bg_mass[0]);
        if (c->deck.gas == 2) {   /* Calculate_2GasPara, src/cuda_Field.cu:669-678 */
            c->BackDen2[g] = (real)(NPERTORR * (double)c->bg_pres[1] / (T * 8.6142e-5 + DBLMIN));
            c->BackVel2[g] = (real)sqrt(T * 1.38e-23 / (double)c->bg_mass[1]);
        }
    }
    free(b); free(x);
    return 0;
}

/* ---------- fluid (metastable) densities: src/Fluid.c:3-68 (uniform load)
 *            and the cell->node copy src/cuda_Fluid.cu:321-339 ---------- */
static void fluid_densities(orc_case *c) {
    int isp, i, j;
    c->FluidN = xcalloc((size_t)c->nfsp * c->Gsize, sizeof(real));
    for (isp = 0; isp < c->nfsp; ++isp)
        for (i = 0; i < c->ncx; ++i)
            for (j = 0; j < c->ncy; ++j) {
                real den = 0;
                if (c->PlasmaRegion[i * c->ncy + j] != 0 && c->deck.fl_loadtype == 0) den = c->fl_dens[isp];
                c->FluidN[(size_t)isp * c->Gsize + i * c->ngy + j] = den;
                if (i == c->ncx - 1 || j == c->ncy - 1)
                    c->FluidN[(size_t)isp * c->Gsize + (i + 1) * c->ngy + j + 1] = den;
            }
}

/* ---------- cross sections: src/PhysicsData.c:638-705 (FUNC_decay),
 *            1579-1718 (Argon_CrossSectionSET) ---------- */
static real func_decay(const float *mx, const float *my, int n, real in_x) {
    if (n > 1) {
        if (in_x < mx[0]) return my[0];
        if (in_x > mx[n - 1]) {
            if (sizeof(real) == sizeof(float)) {
                float inc = (float)(log10((double)(my[n - 2] / my[n - 1])) / log10((double)(mx[n - 1] / mx[n - 2])));
                return (float)((double)my[n - 1] * pow((double)mx[n - 1], (double)inc) / pow((double)in_x, (double)inc));
            } else {
                double inc = log10((double)my[n - 2] / (double)my[n - 1]) / log10((double)mx[n - 1] / (double)mx[n - 2]);
                return (real)((double)my[n - 1] * pow((double)mx[n - 1], inc) / pow((double)in_x, inc));
            }
        }
        int lo = 0, hi = n - 1, k;
        while (hi - lo > 1) { k = (hi + lo) >> 1; if (mx[k] > in_x) hi = k; else lo = k; }
        real h = (real)mx[hi] - (real)mx[lo];
        real a = ((real)mx[hi] - in_x) / h;
        real b = (in_x - (real)mx[lo]) / h;
        return a * (real)my[lo] + b * (real)my[hi];
    }
    return my[0];
}

static void argon_tables(orc_case *c) {
    int i, r;
    c->N_LOGX = 81;
    c->TnRct = 10;
    c->idLOGX = (real)1 / (real)0.1f;
    if (sizeof(real) != sizeof(float)) c->idLOGX = 10.0;
    c->xe = xcalloc(81, sizeof(real)); c->xee = xcalloc(81, sizeof(real));
    c->cx = xcalloc(81 * 10, sizeof(real));
    for (i = 0; i < 81; ++i) {
        if (sizeof(real) == sizeof(float)) c->xe[i] = (real)(0.1f * (float)(i - 40));
        else c->xe[i] = 0.1 * (double)(i - 40);
        c->xee[i] = (real)pow(10.0, (double)c->xe[i]);
    }
    /* reaction -> (data set, subtracted set, threshold) */
    const int set[7] = {1, 2, 3, 4, 5, 7, 9};
    const int sub[7] = {-1, 3, -1, -1, -1, -1, -1};
    const float th[7] = {0.0f, 11.55f, 11.548f, 15.76f, 4.22f, 0.0f, 0.0f};
    const double *fl = c->deck.coll_flag;
    for (r = 0; r < 10; ++r) { c->Flag[r] = (real)fl[r]; if (c->Flag[r] < 0) c->Flag[r] = 0; }
    for (r = 0; r < 7; ++r) {
        c->Th_e[r] = (real)th[r];
        for (i = 0; i < 81; ++i) {
            real v;
            if (c->xee[i] < c->Th_e[r]) v = 0;
            else {
                v = c->Flag[r] * func_decay(pic_xsec_x[set[r]], pic_xsec_y[set[r]], pic_xsec_n[set[r]], c->xee[i]);
                if (sub[r] >= 0)
                    v = v - c->Flag[r] * func_decay(pic_xsec_x[sub[r]], pic_xsec_y[sub[r]], pic_xsec_n[sub[r]], c->xee[i]);
            }
            c->cx[i * 10 + r] = v;
        }
    }
    real me = c->mass[0], mi = c->mass[1], mAr = c->bg_mass[0], mAs = c->fl_mass[0];
    c->mofM[0] = me / mAr; c->mofM[1] = me / mAr; c->mofM[2] = me / mAr; c->mofM[3] = me / mAr;
    c->mofM[4] = me / mAs; c->mofM[5] = mAr / mi; c->mofM[6] = mAr / mi; c->mofM[7] = me / mAs;
    c->mofM[8] = mAs / mAr; c->mofM[9] = mAr / mAs;
    c->RR[7] = c->Flag[7] * (real)2.0e-13; c->RR[8] = c->Flag[8] * (real)6.2e-16; c->RR[9] = c->Flag[9] * (real)3.0e-21;
}

/* ---------- O2 and Ar/O2 rate constants (src/PhysicsData.c:2406-2414,
 *            3256-3275): the reactions after the tabulated ones (58 of O2,
 *            68 of Ar/O2) have a rate constant, RR = Flag x k.  The oracle
 *            keeps only these for the two gases: it restates their fluid
 *            (orc_fluid_diag); their collision step is orc_mcc_o2. ---------- */
static void oxygen_rates(orc_case *c) {
    static const double k_o2[9] = {6.90e-16, 8.00e-18, 2.56e-17, 1.00e-18, 4.20e-17, 1.30e-22, 2.20e-24,
                                   5.50e-29, 2.20e-24};
    static const double k_aro2[20] = {2.00e-13, 2.70e-13, 6.90e-16, 8.00e-18, 2.56e-17, 1.00e-18, 4.20e-17,
                                      1.30e-22, 2.20e-24, 5.50e-29, 2.20e-24, 6.40e-18, 5.10e-17, 6.20e-16,
                                      3.00e-21, 8.10e-18, 4.10e-17, 1.01e-16, 1.14e-16, 1.12e-15};
    int o2 = c->deck.gas == 1, r;
    int ntab = o2 ? 58 : 68;
    c->TnRct = o2 ? 67 : 88;
    for (r = 0; r < c->TnRct; ++r) { c->Flag[r] = (real)c->deck.coll_flag[r]; if (c->Flag[r] < 0) c->Flag[r] = 0; }
    for (r = ntab; r < c->TnRct; ++r) c->RR[r] = (real)((double)c->Flag[r] * (o2 ? k_o2 : k_aro2)[r - ntab]);
}

/* ---------- null-collision constants: src/cuda_mcc.cu:202-291 ---------- */
static void null_collision(orc_case *c) {
    int i, g;
    real sv[3] = {0, 0, 0};
    for (i = 0; i < c->N_LOGX; ++i) {
        const real *row = c->cx + i * 10;
        double e = (double)c->xee[i];
        real s0 = row[0] + row[1] + row[2] + row[3];
        real s1 = row[4];
        real s2 = row[5] + row[6];
        real v0 = (real)(sqrt(2 * 1.602e-19 * e / (double)c->mass[0]) * (double)s0);
        real v1 = (real)(sqrt(2 * 1.602e-19 * e / (double)c->mass[0]) * (double)s1);
        real v2 = (real)(sqrt(2 * 1.602e-19 * e / (double)c->mass[1]) * (double)s2);
        if (v0 > sv[0]) sv[0] = v0;
        if (v1 > sv[1]) sv[1] = v1;
        if (v2 > sv[2]) sv[2] = v2;
    }
    c->sigv[0] = sv[0]; c->sigv[1] = sv[1]; c->sigv[2] = sv[2];
    double t0 = 1.0 - exp(-1 * (double)c->dt * (double)sv[0] * (double)c->bg_dens[0]);
    double t1 = 1.0 - exp(-1 * (double)c->dt * (double)sv[1] * (double)c->fl_dens[0]);
    real colprob = (real)t0 + (real)t1;
    if (colprob > (real)0.08) {
        c->MCCn = (int)ceil((double)(colprob / (real)0.08));
        real ratio = (real)1 / (real)c->MCCn;
        c->dt_mcc = c->dt * ratio;
    } else {
        c->MCCn = 1; c->dt_mcc = c->dt;
    }
    /* per-node probabilities (src/cuda_mccAr.cu:55-57, 187) evaluated once */
    c->P1 = xcalloc(c->Gsize, sizeof(real)); c->P2 = xcalloc(c->Gsize, sizeof(real));
    c->Pion = xcalloc(c->Gsize, sizeof(real));
    for (g = 0; g < c->Gsize; ++g) {
        double dtm = (double)c->dt_mcc, dtd = (double)c->dt;
        real p1 = (real)(1.0 - dm_exp(-dtm * (double)c->sigv[0] * (double)c->BackDen1[g]));
        real p2b = (real)(1.0 - dm_exp(-dtm * (double)c->sigv[1] * (double)c->FluidN[g]));
        c->P1[g] = p1;
        c->P2[g] = p1 + p2b;
        c->Pion[g] = (real)(1.0 - dm_exp(-dtd * (double)c->sigv[2] * (double)c->BackDen1[g]));
    }
}

/* ---------- Maxwellian table and loader: src/Particle.c:3-50, 176-230 ---------- */
typedef struct { long seed; } orc_lcg;
static float orc_frand(orc_lcg *g) {
    long a = 16807, m = 2147483647, q = 127773, r = 2836;
    long hi = g->seed / q, lo = g->seed - q * hi;
    g->seed = a * lo - r * hi;
    if (g->seed <= 0) g->seed = g->seed + m;
    return (float)(g->seed / 2147483646.0);
}
static float orc_F(float v) {
    /* float products as in src/Particle.c:227-230 */
    float m2v = -2 * v, mvv = -v * v;
    return (float)((double)m2v * exp((double)mvv) / sqrt(ORC_PI) + erf((double)v));
}
static void maxwell_table(orc_case *c) {
    int nvel = (int)(1. / (double)(1 - orc_F((float)NVTS)));
    float dv = (float)(sqrt(ORC_PI) / (4.0 * nvel));
    c->nvel = nvel;
    c->vsave = xcalloc(nvel, sizeof(real));
    int i = 0, n;
    for (n = 0; n < nvel; ++n) {
        float rr = (float)((1.0 * n) / nvel);
        while (orc_F((float)i * dv) < rr) i++;
        c->vsave[n] = (real)(float)(sqrt(2.0) * i * (double)dv);
    }
}
/* one maxwellv() draw: three frand() calls in the reference's order */
static void orc_maxwellv(orc_case *c, orc_lcg *g, float vti, float *vx, float *vy, float *vz) {
    int n = (int)((c->nvel - 1) * orc_frand(g));
    float aphi = (float)(2 * ORC_PI * orc_frand(g));
    float costheta = 1 - 2 * orc_frand(g);
    float sintheta = (float)sqrt((double)(1 - costheta * costheta));
    float vs = (float)c->vsave[n];
    *vx = (float)((double)(vti * vs * sintheta) * cos((double)aphi));
    *vy = (float)((double)(vti * vs * sintheta) * sin((double)aphi));
    *vz = vti * vs * costheta;
}

static int load_particles(orc_case *c) {
    const orc_deck *d = &c->deck;
    orc_lcg g; g.seed = (long)d->seed;
    int isp, i, j, n;
    if (d->sp_loadtype != 0) return -21;   /* UNIFORM only (the configs use it) */
    for (isp = 0; isp < c->nsp; ++isp) {
        orc_species *s = &c->sp[isp];
        float dxf = (float)c->dx, dyf = (float)c->dy, zl = (float)c->zlength;
        float ninit = (float)d->sp_dens[isp];
        int npc = (int)(ninit * dxf * dyf * zl / (float)d->sp_np2c[isp]);
        float y0 = (float)d->sp_y0 / dyf, y1 = (float)d->sp_y1 / dyf - 1;
        float x0 = (float)d->sp_x0 / dxf, x1 = (float)d->sp_x1 / dxf - 1;
        /* count first */
        long total = 0;
        for (i = 0; i < c->ncx; ++i)
            for (j = 0; j < c->ncy; ++j)
                if (c->PlasmaRegion[i * c->ncy + j] == 1 && i >= x0 && i <= x1 && j >= y0 && j <= y1) total += npc;
        if (total > d->np_limit) return -22;
        int cap = d->sp_maxnp[isp] > total ? d->sp_maxnp[isp] : (int)total;
        s->cap = cap;
        s->x = xcalloc(cap, sizeof(real)); s->y = xcalloc(cap, sizeof(real));
        s->vx = xcalloc(cap, sizeof(real)); s->vy = xcalloc(cap, sizeof(real)); s->vz = xcalloc(cap, sizeof(real));
        s->id = xcalloc(cap, sizeof(uint32_t));
        s->cell_start = xcalloc(c->Gsize + 1, sizeof(int));
        int *cnt = xcalloc(c->Gsize, sizeof(int));
        int *cellof = xcalloc(total ? total : 1, sizeof(int));
        float *tx = xcalloc(total ? total : 1, sizeof(float)), *ty = xcalloc(total ? total : 1, sizeof(float));
        float *tvx = xcalloc(total ? total : 1, sizeof(float)), *tvy = xcalloc(total ? total : 1, sizeof(float));
        float *tvz = xcalloc(total ? total : 1, sizeof(float));
        int idx = 0;
        float vti = (float)c->vti[isp];
        for (i = 0; i < c->ncx; ++i)
            for (j = 0; j < c->ncy; ++j) {
                if (c->PlasmaRegion[i * c->ncy + j] != 1) continue;
                int num = (i >= x0 && i <= x1 && j >= y0 && j <= y1) ? npc : 0;
                for (n = 0; n < num; ++n) {
                    float vx, vy, vz;
                    orc_maxwellv(c, &g, vti, &vx, &vy, &vz);
                    float px = (float)i + orc_frand(&g);
                    float py = (float)j + orc_frand(&g);
                    /* Copy_HCPtoGCP (src/cuda_Particle.cu:41-56): bucket of
                     * (int)x, offset x - XID.  A position that rounds onto the
                     * far wall would index a node outside the cell range; it is
                     * kept in the last cell with offset 1. */
                    int xid = (int)px, yid = (int)py;
                    if (xid > c->ncx - 1) xid = c->ncx - 1;
                    if (yid > c->ncy - 1) yid = c->ncy - 1;
                    tx[idx] = px - (float)xid; ty[idx] = py - (float)yid;
                    tvx[idx] = vx; tvy[idx] = vy; tvz[idx] = vz;
                    cellof[idx] = xid * c->ngy + yid;
                    cnt[cellof[idx]]++;
                    idx++;
                }
            }
        s->cell_start[0] = 0;
        for (i = 0; i < c->Gsize; ++i) s->cell_start[i + 1] = s->cell_start[i] + cnt[i];
        /* stable placement by cell in load order (= the reference bucket order) */
        for (i = 0; i < c->Gsize; ++i) cnt[i] = s->cell_start[i];
        for (n = 0; n < idx; ++n) {
            int dst = cnt[cellof[n]]++;
            s->x[dst] = (real)tx[n]; s->y[dst] = (real)ty[n];
            s->vx[dst] = (real)tvx[n]; s->vy[dst] = (real)tvy[n]; s->vz[dst] = (real)tvz[n];
            s->id[dst] = (uint32_t)n;
        }
        s->n = idx;
        s->next_id = (uint32_t)idx;
        free(cellof); free(tx); free(ty); free(tvx); free(tvy); free(tvz);
        free(cnt);
    }
    return 0;
}

/* ================= metastable fluid: src/Fluid.c:3-256, src/cuda_Fluid.cu:3-33, 51-90, 179-301 ===== */
/* Ar_meta_omega (src/cuda_Fluid.cu:261-283), loop without a break included */
static float meta_omega(float value) {
    static const float ts[27] = {0.250f, 0.267f, 0.286f, 0.308f, 0.333f, 0.364f, 0.400f, 0.444f, 0.500f,
                                 0.571f, 0.667f, 0.800f, 1.000f, 1.330f, 2.000f, 2.500f, 2.860f, 3.330f,
                                 4.000f, 5.000f, 6.670f, 10.00f, 12.50f, 16.67f, 25.00f, 50.00f, 100.0f};
    static const float om[27] = {1.8714f, 1.8344f, 1.7962f, 1.7560f, 1.7142f, 1.6702f, 1.6238f, 1.5748f, 1.5228f,
                                 1.4676f, 1.4090f, 1.3466f, 1.2802f, 1.2096f, 1.1348f, 1.1042f, 1.0888f, 1.0738f,
                                 1.0592f, 1.0450f, 1.0318f, 1.0198f, 1.0154f, 1.0112f, 1.0072f, 1.0034f, 1.000f};
    float t, r = (float)DBL_MIN;
    int i;
    if (value < ts[0]) r = om[0];
    else if (value > ts[26]) r = om[26];
    else
        for (i = 1; i < 27; i++)
            if (value < ts[i]) { t = (value - ts[i - 1]) / (ts[i] - ts[i - 1]); r = om[i] * t + om[i - 1] * (1 - t); }
    return r;
}
float orc_meta_omega(float v) { return meta_omega(v); }

/* Cal_XRegion_check / Cal_YRegion_check / Set_Con_Region (src/Fluid.c:188-256) */
static void fluid_lines(orc_case *c) {
    int i, j, k, ncx = c->ncx, ncy = c->ncy;
#define PL(i, j) (c->PlasmaRegion[(i) * ncy + (j)] == 1)
    c->nlx = 0;
    for (j = 0; j < ncy; j++)
        for (i = 0; i < ncx; i++)
            if (PL(i, j) && (i == ncx - 1 || !PL(i + 1, j))) c->nlx++;
    c->nly = 0;
    for (i = 0; i < ncx; i++)
        for (j = 0; j < ncy; j++)
            if (PL(i, j) && (j == ncy - 1 || !PL(i, j + 1))) c->nly++;
    c->lx_x1 = xcalloc(c->nlx + 1, sizeof(int)); c->lx_x2 = xcalloc(c->nlx + 1, sizeof(int));
    c->lx_y = xcalloc(c->nlx + 1, sizeof(int));
    c->ly_x = xcalloc(c->nly + 1, sizeof(int)); c->ly_y1 = xcalloc(c->nly + 1, sizeof(int));
    c->ly_y2 = xcalloc(c->nly + 1, sizeof(int));
    k = 0;
    for (j = 0; j < ncy; j++)
        for (i = 0; i < ncx; i++) {
            if (!PL(i, j)) continue;
            if (i == 0 || !PL(i - 1, j)) { c->lx_x1[k] = i; c->lx_y[k] = j; }
            if (i == ncx - 1 || !PL(i + 1, j)) { c->lx_x2[k] = i; k++; }
        }
    k = 0;
    for (i = 0; i < ncx; i++)
        for (j = 0; j < ncy; j++) {
            if (!PL(i, j)) continue;
            if (j == 0 || !PL(i, j - 1)) { c->ly_x[k] = i; c->ly_y1[k] = j; }
            if (j == ncy - 1 || !PL(i, j + 1)) { c->ly_y2[k] = j; k++; }
        }
#undef PL
}

/* Set_Con_Boundary (src/Fluid.c:113-186): every branch gives Gamma1 except a
 * Neumann node at a line end on the domain edge, which gives 0 */
static void fluid_ends(orc_case *c) {
    int f, k, ngy = c->ngy;
    c->lx_f1 = xcalloc((size_t)c->nfsp * c->nlx + 1, sizeof(real)); c->lx_f2 = xcalloc((size_t)c->nfsp * c->nlx + 1, sizeof(real));
    c->ly_f1 = xcalloc((size_t)c->nfsp * c->nly + 1, sizeof(real)); c->ly_f2 = xcalloc((size_t)c->nfsp * c->nly + 1, sizeof(real));
    for (f = 0; f < c->nfsp; f++) {
        real g = (real)(float)c->deck.fl_quench[f];
        for (k = 0; k < c->nlx; k++) {
            int x1 = c->lx_x1[k], x2 = c->lx_x2[k], yy = c->lx_y[k];
            int a = c->Boundary[x1 * ngy + yy] == ORC_NEUMANN || c->Boundary[x1 * ngy + yy + 1] == ORC_NEUMANN;
            int b = c->Boundary[(x2 + 1) * ngy + yy] == ORC_NEUMANN || c->Boundary[(x2 + 1) * ngy + yy + 1] == ORC_NEUMANN;
            c->lx_f1[f * c->nlx + k] = (a && x1 == 0) ? 0 : g;
            c->lx_f2[f * c->nlx + k] = (b && x2 == c->ncx - 1) ? 0 : g;
        }
        for (k = 0; k < c->nly; k++) {
            int xx = c->ly_x[k], y1 = c->ly_y1[k], y2 = c->ly_y2[k];
            int a = c->Boundary[xx * ngy + y1] == ORC_NEUMANN || c->Boundary[(xx + 1) * ngy + y1] == ORC_NEUMANN;
            int b = c->Boundary[xx * ngy + y2 + 1] == ORC_NEUMANN || c->Boundary[(xx + 1) * ngy + y2 + 1] == ORC_NEUMANN;
            c->ly_f1[f * c->nly + k] = (a && y1 == 0) ? 0 : g;
            c->ly_f2[f * c->nly + k] = (b && y2 == c->ncy - 1) ? 0 : g;
        }
    }
}

/* Cal_D_Argon / Cal_D_Oxygen / Cal_D_ArO2 (src/cuda_Fluid.cu:199-260) */
static void fluid_diffusion(orc_case *c) {
    int f, i, j, ncx = c->ncx, ncy = c->ncy, ngy = c->ngy;
    real P = (real)c->deck.total_pres;
    c->fD = xcalloc((size_t)c->nfsp * ncx * ncy, sizeof(real));
    for (f = 0; f < c->nfsp; f++)
        for (i = 0; i < ncx; i++)
            for (j = 0; j < ncy; j++) {
                int G = i * ngy + j;
                real T = (real)(0.25 * (c->Temp[G] + c->Temp[G + 1] + c->Temp[G + ngy] + c->Temp[G + ngy + 1]));
                double D;
                int kind = c->deck.gas == 0 ? 0 : c->deck.gas == 1 ? (f <= 1 ? 1 : 2) : (f == 0 ? 3 : f <= 2 ? 1 : 2);
                if (kind == 0) {
                    real omega = (real)meta_omega((float)(T / (real)124.0f));
                    D = 1.997279e-4 * sqrt(T / 39.948) * T / (P * (real)11.682724f * omega) * 20;
                } else if (kind == 1) {
                    D = 2.63e-3 * sqrt((double)(T * T * T * 24 / 32 / 32)) / P * (real)0.08650f;
                } else if (kind == 2) {
                    D = 2.63e-3 * sqrt((double)(T * T * T * 24 / 16 / 16)) / P * (real)0.10892f;
                } else {
                    D = 2.63e-3 * sqrt((double)(T * T * T * 24) / 39.948 / 39.948) / P * (real)0.08559f;
                }
                c->fD[(size_t)f * ncx * ncy + i * ncy + j] = (real)D;
            }
}

/* calculate_gummel_coef_x/_y (src/cuda_Fluid.cu:179-198): interior x faces
 * average D[i-1] with itself, as the reference does */
static void fluid_gummel(orc_case *c) {
    int f, k, i, j, ncx = c->ncx, ncy = c->ncy, ngx = c->ngx, ngy = c->ngy;
    size_t NX = (size_t)ngx * ncy, NY = (size_t)ncx * ngy, NC = (size_t)ncx * ncy;
    c->fax = xcalloc(c->nfsp * NX, sizeof(real)); c->fbx = xcalloc(c->nfsp * NX, sizeof(real));
    c->fay = xcalloc(c->nfsp * NY, sizeof(real)); c->fby = xcalloc(c->nfsp * NY, sizeof(real));
    for (f = 0; f < c->nfsp; f++) {
        real *ax = c->fax + f * NX, *bx = c->fbx + f * NX, *ay = c->fay + f * NY, *by = c->fby + f * NY;
        const real *D = c->fD + f * NC;
        double m = (double)c->fl_mass[f];
        for (k = 0; k < c->nlx; k++) {
            int x1 = c->lx_x1[k], x2 = c->lx_x2[k], yy = c->lx_y[k];
            bx[x1 * ncy + yy] = (real)(-c->lx_f1[f * c->nlx + k] * (-0.25 * sqrt(c->Temp[x1 * ngy + yy] * 1.38e-23 / m)));
            ax[x1 * ncy + yy] = 0;
            for (i = x1 + 1; i <= x2; i++) {
                ax[i * ncy + yy] = (D[(i - 1) * ncy + yy] + D[(i - 1) * ncy + yy]) / 2 / c->dx;
                bx[i * ncy + yy] = ax[i * ncy + yy];
            }
            ax[(x2 + 1) * ncy + yy] = (real)(c->lx_f2[f * c->nlx + k] * (0.25 * sqrt(c->Temp[x2 * ngy + yy] * 1.38e-23 / m)));
            bx[(x2 + 1) * ncy + yy] = 0;
        }
        for (k = 0; k < c->nly; k++) {
            int xx = c->ly_x[k], y1 = c->ly_y1[k], y2 = c->ly_y2[k];
            by[xx * ngy + y1] = (real)(-c->ly_f1[f * c->nly + k] * (-0.25 * sqrt(c->Temp[xx * ngy + y1] * 1.38e-23 / m)));
            ay[xx * ngy + y1] = 0;
            for (j = y1 + 1; j <= y2; j++) {
                ay[xx * ngy + j] = (D[xx * ncy + j - 1] + D[xx * ncy + j]) / 2 / c->dy;
                by[xx * ngy + j] = ay[xx * ngy + j];
            }
            ay[xx * ngy + y2 + 1] = (real)(c->ly_f2[f * c->nly + k] * (0.25 * sqrt(c->Temp[xx * ngy + y2] * 1.38e-23 / m)));
            by[xx * ngy + y2 + 1] = 0;
        }
    }
}

void orc_fluid_flux_x(orc_case *c, int f);
void orc_fluid_flux_y(orc_case *c, int f);

static void fluid_setup(orc_case *c) {
    int f, i, j, ncx = c->ncx, ncy = c->ncy;
    size_t NC = (size_t)ncx * ncy;
    c->dtc = c->dt * (real)c->deck.dt_conti;
    fluid_lines(c);
    fluid_ends(c);
    fluid_diffusion(c);
    fluid_gummel(c);
    c->fden = xcalloc(c->nfsp * NC, sizeof(real));
    c->fsrc = xcalloc(c->nfsp * NC, sizeof(real));
    c->fflx = xcalloc((size_t)c->nfsp * c->ngx * ncy, sizeof(real));
    c->ffly = xcalloc((size_t)c->nfsp * ncx * c->ngy, sizeof(real));
    for (f = 0; f < c->nfsp; f++)
        for (i = 0; i < ncx; i++)
            for (j = 0; j < ncy; j++)
                if (c->PlasmaRegion[i * ncy + j] != 0 && c->deck.fl_loadtype == 0) c->fden[f * NC + i * ncy + j] = c->fl_dens[f];
    for (f = 0; f < c->nfsp; f++) { orc_fluid_flux_x(c, f); orc_fluid_flux_y(c, f); }
    c->win_den = xcalloc(c->Gsize, sizeof(real));
    c->mccr_mark = xcalloc((size_t)c->Gsize * (c->TnRct > 0 ? c->TnRct : 1), sizeof(int64_t));
}

int orc_setup(orc_case *c) {
    int e = derive_constants(c);
    if (e) return e;
    geometry(c);
    field_matrix(c);
    e = laplace_presolves(c);
    if (e) return e;
    fluid_densities(c);
    if (c->deck.gas != 0) {
        /* oxygen gases: backgrounds, fluid and rate constants (no particles) */
        oxygen_rates(c);
        fluid_setup(c);
        return 0;
    }
    argon_tables(c);
    null_collision(c);
    maxwell_table(c);
    fluid_setup(c);
    e = load_particles(c);
    if (e) return e;
    return 0;
}
