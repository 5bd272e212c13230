// Metastable fluid on the device (the reference's host code, run during its
// warm-up loop: src/cuda_Fluid.cu:34-178 continuity solve,
// src/cuda_Diagnostic.cu:37-110, 534-560, 634-642 source update and sync).
//
// The continuity solve is alternating-direction implicit with
// Scharfetter-Gummel face fluxes: one tridiagonal system per line of plasma
// cells, first along x (fixed y) then along y.  One thread takes one line
// of one fluid species: it forms the system's rows on the fly, runs the
// Thomas recurrence with its intermediate values in a scratch array laid
// out [row][line] (adjacent lines' threads touch adjacent words), writes the
// line's densities and then the line's face fluxes.  The lines of a sweep
// are independent; the arithmetic is the reference's, operation for
// operation, in the case's precision.
//
// The source update averages the event rates and the electron density over
// the reference's window (N_ave accumulations, then an averaging call that
// adds its own step), forms every fluid species' source per node and hands
// it to the cells; a one-thread kernel then advances the window counter and
// switches the solve on.  Once on, the cell densities go to the nodes every
// step; for argon the electron null-collision probability P2 of every node
// follows them, and the oxygen collision steps read the node densities of
// their fluid species (O2A, O2B, OP, OD; Ar/O2 with Ar* first) directly.
//
// The oxygen sources are table driven: which species' Denscale scales each
// reaction's averaged count, the rate-constant reactions as products of two
// node densities, and per fluid species its production and loss terms in
// the reference's order of summation (src/cuda_Diagnostic.cu:561-744).
#include <hip/hip_runtime.h>

#include "detmath.h"
#include "pic_kernels.h"

namespace pic {

namespace {

constexpr int FL_BLOCK = 64;

template <typename R>
__global__ __launch_bounds__(FL_BLOCK) void k_fl_sweep_x(FluidDev<R> d) {
    if (!d.state->conti) return;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int L = d.nx * d.nf;
    if (t >= L) return;
    const int f = t / d.nx, k = t % d.nx;
    const int ncy = d.ncy, ngy = d.ngy;
    const size_t NC = (size_t)d.ncx * ncy, NX = (size_t)d.ngx * ncy, NY = (size_t)d.ncx * ngy;
    R *n = d.den + f * NC;
    const R *ax = d.ax + f * NX, *bx = d.bx + f * NX, *fy = d.flux_y + f * NY, *src = d.src + f * NC;
    R *fx = d.flux_x + f * NX;
    const int x1 = d.x_x1[k], x2 = d.x_x2[k], yy = d.x_y[k], num = x2 - x1 + 1;
    const R del_f = d.dtc / d.dx;
    R *xs = d.xs, *gs = d.gs;
    // forward recurrence, rows formed on the fly (Solve_Density_x, tridiag)
    R bet = 0, xprev = 0, cprev = 0;
    bool ok = true;
    for (int i = 0; i < num; ++i) {
        const int xx = i + x1;
        const R ga = -ax[xx * ncy + yy] * del_f;
        const R gb = 1 + (ax[(xx + 1) * ncy + yy] + bx[xx * ncy + yy]) * del_f;
        const R gc = -bx[(xx + 1) * ncy + yy] * del_f;
        const R rhs = n[xx * ncy + yy] + src[xx * ncy + yy] * d.dtc - (fy[xx * ngy + yy + 1] - fy[xx * ngy + yy]) * d.dtc / d.dy;
        R xi;
        if (i == 0) {
            bet = gb;
            xi = rhs / bet;
        } else {
            gs[(size_t)(i - 1) * L + t] = cprev / bet;
            bet = gb - ga * cprev / bet;
            xi = (rhs - ga * xprev) / bet;
            if (bet == 0) { ok = false; break; }
        }
        xs[(size_t)i * L + t] = xi;
        xprev = xi;
        cprev = gc;
    }
    if (ok) {
        // back substitution, then the line's densities
        R xn = xs[(size_t)(num - 1) * L + t];
        for (int i = num - 2; i >= 0; --i) {
            R xi = xs[(size_t)i * L + t];
            xi -= gs[(size_t)i * L + t] * xn;
            xs[(size_t)i * L + t] = xi;
            xn = xi;
        }
        for (int i = 0; i < num; ++i) {
            const int c = (i + x1) * ncy + yy;
            const R xi = xs[(size_t)i * L + t];
            n[c] = (d.plasma[c] == 1 && xi > (R)1e9) ? xi : (R)0;
        }
    }
    // face fluxes of the line (Calculate_Flux_x)
    const R g1 = d.x_g1[(size_t)f * d.nx + k], g2 = d.x_g2[(size_t)f * d.nx + k];
    fx[x1 * ncy + yy] = -g1 * bx[x1 * ncy + yy] * n[x1 * ncy + yy];
    for (int i = x1 + 1; i <= x2; ++i) {
        const R left = n[(i - 1) * ncy + yy], right = n[i * ncy + yy];
        fx[i * ncy + yy] = ax[i * ncy + yy] * left - bx[i * ncy + yy] * right;
    }
    fx[(x2 + 1) * ncy + yy] = g2 * ax[(x2 + 1) * ncy + yy] * n[x2 * ncy + yy];
}

template <typename R>
__global__ __launch_bounds__(FL_BLOCK) void k_fl_sweep_y(FluidDev<R> d) {
    if (!d.state->conti) return;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int L = d.ny * d.nf;
    if (t >= L) return;
    const int f = t / d.ny, k = t % d.ny;
    const int ncy = d.ncy, ngy = d.ngy;
    const size_t NC = (size_t)d.ncx * ncy, NX = (size_t)d.ngx * ncy, NY = (size_t)d.ncx * ngy;
    R *n = d.den + f * NC;
    const R *ay = d.ay + f * NY, *by = d.by + f * NY, *fx = d.flux_x + f * NX, *src = d.src + f * NC;
    R *fy = d.flux_y + f * NY;
    const int xx = d.y_x[k], y1 = d.y_y1[k], y2 = d.y_y2[k], num = y2 - y1 + 1;
    const R del_f = d.dtc / d.dy;
    R *xs = d.xs, *gs = d.gs;
    R bet = 0, xprev = 0, cprev = 0;
    bool ok = true;
    for (int j = 0; j < num; ++j) {
        const int yy = j + y1;
        const R ga = -ay[xx * ngy + yy] * del_f;
        const R gb = 1 + (ay[xx * ngy + yy + 1] + by[xx * ngy + yy]) * del_f;
        const R gc = -by[xx * ngy + yy + 1] * del_f;
        const R rhs = n[xx * ncy + yy] + src[xx * ncy + yy] * d.dtc - (fx[(xx + 1) * ncy + yy] - fx[xx * ncy + yy]) * d.dtc / d.dx;
        R xj;
        if (j == 0) {
            bet = gb;
            xj = rhs / bet;
        } else {
            gs[(size_t)(j - 1) * L + t] = cprev / bet;
            bet = gb - ga * cprev / bet;
            xj = (rhs - ga * xprev) / bet;
            if (bet == 0) { ok = false; break; }
        }
        xs[(size_t)j * L + t] = xj;
        xprev = xj;
        cprev = gc;
    }
    if (ok) {
        R xn = xs[(size_t)(num - 1) * L + t];
        for (int j = num - 2; j >= 0; --j) {
            R xj = xs[(size_t)j * L + t];
            xj -= gs[(size_t)j * L + t] * xn;
            xs[(size_t)j * L + t] = xj;
            xn = xj;
        }
        for (int j = 0; j < num; ++j) {
            const int c = xx * ncy + j + y1;
            const R xj = xs[(size_t)j * L + t];
            n[c] = (d.plasma[c] == 1 && xj > (R)1e9) ? xj : (R)0;
        }
    }
    const R g1 = d.y_g1[(size_t)f * d.ny + k], g2 = d.y_g2[(size_t)f * d.ny + k];
    fy[xx * ngy + y1] = -g1 * by[xx * ngy + y1] * n[xx * ncy + y1];
    for (int j = y1 + 1; j <= y2; ++j) {
        const R lower = n[xx * ncy + j - 1], upper = n[xx * ncy + j];
        fy[xx * ngy + j] = ay[xx * ngy + j] * lower - by[xx * ngy + j] * upper;
    }
    fy[xx * ngy + y2 + 1] = g2 * ay[xx * ngy + y2 + 1] * n[xx * ncy + y2];
}

// ---------------------------------------------------------------- tables
// averaged count ranges: reactions [r0, r1) scaled by (Denscale[sp] / N_ave)
// / dt; late entries (one reaction) by Denscale[sp] x (1 / N_ave / dt)
struct FlRange { short r0, r1, sp, late; };
__constant__ FlRange kFlRanges[2][11] = {
    // O2 (src/cuda_Diagnostic.cu:561-590)
    {{0, 41, 0, 0}, {41, 47, 3, 0}, {47, 53, 1, 0}, {53, 58, 2, 0}, {58, 59, 3, 1}, {-1, 0, 0, 0}},
    // Ar/O2 (src/cuda_Diagnostic.cu:591-633)
    {{0, 46, 0, 0}, {46, 52, 4, 0}, {52, 60, 2, 0}, {60, 65, 3, 0}, {65, 68, 1, 0}, {69, 70, 3, 1}, {70, 71, 3, 1},
     {79, 80, 1, 1}, {80, 81, 1, 1}, {81, 82, 0, 1}, {-1, 0, 0, 0}},
};
// rate-constant reactions: M[r] = (RR[r] x a) x b, with a, b: fluid species
// 0-4, FL_B1 / FL_B2 the background densities, FL_EDEN 0.25 x the averaged
// electron density
constexpr short FL_B1 = 8, FL_B2 = 9, FL_EDEN = 10;
struct FlRc { short r, a, b; };
__constant__ FlRc kFlRc[2][16] = {
    {{59, 2, 3}, {60, 3, FL_B1}, {61, 3, FL_B1}, {62, 3, FL_B1}, {63, 0, 2}, {64, 0, FL_B1}, {65, 0, 0},
     {66, 1, FL_B1}, {-1, 0, 0}},
    {{68, FL_EDEN, 0}, {71, 3, 4}, {72, 4, FL_B2}, {73, 4, FL_B2}, {74, 4, FL_B2}, {75, 1, 3}, {76, 1, FL_B2},
     {77, 1, 1}, {78, 2, FL_B2}, {82, 0, FL_B1}, {83, 0, 3}, {84, 0, 3}, {85, 0, FL_B2}, {86, 0, FL_B2},
     {87, 0, FL_B2}, {-1, 0, 0}},
};
// per fluid species: production terms, then loss terms; an entry is the
// reaction, | FL_X2 for a term counted twice; -1 ends a list
constexpr short FL_X2 = 0x100;
constexpr int FL_TERMS = 34;
__constant__ short kFlSrc[2][5][2][FL_TERMS] = {
    {   // O2: O2A, O2B, OP, OD
        {{3, 61, -1},
         {14, 15, 16, 17, 18, 19, 20, 21, 46, 51, 56, 63, 64, 65 | FL_X2, -1}},
        {{4, 62, -1},
         {22, 23, 24, 25, 26, 27, 28, 29, 52, 57, 58, 66, -1}},
        {{6, 7 | FL_X2, 8, 11, 13, 15, 18 | FL_X2, 19, 21, 23, 26 | FL_X2, 27, 29, 30, 31, 40, 42, 44, 45 | FL_X2, 46,
          50, 53, 56, 57, 58, 59, 60, 61, 62, -1},
         {33, 34, 35, 36, 37, 38, 43, 47, -1}},
        {{8, 9 | FL_X2, 19, 20 | FL_X2, 27, 28 | FL_X2, 31, 33, -1},
         {39, 40, 59, 60, 61, 62, -1}},
        {{-1}, {-1}},
    },
    {   // Ar/O2: Ar*, O2A, O2B, OP, OD
        {{2, -1},
         {4, 68, 81 | FL_X2, 82, 83, 84, 85, 86, 87, -1}},
        {{8, 73, -1},
         {19, 20, 21, 22, 23, 24, 25, 26, 51, 56, 30, 75, 76, 77 | FL_X2, -1}},
        {{9, 74, -1},
         {27, 28, 29, 30, 31, 32, 33, 34, 57, 64, 70, 78, -1}},
        {{11, 12 | FL_X2, 13, 16, 18, 20, 23 | FL_X2, 24, 26, 28, 31 | FL_X2, 32, 34, 35, 36, 45, 47, 49, 50 | FL_X2,
          51, 55, 60, 63, 64, 70, 71, 72, 73, 74, 69, 85 | FL_X2, 86, -1},
         {38, 39, 40, 41, 42, 43, 48, 52, 79, 83, -1}},
        {{13, 14 | FL_X2, 24, 25 | FL_X2, 32, 33 | FL_X2, 36, 38, 83, 86, -1},
         {44, 45, 71, 72, 73, 74, -1}},
    },
};
constexpr int FL_MAXRCT = 88;

template <typename R>
__device__ inline R fl_sum(const short *list, const R *M) {
    R acc = 0;
    for (int i = 0; i < FL_TERMS && list[i] >= 0; ++i) {
        const R t = (list[i] & FL_X2) ? (R)2 * M[list[i] & 0xff] : M[list[i] & 0xff];
        acc = i == 0 ? t : acc + t;
    }
    return acc;
}

// per node: electron density into the window; on an averaging step the
// averaged rates and density, every fluid species' source of the node's cell
template <typename R>
__global__ __launch_bounds__(256) void k_fl_nodes(FluidDev<R> d, const R *den_e, const unsigned long long *mccr) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= d.G) return;
    const bool avg = d.state->nave == d.n_ave;
    R w = d.win[g] + den_e[g];
    if (!avg) {
        d.win[g] = w;
        return;
    }
    const R ave_e = d.denscale_e / (R)d.n_ave * w;
    d.win[g] = 0;
    const int T = d.TnRct;
    const size_t o = (size_t)g * T;
    const R Eden = (R)(0.25 * (double)ave_e);
    const int x = g / d.ngy, y = g % d.ngy;
    const bool cell = x < d.ncx && y < d.ncy;
    if (d.gas == 0) {
        const R rate = d.denscale_e / (R)d.n_ave / d.dt;
        R Rr[9];
#pragma unroll
        for (int r = 0; r < 9; ++r) Rr[r] = r == 7 ? (R)0 : rate * (R)(long long)(mccr[o + r] - d.mark[o + r]);
        for (int r = 0; r < T; ++r) d.mark[o + r] = mccr[o + r];
        const R nm = d.nodeN[g];
        const R k7 = d.RR7 * Eden * nm;
        const R k9 = d.RR9 * d.BackDen1[g] * nm;
        const R S = Rr[2] - Rr[4] - k7 - 2 * Rr[8] - k9;
        if (cell) d.src[x * d.ncy + y] = S;
        return;
    }
    const int gi = d.gas - 1;
    R M[FL_MAXRCT];
    for (int r = 0; r < FL_MAXRCT; ++r) M[r] = 0;
    for (int e = 0; e < 11 && kFlRanges[gi][e].r0 >= 0; ++e) {
        const FlRange q = kFlRanges[gi][e];
        if (q.late) {
            const R t = (R)1 / (R)d.n_ave / d.dt;
            M[q.r0] = d.dsc[q.sp] * t * (R)(long long)(mccr[o + q.r0] - d.mark[o + q.r0]);
        } else {
            const R t = d.dsc[q.sp] / (R)d.n_ave / d.dt;
            for (int r = q.r0; r < q.r1; ++r) M[r] = t * (R)(long long)(mccr[o + r] - d.mark[o + r]);
        }
    }
    for (int r = 0; r < T; ++r) d.mark[o + r] = mccr[o + r];
    const size_t G = d.G;
    auto fac = [&](int a) -> R {
        return a == FL_B1 ? d.BackDen1[g] : a == FL_B2 ? d.BackDen2[g] : a == FL_EDEN ? Eden : d.nodeN[a * G + g];
    };
    for (int e = 0; e < 16 && kFlRc[gi][e].r >= 0; ++e) {
        const FlRc q = kFlRc[gi][e];
        M[q.r] = d.RR[q.r] * fac(q.a) * fac(q.b);
    }
    if (!cell) return;
    const size_t NC = (size_t)d.ncx * d.ncy;
    for (int f = 0; f < d.nf; ++f) {
        const R sum = fl_sum<R>(kFlSrc[gi][f][0], M);
        const R diff = fl_sum<R>(kFlSrc[gi][f][1], M);
        d.src[f * NC + x * d.ncy + y] = sum - diff;
    }
}

__global__ void k_fl_state(FluidState *s, int n_ave) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    if (s->nave == n_ave) {
        s->nave = 0;
        s->conti = 1;
    } else {
        s->nave += 1;
    }
}

// cells -> nodes for every fluid species (Sync_Fluid_GFCtoGFG_forDen: node
// (i, j) takes cell (i, j); the last cell of a row or column also sets the
// node diagonally above it) and, for argon, the node's P2
template <typename R>
__global__ __launch_bounds__(256) void k_fl_sync(FluidDev<R> d, const R *P1, R *P2) {
    if (!d.state->conti) return;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= d.G * d.nf) return;
    const int f = t / d.G, g = t % d.G;
    const int x = g / d.ngy, y = g % d.ngy;
    const R *den = d.den + (size_t)f * d.ncx * d.ncy;
    R v = d.nodeN[t];
    if (x < d.ncx && y < d.ncy) v = den[x * d.ncy + y];
    else if (x >= 1 && y >= 1 && (x - 1 == d.ncx - 1 || y - 1 == d.ncy - 1)) v = den[(x - 1) * d.ncy + y - 1];
    d.nodeN[t] = v;
    if (d.gas == 0) P2[g] = P1[g] + (R)(1.0 - dm_exp(-d.dtm * d.sigv1 * (double)v));
}

}  // namespace

template <typename R>
void launch_fluid_step(const FluidDev<R> &d, const R *den_e, const unsigned long long *mccr, const R *P1, R *P2,
                       hipStream_t s) {
    const int lx = d.nx * d.nf, ly = d.ny * d.nf;
    if (lx > 0) hipLaunchKernelGGL(k_fl_sweep_x<R>, dim3((lx + FL_BLOCK - 1) / FL_BLOCK), dim3(FL_BLOCK), 0, s, d);
    if (ly > 0) hipLaunchKernelGGL(k_fl_sweep_y<R>, dim3((ly + FL_BLOCK - 1) / FL_BLOCK), dim3(FL_BLOCK), 0, s, d);
    const int nb = (d.G + 255) / 256;
    hipLaunchKernelGGL(k_fl_nodes<R>, dim3(nb), dim3(256), 0, s, d, den_e, mccr);
    hipLaunchKernelGGL(k_fl_state, dim3(1), dim3(64), 0, s, d.state, d.n_ave);
    hipLaunchKernelGGL(k_fl_sync<R>, dim3((d.G * d.nf + 255) / 256), dim3(256), 0, s, d, P1, P2);
}

template void launch_fluid_step<float>(const FluidDev<float> &, const float *, const unsigned long long *,
                                       const float *, float *, hipStream_t);
template void launch_fluid_step<double>(const FluidDev<double> &, const double *, const unsigned long long *,
                                        const double *, double *, hipStream_t);

}  // namespace pic
