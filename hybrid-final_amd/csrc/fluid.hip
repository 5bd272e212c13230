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
// adds its own step), forms the Ar* source per node and hands it to the
// cells; a one-thread kernel then advances the window counter and switches
// the solve on.  Once on, the cell densities go to the nodes every step and
// the electron null-collision probability P2 of every node follows them.
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

// per node: electron density into the window; on an averaging step the
// averaged rates and density, the Ar* source of the node's cell
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
    const R rate = d.denscale_e / (R)d.n_ave / d.dt;
    const int T = d.TnRct;
    R Rr[9];
#pragma unroll
    for (int r = 0; r < 9; ++r)
        Rr[r] = r == 7 ? (R)0 : rate * (R)(long long)(mccr[(size_t)g * T + r] - d.mark[(size_t)g * T + r]);
    for (int r = 0; r < T; ++r) d.mark[(size_t)g * T + r] = mccr[(size_t)g * T + r];
    const R Eden = (R)(0.25 * (double)ave_e);
    const R nm = d.nodeN[g];
    const R k7 = d.RR7 * Eden * nm;
    const R k9 = d.RR9 * d.BackDen1[g] * nm;
    const R S = Rr[2] - Rr[4] - k7 - 2 * Rr[8] - k9;
    const int x = g / d.ngy, y = g % d.ngy;
    if (x < d.ncx && y < d.ncy) d.src[x * d.ncy + y] = S;
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

// cells -> nodes (Sync_Fluid_GFCtoGFG_forDen: node (i, j) takes cell (i, j);
// the last cell of a row or column also sets the node diagonally above it)
// and the node's P2
template <typename R>
__global__ __launch_bounds__(256) void k_fl_sync(FluidDev<R> d, const R *P1, R *P2) {
    if (!d.state->conti) return;
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= d.G) return;
    const int x = g / d.ngy, y = g % d.ngy;
    R v = d.nodeN[g];
    if (x < d.ncx && y < d.ncy) v = d.den[x * d.ncy + y];
    else if (x >= 1 && y >= 1 && (x - 1 == d.ncx - 1 || y - 1 == d.ncy - 1)) v = d.den[(x - 1) * d.ncy + y - 1];
    d.nodeN[g] = v;
    P2[g] = P1[g] + (R)(1.0 - dm_exp(-d.dtm * d.sigv1 * (double)v));
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
    hipLaunchKernelGGL(k_fl_sync<R>, dim3(nb), dim3(256), 0, s, d, P1, P2);
}

template void launch_fluid_step<float>(const FluidDev<float> &, const float *, const unsigned long long *,
                                       const float *, float *, hipStream_t);
template void launch_fluid_step<double>(const FluidDev<double> &, const double *, const unsigned long long *,
                                        const double *, double *, hipStream_t);

}  // namespace pic
