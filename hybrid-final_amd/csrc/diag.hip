// Per-step diagnostics of one case, in FP64 (src/cuda_Diagnostic.cu:37-267):
// particle counts, kinetic energy per species, the surface charge of every
// conductor (the reference's Surf_charge: sum of Sigma x Area over the
// conductor's nodes, src/cuda_Diagnostic.cu:11-35, 136-179), the particles
// each conductor absorbed since the previous sample (the reference's
// per-step CondCharge, from which it forms the conduction current
// q_density CondCharge / dt, src/cuda_Diagnostic.cu:181-190) and the
// conductor voltages.  One sample appends one record to a ring kept on the
// device; nothing is read back until the caller asks for the history.
//
// Two launches per sample: one wave per cell sums m v^2 / 2 of its
// particles in FP64 (one grid row per species), then one block adds the
// per-cell partials, the conductor charges and the absorption counts, each
// in a fixed order, so a record has the same bits in every run.
#include <hip/hip_runtime.h>

#include "pic_kernels.h"

namespace pic {

namespace {

constexpr int DIAG_BLOCK = 256;
// one thread per (species, conductor) absorption counter in k_diag_record
static_assert(5 * MAX_SRC <= DIAG_BLOCK, "absorption counters exceed the record block");

__device__ inline double wave_sum_d(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

template <typename R>
__global__ __launch_bounds__(256) void k_diag_cell_energy(DiagConst d, PSet<R> ps, double *part) {
    int lane;
    const int c = wave_cell(d.ncy, d.ngy, d.Csize, lane);
    if (c < 0) return;
    const int w = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    const PView<R> &p = ps.s[blockIdx.y];
    int k0, k1;
    cell_range(p, c, k0, k1);
    double e = 0.0;
    for (int k = k0 + lane; k < k1; k += 64) {
        const double vx = (double)p.vx[k], vy = (double)p.vy[k], vz = (double)p.vz[k];
        e += (vx * vx + vy * vy) + vz * vz;
    }
    e = wave_sum_d(e);
    if (lane == 0) part[(size_t)blockIdx.y * d.Csize + w] = e;
}

// fixed-tree block sum; every thread returns the total
__device__ inline double block_sum_d(double v, double *sh) {
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int off = DIAG_BLOCK / 2; off > 0; off >>= 1) {
        if (t < off) sh[t] += sh[t + off];
        __syncthreads();
    }
    const double r = sh[0];
    __syncthreads();
    return r;
}

// Record layout (doubles): t, step, np[nsp], ke[nsp], q_cond[ncond],
// absorbed[nsp][ncond] since the previous sample, v_cond[ncond].
template <typename R>
__global__ __launch_bounds__(DIAG_BLOCK) void k_diag_record(DiagConst d, PSet<R> ps, const double *part,
                                                            const R *Sigma, const R *Area, const int *CondID,
                                                            const int *condcharge, int *cc_prev, const R *V,
                                                            long long step, double *rec) {
    __shared__ double sh[DIAG_BLOCK];
    const int t = threadIdx.x;
    int o = 0;
    if (t == 0) {
        rec[0] = (double)step * d.dt;
        rec[1] = (double)step;
    }
    o = 2;
    for (int s = 0; s < d.nsp; ++s) {
        if (t == 0) rec[o + s] = (double)min(ps.s[s].cell_start[d.G], ps.s[s].cap);
    }
    o += d.nsp;
    for (int s = 0; s < d.nsp; ++s) {
        double a = 0.0;
        for (int c = t; c < d.Csize; c += DIAG_BLOCK) a += part[(size_t)s * d.Csize + c];
        const double sum = block_sum_d(a, sh);
        if (t == 0) rec[o + s] = 0.5 * d.mass[s] * d.np2c[s] * sum;
    }
    o += d.nsp;
    for (int i = 0; i < d.ncond; ++i) {
        double a = 0.0;
        for (int g = t; g < d.G; g += DIAG_BLOCK)
            if (CondID[g] == i + 1) a += (double)Sigma[g] * (double)Area[g];
        const double sum = block_sum_d(a, sh);
        if (t == 0) rec[o + i] = sum;
    }
    o += d.ncond;
    if (t < d.nsp * d.ncond) {
        const int now = condcharge[t];
        rec[o + t] = (double)(now - cc_prev[t]);
        cc_prev[t] = now;
    }
    o += d.nsp * d.ncond;
    if (t < d.ncond) rec[o + t] = (double)V[t];
}

}  // namespace

int diag_record_len(int nsp, int ncond) { return 2 + 2 * nsp + ncond + nsp * ncond + ncond; }

template <typename R>
void launch_diag_sample(const DiagConst &d, const PSet<R> &ps, double *part, const R *Sigma, const R *Area,
                        const int *CondID, const int *condcharge, int *cc_prev, const R *V, long long step,
                        double *rec, hipStream_t s) {
    hipLaunchKernelGGL(k_diag_cell_energy<R>, dim3((d.Csize + 3) / 4, d.nsp), dim3(256), 0, s, d, ps, part);
    hipLaunchKernelGGL(k_diag_record<R>, dim3(1), dim3(DIAG_BLOCK), 0, s, d, ps, (const double *)part, Sigma, Area,
                       CondID, condcharge, cc_prev, V, step, rec);
}

#define INST(R)                                                                                                       \
    template void launch_diag_sample<R>(const DiagConst &, const PSet<R> &, double *, const R *, const R *,          \
                                        const int *, const int *, int *, const R *, long long, double *,              \
                                        hipStream_t);
INST(float)
INST(double)
#undef INST

}  // namespace pic
