// Field solve for separable operators: FP64 fast-Poisson inverse inside FP64
// iterative refinement.
//
// When the unknowns form a rectangle with uniform coefficients (every CCP
// deck with electrodes on two opposite walls), the field operator is
// A = cx Tx (x) I + I (x) cy Ty and
//     A^-1 R = Vx [ (Vx^-1 R Vy^-T) ./ (cx lx_a + cy ly_b) ] Vy^T,
// i.e. four small dense GEMMs (nx x nx x ny and nx x ny x ny) and an
// elementwise scale.  Each refinement cycle is
//     r = b - A x   (FP64 stencil, deterministic reduction of ||r||, ||b||)
//     x += A^-1 r   (FP64 GEMMs)
// and the solve stops on ||b - A x|| <= tol ||b||.  When the operator equals
// the separable model exactly, one cycle reaches ~1e-14 from any start; the
// cycle count is capped by the caller.  All kernels read a device-side done
// flag, so the solve is a fixed launch sequence (graph-capturable) that idles
// once converged.
//
// At these sizes (n <= ~1000 per side) the solve is bound by launch and
// memory latency, not by FLOPs: each GEMM block computes one 16x16 output
// tile on the FP64 matrix cores with K split over its 4 waves, so a 256^2
// problem runs ~256 blocks of 4 waves from L2-resident operands.
#include <hip/hip_runtime.h>

#include "dgemm_mfma.h"
#include "multi.h"
#include "pic_kernels.h"

namespace pic {

namespace {

constexpr int RES_BLOCK = 256;
constexpr int GT = 16;   // GEMM output tile edge

// r = b - A x in FP64; per-block partial sums of r.r and b.b.  The first
// residual of a solve (b32 != null) also starts the solve: it converts the
// right-hand side (FP32 or FP64) into b and resets the solve state.
template <typename RB>
struct ResidJob {
    PcgOperator A;
    const RB *b32;
    double *b;
    const double *x;
    double *r, *partials;
    SpecState *s;
};

template <typename RB>
__device__ __forceinline__ void residual64_body(const ResidJob<RB> &J, int bx) {
    const PcgOperator &A = J.A;
    const RB *b32 = J.b32;
    double *b = J.b, *r = J.r, *partials = J.partials;
    const double *x = J.x;
    SpecState *s = J.s;

    if (b32) {
        if (bx == 0 && threadIdx.x == 0) { s->done = 0; s->outer = 0; s->rel = 0.0; s->status = 0; }
    } else if (s->done) {
        return;
    }
    __shared__ double sh[2 * RES_BLOCK];
    const int i = bx * RES_BLOCK + threadIdx.x;
    double rr = 0.0, bb = 0.0;
    if (i < A.N) {
        if (b32) b[i] = (double)b32[i];
        const int o = 5 * i;
        double ax = A.vald[o] * x[A.col[o]];
        ax += A.vald[o + 1] * x[A.col[o + 1]];
        ax += A.vald[o + 2] * x[A.col[o + 2]];
        ax += A.vald[o + 3] * x[A.col[o + 3]];
        ax += A.vald[o + 4] * x[A.col[o + 4]];
        const double ri = b[i] - ax;
        r[i] = ri;
        rr = ri * ri;
        bb = b[i] * b[i];
    }
    sh[threadIdx.x] = rr;
    sh[RES_BLOCK + threadIdx.x] = bb;
    __syncthreads();
    for (int off = RES_BLOCK / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) {
            sh[threadIdx.x] += sh[threadIdx.x + off];
            sh[RES_BLOCK + threadIdx.x] += sh[RES_BLOCK + threadIdx.x + off];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partials[2 * bx] = sh[0];
        partials[2 * bx + 1] = sh[RES_BLOCK];
    }
}

template <typename RB>
__global__ __launch_bounds__(RES_BLOCK) void k_residual64(ResidJob<RB> J) { residual64_body<RB>(J, blockIdx.x); }
template <typename RB>
__global__ __launch_bounds__(RES_BLOCK) void k_residual64_m(const ResidJob<RB> *J, const uint32_t *map) {
    int lb;
    const ResidJob<RB> &a = J[multi_job(map, lb)];
    residual64_body<RB>(a, lb);
}


// one block: fixed-order sum of the partials, convergence decision
struct FinalizeJob {
    const double *partials;
    int nblk;
    double tol;
    int last;
    SpecState *s;
    PcgResult *res;
};

__device__ __forceinline__ void spec_finalize_body(const FinalizeJob &J) {
    const double *partials = J.partials;
    const int nblk = J.nblk, last = J.last;
    const double tol = J.tol;
    SpecState *s = J.s;
    PcgResult *res = J.res;

    if (s->done) return;
    __shared__ double sh[2 * RES_BLOCK];
    double rr = 0.0, bb = 0.0;
    for (int b = threadIdx.x; b < nblk; b += RES_BLOCK) {
        rr += partials[2 * b];
        bb += partials[2 * b + 1];
    }
    sh[threadIdx.x] = rr;
    sh[RES_BLOCK + threadIdx.x] = bb;
    __syncthreads();
    for (int off = RES_BLOCK / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) {
            sh[threadIdx.x] += sh[threadIdx.x + off];
            sh[RES_BLOCK + threadIdx.x] += sh[RES_BLOCK + threadIdx.x + off];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        rr = sh[0];
        bb = sh[RES_BLOCK];
        double rel = bb > 0.0 ? sqrt(rr / bb) : sqrt(rr);
        bool conv = bb > 0.0 ? rr <= tol * tol * bb : rr <= 1e-300;
        s->rel = rel;
        if (conv || last) {
            s->done = 1;
            s->status = conv ? 0 : 1;
            res->iters = s->outer;
            res->outer = s->outer;
            res->rel_res = rel;
            res->status = s->status;
        } else {
            s->outer += 1;
        }
    }
}

__global__ __launch_bounds__(RES_BLOCK) void k_spec_finalize(FinalizeJob J) { spec_finalize_body(J); }
__global__ __launch_bounds__(RES_BLOCK) void k_spec_finalize_m(const FinalizeJob *J, const uint32_t *map) {
    int lb;
    spec_finalize_body(J[multi_job(map, lb)]);
}


// C = A B (row-major, A: M x K, B: K x N) in FP64 on the matrix cores
// (dgemm_tile_mfma), one 16x16 output tile per 256-thread block with K split
// over its 4 waves.  MODE 0: store; MODE 1: store C .* scale; MODE 2:
// x64 += C.
struct GemmJob {
    int M, N, K;
    const double *A, *B;
    double *C;
    const double *scale;
    double *x64;
    const SpecState *s;
    int gx;   // blocks along x of the job's own grid
};

template <int MODE>
__device__ __forceinline__ void dgemm_body(const GemmJob &J, int bx, int by) {
    const int M = J.M, N = J.N, K = J.K;
    const double *__restrict__ A = J.A;
    const double *__restrict__ B = J.B;
    double *__restrict__ C = J.C;
    const double *__restrict__ scale = J.scale;
    double *__restrict__ x64 = J.x64;
    const SpecState *s = J.s;

    if (s->done) return;   // uniform over the grid: no thread reaches a barrier alone
    __shared__ double part[4][256];
    const int r0 = by * GT, c0 = bx * GT;
    const v4d acc = dgemm_tile_mfma(M, N, K, A, B, r0, c0);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int v = 0; v < 4; ++v) part[wv][lane * 4 + v] = acc[v];
    __syncthreads();
    // thread t adds the four waves' accumulator slot t (lane t/4, register t%4)
    const int t = threadIdx.x;
    const int l = t >> 2, v = t & 3;
    const double sum = ((part[0][t] + part[1][t]) + part[2][t]) + part[3][t];
    const int r = r0 + 4 * v + (l >> 4), c = c0 + (l & 15);
    if (r >= M || c >= N) return;
    const size_t o = (size_t)r * N + c;
    if (MODE == 0) C[o] = sum;
    else if (MODE == 1) C[o] = sum * scale[o];
    else x64[o] += sum;
}

template <int MODE>
__global__ __launch_bounds__(256) void k_dgemm(GemmJob J) { dgemm_body<MODE>(J, blockIdx.x, blockIdx.y); }
template <int MODE>
__global__ __launch_bounds__(256) void k_dgemm_m(const GemmJob *J, const uint32_t *map) {
    int lb;
    const GemmJob &a = J[multi_job(map, lb)];
    dgemm_body<MODE>(a, lb % a.gx, lb / a.gx);
}


}  // namespace

template <typename RB>
void spectral_solve_refined(const PcgOperator &A, const SpecDev &sp, const RB *b32, double *b, double *x,
                            double tol, int max_cycles, SpecState *state, double *partials, PcgResult *res,
                            hipStream_t st) {
    const int N = A.N;
    const int nblk = (N + RES_BLOCK - 1) / RES_BLOCK;
    const dim3 g((sp.ny + GT - 1) / GT, (sp.nx + GT - 1) / GT);
    const dim3 blk(GT * GT);
    const SpecState *cs = state;
    Recorder *rec = launch_recorder();
    for (int k = 0; k <= max_cycles; ++k) {
        const ResidJob<RB> Jr{A, k == 0 ? b32 : (const RB *)nullptr, b, (const double *)x, sp.r, partials, state};
        const FinalizeJob Jf{(const double *)partials, nblk, tol, k == max_cycles ? 1 : 0, state, res};
        if (rec) {
            rec->add(&launch_m<ResidJob<RB>, k_residual64_m<RB>, RES_BLOCK>, Jr, nblk);
            rec->add(&launch_m<FinalizeJob, k_spec_finalize_m, RES_BLOCK>, Jf, 1);
        } else {
            hipLaunchKernelGGL(k_residual64<RB>, dim3(nblk), dim3(RES_BLOCK), 0, st, Jr);
            hipLaunchKernelGGL(k_spec_finalize, dim3(1), dim3(RES_BLOCK), 0, st, Jf);
        }
        if (k == max_cycles) break;
        // T1 = Vx^-1 R ; T2 = (T1 Vy^-T) .* dinv ; T1 = Vx T2 ; x += T1 Vy^T
        const int gx = (int)g.x;
        const GemmJob J0{sp.nx, sp.ny, sp.nx, sp.Vxi, (const double *)sp.r, sp.T1, nullptr, nullptr, cs, gx};
        const GemmJob J1{sp.nx, sp.ny, sp.ny, (const double *)sp.T1, sp.VyiT, sp.T2, sp.dinv, nullptr, cs, gx};
        const GemmJob J2{sp.nx, sp.ny, sp.nx, sp.Vx, (const double *)sp.T2, sp.T1, nullptr, nullptr, cs, gx};
        const GemmJob J3{sp.nx, sp.ny, sp.ny, (const double *)sp.T1, sp.VyT, nullptr, nullptr, x, cs, gx};
        if (rec) {
            const int nb = (int)(g.x * g.y);
            rec->add(&launch_m<GemmJob, k_dgemm_m<0>, 256>, J0, nb);
            rec->add(&launch_m<GemmJob, k_dgemm_m<1>, 256>, J1, nb);
            rec->add(&launch_m<GemmJob, k_dgemm_m<0>, 256>, J2, nb);
            rec->add(&launch_m<GemmJob, k_dgemm_m<2>, 256>, J3, nb);
        } else {
            hipLaunchKernelGGL(k_dgemm<0>, g, blk, 0, st, J0);
            hipLaunchKernelGGL(k_dgemm<1>, g, blk, 0, st, J1);
            hipLaunchKernelGGL(k_dgemm<0>, g, blk, 0, st, J2);
            hipLaunchKernelGGL(k_dgemm<2>, g, blk, 0, st, J3);
        }
    }
}
template void spectral_solve_refined<float>(const PcgOperator &, const SpecDev &, const float *, double *, double *,
                                            double, int, SpecState *, double *, PcgResult *, hipStream_t);
template void spectral_solve_refined<double>(const PcgOperator &, const SpecDev &, const double *, double *, double *,
                                             double, int, SpecState *, double *, PcgResult *, hipStream_t);

}  // namespace pic
