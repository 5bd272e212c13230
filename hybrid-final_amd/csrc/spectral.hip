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
// memory latency, not by FLOPs: each GEMM block stages its operand panels
// through LDS in one round trip and computes one 16x16 output tile, so a
// 256^2 problem fills the chip with ~256 blocks.
#include <hip/hip_runtime.h>

#include "pic_kernels.h"

namespace pic {

namespace {

constexpr int RES_BLOCK = 256;
constexpr int GT = 16;   // GEMM output tile edge

// r = b - A x in FP64; per-block partial sums of r.r and b.b.  The first
// residual of a solve (b32 != null) also starts the solve: it converts the
// FP32 right-hand side into b and resets the solve state.
__global__ __launch_bounds__(RES_BLOCK) void k_residual64(PcgOperator A, const float *b32, double *b, const double *x,
                                                          double *r, double *partials, SpecState *s) {
    if (b32) {
        if (blockIdx.x == 0 && threadIdx.x == 0) { s->done = 0; s->outer = 0; s->rel = 0.0; s->status = 0; }
    } else if (s->done) {
        return;
    }
    __shared__ double sh[2 * RES_BLOCK];
    const int i = blockIdx.x * RES_BLOCK + threadIdx.x;
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
        partials[2 * blockIdx.x] = sh[0];
        partials[2 * blockIdx.x + 1] = sh[RES_BLOCK];
    }
}

// one block: fixed-order sum of the partials, convergence decision
__global__ __launch_bounds__(RES_BLOCK) void k_spec_finalize(const double *partials, int nblk, double tol, int last,
                                                             SpecState *s, PcgResult *res) {
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

// C = A B (row-major, A: M x K, B: K x N) in FP64, one 16x16 output tile per
// 256-thread block, one output per thread.  The block's A rows and B columns
// are staged through LDS in K-panels of GK: every thread issues all its
// panel loads at once (one memory round trip per panel), then accumulates
// from LDS (A broadcast along a tile row, B conflict-free along a row).
// MODE 0: store; MODE 1: store C .* scale; MODE 2: x64 += C.
constexpr int GK = 272;   // GT * GK is a multiple of the block size
template <int MODE>
__global__ __launch_bounds__(GT *GT) void k_dgemm(int M, int N, int K, const double *__restrict__ A,
                                                 const double *__restrict__ B, double *__restrict__ C,
                                                 const double *__restrict__ scale, double *__restrict__ x64,
                                                 const SpecState *s) {
    if (s->done) return;   // uniform over the grid: no thread reaches a barrier alone
    __shared__ double As[GT][GK + 1];
    __shared__ double Bs[GK][GT];
    const int tx = threadIdx.x % GT, ty = threadIdx.x / GT;
    const int r0 = blockIdx.y * GT, c0 = blockIdx.x * GT;
    double acc0 = 0.0, acc1 = 0.0;
    for (int kb = 0; kb < K; kb += GK) {
        const int kn = min(GK, K - kb);
        // A panel: GT rows x kn, B panel: kn rows x GT (zero outside the
        // matrix); fixed trip counts so that all loads issue before the
        // first LDS store
        static_assert((GT * GK) % (GT * GT) == 0, "panel size");
        constexpr int PER = (GT * GK) / (GT * GT);
        double ta[PER], tb[PER];
#pragma unroll
        for (int m = 0; m < PER; ++m) {
            const int e = (int)threadIdx.x + m * GT * GT;
            const int rr = e / GK, kk = e % GK, gr = r0 + rr;
            ta[m] = (kk < kn && gr < M) ? A[(size_t)gr * K + kb + kk] : 0.0;
            const int kb2 = e / GT, cc = e % GT, gc = c0 + cc;
            tb[m] = (kb2 < kn && gc < N) ? B[(size_t)(kb + kb2) * N + gc] : 0.0;
        }
#pragma unroll
        for (int m = 0; m < PER; ++m) {
            const int e = (int)threadIdx.x + m * GT * GT;
            As[e / GK][e % GK] = ta[m];
            Bs[e / GT][e % GT] = tb[m];
        }
        __syncthreads();
        int kk = 0;
        for (; kk + 2 <= kn; kk += 2) {
            acc0 = __builtin_fma(As[ty][kk], Bs[kk][tx], acc0);
            acc1 = __builtin_fma(As[ty][kk + 1], Bs[kk + 1][tx], acc1);
        }
        if (kk < kn) acc0 = __builtin_fma(As[ty][kk], Bs[kk][tx], acc0);
        __syncthreads();
    }
    const int r = r0 + ty, c = c0 + tx;
    if (r >= M || c >= N) return;
    const double v = acc0 + acc1;
    const size_t o = (size_t)r * N + c;
    if (MODE == 0) C[o] = v;
    else if (MODE == 1) C[o] = v * scale[o];
    else x64[o] += v;
}

}  // namespace

void spectral_solve_refined(const PcgOperator &A, const SpecDev &sp, const float *b32, double *b, double *x,
                            double tol, int max_cycles, SpecState *state, double *partials, PcgResult *res,
                            hipStream_t st) {
    const int N = A.N;
    const int nblk = (N + RES_BLOCK - 1) / RES_BLOCK;
    const dim3 g((sp.ny + GT - 1) / GT, (sp.nx + GT - 1) / GT);
    const dim3 blk(GT * GT);
    const SpecState *cs = state;
    for (int k = 0; k <= max_cycles; ++k) {
        hipLaunchKernelGGL(k_residual64, dim3(nblk), dim3(RES_BLOCK), 0, st, A, k == 0 ? b32 : (const float *)nullptr,
                           b, (const double *)x, sp.r, partials, state);
        hipLaunchKernelGGL(k_spec_finalize, dim3(1), dim3(RES_BLOCK), 0, st, (const double *)partials, nblk, tol,
                           k == max_cycles ? 1 : 0, state, res);
        if (k == max_cycles) break;
        // T1 = Vx^-1 R ; T2 = (T1 Vy^-T) .* dinv ; T1 = Vx T2 ; x += T1 Vy^T
        hipLaunchKernelGGL(k_dgemm<0>, g, blk, 0, st, sp.nx, sp.ny, sp.nx, sp.Vxi, (const double *)sp.r, sp.T1,
                           (const double *)nullptr, (double *)nullptr, cs);
        hipLaunchKernelGGL(k_dgemm<1>, g, blk, 0, st, sp.nx, sp.ny, sp.ny, (const double *)sp.T1, sp.VyiT, sp.T2,
                           sp.dinv, (double *)nullptr, cs);
        hipLaunchKernelGGL(k_dgemm<0>, g, blk, 0, st, sp.nx, sp.ny, sp.nx, sp.Vx, (const double *)sp.T2, sp.T1,
                           (const double *)nullptr, (double *)nullptr, cs);
        hipLaunchKernelGGL(k_dgemm<2>, g, blk, 0, st, sp.nx, sp.ny, sp.ny, (const double *)sp.T1, sp.VyT,
                           (double *)nullptr, (const double *)nullptr, x, cs);
    }
}

}  // namespace pic
