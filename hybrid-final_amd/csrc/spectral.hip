// Refined field solve for separable operators: FP64 iterative refinement
// around an FP32 fast-Poisson inverse.
//
// When the unknowns form a rectangle with uniform coefficients (every CCP
// deck with electrodes on two opposite walls), the field operator is
// A = cx Tx (x) I + I (x) cy Ty and
//     A^-1 R = Vx [ (Vx^-1 R Vy^-T) ./ (cx lx_a + cy ly_b) ] Vy^T,
// i.e. four small dense GEMMs (nx x nx x ny and nx x ny x ny) and an
// elementwise scale.  Each refinement cycle is
//     r = b - A x   (FP64 stencil, deterministic reduction of ||r||, ||b||)
//     d = A^-1 r    (FP32 GEMMs)
//     x += d        (FP64)
// and stops on ||b - A x|| <= tol ||b||.  An exact FP32 inverse gains ~6
// digits per cycle, so two or three cycles reach 1e-10 from a warm start.
// All kernels read a device-side done flag, so the whole solve is a fixed
// launch sequence (graph-capturable) that idles once converged.
#include <hip/hip_runtime.h>

#include "pic_kernels.h"

namespace pic {

namespace {

constexpr int RES_BLOCK = 256;
constexpr int TILE = 32;
constexpr int BK = 16;

__global__ void k_spec_begin(SpecState *s) {
    if (threadIdx.x == 0) { s->done = 0; s->outer = 0; s->rel = 0.0; s->status = 0; }
}

// r = b - A x in FP64; rf = (float) r; per-block partial sums of r.r and b.b
__global__ __launch_bounds__(RES_BLOCK) void k_residual64(PcgOperator A, const double *b, const double *x, float *rf,
                                                          double *partials, const SpecState *s) {
    if (s->done) return;
    __shared__ double sh[2 * RES_BLOCK];
    const int i = blockIdx.x * RES_BLOCK + threadIdx.x;
    double rr = 0.0, bb = 0.0;
    if (i < A.N) {
        const int o = 5 * i;
        double ax = A.vald[o] * x[A.col[o]];
        ax += A.vald[o + 1] * x[A.col[o + 1]];
        ax += A.vald[o + 2] * x[A.col[o + 2]];
        ax += A.vald[o + 3] * x[A.col[o + 3]];
        ax += A.vald[o + 4] * x[A.col[o + 4]];
        double r = b[i] - ax;
        rf[i] = (float)r;
        rr = r * r;
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

// C = A B (row-major, A: M x K, B: K x N) in FP32 with explicit FMAs.
// MODE 0: store; MODE 1: store C * scale (elementwise); MODE 2: x64 += C.
template <int MODE>
__global__ __launch_bounds__(256) void k_sgemm(int M, int N, int K, const float *A, const float *B, float *C,
                                               const float *scale, double *x64, const SpecState *s) {
    if (s->done) return;
    __shared__ float As[BK][TILE + 1];
    __shared__ float Bs[BK][TILE + 1];
    const int tx = threadIdx.x % 16, ty = threadIdx.x / 16;
    const int row0 = blockIdx.y * TILE, col0 = blockIdx.x * TILE;
    float acc[2][2] = {{0.f, 0.f}, {0.f, 0.f}};
    for (int k0 = 0; k0 < K; k0 += BK) {
        for (int e = threadIdx.x; e < TILE * BK; e += 256) {
            int r = e / BK, kk = e % BK;
            int gr = row0 + r, gk = k0 + kk;
            As[kk][r] = (gr < M && gk < K) ? A[(size_t)gr * K + gk] : 0.f;
            int kb = e / TILE, c = e % TILE;
            int gkb = k0 + kb, gc = col0 + c;
            Bs[kb][c] = (gkb < K && gc < N) ? B[(size_t)gkb * N + gc] : 0.f;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < BK; ++kk) {
            float a0 = As[kk][ty * 2], a1 = As[kk][ty * 2 + 1];
            float b0 = Bs[kk][tx * 2], b1 = Bs[kk][tx * 2 + 1];
            acc[0][0] = __builtin_fmaf(a0, b0, acc[0][0]);
            acc[0][1] = __builtin_fmaf(a0, b1, acc[0][1]);
            acc[1][0] = __builtin_fmaf(a1, b0, acc[1][0]);
            acc[1][1] = __builtin_fmaf(a1, b1, acc[1][1]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            int r = row0 + ty * 2 + i, c = col0 + tx * 2 + j;
            if (r < M && c < N) {
                size_t o = (size_t)r * N + c;
                if (MODE == 0) C[o] = acc[i][j];
                else if (MODE == 1) C[o] = acc[i][j] * scale[o];
                else x64[o] += (double)acc[i][j];
            }
        }
}

}  // namespace

void spectral_solve_refined(const PcgOperator &A, const SpecDev &sp, const double *b, double *x, double tol,
                            int max_cycles, SpecState *state, double *partials, float *rf, PcgResult *res,
                            hipStream_t st) {
    const int N = A.N;
    const int nblk = (N + RES_BLOCK - 1) / RES_BLOCK;
    const dim3 g1((sp.ny + TILE - 1) / TILE, (sp.nx + TILE - 1) / TILE);
    hipLaunchKernelGGL(k_spec_begin, dim3(1), dim3(64), 0, st, state);
    for (int k = 0; k <= max_cycles; ++k) {
        hipLaunchKernelGGL(k_residual64, dim3(nblk), dim3(RES_BLOCK), 0, st, A, b, (const double *)x, rf, partials,
                           (const SpecState *)state);
        hipLaunchKernelGGL(k_spec_finalize, dim3(1), dim3(RES_BLOCK), 0, st, (const double *)partials, nblk, tol,
                           k == max_cycles ? 1 : 0, state, res);
        if (k == max_cycles) break;
        // T1 = Vx^-1 R ; T2 = (T1 Vy^-T) .* dinv ; T1 = Vx T2 ; x += T1 Vy^T
        hipLaunchKernelGGL(k_sgemm<0>, g1, dim3(256), 0, st, sp.nx, sp.ny, sp.nx, sp.Vxi, (const float *)rf, sp.T1,
                           (const float *)nullptr, (double *)nullptr, (const SpecState *)state);
        hipLaunchKernelGGL(k_sgemm<1>, g1, dim3(256), 0, st, sp.nx, sp.ny, sp.ny, (const float *)sp.T1, sp.VyiT, sp.T2,
                           sp.dinv, (double *)nullptr, (const SpecState *)state);
        hipLaunchKernelGGL(k_sgemm<0>, g1, dim3(256), 0, st, sp.nx, sp.ny, sp.nx, sp.Vx, (const float *)sp.T2, sp.T1,
                           (const float *)nullptr, (double *)nullptr, (const SpecState *)state);
        hipLaunchKernelGGL(k_sgemm<2>, g1, dim3(256), 0, st, sp.nx, sp.ny, sp.ny, (const float *)sp.T1, sp.VyT,
                           (float *)nullptr, (const float *)nullptr, x, (const SpecState *)state);
    }
}

}  // namespace pic
