// FP64 GEMM tiles on the CDNA matrix cores (v_mfma_f64_16x16x4f64).
//
// One 16x16 tile of C = A B (row-major, A: M x K, B: K x N) per 256-thread
// block: wave w of the block accumulates the products over its quarter of K
// (a multiple of 4), 16 MFMA steps of K=4 per round with all of the round's
// operand loads issued first.  The operands of the solvers that use this
// (a few hundred rows, some MB) stay in L2, so each lane loads its
// elements straight from there: A[r0 + l%16][k + l/16] and
// B[k + l/16][c0 + l%16] for lane l.  Register v (0..3) of lane l's
// accumulator holds C[r0 + 4*v + l/16][c0 + l%16] (measured on gfx950 with a
// known-answer tile).  Elements outside the matrix
// load as 0, so partial tiles and any K are exact.  The caller adds the four
// waves' accumulators in a fixed order, so results do not depend on timing.
#pragma once
#include <hip/hip_runtime.h>

namespace pic {

typedef double v4d __attribute__((ext_vector_type(4)));

__device__ inline v4d dgemm_tile_mfma(int M, int N, int K, const double *__restrict__ A,
                                      const double *__restrict__ B, int r0, int c0) {
    constexpr int U = 16;   // MFMA steps per round
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int kq = lane >> 4;
    const int ar = r0 + (lane & 15), bc = c0 + (lane & 15);
    const bool arow = ar < M, bcol = bc < N;
    const double *__restrict__ Ap = A + (size_t)(arow ? ar : 0) * K;
    // quarter of K for this wave, in steps of 4
    const int steps = (K + 3) / 4;
    const int per = (steps + 3) / 4;
    const int kbeg = 4 * min(steps, wv * per), kend = 4 * min(steps, (wv + 1) * per);
    v4d acc = {0.0, 0.0, 0.0, 0.0};
    for (int k0 = kbeg; k0 < kend; k0 += 4 * U) {
        double a[U], b[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int k = k0 + 4 * u + kq;
            const bool in = k < kend && k < K;
            a[u] = (arow && in) ? Ap[k] : 0.0;
            b[u] = (bcol && in) ? B[(size_t)k * N + bc] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a[u], b[u], acc, 0, 0, 0);
    }
    return acc;
}

}  // namespace pic
