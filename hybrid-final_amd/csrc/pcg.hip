// Field solve: Jacobi-preconditioned conjugate gradients as one persistent
// kernel per solve (the reference's PCG_float,
// src/cuda_Field.cu:942-1072, uses the same single-launch structure).
//
//  * plain FP64 path: vectors, operator and dot products in FP64;
//  * refined path: FP64 outer refinement (residual b - A x recomputed in FP64,
//    x accumulated in FP64) around an FP32 inner PCG whose vectors and
//    stencil sweeps are FP32 and whose global dot products are FP64.
//
// Both stop on the FP64 relative residual ||b - A x||_2 <= tol ||b||_2.
// Global reductions are deterministic: each block writes its partial sum,
// and after the grid barrier every block adds all partials in the same
// order.  The recurrence q = A p is carried as q <- A z + beta q so each
// iteration needs two grid barriers.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <map>
#include <mutex>
#include <vector>

#include "pic_kernels.h"

namespace pic {

constexpr int PCG_BLOCK = 256;      // default threads per block
constexpr int PCG_MAX_BLOCK = 512;  // largest block the tuner tries (and the largest grid: one partial per thread)
constexpr int PCG_NRED = 4;         // values reduced per barrier

struct GridSync {
    unsigned int count;
    unsigned int gen;
    unsigned int abort;   // set when a barrier wait ran out of time; every block then leaves
    unsigned int pad;
};

struct PcgWork {
    int N = 0, nblk = 0;
    int bs = PCG_BLOCK;   // threads per block
    int ncu = 0;
    double *r = nullptr, *z = nullptr, *p = nullptr, *q = nullptr;   // FP64 vectors (plain path / outer residual)
    float *rf = nullptr, *zf = nullptr, *pf = nullptr, *qf = nullptr, *df = nullptr;
    double *partials = nullptr;   // 2 sets x PCG_MAX_BLOCK x PCG_NRED
    GridSync *gs = nullptr;
    unsigned long long spin_limit = 0;   // barrier wait limit, in ticks of the 100 MHz constant clock
};

namespace {

// Grid barrier of the persistent solver.  The solver is launched as an
// ordinary kernel with no more blocks than the device holds at once, so every
// block becomes resident once the work ahead of it drains; a wait is still
// bounded: after spin_limit ticks of the 100 MHz constant clock the waiting
// block raises gs->abort, and every block that sees it leaves the solve
// (status 3).  Returns false when the solve is being abandoned.
__device__ inline bool grid_sync(GridSync *gs, unsigned nblk, unsigned long long spin_limit) {
    __shared__ int ok;
    // the block's stores happen-before thread 0's agent-scope release via the
    // workgroup barrier; thread 0's acquire makes every block's stores
    // visible to the whole block after the second barrier
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
        int good = __hip_atomic_load(&gs->abort, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == 0u;
        if (good) {
            unsigned g = __hip_atomic_load(&gs->gen, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
            unsigned arrived = __hip_atomic_fetch_add(&gs->count, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
            if (arrived == nblk - 1) {
                __hip_atomic_store(&gs->count, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_fetch_add(&gs->gen, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            } else {
                const unsigned long long t0 = __builtin_amdgcn_s_memrealtime();
                while (__hip_atomic_load(&gs->gen, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == g) {
                    if (__hip_atomic_load(&gs->abort, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) != 0u) {
                        good = 0;
                        break;
                    }
                    if (__builtin_amdgcn_s_memrealtime() - t0 > spin_limit) {
                        __hip_atomic_store(&gs->abort, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
                        good = 0;
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                }
            }
        }
        __threadfence();
        ok = good;
    }
    __syncthreads();
    const bool r = ok != 0;
    __syncthreads();   // ok is rewritten by the next barrier
    return r;
}

// block-level sum of NV doubles with a fixed tree; result valid in all threads
template <int NV, int BS>
__device__ inline void block_sum(double (&v)[NV], double *sh) {
    const int t = threadIdx.x;
#pragma unroll
    for (int j = 0; j < NV; ++j) sh[j * BS + t] = v[j];
    __syncthreads();
    for (int off = BS / 2; off > 0; off >>= 1) {
        if (t < off) {
#pragma unroll
            for (int j = 0; j < NV; ++j) sh[j * BS + t] += sh[j * BS + t + off];
        }
        __syncthreads();
    }
#pragma unroll
    for (int j = 0; j < NV; ++j) v[j] = sh[j * BS];
    __syncthreads();
}

// publish block partials, barrier, and sum all partials in block order
template <int NV, int BS>
__device__ inline bool grid_sum(double (&v)[NV], double *sh, double *partials, int &phase, GridSync *gs, unsigned nblk,
                                unsigned long long spin_limit) {
    block_sum<NV, BS>(v, sh);
    double *set = partials + (size_t)(phase & 1) * nblk * PCG_NRED;
    if (threadIdx.x == 0) {
#pragma unroll
        for (int j = 0; j < NV; ++j) set[blockIdx.x * PCG_NRED + j] = v[j];
    }
    if (!grid_sync(gs, nblk, spin_limit)) return false;
    double acc[NV];
#pragma unroll
    for (int j = 0; j < NV; ++j) acc[j] = 0.0;
    for (int b = threadIdx.x; b < (int)nblk; b += BS) {
#pragma unroll
        for (int j = 0; j < NV; ++j) acc[j] = set[b * PCG_NRED + j];
    }
    // nblk <= BS: thread b holds partial b; fixed-tree block sum
    block_sum<NV, BS>(acc, sh);
#pragma unroll
    for (int j = 0; j < NV; ++j) v[j] = acc[j];
    ++phase;
    return true;
}

template <typename V>
__device__ inline V spmv_row(const int *col, const V *val, const V *x, int i) {
    const int o = 5 * i;
    V s = val[o] * x[col[o]];
    s += val[o + 1] * x[col[o + 1]];
    s += val[o + 2] * x[col[o + 2]];
    s += val[o + 3] * x[col[o + 3]];
    s += val[o + 4] * x[col[o + 4]];
    return s;
}

// a barrier wait ran out of time: the solve is abandoned (status 3)
__device__ inline void pcg_abandon(PcgResult *res) {
    if (blockIdx.x == 0 && threadIdx.x == 0) res->status = 3;
}
#define PCG_SUM(NV, v)                                                                                               \
    do {                                                                                                             \
        if (!grid_sum<NV, BS>(v, sh, partials, phase, gs, nblk, spin_limit)) {                                       \
            pcg_abandon(res);                                                                                        \
            return;                                                                                                  \
        }                                                                                                            \
    } while (0)

// ------------------------------------------------------------ plain FP64
template <int BS>
__global__ __launch_bounds__(BS) void k_pcg_plain(PcgOperator A, const double *b, double *x, double *r,
                                                        double *z, double *p, double *q, double *partials,
                                                        GridSync *gs, double tol, int maxit, PcgResult *res,
                                                        unsigned long long spin_limit) {
    __shared__ double sh[PCG_NRED * BS];
    const unsigned nblk = gridDim.x;
    const int N = A.N;
    const int stride = nblk * BS;
    const int t0 = blockIdx.x * BS + threadIdx.x;
    int phase = 0;
    double v3[3] = {0, 0, 0};
    for (int i = t0; i < N; i += stride) {
        double ri = b[i] - spmv_row<double>(A.col, A.vald, x, i);
        double zi = A.mdiagd[i] * ri;
        r[i] = ri; z[i] = zi; p[i] = zi;
        v3[0] += ri * zi; v3[1] += ri * ri; v3[2] += b[i] * b[i];
    }
    PCG_SUM(3, v3);
    double rz = v3[0], rr = v3[1], bb = v3[2];
    const double tol2 = tol * tol * bb;
    int it = 0;
    int status = 0;
    if (bb == 0.0) {
        for (int i = t0; i < N; i += stride) x[i] = 0.0;
        rr = 0.0;
    } else if (rr > tol2) {
        double v1[1] = {0};
        for (int i = t0; i < N; i += stride) {
            double qi = spmv_row<double>(A.col, A.vald, p, i);
            q[i] = qi;
            v1[0] += p[i] * qi;
        }
        PCG_SUM(1, v1);
        double pq = v1[0];
        for (;;) {
            double alpha = pq != 0.0 ? rz / pq : 0.0;
            double v2[2] = {0, 0};
            for (int i = t0; i < N; i += stride) {
                x[i] += alpha * p[i];
                double ri = r[i] - alpha * q[i];
                double zi = A.mdiagd[i] * ri;
                r[i] = ri; z[i] = zi;
                v2[0] += ri * zi; v2[1] += ri * ri;
            }
            PCG_SUM(2, v2);
            ++it;
            double rz_new = v2[0];
            rr = v2[1];
            if (rr <= tol2) break;
            if (it >= maxit) { status = 1; break; }
            double beta = rz != 0.0 ? rz_new / rz : 0.0;
            rz = rz_new;
            v1[0] = 0.0;
            for (int i = t0; i < N; i += stride) {
                double pi = z[i] + beta * p[i];
                double qi = spmv_row<double>(A.col, A.vald, z, i) + beta * q[i];
                p[i] = pi; q[i] = qi;
                v1[0] += pi * qi;
            }
            PCG_SUM(1, v1);
            pq = v1[0];
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        res->iters = it;
        res->outer = 1;
        res->rel_res = bb > 0.0 ? sqrt(rr / bb) : 0.0;
        res->status = status;
    }
}

// ------------------------------------------------------------ refined
template <int BS>
__global__ __launch_bounds__(BS) void k_pcg_refined(PcgOperator A, const double *b, double *x, double *r64,
                                                          float *r, float *z, float *p, float *q, float *d,
                                                          double *partials, GridSync *gs, double tol, double inner_tol,
                                                          int maxit, PcgResult *res, unsigned long long spin_limit) {
    __shared__ double sh[PCG_NRED * BS];
    const unsigned nblk = gridDim.x;
    const int N = A.N;
    const int stride = nblk * BS;
    const int t0 = blockIdx.x * BS + threadIdx.x;
    int phase = 0;
    int it = 0, outer = 0, status = 0;
    double bb = -1.0, rr64 = 0.0;
    for (;;) {
        // FP64 residual of the current iterate
        double v2[2] = {0, 0};
        for (int i = t0; i < N; i += stride) {
            double ri = b[i] - spmv_row<double>(A.col, A.vald, x, i);
            r64[i] = ri;
            v2[0] += ri * ri;
            if (bb < 0.0) v2[1] += b[i] * b[i];
        }
        PCG_SUM(2, v2);
        rr64 = v2[0];
        if (bb < 0.0) bb = v2[1];
        if (bb == 0.0) {
            for (int i = t0; i < N; i += stride) x[i] = 0.0;
            rr64 = 0.0;
            break;
        }
        if (rr64 <= tol * tol * bb) break;
        if (it >= maxit) { status = 1; break; }
        ++outer;
        // inner FP32 PCG on A d = r, d0 = 0
        double v3[2] = {0, 0};
        for (int i = t0; i < N; i += stride) {
            float ri = (float)r64[i];
            float zi = A.mdiagf[i] * ri;
            r[i] = ri; z[i] = zi; p[i] = zi; d[i] = 0.0f;
            v3[0] += (double)ri * (double)zi;
            v3[1] += (double)ri * (double)ri;
        }
        PCG_SUM(2, v3);
        double rz = v3[0];
        const double stop2 = inner_tol * inner_tol * v3[1];
        double v1[1] = {0};
        for (int i = t0; i < N; i += stride) {
            float qi = spmv_row<float>(A.col, A.valf, p, i);
            q[i] = qi;
            v1[0] += (double)p[i] * (double)qi;
        }
        PCG_SUM(1, v1);
        double pq = v1[0];
        for (;;) {
            float alpha = (float)(pq != 0.0 ? rz / pq : 0.0);
            double w2[2] = {0, 0};
            for (int i = t0; i < N; i += stride) {
                d[i] += alpha * p[i];
                float ri = r[i] - alpha * q[i];
                float zi = A.mdiagf[i] * ri;
                r[i] = ri; z[i] = zi;
                w2[0] += (double)ri * (double)zi;
                w2[1] += (double)ri * (double)ri;
            }
            PCG_SUM(2, w2);
            ++it;
            double rz_new = w2[0];
            if (w2[1] <= stop2 || it >= maxit) break;
            float beta = (float)(rz != 0.0 ? rz_new / rz : 0.0);
            rz = rz_new;
            v1[0] = 0.0;
            for (int i = t0; i < N; i += stride) {
                float pi = z[i] + beta * p[i];
                float qi = spmv_row<float>(A.col, A.valf, z, i) + beta * q[i];
                p[i] = pi; q[i] = qi;
                v1[0] += (double)pi * (double)qi;
            }
            PCG_SUM(1, v1);
            pq = v1[0];
        }
        for (int i = t0; i < N; i += stride) x[i] += (double)d[i];
        // x is read by other blocks in the residual sweep
        if (!grid_sync(gs, nblk, spin_limit)) {
            pcg_abandon(res);
            return;
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        res->iters = it;
        res->outer = outer;
        res->rel_res = bb > 0.0 ? sqrt(rr64 / bb) : 0.0;
        res->status = status;
    }
}

#undef PCG_SUM

__global__ void k_pcg_reset(GridSync *gs) {
    if (threadIdx.x == 0) { gs->count = 0; gs->gen = 0; gs->abort = 0; gs->pad = 0; }
}

}  // namespace

// blocks of BS threads of the plain and refined kernels one CU holds at once
static int pcg_resident_per_cu(int bs) {
    int a = 0, b = 0;
    hipError_t e1, e2;
    switch (bs) {
    case 128:
        e1 = hipOccupancyMaxActiveBlocksPerMultiprocessor(&a, reinterpret_cast<const void *>(k_pcg_refined<128>), 128, 0);
        e2 = hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, reinterpret_cast<const void *>(k_pcg_plain<128>), 128, 0);
        break;
    case 512:
        e1 = hipOccupancyMaxActiveBlocksPerMultiprocessor(&a, reinterpret_cast<const void *>(k_pcg_refined<512>), 512, 0);
        e2 = hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, reinterpret_cast<const void *>(k_pcg_plain<512>), 512, 0);
        break;
    default:
        e1 = hipOccupancyMaxActiveBlocksPerMultiprocessor(&a, reinterpret_cast<const void *>(k_pcg_refined<256>), 256, 0);
        e2 = hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, reinterpret_cast<const void *>(k_pcg_plain<256>), 256, 0);
        break;
    }
    if (e1 != hipSuccess || e2 != hipSuccess) return -1;
    return a < b ? a : b;
}

PcgWork *pcg_work_create(int N, int *err, int want_blocks, hipStream_t stream) {
    PcgWork *w = new PcgWork();
    w->N = N;
    int dev = 0, ncu = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) {
        if (err) *err = -40;
        delete w;
        return nullptr;
    }
    const int per_cu = pcg_resident_per_cu(PCG_BLOCK);
    if (per_cu < 0) {
        if (err) *err = -40;
        delete w;
        return nullptr;
    }
    w->ncu = ncu;
    int resident = ncu * (per_cu > 0 ? per_cu : 1);
    int need = (N + PCG_BLOCK - 1) / PCG_BLOCK;
    // Barrier latency grows with the number of participating blocks while
    // the per-iteration work of a <= 512^2 problem is tiny: a quarter of the
    // CUs (64 on MI355X) measured fastest for 65k unknowns (31 ms vs 110 ms
    // per 1400-iteration solve with one block per CU).  pcg_autotune refines
    // the choice per operator size.
    int nblk = need < ncu / 4 ? need : ncu / 4;
    if (nblk < 1) nblk = 1;
    if (want_blocks > 0) nblk = want_blocks < need ? want_blocks : need;
    if (nblk > resident) nblk = resident;
    if (nblk > PCG_BLOCK) nblk = PCG_BLOCK;      // grid_sum keeps one partial per thread
    if (nblk < 1) nblk = 1;
    w->nblk = nblk;
    w->bs = PCG_BLOCK;
    // a barrier wait normally takes microseconds; 2 s means the grid never
    // became resident together (PIC_PCG_SPIN_US overrides, for tests of the
    // abandon path)
    w->spin_limit = 200000000ull;
    if (const char *sl = getenv("PIC_PCG_SPIN_US")) w->spin_limit = 100ull * strtoull(sl, nullptr, 10);
    size_t n = (size_t)N;
    hipError_t e = hipSuccess;
    e = (hipError_t)(e | hipMalloc(&w->r, n * sizeof(double)));
    e = (hipError_t)(e | hipMalloc(&w->z, n * sizeof(double)));
    e = (hipError_t)(e | hipMalloc(&w->p, n * sizeof(double)));
    e = (hipError_t)(e | hipMalloc(&w->q, n * sizeof(double)));
    e = (hipError_t)(e | hipMalloc(&w->rf, n * sizeof(float)));
    e = (hipError_t)(e | hipMalloc(&w->zf, n * sizeof(float)));
    e = (hipError_t)(e | hipMalloc(&w->pf, n * sizeof(float)));
    e = (hipError_t)(e | hipMalloc(&w->qf, n * sizeof(float)));
    e = (hipError_t)(e | hipMalloc(&w->df, n * sizeof(float)));
    e = (hipError_t)(e | hipMalloc(&w->partials, 2 * (size_t)PCG_MAX_BLOCK * PCG_NRED * sizeof(double)));
    e = (hipError_t)(e | hipMalloc(&w->gs, sizeof(GridSync)));
    if (e != hipSuccess) { *err = -40; pcg_work_destroy(w); return nullptr; }
    // zero the barrier state on the solver's stream and wait: a persistent
    // kernel must never see a stale counter
    if (hipMemsetAsync(w->gs, 0, sizeof(GridSync), stream) != hipSuccess ||
        hipMemsetAsync(w->r, 0, n * sizeof(double), stream) != hipSuccess ||
        hipStreamSynchronize(stream) != hipSuccess) {
        *err = -40;
        pcg_work_destroy(w);
        return nullptr;
    }
    *err = 0;
    return w;
}

void pcg_work_destroy(PcgWork *w) {
    if (!w) return;
    void *bufs[] = {w->r, w->z, w->p, w->q, w->rf, w->zf, w->pf, w->qf, w->df, w->partials, w->gs};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    delete w;
}

int pcg_grid_blocks(PcgWork *w) { return w ? w->nblk : 0; }

hipError_t pcg_work_reset(PcgWork *w, hipStream_t s) {
    hipLaunchKernelGGL(k_pcg_reset, dim3(1), dim3(64), 0, s, w->gs);
    return hipGetLastError();
}

template <int BS>
static void launch_plain(const PcgOperator &A, const double *b, double *x, double tol, int maxit, PcgWork *w,
                         PcgResult *res, hipStream_t s) {
    hipLaunchKernelGGL(k_pcg_plain<BS>, dim3(w->nblk), dim3(BS), 0, s, A, b, x, w->r, w->z, w->p, w->q, w->partials,
                       w->gs, tol, maxit, res, w->spin_limit);
}
template <int BS>
static void launch_refined(const PcgOperator &A, const double *b, double *x, double tol, double inner_tol, int maxit,
                           PcgWork *w, PcgResult *res, hipStream_t s) {
    hipLaunchKernelGGL(k_pcg_refined<BS>, dim3(w->nblk), dim3(BS), 0, s, A, b, x, w->r, w->rf, w->zf, w->pf, w->qf,
                       w->df, w->partials, w->gs, tol, inner_tol, maxit, res, w->spin_limit);
}

hipError_t pcg_solve_plain_fp64(const PcgOperator &A, const double *b, double *x, double tol, int maxit, PcgWork *w,
                                PcgResult *res, hipStream_t s) {
    switch (w->bs) {
    case 128: launch_plain<128>(A, b, x, tol, maxit, w, res, s); break;
    case 512: launch_plain<512>(A, b, x, tol, maxit, w, res, s); break;
    default: launch_plain<256>(A, b, x, tol, maxit, w, res, s); break;
    }
    return hipGetLastError();
}

hipError_t pcg_solve_refined(const PcgOperator &A, const double *b, double *x, double tol, double inner_tol,
                             int maxit, PcgWork *w, PcgResult *res, hipStream_t s) {
    switch (w->bs) {
    case 128: launch_refined<128>(A, b, x, tol, inner_tol, maxit, w, res, s); break;
    case 512: launch_refined<512>(A, b, x, tol, inner_tol, maxit, w, res, s); break;
    default: launch_refined<256>(A, b, x, tol, inner_tol, maxit, w, res, s); break;
    }
    return hipGetLastError();
}

// Autotuner of the persistent solver's launch shape: threads per block
// (128, 256, 512) and blocks in the grid.  Each candidate that fits the
// device at once runs a fixed number of iterations of the real operator
// (right-hand side of ones, tolerance 0 so it never stops early) after one
// untimed run; the fastest per iteration wins.  Results are kept per
// (unknowns, path) for the process, so cases of the same size tune once.
struct TuneKey {
    int N, refined;
    bool operator<(const TuneKey &o) const { return N < o.N || (N == o.N && refined < o.refined); }
};
static std::mutex g_tune_mu;
static std::map<TuneKey, std::pair<int, int>> g_tune;

int pcg_autotune(PcgWork *w, const PcgOperator &A, int refined, hipStream_t st, double *best_us_per_iter) {
    const TuneKey key{w->N, refined ? 1 : 0};
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto it = g_tune.find(key);
        if (it != g_tune.end()) {
            w->bs = it->second.first;
            w->nblk = it->second.second;
            if (best_us_per_iter) *best_us_per_iter = 0.0;
            return 0;
        }
    }
    const int N = w->N;
    double *b = nullptr, *x = nullptr;
    PcgResult *res = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = 0;
    if (hipMalloc(&b, (size_t)N * sizeof(double)) != hipSuccess || hipMalloc(&x, (size_t)N * sizeof(double)) != hipSuccess ||
        hipMalloc(&res, sizeof(PcgResult)) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
        hipEventCreate(&e1) != hipSuccess) {
        rc = -40;
    }
    const int iters = 40;
    int best_bs = w->bs, best_nb = w->nblk;
    float best = 1e30f;
    if (rc == 0) {
        std::vector<double> ones((size_t)N, 1.0);
        if (hipMemcpyAsync(b, ones.data(), (size_t)N * sizeof(double), hipMemcpyHostToDevice, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess)
            rc = -40;
    }
    const int sizes[3] = {128, 256, 512};
    const int grids[9] = {4, 8, 16, 32, 48, 64, 96, 128, 256};
    const int bs0 = w->bs, nb0 = w->nblk;
    for (int si = 0; si < 3 && rc == 0; ++si) {
        const int bs = sizes[si];
        const int per_cu = pcg_resident_per_cu(bs);
        if (per_cu <= 0) continue;
        const int resident = per_cu * w->ncu;
        const int need = (N + bs - 1) / bs;
        for (int gi = 0; gi < 9 && rc == 0; ++gi) {
            const int nb = grids[gi];
            if (nb > need || nb > resident || nb > bs) continue;
            w->bs = bs;
            w->nblk = nb;
            float ms = 0.0f;
            for (int rep = 0; rep < 2 && rc == 0; ++rep) {
                if (hipMemsetAsync(x, 0, (size_t)N * sizeof(double), st) != hipSuccess ||
                    hipEventRecord(e0, st) != hipSuccess) { rc = -40; break; }
                const hipError_t le = refined ? pcg_solve_refined(A, b, x, 0.0, 0.0, iters, w, res, st)
                                              : pcg_solve_plain_fp64(A, b, x, 0.0, iters, w, res, st);
                if (le != hipSuccess || hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                    hipEventElapsedTime(&ms, e0, e1) != hipSuccess) { rc = -40; break; }
            }
            PcgResult r{};
            if (rc == 0 && (hipMemcpyAsync(&r, res, sizeof r, hipMemcpyDeviceToHost, st) != hipSuccess ||
                            hipStreamSynchronize(st) != hipSuccess)) rc = -40;
            if (rc == 0 && r.status != 3 && ms < best) {   // an abandoned solve is no candidate
                best = ms;
                best_bs = bs;
                best_nb = nb;
            }
            if (rc == 0 && r.status == 3 && pcg_work_reset(w, st) != hipSuccess) rc = -40;
        }
    }
    if (rc == 0 && best < 1e29f) {
        w->bs = best_bs;
        w->nblk = best_nb;
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_tune[key] = {best_bs, best_nb};
        if (best_us_per_iter) *best_us_per_iter = 1e3 * (double)best / iters;
    } else {
        w->bs = bs0;
        w->nblk = nb0;
    }
    if (b) (void)hipFree(b);
    if (x) (void)hipFree(x);
    if (res) (void)hipFree(res);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    return rc;
}

int pcg_launch_shape(const PcgWork *w, int *blocks, int *threads) {
    if (!w) return -1;
    if (blocks) *blocks = w->nblk;
    if (threads) *threads = w->bs;
    return 0;
}

}  // namespace pic
