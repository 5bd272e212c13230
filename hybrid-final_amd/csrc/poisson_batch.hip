// Batched 5-point Poisson solves on one GPU (many independent problems of
// different sizes in one launch sequence).
//
// Problem p has nx x ny unknowns on a unit-spacing grid with each side
// Neumann (mirror) or Dirichlet (0), operator A = Tx (x) I + I (x) Ty with
// Tx, Ty the 1-D second-difference matrices (interior rows [-1 2 -1], a
// Neumann end row [2 -2], a Dirichlet end row [2 -1]); unknown (a, b) is at
// a * ny + b of the problem's block.  Every such operator is separable, so
// each problem is solved with the exact FP64 fast-Poisson inverse
//     x = Vx [ (Vx^-1 R Vy^-T) ./ (lx_a + ly_b) ] Vy^T
// (four GEMMs per problem, all problems batched into one launch each,
// closed-form eigenbases), wrapped in FP64 iterative refinement that stops
// on ||b - A x|| <= tol ||b|| per problem.  The eigenbasis inverses come from
// weighted orthogonality (T = D^-1 S with S symmetric, D = 1/2 on Neumann end
// rows), so no matrix is inverted numerically.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <mutex>
#include <cstdio>
#include <cmath>
#include <cstring>
#include <map>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/picmcc.h"
#include "dgemm_mfma.h"

namespace pic {
namespace {

constexpr int GT = 16;    // GEMM output tile edge
constexpr int RB = 256;   // residual block

struct GemmTask {          // one 16x16 output tile of one problem's GEMM
    int prob, r0, c0, pad;
};
struct ProbDev {           // device view of one problem
    int nx, ny;
    long long off;         // offset of its unknowns in the flat vectors
    int bc[4];             // left(a=0), right(a=nx-1), bottom(b=0), top(b=ny-1): 0 Neumann, 1 Dirichlet
    const double *Vx, *Vxi, *VyT, *VyiT, *dinv;
};

// A x for unknown (a, b) of problem q
__device__ inline double apply_A(const ProbDev &q, const double *x, int a, int b) {
    const double *u = x + q.off;
    const int nx = q.nx, ny = q.ny;
    const double c = u[(size_t)a * ny + b];
    double ax = 4.0 * c;
    // x direction
    if (a > 0 && a < nx - 1) ax -= u[(size_t)(a - 1) * ny + b] + u[(size_t)(a + 1) * ny + b];
    else if (nx == 1) ax -= (q.bc[0] == 0 ? c : 0.0) + (q.bc[1] == 0 ? c : 0.0);
    else if (a == 0) ax -= (q.bc[0] == 0 ? 2.0 : 1.0) * u[(size_t)1 * ny + b];
    else ax -= (q.bc[1] == 0 ? 2.0 : 1.0) * u[(size_t)(nx - 2) * ny + b];
    // y direction
    if (b > 0 && b < ny - 1) ax -= u[(size_t)a * ny + b - 1] + u[(size_t)a * ny + b + 1];
    else if (ny == 1) ax -= (q.bc[2] == 0 ? c : 0.0) + (q.bc[3] == 0 ? c : 0.0);
    else if (b == 0) ax -= (q.bc[2] == 0 ? 2.0 : 1.0) * u[(size_t)a * ny + 1];
    else ax -= (q.bc[3] == 0 ? 2.0 : 1.0) * u[(size_t)a * ny + ny - 2];
    return ax;
}

// r = b - A x per block of one problem; per-block partial sums of r.r, b.b
__global__ __launch_bounds__(RB) void k_pb_residual(const ProbDev *P, const int2 *blocks, const double *b,
                                                    const double *x, double *r, double *part, const int *done) {
    const int2 bl = blocks[blockIdx.x];   // (problem, first unknown of the block within the problem)
    const ProbDev q = P[bl.x];
    __shared__ double sh[2 * RB];
    double rr = 0.0, bb = 0.0;
    if (!done[bl.x]) {
        const int i = bl.y + (int)threadIdx.x;
        if (i < q.nx * q.ny) {
            const int a = i / q.ny, c = i % q.ny;
            const double bi = b[q.off + i];
            const double ri = bi - apply_A(q, x, a, c);
            r[q.off + i] = ri;
            rr = ri * ri;
            bb = bi * bi;
        }
    }
    sh[threadIdx.x] = rr;
    sh[RB + threadIdx.x] = bb;
    __syncthreads();
    for (int o = RB / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            sh[threadIdx.x] += sh[threadIdx.x + o];
            sh[RB + threadIdx.x] += sh[RB + threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) { part[2 * blockIdx.x] = sh[0]; part[2 * blockIdx.x + 1] = sh[RB]; }
}

// per problem: fixed-order sum of its blocks' partials, convergence decision
__global__ void k_pb_finalize(int nprob, const int *bstart, const double *part, double tol, int last, int *done,
                              double *rel, int *status) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nprob || done[p]) return;
    double rr = 0.0, bb = 0.0;
    for (int k = bstart[p]; k < bstart[p + 1]; ++k) { rr += part[2 * k]; bb += part[2 * k + 1]; }
    const double rl = bb > 0.0 ? sqrt(rr / bb) : sqrt(rr);
    const bool conv = bb > 0.0 ? rr <= tol * tol * bb : rr <= 1e-300;
    rel[p] = rl;
    if (conv || last) { done[p] = 1; status[p] = conv ? 0 : 1; }
}

// batched C = A B per problem, one 16x16 tile per block on the FP64 matrix
// cores (dgemm_tile_mfma, K split over the block's 4 waves); MODE selects
// the operands:
// 0: T1 = Vx^-1 R, 1: T2 = (T1 Vy^-T) .* dinv, 2: T1 = Vx T2, 3: x += T1 Vy^T
template <int MODE>
__global__ __launch_bounds__(GT *GT) void k_pb_gemm(const ProbDev *P, const GemmTask *tasks, const double *R,
                                                    double *T1, double *T2, double *x, const int *done) {
    const GemmTask tk = tasks[blockIdx.x];
    if (done[tk.prob]) return;   // uniform over the block
    const ProbDev q = P[tk.prob];
    const int M = q.nx, N = q.ny;
    const int K = (MODE == 0 || MODE == 2) ? q.nx : q.ny;
    const double *A = MODE == 0 ? q.Vxi : MODE == 1 ? T1 + q.off : MODE == 2 ? q.Vx : T1 + q.off;
    const double *B = MODE == 0 ? R + q.off : MODE == 1 ? q.VyiT : MODE == 2 ? T2 + q.off : q.VyT;
    __shared__ double part[4][256];
    const v4d acc = dgemm_tile_mfma(M, N, K, A, B, tk.r0, tk.c0);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int j = 0; j < 4; ++j) part[wv][lane * 4 + j] = acc[j];
    __syncthreads();
    // thread t adds the four waves' accumulator slot t (lane t/4, register t%4)
    const int t = threadIdx.x, l = t >> 2;
    const double v = ((part[0][t] + part[1][t]) + part[2][t]) + part[3][t];
    const int r = tk.r0 + 4 * (t & 3) + (l >> 4), c = tk.c0 + (l & 15);
    if (r >= M || c >= N) return;
    const size_t o = (size_t)q.off + (size_t)r * N + c;
    if (MODE == 0) T1[o] = v;
    else if (MODE == 1) T2[o] = v * q.dinv[(size_t)r * N + c];
    else if (MODE == 2) T1[o] = v;
    else x[o] += v;
}

__global__ void k_pb_begin(int nprob, int *done, int *status) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < nprob) { done[p] = 0; status[p] = 0; }
}
__global__ void k_pb_zero(long long n, double *x) {
    const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (i < n) x[i] = 0.0;
}

// closed-form eigenbasis of the n x n second difference with the given ends
// (0 Neumann, 1 Dirichlet): V (columns = eigenvectors), lambda, and V^-1
// from the D-weighted orthogonality of the eigenvectors
void eigen_1d(int n, int left, int right, std::vector<double> &V, std::vector<double> &Vi, std::vector<double> &lam) {
    const double PI = 3.14159265358979323846;
    V.assign((size_t)n * n, 0.0);
    Vi.assign((size_t)n * n, 0.0);
    lam.assign(n, 0.0);
    for (int k = 0; k < n; ++k) {
        double th;
        if (left == 0 && right == 0) th = n > 1 ? PI * k / (n - 1) : 0.0;
        else if (left == 1 && right == 1) th = PI * (k + 1) / (n + 1);
        else th = PI * (2 * k + 1) / (2.0 * n);
        for (int a = 0; a < n; ++a) {
            double v;
            if (left == 0 && right == 0) v = std::cos(th * a);
            else if (left == 1 && right == 1) v = std::sin(th * (a + 1));
            else if (left == 0) v = std::cos(th * a);
            else v = std::cos(th * (n - 1 - a));
            V[(size_t)a * n + k] = v;
        }
        lam[k] = 2.0 - 2.0 * std::cos(th);
        if (n == 1) lam[k] = (left == 0 ? 0.0 : 1.0) + (right == 0 ? 0.0 : 1.0);
    }
    // weights: 1/2 on a Neumann end row (n > 1), 1 elsewhere
    std::vector<double> w(n, 1.0);
    if (n > 1) {
        if (left == 0) w[0] = 0.5;
        if (right == 0) w[n - 1] = 0.5;
    }
    for (int k = 0; k < n; ++k) {
        double d = 0.0;
        for (int a = 0; a < n; ++a) d += w[a] * V[(size_t)a * n + k] * V[(size_t)a * n + k];
        for (int a = 0; a < n; ++a) Vi[(size_t)k * n + a] = w[a] * V[(size_t)a * n + k] / d;
    }
}

}  // namespace
}  // namespace pic

using namespace pic;

struct pic_pbatch {
    int nprob = 0;
    long long N = 0;
    int dev = 0;
    hipStream_t st = nullptr;
    std::vector<ProbDev> host;
    std::vector<long long> offs;
    ProbDev *P = nullptr;
    GemmTask *tasks = nullptr;
    int ntask = 0;
    int2 *blocks = nullptr;
    int nblk = 0;
    int *bstart = nullptr;
    double *b = nullptr, *x = nullptr, *r = nullptr, *T1 = nullptr, *T2 = nullptr, *part = nullptr, *rel = nullptr;
    int *done = nullptr, *status = nullptr;
    std::vector<void *> owned;
    std::string err;
};

static thread_local std::string g_pb_err;

#define PBCHK(x)                                                                                                       \
    do {                                                                                                               \
        hipError_t e_ = (x);                                                                                           \
        if (e_ != hipSuccess) {                                                                                        \
            g_pb_err = std::string("HIP error: ") + hipGetErrorString(e_) + " at " #x;                                \
            return -1;                                                                                                 \
        }                                                                                                              \
    } while (0)

template <typename T>
static hipError_t pb_upload(pic_pbatch *B, T **p, const std::vector<T> &v) {
    hipError_t e = hipMalloc((void **)p, std::max<size_t>(1, v.size()) * sizeof(T));
    if (e != hipSuccess) return e;
    B->owned.push_back(*p);
    if (v.empty()) return hipSuccess;
    e = hipMemcpyAsync(*p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, B->st);
    if (e == hipSuccess) e = hipStreamSynchronize(B->st);
    return e;
}
template <typename T>
static hipError_t pb_alloc(pic_pbatch *B, T **p, size_t n) {
    hipError_t e = hipMalloc((void **)p, std::max<size_t>(1, n) * sizeof(T));
    if (e == hipSuccess) B->owned.push_back(*p);
    return e;
}

static int pb_destroy(pic_pbatch *B) {
    if (!B) return 0;
    if (B->st) {
        const hipError_t e = hipStreamSynchronize(B->st);
        if (e != hipSuccess) fprintf(stderr, "picmcc: Poisson batch teardown: %s\n", hipGetErrorString(e));
    }
    for (void *p : B->owned) (void)hipFree(p);
    if (B->st) (void)hipStreamDestroy(B->st);
    delete B;
    return 0;
}

// Host-side tables of a batch: the problems, their GEMM tiles and residual
// blocks, and the matrices each problem reads (eigenbases shared between
// problems with the same size and ends, and each problem's 1/(lx + ly)).
// ProbDev's matrix pointers are filled by bind() once the matrices are on
// the device.
struct PbPlan {
    std::vector<ProbDev> probs;
    std::vector<std::vector<double>> mats;
    std::vector<std::array<int, 5>> refs;   // per problem: Vx, Vxi, VyT, VyiT, dinv (indices into mats)
    std::vector<GemmTask> tasks;
    std::vector<int2> blocks;
    std::vector<int> bstart;
    long long N = 0;
    void bind(const std::vector<const double *> &dev) {
        for (size_t p = 0; p < probs.size(); ++p) {
            probs[p].Vx = dev[refs[p][0]];
            probs[p].Vxi = dev[refs[p][1]];
            probs[p].VyT = dev[refs[p][2]];
            probs[p].VyiT = dev[refs[p][3]];
            probs[p].dinv = dev[refs[p][4]];
        }
    }
};

static int pb_plan(int nprob, const int *nx, const int *ny, const int *bc, PbPlan &pl) {
    if (nprob <= 0 || !nx || !ny || !bc) { g_pb_err = "invalid arguments"; return -2; }
    for (int p = 0; p < nprob; ++p) {
        if (nx[p] < 1 || ny[p] < 1) { g_pb_err = "problem sizes must be positive"; return -2; }
        const int *c = bc + 4 * p;
        for (int s = 0; s < 4; ++s)
            if (c[s] != 0 && c[s] != 1) { g_pb_err = "boundary codes are 0 (Neumann) or 1 (Dirichlet)"; return -2; }
        if (!(c[0] || c[1] || c[2] || c[3])) { g_pb_err = "a pure Neumann problem is singular"; return -2; }
    }
    std::map<std::tuple<int, int, int, int>, std::pair<int, int>> cache;   // (axis, n, ends) -> (V, Vi)
    auto basis = [&](int axis, int n, int l, int r, int &iv, int &ivi, std::vector<double> &lam) {
        std::vector<double> v, vi;
        eigen_1d(n, l, r, v, vi, lam);
        const auto key = std::make_tuple(axis, n, l, r);
        const auto it = cache.find(key);
        if (it != cache.end()) { iv = it->second.first; ivi = it->second.second; return; }
        if (axis == 1) {   // the y bases are used as V^T and (V^-1)^T
            std::vector<double> vt((size_t)n * n), vit((size_t)n * n);
            for (int i = 0; i < n; ++i)
                for (int j = 0; j < n; ++j) { vt[(size_t)i * n + j] = v[(size_t)j * n + i]; vit[(size_t)i * n + j] = vi[(size_t)j * n + i]; }
            v.swap(vt);
            vi.swap(vit);
        }
        iv = (int)pl.mats.size();
        pl.mats.push_back(std::move(v));
        ivi = (int)pl.mats.size();
        pl.mats.push_back(std::move(vi));
        cache[key] = {iv, ivi};
    };
    long long off = 0;
    pl.bstart.assign(nprob + 1, 0);
    for (int p = 0; p < nprob; ++p) {
        ProbDev q{};
        q.nx = nx[p]; q.ny = ny[p]; q.off = off;
        for (int s = 0; s < 4; ++s) q.bc[s] = bc[4 * p + s];
        std::vector<double> lx, ly;
        std::array<int, 5> ref{};
        basis(0, q.nx, q.bc[0], q.bc[1], ref[0], ref[1], lx);
        basis(1, q.ny, q.bc[2], q.bc[3], ref[2], ref[3], ly);
        std::vector<double> dinv((size_t)q.nx * q.ny);
        for (int a = 0; a < q.nx; ++a)
            for (int c = 0; c < q.ny; ++c) {
                const double d = lx[a] + ly[c];
                dinv[(size_t)a * q.ny + c] = d > 1e-300 ? 1.0 / d : 0.0;
            }
        ref[4] = (int)pl.mats.size();
        pl.mats.push_back(std::move(dinv));
        for (int r0 = 0; r0 < q.nx; r0 += GT)
            for (int c0 = 0; c0 < q.ny; c0 += GT) pl.tasks.push_back(GemmTask{p, r0, c0, 0});
        pl.bstart[p] = (int)pl.blocks.size();
        for (int i0 = 0; i0 < q.nx * q.ny; i0 += RB) pl.blocks.push_back(make_int2(p, i0));
        pl.probs.push_back(q);
        pl.refs.push_back(ref);
        off += (long long)q.nx * q.ny;
    }
    pl.bstart[nprob] = (int)pl.blocks.size();
    pl.N = off;
    return 0;
}

// Work space of pic_pbatch_solve_noalloc: one array per device, reached
// from host code only through the address hipGetSymbolAddress returns for
// the current device (the symbol's name is not a device address on the
// host).  Every call on a device shares it.
constexpr size_t PB_SCRATCH_BYTES = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[PB_SCRATCH_BYTES];
constexpr int PB_MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static void *g_scratch_ptr[PB_MAX_DEVICES];

// address of this device's scratch array (looked up once per device)
static int pb_scratch(int device, unsigned char **out) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) {
            g_pb_err = std::string("cannot locate the solver work space: ") + hipGetErrorString(e);
            return -1;
        }
        g_scratch_ptr[device] = ptr;
    }
    *out = (unsigned char *)g_scratch_ptr[device];
    return 0;
}

// The launch sequence of one batched solve from x = 0: residual and
// convergence test per cycle, then the four GEMMs of the fast inverse on the
// problems not yet converged.  B supplies the problem tables and work
// vectors; b and x are the right-hand sides and the solution.
static void pb_enqueue(const pic_pbatch &B, const double *b, double *x, double tol, int max_cycles, hipStream_t st) {
    const int tb = 256;
    hipLaunchKernelGGL(k_pb_begin, dim3((B.nprob + tb - 1) / tb), dim3(tb), 0, st, B.nprob, B.done, B.status);
    hipLaunchKernelGGL(k_pb_zero, dim3((unsigned)((B.N + tb - 1) / tb)), dim3(tb), 0, st, B.N, x);
    for (int k = 0; k <= max_cycles; ++k) {
        hipLaunchKernelGGL(k_pb_residual, dim3(B.nblk), dim3(RB), 0, st, (const ProbDev *)B.P,
                           (const int2 *)B.blocks, b, (const double *)x, B.r, B.part, (const int *)B.done);
        hipLaunchKernelGGL(k_pb_finalize, dim3((B.nprob + tb - 1) / tb), dim3(tb), 0, st, B.nprob,
                           (const int *)B.bstart, (const double *)B.part, tol, k == max_cycles ? 1 : 0, B.done,
                           B.rel, B.status);
        if (k == max_cycles) break;
        const dim3 g(B.ntask), blk(GT * GT);
        hipLaunchKernelGGL(k_pb_gemm<0>, g, blk, 0, st, (const ProbDev *)B.P, (const GemmTask *)B.tasks,
                           (const double *)B.r, B.T1, B.T2, x, (const int *)B.done);
        hipLaunchKernelGGL(k_pb_gemm<1>, g, blk, 0, st, (const ProbDev *)B.P, (const GemmTask *)B.tasks,
                           (const double *)B.r, B.T1, B.T2, x, (const int *)B.done);
        hipLaunchKernelGGL(k_pb_gemm<2>, g, blk, 0, st, (const ProbDev *)B.P, (const GemmTask *)B.tasks,
                           (const double *)B.r, B.T1, B.T2, x, (const int *)B.done);
        hipLaunchKernelGGL(k_pb_gemm<3>, g, blk, 0, st, (const ProbDev *)B.P, (const GemmTask *)B.tasks,
                           (const double *)B.r, B.T1, B.T2, x, (const int *)B.done);
    }
}

extern "C" {

const char *pic_pbatch_last_error(void) { return g_pb_err.c_str(); }

int pic_pbatch_create(int nprob, const int *nx, const int *ny, const int *bc, int device, pic_pbatch **out) {
    if (!out) { g_pb_err = "invalid arguments"; return -2; }
    *out = nullptr;
    PbPlan pl;
    if (const int rc = pb_plan(nprob, nx, ny, bc, pl)) return rc;
    pic_pbatch *B = new pic_pbatch();
    B->nprob = nprob;
    B->dev = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&B->st, hipStreamNonBlocking) != hipSuccess) {
        g_pb_err = "cannot open the device";
        pb_destroy(B);
        return -1;
    }
    std::vector<const double *> mdev(pl.mats.size());
    for (size_t m = 0; m < pl.mats.size(); ++m) {
        double *d;
        if (pb_upload(B, &d, pl.mats[m]) != hipSuccess) { g_pb_err = "device allocation failed"; pb_destroy(B); return -1; }
        mdev[m] = d;
    }
    pl.bind(mdev);
    B->host = pl.probs;
    B->N = pl.N;
    B->ntask = (int)pl.tasks.size();
    B->nblk = (int)pl.blocks.size();
    if (pb_upload(B, &B->P, pl.probs) != hipSuccess || pb_upload(B, &B->tasks, pl.tasks) != hipSuccess ||
        pb_upload(B, &B->blocks, pl.blocks) != hipSuccess || pb_upload(B, &B->bstart, pl.bstart) != hipSuccess ||
        pb_alloc(B, &B->b, pl.N) != hipSuccess || pb_alloc(B, &B->x, pl.N) != hipSuccess ||
        pb_alloc(B, &B->r, pl.N) != hipSuccess || pb_alloc(B, &B->T1, pl.N) != hipSuccess ||
        pb_alloc(B, &B->T2, pl.N) != hipSuccess || pb_alloc(B, &B->part, 2 * (size_t)B->nblk) != hipSuccess ||
        pb_alloc(B, &B->rel, nprob) != hipSuccess || pb_alloc(B, &B->done, nprob) != hipSuccess ||
        pb_alloc(B, &B->status, nprob) != hipSuccess) {
        g_pb_err = "device allocation failed";
        pb_destroy(B);
        return -1;
    }
    *out = B;
    return 0;
}

int pic_pbatch_destroy(pic_pbatch *B) { return pb_destroy(B); }

long long pic_pbatch_unknowns(pic_pbatch *B) { return B ? B->N : -1; }

int pic_pbatch_set_rhs(pic_pbatch *B, const double *b) {
    if (!B || !b) { g_pb_err = "invalid arguments"; return -2; }
    PBCHK(hipMemcpyAsync(B->b, b, (size_t)B->N * sizeof(double), hipMemcpyHostToDevice, B->st));
    PBCHK(hipStreamSynchronize(B->st));
    return 0;
}

// Solve all problems from x = 0 with at most max_cycles refinement cycles;
// enqueued on the batch's stream (no host synchronisation).
int pic_pbatch_solve(pic_pbatch *B, double tol, int max_cycles) {
    if (!B || tol <= 0 || max_cycles < 1) { g_pb_err = "invalid arguments"; return -2; }
    pb_enqueue(*B, B->b, B->x, tol, max_cycles, B->st);
    PBCHK(hipGetLastError());
    return 0;
}

// waits for the solve; copies x (may be null), per-problem relative residuals
// (may be null) and returns the number of problems that did not converge
int pic_pbatch_result(pic_pbatch *B, double *x, double *rel_res) {
    if (!B) { g_pb_err = "invalid arguments"; return -2; }
    PBCHK(hipStreamSynchronize(B->st));
    if (x) PBCHK(hipMemcpyAsync(x, B->x, (size_t)B->N * sizeof(double), hipMemcpyDeviceToHost, B->st));
    if (rel_res) PBCHK(hipMemcpyAsync(rel_res, B->rel, (size_t)B->nprob * sizeof(double), hipMemcpyDeviceToHost, B->st));
    std::vector<int> stt(B->nprob);
    PBCHK(hipMemcpyAsync(stt.data(), B->status, (size_t)B->nprob * sizeof(int), hipMemcpyDeviceToHost, B->st));
    PBCHK(hipStreamSynchronize(B->st));
    int bad = 0;
    for (int v : stt) bad += v != 0;
    return bad;
}

int pic_pbatch_sync(pic_pbatch *B) {
    if (!B) { g_pb_err = "invalid arguments"; return -2; }
    PBCHK(hipStreamSynchronize(B->st));
    return 0;
}

// The batch of pic_pbatch_create + pic_pbatch_solve + pic_pbatch_result in
// one call that allocates no device memory: tables, eigenbases and work
// vectors are laid out in the device's scratch array (PB_SCRATCH_BYTES),
// right-hand sides and solution are the caller's device arrays.  The work
// is ordered on `stream` and the call waits for it.  Calls that overlap on
// one device race on the shared array.
int pic_pbatch_solve_noalloc(int nprob, const int *nx, const int *ny, const int *bc, int device, const double *b_dev,
                             double *x_dev, double tol, int max_cycles, double *rel_res, void *stream) {
    if (!b_dev || !x_dev || tol <= 0 || max_cycles < 1) { g_pb_err = "invalid arguments"; return -2; }
    if (device < 0 || device >= PB_MAX_DEVICES) { g_pb_err = "device ordinal out of range"; return -2; }
    PbPlan pl;
    if (const int rc = pb_plan(nprob, nx, ny, bc, pl)) return rc;
    // layout: tables first, then matrices, then the work vectors; every
    // region starts on a 256-byte boundary
    size_t at = 0;
    auto region = [&](size_t bytes) { const size_t o = at; at += (bytes + 255) & ~(size_t)255; return o; };
    const size_t oP = region(pl.probs.size() * sizeof(ProbDev));
    const size_t oT = region(pl.tasks.size() * sizeof(GemmTask));
    const size_t oB = region(pl.blocks.size() * sizeof(int2));
    const size_t oS = region(pl.bstart.size() * sizeof(int));
    std::vector<size_t> oM(pl.mats.size());
    for (size_t m = 0; m < pl.mats.size(); ++m) oM[m] = region(pl.mats[m].size() * sizeof(double));
    const size_t host_bytes = at;   // everything above is copied from the host
    const size_t oR = region((size_t)pl.N * sizeof(double));
    const size_t o1 = region((size_t)pl.N * sizeof(double));
    const size_t o2 = region((size_t)pl.N * sizeof(double));
    const size_t oPart = region(2 * pl.blocks.size() * sizeof(double));
    const size_t oRel = region((size_t)nprob * sizeof(double));
    const size_t oDone = region((size_t)nprob * sizeof(int));
    const size_t oStat = region((size_t)nprob * sizeof(int));
    if (at > PB_SCRATCH_BYTES) {
        g_pb_err = "batch needs " + std::to_string(at) + " bytes of work space, more than the " +
                   std::to_string(PB_SCRATCH_BYTES) + " of the no-allocation solve";
        return -3;
    }
    PBCHK(hipSetDevice(device));
    unsigned char *base = nullptr;
    if (pb_scratch(device, &base)) return -1;
    std::vector<const double *> mdev(pl.mats.size());
    for (size_t m = 0; m < pl.mats.size(); ++m) mdev[m] = (const double *)(base + oM[m]);
    pl.bind(mdev);
    std::vector<unsigned char> host(host_bytes, 0);
    auto put = [&](size_t o, const void *src, size_t n) { if (n) memcpy(host.data() + o, src, n); };
    put(oP, pl.probs.data(), pl.probs.size() * sizeof(ProbDev));
    put(oT, pl.tasks.data(), pl.tasks.size() * sizeof(GemmTask));
    put(oB, pl.blocks.data(), pl.blocks.size() * sizeof(int2));
    put(oS, pl.bstart.data(), pl.bstart.size() * sizeof(int));
    for (size_t m = 0; m < pl.mats.size(); ++m) put(oM[m], pl.mats[m].data(), pl.mats[m].size() * sizeof(double));
    const hipStream_t st = (hipStream_t)stream;
    PBCHK(hipMemcpyAsync(base, host.data(), host_bytes, hipMemcpyHostToDevice, st));
    pic_pbatch B;   // a view of the scratch layout (owns nothing)
    B.nprob = nprob;
    B.N = pl.N;
    B.dev = device;
    B.P = (ProbDev *)(base + oP);
    B.tasks = (GemmTask *)(base + oT);
    B.ntask = (int)pl.tasks.size();
    B.blocks = (int2 *)(base + oB);
    B.nblk = (int)pl.blocks.size();
    B.bstart = (int *)(base + oS);
    B.r = (double *)(base + oR);
    B.T1 = (double *)(base + o1);
    B.T2 = (double *)(base + o2);
    B.part = (double *)(base + oPart);
    B.rel = (double *)(base + oRel);
    B.done = (int *)(base + oDone);
    B.status = (int *)(base + oStat);
    pb_enqueue(B, b_dev, x_dev, tol, max_cycles, st);
    PBCHK(hipGetLastError());
    std::vector<int> stt(nprob);
    if (rel_res) PBCHK(hipMemcpyAsync(rel_res, B.rel, (size_t)nprob * sizeof(double), hipMemcpyDeviceToHost, st));
    PBCHK(hipMemcpyAsync(stt.data(), B.status, (size_t)nprob * sizeof(int), hipMemcpyDeviceToHost, st));
    PBCHK(hipStreamSynchronize(st));
    int bad = 0;
    for (int v : stt) bad += v != 0;
    return bad;
}

}  // extern "C"
