// Particle and node kernels of the PIC-MCC step for gfx950.
//
// Particles of one species live in structure-of-arrays form, sorted by cell
// (a cell is named by its lower-left node, as in the reference), with a CSR
// offset array cell_start[Gsize+1].  Every particle kernel runs one 64-lane
// wavefront per cell: the cell's four E nodes and its 9-entry route table are
// wave-uniform, and ranks inside the cell come from ballots, so the new order
// is fixed without atomics:
//   new bucket(d) = stayers of d (old order)
//                 ++ arrivals grouped by source cell ascending (old order)
//                 ++ collision products in parent order.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "multi.h"
#include "pic_device.h"
#include "pic_kernels.h"
#include "pic_setup.h"

namespace pic {

Recorder *&launch_recorder() {
    static thread_local Recorder *r = nullptr;
    return r;
}

namespace {

template <typename R>
__device__ inline R wave_sum(R v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

template <typename R>
struct CellE {
    R ews, ewn, ees, een, yws, ywn, yes, yen;
    __device__ void load(const R *Ex, const R *Ey, int c, int ngy) {
        ews = Ex[c]; ewn = Ex[c + 1]; ees = Ex[c + ngy]; een = Ex[c + ngy + 1];
        yws = Ey[c]; ywn = Ey[c + 1]; yes = Ey[c + ngy]; yen = Ey[c + ngy + 1];
    }
};

// voltage of conductor i at time t (src/cuda_Field.cu:10-20)
template <typename R>
__device__ inline R cond_voltage(const VoltConst &v, int i, double t) {
    R acc = 0;
    if (v.flag[i]) {
        for (int j = 0; j < v.nsrc; ++j)
            if (v.mid[j] == i + 1)
                acc = (R)((double)acc + ((double)(R)v.dc[j] + (double)(R)v.ac[j] * dm_sin((double)(R)v.w[j] * t + (double)(R)v.ph[j])));
    }
    return acc;
}

}  // namespace

// ============================================================ staged: Move
template <typename R>
__global__ __launch_bounds__(256) void k_push(DevConst<R> k, PView<R> p, int isp, uint8_t *tag) {
    int lane;
    int c = wave_cell(k.ncy, k.ngy, k.Csize, lane);
    if (c < 0) return;
    int k0, k1;
    cell_range(p, c, k0, k1);
    if (k0 == k1) return;
    CellE<R> e;
    e.load(k.Ex, k.Ey, c, k.ngy);
    const R As = k.Ascale[isp];
    for (int kk = k0 + lane; kk < k1; kk += 64) {
        R x = p.x[kk], y = p.y[kk], vx = p.vx[kk], vy = p.vy[kk];
        int code = push_particle<R>(e.ews, e.ewn, e.ees, e.een, e.yws, e.ywn, e.yes, e.yen, As, k.dt_dx, k.dt_dy,
                                    x, y, vx, vy);
        p.x[kk] = x; p.y[kk] = y; p.vx[kk] = vx; p.vy[kk] = vy;
        tag[kk] = (uint8_t)code;
    }
}

// side effects of a particle that leaves through a wall
template <typename R>
__device__ inline void wall_effects(const DevConst<R> &k, const Route &rt, int isp, int c, int code, R x, R y, R vx,
                                    R vy, int *condcharge) {
    if (rt.dest == -2) {
        atomicAdd(&condcharge[isp * k.CondNUMR + rt.info], 1);
    } else if (rt.dest == -1) {
        atomicAdd(&k.lost[isp], 1);
    } else if (rt.dest == -3) {
        int mcid = code_mcid(code, k.ngy);
        int sm = padded_index(c, mcid, k.ngy);
        int pa, pb;
        R a, b;
        dielectric_split<R>(k.StructIdx, k.ngy, c, mcid, sm, x, y, vx, vy, pa, a, pb, b);
        atomicAdd(&k.sigma_hits[(size_t)isp * k.Gsize + pa], sigma_fixed((double)a));
        atomicAdd(&k.sigma_hits[(size_t)isp * k.Gsize + pb], sigma_fixed((double)b));
    }
}

// ============================================================ staged: Sort count
// key[k] = slot << 24 | rank within (cell, slot); -1 when the particle leaves
template <typename R>
__global__ __launch_bounds__(256) void k_sort_count(DevConst<R> k, PView<R> p, int isp, const uint8_t *tag, int *key,
                                                    int *cnt, int *condcharge) {
    int lane;
    int c = wave_cell(k.ncy, k.ngy, k.Csize, lane);
    if (c < 0) return;
    int k0, k1;
    cell_range(p, c, k0, k1);
    const unsigned long long lt = lanes_below(lane);
    int run[ROUTE_SLOTS];
#pragma unroll
    for (int s = 0; s < ROUTE_SLOTS; ++s) run[s] = 0;
    for (int base = k0; base < k1; base += 64) {
        int kk = base + lane;
        int slot = -1;
        if (kk < k1) {
            int code = tag[kk];
            Route rt = k.route[c * ROUTE_SLOTS + code];
            if (rt.dest >= 0) slot = rt.info & 15;
            else wall_effects<R>(k, rt, isp, c, code, p.x[kk], p.y[kk], p.vx[kk], p.vy[kk], condcharge);
        }
        int rank = -1;
#pragma unroll
        for (int s = 0; s < ROUTE_SLOTS; ++s) {
            unsigned long long m = __ballot(slot == s);
            if (slot == s) rank = run[s] + __popcll(m & lt);
            run[s] += __popcll(m);
        }
        if (kk < k1) key[kk] = slot < 0 ? -1 : ((slot << 24) | rank);
    }
    if (lane == 0) {
#pragma unroll
        for (int s = 0; s < ROUTE_SLOTS; ++s) cnt[c * ROUTE_SLOTS + s] = run[s];
    }
}

// ============================================================ staged: Move + Sort count
// k_push and k_sort_count in one pass while the particle is in registers
// (whole staged steps), every species in one launch (grid row = species):
// the same positions, tags, keys, counts and wall side effects, one read of
// the particle instead of two.
template <typename R>
__global__ __launch_bounds__(256) void k_push_count(DevConst<R> k, StagedSet<R> set, int *condcharge) {
    int lane;
    int c = wave_cell(k.ncy, k.ngy, k.Csize, lane);
    if (c < 0) return;
    const int isp = blockIdx.y;
    const PView<R> p = set.p[isp];
    uint8_t *tag = set.tag[isp];
    int *key = set.key[isp], *cnt = set.cnt[isp];
    int k0, k1;
    cell_range(p, c, k0, k1);
    const unsigned long long lt = lanes_below(lane);
    int run = 0;   // lane s < ROUTE_SLOTS: survivors of slot s so far
    if (k0 < k1) {
        CellE<R> e;
        e.load(k.Ex, k.Ey, c, k.ngy);
        const R As = k.Ascale[isp];
        for (int base = k0; base < k1; base += 64) {
            const int kk = base + lane;
            int slot = -1;
            if (kk < k1) {
                R x = p.x[kk], y = p.y[kk], vx = p.vx[kk], vy = p.vy[kk];
                const int code = push_particle<R>(e.ews, e.ewn, e.ees, e.een, e.yws, e.ywn, e.yes, e.yen, As, k.dt_dx,
                                                  k.dt_dy, x, y, vx, vy);
                p.x[kk] = x; p.y[kk] = y; p.vx[kk] = vx; p.vy[kk] = vy;
                tag[kk] = (uint8_t)code;
                const Route rt = k.route[c * ROUTE_SLOTS + code];
                if (rt.dest >= 0) slot = rt.info & 15;
                else wall_effects<R>(k, rt, isp, c, code, x, y, vx, vy, condcharge);
            }
            // ranks over the slots present in these 64 (usually 2-4 of
            // the 9): lane s holds slot s's running count
            int rank = -1;
            unsigned long long pend = __ballot(slot >= 0);
            while (pend) {
                const int first = __builtin_ctzll(pend);
                const int sl = __builtin_amdgcn_readfirstlane(__shfl(slot, first, 64));
                const unsigned long long m = __ballot(slot == sl);
                const int before = __builtin_amdgcn_readlane(run, sl);
                if (slot == sl) rank = before + __popcll(m & lt);
                if (lane == sl) run += __popcll(m);
                pend &= ~m;
            }
            if (kk < k1) key[kk] = slot < 0 ? -1 : ((slot << 24) | rank);
        }
    }
    if (lane < ROUTE_SLOTS) cnt[c * ROUTE_SLOTS + lane] = run;
}

// ============================================================ bases
// Per destination cell d: new total, then the base offset of every
// (source cell, slot) entry feeding d in canonical order.  prod_cnt may be
// null (no products).  Products of d follow all survivors of d.
__global__ void k_totals(int Gsize, const int *inv_start, const int *inv_ent, const int *cnt, const int *prod_cnt,
                         int *total, int *alive_total, int *prod_total) {
    int d = blockIdx.x * blockDim.x + threadIdx.x;
    if (d >= Gsize) return;
    int a = 0, pr = 0;
    for (int e = inv_start[d]; e < inv_start[d + 1]; ++e) {
        int ent = inv_ent[e];
        a += cnt[ent];
        if (prod_cnt) pr += prod_cnt[ent];
    }
    total[d] = a + pr;
    alive_total[d] = a;
    if (prod_total) prod_total[d] = pr;
}

__global__ void k_assign_bases(int Gsize, const int *inv_start, const int *inv_ent, const int *cnt, const int *prod_cnt,
                               const int *start, const int *alive_total, const int *prod_start, int *base,
                               int *pbase, int *pid_off) {
    int d = blockIdx.x * blockDim.x + threadIdx.x;
    if (d >= Gsize) return;
    int run = start[d];
    int prun = start[d] + alive_total[d];
    int irun = prod_start ? prod_start[d] : 0;
    for (int e = inv_start[d]; e < inv_start[d + 1]; ++e) {
        int ent = inv_ent[e];
        base[ent] = run;
        run += cnt[ent];
        if (prod_cnt) {
            pbase[ent] = prun;
            pid_off[ent] = irun;
            prun += prod_cnt[ent];
            irun += prod_cnt[ent];
        }
    }
}

// Exclusive scans of up to 3 int arrays of length n (out[n] = total) in two
// launches: per-tile sums, then each tile adds the sums of the tiles before
// it and scans locally.  Tiles are SCAN_TILE = 4 * 1024 elements.
constexpr int SCAN_TILE = 4096;
struct ScanJobs {
    const int *in[3];
    int *out[3];
    int *sums[3];
};

__device__ inline int block_sum_1024(int v, int *sh) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if (lane == 0) sh[w] = v;
    __syncthreads();
    int t = 0;
    if (threadIdx.x < 64) {
        t = threadIdx.x < 16 ? sh[threadIdx.x] : 0;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) t += __shfl_xor(t, off, 64);
    }
    return t;   // valid in thread 0
}

__global__ __launch_bounds__(1024) void k_scan_tile_sums(ScanJobs J, int n) {
    __shared__ int sh[16];
    const int a = blockIdx.y;
    const int *in = J.in[a];
    const int base = blockIdx.x * SCAN_TILE + threadIdx.x * 4;
    int v = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q)
        if (base + q < n) v += in[base + q];
    v = block_sum_1024(v, sh);
    if (threadIdx.x == 0) J.sums[a][blockIdx.x] = v;
}

__global__ __launch_bounds__(1024) void k_scan_tiles(ScanJobs J, int n, int ntiles) {
    __shared__ int wtot[16];
    __shared__ int sh_off, sh_tot;
    const int a = blockIdx.y;
    const int *in = J.in[a];
    int *out = J.out[a];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (w == 0) {
        int before = 0, all = 0;
        for (int t = lane; t < ntiles; t += 64) {
            const int v = J.sums[a][t];
            all += v;
            if (t < (int)blockIdx.x) before += v;
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            before += __shfl_xor(before, off, 64);
            all += __shfl_xor(all, off, 64);
        }
        if (lane == 0) { sh_off = before; sh_tot = all; }
    }
    const int base = blockIdx.x * SCAN_TILE + threadIdx.x * 4;
    int v[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) v[q] = base + q < n ? in[base + q] : 0;
    const int tsum = v[0] + v[1] + v[2] + v[3];
    int incl = tsum;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int u = __shfl_up(incl, off, 64);
        if (lane >= off) incl += u;
    }
    if (lane == 63) wtot[w] = incl;
    __syncthreads();
    int wbefore = 0;
    for (int q = 0; q < w; ++q) wbefore += wtot[q];
    int run = sh_off + wbefore + incl - tsum;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        if (base + q < n) out[base + q] = run;
        run += v[q];
    }
    if (blockIdx.x == ntiles - 1 && threadIdx.x == 0) out[n] = sh_tot;
}

// ============================================================ fused step: bases of both species
// Between the two passes of the fused step, for electrons and ions at once:
// each destination cell's total (survivors routed to it, plus its ionisation
// products: one electron and one ion per event), the exclusive scans of the
// three totals (electron and ion cell_start, global product order), and the
// base of every incoming (source cell, slot) entry in the destination's new
// bucket (the destination's start plus the counts of the entries before it
// in canonical order; products follow the survivors).  Two launches over
// tiles of BASES_TILE destinations, one per thread.
constexpr int BASES_TILE = 256;

// a destination's (at most 9) incoming entries and their counts, all loads
// issued together
struct DestEntries {
    int n;
    int ent[ROUTE_SLOTS], ce[ROUTE_SLOTS], ci[ROUTE_SLOTS], cp[ROUTE_SLOTS];
    __device__ void load(int d, const int *inv_start, const int *inv_ent, const int *cnt_e, const int *cnt_i,
                         const int *prod_cnt) {
        const int e0 = inv_start[d];
        n = inv_start[d + 1] - e0;   // <= 9, checked at set-up
#pragma unroll
        for (int q = 0; q < ROUTE_SLOTS; ++q) ent[q] = q < n ? inv_ent[e0 + q] : -1;
#pragma unroll
        for (int q = 0; q < ROUTE_SLOTS; ++q) {
            const bool v = ent[q] >= 0;
            ce[q] = v ? cnt_e[ent[q]] : 0;
            ci[q] = v ? cnt_i[ent[q]] : 0;
            cp[q] = v ? prod_cnt[ent[q]] : 0;
        }
    }
};

// sum over the block's 4 waves, valid in every thread
__device__ inline int block_sum_256(int v, int *sh) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if (lane == 0) sh[w] = v;
    __syncthreads();
    const int t = sh[0] + sh[1] + sh[2] + sh[3];
    __syncthreads();
    return t;
}

struct PairTotalsJob {
    int G, nt;
    const int *inv_start, *inv_ent, *cnt_e, *cnt_i, *prod_cnt;
    int *total_e, *alive_e, *total_i, *alive_i, *prod_total, *sums;
};

__device__ __forceinline__ void pair_totals_body(const PairTotalsJob &J, int bx) {
    const int G = J.G, nt = J.nt;
    const int *inv_start = J.inv_start, *inv_ent = J.inv_ent, *cnt_e = J.cnt_e, *cnt_i = J.cnt_i;
    const int *prod_cnt = J.prod_cnt;
    int *total_e = J.total_e, *alive_e = J.alive_e, *total_i = J.total_i, *alive_i = J.alive_i;
    int *prod_total = J.prod_total, *sums = J.sums;

    __shared__ int sh[4];
    const int d = bx * BASES_TILE + threadIdx.x;
    int te = 0, ti = 0, tp = 0;
    if (d < G) {
        DestEntries E;
        E.load(d, inv_start, inv_ent, cnt_e, cnt_i, prod_cnt);
        int ae = 0, ai = 0, pr = 0;
#pragma unroll
        for (int q = 0; q < ROUTE_SLOTS; ++q) { ae += E.ce[q]; ai += E.ci[q]; pr += E.cp[q]; }
        total_e[d] = ae + pr; alive_e[d] = ae;
        total_i[d] = ai + pr; alive_i[d] = ai;
        prod_total[d] = pr;
        te = ae + pr; ti = ai + pr; tp = pr;
    }
    te = block_sum_256(te, sh);
    ti = block_sum_256(ti, sh);
    tp = block_sum_256(tp, sh);
    if (threadIdx.x == 0) { sums[bx] = te; sums[nt + bx] = ti; sums[2 * nt + bx] = tp; }
}

__global__ __launch_bounds__(BASES_TILE) void k_pair_totals(PairTotalsJob J) { pair_totals_body(J, blockIdx.x); }
__global__ __launch_bounds__(BASES_TILE) void k_pair_totals_m(const PairTotalsJob *J, const uint32_t *map) {
    int lb;
    const PairTotalsJob &a = J[multi_job(map, lb)];
    pair_totals_body(a, lb);
}


struct PairBasesJob {
    int G, nt;
    const int *sums, *total_e, *total_i, *prod_total, *alive_e, *alive_i;
    const int *inv_start, *inv_ent, *cnt_e, *cnt_i, *prod_cnt;
    int *start_e, *start_i, *prod_start, *base_e, *base_i, *pbase_e, *pbase_i, *pid_off;
};

__device__ __forceinline__ void pair_scan_bases_body(const PairBasesJob &J, int bx) {
    const int G = J.G, nt = J.nt;
    const int *sums = J.sums, *total_e = J.total_e, *total_i = J.total_i, *prod_total = J.prod_total;
    const int *alive_e = J.alive_e, *alive_i = J.alive_i;
    const int *inv_start = J.inv_start, *inv_ent = J.inv_ent, *cnt_e = J.cnt_e, *cnt_i = J.cnt_i;
    const int *prod_cnt = J.prod_cnt;
    int *start_e = J.start_e, *start_i = J.start_i, *prod_start = J.prod_start, *base_e = J.base_e;
    int *base_i = J.base_i, *pbase_e = J.pbase_e, *pbase_i = J.pbase_i, *pid_off = J.pid_off;

    __shared__ int wtot[3][4];
    __shared__ int sh_off[3], sh_tot[3];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (w == 0) {
        // the tiles before this one, and all tiles, per array
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            int before = 0, all = 0;
            for (int t = lane; t < nt; t += 64) {
                const int v = sums[a * nt + t];
                all += v;
                if (t < (int)bx) before += v;
            }
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) {
                before += __shfl_xor(before, off, 64);
                all += __shfl_xor(all, off, 64);
            }
            if (lane == 0) { sh_off[a] = before; sh_tot[a] = all; }
        }
    }
    const int d = bx * BASES_TILE + threadIdx.x;
    const bool in = d < G;
    DestEntries E;
    if (in) E.load(d, inv_start, inv_ent, cnt_e, cnt_i, prod_cnt);
    const int v[3] = {in ? total_e[d] : 0, in ? total_i[d] : 0, in ? prod_total[d] : 0};
    int incl[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        incl[a] = v[a];
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const int u = __shfl_up(incl[a], off, 64);
            if (lane >= off) incl[a] += u;
        }
        if (lane == 63) wtot[a][w] = incl[a];
    }
    __syncthreads();
    int run[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        int wbefore = 0;
        for (int q = 0; q < w; ++q) wbefore += wtot[a][q];
        run[a] = sh_off[a] + wbefore + incl[a] - v[a];
    }
    if (in) {
        start_e[d] = run[0]; start_i[d] = run[1]; prod_start[d] = run[2];
        const int pe0 = run[0] + alive_e[d], pi0 = run[1] + alive_i[d];
        int se = 0, si = 0, sp = 0;
#pragma unroll
        for (int q = 0; q < ROUTE_SLOTS; ++q) {
            if (q < E.n) {
                const int e = E.ent[q];
                base_e[e] = run[0] + se; base_i[e] = run[1] + si;
                pbase_e[e] = pe0 + sp; pbase_i[e] = pi0 + sp; pid_off[e] = run[2] + sp;
            }
            se += E.ce[q]; si += E.ci[q]; sp += E.cp[q];
        }
    }
    if (bx == nt - 1 && threadIdx.x == 0) {
        start_e[G] = sh_tot[0]; start_i[G] = sh_tot[1]; prod_start[G] = sh_tot[2];
    }
}

__global__ __launch_bounds__(BASES_TILE) void k_pair_scan_bases(PairBasesJob J) { pair_scan_bases_body(J, blockIdx.x); }
__global__ __launch_bounds__(BASES_TILE) void k_pair_scan_bases_m(const PairBasesJob *J, const uint32_t *map) {
    int lb;
    const PairBasesJob &a = J[multi_job(map, lb)];
    pair_scan_bases_body(a, lb);
}


// Totals and bases of the staged sort for every species in one launch each
// (no products; grid row = species): the same per-destination sums and
// entry bases as k_totals and k_assign_bases.
__global__ void k_totals_sp(int Gsize, const int *inv_start, const int *inv_ent, SortSpecies j) {
    const int d = blockIdx.x * blockDim.x + threadIdx.x;
    const int s = blockIdx.y;
    if (d >= Gsize) return;
    const int *cnt = j.cnt[s];
    int a = 0;
    for (int e = inv_start[d]; e < inv_start[d + 1]; ++e) a += cnt[inv_ent[e]];
    j.total[s][d] = a;
}

__global__ void k_assign_bases_sp(int Gsize, const int *inv_start, const int *inv_ent, SortSpecies j) {
    const int d = blockIdx.x * blockDim.x + threadIdx.x;
    const int s = blockIdx.y;
    if (d >= Gsize) return;
    const int *cnt = j.cnt[s];
    int *base = j.base[s];
    int run = j.start[s][d];
    for (int e = inv_start[d]; e < inv_start[d + 1]; ++e) {
        const int ent = inv_ent[e];
        base[ent] = run;
        run += cnt[ent];
    }
}

// ============================================================ staged: Sort scatter
template <typename R>
__global__ __launch_bounds__(256) void k_sort_scatter(DevConst<R> k, PView<R> p, PView<R> q, const uint8_t *tag,
                                                      const int *key, const int *base, int cap, int *err) {
    int lane;
    int c = wave_cell(k.ncy, k.ngy, k.Csize, lane);
    if (c < 0) return;
    int k0, k1;
    cell_range(p, c, k0, k1);
    for (int kk = k0 + lane; kk < k1; kk += 64) {
        int ky = key[kk];
        if (ky < 0) continue;
        int slot = ky >> 24, rank = ky & 0xffffff;
        int code = tag[kk];
        Route rt = k.route[c * ROUTE_SLOTS + code];
        R x = p.x[kk], y = p.y[kk], vx = p.vx[kk], vy = p.vy[kk];
        apply_reflect<R>(rt.info, x, y, vx, vy);
        int pos = base[c * ROUTE_SLOTS + slot] + rank;
        if (pos >= cap) { atomicOr(err, 1); continue; }
        q.x[pos] = x; q.y[pos] = y; q.vx[pos] = vx; q.vy[pos] = vy; q.vz[pos] = p.vz[kk]; q.id[pos] = p.id[kk];
    }
}

// k_sort_scatter for every species in one launch (grid row = species)
template <typename R>
__global__ __launch_bounds__(256) void k_sort_scatter_all(DevConst<R> k, StagedSet<R> set, int *err) {
    int lane;
    int c = wave_cell(k.ncy, k.ngy, k.Csize, lane);
    if (c < 0) return;
    const int isp = blockIdx.y;
    const PView<R> p = set.p[isp], q = set.q[isp];
    const uint8_t *tag = set.tag[isp];
    const int *key = set.key[isp], *base = set.base[isp];
    const int cap = q.cap;
    int k0, k1;
    cell_range(p, c, k0, k1);
    for (int kk = k0 + lane; kk < k1; kk += 64) {
        int ky = key[kk];
        if (ky < 0) continue;
        int slot = ky >> 24, rank = ky & 0xffffff;
        int code = tag[kk];
        Route rt = k.route[c * ROUTE_SLOTS + code];
        R x = p.x[kk], y = p.y[kk], vx = p.vx[kk], vy = p.vy[kk];
        apply_reflect<R>(rt.info, x, y, vx, vy);
        int pos = base[c * ROUTE_SLOTS + slot] + rank;
        if (pos >= cap) { atomicOr(err, 1); continue; }
        q.x[pos] = x; q.y[pos] = y; q.vx[pos] = vx; q.vy[pos] = vy; q.vz[pos] = p.vz[kk]; q.id[pos] = p.id[kk];
    }
}

// ============================================================ staged: MCC
// Collisions in place; ionisation products are staged per parent slot.
// prank[k] = rank of the parent's product within its cell (-1: none).
template <typename R>
__global__ __launch_bounds__(256) void k_mcc_collide(DevConst<R> k, PView<R> p, int isp, uint32_t step,
                                                     unsigned long long *mccr, int *pcnt, int *prank, R *stage) {
    int lane;
    int c = wave_cell(k.ncy, k.ngy, k.Csize, lane);
    if (c < 0) return;
    int k0, k1;
    cell_range(p, c, k0, k1);
    const unsigned long long lt = lanes_below(lane);
    int run = 0;
    const size_t cap = (size_t)p.cap;
    for (int base = k0; base < k1; base += 64) {
        int kk = base + lane;
        bool prod = false;
        if (kk < k1) {
            R vx = p.vx[kk], vy = p.vy[kk], vz = p.vz[kk];
            int ev;
            if (isp == 0) {
                R evx = 0, evy = 0, evz = 0, ivx = 0, ivy = 0, ivz = 0;
                ev = mcc_ar_electron<R>(k, c, k.P1[c], k.P2[c], p.id[kk], step, vx, vy, vz, evx, evy, evz, ivx, ivy, ivz);
                if (ev == 3 || ev == 4) {
                    prod = true;
                    stage[kk] = evx; stage[cap + kk] = evy; stage[2 * cap + kk] = evz;
                    stage[3 * cap + kk] = ivx; stage[4 * cap + kk] = ivy; stage[5 * cap + kk] = ivz;
                }
            } else {
                ev = mcc_ar_ion<R>(k, c, k.Pion[c], p.id[kk], step, vx, vy, vz);
            }
            if (ev >= 0) {
                atomicAdd(&mccr[(size_t)c * k.TnRct + ev], 1ull);
                p.vx[kk] = vx; p.vy[kk] = vy; p.vz[kk] = vz;
            }
        }
        unsigned long long m = __ballot(prod);
        if (kk < k1) prank[kk] = prod ? run + __popcll(m & lt) : -1;
        run += __popcll(m);
    }
    if (lane == 0) pcnt[c] = run;
}

// new bucket d = old bucket d ++ products of d; ids of products are
// next_id + global product rank
template <typename R>
__global__ __launch_bounds__(256) void k_mcc_insert(DevConst<R> k, PView<R> src, PView<R> dst, const int *new_start,
                                                    const int *prod_start, const int *prod_parent_start,
                                                    const PView<R> parent, const int *prank, const R *stage,
                                                    int stage_off, const uint32_t *next_id, int cap, int *err) {
    int lane;
    int c = wave_cell(k.ncy, k.ngy, k.Csize, lane);
    if (c < 0) return;
    int k0, k1;
    cell_range(src, c, k0, k1);
    int ns = new_start[c];
    for (int kk = k0 + lane; kk < k1; kk += 64) {
        int pos = ns + (kk - k0);
        if (pos >= cap) { atomicOr(err, 1); continue; }
        dst.x[pos] = src.x[kk]; dst.y[pos] = src.y[kk];
        dst.vx[pos] = src.vx[kk]; dst.vy[pos] = src.vy[kk]; dst.vz[pos] = src.vz[kk]; dst.id[pos] = src.id[kk];
    }
    if (!prank) return;
    int m = k1 - k0;
    int q0, q1;
    cell_range(parent, c, q0, q1);
    const size_t pcap = (size_t)parent.cap;
    for (int kk = q0 + lane; kk < q1; kk += 64) {
        int r = prank[kk];
        if (r < 0) continue;
        int pos = ns + m + r;
        if (pos >= cap) { atomicOr(err, 1); continue; }
        dst.x[pos] = parent.x[kk]; dst.y[pos] = parent.y[kk];
        dst.vx[pos] = stage[(size_t)stage_off * pcap + kk];
        dst.vy[pos] = stage[(size_t)(stage_off + 1) * pcap + kk];
        dst.vz[pos] = stage[(size_t)(stage_off + 2) * pcap + kk];
        dst.id[pos] = *next_id + (uint32_t)(prod_start[c] + r);
    }
    (void)prod_parent_start;
}

// ============================================================ fused step
// Pass 1 (count) pushes every particle, applies the walls and the
// null-collision test, and counts survivors per (cell, slot).  Colliding
// particles are queued per wave and collided 64 at a time, so the long
// collision code runs on full wavefronts; their final state is written back
// in place, and a per-particle flag byte records the move code, whether the
// particle collided and whether it made an ionisation pair (whose velocities
// go to a stash indexed like the particle).  Pass 2 (scatter) re-pushes only
// the particles that did not collide (same bits), ranks everything with
// ballots in particle order and writes the new, sorted arrays; it draws no
// random numbers.
//
// Pass 2 stages each wave's cell in chunks of SCHUNK particles through LDS:
// every lane issues the loads of all its particles of the chunk at once (one
// memory round trip per chunk), then walks them 64 at a time.  The cell's 9
// route entries are loaded once per wave and fetched per particle with a
// lane shuffle.
constexpr int SCHUNK = 256;   // pass 2 chunk (larger chunks measured faster: more loads in flight)
constexpr int FWAVES = 4;   // waves per 256-thread block
enum { KIND_AR_ELECTRON = 0, KIND_AR_ION = 1 };
enum { FLAG_CODE = 15, FLAG_DONE = 16, FLAG_PROD = 32 };
// Collision events are counted per (wave, route code, reaction) in LDS and
// added to the per-node counters once per wave at the end of the pass
// (integer sums: the totals do not depend on the order); reactions beyond
// EV_MAX go to the node counters directly.
constexpr int EV_MAX = 10;

template <typename R, int C = SCHUNK>
struct FStage {
    R x[C], y[C], vx[C], vy[C], vz[C];
    uint32_t id[C];
};

// Orders the LDS accesses of one wavefront (the hardware already executes a
// wave's LDS instructions in order; this keeps the compiler from moving
// them) between a phase that writes slots and one that reads other lanes'.
__device__ inline void wave_lds_fence() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// global -> LDS for particles [b0, b0+n), n >= 1.  Lanes past the end load the
// last particle again (same cache line, never read back), so the 6*C/64
// loads of a lane form one straight-line block the compiler can batch.
template <typename R, int C>
__device__ inline void stage_chunk(FStage<R, C> &S, const PView<R> &p, int b0, int n, int lane) {
    R tx[C / 64], ty[C / 64], tvx[C / 64], tvy[C / 64], tvz[C / 64];
    uint32_t tid[C / 64];
#pragma unroll
    for (int j = 0; j < C / 64; ++j) {
        int i = b0 + min(j * 64 + lane, n - 1);
        tx[j] = p.x[i]; ty[j] = p.y[i]; tvx[j] = p.vx[i]; tvy[j] = p.vy[i]; tvz[j] = p.vz[i]; tid[j] = p.id[i];
    }
#pragma unroll
    for (int j = 0; j < C / 64; ++j) {
        int i = j * 64 + lane;
        S.x[i] = tx[j]; S.y[i] = ty[j]; S.vx[i] = tvx[j]; S.vy[i] = tvy[j]; S.vz[i] = tvz[j]; S.id[i] = tid[j];
    }
}

// null-collision test of one particle at its destination node: the index of
// the colliding draw, or -1
template <typename R, int KIND>
__device__ inline int null_test(const DevConst<R> &k, R pb, uint32_t id, uint32_t step) {
    if (KIND == KIND_AR_ELECTRON) return mcc_ar_electron_null<R>(k, pb, id, step);
    return mcc_ar_ion_null<R>(k, pb, id, step) ? 0 : -1;
}

// collision of a queued particle; returns the reaction index or -1
template <typename R, int KIND>
__device__ inline int collide(const DevConst<R> &k, int dest, R pa, uint32_t id, uint32_t step, int jc, R &vx, R &vy,
                              R &vz, R &evx, R &evy, R &evz, R &ivx, R &ivy, R &ivz) {
    if (KIND == KIND_AR_ELECTRON)
        return mcc_ar_electron_collide<R>(k, dest, pa, id, step, jc, vx, vy, vz, evx, evy, evz, ivx, ivy, ivz);
    Stream st;
    st.init(k.seed, id, step, mcc_tag(1));
    st.seek(1);
    return mcc_ar_ion_collide<R>(k, dest, st, vx, vy, vz);
}

// Pass 1.  flag[k] = move code | FLAG_DONE (collided: final state written
// back in place) | FLAG_PROD (made an ionisation pair; the new electron's
// and ion's velocities in stash[q * stash_stride + k], q = 0..5).
//
// A wave takes CPW consecutive cells.  Their particles are one contiguous
// range (cells of one column are consecutive nodes, and the node between
// two columns starts no cell), so the wave walks the range 64 particles at a
// time with no idle lanes at cell ends, the next 64 particles' loads in
// flight while the current ones are pushed.  A lane finds its cell by
// comparing its index with the cells' starts; the cells' E corners and route
// entries sit in a per-wave LDS table.  Colliding particles go to a per-wave
// LDS queue with their post-push state, and whenever 64 are queued they
// collide together, one per lane, so the long collision code runs on full
// wavefronts.  The waves of a block never wait for each other.
constexpr int CPW = 4;          // cells per wave (at most 16: the queue keeps the cell in 4 bits)
constexpr int CQCAP = 128;      // queued colliders per wave (< 64 left after each collide, + 64 per walk step)

template <typename R>
struct CountWave {
    // collider queue: post-push state, particle index and (draw << 8 | cell << 4 | move code)
    R qx[CQCAP], qy[CQCAP], qvx[CQCAP], qvy[CQCAP], qvz[CQCAP];
    uint32_t qid[CQCAP];
    int qkk[CQCAP], qmeta[CQCAP];
    R e[CPW][8];                                 // E corners of each cell (ews ewn ees een yws ywn yes yen)
    int dest[CPW * ROUTE_SLOTS], info[CPW * ROUTE_SLOTS];
    R pa[CPW * ROUTE_SLOTS], pb[CPW * ROUTE_SLOTS];
    int cnt[CPW * ROUTE_SLOTS], prun[CPW * ROUTE_SLOTS];
    int ev[CPW * ROUTE_SLOTS][EV_MAX];           // collision events per (cell, route slot), flushed at the end
};

template <typename R>
struct CountJob {
    DevConst<R> k;
    PView<R> p;
    int isp;
    const StepClock *clk;
    int *cnt, *pcnt;
    unsigned long long *mccr;
    int *condcharge;
    uint8_t *flag;
    R *stash;
    size_t stash_stride;
    int *err;
};

template <typename R, int KIND>
__device__ __forceinline__ void fused_count_body(const CountJob<R> &J, int bx) {
    const DevConst<R> &k = J.k;
    const PView<R> &p = J.p;
    const int isp = J.isp;
    const StepClock *clk = J.clk;
    int *cnt = J.cnt, *pcnt = J.pcnt, *condcharge = J.condcharge, *err = J.err;
    unsigned long long *mccr = J.mccr;
    uint8_t *flag = J.flag;
    R *stash = J.stash;
    const size_t stash_stride = J.stash_stride;

    __shared__ CountWave<R> waves[FWAVES];
    CountWave<R> &W = waves[threadIdx.x >> 6];
    const uint32_t step = clk->step;
    const int lane = threadIdx.x & 63;
    const int w0 = __builtin_amdgcn_readfirstlane((int)((bx * blockDim.x + threadIdx.x) >> 6)) * CPW;
    const unsigned long long lt = lanes_below(lane);
    // cells of this wave (-1 past the last cell) and their particle ranges
    int cell[CPW], lo[CPW + 1];
#pragma unroll
    for (int q = 0; q < CPW; ++q) {
        const int wq = w0 + q;
        cell[q] = wq < k.Csize ? (wq / k.ncy) * k.ngy + (wq % k.ncy) : -1;
    }
    {
        int a0 = 0, a1 = 0;
        if (cell[0] >= 0) cell_range(p, cell[0], a0, a1);
        lo[0] = a0;
        lo[1] = a1;
#pragma unroll
        for (int q = 1; q < CPW; ++q) {
            int b0 = lo[q], b1 = lo[q];
            if (cell[q] >= 0) cell_range(p, cell[q], b0, b1);
            lo[q + 1] = max(b1, lo[q]);
        }
    }
    const int hi = lo[CPW];
    // per-wave tables
    for (int t = lane; t < CPW * ROUTE_SLOTS; t += 64) {
        const int q = t / ROUTE_SLOTS;
        int d = -1, inf = 0;
        R a = 0, b = 0;
        if (cell[q] >= 0) {
            const Route r = k.route[cell[q] * ROUTE_SLOTS + (t % ROUTE_SLOTS)];
            d = r.dest;
            inf = r.info;
            if (d >= 0) {
                if (KIND == KIND_AR_ELECTRON) { a = k.P1[d]; b = k.P2[d]; }
                else b = k.Pion[d];
            }
        }
        W.dest[t] = d; W.info[t] = inf; W.pa[t] = a; W.pb[t] = b;
        W.cnt[t] = 0; W.prun[t] = 0;
#pragma unroll
        for (int r = 0; r < EV_MAX; ++r) W.ev[t][r] = 0;
    }
    if (lane < CPW * 8) {
        const int q = lane >> 3, j = lane & 7;
        R v = 0;
        if (cell[q] >= 0 && lo[q + 1] > lo[q]) {
            const int c = cell[q];
            const int at = c + ((j & 3) == 1 ? 1 : (j & 3) == 2 ? k.ngy : (j & 3) == 3 ? k.ngy + 1 : 0);
            v = j < 4 ? k.Ex[at] : k.Ey[at];
        }
        W.e[q][j] = v;
    }
    wave_lds_fence();
    const R As = k.Ascale[isp];
    int qn = 0;   // wave-uniform queue length

    // collide queue entries [q0, q0 + n), one per lane
    auto collide_queue = [&](int q0, int n) {
        if (lane < n) {
            const int e = q0 + lane;
            const int meta = W.qmeta[e];
            const int code = meta & 15, kq = (meta >> 4) & 15, jc = meta >> 8;
            const int ent = kq * ROUTE_SLOTS + code;
            const int dest = W.dest[ent], info = W.info[ent];
            const R pa = W.pa[ent];
            R vx = W.qvx[e], vy = W.qvy[e], vz = W.qvz[e];
            R evx, evy, evz, ivx, ivy, ivz;
            const int ev = collide<R, KIND>(k, dest, pa, W.qid[e], step, jc, vx, vy, vz, evx, evy, evz, ivx, ivy, ivz);
            if (ev >= 0 && ev < EV_MAX) atomicAdd(&W.ev[ent][ev], 1);
            else if (ev >= 0) atomicAdd(&mccr[(size_t)dest * k.TnRct + ev], 1ull);
            const bool prod = (KIND == KIND_AR_ELECTRON) && (ev == 3 || ev == 4);
            const int kk = W.qkk[e];
            p.x[kk] = W.qx[e]; p.y[kk] = W.qy[e]; p.vx[kk] = vx; p.vy[kk] = vy; p.vz[kk] = vz;
            flag[kk] = (uint8_t)(code | FLAG_DONE | (prod ? FLAG_PROD : 0));
            if (prod) {
                atomicAdd(&W.prun[kq * ROUTE_SLOTS + (info & 15)], 1);
                stash[kk] = evx; stash[stash_stride + kk] = evy; stash[2 * stash_stride + kk] = evz;
                stash[3 * stash_stride + kk] = ivx; stash[4 * stash_stride + kk] = ivy;
                stash[5 * stash_stride + kk] = ivz;
            }
        }
    };

    // the walk, 64 particles at a time; lanes past the end reload the last
    // particle (never used)
    R x = 0, y = 0, vx = 0, vy = 0, vz = 0;
    uint32_t id = 0;
    if (lo[0] < hi) {
        const int i = min(lo[0] + lane, hi - 1);
        x = p.x[i]; y = p.y[i]; vx = p.vx[i]; vy = p.vy[i]; vz = p.vz[i]; id = p.id[i];
    }
    for (int b = lo[0]; b < hi; b += 64) {
        R nx = 0, ny = 0, nvx = 0, nvy = 0, nvz = 0;
        uint32_t nid = 0;
        if (b + 64 < hi) {
            const int i = min(b + 64 + lane, hi - 1);
            nx = p.x[i]; ny = p.y[i]; nvx = p.vx[i]; nvy = p.vy[i]; nvz = p.vz[i]; nid = p.id[i];
        }
        const int i = b + lane;
        const bool act = i < hi;
        int kq = 0;
#pragma unroll
        for (int q = 1; q < CPW; ++q) kq += i >= lo[q] ? 1 : 0;
        // a stored in-cell offset is in [0,1), or just outside after a move
        // of more than one cell (not handled, as in the reference); anything
        // else is a corrupt particle record
        if (act && !(x > (R)-4 && x < (R)4 && y > (R)-4 && y < (R)4)) atomicOr(err, 2);
        const int code = push_particle<R>(W.e[kq][0], W.e[kq][1], W.e[kq][2], W.e[kq][3], W.e[kq][4], W.e[kq][5],
                                          W.e[kq][6], W.e[kq][7], As, k.dt_dx, k.dt_dy, x, y, vx, vy);
        const int ent = kq * ROUTE_SLOTS + code;
        const int dest = W.dest[ent], info = W.info[ent];
        int slot = -1, jc = -1;
        if (act) {
            if (dest >= 0) {
                apply_reflect<R>(info, x, y, vx, vy);
                slot = info & 15;
                jc = null_test<R, KIND>(k, W.pb[ent], id, step);
            } else {
                int c = cell[0];
#pragma unroll
                for (int q = 1; q < CPW; ++q) c = kq == q ? cell[q] : c;
                wall_effects<R>(k, Route{dest, info}, isp, c, code, x, y, vx, vy, condcharge);
            }
        }
        // survivors per (cell, slot): the lanes of one cell are a contiguous run
        const int kfirst = __builtin_amdgcn_readfirstlane(kq);
        const int klast = __builtin_amdgcn_readfirstlane(__shfl(kq, min(63, hi - 1 - b), 64));
        for (int q = kfirst; q <= klast; ++q) {
            const unsigned long long mk = __ballot(act && kq == q);
#pragma unroll
            for (int s2 = 0; s2 < ROUTE_SLOTS; ++s2) {
                const int n = __popcll(__ballot(slot == s2) & mk);
                if (n && lane == 0) W.cnt[q * ROUTE_SLOTS + s2] += n;
            }
        }
        const unsigned long long cm = __ballot(jc >= 0);
        if (jc >= 0) {
            const int pos = qn + __popcll(cm & lt);
            W.qx[pos] = x; W.qy[pos] = y; W.qvx[pos] = vx; W.qvy[pos] = vy; W.qvz[pos] = vz;
            W.qid[pos] = id;
            W.qkk[pos] = i;
            W.qmeta[pos] = (jc << 8) | (kq << 4) | code;
        } else if (act) {
            flag[i] = (uint8_t)code;
        }
        qn += __popcll(cm);
        wave_lds_fence();
        if (qn >= 64) {
            qn -= 64;
            collide_queue(qn, 64);   // the newest 64; older entries stay in place
            wave_lds_fence();
        }
        x = nx; y = ny; vx = nvx; vy = nvy; vz = nvz; id = nid;
    }
    if (qn > 0) collide_queue(0, qn);
    wave_lds_fence();
    // counts of this wave's cells, event counts per (cell, route slot, reaction)
    for (int t = lane; t < CPW * ROUTE_SLOTS; t += 64) {
        const int q = t / ROUTE_SLOTS;
        if (cell[q] >= 0) {
            const int o = cell[q] * ROUTE_SLOTS + (t % ROUTE_SLOTS);
            cnt[o] = W.cnt[t];
            if (pcnt) pcnt[o] = W.prun[t];
        }
    }
    for (int e = lane; e < CPW * ROUTE_SLOTS * EV_MAX; e += 64) {
        const int t = e / EV_MAX, r = e % EV_MAX;
        const int n = W.ev[t][r];
        if (n) atomicAdd(&mccr[(size_t)W.dest[t] * k.TnRct + r], (unsigned long long)n);
    }
}

template <typename R, int KIND>
__global__ __launch_bounds__(256) void k_fused_count(CountJob<R> J) { fused_count_body<R, KIND>(J, blockIdx.x); }
template <typename R, int KIND>
__global__ __launch_bounds__(256) void k_fused_count_m(const CountJob<R> *J, const uint32_t *map) {
    int lb;
    const CountJob<R> &a = J[multi_job(map, lb)];
    fused_count_body<R, KIND>(a, lb);
}


// Pass 2: write each particle and product to its final slot.  For
// electrons, ionisation products go to both the electron and the ion arrays.
// Particles that did not collide are pushed again (same bits as pass 1);
// collided ones already hold their final state.  The flag bytes are staged
// with the particles.
template <typename R>
struct ScatterJob {
    DevConst<R> k;
    PView<R> p, q, qion;
    int isp;
    const uint8_t *flag;
    const R *stash;
    size_t stash_stride;
    const int *base, *pbase_e, *pbase_i, *pid_off;
    const uint32_t *next_id;
    int cap, cap_ion;
    int *err;
};

template <typename R, int KIND>
__device__ __forceinline__ void fused_scatter_body(const ScatterJob<R> &J, int bx) {
    const DevConst<R> &k = J.k;
    const PView<R> &p = J.p, &q = J.q, &qion = J.qion;
    const int isp = J.isp, cap = J.cap, cap_ion = J.cap_ion;
    const uint8_t *flag = J.flag;
    const R *stash = J.stash;
    const size_t stash_stride = J.stash_stride;
    const int *base = J.base, *pbase_e = J.pbase_e, *pbase_i = J.pbase_i, *pid_off = J.pid_off;
    const uint32_t *next_id = J.next_id;
    int *err = J.err;

    __shared__ FStage<R, SCHUNK> stage_all[FWAVES];
    __shared__ uint8_t flag_all[FWAVES][SCHUNK];
    int lane;
    int c = wave_cell_b(k.ncy, k.ngy, k.Csize, lane, bx);
    if (c < 0) return;
    int k0, k1;
    cell_range(p, c, k0, k1);
    if (k0 == k1) return;
    FStage<R, SCHUNK> &S = stage_all[threadIdx.x >> 6];
    uint8_t *F = flag_all[threadIdx.x >> 6];
    CellE<R> e;
    e.load(k.Ex, k.Ey, c, k.ngy);
    const Route myr = lane < ROUTE_SLOTS ? k.route[c * ROUTE_SLOTS + lane] : Route{-1, 0};
    // per-slot output bases of this cell, held by lane = slot
    const int my_base = lane < ROUTE_SLOTS ? base[c * ROUTE_SLOTS + lane] : 0;
    const R As = k.Ascale[isp];
    const unsigned long long lt = lanes_below(lane);
    int run[ROUTE_SLOTS], prun[ROUTE_SLOTS];
#pragma unroll
    for (int s = 0; s < ROUTE_SLOTS; ++s) { run[s] = 0; prun[s] = 0; }
    for (int b0 = k0; b0 < k1; b0 += SCHUNK) {
        const int n = min(SCHUNK, k1 - b0);
        {
            uint8_t tf[SCHUNK / 64];
#pragma unroll
            for (int j = 0; j < SCHUNK / 64; ++j) tf[j] = flag[b0 + min(j * 64 + lane, n - 1)];
            stage_chunk<R, SCHUNK>(S, p, b0, n, lane);
#pragma unroll
            for (int j = 0; j < SCHUNK / 64; ++j) F[j * 64 + lane] = tf[j];
        }
        for (int j = 0; j * 64 < n; ++j) {
            const int i = j * 64 + lane;
            const int f = F[i];
            const int code = f & FLAG_CODE;
            const int dest = __shfl(myr.dest, code, 64);
            const int info = __shfl(myr.info, code, 64);
            R x = S.x[i], y = S.y[i], vx = S.vx[i], vy = S.vy[i];
            if (!(f & FLAG_DONE)) {
                push_particle<R>(e.ews, e.ewn, e.ees, e.een, e.yws, e.ywn, e.yes, e.yen, As, k.dt_dx, k.dt_dy, x, y,
                                 vx, vy);
                if (dest >= 0) apply_reflect<R>(info, x, y, vx, vy);
            }
            const int slot = (i < n && dest >= 0) ? (info & 15) : -1;
            const bool prod = (KIND == KIND_AR_ELECTRON) && slot >= 0 && (f & FLAG_PROD);
            int pos = -1, prank = 0;
#pragma unroll
            for (int s = 0; s < ROUTE_SLOTS; ++s) {
                const unsigned long long m = __ballot(slot == s);
                if (slot == s) pos = run[s] + __popcll(m & lt);
                run[s] += __popcll(m);
            }
            // product ranks: most groups of 64 hold no ionisation pair
            if (KIND == KIND_AR_ELECTRON && __ballot(prod)) {
#pragma unroll
                for (int s = 0; s < ROUTE_SLOTS; ++s) {
                    const unsigned long long pm = __ballot(prod && slot == s);
                    if (prod && slot == s) prank = prun[s] + __popcll(pm & lt);
                    prun[s] += __popcll(pm);
                }
            }
            const int sbase = __shfl(my_base, slot < 0 ? 0 : slot, 64);
            if (slot >= 0) {
                pos += sbase;
                if (pos < cap) {
                    q.x[pos] = x; q.y[pos] = y; q.vx[pos] = vx; q.vy[pos] = vy; q.vz[pos] = S.vz[i]; q.id[pos] = S.id[i];
                } else {
                    atomicOr(err, 1);
                }
            }
            if (prod) {
                const int kk = b0 + i;
                const int ent = c * ROUTE_SLOTS + slot;
                const int off = pid_off[ent] + prank;
                const int ppos = pbase_e[ent] + prank;
                if (ppos < cap) {
                    q.x[ppos] = x; q.y[ppos] = y;
                    q.vx[ppos] = stash[kk]; q.vy[ppos] = stash[stash_stride + kk]; q.vz[ppos] = stash[2 * stash_stride + kk];
                    q.id[ppos] = next_id[0] + (uint32_t)off;
                } else {
                    atomicOr(err, 1);
                }
                const int ipos = pbase_i[ent] + prank;
                if (ipos < cap_ion) {
                    qion.x[ipos] = x; qion.y[ipos] = y; qion.vx[ipos] = stash[3 * stash_stride + kk];
                    qion.vy[ipos] = stash[4 * stash_stride + kk]; qion.vz[ipos] = stash[5 * stash_stride + kk];
                    qion.id[ipos] = next_id[1] + (uint32_t)off;
                } else {
                    atomicOr(err, 1);
                }
            }
        }
    }
}

template <typename R, int KIND>
__global__ __launch_bounds__(256) void k_fused_scatter(ScatterJob<R> J) { fused_scatter_body<R, KIND>(J, blockIdx.x); }
template <typename R, int KIND>
__global__ __launch_bounds__(256) void k_fused_scatter_m(const ScatterJob<R> *J, const uint32_t *map) {
    int lb;
    const ScatterJob<R> &a = J[multi_job(map, lb)];
    fused_scatter_body<R, KIND>(a, lb);
}


// ============================================================ deposit
// CIC weights summed per cell with a wave reduction (src/cuda_Deposit.cu:289-317);
// blockIdx.y = species.  A lane's particles of a cell are loaded at once
// (up to 4 per lane per chunk) and summed in the same order as a plain
// 64-strided walk.
template <typename R>
struct DepCellsJob {
    DevConst<R> k;
    PSet<R> ps;
    R *part;
    int gx;   // blocks along x of the job's own grid
};

template <typename R>
__device__ __forceinline__ void deposit_cells_body(const DepCellsJob<R> &J, int bx, int by) {
    const DevConst<R> &k = J.k;
    const PSet<R> &ps = J.ps;
    R *part = J.part;

    int lane;
    int c = wave_cell_b(k.ncy, k.ngy, k.Csize, lane, bx);
    if (c < 0) return;
    const PView<R> &p = ps.s[by];
    R *pp = part + (size_t)by * k.Gsize * 4;
    int k0, k1;
    cell_range(p, c, k0, k1);
    R WS = 0, WN = 0, ES = 0, EN = 0;
    for (int b0 = k0; b0 < k1; b0 += 256) {
        const int n = min(256, k1 - b0);
        R lx[4], ly[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int i = b0 + min(j * 64 + lane, n - 1);
            lx[j] = p.x[i];
            ly[j] = p.y[i];
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            if (j * 64 + lane < n) {
                WS += (1 - lx[j]) * (1 - ly[j]);
                WN += (1 - lx[j]) * ly[j];
                ES += lx[j] * (1 - ly[j]);
                EN += lx[j] * ly[j];
            }
        }
    }
    WS = wave_sum(WS); WN = wave_sum(WN); ES = wave_sum(ES); EN = wave_sum(EN);
    if (lane == 0) {
        pp[(size_t)c * 4 + 0] = WS; pp[(size_t)c * 4 + 1] = WN;
        pp[(size_t)c * 4 + 2] = ES; pp[(size_t)c * 4 + 3] = EN;
    }
}

template <typename R>
__global__ __launch_bounds__(256) void k_deposit_cells(DepCellsJob<R> J) {
    deposit_cells_body<R>(J, blockIdx.x, blockIdx.y);
}
template <typename R>
__global__ __launch_bounds__(256) void k_deposit_cells_m(const DepCellsJob<R> *J, const uint32_t *map) {
    int lb;
    const DepCellsJob<R> &a = J[multi_job(map, lb)];
    deposit_cells_body<R>(a, lb % a.gx, lb / a.gx);
}


// Node density from the four adjacent cells, boundary factors
// (src/cuda_Deposit.cu:277-288), then charge density, source and PCG rhs
// (src/cuda_Deposit.cu:30-104).  part holds nsp x Gsize x 4.  This step's
// dielectric hits (fixed point, see sigma_fixed) are first added to the
// per-species surface charge and cleared.
// With ex.on (fused step) the same thread also adds this step's node fields
// to the FP64 profile sums (src/cuda_Diagnostic.cu:505-547) and thread 0
// advances the particle id counters past this step's ionisation products.
template <typename R>
struct DepNodesJob {
    NodeConst<R> n;
    const R *part;
    R *den, *sigma_sp, *Sigma, *Source, *b;
    DepositExtras<R> ex;
};

template <typename R>
__device__ __forceinline__ void deposit_nodes_body(const DepNodesJob<R> &J, int bx) {
    const NodeConst<R> &n = J.n;
    const R *part = J.part;
    R *den = J.den, *sigma_sp = J.sigma_sp, *Sigma = J.Sigma, *Source = J.Source, *b = J.b;
    const DepositExtras<R> &ex = J.ex;

    int g = bx * blockDim.x + threadIdx.x;
    if (ex.on && g == 0) {
        const uint32_t add = (uint32_t)*ex.prod_total_end;
        for (int s = 0; s < n.nsp; ++s) ex.next_id[s] += add;
        // the next step: same arithmetic as the host's t = tstep * dt
        const uint32_t nstep = ex.clk->step + 1u;
        ex.clk->step = nstep;
        ex.clk->t = (double)nstep * ex.dt;
    }
    if (g >= n.Gsize) return;
    if (n.sigma_hits) {
        for (int i = 0; i < n.nsp; ++i) {
            const size_t at = (size_t)i * n.Gsize + g;
            const long long h = (long long)n.sigma_hits[at];
            if (h) {
                sigma_sp[at] += (R)((double)h * (1.0 / SIGMA_FIX));
                n.sigma_hits[at] = 0;
            }
        }
    }
    const int ngy = n.ngy, ngx = n.ngx;
    const int x = g / ngy, y = g % ngy;
    const bool c00 = x < ngx - 1 && y < ngy - 1;   // cell g
    const bool c01 = x < ngx - 1 && y > 0;         // cell g-1 (node is its north-west corner)
    const bool c10 = x > 0 && y < ngy - 1;         // cell g-ngy (south-east corner)
    const bool c11 = x > 0 && y > 0;               // cell g-ngy-1 (north-east corner)
    const int f = n.Face[g], bnd = n.Boundary[g];
    R ss = 0;
    for (int isp = 0; isp < n.nsp; ++isp) {
        const R *pp = part + (size_t)isp * n.Gsize * 4;
        R d = 0;   // cells in ascending index order
        if (c11) d += pp[(size_t)(g - ngy - 1) * 4 + 3];
        if (c10) d += pp[(size_t)(g - ngy) * 4 + 2];
        if (c01) d += pp[(size_t)(g - 1) * 4 + 1];
        if (c00) d += pp[(size_t)g * 4 + 0];
        if (f == F_LEFT || f == F_RIGHT || f == F_UP || f == F_DOWN) d *= 2;
        else if (f == F_UL || f == F_LL || f == F_UR || f == F_LR) d *= 4;
        if (bnd == NEUMANN) d *= 2;
        else if (bnd == DIRICHLET && f != F_NOFACE && n.CondID[g] != 0) d *= 2;
        den[(size_t)isp * n.Gsize + g] = d;
    }
    int nb = -1;
    if (x == 0) nb = g + ngy;
    else if (x == ngx - 1) nb = g - ngy;
    else if (y == 0) nb = g + 1;
    else if (y == ngy - 1) nb = g - 1;
    int nbnd = nb >= 0 ? n.Boundary[nb] : bnd;
    R fac = nb >= 0 ? (R)2 : (R)1;
    R SD = 0;
    if (nbnd == DIELECTRIC) {
        for (int i = 0; i < n.nsp; ++i) ss += sigma_sp[(size_t)i * n.Gsize + g] * n.q_density[i];
        Sigma[g] = fac * ss / n.Area[g];
        SD = ss;
    } else if (nbnd == CONDUCTOR) {
        for (int i = 0; i < n.nsp; ++i) sigma_sp[(size_t)i * n.Gsize + g] = 0;
        SD = 0;
    }
    ss = 0;
    for (int i = 0; i < n.nsp; ++i) ss += den[(size_t)i * n.Gsize + g] * n.q_density[i];
    R src = (R)((double)(ss + SD) / 8.8542e-12);
    Source[g] = src;
    int ai = n.A_idx[g];
    if (ai) b[ai - 1] = src;
    if (ex.on) {
        for (int i = 0; i < n.nsp; ++i) ex.sden[(size_t)i * n.Gsize + g] += (double)den[(size_t)i * n.Gsize + g];
        ex.sphi[g] += (double)ex.tot[g];
        ex.sex[g] += (double)ex.Ex[g];
        ex.sey[g] += (double)ex.Ey[g];
    }
}

template <typename R>
__global__ __launch_bounds__(256) void k_deposit_nodes(DepNodesJob<R> J) { deposit_nodes_body<R>(J, blockIdx.x); }
template <typename R>
__global__ __launch_bounds__(256) void k_deposit_nodes_m(const DepNodesJob<R> *J, const uint32_t *map) {
    int lb;
    const DepNodesJob<R> &a = J[multi_job(map, lb)];
    deposit_nodes_body<R>(a, lb);
}


// Fused step: node potential straight from the field solution, with the
// electrode voltages of this time evaluated per thread:
// phi = x[A_idx-1] (0 on fixed nodes), tot = sum_c V_c(t) lap_c + phi.
template <typename R>
struct PhiPotJob {
    NodeConst<R> n;
    VoltConst v;
    const StepClock *clk;
    const double *x;
    const R *lap;
    R *V, *phi, *tot;
};

template <typename R>
__device__ __forceinline__ void phi_potential_body(const PhiPotJob<R> &J, int bx) {
    const NodeConst<R> &n = J.n;
    const VoltConst &v = J.v;
    const StepClock *clk = J.clk;
    const double *x = J.x;
    const R *lap = J.lap;
    R *V = J.V, *phi = J.phi, *tot = J.tot;

    int g = bx * blockDim.x + threadIdx.x;
    if (g >= n.Gsize) return;
    const double t = clk->t;
    const int a = n.A_idx[g];
    const R ph = a ? (R)x[a - 1] : (R)0;
    phi[g] = ph;
    R tp = 0;
    for (int i = 0; i < n.CondNUMR; ++i) {
        const R vi = cond_voltage<R>(v, i, t);
        if (g == 0) V[i] = vi;
        tp += vi * lap[(size_t)i * n.Gsize + g];
    }
    tp += (R)1 * ph;
    tot[g] = tp;
}

template <typename R>
__global__ __launch_bounds__(256) void k_phi_potential(PhiPotJob<R> J) { phi_potential_body<R>(J, blockIdx.x); }
template <typename R>
__global__ __launch_bounds__(256) void k_phi_potential_m(const PhiPotJob<R> *J, const uint32_t *map) {
    int lb;
    const PhiPotJob<R> &a = J[multi_job(map, lb)];
    phi_potential_body<R>(a, lb);
}


// ============================================================ EEDF
// Electron energy histogram: E = Escale v^2 (the MCC's energy, in eV) in
// EEDF_BINS equal bins over [0, emax); energies at or above emax go to the
// last bin.  Per-block LDS counts, then one 64-bit add per bin: the counts
// are integers, so the result does not depend on the order.
template <typename R>
__global__ __launch_bounds__(256) void k_eedf(PView<R> p, int ncell_nodes, R escale, R emax, unsigned long long *hist) {
    __shared__ unsigned int h[EEDF_BINS];
    for (int b = threadIdx.x; b < EEDF_BINS; b += blockDim.x) h[b] = 0u;
    __syncthreads();
    const int np = min(p.cell_start[ncell_nodes], p.cap);
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < np; k += gridDim.x * blockDim.x) {
        const R vx = p.vx[k], vy = p.vy[k], vz = p.vz[k];
        const R e = escale * (vx * vx + vy * vy + vz * vz);
        int b = (int)(e / emax * (R)EEDF_BINS);
        b = b < 0 ? 0 : (b >= EEDF_BINS ? EEDF_BINS - 1 : b);
        atomicAdd(&h[b], 1u);
    }
    __syncthreads();
    for (int b = threadIdx.x; b < EEDF_BINS; b += blockDim.x)
        if (h[b]) atomicAdd(&hist[b], (unsigned long long)h[b]);
}

// ============================================================ E field
// Total potential = sum_c V_c phi_c + phi_Poisson (src/cuda_Field.cu:22-33)
template <typename R>
__global__ void k_potential(NodeConst<R> n, const R *lap, const R *V, const R *phi, R *tot) {
    int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n.Gsize) return;
    R tp = 0;
    for (int i = 0; i < n.CondNUMR; ++i) tp += V[i] * lap[(size_t)i * n.Gsize + g];
    tp += (R)1 * phi[g];
    tot[g] = tp;
}

template <typename R>
__device__ inline void eps_idx(int TID, int ngx, int ngy, int &ET, int &EL, int &ED, int &EC) {
    int x = TID / ngy, y = TID % ngy;
    ET = x * (ngy - 1) + y;
    EL = x ? ET - ngy + 1 : ET;
    ED = y ? ET - 1 : ET;
    EC = EL + ED - ET;
    int xo = (x == ngx - 1) ? ngy - 1 : 0;
    int yo = (y == ngy - 1) ? 1 : 0;
    ET -= (xo + yo); EL -= yo; ED -= xo;
}

// Surface charge of conductor/Dirichlet faces (src/cuda_Field.cu:500-558,
// from the Poisson potential and the last deposit) and node E
// (src/cuda_Field.cu:698-846, from the total potential).
template <typename R>
struct EfieldJob {
    NodeConst<R> n;
    const R *phi, *tot, *den;
    R *Sigma, *Ex, *Ey;
};

template <typename R>
__device__ __forceinline__ void efield_body(const EfieldJob<R> &J, int bx) {
    const NodeConst<R> &n = J.n;
    const R *phi = J.phi, *tot = J.tot, *den = J.den;
    R *Sigma = J.Sigma, *Ex = J.Ex, *Ey = J.Ey;

    int TID = bx * blockDim.x + threadIdx.x;
    if (TID >= n.Gsize) return;
    const int ngx = n.ngx, ngy = n.ngy;
    const int x = TID / ngy, y = TID % ngy;
    int ET, EL, ED, EC;
    eps_idx<R>(TID, ngx, ngy, ET, EL, ED, EC);
    const int pl = x ? TID - ngy : TID, pr = (x == ngx - 1) ? TID : TID + ngy;
    const int pd = y ? TID - 1 : TID, pu = (y == ngy - 1) ? TID : TID + 1;
    const int b = n.Boundary[TID], f = n.Face[TID];
    const R *e = n.eps;
    const R hdx = n.hdx, hdy = n.hdy, idx = n.idx, idy = n.idy, idxy = n.idxy;
    // --- Cond_Sigma (Poisson part only)
    R sig = Sigma[TID];
    if ((b == CONDUCTOR || b == DIRICHLET) && f != F_NOFACE) {
        R ss = 0;
        for (int i = 0; i < n.nsp; ++i) ss += den[TID + i * n.Gsize] * n.q_density[i];
        const R P = phi[TID];
        switch (f) {
        case F_UP: sig = hdy * (e[ET] + e[EL]) * (P - phi[pu]) - hdx * ss; break;
        case F_DOWN: sig = hdy * (e[ED] + e[EC]) * (P - phi[pd]) - hdx * ss; break;
        case F_LEFT: sig = hdx * (e[EL] + e[EC]) * (P - phi[pl]) - hdy * ss; break;
        case F_RIGHT: sig = hdx * (e[ET] + e[ED]) * (P - phi[pr]) - hdy * ss; break;
        case F_UL: sig = hdy * (e[ET] + e[EL]) * (P - phi[pu]) + hdx * (e[EL] + e[EC]) * (P - phi[pl]) - idxy * ss; break;
        case F_UR: sig = hdy * (e[ET] + e[EL]) * (P - phi[pu]) + hdx * (e[ET] + e[ED]) * (P - phi[pr]) - idxy * ss; break;
        case F_LL: sig = hdy * (e[ED] + e[EC]) * (P - phi[pd]) + hdx * (e[EL] + e[EC]) * (P - phi[pl]) - idxy * ss; break;
        case F_LR: sig = hdy * (e[ED] + e[EC]) * (P - phi[pd]) + hdx * (e[ET] + e[ED]) * (P - phi[pr]) - idxy * ss; break;
        default: break;
        }
        Sigma[TID] = sig;
    }
    // --- VtoEfield
    const R *Phi = tot;
    const R P = Phi[TID];
    R ex = Ex[TID], ey = Ey[TID];
    const R hxy = n.hxy, hyx = n.hyx;
    const bool wall = (b == CONDUCTOR || b == NEUMANN || b == DIRICHLET);
    const bool die = (b == DIELECTRIC || b >= 100);
    const bool hard = (b == CONDUCTOR || b == DIRICHLET || b == DIELECTRIC);
    if (f == F_LEFT) {
        if (wall) { ex = -2 * sig / (e[EL] + e[EC]); ey = 0; }
        else if (die) { ex = idx * (Phi[pl] - P); ey = hdy * (Phi[pd] - Phi[pu]); }
        else {
            ex = (R)0.5 * (e[ET] + e[ED]) * (Phi[pr] - P) * idx;
            ex -= hxy * e[ET] * (P - Phi[pu]) * idy;
            ex -= hxy * e[ED] * (P - Phi[pd]) * idy;
            ex /= (R)-0.5 * (e[EL] + e[EC]);
            ey = hdy * (Phi[pd] - Phi[pu]);
        }
    } else if (f == F_RIGHT) {
        if (wall) { ex = 2 * sig / (e[ET] + e[ED]); ey = 0; }
        else if (die) { ex = idx * (P - Phi[pr]); ey = hdy * (Phi[pd] - Phi[pu]); }
        else {
            ex = (R)0.5 * (e[EL] + e[EC]) * (Phi[pl] - P) * idx;
            ex -= hxy * e[EL] * (P - Phi[pu]) * idy;
            ex -= hxy * e[EC] * (P - Phi[pd]) * idy;
            ex /= (R)0.5 * (e[ET] + e[ED]);
            ey = hdy * (Phi[pd] - Phi[pu]);
        }
    } else if (f == F_UP) {
        if (wall) { ex = 0; ey = 2 * sig / (e[ET] + e[EL]); }
        else if (die) { ex = hdx * (Phi[pl] - Phi[pr]); ey = idy * (P - Phi[pu]); }
        else {
            ex = hdx * (Phi[pl] - Phi[pr]);
            ey = (R)0.5 * (e[ED] + e[EC]) * (Phi[pd] - P) * idy;
            ey -= hyx * e[ED] * (P - Phi[pr]) * idx;
            ey -= hyx * e[EC] * (P - Phi[pl]) * idx;
            ey /= (R)0.5 * (e[ET] + e[EL]);
        }
    } else if (f == F_DOWN) {
        if (wall) { ex = 0; ey = -2 * sig / (e[ED] + e[EC]); }
        else if (die) { ex = hdx * (Phi[pl] - Phi[pr]); ey = idy * (Phi[pd] - P); }
        else {
            ex = hdx * (Phi[pl] - Phi[pr]);
            ey = (R)0.5 * (e[ET] + e[EL]) * (Phi[pu] - P) * idy;
            ey -= hyx * e[ET] * (P - Phi[pr]) * idx;
            ey -= hyx * e[EL] * (P - Phi[pl]) * idx;
            ey /= (R)-0.5 * (e[ED] + e[EC]);
        }
    } else if (f == F_UL) {
        if (hard) { ex = idx * (Phi[pl] - P); ey = idy * (P - Phi[pu]); }
        else { ex = hdx * (Phi[pl] - Phi[pr]); ey = hdy * (Phi[pd] - Phi[pu]); }
    } else if (f == F_UR) {
        if (hard) { ex = idx * (P - Phi[pr]); ey = idy * (P - Phi[pu]); }
        else { ex = hdx * (Phi[pl] - Phi[pr]); ey = hdy * (Phi[pd] - Phi[pu]); }
    } else if (f == F_LL) {
        if (hard) { ex = idx * (Phi[pl] - P); ey = idy * (Phi[pd] - P); }
        else { ex = hdx * (Phi[pl] - Phi[pr]); ey = hdy * (Phi[pd] - Phi[pu]); }
    } else if (f == F_LR) {
        if (hard) { ex = idx * (P - Phi[pr]); ey = idy * (Phi[pd] - P); }
        else { ex = hdx * (Phi[pl] - Phi[pr]); ey = hdy * (Phi[pd] - Phi[pu]); }
    } else if (b == NEUMANN) {
        if (x == 0 || x == ngx - 1) { ex = 0; ey = hdy * (Phi[pd] - Phi[pu]); }
        else if (y == 0 || y == ngy - 1) { ex = hdx * (Phi[pl] - Phi[pr]); ey = 0; }
    } else {
        ex = hdx * (Phi[pl] - Phi[pr]);
        ey = hdy * (Phi[pd] - Phi[pu]);
    }
    Ex[TID] = ex;
    Ey[TID] = ey;
}

template <typename R>
__global__ __launch_bounds__(256) void k_efield(EfieldJob<R> J) { efield_body<R>(J, blockIdx.x); }
template <typename R>
__global__ __launch_bounds__(256) void k_efield_m(const EfieldJob<R> *J, const uint32_t *map) {
    int lb;
    const EfieldJob<R> &a = J[multi_job(map, lb)];
    efield_body<R>(a, lb);
}


// Conductor voltages at t = step * dt: V = sum DC + AC sin(w t + phase)
template <typename R>
__global__ void k_voltages(VoltConst v, double t, R *V) {
    int i = threadIdx.x;
    if (i >= v.ncond) return;
    V[i] = cond_voltage<R>(v, i, t);
}

// profile accumulators in FP64 (src/cuda_Diagnostic.cu:505-547)
template <typename R>
__global__ void k_accumulate(int Gsize, int nsp, const R *den, const R *tot, const R *Ex, const R *Ey, double *sden,
                             double *sphi, double *sex, double *sey) {
    int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= Gsize) return;
    for (int i = 0; i < nsp; ++i) sden[(size_t)i * Gsize + g] += (double)den[(size_t)i * Gsize + g];
    sphi[g] += (double)tot[g];
    sex[g] += (double)Ex[g];
    sey[g] += (double)Ey[g];
}

// ============================================================ small helpers
// new_total[d] = particles already in d + products appended to d
__global__ void k_cell_counts(int Gsize, const int *cell_start, const int *pcnt, int *total) {
    int d = blockIdx.x * blockDim.x + threadIdx.x;
    if (d >= Gsize) return;
    total[d] = cell_start[d + 1] - cell_start[d] + (pcnt ? pcnt[d] : 0);
}

// v into n ints (a kernel, not a memset: usable where the launches are captured)
__global__ void k_fill_int(int *p, int v, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// next_id[s] += number of products created this step (same count for the
// electron and the ion array: one of each per ionisation)
__global__ void k_bump_ids(uint32_t *next_id, const int *prod_total_end, int nsp) {
    if (threadIdx.x == 0) {
        uint32_t add = (uint32_t)*prod_total_end;
        for (int s = 0; s < nsp; ++s) next_id[s] += add;
    }
}

template <typename R>
__global__ void k_rhs_to_fp64(int N, const R *b, double *b64) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < N) b64[i] = (double)b[i];
}

template <typename R>
__global__ void k_x_to_phi(int Gsize, const int *A_idx, const double *x, R *phi) {
    int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= Gsize) return;
    int a = A_idx[g];
    phi[g] = a ? (R)x[a - 1] : (R)0;
}

// ============================================================ launchers
static inline int cell_blocks(int Csize) { return (Csize + 3) / 4; }
// count pass: CPW cells per wave, FWAVES waves per block
static inline int count_blocks(int Csize) { return (Csize + CPW * FWAVES - 1) / (CPW * FWAVES); }
static inline int node_blocks(int n) { return (n + 255) / 256; }

template <typename R>
void launch_push(const DevConst<R> &k, const PView<R> &p, int isp, uint8_t *tag, hipStream_t s) {
    hipLaunchKernelGGL(k_push<R>, dim3(cell_blocks(k.Csize)), dim3(256), 0, s, k, p, isp, tag);
}
template <typename R>
void launch_sort_count(const DevConst<R> &k, const PView<R> &p, int isp, const uint8_t *tag, int *key, int *cnt,
                       int *cc, hipStream_t s) {
    hipLaunchKernelGGL(k_sort_count<R>, dim3(cell_blocks(k.Csize)), dim3(256), 0, s, k, p, isp, tag, key, cnt, cc);
}
template <typename R>
void launch_push_count(const DevConst<R> &k, const StagedSet<R> &set, int nsp, int *cc, hipStream_t s) {
    hipLaunchKernelGGL(k_push_count<R>, dim3(cell_blocks(k.Csize), nsp), dim3(256), 0, s, k, set, cc);
}
template <typename R>
void launch_sort_scatter_all(const DevConst<R> &k, const StagedSet<R> &set, int nsp, int *err, hipStream_t s) {
    hipLaunchKernelGGL(k_sort_scatter_all<R>, dim3(cell_blocks(k.Csize), nsp), dim3(256), 0, s, k, set, err);
}
template <typename R>
void launch_sort_scatter(const DevConst<R> &k, const PView<R> &p, const PView<R> &q, const uint8_t *tag,
                         const int *key, const int *base, int cap, int *err, hipStream_t s) {
    hipLaunchKernelGGL(k_sort_scatter<R>, dim3(cell_blocks(k.Csize)), dim3(256), 0, s, k, p, q, tag, key, base, cap, err);
}
template <typename R>
void launch_mcc_collide(const DevConst<R> &k, const PView<R> &p, int isp, uint32_t step, unsigned long long *mccr,
                        int *pcnt, int *prank, R *stage, hipStream_t s) {
    hipLaunchKernelGGL(k_mcc_collide<R>, dim3(cell_blocks(k.Csize)), dim3(256), 0, s, k, p, isp, step, mccr, pcnt,
                       prank, stage);
}
template <typename R>
void launch_mcc_insert(const DevConst<R> &k, const PView<R> &src, const PView<R> &dst, const int *new_start,
                       const int *prod_start, const PView<R> &parent, const int *prank, const R *stage, int stage_off,
                       const uint32_t *next_id, int cap, int *err, hipStream_t s) {
    hipLaunchKernelGGL(k_mcc_insert<R>, dim3(cell_blocks(k.Csize)), dim3(256), 0, s, k, src, dst, new_start,
                       prod_start, (const int *)nullptr, parent, prank, stage, stage_off, next_id, cap, err);
}
template <typename R>
void launch_fused_count(const DevConst<R> &k, const PView<R> &p, int isp, const StepClock *clk, int *cnt, int *pcnt,
                        unsigned long long *mccr, int *cc, uint8_t *flag, R *stash, size_t stash_stride, int *err,
                        hipStream_t s) {
    const CountJob<R> J{k, p, isp, clk, cnt, pcnt, mccr, cc, flag, stash, stash_stride, err};
    const int nb = count_blocks(k.Csize);
    if (Recorder *r = launch_recorder()) {
        r->add(isp == 0 ? &launch_m<CountJob<R>, k_fused_count_m<R, KIND_AR_ELECTRON>, 256>
                        : &launch_m<CountJob<R>, k_fused_count_m<R, KIND_AR_ION>, 256>,
               J, nb);
        return;
    }
    if (isp == 0)
        hipLaunchKernelGGL((k_fused_count<R, KIND_AR_ELECTRON>), dim3(nb), dim3(256), 0, s, J);
    else
        hipLaunchKernelGGL((k_fused_count<R, KIND_AR_ION>), dim3(nb), dim3(256), 0, s, J);
}
template <typename R>
void launch_fused_scatter(const DevConst<R> &k, const PView<R> &p, const PView<R> &q, const PView<R> &qion, int isp,
                          const uint8_t *flag, const R *stash, size_t stash_stride, const int *base,
                          const int *pbase_e, const int *pbase_i, const int *pid_off, const uint32_t *next_id,
                          int cap, int cap_ion, int *err, hipStream_t s) {
    const ScatterJob<R> J{k, p, q, qion, isp, flag, stash, stash_stride, base, pbase_e, pbase_i, pid_off, next_id,
                          cap, cap_ion, err};
    const int nb = cell_blocks(k.Csize);
    if (Recorder *r = launch_recorder()) {
        r->add(isp == 0 ? &launch_m<ScatterJob<R>, k_fused_scatter_m<R, KIND_AR_ELECTRON>, 256>
                        : &launch_m<ScatterJob<R>, k_fused_scatter_m<R, KIND_AR_ION>, 256>,
               J, nb);
        return;
    }
    if (isp == 0)
        hipLaunchKernelGGL((k_fused_scatter<R, KIND_AR_ELECTRON>), dim3(nb), dim3(256), 0, s, J);
    else
        hipLaunchKernelGGL((k_fused_scatter<R, KIND_AR_ION>), dim3(nb), dim3(256), 0, s, J);
}
template <typename R>
void launch_deposit(const DevConst<R> &k, const NodeConst<R> &n, const PView<R> *ps, R *part, R *den, R *sigma_sp,
                    R *Sigma, R *Source, R *b, const DepositExtras<R> *ex, hipStream_t s) {
    PSet<R> set{};
    for (int isp = 0; isp < n.nsp; ++isp) set.s[isp] = ps[isp];
    DepositExtras<R> e{};
    if (ex) e = *ex;
    const DepCellsJob<R> Jc{k, set, part, cell_blocks(k.Csize)};
    const DepNodesJob<R> Jn{n, (const R *)part, den, sigma_sp, Sigma, Source, b, e};
    if (Recorder *r = launch_recorder()) {
        r->add(&launch_m<DepCellsJob<R>, k_deposit_cells_m<R>, 256>, Jc, Jc.gx * n.nsp);
        r->add(&launch_m<DepNodesJob<R>, k_deposit_nodes_m<R>, 256>, Jn, node_blocks(n.Gsize));
        return;
    }
    hipLaunchKernelGGL(k_deposit_cells<R>, dim3(Jc.gx, n.nsp), dim3(256), 0, s, Jc);
    hipLaunchKernelGGL(k_deposit_nodes<R>, dim3(node_blocks(n.Gsize)), dim3(256), 0, s, Jn);
}
template <typename R>
void launch_eedf(const PView<R> &p, int Gsize, R escale, R emax, unsigned long long *hist, hipStream_t s) {
    hipLaunchKernelGGL(k_eedf<R>, dim3(1024), dim3(256), 0, s, p, Gsize, escale, emax, hist);
}
template <typename R>
void launch_phi_potential(const NodeConst<R> &n, const VoltConst &v, const StepClock *clk, const double *x,
                          const R *lap, R *V, R *phi, R *tot, hipStream_t s) {
    const PhiPotJob<R> J{n, v, clk, x, lap, V, phi, tot};
    if (Recorder *r = launch_recorder()) {
        r->add(&launch_m<PhiPotJob<R>, k_phi_potential_m<R>, 256>, J, node_blocks(n.Gsize));
        return;
    }
    hipLaunchKernelGGL(k_phi_potential<R>, dim3(node_blocks(n.Gsize)), dim3(256), 0, s, J);
}
template <typename R>
void launch_efield_only(const NodeConst<R> &n, const R *phi, const R *tot, const R *den, R *Sigma, R *Ex, R *Ey,
                        hipStream_t s) {
    const EfieldJob<R> J{n, phi, tot, den, Sigma, Ex, Ey};
    if (Recorder *r = launch_recorder()) {
        r->add(&launch_m<EfieldJob<R>, k_efield_m<R>, 256>, J, node_blocks(n.Gsize));
        return;
    }
    hipLaunchKernelGGL(k_efield<R>, dim3(node_blocks(n.Gsize)), dim3(256), 0, s, J);
}
template <typename R>
void launch_efield(const NodeConst<R> &n, const VoltConst &v, double t, const R *lap, R *V, const R *phi, R *tot,
                   const R *den, R *Sigma, R *Ex, R *Ey, hipStream_t s) {
    hipLaunchKernelGGL(k_voltages<R>, dim3(1), dim3(64), 0, s, v, t, V);
    hipLaunchKernelGGL(k_potential<R>, dim3(node_blocks(n.Gsize)), dim3(256), 0, s, n, lap, (const R *)V, phi, tot);
    hipLaunchKernelGGL(k_efield<R>, dim3(node_blocks(n.Gsize)), dim3(256), 0, s,
                       EfieldJob<R>{n, phi, (const R *)tot, den, Sigma, Ex, Ey});
}
template <typename R>
void launch_accumulate(int Gsize, int nsp, const R *den, const R *tot, const R *Ex, const R *Ey, double *sden,
                       double *sphi, double *sex, double *sey, hipStream_t s) {
    hipLaunchKernelGGL(k_accumulate<R>, dim3(node_blocks(Gsize)), dim3(256), 0, s, Gsize, nsp, den, tot, Ex, Ey, sden,
                       sphi, sex, sey);
}

void launch_cell_counts(int Gsize, const int *cell_start, const int *pcnt, int *total, hipStream_t s) {
    hipLaunchKernelGGL(k_cell_counts, dim3(node_blocks(Gsize)), dim3(256), 0, s, Gsize, cell_start, pcnt, total);
}
void launch_fill_int(int *p, int v, int n, hipStream_t s) {
    hipLaunchKernelGGL(k_fill_int, dim3(node_blocks(n)), dim3(256), 0, s, p, v, n);
}
void launch_bump_ids(uint32_t *next_id, const int *prod_total_end, int nsp, hipStream_t s) {
    hipLaunchKernelGGL(k_bump_ids, dim3(1), dim3(64), 0, s, next_id, prod_total_end, nsp);
}
template <typename R>
void launch_rhs_to_fp64(int N, const R *b, double *b64, hipStream_t s) {
    hipLaunchKernelGGL(k_rhs_to_fp64<R>, dim3(node_blocks(N)), dim3(256), 0, s, N, b, b64);
}
template <typename R>
void launch_x_to_phi(int Gsize, const int *A_idx, const double *x, R *phi, hipStream_t s) {
    hipLaunchKernelGGL(k_x_to_phi<R>, dim3(node_blocks(Gsize)), dim3(256), 0, s, Gsize, A_idx, x, phi);
}
template void launch_rhs_to_fp64<float>(int, const float *, double *, hipStream_t);
template void launch_rhs_to_fp64<double>(int, const double *, double *, hipStream_t);
template void launch_x_to_phi<float>(int, const int *, const double *, float *, hipStream_t);
template void launch_x_to_phi<double>(int, const int *, const double *, double *, hipStream_t);

void launch_totals(int Gsize, const int *inv_start, const int *inv_ent, const int *cnt, const int *prod_cnt, int *total,
                   int *alive_total, int *prod_total, hipStream_t s) {
    hipLaunchKernelGGL(k_totals, dim3(node_blocks(Gsize)), dim3(256), 0, s, Gsize, inv_start, inv_ent, cnt, prod_cnt,
                       total, alive_total, prod_total);
}
void launch_assign_bases(int Gsize, const int *inv_start, const int *inv_ent, const int *cnt, const int *prod_cnt,
                         const int *start, const int *alive_total, const int *prod_start, int *base, int *pbase,
                         int *pid_off, hipStream_t s) {
    hipLaunchKernelGGL(k_assign_bases, dim3(node_blocks(Gsize)), dim3(256), 0, s, Gsize, inv_start, inv_ent, cnt,
                       prod_cnt, start, alive_total, prod_start, base, pbase, pid_off);
}
void launch_bases_pair(int G, const int *inv_start, const int *inv_ent, const int *cnt_e, const int *cnt_i,
                       const int *prod_cnt, int *total_e, int *alive_e, int *total_i, int *alive_i, int *prod_total,
                       int *start_e, int *start_i, int *prod_start, int *base_e, int *base_i, int *pbase_e,
                       int *pbase_i, int *pid_off, int *tmp, hipStream_t s) {
    const int nt = std::max(1, (G + BASES_TILE - 1) / BASES_TILE);
    const PairTotalsJob Jt{G, nt, inv_start, inv_ent, cnt_e, cnt_i, prod_cnt, total_e, alive_e, total_i, alive_i,
                           prod_total, tmp};
    const PairBasesJob Jb{G, nt, tmp, total_e, total_i, prod_total, alive_e, alive_i, inv_start, inv_ent, cnt_e,
                          cnt_i, prod_cnt, start_e, start_i, prod_start, base_e, base_i, pbase_e, pbase_i, pid_off};
    if (Recorder *r = launch_recorder()) {
        r->add(&launch_m<PairTotalsJob, k_pair_totals_m, BASES_TILE>, Jt, nt);
        r->add(&launch_m<PairBasesJob, k_pair_scan_bases_m, BASES_TILE>, Jb, nt);
        return;
    }
    hipLaunchKernelGGL(k_pair_totals, dim3(nt), dim3(BASES_TILE), 0, s, Jt);
    hipLaunchKernelGGL(k_pair_scan_bases, dim3(nt), dim3(BASES_TILE), 0, s, Jb);
}
// tile sums of up to 3 arrays: for launch_scan_multi (SCAN_TILE tiles) and
// launch_bases_pair (BASES_TILE tiles)
int scan_tmp_ints(int n) { return 3 * ((n + BASES_TILE - 1) / BASES_TILE) + 3; }
// exclusive scans of k <= 3 arrays of n ints: out[a][0..n], out[a][n] = total;
// tmp holds scan_tmp_ints(n) ints
void launch_totals_sp(int Gsize, int nsp, const int *inv_start, const int *inv_ent, const SortSpecies &j,
                      hipStream_t s) {
    hipLaunchKernelGGL(k_totals_sp, dim3(node_blocks(Gsize), nsp), dim3(256), 0, s, Gsize, inv_start, inv_ent, j);
}
void launch_assign_bases_sp(int Gsize, int nsp, const int *inv_start, const int *inv_ent, const SortSpecies &j,
                            hipStream_t s) {
    hipLaunchKernelGGL(k_assign_bases_sp, dim3(node_blocks(Gsize), nsp), dim3(256), 0, s, Gsize, inv_start, inv_ent,
                       j);
}
void launch_scan_multi(int k, const int *const *in, int *const *out, int n, int *tmp, hipStream_t s) {
    const int nt = std::max(1, (n + SCAN_TILE - 1) / SCAN_TILE);
    ScanJobs J{};
    for (int a = 0; a < k; ++a) { J.in[a] = in[a]; J.out[a] = out[a]; J.sums[a] = tmp + a * nt; }
    hipLaunchKernelGGL(k_scan_tile_sums, dim3(nt, k), dim3(1024), 0, s, J, n);
    hipLaunchKernelGGL(k_scan_tiles, dim3(nt, k), dim3(1024), 0, s, J, n, nt);
}
void launch_scan(const int *in, int *out, int n, int *tmp, hipStream_t s) {
    launch_scan_multi(1, &in, &out, n, tmp, s);
}

#define INST(R)                                                                                                         \
    template void launch_push<R>(const DevConst<R> &, const PView<R> &, int, uint8_t *, hipStream_t);                 \
    template void launch_push_count<R>(const DevConst<R> &, const StagedSet<R> &, int, int *, hipStream_t);          \
    template void launch_sort_scatter_all<R>(const DevConst<R> &, const StagedSet<R> &, int, int *, hipStream_t);    \
    template void launch_sort_count<R>(const DevConst<R> &, const PView<R> &, int, const uint8_t *, int *, int *,      \
                                       int *, hipStream_t);                                                            \
    template void launch_sort_scatter<R>(const DevConst<R> &, const PView<R> &, const PView<R> &, const uint8_t *,     \
                                         const int *, const int *, int, int *, hipStream_t);                           \
    template void launch_mcc_collide<R>(const DevConst<R> &, const PView<R> &, int, uint32_t, unsigned long long *,    \
                                        int *, int *, R *, hipStream_t);                                               \
    template void launch_mcc_insert<R>(const DevConst<R> &, const PView<R> &, const PView<R> &, const int *,           \
                                       const int *, const PView<R> &, const int *, const R *, int, const uint32_t *,   \
                                       int, int *, hipStream_t);                                                       \
    template void launch_fused_count<R>(const DevConst<R> &, const PView<R> &, int, const StepClock *, int *, int *,   \
                                        unsigned long long *, int *, uint8_t *, R *, size_t, int *, hipStream_t);      \
    template void launch_fused_scatter<R>(const DevConst<R> &, const PView<R> &, const PView<R> &, const PView<R> &,   \
                                          int, const uint8_t *, const R *, size_t, const int *, const int *,           \
                                          const int *, const int *, const uint32_t *, int, int, int *, hipStream_t);   \
    template void launch_deposit<R>(const DevConst<R> &, const NodeConst<R> &, const PView<R> *, R *, R *, R *, R *,   \
                                    R *, R *, const DepositExtras<R> *, hipStream_t);                                  \
    template void launch_phi_potential<R>(const NodeConst<R> &, const VoltConst &, const StepClock *, const double *,  \
                                          const R *, R *, R *, R *, hipStream_t);                                      \
    template void launch_eedf<R>(const PView<R> &, int, R, R, unsigned long long *, hipStream_t);               \
    template void launch_efield_only<R>(const NodeConst<R> &, const R *, const R *, const R *, R *, R *, R *,          \
                                        hipStream_t);                                                            \
    template void launch_efield<R>(const NodeConst<R> &, const VoltConst &, double, const R *, R *, const R *, R *,    \
                                   const R *, R *, R *, R *, hipStream_t);                                             \
    template void launch_accumulate<R>(int, int, const R *, const R *, const R *, const R *, double *, double *,       \
                                       double *, double *, hipStream_t);
INST(float)
INST(double)
#undef INST

}  // namespace pic
