// Collision step of the oxygen chemistries (O2, gas type 1; Ar/O2, gas
// type 2) for gfx950.
//
// Five launches per step:
//   probs    per node, the cumulative target probabilities of every species
//            (20 or 25 values; particle targets use this step's den)
//   collide  every species in one launch (one grid row each), one 64-lane
//            wave per four consecutive cells walked as one particle range:
//            null-collision test per particle; the colliders of a wave are
//            queued in LDS and collided 64 at a time on full wavefronts
//            (target, reaction,
//            outcome), written in place (survivors) or as a fate byte and up
//            to two products (velocities in a stash indexed like the parent);
//            event, product and detachment counts gather in LDS and reach
//            global memory once per wave
//   victims  per cell with electron detachments: removes that many O- of
//            the cell, smallest Philox keys first
//   totals   per node: survivors plus products of every species (one grid
//            row per species, one wave per four cells; then scans)
//   scatter  every species in one launch (one grid row per target species,
//            one wave per four cells): survivors in old order, then products
//            by (source species, parent, slot); new ids follow the global
//            (cell, source species, parent, slot) order
// Every count is an integer atomic or a ballot, so the result is the same
// for any schedule.  Reference: src/cuda_mcc.cu:21-96 (MCC_ArO2_Basic,
// MCC_O2_Basic), src/cuda_mccO2.cu:2-977, src/cuda_mccArO2.cu:51-1265
// (handlers and target tests).
#include <hip/hip_runtime.h>

#include "mcc_o2.h"

namespace pic {

namespace {

struct ChemRow {
    int r, sum, type, emode;
    double gain;
    int p0, p1, th;
};
struct ChemGroup {
    int sp, flag, sig, first, n, vsrc;
};
struct ChemTarget {
    int sp, flag, sig, kind, idx;
};

// both chemistries in one form: O2 rows subtract their own threshold, and
// its O2+ and O+ collide in the frame of a neutral at BackVel1
#define PIC_O2_ROW(r, sm, ty, em, g, a, b) {r, sm, ty, em, g, a, b, r},
#define PIC_ARO2_ROW(r, sm, ty, em, g, a, b, th) {r, sm, ty, em, g, a, b, th},
#define PIC_O2_GRP(s, f, sg, fi, n) {s, f, sg, fi, n, (s) == 1 || (s) == 2 ? 1 : 0},
#define PIC_ARO2_GRP(s, f, sg, fi, n, v) {s, f, sg, fi, n, v},
#define PIC_TGT(s, f, sg, kd, ix) {s, f, sg, kd, ix},
__constant__ ChemRow kRow[2][CHEM_MAXROWS] = {{O2_REACTIONS(PIC_O2_ROW)}, {ARO2_REACTIONS(PIC_ARO2_ROW)}};
__constant__ ChemGroup kGrp[2][CHEM_MAXGROUPS] = {{O2_GROUPS(PIC_O2_GRP)}, {ARO2_GROUPS(PIC_ARO2_GRP)}};
__constant__ ChemTarget kTgt[2][CHEM_MAXTARGETS] = {{O2_TARGETS(PIC_TGT)}, {ARO2_TARGETS(PIC_TGT)}};
__constant__ int kGroup0[2][CHEM_MAXSP] = {O2_GROUP0, ARO2_GROUP0};
__constant__ int kNGroup[2][CHEM_MAXSP] = {O2_NGROUP, ARO2_NGROUP};
#undef PIC_O2_ROW
#undef PIC_ARO2_ROW
#undef PIC_O2_GRP
#undef PIC_ARO2_GRP
#undef PIC_TGT

// Target probabilities: P_j = P_{j-1} + 1 - exp(-dt_s sigv n), summed in
// that order as the reference writes it (src/cuda_mccO2.cu:880-930,
// src/cuda_mccArO2.cu:1144-1205).
template <typename R>
__global__ __launch_bounds__(256) void k_o2_probs(O2Const<R> k) {
    const int g = blockIdx.x * 256 + threadIdx.x;
    if (g >= k.G) return;
    R acc = 0;
    int sp = -1;
    for (int t = 0; t < k.ntgt; ++t) {
        const ChemTarget T = kTgt[k.chem][t];
        if (T.sp != sp) { sp = T.sp; acc = 0; }
        R x = (((k.sub_mask >> T.sp) & 1) ? k.dtm : k.dt) * k.sigv[T.sig];
        if (T.kind == 0) x = x * (T.idx ? k.BackDen2[g] : k.BackDen1[g]);
        else if (T.kind == 1) x = x * k.FluidN[(size_t)T.idx * k.G + g];
        else x = x * k.den[(size_t)T.idx * k.G + g] * k.np2c[T.idx] * k.idx * k.idy;
        acc = (acc + (R)1) - (R)dm_exp(-(double)x);
        k.prob[(size_t)g * k.ntgt + t] = acc;
    }
}

template <typename R>
__device__ inline void o2_maxwell(const O2Const<R> &k, R vti, Stream &st, R &vx, R &vy, R &vz) {
    const int n = (int)((R)(k.nvel - 1) * st.u<R>());
    const R rphi = st.u<R>(), rthe = st.u<R>();
    maxwell_draw<R>(vx, vy, vz, k.vsave[n], vti, rphi, rthe);
}

// One collision of a particle of species isp with the target of group grp.
// Returns the reaction (or -1 for a null collision); vx..vz are updated for
// a survivor, and dead / products / detach describe the rest.
template <typename R>
__device__ inline int o2_handle(const O2Const<R> &k, int isp, int grp, int c, Stream &st, R &vx, R &vy, R &vz,
                                int &dead, int ps[2], R pv[2][3], int &np, int &detach) {
    const ChemGroup g = kGrp[k.chem][grp];
    R vnx = 0, vny = 0, vnz = 0;
    const bool heavy = g.vsrc != 0;   // positive ions collide in the neutral's frame
    if (heavy) o2_maxwell<R>(k, g.vsrc == 2 ? k.BackVel2[c] : k.BackVel1[c], st, vnx, vny, vnz);
    R VX = vx - vnx, VY = vy - vny, VZ = vz - vnz;
    const R dum = VX * VX + VY * VY + VZ * VZ;
    R vel = rsqrt_ieee<R>(dum);
    if (!heavy) { VX /= vel; VY /= vel; VZ /= vel; }
    R engy = k.Escale[isp] * dum;
    const R R1 = st.u<R>() * k.sigv[g.sig] / vel;
    ArXs<R> xs;
    xs.init(k, engy);
    R sum = 0;
    int hit = -1;
    for (int q = g.first; q < g.first + g.n; ++q) {
        const int r = kRow[k.chem][q].r;
        if (engy > k.Th_e[r]) {
            const R sg = xs.sigma(k, r);
            sum = kRow[k.chem][q].sum ? sum + sg : sg;
            if (R1 <= sum) { hit = q; break; }
        }
    }
    if (hit < 0) return -1;
    const ChemRow row = kRow[k.chem][hit];
    const R Es = k.Escale[isp];
    switch (row.type) {
    case O2T_LOSS:
    case O2T_ION:
    case O2T_DETACH: {
        if (row.emode == O2E_LOSS) {
            engy -= k.Th_e[row.th];
            vel = rsqrt_ieee<R>(rabs<R>(engy) / Es);
        } else if (row.emode == O2E_GAIN) {
            engy += (R)row.gain;
            vel = rsqrt_ieee<R>(rabs<R>(engy) / Es);
        }
        if (row.type == O2T_ION) {
            for (int slot = 0; slot < 2; ++slot) {
                const int sp = slot ? row.p1 : row.p0;
                ps[np] = sp;
                if (sp == 0) { pv[np][0] = VX; pv[np][1] = VY; pv[np][2] = VZ; }
                else o2_maxwell<R>(k, k.BackVel1[c], st, pv[np][0], pv[np][1], pv[np][2]);
                ++np;
            }
        } else if (row.type == O2T_DETACH) {
            ps[np] = 0; pv[np][0] = VX; pv[np][1] = VY; pv[np][2] = VZ; ++np;
            detach = 1;
        }
        const R a = st.u<R>(), b = st.u<R>();
        anewvel<R>(engy, vel, VX, VY, VZ, a, b);
        vx = VX; vy = VY; vz = VZ;
        break;
    }
    case O2T_IONSHARE: {
        engy -= k.Th_e[row.th];
        const R u = st.u<R>();
        const R rengy = (R)10 * PMath<R>::tan(u * PMath<R>::atan(engy / (R)20));
        engy -= rengy;
        const R velr = rsqrt_ieee<R>(rabs<R>(rengy) / Es);
        const R vel2 = rsqrt_ieee<R>(rabs<R>(engy) / Es);
        for (int slot = 0; slot < 2; ++slot) {
            const int sp = slot ? row.p1 : row.p0;
            ps[np] = sp;
            if (sp == 0) {
                R ex = VX, ey = VY, ez = VZ;
                const R a = st.u<R>(), b = st.u<R>();
                anewvel<R>(rengy, velr, ex, ey, ez, a, b);
                pv[np][0] = ex; pv[np][1] = ey; pv[np][2] = ez;
            } else {
                o2_maxwell<R>(k, k.BackVel1[c], st, pv[np][0], pv[np][1], pv[np][2]);
            }
            ++np;
        }
        const R a = st.u<R>(), b = st.u<R>();
        anewvel<R>(engy, vel2, VX, VY, VZ, a, b);
        vx = VX; vy = VY; vz = VZ;
        break;
    }
    case O2T_ATTACH:
        dead = 1;
        ps[np] = row.p0;
        o2_maxwell<R>(k, k.BackVel1[c], st, pv[np][0], pv[np][1], pv[np][2]);
        ++np;
        break;
    case O2T_DEL:
        dead = 1;
        break;
    case O2T_SCAT: {
        const R a = st.u<R>(), b = st.u<R>();
        newvel_ionsc<R>(VX, VY, VZ, vel, a, b);
        if (heavy) { vx = VX + vnx; vy = VY + vny; vz = VZ + vnz; }
        else { vx = VX; vy = VY; vz = VZ; }
        break;
    }
    case O2T_CXKEEP:
        vx = vnx; vy = vny; vz = vnz;
        break;
    case O2T_CXTO:
        dead = 1;
        ps[np] = row.p0; pv[np][0] = vnx; pv[np][1] = vny; pv[np][2] = vnz; ++np;
        break;
    default:   // O2T_NEGDET
        dead = 1;
        ps[np] = 0; pv[np][0] = VX; pv[np][1] = VY; pv[np][2] = VZ; ++np;
        break;
    }
    return row.r;
}

constexpr int O2_WAVES = 4;     // waves per block, one cell each
constexpr int O2_QCAP = 128;    // queued colliders per wave (< 64 wait while a chunk of 64 is tested)

// Every species in one launch when isp < 0 (grid row = species).  A wave
// tests 64 particles at a time; the colliders go to its LDS queue, and
// whenever 64 are queued they collide together, one per lane, so the long
// collision code runs on full wavefronts.
// A wave takes O2_CPW consecutive cells of one species: their particles are
// one contiguous range (cells of a column are consecutive nodes; the node
// between two columns starts no cell), walked 64 at a time, so species with
// few particles per cell still fill the wave.  Queue entries carry the
// cell's index within the wave; event, product and detachment counts are
// kept per (cell, row) in LDS.
constexpr int O2_CPW = 4;

template <typename R>
__global__ __launch_bounds__(256) void k_o2_collide(O2Const<R> k, O2Set<R> set, int isp, uint32_t step,
                                                    unsigned long long *mccr, int *pcount, int *delreq) {
    __shared__ int q_kk[O2_WAVES][O2_QCAP];
    __shared__ unsigned char q_j[O2_WAVES][O2_QCAP];
    __shared__ unsigned char q_c[O2_WAVES][O2_QCAP];
    __shared__ int c_r[O2_WAVES][O2_CPW][CHEM_MAXROWS + 1];   // events per (cell, row); last = detachments
    __shared__ int c_p[O2_WAVES][O2_CPW][CHEM_MAXSP];         // products per (cell, species)
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (isp < 0) isp = blockIdx.y;
    const int w0 = __builtin_amdgcn_readfirstlane((int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6)) * O2_CPW;
    if (w0 >= k.Csize) return;
    const PView<R> p = set.p[isp];
    uint8_t *fate = set.fate[isp];
    R *stash = set.stash[isp];
    int cell[O2_CPW], lo[O2_CPW + 1];
#pragma unroll
    for (int q = 0; q < O2_CPW; ++q) {
        const int wq = w0 + q;
        cell[q] = wq < k.Csize ? (wq / k.ncy) * k.ngy + (wq % k.ncy) : -1;
    }
    {
        int a0 = 0, a1 = 0;
        cell_range(p, cell[0], a0, a1);
        lo[0] = a0;
        lo[1] = a1;
#pragma unroll
        for (int q = 1; q < O2_CPW; ++q) {
            int b0 = lo[q], b1 = lo[q];
            if (cell[q] >= 0) cell_range(p, cell[q], b0, b1);
            lo[q + 1] = max(b1, lo[q]);
        }
    }
    const int hi = lo[O2_CPW];
    if (isp != 0 && !k.ions) {
        for (int kk = lo[0] + lane; kk < hi; kk += 64) fate[kk] = 0;
        return;
    }
    auto cell_of = [&](int q) {
        int c = cell[0];
#pragma unroll
        for (int r = 1; r < O2_CPW; ++r) c = q == r ? cell[r] : c;
        return c;
    };
    for (int i = lane; i < O2_CPW * (CHEM_MAXROWS + 1); i += 64) (&c_r[wv][0][0])[i] = 0;
    for (int i = lane; i < O2_CPW * CHEM_MAXSP; i += 64) (&c_p[wv][0][0])[i] = 0;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    const size_t cap = (size_t)p.cap;
    const int g0 = kGroup0[k.chem][isp], ng = kNGroup[k.chem][isp];
    const int draws = ((k.sub_mask >> isp) & 1) ? k.MCCn : 1;
    const uint32_t tag = mcc_tag(isp);
    const unsigned long long lt = lanes_below(lane);
    int qn = 0;   // wave-uniform queue length

    // collide queue entries [q0, q0 + n), one per lane
    auto collide = [&](int q0, int n) {
        if (lane >= n) return;
        const int kk = q_kk[wv][q0 + lane];
        const int jc = q_j[wv][q0 + lane];
        const int kq = q_c[wv][q0 + lane];
        const int c = cell_of(kq);
        const R *P = k.prob + (size_t)c * k.ntgt + g0;   // target entries and handler groups share their order
        Stream st;
        st.init(k.seed, p.id[kk], step, tag);
        st.seek(jc);
        const R R1 = st.u<R>();
        int flag = 0;
        while (flag < ng - 1 && R1 > P[flag]) ++flag;
        R vx = p.vx[kk], vy = p.vy[kk], vz = p.vz[kk];
        int dead = 0, np = 0, detach = 0, ps[2] = {-1, -1};
        R pv[2][3];
        const int r = o2_handle<R>(k, isp, g0 + flag, c, st, vx, vy, vz, dead, ps, pv, np, detach);
        uint8_t f = 0;
        if (r >= 0) {
            atomicAdd(&c_r[wv][kq][r], 1);
            if (dead) f |= O2F_DEAD;
            else { p.vx[kk] = vx; p.vy[kk] = vy; p.vz[kk] = vz; }
            for (int q = 0; q < np; ++q) {
                f |= (uint8_t)((ps[q] + 1) << (1 + 3 * q));
                stash[(3 * q) * cap + kk] = pv[q][0];
                stash[(3 * q + 1) * cap + kk] = pv[q][1];
                stash[(3 * q + 2) * cap + kk] = pv[q][2];
                atomicAdd(&c_p[wv][kq][ps[q]], 1);
            }
            if (detach) atomicAdd(&c_r[wv][kq][CHEM_MAXROWS], 1);
        }
        fate[kk] = f;
    };

    for (int b = lo[0]; b < hi; b += 64) {
        const int kk = b + lane;
        int kq = 0;
#pragma unroll
        for (int q = 1; q < O2_CPW; ++q) kq += kk >= lo[q] ? 1 : 0;
        int jc = -1;
        if (kk < hi) {
            const R Tprob = k.prob[(size_t)cell_of(kq) * k.ntgt + g0 + ng - 1];
            Stream st;
            st.init(k.seed, p.id[kk], step, tag);
            for (int j = 0; j < draws; ++j)
                if (st.u<R>() < Tprob) { jc = j; break; }
            if (jc < 0) fate[kk] = 0;
        }
        const unsigned long long m = __ballot(jc >= 0);
        if (jc >= 0) {
            const int pos = qn + __popcll(m & lt);
            q_kk[wv][pos] = kk;
            q_j[wv][pos] = (unsigned char)jc;
            q_c[wv][pos] = (unsigned char)kq;
        }
        qn += __popcll(m);
        __builtin_amdgcn_wave_barrier();
        if (qn >= 64) {
            qn -= 64;
            collide(qn, 64);   // the newest 64; older entries stay in place
            __builtin_amdgcn_wave_barrier();
        }
    }
    if (qn > 0) collide(0, qn);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    for (int t = lane; t < O2_CPW * CHEM_MAXROWS; t += 64) {
        const int q = t / CHEM_MAXROWS, i = t % CHEM_MAXROWS;
        const int n = c_r[wv][q][i];
        if (n) atomicAdd(&mccr[(size_t)cell_of(q) * k.TnRct + i], (unsigned long long)n);
    }
    for (int t = lane; t < O2_CPW * k.nsp; t += 64) {
        const int q = t / k.nsp, sp = t % k.nsp;
        const int n = c_p[wv][q][sp];
        if (n) atomicAdd(&pcount[(size_t)sp * k.G + cell_of(q)], n);
    }
    if (lane < O2_CPW) {
        const int n = c_r[wv][lane][CHEM_MAXROWS];
        if (n) atomicAdd(&delreq[cell_of(lane)], n);
    }
}

// The reference removes a random O- of the cell for each detachment
// (src/cuda_mccO2.cu:398-424, src/cuda_mccArO2.cu:533-551, racing with the
// O- handler of that cell).  Here the candidates are the cell's O- that
// survived their own collision; min(requests, candidates) of them go, in
// increasing order of a Philox key of (seed, id, step), ties by bucket
// position.  The keys of up to O2_VKEYS candidates are computed once into
// LDS; larger cells recompute them per comparison chunk.
constexpr int O2_VKEYS = 256;
template <typename R>
__global__ __launch_bounds__(256) void k_o2_victims(O2Const<R> k, O2Set<R> set, uint32_t step, const int *delreq,
                                                    unsigned long long *delcnt) {
    __shared__ uint32_t s_key[O2_WAVES][O2_VKEYS];
    __shared__ unsigned char s_ok[O2_WAVES][O2_VKEYS];
    int lane;
    const int c = wave_cell(k.ncy, k.ngy, k.Csize, lane);
    if (c < 0) return;
    const int req = delreq[c];
    if (req == 0) return;
    const int wv = threadIdx.x >> 6;
    const PView<R> p = set.p[k.neg];
    uint8_t *fate = set.fate[k.neg];
    int k0, k1;
    cell_range(p, c, k0, k1);
    int ncand = 0;
    for (int b = k0; b < k1; b += 64) {
        const int kk = b + lane;
        ncand += __popcll(__ballot(kk < k1 && !(fate[kk] & O2F_DEAD)));
    }
    const int kdel = min(req, ncand);
    if (lane == 0 && kdel > 0) atomicAdd(&delcnt[c], (unsigned long long)kdel);
    if (kdel == 0) return;
    const uint32_t s0 = (uint32_t)k.seed, s1 = (uint32_t)(k.seed >> 32);
    auto key_of = [&](int j, bool &ok) -> uint32_t {
        ok = j < k1 && !(fate[j] & O2F_DEAD);
        if (!ok) return 0u;
        uint32_t o[4];
        philox4x32_10(p.id[j], step, k.victim_tag, 0u, s0, s1, o);
        return o[0];
    };
    const int n = k1 - k0;
    const bool cached = n <= O2_VKEYS;
    if (cached) {
        for (int j = lane; j < n; j += 64) {
            bool ok;
            s_key[wv][j] = key_of(k0 + j, ok);
            s_ok[wv][j] = ok ? 1 : 0;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
    }
    for (int a0 = k0; a0 < k1; a0 += 64) {
        const int i = a0 + lane;
        bool ai;
        uint32_t ki;
        if (cached) {
            ai = i < k1 && s_ok[wv][i - k0];
            ki = ai ? s_key[wv][i - k0] : 0u;
        } else {
            ki = key_of(i, ai);
        }
        int rank = 0;
        if (cached) {
            if (ai)
                for (int j = 0; j < n; ++j)
                    if (s_ok[wv][j]) {
                        const uint32_t kj = s_key[wv][j];
                        if (kj < ki || (kj == ki && k0 + j < i)) ++rank;
                    }
        } else {
            for (int b0 = k0; b0 < k1; b0 += 64) {
                bool aj;
                const uint32_t kj = key_of(b0 + lane, aj);
                for (int q = 0; q < 64; ++q) {
                    const uint32_t kq = (uint32_t)__shfl((int)kj, q, 64);
                    const int aq = __shfl((int)aj, q, 64);
                    if (aq && (kq < ki || (kq == ki && b0 + q < i))) ++rank;
                }
            }
        }
        // only bit O2F_DEAD decides candidacy, so marking here leaves the
        // candidate set of the other lanes unchanged
        if (ai && rank < kdel) fate[i] = (uint8_t)(fate[i] | O2F_VICTIM);
    }
}

// per node and species: survivors plus products (total is nsp x G; nodes
// that start no cell keep the value already there)
// One wave per (O2_CPW consecutive cells, species t): the survivors of t in
// those cells, walked as one range and counted per cell, plus the cell's
// products that become t.
template <typename R>
__global__ __launch_bounds__(256) void k_o2_totals(O2Const<R> k, O2Set<R> set, const int *pcount, int *total) {
    const int lane = threadIdx.x & 63;
    const int t = blockIdx.y;
    const int w0 = __builtin_amdgcn_readfirstlane((int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6)) * O2_CPW;
    if (w0 >= k.Csize || t >= k.nsp) return;
    const PView<R> p = set.p[t];
    const uint8_t *fate = set.fate[t];
    int cell[O2_CPW], lo[O2_CPW + 1], n[O2_CPW];
#pragma unroll
    for (int q = 0; q < O2_CPW; ++q) {
        const int wq = w0 + q;
        cell[q] = wq < k.Csize ? (wq / k.ncy) * k.ngy + (wq % k.ncy) : -1;
        n[q] = 0;
    }
    {
        int a0 = 0, a1 = 0;
        cell_range(p, cell[0], a0, a1);
        lo[0] = a0;
        lo[1] = a1;
#pragma unroll
        for (int q = 1; q < O2_CPW; ++q) {
            int b0 = lo[q], b1 = lo[q];
            if (cell[q] >= 0) cell_range(p, cell[q], b0, b1);
            lo[q + 1] = max(b1, lo[q]);
        }
    }
    const int hi = lo[O2_CPW];
    for (int b = lo[0]; b < hi; b += 64) {
        const int kk = b + lane;
        int kq = 0;
#pragma unroll
        for (int q = 1; q < O2_CPW; ++q) kq += kk >= lo[q] ? 1 : 0;
        const bool alive = kk < hi && !(fate[kk] & (O2F_DEAD | O2F_VICTIM));
#pragma unroll
        for (int q = 0; q < O2_CPW; ++q) n[q] += __popcll(__ballot(alive && kq == q));
    }
    if (lane == 0) {
#pragma unroll
        for (int q = 0; q < O2_CPW; ++q)
            if (cell[q] >= 0) total[(size_t)t * k.G + cell[q]] = n[q] + pcount[(size_t)t * k.G + cell[q]];
    }
}

// Every species in one pass per cell: survivors of each species, then one
// walk over the parents of every source species, whose products are placed
// into all target species with one wave scan each: new bucket = survivors in
// old order, then products by (source species, parent, slot).
// One wave per (O2_CPW consecutive cells, target species t): the survivors
// of t in those cells, walked as one range, then the products that become
// t, by source species, parent and slot, for each of the cells that has any
// (the product scan tells which).  Survivors keep their order; products
// follow them in (source species, parent, slot) order, with ids from the
// global product scan.
template <typename R>
__global__ __launch_bounds__(256) void k_o2_scatter_all(O2Const<R> k, O2Set<R> set, O2Set<R> dst, const int *pstart,
                                                        const uint32_t *next_id, int *err) {
    const int lane = threadIdx.x & 63;
    const int t = blockIdx.y;
    const int w0 = __builtin_amdgcn_readfirstlane((int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6)) * O2_CPW;
    if (w0 >= k.Csize || t >= k.nsp) return;
    const unsigned long long lt = lanes_below(lane);
    const int nsp = k.nsp;
    int cell[O2_CPW], lo[O2_CPW + 1], run[O2_CPW], ns[O2_CPW];
#pragma unroll
    for (int q = 0; q < O2_CPW; ++q) {
        const int wq = w0 + q;
        cell[q] = wq < k.Csize ? (wq / k.ncy) * k.ngy + (wq % k.ncy) : -1;
        run[q] = 0;
    }
    const PView<R> p = set.p[t];
    const PView<R> d = dst.p[t];
    const uint8_t *ft = set.fate[t];
    {
        int a0 = 0, a1 = 0;
        cell_range(p, cell[0], a0, a1);
        lo[0] = a0;
        lo[1] = a1;
#pragma unroll
        for (int q = 1; q < O2_CPW; ++q) {
            int b0 = lo[q], b1 = lo[q];
            if (cell[q] >= 0) cell_range(p, cell[q], b0, b1);
            lo[q + 1] = max(b1, lo[q]);
        }
    }
#pragma unroll
    for (int q = 0; q < O2_CPW; ++q) ns[q] = cell[q] >= 0 ? d.cell_start[cell[q]] : 0;
    const int hi = lo[O2_CPW];
    // survivors of t; the lanes of one cell are a contiguous run
    for (int b = lo[0]; b < hi; b += 64) {
        const int kk = b + lane;
        int kq = 0;
#pragma unroll
        for (int q = 1; q < O2_CPW; ++q) kq += kk >= lo[q] ? 1 : 0;
        const bool keep = kk < hi && !(ft[kk] & (O2F_DEAD | O2F_VICTIM));
        int pos = 0;
#pragma unroll
        for (int q = 0; q < O2_CPW; ++q) {
            const unsigned long long m = __ballot(keep && kq == q);
            if (keep && kq == q) pos = ns[q] + run[q] + __popcll(m & lt);
            run[q] += __popcll(m);
        }
        if (keep) {
            if (pos >= d.cap) atomicOr(err, 1);
            else {
                d.x[pos] = p.x[kk]; d.y[pos] = p.y[kk];
                d.vx[pos] = p.vx[kk]; d.vy[pos] = p.vy[kk]; d.vz[pos] = p.vz[kk]; d.id[pos] = p.id[kk];
            }
        }
    }
    // products that become t
    const int *ps = pstart + (size_t)t * (k.G + 1);
    for (int q = 0; q < O2_CPW; ++q) {
        const int c = cell[q];
        if (c < 0) continue;
        const int p0 = ps[c];
        if (ps[c + 1] == p0) continue;   // none in this cell
        const int base = ns[q] + run[q];
        const uint32_t id0 = next_id[t] + (uint32_t)p0;
        int prun = 0;
        for (int s2 = 0; s2 < nsp; ++s2) {
            const PView<R> src = set.p[s2];
            const uint8_t *fq = set.fate[s2];
            const R *sq = set.stash[s2];
            const size_t qcap = (size_t)src.cap;
            int q0, q1;
            cell_range(src, c, q0, q1);
            for (int b = q0; b < q1; b += 64) {
                const int kk = b + lane;
                int n0 = 0, n1 = 0;
                if (kk < q1) {
                    const uint8_t f = fq[kk];
                    n0 = o2_fate_prod(f, 0) == t;
                    n1 = o2_fate_prod(f, 1) == t;
                }
                const int cnt = n0 + n1;
                if (!__ballot(cnt > 0)) continue;
                int incl = cnt;
#pragma unroll
                for (int off = 1; off < 64; off <<= 1) {
                    const int u = __shfl_up(incl, off, 64);
                    if (lane >= off) incl += u;
                }
                const int tot = __shfl(incl, 63, 64);
                int r = prun + incl - cnt;
                for (int slot = 0; slot < 2; ++slot) {
                    if (!(slot ? n1 : n0)) continue;
                    const int pos = base + r;
                    if (pos >= d.cap) atomicOr(err, 1);
                    else {
                        d.x[pos] = src.x[kk]; d.y[pos] = src.y[kk];
                        d.vx[pos] = sq[(3 * slot) * qcap + kk];
                        d.vy[pos] = sq[(3 * slot + 1) * qcap + kk];
                        d.vz[pos] = sq[(3 * slot + 2) * qcap + kk];
                        d.id[pos] = id0 + (uint32_t)r;
                    }
                    ++r;
                }
                prun += tot;
            }
        }
    }
}

__global__ void k_o2_bump(uint32_t *next_id, const int *pstart, int G, int nsp) {
    const int t = threadIdx.x;
    if (t < nsp) next_id[t] += (uint32_t)pstart[(size_t)t * (G + 1) + G];
}

inline int wave_blocks(int Csize) { return (Csize + O2_WAVES - 1) / O2_WAVES; }

}  // namespace

template <typename R>
void launch_o2_probs(const O2Const<R> &k, hipStream_t s) {
    hipLaunchKernelGGL(k_o2_probs<R>, dim3((k.G + 255) / 256), dim3(256), 0, s, k);
}
template <typename R>
void launch_o2_collide(const O2Const<R> &k, const O2Set<R> &set, int isp, uint32_t step, unsigned long long *mccr,
                       int *pcount, int *delreq, hipStream_t s) {
    const int nb = (k.Csize + O2_CPW * O2_WAVES - 1) / (O2_CPW * O2_WAVES);
    hipLaunchKernelGGL(k_o2_collide<R>, dim3(nb, isp < 0 ? k.nsp : 1), dim3(64 * O2_WAVES), 0, s, k, set, isp, step,
                       mccr, pcount, delreq);
}
template <typename R>
void launch_o2_victims(const O2Const<R> &k, const O2Set<R> &set, uint32_t step, const int *delreq,
                       unsigned long long *delcnt, hipStream_t s) {
    hipLaunchKernelGGL(k_o2_victims<R>, dim3(wave_blocks(k.Csize)), dim3(64 * O2_WAVES), 0, s, k, set, step, delreq,
                       delcnt);
}
template <typename R>
void launch_o2_totals(const O2Const<R> &k, const O2Set<R> &set, const int *pcount, int *total, hipStream_t s) {
    const int nb = (k.Csize + O2_CPW * O2_WAVES - 1) / (O2_CPW * O2_WAVES);
    hipLaunchKernelGGL(k_o2_totals<R>, dim3(nb, k.nsp), dim3(64 * O2_WAVES), 0, s, k, set, pcount, total);
}
template <typename R>
void launch_o2_scatter_all(const O2Const<R> &k, const O2Set<R> &set, const O2Set<R> &dst, const int *pstart,
                           const uint32_t *next_id, int *err, hipStream_t s) {
    const int nb = (k.Csize + O2_CPW * O2_WAVES - 1) / (O2_CPW * O2_WAVES);
    hipLaunchKernelGGL(k_o2_scatter_all<R>, dim3(nb, k.nsp), dim3(64 * O2_WAVES), 0, s, k, set, dst, pstart, next_id,
                       err);
}
void launch_o2_bump_ids(uint32_t *next_id, const int *pstart, int G, int nsp, hipStream_t s) {
    hipLaunchKernelGGL(k_o2_bump, dim3(1), dim3(64), 0, s, next_id, pstart, G, nsp);
}

#define INST(R)                                                                                                       \
    template void launch_o2_probs<R>(const O2Const<R> &, hipStream_t);                                               \
    template void launch_o2_collide<R>(const O2Const<R> &, const O2Set<R> &, int, uint32_t, unsigned long long *,     \
                                       int *, int *, hipStream_t);                                                   \
    template void launch_o2_victims<R>(const O2Const<R> &, const O2Set<R> &, uint32_t, const int *,                  \
                                       unsigned long long *, hipStream_t);                                           \
    template void launch_o2_totals<R>(const O2Const<R> &, const O2Set<R> &, const int *, int *, hipStream_t);        \
    template void launch_o2_scatter_all<R>(const O2Const<R> &, const O2Set<R> &, const O2Set<R> &, const int *,      \
                                           const uint32_t *, int *, hipStream_t);
INST(float)
INST(double)
#undef INST

}  // namespace pic
