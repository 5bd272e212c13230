// Device-side building blocks of the PIC-MCC step: Philox stream, per-cell
// routing tables, the push of src/cuda_Move.cu:181-270, the wall handling of
// src/cuda_Sortboundary.cu:41-258 and the argon null-collision physics of
// src/cuda_mccAr.cu:41-232,707-777 and src/cuda_mcc.cu:139-201.
//
// Arithmetic is written in a fixed evaluation order and the library is built
// with -ffp-contract=off, so the CPU oracle and these kernels produce the
// same bits.  Transcendentals go through detmath.h.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "detmath.h"

namespace pic {

// ---------------------------------------------------------------- Philox4x32-10
// Salmon et al., SC'11.  key = 64-bit seed; counter = (id, step, tag, block).
__host__ __device__ inline void philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                              uint32_t k0, uint32_t k1, uint32_t out[4]) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c1 = (uint32_t)p1;
        c3 = (uint32_t)p0;
        c0 = n0;
        c2 = n2;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

// uniform in (0,1) from one word: (w>>9 + 1/2) / 2^23.  The sum needs 24
// significant bits and the scale is a power of two, so the value is exact in
// float and in double (both evaluations give the same bits).
template <typename R> __host__ __device__ inline R philox_uniform(uint32_t w);
template <> __host__ __device__ inline float philox_uniform<float>(uint32_t w) {
    return ((float)(w >> 9) + 0.5f) * (1.0f / 8388608.0f);
}
template <> __host__ __device__ inline double philox_uniform<double>(uint32_t w) {
    return ((double)(w >> 9) + 0.5) * (1.0 / 8388608.0);
}

struct Stream {
    uint32_t k0, k1, id, step, tag, block;
    uint32_t b0, b1, b2, b3;   // current Philox block, consumed b0 first
    int used;
    __device__ void init(uint64_t seed, uint32_t pid, uint32_t st, uint32_t tg) {
        k0 = (uint32_t)seed; k1 = (uint32_t)(seed >> 32);
        id = pid; step = st; tag = tg; block = 0; used = 4;
    }
    // position the stream after n draws (n >= 0), as if u() had been called n times
    __device__ void seek(int n) {
        block = (uint32_t)(n >> 2);
        used = 4;
        if (n & 3) {
            uint32_t o[4];
            philox4x32_10(id, step, tag, block, k0, k1, o);
            b0 = o[0]; b1 = o[1]; b2 = o[2]; b3 = o[3];
            ++block;
            used = n & 3;
        }
    }
    template <typename R>
    __device__ R u() {
        if (used == 4) {
            uint32_t o[4];
            philox4x32_10(id, step, tag, block, k0, k1, o);
            b0 = o[0]; b1 = o[1]; b2 = o[2]; b3 = o[3];
            ++block;
            used = 0;
        }
        // register select instead of an indexed buffer (keeps it out of LDS/scratch)
        uint32_t w = used == 0 ? b0 : used == 1 ? b1 : used == 2 ? b2 : b3;
        ++used;
        return philox_uniform<R>(w);
    }
};

__host__ __device__ constexpr uint32_t mcc_tag(int isp) { return ((uint32_t)isp << 8) | 1u; }

// ---------------------------------------------------------------- precision helpers
template <typename R> __device__ inline R rsqrt_ieee(R v);
template <> __device__ inline float rsqrt_ieee<float>(float v) { return __builtin_sqrtf(v); }
template <> __device__ inline double rsqrt_ieee<double>(double v) { return __builtin_sqrt(v); }
template <typename R> __device__ inline R rabs(R v) { return v < 0 ? -v : v; }

// ---------------------------------------------------------------- routing
// Move direction code: (dx+1)*3 + (dy+1); 4 = stays in the cell.
__host__ __device__ inline int dir_code(int mcid, int ngy) {
    int dxs = 0, dys = 0;
    if (mcid == 0) return 4;
    // mcid = dys + dxs*ngy with dys,dxs in {-1,0,1}
    if (mcid >= ngy - 1) { dxs = 1; dys = mcid - ngy; }
    else if (mcid <= -ngy + 1) { dxs = -1; dys = mcid + ngy; }
    else { dxs = 0; dys = mcid; }
    return (dxs + 1) * 3 + (dys + 1);
}
__host__ __device__ inline int code_mcid(int code, int ngy) { return (code / 3 - 1) * ngy + (code % 3 - 1); }

// Per-cell route entry for each direction code:
//   dest >= 0 : surviving particle lands in node-indexed cell dest
//   dest == -1: deleted (Dirichlet wall or undefined structure code)
//   dest == -2: absorbed by conductor (info = conductor index)
//   dest == -3: dielectric surface charge
// info bits 0-3: slot (distinct destination index, 0 = stayers);
//      bit 4: x -> 1-x, bit 5: y -> 1-y, bit 6: vx -> -vx, bit 7: vy -> -vy
struct Route {
    int dest;
    int info;
};
constexpr int ROUTE_SLOTS = 9;

// ---------------------------------------------------------------- case constants
template <typename R>
struct DevConst {
    int ngx, ngy, ncx, ncy, Gsize, Csize, nsp, CondNUMR, TnRct, nvel, MCCn, nlogx;
    R dt_dx, dt_dy, idLOGX;
    R Ascale[5], Escale[5];
    R sigv[3], Th_e[16];
    uint64_t seed;
    // node / table arrays
    const R *Ex, *Ey;
    const R *P1, *P2, *Pion, *BackVel1, *vsave, *xe, *cx;
    const int *StructIdx;
    const Route *route;        // Csize-independent: indexed by node cell * 9 + code
    int *lost;                 // per species: particles deleted at a wall (no conductor or dielectric there)
    unsigned long long *sigma_hits;   // nsp x Gsize: this step's dielectric hits, fixed point (sigma_fixed)
};

// Dielectric surface hits are summed in signed 64-bit fixed point, 2^-40 of a
// particle per unit.  Integer adds commute, so the per-node sums, and every
// bit of the surface charge folded from them, do not depend on the order in
// which the atomics land: repeated runs and the fused and staged steps agree
// bit for bit.  A node takes fewer than 2^23 hits per step: a particle moves
// at most one cell per step, so only particles of the four cells around the
// node can reach it.  The folding kernel adds the sums to the per-species
// surface charge once per step and clears them.
constexpr double SIGMA_FIX = 1099511627776.0;   // 2^40
__device__ inline unsigned long long sigma_fixed(double w) {
    return (unsigned long long)__double2ll_rn(w * SIGMA_FIX);
}

// ---------------------------------------------------------------- push
template <typename R> __host__ __device__ constexpr R one_minus_ulp();
template <> __host__ __device__ constexpr float one_minus_ulp<float>() { return 0.99999994f; }           // 1 - 2^-24
template <> __host__ __device__ constexpr double one_minus_ulp<double>() { return 0.99999999999999989; } // 1 - 2^-53

// Bilinear E gather and leapfrog push in cell units.  Returns the move code.
template <typename R>
__device__ inline int push_particle(R ews, R ewn, R ees, R een, R yws, R ywn, R yes, R yen, R As, R dtdx, R dtdy,
                                    R &x, R &y, R &vx, R &vy) {
    R lx = x, ly = y;
    R dvx = ews * (1 - lx) * (1 - ly) + ewn * (1 - lx) * ly + ees * lx * (1 - ly) + een * lx * ly;
    R dvy = yws * (1 - lx) * (1 - ly) + ywn * (1 - lx) * ly + yes * lx * (1 - ly) + yen * lx * ly;
    R mx = vx + dvx * As;
    R my = vy + dvy * As;
    lx += mx * dtdx;
    ly += my * dtdy;
    // at most one cell per axis and step, as in the reference; an offset
    // still outside [0,1) (a faster move) is clamped into the destination
    // cell instead of being left outside it.  Written as selects (no
    // branches): for an offset inside [0,1) every select keeps its value.
    const int dys = ly >= 1 ? 1 : (ly < 0 ? -1 : 0);
    const int dxs = lx >= 1 ? 1 : (lx < 0 ? -1 : 0);
    ly = ly >= 1 ? ly - 1 : (ly < 0 ? ly + 1 : ly);
    lx = lx >= 1 ? lx - 1 : (lx < 0 ? lx + 1 : lx);
    const R top = one_minus_ulp<R>();
    lx = lx >= 1 ? top : (lx < 0 ? (R)0 : lx);
    ly = ly >= 1 ? top : (ly < 0 ? (R)0 : ly);
    x = lx; y = ly; vx = mx; vy = my;
    return (dxs + 1) * 3 + (dys + 1);
}

template <typename R>
__device__ inline void apply_reflect(int info, R &x, R &y, R &vx, R &vy) {
    // a mirrored offset of 0 (a particle clamped onto the wall) would be 1:
    // keep it inside the cell like every other offset
    const R mxr = 1 - x, myr = 1 - y;
    x = (info & 16) ? (mxr >= 1 ? one_minus_ulp<R>() : mxr) : x;
    y = (info & 32) ? (myr >= 1 ? one_minus_ulp<R>() : myr) : y;
    vx = (info & 64) ? -vx : vx;
    vy = (info & 128) ? -vy : vy;
}

// Dielectric surface-charge split (src/cuda_Sortboundary.cu:91-211).
template <typename R>
__device__ inline void dielectric_split(const int *S, int ngy, int ID, int MCID, int SM, R X, R Y, R VX, R VY,
                                        int &pa, R &a, int &pb, R &b) {
    int CID = ID + MCID;
    R delta;
    pa = 0; pb = 0; a = 0; b = 0;
    if (MCID == 1) { delta = X - VX / VY * Y; pa = CID; a = 1 - delta; pb = CID + ngy; b = delta; }
    else if (MCID == -1) { delta = X + VX / VY * (1 - Y); pa = ID; a = 1 - delta; pb = ID + ngy; b = delta; }
    else if (MCID == ngy) { delta = Y - VY / VX * X; pa = CID; a = 1 - delta; pb = CID + 1; b = delta; }
    else if (MCID == -ngy) { delta = Y + VY / VX * (1 - X); pa = ID; a = 1 - delta; pb = ID + 1; b = delta; }
    else if (MCID == ngy + 1) {
        delta = X - VX / VY * Y;
        if (delta >= 0) {
            if (S[SM - 1]) { delta = Y - VY / VX * X; pa = CID - 1; a = -delta; pb = CID; b = 1 + delta; }
            else { pa = CID; a = 1 - delta; pb = CID + ngy; b = delta; }
        } else {
            if (S[SM - ngy - 1]) { pa = ID + 1; a = -delta; pb = CID; b = 1 + delta; }
            else { delta = Y - VY / VX * X; pa = CID; a = 1 - delta; pb = CID + 1; b = delta; }
        }
    } else if (MCID == ngy - 1) {
        delta = X + VX / VY * (1 - Y);
        if (delta >= 0) {
            if (S[SM + 1]) { delta = Y - VY / VX * X; pa = CID + 1; a = 2 - delta; pb = CID + 2; b = delta - 1; }
            else { pa = CID + 1; a = 1 - delta; pb = CID + 1 + ngy; b = delta; }
        } else {
            if (S[SM - ngy - 1]) { pa = ID; a = -delta; pb = CID + 1; b = 1 + delta; }
            else { delta = Y - VY / VX * X; pa = CID; a = 1 - delta; pb = CID + 1; b = delta; }
        }
    } else if (MCID == -ngy + 1) {
        delta = X - VX / VY * Y;
        if (delta < 1) {
            if (S[SM - 1]) { delta = Y + VY / VX * (1 - X); pa = CID + ngy - 1; a = -delta; pb = CID + ngy; b = 1 + delta; }
            else { pa = CID; a = 1 - delta; pb = CID + ngy; b = delta; }
        } else {
            if (S[SM + ngy + 1]) { pa = ID + 1; a = 2 - delta; pb = ID + 1 + ngy; b = delta - 1; }
            else { delta = Y + VY / VX * (1 - X); pa = ID + 1; a = 1 - delta; pb = ID + 2; b = delta; }
        }
    } else if (MCID == -ngy - 1) {
        delta = X + VX / VY * (1 - Y);
        if (delta < 1) {
            if (S[SM + 1]) { delta = Y + VY / VX * (1 - X); pa = ID; a = 2 - delta; pb = ID + 1; b = delta - 1; }
            else { pa = CID; a = 1 - delta; pb = ID; b = delta; }
        } else {
            if (S[SM + ngy + 1]) { pa = ID; a = 2 - delta; pb = ID + ngy; b = delta - 1; }
            else { delta = Y + VY / VX * (1 - X); pa = ID - 1; a = 1 - delta; pb = ID; b = delta; }
        }
    }
}

__host__ __device__ inline int padded_index(int ID, int MCID, int ngy) {
    int x = ID / ngy;
    if (MCID > 1) return (ID + ngy + 2 + x) + (MCID + 1);
    if (MCID < -1) return (ID + ngy + 2 + x) + (MCID - 1);
    return (ID + ngy + 2 + x) + MCID;
}

// ---------------------------------------------------------------- elementary functions
// In the path's precision: the FP32 path uses the single-precision
// deterministic routines of detmath.h, the FP64 path the double ones.
template <typename R> struct PMath;
template <> struct PMath<float> {
    static __device__ float log10(float x) { return dmf_log10(x); }
    static __device__ float pow(float a, float b) { return dmf_pow(a, b); }
    static __device__ void sincos(float x, float &s, float &c) { dmf_sincos(x, &s, &c); }
    static __device__ float tan(float x) { return dmf_tan(x); }
    static __device__ float atan(float x) { return dmf_atan(x); }
};
template <> struct PMath<double> {
    static __device__ double log10(double x) { return dm_log10(x); }
    static __device__ double pow(double a, double b) { return dm_pow(a, b); }
    static __device__ void sincos(double x, double &s, double &c) { s = dm_sin(x); c = dm_cos(x); }
    static __device__ double tan(double x) { return dm_tan(x); }
    static __device__ double atan(double x) { return dm_atan(x); }
};

// ---------------------------------------------------------------- argon MCC
// Cross section of reaction r at one energy (Argon_CrossSection,
// src/cuda_mccAr.cu:707-777): log-spaced table, linear inside, a decaying
// tail above it, the first row below it.  The table position depends only on
// the energy, so it is found once (init) and shared by all the reactions a
// collision compares; each sigma() is the same arithmetic as a lookup from
// scratch.
template <typename R>
struct ArXs {
    int mode;       // 0 zero energy, 1 below the table, 2 above it, 3 last row, 4 inside
    int ee2;
    R a1, a2;
    double tail;    // dm_exp(-(lengy - xe[N-1])) above the table
    template <class K>
    __device__ void init(const K &k, R engy) {
        const int N = k.nlogx;
        if (engy == 0) { mode = 0; return; }
        const R lengy = PMath<R>::log10(engy);
        const R ee1 = k.idLOGX * (lengy - k.xe[0]);
        ee2 = (int)ee1;
        a1 = ee1 - (R)ee2;
        a2 = 1 - a1;
        if (lengy < k.xe[0]) mode = 1;
        else if (lengy > k.xe[N - 1]) { mode = 2; tail = dm_exp(-(double)(lengy - k.xe[N - 1])); }
        else if (ee2 >= N - 1) mode = 3;
        else mode = 4;
    }
    template <class K>
    __device__ R sigma(const K &k, int r) const {
        const int N = k.nlogx, T = k.TnRct;
        switch (mode) {
        case 0: return 0;
        case 1: return k.cx[r];
        case 2: return (R)((double)k.cx[(N - 1) * T + r] * 0.1 * tail);
        case 3: return k.cx[(N - 1) * T + r];
        default: return a2 * k.cx[ee2 * T + r] + a1 * k.cx[(ee2 + 1) * T + r];
        }
    }
};
template <typename R>
__device__ inline R ar_sigma(const DevConst<R> &k, int r, R engy) {
    ArXs<R> x;
    x.init(k, engy);
    return x.sigma(k, r);
}

#define PIC_TWO_PI 6.28318530717958647692

template <typename R>
__device__ inline void anewvel(R energy, R vel, R &nvx, R &nvy, R &nvz, R r1, R r2) {
    R coschi;
    if (energy < (R)1e-30) coschi = 1;
    else coschi = (energy + 2 - 2 * PMath<R>::pow(energy + 1, r1)) / energy;
    R sinchi = rsqrt_ieee<R>(rabs<R>(1 - coschi * coschi));
    R phi1 = (R)PIC_TWO_PI * r2;
    R cosphi, sinphi;
    PMath<R>::sincos(phi1, sinphi, cosphi);
    R r13 = nvx, r23 = nvy, r33 = nvz, up1, up2, up3;
    if (r33 == 1) { up1 = 0; up2 = 1; up3 = 0; } else { up1 = 0; up2 = 0; up3 = 1; }
    R r12 = r23 * up3 - r33 * up2;
    R r22 = r33 * up1 - r13 * up3;
    R r32 = r13 * up2 - r23 * up1;
    R mag = rsqrt_ieee<R>(r12 * r12 + r22 * r22 + r32 * r32);
    r12 /= mag; r22 /= mag; r32 /= mag;
    R r11 = r22 * r33 - r32 * r23;
    R r21 = r32 * r13 - r12 * r33;
    R r31 = r12 * r23 - r22 * r13;
    nvx = vel * (r11 * sinchi * cosphi + r12 * sinchi * sinphi + r13 * coschi);
    nvy = vel * (r21 * sinchi * cosphi + r22 * sinchi * sinphi + r23 * coschi);
    nvz = vel * (r31 * sinchi * cosphi + r32 * sinchi * sinphi + r33 * coschi);
}

template <typename R>
__device__ inline void newvel_ionsc(R &vx, R &vy, R &vz, R vel, R r1, R r2) {
    R coschi = rsqrt_ieee<R>(r1);
    R sinchi = rsqrt_ieee<R>(rabs<R>(1 - coschi * coschi));
    R phi1 = (R)PIC_TWO_PI * r2;
    R cosphi, sinphi;
    PMath<R>::sincos(phi1, sinphi, cosphi);
    R r13 = vx / vel, r23 = vy / vel, r33 = vz / vel, up1, up2, up3;
    if (r33 == 1) { up1 = 0; up2 = 1; up3 = 0; } else { up1 = 0; up2 = 0; up3 = 1; }
    R r12 = r23 * up3 - r33 * up2;
    R r22 = r33 * up1 - r13 * up3;
    R r32 = r13 * up2 - r23 * up1;
    R mag = rsqrt_ieee<R>(r12 * r12 + r22 * r22 + r32 * r32);
    r12 /= mag; r22 /= mag; r32 /= mag;
    R r11 = r22 * r33 - r32 * r23;
    R r21 = r32 * r13 - r12 * r33;
    R r31 = r12 * r23 - r22 * r13;
    vx = vel * coschi * (r11 * sinchi * cosphi + r12 * sinchi * sinphi + r13 * coschi);
    vy = vel * coschi * (r21 * sinchi * cosphi + r22 * sinchi * sinphi + r23 * coschi);
    vz = vel * coschi * (r31 * sinchi * cosphi + r32 * sinchi * sinphi + r33 * coschi);
}

template <typename R>
__device__ inline void maxwell_draw(R &vx, R &vy, R &vz, R vsaven, R vti, R rphi, R rthe) {
    R aphi = (R)PIC_TWO_PI * rphi;
    R costheta = 1 - 2 * rthe;
    R sintheta = rsqrt_ieee<R>(1 - costheta * costheta);
    R cosa, sina;
    PMath<R>::sincos(aphi, sina, cosa);
    vx = vti * vsaven * sintheta * cosa;
    vy = vti * vsaven * sintheta * sina;
    vz = vti * vsaven * costheta;
}

// The collision functions come in two parts so kernels can run the cheap
// null-collision test on every particle and the collision itself only on the
// (few) particles that collide, gathered into full wavefronts.  The stream
// of a particle is fixed by (seed, id, step, tag); a collision resumes it
// with Stream::seek at the draw that follows the null test.

// Ion null test (Direct_Argon_ArIon, first draw): true when it collides.
template <typename R>
__device__ inline bool mcc_ar_ion_null(const DevConst<R> &k, R pion, uint32_t pid, uint32_t step) {
    Stream st;
    st.init(k.seed, pid, step, mcc_tag(1));
    return !(st.u<R>() > pion);
}

// Ion collision after a positive null test; st is positioned after draw 0.
// Returns the reaction index (5, 6) or -1.
template <typename R>
__device__ inline int mcc_ar_ion_collide(const DevConst<R> &k, int node, Stream &st, R &vx, R &vy, R &vz) {
    int n = (int)((R)(k.nvel - 1) * st.u<R>());
    R rphi = st.u<R>(), rthe = st.u<R>();
    R vnx, vny, vnz;
    maxwell_draw<R>(vnx, vny, vnz, k.vsave[n], k.BackVel1[node], rphi, rthe);
    R VX = vx - vnx, VY = vy - vnz, VZ = vz - vny;
    R dum = VX * VX + VY * VY + VZ * VZ;
    R engy = k.Escale[1] * dum;
    R vel = rsqrt_ieee<R>(dum);
    R R1 = st.u<R>() * k.sigv[2] / vel;
    ArXs<R> xs;
    xs.init(k, engy);
    R sum = xs.sigma(k, 5);
    if (R1 <= sum) {
        vx = vnx; vy = vny; vz = vnz;
        return 5;
    }
    if (R1 <= (sum += xs.sigma(k, 6))) {
        R a = st.u<R>(), b = st.u<R>();
        newvel_ionsc<R>(VX, VY, VZ, vel, a, b);
        vx = VX + vnx; vy = VY + vny; vz = VZ + vnz;
        return 6;
    }
    return -1;
}

// Ion collision (Direct_Argon_ArIon).  Returns the reaction index (5, 6) or -1.
// pion = k.Pion[node].
template <typename R>
__device__ inline int mcc_ar_ion(const DevConst<R> &k, int node, R pion, uint32_t pid, uint32_t step, R &vx, R &vy,
                                 R &vz) {
    if (!mcc_ar_ion_null<R>(k, pion, pid, step)) return -1;
    Stream st;
    st.init(k.seed, pid, step, mcc_tag(1));
    st.seek(1);
    return mcc_ar_ion_collide<R>(k, node, st, vx, vy, vz);
}

// Electron null test (Direct_Argon_Electron, up to MCCn draws against P2):
// returns the index of the draw that collided, or -1.
template <typename R>
__device__ inline int mcc_ar_electron_null(const DevConst<R> &k, R P2, uint32_t pid, uint32_t step) {
    Stream st;
    st.init(k.seed, pid, step, mcc_tag(0));
    for (int j = 0; j < k.MCCn; ++j)
        if (st.u<R>() < P2) return j;
    return -1;
}

// Electron collision after a positive null test at draw jcoll.  Returns the
// reaction index (0..4) or -1; for ionization (3, 4) fills the new electron
// and ion.  P1 = k.P1[node].
template <typename R>
__device__ inline int mcc_ar_electron_collide(const DevConst<R> &k, int node, R P1, uint32_t pid, uint32_t step,
                                              int jcoll, R &vx, R &vy, R &vz, R &evx, R &evy, R &evz, R &ivx,
                                              R &ivy, R &ivz) {
    Stream st;
    st.init(k.seed, pid, step, mcc_tag(0));
    st.seek(jcoll);
    const R R1 = st.u<R>();
    int target = (R1 <= P1) ? 0 : 1;
    R VX = vx, VY = vy, VZ = vz;
    R dum = VX * VX + VY * VY + VZ * VZ;
    R vel = rsqrt_ieee<R>(dum), vel2 = 0, rengy = 0;
    VX /= vel; VY /= vel; VZ /= vel;
    R engy = k.Escale[0] * dum;
    R R2, sum;
    int type = 1, ev = -1;
    const R Es = k.Escale[0];
    ArXs<R> xs;
    xs.init(k, engy);   // every comparison below uses this energy
    if (target == 0) {
        R2 = st.u<R>() * k.sigv[0] / vel;
        sum = xs.sigma(k, 0);
        if (R2 <= sum) { type = 0; ev = 0; }
        else if (engy > k.Th_e[1] && R2 <= (sum += xs.sigma(k, 1))) {
            type = 0; ev = 1; engy -= k.Th_e[1]; vel = rsqrt_ieee<R>(rabs<R>(engy) / Es);
        } else if (engy > k.Th_e[2] && R2 <= (sum += xs.sigma(k, 2))) {
            type = 0; ev = 2; engy -= k.Th_e[2]; vel = rsqrt_ieee<R>(rabs<R>(engy) / Es);
        } else if (engy > k.Th_e[3] && R2 <= (sum += xs.sigma(k, 3))) {
            type = 2; ev = 3; engy -= k.Th_e[3];
            R u = st.u<R>();
            rengy = (R)10 * PMath<R>::tan(u * PMath<R>::atan(engy / (R)20));
            engy -= rengy;
            vel = rsqrt_ieee<R>(rabs<R>(rengy) / Es);
            vel2 = rsqrt_ieee<R>(rabs<R>(engy) / Es);
        }
    } else {
        R2 = st.u<R>() * k.sigv[1] / vel;
        sum = xs.sigma(k, 4);
        if (engy > k.Th_e[4] && R2 <= sum) {
            type = 2; ev = 4; engy -= k.Th_e[4];
            R u = st.u<R>();
            rengy = (R)10 * PMath<R>::tan(u * PMath<R>::atan(engy / (R)20));
            engy -= rengy;
            vel = rsqrt_ieee<R>(rabs<R>(rengy) / Es);
            vel2 = rsqrt_ieee<R>(rabs<R>(engy) / Es);
        }
    }
    if (type == 0) {
        R a = st.u<R>(), b = st.u<R>();
        anewvel<R>(engy, vel, VX, VY, VZ, a, b);
        vx = VX; vy = VY; vz = VZ;
    } else if (type == 2) {
        evx = VX; evy = VY; evz = VZ;
        R a = st.u<R>(), b = st.u<R>();
        anewvel<R>(rengy, vel, evx, evy, evz, a, b);
        int n = (int)((R)(k.nvel - 1) * st.u<R>());
        R rphi = st.u<R>(), rthe = st.u<R>();
        maxwell_draw<R>(ivx, ivy, ivz, k.vsave[n], k.BackVel1[node], rphi, rthe);
        a = st.u<R>(); b = st.u<R>();
        anewvel<R>(engy, vel2, VX, VY, VZ, a, b);
        vx = VX; vy = VY; vz = VZ;
    }
    return ev;
}

// Electron collision (Direct_Argon_Electron): null test + collision.
template <typename R>
__device__ inline int mcc_ar_electron(const DevConst<R> &k, int node, R P1, R P2, uint32_t pid, uint32_t step,
                                      R &vx, R &vy, R &vz, R &evx, R &evy, R &evz, R &ivx, R &ivy, R &ivz) {
    const int jc = mcc_ar_electron_null<R>(k, P2, pid, step);
    if (jc < 0) return -1;
    return mcc_ar_electron_collide<R>(k, node, P1, pid, step, jc, vx, vy, vz, evx, evy, evz, ivx, ivy, ivz);
}

}  // namespace pic
