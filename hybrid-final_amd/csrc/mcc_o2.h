// Collision step of the oxygen chemistries, O2 (gas type 1) and Ar/O2
// (gas type 2): launch interface.
//
// The reference runs one thread per (species, node) that first moves the
// colliding particles of its bucket to the top and tags their target
// (O2Collision_Check, src/cuda_mccO2.cu:867-977; ArO2_Collision_Check,
// src/cuda_mccArO2.cu:1130-1265), then applies the species' handler with
// atomic appends into the buckets of every species.  Here each particle
// draws from its own Philox stream, the outcome of every collision is
// recorded per particle (a fate byte and up to two products), and the new
// buckets are rebuilt from those records in a fixed order, so the result
// does not depend on thread timing.  The two chemistries differ only in
// their tables (o2_chem.h, aro2_chem.h).
#pragma once
#include "aro2_chem.h"
#include "o2_chem.h"
#include "pic_kernels.h"

namespace pic {

enum { CHEM_O2 = 0, CHEM_ARO2 = 1 };

template <typename R>
struct O2Const {
    int G, ncy, ngy, Csize, TnRct, nlogx, nvel, MCCn;
    int chem;       // CHEM_O2 or CHEM_ARO2: which tables
    int nsp;        // 4 (O2) or 5 (Ar/O2)
    int ntgt;       // target-test entries per node (20 or 25)
    int neg;        // the O- species
    int sub_mask;   // species that use dt / MCCn and up to MCCn target draws
    uint32_t victim_tag;
    int ions;       // 0: only electrons collide, as the reference's MCC_O2_Basic (src/cuda_mcc.cu:76)
    R idLOGX, dt, dtm, idx, idy;
    R np2c[CHEM_MAXSP], Escale[CHEM_MAXSP];
    R sigv[CHEM_MAXSIGV];
    R Th_e[CHEM_MAXROWS];
    uint64_t seed;
    const R *xe, *cx, *vsave, *BackDen1, *BackDen2, *BackVel1, *BackVel2, *FluidN, *den;
    R *prob;   // G x ntgt: cumulative target probabilities per species
};

// fate byte of a particle after the collision step
constexpr int O2F_DEAD = 1;       // removed by its own collision
constexpr int O2F_VICTIM = 0x80;  // O- removed by an electron detachment in its cell
__host__ __device__ inline int o2_fate_prod(uint8_t f, int slot) { return ((f >> (1 + 3 * slot)) & 7) - 1; }

// the species' buckets and collision records
template <typename R>
struct O2Set {
    PView<R> p[CHEM_MAXSP];
    uint8_t *fate[CHEM_MAXSP];
    R *stash[CHEM_MAXSP];     // 6 arrays of p[s].cap: product 0 velocity, product 1 velocity
};

template <typename R> void launch_o2_probs(const O2Const<R> &k, hipStream_t s);
// collisions of species isp (every species when isp < 0), in place;
// counts products per (species, node) into pcount (nsp x G), O- removal
// requests per node into delreq, and the O- actually removed for them per
// node into delcnt (by launch_o2_victims)
template <typename R>
void launch_o2_collide(const O2Const<R> &k, const O2Set<R> &set, int isp, uint32_t step, unsigned long long *mccr,
                       int *pcount, int *delreq, hipStream_t s);
// marks the O- removed by electron detachment (smallest Philox keys first)
// and adds their number per node to delcnt
template <typename R>
void launch_o2_victims(const O2Const<R> &k, const O2Set<R> &set, uint32_t step, const int *delreq,
                       unsigned long long *delcnt, hipStream_t s);
// per node and species: survivors plus products, into total (nsp x G;
// nodes that start no cell keep the value already there)
template <typename R>
void launch_o2_totals(const O2Const<R> &k, const O2Set<R> &set, const int *pcount, int *total, hipStream_t s);
// new buckets of every species (dst.p[t], cell_start already set):
// survivors in order, then products by (source species, parent, slot);
// pstart is nsp x (G+1)
template <typename R>
void launch_o2_scatter_all(const O2Const<R> &k, const O2Set<R> &set, const O2Set<R> &dst, const int *pstart,
                           const uint32_t *next_id, int *err, hipStream_t s);
// next_id[t] += product total of species t (pstart is nsp x (G+1))
void launch_o2_bump_ids(uint32_t *next_id, const int *pstart, int G, int nsp, hipStream_t s);

}  // namespace pic
