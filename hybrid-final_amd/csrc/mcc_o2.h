// Collision step of the O2 chemistry (gas type 1): launch interface.
//
// The reference runs one thread per (species, node) that first moves the
// colliding particles of its bucket to the top and tags their target
// (O2Collision_Check, src/cuda_mccO2.cu:867-977), then applies the
// species' handler (src/cuda_mccO2.cu:2-866) with atomic appends into the
// buckets of every species.  Here each particle draws from its own Philox
// stream, the outcome of every collision is recorded per particle (a fate
// byte and up to two products), and the new buckets are rebuilt from those
// records in a fixed order, so the result does not depend on thread timing.
#pragma once
#include "pic_kernels.h"
#include "o2_chem.h"

namespace pic {

template <typename R>
struct O2Const {
    int G, ncy, ngy, Csize, TnRct, nlogx, nvel, MCCn;
    int ions;   // 0: only electrons collide, as the reference's MCC_O2_Basic (src/cuda_mcc.cu:76)
    R idLOGX, dt, dtm, idx, idy;
    R np2c[O2_NSP], Escale[O2_NSP];
    R sigv[O2_NSIGV];
    R Th_e[O2_NROWS];
    uint64_t seed;
    const R *xe, *cx, *vsave, *BackDen1, *BackVel1, *FluidN, *den;
    R *prob;   // G x O2_NTARGETS: cumulative target probabilities per species
};

// fate byte of a particle after the collision step
constexpr int O2F_DEAD = 1;       // removed by its own collision
constexpr int O2F_VICTIM = 0x80;  // O- removed by an electron detachment in its cell
__host__ __device__ inline int o2_fate_prod(uint8_t f, int slot) { return ((f >> (1 + 3 * slot)) & 7) - 1; }

// the four species' buckets and collision records
template <typename R>
struct O2Set {
    PView<R> p[O2_NSP];
    uint8_t *fate[O2_NSP];
    R *stash[O2_NSP];     // 6 arrays of p[s].cap: product 0 velocity, product 1 velocity
};

template <typename R> void launch_o2_probs(const O2Const<R> &k, hipStream_t s);
// collisions of species isp (all four species when isp < 0), in place;
// counts products per (species, node) into pcount (4 x G) and O- removal
// requests per node into delreq
template <typename R>
void launch_o2_collide(const O2Const<R> &k, const O2Set<R> &set, int isp, uint32_t step, unsigned long long *mccr,
                       int *pcount, int *delreq, hipStream_t s);
// marks the O- removed by electron detachment (smallest Philox keys first)
template <typename R>
void launch_o2_victims(const O2Const<R> &k, const O2Set<R> &set, uint32_t step, const int *delreq, hipStream_t s);
// per node and species: survivors plus products, into total (4 x G; nodes
// that start no cell keep the value already there)
template <typename R>
void launch_o2_totals(const O2Const<R> &k, const O2Set<R> &set, const int *pcount, int *total, hipStream_t s);
// new buckets of all four species (dst.p[t], cell_start already set):
// survivors in order, then products by (source species, parent, slot);
// pstart is 4 x (G+1)
template <typename R>
void launch_o2_scatter_all(const O2Const<R> &k, const O2Set<R> &set, const O2Set<R> &dst, const int *pstart,
                           const uint32_t *next_id, int *err, hipStream_t s);
// next_id[t] += product total of species t (pstart is 4 x (G+1))
void launch_o2_bump_ids(uint32_t *next_id, const int *pstart, int G, hipStream_t s);

}  // namespace pic
