/* O2 chemistry (gas type 1) of the collision step, as data.
 *
 * Species: 0 e, 1 O2+, 2 O+, 3 O- (src/start.c:677-726).  A projectile
 * first picks a target with the null-collision test over its pairs
 * (src/cuda_mccO2.cu:867-977); then the handler of that (species, target)
 * group compares R * sigv[sig] / |v| with the running sum of the group's
 * cross sections and applies the first reaction that matches
 * (src/cuda_mccO2.cu:2-866).  This file lists both as tables, written as
 * X-macros so that C and HIP code build their own arrays from one list.
 *
 * Plain C: the oracle includes it too.
 */
#ifndef PIC_O2_CHEM_H
#define PIC_O2_CHEM_H

/* outcome of a matched reaction */
#define O2T_LOSS 1    /* electron: energy change, new direction (anewvel)                          */
#define O2T_ATTACH 2  /* electron removed; an O- appears, Maxwellian at the gas temperature        */
#define O2T_ION 3     /* electron: energy loss, products p0 and p1, then the electron scatters     */
#define O2T_DEL 4     /* projectile removed                                                        */
#define O2T_DETACH 5  /* e + O-: a second electron, energy loss, one O- of the cell removed         */
#define O2T_SCAT 6    /* ion: the relative velocity scatters (newvel_IONSC), neutral added back    */
#define O2T_CXKEEP 7  /* ion: takes the neutral's velocity                                          */
#define O2T_CXTO 8    /* ion: removed; a particle of species p0 appears with the neutral's velocity */
#define O2T_NEGDET 9  /* O-: removed; an electron appears                                           */

/* energy change of an electron reaction */
#define O2E_NONE 0
#define O2E_LOSS 1    /* minus the reaction threshold */
#define O2E_GAIN 2    /* plus the listed energy (superelastic) */

/* sum mode: 0 = the running sum restarts at this reaction's cross section,
 * 1 = adds to it.  Each group starts with 0; the reference also restarts
 * the sum at reactions 49, 50 and 54 (src/cuda_mccO2.cu:487-498, 623-629). */

/* X(r, sum, type, emode, gain_eV, p0, p1) in group order */
#define O2_REACTIONS(X) \
    /* e + O2 */ \
    X(0, 0, O2T_LOSS, O2E_NONE, 0.0f, -1, -1) \
    X(1, 1, O2T_LOSS, O2E_LOSS, 0.0f, -1, -1) \
    X(2, 1, O2T_LOSS, O2E_LOSS, 0.0f, -1, -1) \
    X(3, 1, O2T_LOSS, O2E_LOSS, 0.0f, -1, -1) \
    X(4, 1, O2T_LOSS, O2E_LOSS, 0.0f, -1, -1) \
    X(5, 1, O2T_LOSS, O2E_LOSS, 0.0f, -1, -1) \
    X(6, 1, O2T_ATTACH, O2E_NONE, 0.0f, 3, -1) \
    X(7, 1, O2T_LOSS, O2E_LOSS, 0.0f, -1, -1) \
    X(8, 1, O2T_LOSS, O2E_LOSS, 0.0f, -1, -1) \
    X(9, 1, O2T_LOSS, O2E_LOSS, 0.0f, -1, -1) \
    X(10, 1, O2T_ION, O2E_LOSS, 0.0f, 0, 1) \
    X(11, 1, O2T_LOSS, O2E_LOSS, 0.0f, -1, -1) \
    X(12, 1, O2T_ION, O2E_LOSS, 0.0f, 2, 3) \
    X(13, 1, O2T_ION, O2E_LOSS, 0.0f, 0, 2) \
    /* e + O2(a) */ \
    X(14, 0, O2T_ION, O2E_LOSS, 0.0f, 0, 1) \
    X(15, 1, O2T_ATTACH, O2E_NONE, 0.0f, 3, -1) \
    X(16, 1, O2T_LOSS, O2E_GAIN, 0.977, -1, -1) \
    X(17, 1, O2T_LOSS, O2E_GAIN, 0.977, -1, -1) \
    X(18, 1, O2T_LOSS, O2E_LOSS, 0.0f, -1, -1) \
    X(19, 1, O2T_LOSS, O2E_LOSS, 0.0f, -1, -1) \
    X(20, 1, O2T_LOSS, O2E_LOSS, 0.0f, -1, -1) \
    X(21, 1, O2T_ION, O2E_LOSS, 0.0f, 0, 2) \
    /* e + O2(b) */ \
    X(22, 0, O2T_ION, O2E_LOSS, 0.0f, 0, 1) \
    X(23, 1, O2T_ATTACH, O2E_NONE, 0.0f, 3, -1) \
    X(24, 1, O2T_LOSS, O2E_GAIN, 1.627, -1, -1) \
    X(25, 1, O2T_LOSS, O2E_GAIN, 1.627, -1, -1) \
    X(26, 1, O2T_LOSS, O2E_LOSS, 0.0f, -1, -1) \
    X(27, 1, O2T_LOSS, O2E_LOSS, 0.0f, -1, -1) \
    X(28, 1, O2T_LOSS, O2E_LOSS, 0.0f, -1, -1) \
    X(29, 1, O2T_ION, O2E_LOSS, 0.0f, 0, 2) \
    /* e + O- */ \
    X(30, 0, O2T_DETACH, O2E_LOSS, 0.0f, 0, -1) \
    /* e + O2+ */ \
    X(31, 0, O2T_DEL, O2E_NONE, 0.0f, -1, -1) \
    /* e + O(3P) */ \
    X(32, 0, O2T_LOSS, O2E_LOSS, 0.0f, -1, -1) \
    X(33, 1, O2T_LOSS, O2E_LOSS, 0.0f, -1, -1) \
    X(34, 1, O2T_LOSS, O2E_LOSS, 0.0f, -1, -1) \
    X(35, 1, O2T_LOSS, O2E_LOSS, 0.0f, -1, -1) \
    X(36, 1, O2T_LOSS, O2E_LOSS, 0.0f, -1, -1) \
    X(37, 1, O2T_ION, O2E_LOSS, 0.0f, 0, 2) \
    X(38, 1, O2T_LOSS, O2E_LOSS, 0.0f, -1, -1) \
    /* e + O(1D) */ \
    X(39, 0, O2T_ION, O2E_LOSS, 0.0f, 0, 2) \
    X(40, 1, O2T_LOSS, O2E_GAIN, 1.96, -1, -1) \
    /* O2+ + O(3P), O2, O2(a), O2(b) */ \
    X(47, 0, O2T_CXTO, O2E_NONE, 0.0f, 2, -1) \
    X(48, 0, O2T_CXKEEP, O2E_NONE, 0.0f, -1, -1) \
    X(49, 0, O2T_SCAT, O2E_NONE, 0.0f, -1, -1) \
    X(50, 0, O2T_CXTO, O2E_NONE, 0.0f, 2, -1) \
    X(51, 0, O2T_CXKEEP, O2E_NONE, 0.0f, -1, -1) \
    X(52, 0, O2T_CXKEEP, O2E_NONE, 0.0f, -1, -1) \
    /* O+ + O2, O(3P), O2(a), O2(b) */ \
    X(53, 0, O2T_CXTO, O2E_NONE, 0.0f, 1, -1) \
    X(54, 0, O2T_SCAT, O2E_NONE, 0.0f, -1, -1) \
    X(55, 0, O2T_CXKEEP, O2E_NONE, 0.0f, -1, -1) \
    X(56, 0, O2T_CXTO, O2E_NONE, 0.0f, 1, -1) \
    X(57, 0, O2T_CXTO, O2E_NONE, 0.0f, 1, -1) \
    /* O- + O2, O(3P), O2+, O+, O2(a) */ \
    X(41, 0, O2T_SCAT, O2E_NONE, 0.0f, -1, -1) \
    X(42, 1, O2T_NEGDET, O2E_NONE, 0.0f, 0, -1) \
    X(43, 0, O2T_NEGDET, O2E_NONE, 0.0f, 0, -1) \
    X(44, 0, O2T_DEL, O2E_NONE, 0.0f, -1, -1) \
    X(45, 0, O2T_DEL, O2E_NONE, 0.0f, -1, -1) \
    X(46, 0, O2T_NEGDET, O2E_NONE, 0.0f, 0, -1)

#define O2_NROWS 58

/* handler groups: G(species, target flag, sigv index used by the handler,
 * first row, rows).  The O- handler reads sigv[9 + flag] where its target
 * test used sigv[7 + flag] (src/cuda_mccO2.cu:758-820 against 959-966). */
#define O2_GROUPS(G) \
    G(0, 0, 0, 0, 14) G(0, 1, 1, 14, 8) G(0, 2, 2, 22, 8) G(0, 3, 3, 30, 1) G(0, 4, 4, 31, 1) \
    G(0, 5, 5, 32, 7) G(0, 6, 6, 39, 2) \
    G(1, 0, 12, 41, 1) G(1, 1, 13, 42, 3) G(1, 2, 14, 45, 1) G(1, 3, 15, 46, 1) \
    G(2, 0, 16, 47, 2) G(2, 1, 17, 49, 1) G(2, 2, 18, 50, 1) G(2, 3, 19, 51, 1) \
    G(3, 0, 9, 52, 2) G(3, 1, 10, 54, 1) G(3, 2, 11, 55, 1) G(3, 3, 12, 56, 1) G(3, 4, 13, 57, 1)

#define O2_NGROUPS 20
/* first group and group count per species */
#define O2_GROUP0 {0, 7, 11, 15}
#define O2_NGROUP {7, 4, 4, 5}

/* target test: T(species, flag, sigv index, kind, index) with kind 0 the
 * background gas density, 1 fluid species `index` (O2(a), O2(b), O(3P),
 * O(1D)), 2 particle species `index` (den * np2c / (dx dy)).  Electrons and
 * O2+ use dt / MCCn and up to MCCn draws; O+ and O- use dt and one draw. */
#define O2_TARGETS(T) \
    T(0, 0, 0, 0, 0) T(0, 1, 1, 1, 0) T(0, 2, 2, 1, 1) T(0, 3, 3, 2, 3) T(0, 4, 4, 2, 1) T(0, 5, 5, 1, 2) \
    T(0, 6, 6, 1, 3) \
    T(1, 0, 12, 1, 2) T(1, 1, 13, 0, 0) T(1, 2, 14, 1, 0) T(1, 3, 15, 1, 1) \
    T(2, 0, 16, 0, 0) T(2, 1, 17, 1, 2) T(2, 2, 18, 1, 0) T(2, 3, 19, 1, 1) \
    T(3, 0, 7, 0, 0) T(3, 1, 8, 1, 2) T(3, 2, 9, 2, 1) T(3, 3, 10, 2, 2) T(3, 4, 11, 1, 0)

#define O2_NTARGETS 20
#define O2_NSP 4
#define O2_NSIGV 20

/* stream tag of the key that orders the O- of a cell for removal by
 * electron detachment (collision streams use (species << 8) | 1) */
#define O2_VICTIM_TAG ((3u << 8) | 2u)

#endif
