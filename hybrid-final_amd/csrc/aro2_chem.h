/* Ar/O2 chemistry (gas type 2) of the collision step, as data.
 *
 * Species: 0 e, 1 Ar+, 2 O2+, 3 O+, 4 O- (src/start.c:677-726).  A
 * projectile first picks a target with the null-collision test over its
 * pairs (ArO2_Collision_Check, src/cuda_mccArO2.cu:1130-1265); then the
 * handler of that (species, target) group compares R * sigv[sig] / |v| with
 * the running sum of the group's cross sections and applies the first
 * reaction that matches (ArO2_Electron 51-560, ArO2_Ar_ion 561-671,
 * ArO2_O2_ion 672-842, ArO2_O_ion 843-985, ArO2_O_negative 986-1129).
 * The outcome codes are those of o2_chem.h, plus O2T_IONSHARE.
 *
 * Plain C: the oracle includes it too.
 */
#ifndef PIC_ARO2_CHEM_H
#define PIC_ARO2_CHEM_H

#include "o2_chem.h"

/* electron: energy loss, then the remaining energy is shared with the new
 * electron by 10 tan(R atan(E / 20)); an electron product scatters with its
 * share (anewvel), a heavy product is Maxwellian at the gas temperature,
 * then the projectile scatters with the rest (src/cuda_mccArO2.cu:470-510) */
#define O2T_IONSHARE 10

/* X(r, sum, type, emode, gain_eV, p0, p1, th) in group order; th = the
 * reaction whose threshold an O2E_LOSS subtracts.  It is r itself except
 * for reaction 16, which tests its own threshold and subtracts that of
 * reaction 11 (src/cuda_mccArO2.cu:204-208). */
#define ARO2_REACTIONS(X) \
    /* e + Ar */ \
    X(0, 0, O2T_LOSS, O2E_NONE, 0.0f, -1, -1, 0) \
    X(1, 1, O2T_LOSS, O2E_LOSS, 0.0f, -1, -1, 1) \
    X(2, 1, O2T_LOSS, O2E_LOSS, 0.0f, -1, -1, 2) \
    X(3, 1, O2T_IONSHARE, O2E_LOSS, 0.0f, 0, 1, 3) \
    /* e + Ar* */ \
    X(4, 0, O2T_IONSHARE, O2E_LOSS, 0.0f, 0, 1, 4) \
    /* e + O2 */ \
    X(5, 0, O2T_LOSS, O2E_NONE, 0.0f, -1, -1, 5) \
    X(6, 1, O2T_LOSS, O2E_LOSS, 0.0f, -1, -1, 6) \
    X(7, 1, O2T_LOSS, O2E_LOSS, 0.0f, -1, -1, 7) \
    X(8, 1, O2T_LOSS, O2E_LOSS, 0.0f, -1, -1, 8) \
    X(9, 1, O2T_LOSS, O2E_LOSS, 0.0f, -1, -1, 9) \
    X(10, 1, O2T_LOSS, O2E_LOSS, 0.0f, -1, -1, 10) \
    X(11, 1, O2T_ATTACH, O2E_NONE, 0.0f, 4, -1, 11) \
    X(12, 1, O2T_LOSS, O2E_LOSS, 0.0f, -1, -1, 12) \
    X(13, 1, O2T_LOSS, O2E_LOSS, 0.0f, -1, -1, 13) \
    X(14, 1, O2T_LOSS, O2E_LOSS, 0.0f, -1, -1, 14) \
    X(15, 1, O2T_IONSHARE, O2E_LOSS, 0.0f, 0, 2, 15) \
    X(16, 1, O2T_LOSS, O2E_LOSS, 0.0f, -1, -1, 11) \
    X(17, 1, O2T_IONSHARE, O2E_LOSS, 0.0f, 3, 4, 17) \
    X(18, 1, O2T_IONSHARE, O2E_LOSS, 0.0f, 0, 3, 18) \
    /* e + O2(a) */ \
    X(19, 0, O2T_IONSHARE, O2E_LOSS, 0.0f, 0, 2, 19) \
    X(20, 1, O2T_ATTACH, O2E_NONE, 0.0f, 4, -1, 20) \
    X(21, 1, O2T_LOSS, O2E_GAIN, 0.977, -1, -1, 21) \
    X(22, 1, O2T_LOSS, O2E_GAIN, 0.977, -1, -1, 22) \
    X(23, 1, O2T_LOSS, O2E_LOSS, 0.0f, -1, -1, 23) \
    X(24, 1, O2T_LOSS, O2E_LOSS, 0.0f, -1, -1, 24) \
    X(25, 1, O2T_LOSS, O2E_LOSS, 0.0f, -1, -1, 25) \
    X(26, 1, O2T_IONSHARE, O2E_LOSS, 0.0f, 0, 3, 26) \
    /* e + O2(b) */ \
    X(27, 0, O2T_IONSHARE, O2E_LOSS, 0.0f, 0, 2, 27) \
    X(28, 1, O2T_ATTACH, O2E_NONE, 0.0f, 4, -1, 28) \
    X(29, 1, O2T_LOSS, O2E_GAIN, 1.627, -1, -1, 29) \
    X(30, 1, O2T_LOSS, O2E_GAIN, 1.627, -1, -1, 30) \
    X(31, 1, O2T_LOSS, O2E_LOSS, 0.0f, -1, -1, 31) \
    X(32, 1, O2T_LOSS, O2E_LOSS, 0.0f, -1, -1, 32) \
    X(33, 1, O2T_LOSS, O2E_LOSS, 0.0f, -1, -1, 33) \
    X(34, 1, O2T_IONSHARE, O2E_LOSS, 0.0f, 0, 3, 34) \
    /* e + O- */ \
    X(35, 0, O2T_DETACH, O2E_LOSS, 0.0f, 0, -1, 35) \
    /* e + O2+ */ \
    X(36, 0, O2T_DEL, O2E_NONE, 0.0f, -1, -1, 36) \
    /* e + O(3P) */ \
    X(37, 0, O2T_LOSS, O2E_LOSS, 0.0f, -1, -1, 37) \
    X(38, 1, O2T_LOSS, O2E_LOSS, 0.0f, -1, -1, 38) \
    X(39, 1, O2T_LOSS, O2E_LOSS, 0.0f, -1, -1, 39) \
    X(40, 1, O2T_LOSS, O2E_LOSS, 0.0f, -1, -1, 40) \
    X(41, 1, O2T_LOSS, O2E_LOSS, 0.0f, -1, -1, 41) \
    X(42, 1, O2T_IONSHARE, O2E_LOSS, 0.0f, 0, 3, 42) \
    X(43, 1, O2T_LOSS, O2E_LOSS, 0.0f, -1, -1, 43) \
    /* e + O(1D) */ \
    X(44, 0, O2T_IONSHARE, O2E_LOSS, 0.0f, 0, 3, 44) \
    X(45, 1, O2T_LOSS, O2E_GAIN, 1.96, -1, -1, 45) \
    /* Ar+ + Ar, O2 */ \
    X(65, 0, O2T_CXKEEP, O2E_NONE, 0.0f, -1, -1, 65) \
    X(66, 0, O2T_SCAT, O2E_NONE, 0.0f, -1, -1, 66) \
    X(67, 0, O2T_SCAT, O2E_NONE, 0.0f, -1, -1, 67) \
    /* O2+ + O(3P), O2, O2(a), O2(b), Ar */ \
    X(52, 0, O2T_CXTO, O2E_NONE, 0.0f, 3, -1, 52) \
    X(53, 0, O2T_CXKEEP, O2E_NONE, 0.0f, -1, -1, 53) \
    X(54, 0, O2T_SCAT, O2E_NONE, 0.0f, -1, -1, 54) \
    X(55, 0, O2T_CXTO, O2E_NONE, 0.0f, 3, -1, 55) \
    X(56, 0, O2T_CXKEEP, O2E_NONE, 0.0f, -1, -1, 56) \
    X(57, 0, O2T_CXKEEP, O2E_NONE, 0.0f, -1, -1, 57) \
    X(58, 0, O2T_CXTO, O2E_NONE, 0.0f, 1, -1, 58) \
    X(59, 0, O2T_CXKEEP, O2E_NONE, 0.0f, -1, -1, 59) \
    /* O+ + O2, O(3P), O2(a), O2(b) */ \
    X(60, 0, O2T_CXTO, O2E_NONE, 0.0f, 2, -1, 60) \
    X(61, 0, O2T_SCAT, O2E_NONE, 0.0f, -1, -1, 61) \
    X(62, 0, O2T_CXKEEP, O2E_NONE, 0.0f, -1, -1, 62) \
    X(63, 0, O2T_CXTO, O2E_NONE, 0.0f, 2, -1, 63) \
    X(64, 0, O2T_CXTO, O2E_NONE, 0.0f, 2, -1, 64) \
    /* O- + O2, O(3P), O2+, O+, O2(a) */ \
    X(46, 0, O2T_SCAT, O2E_NONE, 0.0f, -1, -1, 46) \
    X(47, 1, O2T_NEGDET, O2E_NONE, 0.0f, 0, -1, 47) \
    X(48, 0, O2T_NEGDET, O2E_NONE, 0.0f, 0, -1, 48) \
    X(49, 0, O2T_DEL, O2E_NONE, 0.0f, -1, -1, 49) \
    X(50, 0, O2T_DEL, O2E_NONE, 0.0f, -1, -1, 50) \
    X(51, 0, O2T_NEGDET, O2E_NONE, 0.0f, 0, -1, 51)

#define ARO2_NROWS 68

/* handler groups: G(species, target flag, sigv index used by the handler,
 * first row, rows, neutral).  neutral: the velocity the projectile collides
 * in the frame of: 0 none (electron, O-), 1 a Maxwellian at BackVel1 (the
 * argon temperature speed), 2 at BackVel2 (the O2 one).  Ar+ takes the
 * speed of its target gas; O2+ always BackVel1 and O+ always BackVel2, as
 * the handlers are written (src/cuda_mccArO2.cu:584-586, 695, 866). */
#define ARO2_GROUPS(G) \
    G(0, 0, 0, 0, 4, 0) G(0, 1, 1, 4, 1, 0) G(0, 2, 2, 5, 14, 0) G(0, 3, 3, 19, 8, 0) G(0, 4, 4, 27, 8, 0) \
    G(0, 5, 5, 35, 1, 0) G(0, 6, 6, 36, 1, 0) G(0, 7, 7, 37, 7, 0) G(0, 8, 8, 44, 2, 0) \
    G(1, 0, 23, 46, 2, 1) G(1, 1, 24, 48, 1, 2) \
    G(2, 0, 14, 49, 1, 1) G(2, 1, 15, 50, 3, 1) G(2, 2, 16, 53, 1, 1) G(2, 3, 17, 54, 1, 1) G(2, 4, 18, 55, 2, 1) \
    G(3, 0, 19, 57, 2, 2) G(3, 1, 20, 59, 1, 2) G(3, 2, 21, 60, 1, 2) G(3, 3, 22, 61, 1, 2) \
    G(4, 0, 9, 62, 2, 0) G(4, 1, 10, 64, 1, 0) G(4, 2, 11, 65, 1, 0) G(4, 3, 12, 66, 1, 0) G(4, 4, 13, 67, 1, 0)

#define ARO2_NGROUPS 25
#define ARO2_GROUP0 {0, 9, 11, 16, 20}
#define ARO2_NGROUP {9, 2, 5, 4, 5}

/* target test: T(species, flag, sigv index, kind, index) with kind 0 the
 * background gas `index` (0 Ar: BackDen1, 1 O2: BackDen2), 1 fluid species
 * `index` (Ar*, O2(a), O2(b), O(3P), O(1D)), 2 particle species `index`
 * (den * np2c / (dx dy)).  Electrons use dt / MCCn and up to MCCn draws,
 * every ion dt and one draw. */
#define ARO2_TARGETS(T) \
    T(0, 0, 0, 0, 0) T(0, 1, 1, 1, 0) T(0, 2, 2, 0, 1) T(0, 3, 3, 1, 1) T(0, 4, 4, 1, 2) T(0, 5, 5, 2, 4) \
    T(0, 6, 6, 2, 2) T(0, 7, 7, 1, 3) T(0, 8, 8, 1, 4) \
    T(1, 0, 23, 0, 0) T(1, 1, 24, 0, 1) \
    T(2, 0, 14, 1, 3) T(2, 1, 15, 0, 1) T(2, 2, 16, 1, 1) T(2, 3, 17, 1, 2) T(2, 4, 18, 0, 0) \
    T(3, 0, 19, 0, 1) T(3, 1, 20, 1, 3) T(3, 2, 21, 1, 1) T(3, 3, 22, 1, 2) \
    T(4, 0, 9, 0, 1) T(4, 1, 10, 1, 3) T(4, 2, 11, 2, 2) T(4, 3, 12, 2, 3) T(4, 4, 13, 1, 1)

#define ARO2_NTARGETS 25
#define ARO2_NSP 5
#define ARO2_NSIGV 25
/* species that use dt / MCCn and up to MCCn target draws (bit per species) */
#define ARO2_SUBCYCLED 0x1
/* the O- species, which electron detachment removes from */
#define ARO2_NEG 4
#define ARO2_VICTIM_TAG ((4u << 8) | 2u)

/* the same facts of the O2 chemistry in this form */
#define O2_SUBCYCLED 0x3
#define O2_NEG 3

/* largest table sizes of the two oxygen chemistries */
#define CHEM_MAXSP 5
#define CHEM_MAXROWS 68
#define CHEM_MAXGROUPS 25
#define CHEM_MAXTARGETS 25
#define CHEM_MAXSIGV 25

#endif
