// Oxygen and argon/oxygen cross sections on the log10-energy grid.
//
// The oxygen set is J. T. Gudmundsson's analytic fits to measured and
// computed cross sections (references per fit below), as the reference
// evaluates them in src/PhysicsData.c:787-1578, assembled into the
// 81-row tables of Oxygen_CrossSectionSET (src/PhysicsData.c:1719-2415,
// 58 tabulated reactions + 9 rate constants) and ArO2_CrossSectionSET
// (src/PhysicsData.c:2416-3275, 68 + 20), and the per-(projectile, target)
// maxima of sigma*v of Set_NullCollisionTime_cuda (src/cuda_mcc.cu:291-501).
//
// Template parameter R is the table precision.  With R = float every fit
// reproduces the reference's arithmetic bit for bit: the energy comes in as
// float, the fit's polynomial runs in double with the reference's term
// order, the powers of E that the reference keeps in float temporaries are
// rounded to float, and the result is rounded to float.  With R = double
// the same forms run without the float roundings.
#pragma once
#include <cmath>
#include <initializer_list>

namespace pic {
namespace oxy {

// Sum of c[k] * x^(n-k), highest power first.  Each term is formed as
// c * x * x * ... left to right and the terms are added left to right,
// which is the evaluation order of the fits as written.
inline double xpoly(double x, std::initializer_list<double> c) {
    const int n = (int)c.size() - 1;
    double acc = 0.0;
    int k = 0;
    for (double ck : c) {
        double t = ck;
        for (int j = 0; j < n - k; ++j) t *= x;
        acc = k ? acc + t : t;
        ++k;
    }
    return acc;
}

// The same sum with the powers of E kept in R: E^2 = E*E, E^k = E*E^(k-1),
// each rounded to R, and the coefficient products and sums in double.
template <typename R>
inline double epoly(R E, std::initializer_list<double> c) {
    const int n = (int)c.size() - 1;
    R p[8];
    p[1] = E;
    for (int k = 2; k <= n; ++k) p[k] = E * p[k - 1];
    double acc = 0.0;
    int k = 0;
    for (double ck : c) {
        const int pw = n - k;
        const double t = pw ? ck * (double)p[pw] : ck;
        acc = k ? acc + t : t;
        ++k;
    }
    return acc;
}

// 10^poly(log10 E)
template <typename R>
inline R pow10_log(R E, std::initializer_list<double> c) {
    return (R)std::pow(10.0, xpoly(std::log10((double)E), c));
}

// polynomial in E stored in R, scaled by 1e-20 m^2 in R, optionally
// clamped at zero
template <typename R>
inline R epoly20(R E, std::initializer_list<double> c, bool clamp) {
    R t = (R)epoly<R>(E, c);
    t = (R)((double)t * 1e-20);
    if (clamp && t <= 0.0) t = 0;
    return t;
}

// Lindsay et al. charge-exchange form (O+ + O2), in units of 1e-20 m^2
// before the final scale
inline double lindsay(double E) {
    const double a = 1 - std::exp(-0.250 / E);
    const double b = 2.85 + 0.103 * std::log(E);
    const double c = 1 - std::exp(-E / 0.144);
    return 2.16 * a * a + b * b * c * c * c * c;
}

enum Fit {
    F_EL,       // e + O2 elastic (Itikawa 2009 fit)
    F_ROT,      // e + O2 rotational excitation
    F_VIB1,     // vibrational v=1 (0.19 eV)
    F_VIB2,     // v=2 (0.38 eV)
    F_VIB3,     // v=3 (0.57 eV)
    F_VIB4,     // v=4 (0.75 eV)
    F_SD,       // a 1Delta_g metastable (0.977 eV)
    F_SS,       // b 1Sigma_g+ metastable (1.627 eV)
    F_L1,       // 4.05 eV excitation (Shyn and Sweeney; Green et al.)
    F_L2,       // dissociation 6.0 eV
    F_L3,       // dissociation 8.4 eV
    F_L4,       // dissociation 10.0 eV
    F_ION,      // ionisation e + O2 -> O2+ (Krishnakumar and Srivastava)
    F_LE,       // dissociative excitation 14.7 eV
    F_DA,       // dissociative attachment (Rapp and Briglia)
    F_PD,       // polar dissociation e + O2 -> O+ + O- + e
    F_DISSIZ,   // dissociative ionisation e + O2 -> O+ + O + 2e
    F_EO,       // e + O elastic (Itikawa and Ichimura; Thomas and Nesbet)
    F_E1D,      // O(3P) -> O(1D) (Laher and Gilmore)
    F_E1S,      // O(3P) -> O(1S)
    F_E3P0,     // O(3P) -> 3P0
    F_E5S0,     // O(3P) -> 5S0
    F_E3S0,     // O(3P) -> 3S0
    F_IO,       // e + O ionisation (Kim and Desclaux)
    F_ED,       // electron-impact detachment e + O- (Vejby-Christensen et al.)
    F_NED,      // detachment O- + O2 (Comer and Schulz)
    F_NEGOD,    // detachment O- + O (Belostotsky et al.)
    F_DR,       // dissociative recombination e + O2+ (Mul and McGowan; Peverall et al.)
    F_MN,       // mutual neutralisation O- + O2+ (Olson, scaled)
    F_MNO,      // mutual neutralisation O- + O+
    F_CE,       // O2+ + O2 charge exchange (Ellis et al.; Baer et al.; Wilcox and Moran)
    F_CEOO,     // O+ + O2 charge exchange (Lindsay et al.)
    F_CEO,      // O+ + O charge exchange (Stebbings et al.)
    F_CEO2O,    // O2+ + O charge exchange
    F_FRAG,     // O2+ + O2 fragmentation
    F_NEE,      // O- + O2 scattering (Muschlitz)
    F_NEUE,     // O + O2 scattering (Brunetti et al.)
    F_POE,      // O2+ + O2 scattering (half the charge exchange)
    F_POOE,     // O+ + O2 scattering (half the Lindsay charge exchange)
    F_O2O2,     // O2 + O2 gas-kinetic
    F_OO,       // O + O gas-kinetic
    N_FIT
};

// sigma(E) in m^2 of fit f (src/PhysicsData.c:787-1578)
template <typename R>
R fit(int f, R E) {
    switch (f) {
    case F_EL:
        if (0.0 <= E && E <= 5.13) return pow10_log<R>(E, {-0.001478, -0.0354, 0.1556, -19.2254});
        return pow10_log<R>(E, {0.0933, -0.7463, 1.3218, -19.7293});
    case F_ROT:
        if (.07 <= E && E <= 1.67) return epoly20<R>(E, {-.0859, .4233, -.7366, .5205, -.1537, .0604, -.0022}, false);
        return 0;
    case F_VIB1:
        if (.19 <= E && E <= 1.0)
            return epoly20<R>(E, {-1.508395, 6.521786, -9.574636, 5.092031, -0.41602, -0.066398}, true);
        if (1.0 < E && E <= 1.67) return (R)(-7.2193e-22 * ((double)E - 1.67));
        if (4.0 <= E && E <= 5.5) return (R)(4.572852e-22 * ((double)E - 4.0));
        if (5.5 < E && E <= 16.0)
            return epoly20<R>(E, {1.0835e-6, -9.229e-5, 0.0030853, -0.050981, 0.427934, -1.6682, 2.3919}, true);
        if (16.0 < E && E <= 25.0) return (R)(-4.098144e-23 * ((double)E - 25.0));
        return 0;
    case F_VIB2:
        if (.38 <= E && E <= 1.67) return epoly20<R>(E, {-0.606022, 3.157773, -5.933895, 4.664064, -1.233443}, true);
        if (4.0 <= E && E <= 14.5)
            return epoly20<R>(E, {-3.1339358e-6, 2.0994236e-4, -.00503577, 0.0515, -0.2074798, 0.279}, true);
        if (14.5 < E && E <= 25.0) return (R)(-1.71326e-23 * ((double)E - 25.0));
        return 0;
    case F_VIB3:
        if (.57 <= E && E <= 1.67) return epoly20<R>(E, {-.055083, .12457, -.057531}, true);
        if (4.0 <= E && E <= 15.0)
            return epoly20<R>(E, {-7.969385e-6, 4.78119632e-4, -.0107124, 0.1095564, -0.4962553, 0.80444}, true);
        if (15.0 < E && E <= 20.0) return (R)(-1.76e-23 * ((double)E - 20.0));
        return 0;
    case F_VIB4:
        if (.75 <= E && E <= 0.85) return (R)(2.795e-25 * ((double)E - 0.75));
        if (.85 <= E && E <= 1.67) return epoly20<R>(E, {-0.049346, 0.16616, -0.174061, 0.058213}, true);
        if (6.0 <= E && E <= 15.0) return epoly20<R>(E, {-1.3846154e-5, 8.8449e-4, -0.020271, 0.19111, -0.589505}, true);
        return 0;
    case F_SD:
        if (0.977 <= E && E <= 10.5)
            return epoly20<R>(E, {-3.0913e-6, 1.436e-4, -.0022876, .0133286, -.0100266, -.0015636}, true);
        if (10.5 < E && E <= 45.0) return epoly20<R>(E, {-1.0959e-6, 1.349e-4, -.005984, .1079}, false);
        if (45 < E && E <= 100.0) return (R)((double)(R)(-2.18e-4 * (double)E + 2.18e-2) * 1e-20);
        return 0;
    case F_SS:
        if (1.627 <= E && E <= 25.0)
            return epoly20<R>(E, {5.34245e-8, -4.7117e-6, 1.581e-4, -2.4783e-3, 1.70373e-2, -2.2343e-2}, true);
        if (25.0 < E && E <= 45.0) return epoly20<R>(E, {5.445e-7, -5.674e-5, 1.7502e-3, -1.0375e-2}, false);
        if (45 < E && E <= 100.0) return (R)((double)(R)(-5.64e-5 * (double)E + 5.64e-3) * 1e-20);
        return 0;
    case F_L1:
        if (4.05 < E && E <= 10) return (R)xpoly((double)E, {1.1933e-23, -2.4345e-22, 1.7726e-21, -3.9778e-21});
        if (10 < E && E < 40) return pow10_log<R>(E, {-16.7249, 79.2236, -143.5578, 116.2163, -56.0342});
        return 0;
    case F_L2:
        if (6.0 <= E && E <= 20.) return epoly20<R>(E, {-1.1894e-4, 6.8655e-3, -.143425, 1.26276, -3.7338513}, true);
        if (20. < E && E <= 100.) return epoly20<R>(E, {9.9341e-6, -1.7857e-3, 7.924e-2}, true);
        return 0;
    case F_L3:
        if (8.4 <= E && E <= 9.4) {
            R t = (R)((double)E - 8.4);
            t = (R)((double)t * 1e-20);
            return t <= 0.0 ? (R)0 : t;
        }
        if (9.4 < E && E <= 100.) return epoly20<R>(E, {-1.08852e-4, 1.10145e-2, .92302246}, true);
        if (100. < E) return (R)9.4e-21;
        return 0;
    case F_L4:
        if (10. <= E && E <= 100.)
            return epoly20<R>(E, {1.1656e-9, -3.1555e-7, 1.8544e-5, 9.464e-4, -.0110422}, true);
        if (100. < E) return (R)7e-22;
        return 0;
    case F_ION:
        if (12.06 <= E && E <= 180.)
            return pow10_log<R>(E, {-2.5050, 21.1338, -52.8155, -37.2837, 434.1281, -874.6800, 772.1372, -283.5866});
        if (180 < E) return pow10_log<R>(E, {-0.2981, 1.0264, -20.5514});
        return 0;
    case F_LE:
        if (14.7 <= E && E <= 100.) return epoly20<R>(E, {1.21665e-7, -3.0483e-5, 2.51713e-3, -.030335}, true);
        if (100. < E) return (R)3.8e-22;
        return 0;
    case F_DA:
        if (4.2 <= E && E <= 10)
            return pow10_log<R>(E, {5.18296e3, -2.091608e4, 3.361400e4, -2.694881e4, 1.080299e4, -1.758524e3});
        return 0;
    case F_PD:
        if (E < 15) return 0;
        if (E < 50) return pow10_log<R>(E, {-34.9127, 243.9555, -634.9551, 729.7911, -334.66});
        return (R)8.419e-23;
    case F_DISSIZ:
        if (18.73 <= E && E <= 150)
            return pow10_log<R>(E, {5.1188, -53.1537, 220.3817, -457.3868, 476.8381, -220.5401});
        if (150 < E) return pow10_log<R>(E, {-0.9987, 4.3037, -24.6189});
        return 0;
    case F_EO:
        return pow10_log<R>(E, {-0.2637, 0.3825, -19.2932});
    case F_E1D:
        if (E < 1.96) return 0;
        if (E < 200) return pow10_log<R>(E, {-2.6057, 22.3829, -76.8464, 134.4404, -127.0659, 61.2974, -32.2819});
        return 0;
    case F_E1S:
        if (E < 4.18) return 0;
        if (E < 200) return pow10_log<R>(E, {-3.8507, 36.2079, -138.9151, 277.8211, -305.7224, 175.3352, -62.3533});
        return 0;
    case F_E3P0:
        if (E < 15.65) return 0;
        if (E < 200) return pow10_log<R>(E, {-5.8339, 45.9290, -136.0395, 178.6948, -108.4334});
        return 0;
    case F_E5S0:
        if (E < 9.14) return 0;
        if (E < 200) return pow10_log<R>(E, {-4.6176, 34.2631, -94.4228, 111.5283, -69.0001});
        return 0;
    case F_E3S0:
        if (E < 9.51) return 0;
        if (E < 200) return pow10_log<R>(E, {-3.3173, 22.9180, -58.4781, 64.9565, -47.4447});
        return 0;
    case F_IO:
        if (13.62 < E && E < 30) return pow10_log<R>(E, {49.7958, -215.5521, 313.4800, -173.5139});
        if (30 <= E && E <= 5000)
            return pow10_log<R>(E, {-0.0393, 0.6706, -4.7600, 18.1277, -39.5597, 47.0034, -43.4784});
        return 0;
    case F_ED:
        if (E < 1.465) return 0;
        if (E < 20) return (R)xpoly((double)E, {7.2763e-24, -4.6446e-22, 9.703e-21, -1.3458e-20});
        return (R)5.4193e-20;
    case F_NED:
        if (E < 1.465) return 0;
        if (E < 15) return pow10_log<R>(E, {1.8184, -5.6529, 6.2460, -22.2561});
        return (R)1.694e-20;
    case F_NEGOD: {
        if (E == 0) E = (R)0.0001;
        if (E < 0.184) return (R)(2.1034e-19 / std::sqrt((double)E));
        return (R)4.9e-19;
    }
    case F_DR:
        if (E == 0) E = (R)0.0001;
        return pow10_log<R>(E, {-1.0993, -19.6146});
    case F_MN:
        if (E == 0) E = (R)0.0001;
        if (E < 100) return pow10_log<R>(E, {0.1544, -0.5873, -17.6338});
        return (R)6.4472e-19;
    case F_MNO:
        if (E == 0) E = (R)0.0001;
        if (E < 100) return pow10_log<R>(E, {0.1544, -0.5873, -17.7410});
        return (R)5.0368e-19;
    case F_CE:
        if (E == 0) E = (R)0.0001;
        return pow10_log<R>(E, {0.0503, -0.3502, -18.5476});
    case F_CEOO:
        return (R)(lindsay((double)E) * 1e-20);
    case F_CEO: {
        if (E == 0) E = (R)0.0001;
        const double s = 5.95e-10 - 0.63e-10 * std::log((double)E);
        return (R)(s * s);
    }
    case F_CEO2O:
        if (E < 1.42) return 0;
        if (E < 82.035) return (R)(3.9663e-20 * (1 - 1.42 / (double)E));
        return pow10_log<R>(E, {0.1476, -19.6917});
    case F_FRAG:
        if (E <= 13.8) return 0;
        return (R)(4.9e-19 * (1 - 13.8 / (double)E));
    case F_NEE:
        if (E <= 300) return (R)(-3.101e-22 * (double)E + 8.6525e-21 * std::sqrt((double)E) + 3.2995e-20);
        return (R)9e-20;
    case F_NEUE:
        if (E == 0) E = (R)0.0001;
        return (R)(1.4166e-20 / std::sqrt((double)E) + 8.8358e-20);
    case F_POE:
        if (E == 0) E = (R)0.0001;
        return (R)(0.5 * std::pow(10.0, xpoly(std::log10((double)E), {0.0503, -0.3502, -18.5476})));
    case F_POOE:
        return (R)(0.5 * lindsay((double)E) * 1e-20);
    case F_O2O2:
        return (R)4.9e-19;
    case F_OO:
        return (R)2.9e-19;
    default:
        return 0;
    }
}

// One tabulated reaction: threshold (minus a metastable's excitation
// energy when the target is excited) and how sigma is formed.
enum Kind {
    K_FIT,        // Flag * fit(E + shift)
    K_FIT_X35,    // Flag * 3.5 * fit(E + shift)  (O2(a) dissociative attachment)
    K_VIB,        // Flag * (v1 + v2 + v3 + v4)(E)
    K_DATA,       // Flag * FUNC_decay(data set a, E)
    K_DATA_DIFF   // Flag * FUNC_decay(a, E) - Flag * FUNC_decay(b, E)
};
struct Row {
    double th;   // threshold before the shift (eV)
    int shift;   // 0 none, 1: O2(a) 0.977 eV, 2: O2(b) 1.627 eV, 3: O(1D) 1.96 eV
    int kind;
    int a, b;    // fit index, or data set number(s)
};
constexpr float kShift[4] = {0.0f, 0.977f, 1.627f, 1.96f};

// Oxygen_CrossSectionSET, reactions 0-57 (src/PhysicsData.c:1748-2404)
constexpr Row kO2Rows[58] = {
    {0.0, 0, K_FIT, F_EL, 0},       {0.07, 0, K_FIT, F_ROT, 0},     {0.19, 0, K_VIB, 0, 0},
    {0.977, 0, K_FIT, F_SD, 0},     {1.627, 0, K_FIT, F_SS, 0},     {4.05, 0, K_FIT, F_L1, 0},
    {4.2, 0, K_FIT, F_DA, 0},       {6.0, 0, K_FIT, F_L2, 0},       {8.4, 0, K_FIT, F_L3, 0},
    {10.0, 0, K_FIT, F_L4, 0},      {12.06, 0, K_FIT, F_ION, 0},    {14.7, 0, K_FIT, F_LE, 0},
    {15.0, 0, K_FIT, F_PD, 0},      {18.73, 0, K_FIT, F_DISSIZ, 0},
    // e + O2(a)
    {12.06, 1, K_FIT, F_ION, 0},    {4.2, 1, K_FIT_X35, F_DA, 0},   {0.0, 0, K_DATA, 10, 0},
    {4.05, 1, K_FIT, F_L1, 0},      {6.0, 1, K_FIT, F_L2, 0},       {8.4, 1, K_FIT, F_L3, 0},
    {10.0, 1, K_FIT, F_L4, 0},      {18.73, 1, K_FIT, F_DISSIZ, 0},
    // e + O2(b)
    {12.06, 2, K_FIT, F_ION, 0},    {4.2, 2, K_FIT, F_DA, 0},       {0.0, 0, K_DATA, 11, 0},
    {4.05, 2, K_FIT, F_L1, 0},      {6.0, 2, K_FIT, F_L2, 0},       {8.4, 2, K_FIT, F_L3, 0},
    {10.0, 2, K_FIT, F_L4, 0},      {18.73, 2, K_FIT, F_DISSIZ, 0},
    // e + O-, e + O2+, e + O(3P), e + O(1D)
    {1.465, 0, K_FIT, F_ED, 0},     {0.0, 0, K_FIT, F_DR, 0},       {0.0, 0, K_FIT, F_EO, 0},
    {1.96, 0, K_FIT, F_E1D, 0},     {4.18, 0, K_FIT, F_E1S, 0},     {9.14, 0, K_FIT, F_E5S0, 0},
    {9.51, 0, K_FIT, F_E3S0, 0},    {13.62, 0, K_FIT, F_IO, 0},     {15.65, 0, K_FIT, F_E3P0, 0},
    {13.62, 3, K_FIT, F_IO, 0},     {0.0, 0, K_DATA, 12, 0},
    // O- + O2, O(3P), O2+, O+, O2(a)
    {0.0, 0, K_FIT, F_NEE, 0},      {1.465, 0, K_FIT, F_NED, 0},    {0.0, 0, K_FIT, F_NEGOD, 0},
    {0.0, 0, K_FIT, F_MN, 0},       {0.0, 0, K_FIT, F_MNO, 0},      {0.0, 0, K_DATA, 13, 0},
    // O2+ + O(3P), O2, O2(a), O2(b)
    {1.42, 0, K_FIT, F_CEO2O, 0},   {0.0, 0, K_FIT, F_CE, 0},       {0.0, 0, K_FIT, F_POE, 0},
    {13.8, 0, K_FIT, F_FRAG, 0},    {0.0, 0, K_FIT, F_CE, 0},       {0.0, 0, K_FIT, F_CE, 0},
    // O+ + O2, O(3P), O2(a), O2(b)
    {0.0, 0, K_FIT, F_CEOO, 0},     {0.0, 0, K_FIT, F_POOE, 0},     {0.0, 0, K_FIT, F_CEO, 0},
    {0.0, 0, K_FIT, F_CEOO, 0},     {0.0, 0, K_FIT, F_CEOO, 0},
};
// rate constants of reactions 58-66 (m^3/s; src/PhysicsData.c:2405-2414)
constexpr double kO2Rates[9] = {6.90e-16, 8.00e-18, 2.56e-17, 1.00e-18, 4.20e-17,
                                1.30e-22, 2.20e-24, 5.50e-29, 2.20e-24};

// ArO2_CrossSectionSET, reactions 0-67 (src/PhysicsData.c:2446-3254)
constexpr Row kArO2Rows[68] = {
    // e + Ar, e + Ar*
    {0.0, 0, K_DATA, 1, 0},         {11.55, 0, K_DATA_DIFF, 2, 3},  {11.548, 0, K_DATA, 3, 0},
    {15.76, 0, K_DATA, 4, 0},       {4.22, 0, K_DATA, 5, 0},
    // e + O2
    {0.0, 0, K_FIT, F_EL, 0},       {0.07, 0, K_FIT, F_ROT, 0},     {0.19, 0, K_VIB, 0, 0},
    {0.977, 0, K_FIT, F_SD, 0},     {1.627, 0, K_FIT, F_SS, 0},     {4.05, 0, K_FIT, F_L1, 0},
    {4.2, 0, K_FIT, F_DA, 0},       {6.0, 0, K_FIT, F_L2, 0},       {8.4, 0, K_FIT, F_L3, 0},
    {10.0, 0, K_FIT, F_L4, 0},      {12.06, 0, K_FIT, F_ION, 0},    {14.7, 0, K_FIT, F_LE, 0},
    {15.0, 0, K_FIT, F_PD, 0},      {18.73, 0, K_FIT, F_DISSIZ, 0},
    // e + O2(a)
    {12.06, 1, K_FIT, F_ION, 0},    {4.2, 1, K_FIT_X35, F_DA, 0},   {0.0, 0, K_DATA, 10, 0},
    {4.05, 1, K_FIT, F_L1, 0},      {6.0, 1, K_FIT, F_L2, 0},       {8.4, 1, K_FIT, F_L3, 0},
    {10.0, 1, K_FIT, F_L4, 0},      {18.73, 1, K_FIT, F_DISSIZ, 0},
    // e + O2(b)
    {12.06, 2, K_FIT, F_ION, 0},    {4.2, 2, K_FIT, F_DA, 0},       {0.0, 0, K_DATA, 11, 0},
    {4.05, 2, K_FIT, F_L1, 0},      {6.0, 2, K_FIT, F_L2, 0},       {8.4, 2, K_FIT, F_L3, 0},
    {10.0, 2, K_FIT, F_L4, 0},      {18.73, 2, K_FIT, F_DISSIZ, 0},
    // e + O-, e + O2+, e + O(3P), e + O(1D)
    {1.465, 0, K_FIT, F_ED, 0},     {0.0, 0, K_FIT, F_DR, 0},       {0.0, 0, K_FIT, F_EO, 0},
    {1.96, 0, K_FIT, F_E1D, 0},     {4.18, 0, K_FIT, F_E1S, 0},     {9.14, 0, K_FIT, F_E5S0, 0},
    {9.51, 0, K_FIT, F_E3S0, 0},    {13.62, 0, K_FIT, F_IO, 0},     {15.65, 0, K_FIT, F_E3P0, 0},
    {13.62, 3, K_FIT, F_IO, 0},     {0.0, 0, K_DATA, 12, 0},
    // O- + O2, O(3P), O2+, O+, O2(a)
    {0.0, 0, K_FIT, F_NEE, 0},      {1.465, 0, K_FIT, F_NED, 0},    {0.0, 0, K_FIT, F_NEGOD, 0},
    {0.0, 0, K_FIT, F_MN, 0},       {0.0, 0, K_FIT, F_MNO, 0},      {0.0, 0, K_DATA, 13, 0},
    // O2+ + O(3P), O2, O2(a), O2(b), Ar
    {1.42, 0, K_FIT, F_CEO2O, 0},   {0.0, 0, K_FIT, F_CE, 0},       {0.0, 0, K_FIT, F_POE, 0},
    {13.8, 0, K_FIT, F_FRAG, 0},    {0.0, 0, K_FIT, F_CE, 0},       {0.0, 0, K_FIT, F_CE, 0},
    {0.0, 0, K_DATA, 14, 0},        {0.0, 0, K_DATA, 15, 0},
    // O+ + O2, O(3P), O2(a), O2(b)
    {0.0, 0, K_FIT, F_CEOO, 0},     {0.0, 0, K_FIT, F_POOE, 0},     {0.0, 0, K_FIT, F_CEO, 0},
    {0.0, 0, K_FIT, F_CEOO, 0},     {0.0, 0, K_FIT, F_CEOO, 0},
    // Ar+ + Ar, Ar+ + O2
    {0.0, 0, K_DATA, 7, 0},         {0.0, 0, K_DATA, 9, 0},         {0.0, 0, K_DATA, 16, 0},
};
// rate constants of reactions 68-87 (src/PhysicsData.c:3255-3274)
constexpr double kArO2Rates[20] = {2.00e-13, 2.70e-13, 6.90e-16, 8.00e-18, 2.56e-17, 1.00e-18, 4.20e-17,
                                   1.30e-22, 2.20e-24, 5.50e-29, 2.20e-24, 6.40e-18, 5.10e-17, 6.20e-16,
                                   3.00e-21, 8.10e-18, 4.10e-17, 1.01e-16, 1.14e-16, 1.12e-15};

// One (projectile, target) pair of the null-collision method: the table
// columns whose sum is the pair's total sigma, in the reference's summation
// groups (each group is summed left to right, then added to the running
// total; -1 closes a group, -2 ends the list), the projectile species and
// where the target density comes from (src/cuda_mcc.cu:291-489).
enum TKind { T_SP = 0, T_BG = 1, T_FG = 2 };
struct Pair {
    int proj;
    int tkind, tidx;
    int cols[20];
};
constexpr Pair kO2Pairs[20] = {
    {0, T_BG, 0, {0, 1, 2, 3, 4, -1, 5, 6, 7, 8, 9, -1, 10, 11, 12, 13, -2}},   // e + O2
    {0, T_FG, 0, {14, 15, 16, 17, 18, -1, 19, 20, 21, -2}},                     // e + O2(a)
    {0, T_FG, 1, {22, 23, 24, 25, 26, -1, 27, 28, 29, -2}},                     // e + O2(b)
    {0, T_SP, 3, {30, -2}},                                                     // e + O-
    {0, T_SP, 1, {31, -2}},                                                     // e + O2+
    {0, T_FG, 2, {32, 33, 34, 35, 36, -1, 37, 38, -2}},                         // e + O(3P)
    {0, T_FG, 3, {39, 40, -2}},                                                 // e + O(1D)
    {3, T_BG, 0, {41, 42, -2}},                                                 // O- + O2
    {3, T_FG, 2, {43, -2}},                                                     // O- + O(3P)
    {3, T_SP, 1, {44, -2}},                                                     // O- + O2+
    {3, T_SP, 2, {45, -2}},                                                     // O- + O+
    {3, T_FG, 0, {46, -2}},                                                     // O- + O2(a)
    {1, T_FG, 2, {47, -2}},                                                     // O2+ + O(3P)
    {1, T_BG, 0, {48, 49, 50, -2}},                                             // O2+ + O2
    {1, T_FG, 0, {51, -2}},                                                     // O2+ + O2(a)
    {1, T_FG, 1, {52, -2}},                                                     // O2+ + O2(b)
    {2, T_BG, 0, {53, 54, -2}},                                                 // O+ + O2
    {2, T_FG, 2, {55, -2}},                                                     // O+ + O(3P)
    {2, T_FG, 0, {56, -2}},                                                     // O+ + O2(a)
    {2, T_FG, 1, {57, -2}},                                                     // O+ + O2(b)
};
constexpr Pair kArO2Pairs[25] = {
    {0, T_BG, 0, {0, 1, 2, 3, -2}},                                             // e + Ar
    {0, T_FG, 0, {4, -2}},                                                      // e + Ar*
    {0, T_BG, 1, {5, 6, 7, 8, 9, -1, 10, 11, 12, 13, 14, -1, 15, 16, 17, 18, -2}},   // e + O2
    {0, T_FG, 1, {19, 20, 21, 22, 23, -1, 24, 25, 26, -2}},                     // e + O2(a)
    {0, T_FG, 2, {27, 28, 29, 30, 31, -1, 32, 33, 34, -2}},                     // e + O2(b)
    {0, T_SP, 4, {35, -2}},                                                     // e + O-
    {0, T_SP, 2, {36, -2}},                                                     // e + O2+
    {0, T_FG, 3, {37, 38, 39, 40, 41, -1, 42, 43, -2}},                         // e + O(3P)
    {0, T_FG, 4, {44, 45, -2}},                                                 // e + O(1D)
    {4, T_BG, 1, {46, 47, -2}},                                                 // O- + O2
    {4, T_FG, 3, {48, -2}},                                                     // O- + O(3P)
    {4, T_SP, 2, {49, -2}},                                                     // O- + O2+
    {4, T_SP, 3, {50, -2}},                                                     // O- + O+
    {4, T_FG, 1, {51, -2}},                                                     // O- + O2(a)
    {2, T_FG, 3, {52, -2}},                                                     // O2+ + O(3P)
    {2, T_BG, 1, {53, 54, 55, -2}},                                             // O2+ + O2
    {2, T_FG, 1, {56, -2}},                                                     // O2+ + O2(a)
    {2, T_FG, 2, {57, -2}},                                                     // O2+ + O2(b)
    {2, T_BG, 0, {58, 59, -2}},                                                 // O2+ + Ar
    {3, T_BG, 1, {60, 61, -2}},                                                 // O+ + O2
    {3, T_FG, 3, {62, -2}},                                                     // O+ + O(3P)
    {3, T_FG, 1, {63, -2}},                                                     // O+ + O2(a)
    {3, T_FG, 2, {64, -2}},                                                     // O+ + O2(b)
    {1, T_BG, 0, {65, 66, -2}},                                                 // Ar+ + Ar
    {1, T_BG, 1, {67, -2}},                                                     // Ar+ + O2
};

}  // namespace oxy
}  // namespace pic
