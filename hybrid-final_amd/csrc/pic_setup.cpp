// Host-side case setup; see pic_setup.h for the reference lines restated.
#include "pic_setup.h"

#include <cfloat>
#include <cmath>
#include <algorithm>
#include <cstring>
#include <type_traits>

#include "detmath.h"
#include "json.h"
#include "xsec_data.inc"
#include "coll_keys.inc"
#include "xsec_oxygen.h"

namespace pic {

namespace {
constexpr double kEps0 = 8.8542e-12;
constexpr double kCQ = 1.602e-19;
constexpr double kAMU = 1.66053e-27;
constexpr double kNperTorr = 8.3221e20;
constexpr double kDblMin = 1e-200;
constexpr double kPi = 3.14159265358979323846;

// per gas type: the deck object with the reaction switches, their names
// and count, and each reaction's mass-ratio operands
const char *const kCollObj[3] = {pic_coll_obj0, pic_coll_obj1, pic_coll_obj2};
const char *const *const kCollKeys[3] = {pic_coll_keys0, pic_coll_keys1, pic_coll_keys2};
const int kCollN[3] = {pic_coll_n0, pic_coll_n1, pic_coll_n2};
const signed char (*const kCollMofM[3])[4] = {pic_coll_mofm0, pic_coll_mofm1, pic_coll_mofm2};
}  // namespace

std::string reaction_names(int gas) {
    std::string out = kCollObj[gas];
    for (int r = 0; r < kCollN[gas]; ++r) out += std::string("\n") + kCollKeys[gas][r];
    return out;
}

int parse_deck(const std::string &text, Deck &d, std::string &err) {
    JValue root;
    try {
        root = JParser(text).parse();
    } catch (const std::exception &e) {
        err = e.what();
        return -2;
    }
    const JValue &geo = root.object("Geometry");
    const JValue &ss = geo.object("SystemSpec");
    d.xlength = ss.number("X_length(m)");
    d.ylength = ss.number("Y_length(m)");
    d.zlength = ss.number("Z_length(m)");
    d.ngx = (int)ss.number("NumGridx");
    d.ngy = (int)ss.number("NumGridy");
    for (const JValue &b : geo.array("BoundaryCondition")) {
        Line l;
        l.x0 = (int)b.number("X0"); l.x1 = (int)b.number("X1");
        l.y0 = (int)b.number("Y0"); l.y1 = (int)b.number("Y1");
        l.bc = (int)b.number("B.C"); l.temp = b.number("Temp");
        d.bnd.push_back(l);
    }
    for (const JValue &c : geo.array("ConductorSpec")) {
        Line l;
        l.id = (int)c.number("M_ID");
        l.x0 = (int)c.number("X0"); l.x1 = (int)c.number("X1");
        l.y0 = (int)c.number("Y0"); l.y1 = (int)c.number("Y1");
        l.temp = c.number("Temp"); l.a = c.number("R(Ohm)"); l.b = c.number("L(H)"); l.c = c.number("C(F)");
        d.cond.push_back(l);
    }
    for (const JValue &s : geo.array("Source")) {
        Src x;
        x.mid = (int)s.number("M_ID"); x.dc = s.number("DC(V)"); x.pow = s.number("Power(W)");
        x.ac = s.number("AC(V)"); x.freq = s.number("Freq(1/s)"); x.phase = s.number("Phase(deg)");
        d.src.push_back(x);
    }
    for (const JValue &e : geo.array("DielectricSpec")) {
        Line l;
        l.id = (int)e.number("M_ID");
        l.x0 = (int)e.number("X0"); l.x1 = (int)e.number("X1");
        l.y0 = (int)e.number("Y0"); l.y1 = (int)e.number("Y1");
        l.a = e.number("Epsilon");
        d.diel.push_back(l);
    }
    const JValue &gas = root.object("GasSpecies");
    d.gas = (int)gas.number("Type(0:Ar,1:O2,2:Ar/O2)");
    d.total_pres = gas.number("TotalPres(Torr)");
    const JValue &bg = gas.object("Background");
    d.bg_ratio[0] = bg.object("Argon").number("Ratio");
    d.bg_temp[0] = bg.object("Argon").number("Temp(eV)");
    d.bg_ratio[1] = bg.object("Oxygen").number("Ratio");
    d.bg_temp[1] = bg.object("Oxygen").number("Temp(eV)");
    const JValue &ns = gas.object("NeutralSpecies");
    d.fl_loadtype = (int)ns.number("Loadtype");
    std::vector<const char *> fln, spn;
    if (d.gas == 0) { fln = {"Ar*"}; spn = {"Electron", "Ar+"}; }
    else if (d.gas == 1) { fln = {"O2A", "O2B", "OP", "OD"}; spn = {"Electron", "O2+", "O+", "O-"}; }
    else if (d.gas == 2) { fln = {"Ar*", "O2A", "O2B", "OP", "OD"}; spn = {"Electron", "Ar+", "O2+", "O+", "O-"}; }
    for (size_t k = 0; k < fln.size(); ++k) {
        const JValue &o = ns.object(fln[k]);
        d.fl_dens[k] = o.number("Density"); d.fl_temp[k] = o.number("Temp(eV)"); d.fl_quench[k] = o.number("Quenching");
    }
    const JValue &cs = gas.object("ChargeSpecies");
    d.sp_loadtype = (int)cs.number("Loadtype");
    const JValue &lp = cs.object("LoadPosition(m)");
    d.sp_x0 = lp.number("X0"); d.sp_x1 = lp.number("X1"); d.sp_y0 = lp.number("Y0"); d.sp_y1 = lp.number("Y1");
    for (size_t k = 0; k < spn.size(); ++k) {
        const JValue &o = cs.object(spn[k]);
        d.sp_dens[k] = o.number("Density"); d.sp_temp[k] = o.number("Temp(eV)"); d.sp_np2c[k] = o.number("np2c");
        d.sp_ratio[k] = (int)o.number("Ratio"); d.sp_maxnp[k] = (int)o.number("Max_np");
    }
    const JValue &sm = root.object("SimulationMethod");
    d.dt_pic = (int)sm.number("TimeStep_PIC");
    d.dt_conti = (int)sm.number("TimeStep_Conti");
    d.pcg_tol = sm.number("PCGMarginOfError");
    d.n_ave = (int)sm.number("AverageIter");
    d.n_smt = (int)sm.number("DensitySmoothing");
    d.np_limit = (int)sm.number("ParticleLimit");
    d.seed = (int64_t)sm.number("LoaderSeed");
    if (d.seed == 0) d.seed = 1;
    d.rng_seed = (int64_t)sm.number("PhiloxSeed");
    d.rf_cycles = (int)sm.number("RFCycles");
    d.o2_ion_coll = (int)sm.number("O2IonCollisions");
    if (d.gas >= 0 && d.gas <= 2) {
        const JValue &ac = root.object("Collision_On_Off").object(kCollObj[d.gas]);
        for (int r = 0; r < kCollN[d.gas]; ++r) d.coll_flag[r] = ac.number(kCollKeys[d.gas][r]);
    }
    return 0;
}

// ---------------------------------------------------------------- BandLU
template <typename R>
int BandLU::factor(int n, const std::vector<int> &Ai, const std::vector<int> &Aj, const std::vector<R> &val) {
    N = n;
    bw = 0;
    for (int r = 0; r < n; ++r)
        for (int k = Ai[r]; k < Ai[r + 1]; ++k) {
            int dc = std::abs(Aj[k] - r);
            if (dc > bw) bw = dc;
        }
    const int W = 2 * bw + 1;
    a.assign((size_t)n * W, 0.0);
    auto at = [&](int rr, int cc) -> double & { return a[(size_t)rr * W + (cc - rr + bw)]; };
    for (int r = 0; r < n; ++r)
        for (int k = Ai[r]; k < Ai[r + 1]; ++k) at(r, Aj[k]) += (double)val[k];
    for (int k = 0; k < n; ++k) {
        double piv = at(k, k);
        int rmax = k + bw < n - 1 ? k + bw : n - 1;
        for (int i2 = k + 1; i2 <= rmax; ++i2) {
            double f = at(i2, k);
            if (f == 0.0) continue;
            f /= piv;
            at(i2, k) = f;
            for (int j2 = k + 1; j2 <= rmax; ++j2) at(i2, j2) -= f * at(k, j2);
        }
    }
    return 0;
}

void BandLU::solve(const double *b, double *x) const {
    const int W = 2 * bw + 1;
    auto at = [&](int rr, int cc) { return a[(size_t)rr * W + (cc - rr + bw)]; };
    for (int r = 0; r < N; ++r) {
        double s = b[r];
        int k0 = r - bw > 0 ? r - bw : 0;
        for (int k = k0; k < r; ++k) s -= at(r, k) * x[k];
        x[r] = s;
    }
    for (int r = N - 1; r >= 0; --r) {
        double s = x[r];
        int k1 = r + bw < N - 1 ? r + bw : N - 1;
        for (int k = r + 1; k <= k1; ++k) s -= at(r, k) * x[k];
        x[r] = s / at(r, r);
    }
}

namespace {

template <typename R>
constexpr bool is_f32() { return std::is_same<R, float>::value; }

// ------------------------------------------------ constants (start.c, Efield.c)
template <typename R>
int constants(const Deck &d, Setup<R> &s, std::string &err) {
    s.ngx = d.ngx; s.ngy = d.ngy;
    if (d.bnd.size() < 4) { err = "\"BoundaryCondition\" is too small!(BoundaryNum>=4)"; return -10; }
    if (d.cond.size() < 2) { err = "\"Conductor\" is too small!(CondNum>=2)"; return -11; }
    if (d.cond.size() > 100) { err = "\"Conductor\" is too large!(CondNum<100)"; return -11; }
    if (d.ngx < 2 || d.ngy < 2) { err = "NumGridx/NumGridy must be >= 2"; return -12; }
    for (auto &b : d.bnd)
        if (b.x0 > b.x1 || b.x1 >= d.ngx || b.y0 > b.y1 || b.y1 >= d.ngy || b.bc > 4) {
            err = "bad line in \"BoundaryCondition\""; return -12;
        }
    {
        std::vector<int> ids;
        for (auto &c : d.cond) {
            if (c.id == 0) { err = "Error : \"M_ID\" > 0"; return -13; }
            if (ids.empty()) {
                if (c.id != 1) { err = "Error : First \"M_ID\" must be 1"; return -13; }
                ids.push_back(c.id);
                continue;
            }
            if (ids.back() > c.id) { err = "Error : \"M_ID\" must be written in order."; return -13; }
            bool dup = false;
            for (int m : ids) dup |= (m == c.id);
            if (!dup) ids.push_back(c.id);
        }
        s.CondNUMR = (int)ids.size();
    }
    for (auto &c : d.cond)
        if (c.x0 > c.x1 || c.x1 >= d.ngx || c.y0 > c.y1 || c.y1 >= d.ngy) { err = "bad line in \"ConductorSpec\""; return -14; }
    float max_freq = 0.0f;
    double max_freq_d = 0.0;
    for (auto &x : d.src) {
        bool ok = false;
        for (auto &c : d.cond) ok |= (c.id == x.mid);
        if (!ok) { err = "Source is no matching Conductor"; return -15; }
        if ((float)x.freq != 0.0f && (float)x.freq > max_freq) max_freq = (float)x.freq;
        if (x.freq > max_freq_d) max_freq_d = x.freq;
        if ((float)x.dc == 0.0f && (float)x.pow == 0.0f && (float)x.ac == 0.0f) { err = "Source is error(delete Line)"; return -15; }
    }
    if (max_freq <= 0.0f) { err = "no source frequency"; return -16; }
    if (d.dt_pic == 0) { err = "ERROR! - \"TimeStep_PIC\" = 0"; return -17; }
    if (d.gas < 0 || d.gas > 2) { err = "\"Type\" is error in GasSpecies."; return -18; }

    s.Gsize = s.ngx * s.ngy;
    s.ncx = s.ngx - 1; s.ncy = s.ngy - 1;
    s.Csize = s.ncx * s.ncy;
    R xl = (R)d.xlength, yl = (R)d.ylength;
    s.zl = (R)d.zlength;
    s.dx = xl / s.ncx;
    s.dy = yl / s.ncy;
    s.idx = (R)1 / s.dx;
    s.idy = (R)1 / s.dy;
    s.idxy = (R)(1.5 / (double)(s.dx + s.dy));
    s.hdx = (R)(0.5 / (double)s.dx);
    s.hdy = (R)(0.5 / (double)s.dy);
    if (is_f32<R>()) s.dt = (R)(1.0 / (double)max_freq / (double)(float)d.dt_pic);
    else s.dt = (R)(1.0 / max_freq_d / (double)d.dt_pic);
    s.dt_dx = s.dt / s.dx;
    s.dt_dy = s.dt / s.dy;

    // gas species (src/start.c:303-728)
    const R me = (R)(5.485799486e-4 * kAMU), mAr = (R)(39.950 * kAMU), mO2 = (R)(32.0000 * kAMU),
            mO = (R)(16.000 * kAMU), qe = (R)(1.0 * kCQ);
    auto bg = [&](int b, R mass, R pres, double temp) {
        s.bg_mass[b] = mass;
        s.bg_pres[b] = pres;
        s.bg_dens[b] = (R)(kNperTorr * (double)pres / ((double)(R)temp + kDblMin));
    };
    if (d.gas == 0) {
        s.nsp = 2; s.nfsp = 1; s.nbg = 1;
        bg(0, mAr, (R)d.total_pres, d.bg_temp[0]);
        s.fl_mass[0] = mAr;
        s.mass[0] = me; s.q[0] = -qe;
        s.mass[1] = mAr; s.q[1] = qe;
    } else if (d.gas == 1) {
        s.nsp = 4; s.nfsp = 4; s.nbg = 1;
        bg(0, mO2, (R)d.total_pres, d.bg_temp[1]);
        s.fl_mass[0] = mO2; s.fl_mass[1] = mO2; s.fl_mass[2] = mO; s.fl_mass[3] = mO;
        s.mass[0] = me; s.q[0] = -qe;
        s.mass[1] = mO2; s.q[1] = qe;
        s.mass[2] = mO; s.q[2] = qe;
        s.mass[3] = mO; s.q[3] = -qe;
    } else {
        // background ratios: Ar takes 1 - (the O2 ratio) of the total pressure
        s.nsp = 5; s.nfsp = 5; s.nbg = 2;
        const R o2 = (R)d.bg_ratio[1];
        const R ar = (R)(1.0 - (double)o2);
        const R tot = (R)d.total_pres;
        bg(0, mAr, tot * ar, d.bg_temp[0]);
        bg(1, mO2, tot * o2, d.bg_temp[1]);
        s.fl_mass[0] = mAr; s.fl_mass[1] = mO2; s.fl_mass[2] = mO2; s.fl_mass[3] = mO; s.fl_mass[4] = mO;
        s.mass[0] = me; s.q[0] = -qe;
        s.mass[1] = mAr; s.q[1] = qe;
        s.mass[2] = mO2; s.q[2] = qe;
        s.mass[3] = mO; s.q[3] = qe;
        s.mass[4] = mO; s.q[4] = -qe;
    }
    for (int f = 0; f < s.nfsp; ++f) s.fl_dens[f] = (R)d.fl_dens[f];
    for (int i = 0; i < s.nsp; ++i) {
        s.np2c[i] = (R)d.sp_np2c[i];
        R temp = (R)d.sp_temp[i];
        s.q_density[i] = s.q[i] * s.np2c[i];
        s.vti[i] = (R)std::sqrt(kCQ * (double)temp / (double)s.mass[i]);
        s.qm[i] = s.q[i] / s.mass[i];
        s.Escale[i] = (R)(0.5 * (double)s.mass[i] / kCQ);
        s.Ascale[i] = s.qm[i] * s.dt;
        s.Denscale[i] = s.np2c[i] / s.dx / s.dy;
    }
    s.Efield_Flag.assign(s.CondNUMR, 0);
    for (int i = 0; i < s.CondNUMR; ++i) {
        int cnt = 0;
        for (auto &x : d.src) cnt += (x.mid == i + 1);
        s.Efield_Flag[i] = (cnt > 0 && cnt < 4) ? 1 : 0;
    }
    return 0;
}

// ------------------------------------------------ geometry (start.c:1464-1851)
template <typename R>
void geometry(const Deck &d, Setup<R> &s) {
    const int ngx = s.ngx, ngy = s.ngy, ncx = s.ncx, ncy = s.ncy;
    const R dx = s.dx, dy = s.dy, zl = s.zl;
    s.Boundary.assign(s.Gsize, 0); s.CondID.assign(s.Gsize, 0);
    s.Face.assign(s.Gsize, F_NOFACE); s.DensRegion.assign(s.Gsize, 0);
    s.Area.assign(s.Gsize, dy * zl); s.Temp.assign(s.Gsize, 0);
    s.PlasmaRegion.assign(s.Csize, 1); s.eps_r.assign(s.Csize, 1); s.eps.assign(s.Csize, 0);
    const R corner_area = (R)(0.5 * (double)(dx + dy) * (double)zl);
    for (int j = 0; j < ngy; ++j)
        for (int i = 0; i < ngx; ++i) {
            int ID = i * ngy + j;
            for (auto &b : d.bnd) {
                if (i >= b.x0 && i <= b.x1 && j >= b.y0 && j <= b.y1) {
                    if (b.bc == DIRICHLET) { s.Boundary[ID] = DIRICHLET; s.Temp[ID] = (R)b.temp; }
                    else if (s.Boundary[ID] != DIRICHLET && b.bc == NEUMANN) { s.Boundary[ID] = NEUMANN; s.Temp[ID] = (R)b.temp; }
                }
            }
        }
    for (auto &e : d.diel)
        for (int i = e.x0; i <= e.x1; ++i)
            for (int j = e.y0; j <= e.y1; ++j) {
                int ID = i * ngy + j;
                if (s.Boundary[ID] == 0) s.Boundary[ID] = DIELECTRIC;
                if (i < e.x1 && j < e.y1) { s.PlasmaRegion[i * ncy + j] = 0; s.eps_r[i * ncy + j] = (R)e.a; }
            }
    for (auto &c : d.cond)
        for (int i = c.x0; i <= c.x1; ++i)
            for (int j = c.y0; j <= c.y1; ++j) {
                int ID = i * ngy + j;
                s.CondID[ID] = c.id;
                if (s.Boundary[ID] == 0 || s.Boundary[ID] == DIELECTRIC) { s.Boundary[ID] = CONDUCTOR; s.Temp[ID] = (R)c.temp; }
                if (i < c.x1 && j < c.y1) { s.PlasmaRegion[i * ncy + j] = 0; s.eps_r[i * ncy + j] = 0; }
            }
    for (int i = 0; i < ncx; ++i)
        for (int j = 0; j < ncy; ++j)
            if (s.PlasmaRegion[i * ncy + j] == 1) {
                int ID = i * ngy + j;
                s.DensRegion[ID] = s.DensRegion[ID + 1] = s.DensRegion[ID + ngy] = s.DensRegion[ID + ngy + 1] = 1;
            }
    auto face_area = [&](int f) -> R {
        if (f == F_LEFT || f == F_RIGHT) return dy * zl;
        if (f == F_UP || f == F_DOWN) return dx * zl;
        return corner_area;
    };
    for (int i = 1; i < ngx - 1; ++i)
        for (int j = 1; j < ngy - 1; ++j) {
            int ID = i * ngy + j, CID = i * ncy + j;
            int P1 = (int)s.eps_r[CID], P2 = (int)s.eps_r[CID - ncy];
            int P3 = (int)s.eps_r[CID - 1], P4 = (int)s.eps_r[CID - ncy - 1];
            int b = s.Boundary[ID];
            if (b != DIELECTRIC && b != CONDUCTOR) continue;
            int f = -1;
            if ((P2 == 1 && P1 == 1 && P4 == 1) || (P2 != 0 && P1 != 0 && P4 != 0 && P3 == 0)) f = F_UL;
            else if ((P2 == 1 && P1 == 1 && P3 == 1) || (P2 != 0 && P1 != 0 && P3 != 0 && P4 == 0)) f = F_UR;
            else if ((P2 == 1 && P1 == 1 && P4 != 1 && P3 != 1) || (P2 != 0 && P1 != 0 && P4 == 0 && P3 == 0)) f = F_UP;
            else if ((P4 == 1 && P3 == 1 && P2 == 1) || (P2 != 0 && P4 != 0 && P3 != 0 && P1 == 0)) f = F_LL;
            else if ((P4 == 1 && P3 == 1 && P1 == 1) || (P1 != 0 && P4 != 0 && P3 != 0 && P2 == 0)) f = F_LR;
            else if ((P4 == 1 && P3 == 1 && P1 != 1 && P2 != 1) || (P4 != 0 && P3 != 0 && P1 == 0 && P2 == 0)) f = F_DOWN;
            else if ((P4 == 1 && P2 == 1 && P1 != 1 && P3 != 1) || (P4 != 0 && P2 != 0 && P1 == 0 && P3 == 0)) f = F_LEFT;
            else if ((P4 != 1 && P2 != 1 && P1 == 1 && P3 == 1) || (P4 == 0 && P2 == 0 && P1 != 0 && P3 != 0)) f = F_RIGHT;
            if (f < 0) continue;
            s.Face[ID] = f;
            s.Area[ID] = face_area(f);
        }
    // wall nodes; a CondID lookup past either end of the node array reads 0
    auto cid = [&](int n) { return (n >= 0 && n < s.Gsize) ? s.CondID[n] : 0; };
    for (int j = 1; j < ncy; ++j)
        for (int side = 0; side < 2; ++side) {
            int ID = side ? ncx * ngy + j : j;
            int in = side ? ID - ngy : ngy + ID;
            int up = cid(ID + 1), dn = cid(ID - 1), inw = cid(in);
            int wallface = side ? F_LEFT : F_RIGHT;
            if (s.Boundary[ID] == DIRICHLET) {
                s.Face[ID] = wallface; s.Area[ID] = dy * zl;
                if (up == 0 && dn != 0 && inw != 0) { s.Face[ID] = F_UP; s.Area[ID] = dx * zl; }
                else if (up != 0 && dn == 0 && inw != 0) { s.Face[ID] = F_DOWN; s.Area[ID] = dx * zl; }
                else if (up != 0 && dn != 0 && inw != 0) { s.Face[ID] = F_NOFACE; s.Area[ID] = dy * zl; }
                else if (up == 0 && dn == 0 && inw != 0) { s.Face[ID] = F_NOFACE; s.Area[ID] = dy * zl; }
            } else if (s.Boundary[ID] == NEUMANN) {
                s.Face[ID] = F_NOFACE; s.Area[ID] = dy * zl;
                if (up == 0 && dn != 0 && inw != 0) { s.Face[ID] = F_UP; s.Area[ID] = dx * zl; }
                else if (up != 0 && dn == 0 && inw != 0) { s.Face[ID] = F_DOWN; s.Area[ID] = dx * zl; }
                else if (up != 0 && dn != 0 && inw == 0) { s.Face[ID] = wallface; s.Area[ID] = dy * zl; }
            }
        }
    for (int i = 1; i < ncx; ++i)
        for (int side = 0; side < 2; ++side) {
            int ID = side ? i * ngy + ncy : i * ngy;
            int in = side ? ID - 1 : ID + 1;
            int rt = cid(ID + ngy), lf = cid(ID - ngy), inw = cid(in);
            int wallface = side ? F_DOWN : F_UP;
            if (s.Boundary[ID] == DIRICHLET) {
                s.Face[ID] = wallface; s.Area[ID] = dx * zl;
                if (rt == 0 && lf != 0 && inw != 0) { s.Face[ID] = F_RIGHT; s.Area[ID] = dy * zl; }
                else if (rt != 0 && lf == 0 && inw != 0) { s.Face[ID] = F_LEFT; s.Area[ID] = dy * zl; }
                else if (rt != 0 && lf != 0 && inw != 0) { s.Face[ID] = F_NOFACE; s.Area[ID] = dy * zl; }
                else if (rt == 0 && lf == 0 && inw != 0) { s.Face[ID] = F_NOFACE; s.Area[ID] = dy * zl; }
            } else if (s.Boundary[ID] == NEUMANN) {
                s.Face[ID] = F_NOFACE; s.Area[ID] = dy * zl;
                if (rt == 0 && lf != 0 && inw != 0) { s.Face[ID] = F_RIGHT; s.Area[ID] = dy * zl; }
                else if (rt != 0 && lf == 0 && inw != 0) { s.Face[ID] = F_LEFT; s.Area[ID] = dy * zl; }
                else if (rt != 0 && lf != 0 && inw == 0) { s.Face[ID] = wallface; s.Area[ID] = dx * zl; }
            }
        }
    {
        const int corners[4] = {0, ncx * ngy, ncy, ncx * ngy + ncy};
        const int faces[4] = {F_UR, F_UL, F_LR, F_LL};
        for (int k = 0; k < 4; ++k) {
            int ID = corners[k];
            if (s.Boundary[ID] == DIRICHLET && cid(ID + 1) == 0) { s.Face[ID] = faces[k]; s.Area[ID] = corner_area; }
        }
    }
    // padded structure map: 0 plasma, 1..99 dielectric, >=100 conductor, <0 walls
    const int W = ncy + 2;
    s.StructIdx.assign((size_t)(ncx + 2) * W, 0);
    auto SI = [&](int a, int b) -> int & { return s.StructIdx[(size_t)a * W + b]; };
    for (int i = 0; i < ncx; ++i)
        for (int j = 0; j < ncy; ++j) {
            SI(i + 1, j + 1) = 0;
            for (auto &e : d.diel)
                if (i >= e.x0 && i < e.x1 && j >= e.y0 && j < e.y1) SI(i + 1, j + 1) = e.id;
            for (auto &c : d.cond)
                if (i >= c.x0 && i < c.x1 && j >= c.y0 && j < c.y1) SI(i + 1, j + 1) = 100 + c.id;
        }
    const auto &B = s.Boundary;
    for (int i = 1; i < ngx; ++i) {
        if ((B[(i - 1) * ngy] == NEUMANN || B[i * ngy] == NEUMANN) && SI(i, 1) == 0) SI(i, 0) = -B[(i - 1) * ngy] - 2;
        else SI(i, 0) = SI(i, 1) ? SI(i, 1) : -1;
        if ((B[(i - 1) * ngy + ncy] == NEUMANN || B[i * ngy + ncy] == NEUMANN) && SI(i, ngy - 1) == 0)
            SI(i, ngy) = -B[(i - 1) * ngy + ncy] - 6;
        else SI(i, ngy) = SI(i, ngy - 1) ? SI(i, ngy - 1) : -1;
    }
    for (int j = 1; j < ngy; ++j) {
        if ((B[j - 1] == NEUMANN || B[j] == NEUMANN) && SI(1, j) == 0) SI(0, j) = -B[j - 1];
        else SI(0, j) = SI(1, j) ? SI(1, j) : -1;
        if ((B[ncx * ngy + j - 1] == NEUMANN || B[ncx * ngy + j] == NEUMANN) && SI(ngx - 1, j) == 0)
            SI(ngx, j) = -B[ncx * ngy + j] - 4;
        else SI(ngx, j) = SI(ngx - 1, j) ? SI(ngx - 1, j) : -1;
    }
    SI(0, 0) = -B[0];
    SI(ncx + 1, 0) = -B[(ncx - 1) * ngy];
    SI(0, ncy + 1) = -B[ngy + ncy];
    // top-right: the reference reads node (ncy-1, ncy) (src/start.c:1848), a
    // top-wall node while ncy <= ncx and past the node array when ncy > ncx;
    // there the top-wall node (ncx-1, ncy) is used instead
    SI(ncx + 1, ncy + 1) = -B[(size_t)(ncy <= ncx ? ncy - 1 : ncx - 1) * ngy + ncy];
    for (int i = 0; i < s.Csize; ++i) s.eps[i] = (R)(kEps0 * (double)s.eps_r[i]);
}

// ------------------------------------------------ field operator (load.c:121-245)
template <typename R>
void field_matrix(Setup<R> &s) {
    const int ngx = s.ngx, ngy = s.ngy, ncx = s.ncx, ncy = s.ncy;
    s.A_idx.assign(s.Gsize, 0);
    int n = 1;
    for (int i = 0; i < ngx; ++i)
        for (int j = 0; j < ngy; ++j) {
            int g = i * ngy + j;
            if (!s.CondID[g] && s.Boundary[g] != DIRICHLET) s.A_idx[g] = n++;
        }
    const int N = n - 1;
    s.A_size = N;
    s.Ai.assign(N + 1, 0);
    s.Aj.clear(); s.Aval.clear(); s.Atemp.clear();
    s.Mdiag.assign(N, 0);
    s.cond_b.assign((size_t)s.CondNUMR * N, 0);
    s.temp_b.assign(N, 0);
    const R dydx = (R)(0.5 * (double)s.dy / (double)s.dx);
    const R dxdy = (R)(0.5 * (double)s.dx / (double)s.dy);
    const auto &er = s.eps_r;
    int l = 0;
    for (int i = 0; i < ngx; ++i)
        for (int j = 0; j < ngy; ++j) {
            int g = i * ngy + j;
            if (!s.A_idx[g]) continue;
            int x1 = (i == 0) ? i : i - 1, y1 = (j == ncy) ? j - 1 : j, y2 = (j == 0) ? j : j - 1;
            R L = dydx * (er[x1 * ncy + y1] + er[x1 * ncy + y2]);
            x1 = (i == ncx) ? i - 1 : i;
            R Rc = dydx * (er[x1 * ncy + y1] + er[x1 * ncy + y2]);
            int x2 = (i == 0) ? i : i - 1;
            y1 = (j == ncy) ? j - 1 : j;
            R U = dxdy * (er[x1 * ncy + y1] + er[x2 * ncy + y1]);
            y1 = (j == 0) ? j : j - 1;
            R D = dxdy * (er[x1 * ncy + y1] + er[x2 * ncy + y1]);
            const int LN = (i == 0) ? 2 : 1, RN = (i == ncx) ? 2 : 1, UN = (j == ncy) ? 2 : 1, DN = (j == 0) ? 2 : 1;
            auto neighbour = [&](int nb, R coef, R aval, R atemp) {
                if (s.CondID[nb]) {
                    s.cond_b[(size_t)(s.CondID[nb] - 1) * N + l] += coef;
                    s.temp_b[l] += coef * s.Temp[nb];
                } else if (s.Boundary[nb] == DIRICHLET) {
                    s.temp_b[l] += coef * s.Temp[nb];
                } else {
                    s.Aval.push_back(aval); s.Atemp.push_back(atemp); s.Aj.push_back(s.A_idx[nb] - 1);
                }
            };
            if (i) neighbour((i - 1) * ngy + j, L, (-L) * RN, (-L) * RN);
            if (j) neighbour(i * ngy + j - 1, D, (-D) * UN, (-D));
            s.Aval.push_back(L + D + U + Rc);
            s.Mdiag[l] = (R)1 / s.Aval.back();
            s.Atemp.push_back(L + D + U + Rc);
            s.Aj.push_back(s.A_idx[g] - 1);
            if (j != ncy) neighbour(i * ngy + j + 1, U, (-U) * DN, (-U));
            if (i != ncx) neighbour((i + 1) * ngy + j, Rc, (-Rc) * LN, (-Rc));
            ++l;
            s.Ai[l] = (int)s.Aval.size();
        }
}

// ------------------------------------------------ Laplace + gas fields
template <typename R>
int presolves(Setup<R> &s, std::string &err) {
    const int N = s.A_size;
    if (N <= 0) { err = "field operator has no unknowns"; return -20; }
    BandLU lu;
    lu.factor(N, s.Ai, s.Aj, s.Aval);
    std::vector<double> b(N), x(N);
    s.lap_phi.assign((size_t)s.CondNUMR * s.Gsize, 0);
    for (int k = 0; k < s.CondNUMR; ++k) {
        for (int r = 0; r < N; ++r) b[r] = (double)s.cond_b[(size_t)k * N + r];
        lu.solve(b.data(), x.data());
        R *phi = s.lap_phi.data() + (size_t)k * s.Gsize;
        for (int g = 0; g < s.Gsize; ++g) {
            if (s.A_idx[g]) phi[g] = (R)x[s.A_idx[g] - 1];
            else if (s.CondID[g] == k + 1) phi[g] = 1;
        }
    }
    BandLU lut;
    lut.factor(N, s.Ai, s.Aj, s.Atemp);
    for (int r = 0; r < N; ++r) b[r] = (double)s.temp_b[r];
    lut.solve(b.data(), x.data());
    for (int g = 0; g < s.Gsize; ++g)
        if (s.A_idx[g]) s.Temp[g] = (R)x[s.A_idx[g] - 1];
    // Calculate_1GasPara / Calculate_2GasPara (src/cuda_Field.cu:661-678)
    s.BackDen1.assign(s.Gsize, 0);
    s.BackVel1.assign(s.Gsize, 0);
    s.BackDen2.assign(s.nbg > 1 ? s.Gsize : 0, 0);
    s.BackVel2.assign(s.nbg > 1 ? s.Gsize : 0, 0);
    for (int g = 0; g < s.Gsize; ++g) {
        if (!s.DensRegion[g]) continue;
        double T = (double)s.Temp[g];
        s.BackDen1[g] = (R)(kNperTorr * (double)s.bg_pres[0] / (T * 8.6142e-5 + kDblMin));
        s.BackVel1[g] = (R)std::sqrt(T * 1.38e-23 / (double)s.bg_mass[0]);
        if (s.nbg > 1) {
            s.BackDen2[g] = (R)(kNperTorr * (double)s.bg_pres[1] / (T * 8.6142e-5 + kDblMin));
            s.BackVel2[g] = (R)std::sqrt(T * 1.38e-23 / (double)s.bg_mass[1]);
        }
    }
    // metastable density at nodes (src/Fluid.c:3-68, src/cuda_Fluid.cu:321-339)
    s.FluidN.assign((size_t)s.nfsp * s.Gsize, 0);
    for (int f = 0; f < s.nfsp; ++f)
        for (int i = 0; i < s.ncx; ++i)
            for (int j = 0; j < s.ncy; ++j) {
                R den = 0;
                if (s.PlasmaRegion[i * s.ncy + j] != 0 && s.deck.fl_loadtype == 0) den = s.fl_dens[f];
                s.FluidN[(size_t)f * s.Gsize + i * s.ngy + j] = den;
                if (i == s.ncx - 1 || j == s.ncy - 1) s.FluidN[(size_t)f * s.Gsize + (i + 1) * s.ngy + j + 1] = den;
            }
    return 0;
}

// ------------------------------------------------ cross sections (PhysicsData.c)
template <typename R>
R func_decay(const float *mx, const float *my, int n, R in_x) {
    if (n > 1) {
        if (in_x < mx[0]) return my[0];
        if (in_x > mx[n - 1]) {
            if (is_f32<R>()) {
                float inc = (float)(std::log10((double)(my[n - 2] / my[n - 1])) / std::log10((double)(mx[n - 1] / mx[n - 2])));
                return (R)((double)my[n - 1] * std::pow((double)mx[n - 1], (double)inc) / std::pow((double)in_x, (double)inc));
            }
            double inc = std::log10((double)my[n - 2] / (double)my[n - 1]) / std::log10((double)mx[n - 1] / (double)mx[n - 2]);
            return (R)((double)my[n - 1] * std::pow((double)mx[n - 1], inc) / std::pow((double)in_x, inc));
        }
        int lo = 0, hi = n - 1;
        while (hi - lo > 1) {
            int k = (hi + lo) >> 1;
            if (mx[k] > in_x) hi = k; else lo = k;
        }
        R h = (R)mx[hi] - (R)mx[lo];
        R a = ((R)mx[hi] - in_x) / h;
        R b = (in_x - (R)mx[lo]) / h;
        return a * (R)my[lo] + b * (R)my[hi];
    }
    return my[0];
}

template <typename R>
void argon_tables(Setup<R> &s) {
    s.TnRct = AR_NRCT;
    s.idLOGX = is_f32<R>() ? (R)((R)1 / (R)0.1f) : (R)10.0;
    s.xe.assign(N_LOGX, 0); s.xee.assign(N_LOGX, 0); s.cx.assign(N_LOGX * AR_NRCT, 0);
    for (int i = 0; i < N_LOGX; ++i) {
        if (is_f32<R>()) s.xe[i] = (R)(0.1f * (float)(i - 40));
        else s.xe[i] = (R)(0.1 * (double)(i - 40));
        s.xee[i] = (R)std::pow(10.0, (double)s.xe[i]);
    }
    const int set[7] = {1, 2, 3, 4, 5, 7, 9};
    const int sub[7] = {-1, 3, -1, -1, -1, -1, -1};
    const float th[7] = {0.0f, 11.55f, 11.548f, 15.76f, 4.22f, 0.0f, 0.0f};
    for (int r = 0; r < AR_NRCT; ++r) { s.Flag[r] = (R)s.deck.coll_flag[r]; if (s.Flag[r] < 0) s.Flag[r] = 0; }
    for (int r = 0; r < 7; ++r) {
        s.Th_e[r] = (R)th[r];
        for (int i = 0; i < N_LOGX; ++i) {
            R v = 0;
            if (!(s.xee[i] < s.Th_e[r])) {
                v = s.Flag[r] * func_decay<R>(pic_xsec_x[set[r]], pic_xsec_y[set[r]], pic_xsec_n[set[r]], s.xee[i]);
                if (sub[r] >= 0)
                    v = v - s.Flag[r] * func_decay<R>(pic_xsec_x[sub[r]], pic_xsec_y[sub[r]], pic_xsec_n[sub[r]], s.xee[i]);
            }
            s.cx[i * AR_NRCT + r] = v;
        }
    }
    s.RR[7] = s.Flag[7] * (R)2.0e-13; s.RR[8] = s.Flag[8] * (R)6.2e-16; s.RR[9] = s.Flag[9] * (R)3.0e-21;
}

// O2 and Ar/O2 tables (src/PhysicsData.c:1719-3275): the same grid as
// argon, 58 or 68 tabulated reactions from the fits and data sets of
// xsec_oxygen.h, then the rate constants of the remaining reactions.
template <typename R>
void oxygen_tables(Setup<R> &s) {
    using namespace oxy;
    const bool o2 = s.deck.gas == 1;
    const Row *rows = o2 ? kO2Rows : kArO2Rows;
    const int ntab = o2 ? 58 : 68;
    const double *rates = o2 ? kO2Rates : kArO2Rates;
    s.TnRct = kCollN[s.deck.gas];
    s.idLOGX = is_f32<R>() ? (R)((R)1 / (R)0.1f) : (R)10.0;
    s.xe.assign(N_LOGX, 0); s.xee.assign(N_LOGX, 0); s.cx.assign((size_t)N_LOGX * s.TnRct, 0);
    for (int i = 0; i < N_LOGX; ++i) {
        if (is_f32<R>()) s.xe[i] = (R)(0.1f * (float)(i - 40));
        else s.xe[i] = (R)(0.1 * (double)(i - 40));
        s.xee[i] = (R)std::pow(10.0, (double)s.xe[i]);
    }
    for (int r = 0; r < s.TnRct; ++r) { s.Flag[r] = (R)s.deck.coll_flag[r]; if (s.Flag[r] < 0) s.Flag[r] = 0; }
    const double shift_d[4] = {0.0, 0.977, 1.627, 1.96};
    for (int r = 0; r < ntab; ++r) {
        const Row &w = rows[r];
        const R sh = is_f32<R>() ? (R)kShift[w.shift] : (R)shift_d[w.shift];
        // threshold: the constant minus the float shift, in double, rounded to R
        s.Th_e[r] = w.shift ? (R)(w.th - (double)sh) : (R)w.th;
        for (int i = 0; i < N_LOGX; ++i) {
            const R E = s.xee[i];
            R v = 0;
            if (!(E < s.Th_e[r])) {
                switch (w.kind) {
                case K_FIT: v = s.Flag[r] * fit<R>(w.a, (R)(E + sh)); break;
                case K_FIT_X35: v = (R)((double)s.Flag[r] * 3.5 * (double)fit<R>(w.a, (R)(E + sh))); break;
                case K_VIB: v = s.Flag[r] * (fit<R>(F_VIB1, E) + fit<R>(F_VIB2, E) + fit<R>(F_VIB3, E) + fit<R>(F_VIB4, E)); break;
                case K_DATA: v = s.Flag[r] * func_decay<R>(pic_xsec_x[w.a], pic_xsec_y[w.a], pic_xsec_n[w.a], E); break;
                default:
                    v = s.Flag[r] * func_decay<R>(pic_xsec_x[w.a], pic_xsec_y[w.a], pic_xsec_n[w.a], E);
                    v = v - s.Flag[r] * func_decay<R>(pic_xsec_x[w.b], pic_xsec_y[w.b], pic_xsec_n[w.b], E);
                }
            }
            s.cx[(size_t)i * s.TnRct + r] = v;
        }
    }
    for (int r = ntab; r < s.TnRct; ++r) s.RR[r] = (R)((double)s.Flag[r] * rates[r - ntab]);
}

// mass ratio target/projectile of every reaction (src/start.c:1003-1420)
template <typename R>
void mass_ratios(Setup<R> &s) {
    const signed char(*mm)[4] = kCollMofM[s.deck.gas];
    auto mass = [&](int kind, int i) { return kind == 0 ? s.mass[i] : kind == 1 ? s.bg_mass[i] : s.fl_mass[i]; };
    for (int r = 0; r < kCollN[s.deck.gas]; ++r) s.mofM[r] = mass(mm[r][0], mm[r][1]) / mass(mm[r][2], mm[r][3]);
}

// Null-collision constants for O2 and Ar/O2 (src/cuda_mcc.cu:291-501): the
// largest sigma*v over the grid per (projectile, target) pair, the per-step
// collision probability of each projectile species from the initial target
// densities, and the electron MCC sub-cycle count from the electrons'.
template <typename R>
void oxygen_null_collision(Setup<R> &s) {
    using namespace oxy;
    const bool o2 = s.deck.gas == 1;
    const Pair *pairs = o2 ? kO2Pairs : kArO2Pairs;
    const int np = o2 ? 20 : 25;
    s.nsigv = np;
    for (int k = 0; k < np; ++k) s.sigv[k] = 0;
    for (int i = 0; i < N_LOGX; ++i) {
        const R *row = s.cx.data() + (size_t)i * s.TnRct;
        const double e = (double)s.xee[i];
        for (int k = 0; k < np; ++k) {
            R sig = 0, grp = 0;
            bool first = true, open = false;
            for (const int *c = pairs[k].cols; *c != -2; ++c) {
                if (*c == -1) { sig = first ? grp : sig + grp; first = false; open = false; continue; }
                grp = open ? grp + row[*c] : row[*c];
                open = true;
            }
            if (open) sig = first ? grp : sig + grp;
            const R v = (R)(std::sqrt(2 * 1.602e-19 * e / (double)s.mass[pairs[k].proj]) * (double)sig);
            if (v > s.sigv[k]) s.sigv[k] = v;
        }
    }
    for (int p = 0; p < MAXSP; ++p) s.colprob[p] = 0;
    for (int k = 0; k < np; ++k) {
        const Pair &w = pairs[k];
        const double dens = w.tkind == T_BG ? (double)s.bg_dens[w.tidx]
                          : w.tkind == T_FG ? (double)s.fl_dens[w.tidx] : (double)(R)s.deck.sp_dens[w.tidx];
        const R t = (R)(1.0 - std::exp(-1 * (double)s.dt * (double)s.sigv[k] * dens));
        s.colprob[w.proj] = s.colprob[w.proj] + t;
    }
    const R colprob = s.colprob[0];
    if (colprob > (R)0.08) {
        s.MCCn = (int)std::ceil((double)(colprob / (R)0.08));
        R ratio = (R)1 / (R)s.MCCn;
        s.dt_mcc = s.dt * ratio;
    } else {
        s.MCCn = 1;
        s.dt_mcc = s.dt;
    }
}

template <typename R>
void null_collision(Setup<R> &s) {
    R sv[3] = {0, 0, 0};
    for (int i = 0; i < N_LOGX; ++i) {
        const R *row = s.cx.data() + i * AR_NRCT;
        double e = (double)s.xee[i];
        R s0 = row[0] + row[1] + row[2] + row[3];
        R s1 = row[4];
        R s2 = row[5] + row[6];
        R v0 = (R)(std::sqrt(2 * 1.602e-19 * e / (double)s.mass[0]) * (double)s0);
        R v1 = (R)(std::sqrt(2 * 1.602e-19 * e / (double)s.mass[0]) * (double)s1);
        R v2 = (R)(std::sqrt(2 * 1.602e-19 * e / (double)s.mass[1]) * (double)s2);
        if (v0 > sv[0]) sv[0] = v0;
        if (v1 > sv[1]) sv[1] = v1;
        if (v2 > sv[2]) sv[2] = v2;
    }
    s.sigv[0] = sv[0]; s.sigv[1] = sv[1]; s.sigv[2] = sv[2];
    s.nsigv = 3;
    double t0 = 1.0 - std::exp(-1 * (double)s.dt * (double)sv[0] * (double)s.bg_dens[0]);
    double t1 = 1.0 - std::exp(-1 * (double)s.dt * (double)sv[1] * (double)s.fl_dens[0]);
    R colprob = (R)t0 + (R)t1;
    s.colprob[0] = colprob;
    s.colprob[1] = (R)(1.0 - std::exp(-1 * (double)s.dt * (double)sv[2] * (double)s.bg_dens[0]));
    if (colprob > (R)0.08) {
        s.MCCn = (int)std::ceil((double)(colprob / (R)0.08));
        R ratio = (R)1 / (R)s.MCCn;
        s.dt_mcc = s.dt * ratio;
    } else {
        s.MCCn = 1;
        s.dt_mcc = s.dt;
    }
    s.P1.assign(s.Gsize, 0); s.P2.assign(s.Gsize, 0); s.Pion.assign(s.Gsize, 0);
    for (int g = 0; g < s.Gsize; ++g) {
        double dtm = (double)s.dt_mcc, dtd = (double)s.dt;
        R p1 = (R)(1.0 - dm_exp(-dtm * (double)s.sigv[0] * (double)s.BackDen1[g]));
        R p2b = (R)(1.0 - dm_exp(-dtm * (double)s.sigv[1] * (double)s.FluidN[g]));
        s.P1[g] = p1;
        s.P2[g] = p1 + p2b;
        s.Pion[g] = (R)(1.0 - dm_exp(-dtd * (double)s.sigv[2] * (double)s.BackDen1[g]));
    }
}

// ------------------------------------------------ Maxwellian table + loader
struct Lcg {
    long seed;
    float next() {   // Park-Miller minimal standard, Schrage's method
        long a = 16807, m = 2147483647, q = 127773, r = 2836;
        long hi = seed / q, lo = seed - q * hi;
        seed = a * lo - r * hi;
        if (seed <= 0) seed += m;
        return (float)(seed / 2147483646.0);
    }
};

// F(v) = erf(v) - 2 v exp(-v^2)/sqrt(pi), with the float products of
// src/Particle.c:227-230
inline float maxwell_cdf(float v) {
    float m2v = -2 * v, mvv = -v * v;
    return (float)((double)m2v * std::exp((double)mvv) / std::sqrt(kPi) + std::erf((double)v));
}

template <typename R>
void maxwell_table(Setup<R> &s) {
    int nvel = (int)(1. / (double)(1 - maxwell_cdf(3.3f)));
    float dv = (float)(std::sqrt(kPi) / (4.0 * nvel));
    s.nvel = nvel;
    s.vsave.assign(nvel, 0);
    int i = 0;
    for (int n = 0; n < nvel; ++n) {
        float rr = (float)((1.0 * n) / nvel);
        while (maxwell_cdf((float)i * dv) < rr) ++i;
        s.vsave[n] = (R)(float)(std::sqrt(2.0) * i * (double)dv);
    }
}

template <typename R>
int load(Setup<R> &s, std::string &err) {
    const Deck &d = s.deck;
    if (d.sp_loadtype != 0) { err = "only the UNIFORM particle load is available"; return -21; }
    Lcg g{(long)d.seed};
    for (int isp = 0; isp < s.nsp; ++isp) {
        SpeciesInit<R> &sp = s.sp[isp];
        float dxf = (float)s.dx, dyf = (float)s.dy, zl = (float)s.zl;
        float ninit = (float)d.sp_dens[isp];
        int npc = (int)(ninit * dxf * dyf * zl / (float)d.sp_np2c[isp]);
        float y0 = (float)d.sp_y0 / dyf, y1 = (float)d.sp_y1 / dyf - 1;
        float x0 = (float)d.sp_x0 / dxf, x1 = (float)d.sp_x1 / dxf - 1;
        long total = 0;
        for (int i = 0; i < s.ncx; ++i)
            for (int j = 0; j < s.ncy; ++j)
                if (s.PlasmaRegion[i * s.ncy + j] == 1 && i >= x0 && i <= x1 && j >= y0 && j <= y1) total += npc;
        if (total > d.np_limit) { err = "LOAD: too many particles"; return -22; }
        sp.cap = d.sp_maxnp[isp] > total ? d.sp_maxnp[isp] : (int)total;
        std::vector<int> cellof(total), cnt(s.Gsize, 0);
        std::vector<float> tx(total), ty(total), tvx(total), tvy(total), tvz(total);
        const float vti = (float)s.vti[isp];
        long idx = 0;
        for (int i = 0; i < s.ncx; ++i)
            for (int j = 0; j < s.ncy; ++j) {
                if (s.PlasmaRegion[i * s.ncy + j] != 1) continue;
                int num = (i >= x0 && i <= x1 && j >= y0 && j <= y1) ? npc : 0;
                for (int n = 0; n < num; ++n) {
                    int vi = (int)((s.nvel - 1) * g.next());
                    float aphi = (float)(2 * kPi * g.next());
                    float costheta = 1 - 2 * g.next();
                    float sintheta = (float)std::sqrt((double)(1 - costheta * costheta));
                    float vs = (float)s.vsave[vi];
                    tvx[idx] = (float)((double)(vti * vs * sintheta) * std::cos((double)aphi));
                    tvy[idx] = (float)((double)(vti * vs * sintheta) * std::sin((double)aphi));
                    tvz[idx] = vti * vs * costheta;
                    float px = (float)i + g.next();
                    float py = (float)j + g.next();
                    int xid = (int)px, yid = (int)py;
                    if (xid > s.ncx - 1) xid = s.ncx - 1;
                    if (yid > s.ncy - 1) yid = s.ncy - 1;
                    tx[idx] = px - (float)xid;
                    ty[idx] = py - (float)yid;
                    cellof[idx] = xid * s.ngy + yid;
                    cnt[cellof[idx]]++;
                    ++idx;
                }
            }
        sp.cell_start.assign(s.Gsize + 1, 0);
        for (int c = 0; c < s.Gsize; ++c) sp.cell_start[c + 1] = sp.cell_start[c] + cnt[c];
        sp.x.assign(idx, 0); sp.y.assign(idx, 0); sp.vx.assign(idx, 0); sp.vy.assign(idx, 0); sp.vz.assign(idx, 0);
        sp.id.assign(idx, 0);
        for (int c = 0; c < s.Gsize; ++c) cnt[c] = sp.cell_start[c];
        for (long n = 0; n < idx; ++n) {
            int dst = cnt[cellof[n]]++;
            sp.x[dst] = (R)tx[n]; sp.y[dst] = (R)ty[n];
            sp.vx[dst] = (R)tvx[n]; sp.vy[dst] = (R)tvy[n]; sp.vz[dst] = (R)tvz[n];
            sp.id[dst] = (uint32_t)n;
        }
        sp.n = (int)idx;
    }
    return 0;
}

}  // namespace

// ------------------------------------------------ metastable fluid set-up
// Collision integral of Ar* diffusion against T* = kT/epsilon, by the
// reference's table and loop (src/cuda_Fluid.cu:261-283).  That loop has no
// break, so inside the table every T* takes the last interval's line,
// omega[25..26]; the same is done here.
static float ar_meta_omega(float value) {
    static const float tstar[27] = {0.250f, 0.267f, 0.286f, 0.308f, 0.333f, 0.364f, 0.400f, 0.444f, 0.500f,
                                    0.571f, 0.667f, 0.800f, 1.000f, 1.330f, 2.000f, 2.500f, 2.860f, 3.330f,
                                    4.000f, 5.000f, 6.670f, 10.00f, 12.50f, 16.67f, 25.00f, 50.00f, 100.0f};
    static const float omega[27] = {1.8714f, 1.8344f, 1.7962f, 1.7560f, 1.7142f, 1.6702f, 1.6238f, 1.5748f, 1.5228f,
                                    1.4676f, 1.4090f, 1.3466f, 1.2802f, 1.2096f, 1.1348f, 1.1042f, 1.0888f, 1.0738f,
                                    1.0592f, 1.0450f, 1.0318f, 1.0198f, 1.0154f, 1.0112f, 1.0072f, 1.0034f, 1.000f};
    float r = (float)DBL_MIN;
    if (value < tstar[0]) return omega[0];
    if (value > tstar[26]) return omega[26];
    for (int i = 1; i < 27; ++i)
        if (value < tstar[i]) {
            const float t = (value - tstar[i - 1]) / (tstar[i] - tstar[i - 1]);
            r = omega[i] * t + omega[i - 1] * (1 - t);
        }
    return r;
}

template <typename R>
void fluid_setup(Setup<R> &s) {
    const int ncx = s.ncx, ncy = s.ncy, ngx = s.ngx, ngy = s.ngy, nf = s.nfsp;
    s.dtc = s.dt * (R)s.deck.dt_conti;
    auto plasma = [&](int i, int j) { return s.PlasmaRegion[i * ncy + j] == 1; };
    // lines (src/Fluid.c:188-256): along x, y outer; along y, x outer
    s.fx_x1.clear(); s.fx_x2.clear(); s.fx_y.clear(); s.fy_x.clear(); s.fy_y1.clear(); s.fy_y2.clear();
    for (int j = 0; j < ncy; ++j)
        for (int i = 0; i < ncx; ++i) {
            if (!plasma(i, j)) continue;
            if (i == 0 || !plasma(i - 1, j)) { s.fx_x1.push_back(i); s.fx_y.push_back(j); }
            if (i == ncx - 1 || !plasma(i + 1, j)) s.fx_x2.push_back(i);
        }
    for (int i = 0; i < ncx; ++i)
        for (int j = 0; j < ncy; ++j) {
            if (!plasma(i, j)) continue;
            if (j == 0 || !plasma(i, j - 1)) { s.fy_x.push_back(i); s.fy_y1.push_back(j); }
            if (j == ncy - 1 || !plasma(i, j + 1)) s.fy_y2.push_back(j);
        }
    s.fx_n = (int)s.fx_x2.size();
    s.fy_n = (int)s.fy_y2.size();
    // wall flux factors (src/Fluid.c:113-186): 0 at a Neumann node on the
    // domain edge, the species' quenching probability everywhere else (a
    // float in the reference, FG.Gamma1, in FP64 cases too)
    s.fx_g1.assign((size_t)nf * s.fx_n, 0); s.fx_g2.assign((size_t)nf * s.fx_n, 0);
    s.fy_g1.assign((size_t)nf * s.fy_n, 0); s.fy_g2.assign((size_t)nf * s.fy_n, 0);
    auto neu = [&](int g) { return s.Boundary[g] == NEUMANN; };
    for (int f = 0; f < nf; ++f) {
        const R gam = (R)(float)s.deck.fl_quench[f];
        for (int k = 0; k < s.fx_n; ++k) {
            const int x1 = s.fx_x1[k], x2 = s.fx_x2[k], yy = s.fx_y[k];
            const bool n1 = neu(x1 * ngy + yy) || neu(x1 * ngy + yy + 1);
            const bool n2 = neu((x2 + 1) * ngy + yy) || neu((x2 + 1) * ngy + yy + 1);
            s.fx_g1[(size_t)f * s.fx_n + k] = (n1 && x1 == 0) ? (R)0 : gam;
            s.fx_g2[(size_t)f * s.fx_n + k] = (n2 && x2 == ncx - 1) ? (R)0 : gam;
        }
        for (int k = 0; k < s.fy_n; ++k) {
            const int xx = s.fy_x[k], y1 = s.fy_y1[k], y2 = s.fy_y2[k];
            const bool n1 = neu(xx * ngy + y1) || neu((xx + 1) * ngy + y1);
            const bool n2 = neu(xx * ngy + y2 + 1) || neu((xx + 1) * ngy + y2 + 1);
            s.fy_g1[(size_t)f * s.fy_n + k] = (n1 && y1 == 0) ? (R)0 : gam;
            s.fy_g2[(size_t)f * s.fy_n + k] = (n2 && y2 == ncy - 1) ? (R)0 : gam;
        }
    }
    // diffusion coefficients per cell (src/cuda_Fluid.cu:199-260), with the
    // reference's mixed float/double arithmetic
    const size_t C = (size_t)ncx * ncy;
    s.fl_D.assign((size_t)nf * C, 0);
    const R tp = (R)s.deck.total_pres;
    for (int f = 0; f < nf; ++f)
        for (int i = 0; i < ncx; ++i)
            for (int j = 0; j < ncy; ++j) {
                const int g = i * ngy + j;
                const R T = (R)(0.25 * (double)(s.Temp[g] + s.Temp[g + 1] + s.Temp[g + ngy] + s.Temp[g + ngy + 1]));
                double D;
                const double T3 = (double)(T * T * T * (R)24);
                if (s.deck.gas == 0) {
                    const R om = (R)ar_meta_omega((float)(T / (R)124.0f));
                    const R den = tp * (R)11.682724f * om;
                    D = 1.997279e-4 * std::sqrt((double)T / 39.948) * (double)T / (double)den * 20;
                } else {
                    // species: O2 mixtures O2A, O2B (O2), OP, OD (O); Ar/O2 Ar* first
                    const int k = s.deck.gas == 2 ? f - 1 : f;   // -1: Ar*
                    if (k < 0)
                        D = 2.63e-3 * std::sqrt(T3 / 39.948 / 39.948) / (double)tp * (double)(R)0.08559f;
                    else if (k <= 1)
                        D = 2.63e-3 * std::sqrt((double)(T * T * T * (R)24 / (R)32 / (R)32)) / (double)tp * (double)(R)0.08650f;
                    else
                        D = 2.63e-3 * std::sqrt((double)(T * T * T * (R)24 / (R)16 / (R)16)) / (double)tp * (double)(R)0.10892f;
                }
                s.fl_D[(size_t)f * C + (size_t)i * ncy + j] = (R)D;
            }
    // Scharfetter-Gummel coefficients (src/cuda_Fluid.cu:179-198)
    const size_t FX = (size_t)ngx * ncy, FY = (size_t)ncx * ngy;
    s.fl_ax.assign((size_t)nf * FX, 0); s.fl_bx.assign((size_t)nf * FX, 0);
    s.fl_ay.assign((size_t)nf * FY, 0); s.fl_by.assign((size_t)nf * FY, 0);
    for (int f = 0; f < nf; ++f) {
        const R *D = s.fl_D.data() + (size_t)f * C;
        R *ax = s.fl_ax.data() + (size_t)f * FX, *bx = s.fl_bx.data() + (size_t)f * FX;
        R *ay = s.fl_ay.data() + (size_t)f * FY, *by = s.fl_by.data() + (size_t)f * FY;
        const double m = (double)s.fl_mass[f];
        auto vth = [&](int g) { return 0.25 * std::sqrt((double)s.Temp[g] * 1.38e-23 / m); };
        for (int k = 0; k < s.fx_n; ++k) {
            const int x1 = s.fx_x1[k], x2 = s.fx_x2[k], yy = s.fx_y[k];
            const double g1 = (double)s.fx_g1[(size_t)f * s.fx_n + k], g2 = (double)s.fx_g2[(size_t)f * s.fx_n + k];
            bx[x1 * ncy + yy] = (R)(-g1 * (-vth(x1 * ngy + yy)));
            ax[x1 * ncy + yy] = 0;
            for (int i = x1 + 1; i <= x2; ++i) {
                // the reference averages D[i-1] with itself here
                ax[i * ncy + yy] = (D[(i - 1) * ncy + yy] + D[(i - 1) * ncy + yy]) / (R)2 / s.dx;
                bx[i * ncy + yy] = ax[i * ncy + yy];
            }
            ax[(x2 + 1) * ncy + yy] = (R)(g2 * vth(x2 * ngy + yy));
            bx[(x2 + 1) * ncy + yy] = 0;
        }
        for (int k = 0; k < s.fy_n; ++k) {
            const int xx = s.fy_x[k], y1 = s.fy_y1[k], y2 = s.fy_y2[k];
            const double g1 = (double)s.fy_g1[(size_t)f * s.fy_n + k], g2 = (double)s.fy_g2[(size_t)f * s.fy_n + k];
            by[xx * ngy + y1] = (R)(-g1 * (-vth(xx * ngy + y1)));
            ay[xx * ngy + y1] = 0;
            for (int j = y1 + 1; j <= y2; ++j) {
                ay[xx * ngy + j] = (D[xx * ncy + j - 1] + D[xx * ncy + j]) / (R)2 / s.dy;
                by[xx * ngy + j] = ay[xx * ngy + j];
            }
            ay[xx * ngy + y2 + 1] = (R)(g2 * vth(xx * ngy + y2));
            by[xx * ngy + y2 + 1] = 0;
        }
    }
    // initial density (uniform load over the plasma, src/Fluid.c:35-67) and
    // face fluxes (src/cuda_Fluid.cu:51-90)
    s.fl_den.assign((size_t)nf * C, 0);
    for (int f = 0; f < nf; ++f)
        for (int i = 0; i < ncx; ++i)
            for (int j = 0; j < ncy; ++j)
                if (s.PlasmaRegion[i * ncy + j] != 0 && s.deck.fl_loadtype == 0)
                    s.fl_den[(size_t)f * C + (size_t)i * ncy + j] = s.fl_dens[f];
    s.fl_flux_x.assign((size_t)nf * FX, 0);
    s.fl_flux_y.assign((size_t)nf * FY, 0);
    for (int f = 0; f < nf; ++f) {
        const R *n = s.fl_den.data() + (size_t)f * C;
        const R *ax = s.fl_ax.data() + (size_t)f * FX, *bx = s.fl_bx.data() + (size_t)f * FX;
        const R *ay = s.fl_ay.data() + (size_t)f * FY, *by = s.fl_by.data() + (size_t)f * FY;
        R *fx = s.fl_flux_x.data() + (size_t)f * FX, *fy = s.fl_flux_y.data() + (size_t)f * FY;
        for (int k = 0; k < s.fx_n; ++k) {
            const int x1 = s.fx_x1[k], x2 = s.fx_x2[k], yy = s.fx_y[k];
            const R g1 = s.fx_g1[(size_t)f * s.fx_n + k], g2 = s.fx_g2[(size_t)f * s.fx_n + k];
            fx[x1 * ncy + yy] = -g1 * bx[x1 * ncy + yy] * n[x1 * ncy + yy];
            for (int i = x1 + 1; i <= x2; ++i)
                fx[i * ncy + yy] = ax[i * ncy + yy] * n[(i - 1) * ncy + yy] - bx[i * ncy + yy] * n[i * ncy + yy];
            fx[(x2 + 1) * ncy + yy] = g2 * ax[(x2 + 1) * ncy + yy] * n[x2 * ncy + yy];
        }
        for (int k = 0; k < s.fy_n; ++k) {
            const int xx = s.fy_x[k], y1 = s.fy_y1[k], y2 = s.fy_y2[k];
            const R g1 = s.fy_g1[(size_t)f * s.fy_n + k], g2 = s.fy_g2[(size_t)f * s.fy_n + k];
            fy[xx * ngy + y1] = -g1 * by[xx * ngy + y1] * n[xx * ncy + y1];
            for (int j = y1 + 1; j <= y2; ++j)
                fy[xx * ngy + j] = ay[xx * ngy + j] * n[xx * ncy + j - 1] - by[xx * ngy + j] * n[xx * ncy + j];
            fy[xx * ngy + y2 + 1] = g2 * ay[xx * ngy + y2 + 1] * n[xx * ncy + y2];
        }
    }
}

template <typename R>
int build_setup(const Deck &d, Setup<R> &s, std::string &err, bool do_load) {
    s.deck = d;
    int e = constants(d, s, err);
    if (e) return e;
    geometry(d, s);
    field_matrix(s);
    e = presolves(s, err);
    if (e) return e;
    if (d.gas == 0) {
        argon_tables(s);
        null_collision(s);
    } else {
        oxygen_tables(s);
        oxygen_null_collision(s);
    }
    mass_ratios(s);
    maxwell_table(s);
    fluid_setup(s);
    if (do_load) {
        e = load(s, err);
        if (e) return e;
    }
    return 0;
}

// ---------------------------------------------------------------- spectral
namespace {
enum EndType { END_N, END_D };

// eigenvectors (columns) and eigenvalues of the n x n second-difference
// matrix with the given ends: interior rows [-1 2 -1], a Neumann end row
// [2 -2] (mirror), a Dirichlet end row [2 -1]
void tridiag_eigen(int n, EndType left, EndType right, std::vector<double> &V, std::vector<double> &lam) {
    const double PI = 3.14159265358979323846;
    V.assign((size_t)n * n, 0.0);
    lam.assign(n, 0.0);
    for (int k = 0; k < n; ++k) {
        double th;
        for (int a = 0; a < n; ++a) {
            double v;
            if (left == END_N && right == END_N) { th = n > 1 ? PI * k / (n - 1) : 0.0; v = std::cos(th * a); }
            else if (left == END_D && right == END_D) { th = PI * (k + 1) / (n + 1); v = std::sin(th * (a + 1)); }
            else if (left == END_N) { th = PI * (2 * k + 1) / (2.0 * n); v = std::cos(th * a); }
            else { th = PI * (2 * k + 1) / (2.0 * n); v = std::cos(th * (n - 1 - a)); }
            V[(size_t)a * n + k] = v;
        }
        if (left == END_N && right == END_N) th = n > 1 ? PI * k / (n - 1) : 0.0;
        else if (left == END_D && right == END_D) th = PI * (k + 1) / (n + 1);
        else th = PI * (2 * k + 1) / (2.0 * n);
        lam[k] = 2.0 - 2.0 * std::cos(th);
    }
}

void tridiag_matrix(int n, EndType left, EndType right, std::vector<double> &T) {
    T.assign((size_t)n * n, 0.0);
    for (int a = 0; a < n; ++a) {
        T[(size_t)a * n + a] = 2.0;
        if (a > 0) T[(size_t)a * n + a - 1] = -1.0;
        if (a < n - 1) T[(size_t)a * n + a + 1] = -1.0;
    }
    if (n > 1 && left == END_N) T[1] = -2.0;
    if (n > 1 && right == END_N) T[(size_t)(n - 1) * n + n - 2] = -2.0;
}

bool invert(int n, const std::vector<double> &A, std::vector<double> &Ai) {
    std::vector<double> M(A);
    Ai.assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) Ai[(size_t)i * n + i] = 1.0;
    for (int c = 0; c < n; ++c) {
        int piv = c;
        for (int r = c + 1; r < n; ++r)
            if (std::fabs(M[(size_t)r * n + c]) > std::fabs(M[(size_t)piv * n + c])) piv = r;
        if (M[(size_t)piv * n + c] == 0.0) return false;
        if (piv != c)
            for (int k = 0; k < n; ++k) {
                std::swap(M[(size_t)c * n + k], M[(size_t)piv * n + k]);
                std::swap(Ai[(size_t)c * n + k], Ai[(size_t)piv * n + k]);
            }
        double d = 1.0 / M[(size_t)c * n + c];
        for (int k = 0; k < n; ++k) { M[(size_t)c * n + k] *= d; Ai[(size_t)c * n + k] *= d; }
        for (int r = 0; r < n; ++r) {
            if (r == c) continue;
            double f = M[(size_t)r * n + c];
            if (f == 0.0) continue;
            for (int k = 0; k < n; ++k) {
                M[(size_t)r * n + k] -= f * M[(size_t)c * n + k];
                Ai[(size_t)r * n + k] -= f * Ai[(size_t)c * n + k];
            }
        }
    }
    return true;
}
}  // namespace

template <typename R>
bool build_spectral(const Setup<R> &S, Spectral &sp, std::string &why) {
    sp = Spectral();
    const int ngy = S.ngy;
    int i0 = S.ngx, i1 = -1, j0 = S.ngy, j1 = -1;
    for (int i = 0; i < S.ngx; ++i)
        for (int j = 0; j < S.ngy; ++j)
            if (S.A_idx[i * ngy + j]) { i0 = std::min(i0, i); i1 = std::max(i1, i); j0 = std::min(j0, j); j1 = std::max(j1, j); }
    if (i1 < 0) { why = "no unknowns"; return false; }
    const int nx = i1 - i0 + 1, ny = j1 - j0 + 1;
    if ((long)nx * ny != S.A_size) { why = "unknowns do not form a rectangle"; return false; }
    for (int i = i0; i <= i1; ++i)
        for (int j = j0; j <= j1; ++j)
            if (S.A_idx[i * ngy + j] != (i - i0) * ny + (j - j0) + 1) { why = "unknown numbering is not row-major"; return false; }
    for (size_t c = 0; c < S.eps_r.size(); ++c)
        if (S.eps_r[c] != S.eps_r[0]) { why = "non-uniform permittivity"; return false; }
    const EndType xl = i0 == 0 ? END_N : END_D, xr = i1 == S.ncx ? END_N : END_D;
    const EndType yl = j0 == 0 ? END_N : END_D, yr = j1 == S.ncy ? END_N : END_D;
    if (xl == END_N && xr == END_N && yl == END_N && yr == END_N) { why = "pure Neumann operator"; return false; }
    const double cx = (double)(R)((R)(0.5 * (double)S.dy / (double)S.dx) * (S.eps_r[0] + S.eps_r[0]));
    const double cy = (double)(R)((R)(0.5 * (double)S.dx / (double)S.dy) * (S.eps_r[0] + S.eps_r[0]));
    std::vector<double> Tx, Ty;
    tridiag_matrix(nx, xl, xr, Tx);
    tridiag_matrix(ny, yl, yr, Ty);
    // every matrix entry must match cx Tx (x) I + I (x) cy Ty
    const double tol = std::is_same<R, float>::value ? 1e-6 : 1e-13;
    bool exact = true;
    for (int a = 0; a < nx; ++a)
        for (int b = 0; b < ny; ++b) {
            const int row = a * ny + b;
            int seen = 0;
            for (int q = S.Ai[row]; q < S.Ai[row + 1]; ++q) {
                const int col = S.Aj[q], a2 = col / ny, b2 = col % ny;
                double model;
                if (col == row) model = cx * Tx[(size_t)a * nx + a] + cy * Ty[(size_t)b * ny + b];
                else if (b2 == b) model = cx * Tx[(size_t)a * nx + a2];
                else if (a2 == a) model = cy * Ty[(size_t)b * ny + b2];
                else { why = "coupling outside the 5-point stencil"; return false; }
                const double v = (double)S.Aval[q];
                if (std::fabs(v - model) > tol * (std::fabs(model) + cx + cy)) { why = "coefficients differ from the separable model"; return false; }
                if (v != model) exact = false;
                ++seen;
            }
            int expect = 1 + (a > 0) + (a < nx - 1) + (b > 0) + (b < ny - 1);
            if (seen != expect) { why = "stencil entry count differs from the separable model"; return false; }
        }
    sp.nx = nx; sp.ny = ny; sp.cx = cx; sp.cy = cy; sp.exact = exact;
    tridiag_eigen(nx, xl, xr, sp.Vx, sp.lx);
    tridiag_eigen(ny, yl, yr, sp.Vy, sp.ly);
    if (!invert(nx, sp.Vx, sp.Vxi) || !invert(ny, sp.Vy, sp.Vyi)) { why = "singular eigenbasis"; return false; }
    // check T V = V diag(lambda)
    auto check = [](int n, const std::vector<double> &T, const std::vector<double> &V, const std::vector<double> &l) {
        double err = 0, scale = 0;
        for (int a = 0; a < n; ++a)
            for (int k = 0; k < n; ++k) {
                double s = 0;
                for (int m = 0; m < n; ++m) s += T[(size_t)a * n + m] * V[(size_t)m * n + k];
                err = std::max(err, std::fabs(s - l[k] * V[(size_t)a * n + k]));
                scale = std::max(scale, std::fabs(V[(size_t)a * n + k]));
            }
        return err <= 1e-9 * 4.0 * scale;
    };
    if (!check(nx, Tx, sp.Vx, sp.lx) || !check(ny, Ty, sp.Vy, sp.ly)) { why = "eigenbasis check failed"; return false; }
    sp.ok = true;
    return true;
}

template bool build_spectral<float>(const Setup<float> &, Spectral &, std::string &);
template bool build_spectral<double>(const Setup<double> &, Spectral &, std::string &);

template int build_setup<float>(const Deck &, Setup<float> &, std::string &, bool);
template int build_setup<double>(const Deck &, Setup<double> &, std::string &, bool);
template int BandLU::factor<float>(int, const std::vector<int> &, const std::vector<int> &, const std::vector<float> &);
template int BandLU::factor<double>(int, const std::vector<int> &, const std::vector<int> &, const std::vector<double> &);

}  // namespace pic
