// Host orchestration of one discharge case on one GPU, and the C ABI of
// include/picmcc.h.  A case owns every device buffer (no globals, unlike the
// reference's src/Variables.cu) and runs on its own HIP stream, so several
// cases can share a device.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/picmcc.h"
#include "mcc_o2.h"
#include "multi.h"
#include "pic_kernels.h"
#include "pic_setup.h"

namespace pic {

static thread_local std::string g_err;
static int fail(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIPCHK(expr)                                                                              \
    do {                                                                                          \
        hipError_t _e = (expr);                                                                   \
        if (_e != hipSuccess) return fail(-50, "%s: %s (%s:%d)", #expr, hipGetErrorString(_e), __FILE__, __LINE__); \
    } while (0)

template <typename T>
static hipError_t dalloc(T **p, size_t n) {
    return hipMalloc((void **)p, (n ? n : 1) * sizeof(T));
}
// allocate and copy on the given stream, then wait: the host vector may be
// freed as soon as this returns
template <typename T>
static hipError_t upload(T **p, const std::vector<T> &v, hipStream_t s) {
    hipError_t e = dalloc(p, v.size());
    if (e != hipSuccess) return e;
    if (!v.empty()) {
        e = hipMemcpyAsync(*p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
    }
    return e;
}

struct CaseBase {
    virtual ~CaseBase() {}
    virtual int advance_time() = 0;
    virtual int field_solve() = 0;
    virtual int efield() = 0;
    virtual int move() = 0;
    virtual int sort_boundary() = 0;
    virtual int mcc() = 0;
    virtual int deposit() = 0;
    virtual int diagnostic() = 0;
    virtual int eedf_sample(double emax) = 0;
    virtual int diag_sample() = 0;
    virtual int diag_history(double *out, int max_records, int *nrec, int *rec_len) = 0;
    virtual int dump_save(const char *path) = 0;
    virtual int fluid_enable(int on) = 0;
    virtual int dump_load(const char *path) = 0;
    virtual int step_fused(int n) = 0;
    virtual int sync() = 0;
    virtual int query() = 0;
    virtual int info(pic_info *) = 0;
    virtual int get_particles(int, void *, void *, void *, void *, void *, uint32_t *, int32_t *) = 0;
    virtual int set_particles(int, long long, const void *, const void *, const void *, const void *, const void *,
                              const uint32_t *, const int32_t *, uint32_t) = 0;
    virtual int array_info(const char *, size_t *, int *, int *) = 0;
    virtual int get_array(const char *, void *, size_t) = 0;
    virtual int set_array(const char *, const void *, size_t) = 0;
    virtual int set_step(long long) = 0;
    virtual int last_solve(int *, int *, double *) = 0;
    virtual int solver_shape(int *, int *) = 0;
    virtual int set_timing(int) = 0;
    virtual int get_timing(double *, int, long long *) = 0;
    int rf_cycles = 0, dt_pic = 0;
    int batch_refs = 0;    // batches this case belongs to (it may not be destroyed meanwhile)
    int solver_kind = 0;   // 0 Jacobi PCG, 1 separable fast solve
    int gas = 0;           // 0 Ar, 1 O2, 2 Ar/O2 (selects the collision step)
};

template <typename R>
struct Case : CaseBase {
    Setup<R> S;
    int precision = 0;
    int device = 0;
    double inner_tol = 1e-4;
    int pcg_maxit = 100000;
    hipStream_t st = nullptr;
    int G = 0, nsp = 2;
    DevConst<R> kc{};
    NodeConst<R> nc{};
    VoltConst vc{};
    // node / setup arrays
    R *Ex = nullptr, *Ey = nullptr, *P1 = nullptr, *P2 = nullptr, *Pion = nullptr, *BackVel1 = nullptr;
    R *vsave = nullptr, *xe = nullptr, *cx = nullptr, *lap = nullptr, *phi = nullptr, *tot = nullptr;
    R *Sigma = nullptr, *Source = nullptr, *den = nullptr, *sigma_sp = nullptr, *V = nullptr, *part = nullptr;
    R *brhs = nullptr, *Area = nullptr, *eps = nullptr, *stage = nullptr;
    int *StructIdx = nullptr, *Boundary = nullptr, *Face = nullptr, *CondID = nullptr, *A_idx = nullptr;
    Route *route = nullptr;
    int *inv_start = nullptr, *inv_ent = nullptr;
    // field solve
    int *pcol = nullptr;
    float *pvalf = nullptr, *pmdf = nullptr;
    double *pvald = nullptr, *pmdd = nullptr, *b64 = nullptr, *x64 = nullptr;
    PcgWork *pw = nullptr;
    PcgResult *d_res = nullptr;
    PcgOperator op{};
    // separable fast solve (refined path)
    bool spectral = false;
    int field_solver = 0;
    SpecDev spec{};
    SpecState *spec_state = nullptr;
    double *spec_part = nullptr;
    std::vector<double *> spec_bufs;
    int spec_cycles = 1;
    // particles: two buffers per species
    PView<R> cur[MAXSP]{}, nxt[MAXSP]{};
    uint8_t *tag[MAXSP] = {nullptr};
    int *key[MAXSP] = {nullptr}, *prank = nullptr;
    int *cnt[MAXSP] = {nullptr}, *base[MAXSP] = {nullptr}, *pbase[MAXSP] = {nullptr};
    int *pcnt = nullptr, *pid_off = nullptr, *total = nullptr, *alive = nullptr, *ptotal = nullptr;
    int *total2 = nullptr, *alive2 = nullptr;
    StepClock *clk = nullptr;   // device clock of the fused step (the step about to run)
    unsigned long long *eedf = nullptr;   // electron energy histogram (EEDF_BINS counts)
    // diagnostics history (diag.hip): per-cell energy partials, absorption
    // counts at the previous sample, a ring of DIAG_RING records
    static constexpr int DIAG_RING = 4096;
    double *diag_part = nullptr, *diag_hist = nullptr;
    int *diag_cc_prev = nullptr;
    long long diag_n = 0;
    long long tstep_base = 0;   // step at which the counters (mccr, CondCharge, history) last started
    // metastable fluid (fluid.hip), allocated when first switched on
    bool fluid_on = false;
    FluidDev<R> fl{};
    std::vector<void *> fl_bufs;
    double eedf_emax = 0;
    long long eedf_n = 0;
    bool clk_dirty = true;      // host tstep changed outside the fused path
    // captured fused steps, one per buffer parity (key = electron x buffer in use)
    hipGraphExec_t gexec[2] = {nullptr, nullptr};
    const void *gkey[2] = {nullptr, nullptr};
    bool use_graph = true;
    long long stage_stride = 0;   // stride of the 6 arrays of `stage`
    int *pstart = nullptr, *scan_tmp = nullptr, *cellp = nullptr;
    uint32_t *next_id = nullptr;
    unsigned long long *mccr = nullptr;
    int *condcharge = nullptr, *d_err = nullptr, *lost = nullptr;
    unsigned long long *sigma_hits = nullptr;   // nsp x G, this step's dielectric hits in fixed point
    bool diel_routes = false;                   // some cell border leads into a dielectric
    // oxygen chemistries (gas types 1 and 2): constants and buffers of their
    // collision step
    O2Const<R> oc{};
    R *o2prob = nullptr, *o2fluid = nullptr, *o2bg = nullptr, *o2bg2 = nullptr, *o2bv2 = nullptr;
    R *o2stash[CHEM_MAXSP] = {nullptr};
    uint8_t *o2fate[CHEM_MAXSP] = {nullptr};
    int *o2pcount = nullptr, *o2pstart = nullptr, *o2delreq = nullptr, *o2total = nullptr;
    unsigned long long *o2delcnt = nullptr;   // O- removed by electron detachment, per node, cumulative
    double *sum_den = nullptr, *sum_phi = nullptr, *sum_ex = nullptr, *sum_ey = nullptr;
    long long n_accum = 0;
    long long tstep = 0;
    double t = 0.0;
    int tag_valid = 0;
    // optional stage timers: HIP events on the case stream, recorded into a
    // ring of RING steps so that timing never stalls the launch pipeline; a
    // slot is read back only when the ring wraps or the times are requested.
    // Stage k of a step = events k .. k+1: field, efield, count_e, count_i,
    // bases, scatter_e, scatter_i, deposit, accumulate.
    static constexpr int NT = 9, RING = 32;
    bool timing = false;
    hipEvent_t ev[RING][NT + 1] = {};
    bool ev_pending[RING] = {};
    int ev_slot = 0;
    long long ev_rec = 0;
    double t_ms[NT] = {0};
    long long t_n = 0;
    void mark(int k) {
        if (timing) (void)hipEventRecord(ev[ev_slot][k], st);
    }
    int flush_slot(int sl) {
        if (!ev_pending[sl]) return 0;
        HIPCHK(hipEventSynchronize(ev[sl][NT]));
        for (int k = 0; k < NT; ++k) {
            float ms = 0.f;
            HIPCHK(hipEventElapsedTime(&ms, ev[sl][k], ev[sl][k + 1]));
            t_ms[k] += ms;
        }
        ++t_n;
        ev_pending[sl] = false;
        return 0;
    }

    ~Case() override {
        if (st) {
            const hipError_t e = hipStreamSynchronize(st);
            if (e != hipSuccess) fprintf(stderr, "picmcc: case teardown: %s\n", hipGetErrorString(e));
        }
        void *ptrs[] = {Ex, Ey, P1, P2, Pion, BackVel1, vsave, xe, cx, lap, phi, tot, Sigma, Source, den, sigma_sp,
                        V, part, brhs, Area, eps, stage, StructIdx, Boundary, Face, CondID, A_idx, route, inv_start,
                        inv_ent, pcol, pvalf, pmdf, pvald, pmdd, b64, x64, d_res, prank, pcnt, pid_off, total, alive,
                        ptotal, pstart, scan_tmp, cellp, next_id, mccr, condcharge, d_err, sum_den, sum_phi, sum_ex,
                        sum_ey, total2, alive2, clk, eedf, lost, sigma_hits, o2prob, o2fluid, o2bg, o2bg2, o2bv2,
                        o2pcount, o2pstart, o2delreq, o2total, o2delcnt, diag_part, diag_hist, diag_cc_prev};
        for (void *p : ptrs)
            if (p) (void)hipFree(p);
        for (void *p : fl_bufs)
            if (p) (void)hipFree(p);
        for (int s = 0; s < CHEM_MAXSP; ++s) {
            if (o2fate[s]) (void)hipFree(o2fate[s]);
            if (o2stash[s]) (void)hipFree(o2stash[s]);
        }
        for (int s = 0; s < MAXSP; ++s) {
            PView<R> *vs[2] = {&cur[s], &nxt[s]};
            for (PView<R> *v : vs) {
                void *q[] = {v->x, v->y, v->vx, v->vy, v->vz, v->id, v->cell_start};
                for (void *p : q)
                    if (p) (void)hipFree(p);
            }
            if (tag[s]) (void)hipFree(tag[s]);
            if (key[s]) (void)hipFree(key[s]);
            if (cnt[s]) (void)hipFree(cnt[s]);
            if (base[s]) (void)hipFree(base[s]);
            if (pbase[s]) (void)hipFree(pbase[s]);
        }
        pcg_work_destroy(pw);
        for (int q = 0; q < 2; ++q)
            if (gexec[q]) (void)hipGraphExecDestroy(gexec[q]);
        for (int r = 0; r < RING; ++r)
            for (int k = 0; k <= NT; ++k)
                if (ev[r][k]) (void)hipEventDestroy(ev[r][k]);
        for (double *p : spec_bufs) (void)hipFree(p);
        if (spec_state) (void)hipFree(spec_state);
        if (spec_part) (void)hipFree(spec_part);
        if (st) (void)hipStreamDestroy(st);
    }

    // ------------------------------------------------------------ routing tables
    // Destination of a particle leaving cell c in each direction, restating
    // the structure-map lookup of src/cuda_Sortboundary.cu:68-257.
    void build_routes(std::vector<Route> &rt, std::vector<int> &istart, std::vector<int> &ient) {
        const int ngy = S.ngy, ncx = S.ncx, ncy = S.ncy;
        rt.assign((size_t)G * ROUTE_SLOTS, Route{-1, 0});
        std::vector<std::vector<std::pair<int, int>>> inv(G);   // (source cell, entry)
        for (int i = 0; i < ncx; ++i)
            for (int j = 0; j < ncy; ++j) {
                const int c = i * ngy + j;
                int dests[ROUTE_SLOTS];
                int nslot = 0;
                for (int code = 0; code < ROUTE_SLOTS; ++code) {
                    Route r{-1, 0};
                    if (code == 4) {
                        r.dest = c;
                        r.info = 0;
                        rt[(size_t)c * ROUTE_SLOTS + code] = r;
                        continue;
                    }
                    const int mcid = code_mcid(code, ngy);
                    const int SI = S.StructIdx[padded_index(c, mcid, ngy)];
                    const int cid = c + mcid;
                    int flags = 0;
                    if (SI == 0) r.dest = cid;
                    else if (SI >= 1 && SI < 100) { r.dest = -3; diel_routes = true; }
                    else if (SI >= 100) { r.dest = -2; r.info = SI - 101; }
                    else {
                        int sum = 0, ok = 1;
                        switch (SI) {
                        case -2: sum = ngy; flags = 16 | 64; break;
                        case -6: sum = -ngy; flags = 16 | 64; break;
                        case -4: sum = 1; flags = 32 | 128; break;
                        case -8: sum = -1; flags = 32 | 128; break;
                        case -12: sum = ngy + 1; flags = 16 | 32 | 64 | 128; break;
                        case -16: sum = 1 - ngy; flags = 16 | 32 | 64 | 128; break;
                        case -20: sum = ngy - 1; flags = 16 | 32 | 64 | 128; break;
                        case -24: sum = -ngy - 1; flags = 16 | 32 | 64 | 128; break;
                        default: ok = 0;
                        }
                        r.dest = ok ? cid + sum : -1;
                    }
                    if (r.dest >= 0) {
                        int slot = -1;
                        for (int q = 0; q < nslot; ++q)
                            if (dests[q] == r.dest) slot = q + 1;
                        if (slot < 0) {
                            dests[nslot++] = r.dest;
                            slot = nslot;
                            inv[r.dest].push_back({c, c * ROUTE_SLOTS + slot});
                        }
                        r.info = slot | flags;
                    }
                    rt[(size_t)c * ROUTE_SLOTS + code] = r;
                }
            }
        istart.assign(G + 1, 0);
        ient.clear();
        for (int d = 0; d < G; ++d) {
            const int i = d / ngy, j = d % ngy;
            if (i < ncx && j < ncy) ient.push_back(d * ROUTE_SLOTS + 0);
            auto &v = inv[d];
            std::stable_sort(v.begin(), v.end(), [](const std::pair<int, int> &a, const std::pair<int, int> &b) {
                return a.first < b.first;
            });
            for (auto &e : v) ient.push_back(e.second);
            istart[d + 1] = (int)ient.size();
        }
    }

    int alloc_species(int s, int cap) {
        for (PView<R> *v : {&cur[s], &nxt[s]}) {
            HIPCHK(dalloc(&v->x, cap)); HIPCHK(dalloc(&v->y, cap)); HIPCHK(dalloc(&v->vx, cap));
            HIPCHK(dalloc(&v->vy, cap)); HIPCHK(dalloc(&v->vz, cap)); HIPCHK(dalloc(&v->id, cap));
            HIPCHK(dalloc(&v->cell_start, G + 1));
            HIPCHK(hipMemsetAsync(v->cell_start, 0, (G + 1) * sizeof(int), st));
            v->cap = cap;
        }
        HIPCHK(dalloc(&tag[s], cap));
        HIPCHK(dalloc(&key[s], cap));
        HIPCHK(dalloc(&cnt[s], (size_t)G * ROUTE_SLOTS));
        HIPCHK(dalloc(&base[s], (size_t)G * ROUTE_SLOTS));
        HIPCHK(dalloc(&pbase[s], (size_t)G * ROUTE_SLOTS));
        return 0;
    }

    int init(const Deck &d, const pic_options &o) {
        precision = o.precision;
        device = o.device;
        if (o.inner_tol > 0) inner_tol = o.inner_tol;
        if (o.pcg_maxit > 0) pcg_maxit = o.pcg_maxit;
        field_solver = o.field_solver;
        HIPCHK(hipSetDevice(device));
        std::string err;
        int e = build_setup<R>(d, S, err, o.load != 0);
        if (e) return fail(e, "%s", err.c_str());
        rf_cycles = d.rf_cycles;
        dt_pic = d.dt_pic;
        G = S.Gsize;
        nsp = S.nsp;
        // every transfer and fill below is enqueued on the case stream, so
        // kernels on that stream are ordered after them (a non-blocking
        // stream is not ordered after null-stream work)
        HIPCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        // node arrays
        HIPCHK(dalloc(&Ex, G)); HIPCHK(dalloc(&Ey, G)); HIPCHK(dalloc(&phi, G)); HIPCHK(dalloc(&tot, G));
        HIPCHK(dalloc(&Sigma, G)); HIPCHK(dalloc(&Source, G));
        HIPCHK(dalloc(&den, (size_t)nsp * G)); HIPCHK(dalloc(&sigma_sp, (size_t)nsp * G));
        HIPCHK(dalloc(&part, (size_t)nsp * G * 4)); HIPCHK(dalloc(&V, 64));
        HIPCHK(dalloc(&brhs, S.A_size));
        const void *zero_targets[] = {};
        (void)zero_targets;
        HIPCHK(hipMemsetAsync(Ex, 0, G * sizeof(R), st)); HIPCHK(hipMemsetAsync(Ey, 0, G * sizeof(R), st));
        HIPCHK(hipMemsetAsync(phi, 0, G * sizeof(R), st)); HIPCHK(hipMemsetAsync(tot, 0, G * sizeof(R), st));
        HIPCHK(hipMemsetAsync(Sigma, 0, G * sizeof(R), st)); HIPCHK(hipMemsetAsync(Source, 0, G * sizeof(R), st));
        HIPCHK(hipMemsetAsync(den, 0, (size_t)nsp * G * sizeof(R), st));
        HIPCHK(hipMemsetAsync(sigma_sp, 0, (size_t)nsp * G * sizeof(R), st));
        HIPCHK(hipMemsetAsync(part, 0, (size_t)nsp * G * 4 * sizeof(R), st));
        HIPCHK(hipMemsetAsync(brhs, 0, std::max(1, S.A_size) * sizeof(R), st));
        HIPCHK(upload(&P1, S.P1, st)); HIPCHK(upload(&P2, S.P2, st)); HIPCHK(upload(&Pion, S.Pion, st));
        HIPCHK(upload(&BackVel1, S.BackVel1, st)); HIPCHK(upload(&vsave, S.vsave, st));
        HIPCHK(upload(&xe, S.xe, st)); HIPCHK(upload(&cx, S.cx, st)); HIPCHK(upload(&lap, S.lap_phi, st));
        HIPCHK(upload(&StructIdx, S.StructIdx, st)); HIPCHK(upload(&Boundary, S.Boundary, st));
        HIPCHK(upload(&Face, S.Face, st)); HIPCHK(upload(&CondID, S.CondID, st)); HIPCHK(upload(&A_idx, S.A_idx, st));
        HIPCHK(upload(&Area, S.Area, st)); HIPCHK(upload(&eps, S.eps, st));
        std::vector<Route> rt;
        std::vector<int> istart, ient;
        build_routes(rt, istart, ient);
        HIPCHK(upload(&route, rt, st)); HIPCHK(upload(&inv_start, istart, st)); HIPCHK(upload(&inv_ent, ient, st));
        // the bases kernels walk at most 9 incoming entries per cell
        for (int d = 0; d < G; ++d)
            if (istart[d + 1] - istart[d] > ROUTE_SLOTS)
                return fail(-43, "a cell receives particles from more than 9 source entries");
        // field operator in padded 5-entry rows
        {
            const int N = S.A_size;
            std::vector<int> col(5 * (size_t)N);
            std::vector<float> vf(5 * (size_t)N, 0.0f), mf(N);
            std::vector<double> vd(5 * (size_t)N, 0.0), md(N);
            for (int r = 0; r < N; ++r) {
                int k = 0;
                for (int q = S.Ai[r]; q < S.Ai[r + 1]; ++q, ++k) {
                    col[5 * r + k] = S.Aj[q];
                    vf[5 * r + k] = (float)S.Aval[q];
                    vd[5 * r + k] = (double)S.Aval[q];
                }
                for (; k < 5; ++k) col[5 * r + k] = r;
                mf[r] = (float)S.Mdiag[r];
                md[r] = (double)S.Mdiag[r];
            }
            HIPCHK(upload(&pcol, col, st)); HIPCHK(upload(&pvalf, vf, st)); HIPCHK(upload(&pvald, vd, st));
            HIPCHK(upload(&pmdf, mf, st)); HIPCHK(upload(&pmdd, md, st));
            HIPCHK(dalloc(&b64, N)); HIPCHK(dalloc(&x64, N));
            HIPCHK(hipMemsetAsync(x64, 0, std::max(1, N) * sizeof(double), st));
            HIPCHK(hipMemsetAsync(b64, 0, std::max(1, N) * sizeof(double), st));
            op.N = N; op.col = pcol; op.valf = pvalf; op.vald = pvald; op.mdiagf = pmdf; op.mdiagd = pmdd;
            int werr = 0;
            const char *nb_env = getenv("PIC_PCG_BLOCKS");
            pw = pcg_work_create(N, &werr, nb_env ? atoi(nb_env) : 0, st);
            if (!pw) return fail(-40, "field solver workspace allocation failed");
            HIPCHK(dalloc(&d_res, 1));
            HIPCHK(hipMemsetAsync(d_res, 0, sizeof(PcgResult), st));
            if ((precision == PIC_FP32_REFINED && field_solver == 0) || field_solver == 2) {
                Spectral sp;
                std::string why;
                if (build_spectral<R>(S, sp, why)) {
                    const int nx = sp.nx, ny = sp.ny;
                    std::vector<double> VyiT((size_t)ny * ny), VyT((size_t)ny * ny), dinv((size_t)nx * ny);
                    for (int a = 0; a < ny; ++a)
                        for (int b2 = 0; b2 < ny; ++b2) {
                            VyiT[(size_t)a * ny + b2] = sp.Vyi[(size_t)b2 * ny + a];
                            VyT[(size_t)a * ny + b2] = sp.Vy[(size_t)b2 * ny + a];
                        }
                    for (int a = 0; a < nx; ++a)
                        for (int b2 = 0; b2 < ny; ++b2) {
                            double den = sp.cx * sp.lx[a] + sp.cy * sp.ly[b2];
                            dinv[(size_t)a * ny + b2] = den > 1e-300 ? 1.0 / den : 0.0;
                        }
                    double *p;
                    const std::vector<double> *srcs[5] = {&sp.Vxi, &VyiT, &sp.Vx, &VyT, &dinv};
                    const double *dst[5];
                    for (int q = 0; q < 5; ++q) {
                        HIPCHK(upload(&p, *srcs[q], st));
                        spec_bufs.push_back(p);
                        dst[q] = p;
                    }
                    spec.nx = nx; spec.ny = ny;
                    spec.Vxi = dst[0]; spec.VyiT = dst[1]; spec.Vx = dst[2]; spec.VyT = dst[3]; spec.dinv = dst[4];
                    HIPCHK(dalloc(&p, (size_t)nx * ny)); spec_bufs.push_back(p); spec.T1 = p;
                    HIPCHK(dalloc(&p, (size_t)nx * ny)); spec_bufs.push_back(p); spec.T2 = p;
                    HIPCHK(dalloc(&p, (size_t)nx * ny)); spec_bufs.push_back(p); spec.r = p;
                    HIPCHK(dalloc(&spec_state, 1));
                    HIPCHK(dalloc(&spec_part, 2 * ((N + 255) / 256) + 2));
                    // an exact separable model converges in one FP64 cycle; otherwise
                    // the model error (~1e-7 relative) needs a few refinement cycles
                    spec_cycles = sp.exact ? 1 : 4;
                    spectral = true;
                    solver_kind = 1;
                }
            }
        }
        // the persistent solver's launch shape, tuned on this operator when
        // the PCG is the field solver (PIC_PCG_BLOCKS fixes the grid instead)
        if (!spectral && !getenv("PIC_PCG_BLOCKS") && S.A_size > 0) {
            if (int e2 = pcg_autotune(pw, op, precision == PIC_FP32_REFINED, st, nullptr))
                return fail(e2, "field solver autotuning failed");
        }
        // particles
        int capmax = 0;
        for (int s = 0; s < nsp; ++s) {
            int cap = std::max(S.sp[s].cap, 1);
            capmax = std::max(capmax, cap);
            if (int e2 = alloc_species(s, cap)) return e2;
        }
        HIPCHK(dalloc(&stage, (size_t)6 * capmax));
        stage_stride = capmax;
        HIPCHK(dalloc(&prank, capmax));
        HIPCHK(dalloc(&pcnt, (size_t)G * ROUTE_SLOTS));
        HIPCHK(dalloc(&pid_off, (size_t)G * ROUTE_SLOTS));
        HIPCHK(dalloc(&total, G)); HIPCHK(dalloc(&alive, G)); HIPCHK(dalloc(&ptotal, G));
        HIPCHK(dalloc(&pstart, G + 1)); HIPCHK(dalloc(&cellp, G));
        HIPCHK(dalloc(&total2, G)); HIPCHK(dalloc(&alive2, G));
        HIPCHK(dalloc(&scan_tmp, scan_tmp_ints(G)));
        HIPCHK(dalloc(&next_id, MAXSP));
        HIPCHK(dalloc(&mccr, (size_t)G * S.TnRct));
        HIPCHK(hipMemsetAsync(mccr, 0, (size_t)G * S.TnRct * sizeof(unsigned long long), st));
        HIPCHK(dalloc(&condcharge, nsp * S.CondNUMR));
        HIPCHK(hipMemsetAsync(condcharge, 0, nsp * S.CondNUMR * sizeof(int), st));
        HIPCHK(dalloc(&lost, nsp));
        HIPCHK(hipMemsetAsync(lost, 0, nsp * sizeof(int), st));
        HIPCHK(dalloc(&sigma_hits, (size_t)nsp * G));
        HIPCHK(hipMemsetAsync(sigma_hits, 0, (size_t)nsp * G * sizeof(unsigned long long), st));
        HIPCHK(dalloc(&d_err, 1));
        HIPCHK(hipMemsetAsync(d_err, 0, sizeof(int), st));
        HIPCHK(dalloc(&clk, 1));
        clk_dirty = true;
        if (const char *ng = getenv("PIC_NO_GRAPH")) use_graph = atoi(ng) == 0;
        HIPCHK(dalloc(&sum_den, (size_t)nsp * G)); HIPCHK(dalloc(&sum_phi, G));
        HIPCHK(dalloc(&sum_ex, G)); HIPCHK(dalloc(&sum_ey, G));
        HIPCHK(hipMemsetAsync(sum_den, 0, (size_t)nsp * G * sizeof(double), st));
        HIPCHK(hipMemsetAsync(sum_phi, 0, G * sizeof(double), st));
        HIPCHK(hipMemsetAsync(sum_ex, 0, G * sizeof(double), st));
        HIPCHK(hipMemsetAsync(sum_ey, 0, G * sizeof(double), st));
        {
            uint32_t nid[MAXSP] = {0};
            for (int s = 0; s < nsp; ++s) nid[s] = (uint32_t)S.sp[s].n;
            HIPCHK(hipMemcpyAsync(next_id, nid, sizeof nid, hipMemcpyHostToDevice, st));
            HIPCHK(hipStreamSynchronize(st));   // nid lives on this stack frame
        }
        for (int s = 0; s < nsp; ++s) {
            SpeciesInit<R> &si = S.sp[s];
            if (si.n > 0) {
                HIPCHK(hipMemcpyAsync(cur[s].x, si.x.data(), si.n * sizeof(R), hipMemcpyHostToDevice, st));
                HIPCHK(hipMemcpyAsync(cur[s].y, si.y.data(), si.n * sizeof(R), hipMemcpyHostToDevice, st));
                HIPCHK(hipMemcpyAsync(cur[s].vx, si.vx.data(), si.n * sizeof(R), hipMemcpyHostToDevice, st));
                HIPCHK(hipMemcpyAsync(cur[s].vy, si.vy.data(), si.n * sizeof(R), hipMemcpyHostToDevice, st));
                HIPCHK(hipMemcpyAsync(cur[s].vz, si.vz.data(), si.n * sizeof(R), hipMemcpyHostToDevice, st));
                HIPCHK(hipMemcpyAsync(cur[s].id, si.id.data(), si.n * sizeof(uint32_t), hipMemcpyHostToDevice, st));
            }
            if (!si.cell_start.empty())
                HIPCHK(hipMemcpyAsync(cur[s].cell_start, si.cell_start.data(), (G + 1) * sizeof(int), hipMemcpyHostToDevice, st));
            // free the host copy of the load once the copies have landed
            HIPCHK(hipStreamSynchronize(st));
            std::vector<R>().swap(si.x); std::vector<R>().swap(si.y); std::vector<R>().swap(si.vx);
            std::vector<R>().swap(si.vy); std::vector<R>().swap(si.vz); std::vector<uint32_t>().swap(si.id);
        }
        // constants
        kc.ngx = S.ngx; kc.ngy = S.ngy; kc.ncx = S.ncx; kc.ncy = S.ncy; kc.Gsize = G; kc.Csize = S.Csize;
        kc.nsp = nsp; kc.CondNUMR = S.CondNUMR; kc.TnRct = S.TnRct; kc.nvel = S.nvel; kc.MCCn = S.MCCn;
        kc.nlogx = N_LOGX; kc.dt_dx = S.dt_dx; kc.dt_dy = S.dt_dy; kc.idLOGX = S.idLOGX;
        for (int s = 0; s < MAXSP; ++s) { kc.Ascale[s] = S.Ascale[s]; kc.Escale[s] = S.Escale[s]; }
        for (int r = 0; r < 3; ++r) kc.sigv[r] = S.sigv[r];
        for (int r = 0; r < 16; ++r) kc.Th_e[r] = S.Th_e[r];
        kc.seed = (uint64_t)d.rng_seed;
        kc.Ex = Ex; kc.Ey = Ey; kc.P1 = P1; kc.P2 = P2; kc.Pion = Pion; kc.BackVel1 = BackVel1;
        kc.vsave = vsave; kc.xe = xe; kc.cx = cx; kc.StructIdx = StructIdx; kc.route = route; kc.lost = lost;
        kc.sigma_hits = sigma_hits;
        gas = d.gas;
        if (gas == 1 || gas == 2) {
            const bool aro2 = gas == 2;
            const int want_nsp = aro2 ? ARO2_NSP : O2_NSP, want_sigv = aro2 ? ARO2_NSIGV : O2_NSIGV;
            const int want_rows = aro2 ? ARO2_NROWS : O2_NROWS;
            if (nsp != want_nsp || S.nsigv != want_sigv || S.TnRct < want_rows || (aro2 && S.BackDen2.empty()))
                return fail(-19, "oxygen setup does not match the collision tables");
            oc.chem = aro2 ? CHEM_ARO2 : CHEM_O2;
            oc.nsp = nsp;
            oc.ntgt = aro2 ? ARO2_NTARGETS : O2_NTARGETS;
            oc.neg = aro2 ? ARO2_NEG : O2_NEG;
            oc.sub_mask = aro2 ? ARO2_SUBCYCLED : O2_SUBCYCLED;
            oc.victim_tag = aro2 ? ARO2_VICTIM_TAG : O2_VICTIM_TAG;
            HIPCHK(dalloc(&o2prob, (size_t)G * oc.ntgt));
            HIPCHK(upload(&o2fluid, S.FluidN, st)); HIPCHK(upload(&o2bg, S.BackDen1, st));
            if (aro2) { HIPCHK(upload(&o2bg2, S.BackDen2, st)); HIPCHK(upload(&o2bv2, S.BackVel2, st)); }
            HIPCHK(dalloc(&o2pcount, (size_t)nsp * G)); HIPCHK(dalloc(&o2pstart, (size_t)nsp * (G + 1)));
            HIPCHK(dalloc(&o2delreq, G));
            HIPCHK(dalloc(&o2total, (size_t)nsp * G));
            HIPCHK(dalloc(&o2delcnt, G));
            HIPCHK(hipMemsetAsync(o2delcnt, 0, (size_t)G * sizeof(unsigned long long), st));
            for (int s = 0; s < nsp; ++s) {
                HIPCHK(dalloc(&o2fate[s], cur[s].cap));
                HIPCHK(dalloc(&o2stash[s], (size_t)6 * cur[s].cap));
            }
            oc.G = G; oc.ncy = S.ncy; oc.ngy = S.ngy; oc.Csize = S.Csize; oc.TnRct = S.TnRct; oc.nlogx = N_LOGX;
            oc.nvel = S.nvel; oc.MCCn = S.MCCn;
            // the Ar/O2 dispatch runs every species' handler (src/cuda_mcc.cu:25); the O2 one only
            // the electrons' unless the deck asks for the ions too
            oc.ions = aro2 || d.o2_ion_coll != 0;
            oc.idLOGX = S.idLOGX; oc.dt = S.dt; oc.dtm = S.dt_mcc; oc.idx = S.idx; oc.idy = S.idy;
            for (int s = 0; s < nsp; ++s) { oc.np2c[s] = S.np2c[s]; oc.Escale[s] = S.Escale[s]; }
            for (int j = 0; j < want_sigv; ++j) oc.sigv[j] = S.sigv[j];
            for (int r = 0; r < want_rows; ++r) oc.Th_e[r] = S.Th_e[r];
            oc.seed = (uint64_t)d.rng_seed;
            oc.xe = xe; oc.cx = cx; oc.vsave = vsave; oc.BackDen1 = o2bg; oc.BackVel1 = BackVel1; oc.FluidN = o2fluid;
            oc.BackDen2 = aro2 ? o2bg2 : o2bg; oc.BackVel2 = aro2 ? o2bv2 : BackVel1;
            oc.den = den; oc.prob = o2prob;
        }
        nc.Gsize = G; nc.ngx = S.ngx; nc.ngy = S.ngy; nc.nsp = nsp; nc.CondNUMR = S.CondNUMR;
        nc.hdx = S.hdx; nc.hdy = S.hdy; nc.idx = S.idx; nc.idy = S.idy; nc.idxy = S.idxy;
        nc.hxy = (R)(0.5 * (double)S.dx / (double)S.dy);
        nc.hyx = (R)(0.5 * (double)S.dy / (double)S.dx);
        for (int s = 0; s < MAXSP; ++s) nc.q_density[s] = S.q_density[s];
        nc.Boundary = Boundary; nc.Face = Face; nc.CondID = CondID; nc.A_idx = A_idx; nc.Area = Area; nc.eps = eps;
        nc.sigma_hits = diel_routes ? sigma_hits : nullptr;
        if (S.CondNUMR > MAX_SRC || (int)d.src.size() > MAX_SRC) return fail(-41, "too many conductors or sources");
        vc.ncond = S.CondNUMR;
        vc.nsrc = (int)d.src.size();
        for (int i = 0; i < S.CondNUMR; ++i) vc.flag[i] = S.Efield_Flag[i];
        const double PI = 3.14159265358979323846;
        for (int j = 0; j < vc.nsrc; ++j) {
            vc.mid[j] = d.src[j].mid;
            vc.dc[j] = (double)(R)d.src[j].dc;
            vc.ac[j] = (double)(R)d.src[j].ac;
            vc.w[j] = (double)(R)(2 * PI * (double)(R)d.src[j].freq);
            vc.ph[j] = (double)(R)(PI / 180 * (double)(R)d.src[j].phase);
        }
        // the initial deposit (src/cuda_main.cu:28)
        if (int e3 = deposit()) return e3;
        HIPCHK(hipStreamSynchronize(st));
        return 0;
    }

    void swap_species(int s) { std::swap(cur[s], nxt[s]); }

    // ------------------------------------------------------------ stages
    void host_tick() {
        ++tstep;
        t = (double)tstep * (double)S.dt;
    }
    int advance_time() override {
        host_tick();
        clk_dirty = true;
        return 0;
    }
    // bring the device clock to the host's step count before fused steps
    int ensure_clock() {
        if (!clk_dirty) return 0;
        StepClock h{(uint32_t)(tstep + 1), 0u, (double)(tstep + 1) * (double)S.dt};
        HIPCHK(hipMemcpyAsync(clk, &h, sizeof h, hipMemcpyHostToDevice, st));
        HIPCHK(hipStreamSynchronize(st));   // h lives on this stack frame
        clk_dirty = false;
        return 0;
    }
    // the solve itself: FP64 solution in x64 (warm start) from the rhs brhs
    int solve_x64() {
        const int N = S.A_size;
        const double tol = S.deck.pcg_tol > 0 ? S.deck.pcg_tol : 1e-10;
        if (spectral) {
            // the first residual converts the rhs into b64
            spectral_solve_refined<R>(op, spec, brhs, b64, x64, tol, spec_cycles, spec_state, spec_part, d_res, st);
        } else {
            launch_rhs_to_fp64<R>(N, brhs, b64, st);
            if (precision == PIC_FP64_PLAIN) HIPCHK(pcg_solve_plain_fp64(op, b64, x64, tol, pcg_maxit, pw, d_res, st));
            else HIPCHK(pcg_solve_refined(op, b64, x64, tol, inner_tol, pcg_maxit, pw, d_res, st));
        }
        HIPCHK(hipGetLastError());
        return 0;
    }
    int field_solve() override {
        if (int e = solve_x64()) return e;
        launch_x_to_phi<R>(G, A_idx, x64, phi, st);
        return 0;
    }
    // fused step: solve, then phi + voltages + potential in one kernel and E
    int field_efield_fused() {
        if (int e = solve_x64()) return e;
        mark(1);
        launch_phi_potential<R>(nc, vc, clk, x64, lap, V, phi, tot, st);
        launch_efield_only<R>(nc, phi, tot, den, Sigma, Ex, Ey, st);
        HIPCHK(hipGetLastError());
        return 0;
    }
    int efield() override {
        launch_efield<R>(nc, vc, t, lap, V, phi, tot, den, Sigma, Ex, Ey, st);
        HIPCHK(hipGetLastError());
        return 0;
    }
    int move() override {
        for (int s = 0; s < nsp; ++s) launch_push<R>(kc, cur[s], s, tag[s], st);
        tag_valid = 1;
        HIPCHK(hipGetLastError());
        return 0;
    }
    int sort_boundary() override {
        if (!tag_valid) return fail(-42, "SortBounndary_cuda called without a preceding Move_cuda");
        for (int s = 0; s < nsp; ++s) {
            launch_sort_count<R>(kc, cur[s], s, tag[s], key[s], cnt[s], condcharge, st);
            launch_totals(G, inv_start, inv_ent, cnt[s], nullptr, total, alive, nullptr, st);
            launch_scan(total, nxt[s].cell_start, G, scan_tmp, st);
            launch_assign_bases(G, inv_start, inv_ent, cnt[s], nullptr, nxt[s].cell_start, alive, nullptr, base[s],
                                nullptr, nullptr, st);
            launch_sort_scatter<R>(kc, cur[s], nxt[s], tag[s], key[s], base[s], nxt[s].cap, d_err, st);
            swap_species(s);
        }
        tag_valid = 0;
        HIPCHK(hipGetLastError());
        return 0;
    }
    // Move and SortBounndary of a whole step: per species the push and the
    // destination count in one pass, then the rest of the sort.  The same
    // result as move() then sort_boundary() (no species' sort depends on
    // another's push); the stage entry points keep the two passes apart.
    // The bookkeeping between the two passes runs for all species at once:
    // one totals launch, the scans three arrays per pass, one bases launch
    // (the per-species totals live in o2total, which the collision step
    // clears before its own use).
    int move_sort() {
        if (nsp > 5) return fail(-19, "staged whole steps take at most 5 species");
        SortSpecies j{};
        StagedSet<R> set{};
        const int *in[CHEM_MAXSP];
        int *out[CHEM_MAXSP];
        for (int s = 0; s < nsp; ++s) {
            set.p[s] = cur[s]; set.q[s] = nxt[s]; set.tag[s] = tag[s]; set.key[s] = key[s]; set.cnt[s] = cnt[s];
            set.base[s] = base[s];
            j.cnt[s] = cnt[s];
            j.total[s] = o2total + (size_t)s * G;
            j.start[s] = nxt[s].cell_start;
            j.base[s] = base[s];
            in[s] = j.total[s];
            out[s] = nxt[s].cell_start;
        }
        launch_push_count<R>(kc, set, nsp, condcharge, st);
        launch_totals_sp(G, nsp, inv_start, inv_ent, j, st);
        for (int a = 0; a < nsp; a += 3) launch_scan_multi(std::min(3, nsp - a), in + a, out + a, G, scan_tmp, st);
        launch_assign_bases_sp(G, nsp, inv_start, inv_ent, j, st);
        launch_sort_scatter_all<R>(kc, set, nsp, d_err, st);
        for (int s = 0; s < nsp; ++s) swap_species(s);
        tag_valid = 0;
        HIPCHK(hipGetLastError());
        return 0;
    }
    // oxygen chemistries: collide every species, remove the O- taken by
    // electron detachment, then rebuild every bucket with the products
    int mcc_o2() {
        const uint32_t step = (uint32_t)tstep;
        O2Set<R> set;
        for (int s = 0; s < nsp; ++s) { set.p[s] = cur[s]; set.fate[s] = o2fate[s]; set.stash[s] = o2stash[s]; }
        launch_o2_probs<R>(oc, st);
        launch_fill_int(o2pcount, 0, nsp * G, st);
        launch_fill_int(o2delreq, 0, G, st);
        launch_o2_collide<R>(oc, set, -1, step, mccr, o2pcount, o2delreq, st);   // all species
        launch_o2_victims<R>(oc, set, step, o2delreq, o2delcnt, st);
        O2Set<R> dst = set;
        launch_fill_int(o2total, 0, nsp * G, st);
        launch_o2_totals<R>(oc, set, o2pcount, o2total, st);
        // 2 nsp exclusive scans (new cell_start and product order of each
        // species), three arrays per pass
        const int nscan = 2 * nsp;
        const int *in[2 * CHEM_MAXSP];
        int *out[2 * CHEM_MAXSP];
        for (int t = 0; t < nsp; ++t) {
            in[t] = o2total + (size_t)t * G; out[t] = nxt[t].cell_start;
            in[nsp + t] = o2pcount + (size_t)t * G; out[nsp + t] = o2pstart + (size_t)t * (G + 1);
            dst.p[t] = nxt[t];
        }
        for (int a = 0; a < nscan; a += 3) launch_scan_multi(std::min(3, nscan - a), in + a, out + a, G, scan_tmp, st);
        launch_o2_scatter_all<R>(oc, set, dst, o2pstart, next_id, d_err, st);
        for (int s = 0; s < nsp; ++s) swap_species(s);
        launch_o2_bump_ids(next_id, o2pstart, G, nsp, st);
        HIPCHK(hipGetLastError());
        return 0;
    }
    int mcc() override {
        if (gas == 1 || gas == 2) return mcc_o2();
        const uint32_t step = (uint32_t)tstep;
        // electrons first in the event order is irrelevant: each particle owns its stream
        launch_mcc_collide<R>(kc, cur[1], 1, step, mccr, cellp, prank, stage, st);   // ions (no products)
        // product counts are written per cell; the nodes that start no cell
        // (last row and column) take part in the scan as zeros
        launch_fill_int(pcnt, 0, G, st);
        launch_mcc_collide<R>(kc, cur[0], 0, step, mccr, pcnt, prank, stage, st);    // electrons
        launch_scan(pcnt, pstart, G, scan_tmp, st);
        for (int s = 0; s < nsp; ++s) {
            launch_cell_counts(G, cur[s].cell_start, pcnt, total, st);
            launch_scan(total, nxt[s].cell_start, G, scan_tmp, st);
            launch_mcc_insert<R>(kc, cur[s], nxt[s], nxt[s].cell_start, pstart, cur[0], prank, stage, s == 0 ? 0 : 3,
                                 next_id + s, nxt[s].cap, d_err, st);
        }
        for (int s = 0; s < nsp; ++s) swap_species(s);
        launch_bump_ids(next_id, pstart + G, nsp, st);
        HIPCHK(hipGetLastError());
        return 0;
    }
    int deposit() override {
        launch_deposit<R>(kc, nc, cur, part, den, sigma_sp, Sigma, Source, brhs, nullptr, st);
        HIPCHK(hipGetLastError());
        return 0;
    }
    // adds the current electron energy histogram to the EEDF accumulator;
    // a new emax starts a new accumulation
    int eedf_sample(double emax) override {
        if (!(emax > 0)) return fail(-2, "EEDF energy range must be positive");
        if (!eedf) {
            HIPCHK(dalloc(&eedf, EEDF_BINS));
            eedf_emax = 0;
        }
        if (emax != eedf_emax) {
            HIPCHK(hipMemsetAsync(eedf, 0, EEDF_BINS * sizeof(unsigned long long), st));
            eedf_emax = emax;
            eedf_n = 0;
        }
        launch_eedf<R>(cur[0], G, S.Escale[0], (R)emax, eedf, st);
        HIPCHK(hipGetLastError());
        ++eedf_n;
        return 0;
    }
    int diag_sample() override {
        const int ncond = S.CondNUMR, len = diag_record_len(nsp, ncond);
        if (!diag_hist) {
            HIPCHK(dalloc(&diag_part, (size_t)nsp * S.Csize));
            HIPCHK(dalloc(&diag_hist, (size_t)DIAG_RING * len));
            HIPCHK(dalloc(&diag_cc_prev, (size_t)nsp * ncond));
            HIPCHK(hipMemsetAsync(diag_cc_prev, 0, (size_t)nsp * ncond * sizeof(int), st));
        }
        DiagConst d{};
        d.G = G; d.ngy = S.ngy; d.ncy = S.ncy; d.Csize = S.Csize; d.nsp = nsp; d.ncond = ncond;
        d.dt = (double)S.dt;
        for (int sp = 0; sp < nsp; ++sp) { d.mass[sp] = (double)S.mass[sp]; d.np2c[sp] = (double)S.np2c[sp]; }
        PSet<R> ps{};
        for (int sp = 0; sp < nsp; ++sp) ps.s[sp] = cur[sp];
        double *rec = diag_hist + (size_t)(diag_n % DIAG_RING) * len;
        launch_diag_sample<R>(d, ps, diag_part, Sigma, Area, CondID, condcharge, diag_cc_prev, V, tstep, rec, st);
        HIPCHK(hipGetLastError());
        ++diag_n;
        return 0;
    }
    // the retained records, oldest first
    int diag_history(double *out, int max_records, int *nrec, int *rec_len) override {
        const int len = diag_record_len(nsp, S.CondNUMR);
        if (rec_len) *rec_len = len;
        const long long have = diag_n < DIAG_RING ? diag_n : DIAG_RING;
        const int n = (int)(have < max_records ? have : max_records);
        if (nrec) *nrec = n;
        if (!out || n == 0) return 0;
        HIPCHK(hipStreamSynchronize(st));
        const long long first = diag_n - n;   // the newest n records
        for (int r = 0; r < n; ++r) {
            const long long q = (first + r) % DIAG_RING;
            HIPCHK(hipMemcpyAsync(out + (size_t)r * len, diag_hist + (size_t)q * len, len * sizeof(double),
                                  hipMemcpyDeviceToHost, st));
        }
        HIPCHK(hipStreamSynchronize(st));
        return 0;
    }
    // ------------------------------------------------------------ V000 dump
    // The reference's dump format, version 0.0.0 (writer src/cuda_SaveDump.cu:5-106
    // and 142-155, reader src/load.c:5-120): 4-byte ints and floats, t as a
    // double, everything in the order below.
    struct DumpOut {
        FILE *f;
        bool ok = true;
        template <typename T>
        void put(const T *p, size_t n) {
            if (ok && n && fwrite(p, sizeof(T), n, f) != n) ok = false;
        }
        void i32(int v) { put(&v, 1); }
        void f32(float v) { put(&v, 1); }
        void f64(double v) { put(&v, 1); }
        void zeros(size_t n) {
            static const float z[1024] = {0};
            while (n) {
                const size_t k = std::min<size_t>(n, 1024);
                put(z, k);
                n -= k;
            }
        }
    };
    template <typename T>
    int fetch(std::vector<T> &h, const T *d, size_t n) {
        h.resize(n);
        if (n) HIPCHK(hipMemcpy(h.data(), d, n * sizeof(T), hipMemcpyDeviceToHost));
        return 0;
    }
    int dump_save(const char *path) override {
        HIPCHK(hipStreamSynchronize(st));
        const int ncx = S.ncx, ncy = S.ncy, ngx = S.ngx, ngy = S.ngy, nc = S.CondNUMR;
        // everything the file needs, read back before the file is opened
        std::vector<std::vector<uint32_t>> recs(nsp);
        for (int s = 0; s < nsp; ++s) {
            long long n = 0;
            if (int e = np_of(s, n)) return e;
            recs[s].resize((size_t)n * 6);
            if (n == 0) continue;
            uint32_t *d = nullptr;
            HIPCHK(hipMalloc(&d, (size_t)n * 24));
            launch_dump_pack<R>(cur[s], ncy, ngy, S.Csize, d, st);
            hipError_t e = hipGetLastError();
            if (e == hipSuccess) e = hipMemcpyAsync(recs[s].data(), d, (size_t)n * 24, hipMemcpyDeviceToHost, st);
            if (e == hipSuccess) e = hipStreamSynchronize(st);
            (void)hipFree(d);
            if (e != hipSuccess) return fail(-50, "dump: particle records: %s", hipGetErrorString(e));
        }
        std::vector<R> hden, hsig, htot, hphi, hex, hey, hsrc, hSig;
        std::vector<double> hsd, hsp, hsx, hsy;
        std::vector<unsigned long long> hmccr;
        if (int e = fetch(hden, den, (size_t)nsp * G)) return e;
        if (int e = fetch(hsig, sigma_sp, (size_t)nsp * G)) return e;
        if (int e = fetch(htot, tot, G)) return e;
        if (int e = fetch(hphi, phi, G)) return e;
        if (int e = fetch(hex, Ex, G)) return e;
        if (int e = fetch(hey, Ey, G)) return e;
        if (int e = fetch(hsrc, Source, G)) return e;
        if (int e = fetch(hSig, Sigma, G)) return e;
        if (int e = fetch(hsd, sum_den, (size_t)nsp * G)) return e;
        if (int e = fetch(hsp, sum_phi, G)) return e;
        if (int e = fetch(hsx, sum_ex, G)) return e;
        if (int e = fetch(hsy, sum_ey, G)) return e;
        if (int e = fetch(hmccr, mccr, (size_t)G * S.TnRct)) return e;
        int len = 0, nrec = 0;
        if (int e = diag_history(nullptr, DIAG_RING, &nrec, &len)) return e;
        std::vector<double> hist((size_t)nrec * len);
        if (nrec)
            if (int e = diag_history(hist.data(), nrec, &nrec, &len)) return e;

        FILE *f = fopen(path, "wb");
        if (!f) return fail(-60, "dump: cannot open '%s' for writing", path);
        DumpOut w{f};
        w.i32(0); w.i32(0); w.i32(0);   // version KEY2.KEY1.KEY0
        // time (cstep counts completed RF cycles, src/cuda_main.cu:116)
        w.f64(t);
        w.i32((int)tstep);
        w.i32(dt_pic > 0 ? (int)(tstep / dt_pic) : 0);
        w.i32(gas);
        for (int s = 0; s < nsp; ++s) {
            w.f32((float)S.np2c[s]);
            w.i32((int)(recs[s].size() / 6));
            w.put(recs[s].data(), recs[s].size());
        }
        // fluid species: with the fluid on, its live density, source, face
        // fluxes and continuity flag; otherwise the set-up's densities and
        // fluxes, sources zero
        std::vector<R> fden = S.fl_den, fsrc(S.fl_den.size(), (R)0), ffx = S.fl_flux_x, ffy = S.fl_flux_y;
        FluidState fst{0, 0};
        if (fl.state) {
            if (int e = fetch(fden, fl.den, S.fl_den.size())) return e;
            if (int e = fetch(fsrc, (const R *)fl.src, S.fl_den.size())) return e;
            if (int e = fetch(ffx, (const R *)fl.flux_x, S.fl_flux_x.size())) return e;
            if (int e = fetch(ffy, (const R *)fl.flux_y, S.fl_flux_y.size())) return e;
            HIPCHK(hipMemcpy(&fst, fl.state, sizeof fst, hipMemcpyDeviceToHost));
        }
        w.i32(fst.conti);
        std::vector<float> row((size_t)std::max(ngy, ngx) * std::max(ngx, ngy));
        for (int fs = 0; fs < S.nfsp; ++fs) {
            const size_t o = (size_t)fs * ncx * ncy;
            for (int i = 0; i < ncx; ++i) {
                for (int j = 0; j < ncy; ++j) row[j] = (float)fden[o + (size_t)i * ncy + j];
                w.put(row.data(), ncy);
                for (int j = 0; j < ncy; ++j) row[j] = (float)fsrc[o + (size_t)i * ncy + j];
                w.put(row.data(), ncy);
            }
            const size_t ox = (size_t)fs * ngx * ncy, oy = (size_t)fs * ncx * ngy;
            for (size_t k = 0; k < (size_t)ngx * ncy; ++k) row[k] = (float)ffx[ox + k];
            w.put(row.data(), (size_t)ngx * ncy);
            for (size_t k = 0; k < (size_t)ncx * ngy; ++k) row[k] = (float)ffy[oy + k];
            w.put(row.data(), (size_t)ncx * ngy);
        }
        // history: the retained diagnostics records (pic_diag_sample), oldest first
        const int o_np = 2, o_q = 2 + 2 * nsp, o_ab = o_q + nc, o_v = o_ab + nsp * nc;
        auto rec = [&](int r, int k) { return hist[(size_t)r * len + k]; };
        auto spacing = [&](int r) {
            const double prev = r > 0 ? rec(r - 1, 1) : (double)tstep_base;
            return std::max(1.0, rec(r, 1) - prev);
        };
        w.i32(nrec);
        w.i32(nrec >= 2 ? (int)spacing(nrec - 1) : 1);
        for (int r = 0; r < nrec; ++r) w.f32((float)rec(r, 0));
        for (int s = 0; s < nsp; ++s)
            for (int r = 0; r < nrec; ++r) w.f32((float)rec(r, o_np + s));
        for (int fs = 0; fs < S.nfsp; ++fs) {
            double a = 0;
            int m = 0;
            for (int c = 0; c < S.Csize; ++c)
                if (S.PlasmaRegion[c]) { a += (double)S.FluidN[(size_t)fs * G + (c / ncy) * ngy + c % ncy]; ++m; }
            const float mean = m ? (float)(a / m) : 0.f;
            for (int r = 0; r < nrec; ++r) w.f32(mean);
        }
        w.zeros(nrec);   // field iterations are not kept per record
        const double dt = (double)S.dt;
        for (int i = 0; i < nc; ++i) {
            // current into conductor i: displacement (backward difference of
            // its charge) minus conduction (absorbed charge per unit time),
            // src/cuda_Diagnostic.cu:181-190
            for (int r = 0; r < nrec; ++r) {
                const double h = spacing(r) * dt;
                double cond = 0;
                for (int s = 0; s < nsp; ++s) cond += (double)S.q_density[s] * rec(r, o_ab + s * nc + i) / h;
                double dis = 0;
                if (r >= 2 && spacing(r) == spacing(r - 1))
                    dis = 0.5 * (3 * rec(r, o_q + i) - 4 * rec(r - 1, o_q + i) + rec(r - 2, o_q + i)) / h;
                else if (r >= 1)
                    dis = (rec(r, o_q + i) - rec(r - 1, o_q + i)) / h;
                w.f32((float)(dis - cond));
            }
            for (int s = 0; s < nsp; ++s)
                for (int r = 0; r < nrec; ++r)
                    w.f32((float)((double)S.q_density[s] * rec(r, o_ab + s * nc + i) / (spacing(r) * dt)));
            for (int r = 0; r < nrec; ++r) w.f32((float)rec(r, o_v + i));
            for (int r = 0; r < nrec; ++r) w.f32((float)rec(r, o_v + i));
            for (int r = 0; r < nrec; ++r) w.f32((float)rec(r, o_q + i));
        }
        w.i32(0);   // steady-state history: none
        // per species and node: density, its running average, surface charge
        for (int s = 0; s < nsp; ++s)
            for (int g = 0; g < G; ++g) {
                const size_t a = (size_t)s * G + g;
                w.f32((float)hden[a]);
                w.f32(n_accum ? (float)(hsd[a] / (double)n_accum) : 0.f);
                w.f32((float)hsig[a]);
            }
        // per node: background gas, Laplace and Poisson potentials, E
        for (int g = 0; g < G; ++g) {
            w.f32((float)S.Temp[g]);
            w.f32((float)S.BackVel1[g]);
            w.f32((float)S.BackDen1[g]);
            w.f32(S.BackVel2.empty() ? 0.f : (float)S.BackVel2[g]);
            w.f32(S.BackDen2.empty() ? 0.f : (float)S.BackDen2[g]);
            w.f32((float)(htot[g] - hphi[g]));
            w.f32((float)hphi[g]);
            w.f32((float)hex[g]);
            w.f32((float)hey[g]);
        }
        // potential and the averages (source and surface charge: current values)
        std::vector<float> fl(G);
        auto put_r = [&](const std::vector<R> &v) { for (int g = 0; g < G; ++g) fl[g] = (float)v[g]; w.put(fl.data(), G); };
        auto put_avg = [&](const std::vector<double> &v) {
            for (int g = 0; g < G; ++g) fl[g] = n_accum ? (float)(v[g] / (double)n_accum) : 0.f;
            w.put(fl.data(), G);
        };
        put_r(htot);
        put_avg(hsp);
        put_r(hsrc);
        put_r(hSig);
        put_avg(hsx);
        put_avg(hsy);
        // collision rates per node and reaction (src/cuda_Diagnostic.cu:548-560)
        const double steps = (double)std::max(1LL, tstep - tstep_base);
        const double scale = (double)S.Denscale[0] / steps / dt;
        std::vector<float> mr(hmccr.size());
        for (size_t k = 0; k < hmccr.size(); ++k) mr[k] = (float)(scale * (double)hmccr[k]);
        w.put(mr.data(), mr.size());
        const bool ok = w.ok;
        if (fclose(f) != 0 || !ok) return fail(-60, "dump: writing '%s' failed", path);
        return 0;
    }
    struct DumpIn {
        FILE *f;
        long size = 0;   // file length: a skip may not pass it
        bool ok = true;
        template <typename T>
        void get(T *p, size_t n) {
            if (ok && n && fread(p, sizeof(T), n, f) != n) ok = false;
        }
        int i32() { int v = 0; get(&v, 1); return v; }
        void skip(size_t nfloats) {
            if (!ok || !nfloats) return;
            const long at = ftell(f);
            if (at < 0 || (size_t)(size - at) < nfloats * 4 || fseek(f, (long)(nfloats * 4), SEEK_CUR) != 0) ok = false;
        }
    };
    int dump_load(const char *path) override {
        FILE *f = fopen(path, "rb");
        if (!f) return fail(-60, "dump: cannot open '%s'", path);
        std::unique_ptr<FILE, int (*)(FILE *)> guard(f, fclose);
        DumpIn r{f};
        if (fseek(f, 0, SEEK_END) != 0 || (r.size = ftell(f)) < 0 || fseek(f, 0, SEEK_SET) != 0)
            return fail(-60, "dump: cannot read '%s'", path);
        const int ncx = S.ncx, ncy = S.ncy, ngx = S.ngx, ngy = S.ngy, nc = S.CondNUMR;
        auto early = [&](const char *where) { return fail(-61, "dump '%s' ends early (%s)", path, where); };
        const int k2 = r.i32(), k1 = r.i32(), k0 = r.i32();
        if (!r.ok) return early("version");
        if (k2 < 0 || k1 < 0 || k0 < 0) return fail(-62, "dump version %d.%d.%d is not readable", k2, k1, k0);
        double tt = 0;
        r.get(&tt, 1);
        const int ts = r.i32();
        (void)r.i32();   // completed RF cycles, derived from the step
        const int g_dump = r.i32();
        if (!r.ok) return early("time");
        if (g_dump != gas) return fail(-63, "gas type of the dump (%d) differs from the deck's (%d)", g_dump, gas);
        if (ts < 0) return fail(-62, "dump step %d is negative", ts);
        std::vector<std::vector<uint32_t>> recs(nsp);
        for (int s = 0; s < nsp; ++s) {
            float np2c = 0;
            r.get(&np2c, 1);
            const int n = r.i32();
            if (!r.ok) return early("species header");
            if (np2c != (float)S.np2c[s])
                return fail(-64, "species %d: np2c of the dump (%g) differs from the deck's (%g)", s, (double)np2c,
                            (double)(float)S.np2c[s]);
            if (n < 0 || n > cur[s].cap)
                return fail(-30, "species %d: dump holds %d particles, capacity %d", s, n, cur[s].cap);
            recs[s].resize((size_t)n * 6);
            r.get(recs[s].data(), recs[s].size());
            if (!r.ok) return early("particles");
        }
        // fluid: restored when the case runs the fluid, else read past
        const int conti = r.i32();
        std::vector<float> fden((size_t)S.nfsp * ncx * ncy), fsrc(fden.size());
        std::vector<float> ffx((size_t)S.nfsp * ngx * ncy), ffy((size_t)S.nfsp * ncx * ngy);
        for (int fs = 0; fs < S.nfsp; ++fs) {
            for (int i = 0; i < ncx; ++i) {
                r.get(fden.data() + ((size_t)fs * ncx + i) * ncy, ncy);
                r.get(fsrc.data() + ((size_t)fs * ncx + i) * ncy, ncy);
            }
            r.get(ffx.data() + (size_t)fs * ngx * ncy, (size_t)ngx * ncy);
            r.get(ffy.data() + (size_t)fs * ncx * ngy, (size_t)ncx * ngy);
        }
        const int hc = r.i32();
        (void)r.i32();
        if (!r.ok) return early("fluid");
        if (hc < 0) return fail(-62, "dump history length %d is negative", hc);
        r.skip((size_t)hc * (2 + nsp + S.nfsp + (size_t)nc * (4 + nsp)));
        const int hac = r.i32();
        if (!r.ok) return early("history");
        if (hac < 0) return fail(-62, "dump steady-state history length %d is negative", hac);
        r.skip((size_t)hac * (1 + 2 * nsp));
        std::vector<float> g3((size_t)nsp * G * 3), g9((size_t)G * 9), fpot(G);
        r.get(g3.data(), g3.size());
        r.get(g9.data(), g9.size());
        r.get(fpot.data(), G);
        r.skip((size_t)5 * G + (size_t)S.TnRct * G);
        if (!r.ok) return early("node arrays");

        // particles: bucketed by the cell of their absolute position, in file
        // order within a cell (src/cuda_Particle.cu:27-58); the in-cell offset
        // is the position minus the cell's corner, which float subtraction
        // gives exactly.  A position on the far wall (x == ncx or y == ncy,
        // reached when a rounded sum lands on it) is kept in the last cell at
        // the largest offset below 1.
        const R below_one = (R)1 - (sizeof(R) == 4 ? (R)FLT_EPSILON / 2 : (R)DBL_EPSILON / 2);
        for (int s = 0; s < nsp; ++s) {
            const size_t n = recs[s].size() / 6;
            std::vector<int> node(n);
            std::vector<R> ox(n), oy(n);
            std::vector<int> cs(G + 1, 0);
            for (size_t k = 0; k < n; ++k) {
                float x, y;
                memcpy(&x, &recs[s][k * 6 + 1], 4);
                memcpy(&y, &recs[s][k * 6 + 2], 4);
                if (!(x >= 0.f && x <= (float)ncx && y >= 0.f && y <= (float)ncy))
                    return fail(-65, "species %d particle %zu lies outside the grid (%g, %g)", s, k, (double)x,
                                (double)y);
                int xi = (int)x, yi = (int)y;
                ox[k] = xi < ncx ? (R)(x - (float)xi) : below_one;
                oy[k] = yi < ncy ? (R)(y - (float)yi) : below_one;
                xi = std::min(xi, ncx - 1);
                yi = std::min(yi, ncy - 1);
                node[k] = xi * ngy + yi;
                ++cs[node[k] + 1];
            }
            for (int c = 0; c < G; ++c) cs[c + 1] += cs[c];
            std::vector<int> fill(cs.begin(), cs.end() - 1);
            std::vector<R> px(n), py(n), pvx(n), pvy(n), pvz(n);
            std::vector<uint32_t> pid(n);
            for (size_t k = 0; k < n; ++k) {
                const int at = fill[node[k]]++;
                float v[3];
                memcpy(v, &recs[s][k * 6 + 3], 12);
                px[at] = ox[k]; py[at] = oy[k];
                pvx[at] = (R)v[0]; pvy[at] = (R)v[1]; pvz[at] = (R)v[2];
            }
            for (size_t k = 0; k < n; ++k) pid[k] = (uint32_t)k;
            if (int e = set_particles(s, (long long)n, px.data(), py.data(), pvx.data(), pvy.data(), pvz.data(),
                                      pid.data(), cs.data(), (uint32_t)n))
                return e;
        }
        // surface charge per species; Poisson potential, E and the total potential
        std::vector<R> hs((size_t)nsp * G), hp(G), hx(G), hy(G), ht(G);
        for (size_t a = 0; a < hs.size(); ++a) hs[a] = (R)g3[a * 3 + 2];
        for (int g = 0; g < G; ++g) {
            hp[g] = (R)g9[(size_t)g * 9 + 6];
            hx[g] = (R)g9[(size_t)g * 9 + 7];
            hy[g] = (R)g9[(size_t)g * 9 + 8];
            ht[g] = (R)fpot[g];
        }
        std::vector<double> hx64(S.A_size, 0.0);
        for (int g = 0; g < G; ++g)
            if (S.A_idx[g] >= 0 && S.A_idx[g] < S.A_size) hx64[S.A_idx[g]] = (double)hp[g];
        HIPCHK(hipMemcpy(sigma_sp, hs.data(), hs.size() * sizeof(R), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(phi, hp.data(), G * sizeof(R), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(Ex, hx.data(), G * sizeof(R), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(Ey, hy.data(), G * sizeof(R), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(tot, ht.data(), G * sizeof(R), hipMemcpyHostToDevice));
        if (S.A_size) HIPCHK(hipMemcpy(x64, hx64.data(), hx64.size() * sizeof(double), hipMemcpyHostToDevice));
        if (fl.state) {
            auto up = [&](R *dst, const std::vector<float> &v) {
                std::vector<R> t(v.begin(), v.end());
                return hipMemcpy(dst, t.data(), t.size() * sizeof(R), hipMemcpyHostToDevice);
            };
            HIPCHK(up(fl.den, fden));
            HIPCHK(up(fl.src, fsrc));
            HIPCHK(up(fl.flux_x, ffx));
            HIPCHK(up(fl.flux_y, ffy));
            const FluidState fst{0, conti ? 1 : 0};
            HIPCHK(hipMemcpy(fl.state, &fst, sizeof fst, hipMemcpyHostToDevice));
            HIPCHK(hipMemset(fl.win, 0, (size_t)G * sizeof(R)));
        }
        // counters start again at the dump's step
        HIPCHK(hipMemsetAsync(mccr, 0, (size_t)G * S.TnRct * sizeof(unsigned long long), st));
        if (fl.state) HIPCHK(hipMemsetAsync(fl.mark, 0, (size_t)G * S.TnRct * sizeof(unsigned long long), st));
        HIPCHK(hipMemsetAsync(condcharge, 0, (size_t)nsp * nc * sizeof(int), st));
        HIPCHK(hipMemsetAsync(lost, 0, (size_t)nsp * sizeof(int), st));
        HIPCHK(hipMemsetAsync(sigma_hits, 0, (size_t)nsp * G * sizeof(unsigned long long), st));
        HIPCHK(hipMemsetAsync(sum_den, 0, (size_t)nsp * G * sizeof(double), st));
        HIPCHK(hipMemsetAsync(sum_phi, 0, (size_t)G * sizeof(double), st));
        HIPCHK(hipMemsetAsync(sum_ex, 0, (size_t)G * sizeof(double), st));
        HIPCHK(hipMemsetAsync(sum_ey, 0, (size_t)G * sizeof(double), st));
        if (o2delcnt) HIPCHK(hipMemsetAsync(o2delcnt, 0, (size_t)G * sizeof(unsigned long long), st));
        if (diag_cc_prev) HIPCHK(hipMemsetAsync(diag_cc_prev, 0, (size_t)nsp * nc * sizeof(int), st));
        n_accum = 0;
        diag_n = 0;
        set_step(ts);
        tstep_base = ts;
        // the deposit that starts a run (src/cuda_main.cu:28): densities,
        // source and the dielectric surface charge of the loaded state
        if (int e = deposit()) return e;
        HIPCHK(hipStreamSynchronize(st));
        return 0;
    }
    // ------------------------------------------------------------ metastable fluid
    template <typename T>
    int fl_upload(T *&dst, const std::vector<T> &src, size_t n = 0) {
        const size_t c = n ? n : src.size();
        T *p = nullptr;
        HIPCHK(dalloc(&p, c ? c : 1));
        fl_bufs.push_back(p);
        if (!src.empty()) HIPCHK(hipMemcpy(p, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice));
        else HIPCHK(hipMemset(p, 0, (c ? c : 1) * sizeof(T)));
        dst = p;
        return 0;
    }
    int fluid_alloc() {
        if (fl.state) return 0;
        HIPCHK(hipStreamSynchronize(st));
        FluidDev<R> d{};
        d.ncx = S.ncx; d.ncy = S.ncy; d.ngx = S.ngx; d.ngy = S.ngy; d.G = G; d.nf = S.nfsp;
        d.nx = S.fx_n; d.ny = S.fy_n; d.n_ave = S.deck.n_ave; d.TnRct = S.TnRct; d.gas = gas;
        d.dtc = S.dtc; d.dx = S.dx; d.dy = S.dy; d.dt = S.dt; d.denscale_e = S.Denscale[0];
        for (int k = 0; k < nsp && k < 8; ++k) d.dsc[k] = S.Denscale[k];
        d.RR7 = S.RR[7]; d.RR9 = S.RR[9];
        d.dtm = (double)S.dt_mcc; d.sigv1 = (double)S.sigv[1];
        int *ip = nullptr;
        R *rp = nullptr;
        if (int e = fl_upload(ip, S.fx_x1)) return e; d.x_x1 = ip;
        if (int e = fl_upload(ip, S.fx_x2)) return e; d.x_x2 = ip;
        if (int e = fl_upload(ip, S.fx_y)) return e; d.x_y = ip;
        if (int e = fl_upload(ip, S.fy_x)) return e; d.y_x = ip;
        if (int e = fl_upload(ip, S.fy_y1)) return e; d.y_y1 = ip;
        if (int e = fl_upload(ip, S.fy_y2)) return e; d.y_y2 = ip;
        if (int e = fl_upload(ip, S.PlasmaRegion)) return e; d.plasma = ip;
        if (int e = fl_upload(rp, S.fx_g1)) return e; d.x_g1 = rp;
        if (int e = fl_upload(rp, S.fx_g2)) return e; d.x_g2 = rp;
        if (int e = fl_upload(rp, S.fy_g1)) return e; d.y_g1 = rp;
        if (int e = fl_upload(rp, S.fy_g2)) return e; d.y_g2 = rp;
        if (int e = fl_upload(rp, S.fl_ax)) return e; d.ax = rp;
        if (int e = fl_upload(rp, S.fl_bx)) return e; d.bx = rp;
        if (int e = fl_upload(rp, S.fl_ay)) return e; d.ay = rp;
        if (int e = fl_upload(rp, S.fl_by)) return e; d.by = rp;
        if (int e = fl_upload(rp, S.BackDen1)) return e; d.BackDen1 = rp;
        if (int e = fl_upload(rp, S.BackDen2.empty() ? S.BackDen1 : S.BackDen2)) return e; d.BackDen2 = rp;
        if (int e = fl_upload(rp, std::vector<R>(S.RR, S.RR + 100))) return e; d.RR = rp;
        if (int e = fl_upload(d.den, S.fl_den)) return e;
        if (int e = fl_upload(d.src, std::vector<R>(), S.fl_den.size())) return e;
        if (int e = fl_upload(d.flux_x, S.fl_flux_x)) return e;
        if (int e = fl_upload(d.flux_y, S.fl_flux_y)) return e;
        const size_t scratch = std::max((size_t)S.fx_n * S.nfsp * S.ncx, (size_t)S.fy_n * S.nfsp * S.ncy);
        if (int e = fl_upload(d.xs, std::vector<R>(), scratch)) return e;
        if (int e = fl_upload(d.gs, std::vector<R>(), scratch)) return e;
        if (int e = fl_upload(d.win, std::vector<R>(), (size_t)G)) return e;
        // the node densities: argon's own array (its MCC reads P2); the
        // oxygen collision steps read theirs from the array the fluid updates
        if (gas == 0) {
            if (int e = fl_upload(d.nodeN, std::vector<R>(S.FluidN.begin(), S.FluidN.begin() + G))) return e;
        } else {
            d.nodeN = o2fluid;
        }
        if (int e = fl_upload(d.mark, std::vector<unsigned long long>(), (size_t)G * S.TnRct)) return e;
        // the event counts so far are not part of the first window
        HIPCHK(hipMemcpy(d.mark, mccr, (size_t)G * S.TnRct * sizeof(unsigned long long), hipMemcpyDeviceToDevice));
        if (int e = fl_upload(d.state, std::vector<FluidState>(1, FluidState{0, 0}))) return e;
        fl = d;
        return 0;
    }
    int fluid_enable(int on) override {
        if (on && S.deck.n_ave <= 0) return fail(-75, "the metastable fluid needs AverageIter > 0");
        if (on)
            if (int e = fluid_alloc()) return e;
        if (fluid_on != (on != 0)) {
            // the captured step graphs hold the old launch sequence
            HIPCHK(hipStreamSynchronize(st));
            for (int q = 0; q < 2; ++q)
                if (gexec[q]) { (void)hipGraphExecDestroy(gexec[q]); gexec[q] = nullptr; gkey[q] = nullptr; }
        }
        fluid_on = on != 0;
        return 0;
    }
    int diagnostic() override {
        // the continuity solve and the fluid's averaging come between the
        // deposit and the diagnostic (src/cuda_main.cu:84-92)
        if (fluid_on) launch_fluid_step<R>(fl, den, mccr, P1, P2, st);
        launch_accumulate<R>(G, nsp, den, tot, Ex, Ey, sum_den, sum_phi, sum_ex, sum_ey, st);
        ++n_accum;
        HIPCHK(hipGetLastError());
        return 0;
    }
    int particle_stage_fused() {
        // flag bytes reuse the staged path's per-particle move tags; the
        // ionisation stash reuses its product staging buffer (stride = the
        // largest capacity)
        const size_t sstride = (size_t)stage_stride;
        launch_fused_count<R>(kc, cur[0], 0, clk, cnt[0], pcnt, mccr, condcharge, tag[0], stage, sstride, d_err, st);
        mark(3);
        launch_fused_count<R>(kc, cur[1], 1, clk, cnt[1], nullptr, mccr, condcharge, tag[1], stage, sstride, d_err,
                              st);
        mark(4);
        // survivors + ionisation products per destination cell for both
        // species, their three scans (electron and ion cell_start, global
        // product order), then the base of every (cell, slot) entry
        launch_bases_pair(G, inv_start, inv_ent, cnt[0], cnt[1], pcnt, total, alive, total2, alive2, ptotal,
                          nxt[0].cell_start, nxt[1].cell_start, pstart, base[0], base[1], pbase[0], pbase[1], pid_off,
                          scan_tmp, st);
        mark(5);
        launch_fused_scatter<R>(kc, cur[0], nxt[0], nxt[1], 0, tag[0], stage, sstride, base[0], pbase[0], pbase[1],
                                pid_off, next_id, nxt[0].cap, nxt[1].cap, d_err, st);
        mark(6);
        launch_fused_scatter<R>(kc, cur[1], nxt[1], nxt[1], 1, tag[1], stage, sstride, base[1], pbase[1], pbase[1],
                                pid_off, next_id, nxt[1].cap, nxt[1].cap, d_err, st);
        // the id counters advance in the deposit that follows (after both scatters)
        swap_species(0);
        swap_species(1);
        HIPCHK(hipGetLastError());
        return 0;
    }
    // one fused step enqueued on the stream (no host synchronisation); the
    // host buffer parity flips, the host clock does not move
    int enqueue_step() {
        if (int e = field_efield_fused()) return e;   // marks 1 between solve and E field
        mark(2);
        if (int e = particle_stage_fused()) return e;
        mark(7);
        // deposit + profile sums + id counter and clock advance in one node kernel
        DepositExtras<R> ex{1, sum_den, sum_phi, sum_ex, sum_ey, tot, Ex, Ey, next_id, pstart + G, clk,
                            (double)S.dt};
        launch_deposit<R>(kc, nc, cur, part, den, sigma_sp, Sigma, Source, brhs, &ex, st);
        if (fluid_on) launch_fluid_step<R>(fl, den, mccr, P1, P2, st);
        HIPCHK(hipGetLastError());
        mark(8);
        mark(9);
        return 0;
    }
    // the graph of one fused step for the current buffer parity (captured on
    // first use; the capture itself runs nothing)
    int step_graph(hipGraphExec_t &out) {
        const void *key = cur[0].x;
        for (int q = 0; q < 2; ++q)
            if (gexec[q] && gkey[q] == key) { out = gexec[q]; return 0; }
        const int q = gexec[0] ? 1 : 0;
        HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        const int e = enqueue_step();
        hipGraph_t g = nullptr;
        const hipError_t ce = hipStreamEndCapture(st, &g);
        // the capture flipped the host parity without running: flip back
        for (int s = 0; s < nsp; ++s) swap_species(s);
        if (e || ce != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            if (e) return e;
            HIPCHK(ce);
        }
        const hipError_t ie = hipGraphInstantiate(&gexec[q], g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        HIPCHK(ie);
        gkey[q] = key;
        out = gexec[q];
        return 0;
    }
    int step_fused(int n) override {
        if (gas == 1 || gas == 2) {
            // the fused two-pass step is built for the argon chemistry; O2
            // and Ar/O2 cases run the stages one by one
            for (int i = 0; i < n; ++i) {
                int e;
                if ((e = advance_time()) || (e = field_solve()) || (e = efield()) || (e = move_sort()) ||
                    (e = mcc()) || (e = deposit()) || (e = diagnostic()))
                    return e;
            }
            return 0;
        }
        if (int e = ensure_clock()) return e;
        // graphs need a fixed launch sequence: the separable field solve, no timers
        const bool graph = use_graph && spectral && !timing;
        for (int i = 0; i < n; ++i) {
            if (timing) {
                ev_slot = (int)(ev_rec % RING);
                if (int e = flush_slot(ev_slot)) return e;
            }
            host_tick();
            if (graph) {
                hipGraphExec_t ge = nullptr;
                if (int e = step_graph(ge)) return e;
                HIPCHK(hipGraphLaunch(ge, st));
                for (int s = 0; s < nsp; ++s) swap_species(s);
            } else {
                mark(0);
                if (int e = enqueue_step()) return e;
            }
            ++n_accum;
            if (timing) {
                ev_pending[ev_slot] = true;
                ++ev_rec;
            }
        }
        return 0;
    }
    int set_timing(int on) override {
        if (on && !ev[0][0])
            for (int r = 0; r < RING; ++r)
                for (int k = 0; k <= NT; ++k) HIPCHK(hipEventCreate(&ev[r][k]));
        for (int r = 0; r < RING; ++r) ev_pending[r] = false;
        timing = on != 0;
        for (int k = 0; k < NT; ++k) t_ms[k] = 0;
        t_n = 0;
        ev_rec = 0;
        return 0;
    }
    int get_timing(double *ms, int n, long long *steps) override {
        if (timing)
            for (long long q = 0; q < RING; ++q)
                if (int e = flush_slot((int)((ev_rec + q) % RING))) return e;
        for (int k = 0; k < n && k < NT; ++k) ms[k] = t_ms[k];
        if (steps) *steps = t_n;
        return 0;
    }
    int sync() override {
        HIPCHK(hipStreamSynchronize(st));
        int e = 0;
        HIPCHK(hipMemcpyAsync(&e, d_err, sizeof(int), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        if (e & 1) return fail(-30, "particle capacity exceeded (Max_np too small for this case)");
        if (e & 2) return fail(-33, "corrupt particle record (position far outside its cell)");
        PcgResult r{};
        HIPCHK(hipMemcpyAsync(&r, d_res, sizeof r, hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        if (r.status == 1) return fail(-31, "field solve hit the iteration limit (rel. residual %g)", r.rel_res);
        if (r.status == 3) {
            // the abandoned solve has finished: clear its barrier and status for the next solve
            HIPCHK(pcg_work_reset(pw, st));
            const PcgResult z{};
            HIPCHK(hipMemcpyAsync(d_res, &z, sizeof z, hipMemcpyHostToDevice, st));
            HIPCHK(hipStreamSynchronize(st));
            return fail(-34, "field solve abandoned: a grid barrier wait of the persistent solver timed out");
        }
        if (std::isnan(r.rel_res)) return fail(-32, "Field solver Error! (NaN residual)");
        return 0;
    }
    // 1 when all work enqueued on the case stream has finished, 0 while it runs
    int query() override {
        const hipError_t q = hipStreamQuery(st);
        if (q == hipSuccess) return 1;
        if (q == hipErrorNotReady) return 0;
        return fail(-1, "HIP error %d (%s) at stream query", (int)q, hipGetErrorString(q));
    }
    int np_of(int s, long long &n) {
        int v = 0;
        HIPCHK(hipMemcpyAsync(&v, cur[s].cell_start + G, sizeof(int), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        n = std::min<long long>(v, cur[s].cap);   // after an overflow only cap slots exist
        return 0;
    }
    int info(pic_info *o) override {
        HIPCHK(hipStreamSynchronize(st));
        o->ngx = S.ngx; o->ngy = S.ngy; o->nsp = nsp; o->A_size = S.A_size; o->CondNUMR = S.CondNUMR;
        o->TnRct = S.TnRct; o->MCCn = S.MCCn; o->nvel = S.nvel;
        o->dt = (double)S.dt; o->dx = (double)S.dx; o->dy = (double)S.dy; o->t = t; o->tstep = tstep;
        o->precision = precision;
        for (int s = 0; s < 5; ++s) {
            long long n = 0;
            if (s < nsp)
                if (int e = np_of(s, n)) return e;
            o->np[s] = n;
            o->cap[s] = s < nsp ? cur[s].cap : 0;
        }
        return 0;
    }
    int get_particles(int s, void *x, void *y, void *vx, void *vy, void *vz, uint32_t *id, int32_t *cs) override {
        if (s < 0 || s >= nsp) return fail(-3, "bad species %d", s);
        HIPCHK(hipStreamSynchronize(st));
        long long n = 0;
        if (int e = np_of(s, n)) return e;
        const size_t b = (size_t)n * sizeof(R);
        if (x) HIPCHK(hipMemcpyAsync(x, cur[s].x, b, hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        if (y) HIPCHK(hipMemcpyAsync(y, cur[s].y, b, hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        if (vx) HIPCHK(hipMemcpyAsync(vx, cur[s].vx, b, hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        if (vy) HIPCHK(hipMemcpyAsync(vy, cur[s].vy, b, hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        if (vz) HIPCHK(hipMemcpyAsync(vz, cur[s].vz, b, hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        if (id) HIPCHK(hipMemcpyAsync(id, cur[s].id, (size_t)n * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        if (cs) HIPCHK(hipMemcpyAsync(cs, cur[s].cell_start, (G + 1) * sizeof(int), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        return 0;
    }
    int set_particles(int s, long long n, const void *x, const void *y, const void *vx, const void *vy, const void *vz,
                      const uint32_t *id, const int32_t *cs, uint32_t nid) override {
        if (s < 0 || s >= nsp) return fail(-3, "bad species %d", s);
        if (n < 0 || n > cur[s].cap) return fail(-30, "particle count %lld exceeds capacity %d", n, cur[s].cap);
        if (cs[0] != 0 || cs[G] != n) return fail(-3, "cell_start must run from 0 to n");
        for (int c = 0; c < G; ++c)
            if (cs[c + 1] < cs[c]) return fail(-3, "cell_start must be non-decreasing");
        HIPCHK(hipStreamSynchronize(st));
        const size_t b = (size_t)n * sizeof(R);
        HIPCHK(hipMemcpyAsync(cur[s].x, x, b, hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(cur[s].y, y, b, hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(cur[s].vx, vx, b, hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(cur[s].vy, vy, b, hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(cur[s].vz, vz, b, hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(cur[s].id, id, (size_t)n * sizeof(uint32_t), hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(cur[s].cell_start, cs, (G + 1) * sizeof(int), hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(next_id + s, &nid, sizeof nid, hipMemcpyHostToDevice, st));
        HIPCHK(hipStreamSynchronize(st));
        tag_valid = 0;
        return 0;
    }

    // ------------------------------------------------------------ named arrays
    struct Arr {
        void *dev;
        const void *host;
        size_t count;
        int elsize;
        int is_float;
        bool writable;
    };
    bool lookup(const std::string &n, Arr &a) {
        const int rs = (int)sizeof(R);
        auto D = [&](void *p, size_t c, int es, int fl, bool w = true) { a = Arr{p, nullptr, c, es, fl, w}; return true; };
        auto H = [&](const void *p, size_t c, int es, int fl) { a = Arr{nullptr, p, c, es, fl, false}; return true; };
        if (n == "Ex") return D(Ex, G, rs, 1);
        if (n == "Ey") return D(Ey, G, rs, 1);
        if (n == "phi") return D(phi, G, rs, 1);
        if (n == "TotPot") return D(tot, G, rs, 1);
        if (n == "Sigma") return D(Sigma, G, rs, 1);
        if (n == "Source") return D(Source, G, rs, 1);
        if (n == "den") return D(den, (size_t)nsp * G, rs, 1);
        if (n == "sigma_sp") return D(sigma_sp, (size_t)nsp * G, rs, 1);
        if (n == "b") return D(brhs, S.A_size, rs, 1);
        if (n == "xsol") return D(x64, S.A_size, 8, 1);
        if (n == "V_cond") return D(V, S.CondNUMR, rs, 1);
        if (n == "mccr") return D(mccr, (size_t)G * S.TnRct, 8, 0);
        if (n == "CondCharge") return D(condcharge, (size_t)nsp * S.CondNUMR, 4, 0);
        if (n == "Lost") return D(lost, nsp, 4, 0);
        if (n == "sigma_hits") return D(sigma_hits, (size_t)nsp * G, 8, 0);
        if (n == "sum_den") return D(sum_den, (size_t)nsp * G, 8, 1);
        if (n == "sum_phi") return D(sum_phi, G, 8, 1);
        if (n == "sum_ex") return D(sum_ex, G, 8, 1);
        if (n == "sum_ey") return D(sum_ey, G, 8, 1);
        if (n == "next_id") return D(next_id, nsp, 4, 0);
        if (fl.state) {
            const size_t NC = (size_t)S.nfsp * S.ncx * S.ncy;
            if (n == "fl_den") return D(fl.den, NC, rs, 1);
            if (n == "fl_src") return D(fl.src, NC, rs, 1);
            if (n == "fl_flux_x") return D(fl.flux_x, (size_t)S.nfsp * S.ngx * S.ncy, rs, 1);
            if (n == "fl_flux_y") return D(fl.flux_y, (size_t)S.nfsp * S.ncx * S.ngy, rs, 1);
            if (n == "fl_win") return D(fl.win, G, rs, 1);
            if (n == "FluidN") return D(fl.nodeN, (size_t)(gas == 0 ? 1 : S.nfsp) * G, rs, 1, false);
            if (n == "P2") return D(P2, G, rs, 1, false);
            if (n == "fl_state") return D(fl.state, 2, 4, 0, false);
        }
        if (n == "detach_removed") {
            if (!o2delcnt) return false;
            return D(o2delcnt, G, 8, 0);
        }
        if (n == "eedf") {
            if (!eedf) return false;
            return D(eedf, EEDF_BINS, 8, 0, false);
        }
        if (n == "Boundary") return H(S.Boundary.data(), G, 4, 0);
        if (n == "CondID") return H(S.CondID.data(), G, 4, 0);
        if (n == "Face") return H(S.Face.data(), G, 4, 0);
        if (n == "DensRegion") return H(S.DensRegion.data(), G, 4, 0);
        if (n == "Area") return H(S.Area.data(), G, rs, 1);
        if (n == "Temp") return H(S.Temp.data(), G, rs, 1);
        if (n == "PlasmaRegion") return H(S.PlasmaRegion.data(), S.Csize, 4, 0);
        if (n == "eps_r") return H(S.eps_r.data(), S.Csize, rs, 1);
        if (n == "eps") return H(S.eps.data(), S.Csize, rs, 1);
        if (n == "StructIdx") return H(S.StructIdx.data(), S.StructIdx.size(), 4, 0);
        if (n == "A_idx") return H(S.A_idx.data(), G, 4, 0);
        if (n == "Ai") return H(S.Ai.data(), S.Ai.size(), 4, 0);
        if (n == "Aj") return H(S.Aj.data(), S.Aj.size(), 4, 0);
        if (n == "Aval") return H(S.Aval.data(), S.Aval.size(), rs, 1);
        if (n == "Atemp") return H(S.Atemp.data(), S.Atemp.size(), rs, 1);
        if (n == "Mdiag") return H(S.Mdiag.data(), S.Mdiag.size(), rs, 1);
        if (n == "cond_b") return H(S.cond_b.data(), S.cond_b.size(), rs, 1);
        if (n == "temp_b") return H(S.temp_b.data(), S.temp_b.size(), rs, 1);
        if (n == "lap_phi") return H(S.lap_phi.data(), S.lap_phi.size(), rs, 1);
        if (n == "BackDen1") return H(S.BackDen1.data(), G, rs, 1);
        if (n == "BackVel1") return H(S.BackVel1.data(), G, rs, 1);
        if (n == "BackDen2" && !S.BackDen2.empty()) return H(S.BackDen2.data(), G, rs, 1);
        if (n == "BackVel2" && !S.BackVel2.empty()) return H(S.BackVel2.data(), G, rs, 1);
        if (n == "FluidN") return H(S.FluidN.data(), S.FluidN.size(), rs, 1);
        if (n == "xe") return H(S.xe.data(), S.xe.size(), rs, 1);
        if (n == "xee") return H(S.xee.data(), S.xee.size(), rs, 1);
        if (n == "cx") return H(S.cx.data(), S.cx.size(), rs, 1);
        if (n == "sigv") return H(S.sigv, S.nsigv, rs, 1);
        if (n == "P1") return H(S.P1.data(), G, rs, 1);
        if (n == "P2") return H(S.P2.data(), G, rs, 1);
        if (n == "Pion") return H(S.Pion.data(), G, rs, 1);
        if (n == "vsave") return H(S.vsave.data(), S.vsave.size(), rs, 1);
        return false;
    }
    int array_info(const char *name, size_t *count, int *elsize, int *is_float) override {
        Arr a;
        if (!lookup(name, a)) return fail(-4, "unknown array '%s'", name);
        *count = a.count; *elsize = a.elsize; *is_float = a.is_float;
        return 0;
    }
    int get_array(const char *name, void *host, size_t bytes) override {
        Arr a;
        if (!lookup(name, a)) return fail(-4, "unknown array '%s'", name);
        const size_t need = a.count * a.elsize;
        if (bytes < need) return fail(-5, "buffer too small for '%s' (%zu < %zu)", name, bytes, need);
        if (a.dev) {
            HIPCHK(hipStreamSynchronize(st));
            HIPCHK(hipMemcpyAsync(host, a.dev, need, hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        } else if (need) {
            std::memcpy(host, a.host, need);
        }
        return 0;
    }
    int set_array(const char *name, const void *host, size_t bytes) override {
        Arr a;
        if (!lookup(name, a)) return fail(-4, "unknown array '%s'", name);
        if (!a.writable) return fail(-6, "array '%s' is read-only", name);
        const size_t need = a.count * a.elsize;
        if (bytes != need) return fail(-5, "size mismatch for '%s' (%zu != %zu)", name, bytes, need);
        HIPCHK(hipStreamSynchronize(st));
        HIPCHK(hipMemcpyAsync(a.dev, host, need, hipMemcpyHostToDevice, st));
        HIPCHK(hipStreamSynchronize(st));
        return 0;
    }
    int set_step(long long s) override {
        tstep = s;
        clk_dirty = true;
        t = (double)s * (double)S.dt;
        return 0;
    }
    // launch shape of the persistent PCG (0, 0 when the separable solve is used)
    int solver_shape(int *blocks, int *threads) override {
        if (spectral || !pw) {
            if (blocks) *blocks = 0;
            if (threads) *threads = 0;
            return 0;
        }
        return pcg_launch_shape(pw, blocks, threads);
    }
    int last_solve(int *iters, int *outer, double *rel) override {
        PcgResult r{};
        HIPCHK(hipStreamSynchronize(st));
        HIPCHK(hipMemcpyAsync(&r, d_res, sizeof r, hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        if (iters) *iters = r.iters;
        if (outer) *outer = r.outer;
        if (rel) *rel = r.rel_res;
        return 0;
    }
};

}  // namespace pic

namespace pic {
// Host-only view of a case setup (no device needed): lets the CPU test suite
// check deck parsing, geometry, operator, tables and the initial load.
template <typename R>
static bool setup_lookup(Setup<R> &S, const std::string &n, const void *&p, size_t &count, int &es, int &fl) {
    const int rs = (int)sizeof(R);
    auto H = [&](const void *q, size_t c, int e, int f) { p = q; count = c; es = e; fl = f; return true; };
    const size_t G = (size_t)S.Gsize;
    if (n == "Boundary") return H(S.Boundary.data(), G, 4, 0);
    if (n == "CondID") return H(S.CondID.data(), G, 4, 0);
    if (n == "Face") return H(S.Face.data(), G, 4, 0);
    if (n == "DensRegion") return H(S.DensRegion.data(), G, 4, 0);
    if (n == "Area") return H(S.Area.data(), G, rs, 1);
    if (n == "Temp") return H(S.Temp.data(), G, rs, 1);
    if (n == "PlasmaRegion") return H(S.PlasmaRegion.data(), S.PlasmaRegion.size(), 4, 0);
    if (n == "eps_r") return H(S.eps_r.data(), S.eps_r.size(), rs, 1);
    if (n == "eps") return H(S.eps.data(), S.eps.size(), rs, 1);
    if (n == "StructIdx") return H(S.StructIdx.data(), S.StructIdx.size(), 4, 0);
    if (n == "A_idx") return H(S.A_idx.data(), G, 4, 0);
    if (n == "Ai") return H(S.Ai.data(), S.Ai.size(), 4, 0);
    if (n == "Aj") return H(S.Aj.data(), S.Aj.size(), 4, 0);
    if (n == "Aval") return H(S.Aval.data(), S.Aval.size(), rs, 1);
    if (n == "Atemp") return H(S.Atemp.data(), S.Atemp.size(), rs, 1);
    if (n == "Mdiag") return H(S.Mdiag.data(), S.Mdiag.size(), rs, 1);
    if (n == "cond_b") return H(S.cond_b.data(), S.cond_b.size(), rs, 1);
    if (n == "temp_b") return H(S.temp_b.data(), S.temp_b.size(), rs, 1);
    if (n == "lap_phi") return H(S.lap_phi.data(), S.lap_phi.size(), rs, 1);
    if (n == "BackDen1") return H(S.BackDen1.data(), S.BackDen1.size(), rs, 1);
    if (n == "BackVel1") return H(S.BackVel1.data(), S.BackVel1.size(), rs, 1);
    if (n == "BackDen2") return H(S.BackDen2.data(), S.BackDen2.size(), rs, 1);
    if (n == "BackVel2") return H(S.BackVel2.data(), S.BackVel2.size(), rs, 1);
    if (n == "FluidN") return H(S.FluidN.data(), S.FluidN.size(), rs, 1);
    if (n == "xe") return H(S.xe.data(), S.xe.size(), rs, 1);
    if (n == "xee") return H(S.xee.data(), S.xee.size(), rs, 1);
    if (n == "cx") return H(S.cx.data(), S.cx.size(), rs, 1);
    if (n == "sigv") return H(S.sigv, S.nsigv, rs, 1);
    if (n == "Th_e") return H(S.Th_e, S.TnRct, rs, 1);
    if (n == "Flag") return H(S.Flag, S.TnRct, rs, 1);
    if (n == "mofM") return H(S.mofM, S.TnRct, rs, 1);
    if (n == "RR") return H(S.RR, S.TnRct, rs, 1);
    if (n == "colprob") return H(S.colprob, S.nsp, rs, 1);
    if (n == "P1") return H(S.P1.data(), S.P1.size(), rs, 1);
    if (n == "P2") return H(S.P2.data(), S.P2.size(), rs, 1);
    if (n == "Pion") return H(S.Pion.data(), S.Pion.size(), rs, 1);
    if (n == "vsave") return H(S.vsave.data(), S.vsave.size(), rs, 1);
    if (n == "fl_x1") return H(S.fx_x1.data(), S.fx_x1.size(), 4, 0);
    if (n == "fl_x2") return H(S.fx_x2.data(), S.fx_x2.size(), 4, 0);
    if (n == "fl_xy") return H(S.fx_y.data(), S.fx_y.size(), 4, 0);
    if (n == "fl_yx") return H(S.fy_x.data(), S.fy_x.size(), 4, 0);
    if (n == "fl_y1") return H(S.fy_y1.data(), S.fy_y1.size(), 4, 0);
    if (n == "fl_y2") return H(S.fy_y2.data(), S.fy_y2.size(), 4, 0);
    if (n == "fl_xg1") return H(S.fx_g1.data(), S.fx_g1.size(), rs, 1);
    if (n == "fl_xg2") return H(S.fx_g2.data(), S.fx_g2.size(), rs, 1);
    if (n == "fl_yg1") return H(S.fy_g1.data(), S.fy_g1.size(), rs, 1);
    if (n == "fl_yg2") return H(S.fy_g2.data(), S.fy_g2.size(), rs, 1);
    if (n == "fl_D") return H(S.fl_D.data(), S.fl_D.size(), rs, 1);
    if (n == "fl_den") return H(S.fl_den.data(), S.fl_den.size(), rs, 1);
    if (n == "fl_ax") return H(S.fl_ax.data(), S.fl_ax.size(), rs, 1);
    if (n == "fl_bx") return H(S.fl_bx.data(), S.fl_bx.size(), rs, 1);
    if (n == "fl_ay") return H(S.fl_ay.data(), S.fl_ay.size(), rs, 1);
    if (n == "fl_by") return H(S.fl_by.data(), S.fl_by.size(), rs, 1);
    if (n == "fl_flux_x") return H(S.fl_flux_x.data(), S.fl_flux_x.size(), rs, 1);
    if (n == "fl_flux_y") return H(S.fl_flux_y.data(), S.fl_flux_y.size(), rs, 1);
    if (n.size() > 4 && n.compare(0, 2, "sp") == 0 && n[3] == '.') {
        int s = n[2] - '0';
        if (s < 0 || s >= S.nsp) return false;
        SpeciesInit<R> &sp = S.sp[s];
        std::string f = n.substr(4);
        if (f == "x") return H(sp.x.data(), sp.n, rs, 1);
        if (f == "y") return H(sp.y.data(), sp.n, rs, 1);
        if (f == "vx") return H(sp.vx.data(), sp.n, rs, 1);
        if (f == "vy") return H(sp.vy.data(), sp.n, rs, 1);
        if (f == "vz") return H(sp.vz.data(), sp.n, rs, 1);
        if (f == "id") return H(sp.id.data(), sp.n, 4, 0);
        if (f == "cell_start") return H(sp.cell_start.data(), sp.cell_start.size(), 4, 0);
    }
    return false;
}

template <typename R>
static bool setup_scalar(Setup<R> &S, const std::string &n, double &v) {
    if (n == "dt") { v = (double)S.dt; return true; }
    if (n == "dx") { v = (double)S.dx; return true; }
    if (n == "dy") { v = (double)S.dy; return true; }
    if (n == "A_size") { v = S.A_size; return true; }
    if (n == "CondNUMR") { v = S.CondNUMR; return true; }
    if (n == "MCCn") { v = S.MCCn; return true; }
    if (n == "dt_mcc") { v = (double)S.dt_mcc; return true; }
    if (n == "nvel") { v = S.nvel; return true; }
    if (n == "nsp") { v = S.nsp; return true; }
    if (n == "TnRct") { v = S.TnRct; return true; }
    if (n == "nbg") { v = S.nbg; return true; }
    if (n == "bg_dens0") { v = (double)S.bg_dens[0]; return true; }
    if (n == "bg_dens1") { v = (double)S.bg_dens[1]; return true; }
    if (n == "idLOGX") { v = (double)S.idLOGX; return true; }
    if (n == "idx") { v = (double)S.idx; return true; }
    if (n == "idy") { v = (double)S.idy; return true; }
    if (n == "dtc") { v = (double)S.dtc; return true; }
    if (n == "fl_nx") { v = S.fx_n; return true; }
    if (n == "fl_ny") { v = S.fy_n; return true; }
    if (n == "nfsp") { v = S.nfsp; return true; }
    if (n.compare(0, 4, "np2c") == 0 && n.size() == 5) { v = (double)S.np2c[n[4] - '0']; return true; }
    if (n.compare(0, 2, "np") == 0 && n.size() == 3) { v = S.sp[n[2] - '0'].n; return true; }
    if (n.compare(0, 6, "Ascale") == 0 && n.size() == 7) { v = (double)S.Ascale[n[6] - '0']; return true; }
    if (n.compare(0, 6, "Escale") == 0 && n.size() == 7) { v = (double)S.Escale[n[6] - '0']; return true; }
    if (n.compare(0, 3, "vti") == 0 && n.size() == 4) { v = (double)S.vti[n[3] - '0']; return true; }
    if (n.compare(0, 9, "q_density") == 0 && n.size() == 10) { v = (double)S.q_density[n[9] - '0']; return true; }
    return false;
}

struct HostSetupBase {
    virtual ~HostSetupBase() {}
    virtual bool lookup(const std::string &, const void *&, size_t &, int &, int &) = 0;
    virtual bool scalar(const std::string &, double &) = 0;
};
template <typename R>
struct HostSetup : HostSetupBase {
    Setup<R> S;
    bool lookup(const std::string &n, const void *&p, size_t &c, int &e, int &f) override {
        return setup_lookup<R>(S, n, p, c, e, f);
    }
    bool scalar(const std::string &n, double &v) override { return setup_scalar<R>(S, n, v); }
};

// ------------------------------------------------------------------ batches
// A batch steps several cases of one precision together: every kernel of
// the fused step runs once for all of them, over a grid that concatenates
// the cases' own grids (multi.h).  The launches of one step of each case are
// recorded twice at set-up (once per particle buffer parity), the records of
// the same launch placed side by side in device memory, and each parity's
// merged sequence captured in a graph.  A batch step then is one graph
// launch on the batch stream, ordered after the work already enqueued on the
// cases' streams and before work enqueued on them later.  The cases give
// the same bits as when stepped alone.
struct BatchBase {
    virtual ~BatchBase() {}
    virtual int step(int n) = 0;
    virtual int sync() = 0;
};

template <typename R>
struct Batch : BatchBase {
    struct Merged {
        MultiLaunchFn fn;
        const void *jobs;
        const uint32_t *map;
        int total;
    };
    std::vector<Case<R> *> cs;
    hipStream_t st = nullptr;
    int device = 0;
    unsigned char *dmem = nullptr;
    std::vector<Merged> seq[2];
    hipGraphExec_t gexec[2] = {nullptr, nullptr};
    std::vector<const void *> expect[2];   // each case's electron x buffer at the batch's parity
    int parity = 0;
    std::vector<hipEvent_t> ev;
    hipEvent_t done = nullptr;

    ~Batch() override {
        if (st) (void)hipStreamSynchronize(st);
        for (auto &g : gexec)
            if (g) (void)hipGraphExecDestroy(g);
        for (auto e : ev) (void)hipEventDestroy(e);
        if (done) (void)hipEventDestroy(done);
        if (dmem) (void)hipFree(dmem);
        if (st) (void)hipStreamDestroy(st);
        for (auto *c : cs) --c->batch_refs;
    }

    int build(const std::vector<Case<R> *> &cases) {
        cs = cases;
        for (auto *c : cs) ++c->batch_refs;
        device = cs[0]->device;
        for (size_t i = 0; i < cs.size(); ++i) {
            Case<R> *c = cs[i];
            if (c->device != device) return fail(-70, "batch: case %zu is on another device", i);
            if (c->gas != 0) return fail(-70, "batch: case %zu is not an argon case (its step is not the fused one)", i);
            if (!c->spectral) return fail(-70, "batch: case %zu does not use the separable field solve", i);
            if (c->timing) return fail(-70, "batch: case %zu has stage timers on", i);
            if (c->fluid_on) return fail(-70, "batch: case %zu runs the metastable fluid", i);
            if (c->spec_cycles != cs[0]->spec_cycles)
                return fail(-70, "batch: case %zu's field solve runs %d cycles, case 0's %d", i, c->spec_cycles,
                            cs[0]->spec_cycles);
        }
        HIPCHK(hipSetDevice(device));
        HIPCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        HIPCHK(hipEventCreateWithFlags(&done, hipEventDisableTiming));
        ev.resize(cs.size(), nullptr);
        for (auto &e : ev) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        // record one step of every case at each parity (recording launches
        // nothing; each recorded step flips the case's host parity, so two
        // leave it where it was)
        std::vector<std::vector<RecordedLaunch>> rec[2];
        for (int q = 0; q < 2; ++q) {
            rec[q].resize(cs.size());
            for (size_t i = 0; i < cs.size(); ++i) {
                Case<R> *c = cs[i];
                if (q == 0) {
                    HIPCHK(hipStreamSynchronize(c->st));
                    if (int e = c->ensure_clock()) return e;
                    HIPCHK(hipStreamSynchronize(c->st));
                }
                expect[q].push_back(c->cur[0].x);
                Recorder r;
                launch_recorder() = &r;
                const int e = c->enqueue_step();
                launch_recorder() = nullptr;
                if (e) return e;
                rec[q][i] = std::move(r.launches);
            }
        }
        // lay out the job records and block maps of both parities
        struct Place { size_t jobs, map; int total; };
        std::vector<Place> places[2];
        size_t bytes = 0;
        for (int q = 0; q < 2; ++q) {
            const auto &r0 = rec[q][0];
            for (size_t i = 1; i < cs.size(); ++i)
                if (rec[q][i].size() != r0.size()) return fail(-70, "batch: case %zu's step has another launch sequence", i);
            for (size_t l = 0; l < r0.size(); ++l) {
                Place pl{};
                const size_t stride = r0[l].job_size;   // the kernel indexes the records as an array
                bytes = (bytes + 255) / 256 * 256;
                pl.jobs = bytes;
                bytes += stride * cs.size();
                bytes = (bytes + 255) / 256 * 256;
                pl.map = bytes;
                int total = 0;
                for (size_t i = 0; i < cs.size(); ++i) {
                    const RecordedLaunch &x = rec[q][i][l];
                    if (x.fn != r0[l].fn || x.job_size != r0[l].job_size)
                        return fail(-70, "batch: case %zu's step has another launch sequence", i);
                    if (x.blocks >= (1 << MULTI_LB_BITS)) return fail(-70, "batch: a launch of case %zu is too large", i);
                    total += x.blocks;
                }
                pl.total = total;
                bytes += (size_t)total * sizeof(uint32_t);
                places[q].push_back(pl);
            }
        }
        if (cs.size() >= (1u << (32 - MULTI_LB_BITS))) return fail(-70, "batch: too many cases");
        std::vector<unsigned char> host(bytes, 0);
        for (int q = 0; q < 2; ++q)
            for (size_t l = 0; l < places[q].size(); ++l) {
                const Place &pl = places[q][l];
                const size_t stride = rec[q][0][l].job_size;
                uint32_t *map = (uint32_t *)(host.data() + pl.map);
                int b = 0;
                for (size_t i = 0; i < cs.size(); ++i) {
                    const RecordedLaunch &x = rec[q][i][l];
                    memcpy(host.data() + pl.jobs + i * stride, x.job.data(), x.job_size);
                    for (int k = 0; k < x.blocks; ++k) map[b++] = ((uint32_t)i << MULTI_LB_BITS) | (uint32_t)k;
                }
            }
        HIPCHK(hipMalloc(&dmem, bytes ? bytes : 1));
        HIPCHK(hipMemcpy(dmem, host.data(), bytes, hipMemcpyHostToDevice));
        for (int q = 0; q < 2; ++q) {
            for (size_t l = 0; l < places[q].size(); ++l) {
                const Place &pl = places[q][l];
                seq[q].push_back(Merged{rec[q][0][l].fn, dmem + pl.jobs, (const uint32_t *)(dmem + pl.map), pl.total});
            }
        }
        // one graph per parity
        for (int q = 0; q < 2; ++q) {
            HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            for (const Merged &m : seq[q]) m.fn(m.jobs, m.map, m.total, st);
            hipGraph_t g = nullptr;
            const hipError_t ce = hipStreamEndCapture(st, &g);
            if (ce != hipSuccess) {
                if (g) (void)hipGraphDestroy(g);
                HIPCHK(ce);
            }
            const hipError_t ie = hipGraphInstantiate(&gexec[q], g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            HIPCHK(ie);
        }
        return 0;
    }

    int step(int n) override {
        if (n <= 0) return 0;
        HIPCHK(hipSetDevice(device));
        for (size_t i = 0; i < cs.size(); ++i) {
            if (cs[i]->cur[0].x != expect[parity][i])
                return fail(-71, "batch: case %zu was stepped outside the batch; make a new batch", i);
            HIPCHK(hipEventRecord(ev[i], cs[i]->st));
            HIPCHK(hipStreamWaitEvent(st, ev[i], 0));
        }
        for (int k = 0; k < n; ++k) {
            HIPCHK(hipGraphLaunch(gexec[parity], st));
            parity ^= 1;
            for (auto *c : cs) {
                for (int s = 0; s < c->nsp; ++s) c->swap_species(s);
                c->host_tick();
                ++c->n_accum;
            }
        }
        HIPCHK(hipEventRecord(done, st));
        for (auto *c : cs) HIPCHK(hipStreamWaitEvent(c->st, done, 0));
        return 0;
    }
    int sync() override {
        HIPCHK(hipStreamSynchronize(st));
        for (auto *c : cs)
            if (int e = c->sync()) return e;
        return 0;
    }
};
}  // namespace pic

// ================================================================== C ABI
using namespace pic;

struct pic_ctx {
    std::unique_ptr<CaseBase> c;
};

extern "C" {

const char *pic_last_error(void) { return g_err.c_str(); }

int pic_create(const char *deck_json, const pic_options *opt, pic_ctx **out) {
    if (!deck_json || !out) return fail(-1, "null argument");
    *out = nullptr;
    pic_options o{};
    o.load = 1;
    if (opt) o = *opt;
    Deck d;
    std::string err;
    if (int e = parse_deck(deck_json, d, err)) return fail(e, "%s", err.c_str());
    std::unique_ptr<pic_ctx> h(new pic_ctx());
    int e = 0;
    try {
        if (o.precision == PIC_FP64_PLAIN) {
            auto *c = new Case<double>();
            h->c.reset(c);
            e = c->init(d, o);
        } else if (o.precision == PIC_FP32_REFINED) {
            auto *c = new Case<float>();
            h->c.reset(c);
            e = c->init(d, o);
        } else {
            return fail(-1, "unknown precision %d", o.precision);
        }
    } catch (const std::bad_alloc &) {
        return fail(-7, "host out of memory during setup");
    }
    if (e) return e;
    *out = h.release();
    return 0;
}

int pic_create_from_file(const char *path, const pic_options *opt, pic_ctx **out) {
    if (!path) return fail(-1, "null path");
    std::ifstream f(path);
    if (!f) return fail(-1, "cannot open deck '%s'", path);
    std::stringstream ss;
    ss << f.rdbuf();
    std::string s = ss.str();
    return pic_create(s.c_str(), opt, out);
}

int pic_destroy(pic_ctx *ctx) {
    if (ctx && ctx->c && ctx->c->batch_refs > 0)
        return fail(-72, "the case belongs to a batch: destroy the batch first");
    delete ctx;
    return 0;
}

struct pic_batch {
    std::unique_ptr<BatchBase> b;
};

int pic_batch_create(pic_ctx *const *cases, int n, pic_batch **out) {
    if (!cases || n <= 0 || !out) return fail(-1, "batch: no cases");
    *out = nullptr;
    std::vector<Case<float> *> f;
    std::vector<Case<double> *> d;
    for (int i = 0; i < n; ++i) {
        if (!cases[i] || !cases[i]->c) return fail(-1, "batch: case %d is null", i);
        for (int j = 0; j < i; ++j)
            if (cases[j] == cases[i]) return fail(-70, "batch: case %d is listed twice", i);
        if (auto *cf = dynamic_cast<Case<float> *>(cases[i]->c.get())) f.push_back(cf);
        else if (auto *cd = dynamic_cast<Case<double> *>(cases[i]->c.get())) d.push_back(cd);
    }
    if (!f.empty() && !d.empty()) return fail(-70, "batch: the cases have different precisions");
    std::unique_ptr<pic_batch> h(new pic_batch());
    int e = 0;
    if (!f.empty()) {
        auto *b = new Batch<float>();
        h->b.reset(b);
        e = b->build(f);
    } else {
        auto *b = new Batch<double>();
        h->b.reset(b);
        e = b->build(d);
    }
    if (e) return e;
    *out = h.release();
    return 0;
}
int pic_batch_step(pic_batch *b, int nsteps) {
    if (!b || !b->b) return fail(-1, "null batch");
    return b->b->step(nsteps);
}
int pic_batch_sync(pic_batch *b) {
    if (!b || !b->b) return fail(-1, "null batch");
    return b->b->sync();
}
int pic_batch_destroy(pic_batch *b) {
    delete b;
    return 0;
}

#define CTX_OR_FAIL(ctx) \
    if (!(ctx) || !(ctx)->c) return fail(-1, "null context")

int pic_info_get(pic_ctx *ctx, pic_info *info) {
    CTX_OR_FAIL(ctx);
    if (!info) return fail(-1, "null info");
    return ctx->c->info(info);
}
int pic_sync(pic_ctx *ctx) { CTX_OR_FAIL(ctx); return ctx->c->sync(); }
int pic_query(pic_ctx *ctx) { CTX_OR_FAIL(ctx); return ctx->c->query(); }
int pic_device_sync(int device) {
    int cur = 0;
    HIPCHK(hipGetDevice(&cur));
    HIPCHK(hipSetDevice(device));
    const hipError_t e = hipDeviceSynchronize();
    const hipError_t l = hipGetLastError();
    HIPCHK(hipSetDevice(cur));
    HIPCHK(e);
    HIPCHK(l);
    return 0;
}
int pic_device_memory(int device, size_t *free_bytes, size_t *total_bytes) {
    if (!free_bytes || !total_bytes) return fail(-2, "pic_device_memory: null output");
    int cur = 0;
    HIPCHK(hipGetDevice(&cur));
    HIPCHK(hipSetDevice(device));
    const hipError_t e = hipMemGetInfo(free_bytes, total_bytes);
    HIPCHK(hipSetDevice(cur));
    HIPCHK(e);
    return 0;
}
int pic_eedf_sample(pic_ctx *ctx, double emax_ev) { CTX_OR_FAIL(ctx); return ctx->c->eedf_sample(emax_ev); }
int pic_fluid_enable(pic_ctx *ctx, int on) { CTX_OR_FAIL(ctx); return ctx->c->fluid_enable(on); }
int pic_dump_save(pic_ctx *ctx, const char *path) {
    CTX_OR_FAIL(ctx);
    if (!path) return fail(-3, "dump: no path");
    return ctx->c->dump_save(path);
}
int pic_dump_load(pic_ctx *ctx, const char *path) {
    CTX_OR_FAIL(ctx);
    if (!path) return fail(-3, "dump: no path");
    return ctx->c->dump_load(path);
}
int pic_diag_sample(pic_ctx *ctx) { CTX_OR_FAIL(ctx); return ctx->c->diag_sample(); }
int pic_diag_history(pic_ctx *ctx, double *out, int max_records, int *nrec, int *rec_len) {
    CTX_OR_FAIL(ctx);
    return ctx->c->diag_history(out, max_records, nrec, rec_len);
}
int pic_advance_time(pic_ctx *ctx) { CTX_OR_FAIL(ctx); return ctx->c->advance_time(); }
int pic_field_solve(pic_ctx *ctx) { CTX_OR_FAIL(ctx); return ctx->c->field_solve(); }
int pic_efield(pic_ctx *ctx) { CTX_OR_FAIL(ctx); return ctx->c->efield(); }
int pic_move(pic_ctx *ctx) { CTX_OR_FAIL(ctx); return ctx->c->move(); }
int pic_sort_boundary(pic_ctx *ctx) { CTX_OR_FAIL(ctx); return ctx->c->sort_boundary(); }
int pic_mcc(pic_ctx *ctx) { CTX_OR_FAIL(ctx); return ctx->c->mcc(); }
int pic_deposit(pic_ctx *ctx) { CTX_OR_FAIL(ctx); return ctx->c->deposit(); }
int pic_diagnostic(pic_ctx *ctx) { CTX_OR_FAIL(ctx); return ctx->c->diagnostic(); }

int pic_run_staged(pic_ctx *ctx, int nsteps) {
    CTX_OR_FAIL(ctx);
    CaseBase *c = ctx->c.get();
    for (int i = 0; i < nsteps; ++i) {
        int e;
        if ((e = c->advance_time()) || (e = c->field_solve()) || (e = c->efield()) || (e = c->move()) ||
            (e = c->sort_boundary()) || (e = c->mcc()) || (e = c->deposit()) || (e = c->diagnostic()))
            return e;
    }
    return 0;
}

int pic_step(pic_ctx *ctx, int nsteps) {
    CTX_OR_FAIL(ctx);
    return ctx->c->step_fused(nsteps);
}

int pic_get_particles(pic_ctx *ctx, int isp, void *x, void *y, void *vx, void *vy, void *vz, uint32_t *id,
                      int32_t *cell_start) {
    CTX_OR_FAIL(ctx);
    return ctx->c->get_particles(isp, x, y, vx, vy, vz, id, cell_start);
}
int pic_set_particles(pic_ctx *ctx, int isp, long long n, const void *x, const void *y, const void *vx,
                      const void *vy, const void *vz, const uint32_t *id, const int32_t *cell_start,
                      uint32_t next_id) {
    CTX_OR_FAIL(ctx);
    if (!x || !y || !vx || !vy || !vz || !id || !cell_start) return fail(-1, "null particle buffer");
    return ctx->c->set_particles(isp, n, x, y, vx, vy, vz, id, cell_start, next_id);
}
int pic_array_info(pic_ctx *ctx, const char *name, size_t *count, int *elsize, int *is_float) {
    CTX_OR_FAIL(ctx);
    return ctx->c->array_info(name, count, elsize, is_float);
}
int pic_get_array(pic_ctx *ctx, const char *name, void *host, size_t bytes) {
    CTX_OR_FAIL(ctx);
    return ctx->c->get_array(name, host, bytes);
}
int pic_set_array(pic_ctx *ctx, const char *name, const void *host, size_t bytes) {
    CTX_OR_FAIL(ctx);
    return ctx->c->set_array(name, host, bytes);
}
const char *pic_array_names(void) {
    return "Ex Ey phi TotPot Sigma Source den sigma_sp sigma_hits b xsol V_cond mccr CondCharge sum_den sum_phi sum_ex sum_ey "
           "eedf Lost next_id detach_removed Boundary CondID Face DensRegion Area Temp PlasmaRegion eps_r eps StructIdx A_idx Ai "
           "Aj Aval Atemp Mdiag cond_b temp_b lap_phi BackDen1 BackVel1 BackDen2 BackVel2 FluidN xe xee cx sigv P1 P2 Pion "
           "vsave";
}
int pic_set_step(pic_ctx *ctx, long long tstep) { CTX_OR_FAIL(ctx); return ctx->c->set_step(tstep); }
int pic_last_field_solve(pic_ctx *ctx, int *iters, int *outer, double *rel_res) {
    CTX_OR_FAIL(ctx);
    return ctx->c->last_solve(iters, outer, rel_res);
}

int pic_stage_timing(pic_ctx *ctx, int enable) {
    CTX_OR_FAIL(ctx);
    return ctx->c->set_timing(enable);
}
int pic_stage_times(pic_ctx *ctx, double *ms, int n, long long *steps) {
    CTX_OR_FAIL(ctx);
    if (!ms) return fail(-1, "null buffer");
    return ctx->c->get_timing(ms, n, steps);
}

int pic_field_solver_kind(pic_ctx *ctx) {
    CTX_OR_FAIL(ctx);
    return ctx->c->solver_kind;
}
int pic_field_solver_shape(pic_ctx *ctx, int *blocks, int *threads) {
    CTX_OR_FAIL(ctx);
    return ctx->c->solver_shape(blocks, threads);
}

// ------------------------------------------------------------------ host-only setup
struct pic_setup_h {
    std::unique_ptr<HostSetupBase> h;
};

int pic_setup_open(const char *deck_json, int precision, int load, pic_setup_h **out) {
    if (!deck_json || !out) return fail(-1, "null argument");
    *out = nullptr;
    Deck d;
    std::string err;
    if (int e = parse_deck(deck_json, d, err)) return fail(e, "%s", err.c_str());
    std::unique_ptr<pic_setup_h> h(new pic_setup_h());
    int e;
    if (precision == PIC_FP64_PLAIN) {
        auto *s = new HostSetup<double>();
        h->h.reset(s);
        e = build_setup<double>(d, s->S, err, load != 0);
    } else {
        auto *s = new HostSetup<float>();
        h->h.reset(s);
        e = build_setup<float>(d, s->S, err, load != 0);
    }
    if (e) return fail(e, "%s", err.c_str());
    *out = h.release();
    return 0;
}
int pic_setup_array_info(pic_setup_h *h, const char *name, size_t *count, int *elsize, int *is_float) {
    if (!h) return fail(-1, "null handle");
    const void *p;
    if (!h->h->lookup(name, p, *count, *elsize, *is_float)) return fail(-4, "unknown array '%s'", name);
    return 0;
}
int pic_setup_get_array(pic_setup_h *h, const char *name, void *host, size_t bytes) {
    if (!h) return fail(-1, "null handle");
    const void *p;
    size_t c;
    int e, f;
    if (!h->h->lookup(name, p, c, e, f)) return fail(-4, "unknown array '%s'", name);
    if (bytes < c * e) return fail(-5, "buffer too small for '%s'", name);
    if (c) std::memcpy(host, p, c * e);
    return 0;
}
int pic_setup_scalar(pic_setup_h *h, const char *name, double *v) {
    if (!h) return fail(-1, "null handle");
    if (!h->h->scalar(name, *v)) return fail(-4, "unknown scalar '%s'", name);
    return 0;
}
void pic_setup_close(pic_setup_h *h) { delete h; }
const char *pic_reaction_names(int gas) {
    static const std::string names[3] = {reaction_names(0), reaction_names(1), reaction_names(2)};
    if (gas < 0 || gas > 2) return nullptr;
    return names[gas].c_str();
}

// ------------------------------------------------------------------ shims
static pic_ctx *g_default = nullptr;

static void shim_check(int e, const char *where) {
    if (e) {
        fprintf(stderr, "\n%s: %s (status %d)\n", where, pic_last_error(), e);
        exit(1);
    }
}
static CaseBase *shim_case(const char *where) {
    if (!g_default) {
        fprintf(stderr, "\n%s: no case open (call pic_shim_open first)\n", where);
        exit(1);
    }
    return g_default->c.get();
}

int pic_shim_open(const char *deck_path, int precision) {
    pic_shim_close();
    pic_options o{};
    o.precision = precision;
    o.load = 1;
    return pic_create_from_file(deck_path, &o, &g_default);
}
void pic_shim_close(void) {
    if (g_default) pic_destroy(g_default);
    g_default = nullptr;
}
pic_ctx *pic_shim_ctx(void) { return g_default; }
void info_Device(void) {
    int dev = 0;
    hipDeviceProp_t p;
    (void)hipGetDevice(&dev);
    if (hipGetDeviceProperties(&p, dev) == hipSuccess)
        printf(" Device %d: %s, %d CUs, %.1f GB\n", dev, p.name, p.multiProcessorCount, p.totalGlobalMem / 1e9);
}
// the setup stages run inside pic_shim_open; these keep the reference's call sequence valid
void start_cuda(void) { shim_case("start_cuda"); }
void Set_Device_Parameter(void) { shim_case("Set_Device_Parameter"); }
void Set_Particle_cuda(void) { shim_case("Set_Particle_cuda"); }
void Set_NullCollisionTime_cuda(void) { shim_case("Set_NullCollisionTime_cuda"); }
void Set_SortBoundary_cuda(void) { shim_case("Set_SortBoundary_cuda"); }
void Set_MatrixPCG_cuda(void) { shim_case("Set_MatrixPCG_cuda"); }
void PCG_SOLVER_Laplace(void) { shim_case("PCG_SOLVER_Laplace"); }
void PCG_SOLVER(void) { shim_check(shim_case("PCG_SOLVER")->field_solve(), "PCG_SOLVER"); }
void Efield_cuda_Basic(void) { shim_check(shim_case("Efield_cuda_Basic")->efield(), "Efield_cuda_Basic"); }
void Move_cuda(void) { shim_check(shim_case("Move_cuda")->move(), "Move_cuda"); }
void SortBounndary_cuda(void) { shim_check(shim_case("SortBounndary_cuda")->sort_boundary(), "SortBounndary_cuda"); }
void MCC_Ar_cuda(void) {
    CaseBase *c = shim_case("MCC_Ar_cuda");
    if (c->gas != 0) shim_check(fail(-19, "the case is not an argon case"), "MCC_Ar_cuda");
    shim_check(c->mcc(), "MCC_Ar_cuda");
}
void MCC_O2_cuda(void) {
    CaseBase *c = shim_case("MCC_O2_cuda");
    if (c->gas != 1) shim_check(fail(-19, "the case is not an O2 case"), "MCC_O2_cuda");
    shim_check(c->mcc(), "MCC_O2_cuda");
}
void MCC_ArO2_cuda(void) {
    CaseBase *c = shim_case("MCC_ArO2_cuda");
    if (c->gas != 2) shim_check(fail(-19, "the case is not an Ar/O2 case"), "MCC_ArO2_cuda");
    shim_check(c->mcc(), "MCC_ArO2_cuda");
}
void Deposit_cuda(void) { shim_check(shim_case("Deposit_cuda")->deposit(), "Deposit_cuda"); }
void Deposit_Basic(void) { shim_check(shim_case("Deposit_Basic")->deposit(), "Deposit_Basic"); }
void Diagnostic(void) { shim_check(shim_case("Diagnostic")->diagnostic(), "Diagnostic"); }
void main_cuda(void) {
    CaseBase *c = shim_case("main_cuda");
    long long nsteps = (long long)c->rf_cycles * c->dt_pic;
    printf("-------------GPU START-------------\n");
    for (long long i = 0; i < nsteps; ++i) {
        shim_check(c->step_fused(1), "main_cuda");
        if ((i + 1) % c->dt_pic == 0) shim_check(c->sync(), "main_cuda");
    }
    shim_check(c->sync(), "main_cuda");
}

}  // extern "C"
