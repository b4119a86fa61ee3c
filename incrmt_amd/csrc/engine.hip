// Host engine of the MI355X IncRMT hot path and its C ABI (include/incrmt.h).
//
// FluidEngine keeps the whole fluid_2d state resident on one GPU in the
// structure-of-arrays layout of common.hh and runs fluid_2d::step_forward
// (fluid_2d.cc:298-369) as a sequence of kernels on one HIP stream: the
// stencil sweeps of step.hip, the multigrid solves of mg.hip and the level
// set of levelset.hip. Scalars the reference computes on the host once per
// step (the anchor rotation of pre_stress_setup, the time) are computed here
// with the same host arithmetic and passed by value.
#include <cmath>
#include <cstring>
#include <cstdio>
#include <string>
#include <vector>
#include <limits>
#include <atomic>
#include <chrono>
#include "step.hh"
#include "mg.hh"
#include "levelset.hh"
#include "tracers.hh"
#include "unet.hh"
#include "../../include/incrmt.h"

namespace incrmt {

static thread_local std::string g_last_error;

int hip_fail(hipError_t e, const char *what, const char *file, int line) {
    char buf[512];
    snprintf(buf, sizeof buf, "%s failed: %s (%s:%d)", what, hipGetErrorString(e), file, line);
    g_last_error = buf;
    return INCRMT_ERR_HIP;
}
static int fail(int code, const char *msg) {
    g_last_error = msg;
    return code;
}
int set_error(int code, const char *msg) { return fail(code, msg); }

namespace {

constexpr int BXe = 64, BYe = 4;
inline int round_up(int a, int b) { return (a + b - 1) / b * b; }

// Trapezoid-weighted mean of the nodal pressure (fluid_2d.cc:1332-1348), as
// a deterministic two-pass reduction: weights 1/2 on the boundary rows and
// columns of non-periodic directions, the sum divided by m*n.
__global__ void k_pavg_part(int mf, int nf, int px, int py, const double *z, int zp, double *part) {
    __shared__ double s[256];
    double a = 0;
    const int N = mf * nf;
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < N; p += gridDim.x * blockDim.x) {
        const int i = p % mf, j = p / mf;
        double w = (!px && (i == 0 || i == mf - 1)) ? 0.5 : 1.0;
        if (!py && (j == 0 || j == nf - 1)) w *= 0.5;
        a += w * z[j * zp + i];
    }
    s[threadIdx.x] = a;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}
__global__ void k_pavg_final(const double *part, int nb, double scale, double *out) {
    __shared__ double s[256];
    double a = 0;
    for (int p = threadIdx.x; p < nb; p += blockDim.x) a += part[p];
    s[threadIdx.x] = a;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = s[0] * scale;
}

// The level-set channel of the learned projection guess on a solver's
// level-0 grid: the smallest solid phi at each MAC cell centre or, for the
// FEM, at each node (the mean of its four cells), over six cell widths,
// clamped to [-1, 1]; 1 without solids.
__global__ void k_guess_feature(Grid g, const ObjArrays *oa, int nobj, int fem, int nx, int ny, int zp,
                                double *feat) {
    const int i = blockIdx.x * BXe + threadIdx.x, j = blockIdx.y * BYe + threadIdx.y;
    if (i >= nx || j >= ny) return;
    double v = 1e300;
    for (int k = 0; k < nobj; k++) {
        const double *ph = oa[k].phi;
        const double p = fem ? 0.25 * (ph[g.at(i - 1, j - 1)] + ph[g.at(i, j - 1)] + ph[g.at(i - 1, j)] + ph[g.at(i, j)])
                             : ph[g.at(i, j)];
        v = fmin(v, p);
    }
    const double h = fmin(g.dx, g.dy);
    feat[(size_t)j * zp + i] = nobj ? fmax(-1.0, fmin(1.0, v / (6 * h))) : 1.0;
}

}  // namespace

struct Solid {
    ObjConst c;
    ObjArrays a;
    LsWork w;
    double r_scale = 0;  // obj_simple_spin's r_scale (its tracer starts at 0.9 r_scale)
    // steps left whose sweeps cover the whole grid: after an upload of the
    // solid's fields, one full step restores X = nX outside the band,
    // which windowed sweeps rely on
    int full_steps = 0;
};

struct FluidEngine {
    Grid g{};
    double visc = 0, rhof = 1, tw = 2.5, pinw = 2.5, eps = 0, w_scale = 3, w_force = 20;
    double dt_reg = 0, time = 0;
    bool const_rho = true, initialized = false;
    std::vector<Solid> obj;
    ObjConst *d_oc = nullptr;
    ObjArrays *d_oa = nullptr;
    SolidWin *d_win = nullptr;  // per-solid sweep windows (device)
    const int **d_bbox = nullptr;  // the solids' band boxes (device array of their LsWork::bbox)
    bool windows = true;        // windowed per-solid sweeps and batched band passes (else
                                // whole-grid sweeps and one solid's passes at a time)
    LsBatch lsb{};              // batched band passes of the solids
    // the frame loop's diagnostics (fluid_2d::solve, fluid_2d.cc:259-296):
    // frames written so far, and per projection the wall seconds, V-cycles
    // and solves since the last frame line
    int f_num = 0;
    double wc[2] = {0, 0};
    long vcyc[2] = {0, 0}, nsolve[2] = {0, 0};
    Fields f{};
    Phys ph{};
    std::vector<void *> allocs;
    double *d_xrow = nullptr, *d_pavg = nullptr, *d_part = nullptr;
    Multigrid mac, fem;
    hipStream_t s = nullptr;
    // The end-of-step band rebuild of a lone solid runs on a second stream
    // beside the FEM projection and the velocity update (fork after the FEM
    // set-up, join before the window and density updates); overlap = false
    // keeps it on s, after the velocity update.
    bool overlap = true;
    hipStream_t s2 = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    double mg_tol = 1e-8;
    int max_cycles = 100, precision = INCRMT_PREC_FP64;
    incrmt_stats st{};
    int ntrace_req = 0;   // fluid tracers asked for before initialize
    // learned projection guess: one network per projection (MAC cells, FEM
    // nodes) and the level-set channel on each solver's grid
    int guess_mode = INCRMT_GUESS_PREVIOUS;
    bool guess_train = false;
    float guess_lr = 1e-3f;
    double guess_tol = 1e-4;
    UNet *gnet[2] = {nullptr, nullptr};
    double *gfeat[2] = {nullptr, nullptr};
    int ensure_guess();
    int guess_features(Multigrid &mg, int which);
    // HIP events around the two projections of every step (their learned
    // starts and solves), summed into proj_ms when the stats are read
    bool time_proj = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pev;
    size_t pev_used = 0;
    double proj_ms = 0;
    int proj_begin() {
        if (!time_proj) return 0;
        if (pev_used == pev.size()) {
            hipEvent_t a, b;
            HIP_CHECK(hipEventCreate(&a));
            HIP_CHECK(hipEventCreate(&b));
            pev.push_back({a, b});
        }
        HIP_CHECK(hipEventRecord(pev[pev_used].first, s));
        return 0;
    }
    int proj_end() {
        if (!time_proj) return 0;
        HIP_CHECK(hipEventRecord(pev[pev_used++].second, s));
        return 0;
    }
    int proj_collect() {
        for (size_t k = 0; k < pev_used; k++) {
            HIP_CHECK(hipEventSynchronize(pev[k].second));
            float ms = 0;
            HIP_CHECK(hipEventElapsedTime(&ms, pev[k].first, pev[k].second));
            proj_ms += ms;
        }
        pev_used = 0;
        return 0;
    }
    int launch_req = -1;  // V-cycle issue mode asked for (-1: the solvers' default by size)
    TracerSet tr;

    ~FluidEngine() {
        for (auto &e : pev) { (void)hipEventDestroy(e.first); (void)hipEventDestroy(e.second); }
        for (UNet *n : gnet) delete n;
        for (auto &o : obj) ls_free(o.w);
        ls_batch_free(lsb);
        for (void *p : allocs) (void)hipFree(p);
        if (ev_fork) (void)hipEventDestroy(ev_fork);
        if (ev_join) (void)hipEventDestroy(ev_join);
        if (s2) (void)hipStreamDestroy(s2);
        if (s) (void)hipStreamDestroy(s);
    }
    int dalloc(double *&p) {
        void *q;
        HIP_CHECK(hipMalloc(&q, (size_t)g.N * sizeof(double)));
        HIP_CHECK(memset_sync(q, 0, (size_t)g.N * sizeof(double), s));
        allocs.push_back(q);
        p = (double *)q;
        return 0;
    }
    Solids solids() const { return Solids{d_oc, d_oa, (int)obj.size(), d_win}; }
    // solid k's sweep window after a rebuild of its band (full: the whole grid)
    std::vector<LsJob> jobs() {
        std::vector<LsJob> j;
        for (auto &o : obj) j.push_back(LsJob{&o.c, &o.a, &o.w});
        return j;
    }
    int update_window(size_t k, bool full = false) {
        return launch_solid_window(s, g, obj[k].w.bbox, obj[k].a.st, d_win + k,
                                   full || !windows || obj[k].full_steps > 0);
    }
    // every solid's window in one launch
    int update_windows() {
        unsigned long long mask = 0;
        for (size_t k = 0; k < obj.size(); k++)
            if (!windows || obj[k].full_steps > 0) mask |= 1ull << k;
        return launch_solid_windows(s, g, d_bbox, d_oa, d_win, (int)obj.size(), mask);
    }

    // fluid_2d::choose_ev_and_dt + obj_field::set_ex_visc (fluid_2d.cc:152-192, obj_field.cc:35-49)
    void choose_ev_and_dt(double ev_mult, double dt_pad, double dt_ev_pad) {
        double hmax = g.dx > g.dy ? g.dx : g.dy, hmin = g.dx > g.dy ? g.dy : g.dx, sws_max = 0, ex_visc_max = 0,
               rhos_div_ev_min = std::numeric_limits<double>::max();
        for (auto &o : obj) {
            double tfac = ev_mult * hmax, sws = std::sqrt(o.c.G / o.c.rhos);
            if (sws > sws_max) sws_max = sws;
            o.c.ex_visc = tfac * o.c.rhos * sws;
            if (o.c.ex_visc > ex_visc_max) ex_visc_max = o.c.ex_visc;
            double r_d_ev = o.c.rhos / (o.c.ex_visc * (o.c.ev_trans_mult + 1));
            if (r_d_ev < rhos_div_ev_min) rhos_div_ev_min = r_d_ev;
        }
        double dt0 = hmin / sws_max * dt_pad, dt1 = 0.5 * rhos_div_ev_min / (g.xxsp + g.yysp) * dt_ev_pad,
               dt2 = 0.5 * rhof / (visc * (g.xxsp + g.yysp)) * dt_pad;
        if (dt0 < dt1) dt_reg = dt0 < dt2 ? dt0 : dt2;
        else dt_reg = dt1 < dt2 ? dt1 : dt2;
    }

    // object::pre_stress_setup at the current time, host arithmetic:
    // obj_simple_spin (object.cc:175-181), obj_flapper (object.cc:291-296),
    // obj_lamprey (object.cc:563-568)
    void pre_stress_setup() {
        for (size_t k = 0; k < obj.size(); k++) {
            const ObjConst &c = obj[k].c;
            if (c.kind == kSpin) {
                double ang = time * c.omega, theta = ang > 2 * M_PI ? 0 : c.th_max * (1 - std::cos(ang));
                ph.pss[2 * k] = std::cos(theta);
                ph.pss[2 * k + 1] = std::sin(theta);
            } else if (c.kind == kFlapper) {
                double a = std::sin(c.omega * time);
                a *= a * a;
                a *= a;
                a *= a;
                ph.pss[2 * k] = -c.amp * a;
                ph.pss[2 * k + 1] = 0;
            } else if (c.kind == kLamprey) {
                double sh = time / c.T;
                const double o = 2 * sh;
                ph.pss[2 * k] = o < 1 ? c.A * o * o * (3 - 2 * o) : c.A;
                sh *= c.la;
                ph.pss[2 * k + 1] = sh;
            }
        }
    }
    int compute_stress() {
        pre_stress_setup();
        return launch_stress(s, g, f, solids(), ph);
    }
    int set_rho() { return launch_set_rho(s, g, f, solids(), rhof); }

    int initialize(double ev_mult, double dt_pad, double dt_ev_pad);
    int init_tracers();
    int step(double dt);
    int pin_all(int *passes);
    int pin_join(LsRebuild &r, LsBatchPending *bp, int *passes);
};

// fluid_2d::init_tracers (fluid_2d.cc:1376-1408) with the reference's
// unseeded rand() sequence, on the host as the reference does it, against
// the level sets just built; and the obj_track of every obj_simple_spin
// (object.hh:456-461: one tracer at (0.9 r_scale, 0), X = x, Y = y,
// recorded every 10 steps).
int FluidEngine::init_tracers() {
    tr.ntrace = ntrace_req;
    GlibcRand rng;
    rng.seed(1);
    const int nt = ntrace_req;
    std::vector<double> tm(4 * (size_t)nt, 0.0);
    if (nt > 0) {
        std::vector<std::vector<double>> phi(obj.size(), std::vector<double>(g.N));
        for (size_t k = 0; k < obj.size(); k++)
            HIP_CHECK(hipMemcpy(phi[k].data(), obj[k].a.phi, g.N * sizeof(double), hipMemcpyDeviceToHost));
        auto inside = [&](double tx, double ty) {
            double x = (tx - g.ax) * g.xsp - 0.5, y = (ty - g.ay) * g.ysp - 0.5;
            int i = static_cast<int>(x), j = static_cast<int>(y);
            if (i < -1) i = -1; else if (i > g.m - 1) i = g.m - 1;
            if (j < -1) j = -1; else if (j > g.n - 1) j = g.n - 1;
            x -= i; y -= j;
            const double fx = 1 - x, fy = 1 - y;
            for (size_t k = 0; k < obj.size(); k++) {
                const double *p = phi[k].data();
                const int a = g.at(i, j), b = g.at(i + 1, j), c = g.at(i, j + 1), d = g.at(i + 1, j + 1);
                if (fy * (p[a] * fx + p[b] * x) + y * (p[c] * fx + p[d] * x) < obj[k].c.eps) return true;
            }
            return false;
        };
        // rejection sampling as fluid_2d::init_tracers (fluid_2d.cc:1376-1386),
        // which loops for ever when the solids leave no fluid (obj_full): give
        // up with an error after a bounded number of draws instead
        long long draws = 0;
        const long long max_draws = 1000LL * nt + 10000000LL;
        for (size_t q = 0; q < tm.size();) {
            if (++draws > max_draws) return fail(INCRMT_ERR_ARG, "no fluid left outside the solids for tracers");
            double tx = rng.rnd(g.ax, g.bx);
            double ty = rng.rnd(g.ay, g.by);
            if (!inside(tx, ty)) { tm[q] = tx; tm[q + 1] = ty; q += 4; }
        }
    }
    std::vector<TrackDev> tk;
    for (size_t k = 0; k < obj.size(); k++) {
        if (obj[k].c.kind != kSpin) continue;
        TrackDev t{};
        t.tk[0] = 0.9 * obj[k].r_scale;
        t.tk[1] = 0;
        t.tk[4] = t.tk[0];
        t.tk[5] = t.tk[1];
        t.obj = (int)k;
        t.count = 1; t.every = 10; t.prev_q = 0; t.wind = 0; t.tcorr = 1; t.nrec = 0;
        tk.push_back(t);
    }
    tr.ntrack = (int)tk.size();
    tr.rec_cap = 16384;
    void *q;
    HIP_CHECK(hipMalloc(&q, sizeof(GlibcRand) + 64));
    allocs.push_back(q);
    tr.rng = (GlibcRand *)q;
    tr.nout = (int *)((char *)q + ((sizeof(GlibcRand) + 15) / 16) * 16);
    HIP_CHECK(hipMemcpy(tr.rng, &rng, sizeof(GlibcRand), hipMemcpyHostToDevice));
    HIP_CHECK(memset_sync(tr.nout, 0, sizeof(int), s));
    if (nt > 0) {
        HIP_CHECK(hipMalloc(&q, tm.size() * sizeof(double)));
        allocs.push_back(q);
        tr.tm = (double *)q;
        HIP_CHECK(hipMemcpy(tr.tm, tm.data(), tm.size() * sizeof(double), hipMemcpyHostToDevice));
    }
    if (tr.ntrack > 0) {
        HIP_CHECK(hipMalloc(&q, tk.size() * sizeof(TrackDev)));
        allocs.push_back(q);
        tr.trk = (TrackDev *)q;
        HIP_CHECK(hipMemcpy(tr.trk, tk.data(), tk.size() * sizeof(TrackDev), hipMemcpyHostToDevice));
        HIP_CHECK(hipMalloc(&q, (size_t)tr.ntrack * tr.rec_cap * 4 * sizeof(double)));
        allocs.push_back(q);
        tr.rec = (double *)q;
    }
    return 0;
}

// The two guess networks (seeded, so runs repeat) and their feature
// arrays, once the solvers exist; the solvers' hooks point at them.
int FluidEngine::ensure_guess() {
    Multigrid *mg[2] = {&mac, &fem};
    for (int w = 0; w < 2; w++) {
        MgLevel &f = mg[w]->fine();
        if (!gnet[w]) {
            gnet[w] = new UNet;
            if (int rc = gnet[w]->init(f.nx, f.ny, 20260105u + (unsigned)w, s)) return rc;
            const size_t len = (size_t)(f.ny + 2) * f.pitch + 64;
            void *q;
            HIP_CHECK(hipMalloc(&q, len * sizeof(double)));
            HIP_CHECK(memset_sync(q, 0, len * sizeof(double), s));
            allocs.push_back(q);
            gfeat[w] = (double *)q + f.pitch + 1;
        }
        mg[w]->net = gnet[w];
        mg[w]->feat = gfeat[w];
        mg[w]->train = guess_train;
        mg[w]->lr = guess_lr;
        mg[w]->guess_tol = guess_tol;
    }
    return 0;
}

int FluidEngine::guess_features(Multigrid &mg, int which) {
    if (!mg.net || !(mg.extrapolate == INCRMT_GUESS_LEARNED || mg.train)) return 0;
    const MgLevel &f = mg.fine();
    const dim3 grid((f.nx + BXe - 1) / BXe, (f.ny + BYe - 1) / BYe);
    k_guess_feature<<<grid, dim3(BXe, BYe), 0, s>>>(g, d_oa, (int)obj.size(), which, f.nx, f.ny, f.pitch,
                                                     gfeat[which]);
    HIP_CHECK(hipGetLastError());
    return 0;
}

int FluidEngine::initialize(double ev_mult, double dt_pad, double dt_ev_pad) {
    if ((int)obj.size() > kMaxObj) return fail(INCRMT_ERR_ARG, "too many solids");
    choose_ev_and_dt(ev_mult, dt_pad, dt_ev_pad);
    const_rho = true;
    for (auto &o : obj)
        if (std::fabs(o.c.rhos - rhof) > 100. * std::numeric_limits<double>::epsilon()) const_rho = false;
    for (auto &o : obj) {
        // obj_simple_spin::start_extra (object.cc:152-165)
        if (o.c.kind == kSpin) {
            o.c.piv_thresh = o.c.piv_r + o.c.eps;
            o.c.piv_thresh *= o.c.piv_thresh;
            o.c.K_stiff = 0.1 / (dt_reg * dt_reg);
        }
    }
    // fields
    double **fl[] = {&f.u, &f.v, &f.p, &f.rho, &f.ux, &f.uy, &f.vx, &f.vy, &f.c0, &f.c1, &f.s11, &f.s12, &f.s21, &f.s22,
                     &f.ulc, &f.vdc};
    for (auto p : fl)
        if (int rc = dalloc(*p)) return rc;
    for (int k = 0; k < 8; k++) {
        if (int rc = dalloc(f.eA[k])) return rc;
        if (int rc = dalloc(f.eB[k])) return rc;
    }
    // x of each column as compute_ustar forms it by repeated addition
    std::vector<double> xr(g.m);
    double x = g.ax + 0.5 * g.dx;
    for (int i = 0; i < g.m; i++, x += g.dx) xr[i] = x;
    HIP_CHECK(hipMalloc((void **)&d_xrow, g.m * sizeof(double)));
    allocs.push_back(d_xrow);
    HIP_CHECK(hipMemcpy(d_xrow, xr.data(), g.m * sizeof(double), hipMemcpyHostToDevice));
    ph.xrow = d_xrow;
    HIP_CHECK(hipMalloc((void **)&d_part, 1024 * sizeof(double)));
    allocs.push_back(d_part);
    HIP_CHECK(memset_sync(d_part, 0, 1024 * sizeof(double), s));
    d_pavg = d_part + 1000;
    ph.visc = visc; ph.rhof = rhof; ph.eps = eps;
    ph.w = w_scale * eps; ph.winv = 1 / ph.w; ph.k_rep = w_force * ph.winv;
    // solids: reference map and level set on the ghosted grid (obj_field::init,
    // fluid_2d.cc:200-207), evaluated on the host as the reference does
    std::vector<double> hX(g.N, 0.0), hY(g.N, 0.0), hphi(g.N, 0.0);
    for (auto &o : obj) {
        double **al[] = {&o.a.X, &o.a.Y, &o.a.nX, &o.a.nY, &o.a.Xx, &o.a.Xy, &o.a.Yx, &o.a.Yy, &o.a.Xl, &o.a.Yl,
                         &o.a.Xr, &o.a.Yr, &o.a.Xd, &o.a.Yd, &o.a.Xu, &o.a.Yu, &o.a.phi};
        for (auto p : al)
            if (int rc = dalloc(*p)) return rc;
        void *q;
        HIP_CHECK(hipMalloc(&q, (size_t)g.N * sizeof(int)));
        HIP_CHECK(memset_sync(q, 0, (size_t)g.N * sizeof(int), s));
        allocs.push_back(q);
        o.a.st = (int *)q;
        if (int rc = ls_alloc(g, o.w, s)) return rc;
        for (int j = -2; j < g.n + 2; j++) {
            double y = g.ay + g.dy * (0.5 + j);
            for (int i = -2; i < g.m + 2; i++) {
                const int ij = g.at(i, j);
                double xx = g.ax + g.dx * (0.5 + i), X, Y;
                if (o.c.kind == kFlapper || o.c.kind == kLamprey) {
                    // translation and rotation (object.hh:270-274, 576-580)
                    const double x0 = xx - o.c.cx, y0 = y - o.c.cy;
                    X = x0 * o.c.cth + y0 * o.c.sth;
                    Y = -x0 * o.c.sth + y0 * o.c.cth;
                } else if (o.c.kind != kSpin) {
                    X = xx - o.c.cx; Y = y - o.c.cy;  // obj_standard::transform
                } else {
                    X = xx; Y = y;
                }
                hX[ij] = X; hY[ij] = Y;
                hphi[ij] = obj_phi(o.c, X, Y);
            }
        }
        HIP_CHECK(hipMemcpy(o.a.X, hX.data(), g.N * sizeof(double), hipMemcpyHostToDevice));
        HIP_CHECK(hipMemcpy(o.a.Y, hY.data(), g.N * sizeof(double), hipMemcpyHostToDevice));
        HIP_CHECK(hipMemcpy(o.a.nX, hX.data(), g.N * sizeof(double), hipMemcpyHostToDevice));
        HIP_CHECK(hipMemcpy(o.a.nY, hY.data(), g.N * sizeof(double), hipMemcpyHostToDevice));
        HIP_CHECK(hipMemcpy(o.a.phi, hphi.data(), g.N * sizeof(double), hipMemcpyHostToDevice));
    }
    // (a fluid-only simulation, such as conv_test's types 0 and 1, has none)
    if (!obj.empty()) {
        HIP_CHECK(hipMalloc((void **)&d_oc, obj.size() * sizeof(ObjConst)));
        HIP_CHECK(hipMalloc((void **)&d_oa, obj.size() * sizeof(ObjArrays)));
        HIP_CHECK(hipMalloc((void **)&d_win, obj.size() * sizeof(SolidWin)));
        HIP_CHECK(hipMalloc((void **)&d_bbox, obj.size() * sizeof(int *)));
        allocs.push_back(d_oc);
        allocs.push_back(d_oa);
        allocs.push_back(d_win);
        allocs.push_back((void *)d_bbox);
        std::vector<const int *> hb;
        for (auto &o : obj) hb.push_back(o.w.bbox);
        HIP_CHECK(hipMemcpy((void *)d_bbox, hb.data(), hb.size() * sizeof(int *), hipMemcpyHostToDevice));
        std::vector<ObjConst> hc;
        std::vector<ObjArrays> ha;
        for (auto &o : obj) { hc.push_back(o.c); ha.push_back(o.a); }
        HIP_CHECK(hipMemcpy(d_oc, hc.data(), hc.size() * sizeof(ObjConst), hipMemcpyHostToDevice));
        HIP_CHECK(hipMemcpy(d_oa, ha.data(), ha.size() * sizeof(ObjArrays), hipMemcpyHostToDevice));
    }
    // projection solvers (fluid_2d::setup_multigrid, fluid_2d.cc:126-142)
    // m_fem = m nodes along a periodic direction (fluid_2d.cc:30-31)
    // m_fem = m nodes along a periodic direction (fluid_2d.cc:30-31); with
    // zero pressure on the right edge the solver's unknowns stop before the
    // fixed node column m (mgs_fem.cc:21)
    int rc0 = mac.init(kMAC, g.m, g.n, g.dx, g.dy, s, g.px, g.py, g.zpr);
    if (!rc0)
        rc0 = fem.init(kFEM, (g.px || g.zpr) ? g.m : g.m + 1, g.py ? g.n : g.n + 1, g.dx, g.dy, s, g.px, g.py, g.zpr);
    if (rc0 == INCRMT_ERR_UNSUPPORTED)
        return fail(rc0, "a singly periodic direction or a zero-pressure edge needs an even cell count whose "
                         "halvings end at most 16 points (a doubly periodic box: at least 8 cells)");
    if (rc0) return rc0;
    if (launch_req >= 0) mac.use_graphs = fem.use_graphs = launch_req == INCRMT_LAUNCH_GRAPH;
    if ((guess_mode == INCRMT_GUESS_LEARNED || guess_train) && (rc0 = ensure_guess())) return rc0;
    // fluid_2d::init_fields (fluid_2d.cc:195-254)
    for (size_t k = 0; k < obj.size(); k++) {
        Solid &o = obj[k];
        int passes;
        if (int rc = ls_rebuild(s, g, o.c, o.a, o.w, &passes)) return rc;
        if (int rc = update_window(k)) return rc;
    }
    if (int rc = launch_init_velocity(s, g, f, solids(), rhof)) return rc;
    if (int rc = launch_set_boundaries(s, g, f, solids())) return rc;
    if (int rc = compute_stress()) return rc;
    if (int rc = set_rho()) return rc;
    HIP_CHECK(hipStreamSynchronize(s));
    if (int rc = init_tracers()) return rc;
    initialized = true;
    return 0;
}

int FluidEngine::pin_all(int *passes) {
    if (obj.size() > 1 && windows) {
        if (!lsb.cap)
            if (int rc = ls_batch_alloc((int)obj.size(), lsb)) return rc;
        int p = 0;
        if (int rc = ls_pin_rebuild_batch(s, g, jobs(), lsb, &p)) return rc;
        if (int rc = update_windows()) return rc;
        if (passes) *passes = p;
        if (!const_rho) return set_rho();
        return 0;
    }
    for (size_t k = 0; k < obj.size(); k++) {
        Solid &o = obj[k];
        int p = 0;
        if (int rc = ls_pin_rebuild(s, g, o.c, o.a, o.w, &p)) return rc;
        if (int rc = update_window(k)) return rc;
        if (passes) *passes = p;
    }
    if (!const_rho) return set_rho();
    return 0;
}

// The join of the overlapped end-of-step rebuild (the rest of what pin_all
// does for one solid): s waits for everything enqueued on s2, also after an
// error, then the window and the density follow the new band.
int FluidEngine::pin_join(LsRebuild &r, LsBatchPending *bp, int *passes) {
    int p = 0;
    int rc;
    if (bp) {
        rc = ls_pin_rebuild_batch_end(s2, g, jobs(), lsb, bp, &p);
    } else {
        Solid &o = obj[0];
        rc = ls_pin_rebuild_end(s2, g, o.c, o.a, o.w, r, &p);
    }
    HIP_CHECK(hipEventRecord(ev_join, s2));
    HIP_CHECK(hipStreamWaitEvent(s, ev_join, 0));
    if (rc) return rc;
    if (int e = bp ? update_windows() : update_window(0)) return e;
    if (passes) *passes = p;
    if (!const_rho) return set_rho();
    return 0;
}

int FluidEngine::step(double dt) {
    if (!initialized) return fail(INCRMT_ERR_ARG, "step before initialize");
    const double ntime = time + dt;
    int rc;
    const Solids o = solids();
    if ((rc = tracers_stage1(s, g, f.u, f.v, tr))) return rc;
    if ((rc = launch_normal_derivs(s, g, f, o, dt))) return rc;
    if ((rc = launch_godunov(s, g, f, o, true))) return rc;
    if ((rc = launch_tang_half(s, g, f, o, ph, dt))) return rc;
    if ((rc = launch_godunov(s, g, f, o, false))) return rc;
    time += 0.5 * dt;
    // MAC projection (fluid_2d.cc:1175-1236)
    MgLevel &ml = mac.fine();
    if ((rc = launch_mac_src(s, g, f, ml.b, ml.pitch))) return rc;
    if (!const_rho && (rc = mac.set_density(g, f.rho))) return rc;
    MgStats ms{};
    mac.t_solve = time;  // the MAC solve's time, t + dt/2 (initial_guess(f.time))
    if ((rc = proj_begin())) return rc;
    if ((rc = guess_features(mac, 0))) return rc;
    auto w0 = std::chrono::steady_clock::now();
    rc = precision == INCRMT_PREC_MIXED ? mac.solve_mixed(mg_tol, max_cycles, &ms) : mac.solve(mg_tol, max_cycles, &ms);
    wc[0] += std::chrono::duration<double>(std::chrono::steady_clock::now() - w0).count();
    vcyc[0] += ms.cycles;
    nsolve[0]++;
    if (!rc) rc = proj_end();
    if (rc == 2) return fail(INCRMT_ERR_NOT_CONVERGED, "V-cycle failed to converge in MAC problem");
    if (rc) return rc;
    st.mac_cycles = ms.cycles; st.mac_rel_res = ms.rel_res; st.total_mac_cycles += ms.cycles;
    if ((rc = launch_mac_correct_predict(s, g, f, o, ml.z, ml.pitch, ml.fw, ml.fs, ml.cp, const_rho, dt))) return rc;
    // update_reference_map (fluid_2d.cc:1242-1273)
    if (obj.size() > 1 && windows) {
        int p = 0;
        if (!lsb.cap && (rc = ls_batch_alloc((int)obj.size(), lsb))) return rc;
        if ((rc = ls_extend_batch(s, g, jobs(), lsb, &p)))
            return rc == 5 ? fail(INCRMT_ERR_SINGULAR_FIT, "singular extension fit") : rc;
        st.extend_passes = p;
    } else {
        for (auto &ob : obj) {
            int p = 0;
            if ((rc = ls_extend(s, g, ob.a, ob.w, &p)))
                return rc == 5 ? fail(INCRMT_ERR_SINGULAR_FIT, "singular extension fit") : rc;
            st.extend_passes = p;
        }
    }
    if ((rc = launch_half_timestep(s, g, o))) return rc;
    int passes = 0;
    if ((rc = pin_all(&passes))) return rc == 3 ? fail(INCRMT_ERR_NEWTON, "Newton iteration failed.") : rc;
    st.reinit_passes = passes;
    // u,v are unchanged since the last set_boundaries, so its call here
    // (fluid_2d.cc:340) would rewrite the same ghost values; the reference
    // maps have moved by half a step, and with solid periodic boundaries
    // their ghost values are its first part (set_boundaries_solid)
    if ((rc = launch_set_boundaries_solid(s, g, o))) return rc;
    if ((rc = compute_stress())) return rc;
    if ((rc = launch_ustar(s, g, f, o, ph, dt))) return rc;
    time = ntime;
    // finite-element projection (fluid_2d.cc:1297-1313)
    MgLevel &fl = fem.fine();
    if ((rc = launch_fem_src(s, g, f, dt, fl.b, fl.pitch))) return rc;
    if (!const_rho && (rc = fem.set_density(g, f.rho))) return rc;
    fem.t_solve = time;  // t + dt
    if ((rc = proj_begin())) return rc;
    if ((rc = guess_features(fem, 1))) return rc;
    // From here to update_velocity nothing on s reads the level set, the
    // band or the windows, and the end-of-step reference map is final in
    // (nX, nY): the end-of-step rebuild (a lone solid's, or the batched one
    // of several) starts now on s2.
    const bool ovl = overlap && (obj.size() == 1 || (obj.size() > 1 && windows));
    if (ovl && !s2) {
        // made on first use only: a stream per fluid that never overlaps
        // (an ensemble's instances) would still take a share of the
        // process's hardware queues from the streams that do the work
        HIP_CHECK(hipStreamCreateWithFlags(&s2, hipStreamNonBlocking));  // default priority, like s
        HIP_CHECK(hipEventCreateWithFlags(&ev_fork, hipEventDisableTiming));
        HIP_CHECK(hipEventCreateWithFlags(&ev_join, hipEventDisableTiming));
    }
    LsRebuild pend;
    LsBatchPending *bpend = nullptr;
    auto drain = [&](int code) {  // an error return while s2 may still run
        if (ovl) (void)hipStreamSynchronize(s2);
        ls_batch_pending_free(bpend);
        return code;
    };
    if (ovl) {
        HIP_CHECK(hipEventRecord(ev_fork, s));
        HIP_CHECK(hipStreamWaitEvent(s2, ev_fork, 0));
        if (obj.size() == 1) {
            Solid &o0 = obj[0];
            if ((rc = ls_pin_rebuild_begin(s2, g, o0.c, o0.a, o0.w, o0.a.nX, o0.a.nY, pend))) return drain(rc);
        } else {
            if (!lsb.cap && (rc = ls_batch_alloc((int)obj.size(), lsb))) return drain(rc);
            if ((rc = ls_pin_rebuild_batch_begin(s2, g, jobs(), lsb, true, &bpend))) return drain(rc);
        }
    }
    w0 = std::chrono::steady_clock::now();
    rc = precision == INCRMT_PREC_MIXED ? fem.solve_mixed(mg_tol, max_cycles, &ms) : fem.solve(mg_tol, max_cycles, &ms);
    wc[1] += std::chrono::duration<double>(std::chrono::steady_clock::now() - w0).count();
    vcyc[1] += ms.cycles;
    nsolve[1]++;
    if (!rc) rc = proj_end();
    if (rc == 2) return drain(fail(INCRMT_ERR_NOT_CONVERGED, "V-cycle failed to converge in FEM problem"));
    if (rc) return drain(rc);
    st.fem_cycles = ms.cycles; st.fem_rel_res = ms.rel_res; st.total_fem_cycles += ms.cycles;
    // copy_pressure + update_velocity (fluid_2d.cc:1317-1374)
    // (no average is subtracted with zero pressure on the right edge,
    // fluid_2d.cc:1318: d_pavg stays zero)
    if (!g.zpr) {
        k_pavg_part<<<256, 256, 0, s>>>(fl.nx, fl.ny, g.px, g.py, fl.z, fl.pitch, d_part);
        if (hipError_t e = hipGetLastError()) return drain(hip_fail(e, "k_pavg_part", __FILE__, __LINE__));
        k_pavg_final<<<1, 256, 0, s>>>(d_part, 256, 1. / (double(g.m) * g.n), d_pavg);
        if (hipError_t e = hipGetLastError()) return drain(hip_fail(e, "k_pavg_final", __FILE__, __LINE__));
    }
    if ((rc = launch_pressure_velocity(s, g, f, o, fl.z, fl.pitch, d_pavg, dt))) return drain(rc);
    rc = ovl ? pin_join(pend, bpend, &passes) : pin_all(&passes);
    if (rc) return rc == 3 ? fail(INCRMT_ERR_NEWTON, "Newton iteration failed.") : rc;
    for (auto &ob : obj)
        if (ob.full_steps > 0) ob.full_steps--;
    if ((rc = launch_set_boundaries(s, g, f, o))) return rc;
    if ((rc = compute_stress())) return rc;
    if ((rc = tracers_stage2(s, g, f.u, f.v, d_oa, tr, dt, time))) return rc;
    st.steps++;
    return 0;
}

}  // namespace incrmt

using namespace incrmt;

struct incrmt_fluid {
    FluidEngine e;
};
struct incrmt_mg {
    Multigrid mg;
    hipStream_t s = nullptr;
    double *x = nullptr, *y = nullptr;
    double *bkeep = nullptr;  // the last solve's right-hand side (timed solves)
    // host row length: the solver's nx, plus the fixed node column of a FEM
    // system with zero pressure on the right edge (read as zero, written
    // as zero)
    int hm = 0;
    ~incrmt_mg() {
        if (x) (void)hipFree(x);
        if (y) (void)hipFree(y);
        if (bkeep) (void)hipFree(bkeep);
        if (s) (void)hipStreamDestroy(s);
    }
};

namespace {

// Copy between a ghosted (m+4)x(n+4) host array and a padded device array.
int copy_field(FluidEngine &e, double *dev, double *host, bool down) {
    const Grid &g = e.g;
    double *d0 = dev + g.ls(0, 0);
    if (down)
        HIP_CHECK(hipMemcpy2DAsync(host, g.ml * sizeof(double), d0, g.P * sizeof(double), g.ml * sizeof(double), g.nl,
                                   hipMemcpyDeviceToHost, e.s));
    else
        HIP_CHECK(hipMemcpy2DAsync(d0, g.P * sizeof(double), host, g.ml * sizeof(double), g.ml * sizeof(double), g.nl,
                                   hipMemcpyHostToDevice, e.s));
    HIP_CHECK(hipStreamSynchronize(e.s));
    return 0;
}

double *field_ptr(FluidEngine &e, const std::string &nm, int k, bool &is_int) {
    is_int = false;
    Fields &f = e.f;
    if (nm == "u") return f.u;
    if (nm == "v") return f.v;
    if (nm == "p") return f.p;
    if (nm == "rho") return f.rho;
    if (nm == "s11") return f.s11;
    if (nm == "s12") return f.s12;
    if (nm == "s21") return f.s21;
    if (nm == "s22") return f.s22;
    if (nm == "c0") return f.c0;    // u* after the step (fluid_2d.cc:1282-1288)
    if (nm == "c1") return f.c1;
    if (nm == "ul") return f.ulc;   // MAC-projected faces (fluid_2d.cc:1193-1235)
    if (nm == "vd") return f.vdc;
    if (k < 0 || k >= (int)e.obj.size()) return nullptr;
    ObjArrays &a = e.obj[k].a;
    if (nm == "X") return a.X;
    if (nm == "Y") return a.Y;
    if (nm == "nX") return a.nX;
    if (nm == "nY") return a.nY;
    if (nm == "phi") return a.phi;
    if (nm == "s") { is_int = true; return (double *)a.st; }
    return nullptr;
}

int add_object(incrmt_fluid *f, const ObjConst &c0) {
    if (!f) return fail(INCRMT_ERR_ARG, "null handle");
    if (f->e.initialized) return fail(INCRMT_ERR_ARG, "add_object after initialize");
    Solid so{};
    so.c = c0;
    so.c.set_velocity = 1;  // mat_const's default
    FluidEngine &e = f->e;
    // obj_field constants (obj_field.cc:14-20)
    so.c.eps = e.eps;
    so.c.pinw = e.pinw * e.g.dx;
    so.c.tf1 = 0.5 / e.eps;
    so.c.tf2 = 0.5 / M_PI;
    so.c.tf3 = M_PI / e.eps;
    so.c.phi_down = -e.eps - 4.5 * e.g.dx;
    so.c.phi_up = e.eps + 4.5 * e.g.dx;
    e.obj.push_back(so);
    return 0;
}

int check_ready(incrmt_fluid *f) {
    if (!f) return fail(INCRMT_ERR_ARG, "null handle");
    if (!f->e.initialized) return fail(INCRMT_ERR_ARG, "call before initialize");
    return 0;
}

}  // namespace

extern "C" {

const char *incrmt_last_error(void) { return g_last_error.c_str(); }

int incrmt_device_info(char *name, int len, double *hbm_gib) {
    hipDeviceProp_t p;
    HIP_CHECK(hipGetDeviceProperties(&p, 0));
    if (name && len > 0) {
        // the marketing name, with the ISA appended (some drivers leave the name empty)
        snprintf(name, len, "%s%s%s (%d CUs, LDS %zu KiB per workgroup)", p.name, p.name[0] ? " " : "", p.gcnArchName,
                 p.multiProcessorCount, (size_t)p.sharedMemPerBlock / 1024);
    }
    if (hbm_gib) *hbm_gib = p.totalGlobalMem / 1073741824.0;
    return 0;
}

int incrmt_copy_probe(long long bytes, int reps, double *ms_per_copy) {
    if (bytes <= 0 || reps <= 0 || !ms_per_copy) return fail(INCRMT_ERR_ARG, "bad copy probe");
    void *a = nullptr, *b = nullptr;
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = 0;
    auto run = [&]() -> int {
        HIP_CHECK(hipMalloc(&a, bytes));
        HIP_CHECK(hipMalloc(&b, bytes));
        HIP_CHECK(hipStreamCreate(&s));
        HIP_CHECK(hipEventCreate(&e0));
        HIP_CHECK(hipEventCreate(&e1));
        HIP_CHECK(hipMemsetAsync(a, 0, bytes, s));
        HIP_CHECK(hipMemcpyAsync(b, a, bytes, hipMemcpyDeviceToDevice, s));
        HIP_CHECK(hipEventRecord(e0, s));
        for (int r = 0; r < reps; r++) HIP_CHECK(hipMemcpyAsync(b, a, bytes, hipMemcpyDeviceToDevice, s));
        HIP_CHECK(hipEventRecord(e1, s));
        HIP_CHECK(hipEventSynchronize(e1));
        float ms = 0;
        HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
        *ms_per_copy = ms / reps;
        return 0;
    };
    rc = run();
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (s) (void)hipStreamDestroy(s);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    return rc;
}

int incrmt_fluid_create(int m, int n, double ax, double bx, double ay, double by, int x_prd, int y_prd,
                        double w_scale, double w_force, double visc, double rhof, double tw, double pinw,
                        incrmt_fluid **out) {
    if (!out) return fail(INCRMT_ERR_ARG, "null output handle");
    *out = nullptr;
    if (m < 8 || n < 8 || m > 65536 || n > 65536) return fail(INCRMT_ERR_ARG, "grid size out of range");
    if (!(bx > ax) || !(by > ay) || !(visc > 0) || !(rhof > 0) || !(tw > 0)) return fail(INCRMT_ERR_ARG, "bad geometry or material");
    incrmt_fluid *f = new incrmt_fluid;
    FluidEngine &e = f->e;
    Grid &g = e.g;
    g.m = m; g.n = n; g.ml = m + 4; g.nl = n + 4;
    g.P = round_up(m + 4, 32);
    g.N = (long long)kOrigin + (long long)(n + 4) * g.P + 64;
    g.ax = ax; g.bx = bx; g.ay = ay; g.by = by;
    g.dx = (bx - ax) / m; g.dy = (by - ay) / n;
    g.xsp = 1 / g.dx; g.ysp = 1 / g.dy; g.xxsp = g.xsp * g.xsp; g.yysp = g.ysp * g.ysp;
    g.lax = ax - 1.5 * g.dx; g.lay = ay - 1.5 * g.dy;
    g.px = x_prd != 0; g.py = y_prd != 0;
    g.vel = kSimRest;
    g.zpr = 0; g.U = 0; g.spbc = 0;
    e.visc = visc; e.rhof = rhof; e.tw = tw; e.pinw = pinw; e.w_scale = w_scale; e.w_force = w_force;
    e.eps = tw * (g.dx < g.dy ? g.dx : g.dy);
    hipError_t er = hipStreamCreateWithFlags(&e.s, hipStreamNonBlocking);
    if (er != hipSuccess) {
        delete f;
        return hip_fail(er, "hipStreamCreate", __FILE__, __LINE__);
    }
    *out = f;
    return 0;
}

int incrmt_fluid_add_spin(incrmt_fluid *f, double r_scale, double a1, double a2, double piv_r, double th_max,
                          double omega, double G, double rhos, double ev_trans_mult) {
    ObjConst c{};
    c.kind = kSpin;
    const double scf = r_scale / (1 + a1 + a2);  // object.hh:449-455
    c.c0 = scf * (1 - a2); c.c1 = scf * a1; c.c2 = 2 * scf * a2;
    c.piv_r = piv_r; c.th_max = th_max; c.omega = omega;
    c.G = G; c.rhos = rhos; c.ev_trans_mult = ev_trans_mult;
    if (int rc = add_object(f, c)) return rc;
    f->e.obj.back().r_scale = r_scale;
    return 0;
}

int incrmt_fluid_add_circle(incrmt_fluid *f, double cx, double cy, double cr, double grav, double G, double rhos,
                            double ev_trans_mult) {
    ObjConst c{};
    c.kind = kCircle;
    c.cx = cx; c.cy = cy; c.cr = cr; c.grav = grav;
    c.G = G; c.rhos = rhos; c.ev_trans_mult = ev_trans_mult;
    return add_object(f, c);
}

int incrmt_fluid_add_square(incrmt_fluid *f, double cx, double cy, double cr, double grav, double G, double rhos,
                            double ev_trans_mult) {
    ObjConst c{};
    c.kind = kSquare;
    c.cx = cx; c.cy = cy; c.cr = cr; c.grav = grav;
    c.G = G; c.rhos = rhos; c.ev_trans_mult = ev_trans_mult;
    return add_object(f, c);
}

int incrmt_fluid_set_solid_velocity(incrmt_fluid *f, int obj, int set_velocity) {
    if (!f) return fail(INCRMT_ERR_ARG, "null handle");
    if (f->e.initialized) return fail(INCRMT_ERR_ARG, "the initial velocity is set by initialize: set it before");
    if (obj < 0 || obj >= (int)f->e.obj.size()) return fail(INCRMT_ERR_ARG, "no such solid");
    f->e.obj[obj].c.set_velocity = set_velocity != 0;
    return 0;
}

int incrmt_fluid_add_full(incrmt_fluid *f, double G, double rhos, double ev_trans_mult) {
    ObjConst c{};
    c.kind = kFull;
    c.G = G; c.rhos = rhos; c.ev_trans_mult = ev_trans_mult;
    return add_object(f, c);
}

int incrmt_fluid_set_solid_periodic(incrmt_fluid *f, int on) {
    if (!f) return fail(INCRMT_ERR_ARG, "null handle");
    if (f->e.initialized) return fail(INCRMT_ERR_ARG, "set solid periodic boundaries before initialize");
    if (on && !(f->e.g.px && f->e.g.py)) return fail(INCRMT_ERR_ARG, "solid periodic boundaries need a doubly periodic box");
    f->e.g.spbc = on != 0;
    return 0;
}

int incrmt_fluid_add_flapper(incrmt_fluid *f, double cx, double cy, double cr, double cl, double ar, double al,
                             double theta, double m_stretch, double T, double G, double rhos, double ev_trans_mult) {
    if (!(m_stretch > 0) || !(ar > 0) || !(T > 0)) return fail(INCRMT_ERR_ARG, "bad flapper parameters");
    ObjConst c{};
    c.kind = kFlapper;
    c.cx = cx; c.cy = cy; c.cr = cr; c.cl = cl; c.ar = ar; c.al = al;
    c.amp = std::log(m_stretch) / ar;  // object.hh:264-267
    c.omega = 2 * M_PI / T;
    c.cth = std::cos(theta); c.sth = std::sin(theta);
    c.G = G; c.rhos = rhos; c.ev_trans_mult = ev_trans_mult;
    return add_object(f, c);
}

int incrmt_fluid_add_lamprey(incrmt_fluid *f, double cx, double cy, double theta, double le, double dis, double w1,
                             double w2, double T, double mlam, double G, double rhos, double ev_trans_mult) {
    if (!(mlam > 0) || !(T > 0) || !(le > dis) || !(w1 != 0)) return fail(INCRMT_ERR_ARG, "bad lamprey parameters");
    ObjConst c{};
    c.kind = kLamprey;
    c.cx = cx; c.cy = cy; c.dis = dis; c.la = le - dis; c.w1 = w1; c.w2 = w2;  // object.hh:567-570
    c.T = T; c.A = std::log(mlam);
    c.cth = std::cos(theta); c.sth = std::sin(theta);
    c.G = G; c.rhos = rhos; c.ev_trans_mult = ev_trans_mult;
    return add_object(f, c);
}

int incrmt_fluid_initialize(incrmt_fluid *f, double ev_mult, double dt_pad, double dt_ev_pad) {
    if (!f) return fail(INCRMT_ERR_ARG, "null handle");
    if (f->e.initialized) return fail(INCRMT_ERR_ARG, "initialize called twice");
    return f->e.initialize(ev_mult, dt_pad, dt_ev_pad);
}

int incrmt_fluid_set_solver(incrmt_fluid *f, double rel_tol, int max_cycles, int precision) {
    if (!f || !(rel_tol > 0) || max_cycles < 1 || (precision != INCRMT_PREC_FP64 && precision != INCRMT_PREC_MIXED))
        return fail(INCRMT_ERR_ARG, "bad solver settings");
    f->e.mg_tol = rel_tol;
    f->e.max_cycles = max_cycles;
    f->e.precision = precision;
    return 0;
}

int incrmt_fluid_step(incrmt_fluid *f, double dt) {
    if (!f || !(dt > 0)) return fail(INCRMT_ERR_ARG, "bad step");
    return f->e.step(dt);
}

int incrmt_fluid_steps(incrmt_fluid *f, double dt, int k) {
    if (!f || !(dt > 0) || k < 0) return fail(INCRMT_ERR_ARG, "bad step");
    for (int q = 0; q < k; q++)
        if (int rc = f->e.step(dt)) return rc;
    return 0;
}

double incrmt_fluid_dt_reg(const incrmt_fluid *f) { return f ? f->e.dt_reg : 0; }
double incrmt_fluid_time(const incrmt_fluid *f) { return f ? f->e.time : 0; }

int incrmt_fluid_timestep_select(const incrmt_fluid *f, double interval, double *adt) {
    if (!f || !adt || !(f->e.dt_reg > 0)) return -1;
    int l = static_cast<int>(interval / f->e.dt_reg) + 1;  // fluid_2d.hh:173-177
    *adt = interval / l;
    return l;
}

int incrmt_fluid_download(incrmt_fluid *f, const char *name, int obj, double *host) {
    if (!f || !name || !host || !f->e.initialized) return fail(INCRMT_ERR_ARG, "bad download");
    bool is_int;
    double *p = field_ptr(f->e, name, obj, is_int);
    if (!p) return fail(INCRMT_ERR_ARG, "unknown field");
    FluidEngine &e = f->e;
    HIP_CHECK(hipStreamSynchronize(e.s));
    if (!is_int) return copy_field(e, p, host, true);
    std::vector<int> tmp(e.g.N);
    HIP_CHECK(hipMemcpy(tmp.data(), p, e.g.N * sizeof(int), hipMemcpyDeviceToHost));
    for (int J = 0; J < e.g.nl; J++)
        for (int I = 0; I < e.g.ml; I++) host[J * e.g.ml + I] = tmp[e.g.ls(I, J)];
    return 0;
}

int incrmt_fluid_upload(incrmt_fluid *f, const char *name, int obj, const double *host) {
    if (!f || !name || !host || !f->e.initialized) return fail(INCRMT_ERR_ARG, "bad upload");
    bool is_int;
    double *p = field_ptr(f->e, name, obj, is_int);
    if (!p) return fail(INCRMT_ERR_ARG, "unknown field");
    FluidEngine &e = f->e;
    HIP_CHECK(hipStreamSynchronize(e.s));
    const std::string nm = name;
    if (nm == "X" || nm == "Y" || nm == "nX" || nm == "nY" || nm == "phi" || nm == "s") {
        ls_reset_window(e.obj[obj].w);  // the next band rebuild of this solid covers the whole grid
        e.obj[obj].full_steps = 1;          // and its sweeps the whole grid for one step
        if (int rc = e.update_window(obj, true)) return rc;
    }
    if (!is_int) return copy_field(e, p, const_cast<double *>(host), false);
    std::vector<int> tmp(e.g.N, 0);
    HIP_CHECK(hipMemcpy(tmp.data(), p, e.g.N * sizeof(int), hipMemcpyDeviceToHost));
    for (int J = 0; J < e.g.nl; J++)
        for (int I = 0; I < e.g.ml; I++) tmp[e.g.ls(I, J)] = (int)host[J * e.g.ml + I];
    HIP_CHECK(hipMemcpy(p, tmp.data(), e.g.N * sizeof(int), hipMemcpyHostToDevice));
    return 0;
}

int incrmt_fluid_set_time(incrmt_fluid *f, double t) {
    if (!f) return fail(INCRMT_ERR_ARG, "null handle");
    f->e.time = t;
    return 0;
}

int incrmt_fluid_sync(incrmt_fluid *f) {
    if (!f) return fail(INCRMT_ERR_ARG, "null handle");
    HIP_CHECK(hipStreamSynchronize(f->e.s));
    return 0;
}

int incrmt_fluid_get_stats(incrmt_fluid *f, incrmt_stats *out) {
    if (!f || !out) return fail(INCRMT_ERR_ARG, "null argument");
    if (int rc = f->e.proj_collect()) return rc;
    *out = f->e.st;
    out->proj_ms = f->e.proj_ms;
    out->smoother_ms = f->e.mac.smoother_ms + f->e.fem.smoother_ms;
    out->smoother_launches = f->e.mac.smoother_launches + f->e.fem.smoother_launches;
    out->smoother_bytes = f->e.mac.smoother_bytes + f->e.fem.smoother_bytes;
    return 0;
}

int incrmt_fluid_time_projections(incrmt_fluid *f, int enable) {
    if (!f) return fail(INCRMT_ERR_ARG, "null handle");
    if (int rc = f->e.proj_collect()) return rc;
    f->e.time_proj = enable != 0;
    f->e.proj_ms = 0;
    return 0;
}

int incrmt_fluid_set_solid_windows(incrmt_fluid *f, int on) {
    if (!f) return fail(INCRMT_ERR_ARG, "null handle");
    f->e.windows = on != 0;
    return 0;
}

int incrmt_fluid_time_smoothers(incrmt_fluid *f, int enable) {
    if (!f) return fail(INCRMT_ERR_ARG, "null handle");
    f->e.mac.time_smoother = f->e.fem.time_smoother = enable != 0;
    f->e.mac.smoother_ms = f->e.fem.smoother_ms = 0;
    f->e.mac.smoother_launches = f->e.fem.smoother_launches = 0;
    f->e.mac.smoother_bytes = f->e.fem.smoother_bytes = 0;
    return 0;
}

// ---- frame output: fluid_2d::write_files (fluid_2d_io.cc:60-179) ----------
// The fields are copied to the host once per frame and written in the
// reference's Gnuplot matrix-binary format (float32; first row: column count
// then the x coordinates; every further row: y then the values).
namespace {

struct HostFrame {
    int m, n, ml, nl;
    double ax, ay, dx, dy, xsp, ysp;
    std::vector<double> u, v, p, rho, c0, c1;
    std::vector<std::vector<double>> phi, X, Y;
    std::vector<std::vector<int>> st;
    double eps;
};

int write_matrix(const std::string &path, const HostFrame &h, int mode, int li, int ui, int lj, int uj) {
    const int si = ui - li, sj = uj - lj;
    if (si <= 0 || sj <= 0) return 0;
    FILE *fp = fopen(path.c_str(), "wb");
    if (!fp) return fail(INCRMT_ERR_ARG, ("cannot open " + path).c_str());
    const bool cen = !((1 << mode) & (8 | 16 | 512 | 2048));  // fluid_2d_io.cc:130-133
    const double lx = h.ax + h.dx * (li - (cen ? 1.5 : 2.)), ly = h.ay + h.dy * (lj - (cen ? 1.5 : 2.));
    std::vector<float> buf(si + 1);
    buf[0] = (float)si;
    for (int i = 0; i < si; i++) buf[i + 1] = (float)(lx + i * h.dx);
    fwrite(buf.data(), sizeof(float), si + 1, fp);
    const int ml = h.ml;
    auto minobj = [&](int ij) {  // min_phi_obj (fluid_2d.cc:1502-1508)
        size_t k = 0;
        double mp = h.phi[0][ij];
        for (size_t q = 1; q < h.phi.size(); q++)
            if (h.phi[q][ij] < mp) { mp = h.phi[q][ij]; k = q; }
        return k;
    };
    auto corner_phi = [&](size_t k, int ij) {  // fluid_2d.hh:266-268
        const double *p = h.phi[k].data() + ij;
        return p[-ml - 1] + p[-ml] + p[-1] + *p;
    };
    for (int j = 0; j < sj; j++) {
        buf[0] = (float)(ly + j * h.dy);
        const int ijr = li + ml * (lj + j);
        for (int i = 0; i < si; i++) {
            const int ij = ijr + i;
            double val = 0;
            switch (mode) {
            case 0: val = h.u[ij]; break;
            case 1: val = h.v[ij]; break;
            case 2: val = sqrt(h.u[ij] * h.u[ij] + h.v[ij] * h.v[ij]); break;
            case 3: val = h.p[ij]; break;
            case 4:  // vorticity at the lower-left corner (fluid_2d.hh:252-255)
                val = 0.5 * (h.xsp * (h.v[ij] - h.v[ij - 1] + h.v[ij - ml] - h.v[ij - ml - 1]) -
                             h.ysp * (h.u[ij] + h.u[ij - 1] - h.u[ij - ml] - h.u[ij - ml - 1]));
                break;
            case 5: val = h.rho[ij]; break;
            case 6: val = h.X[minobj(ij)][ij]; break;
            case 7: val = h.Y[minobj(ij)][ij]; break;
            case 8: val = h.phi[minobj(ij)][ij]; break;
            case 9: {  // vol_change (fluid_2d_io.cc:303-316)
                size_t k = 0;
                double mp = corner_phi(0, ij);
                for (size_t q = 1; q < h.phi.size(); q++)
                    if (corner_phi(q, ij) < mp) { mp = corner_phi(q, ij); k = q; }
                if (corner_phi(k, ij) > 0) { val = 0; break; }
                const double *X = h.X[k].data() + ij, *Y = h.Y[k].data() + ij;
                const double Xx = 0.5 * h.xsp * (X[0] + X[-ml] - X[-1] - X[-ml - 1]),
                             Xy = 0.5 * h.ysp * (X[0] + X[-1] - X[-ml] - X[-ml - 1]),
                             Yx = 0.5 * h.xsp * (Y[0] + Y[-ml] - Y[-1] - Y[-ml - 1]),
                             Yy = 0.5 * h.ysp * (Y[0] + Y[-1] - Y[-ml] - Y[-ml - 1]);
                val = 1. / (Xx * Yy - Xy * Yx) - 1;
                break;
            }
            case 10: val = h.c0[ij]; break;
            case 11: val = h.c1[ij]; break;
            case 12:  // divergence at the lower-left corner (fluid_2d.hh:259-262)
                val = 0.5 * (h.xsp * (h.u[ij] - h.u[ij - 1] + h.u[ij - ml] - h.u[ij - ml - 1]) +
                             h.ysp * (h.v[ij] + h.v[ij - 1] - h.v[ij - ml] - h.v[ij - ml - 1]));
                break;
            }
            buf[i + 1] = (float)val;
        }
        fwrite(buf.data(), sizeof(float), si + 1, fp);
    }
    fclose(fp);
    return 0;
}

}  // namespace

int incrmt_fluid_solve(incrmt_fluid *f, const char *dir, double duration, int frames, unsigned int fflags) {
    if (int rc = check_ready(f)) return rc;
    if (!(duration > 0) || frames < 1) return fail(INCRMT_ERR_ARG, "bad duration or frame count");
    FluidEngine &e = f->e;
    auto now = []() { return std::chrono::steady_clock::now(); };
    auto secs = [](std::chrono::steady_clock::time_point a, std::chrono::steady_clock::time_point b) {
        return std::chrono::duration<double>(b - a).count();
    };
    double adt;
    const int l = incrmt_fluid_timestep_select(f, duration / frames, &adt);
    if (l < 0) return fail(INCRMT_ERR_ARG, "timestep_select failed");
    if (dir) {  // save_header (fluid_2d_io.cc:13-19)
        const std::string hp = std::string(dir) + "/header";
        FILE *fp = std::fopen(hp.c_str(), e.f_num == 0 ? "w" : "a");
        if (!fp) return fail(INCRMT_ERR_ARG, "cannot open the header file");
        std::fprintf(fp, "%g %g %d\n", e.time, e.time + duration, frames);
        std::fclose(fp);
    }
    if (e.f_num == 0) {
        if (dir)
            if (int rc = incrmt_fluid_write_frame(f, dir, 0, fflags)) return rc;
        std::puts("# Output frame 0");
    }
    auto t0 = now();
    for (int k = 1; k <= frames; k++) {
        if (int rc = incrmt_fluid_steps(f, adt, l)) return rc;
        HIP_CHECK(hipStreamSynchronize(e.s));
        const auto t1 = now();
        if (dir)
            if (int rc = incrmt_fluid_write_frame(f, dir, k + e.f_num, fflags)) return rc;
        const auto t2 = now();
        // the reference prints TGMG's V-cycle count and the solver wall
        // time with 8192 (its tgmg_predict counters are not available: the
        // V-cycles run in the frame stand in), else the mean V-cycles per
        // solve
        const double avg0 = e.nsolve[0] ? double(e.vcyc[0]) / e.nsolve[0] : 0;
        const double avg1 = e.nsolve[1] ? double(e.vcyc[1]) / e.nsolve[1] : 0;
        if (fflags & 8192)
            std::printf("%d %d %.8g %.8g  %ld %.8g %ld %.8g\n", k + e.f_num, l, secs(t0, t1), secs(t1, t2), e.vcyc[0],
                        e.wc[0], e.vcyc[1], e.wc[1]);
        else
            std::printf("# Output frame %d [%d, %.8g s, %.8g s] {MAC %.2f, FEM %.2f}\n", k + e.f_num, l, secs(t0, t1),
                        secs(t1, t2), avg0, avg1);
        std::fflush(stdout);
        e.wc[0] = e.wc[1] = 0;
        e.vcyc[0] = e.vcyc[1] = e.nsolve[0] = e.nsolve[1] = 0;
        t0 = t2;
    }
    e.f_num += frames;
    return 0;
}

int incrmt_fluid_write_frame(incrmt_fluid *f, const char *dir, int sn, unsigned int fflags) {
    if (int rc = check_ready(f)) return rc;
    if (!dir) return fail(INCRMT_ERR_ARG, "null directory");
    FluidEngine &e = f->e;
    const Grid &g = e.g;
    HostFrame h;
    h.m = g.m; h.n = g.n; h.ml = g.ml; h.nl = g.nl;
    h.ax = g.ax; h.ay = g.ay; h.dx = g.dx; h.dy = g.dy; h.xsp = g.xsp; h.ysp = g.ysp;
    const size_t N = (size_t)g.ml * g.nl;
    auto get = [&](const char *nm, int k, std::vector<double> &out) -> int {
        out.assign(N, 0.0);
        return incrmt_fluid_download(f, nm, k, out.data());
    };
    int rc;
    if ((rc = get("u", 0, h.u)) || (rc = get("v", 0, h.v)) || (rc = get("p", 0, h.p)) || (rc = get("rho", 0, h.rho)) ||
        (rc = get("c0", 0, h.c0)) || (rc = get("c1", 0, h.c1)))
        return rc;
    const int nobj = (int)e.obj.size();
    h.phi.resize(nobj); h.X.resize(nobj); h.Y.resize(nobj); h.st.resize(nobj);
    for (int k = 0; k < nobj; k++) {
        std::vector<double> s;
        if ((rc = get("phi", k, h.phi[k])) || (rc = get("X", k, h.X[k])) || (rc = get("Y", k, h.Y[k])) ||
            (rc = get("s", k, s)))
            return rc;
        h.st[k].resize(N);
        for (size_t q = 0; q < N; q++) h.st[k][q] = (int)s[q];
    }
    const std::string base = std::string(dir) + "/";
    auto out = [&](const char *prefix, int mode) -> int {
        const std::string path = base + prefix + "." + std::to_string(sn);
        if (fflags & 1024) {  // whole grid (fluid_2d_io.cc:109-111)
            const int lo = mode == 4 ? 1 : 0;
            return write_matrix(path, h, mode, lo, g.ml, lo, g.nl);
        }
        const int dis = (1 << mode) & (8 | 16 | 512 | 2048) ? 1 : 2;
        return write_matrix(path, h, mode, 2, g.ml - dis, 2, g.nl - dis);
    };
    if ((fflags & 1) && (rc = out("u", 0))) return rc;
    if ((fflags & 2) && (rc = out("v", 1))) return rc;
    if ((fflags & 4) && (rc = out("spd", 2))) return rc;
    if ((fflags & 8) && (rc = out("p", 3))) return rc;
    if ((fflags & 16) && (rc = out("w", 4))) return rc;
    if (!e.const_rho && (fflags & 32) && (rc = out("rho", 5))) return rc;
    if ((fflags & 4096) && (rc = out("div", 12))) return rc;
    if (nobj > 0 && (fflags & 8640)) {
        if (fflags & 512) {  // bounding box of the solids' inner band, padded by 2 (fluid_2d_io.cc:75-83)
            int li = g.ml, ui = 0, lj = g.nl, uj = 0;
            for (int k = 0; k < nobj; k++)
                for (int J = 0; J < g.nl; J++)
                    for (int I = 0; I < g.ml; I++)
                        if (h.st[k][J * g.ml + I] < 4) {
                            if (I < li) li = I;
                            if (I >= ui) ui = I + 1;
                            if (J < lj) lj = J;
                            if (J >= uj) uj = J + 1;
                        }
            li -= 2; ui += 2; lj -= 2; uj += 2;
            if (li < 0) li = 0;
            if (ui > g.ml) ui = g.ml;
            if (lj < 0) lj = 0;
            if (uj > g.nl) uj = g.nl;
            const std::string sfx = "." + std::to_string(sn);
            if ((fflags & 64) && (rc = write_matrix(base + "X" + sfx, h, 6, li, ui, lj, uj))) return rc;
            if ((fflags & 128) && (rc = write_matrix(base + "Y" + sfx, h, 7, li, ui, lj, uj))) return rc;
            if ((fflags & 256) && (rc = write_matrix(base + "phi" + sfx, h, 8, li, ui, lj, uj))) return rc;
        } else {
            if ((fflags & 64) && (rc = out("X", 6))) return rc;
            if ((fflags & 128) && (rc = out("Y", 7))) return rc;
            if ((fflags & 256) && (rc = out("phi", 8))) return rc;
        }
        if ((fflags & 8192) && (rc = out("J", 9))) return rc;
    }
    if (fflags & 2048) {
        if ((rc = out("ds", 10)) || (rc = out("dp", 11))) return rc;
    }
    // tracer positions (fluid_2d_io.cc:24-55): x, y as float32 pairs
    if (e.tr.ntrace > 0) {
        std::vector<double> tm(4 * (size_t)e.tr.ntrace);
        if ((rc = incrmt_fluid_get_tracers(f, tm.data()))) return rc;
        std::vector<float> b(2 * (size_t)e.tr.ntrace);
        for (int q = 0; q < e.tr.ntrace; q++) { b[2 * q] = (float)tm[4 * q]; b[2 * q + 1] = (float)tm[4 * q + 1]; }
        FILE *fp = fopen((base + "trace." + std::to_string(sn)).c_str(), "wb");
        if (!fp) return fail(INCRMT_ERR_ARG, "cannot open the tracer file");
        fwrite(b.data(), sizeof(float), b.size(), fp);
        fclose(fp);
    }
    // the objects' own output: a track.dat record, the file rewritten whole
    if (e.tr.ntrack > 0) {
        if ((rc = incrmt_fluid_frame_output(f))) return rc;
        for (size_t k = 0, t = 0; k < e.obj.size(); k++) {
            if (e.obj[k].c.kind != kSpin) continue;
            int nrec = 0;
            if ((rc = incrmt_fluid_track(f, (int)k, nullptr, 0, &nrec))) return rc;
            const int have = nrec < e.tr.rec_cap ? nrec : e.tr.rec_cap;
            std::vector<double> rec(4 * (size_t)have);
            if ((rc = incrmt_fluid_track(f, (int)k, rec.data(), have, &nrec))) return rc;
            const std::string nm = t == 0 ? "track.dat" : "track" + std::to_string(t) + ".dat";
            FILE *fp = fopen((base + nm).c_str(), "wb");
            if (!fp) return fail(INCRMT_ERR_ARG, "cannot open track.dat");
            fwrite(rec.data(), sizeof(double), rec.size(), fp);
            fclose(fp);
            t++;
        }
    }
    return 0;
}

int incrmt_fluid_time_fine_smoother(incrmt_fluid *f, int kind, int sweeps, int precision, double *ms_per_launch,
                                    int *launches) {
    if (int rc = check_ready(f)) return rc;
    if ((kind != INCRMT_MG_MAC && kind != INCRMT_MG_FEM) || sweeps < 1) return fail(INCRMT_ERR_ARG, "bad arguments");
    Multigrid &mg = kind == INCRMT_MG_MAC ? f->e.mac : f->e.fem;
    long n = 0;
    int rc = mg.time_fine_smoother(sweeps, precision == INCRMT_PREC_MIXED, ms_per_launch, &n);
    if (launches) *launches = (int)n;
    return rc;
}

int incrmt_fluid_set_guess(incrmt_fluid *f, int mode) {
    if (!f) return fail(INCRMT_ERR_ARG, "null handle");
    if (mode != INCRMT_GUESS_PREVIOUS && mode != INCRMT_GUESS_EXTRAPOLATE && mode != INCRMT_GUESS_QUADRATIC &&
        mode != INCRMT_GUESS_LEARNED)
        return fail(INCRMT_ERR_ARG, "bad guess mode");
    f->e.guess_mode = mode;
    f->e.mac.set_extrapolate(mode);
    f->e.fem.set_extrapolate(mode);
    if (mode == INCRMT_GUESS_LEARNED && f->e.initialized) return f->e.ensure_guess();
    return 0;
}

int incrmt_fluid_set_guess_training(incrmt_fluid *f, int train, float lr, double fallback_tol) {
    if (!f) return fail(INCRMT_ERR_ARG, "null handle");
    if (!(lr > 0) || !(fallback_tol >= 0)) return fail(INCRMT_ERR_ARG, "bad learning rate or tolerance");
    FluidEngine &e = f->e;
    e.guess_train = train != 0;
    e.guess_lr = lr;
    e.guess_tol = fallback_tol;
    e.mac.policy.reset();
    e.fem.policy.reset();
    if (e.initialized && (e.guess_train || e.guess_mode == INCRMT_GUESS_LEARNED)) return e.ensure_guess();
    e.mac.train = e.fem.train = e.guess_train;
    return 0;
}

static int guess_net(incrmt_fluid *f, int kind, UNet **out) {
    if (!f) return fail(INCRMT_ERR_ARG, "null handle");
    if (kind != INCRMT_MG_MAC && kind != INCRMT_MG_FEM) return fail(INCRMT_ERR_ARG, "bad solver kind");
    if (!f->e.initialized) return fail(INCRMT_ERR_ARG, "the guess networks are made by initialize");
    if (int rc = f->e.ensure_guess()) return rc;
    *out = f->e.gnet[kind];
    return 0;
}

int incrmt_fluid_guess_nparams(incrmt_fluid *f, int kind, long long *n) {
    UNet *net;
    if (int rc = guess_net(f, kind, &net)) return rc;
    if (!n) return fail(INCRMT_ERR_ARG, "null argument");
    *n = (long long)net->nparams();
    return 0;
}

int incrmt_fluid_get_guess_params(incrmt_fluid *f, int kind, float *params) {
    UNet *net;
    if (int rc = guess_net(f, kind, &net)) return rc;
    if (!params) return fail(INCRMT_ERR_ARG, "null argument");
    return net->get_params(params);
}

int incrmt_fluid_set_guess_params(incrmt_fluid *f, int kind, const float *params) {
    UNet *net;
    if (int rc = guess_net(f, kind, &net)) return rc;
    if (!params) return fail(INCRMT_ERR_ARG, "null argument");
    (kind == INCRMT_MG_MAC ? f->e.mac : f->e.fem).policy.reset();
    return net->set_params(params);
}

int incrmt_fluid_guess_stats(incrmt_fluid *f, int kind, double *out8) {
    if (!f || !out8) return fail(INCRMT_ERR_ARG, "null argument");
    double *out6 = out8;
    if (kind != INCRMT_MG_MAC && kind != INCRMT_MG_FEM) return fail(INCRMT_ERR_ARG, "bad solver kind");
    Multigrid &mg = kind == INCRMT_MG_MAC ? f->e.mac : f->e.fem;
    double loss = 0;
    if (mg.net && mg.train_steps > 0)
        if (int rc = mg.read_train_loss(&loss)) return rc;
    out6[0] = (double)mg.learned_solves;
    out6[1] = (double)mg.guess_fallbacks;
    out6[2] = (double)mg.train_steps;
    out6[3] = loss;
    out6[4] = mg.policy.off ? 1 : 0;
    out6[5] = (double)mg.policy.misses;
    out8[6] = mg.last_rel0;
    out8[7] = mg.last_rel1;
    return 0;
}

int incrmt_fluid_set_guess_cost(incrmt_fluid *f, int mode, double cycles) {
    if (!f || mode < 0 || mode > 2 || (mode == 2 && !(cycles >= 0))) return fail(INCRMT_ERR_ARG, "bad arguments");
    for (Multigrid *mg : {&f->e.mac, &f->e.fem}) {
        mg->policy.reset();
        mg->policy.cost_aware = mode != 0;
        mg->policy.cost = mode == 2 ? cycles : -1;
    }
    return 0;
}

int incrmt_fluid_guess_cost_stats(incrmt_fluid *f, int kind, double *out6) {
    if (!f || !out6) return fail(INCRMT_ERR_ARG, "null argument");
    if (kind != INCRMT_MG_MAC && kind != INCRMT_MG_FEM) return fail(INCRMT_ERR_ARG, "bad solver kind");
    const GuessPolicy &p = (kind == INCRMT_MG_MAC ? f->e.mac : f->e.fem).policy;
    out6[0] = p.cost;
    out6[1] = p.ema_learn;
    out6[2] = p.ema_quad;
    out6[3] = p.benefit_off ? 1 : 0;
    out6[4] = p.nlearn;
    out6[5] = p.nquad;
    return 0;
}

int incrmt_guess_policy_run(const double *rel, int n, double tol, int *decision) {
    if (!rel || !decision || n < 0) return fail(INCRMT_ERR_ARG, "null argument");
    GuessPolicy p;
    for (int k = 0; k < n; k++) decision[k] = !p.use() ? -1 : (p.accept(rel[k], tol) ? 1 : 0);
    return 0;
}

int incrmt_fluid_set_sim_velocity(incrmt_fluid *f, int kind) {
    if (!f) return fail(INCRMT_ERR_ARG, "null handle");
    if (f->e.initialized) return fail(INCRMT_ERR_ARG, "the initial velocity is set by initialize: set it before");
    if (kind != INCRMT_SIM_REST && kind != INCRMT_SIM_FOUR_ROLLS && kind != INCRMT_SIM_VELOCITY_PULSES)
        return fail(INCRMT_ERR_ARG, "bad simulation velocity");
    f->e.g.vel = kind;
    return 0;
}

int incrmt_fluid_set_horizontal_flow(incrmt_fluid *f, double U) {
    if (!f) return fail(INCRMT_ERR_ARG, "null handle");
    if (f->e.initialized) return fail(INCRMT_ERR_ARG, "the flow is set up by initialize: set it before");
    if (f->e.g.px) return fail(INCRMT_ERR_ARG, "horizontal flow needs a non-periodic x direction");
    if (!std::isfinite(U)) return fail(INCRMT_ERR_ARG, "bad flow velocity");
    Grid &g = f->e.g;
    g.zpr = 1; g.U = U; g.vel = kSimHorizFlow;
    return 0;
}

int incrmt_fluid_set_overlap(incrmt_fluid *f, int on) {
    if (!f) return fail(INCRMT_ERR_ARG, "null handle");
    f->e.overlap = on != 0;
    return 0;
}

int incrmt_fluid_set_launch(incrmt_fluid *f, int mode) {
    if (!f) return fail(INCRMT_ERR_ARG, "null handle");
    if (mode != INCRMT_LAUNCH_DIRECT && mode != INCRMT_LAUNCH_GRAPH) return fail(INCRMT_ERR_ARG, "bad launch mode");
    f->e.launch_req = mode;
    // before initialize the request is kept and applied after the solvers'
    // set-up (which otherwise picks the mode by size)
    if (f->e.initialized) f->e.mac.use_graphs = f->e.fem.use_graphs = mode == INCRMT_LAUNCH_GRAPH;
    return 0;
}

int incrmt_fluid_set_ntrace(incrmt_fluid *f, int ntrace) {
    if (!f) return fail(INCRMT_ERR_ARG, "null handle");
    if (f->e.initialized) return fail(INCRMT_ERR_ARG, "tracers are placed by initialize: set them before");
    if (ntrace < 0 || ntrace > (1 << 24)) return fail(INCRMT_ERR_ARG, "bad tracer count");
    f->e.ntrace_req = ntrace;
    return 0;
}

int incrmt_fluid_get_tracers(incrmt_fluid *f, double *xy4) {
    if (int rc = check_ready(f)) return rc;
    if (!xy4) return fail(INCRMT_ERR_ARG, "null argument");
    FluidEngine &e = f->e;
    if (e.tr.ntrace == 0) return 0;
    HIP_CHECK(hipStreamSynchronize(e.s));
    HIP_CHECK(hipMemcpy(xy4, e.tr.tm, 4 * (size_t)e.tr.ntrace * sizeof(double), hipMemcpyDeviceToHost));
    return 0;
}

int incrmt_fluid_set_tracers(incrmt_fluid *f, const double *xy4) {
    if (int rc = check_ready(f)) return rc;
    if (!xy4) return fail(INCRMT_ERR_ARG, "null argument");
    FluidEngine &e = f->e;
    if (e.tr.ntrace == 0) return 0;
    HIP_CHECK(hipStreamSynchronize(e.s));
    HIP_CHECK(hipMemcpy(e.tr.tm, xy4, 4 * (size_t)e.tr.ntrace * sizeof(double), hipMemcpyHostToDevice));
    return 0;
}

// track state: x, y, stage u, stage v, X, Y, count, prev_q, wind, correction on
static int find_track(incrmt_fluid *f, int obj, int *k) {
    FluidEngine &e = f->e;
    if (obj < 0 || obj >= (int)e.obj.size() || e.obj[obj].c.kind != kSpin) return fail(INCRMT_ERR_ARG, "object has no track");
    int q = 0;
    for (int i = 0; i < obj; i++) q += e.obj[i].c.kind == kSpin;
    *k = q;
    return 0;
}

int incrmt_fluid_get_track_state(incrmt_fluid *f, int obj, double *st10) {
    if (int rc = check_ready(f)) return rc;
    int k;
    if (int rc = find_track(f, obj, &k)) return rc;
    FluidEngine &e = f->e;
    TrackDev t;
    HIP_CHECK(hipStreamSynchronize(e.s));
    HIP_CHECK(hipMemcpy(&t, e.tr.trk + k, sizeof t, hipMemcpyDeviceToHost));
    for (int q = 0; q < 6; q++) st10[q] = t.tk[q];
    st10[6] = t.count; st10[7] = t.prev_q; st10[8] = t.wind; st10[9] = t.tcorr;
    return 0;
}

int incrmt_fluid_set_track_state(incrmt_fluid *f, int obj, const double *st10) {
    if (int rc = check_ready(f)) return rc;
    int k;
    if (int rc = find_track(f, obj, &k)) return rc;
    FluidEngine &e = f->e;
    TrackDev t;
    HIP_CHECK(hipStreamSynchronize(e.s));
    HIP_CHECK(hipMemcpy(&t, e.tr.trk + k, sizeof t, hipMemcpyDeviceToHost));
    for (int q = 0; q < 6; q++) t.tk[q] = st10[q];
    t.count = (int)st10[6]; t.prev_q = (int)st10[7]; t.wind = (int)st10[8]; t.tcorr = st10[9] != 0;
    HIP_CHECK(hipMemcpy(e.tr.trk + k, &t, sizeof t, hipMemcpyHostToDevice));
    return 0;
}

int incrmt_fluid_frame_output(incrmt_fluid *f) {
    if (int rc = check_ready(f)) return rc;
    FluidEngine &e = f->e;
    return tracers_frame_output(e.s, e.g, e.d_oa, e.tr, e.time);
}

int incrmt_fluid_track(incrmt_fluid *f, int obj, double *rec, int cap, int *nrec) {
    if (int rc = check_ready(f)) return rc;
    int k;
    if (int rc = find_track(f, obj, &k)) return rc;
    FluidEngine &e = f->e;
    TrackDev t;
    HIP_CHECK(hipStreamSynchronize(e.s));
    HIP_CHECK(hipMemcpy(&t, e.tr.trk + k, sizeof t, hipMemcpyDeviceToHost));
    const int have = t.nrec < e.tr.rec_cap ? t.nrec : e.tr.rec_cap;
    if (nrec) *nrec = t.nrec;
    const int n = have < cap ? have : cap;
    if (n > 0 && rec)
        HIP_CHECK(hipMemcpy(rec, e.tr.rec + (size_t)k * e.tr.rec_cap * 4, 4 * (size_t)n * sizeof(double),
                            hipMemcpyDeviceToHost));
    return 0;
}

int incrmt_fluid_describe(incrmt_fluid *f, char *buf, int len) {
    if (int rc = check_ready(f)) return rc;
    if (!buf || len <= 0) return fail(INCRMT_ERR_ARG, "null buffer");
    const std::string d = "MAC " + f->e.mac.shapes() + "; FEM " + f->e.fem.shapes() + "; launch MAC " +
                          (f->e.mac.use_graphs ? "graph" : "direct") + ", FEM " +
                          (f->e.fem.use_graphs ? "graph" : "direct");
    snprintf(buf, len, "%s", d.c_str());
    return 0;
}

void incrmt_fluid_destroy(incrmt_fluid *f) { delete f; }

// ------------------------------------------------------------ standalone MG
int incrmt_mg_create(int kind, int m, int n, int x_prd, int y_prd, double dx, double dy, incrmt_mg **out) {
    return incrmt_mg_create_bc(kind, m, n, x_prd, y_prd, 0, dx, dy, out);
}

int incrmt_mg_create_bc(int kind, int m, int n, int x_prd, int y_prd, int zpr, double dx, double dy,
                        incrmt_mg **out) {
    if (!out) return fail(INCRMT_ERR_ARG, "null output handle");
    *out = nullptr;
    if (kind != INCRMT_MG_MAC && kind != INCRMT_MG_FEM) return fail(INCRMT_ERR_ARG, "bad solver kind");
    if (m < 2 || n < 2 || !(dx > 0) || !(dy > 0)) return fail(INCRMT_ERR_ARG, "bad grid");
    if (zpr && x_prd) return fail(INCRMT_ERR_ARG, "a zero-pressure right edge needs a non-periodic x direction");
    const bool fixcol = zpr && kind == INCRMT_MG_FEM;
    incrmt_mg *h = new incrmt_mg;
    hipError_t er = hipStreamCreateWithFlags(&h->s, hipStreamNonBlocking);
    if (er != hipSuccess) {
        delete h;
        return hip_fail(er, "hipStreamCreate", __FILE__, __LINE__);
    }
    h->hm = m;
    int rc = h->mg.init(kind == INCRMT_MG_MAC ? kMAC : kFEM, fixcol ? m - 1 : m, n, dx, dy, h->s, x_prd != 0,
                        y_prd != 0, zpr != 0);
    if (rc == INCRMT_ERR_UNSUPPORTED)
        rc = fail(rc, "a singly periodic direction or a zero-pressure edge needs an even size whose halvings end "
                      "at most 16 points (a doubly periodic box: at least 8 points)");
    if (!rc) {
        const MgLevel &l = h->mg.fine();
        const size_t len = (size_t)(l.ny + 2) * l.pitch + 64;
        if (hipMalloc((void **)&h->x, len * sizeof(double)) != hipSuccess ||
            hipMalloc((void **)&h->y, len * sizeof(double)) != hipSuccess ||
            memset_sync(h->x, 0, len * sizeof(double), h->s) != hipSuccess ||
            memset_sync(h->y, 0, len * sizeof(double), h->s) != hipSuccess)
            rc = fail(INCRMT_ERR_HIP, "allocation failed");
    }
    if (rc) {
        delete h;
        return rc;
    }
    *out = h;
    return 0;
}

int incrmt_mg_set_irho(incrmt_mg *h, const double *irho) {
    if (!h) return fail(INCRMT_ERR_ARG, "null handle");
    int rc = h->mg.set_irho_host(irho);
    if (rc) return rc;
    HIP_CHECK(hipStreamSynchronize(h->s));
    return 0;
}

static int mg_copy(incrmt_mg *h, double *dev, const double *host_in, double *host_out) {
    const MgLevel &l = h->mg.fine();
    const size_t hp = (size_t)h->hm * sizeof(double);
    if (host_in)
        HIP_CHECK(hipMemcpy2DAsync(dev, l.pitch * sizeof(double), host_in, hp, l.nx * sizeof(double), l.ny,
                                   hipMemcpyHostToDevice, h->s));
    else
        HIP_CHECK(hipMemcpy2DAsync(host_out, hp, dev, l.pitch * sizeof(double), l.nx * sizeof(double), l.ny,
                                   hipMemcpyDeviceToHost, h->s));
    HIP_CHECK(hipStreamSynchronize(h->s));
    if (host_out)
        for (int j = 0; j < l.ny; j++)
            for (int i = l.nx; i < h->hm; i++) host_out[(size_t)j * h->hm + i] = 0;  // the fixed column
    return 0;
}

int incrmt_mg_solve(incrmt_mg *h, const double *b, double *z, double rel_tol, int max_cycles, int precision,
                    int *cycles, double *rel_res) {
    if (!h || !b || !z || !(rel_tol > 0) || max_cycles < 1) return fail(INCRMT_ERR_ARG, "bad solve arguments");
    MgLevel &l = h->mg.fine();
    if (int rc = mg_copy(h, l.b, b, nullptr)) return rc;
    if (int rc = mg_copy(h, l.z, z, nullptr)) return rc;
    const size_t len = (size_t)(l.ny + 2) * l.pitch;
    if (!h->bkeep) HIP_CHECK(hipMalloc((void **)&h->bkeep, len * sizeof(double)));
    HIP_CHECK(hipMemcpyAsync(h->bkeep, l.b - l.pitch - 1, len * sizeof(double), hipMemcpyDeviceToDevice, h->s));
    MgStats st{};
    int rc = precision == INCRMT_PREC_MIXED ? h->mg.solve_mixed(rel_tol, max_cycles, &st)
                                            : h->mg.solve(rel_tol, max_cycles, &st);
    if (rc && rc != 2) return rc;
    if (int r2 = mg_copy(h, l.z, nullptr, z)) return r2;
    if (cycles) *cycles = st.cycles;
    if (rel_res) *rel_res = st.rel_res;
    return rc == 2 ? fail(INCRMT_ERR_NOT_CONVERGED, "V-cycle failed to converge") : 0;
}

int incrmt_mg_time_solve(incrmt_mg *h, double rel_tol, int max_cycles, int precision, int reps,
                         double *ms_per_solve, int *cycles, double *rel_res) {
    if (!h || !(rel_tol > 0) || max_cycles < 1 || reps < 1) return fail(INCRMT_ERR_ARG, "bad arguments");
    if (!h->bkeep) return fail(INCRMT_ERR_ARG, "no right-hand side: solve once first");
    MgLevel &l = h->mg.fine();
    const size_t len = (size_t)(l.ny + 2) * l.pitch;
    hipEvent_t e0, e1;
    HIP_CHECK(hipEventCreate(&e0));
    HIP_CHECK(hipEventCreate(&e1));
    double total = 0;
    MgStats st{};
    int rc = 0;
    for (int r = 0; r < reps && !rc; r++) {
        // the right-hand side as uploaded and a zero start, then the solve
        // alone between the two events
        HIP_CHECK(hipMemcpyAsync(l.b - l.pitch - 1, h->bkeep, len * sizeof(double), hipMemcpyDeviceToDevice, h->s));
        HIP_CHECK(hipMemsetAsync(l.z - l.pitch - 1, 0, len * sizeof(double), h->s));
        HIP_CHECK(hipEventRecord(e0, h->s));
        rc = precision == INCRMT_PREC_MIXED ? h->mg.solve_mixed(rel_tol, max_cycles, &st)
                                            : h->mg.solve(rel_tol, max_cycles, &st);
        HIP_CHECK(hipEventRecord(e1, h->s));
        HIP_CHECK(hipEventSynchronize(e1));
        float ms = 0;
        HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
        total += ms;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc && rc != 2) return rc;
    if (ms_per_solve) *ms_per_solve = total / reps;
    if (cycles) *cycles = st.cycles;
    if (rel_res) *rel_res = st.rel_res;
    return rc == 2 ? fail(INCRMT_ERR_NOT_CONVERGED, "V-cycle failed to converge") : 0;
}

int incrmt_mg_set_krylov(incrmt_mg *h, int on) {
    if (!h) return fail(INCRMT_ERR_ARG, "null handle");
    h->mg.krylov = on != 0;
    return 0;
}

int incrmt_fluid_set_krylov(incrmt_fluid *f, int on) {
    if (!f) return fail(INCRMT_ERR_ARG, "null handle");
    f->e.mac.krylov = f->e.fem.krylov = on != 0;
    return 0;
}

int incrmt_mg_set_launch(incrmt_mg *h, int mode) {
    if (!h) return fail(INCRMT_ERR_ARG, "null handle");
    if (mode != INCRMT_LAUNCH_DIRECT && mode != INCRMT_LAUNCH_GRAPH) return fail(INCRMT_ERR_ARG, "bad launch mode");
    h->mg.use_graphs = mode == INCRMT_LAUNCH_GRAPH;
    return 0;
}

int incrmt_mg_smooth(incrmt_mg *h, const double *b, double *z, int sweeps, int shape) {
    return incrmt_mg_smooth_prec(h, b, z, sweeps, shape, INCRMT_PREC_FP64);
}

int incrmt_mg_smooth_prec(incrmt_mg *h, const double *b, double *z, int sweeps, int shape, int precision) {
    if (!h || !b || !z || sweeps < 1 || (precision != INCRMT_PREC_FP64 && precision != INCRMT_PREC_MIXED))
        return fail(INCRMT_ERR_ARG, "bad smooth arguments");
    MgLevel &l = h->mg.fine();
    if (int rc = mg_copy(h, l.b, b, nullptr)) return rc;
    if (int rc = mg_copy(h, l.z, z, nullptr)) return rc;
    if (int rc = h->mg.smooth_check(sweeps, shape, precision == INCRMT_PREC_MIXED))
        return rc == 1 ? fail(INCRMT_ERR_ARG, "unknown shape") : rc;
    return mg_copy(h, l.z, nullptr, z);
}

int incrmt_mg_apply(incrmt_mg *h, const double *x, double *y) {
    if (!h || !x || !y) return fail(INCRMT_ERR_ARG, "null argument");
    const MgLevel &l = h->mg.fine();
    double *xd = h->x + l.pitch + 1, *yd = h->y + l.pitch + 1;
    if (int rc = mg_copy(h, xd, x, nullptr)) return rc;
    if (int rc = h->mg.apply(xd, yd)) return rc;
    return mg_copy(h, yd, nullptr, y);
}

int incrmt_mg_time_smoother(incrmt_mg *h, int sweeps, int precision, double *ms_per_launch, int *launches) {
    return incrmt_mg_time_shape(h, sweeps, precision, -1, ms_per_launch, launches);
}

int incrmt_mg_time_shape(incrmt_mg *h, int sweeps, int precision, int shape, double *ms_per_launch, int *launches) {
    if (!h || sweeps < 1) return fail(INCRMT_ERR_ARG, "bad arguments");
    long n = 0;
    int rc = h->mg.time_fine_smoother(sweeps, precision == INCRMT_PREC_MIXED, ms_per_launch, &n, shape);
    if (rc == 1) return fail(INCRMT_ERR_ARG, "unknown shape");
    if (launches) *launches = (int)n;
    return rc;
}

void incrmt_mg_destroy(incrmt_mg *h) { delete h; }

// ------------------------------------------------------- standalone level set
namespace {
struct LsStandalone {
    Grid g{};
    ObjConst c{};
    ObjArrays a{};
    LsWork w{};
    hipStream_t s = nullptr;
    std::vector<void *> allocs;
    ~LsStandalone() {
        ls_free(w);
        for (void *p : allocs) (void)hipFree(p);
        if (s) (void)hipStreamDestroy(s);
    }
    // an m x n level-set grid is the ghosted grid of (m-4) x (n-4) cells
    int init(int m, int n, double ax, double ay, double dx, double dy, double pd, double pu) {
        g.m = m - 4; g.n = n - 4; g.ml = m; g.nl = n;
        g.P = round_up(m, 32);
        g.N = (long long)kOrigin + (long long)n * g.P + 64;
        g.dx = dx; g.dy = dy; g.ax = ax + 1.5 * dx; g.ay = ay + 1.5 * dy;
        g.lax = ax; g.lay = ay;
        g.bx = g.ax + g.m * dx; g.by = g.ay + g.n * dy;
        g.xsp = 1 / dx; g.ysp = 1 / dy; g.xxsp = g.xsp * g.xsp; g.yysp = g.ysp * g.ysp;
        c.phi_down = pd; c.phi_up = pu;
        HIP_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        double **al[] = {&a.phi, &a.nX, &a.nY};
        for (auto p : al) {
            void *q;
            HIP_CHECK(hipMalloc(&q, g.N * sizeof(double)));
            HIP_CHECK(memset_sync(q, 0, g.N * sizeof(double), s));
            allocs.push_back(q);
            *p = (double *)q;
        }
        void *q;
        HIP_CHECK(hipMalloc(&q, g.N * sizeof(int)));
        HIP_CHECK(memset_sync(q, 0, g.N * sizeof(int), s));
        allocs.push_back(q);
        a.st = (int *)q;
        return ls_alloc(g, w, s);
    }
    int put(double *dev, const double *host) {
        HIP_CHECK(hipMemcpy2D(dev + g.ls(0, 0), g.P * sizeof(double), host, g.ml * sizeof(double),
                              g.ml * sizeof(double), g.nl, hipMemcpyHostToDevice));
        return 0;
    }
    int get(const double *dev, double *host) {
        HIP_CHECK(hipMemcpy2D(host, g.ml * sizeof(double), dev + g.ls(0, 0), g.P * sizeof(double),
                              g.ml * sizeof(double), g.nl, hipMemcpyDeviceToHost));
        return 0;
    }
    int get_status(int *host) {
        HIP_CHECK(hipMemcpy2D(host, g.ml * sizeof(int), a.st + g.ls(0, 0), g.P * sizeof(int), g.ml * sizeof(int),
                              g.nl, hipMemcpyDeviceToHost));
        return 0;
    }
};
int ls_args_ok(int m, int n, double dx, double dy, double pd, double pu) {
    if (m < 8 || n < 8 || !(dx > 0) || !(dy > 0) || !(pd < 0) || !(pu > 0))
        return fail(INCRMT_ERR_ARG, "bad level-set arguments");
    return 0;
}

// ---- the band rebuild without allocation ---------------------------------
// Everything the rebuild needs beyond the caller's phi and status arrays
// (the padded working copies, the fixed-point buffers, the point and tile
// lists, the counters) lives in `scratch`, one file-scope device array sized
// for level-set grids of up to kNoAllocSide x kNoAllocSide points (12 MiB).
// Each GPU holds its own copy until the process ends. Host code never uses
// the array's name as an address: the address hipGetSymbolAddress returns on
// each device is looked up once per device and kept in g_scratch_dev.
constexpr int kNoAllocSide = 520;
constexpr long long kNoAllocPitch = (kNoAllocSide + 31) / 32 * 32;
constexpr long long kNoAllocN = kOrigin + kNoAllocSide * kNoAllocPitch + 64;
constexpr long long kNoAllocL = (long long)kNoAllocSide * kNoAllocSide + 64;
constexpr long long kNoAllocT =
    (long long)((kNoAllocSide + kTileX - 1) / kTileX) * ((kNoAllocSide + kTileY - 1) / kTileY) + 64;
constexpr size_t a256(long long b) { return (size_t)((b + 255) / 256 * 256); }
constexpr size_t kNoAllocBytes = 3 * a256(kNoAllocN * 8) + a256(kNoAllocN * 4) + 2 * a256(kNoAllocN) +
                                 3 * a256(kNoAllocL * 4) + 3 * a256(kNoAllocT * 4) +
                                 a256((kCntIter + kIterSlots) * 4) + a256(16);
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

constexpr int kMaxDevices = 64;
std::atomic<unsigned char *> g_scratch_dev[kMaxDevices];
// host mirror of the rebuild's counters (shared by every call in the process)
int g_noalloc_hcnt[kCntIter + kIterSlots];

int scratch_address(unsigned char **out) {
    int dev = 0;
    HIP_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(INCRMT_ERR_ARG, "device index beyond the scratch table");
    unsigned char *p = g_scratch_dev[dev].load();
    if (!p) {
        void *q = nullptr;
        const hipError_t e = hipGetSymbolAddress(&q, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)", __FILE__, __LINE__);
        p = (unsigned char *)q;
        g_scratch_dev[dev].store(p);
    }
    *out = p;
    return 0;
}

__global__ void k_fill_words(unsigned *p, long long n, unsigned v) {
    for (long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x; k < n; k += (long long)gridDim.x * blockDim.x)
        p[k] = v;
}

int device_array(const void *p) {
    hipPointerAttribute_t at{};
    if (hipPointerGetAttributes(&at, p) != hipSuccess || at.type != hipMemoryTypeDevice) {
        (void)hipGetLastError();  // a host pointer leaves an error behind: clear it
        return fail(INCRMT_ERR_ARG, "phi and s must be device arrays");
    }
    return 0;
}
}  // namespace

int incrmt_levelset_build_band(int m, int n, double ax, double ay, double dx, double dy, double phi_down,
                               double phi_up, double *phi, int *s, int *passes) {
    if (int rc = ls_args_ok(m, n, dx, dy, phi_down, phi_up)) return rc;
    if (!phi || !s) return fail(INCRMT_ERR_ARG, "null array");
    LsStandalone L;
    if (int rc = L.init(m, n, ax, ay, dx, dy, phi_down, phi_up)) return rc;
    if (int rc = L.put(L.a.phi, phi)) return rc;
    int p = 0;
    int rc = ls_rebuild(L.s, L.g, L.c, L.a, L.w, &p);
    if (rc == 3) return fail(INCRMT_ERR_NEWTON, "Newton iteration failed.");
    if (rc) return rc;
    HIP_CHECK(hipStreamSynchronize(L.s));
    if (passes) *passes = p;
    if (int r2 = L.get(L.a.phi, phi)) return r2;
    return L.get_status(s);
}

int incrmt_levelset_build_band_noalloc(int m, int n, double ax, double ay, double dx, double dy, double phi_down,
                                       double phi_up, double *phi, int *s, int *passes) {
    if (int rc = ls_args_ok(m, n, dx, dy, phi_down, phi_up)) return rc;
    if (!phi || !s) return fail(INCRMT_ERR_ARG, "null array");
    if (m > kNoAllocSide || n > kNoAllocSide)
        return fail(INCRMT_ERR_ARG, "level-set grid larger than the no-allocation limit of 520 x 520 points");
    if (int rc = device_array(phi)) return rc;
    if (int rc = device_array(s)) return rc;
    unsigned char *base = nullptr;
    if (int rc = scratch_address(&base)) return rc;
    // the grid as incrmt_levelset_build_band lays it out
    Grid g{};
    g.m = m - 4; g.n = n - 4; g.ml = m; g.nl = n;
    g.P = round_up(m, 32);
    g.N = (long long)kOrigin + (long long)n * g.P + 64;
    g.dx = dx; g.dy = dy; g.ax = ax + 1.5 * dx; g.ay = ay + 1.5 * dy;
    g.lax = ax; g.lay = ay;
    g.bx = g.ax + g.m * dx; g.by = g.ay + g.n * dy;
    g.xsp = 1 / dx; g.ysp = 1 / dy; g.xxsp = g.xsp * g.xsp; g.yysp = g.ysp * g.ysp;
    ObjConst c{};
    c.phi_down = phi_down; c.phi_up = phi_up;
    ObjArrays a{};
    LsWork w{};
    w.ntx = (g.ml + kTileX - 1) / kTileX;
    w.nty = (g.nl + kTileY - 1) / kTileY;
    const long long N = g.N, L = (long long)m * n + 64, T = (long long)w.ntx * w.nty + 64;
    // carve the scratch (every piece 256-B aligned); sizes are within the
    // maxima the array was sized for, since m, n <= kNoAllocSide
    size_t off = 0;
    auto take = [&](long long bytes) { unsigned char *p = base + off; off += a256(bytes); return p; };
    a.phi = (double *)take(N * 8);
    w.va = (double *)take(N * 8);
    w.vb = (double *)take(N * 8);
    a.st = (int *)take(N * 4);
    w.mark = take(N);
    w.cflag = take(N);
    w.iface = (int *)take(L * 4);
    w.cand = (int *)take(L * 4);
    w.pband = (int *)take(L * 4);
    w.tflag = (int *)take(T * 4);
    w.tiles_fim = (int *)take(T * 4);
    w.tiles_ext = (int *)take(T * 4);
    w.cnt = (int *)take((kCntIter + kIterSlots) * 4);
    w.bbox = (int *)take(16);
    if (off > kNoAllocBytes) return fail(INCRMT_ERR_ARG, "scratch layout exceeds the array");
    w.hcnt = g_noalloc_hcnt;  // no window is kept: no host box, no event
    hipStream_t st = nullptr;
    // the scratch keeps the previous call's contents: clear what the rebuild
    // expects to start at zero (with a kernel, not a memset node)
    auto clear = [&](void *p, long long bytes) -> int {
        k_fill_words<<<256, 256, 0, st>>>((unsigned *)p, (long long)(a256(bytes) / 4), 0u);
        HIP_CHECK(hipGetLastError());
        return 0;
    };
    int rc = 0;
    if ((rc = clear(a.phi, N * 8)) || (rc = clear(w.va, N * 8)) || (rc = clear(w.vb, N * 8)) ||
        (rc = clear(a.st, N * 4)) || (rc = clear(w.mark, N)) || (rc = clear(w.cflag, N)) || (rc = clear(w.tflag, T * 4)))
        return rc;
    HIP_CHECK(hipMemcpy2DAsync(a.phi + g.ls(0, 0), g.P * sizeof(double), phi, m * sizeof(double), m * sizeof(double), n,
                               hipMemcpyDeviceToDevice, st));
    int p = 0;
    rc = ls_rebuild(st, g, c, a, w, &p);
    if (rc == 3) return fail(INCRMT_ERR_NEWTON, "Newton iteration failed.");
    if (rc) return rc;
    HIP_CHECK(hipMemcpy2DAsync(phi, m * sizeof(double), a.phi + g.ls(0, 0), g.P * sizeof(double), m * sizeof(double), n,
                               hipMemcpyDeviceToDevice, st));
    HIP_CHECK(hipMemcpy2DAsync(s, m * sizeof(int), a.st + g.ls(0, 0), g.P * sizeof(int), m * sizeof(int), n,
                               hipMemcpyDeviceToDevice, st));
    HIP_CHECK(hipStreamSynchronize(st));
    if (passes) *passes = p;
    return 0;
}

int incrmt_levelset_extrapolate(int m, int n, double ax, double ay, double dx, double dy, double phi_down,
                                double phi_up, double *phi, int *s, double *fx, double *fy, int *passes) {
    if (int rc = ls_args_ok(m, n, dx, dy, phi_down, phi_up)) return rc;
    if (!phi || !s || !fx || !fy) return fail(INCRMT_ERR_ARG, "null array");
    LsStandalone L;
    if (int rc = L.init(m, n, ax, ay, dx, dy, phi_down, phi_up)) return rc;
    if (int rc = L.put(L.a.phi, phi)) return rc;
    if (int rc = L.put(L.a.nX, fx)) return rc;
    if (int rc = L.put(L.a.nY, fy)) return rc;
    int p = 0;
    int rc = ls_rebuild(L.s, L.g, L.c, L.a, L.w, &p);
    if (rc == 3) return fail(INCRMT_ERR_NEWTON, "Newton iteration failed.");
    if (rc) return rc;
    rc = ls_extend(L.s, L.g, L.a, L.w, &p);
    if (rc == 5) return fail(INCRMT_ERR_SINGULAR_FIT, "singular extension fit");
    if (rc) return rc;
    HIP_CHECK(hipStreamSynchronize(L.s));
    if (passes) *passes = p;
    if (int r2 = L.get(L.a.phi, phi)) return r2;
    if (int r2 = L.get(L.a.nX, fx)) return r2;
    if (int r2 = L.get(L.a.nY, fy)) return r2;
    return L.get_status(s);
}

}  // extern "C"
