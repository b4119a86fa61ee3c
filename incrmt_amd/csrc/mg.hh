// Geometric multigrid for the two projection systems of the step:
//   MAC: cell-centred 5-point  div((1/rho) grad q) (mgs_mac.hh:14-102)
//   FEM: vertex-centred 9-point bilinear-FEM stiffness with cell 1/rho
//        (mgs_fem.hh:15-136)
// on pure-Neumann boxes, periodic in x and/or y or not, or with the
// zero-pressure right edge of an outflow (zpr, mgs_common.hh:33). The
// reference's engine (TGMG) is an external library absent from the reference tree; this solver honours its
// contract (solve A z = b in place, z warm-started from the previous solve)
// with its own algorithm: colored Gauss-Seidel V-cycles over rediscretized
// coarse levels, stopped on the relative residual ||b - A z||_2 / ||b||_2.
#ifndef INCRMT_MG_HH
#define INCRMT_MG_HH

#include <cmath>
#include <cstdint>
#include <functional>
#include <string>
#include <vector>
#include "common.hh"

#ifndef INCRMT_MG_NU1
#define INCRMT_MG_NU1 2
#endif
#ifndef INCRMT_MG_NU2
#define INCRMT_MG_NU2 2
#endif

namespace incrmt {

class UNet;

enum MgKind { kMAC = 0, kFEM = 1 };

// When a learned start is used and when it is given up (host logic): a start
// is accepted when the relative residual after its first V-cycle is finite
// and at most the tolerance. After kMaxMisses rejections in a row the
// learned start is switched off for kRetry solves, then probed once: one
// accepted start switches it back on, one more rejection off again.
struct GuessPolicy {
    static constexpr int kMaxMisses = 8, kRetry = 32;
    int misses = 0, cooldown = 0;
    bool off = false;
    // whether this solve may use the learned start (counts the solves off)
    bool use() {
        if (!off) return true;
        if (--cooldown > 0) return false;
        off = false;
        misses = kMaxMisses - 1;  // a probe: one more miss switches it off again
        return true;
    }
    bool accept(double rel, double tol) {
        const bool ok = std::isfinite(rel) && rel <= tol;
        misses = ok ? 0 : misses + 1;
        if (misses >= kMaxMisses) { off = true; cooldown = kRetry; }
        return ok;
    }
    void reset() {
        misses = 0; cooldown = 0; off = false;
        ema_learn = ema_quad = -1; nlearn = nquad = since_control = 0; benefit_off = false; benefit_cool = 0;
    }

    // Cost-aware use (when enabled): the learned start pays when the
    // V-cycles it saves against the quadratic start exceed its own cost in
    // V-cycles (the network's forward pass timed against one V-cycle). One
    // solve in kControlEvery runs the quadratic start as a control; after
    // kMinLearn learned and kMinQuad control solves, a learned start whose
    // mean V-cycles (exponential average) are not below the controls' by
    // more than its cost is switched off for kBenefitRetry solves, then
    // probed again from fresh averages.
    static constexpr int kControlEvery = 4, kMinLearn = 3, kMinQuad = 1, kBenefitRetry = 256;
    static constexpr double kEma = 0.25;
    bool cost_aware = false;
    double cost = -1;  // the learned start's cost in V-cycles (< 0: not measured yet)
    double ema_learn = -1, ema_quad = -1;
    int nlearn = 0, nquad = 0, since_control = 0, benefit_cool = 0;
    bool benefit_off = false;
    // whether this solve is a control (quadratic start) or off for cost
    bool skip_for_cost() {
        if (!cost_aware) return false;
        if (benefit_off) {
            if (--benefit_cool > 0) return true;
            benefit_off = false;
            ema_learn = ema_quad = -1;
            nlearn = nquad = since_control = 0;
        }
        if (++since_control >= kControlEvery) {
            since_control = 0;
            return true;
        }
        return false;
    }
    // V-cycles of a solve that could have used the learned start (learned:
    // whether it did)
    void record(bool learned, int cycles) {
        if (!cost_aware || benefit_off) return;
        double &e = learned ? ema_learn : ema_quad;
        e = e < 0 ? cycles : (1 - kEma) * e + kEma * cycles;
        (learned ? nlearn : nquad)++;
        if (nlearn >= kMinLearn && nquad >= kMinQuad && cost >= 0 && ema_quad - ema_learn <= cost) {
            benefit_off = true;
            benefit_cool = kBenefitRetry;
        }
    }
};

// One level: nx x ny unknowns stored with a one-point zero ghost ring, so
// stencils read neighbours without branches. Element (i,j) lives at
// base[(j+1)*pitch + i+1]; the z/b/r pointers below point at (0,0).
struct MgLevel {
    int nx, ny, pitch;
    double *z, *b, *r;
    float *zf, *bf, *rf;  // single-precision copies for the mixed cycle
    double *z2;           // second iterate buffer: the fused smoother reads one and writes the other
    float *zf2;
    // MAC: face inverse densities, zero on walls: fw(i,j) face between cells
    // (i-1,j),(i,j), i in [0,nx]; fs(i,j) face between (i,j-1),(i,j).
    // FEM: kc(i,j) = 1/rho of cell (i,j), i in [0,nx-2], zero ring around.
    double *fw, *fs, *kc;
    float *fwf, *fsf, *kcf;
    int cp;       // coefficient pitch
    double kx, ky;  // MAC 1/dx^2, 1/dy^2 of this level
    double hx, hy;  // FEM node spacing of this level
    int shape64 = 0, shape32 = 0;  // fused-smoother launch shape (FP64, FP32 sweeps)
    bool odd = false;              // odd size along a periodic direction: nine-colour sweeps
    // periodic directions: the ghost ring of every iterate holds the
    // opposite edge (each write of an edge point also writes its mirrors),
    // MAC faces 0 and nx are the two sides of the seam face, FEM has nx
    // cells with the ring cell mirroring cell nx-1
    int px = 0, py = 0;
    // zero-pressure right edge (the reference's zpr): the point column nx
    // past the last unknown is held at zero (the ring column). MAC: face nx
    // carries twice its inverse density (mgs_mac.hh:50, 95); FEM: the fixed
    // node column of mgs_fem_base::pfix is that ring column, so the level
    // has nx cells per row. The system is then definite, not singular.
    int dr = 0;
};

// Transfer tables between a FEM level and a uniform coarse level whose
// nodes do not coincide with every other fine node (node counts that are
// not 2^k+1). Per direction: fine node i lies in coarse interval i0[i] at
// fraction t[i]; coarse node I gathers fine nodes lo[I]..hi[I]; coarse cell
// I overlaps fine cells clo[I]..chi[I].
struct Resample {
    int *i0, *lo, *hi, *clo, *chi;
    double *t;
    int nf, nc;
    double hf, hc;
};

struct MgStats {
    int cycles;
    double rel_res, bnorm, rnorm;
};

class Multigrid {
  public:
    Multigrid() {}
    ~Multigrid();
    // Allocate a hierarchy for an m x n system (cells for MAC, nodes for FEM),
    // periodic in x and/or y (mgs_mac.hh:45-56, mgs_fem.hh:47-52: a periodic
    // FEM direction has m nodes, node m being node 0). dr: zero-pressure
    // right edge (x not periodic); for FEM, m counts the unknown nodes, the
    // fixed column not included.
    int init(MgKind kind, int m, int n, double dx, double dy, hipStream_t s, bool px = false, bool py = false,
             bool dr = false);
    // Refresh the periodic ghost ring of a level-0 array (after a host upload).
    int wrap_fine(double *z);
    // Fine-level coefficients from the ghosted cell densities (fluid_2d.cc:
    // 1140-1170 build the same values), then coarsen.
    int set_density(const Grid &g, const double *rho);
    // Fine-level coefficients from the reference's irho layouts on the host
    // (mgs_mac_varying_rho::irho 2mn+n+m; mgs_fem_varying_rho::irho m*n, m
    // counting the fixed column with dr), or constant density when irho is
    // null.
    int set_irho_host(const double *irho);
    // Solve with the right-hand side in level 0's b and the warm start in z.
    int solve(double tol, int max_cycles, MgStats *st);
    // Mixed-precision solve: FP64 residual and correction accumulation,
    // FP32 V-cycles (iterative refinement); with krylov set, conjugate
    // gradients in FP64 preconditioned by one FP32 V-cycle per iteration
    // (on walled levels without a zero-pressure edge; elsewhere the
    // refinement).
    int solve_mixed(double tol, int max_cycles, MgStats *st);
    bool krylov = false;
    // Launch `sweeps` fine-level smoothing sweeps (for timing the smoother).
    int smooth_fine(int sweeps, bool fp32);
    // `sweeps` forward sweeps on level 0 (FP64, or FP32 on the rounded
    // iterate with f32) with fused launch shape `shape`, or with one launch
    // per colour when shape < 0 (for checking that every form gives the same
    // sweep)
    int smooth_check(int sweeps, int shape, bool f32 = false);
    // y = A x on level 0; x and y are padded level arrays (pointers at (0,0)).
    int apply(const double *x, double *y);
    // Time `sweeps` fine-level sweeps with HIP events: mean ms per launch.
    // shape >= 0: time that launch shape instead of the level's own
    int time_fine_smoother(int sweeps, bool fp32, double *ms_per_launch, long *launches, int shape = -1);
    MgLevel &fine() { return L[0]; }
    const std::vector<MgLevel> &levels() const { return L; }
    MgKind kind() const { return kind_; }
    // a pure-Neumann system (constants in the null space): right-hand sides
    // and coarsest solutions are made mean-free
    bool singular() const { return !dr_; }
    // pre- and post-smoothing sweeps per level and cycle (a build parameter)
    int nu1 = INCRMT_MG_NU1, nu2 = INCRMT_MG_NU2;
    size_t tail_level() const { return tail_; }
    // timing of fine-level smoothing launches (HIP events on the stream)
    bool time_smoother = false;
    // start each solve from the linear extrapolation in time of the last two
    // solutions (mgs_base::initial_guess with mg_lin_t_interp), the solve's
    // time being t_solve (the step's projections set both)
    // 0: previous solution; 1: linear (mg_lin_t_interp); 2: quadratic in
    // time through the last three solutions (linear while two exist)
    int extrapolate = 0;
    double t_solve = 0;
    // change the starting-iterate mode; a change clears the solution history
    void set_extrapolate(int mode);
    // training loss of the last training step (waits for the stream)
    int read_train_loss(double *loss);
    // Learned start (extrapolate == 3): the quadratic start plus alpha d,
    // where d is the network's prediction from the starting residual, the
    // feature channel and the source, and alpha = r.d / d.Ad the step along
    // d that minimises the error's energy norm (computable from the
    // residual r alone). If the relative residual after the first
    // V-cycle exceeds guess_tol (or is not finite) the solve restarts from
    // the quadratic start with FP64 V-cycles (the fallback). With train set,
    // every solve that converges also takes one Adam step on its (input,
    // starting residual) pair: the loss is the relative energy norm of the
    // error left after that step, the error being the solution minus the
    // quadratic start.
    UNet *net = nullptr;
    const double *feat = nullptr;  // level-0 array: the level-set channel (null: zero)
    double guess_tol = 1e-4;
    bool train = false;
    float lr = 1e-3f;
    long learned_solves = 0, guess_fallbacks = 0, train_steps = 0;
    double last_rel0 = 0, last_rel1 = 0;  // relative residual of the last learned start, after its first cycle
    GuessPolicy policy;
    hipEvent_t cev_[3] = {nullptr, nullptr, nullptr};  // the learned start's cost measurement
    // replay captured V-cycles (HIP graphs) instead of launching them; init
    // sets it for hierarchies whose finest level is below kGraphMaxPoints
    bool use_graphs = true;
    static constexpr long long kGraphMaxPoints = 512LL * 512;
    double smoother_ms = 0;
    long smoother_launches = 0;
    double smoother_bytes = 0;  // algorithmic bytes of the timed sweeps
    double *d_red = nullptr;   // reduction partials
    double *h_red = nullptr;   // pinned host mirror
  private:
    int vcycle(size_t l);
    int vcycle_f(size_t l);
    int coarsen_coefs(bool f32 = false);
    int reduce_norms(double *out3);  // ||b||^2, ||r||^2, sum b of level 0
    int reduce_dev();                // the same into d_red, no host copy
    template <class Cycle>
    int solve_loop(double tol, int max_cycles, MgStats *st, Cycle cycle);
    int predict_ = 0;                // V-cycles run before the first residual check
    double *zprev_ = nullptr;        // the solution before the last one (extrapolation)
    bool f32_stale_ = true;          // coefficient copies for the FP32 cycle out of date
    bool want_f32_ = false;          // a mixed solve has run: coefficient updates write FP32 copies too
    int make_f32_coefs();
    int wrap_coefs(size_t l, bool f32 = false);
    hipGraphExec_t graph_exec_[3] = {nullptr, nullptr, nullptr};  // FP64 cycle, mixed cycle, Krylov iteration
    std::vector<uintptr_t> graph_key_[3];                          // launch state each graph was captured in
    // Krylov solve: residual, direction, preconditioned residual, A times
    // the direction (FP64 level-0 arrays) and its device scalars
    double *kr_ = nullptr, *kp_ = nullptr, *kc_ = nullptr, *kq_ = nullptr, *ks_ = nullptr;
    int krylov_buffers();
    int solve_krylov(double tol, int max_cycles, MgStats *st);
    std::vector<uintptr_t> launch_state() const;
    void restore_buffers(const std::vector<uintptr_t> &key);
    int run_cycle(int which, const std::function<int()> &cycle);
    double *zs_ = nullptr, *dz_ = nullptr;  // learned start: the quadratic start, the network's direction
    double *eprev_ = nullptr;               // the last solve's correction to its quadratic start
    int guess_buffers();
    float *tgt_ = nullptr;                  // training: d loss / d prediction (dense)
    double *d_gs_ = nullptr;                // r.d, d.Ad, |r|^2 or e.r, training loss; |eprev|^2
    int learned_start(bool apply);
    int train_step();
    int scount_ = 0;                 // solves recorded for the extrapolation (0, 1, 2)
    int nq_ = 0;                     // solutions known to the quadratic start (0..3)
    double tq_[3] = {0, 0, 0};       // their times, newest first
    double *zq1_ = nullptr, *zq2_ = nullptr;  // the second and third newest (the newest is z)
    double t0_ = 0, tprev_ = 0;      // times of the last two solves
    bool resampled_ = false;         // level 1 is a uniform resampling of level 0
    bool rs_per_ = false;            // ... of a doubly periodic level 0
    Resample rx_{}, ry_{};
    int make_resample(int nf, double hf, int nc, Resample &r);
    // periodic: nf and nc points (cell centres, off = 0.5, or nodes, off = 0)
    // spread evenly over the same period
    int make_resample_per(int nf, double hf, int nc, double off, Resample &r);
    template <class T>
    int smooth_colours(size_t l, int sweeps, bool reverse);
    // restriction weight of a periodic resampling: FEM the transpose of the
    // prolongation, MAC its average (nc/nf per direction)
    double rs_scale() const {
        return kind_ == kMAC ? (double)rx_.nc / rx_.nf * ((double)ry_.nc / ry_.nf) : 1.0;
    }
  public:
    // Time every fused-smoother launch shape on each level large enough to
    // matter and keep the fastest (results do not depend on the shape).
    int autotune();
    bool tuned = false;
    // "nx x ny: SXxSY hH / SXxSY hH" per fused level (FP64 / FP32 shapes)
    std::string shapes() const;
  private:
    // correct: add level l+1's correction (prolonged) as the sweep loads z
    int smooth(size_t l, int sweeps, bool reverse, bool correct = false);
    int smooth_f(size_t l, int sweeps, bool reverse, bool correct = false);
    MgKind kind_ = kMAC;
    bool dr_ = false;
    std::vector<MgLevel> L;
    size_t tail_ = 0;  // first level handled by the single-workgroup tail cycle
    std::vector<void *> allocs;
    hipStream_t s_ = nullptr;
    double dx_ = 1, dy_ = 1;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev_;
    size_t ev_used_ = 0;
    int record_begin();
    int record_end();
    hipEvent_t tev_[2] = {nullptr, nullptr};  // the timing pair of the next timed smoother call
    int collect_events();
};

}  // namespace incrmt

#endif
