/*
 * incrmt.h -- C ABI of the MI355X IncRMT hot path (libincrmt_hip.so).
 *
 * The reference (chr1shr/incrmt) is a C++ library with no FFI of its own:
 * its drivers call fluid_2d and the mgs_* projection solvers directly. These
 * entry points are what a C-ABI binding of exactly those calls needs; each
 * cites the reference interface it replaces. Device state is resident on the
 * GPU between calls; host arrays cross only in the upload/download calls.
 *
 * Conventions: every function returns an int status (INCRMT_OK on success),
 * never exits the process (the reference exits on solver failure: callers
 * that want that behaviour test the status). Handles are single-threaded:
 * one HIP stream per handle, any number of handles per process.
 */
#ifndef INCRMT_H
#define INCRMT_H

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define INCRMT_OK 0
#define INCRMT_ERR_ARG 1           /* bad argument or call order */
#define INCRMT_ERR_NOT_CONVERGED 2 /* "V-cycle failed to converge" (mgs_mac.cc:67-70, mgs_fem.cc:116-119) */
#define INCRMT_ERR_NEWTON 3        /* "Newton iteration failed." (levelset.cc:275,282) */
#define INCRMT_ERR_BAND 4          /* narrow band overflow (levelset.cc:509) */
#define INCRMT_ERR_SINGULAR_FIT 5  /* singular extension fit (traverse.hh:211-222) */
#define INCRMT_ERR_HIP 6           /* HIP runtime error; see incrmt_last_error() */
#define INCRMT_ERR_UNSUPPORTED 7   /* option outside the implemented path */

/* projection solver kinds (mgs_mac_* and mgs_fem_*) */
#define INCRMT_MG_MAC 0
#define INCRMT_MG_FEM 1

/* solver arithmetic */
#define INCRMT_PREC_FP64 0  /* FP64 V-cycles */
#define INCRMT_PREC_MIXED 1 /* FP32 V-cycles inside FP64 iterative refinement */

typedef struct incrmt_fluid incrmt_fluid;
typedef struct incrmt_mg incrmt_mg;

typedef struct incrmt_stats {
    int mac_cycles, fem_cycles;       /* V-cycles of the last step's solves */
    double mac_rel_res, fem_rel_res;  /* their final ||b-Az||/||b|| */
    long total_mac_cycles, total_fem_cycles, steps;
    int reinit_passes, extend_passes; /* fixed-point passes of the last step */
    double smoother_ms;               /* summed fine-level smoother time (timing on) */
    long smoother_launches;           /* fine-level smoother kernel launches timed */
    double smoother_bytes;            /* their algorithmic bytes (z, b, coefficients read; z written) */
    double proj_ms;                   /* summed time of the two projections per step (timing on) */
} incrmt_stats;

/* Last error message of the calling thread. */
const char *incrmt_last_error(void);
/* Device name and HBM size of device 0, for reports. */
int incrmt_device_info(char *name, int len, double *hbm_gib);
/* No reference counterpart: `reps` device-to-device copies of `bytes` bytes
 * timed with HIP events (a copy of known size, for calibrating memory
 * counters and for the achievable-bandwidth figure). */
int incrmt_copy_probe(long long bytes, int reps, double *ms_per_copy);

/* ---- fluid_2d ------------------------------------------------------------
 * fluid_2d::fluid_2d(m,n,sim_type&,visc,rhof,tw,pinw,ntrace,fflags,dir)
 * (fluid_2d.cc:26-43) with sim_type(ax,bx,ay,by,x_prd,y_prd,w_scale,w_force)
 * (sim_type.hh:63-71): no-slip walls, or periodic along x_prd / y_prd
 * (fluid_2d.cc:689-872, 1193-1235, 1317-1348). A box periodic in one
 * direction only needs an even cell count there whose halvings end at 2 or
 * 3; a doubly periodic box takes any size of at least 8 (conv_test's ladder
 * 280..2520 included): when halving does not end on a tiny level the
 * solvers resample to a 2^k or 3 2^k grid (checked at initialize).
 * The outflow (zpr) and periodic-solid (solid_prd_bc) variants are set by
 * incrmt_fluid_set_horizontal_flow and incrmt_fluid_set_solid_periodic. */
int incrmt_fluid_create(int m, int n, double ax, double bx, double ay, double by, int x_prd, int y_prd,
                        double w_scale, double w_force, double visc, double rhof, double tw, double pinw,
                        incrmt_fluid **out);
/* fluid_2d::add_object(new obj_simple_spin(r_scale,a1,a2,piv_r,th_max,omega),
 * mat_const(G,rhos,ev_trans_mult)) (fluid_2d.cc:74-90, object.hh:449-455). */
int incrmt_fluid_add_spin(incrmt_fluid *f, double r_scale, double a1, double a2, double piv_r, double th_max,
                          double omega, double G, double rhos, double ev_trans_mult);
/* fluid_2d::add_object(new obj_circle(cx,cy,cr,grav), mat_const) (object.hh:128-138). */
int incrmt_fluid_add_circle(incrmt_fluid *f, double cx, double cy, double cr, double grav, double G, double rhos,
                            double ev_trans_mult);
/* fluid_2d::add_object(new obj_square(cx,cy,cr,grav), mat_const) (object.hh:140-150,
 * object.cc:41-44, 83-87): the square of half-width cr, at rest. */
int incrmt_fluid_add_square(incrmt_fluid *f, double cx, double cy, double cr, double grav, double G, double rhos,
                            double ev_trans_mult);
/* fluid_2d::add_object(new obj_full(), mat_const) (object.hh:83-86): a
 * solid filling the whole box (phi = -1), as conv_test type 2 and ftest's
 * solid-only case use with solid periodic boundaries. */
int incrmt_fluid_add_full(incrmt_fluid *f, double G, double rhos, double ev_trans_mult);
/* fluid_2d::solid_prd_bc (fluid_2d.hh:124; conv_test.cc:120, ftest.cc:233):
 * every solid's reference map is periodic up to the box lengths in its
 * ghost layers (set_boundaries_solid, fluid_2d.cc:818-835) and across the
 * edge upwinding (edge_boundary_conditions_solid_pre/_post,
 * fluid_2d.cc:784-815), and solids feel neither wall forces nor their own
 * accelerations (fluid_2d.cc:1037). Needs x_prd = y_prd = 1. Before
 * initialize. */
int incrmt_fluid_set_solid_periodic(incrmt_fluid *f, int on);
/* mat_const::set_velocity (obj_field.hh:25-28, obj_field.cc:62-75) of solid
 * obj (in the order added): 1 (the default) blends the solid's own initial
 * velocity into init_fields, 0 leaves the sim_type's velocity there
 * (conv_test constructs its solids with 0, conv_test.cc:108). Before
 * initialize. */
int incrmt_fluid_set_solid_velocity(incrmt_fluid *f, int obj, int set_velocity);
/* fluid_2d::add_object(new obj_flapper(cx,cy,cr,cl,ar,al,theta,m_stretch,T),
 * mat_const) (object.hh:245-289, object.cc:284-308): a rounded rod of
 * half-length cl and cap radius cr, rotated by theta, whose actuated region
 * (half-length al, radius ar) stretches its deformation gradient by
 * exp(con Y H(phi)) with con = -log(m_stretch)/ar sin^12(2 pi t / T)
 * (pre_stress_setup at each stress evaluation). */
int incrmt_fluid_add_flapper(incrmt_fluid *f, double cx, double cy, double cr, double cl, double ar, double al,
                             double theta, double m_stretch, double T, double G, double rhos, double ev_trans_mult);
/* fluid_2d::add_object(new obj_lamprey(cx,cy,theta,le,dis,w1,w2,T,mlam),
 * mat_const) (object.hh:542-595, object.cc:554-580): the lamprey of Tytell
 * et al. (head length dis, body length le-dis, half-widths w1, w2), rotated
 * by theta, swimming by a contraction wave of period T and maximum stretch
 * mlam ramped in over the first half period. */
int incrmt_fluid_add_lamprey(incrmt_fluid *f, double cx, double cy, double theta, double le, double dis, double w1,
                             double w2, double T, double mlam, double G, double rhos, double ev_trans_mult);
/* The sim_type's initial fluid velocity, which init_fields blends with the
 * solids' momenta (fluid_2d.cc:210-238): at rest (sim_type.hh:67-70, the
 * default), sim_four_rolls (sim_type.cc:30-35) or sim_velocity_pulses
 * (sim_type.cc:43-47, sim_type.hh:99-114; conv_test's initial field).
 * Set before initialize. */
#define INCRMT_SIM_REST 0
#define INCRMT_SIM_FOUR_ROLLS 1
#define INCRMT_SIM_VELOCITY_PULSES 2
int incrmt_fluid_set_sim_velocity(incrmt_fluid *f, int kind);
/* sim_horiz_flow(U,ax,bx,ay,by) (sim_type.hh:116-145, sim_type.cc:51-65):
 * initial velocity (U,0) blended with the solids' momenta; inflow (U,0) in
 * the two left ghost columns of every row after each set_boundaries; zero
 * pressure on the right edge (fluid_2d::zpr): linearly extrapolated right
 * ghost columns (fluid_2d.cc:701-704), the edge conditions of
 * fluid_2d.cc:745-749 and 847-851, the MAC face past the last cell held at
 * zero q (mgs_mac.hh:50, 95), the FEM node column m fixed at zero
 * (mgs_fem.cc:21, 85, 112) and no pressure average removed
 * (fluid_2d.cc:1318). The fluid must be non-periodic in x (the reference's
 * sim_horiz_flow is periodic in y; walls in y are accepted as well). Set
 * before initialize. */
int incrmt_fluid_set_horizontal_flow(incrmt_fluid *f, double U);
/* fluid_2d::initialize(ev_mult,dt_pad,dt_ev_pad) (fluid_2d.cc:101-122). */
int incrmt_fluid_initialize(incrmt_fluid *f, double ev_mult, double dt_pad, double dt_ev_pad);
/* Projection stopping rule: relative residual ||b-Az||/||b|| (the
 * reference's TGMG rule is not available; see mgs_common.hh:37). */
int incrmt_fluid_set_solver(incrmt_fluid *f, double rel_tol, int max_cycles, int precision);
/* Starting iterate of the two projections: the previous step's solution
 * (INCRMT_GUESS_PREVIOUS, the reference's warm start, mgs_common.hh:48-50;
 * the default) or the linear extrapolation in time of the last two at the
 * solve times t + dt/2 (MAC) and t + dt (FEM) (INCRMT_GUESS_EXTRAPOLATE:
 * mgs_base::initial_guess with mg_lin_t_interp, mgs_common.cc:33-66, an
 * option the reference compiles off, mgs_common.hh:12). Either way each
 * solve stops at the same relative residual. */
#define INCRMT_GUESS_PREVIOUS 0
#define INCRMT_GUESS_EXTRAPOLATE 1
/* No reference counterpart: the quadratic (Lagrange) extrapolation in time
 * of the last three solutions (linear while only two exist), O(dt^3) from
 * the new one instead of O(dt^2). */
#define INCRMT_GUESS_QUADRATIC 2
/* No reference counterpart: the quadratic start corrected by the learned
 * guess network (see "learned projection guess" below) along its
 * prediction, by the step that minimises the error's energy norm; when the
 * relative residual after one V-cycle exceeds the fallback tolerance (1e-4
 * by default) the solve restarts from the quadratic start with FP64
 * V-cycles; after 8 such misses in a row the learned start is switched off
 * for 32 solves, then probed again (new parameters or a training call switch
 * it back on at once). The learned start is used once the quadratic start
 * has its three solutions (from the third solve on). */
#define INCRMT_GUESS_LEARNED 3
int incrmt_fluid_set_guess(incrmt_fluid *f, int mode);
/* No reference counterpart: train the guess networks online, one Adam step
 * (learning rate lr) per converged projection on its own start: the loss is
 * the relative energy norm of the error left after the learned start's
 * step along the prediction d, 1 - (e.Ad)^2 / ((e.Ae)(d.Ad)), e being the
 * solution minus the quadratic start; fallback_tol as above. Works in any
 * guess mode (the networks see every solve's start either way). */
int incrmt_fluid_set_guess_training(incrmt_fluid *f, int train, float lr, double fallback_tol);
/* The guess network of one projection (INCRMT_MG_MAC on the m x n cells,
 * INCRMT_MG_FEM on the nodes): packed parameter count, get, set (after
 * initialize). */
int incrmt_fluid_guess_nparams(incrmt_fluid *f, int kind, long long *n);
int incrmt_fluid_get_guess_params(incrmt_fluid *f, int kind, float *params);
int incrmt_fluid_set_guess_params(incrmt_fluid *f, int kind, const float *params);
/* learned starts used, fallbacks, training steps, last training loss,
 * learned start switched off (0/1), misses in a row, the relative residual
 * of the last learned start and after its first V-cycle */
int incrmt_fluid_guess_stats(incrmt_fluid *f, int kind, double *out8);
/* No reference counterpart: the learned start used only where it pays.
 * mode 0: always (the default); 1: its cost in V-cycles is measured once
 * (HIP events around the network's start and the first V-cycle after it);
 * 2: the cost is `cycles`. One solve in four then runs the quadratic start
 * as a control, and once three learned and one control solves are averaged
 * (exponentially), a learned start that saves no more V-cycles than its
 * cost is switched off for 256 solves, then tried again. */
int incrmt_fluid_set_guess_cost(incrmt_fluid *f, int mode, double cycles);
/* cost in V-cycles (< 0 unmeasured), mean V-cycles of learned and of
 * control solves, switched off for cost (0/1), learned and control solves
 * counted since the last decision */
int incrmt_fluid_guess_cost_stats(incrmt_fluid *f, int kind, double *out6);
/* No reference counterpart (host logic, no GPU): the learned start's
 * accept/fallback rule run over a sequence of relative residuals after the
 * first V-cycle: decision 1 accepted, 0 fallback, -1 learned start off. */
int incrmt_guess_policy_run(const double *rel, int n, double tol, int *decision);
/* How the projections' V-cycles are issued: captured once into a HIP graph
 * and replayed (INCRMT_LAUNCH_GRAPH) or launched kernel by kernel
 * (INCRMT_LAUNCH_DIRECT). Results are bit-identical. The default is graphs
 * for solvers whose finest level has fewer than 512^2 unknowns (launch
 * latency dominates) and direct launches above (a 1024^2 step is 1% faster
 * so). A call before initialize is kept and applied there. No reference
 * counterpart (the reference's V-cycles are host loops in TGMG). */
#define INCRMT_LAUNCH_DIRECT 0
#define INCRMT_LAUNCH_GRAPH 1
int incrmt_fluid_set_launch(incrmt_fluid *f, int mode);
/* Whether the end-of-step band rebuild (the second pin of
 * fluid_2d::step_forward, fluid_2d.cc:359; a lone solid's, or the shared
 * one of several solids with windowed sweeps) runs on a second stream
 * beside the FEM projection and the velocity update (on, the default) or
 * after them on the step's stream. It pins to the end-of-step reference map
 * (nX, nY), which update_velocity copies into (X, Y). Results are
 * bit-identical. No reference counterpart (the reference's step is serial
 * host code). */
int incrmt_fluid_set_overlap(incrmt_fluid *f, int on);
/* fluid_2d::step_forward(dt) (fluid_2d.cc:298-369). */
int incrmt_fluid_step(incrmt_fluid *f, double dt);
/* k calls of step_forward(dt), the inner loop of fluid_2d::solve (fluid_2d.cc:274). */
int incrmt_fluid_steps(incrmt_fluid *f, double dt, int k);
/* fluid_2d::dt_reg, fluid_2d::time, fluid_2d::timestep_select (fluid_2d.hh:173-177);
 * timestep_select returns the number of steps l (adt = interval/l), or -1
 * before initialize. */
double incrmt_fluid_dt_reg(const incrmt_fluid *f);
double incrmt_fluid_time(const incrmt_fluid *f);
int incrmt_fluid_timestep_select(const incrmt_fluid *f, double interval, double *adt);
/* State on the reference's ghosted (m+4)x(n+4) grid, row-major, fbase /
 * sbase indexing (fluid_2d.cc:37, obj_field.cc:18). Fluid names: u v p rho
 * s11 s12 s21 s22, c0 c1 (u* of the last step), ul vd (its MAC-projected
 * left/down faces); solid names (object index obj): X Y nX nY phi s. */
int incrmt_fluid_download(incrmt_fluid *f, const char *name, int obj, double *host);
int incrmt_fluid_upload(incrmt_fluid *f, const char *name, int obj, const double *host);
/* The current simulation time (fluid_2d::time) may be set for restarts. */
int incrmt_fluid_set_time(incrmt_fluid *f, double t);
int incrmt_fluid_sync(incrmt_fluid *f);
int incrmt_fluid_get_stats(incrmt_fluid *f, incrmt_stats *out);
/* fluid_2d::write_files(sn) (fluid_2d_io.cc:60-102): the fields selected by
 * the reference's fflags bits (1 u, 2 v, 4 spd, 8 p, 16 w, 32 rho, 64 X,
 * 128 Y, 256 phi, 512 solid fields on the solids' bounding box, 1024 whole
 * grid, 2048 ds/dp, 4096 div, 8192 J) as "<dir>/<name>.<sn>" in the
 * reference's float32 Gnuplot matrix format, the tracers as "trace.<sn>",
 * and the objects' output (a track.dat record; the file is rewritten whole). */
int incrmt_fluid_write_frame(incrmt_fluid *f, const char *dir, int sn, unsigned int fflags);
/* fluid_2d::solve(duration, frames) (fluid_2d.cc:259-296): the header file
 * (save_header, fluid_2d_io.cc:13-19), frame 0 on the first call, then per
 * frame timestep_select's l steps, write_frame (when dir is not null) and
 * one line on stdout: with fflags & 8192 the machine-readable "%d %d %.8g
 * %.8g  %d %.8g %d %.8g" (frame, steps, step and output wall seconds, then
 * per projection its V-cycles in the frame, standing in for TGMG's
 * vcount_extra, and its solver wall seconds), else "# Output frame k [l,
 * step s, output s] {MAC v, FEM v}" with mean V-cycles per solve. */
int incrmt_fluid_solve(incrmt_fluid *f, const char *dir, double duration, int frames, unsigned int fflags);
/* No reference counterpart: the fused-smoother launch shape the autotuner
 * picked for each multigrid level ("nx x ny: SXxSY hH/SXxSY hH", FP64/FP32
 * sweeps) and each solver's V-cycle issue mode ("launch MAC graph, FEM
 * direct"), for reports. */
int incrmt_fluid_describe(incrmt_fluid *f, char *buf, int len);
/* No reference counterpart: time `sweeps` fine-level smoother sweeps of one
 * projection system (INCRMT_MG_MAC / INCRMT_MG_FEM) back to back between two
 * HIP events on the handle's stream (the sweeps act on that system's current
 * iterate); returns the mean milliseconds per launch. */
int incrmt_fluid_time_fine_smoother(incrmt_fluid *f, int kind, int sweeps, int precision, double *ms_per_launch,
                                    int *launches);
/* No reference counterpart: record HIP events around each step's two
 * projections, their starts included (stats.proj_ms; enabling resets it). */
int incrmt_fluid_time_projections(incrmt_fluid *f, int enable);
/* Record HIP events around every fine-level smoother launch (stats.smoother_ms). */
int incrmt_fluid_time_smoothers(incrmt_fluid *f, int enable);
/* No reference counterpart (the reference sweeps every solid over the whole
 * grid, fluid_2d.cc:915-950, 1082-1108, 1242-1273, 510-560, and rebuilds
 * one solid's band after another): with `on` (the default) each solid's
 * per-step sweeps cover only its band's bounding box widened by a few
 * cells, and the band passes of all solids run in shared launches, which
 * gives the same fields bit for bit; off, the sweeps cover the whole grid
 * and each solid's band is rebuilt alone. Takes effect at the next band
 * rebuild. */
int incrmt_fluid_set_solid_windows(incrmt_fluid *f, int on);
void incrmt_fluid_destroy(incrmt_fluid *f);

/* ---- tracers -------------------------------------------------------------
 * fluid_2d's ntrace (fluid_2d.cc:26-43): passive tracers placed by
 * initialize at random positions outside the solids, drawn from the
 * reference's unseeded glibc rand() sequence (fluid_2d.cc:1376-1386), and
 * moved by the improved Euler method on the bicubic velocity every step
 * (fluid_2d.cc:1464-1484). Call before initialize. */
int incrmt_fluid_set_ntrace(incrmt_fluid *f, int ntrace);
/* fluid_2d::tm: per tracer x, y and the stage velocity (4 doubles each). */
int incrmt_fluid_get_tracers(incrmt_fluid *f, double *xy4);
int incrmt_fluid_set_tracers(incrmt_fluid *f, const double *xy4);
/* The track tracer of an obj_simple_spin solid (object.hh:456-461,
 * object.cc:351-417): state = x, y, stage u, stage v, anchored X, Y, step
 * count, previous quadrant, winding number, correction still on. */
int incrmt_fluid_get_track_state(incrmt_fluid *f, int obj, double *st10);
int incrmt_fluid_set_track_state(incrmt_fluid *f, int obj, const double *st10);
/* The object part of fluid_2d::write_files (fluid_2d_io.cc:101,
 * obj_track::write, object.cc:330-336): one track.dat record per track. */
int incrmt_fluid_frame_output(incrmt_fluid *f);
/* track.dat of solid obj: up to cap records (time, x, y, unwound angle) into
 * rec (4 doubles each); *nrec = records written so far (the device keeps
 * the first 16384). */
int incrmt_fluid_track(incrmt_fluid *f, int obj, double *rec, int cap, int *nrec);

/* ---- mgs_mac_* / mgs_fem_* ------------------------------------------------
 * Standalone solver, the constructor mgs_*_const_rho(m,n,x_prd,y_prd,dx,dy,src)
 * (mgs_mac.cc:20-25, mgs_fem.cc:39-43): m x n cells (MAC) or nodes (FEM).
 * A periodic direction (mgs_mac.hh:45-56, mgs_fem.hh:47-52; a periodic FEM
 * direction has m nodes, node m being node 0) needs an even size whose
 * repeated halving ends on a coarsest level of at most 16 points (e.g.
 * 2^k or 3 * 2^k) unless both directions are periodic: then any size of at
 * least 8 is taken (the first coarse level resamples to 2^k or 3 * 2^k
 * points per direction, an odd finest size is smoothed in nine colours);
 * others return INCRMT_ERR_UNSUPPORTED. */
int incrmt_mg_create(int kind, int m, int n, int x_prd, int y_prd, double dx, double dy, incrmt_mg **out);
/* The same with the zero-pressure right edge of mgs_base(m,n,x_prd,y_prd,zpr)
 * (mgs_common.cc:15-17; the reference sets it through fluid_2d::zpr): MAC
 * a_cc of the last column takes twice the right-edge inverse density
 * (mgs_mac.hh:50, 95); FEM node column m-1 is fixed at zero (pfix,
 * mgs_fem.cc:21): arrays keep it, its b and x entries are read as zero and
 * its z and y entries are written as zero. x must not be periodic. The
 * system is definite: b is used as it is. */
int incrmt_mg_create_bc(int kind, int m, int n, int x_prd, int y_prd, int zpr, double dx, double dy,
                        incrmt_mg **out);
/* Varying density: the reference's irho layouts (mgs_mac_varying_rho::irho,
 * 2mn+n+m edge values, fluid_2d.cc:1140-1160; mgs_fem_varying_rho::irho, m*n
 * corner values, fluid_2d.cc:1163-1170). NULL restores constant density. */
int incrmt_mg_set_irho(incrmt_mg *h, const double *irho);
/* mgs_base::solve (mgs_common.hh:72-77): solve A z = b, z warm-started from
 * its input, to ||b-Az|| <= rel_tol ||b|| (b's mean is projected out first,
 * except with a zero-pressure edge). */
int incrmt_mg_solve(incrmt_mg *h, const double *b, double *z, double rel_tol, int max_cycles, int precision,
                    int *cycles, double *rel_res);
/* No reference counterpart (for measuring the solver on the device): `reps`
 * solves of the last incrmt_mg_solve's right-hand side, kept on the
 * device, each from a zero start; *ms_per_solve is the mean time between
 * HIP events recorded on the solver's stream around the solve alone (no
 * host copy of b or z inside). */
int incrmt_mg_time_solve(incrmt_mg *h, double rel_tol, int max_cycles, int precision, int reps,
                         double *ms_per_solve, int *cycles, double *rel_res);
/* V-cycle issue mode of incrmt_mg_solve (INCRMT_LAUNCH_GRAPH or
 * INCRMT_LAUNCH_DIRECT, default by size as for incrmt_fluid_set_launch). No
 * reference counterpart. */
int incrmt_mg_set_launch(incrmt_mg *h, int mode);
/* Mixed-precision solves (INCRMT_PREC_MIXED) as conjugate gradients in FP64
 * preconditioned by one FP32 V-cycle per iteration (on, flexible
 * Polak-Ribiere form), or as V-cycle refinement (off, the default). The
 * stopping test is the same relative residual. Walled systems without a
 * zero-pressure edge only; the others keep the refinement. No reference
 * counterpart (the reference's TGMG is a multigrid iteration). */
int incrmt_mg_set_krylov(incrmt_mg *h, int on);
int incrmt_fluid_set_krylov(incrmt_fluid *f, int on);
/* No reference counterpart (for checking the smoother): `sweeps` forward
 * colored Gauss-Seidel sweeps of A z = b on the host arrays, with fused
 * launch shape `shape` (0..16; see incrmt_fluid_describe; the FEM form of
 * 16 fits in LDS only in single precision) or one launch per colour when
 * shape < 0. */
int incrmt_mg_smooth(incrmt_mg *h, const double *b, double *z, int sweeps, int shape);
/* The same in INCRMT_PREC_FP64 or, with INCRMT_PREC_MIXED, in single
 * precision: b and z rounded to FP32, the FP32 sweeps of the mixed
 * V-cycle, z widened back. */
int incrmt_mg_smooth_prec(incrmt_mg *h, const double *b, double *z, int sweeps, int shape, int precision);
/* y = A x with the solver's operator (mgs_*::mul_a + a_cc, full row). */
int incrmt_mg_apply(incrmt_mg *h, const double *x, double *y);
/* Time `sweeps` fine-level smoother sweeps on the resident problem with HIP
 * events; returns the mean milliseconds per smoother launch. */
int incrmt_mg_time_smoother(incrmt_mg *h, int sweeps, int precision, double *ms_per_launch, int *launches);
/* The same with fused launch shape `shape` (as for incrmt_mg_smooth) in
 * place of the one the autotuner picked, for measuring every shape. */
int incrmt_mg_time_shape(incrmt_mg *h, int sweeps, int precision, int shape, double *ms_per_launch, int *launches);
void incrmt_mg_destroy(incrmt_mg *h);

/* ---- levelset -------------------------------------------------------------
 * levelset(m,n,ax,ay,dx,dy,phi_down,phi_up) + levelset::build_band()
 * (levelset.cc:13-26, 199-335) on host arrays: phi (m*n, row-major) in: the
 * pinned field, out: the narrow band; s out: status codes 0..7. passes: the
 * fixed-point passes the rebuild took. */
int incrmt_levelset_build_band(int m, int n, double ax, double ay, double dx, double dy, double phi_down,
                               double phi_up, double *phi, int *s, int *passes);
/* No reference counterpart: incrmt_levelset_build_band for callers that
 * cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid allocation). phi (in/out) and s
 * (out) are DEVICE arrays of m*n points, row-major; everything else the
 * rebuild needs lives in one static per-device scratch array of 12 MiB,
 * so m and n may be at most 520 (INCRMT_ERR_ARG beyond). Runs on the
 * default stream of the current device and returns when the result is in
 * phi and s. Results are identical to incrmt_levelset_build_band's. Every
 * call on a device shares that device's scratch: calls that overlap on one
 * GPU (from several threads) race. */
int incrmt_levelset_build_band_noalloc(int m, int n, double ax, double ay, double dx, double dy, double phi_down,
                                       double phi_up, double *phi, int *s, int *passes);
/* build_band followed by levelset::extrapolate_fields<E>() of two fields
 * (traverse.hh:235-260), as obj_field::extrapolate does for (nX,nY). */
int incrmt_levelset_extrapolate(int m, int n, double ax, double ay, double dx, double dy, double phi_down,
                                double phi_up, double *phi, int *s, double *fx, double *fy, int *passes);

/* ---- learned projection guess ----------------------------------------------
 * No reference counterpart. A 4-level U-Net on an nx x ny node grid (levels of
 * (n+1)/2 nodes; 16/32/48/64 channels; 3x3 convolutions; BF16 weights and
 * activations on the MFMA units with FP32 accumulation; ReLU; stride-2
 * down-sampling, vertex-centred bilinear up-sampling, skip connections by
 * concatenation; a linear one-channel output) that maps a projection's
 * starting residual, level set and source to a correction of its starting
 * iterate: the hook of mgs_base::initial_guess (mgs_common.cc:33-66).
 * Parameters travel as one packed float vector: per layer, in forward order
 * (E0a 4->16, E0b 16->16, D1 16->32 /2, E1 32->32, D2 32->48 /2, E2 48->48,
 * D3 48->64 /2, B3 64->64, U2 (64+48)->48, U1 (48+32)->32, U0 (32+16)->16,
 * Out 16->1), the weights [cout][tap = 3 (dy+1) + (dx+1)][cin] then the bias
 * [cout]. Host arrays: in3 = nx*ny nodes x 3 channels (node-major), out and
 * target = nx*ny; the loss is 0.5 mean((out - target)^2). */
typedef struct incrmt_unet incrmt_unet;
int incrmt_unet_create(int nx, int ny, unsigned seed, incrmt_unet **out);
int incrmt_unet_nparams(const incrmt_unet *h, long long *n);
int incrmt_unet_get_params(incrmt_unet *h, float *params);
int incrmt_unet_set_params(incrmt_unet *h, const float *params);
int incrmt_unet_forward(incrmt_unet *h, const float *in3, float *out);
/* loss and d loss / d params (packed) of one sample, parameters unchanged */
int incrmt_unet_gradient(incrmt_unet *h, const float *in3, const float *target, double *loss, float *grads);
/* one Adam step (0.9, 0.999, 1e-8) on one sample */
int incrmt_unet_train_step(incrmt_unet *h, const float *in3, const float *target, float lr, double *loss);
void incrmt_unet_destroy(incrmt_unet *h);

#ifdef __cplusplus
}
#endif

#endif /* INCRMT_H */
