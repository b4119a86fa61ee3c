// CPU oracle for the IncRMT per-timestep hot path.
//
// TEST INFRASTRUCTURE ONLY. This is the project's own C++ restatement of the
// reference algorithm (chr1shr/incrmt), used by tests/ as the checker, by
// __graft_entry__.smoke() as the checker, and by bench.py's cpu_baseline leg
// as the timed CPU baseline. The product (incrmt_amd/) never links, loads or
// calls it, and must fail loudly when its HIP library is missing.
//
// Each function cites the reference file:line it follows. Parity pins (what
// checks each part against something independent of this file) are listed in
// the test files named beside each section.
#ifndef INCRMT_ORACLE_HH
#define INCRMT_ORACLE_HH

#include <vector>
#include <cstddef>

namespace orc {

// ---------------------------------------------------------------- level set
// A level-set grid: m x n points at (ax + i dx, ay + j dy), band cut-offs.
// Follows levelset::levelset (levelset/levelset.cc:13-26).
struct ls_grid {
    int m, n, mn;
    double ax, ay, dx, dy, xsp, ysp;
    double phi_down, phi_up;
    ls_grid(int m_, int n_, double ax_, double ay_, double dx_, double dy_,
            double pd = -1e30, double pu = 1e30);
};

// Status codes of the narrow band (levelset.cc:199-263, heap.cc:525-526):
// 0 far inside, 1 inside trial, 2 inside accepted, 3 inside interface,
// 4 outside interface, 5 outside accepted, 6 outside trial, 7 far outside.

// Bicubic interpolation with gradient (levelset.cc:63-79, built/bicubic.cc).
double bicubic(const ls_grid &g, const double *phi, double x, double y,
               double &phix, double &phiy);
// Chopp's modified Newton closest point (levelset.cc:107-127) with the
// compiled reference's integer abs (8 fixed iterations, |d|<1 acceptance).
bool newton(const ls_grid &g, const double *phi, const int *s, int i, int ij,
            double &x0, double &y0, double &x, double &y);
bool newton_failsafe(const ls_grid &g, const double *phi, const int *s, int i,
                     int ij, double x0, double y0, double &x, double &y);

// levelset::build_band (levelset.cc:199-335). phi in/out, s out; band out is
// the traversal order used by extension (ascending phi, inside part first).
void build_band(const ls_grid &g, double *phi, int *s, std::vector<int> &band);

// levelset::extrapolate_fields<obj_field> (traverse.hh:169-260) for two fields.
// Returns false if a least-squares fit stays singular up to radius 7.
bool extrapolate2(const ls_grid &g, const double *phi, const int *s,
                  const std::vector<int> &band, double *fx, double *fy);

// --------------------------------------------------------------- multigrid
// Projection systems (mgs_mac.hh:64-102, mgs_fem.hh:86-136), walls,
// periodic directions or a zero-pressure right edge. irho layouts are the
// reference's own (fluid_2d.cc:1140-1170).
struct mac_system {
    int m, n;
    double xxsp, yysp;
    const double *irho;  // 2mn+n+m edge inverse densities, or null (const rho)
    int px = 0, py = 0;  // periodic directions (mgs_mac.hh:45-56)
    int zpr = 0;         // zero pressure past the right edge (mgs_mac.hh:50, 95)
};
struct fem_system {
    int m, n;            // nodes (a periodic direction: m nodes, m cells)
    double dx, dy;
    const double *irho;  // m*n (+m padding row) corner inverse densities, or null
    int px = 0, py = 0;  // periodic directions (mgs_fem.hh:47-52)
    int zpr = 0;         // node column m-1 fixed at zero (pfix, mgs_fem.cc:21)
};
// y = A x (full operator, diagonal included)
void mac_apply(const mac_system &s, const double *x, double *y);
void fem_apply(const fem_system &s, const double *x, double *y);
// Solve A z = b to ||b - A z||_2 <= tol ||b||_2 (b's mean is projected out
// first, unless zpr makes the system definite; a FEM zpr solve zeroes b on
// the fixed column as mgs_fem.cc:85 does). z is the warm start on entry.
// Returns V-cycles, or -1 on failure.
int mac_solve(const mac_system &s, const double *b, double *z, double tol,
              double *relres);
int fem_solve(const fem_system &s, const double *b, double *z, double tol,
              double *relres);

// ------------------------------------------------------------ fluid step
enum obj_kind { OBJ_SIMPLE_SPIN = 0, OBJ_CIRCLE = 1, OBJ_SQUARE = 2, OBJ_FLAPPER = 3, OBJ_LAMPREY = 4, OBJ_FULL = 5 };

struct object_spec {
    int kind;
    // obj_simple_spin(r_scale,a1,a2,piv_r,th_max,omega) (object.hh:422-467)
    // or obj_circle(cx,cy,cr,grav) (object.hh:128-138)
    // or obj_flapper(cx,cy,cr,cl,ar,al,theta,m_stretch,T) (object.hh:262-267)
    // or obj_lamprey(cx,cy,theta,le,dis,w1,w2,T,mlam) (object.hh:567-570)
    double p[9];
    double G, rhos, ev_trans_mult;
    int set_velocity = 1;  // mat_const::set_velocity (obj_field.hh:25-28)
};

struct sim_spec {
    int m, n;
    double ax, bx, ay, by;
    double visc, rhof, tw, pinw;
    double w_scale, w_force;
    double ev_mult, dt_pad, dt_ev_pad;
    double mg_tol;  // relative residual for both projections
    int x_prd = 0, y_prd = 0;  // periodic directions (sim_type.hh:63-71)
    int vel = 0;    // sim_type::velocity: 0 rest, 1 four_rolls, 2 velocity_pulses
    // sim_horiz_flow (sim_type.hh:116-145): zero pressure on the right edge,
    // inflow (U, 0) on the left, initial velocity (U, 0)
    int zpr = 0;
    double U = 0;
    int solid_prd_bc = 0;  // fluid_2d::solid_prd_bc (fluid_2d.hh:124)
};

class fluid;  // opaque

fluid *fluid_create(const sim_spec &sp, const std::vector<object_spec> &objs);
// One solid's object::phi at (X,Y) and object::actuate applied to F4 =
// (a,b,c,d) after pre_stress_setup(time), for a transition half-width eps.
void object_eval(const object_spec &o, double eps, double time, double X, double Y, double *phi, double *F4);
void fluid_destroy(fluid *f);
double fluid_dt_reg(fluid *f);
double fluid_time(fluid *f);
// fluid_2d::step_forward (fluid_2d.cc:298-369). Returns 0 or an error code.
int fluid_step(fluid *f, double dt);
// Field access on the ghosted (m+4)x(n+4) grid, row-major, the reference's
// own fbase/sbase indexing (fluid_2d.cc:37, obj_field.cc:18).
// name: u v p rho  (fluid)   X Y nX nY phi s  (object k)
int fluid_get(fluid *f, const char *name, int k, double *out);
int fluid_set(fluid *f, const char *name, int k, const double *in);
void fluid_solver_stats(fluid *f, double *out4);
// One named phase of step_forward on the current state (names and the
// reference lines they follow are listed at the definition); in/out carry
// the FEM solution or source where a phase takes or makes one.
int fluid_phase(fluid *f, const char *name, double dt, const double *in, double *out);
void fluid_set_time(fluid *f, double t);
// Tracers: fluid_2d::init_tracers with glibc's unseeded rand() sequence
// (fluid_2d.cc:1376-1386); the per-step improved-Euler update is part of
// fluid_step. frame_output is the object part of write_files (track.dat).
void fluid_init_tracers(fluid *f, int ntrace);
void fluid_frame_output(fluid *f);
int fluid_ntrace(fluid *f);
void fluid_get_tracers(fluid *f, double *out);
void fluid_set_tracers(fluid *f, const double *in);
int fluid_get_track(fluid *f, int k, double *st10);
int fluid_set_track(fluid *f, int k, const double *st10);
int fluid_track_records(fluid *f, int k, double *out, int cap);
// bicubic_interp::f and f_grad_f (bi_interp.cc) of two fields on an m x n grid
void interp2(int m, int n, double ax, double bx, double ay, double by, const double *fx, const double *fy, double x,
             double y, double *out6);
int glibc_rand_probe(int n, int *out);

}  // namespace orc

#endif
