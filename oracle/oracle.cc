// C-ABI of the CPU oracle, for tests/ (ctypes) and bench.py's CPU baseline.
// TEST INFRASTRUCTURE ONLY: the product never links or loads this library.
#include <cstring>
#include <vector>
#include "oracle.hh"

using namespace orc;

extern "C" {

// levelset::build_band (levelset.cc:199-335). order/norder: band traversal.
int orc_build_band(int m, int n, double ax, double ay, double dx, double dy,
                   double phi_down, double phi_up, double *phi, int *s,
                   int *order, int *norder) {
    ls_grid g(m, n, ax, ay, dx, dy, phi_down, phi_up);
    std::vector<int> band;
    build_band(g, phi, s, band);
    if (order) {
        std::memcpy(order, band.data(), sizeof(int) * band.size());
        *norder = (int)band.size();
    }
    return 0;
}

// build_band then the two-field extension (traverse.hh:235-260).
int orc_build_band_extrapolate(int m, int n, double ax, double ay, double dx,
                               double dy, double phi_down, double phi_up,
                               double *phi, int *s, double *fx, double *fy) {
    ls_grid g(m, n, ax, ay, dx, dy, phi_down, phi_up);
    std::vector<int> band;
    build_band(g, phi, s, band);
    return extrapolate2(g, phi, s, band, fx, fy) ? 0 : 1;
}

// Extension alone, given a band state (phi, s) and its traversal order.
int orc_extrapolate(int m, int n, double ax, double ay, double dx, double dy,
                    const double *phi, const int *s, const int *order,
                    int norder, double *fx, double *fy) {
    ls_grid g(m, n, ax, ay, dx, dy);
    std::vector<int> band(order, order + norder);
    return extrapolate2(g, phi, s, band, fx, fy) ? 0 : 1;
}

int orc_newton(int m, int n, double ax, double ay, double dx, double dy,
               const double *phi, const int *s, int i, int ij, double *out4) {
    ls_grid g(m, n, ax, ay, dx, dy);
    double x0, y0, x, y;
    bool ok = newton(g, phi, s, i, ij, x0, y0, x, y);
    out4[0] = x0; out4[1] = y0; out4[2] = x; out4[3] = y;
    return ok ? 1 : 0;
}

double orc_bicubic(int m, int n, double ax, double ay, double dx, double dy,
                   const double *phi, double x, double y, double *grad2) {
    ls_grid g(m, n, ax, ay, dx, dy);
    return bicubic(g, phi, x, y, grad2[0], grad2[1]);
}

}

extern "C" {

// Projection operators and solves (mgs_mac.hh:64-102, mgs_fem.hh:86-136),
// periodic in x and/or y when px / py are set (mgs_mac.hh:45-56,
// mgs_fem.hh:47-52). irho may be null for the constant-density systems.
void orc_mac_apply(int m, int n, double xxsp, double yysp, const double *irho,
                   const double *x, double *y, int px, int py, int zpr) {
    mac_system s{m, n, xxsp, yysp, irho, px, py, zpr};
    mac_apply(s, x, y);
}
void orc_fem_apply(int m, int n, double dx, double dy, const double *irho,
                   const double *x, double *y, int px, int py, int zpr) {
    fem_system s{m, n, dx, dy, irho, px, py, zpr};
    fem_apply(s, x, y);
}
int orc_mac_solve(int m, int n, double xxsp, double yysp, const double *irho,
                  const double *b, double *z, double tol, double *relres, int px, int py, int zpr) {
    mac_system s{m, n, xxsp, yysp, irho, px, py, zpr};
    return mac_solve(s, b, z, tol, relres);
}
int orc_fem_solve(int m, int n, double dx, double dy, const double *irho,
                  const double *b, double *z, double tol, double *relres, int px, int py, int zpr) {
    fem_system s{m, n, dx, dy, irho, px, py, zpr};
    return fem_solve(s, b, z, tol, relres);
}

}

extern "C" {

// fluid_2d construction (fluid_2d.cc:26-43, 101-122) for a box with no-slip
// walls or periodic directions, at rest or with a sim_type's velocity, or
// sim_horiz_flow's inflow and zero-pressure outflow (zpr, U).
// params: 13 doubles per object: 9 shape parameters, G, rhos, ev_trans_mult,
// set_velocity (0 or 1).
void *orc_fluid_create(int m, int n, double ax, double bx, double ay, double by,
                       double visc, double rhof, double tw, double pinw,
                       double w_scale, double w_force, double ev_mult,
                       double dt_pad, double dt_ev_pad, double mg_tol, int nobj,
                       const int *kinds, const double *params, int x_prd, int y_prd, int vel, int zpr,
                       double U, int solid_prd_bc) {
    sim_spec sp{m, n, ax, bx, ay, by, visc, rhof, tw, pinw, w_scale, w_force,
                ev_mult, dt_pad, dt_ev_pad, mg_tol, x_prd, y_prd, vel, zpr, U, solid_prd_bc};
    std::vector<object_spec> os(nobj);
    for (int k = 0; k < nobj; k++) {
        os[k].kind = kinds[k];
        for (int q = 0; q < 9; q++) os[k].p[q] = params[13 * k + q];
        os[k].G = params[13 * k + 9];
        os[k].rhos = params[13 * k + 10];
        os[k].ev_trans_mult = params[13 * k + 11];
        os[k].set_velocity = params[13 * k + 12] != 0;
    }
    return fluid_create(sp, os);
}
void orc_fluid_destroy(void *h) { fluid_destroy((fluid *)h); }
// object_eval for one solid given as kind + 9 constructor parameters
void orc_object_eval(int kind, const double *p9, double eps, double time, double X, double Y, double *phi,
                     double *F4) {
    object_spec o{};
    o.kind = kind;
    for (int q = 0; q < 9; q++) o.p[q] = p9[q];
    object_eval(o, eps, time, X, Y, phi, F4);
}
double orc_fluid_dt_reg(void *h) { return fluid_dt_reg((fluid *)h); }
double orc_fluid_time(void *h) { return fluid_time((fluid *)h); }
int orc_fluid_step(void *h, double dt) { return fluid_step((fluid *)h, dt); }
int orc_fluid_get(void *h, const char *name, int k, double *out) { return fluid_get((fluid *)h, name, k, out); }
int orc_fluid_set(void *h, const char *name, int k, const double *in) { return fluid_set((fluid *)h, name, k, in); }
void orc_fluid_stats(void *h, double *out4) { fluid_solver_stats((fluid *)h, out4); }
int orc_fluid_phase(void *h, const char *name, double dt, const double *in, double *out) {
    return fluid_phase((fluid *)h, name, dt, in, out);
}
void orc_fluid_set_time(void *h, double t) { fluid_set_time((fluid *)h, t); }
void orc_fluid_init_tracers(void *h, int ntrace) { fluid_init_tracers((fluid *)h, ntrace); }
void orc_fluid_frame_output(void *h) { fluid_frame_output((fluid *)h); }
int orc_fluid_ntrace(void *h) { return fluid_ntrace((fluid *)h); }
void orc_fluid_get_tracers(void *h, double *out) { fluid_get_tracers((fluid *)h, out); }
void orc_fluid_set_tracers(void *h, const double *in) { fluid_set_tracers((fluid *)h, in); }
int orc_fluid_get_track(void *h, int k, double *st10) { return fluid_get_track((fluid *)h, k, st10); }
int orc_fluid_set_track(void *h, int k, const double *st10) { return fluid_set_track((fluid *)h, k, st10); }
int orc_fluid_track_records(void *h, int k, double *out, int cap) { return fluid_track_records((fluid *)h, k, out, cap); }
void orc_interp2(int m, int n, double ax, double bx, double ay, double by, const double *fx, const double *fy, double x,
                 double y, double *out6) {
    interp2(m, n, ax, bx, ay, by, fx, fy, x, y, out6);
}
int orc_glibc_rand(int n, int *out) { return glibc_rand_probe(n, out); }

}

namespace orc {
double mono_diff_probe(double, double, double, double, double);
}
extern "C" {
// fluid_2d::mono_diff (fluid_2d.cc:604-634), for tests
double orc_mono_diff(double f0, double f1, double f2, double f3, double f4) {
    return orc::mono_diff_probe(f0, f1, f2, f3, f4);
}
}
