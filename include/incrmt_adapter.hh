// C++ adapters that put the C ABI of incrmt.h behind the reference's own
// class interfaces, for a reference driver that moves its step loop and
// projection solves to the GPU (header-only, C++11, no HIP types).
//
//   fluid_2d_gpu<sim_type, mat_const>  fluid_2d's set-up and step loop
//                                      (fluid_2d.hh:131-177, fluid_2d.cc:26-122,
//                                      260-369)
//   mgs_gpu<fluid_2d>                  mgs_base's solver surface
//                                      (mgs_common.hh:20-77): solve(f) ->
//                                      solve_internal(f) on f.src into z
//
// The template parameters are the reference's own types: any types with
// the members used below (sim_type: ax, bx, ay, by, x_prd, y_prd, w_scale,
// w_force; mat_const: G, rhos, ev_trans_mult, set_velocity; fluid_2d: src,
// time, set_edge_irho, set_corner_irho). Errors end the process with the
// library's message, as the reference's own failed solves do
// (mgs_mac.cc:31-34, mgs_fem.cc:89-92).
#ifndef INCRMT_ADAPTER_HH
#define INCRMT_ADAPTER_HH

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "incrmt.h"

namespace incrmt_ref {

inline void check(int rc) {
    if (rc != INCRMT_OK) {
        std::fprintf(stderr, "%s\n", incrmt_last_error());
        std::exit(1);
    }
}

template <class SimType, class MatConst>
class fluid_2d_gpu {
  public:
    // fluid_2d(m,n,sim_type&,visc,rhof,tw,pinw,...) (fluid_2d.cc:26-43)
    fluid_2d_gpu(int m, int n, const SimType &st, double visc, double rhof, double tw, double pinw) {
        check(incrmt_fluid_create(m, n, st.ax, st.bx, st.ay, st.by, st.x_prd, st.y_prd, st.w_scale, st.w_force, visc,
                                  rhof, tw, pinw, &f));
    }
    ~fluid_2d_gpu() { incrmt_fluid_destroy(f); }
    fluid_2d_gpu(const fluid_2d_gpu &) = delete;
    fluid_2d_gpu &operator=(const fluid_2d_gpu &) = delete;
    // add_object(new obj_simple_spin(...), mc) (fluid_2d.cc:74-90)
    void add_simple_spin(double r_scale, double a1, double a2, double piv_r, double th_max, double omega,
                         const MatConst &mc) {
        check(incrmt_fluid_add_spin(f, r_scale, a1, a2, piv_r, th_max, omega, mc.G, mc.rhos, mc.ev_trans_mult));
        set_velocity(mc);
    }
    void add_circle(double cx, double cy, double cr, double grav, const MatConst &mc) {
        check(incrmt_fluid_add_circle(f, cx, cy, cr, grav, mc.G, mc.rhos, mc.ev_trans_mult));
        set_velocity(mc);
    }
    // obj_flapper / obj_full (object.hh:83-86, 245-289)
    void add_flapper(double cx, double cy, double cr, double cl, double ar, double al, double theta,
                     double m_stretch, double T, const MatConst &mc) {
        check(incrmt_fluid_add_flapper(f, cx, cy, cr, cl, ar, al, theta, m_stretch, T, mc.G, mc.rhos,
                                       mc.ev_trans_mult));
        set_velocity(mc);
    }
    void add_full(const MatConst &mc) {
        check(incrmt_fluid_add_full(f, mc.G, mc.rhos, mc.ev_trans_mult));
        set_velocity(mc);
    }
    // sim_horiz_flow::start sets zpr (sim_type.cc:59-65); conv_test sets
    // solid_prd_bc (conv_test.cc:120)
    void horizontal_flow(double U) { check(incrmt_fluid_set_horizontal_flow(f, U)); }
    void solid_prd_bc(bool on) { check(incrmt_fluid_set_solid_periodic(f, on)); }
    void initialize(double ev_mult, double dt_pad, double dt_ev_pad) {  // fluid_2d.cc:101-122
        check(incrmt_fluid_initialize(f, ev_mult, dt_pad, dt_ev_pad));
    }
    void step_forward(const double &dt) { check(incrmt_fluid_step(f, dt)); }  // fluid_2d.cc:298-369
    // fluid_2d::solve's inner loop: l steps of adt (fluid_2d.cc:260-296)
    void steps(double adt, int l) { check(incrmt_fluid_steps(f, adt, l)); }
    int timestep_select(double interval, double &adt) {  // fluid_2d.hh:172-177
        const int l = incrmt_fluid_timestep_select(f, interval, &adt);  // the step count, -1 before initialize
        if (l < 0) check(INCRMT_ERR_ARG);
        return l;
    }
    // fields back on the host, fbase layout (fluid_2d.cc:37), for output()
    void download(const char *name, double *host, int obj = 0) { check(incrmt_fluid_download(f, name, obj, host)); }
    double dt_reg() const { return incrmt_fluid_dt_reg(f); }
    double time() const { return incrmt_fluid_time(f); }
    incrmt_fluid *handle() { return f; }

  private:
    incrmt_fluid *f = nullptr;
    int nobj = 0;
    // mat_const's fourth argument (obj_field.hh:25-28) for the solid just added
    void set_velocity(const MatConst &mc) { check(incrmt_fluid_set_solid_velocity(f, nobj++, mc.set_velocity)); }
};

// mgs_base (mgs_common.hh:20-77) over the standalone solver: the same member
// names (m, n, mn, x_prd, y_prd, zpr, z, wc_time, scount, avg_v_cycles,
// solve, solve_internal). kind is INCRMT_MG_MAC (mgs_mac_*) or INCRMT_MG_FEM
// (mgs_fem_*); varying selects the varying-density variant, whose
// solve_internal first takes the inverse densities from the fluid
// (set_edge_irho, mgs_mac.cc:58-71; set_corner_irho, mgs_fem.cc:104-121).
// The solve stops at relative residual rel_tol (TGMG's own threshold,
// mgs_base::acc, belongs to the absent library).
template <class Fluid>
struct mgs_gpu {
    const int m, n, mn;
    const bool x_prd, y_prd;
    bool zpr;
    const int kind;
    const bool varying;
    double rel_tol;
    int max_cycles, precision;
    double wc_time = 0;
    double *const z;
    int scount = 0;

    mgs_gpu(int kind_, bool varying_, int m_, int n_, bool x_prd_, bool y_prd_, bool zpr_, double dx, double dy,
            double rel_tol_ = 1e-10, int max_cycles_ = 200, int precision_ = INCRMT_PREC_FP64)
        : m(m_), n(n_), mn(m_ * n_), x_prd(x_prd_), y_prd(y_prd_), zpr(zpr_), kind(kind_), varying(varying_),
          rel_tol(rel_tol_), max_cycles(max_cycles_), precision(precision_), z(new double[(size_t)m_ * n_]()) {
        check(incrmt_mg_create_bc(kind, m, n, x_prd, y_prd, zpr, dx, dy, &h));
        if (varying) irho.assign(kind == INCRMT_MG_MAC ? (size_t)2 * mn + n + m : (size_t)mn, 0.0);
    }
    virtual ~mgs_gpu() {
        incrmt_mg_destroy(h);
        delete[] z;
    }
    mgs_gpu(const mgs_gpu &) = delete;
    mgs_gpu &operator=(const mgs_gpu &) = delete;
    // mean V-cycles per solve since the last call; resets the counters
    float avg_v_cycles() {
        wc_time = 0;
        const float a = nsolve ? float(ncycles) / nsolve : 0.f;
        ncycles = nsolve = 0;
        return a;
    }
    // mgs_base::solve: solve_internal, timed
    void solve(Fluid &f) {
        const auto t0 = std::chrono::steady_clock::now();
        solve_internal(f);
        wc_time += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
    // A z = f.src, z warm-started from its last value (mgs_common.hh:48-50)
    virtual void solve_internal(Fluid &f) {
        if (varying) {
            if (kind == INCRMT_MG_MAC) f.set_edge_irho(irho.data());
            else f.set_corner_irho(irho.data());
            check(incrmt_mg_set_irho(h, irho.data()));
        }
        if (zpr && kind == INCRMT_MG_FEM)
            for (int ij = m - 1; ij < mn; ij += m) f.src[ij] = 0;  // mgs_fem.cc:85
        int cyc = 0;
        double res = 0;
        check(incrmt_mg_solve(h, f.src, z, rel_tol, max_cycles, precision, &cyc, &res));
        ncycles += cyc;
        nsolve++;
        scount++;
    }

  private:
    incrmt_mg *h = nullptr;
    std::vector<double> irho;
    long ncycles = 0, nsolve = 0;
};

}  // namespace incrmt_ref

#endif
