// Oracle restatement of fluid_2d::step_forward (fluid_2d.cc:298-369) and
// everything it calls, for non-periodic no-slip boxes with
// obj_simple_spin / obj_circle solids. TEST INFRASTRUCTURE ONLY.
//
// Arithmetic order follows the reference expression by expression, so that
// on identical inputs it agrees with it to the last bit wherever the
// reference's own result is defined (the two projection solves are solved
// to a stated tolerance instead: the reference's TGMG is absent).
// Sweeps whose results the reference never reads are left out:
//   - the reference-map edge extrapolations in normal_derivatives_refmap and
//     the reference-map copies in godunov_upwinding_set<true>: all are
//     overwritten by compute_half_time_edge_reference_map before any read;
//   - the r0..r3 tangential terms of compute_tangential_stability_derivatives:
//     computed but never read anywhere in the reference.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>
#include <limits>
#include "oracle.hh"

namespace orc {

namespace {

const double small_number = 100. * std::numeric_limits<double>::epsilon();
const double rand_max = 2147483647.;  // RAND_MAX of glibc

// glibc's rand(), i.e. random(3) with its default TYPE_3 state: an additive
// lagged-Fibonacci generator r[k] = r[k-31] + r[k-3] (mod 2^32), output
// r[k] >> 1, whose 31 words are seeded by the Park-Miller sequence
// 16807^k * seed mod (2^31 - 1) and then stepped 310 times. The reference
// never calls srand(), which is seed 1 (common.hh:28-30 draws from it).
struct glibc_rand {
    int32_t s[31];
    int f = 3, r = 0;
    explicit glibc_rand(uint32_t seed = 1) {
        int32_t w = seed ? (int32_t)seed : 1;
        s[0] = w;
        for (int i = 1; i < 31; i++) {  // Schrage's form of 16807 w mod (2^31-1)
            const int32_t hi = w / 127773, lo = w % 127773;
            w = 16807 * lo - 2836 * hi;
            if (w < 0) w += 2147483647;
            s[i] = w;
        }
        for (int i = 0; i < 310; i++) next();
    }
    int next() {
        const uint32_t v = (uint32_t)s[f] + (uint32_t)s[r];
        s[f] = (int32_t)v;
        f = f == 30 ? 0 : f + 1;
        r = r == 30 ? 0 : r + 1;
        return (int)(v >> 1);
    }
    double rnd(double a, double b) { return a + ((b - a) / rand_max) * static_cast<double>(next()); }  // common.hh:28-30
};

// Two-component value of bi_interp.cc's vec arithmetic, evaluated component
// by component in the reference's operator order.
struct v2 {
    double x, y;
};
inline v2 operator+(v2 a, v2 b) { return {a.x + b.x, a.y + b.y}; }
inline v2 operator-(v2 a, v2 b) { return {a.x - b.x, a.y - b.y}; }
inline v2 operator*(double e, v2 a) { return {e * a.x, e * a.y}; }

// bicubic_interp (bi_interp.cc:4-194) over an m x n point grid spanning
// [ax,bx] x [ay,by]: cubic in the bulk, quadratic in the first and last
// intervals of each direction. val(ij) gives the two components at point ij.
template <class F>
struct bicubic2 {
    int m, n;
    double ax, ay, xsp, ysp;
    F val;
    v2 a[16];
    bicubic2(int m_, int n_, double ax_, double bx_, double ay_, double by_, F v)
        : m(m_), n(n_), ax(ax_), ay(ay_), xsp((m - 1) / (bx_ - ax)), ysp((n - 1) / (by_ - ay)), val(v) {}
    void column(int i, int ij, v2 &c0, v2 &c1, v2 &c2, v2 &c3) {  // bi_interp.cc:105-171
        c0 = val(ij);
        if (i == 0) {
            c1 = (-1.5 * val(ij) + 2.0 * val(ij + 1)) - 0.5 * val(ij + 2);
            c2 = (0.5 * val(ij) - val(ij + 1)) + 0.5 * val(ij + 2);
            c3 = {0, 0};
        } else if (i == m - 2) {
            c1 = -0.5 * val(ij - 1) + 0.5 * val(ij + 1);
            c2 = (0.5 * val(ij - 1) - val(ij)) + 0.5 * val(ij + 1);
            c3 = {0, 0};
        } else {
            c1 = -0.5 * val(ij - 1) + 0.5 * val(ij + 1);
            c2 = ((val(ij - 1) - 2.5 * val(ij)) + 2.0 * val(ij + 1)) - 0.5 * val(ij + 2);
            c3 = ((-0.5 * val(ij - 1) + 1.5 * val(ij)) - 1.5 * val(ij + 1)) + 0.5 * val(ij + 2);
        }
    }
    static void fill_ad(v2 *p, v2 c1, v2 c2, v2 c3) {  // bi_interp.cc:153-158
        p[0] = c1;
        p[1] = (-1.5 * c1 + 2.0 * c2) - 0.5 * c3;
        p[2] = (0.5 * c1 - c2) + 0.5 * c3;
        p[3] = {0, 0};
    }
    static void fill_au(v2 *p, v2 c0, v2 c1, v2 c2) {  // bi_interp.cc:165-170
        p[0] = c1;
        p[1] = -0.5 * c0 + 0.5 * c2;
        p[2] = (0.5 * c0 - c1) + 0.5 * c2;
        p[3] = {0, 0};
    }
    static void fill_a(v2 *p, v2 c0, v2 c1, v2 c2, v2 c3) {  // bi_interp.cc:178-183
        p[0] = c1;
        p[1] = -0.5 * c0 + 0.5 * c2;
        p[2] = ((c0 - 2.5 * c1) + 2.0 * c2) - 0.5 * c3;
        p[3] = ((-0.5 * c0 + 1.5 * c1) - 1.5 * c2) + 0.5 * c3;
    }
    void index(double &x, double &y) {  // bi_interp.cc:30-44
        x = (x - ax) * xsp;
        y = (y - ay) * ysp;
        int i = static_cast<int>(x), j = static_cast<int>(y);
        if (i < 0) i = 0; else if (i > m - 2) i = m - 2;
        if (j < 0) j = 0; else if (j > n - 2) j = n - 2;
        const int ij = i + m * j;
        v2 c[4][4];  // c[q][row]: the q-th polynomial coefficient of row j-1..j+2
        column(i, ij, c[0][1], c[1][1], c[2][1], c[3][1]);
        column(i, ij + m, c[0][2], c[1][2], c[2][2], c[3][2]);
        if (j == 0) {
            column(i, ij + 2 * m, c[0][3], c[1][3], c[2][3], c[3][3]);
            for (int q = 0; q < 4; q++) fill_ad(a + 4 * q, c[q][1], c[q][2], c[q][3]);
        } else if (j == n - 2) {
            column(i, ij - m, c[0][0], c[1][0], c[2][0], c[3][0]);
            for (int q = 0; q < 4; q++) fill_au(a + 4 * q, c[q][0], c[q][1], c[q][2]);
        } else {
            column(i, ij - m, c[0][0], c[1][0], c[2][0], c[3][0]);
            column(i, ij + 2 * m, c[0][3], c[1][3], c[2][3], c[3][3]);
            for (int q = 0; q < 4; q++) fill_a(a + 4 * q, c[q][0], c[q][1], c[q][2], c[q][3]);
        }
        x -= i;
        y -= j;
    }
    static v2 yl(const v2 *p, double y) { return p[0] + y * (p[1] + y * (p[2] + y * p[3])); }
    static v2 dyl(const v2 *p, double y) { return p[1] + y * (2.0 * p[2] + (3.0 * y) * p[3]); }
    v2 f(double x, double y) {  // bi_interp.cc:137-140
        index(x, y);
        return yl(a, y) + x * (yl(a + 4, y) + x * (yl(a + 8, y) + x * yl(a + 12, y)));
    }
    v2 f_grad_f(double x, double y, v2 &fx, v2 &fy) {  // bi_interp.cc:146-151
        index(x, y);
        fx = xsp * (yl(a + 4, y) + x * (2.0 * yl(a + 8, y) + (3 * x) * yl(a + 12, y)));
        fy = ysp * (dyl(a, y) + x * (dyl(a + 4, y) + x * (dyl(a + 8, y) + x * dyl(a + 12, y))));
        return yl(a, y) + x * (yl(a + 4, y) + x * (yl(a + 8, y) + x * yl(a + 12, y)));
    }
};
template <class F>
bicubic2<F> make_bicubic(int m, int n, double ax, double bx, double ay, double by, F v) {
    return bicubic2<F>(m, n, ax, bx, ay, by, v);
}

struct cellf {  // fields.hh:7-49 (the members that the step reads)
    double u, v, p, rho, ux, uy, vx, vy;
    double ud, vd, ul, vl, ur, vr, uu, vu;
    double c0, c1, c2, c3, s11, s12, s21, s22;
};
struct scell {  // fields.hh:51-79
    double X, Y, nX, nY, Xx, Xy, Yx, Yy;
    double Xd, Yd, Xl, Yl, Xr, Yr, Xu, Yu;
};

struct solid {
    object_spec sp;
    double eps, pinw, G, rhos, ev_trans_mult, ex_visc, tf1, tf2, tf3;
    // obj_simple_spin constants (object.hh:441-448, object.cc:152-165)
    double c0, c1, c2, piv_r, th_max, omega, K_stiff, piv_thresh, cth, sth;
    // obj_circle constants (object.hh:128-138)
    double cx, cy, cr, grav;
    // obj_flapper (object.hh:245-289) and obj_lamprey (object.hh:542-595):
    // shape, actuation, the transform's rotation (rc, rs) and the state of
    // pre_stress_setup (con, sh)
    double cl = 0, ar = 0, al = 0, amp = 0, fomega = 0;
    double dis = 0, la = 0, w1 = 0, w2 = 0, T = 0, A = 0;
    double rc = 1, rs = 0, con = 0, sh = 0;
    std::vector<scell> sb;
    std::vector<double> phi;  // level set on the ghosted grid
    std::vector<int> st;      // band status
    std::vector<int> band;    // band traversal order
    ls_grid g;
    // obj_track(1,10,...) of obj_simple_spin (object.hh:314-318, 456-461):
    // position, improved-Euler stage velocity, anchored reference map
    bool has_track = false, t_correction = true;
    double tk[6] = {0, 0, 0, 0, 0, 0};
    int count = 1, every = 10, prev_q = 0, wind = 0;
    std::vector<double> rec;  // track.dat records: time, x, y, unwound angle
    solid(const ls_grid &g_) : g(g_) {}

    double phi_of(double X, double Y) const {
        if (sp.kind == OBJ_FULL) return -1;  // object.hh:85
        if (sp.kind == OBJ_SIMPLE_SPIN) {  // object.cc:167-171
            double th = std::atan2(Y, X), c = std::cos(3 * th);
            return std::sqrt(X * X + Y * Y) - c0 - c * (c1 + c2 * c);
        }
        if (sp.kind == OBJ_SQUARE) {  // object.cc:83-87
            X = std::fabs(X) - cr;
            Y = std::fabs(Y) - cr;
            return X < 0 || Y < 0 ? std::max(X, Y) : std::sqrt(X * X + Y * Y);
        }
        if (sp.kind == OBJ_FLAPPER) {  // object.cc:284-287
            double fX = std::fabs(X) - cl;
            return (fX < 0 ? std::fabs(Y) : std::sqrt(fX * fX + Y * Y)) - cr;
        }
        if (sp.kind == OBJ_LAMPREY) {  // object.cc:554-558
            if (X > 0) return dis * Y * Y / (w1 * w1) + X - dis;
            double wl = w1 - (w2 - w1) * X / la;
            return std::max(-la - X, Y * Y / (wl * wl) - 1);
        }
        return std::sqrt(X * X + Y * Y) - cr;  // object.cc:22-24
    }
    bool actuated() const { return sp.kind == OBJ_FLAPPER || sp.kind == OBJ_LAMPREY; }
    void transform(double x, double y, double &X, double &Y) const {
        if (actuated()) {  // object.hh:270-274, 576-580
            x -= cx; y -= cy;
            X = x * rc + y * rs;
            Y = -x * rs + y * rc;
        } else if (sp.kind != OBJ_SIMPLE_SPIN && sp.kind != OBJ_FULL) {
            X = x - cx; Y = y - cy;  // obj_standard, object.hh:116-118
        } else { X = x; Y = y; }     // object.hh:26-28
    }
    void velocity(double X, double Y, double &u, double &v) const {
        if (sp.kind == OBJ_CIRCLE || sp.kind == OBJ_SQUARE) { u = 0 - 0 * Y; v = 0 + 0 * X; }  // object.cc:13-16
        else { u = v = 0.; }  // object.hh:31-33
    }
    // smoothed Heaviside and helpers (obj_field.hh:97-121)
    double trans_func_in(double x) const { return x < -eps ? 1 : 0.5 - x * tf1 - tf2 * std::sin(tf3 * x); }
    double trans_func(double x) const { return x > eps ? 0 : trans_func_in(x); }
    double tderiv_func_in(double x) const { return x < -eps ? 0 : 0.5 * (1 + std::cos(M_PI * x / eps)); }
    double coll_func(double x) const { return x > eps ? 0 : 0.5 - x * tf1; }
    void pre_stress_setup(double time) {
        if (sp.kind == OBJ_FLAPPER) {  // object.cc:291-296
            double a = std::sin(fomega * time);
            a *= a * a; a *= a; a *= a;
            con = -amp * a;
            return;
        }
        if (sp.kind == OBJ_LAMPREY) {  // object.cc:563-568
            sh = time / T;
            double o = 2 * sh;
            con = o < 1 ? A * o * o * (3 - 2 * o) : A;
            sh *= la;
            return;
        }
        if (sp.kind != OBJ_SIMPLE_SPIN) return;  // object.cc:175-181
        double ang = time * omega, theta = ang > 2 * M_PI ? 0 : th_max * (1 - std::cos(ang));
        cth = std::cos(theta); sth = std::sin(theta);
    }
    // object::actuate (object.cc:301-308, 573-580) on F = (a b; c d)
    void actuate(double X, double Y, double &a, double &b, double &c, double &d) const {
        if (sp.kind == OBJ_FLAPPER) {
            double fX = std::fabs(X) - al, ph = (fX < 0 ? std::fabs(Y) : std::sqrt(fX * fX + Y * Y)) - ar;
            if (ph < eps) {
                double fac = std::exp(con * Y * trans_func_in(ph));
                a *= fac; b *= 1. / fac;
                c *= fac; d *= 1. / fac;
            }
        } else if (sp.kind == OBJ_LAMPREY) {
            double dx = X + (Y > 0 ? 1.5 * la : la) + sh, wl = w1 - (w2 - w1) * X / la, fac = std::fabs(Y) / wl,
                   st = std::exp(con * (fac < 1 ? fac * (2 - fac) : 1)
                                 * trans_func(-std::min(-X, la * 0.15 - std::fabs(dx - la * (int(dx / la) + 0.5)))));
            a *= st; b *= 1. / st;
            c *= st; d *= 1. / st;
        }
    }
    void accel(double x, double y, double X, double Y, double phiv, double &ax, double &ay) const {
        if (sp.kind == OBJ_SIMPLE_SPIN) {  // object.cc:188-197
            double rsq = x * x + y * y;
            if (rsq < piv_thresh) {
                double rx = X * cth - Y * sth, ry = X * sth + Y * cth,
                       K = K_stiff * trans_func_in(std::sqrt(rsq) - piv_r);
                ax -= K * (x - rx);
                ay -= K * (y - ry);
            }
        } else if (sp.kind == OBJ_CIRCLE || sp.kind == OBJ_SQUARE) {
            ay -= grav * trans_func_in(phiv);  // object.cc:31-34
        }
    }
};

// The constants of one solid from its constructor's arguments and the
// grid's transition half-width eps (obj_field.cc:14-20).
void setup_solid(solid &so, const object_spec &o, double eps) {
    so.sp = o;
    so.eps = eps;
    so.G = o.G; so.rhos = o.rhos; so.ev_trans_mult = o.ev_trans_mult;
    so.tf1 = 0.5 / eps; so.tf2 = 0.5 / M_PI; so.tf3 = M_PI / eps;
    if (o.kind == OBJ_SIMPLE_SPIN) {  // object.hh:449-455
        const double rs = o.p[0], a1 = o.p[1], a2 = o.p[2];
        const double scf = rs / (1 + a1 + a2);
        so.c0 = scf * (1 - a2); so.c1 = scf * a1; so.c2 = 2 * scf * a2;
        so.piv_r = o.p[3]; so.th_max = o.p[4]; so.omega = o.p[5];
        so.cth = 1; so.sth = 0;
        so.has_track = true;  // tracer at (0.9 r_scale, 0), X = x, Y = y (object.hh:456-461)
        so.tk[0] = 0.9 * rs; so.tk[1] = 0;
        so.tk[4] = so.tk[0]; so.tk[5] = so.tk[1];
    } else if (o.kind == OBJ_FLAPPER) {  // object.hh:262-267
        so.cx = o.p[0]; so.cy = o.p[1]; so.cr = o.p[2]; so.cl = o.p[3];
        so.ar = o.p[4]; so.al = o.p[5];
        so.amp = std::log(o.p[7]) / so.ar; so.fomega = 2 * M_PI / o.p[8];
        so.rc = std::cos(o.p[6]); so.rs = std::sin(o.p[6]);
        so.grav = 0;
    } else if (o.kind == OBJ_LAMPREY) {  // object.hh:567-570
        so.cx = o.p[0]; so.cy = o.p[1];
        so.rc = std::cos(o.p[2]); so.rs = std::sin(o.p[2]);
        so.dis = o.p[4]; so.la = o.p[3] - o.p[4]; so.w1 = o.p[5]; so.w2 = o.p[6];
        so.T = o.p[7]; so.A = std::log(o.p[8]);
        so.grav = 0;
    } else if (o.kind == OBJ_FULL) {
        so.grav = 0;
    } else {
        so.cx = o.p[0]; so.cy = o.p[1]; so.cr = o.p[2]; so.grav = o.p[3];
    }
}

}  // namespace

class fluid {
  public:
    sim_spec sp;
    int m, n, mn, m_fem, n_fem, ml, nl;
    double ax, bx, ay, by, dx, dy, xsp, ysp, xxsp, yysp, visc, rhof, eps, epsinv;
    double w_scale, w_force, dt_reg, time;
    bool const_rho;
    std::vector<cellf> fb;
    std::vector<solid> obj;
    std::vector<double> src, zmac, zfem, irho_mac, irho_fem;
    int mac_cycles = 0, fem_cycles = 0;
    double mac_res = 0, fem_res = 0;
    int ntrace = 0;
    std::vector<double> tm;  // fluid tracers: x, y, stage velocity (fluid_2d.cc:1376-1386)
    glibc_rand rng;

    int at(int i, int j) const { return (j + 2) * ml + (i + 2); }
    cellf &F(int i, int j) { return fb[at(i, j)]; }

    fluid(const sim_spec &s, const std::vector<object_spec> &os);
    void choose_ev_and_dt();
    void init_fields();
    void set_boundaries();
    void set_rho();
    void compute_stress();
    void solid_stress(bool left, solid &o, int i, int j, double &s1s, double &s2s, double &sfrac);
    void fluid_stress(bool left, int i, int j, double &s1f, double &s2f);
    void collision_stress(bool left, int i, int j, double &s1, double &s2, std::vector<solid *> &ct);
    static double mono_diff(double f0, double f1, double f2, double f3, double f4);
    static double delta_lim(double f0, double f1, double f2);
    static double delta_f(double f0, double f1, double f2);
    void normal_derivatives(double dt);
    void godunov(bool tang);
    void edge_bc();
    void set_lr_tang(int ij);
    void set_du_tang(int ij);
    void sim_velocity(double x, double y, double &u, double &v) const;
    void tangential();
    void half_time_edges(double dt);
    void acceleration(int i, int j, double &accx, double &accy, double x, double y, bool pressure);
    void wall_force(double x, double y, double phiv, double &fx, double &fy);
    int mac_projection();
    void update_reference_map(double dt);
    int pin();
    void compute_ustar(double dt);
    int fem_projection(double dt);
    void fem_source(double dt);
    void copy_pressure();
    void update_velocity(double dt);
    int step(double dt);
    // tracers (fluid_2d.cc:1376-1484, object.cc:313-417)
    void init_tracers(int nt);
    bool inside_object(double tx, double ty);
    v2 vel(double x, double y);
    void tracer_update_1(double *tp);
    void tracer_update_2(double *tp, double dt);
    void tracer_remap(double &x, double &y);
    void update_tracers(int stage, double dt);
    bool correct_tracer(solid &o, double xref, double yref, double &xx, double &yy);
    void track_output(solid &o);
    void frame_output();
};

fluid::fluid(const sim_spec &s, const std::vector<object_spec> &os)
    : sp(s), m(s.m), n(s.n), mn(s.m * s.n), m_fem(s.x_prd ? s.m : s.m + 1), n_fem(s.y_prd ? s.n : s.n + 1),
      ml(s.m + 4), nl(s.n + 4),
      ax(s.ax), bx(s.bx), ay(s.ay), by(s.by), dx((s.bx - s.ax) / s.m), dy((s.by - s.ay) / s.n),
      xsp(1 / dx), ysp(1 / dy), xxsp(xsp * xsp), yysp(ysp * ysp), visc(s.visc), rhof(s.rhof),
      eps(s.tw * (dx < dy ? dx : dy)), epsinv(1 / eps), w_scale(s.w_scale), w_force(s.w_force),
      dt_reg(0), time(0) {
    fb.assign(ml * nl, cellf());
    std::memset(fb.data(), 0, sizeof(cellf) * fb.size());
    const ls_grid g(ml, nl, ax - 1.5 * dx, ay - 1.5 * dy, dx, dy, -eps - 4.5 * dx, eps + 4.5 * dx);
    for (const auto &o : os) {
        solid so(g);
        setup_solid(so, o, eps);
        so.pinw = s.pinw * dx;  // obj_field.cc:14-20
        so.sb.assign(ml * nl, scell());
        std::memset(so.sb.data(), 0, sizeof(scell) * so.sb.size());
        so.phi.assign(ml * nl, 0.0);
        so.st.assign(ml * nl, 0);
        obj.push_back(so);
    }
    const_rho = true;
    for (auto &o : obj)
        if (std::fabs(o.rhos - rhof) > small_number) const_rho = false;
    src.assign(m_fem * n_fem, 0.0);
    zmac.assign(mn, 0.0);
    zfem.assign(m_fem * n_fem, 0.0);
    irho_mac.assign(2 * mn + n + m, 0.0);
    irho_fem.assign(m_fem * n_fem + m_fem, 0.0);
    choose_ev_and_dt();
    for (auto &o : obj)
        if (o.sp.kind == OBJ_SIMPLE_SPIN) {  // object.cc:152-165
            o.piv_thresh = o.piv_r + o.eps;
            o.piv_thresh *= o.piv_thresh;
            o.K_stiff = 0.1 / (dt_reg * dt_reg);
        }
    init_fields();
}

void fluid::choose_ev_and_dt() {  // fluid_2d.cc:152-192, obj_field.cc:35-49
    double hmax = dx > dy ? dx : dy, hmin = dx > dy ? dy : dx, sws_max = 0, ex_visc_max = 0,
           rhos_div_ev_min = std::numeric_limits<double>::max();
    for (auto &o : obj) {
        double tfac = sp.ev_mult * hmax;
        double sws = std::sqrt(o.G / o.rhos);
        if (sws > sws_max) sws_max = sws;
        o.ex_visc = tfac * o.rhos * sws;
        if (o.ex_visc > ex_visc_max) ex_visc_max = o.ex_visc;
        double r_d_ev = o.rhos / (o.ex_visc * (o.ev_trans_mult + 1));
        if (r_d_ev < rhos_div_ev_min) rhos_div_ev_min = r_d_ev;
    }
    double dt0 = hmin / sws_max * sp.dt_pad, dt1 = 0.5 * rhos_div_ev_min / (xxsp + yysp) * sp.dt_ev_pad,
           dt2 = 0.5 * rhof / (visc * (xxsp + yysp)) * sp.dt_pad;
    if (dt0 < dt1) dt_reg = dt0 < dt2 ? dt0 : dt2;
    else dt_reg = dt1 < dt2 ? dt1 : dt2;
}

void fluid::init_fields() {  // fluid_2d.cc:195-254
    for (auto &o : obj) {
        for (int j = -2; j < n + 2; j++) {
            double y = ay + dy * (0.5 + j);
            for (int i = -2; i < m + 2; i++) {
                scell &s = o.sb[at(i, j)];
                o.transform(ax + dx * (0.5 + i), y, s.X, s.Y);
                // nX starts as X: the reference leaves it unset, and it is
                // only ever read where the step has written it first
                s.nX = s.X; s.nY = s.Y;
                o.phi[at(i, j)] = o.phi_of(s.X, s.Y);
            }
        }
        build_band(o.g, o.phi.data(), o.st.data(), o.band);
    }
    for (int j = 0; j < n; j++) {
        for (int i = 0; i < m; i++) {
            cellf &f = F(i, j);
            double mu = 0, mv = 0, rho = 0, sfrac = 0;
            for (auto &o : obj) {  // obj_field.cc:62-75
                if (!o.sp.set_velocity) continue;
                const scell &s = o.sb[at(i, j)];
                double tf = o.trans_func(o.phi[at(i, j)]), uu, vv;
                o.velocity(s.X, s.Y, uu, vv);
                sfrac += tf;
                rho += tf * o.rhos;
                mu += tf * o.rhos * uu;
                mv += tf * o.rhos * vv;
            }
            if (sfrac >= 1) {
                f.u = mu / rho;
                f.v = mv / rho;
            } else {
                double uu = 0, vv = 0, irho = 1 / ((1 - sfrac) * rhof + rho);
                sim_velocity(ax + dx * (0.5 + i), ay + dy * (0.5 + j), uu, vv);
                f.u = ((1 - sfrac) * rhof * uu + mu) * irho;
                f.v = ((1 - sfrac) * rhof * vv + mv) * irho;
            }
            f.p = 0;
        }
        F(m, j).p = 0;
    }
    for (int i = 0; i <= m; i++) F(i, n).p = 0;
    set_boundaries();
    compute_stress();
    set_rho();
}

// sim_type::velocity (sim_type.hh:67-70; sim_four_rolls, sim_type.cc:30-35;
// sim_velocity_pulses, sim_type.cc:43-47 with pulse(), sim_type.hh:109-113;
// sim_horiz_flow, sim_type.hh:129-135)
void fluid::sim_velocity(double x, double y, double &u, double &v) const {
    u = v = 0;
    if (sp.zpr) {
        u = sp.U;
        return;
    }
    if (sp.vel == 1) {
        u = std::sin(M_PI * x) * std::cos(M_PI * y);
        v = -std::cos(M_PI * x) * std::sin(M_PI * y);
    } else if (sp.vel == 2) {
        for (double k = 2, s = 1, z = -5 / 6.; z < 1; z += 1 / 3., k += 2, s = -s) {
            const double xx = x - z, yy = y - z;
            const double f = s * std::exp(-k * (2 - std::cos(xx * M_PI) - std::cos(yy * M_PI)));
            u -= std::sin(yy * M_PI) * f;
            v += std::sin(xx * M_PI) * f;
        }
    }
}

void fluid::set_boundaries() {  // fluid_2d.cc:689-726: no-slip walls or periodic copies
    if (sp.solid_prd_bc) {  // set_boundaries_solid (fluid_2d.cc:818-835), s_field::prd_bc
        const double lx = bx - ax, ly = by - ay;
        for (auto &o : obj) {
            auto S = [&](int i, int j) -> scell & { return o.sb[at(i, j)]; };
            auto pb = [&](scell &d, const scell &s, double ex, double ey) { d.X = s.X + ex; d.Y = s.Y + ey; };
            for (int j = 0; j < n; j++) {
                pb(S(-2, j), S(m - 2, j), -lx, 0); pb(S(-1, j), S(m - 1, j), -lx, 0);
                pb(S(m, j), S(0, j), lx, 0); pb(S(m + 1, j), S(1, j), lx, 0);
            }
            for (int i = -2; i < m + 2; i++) {
                pb(S(i, -2), S(i, n - 2), 0, -ly); pb(S(i, -1), S(i, n - 1), 0, -ly);
                pb(S(i, n), S(i, 0), 0, ly); pb(S(i, n + 1), S(i, 1), 0, ly);
            }
        }
    }
    for (int j = 0; j < n; j++) {
        if (sp.x_prd) {  // field::prd_bc (fields.hh:39-41)
            F(-2, j).u = F(m - 2, j).u; F(-2, j).v = F(m - 2, j).v;
            F(-1, j).u = F(m - 1, j).u; F(-1, j).v = F(m - 1, j).v;
            F(m, j).u = F(0, j).u; F(m, j).v = F(0, j).v;
            F(m + 1, j).u = F(1, j).u; F(m + 1, j).v = F(1, j).v;
            continue;
        }
        F(-2, j).u = -F(1, j).u; F(-2, j).v = -F(1, j).v;
        F(-1, j).u = -F(0, j).u; F(-1, j).v = -F(0, j).v;
        if (sp.zpr) {  // field::extrap (fields.hh:45-48)
            F(m, j).u = 2 * F(m - 1, j).u - F(m - 2, j).u; F(m, j).v = 2 * F(m - 1, j).v - F(m - 2, j).v;
            F(m + 1, j).u = 2 * F(m, j).u - F(m - 1, j).u; F(m + 1, j).v = 2 * F(m, j).v - F(m - 1, j).v;
            continue;
        }
        F(m, j).u = -F(m - 1, j).u; F(m, j).v = -F(m - 1, j).v;
        F(m + 1, j).u = -F(m - 2, j).u; F(m + 1, j).v = -F(m - 2, j).v;
    }
    for (int i = -2; i < m + 2; i++) {
        if (sp.y_prd) {
            F(i, -2).u = F(i, n - 2).u; F(i, -2).v = F(i, n - 2).v;
            F(i, -1).u = F(i, n - 1).u; F(i, -1).v = F(i, n - 1).v;
            F(i, n).u = F(i, 0).u; F(i, n).v = F(i, 0).v;
            F(i, n + 1).u = F(i, 1).u; F(i, n + 1).v = F(i, 1).v;
            continue;
        }
        F(i, -2).u = -F(i, 1).u; F(i, -2).v = -F(i, 1).v;
        F(i, -1).u = -F(i, 0).u; F(i, -1).v = -F(i, 0).v;
        F(i, n).u = -F(i, n - 1).u; F(i, n).v = -F(i, n - 1).v;
        F(i, n + 1).u = -F(i, n - 2).u; F(i, n + 1).v = -F(i, n - 2).v;
    }
    if (sp.zpr)  // sim_horiz_flow::boundary (sim_type.cc:51-55)
        for (int j = -2; j < n + 2; j++) {
            F(-1, j).u = F(-2, j).u = sp.U;
            F(-1, j).v = F(-2, j).v = 0;
        }
}

void fluid::set_rho() {  // fluid_2d.cc:1124-1137, obj_field.cc:85-90
#pragma omp parallel for
    for (int ij = 0; ij < ml * nl; ij++) {
        double rho = 0, sfrac = 0;
        for (auto &o : obj) {
            double tf = o.trans_func(o.phi[ij]);
            sfrac += tf;
            rho += tf * o.rhos;
        }
        fb[ij].rho = sfrac >= 1 ? rho / sfrac : (1 - sfrac) * rhof + rho;
    }
}

// fluid_2d.cc:510-560 (left: the edge left of cell (i,j); else the edge below)
void fluid::solid_stress(bool left, solid &o, int i, int j, double &s1s, double &s2s, double &sfrac) {
    const int ij = at(i, j), d = left ? 1 : ml;
    double phiv = 0.5 * (o.phi[ij] + o.phi[ij - d]);
    if (phiv > o.eps) return;
    const scell *s = &o.sb[ij];
    double Xx, Yx, Xy, Yy, X, Y;
    if (left) {
        X = 0.5 * (s->X + s[-1].X);
        Y = 0.5 * (s->Y + s[-1].Y);
        Xx = xsp * (s->X - s[-1].X);
        Yx = xsp * (s->Y - s[-1].Y);
        Xy = 0.25 * ysp * (s[ml].X + s[ml - 1].X - s[-ml].X - s[-ml - 1].X);
        Yy = 0.25 * ysp * (s[ml].Y + s[ml - 1].Y - s[-ml].Y - s[-ml - 1].Y);
    } else {
        X = 0.5 * (s->X + s[-ml].X);
        Y = 0.5 * (s->Y + s[-ml].Y);
        Xx = 0.25 * xsp * (s[1].X + s[-ml + 1].X - s[-1].X - s[-ml - 1].X);
        Yx = 0.25 * xsp * (s[1].Y + s[-ml + 1].Y - s[-1].Y - s[-ml - 1].Y);
        Xy = ysp * (s->X - s[-ml].X);
        Yy = ysp * (s->Y - s[-ml].Y);
    }
    // F = [Yy,-Xy;-Yx,Xx] with the object's actuation, sigma = F F^T
    // (fluid_2d.cc:533-541, mat.hh:159-161)
    double Fa = Yy, Fb = -Xy, Fc = -Yx, Fd = Xx;
    o.actuate(X, Y, Fa, Fb, Fc, Fd);
    const double sa = Fa * Fa + Fb * Fb, sb = Fa * Fc + Fb * Fd, sd = Fc * Fc + Fd * Fd;
    double sfac = (1 / 3.0) * ((sa + sd) + 1), tf = o.trans_func_in(phiv), G = o.G * tf,
           ex_visc = o.ex_visc * tf * (1 + o.ev_trans_mult * o.tderiv_func_in(phiv));
    const cellf *fp = &fb[ij];
    sfrac += tf;
    if (left) {
        s1s += G * (sa - sfac);
        s2s += G * sb;
        s1s += ex_visc * xsp * (fp->u - fp[-1].u);
        s2s += ex_visc * xsp * (fp->v - fp[-1].v);
    } else {
        s1s += G * sb;
        s2s += G * (sd - sfac);
        s1s += ex_visc * ysp * (fp->u - fp[-ml].u);
        s2s += ex_visc * ysp * (fp->v - fp[-ml].v);
    }
}

void fluid::fluid_stress(bool left, int i, int j, double &s1f, double &s2f) {  // fluid_2d.cc:568-598
    const cellf *fp = &fb[at(i, j)];
    double ux, vx, uy, vy;
    if (left) {
        ux = xsp * (fp->u - fp[-1].u);
        vx = xsp * (fp->v - fp[-1].v);
        uy = 0.25 * ysp * (fp[ml].u + fp[ml - 1].u - fp[-ml].u - fp[-ml - 1].u);
        s1f = 2 * visc * ux;
        s2f = visc * (vx + uy);
    } else {
        vx = 0.25 * xsp * (fp[1].v + fp[-ml + 1].v - fp[-1].v - fp[-ml - 1].v);
        uy = ysp * (fp->u - fp[-ml].u);
        vy = ysp * (fp->v - fp[-ml].v);
        s1f = visc * (vx + uy);
        s2f = 2 * visc * vy;
    }
}

void fluid::collision_stress(bool left, int i, int j, double &s1, double &s2, std::vector<solid *> &ct) {
    // fluid_2d.cc:455-500
    const int ij = at(i, j), k = (int)ct.size();
    auto edge = [&](const solid *o, double &p, double &px, double &py) {
        const double *ph = &o->phi[ij];
        if (left) {
            p = 0.5 * (*ph + ph[-1]);
            px = xsp * (*ph - ph[-1]);
            py = 0.25 * ysp * (ph[ml] + ph[ml - 1] - ph[-ml] - ph[-ml - 1]);
        } else {
            p = 0.5 * (*ph + ph[-ml]);
            px = 0.25 * xsp * (ph[1] + ph[-ml + 1] - ph[-1] - ph[-ml - 1]);
            py = ysp * (*ph - ph[-ml]);
        }
    };
    for (int a = 0; a < k - 1; a++) {
        double p1, p1x, p1y;
        edge(ct[a], p1, p1x, p1y);
        for (int b = a + 1; b < k; b++) {
            double p2, p2x, p2y;
            edge(ct[b], p2, p2x, p2y);
            double norx = (p2x - p1x), nory = (p2y - p1y), nn = norx * norx + nory * nory,
                   fac = 4. * std::min(ct[a]->coll_func(p1), ct[b]->coll_func(p2)) * (ct[a]->G + ct[b]->G);
            fac = nn < small_number ? 0. : fac / nn;
            if (left) {
                s1 -= fac * 0.5 * (norx * norx - nory * nory);
                s2 -= fac * norx * nory;
            } else {
                s1 -= fac * norx * nory;
                s2 -= fac * 0.5 * (nory * nory - norx * norx);
            }
        }
    }
}

void fluid::compute_stress() {  // fluid_2d.cc:381-445
    for (auto &o : obj) o.pre_stress_setup(time);
#pragma omp parallel for
    for (int j = 0; j < n + 1; j++) {
        std::vector<solid *> ct;
        for (int i = 0; i < m + 1; i++) {
            cellf &f = F(i, j);
            for (int left = 1; left >= 0; left--) {
                ct.clear();
                double s1s = 0, s2s = 0, sfrac = 0, a, b;
                for (auto &o : obj) {
                    double before = sfrac;
                    solid_stress(left, o, i, j, s1s, s2s, sfrac);
                    // an object joins the collision table when its edge phi <= eps
                    const int d = left ? 1 : ml;
                    if (0.5 * (o.phi[at(i, j)] + o.phi[at(i, j) - d]) <= o.eps) ct.push_back(&o);
                    (void)before;
                }
                if (sfrac >= 1) {
                    a = s1s / sfrac;
                    b = s2s / sfrac;
                } else {
                    double s1f, s2f;
                    fluid_stress(left, i, j, s1f, s2f);
                    a = (1 - sfrac) * s1f + s1s;
                    b = (1 - sfrac) * s2f + s2s;
                }
                if (ct.size() > 1) collision_stress(left, i, j, a, b, ct);
                if (left) { f.s11 = a; f.s21 = b; }
                else { f.s12 = a; f.s22 = b; }
            }
        }
    }
}

// fluid_2d.cc:604-634
double fluid::delta_lim(double f0, double f1, double f2) {
    double drm = f1 - f0, drp = f2 - f1;
    return drm > 0 ? (drp > 0 ? 2 * std::min(drm, drp) : 0) : (drp > 0 ? 0 : -2 * std::max(drm, drp));
}
double fluid::delta_f(double f0, double f1, double f2) {
    double drc = 0.5 * (f2 - f0);
    return drc > 0 ? std::min(drc, delta_lim(f0, f1, f2)) : std::max(drc, -delta_lim(f0, f1, f2));
}
double fluid::mono_diff(double f0, double f1, double f2, double f3, double f4) {
    double tdrc = f3 - f1;
    double s = (1 / 6.) * (4 * tdrc - delta_f(f0, f1, f2) - delta_f(f2, f3, f4));
    double t = delta_lim(f1, f2, f3);
    return tdrc > 0 ? std::min(s, t) : std::max(s, -t);
}

void fluid::normal_derivatives(double dt) {  // fluid_2d.cc:877-950
#pragma omp parallel for
    for (int j = 0; j < n; j++)
        for (int i = 0; i < m; i++) {
            cellf *fp = &F(i, j);
            cellf &f = *fp;
            f.ux = xsp * mono_diff(fp[-2].u, fp[-1].u, f.u, fp[1].u, fp[2].u);
            f.vx = xsp * mono_diff(fp[-2].v, fp[-1].v, f.v, fp[1].v, fp[2].v);
            f.uy = ysp * mono_diff(fp[-2 * ml].u, fp[-ml].u, f.u, fp[ml].u, fp[2 * ml].u);
            f.vy = ysp * mono_diff(fp[-2 * ml].v, fp[-ml].v, f.v, fp[ml].v, fp[2 * ml].v);
            f.ul = f.u - 0.5 * (dx + dt * f.u) * f.ux;
            f.vl = f.v - 0.5 * (dx + dt * f.u) * f.vx;
            f.ur = f.u + 0.5 * (dx - dt * f.u) * f.ux;
            f.vr = f.v + 0.5 * (dx - dt * f.u) * f.vx;
            f.ud = f.u - 0.5 * (dy + dt * f.v) * f.uy;
            f.vd = f.v - 0.5 * (dy + dt * f.v) * f.vy;
            f.uu = f.u + 0.5 * (dy - dt * f.v) * f.uy;
            f.vu = f.v + 0.5 * (dy - dt * f.v) * f.vy;
            for (auto &o : obj) {
                scell *sp2 = &o.sb[at(i, j)];
                scell &s = *sp2;
                s.Xx = xsp * mono_diff(sp2[-2].X, sp2[-1].X, s.X, sp2[1].X, sp2[2].X);
                s.Yx = xsp * mono_diff(sp2[-2].Y, sp2[-1].Y, s.Y, sp2[1].Y, sp2[2].Y);
                s.Xy = ysp * mono_diff(sp2[-2 * ml].X, sp2[-ml].X, s.X, sp2[ml].X, sp2[2 * ml].X);
                s.Yy = ysp * mono_diff(sp2[-2 * ml].Y, sp2[-ml].Y, s.Y, sp2[ml].Y, sp2[2 * ml].Y);
            }
        }
}

// godunov_set_lr / godunov_set_du (fluid_2d.cc:642-686) over interior faces
void fluid::godunov(bool tang) {
#pragma omp parallel for
    for (int j = 0; j < n; j++)
        for (int i = 0; i < m; i++) {
            const int ij = at(i, j);
            if (i < m - 1) {
                cellf &f = fb[ij], &fr = fb[ij + 1];
                if (f.ur + fr.ul > 0 && f.ur > 0) {
                    fr.ul = f.ur; fr.vl = f.vr;
                    if (!tang)
                        for (auto &o : obj) { o.sb[ij + 1].Xl = o.sb[ij].Xr; o.sb[ij + 1].Yl = o.sb[ij].Yr; }
                } else if (f.ur + fr.ul < 0 && fr.ul < 0) {
                    if (tang) f.ur = fr.ul;
                } else {
                    if (tang) f.ur = 0.5 * (f.ur + fr.ul);
                    fr.ul = 0;
                    fr.vl = 0.5 * (f.vr + fr.vl);
                    if (!tang)
                        for (auto &o : obj) {
                            o.sb[ij + 1].Xl = 0.5 * (o.sb[ij].Xr + o.sb[ij + 1].Xl);
                            o.sb[ij + 1].Yl = 0.5 * (o.sb[ij].Yr + o.sb[ij + 1].Yl);
                        }
                }
            }
            if (j < n - 1) {
                cellf &f = fb[ij], &fu = fb[ij + ml];
                if (f.vu + fu.vd > 0 && f.vu > 0) {
                    fu.vd = f.vu; fu.ud = f.uu;
                    if (!tang)
                        for (auto &o : obj) { o.sb[ij + ml].Xd = o.sb[ij].Xu; o.sb[ij + ml].Yd = o.sb[ij].Yu; }
                } else if (f.vu + fu.vd < 0 && fu.vd < 0) {
                    if (tang) f.vu = fu.vd;
                } else {
                    if (tang) f.vu = 0.5 * (f.vu + fu.vd);
                    fu.vd = 0;
                    fu.ud = 0.5 * (f.uu + fu.ud);
                    if (!tang)
                        for (auto &o : obj) {
                            o.sb[ij + ml].Xd = 0.5 * (o.sb[ij].Xu + o.sb[ij + ml].Xd);
                            o.sb[ij + ml].Yd = 0.5 * (o.sb[ij].Yu + o.sb[ij + ml].Yd);
                        }
                }
            }
        }
}

// godunov_set_lr<true> / godunov_set_du<true> on one face (fluid_2d.cc:642-686)
void fluid::set_lr_tang(int ij) {
    cellf &f = fb[ij], &fr = fb[ij + 1];
    if (f.ur + fr.ul > 0 && f.ur > 0) {
        fr.ul = f.ur; fr.vl = f.vr;
        for (auto &o : obj) { o.sb[ij + 1].Xl = o.sb[ij].Xr; o.sb[ij + 1].Yl = o.sb[ij].Yr; }
    } else if (f.ur + fr.ul < 0 && fr.ul < 0) {
        f.ur = fr.ul;
    } else {
        f.ur = 0.5 * (f.ur + fr.ul);
        fr.ul = 0;
        fr.vl = 0.5 * (f.vr + fr.vl);
        for (auto &o : obj) {
            o.sb[ij + 1].Xl = 0.5 * (o.sb[ij].Xr + o.sb[ij + 1].Xl);
            o.sb[ij + 1].Yl = 0.5 * (o.sb[ij].Yr + o.sb[ij + 1].Yl);
        }
    }
}
void fluid::set_du_tang(int ij) {
    cellf &f = fb[ij], &fu = fb[ij + ml];
    if (f.vu + fu.vd > 0 && f.vu > 0) {
        fu.vd = f.vu; fu.ud = f.uu;
        for (auto &o : obj) { o.sb[ij + ml].Xd = o.sb[ij].Xu; o.sb[ij + ml].Yd = o.sb[ij].Yu; }
    } else if (f.vu + fu.vd < 0 && fu.vd < 0) {
        f.vu = fu.vd;
    } else {
        f.vu = 0.5 * (f.vu + fu.vd);
        fu.vd = 0;
        fu.ud = 0.5 * (f.uu + fu.ud);
        for (auto &o : obj) {
            o.sb[ij + ml].Xd = 0.5 * (o.sb[ij].Xu + o.sb[ij + ml].Xd);
            o.sb[ij + ml].Yd = 0.5 * (o.sb[ij].Yu + o.sb[ij + ml].Yd);
        }
    }
}

void fluid::edge_bc() {  // fluid_2d.cc:729-779: no-slip walls or periodic seams
    const double lx = bx - ax, ly = by - ay;
    if (sp.solid_prd_bc)  // edge_boundary_conditions_solid_pre (fluid_2d.cc:784-797)
        for (auto &o : obj) {
            for (int j = 0; j < n; j++) {
                o.sb[at(-1, j)].Xr = o.sb[at(m - 1, j)].Xr - lx;
                o.sb[at(-1, j)].Yr = o.sb[at(m - 1, j)].Yr;
            }
            for (int i = 0; i < m; i++) {
                o.sb[at(i, -1)].Xu = o.sb[at(i, n - 1)].Xu;
                o.sb[at(i, -1)].Yu = o.sb[at(i, n - 1)].Yu - ly;
            }
        }
    for (int j = 0; j < n; j++) {
        if (sp.x_prd) {
            F(-1, j).ur = F(m - 1, j).ur;
            F(-1, j).vr = F(m - 1, j).vr;
            set_lr_tang(at(-1, j));
            F(m, j).ul = F(0, j).ul;
            F(m, j).vl = F(0, j).vl;
            F(m - 1, j).ur = F(-1, j).ur;
            continue;
        }
        if (sp.zpr) {  // fluid_2d.cc:745-749
            F(-1, j).ur = F(0, j).ul;
            F(m, j).ul = F(m - 1, j).ur;
            F(m, j).vl = F(m - 1, j).vr;
            continue;
        }
        F(0, j).ul = F(0, j).vl = 0;
        F(-1, j).ur = 0;
        F(m, j).ul = F(m, j).vl = 0;
        F(m - 1, j).ur = 0;
    }
    for (int i = 0; i < m; i++) {
        if (sp.y_prd) {
            F(i, -1).vu = F(i, n - 1).vu;
            F(i, -1).uu = F(i, n - 1).uu;
            set_du_tang(at(i, -1));
            F(i, n).vd = F(i, 0).vd;
            F(i, n).ud = F(i, 0).ud;
            F(i, n - 1).vu = F(i, -1).vu;
            continue;
        }
        F(i, 0).ud = F(i, 0).vd = 0;
        F(i, -1).vu = 0;
        F(i, n).ud = F(i, n).vd = 0;
        F(i, n - 1).vu = 0;
    }
    if (sp.solid_prd_bc)  // edge_boundary_conditions_solid_post (fluid_2d.cc:802-815)
        for (auto &o : obj) {
            for (int j = 0; j < n; j++) {
                o.sb[at(m, j)].Xl = o.sb[at(0, j)].Xl + lx;
                o.sb[at(m, j)].Yl = o.sb[at(0, j)].Yl;
            }
            for (int i = 0; i < m; i++) {
                o.sb[at(i, n)].Xd = o.sb[at(i, 0)].Xd;
                o.sb[at(i, n)].Yd = o.sb[at(i, 0)].Yd + ly;
            }
        }
}

void fluid::tangential() {  // fluid_2d.cc:953-975 (velocity part)
#pragma omp parallel for
    for (int j = 0; j < n; j++)
        for (int i = 0; i < m; i++) {
            cellf *fp = &F(i, j);
            cellf &f = *fp, &fl = fp[-1], &fr = fp[1], &fu = fp[ml], &fd = fp[-ml];
            double u_adv = 0.5 * xsp * (fr.ul + f.ul), v_adv = 0.5 * ysp * (fu.vd + f.vd);
            f.c0 = u_adv * (f.ur - fl.ur);
            f.c1 = u_adv * (fr.vl - f.vl);
            f.c2 = v_adv * (fu.ud - f.ud);
            f.c3 = v_adv * (f.vu - fd.vu);
        }
}

void fluid::wall_force(double x, double y, double phiv, double &fx, double &fy) {  // fluid_2d.cc:1045-1078
    const double w = w_scale * eps, winv = 1 / w, k_rep = w_force * winv;
    auto dfw = [](double q) { return (1. - q) / q; };
    if (x <= ax + w) {
        fx += k_rep * dfw((x - ax) * winv) * (phiv > -dx ? (phiv > dx ? 0 : 0.5 * (1 - phiv * xsp)) : 1);
    } else if (x >= bx - w) {
        fx -= k_rep * dfw((bx - x) * winv) * (phiv > -dx ? (phiv > dx ? 0 : 0.5 * (1 - phiv * xsp)) : 1);
    }
    if (y <= ay + w) {
        fy += k_rep * dfw((y - ay) * winv) * (phiv > -dy ? (phiv > dy ? 0 : 0.5 * (1 - phiv * ysp)) : 1);
    } else if (y >= by - w) {
        fy -= k_rep * dfw((by - y) * winv) * (phiv > -dy ? (phiv > dy ? 0 : 0.5 * (1 - phiv * ysp)) : 1);
    }
}

void fluid::acceleration(int i, int j, double &accx, double &accy, double x, double y, bool pressure) {
    // fluid_2d.cc:1019-1039
    const int ij = at(i, j);
    cellf *fp = &fb[ij];
    cellf &f = *fp, &fr = fp[1], &fu = fp[ml];
    double accx_ = 0, accy_ = 0;
    accx = xsp * (fr.s11 - f.s11) + ysp * (fu.s12 - f.s12);
    accy = xsp * (fr.s21 - f.s21) + ysp * (fu.s22 - f.s22);
    if (pressure) {
        accx -= 0.5 * xsp * (fp[ml + 1].p + fr.p - fu.p - f.p);
        accy -= 0.5 * ysp * (fp[ml + 1].p - fr.p + fu.p - f.p);
    }
    for (auto &o : obj) {
        const double phiv = o.phi[ij];
        if (phiv < o.eps && !sp.solid_prd_bc) {  // fluid_2d.cc:1037
            wall_force(x, y, phiv, accx, accy);
            o.accel(x, y, o.sb[ij].X, o.sb[ij].Y, phiv, accx_, accy_);
        }
    }
    accx /= f.rho; accy /= f.rho;
    accx += accx_;
    accy += accy_;
}

void fluid::half_time_edges(double dt) {  // fluid_2d.cc:979-1014, 1082-1108
#pragma omp parallel for
    for (int j = 0; j < n; j++) {
        double y = ay + dy * (0.5 + j);
        for (int i = 0; i < m; i++) {
            double x = ax + dx * (0.5 + i), accx, accy;
            cellf &f = F(i, j);
            acceleration(i, j, accx, accy, x, y, true);
            f.ul = f.u + 0.5 * (-dx * f.ux + dt * (-f.u * f.ux - f.c2 + accx));
            f.vl = f.v + 0.5 * (-dx * f.vx + dt * (-f.u * f.vx - f.c3 + accy));
            f.ur = f.u + 0.5 * (dx * f.ux + dt * (-f.u * f.ux - f.c2 + accx));
            f.vr = f.v + 0.5 * (dx * f.vx + dt * (-f.u * f.vx - f.c3 + accy));
            f.ud = f.u + 0.5 * (-dy * f.uy + dt * (-f.v * f.uy - f.c0 + accx));
            f.vd = f.v + 0.5 * (-dy * f.vy + dt * (-f.v * f.vy - f.c1 + accy));
            f.uu = f.u + 0.5 * (dy * f.uy + dt * (-f.v * f.uy - f.c0 + accx));
            f.vu = f.v + 0.5 * (dy * f.vy + dt * (-f.v * f.vy - f.c1 + accy));
        }
    }
    for (auto &o : obj) {
#pragma omp parallel for
        for (int j = 0; j < n; j++)
            for (int i = 0; i < m; i++) {
                scell &s = o.sb[at(i, j)];
                cellf &f = F(i, j);
                double Xdt = dt * (-f.u * s.Xx - f.v * s.Xy), Ydt = dt * (-f.u * s.Yx - f.v * s.Yy);
                s.Xl = s.X + 0.5 * (-dx * s.Xx + Xdt);
                s.Yl = s.Y + 0.5 * (-dx * s.Yx + Ydt);
                s.Xr = s.X + 0.5 * (dx * s.Xx + Xdt);
                s.Yr = s.Y + 0.5 * (dx * s.Yx + Ydt);
                s.Xd = s.X + 0.5 * (-dy * s.Xy + Xdt);
                s.Yd = s.Y + 0.5 * (-dy * s.Yy + Ydt);
                s.Xu = s.X + 0.5 * (dy * s.Xy + Xdt);
                s.Yu = s.Y + 0.5 * (dy * s.Yy + Ydt);
            }
    }
}

int fluid::mac_projection() {  // fluid_2d.cc:1175-1236, mgs_mac.cc:58-71
    for (int j = 0; j < n; j++)
        for (int i = 0; i < m; i++) {
            cellf *fp = &F(i, j);
            src[j * m + i] = xsp * (fp[1].ul - fp->ul) + ysp * (fp[ml].vd - fp->vd);
        }
    const double *ir = nullptr;
    if (!const_rho) {
        for (int j = 0; j < n; j++)
            for (int i = 0; i < m; i++) {
                const cellf *fp = &F(i, j);
                irho_mac[2 * (j * m + i)] = 0.5 / (fp->rho) + 0.5 / (fp[-1].rho);
                irho_mac[2 * (j * m + i) + 1] = 0.5 / (fp->rho) + 0.5 / (fp[-ml].rho);
            }
        int q = 2 * mn;
        for (int j = 0; j < n; j++) irho_mac[q++] = 0.5 / F(m - 1, j).rho + 0.5 / F(m, j).rho;
        for (int i = 0; i < m; i++) irho_mac[q++] = 0.5 / F(i, n - 1).rho + 0.5 / F(i, n).rho;
        ir = irho_mac.data();
    }
    mac_system s{m, n, xxsp, yysp, ir, sp.x_prd, sp.y_prd, sp.zpr};
    mac_cycles = mac_solve(s, src.data(), zmac.data(), sp.mg_tol, &mac_res);
    if (mac_cycles < 0) return 2;
    for (int j = 0; j < n; j++)
        for (int i = 0; i < m; i++) {
            cellf &f = F(i, j);
            const double *so = &zmac[j * m + i];
            const double ed = const_rho ? 1 : irho_mac[2 * (j * m + i) + 1], el = const_rho ? 1 : irho_mac[2 * (j * m + i)];
            // a periodic seam is corrected on row 0 / column 0 with the
            // opposite edge's q, and the ghost row n / column m repeats it
            if (const_rho) {
                if (j > 0) f.vd -= (*so - so[-m]) * ysp;
                else if (sp.y_prd) { f.vd -= (*so - so[mn - m]) * ysp; F(i, n).vd = f.vd; }
                if (i > 0) f.ul -= (*so - so[-1]) * xsp;
                else if (sp.x_prd) { f.ul -= (*so - so[m - 1]) * xsp; F(m, j).ul = f.ul; }
            } else {
                if (j > 0) f.vd -= ed * (*so - so[-m]) * ysp;
                else if (sp.y_prd) { f.vd -= ed * (*so - so[mn - m]) * ysp; F(i, n).vd = f.vd; }
                if (i > 0) f.ul -= el * (*so - so[-1]) * xsp;
                else if (sp.x_prd) { f.ul -= el * (*so - so[m - 1]) * xsp; F(m, j).ul = f.ul; }
            }
        }
    return 0;
}

void fluid::update_reference_map(double dt) {  // fluid_2d.cc:1242-1273
    double hx = 0.5 * dt * xsp, hy = 0.5 * dt * ysp;
    for (auto &o : obj) {
#pragma omp parallel for
        for (int j = 0; j < n; j++)
            for (int i = 0; i < m; i++) {
                const int ij = at(i, j);
                if (o.st[ij] >= 4) continue;
                cellf &f = fb[ij], &fr = fb[ij + 1], &fu = fb[ij + ml];
                scell &s = o.sb[ij], &sr = o.sb[ij + 1], &su = o.sb[ij + ml];
                double cX = -hx * (fr.ul + f.ul) * (sr.Xl - s.Xl) - hy * (fu.vd + f.vd) * (su.Xd - s.Xd);
                double cY = -hx * (fr.ul + f.ul) * (sr.Yl - s.Yl) - hy * (fu.vd + f.vd) * (su.Yd - s.Yd);
                s.nX = s.X + cX;
                s.nY = s.Y + cY;
            }
    }
    for (auto &o : obj) {
        std::vector<double> fx(ml * nl), fy(ml * nl);
        for (int q = 0; q < ml * nl; q++) { fx[q] = o.sb[q].nX; fy[q] = o.sb[q].nY; }
        extrapolate2(o.g, o.phi.data(), o.st.data(), o.band, fx.data(), fy.data());
        for (int q = 0; q < ml * nl; q++) { o.sb[q].nX = fx[q]; o.sb[q].nY = fy[q]; }
    }
    for (auto &o : obj)
        for (int j = 0; j < n; j++)
            for (int i = 0; i < m; i++) {
                const int ij = at(i, j);
                if (o.st[ij] < 6) {
                    scell &s = o.sb[ij];
                    s.X = 0.5 * (s.X + s.nX);
                    s.Y = 0.5 * (s.Y + s.nY);
                }
            }
}

int fluid::pin() {  // fluid_2d.cc:373-377, obj_field.cc:92-97
    for (auto &o : obj) {
        for (int ij = 0; ij < ml * nl; ij++) {
            double &p = o.phi[ij];
            if (p < o.pinw && p > -o.pinw) p = o.phi_of(o.sb[ij].X, o.sb[ij].Y);
        }
        build_band(o.g, o.phi.data(), o.st.data(), o.band);
    }
    if (!const_rho) set_rho();
    return 0;
}

void fluid::compute_ustar(double dt) {  // fluid_2d.cc:1278-1291
    double hx = 0.5 * dt * xsp, hy = 0.5 * dt * ysp;
#pragma omp parallel for
    for (int j = 0; j < n; j++) {
        double y = ay + dy * (.5 + j), x = ax + 0.5 * dx, accx, accy;
        for (int i = 0; i < m; i++, x += dx) {
            cellf *fp = &F(i, j);
            cellf &f = *fp, &fr = fp[1], &fu = fp[ml];
            acceleration(i, j, accx, accy, x, y, false);
            f.c0 = f.u - hx * (fr.ul + f.ul) * (fr.ul - f.ul) - hy * (fu.vd + f.vd) * (fu.ud - f.ud) + dt * accx;
            f.c1 = f.v - hx * (fr.ul + f.ul) * (fr.vl - f.vl) - hy * (fu.vd + f.vd) * (fu.vd - f.vd) + dt * accy;
        }
    }
}

int fluid::fem_projection(double dt) {  // fluid_2d.cc:1297-1313; mgs_fem.cc:104-121
    fem_source(dt);
    const double *ir = nullptr;
    if (!const_rho) {
        for (int j = 0; j < n_fem; j++)
            for (int i = 0; i < m_fem; i++) irho_fem[j * m_fem + i] = 1. / F(i - 1, j - 1).rho;
        ir = irho_fem.data();
    }
    fem_system s{m_fem, n_fem, dx, dy, ir, sp.x_prd, sp.y_prd, sp.zpr};
    fem_cycles = fem_solve(s, src.data(), zfem.data(), sp.mg_tol, &fem_res);
    return fem_cycles < 0 ? 3 : 0;
}

void fluid::fem_source(double dt) {  // fluid_2d.cc:1297-1309, 839-872
    for (int j = 0; j < n; j++) {
        if (sp.x_prd) {
            F(-1, j).c0 = F(m - 1, j).c0;
            F(-1, j).c1 = F(m - 1, j).c1;
        } else if (sp.zpr) {  // fluid_2d.cc:847-851
            F(-1, j).c0 = F(-1, j).u; F(m, j).c0 = 0;
            F(-1, j).c1 = F(-1, j).v; F(m, j).c1 = 0;
        } else {
            F(-1, j).c0 = F(m, j).c0 = 0;
            F(-1, j).c1 = F(m, j).c1 = 0;
        }
    }
    const int xl = sp.x_prd ? m : m + 1;
    for (int i = -1; i < xl; i++) {
        if (sp.y_prd) {
            F(i, -1).c0 = F(i, n - 1).c0;
            F(i, -1).c1 = F(i, n - 1).c1;
        } else {
            F(i, -1).c0 = F(i, n).c0 = 0;
            F(i, -1).c1 = F(i, n).c1 = 0;
        }
    }
    double hx = 0.5 * dy / dt, hy = 0.5 * dx / dt;
    for (int j = 0; j < n_fem; j++)
        for (int i = 0; i < m_fem; i++) {
            cellf *fp = &F(i, j);
            src[j * m_fem + i] = hx * (fp[-ml - 1].c0 + fp[-1].c0 - fp[-ml].c0 - fp->c0)
                                 + hy * (fp[-ml - 1].c1 - fp[-1].c1 + fp[-ml].c1 - fp->c1);
        }
}

void fluid::copy_pressure() {  // fluid_2d.cc:1317-1348
    double pavg = 0;
    for (int j = 0; j < n_fem; j++) {
        const double *so = &zfem[j * m_fem];
        double prow = so[0] * (sp.x_prd ? 1 : 0.5);
        for (int i = 1; i < m_fem - 1; i++) prow += so[i];
        prow += so[m_fem - 1] * (sp.x_prd ? 1 : 0.5);
        if (!sp.y_prd && (j == 0 || j == n)) prow *= 0.5;
        pavg += prow;
    }
    pavg = sp.zpr ? 0. : pavg * (1. / mn);  // fluid_2d.cc:1318
    // node m (row n) of a periodic direction is node 0 (row 0)
    for (int j = 0; j < n + 1; j++) {
        const int jr = sp.y_prd && j == n ? 0 : j;
        for (int i = 0; i < m; i++) F(i, j).p = zfem[jr * m_fem + i] - pavg;
        F(m, j).p = sp.x_prd ? F(0, j).p : zfem[jr * m_fem + m] - pavg;
    }
}

void fluid::update_velocity(double dt) {  // fluid_2d.cc:1354-1374
#pragma omp parallel for
    for (int j = 0; j < n; j++)
        for (int i = 0; i < m; i++) {
            cellf *fp = &F(i, j);
            const double rhoinv = 1. / fp->rho, u_star = fp->c0, v_star = fp->c1;
            fp->u = u_star - 0.5 * dt * rhoinv * xsp * (fp[ml + 1].p + fp[1].p - fp[ml].p - fp->p);
            fp->v = v_star - 0.5 * dt * rhoinv * ysp * (fp[ml + 1].p - fp[1].p + fp[ml].p - fp->p);
        }
    for (auto &o : obj)
        for (int j = 0; j < n; j++)
            for (int i = 0; i < m; i++) {
                scell &s = o.sb[at(i, j)];
                s.X = s.nX; s.Y = s.nY;
            }
}

void fluid::init_tracers(int nt) {  // fluid_2d.cc:1376-1386
    ntrace = nt;
    tm.assign(4 * (size_t)nt, 0.0);
    for (size_t q = 0; q < tm.size();) {
        double tx = rng.rnd(ax, bx);
        double ty = rng.rnd(ay, by);
        if (!inside_object(tx, ty)) { tm[q] = tx; tm[q + 1] = ty; q += 4; }
    }
}

bool fluid::inside_object(double tx, double ty) {  // fluid_2d.cc:1393-1430
    double x = (tx - ax) * xsp - 0.5, y = (ty - ay) * ysp - 0.5;
    int i = static_cast<int>(x), j = static_cast<int>(y);
    if (i < -1) i = -1; else if (i > m - 1) i = m - 1;
    if (j < -1) j = -1; else if (j > n - 1) j = n - 1;
    x -= i; y -= j;
    const double fx = 1 - x, fy = 1 - y;
    for (auto &o : obj) {
        const double *p = o.phi.data() + at(i, j);
        if (fy * (*p * fx + p[1] * x) + y * (p[ml] * fx + p[ml + 1] * x) < o.eps) return true;
    }
    return false;
}

v2 fluid::vel(double x, double y) {  // fluid_2d.cc:43, 1435-1438
    const cellf *f = fb.data();
    auto b = make_bicubic(ml, nl, ax - 1.5 * dx, bx + 1.5 * dx, ay - 1.5 * dy, by + 1.5 * dy,
                          [f](int ij) { return v2{f[ij].u, f[ij].v}; });
    return b.f(x, y);
}

void fluid::tracer_update_1(double *tp) {  // fluid_2d.cc:1445-1447
    v2 v = vel(tp[0], tp[1]);
    tp[2] = v.x; tp[3] = v.y;
}

void fluid::tracer_update_2(double *tp, double dt) {  // fluid_2d.cc:1454-1459
    double tx = tp[0] + dt * tp[2], ty = tp[1] + dt * tp[3];
    v2 v = vel(tx, ty);
    tp[0] += 0.5 * dt * (tp[2] + v.x);
    tp[1] += 0.5 * dt * (tp[3] + v.y);
}

void fluid::tracer_remap(double &x, double &y) {  // sim_type.cc:12-17
    if (x < ax) x += (bx - ax), y = rng.rnd(ay, by);
    if (x > bx) x -= (bx - ax), y = rng.rnd(ay, by);
    if (y < ay) y += (by - ay), x = rng.rnd(ax, bx);
    if (y > by) y -= (by - ay), x = rng.rnd(ax, bx);
}

bool fluid::correct_tracer(solid &o, double xref, double yref, double &xx, double &yy) {  // object.cc:383-417
    const double tolnewt = small_number * small_number * 100, bigstep = 1;
    const scell *sb = o.sb.data();
    auto b = make_bicubic(ml, nl, ax - 1.5 * dx, bx + 1.5 * dx, ay - 1.5 * dy, by + 1.5 * dy,
                          [sb](int ij) { return v2{sb[ij].X, sb[ij].Y}; });
    int count = 0, iter = 0;
    while (count < 4) {
        if (iter > 100) return false;
        v2 fx, fy, f = b.f_grad_f(xx, yy, fx, fy);
        f = f - v2{xref, yref};
        if (f.x * f.x + f.y * f.y < tolnewt) count++;
        const double det = fx.x * fy.y - fy.x * fx.y;
        if (det < small_number && det > -small_number) return false;
        const double delx = (fy.y * f.x - fy.x * f.y) / det, dely = (-fx.y * f.x + fx.x * f.y) / det;
        if (delx * delx + dely * dely > bigstep) return false;
        xx -= delx; yy -= dely; iter++;
    }
    return true;
}

void fluid::track_output(solid &o) {  // obj_simple_spin::file_output (object.cc:202-219)
    const double ang = std::atan2(o.tk[1], o.tk[0]);
    if (o.t_correction && !correct_tracer(o, o.tk[4], o.tk[5], o.tk[0], o.tk[1]))  // object.cc:367-377
        o.t_correction = false;
    const int q = std::fabs(ang) > 0.5 * M_PI ? (ang > 0 ? 1 : -1) : 0;
    if (q == 1 && o.prev_q == -1) o.wind--;
    else if (q == -1 && o.prev_q == 1) o.wind++;
    o.prev_q = q;
    o.rec.push_back(time);
    o.rec.push_back(o.tk[0]);
    o.rec.push_back(o.tk[1]);
    o.rec.push_back(ang + 2 * M_PI * o.wind);
    o.count = 1;
}

void fluid::frame_output() {  // the object part of write_files (fluid_2d_io.cc:101), obj_track::write
    for (auto &o : obj)
        if (o.has_track) track_output(o);
}

void fluid::update_tracers(int stage, double dt) {  // fluid_2d.cc:1464-1484, object.cc:351-362
    if (stage == 1) {
        for (int q = 0; q < ntrace; q++) tracer_update_1(&tm[4 * q]);
        for (auto &o : obj)
            if (o.has_track) tracer_update_1(o.tk);
    } else {
        for (int q = 0; q < ntrace; q++) {
            tracer_update_2(&tm[4 * q], dt);
            tracer_remap(tm[4 * q], tm[4 * q + 1]);
        }
        for (auto &o : obj)
            if (o.has_track) {
                tracer_update_2(o.tk, dt);
                if (o.count == o.every) track_output(o);
                else o.count++;
            }
    }
}

int fluid::step(double dt) {  // fluid_2d.cc:298-369
    double ntime = time + dt;
    int rc;
    update_tracers(1, dt);
    normal_derivatives(dt);
    godunov(true);
    edge_bc();
    tangential();
    half_time_edges(dt);
    godunov(false);
    edge_bc();
    time += 0.5 * dt;
    if ((rc = mac_projection())) return rc;
    update_reference_map(dt);
    if ((rc = pin())) return rc;
    set_boundaries();
    compute_stress();
    compute_ustar(dt);
    time = ntime;
    if ((rc = fem_projection(dt))) return rc;
    copy_pressure();
    update_velocity(dt);
    if ((rc = pin())) return rc;
    set_boundaries();
    compute_stress();
    update_tracers(2, dt);
    return 0;
}

void fluid_init_tracers(fluid *f, int ntrace) { f->init_tracers(ntrace); }
void fluid_frame_output(fluid *f) { f->frame_output(); }
int fluid_ntrace(fluid *f) { return f->ntrace; }
void fluid_get_tracers(fluid *f, double *out) { std::memcpy(out, f->tm.data(), sizeof(double) * f->tm.size()); }
void fluid_set_tracers(fluid *f, const double *in) { std::memcpy(f->tm.data(), in, sizeof(double) * f->tm.size()); }
// track state of object k: x, y, stage u, stage v, X, Y, count, prev_q, wind, t_correction
int fluid_get_track(fluid *f, int k, double *st10) {
    if (k < 0 || k >= (int)f->obj.size() || !f->obj[k].has_track) return 1;
    const solid &o = f->obj[k];
    for (int q = 0; q < 6; q++) st10[q] = o.tk[q];
    st10[6] = o.count; st10[7] = o.prev_q; st10[8] = o.wind; st10[9] = o.t_correction;
    return 0;
}
int fluid_set_track(fluid *f, int k, const double *st10) {
    if (k < 0 || k >= (int)f->obj.size() || !f->obj[k].has_track) return 1;
    solid &o = f->obj[k];
    for (int q = 0; q < 6; q++) o.tk[q] = st10[q];
    o.count = (int)st10[6]; o.prev_q = (int)st10[7]; o.wind = (int)st10[8]; o.t_correction = st10[9] != 0;
    return 0;
}
int fluid_track_records(fluid *f, int k, double *out, int cap) {
    if (k < 0 || k >= (int)f->obj.size()) return -1;
    const std::vector<double> &r = f->obj[k].rec;
    const int nrec = (int)(r.size() / 4);
    for (int q = 0; q < 4 * std::min(nrec, cap); q++) out[q] = r[q];
    return nrec;
}
void interp2(int m, int n, double ax, double bx, double ay, double by, const double *fx, const double *fy, double x,
             double y, double *out6) {
    auto b = make_bicubic(m, n, ax, bx, ay, by, [fx, fy](int ij) { return v2{fx[ij], fy[ij]}; });
    v2 gx, gy, v = b.f_grad_f(x, y, gx, gy);
    v2 w = b.f(x, y);
    out6[0] = w.x; out6[1] = w.y; out6[2] = gx.x; out6[3] = gx.y; out6[4] = gy.x; out6[5] = gy.y;
    (void)v;
}
int glibc_rand_probe(int n, int *out) {
    glibc_rand r;
    for (int q = 0; q < n; q++) out[q] = r.next();
    return n;
}

double mono_diff_probe(double f0, double f1, double f2, double f3, double f4) {
    return fluid::mono_diff(f0, f1, f2, f3, f4);
}

void object_eval(const object_spec &o, double eps, double time, double X, double Y, double *phi, double *F4) {
    const ls_grid g(4, 4, 0, 0, 1, 1);
    solid so(g);
    setup_solid(so, o, eps);
    so.pre_stress_setup(time);
    *phi = so.phi_of(X, Y);
    so.actuate(X, Y, F4[0], F4[1], F4[2], F4[3]);
}

fluid *fluid_create(const sim_spec &sp, const std::vector<object_spec> &objs) { return new fluid(sp, objs); }
void fluid_destroy(fluid *f) { delete f; }
double fluid_dt_reg(fluid *f) { return f->dt_reg; }
double fluid_time(fluid *f) { return f->time; }
int fluid_step(fluid *f, double dt) { return f->step(dt); }
void fluid_solver_stats(fluid *f, double *o) {
    o[0] = f->mac_cycles; o[1] = f->mac_res; o[2] = f->fem_cycles; o[3] = f->fem_res;
}

static bool field_ptr(fluid *f, const std::string &nm, int k, size_t &off, bool &solidf, bool &isint) {
    solidf = false; isint = false;
    if (nm == "u") off = offsetof(cellf, u);
    else if (nm == "v") off = offsetof(cellf, v);
    else if (nm == "p") off = offsetof(cellf, p);
    else if (nm == "rho") off = offsetof(cellf, rho);
    else if (nm == "s11") off = offsetof(cellf, s11);
    else if (nm == "s12") off = offsetof(cellf, s12);
    else if (nm == "s21") off = offsetof(cellf, s21);
    else if (nm == "s22") off = offsetof(cellf, s22);
    else if (nm == "c0") off = offsetof(cellf, c0);
    else if (nm == "c1") off = offsetof(cellf, c1);
    else if (nm == "ul") off = offsetof(cellf, ul);
    else if (nm == "vd") off = offsetof(cellf, vd);
    else if (nm == "vl") off = offsetof(cellf, vl);
    else if (nm == "ur") off = offsetof(cellf, ur);
    else if (nm == "vr") off = offsetof(cellf, vr);
    else if (nm == "ud") off = offsetof(cellf, ud);
    else if (nm == "uu") off = offsetof(cellf, uu);
    else if (nm == "vu") off = offsetof(cellf, vu);
    else if (nm == "ux") off = offsetof(cellf, ux);
    else if (nm == "uy") off = offsetof(cellf, uy);
    else if (nm == "vx") off = offsetof(cellf, vx);
    else if (nm == "vy") off = offsetof(cellf, vy);
    else if (nm == "c2") off = offsetof(cellf, c2);
    else if (nm == "c3") off = offsetof(cellf, c3);
    else {
        if (k < 0 || k >= (int)f->obj.size()) return false;
        solidf = true;
        if (nm == "X") off = offsetof(scell, X);
        else if (nm == "Y") off = offsetof(scell, Y);
        else if (nm == "nX") off = offsetof(scell, nX);
        else if (nm == "nY") off = offsetof(scell, nY);
        else if (nm == "Xx") off = offsetof(scell, Xx);
        else if (nm == "Xy") off = offsetof(scell, Xy);
        else if (nm == "Yx") off = offsetof(scell, Yx);
        else if (nm == "Yy") off = offsetof(scell, Yy);
        else if (nm == "Xl") off = offsetof(scell, Xl);
        else if (nm == "Xr") off = offsetof(scell, Xr);
        else if (nm == "Xd") off = offsetof(scell, Xd);
        else if (nm == "Xu") off = offsetof(scell, Xu);
        else if (nm == "Yl") off = offsetof(scell, Yl);
        else if (nm == "Yr") off = offsetof(scell, Yr);
        else if (nm == "Yd") off = offsetof(scell, Yd);
        else if (nm == "Yu") off = offsetof(scell, Yu);
        else if (nm == "phi") off = (size_t)-1;
        else if (nm == "s") { off = (size_t)-1; isint = true; }
        else return false;
    }
    return true;
}

// One phase of fluid_2d::step_forward on the current state, so that tests
// can pin a single sweep by a closed form. The names and the reference
// code each runs:
//   normal_derivatives  fluid_2d.cc:877-950      godunov_tang  :642-686 (tang)
//   edge_bc             :729-779                 tangential    :953-975
//   half_time_edges     :979-1108                godunov       :1112-1121
//   mac_projection      :1175-1236               update_reference_map :1242-1273
//   stress              :381-600                 ustar         :1278-1291
//   fem_source          :1297-1309, 839-872 (out: the m_fem x n_fem source)
//   copy_pressure       :1317-1348 (in: the m_fem x n_fem solution)
//   update_velocity     :1354-1374               set_boundaries :689-726
//   set_rho             :1124-1137
// Returns 0, or 1 for an unknown name.
int fluid_phase(fluid *f, const char *name, double dt, const double *in, double *out) {
    const std::string nm(name);
    if (nm == "normal_derivatives") f->normal_derivatives(dt);
    else if (nm == "godunov_tang") f->godunov(true);
    else if (nm == "edge_bc") f->edge_bc();
    else if (nm == "tangential") f->tangential();
    else if (nm == "half_time_edges") f->half_time_edges(dt);
    else if (nm == "godunov") f->godunov(false);
    else if (nm == "mac_projection") return f->mac_projection() ? 2 : 0;
    else if (nm == "update_reference_map") f->update_reference_map(dt);
    else if (nm == "stress") f->compute_stress();
    else if (nm == "ustar") f->compute_ustar(dt);
    else if (nm == "fem_source") {
        f->fem_source(dt);
        std::memcpy(out, f->src.data(), sizeof(double) * f->m_fem * f->n_fem);
    } else if (nm == "copy_pressure") {
        std::memcpy(f->zfem.data(), in, sizeof(double) * f->m_fem * f->n_fem);
        f->copy_pressure();
    } else if (nm == "update_velocity") f->update_velocity(dt);
    else if (nm == "set_boundaries") f->set_boundaries();
    else if (nm == "set_rho") f->set_rho();
    else return 1;
    return 0;
}

void fluid_set_time(fluid *f, double t) { f->time = t; }

int fluid_get(fluid *f, const char *name, int k, double *out) {
    size_t off; bool so, isint;
    if (!field_ptr(f, name, k, off, so, isint)) return 1;
    const int N = f->ml * f->nl;
    for (int q = 0; q < N; q++) {
        if (!so) out[q] = *reinterpret_cast<const double *>(reinterpret_cast<const char *>(&f->fb[q]) + off);
        else if (isint) out[q] = f->obj[k].st[q];
        else if (off == (size_t)-1) out[q] = f->obj[k].phi[q];
        else out[q] = *reinterpret_cast<const double *>(reinterpret_cast<const char *>(&f->obj[k].sb[q]) + off);
    }
    return 0;
}

int fluid_set(fluid *f, const char *name, int k, const double *in) {
    size_t off; bool so, isint;
    if (!field_ptr(f, name, k, off, so, isint)) return 1;
    const int N = f->ml * f->nl;
    for (int q = 0; q < N; q++) {
        if (!so) *reinterpret_cast<double *>(reinterpret_cast<char *>(&f->fb[q]) + off) = in[q];
        else if (isint) f->obj[k].st[q] = (int)in[q];
        else if (off == (size_t)-1) f->obj[k].phi[q] = in[q];
        else *reinterpret_cast<double *>(reinterpret_cast<char *>(&f->obj[k].sb[q]) + off) = in[q];
    }
    return 0;
}

}  // namespace orc
