// Oracle restatement of the reference's narrow-band level set: bicubic
// closest-point Newton, second-order heap fast marching, and the
// least-squares extension of the reference map. TEST INFRASTRUCTURE ONLY.
//
// Pinned bit-for-bit against the reference's own level-set library compiled
// from its sources (oracle/_ref, tests/test_oracle_levelset.py) and against
// closed forms (signed distance to a circle; linear fields extend exactly).
#include <cmath>
#include <cstdlib>
#include <algorithm>
#include "oracle.hh"

namespace orc {

ls_grid::ls_grid(int m_, int n_, double ax_, double ay_, double dx_, double dy_,
                 double pd, double pu)
    : m(m_), n(n_), mn(m_ * n_), ax(ax_), ay(ay_), dx(dx_), dy(dy_),
      xsp(1 / dx_), ysp(1 / dy_), phi_down(pd), phi_up(pu) {}

namespace {

// The compiled reference binds abs(double) to C's int abs(int)
// (levelset.cc:124,126 under g++ -ansi): the value is truncated to int first.
inline double int_abs(double v) { return static_cast<double>(std::abs(static_cast<int>(v))); }

// One row of the bicubic: cubic Hermite in x through f0 (left node) and f1
// (right node) with end slopes pl, pr, and its derivative in e. The three
// variants are the interior form and the one-sided forms used at the grid
// edges (built/bicubic.cc). Evaluation order follows the reference so the
// oracle agrees with it bit-for-bit.
struct xrow { double v, d; };
inline xrow herm_mid(double pl, double f0, double f1, double pr, double e, double me) {
    xrow r;
    r.v = pl * e * me * me + f0 * me * me * (1 + 2 * e) + f1 * e * e * (3 - 2 * e) - pr * e * e * me;
    r.d = pl * me * (1 - 3 * e) - f0 * 6 * e * me + f1 * 6 * e * me + pr * e * (3 * e - 2);
    return r;
}
inline xrow herm_left(double f0, double f1, double pr, double e, double me) {
    xrow r;
    r.v = f0 * me * me + f1 * e * (2 - e) - pr * e * me;
    r.d = -f0 * 2 * me + f1 * 2 * me + pr * (2 * e - 1);
    return r;
}
inline xrow herm_right(double pl, double f0, double f1, double e, double me) {
    xrow r;
    r.v = pl * e * me + f0 * (1 - e * e) + f1 * e * e;
    r.d = pl * (1 - 2 * e) - f0 * 2 * e + f1 * 2 * e;
    return r;
}

}  // namespace

double bicubic(const ls_grid &g, const double *phi, double x, double y,
               double &phix, double &phiy) {
    const int m = g.m, n = g.n;
    const double dxf = 0.5, dyf = 0.5, d2f = 0.25;
    double e = (x - g.ax) * g.xsp, f = (y - g.ay) * g.ysp;
    int i = int(e), j = int(f);
    // x-position class: 0 left edge, 1 interior, 2 right edge
    int xc, yc, ic, jc;
    if (i < 1) { xc = 0; ic = 0; }
    else if (i >= m - 2) { xc = 2; ic = m - 2; e -= double(m - 2); }
    else { xc = 1; ic = i; e -= double(i); }
    if (j < 1) { yc = 0; jc = 0; }
    else if (j >= n - 2) { yc = 2; jc = n - 2; f -= double(n - 2); }
    else { yc = 1; jc = j; f -= double(j); }
    const double me = 1 - e, mf = 1 - f;
    const double *p = phi + (jc * m + ic);

    // Hermite rows: r0 = y-slopes on row j, r1 = values on row j,
    // r2 = values on row j+1, r3 = y-slopes on row j+1.
    auto row = [&](double l, double v0, double v1, double r) -> xrow {
        return xc == 1 ? herm_mid(l, v0, v1, r, e, me)
             : xc == 0 ? herm_left(v0, v1, r, e, me)
                       : herm_right(l, v0, v1, e, me);
    };
    double plbx = 0, pltx = 0, prbx = 0, prtx = 0;
    if (xc != 0) {
        plbx = (p[1] - p[-1]) * dxf;
        pltx = (p[m + 1] - p[m - 1]) * dxf;
    }
    if (xc != 2) {
        prbx = (p[2] - p[0]) * dxf;
        prtx = (p[m + 2] - p[m]) * dxf;
    }
    xrow r1 = row(plbx, p[0], p[1], prbx);
    xrow r2 = row(pltx, p[m], p[m + 1], prtx);
    xrow r0 = {0, 0}, r3 = {0, 0};
    if (yc != 0) {
        double plby = (p[m] - p[-m]) * dyf, prby = (p[m + 1] - p[-m + 1]) * dyf;
        double plbxy = 0, prbxy = 0;
        if (xc != 0) plbxy = (p[m + 1] - p[-m + 1] - p[m - 1] + p[-m - 1]) * d2f;
        if (xc != 2) prbxy = (p[m + 2] - p[-m + 2] - p[m] + p[-m]) * d2f;
        r0 = row(plbxy, plby, prby, prbxy);
    }
    if (yc != 2) {
        double plty = (p[2 * m] - p[0]) * dyf, prty = (p[2 * m + 1] - p[1]) * dyf;
        double pltxy = 0, prtxy = 0;
        if (xc != 0) pltxy = (p[2 * m + 1] - p[1] - p[2 * m - 1] + p[-1]) * d2f;
        if (xc != 2) prtxy = (p[2 * m + 2] - p[2] - p[2 * m] + p[0]) * d2f;
        r3 = row(pltxy, plty, prty, prtxy);
    }
    if (yc == 1) {
        phix = g.xsp * (f * mf * mf * r0.d + mf * mf * (1 + 2 * f) * r1.d
                        + f * f * (3 - 2 * f) * r2.d - f * f * mf * r3.d);
        phiy = g.ysp * (mf * (1 - 3 * f) * r0.v - 6 * f * mf * r1.v + 6 * f * mf * r2.v
                        + f * (3 * f - 2) * r3.v);
        return f * mf * mf * r0.v + mf * mf * (1 + 2 * f) * r1.v + f * f * (3 - 2 * f) * r2.v
               - f * f * mf * r3.v;
    } else if (yc == 0) {
        phix = g.xsp * (mf * mf * r1.d + f * (2 - f) * r2.d - f * mf * r3.d);
        phiy = g.ysp * (-2 * mf * r1.v + 2 * mf * r2.v + (2 * f - 1) * r3.v);
        return mf * mf * r1.v + f * (2 - f) * r2.v - f * mf * r3.v;
    } else {
        phix = g.xsp * (f * mf * r0.d + (1 - f * f) * r1.d + f * f * r2.d);
        phiy = g.ysp * ((1 - 2 * f) * r0.v - 2 * f * r1.v + 2 * f * r2.v);
        return f * mf * r0.v + (1 - f * f) * r1.v + f * f * r2.v;
    }
}

bool newton_failsafe(const ls_grid &g, const double *phi, const int *s, int i,
                     int ij, double x0, double y0, double &x, double &y) {
    // First-order crossing fractions along the axes to the opposite
    // interface layer (levelset.cc:137-195).
    const int m = g.m, mn = g.mn, ts = 7 - s[ij];
    const double cphi = phi[ij];
    bool hf = false, vf = false;
    double fx = 0, fy = 0, t;
    if (i > 0 && s[ij - 1] == ts) {
        fx = cphi / (phi[ij - 1] - cphi);
        if (i < m - 1 && s[ij + 1] == ts) {
            t = cphi / (cphi - phi[ij + 1]);
            if (t < -fx) fx = t;
        }
        hf = true;
    } else if (i < m - 1 && s[ij + 1] == ts) {
        fx = cphi / (cphi - phi[ij + 1]);
        hf = true;
    }
    if (ij >= m && s[ij - m] == ts) {
        fy = cphi / (phi[ij - m] - cphi);
        if (ij < mn - m && s[ij + m] == ts) {
            t = cphi / (cphi - phi[ij + m]);
            if (t < -fy) fy = t;
        }
        vf = true;
    } else if (ij < mn - m && s[ij + m] == ts) {
        fy = cphi / (cphi - phi[ij + m]);
        vf = true;
    }
    if (hf && vf) {
        double yy = fy * g.dy, tx = fx * g.dx;
        double c = tx * yy / (tx * tx + yy * yy);
        x = yy * c + x0;
        y = tx * c + y0;
        return true;
    }
    if (hf) { x = fx * g.dx + x0; y = y0; return true; }
    if (vf) { x = x0; y = fy * g.dy + y0; return true; }
    return false;
}

bool newton(const ls_grid &g, const double *phi, const int *s, int i, int ij,
            double &x0, double &y0, double &x, double &y) {
    // levelset.cc:107-127; y0 uses (ij-i)*(dy/m), not j*dy.
    const double dyy = g.dy / double(g.m), step_limit = std::sqrt(g.dx * g.dx + g.dy * g.dy);
    int iter = 0, cutoff = 0;
    double ph, px, py;
    x0 = x = g.ax + i * g.dx;
    y0 = y = g.ay + (ij - i) * dyy;
    ph = bicubic(g, phi, x, y, px, py);
    do {
        double delx = x0 - x, dely = y0 - y;
        double m0 = delx * px + dely * py + ph;
        double m1 = px * px + py * py;
        iter++;
        m0 /= m1;
        if (iter > 100 || !std::isfinite(m0) || m0 > step_limit)
            return newton_failsafe(g, phi, s, i, ij, x0, y0, x, y);
        x += delx - m0 * px;
        y += dely - m0 * py;
        ph = bicubic(g, phi, x, y, px, py);
        if (int_abs(ph) < 1e-10) cutoff++; else cutoff = 0;
    } while (cutoff < 8);
    if (int_abs(x0 - x) < 3 * g.dx && int_abs(y0 - y) < 3 * g.dy) return true;
    return newton_failsafe(g, phi, s, i, ij, x0, y0, x, y);
}

namespace {

// ---- the reference's in-place quicksort of band indices by phi
// (sorting.cc:3-63); its tie order decides heap insertion order.
void qsort_band(const double *phi, int *a, int k);
void qsort_split(const double *phi, int *a, int k) {
    k--;
    const int pv = k >> 1;
    int x = a[pv], i = 0, j = k - 1, h;
    const double px = phi[x];
    a[pv] = a[k];
    while (phi[a[i]] < px && i < j - 1) i++;
    while (phi[a[j]] > px && i + 1 < j) j--;
    while (i + 1 < j) {
        h = a[i]; a[i++] = a[j]; a[j--] = h;
        while (phi[a[i]] < px && i < j - 1) i++;
        while (phi[a[j]] > px && i + 1 < j) j--;
    }
    if (i == j) {
        if (phi[a[j]] < px) j++;
        a[k] = a[j]; a[j] = x;
    } else if (phi[a[i]] < px) {
        if (phi[a[j]] < px) j++;
        a[k] = a[j]; a[j] = x;
    } else if (phi[a[j]] < px) {
        a[k] = a[i]; a[i] = a[j]; a[j] = x;
    } else {
        a[k] = a[i]; a[i] = x; j = i;
    }
    qsort_band(phi, a, j);
    qsort_band(phi, a + (j + 1), k - j);
}
void qsort_band(const double *phi, int *a, int k) {
    if (k < 3) {
        if (k == 2 && phi[a[0]] > phi[a[1]]) std::swap(a[0], a[1]);
        return;
    }
    if (k > 3) { qsort_split(phi, a, k); return; }
    const int a1 = a[0], a2 = a[1], a3 = a[2];
    const double p1 = phi[a1], p2 = phi[a2], p3 = phi[a3];
    if (p1 < p2) {
        if (p3 < p2) {
            if (p1 < p3) a[1] = a3;
            else { a[0] = a3; a[1] = a1; }
            a[2] = a2;
        }
    } else if (p3 > p1) {
        a[0] = a2; a[1] = a1; a[2] = a3;
    } else {
        if (p2 > p3) a[0] = a3;
        else { a[0] = a2; a[1] = a3; }
        a[2] = a1;
    }
}

// ---- second-order fast marching on one side of the interface
// (heap.cc). POS marches outward (phi increasing), NEG inward.
template <bool POS>
struct marcher {
    const ls_grid &g;
    double *phi;
    int *s;
    std::vector<int> hp;       // 1-based binary heap of grid indices
    std::vector<unsigned> bp;  // back pointers: heap slot of each trial point
    std::vector<int> &acc;     // accepted points, in acceptance order
    int w;
    const double dx2, dy2, dxbydy, dx2bydy2;
    // direction-dependent primitives (heap.hh:59-83)
    static bool cm(double a, double b) { return POS ? a < b : a > b; }
    static double ct(double a, double b) { return POS ? a + b : a - b; }
    bool in(int ij) const { return POS ? s[ij] < 6 : s[ij] > 1; }
    static const int st_acc = POS ? 5 : 2, st_trial = POS ? 6 : 1, st_far = POS ? 7 : 0;

    marcher(const ls_grid &g_, double *phi_, int *s_, std::vector<int> &acc_)
        : g(g_), phi(phi_), s(s_), hp(1, 0), bp(g_.mn, 0), acc(acc_), w(1),
          dx2(g_.dx * g_.dx), dy2(g_.dy * g_.dy), dxbydy(g_.dx / g_.dy),
          dx2bydy2(dxbydy * dxbydy) {}

    // Upwind neighbour value, second order when the next point out is known
    // and further upwind (heap.cc:442-487).
    bool look(int ij, int d1, int d2ok, double &e) const {
        if (d2ok && in(ij + 2 * d1) && cm(phi[ij + 2 * d1], phi[ij + d1])) {
            e = 2 * phi[ij + d1] - 0.5 * phi[ij + 2 * d1];
            return true;
        }
        e = phi[ij + d1];
        return false;
    }
    bool look_l(int i, int ij, double &e) const { return look(ij, -1, i > 1, e); }
    bool look_r(int i, int ij, double &e) const { return look(ij, 1, i < g.m - 2, e); }
    bool look_d(int ij, double &e) const { return look(ij, -g.m, ij >= 2 * g.m, e); }
    bool look_u(int ij, double &e) const { return look(ij, g.m, ij < g.mn - 2 * g.m, e); }

    double corner_bail(bool h, bool v, double a, double b) const {
        const double ti = 2.0 / 3.0;
        double ht = h ? ct(a, g.dx) * ti : ct(a, g.dx);
        double vt = v ? ct(b, g.dy) * ti : ct(b, g.dy);
        return cm(ht, vt) ? ht : vt;
    }
    // Two-direction quadratic update with first/second-order stencils in
    // each direction, falling back to one direction when the two-direction
    // solution is not upwind (heap.cc:349-380).
    double corner(bool h, bool v, double a, double b) const {
        const double f2a = 2.25 * (dy2 + dx2), f2b = 2.25 * dy2 + dx2,
                     f2c = dy2 + 2.25 * dx2, f2d = dy2 + dx2;
        const double b1 = 2.25 * dx2 > dy2 ? 1 + 2.25 * dx2bydy2 : 1 + dy2 / (2.25 * dx2);
        const double b2 = dx2 > dy2 ? 1 + dx2bydy2 : 1 + dy2 / dx2;
        const double b3 = dx2 > 2.25 * dy2 ? 1 + dx2bydy2 / 2.25 : 1 + 2.25 * dy2 / dx2;
        double f;
        if (h && v) {
            f = 1.5 * (a - b); f *= f;
            return f * b2 < f2a ? ct(a * 1.5 + b * 1.5 * dx2bydy2, dxbydy * std::sqrt(f2a - f)) * (dy2 / f2a)
                                : corner_bail(true, true, a, b);
        } else if (h) {
            f = a - 1.5 * b; f *= f;
            return f * b3 < f2b ? ct(a * 1.5 + b * dx2bydy2, dxbydy * std::sqrt(f2b - f)) * (dy2 / f2b)
                                : corner_bail(true, false, a, b);
        } else if (v) {
            f = 1.5 * a - b; f *= f;
            return f * b1 < f2c ? ct(a + b * 1.5 * dx2bydy2, dxbydy * std::sqrt(f2c - f)) * (dy2 / f2c)
                                : corner_bail(false, true, a, b);
        }
        f = a - b; f *= f;
        return f * b2 < f2d ? ct(a + b * dx2bydy2, dxbydy * std::sqrt(f2d - f)) * (dy2 / f2d)
                            : corner_bail(false, false, a, b);
    }
    void clamp0(double &v) const { if (cm(v, 0)) v = 0; }
    bool small(double gg, double hh) const { return gg * gg + hh * hh * dx2bydy2 < dx2; }
    // (heap.cc:264-292) a,b first-order (h,v); c,d second-order (h,v)
    double fourway(double a, double b, double c, double d) const {
        double e = 2 * (c - a), f = 2 * (d - b), gg, hh;
        if (cm(e, f)) {
            gg = e - a; hh = e - b; clamp0(gg); clamp0(hh);
            if (small(gg, hh)) {
                gg = 1.5 * f - c; hh = f - b; clamp0(gg); clamp0(hh);
                return small(gg, hh) ? corner(true, true, c, d) : corner(true, false, c, b);
            }
            return corner(false, false, a, b);
        }
        gg = f - a; hh = f - b; clamp0(gg); clamp0(hh);
        if (small(gg, hh)) {
            gg = e - a; hh = 1.5 * e - d; clamp0(gg); clamp0(hh);
            return small(gg, hh) ? corner(true, true, c, d) : corner(false, true, a, d);
        }
        return corner(false, false, a, b);
    }
    // (heap.cc:295-321) two horizontal candidates a (left) c (right), one vertical b
    double threeway1(bool h1, bool h2, bool v, double a, double c, double b) const {
        if (h1 && h2) return corner(true, v, cm(a, c) ? a : c, b);
        if (!h1 && !h2) return corner(false, v, cm(a, c) ? a : c, b);
        double d, e, f;
        if (h1) {
            d = 2 * (a - c); e = d - c; f = (v ? 1.5 * d : d) - b;
            clamp0(e); clamp0(f);
            return e * e + f * f * dx2bydy2 < dx2 ? corner(true, v, a, b) : corner(false, v, c, b);
        }
        d = 2 * (c - a); e = d - a; f = (v ? 1.5 * d : d) - b;
        clamp0(e); clamp0(f);
        return e * e + f * f * dx2bydy2 < dx2 ? corner(true, v, c, b) : corner(false, v, a, b);
    }
    // (heap.cc:324-346) one horizontal candidate b, two vertical a, c
    double threeway2(bool h, bool v1, bool v2, double b, double a, double c) const {
        if (v1 && v2) return corner(h, true, b, cm(a, c) ? a : c);
        if (!v1 && !v2) return corner(h, false, b, cm(a, c) ? a : c);
        double d, e, f;
        if (v1) {
            d = 2 * (a - c); e = d - c; f = (h ? 1.5 * d : d) - b;
            clamp0(e); clamp0(f);
            return f * f + e * e * dx2bydy2 < dx2 ? corner(h, true, b, a) : corner(h, false, b, c);
        }
        d = 2 * (c - a); e = d - a; f = (h ? 1.5 * d : d) - b;
        clamp0(e); clamp0(f);
        return f * f + e * e * dx2bydy2 < dx2 ? corner(h, true, b, c) : corner(h, false, b, a);
    }
    // Local eikonal update from the known neighbours (heap.cc:177-262).
    double calc(int i, int ij) const {
        const int m = g.m, mn = g.mn;
        const double ti = 2.0 / 3.0;
        double a, b, c, d;
        const bool hd = ij >= m && in(ij - m), hu = ij < mn - m && in(ij + m);
        const bool hl = i > 0 && in(ij - 1), hr = i < m - 1 && in(ij + 1);
        if (hd) {
            if (hu) {
                if (hr) {
                    if (hl) {
                        bool sd = look_d(ij, a), su = look_u(ij, c);
                        if (sd && su) {
                            d = cm(a, c) ? a : c;
                            bool l2 = look_l(i, ij, a), r2 = look_r(i, ij, c);
                            return threeway1(l2, r2, true, a, c, d);
                        }
                        if (sd) {  // down second order, up first order (value c)
                            bool l2 = look_l(i, ij, b), r2 = look_r(i, ij, d);
                            if (l2 && r2) return threeway2(true, true, false, cm(b, d) ? b : d, a, c);
                            if (l2) return fourway(d, c, b, a);
                            if (r2) return fourway(b, c, d, a);
                            return threeway2(false, true, false, cm(b, d) ? b : d, a, c);
                        }
                        if (su) {  // up second order (value c), down first order (a)
                            bool l2 = look_l(i, ij, b), r2 = look_r(i, ij, d);
                            if (l2 && r2) return threeway2(true, true, false, cm(b, d) ? b : d, c, a);
                            if (l2) return fourway(d, a, b, c);
                            if (r2) return fourway(b, a, d, c);
                            return threeway2(false, true, false, cm(b, d) ? b : d, c, a);
                        }
                        d = cm(a, c) ? a : c;
                        bool l2 = look_l(i, ij, a), r2 = look_r(i, ij, c);
                        return threeway1(l2, r2, false, a, c, d);
                    }
                    bool r2 = look_r(i, ij, b), u2 = look_u(ij, a), d2 = look_d(ij, c);
                    return threeway2(r2, u2, d2, b, a, c);
                }
                if (hl) {
                    bool l2 = look_l(i, ij, b), u2 = look_u(ij, a), d2 = look_d(ij, c);
                    return threeway2(l2, u2, d2, b, a, c);
                }
                c = look_d(ij, a) ? ct(a, g.dy) * ti : ct(a, g.dy);
                d = look_u(ij, b) ? ct(b, g.dy) * ti : ct(b, g.dy);
                return cm(c, d) ? c : d;
            }
            if (hr) {
                if (hl) {
                    bool l2 = look_l(i, ij, a), r2 = look_r(i, ij, c), d2 = look_d(ij, b);
                    return threeway1(l2, r2, d2, a, c, b);
                }
                bool r2 = look_r(i, ij, a), d2 = look_d(ij, b);
                return corner(r2, d2, a, b);
            }
            if (hl) {
                bool l2 = look_l(i, ij, a), d2 = look_d(ij, b);
                return corner(l2, d2, a, b);
            }
            return look_d(ij, a) ? ct(a, g.dy) * ti : ct(a, g.dy);
        }
        if (hu) {
            if (hr) {
                if (hl) {
                    bool l2 = look_l(i, ij, a), r2 = look_r(i, ij, c), u2 = look_u(ij, b);
                    return threeway1(l2, r2, u2, a, c, b);
                }
                bool r2 = look_r(i, ij, a), u2 = look_u(ij, b);
                return corner(r2, u2, a, b);
            }
            if (hl) {
                bool l2 = look_l(i, ij, a), u2 = look_u(ij, b);
                return corner(l2, u2, a, b);
            }
            return look_u(ij, a) ? ct(a, g.dy) * ti : ct(a, g.dy);
        }
        if (hr) {
            if (hl) {
                c = look_l(i, ij, a) ? ct(a, g.dx) * ti : ct(a, g.dx);
                d = look_r(i, ij, b) ? ct(b, g.dx) * ti : ct(b, g.dx);
                return cm(c, d) ? c : d;
            }
            return look_r(i, ij, a) ? ct(a, g.dx) * ti : ct(a, g.dx);
        }
        if (hl) return look_l(i, ij, a) ? ct(a, g.dx) * ti : ct(a, g.dx);
        return 0;
    }

    // ---- heap primitives (heap.cc:41-160)
    void check_point(int i, int ij) {  // stack_test target
        if (s[ij] == st_far) {
            s[ij] = st_trial;
            phi[ij] = calc(i, ij);
            hp.push_back(ij);
            w++;
        }
    }
    void seed(int ij) {
        const int i = ij % g.m;
        if (i > 0) check_point(i - 1, ij - 1);
        if (i < g.m - 1) check_point(i + 1, ij + 1);
        if (ij >= g.m) check_point(i, ij - g.m);
        if (ij < g.mn - g.m) check_point(i, ij + g.m);
    }
    // Floyd heap construction with the reference's comparison sequence.
    void heapify() {
        const int last = w - 1;  // number of entries
        for (int bc = last >> 1; bc >= 1; bc--) {
            int tc = bc, c = tc << 1;
            const bool odd_tail = (w & 1) != 0;  // one node with a single child
            const int lim = odd_tail ? w - 1 : w;
            while (c < lim) {
                int ij = hp[tc];
                double tphi = phi[ij], tphi2 = phi[hp[c]];
                int ij3 = hp[c + 1];
                if (cm(tphi, tphi2)) {
                    if (cm(tphi, phi[ij3])) break;
                    c++; hp[tc] = ij3;
                } else if (cm(phi[ij3], tphi2)) {
                    c++; hp[tc] = ij3;
                } else {
                    hp[tc] = hp[c];
                }
                hp[c] = ij; tc = c; c = tc << 1;
            }
            if (odd_tail && c == w - 1) {
                int ij = hp[tc], ij2 = hp[c];
                if (cm(phi[ij2], phi[ij])) { hp[tc] = ij2; hp[c] = ij; }
            }
        }
        for (int c = 1; c < w; c++) bp[hp[c]] = c;
    }
    // Move a trial value up (heap.cc:120-135; values only move up).
    void raise(int c, int ij, double tphi) {
        int bc = c >> 1;
        if (bc >= 1 && cm(tphi, phi[hp[bc]])) {
            do {
                int ijt = hp[bc];
                bp[ijt] = c; hp[c] = ijt;
                c = bc; bc = c >> 1;
            } while (bc >= 1 && cm(tphi, phi[hp[bc]]));
            bp[ij] = c; hp[c] = ij;
        }
        phi[ij] = tphi;
    }
    void insert(int ij, double tphi) {  // heap.cc:105-118
        if ((int)hp.size() <= w) hp.resize(w + 1);
        int c = w, bc = c >> 1;
        while (bc >= 1 && cm(tphi, phi[hp[bc]])) {
            int ijt = hp[bc];
            bp[ijt] = c; hp[c] = ijt;
            c = bc; bc = c >> 1;
        }
        bp[ij] = c; phi[ij] = tphi; hp[c] = ij;
        w++;
    }
    void update(int i, int ij) {
        if (s[ij] == st_far) { s[ij] = st_trial; insert(ij, calc(i, ij)); }
        else if (s[ij] == st_trial) raise(bp[ij], ij, calc(i, ij));
    }
    void rescan(int i, int ij) {
        if (s[ij] == st_trial) raise(bp[ij], ij, calc(i, ij));
    }
    // Accept heap minima until the band limit (heap.cc:489-523).
    void march(double limit) {
        const int m = g.m, mn = g.mn;
        while (w > 1 && cm(phi[hp[1]], limit)) {
            const int ij = hp[1], i = ij % m;
            acc.push_back(ij);
            s[ij] = st_acc;
            w--;
            const int ijt = hp[w];
            const double tphi = phi[ijt];
            int bc = 1, c = 2;
            while (c + 1 < w) {
                int ij2 = hp[c], ij3 = hp[c + 1];
                double tphi2 = phi[ij2];
                if (cm(tphi2, tphi)) {
                    if (cm(phi[ij3], tphi2)) { hp[bc] = ij3; bp[ij3] = bc; bc = c + 1; }
                    else { hp[bc] = ij2; bp[ij2] = bc; bc = c; }
                } else if (cm(phi[ij3], tphi)) {
                    hp[bc] = ij3; bp[ij3] = bc; bc = c + 1;
                } else break;
                c = bc << 1;
            }
            if (c + 1 == w) {
                int ij2 = hp[c];
                if (cm(phi[ij2], tphi)) { hp[bc] = ij2; bp[ij2] = bc; bc = c; }
            }
            hp[bc] = ijt; bp[ijt] = bc;
            if (ij >= 2 * m) { rescan(i, ij - 2 * m); update(i, ij - m); }
            else if (ij >= m) update(i, ij - m);
            if (i > 1) { rescan(i - 2, ij - 2); update(i - 1, ij - 1); }
            else if (i > 0) update(i - 1, ij - 1);
            if (i < m - 2) { update(i + 1, ij + 1); rescan(i + 2, ij + 2); }
            else if (i < m - 1) update(i + 1, ij + 1);
            if (ij < mn - 2 * m) { update(i, ij + m); rescan(i, ij + 2 * m); }
            else if (ij < mn - m) update(i, ij + m);
        }
    }
};

// Locate the merge start: first slot in [lo,uv) whose phi is not below
// tphi, by galloping down from uv then bisecting (levelset.cc:394-400, 557-564).
int bisect(const double *phi, const std::vector<int> &u, int lu, int uu, double t) {
    while (uu - lu > 1) {
        int i = (lu + uu) >> 1;
        if (phi[u[i]] < t) lu = i; else uu = i;
    }
    return uu;
}
int search_down(const double *phi, const std::vector<int> &u, int uv, int lo, double t) {
    int k = 0, l = 1;
    while (uv - l > lo) {
        if (phi[u[uv - l]] < t) return bisect(phi, u, uv - l, uv - k, t);
        k = l; l <<= 1;
    }
    return phi[u[lo]] < t ? bisect(phi, u, lo, uv - k, t) : lo;
}
// Merge sorted runs u[lo,uv) and u[uv,hi) in place, ties to the second run
// (levelset.cc:337-365).
void merge_runs(const double *phi, std::vector<int> &u, int uv, int lo, int hi) {
    if (uv == hi) return;
    int k = search_down(phi, u, uv, lo, phi[u[uv]]), us = k, l = uv;
    std::vector<int> t;
    while (k < uv) {
        if (phi[u[k]] < phi[u[l]]) t.push_back(u[k++]);
        else {
            t.push_back(u[l++]);
            if (l == hi) {
                while (k < uv) t.push_back(u[k++]);
                break;
            }
        }
    }
    for (size_t q = 0; q < t.size(); q++) u[us + q] = t[q];
}

}  // namespace

void build_band(const ls_grid &g, double *phi, int *s, std::vector<int> &band) {
    const int m = g.m, n = g.n, mn = g.mn;
    // pos: outside interface points in insertion order; neg: inside ones in
    // insertion order (the reference prepends them, so its memory order is
    // the reverse).
    std::vector<int> pos, neg;
    auto addp = [&](int ij) { pos.push_back(ij); };
    auto addn = [&](int ij) { neg.push_back(ij); };

    // Sign-change scan (levelset.cc:211-263).
    int ij;
    s[0] = phi[0] > 0 ? 7 : 0;
    for (ij = 1; ij < m; ij++) {
        if (phi[ij] > 0) {
            if (s[ij - 1] < 4) {
                if (s[ij - 1] == 0) { s[ij - 1] = 3; addn(ij - 1); }
                s[ij] = 4; addp(ij);
            } else s[ij] = 7;
        } else {
            if (s[ij - 1] < 4) s[ij] = 0;
            else {
                if (s[ij - 1] == 7) { s[ij - 1] = 4; addp(ij - 1); }
                s[ij] = 3; addn(ij);
            }
        }
    }
    for (int j = 1; j < n; j++) {
        // first column: compare with the point below only
        if (phi[ij] > 0) {
            if (s[ij - m] < 4) {
                if (s[ij - m] == 0) { s[ij - m] = 3; addn(ij - m); }
                s[ij] = 4; addp(ij);
            } else s[ij] = 7;
        } else {
            if (s[ij - m] < 4) s[ij] = 0;
            else {
                if (s[ij - m] == 7) { s[ij - m] = 4; addp(ij - m); }
                s[ij] = 3; addn(ij);
            }
        }
        ij++;
        for (int i = 1; i < m; i++, ij++) {
            if (phi[ij] > 0) {
                if (s[ij - 1] < 4) {
                    if (s[ij - 1] == 0) { s[ij - 1] = 3; addn(ij - 1); }
                    if (s[ij - m] == 0) { s[ij - m] = 3; addn(ij - m); }
                    s[ij] = 4; addp(ij);
                } else if (s[ij - m] < 4) {
                    if (s[ij - m] == 0) { s[ij - m] = 3; addn(ij - m); }
                    s[ij] = 4; addp(ij);
                } else s[ij] = 7;
            } else {
                if (s[ij - 1] >= 4) {
                    if (s[ij - 1] == 7) { s[ij - 1] = 4; addp(ij - 1); }
                    if (s[ij - m] == 7) { s[ij - m] = 4; addp(ij - m); }
                    s[ij] = 3; addn(ij);
                } else if (s[ij - m] >= 4) {
                    if (s[ij - m] == 7) { s[ij - m] = 4; addp(ij - m); }
                    s[ij] = 3; addn(ij);
                } else s[ij] = 0;
            }
        }
    }

    // Exact distances on both interface layers from the original phi.
    std::vector<double> dpos(pos.size()), dneg(neg.size());
    for (size_t k = 0; k < pos.size(); k++) {
        double x0, y0, x, y;
        int q = pos[k];
        if (!newton(g, phi, s, q % m, q, x0, y0, x, y)) std::abort();
        x -= x0; y -= y0;
        dpos[k] = std::sqrt(x * x + y * y);
    }
    for (size_t k = 0; k < neg.size(); k++) {
        double x0, y0, x, y;
        int q = neg[k];
        if (!newton(g, phi, s, q % m, q, x0, y0, x, y)) std::abort();
        x -= x0; y -= y0;
        dneg[k] = -std::sqrt(x * x + y * y);
    }
    for (size_t k = 0; k < pos.size(); k++) phi[pos[k]] = dpos[k];
    for (size_t k = 0; k < neg.size(); k++) phi[neg[k]] = dneg[k];

    // Sort each layer: positive in insertion order, negative in the
    // reference's memory order (reverse insertion).
    std::vector<int> up(pos), dn(neg.rbegin(), neg.rend());
    if (!up.empty()) qsort_band(phi, up.data(), (int)up.size());
    if (!dn.empty()) qsort_band(phi, dn.data(), (int)dn.size());

    // Outward march; accepted points follow the interface layer.
    std::vector<int> acc_up, acc_dn;
    {
        marcher<true> mk(g, phi, s, acc_up);
        for (int q : up) mk.seed(q);
        mk.heapify();
        mk.march(g.phi_up);
    }
    {
        marcher<false> mk(g, phi, s, acc_dn);
        for (int q : dn) mk.seed(q);
        mk.heapify();
        mk.march(g.phi_down);
    }
    // Negative run in memory order: marched (reverse acceptance), then layer.
    std::vector<int> nu(acc_dn.rbegin(), acc_dn.rend());
    const int nmarch = (int)nu.size();
    nu.insert(nu.end(), dn.begin(), dn.end());
    merge_runs(phi, nu, nmarch, 0, (int)nu.size());
    std::vector<int> pu(up);
    const int nlayer = (int)pu.size();
    pu.insert(pu.end(), acc_up.begin(), acc_up.end());
    merge_runs(phi, pu, nlayer, 0, (int)pu.size());

    // Clamp outside the band (levelset.cc:408-413).
    for (int q = 0; q < mn; q++) {
        if (s[q] < 2) phi[q] = g.phi_down;
        if (s[q] > 5) phi[q] = g.phi_up;
    }
    band.clear();
    band.insert(band.end(), nu.begin(), nu.end());
    band.insert(band.end(), pu.begin(), pu.end());
}

bool extrapolate2(const ls_grid &g, const double *phi, const int *s,
                  const std::vector<int> &band, double *fx, double *fy) {
    const int m = g.m, n = g.n;
    const double s_point = -1.2 * std::sqrt(g.dx * g.dx + g.dy * g.dy);
    std::vector<char> known(g.mn, 0);
    size_t l = 0;
    if (band.empty()) return true;
    while (phi[band[l]] < s_point) {
        known[band[l]] = 1;
        if (++l == band.size()) return true;
    }
    for (; l < band.size(); l++) {
        const int ij = band[l], i = ij % m, j = ij / m;
        if (s[ij] < 4) { known[ij] = 1; continue; }
        double S, sx, sy, sxx, sxy, syy, det, p[6];
        int r = 3;
        do {
            const int ilo = i <= r ? -i : -r, ihi = i + r >= m ? m - 1 - i : r,
                      jlo = j <= r ? -j : -r, jhi = j + r >= n ? n - 1 - j : r;
            S = sx = sy = sxx = sxy = syy = 0;
            for (int k = 0; k < 6; k++) p[k] = 0;
            for (int dj = jlo; dj <= jhi; dj++)
                for (int di = ilo; di <= ihi; di++) {
                    const int q = ij + dj * m + di;
                    if (!known[q]) continue;
                    double w = phi[q];
                    if (!(w < phi[ij])) continue;
                    w = phi[ij] - w;
                    S += w; sx += di * w; sy += dj * w;
                    sxx += di * di * w; sxy += di * dj * w; syy += dj * dj * w;
                    const double a = fx[q], b = fy[q], x = di, y = dj;
                    p[0] += w * a; p[1] += w * x * a; p[2] += w * y * a;
                    p[3] += w * b; p[4] += w * x * b; p[5] += w * y * b;
                }
            det = S * (sxy * sxy - sxx * syy) + sx * sx * syy - 2 * sx * sxy * sy + sy * sy * sxx;
            r++;
        } while (std::fabs(det) < 1e-10 && r <= 7);
        if (std::fabs(det) < 1e-10) return false;
        det = 1 / det;
        const double al = (sxy * sxy - sxx * syy) * det, be = (sx * syy - sxy * sy) * det,
                     ga = (sxx * sy - sxy * sx) * det;
        fx[ij] = p[0] * al + p[1] * be + p[2] * ga;
        fy[ij] = p[3] * al + p[4] * be + p[5] * ga;
        known[ij] = 1;
    }
    return true;
}

}  // namespace orc
