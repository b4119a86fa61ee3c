// Local second-order upwind eikonal update of the reference's fast-marching
// method (levelset/heap.cc:177-388): given which 4-neighbours of a point are
// known and their values, return the point's distance estimate. The branch
// structure and every expression follow the reference, so the GPU band
// rebuild solves the same discrete equations the reference's heap march
// solves; what differs is only how the known set is found (a parallel
// fixed-point iteration instead of a serial heap).
//
// The accessor A supplies in(q) (is neighbour q known) and val(q).
// POS selects the outward march (values increasing) or the inward one.
#ifndef INCRMT_EIKONAL_HH
#define INCRMT_EIKONAL_HH

#ifndef __HIPCC__
#ifndef __host__
#define __host__
#define __device__
#endif
#endif
#include <cmath>

namespace incrmt {

struct EikGeom {
    double dx, dy, dx2, dy2, dxbydy, dx2bydy2;
    // corner() constants (heap.cc:349-356)
    double f2a, f2b, f2c, f2d, b1, b2, b3, i2a, i2b, i2c, i2d;
    __host__ __device__ void init(double dx_, double dy_) {
        dx = dx_; dy = dy_; dx2 = dx * dx; dy2 = dy * dy;
        dxbydy = dx / dy; dx2bydy2 = dxbydy * dxbydy;
        f2a = 2.25 * (dy2 + dx2); f2b = 2.25 * dy2 + dx2; f2c = dy2 + 2.25 * dx2; f2d = dy2 + dx2;
        b1 = 2.25 * dx2 > dy2 ? 1 + 2.25 * dx2bydy2 : 1 + dy2 / (2.25 * dx2);
        b2 = dx2 > dy2 ? 1 + dx2bydy2 : 1 + dy2 / dx2;
        b3 = dx2 > 2.25 * dy2 ? 1 + dx2bydy2 / 2.25 : 1 + 2.25 * dy2 / dx2;
        i2a = dy2 / f2a; i2b = dy2 / f2b; i2c = dy2 / f2c; i2d = dy2 / f2d;
    }
};

template <bool POS, class A>
struct Eik {
    const A &a;
    const EikGeom &g;
    // m, n: grid extents; P: row stride
    int m, n, P;
    __host__ __device__ static bool cm(double x, double y) { return POS ? x < y : x > y; }
    __host__ __device__ static double ct(double x, double y) { return POS ? x + y : x - y; }
    __host__ __device__ void clamp0(double &v) const { if (cm(v, 0)) v = 0; }
    __host__ __device__ bool small(double gg, double hh) const { return gg * gg + hh * hh * g.dx2bydy2 < g.dx2; }

    // upwind neighbour value, second order when the next point out is known
    // and further upwind (heap.cc:442-487)
    __host__ __device__ bool look(int ij, int d1, bool d2ok, double &e) const {
        if (d2ok && a.in(ij + 2 * d1) && cm(a.val(ij + 2 * d1), a.val(ij + d1))) {
            e = 2 * a.val(ij + d1) - 0.5 * a.val(ij + 2 * d1);
            return true;
        }
        e = a.val(ij + d1);
        return false;
    }
    __host__ __device__ double corner_bail(bool h, bool v, double x, double y) const {
        const double ti = 2.0 / 3.0;
        double ht = h ? ct(x, g.dx) * ti : ct(x, g.dx);
        double vt = v ? ct(y, g.dy) * ti : ct(y, g.dy);
        return cm(ht, vt) ? ht : vt;
    }
    __host__ __device__ double corner(bool h, bool v, double x, double y) const {
        double f;
        if (h && v) {
            f = 1.5 * (x - y); f *= f;
            return f * g.b2 < g.f2a ? ct(x * 1.5 + y * 1.5 * g.dx2bydy2, g.dxbydy * sqrt(g.f2a - f)) * g.i2a
                                    : corner_bail(true, true, x, y);
        } else if (h) {
            f = x - 1.5 * y; f *= f;
            return f * g.b3 < g.f2b ? ct(x * 1.5 + y * g.dx2bydy2, g.dxbydy * sqrt(g.f2b - f)) * g.i2b
                                    : corner_bail(true, false, x, y);
        } else if (v) {
            f = 1.5 * x - y; f *= f;
            return f * g.b1 < g.f2c ? ct(x + y * 1.5 * g.dx2bydy2, g.dxbydy * sqrt(g.f2c - f)) * g.i2c
                                    : corner_bail(false, true, x, y);
        }
        f = x - y; f *= f;
        return f * g.b2 < g.f2d ? ct(x + y * g.dx2bydy2, g.dxbydy * sqrt(g.f2d - f)) * g.i2d
                                : corner_bail(false, false, x, y);
    }
    __host__ __device__ double fourway(double x1, double y1, double x2, double y2) const {
        // x1,y1 first-order (h,v); x2,y2 second-order (h,v) (heap.cc:264-292)
        double e = 2 * (x2 - x1), f = 2 * (y2 - y1), gg, hh;
        if (cm(e, f)) {
            gg = e - x1; hh = e - y1; clamp0(gg); clamp0(hh);
            if (small(gg, hh)) {
                gg = 1.5 * f - x2; hh = f - y1; clamp0(gg); clamp0(hh);
                return small(gg, hh) ? corner(true, true, x2, y2) : corner(true, false, x2, y1);
            }
            return corner(false, false, x1, y1);
        }
        gg = f - x1; hh = f - y1; clamp0(gg); clamp0(hh);
        if (small(gg, hh)) {
            gg = e - x1; hh = 1.5 * e - y2; clamp0(gg); clamp0(hh);
            return small(gg, hh) ? corner(true, true, x2, y2) : corner(false, true, x1, y2);
        }
        return corner(false, false, x1, y1);
    }
    __host__ __device__ double threeway1(bool h1, bool h2, bool v, double l, double r, double b) const {
        if (h1 && h2) return corner(true, v, cm(l, r) ? l : r, b);
        if (!h1 && !h2) return corner(false, v, cm(l, r) ? l : r, b);
        double d, e, f;
        if (h1) {
            d = 2 * (l - r); e = d - r; f = (v ? 1.5 * d : d) - b;
            clamp0(e); clamp0(f);
            return e * e + f * f * g.dx2bydy2 < g.dx2 ? corner(true, v, l, b) : corner(false, v, r, b);
        }
        d = 2 * (r - l); e = d - l; f = (v ? 1.5 * d : d) - b;
        clamp0(e); clamp0(f);
        return e * e + f * f * g.dx2bydy2 < g.dx2 ? corner(true, v, r, b) : corner(false, v, l, b);
    }
    __host__ __device__ double threeway2(bool h, bool v1, bool v2, double b, double x, double y) const {
        if (v1 && v2) return corner(h, true, b, cm(x, y) ? x : y);
        if (!v1 && !v2) return corner(h, false, b, cm(x, y) ? x : y);
        double d, e, f;
        if (v1) {
            d = 2 * (x - y); e = d - y; f = (h ? 1.5 * d : d) - b;
            clamp0(e); clamp0(f);
            return f * f + e * e * g.dx2bydy2 < g.dx2 ? corner(h, true, b, x) : corner(h, false, b, y);
        }
        d = 2 * (y - x); e = d - x; f = (h ? 1.5 * d : d) - b;
        clamp0(e); clamp0(f);
        return f * f + e * e * g.dx2bydy2 < g.dx2 ? corner(h, true, b, y) : corner(h, false, b, x);
    }
    // The local update at grid point (I,J), linear index ij (heap.cc:177-262).
    __host__ __device__ double calc(int I, int J, int ij) const {
        const double ti = 2.0 / 3.0;
        double va, vb, vc, vd;
        const bool hd = J > 0 && a.in(ij - P), hu = J < n - 1 && a.in(ij + P);
        const bool hl = I > 0 && a.in(ij - 1), hr = I < m - 1 && a.in(ij + 1);
        const bool okl = I > 1, okr = I < m - 2, okd = J > 1, oku = J < n - 2;
        auto L = [&](double &e) { return look(ij, -1, okl, e); };
        auto R = [&](double &e) { return look(ij, 1, okr, e); };
        auto D = [&](double &e) { return look(ij, -P, okd, e); };
        auto U = [&](double &e) { return look(ij, P, oku, e); };
        if (hd) {
            if (hu) {
                if (hr) {
                    if (hl) {
                        bool sd = D(va), su = U(vc);
                        if (sd && su) {
                            vd = cm(va, vc) ? va : vc;
                            bool l2 = L(va), r2 = R(vc);
                            return threeway1(l2, r2, true, va, vc, vd);
                        }
                        if (sd) {
                            bool l2 = L(vb), r2 = R(vd);
                            if (l2 && r2) return threeway2(true, true, false, cm(vb, vd) ? vb : vd, va, vc);
                            if (l2) return fourway(vd, vc, vb, va);
                            if (r2) return fourway(vb, vc, vd, va);
                            return threeway2(false, true, false, cm(vb, vd) ? vb : vd, va, vc);
                        }
                        if (su) {
                            bool l2 = L(vb), r2 = R(vd);
                            if (l2 && r2) return threeway2(true, true, false, cm(vb, vd) ? vb : vd, vc, va);
                            if (l2) return fourway(vd, va, vb, vc);
                            if (r2) return fourway(vb, va, vd, vc);
                            return threeway2(false, true, false, cm(vb, vd) ? vb : vd, vc, va);
                        }
                        vd = cm(va, vc) ? va : vc;
                        bool l2 = L(va), r2 = R(vc);
                        return threeway1(l2, r2, false, va, vc, vd);
                    }
                    bool r2 = R(vb), u2 = U(va), d2 = D(vc);
                    return threeway2(r2, u2, d2, vb, va, vc);
                }
                if (hl) {
                    bool l2 = L(vb), u2 = U(va), d2 = D(vc);
                    return threeway2(l2, u2, d2, vb, va, vc);
                }
                vc = D(va) ? ct(va, g.dy) * ti : ct(va, g.dy);
                vd = U(vb) ? ct(vb, g.dy) * ti : ct(vb, g.dy);
                return cm(vc, vd) ? vc : vd;
            }
            if (hr) {
                if (hl) {
                    bool l2 = L(va), r2 = R(vc), d2 = D(vb);
                    return threeway1(l2, r2, d2, va, vc, vb);
                }
                bool r2 = R(va), d2 = D(vb);
                return corner(r2, d2, va, vb);
            }
            if (hl) {
                bool l2 = L(va), d2 = D(vb);
                return corner(l2, d2, va, vb);
            }
            return D(va) ? ct(va, g.dy) * ti : ct(va, g.dy);
        }
        if (hu) {
            if (hr) {
                if (hl) {
                    bool l2 = L(va), r2 = R(vc), u2 = U(vb);
                    return threeway1(l2, r2, u2, va, vc, vb);
                }
                bool r2 = R(va), u2 = U(vb);
                return corner(r2, u2, va, vb);
            }
            if (hl) {
                bool l2 = L(va), u2 = U(vb);
                return corner(l2, u2, va, vb);
            }
            return U(va) ? ct(va, g.dy) * ti : ct(va, g.dy);
        }
        if (hr) {
            if (hl) {
                vc = L(va) ? ct(va, g.dx) * ti : ct(va, g.dx);
                vd = R(vb) ? ct(vb, g.dx) * ti : ct(vb, g.dx);
                return cm(vc, vd) ? vc : vd;
            }
            return R(va) ? ct(va, g.dx) * ti : ct(va, g.dx);
        }
        if (hl) return L(va) ? ct(va, g.dx) * ti : ct(va, g.dx);
        return 0;
    }

    // ---- the same update with the memory reads and the square root taken
    // out of the decision tree (for SIMD execution): every branch of calc
    // ends in corner(h,v,x,y) or in a pair of one-sided values (corner_bail
    // with per-side spacings), so the tree only picks those arguments; the
    // one arithmetic tail below then evaluates them with selects. Each
    // expression keeps the operation order of corner() and corner_bail(),
    // multiplying by 1 where a case has no factor (exact in IEEE
    // arithmetic), so the result is bit-identical to calc().
    struct Pick {
        bool corner;       // corner(h,v,x,y), else the two one-sided values
        bool h, v;         // second-order flags
        double x, y;       // arguments
        double sx, sy;     // spacings of the one-sided values
    };
    __host__ __device__ static double far() { return POS ? INFINITY : -INFINITY; }
    __host__ __device__ void pick_corner(Pick &p, bool h, bool v, double x, double y) const {
        p.corner = true; p.h = h; p.v = v; p.x = x; p.y = y; p.sx = g.dx; p.sy = g.dy;
    }
    __host__ __device__ void pick_sides(Pick &p, bool h, double x, double sx, bool v, double y, double sy) const {
        p.corner = false; p.h = h; p.v = v; p.x = x; p.y = y; p.sx = sx; p.sy = sy;
    }
    __host__ __device__ double eval(const Pick &p) const {
        const double ti = 2.0 / 3.0;
        // one-sided values (corner_bail, heap.cc:371-388)
        const double hb = ct(p.x, p.sx), vb = ct(p.y, p.sy);
        const double ht = p.h ? hb * ti : hb, vt = p.v ? vb * ti : vb;
        const double bail = cm(ht, vt) ? ht : vt;
        if (!p.corner) return bail;
        // corner (heap.cc:349-370): constants of the four cases
        const bool h = p.h, v = p.v;
        const double fs = h && v ? 1.5 : 1.0, fx = !h && v ? 1.5 : 1.0, fy = h && !v ? 1.5 : 1.0;
        const double nx = h ? 1.5 : 1.0, ny = v ? 1.5 : 1.0;
        const double B = h ? (v ? g.b2 : g.b3) : (v ? g.b1 : g.b2);
        const double F = h ? (v ? g.f2a : g.f2b) : (v ? g.f2c : g.f2d);
        const double I = h ? (v ? g.i2a : g.i2b) : (v ? g.i2c : g.i2d);
        double f = fs * (p.x * fx - p.y * fy);
        f *= f;
        const bool ok = f * B < F;
        const double r = ct(p.x * nx + (p.y * ny) * g.dx2bydy2, g.dxbydy * sqrt(ok ? F - f : 0.0)) * I;
        return ok ? r : bail;
    }
    __host__ __device__ void fourway_p(Pick &p, double x1, double y1, double x2, double y2) const {
        double e = 2 * (x2 - x1), f = 2 * (y2 - y1), gg, hh;
        if (cm(e, f)) {
            gg = e - x1; hh = e - y1; clamp0(gg); clamp0(hh);
            if (small(gg, hh)) {
                gg = 1.5 * f - x2; hh = f - y1; clamp0(gg); clamp0(hh);
                if (small(gg, hh)) pick_corner(p, true, true, x2, y2);
                else pick_corner(p, true, false, x2, y1);
                return;
            }
            pick_corner(p, false, false, x1, y1);
            return;
        }
        gg = f - x1; hh = f - y1; clamp0(gg); clamp0(hh);
        if (small(gg, hh)) {
            gg = e - x1; hh = 1.5 * e - y2; clamp0(gg); clamp0(hh);
            if (small(gg, hh)) pick_corner(p, true, true, x2, y2);
            else pick_corner(p, false, true, x1, y2);
            return;
        }
        pick_corner(p, false, false, x1, y1);
    }
    __host__ __device__ void threeway1_p(Pick &p, bool h1, bool h2, bool v, double l, double r, double b) const {
        if (h1 && h2) { pick_corner(p, true, v, cm(l, r) ? l : r, b); return; }
        if (!h1 && !h2) { pick_corner(p, false, v, cm(l, r) ? l : r, b); return; }
        double d, e, f;
        if (h1) {
            d = 2 * (l - r); e = d - r; f = (v ? 1.5 * d : d) - b;
            clamp0(e); clamp0(f);
            if (e * e + f * f * g.dx2bydy2 < g.dx2) pick_corner(p, true, v, l, b);
            else pick_corner(p, false, v, r, b);
            return;
        }
        d = 2 * (r - l); e = d - l; f = (v ? 1.5 * d : d) - b;
        clamp0(e); clamp0(f);
        if (e * e + f * f * g.dx2bydy2 < g.dx2) pick_corner(p, true, v, r, b);
        else pick_corner(p, false, v, l, b);
    }
    __host__ __device__ void threeway2_p(Pick &p, bool h, bool v1, bool v2, double b, double x, double y) const {
        if (v1 && v2) { pick_corner(p, h, true, b, cm(x, y) ? x : y); return; }
        if (!v1 && !v2) { pick_corner(p, h, false, b, cm(x, y) ? x : y); return; }
        double d, e, f;
        if (v1) {
            d = 2 * (x - y); e = d - y; f = (h ? 1.5 * d : d) - b;
            clamp0(e); clamp0(f);
            if (f * f + e * e * g.dx2bydy2 < g.dx2) pick_corner(p, h, true, b, x);
            else pick_corner(p, h, false, b, y);
            return;
        }
        d = 2 * (y - x); e = d - x; f = (h ? 1.5 * d : d) - b;
        clamp0(e); clamp0(f);
        if (f * f + e * e * g.dx2bydy2 < g.dx2) pick_corner(p, h, true, b, y);
        else pick_corner(p, h, false, b, x);
    }
    // calc() with the neighbour reads done first: kX known, sX second order,
    // eX the value look() gives (heap.cc:177-262)
    __host__ __device__ double calc_fast(int I, int J, int ij) const {
        const bool hd = J > 0 && a.in(ij - P), hu = J < n - 1 && a.in(ij + P);
        const bool hl = I > 0 && a.in(ij - 1), hr = I < m - 1 && a.in(ij + 1);
        double eL = 0, eR = 0, eD = 0, eU = 0;
        const bool sL = hl && look(ij, -1, I > 1, eL), sR = hr && look(ij, 1, I < m - 2, eR);
        const bool sD = hd && look(ij, -P, J > 1, eD), sU = hu && look(ij, P, J < n - 2, eU);
        Pick p;
        pick_sides(p, false, far(), g.dx, false, far(), g.dy);
        if (hd) {
            if (hu) {
                if (hr) {
                    if (hl) {
                        if (sD && sU) threeway1_p(p, sL, sR, true, eL, eR, cm(eD, eU) ? eD : eU);
                        else if (sD) {
                            const double mlr = cm(eL, eR) ? eL : eR;
                            if (sL && sR) threeway2_p(p, true, true, false, mlr, eD, eU);
                            else if (sL) fourway_p(p, eR, eU, eL, eD);
                            else if (sR) fourway_p(p, eL, eU, eR, eD);
                            else threeway2_p(p, false, true, false, mlr, eD, eU);
                        } else if (sU) {
                            const double mlr = cm(eL, eR) ? eL : eR;
                            if (sL && sR) threeway2_p(p, true, true, false, mlr, eU, eD);
                            else if (sL) fourway_p(p, eR, eD, eL, eU);
                            else if (sR) fourway_p(p, eL, eD, eR, eU);
                            else threeway2_p(p, false, true, false, mlr, eU, eD);
                        } else threeway1_p(p, sL, sR, false, eL, eR, cm(eD, eU) ? eD : eU);
                    } else threeway2_p(p, sR, sU, sD, eR, eU, eD);
                } else if (hl) threeway2_p(p, sL, sU, sD, eL, eU, eD);
                else pick_sides(p, sD, eD, g.dy, sU, eU, g.dy);
            } else if (hr) {
                if (hl) threeway1_p(p, sL, sR, sD, eL, eR, eD);
                else pick_corner(p, sR, sD, eR, eD);
            } else if (hl) pick_corner(p, sL, sD, eL, eD);
            else pick_sides(p, sD, eD, g.dy, false, far(), g.dy);
        } else if (hu) {
            if (hr) {
                if (hl) threeway1_p(p, sL, sR, sU, eL, eR, eU);
                else pick_corner(p, sR, sU, eR, eU);
            } else if (hl) pick_corner(p, sL, sU, eL, eU);
            else pick_sides(p, sU, eU, g.dy, false, far(), g.dy);
        } else if (hr) {
            if (hl) pick_sides(p, sL, eL, g.dx, sR, eR, g.dx);
            else pick_sides(p, sR, eR, g.dx, false, far(), g.dx);
        } else if (hl) pick_sides(p, sL, eL, g.dx, false, far(), g.dx);
        else return 0;
        return eval(p);
    }
};

}  // namespace incrmt

#endif