// Test infrastructure only: a C-ABI wrapper around the reference's level-set
// library (levelset/levelset.cc, heap.cc, sorting.cc, traverse.hh), compiled
// from the sources where they lie by oracle/Makefile into oracle/_ref/.
// The oracle's own restatement (oracle/levelset_oracle.cc) is checked against
// these entry points; nothing in the product links or loads this file.
#include <cstring>
#include <cmath>
#include "level++.hh"

namespace {

// Two-field extension target with the same acc/set contract as the
// reference's obj_field (obj_field.hh:82-90): it extends (nX,nY).
struct two_field {
    static const int c = 2;
    double *fx, *fy;
    inline void acc(int ij, double *p, double w, double x, double y) {
        double a = fx[ij], b = fy[ij];
        *p += w * a; p[1] += w * x * a; p[2] += w * y * a;
        p[3] += w * b; p[4] += w * x * b; p[5] += w * y * b;
    }
    inline void set(int ij, double *p, double al, double be, double ga) {
        fx[ij] = *p * al + p[1] * be + p[2] * ga;
        fy[ij] = p[3] * al + p[4] * be + p[5] * ga;
    }
};

void copy_out(levelset &ls, double *phi, int *s, int *order, int *norder) {
    std::memcpy(phi, ls.phi, sizeof(double) * ls.mn);
    if (s) std::memcpy(s, ls.s, sizeof(int) * ls.mn);
    if (order) {
        int k = 0, l = ls.u_start;
        if (ls.u_start != ls.u_end || ls.u_end != 0) {
            // The band is a circular list from u_start to u_end.
            do {
                order[k++] = ls.u[l];
                l = (l == ls.u_mem - 1) ? 0 : l + 1;
            } while (l != ls.u_end);
        }
        *norder = k;
    }
}

}  // namespace

template class traverse<tra_positive, two_field>;

extern "C" {

// levelset::build_band (levelset.cc:199-335) on an m x n grid.
// phi: in (pinned values) / out (narrow band). s: out status codes.
// order: out band list in traversal order (ascending phi); norder: its length.
int ref_build_band(int m, int n, double ax, double ay, double dx, double dy,
                   double phi_down, double phi_up, double *phi, int *s,
                   int *order, int *norder) {
    levelset ls(m, n, ax, ay, dx, dy, phi_down, phi_up);
    std::memcpy(ls.phi, phi, sizeof(double) * m * n);
    ls.build_band();
    copy_out(ls, phi, s, order, norder);
    return 0;
}

// build_band followed by levelset::extrapolate_fields (traverse.hh:235-260)
// of two fields, as obj_field::extrapolate does for (nX,nY).
int ref_build_band_extrapolate(int m, int n, double ax, double ay, double dx,
                               double dy, double phi_down, double phi_up,
                               double *phi, int *s, double *fx, double *fy) {
    levelset ls(m, n, ax, ay, dx, dy, phi_down, phi_up);
    std::memcpy(ls.phi, phi, sizeof(double) * m * n);
    ls.build_band();
    two_field e;
    e.fx = fx; e.fy = fy;
    ls.extrapolate_fields(e);
    copy_out(ls, phi, s, 0, 0);
    return 0;
}

// levelset::newton (levelset.cc:107-127) for one grid point, after setting the
// status array by build_band's sign scan (so the failsafe has what it reads).
int ref_newton(int m, int n, double ax, double ay, double dx, double dy,
               const double *phi, const int *s, int i, int ij, double *out4) {
    levelset ls(m, n, ax, ay, dx, dy);
    std::memcpy(ls.phi, phi, sizeof(double) * m * n);
    std::memcpy(ls.s, s, sizeof(int) * m * n);
    double x0, y0, x, y;
    bool ok = ls.newton(i, ij, x0, y0, x, y);
    out4[0] = x0; out4[1] = y0; out4[2] = x; out4[3] = y;
    return ok ? 1 : 0;
}

// levelset::bicubic with gradient (levelset.cc:63-79, built/bicubic.cc).
double ref_bicubic(int m, int n, double ax, double ay, double dx, double dy,
                   const double *phi, double x, double y, double *grad2) {
    levelset ls(m, n, ax, ay, dx, dy);
    std::memcpy(ls.phi, phi, sizeof(double) * m * n);
    double px, py;
    double v = ls.bicubic(x, y, px, py);
    grad2[0] = px; grad2[1] = py;
    return v;
}

}
