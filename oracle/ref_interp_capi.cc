// Test infrastructure only: a C-ABI wrapper around the reference's bicubic
// interpolation (bi_interp.cc, fields.hh, vec.hh), compiled from the sources
// where they lie by oracle/Makefile into oracle/_ref/. The oracle's own
// restatement of the tracer interpolation is checked against it; nothing in
// the product links or loads this file.
#include <cstring>
#include <vector>
#include "bi_interp.hh"

extern "C" {

// Interpolates (fx, fy), given on an m x n grid spanning [ax,bx] x [ay,by],
// at (x,y): out6 = f (2), d/dx f (2), d/dy f (2). use_solid selects the
// reference-map path of the reference (s_field X,Y, f_grad_f), otherwise the
// velocity path (field u,v, f); both are evaluated and returned.
void ref_interp2(int m, int n, double ax, double bx, double ay, double by, const double *fx, const double *fy,
                 double x, double y, int use_solid, double *out6) {
    const int mn = m * n;
    vec v, gx, gy;
    if (use_solid) {
        std::vector<s_field> sf(mn);
        std::memset(static_cast<void *>(sf.data()), 0, sizeof(s_field) * mn);
        for (int q = 0; q < mn; q++) { sf[q].X = fx[q]; sf[q].Y = fy[q]; }
        bicubic_interp b(m, n, ax, bx, ay, by, NULL, sf.data());
        v = b.f_grad_f(x, y, gx, gy);
    } else {
        std::vector<field> ff(mn);
        std::memset(static_cast<void *>(ff.data()), 0, sizeof(field) * mn);
        for (int q = 0; q < mn; q++) { ff[q].u = fx[q]; ff[q].v = fy[q]; }
        bicubic_interp b(m, n, ax, bx, ay, by, ff.data(), NULL);
        v = b.f(x, y);
        gx = vec(0.); gy = vec(0.);
    }
    out6[0] = v.x; out6[1] = v.y; out6[2] = gx.x; out6[3] = gx.y; out6[4] = gy.x; out6[5] = gy.y;
}

}
