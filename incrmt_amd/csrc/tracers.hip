// Tracer kernels (fluid_2d.cc:1376-1484, object.cc:202-219, 313-417,
// bi_interp.cc). A few hundred tracers per step: one thread per tracer for
// the interpolations, one thread for the parts that the reference performs
// in sequence (the rand() draws of tracer_remap, in tracer order).
#include "tracers.hh"

namespace incrmt {

namespace {

#define KCHECK() HIP_CHECK(hipGetLastError())

struct Pt2 {
    double x, y;
};
__device__ inline Pt2 operator+(Pt2 a, Pt2 b) { return {a.x + b.x, a.y + b.y}; }
__device__ inline Pt2 operator-(Pt2 a, Pt2 b) { return {a.x - b.x, a.y - b.y}; }
__device__ inline Pt2 operator*(double e, Pt2 a) { return {e * a.x, e * a.y}; }

// bicubic_interp (bi_interp.cc:4-194) of two fields on the ghosted grid:
// cubic convolution in the bulk, quadratic in the first and last interval
// of each direction, every expression in the reference's operator order.
struct Bic {
    const Grid &g;
    const BicGeom &b;
    const double *fx, *fy;
    Pt2 a[16];
    __device__ Pt2 val(int i, int j) const {
        const int q = g.ls(i, j);
        return {fx[q], fy[q]};
    }
    __device__ void column(int i, int j, Pt2 &c0, Pt2 &c1, Pt2 &c2, Pt2 &c3) const {  // bi_interp.cc:105-171
        c0 = val(i, j);
        if (i == 0) {
            c1 = (-1.5 * val(i, j) + 2.0 * val(i + 1, j)) - 0.5 * val(i + 2, j);
            c2 = (0.5 * val(i, j) - val(i + 1, j)) + 0.5 * val(i + 2, j);
            c3 = {0, 0};
        } else if (i == b.m - 2) {
            c1 = -0.5 * val(i - 1, j) + 0.5 * val(i + 1, j);
            c2 = (0.5 * val(i - 1, j) - val(i, j)) + 0.5 * val(i + 1, j);
            c3 = {0, 0};
        } else {
            c1 = -0.5 * val(i - 1, j) + 0.5 * val(i + 1, j);
            c2 = ((val(i - 1, j) - 2.5 * val(i, j)) + 2.0 * val(i + 1, j)) - 0.5 * val(i + 2, j);
            c3 = ((-0.5 * val(i - 1, j) + 1.5 * val(i, j)) - 1.5 * val(i + 1, j)) + 0.5 * val(i + 2, j);
        }
    }
    __device__ static void fill_ad(Pt2 *p, Pt2 c1, Pt2 c2, Pt2 c3) {
        p[0] = c1;
        p[1] = (-1.5 * c1 + 2.0 * c2) - 0.5 * c3;
        p[2] = (0.5 * c1 - c2) + 0.5 * c3;
        p[3] = {0, 0};
    }
    __device__ static void fill_au(Pt2 *p, Pt2 c0, Pt2 c1, Pt2 c2) {
        p[0] = c1;
        p[1] = -0.5 * c0 + 0.5 * c2;
        p[2] = (0.5 * c0 - c1) + 0.5 * c2;
        p[3] = {0, 0};
    }
    __device__ static void fill_a(Pt2 *p, Pt2 c0, Pt2 c1, Pt2 c2, Pt2 c3) {
        p[0] = c1;
        p[1] = -0.5 * c0 + 0.5 * c2;
        p[2] = ((c0 - 2.5 * c1) + 2.0 * c2) - 0.5 * c3;
        p[3] = ((-0.5 * c0 + 1.5 * c1) - 1.5 * c2) + 0.5 * c3;
    }
    __device__ void index(double &x, double &y) {  // bi_interp.cc:30-44
        x = (x - b.ax) * b.xsp;
        y = (y - b.ay) * b.ysp;
        int i = static_cast<int>(x), j = static_cast<int>(y);
        if (i < 0) i = 0; else if (i > b.m - 2) i = b.m - 2;
        if (j < 0) j = 0; else if (j > b.n - 2) j = b.n - 2;
        Pt2 c[4][4];  // c[q][r]: coefficient q of row j-1+r
        column(i, j, c[0][1], c[1][1], c[2][1], c[3][1]);
        column(i, j + 1, c[0][2], c[1][2], c[2][2], c[3][2]);
        if (j == 0) {
            column(i, j + 2, c[0][3], c[1][3], c[2][3], c[3][3]);
            for (int q = 0; q < 4; q++) fill_ad(a + 4 * q, c[q][1], c[q][2], c[q][3]);
        } else if (j == b.n - 2) {
            column(i, j - 1, c[0][0], c[1][0], c[2][0], c[3][0]);
            for (int q = 0; q < 4; q++) fill_au(a + 4 * q, c[q][0], c[q][1], c[q][2]);
        } else {
            column(i, j - 1, c[0][0], c[1][0], c[2][0], c[3][0]);
            column(i, j + 2, c[0][3], c[1][3], c[2][3], c[3][3]);
            for (int q = 0; q < 4; q++) fill_a(a + 4 * q, c[q][0], c[q][1], c[q][2], c[q][3]);
        }
        x -= i;
        y -= j;
    }
    __device__ static Pt2 yl(const Pt2 *p, double y) { return p[0] + y * (p[1] + y * (p[2] + y * p[3])); }
    __device__ static Pt2 dyl(const Pt2 *p, double y) { return p[1] + y * (2.0 * p[2] + (3.0 * y) * p[3]); }
    __device__ Pt2 f(double x, double y) {  // bi_interp.cc:137-140
        index(x, y);
        return yl(a, y) + x * (yl(a + 4, y) + x * (yl(a + 8, y) + x * yl(a + 12, y)));
    }
    __device__ Pt2 f_grad_f(double x, double y, Pt2 &gx, Pt2 &gy) {  // bi_interp.cc:146-151
        index(x, y);
        gx = b.xsp * (yl(a + 4, y) + x * (2.0 * yl(a + 8, y) + (3 * x) * yl(a + 12, y)));
        gy = b.ysp * (dyl(a, y) + x * (dyl(a + 4, y) + x * (dyl(a + 8, y) + x * dyl(a + 12, y))));
        return yl(a, y) + x * (yl(a + 4, y) + x * (yl(a + 8, y) + x * yl(a + 12, y)));
    }
};

// tracer_update_1 / tracer_update_2 (fluid_2d.cc:1445-1459)
__device__ void update_1(const Grid &g, const BicGeom &b, const double *u, const double *v, double *tp) {
    Bic bi{g, b, u, v};
    const Pt2 w = bi.f(tp[0], tp[1]);
    tp[2] = w.x;
    tp[3] = w.y;
}
__device__ void update_2(const Grid &g, const BicGeom &b, const double *u, const double *v, double *tp, double dt) {
    const double tx = tp[0] + dt * tp[2], ty = tp[1] + dt * tp[3];
    Bic bi{g, b, u, v};
    const Pt2 w = bi.f(tx, ty);
    tp[0] += 0.5 * dt * (tp[2] + w.x);
    tp[1] += 0.5 * dt * (tp[3] + w.y);
}

// obj_track::correct_tracer (object.cc:383-417)
__device__ bool correct_tracer(const Grid &g, const BicGeom &b, const ObjArrays &oa, double xref, double yref,
                               double &xx, double &yy) {
    const double small_number = 100. * 2.220446049250313e-16;
    const double tolnewt = small_number * small_number * 100, bigstep = 1;
    Bic bi{g, b, oa.X, oa.Y};
    int count = 0, iter = 0;
    while (count < 4) {
        if (iter > 100) return false;
        Pt2 fx, fy, f = bi.f_grad_f(xx, yy, fx, fy);
        f = f - Pt2{xref, yref};
        if (f.x * f.x + f.y * f.y < tolnewt) count++;
        const double det = fx.x * fy.y - fy.x * fx.y;
        if (det < small_number && det > -small_number) return false;
        const double delx = (fy.y * f.x - fy.x * f.y) / det, dely = (-fx.y * f.x + fx.x * f.y) / det;
        if (delx * delx + dely * dely > bigstep) return false;
        xx -= delx;
        yy -= dely;
        iter++;
    }
    return true;
}

// obj_simple_spin::file_output (object.cc:202-219), the record kept on the device
__device__ void track_output(const Grid &g, const BicGeom &b, const ObjArrays *oa, TrackDev &t, double *rec,
                             int cap, int k, double time) {
    const double ang = atan2(t.tk[1], t.tk[0]);
    if (t.tcorr && !correct_tracer(g, b, oa[t.obj], t.tk[4], t.tk[5], t.tk[0], t.tk[1])) t.tcorr = 0;
    const int q = fabs(ang) > 0.5 * M_PI ? (ang > 0 ? 1 : -1) : 0;
    if (q == 1 && t.prev_q == -1) t.wind--;
    else if (q == -1 && t.prev_q == 1) t.wind++;
    t.prev_q = q;
    if (t.nrec < cap) {
        double *r = rec + ((size_t)k * cap + t.nrec) * 4;
        r[0] = time;
        r[1] = t.tk[0];
        r[2] = t.tk[1];
        r[3] = ang + 2 * M_PI * t.wind;
    }
    t.nrec++;
    t.count = 1;
}

__global__ void k_stage1(Grid g, BicGeom b, const double *u, const double *v, TracerSet t) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < t.ntrace) update_1(g, b, u, v, t.tm + 4 * q);
    else if (q < t.ntrace + t.ntrack) update_1(g, b, u, v, t.trk[q - t.ntrace].tk);
}

__global__ void k_stage2(Grid g, BicGeom b, const double *u, const double *v, TracerSet t, double dt) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= t.ntrace) return;
    double *tp = t.tm + 4 * q;
    update_2(g, b, u, v, tp, dt);
    if (tp[0] < b.bx_lo || tp[0] > b.bx_hi || tp[1] < b.by_lo || tp[1] > b.by_hi) atomicAdd(t.nout, 1);
}

// sim_type::tracer_remap (sim_type.cc:12-17) in tracer order, then the
// tracks' second stage and their every-10-steps records (object.cc:354-361)
__global__ void k_finish(Grid g, BicGeom b, const double *u, const double *v, const ObjArrays *oa, TracerSet t,
                         double dt, double time) {
    if (threadIdx.x == 0 && t.ntrace > 0 && *t.nout > 0) {
        GlibcRand r = *t.rng;
        for (int q = 0; q < t.ntrace; q++) {
            double &x = t.tm[4 * q], &y = t.tm[4 * q + 1];
            if (x < b.bx_lo) x += (b.bx_hi - b.bx_lo), y = r.rnd(b.by_lo, b.by_hi);
            if (x > b.bx_hi) x -= (b.bx_hi - b.bx_lo), y = r.rnd(b.by_lo, b.by_hi);
            if (y < b.by_lo) y += (b.by_hi - b.by_lo), x = r.rnd(b.bx_lo, b.bx_hi);
            if (y > b.by_hi) y -= (b.by_hi - b.by_lo), x = r.rnd(b.bx_lo, b.bx_hi);
        }
        *t.rng = r;
        *t.nout = 0;
    }
    for (int k = threadIdx.x; k < t.ntrack; k += blockDim.x) {
        TrackDev tr = t.trk[k];
        update_2(g, b, u, v, tr.tk, dt);
        if (tr.count == tr.every) track_output(g, b, oa, tr, t.rec, t.rec_cap, k, time);
        else tr.count++;
        t.trk[k] = tr;
    }
}

__global__ void k_frame_output(Grid g, BicGeom b, const ObjArrays *oa, TracerSet t, double time) {
    for (int k = threadIdx.x; k < t.ntrack; k += blockDim.x) {
        TrackDev tr = t.trk[k];
        track_output(g, b, oa, tr, t.rec, t.rec_cap, k, time);
        t.trk[k] = tr;
    }
}

}  // namespace

BicGeom bic_geom(const Grid &g) {
    BicGeom b;
    b.m = g.ml;
    b.n = g.nl;
    const double ddx = 1.5 * g.dx, ddy = 1.5 * g.dy;  // object.cc:322
    const double ax = g.ax - ddx, bx = g.bx + ddx, ay = g.ay - ddy, by = g.by + ddy;
    b.ax = ax;
    b.ay = ay;
    b.xsp = (b.m - 1) / (bx - ax);  // bi_interp.cc:8
    b.ysp = (b.n - 1) / (by - ay);
    b.bx_lo = g.ax; b.bx_hi = g.bx; b.by_lo = g.ay; b.by_hi = g.by;
    return b;
}

int tracers_stage1(hipStream_t s, const Grid &g, const double *u, const double *v, TracerSet &t) {
    const int n = t.ntrace + t.ntrack;
    if (n == 0) return 0;
    k_stage1<<<(n + 255) / 256, 256, 0, s>>>(g, bic_geom(g), u, v, t);
    KCHECK();
    return 0;
}

int tracers_stage2(hipStream_t s, const Grid &g, const double *u, const double *v, const ObjArrays *d_oa,
                   TracerSet &t, double dt, double time) {
    if (t.ntrace + t.ntrack == 0) return 0;
    const BicGeom b = bic_geom(g);
    if (t.ntrace > 0) {
        k_stage2<<<(t.ntrace + 255) / 256, 256, 0, s>>>(g, b, u, v, t, dt);
        KCHECK();
    }
    k_finish<<<1, 64, 0, s>>>(g, b, u, v, d_oa, t, dt, time);
    KCHECK();
    return 0;
}

int tracers_frame_output(hipStream_t s, const Grid &g, const ObjArrays *d_oa, TracerSet &t, double time) {
    if (t.ntrack == 0) return 0;
    k_frame_output<<<1, 64, 0, s>>>(g, bic_geom(g), d_oa, t, time);
    KCHECK();
    return 0;
}

}  // namespace incrmt
