// Shared definitions of the MI355X IncRMT hot path: device layout of the
// ghosted grid, per-solid descriptors, and the pointwise formulas of the
// reference step that several kernels share.
//
// Layout. Every field is its own FP64 array (structure of arrays) over the
// reference's ghosted (m+4) x (n+4) grid (fluid_2d.cc:37: fm = fbase+2ml+2).
// Rows are padded to a pitch P that is a multiple of 32 doubles (256 B) and
// the array origin is shifted so that interior column i = 0 of every row is
// 256 B aligned: each 64-lane wavefront that sweeps a row reads whole 512 B
// segments.
#ifndef INCRMT_COMMON_HH
#define INCRMT_COMMON_HH

#include <hip/hip_runtime.h>
#include <cstdint>
#include <cmath>

namespace incrmt {

constexpr int kMaxObj = 64;
constexpr int kOrigin = 32;  // doubles before the row of ghost column -2, +2

struct Grid {
    int m, n;    // cells
    int ml, nl;  // ghosted extents m+4, n+4 (the level-set grid)
    int P;       // row pitch (doubles)
    long long N; // allocation length (doubles)
    double ax, bx, ay, by, dx, dy, xsp, ysp, xxsp, yysp;
    double lax, lay;  // origin of the level-set grid: ax-1.5dx, ay-1.5dy (obj_field.cc:18)
    int px, py;       // periodic directions (sim_type::x_prd, y_prd)
    int vel;          // initial fluid velocity: one of the SimVelocity kinds
    // horizontal flow (sim_horiz_flow, sim_type.hh:116-145): zero pressure on
    // the right edge (fluid_2d::zpr) and inflow U through the left ghost
    // columns (sim_type.cc:51-55)
    int zpr;
    double U;
    // fluid_2d::solid_prd_bc: the reference maps are periodic up to the box
    // lengths (set_boundaries_solid, edge_boundary_conditions_solid_pre /
    // _post, fluid_2d.cc:784-835) and solids feel no wall force
    int spbc;
    // cell (i,j), i in [-2,m+1], j in [-2,n+1]
    __host__ __device__ inline int at(int i, int j) const { return kOrigin + (j + 2) * P + i; }
    // level-set point (I,J) = cell (I-2,J-2)
    __host__ __device__ inline int ls(int I, int J) const { return kOrigin - 2 + J * P + I; }
};

enum ObjKind : int { kSpin = 0, kCircle = 1, kSquare = 2, kFlapper = 3, kLamprey = 4, kFull = 5 };

// sim_type::velocity of the reference's simulation types: at rest (the
// default, sim_type.hh:67-70), four_rolls (sim_type.cc:30-35),
// velocity_pulses (sim_type.cc:43-47, sim_type.hh:109-113) and horiz_flow
// (u = U, sim_type.hh:129-135).
enum SimVelocity : int { kSimRest = 0, kSimFourRolls = 1, kSimPulses = 2, kSimHorizFlow = 3 };
__host__ __device__ inline void sim_velocity(int kind, double x, double y, double &u, double &v, double U) {
    u = v = 0;
    if (kind == kSimHorizFlow) {
        u = U;
    } else if (kind == kSimFourRolls) {
        u = sin(M_PI * x) * cos(M_PI * y);
        v = -cos(M_PI * x) * sin(M_PI * y);
    } else if (kind == kSimPulses) {
        for (double k = 2, s = 1, z = -5 / 6.; z < 1; z += 1 / 3., k += 2, s = -s) {
            const double xx = x - z, yy = y - z;
            const double f = s * exp(-k * (2 - cos(xx * M_PI) - cos(yy * M_PI)));
            u -= sin(yy * M_PI) * f;
            v += sin(xx * M_PI) * f;
        }
    }
}

// Constants of one solid (obj_field.hh:32-135, object.hh:128-138, 245-289,
// 422-467, 542-595).
struct ObjConst {
    int kind;
    int set_velocity;  // mat_const::set_velocity (obj_field.hh:25-28): the solid sets its initial velocity
    double eps, pinw, G, rhos, ev_trans_mult, ex_visc, tf1, tf2, tf3;
    double c0, c1, c2, piv_r, th_max, omega, K_stiff, piv_thresh;  // simple spin
    double cx, cy, cr, grav;                                       // circle, square; centre of all
    double cl, ar, al, amp;                                        // flapper (omega shared)
    double dis, la, w1, w2, T, A;                                  // lamprey
    double cth, sth;                                               // flapper / lamprey rotation
    double phi_down, phi_up;                                       // band cut-offs
};

// Device arrays of one solid (all on the ghosted grid).
struct ObjArrays {
    double *X, *Y, *nX, *nY;
    double *Xx, *Xy, *Yx, *Yy;
    double *Xl, *Yl, *Xr, *Yr, *Xd, *Yd, *Xu, *Yu;
    double *phi;
    int *st;     // narrow-band status 0..7 (levelset.cc:199-263, heap.cc:525-526)
};

// Pointwise device helpers shared by the kernels ------------------------

__device__ inline double trans_func_in(const ObjConst &o, double x) {
    return x < -o.eps ? 1 : 0.5 - x * o.tf1 - o.tf2 * sin(o.tf3 * x);  // obj_field.hh:107-112
}
__device__ inline double trans_func(const ObjConst &o, double x) {
    return x > o.eps ? 0 : trans_func_in(o, x);  // obj_field.hh:102-104
}
__device__ inline double tderiv_func_in(const ObjConst &o, double x) {
    return x < -o.eps ? 0 : 0.5 * (1 + cos(M_PI * x / o.eps));  // obj_field.hh:117-120
}
__device__ inline double coll_func(const ObjConst &o, double x) {
    return x > o.eps ? 0 : 0.5 - x * o.tf1;  // obj_field.hh:97-99
}
// object::actuate of the actuated solids: the deformation gradient
// (Fa Fb; Fc Fd) scaled by an active stretch st, (a, c) by st and (b, d) by
// 1/st. obj_flapper (object.cc:301-308): st = exp(con Y H(ph)) on the
// rounded rod ph of the actuated region; obj_lamprey (object.cc:573-580): a
// contraction wave of shift sh along the body. s0, s1: the solid's
// pre_stress_setup state (con, sh).
__device__ inline void actuate(const ObjConst &o, double s0, double s1, double X, double Y, double &Fa, double &Fb,
                               double &Fc, double &Fd) {
    if (o.kind == kFlapper) {
        const double fX = fabs(X) - o.al, ph = (fX < 0 ? fabs(Y) : sqrt(fX * fX + Y * Y)) - o.ar;
        if (ph < o.eps) {
            const double fac = exp(s0 * Y * trans_func_in(o, ph)), inv = 1. / fac;
            Fa *= fac; Fb *= inv;
            Fc *= fac; Fd *= inv;
        }
    } else if (o.kind == kLamprey) {
        const double dx = X + (Y > 0 ? 1.5 * o.la : o.la) + s1, wl = o.w1 - (o.w2 - o.w1) * X / o.la,
                     fac = fabs(Y) / wl;
        const double a = -X, b = o.la * 0.15 - fabs(dx - o.la * (int(dx / o.la) + 0.5));
        const double st = exp(s0 * (fac < 1 ? fac * (2 - fac) : 1) * trans_func(o, -(b < a ? b : a))), inv = 1. / st;
        Fa *= st; Fb *= inv;
        Fc *= st; Fd *= inv;
    }
}
// object::phi for the supported solids (object.hh:85, object.cc:22-24,
// 83-87, 167-171, 284-287, 554-558)
__host__ __device__ inline double obj_phi(const ObjConst &o, double X, double Y) {
    if (o.kind == kFull) return -1;  // obj_full fills the box (object.hh:83-86)
    if (o.kind == kSpin) {
        double th = atan2(Y, X), c = cos(3 * th);
        return sqrt(X * X + Y * Y) - o.c0 - c * (o.c1 + o.c2 * c);
    }
    if (o.kind == kFlapper) {  // a rounded rod
        const double fX = fabs(X) - o.cl;
        return (fX < 0 ? fabs(Y) : sqrt(fX * fX + Y * Y)) - o.cr;
    }
    if (o.kind == kLamprey) {
        if (X > 0) return o.dis * Y * Y / (o.w1 * o.w1) + X - o.dis;
        const double wl = o.w1 - (o.w2 - o.w1) * X / o.la, a = -o.la - X, b = Y * Y / (wl * wl) - 1;
        return a < b ? b : a;  // std::max
    }
    if (o.kind == kSquare) {
        X = fabs(X) - o.cr;
        Y = fabs(Y) - o.cr;
        return X < 0 || Y < 0 ? fmax(X, Y) : sqrt(X * X + Y * Y);
    }
    return sqrt(X * X + Y * Y) - o.cr;
}

}  // namespace incrmt

#define HIP_CHECK(x)                                                      \
    do {                                                                  \
        hipError_t e_ = (x);                                              \
        if (e_ != hipSuccess) return ::incrmt::hip_fail(e_, #x, __FILE__, __LINE__); \
    } while (0)

namespace incrmt {
int hip_fail(hipError_t e, const char *what, const char *file, int line);
// Record msg as the calling thread's last error and return code.
int set_error(int code, const char *msg);
// Clear device memory on the owner's stream s and wait for that stream
// only. Nothing here touches the null stream or the whole device: another
// host thread may be capturing a graph on its own stream meanwhile, and a
// device-wide synchronisation invalidates that capture.
inline hipError_t memset_sync(void *p, int v, size_t bytes, hipStream_t s) {
    const hipError_t e = hipMemsetAsync(p, v, bytes, s);
    return e != hipSuccess ? e : hipStreamSynchronize(s);
}
}

#endif
