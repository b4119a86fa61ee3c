// Tracers of the IncRMT step on the GPU: the passive fluid tracers of
// fluid_2d (fluid_2d.cc:1376-1484) and the arm tracer that obj_simple_spin
// records to track.dat (object.hh:289-365, 422-467; object.cc:202-219,
// 313-417). Both move by the improved Euler method on the bicubic
// interpolation of the velocity (bi_interp.cc); the arm tracer is snapped
// back onto its anchored reference-map point by Newton's method on the
// bicubic interpolation of the reference map whenever it is recorded.
#ifndef INCRMT_TRACERS_HH
#define INCRMT_TRACERS_HH

#include <cstdint>
#include "common.hh"

namespace incrmt {

// glibc's rand(): random(3) with its default TYPE_3 state, an additive
// lagged-Fibonacci generator r[k] = r[k-31] + r[k-3] (mod 2^32) returning
// r[k] >> 1, whose 31 words are seeded by the Park-Miller sequence
// 16807^k seed mod (2^31-1) and then stepped 310 times. The reference never
// seeds it (seed 1) and draws from it in common.hh:28-30.
struct GlibcRand {
    int32_t s[31];
    int f, r;
    __host__ __device__ void seed(uint32_t sd) {
        int32_t w = sd ? (int32_t)sd : 1;
        s[0] = w;
        for (int i = 1; i < 31; i++) {  // Schrage's form of 16807 w mod (2^31-1)
            const int32_t hi = w / 127773, lo = w % 127773;
            w = 16807 * lo - 2836 * hi;
            if (w < 0) w += 2147483647;
            s[i] = w;
        }
        f = 3;
        r = 0;
        for (int i = 0; i < 310; i++) next();
    }
    __host__ __device__ int next() {
        const uint32_t v = (uint32_t)s[f] + (uint32_t)s[r];
        s[f] = (int32_t)v;
        f = f == 30 ? 0 : f + 1;
        r = r == 30 ? 0 : r + 1;
        return (int)(v >> 1);
    }
    // rnd(a,b) of common.hh:28-30 (RAND_MAX = 2^31-1)
    __host__ __device__ double rnd(double a, double b) {
        return a + ((b - a) / 2147483647.) * static_cast<double>(next());
    }
};

// Geometry of the bicubic interpolants of fluid_2d and obj_track: the
// ghosted (m+4) x (n+4) grid spanning [ax-1.5dx, bx+1.5dx] x [ay-1.5dy,
// by+1.5dy] (fluid_2d.cc:43, object.cc:321-324).
struct BicGeom {
    int m, n;           // points: m+4, n+4
    double ax, ay;      // origin of point (0,0)
    double xsp, ysp;    // (points-1)/(extent), as bi_interp.cc:8 forms them
    double bx_lo, bx_hi, by_lo, by_hi;  // the box, for tracer_remap
};

// One obj_track with one tracer (obj_simple_spin: every 10 steps).
struct TrackDev {
    double tk[6];  // position, stage velocity, anchored reference map point
    int obj;       // solid whose reference map anchors it
    int count, every, prev_q, wind, tcorr;
    int nrec;      // records written (may exceed the buffer; later ones are dropped)
};

struct TracerSet {
    int ntrace = 0;
    double *tm = nullptr;      // x, y, stage u, stage v per tracer
    GlibcRand *rng = nullptr;  // the generator's state after initialisation
    int *nout = nullptr;       // tracers outside the box after a step
    int ntrack = 0;
    TrackDev *trk = nullptr;
    double *rec = nullptr;     // ntrack x rec_cap records of 4 doubles
    int rec_cap = 0;
};

BicGeom bic_geom(const Grid &g);
// update_tracers(1,dt) (fluid_2d.cc:1466-1472, object.cc:351-353)
int tracers_stage1(hipStream_t s, const Grid &g, const double *u, const double *v, TracerSet &t);
// update_tracers(2,dt) (fluid_2d.cc:1473-1483, object.cc:354-361); time is
// the end-of-step time the records carry.
int tracers_stage2(hipStream_t s, const Grid &g, const double *u, const double *v, const ObjArrays *d_oa,
                   TracerSet &t, double dt, double time);
// obj_track::write for every track (fluid_2d_io.cc:101, object.cc:330-336)
int tracers_frame_output(hipStream_t s, const Grid &g, const ObjArrays *d_oa, TracerSet &t, double time);

}  // namespace incrmt

#endif
