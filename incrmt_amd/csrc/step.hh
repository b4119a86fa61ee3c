// Kernel launchers of the per-timestep sweeps (step.hip) and the host-side
// argument bundles they take.
#ifndef INCRMT_STEP_HH
#define INCRMT_STEP_HH

#include "common.hh"

namespace incrmt {

// The fluid fields (fields.hh:7-49) as separate device arrays. Two sets of
// edge values: A holds the normal-derivative extrapolations and their first
// upwinding (the tangential-stability pass), B the half-time edge values
// that the projections and the updates read.
struct Fields {
    double *u, *v, *p, *rho;
    double *ux, *uy, *vx, *vy;
    double *c0, *c1;
    double *s11, *s12, *s21, *s22;
    double *eA[8];  // ul vl ur vr ud vd uu vu
    double *eB[8];
    // MAC-corrected left and down faces (written beside eB, so that the
    // fused correction never overwrites a face a neighbour still reads)
    double *ulc, *vdc;
};
enum { UL = 0, VL, UR, VR, UD, VD, UU, VU };

struct Phys {
    double visc, rhof, eps;
    double w, winv, k_rep;  // wall repulsion (fluid_2d.cc:1049)
    // per-solid state of object::pre_stress_setup at the current time:
    // simple spin (cos, sin) of the anchor rotation; flapper (con, 0);
    // lamprey (con, sh)
    double pss[2 * kMaxObj];
    const double *xrow;       // x of cell i formed as compute_ustar forms it
};

struct Solids {
    const ObjConst *oc;  // device
    const ObjArrays *oa; // device
    int n;
};

int launch_normal_derivs(hipStream_t s, const Grid &g, const Fields &f, Solids o, double dt);
int launch_godunov(hipStream_t s, const Grid &g, const Fields &f, Solids o, bool tang);
int launch_tang_half(hipStream_t s, const Grid &g, const Fields &f, Solids o, const Phys &ph, double dt);
int launch_mac_src(hipStream_t s, const Grid &g, const Fields &f, double *b, int bp);
int launch_mac_correct_predict(hipStream_t s, const Grid &g, const Fields &f, Solids o,
                               const double *z, int zp, const double *fw, const double *fs, int cp,
                               bool const_rho, double dt);
int launch_half_timestep(hipStream_t s, const Grid &g, Solids o);
int launch_set_rho(hipStream_t s, const Grid &g, const Fields &f, Solids o, double rhof);
int launch_stress(hipStream_t s, const Grid &g, const Fields &f, Solids o, const Phys &ph);
int launch_ustar(hipStream_t s, const Grid &g, const Fields &f, Solids o, const Phys &ph, double dt);
int launch_fem_src(hipStream_t s, const Grid &g, const Fields &f, double dt, double *b, int bp);
int launch_pressure_velocity(hipStream_t s, const Grid &g, const Fields &f, Solids o,
                             const double *z, int zp, const double *pavg, double dt);
// set_boundaries (fluid_2d.cc:689-726) with set_boundaries_solid first when
// the grid has solid periodic boundaries
int launch_set_boundaries(hipStream_t s, const Grid &g, const Fields &f, Solids o);
int launch_set_boundaries_solid(hipStream_t s, const Grid &g, Solids o);
int launch_init_velocity(hipStream_t s, const Grid &g, const Fields &f, Solids o, double rhof);

}  // namespace incrmt

#endif
