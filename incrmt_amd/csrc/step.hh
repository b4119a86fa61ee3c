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

// A solid's sweep window: the cells [i0, i1) x [j0, j1) (ghost cells are
// -2, -1 and m, m+1) outside which the solid leaves nothing to compute.
// Written on the device after every band rebuild (launch_solid_window):
// the band's bounding box (statuses 1..6) widened by kWinMargin cells; where
// that crosses the grid's edges it is clipped to the ghosted grid when the
// solid lies on the side of its interface the box encloses, else it is the
// whole ghosted grid (two opposite edges crossed, the solid on the far side,
// a periodic grid, no band). Inside a closed interface the solid's interior
// lies in that box, so outside the window every point of
// the solid is far outside it (status 7, phi at the positive cut-off, above
// the transition half-width): no transition function, wall force,
// stress, collision, reference-map update or density term of the solid is
// non-zero there.
struct SolidWin {
    int i0, i1, j0, j1;
};
constexpr int kWinMargin = 6;
__device__ inline bool in_win(const SolidWin &w, int i, int j) {
    return i >= w.i0 && i < w.i1 && j >= w.j0 && j < w.j1;
}

struct Solids {
    const ObjConst *oc;  // device
    const ObjArrays *oa; // device
    int n;
    const SolidWin *win; // device, n windows
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
// solid k's window from the band box `bbox` (level-set coordinates, as the
// rebuild leaves it on the device), or the whole grid (full)
// (st: the solid's status codes, which tell on which side of an interface
// cut by the grid's edge the solid lies)
int launch_solid_window(hipStream_t s, const Grid &g, const int *bbox, const int *st, SolidWin *win, bool full);
// every solid's window (n <= 64 solids; bbox: device array of the solids'
// band boxes; oa: their fields; bit k of full: solid k takes the whole grid)
int launch_solid_windows(hipStream_t s, const Grid &g, const int *const *bbox, const ObjArrays *oa, SolidWin *win,
                         int n, unsigned long long full);

}  // namespace incrmt

#endif
