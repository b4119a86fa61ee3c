// Per-timestep sweeps of fluid_2d::step_forward (fluid_2d.cc:298-369) as
// HIP kernels for gfx950. One thread per cell (or edge, or node), 64-wide
// rows so that each wavefront reads contiguous 512 B row segments of every
// field it touches. All arithmetic is FP64 and follows the reference's
// expression order (the library is compiled without FP contraction), so a
// sweep reproduces the reference's values bit for bit on identical inputs
// wherever only +,-,*,/,sqrt and min/max are involved.
#include "step.hh"

namespace incrmt {

namespace {

constexpr int BX = 64, BY = 4;

inline dim3 grid2(int nx, int ny) { return dim3((nx + BX - 1) / BX, (ny + BY - 1) / BY); }

// The solids whose windows meet this wavefront's cells: a wavefront is one
// row j of a BX x BY workgroup (BX = 64 lanes), the cells [i0, i0 + 64) in
// window coordinates; lane k tests solid k (n <= kMaxObj = 64), one ballot.
// Every lane of the workgroup must call it (before any early return). A
// loop over the set bits in increasing order visits the same solids, in the
// same order, as a loop over all solids that skips those outside their
// windows, since a solid missing from the mask is outside its window at
// every cell of the row; the per-cell window test stays.
static_assert(BX == 64, "one wavefront per row of the workgroup");
__device__ inline unsigned long long row_solids(const Solids &o, int i0, int j) {
    const int k = threadIdx.x;
    bool hit = false;
    if (k < o.n) {
        const SolidWin w = o.win[k];
        hit = j >= w.j0 && j < w.j1 && w.i0 < i0 + BX && w.i1 > i0;
    }
    return __ballot(hit);
}
__device__ inline int low_bit(unsigned long long m) { return __ffsll((long long)m) - 1; }

// ---- monotonicity-limited fourth-order slope (fluid_2d.cc:604-634)
__device__ inline double delta_lim(double f0, double f1, double f2) {
    double drm = f1 - f0, drp = f2 - f1;
    return drm > 0 ? (drp > 0 ? 2 * fmin(drm, drp) : 0) : (drp > 0 ? 0 : -2 * fmax(drm, drp));
}
__device__ inline double delta_f(double f0, double f1, double f2) {
    double drc = 0.5 * (f2 - f0);
    return drc > 0 ? fmin(drc, delta_lim(f0, f1, f2)) : fmax(drc, -delta_lim(f0, f1, f2));
}
__device__ inline double mono_diff(double f0, double f1, double f2, double f3, double f4) {
    double tdrc = f3 - f1;
    double s = (1 / 6.) * (4 * tdrc - delta_f(f0, f1, f2) - delta_f(f2, f3, f4));
    double t = delta_lim(f1, f2, f3);
    return tdrc > 0 ? fmin(s, t) : fmax(s, -t);
}

// ---- fluid_2d::normal_derivatives_velocity / _refmap (fluid_2d.cc:877-950).
// The reference-map edge extrapolations of the second routine are never
// read (compute_half_time_edge_reference_map overwrites them), so only its
// slopes are computed.
__global__ void k_normal_derivs(Grid g, Fields f, Solids o, double dt) {
    const int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y;
    if (i >= g.m || j >= g.n) return;
    const int ij = g.at(i, j), P = g.P;
    const double *u = f.u, *v = f.v;
    const double uc = u[ij], vc = v[ij];
    const double ux = g.xsp * mono_diff(u[ij - 2], u[ij - 1], uc, u[ij + 1], u[ij + 2]);
    const double vx = g.xsp * mono_diff(v[ij - 2], v[ij - 1], vc, v[ij + 1], v[ij + 2]);
    const double uy = g.ysp * mono_diff(u[ij - 2 * P], u[ij - P], uc, u[ij + P], u[ij + 2 * P]);
    const double vy = g.ysp * mono_diff(v[ij - 2 * P], v[ij - P], vc, v[ij + P], v[ij + 2 * P]);
    f.ux[ij] = ux; f.vx[ij] = vx; f.uy[ij] = uy; f.vy[ij] = vy;
    f.eA[UL][ij] = uc - 0.5 * (g.dx + dt * uc) * ux;
    f.eA[VL][ij] = vc - 0.5 * (g.dx + dt * uc) * vx;
    f.eA[UR][ij] = uc + 0.5 * (g.dx - dt * uc) * ux;
    f.eA[VR][ij] = vc + 0.5 * (g.dx - dt * uc) * vx;
    f.eA[UD][ij] = uc - 0.5 * (g.dy + dt * vc) * uy;
    f.eA[VD][ij] = vc - 0.5 * (g.dy + dt * vc) * vy;
    f.eA[UU][ij] = uc + 0.5 * (g.dy - dt * vc) * uy;
    f.eA[VU][ij] = vc + 0.5 * (g.dy - dt * vc) * vy;
}

// Per-solid sweeps over the solid's window only: workgroup (x, z) walks the
// BX x BY tiles x, x + gridDim.x, ... of solid z's window (clipped to the
// cells [0, m) x [0, n)). The windows live on the device, so the launch is
// sized by the solid count: about 4096 workgroups in all (a lone solid's
// window of a large grid gets one or two tiles per workgroup, many solids
// at least 96 workgroups each); a workgroup past its window's tiles does
// nothing. Never more than the grid's own tile count.
constexpr int kWinBlocks = 96, kWinTotal = 4096;
inline int win_blocks(const Grid &g, int n) {
    int b = n > 0 && kWinTotal / n > kWinBlocks ? kWinTotal / n : kWinBlocks;
    const int tiles = ((g.m + BX - 1) / BX) * ((g.n + BY - 1) / BY);
    if (b > tiles) b = tiles;
    return b < 1 ? 1 : b;
}
template <class F>
__device__ inline void window_cells(const Grid &g, const SolidWin &w0, F f) {
    const int i0 = max(w0.i0, 0), i1 = min(w0.i1, g.m), j0 = max(w0.j0, 0), j1 = min(w0.j1, g.n);
    const int tw = (i1 - i0 + BX - 1) / BX, th = (j1 - j0 + BY - 1) / BY;
    for (int t = blockIdx.x; t < tw * th; t += gridDim.x) {
        const int i = i0 + (t % tw) * BX + threadIdx.x, j = j0 + (t / tw) * BY + threadIdx.y;
        if (i < i1 && j < j1) f(i, j);
    }
}

// fluid_2d::normal_derivatives_refmap's slopes (fluid_2d.cc:915-950)
__global__ void k_refmap_slopes(Grid g, Solids o) {
    const int k = blockIdx.z, P = g.P;
    const ObjArrays a = o.oa[k];
    window_cells(g, o.win[k], [&](int i, int j) {
        const int ij = g.at(i, j);
        const double *X = a.X, *Y = a.Y;
        a.Xx[ij] = g.xsp * mono_diff(X[ij - 2], X[ij - 1], X[ij], X[ij + 1], X[ij + 2]);
        a.Yx[ij] = g.xsp * mono_diff(Y[ij - 2], Y[ij - 1], Y[ij], Y[ij + 1], Y[ij + 2]);
        a.Xy[ij] = g.ysp * mono_diff(X[ij - 2 * P], X[ij - P], X[ij], X[ij + P], X[ij + 2 * P]);
        a.Yy[ij] = g.ysp * mono_diff(Y[ij - 2 * P], Y[ij - P], Y[ij], Y[ij + P], Y[ij + 2 * P]);
    });
}

// ---- fluid_2d::godunov_upwinding_set<tang> + edge_boundary_conditions
// (fluid_2d.cc:642-686, 729-779, 1112-1121), walls or seams. Thread (i,j)
// owns the face to its right and the face above it; every value it reads or
// writes belongs to those two faces, so the faces are independent. The wall
// faces that the boundary conditions zero are disjoint from the faces the
// upwinding touches, so both run in one pass. With tang the reference-map
// copies are skipped: their values are overwritten before anything reads them.
template <bool TANG>
__global__ void k_godunov(Grid g, Fields f, Solids o) {
    const int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y;
    if (i >= g.m || j >= g.n) return;
    const int ij = g.at(i, j), P = g.P;
    double *const *e = TANG ? f.eA : f.eB;
    if (i < g.m - 1) {
        const double ur = e[UR][ij], ulr = e[UL][ij + 1];
        if (ur + ulr > 0 && ur > 0) {
            e[UL][ij + 1] = ur;
            e[VL][ij + 1] = e[VR][ij];
        } else if (ur + ulr < 0 && ulr < 0) {
            if (TANG) e[UR][ij] = ulr;
        } else {
            if (TANG) e[UR][ij] = 0.5 * (ur + ulr);
            e[UL][ij + 1] = 0;
            e[VL][ij + 1] = 0.5 * (e[VR][ij] + e[VL][ij + 1]);
        }
    }
    if (j < g.n - 1) {
        const double vu = e[VU][ij], vdu = e[VD][ij + P];
        if (vu + vdu > 0 && vu > 0) {
            e[VD][ij + P] = vu;
            e[UD][ij + P] = e[UU][ij];
        } else if (vu + vdu < 0 && vdu < 0) {
            if (TANG) e[VU][ij] = vdu;
        } else {
            if (TANG) e[VU][ij] = 0.5 * (vu + vdu);
            e[VD][ij + P] = 0;
            e[UD][ij + P] = 0.5 * (e[UU][ij] + e[UD][ij + P]);
        }
    }
    // no-slip wall faces (fluid_2d.cc:754-776)
    if (g.zpr) {
        // horizontal flow (fluid_2d.cc:745-749): the inflow face keeps cell
        // 0's own left state and the ghost cell takes it as its right state;
        // the outflow ghost cell takes cell m-1's right state. No thread of
        // the upwinding writes cell 0's left or cell m-1's right state.
        if (i == 0) e[UR][ij - 1] = e[UL][ij];
        if (i == g.m - 1) { e[UL][ij + 1] = e[UR][ij]; e[VL][ij + 1] = e[VR][ij]; }
    } else if (!g.px) {
        if (i == 0) { e[UL][ij] = 0; e[VL][ij] = 0; e[UR][ij - 1] = 0; }
        if (i == g.m - 1) { e[UL][ij + 1] = 0; e[VL][ij + 1] = 0; e[UR][ij] = 0; }
    } else if (i == g.m - 1) {
        // periodic seam (fluid_2d.cc:733-744): ghost cell -1 takes cell
        // m-1's right state, godunov_set_lr<true> runs on the face between
        // it and cell 0, ghost cell m takes cell 0's new left state and cell
        // m-1 the ghost's right state. Nothing else in the sweep reads or
        // writes these values, so the seam of row j is this thread's.
        const int r0 = g.at(0, j), gl = r0 - 1, gr = ij + 1;
        if (!TANG && g.spbc)  // edge_boundary_conditions_solid_pre (fluid_2d.cc:784-797)
            for (int k = 0; k < o.n; k++) {
                const ObjArrays &a = o.oa[k];
                a.Xr[gl] = a.Xr[ij] - (g.bx - g.ax); a.Yr[gl] = a.Yr[ij];
            }
        double ur = e[UR][ij], vr = e[VR][ij], ul0 = e[UL][r0], vl0 = e[VL][r0];
        if (ur + ul0 > 0 && ur > 0) {
            ul0 = ur;
            vl0 = vr;
            if (!TANG)
                for (int k = 0; k < o.n; k++) {
                    const ObjArrays &a = o.oa[k];
                    a.Xl[r0] = a.Xr[gl]; a.Yl[r0] = a.Yr[gl];
                }
        } else if (ur + ul0 < 0 && ul0 < 0) {
            ur = ul0;
        } else {
            ur = 0.5 * (ur + ul0);
            ul0 = 0;
            vl0 = 0.5 * (vr + vl0);
            if (!TANG)
                for (int k = 0; k < o.n; k++) {
                    const ObjArrays &a = o.oa[k];
                    a.Xl[r0] = 0.5 * (a.Xr[gl] + a.Xl[r0]);
                    a.Yl[r0] = 0.5 * (a.Yr[gl] + a.Yl[r0]);
                }
        }
        e[UR][gl] = ur; e[VR][gl] = vr;
        e[UL][r0] = ul0; e[VL][r0] = vl0;
        e[UL][gr] = ul0; e[VL][gr] = vl0;
        e[UR][ij] = ur;
        if (!TANG && g.spbc)  // edge_boundary_conditions_solid_post (fluid_2d.cc:802-815)
            for (int k = 0; k < o.n; k++) {
                const ObjArrays &a = o.oa[k];
                a.Xl[gr] = a.Xl[r0] + (g.bx - g.ax); a.Yl[gr] = a.Yl[r0];
            }
    }
    if (!g.py) {
        if (j == 0) { e[UD][ij] = 0; e[VD][ij] = 0; e[VU][ij - P] = 0; }
        if (j == g.n - 1) { e[UD][ij + P] = 0; e[VD][ij + P] = 0; e[VU][ij] = 0; }
    } else if (j == g.n - 1) {
        // periodic seam of column i (fluid_2d.cc:760-769), as above
        const int r0 = g.at(i, 0), gd = r0 - P, gu = ij + P;
        if (!TANG && g.spbc)
            for (int k = 0; k < o.n; k++) {
                const ObjArrays &a = o.oa[k];
                a.Xu[gd] = a.Xu[ij]; a.Yu[gd] = a.Yu[ij] - (g.by - g.ay);
            }
        double vu = e[VU][ij], uu = e[UU][ij], vd0 = e[VD][r0], ud0 = e[UD][r0];
        if (vu + vd0 > 0 && vu > 0) {
            vd0 = vu;
            ud0 = uu;
            if (!TANG)
                for (int k = 0; k < o.n; k++) {
                    const ObjArrays &a = o.oa[k];
                    a.Xd[r0] = a.Xu[gd]; a.Yd[r0] = a.Yu[gd];
                }
        } else if (vu + vd0 < 0 && vd0 < 0) {
            vu = vd0;
        } else {
            vu = 0.5 * (vu + vd0);
            vd0 = 0;
            ud0 = 0.5 * (uu + ud0);
            if (!TANG)
                for (int k = 0; k < o.n; k++) {
                    const ObjArrays &a = o.oa[k];
                    a.Xd[r0] = 0.5 * (a.Xu[gd] + a.Xd[r0]);
                    a.Yd[r0] = 0.5 * (a.Yu[gd] + a.Yd[r0]);
                }
        }
        e[VU][gd] = vu; e[UU][gd] = uu;
        e[VD][r0] = vd0; e[UD][r0] = ud0;
        e[VD][gu] = vd0; e[UD][gu] = ud0;
        e[VU][ij] = vu;
        if (!TANG && g.spbc)
            for (int k = 0; k < o.n; k++) {
                const ObjArrays &a = o.oa[k];
                a.Xd[gu] = a.Xd[r0]; a.Yd[gu] = a.Yd[r0] + (g.by - g.ay);
            }
    }
}

// The reference-map half of godunov_upwinding_set<false> on the interior
// faces (fluid_2d.cc:642-686, 1112-1121): per solid over its window, run
// before the velocity pass so that it reads the same unmodified face
// velocities that pass decides on.
__global__ void k_refmap_upwind(Grid g, Fields f, Solids o) {
    const int k = blockIdx.z, P = g.P;
    const ObjArrays a = o.oa[k];
    double *const *e = f.eB;
    window_cells(g, o.win[k], [&](int i, int j) {
        const int ij = g.at(i, j);
        if (i < g.m - 1) {
            const double ur = e[UR][ij], ulr = e[UL][ij + 1];
            if (ur + ulr > 0 && ur > 0) {
                a.Xl[ij + 1] = a.Xr[ij]; a.Yl[ij + 1] = a.Yr[ij];
            } else if (!(ur + ulr < 0 && ulr < 0)) {
                a.Xl[ij + 1] = 0.5 * (a.Xr[ij] + a.Xl[ij + 1]);
                a.Yl[ij + 1] = 0.5 * (a.Yr[ij] + a.Yl[ij + 1]);
            }
        }
        if (j < g.n - 1) {
            const double vu = e[VU][ij], vdu = e[VD][ij + P];
            if (vu + vdu > 0 && vu > 0) {
                a.Xd[ij + P] = a.Xu[ij]; a.Yd[ij + P] = a.Yu[ij];
            } else if (!(vu + vdu < 0 && vdu < 0)) {
                a.Xd[ij + P] = 0.5 * (a.Xu[ij] + a.Xd[ij + P]);
                a.Yd[ij + P] = 0.5 * (a.Yu[ij] + a.Yd[ij + P]);
            }
        }
    });
}

// ---- fluid_2d::acceleration + wall_force + object::accel
// (fluid_2d.cc:1019-1078, object.cc:31-34, 188-197)
__device__ inline void wall_force(const Grid &g, const Phys &ph, double x, double y, double phiv,
                                  double &fx, double &fy) {
    auto dfw = [](double q) { return (1. - q) / q; };
    if (x <= g.ax + ph.w) {
        fx += ph.k_rep * dfw((x - g.ax) * ph.winv) * (phiv > -g.dx ? (phiv > g.dx ? 0 : 0.5 * (1 - phiv * g.xsp)) : 1);
    } else if (x >= g.bx - ph.w) {
        fx -= ph.k_rep * dfw((g.bx - x) * ph.winv) * (phiv > -g.dx ? (phiv > g.dx ? 0 : 0.5 * (1 - phiv * g.xsp)) : 1);
    }
    if (y <= g.ay + ph.w) {
        fy += ph.k_rep * dfw((y - g.ay) * ph.winv) * (phiv > -g.dy ? (phiv > g.dy ? 0 : 0.5 * (1 - phiv * g.ysp)) : 1);
    } else if (y >= g.by - ph.w) {
        fy -= ph.k_rep * dfw((g.by - y) * ph.winv) * (phiv > -g.dy ? (phiv > g.dy ? 0 : 0.5 * (1 - phiv * g.ysp)) : 1);
    }
}

__device__ inline void acceleration(const Grid &g, const Fields &f, const Solids &o, const Phys &ph,
                                    unsigned long long near, int i, int j, int ij, double &accx, double &accy,
                                    double x, double y, bool pressure) {
    const int P = g.P;
    double accx_ = 0, accy_ = 0;
    accx = g.xsp * (f.s11[ij + 1] - f.s11[ij]) + g.ysp * (f.s12[ij + P] - f.s12[ij]);
    accy = g.xsp * (f.s21[ij + 1] - f.s21[ij]) + g.ysp * (f.s22[ij + P] - f.s22[ij]);
    if (pressure) {
        const double *p = f.p;
        accx -= 0.5 * g.xsp * (p[ij + P + 1] + p[ij + 1] - p[ij + P] - p[ij]);
        accy -= 0.5 * g.ysp * (p[ij + P + 1] - p[ij + 1] + p[ij + P] - p[ij]);
    }
    for (unsigned long long mk = near; mk; mk &= mk - 1) {
        const int k = low_bit(mk);
        if (!in_win(o.win[k], i, j)) continue;
        const ObjConst &c = o.oc[k];
        const double phiv = o.oa[k].phi[ij];
        if (phiv < c.eps && !g.spbc) {  // fluid_2d.cc:1037
            wall_force(g, ph, x, y, phiv, accx, accy);
            if (c.kind == kSpin) {
                double rsq = x * x + y * y;
                if (rsq < c.piv_thresh) {
                    const double X = o.oa[k].X[ij], Y = o.oa[k].Y[ij];
                    const double cth = ph.pss[2 * k], sth = ph.pss[2 * k + 1];
                    double rx = X * cth - Y * sth, ry = X * sth + Y * cth,
                           K = c.K_stiff * trans_func_in(c, sqrt(rsq) - c.piv_r);
                    accx_ -= K * (x - rx);
                    accy_ -= K * (y - ry);
                }
            } else {
                accy_ -= c.grav * trans_func_in(c, phiv);
            }
        }
    }
    const double rho = f.rho[ij];
    accx /= rho; accy /= rho;
    accx += accx_;
    accy += accy_;
}

// ---- compute_tangential_stability_derivatives fused with
// compute_half_time_edge_velocities and _reference_map
// (fluid_2d.cc:953-1014, 1082-1108). The tangential terms c0..c3 of a cell
// only feed that cell's half-time extrapolation, so they live in registers;
// reading the A edges and writing the B edges keeps neighbours race-free.
__global__ void k_tang_half(Grid g, Fields f, Solids o, Phys ph, double dt) {
    const int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y;
    const unsigned long long near = row_solids(o, blockIdx.x * BX, j);
    if (i >= g.m || j >= g.n) return;
    const int ij = g.at(i, j), P = g.P;
    double *const *A = f.eA;
    const double u_adv = 0.5 * g.xsp * (A[UL][ij + 1] + A[UL][ij]);
    const double v_adv = 0.5 * g.ysp * (A[VD][ij + P] + A[VD][ij]);
    const double c0 = u_adv * (A[UR][ij] - A[UR][ij - 1]);
    const double c1 = u_adv * (A[VL][ij + 1] - A[VL][ij]);
    const double c2 = v_adv * (A[UD][ij + P] - A[UD][ij]);
    const double c3 = v_adv * (A[VU][ij] - A[VU][ij - P]);
    const double y = g.ay + g.dy * (0.5 + j), x = g.ax + g.dx * (0.5 + i);
    double accx, accy;
    acceleration(g, f, o, ph, near, i, j, ij, accx, accy, x, y, true);
    const double u = f.u[ij], v = f.v[ij], ux = f.ux[ij], vx = f.vx[ij], uy = f.uy[ij], vy = f.vy[ij];
    double *const *B = f.eB;
    B[UL][ij] = u + 0.5 * (-g.dx * ux + dt * (-u * ux - c2 + accx));
    B[VL][ij] = v + 0.5 * (-g.dx * vx + dt * (-u * vx - c3 + accy));
    B[UR][ij] = u + 0.5 * (g.dx * ux + dt * (-u * ux - c2 + accx));
    B[VR][ij] = v + 0.5 * (g.dx * vx + dt * (-u * vx - c3 + accy));
    B[UD][ij] = u + 0.5 * (-g.dy * uy + dt * (-v * uy - c0 + accx));
    B[VD][ij] = v + 0.5 * (-g.dy * vy + dt * (-v * vy - c1 + accy));
    B[UU][ij] = u + 0.5 * (g.dy * uy + dt * (-v * uy - c0 + accx));
    B[VU][ij] = v + 0.5 * (g.dy * vy + dt * (-v * vy - c1 + accy));
}

// compute_half_time_edge_reference_map (fluid_2d.cc:1082-1108), per solid
// over its window
__global__ void k_refmap_half(Grid g, Fields f, Solids o, double dt) {
    const int k = blockIdx.z;
    const ObjArrays a = o.oa[k];
    window_cells(g, o.win[k], [&](int i, int j) {
        const int ij = g.at(i, j);
        const double u = f.u[ij], v = f.v[ij];
        const double X = a.X[ij], Y = a.Y[ij], Xx = a.Xx[ij], Xy = a.Xy[ij], Yx = a.Yx[ij], Yy = a.Yy[ij];
        double Xdt = dt * (-u * Xx - v * Xy), Ydt = dt * (-u * Yx - v * Yy);
        a.Xl[ij] = X + 0.5 * (-g.dx * Xx + Xdt);
        a.Yl[ij] = Y + 0.5 * (-g.dx * Yx + Ydt);
        a.Xr[ij] = X + 0.5 * (g.dx * Xx + Xdt);
        a.Yr[ij] = Y + 0.5 * (g.dx * Yx + Ydt);
        a.Xd[ij] = X + 0.5 * (-g.dy * Xy + Xdt);
        a.Yd[ij] = Y + 0.5 * (-g.dy * Yy + Ydt);
        a.Xu[ij] = X + 0.5 * (g.dy * Xy + Xdt);
        a.Yu[ij] = Y + 0.5 * (g.dy * Yy + Ydt);
    });
}

// ---- MAC source: the divergence of the half-time face velocities
// (fluid_2d.cc:1179-1187), written into the solver's right-hand side.
__global__ void k_mac_src(Grid g, Fields f, double *b, int bp) {
    const int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y;
    if (i >= g.m || j >= g.n) return;
    const int ij = g.at(i, j);
    b[j * bp + i] = g.xsp * (f.eB[UL][ij + 1] - f.eB[UL][ij]) + g.ysp * (f.eB[VD][ij + g.P] - f.eB[VD][ij]);
}

// ---- MAC correction (fluid_2d.cc:1193-1235) fused with the
// update_reference_map predictor (fluid_2d.cc:1244-1256). The predictor
// needs the corrected right and upper faces, which belong to the
// neighbouring threads; it recomputes them with the same expression instead
// of reading them, so every thread writes only its own faces.
__global__ void k_mac_correct_predict(Grid g, Fields f, Solids o, const double *z, int zp,
                                      const double *fw, const double *fs, int cp, bool const_rho,
                                      double dt) {
    const int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y;
    const unsigned long long near = row_solids(o, blockIdx.x * BX, j);
    if (i >= g.m || j >= g.n) return;
    const int ij = g.at(i, j), P = g.P, q = j * zp + i;
    double *ul = f.eB[UL], *vd = f.eB[VD];
    // corrected left face of cell (ii,jj); a periodic seam is corrected on
    // cell 0 with the opposite edge's q, and ghost cell m repeats it
    auto corr_l = [&](int ii, int jj) -> double {
        if (g.px && ii == g.m) ii = 0;
        const int c = g.at(ii, jj);
        if (!g.px && (ii <= 0 || ii >= g.m)) return ul[c];
        const int im = ii == 0 ? g.m - 1 : ii - 1;
        const double dz = z[jj * zp + ii] - z[jj * zp + im];
        return const_rho ? ul[c] - dz * g.xsp : ul[c] - fw[jj * cp + ii] * dz * g.xsp;
    };
    auto corr_d = [&](int ii, int jj) -> double {
        if (g.py && jj == g.n) jj = 0;
        const int c = g.at(ii, jj);
        if (!g.py && (jj <= 0 || jj >= g.n)) return vd[c];
        const int jm = jj == 0 ? g.n - 1 : jj - 1;
        const double dz = z[jj * zp + ii] - z[jm * zp + ii];
        return const_rho ? vd[c] - dz * g.ysp : vd[c] - fs[jj * cp + ii] * dz * g.ysp;
    };
    const double ulc = corr_l(i, j), vdc = corr_d(i, j);
    const double ulr = corr_l(i + 1, j), vdu = corr_d(i, j + 1);
    (void)q;
    const double hx = 0.5 * dt * g.xsp, hy = 0.5 * dt * g.ysp;
    for (unsigned long long mk = near; mk; mk &= mk - 1) {
        const int k = low_bit(mk);
        if (!in_win(o.win[k], i, j)) continue;
        const ObjArrays &a = o.oa[k];
        if (a.st[ij] >= 4) continue;
        double cX = -hx * (ulr + ulc) * (a.Xl[ij + 1] - a.Xl[ij]) - hy * (vdu + vdc) * (a.Xd[ij + P] - a.Xd[ij]);
        double cY = -hx * (ulr + ulc) * (a.Yl[ij + 1] - a.Yl[ij]) - hy * (vdu + vdc) * (a.Yd[ij + P] - a.Yd[ij]);
        a.nX[ij] = a.X[ij] + cX;
        a.nY[ij] = a.Y[ij] + cY;
    }
    f.ulc[ij] = ulc;
    f.vdc[ij] = vdc;
    if (i == g.m - 1) f.ulc[ij + 1] = g.px ? ulr : ul[ij + 1];  // right wall face, or the seam again
    if (j == g.n - 1) f.vdc[ij + P] = g.py ? vdu : vd[ij + P];  // top wall face, or the seam again
}

// ---- update_reference_map's last loop (fluid_2d.cc:1265-1272)
__global__ void k_half_timestep(Grid g, Solids o) {
    const int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y;
    const unsigned long long near = row_solids(o, blockIdx.x * BX, j);
    if (i >= g.m || j >= g.n) return;
    const int ij = g.at(i, j);
    for (unsigned long long mk = near; mk; mk &= mk - 1) {
        const int k = low_bit(mk);
        if (!in_win(o.win[k], i, j)) continue;
        const ObjArrays &a = o.oa[k];
        if (a.st[ij] < 6) {
            a.X[ij] = 0.5 * (a.X[ij] + a.nX[ij]);
            a.Y[ij] = 0.5 * (a.Y[ij] + a.nY[ij]);
        }
    }
}

// ---- fluid_2d::set_rho over the ghosted grid (fluid_2d.cc:1124-1137)
__global__ void k_set_rho(Grid g, Fields f, Solids o, double rhof) {
    const int I = blockIdx.x * BX + threadIdx.x, J = blockIdx.y * BY + threadIdx.y;
    const unsigned long long near = row_solids(o, blockIdx.x * BX - 2, J - 2);
    if (I >= g.ml || J >= g.nl) return;
    const int ij = g.ls(I, J);
    double rho = 0, sfrac = 0;
    for (unsigned long long mk = near; mk; mk &= mk - 1) {
        const int k = low_bit(mk);
        if (!in_win(o.win[k], I - 2, J - 2)) continue;
        double tf = trans_func(o.oc[k], o.oa[k].phi[ij]);
        sfrac += tf;
        rho += tf * o.oc[k].rhos;
    }
    f.rho[ij] = sfrac >= 1 ? rho / sfrac : (1 - sfrac) * rhof + rho;
}

// ---- fluid_2d::compute_stress (fluid_2d.cc:381-600) on the (m+1)x(n+1)
// left and down edges, with solid_stress, fluid_stress and collision_stress.
// (a solid takes part at an edge only inside its window: the edge's own
// cell (i, j) is in it, and with it, for any edge where the solid is
// present, the neighbour across the edge)
template <bool LEFT>
__device__ inline void edge_stress(const Grid &g, const Fields &f, const Solids &o, const Phys &ph,
                                   unsigned long long near, int i, int j, int ij, double &a, double &b) {
    const int P = g.P, d = LEFT ? 1 : P;
    double s1s = 0, s2s = 0, sfrac = 0;
    int ncol = 0;
    for (unsigned long long mk = near; mk; mk &= mk - 1) {
        const int k = low_bit(mk);
        if (!in_win(o.win[k], i, j)) continue;
        const ObjConst &c = o.oc[k];
        const ObjArrays &A = o.oa[k];
        double phiv = 0.5 * (A.phi[ij] + A.phi[ij - d]);
        if (phiv > c.eps) continue;
        ncol++;
        const double *X = A.X, *Y = A.Y;
        double Xx, Yx, Xy, Yy;
        if (LEFT) {
            Xx = g.xsp * (X[ij] - X[ij - 1]);
            Yx = g.xsp * (Y[ij] - Y[ij - 1]);
            Xy = 0.25 * g.ysp * (X[ij + P] + X[ij + P - 1] - X[ij - P] - X[ij - P - 1]);
            Yy = 0.25 * g.ysp * (Y[ij + P] + Y[ij + P - 1] - Y[ij - P] - Y[ij - P - 1]);
        } else {
            Xx = 0.25 * g.xsp * (X[ij + 1] + X[ij - P + 1] - X[ij - 1] - X[ij - P - 1]);
            Yx = 0.25 * g.xsp * (Y[ij + 1] + Y[ij - P + 1] - Y[ij - 1] - Y[ij - P - 1]);
            Xy = g.ysp * (X[ij] - X[ij - P]);
            Yy = g.ysp * (Y[ij] - Y[ij - P]);
        }
        // F = [Yy,-Xy;-Yx,Xx] with the solid's active part (object::actuate
        // at the edge's reference map), sigma = F F^T (mat.hh:159-161)
        double Fa = Yy, Fb = -Xy, Fc = -Yx, Fd = Xx;
        if (c.kind == kFlapper || c.kind == kLamprey) {
            const double Xe = LEFT ? 0.5 * (X[ij] + X[ij - 1]) : 0.5 * (X[ij] + X[ij - P]);
            const double Ye = LEFT ? 0.5 * (Y[ij] + Y[ij - 1]) : 0.5 * (Y[ij] + Y[ij - P]);
            actuate(c, ph.pss[2 * k], ph.pss[2 * k + 1], Xe, Ye, Fa, Fb, Fc, Fd);
        }
        const double sa = Fa * Fa + Fb * Fb, sb = Fa * Fc + Fb * Fd, sd = Fc * Fc + Fd * Fd;
        double sfac = (1 / 3.0) * ((sa + sd) + 1), tf = trans_func_in(c, phiv), G = c.G * tf,
               ex_visc = c.ex_visc * tf * (1 + c.ev_trans_mult * tderiv_func_in(c, phiv));
        sfrac += tf;
        if (LEFT) {
            s1s += G * (sa - sfac);
            s2s += G * sb;
            s1s += ex_visc * g.xsp * (f.u[ij] - f.u[ij - 1]);
            s2s += ex_visc * g.xsp * (f.v[ij] - f.v[ij - 1]);
        } else {
            s1s += G * sb;
            s2s += G * (sd - sfac);
            s1s += ex_visc * g.ysp * (f.u[ij] - f.u[ij - P]);
            s2s += ex_visc * g.ysp * (f.v[ij] - f.v[ij - P]);
        }
    }
    if (sfrac >= 1) {
        a = s1s / sfrac;
        b = s2s / sfrac;
    } else {
        const double *u = f.u, *v = f.v;
        double s1f, s2f;
        if (LEFT) {
            double ux = g.xsp * (u[ij] - u[ij - 1]);
            double vx = g.xsp * (v[ij] - v[ij - 1]);
            double uy = 0.25 * g.ysp * (u[ij + P] + u[ij + P - 1] - u[ij - P] - u[ij - P - 1]);
            s1f = 2 * ph.visc * ux;
            s2f = ph.visc * (vx + uy);
        } else {
            double vx = 0.25 * g.xsp * (v[ij + 1] + v[ij - P + 1] - v[ij - 1] - v[ij - P - 1]);
            double uy = g.ysp * (u[ij] - u[ij - P]);
            double vy = g.ysp * (v[ij] - v[ij - P]);
            s1f = ph.visc * (vx + uy);
            s2f = 2 * ph.visc * vy;
        }
        a = (1 - sfrac) * s1f + s1s;
        b = (1 - sfrac) * s2f + s2s;
    }
    if (ncol > 1) {
        // collision stress over pairs of solids present at this edge
        // (fluid_2d.cc:455-500); table order is solid order
        auto edge = [&](int k, double &p, double &px, double &py) {
            const double *h = o.oa[k].phi;
            if (LEFT) {
                p = 0.5 * (h[ij] + h[ij - 1]);
                px = g.xsp * (h[ij] - h[ij - 1]);
                py = 0.25 * g.ysp * (h[ij + P] + h[ij + P - 1] - h[ij - P] - h[ij - P - 1]);
            } else {
                p = 0.5 * (h[ij] + h[ij - P]);
                px = 0.25 * g.xsp * (h[ij + 1] + h[ij - P + 1] - h[ij - 1] - h[ij - P - 1]);
                py = g.ysp * (h[ij] - h[ij - P]);
            }
        };
        const double small_number = 100. * 2.220446049250313e-16;
        for (unsigned long long m1 = near; m1; m1 &= m1 - 1) {
            const int k1 = low_bit(m1);
            if (!in_win(o.win[k1], i, j)) continue;
            if (0.5 * (o.oa[k1].phi[ij] + o.oa[k1].phi[ij - d]) > o.oc[k1].eps) continue;
            double p1, p1x, p1y;
            edge(k1, p1, p1x, p1y);
            for (unsigned long long m2 = m1 & (m1 - 1); m2; m2 &= m2 - 1) {
                const int k2 = low_bit(m2);
                if (!in_win(o.win[k2], i, j)) continue;
                if (0.5 * (o.oa[k2].phi[ij] + o.oa[k2].phi[ij - d]) > o.oc[k2].eps) continue;
                double p2, p2x, p2y;
                edge(k2, p2, p2x, p2y);
                double norx = (p2x - p1x), nory = (p2y - p1y), nn = norx * norx + nory * nory,
                       fac = 4. * fmin(coll_func(o.oc[k1], p1), coll_func(o.oc[k2], p2)) * (o.oc[k1].G + o.oc[k2].G);
                fac = nn < small_number ? 0. : fac / nn;
                if (LEFT) {
                    a -= fac * 0.5 * (norx * norx - nory * nory);
                    b -= fac * norx * nory;
                } else {
                    a -= fac * norx * nory;
                    b -= fac * 0.5 * (nory * nory - norx * norx);
                }
            }
        }
    }
}

__global__ void k_stress(Grid g, Fields f, Solids o, Phys ph) {
    const int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y;
    const unsigned long long near = row_solids(o, blockIdx.x * BX, j);
    if (i > g.m || j > g.n) return;
    const int ij = g.at(i, j);
    double a, b;
    edge_stress<true>(g, f, o, ph, near, i, j, ij, a, b);
    f.s11[ij] = a; f.s21[ij] = b;
    edge_stress<false>(g, f, o, ph, near, i, j, ij, a, b);
    f.s12[ij] = a; f.s22[ij] = b;
}

// ---- fluid_2d::compute_ustar (fluid_2d.cc:1278-1291). The reference walks
// x along the row by repeated addition; xrow holds that running sum (formed
// once on the host) so that the wall and anchor terms see the same x.
__global__ void k_ustar(Grid g, Fields f, Solids o, Phys ph, double dt) {
    const int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y;
    const unsigned long long near = row_solids(o, blockIdx.x * BX, j);
    if (i >= g.m || j >= g.n) return;
    const int ij = g.at(i, j), P = g.P;
    const double y = g.ay + g.dy * (.5 + j), x = ph.xrow[i];
    double accx, accy;
    acceleration(g, f, o, ph, near, i, j, ij, accx, accy, x, y, false);
    const double hx = 0.5 * dt * g.xsp, hy = 0.5 * dt * g.ysp;
    const double *ul = f.ulc, *vl = f.eB[VL], *ud = f.eB[UD], *vd = f.vdc;
    f.c0[ij] = f.u[ij] - hx * (ul[ij + 1] + ul[ij]) * (ul[ij + 1] - ul[ij]) - hy * (vd[ij + P] + vd[ij]) * (ud[ij + P] - ud[ij]) + dt * accx;
    f.c1[ij] = f.v[ij] - hx * (ul[ij + 1] + ul[ij]) * (vl[ij + 1] - vl[ij]) - hy * (vd[ij + P] + vd[ij]) * (vd[ij + P] - vd[ij]) + dt * accy;
}

// ---- FEM source at the (m+1)x(n+1) corner nodes (fluid_2d.cc:1297-1309)
// with the no-slip ghost values c0 = c1 = 0 of fem_source_term_conditions
// (fluid_2d.cc:839-872) applied as predicates.
__global__ void k_fem_src(Grid g, Fields f, double dt, double *b, int bp) {
    const int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y;
    // m_fem = m nodes along a periodic direction, m+1 otherwise; with zero
    // pressure on the right edge node m is fixed (its source is zeroed,
    // mgs_fem.cc:85, 112) and not an unknown of the solver
    if (i >= g.m + ((g.px || g.zpr) ? 0 : 1) || j >= g.n + (g.py ? 0 : 1)) return;
    const int ij = g.at(i, j), P = g.P;
    // ghost cells: zero beyond a wall, the opposite edge across a seam, the
    // ghost velocity left of an inflow (fluid_2d.cc:847-851)
    auto c = [&](const double *a, const double *vel, int ii, int jj, int q) -> double {
        if (g.px && ii < 0) { ii += g.m; q += g.m; }
        if (g.py && jj < 0) { jj += g.n; q += g.n * P; }
        if (jj < 0 || jj >= g.n || ii >= g.m) return 0.0;
        if (ii < 0) return g.zpr ? vel[q] : 0.0;
        return a[q];
    };
    const double hx = 0.5 * g.dy / dt, hy = 0.5 * g.dx / dt;
    const double *c0 = f.c0, *c1 = f.c1, *u = f.u, *v = f.v;
    b[j * bp + i] = hx * (c(c0, u, i - 1, j - 1, ij - P - 1) + c(c0, u, i - 1, j, ij - 1) - c(c0, u, i, j - 1, ij - P)
                          - c(c0, u, i, j, ij))
                    + hy * (c(c1, v, i - 1, j - 1, ij - P - 1) - c(c1, v, i - 1, j, ij - 1)
                            + c(c1, v, i, j - 1, ij - P) - c(c1, v, i, j, ij));
}

// ---- copy_pressure + update_velocity + set_full_timestep
// (fluid_2d.cc:1317-1374). Thread (i,j) writes the corner pressure p(i,j)
// and updates cell (i,j) from its four corners, each formed as z - pavg
// exactly as copy_pressure stores it.
// Along a periodic direction node m is node 0: the solver's ghost ring
// mirrors it, so reading z there is copy_pressure's fp[-m].p. With zero
// pressure on the right edge node m is the solver's zero ring column.
__global__ void k_pressure_velocity(Grid g, Fields f, Solids o, const double *z, int zp,
                                    const double *pavg_p, double dt) {
    const int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y;
    const unsigned long long near = row_solids(o, blockIdx.x * BX, j);
    if (i > g.m || j > g.n) return;
    const int ij = g.at(i, j);
    const double pavg = *pavg_p;
    const double p00 = z[j * zp + i] - pavg;
    f.p[ij] = p00;
    if (i == g.m || j == g.n) return;
    const double p10 = z[j * zp + i + 1] - pavg, p01 = z[(j + 1) * zp + i] - pavg,
                 p11 = z[(j + 1) * zp + i + 1] - pavg;
    const double rhoinv = 1. / f.rho[ij], u_star = f.c0[ij], v_star = f.c1[ij];
    f.u[ij] = u_star - 0.5 * dt * rhoinv * g.xsp * (p11 + p10 - p01 - p00);
    f.v[ij] = v_star - 0.5 * dt * rhoinv * g.ysp * (p11 - p10 + p01 - p00);
    // (outside a solid's window X = nX already: neither changes there)
    for (unsigned long long mk = near; mk; mk &= mk - 1) {
        const int k = low_bit(mk);
        if (!in_win(o.win[k], i, j)) continue;
        const ObjArrays &a = o.oa[k];
        a.X[ij] = a.nX[ij];
        a.Y[ij] = a.nY[ij];
    }
}

// ---- fluid_2d::set_boundaries (fluid_2d.cc:689-726), no-slip walls or
// periodic copies, and sim_horiz_flow::boundary (sim_type.cc:51-55): the
// left ghost columns of every row, ghost rows included, carry the inflow
// (U, 0); the right ghost columns extrapolate linearly (field::extrap,
// fields.hh:45-48).
// Phase 0 sets the left/right ghost columns of the interior rows, phase 1
// the bottom/top ghost rows over the full ghosted width (which reads the
// ghost columns phase 0 wrote, as in the reference).
__global__ void k_bc_cols(Grid g, Fields f) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= g.n) return;
    double *u = f.u, *v = f.v;
    const int r = g.at(0, j), m = g.m;
    if (g.px) {  // field::prd_bc (fields.hh:39-41)
        u[r - 2] = u[r + m - 2]; v[r - 2] = v[r + m - 2];
        u[r - 1] = u[r + m - 1]; v[r - 1] = v[r + m - 1];
        u[r + m] = u[r]; v[r + m] = v[r];
        u[r + m + 1] = u[r + 1]; v[r + m + 1] = v[r + 1];
        return;
    }
    if (g.zpr) {
        u[r - 2] = g.U; v[r - 2] = 0;
        u[r - 1] = g.U; v[r - 1] = 0;
        u[r + m] = 2 * u[r + m - 1] - u[r + m - 2]; v[r + m] = 2 * v[r + m - 1] - v[r + m - 2];
        u[r + m + 1] = 2 * u[r + m] - u[r + m - 1]; v[r + m + 1] = 2 * v[r + m] - v[r + m - 1];
        return;
    }
    u[r - 2] = -u[r + 1]; v[r - 2] = -v[r + 1];
    u[r - 1] = -u[r]; v[r - 1] = -v[r];
    u[r + m] = -u[r + m - 1]; v[r + m] = -v[r + m - 1];
    u[r + m + 1] = -u[r + m - 2]; v[r + m + 1] = -v[r + m - 2];
}
__global__ void k_bc_rows(Grid g, Fields f) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x - 2;
    if (i >= g.m + 2) return;
    double *u = f.u, *v = f.v;
    const int P = g.P, b = g.at(i, 0), t = g.at(i, g.n - 1);
    if (g.py) {
        u[b - 2 * P] = u[t - P]; v[b - 2 * P] = v[t - P];
        u[b - P] = u[t]; v[b - P] = v[t];
        u[t + P] = u[b]; v[t + P] = v[b];
        u[t + 2 * P] = u[b + P]; v[t + 2 * P] = v[b + P];
        return;
    }
    if (g.zpr && i < 0) {  // the inflow columns of the ghost rows
        u[b - 2 * P] = u[b - P] = u[t + P] = u[t + 2 * P] = g.U;
        v[b - 2 * P] = v[b - P] = v[t + P] = v[t + 2 * P] = 0;
        return;
    }
    u[b - 2 * P] = -u[b + P]; v[b - 2 * P] = -v[b + P];
    u[b - P] = -u[b]; v[b - P] = -v[b];
    u[t + P] = -u[t]; v[t + P] = -v[t];
    u[t + 2 * P] = -u[t - P]; v[t + 2 * P] = -v[t - P];
}

// ---- fluid_2d::set_boundaries_solid (fluid_2d.cc:818-835): each solid's
// reference map in the ghost layers is its periodic image shifted by the
// box length (s_field::prd_bc). Columns first, then the rows over the full
// ghosted width, as for the velocities.
__global__ void k_bc_solid_cols(Grid g, Solids o) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= g.n) return;
    const int r = g.at(0, j), m = g.m;
    const double lx = g.bx - g.ax;
    for (int k = 0; k < o.n; k++) {
        double *X = o.oa[k].X, *Y = o.oa[k].Y;
        X[r - 2] = X[r + m - 2] + -lx; Y[r - 2] = Y[r + m - 2] + 0;
        X[r - 1] = X[r + m - 1] + -lx; Y[r - 1] = Y[r + m - 1] + 0;
        X[r + m] = X[r] + lx; Y[r + m] = Y[r] + 0;
        X[r + m + 1] = X[r + 1] + lx; Y[r + m + 1] = Y[r + 1] + 0;
    }
}
__global__ void k_bc_solid_rows(Grid g, Solids o) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x - 2;
    if (i >= g.m + 2) return;
    const int P = g.P, b = g.at(i, 0), t = g.at(i, g.n - 1);
    const double ly = g.by - g.ay;
    for (int k = 0; k < o.n; k++) {
        double *X = o.oa[k].X, *Y = o.oa[k].Y;
        X[b - 2 * P] = X[t - P] + 0; Y[b - 2 * P] = Y[t - P] + -ly;
        X[b - P] = X[t] + 0; Y[b - P] = Y[t] + -ly;
        X[t + P] = X[b] + 0; Y[t + P] = Y[b] + ly;
        X[t + 2 * P] = X[b + P] + 0; Y[t + 2 * P] = Y[b + P] + ly;
    }
}

// ---- initial velocity from the solids' momenta (fluid_2d.cc:210-238)
__global__ void k_init_velocity(Grid g, Fields f, Solids o, double rhof) {
    const int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y;
    if (i > g.m || j > g.n) return;
    const int ij = g.at(i, j);
    f.p[ij] = 0;
    if (i == g.m || j == g.n) return;
    double mu = 0, mv = 0, rho = 0, sfrac = 0;
    for (int k = 0; k < o.n; k++) {
        const ObjConst &c = o.oc[k];
        if (!c.set_velocity) continue;  // obj_field::mom_contrib (obj_field.cc:62-75)
        double tf = trans_func(c, o.oa[k].phi[ij]), uu, vv;
        if (c.kind != kSpin) {  // obj_standard::velocity with zero speed
            uu = 0 - 0 * o.oa[k].Y[ij];
            vv = 0 + 0 * o.oa[k].X[ij];
        } else {
            uu = vv = 0.;
        }
        sfrac += tf;
        rho += tf * c.rhos;
        mu += tf * c.rhos * uu;
        mv += tf * c.rhos * vv;
    }
    if (sfrac >= 1) {
        f.u[ij] = mu / rho;
        f.v[ij] = mv / rho;
    } else {
        double uu, vv, irho = 1 / ((1 - sfrac) * rhof + rho);
        sim_velocity(g.vel, g.ax + g.dx * (0.5 + i), g.ay + g.dy * (0.5 + j), uu, vv, g.U);
        f.u[ij] = ((1 - sfrac) * rhof * uu + mu) * irho;
        f.v[ij] = ((1 - sfrac) * rhof * vv + mv) * irho;
    }
}

#define LAUNCH_CHECK() HIP_CHECK(hipGetLastError())

}  // namespace

int launch_normal_derivs(hipStream_t s, const Grid &g, const Fields &f, Solids o, double dt) {
    k_normal_derivs<<<grid2(g.m, g.n), dim3(BX, BY), 0, s>>>(g, f, o, dt);
    LAUNCH_CHECK();
    if (o.n) {
        k_refmap_slopes<<<dim3(win_blocks(g, o.n), 1, o.n), dim3(BX, BY), 0, s>>>(g, o);
        LAUNCH_CHECK();
    }
    return 0;
}
int launch_godunov(hipStream_t s, const Grid &g, const Fields &f, Solids o, bool tang) {
    if (!tang && o.n) {
        k_refmap_upwind<<<dim3(win_blocks(g, o.n), 1, o.n), dim3(BX, BY), 0, s>>>(g, f, o);
        LAUNCH_CHECK();
    }
    if (tang) k_godunov<true><<<grid2(g.m, g.n), dim3(BX, BY), 0, s>>>(g, f, o);
    else k_godunov<false><<<grid2(g.m, g.n), dim3(BX, BY), 0, s>>>(g, f, o);
    LAUNCH_CHECK();
    return 0;
}
int launch_tang_half(hipStream_t s, const Grid &g, const Fields &f, Solids o, const Phys &ph, double dt) {
    k_tang_half<<<grid2(g.m, g.n), dim3(BX, BY), 0, s>>>(g, f, o, ph, dt);
    LAUNCH_CHECK();
    if (o.n) {
        k_refmap_half<<<dim3(win_blocks(g, o.n), 1, o.n), dim3(BX, BY), 0, s>>>(g, f, o, dt);
        LAUNCH_CHECK();
    }
    return 0;
}

// ---- a solid's sweep window from its band's box (see SolidWin)
// A box widened past one edge (or two adjacent ones) still holds the whole
// solid when the solid's interior is the side of the interface the box
// encloses: outside the box there is no interface, and with no two opposite
// edges crossed the rest of the grid is connected, so one status sampled
// just past a side that crosses no edge tells (7, far outside: the window
// is the box clipped to the ghosted grid; 0, inside: the whole grid).
__device__ void solid_window(const Grid &g, const int *bbox, const int *st, SolidWin *win, bool full) {
    const SolidWin all{-2, g.m + 2, -2, g.n + 2};
    if (full || g.px || g.py || bbox[1] < 0) {
        *win = all;
        return;
    }
    // level-set coordinates are cell coordinates + 2
    const int i0 = bbox[0] - 2 - kWinMargin, i1 = bbox[1] - 2 + 1 + kWinMargin;
    const int j0 = bbox[2] - 2 - kWinMargin, j1 = bbox[3] - 2 + 1 + kWinMargin;
    const bool l = i0 < 0, r = i1 > g.m, d = j0 < 0, u = j1 > g.n;
    if (!l && !r && !d && !u) {
        *win = SolidWin{i0, i1, j0, j1};
        return;
    }
    if ((l && r) || (d && u)) {
        *win = all;
        return;
    }
    // a level-set point just past the box's right or left side (one of the
    // two crosses no edge)
    const int I = !r ? bbox[1] + 1 : bbox[0] - 1, J = bbox[2];
    if (st[g.ls(I, J)] != 7) {
        *win = all;
        return;
    }
    *win = SolidWin{l ? -2 : i0, r ? g.m + 2 : i1, d ? -2 : j0, u ? g.n + 2 : j1};
}
__global__ void k_solid_window(Grid g, const int *bbox, const int *st, SolidWin *win, int full) {
    if (threadIdx.x == 0) solid_window(g, bbox, st, win, full != 0);
}
// every solid's window in one launch (workgroup k: solid k)
__global__ void k_solid_windows(Grid g, const int *const *bbox, const ObjArrays *oa, SolidWin *win,
                                unsigned long long full) {
    const int k = blockIdx.x;
    if (threadIdx.x == 0) solid_window(g, bbox[k], oa[k].st, win + k, (full >> k) & 1);
}
int launch_solid_window(hipStream_t s, const Grid &g, const int *bbox, const int *st, SolidWin *win, bool full) {
    k_solid_window<<<1, 64, 0, s>>>(g, bbox, st, win, full);
    LAUNCH_CHECK();
    return 0;
}
int launch_solid_windows(hipStream_t s, const Grid &g, const int *const *bbox, const ObjArrays *oa, SolidWin *win,
                         int n, unsigned long long full) {
    if (n <= 0) return 0;
    k_solid_windows<<<n, 64, 0, s>>>(g, bbox, oa, win, full);
    LAUNCH_CHECK();
    return 0;
}
int launch_mac_src(hipStream_t s, const Grid &g, const Fields &f, double *b, int bp) {
    k_mac_src<<<grid2(g.m, g.n), dim3(BX, BY), 0, s>>>(g, f, b, bp);
    LAUNCH_CHECK();
    return 0;
}
int launch_mac_correct_predict(hipStream_t s, const Grid &g, const Fields &f, Solids o, const double *z,
                               int zp, const double *fw, const double *fs, int cp, bool const_rho, double dt) {
    k_mac_correct_predict<<<grid2(g.m, g.n), dim3(BX, BY), 0, s>>>(g, f, o, z, zp, fw, fs, cp, const_rho, dt);
    LAUNCH_CHECK();
    return 0;
}
int launch_half_timestep(hipStream_t s, const Grid &g, Solids o) {
    k_half_timestep<<<grid2(g.m, g.n), dim3(BX, BY), 0, s>>>(g, o);
    LAUNCH_CHECK();
    return 0;
}
int launch_set_rho(hipStream_t s, const Grid &g, const Fields &f, Solids o, double rhof) {
    k_set_rho<<<grid2(g.ml, g.nl), dim3(BX, BY), 0, s>>>(g, f, o, rhof);
    LAUNCH_CHECK();
    return 0;
}
int launch_stress(hipStream_t s, const Grid &g, const Fields &f, Solids o, const Phys &ph) {
    k_stress<<<grid2(g.m + 1, g.n + 1), dim3(BX, BY), 0, s>>>(g, f, o, ph);
    LAUNCH_CHECK();
    return 0;
}
int launch_ustar(hipStream_t s, const Grid &g, const Fields &f, Solids o, const Phys &ph, double dt) {
    k_ustar<<<grid2(g.m, g.n), dim3(BX, BY), 0, s>>>(g, f, o, ph, dt);
    LAUNCH_CHECK();
    return 0;
}
int launch_fem_src(hipStream_t s, const Grid &g, const Fields &f, double dt, double *b, int bp) {
    k_fem_src<<<grid2(g.m + 1, g.n + 1), dim3(BX, BY), 0, s>>>(g, f, dt, b, bp);
    LAUNCH_CHECK();
    return 0;
}
int launch_pressure_velocity(hipStream_t s, const Grid &g, const Fields &f, Solids o, const double *z, int zp,
                             const double *pavg, double dt) {
    k_pressure_velocity<<<grid2(g.m + 1, g.n + 1), dim3(BX, BY), 0, s>>>(g, f, o, z, zp, pavg, dt);
    LAUNCH_CHECK();
    return 0;
}
int launch_set_boundaries_solid(hipStream_t s, const Grid &g, Solids o) {
    if (!g.spbc || !o.n) return 0;
    k_bc_solid_cols<<<(g.n + 255) / 256, 256, 0, s>>>(g, o);
    LAUNCH_CHECK();
    k_bc_solid_rows<<<(g.m + 4 + 255) / 256, 256, 0, s>>>(g, o);
    LAUNCH_CHECK();
    return 0;
}
int launch_set_boundaries(hipStream_t s, const Grid &g, const Fields &f, Solids o) {
    if (int rc = launch_set_boundaries_solid(s, g, o)) return rc;
    k_bc_cols<<<(g.n + 255) / 256, 256, 0, s>>>(g, f);
    LAUNCH_CHECK();
    k_bc_rows<<<(g.m + 4 + 255) / 256, 256, 0, s>>>(g, f);
    LAUNCH_CHECK();
    return 0;
}
int launch_init_velocity(hipStream_t s, const Grid &g, const Fields &f, Solids o, double rhof) {
    k_init_velocity<<<grid2(g.m + 1, g.n + 1), dim3(BX, BY), 0, s>>>(g, f, o, rhof);
    LAUNCH_CHECK();
    return 0;
}

}  // namespace incrmt
