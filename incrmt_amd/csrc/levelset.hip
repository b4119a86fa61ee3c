// GPU narrow-band level set (see levelset.hh).
//
// Rebuild. The reference rebuilds the band serially: a sign-change scan, an
// exact closest-point distance for the two interface layers (Newton on a
// bicubic), then a heap-ordered second-order fast march out to the band
// limit (levelset.cc:199-335, heap.cc). Here the scan and the Newton solves
// are one thread per point, and the march is replaced by a Jacobi fixed-point
// iteration of the SAME local update (eikonal.hh) over the points near the
// interface: a neighbour counts as known when its current value is upwind of
// the point's own, which is exactly the set the heap has accepted when it
// accepts the point. Where the heap's order is monotone and free of exact
// ties the fixed point equals the heap's result bit for bit.
//
// Extension. The reference extends (nX,nY) outward in increasing-phi order,
// each point by a weighted least-squares fit over the neighbours with
// smaller phi (traverse.hh:169-260). The set a point reads depends on phi
// alone, so a Jacobi iteration over the outside band reaches the sequential
// result exactly once every dependency chain has been walked; it stops when
// an iteration changes no value.
#include <algorithm>
#include <cstdio>
#include "levelset.hh"
#include "eikonal.hh"

namespace incrmt {

namespace {

constexpr int BX = 64, BY = 4;

// Wavefront-aggregated list append: the lanes of a wavefront (among those
// executing the call) with pred set take consecutive slots of `list` with
// one atomic on `counter`. The order inside a list carries no meaning: each
// entry is processed on its own.
__device__ inline void wave_append(bool pred, int *counter, int *list, int value) {
    const unsigned long long m = __ballot(pred);
    if (!m) return;
    const int lane = __lane_id(), leader = __ffsll((long long)m) - 1;
    int base = 0;
    if (lane == leader) base = atomicAdd(counter, __popcll(m));
    base = __shfl(base, leader, 64);
    if (pred) list[base + __popcll(m & ((1ull << lane) - 1))] = value;
}
// Wavefront-aggregated flag set: one atomicOr per distinct tile index.
__device__ inline void wave_flag(bool pred, int *flags, int tile, int bit) {
    unsigned long long m = __ballot(pred);
    const int lane = __lane_id();
    while (m) {
        const int l = __ffsll((long long)m) - 1;
        const int t = __shfl(tile, l, 64);
        const unsigned long long same = __ballot(pred && tile == t);
        if (lane == l) atomicOr(&flags[t], bit);
        m &= ~same;
    }
}

inline dim3 grid_win(const Win &v) { return dim3((v.I1 - v.I0 + BX - 1) / BX, (v.J1 - v.J0 + BY - 1) / BY); }
constexpr int kListBlocks = 1024, kListThreads = 256;
constexpr int kTileBlocks = 512;

inline dim3 grid2(int nx, int ny) { return dim3((nx + BX - 1) / BX, (ny + BY - 1) / BY); }

__device__ inline void ls_coords(const Grid &g, int ij, int &I, int &J) {
    const int r = ij - (kOrigin - 2);
    J = r / g.P;
    I = r - J * g.P;
}

// ---- bicubic with gradient on the ghosted grid (levelset.cc:63-79,
// built/bicubic.cc); evaluation order as in the reference.
struct xrow { double v, d; };
__device__ inline xrow herm_mid(double pl, double f0, double f1, double pr, double e, double me) {
    xrow r;
    r.v = pl * e * me * me + f0 * me * me * (1 + 2 * e) + f1 * e * e * (3 - 2 * e) - pr * e * e * me;
    r.d = pl * me * (1 - 3 * e) - f0 * 6 * e * me + f1 * 6 * e * me + pr * e * (3 * e - 2);
    return r;
}
__device__ inline xrow herm_left(double f0, double f1, double pr, double e, double me) {
    xrow r;
    r.v = f0 * me * me + f1 * e * (2 - e) - pr * e * me;
    r.d = -f0 * 2 * me + f1 * 2 * me + pr * (2 * e - 1);
    return r;
}
__device__ inline xrow herm_right(double pl, double f0, double f1, double e, double me) {
    xrow r;
    r.v = pl * e * me + f0 * (1 - e * e) + f1 * e * e;
    r.d = pl * (1 - 2 * e) - f0 * 2 * e + f1 * 2 * e;
    return r;
}

struct LsGeom {
    int m, n, P;
    double ax, ay, dx, dy, xsp, ysp;
};

__device__ double bicubic(const LsGeom &g, const double *base, double x, double y, double &phix, double &phiy) {
    const int m = g.m, n = g.n, P = g.P;
    const double dxf = 0.5, dyf = 0.5, d2f = 0.25;
    double e = (x - g.ax) * g.xsp, f = (y - g.ay) * g.ysp;
    int i = int(e), j = int(f);
    int xc, yc, ic, jc;
    if (i < 1) { xc = 0; ic = 0; }
    else if (i >= m - 2) { xc = 2; ic = m - 2; e -= double(m - 2); }
    else { xc = 1; ic = i; e -= double(i); }
    if (j < 1) { yc = 0; jc = 0; }
    else if (j >= n - 2) { yc = 2; jc = n - 2; f -= double(n - 2); }
    else { yc = 1; jc = j; f -= double(j); }
    const double me = 1 - e, mf = 1 - f;
    const double *p = base + (jc * P + ic);
    auto row = [&](double l, double v0, double v1, double r) -> xrow {
        return xc == 1 ? herm_mid(l, v0, v1, r, e, me) : xc == 0 ? herm_left(v0, v1, r, e, me) : herm_right(l, v0, v1, e, me);
    };
    double plbx = 0, pltx = 0, prbx = 0, prtx = 0;
    if (xc != 0) { plbx = (p[1] - p[-1]) * dxf; pltx = (p[P + 1] - p[P - 1]) * dxf; }
    if (xc != 2) { prbx = (p[2] - p[0]) * dxf; prtx = (p[P + 2] - p[P]) * dxf; }
    xrow r1 = row(plbx, p[0], p[1], prbx);
    xrow r2 = row(pltx, p[P], p[P + 1], prtx);
    xrow r0 = {0, 0}, r3 = {0, 0};
    if (yc != 0) {
        double plby = (p[P] - p[-P]) * dyf, prby = (p[P + 1] - p[-P + 1]) * dyf, plbxy = 0, prbxy = 0;
        if (xc != 0) plbxy = (p[P + 1] - p[-P + 1] - p[P - 1] + p[-P - 1]) * d2f;
        if (xc != 2) prbxy = (p[P + 2] - p[-P + 2] - p[P] + p[-P]) * d2f;
        r0 = row(plbxy, plby, prby, prbxy);
    }
    if (yc != 2) {
        double plty = (p[2 * P] - p[0]) * dyf, prty = (p[2 * P + 1] - p[1]) * dyf, pltxy = 0, prtxy = 0;
        if (xc != 0) pltxy = (p[2 * P + 1] - p[1] - p[2 * P - 1] + p[-1]) * d2f;
        if (xc != 2) prtxy = (p[2 * P + 2] - p[2] - p[2 * P] + p[0]) * d2f;
        r3 = row(pltxy, plty, prty, prtxy);
    }
    if (yc == 1) {
        phix = g.xsp * (f * mf * mf * r0.d + mf * mf * (1 + 2 * f) * r1.d + f * f * (3 - 2 * f) * r2.d - f * f * mf * r3.d);
        phiy = g.ysp * (mf * (1 - 3 * f) * r0.v - 6 * f * mf * r1.v + 6 * f * mf * r2.v + f * (3 * f - 2) * r3.v);
        return f * mf * mf * r0.v + mf * mf * (1 + 2 * f) * r1.v + f * f * (3 - 2 * f) * r2.v - f * f * mf * r3.v;
    } else if (yc == 0) {
        phix = g.xsp * (mf * mf * r1.d + f * (2 - f) * r2.d - f * mf * r3.d);
        phiy = g.ysp * (-2 * mf * r1.v + 2 * mf * r2.v + (2 * f - 1) * r3.v);
        return mf * mf * r1.v + f * (2 - f) * r2.v - f * mf * r3.v;
    }
    phix = g.xsp * (f * mf * r0.d + (1 - f * f) * r1.d + f * f * r2.d);
    phiy = g.ysp * ((1 - 2 * f) * r0.v - 2 * f * r1.v + 2 * f * r2.v);
    return f * mf * r0.v + (1 - f * f) * r1.v + f * f * r2.v;
}

// The compiled reference truncates before abs (levelset.cc:124,126).
__device__ inline double int_abs(double v) { return double(abs(int(v))); }

// levelset::newton_failsafe (levelset.cc:137-195)
__device__ bool newton_failsafe(const LsGeom &g, const double *phi, const int *st, int I, int J, int ij, double x0,
                                double y0, double &x, double &y) {
    const int ts = 7 - st[ij], P = g.P;
    const double cphi = phi[ij];
    bool hf = false, vf = false;
    double fx = 0, fy = 0, t;
    if (I > 0 && st[ij - 1] == ts) {
        fx = cphi / (phi[ij - 1] - cphi);
        if (I < g.m - 1 && st[ij + 1] == ts) {
            t = cphi / (cphi - phi[ij + 1]);
            if (t < -fx) fx = t;
        }
        hf = true;
    } else if (I < g.m - 1 && st[ij + 1] == ts) {
        fx = cphi / (cphi - phi[ij + 1]);
        hf = true;
    }
    if (J > 0 && st[ij - P] == ts) {
        fy = cphi / (phi[ij - P] - cphi);
        if (J < g.n - 1 && st[ij + P] == ts) {
            t = cphi / (cphi - phi[ij + P]);
            if (t < -fy) fy = t;
        }
        vf = true;
    } else if (J < g.n - 1 && st[ij + P] == ts) {
        fy = cphi / (cphi - phi[ij + P]);
        vf = true;
    }
    if (hf && vf) {
        double yy = fy * g.dy, tx = fx * g.dx;
        double cc = tx * yy / (tx * tx + yy * yy);
        x = yy * cc + x0;
        y = tx * cc + y0;
        return true;
    }
    if (hf) { x = fx * g.dx + x0; y = y0; return true; }
    if (vf) { x = x0; y = fy * g.dy + y0; return true; }
    return false;
}

// levelset::newton (levelset.cc:107-127). b0 and st point at ghosted point
// (0,0) and ij = J*P + I indexes both; y0 is formed from the reference's own
// linear index J*m as (J*m)*(dy/m).
__device__ bool newton(const LsGeom &g, const double *b0, const int *st, int I, int J, int ij, double &x0,
                       double &y0, double &x, double &y) {
    const double dyy = g.dy / double(g.m), step_limit = sqrt(g.dx * g.dx + g.dy * g.dy);
    int iter = 0, cutoff = 0;
    double ph, px, py;
    x0 = x = g.ax + I * g.dx;
    y0 = y = g.ay + double(J * g.m) * dyy;
    ph = bicubic(g, b0, x, y, px, py);
    do {
        double delx = x0 - x, dely = y0 - y;
        double m0 = delx * px + dely * py + ph;
        double m1 = px * px + py * py;
        iter++;
        m0 /= m1;
        if (iter > 100 || !isfinite(m0) || m0 > step_limit)
            return newton_failsafe(g, b0, st, I, J, ij, x0, y0, x, y);
        x += delx - m0 * px;
        y += dely - m0 * py;
        ph = bicubic(g, b0, x, y, px, py);
        if (int_abs(ph) < 1e-10) cutoff++; else cutoff = 0;
    } while (cutoff < 8);
    if (int_abs(x0 - x) < 3 * g.dx && int_abs(y0 - y) < 3 * g.dy) return true;
    return newton_failsafe(g, b0, st, I, J, ij, x0, y0, x, y);
}

LsGeom geom_of(const Grid &g) {
    LsGeom l;
    l.m = g.ml; l.n = g.nl; l.P = g.P;
    l.ax = g.lax; l.ay = g.lay;
    l.dx = g.dx; l.dy = g.dy; l.xsp = g.xsp; l.ysp = g.ysp;
    return l;
}

// ------------------------------------------------------------------ kernels
__global__ void k_zero_counts(int *cnt, int n) {
    for (int k = threadIdx.x; k < n; k += blockDim.x) cnt[k] = 0;
}

// obj_field::pin (obj_field.cc:92-97) over the ghosted grid
// (ix, iy), when given, stand in for (X, Y) on the interior cells: the
// reference map update_velocity will copy there (X = nX, Y = nY), so a
// rebuild can pin to the end-of-step map before that copy has run.
__device__ __forceinline__ void pin_body(const Grid &g, const ObjConst &c, const ObjArrays &a, const Win &wn,
                                         const double *ix, const double *iy) {
    const int I = wn.I0 + blockIdx.x * BX + threadIdx.x, J = wn.J0 + blockIdx.y * BY + threadIdx.y;
    if (I >= wn.I1 || J >= wn.J1) return;
    const int ij = g.ls(I, J);
    const bool inner = ix && I >= 2 && I < g.m + 2 && J >= 2 && J < g.n + 2;
    const double p = a.phi[ij];
    if (p < c.pinw && p > -c.pinw) a.phi[ij] = obj_phi(c, inner ? ix[ij] : a.X[ij], inner ? iy[ij] : a.Y[ij]);
}
__global__ void k_pin(Grid g, ObjConst c, ObjArrays a, Win wn, const double *ix, const double *iy) {
    pin_body(g, c, a, wn, ix, iy);
}

// Sign-change scan (levelset.cc:211-263): a point is on an interface layer
// when a 4-neighbour has the other sign. Interface points are listed for
// the Newton solves; the fixed-point buffers start at +-inf elsewhere.
__device__ __forceinline__ void scan_body(const Grid &g, const ObjArrays &a, const LsWork &w, const Win &wn) {
    const int I = wn.I0 + blockIdx.x * BX + threadIdx.x, J = wn.J0 + blockIdx.y * BY + threadIdx.y;
    if (I >= wn.I1 || J >= wn.J1) return;
    const int ij = g.ls(I, J), P = g.P;
    const double *phi = a.phi;
    // the four neighbours read together (a neighbour off the grid reads the
    // point itself, which never differs in sign), then tested
    const double p0 = phi[ij], pl = phi[I > 0 ? ij - 1 : ij], pr = phi[I < g.ml - 1 ? ij + 1 : ij],
                 pd = phi[J > 0 ? ij - P : ij], pu = phi[J < g.nl - 1 ? ij + P : ij];
    const bool pos = p0 > 0;
    const bool iface = ((pl > 0) != pos) | ((pr > 0) != pos) | ((pd > 0) != pos) | ((pu > 0) != pos);
    a.st[ij] = pos ? (iface ? 4 : 7) : (iface ? 3 : 0);
    wave_append(iface, &w.cnt[kCntIface], w.iface, ij);
    if (!iface) {
        const double inf = pos ? INFINITY : -INFINITY;
        w.va[ij] = inf;
        w.vb[ij] = inf;
    }
}
__global__ void k_scan(Grid g, ObjArrays a, LsWork w, Win wn) { scan_body(g, a, w, wn); }

// Exact closest-point distances of both interface layers, from the pinned
// phi (levelset.cc:266-291). The results go to the fixed-point buffers; phi
// itself keeps the pinned values until the band is finalised, since every
// Newton solve reads the pinned field.
__device__ __forceinline__ void newton_body(const Grid &g, const LsGeom &lg, const ObjArrays &a, const LsWork &w) {
    const int n = w.cnt[kCntIface];
    const double *b0 = a.phi + g.ls(0, 0);
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        const int ij = w.iface[k];
        int I, J;
        ls_coords(g, ij, I, J);
        double x0, y0, x, y;
        if (!newton(lg, b0, a.st + g.ls(0, 0), I, J, ij - g.ls(0, 0), x0, y0, x, y)) {
            atomicOr(&w.cnt[kCntErr], kErrNewton);
            continue;
        }
        x -= x0; y -= y0;
        const double d = sqrt(x * x + y * y);
        const double v = a.st[ij] == 4 ? d : -d;
        w.va[ij] = v;
        w.vb[ij] = v;
    }
}
__global__ void k_newton(Grid g, LsGeom lg, ObjArrays a, LsWork w) { newton_body(g, lg, a, w); }

// Candidate points: non-interface points within R cells (max norm) of an
// interface point, R covering the band cut-offs; separable dilation, rows
// then columns.
__device__ __forceinline__ void dilate_rows_body(const Grid &g, const ObjArrays &a, const LsWork &w, int R,
                                                 const Win &wn) {
    const int I = wn.I0 + blockIdx.x * BX + threadIdx.x, J = wn.J0 + blockIdx.y * BY + threadIdx.y;
    if (I >= wn.I1 || J >= wn.J1) return;
    const int ij = g.ls(I, J);
    // every read of the range issued without waiting on the one before (the
    // OR over it is the same with or without stopping at the first hit)
    const int d0 = I - R < 0 ? -I : -R, d1 = I + R >= g.ml ? g.ml - 1 - I : R;
    bool hit = false;
#pragma unroll 4
    for (int d = d0; d <= d1; d++) {
        const int s = a.st[ij + d];
        hit |= s == 3 || s == 4;
    }
    w.mark[ij] = hit;
}
__global__ void k_dilate_rows(Grid g, ObjArrays a, LsWork w, int R, Win wn) { dilate_rows_body(g, a, w, R, wn); }
// The column pass: a thread takes kColSeg consecutive rows of one column and
// keeps the count of marked points in the column window [J - R, J + R]
// (clipped to the grid) as it slides down, one point in and one out per row,
// instead of reading the 2R + 1 marks of every point.
constexpr int kColSeg = 8;
static_assert(kTileY >= kColSeg, "a thread's rows span at most two tile rows");
inline dim3 grid_win_cols(const Win &v) {
    return dim3((v.I1 - v.I0 + BX - 1) / BX, (v.J1 - v.J0 + BY * kColSeg - 1) / (BY * kColSeg));
}
__device__ __forceinline__ void dilate_cols_body(const Grid &g, const ObjArrays &a, const LsWork &w, int R,
                                                 const Win &wn) {
    const int I = wn.I0 + blockIdx.x * BX + threadIdx.x;
    const int J0 = wn.J0 + (blockIdx.y * BY + threadIdx.y) * kColSeg;
    const bool col = I < wn.I1 && J0 < wn.J1;
    const unsigned char *mk = w.mark + (col ? g.ls(I, 0) : 0);
    int cnt = 0;
    if (col) {
        const int lo = J0 - R < 0 ? 0 : J0 - R, hi = J0 + R >= g.nl ? g.nl - 1 : J0 + R;
#pragma unroll 4
        for (int J = lo; J <= hi; J++) cnt += mk[J * g.P] != 0;
    }
    unsigned hm = 0;  // bit k: row J0 + k is a candidate
#pragma unroll
    for (int k = 0; k < kColSeg; k++) {
        const int J = J0 + k;
        const bool in = col && J < wn.J1;
        if (in && k > 0) {
            // the window moves from [J-1-R, J-1+R] to [J-R, J+R]
            if (J + R < g.nl) cnt += mk[(J + R) * g.P] != 0;
            if (J - 1 - R >= 0) cnt -= mk[(J - 1 - R) * g.P] != 0;
        }
        if (in) {
            const int ij = g.ls(I, J), s = a.st[ij];
            const bool hit = s != 3 && s != 4 && cnt > 0;
            w.cflag[ij] = hit;
            hm |= (unsigned)hit << k;
        }
    }
    // the candidates of the whole wavefront appended with one atomic (an
    // inclusive prefix count across the lanes), and the tile flags set once
    // per wavefront and tile row: a thread's rows span at most two tile rows
    const int lane = __lane_id(), nh = __popc(hm);
    int incl = nh;
    for (int o = 1; o < 64; o <<= 1) {
        const int v = __shfl_up(incl, o, 64);
        if (lane >= o) incl += v;
    }
    const int total = __shfl(incl, 63, 64);
    int base = 0;
    if (lane == 63 && total > 0) base = atomicAdd(&w.cnt[kCntCand], total);
    base = __shfl(base, 63, 64);
    int pos = base + incl - nh;
#pragma unroll
    for (int k = 0; k < kColSeg; k++)
        if ((hm >> k) & 1) w.cand[pos++] = g.ls(I, J0 + k);
    const int ta = J0 / kTileY, tb = (J0 + kColSeg - 1) / kTileY;
    unsigned ma = 0;  // rows of tile row ta
#pragma unroll
    for (int k = 0; k < kColSeg; k++)
        if ((J0 + k) / kTileY == ta) ma |= 1u << k;
    wave_flag((hm & ma) != 0, w.tflag, ta * w.ntx + I / kTileX, 1);
    if (tb != ta) wave_flag((hm & ~ma) != 0, w.tflag, tb * w.ntx + I / kTileX, 1);
}
__global__ void k_dilate_cols(Grid g, ObjArrays a, LsWork w, int R, Win wn) { dilate_cols_body(g, a, w, R, wn); }

__global__ void k_bbox_init(LsWork w) {
    if (threadIdx.x == 0) {
        w.bbox[0] = w.bbox[2] = 1 << 30;
        w.bbox[1] = w.bbox[3] = -1;
    }
}

// Collect the tiles whose flag has `bit` set into a list, clearing the bit.
__device__ __forceinline__ void compact_tiles_body(const LsWork &w, int bit, int *list, int slot) {
    const int n = w.ntx * w.nty;
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) {
        const int f = w.tflag[t];
        if (f & bit) {
            const int k = atomicAdd(&w.cnt[slot], 1);
            list[k] = t;
            w.tflag[t] = f & ~bit;
        }
    }
}
__global__ void k_compact_tiles(LsWork w, int bit, int *list, int slot) { compact_tiles_body(w, bit, list, slot); }

// Known-neighbour test of the heap march (heap.hh:59-83) at the fixed point:
// the other side and the interface layer are always known; a same-side
// point is known when it has been reached, lies within the band, and is
// upwind of the point being updated.
struct FimAcc {
    const double *v;
    const int *st;
    double pval, limit;  // limit: the band cut-off of this side
    bool pos;
    __device__ bool in(int q) const {
        const int s = st[q];
        if (pos) return s <= 4 || (v[q] < pval && v[q] < limit);
        return s >= 3 || (v[q] > pval && v[q] > limit);
    }
    __device__ double val(int q) const { return v[q]; }
};

// One pass of the band march over the active tiles. A workgroup loads its
// tile and a 2-point halo (the reach of the local update) into LDS and
// repeats the Jacobi update on the tile until it stops changing, the halo
// held at the previous pass's values; passes repeat until no tile changes.
// Any order of such relaxations reaches the same fixed point.
constexpr int kFH = 2, kFW = kTileX + 2 * kFH, kFHt = kTileY + 2 * kFH, kFN = kFW * kFHt;

// Known-set test on the LDS tile. val() hands the update values with the
// inward march's sign flipped, so both marches run the outward update:
// negating every input of the inward update negates its result exactly
// (IEEE rounding is symmetric), which keeps one code path per wavefront.
struct LdsAcc {
    const double *v;
    const int *st;
    double pval, limit, sg;
    bool pos;
    __device__ bool in(int q) const {
        const int s = st[q];
        if (s < 0) return false;
        if (pos) return s <= 4 || (v[q] < pval && v[q] < limit);
        return s >= 3 || (v[q] > pval && v[q] > limit);
    }
    __device__ double val(int q) const { return sg * v[q]; }
};

// The OR over a workgroup of one predicate per thread, with one barrier:
// the wavefronts post their votes to one of three LDS slots used in turn,
// and the slot two rounds ahead is cleared between this round's barrier and
// the next (its last readers have passed this barrier, its next writers
// have not passed the next one). The slots are zero before the first round,
// behind a barrier.
struct VoteOr {
    int *slot;
    int k;
    __device__ int operator()(int pred) {
        if (__any(pred) && (threadIdx.x & 63) == 0) slot[k] = 1;
        __syncthreads();
        const int r = slot[k];
        if (threadIdx.x == 0) slot[k == 0 ? 2 : k - 1] = 0;
        k = k == 2 ? 0 : k + 1;
        return r;
    }
};

// The same test on one point's radius-2 cross stencil held in registers
// (slots: the point, x -1 +1 -2 +2, y -1 +1 -2 +2). The update's reads,
// taken one by one from LDS, form a chain of dependent round trips; here
// the statuses are read once per tile and the values once per iteration,
// all issued together. The offsets the update asks for are compile-time
// constants after inlining, so the slot lookups fold away.
struct RegAcc {
    double v[9];
    int st[9];
    int base;
    double pval, limit, sg;
    bool pos;
    __device__ static int slot(int d) {
        return d == -1 ? 1 : d == 1 ? 2 : d == -2 ? 3 : d == 2 ? 4 : d == -kFW ? 5 : d == kFW ? 6
             : d == -2 * kFW ? 7 : d == 2 * kFW ? 8 : 0;
    }
    __device__ bool in(int q) const {
        const int k = slot(q - base), s = st[k];
        if (s < 0) return false;
        if (pos) return s <= 4 || (v[k] < pval && v[k] < limit);
        return s >= 3 || (v[k] > pval && v[k] > limit);
    }
    __device__ double val(int q) const { return sg * v[slot(q - base)]; }
};
__device__ constexpr int kStencil[9] = {0, -1, 1, -2, 2, -kFW, kFW, -2 * kFW, 2 * kFW};

__device__ __forceinline__ void fim_tiles(const Grid &g, const EikGeom &eg, const ObjConst &c, const ObjArrays &a,
                                          const LsWork &w, const double *src, double *dst, int slot, int max_local) {
    __shared__ double sv[2][kFN];
    __shared__ int sst[kFN];
    __shared__ int vslot[3];
    if (slot > 0 && w.cnt[kCntIter + slot - 1] == 0) return;  // fixed point already reached
    const int nt = w.cnt[kCntTileFim];
    if (threadIdx.x < 3) vslot[threadIdx.x] = 0;
    VoteOr vote{vslot, 0};
    int changed_any = 0;
    constexpr int kRounds = (kFN + kTileThreads - 1) / kTileThreads;
    for (int t = blockIdx.x; t < nt; t += gridDim.x) {
        const int tile = w.tiles_fim[t], tx = tile % w.ntx, ty = tile / w.ntx;
        const int I0 = tx * kTileX - kFH, J0 = ty * kTileY - kFH;
        // the tile and its halo: every global read issued before any wait
        double rv[kRounds];
        int rs[kRounds];
#pragma unroll
        for (int u = 0; u < kRounds; u++) {
            const int e = threadIdx.x + u * kTileThreads, I = I0 + e % kFW, J = J0 + e / kFW;
            rv[u] = 0;
            rs[u] = -1;
            if (e < kFN && I >= 0 && I < g.ml && J >= 0 && J < g.nl) {
                const int q = g.ls(I, J);
                rv[u] = src[q];
                rs[u] = a.st[q];
            }
        }
        const int li = threadIdx.x % kTileX + kFH, lj = threadIdx.x / kTileX + kFH, e = lj * kFW + li;
        const int I = I0 + li, J = J0 + lj;
        const bool mine = I < g.ml && J < g.nl && w.cflag[g.ls(I, J)];
#pragma unroll
        for (int u = 0; u < kRounds; u++) {
            const int f = threadIdx.x + u * kTileThreads;
            if (f < kFN) {
                sv[0][f] = sv[1][f] = rv[u];
                sst[f] = rs[u];
            }
        }
        __syncthreads();
        const bool pos = mine && sst[e] == 7;
        const double limit = pos ? c.phi_up : c.phi_down;
        // every stencil point of a tile point lies in the tile's 2-point halo
        RegAcc acc;
        acc.base = e; acc.limit = limit; acc.sg = pos ? 1.0 : -1.0; acc.pos = pos;
#pragma unroll
        for (int k = 0; k < 9; k++) acc.st[k] = sst[e + kStencil[k]];
        int cur = 0;
        for (int it = 0; it < max_local; it++) {
            int changed = 0;
            if (mine) {
#pragma unroll
                for (int k = 0; k < 9; k++) acc.v[k] = sv[cur][e + kStencil[k]];
                acc.pval = acc.v[0];
                const bool any = (I > 0 && acc.in(e - 1)) || (I < g.ml - 1 && acc.in(e + 1)) ||
                                 (J > 0 && acc.in(e - kFW)) || (J < g.nl - 1 && acc.in(e + kFW));
                double r;
                if (!any) r = pos ? INFINITY : -INFINITY;
                else {
                    Eik<true, RegAcc> k{acc, eg, g.ml, g.nl, kFW};
                    r = k.calc_fast(I, J, e);
                    if (!pos) r = -r;
                }
                sv[cur ^ 1][e] = r;
                changed = __double_as_longlong(r) != __double_as_longlong(acc.v[0]);
            }
            const int ch = vote(changed);
            cur ^= 1;
            if (!ch) break;
        }
        if (mine) {
            const int q = g.ls(I, J);
            const double v = sv[cur][e];
            dst[q] = v;
            if (__double_as_longlong(v) != __double_as_longlong(src[q])) changed_any = 1;
        }
        __syncthreads();
    }
    if (__any(changed_any) && (threadIdx.x & 63) == 0) atomicAdd(&w.cnt[kCntIter + slot], 1);
}
__global__ __launch_bounds__(kTileThreads) void k_fim_tiles(Grid g, EikGeom eg, ObjConst c, ObjArrays a, LsWork w,
                                                   const double *src, double *dst, int slot, int max_local) {
    fim_tiles(g, eg, c, a, w, src, dst, slot, max_local);
}
// The march passes of several solids in one launch: workgroup (x, y) runs
// the pass of solid y with x as its block index, on its own buffers and
// counters (pass p reads buf[p & 1], writes the other), so each solid's
// passes are exactly those it runs alone.
struct FimItem {
    ObjConst c;
    ObjArrays a;
    LsWork w;
    double *buf[2];
};
__global__ __launch_bounds__(kTileThreads) void k_fim_tiles_batch(Grid g, EikGeom eg, const FimItem *it, int p,
                                                         int slot, int max_local) {
    const FimItem &x = it[blockIdx.y];
    fim_tiles(g, eg, x.c, x.a, x.w, x.buf[p & 1], x.buf[(p + 1) & 1], slot, max_local);
}

// Final band (levelset.cc:300-335, 408-413): accepted points get their
// fixed-point values and status 5/2, points the heap would have left in its
// queue status 6/1, everything outside the band the cut-off value. Outside
// band points are listed for the extension.
__device__ void finalize_point(Grid g, ObjConst c, ObjArrays a, LsWork w, const double *v, int I, int J,
                               int (*bb)[BX * BY], int lane);
__device__ __forceinline__ void finalize_body(const Grid &g, const ObjConst &c, const ObjArrays &a, const LsWork &w,
                                              const double *v, const Win &wn) {
    __shared__ int bb[4][BX * BY];
    const int t = threadIdx.y * BX + threadIdx.x;
    bb[0][t] = bb[2][t] = 1 << 30;
    bb[1][t] = bb[3][t] = -1;
    const int I = wn.I0 + blockIdx.x * BX + threadIdx.x, J = wn.J0 + blockIdx.y * BY + threadIdx.y;
    if (I < wn.I1 && J < wn.J1) finalize_point(g, c, a, w, v, I, J, bb, t);
    // bounding box of the band and its trial points, for the next window
    __syncthreads();
    for (int h = BX * BY / 2; h > 0; h >>= 1) {
        if (t < h) {
            bb[0][t] = min(bb[0][t], bb[0][t + h]);
            bb[1][t] = max(bb[1][t], bb[1][t + h]);
            bb[2][t] = min(bb[2][t], bb[2][t + h]);
            bb[3][t] = max(bb[3][t], bb[3][t + h]);
        }
        __syncthreads();
    }
    if (t == 0 && bb[1][0] >= 0) {
        atomicMin(&w.bbox[0], bb[0][0]);
        atomicMax(&w.bbox[1], bb[1][0]);
        atomicMin(&w.bbox[2], bb[2][0]);
        atomicMax(&w.bbox[3], bb[3][0]);
    }
}
__global__ void k_finalize(Grid g, ObjConst c, ObjArrays a, LsWork w, const double *v, Win wn) {
    finalize_body(g, c, a, w, v, wn);
}
__device__ void finalize_point(Grid g, ObjConst c, ObjArrays a, LsWork w, const double *v, int I, int J,
                               int (*bb)[BX * BY], int lane) {
    const int ij = g.ls(I, J), P = g.P, s = a.st[ij];
    const double lim = c.phi_up, dlim = c.phi_down;
    // a point outside the band tests its four neighbours: their statuses
    // and values are read together (a neighbour off the grid reads the point
    // itself and is masked out)
    const bool hl = I > 0, hr = I < g.ml - 1, hd = J > 0, hu = J < g.nl - 1;
    const int ql = hl ? ij - 1 : ij, qr = hr ? ij + 1 : ij, qd = hd ? ij - P : ij, qu = hu ? ij + P : ij;
    auto any_nb = [&](auto acc) {
        const int tl = a.st[ql], tr = a.st[qr], td = a.st[qd], tu = a.st[qu];
        const double vl = v[ql], vr = v[qr], vd = v[qd], vu = v[qu];
        return (hl && acc(tl, vl)) | (hr && acc(tr, vr)) | (hd && acc(td, vd)) | (hu && acc(tu, vu));
    };
    auto acc_pos = [&](int t, double x) { return t == 4 || (t == 7 && x < lim); };
    auto acc_neg = [&](int t, double x) { return t == 3 || (t == 0 && x > dlim); };
    const double v0 = v[ij];
    int ns = s;
    double phi;
    if (s == 3 || s == 4) {
        phi = v0;
    } else if (s == 7) {
        if (v0 < lim) { ns = 5; phi = v0; }
        else {
            ns = any_nb(acc_pos) ? 6 : 7;
            phi = lim;
        }
    } else {
        if (v0 > dlim) { ns = 2; phi = v0; }
        else {
            ns = any_nb(acc_neg) ? 1 : 0;
            phi = dlim;
        }
    }
    w.mark[ij] = (unsigned char)ns;  // statuses go to st after every thread read the old ones
    if (ns >= 1 && ns <= 6) {
        bb[0][lane] = I; bb[1][lane] = I; bb[2][lane] = J; bb[3][lane] = J;
    }
    a.phi[ij] = phi;
    const bool pb = ns == 4 || ns == 5;
    wave_append(pb, &w.cnt[kCntPband], w.pband, ij);
    wave_flag(pb, w.tflag, (J / kTileY) * w.ntx + I / kTileX, 2);
}
__device__ __forceinline__ void commit_status_body(const Grid &g, const ObjArrays &a, const LsWork &w, const Win &wn) {
    const int I = wn.I0 + blockIdx.x * BX + threadIdx.x, J = wn.J0 + blockIdx.y * BY + threadIdx.y;
    if (I >= wn.I1 || J >= wn.J1) return;
    const int ij = g.ls(I, J);
    a.st[ij] = w.mark[ij];
}
__global__ void k_commit_status(Grid g, ObjArrays a, LsWork w, Win wn) { commit_status_body(g, a, w, wn); }

// One pass of the extension (traverse.hh:169-231) over the active tiles:
// outside band points take a weighted least-squares linear fit of (nX,nY)
// over band neighbours with smaller phi, in a (2r+1)^2 box, r = 3 growing to
// 7 while the fit is singular. A workgroup holds its tile and a 7-point halo
// in LDS and repeats the fit until the tile stops changing; outside points
// read the previous values (sx,sy), inside points their predictor values.
// A point's fit reads only neighbours of strictly smaller phi, so the
// equations form an acyclic system with one solution; the tile therefore
// updates in place (each fit reads whatever its neighbours hold, some
// already updated in the same round), and a round in which nothing changes
// leaves every point satisfying its own equation: that solution, bit for
// bit, whatever order the wavefronts ran in.
constexpr int kEH = 7, kEW = kTileX + 2 * kEH, kEHt = kTileY + 2 * kEH, kEN = kEW * kEHt;

__device__ __forceinline__ void extend_tiles(const Grid &g, const ObjArrays &a, const LsWork &w, const double *srcx,
                                             const double *srcy, double *dstx, double *dsty, int slot, int max_local) {
    __shared__ double sphi[kEN], sx[kEN], sy[kEN];
    __shared__ int sst[kEN];
    __shared__ int vslot[3];
    if (slot > 0 && w.cnt[kCntIter + slot - 1] == 0) return;  // fixed point already reached
    const int nt = w.cnt[kCntTileExt], m = g.ml, nn = g.nl;
    if (threadIdx.x < 3) vslot[threadIdx.x] = 0;
    VoteOr vote{vslot, 0};
    int changed_any = 0;
    constexpr int kRounds = (kEN + kTileThreads - 1) / kTileThreads;
    for (int t = blockIdx.x; t < nt; t += gridDim.x) {
        const int tile = w.tiles_ext[t], tx = tile % w.ntx, ty = tile / w.ntx;
        const int I0 = tx * kTileX - kEH, J0 = ty * kTileY - kEH;
        // the tile and its halo: both candidate values of a point are read,
        // so no read waits on the status, and every read is issued before
        // any wait
        double rph[kRounds], rsx[kRounds], rnx[kRounds], rsy[kRounds], rny[kRounds];
        int rst[kRounds];
#pragma unroll
        for (int u = 0; u < kRounds; u++) {
            const int e = threadIdx.x + u * kTileThreads, I = I0 + e % kEW, J = J0 + e / kEW;
            rst[u] = -1;
            rph[u] = rsx[u] = rnx[u] = rsy[u] = rny[u] = 0;
            if (e < kEN && I >= 0 && I < m && J >= 0 && J < nn) {
                const int q = g.ls(I, J);
                rst[u] = a.st[q];
                rph[u] = a.phi[q];
                rsx[u] = srcx[q]; rnx[u] = a.nX[q];
                rsy[u] = srcy[q]; rny[u] = a.nY[q];
            }
        }
#pragma unroll
        for (int u = 0; u < kRounds; u++) {
            const int e = threadIdx.x + u * kTileThreads;
            if (e < kEN) {
                const int st = rst[u];
                sst[e] = st;
                sphi[e] = rph[u];
                const double vx = st >= 4 ? rsx[u] : rnx[u], vy = st >= 4 ? rsy[u] : rny[u];
                sx[e] = vx;
                sy[e] = vy;
            }
        }
        __syncthreads();
        const int li = threadIdx.x % kTileX + kEH, lj = threadIdx.x / kTileX + kEH, e = lj * kEW + li;
        const int i = I0 + li, j = J0 + lj;
        const bool mine = i < m && j < nn && (sst[e] == 4 || sst[e] == 5);
        const double pp = sphi[e];
        // The fit's weights and moments depend on phi and the status only,
        // which the extension does not change: for the usual radius 3 they
        // are formed once, the contributing neighbours kept as a 7x7 bit
        // mask, and every iteration only re-accumulates the value sums in
        // the same neighbour order (bit-identical to forming everything).
        unsigned long long mask = 0;
        double al = 0, be = 0, ga = 0;
        bool wide = false, singular = false;
        int rw = 3;  // the fit's radius
        if (mine) {
            // the radius-3 window, unrolled a row at a time like the
            // iterations below; window points outside the grid hold status
            // -1 in the tile, which the status test rejects, so the window
            // is the clipped one of traverse.hh:180-203
            double S = 0, Sx = 0, Sy = 0, Sxx = 0, Sxy = 0, Syy = 0;
#pragma unroll
            for (int dj = -3; dj <= 3; dj++) {
                int rs[7];
                double rp[7];
#pragma unroll
                for (int di = -3; di <= 3; di++) {
                    rs[di + 3] = sst[e + dj * kEW + di];
                    rp[di + 3] = sphi[e + dj * kEW + di];
                }
#pragma unroll
                for (int di = -3; di <= 3; di++) {
                    const int s = rs[di + 3];
                    if (s < 2 || s > 5) continue;
                    double wgt = rp[di + 3];
                    if (!(wgt < pp)) continue;
                    wgt = pp - wgt;
                    S += wgt; Sx += di * wgt; Sy += dj * wgt;
                    Sxx += di * di * wgt; Sxy += di * dj * wgt; Syy += dj * dj * wgt;
                    mask |= 1ull << ((dj + 3) * 7 + di + 3);
                }
            }
            double det = S * (Sxy * Sxy - Sxx * Syy) + Sx * Sx * Syy - 2 * Sx * Sxy * Sy + Sy * Sy * Sxx;
            if (fabs(det) < 1e-10) {
                // the radius grows (traverse.hh:179-207) while the fit is
                // singular, to 7 at most; like the weights, the radius it
                // stops at and its coefficients depend on phi and the
                // status only, so they are found once here
                wide = true;
                int rr = 4;
                do {
                    const int ilo = i <= rr ? -i : -rr, ihi = i + rr >= m ? m - 1 - i : rr,
                              jlo = j <= rr ? -j : -rr, jhi = j + rr >= nn ? nn - 1 - j : rr;
                    S = Sx = Sy = Sxx = Sxy = Syy = 0;
                    for (int dj = jlo; dj <= jhi; dj++)
                        for (int di = ilo; di <= ihi; di++) {
                            const int q = e + dj * kEW + di, s = sst[q];
                            if (s < 2 || s > 5) continue;
                            double wgt = sphi[q];
                            if (!(wgt < pp)) continue;
                            wgt = pp - wgt;
                            S += wgt; Sx += di * wgt; Sy += dj * wgt;
                            Sxx += di * di * wgt; Sxy += di * dj * wgt; Syy += dj * dj * wgt;
                        }
                    det = S * (Sxy * Sxy - Sxx * Syy) + Sx * Sx * Syy - 2 * Sx * Sxy * Sy + Sy * Sy * Sxx;
                    rr++;
                } while (fabs(det) < 1e-10 && rr <= 7);
                rw = rr - 1;
                if (fabs(det) < 1e-10) singular = true;
            }
            if (!singular) {
                det = 1 / det;
                al = (Sxy * Sxy - Sxx * Syy) * det;
                be = (Sx * Syy - Sxy * Sy) * det;
                ga = (Sxx * Sy - Sxy * Sx) * det;
            }
        }
        for (int it = 0; it < max_local; it++) {
            int changed = 0;
            if (mine) {
                const double *cx = sx, *cy = sy;
                double p0 = 0, p1 = 0, p2 = 0, p3 = 0, p4 = 0, p5 = 0;
                double a0 = al, b0 = be, g0 = ga;
                if (!wide) {
                    // the 7x7 window unrolled in mask-bit order, a row at a
                    // time: the row's 21 LDS reads are issued together
                    // (every window point lies in the tile's halo), then
                    // its neighbours are summed in order. Offsets and
                    // coordinates are constants, so no mask bit is decoded
                    // at run time.
#pragma unroll
                    for (int dj = -3; dj <= 3; dj++) {
                        double rp[7], ra[7], rb[7];
#pragma unroll
                        for (int di = -3; di <= 3; di++) {
                            const int q = e + dj * kEW + di;
                            rp[di + 3] = sphi[q];
                            ra[di + 3] = cx[q];
                            rb[di + 3] = cy[q];
                        }
#pragma unroll
                        for (int di = -3; di <= 3; di++)
                            if ((mask >> ((dj + 3) * 7 + di + 3)) & 1) {
                                const double wgt = pp - rp[di + 3], fa = ra[di + 3], fb = rb[di + 3], xx = di, yy = dj;
                                p0 += wgt * fa; p1 += wgt * xx * fa; p2 += wgt * yy * fa;
                                p3 += wgt * fb; p4 += wgt * xx * fb; p5 += wgt * yy * fb;
                            }
                    }
                } else if (!singular) {
                    // the value sums over the window of radius rw, in the
                    // order the moments were formed
                    const int ilo = i <= rw ? -i : -rw, ihi = i + rw >= m ? m - 1 - i : rw,
                              jlo = j <= rw ? -j : -rw, jhi = j + rw >= nn ? nn - 1 - j : rw;
                    for (int dj = jlo; dj <= jhi; dj++)
                        for (int di = ilo; di <= ihi; di++) {
                            const int q = e + dj * kEW + di, s = sst[q];
                            if (s < 2 || s > 5) continue;
                            double wgt = sphi[q];
                            if (!(wgt < pp)) continue;
                            wgt = pp - wgt;
                            const double fa = cx[q], fb = cy[q], xx = di, yy = dj;
                            p0 += wgt * fa; p1 += wgt * xx * fa; p2 += wgt * yy * fa;
                            p3 += wgt * fb; p4 += wgt * xx * fb; p5 += wgt * yy * fb;
                        }
                }
                if (singular) {
                    atomicOr(&w.cnt[kCntErr], kErrSingular);  // the value stays
                } else {
                    const double vx = p0 * a0 + p1 * b0 + p2 * g0, vy = p3 * a0 + p4 * b0 + p5 * g0;
                    changed = __double_as_longlong(vx) != __double_as_longlong(cx[e]) ||
                              __double_as_longlong(vy) != __double_as_longlong(cy[e]);
                    if (changed) {  // only this thread writes its point
                        sx[e] = vx;
                        sy[e] = vy;
                    }
                }
            }
            if (!vote(changed)) break;
        }
        if (mine) {
            const int q = g.ls(i, j);
            const double vx = sx[e], vy = sy[e];
            dstx[q] = vx;
            dsty[q] = vy;
            if (__double_as_longlong(vx) != __double_as_longlong(srcx[q]) ||
                __double_as_longlong(vy) != __double_as_longlong(srcy[q]))
                changed_any = 1;
        }
        __syncthreads();
    }
    if (__any(changed_any) && (threadIdx.x & 63) == 0) atomicAdd(&w.cnt[kCntIter + slot], 1);
}
__global__ __launch_bounds__(kTileThreads) void k_extend_tiles(Grid g, ObjArrays a, LsWork w, const double *srcx,
                                                      const double *srcy, double *dstx, double *dsty, int slot,
                                                      int max_local) {
    extend_tiles(g, a, w, srcx, srcy, dstx, dsty, slot, max_local);
}
// Extension passes of several solids in one launch (as k_fim_tiles_batch)
struct ExtItem {
    ObjArrays a;
    LsWork w;
    double *bx[2], *by[2];
};
__global__ __launch_bounds__(kTileThreads) void k_extend_tiles_batch(Grid g, const ExtItem *it, int p, int slot,
                                                            int max_local) {
    const ExtItem &x = it[blockIdx.y];
    extend_tiles(g, x.a, x.w, x.bx[p & 1], x.by[p & 1], x.bx[(p + 1) & 1], x.by[(p + 1) & 1], slot, max_local);
}
// Counters of a batch of solids: clear the pass slots of every solid, or
// gather its error flags and the first `batch` slots into out (batch + 1
// ints per solid)
template <class Item>
__global__ void k_batch_zero(const Item *it, int n) {
    for (int k = blockIdx.x; k < n; k += gridDim.x)
        for (int e = threadIdx.x; e < kIterSlots; e += blockDim.x) it[k].w.cnt[kCntIter + e] = 0;
}
template <class Item>
__global__ void k_batch_gather(const Item *it, int n, int batch, int *out) {
    for (int k = blockIdx.x; k < n; k += gridDim.x)
        for (int e = threadIdx.x; e <= batch; e += blockDim.x)
            out[k * (batch + 1) + e] = e == 0 ? it[k].w.cnt[kCntErr] : it[k].w.cnt[kCntIter + e - 1];
}

// Seed both extension buffers with the current nX,nY of the outside band,
// or copy the converged buffer back into nX,nY.
__device__ __forceinline__ void ext_copy_body(const ObjArrays &a, const LsWork &w, const double *fx, const double *fy,
                                              int to_obj) {
    const int n = w.cnt[kCntPband];
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        const int ij = w.pband[k];
        if (to_obj) { a.nX[ij] = fx[ij]; a.nY[ij] = fy[ij]; }
        else { w.ex[ij] = w.fx[ij] = a.nX[ij]; w.ey[ij] = w.fy[ij] = a.nY[ij]; }
    }
}
__global__ void k_ext_copy(ObjArrays a, LsWork w, const double *fx, const double *fy, int to_obj) {
    ext_copy_body(a, w, fx, fy, to_obj);
}

// ---- the set-up and finish of several solids' rebuilds (and the copies
// around their extensions) in shared launches: workgroup (x, y, z) does
// what workgroup (x, y) of the single-solid launch does for solid z, whose
// windows and parameters are item z (grids sized for the largest window;
// workgroups past a solid's window do nothing). The solids' arrays and
// counters are their own, so each gets exactly what its own launches give.
struct SetupItem {
    ObjConst c;
    ObjArrays a;
    LsWork w;
    Win wn, wr;                   // the rebuild's window, the row dilation's
    int R;                        // dilation radius
    int pin;                      // pin before the rebuild
    const double *pin_x, *pin_y;  // (nX, nY) standing in for (X, Y) inside, or null
    const double *v;              // finish: the fixed-point buffer of the last pass
    const double *fx, *fy;        // extension copy-out: the buffers of the last pass
};
__global__ void k_pin_b(Grid g, const SetupItem *it) {
    const SetupItem &x = it[blockIdx.z];
    if (x.pin) pin_body(g, x.c, x.a, x.wn, x.pin_x, x.pin_y);
}
__global__ void k_counts_b(const SetupItem *it, int from, int n) {
    int *cnt = it[blockIdx.z].w.cnt + from;
    for (int k = threadIdx.x; k < n; k += blockDim.x) cnt[k] = 0;
}
__global__ void k_scan_b(Grid g, const SetupItem *it) {
    const SetupItem &x = it[blockIdx.z];
    scan_body(g, x.a, x.w, x.wn);
}
__global__ void k_newton_b(Grid g, LsGeom lg, const SetupItem *it) {
    const SetupItem &x = it[blockIdx.z];
    newton_body(g, lg, x.a, x.w);
}
__global__ void k_dilate_rows_b(Grid g, const SetupItem *it) {
    const SetupItem &x = it[blockIdx.z];
    dilate_rows_body(g, x.a, x.w, x.R < g.ml ? x.R : g.ml, x.wr);
}
__global__ void k_dilate_cols_b(Grid g, const SetupItem *it) {
    const SetupItem &x = it[blockIdx.z];
    dilate_cols_body(g, x.a, x.w, x.R < g.nl ? x.R : g.nl, x.wn);
}
__global__ void k_compact_tiles_b(const SetupItem *it, int ext) {
    const LsWork &w = it[blockIdx.z].w;
    if (ext) compact_tiles_body(w, 2, w.tiles_ext, kCntTileExt);
    else compact_tiles_body(w, 1, w.tiles_fim, kCntTileFim);
}
__global__ void k_bbox_init_b(const SetupItem *it) {
    const LsWork &w = it[blockIdx.z].w;
    if (threadIdx.x == 0) {
        w.cnt[kCntPband] = 0;
        w.bbox[0] = w.bbox[2] = 1 << 30;
        w.bbox[1] = w.bbox[3] = -1;
    }
}
__global__ void k_finalize_b(Grid g, const SetupItem *it) {
    const SetupItem &x = it[blockIdx.z];
    finalize_body(g, x.c, x.a, x.w, x.v, x.wn);
}
__global__ void k_commit_status_b(Grid g, const SetupItem *it) {
    const SetupItem &x = it[blockIdx.z];
    commit_status_body(g, x.a, x.w, x.wn);
}
__global__ void k_ext_copy_b(const SetupItem *it, int to_obj) {
    const SetupItem &x = it[blockIdx.z];
    ext_copy_body(x.a, x.w, x.fx, x.fy, to_obj);
}

#define KCHECK() HIP_CHECK(hipGetLastError())

// Run a fixed-point iteration in batches of passes, reading the change
// counts back once per batch. A pass whose predecessor in the batch changed
// nothing returns at once (the fixed point is reached and both ping-pong
// buffers hold it), so a batch may safely run past convergence; its length
// is the number of passes the previous iteration needed, which usually
// makes one host round trip per iteration. *passes_out: passes up to and
// including the first one that changed nothing. Returns 0, 7 when
// max_passes pass without convergence, or a HIP error.
// Its two halves: enqueue one batch of passes (starting at pass `passes`)
// with the read-back of its counters, and, once the stream has run them,
// decide from the counters whether the iteration has ended.
template <class F>
int iter_enqueue(hipStream_t s, LsWork &w, int predict, int passes, F pass, int &batch) {
    batch = predict < 2 ? 2 : predict;
    if (batch > kIterSlots) batch = kIterSlots;
    k_zero_counts<<<1, 64, 0, s>>>(w.cnt + kCntIter, kIterSlots);
    KCHECK();
    for (int b = 0; b < batch; b++)
        if (int rc = pass(passes + b, b)) return rc;
    HIP_CHECK(hipMemcpyAsync(w.hcnt, w.cnt, (kCntIter + kIterSlots) * sizeof(int), hipMemcpyDeviceToHost, s));
    return 0;
}
bool iter_done(LsWork &w, int &predict, int &passes, int batch, int *passes_out) {
    if (w.hcnt[kCntErr]) {
        *passes_out = passes + batch;
        return true;
    }
    for (int b = 0; b < batch; b++)
        if (w.hcnt[kCntIter + b] == 0) {
            *passes_out = passes + b + 1;
            predict = passes + b + 1;
            return true;
        }
    passes += batch;
    predict = passes + 2;
    return false;
}
template <class F>
int iterate(hipStream_t s, LsWork &w, int &predict, int max_passes, F pass, int *passes_out) {
    int passes = 0;
    while (passes < max_passes) {
        int batch;
        if (int rc = iter_enqueue(s, w, predict, passes, pass, batch)) return rc;
        HIP_CHECK(hipStreamSynchronize(s));
        if (iter_done(w, predict, passes, batch, passes_out)) return 0;
    }
    *passes_out = passes;
    return 7;
}

// The same iteration for several solids at once: every pass is one launch
// for all solids not yet at their fixed point (launch(p, slot, items, n)),
// one gather of their counters per batch. Each solid's passes, results,
// pass count and prediction are those of iterate on it alone; its error
// flags go to its host mirror (hcnt[kCntErr]).
// Its state between the halves (enqueue one batch of passes with the
// gather and read-back of the counters; once the stream has run them,
// decide per solid), so that a caller can overlap the first batch with other
// work.
template <class Item>
struct BatchIter {
    std::vector<Item> items;
    std::vector<int *> predict;
    std::vector<size_t> idx;  // solids still going
    std::vector<int> passes_out;
    int passes = 0, batch = 0;
    void start(std::vector<Item> it, std::vector<int *> pr) {
        items = std::move(it);
        predict = std::move(pr);
        idx.resize(items.size());
        for (size_t k = 0; k < idx.size(); k++) idx[k] = k;
        passes_out.assign(items.size(), 0);
        passes = batch = 0;
    }
};
template <class Item, class F>
int batch_enqueue(hipStream_t s, LsBatch &bt, BatchIter<Item> &it, F launch) {
    Item *d_items = (Item *)bt.d_items, *h_items = (Item *)bt.h_items;
    const int n = (int)it.idx.size();
    int batch = 2;
    for (size_t k : it.idx) batch = std::max(batch, *it.predict[k] < 2 ? 2 : *it.predict[k]);
    if (batch > kIterSlots) batch = kIterSlots;
    it.batch = batch;
    for (int k = 0; k < n; k++) h_items[k] = it.items[it.idx[k]];
    HIP_CHECK(hipMemcpyAsync(d_items, h_items, n * sizeof(Item), hipMemcpyHostToDevice, s));
    k_batch_zero<Item><<<n < 256 ? n : 256, 64, 0, s>>>(d_items, n);
    KCHECK();
    for (int b = 0; b < batch; b++)
        if (int rc = launch(it.passes + b, b, d_items, n)) return rc;
    k_batch_gather<Item><<<n < 256 ? n : 256, 64, 0, s>>>(d_items, n, batch, bt.d_cnt);
    KCHECK();
    HIP_CHECK(hipMemcpyAsync(bt.h_cnt, bt.d_cnt, (size_t)n * (batch + 1) * sizeof(int), hipMemcpyDeviceToHost, s));
    return 0;
}
template <class Item>
void batch_check(LsBatch &bt, BatchIter<Item> &it) {
    const int n = (int)it.idx.size(), batch = it.batch;
    std::vector<size_t> rest;
    for (int k = 0; k < n; k++) {
        const int *c = bt.h_cnt + (size_t)k * (batch + 1);
        const size_t id = it.idx[k];
        it.items[id].w.hcnt[kCntErr] = c[0];
        if (c[0]) {
            it.passes_out[id] = it.passes + batch;
            continue;
        }
        int done = -1;
        for (int b = 0; b < batch && done < 0; b++)
            if (c[1 + b] == 0) done = b;
        if (done >= 0) {
            it.passes_out[id] = it.passes + done + 1;
            *it.predict[id] = it.passes + done + 1;
        } else {
            rest.push_back(id);
        }
    }
    it.passes += batch;
    for (size_t id : rest) *it.predict[id] = it.passes + 2;
    it.idx.swap(rest);
}
// the loop from the state left by batch_enqueue's first batch, or from the start
template <class Item, class F>
int batch_finish(hipStream_t s, LsBatch &bt, BatchIter<Item> &it, int max_passes, F launch, bool pending) {
    if (pending) {
        HIP_CHECK(hipStreamSynchronize(s));
        batch_check(bt, it);
    }
    while (!it.idx.empty()) {
        if (it.passes >= max_passes) {
            for (size_t k : it.idx) it.passes_out[k] = it.passes;
            return 7;
        }
        if (int rc = batch_enqueue(s, bt, it, launch)) return rc;
        HIP_CHECK(hipStreamSynchronize(s));
        batch_check(bt, it);
    }
    return 0;
}
template <class Item, class F>
int iterate_batch(hipStream_t s, LsBatch &bt, std::vector<Item> items, std::vector<int *> predict, int max_passes,
                  F launch, std::vector<int> &passes_out) {
    BatchIter<Item> it;
    it.start(std::move(items), std::move(predict));
    const int rc = batch_finish(s, bt, it, max_passes, launch, false);
    passes_out = it.passes_out;
    return rc;
}

}  // namespace

int ls_batch_alloc(int cap, LsBatch &bt) {
    const size_t bytes = (size_t)cap * std::max(sizeof(FimItem), sizeof(ExtItem));
    HIP_CHECK(hipMalloc(&bt.d_items, bytes));
    HIP_CHECK(hipHostMalloc(&bt.h_items, bytes, hipHostMallocDefault));
    HIP_CHECK(hipMalloc((void **)&bt.d_cnt, (size_t)cap * (kIterSlots + 1) * sizeof(int)));
    HIP_CHECK(hipHostMalloc((void **)&bt.h_cnt, (size_t)cap * (kIterSlots + 1) * sizeof(int), hipHostMallocDefault));
    bt.setup_stride = ((size_t)cap * sizeof(SetupItem) + 255) / 256 * 256;
    HIP_CHECK(hipMalloc(&bt.d_setup, 4 * bt.setup_stride));
    HIP_CHECK(hipHostMalloc(&bt.h_setup, 4 * bt.setup_stride, hipHostMallocDefault));
    for (hipEvent_t &e : bt.ev_setup) HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    bt.cap = cap;
    return 0;
}

void ls_batch_free(LsBatch &bt) {
    if (bt.d_items) (void)hipFree(bt.d_items);
    if (bt.h_items) (void)hipHostFree(bt.h_items);
    if (bt.d_cnt) (void)hipFree(bt.d_cnt);
    if (bt.h_cnt) (void)hipHostFree(bt.h_cnt);
    if (bt.d_setup) (void)hipFree(bt.d_setup);
    if (bt.h_setup) (void)hipHostFree(bt.h_setup);
    for (hipEvent_t e : bt.ev_setup)
        if (e) (void)hipEventDestroy(e);
    bt = LsBatch{};
}

int ls_alloc(const Grid &g, LsWork &w, hipStream_t s) {
    const size_t N = (size_t)g.N;
    double **d[] = {&w.va, &w.vb, &w.ex, &w.ey, &w.fx, &w.fy};
    for (auto p : d) {
        HIP_CHECK(hipMalloc((void **)p, N * sizeof(double)));
        HIP_CHECK(memset_sync(*p, 0, N * sizeof(double), s));
    }
    const size_t L = (size_t)g.ml * g.nl + 64;
    HIP_CHECK(hipMalloc((void **)&w.iface, L * sizeof(int)));
    HIP_CHECK(hipMalloc((void **)&w.cand, L * sizeof(int)));
    HIP_CHECK(hipMalloc((void **)&w.pband, L * sizeof(int)));
    HIP_CHECK(hipMalloc((void **)&w.mark, N));
    HIP_CHECK(memset_sync(w.mark, 0, N, s));
    HIP_CHECK(hipMalloc((void **)&w.cflag, N));
    HIP_CHECK(memset_sync(w.cflag, 0, N, s));
    w.ntx = (g.ml + kTileX - 1) / kTileX;
    w.nty = (g.nl + kTileY - 1) / kTileY;
    const size_t T = (size_t)w.ntx * w.nty + 64;
    HIP_CHECK(hipMalloc((void **)&w.tflag, T * sizeof(int)));
    HIP_CHECK(memset_sync(w.tflag, 0, T * sizeof(int), s));
    HIP_CHECK(hipMalloc((void **)&w.tiles_fim, T * sizeof(int)));
    HIP_CHECK(hipMalloc((void **)&w.tiles_ext, T * sizeof(int)));
    HIP_CHECK(hipMalloc((void **)&w.cnt, (kCntIter + kIterSlots) * sizeof(int)));
    HIP_CHECK(hipHostMalloc((void **)&w.hcnt, (kCntIter + kIterSlots) * sizeof(int), hipHostMallocDefault));
    HIP_CHECK(hipMalloc((void **)&w.bbox, 4 * sizeof(int)));
    HIP_CHECK(hipHostMalloc((void **)&w.hbbox, 4 * sizeof(int), hipHostMallocDefault));
    HIP_CHECK(hipEventCreateWithFlags(&w.ev_bbox, hipEventDisableTiming));
    return 0;
}

void ls_free(LsWork &w) {
    void *p[] = {w.va, w.vb, w.ex, w.ey, w.fx, w.fy, w.iface, w.cand, w.pband, w.mark, w.cflag, w.tflag,
                 w.tiles_fim, w.tiles_ext, w.cnt, w.bbox};
    for (auto q : p)
        if (q) (void)hipFree(q);
    if (w.hcnt) (void)hipHostFree(w.hcnt);
    if (w.hbbox) (void)hipHostFree(w.hbbox);
    if (w.ev_bbox) (void)hipEventDestroy(w.ev_bbox);
    w = LsWork{};
}

// The points a rebuild may change lie within the reach of the march from
// the previous band. The window is the union of every earlier window and
// the previous band's bounding box (statuses 1..6) widened by that reach
// plus two, aligned outward to whole tiles; it never shrinks, so every
// point outside it still holds what the first, whole-grid rebuild left
// there (status 0 or 7, phi at the cut-off, +-inf in the march buffers).
static Win next_window(const Grid &g, LsWork &w, int R) {
    const Win full{0, g.ml, 0, g.nl};
    if (!w.have_win) return full;
    Win v{w.win[0], w.win[1], w.win[2], w.win[3]};
    if (w.bbox_pending) {
        (void)hipEventSynchronize(w.ev_bbox);
        w.bbox_pending = false;
        const int *b = w.hbbox;
        if (b[1] >= 0) {
            const int m = R + 2;
            int I0 = (b[0] - m) / kTileX * kTileX, I1 = ((b[1] + 1 + m + kTileX - 1) / kTileX) * kTileX;
            int J0 = (b[2] - m) / kTileY * kTileY, J1 = ((b[3] + 1 + m + kTileY - 1) / kTileY) * kTileY;
            if (b[0] - m < 0) I0 = 0;
            if (b[2] - m < 0) J0 = 0;
            if (I1 > g.ml) I1 = g.ml;
            if (J1 > g.nl) J1 = g.nl;
            if (v.I1 <= v.I0) v = Win{I0, I1, J0, J1};
            else v = Win{I0 < v.I0 ? I0 : v.I0, I1 > v.I1 ? I1 : v.I1, J0 < v.J0 ? J0 : v.J0, J1 > v.J1 ? J1 : v.J1};
        }
    }
    return v;
}

// A rebuild in three parts: the set-up (pinning, interface list, Newton
// closest points, dilation to the march region, active tiles), the march
// passes, and the finish (final band and statuses, the extension's tiles).
// *live = false: nothing of the solid is left on the grid (nothing to do).
// The host side of a rebuild's set-up: its window wn, the row dilation's
// window wr, the dilation radius R; false when the window is empty.
static bool rebuild_window(const Grid &g, const ObjConst &c, LsWork &w, Win &wn, Win &wr, bool &whole, int &R) {
    // the march cannot accept a point further than limit/h + 2 cells (max
    // norm) from the interface layer: beyond it every path is longer
    const double reach = fmax(-c.phi_down, c.phi_up) / fmin(g.dx, g.dy) + 3;
    R = reach > g.ml + g.nl ? g.ml + g.nl : (int)reach;
    whole = !w.have_win;
    wn = next_window(g, w, R);
    if (wn.I1 <= wn.I0 || wn.J1 <= wn.J0) return false;
    // the row dilation also covers R rows either side, which the column pass reads
    wr = Win{wn.I0, wn.I1, wn.J0 - R < 0 ? 0 : wn.J0 - R, wn.J1 + R > g.nl ? g.nl : wn.J1 + R};
    return true;
}

static int rebuild_setup(hipStream_t s, const Grid &g, const ObjConst &c, const ObjArrays &a, LsWork &w, Win &wn,
                         bool &whole, bool *live, const double *pin_x = nullptr, const double *pin_y = nullptr) {
    const dim3 bb(BX, BY);
    Win wr;
    int R;
    *live = rebuild_window(g, c, w, wn, wr, whole, R);
    if (!*live) return 0;
    if (w.pin_first) {
        k_pin<<<grid_win(wn), bb, 0, s>>>(g, c, a, wn, pin_x, pin_y);
        KCHECK();
    }
    k_zero_counts<<<1, 64, 0, s>>>(w.cnt, kCntIter + kIterSlots);
    KCHECK();
    k_scan<<<grid_win(wn), bb, 0, s>>>(g, a, w, wn);
    KCHECK();
    k_newton<<<kListBlocks, kListThreads, 0, s>>>(g, geom_of(g), a, w);
    KCHECK();
    k_dilate_rows<<<grid_win(wr), bb, 0, s>>>(g, a, w, R < g.ml ? R : g.ml, wr);
    KCHECK();
    k_dilate_cols<<<grid_win_cols(wn), bb, 0, s>>>(g, a, w, R < g.nl ? R : g.nl, wn);
    KCHECK();
    k_compact_tiles<<<64, 256, 0, s>>>(w, 1, w.tiles_fim, kCntTileFim);
    KCHECK();
    return 0;
}

// The host side of a rebuild's finish (after the finalisation): the band
// box's read-back and the window the next rebuild grows from.
static int finish_window(hipStream_t s, LsWork &w, const Win &wn, bool whole) {
    if (w.ev_bbox) {  // a work set without the event keeps no window (one-off rebuilds)
        HIP_CHECK(hipMemcpyAsync(w.hbbox, w.bbox, 4 * sizeof(int), hipMemcpyDeviceToHost, s));
        HIP_CHECK(hipEventRecord(w.ev_bbox, s));
        w.bbox_pending = true;
    }
    w.have_win = true;
    if (whole) {  // windows grow from the band's box, not from the whole grid
        w.win[0] = w.win[1] = w.win[2] = w.win[3] = 0;
    } else {
        w.win[0] = wn.I0; w.win[1] = wn.I1; w.win[2] = wn.J0; w.win[3] = wn.J1;
    }
    return 0;
}

static int rebuild_finish(hipStream_t s, const Grid &g, const ObjConst &c, const ObjArrays &a, LsWork &w,
                          const Win &wn, bool whole, int passes) {
    const dim3 bb(BX, BY);
    double *bufs[2] = {w.va, w.vb};
    k_zero_counts<<<1, 64, 0, s>>>(w.cnt + kCntPband, 1);
    KCHECK();
    k_bbox_init<<<1, 64, 0, s>>>(w);
    KCHECK();
    k_finalize<<<grid_win(wn), bb, 0, s>>>(g, c, a, w, bufs[passes & 1], wn);
    KCHECK();
    if (int rc = finish_window(s, w, wn, whole)) return rc;
    k_compact_tiles<<<64, 256, 0, s>>>(w, 2, w.tiles_ext, kCntTileExt);
    KCHECK();
    k_commit_status<<<grid_win(wn), bb, 0, s>>>(g, a, w, wn);
    KCHECK();
    return 0;
}

constexpr int kMaxFimPasses = 4096;

// One march pass of a single solid
static int fim_pass(hipStream_t s, const Grid &g, const EikGeom &eg, const ObjConst &c, const ObjArrays &a,
                    const LsWork &w, int p, int slot) {
    double *bufs[2] = {w.va, w.vb};
    k_fim_tiles<<<kTileBlocks, kTileThreads, 0, s>>>(g, eg, c, a, w, bufs[p & 1], bufs[(p + 1) & 1], slot, 64);
    KCHECK();
    return 0;
}

static int rebuild_common(hipStream_t s, const Grid &g, const ObjConst &c, const ObjArrays &a, LsWork &w,
                          int *fim_iters) {
    Win wn;
    bool whole, live;
    if (int rc = rebuild_setup(s, g, c, a, w, wn, whole, &live)) return rc;
    if (!live) return 0;
    EikGeom eg;
    eg.init(g.dx, g.dy);
    int passes = 0;
    int rc = iterate(s, w, w.pred_fim, kMaxFimPasses,
                     [&](int p, int slot) { return fim_pass(s, g, eg, c, a, w, p, slot); }, &passes);
    if (rc) return rc;
    if (w.hcnt[kCntErr] & kErrNewton) return 3;
    if (fim_iters) *fim_iters = passes;
    return rebuild_finish(s, g, c, a, w, wn, whole, passes);
}

// Workgroups per solid of a batched pass: the single-solid count shared out,
// at least 64
static int batch_blocks(int n) { return std::max(64, kTileBlocks / std::max(1, (n + 3) / 4)); }

// Upload the items of one region (0 set-up, 1 finish, 2 copy-in, 3
// copy-out) after its last upload has run; the device copy of the items.
static const SetupItem *upload_items(hipStream_t s, LsBatch &bt, int region, const std::vector<SetupItem> &v,
                                     int *rc) {
    *rc = 0;
    if (bt.setup_used[region]) {
        const hipError_t e = hipEventSynchronize(bt.ev_setup[region]);
        if (e != hipSuccess) {
            *rc = hip_fail(e, "hipEventSynchronize", __FILE__, __LINE__);
            return nullptr;
        }
    }
    SetupItem *h = (SetupItem *)((char *)bt.h_setup + region * bt.setup_stride);
    SetupItem *d = (SetupItem *)((char *)bt.d_setup + region * bt.setup_stride);
    std::copy(v.begin(), v.end(), h);
    hipError_t e = hipMemcpyAsync(d, h, v.size() * sizeof(SetupItem), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipEventRecord(bt.ev_setup[region], s);
    if (e != hipSuccess) {
        *rc = hip_fail(e, "upload of batch items", __FILE__, __LINE__);
        return nullptr;
    }
    bt.setup_used[region] = true;
    return d;
}
// a grid covering the largest of the items' windows (rows: their row windows)
static dim3 grid_items(const std::vector<SetupItem> &v, bool rows, bool cols = false) {
    unsigned gx = 1, gy = 1;
    for (const SetupItem &x : v) {
        const dim3 d = cols ? grid_win_cols(x.wn) : grid_win(rows ? x.wr : x.wn);
        gx = std::max(gx, d.x);
        gy = std::max(gy, d.y);
    }
    return dim3(gx, gy, (unsigned)v.size());
}
static int list_blocks(size_t n) { return std::max(16, kListBlocks / (int)std::max<size_t>(1, n)); }

// A batched rebuild between its halves.
struct LsBatchPending {
    std::vector<Win> wn;
    std::vector<char> whole;
    std::vector<SetupItem> su;
    std::vector<size_t> who;
    BatchIter<FimItem> it;
    EikGeom eg;
    bool live = false;
};
void ls_batch_pending_free(LsBatchPending *p) { delete p; }

int ls_pin_rebuild_batch_begin(hipStream_t s, const Grid &g, const std::vector<LsJob> &jobs, LsBatch &bt,
                               bool pin_new, LsBatchPending **out) {
    *out = nullptr;
    const size_t n = jobs.size();
    if (n > (size_t)bt.cap) return 7;
    LsBatchPending *p = new LsBatchPending;
    *out = p;
    p->wn.resize(n);
    p->whole.resize(n);
    std::vector<FimItem> items;
    std::vector<int *> pred;
    for (size_t k = 0; k < n; k++) {
        LsWork &w = *jobs[k].w;
        w.pin_first = true;
        Win wr;
        int R;
        bool wh;
        const bool lv = rebuild_window(g, *jobs[k].c, w, p->wn[k], wr, wh, R);
        p->whole[k] = wh;
        if (!lv) continue;
        const ObjArrays &a = *jobs[k].a;
        p->su.push_back(SetupItem{*jobs[k].c, a, w, p->wn[k], wr, R, 1, pin_new ? a.nX : nullptr,
                                  pin_new ? a.nY : nullptr, nullptr, nullptr, nullptr});
        items.push_back(FimItem{*jobs[k].c, a, w, {w.va, w.vb}});
        pred.push_back(&w.pred_fim);
        p->who.push_back(k);
    }
    if (items.empty()) return 0;
    p->live = true;
    // the set-up of every live solid (the single-solid rebuild_setup's
    // launches, one per kernel for all of them)
    int rc = 0;
    const SetupItem *d = upload_items(s, bt, 0, p->su, &rc);
    if (rc) return rc;
    const dim3 bb(BX, BY), gw = grid_items(p->su, false), gr = grid_items(p->su, true),
               gc = grid_items(p->su, false, true);
    const unsigned m = (unsigned)p->su.size();
    k_pin_b<<<gw, bb, 0, s>>>(g, d);
    KCHECK();
    k_counts_b<<<dim3(1, 1, m), 64, 0, s>>>(d, 0, kCntIter + kIterSlots);
    KCHECK();
    k_scan_b<<<gw, bb, 0, s>>>(g, d);
    KCHECK();
    k_newton_b<<<dim3(list_blocks(m), 1, m), kListThreads, 0, s>>>(g, geom_of(g), d);
    KCHECK();
    k_dilate_rows_b<<<gr, bb, 0, s>>>(g, d);
    KCHECK();
    k_dilate_cols_b<<<gc, bb, 0, s>>>(g, d);
    KCHECK();
    k_compact_tiles_b<<<dim3(64, 1, m), 256, 0, s>>>(d, 0);
    KCHECK();
    // the predicted batch of march passes, its counters read back later
    p->eg.init(g.dx, g.dy);
    p->it.start(std::move(items), std::move(pred));
    const int nb = batch_blocks((int)p->it.items.size());
    const EikGeom eg = p->eg;
    return batch_enqueue(s, bt, p->it, [&](int q, int slot, const FimItem *dd, int mm) -> int {
        k_fim_tiles_batch<<<dim3(nb, mm), kTileThreads, 0, s>>>(g, eg, dd, q, slot, 64);
        KCHECK();
        return 0;
    });
}

int ls_pin_rebuild_batch_end(hipStream_t s, const Grid &g, const std::vector<LsJob> &jobs, LsBatch &bt,
                             LsBatchPending *p, int *fim_iters) {
    if (!p->live) {
        delete p;
        return 0;
    }
    const int nb = batch_blocks((int)p->it.items.size());
    const EikGeom eg = p->eg;
    int rc = batch_finish(s, bt, p->it, kMaxFimPasses, [&](int q, int slot, const FimItem *dd, int mm) -> int {
        k_fim_tiles_batch<<<dim3(nb, mm), kTileThreads, 0, s>>>(g, eg, dd, q, slot, 64);
        KCHECK();
        return 0;
    }, true);
    if (!rc) rc = [&]() -> int {
        int most = 0;
        for (size_t t = 0; t < p->who.size(); t++) {
            LsWork &w = *jobs[p->who[t]].w;
            if (w.hcnt[kCntErr] & kErrNewton) return 3;
            most = std::max(most, p->it.passes_out[t]);
            p->su[t].v = p->it.passes_out[t] & 1 ? w.vb : w.va;
        }
        // the finish of every live solid (rebuild_finish's launches, shared)
        int e = 0;
        const SetupItem *d = upload_items(s, bt, 1, p->su, &e);
        if (e) return e;
        const dim3 bb(BX, BY), gw = grid_items(p->su, false);
        const unsigned m = (unsigned)p->su.size();
        k_bbox_init_b<<<dim3(1, 1, m), 64, 0, s>>>(d);
        KCHECK();
        k_finalize_b<<<gw, bb, 0, s>>>(g, d);
        KCHECK();
        for (size_t t = 0; t < p->who.size(); t++)
            if ((e = finish_window(s, *jobs[p->who[t]].w, p->wn[p->who[t]], p->whole[p->who[t]]))) return e;
        k_compact_tiles_b<<<dim3(64, 1, m), 256, 0, s>>>(d, 1);
        KCHECK();
        k_commit_status_b<<<gw, bb, 0, s>>>(g, d);
        KCHECK();
        if (fim_iters) *fim_iters = most;
        return 0;
    }();
    delete p;
    return rc;
}

int ls_pin_rebuild_batch(hipStream_t s, const Grid &g, const std::vector<LsJob> &jobs, LsBatch &bt,
                         int *fim_iters) {
    LsBatchPending *p = nullptr;
    if (int rc = ls_pin_rebuild_batch_begin(s, g, jobs, bt, false, &p)) {
        if (p) (void)hipStreamSynchronize(s);
        delete p;
        return rc;
    }
    return ls_pin_rebuild_batch_end(s, g, jobs, bt, p, fim_iters);
}

int ls_rebuild(hipStream_t s, const Grid &g, const ObjConst &c, const ObjArrays &a, LsWork &w, int *fim_iters) {
    ls_reset_window(w);  // a new field: the whole grid
    w.pin_first = false;
    return rebuild_common(s, g, c, a, w, fim_iters);
}

int ls_pin_rebuild(hipStream_t s, const Grid &g, const ObjConst &c, const ObjArrays &a, LsWork &w,
                   int *fim_iters) {
    w.pin_first = true;
    return rebuild_common(s, g, c, a, w, fim_iters);
}

int ls_pin_rebuild_begin(hipStream_t s, const Grid &g, const ObjConst &c, const ObjArrays &a, LsWork &w,
                         const double *pin_x, const double *pin_y, LsRebuild &r) {
    w.pin_first = true;
    r = LsRebuild{};
    if (int rc = rebuild_setup(s, g, c, a, w, r.wn, r.whole, &r.live, pin_x, pin_y)) return rc;
    if (!r.live) return 0;
    r.eg.init(g.dx, g.dy);
    return iter_enqueue(s, w, w.pred_fim, 0,
                        [&](int p, int slot) { return fim_pass(s, g, r.eg, c, a, w, p, slot); }, r.batch);
}

int ls_pin_rebuild_end(hipStream_t s, const Grid &g, const ObjConst &c, const ObjArrays &a, LsWork &w, LsRebuild &r,
                       int *fim_iters) {
    if (!r.live) return 0;
    HIP_CHECK(hipStreamSynchronize(s));
    int passes = 0;
    // the rest of iterate from the batch begin enqueued
    bool done = iter_done(w, w.pred_fim, r.passes, r.batch, &passes);
    while (!done && r.passes < kMaxFimPasses) {
        if (int rc = iter_enqueue(s, w, w.pred_fim, r.passes,
                                  [&](int p, int slot) { return fim_pass(s, g, r.eg, c, a, w, p, slot); }, r.batch))
            return rc;
        HIP_CHECK(hipStreamSynchronize(s));
        done = iter_done(w, w.pred_fim, r.passes, r.batch, &passes);
    }
    if (!done) return 7;
    if (w.hcnt[kCntErr] & kErrNewton) return 3;
    if (fim_iters) *fim_iters = passes;
    return rebuild_finish(s, g, c, a, w, r.wn, r.whole, passes);
}

void ls_reset_window(LsWork &w) {
    if (w.bbox_pending) (void)hipEventSynchronize(w.ev_bbox);
    w.bbox_pending = false;
    w.have_win = false;
}

int ls_extend(hipStream_t s, const Grid &g, const ObjArrays &a, LsWork &w, int *ext_iters) {
    k_ext_copy<<<kListBlocks, kListThreads, 0, s>>>(a, w, w.ex, w.ey, 0);
    KCHECK();
    double *bx[2] = {w.ex, w.fx}, *by[2] = {w.ey, w.fy};
    int passes = 0;
    int rc = iterate(s, w, w.pred_ext, 4096, [&](int p, int slot) -> int {
        k_extend_tiles<<<kTileBlocks, kTileThreads, 0, s>>>(g, a, w, bx[p & 1], by[p & 1], bx[(p + 1) & 1], by[(p + 1) & 1],
                                                   slot, 128);
        KCHECK();
        return 0;
    }, &passes);
    if (rc) return rc;
    if (w.hcnt[kCntErr] & kErrSingular) return 5;
    if (ext_iters) *ext_iters = passes;
    k_ext_copy<<<kListBlocks, kListThreads, 0, s>>>(a, w, bx[passes & 1], by[passes & 1], 1);
    KCHECK();
    return 0;
}

int ls_extend_batch(hipStream_t s, const Grid &g, const std::vector<LsJob> &jobs, LsBatch &bt, int *ext_iters) {
    const size_t n = jobs.size();
    if (n > (size_t)bt.cap) return 7;
    if (!n) return 0;
    std::vector<ExtItem> items;
    std::vector<SetupItem> cp;
    std::vector<int *> pred;
    for (size_t k = 0; k < n; k++) {
        LsWork &w = *jobs[k].w;
        cp.push_back(SetupItem{*jobs[k].c, *jobs[k].a, w, {}, {}, 0, 0, nullptr, nullptr, nullptr, w.ex, w.ey});
        items.push_back(ExtItem{*jobs[k].a, w, {w.ex, w.fx}, {w.ey, w.fy}});
        pred.push_back(&w.pred_ext);
    }
    int rc = 0;
    const SetupItem *d = upload_items(s, bt, 2, cp, &rc);
    if (rc) return rc;
    k_ext_copy_b<<<dim3(list_blocks(n), 1, (unsigned)n), kListThreads, 0, s>>>(d, 0);
    KCHECK();
    const int nb = batch_blocks((int)items.size());
    std::vector<int> passes;
    rc = iterate_batch(s, bt, items, pred, 4096, [&](int p, int slot, const ExtItem *dd, int m) -> int {
        k_extend_tiles_batch<<<dim3(nb, m), kTileThreads, 0, s>>>(g, dd, p, slot, 128);
        KCHECK();
        return 0;
    }, passes);
    if (rc) return rc;
    int most = 0;
    for (size_t k = 0; k < n; k++) {
        LsWork &w = *jobs[k].w;
        if (w.hcnt[kCntErr] & kErrSingular) return 5;
        most = std::max(most, passes[k]);
        const int q = passes[k];
        cp[k].fx = q & 1 ? w.fx : w.ex;
        cp[k].fy = q & 1 ? w.fy : w.ey;
    }
    d = upload_items(s, bt, 3, cp, &rc);
    if (rc) return rc;
    k_ext_copy_b<<<dim3(list_blocks(n), 1, (unsigned)n), kListThreads, 0, s>>>(d, 1);
    KCHECK();
    if (ext_iters) *ext_iters = most;
    return 0;
}

}  // namespace incrmt
