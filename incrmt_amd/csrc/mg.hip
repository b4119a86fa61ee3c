// Geometric multigrid for the MAC and FEM projection systems (see mg.hh).
//
// Design for gfx950: every level is a padded 2D array with a zero ghost
// ring, so the smoothers, residuals and transfers are branch-free 5/9-point
// stencils over rows of 64 consecutive doubles per wavefront. Fine-level
// smoothing is the HBM-bound hot loop: per sweep it reads z, b and the
// level's coefficients and writes z once. Reductions (norms, means) are
// deterministic two-pass block reductions, so a solve is reproducible
// launch to launch.
#include <algorithm>
#include <functional>
#include <cstring>
#include <cstdio>
#include <cmath>
#include <limits>
#include <type_traits>
#include <hip/hip_ext.h>
#include "mg.hh"
#include "unet.hh"

namespace incrmt {

namespace {

constexpr int BX = 64, BY = 4;
constexpr int INCRMT_ERR_ARG_SHAPE = 1;  // unknown smoother shape
constexpr int INCRMT_ERR_GRID = 7;       // grid outside what the kernels index exactly
constexpr int kRedBlocks = 512;   // partial sums per reduction
constexpr int kRedThreads = 256;

inline int round_up(int a, int b) { return (a + b - 1) / b * b; }
inline dim3 grid2(int nx, int ny) { return dim3((nx + BX - 1) / BX, (ny + BY - 1) / BY); }

// ------------------------------------------------------------ operators
// MAC: A z = kx (fw_l (z_l - z) + fw_r (z_r - z)) + ky (fs_d (z_d - z) + fs_u (z_u - z))
template <class T>
__device__ inline T mac_offdiag(const T *z, const T *fw, const T *fs, int q, int c, int zp, int cp, T kx, T ky,
                                T &diag) {
    const T wl = fw[c], wr = fw[c + 1], sd = fs[c], su = fs[c + cp];
    diag = -(kx * (wl + wr) + ky * (sd + su));
    return kx * (wl * z[q - 1] + wr * z[q + 1]) + ky * (sd * z[q - zp] + su * z[q + zp]);
}

// FEM: bilinear-element stiffness with cell weights kc (mgs_fem.hh:86-136)
struct FemC {
    double fm, fc, hex, hey;
};
template <class T>
__device__ inline T fem_offdiag(const T *z, const T *kc, int q, int c, int zp, int cp, const FemC &k, T &diag) {
    // cells around node (i,j): dl=(i-1,j-1) dr=(i,j-1) ul=(i-1,j) ur=(i,j)
    const T kdl = kc[c - cp - 1], kdr = kc[c - cp], kul = kc[c - 1], kur = kc[c];
    diag = T(k.fm) * (kdl + kdr + kul + kur);
    return T(k.fc) * (kdl * z[q - zp - 1] + kdr * z[q - zp + 1] + kul * z[q + zp - 1] + kur * z[q + zp + 1])
           + T(k.hey) * ((kdl + kdr) * z[q - zp] + (kul + kur) * z[q + zp])
           + T(k.hex) * ((kdl + kul) * z[q - 1] + (kdr + kur) * z[q + 1]);
}

struct LevelArgs {
    int nx, ny, zp, cp;
    double kx, ky;
    int px, py;  // periodic directions
    int dr;      // zero-pressure right edge (ring column nx held at zero)
};

// Periodic directions keep the ghost ring of an iterate equal to the
// opposite edge: writing edge point (i,j) also writes the ring points that
// mirror it (column -1 for i = nx-1, column nx for i = 0, rows likewise,
// and the corner).
template <class T>
__device__ inline void put_mirrors(const LevelArgs &a, T *z, int i, int j, T v) {
    const bool xe = a.px && (i == 0 || i == a.nx - 1), ye = a.py && (j == 0 || j == a.ny - 1);
    if (!xe && !ye) return;
    const int mi = i == 0 ? a.nx : -1, mj = j == 0 ? a.ny : -1;
    if (xe) z[j * a.zp + mi] = v;
    if (ye) z[mj * a.zp + i] = v;
    if (xe && ye) z[mj * a.zp + mi] = v;
}
template <class T>
__device__ inline void put_z(const LevelArgs &a, T *z, int i, int j, T v) {
    z[j * a.zp + i] = v;
    put_mirrors(a, z, i, j, v);
}
// periodic index of a point outside [0, n)
__device__ inline int wrap_idx(int i, int n) { return ((i % n) + n) % n; }

// --------------------------------------------------------------- smoothers
// Red-black Gauss-Seidel for the 5-point MAC operator: thread t of row j
// updates i = 2t + ((j + color) & 1).
template <class T>
__global__ void k_mac_smooth(LevelArgs a, T *z, const T *b, const T *fw, const T *fs, int color) {
    const int t = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y;
    const int i = 2 * t + ((j + color) & 1);
    if (i >= a.nx || j >= a.ny) return;
    const int q = j * a.zp + i, c = j * a.cp + i;
    T diag;
    const T off = mac_offdiag<T>(z, fw, fs, q, c, a.zp, a.cp, T(a.kx), T(a.ky), diag);
    if (diag != T(0)) put_z(a, z, i, j, (b[q] - off) / diag);
}

// Four-colour Gauss-Seidel for the 9-point FEM operator (colours (i%2,j%2)
// are a valid Gauss-Seidel ordering for a 9-point stencil).
template <class T>
__global__ void k_fem_smooth(LevelArgs a, FemC k, T *z, const T *b, const T *kc, int color) {
    const int i = 2 * (blockIdx.x * BX + threadIdx.x) + (color & 1);
    const int j = 2 * (blockIdx.y * BY + threadIdx.y) + (color >> 1);
    if (i >= a.nx || j >= a.ny) return;
    const int q = j * a.zp + i, c = j * a.cp + i;
    T diag;
    const T off = fem_offdiag<T>(z, kc, q, c, a.zp, a.cp, k, diag);
    if (diag != T(0)) put_z(a, z, i, j, (b[q] - off) / diag);
}

// Nine-colour Gauss-Seidel for a level of odd size along a periodic
// direction, where the two colours of i % 2 meet across the seam: the last
// point of such a direction takes a third colour, so no two neighbours of
// the 5- or 9-point stencil share a colour (i, j) -> cx + 3 cy. Colours
// with no points return at once.
template <class T>
__global__ void k_smooth_colour9(int kind, LevelArgs a, FemC k, T *z, const T *b, const T *c1, const T *c2,
                                 int colour) {
    const int cx = colour % 3, cy = colour / 3;
    const int t = blockIdx.x * BX + threadIdx.x, u = blockIdx.y * BY + threadIdx.y;
    const bool ox = a.px && (a.nx & 1), oy = a.py && (a.ny & 1);
    int i, j;
    if (cx == 2) {
        if (!ox || t) return;
        i = a.nx - 1;
    } else {
        i = 2 * t + cx;
        if (i >= a.nx - (ox ? 1 : 0)) return;
    }
    if (cy == 2) {
        if (!oy || u) return;
        j = a.ny - 1;
    } else {
        j = 2 * u + cy;
        if (j >= a.ny - (oy ? 1 : 0)) return;
    }
    const int q = j * a.zp + i, c = j * a.cp + i;
    T diag, off;
    if (kind == kMAC) off = mac_offdiag<T>(z, c1, c2, q, c, a.zp, a.cp, T(a.kx), T(a.ky), diag);
    else off = fem_offdiag<T>(z, c1, q, c, a.zp, a.cp, k, diag);
    if (diag != T(0)) put_z(a, z, i, j, (b[q] - off) / diag);
}

// Fused colored Gauss-Seidel: a whole multi-colour sweep (MAC: two sweeps)
// in one launch. A tile plus a halo of one point per colour phase is loaded
// into LDS with coalesced row reads; phase p updates its colour on the
// region shrunk by p points, so after the last phase the tile interior
// holds exactly what the sequence of single-colour launches would give. The
// launch reads z and writes zout (a second buffer): no workgroup can see a
// neighbour's result while loading its halo.
// Coarse-grid correction of a nested level, as the prolongation kernels
// below add it (same expressions, so fusing it into a smoother's load is
// bit-identical to a separate prolongation launch).
struct ProArgs {
    const void *e;   // coarse correction at (0,0), or null for none
    int ezp, cnx, cny;
    int px, py;      // periodic directions
    int dr;          // zero-pressure right edge
};
// MAC: bilinear cell-centred interpolation, the neighbour clamped at walls
// (constant extrapolation), wrapped across periodic seams and mirrored with
// the opposite sign past a zero-pressure edge (the correction vanishes on
// that face). FEM: bilinear
// vertex interpolation (the next coarse node of a periodic direction's last
// interval is node 0, which the coarse ring mirrors).
template <class T, int KIND>
__device__ inline T prolong_at(const ProArgs &pa, int i, int j) {
    const T *e = (const T *)pa.e;
    const int I = i >> 1, J = j >> 1, p = pa.ezp;
    if (KIND == kMAC) {
        int I2 = I + ((i & 1) ? 1 : -1), J2 = J + ((j & 1) ? 1 : -1);
        const bool flip = pa.dr && I2 >= pa.cnx;
        if (I2 < 0 || I2 >= pa.cnx) I2 = pa.px ? wrap_idx(I2, pa.cnx) : I;
        if (J2 < 0 || J2 >= pa.cny) J2 = pa.py ? wrap_idx(J2, pa.cny) : J;
        T e2 = e[J * p + I2], e22 = e[J2 * p + I2];
        if (flip) { e2 = -e2; e22 = -e22; }
        return T(0.5625) * e[J * p + I] + T(0.1875) * e2 + T(0.1875) * e[J2 * p + I] + T(0.0625) * e22;
    }
    const int q = J * p + I;
    if (!(i & 1) && !(j & 1)) return e[q];
    if (!(j & 1)) return T(0.5) * (e[q] + e[q + 1]);
    if (!(i & 1)) return T(0.5) * (e[q] + e[q + p]);
    return T(0.25) * (e[q] + e[q + 1] + e[q + p] + e[q + p + 1]);
}

// Threads per workgroup of a fused launch: one per tile point up to 1024.
template <int SX, int SY>
constexpr int smooth_threads() { return SX * SY < 1024 ? SX * SY : 1024; }

// LDS bytes of a fused launch (the periodic seam's extra rows and columns
// included), against gfx950's 160 KiB per workgroup
constexpr long smooth_lds(int kind, int h, int sx, int sy, int tsize) {
    return (long)(sx + 2 * h + 1) * (sy + 2 * h + 1) * (kind == kMAC ? 3 : 2) * tsize
           + (kind == kMAC ? (long)(sx + sy + 4 * h + 2) * (tsize + 1) : 0);
}
// Waves per SIMD a fused launch can keep resident as LDS allows (gfx950:
// 160 KiB of LDS and 32 waves per CU), lowered where the threads' register
// groups (below) and the working registers (about 30, 45 with periodic
// wrapping) exceed a lane's share of the 512 registers at that residency:
// the register budget the compiler is held to.
template <int KIND, int H, int SX, int SY, int TSIZE, bool PER>
constexpr int smooth_waves() {
    const int nt = smooth_threads<SX, SY>(), waves = nt / 64, rw = SX + 2 * H, rh = SY + 2 * H;
    int wgs = (int)(160L * 1024 / smooth_lds(KIND, H, SX, SY, TSIZE));
    if (wgs > 32 / waves) wgs = 32 / waves;
    int w = wgs * waves / 4;
    const int nph = KIND == kMAC ? (rw / 2) * rh : (rw / 2) * (rh / 2), kp = (nph + nt - 1) / nt;
    const int data = kp * (KIND == kMAC ? 2 : 17) * TSIZE / 4;  // registers of the groups' values
    const int work = PER ? 45 : 30;
    if (data + work > 512 / (w < 1 ? 1 : w)) w = 512 / (data + work + 10);
    return w < 1 ? 1 : w > 8 ? 8 : w;
}

// Register-group updates of the fused FEM smoother, with the expressions
// of fem_offdiag (so every form of the sweep is bit-identical). A
// neighbour inside the thread's 2x2 block comes from the block's
// registers, any other from the LDS tile.
template <int PW, int DX, int DY, int OX, int OY, class T>
__device__ inline T fem_grp_z(const T (&zo)[4], const T *sz, int q) {
    constexpr int bx = DX + OX, by = DY + OY;
    if constexpr (bx >= 0 && bx <= 1 && by >= 0 && by <= 1) return zo[bx + 2 * by];
    else return sz[q + OY * PW + OX];
}
// FEM: point (DX, DY) of a 2x2 block; kq holds the block's 3x3 cell
// weights (cell (cx, cy) at kq[3 cy + cx], cell 0 left of / below the block)
template <int PW, int DX, int DY, class T>
__device__ inline void fem_grp_update(T (&zo)[4], const T (&kq)[9], const T (&bq)[4], T *sz, int q, const FemC &k) {
    const T kdl = kq[3 * DY + DX], kdr = kq[3 * DY + DX + 1], kul = kq[3 * DY + 3 + DX], kur = kq[3 * DY + 4 + DX];
    const T diag = T(k.fm) * (kdl + kdr + kul + kur);
    const T off = T(k.fc) * (kdl * fem_grp_z<PW, DX, DY, -1, -1>(zo, sz, q) + kdr * fem_grp_z<PW, DX, DY, 1, -1>(zo, sz, q)
                             + kul * fem_grp_z<PW, DX, DY, -1, 1>(zo, sz, q) + kur * fem_grp_z<PW, DX, DY, 1, 1>(zo, sz, q))
                  + T(k.hey) * ((kdl + kdr) * fem_grp_z<PW, DX, DY, 0, -1>(zo, sz, q)
                                + (kul + kur) * fem_grp_z<PW, DX, DY, 0, 1>(zo, sz, q))
                  + T(k.hex) * ((kdl + kul) * fem_grp_z<PW, DX, DY, -1, 0>(zo, sz, q)
                                + (kdr + kur) * fem_grp_z<PW, DX, DY, 1, 0>(zo, sz, q));
    if (diag != T(0)) {
        const T v = (bq[DX + 2 * DY] - off) / diag;
        zo[DX + 2 * DY] = v;
        sz[q] = v;
    }
}
// PER: the level has a periodic direction or a zero-pressure edge (compiled
// separately, so a walled level runs none of the wrapping).
//
// Each thread owns one group of points per NT-th colour enumeration index:
// a 2x2 FEM block holding one point of each colour, or a horizontal MAC
// pair holding one of each, and keeps the group's right-hand sides in
// registers. A FEM thread also keeps the block's iterate values and cell
// weights there, so a phase reads from LDS only the 5 neighbours outside
// the block and writes each new value to LDS and to its register (for the
// 5-point MAC stencil the extra registers cost more than the LDS reads they
// save). The LDS tile holds the iterate and the coefficients (two FEM or
// three MAC arrays), which lets two workgroups of the common shapes share a
// CU.
template <class T, int KIND, int H, int kSX, int kSY, bool PER>
__global__ __launch_bounds__((smooth_threads<kSX, kSY>()))
__attribute__((amdgpu_waves_per_eu(smooth_waves<KIND, H, kSX, kSY, sizeof(T), PER>()))) void k_smooth_fused(LevelArgs a, FemC k, const T *z, T *zout,
                                                                            const T *b, const T *c1, const T *c2,
                                                                            int phases, int first_color, int dir,
                                                                            ProArgs pa) {
    if (!PER) { a.px = a.py = 0; pa.dr = 0; }
    constexpr int RW = kSX + 2 * H, RH = kSY + 2 * H, PW = RW + 1, PN = PW * (RH + 1);
    constexpr int NT = smooth_threads<kSX, kSY>();
    __shared__ T sz[PN], s1[PN], s2[KIND == kMAC ? PN : 1];
    // A MAC seam's far faces: face nx (right of the last cell) per region
    // row and face ny (above the last row) per region column, used by the
    // region points whose wrapped index is the last one (the face slot of
    // their wrapped neighbour holds face 0, the seam's other side).
    constexpr bool SEAM = PER && KIND == kMAC;
    __shared__ T sfr[SEAM ? RH + 1 : 1], sft[SEAM ? RW + 1 : 1];
    __shared__ unsigned char lastc[SEAM ? RW + 1 : 1], lastr[SEAM ? RH + 1 : 1];
    const int i0 = blockIdx.x * kSX - H, j0 = blockIdx.y * kSY - H;
    const int tid = threadIdx.x;
    // region point (li, lj) -> level point (i, j); periodic directions read
    // the region's outside points from the opposite edge: they are real
    // points, updated like the inside ones (a MAC seam's far face comes
    // from sfr / sft below). inpad: inside the level or its ghost ring.
    auto level_pt = [&](int li, int lj, int &i, int &j) -> bool {
        i = i0 + li;
        j = j0 + lj;
        if (a.px && (i < 0 || i >= a.nx)) i = wrap_idx(i, a.nx);
        if (a.py && (j < 0 || j >= a.ny)) j = wrap_idx(j, a.ny);
        return i >= -1 && i <= a.nx && j >= -1 && j <= a.ny;
    };
    // i0, j0, RW and RH are even, so a point's colour follows from its
    // region coordinates. Group g of a thread: FEM block at
    // (2 (g % HW), 2 (g / HW)), point of colour col at offset (col & 1,
    // col >> 1); MAC pair at (2 (g % HW), g / HW), the point of colour col
    // at offset (col + lj) & 1.
    constexpr int ncol = KIND == kMAC ? 2 : 4, HW = RW / 2;
    constexpr int NPH = KIND == kMAC ? HW * RH : HW * (RH / 2);
    constexpr int KP = (NPH + NT - 1) / NT;
    T bq[KP][ncol], zo[KP][ncol], kq[KP][9];
    int gli[KP], glj[KP];
#pragma unroll
    for (int m = 0; m < KP; m++) {
        const int g = tid + m * NT;
        gli[m] = 2 * (g % HW);
        glj[m] = KIND == kMAC ? g / HW : 2 * (g / HW);
#pragma unroll
        for (int s = 0; s < ncol; s++) {
            int i, j;
            const bool in = g < NPH && level_pt(gli[m] + (s & 1), glj[m] + (KIND == kMAC ? 0 : s >> 1), i, j);
            bq[m][s] = in ? b[j * a.zp + i] : T(0);
        }
    }
    // region points (li,lj) in [0,RW]x[0,RH]: the extra row/column holds the
    // faces to the right/top of the region; element e = tid + n NT walks
    // the rows incrementally
    {
        constexpr int DL = NT % PW, DJ = NT / PW;
        int li = tid % PW, lj = tid / PW;
        for (int e = tid; e < PN; e += NT) {
            int i, j;
            const bool inpad = level_pt(li, lj, i, j);
            const int q = j * a.zp + i, c = j * a.cp + i;
            T zv = inpad ? z[q] : T(0);
            if (pa.e && i >= 0 && i < a.nx && j >= 0 && j < a.ny) zv += prolong_at<T, KIND>(pa, i, j);
            sz[e] = zv;
            s1[e] = inpad ? c1[c] : T(0);
            if (KIND == kMAC) s2[e] = inpad ? c2[c] : T(0);
            li += DL;
            lj += DJ;
            if (li >= PW) { li -= PW; lj++; }
        }
    }
    if (SEAM) {
        for (int e = tid; e < RH + 1; e += NT) {
            int j = j0 + e;
            if (a.py) j = wrap_idx(j, a.ny);
            const bool in = j >= 0 && j < a.ny;
            sfr[e] = (a.px && in) ? c1[j * a.cp + a.nx] : T(0);
            lastr[e] = a.py && j == a.ny - 1;
        }
        for (int e = tid; e < RW + 1; e += NT) {
            int i = i0 + e;
            if (a.px) i = wrap_idx(i, a.nx);
            const bool in = i >= 0 && i < a.nx;
            sft[e] = (a.py && in) ? c2[a.ny * a.cp + i] : T(0);
            lastc[e] = a.px && i == a.nx - 1;
        }
    }
    __syncthreads();
    // FEM blocks take their iterate values and cell weights (cells left of /
    // below a block at the region's first column or row belong to points
    // no phase updates: clamped into the tile)
    if constexpr (KIND == kFEM) {
#pragma unroll
        for (int m = 0; m < KP; m++) {
            const bool ok = tid + m * NT < NPH;
            const int q0 = glj[m] * PW + gli[m];
#pragma unroll
            for (int s = 0; s < 4; s++) zo[m][s] = ok ? sz[q0 + (s >> 1) * PW + (s & 1)] : T(0);
#pragma unroll
            for (int c = 0; c < 9; c++) {
                const int li = max(gli[m] - 1 + c % 3, 0), lj = max(glj[m] - 1 + c / 3, 0);
                kq[m][c] = ok ? s1[lj * PW + li] : T(0);
            }
        }
    }
    // A region that reaches the ghost column (row) past the level's last
    // point in a walled direction has a fixed edge there: nothing outside
    // the region can change its points, so it is not shrunk on that side
    // (this lets the last tile of a row also finish the up to H-1 points
    // past its interior, see smooth_tiles).
    const bool openx = a.px || i0 + RW <= a.nx, openy = a.py || j0 + RH <= a.ny;
    for (int p = 1; p <= phases; p++) {
        // colour of phase p: forward or backward through the colours
        const int step = p - 1;
        const int col = dir > 0 ? (first_color + step) % ncol : ((first_color - step) % ncol + ncol) % ncol;
        const int rx = openx ? RW - p : RW, ry = openy ? RH - p : RH;
#pragma unroll
        for (int m = 0; m < KP; m++) {
            if (KP * NT > NPH && tid + m * NT >= NPH) continue;
            // slot of the colour in the thread's group
            const int sl = KIND == kMAC ? ((col + glj[m]) & 1) : col;
            const int li = gli[m] + (sl & 1), lj = glj[m] + (KIND == kMAC ? 0 : sl >> 1);
            if (li < p || li >= rx || lj < p || lj >= ry) continue;
            const int i = i0 + li, j = j0 + lj;
            if ((!a.px && (i < 0 || i >= a.nx)) || (!a.py && (j < 0 || j >= a.ny))) continue;
            const int q = lj * PW + li;
            if constexpr (KIND == kFEM) {
                switch (sl) {
                    case 0: fem_grp_update<PW, 0, 0>(zo[m], kq[m], bq[m], sz, q, k); break;
                    case 1: fem_grp_update<PW, 1, 0>(zo[m], kq[m], bq[m], sz, q, k); break;
                    case 2: fem_grp_update<PW, 0, 1>(zo[m], kq[m], bq[m], sz, q, k); break;
                    default: fem_grp_update<PW, 1, 1>(zo[m], kq[m], bq[m], sz, q, k); break;
                }
            } else {
                T diag, off;
                if (SEAM) {
                    // mac_offdiag with the seam's far faces where they apply
                    const T wl = s1[q], sd = s2[q];
                    const T wr = lastc[li] ? sfr[lj] : s1[q + 1];
                    const T su = lastr[lj] ? sft[li] : s2[q + PW];
                    const T kx = T(a.kx), ky = T(a.ky);
                    diag = -(kx * (wl + wr) + ky * (sd + su));
                    off = kx * (wl * sz[q - 1] + wr * sz[q + 1]) + ky * (sd * sz[q - PW] + su * sz[q + PW]);
                } else off = mac_offdiag<T>(sz, s1, s2, q, q, PW, PW, T(a.kx), T(a.ky), diag);
                // a select of values (a select of the two addresses would
                // index the register array dynamically)
                const T b0 = bq[m][0], b1 = bq[m][1], bv = sl ? b1 : b0;
                if (diag != T(0)) sz[q] = (bv - off) / diag;
            }
        }
        __syncthreads();
    }
    auto put = [&](int li, int lj) {
        const int i = i0 + li, j = j0 + lj;
        if (i < a.nx && j < a.ny) {
            if (PER) put_z(a, zout, i, j, sz[lj * PW + li]);
            else zout[j * a.zp + i] = sz[lj * PW + li];
        }
    };
    for (int e = tid; e < kSX * kSY; e += NT) put(e % kSX + H, e / kSX + H);
    // the last tile of a walled direction also writes the points past its
    // interior (at most H-1 columns / rows, all inside the region)
    const int ex = openx ? 0 : a.nx - (i0 + H + kSX), ey = openy ? 0 : a.ny - (j0 + H + kSY);
    if (ex > 0)
        for (int e = tid; e < ex * (kSY + (ey > 0 ? ey : 0)); e += NT) put(H + kSX + e % ex, H + e / ex);
    if (ey > 0)
        for (int e = tid; e < kSX * ey; e += NT) put(H + e % kSX, H + kSY + e / kSX);
}

// ------------------------------------------------- residual and transfers
// MAC: fused residual of the four children and averaging restriction.
template <class T>
__global__ void k_mac_restrict(LevelArgs f, LevelArgs cl, const T *z, const T *b, const T *fw, const T *fs,
                               T *bc, T *zc) {
    const int I = blockIdx.x * BX + threadIdx.x, J = blockIdx.y * BY + threadIdx.y;
    if (I >= cl.nx || J >= cl.ny) return;
    T s = 0;
    for (int dj = 0; dj < 2; dj++)
        for (int di = 0; di < 2; di++) {
            const int i = 2 * I + di, j = 2 * J + dj, q = j * f.zp + i, c = j * f.cp + i;
            T diag;
            const T off = mac_offdiag<T>(z, fw, fs, q, c, f.zp, f.cp, T(f.kx), T(f.ky), diag);
            s += b[q] - off - diag * z[q];
        }
    bc[J * cl.zp + I] = T(0.25) * s;
    put_z(cl, zc, I, J, T(0));
}

// FEM: residual into r, then full-weighting restriction (the transpose of
// bilinear interpolation, weights 1, 1/2, 1/4).
template <class T>
__global__ void k_fem_residual(LevelArgs a, FemC k, const T *z, const T *b, const T *kc, T *r) {
    const int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y;
    if (i >= a.nx || j >= a.ny) return;
    const int q = j * a.zp + i, c = j * a.cp + i;
    T diag;
    const T off = fem_offdiag<T>(z, kc, q, c, a.zp, a.cp, k, diag);
    r[q] = b[q] - off - diag * z[q];
}
template <class T>
__global__ void k_fem_restrict(LevelArgs f, LevelArgs cl, const T *r, T *bc, T *zc) {
    const int I = blockIdx.x * BX + threadIdx.x, J = blockIdx.y * BY + threadIdx.y;
    if (I >= cl.nx || J >= cl.ny) return;
    const int q = 2 * J * f.zp + 2 * I, p = f.zp;
    // ghost-ring entries of r are zero, so boundary children drop out
    const T s = r[q] + T(0.5) * (r[q - 1] + r[q + 1] + r[q - p] + r[q + p])
                + T(0.25) * (r[q - p - 1] + r[q - p + 1] + r[q + p - 1] + r[q + p + 1]);
    bc[J * cl.zp + I] = s;
    put_z(cl, zc, I, J, T(0));
}

// MAC: bilinear cell-centred prolongation with constant extrapolation at
// the walls, added to the fine iterate.
template <class T>
__global__ void k_mac_prolong(LevelArgs f, LevelArgs cl, T *z, const T *e) {
    const int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y;
    if (i >= f.nx || j >= f.ny) return;
    const ProArgs pa{e, cl.zp, cl.nx, cl.ny, cl.px, cl.py, cl.dr};
    put_z(f, z, i, j, z[j * f.zp + i] + prolong_at<T, kMAC>(pa, i, j));
}

// FEM: bilinear vertex prolongation, added to the fine iterate.
template <class T>
__global__ void k_fem_prolong(LevelArgs f, LevelArgs cl, T *z, const T *e) {
    const int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y;
    if (i >= f.nx || j >= f.ny) return;
    const ProArgs pa{e, cl.zp, cl.nx, cl.ny, cl.px, cl.py, cl.dr};
    put_z(f, z, i, j, z[j * f.zp + i] + prolong_at<T, kFEM>(pa, i, j));
}

// FEM residual and full-weighting restriction in one pass: each coarse
// node forms the nine fine residuals it gathers (the same expressions as
// k_fem_residual, summed as k_fem_restrict sums them). Residuals outside
// the level are the zero ghost ring's.
template <class T>
__global__ void k_fem_restrict_fused(LevelArgs f, LevelArgs cl, FemC k, const T *z, const T *b, const T *kc, T *bc,
                                     T *zc) {
    const int I = blockIdx.x * BX + threadIdx.x, J = blockIdx.y * BY + threadIdx.y;
    if (I >= cl.nx || J >= cl.ny) return;
    auto res = [&](int i, int j) -> T {
        // children across a periodic seam are the opposite edge's nodes
        if (f.px) i = wrap_idx(i, f.nx);
        if (f.py) j = wrap_idx(j, f.ny);
        if (i < 0 || i >= f.nx || j < 0 || j >= f.ny) return T(0);
        const int q = j * f.zp + i, c = j * f.cp + i;
        T diag;
        const T off = fem_offdiag<T>(z, kc, q, c, f.zp, f.cp, k, diag);
        return b[q] - off - diag * z[q];
    };
    const int i = 2 * I, j = 2 * J;
    const T s = res(i, j) + T(0.5) * (res(i - 1, j) + res(i + 1, j) + res(i, j - 1) + res(i, j + 1))
                + T(0.25) * (res(i - 1, j - 1) + res(i + 1, j - 1) + res(i - 1, j + 1) + res(i + 1, j + 1));
    bc[J * cl.zp + I] = s;
    put_z(cl, zc, I, J, T(0));
}

// FEM transfers to and from a uniform coarse level that is not nested in
// the fine one (fine node counts other than 2^k+1): tensor-product linear
// interpolation by node position, and its transpose as the restriction.
struct RsArgs {
    const int *i0, *lo, *hi;
    const double *t;
};
template <class T>
__global__ void k_fem_restrict_rs(LevelArgs f, LevelArgs cl, const T *r, T *bc, T *zc, RsArgs x, RsArgs y) {
    const int I = blockIdx.x * BX + threadIdx.x, J = blockIdx.y * BY + threadIdx.y;
    if (I >= cl.nx || J >= cl.ny) return;
    T s = 0;
    for (int j = y.lo[J]; j <= y.hi[J]; j++) {
        const T wy = y.i0[j] == J ? T(1 - y.t[j]) : (y.i0[j] == J - 1 ? T(y.t[j]) : T(0));
        T sr = 0;
        for (int i = x.lo[I]; i <= x.hi[I]; i++) {
            const T wx = x.i0[i] == I ? T(1 - x.t[i]) : (x.i0[i] == I - 1 ? T(x.t[i]) : T(0));
            sr += wx * r[j * f.zp + i];
        }
        s += wy * sr;
    }
    bc[J * cl.zp + I] = s;
    zc[J * cl.zp + I] = 0;
}
template <class T>
__global__ void k_fem_prolong_rs(LevelArgs f, LevelArgs cl, T *z, const T *e, RsArgs x, RsArgs y) {
    const int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y;
    if (i >= f.nx || j >= f.ny) return;
    const int I = x.i0[i], J = y.i0[j], p = cl.zp;
    const T tx = T(x.t[i]), ty = T(y.t[j]);
    const T lo = (T(1) - tx) * e[J * p + I] + tx * e[J * p + I + 1];
    const T up = (T(1) - tx) * e[(J + 1) * p + I] + tx * e[(J + 1) * p + I + 1];
    z[j * f.zp + i] += (T(1) - ty) * lo + ty * up;
}
// Transfers between a doubly periodic level and its resampling: nf fine and
// nc coarse points (cell centres or nodes) spread evenly over one period.
// Fine point i sits at coarse coordinate i0[i] + t[i] (i0 wrapped into
// [0, nc)): prolongation is linear interpolation between coarse points i0
// and i0 + 1 (wrapped), restriction its transpose, gathered per coarse
// point over the fine window lo..hi (indices wrapped), times `scale`
// (FEM: 1, the transpose itself; MAC: nc/nf per direction, an average).
template <class T>
__global__ void k_restrict_rs_per(LevelArgs f, LevelArgs cl, const T *r, T *bc, T *zc, RsArgs x, RsArgs y, T scale) {
    const int I = blockIdx.x * BX + threadIdx.x, J = blockIdx.y * BY + threadIdx.y;
    if (I >= cl.nx || J >= cl.ny) return;
    auto w = [](const RsArgs &d, int i, int I, int nc) -> T {
        const int a = d.i0[i];
        return a == I ? T(1 - d.t[i]) : ((a + 1 == nc ? 0 : a + 1) == I ? T(d.t[i]) : T(0));
    };
    T s = 0;
    for (int jj = y.lo[J]; jj <= y.hi[J]; jj++) {
        const int j = wrap_idx(jj, f.ny);
        const T wy = w(y, j, J, cl.ny);
        if (wy == T(0)) continue;
        T sr = 0;
        for (int ii = x.lo[I]; ii <= x.hi[I]; ii++) {
            const int i = wrap_idx(ii, f.nx);
            sr += w(x, i, I, cl.nx) * r[j * f.zp + i];
        }
        s += wy * sr;
    }
    bc[J * cl.zp + I] = scale * s;
    put_z(cl, zc, I, J, T(0));
}
template <class T>
__global__ void k_prolong_rs_per(LevelArgs f, LevelArgs cl, T *z, const T *e, RsArgs x, RsArgs y) {
    const int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y;
    if (i >= f.nx || j >= f.ny) return;
    const int I = x.i0[i], J = y.i0[j], p = cl.zp;
    const int I1 = I + 1 == cl.nx ? 0 : I + 1, J1 = J + 1 == cl.ny ? 0 : J + 1;
    const T tx = T(x.t[i]), ty = T(y.t[j]);
    const T lo = (T(1) - tx) * e[J * p + I] + tx * e[J * p + I1];
    const T up = (T(1) - tx) * e[J1 * p + I] + tx * e[J1 * p + I1];
    put_z(f, z, i, j, z[j * f.zp + i] + ((T(1) - ty) * lo + ty * up));
}
// MAC residual r = b - A z of one level
template <class T>
__global__ void k_mac_residual(LevelArgs a, const T *z, const T *b, const T *fw, const T *fs, T *r) {
    const int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y;
    if (i >= a.nx || j >= a.ny) return;
    const int q = j * a.zp + i, c = j * a.cp + i;
    T diag;
    const T off = mac_offdiag<T>(z, fw, fs, q, c, a.zp, a.cp, T(a.kx), T(a.ky), diag);
    r[q] = b[q] - off - diag * z[q];
}
// Overlap of [a0, a1) and [b0, b1)
__device__ inline double overlap(double a0, double a1, double b0, double b1) {
    const double v = fmin(a1, b1) - fmax(a0, b0);
    return v > 0 ? v : 0;
}
// Resampled coefficients of a doubly periodic level, in units of one
// period per direction (fine spacing 1/f.nx, coarse 1/c.nx). FEM: a coarse
// cell's weight is its fine cells' average weighted by overlap. MAC: a
// coarse face's inverse density is the fine faces' value interpolated
// linearly along the face normal and averaged with overlap along the face.
__global__ void k_fem_coef_rs_per(LevelArgs f, LevelArgs c, const double *kf, double *kcc) {
    const int I = blockIdx.x * BX + threadIdx.x, J = blockIdx.y * BY + threadIdx.y;
    if (I >= c.nx || J >= c.ny) return;
    const double hfx = 1.0 / f.nx, hfy = 1.0 / f.ny, hcx = 1.0 / c.nx, hcy = 1.0 / c.ny;
    const int i0 = (int)(I * hcx / hfx), i1 = min(f.nx - 1, (int)((I + 1) * hcx / hfx));
    const int j0 = (int)(J * hcy / hfy), j1 = min(f.ny - 1, (int)((J + 1) * hcy / hfy));
    double s = 0;
    for (int j = j0; j <= j1; j++) {
        const double oy = overlap(j * hfy, (j + 1) * hfy, J * hcy, (J + 1) * hcy);
        for (int i = i0; i <= i1; i++) s += oy * overlap(i * hfx, (i + 1) * hfx, I * hcx, (I + 1) * hcx) * kf[j * f.cp + i];
    }
    kcc[J * c.cp + I] = s / (hcx * hcy);
}
__global__ void k_mac_coef_rs_per(LevelArgs f, LevelArgs c, const double *fwf, const double *fsf, double *fwc,
                                  double *fsc) {
    const int I = blockIdx.x * BX + threadIdx.x, J = blockIdx.y * BY + threadIdx.y;
    if (I > c.nx || J > c.ny) return;
    const double hfx = 1.0 / f.nx, hfy = 1.0 / f.ny, hcx = 1.0 / c.nx, hcy = 1.0 / c.ny;
    if (J < c.ny) {  // x-face I of row J (faces 0..nx, face nx the seam's copy of face 0)
        const double u = I * hcx / hfx;
        int a = (int)u;
        if (a > f.nx) a = f.nx;
        const double t = u - a;
        const int b = a < f.nx ? a + 1 : a;
        const int j0 = (int)(J * hcy / hfy), j1 = min(f.ny - 1, (int)((J + 1) * hcy / hfy));
        double s = 0;
        for (int j = j0; j <= j1; j++)
            s += overlap(j * hfy, (j + 1) * hfy, J * hcy, (J + 1) * hcy) *
                 ((1 - t) * fwf[j * f.cp + a] + t * fwf[j * f.cp + b]);
        fwc[J * c.cp + I] = s / hcy;
    }
    if (I < c.nx) {  // y-face J of column I
        const double u = J * hcy / hfy;
        int a = (int)u;
        if (a > f.ny) a = f.ny;
        const double t = u - a;
        const int b = a < f.ny ? a + 1 : a;
        const int i0 = (int)(I * hcx / hfx), i1 = min(f.nx - 1, (int)((I + 1) * hcx / hfx));
        double s = 0;
        for (int i = i0; i <= i1; i++)
            s += overlap(i * hfx, (i + 1) * hfx, I * hcx, (I + 1) * hcx) *
                 ((1 - t) * fsf[a * f.cp + i] + t * fsf[b * f.cp + i]);
        fsc[J * c.cp + I] = s / hcx;
    }
}

// coarse cell coefficient: fine cell values averaged with their overlap
__global__ void k_fem_coef_rs(LevelArgs f, LevelArgs c, const double *kf, double *kcc, int *cxlo, int *cxhi,
                              int *cylo, int *cyhi, double hfx, double hcx, double hfy, double hcy) {
    const int I = blockIdx.x * BX + threadIdx.x, J = blockIdx.y * BY + threadIdx.y;
    if (I >= c.nx - 1 || J >= c.ny - 1) return;
    double s = 0;
    for (int j = cylo[J]; j <= cyhi[J]; j++) {
        const double oy = fmin((j + 1) * hfy, (J + 1) * hcy) - fmax(j * hfy, J * hcy);
        if (!(oy > 0)) continue;
        for (int i = cxlo[I]; i <= cxhi[I]; i++) {
            const double ox = fmin((i + 1) * hfx, (I + 1) * hcx) - fmax(i * hfx, I * hcx);
            if (ox > 0) s += ox * oy * kf[j * f.cp + i];
        }
    }
    kcc[J * c.cp + I] = s / (hcx * hcy);
}

// Coarsest level: one workgroup, mean-free right-hand side, colored
// Gauss-Seidel sweeps until the residual has dropped by `reduce` (checked
// every 8 sweeps, at most max_sweeps), mean-free result (the systems are
// pure Neumann).
template <class T>
__device__ double block_sum(double v, double *red) {
    const int tid = threadIdx.x;
    red[tid] = v;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (tid < w) red[tid] += red[tid + w];
        __syncthreads();
    }
    const double s = red[0];
    __syncthreads();
    return s;
}
template <class T>
__global__ void k_coarse_solve(int kind, LevelArgs a, FemC k, T *z, T *b, const T *fw, const T *fs, const T *kc,
                               double reduce, int max_sweeps, int sing) {
    __shared__ double red[256];
    const int tid = threadIdx.x, N = a.nx * a.ny;
    auto Q = [&](int p) { return (p / a.nx) * a.zp + p % a.nx; };
    double s = 0;
    for (int p = tid; p < N; p += blockDim.x) s += b[Q(p)];
    // a definite system (zero-pressure edge) keeps its right-hand side and
    // solution as they are
    const T mb = sing ? T(block_sum<T>(s, red) / N) : T(0);
    // z is the correction the restriction zeroed, or, when the finest level
    // is also the coarsest, the current iterate: it is not reset here
    for (int p = tid; p < N; p += blockDim.x) b[Q(p)] -= mb;
    __syncthreads();
    auto resid2 = [&]() -> double {
        double r2 = 0;
        for (int p = tid; p < N; p += blockDim.x) {
            const int q = Q(p), c = (p / a.nx) * a.cp + p % a.nx;
            T diag, off;
            if (kind == kMAC) off = mac_offdiag<T>(z, fw, fs, q, c, a.zp, a.cp, T(a.kx), T(a.ky), diag);
            else off = fem_offdiag<T>(z, kc, q, c, a.zp, a.cp, k, diag);
            const double r = double(b[q] - off - diag * z[q]);
            r2 += r * r;
        }
        return block_sum<T>(r2, red);
    };
    const double r0 = resid2();
    const int ncol = kind == kMAC ? 2 : 4;
    for (int it = 0; it < max_sweeps && r0 > 0; it++) {
        for (int col = 0; col < ncol; col++) {
            for (int p = tid; p < N; p += blockDim.x) {
                const int i = p % a.nx, j = p / a.nx;
                const bool mine = kind == kMAC ? (((i + j) & 1) == col) : ((i & 1) + 2 * (j & 1) == col);
                if (!mine) continue;
                const int q = j * a.zp + i, c = j * a.cp + i;
                T diag, off;
                if (kind == kMAC) off = mac_offdiag<T>(z, fw, fs, q, c, a.zp, a.cp, T(a.kx), T(a.ky), diag);
                else off = fem_offdiag<T>(z, kc, q, c, a.zp, a.cp, k, diag);
                if (diag != T(0)) z[q] = (b[q] - off) / diag;
            }
            __syncthreads();
        }
        if ((it & 7) == 7 && resid2() <= reduce * reduce * r0) break;
    }
    s = 0;
    for (int p = tid; p < N; p += blockDim.x) s += z[Q(p)];
    const T mz = sing ? T(block_sum<T>(s, red) / N) : T(0);
    for (int p = tid; p < N; p += blockDim.x) z[Q(p)] -= mz;
}

// ------------------------------------------------------------ tail cycle
// The coarse end of a V-cycle (every level from `top` down, when they fit in
// LDS together) runs in ONE workgroup: the levels' iterates, right-hand
// sides, residuals and coefficients live in LDS and each smoothing colour,
// transfer and coarse sweep is a barrier, not a kernel launch. The levels of
// at most kTailWavePts points (the bottom of the cycle) run in the first
// wavefront alone, whose lanes synchronise without a workgroup barrier, and
// a coarsest level of at most kTailDirect points is solved exactly by dense
// elimination instead of by sweeps.
constexpr int kTailMaxLev = 12;
constexpr int kTailThreads = 1024;
constexpr int kTailArrays = 4;    // z, b, coef1, and coef2 (MAC) or r (FEM)
constexpr int kTailWavePts = 128;
constexpr int kTailDirect = 16;

struct TailLev {
    int nx, ny, zp, off;   // LDS pitch (nx+2) and offset of the padded array
    int gcp;               // global pitch
    unsigned mag;          // ceil(2^20 / ((nx+1)/2)): k -> k / ((nx+1)/2) as (k * mag) >> 20
    const void *c1, *c2;   // global coefficient arrays at (0,0): fw,fs (MAC) or kc (FEM)
    double kx, ky;
};
struct TailArgs {
    int kind, nlev, nu1, nu2, total;
    int wave_from;         // first level run by the first wavefront alone (nlev: none)
    int px, py;            // periodic directions (the LDS iterates keep mirrored rings)
    int dr;                // zero-pressure right edge (MAC prolongation mirrors with a sign)
    int sing;              // pure-Neumann system: mean-free coarsest level, one point pinned
    FemC k;
    TailLev lv[kTailMaxLev];
    void *z_top;
    const void *b_top;
    int top_pitch;
    double reduce;
    int max_sweeps;
};

template <class T>
__device__ inline double tail_wave_sum(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
template <class T>
__device__ double tail_block_sum(double v, double *red) {
    v = tail_wave_sum<T>(v);
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63, nw = blockDim.x >> 6;
    if (lane == 0) red[w] = v;
    __syncthreads();
    double s = 0;
    for (int k = 0; k < nw; k++) s += red[k];
    __syncthreads();
    return s;
}

// LDS writes of a wavefront's lanes visible to its other lanes
__device__ inline void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <class T, bool PER>
__global__ __launch_bounds__(kTailThreads) void k_tail(TailArgs A) {
    const int px = PER ? A.px : 0, py = PER ? A.py : 0;
    extern __shared__ __align__(16) unsigned char smem[];
    T *pool = (T *)smem;
    __shared__ double red[16];
    __shared__ T dm[kTailDirect * kTailDirect];  // reduced system of the direct coarsest solve
    const int tid = threadIdx.x, nt = blockDim.x, lane = tid & 63;
    auto arr = [&](int a, int l) -> T * { return pool + a * A.total + A.lv[l].off + A.lv[l].zp + 1; };
    // a level's LDS arrays as a LevelArgs, for put_z (mirrored periodic rings)
    auto lds = [&](int l) -> LevelArgs {
        const TailLev &v = A.lv[l];
        return LevelArgs{v.nx, v.ny, v.zp, v.zp, v.kx, v.ky, px, py};
    };
    // an iterate write, with the periodic mirrors when the levels have them
    auto tput = [&](const LevelArgs &la, T *arr_, int i, int j, T v) {
        if (PER) put_z(la, arr_, i, j, v);
        else arr_[j * la.zp + i] = v;
    };
    for (int p = tid; p < kTailArrays * A.total; p += nt) pool[p] = 0;
    __syncthreads();
    // coefficients with their zero ring (all levels as one flat range, two
    // points per thread in flight), and the top level's z and b
    for (int base = 0; base < A.total; base += 2 * nt) {
        T v1[2], v2[2];
        int dst[2];
#pragma unroll
        for (int u = 0; u < 2; u++) {
            const int p = base + u * nt + tid;
            dst[u] = -1;
            v1[u] = v2[u] = 0;
            if (p < A.total) {
                int l = 0;
                while (l + 1 < A.nlev && p >= A.lv[l + 1].off) l++;
                const TailLev &v = A.lv[l];
                const int q = p - v.off, W = v.nx + 2, jj = q / W, ii = q - jj * W;
                const int g = (jj - 1) * v.gcp + (ii - 1);
                v1[u] = ((const T *)v.c1)[g];
                if (A.kind == kMAC) v2[u] = ((const T *)v.c2)[g];
                dst[u] = p;
            }
        }
#pragma unroll
        for (int u = 0; u < 2; u++)
            if (dst[u] >= 0) {
                pool[2 * A.total + dst[u]] = v1[u];
                if (A.kind == kMAC) pool[3 * A.total + dst[u]] = v2[u];
            }
    }
    {
        const TailLev &v = A.lv[0];
        const T *gb = (const T *)A.b_top, *gz = (const T *)A.z_top;
        for (int p = tid; p < v.nx * v.ny; p += nt) {
            const int i = p % v.nx, j = p / v.nx;
            const T bb = gb[j * A.top_pitch + i], zz = gz[j * A.top_pitch + i];
            arr(1, 0)[j * v.zp + i] = bb;
            tput(lds(0), arr(0, 0), i, j, zz);
        }
    }
    __syncthreads();
    // w: the first wavefront alone (lanes 0..63, wave_sync) instead of the
    // whole workgroup (__syncthreads)
    auto sync = [&](bool w) {
        if (w) wave_sync();
        else __syncthreads();
    };
    // A level's parameters and LDS arrays, read from the kernel arguments
    // once per phase: a level index known only at run time keeps the
    // compiler from hoisting those reads out of the conditional point
    // updates, where each would wait on a scalar load.
    struct Lv {
        TailLev v;
        T *z, *b, *c1, *c2;
        LevelArgs la;
    };
    auto lvl = [&](int l) -> Lv {
        Lv L;
        L.v = A.lv[l];
        L.z = arr(0, l); L.b = arr(1, l); L.c1 = arr(2, l); L.c2 = arr(3, l);
        L.la = LevelArgs{L.v.nx, L.v.ny, L.v.zp, L.v.zp, L.v.kx, L.v.ky, px, py};
        return L;
    };
    const int kind = A.kind;
    const FemC kk = A.k;
    auto offd = [&](const Lv &L, int i, int j, T &diag) -> T {
        const int q = j * L.v.zp + i;
        if (kind == kMAC) return mac_offdiag<T>(L.z, L.c1, L.c2, q, q, L.v.zp, L.v.zp, T(L.v.kx), T(L.v.ky), diag);
        return fem_offdiag<T>(L.z, L.c1, q, q, L.v.zp, L.v.zp, kk, diag);
    };
    // colour phases enumerate only their own points: the k-th point of a
    // colour is in colour row r = k / ha, column a = k - r ha
    auto smooth = [&](int l, int sweeps, bool reverse, bool w) {
        const Lv L = lvl(l);
        const TailLev &v = L.v;
        const int ncol = kind == kMAC ? 2 : 4, ha = (v.nx + 1) >> 1;
        const int N = ha * (kind == kMAC ? v.ny : (v.ny + 1) >> 1);
        const int t0 = w ? lane : tid, st = w ? 64 : nt;
        for (int sw = 0; sw < sweeps; sw++)
            for (int cc = 0; cc < ncol; cc++) {
                const int col = reverse ? ncol - 1 - cc : cc;
                for (int k = t0; k < N; k += st) {
                    const int r = (int)(((unsigned)k * v.mag) >> 20), a = k - r * ha;
                    const int j = kind == kMAC ? r : 2 * r + (col >> 1);
                    const int i = kind == kMAC ? 2 * a + ((col + j) & 1) : 2 * a + (col & 1);
                    if (i >= v.nx || j >= v.ny) continue;
                    T diag;
                    const T off = offd(L, i, j, diag);
                    if (diag != T(0)) tput(L.la, L.z, i, j, (L.b[j * v.zp + i] - off) / diag);
                }
                sync(w);
            }
    };
    auto residual = [&](const Lv &L, int i, int j) -> T {
        T diag;
        const T off = offd(L, i, j, diag);
        return L.b[j * L.v.zp + i] - off - diag * L.z[j * L.v.zp + i];
    };
    // pre-smooth level l, restrict its residual to l+1's right-hand side and
    // clear l+1's iterate
    auto down = [&](int l, bool w) {
        smooth(l, A.nu1, false, w);
        const Lv F = lvl(l), C = lvl(l + 1);
        const TailLev &f = F.v, &c = C.v;
        const int t0 = w ? lane : tid, st = w ? 64 : nt;
        if (kind == kMAC) {
            for (int p = t0; p < c.nx * c.ny; p += st) {
                const int I = p % c.nx, J = p / c.nx;
                T s = 0;
                for (int dj = 0; dj < 2; dj++)
                    for (int di = 0; di < 2; di++) s += residual(F, 2 * I + di, 2 * J + dj);
                C.b[J * c.zp + I] = T(0.25) * s;
                tput(C.la, C.z, I, J, T(0));
            }
        } else {
            for (int p = t0; p < f.nx * f.ny; p += st) {
                const int i = p % f.nx, j = p / f.nx;
                tput(F.la, F.c2, i, j, residual(F, i, j));
            }
            sync(w);
            const T *r = F.c2;
            const int pz = f.zp;
            for (int p = t0; p < c.nx * c.ny; p += st) {
                const int I = p % c.nx, J = p / c.nx, q = 2 * J * pz + 2 * I;
                C.b[J * c.zp + I] = r[q] + T(0.5) * (r[q - 1] + r[q + 1] + r[q - pz] + r[q + pz])
                                    + T(0.25) * (r[q - pz - 1] + r[q - pz + 1] + r[q + pz - 1] + r[q + pz + 1]);
                tput(C.la, C.z, I, J, T(0));
            }
        }
        sync(w);
    };
    // add level l+1's correction to level l and post-smooth level l
    auto up = [&](int l, bool w) {
        const Lv F = lvl(l), C = lvl(l + 1);
        const TailLev &f = F.v, &c = C.v;
        const T *e = C.z;
        const int p2 = c.zp, t0 = w ? lane : tid, st = w ? 64 : nt;
        const bool dr = A.dr;
        for (int p = t0; p < f.nx * f.ny; p += st) {
            const int i = p % f.nx, j = p / f.nx, I = i >> 1, J = j >> 1;
            T add;
            if (kind == kMAC) {
                int I2 = I + ((i & 1) ? 1 : -1), J2 = J + ((j & 1) ? 1 : -1);
                const bool flip = dr && I2 >= c.nx;
                if (I2 < 0 || I2 >= c.nx) I2 = px ? wrap_idx(I2, c.nx) : I;
                if (J2 < 0 || J2 >= c.ny) J2 = py ? wrap_idx(J2, c.ny) : J;
                T e2 = e[J * p2 + I2], e22 = e[J2 * p2 + I2];
                if (flip) { e2 = -e2; e22 = -e22; }
                add = T(0.5625) * e[J * p2 + I] + T(0.1875) * e2 + T(0.1875) * e[J2 * p2 + I] + T(0.0625) * e22;
            } else {
                const int q = J * p2 + I;
                if (!(i & 1) && !(j & 1)) add = e[q];
                else if (!(j & 1)) add = T(0.5) * (e[q] + e[q + 1]);
                else if (!(i & 1)) add = T(0.5) * (e[q] + e[q + p2]);
                else add = T(0.25) * (e[q] + e[q + 1] + e[q + p2] + e[q + p2 + 1]);
            }
            tput(F.la, F.z, i, j, F.z[j * f.zp + i] + add);
        }
        sync(w);
        smooth(l, A.nu2, true, w);
    };
    // Exact solve of a coarsest level of N <= kTailDirect points by the first
    // wavefront. A pure-Neumann level has a mean-free, hence compatible,
    // right-hand side and a symmetric operator with the constants as its null
    // space, so fixing point 0 at zero leaves a definite system in points
    // 1..N-1 (o = 1); a definite level keeps all N points (o = 0). Eliminated
    // without pivoting (row r of dm is point r+o). False (nothing written)
    // when a pivot is zero, not finite or of the other sign.
    auto direct = [&](int l) -> bool {
        const Lv L = lvl(l);
        const TailLev &v = L.v;
        const int o = A.sing ? 1 : 0, N = v.nx * v.ny, n = N - o;
        if (lane < n) {
            const int p = lane + o, i = p % v.nx, j = p / v.nx, q = j * v.zp + i;
            T *row = dm + lane * kTailDirect;
            for (int c = 0; c < kTailDirect; c++) row[c] = 0;
            auto put = [&](int di, int dj, T a) {
                int ii = i + di, jj = j + dj;
                if (px) ii = wrap_idx(ii, v.nx);
                if (py) jj = wrap_idx(jj, v.ny);
                if (ii < 0 || ii >= v.nx || jj < 0 || jj >= v.ny) return;
                const int pn = jj * v.nx + ii;
                if (pn >= o) row[pn - o] += a;
            };
            if (kind == kMAC) {
                const T *fw = L.c1, *fs = L.c2, kx = T(v.kx), ky = T(v.ky);
                const T wl = fw[q], wr = fw[q + 1], sd = fs[q], su = fs[q + v.zp];
                put(0, 0, -(kx * (wl + wr) + ky * (sd + su)));
                put(-1, 0, kx * wl);
                put(1, 0, kx * wr);
                put(0, -1, ky * sd);
                put(0, 1, ky * su);
            } else {
                const T *kc = L.c1;
                const T kdl = kc[q - v.zp - 1], kdr = kc[q - v.zp], kul = kc[q - 1], kur = kc[q];
                put(0, 0, T(kk.fm) * (kdl + kdr + kul + kur));
                put(-1, -1, T(kk.fc) * kdl);
                put(1, -1, T(kk.fc) * kdr);
                put(-1, 1, T(kk.fc) * kul);
                put(1, 1, T(kk.fc) * kur);
                put(0, -1, T(kk.hey) * (kdl + kdr));
                put(0, 1, T(kk.hey) * (kul + kur));
                put(-1, 0, T(kk.hex) * (kdl + kul));
                put(1, 0, T(kk.hex) * (kdr + kur));
            }
        }
        // lane r keeps row r's right-hand side in a register
        T rhs = 0;
        if (lane < n) {
            const int p = lane + o;
            rhs = L.b[(p / v.nx) * v.zp + p % v.nx];
        }
        wave_sync();
        bool ok = true;
        T s0 = 0;
        for (int k = 0; k < n; k++) {
            const T piv = dm[k * kTailDirect + k];
            if (k == 0) s0 = piv;
            if (!(piv == piv) || piv == T(0) || (piv > T(0)) != (s0 > T(0)) || !(piv * T(0) == T(0))) {
                ok = false;
                break;
            }
            const T rk = __shfl(rhs, k, 64);
            if (lane > k && lane < n) {
                const T f = dm[lane * kTailDirect + k] / piv;
                for (int c = k + 1; c < n; c++) dm[lane * kTailDirect + c] -= f * dm[k * kTailDirect + c];
                rhs -= f * rk;
            }
            wave_sync();
        }
        if (!ok) return false;
        // back substitution, one unknown at a time from the last
        T x = 0;
        for (int k = n - 1; k >= 0; k--) {
            // lane k owns row k; the unknowns after it are already in x of their lanes
            T acc = 0;
            for (int c = k + 1; c < n; c++) acc += dm[k * kTailDirect + c] * __shfl(x, c, 64);
            const T xk = (__shfl(rhs, k, 64) - acc) / dm[k * kTailDirect + k];
            if (lane == k) x = xk;
        }
        if (lane < n) {
            const int p = lane + o;
            tput(L.la, L.z, p % v.nx, p / v.nx, x);
        }
        // the pinned point 0 (a warm-started single-level tail holds an
        // iterate there)
        if (o && lane == 0) tput(L.la, L.z, 0, 0, T(0));
        wave_sync();
        return true;
    };
    // coarsest: mean-free right-hand side, then the direct solve or sweeps to
    // a residual reduction, then a mean-free iterate (pure Neumann; a
    // definite level skips both projections)
    auto coarsest = [&](bool w) {
        const int l = A.nlev - 1;
        const Lv L = lvl(l);
        const TailLev &v = L.v;
        const int N = v.nx * v.ny, t0 = w ? lane : tid, st = w ? 64 : nt;
        auto total = [&](double a) -> double { return w ? tail_wave_sum<T>(a) : tail_block_sum<T>(a, red); };
        double sb = 0;
        for (int p = t0; p < N; p += st) sb += double(L.b[(p / v.nx) * v.zp + p % v.nx]);
        const T mb = A.sing ? T(total(sb) / N) : T(0);
        for (int p = t0; p < N; p += st) L.b[(p / v.nx) * v.zp + p % v.nx] -= mb;
        sync(w);
        const bool solved = w && N <= kTailDirect && N > 1 && direct(l);
        if (!solved) {
            auto r2 = [&]() -> double {
                double a = 0;
                for (int p = t0; p < N; p += st) {
                    const double r = double(residual(L, p % v.nx, p / v.nx));
                    a += r * r;
                }
                return total(a);
            };
            const double r0 = r2();
            for (int it = 0; it < A.max_sweeps && r0 > 0; it++) {
                smooth(l, 1, false, w);
                if ((it & 7) == 7 && r2() <= A.reduce * A.reduce * r0) break;
            }
        }
        double sz = 0;
        for (int p = t0; p < N; p += st) sz += double(L.z[(p / v.nx) * v.zp + p % v.nx]);
        const T mz = A.sing ? T(total(sz) / N) : T(0);
        // (total() synchronises: every lane has read z before it changes)
        for (int p = t0; p < N; p += st) {
            const int i = p % v.nx, j = p / v.nx;
            tput(L.la, L.z, i, j, L.z[j * v.zp + i] - mz);
        }
        sync(w);
    };
    const int last = A.nlev - 1, ws = A.wave_from <= last ? A.wave_from : last + 1;
    for (int l = 0; l < ws && l < last; l++) down(l, false);
    if (ws <= last) {
        if (tid < 64) {
            for (int l = ws; l < last; l++) down(l, true);
            coarsest(true);
            for (int l = last - 1; l >= ws; l--) up(l, true);
        }
        __syncthreads();
    } else {
        coarsest(false);
    }
    for (int l = (ws < last ? ws : last) - 1; l >= 0; l--) up(l, false);
    {
        // the top level's iterate back, with its periodic ring mirrors
        const TailLev &v = A.lv[0];
        T *gz = (T *)A.z_top;
        const LevelArgs ga{v.nx, v.ny, A.top_pitch, A.top_pitch, v.kx, v.ky, px, py};
        for (int p = tid; p < v.nx * v.ny; p += nt) {
            const int i = p % v.nx, j = p / v.nx;
            tput(ga, gz, i, j, arr(0, 0)[j * v.zp + i]);
        }
    }
}

// ------------------------------------------------------ coefficient setup
// MAC fine faces from cell densities: 0.5/rho_c + 0.5/rho_nb
// (fluid_2d.cc:1140-1160), zeroed on the walls. Across a periodic seam the
// two sides are the reference's left-edge and right-edge values: face 0
// from cells 0 and -1, face nx from cells nx-1 and nx (ghost densities).
// fwf, fsf (kcf below): single-precision copies written alongside, or null
__global__ void k_mac_coef_fine(Grid g, const double *rho, LevelArgs a, double *fw, double *fs, float *fwf,
                                float *fsf) {
    const int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y;
    if (i > a.nx || j > a.ny) return;
    const int c = j * a.cp + i;
    if (j < a.ny) {
        const int ij = g.at(i, j);
        // a zero-pressure right edge doubles its face (mgs_mac.hh:95)
        const double v = (!a.px && (i == 0 || i == a.nx))
                             ? (a.dr && i == a.nx ? 2 * (0.5 / rho[ij] + 0.5 / rho[ij - 1]) : 0.0)
                             : 0.5 / rho[ij] + 0.5 / rho[ij - 1];
        fw[c] = v;
        if (fwf) fwf[c] = float(v);
    }
    if (i < a.nx) {
        const int ij = g.at(i, j);
        const double v = (!a.py && (j == 0 || j == a.ny)) ? 0.0 : 0.5 / rho[ij] + 0.5 / rho[ij - g.P];
        fs[c] = v;
        if (fsf) fsf[c] = float(v);
    }
}
// FEM fine cell weights 1/rho (fluid_2d.cc:1163-1170). A periodic
// direction has nx cells; the seam cell's weight is the ghost cell -1's,
// as mgs_fem_varying_rho wraps its corner lookup (mgs_fem.hh:99-106). A
// zero-pressure edge also has nx cells, the last one between the last
// unknown node and the fixed column.
__global__ void k_fem_coef_fine(Grid g, const double *rho, LevelArgs a, double *kc, float *kcf) {
    const int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y;
    if (i >= a.nx - ((a.px || a.dr) ? 0 : 1) || j >= a.ny - (a.py ? 0 : 1)) return;
    const int ci = a.px && i == a.nx - 1 ? -1 : i, cj = a.py && j == a.ny - 1 ? -1 : j;
    const double v = 1. / rho[g.at(ci, cj)];
    kc[j * a.cp + i] = v;
    if (kcf) kcf[j * a.cp + i] = float(v);
}
__global__ void k_mac_coef_coarsen(LevelArgs f, LevelArgs c, const double *fwf, const double *fsf, double *fwc,
                                   double *fsc, float *fwc32, float *fsc32) {
    const int I = blockIdx.x * BX + threadIdx.x, J = blockIdx.y * BY + threadIdx.y;
    if (I > c.nx || J > c.ny) return;
    if (J < c.ny) {
        const double v = 0.5 * (fwf[2 * J * f.cp + 2 * I] + fwf[(2 * J + 1) * f.cp + 2 * I]);
        fwc[J * c.cp + I] = v;
        if (fwc32) fwc32[J * c.cp + I] = float(v);
    }
    if (I < c.nx) {
        const double v = 0.5 * (fsf[2 * J * f.cp + 2 * I] + fsf[2 * J * f.cp + 2 * I + 1]);
        fsc[J * c.cp + I] = v;
        if (fsc32) fsc32[J * c.cp + I] = float(v);
    }
}
__global__ void k_fem_coef_coarsen(LevelArgs f, LevelArgs c, const double *kf, double *kcc, float *kcc32) {
    const int I = blockIdx.x * BX + threadIdx.x, J = blockIdx.y * BY + threadIdx.y;
    if (I >= c.nx - ((c.px || c.dr) ? 0 : 1) || J >= c.ny - (c.py ? 0 : 1)) return;
    const int q = 2 * J * f.cp + 2 * I;
    const double v = 0.25 * (kf[q] + kf[q + 1] + kf[q + f.cp] + kf[q + f.cp + 1]);
    kcc[J * c.cp + I] = v;
    if (kcc32) kcc32[J * c.cp + I] = float(v);
}
__global__ void k_to_f64(const float *a, double *b, long long n) {
    long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) b[k] = double(a[k]);
}
__global__ void k_to_f32(const double *a, float *b, long long n) {
    long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) b[k] = float(a[k]);
}

__global__ void k_apply(int kind, LevelArgs a, FemC k, const double *x, const double *fw, const double *fs,
                        const double *kc, double *y) {
    const int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y;
    if (i >= a.nx || j >= a.ny) return;
    const int q = j * a.zp + i, c = j * a.cp + i;
    double diag, off;
    if (kind == kMAC) off = mac_offdiag<double>(x, fw, fs, q, c, a.zp, a.cp, a.kx, a.ky, diag);
    else off = fem_offdiag<double>(x, kc, q, c, a.zp, a.cp, k, diag);
    y[q] = off + diag * x[q];
}

// ------------------------------------------------------------- reductions
// Partial sums per block of (b^2, r^2, b) over level 0 with r = b - A z.
__global__ void k_norms(int kind, LevelArgs a, FemC k, const double *z, const double *b, const double *fw,
                        const double *fs, const double *kc, double *part) {
    __shared__ double s0[kRedThreads], s1[kRedThreads], s2[kRedThreads];
    double bb = 0, rr = 0, bs = 0;
    const int N = a.nx * a.ny, S = gridDim.x * blockDim.x, di = S % a.nx, dj = S / a.nx;
    // point p = j nx + i, stepped by S without a division per point; the
    // unrolled iterations' loads are in flight together (the sums keep
    // their order)
    const int p0 = blockIdx.x * blockDim.x + threadIdx.x;
    int i = p0 % a.nx, j = p0 / a.nx;
#pragma unroll 4
    for (int p = p0; p < N; p += S) {
        const int q = j * a.zp + i, c = j * a.cp + i;
        double diag, off;
        if (kind == kMAC) off = mac_offdiag<double>(z, fw, fs, q, c, a.zp, a.cp, a.kx, a.ky, diag);
        else off = fem_offdiag<double>(z, kc, q, c, a.zp, a.cp, k, diag);
        const double bq = b[q], r = bq - off - diag * z[q];
        bb += bq * bq; rr += r * r; bs += bq;
        i += di;
        j += dj;
        if (i >= a.nx) { i -= a.nx; j++; }
    }
    s0[threadIdx.x] = bb; s1[threadIdx.x] = rr; s2[threadIdx.x] = bs;
    __syncthreads();
    for (int w = kRedThreads / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            s0[threadIdx.x] += s0[threadIdx.x + w];
            s1[threadIdx.x] += s1[threadIdx.x + w];
            s2[threadIdx.x] += s2[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[3 * blockIdx.x] = s0[0]; part[3 * blockIdx.x + 1] = s1[0]; part[3 * blockIdx.x + 2] = s2[0];
    }
}
// Fixed-order sum of the partials into out[0..2].
__global__ void k_sum_partials(const double *part, int nb, double *out) {
    __shared__ double s[3][kRedThreads];
    for (int k = 0; k < 3; k++) {
        double v = 0;
        for (int p = threadIdx.x; p < nb; p += blockDim.x) v += part[3 * p + k];
        s[k][threadIdx.x] = v;
    }
    __syncthreads();
    for (int w = kRedThreads / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w)
            for (int k = 0; k < 3; k++) s[k][threadIdx.x] += s[k][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int k = 0; k < 3; k++) out[k] = s[k][0];
}
__global__ void k_sub_mean(LevelArgs a, double *b, const double *sums) {
    const int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y;
    if (i >= a.nx || j >= a.ny) return;
    b[j * a.zp + i] -= sums[2] / (double(a.nx) * a.ny);
}
// Mixed precision: r = b - A z in FP64, handed to the FP32 cycle; z += e.
__global__ void k_residual_f32(int kind, LevelArgs a, FemC k, const double *z, const double *b, const double *fw,
                               const double *fs, const double *kc, float *bf, float *zf) {
    const int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y;
    if (i >= a.nx || j >= a.ny) return;
    const int q = j * a.zp + i, c = j * a.cp + i;
    double diag, off;
    if (kind == kMAC) off = mac_offdiag<double>(z, fw, fs, q, c, a.zp, a.cp, a.kx, a.ky, diag);
    else off = fem_offdiag<double>(z, kc, q, c, a.zp, a.cp, k, diag);
    bf[q] = float(b[q] - off - diag * z[q]);
    put_z(a, zf, i, j, 0.0f);
}
// ---- Conjugate gradients preconditioned by the FP32 V-cycle (mixed
// solves, flexible form). Vectors in FP64 on level 0: r the residual, c the
// V-cycle's correction for r, p the direction, q = A p. Scalars on the
// device (ks): the step's r.c, beta, alpha, and a flag for the first
// direction. Reductions are fixed-order partial sums (kRedBlocks blocks,
// then k_sum_partials), so a solve is deterministic.
enum { kKrRho = 0, kKrBeta, kKrAlpha, kKrFirst, kKrSums = 4 };  // sums: 3 doubles at ks + kKrSums

// block sum of three values into part[3 * blockIdx.x + k]
__device__ void block_sum3(double v0, double v1, double v2, double *part) {
    __shared__ double s0[kRedThreads], s1[kRedThreads], s2[kRedThreads];
    s0[threadIdx.x] = v0; s1[threadIdx.x] = v1; s2[threadIdx.x] = v2;
    __syncthreads();
    for (int w = kRedThreads / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            s0[threadIdx.x] += s0[threadIdx.x + w];
            s1[threadIdx.x] += s1[threadIdx.x + w];
            s2[threadIdx.x] += s2[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[3 * blockIdx.x] = s0[0];
        part[3 * blockIdx.x + 1] = s1[0];
        part[3 * blockIdx.x + 2] = s2[0];
    }
}
// r = b - A z, its FP32 copy for the V-cycle, zf = 0; the next direction is
// the first
__global__ void k_kr_init(int kind, LevelArgs a, FemC k, const double *z, const double *b, const double *fw,
                          const double *fs, const double *kc, double *r, float *bf, float *zf, double *ks) {
    const int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y;
    if (i == 0 && j == 0) ks[kKrFirst] = 1;
    if (i >= a.nx || j >= a.ny) return;
    const int q = j * a.zp + i, c = j * a.cp + i;
    double diag, off;
    if (kind == kMAC) off = mac_offdiag<double>(z, fw, fs, q, c, a.zp, a.cp, a.kx, a.ky, diag);
    else off = fem_offdiag<double>(z, kc, q, c, a.zp, a.cp, k, diag);
    const double rq = b[q] - off - diag * z[q];
    r[q] = rq;
    bf[q] = float(rq);
    put_z(a, zf, i, j, 0.0f);
}
// c <- the V-cycle's correction in FP64; partial sums of r.c (new) and r.c (old)
__global__ void k_kr_dots(LevelArgs a, const float *zf, const double *r, double *c, double *part) {
    double rn = 0, ro = 0;
    const int N = a.nx * a.ny;
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < N; p += gridDim.x * blockDim.x) {
        const int q = (p / a.nx) * a.zp + p % a.nx;
        const double cn = double(zf[q]), rq = r[q];
        rn += rq * cn;
        ro += rq * c[q];
        c[q] = cn;
    }
    block_sum3(rn, ro, 0.0, part);
}
// beta = r.(c - c_old) / (r.c)_old (Polak-Ribiere, clamped at 0: a restart),
// 0 for the first direction
__global__ void k_kr_beta(double *ks) {
    const double rn = ks[kKrSums], ro = ks[kKrSums + 1], rho = ks[kKrRho];
    double beta = 0;
    if (ks[kKrFirst] == 0 && rho != 0) beta = (rn - ro) / rho;
    if (!(beta > 0) || !isfinite(beta)) beta = 0;
    ks[kKrBeta] = beta;
    ks[kKrRho] = rn;
    ks[kKrFirst] = 0;
}
// p <- c + beta p
__global__ void k_kr_dir(LevelArgs a, const double *c, double *p, const double *ks) {
    const int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y;
    if (i >= a.nx || j >= a.ny) return;
    const int q = j * a.zp + i;
    put_z(a, p, i, j, c[q] + ks[kKrBeta] * p[q]);
}
// q = A p; partial sums of p.q
__global__ void k_kr_apply(int kind, LevelArgs a, FemC k, const double *p, const double *fw, const double *fs,
                           const double *kc, double *qv, double *part) {
    double pq = 0;
    const int N = a.nx * a.ny;
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < N; t += gridDim.x * blockDim.x) {
        const int i = t % a.nx, j = t / a.nx, q = j * a.zp + i, c = j * a.cp + i;
        double diag, off;
        if (kind == kMAC) off = mac_offdiag<double>(p, fw, fs, q, c, a.zp, a.cp, a.kx, a.ky, diag);
        else off = fem_offdiag<double>(p, kc, q, c, a.zp, a.cp, k, diag);
        const double ap = off + diag * p[q];
        qv[q] = ap;
        pq += p[q] * ap;
    }
    block_sum3(pq, 0.0, 0.0, part);
}
// alpha = (r.c) / (p.q), 0 when that breaks down
__global__ void k_kr_alpha(double *ks) {
    const double pq = ks[kKrSums];
    double al = pq != 0 ? ks[kKrRho] / pq : 0;
    if (!isfinite(al)) al = 0;
    ks[kKrAlpha] = al;
}
// z += alpha p, r -= alpha q; the new residual's FP32 copy, zf = 0
__global__ void k_kr_update(LevelArgs a, double *z, double *r, const double *p, const double *qv, const double *ks,
                            float *bf, float *zf) {
    const int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y;
    if (i >= a.nx || j >= a.ny) return;
    const int q = j * a.zp + i;
    const double al = ks[kKrAlpha];
    put_z(a, z, i, j, z[q] + al * p[q]);
    const double rn = r[q] - al * qv[q];
    r[q] = rn;
    bf[q] = float(rn);
    put_z(a, zf, i, j, 0.0f);
}

// Starting iterate from the two previous solutions: z <- cb z - cc zp, zp <- z.
__global__ void k_extrapolate(LevelArgs a, double *z, double *zp, double cb, double cc) {
    const int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y;
    if (i >= a.nx || j >= a.ny) return;
    const int q = j * a.zp + i;
    const double zz = z[q];
    put_z(a, z, i, j, cb * zz - cc * zp[q]);
    zp[q] = zz;
}
// Quadratic start: z <- w0 z + w1 r1 + w2 r2, then the history shifts
// (r2 <- r1, r1 <- the old z).
__global__ void k_extrapolate2(LevelArgs a, double *z, double *r1, double *r2, double w0, double w1, double w2) {
    const int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y;
    if (i >= a.nx || j >= a.ny) return;
    const int q = j * a.zp + i;
    const double z0 = z[q], z1 = r1[q], z2 = r2[q];
    put_z(a, z, i, j, w0 * z0 + w1 * z1 + w2 * z2);
    r2[q] = z1;
    r1[q] = z0;
}
__global__ void k_copy_level(LevelArgs a, const double *z, double *zp) {
    const int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y;
    if (i >= a.nx || j >= a.ny) return;
    zp[j * a.zp + i] = z[j * a.zp + i];
}
__global__ void k_add_f32(LevelArgs a, double *z, const float *e) {
    const int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y;
    if (i >= a.nx || j >= a.ny) return;
    const int q = j * a.zp + i;
    put_z(a, z, i, j, z[q] + double(e[q]));
}

// Periodic ghost ring of a level array from its edges (one thread per ring
// point; corners from the opposite corner).
template <class T>
__global__ void k_wrap_ring(LevelArgs a, T *z) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int W = a.nx + 2, H = a.ny + 2, n = 2 * W + 2 * H;
    if (t >= n) return;
    int i, j;
    if (t < W) { i = t - 1; j = -1; }
    else if (t < 2 * W) { i = t - W - 1; j = a.ny; }
    else if (t < 2 * W + H) { i = -1; j = t - 2 * W - 1; }
    else { i = a.nx; j = t - 2 * W - H - 1; }
    int si = i, sj = j;
    if (a.px) si = wrap_idx(i, a.nx);
    if (a.py) sj = wrap_idx(j, a.ny);
    if (si < 0 || si >= a.nx || sj < 0 || sj >= a.ny) return;  // a wall side keeps its ring
    z[j * a.zp + i] = z[sj * a.zp + si];
}

// ------------------------------------------------------- learned start
// Network input at level 0 (channels-last BF16, 4 channels): the starting
// residual r = b - A z and the source b, each scaled to unit root mean
// square (sums: |b|^2, |r|^2 over the level), the feature channel, and the
// previous solve's correction to its quadratic start scaled likewise
// (esum: its sum of squares).
__global__ void k_guess_input(int kind, LevelArgs a, FemC k, const double *z, const double *b, const double *fw,
                              const double *fs, const double *kc, const double *feat, const double *sums,
                              const double *eprev, const double *esum, bf16_t *in) {
    const int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y;
    if (i >= a.nx || j >= a.ny) return;
    const int q = j * a.zp + i, c = j * a.cp + i;
    const double N = double(a.nx) * a.ny;
    const double sr = sums[1] > 0 ? sqrt(N / sums[1]) : 0., sb = sums[0] > 0 ? sqrt(N / sums[0]) : 0.;
    double diag, off;
    if (kind == kMAC) off = mac_offdiag<double>(z, fw, fs, q, c, a.zp, a.cp, a.kx, a.ky, diag);
    else off = fem_offdiag<double>(z, kc, q, c, a.zp, a.cp, k, diag);
    const double r = b[q] - off - diag * z[q];
    bf16_t *o = in + 4 * ((size_t)j * a.nx + i);
    o[0] = f2bf((float)(r * sr));
    o[1] = f2bf(feat ? (float)feat[q] : 0.f);
    o[2] = f2bf((float)(b[q] * sb));
    const double se = esum[0] > 0 ? sqrt(N / esum[0]) : 0.;
    o[3] = f2bf((float)(eprev[q] * se));
}
// the correction this solve made to its quadratic start, kept as the next
// solve's input, and per-block partial sums of its square
__global__ void k_guess_eprev(LevelArgs a, const double *z, const double *zs, double *e, double *part) {
    __shared__ double s0[kRedThreads];
    double t2 = 0;
    const int N = a.nx * a.ny;
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < N; p += gridDim.x * blockDim.x) {
        const int i = p % a.nx, j = p / a.nx, q = j * a.zp + i;
        const double t = z[q] - zs[q];
        put_z(a, e, i, j, t);
        t2 += t * t;
    }
    s0[threadIdx.x] = t2;
    __syncthreads();
    for (int w = kRedThreads / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) s0[threadIdx.x] += s0[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) { part[3 * blockIdx.x] = s0[0]; part[3 * blockIdx.x + 1] = 0; part[3 * blockIdx.x + 2] = 0; }
}
// the network's dense FP32 prediction into a level array
__global__ void k_guess_dir(LevelArgs a, const float *d, double *dz) {
    const int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y;
    if (i >= a.nx || j >= a.ny) return;
    put_z(a, dz, i, j, (double)d[(size_t)j * a.nx + i]);
}
// partial sums per block: (r . dz, dz . A dz, |r|^2) with r = b - A z
__global__ void k_guess_dots(int kind, LevelArgs a, FemC k, const double *z, const double *b, const double *dz,
                             const double *fw, const double *fs, const double *kc, double *part) {
    __shared__ double s0[kRedThreads], s1[kRedThreads], s2[kRedThreads];
    double ra = 0, aa = 0, rr = 0;
    const int N = a.nx * a.ny;
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < N; p += gridDim.x * blockDim.x) {
        const int i = p % a.nx, j = p / a.nx, q = j * a.zp + i, c = j * a.cp + i;
        double diag, off, dd, od;
        if (kind == kMAC) {
            off = mac_offdiag<double>(z, fw, fs, q, c, a.zp, a.cp, a.kx, a.ky, diag);
            od = mac_offdiag<double>(dz, fw, fs, q, c, a.zp, a.cp, a.kx, a.ky, dd);
        } else {
            off = fem_offdiag<double>(z, kc, q, c, a.zp, a.cp, k, diag);
            od = fem_offdiag<double>(dz, kc, q, c, a.zp, a.cp, k, dd);
        }
        const double r = b[q] - off - diag * z[q], ad = od + dd * dz[q];
        ra += r * dz[q];
        aa += dz[q] * ad;
        rr += r * r;
    }
    s0[threadIdx.x] = ra; s1[threadIdx.x] = aa; s2[threadIdx.x] = rr;
    __syncthreads();
    for (int w = kRedThreads / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            s0[threadIdx.x] += s0[threadIdx.x + w];
            s1[threadIdx.x] += s1[threadIdx.x + w];
            s2[threadIdx.x] += s2[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[3 * blockIdx.x] = s0[0]; part[3 * blockIdx.x + 1] = s1[0]; part[3 * blockIdx.x + 2] = s2[0];
    }
}
// z += alpha dz with alpha = (r . dz) / (dz . A dz), the step that minimises
// the energy norm of the error along dz (0 for a null direction)
__global__ void k_guess_apply(LevelArgs a, double *z, const double *dz, const double *s) {
    const int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y;
    if (i >= a.nx || j >= a.ny) return;
    const double alpha = s[1] != 0 && isfinite(s[0] / s[1]) ? s[0] / s[1] : 0.;
    const int q = j * a.zp + i;
    put_z(a, z, i, j, z[q] + alpha * dz[q]);
}
// Training loss of the learned start: the relative energy norm of the
// error left after the optimal step along the prediction d,
//   L = |e - alpha d|_A^2 / |e|_A^2 = 1 - (e.Ad)^2 / ((e.Ae)(d.Ad)),
// where e = z - zs is the correction the solve made to the quadratic start
// zs, r = b - A zs = A e its starting residual (to the solve's tolerance),
// so e.Ad = r.d, e.Ae = e.r and alpha = r.d / d.Ad is the step the learned
// start takes. dL/dd = -2 (r.d) / ((e.r)(d.Ad)) (r - alpha A d).
// Partial sums per block: (r.d, d.Ad, e.r).
__global__ void k_guess_edots(int kind, LevelArgs a, FemC k, const double *z, const double *zs, const double *b,
                              const double *dz, const double *fw, const double *fs, const double *kc, double *part) {
    __shared__ double s0[kRedThreads], s1[kRedThreads], s2[kRedThreads];
    double rd = 0, dad = 0, er = 0;
    const int N = a.nx * a.ny;
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < N; p += gridDim.x * blockDim.x) {
        const int i = p % a.nx, j = p / a.nx, q = j * a.zp + i, c = j * a.cp + i;
        double diag, off, dd, od;
        if (kind == kMAC) {
            off = mac_offdiag<double>(zs, fw, fs, q, c, a.zp, a.cp, a.kx, a.ky, diag);
            od = mac_offdiag<double>(dz, fw, fs, q, c, a.zp, a.cp, a.kx, a.ky, dd);
        } else {
            off = fem_offdiag<double>(zs, kc, q, c, a.zp, a.cp, k, diag);
            od = fem_offdiag<double>(dz, kc, q, c, a.zp, a.cp, k, dd);
        }
        const double r = b[q] - off - diag * zs[q], ad = od + dd * dz[q];
        rd += r * dz[q];
        dad += dz[q] * ad;
        er += (z[q] - zs[q]) * r;
    }
    s0[threadIdx.x] = rd; s1[threadIdx.x] = dad; s2[threadIdx.x] = er;
    __syncthreads();
    for (int w = kRedThreads / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            s0[threadIdx.x] += s0[threadIdx.x + w];
            s1[threadIdx.x] += s1[threadIdx.x + w];
            s2[threadIdx.x] += s2[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[3 * blockIdx.x] = s0[0]; part[3 * blockIdx.x + 1] = s1[0]; part[3 * blockIdx.x + 2] = s2[0];
    }
}
// the loss gradient with respect to the (dense FP32) prediction, and the
// loss itself into sums[3]
__global__ void k_guess_egrad(int kind, LevelArgs a, FemC k, const double *zs, const double *b, const double *dz,
                              const double *fw, const double *fs, const double *kc, double *sums, float *g) {
    const int i = blockIdx.x * BX + threadIdx.x, j = blockIdx.y * BY + threadIdx.y;
    if (i >= a.nx || j >= a.ny) return;
    // the MAC operator is negative definite, the FEM one positive: the
    // energies e.Ae and d.Ad share the operator's sign
    const double rd = sums[0], dad = sums[1], er = sums[2];
    const bool ok = dad * er > 0;
    const double alpha = ok ? rd / dad : 0., f = ok ? -2. * rd / (er * dad) : 0.;
    const int q = j * a.zp + i, c = j * a.cp + i;
    double diag, off, dd, od;
    if (kind == kMAC) {
        off = mac_offdiag<double>(zs, fw, fs, q, c, a.zp, a.cp, a.kx, a.ky, diag);
        od = mac_offdiag<double>(dz, fw, fs, q, c, a.zp, a.cp, a.kx, a.ky, dd);
    } else {
        off = fem_offdiag<double>(zs, kc, q, c, a.zp, a.cp, k, diag);
        od = fem_offdiag<double>(dz, kc, q, c, a.zp, a.cp, k, dd);
    }
    g[(size_t)j * a.nx + i] = (float)(f * ((b[q] - off - diag * zs[q]) - alpha * (od + dd * dz[q])));
    if (i == 0 && j == 0) sums[3] = ok ? 1. - rd * rd / (er * dad) : 1.;
}

FemC fem_consts(double dx, double dy) {
    // mgs_fem.cc:12-21
    const double dydx = dy / dx, dxdy = dx / dy;
    FemC k;
    k.fm = 1. / 3. * (dxdy + dydx);
    const double fey = 1. / 3. * (-2 * dxdy + dydx), fex = 1. / 3. * (-2 * dydx + dxdy);
    k.hey = 0.5 * fey;
    k.hex = 0.5 * fex;
    k.fc = -1. / 6. * (dxdy + dydx);
    return k;
}

LevelArgs args_of(const MgLevel &l) { return LevelArgs{l.nx, l.ny, l.pitch, l.cp, l.kx, l.ky, l.px, l.py, l.dr}; }
RsArgs rs_args(const Resample &r) { return RsArgs{r.i0, r.lo, r.hi, r.t}; }

}  // namespace

#define KCHECK() HIP_CHECK(hipGetLastError())

// The first level from which the remaining levels fit in LDS together.
static size_t tail_start(const std::vector<MgLevel> &L) {
    size_t t = L.size();
    long long need = 0;
    const long long budget = 60 * 1024 / (kTailArrays * sizeof(double));
    while (t > 0) {
        const long long sz = (long long)(L[t - 1].nx + 2) * (L[t - 1].ny + 2);
        if (need + sz > budget || L.size() - (t - 1) > (size_t)kTailMaxLev) break;
        need += sz;
        t--;
    }
    return t;
}

template <class T>
static TailArgs tail_args(const std::vector<MgLevel> &L, size_t t, MgKind kind, FemC k, int nu1, int nu2, bool f32,
                          size_t *shmem) {
    TailArgs A{};
    A.kind = kind; A.nlev = (int)(L.size() - t); A.nu1 = nu1; A.nu2 = nu2; A.k = k;
    A.px = L[t].px; A.py = L[t].py; A.dr = L[t].dr; A.sing = L[t].dr ? 0 : 1;
    int off = 0;
    A.wave_from = A.nlev;
    for (int q = 0; q < A.nlev; q++) {
        const MgLevel &v = L[t + q];
        TailLev &d = A.lv[q];
        d.nx = v.nx; d.ny = v.ny; d.zp = v.nx + 2; d.off = off; d.gcp = v.cp; d.kx = v.kx; d.ky = v.ky;
        const unsigned ha = (unsigned)(v.nx + 1) / 2;
        d.mag = ((1u << 20) + ha - 1) / ha;
        if (A.wave_from == A.nlev && (long long)v.nx * v.ny <= kTailWavePts) A.wave_from = q;
        if (kind == kMAC) { d.c1 = f32 ? (const void *)v.fwf : v.fw; d.c2 = f32 ? (const void *)v.fsf : v.fs; }
        else { d.c1 = f32 ? (const void *)v.kcf : v.kc; d.c2 = nullptr; }
        off += (v.nx + 2) * (v.ny + 2);
    }
    A.total = off;
    A.z_top = f32 ? (void *)L[t].zf : (void *)L[t].z;
    A.b_top = f32 ? (const void *)L[t].bf : (const void *)L[t].b;
    A.top_pitch = L[t].pitch;
    A.reduce = f32 ? 1e-5 : 1e-6;
    A.max_sweeps = 20000;
    *shmem = (size_t)kTailArrays * off * sizeof(T);
    return A;
}


Multigrid::~Multigrid() {
    for (auto &g : graph_exec_)
        if (g) (void)hipGraphExecDestroy(g);
    for (auto &e : ev_) { (void)hipEventDestroy(e.first); (void)hipEventDestroy(e.second); }
    for (hipEvent_t e : cev_)
        if (e) (void)hipEventDestroy(e);
    for (void *p : allocs) (void)hipFree(p);
    if (h_red) (void)hipHostFree(h_red);
}

int Multigrid::make_resample(int nf, double hf, int nc, Resample &r) {
    // the nc nodes span the nf fine nodes' extent; when nf = 2nc-1 the levels
    // are nested and these tables reduce to the usual 1/2 weights
    const double L0 = (nf - 1) * hf, hc = L0 / (nc - 1);
    std::vector<int> i0(nf), lo(nc, nf), hi(nc, -1), clo(nc, nf), chi(nc, -1);
    std::vector<double> t(nf);
    for (int i = 0; i < nf; i++) {
        const double x = i * hf / hc;
        int I = (int)x;
        if (I > nc - 2) I = nc - 2;
        i0[i] = I;
        t[i] = x - I;
        if (i == nf - 1) { i0[i] = nc - 2; t[i] = 1.0; }
        for (int J = i0[i]; J <= i0[i] + 1; J++) {
            lo[J] = std::min(lo[J], i);
            hi[J] = std::max(hi[J], i);
        }
    }
    for (int i = 0; i + 1 < nf; i++) {  // fine cell i = [i hf, (i+1) hf]
        const int a = (int)(i * hf / hc), b = std::min((int)((i + 1) * hf / hc), nc - 2);
        for (int J = std::max(a - 1, 0); J <= std::min(b + 1, nc - 2); J++) {
            clo[J] = std::min(clo[J], i);
            chi[J] = std::max(chi[J], i);
        }
    }
    auto up = [&](auto &v, auto *&d) -> int {
        using E = typename std::remove_reference<decltype(v[0])>::type;
        void *q;
        HIP_CHECK(hipMalloc(&q, v.size() * sizeof(E)));
        allocs.push_back(q);
        HIP_CHECK(hipMemcpy(q, v.data(), v.size() * sizeof(E), hipMemcpyHostToDevice));
        d = (decltype(d))q;
        return 0;
    };
    int rc;
    if ((rc = up(i0, r.i0)) || (rc = up(lo, r.lo)) || (rc = up(hi, r.hi)) || (rc = up(clo, r.clo)) ||
        (rc = up(chi, r.chi)) || (rc = up(t, r.t)))
        return rc;
    r.nf = nf; r.nc = nc; r.hf = hf; r.hc = hc;
    return 0;
}

int Multigrid::make_resample_per(int nf, double hf, int nc, double off, Resample &r) {
    const double ratio = (double)nc / nf;  // coarse points per fine point
    std::vector<int> i0(nf), lo(nc), hi(nc);
    std::vector<double> t(nf);
    for (int i = 0; i < nf; i++) {
        const double u = (i + off) * ratio - off, fl = std::floor(u);
        t[i] = u - fl;
        int I = (int)fl % nc;
        if (I < 0) I += nc;
        i0[i] = I;
    }
    // the fine points within one coarse spacing of coarse point I
    for (int I = 0; I < nc; I++) {
        lo[I] = (int)std::floor((I + off - 1) / ratio - off);
        hi[I] = (int)std::ceil((I + off + 1) / ratio - off);
        if (hi[I] - lo[I] + 1 > nf) return INCRMT_ERR_GRID;  // a window wrapping onto itself
    }
    auto up = [&](auto &v, auto *&d) -> int {
        using E = typename std::remove_reference<decltype(v[0])>::type;
        void *q;
        HIP_CHECK(hipMalloc(&q, v.size() * sizeof(E)));
        allocs.push_back(q);
        HIP_CHECK(hipMemcpy(q, v.data(), v.size() * sizeof(E), hipMemcpyHostToDevice));
        d = (decltype(d))q;
        return 0;
    };
    int rc;
    if ((rc = up(i0, r.i0)) || (rc = up(lo, r.lo)) || (rc = up(hi, r.hi)) || (rc = up(t, r.t))) return rc;
    r.clo = r.chi = nullptr;
    r.nf = nf; r.nc = nc; r.hf = hf; r.hc = hf * nf / nc;
    return 0;
}

int Multigrid::init(MgKind kind, int m, int n, double dx, double dy, hipStream_t s, bool px, bool py, bool dr) {
    kind_ = kind; s_ = s; dx_ = dx; dy_ = dy; dr_ = dr;
    // graph replay pays where launch latency dominates; a large finest level
    // keeps the queue full with direct launches and skips the replay cost
    use_graphs = (long long)m * n < kGraphMaxPoints;
    int nx = m, ny = n;
    double kx = 1. / (dx * dx), ky = 1. / (dy * dy), hx = dx, hy = dy;
    // a periodic direction is coloured across its seam: it needs an even
    // size on every level that is smoothed; a zero-pressure edge closes a
    // non-periodic x direction
    // (an odd size is taken on a doubly periodic finest level that is
    // resampled below, and smoothed there in nine colours)
    if (dr && px) return INCRMT_ERR_GRID;
    if (((px && nx % 2) || (py && ny % 2)) && !(px && py)) return INCRMT_ERR_GRID;
    while (true) {
        MgLevel l{};
        l.nx = nx; l.ny = ny;
        l.px = px; l.py = py; l.dr = dr;
        l.pitch = round_up(nx + 2, 32);
        l.cp = l.pitch;
        l.kx = kx; l.ky = ky;
        const size_t len = (size_t)(ny + 2) * l.pitch + 64;
        auto dalloc = [&](double *&p) -> int {
            void *q;
            HIP_CHECK(hipMalloc(&q, len * sizeof(double)));
            HIP_CHECK(memset_sync(q, 0, len * sizeof(double), s_));
            allocs.push_back(q);
            p = (double *)q + l.pitch + 1;
            return 0;
        };
        auto falloc = [&](float *&p) -> int {
            void *q;
            HIP_CHECK(hipMalloc(&q, len * sizeof(float)));
            HIP_CHECK(memset_sync(q, 0, len * sizeof(float), s_));
            allocs.push_back(q);
            p = (float *)q + l.pitch + 1;
            return 0;
        };
        int rc;
        if ((rc = dalloc(l.z)) || (rc = dalloc(l.z2)) || (rc = dalloc(l.b)) || (rc = dalloc(l.r))) return rc;
        if ((rc = falloc(l.zf)) || (rc = falloc(l.zf2)) || (rc = falloc(l.bf)) || (rc = falloc(l.rf))) return rc;
        if (kind == kMAC) {
            if ((rc = dalloc(l.fw)) || (rc = dalloc(l.fs))) return rc;
            if ((rc = falloc(l.fwf)) || (rc = falloc(l.fsf))) return rc;
        } else {
            if ((rc = dalloc(l.kc))) return rc;
            if ((rc = falloc(l.kcf))) return rc;
        }
        l.hx = hx; l.hy = hy;
        L.push_back(l);
        if (kind == kFEM && !px && !py && !dr && L.size() == 1 && nx >= 5 && ny >= 5 && (nx % 2 == 0 || ny % 2 == 0)) {
            // node counts that vertex coarsening cannot halve: the next level
            // is the largest uniform 2^k+1 grid on the same box
            auto pow2p1 = [](int n) { int c = 2; while (2 * (c - 1) + 1 <= n - 1) c = 2 * (c - 1) + 1; return c; };
            const int cx = nx % 2 ? (nx + 1) / 2 : pow2p1(nx), cy = ny % 2 ? (ny + 1) / 2 : pow2p1(ny);
            if ((rc = make_resample(nx, hx, cx, rx_)) || (rc = make_resample(ny, hy, cy, ry_))) return rc;
            resampled_ = true;
            hx = rx_.hc; hy = ry_.hc;
            nx = cx; ny = cy;
            continue;
        }
        if (px && py && !dr && L.size() == 1) {
            // a doubly periodic level whose halving does not end in a tiny
            // (directly solved) level: the next level is the largest 2^k or
            // 3 2^k grid of at most half the points on the same period
            // (conv_test's ladder 280..2520 halves to 35, 45, 63, 105 or 315)
            auto reaches = [](int a, int b) {
                while (a % 2 == 0 && b % 2 == 0 && a >= 4 && b >= 4) { a /= 2; b /= 2; }
                return (long long)a * b <= 16;
            };
            auto target = [](int k) {
                int best = 0;
                for (int c = 4; c <= k / 2; c *= 2) best = c;
                for (int c = 6; c <= k / 2; c *= 2) best = std::max(best, c);
                return best;
            };
            const int cx = target(nx), cy = target(ny);
            if (!reaches(nx, ny) && cx >= 4 && cy >= 4 && reaches(cx, cy)) {
                const double off = kind == kMAC ? 0.5 : 0.0;
                if ((rc = make_resample_per(nx, hx, cx, off, rx_)) || (rc = make_resample_per(ny, hy, cy, off, ry_)))
                    return rc;
                resampled_ = rs_per_ = true;
                hx = rx_.hc; hy = ry_.hc;
                kx = 1. / (hx * hx); ky = 1. / (hy * hy);
                nx = cx; ny = cy;
                continue;
            }
        }
        // MAC halves even cell counts; FEM halves 2^k+1 node counts, or even
        // node counts along a periodic direction or up to a zero-pressure
        // edge (nodes 2I are kept, the fixed column stays the ring)
        auto fem_more = [](int k, bool p) { return p ? k % 2 == 0 && k >= 4 : k % 2 == 1 && k >= 5; };
        bool more = kind == kMAC ? (nx % 2 == 0 && ny % 2 == 0 && nx >= 4 && ny >= 4)
                                 : (fem_more(nx, px || dr) && fem_more(ny, py));
        if (!more) break;
        if (kind == kMAC) { nx /= 2; ny /= 2; kx *= 0.25; ky *= 0.25; }
        else { nx = (px || dr) ? nx / 2 : (nx + 1) / 2; ny = py ? ny / 2 : (ny + 1) / 2; hx *= 2; hy *= 2; }
    }
    L[0].odd = (px && m % 2) || (py && n % 2);
    if (L[0].odd && !rs_per_) return INCRMT_ERR_GRID;
    // a periodic or zero-pressure coarsest level is solved exactly (dense
    // elimination in the tail cycle), so it must be tiny
    if ((px || py || dr) && (long long)L.back().nx * L.back().ny > 16) return INCRMT_ERR_GRID;
    void *q;
    HIP_CHECK(hipMalloc(&q, (3 * kRedBlocks + 8) * sizeof(double)));
    allocs.push_back(q);
    d_red = (double *)q;
    HIP_CHECK(hipHostMalloc((void **)&h_red, 8 * sizeof(double), hipHostMallocDefault));
    tail_ = tail_start(L);
    // the tail kernel divides colour indices k < ha * rows by ha = (nx+1)/2
    // as (k * ceil(2^20/ha)) >> 20: check it is exact on every tail level
    for (size_t l = tail_; l < L.size(); l++) {
        const unsigned ha = (unsigned)(L[l].nx + 1) / 2, mag = ((1u << 20) + ha - 1) / ha;
        for (unsigned k = 0; k < ha * (unsigned)L[l].ny; k++)
            if (((unsigned long long)k * mag >> 20) != k / ha || (unsigned long long)k * mag >= (1ull << 32))
                return INCRMT_ERR_GRID;
    }
    if (resampled_ && tail_ == 0) tail_ = 1;  // the tail's transfers assume nested levels
    if (int rc = set_irho_host(nullptr)) return rc;
    return autotune();
}


int Multigrid::coarsen_coefs(bool f32) {
    // f32: the caller wrote level 0's single-precision copies with its
    // coefficients; every coarse level writes its copies too
    if (resampled_) f32 = false;
    for (size_t l = 0; l + 1 < L.size(); l++) {
        const MgLevel &f = L[l];
        MgLevel &c = L[l + 1];
        if (kind_ == kMAC && l == 0 && rs_per_) {
            k_mac_coef_rs_per<<<grid2(c.nx + 1, c.ny + 1), dim3(BX, BY), 0, s_>>>(args_of(f), args_of(c), f.fw, f.fs,
                                                                                c.fw, c.fs);
        } else if (kind_ == kMAC) {
            k_mac_coef_coarsen<<<grid2(c.nx + 1, c.ny + 1), dim3(BX, BY), 0, s_>>>(
                args_of(f), args_of(c), f.fw, f.fs, c.fw, c.fs, f32 ? c.fwf : nullptr, f32 ? c.fsf : nullptr);
        } else if (l == 0 && rs_per_) {
            k_fem_coef_rs_per<<<grid2(c.nx, c.ny), dim3(BX, BY), 0, s_>>>(args_of(f), args_of(c), f.kc, c.kc);
        } else if (l == 0 && resampled_) {
            k_fem_coef_rs<<<grid2(c.nx, c.ny), dim3(BX, BY), 0, s_>>>(args_of(f), args_of(c), f.kc, c.kc, rx_.clo,
                                                                      rx_.chi, ry_.clo, ry_.chi, rx_.hf, rx_.hc,
                                                                      ry_.hf, ry_.hc);
        } else {
            k_fem_coef_coarsen<<<grid2(c.nx, c.ny), dim3(BX, BY), 0, s_>>>(args_of(f), args_of(c), f.kc, c.kc,
                                                                           f32 ? c.kcf : nullptr);
        }
        KCHECK();
        if (int rc = wrap_coefs(l + 1, f32)) return rc;
    }
    // otherwise the single-precision copies are made when a mixed solve needs them
    f32_stale_ = !f32;
    return 0;
}

int Multigrid::make_f32_coefs() {
    f32_stale_ = false;
    for (auto &l : L) {
        const long long len = (long long)(l.ny + 2) * l.pitch;
        const int nb = (int)((len + 255) / 256);
        if (kind_ == kMAC) {
            k_to_f32<<<nb, 256, 0, s_>>>(l.fw - l.pitch - 1, l.fwf - l.pitch - 1, len);
            k_to_f32<<<nb, 256, 0, s_>>>(l.fs - l.pitch - 1, l.fsf - l.pitch - 1, len);
        } else {
            k_to_f32<<<nb, 256, 0, s_>>>(l.kc - l.pitch - 1, l.kcf - l.pitch - 1, len);
        }
        KCHECK();
    }
    return 0;
}

int Multigrid::set_density(const Grid &g, const double *rho) {
    MgLevel &f = L[0];
    // once a mixed solve has run, the single-precision copies are written
    // with the coefficients instead of converted afterwards
    const bool f32 = want_f32_ && !resampled_;
    if (kind_ == kMAC)
        k_mac_coef_fine<<<grid2(f.nx + 1, f.ny + 1), dim3(BX, BY), 0, s_>>>(g, rho, args_of(f), f.fw, f.fs,
                                                                          f32 ? f.fwf : nullptr, f32 ? f.fsf : nullptr);
    else
        k_fem_coef_fine<<<grid2(f.nx, f.ny), dim3(BX, BY), 0, s_>>>(g, rho, args_of(f), f.kc, f32 ? f.kcf : nullptr);
    KCHECK();
    if (int rc = wrap_coefs(0, f32)) return rc;
    return coarsen_coefs(f32);
}

// FEM cell weights of a periodic level: ring cells mirror the last cells
int Multigrid::wrap_coefs(size_t l, bool f32) {
    const MgLevel &v = L[l];
    if (kind_ != kFEM || !(v.px || v.py)) return 0;
    const int n = 2 * (v.nx + 2) + 2 * (v.ny + 2);
    k_wrap_ring<double><<<(n + 255) / 256, 256, 0, s_>>>(args_of(v), v.kc);
    KCHECK();
    if (f32) {
        k_wrap_ring<float><<<(n + 255) / 256, 256, 0, s_>>>(args_of(v), v.kcf);
        KCHECK();
    }
    return 0;
}

int Multigrid::wrap_fine(double *z) {
    const MgLevel &v = L[0];
    if (!(v.px || v.py)) return 0;
    const int n = 2 * (v.nx + 2) + 2 * (v.ny + 2);
    k_wrap_ring<double><<<(n + 255) / 256, 256, 0, s_>>>(args_of(v), z);
    KCHECK();
    return 0;
}

int Multigrid::set_irho_host(const double *irho) {
    MgLevel &f = L[0];
    const int nx = f.nx, ny = f.ny;
    std::vector<double> h((size_t)(ny + 2) * f.pitch, 0.0);
    auto at = [&](int i, int j) -> double & { return h[(size_t)(j + 1) * f.pitch + i + 1]; };
    const bool px = f.px, py = f.py;
    if (kind_ == kMAC) {
        // fw from the left-edge entries, walls zero (mgs_mac.hh:64-102); a
        // periodic seam's right side is the right-edge entry irhor, a
        // zero-pressure edge twice that entry (mgs_mac.hh:50, 95)
        const long long mn = (long long)nx * ny;
        for (int j = 0; j < ny; j++) {
            for (int i = px ? 0 : 1; i < nx; i++) at(i, j) = irho ? irho[2 * (j * nx + i)] : 1.0;
            if (px) at(nx, j) = irho ? irho[2 * mn + j] : 1.0;
            if (f.dr) at(nx, j) = 2 * (irho ? irho[2 * mn + j] : 1.0);
        }
        HIP_CHECK(hipMemcpy(f.fw - f.pitch - 1, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
        std::fill(h.begin(), h.end(), 0.0);
        for (int j = py ? 0 : 1; j < ny; j++)
            for (int i = 0; i < nx; i++) at(i, j) = irho ? irho[2 * (j * nx + i) + 1] : 1.0;
        if (py)
            for (int i = 0; i < nx; i++) at(i, ny) = irho ? irho[2 * mn + ny + i] : 1.0;
        HIP_CHECK(hipMemcpy(f.fs - f.pitch - 1, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
    } else {
        // cell (i,j) weight = irho of its upper-right node (mgs_fem.hh:99-106),
        // node nx being node 0 along a periodic direction; with a
        // zero-pressure edge the reference's rows hold the fixed node too
        const int rl = f.dr ? nx + 1 : nx;
        for (int j = 0; j < ny - (py ? 0 : 1); j++)
            for (int i = 0; i < nx - ((px || f.dr) ? 0 : 1); i++)
                at(i, j) = irho ? irho[(size_t)((j + 1) % ny) * rl + (i + 1) % rl] : 1.0;
        HIP_CHECK(hipMemcpy(f.kc - f.pitch - 1, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
        if (int rc = wrap_coefs(0)) return rc;
    }
    return coarsen_coefs();
}

// The next pair of timing events, recorded by the timed smoother launches
// themselves (their first dispatch's start, their last dispatch's end).
int Multigrid::record_begin() {
    if (!time_smoother) return 0;
    if (ev_used_ == ev_.size()) {
        hipEvent_t a, b;
        HIP_CHECK(hipEventCreate(&a));
        HIP_CHECK(hipEventCreate(&b));
        ev_.push_back({a, b});
    }
    tev_[0] = ev_[ev_used_].first;
    tev_[1] = ev_[ev_used_].second;
    return 0;
}
int Multigrid::record_end() {
    if (!time_smoother) return 0;
    ev_used_++;
    return 0;
}
int Multigrid::collect_events() {
    for (size_t k = 0; k < ev_used_; k++) {
        float ms = 0;
        HIP_CHECK(hipEventSynchronize(ev_[k].second));
        HIP_CHECK(hipEventElapsedTime(&ms, ev_[k].first, ev_[k].second));
        smoother_ms += ms;
    }
    ev_used_ = 0;
    return 0;
}

// The fused smoother's launch shapes: a tile of SX x SY points per
// workgroup and a halo of H points, i.e. H colour phases per launch (MAC:
// H/2 red-black sweeps, FEM: H/4 four-colour sweeps). Every shape gives
// bit-identical results; Multigrid::autotune picks one per level and
// arithmetic by timing them on the level itself.
struct SmoothShape {
    int sx, sy, h;
};
// gfx950 gives a CU 160 KiB of LDS; the larger tiles halve the halo
// re-reads of the small ones, the smaller ones let two workgroups share a
// CU (up to 80 KiB each). The tile data of a shape is (SX+2H+1)(SY+2H+1)
// points of 3 MAC or 2 FEM arrays; the largest FEM tiles fit only in
// single precision.
constexpr SmoothShape kShapes[] = {{64, 16, 4}, {64, 8, 4},   {64, 4, 4},   {64, 16, 2}, {64, 8, 2},
                                   {64, 4, 2},  {64, 8, 8},   {64, 4, 8},   {64, 32, 4}, {128, 16, 4},
                                   {64, 32, 8}, {128, 16, 8}, {128, 8, 8},  {64, 64, 8}, {128, 32, 4},
                                   {128, 32, 8}, {128, 64, 8}};
constexpr int kNumShapes = sizeof(kShapes) / sizeof(kShapes[0]);

constexpr long kLdsBudget = 160 * 1024 - 1024;
inline bool shape_valid(MgKind kind, const SmoothShape &v, bool f32 = false) {
    return (kind == kMAC ? v.h <= 4 : v.h >= 4) && smooth_lds(kind, v.h, v.sx, v.sy, f32 ? 4 : 8) <= kLdsBudget;
}

// Tiles of a fused launch along a direction of n points: ceil(n / S), but a
// last tile of at most H-1 points in a walled direction is merged into the
// one before it, whose halo then reaches the ghost ring (for 2^k + 1 FEM
// nodes this drops a row and a column of nearly empty tiles).
inline int smooth_tiles(int n, int S, int H, bool periodic) {
    int g = (n + S - 1) / S;
    if (!periodic && g > 1 && n - (g - 1) * S <= H - 1) g--;
    return g;
}

// tev: when given, the call's first launch records tev[0] at its start and
// its last launch tev[1] at its end (events of the dispatches themselves,
// hipExtLaunchKernelGGL, so the span is the kernels' own time)
template <class T, int KIND, int H, int SX, int SY>
static int launch_shape(const LevelArgs &a, FemC k, T **z, T **z2, const T *b, const T *c1, const T *c2, int sweeps,
                        bool reverse, hipStream_t s, int *nlaunch, ProArgs pa, const hipEvent_t *tev) {
    const dim3 grid(smooth_tiles(a.nx, SX, H, a.px), smooth_tiles(a.ny, SY, H, a.py)), block(smooth_threads<SX, SY>());
    const int ncol = KIND == kMAC ? 2 : 4, per = H / ncol;  // sweeps per launch
    for (int sw = 0; sw < sweeps; sw += per) {
        const int nsw = sweeps - sw < per ? sweeps - sw : per;
        const ProArgs p = sw == 0 ? pa : ProArgs{nullptr, 0, 0, 0};  // the correction enters once
        const bool edges = a.px || a.py || a.dr;
        if (tev) {
            const hipEvent_t e0 = sw == 0 ? tev[0] : nullptr, e1 = sw + per >= sweeps ? tev[1] : nullptr;
            if (edges)
                hipExtLaunchKernelGGL(k_smooth_fused<T, KIND, H, SX, SY, true>, grid, block, 0, s, e0, e1, 0, a, k, *z, *z2,
                                      b, c1, c2, nsw * ncol, reverse ? ncol - 1 : 0, reverse ? -1 : 1, p);
            else
                hipExtLaunchKernelGGL(k_smooth_fused<T, KIND, H, SX, SY, false>, grid, block, 0, s, e0, e1, 0, a, k, *z,
                                      *z2, b, c1, c2, nsw * ncol, reverse ? ncol - 1 : 0, reverse ? -1 : 1, p);
        } else if (edges)
            k_smooth_fused<T, KIND, H, SX, SY, true><<<grid, block, 0, s>>>(a, k, *z, *z2, b, c1, c2, nsw * ncol,
                                                                            reverse ? ncol - 1 : 0, reverse ? -1 : 1, p);
        else
            k_smooth_fused<T, KIND, H, SX, SY, false><<<grid, block, 0, s>>>(a, k, *z, *z2, b, c1, c2, nsw * ncol,
                                                                             reverse ? ncol - 1 : 0, reverse ? -1 : 1, p);
        T *t = *z;  // the result is the level's iterate from now on
        *z = *z2;
        *z2 = t;
        HIP_CHECK(hipGetLastError());
        ++*nlaunch;
    }
    return 0;
}

template <class T, int KIND>
static int launch_kind(int shape, const LevelArgs &a, FemC k, T **z, T **z2, const T *b, const T *c1, const T *c2,
                       int sweeps, bool reverse, hipStream_t s, int *nl, ProArgs pa, const hipEvent_t *tev) {
#define SHAPE(id, H, SX, SY)                                                                          \
    case id:                                                                                          \
        if constexpr (smooth_lds(KIND, H, SX, SY, sizeof(T)) <= kLdsBudget)                          \
            return launch_shape<T, KIND, H, SX, SY>(a, k, z, z2, b, c1, c2, sweeps, reverse, s, nl, pa, tev); \
        else                                                                                          \
            return INCRMT_ERR_ARG_SHAPE;
    switch (shape) {
        SHAPE(0, 4, 64, 16)
        SHAPE(1, 4, 64, 8)
        SHAPE(2, 4, 64, 4)
        SHAPE(8, 4, 64, 32)
        SHAPE(9, 4, 128, 16)
        SHAPE(14, 4, 128, 32)
    }
    if constexpr (KIND == kMAC) {
        switch (shape) {
            SHAPE(3, 2, 64, 16)
            SHAPE(4, 2, 64, 8)
            SHAPE(5, 2, 64, 4)
        }
    } else {
        switch (shape) {
            SHAPE(6, 8, 64, 8)
            SHAPE(7, 8, 64, 4)
            SHAPE(10, 8, 64, 32)
            SHAPE(11, 8, 128, 16)
            SHAPE(12, 8, 128, 8)
            SHAPE(13, 8, 64, 64)
            SHAPE(15, 8, 128, 32)
            SHAPE(16, 8, 128, 64)
        }
    }
#undef SHAPE
    return INCRMT_ERR_ARG_SHAPE;
}

// z, z2: the level's iterate and its second buffer, swapped after every launch
template <class T>
static int launch_smooth(MgKind kind, int shape, const LevelArgs &a, FemC k, T **z, T **z2, const T *b, const T *c1,
                         const T *c2, int sweeps, bool reverse, hipStream_t s, int *nlaunch,
                         ProArgs pa = ProArgs{nullptr, 0, 0, 0}, const hipEvent_t *tev = nullptr) {
    *nlaunch = 0;
    if (kind == kMAC) return launch_kind<T, kMAC>(shape, a, k, z, z2, b, c1, c2, sweeps, reverse, s, nlaunch, pa, tev);
    return launch_kind<T, kFEM>(shape, a, k, z, z2, b, c1, c2, sweeps, reverse, s, nlaunch, pa, tev);
}

// Launch shapes measured on gfx950 (MI355X) for the walled levels of the
// benchmark ladders (MAC m x m cells, FEM (m+1)^2 nodes), by
// incrmt_amd/tune_shapes.py: the median launch time over five repeats,
// lowest time per sweep. A level listed here takes its shape from the
// table, so a run's choice does not depend on timing noise; any other
// level is timed by autotune.
struct TunedShape {
    MgKind kind;
    int f32, nx, ny, shape;
};
constexpr TunedShape kTuned[] = {
    {kMAC, 0, 1024, 1024, 14}, {kMAC, 1, 1024, 1024, 9}, {kFEM, 0, 1025, 1025, 15}, {kFEM, 1, 1025, 1025, 15},
    {kMAC, 0, 512, 512, 0},    {kMAC, 1, 512, 512, 0},   {kFEM, 0, 513, 513, 12},   {kFEM, 1, 513, 513, 12},
    {kMAC, 0, 256, 256, 0},    {kMAC, 1, 256, 256, 0},   {kFEM, 0, 257, 257, 6},    {kFEM, 1, 257, 257, 6},
    {kMAC, 0, 128, 128, 0},    {kMAC, 1, 128, 128, 0},   {kFEM, 0, 129, 129, 6},    {kFEM, 1, 129, 129, 6},
    {kMAC, 0, 64, 64, 0},      {kMAC, 1, 64, 64, 0},     {kFEM, 0, 65, 65, 6},      {kFEM, 1, 65, 65, 6},
};

int Multigrid::autotune() {
    if (f32_stale_)
        if (int rc = make_f32_coefs()) return rc;
    hipEvent_t e0, e1;
    HIP_CHECK(hipEventCreate(&e0));
    HIP_CHECK(hipEventCreate(&e1));
    int rc = 0;
    for (size_t l = 0; l < L.size() && l < tail_ && !rc; l++) {
        MgLevel &v = L[l];
        if ((long long)v.nx * v.ny < 64 * 64 || v.odd) continue;
        const LevelArgs a = args_of(v);
        const FemC k = fem_consts(v.hx, v.hy);
        for (int prec = 0; prec < 2 && !rc; prec++) {
            int pick = -1;
            if (!a.px && !a.py && !a.dr)
                for (const TunedShape &t : kTuned)
                    if (t.kind == kind_ && t.f32 == prec && t.nx == v.nx && t.ny == v.ny &&
                        shape_valid(kind_, kShapes[t.shape], prec == 1))
                        pick = t.shape;
            if (pick >= 0) {
                (prec == 0 ? v.shape64 : v.shape32) = pick;
                continue;
            }
            float best = 1e30f;
            pick = 0;
            for (int sh = 0; sh < kNumShapes && !rc; sh++) {
                if (!shape_valid(kind_, kShapes[sh], prec == 1)) continue;
                // scratch iterate: the level's residual array (always written
                // before it is read); b and the coefficients are the level's
                auto run = [&]() -> int {
                    int nl;
                    double *d0 = v.r, *d1 = v.z2;
                    float *f0 = v.rf, *f1 = v.zf2;
                    if (prec == 0)
                        return launch_smooth<double>(kind_, sh, a, k, &d0, &d1, v.b, kind_ == kMAC ? v.fw : v.kc, v.fs,
                                                     nu1, false, s_, &nl);
                    return launch_smooth<float>(kind_, sh, a, k, &f0, &f1, v.bf, kind_ == kMAC ? v.fwf : v.kcf, v.fsf,
                                                nu1, false, s_, &nl);
                };
                if ((rc = run())) break;
                // median of five timed runs
                float t[5];
                for (int r = 0; r < 5 && !rc; r++) {
                    HIP_CHECK(hipEventRecord(e0, s_));
                    rc = run();
                    HIP_CHECK(hipEventRecord(e1, s_));
                    HIP_CHECK(hipEventSynchronize(e1));
                    HIP_CHECK(hipEventElapsedTime(&t[r], e0, e1));
                }
                std::sort(t, t + 5);
                if (t[2] < best) { best = t[2]; pick = sh; }
            }
            (prec == 0 ? v.shape64 : v.shape32) = pick;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    tuned = true;
    return rc;
}

std::string Multigrid::shapes() const {
    std::string s;
    char buf[96];
    for (size_t l = 0; l < L.size() && l < tail_; l++) {
        const SmoothShape &a = kShapes[L[l].shape64], &b = kShapes[L[l].shape32];
        snprintf(buf, sizeof buf, "%s%dx%d:%dx%dh%d/%dx%dh%d", s.empty() ? "" : " ", L[l].nx, L[l].ny, a.sx, a.sy, a.h,
                 b.sx, b.sy, b.h);
        s += buf;
    }
    return s;
}

int Multigrid::smooth_check(int sweeps, int shape, bool f32) {
    MgLevel &v = L[0];
    const LevelArgs a = args_of(v);
    const FemC k = fem_consts(v.hx, v.hy);
    if (int rc = wrap_fine(v.z)) return rc;
    if (f32) {
        // the same sweeps in single precision: z and b rounded to FP32 (rings
        // included), the result widened back into z
        if (shape >= kNumShapes || (shape >= 0 && !shape_valid(kind_, kShapes[shape], true))) return INCRMT_ERR_ARG_SHAPE;
        if (f32_stale_)
            if (int rc = make_f32_coefs()) return rc;
        const long long len = (long long)(v.ny + 2) * v.pitch;
        const int nb = (int)((len + 255) / 256);
        k_to_f32<<<nb, 256, 0, s_>>>(v.z - v.pitch - 1, v.zf - v.pitch - 1, len);
        k_to_f32<<<nb, 256, 0, s_>>>(v.b - v.pitch - 1, v.bf - v.pitch - 1, len);
        KCHECK();
        if (shape < 0) {
            const int ncol = kind_ == kMAC ? 2 : 4;
            for (int sw = 0; sw < sweeps; sw++)
                for (int c = 0; c < ncol; c++) {
                    if (kind_ == kMAC) {
                        const dim3 grid((v.nx / 2 + 1 + BX - 1) / BX, (v.ny + BY - 1) / BY);
                        k_mac_smooth<float><<<grid, dim3(BX, BY), 0, s_>>>(a, v.zf, v.bf, v.fwf, v.fsf, c);
                    } else {
                        const dim3 grid(((v.nx + 1) / 2 + BX - 1) / BX, ((v.ny + 1) / 2 + BY - 1) / BY);
                        k_fem_smooth<float><<<grid, dim3(BX, BY), 0, s_>>>(a, k, v.zf, v.bf, v.kcf, c);
                    }
                    KCHECK();
                }
        } else {
            int nl;
            if (int rc = launch_smooth<float>(kind_, shape, a, k, &v.zf, &v.zf2, v.bf, kind_ == kMAC ? v.fwf : v.kcf,
                                              v.fsf, sweeps, false, s_, &nl))
                return rc;
        }
        k_to_f64<<<nb, 256, 0, s_>>>(v.zf - v.pitch - 1, v.z - v.pitch - 1, len);
        KCHECK();
        return 0;
    }
    if (shape < 0) {  // one launch per colour: the reference form of the sweep
        const int ncol = kind_ == kMAC ? 2 : 4;
        for (int sw = 0; sw < sweeps; sw++)
            for (int c = 0; c < ncol; c++) {
                if (kind_ == kMAC) {
                    const dim3 grid((v.nx / 2 + 1 + BX - 1) / BX, (v.ny + BY - 1) / BY);
                    k_mac_smooth<double><<<grid, dim3(BX, BY), 0, s_>>>(a, v.z, v.b, v.fw, v.fs, c);
                } else {
                    const dim3 grid(((v.nx + 1) / 2 + BX - 1) / BX, ((v.ny + 1) / 2 + BY - 1) / BY);
                    k_fem_smooth<double><<<grid, dim3(BX, BY), 0, s_>>>(a, k, v.z, v.b, v.kc, c);
                }
                KCHECK();
            }
        return 0;
    }
    if (shape >= kNumShapes || !shape_valid(kind_, kShapes[shape])) return INCRMT_ERR_ARG_SHAPE;
    int nl;
    return launch_smooth<double>(kind_, shape, a, k, &v.z, &v.z2, v.b, kind_ == kMAC ? v.fw : v.kc, v.fs, sweeps,
                                 false, s_, &nl);
}

template <class T>
int Multigrid::smooth_colours(size_t l, int sweeps, bool reverse) {
    MgLevel &v = L[l];
    const bool f32 = std::is_same<T, float>::value;
    T *z = f32 ? (T *)v.zf : (T *)v.z;
    const T *b = f32 ? (const T *)v.bf : (const T *)v.b;
    const T *c1 = f32 ? (const T *)(kind_ == kMAC ? v.fwf : v.kcf) : (const T *)(kind_ == kMAC ? v.fw : v.kc);
    const T *c2 = f32 ? (const T *)v.fsf : (const T *)v.fs;
    const dim3 grid((v.nx + 1) / 2 / BX + 1, (v.ny + 1) / 2 / BY + 1);
    for (int sw = 0; sw < sweeps; sw++)
        for (int cc = 0; cc < 9; cc++) {
            const int col = reverse ? 8 - cc : cc;
            k_smooth_colour9<T><<<grid, dim3(BX, BY), 0, s_>>>(kind_, args_of(v), fem_consts(v.hx, v.hy), z, b, c1, c2,
                                                               col);
            KCHECK();
        }
    return 0;
}

int Multigrid::smooth(size_t l, int sweeps, bool reverse, bool correct) {
    const MgLevel &v = L[l];
    if (v.odd) return smooth_colours<double>(l, sweeps, reverse);
    const bool timed = l == 0;
    if (timed && record_begin()) return 6;
    int nl = 0;
    ProArgs pa{nullptr, 0, 0, 0};
    if (correct)
        pa = ProArgs{L[l + 1].z, L[l + 1].pitch, L[l + 1].nx, L[l + 1].ny, L[l + 1].px, L[l + 1].py, L[l + 1].dr};
    MgLevel &w = L[l];
    int rc = launch_smooth<double>(kind_, v.shape64, args_of(v), fem_consts(v.hx, v.hy), &w.z, &w.z2, v.b,
                                   kind_ == kMAC ? v.fw : v.kc, v.fs, sweeps, reverse, s_, &nl, pa,
                                   timed && time_smoother ? tev_ : nullptr);
    if (rc) return rc;
    if (timed && time_smoother) {
        if (nl > 0 && record_end()) return 6;  // (no launch recorded no event)
        // z, b, coefficients read and z written per sweep; a fused coarse
        // correction adds the coarse level's read (a quarter point per point)
        smoother_bytes += (kind_ == kMAC ? 40.0 : 32.0) * v.nx * v.ny * sweeps + (correct ? 2.0 * v.nx * v.ny : 0.0);
        smoother_launches += nl;
    }
    return 0;
}

int Multigrid::smooth_f(size_t l, int sweeps, bool reverse, bool correct) {
    const MgLevel &v = L[l];
    if (v.odd) return smooth_colours<float>(l, sweeps, reverse);
    const bool timed = l == 0;
    if (timed && record_begin()) return 6;
    int nl = 0;
    ProArgs pa{nullptr, 0, 0, 0};
    if (correct)
        pa = ProArgs{L[l + 1].zf, L[l + 1].pitch, L[l + 1].nx, L[l + 1].ny, L[l + 1].px, L[l + 1].py, L[l + 1].dr};
    MgLevel &w = L[l];
    int rc = launch_smooth<float>(kind_, v.shape32, args_of(v), fem_consts(v.hx, v.hy), &w.zf, &w.zf2, v.bf,
                                  kind_ == kMAC ? v.fwf : v.kcf, v.fsf, sweeps, reverse, s_, &nl, pa,
                                  timed && time_smoother ? tev_ : nullptr);
    if (rc) return rc;
    if (timed && time_smoother) {
        if (nl > 0 && record_end()) return 6;  // (no launch recorded no event)
        smoother_bytes += (kind_ == kMAC ? 20.0 : 16.0) * v.nx * v.ny * sweeps + (correct ? 1.0 * v.nx * v.ny : 0.0);
        smoother_launches += nl;
    }
    return 0;
}

int Multigrid::apply(const double *x, double *y) {
    const MgLevel &f = L[0];
    if (int rc = wrap_fine(const_cast<double *>(x))) return rc;
    k_apply<<<grid2(f.nx, f.ny), dim3(BX, BY), 0, s_>>>(kind_, args_of(f), fem_consts(dx_, dy_), x, f.fw, f.fs, f.kc, y);
    KCHECK();
    return 0;
}

int Multigrid::time_fine_smoother(int sweeps, bool fp32, double *ms_per_launch, long *launches, int shape) {
    if (shape >= kNumShapes || (shape >= 0 && !shape_valid(kind_, kShapes[shape], fp32))) return INCRMT_ERR_ARG_SHAPE;
    if (fp32 && f32_stale_)
        if (int rc = make_f32_coefs()) return rc;
    const bool was = time_smoother;
    const double ms0 = smoother_ms, b0 = smoother_bytes;
    const long n0 = smoother_launches;
    int &own = fp32 ? L[0].shape32 : L[0].shape64;
    const int keep = own;
    if (shape >= 0) own = shape;
    time_smoother = true;
    int rc = smooth_fine(sweeps, fp32);
    if (!rc) rc = collect_events();
    time_smoother = was;
    own = keep;
    if (rc) return rc;
    const long n = smoother_launches - n0;
    if (launches) *launches = n;
    if (ms_per_launch) *ms_per_launch = n ? (smoother_ms - ms0) / n : 0;
    smoother_ms = ms0;
    smoother_launches = n0;
    smoother_bytes = b0;
    return 0;
}

int Multigrid::smooth_fine(int sweeps, bool fp32) { return fp32 ? smooth_f(0, sweeps, false) : smooth(0, sweeps, false); }

int Multigrid::vcycle(size_t l) {
    const MgLevel &f = L[l];
    const FemC k = fem_consts(f.hx, f.hy);
    if (l == tail_) {
        size_t sh;
        const TailArgs A = tail_args<double>(L, l, kind_, k, nu1, nu2, false, &sh);
        if (A.px || A.py) k_tail<double, true><<<1, kTailThreads, sh, s_>>>(A);
        else k_tail<double, false><<<1, kTailThreads, sh, s_>>>(A);
        KCHECK();
        return 0;
    }
    if (l + 1 == L.size()) {
        k_coarse_solve<double><<<1, 256, 0, s_>>>(kind_, args_of(f), k, f.z, f.b, f.fw, f.fs, f.kc, 1e-6, 20000,
                                                  singular());
        KCHECK();
        return 0;
    }
    int rc;
    if ((rc = smooth(l, nu1, false))) return rc;
    const MgLevel &c = L[l + 1];
    if (l == 0 && rs_per_) {
        if (kind_ == kMAC)
            k_mac_residual<double><<<grid2(f.nx, f.ny), dim3(BX, BY), 0, s_>>>(args_of(f), f.z, f.b, f.fw, f.fs, f.r);
        else
            k_fem_residual<double><<<grid2(f.nx, f.ny), dim3(BX, BY), 0, s_>>>(args_of(f), k, f.z, f.b, f.kc, f.r);
        KCHECK();
        k_restrict_rs_per<double><<<grid2(c.nx, c.ny), dim3(BX, BY), 0, s_>>>(args_of(f), args_of(c), f.r, c.b, c.z,
                                                                             rs_args(rx_), rs_args(ry_), rs_scale());
    } else if (kind_ == kMAC) {
        k_mac_restrict<double><<<grid2(c.nx, c.ny), dim3(BX, BY), 0, s_>>>(args_of(f), args_of(c), f.z, f.b, f.fw,
                                                                            f.fs, c.b, c.z);
    } else {
        if (l == 0 && resampled_) {
            k_fem_residual<double><<<grid2(f.nx, f.ny), dim3(BX, BY), 0, s_>>>(args_of(f), k, f.z, f.b, f.kc, f.r);
            KCHECK();
            k_fem_restrict_rs<double><<<grid2(c.nx, c.ny), dim3(BX, BY), 0, s_>>>(args_of(f), args_of(c), f.r, c.b,
                                                                               c.z, rs_args(rx_), rs_args(ry_));
        } else {
            k_fem_restrict_fused<double><<<grid2(c.nx, c.ny), dim3(BX, BY), 0, s_>>>(args_of(f), args_of(c), k, f.z, f.b,
                                                                                  f.kc, c.b, c.z);
        }
    }
    KCHECK();
    if ((rc = vcycle(l + 1))) return rc;
    // the coarse-grid correction enters in the post-smoother's load, except
    // across a resampled (non-nested) level pair
    if (l == 0 && resampled_) {
        if (rs_per_)
            k_prolong_rs_per<double><<<grid2(f.nx, f.ny), dim3(BX, BY), 0, s_>>>(args_of(f), args_of(c), f.z, c.z,
                                                                              rs_args(rx_), rs_args(ry_));
        else
            k_fem_prolong_rs<double><<<grid2(f.nx, f.ny), dim3(BX, BY), 0, s_>>>(args_of(f), args_of(c), f.z, c.z,
                                                                              rs_args(rx_), rs_args(ry_));
        KCHECK();
        return smooth(l, nu2, true, false);
    }
    if (nu2 == 0) {
        if (kind_ == kMAC) k_mac_prolong<double><<<grid2(f.nx, f.ny), dim3(BX, BY), 0, s_>>>(args_of(f), args_of(c), f.z, c.z);
        else k_fem_prolong<double><<<grid2(f.nx, f.ny), dim3(BX, BY), 0, s_>>>(args_of(f), args_of(c), f.z, c.z);
        KCHECK();
        return 0;
    }
    return smooth(l, nu2, true, true);
}

int Multigrid::vcycle_f(size_t l) {
    const MgLevel &f = L[l];
    const FemC k = fem_consts(f.hx, f.hy);
    if (l == tail_) {
        size_t sh;
        const TailArgs A = tail_args<float>(L, l, kind_, k, nu1, nu2, true, &sh);
        if (A.px || A.py) k_tail<float, true><<<1, kTailThreads, sh, s_>>>(A);
        else k_tail<float, false><<<1, kTailThreads, sh, s_>>>(A);
        KCHECK();
        return 0;
    }
    if (l + 1 == L.size()) {
        k_coarse_solve<float><<<1, 256, 0, s_>>>(kind_, args_of(f), k, f.zf, f.bf, f.fwf, f.fsf, f.kcf, 1e-5, 20000,
                                                 singular());
        KCHECK();
        return 0;
    }
    int rc;
    if ((rc = smooth_f(l, nu1, false))) return rc;
    const MgLevel &c = L[l + 1];
    if (l == 0 && rs_per_) {
        if (kind_ == kMAC)
            k_mac_residual<float><<<grid2(f.nx, f.ny), dim3(BX, BY), 0, s_>>>(args_of(f), f.zf, f.bf, f.fwf, f.fsf, f.rf);
        else
            k_fem_residual<float><<<grid2(f.nx, f.ny), dim3(BX, BY), 0, s_>>>(args_of(f), k, f.zf, f.bf, f.kcf, f.rf);
        KCHECK();
        k_restrict_rs_per<float><<<grid2(c.nx, c.ny), dim3(BX, BY), 0, s_>>>(
            args_of(f), args_of(c), f.rf, c.bf, c.zf, rs_args(rx_), rs_args(ry_), float(rs_scale()));
    } else if (kind_ == kMAC) {
        k_mac_restrict<float><<<grid2(c.nx, c.ny), dim3(BX, BY), 0, s_>>>(args_of(f), args_of(c), f.zf, f.bf, f.fwf,
                                                                           f.fsf, c.bf, c.zf);
    } else {
        if (l == 0 && resampled_) {
            k_fem_residual<float><<<grid2(f.nx, f.ny), dim3(BX, BY), 0, s_>>>(args_of(f), k, f.zf, f.bf, f.kcf, f.rf);
            KCHECK();
            k_fem_restrict_rs<float><<<grid2(c.nx, c.ny), dim3(BX, BY), 0, s_>>>(args_of(f), args_of(c), f.rf, c.bf,
                                                                               c.zf, rs_args(rx_), rs_args(ry_));
        } else {
            k_fem_restrict_fused<float><<<grid2(c.nx, c.ny), dim3(BX, BY), 0, s_>>>(args_of(f), args_of(c), k, f.zf, f.bf,
                                                                                  f.kcf, c.bf, c.zf);
        }
    }
    KCHECK();
    if ((rc = vcycle_f(l + 1))) return rc;
    // the coarse-grid correction enters in the post-smoother's load, except
    // across a resampled (non-nested) level pair
    if (l == 0 && resampled_) {
        if (rs_per_)
            k_prolong_rs_per<float><<<grid2(f.nx, f.ny), dim3(BX, BY), 0, s_>>>(args_of(f), args_of(c), f.zf, c.zf,
                                                                             rs_args(rx_), rs_args(ry_));
        else
            k_fem_prolong_rs<float><<<grid2(f.nx, f.ny), dim3(BX, BY), 0, s_>>>(args_of(f), args_of(c), f.zf, c.zf,
                                                                             rs_args(rx_), rs_args(ry_));
        KCHECK();
        return smooth_f(l, nu2, true, false);
    }
    if (nu2 == 0) {
        if (kind_ == kMAC) k_mac_prolong<float><<<grid2(f.nx, f.ny), dim3(BX, BY), 0, s_>>>(args_of(f), args_of(c), f.zf, c.zf);
        else k_fem_prolong<float><<<grid2(f.nx, f.ny), dim3(BX, BY), 0, s_>>>(args_of(f), args_of(c), f.zf, c.zf);
        KCHECK();
        return 0;
    }
    return smooth_f(l, nu2, true, true);
}

int Multigrid::reduce_dev() {
    const MgLevel &f = L[0];
    k_norms<<<kRedBlocks, kRedThreads, 0, s_>>>(kind_, args_of(f), fem_consts(dx_, dy_), f.z, f.b, f.fw, f.fs, f.kc,
                                                d_red + 8);
    KCHECK();
    k_sum_partials<<<1, kRedThreads, 0, s_>>>(d_red + 8, kRedBlocks, d_red);
    KCHECK();
    return 0;
}

int Multigrid::reduce_norms(double *out3) {
    if (int rc = reduce_dev()) return rc;
    HIP_CHECK(hipMemcpyAsync(h_red, d_red, 3 * sizeof(double), hipMemcpyDeviceToHost, s_));
    HIP_CHECK(hipStreamSynchronize(s_));
    for (int k = 0; k < 3; k++) out3[k] = h_red[k];
    return 0;
}

// The solve loop shared by both arithmetics. The right-hand side's mean is
// projected out on the device (pure Neumann system) and the starting
// residual is checked. Then, as TGMG's tgmg_predict does for the reference
// (mgs_common.hh:37-77), the number of V-cycles the last solve needed, less
// one, is run before the next residual check, so a solve costs two or three
// host round trips instead of one per cycle; after that the residual is
// checked after every cycle.
template <class Cycle>
int Multigrid::solve_loop(double tol, int max_cycles, MgStats *st, Cycle cycle) {
    const MgLevel &f = L[0];
    int rc;
    if ((rc = wrap_fine(f.z))) return rc;  // the iterate may come from outside
    const bool quad_start = extrapolate >= 2 && nq_ >= 3;  // this solve starts from a quadratic extrapolation
    if (extrapolate >= 2) {
        // Lagrange extrapolation in time through the newest solutions: z
        // holds the newest (time tq_[0]), zq1_ and zq2_ the two before.
        const double t = t_solve;
        if (!zq1_) {
            const size_t len = (size_t)(f.ny + 2) * f.pitch + 64;
            for (double **p : {&zq1_, &zq2_}) {
                void *q;
                HIP_CHECK(hipMalloc(&q, len * sizeof(double)));
                HIP_CHECK(memset_sync(q, 0, len * sizeof(double), s_));
                allocs.push_back(q);
                *p = (double *)q + f.pitch + 1;
            }
        }
        const double *T = tq_;
        const double eps = 100. * std::numeric_limits<double>::epsilon();
        double w0 = 1, w1 = 0, w2 = 0;
        bool ok = true;
        if (nq_ >= 3) {
            ok = std::fabs(T[0] - T[1]) > eps && std::fabs(T[0] - T[2]) > eps && std::fabs(T[1] - T[2]) > eps;
            if (ok) {
                w0 = (t - T[1]) * (t - T[2]) / ((T[0] - T[1]) * (T[0] - T[2]));
                w1 = (t - T[0]) * (t - T[2]) / ((T[1] - T[0]) * (T[1] - T[2]));
                w2 = (t - T[0]) * (t - T[1]) / ((T[2] - T[0]) * (T[2] - T[1]));
            }
        } else if (nq_ == 2) {
            ok = std::fabs(T[0] - T[1]) > eps;
            if (ok) {
                w0 = (t - T[1]) / (T[0] - T[1]);
                w1 = (t - T[0]) / (T[1] - T[0]);
            }
        }
        if (!ok) { w0 = 1; w1 = w2 = 0; }  // repeated times: keep the newest as it is
        if (nq_ >= 1) {
            k_extrapolate2<<<grid2(f.nx, f.ny), dim3(BX, BY), 0, s_>>>(args_of(f), f.z, zq1_, zq2_, w0, w1, w2);
            KCHECK();
            tq_[2] = tq_[1];
            tq_[1] = tq_[0];
        }
        tq_[0] = t;  // the time of the solution this solve leaves in z
        if (nq_ < 3) nq_++;
    } else if (extrapolate) {
        // mgs_base::initial_guess (mgs_common.cc:33-66): start from the
        // linear interpolation in time of the last two solutions, at the
        // solve time t_solve: the first call records its time, the second
        // keeps its starting iterate (the first solution) and time, later
        // calls extrapolate and shift. The start is O(dt^2) from the new
        // solution instead of O(dt).
        const double t = t_solve;
        if (scount_ < 2) {
            if (scount_ == 1) {
                if (!zprev_) {
                    void *q;
                    const size_t len = (size_t)(f.ny + 2) * f.pitch + 64;
                    HIP_CHECK(hipMalloc(&q, len * sizeof(double)));
                    HIP_CHECK(memset_sync(q, 0, len * sizeof(double), s_));
                    allocs.push_back(q);
                    zprev_ = (double *)q + f.pitch + 1;
                }
                k_copy_level<<<grid2(f.nx, f.ny), dim3(BX, BY), 0, s_>>>(args_of(f), f.z, zprev_);
                KCHECK();
                t0_ = t;
            } else {
                tprev_ = t;
            }
            scount_++;
        } else if (std::fabs(t0_ - tprev_) < 100. * std::numeric_limits<double>::epsilon()) {
            // "Multigrid routine called for duplicate times": the start is
            // left as it is
        } else {
            const double ia = 1. / (t0_ - tprev_), cb = (t - tprev_) * ia, cc = (t - t0_) * ia;
            k_extrapolate<<<grid2(f.nx, f.ny), dim3(BX, BY), 0, s_>>>(args_of(f), f.z, zprev_, cb, cc);
            KCHECK();
            tprev_ = t0_;
            t0_ = t;
        }
    }
    if (extrapolate && singular()) {
        // an extrapolated start also carries the constant null-space part
        // of the iterates forward, and the quadratic weights (3,-3,1 at
        // equal steps) amplify any drift of it from solve to solve: the
        // start is made mean-free (the caller uses only gradients and
        // mean-free values, so the solution is unchanged)
        k_norms<<<kRedBlocks, kRedThreads, 0, s_>>>(kind_, args_of(f), fem_consts(dx_, dy_), f.z, f.z, f.fw, f.fs,
                                                    f.kc, d_red + 8);
        KCHECK();
        k_sum_partials<<<1, kRedThreads, 0, s_>>>(d_red + 8, kRedBlocks, d_red);
        KCHECK();
        k_sub_mean<<<grid2(f.nx, f.ny), dim3(BX, BY), 0, s_>>>(args_of(f), f.z, d_red);
        KCHECK();
    }
    if (singular()) {
        // pure Neumann: only the mean-free part of b is solvable
        if ((rc = reduce_dev())) return rc;
        k_sub_mean<<<grid2(f.nx, f.ny), dim3(BX, BY), 0, s_>>>(args_of(f), f.b, d_red);
        KCHECK();
    }
    // the learned start (mgs_base::initial_guess's hook, mgs_common.cc:33-66);
    // with training on, a solve without it still runs the network's forward
    // pass on its input (the pair it will train on)
    // (only once the quadratic start has its three solutions: the first
    // two solves start from the previous solution and a linear
    // extrapolation, which the network was not trained on)
    const bool eligible = extrapolate == 3 && net && quad_start;
    bool learned = eligible && policy.use();
    if (learned && policy.skip_for_cost()) learned = false;  // a control solve, or off for its cost
    // the learned start's cost in V-cycles, measured once: events around
    // the network's start and around the first V-cycle after it
    const bool measure = learned && policy.cost_aware && policy.cost < 0;
    if (measure && !cev_[0])
        for (hipEvent_t &e : cev_) HIP_CHECK(hipEventCreate(&e));
    const bool sample = train && net;
    // the network's inputs (the start before the learned correction, the
    // last correction) are kept whenever a network is attached and used
    const bool keep = net && (extrapolate >= 2 || train);
    if (keep) {
        if ((rc = guess_buffers())) return rc;
        k_copy_level<<<grid2(f.nx, f.ny), dim3(BX, BY), 0, s_>>>(args_of(f), f.z, zs_);
        KCHECK();
    }
    if (measure) HIP_CHECK(hipEventRecord(cev_[0], s_));
    if ((learned || sample) && (rc = learned_start(learned))) return rc;
    if (measure) HIP_CHECK(hipEventRecord(cev_[1], s_));
    double nr[3];
    if ((rc = reduce_norms(nr))) return rc;
    const double bn = std::sqrt(nr[0]);
    double rel = bn > 0 ? std::sqrt(nr[1]) / bn : 0;
    int cycles = 0;
    bool fallback = false;
    auto fp64_cycle = [&]() -> int { return run_cycle(0, [&]() { return vcycle(0); }); };
    auto next = [&]() -> int { return fallback ? fp64_cycle() : cycle(); };
    if (learned && rel > tol && bn > 0 && max_cycles > 0) {
        // one V-cycle, then the judgement in FP64: a learned start that has
        // not reached guess_tol is dropped for the quadratic start and plain
        // FP64 V-cycles
        last_rel0 = rel;
        if ((rc = cycle())) return rc;
        if (measure) {
            HIP_CHECK(hipEventRecord(cev_[2], s_));
            HIP_CHECK(hipEventSynchronize(cev_[2]));
            float t_net = 0, t_cycle = 0;
            HIP_CHECK(hipEventElapsedTime(&t_net, cev_[0], cev_[1]));
            HIP_CHECK(hipEventElapsedTime(&t_cycle, cev_[1], cev_[2]));
            policy.cost = t_cycle > 0 ? t_net / t_cycle : 0;
        }
        cycles = 1;
        if ((rc = reduce_norms(nr))) return rc;
        rel = std::sqrt(nr[1]) / bn;
        last_rel1 = rel;
        learned_solves++;
        if (!policy.accept(rel, guess_tol)) {
            guess_fallbacks++;
            fallback = true;
            k_copy_level<<<grid2(f.nx, f.ny), dim3(BX, BY), 0, s_>>>(args_of(f), zs_, f.z);
            KCHECK();
            if ((rc = wrap_fine(f.z))) return rc;
            if ((rc = reduce_norms(nr))) return rc;
            rel = std::sqrt(nr[1]) / bn;
        }
    }
    if (rel > tol && bn > 0 && predict_ > cycles) {
        const int ahead = predict_ < max_cycles ? predict_ : max_cycles;
        for (; cycles < ahead; cycles++)
            if ((rc = next())) return rc;
        if ((rc = reduce_norms(nr))) return rc;
        rel = std::sqrt(nr[1]) / bn;
    }
    while (rel > tol && bn > 0) {
        if (cycles >= max_cycles) break;
        if ((rc = next())) return rc;
        cycles++;
        if ((rc = reduce_norms(nr))) return rc;
        rel = std::sqrt(nr[1]) / bn;
    }
    if (cycles > 0) predict_ = cycles - 1;
    if (eligible) policy.record(learned, cycles);
    if (sample && rel <= tol && bn > 0 && (rc = train_step())) return rc;
    if (keep && rel <= tol && bn > 0) {
        k_guess_eprev<<<kRedBlocks, kRedThreads, 0, s_>>>(args_of(f), f.z, zs_, eprev_, d_red + 8);
        KCHECK();
        k_sum_partials<<<1, kRedThreads, 0, s_>>>(d_red + 8, kRedBlocks, d_gs_ + 4);
        KCHECK();
    }
    if (time_smoother && (rc = collect_events())) return rc;
    if (st) { st->cycles = cycles; st->rel_res = rel; st->bnorm = bn; st->rnorm = rel * bn; }
    return rel <= tol || bn == 0 ? 0 : 2;
}

// One V-cycle as a HIP graph. A cycle's launch sequence is fixed by the
// hierarchy (kernels, launch shapes, sweep counts) and by which of each
// level's two iterate buffers is current, so it is captured once and
// replayed while those stay the same; a smoothing or timing call between
// solves that leaves the other buffer current makes the next solve capture
// again. Each level's fused smoother swaps its buffers an even number of
// times per cycle (as many pre- as post-smoothing launches), so a cycle ends
// in the buffer state it began in; a hierarchy where that fails launches its
// cycles directly. Nothing in a cycle sets memory with a memset node (every
// array a cycle clears is cleared by a kernel), and timing runs (HIP events
// around smoother launches) always launch directly.
std::vector<uintptr_t> Multigrid::launch_state() const {
    std::vector<uintptr_t> k{(uintptr_t)nu1, (uintptr_t)nu2, (uintptr_t)tail_};
    for (const MgLevel &v : L) {
        k.insert(k.end(), {(uintptr_t)v.z, (uintptr_t)v.z2, (uintptr_t)v.zf, (uintptr_t)v.zf2});
        k.insert(k.end(), {(uintptr_t)v.shape64, (uintptr_t)v.shape32});
    }
    return k;
}

void Multigrid::restore_buffers(const std::vector<uintptr_t> &key) {
    size_t i = 3;
    for (MgLevel &v : L) {
        v.z = (double *)key[i]; v.z2 = (double *)key[i + 1];
        v.zf = (float *)key[i + 2]; v.zf2 = (float *)key[i + 3];
        i += 6;
    }
}

int Multigrid::run_cycle(int which, const std::function<int()> &cycle) {
    if (!use_graphs || time_smoother) return cycle();
    const std::vector<uintptr_t> key = launch_state();
    if (graph_exec_[which] && graph_key_[which] != key) {
        (void)hipGraphExecDestroy(graph_exec_[which]);
        graph_exec_[which] = nullptr;
    }
    if (!graph_exec_[which]) {
        hipGraph_t g = nullptr;
        HIP_CHECK(hipStreamBeginCapture(s_, hipStreamCaptureModeThreadLocal));
        const int rc = cycle();
        const hipError_t e = hipStreamEndCapture(s_, &g);
        if (rc || e != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            restore_buffers(key);  // nothing ran
            return rc ? rc : hip_fail(e, "hipStreamEndCapture", __FILE__, __LINE__);
        }
        if (launch_state() != key) {
            // the cycle leaves other buffers current: a replay would start
            // from the wrong ones, so this hierarchy launches directly
            (void)hipGraphDestroy(g);
            restore_buffers(key);
            use_graphs = false;
            return cycle();
        }
        const hipError_t e2 = hipGraphInstantiate(&graph_exec_[which], g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e2 != hipSuccess) {
            graph_exec_[which] = nullptr;
            return hip_fail(e2, "hipGraphInstantiate", __FILE__, __LINE__);
        }
        graph_key_[which] = key;
    }
    HIP_CHECK(hipGraphLaunch(graph_exec_[which], s_));
    return 0;
}

int Multigrid::guess_buffers() {
    if (zs_) return 0;
    const MgLevel &f = L[0];
    const size_t len = (size_t)(f.ny + 2) * f.pitch + 64;
    for (double **p : {&zs_, &dz_, &eprev_}) {
        void *q;
        HIP_CHECK(hipMalloc(&q, len * sizeof(double)));
        HIP_CHECK(memset_sync(q, 0, len * sizeof(double), s_));
        allocs.push_back(q);
        *p = (double *)q + f.pitch + 1;
    }
    void *q;
    HIP_CHECK(hipMalloc(&q, (size_t)f.nx * f.ny * sizeof(float)));
    allocs.push_back(q);
    tgt_ = (float *)q;
    HIP_CHECK(hipMalloc(&q, 8 * sizeof(double)));
    HIP_CHECK(memset_sync(q, 0, 8 * sizeof(double), s_));
    allocs.push_back(q);
    d_gs_ = (double *)q;
    return 0;
}

int Multigrid::learned_start(bool apply) {
    const MgLevel &f = L[0];
    const FemC k = fem_consts(dx_, dy_);
    if (int rc = reduce_dev()) return rc;
    k_guess_input<<<grid2(f.nx, f.ny), dim3(BX, BY), 0, s_>>>(kind_, args_of(f), k, f.z, f.b, f.fw, f.fs, f.kc, feat,
                                                              d_red, eprev_, d_gs_ + 4, net->input_buffer());
    KCHECK();
    if (int rc = net->forward(net->input_buffer(), nullptr)) return rc;
    if (!apply) return 0;
    k_guess_dir<<<grid2(f.nx, f.ny), dim3(BX, BY), 0, s_>>>(args_of(f), net->output_buffer(), dz_);
    KCHECK();
    k_guess_dots<<<kRedBlocks, kRedThreads, 0, s_>>>(kind_, args_of(f), k, f.z, f.b, dz_, f.fw, f.fs, f.kc,
                                                    d_red + 8);
    KCHECK();
    k_sum_partials<<<1, kRedThreads, 0, s_>>>(d_red + 8, kRedBlocks, d_gs_);
    KCHECK();
    k_guess_apply<<<grid2(f.nx, f.ny), dim3(BX, BY), 0, s_>>>(args_of(f), f.z, dz_, d_gs_);
    KCHECK();
    if (singular()) {  // keep the start mean-free
        k_norms<<<kRedBlocks, kRedThreads, 0, s_>>>(kind_, args_of(f), k, f.z, f.z, f.fw, f.fs, f.kc, d_red + 8);
        KCHECK();
        k_sum_partials<<<1, kRedThreads, 0, s_>>>(d_red + 8, kRedBlocks, d_red);
        KCHECK();
        k_sub_mean<<<grid2(f.nx, f.ny), dim3(BX, BY), 0, s_>>>(args_of(f), f.z, d_red);
        KCHECK();
    }
    return 0;
}

int Multigrid::train_step() {
    const MgLevel &f = L[0];
    const FemC k = fem_consts(dx_, dy_);
    k_guess_dir<<<grid2(f.nx, f.ny), dim3(BX, BY), 0, s_>>>(args_of(f), net->output_buffer(), dz_);
    KCHECK();
    k_guess_edots<<<kRedBlocks, kRedThreads, 0, s_>>>(kind_, args_of(f), k, f.z, zs_, f.b, dz_, f.fw, f.fs, f.kc,
                                                     d_red + 8);
    KCHECK();
    k_sum_partials<<<1, kRedThreads, 0, s_>>>(d_red + 8, kRedBlocks, d_gs_);
    KCHECK();
    k_guess_egrad<<<grid2(f.nx, f.ny), dim3(BX, BY), 0, s_>>>(kind_, args_of(f), k, zs_, f.b, dz_, f.fw, f.fs, f.kc,
                                                              d_gs_, tgt_);
    KCHECK();
    if (int rc = net->backward_grad(tgt_)) return rc;
    if (int rc = net->adam(lr)) return rc;
    train_steps++;
    return 0;
}

int Multigrid::read_train_loss(double *loss) {
    if (!d_gs_) { *loss = 0; return 0; }
    HIP_CHECK(hipMemcpyAsync(h_red + 4, d_gs_ + 3, sizeof(double), hipMemcpyDeviceToHost, s_));
    HIP_CHECK(hipStreamSynchronize(s_));
    *loss = h_red[4];
    return 0;
}

void Multigrid::set_extrapolate(int mode) {
    if (mode != extrapolate) {
        // the history of one mode means nothing to another: start afresh
        scount_ = 0;
        nq_ = 0;
    }
    extrapolate = mode;
}

int Multigrid::solve(double tol, int max_cycles, MgStats *st) {
    return solve_loop(tol, max_cycles, st, [&]() { return run_cycle(0, [&]() { return vcycle(0); }); });
}

int Multigrid::krylov_buffers() {
    if (kr_) return 0;
    const MgLevel &f = L[0];
    const size_t len = (size_t)(f.ny + 2) * f.pitch + 64;
    for (double **p : {&kr_, &kp_, &kc_, &kq_}) {
        void *q;
        HIP_CHECK(hipMalloc(&q, len * sizeof(double)));
        HIP_CHECK(memset_sync(q, 0, len * sizeof(double), s_));  // finite rings and a finite first p
        allocs.push_back(q);
        *p = (double *)q + f.pitch + 1;
    }
    void *q;
    const size_t ns = kKrSums + 3 + 3 * (size_t)kRedBlocks;  // scalars, sums, partials
    HIP_CHECK(hipMalloc(&q, ns * sizeof(double)));
    HIP_CHECK(memset_sync(q, 0, ns * sizeof(double), s_));
    allocs.push_back(q);
    ks_ = (double *)q;
    return 0;
}

// The Krylov form of solve_mixed: one iteration is one FP32 V-cycle on the
// current residual, then the conjugate-gradient step (the solve loop's
// "cycle"). The stopping test is the same true relative residual.
int Multigrid::solve_krylov(double tol, int max_cycles, MgStats *st) {
    if (int rc = krylov_buffers()) return rc;
    const FemC k = fem_consts(dx_, dy_);
    double *part = ks_ + kKrSums + 3;
    bool started = false;
    return solve_loop(tol, max_cycles, st, [&]() -> int {
        const MgLevel &f = L[0];
        if (!started) {  // the residual of the solve's (possibly learned) start
            k_kr_init<<<grid2(f.nx, f.ny), dim3(BX, BY), 0, s_>>>(kind_, args_of(f), k, f.z, f.b, f.fw, f.fs, f.kc,
                                                                   kr_, f.bf, f.zf, ks_);
            KCHECK();
            started = true;
        }
        return run_cycle(2, [&]() -> int {
            const MgLevel &g = L[0];
            const LevelArgs a = args_of(g);
            if (int rc = vcycle_f(0)) return rc;
            k_kr_dots<<<kRedBlocks, kRedThreads, 0, s_>>>(a, g.zf, kr_, kc_, part);
            KCHECK();
            k_sum_partials<<<1, kRedThreads, 0, s_>>>(part, kRedBlocks, ks_ + kKrSums);
            KCHECK();
            k_kr_beta<<<1, 1, 0, s_>>>(ks_);
            KCHECK();
            k_kr_dir<<<grid2(g.nx, g.ny), dim3(BX, BY), 0, s_>>>(a, kc_, kp_, ks_);
            KCHECK();
            k_kr_apply<<<kRedBlocks, kRedThreads, 0, s_>>>(kind_, a, k, kp_, g.fw, g.fs, g.kc, kq_, part);
            KCHECK();
            k_sum_partials<<<1, kRedThreads, 0, s_>>>(part, kRedBlocks, ks_ + kKrSums);
            KCHECK();
            k_kr_alpha<<<1, 1, 0, s_>>>(ks_);
            KCHECK();
            k_kr_update<<<grid2(g.nx, g.ny), dim3(BX, BY), 0, s_>>>(a, g.z, kr_, kp_, kq_, ks_, g.bf, g.zf);
            KCHECK();
            return 0;
        });
    });
}

int Multigrid::solve_mixed(double tol, int max_cycles, MgStats *st) {
    const MgLevel &f = L[0];
    want_f32_ = true;
    if (f32_stale_)
        if (int rc = make_f32_coefs()) return rc;
    if (krylov && !f.px && !f.py && !f.dr && !resampled_) return solve_krylov(tol, max_cycles, st);
    const FemC k = fem_consts(dx_, dy_);
    // FP64 residual -> FP32 right-hand side, FP32 V-cycle, FP64 update
    return solve_loop(tol, max_cycles, st, [&]() -> int {
        return run_cycle(1, [&]() -> int {
            k_residual_f32<<<grid2(f.nx, f.ny), dim3(BX, BY), 0, s_>>>(kind_, args_of(f), k, f.z, f.b, f.fw, f.fs,
                                                                        f.kc, f.bf, f.zf);
            KCHECK();
            if (int rc = vcycle_f(0)) return rc;
            k_add_f32<<<grid2(f.nx, f.ny), dim3(BX, BY), 0, s_>>>(args_of(f), f.z, f.zf);
            KCHECK();
            return 0;
        });
    });
}

}  // namespace incrmt
