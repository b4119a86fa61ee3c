// Oracle solver for the two projection systems. TEST INFRASTRUCTURE ONLY.
//
// The reference solves these systems with TGMG, an external multigrid
// library that is absent here and pinned to no version, so the oracle does
// not restate a solver: it restates the OPERATORS exactly
// (mgs_mac.hh:64-102 + mgs_mac.cc:43-49; mgs_fem.hh:86-136 +
// mgs_fem.cc:139-162) and drives the residual ||b - A z||_2 / ||b||_2 down to
// the requested tolerance with a plain geometric multigrid of its own
// (colored Gauss-Seidel, rediscretized coarse levels). Tests pin the
// solutions against dense direct solves (tests/test_oracle_mg.py).
#include <cmath>
#include <vector>
#include <algorithm>
#include "oracle.hh"

namespace orc {

namespace {

// A level: nx x ny unknowns with an explicit 9-point stencil
// c[k][p], k = 0..8 ordered (dl,dc,dr,cl,cc,cr,ul,uc,ur).
struct level {
    int nx, ny;
    bool cell;               // cell-centred (MAC) or vertex (FEM)
    bool px = false, py = false;  // periodic directions: neighbours wrap
    // zero pressure on the right edge: MAC doubles the face past the last
    // cell into the diagonal; FEM fixes node column nx-1 (pfix rows hold
    // only their diagonal and a zero right-hand side)
    bool zpr = false;
    std::vector<double> c[9];
    std::vector<double> z, b, r;
    // coefficients that define the operator and are averaged to coarsen
    std::vector<double> fw, fs;  // MAC: face inverse densities (nx+1)*ny, nx*(ny+1)
    std::vector<double> kc;      // FEM: cell inverse densities cx()*cy()
    double kx, ky;               // MAC scalings 1/dx^2, 1/dy^2
    double dxdy, dydx;           // FEM ratios
    // a resampling of the doubly periodic level above it: fine point i of
    // that level sits at coarse coordinate ix[i] + tx[i] (ix wrapped)
    bool rs = false;
    std::vector<int> ix, iy;
    std::vector<double> tx, ty;
    int cx() const { return px ? nx : nx - 1; }  // FEM cells per row
    int cy() const { return py ? ny : ny - 1; }
    void alloc() {
        const size_t N = (size_t)nx * ny;
        for (auto &v : c) v.assign(N, 0.0);
        z.assign(N, 0.0); b.assign(N, 0.0); r.assign(N, 0.0);
    }
};

void mac_stencil(level &L) {
    const int nx = L.nx, ny = L.ny;
    for (int j = 0; j < ny; j++)
        for (int i = 0; i < nx; i++) {
            const int p = j * nx + i;
            const double el = L.fw[j * (nx + 1) + i], er = L.fw[j * (nx + 1) + i + 1];
            const double ed = L.fs[j * nx + i], eu = L.fs[(j + 1) * nx + i];
            // a periodic direction keeps its seam faces: face 0 on the left
            // of cell 0, face nx (the right-edge entry) right of cell nx-1
            const bool hl = L.px || i > 0, hr = L.px || i < nx - 1, hd = L.py || j > 0, hu = L.py || j < ny - 1;
            const double cl = hl ? L.kx * el : 0, cr = hr ? L.kx * er : 0;
            const double dc = hd ? L.ky * ed : 0, uc = hu ? L.ky * eu : 0;
            L.c[3][p] = cl; L.c[5][p] = cr; L.c[1][p] = dc; L.c[7][p] = uc;
            // mgs_mac.hh:50 (const rho, -3 xxsp) and 95 (2 er) at a zero-pressure edge
            L.c[4][p] = -L.kx * ((hl ? el : 0) + (hr ? er : (L.zpr ? 2 * er : 0))) - L.ky * ((hd ? ed : 0) + (hu ? eu : 0));
        }
}

void fem_stencil(level &L) {
    const int nx = L.nx, ny = L.ny;
    const double fm = 1. / 3. * (L.dxdy + L.dydx), fey = 1. / 3. * (-2 * L.dxdy + L.dydx);
    const double hey = 0.5 * fey, fex = 1. / 3. * (-2 * L.dydx + L.dxdy), hex = 0.5 * fex;
    const double fc = -1. / 6. * (L.dxdy + L.dydx);
    const int cx = L.cx(), cy = L.cy();
    auto K = [&](int ci, int cj) -> double {
        if (L.px) ci = (ci + nx) % nx;
        if (L.py) cj = (cj + ny) % ny;
        return (ci >= 0 && ci < cx && cj >= 0 && cj < cy) ? L.kc[cj * cx + ci] : 0.0;
    };
    for (int j = 0; j < ny; j++)
        for (int i = 0; i < nx; i++) {
            const int p = j * nx + i;
            const double kdl = K(i - 1, j - 1), kdr = K(i, j - 1), kul = K(i - 1, j), kur = K(i, j);
            L.c[0][p] = fc * kdl; L.c[2][p] = fc * kdr; L.c[6][p] = fc * kul; L.c[8][p] = fc * kur;
            L.c[1][p] = hey * (kdl + kdr); L.c[7][p] = hey * (kul + kur);
            L.c[3][p] = hex * (kdl + kul); L.c[5][p] = hex * (kdr + kur);
            L.c[4][p] = fm * (kdl + kdr + kul + kur);
            if (L.zpr && i == nx - 1)  // pfix row (mgs_fem.hh:65-72, 107-126)
                for (int k = 0; k < 9; k++)
                    if (k != 4) L.c[k][p] = 0;
        }
}

double stencil_at(const level &L, const std::vector<double> &x, int i, int j, bool offdiag_only) {
    const int nx = L.nx, ny = L.ny, p = j * nx + i;
    double s = 0;
    for (int dj = -1; dj <= 1; dj++) {
        int jj = j + dj;
        if (L.py) jj = (jj + ny) % ny;
        if (jj < 0 || jj >= ny) continue;
        for (int di = -1; di <= 1; di++) {
            int ii = i + di;
            if (L.px) ii = (ii + nx) % nx;
            if (ii < 0 || ii >= nx) continue;
            if (offdiag_only && di == 0 && dj == 0) continue;
            s += L.c[(dj + 1) * 3 + di + 1][p] * x[jj * nx + ii];
        }
    }
    return s;
}

void apply(const level &L, const std::vector<double> &x, std::vector<double> &y) {
#pragma omp parallel for
    for (int j = 0; j < L.ny; j++)
        for (int i = 0; i < L.nx; i++) y[j * L.nx + i] = stencil_at(L, x, i, j, false);
}

void residual(level &L) {
#pragma omp parallel for
    for (int j = 0; j < L.ny; j++)
        for (int i = 0; i < L.nx; i++) {
            const int p = j * L.nx + i;
            L.r[p] = L.b[p] - stencil_at(L, L.z, i, j, false);
        }
}

void smooth_serial(level &L);

// Colored Gauss-Seidel: colors (i%2, j%2), a valid GS ordering for 9 points
// (a level of odd size along a periodic direction, where two points of one
// colour meet across the seam, is swept in lexicographic order instead).
void smooth(level &L, int sweeps, bool reverse) {
    if ((L.px && L.nx % 2) || (L.py && L.ny % 2)) {
        for (int s = 0; s < sweeps; s++) smooth_serial(L);
        return;
    }
    for (int s = 0; s < sweeps; s++)
        for (int cc = 0; cc < 4; cc++) {
            const int col = reverse ? 3 - cc : cc, ci = col & 1, cj = col >> 1;
#pragma omp parallel for
            for (int j = cj; j < L.ny; j += 2)
                for (int i = ci; i < L.nx; i += 2) {
                    const int p = j * L.nx + i;
                    const double d = L.c[4][p];
                    if (d != 0) L.z[p] = (L.b[p] - stencil_at(L, L.z, i, j, true)) / d;
                }
        }
}

// Lexicographic Gauss-Seidel, serial: the coarsest level of a periodic
// direction may have an odd size, where the colours are not independent.
void smooth_serial(level &L) {
    for (int j = 0; j < L.ny; j++)
        for (int i = 0; i < L.nx; i++) {
            const int p = j * L.nx + i;
            const double d = L.c[4][p];
            if (d != 0) L.z[p] = (L.b[p] - stencil_at(L, L.z, i, j, true)) / d;
        }
}

double mean(const std::vector<double> &v) {
    double s = 0;
    for (double x : v) s += x;
    return v.empty() ? 0 : s / v.size();
}

// Periodic resampling weights: nf fine points (cell centres, off = 0.5, or
// nodes, off = 0) onto nc coarse points over the same period.
void resample_weights(int nf, int nc, double off, std::vector<int> &i0, std::vector<double> &t) {
    i0.resize(nf);
    t.resize(nf);
    for (int i = 0; i < nf; i++) {
        const double u = (i + off) * nc / (double)nf - off, fl = std::floor(u);
        t[i] = u - fl;
        i0[i] = (((int)fl % nc) + nc) % nc;
    }
}

// Restriction onto a resampled level: the transpose of linear
// interpolation (FEM), times nc/nf per direction (MAC, an average).
void restrict_rs(const level &F, level &C) {
    std::fill(C.b.begin(), C.b.end(), 0.0);
    const double scale = F.cell ? (double)C.nx / F.nx * ((double)C.ny / F.ny) : 1.0;
    for (int j = 0; j < F.ny; j++) {
        const int J0 = C.iy[j], J1 = (J0 + 1) % C.ny;
        const double wy[2] = {1 - C.ty[j], C.ty[j]};
        for (int i = 0; i < F.nx; i++) {
            const int I0 = C.ix[i], I1 = (I0 + 1) % C.nx;
            const double wx[2] = {1 - C.tx[i], C.tx[i]}, r = scale * F.r[j * F.nx + i];
            C.b[J0 * C.nx + I0] += wy[0] * wx[0] * r;
            C.b[J0 * C.nx + I1] += wy[0] * wx[1] * r;
            C.b[J1 * C.nx + I0] += wy[1] * wx[0] * r;
            C.b[J1 * C.nx + I1] += wy[1] * wx[1] * r;
        }
    }
}

void prolong_rs(level &F, const level &C) {
    for (int j = 0; j < F.ny; j++) {
        const int J0 = C.iy[j], J1 = (J0 + 1) % C.ny;
        const double ty = C.ty[j];
        for (int i = 0; i < F.nx; i++) {
            const int I0 = C.ix[i], I1 = (I0 + 1) % C.nx;
            const double tx = C.tx[i];
            F.z[j * F.nx + i] += (1 - ty) * ((1 - tx) * C.z[J0 * C.nx + I0] + tx * C.z[J0 * C.nx + I1])
                                 + ty * ((1 - tx) * C.z[J1 * C.nx + I0] + tx * C.z[J1 * C.nx + I1]);
        }
    }
}

double overlap(double a0, double a1, double b0, double b1) {
    const double v = std::min(a1, b1) - std::max(a0, b0);
    return v > 0 ? v : 0;
}

// The coefficients of a resampled doubly periodic level, in units of one
// period: FEM cell weights averaged over the fine cells with their overlap;
// MAC face inverse densities interpolated linearly along the face normal
// and averaged with overlap along the face.
void resample_coefs(const level &F, level &C) {
    const double hfx = 1.0 / F.nx, hfy = 1.0 / F.ny, hcx = 1.0 / C.nx, hcy = 1.0 / C.ny;
    if (!F.cell) {
        C.kc.assign((size_t)C.nx * C.ny, 0.0);
        for (int J = 0; J < C.ny; J++)
            for (int I = 0; I < C.nx; I++) {
                double s = 0;
                for (int j = 0; j < F.ny; j++) {
                    const double oy = overlap(j * hfy, (j + 1) * hfy, J * hcy, (J + 1) * hcy);
                    if (oy == 0) continue;
                    for (int i = 0; i < F.nx; i++)
                        s += oy * overlap(i * hfx, (i + 1) * hfx, I * hcx, (I + 1) * hcx) * F.kc[j * F.nx + i];
                }
                C.kc[J * C.nx + I] = s / (hcx * hcy);
            }
        return;
    }
    C.fw.assign((size_t)(C.nx + 1) * C.ny, 0.0);
    C.fs.assign((size_t)C.nx * (C.ny + 1), 0.0);
    for (int J = 0; J < C.ny; J++)
        for (int I = 0; I <= C.nx; I++) {
            const double u = I * (double)F.nx / C.nx;
            const int a = std::min((int)std::floor(u), F.nx), b = std::min(a + 1, F.nx);
            const double t = u - a;
            double s = 0;
            for (int j = 0; j < F.ny; j++) {
                const double oy = overlap(j * hfy, (j + 1) * hfy, J * hcy, (J + 1) * hcy);
                if (oy > 0) s += oy * ((1 - t) * F.fw[j * (F.nx + 1) + a] + t * F.fw[j * (F.nx + 1) + b]);
            }
            C.fw[J * (C.nx + 1) + I] = s / hcy;
        }
    for (int J = 0; J <= C.ny; J++)
        for (int I = 0; I < C.nx; I++) {
            const double u = J * (double)F.ny / C.ny;
            const int a = std::min((int)std::floor(u), F.ny), b = std::min(a + 1, F.ny);
            const double t = u - a;
            double s = 0;
            for (int i = 0; i < F.nx; i++) {
                const double ox = overlap(i * hfx, (i + 1) * hfx, I * hcx, (I + 1) * hcx);
                if (ox > 0) s += ox * ((1 - t) * F.fs[a * F.nx + i] + t * F.fs[b * F.nx + i]);
            }
            C.fs[J * C.nx + I] = s / hcx;
        }
}

void restrict_to(const level &F, level &C) {
    if (C.rs) {
        restrict_rs(F, C);
        return;
    }
    if (F.cell) {
        for (int J = 0; J < C.ny; J++)
            for (int I = 0; I < C.nx; I++) {
                const int p = 2 * J * F.nx + 2 * I;
                C.b[J * C.nx + I] = 0.25 * (F.r[p] + F.r[p + 1] + F.r[p + F.nx] + F.r[p + F.nx + 1]);
            }
    } else {
        for (int J = 0; J < C.ny; J++)
            for (int I = 0; I < C.nx; I++) {
                double s = 0;
                for (int dj = -1; dj <= 1; dj++) {
                    int j = 2 * J + dj;
                    if (F.py) j = (j + F.ny) % F.ny;
                    if (j < 0 || j >= F.ny) continue;
                    for (int di = -1; di <= 1; di++) {
                        int i = 2 * I + di;
                        if (F.px) i = (i + F.nx) % F.nx;
                        if (i < 0 || i >= F.nx) continue;
                        s += (di ? 0.5 : 1.0) * (dj ? 0.5 : 1.0) * F.r[j * F.nx + i];
                    }
                }
                C.b[J * C.nx + I] = s;
            }
        if (F.zpr)  // the fixed column stays at zero
            for (int J = 0; J < C.ny; J++) C.b[J * C.nx + C.nx - 1] = 0;
    }
}

void prolong_add(level &F, const level &C) {
    if (C.rs) {
        prolong_rs(F, C);
        return;
    }
    if (F.cell) {
        for (int j = 0; j < F.ny; j++)
            for (int i = 0; i < F.nx; i++) {
                const int I = i / 2, J = j / 2;
                int I2 = I + ((i & 1) ? 1 : -1), J2 = J + ((j & 1) ? 1 : -1);
                // past a zero-pressure edge the correction is mirrored with
                // the opposite sign
                const double sx = F.zpr && I2 >= C.nx ? -1.0 : 1.0;
                if (I2 < 0 || I2 >= C.nx) I2 = F.px ? (I2 + C.nx) % C.nx : I;
                if (J2 < 0 || J2 >= C.ny) J2 = F.py ? (J2 + C.ny) % C.ny : J;
                F.z[j * F.nx + i] += 0.5625 * C.z[J * C.nx + I] + sx * 0.1875 * C.z[J * C.nx + I2]
                                     + 0.1875 * C.z[J2 * C.nx + I] + sx * 0.0625 * C.z[J2 * C.nx + I2];
            }
    } else {
        for (int j = 0; j < F.ny; j++)
            for (int i = 0; i < F.nx; i++) {
                const int I = i / 2, J = j / 2;
                const bool oi = i & 1, oj = j & 1;
                double v;
                auto e = [&](int a, int b) { return C.z[(b % C.ny) * C.nx + a % C.nx]; };
                if (!oi && !oj) v = e(I, J);
                else if (oi && !oj) v = 0.5 * (e(I, J) + e(I + 1, J));
                else if (!oi && oj) v = 0.5 * (e(I, J) + e(I, J + 1));
                else v = 0.25 * (e(I, J) + e(I + 1, J) + e(I, J + 1) + e(I + 1, J + 1));
                F.z[j * F.nx + i] += v;
            }
    }
}

struct hierarchy {
    std::vector<level> L;
    void coarsen() {
        while (true) {
            const level &F = L.back();
            level C;
            C.cell = F.cell;
            C.px = F.px; C.py = F.py; C.zpr = F.zpr;
            if (F.px && F.py && !F.zpr && L.size() == 1) {
                // a doubly periodic level whose halving does not end on a
                // tiny level: the next level resamples it onto the largest
                // 2^k or 3 2^k grid of at most half the points
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
                const int cx = target(F.nx), cy = target(F.ny);
                if (!reaches(F.nx, F.ny) && cx >= 4 && cy >= 4 && reaches(cx, cy)) {
                    C.nx = cx; C.ny = cy;
                    C.rs = true;
                    const double off = F.cell ? 0.5 : 0.0;
                    resample_weights(F.nx, cx, off, C.ix, C.tx);
                    resample_weights(F.ny, cy, off, C.iy, C.ty);
                    resample_coefs(F, C);
                    if (F.cell) {
                        C.kx = F.kx * ((double)cx / F.nx) * ((double)cx / F.nx);
                        C.ky = F.ky * ((double)cy / F.ny) * ((double)cy / F.ny);
                        C.alloc();
                        mac_stencil(C);
                    } else {
                        const double rx = (double)F.nx / cx, ry = (double)F.ny / cy;  // spacing growth
                        C.dxdy = F.dxdy * rx / ry; C.dydx = F.dydx * ry / rx;
                        C.alloc();
                        fem_stencil(C);
                    }
                    L.push_back(std::move(C));
                    continue;
                }
            }
            if (F.cell) {
                if ((F.nx & 1) || (F.ny & 1) || F.nx < 4 || F.ny < 4) break;
                C.nx = F.nx / 2; C.ny = F.ny / 2;
                C.kx = F.kx * 0.25; C.ky = F.ky * 0.25;
                C.fw.resize((C.nx + 1) * C.ny); C.fs.resize(C.nx * (C.ny + 1));
                for (int J = 0; J < C.ny; J++)
                    for (int I = 0; I <= C.nx; I++)
                        C.fw[J * (C.nx + 1) + I] = 0.5 * (F.fw[2 * J * (F.nx + 1) + 2 * I]
                                                          + F.fw[(2 * J + 1) * (F.nx + 1) + 2 * I]);
                for (int J = 0; J <= C.ny; J++)
                    for (int I = 0; I < C.nx; I++)
                        C.fs[J * C.nx + I] = 0.5 * (F.fs[2 * J * F.nx + 2 * I] + F.fs[2 * J * F.nx + 2 * I + 1]);
                C.alloc();
                mac_stencil(C);
            } else {
                // 2^k+1 node counts keep every other node; a periodic
                // direction halves an even node count
                auto more = [](int k, bool p) { return p ? (k % 2 == 0 && k >= 4) : (k % 2 == 1 && k >= 5); };
                if (!more(F.nx, F.px) || !more(F.ny, F.py)) break;
                C.nx = F.px ? F.nx / 2 : (F.nx + 1) / 2; C.ny = F.py ? F.ny / 2 : (F.ny + 1) / 2;
                C.dxdy = F.dxdy; C.dydx = F.dydx;
                C.kc.resize((size_t)C.cx() * C.cy());
                for (int J = 0; J < C.cy(); J++)
                    for (int I = 0; I < C.cx(); I++) {
                        const int w = F.cx(), p = 2 * J * w + 2 * I;
                        C.kc[J * C.cx() + I] = 0.25 * (F.kc[p] + F.kc[p + 1] + F.kc[p + w] + F.kc[p + w + 1]);
                    }
                C.alloc();
                fem_stencil(C);
            }
            L.push_back(std::move(C));
        }
    }
    void coarse_solve(level &C) {
        const double mb = C.zpr ? 0 : mean(C.b);  // a definite system keeps b
        for (double &v : C.b) v -= mb;
        std::fill(C.z.begin(), C.z.end(), 0.0);
        double bn = 0;
        for (double v : C.b) bn += v * v;
        if (bn == 0) return;
        for (int it = 0; it < 4000; it++) {
            if (C.px || C.py) smooth_serial(C);
            else smooth(C, 1, false);
            if ((it & 15) == 15) {
                residual(C);
                double rn = 0;
                for (double v : C.r) rn += v * v;
                if (rn <= 1e-30 * bn) break;
            }
        }
        const double mz = C.zpr ? 0 : mean(C.z);
        for (double &v : C.z) v -= mz;
    }
    void vcycle(size_t l) {
        level &F = L[l];
        if (l + 1 == L.size()) { coarse_solve(F); return; }
        smooth(F, 2, false);
        residual(F);
        level &C = L[l + 1];
        restrict_to(F, C);
        std::fill(C.z.begin(), C.z.end(), 0.0);
        vcycle(l + 1);
        prolong_add(F, C);
        smooth(F, 2, true);
    }
    // Solve L[0] with z warm start; returns V-cycles or -1.
    int solve(const double *b, double *z, double tol, double *relres) {
        level &F = L[0];
        const size_t N = (size_t)F.nx * F.ny;
        F.b.assign(b, b + N);
        const double mb = F.zpr ? 0 : mean(F.b);
        for (double &v : F.b) v -= mb;
        if (F.zpr && !F.cell)  // mgs_fem.cc:85, 112
            for (int j = 0; j < F.ny; j++) F.b[(size_t)j * F.nx + F.nx - 1] = 0;
        F.z.assign(z, z + N);
        double bn = 0;
        for (double v : F.b) bn += v * v;
        bn = std::sqrt(bn);
        int cycles = 0;
        double rr = 0;
        while (true) {
            residual(F);
            double rn = 0;
            for (double v : F.r) rn += v * v;
            rr = bn > 0 ? std::sqrt(rn) / bn : 0;
            if (rr <= tol || bn == 0) break;
            if (cycles >= 200) { cycles = -1; break; }
            vcycle(0);
            cycles++;
        }
        std::copy(F.z.begin(), F.z.end(), z);
        if (relres) *relres = rr;
        return cycles;
    }
};

void fill_mac_fine(const mac_system &s, level &F) {
    const int m = s.m, n = s.n, mn = m * n;
    F.cell = true; F.nx = m; F.ny = n; F.kx = s.xxsp; F.ky = s.yysp;
    F.px = s.px; F.py = s.py; F.zpr = s.zpr;
    F.fw.assign((m + 1) * n, 0.0); F.fs.assign(m * (n + 1), 0.0);
    // reference irho layout: [2ij] left edge, [2ij+1] down edge, then the
    // right column (n) and top row (m) (fluid_2d.cc:1140-1160)
    for (int j = 0; j < n; j++)
        for (int i = 0; i < m; i++) {
            const int ij = j * m + i;
            F.fw[j * (m + 1) + i] = s.irho ? s.irho[2 * ij] : 1.0;
            F.fs[j * m + i] = s.irho ? s.irho[2 * ij + 1] : 1.0;
        }
    for (int j = 0; j < n; j++) F.fw[j * (m + 1) + m] = s.irho ? s.irho[2 * mn + j] : 1.0;
    for (int i = 0; i < m; i++) F.fs[n * m + i] = s.irho ? s.irho[2 * mn + n + i] : 1.0;
    F.alloc();
    mac_stencil(F);
}

void fill_fem_fine(const fem_system &s, level &F) {
    const int m = s.m, n = s.n;
    F.cell = false; F.nx = m; F.ny = n; F.dxdy = s.dx / s.dy; F.dydx = s.dy / s.dx;
    F.px = s.px; F.py = s.py; F.zpr = s.zpr;
    const int cx = F.cx(), cy = F.cy();
    F.kc.assign((size_t)cx * cy, 1.0);
    // reference corner layout: irho[j*m+i] = 1/rho of the cell below-left of
    // node (i,j) (fluid_2d.cc:1163-1170); across a periodic seam node m is
    // node 0 (mgs_fem.hh:99-106)
    if (s.irho)
        for (int cj = 0; cj < cy; cj++)
            for (int ci = 0; ci < cx; ci++) F.kc[cj * cx + ci] = s.irho[((cj + 1) % n) * m + (ci + 1) % m];
    F.alloc();
    fem_stencil(F);
}

}  // namespace

void mac_apply(const mac_system &s, const double *x, double *y) {
    level F;
    fill_mac_fine(s, F);
    std::vector<double> xv(x, x + (size_t)s.m * s.n), yv(xv.size());
    apply(F, xv, yv);
    std::copy(yv.begin(), yv.end(), y);
}

void fem_apply(const fem_system &s, const double *x, double *y) {
    level F;
    fill_fem_fine(s, F);
    std::vector<double> xv(x, x + (size_t)s.m * s.n), yv(xv.size());
    apply(F, xv, yv);
    std::copy(yv.begin(), yv.end(), y);
}

int mac_solve(const mac_system &s, const double *b, double *z, double tol, double *relres) {
    hierarchy h;
    h.L.resize(1);
    fill_mac_fine(s, h.L[0]);
    h.coarsen();
    return h.solve(b, z, tol, relres);
}

int fem_solve(const fem_system &s, const double *b, double *z, double tol, double *relres) {
    hierarchy h;
    h.L.resize(1);
    fill_fem_fine(s, h.L[0]);
    h.coarsen();
    return h.solve(b, z, tol, relres);
}

}  // namespace orc
