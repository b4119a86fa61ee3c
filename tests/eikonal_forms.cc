// Host-side check of incrmt_amd/csrc/eikonal.hh: the branch-reduced local
// update (calc_fast) equals the reference-shaped one (calc) bit for bit, and
// the inward update equals the negated outward update on negated values, on
// random neighbourhoods (random known sets, values near a plane with noise
// and exact ties, anisotropic spacings, every grid-edge position).
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include "../incrmt_amd/csrc/eikonal.hh"
using namespace incrmt;

struct Acc {
    const double *v;
    const int *k;
    double sg;
    bool in(int q) const { return k[q]; }
    double val(int q) const { return sg * v[q]; }
};

int main(int argc, char **argv) {
    const long trials = argc > 1 ? atol(argv[1]) : 1000000;
    std::mt19937_64 rng(7);
    std::uniform_real_distribution<double> U(0, 1);
    const int m = 7, n = 7, P = 7;
    double v[49];
    int k[49];
    long bad_fast = 0, bad_flip = 0;
    for (long t = 0; t < trials; t++) {
        const double dx = 0.01 + 0.02 * U(rng), dy = (t % 3 == 0) ? dx : 0.01 + 0.02 * U(rng);
        EikGeom g;
        g.init(dx, dy);
        const bool pos = t & 1;
        const double base = pos ? U(rng) * 0.05 : -U(rng) * 0.05;
        for (int q = 0; q < 49; q++) {
            const double r = U(rng), a = 0.3 * U(rng), b = 0.3 * U(rng);
            const int i = q % 7, j = q / 7;
            v[q] = base + (pos ? 1 : -1) * (dx * (i - 3) * a + dy * (j - 3) * b) +
                   (r < 0.1 ? 0 : 0.3 * dx * (U(rng) - 0.5));
            if (r > 0.97) v[q] = v[(q + 1) % 49];
            k[q] = U(rng) < 0.6;
        }
        const int I = t % 7, J = (t / 7) % 7, ij = J * P + I;
        bool anyk = (I > 0 && k[ij - 1]) || (I < m - 1 && k[ij + 1]) || (J > 0 && k[ij - P]) || (J < n - 1 && k[ij + P]);
        Acc A{v, k, 1.0}, N{v, k, -1.0};
        double r1, r2;
        if (pos) {
            Eik<true, Acc> e{A, g, m, n, P};
            r1 = e.calc(I, J, ij);
            r2 = e.calc_fast(I, J, ij);
        } else {
            Eik<false, Acc> e{A, g, m, n, P};
            r1 = e.calc(I, J, ij);
            r2 = e.calc_fast(I, J, ij);
            Eik<true, Acc> f{N, g, m, n, P};
            const double r3 = -f.calc_fast(I, J, ij);
            if (anyk && std::memcmp(&r1, &r3, 8) != 0) bad_flip++;
        }
        if (std::memcmp(&r1, &r2, 8) != 0) bad_fast++;
    }
    printf("%ld %ld\n", bad_fast, bad_flip);
    return bad_fast || bad_flip;
}
