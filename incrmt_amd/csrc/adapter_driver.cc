// A reference-style driver written against the adapters of
// include/incrmt_adapter.hh, as a reference driver would use them:
// `ftest simple-spin m` (ftest.cc:116-143: G=48, rhos=3, visc=1e-2, the
// rotor obj_simple_spin(0.8,0.5,0.125,0.2,pi,1) in a no-slip [-1,1]^2 box,
// initialize(0.8,0.4,0.8)) for `steps` steps of the frame timestep
// (fluid_2d::solve with 240 frames over 4 pi), then one standalone MAC
// solve through mgs_gpu of a smooth mean-free source. Writes u, v, p (the
// ghosted (m+4)^2 grid), the solve's source and its z, all float64, to the
// output file.
//
//     adapter_driver m steps out.bin
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "incrmt_adapter.hh"

namespace {

// the members of the reference's sim_type (sim_type.hh:63-71) and
// mat_const (obj_field.hh:17-28) that the adapters read
struct sim_type {
    double ax, bx, ay, by;
    bool x_prd, y_prd;
    double w_scale, w_force;
};
struct mat_const {
    double G, rhos, ev_trans_mult;
    bool set_velocity;
};
// the members of fluid_2d that mgs_gpu reads
struct source_fluid {
    double *src;
    double time;
    void set_edge_irho(double *) {}
    void set_corner_irho(double *) {}
};

}  // namespace

int main(int argc, char **argv) {
    if (argc != 4) {
        std::fprintf(stderr, "usage: adapter_driver m steps out.bin\n");
        return 2;
    }
    const int m = std::atoi(argv[1]), steps = std::atoi(argv[2]);
    if (m < 8 || steps < 0) return 2;
    const sim_type st{-1, 1, -1, 1, false, false, 3.0, 20.0};
    const mat_const mc{48.0, 3.0, 1.0, true};
    incrmt_ref::fluid_2d_gpu<sim_type, mat_const> f(m, m, st, 1e-2, 1.0, 2.5, 2.5);
    f.add_simple_spin(0.8, 0.5, 0.125, 0.2, M_PI, 1.0, mc);
    incrmt_ref::check(incrmt_fluid_set_solver(f.handle(), 1e-8, 100, INCRMT_PREC_FP64));
    f.initialize(0.8, 0.4, 0.8);
    double adt = 0;
    f.timestep_select(4 * M_PI / 240, adt);
    f.steps(adt, steps);
    const size_t N = (size_t)(m + 4) * (m + 4);
    std::vector<double> out(3 * N + 2 * (size_t)m * m);
    f.download("u", out.data());
    f.download("v", out.data() + N);
    f.download("p", out.data() + 2 * N);

    // mgs_base::solve of a MAC system (mgs_mac_const_rho's standalone form)
    const double h = 2.0 / m;
    std::vector<double> src((size_t)m * m);
    double mean = 0;
    for (int j = 0; j < m; j++)
        for (int i = 0; i < m; i++) {
            const double x = -1 + (i + 0.5) * h, y = -1 + (j + 0.5) * h;
            src[(size_t)j * m + i] = std::cos(M_PI * x) * std::cos(2 * M_PI * y) + 0.3 * x * y;
            mean += src[(size_t)j * m + i];
        }
    mean /= (double)m * m;
    for (double &v : src) v -= mean;
    for (int k = 0; k < m * m; k++) out[3 * N + k] = src[k];
    source_fluid sf{src.data(), 0.0};
    incrmt_ref::mgs_gpu<source_fluid> ms(INCRMT_MG_MAC, false, m, m, false, false, false, h, h, 1e-10);
    ms.solve(sf);
    for (int k = 0; k < m * m; k++) out[3 * N + (size_t)m * m + k] = ms.z[k];

    FILE *fp = std::fopen(argv[3], "wb");
    if (!fp || std::fwrite(out.data(), sizeof(double), out.size(), fp) != out.size()) return 3;
    std::fclose(fp);
    std::printf("steps %d adt %.17g time %.17g mgs v-cycles %.1f\n", steps, adt, f.time(), ms.avg_v_cycles());
    return 0;
}
