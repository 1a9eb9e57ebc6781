// Drives the C++ call surface (include/alens/SylinderSystem.hpp) the way TubuleSystem::step drives
// SimToolbox (prepareStep -> runStep -> stepEuler) on a small seeded rod system and prints a summary
// line: "rods M iters converged steps".  Rods are read from a rod file given as argv[1] when present.
#include <cmath>
#include <cstdio>
#include <fstream>
#include <random>
#include <sstream>

#include "alens/SylinderSystem.hpp"

int main(int argc, char **argv) {
    alens_config cfg{};
    for (int k = 0; k < 3; ++k) {
        cfg.box_lo[k] = 0.0;
        cfg.box_hi[k] = 3.4;
        cfg.pbc[k] = 1;
    }
    cfg.col_buf = 0.025;
    cfg.viscosity = 0.01;
    cfg.kT = 0.00411;
    cfg.dt = 1e-4;
    cfg.tol = 1e-5;
    cfg.diam_col_ratio = 1.0;
    cfg.len_col_ratio = 1.0;
    cfg.max_ite = 100000;
    cfg.seed = 1234;
    alens::Rods r;
    std::mt19937_64 rng(5);
    std::uniform_real_distribution<double> U(0.0, 1.0);
    const int n = argc > 2 ? std::atoi(argv[2]) : 600;
    for (int i = 0; i < n; ++i) {
        r.gid.push_back(i);
        for (int k = 0; k < 3; ++k) r.center.push_back(3.4 * U(rng));
        const double z = 2 * U(rng) - 1, ph = 6.283185307179586 * U(rng), s = std::sqrt(1 - z * z);
        r.direction.insert(r.direction.end(), {s * std::cos(ph), s * std::sin(ph), z});
        r.length.push_back(0.5 + U(rng));
        r.radius.push_back(0.0125);
    }
    try {
        alens::SylinderSystem sys(cfg, 4);
        sys.initialize(r);
        int64_t M = 0;
        int iters = 0, conv = 1;
        for (int step = 0; step < 3; ++step) {
            sys.prepareStep();
            auto &pool = sys.getConstraintPoolNonConst();
            if (pool.size() != 4) return 3;
            sys.runStep();
            M = (int64_t)sys.getCollisionBlocks().size();
            iters = sys.getReport().iters;
            conv &= sys.getReport().converged;
            sys.stepEuler();
        }
        std::printf("rods %d M %lld iters %d converged %d steps %lld\n", n, (long long)M, iters, conv,
                    (long long)sys.getStepCount());
        return conv ? 0 : 1;
    } catch (const alens::Error &e) {
        std::printf("error %d: %s\n", (int)e.status, e.what());
        return 2;
    }
}
