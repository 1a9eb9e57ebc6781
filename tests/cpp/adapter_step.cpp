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
        int64_t M = 0, pooled = 0;
        int iters = 0, conv = 1, springs_pull = 1;
        double col[9], pst[9];
        for (int step = 0; step < 3; ++step) {
            sys.prepareStep();
            auto &pool = sys.getConstraintPoolNonConst();
            if (pool.size() != 4) return 3;
            // springs between rods (2q, 2q + 1), emplaced into thread queues the way
            // setProteinConstraints does (SRC/TubuleSystem.cpp:713-742): anchors at the rod centres
            std::vector<double> c, t;
            sys.getPositions(c, t);
            for (int q = 0; q < 3; ++q) {
                const int i = 2 * q, j = 2 * q + 1;
                double d[3], len = 0.0;
                for (int k = 0; k < 3; ++k) {
                    d[k] = c[3 * i + k] - c[3 * j + k];
                    len += d[k] * d[k];
                }
                len = std::sqrt(len);
                const double nI[3] = {d[0] / len, d[1] / len, d[2] / len}, nJ[3] = {-nI[0], -nI[1], -nI[2]};
                const double zero[3] = {0.0, 0.0, 0.0};
                const double kappa = 10.0, delta0 = len - 0.05;
                pool[q + 1].emplace_back(delta0, -delta0 * kappa, i, j, i, j, nI, nJ, zero, zero, &c[3 * i], &c[3 * j],
                                         false, true, kappa, 0.0);
                const double S[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
                pool[q + 1].back().setStress(S);
            }
            sys.runStep();
            const auto blocks = sys.getCollisionBlocks();
            M = (int64_t)blocks.size();
            pooled = 0;
            for (const auto &b : blocks)
                if (b.bilateral) {
                    ++pooled;
                    springs_pull &= b.gamma < 0.0;
                }
            iters = sys.getReport().iters;
            conv &= sys.getReport().converged;
            sys.calcConStress(col, pst);
            sys.stepEuler();
        }
        std::printf("rods %d M %lld iters %d converged %d steps %lld pooled %lld pull %d trace %.6e\n", n,
                    (long long)M, iters, conv, (long long)sys.getStepCount(), (long long)pooled, springs_pull,
                    pst[0] + pst[4] + pst[8]);
        return conv ? 0 : 1;
    } catch (const alens::Error &e) {
        std::printf("error %d: %s\n", (int)e.status, e.what());
        return 2;
    }
}
