// Rod files and restart through the C++ call surface (include/alens/SylinderSystem.hpp).
//   adapter_restart read <rodfile>        : parse a rod file, print "rods N stationary S group0 G" (no GPU)
//   adapter_restart run <folder> [n]      : system A steps 3 times, writes a snapshot, steps once more;
//                                           system B resumes from the snapshot and steps once; prints both
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <random>
#include <string>
#include <tuple>

#include "alens/SylinderSystem.hpp"

using Row = std::tuple<int, int, double>;

static std::vector<Row> rows_of(alens::SylinderSystem &s) {
    std::vector<Row> out;
    for (const auto &b : s.getCollisionBlocks())
        out.emplace_back(std::min(b.gidI, b.gidJ), std::max(b.gidI, b.gidJ), b.gamma);
    std::sort(out.begin(), out.end());
    return out;
}

int main(int argc, char **argv) {
    if (argc < 3) return 4;
    const std::string mode = argv[1];
    try {
        if (mode == "read") {
            const alens::Rods r = alens::readRodFile(argv[2]);
            int st = 0;
            for (auto v : r.immovable) st += v;
            std::printf("rods %d stationary %d group0 %d\n", r.size(), st, r.size() ? r.group[0] : 0);
            return 0;
        }
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
        cfg.seed = 77;
        alens::Rods r;
        std::mt19937_64 rng(9);
        std::uniform_real_distribution<double> U(0.0, 1.0);
        const int n = argc > 3 ? std::atoi(argv[3]) : 600;
        for (int i = 0; i < n; ++i) {
            r.gid.push_back(1000 + i);
            for (int k = 0; k < 3; ++k) r.center.push_back(3.4 * U(rng));
            const double z = 2 * U(rng) - 1, ph = 6.283185307179586 * U(rng), s = std::sqrt(1 - z * z);
            r.direction.insert(r.direction.end(), {s * std::cos(ph), s * std::sin(ph), z});
            r.length.push_back(0.5 + U(rng));
            r.radius.push_back(0.0125);
            r.group.push_back(i % 3);
        }
        const std::string folder = argv[2];
        alens::SylinderSystem a(cfg, 2);
        a.initialize(r);
        for (int step = 0; step < 3; ++step) {
            a.prepareStep();
            a.runStep();
            a.stepEuler();
        }
        a.writeResult(folder, 3 * cfg.dt);
        a.prepareStep();
        a.runStep();
        const auto ra = rows_of(a);
        const int ia = a.getReport().iters;

        alens::SylinderSystem b(cfg, 2);
        b.reinitialize(folder + "/TimeStepInfo.txt");
        b.prepareStep();
        b.runStep();
        const auto rb = rows_of(b);
        const int ib = b.getReport().iters;
        bool same = ra.size() == rb.size();
        double gmax = 0.0, dmax = 0.0;
        for (size_t k = 0; same && k < ra.size(); ++k) {
            same = std::get<0>(ra[k]) == std::get<0>(rb[k]) && std::get<1>(ra[k]) == std::get<1>(rb[k]);
            gmax = std::max(gmax, std::fabs(std::get<2>(ra[k])));
            dmax = std::max(dmax, std::fabs(std::get<2>(ra[k]) - std::get<2>(rb[k])));
        }
        std::printf("M %zu same_pairs %d iters_a %d iters_b %d rel %.3e snap %d step %lld\n", ra.size(), same ? 1 : 0,
                    ia, ib, gmax > 0 ? dmax / gmax : 0.0, b.getSnapID(), (long long)b.getStepCount());
        return 0;
    } catch (const alens::Error &e) {
        std::printf("error %d: %s\n", (int)e.status, e.what());
        return 2;
    }
}
