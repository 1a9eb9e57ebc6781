// The reference's start-up path through the C++ adapter: SylinderSystem::initialize(RunConfig.yaml,
// TubuleInitial.dat, argc, argv) (SRC/TubuleSystem.cpp:79), then prepareStep / runStep / stepEuler
// steps and writeResult (rod snapshot, ConBlock snapshot, restart record).
// usage: adapter_config RunConfig.yaml posFile outFolder steps
// prints "rods N M m iters k converged c dt x tol y monolayer b pbc px py pz steps s" and writes the
// last step's impulses (m doubles) to outFolder/gamma.bin.
#include <cstdio>
#include <cstdlib>
#include <string>

#include "alens/SylinderSystem.hpp"

int main(int argc, char **argv) {
    if (argc < 5) return 2;
    try {
        alens::SylinderSystem sys(2);
        sys.initialize(argv[1], argv[2], argc, argv);
        const int steps = std::atoi(argv[4]);
        std::vector<alens::ConstraintBlock> blocks;
        size_t nrods = 0;
        for (int s = 0; s < steps; ++s) {
            sys.prepareStep();
            sys.runStep();
            blocks = sys.getCollisionBlocks();
            nrods = sys.getVelocity().size() / 6;
            sys.stepEuler();
        }
        sys.writeResult(argv[3], sys.runConfig.dt * (double)sys.getStepCount());
        std::FILE *f = std::fopen((std::string(argv[3]) + "/gamma.bin").c_str(), "wb");
        if (!f) return 3;
        for (const auto &b : blocks) std::fwrite(&b.gamma, sizeof(double), 1, f);
        std::fclose(f);
        const auto &rc = sys.runConfig;
        std::printf("rods %zu M %zu iters %d converged %d dt %.17g tol %.17g monolayer %d pbc %d %d %d steps %lld\n",
                    nrods, blocks.size(),
                    sys.getReport().iters, sys.getReport().converged, rc.dt, rc.conResTol, (int)rc.monolayer,
                    (int)rc.simBoxPBC[0], (int)rc.simBoxPBC[1], (int)rc.simBoxPBC[2], (long long)sys.getStepCount());
        return 0;
    } catch (const alens::Error &e) {
        std::printf("error %d: %s\n", (int)e.status, e.what());
        return 1;
    }
}
