// Prints what include/alens/RunConfig.hpp reads from a RunConfig.yaml file, as one JSON object: the
// SylinderConfig fields, the raw blocks' keys and the engine configuration (alens_config) they map
// to.  On a parse error it prints {"error": status, "message": "..."} and exits with 1.
#include <cstdio>
#include <string>

#include "alens/RunConfig.hpp"

static void d3(const char *k, const double *v) { std::printf("\"%s\": [%.17g, %.17g, %.17g], ", k, v[0], v[1], v[2]); }
static void b3(const char *k, const bool *v) {
    std::printf("\"%s\": [%s, %s, %s], ", k, v[0] ? "true" : "false", v[1] ? "true" : "false", v[2] ? "true" : "false");
}

int main(int argc, char **argv) {
    if (argc < 2) return 2;
    try {
        const alens::RunConfig rc = alens::readRunConfig(argv[1]);
        alens_config base{};
        base.precision = 1;
        base.device = 3;
        const alens_config c = rc.toEngineConfig(base);
        std::printf("{");
        std::printf("\"rngSeed\": %u, \"logLevel\": %d, \"timerLevel\": %d, ", rc.rngSeed, rc.logLevel, rc.timerLevel);
        d3("simBoxLow", rc.simBoxLow);
        d3("simBoxHigh", rc.simBoxHigh);
        b3("simBoxPBC", rc.simBoxPBC);
        std::printf("\"monolayer\": %s, ", rc.monolayer ? "true" : "false");
        d3("initBoxLow", rc.initBoxLow);
        d3("initBoxHigh", rc.initBoxHigh);
        d3("initOrient", rc.initOrient);
        std::printf("\"initCircularX\": %s, \"initPreSteps\": %d, ", rc.initCircularX ? "true" : "false", rc.initPreSteps);
        std::printf("\"viscosity\": %.17g, \"KBT\": %.17g, \"linkKappa\": %.17g, \"linkGap\": %.17g, ", rc.viscosity,
                    rc.KBT, rc.linkKappa, rc.linkGap);
        std::printf("\"sylinderFixed\": %s, \"sylinderNumber\": %d, ", rc.sylinderFixed ? "true" : "false", rc.sylinderNumber);
        std::printf("\"sylinderLength\": %.17g, \"sylinderLengthSigma\": %.17g, \"sylinderDiameter\": %.17g, ",
                    rc.sylinderLength, rc.sylinderLengthSigma, rc.sylinderDiameter);
        std::printf("\"sylinderDiameterColRatio\": %.17g, \"sylinderLengthColRatio\": %.17g, \"sylinderColBuf\": %.17g, ",
                    rc.sylinderDiameterColRatio, rc.sylinderLengthColRatio, rc.sylinderColBuf);
        std::printf("\"dt\": %.17g, \"timeTotal\": %.17g, \"timeSnap\": %.17g, ", rc.dt, rc.timeTotal, rc.timeSnap);
        std::printf("\"conResTol\": %.17g, \"conMaxIte\": %d, \"conSolverChoice\": %d, ", rc.conResTol, rc.conMaxIte,
                    rc.conSolverChoice);
        std::printf("\"keys\": [");
        bool first = true;
        for (const auto &kv : rc.values) {
            std::printf("%s\"%s\"", first ? "" : ", ", kv.first.c_str());
            first = false;
        }
        std::printf("], \"blocks\": [");
        first = true;
        for (const auto &kv : rc.blocks) {
            std::printf("%s\"%s\"", first ? "" : ", ", kv.first.c_str());
            first = false;
        }
        std::printf("], ");
        std::printf("\"engine\": {");
        d3("box_lo", c.box_lo);
        d3("box_hi", c.box_hi);
        std::printf("\"pbc\": [%d, %d, %d], \"monolayer\": %d, \"col_buf\": %.17g, \"viscosity\": %.17g, \"kT\": %.17g, "
                    "\"dt\": %.17g, \"tol\": %.17g, \"diam_col_ratio\": %.17g, \"len_col_ratio\": %.17g, "
                    "\"max_ite\": %d, \"solver\": %d, \"seed\": %llu, \"precision\": %d, \"device\": %d}}\n",
                    c.pbc[0], c.pbc[1], c.pbc[2], c.monolayer, c.col_buf, c.viscosity, c.kT, c.dt, c.tol,
                    c.diam_col_ratio, c.len_col_ratio, c.max_ite, c.solver, (unsigned long long)c.seed, c.precision,
                    c.device);
    } catch (const alens::Error &e) {
        std::string m = e.what();
        for (auto &ch : m)
            if (ch == '"' || ch == '\\') ch = '\'';
        std::printf("{\"error\": %d, \"message\": \"%s\"}\n", (int)e.status, m.c_str());
        return 1;
    }
    return 0;
}
