// SylinderSystem.hpp -- C++ call surface of the reference's rod engine over libalens.so's C ABI.
//
// The reference calls into SimToolbox's SylinderSystem / ConstraintCollector / BCQPSolver from
// TubuleSystem (SRC/TubuleSystem.cpp:79, 120, 223, 55, 697-742).  These header-only classes keep those
// names and call shapes so that code written against them moves onto the GPU engine unchanged:
//   SylinderSystem::initialize(configFile, posFile, argc, argv) / reinitialize(configFile, restartFile,
//   argc, argv, eulerStep) / runConfig / prepareStep / runStep / calcConStress / stepEuler /
//   getConstraintPoolNonConst / getStepCount / getSnapID, the per-thread ConstraintBlockPool
//   (pool.size() == thread count, one queue per thread), ConstraintBlock's 16-argument constructor and
//   setStress, and BCQPSolver::solveBBPGD; RunConfig.yaml (RunConfig.hpp), rod files
//   (TubuleInitial.dat / SylinderAscii_*.dat), constraint-block snapshots (ConBlock_*.pvtp,
//   ConBlockVTK.hpp) and the restart record.
// Errors surface as alens::Error exceptions carrying the alens_status and message.
#pragma once
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <random>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../alens_capi.h"
#include "ConBlockVTK.hpp"
#include "RunConfig.hpp"

namespace alens {

inline void check(const alens_ctx *ctx, alens_status s, const char *what) {
    if (s != ALENS_OK) throw Error(s, std::string(what) + ": " + alens_last_error(ctx));
}

// Field order of the reference's ConstraintBlock constructor (SRC/TubuleSystem.cpp:727-736).
struct ConstraintBlock {
    double delta0 = 0, gamma = 0;
    int gidI = 0, gidJ = 0, globalIndexI = 0, globalIndexJ = 0;
    double normI[3] = {}, normJ[3] = {}, posI[3] = {}, posJ[3] = {}, labI[3] = {}, labJ[3] = {};
    bool oneSide = false, bilateral = false;
    double kappa = 0, gammaLB = 0;
    double stress[9] = {};

    ConstraintBlock() = default;
    ConstraintBlock(double delta0_, double gamma_, int gidI_, int gidJ_, int globalIndexI_, int globalIndexJ_,
                    const double normI_[3], const double normJ_[3], const double posI_[3], const double posJ_[3],
                    const double labI_[3], const double labJ_[3], bool oneSide_, bool bilateral_, double kappa_,
                    double gammaLB_)
        : delta0(delta0_), gamma(gamma_), gidI(gidI_), gidJ(gidJ_), globalIndexI(globalIndexI_),
          globalIndexJ(globalIndexJ_), oneSide(oneSide_), bilateral(bilateral_), kappa(kappa_), gammaLB(gammaLB_) {
        std::memcpy(normI, normI_, sizeof normI);
        std::memcpy(normJ, normJ_, sizeof normJ);
        std::memcpy(posI, posI_, sizeof posI);
        std::memcpy(posJ, posJ_, sizeof posJ);
        std::memcpy(labI, labI_, sizeof labI);
        std::memcpy(labJ, labJ_, sizeof labJ);
    }
    // row-major 3 x 3 (Emat3 in the reference, SRC/TubuleSystem.cpp:742)
    void setStress(const double s[9]) { std::memcpy(stress, s, sizeof stress); }

    alens_con_block toC() const {
        alens_con_block c{};
        c.delta0 = delta0;
        c.gamma = gamma;
        c.kappa = kappa;
        c.gamma_lb = gammaLB;
        c.gid_i = gidI;
        c.gid_j = gidJ;
        c.gidx_i = globalIndexI;
        c.gidx_j = globalIndexJ;
        std::memcpy(c.norm_i, normI, sizeof normI);
        std::memcpy(c.norm_j, normJ, sizeof normJ);
        std::memcpy(c.pos_i, posI, sizeof posI);
        std::memcpy(c.pos_j, posJ, sizeof posJ);
        std::memcpy(c.lab_i, labI, sizeof labI);
        std::memcpy(c.lab_j, labJ, sizeof labJ);
        std::memcpy(c.stress, stress, sizeof stress);
        c.one_side = oneSide ? 1 : 0;
        c.bilateral = bilateral ? 1 : 0;
        return c;
    }
    static ConstraintBlock fromC(const alens_con_block &c) {
        ConstraintBlock b(c.delta0, c.gamma, c.gid_i, c.gid_j, c.gidx_i, c.gidx_j, c.norm_i, c.norm_j, c.pos_i,
                          c.pos_j, c.lab_i, c.lab_j, c.one_side != 0, c.bilateral != 0, c.kappa, c.gamma_lb);
        b.setStress(c.stress);
        return b;
    }
};

using ConstraintBlockQue = std::vector<ConstraintBlock>;
using ConstraintBlockPool = std::vector<ConstraintBlockQue>;

struct Rods {
    std::vector<int32_t> gid;
    std::vector<double> center, direction, length, radius;  // center/direction: 3 per rod
    std::vector<uint8_t> immovable;                          // empty = all movable
    std::vector<int32_t> group;                              // empty = -1 for every rod
    int32_t size() const { return (int32_t)gid.size(); }
};

// Rod text files (docs/source/quickstart.md:215-221): two header lines (count, time), then
// "C|S gid radius minus(3) plus(3) [group]" per rod; 'S' = stationary.  Lines that do not start with
// C or S (comments, links) are skipped.
inline Rods readRodFile(const std::string &path) {
    std::ifstream in(path);
    if (!in) throw Error(ALENS_E_ARG, "cannot open rod file " + path);
    Rods r;
    std::string line;
    while (std::getline(in, line)) {
        std::istringstream ls(line);
        std::string type;
        int32_t gid = 0;
        double rad = 0, e[6];
        if (!(ls >> type) || (type != "C" && type != "S")) continue;
        if (!(ls >> gid >> rad >> e[0] >> e[1] >> e[2] >> e[3] >> e[4] >> e[5])) continue;
        int32_t grp = -1;
        if (!(ls >> grp)) grp = -1;
        const double v[3] = {e[3] - e[0], e[4] - e[1], e[5] - e[2]};
        const double L = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        r.gid.push_back(gid);
        r.radius.push_back(rad);
        r.length.push_back(L);
        for (int k = 0; k < 3; ++k) r.center.push_back(0.5 * (e[k] + e[3 + k]));
        if (L > 0)
            r.direction.insert(r.direction.end(), {v[0] / L, v[1] / L, v[2] / L});
        else
            r.direction.insert(r.direction.end(), {0.0, 0.0, 1.0});
        r.immovable.push_back(type == "S" ? 1 : 0);
        r.group.push_back(grp);
    }
    return r;
}

inline void writeRodFile(const std::string &path, const Rods &r, double time) {
    std::FILE *f = std::fopen(path.c_str(), "w");
    if (!f) throw Error(ALENS_E_ARG, "cannot write rod file " + path);
    std::fprintf(f, "%d\n%.17g\n", r.size(), time);
    for (int32_t i = 0; i < r.size(); ++i) {
        const double *c = &r.center[3 * (size_t)i], *t = &r.direction[3 * (size_t)i], h = 0.5 * r.length[i];
        std::fprintf(f, "%c %d %.17g %.17g %.17g %.17g %.17g %.17g %.17g %d\n",
                     (!r.immovable.empty() && r.immovable[i]) ? 'S' : 'C', r.gid[i], r.radius[i], c[0] - h * t[0],
                     c[1] - h * t[1], c[2] - h * t[2], c[0] + h * t[0], c[1] + h * t[1], c[2] + h * t[2],
                     r.group.empty() ? -1 : r.group[i]);
    }
    std::fclose(f);
}

inline bool fileExists(const std::string &path) {
    std::ifstream f(path);
    return (bool)f;
}

// Rods drawn from the run configuration when no rod file exists (what the reference's initialize
// does without its posFile): sylinderNumber rods, centres uniform in initBox (with initCircularX, in
// the disc of the box's y-z cross-section), direction initOrient, or uniform on the sphere when any
// component lies outside [-1, 1] (in the x-y plane for a monolayer, whose rods sit at the box's
// mid-height), length sylinderLength or lognormal with that mean and sylinderLengthSigma, radius
// sylinderDiameter / 2.  The random stream (std::mt19937_64 seeded with rngSeed) is this engine's own.
inline Rods initialRodsFromConfig(const RunConfig &rc) {
    Rods r;
    std::mt19937_64 rng((uint64_t)rc.rngSeed);
    std::uniform_real_distribution<double> U(0.0, 1.0);
    const double *lo = rc.initBoxLow, *hi = rc.initBoxHigh;
    bool fixedDir = true;
    for (int k = 0; k < 3; ++k) fixedDir = fixedDir && rc.initOrient[k] >= -1.0 && rc.initOrient[k] <= 1.0;
    const double sig = rc.sylinderLengthSigma;
    std::lognormal_distribution<double> LN(std::log(rc.sylinderLength > 0 ? rc.sylinderLength : 1.0) - 0.5 * sig * sig,
                                           sig > 0 ? sig : 1.0);
    for (int i = 0; i < rc.sylinderNumber; ++i) {
        double c[3];
        for (int k = 0; k < 3; ++k) c[k] = lo[k] + (hi[k] - lo[k]) * U(rng);
        if (rc.initCircularX) {
            const double yc = 0.5 * (lo[1] + hi[1]), zc = 0.5 * (lo[2] + hi[2]);
            const double R = 0.5 * std::min(hi[1] - lo[1], hi[2] - lo[2]);
            const double rr = R * std::sqrt(U(rng)), th = 6.283185307179586 * U(rng);
            c[1] = yc + rr * std::cos(th);
            c[2] = zc + rr * std::sin(th);
        }
        if (rc.monolayer) c[2] = 0.5 * (lo[2] + hi[2]);
        double t[3];
        if (fixedDir) {
            for (int k = 0; k < 3; ++k) t[k] = rc.initOrient[k];
        } else if (rc.monolayer) {
            const double th = 6.283185307179586 * U(rng);
            t[0] = std::cos(th);
            t[1] = std::sin(th);
            t[2] = 0.0;
        } else {
            const double z = 2.0 * U(rng) - 1.0, ph = 6.283185307179586 * U(rng), s = std::sqrt(1.0 - z * z);
            t[0] = s * std::cos(ph);
            t[1] = s * std::sin(ph);
            t[2] = z;
        }
        const double tn = std::sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
        if (!(tn > 0)) throw Error(ALENS_E_ARG, "RunConfig: initOrient is the zero vector");
        r.gid.push_back(i);
        r.center.insert(r.center.end(), {c[0], c[1], c[2]});
        r.direction.insert(r.direction.end(), {t[0] / tn, t[1] / tn, t[2] / tn});
        r.length.push_back(sig > 0 ? LN(rng) : rc.sylinderLength);
        r.radius.push_back(0.5 * rc.sylinderDiameter);
    }
    return r;
}

// Engine settings that are not keys of RunConfig.yaml: precision form, warm start, device and
// cell-list edge (alens_config fields of the same names).
struct EngineOptions {
    int32_t precision = 0, warm_start = 0, device = 0;
    double cell_size = 0.0;
};

class SylinderSystem {
  public:
    // The reference constructs SylinderSystem empty and calls initialize(configFile, posFile, ...);
    // nThreads is the size of the constraint pool (one queue per thread).
    explicit SylinderSystem(int nThreads = 1, EngineOptions opt = EngineOptions())
        : opt_(opt), pool_((size_t)(nThreads > 0 ? nThreads : 1)) {}
    SylinderSystem(const alens_config &cfg, int nThreads = 1) : pool_((size_t)(nThreads > 0 ? nThreads : 1)) {
        create(cfg);
    }
    ~SylinderSystem() {
        if (ctx_) alens_destroy(ctx_);
    }
    SylinderSystem(const SylinderSystem &) = delete;
    SylinderSystem &operator=(const SylinderSystem &) = delete;

    // rodSystem.runConfig: the configuration read by initialize / reinitialize (SRC/TubuleSystem.cpp:112-113)
    RunConfig runConfig;

    // initialize(configFile, posFile, argc, argv) (SRC/TubuleSystem.cpp:79): read RunConfig.yaml, create
    // the engine, and load the rods of posFile, or draw them from the configuration when that file
    // does not exist; sylinderFixed makes every rod immovable.  argc / argv are accepted for the
    // reference's signature (it forwards them to its MPI / Trilinos set-up) and not used.
    void initialize(const std::string &configFile, const std::string &posFile, int argc = 0, char **argv = nullptr) {
        (void)argc;
        (void)argv;
        runConfig = readRunConfig(configFile);
        create(runConfig.toEngineConfig(engineBase()));
        Rods r = fileExists(posFile) ? readRodFile(posFile) : initialRodsFromConfig(runConfig);
        if (runConfig.sylinderFixed) r.immovable.assign((size_t)r.size(), 1);
        initialize(r);
    }
    // reinitialize(configFile, restartFile, argc, argv, eulerStep) (SRC/TubuleSystem.cpp:38): read the
    // configuration, then resume from the restart record.  The snapshot holds the positions after the
    // last Euler move (writeResult runs after stepEuler here), so the next stepEuler before any
    // runStep has nothing left to apply and returns at once; eulerStep is accepted for the signature.
    void reinitialize(const std::string &configFile, const std::string &restartFile, int argc, char **argv,
                      bool eulerStep) {
        (void)argc;
        (void)argv;
        (void)eulerStep;
        runConfig = readRunConfig(configFile);
        create(runConfig.toEngineConfig(engineBase()));
        reinitialize(restartFile);
        if (runConfig.sylinderFixed) {
            rods_.immovable.assign((size_t)n_, 1);
            std::vector<double> c, t;
            getPositions(c, t);
            check(ctx_, alens_set_rods(ctx_, n_, rods_.gid.data(), c.data(), t.data(), rods_.length.data(),
                                       rods_.radius.data(), rods_.immovable.data()),
                  "set_rods");
        }
    }

    // initialize(): load rods; spatialSort reorders them the way FDPS keeps its particle arrays
    void initialize(const Rods &r, bool spatialSort = true) {
        if (!ctx_) throw Error(ALENS_E_STATE, "SylinderSystem: no engine configuration (construct with an "
                                              "alens_config or call initialize(configFile, posFile))");
        n_ = r.size();
        check(ctx_, alens_set_rods(ctx_, n_, r.gid.data(), r.center.data(), r.direction.data(), r.length.data(),
                                   r.radius.data(), r.immovable.empty() ? nullptr : r.immovable.data()),
              "set_rods");
        rods_ = r;
        if (spatialSort && n_ > 0) {
            std::vector<int32_t> perm((size_t)n_);
            check(ctx_, alens_sort_rods(ctx_, perm.data()), "sort_rods");
            Rods s;
            for (int32_t i : perm) {
                s.gid.push_back(r.gid[i]);
                s.length.push_back(r.length[i]);
                s.radius.push_back(r.radius[i]);
                if (!r.immovable.empty()) s.immovable.push_back(r.immovable[i]);
                if (!r.group.empty()) s.group.push_back(r.group[i]);
            }
            rods_ = std::move(s);
        }
    }
    // rods from a rod file only (the engine already configured)
    void initializeFromFile(const std::string &posFile, bool spatialSort = true) {
        initialize(readRodFile(posFile), spatialSort);
    }

    // writeResult(): SylinderAscii_<snapID>.dat with the current rods (17 significant digits), the
    // constraint blocks of the last step with their impulses as ConBlock_r0_<snapID>.vtp and
    // ConBlock_<snapID>.pvtp (when a step has run since the last Euler move, i.e. when called between
    // runStep and stepEuler; otherwise the blocks of the step before are kept from that call), and the
    // restart record TimeStepInfo.txt ("key value" lines: snapID, stepCount, time, snapFile) in folder.
    // For a restart that continues the trajectory, call it after stepEuler, so the rod file holds the
    // positions the next step starts from.
    void writeResult(const std::string &folder, double time) {
        Rods r = rods_;
        getPositions(r.center, r.direction);
        const std::string name = "SylinderAscii_" + std::to_string(snapID_) + ".dat";
        writeRodFile(folder + "/" + name, r, time);
        if (stepped_) lastBlocks_ = getCollisionBlocksC();
        if (!vtk::writeConBlocks(folder, snapID_, lastBlocks_))
            throw Error(ALENS_E_ARG, "cannot write ConBlock files in " + folder);
        std::ofstream info(folder + "/TimeStepInfo.txt");
        info << "snapID " << snapID_ << "\nstepCount " << (getStepCount() - 1) << "\ntime ";
        char buf[40];
        std::snprintf(buf, sizeof buf, "%.17g", time);
        info << buf << "\nsnapFile " << name << "\n";
        if (!info) throw Error(ALENS_E_ARG, "cannot write " + folder + "/TimeStepInfo.txt");
        ++snapID_;
    }
    // reinitialize(): resume from the snapshot named by a restart record (SRC/main.cpp:28-33,
    // SRC/TubuleSystem.cpp:38): rods from its file, step count and snapshot id one past the record's.
    // The file keeps the rods in the order the writing system held them (already spatially sorted),
    // so by default they are not sorted again and the constraint rows come out in the same order.
    void reinitialize(const std::string &restartFile, bool spatialSort = false) {
        std::ifstream in(restartFile);
        if (!in) throw Error(ALENS_E_ARG, "cannot open restart record " + restartFile);
        std::map<std::string, std::string> kv;
        std::string key, val;
        while (in >> key >> val) kv[key] = val;
        for (const char *k : {"snapID", "stepCount", "snapFile"})
            if (!kv.count(k)) throw Error(ALENS_E_ARG, restartFile + ": missing '" + k + "'");
        const size_t slash = restartFile.find_last_of('/');
        const std::string folder = slash == std::string::npos ? "." : restartFile.substr(0, slash);
        initialize(readRodFile(folder + "/" + kv["snapFile"]), spatialSort);
        check(ctx_, alens_set_step(ctx_, std::stoll(kv["stepCount"]) + 1), "set_step");
        snapID_ = std::stoi(kv["snapID"]) + 1;
    }
    int getSnapID() const { return snapID_; }
    // prepareStep(): clears the constraint pool for this step's external blocks
    void prepareStep() {
        for (auto &q : pool_) q.clear();
    }
    // runStep(): Brownian + non-constraint velocity, collision collection, BBPGD over the collision
    // rows and the pool's blocks (queues concatenated in thread order), write-back
    void runStep(const double *forceTorque = nullptr) {
        requireEngine();
        check(ctx_, alens_clear_constraints(ctx_), "clear_constraints");
        std::vector<alens_con_block> raw;
        for (const auto &q : pool_)
            for (const auto &b : q) raw.push_back(b.toC());
        if (!raw.empty()) check(ctx_, alens_add_constraints(ctx_, (int64_t)raw.size(), raw.data()), "add_constraints");
        alens_status s = alens_step_constraints(ctx_, forceTorque, &report_);
        if (s != ALENS_OK && s != ALENS_E_NOCONV) check(ctx_, s, "step_constraints");
        stepped_ = true;
    }
    // calcConStress(): gamma-weighted stress sums of the last step (collision rows, pool rows)
    void calcConStress(double collision[9], double pool[9]) {
        check(ctx_, alens_get_con_stress(ctx_, collision, pool), "get_con_stress");
    }
    // stepEuler(): move the rods with the last step's velocities (SRC/TubuleSystem.cpp:55).  Right
    // after (re)initialize, before any runStep, there is no velocity to apply and it returns at once.
    void stepEuler() {
        requireEngine();
        if (!stepped_) return;
        lastBlocks_ = getCollisionBlocksC();
        check(ctx_, alens_step_euler(ctx_), "step_euler");
        stepped_ = false;
    }

    ConstraintBlockPool &getConstraintPoolNonConst() { return pool_; }
    int64_t getStepCount() const { return alens_get_step(ctx_); }
    const alens_solve_report &getReport() const { return report_; }
    alens_ctx *handle() { return ctx_; }

    // this step's constraints with their impulses: collision rows in canonical order, then the pool's
    std::vector<alens_con_block> getCollisionBlocksC() {
        int64_t m = 0;
        check(ctx_, alens_get_num_constraints(ctx_, &m), "get_num_constraints");
        std::vector<alens_con_block> raw((size_t)m);
        if (m > 0) check(ctx_, alens_get_constraints(ctx_, m, raw.data(), &m), "get_constraints");
        return raw;
    }
    std::vector<ConstraintBlock> getCollisionBlocks() {
        const std::vector<alens_con_block> raw = getCollisionBlocksC();
        std::vector<ConstraintBlock> out;
        out.reserve(raw.size());
        for (const auto &c : raw) out.push_back(ConstraintBlock::fromC(c));
        return out;
    }
    // velocity (6 per rod: v, omega) after the step
    std::vector<double> getVelocity() {
        std::vector<double> u(6 * (size_t)n_);
        check(ctx_, alens_get_velocity(ctx_, u.data(), nullptr, nullptr, nullptr), "get_velocity");
        return u;
    }
    void getPositions(std::vector<double> &center, std::vector<double> &direction) {
        center.resize(3 * (size_t)n_);
        direction.resize(3 * (size_t)n_);
        check(ctx_, alens_get_rods(ctx_, center.data(), direction.data()), "get_rods");
    }

  private:
    void create(const alens_config &cfg) {
        if (ctx_) {
            alens_destroy(ctx_);
            ctx_ = nullptr;
        }
        alens_status s = alens_create(&cfg, &ctx_);
        if (s != ALENS_OK) {
            ctx_ = nullptr;
            throw Error(s, "alens_create failed (invalid configuration or no HIP device)");
        }
        stepped_ = false;
        lastBlocks_.clear();
    }
    alens_config engineBase() const {
        alens_config c{};
        c.precision = opt_.precision;
        c.warm_start = opt_.warm_start;
        c.device = opt_.device;
        c.cell_size = opt_.cell_size;
        return c;
    }
    void requireEngine() const {
        if (!ctx_) throw Error(ALENS_E_STATE, "SylinderSystem: not initialized");
    }

    alens_ctx *ctx_ = nullptr;
    EngineOptions opt_;
    bool stepped_ = false;                       // a runStep since the last Euler move
    std::vector<alens_con_block> lastBlocks_;    // blocks of the last step (for writeResult after stepEuler)
    int32_t n_ = 0;
    int snapID_ = 0;
    Rods rods_;  // per-rod data in the context's (sorted) order; positions live on the device
    ConstraintBlockPool pool_;
    alens_solve_report report_{};
};

// BCQPSolver: min 1/2 x^T A x + b^T x, x >= 0, A = dt D^T M D given by constraint rows (normal and the
// two moment arms posI x n, posJ x n) and per-rod mobility {t, 1/zperp, 1/zpar - 1/zperp, 1/zrot}.
class BCQPSolver {
  public:
    BCQPSolver(alens_ctx *ctx, int32_t nRods, std::vector<int32_t> I, std::vector<int32_t> J,
               std::vector<double> normal, std::vector<double> armI, std::vector<double> armJ, std::vector<double> b,
               std::vector<double> mobility)
        : ctx_(ctx), n_(nRods), I_(std::move(I)), J_(std::move(J)), nrm_(std::move(normal)), rI_(std::move(armI)),
          rJ_(std::move(armJ)), b_(std::move(b)), mob_(std::move(mobility)) {}
    // returns the iteration count; xsol holds the start point on entry (empty = cold)
    int solveBBPGD(std::vector<double> &xsol, alens_solve_report *history = nullptr) {
        const int64_t m = (int64_t)I_.size();
        std::vector<double> out((size_t)m);
        alens_solve_report rep{};
        alens_status s = alens_bcqp_solve(ctx_, n_, m, I_.data(), J_.data(), nrm_.data(), rI_.data(), rJ_.data(),
                                          b_.data(), mob_.data(), xsol.size() == (size_t)m ? xsol.data() : nullptr,
                                          out.data(), &rep);
        if (s != ALENS_OK && s != ALENS_E_NOCONV) check(ctx_, s, "bcqp_solve");
        xsol.swap(out);
        if (history) *history = rep;
        return rep.iters;
    }

  private:
    alens_ctx *ctx_;
    int32_t n_;
    std::vector<int32_t> I_, J_;
    std::vector<double> nrm_, rI_, rJ_, b_, mob_;
};

}  // namespace alens
