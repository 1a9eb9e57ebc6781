// SylinderSystem.hpp -- C++ call surface of the reference's rod engine over libalens.so's C ABI.
//
// The reference calls into SimToolbox's SylinderSystem / ConstraintCollector / BCQPSolver from
// TubuleSystem (SRC/TubuleSystem.cpp:79, 120, 223, 55, 697-742).  These header-only classes keep those
// names and call shapes so that code written against them moves onto the GPU engine unchanged:
//   SylinderSystem::initialize / prepareStep / runStep / stepEuler / getConstraintPoolNonConst /
//   getStepCount, the per-thread ConstraintBlockPool (pool.size() == thread count, one queue per
//   thread), ConstraintBlock's 16-argument constructor and setStress, and BCQPSolver::solveBBPGD;
//   rod files (TubuleInitial.dat / SylinderAscii_*.dat), writeResult, reinitialize and getSnapID.
// Errors surface as alens::Error exceptions carrying the alens_status and message.
#pragma once
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../alens_capi.h"

namespace alens {

struct Error : std::runtime_error {
    alens_status status;
    Error(alens_status s, const std::string &m) : std::runtime_error(m), status(s) {}
};

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

class SylinderSystem {
  public:
    SylinderSystem(const alens_config &cfg, int nThreads = 1) : pool_((size_t)(nThreads > 0 ? nThreads : 1)) {
        alens_status s = alens_create(&cfg, &ctx_);
        if (s != ALENS_OK) throw Error(s, "alens_create failed");
    }
    ~SylinderSystem() { alens_destroy(ctx_); }
    SylinderSystem(const SylinderSystem &) = delete;
    SylinderSystem &operator=(const SylinderSystem &) = delete;

    // initialize(): load rods; spatialSort reorders them the way FDPS keeps its particle arrays
    void initialize(const Rods &r, bool spatialSort = true) {
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
    // initialize(posFile): rods from a rod file (SRC/TubuleSystem.cpp:79 passes "TubuleInitial.dat")
    void initialize(const std::string &posFile, bool spatialSort = true) { initialize(readRodFile(posFile), spatialSort); }

    // writeResult(): SylinderAscii_<snapID>.dat with the current rods (17 significant digits) and the
    // restart record TimeStepInfo.txt ("key value" lines: snapID, stepCount, time, snapFile) in folder;
    // call it after stepEuler, so the file holds the positions the next step starts from.
    void writeResult(const std::string &folder, double time) {
        Rods r = rods_;
        getPositions(r.center, r.direction);
        const std::string name = "SylinderAscii_" + std::to_string(snapID_) + ".dat";
        writeRodFile(folder + "/" + name, r, time);
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
        check(ctx_, alens_clear_constraints(ctx_), "clear_constraints");
        std::vector<alens_con_block> raw;
        for (const auto &q : pool_)
            for (const auto &b : q) raw.push_back(b.toC());
        if (!raw.empty()) check(ctx_, alens_add_constraints(ctx_, (int64_t)raw.size(), raw.data()), "add_constraints");
        alens_status s = alens_step_constraints(ctx_, forceTorque, &report_);
        if (s != ALENS_OK && s != ALENS_E_NOCONV) check(ctx_, s, "step_constraints");
    }
    // calcConStress(): gamma-weighted stress sums of the last step (collision rows, pool rows)
    void calcConStress(double collision[9], double pool[9]) {
        check(ctx_, alens_get_con_stress(ctx_, collision, pool), "get_con_stress");
    }
    void stepEuler() { check(ctx_, alens_step_euler(ctx_), "step_euler"); }

    ConstraintBlockPool &getConstraintPoolNonConst() { return pool_; }
    int64_t getStepCount() const { return alens_get_step(ctx_); }
    const alens_solve_report &getReport() const { return report_; }
    alens_ctx *handle() { return ctx_; }

    // this step's constraints with their impulses: collision rows in canonical order, then the pool's
    std::vector<ConstraintBlock> getCollisionBlocks() {
        int64_t m = 0;
        check(ctx_, alens_get_num_constraints(ctx_, &m), "get_num_constraints");
        std::vector<alens_con_block> raw((size_t)m);
        check(ctx_, alens_get_constraints(ctx_, m, raw.data(), &m), "get_constraints");
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
    alens_ctx *ctx_ = nullptr;
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
