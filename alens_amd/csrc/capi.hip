// capi.hip -- the C ABI of libalens.so (include/alens_capi.h): context lifecycle and the per-step
// orchestration of the collision-constraint path on one GPU: plain steps (counts read back between
// phases), graph mode (the whole step captured once and replayed with device-side counts), caller
// constraint blocks, batched systems, the autotuner, the learned warm start and the BCQP entry point.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "alens_dev.hpp"
#include "alens_internal.hpp"

using namespace alens;

struct alens_ctx {
    alens_config cfg{};
    std::string err;
    hipStream_t st = nullptr;
    int dev = 0;
    int ncu = 0;
    int64_t step = 0;
    int32_t n = 0;
    bool have_rods = false;
    bool have_step = false;
    int grid_used = 0;
    // independent systems (1 for alens_set_rods): boxes, rod ranges, cell grids
    int32_t nsys = 0;
    std::vector<alens_box> boxes;
    std::vector<int32_t> rod_off;
    std::vector<SysDev> sysh;
    int64_t ncell = 0;
    int keybits = 1;
    DBuf<SysDev> sysd;
    DBuf<int32_t> rsys, smono, chunk_sys;
    DBuf<uint32_t> con_off, chunk_off;
    std::vector<alens_solve_report> sysrep;

    std::vector<int32_t> h_gid;
    std::unordered_map<int32_t, int32_t> gid_index;  // gid -> rod index (built on demand)
    bool gid_index_ok = false;
    DBuf<int32_t> gid;
    DBuf<double> center, dir, length, radius, hcol, rcol;
    DBuf<uint8_t> immov;
    bool has_immov = false;
    DBuf<double> mob, unc, U, F, Us, fext;

    DBuf<int32_t> clo, chi;
    DBuf<double> blo, bhi;
    DBuf<uint32_t> ecnt, eoff, keys, vals, keys2, vals2, cstart, cend, hist;
    ScanScratch scan;

    DBuf<uint32_t> counter, pI, pJ, pImg, cntI, cntJ, rowStart, jStart, cur, cur2, jlist;
    uint32_t pcap = 0;
    DBuf<int32_t> cI, cJ;
    DBuf<uint32_t> cImg;
    DBuf<double> delta0, geo, posI, posJ, b;
    DBuf<uint32_t> adjStart, sellStart, sellK, slen;
    DBuf<uint32_t> sord, spos;             // rod at each slice position (degree order within windows), inverse
    const uint32_t *sord_cur = nullptr, *spos_cur = nullptr;  // the last step's (NULL: identity)
    DBuf<double> sellW;
    uint64_t epad = 0;  // padded entry count of the sliced adjacency (component stride of sellW)

    DBuf<double2> xg;
    DBuf<double> gamma, part, out;
    // caller rows (the constraint pool): pending for the next step, and those of the last step
    std::vector<alens_con_block> ext_pending, ext_step;
    DBuf<uint32_t> extStart, extEnt;
    DBuf<double> extNrm, extX0, lb, invk;
    DBuf<float> lb32, invk32;
    uint32_t mcol = 0, mext = 0;  // collision rows, caller rows of the last step
    // precision form: FP32 copies of the operator data and the FP32 iterate
    DBuf<float> geo32, sellW32, mob32, b32, U32;
    DBuf<float2> xg32;
    // learned warm start: FP32 master parameters, normalisation, features of the step
    DBuf<float> mlpP, mlpNorm, feat;
    bool mlp_ready = false;
    DBuf<float> adam1, adam2;                // Adam moments of energy training (reset by alens_mlp_set)
    int adam_step = 0;
    DBuf<double2> xprev;  // previous-step gamma matched to this step's rows (x component)
    bool have_xprev = false;
    DBuf<uint32_t> bar;
    // solver schedule (the autotuner's choices; environment variables override them unless tuning)
    int32_t bar_mode = 3, xcd_map = 0, max_blocks = 0, entries = 0, group_width = 0, stream_hint = 0;
    uint32_t gw_rows = 0;  // row count of the step the group width was tuned on (0: not tuned)
    bool tuning = false;
    // device-resident step (graph mode): capacities, device counts, the captured step
    bool capture = false;                  // enqueueing into a graph: no host waits
    bool cap_solve_ev = false;             // the graph already reads the clock at solve start
    uint32_t gstride = 0;                  // row stride of geo (and the solver's row bound without cur_mdev)
    const uint32_t *cur_mdev = nullptr;    // row count of the current step buffers on the device
    const int32_t *cur_nchdev = nullptr;   // their chunk count on the device
    DBuf<double> out32;
    DBuf<uint32_t> dword;                  // [0] valid entries, [1] valid rows, [2] rows of the previous step
    DBuf<uint64_t> stamps;                 // wall-clock readings in a replay: start, collected, solve start, solved
    DBuf<double> tail;                     // a replay's counts, readings and reports, read back in one copy
    bool bar_clean = false;                // the captured step has zeroed the barrier words already
    int64_t clock_khz = 0;                 // rate of that clock
    DBuf<int32_t> nch_dev;
    DBuf<int64_t> step_dev;
    int graph_mode = 0;
    struct Caps {
        uint32_t E = 0, M = 0, epad = 0;
        int32_t nch = 0;
    } caps;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    int graph_key = -1;                    // start-point kind and f_ext presence the graph was captured with
    int64_t graph_steps = 0, graph_captures = 0, graph_capfail = 0, graph_overflows = 0;
    std::string graph_err;
    bool have_counts = false;              // E, m, epad hold a step's counts (planning input)

    DBuf<uint64_t> pkey;
    DBuf<double> pgam;
    uint32_t mprev = 0;
    bool have_prev = false;

    int64_t m = 0, E = 0;
    int32_t nchunks_cur = 0, nsys_cur = 1;
    hipEvent_t ev[6] = {};
};

namespace {

const char *kVersion = "alens-mi355x 0.1 (gfx950, FP64 bit-exact path)";

alens_status fail(alens_ctx *c, alens_status s, const std::string &msg) {
    if (c) c->err = msg;
    return s;
}

// degree-ordered slice positions in the sliced adjacency (ALENS_SLOT_ORDER=0 keeps index order)
static bool slot_order_on() {
    const char *v = getenv("ALENS_SLOT_ORDER");
    return !v || atoi(v) != 0;
}

alens_status hip_fail(alens_ctx *c, const HipErr &e) {
    char buf[512];
    snprintf(buf, sizeof buf, "HIP error %d (%s) in %s at %s:%d", (int)e.code, hipGetErrorString(e.code),
             e.what ? e.what : "?", e.file ? e.file : "?", e.line);
    c->err = buf;
    return e.code == hipErrorOutOfMemory ? ALENS_E_OOM : ALENS_E_HIP;
}

#define TRY_HIP(call)                                                                    \
    do {                                                                                 \
        hipError_t e_ = (call);                                                          \
        if (e_ != hipSuccess) return hip_fail(ctx, HipErr{e_, #call, __FILE__, __LINE__}); \
    } while (0)
#define TRY_K(call)                                   \
    do {                                              \
        HipErr e_ = (call);                           \
        if (!e_.ok()) return hip_fail(ctx, e_);       \
    } while (0)

CfgDev make_cfgdev(const alens_ctx *ctx) {
    CfgDev c{};
    for (int k = 0; k < 3; ++k) {
        c.box_lo[k] = ctx->cfg.box_lo[k];
        c.len[k] = ctx->cfg.box_hi[k] - ctx->cfg.box_lo[k];
        c.pbc[k] = ctx->cfg.pbc[k] ? 1 : 0;
    }
    c.mono = ctx->cfg.monolayer ? 1 : 0;
    c.col_buf = ctx->cfg.col_buf;
    c.dt = ctx->cfg.dt;
    c.kT = ctx->cfg.kT;
    c.mu = ctx->cfg.viscosity;
    c.seed = ctx->cfg.seed;
    c.step = ctx->step;
    return c;
}

// forget a captured step (its buffers may be reallocated, or its settings changed)
void drop_graph(alens_ctx *ctx) {
    if (ctx->gexec) (void)hipGraphExecDestroy(ctx->gexec);
    if (ctx->graph) (void)hipGraphDestroy(ctx->graph);
    ctx->gexec = nullptr;
    ctx->graph = nullptr;
    ctx->graph_key = -1;
}

RodsDev make_rods(const alens_ctx *ctx) {
    RodsDev r{};
    r.n = ctx->n;
    r.rsys = ctx->rsys.p;
    r.sys = ctx->sysd.p;
    r.gid = ctx->gid.p;
    r.center = ctx->center.p;
    r.dir = ctx->dir.p;
    r.hcol = ctx->hcol.p;
    r.rcol = ctx->rcol.p;
    return r;
}

bool finite_all(const double *p, size_t n) {
    for (size_t i = 0; i < n; ++i)
        if (!std::isfinite(p[i])) return false;
    return true;
}

template <typename T>
hipError_t d2h(alens_ctx *ctx, T *dst, const T *src, size_t count) {
    if (!dst || count == 0) return hipSuccess;
    return hipMemcpyAsync(dst, src, sizeof(T) * count, hipMemcpyDeviceToHost, ctx->st);
}

}  // namespace

extern "C" {

const char *alens_version(void) { return kVersion; }

const char *alens_last_error(const alens_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

alens_status alens_create(const alens_config *cfg, alens_ctx **out) {
    if (!cfg || !out) return ALENS_E_ARG;
    *out = nullptr;
    alens_ctx *ctx = new alens_ctx();
    ctx->cfg = *cfg;
    auto bad = [&](const char *m) {
        delete ctx;
        (void)m;
        return ALENS_E_ARG;
    };
    for (int k = 0; k < 3; ++k)
        if (!(cfg->box_hi[k] > cfg->box_lo[k])) return bad("box");
    if (!(cfg->dt > 0.0) || !(cfg->viscosity > 0.0) || !(cfg->tol > 0.0) || cfg->max_ite < 0 ||
        !(cfg->col_buf >= 0.0) || !(cfg->diam_col_ratio > 0.0) || !(cfg->len_col_ratio > 0.0))
        return bad("params");
    if (cfg->solver != 0) return bad("solver");
    if (cfg->precision != 0 && cfg->precision != 1) return bad("precision");
    if (cfg->warm_start < 0 || cfg->warm_start > 2) return bad("warm_start");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
        delete ctx;
        return ALENS_E_HIP;
    }
    if (cfg->device < 0 || cfg->device >= ndev) return bad("device");
    ctx->dev = cfg->device;
    if (hipSetDevice(ctx->dev) != hipSuccess || hipStreamCreate(&ctx->st) != hipSuccess) {
        delete ctx;
        return ALENS_E_HIP;
    }
    if (hipDeviceGetAttribute(&ctx->ncu, hipDeviceAttributeMultiprocessorCount, ctx->dev) != hipSuccess) ctx->ncu = 256;
    for (auto &e : ctx->ev)
        if (hipEventCreate(&e) != hipSuccess) {
            delete ctx;
            return ALENS_E_HIP;
        }
    *out = ctx;
    return ALENS_OK;
}

void alens_destroy(alens_ctx *ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->dev);
    if (ctx->st) (void)hipStreamSynchronize(ctx->st);
    auto rel = [](auto &b) { b.release(); };
    rel(ctx->gid); rel(ctx->center); rel(ctx->dir); rel(ctx->length); rel(ctx->radius); rel(ctx->hcol);
    rel(ctx->rcol); rel(ctx->immov); rel(ctx->mob); rel(ctx->unc); rel(ctx->U); rel(ctx->F); rel(ctx->Us);
    rel(ctx->fext); rel(ctx->clo); rel(ctx->chi); rel(ctx->blo); rel(ctx->bhi); rel(ctx->ecnt);
    rel(ctx->eoff); rel(ctx->keys); rel(ctx->vals); rel(ctx->keys2); rel(ctx->vals2); rel(ctx->cstart);
    rel(ctx->cend); rel(ctx->hist);
    for (auto &l : ctx->scan.lvl) rel(l);
    rel(ctx->counter); rel(ctx->pI); rel(ctx->pJ); rel(ctx->pImg); rel(ctx->cntI); rel(ctx->cntJ);
    rel(ctx->rowStart); rel(ctx->jStart); rel(ctx->cur); rel(ctx->cur2); rel(ctx->sord); rel(ctx->spos); rel(ctx->jlist); rel(ctx->cI); rel(ctx->cJ);
    rel(ctx->cImg); rel(ctx->delta0); rel(ctx->geo); rel(ctx->posI); rel(ctx->posJ); rel(ctx->b);
    rel(ctx->adjStart); rel(ctx->sellStart); rel(ctx->sellK); rel(ctx->slen); rel(ctx->sellW); rel(ctx->xg); rel(ctx->gamma); rel(ctx->part);
    rel(ctx->out); rel(ctx->bar); rel(ctx->pkey); rel(ctx->pgam);
    rel(ctx->out32); rel(ctx->dword); rel(ctx->nch_dev); rel(ctx->step_dev);
    if (ctx->gexec) (void)hipGraphExecDestroy(ctx->gexec);
    if (ctx->graph) (void)hipGraphDestroy(ctx->graph);
    rel(ctx->extStart); rel(ctx->extEnt); rel(ctx->extNrm); rel(ctx->extX0); rel(ctx->lb); rel(ctx->invk);
    rel(ctx->lb32); rel(ctx->invk32); rel(ctx->geo32); rel(ctx->sellW32); rel(ctx->mob32); rel(ctx->b32); rel(ctx->U32); rel(ctx->xg32);
    rel(ctx->mlpP); rel(ctx->mlpNorm); rel(ctx->feat); rel(ctx->xprev);
    rel(ctx->sysd); rel(ctx->rsys); rel(ctx->smono); rel(ctx->chunk_sys); rel(ctx->con_off); rel(ctx->chunk_off);
    for (auto &e : ctx->ev)
        if (e) (void)hipEventDestroy(e);
    if (ctx->st) (void)hipStreamDestroy(ctx->st);
    delete ctx;
}

alens_status alens_set_systems(alens_ctx *ctx, int32_t nsys, const int32_t *rod_offset, const alens_box *boxes,
                               const int32_t *gid, const double *center, const double *dir, const double *length,
                               const double *radius, const uint8_t *immovable) {
    if (ctx) {
        drop_graph(ctx);
        ctx->have_counts = false;
    }
    if (!ctx) return ALENS_E_ARG;
    if (nsys < 1 || nsys > kMaxSys || !rod_offset || !boxes || rod_offset[0] != 0)
        return fail(ctx, ALENS_E_ARG, "set_systems: need 1..256 systems and rod_offset[0] = 0");
    for (int32_t s = 0; s < nsys; ++s)
        if (rod_offset[s + 1] < rod_offset[s]) return fail(ctx, ALENS_E_ARG, "set_systems: rod_offset decreases");
    const int32_t n = rod_offset[nsys];
    if (n < 0 || (n > 0 && (!gid || !center || !dir || !length || !radius)))
        return fail(ctx, ALENS_E_ARG, "set_rods: null array or negative count");
    if (n > (1 << 30)) return fail(ctx, ALENS_E_ARG, "set_rods: too many rods");
    if (!finite_all(center, 3 * (size_t)n) || !finite_all(dir, 3 * (size_t)n) ||
        !finite_all(length, (size_t)n) || !finite_all(radius, (size_t)n))
        return fail(ctx, ALENS_E_NONFINITE, "set_rods: non-finite rod data");
    const alens_config &cf = ctx->cfg;
    std::vector<double> hc(n), rc(n);
    std::vector<int32_t> rs(n);
    std::vector<SysDev> sh(nsys);
    std::vector<int32_t> mono(nsys);
    int64_t ncell = 0;
    for (int32_t s = 0; s < nsys; ++s) {
        const alens_box &bx = boxes[s];
        for (int k = 0; k < 3; ++k)
            if (!(bx.box_hi[k] > bx.box_lo[k])) return fail(ctx, ALENS_E_ARG, "set_systems: empty box");
        double lmax = 0.0, rmax = 0.0, lsum = 0.0, imax = 0.0;
        const int32_t r0 = rod_offset[s], r1 = rod_offset[s + 1];
        for (int32_t i = r0; i < r1; ++i) {
            if (!(length[i] >= 0.0) || !(radius[i] > 0.0))
                return fail(ctx, ALENS_E_ARG, "set_rods: length must be >= 0 and radius > 0");
            hc[i] = (length[i] < radius[i] * 2.0) ? 0.0 : 0.5 * (length[i] * cf.len_col_ratio);
            rc[i] = radius[i] * cf.diam_col_ratio;
            rs[i] = s;
            lmax = std::max(lmax, 2.0 * hc[i]);
            rmax = std::max(rmax, rc[i]);
            lsum += 2.0 * hc[i];
            imax = std::max(imax, rc[i] + 0.5 * cf.col_buf + kAabbMargin);
        }
        for (int k = 0; k < 3; ++k) {
            if (!bx.pbc[k]) continue;
            if (!(bx.box_hi[k] - bx.box_lo[k] >= 2.0 * (lmax + 2.0 * rmax + cf.col_buf)))
                return fail(ctx, ALENS_E_BOX, "periodic box shorter than twice the largest interaction range");
        }
        // cell grid: edge = configured size, or a quarter of the mean collision length (at least twice
        // the largest inflation); capped at 8 cells per rod.  Rods enter only the cells their inflated
        // segment meets, so small cells keep few rods per cell.
        const int32_t ns = r1 - r0;
        double h = cf.cell_size;
        if (!(h > 0.0)) {
            h = ns ? 0.25 * lsum / ns : 1.0;
            if (h < 2.0 * imax) h = 2.0 * imax;
            if (!(h > 0.0)) h = 1.0;
        }
        const double maxcells = std::max<double>(8.0 * ns, 4096.0);
        SysDev &g = sh[s];
        for (;;) {
            double tot = 1.0;
            for (int k = 0; k < 3; ++k) {
                g.box_lo[k] = bx.box_lo[k];
                g.len[k] = bx.box_hi[k] - bx.box_lo[k];
                g.pbc[k] = bx.pbc[k] ? 1 : 0;
                int c = (int)std::floor(g.len[k] / h);
                if (c < 1) c = 1;
                g.nc[k] = c;
                g.h[k] = g.len[k] / c;
                tot *= c;
            }
            if (tot <= maxcells) break;
            h *= 1.25;
        }
        g.mono = bx.monolayer ? 1 : 0;
        g.infl = imax;
        g.cell_off = (uint32_t)ncell;
        g.rod_off = r0;
        g.rod_end = r1;
        mono[s] = g.mono;
        ncell += (int64_t)g.nc[0] * g.nc[1] * g.nc[2];
        if (ncell > (int64_t)0xFFFFFFFF) return fail(ctx, ALENS_E_ARG, "set_systems: too many cells");
    }
    int kb = 1;
    while (((int64_t)1 << kb) < ncell) ++kb;

    TRY_HIP(hipSetDevice(ctx->dev));
    const size_t N = (size_t)std::max(n, 1);
    TRY_HIP(ctx->gid.ensure(N));
    TRY_HIP(ctx->center.ensure(3 * N));
    TRY_HIP(ctx->dir.ensure(3 * N));
    TRY_HIP(ctx->length.ensure(N));
    TRY_HIP(ctx->radius.ensure(N));
    TRY_HIP(ctx->hcol.ensure(N));
    TRY_HIP(ctx->rcol.ensure(N));
    TRY_HIP(ctx->rsys.ensure(N));
    for (auto *bf : {&ctx->mob, &ctx->unc, &ctx->U, &ctx->F, &ctx->fext}) TRY_HIP(bf->ensure(6 * N));
    TRY_HIP(ctx->Us.ensure(kUStride * N));
    TRY_HIP(ctx->clo.ensure(3 * N));
    TRY_HIP(ctx->chi.ensure(3 * N));
    TRY_HIP(ctx->blo.ensure(3 * N));
    TRY_HIP(ctx->bhi.ensure(3 * N));
    TRY_HIP(ctx->ecnt.ensure(N + 1));
    TRY_HIP(ctx->eoff.ensure(N + 1));
    for (auto *bf : {&ctx->cntI, &ctx->cntJ, &ctx->rowStart, &ctx->jStart, &ctx->cur, &ctx->cur2, &ctx->adjStart})
        TRY_HIP(bf->ensure(N + 1));
    TRY_HIP(ctx->cstart.ensure((size_t)ncell));
    TRY_HIP(ctx->cend.ensure((size_t)ncell));
    TRY_HIP(ctx->counter.ensure(4));
    TRY_HIP(ctx->bar.ensure(kBarTotalWords));
    TRY_HIP(ctx->out.ensure(8 * (size_t)nsys + 8));
    TRY_HIP(ctx->sysd.ensure((size_t)nsys));
    TRY_HIP(ctx->smono.ensure((size_t)nsys));
    TRY_HIP(ctx->con_off.ensure((size_t)nsys + 1));
    TRY_HIP(ctx->chunk_off.ensure((size_t)nsys + 1));
    if (n > 0) {
        TRY_HIP(hipMemcpyAsync(ctx->gid.p, gid, sizeof(int32_t) * n, hipMemcpyHostToDevice, ctx->st));
        TRY_HIP(hipMemcpyAsync(ctx->center.p, center, sizeof(double) * 3 * n, hipMemcpyHostToDevice, ctx->st));
        TRY_HIP(hipMemcpyAsync(ctx->dir.p, dir, sizeof(double) * 3 * n, hipMemcpyHostToDevice, ctx->st));
        TRY_HIP(hipMemcpyAsync(ctx->length.p, length, sizeof(double) * n, hipMemcpyHostToDevice, ctx->st));
        TRY_HIP(hipMemcpyAsync(ctx->radius.p, radius, sizeof(double) * n, hipMemcpyHostToDevice, ctx->st));
        TRY_HIP(hipMemcpyAsync(ctx->hcol.p, hc.data(), sizeof(double) * n, hipMemcpyHostToDevice, ctx->st));
        TRY_HIP(hipMemcpyAsync(ctx->rcol.p, rc.data(), sizeof(double) * n, hipMemcpyHostToDevice, ctx->st));
        TRY_HIP(hipMemcpyAsync(ctx->rsys.p, rs.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, ctx->st));
    }
    TRY_HIP(hipMemcpyAsync(ctx->sysd.p, sh.data(), sizeof(SysDev) * nsys, hipMemcpyHostToDevice, ctx->st));
    TRY_HIP(hipMemcpyAsync(ctx->smono.p, mono.data(), sizeof(int32_t) * nsys, hipMemcpyHostToDevice, ctx->st));
    ctx->has_immov = immovable != nullptr;
    if (immovable && n > 0) {
        TRY_HIP(ctx->immov.ensure(N));
        TRY_HIP(hipMemcpyAsync(ctx->immov.p, immovable, n, hipMemcpyHostToDevice, ctx->st));
    }
    // the previous-step impulses stay valid only for the same rods in the same order
    std::vector<int32_t> ng(gid, gid + n);
    if (ng != ctx->h_gid || nsys != ctx->nsys) ctx->have_prev = false;
    ctx->h_gid.swap(ng);
    ctx->gid_index_ok = false;
    TRY_HIP(hipStreamSynchronize(ctx->st));
    ctx->n = n;
    ctx->nsys = nsys;
    ctx->boxes.assign(boxes, boxes + nsys);
    ctx->rod_off.assign(rod_offset, rod_offset + nsys + 1);
    ctx->sysh = sh;
    ctx->ncell = ncell;
    ctx->keybits = kb;
    ctx->sysrep.assign((size_t)nsys, alens_solve_report{});
    ctx->have_rods = true;
    ctx->have_step = false;
    return ALENS_OK;
}

alens_status alens_set_rods(alens_ctx *ctx, int32_t n, const int32_t *gid, const double *center,
                            const double *dir, const double *length, const double *radius,
                            const uint8_t *immovable) {
    if (!ctx) return ALENS_E_ARG;
    if (n < 0) return fail(ctx, ALENS_E_ARG, "set_rods: null array or negative count");
    alens_box bx{};
    for (int k = 0; k < 3; ++k) {
        bx.box_lo[k] = ctx->cfg.box_lo[k];
        bx.box_hi[k] = ctx->cfg.box_hi[k];
        bx.pbc[k] = ctx->cfg.pbc[k];
    }
    bx.monolayer = ctx->cfg.monolayer;
    const int32_t off[2] = {0, n};
    return alens_set_systems(ctx, 1, off, &bx, gid, center, dir, length, radius, immovable);
}

alens_status alens_get_rods(alens_ctx *ctx, double *center, double *dir) {
    if (!ctx) return ALENS_E_ARG;
    if (!ctx->have_rods) return fail(ctx, ALENS_E_STATE, "get_rods before set_rods");
    TRY_HIP(hipSetDevice(ctx->dev));
    if (ctx->n > 0) {
        if (center) TRY_HIP(hipMemcpyAsync(center, ctx->center.p, sizeof(double) * 3 * ctx->n, hipMemcpyDeviceToHost, ctx->st));
        if (dir) TRY_HIP(hipMemcpyAsync(dir, ctx->dir.p, sizeof(double) * 3 * ctx->n, hipMemcpyDeviceToHost, ctx->st));
    }
    TRY_HIP(hipStreamSynchronize(ctx->st));
    return ALENS_OK;
}

alens_status alens_sort_rods(alens_ctx *ctx, int32_t *perm_out) {
    if (ctx) drop_graph(ctx);
    if (!ctx) return ALENS_E_ARG;
    if (!ctx->have_rods) return fail(ctx, ALENS_E_STATE, "sort_rods before set_rods");
    TRY_HIP(hipSetDevice(ctx->dev));
    hipStream_t st = ctx->st;
    const int32_t n = ctx->n;
    if (n <= 1) {
        if (perm_out && n == 1) perm_out[0] = 0;
        return ALENS_OK;
    }
    const size_t N = (size_t)n;
    TRY_HIP(ctx->keys.ensure(N));
    TRY_HIP(ctx->vals.ensure(N));
    TRY_HIP(ctx->keys2.ensure(N));
    TRY_HIP(ctx->vals2.ensure(N));
    const RodsDev rd = make_rods(ctx);
    DBuf<uint32_t> permd;
    TRY_HIP(permd.ensure(N));
    for (int32_t sy = 0; sy < ctx->nsys; ++sy) {
        const int32_t r0 = ctx->rod_off[sy], ns = ctx->rod_off[sy + 1] - r0;
        if (ns == 0) continue;
        TRY_K(launch_morton_keys(st, rd, r0, ns, ctx->keys.p, ctx->vals.p));
        bool alt = false;
        TRY_K(radix_sort_pairs(st, ctx->keys.p, ctx->vals.p, ctx->keys2.p, ctx->vals2.p, ns, 30, ctx->hist,
                               ctx->scan, alt));
        TRY_HIP(hipMemcpyAsync(permd.p + r0, alt ? ctx->vals2.p : ctx->vals.p, sizeof(uint32_t) * ns,
                               hipMemcpyDeviceToDevice, st));
    }
    const uint32_t *perm = permd.p;
    DBuf<int32_t> gid2;
    DBuf<double> center2, dir2, length2, radius2, hcol2, rcol2;
    DBuf<uint8_t> immov2;
    TRY_HIP(gid2.ensure(N));
    TRY_HIP(center2.ensure(3 * N));
    TRY_HIP(dir2.ensure(3 * N));
    TRY_HIP(length2.ensure(N));
    TRY_HIP(radius2.ensure(N));
    TRY_HIP(hcol2.ensure(N));
    TRY_HIP(rcol2.ensure(N));
    if (ctx->has_immov) TRY_HIP(immov2.ensure(N));
    TRY_K(launch_permute_rods(st, n, perm, ctx->gid.p, ctx->center.p, ctx->dir.p, ctx->length.p, ctx->radius.p,
                              ctx->hcol.p, ctx->rcol.p, ctx->has_immov ? ctx->immov.p : nullptr, gid2.p, center2.p,
                              dir2.p, length2.p, radius2.p, hcol2.p, rcol2.p,
                              ctx->has_immov ? immov2.p : nullptr));
    std::vector<uint32_t> hp(N);
    TRY_HIP(hipMemcpyAsync(hp.data(), perm, sizeof(uint32_t) * N, hipMemcpyDeviceToHost, st));
    TRY_HIP(hipStreamSynchronize(st));
    std::swap(ctx->gid, gid2);
    std::swap(ctx->center, center2);
    std::swap(ctx->dir, dir2);
    std::swap(ctx->length, length2);
    std::swap(ctx->radius, radius2);
    std::swap(ctx->hcol, hcol2);
    std::swap(ctx->rcol, rcol2);
    if (ctx->has_immov) std::swap(ctx->immov, immov2);
    gid2.release();
    center2.release();
    dir2.release();
    length2.release();
    radius2.release();
    hcol2.release();
    rcol2.release();
    immov2.release();
    permd.release();
    std::vector<int32_t> ng(N);
    for (size_t i = 0; i < N; ++i) {
        ng[i] = ctx->h_gid[hp[i]];
        if (perm_out) perm_out[i] = (int32_t)hp[i];
    }
    ctx->h_gid.swap(ng);
    ctx->gid_index_ok = false;
    ctx->have_prev = false;
    ctx->have_step = false;
    return ALENS_OK;
}

alens_status alens_set_step(alens_ctx *ctx, int64_t step) {
    if (!ctx) return ALENS_E_ARG;
    ctx->step = step;
    return ALENS_OK;
}
int64_t alens_get_step(const alens_ctx *ctx) { return ctx ? ctx->step : -1; }

}  // extern "C"

// Rod-major adjacency of arbitrary rows (I, J need not be ordered), built on the host in the sliced
// layout the solver reads (see launch_adjacency): each rod's entries in ascending constraint index,
// the I side with (n, r_i), the J side with (-n, -r_j) and the top bit set on its index; geometry
// component-major (geo[q * m + k]).
struct HostAdjacency {
    std::vector<uint32_t> start, sstart, sk;
    std::vector<double> sw, geo;
    uint64_t ep = 0;
    // padded slot count only (no fill), for a capacity check before anything is copied
    static uint64_t slots(int32_t n, int64_t m, const int32_t *I, const int32_t *J, std::vector<uint32_t> &deg) {
        deg.assign((size_t)n, 0);
        for (int64_t k = 0; k < m; ++k) {
            deg[I[k]]++;
            deg[J[k]]++;
        }
        uint64_t ep = 0;
        for (size_t sl = 0; sl * kSlice < (size_t)n; ++sl) {
            uint32_t dmax = 0;
            for (size_t i = sl * kSlice; i < std::min((sl + 1) * kSlice, (size_t)n); ++i) dmax = std::max(dmax, deg[i]);
            ep += (uint64_t)dmax * kSlice;
        }
        return ep;
    }
    void build(int32_t n, int64_t m, const int32_t *I, const int32_t *J, const double *nrm, const double *r_i,
               const double *r_j) {
        const size_t M = (size_t)m;
        start.assign((size_t)n + 1, 0);
        geo.assign(9 * M, 0.0);
        for (int64_t k = 0; k < m; ++k) {
            start[I[k] + 1]++;
            start[J[k] + 1]++;
            for (int a = 0; a < 3; ++a) {
                geo[a * M + k] = nrm[3 * k + a];
                geo[(3 + a) * M + k] = r_i[3 * k + a];
                geo[(6 + a) * M + k] = r_j[3 * k + a];
            }
        }
        for (int32_t i = 0; i < n; ++i) start[i + 1] += start[i];
        const size_t nsl = ((size_t)n + kSlice - 1) / kSlice;
        sstart.assign(nsl + 1, 0);
        for (size_t sl = 0; sl < nsl; ++sl) {
            uint32_t dmax = 0;
            for (size_t i = sl * kSlice; i < std::min((sl + 1) * kSlice, (size_t)n); ++i)
                dmax = std::max(dmax, start[i + 1] - start[i]);
            sstart[sl + 1] = sstart[sl] + dmax * (uint32_t)kSlice;
        }
        ep = sstart[nsl];
        sk.assign(std::max<uint64_t>(ep, 1), 0);
        sw.assign(6 * std::max<uint64_t>(ep, 1), 0.0);
        std::vector<uint32_t> fill((size_t)n, 0);
        // entries go in ascending k per rod; k runs in order, so each rod's slots fill left to right
        for (int64_t k = 0; k < m; ++k) {
            for (int side = 0; side < 2; ++side) {
                const int32_t i = side == 0 ? I[k] : J[k];
                const size_t idx = sstart[(size_t)i / kSlice] + (size_t)fill[i]++ * kSlice + (size_t)i % kSlice;
                sk[idx] = side == 0 ? (uint32_t)k : ((uint32_t)k | 0x80000000u);
                const double *r = side == 0 ? r_i : r_j;
                for (int a = 0; a < 3; ++a) {
                    sw[a * ep + idx] = side == 0 ? nrm[3 * k + a] : -nrm[3 * k + a];
                    sw[(3 + a) * ep + idx] = side == 0 ? r[3 * k + a] : -r[3 * k + a];
                }
            }
        }
    }
};

// Per-system constraint offsets -> reduction-chunk offsets and the chunk -> system table.
static alens_status set_chunk_table(alens_ctx *ctx, const std::vector<uint32_t> &coff) {
    const int32_t ns = (int32_t)coff.size() - 1;
    std::vector<uint32_t> choff((size_t)ns + 1, 0);
    for (int32_t sy = 0; sy < ns; ++sy)
        choff[sy + 1] = choff[sy] + (coff[sy + 1] - coff[sy] + kRedChunk - 1) / kRedChunk;
    std::vector<int32_t> csys(std::max<uint32_t>(choff[ns], 1));
    for (int32_t sy = 0; sy < ns; ++sy)
        for (uint32_t c = choff[sy]; c < choff[sy + 1]; ++c) csys[c] = sy;
    TRY_HIP(ctx->con_off.ensure((size_t)ns + 1));
    TRY_HIP(ctx->chunk_off.ensure((size_t)ns + 1));
    TRY_HIP(ctx->chunk_sys.ensure(csys.size()));
    TRY_HIP(hipMemcpyAsync(ctx->con_off.p, coff.data(), sizeof(uint32_t) * (ns + 1), hipMemcpyHostToDevice, ctx->st));
    TRY_HIP(hipMemcpyAsync(ctx->chunk_off.p, choff.data(), sizeof(uint32_t) * (ns + 1), hipMemcpyHostToDevice,
                           ctx->st));
    TRY_HIP(hipMemcpyAsync(ctx->chunk_sys.p, csys.data(), sizeof(int32_t) * csys.size(), hipMemcpyHostToDevice,
                           ctx->st));
    TRY_HIP(hipStreamSynchronize(ctx->st));
    ctx->nchunks_cur = (int32_t)choff[ns];
    ctx->nsys_cur = ns;
    return ALENS_OK;
}

// Per-system report words of a finished solve -> per-system reports and their aggregate (max
// iterations / residuals, all converged, any fallback, worst status).
// Per-system report words (8 per system, host copy) into the step report and the system reports.
static alens_status parse_reports(alens_ctx *ctx, const double *o, uint32_t abort_word, alens_solve_report *rep,
                                  std::vector<alens_solve_report> *per_sys) {
    const int32_t ns = ctx->nsys_cur;
    if (abort_word != 0) return fail(ctx, ALENS_E_HIP, "BBPGD grid barrier timed out (blocks not co-resident)");
    *rep = alens_solve_report{};
    rep->converged = 1;
    if (per_sys) per_sys->assign((size_t)ns, alens_solve_report{});
    for (int32_t sy = 0; sy < ns; ++sy) {
        const double *w = o + 8 * (size_t)sy;
        alens_solve_report r{};
        r.iters = (int32_t)w[0];
        r.matvecs = (int32_t)w[1];
        r.converged = (int32_t)w[2];
        r.fallback_used = (int32_t)w[3];
        r.status = (int32_t)w[4];
        r.res0 = w[5];
        r.res = w[6];
        r.n_con = (int64_t)w[7];
        if (per_sys) (*per_sys)[sy] = r;
        rep->iters = std::max(rep->iters, r.iters);
        rep->matvecs = std::max(rep->matvecs, r.matvecs);
        rep->converged &= r.converged;
        rep->fallback_used |= r.fallback_used;
        if (r.status == ALENS_E_NONFINITE || (r.status == ALENS_E_NOCONV && rep->status == 0)) rep->status = r.status;
        rep->res0 = std::max(rep->res0, r.res0);
        rep->res = std::max(rep->res, r.res);
    }
    return ALENS_OK;
}

static alens_status read_reports(alens_ctx *ctx, const double *out_dev, const uint32_t *bar, alens_solve_report *rep,
                                 std::vector<alens_solve_report> *per_sys) {
    std::vector<double> o(8 * (size_t)ctx->nsys_cur);
    uint32_t abort_word = 0;
    TRY_HIP(hipMemcpyAsync(o.data(), out_dev, sizeof(double) * o.size(), hipMemcpyDeviceToHost, ctx->st));
    TRY_HIP(hipMemcpyAsync(&abort_word, bar + kBarWords, sizeof abort_word, hipMemcpyDeviceToHost, ctx->st));
    TRY_HIP(hipStreamSynchronize(ctx->st));
    return parse_reports(ctx, o.data(), abort_word, rep, per_sys);
}

// The solver's working memory: chunk partials, per-system report words, barrier words and the chunk
// table (the context's buffers, or a caller's memory for the allocation-free entry point).
struct SolveMem {
    double *part, *out;
    uint32_t *bar;
    const uint32_t *con_off, *chunk_off;
    const int32_t *chunk_sys;
};

template <typename T>
static alens_status run_bbpgd_on(alens_ctx *ctx, SolveArgsT<T> a, const SolveMem &mem, alens_solve_report *rep,
                                 float *ms_solve, std::vector<alens_solve_report> *per_sys);

// Launches one persistent BBPGD solve on the context's stream.  Outside graph capture it waits for
// it, times it and reads the report; while a step is being captured it only enqueues (the FP64
// solve reports into ctx->out, the FP32 one into ctx->out32, read after the graph has run).
template <typename T>
static alens_status run_bbpgd(alens_ctx *ctx, SolveArgsT<T> a, alens_solve_report *rep, float *ms_solve,
                              std::vector<alens_solve_report> *per_sys) {
    const int32_t ns = ctx->nsys_cur;
    TRY_HIP(ctx->part.ensure((size_t)kRedQ * std::max<int32_t>(ctx->nchunks_cur, 1)));
    TRY_HIP(ctx->out.ensure(8 * (size_t)ns + 8));
    TRY_HIP(ctx->out32.ensure(8 * (size_t)ns + 8));
    const SolveMem mem{ctx->part.p, sizeof(T) == 8 ? ctx->out.p : ctx->out32.p, ctx->bar.p,
                       ctx->con_off.p, ctx->chunk_off.p, ctx->chunk_sys.p};
    return run_bbpgd_on<T>(ctx, a, mem, rep, ms_solve, per_sys);
}

template <typename T>
static alens_status run_bbpgd_on(alens_ctx *ctx, SolveArgsT<T> a, const SolveMem &mem, alens_solve_report *rep,
                                 float *ms_solve, std::vector<alens_solve_report> *per_sys) {
    const int32_t ns = ctx->nsys_cur;
    if (ctx->capture && ctx->bar_clean) {
        ctx->bar_clean = false;  // the captured step zeroed them with its other counters
    } else if (ctx->capture) {
        // a later launch of the same captured step (the FP64 refinement after the FP32 solve): zero
        // every word but the abort word, so an abort of the earlier launch reaches the step's tail
        // (and makes this launch leave at once instead of refining a partial iterate)
        ZeroJobs z{};
        z.p[0] = mem.bar;
        z.n[0] = kBarWords;
        z.p[1] = mem.bar + kBarWords + 1;
        z.n[1] = kBarTotalWords - kBarWords - 1;
        z.count = 2;
        TRY_K(launch_zero_many(ctx->st, z));
    } else {
        TRY_K(launch_fill_u32(ctx->st, mem.bar, 0u, kBarTotalWords));
    }
    double *out = mem.out;
    a.part = mem.part;
    a.bar = mem.bar;
    a.out = out;
    a.con_off = mem.con_off;
    a.chunk_off = mem.chunk_off;
    a.chunk_sys = mem.chunk_sys;
    a.nsys = ns;
    a.dt = ctx->cfg.dt;
    a.tol = ctx->cfg.tol;
    a.max_ite = ctx->cfg.max_ite;
    a.nchunks = ctx->nchunks_cur;
    a.nch_dev = ctx->cur_nchdev;
    a.m_dev = ctx->cur_mdev;
    // schedule: the context's (autotuned) choice, or the environment's (ALENS_SOLVER_MAX_BLOCKS,
    // ALENS_BARRIER_MODE, ALENS_XCD_MAP, ALENS_SOLVER_ENTRIES, ALENS_GROUP_WIDTH, ALENS_STREAM_HINT)
    // outside of tuning
    auto env = [&](const char *name, int dflt) {
        const char *v = ctx->tuning ? nullptr : getenv(name);
        return v ? atoi(v) : dflt;
    };
    const int max_blocks = env("ALENS_SOLVER_MAX_BLOCKS", ctx->max_blocks);
    a.bar_mode = env("ALENS_BARRIER_MODE", ctx->bar_mode);
    if (a.bar_mode < 0 || a.bar_mode > 3) a.bar_mode = 0;
    a.xcd_map = env("ALENS_XCD_MAP", ctx->xcd_map) != 0;
    a.split_ok = env("ALENS_SOLVER_SPLIT", 1) != 0;
    a.group_width = env("ALENS_GROUP_WIDTH", ctx->group_width);
    if (a.group_width != 128 && a.group_width != 256) a.group_width = 0;
    // a tuned width holds within 10% of the tuned step's row count (one set by the environment, or by
    // the tuner while it times candidates, holds for every row count)
    const bool gw_band = !ctx->tuning && !getenv("ALENS_GROUP_WIDTH") && ctx->gw_rows > 0;
    a.gw_rows_lo = gw_band ? (uint32_t)(0.9 * ctx->gw_rows) : 0u;
    a.gw_rows_hi = gw_band ? (uint32_t)std::min(4294967295.0, 1.1 * ctx->gw_rows) : 0u;
    // row streams of phase B: non-temporal when an iteration's bytes (per constraint 2 (9w + 8) + 5w,
    // per rod 18w) overflow the GPU's last-level cache, which otherwise keeps them between iterations
    // (a resident step's a.m is its row capacity; its last row count is the host's ctx->m)
    const int hint = env("ALENS_STREAM_HINT", ctx->stream_hint);
    const double rows = a.m_dev ? (double)ctx->m : (double)a.m;
    const double iter_bytes = (double)sizeof(T) * (23.0 * rows + 18.0 * (double)a.n) + 16.0 * rows;
    a.stream_nt = hint == 2 || (hint != 1 && iter_bytes > kLastLevelCacheBytes);
    const int entries = env("ALENS_SOLVER_ENTRIES", ctx->entries);
    // inside a graph the timing points are clock readings: stamp 2 before the first solver launch,
    // stamp 3 after each one (the last reading is the one read after the replay)
    if (!ctx->capture) {
        TRY_HIP(hipEventRecord(ctx->ev[2], ctx->st));
    } else if (!ctx->cap_solve_ev) {
        TRY_K(launch_stamp(ctx->st, ctx->stamps.p, 2));
        ctx->cap_solve_ev = true;
    }
    if constexpr (sizeof(T) == 8)
        TRY_K(launch_bbpgd64(ctx->st, a, max_blocks, entries, &ctx->grid_used));
    else
        TRY_K(launch_bbpgd32(ctx->st, a, max_blocks, entries, &ctx->grid_used));
    if (ctx->capture) {
        TRY_K(launch_stamp(ctx->st, ctx->stamps.p, 3));
        *ms_solve = 0.f;
        return ALENS_OK;
    }
    TRY_HIP(hipEventRecord(ctx->ev[3], ctx->st));
    alens_status s = read_reports(ctx, out, mem.bar, rep, per_sys);
    if (s != ALENS_OK) return s;
    TRY_HIP(hipEventElapsedTime(ms_solve, ctx->ev[2], ctx->ev[3]));
    return ALENS_OK;
}

// FP64 solve on the step buffers; xg holds the start point
static alens_status run_solver(alens_ctx *ctx, int32_t n, uint32_t m, const double *mob, bool warm,
                               alens_solve_report *rep, float *ms_solve, const int32_t *rsys, const int32_t *smono,
                               std::vector<alens_solve_report> *per_sys) {
    SolveArgsT<double> a{};
    a.n = n;
    a.m = ctx->gstride;  // row stride of geo: the row count, or a resident step's row capacity
    (void)m;
    a.adjStart = ctx->adjStart.p;
    a.sellStart = ctx->sellStart.p;
    a.sord = ctx->sord_cur;
    a.sellK = ctx->sellK.p;
    a.sellW = ctx->sellW.p;
    a.epad = ctx->epad;
    a.mob = mob;
    a.cI = ctx->cI.p;
    a.cJ = ctx->cJ.p;
    a.geo = ctx->geo.p;
    a.b = ctx->b.p;
    a.lb = ctx->mext ? ctx->lb.p : nullptr;
    a.invk = ctx->mext ? ctx->invk.p : nullptr;
    a.xg = ctx->xg.p;
    a.U = ctx->Us.p;
    a.rsys = rsys;
    a.smono = smono;
    a.warm = warm ? 1 : 0;
    return run_bbpgd<double>(ctx, a, rep, ms_solve, per_sys);
}

// Precision form: FP32 BBPGD from the start point in xg, then the FP64 solve started from the
// FP32 iterate (its residual is the FP64 residual; it refines until the FP64 residual meets tol).
static alens_status run_solver_mixed(alens_ctx *ctx, int32_t n, uint32_t m, bool warm, alens_solve_report *rep,
                                     float *ms_solve) {
    hipStream_t st = ctx->st;
    m = ctx->gstride;  // rows and row stride of the step buffers (a resident step: the capacity)
    const size_t M = std::max<uint32_t>(m, 1), N = (size_t)std::max(n, 1);
    TRY_HIP(ctx->geo32.ensure(9 * M));
    TRY_HIP(ctx->sellW32.ensure(6 * std::max<uint64_t>(ctx->epad, 1)));
    TRY_HIP(ctx->b32.ensure(M));
    TRY_HIP(ctx->xg32.ensure(M));
    TRY_HIP(ctx->mob32.ensure(6 * N));
    TRY_HIP(ctx->U32.ensure(kUStride * N));
    TRY_K(launch_to_f32(st, ctx->geo.p, ctx->geo32.p, 9 * (size_t)m));
    TRY_K(launch_to_f32(st, ctx->sellW.p, ctx->sellW32.p, 6 * (size_t)ctx->epad));
    TRY_K(launch_to_f32(st, ctx->b.p, ctx->b32.p, (size_t)m));
    TRY_K(launch_to_f32(st, ctx->mob.p, ctx->mob32.p, 6 * (size_t)n));
    if (ctx->mext) {
        TRY_HIP(ctx->lb32.ensure(M));
        TRY_HIP(ctx->invk32.ensure(M));
        TRY_K(launch_to_f32(st, ctx->lb.p, ctx->lb32.p, (size_t)m));
        TRY_K(launch_to_f32(st, ctx->invk.p, ctx->invk32.p, (size_t)m));
    }
    TRY_K(launch_init_xg32(st, ctx->xg32.p, ctx->xg.p, m));
    SolveArgsT<float> a{};
    a.n = n;
    a.m = m;  // row stride
    a.adjStart = ctx->adjStart.p;
    a.sellStart = ctx->sellStart.p;
    a.sord = ctx->sord_cur;
    a.sellK = ctx->sellK.p;
    a.sellW = ctx->sellW32.p;
    a.epad = ctx->epad;
    a.mob = ctx->mob32.p;
    a.cI = ctx->cI.p;
    a.cJ = ctx->cJ.p;
    a.geo = ctx->geo32.p;
    a.b = ctx->b32.p;
    a.lb = ctx->mext ? ctx->lb32.p : nullptr;
    a.invk = ctx->mext ? ctx->invk32.p : nullptr;
    a.xg = ctx->xg32.p;
    a.U = ctx->U32.p;
    a.rsys = ctx->rsys.p;
    a.smono = ctx->smono.p;
    a.warm = warm ? 1 : 0;
    alens_solve_report r32{};
    std::vector<alens_solve_report> s32;
    float ms32 = 0.f, ms64 = 0.f;
    alens_status s = run_bbpgd<float>(ctx, a, &r32, &ms32, &s32);
    if (s != ALENS_OK) return s;
    TRY_K(launch_xg32_to_x64(st, ctx->xg32.p, ctx->gamma.p, m));
    TRY_K(launch_init_xg(st, ctx->xg.p, ctx->gamma.p, m));
    s = run_solver(ctx, n, m, ctx->mob.p, true, rep, &ms64, ctx->rsys.p, ctx->smono.p, &ctx->sysrep);
    if (s != ALENS_OK) return s;
    if (ctx->capture) return ALENS_OK;  // both reports are read after the graph has run
    rep->iters_fp32 = r32.iters;
    for (size_t q = 0; q < s32.size() && q < ctx->sysrep.size(); ++q) ctx->sysrep[q].iters_fp32 = s32[q].iters;
    *ms_solve = ms32 + ms64;
    return ALENS_OK;
}

extern "C" {

// One step with exact buffer sizes: the counts (cell entries, pairs, adjacency slots) are read back
// between the phases.  Caller rows are supported here.
static alens_status step_plain(alens_ctx *ctx, const double *f_ext, alens_solve_report *rep_out) {
    alens_solve_report rep{};
    TRY_HIP(hipSetDevice(ctx->dev));
    hipStream_t st = ctx->st;
    const int32_t n = ctx->n;
    const CfgDev cd = make_cfgdev(ctx);
    const RodsDev rd = make_rods(ctx);
    // caller rows: rods by gid, per-rod entry lists in ascending row order
    const uint32_t Me = (uint32_t)ctx->ext_pending.size();
    std::vector<int32_t> eI(Me), eJ(Me);
    std::vector<double> eD0(Me), eNrm(3 * (size_t)Me), ePI(3 * (size_t)Me), ePJ(3 * (size_t)Me), eX0(Me), eLb(Me),
        eIk(Me);
    std::vector<uint32_t> eStart, eEnt;
    if (Me > 0) {
        if (ctx->nsys > 1) return fail(ctx, ALENS_E_STATE, "caller constraint blocks need a one-system context");
        if (!ctx->gid_index_ok) {
            ctx->gid_index.clear();
            ctx->gid_index.reserve((size_t)n * 2);
            for (int32_t i = 0; i < n; ++i) ctx->gid_index[ctx->h_gid[i]] = i;
            ctx->gid_index_ok = true;
        }
        eStart.assign((size_t)n + 1, 0);
        for (uint32_t e = 0; e < Me; ++e) {
            const alens_con_block &cb = ctx->ext_pending[e];
            auto fi = ctx->gid_index.find(cb.gid_i), fj = ctx->gid_index.find(cb.gid_j);
            if (fi == ctx->gid_index.end() || fj == ctx->gid_index.end() || fi->second == fj->second)
                return fail(ctx, ALENS_E_ARG, "constraint block: unknown gid or gid_i == gid_j");
            eI[e] = fi->second;
            eJ[e] = fj->second;
            bool ok = std::isfinite(cb.delta0) && std::isfinite(cb.gamma) && (cb.bilateral || std::isfinite(cb.gamma_lb));
            for (int a = 0; a < 3; ++a) {
                ok = ok && cb.norm_j[a] == -cb.norm_i[a] && std::isfinite(cb.norm_i[a]) && std::isfinite(cb.pos_i[a]) &&
                     std::isfinite(cb.pos_j[a]);
                eNrm[3 * (size_t)e + a] = cb.norm_i[a];
                ePI[3 * (size_t)e + a] = cb.pos_i[a];
                ePJ[3 * (size_t)e + a] = cb.pos_j[a];
            }
            if (!ok) return fail(ctx, ALENS_E_ARG, "constraint block: non-finite field or norm_j != -norm_i");
            eD0[e] = cb.delta0;
            eX0[e] = cb.gamma;
            eLb[e] = cb.bilateral ? -INFINITY : cb.gamma_lb;
            eIk[e] = (cb.bilateral && cb.kappa > 0.0) ? 1.0 / cb.kappa : 0.0;
            eStart[(size_t)eI[e] + 1]++;
            eStart[(size_t)eJ[e] + 1]++;
        }
        for (int32_t i = 0; i < n; ++i) eStart[(size_t)i + 1] += eStart[i];
        eEnt.resize(2 * (size_t)Me);
        std::vector<uint32_t> cur(eStart.begin(), eStart.end() - 1);
        for (uint32_t e = 0; e < Me; ++e) {
            eEnt[cur[eI[e]]++] = e;
            eEnt[cur[eJ[e]]++] = e | 0x80000000u;
        }
    }
    if (f_ext) {
        if (!finite_all(f_ext, 6 * (size_t)n)) return fail(ctx, ALENS_E_NONFINITE, "non-finite f_ext");
        if (n > 0) TRY_HIP(hipMemcpyAsync(ctx->fext.p, f_ext, sizeof(double) * 6 * n, hipMemcpyHostToDevice, st));
    }
    TRY_HIP(hipEventRecord(ctx->ev[0], st));
    TRY_K(launch_mob_unc(st, rd, ctx->length.p, ctx->radius.p, ctx->has_immov ? ctx->immov.p : nullptr,
                         f_ext ? ctx->fext.p : nullptr, cd, ctx->mob.p, ctx->unc.p));
    // cell list
    TRY_K(launch_aabb(st, rd, cd, ctx->clo.p, ctx->chi.p, ctx->blo.p, ctx->bhi.p, ctx->ecnt.p));
    TRY_K(exclusive_scan_u32(st, ctx->ecnt.p, ctx->eoff.p, n, ctx->scan));
    uint32_t E = 0;
    TRY_HIP(hipMemcpyAsync(&E, ctx->eoff.p + n, sizeof E, hipMemcpyDeviceToHost, st));
    TRY_HIP(hipStreamSynchronize(st));
    ctx->E = E;
    const size_t Ecap = std::max<size_t>(E, 1);
    TRY_HIP(ctx->keys.ensure(Ecap));
    TRY_HIP(ctx->vals.ensure(Ecap));
    TRY_HIP(ctx->keys2.ensure(Ecap));
    TRY_HIP(ctx->vals2.ensure(Ecap));
    TRY_K(launch_fill_entries(st, rd, ctx->clo.p, ctx->chi.p, ctx->eoff.p, ctx->keys.p, ctx->vals.p, E));
    bool alt = false;
    TRY_K(radix_sort_pairs(st, ctx->keys.p, ctx->vals.p, ctx->keys2.p, ctx->vals2.p, E, ctx->keybits,
                           ctx->hist, ctx->scan, alt));
    const uint32_t *skeys = alt ? ctx->keys2.p : ctx->keys.p;
    const uint32_t *svals = alt ? ctx->vals2.p : ctx->vals.p;
    TRY_HIP(hipMemsetAsync(ctx->cstart.p, 0, sizeof(uint32_t) * ctx->ncell, st));
    TRY_HIP(hipMemsetAsync(ctx->cend.p, 0, sizeof(uint32_t) * ctx->ncell, st));
    TRY_K(launch_cell_ranges(st, skeys, E, ctx->cstart.p, ctx->cend.p));
    // pair sweep (re-run once with a larger buffer if the pair buffer overflowed)
    if (ctx->pcap == 0) {
        const size_t c0 = std::max<size_t>(4096, 8 * (size_t)n);
        TRY_HIP(ctx->pI.ensure(c0));
        TRY_HIP(ctx->pJ.ensure(c0));
        TRY_HIP(ctx->pImg.ensure(c0));
        ctx->pcap = (uint32_t)ctx->pI.cap;
    }
    uint32_t M = 0;
    for (int attempt = 0; attempt < 2; ++attempt) {
        TRY_HIP(hipMemsetAsync(ctx->counter.p, 0, sizeof(uint32_t), st));
        TRY_HIP(hipMemsetAsync(ctx->cntI.p, 0, sizeof(uint32_t) * (n + 1), st));
        TRY_HIP(hipMemsetAsync(ctx->cntJ.p, 0, sizeof(uint32_t) * (n + 1), st));
        TRY_K(launch_sweep(st, rd, cd, skeys, E, ctx->cstart.p, ctx->cend.p, svals, ctx->clo.p, ctx->blo.p, ctx->bhi.p,
                           ctx->counter.p, ctx->pcap, ctx->pI.p, ctx->pJ.p, ctx->pImg.p, ctx->cntI.p,
                           ctx->cntJ.p));
        TRY_HIP(hipMemcpyAsync(&M, ctx->counter.p, sizeof M, hipMemcpyDeviceToHost, st));
        TRY_HIP(hipStreamSynchronize(st));
        if (M <= ctx->pcap) break;
        const size_t nc = (size_t)M + M / 4 + 1024;
        TRY_HIP(ctx->pI.ensure(nc));
        TRY_HIP(ctx->pJ.ensure(nc));
        TRY_HIP(ctx->pImg.ensure(nc));
        ctx->pcap = (uint32_t)std::min(ctx->pI.cap, std::min(ctx->pJ.cap, ctx->pImg.cap));
    }
    if (M > ctx->pcap) return fail(ctx, ALENS_E_HIP, "pair buffer overflow after resize");
    const uint32_t Mcol = M;
    M = Mcol + Me;  // rows of the step: collision rows, then the caller's
    ctx->m = M;
    ctx->mcol = Mcol;
    ctx->mext = Me;
    ctx->gstride = M;
    ctx->cur_mdev = nullptr;
    ctx->cur_nchdev = nullptr;
    const size_t Mc = std::max<size_t>(M, 1);
    TRY_HIP(ctx->cI.ensure(Mc));
    TRY_HIP(ctx->cJ.ensure(Mc));
    TRY_HIP(ctx->cImg.ensure(Mc));
    TRY_HIP(ctx->jlist.ensure(Mc));
    TRY_HIP(ctx->delta0.ensure(Mc));
    TRY_HIP(ctx->geo.ensure(9 * Mc));
    TRY_HIP(ctx->posI.ensure(3 * Mc));
    TRY_HIP(ctx->posJ.ensure(3 * Mc));
    TRY_HIP(ctx->b.ensure(Mc));
    TRY_HIP(ctx->xg.ensure(Mc));
    TRY_HIP(ctx->gamma.ensure(Mc));
    // canonical order + rows + adjacency
    TRY_K(exclusive_scan_u32(st, ctx->cntI.p, ctx->rowStart.p, n, ctx->scan));
    TRY_K(exclusive_scan_u32(st, ctx->cntJ.p, ctx->jStart.p, n, ctx->scan));
    TRY_HIP(hipMemsetAsync(ctx->cur.p, 0, sizeof(uint32_t) * (n + 1), st));
    TRY_K(launch_order_I(st, n, Mcol, ctx->pI.p, ctx->pJ.p, ctx->pImg.p, ctx->rowStart.p, ctx->cur.p, ctx->cI.p,
                         ctx->cJ.p, ctx->cImg.p));
    TRY_HIP(hipMemsetAsync(ctx->cur.p, 0, sizeof(uint32_t) * (n + 1), st));
    TRY_K(launch_order_J(st, n, Mcol, ctx->cJ.p, ctx->jStart.p, ctx->cur.p, ctx->jlist.p));
    TRY_K(launch_rows(st, rd, cd, Mcol, M, ctx->cI.p, ctx->cJ.p, ctx->cImg.p, ctx->unc.p, ctx->delta0.p, ctx->geo.p,
                      ctx->posI.p, ctx->posJ.p, ctx->b.p));
    if (Me > 0) {
        TRY_HIP(ctx->extStart.ensure((size_t)n + 1));
        TRY_HIP(ctx->extEnt.ensure(2 * (size_t)Me));
        TRY_HIP(ctx->extNrm.ensure(3 * (size_t)Me));
        TRY_HIP(ctx->extX0.ensure(Me));
        TRY_HIP(ctx->lb.ensure(Mc));
        TRY_HIP(ctx->invk.ensure(Mc));
        TRY_HIP(hipMemcpyAsync(ctx->cI.p + Mcol, eI.data(), sizeof(int32_t) * Me, hipMemcpyHostToDevice, st));
        TRY_HIP(hipMemcpyAsync(ctx->cJ.p + Mcol, eJ.data(), sizeof(int32_t) * Me, hipMemcpyHostToDevice, st));
        TRY_HIP(hipMemsetAsync(ctx->cImg.p + Mcol, 0, sizeof(uint32_t) * Me, st));
        TRY_HIP(hipMemcpyAsync(ctx->delta0.p + Mcol, eD0.data(), sizeof(double) * Me, hipMemcpyHostToDevice, st));
        TRY_HIP(hipMemcpyAsync(ctx->posI.p + 3 * (size_t)Mcol, ePI.data(), sizeof(double) * 3 * Me,
                               hipMemcpyHostToDevice, st));
        TRY_HIP(hipMemcpyAsync(ctx->posJ.p + 3 * (size_t)Mcol, ePJ.data(), sizeof(double) * 3 * Me,
                               hipMemcpyHostToDevice, st));
        TRY_HIP(hipMemcpyAsync(ctx->extNrm.p, eNrm.data(), sizeof(double) * 3 * Me, hipMemcpyHostToDevice, st));
        TRY_HIP(hipMemcpyAsync(ctx->extX0.p, eX0.data(), sizeof(double) * Me, hipMemcpyHostToDevice, st));
        TRY_HIP(hipMemcpyAsync(ctx->extStart.p, eStart.data(), sizeof(uint32_t) * (n + 1), hipMemcpyHostToDevice, st));
        TRY_HIP(hipMemcpyAsync(ctx->extEnt.p, eEnt.data(), sizeof(uint32_t) * 2 * Me, hipMemcpyHostToDevice, st));
        if (Mcol > 0) {
            TRY_HIP(hipMemsetAsync(ctx->lb.p, 0, sizeof(double) * Mcol, st));  // 0.0 = all-zero bytes
            TRY_HIP(hipMemsetAsync(ctx->invk.p, 0, sizeof(double) * Mcol, st));
        }
        TRY_HIP(hipMemcpyAsync(ctx->lb.p + Mcol, eLb.data(), sizeof(double) * Me, hipMemcpyHostToDevice, st));
        TRY_HIP(hipMemcpyAsync(ctx->invk.p + Mcol, eIk.data(), sizeof(double) * Me, hipMemcpyHostToDevice, st));
        TRY_K(launch_rows_ext(st, Me, Mcol, M, ctx->cI.p, ctx->cJ.p, ctx->extNrm.p, ctx->posI.p, ctx->posJ.p,
                              ctx->delta0.p, ctx->unc.p, ctx->cfg.dt, ctx->geo.p, ctx->b.p));
    }
    const uint32_t *xs = Me ? ctx->extStart.p : nullptr;
    {
        // sliced adjacency sizes; constraint range of each system = rowStart at its first rod (rows are
        // grouped by rod I; caller rows exist only with one system and close its range)
        const int64_t nsl = ((int64_t)n + kSlice - 1) / kSlice;
        TRY_HIP(ctx->slen.ensure((size_t)nsl + 1));
        TRY_HIP(ctx->sellStart.ensure((size_t)nsl + 1));
        ctx->sord_cur = ctx->spos_cur = nullptr;
        if (slot_order_on()) {
            TRY_HIP(ctx->sord.ensure((size_t)std::max(n, 1)));
            TRY_HIP(ctx->spos.ensure((size_t)std::max(n, 1)));
            TRY_K(launch_slot_order(st, n, ctx->rowStart.p, ctx->jStart.p, xs, ctx->sord.p, ctx->spos.p,
                                    slot_order_window(n, ctx->ncu)));
            ctx->sord_cur = ctx->sord.p;
            ctx->spos_cur = ctx->spos.p;
        }
        TRY_K(launch_slice_sizes(st, n, ctx->rowStart.p, ctx->jStart.p, xs, ctx->slen.p, ctx->sord_cur));
        TRY_K(exclusive_scan_u32(st, ctx->slen.p, ctx->sellStart.p, nsl, ctx->scan));
        uint32_t ep = 0;
        TRY_HIP(hipMemcpyAsync(&ep, ctx->sellStart.p + nsl, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
        std::vector<uint32_t> coff((size_t)ctx->nsys + 1);
        for (int32_t sy = 0; sy <= ctx->nsys; ++sy)
            TRY_HIP(hipMemcpyAsync(&coff[sy], ctx->rowStart.p + ctx->rod_off[sy], sizeof(uint32_t),
                                   hipMemcpyDeviceToHost, st));
        TRY_HIP(hipStreamSynchronize(st));
        coff[ctx->nsys] += Me;
        ctx->epad = ep;
        TRY_HIP(ctx->sellK.ensure(std::max<size_t>(ep, 1)));
        TRY_HIP(ctx->sellW.ensure(6 * std::max<size_t>(ep, 1)));
        TRY_K(launch_adjacency(st, n, M, ctx->rowStart.p, ctx->jStart.p, ctx->jlist.p, xs, Me ? ctx->extEnt.p : nullptr,
                               Mcol, ctx->geo.p, ctx->sellStart.p, ctx->epad, ctx->adjStart.p, ctx->sellK.p,
                               ctx->sellW.p, ctx->spos_cur));
        alens_status cs = set_chunk_table(ctx, coff);
        if (cs != ALENS_OK) return cs;
    }
    TRY_HIP(hipEventRecord(ctx->ev[1], st));
    // solve: start point = 0 (cold), the previous step's gamma (warm_start 1) or the network's
    // prediction from this step's features and the previous gamma (warm_start 2)
    float ms_solve = 0.0f;
    const int ws = ctx->cfg.warm_start;
    if (ws == 2 && !ctx->mlp_ready) return fail(ctx, ALENS_E_STATE, "warm_start = 2 needs network parameters");
    bool warm = false;
    ctx->have_xprev = false;
    if (M > 0) {
        // start point of the collision rows (caller rows start from their blocks' gamma)
        if (ws >= 1 && ctx->have_prev && Mcol > 0) {
            TRY_K(launch_warm_match(st, Mcol, ctx->cI.p, ctx->cJ.p, ctx->gid.p, ctx->mprev, ctx->pkey.p, ctx->pgam.p,
                                    ctx->xg.p));
            TRY_HIP(ctx->xprev.ensure(Mc));
            TRY_HIP(hipMemcpyAsync(ctx->xprev.p, ctx->xg.p, sizeof(double2) * Mcol, hipMemcpyDeviceToDevice, st));
            ctx->have_xprev = true;
            warm = true;
        } else {
            TRY_K(launch_init_xg(st, ctx->xg.p, nullptr, Mcol));
        }
        if (ws == 2 && Mcol > 0) {
            TRY_HIP(ctx->feat.ensure(16 * Mc));
            TRY_K(launch_features(st, Mcol, M, ctx->cI.p, ctx->cJ.p, ctx->delta0.p, ctx->geo.p, ctx->posI.p,
                                  ctx->posJ.p, ctx->dir.p, ctx->length.p, ctx->unc.p,
                                  ctx->have_xprev ? ctx->xprev.p : nullptr, ctx->cfg.dt, ctx->feat.p));
            TRY_K(launch_init_xg(st, ctx->xg.p, nullptr, Mcol));
            TRY_K(launch_mlp_infer_xg(st, Mcol, ctx->feat.p, ctx->mlpP.p, ctx->mlpNorm.p, ctx->xg.p));
            warm = true;
        }
        if (Me > 0) {
            TRY_K(launch_set_x(st, ctx->xg.p + Mcol, ctx->extX0.p, Me));
            warm = true;
        }
        alens_status s = ctx->cfg.precision == 1
                             ? run_solver_mixed(ctx, n, M, warm, &rep, &ms_solve)
                             : run_solver(ctx, n, M, ctx->mob.p, warm, &rep, &ms_solve, ctx->rsys.p, ctx->smono.p,
                                          &ctx->sysrep);
        if (s != ALENS_OK) return s;
    } else {
        rep.converged = 1;
        for (auto &r : ctx->sysrep) r = alens_solve_report{}, r.converged = 1;
    }
    TRY_K(launch_writeback(st, n, ctx->rsys.p, ctx->sysd.p, ctx->adjStart.p, ctx->sellStart.p, ctx->sellK.p,
                           ctx->sellW.p, ctx->epad, ctx->xg.p, ctx->mob.p, ctx->unc.p, ctx->F.p, ctx->U.p,
                           ctx->gamma.p, M, nullptr, ctx->spos_cur));
    if (ctx->cfg.warm_start >= 1) {
        TRY_HIP(ctx->pkey.ensure(Mc));
        TRY_HIP(ctx->pgam.ensure(Mc));
        TRY_HIP(ctx->dword.ensure(4));
        TRY_K(launch_make_keys(st, Mcol, ctx->cI.p, ctx->cJ.p, ctx->gid.p, ctx->xg.p, ctx->pkey.p, ctx->pgam.p, nullptr,
                               ctx->dword.p + 2));
        ctx->mprev = Mcol;
        ctx->have_prev = true;
    }
    TRY_HIP(hipEventRecord(ctx->ev[4], st));
    TRY_HIP(hipStreamSynchronize(st));
    float ms_collect = 0.f, ms_step = 0.f;
    TRY_HIP(hipEventElapsedTime(&ms_collect, ctx->ev[0], ctx->ev[1]));
    TRY_HIP(hipEventElapsedTime(&ms_step, ctx->ev[0], ctx->ev[4]));
    rep.n_con = M;
    rep.n_entries = E;
    rep.t_collect_ms = ms_collect;
    rep.t_solve_ms = ms_solve;
    rep.t_step_ms = ms_step;
    ctx->have_step = true;
    ctx->have_counts = true;
    ctx->ext_step.swap(ctx->ext_pending);
    ctx->ext_pending.clear();
    if (rep_out) *rep_out = rep;
    if (rep.status == ALENS_E_NONFINITE) return fail(ctx, ALENS_E_NONFINITE, "non-finite BBPGD residual");
    if (rep.status == ALENS_E_NOCONV) return fail(ctx, ALENS_E_NOCONV, "BBPGD did not reach tol within max_ite");
    return ALENS_OK;
}

// ---- device-resident step (graph mode) --------------------------------------------------------
// Capacities from the counts of the last step that sized its buffers exactly, with head room; every
// buffer the step touches is allocated here, so that enqueuing the step allocates nothing.
static alens_status plan_resident(alens_ctx *ctx) {
    const int32_t n = ctx->n, ns = ctx->nsys;
    const size_t N = (size_t)std::max(n, 1);
    alens_ctx::Caps c;
    // head room over the last counts (ALENS_GRAPH_HEADROOM, default 1.5; tests lower it to exercise the
    // fallback of a replay whose counts outgrow the capacities)
    const char *hr = getenv("ALENS_GRAPH_HEADROOM");
    const double f = std::min(8.0, std::max(0.05, hr ? atof(hr) : 1.5));
    const double slack = hr ? 0.0 : 1.0;
    c.E = (uint32_t)std::min<double>(4294967295.0, (double)ctx->E * f + slack * 4096);
    c.M = (uint32_t)std::min<double>(2147483647.0, (double)ctx->m * f + slack * 4096);
    c.epad = (uint32_t)std::min<double>(4294967295.0, (double)ctx->epad * f + slack * 4096);
    c.E = std::max<uint32_t>(c.E, 64);
    c.M = std::max<uint32_t>(c.M, 64);
    c.epad = std::max<uint32_t>(c.epad, 64);
    c.nch = (int32_t)((c.M + kRedChunk - 1) / kRedChunk) + ns + 1;
    for (auto *b : {&ctx->keys, &ctx->vals, &ctx->keys2, &ctx->vals2}) TRY_HIP(b->ensure(c.E));
    TRY_K(ensure_radix_scratch(ctx->hist, ctx->scan, c.E));
    TRY_K(ensure_scan_scratch(ctx->scan, (int64_t)n + 1));
    const int64_t nsl = ((int64_t)n + kSlice - 1) / kSlice;
    TRY_K(ensure_scan_scratch(ctx->scan, nsl + 1));
    for (auto *b : {&ctx->pI, &ctx->pJ, &ctx->pImg, &ctx->cImg, &ctx->jlist}) TRY_HIP(b->ensure(c.M));
    ctx->pcap = c.M;
    for (auto *b : {&ctx->cI, &ctx->cJ}) TRY_HIP(b->ensure(c.M));
    for (auto *b : {&ctx->delta0, &ctx->b, &ctx->gamma}) TRY_HIP(b->ensure(c.M));
    // the previous step's impulse cache is read by the next step: grow it keeping its contents
    auto grow_keep = [&](auto &buf, size_t want) -> hipError_t {
        if (want <= buf.cap && buf.p) return hipSuccess;
        std::remove_reference_t<decltype(buf)> nb;
        hipError_t e = nb.ensure(want);
        if (e != hipSuccess) return e;
        if (buf.p) {
            e = hipMemcpyAsync(nb.p, buf.p, sizeof(*buf.p) * buf.cap, hipMemcpyDeviceToDevice, ctx->st);
            if (e == hipSuccess) e = hipStreamSynchronize(ctx->st);
            if (e != hipSuccess) {
                nb.release();
                return e;
            }
            buf.release();
        }
        buf.p = nb.p;
        buf.cap = nb.cap;
        nb.p = nullptr;
        nb.cap = 0;
        return hipSuccess;
    };
    TRY_HIP(grow_keep(ctx->pgam, c.M));
    TRY_HIP(ctx->geo.ensure(9 * (size_t)c.M));
    TRY_HIP(ctx->posI.ensure(3 * (size_t)c.M));
    TRY_HIP(ctx->posJ.ensure(3 * (size_t)c.M));
    TRY_HIP(ctx->xg.ensure(c.M));
    TRY_HIP(ctx->xprev.ensure(c.M));
    TRY_HIP(grow_keep(ctx->pkey, c.M));
    TRY_HIP(ctx->slen.ensure((size_t)nsl + 1));
    TRY_HIP(ctx->sellStart.ensure((size_t)nsl + 1));
    TRY_HIP(ctx->sord.ensure(N));
    TRY_HIP(ctx->spos.ensure(N));
    TRY_HIP(ctx->sellK.ensure(c.epad));
    TRY_HIP(ctx->sellW.ensure(6 * (size_t)c.epad));
    TRY_HIP(ctx->con_off.ensure((size_t)ns + 1));
    TRY_HIP(ctx->chunk_off.ensure((size_t)ns + 1));
    TRY_HIP(ctx->chunk_sys.ensure((size_t)c.nch));
    TRY_HIP(ctx->part.ensure((size_t)kRedQ * c.nch));
    TRY_HIP(ctx->out.ensure(8 * (size_t)ns + 8));
    TRY_HIP(ctx->out32.ensure(8 * (size_t)ns + 8));
    TRY_HIP(ctx->dword.ensure(4));
    TRY_HIP(ctx->nch_dev.ensure(1));
    TRY_HIP(ctx->step_dev.ensure(1));
    TRY_HIP(ctx->stamps.ensure(4));
    TRY_HIP(ctx->tail.ensure((size_t)kTailHead + 16 * (size_t)ns));
    if (ctx->clock_khz <= 0) {
        int khz = 0;
        TRY_HIP(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, ctx->dev));
        ctx->clock_khz = khz;
    }
    if (ctx->cfg.warm_start == 2) TRY_HIP(ctx->feat.ensure(16 * (size_t)c.M));
    if (ctx->cfg.precision == 1) {
        TRY_HIP(ctx->geo32.ensure(9 * (size_t)c.M));
        TRY_HIP(ctx->sellW32.ensure(6 * (size_t)c.epad));
        TRY_HIP(ctx->b32.ensure(c.M));
        TRY_HIP(ctx->xg32.ensure(c.M));
        TRY_HIP(ctx->mob32.ensure(6 * N));
        TRY_HIP(ctx->U32.ensure(kUStride * N));
    }
    ctx->caps = c;
    return ALENS_OK;
}

// The whole step (U_nc through the warm-start keys) on the stream, sized by the capacities, with
// every count kept on the device and no host waits; buffers come from plan_resident.  warm_prev:
// start from the previous step's impulses.
static alens_status enqueue_resident(alens_ctx *ctx, bool fx, bool warm_prev) {
    hipStream_t st = ctx->st;
    const int32_t n = ctx->n;
    const alens_ctx::Caps c = ctx->caps;
    CfgDev cd = make_cfgdev(ctx);
    cd.step_dev = ctx->step_dev.p;
    const RodsDev rd = make_rods(ctx);
    uint32_t *Ee = ctx->dword.p + 0, *Mv = ctx->dword.p + 1, *mprev = ctx->dword.p + 2;
    TRY_K(launch_stamp(st, ctx->stamps.p, 0));
    TRY_K(launch_mob_unc(st, rd, ctx->length.p, ctx->radius.p, ctx->has_immov ? ctx->immov.p : nullptr,
                         fx ? ctx->fext.p : nullptr, cd, ctx->mob.p, ctx->unc.p));
    TRY_K(launch_aabb(st, rd, cd, ctx->clo.p, ctx->chi.p, ctx->blo.p, ctx->bhi.p, ctx->ecnt.p));
    TRY_K(exclusive_scan_u32(st, ctx->ecnt.p, ctx->eoff.p, n, ctx->scan));
    TRY_K(launch_guard_entries(st, ctx->eoff.p + n, c.E, Ee));
    TRY_K(launch_fill_entries(st, rd, ctx->clo.p, ctx->chi.p, ctx->eoff.p, ctx->keys.p, ctx->vals.p, c.E));
    bool alt = false;
    TRY_K(radix_sort_pairs(st, ctx->keys.p, ctx->vals.p, ctx->keys2.p, ctx->vals2.p, c.E, ctx->keybits, ctx->hist,
                           ctx->scan, alt, Ee));
    const uint32_t *skeys = alt ? ctx->keys2.p : ctx->keys.p;
    const uint32_t *svals = alt ? ctx->vals2.p : ctx->vals.p;
    {
        // every counter and cursor of the step, and the solver's barrier words, in one launch
        ZeroJobs z{};
        uint32_t *bufs[8] = {ctx->cstart.p, ctx->cend.p, ctx->counter.p, ctx->cntI.p,
                             ctx->cntJ.p,   ctx->cur.p,  ctx->cur2.p,    ctx->bar.p};
        const int64_t lens[8] = {ctx->ncell, ctx->ncell, 1, (int64_t)n + 1, (int64_t)n + 1, (int64_t)n + 1,
                                 (int64_t)n + 1, kBarTotalWords};
        for (int q = 0; q < 8; ++q) {
            z.p[q] = bufs[q];
            z.n[q] = lens[q];
        }
        z.count = 8;
        TRY_K(launch_zero_many(st, z));
        ctx->bar_clean = true;
    }
    TRY_K(launch_cell_ranges(st, skeys, c.E, ctx->cstart.p, ctx->cend.p, Ee));
    TRY_K(launch_sweep(st, rd, cd, skeys, c.E, ctx->cstart.p, ctx->cend.p, svals, ctx->clo.p, ctx->blo.p, ctx->bhi.p,
                       ctx->counter.p, c.M, ctx->pI.p, ctx->pJ.p, ctx->pImg.p, ctx->cntI.p, ctx->cntJ.p, Ee));
    TRY_K(launch_guard_pairs(st, ctx->counter.p, c.M, Mv, ctx->cntI.p, ctx->cntJ.p, n));
    TRY_K(exclusive_scan_u32(st, ctx->cntI.p, ctx->rowStart.p, n, ctx->scan));
    TRY_K(exclusive_scan_u32(st, ctx->cntJ.p, ctx->jStart.p, n, ctx->scan));
    TRY_K(launch_order_I(st, n, c.M, ctx->pI.p, ctx->pJ.p, ctx->pImg.p, ctx->rowStart.p, ctx->cur.p, ctx->cI.p,
                         ctx->cJ.p, ctx->cImg.p, Mv));
    TRY_K(launch_order_J(st, n, c.M, ctx->cJ.p, ctx->jStart.p, ctx->cur2.p, ctx->jlist.p, Mv));
    TRY_K(launch_rows(st, rd, cd, c.M, c.M, ctx->cI.p, ctx->cJ.p, ctx->cImg.p, ctx->unc.p, ctx->delta0.p,
                      ctx->geo.p, ctx->posI.p, ctx->posJ.p, ctx->b.p, Mv));
    const int64_t nsl = ((int64_t)n + kSlice - 1) / kSlice;
    ctx->sord_cur = ctx->spos_cur = nullptr;
    if (slot_order_on()) {
        TRY_K(launch_slot_order(st, n, ctx->rowStart.p, ctx->jStart.p, nullptr, ctx->sord.p, ctx->spos.p,
                                slot_order_window(n, ctx->ncu)));
        ctx->sord_cur = ctx->sord.p;
        ctx->spos_cur = ctx->spos.p;
    }
    TRY_K(launch_slice_sizes(st, n, ctx->rowStart.p, ctx->jStart.p, nullptr, ctx->slen.p, ctx->sord_cur));
    TRY_K(exclusive_scan_u32(st, ctx->slen.p, ctx->sellStart.p, nsl, ctx->scan));
    TRY_K(launch_guard_slots(st, ctx->sellStart.p + nsl, c.epad, Mv));
    TRY_K(launch_adjacency(st, n, c.M, ctx->rowStart.p, ctx->jStart.p, ctx->jlist.p, nullptr, nullptr, 0u, ctx->geo.p,
                           ctx->sellStart.p, c.epad, ctx->adjStart.p, ctx->sellK.p, ctx->sellW.p, ctx->spos_cur));
    TRY_K(launch_chunk_table(st, ctx->nsys, n, ctx->sysd.p, ctx->rowStart.p, 0u, ctx->con_off.p, ctx->chunk_off.p,
                             ctx->nch_dev.p, ctx->chunk_sys.p, c.nch, Mv));
    // start point and solve (the solver reads the row and chunk counts from the device)
    TRY_K(launch_stamp(st, ctx->stamps.p, 1));
    ctx->gstride = c.M;
    ctx->epad = c.epad;
    ctx->nchunks_cur = c.nch;
    ctx->nsys_cur = ctx->nsys;
    ctx->cur_mdev = Mv;
    ctx->cur_nchdev = ctx->nch_dev.p;
    const int ws = ctx->cfg.warm_start;
    bool warm = false;
    if (ws >= 1 && warm_prev) {
        TRY_K(launch_warm_match(st, c.M, ctx->cI.p, ctx->cJ.p, ctx->gid.p, 0u, ctx->pkey.p, ctx->pgam.p, ctx->xg.p, Mv,
                                mprev));
        TRY_HIP(hipMemcpyAsync(ctx->xprev.p, ctx->xg.p, sizeof(double2) * c.M, hipMemcpyDeviceToDevice, st));
        warm = true;
    } else {
        TRY_K(launch_init_xg(st, ctx->xg.p, nullptr, c.M));
    }
    if (ws == 2) {
        TRY_K(launch_features(st, c.M, c.M, ctx->cI.p, ctx->cJ.p, ctx->delta0.p, ctx->geo.p, ctx->posI.p, ctx->posJ.p,
                              ctx->dir.p, ctx->length.p, ctx->unc.p, warm ? ctx->xprev.p : nullptr, ctx->cfg.dt,
                              ctx->feat.p, Mv));
        TRY_K(launch_init_xg(st, ctx->xg.p, nullptr, c.M));
        TRY_K(launch_mlp_infer_xg(st, c.M, ctx->feat.p, ctx->mlpP.p, ctx->mlpNorm.p, ctx->xg.p, Mv));
        warm = true;
    }
    alens_solve_report rep{};
    float ms = 0.f;
    alens_status s = ctx->cfg.precision == 1
                         ? run_solver_mixed(ctx, n, c.M, warm, &rep, &ms)
                         : run_solver(ctx, n, c.M, ctx->mob.p, warm, &rep, &ms, ctx->rsys.p, ctx->smono.p, nullptr);
    if (s != ALENS_OK) return s;
    TRY_K(launch_writeback(st, n, ctx->rsys.p, ctx->sysd.p, ctx->adjStart.p, ctx->sellStart.p, ctx->sellK.p,
                           ctx->sellW.p, c.epad, ctx->xg.p, ctx->mob.p, ctx->unc.p, ctx->F.p, ctx->U.p, ctx->gamma.p,
                           c.M, Mv, ctx->spos_cur));
    if (ws >= 1)
        TRY_K(launch_make_keys(st, c.M, ctx->cI.p, ctx->cJ.p, ctx->gid.p, ctx->xg.p, ctx->pkey.p, ctx->pgam.p, Mv,
                               mprev));
    // everything the host reads after a replay, in one buffer
    TRY_K(launch_step_tail(st, ctx->eoff.p + n, ctx->counter.p, ctx->sellStart.p + nsl, Mv, ctx->nch_dev.p,
                           ctx->stamps.p, ctx->bar.p + kBarWords, ctx->out.p,
                           ctx->cfg.precision == 1 ? ctx->out32.p : nullptr, 8 * ctx->nsys, ctx->tail.p));
    return ALENS_OK;
}

// Record the resident step into a graph (not run).  A failure leaves graph mode on plain steps.
static alens_status capture_resident(alens_ctx *ctx, bool fx, bool warm_prev) {
    drop_graph(ctx);
    TRY_HIP(hipStreamSynchronize(ctx->st));
    TRY_HIP(hipStreamBeginCapture(ctx->st, hipStreamCaptureModeThreadLocal));
    ctx->capture = true;
    ctx->cap_solve_ev = false;
    alens_status s = enqueue_resident(ctx, fx, warm_prev);
    ctx->capture = false;
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(ctx->st, &g);
    if (s != ALENS_OK || e != hipSuccess || !g) {
        if (g) (void)hipGraphDestroy(g);
        (void)hipGetLastError();
        return s != ALENS_OK ? s : fail(ctx, ALENS_E_HIP, "step capture failed");
    }
    ctx->graph = g;
    TRY_HIP(hipGraphInstantiate(&ctx->gexec, ctx->graph, nullptr, nullptr, 0));
    ctx->graph_key = (fx ? 1 : 0) | (warm_prev ? 2 : 0);
    ++ctx->graph_captures;
    return ALENS_OK;
}

// Replay the captured step; if a count outgrew its capacity the step is redone with exact sizes
// (and the next step captures again).
static alens_status step_resident(alens_ctx *ctx, const double *f_ext, alens_solve_report *rep_out) {
    hipStream_t st = ctx->st;
    const int32_t n = ctx->n;
    const alens_ctx::Caps c = ctx->caps;
    TRY_HIP(hipMemcpyAsync(ctx->step_dev.p, &ctx->step, sizeof(int64_t), hipMemcpyHostToDevice, st));
    if (f_ext && n > 0)
        TRY_HIP(hipMemcpyAsync(ctx->fext.p, f_ext, sizeof(double) * 6 * n, hipMemcpyHostToDevice, st));
    TRY_HIP(hipEventRecord(ctx->ev[0], st));
    TRY_HIP(hipGraphLaunch(ctx->gexec, st));
    TRY_HIP(hipEventRecord(ctx->ev[4], st));
    const int nrep = 8 * ctx->nsys;
    std::vector<double> tl((size_t)kTailHead + 2 * (size_t)nrep);
    TRY_HIP(hipMemcpyAsync(tl.data(), ctx->tail.p, sizeof(double) * tl.size(), hipMemcpyDeviceToHost, st));
    TRY_HIP(hipStreamSynchronize(st));
    uint32_t cnt[5];  // entries, pairs, slots, valid rows, chunks
    for (int q = 0; q < 5; ++q) cnt[q] = (uint32_t)tl[q];
    if (cnt[0] > c.E || cnt[1] > c.M || cnt[2] > c.epad) {
        // every overflow left zero valid rows downstream, so the replay wrote no impulse cache and
        // the plain redo starts from the same previous-step impulses
        drop_graph(ctx);
        ++ctx->graph_overflows;
        return step_plain(ctx, f_ext, rep_out);
    }
    alens_solve_report rep{};
    const uint32_t abort_word = (uint32_t)tl[9];
    alens_status s = parse_reports(ctx, tl.data() + kTailHead, abort_word, &rep, &ctx->sysrep);
    if (s != ALENS_OK) return s;
    if (ctx->cfg.precision == 1) {
        alens_solve_report r32{};
        std::vector<alens_solve_report> s32;
        s = parse_reports(ctx, tl.data() + kTailHead + nrep, abort_word, &r32, &s32);
        if (s != ALENS_OK) return s;
        rep.iters_fp32 = r32.iters;
        for (size_t q = 0; q < s32.size() && q < ctx->sysrep.size(); ++q) ctx->sysrep[q].iters_fp32 = s32[q].iters;
    }
    const uint64_t stamp[4] = {(uint64_t)tl[5], (uint64_t)tl[6], (uint64_t)tl[7], (uint64_t)tl[8]};
    float ms_step = 0.f;
    TRY_HIP(hipEventElapsedTime(&ms_step, ctx->ev[0], ctx->ev[4]));
    const double tick_ms = ctx->clock_khz > 0 ? 1.0 / (double)ctx->clock_khz : 0.0;
    const double ms_collect = (double)(int64_t)(stamp[1] - stamp[0]) * tick_ms;
    const double ms_solve = (double)(int64_t)(stamp[3] - stamp[2]) * tick_ms;
    ctx->E = cnt[0];
    ctx->m = cnt[3];
    ctx->mcol = cnt[3];
    ctx->mext = 0;
    ctx->have_xprev = ctx->cfg.warm_start >= 1 && (ctx->graph_key & 2);
    if (ctx->cfg.warm_start >= 1) {
        ctx->mprev = cnt[3];
        ctx->have_prev = true;
    }
    ctx->ext_step.clear();
    ctx->have_step = true;
    ++ctx->graph_steps;
    rep.n_con = cnt[3];
    rep.n_entries = cnt[0];
    rep.t_step_ms = ms_step;
    rep.t_collect_ms = ms_collect;  // U_nc, cell list, sweep, rows, adjacency, chunk table
    rep.t_solve_ms = ms_solve;      // the solver launch(es), from clock readings in the graph
    if (rep_out) *rep_out = rep;
    if (rep.status == ALENS_E_NONFINITE) return fail(ctx, ALENS_E_NONFINITE, "non-finite BBPGD residual");
    if (rep.status == ALENS_E_NOCONV) return fail(ctx, ALENS_E_NOCONV, "BBPGD did not reach tol within max_ite");
    return ALENS_OK;
}

alens_status alens_step_constraints(alens_ctx *ctx, const double *f_ext, alens_solve_report *rep_out) {
    if (!ctx) return ALENS_E_ARG;
    if (!ctx->have_rods) return fail(ctx, ALENS_E_STATE, "step_constraints before set_rods");
    TRY_HIP(hipSetDevice(ctx->dev));
    if (!ctx->graph_mode || !ctx->ext_pending.empty()) return step_plain(ctx, f_ext, rep_out);
    if (f_ext && !finite_all(f_ext, 6 * (size_t)ctx->n)) return fail(ctx, ALENS_E_NONFINITE, "non-finite f_ext");
    const bool fx = f_ext != nullptr;
    const bool warm_prev = ctx->cfg.warm_start >= 1 && ctx->have_prev;
    if (!ctx->gexec || ctx->graph_key != ((fx ? 1 : 0) | (warm_prev ? 2 : 0))) {
        // the first step sizes the buffers exactly; later ones plan capacities from the last counts
        // (the previous step's outputs are about to be replaced) and capture the resident step once
        if (!ctx->have_counts) return step_plain(ctx, f_ext, rep_out);
        alens_status p = plan_resident(ctx);
        if (p == ALENS_OK) p = capture_resident(ctx, fx, warm_prev);
        if (p != ALENS_OK) {
            ++ctx->graph_capfail;
            ctx->graph_err = ctx->err;
            drop_graph(ctx);
            return step_plain(ctx, f_ext, rep_out);
        }
    }
    return step_resident(ctx, f_ext, rep_out);
}

alens_status alens_set_graph(alens_ctx *ctx, int32_t enable) {
    if (!ctx) return ALENS_E_ARG;
    ctx->graph_mode = enable ? 1 : 0;
    if (!enable) drop_graph(ctx);
    return ALENS_OK;
}

alens_status alens_get_graph_info(const alens_ctx *ctx, alens_graph_info *info) {
    if (!ctx) return ALENS_E_ARG;
    if (!info) return ALENS_E_ARG;
    *info = alens_graph_info{};
    info->captured = ctx->gexec ? 1 : 0;
    info->replays = ctx->graph_steps;
    info->captures = ctx->graph_captures;
    info->capture_failures = ctx->graph_capfail;
    info->overflows = ctx->graph_overflows;
    snprintf(info->last_failure, sizeof info->last_failure, "%s", ctx->graph_err.c_str());
    return ALENS_OK;
}

alens_status alens_step_euler(alens_ctx *ctx) {
    if (!ctx) return ALENS_E_ARG;
    if (!ctx->have_step) return fail(ctx, ALENS_E_STATE, "step_euler before step_constraints");
    TRY_HIP(hipSetDevice(ctx->dev));
    const CfgDev cd = make_cfgdev(ctx);
    // enqueued only: every later call that reads rods or steps runs on the same stream, and the
    // calls that copy to the host wait for it
    TRY_K(launch_euler(ctx->st, ctx->n, ctx->rsys.p, ctx->sysd.p, ctx->center.p, ctx->dir.p, ctx->U.p, cd));
    ctx->step += 1;
    ctx->have_step = false;
    return ALENS_OK;
}

alens_status alens_get_num_constraints(alens_ctx *ctx, int64_t *m) {
    if (!ctx || !m) return ALENS_E_ARG;
    if (!ctx->have_step) return fail(ctx, ALENS_E_STATE, "no step has been run");
    *m = ctx->m;
    return ALENS_OK;
}

alens_status alens_get_constraint_arrays(alens_ctx *ctx, int32_t *I, int32_t *J, int32_t *img,
                                         double *delta0, double *nrm, double *pos_i, double *pos_j,
                                         double *r_i, double *r_j, double *b) {
    if (!ctx) return ALENS_E_ARG;
    if (!ctx->have_step) return fail(ctx, ALENS_E_STATE, "no step has been run");
    TRY_HIP(hipSetDevice(ctx->dev));
    const size_t m = (size_t)ctx->m;
    TRY_HIP(d2h(ctx, I, ctx->cI.p, m));
    TRY_HIP(d2h(ctx, J, ctx->cJ.p, m));
    TRY_HIP(d2h(ctx, delta0, ctx->delta0.p, m));
    TRY_HIP(d2h(ctx, pos_i, ctx->posI.p, 3 * m));
    TRY_HIP(d2h(ctx, pos_j, ctx->posJ.p, 3 * m));
    TRY_HIP(d2h(ctx, b, ctx->b.p, m));
    std::vector<uint32_t> im;
    std::vector<double> geo;
    if (img && m) {
        im.resize(m);
        TRY_HIP(d2h(ctx, im.data(), ctx->cImg.p, m));
    }
    const size_t gs = ctx->gstride;
    if ((nrm || r_i || r_j) && m) {
        geo.resize(9 * gs);
        TRY_HIP(d2h(ctx, geo.data(), ctx->geo.p, 9 * gs));
    }
    TRY_HIP(hipStreamSynchronize(ctx->st));
    for (size_t k = 0; k < m; ++k) {
        if (img)
            for (int a = 0; a < 3; ++a) img[3 * k + a] = img_get(im[k], a);
        for (int a = 0; a < 3; ++a) {
            if (nrm) nrm[3 * k + a] = geo[a * gs + k];
            if (r_i) r_i[3 * k + a] = geo[(3 + a) * gs + k];
            if (r_j) r_j[3 * k + a] = geo[(6 + a) * gs + k];
        }
    }
    return ALENS_OK;
}

alens_status alens_get_gamma(alens_ctx *ctx, double *gamma) {
    if (!ctx || (!gamma && ctx->m)) return ALENS_E_ARG;
    if (!ctx->have_step) return fail(ctx, ALENS_E_STATE, "no step has been run");
    TRY_HIP(hipSetDevice(ctx->dev));
    TRY_HIP(d2h(ctx, gamma, ctx->gamma.p, (size_t)ctx->m));
    TRY_HIP(hipStreamSynchronize(ctx->st));
    return ALENS_OK;
}

alens_status alens_get_velocity(alens_ctx *ctx, double *u, double *u_nc, double *f_con, double *mob) {
    if (!ctx) return ALENS_E_ARG;
    if (!ctx->have_step) return fail(ctx, ALENS_E_STATE, "no step has been run");
    TRY_HIP(hipSetDevice(ctx->dev));
    const size_t n6 = 6 * (size_t)ctx->n;
    TRY_HIP(d2h(ctx, u, ctx->U.p, n6));
    TRY_HIP(d2h(ctx, u_nc, ctx->unc.p, n6));
    TRY_HIP(d2h(ctx, f_con, ctx->F.p, n6));
    TRY_HIP(d2h(ctx, mob, ctx->mob.p, n6));
    TRY_HIP(hipStreamSynchronize(ctx->st));
    return ALENS_OK;
}

alens_status alens_get_constraints(alens_ctx *ctx, int64_t cap, alens_con_block *out, int64_t *m_out) {
    if (!ctx || !m_out) return ALENS_E_ARG;
    if (!ctx->have_step) return fail(ctx, ALENS_E_STATE, "no step has been run");
    const size_t m = (size_t)ctx->m;
    *m_out = (int64_t)m;
    if (!out || cap < (int64_t)m) return m == 0 ? ALENS_OK : fail(ctx, ALENS_E_ARG, "get_constraints: capacity");
    std::vector<int32_t> I(m), J(m), img(3 * m);
    std::vector<double> d0(m), nrm(3 * m), pI(3 * m), pJ(3 * m), gam(m), cen(3 * (size_t)ctx->n);
    alens_status s = alens_get_constraint_arrays(ctx, I.data(), J.data(), img.data(), d0.data(), nrm.data(),
                                                 pI.data(), pJ.data(), nullptr, nullptr, nullptr);
    if (s != ALENS_OK) return s;
    s = alens_get_gamma(ctx, gam.data());
    if (s != ALENS_OK) return s;
    s = alens_get_rods(ctx, cen.data(), nullptr);
    if (s != ALENS_OK) return s;
    size_t sy = 0;  // rows are grouped by rod I, and rods by system, so the system index only grows
    const size_t mc = ctx->mcol;
    for (size_t k = 0; k < m; ++k) {
        if (k >= mc) {
            // caller row: the block as given, with this step's gamma and rod indices
            alens_con_block &cb = out[k];
            cb = ctx->ext_step[k - mc];
            cb.gamma = gam[k];
            cb.gidx_i = I[k];
            cb.gidx_j = J[k];
            continue;
        }
        while (sy + 1 < ctx->boxes.size() && I[k] >= ctx->rod_off[sy + 1]) ++sy;
        const alens_box &bx = ctx->boxes[sy];
        alens_con_block &cb = out[k];
        memset(&cb, 0, sizeof cb);
        cb.delta0 = d0[k];
        cb.gamma = gam[k];
        cb.kappa = 0.0;
        cb.gamma_lb = 0.0;
        cb.gid_i = ctx->h_gid[I[k]];
        cb.gid_j = ctx->h_gid[J[k]];
        cb.gidx_i = I[k];
        cb.gidx_j = J[k];
        for (int a = 0; a < 3; ++a) {
            const double len = bx.box_hi[a] - bx.box_lo[a];
            const double cj = bx.pbc[a] ? cen[3 * J[k] + a] + (double)img[3 * k + a] * len : cen[3 * J[k] + a];
            cb.norm_i[a] = nrm[3 * k + a];
            cb.norm_j[a] = -nrm[3 * k + a];
            cb.pos_i[a] = pI[3 * k + a];
            cb.pos_j[a] = pJ[3 * k + a];
            cb.lab_i[a] = cen[3 * I[k] + a] + pI[3 * k + a];
            cb.lab_j[a] = cj + pJ[3 * k + a];
        }
        // per-unit-impulse stress: virial of the force pair +n at lab_i, -n at lab_j
        for (int a = 0; a < 3; ++a)
            for (int c = 0; c < 3; ++c) cb.stress[3 * a + c] = (cb.lab_i[a] - cb.lab_j[a]) * cb.norm_i[c];
        cb.one_side = 0;
        cb.bilateral = 0;
    }
    return ALENS_OK;
}

alens_status alens_get_con_stress(alens_ctx *ctx, double collision[9], double external[9]) {
    if (!ctx) return ALENS_E_ARG;
    if (!ctx->have_step) return fail(ctx, ALENS_E_STATE, "no step has been run");
    int64_t m = ctx->m;
    std::vector<alens_con_block> blk((size_t)m);
    alens_status s = alens_get_constraints(ctx, m, blk.data(), &m);
    if (s != ALENS_OK) return s;
    double col[9] = {0}, ext[9] = {0};
    for (size_t k = 0; k < (size_t)m; ++k) {
        double *t = k < ctx->mcol ? col : ext;
        for (int q = 0; q < 9; ++q) t[q] = t[q] + blk[k].gamma * blk[k].stress[q];
    }
    if (collision) memcpy(collision, col, sizeof col);
    if (external) memcpy(external, ext, sizeof ext);
    return ALENS_OK;
}

alens_status alens_add_constraints(alens_ctx *ctx, int64_t m, const alens_con_block *blocks) {
    if (!ctx || m < 0 || (m > 0 && !blocks)) return ALENS_E_ARG;
    if (ctx->ext_pending.size() + (size_t)m > 0x3fffffffu) return fail(ctx, ALENS_E_ARG, "add_constraints: too many");
    ctx->ext_pending.insert(ctx->ext_pending.end(), blocks, blocks + m);
    return ALENS_OK;
}

alens_status alens_clear_constraints(alens_ctx *ctx) {
    if (!ctx) return ALENS_E_ARG;
    ctx->ext_pending.clear();
    return ALENS_OK;
}

alens_status alens_autotune(alens_ctx *ctx, alens_tune_report *rep_out) {
    if (ctx) drop_graph(ctx);  // the schedule is part of a captured step
    if (!ctx) return ALENS_E_ARG;
    if (!ctx->have_step) return fail(ctx, ALENS_E_STATE, "autotune needs a step's constraints (run a step first)");
    if (ctx->m == 0) return fail(ctx, ALENS_E_STATE, "autotune: the last step had no constraints");
    TRY_HIP(hipSetDevice(ctx->dev));
    const int32_t n = ctx->n;
    const uint32_t M = (uint32_t)ctx->m;
    const auto saved_rep = ctx->sysrep;
    // one candidate = one cold solve of the last step's rows with the precision form in use; the
    // fastest of two runs counts (results are identical under every schedule)
    auto run = [&](float *best) -> alens_status {
        *best = INFINITY;
        for (int r = 0; r < 2; ++r) {
            TRY_K(launch_init_xg(ctx->st, ctx->xg.p, nullptr, M));
            alens_solve_report rep{};
            float ms = 0.f;
            alens_status s = ctx->cfg.precision == 1
                                 ? run_solver_mixed(ctx, n, M, false, &rep, &ms)
                                 : run_solver(ctx, n, M, ctx->mob.p, false, &rep, &ms, ctx->rsys.p, ctx->smono.p,
                                              nullptr);
            if (s != ALENS_OK && s != ALENS_E_NOCONV) return s;
            *best = std::min(*best, ms);
        }
        return ALENS_OK;
    };
    constexpr float kTuneMargin = 0.01f;
    alens_tune_report tr{};
    ctx->tuning = true;
    // coordinate descent from the defaults over the switches, one axis at a time
    ctx->bar_mode = 3;
    ctx->xcd_map = 0;
    ctx->max_blocks = 0;
    ctx->entries = 0;
    ctx->group_width = 0;
    ctx->gw_rows = 0;
    ctx->stream_hint = 0;
    float best = 0.f;
    alens_status s = run(&best);
    tr.default_ms = best;
    tr.candidates = 1;
    const int64_t grid = std::max<int64_t>(ctx->grid_used, 8);  // the default schedule's grid
    struct Axis {
        int32_t *field;
        std::vector<int32_t> values;
    };
    std::vector<Axis> axes = {
        {&ctx->entries, {0, 2}},
        {&ctx->bar_mode, {3, 1, 0, 2}},
        {&ctx->xcd_map, {0, 1}},
        {&ctx->max_blocks, {0, (int32_t)std::max<int64_t>(8, grid * 3 / 4), (int32_t)std::max<int64_t>(8, grid / 2)}},
        {&ctx->group_width, {0, 128}},
        {&ctx->stream_hint, {0, 1, 2}},
    };
    for (auto &ax : axes) {
        if (s != ALENS_OK) break;
        const int32_t keep = *ax.field;
        int32_t pick = keep;
        for (int32_t v : ax.values) {
            if (v == keep) continue;
            *ax.field = v;
            float t = 0.f;
            s = run(&t);
            if (s != ALENS_OK) break;
            ++tr.candidates;
            // a value replaces the current one only when it is clearly faster: the automatic
            // values (group width, stream hint) adapt to later steps' row counts, and a fixed value
            // that ties with them on the tuned step can lose on later ones
            if (t < best * (1.0f - kTuneMargin)) {
                best = t;
                pick = v;
            }
        }
        *ax.field = pick;
    }
    ctx->tuning = false;
    ctx->sysrep = saved_rep;
    // 128-lane groups won on this step's rows; later steps with other row counts may not gain (on the
    // 200k slab they won at about 1,070 chunks and lost 12% at 645), so the choice holds near M only
    ctx->gw_rows = ctx->group_width == 128 ? M : 0u;
    if (s != ALENS_OK) return s;
    tr.bar_mode = ctx->bar_mode;
    tr.xcd_map = ctx->xcd_map;
    tr.max_blocks = ctx->max_blocks;
    tr.entries = ctx->entries;
    tr.group_width = ctx->group_width;
    tr.stream_hint = ctx->stream_hint;
    tr.best_ms = best;
    if (rep_out) *rep_out = tr;
    return ALENS_OK;
}

alens_status alens_bcqp_solve(alens_ctx *ctx, int32_t n, int64_t m, const int32_t *I, const int32_t *J,
                              const double *nrm, const double *r_i, const double *r_j, const double *b,
                              const double *mob, const double *x0, double *gamma_out,
                              alens_solve_report *rep_out) {
    if (!ctx) return ALENS_E_ARG;
    if (n < 0 || m < 0 || m > 0x7fffffff || (m > 0 && (!I || !J || !nrm || !r_i || !r_j || !b || !mob || !gamma_out)))
        return fail(ctx, ALENS_E_ARG, "bcqp_solve: bad arguments");
    alens_solve_report rep{};
    if (m == 0) {
        rep.converged = 1;
        if (rep_out) *rep_out = rep;
        return ALENS_OK;
    }
    for (int64_t k = 0; k < m; ++k)
        if (I[k] < 0 || I[k] >= n || J[k] < 0 || J[k] >= n || I[k] == J[k])
            return fail(ctx, ALENS_E_ARG, "bcqp_solve: rod index out of range");
    TRY_HIP(hipSetDevice(ctx->dev));
    ctx->have_step = false;  // the step buffers are about to hold this solve's rows
    drop_graph(ctx);
    ctx->mext = 0;
    ctx->mcol = 0;
    ctx->gstride = (uint32_t)m;
    ctx->cur_mdev = nullptr;
    ctx->cur_nchdev = nullptr;
    // rod-major adjacency in ascending constraint index, built on the host for arbitrary rows, in the
    // sliced layout the solver reads (see launch_adjacency); geometry component-major
    const size_t M = (size_t)m, N = (size_t)std::max(n, 1);
    HostAdjacency h;
    h.build(n, m, I, J, nrm, r_i, r_j);
    const size_t nsl = h.sstart.size() - 1;
    const uint64_t ep = h.ep;
    std::vector<uint32_t> &start = h.start, &sstart = h.sstart, &sk = h.sk;
    std::vector<double> &sw = h.sw, &geo = h.geo;
    ctx->epad = ep;
    ctx->sord_cur = ctx->spos_cur = nullptr;  // rods in index order in this layout
    TRY_HIP(ctx->adjStart.ensure(N + 1));
    TRY_HIP(ctx->sellStart.ensure(nsl + 1));
    TRY_HIP(ctx->sellK.ensure(sk.size()));
    TRY_HIP(ctx->sellW.ensure(sw.size()));
    TRY_HIP(ctx->geo.ensure(9 * M));
    TRY_HIP(ctx->cI.ensure(M));
    TRY_HIP(ctx->cJ.ensure(M));
    TRY_HIP(ctx->b.ensure(M));
    TRY_HIP(ctx->xg.ensure(M));
    TRY_HIP(ctx->gamma.ensure(M));
    TRY_HIP(ctx->Us.ensure(kUStride * N));
    TRY_HIP(ctx->fext.ensure(6 * N));
    TRY_HIP(ctx->bar.ensure(kBarTotalWords));
    TRY_HIP(ctx->out.ensure(16));
    TRY_HIP(ctx->delta0.ensure(M));
    hipStream_t st = ctx->st;
    TRY_HIP(hipMemcpyAsync(ctx->adjStart.p, start.data(), sizeof(uint32_t) * (n + 1), hipMemcpyHostToDevice, st));
    TRY_HIP(hipMemcpyAsync(ctx->sellStart.p, sstart.data(), sizeof(uint32_t) * (nsl + 1), hipMemcpyHostToDevice, st));
    TRY_HIP(hipMemcpyAsync(ctx->sellK.p, sk.data(), sizeof(uint32_t) * sk.size(), hipMemcpyHostToDevice, st));
    TRY_HIP(hipMemcpyAsync(ctx->sellW.p, sw.data(), sizeof(double) * sw.size(), hipMemcpyHostToDevice, st));
    TRY_HIP(hipMemcpyAsync(ctx->geo.p, geo.data(), sizeof(double) * 9 * M, hipMemcpyHostToDevice, st));
    TRY_HIP(hipMemcpyAsync(ctx->cI.p, I, sizeof(int32_t) * M, hipMemcpyHostToDevice, st));
    TRY_HIP(hipMemcpyAsync(ctx->cJ.p, J, sizeof(int32_t) * M, hipMemcpyHostToDevice, st));
    TRY_HIP(hipMemcpyAsync(ctx->b.p, b, sizeof(double) * M, hipMemcpyHostToDevice, st));
    TRY_HIP(hipMemcpyAsync(ctx->fext.p, mob, sizeof(double) * 6 * n, hipMemcpyHostToDevice, st));
    if (x0) TRY_HIP(hipMemcpyAsync(ctx->delta0.p, x0, sizeof(double) * M, hipMemcpyHostToDevice, st));
    TRY_K(launch_init_xg(st, ctx->xg.p, x0 ? ctx->delta0.p : nullptr, (uint32_t)M));
    // one system: every rod in system 0, no monolayer projection
    DBuf<int32_t> rs0, mono0;
    TRY_HIP(rs0.ensure(N));
    TRY_HIP(mono0.ensure(1));
    TRY_HIP(hipMemsetAsync(rs0.p, 0, sizeof(int32_t) * N, st));
    TRY_HIP(hipMemsetAsync(mono0.p, 0, sizeof(int32_t), st));
    alens_status cs = set_chunk_table(ctx, std::vector<uint32_t>{0u, (uint32_t)M});
    if (cs != ALENS_OK) return cs;
    float ms = 0.f;
    alens_status s = run_solver(ctx, n, (uint32_t)M, ctx->fext.p, x0 != nullptr, &rep, &ms, rs0.p, mono0.p, nullptr);
    rs0.release();
    mono0.release();
    if (s != ALENS_OK) return s;
    TRY_K(launch_writeback(st, 0, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, 0, ctx->xg.p, nullptr, nullptr,
                           nullptr, nullptr, ctx->gamma.p, (uint32_t)M));
    TRY_HIP(hipMemcpyAsync(gamma_out, ctx->gamma.p, sizeof(double) * M, hipMemcpyDeviceToHost, st));
    TRY_HIP(hipStreamSynchronize(st));
    rep.n_con = m;
    rep.t_solve_ms = ms;
    rep.t_step_ms = ms;
    ctx->have_step = false;  // the step buffers now hold this solve's rows
    if (rep_out) *rep_out = rep;
    if (rep.status == ALENS_E_NONFINITE) return fail(ctx, ALENS_E_NONFINITE, "non-finite BBPGD residual");
    if (rep.status == ALENS_E_NOCONV) return fail(ctx, ALENS_E_NOCONV, "BBPGD did not reach tol within max_ite");
    return ALENS_OK;
}

}  // extern "C"

// ---- allocation-free BCQP solve: every device buffer lives in one static device array ----
// Sizes: rows, rods and padded adjacency slots of one call.  scratch holds, each part 256-byte
// aligned: adjStart (n + 1 u32), sellStart (n / 64 + 2 u32), sellK (S u32), sellW (6 S f64), geo (9 M f64),
// cI, cJ (M i32), b, x0, gamma (M f64), xg (M x 2 f64), Us and mob (6 N f64), rsys (N i32), smono,
// con_off, chunk_off (2 u32 each), chunk_sys (chunks i32), part (8 x chunks f64), bar (barrier words),
// out (16 f64).
constexpr size_t kStaticMaxRows = size_t(1) << 17;
constexpr size_t kStaticMaxRods = size_t(1) << 16;
constexpr size_t kStaticMaxSlots = 4 * kStaticMaxRows;
constexpr size_t kStaticMaxChunks = (kStaticMaxRows + kRedChunk - 1) / kRedChunk;
static constexpr size_t al256(size_t b) { return (b + 255) & ~size_t(255); }
constexpr size_t kScratchBytes =
    al256(4 * (kStaticMaxRods + 1)) + al256(4 * (kStaticMaxRods / kSlice + 2)) + al256(4 * kStaticMaxSlots) +
    al256(48 * kStaticMaxSlots) + al256(72 * kStaticMaxRows) + 2 * al256(4 * kStaticMaxRows) +
    3 * al256(8 * kStaticMaxRows) + al256(16 * kStaticMaxRows) + 2 * al256(48 * kStaticMaxRods) +
    al256(4 * kStaticMaxRods) + 4 * al256(8) + al256(4 * kStaticMaxChunks) + al256(8 * kRedQ * kStaticMaxChunks) +
    al256(4 * kBarTotalWords) + al256(8 * 16);

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
// The device address of `scratch` on each device, fetched once with that device current (the array's
// name is only a host-side stand-in; kernels and copies get this pointer).
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char *g_scratch_ptr[kMaxDevices] = {};

hipError_t scratch_on(int dev, unsigned char **out) {
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        if (!p) return hipErrorInvalidSymbol;
        g_scratch_ptr[dev] = (unsigned char *)p;
    }
    *out = g_scratch_ptr[dev];
    return hipSuccess;
}

// carves consecutive 256-byte aligned parts out of the scratch array
struct Carve {
    unsigned char *base;
    size_t off = 0;
    template <typename T>
    T *take(size_t count) {
        T *p = (T *)(base + off);
        off += al256(sizeof(T) * count);
        return p;
    }
};
}  // namespace

extern "C" {

alens_status alens_bcqp_solve_static(alens_ctx *ctx, int32_t n, int64_t m, const int32_t *I, const int32_t *J,
                                     const double *nrm, const double *r_i, const double *r_j, const double *b,
                                     const double *mob, const double *x0, double *gamma_out,
                                     alens_solve_report *rep_out) {
    if (!ctx) return ALENS_E_ARG;
    if (n < 0 || m < 0 || (m > 0 && (!I || !J || !nrm || !r_i || !r_j || !b || !mob || !gamma_out)))
        return fail(ctx, ALENS_E_ARG, "bcqp_solve_static: bad arguments");
    if ((size_t)m > kStaticMaxRows || (size_t)n > kStaticMaxRods)
        return fail(ctx, ALENS_E_ARG,
                    "bcqp_solve_static: at most " + std::to_string(kStaticMaxRows) + " rows and " +
                        std::to_string(kStaticMaxRods) + " rods (use alens_bcqp_solve for larger problems)");
    alens_solve_report rep{};
    if (m == 0) {
        rep.converged = 1;
        if (rep_out) *rep_out = rep;
        return ALENS_OK;
    }
    for (int64_t k = 0; k < m; ++k)
        if (I[k] < 0 || I[k] >= n || J[k] < 0 || J[k] >= n || I[k] == J[k])
            return fail(ctx, ALENS_E_ARG, "bcqp_solve_static: rod index out of range");
    std::vector<uint32_t> deg;
    if (HostAdjacency::slots(n, m, I, J, deg) > kStaticMaxSlots)
        return fail(ctx, ALENS_E_ARG,
                    "bcqp_solve_static: the rows' padded adjacency needs more than " + std::to_string(kStaticMaxSlots) +
                        " slots (a few rods with very many rows); use alens_bcqp_solve");
    TRY_HIP(hipSetDevice(ctx->dev));
    unsigned char *base = nullptr;
    TRY_HIP(scratch_on(ctx->dev, &base));
    HostAdjacency h;
    h.build(n, m, I, J, nrm, r_i, r_j);
    const size_t M = (size_t)m, N = (size_t)n, ep = h.ep, nsl = h.sstart.size() - 1;
    const uint32_t nch = (uint32_t)((M + kRedChunk - 1) / kRedChunk);
    Carve cv{base};
    uint32_t *adjStart = cv.take<uint32_t>(kStaticMaxRods + 1);
    uint32_t *sellStart = cv.take<uint32_t>(kStaticMaxRods / kSlice + 2);
    uint32_t *sellK = cv.take<uint32_t>(kStaticMaxSlots);
    double *sellW = cv.take<double>(6 * kStaticMaxSlots);
    double *geo = cv.take<double>(9 * kStaticMaxRows);
    int32_t *cI = cv.take<int32_t>(kStaticMaxRows);
    int32_t *cJ = cv.take<int32_t>(kStaticMaxRows);
    double *bd = cv.take<double>(kStaticMaxRows);
    double *x0d = cv.take<double>(kStaticMaxRows);
    double *gam = cv.take<double>(kStaticMaxRows);
    double2 *xg = cv.take<double2>(kStaticMaxRows);
    double *Us = cv.take<double>(6 * kStaticMaxRods);
    double *mobd = cv.take<double>(6 * kStaticMaxRods);
    int32_t *rsys = cv.take<int32_t>(kStaticMaxRods);
    int32_t *smono = cv.take<int32_t>(2);
    uint32_t *con_off = cv.take<uint32_t>(2);
    uint32_t *chunk_off = cv.take<uint32_t>(2);
    cv.take<uint32_t>(2);
    int32_t *chunk_sys = cv.take<int32_t>(kStaticMaxChunks);
    double *part = cv.take<double>((size_t)kRedQ * kStaticMaxChunks);
    uint32_t *bar = cv.take<uint32_t>(kBarTotalWords);
    double *out = cv.take<double>(16);
    if (cv.off > kScratchBytes) return fail(ctx, ALENS_E_ARG, "bcqp_solve_static: scratch layout overflow");
    hipStream_t st = ctx->st;
    // the previous call's contents stay in the array: every part read below is written first
    const uint32_t coff[2] = {0u, (uint32_t)M}, choff[2] = {0u, nch};
    TRY_HIP(hipMemcpyAsync(adjStart, h.start.data(), sizeof(uint32_t) * (N + 1), hipMemcpyHostToDevice, st));
    TRY_HIP(hipMemcpyAsync(sellStart, h.sstart.data(), sizeof(uint32_t) * (nsl + 1), hipMemcpyHostToDevice, st));
    TRY_HIP(hipMemcpyAsync(sellK, h.sk.data(), sizeof(uint32_t) * ep, hipMemcpyHostToDevice, st));
    for (int c = 0; c < 6; ++c)  // component stride ep (the padded slot count of this call)
        TRY_HIP(hipMemcpyAsync(sellW + c * ep, h.sw.data() + c * ep, sizeof(double) * ep, hipMemcpyHostToDevice, st));
    TRY_HIP(hipMemcpyAsync(geo, h.geo.data(), sizeof(double) * 9 * M, hipMemcpyHostToDevice, st));
    TRY_HIP(hipMemcpyAsync(cI, I, sizeof(int32_t) * M, hipMemcpyHostToDevice, st));
    TRY_HIP(hipMemcpyAsync(cJ, J, sizeof(int32_t) * M, hipMemcpyHostToDevice, st));
    TRY_HIP(hipMemcpyAsync(bd, b, sizeof(double) * M, hipMemcpyHostToDevice, st));
    TRY_HIP(hipMemcpyAsync(mobd, mob, sizeof(double) * 6 * N, hipMemcpyHostToDevice, st));
    if (x0) TRY_HIP(hipMemcpyAsync(x0d, x0, sizeof(double) * M, hipMemcpyHostToDevice, st));
    TRY_HIP(hipMemcpyAsync(con_off, coff, sizeof coff, hipMemcpyHostToDevice, st));
    TRY_HIP(hipMemcpyAsync(chunk_off, choff, sizeof choff, hipMemcpyHostToDevice, st));
    ZeroJobs z{};
    z.p[0] = (uint32_t *)rsys;
    z.n[0] = (int64_t)N;
    z.p[1] = (uint32_t *)smono;
    z.n[1] = 1;
    z.p[2] = (uint32_t *)chunk_sys;
    z.n[2] = nch;
    z.count = 3;
    TRY_K(launch_zero_many(st, z));  // one system: every rod and chunk in system 0, no monolayer
    TRY_K(launch_init_xg(st, xg, x0 ? x0d : nullptr, (uint32_t)M));
    SolveArgsT<double> a{};
    a.n = n;
    a.m = (uint32_t)M;
    a.adjStart = adjStart;
    a.sellStart = sellStart;
    a.sord = nullptr;  // rods in index order
    a.sellK = sellK;
    a.sellW = sellW;
    a.epad = ep;
    a.mob = mobd;
    a.cI = cI;
    a.cJ = cJ;
    a.geo = geo;
    a.b = bd;
    a.xg = xg;
    a.U = Us;
    a.rsys = rsys;
    a.smono = smono;
    a.warm = x0 ? 1 : 0;
    // the solver reads the context's system and chunk counts; this call's are one system, nch chunks
    const int32_t ns_keep = ctx->nsys_cur, nch_keep = ctx->nchunks_cur;
    const bool cap_keep = ctx->capture;
    const uint32_t *md_keep = ctx->cur_mdev;
    const int32_t *nd_keep = ctx->cur_nchdev;
    ctx->nsys_cur = 1;
    ctx->nchunks_cur = (int32_t)nch;
    ctx->capture = false;
    ctx->cur_mdev = nullptr;
    ctx->cur_nchdev = nullptr;
    float ms = 0.f;
    const SolveMem mem{part, out, bar, con_off, chunk_off, chunk_sys};
    alens_status s = run_bbpgd_on<double>(ctx, a, mem, &rep, &ms, nullptr);
    ctx->nsys_cur = ns_keep;
    ctx->nchunks_cur = nch_keep;
    ctx->capture = cap_keep;
    ctx->cur_mdev = md_keep;
    ctx->cur_nchdev = nd_keep;
    if (s != ALENS_OK) return s;
    TRY_K(launch_writeback(st, 0, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, 0, xg, nullptr, nullptr,
                           nullptr, nullptr, gam, (uint32_t)M));
    TRY_HIP(hipMemcpyAsync(gamma_out, gam, sizeof(double) * M, hipMemcpyDeviceToHost, st));
    TRY_HIP(hipStreamSynchronize(st));
    rep.n_con = m;
    rep.t_solve_ms = ms;
    rep.t_step_ms = ms;
    if (rep_out) *rep_out = rep;
    if (rep.status == ALENS_E_NONFINITE) return fail(ctx, ALENS_E_NONFINITE, "non-finite BBPGD residual");
    if (rep.status == ALENS_E_NOCONV) return fail(ctx, ALENS_E_NOCONV, "BBPGD did not reach tol within max_ite");
    return ALENS_OK;
}

alens_status alens_mlp_set(alens_ctx *ctx, const float *params, const float *norm) {
    if (ctx) drop_graph(ctx);
    if (!ctx || !params || !norm) return ALENS_E_ARG;
    for (int i = 0; i < kMlpNorm; ++i)
        if (!std::isfinite(norm[i])) return fail(ctx, ALENS_E_NONFINITE, "mlp_set: non-finite normalisation");
    for (int i = 0; i < kMlpParams; ++i)
        if (!std::isfinite(params[i])) return fail(ctx, ALENS_E_NONFINITE, "mlp_set: non-finite parameter");
    TRY_HIP(hipSetDevice(ctx->dev));
    TRY_HIP(ctx->mlpP.ensure(kMlpParams));
    TRY_HIP(ctx->mlpNorm.ensure(kMlpNorm));
    TRY_HIP(hipMemcpyAsync(ctx->mlpP.p, params, sizeof(float) * kMlpParams, hipMemcpyHostToDevice, ctx->st));
    TRY_HIP(hipMemcpyAsync(ctx->mlpNorm.p, norm, sizeof(float) * kMlpNorm, hipMemcpyHostToDevice, ctx->st));
    TRY_HIP(hipStreamSynchronize(ctx->st));
    ctx->mlp_ready = true;
    ctx->adam_step = 0;
    return ALENS_OK;
}

alens_status alens_mlp_get(alens_ctx *ctx, float *params, float *norm) {
    if (!ctx) return ALENS_E_ARG;
    if (!ctx->mlp_ready) return fail(ctx, ALENS_E_STATE, "no network parameters");
    TRY_HIP(hipSetDevice(ctx->dev));
    if (params) TRY_HIP(hipMemcpyAsync(params, ctx->mlpP.p, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost, ctx->st));
    if (norm) TRY_HIP(hipMemcpyAsync(norm, ctx->mlpNorm.p, sizeof(float) * kMlpNorm, hipMemcpyDeviceToHost, ctx->st));
    TRY_HIP(hipStreamSynchronize(ctx->st));
    return ALENS_OK;
}

alens_status alens_get_features(alens_ctx *ctx, float *X) {
    if (!ctx || (!X && ctx->m)) return ALENS_E_ARG;
    if (!ctx->have_step) return fail(ctx, ALENS_E_STATE, "no step has been run");
    TRY_HIP(hipSetDevice(ctx->dev));
    const uint32_t M = (uint32_t)ctx->m;
    if (M == 0) return ALENS_OK;
    TRY_HIP(ctx->feat.ensure(16 * (size_t)M));
    // collision rows (with the previous gamma when the step had one), then caller rows (previous gamma 0)
    const uint32_t mc = ctx->mcol, me = M - mc;
    TRY_K(launch_features(ctx->st, mc, ctx->gstride, ctx->cI.p, ctx->cJ.p, ctx->delta0.p, ctx->geo.p, ctx->posI.p, ctx->posJ.p,
                          ctx->dir.p, ctx->length.p, ctx->unc.p, ctx->have_xprev ? ctx->xprev.p : nullptr,
                          ctx->cfg.dt, ctx->feat.p));
    TRY_K(launch_features(ctx->st, me, ctx->gstride, ctx->cI.p + mc, ctx->cJ.p + mc, ctx->delta0.p + mc, ctx->geo.p + mc,
                          ctx->posI.p + 3 * (size_t)mc, ctx->posJ.p + 3 * (size_t)mc, ctx->dir.p, ctx->length.p,
                          ctx->unc.p, nullptr, ctx->cfg.dt, ctx->feat.p + 16 * (size_t)mc));
    TRY_HIP(hipMemcpyAsync(X, ctx->feat.p, sizeof(float) * 16 * M, hipMemcpyDeviceToHost, ctx->st));
    TRY_HIP(hipStreamSynchronize(ctx->st));
    return ALENS_OK;
}

alens_status alens_mlp_predict(alens_ctx *ctx, int64_t m, const float *X, float *gamma0) {
    if (!ctx || m < 0 || m > 0x7fffffff || (m > 0 && (!X || !gamma0))) return ALENS_E_ARG;
    if (!ctx->mlp_ready) return fail(ctx, ALENS_E_STATE, "no network parameters");
    if (m == 0) return ALENS_OK;
    TRY_HIP(hipSetDevice(ctx->dev));
    DBuf<float> dX, dY;
    TRY_HIP(dX.ensure(16 * (size_t)m));
    TRY_HIP(dY.ensure((size_t)m));
    TRY_HIP(hipMemcpyAsync(dX.p, X, sizeof(float) * 16 * m, hipMemcpyHostToDevice, ctx->st));
    TRY_K(launch_mlp_infer_f32(ctx->st, (uint32_t)m, dX.p, ctx->mlpP.p, ctx->mlpNorm.p, dY.p));
    TRY_HIP(hipMemcpyAsync(gamma0, dY.p, sizeof(float) * m, hipMemcpyDeviceToHost, ctx->st));
    TRY_HIP(hipStreamSynchronize(ctx->st));
    dX.release();
    dY.release();
    return ALENS_OK;
}

namespace {
struct TrainBufs {
    DBuf<float> XT, Z1, H1, Z2, H2, Z3, H3, dO, dZ1, dZ2, dZ3, loss, lossv, kpart;
    hipError_t ensure(int B) {
        const size_t nb = (size_t)64 * B;
        hipError_t e;
        if ((e = XT.ensure(16 * (size_t)B)) != hipSuccess) return e;
        for (auto *b : {&Z1, &H1, &Z2, &H2, &Z3, &H3, &dZ1, &dZ2, &dZ3})
            if ((e = b->ensure(nb)) != hipSuccess) return e;
        if ((e = dO.ensure((size_t)B)) != hipSuccess) return e;
        if ((e = lossv.ensure((size_t)B)) != hipSuccess) return e;
        if ((e = kpart.ensure(kMlpSplitFloats)) != hipSuccess) return e;
        return loss.ensure(4);
    }
    MlpScratch view() {
        return MlpScratch{XT.p, Z1.p, H1.p, Z2.p, H2.p, Z3.p, H3.p, dO.p, dZ1.p, dZ2.p, dZ3.p, loss.p, lossv.p, kpart.p};
    }
    void release() {
        for (auto *b : {&XT, &Z1, &H1, &Z2, &H2, &Z3, &H3, &dO, &dZ1, &dZ2, &dZ3, &loss, &lossv, &kpart}) b->release();
    }
};
}  // namespace

alens_status alens_mlp_gradient(alens_ctx *ctx, int64_t m, const float *X, const float *target, float *grad,
                                float *loss) {
    if (!ctx || m <= 0 || m % 16 || m > (1 << 26) || !X || !target || !grad) return ALENS_E_ARG;
    if (!ctx->mlp_ready) return fail(ctx, ALENS_E_STATE, "no network parameters");
    TRY_HIP(hipSetDevice(ctx->dev));
    TrainBufs tb;
    DBuf<float> dX, dT, dG;
    TRY_HIP(tb.ensure((int)m));
    TRY_HIP(dX.ensure(16 * (size_t)m));
    TRY_HIP(dT.ensure((size_t)m));
    TRY_HIP(dG.ensure(kMlpParams));
    TRY_HIP(hipMemcpyAsync(dX.p, X, sizeof(float) * 16 * m, hipMemcpyHostToDevice, ctx->st));
    TRY_HIP(hipMemcpyAsync(dT.p, target, sizeof(float) * m, hipMemcpyHostToDevice, ctx->st));
    MlpScratch sv = tb.view();
    TRY_K(mlp_gradient(ctx->st, (int)m, dX.p, dT.p, ctx->mlpP.p, ctx->mlpNorm.p, sv, dG.p));
    TRY_HIP(hipMemcpyAsync(grad, dG.p, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost, ctx->st));
    float l = 0.f;
    TRY_HIP(hipMemcpyAsync(&l, tb.loss.p, sizeof(float), hipMemcpyDeviceToHost, ctx->st));
    TRY_HIP(hipStreamSynchronize(ctx->st));
    if (loss) *loss = l;
    tb.release();
    dX.release();
    dT.release();
    dG.release();
    return ALENS_OK;
}

alens_status alens_mlp_train(alens_ctx *ctx, int64_t m, const float *X, const float *gamma, int32_t epochs,
                             int32_t batch, float lr, uint64_t seed, float *loss_per_epoch) {
    if (ctx) drop_graph(ctx);
    if (!ctx || m <= 0 || m > (1 << 28) || !X || !gamma || epochs < 0 || batch < 16 || batch % 16 || !(lr > 0.f))
        return ALENS_E_ARG;
    if (batch > m) batch = (int32_t)(m / 16 * 16);
    if (batch < 16) return fail(ctx, ALENS_E_ARG, "mlp_train: fewer than 16 samples");
    TRY_HIP(hipSetDevice(ctx->dev));
    hipStream_t st = ctx->st;
    // normalisation from the data unless a network is already loaded
    std::vector<float> norm(kMlpNorm), P(kMlpParams, 0.0f);
    if (!ctx->mlp_ready) {
        for (int f = 0; f < kMlpIn; ++f) {
            double s = 0, s2 = 0;
            for (int64_t i = 0; i < m; ++i) {
                const double v = X[16 * i + f];
                s += v;
                s2 += v * v;
            }
            const double mu = s / m, var = std::max(s2 / m - mu * mu, 0.0);
            norm[f] = (float)mu;
            norm[kMlpIn + f] = (float)(1.0 / std::max(std::sqrt(var), 1e-12));
        }
        double g2 = 0;
        for (int64_t i = 0; i < m; ++i) g2 += (double)gamma[i] * gamma[i];
        norm[2 * kMlpIn] = (float)(g2 > 0 ? std::sqrt(g2 / m) : 1.0);
        // Glorot-uniform weights (xorshift64*), zero biases, output bias at softplus^{-1}(1)
        uint64_t rs = seed ? seed : 0x9E3779B97F4A7C15ull;
        auto urand = [&]() {
            rs ^= rs >> 12;
            rs ^= rs << 25;
            rs ^= rs >> 27;
            return (double)((rs * 2685821657736338717ull) >> 11) * 0x1p-53;
        };
        auto fill = [&](int off, int rows, int cols, int fan_in) {
            const double a = std::sqrt(6.0 / (fan_in + rows));
            for (int r = 0; r < rows; ++r)
                for (int c = 0; c < cols; ++c)
                    P[off + r * cols + c] = (c < fan_in) ? (float)((2.0 * urand() - 1.0) * a) : 0.0f;
        };
        fill(kPW1, 64, 16, kMlpIn);
        fill(kPW2, 64, 64, 64);
        fill(kPW3, 64, 64, 64);
        fill(kPW4, 1, 64, 64);
        P[kPB4] = 0.5413248546f;
        alens_status s = alens_mlp_set(ctx, P.data(), norm.data());
        if (s != ALENS_OK) return s;
    } else {
        alens_status s = alens_mlp_get(ctx, nullptr, norm.data());
        if (s != ALENS_OK) return s;
    }
    const float gs = norm[2 * kMlpIn];
    std::vector<float> t(m);
    for (int64_t i = 0; i < m; ++i) t[i] = gamma[i] / gs;
    TrainBufs tb;
    DBuf<float> dX, dT, dXB, dTB, dG, dM1, dM2, dL;
    DBuf<int32_t> dIdx;
    const int64_t nb = m / batch;
    TRY_HIP(tb.ensure(batch));
    TRY_HIP(dX.ensure(16 * (size_t)m));
    TRY_HIP(dT.ensure((size_t)m));
    TRY_HIP(dXB.ensure(16 * (size_t)batch));
    TRY_HIP(dTB.ensure((size_t)batch));
    TRY_HIP(dG.ensure(kMlpParams));
    TRY_HIP(dM1.ensure(kMlpParams));
    TRY_HIP(dM2.ensure(kMlpParams));
    TRY_HIP(dL.ensure((size_t)std::max<int64_t>(nb, 1)));
    TRY_HIP(dIdx.ensure((size_t)m));
    TRY_HIP(hipMemcpyAsync(dX.p, X, sizeof(float) * 16 * m, hipMemcpyHostToDevice, st));
    TRY_HIP(hipMemcpyAsync(dT.p, t.data(), sizeof(float) * m, hipMemcpyHostToDevice, st));
    TRY_HIP(hipMemsetAsync(dM1.p, 0, sizeof(float) * kMlpParams, st));
    TRY_HIP(hipMemsetAsync(dM2.p, 0, sizeof(float) * kMlpParams, st));
    std::vector<int32_t> idx(m);
    std::vector<float> lh((size_t)std::max<int64_t>(nb, 1));
    MlpScratch sv = tb.view();
    uint64_t rs = seed ^ 0xD1B54A32D192ED03ull;
    int step = 0;
    for (int ep = 0; ep < epochs; ++ep) {
        for (int64_t i = 0; i < m; ++i) idx[i] = (int32_t)i;
        for (int64_t i = m - 1; i > 0; --i) {  // Fisher-Yates with xorshift64*
            rs ^= rs >> 12;
            rs ^= rs << 25;
            rs ^= rs >> 27;
            const int64_t j = (int64_t)(((rs * 2685821657736338717ull) >> 11) % (uint64_t)(i + 1));
            std::swap(idx[i], idx[j]);
        }
        TRY_HIP(hipMemcpyAsync(dIdx.p, idx.data(), sizeof(int32_t) * m, hipMemcpyHostToDevice, st));
        for (int64_t b = 0; b < nb; ++b) {
            TRY_K(launch_gather_batch(st, batch, dIdx.p + b * batch, dX.p, dT.p, dXB.p, dTB.p));
            TRY_K(mlp_gradient(st, batch, dXB.p, dTB.p, ctx->mlpP.p, ctx->mlpNorm.p, sv, dG.p));
            TRY_K(mlp_adam_step(st, ctx->mlpP.p, dG.p, dM1.p, dM2.p, lr, ++step));
            TRY_HIP(hipMemcpyAsync(dL.p + b, tb.loss.p, sizeof(float), hipMemcpyDeviceToDevice, st));
        }
        TRY_HIP(hipMemcpyAsync(lh.data(), dL.p, sizeof(float) * nb, hipMemcpyDeviceToHost, st));
        TRY_HIP(hipStreamSynchronize(st));
        double acc = 0;
        for (int64_t b = 0; b < nb; ++b) acc += lh[b];
        if (loss_per_epoch) loss_per_epoch[ep] = (float)(acc / std::max<int64_t>(nb, 1));
    }
    tb.release();
    for (auto *b : {&dX, &dT, &dXB, &dTB, &dG, &dM1, &dM2, &dL}) b->release();
    dIdx.release();
    return ALENS_OK;
}

}  // extern "C"

// The learned start point's BCQP energy on the last step's collision rows, and optionally its gradient
// with respect to the network parameters (into grad_dev, on the device).
//   gamma0 = network(features of the rows); g = A gamma0 + b from one start-point pass of the solver
//   (max_ite 0, no warm-start test); E = 1/2 gamma0' (g + b) = 1/2 gamma0' A gamma0 + b' gamma0.
//   dE/dgamma0 = g.  With gamma0 = s y (s the network's output scale) the mean-square loss of
//   mlp_gradient, mean((y - t)^2) over B rows, has gradient (2 / B)(y - t) dy; the targets
//   t = y - (B / 2) s g make that s g dy = dE.  Rows are padded to a multiple of 16 with copies of row 0
//   whose target is their own output (zero gradient).
// The solver's iterate buffer is overwritten; the step's outputs (gamma, velocities, rows) are not.
namespace {
struct EnergyBufs {
    TrainBufs tb;
    DBuf<float> X16, tgt, grad;
};
}  // namespace

static alens_status mlp_energy_eval(alens_ctx *ctx, double *energy, EnergyBufs *eb) {
    if (!ctx->have_step) return fail(ctx, ALENS_E_STATE, "mlp energy: no step has been run");
    if (!ctx->mlp_ready) return fail(ctx, ALENS_E_STATE, "mlp energy: no network parameters");
    if (ctx->mext > 0) return fail(ctx, ALENS_E_STATE, "mlp energy: the step had caller constraint blocks");
    const uint32_t M = ctx->mcol;
    if (M == 0) return fail(ctx, ALENS_E_STATE, "mlp energy: the last step had no constraints");
    hipStream_t st = ctx->st;
    TRY_HIP(ctx->feat.ensure(16 * (size_t)M));
    TRY_K(launch_features(st, M, ctx->gstride, ctx->cI.p, ctx->cJ.p, ctx->delta0.p, ctx->geo.p, ctx->posI.p,
                          ctx->posJ.p, ctx->dir.p, ctx->length.p, ctx->unc.p, ctx->have_xprev ? ctx->xprev.p : nullptr,
                          ctx->cfg.dt, ctx->feat.p));
    TRY_K(launch_mlp_infer_xg(st, M, ctx->feat.p, ctx->mlpP.p, ctx->mlpNorm.p, ctx->xg.p));
    const int32_t keep_ite = ctx->cfg.max_ite;
    ctx->cfg.max_ite = 0;
    alens_solve_report rep{};
    float ms = 0.f;
    alens_status s = run_solver(ctx, ctx->n, M, ctx->mob.p, false, &rep, &ms, ctx->rsys.p, ctx->smono.p, nullptr);
    ctx->cfg.max_ite = keep_ite;
    if (s != ALENS_OK && s != ALENS_E_NOCONV) return s;
    std::vector<double2> xg(M);
    std::vector<double> b(M);
    TRY_HIP(hipMemcpyAsync(xg.data(), ctx->xg.p, sizeof(double2) * M, hipMemcpyDeviceToHost, st));
    TRY_HIP(hipMemcpyAsync(b.data(), ctx->b.p, sizeof(double) * M, hipMemcpyDeviceToHost, st));
    float scale = 1.f;
    TRY_HIP(hipMemcpyAsync(&scale, ctx->mlpNorm.p + 2 * kMlpIn, sizeof(float), hipMemcpyDeviceToHost, st));
    TRY_HIP(hipStreamSynchronize(st));
    double e = 0.0;
    for (uint32_t k = 0; k < M; ++k) e += xg[k].x * (xg[k].y + b[k]);
    *energy = 0.5 * e;
    if (!std::isfinite(*energy)) return fail(ctx, ALENS_E_NONFINITE, "mlp energy: non-finite energy");
    if (!eb) return ALENS_OK;
    const int B = (int)((M + 15) / 16 * 16);
    TRY_HIP(eb->tb.ensure(B));
    TRY_HIP(eb->X16.ensure(16 * (size_t)B));
    TRY_HIP(eb->tgt.ensure((size_t)B));
    TRY_HIP(eb->grad.ensure(kMlpParams));
    std::vector<float> t((size_t)B);
    const double half_b = 0.5 * (double)B;
    for (uint32_t k = 0; k < M; ++k) {
        const double y = xg[k].x / (double)scale;
        t[k] = (float)(y - half_b * (double)scale * xg[k].y);
    }
    for (int k = (int)M; k < B; ++k) t[k] = (float)(xg[0].x / (double)scale);
    TRY_HIP(hipMemcpyAsync(eb->X16.p, ctx->feat.p, sizeof(float) * 16 * M, hipMemcpyDeviceToDevice, st));
    for (int k = (int)M; k < B; ++k)
        TRY_HIP(hipMemcpyAsync(eb->X16.p + 16 * (size_t)k, ctx->feat.p, sizeof(float) * 16, hipMemcpyDeviceToDevice, st));
    TRY_HIP(hipMemcpyAsync(eb->tgt.p, t.data(), sizeof(float) * B, hipMemcpyHostToDevice, st));
    MlpScratch sv = eb->tb.view();
    TRY_K(mlp_gradient(st, B, eb->X16.p, eb->tgt.p, ctx->mlpP.p, ctx->mlpNorm.p, sv, eb->grad.p));
    return ALENS_OK;
}

extern "C" {

alens_status alens_mlp_energy(alens_ctx *ctx, double *energy, float *grad) {
    if (!ctx || !energy) return ALENS_E_ARG;
    TRY_HIP(hipSetDevice(ctx->dev));
    EnergyBufs eb;
    alens_status s = mlp_energy_eval(ctx, energy, grad ? &eb : nullptr);
    if (s != ALENS_OK) return s;
    if (grad) {
        TRY_HIP(hipMemcpyAsync(grad, eb.grad.p, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost, ctx->st));
        TRY_HIP(hipStreamSynchronize(ctx->st));
    }
    return ALENS_OK;
}

alens_status alens_mlp_train_energy(alens_ctx *ctx, int32_t iters, float lr, double *energy) {
    if (ctx) drop_graph(ctx);  // the network is part of a captured step
    if (!ctx || iters < 0 || !(lr > 0.f)) return ALENS_E_ARG;
    TRY_HIP(hipSetDevice(ctx->dev));
    EnergyBufs eb;
    TRY_HIP(ctx->adam1.ensure(kMlpParams));
    TRY_HIP(ctx->adam2.ensure(kMlpParams));
    if (ctx->adam_step == 0) {
        TRY_K(launch_fill_u32(ctx->st, reinterpret_cast<uint32_t *>(ctx->adam1.p), 0u, kMlpParams));
        TRY_K(launch_fill_u32(ctx->st, reinterpret_cast<uint32_t *>(ctx->adam2.p), 0u, kMlpParams));
    }
    for (int32_t it = 0; it <= iters; ++it) {
        double e = 0.0;
        alens_status s = mlp_energy_eval(ctx, &e, it < iters ? &eb : nullptr);
        if (s != ALENS_OK) return s;
        if (energy) energy[it] = e;
        if (it < iters)
            TRY_K(mlp_adam_step(ctx->st, ctx->mlpP.p, eb.grad.p, ctx->adam1.p, ctx->adam2.p, lr, ++ctx->adam_step));
    }
    TRY_HIP(hipStreamSynchronize(ctx->st));
    return ALENS_OK;
}

}  // extern "C"

// Streaming copy used by alens_copy_probe: 16-byte words, grid-stride, four loads in flight per
// thread before their stores.  Launched as one 1024-thread block per CU (the fastest of the shapes
// measured on MI355X for 0.25-4 GiB: 4.9-5.3 TB/s).
__global__ __launch_bounds__(1024) void k_copy_probe(const uint4 *__restrict__ src, uint4 *__restrict__ dst,
                                                     size_t nvec) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < nvec; i += 4 * stride) {
        uint4 v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (i + u * stride < nvec) v[u] = src[i + u * stride];
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (i + u * stride < nvec) dst[i + u * stride] = v[u];
    }
}

__global__ void k_fill_probe(uint4 *__restrict__ dst, size_t nvec) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < nvec; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = make_uint4((uint32_t)i, 1u, 2u, 3u);
}

extern "C" {

alens_status alens_copy_probe(int32_t device, int64_t nbytes, int32_t reps, double *gbs) {
    if (nbytes < 16 || reps < 1 || !gbs) return ALENS_E_ARG;
    *gbs = 0.0;
    const size_t nvec = (size_t)nbytes / 16;
    if (hipSetDevice(device) != hipSuccess) return ALENS_E_HIP;
    uint4 *a = nullptr, *b = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    alens_status rc = ALENS_E_HIP;
    float ms = 0.f;
    int ncu = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || ncu < 1) ncu = 256;
    const unsigned grid = (unsigned)ncu, blk = 1024;
    if (hipMalloc((void **)&a, nvec * 16) != hipSuccess || hipMalloc((void **)&b, nvec * 16) != hipSuccess) {
        rc = ALENS_E_OOM;
        goto done;
    }
    if (hipStreamCreate(&st) != hipSuccess || hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
        goto done;
    k_fill_probe<<<grid, blk, 0, st>>>(a, nvec);
    if (hipGetLastError() != hipSuccess || hipEventRecord(e0, st) != hipSuccess) goto done;
    for (int32_t r = 0; r < reps; ++r) k_copy_probe<<<grid, blk, 0, st>>>(a, b, nvec);
    if (hipGetLastError() != hipSuccess || hipEventRecord(e1, st) != hipSuccess) goto done;
    if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess) goto done;
    *gbs = ms > 0.f ? 2.0 * (double)(nvec * 16) * reps / (ms * 1e-3) / 1e9 : 0.0;
    rc = ALENS_OK;
done:
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) (void)hipStreamDestroy(st);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    return rc;
}

alens_status alens_get_num_systems(const alens_ctx *ctx, int32_t *nsys) {
    if (!ctx || !nsys) return ALENS_E_ARG;
    *nsys = ctx->nsys;
    return ALENS_OK;
}

alens_status alens_get_system_reports(alens_ctx *ctx, alens_solve_report *reps) {
    if (!ctx || !reps) return ALENS_E_ARG;
    if (!ctx->have_step) return fail(ctx, ALENS_E_STATE, "no step has been run");
    for (int32_t sy = 0; sy < ctx->nsys; ++sy) reps[sy] = ctx->sysrep[sy];
    return ALENS_OK;
}

}  // extern "C"
