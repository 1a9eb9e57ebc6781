// k_collide.hip -- per-step collision collection on gfx950:
//   per-rod mobility + U_nc (Brownian), the tube cell list (each rod enters the cells its inflated
//   segment meets), LSD radix sort of (cell, rod) entries, cell ranges, pair sweep (FP64 segment
//   distance + cutoff, emission in the cell of the closest-point midpoint), canonical ordering of the
//   pairs, constraint rows, the degree order of the slices and the rod-major sliced adjacency used
//   by the solver.
// Reference roles: SylinderSystem::calcVelocityBrown / calcVelocityNonCon / collectPairCollision and
// ConstraintCollector::buildConstraintMatrixVector, reached from rodSystem.runStep()
// (SRC/TubuleSystem.cpp:223); block layout from SRC/TubuleSystem.cpp:719-736.
#include "alens_dev.hpp"
#include "alens_internal.hpp"

namespace alens {

#define LAUNCH_CHECK()                                                                           \
    do {                                                                                         \
        hipError_t e_ = hipGetLastError();                                                       \
        if (e_ != hipSuccess) return HipErr{e_, "kernel launch", __FILE__, __LINE__};             \
    } while (0)

static inline unsigned nblk(int64_t n, int t) { return (unsigned)((n + t - 1) / t); }

// ------------------------------------------------------------------------------------------------
// mobility + U_nc = M F_ext + U_B.  U_B = sqrt(2 kT/dt) M^{1/2} xi, xi from Philox4x32-10 keyed by
// the seed, counter (gid, step_lo, step_hi, c), Box-Muller on 53-bit uniforms.
// ------------------------------------------------------------------------------------------------
__global__ void k_mob_unc(int32_t n, const int32_t *__restrict__ rsys, const SysDev *__restrict__ sys,
                          const int32_t *__restrict__ gid, const double *__restrict__ dir,
                          const double *__restrict__ length, const double *__restrict__ radius,
                          const uint8_t *__restrict__ immovable, const double *__restrict__ fext,
                          CfgDev c, double *__restrict__ mob, double *__restrict__ unc) {
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int mono = sys[rsys[i]].mono;
    double ipar, iperp, irot;
    drag_inv(length[i], radius[i], c.mu, ipar, iperp, irot);
    double mb[6];
    mb[0] = dir[3 * i];
    mb[1] = dir[3 * i + 1];
    mb[2] = dir[3 * i + 2];
    if (immovable && immovable[i]) {
        mb[3] = 0.0;
        mb[4] = 0.0;
        mb[5] = 0.0;
    } else {
        mb[3] = iperp;
        mb[4] = ipar - iperp;
        mb[5] = irot;
    }
    double ub[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (c.kT > 0.0) {
        const double pre = sqrt(2.0 * c.kT / c.dt);
        double z[6];
        const uint32_t k0 = (uint32_t)c.seed, k1 = (uint32_t)(c.seed >> 32);
        const int64_t step = c.step_dev ? *c.step_dev : c.step;
#pragma unroll
        for (uint32_t q = 0; q < 3; ++q) {
            uint32_t ctr[4] = {(uint32_t)gid[i], (uint32_t)step, (uint32_t)((uint64_t)step >> 32), q};
            philox4x32_10(ctr, k0, k1);
            const double u1 =
                ((double)((((uint64_t)ctr[0]) << 21) ^ (uint64_t)(ctr[1] >> 11)) + 1.0) * 0x1p-53;
            const double u2 = (double)((((uint64_t)ctr[2]) << 21) ^ (uint64_t)(ctr[3] >> 11)) * 0x1p-53;
            const double R = sqrt(-2.0 * alog(u1));
            double sn, cs;
            asincos2pi(u2, sn, cs);
            z[2 * q] = R * cs;
            z[2 * q + 1] = R * sn;
        }
        const double sq = sqrt(mb[3]);
        const double sp = sqrt(mb[3] + mb[4]);
        const double sr = sqrt(mb[5]);
        const double tx = mb[0] * z[0] + mb[1] * z[1] + mb[2] * z[2];
        const double ct = (sp - sq) * tx;
        ub[0] = pre * (sq * z[0] + ct * mb[0]);
        ub[1] = pre * (sq * z[1] + ct * mb[1]);
        ub[2] = pre * (sq * z[2] + ct * mb[2]);
        ub[3] = pre * (sr * z[3]);
        ub[4] = pre * (sr * z[4]);
        ub[5] = pre * (sr * z[5]);
    }
    double u[6];
    if (fext) {
        double fe[6], ue[6];
#pragma unroll
        for (int k = 0; k < 6; ++k) fe[k] = fext[6 * i + k];
        apply_M(mb, fe, mono, ue);
#pragma unroll
        for (int k = 0; k < 6; ++k) u[k] = ue[k] + ub[k];
    } else {
#pragma unroll
        for (int k = 0; k < 6; ++k) u[k] = ub[k];
    }
    if (mono) {
        u[2] = 0.0;
        u[3] = 0.0;
        u[4] = 0.0;
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        mob[6 * i + k] = mb[k];
        unc[6 * i + k] = u[k];
    }
}

HipErr launch_mob_unc(hipStream_t st, const RodsDev &r, const double *length, const double *radius,
                      const uint8_t *immovable, const double *fext, const CfgDev &c, double *mob, double *unc) {
    if (r.n == 0) return hip_ok();
    k_mob_unc<<<nblk(r.n, 256), 256, 0, st>>>(r.n, r.rsys, r.sys, r.gid, r.dir, length, radius, immovable, fext, c,
                                              mob, unc);
    LAUNCH_CHECK();
    return hip_ok();
}

// ------------------------------------------------------------------------------------------------
// AABB (inflated by collision radius + colBuf/2 + margin) and its cell range per axis.  Periodic
// ranges are unwrapped integers; non-periodic ranges are clamped to the grid.
// ------------------------------------------------------------------------------------------------
__device__ inline int pmod_i(int a, int n) {
    const int m = a % n;
    return m < 0 ? m + n : m;
}

// Parameter interval [s0, s1] (within [-hh, hh]) where the segment x(s) = c + s t lies inside the
// slab of cell u on axis k grown by rho; a non-periodic axis's first and last cells reach to -inf /
// +inf (positions outside the box are clamped into them).  Returns false when empty.
__device__ inline bool slab_interval(const SysDev &S, int k, double ck, double tk, double hh, int u, double rho,
                                     double &s0, double &s1) {
    const double lo = (!S.pbc[k] && u <= 0) ? -INFINITY : S.box_lo[k] + (double)u * S.h[k] - rho;
    const double hi = (!S.pbc[k] && u >= S.nc[k] - 1) ? INFINITY : S.box_lo[k] + (double)(u + 1) * S.h[k] + rho;
    if (tk == 0.0 || hh == 0.0) return ck >= lo && ck <= hi;
    double sa = (lo - ck) / tk, sb = (hi - ck) / tk;
    if (sa > sb) {
        const double w = sa;
        sa = sb;
        sb = w;
    }
    s0 = sa > s0 ? sa : s0;
    s1 = sb < s1 ? sb : s1;
    return s0 <= s1;
}

// Calls f(u) for the cells (unwrapped indices, inside the box range a..b) that the segment grown by
// rho meets: for each (y, z) row, the parameter interval inside the row's grown slabs gives the
// segment's x extent there, and the row's cells are those whose grown x range meets it.  A superset
// of the cells within rho of the segment (rho carries a 1e-6 um margin, far above rounding).
template <class F>
__device__ inline void for_tube_cells(const SysDev &S, const double c[3], const double t[3], double hh,
                                      const int a[3], const int b[3], double rho, F f) {
    int u[3];
    for (u[2] = a[2]; u[2] <= b[2]; ++u[2]) {
        double z0 = -hh, z1 = hh;
        if (!slab_interval(S, 2, c[2], t[2], hh, u[2], rho, z0, z1)) continue;
        for (u[1] = a[1]; u[1] <= b[1]; ++u[1]) {
            double s0 = z0, s1 = z1;
            if (!slab_interval(S, 1, c[1], t[1], hh, u[1], rho, s0, s1)) continue;
            const double xa = c[0] + s0 * t[0], xb = c[0] + s1 * t[0];
            const double xlo = xa < xb ? xa : xb, xhi = xa < xb ? xb : xa;
            int ulo = (int)floor((xlo - rho - S.box_lo[0]) / S.h[0]);
            int uhi = (int)floor((xhi + rho - S.box_lo[0]) / S.h[0]);
            ulo = ulo < a[0] ? a[0] : ulo;
            uhi = uhi > b[0] ? b[0] : uhi;
            for (u[0] = ulo; u[0] <= uhi; ++u[0]) f(u);
        }
    }
}

__global__ void k_aabb(RodsDev r, CfgDev c, int32_t *__restrict__ clo, int32_t *__restrict__ chi,
                       double *__restrict__ blo, double *__restrict__ bhi, uint32_t *__restrict__ ecnt) {
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= r.n) return;
    const SysDev &S = r.sys[r.rsys[i]];
    const double hh = r.hcol[i];
    // one inflation per system (the largest radius): a contact's closest-point midpoint then lies
    // inside both rods' boxes and tubes, whatever their radii
    const double infl = S.infl;
    (void)c;
    int a3[3], b3[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const double hw = hh * fabs(r.dir[3 * i + k]);
        const double lo = (r.center[3 * i + k] - hw) - infl;
        const double hi = (r.center[3 * i + k] + hw) + infl;
        int a = (int)floor((lo - S.box_lo[k]) / S.h[k]);
        int b = (int)floor((hi - S.box_lo[k]) / S.h[k]);
        if (!S.pbc[k]) {
            a = a < 0 ? 0 : (a >= S.nc[k] ? S.nc[k] - 1 : a);
            b = b < 0 ? 0 : (b >= S.nc[k] ? S.nc[k] - 1 : b);
        }
        clo[3 * i + k] = a;
        chi[3 * i + k] = b;
        blo[3 * i + k] = lo;
        bhi[3 * i + k] = hi;
        a3[k] = a;
        b3[k] = b;
    }
    // entries: the cells of the box range that the inflated segment meets
    const double ci[3] = {r.center[3 * i], r.center[3 * i + 1], r.center[3 * i + 2]};
    const double ti[3] = {r.dir[3 * i], r.dir[3 * i + 1], r.dir[3 * i + 2]};
    uint32_t cnt = 0;
    for_tube_cells(S, ci, ti, hh, a3, b3, infl, [&](const int *) { ++cnt; });
    ecnt[i] = cnt;
}

HipErr launch_aabb(hipStream_t st, const RodsDev &r, const CfgDev &c, int32_t *clo, int32_t *chi, double *blo,
                   double *bhi, uint32_t *ecnt) {
    if (r.n == 0) return hip_ok();
    k_aabb<<<nblk(r.n, 256), 256, 0, st>>>(r, c, clo, chi, blo, bhi, ecnt);
    LAUNCH_CHECK();
    return hip_ok();
}

// ------------------------------------------------------------------------------------------------
// Exclusive scan of u32 (out has n + 1 entries, out[n] = total).  Tiles of 2048 (256 x 8):
// tile sums -> recursive scan of the sums -> per-tile scan with the tile offset.
// ------------------------------------------------------------------------------------------------
constexpr int kScanT = 256, kScanI = 8, kScanTile = kScanT * kScanI;

__device__ inline uint32_t wave_incl_scan(uint32_t v, int lane) {
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t t = __shfl_up(v, o, 64);
        if (lane >= o) v += t;
    }
    return v;
}

// block-wide exclusive scan of one value per thread; returns exclusive prefix, total via *tot
__device__ inline uint32_t block_excl_scan(uint32_t v, uint32_t *sh, uint32_t *tot) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint32_t inc = wave_incl_scan(v, lane);
    if (lane == 63) sh[w] = inc;
    __syncthreads();
    uint32_t off = 0, all = 0;
#pragma unroll
    for (int q = 0; q < kScanT / 64; ++q) {
        if (q < w) off += sh[q];
        all += sh[q];
    }
    __syncthreads();
    *tot = all;
    return off + inc - v;
}

__global__ void k_scan_tilesum(const uint32_t *__restrict__ in, int64_t n, uint32_t *__restrict__ sums) {
    __shared__ uint32_t sh[kScanT / 64];
    const int64_t base = (int64_t)blockIdx.x * kScanTile;
    uint32_t s = 0;
#pragma unroll
    for (int j = 0; j < kScanI; ++j) {
        const int64_t idx = base + (int64_t)threadIdx.x * kScanI + j;
        if (idx < n) s += in[idx];
    }
    uint32_t tot;
    block_excl_scan(s, sh, &tot);
    if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

__global__ void k_scan_tile(const uint32_t *__restrict__ in, int64_t n,
                            const uint32_t *__restrict__ tileoff, uint32_t *__restrict__ out) {
    __shared__ uint32_t sh[kScanT / 64];
    const int64_t base = (int64_t)blockIdx.x * kScanTile;
    uint32_t v[kScanI];
    uint32_t s = 0;
#pragma unroll
    for (int j = 0; j < kScanI; ++j) {
        const int64_t idx = base + (int64_t)threadIdx.x * kScanI + j;
        v[j] = idx < n ? in[idx] : 0u;
        s += v[j];
    }
    uint32_t tot;
    uint32_t pre = block_excl_scan(s, sh, &tot) + (tileoff ? tileoff[blockIdx.x] : 0u);
#pragma unroll
    for (int j = 0; j < kScanI; ++j) {
        const int64_t idx = base + (int64_t)threadIdx.x * kScanI + j;
        if (idx < n) out[idx] = pre;
        pre += v[j];
    }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == kScanT - 1) out[n] = pre;
}

static HipErr scan_level(hipStream_t st, const uint32_t *in, uint32_t *out, int64_t n,
                         ScanScratch &s, int lvl) {
    const int64_t nt = (n + kScanTile - 1) / kScanTile;
    if (nt <= 1) {
        k_scan_tile<<<1, kScanT, 0, st>>>(in, n, nullptr, out);
        LAUNCH_CHECK();
        return hip_ok();
    }
    if (lvl >= 3) return HipErr{hipErrorInvalidValue, "scan depth", __FILE__, __LINE__};
    ALENS_HIP_TRY(s.lvl[lvl].ensure((size_t)(2 * nt + 2)));
    uint32_t *sums = s.lvl[lvl].p;
    uint32_t *offs = s.lvl[lvl].p + nt + 1;
    k_scan_tilesum<<<(unsigned)nt, kScanT, 0, st>>>(in, n, sums);
    LAUNCH_CHECK();
    HipErr e = scan_level(st, sums, offs, nt, s, lvl + 1);
    if (!e.ok()) return e;
    k_scan_tile<<<(unsigned)nt, kScanT, 0, st>>>(in, n, offs, out);
    LAUNCH_CHECK();
    return hip_ok();
}

// allocations a scan / radix sort of n items makes, done ahead of time (a step captured into a graph
// must not allocate)
HipErr ensure_scan_scratch(ScanScratch &s, int64_t n) {
    for (int lvl = 0; lvl < 3; ++lvl) {
        const int64_t nt = (n + kScanTile - 1) / kScanTile;
        if (nt <= 1) break;
        ALENS_HIP_TRY(s.lvl[lvl].ensure((size_t)(2 * nt + 2)));
        n = nt;
    }
    return hip_ok();
}
HipErr ensure_radix_scratch(DBuf<uint32_t> &hist, ScanScratch &s, int64_t n);

HipErr exclusive_scan_u32(hipStream_t st, const uint32_t *in, uint32_t *out, int64_t n,
                          ScanScratch &s) {
    if (n == 0) return launch_fill_u32(st, out, 0u, 1);  // a kernel, not a memset node, under capture
    return scan_level(st, in, out, n, s, 0);
}


// ------------------------------------------------------------------------------------------------
// cell entries: rod i writes (wrapped cell id, i) for every cell of its range, rod-major, so the
// stable sort below leaves each cell's rods in ascending index order.
// ------------------------------------------------------------------------------------------------
__global__ void k_fill_entries(RodsDev r, const int32_t *__restrict__ clo, const int32_t *__restrict__ chi,
                               const uint32_t *__restrict__ eoff, uint32_t *__restrict__ keys,
                               uint32_t *__restrict__ vals, uint32_t ecap) {
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= r.n || eoff[i + 1] > ecap) return;  // past the capacity: the host sees eoff[n] > ecap
    const SysDev &g = r.sys[r.rsys[i]];
    uint32_t w = eoff[i];
    const uint32_t wend = eoff[i + 1];
    const double ci[3] = {r.center[3 * i], r.center[3 * i + 1], r.center[3 * i + 2]};
    const double ti[3] = {r.dir[3 * i], r.dir[3 * i + 1], r.dir[3 * i + 2]};
    const double hh = r.hcol[i];
    const int a3[3] = {clo[3 * i], clo[3 * i + 1], clo[3 * i + 2]};
    const int b3[3] = {chi[3 * i], chi[3 * i + 1], chi[3 * i + 2]};
    // the same cells, in the same order, as k_aabb counted (never more than its count)
    for_tube_cells(g, ci, ti, hh, a3, b3, g.infl, [&](const int *u) {
        if (w >= wend) return;
        const uint32_t cx = (uint32_t)pmod_i(u[0], g.nc[0]);
        const uint32_t cy = (uint32_t)pmod_i(u[1], g.nc[1]);
        const uint32_t cz = (uint32_t)pmod_i(u[2], g.nc[2]);
        keys[w] = g.cell_off + (cz * (uint32_t)g.nc[1] + cy) * (uint32_t)g.nc[0] + cx;
        vals[w] = (uint32_t)i;
        ++w;
    });
}

HipErr launch_fill_entries(hipStream_t st, const RodsDev &r, const int32_t *clo, const int32_t *chi,
                           const uint32_t *eoff, uint32_t *keys, uint32_t *vals, uint32_t ecap) {
    if (r.n == 0) return hip_ok();
    k_fill_entries<<<nblk(r.n, 256), 256, 0, st>>>(r, clo, chi, eoff, keys, vals, ecap);
    LAUNCH_CHECK();
    return hip_ok();
}

__global__ void k_fill_u32(uint32_t *p, uint32_t v, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
__global__ void k_zero_many(ZeroJobs z) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (int j = 0; j < z.count; ++j)
        if (i < z.n[j]) z.p[j][i] = 0u;
}
HipErr launch_zero_many(hipStream_t st, const ZeroJobs &z) {
    int64_t mx = 0;
    for (int j = 0; j < z.count; ++j) mx = z.n[j] > mx ? z.n[j] : mx;
    if (mx == 0) return hip_ok();
    k_zero_many<<<nblk(mx, 256), 256, 0, st>>>(z);
    LAUNCH_CHECK();
    return hip_ok();
}

HipErr launch_fill_u32(hipStream_t st, uint32_t *p, uint32_t v, int64_t n) {
    if (n == 0) return hip_ok();
    k_fill_u32<<<nblk(n, 256), 256, 0, st>>>(p, v, n);
    LAUNCH_CHECK();
    return hip_ok();
}

// ------------------------------------------------------------------------------------------------
// LSD radix sort of (u32 key, u32 value), 8-bit digits, stable.  Per pass: tile histograms
// (digit-major), one exclusive scan over all (digit, tile) counters, then a stable scatter where
// each 64-lane wave ranks equal digits with 8 ballots and the tile keeps a running offset per
// digit across its rounds.  Tile = 256 threads x 8 rounds; round r covers items r*256 .. r*256+255.
// ------------------------------------------------------------------------------------------------
constexpr int kRsT = 256, kRsR = 8, kRsTile = kRsT * kRsR;

__global__ void k_radix_hist(const uint32_t *__restrict__ keys, int64_t ncap, int shift, int ntiles,
                             uint32_t *__restrict__ hist, const uint32_t *__restrict__ nd) {
    const int64_t n = dcount((uint32_t)ncap, nd);
    __shared__ uint32_t lh[256];
    lh[threadIdx.x] = 0;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * kRsTile;
#pragma unroll
    for (int r = 0; r < kRsR; ++r) {
        const int64_t idx = base + r * kRsT + threadIdx.x;
        if (idx < n) atomicAdd(&lh[(keys[idx] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[(int64_t)threadIdx.x * ntiles + blockIdx.x] = lh[threadIdx.x];
}

__global__ void k_radix_scatter(const uint32_t *__restrict__ keys, const uint32_t *__restrict__ vals,
                                int64_t ncap, int shift, int ntiles, const uint32_t *__restrict__ off,
                                uint32_t *__restrict__ okeys, uint32_t *__restrict__ ovals,
                                const uint32_t *__restrict__ nd) {
    const int64_t n = dcount((uint32_t)ncap, nd);
    __shared__ uint32_t run[256];
    __shared__ uint32_t wc[kRsT / 64][256];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    run[threadIdx.x] = off[(int64_t)threadIdx.x * ntiles + blockIdx.x];
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const int64_t base = (int64_t)blockIdx.x * kRsTile;
    for (int r = 0; r < kRsR; ++r) {
#pragma unroll
        for (int q = 0; q < kRsT / 64; ++q) wc[q][threadIdx.x] = 0;
        __syncthreads();
        const int64_t idx = base + r * kRsT + threadIdx.x;
        const bool valid = idx < n;
        const uint32_t key = valid ? keys[idx] : 0u;
        const uint32_t val = valid ? vals[idx] : 0u;
        const uint32_t d = (key >> shift) & 255u;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int bit = 0; bit < 8; ++bit) {
            const bool on = (d >> bit) & 1u;
            const uint64_t m = __ballot(valid && on);
            peers &= on ? m : ~m;
        }
        const uint32_t rank = (uint32_t)__popcll(peers & lt);
        if (valid && rank == 0) wc[w][d] = (uint32_t)__popcll(peers);
        __syncthreads();
        if (valid) {
            uint32_t pos = run[d] + rank;
            for (int q = 0; q < w; ++q) pos += wc[q][d];
            okeys[pos] = key;
            ovals[pos] = val;
        }
        __syncthreads();
        uint32_t add = 0;
#pragma unroll
        for (int q = 0; q < kRsT / 64; ++q) add += wc[q][threadIdx.x];
        run[threadIdx.x] += add;
        __syncthreads();
    }
}

HipErr ensure_radix_scratch(DBuf<uint32_t> &hist, ScanScratch &s, int64_t n) {
    if (n <= 1) return hip_ok();
    const int64_t ntiles = (n + kRsTile - 1) / kRsTile;
    const int64_t nh = 256 * ntiles;
    ALENS_HIP_TRY(hist.ensure((size_t)(2 * nh + 1)));
    return ensure_scan_scratch(s, nh);
}

HipErr radix_sort_pairs(hipStream_t st, uint32_t *keys, uint32_t *vals, uint32_t *keys_alt,
                        uint32_t *vals_alt, int64_t n, int keybits, DBuf<uint32_t> &hist,
                        ScanScratch &s, bool &result_in_alt, const uint32_t *nd) {
    result_in_alt = false;
    if (n <= 1 || keybits <= 0) return hip_ok();
    const int ntiles = (int)((n + kRsTile - 1) / kRsTile);
    const int64_t nh = (int64_t)256 * ntiles;
    ALENS_HIP_TRY(hist.ensure((size_t)(2 * nh + 1)));
    uint32_t *h = hist.p, *ho = hist.p + nh;
    uint32_t *ka = keys, *va = vals, *kb = keys_alt, *vb = vals_alt;
    for (int shift = 0; shift < keybits; shift += 8) {
        k_radix_hist<<<ntiles, kRsT, 0, st>>>(ka, n, shift, ntiles, h, nd);
        LAUNCH_CHECK();
        HipErr e = exclusive_scan_u32(st, h, ho, nh, s);
        if (!e.ok()) return e;
        k_radix_scatter<<<ntiles, kRsT, 0, st>>>(ka, va, n, shift, ntiles, ho, kb, vb, nd);
        LAUNCH_CHECK();
        uint32_t *t = ka;
        ka = kb;
        kb = t;
        t = va;
        va = vb;
        vb = t;
        result_in_alt = !result_in_alt;
    }
    return hip_ok();
}

// ------------------------------------------------------------------------------------------------
// cell ranges in the sorted entry list (cells without entries keep start = end = 0)
// ------------------------------------------------------------------------------------------------
__global__ void k_cell_ranges(const uint32_t *__restrict__ keys, int64_t Ecap,
                              uint32_t *__restrict__ cstart, uint32_t *__restrict__ cend,
                              const uint32_t *__restrict__ Ed) {
    const int64_t E = dcount((uint32_t)Ecap, Ed);
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= E) return;
    const uint32_t k = keys[i];
    if (i == 0 || keys[i - 1] != k) cstart[k] = (uint32_t)i;
    if (i == E - 1 || keys[i + 1] != k) cend[k] = (uint32_t)(i + 1);
}

HipErr launch_cell_ranges(hipStream_t st, const uint32_t *keys, int64_t E, uint32_t *cstart,
                          uint32_t *cend, const uint32_t *Ed) {
    if (E == 0) return hip_ok();
    k_cell_ranges<<<nblk(E, 256), 256, 0, st>>>(keys, E, cstart, cend, Ed);
    LAUNCH_CHECK();
    return hip_ok();
}

// ------------------------------------------------------------------------------------------------
// Pair sweep.  For entries a < b of one cell (rods i < j): the image shift of j follows from the
// unwrapped cell indices (each rod's box range is shorter than the periodic length), then AABB
// reject, the FP64 segment distance and delta0 < colBuf.  A contact is emitted only in the cell that
// holds the midpoint of its two closest points: that point lies within both rods' inflated tubes
// (half the contact distance is at most the inflation), so both rods have an entry there, and every
// shared cell computes the same midpoint -- each (pair, image) is emitted exactly once.  Per-rod
// counts of both sides feed the ordering.
// ------------------------------------------------------------------------------------------------
// One thread per sorted entry.  The k(k - 1) / 2 pairs of a cell of k entries are dealt round the
// cell: entry r takes partners r + d (mod k) for d = 1 .. (k - 1) / 2, and for even k the entries
// r < k / 2 also take r + k / 2, so every lane of the cell runs (k - 1) / 2 or k / 2 pairs instead of
// k - 1 - r.  Each pair is still evaluated with its lower entry as rod i, by the same expressions.
// Lanes of a wave run the same trip count (the wave maximum) so emission stays wave-aggregated.
// Found pairs are staged in LDS (an LDS counter per block) and reserved in the global list with one
// atomic per block at the end; a block that fills its stage reserves the rest one by one.
constexpr int kSweepT = 256;
constexpr int kStage = 512;
constexpr int kSweepPre = 64;                          // entries staged before the block's own
constexpr int kSweepS = kSweepPre + kSweepT + 64;      // and after them

// Box data of one sweep entry (rod id, unwrapped cell index, inflated box), from the block's LDS
// stage or, outside it, from global memory.  The unwrapped index of a rod in a periodic direction is
// the copy of the entry's cell that lies in the rod's box range: cl + ((c - cl) mod nc) for box range
// start cl and cell c; two rods of one cell differ in it by a whole number of periods, their image
// shift.  It depends on the entry only, so the stage computes it once instead of once per pair.
struct SweepBox {
    uint32_t id;
    int u[3];
    double lo[3], hi[3];
};

// four waves per SIMD (at most 128 VGPRs): measured faster than three with more registers, and than
// five with spills
__global__ __launch_bounds__(kSweepT) __attribute__((amdgpu_waves_per_eu(4)))
void k_sweep_entries(RodsDev r, CfgDev c, const uint32_t *__restrict__ keys, int64_t E,
                     const uint32_t *__restrict__ cstart, const uint32_t *__restrict__ cend,
                     const uint32_t *__restrict__ vals, const int32_t *__restrict__ clo,
                     const double *__restrict__ blo, const double *__restrict__ bhi,
                     uint32_t *__restrict__ pair_count, uint32_t cap, uint32_t *__restrict__ pI,
                     uint32_t *__restrict__ pJ, uint32_t *__restrict__ pImg, uint32_t *__restrict__ cntI,
                     uint32_t *__restrict__ cntJ, const uint32_t *__restrict__ Ed) {
    E = dcount((uint32_t)E, Ed);
    __shared__ uint32_t s_n, s_base;
    __shared__ uint32_t s_pi[kStage], s_pj[kStage], s_pm[kStage];
    // box data of entries e0 - kSweepPre .. e0 + kSweepT + 64, structure of arrays
    __shared__ uint32_t s_id[kSweepS];
    __shared__ int s_u[3][kSweepS];
    __shared__ double s_lo[3][kSweepS], s_hi[3][kSweepS];
    const int lane = threadIdx.x & 63;
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const int64_t e0 = (int64_t)blockIdx.x * blockDim.x;
    const int64_t es = e0 - kSweepPre;  // first staged entry
    const int64_t e = e0 + threadIdx.x;
    // cell coordinates of a cell id in its system's grid
    auto cell_xyz = [](uint32_t lc, const int (&n)[3], int (&x)[3]) {
        x[0] = (int)(lc % (uint32_t)n[0]);
        x[1] = (int)((lc / (uint32_t)n[0]) % (uint32_t)n[1]);
        x[2] = (int)(lc / ((uint32_t)n[0] * (uint32_t)n[1]));
    };
    for (int q = threadIdx.x; q < kSweepS; q += blockDim.x) {
        const int64_t eq = es + q;
        if (eq < 0) continue;
        if (eq >= E) break;
        const uint32_t j = vals[eq];
        const SysDev &g = r.sys[r.rsys[j]];
        const int gn[3] = {g.nc[0], g.nc[1], g.nc[2]};
        int x[3];
        cell_xyz(keys[eq] - g.cell_off, gn, x);
        s_id[q] = j;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const int cl = clo[3 * j + k];
            s_u[k][q] = g.pbc[k] ? cl + pmod_i(x[k] - cl, gn[k]) : x[k];
            s_lo[k][q] = blo[3 * j + k];
            s_hi[k][q] = bhi[3 * j + k];
        }
    }
    if (threadIdx.x == 0) s_n = 0;
    __syncthreads();
    // this lane's entry: cell, place r in the cell of kc entries, pair count
    uint32_t cnt = 0, cs = 0, kc = 1, rr = 0;
    int cc[3] = {0, 0, 0};
    int pbc[3] = {0, 0, 0}, nc[3] = {1, 1, 1};
    double len[3] = {0, 0, 0}, blo0[3] = {0, 0, 0}, hcell[3] = {1, 1, 1};
    // an entry of this lane's cell (its own, or a partner outside the stage)
    auto box_of = [&](int64_t ex) {
        SweepBox w;
        const int64_t q = ex - es;
        if (q >= 0 && q < kSweepS) {
            w.id = s_id[q];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                w.u[k] = s_u[k][q];
                w.lo[k] = s_lo[k][q];
                w.hi[k] = s_hi[k][q];
            }
        } else {
            w.id = vals[ex];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const int cl = clo[3 * w.id + k];
                w.u[k] = pbc[k] ? cl + pmod_i(cc[k] - cl, nc[k]) : cc[k];
                w.lo[k] = blo[3 * w.id + k];
                w.hi[k] = bhi[3 * w.id + k];
            }
        }
        return w;
    };
    SweepBox own{0u, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    if (e < E) {
        const uint32_t cell = keys[e];
        cs = cstart[cell];
        kc = cend[cell] - cs;
        rr = (uint32_t)e - cs;
        cnt = (kc - 1u) / 2u + ((kc % 2u == 0u && rr < kc / 2u) ? 1u : 0u);
        own = box_of(e);  // always staged
        const SysDev &g = r.sys[r.rsys[own.id]];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            pbc[k] = g.pbc[k];
            nc[k] = g.nc[k];
            len[k] = g.len[k];
            blo0[k] = g.box_lo[k];
            hcell[k] = g.h[k];
        }
        cell_xyz(cell - g.cell_off, nc, cc);
    }
    uint32_t maxc = cnt;
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        const uint32_t v = __shfl_xor(maxc, o, 64);
        maxc = v > maxc ? v : maxc;
    }
    for (uint32_t t = 0; t < maxc; ++t) {
        bool pass = false;
        uint32_t i = 0, j = 0;
        int m[3] = {0, 0, 0};
        if (t < cnt) {
            uint32_t r2 = rr + 1u + t;
            if (r2 >= kc) r2 -= kc;
            const int64_t ep = (int64_t)cs + r2;
            const SweepBox pw = box_of(ep);
            // rod i is the pair's lower entry
            const bool lo_own = e < ep;
            const SweepBox &wi = lo_own ? own : pw;
            const SweepBox &wj = lo_own ? pw : own;
            i = wi.id;
            j = wj.id;
            int ui[3];
            bool ok = true;
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                ui[k] = wi.u[k];
                const int du = ui[k] - wj.u[k];  // a whole number of periods (0 if not periodic)
                m[k] = du == 0 ? 0 : du / nc[k];
                const double sh = pbc[k] ? (double)m[k] * len[k] : 0.0;
                ok = ok && (wj.lo[k] + sh <= wi.hi[k]) && (wi.lo[k] <= wj.hi[k] + sh);
            }
            if (ok) {
                double ci[3], ti[3], cj[3], tj[3];
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    ci[k] = r.center[3 * i + k];
                    ti[k] = r.dir[3 * i + k];
                    const double ck = r.center[3 * j + k];
                    cj[k] = pbc[k] ? ck + (double)m[k] * len[k] : ck;
                    tj[k] = r.dir[3 * j + k];
                }
                const double hi_ = r.hcol[i], ri = r.rcol[i], hj = r.hcol[j], rj = r.rcol[j];
                const SegRes sg = segseg(ci[0], ci[1], ci[2], ti[0], ti[1], ti[2], hi_, cj[0], cj[1], cj[2], tj[0], tj[1],
                                         tj[2], hj);
                const double d0 = (sg.dist - ri) - rj;
                pass = d0 < c.col_buf;
                // the cell of the closest-point midpoint, in i's unwrapped frame
                const double D[3] = {sg.D0, sg.D1, sg.D2};
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    const double mk = (ci[k] + sg.s * ti[k]) - 0.5 * D[k];
                    int um = (int)floor((mk - blo0[k]) / hcell[k]);
                    if (!pbc[k]) um = um < 0 ? 0 : (um >= nc[k] ? nc[k] - 1 : um);
                    pass = pass && um == ui[k];
                }
            }
        }
        const uint64_t mask = __ballot(pass);
        if (mask) {
            const int leader = __ffsll((unsigned long long)mask) - 1;
            uint32_t base = 0;
            if (lane == leader) base = atomicAdd(&s_n, (uint32_t)__popcll(mask));
            base = __shfl(base, leader, 64);
            if (pass) {
                const uint32_t q = base + (uint32_t)__popcll(mask & lt);
                const uint32_t im = img_pack(m[0], m[1], m[2]);
                if (q < (uint32_t)kStage) {
                    s_pi[q] = i;
                    s_pj[q] = j;
                    s_pm[q] = im;
                } else {
                    const uint32_t slot = atomicAdd(pair_count, 1u);
                    if (slot < cap) {
                        pI[slot] = i;
                        pJ[slot] = j;
                        pImg[slot] = im;
                    }
                }
                atomicAdd(&cntI[i], 1u);
                atomicAdd(&cntJ[j], 1u);
            }
        }
    }
    __syncthreads();
    const uint32_t ns = s_n < (uint32_t)kStage ? s_n : (uint32_t)kStage;
    if (threadIdx.x == 0) s_base = ns ? atomicAdd(pair_count, ns) : 0u;
    __syncthreads();
    for (uint32_t q = threadIdx.x; q < ns; q += kSweepT) {
        const uint32_t slot = s_base + q;
        if (slot < cap) {
            pI[slot] = s_pi[q];
            pJ[slot] = s_pj[q];
            pImg[slot] = s_pm[q];
        }
    }
}

HipErr launch_sweep(hipStream_t st, const RodsDev &r, const CfgDev &c, const uint32_t *keys, int64_t E,
                    const uint32_t *cstart, const uint32_t *cend, const uint32_t *vals, const int32_t *clo,
                    const double *blo, const double *bhi, uint32_t *pair_count, uint32_t cap, uint32_t *pI,
                    uint32_t *pJ, uint32_t *pImg, uint32_t *cntI, uint32_t *cntJ, const uint32_t *Ed) {
    if (E == 0) return hip_ok();
    k_sweep_entries<<<nblk(E, kSweepT), kSweepT, 0, st>>>(r, c, keys, E, cstart, cend, vals, clo, blo, bhi, pair_count,
                                                          cap, pI, pJ, pImg, cntI, cntJ, Ed);
    LAUNCH_CHECK();
    return hip_ok();
}

// ------------------------------------------------------------------------------------------------
// Canonical order: pairs grouped by I (scatter to rowStart[I] + atomic cursor), then each I-group
// sorted by J by one thread (groups hold a handful of contacts).  The J-side lists hold constraint
// indices, sorted ascending per rod.
// ------------------------------------------------------------------------------------------------
__global__ void k_scatter_I(uint32_t m, const uint32_t *__restrict__ pI, const uint32_t *__restrict__ pJ,
                            const uint32_t *__restrict__ pImg, const uint32_t *__restrict__ rowStart,
                            uint32_t *__restrict__ cur, int32_t *__restrict__ cI,
                            int32_t *__restrict__ cJ, uint32_t *__restrict__ cImg, const uint32_t *__restrict__ md) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= dcount(m, md)) return;
    const uint32_t i = pI[p];
    const uint32_t pos = rowStart[i] + atomicAdd(&cur[i], 1u);
    cI[pos] = (int32_t)i;
    cJ[pos] = (int32_t)pJ[p];
    cImg[pos] = pImg[p];
}

__global__ void k_sort_I(int32_t n, const uint32_t *__restrict__ rowStart, int32_t *__restrict__ cJ,
                         uint32_t *__restrict__ cImg) {
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t s0 = rowStart[i], s1 = rowStart[i + 1];
    for (uint32_t a = s0 + 1; a < s1; ++a) {
        const int32_t kj = cJ[a];
        const uint32_t km = cImg[a];
        uint32_t b = a;
        while (b > s0 && cJ[b - 1] > kj) {
            cJ[b] = cJ[b - 1];
            cImg[b] = cImg[b - 1];
            --b;
        }
        cJ[b] = kj;
        cImg[b] = km;
    }
}

HipErr launch_order_I(hipStream_t st, int32_t n, uint32_t m, const uint32_t *pI,
                      const uint32_t *pJ, const uint32_t *pImg, const uint32_t *rowStart,
                      uint32_t *cur, int32_t *cI, int32_t *cJ, uint32_t *cImg, const uint32_t *md) {
    if (m == 0) return hip_ok();
    k_scatter_I<<<nblk(m, 256), 256, 0, st>>>(m, pI, pJ, pImg, rowStart, cur, cI, cJ, cImg, md);
    LAUNCH_CHECK();
    k_sort_I<<<nblk(n, 256), 256, 0, st>>>(n, rowStart, cJ, cImg);
    LAUNCH_CHECK();
    return hip_ok();
}

__global__ void k_scatter_J(uint32_t m, const int32_t *__restrict__ cJ, const uint32_t *__restrict__ jStart,
                            uint32_t *__restrict__ cur, uint32_t *__restrict__ jlist, const uint32_t *__restrict__ md) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= dcount(m, md)) return;
    const uint32_t j = (uint32_t)cJ[k];
    jlist[jStart[j] + atomicAdd(&cur[j], 1u)] = k;
}

__global__ void k_sort_J(int32_t n, const uint32_t *__restrict__ jStart, uint32_t *__restrict__ jlist) {
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t s0 = jStart[i], s1 = jStart[i + 1];
    for (uint32_t a = s0 + 1; a < s1; ++a) {
        const uint32_t v = jlist[a];
        uint32_t b = a;
        while (b > s0 && jlist[b - 1] > v) {
            jlist[b] = jlist[b - 1];
            --b;
        }
        jlist[b] = v;
    }
}

HipErr launch_order_J(hipStream_t st, int32_t n, uint32_t m, const int32_t *cJ,
                      const uint32_t *jStart, uint32_t *cur, uint32_t *jlist, const uint32_t *md) {
    if (m == 0) return hip_ok();
    k_scatter_J<<<nblk(m, 256), 256, 0, st>>>(m, cJ, jStart, cur, jlist, md);
    LAUNCH_CHECK();
    k_sort_J<<<nblk(n, 256), 256, 0, st>>>(n, jStart, jlist);
    LAUNCH_CHECK();
    return hip_ok();
}

// ------------------------------------------------------------------------------------------------
// Constraint rows: delta0, n (J -> I), posI = s tI, posJ = u tJ, rI = posI x n, rJ = posJ x n,
// b = delta0 + dt (D^T U_nc).  geo component-major: geo[q * m + k], q = 0..8 = {n, rI, rJ}.
// ------------------------------------------------------------------------------------------------
__global__ void k_rows(RodsDev r, CfgDev c, uint32_t m, uint32_t gs, const int32_t *__restrict__ cI,
                       const int32_t *__restrict__ cJ, const uint32_t *__restrict__ cImg,
                       const double *__restrict__ unc, double *__restrict__ delta0,
                       double *__restrict__ geo, double *__restrict__ posI,
                       double *__restrict__ posJ, double *__restrict__ b, const uint32_t *__restrict__ md) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= dcount(m, md)) return;
    const int32_t i = cI[k], j = cJ[k];
    const uint32_t im = cImg[k];
    const SysDev &S = r.sys[r.rsys[i]];
    double cj[3];
#pragma unroll
    for (int a = 0; a < 3; ++a)
        cj[a] = S.pbc[a] ? r.center[3 * j + a] + (double)img_get(im, a) * S.len[a] : r.center[3 * j + a];
    const double tI0 = r.dir[3 * i], tI1 = r.dir[3 * i + 1], tI2 = r.dir[3 * i + 2];
    const double tJ0 = r.dir[3 * j], tJ1 = r.dir[3 * j + 1], tJ2 = r.dir[3 * j + 2];
    const SegRes sg = segseg(r.center[3 * i], r.center[3 * i + 1], r.center[3 * i + 2], tI0, tI1, tI2,
                             r.hcol[i], cj[0], cj[1], cj[2], tJ0, tJ1, tJ2, r.hcol[j]);
    double nk[3];
    contact_normal(sg, tI0, tI1, tI2, tJ0, tJ1, tJ2, nk);
    const double pI[3] = {sg.s * tI0, sg.s * tI1, sg.s * tI2};
    const double pJ[3] = {sg.u * tJ0, sg.u * tJ1, sg.u * tJ2};
    double g9[9];
    g9[0] = nk[0];
    g9[1] = nk[1];
    g9[2] = nk[2];
    g9[3] = pI[1] * nk[2] - pI[2] * nk[1];
    g9[4] = pI[2] * nk[0] - pI[0] * nk[2];
    g9[5] = pI[0] * nk[1] - pI[1] * nk[0];
    g9[6] = pJ[1] * nk[2] - pJ[2] * nk[1];
    g9[7] = pJ[2] * nk[0] - pJ[0] * nk[2];
    g9[8] = pJ[0] * nk[1] - pJ[1] * nk[0];
    const double d0 = (sg.dist - r.rcol[i]) - r.rcol[j];
    delta0[k] = d0;
#pragma unroll
    for (int a = 0; a < 9; ++a) geo[(size_t)a * gs + k] = g9[a];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        posI[3 * (size_t)k + a] = pI[a];
        posJ[3 * (size_t)k + a] = pJ[a];
    }
    b[k] = d0 + c.dt * dt_row(g9, unc + 6 * (size_t)i, unc + 6 * (size_t)j);
}

HipErr launch_rows(hipStream_t st, const RodsDev &r, const CfgDev &c, uint32_t m, uint32_t gs, const int32_t *cI,
                   const int32_t *cJ, const uint32_t *cImg, const double *unc, double *delta0,
                   double *geo, double *posI, double *posJ, double *b, const uint32_t *md) {
    if (m == 0) return hip_ok();
    k_rows<<<nblk(m, 256), 256, 0, st>>>(r, c, m, gs, cI, cJ, cImg, unc, delta0, geo, posI, posJ, b, md);
    LAUNCH_CHECK();
    return hip_ok();
}

// Caller-supplied rows k = k0 + e (the constraint pool): normal n = normI (uploaded into nrm),
// arms posI / posJ given; rI = posI x n, rJ = posJ x n, b = delta0 + dt (D^T U_nc).
__global__ void k_rows_ext(uint32_t me, uint32_t k0, uint32_t gs, const int32_t *__restrict__ cI,
                           const int32_t *__restrict__ cJ, const double *__restrict__ nrm,
                           const double *__restrict__ posI, const double *__restrict__ posJ,
                           const double *__restrict__ delta0, const double *__restrict__ unc, double dt,
                           double *__restrict__ geo, double *__restrict__ b) {
    const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= me) return;
    const uint32_t k = k0 + e;
    const double *nk = nrm + 3 * (size_t)e;
    const double *pI = posI + 3 * (size_t)k, *pJ = posJ + 3 * (size_t)k;
    double g9[9];
    g9[0] = nk[0];
    g9[1] = nk[1];
    g9[2] = nk[2];
    g9[3] = pI[1] * nk[2] - pI[2] * nk[1];
    g9[4] = pI[2] * nk[0] - pI[0] * nk[2];
    g9[5] = pI[0] * nk[1] - pI[1] * nk[0];
    g9[6] = pJ[1] * nk[2] - pJ[2] * nk[1];
    g9[7] = pJ[2] * nk[0] - pJ[0] * nk[2];
    g9[8] = pJ[0] * nk[1] - pJ[1] * nk[0];
#pragma unroll
    for (int a = 0; a < 9; ++a) geo[(size_t)a * gs + k] = g9[a];
    b[k] = delta0[k] + dt * dt_row(g9, unc + 6 * (size_t)cI[k], unc + 6 * (size_t)cJ[k]);
}

HipErr launch_rows_ext(hipStream_t st, uint32_t me, uint32_t k0, uint32_t gs, const int32_t *cI, const int32_t *cJ,
                       const double *nrm, const double *posI, const double *posJ, const double *delta0,
                       const double *unc, double dt, double *geo, double *b) {
    if (me == 0) return hip_ok();
    k_rows_ext<<<nblk(me, 256), 256, 0, st>>>(me, k0, gs, cI, cJ, nrm, posI, posJ, delta0, unc, dt, geo, b);
    LAUNCH_CHECK();
    return hip_ok();
}

// ------------------------------------------------------------------------------------------------
// Rod-major adjacency in ascending constraint index: the J-side list (constraints with J = i all
// have I < i, so they precede i's own rows) followed by i's I-side rows.  Each entry carries the
// rod's column block of D: (n, rI) on the I side, (-n, -rJ) on the J side.  Stored sliced: slice s
// (rods kSlice*s ..) holds kSlice * (largest degree in the slice) slots, entry j of rod i at
// sellStart[s] + j * kSlice + i % kSlice.
// ------------------------------------------------------------------------------------------------
__device__ inline uint32_t rod_degree(const uint32_t *rowStart, const uint32_t *jStart, const uint32_t *extStart,
                                      int32_t i) {
    uint32_t deg = (rowStart[i + 1] - rowStart[i]) + (jStart[i + 1] - jStart[i]);
    if (extStart) deg += extStart[i + 1] - extStart[i];
    return deg;
}

// one block per window of W rods: a bitonic sort in LDS of the keys (descending degree,
// then ascending index); rods past n sort last
template <int W>
__global__ __launch_bounds__(W) void k_slot_order(int32_t n, const uint32_t *__restrict__ rowStart,
                                                  const uint32_t *__restrict__ jStart,
                                                  const uint32_t *__restrict__ extStart, uint32_t *__restrict__ sord,
                                                  uint32_t *__restrict__ spos) {
    __shared__ uint64_t s_key[W];
    const int32_t base = blockIdx.x * W;
    const int t = threadIdx.x;
    const int32_t i = base + t;
    s_key[t] = i < n ? ((uint64_t)(0xffffffffu - rod_degree(rowStart, jStart, extStart, i)) << 32) | (uint32_t)t
                     : ~0ull;
    __syncthreads();
    for (int k = 2; k <= W; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            const int u = t ^ j;
            if (u > t) {
                const uint64_t a = s_key[t], b = s_key[u];
                if ((a > b) == ((t & k) == 0)) {
                    s_key[t] = b;
                    s_key[u] = a;
                }
            }
            __syncthreads();
        }
    }
    if (i < n) {
        const uint32_t r = (uint32_t)base + (uint32_t)(s_key[t] & 0xffffffffu);
        sord[i] = r;         // position i holds rod r
        spos[r] = (uint32_t)i;
    }
}

HipErr launch_slot_order(hipStream_t st, int32_t n, const uint32_t *rowStart, const uint32_t *jStart,
                         const uint32_t *extStart, uint32_t *sord, uint32_t *spos, int win) {
    if (n == 0) return hip_ok();
    if (win == kOrderWinPasses)
        k_slot_order<kOrderWinPasses><<<nblk(n, kOrderWinPasses), kOrderWinPasses, 0, st>>>(n, rowStart, jStart,
                                                                                         extStart, sord, spos);
    else
        k_slot_order<kOrderWin><<<nblk(n, kOrderWin), kOrderWin, 0, st>>>(n, rowStart, jStart, extStart, sord, spos);
    LAUNCH_CHECK();
    return hip_ok();
}

__global__ void k_slice_sizes(int32_t n, const uint32_t *__restrict__ rowStart, const uint32_t *__restrict__ jStart,
                              const uint32_t *__restrict__ extStart, uint32_t *__restrict__ slice_len,
                              const uint32_t *__restrict__ sord) {
    const int32_t p = blockIdx.x * blockDim.x + threadIdx.x;  // one lane per slice position, one wave per slice
    uint32_t deg = 0;
    if (p < n) deg = rod_degree(rowStart, jStart, extStart, sord ? (int32_t)sord[p] : p);
#pragma unroll
    for (int h = 32; h >= 1; h >>= 1) {
        const uint32_t o = __shfl_xor(deg, h, 64);
        deg = o > deg ? o : deg;
    }
    if ((threadIdx.x & 63) == 0 && p < n) slice_len[p >> 6] = deg * (uint32_t)kSlice;
}

// Capacity guards between the phases of a step whose sizes are not read back: a count past its
// capacity becomes 0 downstream (and the per-rod pair counts are cleared), so that no later kernel
// indexes past a buffer; the host compares the raw counts with the capacities after the step.
__global__ void k_guard_entries(const uint32_t *__restrict__ eoff_n, uint32_t ecap, uint32_t *__restrict__ ee) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *ee = *eoff_n <= ecap ? *eoff_n : 0u;
}
__global__ void k_guard_pairs(const uint32_t *__restrict__ count, uint32_t cap, uint32_t *__restrict__ mv,
                              uint32_t *__restrict__ cntI, uint32_t *__restrict__ cntJ, int32_t n) {
    const uint32_t c = *count;
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) *mv = c <= cap ? c : 0u;
    if (c > cap && i <= n) {
        cntI[i] = 0u;
        cntJ[i] = 0u;
    }
}
__global__ void k_guard_slots(const uint32_t *__restrict__ slots, uint32_t cap, uint32_t *__restrict__ mv) {
    if (blockIdx.x == 0 && threadIdx.x == 0 && *slots > cap) *mv = 0u;
}
HipErr launch_guard_slots(hipStream_t st, const uint32_t *slots, uint32_t cap, uint32_t *mv) {
    k_guard_slots<<<1, 64, 0, st>>>(slots, cap, mv);
    LAUNCH_CHECK();
    return hip_ok();
}
HipErr launch_guard_entries(hipStream_t st, const uint32_t *eoff_n, uint32_t ecap, uint32_t *ee) {
    k_guard_entries<<<1, 64, 0, st>>>(eoff_n, ecap, ee);
    LAUNCH_CHECK();
    return hip_ok();
}
HipErr launch_guard_pairs(hipStream_t st, const uint32_t *count, uint32_t cap, uint32_t *mv, uint32_t *cntI,
                          uint32_t *cntJ, int32_t n) {
    k_guard_pairs<<<nblk((int64_t)n + 1, 256), 256, 0, st>>>(count, cap, mv, cntI, cntJ, n);
    LAUNCH_CHECK();
    return hip_ok();
}

// Chunk table: system s owns rows [con_off[s], con_off[s+1]) (rows are grouped by rod I; the caller
// rows close the last system) and chunks [chunk_off[s], chunk_off[s+1]) of 512 rows.
__global__ void k_chunk_table(int32_t nsys, int32_t n, const SysDev *__restrict__ sys,
                              const uint32_t *__restrict__ rowStart, uint32_t me, uint32_t *__restrict__ con_off,
                              uint32_t *__restrict__ chunk_off, int32_t *__restrict__ nchunks,
                              const uint32_t *__restrict__ mv) {
    if (threadIdx.x != 0) return;
    if (mv && *mv == 0u) {  // no rows, or a step past its capacities: nothing to solve
        for (int32_t s = 0; s <= nsys; ++s) con_off[s] = chunk_off[s] = 0u;
        *nchunks = 0;
        return;
    }
    uint32_t acc = 0;
    for (int32_t s = 0; s < nsys; ++s) {
        const uint32_t a = rowStart[sys[s].rod_off];
        const uint32_t b = s + 1 < nsys ? rowStart[sys[s + 1].rod_off] : rowStart[n] + me;
        con_off[s] = a;
        chunk_off[s] = acc;
        acc += (b - a + (uint32_t)kRedChunk - 1u) / (uint32_t)kRedChunk;
    }
    con_off[nsys] = rowStart[n] + me;
    chunk_off[nsys] = acc;
    *nchunks = (int32_t)acc;
}
__global__ void k_chunk_sys(int32_t nsys, const uint32_t *__restrict__ chunk_off, const int32_t *__restrict__ nchunks,
                            int32_t *__restrict__ chunk_sys, int32_t cap) {
    const int32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= cap) return;
    if (c >= *nchunks) {
        chunk_sys[c] = -1;
        return;
    }
    int32_t lo = 0, hi = nsys;  // last s with chunk_off[s] <= c
    while (hi - lo > 1) {
        const int32_t mid = (lo + hi) / 2;
        if (chunk_off[mid] <= (uint32_t)c)
            lo = mid;
        else
            hi = mid;
    }
    chunk_sys[c] = lo;
}
HipErr launch_chunk_table(hipStream_t st, int32_t nsys, int32_t n, const SysDev *sys, const uint32_t *rowStart,
                          uint32_t me, uint32_t *con_off, uint32_t *chunk_off, int32_t *nchunks, int32_t *chunk_sys,
                          int32_t chunk_cap, const uint32_t *mv) {
    k_chunk_table<<<1, 64, 0, st>>>(nsys, n, sys, rowStart, me, con_off, chunk_off, nchunks, mv);
    LAUNCH_CHECK();
    if (chunk_cap > 0) {
        k_chunk_sys<<<nblk(chunk_cap, 256), 256, 0, st>>>(nsys, chunk_off, nchunks, chunk_sys, chunk_cap);
        LAUNCH_CHECK();
    }
    return hip_ok();
}

HipErr launch_slice_sizes(hipStream_t st, int32_t n, const uint32_t *rowStart, const uint32_t *jStart,
                          const uint32_t *extStart, uint32_t *slice_len, const uint32_t *sord) {
    static_assert(kSlice == 64, "one wave per slice");
    static_assert(kOrderWin % kSlice == 0 && kOrderWinPasses % kSlice == 0, "order windows hold whole slices");
    if (n == 0) return hip_ok();
    k_slice_sizes<<<nblk(n, 256), 256, 0, st>>>(n, rowStart, jStart, extStart, slice_len, sord);
    LAUNCH_CHECK();
    return hip_ok();
}

__global__ void k_adjacency(int32_t n, uint32_t m, const uint32_t *__restrict__ rowStart,
                            const uint32_t *__restrict__ jStart, const uint32_t *__restrict__ jlist,
                            const uint32_t *__restrict__ extStart, const uint32_t *__restrict__ extEnt, uint32_t mc,
                            const double *__restrict__ geo, const uint32_t *__restrict__ sellStart, uint64_t epad,
                            uint32_t *__restrict__ adjStart, uint32_t *__restrict__ sellK,
                            double *__restrict__ sellW, const uint32_t *__restrict__ spos) {
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > n) return;
    // past the slot capacity every rod gets degree 0 (the host sees sellStart[last] > epad)
    const bool fits = sellStart[(n + kSlice - 1) / kSlice] <= epad;
    adjStart[i] = fits ? rowStart[i] + jStart[i] + (extStart ? extStart[i] : 0u) : 0u;
    if (i == n || !fits) return;
    // rod-major: one thread per rod reads its own lists in order and writes its slice position's
    // slots (measured faster than walking positions, whose rods' lists are scattered)
    const uint32_t p = spos ? spos[i] : (uint32_t)i;
    const uint32_t j0 = jStart[i], j1 = jStart[i + 1];
    uint64_t idx = sellStart[p >> 6] + (p & (kSlice - 1));
    for (uint32_t t = j0; t < j1; ++t, idx += kSlice) {
        const uint32_t k = jlist[t];
        sellK[idx] = k | 0x80000000u;
        sellW[idx] = -geo[k];
        sellW[epad + idx] = -geo[(size_t)m + k];
        sellW[2 * epad + idx] = -geo[2 * (size_t)m + k];
        sellW[3 * epad + idx] = -geo[6 * (size_t)m + k];
        sellW[4 * epad + idx] = -geo[7 * (size_t)m + k];
        sellW[5 * epad + idx] = -geo[8 * (size_t)m + k];
    }
    for (uint32_t k = rowStart[i]; k < rowStart[i + 1]; ++k, idx += kSlice) {
        sellK[idx] = k;
#pragma unroll
        for (int c = 0; c < 6; ++c) sellW[c * epad + idx] = geo[(size_t)c * m + k];
    }
    // caller rows (k = mc + e >= every collision row), ascending; top bit of the entry = J side
    if (extStart) {
        for (uint32_t t = extStart[i]; t < extStart[i + 1]; ++t, idx += kSlice) {
            const uint32_t ent = extEnt[t];
            const uint32_t k = mc + (ent & 0x7fffffffu);
            if (ent & 0x80000000u) {
                sellK[idx] = k | 0x80000000u;
                sellW[idx] = -geo[k];
                sellW[epad + idx] = -geo[(size_t)m + k];
                sellW[2 * epad + idx] = -geo[2 * (size_t)m + k];
                sellW[3 * epad + idx] = -geo[6 * (size_t)m + k];
                sellW[4 * epad + idx] = -geo[7 * (size_t)m + k];
                sellW[5 * epad + idx] = -geo[8 * (size_t)m + k];
            } else {
                sellK[idx] = k;
#pragma unroll
                for (int c = 0; c < 6; ++c) sellW[c * epad + idx] = geo[(size_t)c * m + k];
            }
        }
    }
}

HipErr launch_adjacency(hipStream_t st, int32_t n, uint32_t m, const uint32_t *rowStart,
                        const uint32_t *jStart, const uint32_t *jlist, const uint32_t *extStart,
                        const uint32_t *extEnt, uint32_t mc, const double *geo, const uint32_t *sellStart,
                        uint64_t epad, uint32_t *adjStart, uint32_t *sellK, double *sellW, const uint32_t *spos) {
    k_adjacency<<<nblk((int64_t)n + 1, 256), 256, 0, st>>>(n, m, rowStart, jStart, jlist, extStart, extEnt, mc, geo,
                                                          sellStart, epad, adjStart, sellK, sellW, spos);
    LAUNCH_CHECK();
    return hip_ok();
}

// ------------------------------------------------------------------------------------------------
// Spatial reordering: key = Morton code of the rod centre's cell (10 bits per axis), stable sort,
// then every per-rod array is gathered into the new order.
// ------------------------------------------------------------------------------------------------
__device__ inline uint32_t spread3(uint32_t v) {
    v &= 0x3ffu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

__global__ void k_morton_keys(RodsDev r, int32_t i0, int32_t n, uint32_t *__restrict__ keys,
                              uint32_t *__restrict__ vals) {
    const int32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int32_t i = i0 + t;
    const SysDev &g = r.sys[r.rsys[i]];
    const double *center = r.center;
    uint32_t q[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        int a = (int)floor((center[3 * i + k] - g.box_lo[k]) / g.h[k]);
        a = g.pbc[k] ? pmod_i(a, g.nc[k]) : (a < 0 ? 0 : (a >= g.nc[k] ? g.nc[k] - 1 : a));
        q[k] = (uint32_t)a;
    }
    keys[t] = spread3(q[0]) | (spread3(q[1]) << 1) | (spread3(q[2]) << 2);
    vals[t] = (uint32_t)i;
}

HipErr launch_morton_keys(hipStream_t st, const RodsDev &r, int32_t i0, int32_t n, uint32_t *keys,
                          uint32_t *vals) {
    if (n == 0) return hip_ok();
    k_morton_keys<<<nblk(n, 256), 256, 0, st>>>(r, i0, n, keys, vals);
    LAUNCH_CHECK();
    return hip_ok();
}

__global__ void k_permute_rods(int32_t n, const uint32_t *__restrict__ perm, const int32_t *__restrict__ gid,
                               const double *__restrict__ center, const double *__restrict__ dir,
                               const double *__restrict__ length, const double *__restrict__ radius,
                               const double *__restrict__ hcol, const double *__restrict__ rcol,
                               const uint8_t *__restrict__ immov, int32_t *__restrict__ gid2,
                               double *__restrict__ center2, double *__restrict__ dir2, double *__restrict__ length2,
                               double *__restrict__ radius2, double *__restrict__ hcol2, double *__restrict__ rcol2,
                               uint8_t *__restrict__ immov2) {
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t s = perm[i];
    gid2[i] = gid[s];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        center2[3 * i + k] = center[3 * (size_t)s + k];
        dir2[3 * i + k] = dir[3 * (size_t)s + k];
    }
    length2[i] = length[s];
    radius2[i] = radius[s];
    hcol2[i] = hcol[s];
    rcol2[i] = rcol[s];
    if (immov) immov2[i] = immov[s];
}

HipErr launch_permute_rods(hipStream_t st, int32_t n, const uint32_t *perm, const int32_t *gid,
                           const double *center, const double *dir, const double *length, const double *radius,
                           const double *hcol, const double *rcol, const uint8_t *immov, int32_t *gid2,
                           double *center2, double *dir2, double *length2, double *radius2, double *hcol2,
                           double *rcol2, uint8_t *immov2) {
    if (n == 0) return hip_ok();
    k_permute_rods<<<nblk(n, 256), 256, 0, st>>>(n, perm, gid, center, dir, length, radius, hcol, rcol, immov, gid2,
                                                 center2, dir2, length2, radius2, hcol2, rcol2, immov2);
    LAUNCH_CHECK();
    return hip_ok();
}

}  // namespace alens
