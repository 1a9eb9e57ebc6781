/*
 * alens_oracle.c -- FP64 CPU restatement of the aLENS/SimToolbox collision-constraint path.
 *
 * TEST INFRASTRUCTURE ONLY: the correctness oracle and the timed CPU baseline.  Nothing in the
 * product loads it.  Build: oracle/build_oracle.sh (gcc -O3 -fopenmp -ffp-contract=off).
 *
 * Reference anchors (the SimToolbox sources that implement this path are absent from the reference
 * tree; the visible call sites and docs fix the conventions):
 *   - rod = segment centre +- (L/2) * direction, direction = q * e_z    (SRC/TubuleBind.hpp:54-67)
 *   - sphere iff length < 2 * radius                                    (SRC/TubuleBind.hpp:99)
 *   - constraint block fields, normal from J to I, pos relative to the
 *     rod centre, lab-frame points, unilateral flags                    (SRC/TubuleSystem.cpp:719-736)
 *   - collision radius/length ratios, colBuf gap threshold, conResTol,
 *     conMaxIte, conSolverChoice 0 = BBPGD                              (docs/source/quickstart.md:340-358)
 *   - KBT < 0 disables thermal noise                                    (Examples/MixMotorSliding/RunConfig.yaml:20)
 *   - per-step order: Brownian -> non-constraint velocity -> collision
 *     collection -> BCQP solve -> write-back                            (SRC/TubuleSystem.cpp:223, rodSystem.runStep)
 *
 * Parity pins (tests/test_oracle_*.py): segment distance against 50-digit mpmath solutions; pair
 * sets against brute force; D against a finite-difference kinematic derivative and its adjoint;
 * mobility against closed forms; Philox against the Random123 known-answer vectors; BBPGD against
 * exact active-set solutions of small LCPs and the single-contact closed form; velocities against
 * the Brownian covariance 2 kT M / dt.
 */
#include "alens_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_PI 3.14159265358979323846
#define ORC_EPS_PAR 1e-30     /* |tI x tJ|^2 <= EPS_PAR * a * c  =>  treated as parallel */
#define ORC_AABB_MARGIN 1e-6  /* um added to every AABB half-width so rounding never drops a pair */
#define ORC_RED_THREADS 256   /* reduction layout: 256 lanes = 4 waves of 64 */
#define ORC_RED_CHUNK 512     /* elements per reduction chunk (2 passes of 256 lanes) */

static int g_threads = 0;

void orc_set_threads(int32_t nthreads) { g_threads = nthreads; }
int32_t orc_get_threads(void) {
#ifdef _OPENMP
    return g_threads > 0 ? g_threads : omp_get_max_threads();
#else
    return 1;
#endif
}
#ifdef _OPENMP
#define ORC_NT num_threads(orc_get_threads())
#endif

/* ------------------------------------------------------------------------------------------------
 * Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as 1, 2, 3", SC'11).
 * ------------------------------------------------------------------------------------------------ */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        if (r > 0) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

/* ------------------------------------------------------------------------------------------------
 * Natural log: x = m * 2^e with m in [sqrt(1/2), sqrt(2)], log m = 2 atanh(z), z = (m-1)/(m+1),
 * atanh series to z^25 (|z| <= 0.1716), ln 2 split into a 32-bit head and a tail.
 * Valid for finite x > 0 (subnormals are pre-scaled).
 * ------------------------------------------------------------------------------------------------ */
double orc_log(double x) {
    uint64_t bits;
    memcpy(&bits, &x, 8);
    int e = (int)((bits >> 52) & 0x7ff);
    int adj = 0;
    if (e == 0) { /* subnormal: scale by 2^54 */
        double y = x * 18014398509481984.0;
        memcpy(&bits, &y, 8);
        e = (int)((bits >> 52) & 0x7ff);
        adj = -54;
    }
    e = e - 1023 + adj;
    uint64_t mb = (bits & 0x000FFFFFFFFFFFFFull) | 0x3FF0000000000000ull;
    double m;
    memcpy(&m, &mb, 8);
    if (m > 1.4142135623730951) {
        m = m * 0.5;
        e = e + 1;
    }
    const double z = (m - 1.0) / (m + 1.0);
    const double z2 = z * z;
    double p = 1.0 / 25.0;
    p = 1.0 / 23.0 + z2 * p;
    p = 1.0 / 21.0 + z2 * p;
    p = 1.0 / 19.0 + z2 * p;
    p = 1.0 / 17.0 + z2 * p;
    p = 1.0 / 15.0 + z2 * p;
    p = 1.0 / 13.0 + z2 * p;
    p = 1.0 / 11.0 + z2 * p;
    p = 1.0 / 9.0 + z2 * p;
    p = 1.0 / 7.0 + z2 * p;
    p = 1.0 / 5.0 + z2 * p;
    p = 1.0 / 3.0 + z2 * p;
    const double lm = 2.0 * z + (2.0 * z) * (z2 * p);
    const double de = (double)e;
    const double ln2_hi = 6.93147180369123816490e-01; /* ln 2 rounded to 32 significant bits */
    const double ln2_lo = 1.90821492927058770002e-10; /* ln 2 - ln2_hi */
    return de * ln2_hi + (lm + de * ln2_lo);
}

/* sin(2 pi u), cos(2 pi u): quadrant q = nearest integer to 4u, reduced angle (4u - q) pi/2 in
 * [-pi/4, pi/4], Taylor polynomials to degree 17 / 18. */
void orc_sincos2pi(double u, double *s_out, double *c_out) {
    const double u4 = 4.0 * u;
    const double qf = floor(u4 + 0.5);
    const double f = u4 - qf;
    const double th = f * 1.57079632679489661923;
    const double t2 = th * th;
    /* sin th = th + th t2 ps,  ps = -1/3! + t2/5! - t2^2/7! + ... + t2^7/17! */
    double ps = 1.0 / 355687428096000.0;
    ps = -1.0 / 1307674368000.0 + t2 * ps;
    ps = 1.0 / 6227020800.0 + t2 * ps;
    ps = -1.0 / 39916800.0 + t2 * ps;
    ps = 1.0 / 362880.0 + t2 * ps;
    ps = -1.0 / 5040.0 + t2 * ps;
    ps = 1.0 / 120.0 + t2 * ps;
    ps = -1.0 / 6.0 + t2 * ps;
    const double sn = th + th * (t2 * ps);
    /* cos th = 1 + t2 pc,  pc = -1/2! + t2/4! - ... - t2^8/18! */
    double pc = -1.0 / 6402373705728000.0;
    pc = 1.0 / 20922789888000.0 + t2 * pc;
    pc = -1.0 / 87178291200.0 + t2 * pc;
    pc = 1.0 / 479001600.0 + t2 * pc;
    pc = -1.0 / 3628800.0 + t2 * pc;
    pc = 1.0 / 40320.0 + t2 * pc;
    pc = -1.0 / 720.0 + t2 * pc;
    pc = 1.0 / 24.0 + t2 * pc;
    pc = -0.5 + t2 * pc;
    const double cs = 1.0 + t2 * pc;
    const int q = ((int)qf) & 3;
    if (q == 0) {
        *s_out = sn;
        *c_out = cs;
    } else if (q == 1) {
        *s_out = cs;
        *c_out = -sn;
    } else if (q == 2) {
        *s_out = -sn;
        *c_out = -cs;
    } else {
        *s_out = -cs;
        *c_out = sn;
    }
}

double orc_sqrt(double x) { return sqrt(x); }

/* ------------------------------------------------------------------------------------------------
 * Segment-segment closest points.  Segment I: cI + s tI, s in [-hI, hI]; segment J: cJ + u tJ,
 * u in [-hJ, hJ] (cJ already shifted to the periodic image).  Minimises |r + s tI - u tJ|^2,
 * r = cI - cJ, by the clamped alternating solution of the 2x2 normal equations (Ericson,
 * Real-Time Collision Detection, sec. 5.1.9): s from the infinite lines (or, for parallel
 * segments, the midpoint of the overlap of I with J's projection), clamp, u = argmin given s,
 * and if u is clamped, s = clamp(argmin given u).
 * ------------------------------------------------------------------------------------------------ */
static inline double clampd(double x, double lo, double hi) { return x < lo ? lo : (x > hi ? hi : x); }

void orc_segseg(const double cI[3], const double tI[3], double hI, const double cJ[3],
                const double tJ[3], double hJ, double out[6]) {
    const double r0 = cI[0] - cJ[0], r1 = cI[1] - cJ[1], r2 = cI[2] - cJ[2];
    const double a = tI[0] * tI[0] + tI[1] * tI[1] + tI[2] * tI[2];
    const double b = tI[0] * tJ[0] + tI[1] * tJ[1] + tI[2] * tJ[2];
    const double c = tJ[0] * tJ[0] + tJ[1] * tJ[1] + tJ[2] * tJ[2];
    const double d = tI[0] * r0 + tI[1] * r1 + tI[2] * r2;
    const double e = tJ[0] * r0 + tJ[1] * r1 + tJ[2] * r2;
    /* Lagrange identity: a c - b^2 = |tI x tJ|^2 and b e - c d = (tI x tJ).(tJ x r); the cross
     * products keep relative accuracy for nearly parallel segments. */
    const double w0 = tI[1] * tJ[2] - tI[2] * tJ[1];
    const double w1 = tI[2] * tJ[0] - tI[0] * tJ[2];
    const double w2 = tI[0] * tJ[1] - tI[1] * tJ[0];
    const double v0 = tJ[1] * r2 - tJ[2] * r1;
    const double v1 = tJ[2] * r0 - tJ[0] * r2;
    const double v2 = tJ[0] * r1 - tJ[1] * r0;
    const double det = w0 * w0 + w1 * w1 + w2 * w2;
    double s;
    if (det > ORC_EPS_PAR * (a * c)) {
        s = (w0 * v0 + w1 * v1 + w2 * v2) / det;
    } else {
        const double sp = (b * hJ - d) / a;
        const double sm = (-(b * hJ) - d) / a;
        double lo = sp < sm ? sp : sm;
        double hi = sp < sm ? sm : sp;
        lo = lo < -hI ? -hI : lo;
        hi = hi > hI ? hI : hi;
        s = 0.5 * (lo + hi);
    }
    s = clampd(s, -hI, hI);
    double u = (b * s + e) / c;
    if (u < -hJ) {
        u = -hJ;
        s = clampd((b * u - d) / a, -hI, hI);
    } else if (u > hJ) {
        u = hJ;
        s = clampd((b * u - d) / a, -hI, hI);
    }
    const double D0 = (r0 + s * tI[0]) - u * tJ[0];
    const double D1 = (r1 + s * tI[1]) - u * tJ[1];
    const double D2 = (r2 + s * tI[2]) - u * tJ[2];
    out[0] = s;
    out[1] = u;
    out[2] = sqrt(D0 * D0 + D1 * D1 + D2 * D2);
    out[3] = D0;
    out[4] = D1;
    out[5] = D2;
}

/* unit normal from J to I; at zero distance the rule falls back to tI x tJ, then to tI x e_axis */
static void contact_normal(const double D[3], double dist, const double tI[3], const double tJ[3],
                           double n[3]) {
    if (dist > 0.0) {
        n[0] = D[0] / dist;
        n[1] = D[1] / dist;
        n[2] = D[2] / dist;
        return;
    }
    double v0 = tI[1] * tJ[2] - tI[2] * tJ[1];
    double v1 = tI[2] * tJ[0] - tI[0] * tJ[2];
    double v2 = tI[0] * tJ[1] - tI[1] * tJ[0];
    double vn = sqrt(v0 * v0 + v1 * v1 + v2 * v2);
    if (!(vn > 0.0)) {
        const double ax = fabs(tI[0]), ay = fabs(tI[1]), az = fabs(tI[2]);
        if (ax <= ay && ax <= az) {
            v0 = 0.0; v1 = tI[2]; v2 = -tI[1];
        } else if (ay <= az) {
            v0 = -tI[2]; v1 = 0.0; v2 = tI[0];
        } else {
            v0 = tI[1]; v1 = -tI[0]; v2 = 0.0;
        }
        vn = sqrt(v0 * v0 + v1 * v1 + v2 * v2);
    }
    n[0] = v0 / vn;
    n[1] = v1 / vn;
    n[2] = v2 / vn;
}

/* ------------------------------------------------------------------------------------------------
 * Fixed reduction order (both the oracle and the device solver sum in exactly this order):
 * element e of a list goes to lane e % 256 as its (e / 256)-th term; each lane accumulates its
 * terms left to right from 0.0; each 64-lane wave folds by halving (lane l += lane l + h for
 * h = 32, 16, ..., 1); the four wave totals are added left to right.  Long vectors are cut into
 * 512-element chunks, each chunk is summed this way, and the chunk totals are summed the same way.
 * ------------------------------------------------------------------------------------------------ */
double orc_block_sum(const double *v, int64_t n) {
    double acc[ORC_RED_THREADS];
    for (int t = 0; t < ORC_RED_THREADS; ++t) {
        double s = 0.0;
        for (int64_t k = t; k < n; k += ORC_RED_THREADS) s = s + v[k];
        acc[t] = s;
    }
    for (int w = 0; w < ORC_RED_THREADS / 64; ++w) {
        double *a = acc + 64 * w;
        for (int h = 32; h >= 1; h >>= 1)
            for (int l = 0; l < h; ++l) a[l] = a[l] + a[l + h];
    }
    return ((acc[0] + acc[64]) + acc[128]) + acc[192];
}

double orc_chunked_sum(const double *v, int64_t n) {
    const int64_t nc = (n + ORC_RED_CHUNK - 1) / ORC_RED_CHUNK;
    if (nc == 0) return orc_block_sum(v, 0);
    double *part = (double *)malloc(sizeof(double) * nc);
    for (int64_t c = 0; c < nc; ++c) {
        int64_t len = n - c * ORC_RED_CHUNK;
        if (len > ORC_RED_CHUNK) len = ORC_RED_CHUNK;
        part[c] = orc_block_sum(v + c * ORC_RED_CHUNK, len);
    }
    double s = orc_block_sum(part, nc);
    free(part);
    return s;
}

/* ------------------------------------------------------------------------------------------------
 * Rod collision geometry
 * ------------------------------------------------------------------------------------------------ */
static inline double col_half(double L, double r, const orc_cfg *cfg) {
    return (L < r * 2.0) ? 0.0 : 0.5 * (L * cfg->len_col_ratio);
}
static inline double col_rad(double r, const orc_cfg *cfg) { return r * cfg->diam_col_ratio; }

static int box_ok(int32_t n, const double *length, const double *radius, const orc_cfg *cfg) {
    double lmax = 0.0, rmax = 0.0;
    for (int32_t i = 0; i < n; ++i) {
        double h = col_half(length[i], radius[i], cfg);
        double r = col_rad(radius[i], cfg);
        if (2.0 * h > lmax) lmax = 2.0 * h;
        if (r > rmax) rmax = r;
    }
    for (int k = 0; k < 3; ++k) {
        if (!cfg->pbc[k]) continue;
        const double len = cfg->box_hi[k] - cfg->box_lo[k];
        if (!(len >= 2.0 * (lmax + 2.0 * rmax + cfg->col_buf))) return 0;
    }
    return 1;
}

/* J's centre moved by img periods along each periodic axis */
static inline void shifted_centre(const double *cj, const int32_t *img, const orc_cfg *cfg,
                                  double out[3]) {
    for (int k = 0; k < 3; ++k) {
        const double len = cfg->box_hi[k] - cfg->box_lo[k];
        out[k] = cfg->pbc[k] ? cj[k] + (double)img[k] * len : cj[k];
    }
}

static inline double pair_delta0(const double *center, const double *dir, const double *length,
                                 const double *radius, const orc_cfg *cfg, int32_t i, int32_t j,
                                 const int32_t img[3]) {
    double cJ[3], o[6];
    shifted_centre(center + 3 * j, img, cfg, cJ);
    orc_segseg(center + 3 * i, dir + 3 * i, col_half(length[i], radius[i], cfg), cJ, dir + 3 * j,
               col_half(length[j], radius[j], cfg), o);
    return (o[2] - col_rad(radius[i], cfg)) - col_rad(radius[j], cfg);
}

/* Brute force over all i < j with the minimal image of the centre difference. */
int64_t orc_pairs_brute(int32_t n, const double *center, const double *dir, const double *length,
                        const double *radius, const orc_cfg *cfg, int64_t cap, int32_t *I,
                        int32_t *J, int32_t *img) {
    if (!box_ok(n, length, radius, cfg)) return -1;
    int64_t *cnt = (int64_t *)calloc((size_t)n + 1, sizeof(int64_t));
    for (int pass = 0; pass < 2; ++pass) {
#pragma omp parallel for schedule(dynamic, 16) ORC_NT
        for (int32_t i = 0; i < n; ++i) {
            int64_t w = pass ? cnt[i] : 0;
            for (int32_t j = i + 1; j < n; ++j) {
                int32_t m[3];
                for (int k = 0; k < 3; ++k) {
                    const double len = cfg->box_hi[k] - cfg->box_lo[k];
                    m[k] = cfg->pbc[k]
                               ? -(int32_t)floor((center[3 * j + k] - center[3 * i + k]) / len + 0.5)
                               : 0;
                }
                if (pair_delta0(center, dir, length, radius, cfg, i, j, m) < cfg->col_buf) {
                    if (pass && w < cap) {
                        I[w] = i;
                        J[w] = j;
                        img[3 * w] = m[0];
                        img[3 * w + 1] = m[1];
                        img[3 * w + 2] = m[2];
                    }
                    ++w;
                }
            }
            if (!pass) cnt[i + 1] = w;
        }
        if (!pass)
            for (int32_t i = 0; i < n; ++i) cnt[i + 1] += cnt[i];
    }
    int64_t total = cnt[n];
    free(cnt);
    return total;
}

/* ------------------------------------------------------------------------------------------------
 * Cell list: every rod is inserted in each cell its inflated AABB touches (inflation = collision
 * radius + colBuf/2 + margin).  In a cell, the pair (i, j) is examined only if, on every axis, the
 * cell is the first cell of i's range or of j's range (unwrapped), which visits each (pair, image)
 * once.  The image shift follows from the unwrapped cell indices.
 * ------------------------------------------------------------------------------------------------ */
typedef struct {
    int nc[3];
    double h[3];
} grid_t;

static void grid_setup(int32_t n, const double *length, const double *radius, const orc_cfg *cfg,
                       grid_t *g) {
    double hmax_infl = 0.0, lsum = 0.0;
    int32_t nrod = 0;
    for (int32_t i = 0; i < n; ++i) {
        const double infl = col_rad(radius[i], cfg) + 0.5 * cfg->col_buf + ORC_AABB_MARGIN;
        if (infl > hmax_infl) hmax_infl = infl;
        lsum += 2.0 * col_half(length[i], radius[i], cfg);
        ++nrod;
    }
    double h = cfg->cell_size;
    if (!(h > 0.0)) {
        h = nrod ? 0.5 * lsum / nrod : 1.0;
        if (h < 2.0 * hmax_infl) h = 2.0 * hmax_infl;
        if (!(h > 0.0)) h = 1.0;
    }
    const double maxcells = (double)(8 * (int64_t)n > 4096 ? 8 * (int64_t)n : 4096);
    for (;;) {
        double tot = 1.0;
        for (int k = 0; k < 3; ++k) {
            const double len = cfg->box_hi[k] - cfg->box_lo[k];
            int c = (int)floor(len / h);
            if (c < 1) c = 1;
            g->nc[k] = c;
            g->h[k] = len / c;
            tot *= c;
        }
        if (tot <= maxcells) break;
        h *= 1.25;
    }
}

static inline int pmod(int a, int n) {
    int r = a % n;
    return r < 0 ? r + n : r;
}

typedef struct {
    int32_t *I, *J, *img;
    int64_t n, cap;
} pairbuf;

static void pb_push(pairbuf *b, int32_t i, int32_t j, const int32_t m[3]) {
    if (b->n == b->cap) {
        b->cap = b->cap ? 2 * b->cap : 1024;
        b->I = (int32_t *)realloc(b->I, sizeof(int32_t) * b->cap);
        b->J = (int32_t *)realloc(b->J, sizeof(int32_t) * b->cap);
        b->img = (int32_t *)realloc(b->img, sizeof(int32_t) * 3 * b->cap);
    }
    b->I[b->n] = i;
    b->J[b->n] = j;
    memcpy(b->img + 3 * b->n, m, 3 * sizeof(int32_t));
    ++b->n;
}

typedef struct {
    int64_t key;
    int32_t idx;
} keyidx;
static int cmp_keyidx(const void *a, const void *b) {
    const keyidx *x = (const keyidx *)a, *y = (const keyidx *)b;
    return x->key < y->key ? -1 : (x->key > y->key ? 1 : 0);
}

int64_t orc_pairs_cells(int32_t n, const double *center, const double *dir, const double *length,
                        const double *radius, const orc_cfg *cfg, int64_t cap, int32_t *I,
                        int32_t *J, int32_t *img) {
    if (!box_ok(n, length, radius, cfg)) return -1;
    grid_t g;
    grid_setup(n, length, radius, cfg, &g);
    const int64_t ncell = (int64_t)g.nc[0] * g.nc[1] * g.nc[2];
    int32_t *clo = (int32_t *)malloc(sizeof(int32_t) * 3 * (size_t)n);
    int32_t *chi = (int32_t *)malloc(sizeof(int32_t) * 3 * (size_t)n);
    double *blo = (double *)malloc(sizeof(double) * 3 * (size_t)n);
    double *bhi = (double *)malloc(sizeof(double) * 3 * (size_t)n);
    int64_t *ecnt = (int64_t *)calloc((size_t)n + 1, sizeof(int64_t));
#pragma omp parallel for schedule(static) ORC_NT
    for (int32_t i = 0; i < n; ++i) {
        const double hh = col_half(length[i], radius[i], cfg);
        const double infl = col_rad(radius[i], cfg) + 0.5 * cfg->col_buf + ORC_AABB_MARGIN;
        int64_t cnt = 1;
        for (int k = 0; k < 3; ++k) {
            const double hw = hh * fabs(dir[3 * i + k]);
            const double lo = (center[3 * i + k] - hw) - infl;
            const double hi = (center[3 * i + k] + hw) + infl;
            int a = (int)floor((lo - cfg->box_lo[k]) / g.h[k]);
            int b = (int)floor((hi - cfg->box_lo[k]) / g.h[k]);
            if (!cfg->pbc[k]) {
                a = a < 0 ? 0 : (a >= g.nc[k] ? g.nc[k] - 1 : a);
                b = b < 0 ? 0 : (b >= g.nc[k] ? g.nc[k] - 1 : b);
            }
            clo[3 * i + k] = a;
            chi[3 * i + k] = b;
            blo[3 * i + k] = lo;
            bhi[3 * i + k] = hi;
            cnt *= (int64_t)(b - a + 1);
        }
        ecnt[i + 1] = cnt;
    }
    for (int32_t i = 0; i < n; ++i) ecnt[i + 1] += ecnt[i];
    const int64_t E = ecnt[n];
    int64_t *ekey = (int64_t *)malloc(sizeof(int64_t) * (size_t)(E ? E : 1));
    int32_t *erod = (int32_t *)malloc(sizeof(int32_t) * (size_t)(E ? E : 1));
#pragma omp parallel for schedule(static) ORC_NT
    for (int32_t i = 0; i < n; ++i) {
        int64_t w = ecnt[i];
        for (int z = clo[3 * i + 2]; z <= chi[3 * i + 2]; ++z)
            for (int y = clo[3 * i + 1]; y <= chi[3 * i + 1]; ++y)
                for (int x = clo[3 * i]; x <= chi[3 * i]; ++x) {
                    const int64_t cx = pmod(x, g.nc[0]), cy = pmod(y, g.nc[1]), cz = pmod(z, g.nc[2]);
                    ekey[w] = (cz * g.nc[1] + cy) * g.nc[0] + cx;
                    erod[w] = i;
                    ++w;
                }
    }
    /* counting sort by cell, stable in rod order */
    int64_t *cstart = (int64_t *)calloc((size_t)ncell + 1, sizeof(int64_t));
    for (int64_t e = 0; e < E; ++e) cstart[ekey[e] + 1]++;
    for (int64_t c = 0; c < ncell; ++c) cstart[c + 1] += cstart[c];
    int64_t *cur = (int64_t *)malloc(sizeof(int64_t) * (size_t)(ncell ? ncell : 1));
    memcpy(cur, cstart, sizeof(int64_t) * (size_t)ncell);
    int32_t *srt = (int32_t *)malloc(sizeof(int32_t) * (size_t)(E ? E : 1));
    for (int64_t e = 0; e < E; ++e) srt[cur[ekey[e]]++] = erod[e];
    free(cur);
    free(ekey);
    free(erod);

    const int nt = orc_get_threads();
    pairbuf *bufs = (pairbuf *)calloc((size_t)nt, sizeof(pairbuf));
#pragma omp parallel ORC_NT
    {
#ifdef _OPENMP
        pairbuf *pb = bufs + omp_get_thread_num();
#else
        pairbuf *pb = bufs;
#endif
#pragma omp for schedule(dynamic, 64)
        for (int64_t c = 0; c < ncell; ++c) {
            const int64_t s0 = cstart[c], s1 = cstart[c + 1];
            if (s1 - s0 < 2) continue;
            const int cc[3] = {(int)(c % g.nc[0]), (int)((c / g.nc[0]) % g.nc[1]),
                               (int)(c / ((int64_t)g.nc[0] * g.nc[1]))};
            for (int64_t pa = s0; pa < s1; ++pa) {
                const int32_t i = srt[pa];
                int ui[3];
                for (int k = 0; k < 3; ++k)
                    ui[k] = cfg->pbc[k] ? clo[3 * i + k] + pmod(cc[k] - clo[3 * i + k], g.nc[k]) : cc[k];
                for (int64_t pbi = pa + 1; pbi < s1; ++pbi) {
                    const int32_t j = srt[pbi];
                    int ok = 1;
                    int32_t m[3];
                    for (int k = 0; k < 3 && ok; ++k) {
                        const int uj = cfg->pbc[k] ? clo[3 * j + k] + pmod(cc[k] - clo[3 * j + k], g.nc[k])
                                                   : cc[k];
                        if (!(ui[k] == clo[3 * i + k] || uj == clo[3 * j + k])) ok = 0;
                        m[k] = cfg->pbc[k] ? (ui[k] - uj) / g.nc[k] : 0;
                        const double len = cfg->box_hi[k] - cfg->box_lo[k];
                        const double sh = cfg->pbc[k] ? (double)m[k] * len : 0.0;
                        if (!(blo[3 * j + k] + sh <= bhi[3 * i + k] && blo[3 * i + k] <= bhi[3 * j + k] + sh))
                            ok = 0;
                    }
                    if (!ok) continue;
                    if (pair_delta0(center, dir, length, radius, cfg, i, j, m) < cfg->col_buf)
                        pb_push(pb, i, j, m);
                }
            }
        }
    }
    int64_t total = 0;
    for (int t = 0; t < nt; ++t) total += bufs[t].n;
    keyidx *ki = (keyidx *)malloc(sizeof(keyidx) * (size_t)(total ? total : 1));
    int32_t *aI = (int32_t *)malloc(sizeof(int32_t) * (size_t)(total ? total : 1));
    int32_t *aJ = (int32_t *)malloc(sizeof(int32_t) * (size_t)(total ? total : 1));
    int32_t *aM = (int32_t *)malloc(sizeof(int32_t) * 3 * (size_t)(total ? total : 1));
    int64_t w = 0;
    for (int t = 0; t < nt; ++t) {
        for (int64_t p = 0; p < bufs[t].n; ++p, ++w) {
            aI[w] = bufs[t].I[p];
            aJ[w] = bufs[t].J[p];
            memcpy(aM + 3 * w, bufs[t].img + 3 * p, 3 * sizeof(int32_t));
            ki[w].key = (int64_t)aI[w] * (int64_t)n + aJ[w];
            ki[w].idx = (int32_t)w;
        }
        free(bufs[t].I);
        free(bufs[t].J);
        free(bufs[t].img);
    }
    free(bufs);
    qsort(ki, (size_t)total, sizeof(keyidx), cmp_keyidx);
    for (int64_t p = 0; p < total && p < cap; ++p) {
        const int32_t q = ki[p].idx;
        I[p] = aI[q];
        J[p] = aJ[q];
        memcpy(img + 3 * p, aM + 3 * q, 3 * sizeof(int32_t));
    }
    free(ki);
    free(aI);
    free(aJ);
    free(aM);
    free(srt);
    free(cstart);
    free(clo);
    free(chi);
    free(blo);
    free(bhi);
    free(ecnt);
    return total;
}

/* ------------------------------------------------------------------------------------------------
 * Mobility.  Per rod the 6x6 block is diag(Mt, Mr) with Mt = t t^T / zpar + (I - t t^T) / zperp,
 * Mr = I / zrot.  Slender-body drag: b = -(1 + 2 ln(r/L)), zpar = 8 pi mu L / (2b),
 * zperp = 8 pi mu L / (b + 2), zrot = 2 pi mu L^3 / (3 (b + 2)); spheres (L < 2r): 6 pi mu r and
 * 8 pi mu r^3.  Stored per rod as {t0, t1, t2, 1/zperp, 1/zpar - 1/zperp, 1/zrot}.
 * ------------------------------------------------------------------------------------------------ */
static void drag_inv(double L, double r, double mu, double *ipar, double *iperp, double *irot) {
    double zpar, zperp, zrot;
    if (L < r * 2.0) {
        zpar = 6.0 * ORC_PI * mu * r;
        zperp = zpar;
        zrot = 8.0 * ORC_PI * mu * r * r * r;
    } else {
        const double bb = -(1.0 + 2.0 * orc_log(r / L));
        zpar = 8.0 * ORC_PI * mu * L / (2.0 * bb);
        zperp = 8.0 * ORC_PI * mu * L / (bb + 2.0);
        zrot = 2.0 * ORC_PI * mu * L * L * L / (3.0 * (bb + 2.0));
    }
    *ipar = 1.0 / zpar;
    *iperp = 1.0 / zperp;
    *irot = 1.0 / zrot;
}

void orc_mobility(int32_t n, const double *dir, const double *length, const double *radius,
                  const uint8_t *immovable, const orc_cfg *cfg, double *mob) {
#pragma omp parallel for schedule(static) ORC_NT
    for (int32_t i = 0; i < n; ++i) {
        double ipar, iperp, irot;
        drag_inv(length[i], radius[i], cfg->viscosity, &ipar, &iperp, &irot);
        double *m = mob + 6 * i;
        m[0] = dir[3 * i];
        m[1] = dir[3 * i + 1];
        m[2] = dir[3 * i + 2];
        if (immovable && immovable[i]) {
            m[3] = 0.0;
            m[4] = 0.0;
            m[5] = 0.0;
        } else {
            m[3] = iperp;
            m[4] = ipar - iperp;
            m[5] = irot;
        }
    }
}

static inline void apply_M(const double *mb, const double *F, int mono, double *U) {
    double f0 = F[0], f1 = F[1], f2 = F[2], q0 = F[3], q1 = F[4], q2 = F[5];
    if (mono) {
        f2 = 0.0;
        q0 = 0.0;
        q1 = 0.0;
    }
    const double tf = mb[0] * f0 + mb[1] * f1 + mb[2] * f2;
    const double c = mb[4] * tf;
    U[0] = mb[3] * f0 + c * mb[0];
    U[1] = mb[3] * f1 + c * mb[1];
    U[2] = mb[3] * f2 + c * mb[2];
    U[3] = mb[5] * q0;
    U[4] = mb[5] * q1;
    U[5] = mb[5] * q2;
    if (mono) {
        U[2] = 0.0;
        U[3] = 0.0;
        U[4] = 0.0;
    }
}

/* U_nc = M F_ext + U_B, U_B = sqrt(2 kT / dt) M^{1/2} xi, xi ~ N(0, I_6) from Philox4x32-10 keyed by
 * (seed) with counter (gid, step_lo, step_hi, c), c = 0, 1, 2, and Box-Muller on 53-bit uniforms. */
static void brownian6(int32_t gid, const orc_cfg *cfg, double z[6]) {
    const uint32_t key[2] = {(uint32_t)cfg->seed, (uint32_t)(cfg->seed >> 32)};
    for (uint32_t c = 0; c < 3; ++c) {
        const uint32_t ctr[4] = {(uint32_t)gid, (uint32_t)cfg->step, (uint32_t)((uint64_t)cfg->step >> 32), c};
        uint32_t o[4];
        orc_philox4x32_10(ctr, key, o);
        const double u1 = ((double)((((uint64_t)o[0]) << 21) ^ (uint64_t)(o[1] >> 11)) + 1.0) * 0x1p-53;
        const double u2 = (double)((((uint64_t)o[2]) << 21) ^ (uint64_t)(o[3] >> 11)) * 0x1p-53;
        const double R = sqrt(-2.0 * orc_log(u1));
        double sn, cs;
        orc_sincos2pi(u2, &sn, &cs);
        z[2 * c] = R * cs;
        z[2 * c + 1] = R * sn;
    }
}

void orc_unc(int32_t n, const int32_t *gid, const double *mob, const double *f_ext,
             const orc_cfg *cfg, double *unc) {
    const int mono = cfg->monolayer;
    const double pre = cfg->kT > 0.0 ? sqrt(2.0 * cfg->kT / cfg->dt) : 0.0;
#pragma omp parallel for schedule(static) ORC_NT
    for (int32_t i = 0; i < n; ++i) {
        const double *mb = mob + 6 * i;
        double ub[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        if (cfg->kT > 0.0) {
            double z[6];
            brownian6(gid[i], cfg, z);
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
        double *u = unc + 6 * i;
        if (f_ext) {
            double ue[6];
            apply_M(mb, f_ext + 6 * i, mono, ue);
            for (int k = 0; k < 6; ++k) u[k] = ue[k] + ub[k];
        } else {
            for (int k = 0; k < 6; ++k) u[k] = ub[k];
        }
        if (mono) {
            u[2] = 0.0;
            u[3] = 0.0;
            u[4] = 0.0;
        }
    }
}

/* ------------------------------------------------------------------------------------------------
 * Constraint rows (one per recorded pair, canonical order = ascending (I, J)):
 *   delta0 = |P - Q| - rI - rJ, n = (P - Q)/|P - Q| (from J to I), posI = P - cI, posJ = Q - cJ',
 *   labI = P, labJ = Q, rI = posI x n, rJ = posJ x n, b = delta0 + dt * (D^T U_nc).
 * D column k: rod I gets (n, rI), rod J gets (-n, -rJ); (D^T U)_k = (n.vI + rI.wI) - (n.vJ + rJ.wJ).
 * ------------------------------------------------------------------------------------------------ */
static inline double dot3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

static inline double dt_row(const double *nk, const double *rik, const double *rjk, const double *uI,
                            const double *uJ) {
    return (dot3(nk, uI) + dot3(rik, uI + 3)) - (dot3(nk, uJ) + dot3(rjk, uJ + 3));
}

void orc_rows(int32_t n, const double *center, const double *dir, const double *length,
              const double *radius, const orc_cfg *cfg, int64_t m, const int32_t *I,
              const int32_t *J, const int32_t *img, const double *unc, double *delta0, double *nrm,
              double *posI, double *posJ, double *rI, double *rJ, double *b) {
    (void)n;
#pragma omp parallel for schedule(static) ORC_NT
    for (int64_t k = 0; k < m; ++k) {
        const int32_t i = I[k], j = J[k];
        double cJ[3], o[6];
        shifted_centre(center + 3 * j, img + 3 * k, cfg, cJ);
        const double *tI = dir + 3 * i, *tJ = dir + 3 * j;
        orc_segseg(center + 3 * i, tI, col_half(length[i], radius[i], cfg), cJ, tJ,
                   col_half(length[j], radius[j], cfg), o);
        const double s = o[0], u = o[1], dist = o[2];
        const double D[3] = {o[3], o[4], o[5]};
        double nk[3];
        contact_normal(D, dist, tI, tJ, nk);
        const double pI[3] = {s * tI[0], s * tI[1], s * tI[2]};
        const double pJ[3] = {u * tJ[0], u * tJ[1], u * tJ[2]};
        const double ri[3] = {pI[1] * nk[2] - pI[2] * nk[1], pI[2] * nk[0] - pI[0] * nk[2],
                              pI[0] * nk[1] - pI[1] * nk[0]};
        const double rj[3] = {pJ[1] * nk[2] - pJ[2] * nk[1], pJ[2] * nk[0] - pJ[0] * nk[2],
                              pJ[0] * nk[1] - pJ[1] * nk[0]};
        const double d0 = (dist - col_rad(radius[i], cfg)) - col_rad(radius[j], cfg);
        delta0[k] = d0;
        for (int a = 0; a < 3; ++a) {
            nrm[3 * k + a] = nk[a];
            posI[3 * k + a] = pI[a];
            posJ[3 * k + a] = pJ[a];
            rI[3 * k + a] = ri[a];
            rJ[3 * k + a] = rj[a];
        }
        b[k] = d0 + cfg->dt * dt_row(nk, ri, rj, unc + 6 * i, unc + 6 * j);
    }
}

/* Rows supplied by the caller (the reference's constraint pool, SRC/TubuleSystem.cpp:694-745): the
 * normal n = normI (normJ = -normI), arms posI / posJ relative to the rod centres, separation
 * delta0.  rI = posI x n, rJ = posJ x n and b = delta0 + dt D^T U_nc as for collision rows. */
void orc_rows_ext(int64_t m, const int32_t *I, const int32_t *J, const double *delta0,
                  const double *nrm, const double *posI, const double *posJ, const double *unc,
                  const orc_cfg *cfg, double *rI, double *rJ, double *b) {
    for (int64_t k = 0; k < m; ++k) {
        const double *nk = nrm + 3 * k, *pI = posI + 3 * k, *pJ = posJ + 3 * k;
        const double ri[3] = {pI[1] * nk[2] - pI[2] * nk[1], pI[2] * nk[0] - pI[0] * nk[2],
                              pI[0] * nk[1] - pI[1] * nk[0]};
        const double rj[3] = {pJ[1] * nk[2] - pJ[2] * nk[1], pJ[2] * nk[0] - pJ[0] * nk[2],
                              pJ[0] * nk[1] - pJ[1] * nk[0]};
        for (int a = 0; a < 3; ++a) {
            rI[3 * k + a] = ri[a];
            rJ[3 * k + a] = rj[a];
        }
        b[k] = delta0[k] + cfg->dt * dt_row(nk, ri, rj, unc + 6 * (size_t)I[k], unc + 6 * (size_t)J[k]);
    }
}

/* Sum over rows, in row order, of gamma_k * S_k (row-major 3 x 3), where S_k is the per-unit-impulse
 * stress of row k: (labI - labJ) (x) normI for collision rows (the virial of the unit force pair
 * +n at labI, -n at labJ), or the caller's matrix where one is given (stress != NULL and
 * use_given[k] != 0). */
void orc_stress_sum(int64_t m, const double *nrm, const double *labI, const double *labJ,
                    const double *gamma, const double *stress, const uint8_t *use_given, double out[9]) {
    for (int q = 0; q < 9; ++q) out[q] = 0.0;
    for (int64_t k = 0; k < m; ++k) {
        double S[9];
        if (stress && use_given && use_given[k]) {
            for (int q = 0; q < 9; ++q) S[q] = stress[9 * k + q];
        } else {
            for (int a = 0; a < 3; ++a)
                for (int c = 0; c < 3; ++c) S[3 * a + c] = (labI[3 * k + a] - labJ[3 * k + a]) * nrm[3 * k + c];
        }
        for (int q = 0; q < 9; ++q) out[q] = out[q] + gamma[k] * S[q];
    }
}

/* ------------------------------------------------------------------------------------------------
 * Operator pieces.  D is applied rod-major: rod i sums its incident constraints in ascending k.
 * ------------------------------------------------------------------------------------------------ */
typedef struct {
    int64_t *start; /* n + 1 */
    int64_t *ent;   /* 2m: k * 2 + side (0 = I, 1 = J) */
} adj_t;

static void adj_build(int32_t n, int64_t m, const int32_t *I, const int32_t *J, adj_t *a) {
    a->start = (int64_t *)calloc((size_t)n + 1, sizeof(int64_t));
    a->ent = (int64_t *)malloc(sizeof(int64_t) * (size_t)(2 * m + 1));
    for (int64_t k = 0; k < m; ++k) {
        a->start[I[k] + 1]++;
        a->start[J[k] + 1]++;
    }
    for (int32_t i = 0; i < n; ++i) a->start[i + 1] += a->start[i];
    int64_t *cur = (int64_t *)malloc(sizeof(int64_t) * ((size_t)n + 1));
    memcpy(cur, a->start, sizeof(int64_t) * ((size_t)n + 1));
    for (int64_t k = 0; k < m; ++k) {
        a->ent[cur[I[k]]++] = 2 * k;
        a->ent[cur[J[k]]++] = 2 * k + 1;
    }
    free(cur);
}
static void adj_free(adj_t *a) {
    free(a->start);
    free(a->ent);
}

static void apply_DM(int32_t n, const adj_t *adj, const double *nrm, const double *rI,
                     const double *rJ, const double *mob, int mono, const double *v, double *F,
                     double *U) {
#pragma omp parallel for schedule(static) ORC_NT
    for (int32_t i = 0; i < n; ++i) {
        double f[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        for (int64_t e = adj->start[i]; e < adj->start[i + 1]; ++e) {
            const int64_t k = adj->ent[e] >> 1;
            const double g = v[k];
            if ((adj->ent[e] & 1) == 0) {
                f[0] = f[0] + g * nrm[3 * k];
                f[1] = f[1] + g * nrm[3 * k + 1];
                f[2] = f[2] + g * nrm[3 * k + 2];
                f[3] = f[3] + g * rI[3 * k];
                f[4] = f[4] + g * rI[3 * k + 1];
                f[5] = f[5] + g * rI[3 * k + 2];
            } else {
                f[0] = f[0] + g * -nrm[3 * k];
                f[1] = f[1] + g * -nrm[3 * k + 1];
                f[2] = f[2] + g * -nrm[3 * k + 2];
                f[3] = f[3] + g * -rJ[3 * k];
                f[4] = f[4] + g * -rJ[3 * k + 1];
                f[5] = f[5] + g * -rJ[3 * k + 2];
            }
        }
        if (F)
            for (int a = 0; a < 6; ++a) F[6 * i + a] = f[a];
        apply_M(mob + 6 * i, f, mono, U + 6 * i);
    }
}

static void apply_DT(int64_t m, const int32_t *I, const int32_t *J, const double *nrm,
                     const double *rI, const double *rJ, const double *U, double *du) {
#pragma omp parallel for schedule(static) ORC_NT
    for (int64_t k = 0; k < m; ++k)
        du[k] = dt_row(nrm + 3 * k, rI + 3 * k, rJ + 3 * k, U + 6 * I[k], U + 6 * J[k]);
}

void orc_apply_D_M(int32_t n, int64_t m, const int32_t *I, const int32_t *J, const double *nrm,
                   const double *rI, const double *rJ, const double *mob, const orc_cfg *cfg,
                   const double *gamma, double *F, double *U) {
    adj_t a;
    adj_build(n, m, I, J, &a);
    apply_DM(n, &a, nrm, rI, rJ, mob, cfg->monolayer, gamma, F, U);
    adj_free(&a);
}

void orc_apply_DT(int64_t m, const int32_t *I, const int32_t *J, const double *nrm,
                  const double *rI, const double *rJ, const double *U, double *du) {
    apply_DT(m, I, J, nrm, rI, rJ, U, du);
}

/* ------------------------------------------------------------------------------------------------
 * BBPGD for  min 1/2 x^T A x + b^T x,  x >= lb,  A = dt D^T M D  (Barzilai & Borwein, IMA J. Numer.
 * Anal. 8 (1988); projected form with alternating BB1/BB2 steps).
 *   P(v) = (v < lb) ? lb : v ;  res = max_k |x_k - P(x_k - g_k)|
 *   x = P(x0); g = A x + b; stop if res <= tol
 *   alpha = g^T g / g^T A g   (1.0 if that is not a positive finite number)
 *   k = 1, 2, ...: x+ = P(x - alpha g); g+ = A x+ + b; stop if res(x+, g+) <= tol;
 *                  s = x+ - x, y = g+ - g; alpha = (k odd) ? s^T s / s^T y : s^T y / y^T y,
 *                  kept from the previous iteration unless positive and finite.
 * Warm start: if res(x0) > res(0) the solve restarts from x = 0, g = b (fallback_used = 1).
 * Sums use orc_chunked_sum's order; the max is order-free.
 * ------------------------------------------------------------------------------------------------ */
static inline double proj(double v, double lb) { return v < lb ? lb : v; }

typedef struct {
    double sum[3];
    double rmax;
    int64_t nonfinite;
} red_t;

/* Sums of three per-element products in the fixed chunk order, plus max and non-finite count. */
static void reduce3(int64_t m, const double *a0, const double *b0, const double *a1,
                    const double *b1, const double *a2, const double *b2, const double *res,
                    red_t *out) {
    const int64_t nc = (m + ORC_RED_CHUNK - 1) / ORC_RED_CHUNK;
    double *part = (double *)malloc(sizeof(double) * 3 * (size_t)(nc ? nc : 1));
    double rmax = 0.0;
    int64_t nf = 0;
#pragma omp parallel for schedule(static) reduction(max : rmax) reduction(+ : nf) ORC_NT
    for (int64_t c = 0; c < nc; ++c) {
        double tmp[ORC_RED_CHUNK];
        int64_t len = m - c * ORC_RED_CHUNK;
        if (len > ORC_RED_CHUNK) len = ORC_RED_CHUNK;
        const int64_t o = c * ORC_RED_CHUNK;
        const double *as[3] = {a0, a1, a2}, *bs[3] = {b0, b1, b2};
        for (int q = 0; q < 3; ++q) {
            if (!as[q]) {
                part[3 * c + q] = 0.0;
                continue;
            }
            for (int64_t e = 0; e < len; ++e) tmp[e] = as[q][o + e] * bs[q][o + e];
            part[3 * c + q] = orc_block_sum(tmp, len);
        }
        for (int64_t e = 0; e < len; ++e) {
            const double r = res[o + e];
            if (!isfinite(r)) nf++;
            else if (r > rmax) rmax = r;
        }
    }
    double *col = (double *)malloc(sizeof(double) * (size_t)(nc ? nc : 1));
    for (int q = 0; q < 3; ++q) {
        for (int64_t c = 0; c < nc; ++c) col[c] = part[3 * c + q];
        out->sum[q] = orc_block_sum(col, nc);
    }
    out->rmax = rmax;
    out->nonfinite = nf;
    free(col);
    free(part);
}

int orc_bbpgd(int32_t n, int64_t m, const int32_t *I, const int32_t *J, const double *nrm,
              const double *rI, const double *rJ, const double *b, const double *lb,
              const double *invk, const double *mob, const orc_cfg *cfg, const double *x0,
              double *gamma, orc_report *rep) {
    memset(rep, 0, sizeof(*rep));
    if (m == 0) {
        rep->converged = 1;
        return 0;
    }
    const double dt = cfg->dt, tol = cfg->tol;
    const int mono = cfg->monolayer;
    adj_t adj;
    adj_build(n, m, I, J, &adj);
    double *x = (double *)malloc(sizeof(double) * m);
    double *g = (double *)malloc(sizeof(double) * m);
    double *xn = (double *)malloc(sizeof(double) * m);
    double *gn = (double *)malloc(sizeof(double) * m);
    double *du = (double *)malloc(sizeof(double) * m);
    double *res = (double *)malloc(sizeof(double) * m);
    double *sv = (double *)malloc(sizeof(double) * m);
    double *yv = (double *)malloc(sizeof(double) * m);
    double *U = (double *)malloc(sizeof(double) * 6 * (size_t)n);
    red_t rd;
    int status = 0;

    /* start point */
#pragma omp parallel for schedule(static) ORC_NT
    for (int64_t k = 0; k < m; ++k) xn[k] = proj(x0 ? x0[k] : 0.0, lb ? lb[k] : 0.0);
    apply_DM(n, &adj, nrm, rI, rJ, mob, mono, xn, NULL, U);
    apply_DT(m, I, J, nrm, rI, rJ, U, du);
    double *rcold = x0 ? (double *)malloc(sizeof(double) * m) : NULL;
#pragma omp parallel for schedule(static) ORC_NT
    for (int64_t k = 0; k < m; ++k) {
        const double l = lb ? lb[k] : 0.0;
        x[k] = xn[k];
        g[k] = dt * du[k] + b[k];
        if (invk) g[k] = g[k] + invk[k] * xn[k];
        res[k] = fabs(xn[k] - proj(xn[k] - g[k], l));
        if (rcold) rcold[k] = fabs(0.0 - proj(0.0 - b[k], l));
    }
    reduce3(m, NULL, NULL, NULL, NULL, NULL, NULL, res, &rd);
    double R0 = rd.rmax;
    int64_t nf0 = rd.nonfinite;
    rep->matvecs = 1;
    if (rcold) {
        red_t rc;
        reduce3(m, NULL, NULL, NULL, NULL, NULL, NULL, rcold, &rc);
        if (nf0 > 0 || rd.rmax > rc.rmax) {
            for (int64_t k = 0; k < m; ++k) {
                x[k] = 0.0;
                g[k] = 0.0 + b[k];
            }
            R0 = rc.rmax;
            nf0 = rc.nonfinite;
            rep->fallback_used = 1;
        }
        free(rcold);
    }
    rep->res0 = R0;
    rep->res = R0;
    if (nf0 > 0) {
        status = 4;
        goto done;
    }
    if (R0 <= tol) {
        rep->converged = 1;
        goto done;
    }
    /* initial step length from the Rayleigh quotient of the gradient */
    apply_DM(n, &adj, nrm, rI, rJ, mob, mono, g, NULL, U);
    apply_DT(m, I, J, nrm, rI, rJ, U, du);
    rep->matvecs++;
#pragma omp parallel for schedule(static) ORC_NT
    for (int64_t k = 0; k < m; ++k) {
        du[k] = dt * du[k];
        if (invk) du[k] = du[k] + invk[k] * g[k];
        res[k] = 0.0;
    }
    reduce3(m, g, g, g, du, NULL, NULL, res, &rd);
    double alpha = rd.sum[0] / rd.sum[1];
    if (!(alpha > 0.0 && alpha < INFINITY)) alpha = 1.0;

    for (int it = 1; it <= cfg->max_ite; ++it) {
#pragma omp parallel for schedule(static) ORC_NT
        for (int64_t k = 0; k < m; ++k) xn[k] = proj(x[k] - alpha * g[k], lb ? lb[k] : 0.0);
        apply_DM(n, &adj, nrm, rI, rJ, mob, mono, xn, NULL, U);
        apply_DT(m, I, J, nrm, rI, rJ, U, du);
        rep->matvecs++;
#pragma omp parallel for schedule(static) ORC_NT
        for (int64_t k = 0; k < m; ++k) {
            const double l = lb ? lb[k] : 0.0;
            gn[k] = dt * du[k] + b[k];
            if (invk) gn[k] = gn[k] + invk[k] * xn[k];
            res[k] = fabs(xn[k] - proj(xn[k] - gn[k], l));
            sv[k] = xn[k] - x[k];
            yv[k] = gn[k] - g[k];
            x[k] = xn[k];
            g[k] = gn[k];
        }
        reduce3(m, sv, sv, sv, yv, yv, yv, res, &rd);
        rep->iters = it;
        rep->res = rd.rmax;
        if (rd.nonfinite > 0) {
            status = 4;
            goto done;
        }
        if (rd.rmax <= tol) {
            rep->converged = 1;
            goto done;
        }
        const double an = (it & 1) ? rd.sum[0] / rd.sum[1] : rd.sum[1] / rd.sum[2];
        if (an > 0.0 && an < INFINITY) alpha = an;
    }
    status = 5;
done:
    memcpy(gamma, x, sizeof(double) * m);
    rep->status = status;
    free(x);
    free(g);
    free(xn);
    free(gn);
    free(du);
    free(res);
    free(sv);
    free(yv);
    free(U);
    adj_free(&adj);
    return status;
}

void orc_writeback(int32_t n, int64_t m, const int32_t *I, const int32_t *J, const double *nrm,
                   const double *rI, const double *rJ, const double *gamma, const double *mob,
                   const double *unc, const orc_cfg *cfg, double *fcon, double *u) {
    adj_t a;
    adj_build(n, m, I, J, &a);
    double *U = (double *)malloc(sizeof(double) * 6 * (size_t)n);
    apply_DM(n, &a, nrm, rI, rJ, mob, cfg->monolayer, gamma, fcon, U);
    for (int64_t q = 0; q < 6 * (int64_t)n; ++q) u[q] = unc[q] + U[q];
    free(U);
    adj_free(&a);
}

/* ------------------------------------------------------------------------------------------------
 * Explicit Euler step: x += v dt; direction rotated by the unit quaternion normalise(1, w dt / 2)
 * and renormalised; periodic coordinates wrapped into [lo, hi).
 * ------------------------------------------------------------------------------------------------ */
void orc_euler(int32_t n, double *center, double *dir, const double *u, const orc_cfg *cfg) {
    const double dt = cfg->dt;
    const double hdt = 0.5 * dt;
#pragma omp parallel for schedule(static) ORC_NT
    for (int32_t i = 0; i < n; ++i) {
        const double *ui = u + 6 * i;
        double *c = center + 3 * i, *t = dir + 3 * i;
        for (int k = 0; k < 3; ++k) c[k] = c[k] + ui[k] * dt;
        double q1 = hdt * ui[3], q2 = hdt * ui[4], q3 = hdt * ui[5];
        const double qn = sqrt(1.0 + (q1 * q1 + q2 * q2 + q3 * q3));
        const double q0 = 1.0 / qn;
        q1 = q1 / qn;
        q2 = q2 / qn;
        q3 = q3 / qn;
        const double a0 = q2 * t[2] - q3 * t[1], a1 = q3 * t[0] - q1 * t[2], a2 = q1 * t[1] - q2 * t[0];
        const double b0 = q2 * a2 - q3 * a1, b1 = q3 * a0 - q1 * a2, b2 = q1 * a1 - q2 * a0;
        const double tq = 2.0 * q0;
        const double n0 = (t[0] + tq * a0) + 2.0 * b0;
        const double n1 = (t[1] + tq * a1) + 2.0 * b1;
        const double n2 = (t[2] + tq * a2) + 2.0 * b2;
        const double nn = sqrt(n0 * n0 + n1 * n1 + n2 * n2);
        t[0] = n0 / nn;
        t[1] = n1 / nn;
        t[2] = n2 / nn;
        for (int k = 0; k < 3; ++k) {
            if (!cfg->pbc[k]) continue;
            const double len = cfg->box_hi[k] - cfg->box_lo[k];
            c[k] = c[k] - floor((c[k] - cfg->box_lo[k]) / len) * len;
        }
    }
}
