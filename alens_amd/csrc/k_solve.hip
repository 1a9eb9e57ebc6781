// k_solve.hip -- BBPGD solve of the collision BCQP as one persistent kernel per solve, the
// write-back of the contact impulses, the Euler step and the warm-start bookkeeping (gfx950).
// Reference roles: BCQPSolver::solveBBPGD, ConstraintSolver::writebackGamma and
// SylinderSystem::stepEuler (SimToolbox, CMakeLists.txt:109-115; called through
// rodSystem.runStep() / stepEuler(), SRC/TubuleSystem.cpp:223, :55).
//
// One iteration = phase A (rod-parallel: x+ = P(x - alpha g) gathered per incident constraint,
// F = D x+, U = M F), grid barrier, phase B (constraint-parallel in 2048-element chunks:
// g+ = dt D^T U + b, residual, s = x+ - x, y = g+ - g, chunk partial sums), grid barrier, and every
// block folds the chunk partials in the same fixed order, so all blocks hold identical step lengths
// and stop together.
#include "alens_dev.hpp"
#include "alens_internal.hpp"

namespace alens {

#define LAUNCH_CHECK()                                                                           \
    do {                                                                                         \
        hipError_t e_ = hipGetLastError();                                                       \
        if (e_ != hipSuccess) return HipErr{e_, "kernel launch", __FILE__, __LINE__};             \
    } while (0)

static inline unsigned nblk(int64_t n, int t) { return (unsigned)((n + t - 1) / t); }

constexpr uint32_t kSpinLimit = 1u << 26;  // barrier wait bound: sets the abort word, never hangs

enum : int { Q_SS = 0, Q_SY = 1, Q_YY = 2, Q_RMAX = 3, Q_NF = 4, Q_RC = 5, Q_NFC = 6 };

// Grid barrier over co-resident blocks (cooperative launch).  bar[0] = arrivals, bar[1] = generation,
// bar[2] = abort.  Returns false once any block gave up waiting.
__device__ inline bool grid_barrier(uint32_t *bar, int *s_flag) {
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t gen = __hip_atomic_load(&bar[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __threadfence();
        const uint32_t prev = __hip_atomic_fetch_add(&bar[0], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        if (prev == gridDim.x - 1) {
            __hip_atomic_store(&bar[0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&bar[1], gen + 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            uint32_t spins = 0;
            while (__hip_atomic_load(&bar[1], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == gen) {
                __builtin_amdgcn_s_sleep(1);
                if (++spins > kSpinLimit) {
                    __hip_atomic_store(&bar[2], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    break;
                }
                if ((spins & 1023u) == 0 &&
                    __hip_atomic_load(&bar[2], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
                    break;
            }
        }
        __threadfence();
        *s_flag = (int)__hip_atomic_load(&bar[2], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
    return *s_flag == 0;
}

// Sum in the fixed order: per-thread partial already accumulated left to right; fold each wave by
// halving (lane l += lane l + h, h = 32 .. 1); add the four wave totals left to right.
__device__ inline double block_sum_fixed(double v, double *sh) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int h = 32; h >= 1; h >>= 1) v = v + __shfl_down(v, h, 64);
    if (lane == 0) sh[w] = v;
    __syncthreads();
    const double t = ((sh[0] + sh[1]) + sh[2]) + sh[3];
    __syncthreads();
    return t;
}

__device__ inline double block_max(double v, double *sh) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int h = 32; h >= 1; h >>= 1) {
        const double o = __shfl_down(v, h, 64);
        v = o > v ? o : v;
    }
    if (lane == 0) sh[w] = v;
    __syncthreads();
    double t = sh[0];
    t = sh[1] > t ? sh[1] : t;
    t = sh[2] > t ? sh[2] : t;
    t = sh[3] > t ? sh[3] : t;
    __syncthreads();
    return t;
}

// total of partial slot q over all chunks, fixed order
__device__ inline double fold_sum(const double *part, int nchunks, int q, double *sh) {
    double acc = 0.0;
    for (int c = threadIdx.x; c < nchunks; c += kRedThreads) acc = acc + part[(size_t)q * nchunks + c];
    return block_sum_fixed(acc, sh);
}
__device__ inline double fold_max(const double *part, int nchunks, int q, double *sh) {
    double acc = 0.0;
    for (int c = threadIdx.x; c < nchunks; c += kRedThreads) {
        const double v = part[(size_t)q * nchunks + c];
        acc = v > acc ? v : acc;
    }
    return block_max(acc, sh);
}

enum : int { A_PROJ = 0, A_GRAD = 1 };

template <int MODE>
__device__ inline void phase_A(const SolveArgs &a, double alpha) {
    const int stride = gridDim.x * blockDim.x;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < a.n; i += stride) {
        double f[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        const uint32_t e1 = a.adjStart[i + 1];
        for (uint32_t e = a.adjStart[i]; e < e1; ++e) {
            const uint32_t k = a.adjK[e] & 0x7fffffffu;
            const double2 xg = a.xg[k];
            double v;
            if (MODE == A_GRAD) {
                v = xg.y;
            } else {
                v = proj(xg.x - alpha * xg.y, a.lb ? a.lb[k] : 0.0);
            }
            const double2 *w2 = reinterpret_cast<const double2 *>(a.adjW + 6 * (size_t)e);
            const double2 w01 = w2[0], w23 = w2[1], w45 = w2[2];
            f[0] = f[0] + v * w01.x;
            f[1] = f[1] + v * w01.y;
            f[2] = f[2] + v * w23.x;
            f[3] = f[3] + v * w23.y;
            f[4] = f[4] + v * w45.x;
            f[5] = f[5] + v * w45.y;
        }
        double mb[6], u[6];
#pragma unroll
        for (int q = 0; q < 6; ++q) mb[q] = a.mob[6 * (size_t)i + q];
        apply_M(mb, f, a.mono, u);
        double2 *U2 = reinterpret_cast<double2 *>(a.U + 6 * (size_t)i);
        U2[0] = make_double2(u[0], u[1]);
        U2[1] = make_double2(u[2], u[3]);
        U2[2] = make_double2(u[4], u[5]);
    }
}

__device__ inline double du_row(const SolveArgs &a, uint32_t k) {
    double gk[9], uI[6], uJ[6];
#pragma unroll
    for (int q = 0; q < 9; ++q) gk[q] = a.geo[9 * (size_t)k + q];
    const size_t I = (size_t)a.cI[k], J = (size_t)a.cJ[k];
    const double2 *UI = reinterpret_cast<const double2 *>(a.U + 6 * I);
    const double2 *UJ = reinterpret_cast<const double2 *>(a.U + 6 * J);
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        const double2 p = UI[q], r = UJ[q];
        uI[2 * q] = p.x;
        uI[2 * q + 1] = p.y;
        uJ[2 * q] = r.x;
        uJ[2 * q + 1] = r.y;
    }
    return dt_row(gk, uI, uJ);
}

enum : int { B_INIT = 0, B_ALPHA0 = 1, B_ITER = 2 };

template <int MODE>
__device__ inline void phase_B(const SolveArgs &a, double alpha, double *sh) {
    for (int c = blockIdx.x; c < a.nchunks; c += gridDim.x) {
        double s0 = 0.0, s1 = 0.0, s2 = 0.0, rmax = 0.0, nf = 0.0, rc = 0.0, nfc = 0.0;
        for (int p = 0; p < kRedPasses; ++p) {
            const uint32_t k = (uint32_t)c * kRedChunk + p * kRedThreads + threadIdx.x;
            if (k >= a.m) break;
            const double2 xg = a.xg[k];
            const double l = a.lb ? a.lb[k] : 0.0;
            const double du = du_row(a, k);
            if (MODE == B_ALPHA0) {
                const double ag = a.dt * du;
                s0 = s0 + xg.y * xg.y;
                s1 = s1 + xg.y * ag;
            } else {
                const double xn = proj(xg.x - alpha * xg.y, l);
                const double bk = a.b[k];
                const double gn = a.dt * du + bk;
                const double res = fabs(xn - proj(xn - gn, l));
                if (!isfinite(res))
                    nf = nf + 1.0;
                else if (res > rmax)
                    rmax = res;
                if (MODE == B_INIT) {
                    if (a.warm) {
                        const double r0 = fabs(0.0 - proj(0.0 - bk, l));
                        if (!isfinite(r0))
                            nfc = nfc + 1.0;
                        else if (r0 > rc)
                            rc = r0;
                    }
                } else {
                    const double s = xn - xg.x;
                    const double y = gn - xg.y;
                    s0 = s0 + s * s;
                    s1 = s1 + s * y;
                    s2 = s2 + y * y;
                }
                a.xg[k] = make_double2(xn, gn);
            }
        }
        double *pc = a.part + c;
        const int nc = a.nchunks;
        if (MODE != B_INIT) {
            const double t0 = block_sum_fixed(s0, sh);
            const double t1 = block_sum_fixed(s1, sh);
            if (threadIdx.x == 0) {
                pc[(size_t)Q_SS * nc] = t0;
                pc[(size_t)Q_SY * nc] = t1;
            }
        }
        if (MODE == B_ITER) {
            const double t2 = block_sum_fixed(s2, sh);
            if (threadIdx.x == 0) pc[(size_t)Q_YY * nc] = t2;
        }
        if (MODE != B_ALPHA0) {
            const double tm = block_max(rmax, sh);
            const double tn = block_sum_fixed(nf, sh);
            if (threadIdx.x == 0) {
                pc[(size_t)Q_RMAX * nc] = tm;
                pc[(size_t)Q_NF * nc] = tn;
            }
        }
        if (MODE == B_INIT && a.warm) {
            const double tm = block_max(rc, sh);
            const double tn = block_sum_fixed(nfc, sh);
            if (threadIdx.x == 0) {
                pc[(size_t)Q_RC * nc] = tm;
                pc[(size_t)Q_NFC * nc] = tn;
            }
        }
    }
}

// out: [0] iters, [1] matvecs, [2] converged, [3] fallback, [4] status, [5] res0, [6] res,
//      [7] alpha at exit
__global__ __launch_bounds__(256) void k_bbpgd64(SolveArgs a) {
    __shared__ double sh[8];
    __shared__ int s_flag;
    double alpha = 0.0;
    int iters = 0, matvecs = 1, converged = 0, fallback = 0, status = 0;
    double res0 = 0.0, res = 0.0;

    // start point: x = P(x0), g = A x + b
    phase_A<A_PROJ>(a, 0.0);
    if (!grid_barrier(a.bar, &s_flag)) return;
    phase_B<B_INIT>(a, 0.0, sh);
    if (!grid_barrier(a.bar, &s_flag)) return;
    double R0 = fold_max(a.part, a.nchunks, Q_RMAX, sh);
    double NF0 = fold_sum(a.part, a.nchunks, Q_NF, sh);
    if (a.warm) {
        const double Rc = fold_max(a.part, a.nchunks, Q_RC, sh);
        const double NFc = fold_sum(a.part, a.nchunks, Q_NFC, sh);
        if (NF0 > 0.0 || R0 > Rc) {
            // warm start rejected: x = 0, g = 0 + b
            for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < a.m; k += gridDim.x * blockDim.x)
                a.xg[k] = make_double2(0.0, 0.0 + a.b[k]);
            R0 = Rc;
            NF0 = NFc;
            fallback = 1;
            if (!grid_barrier(a.bar, &s_flag)) return;
        }
    }
    res0 = R0;
    res = R0;
    if (NF0 > 0.0) {
        status = ALENS_E_NONFINITE;
    } else if (R0 <= a.tol) {
        converged = 1;
    } else {
        // alpha0 = g^T g / g^T A g
        phase_A<A_GRAD>(a, 0.0);
        if (!grid_barrier(a.bar, &s_flag)) return;
        phase_B<B_ALPHA0>(a, 0.0, sh);
        if (!grid_barrier(a.bar, &s_flag)) return;
        ++matvecs;
        const double gg = fold_sum(a.part, a.nchunks, Q_SS, sh);
        const double gag = fold_sum(a.part, a.nchunks, Q_SY, sh);
        alpha = gg / gag;
        if (!(alpha > 0.0 && alpha < INFINITY)) alpha = 1.0;
        status = ALENS_E_NOCONV;
        for (int it = 1; it <= a.max_ite; ++it) {
            phase_A<A_PROJ>(a, alpha);
            if (!grid_barrier(a.bar, &s_flag)) return;
            phase_B<B_ITER>(a, alpha, sh);
            if (!grid_barrier(a.bar, &s_flag)) return;
            ++matvecs;
            const double ss = fold_sum(a.part, a.nchunks, Q_SS, sh);
            const double sy = fold_sum(a.part, a.nchunks, Q_SY, sh);
            const double yy = fold_sum(a.part, a.nchunks, Q_YY, sh);
            const double R = fold_max(a.part, a.nchunks, Q_RMAX, sh);
            const double NF = fold_sum(a.part, a.nchunks, Q_NF, sh);
            iters = it;
            res = R;
            if (NF > 0.0) {
                status = ALENS_E_NONFINITE;
                break;
            }
            if (R <= a.tol) {
                converged = 1;
                status = 0;
                break;
            }
            const double an = (it & 1) ? ss / sy : sy / yy;
            if (an > 0.0 && an < INFINITY) alpha = an;
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        a.out[0] = (double)iters;
        a.out[1] = (double)matvecs;
        a.out[2] = (double)converged;
        a.out[3] = (double)fallback;
        a.out[4] = (double)status;
        a.out[5] = res0;
        a.out[6] = res;
        a.out[7] = alpha;
    }
}

HipErr launch_bbpgd64(hipStream_t st, SolveArgs a, int max_blocks, int *grid_used) {
    int dev = 0, ncu = 0, per_cu = 0;
    ALENS_HIP_TRY(hipGetDevice(&dev));
    ALENS_HIP_TRY(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    ALENS_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)k_bbpgd64, 256, 0));
    if (per_cu < 1) per_cu = 1;
    int64_t need = (int64_t)nblk(a.n, 256);
    if ((int64_t)a.nchunks > need) need = a.nchunks;
    if (need < 1) need = 1;
    int64_t lim = (int64_t)ncu * per_cu;
    if (max_blocks > 0 && max_blocks < lim) lim = max_blocks;
    const int grid = (int)(need < lim ? need : lim);
    *grid_used = grid;
    void *args[] = {&a};
    ALENS_HIP_TRY(hipLaunchCooperativeKernel((const void *)k_bbpgd64, dim3(grid), dim3(256), args, 0, st));
    return hip_ok();
}

// ------------------------------------------------------------------------------------------------
// write-back: F_con = D gamma (gamma = solver x), U = U_nc + M F_con; gamma copied out
// ------------------------------------------------------------------------------------------------
__global__ void k_writeback(int32_t n, const uint32_t *__restrict__ adjStart,
                            const uint32_t *__restrict__ adjK, const double *__restrict__ adjW,
                            const double2 *__restrict__ xg, const double *__restrict__ mob,
                            const double *__restrict__ unc, int mono, double *__restrict__ F,
                            double *__restrict__ U) {
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double f[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (uint32_t e = adjStart[i]; e < adjStart[i + 1]; ++e) {
        const uint32_t k = adjK[e] & 0x7fffffffu;
        const double v = xg[k].x;
        const double *w = adjW + 6 * (size_t)e;
#pragma unroll
        for (int q = 0; q < 6; ++q) f[q] = f[q] + v * w[q];
    }
    double mb[6], u[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) mb[q] = mob[6 * (size_t)i + q];
    apply_M(mb, f, mono, u);
#pragma unroll
    for (int q = 0; q < 6; ++q) {
        F[6 * (size_t)i + q] = f[q];
        U[6 * (size_t)i + q] = unc[6 * (size_t)i + q] + u[q];
    }
}

__global__ void k_copy_gamma(const double2 *__restrict__ xg, double *__restrict__ gamma, uint32_t m) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < m) gamma[k] = xg[k].x;
}

HipErr launch_writeback(hipStream_t st, int32_t n, const uint32_t *adjStart, const uint32_t *adjK,
                        const double *adjW, const double2 *xg, const double *mob,
                        const double *unc, int mono, double *F, double *U, double *gamma,
                        uint32_t m) {
    if (n > 0) {
        k_writeback<<<nblk(n, 256), 256, 0, st>>>(n, adjStart, adjK, adjW, xg, mob, unc, mono, F, U);
        LAUNCH_CHECK();
    }
    if (m > 0) {
        k_copy_gamma<<<nblk(m, 256), 256, 0, st>>>(xg, gamma, m);
        LAUNCH_CHECK();
    }
    return hip_ok();
}

// ------------------------------------------------------------------------------------------------
// Euler step: x += v dt; direction rotated by normalise(1, w dt / 2) and renormalised; periodic
// coordinates wrapped into [lo, hi).
// ------------------------------------------------------------------------------------------------
__global__ void k_euler(int32_t n, double *__restrict__ center, double *__restrict__ dir,
                        const double *__restrict__ u, CfgDev c) {
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double dt = c.dt;
    const double hdt = 0.5 * dt;
    const double *ui = u + 6 * (size_t)i;
    double *cc = center + 3 * (size_t)i, *t = dir + 3 * (size_t)i;
    double x[3], tt[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        x[k] = cc[k] + ui[k] * dt;
        tt[k] = t[k];
    }
    double q1 = hdt * ui[3], q2 = hdt * ui[4], q3 = hdt * ui[5];
    const double qn = sqrt(1.0 + (q1 * q1 + q2 * q2 + q3 * q3));
    const double q0 = 1.0 / qn;
    q1 = q1 / qn;
    q2 = q2 / qn;
    q3 = q3 / qn;
    const double a0 = q2 * tt[2] - q3 * tt[1], a1 = q3 * tt[0] - q1 * tt[2], a2 = q1 * tt[1] - q2 * tt[0];
    const double b0 = q2 * a2 - q3 * a1, b1 = q3 * a0 - q1 * a2, b2 = q1 * a1 - q2 * a0;
    const double tq = 2.0 * q0;
    const double n0 = (tt[0] + tq * a0) + 2.0 * b0;
    const double n1 = (tt[1] + tq * a1) + 2.0 * b1;
    const double n2 = (tt[2] + tq * a2) + 2.0 * b2;
    const double nn = sqrt(n0 * n0 + n1 * n1 + n2 * n2);
    t[0] = n0 / nn;
    t[1] = n1 / nn;
    t[2] = n2 / nn;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        double v = x[k];
        if (c.pbc[k]) v = v - floor((v - c.box_lo[k]) / c.len[k]) * c.len[k];
        cc[k] = v;
    }
}

HipErr launch_euler(hipStream_t st, int32_t n, double *center, double *dir, const double *u,
                    const CfgDev &c) {
    if (n == 0) return hip_ok();
    k_euler<<<nblk(n, 256), 256, 0, st>>>(n, center, dir, u, c);
    LAUNCH_CHECK();
    return hip_ok();
}

// ------------------------------------------------------------------------------------------------
// solver state set-up and the previous-step warm start (match by rod-index pair key I << 32 | J)
// ------------------------------------------------------------------------------------------------
__global__ void k_init_xg(double2 *__restrict__ xg, const double *__restrict__ x0, uint32_t m) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < m) xg[k] = make_double2(x0 ? x0[k] : 0.0, 0.0);
}
HipErr launch_init_xg(hipStream_t st, double2 *xg, const double *x0, uint32_t m) {
    if (m == 0) return hip_ok();
    k_init_xg<<<nblk(m, 256), 256, 0, st>>>(xg, x0, m);
    LAUNCH_CHECK();
    return hip_ok();
}

__global__ void k_warm_match(uint32_t m, const int32_t *__restrict__ cI, const int32_t *__restrict__ cJ,
                             uint32_t mprev, const uint64_t *__restrict__ pkey,
                             const double *__restrict__ pgam, double2 *__restrict__ xg) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= m) return;
    const uint64_t key = ((uint64_t)(uint32_t)cI[k] << 32) | (uint32_t)cJ[k];
    uint32_t lo = 0, hi = mprev;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (pkey[mid] < key)
            lo = mid + 1;
        else
            hi = mid;
    }
    const double g = (lo < mprev && pkey[lo] == key) ? pgam[lo] : 0.0;
    xg[k] = make_double2(g, 0.0);
}
HipErr launch_warm_match(hipStream_t st, uint32_t m, const int32_t *cI, const int32_t *cJ,
                         const int32_t *gid, uint32_t mprev, const uint64_t *prev_key,
                         const double *prev_gamma, double2 *xg) {
    (void)gid;
    if (m == 0) return hip_ok();
    k_warm_match<<<nblk(m, 256), 256, 0, st>>>(m, cI, cJ, mprev, prev_key, prev_gamma, xg);
    LAUNCH_CHECK();
    return hip_ok();
}

__global__ void k_make_keys(uint32_t m, const int32_t *__restrict__ cI, const int32_t *__restrict__ cJ,
                            const double2 *__restrict__ xg, uint64_t *__restrict__ key,
                            double *__restrict__ gam) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= m) return;
    key[k] = ((uint64_t)(uint32_t)cI[k] << 32) | (uint32_t)cJ[k];
    gam[k] = xg[k].x;
}
HipErr launch_make_keys(hipStream_t st, uint32_t m, const int32_t *cI, const int32_t *cJ,
                        const int32_t *gid, const double2 *xg, uint64_t *key, double *gam) {
    (void)gid;
    if (m == 0) return hip_ok();
    k_make_keys<<<nblk(m, 256), 256, 0, st>>>(m, cI, cJ, xg, key, gam);
    LAUNCH_CHECK();
    return hip_ok();
}

}  // namespace alens
