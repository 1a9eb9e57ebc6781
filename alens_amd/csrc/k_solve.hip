// k_solve.hip -- BBPGD solve of the collision BCQP as one persistent kernel per solve, the
// write-back of the contact impulses, the Euler step and the warm-start bookkeeping (gfx950).
// Reference roles: BCQPSolver::solveBBPGD, ConstraintSolver::writebackGamma and
// SylinderSystem::stepEuler (SimToolbox, CMakeLists.txt:109-115; called through
// rodSystem.runStep() / stepEuler(), SRC/TubuleSystem.cpp:223, :55).
//
// One iteration = phase A (rod-parallel: x+ = P(x - alpha g) gathered per incident constraint,
// F = D x+, U = M F), grid barrier, phase B (constraint-parallel in 512-element chunks:
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
constexpr int kBlk = 1024;                  // threads per block: 4 reduction groups of 256
constexpr int kGrp = kBlk / kRedThreads;

enum : int { Q_SS = 0, Q_SY = 1, Q_YY = 2, Q_RMAX = 3, Q_NF = 4, Q_RC = 5, Q_NFC = 6 };

// Grid barrier over co-resident blocks (cooperative launch).  Words (kBarWords apart, one cache
// line each): bar[0] = arrivals, bar[kBarWords] = abort, bar[(2 + q) * kBarWords] = arrivals of
// block group q = blockIdx % kBarGroups; for mode 3, bar[(kBarXccArr + x) * kBarWords] = arrivals
// and bar[(kBarXccSize + x) * kBarWords] = blocks of the group on XCC x, bar[kBarSetup * kBarWords]
// = the set-up count.  Counters only grow: barrier number e completes when the count reaches
// (e + 1) * (number of arrivers).  Thread 0 of each block publishes the block's writes with one
// agent-scope release fence and, after the wait, takes one agent-scope acquire fence.
//   mode 0: every block adds to bar[0] and polls it.
//   mode 1: blocks add to their group's counter; the last of a group adds to bar[0]; everyone polls
//           bar[0] (kBarGroups contenders on bar[0] instead of gridDim.x).
//   mode 2: as 1, with a longer pause between polls.
//   mode 3: groups are the blocks on one XCC (read from the hardware register), and only the last
//           arriver of a group releases.  A block's stores count as done (vmcnt 0, before its
//           arrival) once its XCC's L2 holds them, since the vector L1 writes through; the group's
//           last arriver sees every arrival of its group and then writes back that L2 with its
//           agent-scope release, so one write-back per XCC replaces one per block.  Every block
//           still acquires.  Measured with a stale-read check (2 x 10^5 rows written before each
//           barrier and read by other blocks after it, 300 rounds, 196 and 256 blocks): no stale
//           value, and 25.5 -> 18.2 us per write/barrier/read/barrier round at 196 blocks.
// The wait is bounded: past kSpinLimit polls the abort word is set and every block leaves.

__device__ inline uint32_t xcc_id() {
    uint32_t v;
    asm volatile("s_getreg_b32 %0, hwreg(HW_REG_XCC_ID)" : "=s"(v));
    return v & (uint32_t)(kBarXcc - 1);
}

// Barrier group of this block: {group, blocks in the group, groups}.  Mode 3 counts the blocks of
// each XCC once per launch (one flat, fenced rendezvous); false if that wait gave up.
struct BarGroup {
    uint32_t q, size, n;
};
__device__ inline bool bar_group_setup(uint32_t *bar, int mode, int *s_flag, BarGroup *s_bg) {
    if (threadIdx.x == 0) {
        BarGroup g;
        if (mode != 3) {
            g.q = blockIdx.x % kBarGroups;
            g.size = (gridDim.x - g.q + kBarGroups - 1) / kBarGroups;
            g.n = gridDim.x < (uint32_t)kBarGroups ? gridDim.x : (uint32_t)kBarGroups;
            *s_flag = 0;
        } else {
            g.q = xcc_id();
            __hip_atomic_fetch_add(&bar[(kBarXccSize + g.q) * kBarWords], 1u, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
            __hip_atomic_fetch_add(&bar[kBarSetup * kBarWords], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            uint32_t spins = 0;
            while (__hip_atomic_load(&bar[kBarSetup * kBarWords], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) <
                   gridDim.x) {
                __builtin_amdgcn_s_sleep(1);
                if (++spins > kSpinLimit) {
                    __hip_atomic_store(&bar[kBarWords], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    break;
                }
                if ((spins & 1023u) == 0 &&
                    __hip_atomic_load(&bar[kBarWords], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
                    break;
            }
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            g.n = 0;
            for (int x = 0; x < kBarXcc; ++x)
                g.n += __hip_atomic_load(&bar[(kBarXccSize + x) * kBarWords], __ATOMIC_RELAXED,
                                         __HIP_MEMORY_SCOPE_AGENT) > 0u;
            g.size = __hip_atomic_load(&bar[(kBarXccSize + g.q) * kBarWords], __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
            *s_flag = (int)__hip_atomic_load(&bar[kBarWords], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        *s_bg = g;
    }
    __syncthreads();
    return *s_flag == 0;
}

__device__ inline bool grid_barrier(uint32_t *bar, uint32_t &epoch, int *s_flag, int mode, const BarGroup &bg) {
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (threadIdx.x == 0) {
        if (mode != 3) __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        uint32_t target;
        if (mode == 0) {
            target = (epoch + 1u) * gridDim.x;
            __hip_atomic_fetch_add(&bar[0], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        } else if (mode == 3) {
            target = (epoch + 1u) * bg.n;
            const uint32_t old = __hip_atomic_fetch_add(&bar[(kBarXccArr + bg.q) * kBarWords], 1u,
                                                        __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (old + 1u == (epoch + 1u) * bg.size) {
                // the XCC's last arriver writes back its L2 (all of the group's stores) for everyone
                __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "agent");
                __hip_atomic_fetch_add(&bar[0], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
        } else {
            const uint32_t q = blockIdx.x % kBarGroups;
            const uint32_t ngroups = gridDim.x < (uint32_t)kBarGroups ? gridDim.x : (uint32_t)kBarGroups;
            const uint32_t gsize = (gridDim.x - q + kBarGroups - 1) / kBarGroups;
            target = (epoch + 1u) * ngroups;
            const uint32_t old = __hip_atomic_fetch_add(&bar[(2 + q) * kBarWords], 1u, __ATOMIC_RELAXED,
                                                        __HIP_MEMORY_SCOPE_AGENT);
            if (old + 1u == (epoch + 1u) * gsize) {
                // the group's last arriver passes the group's published writes on to bar[0]
                __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "agent");
                __hip_atomic_fetch_add(&bar[0], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
        }
        uint32_t spins = 0;
        while (__hip_atomic_load(&bar[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
            if (mode == 2)
                __builtin_amdgcn_s_sleep(8);
            else
                __builtin_amdgcn_s_sleep(1);
            if (++spins > kSpinLimit) {
                __hip_atomic_store(&bar[kBarWords], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                break;
            }
            if ((spins & 1023u) == 0 &&
                __hip_atomic_load(&bar[kBarWords], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
                break;
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        *s_flag = (int)__hip_atomic_load(&bar[kBarWords], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    ++epoch;
    __syncthreads();
    return *s_flag == 0;
}

// Fixed-order sums (NS of them) and maxima (NX) within each 256-lane reduction group: each lane's
// partial is already accumulated left to right; each wave folds by halving (lane l += lane l + h,
// h = 32 .. 1); the group's four wave results are added (or compared) left to right, with one LDS
// exchange for all of them.  Called by every thread of the block (all groups together).
// Halving fold of one wave: lane l takes lane l + h for h = 32, 16, ..., 1 (sums: x + o; maxima:
// o > x ? o : x) and the result is read from lane 0.  The partner values come from the gfx950 lane
// swaps (h = 32 and 16) and DPP row shifts (h <= 8), register moves without the LDS round trip of a
// general shuffle; lanes at or above h receive values that never reach lane 0.
__device__ inline double lanes_up32(double x) {
    const unsigned lo = (unsigned)__double2loint(x), hi = (unsigned)__double2hiint(x);
    const auto a = __builtin_amdgcn_permlane32_swap(lo, lo, false, false);
    const auto b = __builtin_amdgcn_permlane32_swap(hi, hi, false, false);
    return __hiloint2double((int)b[1], (int)a[1]);
}
__device__ inline double lanes_up16(double x) {
    const unsigned lo = (unsigned)__double2loint(x), hi = (unsigned)__double2hiint(x);
    const auto a = __builtin_amdgcn_permlane16_swap(lo, lo, false, false);
    const auto b = __builtin_amdgcn_permlane16_swap(hi, hi, false, false);
    return __hiloint2double((int)b[1], (int)a[1]);
}
template <int H>
__device__ inline double lanes_up_row(double x) {  // DPP row_shl:H (H < 16)
    const unsigned lo = (unsigned)__double2loint(x), hi = (unsigned)__double2hiint(x);
    return __hiloint2double((int)__builtin_amdgcn_update_dpp(0u, hi, 0x100 + H, 0xF, 0xF, true),
                            (int)__builtin_amdgcn_update_dpp(0u, lo, 0x100 + H, 0xF, 0xF, true));
}
template <bool MAX>
__device__ inline double fold_step(double x, double o) {
    if (MAX) return o > x ? o : x;
    return x + o;
}
template <bool MAX>
__device__ inline double wave_fold(double x) {
    x = fold_step<MAX>(x, lanes_up32(x));
    x = fold_step<MAX>(x, lanes_up16(x));
    x = fold_step<MAX>(x, lanes_up_row<8>(x));
    x = fold_step<MAX>(x, lanes_up_row<4>(x));
    x = fold_step<MAX>(x, lanes_up_row<2>(x));
    x = fold_step<MAX>(x, lanes_up_row<1>(x));
    return x;
}

template <int NS, int NX>
__device__ inline void group_fold(const double (&vs)[NS], const double (&vx)[NX], double *sh /* kGrp * (NS + NX) * 4 */,
                                  double (&ts)[NS], double (&tx)[NX]) {
    constexpr int NQ = NS + NX;
    const int lane = threadIdx.x & 63, w = (threadIdx.x >> 6) & 3, g = threadIdx.x >> 8;
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        double x = q < NS ? vs[q] : vx[q - NS];
        x = q >= NS ? wave_fold<true>(x) : wave_fold<false>(x);
        if (lane == 0) sh[(g * NQ + q) * 4 + w] = x;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        const double *s4 = sh + (g * NQ + q) * 4;
        if (q >= NS) {
            double t = s4[0];
            t = s4[1] > t ? s4[1] : t;
            t = s4[2] > t ? s4[2] : t;
            t = s4[3] > t ? s4[3] : t;
            tx[q - NS] = t;
        } else {
            ts[q] = ((s4[0] + s4[1]) + s4[2]) + s4[3];
        }
    }
    __syncthreads();
}

// Per-system totals over chunk partials: for each system s (four at a time, one per 256-lane
// group), lane c % 256 of the group accumulates the partials of the system's chunks c, c + 256, ...
// left to right, then the group folds them (sums of slots S[], maxes of slots X[]).  Results land in
// tot[q][s] (q < NS: sums, then the maxes) for every thread; systems not active are skipped.
enum : int { S_ACTIVE = 0, S_CONVERGED = 1, S_NONFINITE = 2, S_NOCONV = 3 };

// The same totals with one wave per system, for many small systems (replicas): lane l of the wave
// holds the virtual lanes l, l + 64, l + 128 and l + 192 of the system's 256-lane group, accumulates
// each one's chunks in order, folds each virtual wave by halving with shuffles and adds the four
// results left to right -- the order above, with no LDS exchange and sixteen systems in flight.
constexpr int kWaveFoldChunks = 4 * kRedThreads;  // most chunks per system the wave fold takes
template <int NS, int NX>
__device__ inline void fold_waves(const double *part, int nchunks, const uint32_t *chunk_off, int nsys,
                                  const int *state, const int (&S)[NS], const int (&X)[NX], double (*tot)[kMaxSys]) {
    constexpr int NQ = NS + NX, kW = kBlk / 64;
    const int wv = threadIdx.x >> 6, wl = threadIdx.x & 63;
    for (int s0 = 0; s0 < nsys; s0 += kW) {
        const int sy = s0 + wv;
        double vs[4][NS], vx[4][NX];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
#pragma unroll
            for (int q = 0; q < NS; ++q) vs[v][q] = 0.0;
#pragma unroll
            for (int q = 0; q < NX; ++q) vx[v][q] = 0.0;
        }
        if (sy < nsys && state[sy] == S_ACTIVE) {
            const int cb = (int)chunk_off[sy], c1 = (int)chunk_off[sy + 1];
            // chunk cb + v * 64 + wl + 256 j is term j of virtual lane v * 64 + wl
            for (int c = cb + wl; c < c1; c += kRedThreads) {
                double ls[4][NS], lx[4][NX];
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int cc = c + v * 64;
#pragma unroll
                    for (int q = 0; q < NS; ++q) ls[v][q] = cc < c1 ? part[(size_t)S[q] * nchunks + cc] : 0.0;
#pragma unroll
                    for (int q = 0; q < NX; ++q) lx[v][q] = cc < c1 ? part[(size_t)X[q] * nchunks + cc] : 0.0;
                }
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    if (c + v * 64 >= c1) continue;
#pragma unroll
                    for (int q = 0; q < NS; ++q) vs[v][q] = vs[v][q] + ls[v][q];
#pragma unroll
                    for (int q = 0; q < NX; ++q) vx[v][q] = lx[v][q] > vx[v][q] ? lx[v][q] : vx[v][q];
                }
            }
        }
        // a virtual wave that holds no chunk of the system (all of its lanes +0.0) folds to +0.0: its
        // fold is skipped (the condition is the same for the whole wave)
        const int nchs = sy < nsys && state[sy] == S_ACTIVE ? (int)(chunk_off[sy + 1] - chunk_off[sy]) : 0;
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            double t[4];
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                double x = q < NS ? vs[v][q] : vx[v][q - NS];
                if (v * 64 < nchs)
                    x = q >= NS ? wave_fold<true>(x) : wave_fold<false>(x);
                else
                    x = 0.0;
                t[v] = x;
            }
            if (wl == 0 && sy < nsys) {
                if (q >= NS) {
                    double m = t[0];
                    m = t[1] > m ? t[1] : m;
                    m = t[2] > m ? t[2] : m;
                    m = t[3] > m ? t[3] : m;
                    tot[q][sy] = m;
                } else {
                    tot[q][sy] = ((t[0] + t[1]) + t[2]) + t[3];
                }
            }
        }
    }
    __syncthreads();
}

template <int NS, int NX>
__device__ inline void fold_systems(const double *part, int nchunks, const uint32_t *chunk_off, int nsys,
                                    const int *state, const int (&S)[NS], const int (&X)[NX], double *sh,
                                    double (*tot)[kMaxSys], bool by_wave) {
    if (by_wave) {
        fold_waves<NS, NX>(part, nchunks, chunk_off, nsys, state, S, X, tot);
        return;
    }
    const int g = threadIdx.x >> 8, lane = threadIdx.x & 255;
    for (int s0 = 0; s0 < nsys; s0 += kGrp) {
        const int sy = s0 + g;
        double vs[NS], vx[NX];
#pragma unroll
        for (int q = 0; q < NS; ++q) vs[q] = 0.0;
#pragma unroll
        for (int q = 0; q < NX; ++q) vx[q] = 0.0;
        if (sy < nsys && state[sy] == S_ACTIVE) {
            // chunks c, c + 256, ... in order; the loads of kAhead of them are issued together
            constexpr int kAhead = 4;
            const int c1 = (int)chunk_off[sy + 1];
            for (int c = (int)chunk_off[sy] + lane; c < c1; c += kAhead * kRedThreads) {
                double ls[kAhead][NS], lx[kAhead][NX];
#pragma unroll
                for (int u = 0; u < kAhead; ++u) {
                    const int cc = c + u * kRedThreads;
#pragma unroll
                    for (int q = 0; q < NS; ++q) ls[u][q] = cc < c1 ? part[(size_t)S[q] * nchunks + cc] : 0.0;
#pragma unroll
                    for (int q = 0; q < NX; ++q) lx[u][q] = cc < c1 ? part[(size_t)X[q] * nchunks + cc] : 0.0;
                }
#pragma unroll
                for (int u = 0; u < kAhead; ++u) {
                    if (c + u * kRedThreads >= c1) break;
#pragma unroll
                    for (int q = 0; q < NS; ++q) vs[q] = vs[q] + ls[u][q];
#pragma unroll
                    for (int q = 0; q < NX; ++q) vx[q] = lx[u][q] > vx[q] ? lx[u][q] : vx[q];
                }
            }
        }
        double ts[NS], tx[NX];
        group_fold<NS, NX>(vs, vx, sh, ts, tx);
        if (lane == 0 && sy < nsys) {
#pragma unroll
            for (int q = 0; q < NS; ++q) tot[q][sy] = ts[q];
#pragma unroll
            for (int q = 0; q < NX; ++q) tot[NS + q][sy] = tx[q];
        }
    }
    __syncthreads();
}

enum : int { A_PROJ = 0, A_GRAD = 1 };

template <typename T>
__device__ inline typename Vec2<T>::type mk2(T x, T y);
template <>
__device__ inline double2 mk2<double>(double x, double y) {
    return make_double2(x, y);
}
template <>
__device__ inline float2 mk2<float>(float x, float y) {
    return make_float2(x, y);
}

// adjacency entries a phase-A lane loads per step (their loads are in flight together): a template
// parameter of the solver, at most what the 128-register budget of the persistent kernel holds
// without spilling (3 in FP64, 4 in FP32); the autotuner picks among the instantiated values
template <typename T>
struct AGroup {
    static constexpr int max = sizeof(T) == 8 ? 3 : 4;
};

// Per-rod data that stays fixed through a solve, for the rod at slice position p: its index,
// system, degree and first slot in the sliced adjacency.  Each thread keeps those of its first
// position in registers.
struct RodMeta {
    uint32_t deg, idx;
    int sy, rod;
};
template <typename T>
__device__ inline RodMeta rod_meta(const SolveArgsT<T> &a, int p) {
    RodMeta r;
    r.rod = a.sord ? (int)a.sord[p] : p;
    r.sy = a.rsys[r.rod];
    r.deg = a.adjStart[r.rod + 1] - a.adjStart[r.rod];
    r.idx = a.sellStart[p >> 6] + (uint32_t)(p & (kSlice - 1));
    return r;
}

// value of the iterate seen by one adjacency entry: x+ = P(x - alpha g) (A_PROJ) or g (A_GRAD),
// from the constraint's (x, g) and lower bound
template <int MODE, typename T>
__device__ inline T a_value(typename Vec2<T>::type xg, T l, T alpha) {
    if (MODE == A_GRAD) return xg.y;
    return proj(xg.x - alpha * xg.y, l);
}

// Work index of this block.  Workgroups are dealt to the XCDs round-robin (workgroup b runs on XCD
// b % 8); with xcd_map set, the blocks of one XCD take consecutive work indices, so each XCD walks
// a contiguous run of rods and constraints and its L2 serves the neighbour gathers.
__device__ inline int work_block(int xcd_map) {
    const int b = blockIdx.x, G = gridDim.x;
    if (!xcd_map || G < kBarGroups) return b;
    const int x = b % kBarGroups, q = b / kBarGroups;
    return x * (G / kBarGroups) + (x < G % kBarGroups ? x : G % kBarGroups) + q;
}

// Walks the adjacency entries j0 <= j < j1 of the rod whose first slot is idx0, in groups of kAG (a
// group's gathers and weights load together, with the constraint indices of the next group).  With
// PROD the products v_k * w_e go to prod[j - j0][.]; otherwise they are added to f one entry at a
// time, in order.  The products are rounded the same way either way (no contraction), so a sum that
// adds stored products in order equals the running sum.
template <int MODE, typename T, int kAG, bool PROD>
__device__ inline void walk_entries(const SolveArgsT<T> &a, size_t idx0, uint32_t j0, uint32_t j1, T alpha,
                                    T (&f)[6], T (*prod)[6]) {
    using V2 = typename Vec2<T>::type;
    const size_t ep = a.epad;
    constexpr uint32_t kMask = 0x7fffffffu;
    size_t idx = idx0 + (size_t)j0 * kSlice;
    uint32_t kc[kAG];
#pragma unroll
    for (int g = 0; g < kAG; ++g) kc[g] = j0 + g < j1 ? a.sellK[idx + g * kSlice] & kMask : 0u;
    for (uint32_t j = j0; j < j1; j += kAG, idx += kAG * kSlice) {
        V2 xv[kAG];
        T lv[kAG], w[kAG][6];
        uint32_t kn[kAG];
#pragma unroll
        for (int g = 0; g < kAG; ++g) {
            const bool on = j + g < j1;
            const size_t ig = idx + g * kSlice;
            xv[g] = on ? a.xg[kc[g]] : mk2<T>(T(0), T(0));
            lv[g] = (on && a.lb) ? a.lb[kc[g]] : T(0);
#pragma unroll
            for (int c = 0; c < 6; ++c) w[g][c] = on ? a.sellW[c * ep + ig] : T(0);
            kn[g] = j + kAG + g < j1 ? a.sellK[idx + (kAG + g) * kSlice] & kMask : 0u;
        }
#pragma unroll
        for (int g = 0; g < kAG; ++g) {
            if (j + g >= j1) break;
            const T v = a_value<MODE, T>(xv[g], lv[g], alpha);
#pragma unroll
            for (int c = 0; c < 6; ++c) {
                if (PROD)
                    prod[j + g - j0][c] = v * w[g][c];
                else
                    f[c] = f[c] + v * w[g][c];
            }
        }
#pragma unroll
        for (int g = 0; g < kAG; ++g) kc[g] = kn[g];
    }
}

// U = M F of a rod
template <typename T>
__device__ inline void store_U(const SolveArgsT<T> &a, const RodMeta &rm, const T (&f)[6]) {
    using V2 = typename Vec2<T>::type;
    const size_t i = (size_t)rm.rod;
    const V2 *m2 = reinterpret_cast<const V2 *>(a.mob + 6 * i);
    const V2 m01 = m2[0], m23 = m2[1], m45 = m2[2];
    const T mb[6] = {m01.x, m01.y, m23.x, m23.y, m45.x, m45.y};
    T u[6];
    apply_M(mb, f, a.smono[rm.sy], u);
    V2 *U2 = reinterpret_cast<V2 *>(a.U + kUStride * i);
    U2[0] = mk2<T>(u[0], u[1]);
    U2[1] = mk2<T>(u[2], u[3]);
    U2[2] = mk2<T>(u[4], u[5]);
}

// Split of heavy rods (one slice position per lane): in a block, the rod at position t < kSplitN whose
// degree exceeds kSplitH gets lane kBlk - 1 - t (the lightest end of the degree-ordered window) as a
// helper when that lane's own rod has at most kSplitLight entries.  The helper forms the products of
// entries kSplitH .. kSplitH + kSplitMax - 1 in LDS; the owner sums its first kSplitH entries, then adds
// the stored products and walks the rest, so each rod's sum keeps its ascending order.
constexpr int kSplitN = 128, kSplitH = 9, kSplitMax = 9;
constexpr uint32_t kSplitLight = 3;
struct SplitMeta {
    uint32_t idx, deg;  // owner's first slot and degree; deg 0 = no helper
    int sy;
};

// phase A: per rod, F = sum over incident constraints (ascending k) of v_k * w_e; U = M F
// (rods of systems that have stopped are skipped).  Lane p % 64 of a wave walks the entries of the
// rod at slice position p, so each step of the 64 lanes reads consecutive words.
// rod metadata of a lane's passes 1 .. kPasses - 1 (pass 0 is kept in registers), fixed through a
// solve: a pass of a stopped system costs one LDS read instead of five dependent global loads
constexpr int kPasses = 6;

// Slice of its block's 1024 positions that wave w walks in pass q.  Slices are in descending degree
// within each degree window (1024 or 128 rods), so with the identity every pass gives the same waves
// the heavy slices (wave 0 alone, or the even waves) and the block waits on them.  Passes
// q = 0, 1, 2, 3 (mod 4) give wave w the slices w, W-1-w, (w + W/2) % W and W-1-(w + W/2) % W, so
// each wave alternates between the heavy and the light ends.  Every position of pass q lies below
// every position of pass q + 1, so a lane whose pass-q position is past the rods has no later pass
// either.
constexpr int kWaves = kBlk / 64;
__device__ inline int pass_pos(int base, int q, int stride, int t) {
    const int w = t >> 6;
    const int v = (q & 2) ? (w + kWaves / 2) % kWaves : w;
    const int s = (q & 1) ? kWaves - 1 - v : v;
    return base + q * stride + s * 64 + (t & 63);
}

template <int MODE, typename T, int kAG, bool SPLIT>
__device__ inline void phase_A(const SolveArgsT<T> &a, const double *s_alpha, const int *s_state,
                               const RodMeta &rm0, bool split, const SplitMeta *s_help,
                               T (*s_prod)[kSplitMax][6], const RodMeta (*s_rmeta)[kBlk]) {
    const int stride = gridDim.x * blockDim.x;
    const int i0 = work_block(a.xcd_map) * blockDim.x + threadIdx.x;
    if (SPLIT && split) {
        // one position per lane (n <= stride): helpers first, then every lane's own rod; owners finish
        // after the block has synchronised.  The lane index passes through an empty asm so that the
        // helper addresses are formed here each time instead of being kept live across the solve.
        int t = threadIdx.x;
        asm volatile("" : "+v"(t));
        const volatile SplitMeta *vh = s_help;  // read per iteration, not kept live
        if (t >= kBlk - kSplitN) {
            const SplitMeta h{vh[kBlk - 1 - t].idx, vh[kBlk - 1 - t].deg, vh[kBlk - 1 - t].sy};
            if (h.deg > 0 && s_state[h.sy] == S_ACTIVE) {
                const uint32_t j1 =
                    h.deg < (uint32_t)(kSplitH + kSplitMax - 1) ? h.deg : (uint32_t)(kSplitH + kSplitMax - 1);
                T none[6];
                walk_entries<MODE, T, kAG, true>(a, h.idx, kSplitH, j1, (T)s_alpha[h.sy], none, s_prod[kBlk - 1 - t]);
            }
        }
        // an owner parks its partial sum in the last product row of its helper slot (rows past the
        // helper's entries are free: the helper takes at most kSplitMax - 1 entries below)
        bool owner = false;
        if (i0 < a.n && s_state[rm0.sy] == S_ACTIVE) {
            T f[6] = {T(0), T(0), T(0), T(0), T(0), T(0)};
            owner = t < kSplitN && vh[t].deg > 0;
            walk_entries<MODE, T, kAG, false>(a, rm0.idx, 0, owner ? (uint32_t)kSplitH : rm0.deg,
                                              (T)s_alpha[rm0.sy], f, nullptr);
            if (owner) {
#pragma unroll
                for (int c = 0; c < 6; ++c) s_prod[t][kSplitMax - 1][c] = f[c];
            } else {
                store_U(a, rm0, f);
            }
        }
        __syncthreads();
        if (owner) {
            T f[6];
#pragma unroll
            for (int c = 0; c < 6; ++c) f[c] = s_prod[t][kSplitMax - 1][c];
            const uint32_t jh = rm0.deg < (uint32_t)(kSplitH + kSplitMax - 1) ? rm0.deg
                                                                               : (uint32_t)(kSplitH + kSplitMax - 1);
            for (uint32_t q = 0; q < jh - kSplitH; ++q) {
#pragma unroll
                for (int c = 0; c < 6; ++c) f[c] = f[c] + s_prod[t][q][c];
            }
            walk_entries<MODE, T, kAG, false>(a, rm0.idx, jh, rm0.deg, (T)s_alpha[rm0.sy], f, nullptr);
            store_U(a, rm0, f);
        }
        return;
    }
    const int base = i0 - (int)threadIdx.x;
    for (int q = 0;; ++q) {
        const int p = pass_pos(base, q, stride, threadIdx.x);
        if (p >= a.n) break;
        const RodMeta rm = q == 0 ? rm0 : (q < kPasses ? s_rmeta[q - 1][threadIdx.x] : rod_meta(a, p));
        if (s_state[rm.sy] != S_ACTIVE) continue;
        T f[6] = {T(0), T(0), T(0), T(0), T(0), T(0)};
        walk_entries<MODE, T, kAG, false>(a, rm.idx, 0, rm.deg, (T)s_alpha[rm.sy], f, nullptr);
        store_U(a, rm, f);
    }
}


// A load of a word that the phase reads once: with NT it carries the non-temporal hint, so that the
// line goes out of the caches first and the rod velocities and the partials stay longer.
template <bool NT, typename X>
__device__ inline X ld_stream(const X *p) {
    if constexpr (NT)
        return __builtin_nontemporal_load(p);
    else
        return *p;
}

template <typename T, bool NT>
__device__ inline T du_row(const SolveArgsT<T> &a, uint32_t k) {
    using V2 = typename Vec2<T>::type;
    T gk[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) gk[q] = ld_stream<NT>(a.geo + (size_t)q * a.m + k);
    const size_t I = (size_t)ld_stream<NT>(a.cI + k), J = (size_t)ld_stream<NT>(a.cJ + k);
    const V2 *UI = reinterpret_cast<const V2 *>(a.U + kUStride * I);
    const V2 *UJ = reinterpret_cast<const V2 *>(a.U + kUStride * J);
    T uI[6], uJ[6];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        const V2 p = UI[q], r = UJ[q];
        uI[2 * q] = p.x;
        uI[2 * q + 1] = p.y;
        uJ[2 * q] = r.x;
        uJ[2 * q + 1] = r.y;
    }
    return dt_row(gk, uI, uJ);
}

enum : int { B_INIT = 0, B_ALPHA0 = 1, B_ITER = 2 };

// Phase-B reduction groups.  A 512-element chunk is reduced by GL lanes (GL = 256, or 128 when the
// schedule asks for it); in the fixed order element e belongs to virtual lane e % 256 (pass
// e / 256), and lane l of a group holds the virtual lanes l + j * GL (j < 256 / GL).  Each physical
// wave folds its virtual waves separately by halving (lane l += lane l + h, h = 32 .. 1) and the four
// virtual-wave totals are added left to right, so the result is the one a 256-lane group gives.
constexpr int kGrpB = kBlk / 128;  // most groups per block (GL = 128)
template <int GL>
struct GroupShape {
    static constexpr int vpt = kRedThreads / GL;  // virtual lanes per lane
    static constexpr int per_block = kBlk / GL;   // groups per block
};

// Folds NS sums (vs) and NX maxima (vx) of every group in one pass: each physical wave folds its
// virtual waves by halving, one LDS exchange, then each group adds (or takes the maximum of) its four
// virtual-wave results left to right.  sh holds groups * (NS + NX) * 4 doubles.
template <int NS, int NX, int GL>
__device__ inline void vgroup_fold(const double (*vs)[GroupShape<GL>::vpt], const double (*vx)[GroupShape<GL>::vpt],
                                   double *sh, double *ts, double *tx) {
    constexpr int VPT = GroupShape<GL>::vpt, NQ = NS + NX;
    const int l = threadIdx.x % GL, g = threadIdx.x / GL, lane = l & 63, pw = l >> 6;
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
#pragma unroll
        for (int j = 0; j < VPT; ++j) {
            double x = q < NS ? vs[q][j] : vx[q - NS][j];
            x = q >= NS ? wave_fold<true>(x) : wave_fold<false>(x);
            if (lane == 0) sh[(g * NQ + q) * 4 + pw + j * (GL / 64)] = x;
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        const double *s4 = sh + (g * NQ + q) * 4;
        if (q >= NS) {
            double t = s4[0];
            t = s4[1] > t ? s4[1] : t;
            t = s4[2] > t ? s4[2] : t;
            t = s4[3] > t ? s4[3] : t;
            tx[q - NS] = t;
        } else {
            ts[q] = ((s4[0] + s4[1]) + s4[2]) + s4[3];
        }
    }
    __syncthreads();
}

// Chunk table of a block, fixed through a solve: for round r (chunk c = work_block * G + g +
// r * gridDim * G, G groups per block) its system and constraint range.  The first kChunkRounds rounds are kept in
// LDS; later rounds (very large systems) read the tables.
constexpr int kChunkRounds = 16;
struct ChunkMeta {
    uint32_t k0, k1;
    int sy;
};
template <typename T>
__device__ inline ChunkMeta chunk_meta(const SolveArgsT<T> &a, int c, int nch) {
    ChunkMeta m{0u, 0u, -1};
    if (c < nch) {
        m.sy = a.chunk_sys[c];
        m.k0 = a.con_off[m.sy] + (uint32_t)(c - (int)a.chunk_off[m.sy]) * kRedChunk;
        const uint32_t k1 = m.k0 + kRedChunk, kend = a.con_off[m.sy + 1];
        m.k1 = k1 < kend ? k1 : kend;
    }
    return m;
}

// Active chunks of many systems, compacted: s_cact[s] = chunks of the active systems before s, so the
// groups of phase B walk only the chunks of systems still iterating (with fixed chunk ownership, a
// round of chunks costs the same while any of its systems is active).  Wave 0 computes it from the
// per-system states; the grid barrier before phase B publishes it to the block.
__device__ inline void compact_chunks(const int *s_state, const uint32_t *s_choff, int ns, uint32_t *s_cact) {
    static_assert(kMaxSys <= 4 * 64, "four systems per lane");
    if (threadIdx.x >= 64) return;
    const int l = threadIdx.x;
    uint32_t cnt[4], sum = 0;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const int q = 4 * l + v;
        cnt[v] = (q < ns && s_state[q] == S_ACTIVE) ? s_choff[q + 1] - s_choff[q] : 0u;
        sum += cnt[v];
    }
    uint32_t inc = sum;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t t = __shfl_up(inc, o, 64);
        if (l >= o) inc += t;
    }
    uint32_t run = inc - sum;
    if (l == 0) s_cact[0] = 0u;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const int q = 4 * l + v;
        run += cnt[v];
        if (q < ns) s_cact[q + 1] = run;
    }
}

// phase B: constraints in 512-element chunks, one chunk per GL-lane group; chunk partials go to
// part[q * nchunks + c].  NT: the row streams (geometry, rod ids, rhs) load non-temporally.
// COMPACT: the groups walk the compacted chunks of the active systems (s_cact), each mapped back to its
// system and chunk; a chunk's elements, order and partial slot are the same either way.
template <int MODE, typename T, int GL, bool NT = false, bool COMPACT = false>
__device__ inline void phase_B(const SolveArgsT<T> &a, const double *s_alpha, const int *s_state, double *sh,
                               const ChunkMeta (*s_cm)[kGrpB], int nch, const uint32_t *s_cact = nullptr,
                               const uint32_t *s_choff = nullptr, int ns = 0) {
    constexpr int kVPT = GroupShape<GL>::vpt, G = GroupShape<GL>::per_block;
    constexpr int NE = kVPT * kRedPasses;  // elements per lane
    const int g = threadIdx.x / GL, l = threadIdx.x % GL;
    const int ngroups = gridDim.x * G;
    const T dt = (T)a.dt;
    int r = 0;
    const int nloop = COMPACT ? (int)s_cact[ns] : nch;
    for (int base = work_block(a.xcd_map) * G; base < nloop; base += ngroups, ++r) {
        int c = base + g;
        ChunkMeta cm{0u, 0u, -1};
        if (COMPACT) {
            if (c < nloop) {
                // the active system holding compacted chunk c: s_cact[lo] <= c < s_cact[hi]
                int lo = 0, hi = ns;
                while (hi - lo > 1) {
                    const int mid = (lo + hi) >> 1;
                    if ((int)s_cact[mid] <= c) lo = mid; else hi = mid;
                }
                const uint32_t j = (uint32_t)c - s_cact[lo];
                cm.sy = lo;
                cm.k0 = a.con_off[lo] + j * kRedChunk;
                const uint32_t k1 = cm.k0 + kRedChunk, kend = a.con_off[lo + 1];
                cm.k1 = k1 < kend ? k1 : kend;
                c = (int)(s_choff[lo] + j);
            }
        } else {
            cm = r < kChunkRounds ? s_cm[r][g] : chunk_meta(a, c, nch);
        }
        // per virtual lane: sums (s^T s, s^T y, y^T y, non-finite count), max residual, and for a warm
        // start the cold-start residual and its non-finite count
        double vs[5][kVPT], vx[2][kVPT];
#pragma unroll
        for (int j = 0; j < kVPT; ++j) {
            vs[0][j] = vs[1][j] = vs[2][j] = vs[3][j] = vs[4][j] = 0.0;
            vx[0][j] = vx[1][j] = 0.0;
        }
        const int sy = cm.sy < 0 ? 0 : cm.sy;
        const bool act = cm.sy >= 0 && s_state[sy] == S_ACTIVE;
        if (act) {
            const T alpha = (T)s_alpha[sy];
            // every load of the round (including the rhs) is issued before the first store to xg,
            // which the compiler could not move them past
            T xo[NE], go[NE], du[NE], bb[NE];
            uint32_t kk[NE];
            bool ok[NE];
#pragma unroll
            for (int j = 0; j < kVPT; ++j) {
#pragma unroll
                for (int p = 0; p < kRedPasses; ++p) {
                    const int q = j * kRedPasses + p;
                    kk[q] = cm.k0 + (uint32_t)(l + j * GL + p * kRedThreads);
                    ok[q] = kk[q] < cm.k1;
                    if (ok[q]) {
                        const typename Vec2<T>::type xg = a.xg[kk[q]];
                        xo[q] = xg.x;
                        go[q] = xg.y;
                        if (MODE != B_ALPHA0) bb[q] = ld_stream<NT>(a.b + kk[q]);
                        du[q] = du_row<T, NT>(a, kk[q]);
                    }
                }
            }
#pragma unroll
            for (int j = 0; j < kVPT; ++j) {
#pragma unroll
                for (int p = 0; p < kRedPasses; ++p) {
                    const int q = j * kRedPasses + p;
                    if (!ok[q]) continue;
                    const uint32_t k = kk[q];
                    const T lo = a.lb ? a.lb[k] : T(0);
                    if (MODE == B_ALPHA0) {
                        T ag = dt * du[q];
                        if (a.invk) ag = ag + a.invk[k] * go[q];
                        vs[0][j] = vs[0][j] + (double)(go[q] * go[q]);
                        vs[1][j] = vs[1][j] + (double)(go[q] * ag);
                    } else {
                        const T xn = proj(xo[q] - alpha * go[q], lo);
                        const T bk = bb[q];
                        T gn = dt * du[q] + bk;
                        if (a.invk) gn = gn + a.invk[k] * xn;
                        const double res = (double)fabs(xn - proj(xn - gn, lo));
                        if (!isfinite(res))
                            vs[3][j] = vs[3][j] + 1.0;
                        else if (res > vx[0][j])
                            vx[0][j] = res;
                        if (MODE == B_INIT) {
                            if (a.warm) {
                                const double r0 = (double)fabs(T(0) - proj(T(0) - bk, lo));
                                if (!isfinite(r0))
                                    vs[4][j] = vs[4][j] + 1.0;
                                else if (r0 > vx[1][j])
                                    vx[1][j] = r0;
                            }
                        } else {
                            const T sv = xn - xo[q];
                            const T yv = gn - go[q];
                            vs[0][j] = vs[0][j] + (double)(sv * sv);
                            vs[1][j] = vs[1][j] + (double)(sv * yv);
                            vs[2][j] = vs[2][j] + (double)(yv * yv);
                        }
                        a.xg[k] = mk2<T>(xn, gn);
                    }
                }
            }
        }
        double ts[5], tx[2];
        if (MODE == B_INIT && a.warm)
            vgroup_fold<5, 2, GL>(vs, vx, sh, ts, tx);
        else
            vgroup_fold<4, 1, GL>(vs, vx, sh, ts, tx);
        double *pc = a.part + c;
        const int nc = a.nchunks;
        if (act && l == 0) {
            if (MODE != B_INIT) {
                pc[(size_t)Q_SS * nc] = ts[0];
                pc[(size_t)Q_SY * nc] = ts[1];
            }
            if (MODE == B_ITER) pc[(size_t)Q_YY * nc] = ts[2];
            if (MODE != B_ALPHA0) {
                pc[(size_t)Q_RMAX * nc] = tx[0];
                pc[(size_t)Q_NF * nc] = ts[3];
            }
        }
        if (MODE == B_INIT && a.warm && act && l == 0) {
            pc[(size_t)Q_RC * nc] = tx[1];
            pc[(size_t)Q_NFC * nc] = ts[4];
        }
    }
}

// Per-system report words in sysout[8 s ..]: iters, matvecs, converged, fallback, status, res0, res,
// alpha at exit.  All blocks hold identical per-system state, so control flow stays uniform.
// MANY: the instance for more systems than reduction groups per block, whose fold may take one wave
// per system (a separate instance, so the single-system kernels keep their register budget).
// NT: the constraint phase's row streams load non-temporally (an instance of its own for the same
// reason: both load forms in one kernel spilled 16 B more per lane and slowed the 200k slab)
template <typename T, int AG, bool SPLIT, bool MANY, bool NT>
__global__ __launch_bounds__(kBlk) void k_bbpgd(SolveArgsT<T> a) {
    __shared__ double sh[kGrpB * 8 * 4];
    __shared__ int s_flag;
    __shared__ double s_alpha[kMaxSys], s_res0[kMaxSys], s_res[kMaxSys];
    __shared__ int s_state[kMaxSys], s_iters[kMaxSys], s_mv[kMaxSys], s_fb[kMaxSys];
    __shared__ double tot[5][kMaxSys];
    __shared__ int s_any;
    __shared__ ChunkMeta s_cm[kChunkRounds][kGrpB];
    const int ns = a.nsys;
    const int bm = a.bar_mode;
    // row and chunk counts: on the device for a step replayed from a captured graph
    const int nch = a.nch_dev ? (*a.nch_dev < a.nchunks ? *a.nch_dev : a.nchunks) : a.nchunks;
    const uint32_t mrows = dcount(a.m, a.m_dev);
    // 256-lane groups (two elements per lane) unless the schedule asks for 128 (four per lane, half the
    // rounds of chunks).  With the wave folds in register moves, 128-lane groups were faster only
    // around 1,000-1,100 chunks on 256 blocks (3.8-6%); at 645 chunks they were 12% slower, and at
    // 1,500-6,300 chunks 1.4-4.8% slower.  The split kernel keeps to 256-lane groups (the 128-lane
    // path would not fit its register budget).
    // A width the autotuner picked holds for row counts near the tuned step's (gw_rows_lo .. hi); the
    // device's row count decides, so a replayed graph step follows it too.
    const int G = kGrpB / 2;
    const bool gw_on = a.gw_rows_hi == 0 || (mrows >= a.gw_rows_lo && mrows <= a.gw_rows_hi);
    const bool wide = SPLIT || !gw_on || a.group_width != 128;
    const int gb = wide ? G : kGrpB;
    if (threadIdx.x < kChunkRounds * gb) {
        const int r = threadIdx.x / gb, g = threadIdx.x % gb;
        s_cm[r][g] = chunk_meta(a, work_block(a.xcd_map) * gb + g + r * gridDim.x * gb, nch);
    }
    const int i0 = work_block(a.xcd_map) * blockDim.x + threadIdx.x;
    const RodMeta rm0 = i0 < a.n ? rod_meta(a, i0) : RodMeta{0u, 0u, 0, 0};
    __shared__ RodMeta s_rmeta[kPasses - 1][kBlk];
    {
        const int stride = gridDim.x * blockDim.x;
        for (int q = 1; q < kPasses; ++q) {
            const int p = pass_pos(i0 - (int)threadIdx.x, q, stride, threadIdx.x);
            if (p >= a.n) break;
            s_rmeta[q - 1][threadIdx.x] = rod_meta(a, p);
        }
    }
    // heavy-rod helpers (phase A), fixed through the solve
    __shared__ SplitMeta s_help[kSplitN];
    __shared__ T s_prod[kSplitN][kSplitMax][6];
    const bool split = SPLIT && a.n <= (int)(gridDim.x * blockDim.x);
    {
        __shared__ uint32_t s_deg[kBlk];
        s_deg[threadIdx.x] = i0 < a.n ? rm0.deg : 0xffffffffu;
        __syncthreads();
        if (threadIdx.x < kSplitN) {
            const uint32_t pd = s_deg[kBlk - 1 - threadIdx.x];
            const bool help = split && i0 < a.n && rm0.deg > (uint32_t)kSplitH && pd <= kSplitLight;
            s_help[threadIdx.x] = help ? SplitMeta{rm0.idx, rm0.deg, rm0.sy} : SplitMeta{0u, 0u, 0};
        }
    }
    // each system's chunk range, fixed through the solve: the fold reads it from LDS instead of
    // starting every fold with a dependent global load
    __shared__ uint32_t s_choff[kMaxSys + 1];
    __shared__ uint32_t s_cact[MANY ? kMaxSys + 1 : 1];
    for (int q = threadIdx.x; q <= ns; q += blockDim.x) s_choff[q] = a.chunk_off[q];
    // many systems, none with more chunks than a wave folds quickly: one wave per system in the fold
    __shared__ int s_big;
    if (threadIdx.x == 0) s_big = 0;
    __syncthreads();
    for (int q = threadIdx.x; q < ns; q += blockDim.x)
        if (s_choff[q + 1] - s_choff[q] > (uint32_t)kWaveFoldChunks) s_big = 1;
    __syncthreads();
    const bool by_wave = MANY && ns > kGrp && !s_big;
    for (int q = threadIdx.x; q < ns; q += blockDim.x) {
        s_alpha[q] = 0.0;
        s_state[q] = S_ACTIVE;
        s_iters[q] = 0;
        s_mv[q] = 1;
        s_fb[q] = 0;
        s_res0[q] = 0.0;
        s_res[q] = 0.0;
    }
    __syncthreads();
    uint32_t epoch = 0;
    __shared__ BarGroup s_bg;
    if (!bar_group_setup(a.bar, bm, &s_flag, &s_bg)) return;
    const BarGroup bg = s_bg;

    // start point: x = P(x0), g = A x + b
    phase_A<A_PROJ, T, AG, SPLIT>(a, s_alpha, s_state, rm0, split, s_help, s_prod, s_rmeta);
    if (!grid_barrier(a.bar, epoch, &s_flag, bm, bg)) return;
    if (wide || SPLIT)
        phase_B<B_INIT, T, 256>(a, s_alpha, s_state, sh, s_cm, nch);
    else
        phase_B<B_INIT, T, 128>(a, s_alpha, s_state, sh, s_cm, nch);
    if (!grid_barrier(a.bar, epoch, &s_flag, bm, bg)) return;
    {
        const int S[1] = {Q_NF}, X[1] = {Q_RMAX};
        fold_systems<1, 1>(a.part, a.nchunks, s_choff, ns, s_state, S, X, sh, tot, by_wave);
    }
    if (a.warm) {
        // tot[0] = NF0, tot[1] = R0 -> keep them in tot[2], tot[3]
        for (int q = threadIdx.x; q < ns; q += blockDim.x) {
            tot[2][q] = tot[0][q];
            tot[3][q] = tot[1][q];
        }
        __syncthreads();
        const int S[1] = {Q_NFC}, X[1] = {Q_RC};
        fold_systems<1, 1>(a.part, a.nchunks, s_choff, ns, s_state, S, X, sh, tot, by_wave);
        // tot[0] = NFc, tot[1] = Rc
        if (threadIdx.x == 0) s_any = 0;
        __syncthreads();
        for (int q = threadIdx.x; q < ns; q += blockDim.x) {
            if (tot[2][q] > 0.0 || tot[3][q] > tot[1][q]) {
                s_fb[q] = 1;
                tot[2][q] = tot[0][q];
                tot[3][q] = tot[1][q];
                s_any = 1;
            }
        }
        __syncthreads();
        if (s_any) {
            // warm start rejected for those systems: x = 0, g = 0 + b
            for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < mrows; k += gridDim.x * blockDim.x)
                if (s_fb[a.rsys[a.cI[k]]]) a.xg[k] = mk2<T>(T(0), T(0) + a.b[k]);
            if (!grid_barrier(a.bar, epoch, &s_flag, bm, bg)) return;
        }
        for (int q = threadIdx.x; q < ns; q += blockDim.x) {
            tot[0][q] = tot[2][q];
            tot[1][q] = tot[3][q];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) s_any = 0;
    __syncthreads();
    for (int q = threadIdx.x; q < ns; q += blockDim.x) {
        s_res0[q] = tot[1][q];
        s_res[q] = tot[1][q];
        if (tot[0][q] > 0.0)
            s_state[q] = S_NONFINITE;
        else if (tot[1][q] <= a.tol)
            s_state[q] = S_CONVERGED;
        else
            s_any = 1;
    }
    __syncthreads();
    if (s_any) {
        // alpha0 = g^T g / g^T A g
        phase_A<A_GRAD, T, AG, SPLIT>(a, s_alpha, s_state, rm0, split, s_help, s_prod, s_rmeta);
        if (!grid_barrier(a.bar, epoch, &s_flag, bm, bg)) return;
        if (wide || SPLIT)
            phase_B<B_ALPHA0, T, 256>(a, s_alpha, s_state, sh, s_cm, nch);
        else
            phase_B<B_ALPHA0, T, 128>(a, s_alpha, s_state, sh, s_cm, nch);
        if (!grid_barrier(a.bar, epoch, &s_flag, bm, bg)) return;
        {
            const int S[2] = {Q_SS, Q_SY}, X[1] = {Q_RMAX};
            fold_systems<2, 1>(a.part, a.nchunks, s_choff, ns, s_state, S, X, sh, tot, by_wave);
        }
        for (int q = threadIdx.x; q < ns; q += blockDim.x) {
            if (s_state[q] != S_ACTIVE) continue;
            ++s_mv[q];
            double al = tot[0][q] / tot[1][q];
            if (!(al > 0.0 && al < INFINITY)) al = 1.0;
            s_alpha[q] = al;
        }
        __syncthreads();
        if (MANY) compact_chunks(s_state, s_choff, ns, s_cact);  // read after the next grid barrier
        for (int it = 1; it <= a.max_ite; ++it) {
            phase_A<A_PROJ, T, AG, SPLIT>(a, s_alpha, s_state, rm0, split, s_help, s_prod, s_rmeta);
            if (!grid_barrier(a.bar, epoch, &s_flag, bm, bg)) return;
            if (wide || SPLIT)
                phase_B<B_ITER, T, 256, NT, MANY>(a, s_alpha, s_state, sh, s_cm, nch, s_cact, s_choff, ns);
            else
                phase_B<B_ITER, T, 128, NT, MANY>(a, s_alpha, s_state, sh, s_cm, nch, s_cact, s_choff, ns);
            if (!grid_barrier(a.bar, epoch, &s_flag, bm, bg)) return;
            {
                // the step length of this iteration needs two of the three sums: s^T s / s^T y
                // (odd) or s^T y / y^T y (even); the fold reads only those, so tot[0] / tot[1] is
                // the step either way, tot[2] the non-finite count and tot[3] the residual
                const int X[1] = {Q_RMAX};
                if (it & 1) {
                    const int S[3] = {Q_SS, Q_SY, Q_NF};
                    fold_systems<3, 1>(a.part, a.nchunks, s_choff, ns, s_state, S, X, sh, tot, by_wave);
                } else {
                    const int S[3] = {Q_SY, Q_YY, Q_NF};
                    fold_systems<3, 1>(a.part, a.nchunks, s_choff, ns, s_state, S, X, sh, tot, by_wave);
                }
            }
            if (threadIdx.x == 0) s_any = 0;
            __syncthreads();
            for (int q = threadIdx.x; q < ns; q += blockDim.x) {
                if (s_state[q] != S_ACTIVE) continue;
                ++s_mv[q];
                s_iters[q] = it;
                s_res[q] = tot[3][q];
                if (tot[2][q] > 0.0) {
                    s_state[q] = S_NONFINITE;
                } else if (tot[3][q] <= a.tol) {
                    s_state[q] = S_CONVERGED;
                } else {
                    const double an = tot[0][q] / tot[1][q];
                    if (an > 0.0 && an < INFINITY) s_alpha[q] = an;
                    s_any = 1;
                }
            }
            __syncthreads();
            if (!s_any) break;
            if (MANY) compact_chunks(s_state, s_choff, ns, s_cact);  // read after the next grid barrier
        }
    }
    if (blockIdx.x == 0) {
        for (int q = threadIdx.x; q < ns; q += blockDim.x) {
            const int st = s_state[q];
            double *o = a.out + 8 * (size_t)q;
            o[0] = (double)s_iters[q];
            o[1] = (double)s_mv[q];
            o[2] = st == S_CONVERGED ? 1.0 : 0.0;
            o[3] = (double)s_fb[q];
            o[4] = st == S_CONVERGED ? 0.0 : (st == S_NONFINITE ? (double)ALENS_E_NONFINITE : (double)ALENS_E_NOCONV);
            o[5] = s_res0[q];
            o[6] = s_res[q];
            o[7] = (double)(a.con_off[q + 1] - a.con_off[q]);  // the system's rows
        }
    }
}

template <typename T, int AG, bool SPLIT, bool MANY, bool NT>
static HipErr launch_bbpgd_k(hipStream_t st, SolveArgsT<T> a, int max_blocks, int *grid_used) {
    int dev = 0, ncu = 0, per_cu = 0;
    ALENS_HIP_TRY(hipGetDevice(&dev));
    ALENS_HIP_TRY(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    ALENS_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)k_bbpgd<T, AG, SPLIT, MANY, NT>, kBlk, 0));
    if (per_cu < 1) per_cu = 1;
    int64_t need = (int64_t)nblk(a.n, kBlk);
    const int64_t needB = ((int64_t)a.nchunks + kGrpB - 1) / kGrpB;
    if (needB > need) need = needB;
    if (need < 1) need = 1;
    int64_t lim = (int64_t)ncu * per_cu;
    if (max_blocks > 0 && max_blocks < lim) lim = max_blocks;
    const int grid = (int)(need < lim ? need : lim);
    *grid_used = grid;
    void *args[] = {&a};
    ALENS_HIP_TRY(hipLaunchCooperativeKernel((const void *)k_bbpgd<T, AG, SPLIT, MANY, NT>, dim3(grid), dim3(kBlk), args, 0, st));
    return hip_ok();
}

// The heavy-rod split variant (FP32 only) when every lane holds at most one rod.  Measured on the
// 200k-rod slab: FP32 solve 4.21 -> 3.99 ms; in FP64 the split kernel sits at the 128-register budget
// with spills and was slower (4.79 -> 4.92 ms).
template <typename T, int AG>
static HipErr launch_bbpgd(hipStream_t st, SolveArgsT<T> a, int max_blocks, int *grid_used) {
    if constexpr (sizeof(T) == 4) {
        int dev = 0, ncu = 0;
        ALENS_HIP_TRY(hipGetDevice(&dev));
        ALENS_HIP_TRY(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
        const int64_t lanes = (int64_t)(max_blocks > 0 && max_blocks < ncu ? max_blocks : ncu) * kBlk;
        if (a.split_ok && (int64_t)a.n <= lanes)
            return launch_bbpgd_k<T, AG, true, false, false>(st, a, max_blocks, grid_used);
    }
    if (a.nsys > kGrp)
        return a.stream_nt ? launch_bbpgd_k<T, AG, false, true, true>(st, a, max_blocks, grid_used)
                           : launch_bbpgd_k<T, AG, false, true, false>(st, a, max_blocks, grid_used);
    return a.stream_nt ? launch_bbpgd_k<T, AG, false, false, true>(st, a, max_blocks, grid_used)
                       : launch_bbpgd_k<T, AG, false, false, false>(st, a, max_blocks, grid_used);
}

HipErr launch_bbpgd64(hipStream_t st, SolveArgsT<double> a, int max_blocks, int entries, int *grid_used) {
    if (entries == 2) return launch_bbpgd<double, 2>(st, a, max_blocks, grid_used);
    return launch_bbpgd<double, AGroup<double>::max>(st, a, max_blocks, grid_used);
}
HipErr launch_bbpgd32(hipStream_t st, SolveArgsT<float> a, int max_blocks, int entries, int *grid_used) {
    if (entries == 2) return launch_bbpgd<float, 2>(st, a, max_blocks, grid_used);
    return launch_bbpgd<float, AGroup<float>::max>(st, a, max_blocks, grid_used);
}

// precision-form conversions
__global__ void k_to_f32(const double *__restrict__ s, float *__restrict__ d, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) d[i] = (float)s[i];
}
HipErr launch_to_f32(hipStream_t st, const double *src, float *dst, size_t count) {
    if (count == 0) return hip_ok();
    k_to_f32<<<nblk((int64_t)count, 256), 256, 0, st>>>(src, dst, count);
    LAUNCH_CHECK();
    return hip_ok();
}
__global__ void k_xg32_to_x64(const float2 *__restrict__ xg, double *__restrict__ x, uint32_t m) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < m) x[k] = (double)xg[k].x;
}
HipErr launch_xg32_to_x64(hipStream_t st, const float2 *xg32, double *x64, uint32_t m) {
    if (m == 0) return hip_ok();
    k_xg32_to_x64<<<nblk(m, 256), 256, 0, st>>>(xg32, x64, m);
    LAUNCH_CHECK();
    return hip_ok();
}
__global__ void k_init_xg32(float2 *__restrict__ xg, const double2 *__restrict__ xg64, uint32_t m) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < m) xg[k] = make_float2((float)xg64[k].x, 0.0f);
}
HipErr launch_init_xg32(hipStream_t st, float2 *xg, const double2 *xg64, uint32_t m) {
    if (m == 0) return hip_ok();
    k_init_xg32<<<nblk(m, 256), 256, 0, st>>>(xg, xg64, m);
    LAUNCH_CHECK();
    return hip_ok();
}

// ------------------------------------------------------------------------------------------------
// write-back: F_con = D gamma (gamma = solver x), U = U_nc + M F_con; gamma copied out
// ------------------------------------------------------------------------------------------------
__global__ void k_writeback(int32_t n, const int32_t *__restrict__ rsys, const SysDev *__restrict__ sys,
                            const uint32_t *__restrict__ adjStart, const uint32_t *__restrict__ sellStart,
                            const uint32_t *__restrict__ sellK, const double *__restrict__ sellW, uint64_t epad,
                            const double2 *__restrict__ xg,
                            const double *__restrict__ mob, const double *__restrict__ unc,
                            double *__restrict__ F, double *__restrict__ U, const uint32_t *__restrict__ spos) {
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int mono = sys[rsys[i]].mono;
    double f[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    const uint32_t deg = adjStart[i + 1] - adjStart[i];
    const uint32_t p = spos ? spos[i] : (uint32_t)i;
    size_t idx = (size_t)sellStart[p >> 6] + (p & (kSlice - 1));
    for (uint32_t j = 0; j < deg; ++j, idx += kSlice) {
        const uint32_t k = sellK[idx] & 0x7fffffffu;
        const double v = xg[k].x;
#pragma unroll
        for (int q = 0; q < 6; ++q) f[q] = f[q] + v * sellW[q * epad + idx];
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

// start point of the caller rows k0 .. k0 + me: x = the blocks' gamma, g = 0
__global__ void k_set_x(double2 *__restrict__ xg, const double *__restrict__ x0, uint32_t me) {
    const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < me) xg[e] = make_double2(x0[e], 0.0);
}
HipErr launch_set_x(hipStream_t st, double2 *xg, const double *x0, uint32_t me) {
    if (me == 0) return hip_ok();
    k_set_x<<<nblk(me, 256), 256, 0, st>>>(xg, x0, me);
    LAUNCH_CHECK();
    return hip_ok();
}

__global__ void k_copy_gamma(const double2 *__restrict__ xg, double *__restrict__ gamma, uint32_t m,
                             const uint32_t *__restrict__ md) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < dcount(m, md)) gamma[k] = xg[k].x;
}

HipErr launch_writeback(hipStream_t st, int32_t n, const int32_t *rsys, const SysDev *sys, const uint32_t *adjStart,
                        const uint32_t *sellStart, const uint32_t *sellK, const double *sellW, uint64_t epad,
                        const double2 *xg, const double *mob, const double *unc, double *F, double *U,
                        double *gamma, uint32_t m, const uint32_t *md, const uint32_t *spos) {
    if (n > 0) {
        k_writeback<<<nblk(n, 256), 256, 0, st>>>(n, rsys, sys, adjStart, sellStart, sellK, sellW, epad, xg, mob, unc,
                                                  F, U, spos);
        LAUNCH_CHECK();
    }
    if (m > 0) {
        k_copy_gamma<<<nblk(m, 256), 256, 0, st>>>(xg, gamma, m, md);
        LAUNCH_CHECK();
    }
    return hip_ok();
}

// ------------------------------------------------------------------------------------------------
// Euler step: x += v dt; direction rotated by normalise(1, w dt / 2) and renormalised; periodic
// coordinates wrapped into [lo, hi).
// ------------------------------------------------------------------------------------------------
__global__ void k_euler(int32_t n, const int32_t *__restrict__ rsys, const SysDev *__restrict__ sys,
                        double *__restrict__ center, double *__restrict__ dir, const double *__restrict__ u,
                        CfgDev c) {
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const SysDev &S = sys[rsys[i]];
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
        if (S.pbc[k]) v = v - floor((v - S.box_lo[k]) / S.len[k]) * S.len[k];
        cc[k] = v;
    }
}

HipErr launch_euler(hipStream_t st, int32_t n, const int32_t *rsys, const SysDev *sys, double *center, double *dir,
                    const double *u, const CfgDev &c) {
    if (n == 0) return hip_ok();
    k_euler<<<nblk(n, 256), 256, 0, st>>>(n, rsys, sys, center, dir, u, c);
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
                             const double *__restrict__ pgam, double2 *__restrict__ xg,
                             const uint32_t *__restrict__ md, const uint32_t *__restrict__ mprev_d) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= dcount(m, md)) return;
    if (mprev_d) mprev = *mprev_d;
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
                         const double *prev_gamma, double2 *xg, const uint32_t *md, const uint32_t *mprev_d) {
    (void)gid;
    if (m == 0) return hip_ok();
    k_warm_match<<<nblk(m, 256), 256, 0, st>>>(m, cI, cJ, mprev, prev_key, prev_gamma, xg, md, mprev_d);
    LAUNCH_CHECK();
    return hip_ok();
}

__global__ void k_fill_u32_one(uint32_t *p, uint32_t v) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *p = v;
}

__global__ void k_step_tail(const uint32_t *__restrict__ entries, const uint32_t *__restrict__ pairs,
                            const uint32_t *__restrict__ slots, const uint32_t *__restrict__ rows,
                            const int32_t *__restrict__ chunks, const uint64_t *__restrict__ stamps,
                            const uint32_t *__restrict__ abort_word, const double *__restrict__ out,
                            const double *__restrict__ out32, int nrep, double *__restrict__ tail) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t == 0) {
        tail[0] = (double)*entries;
        tail[1] = (double)*pairs;
        tail[2] = (double)*slots;
        tail[3] = (double)*rows;
        tail[4] = (double)*chunks;
        for (int q = 0; q < 4; ++q) tail[5 + q] = (double)stamps[q];
        tail[9] = (double)*abort_word;
    }
    for (int i = t; i < nrep; i += gridDim.x * blockDim.x) {
        tail[kTailHead + i] = out[i];
        if (out32) tail[kTailHead + nrep + i] = out32[i];
    }
}
HipErr launch_step_tail(hipStream_t st, const uint32_t *entries, const uint32_t *pairs, const uint32_t *slots,
                        const uint32_t *rows, const int32_t *chunks, const uint64_t *stamps, const uint32_t *abort_word,
                        const double *out, const double *out32, int nrep, double *tail) {
    k_step_tail<<<nblk(nrep > 0 ? nrep : 1, 256), 256, 0, st>>>(entries, pairs, slots, rows, chunks, stamps, abort_word,
                                                                out, out32, nrep, tail);
    LAUNCH_CHECK();
    return hip_ok();
}

// One reading of the device's constant-rate wall clock into t[idx]: the timing points of a captured
// step (an event record inside stream capture is rejected by some HIP runtime builds).
__global__ void k_stamp(uint64_t *t, int idx) {
    if (blockIdx.x == 0 && threadIdx.x == 0) t[idx] = wall_clock64();
}
HipErr launch_stamp(hipStream_t st, uint64_t *t, int idx) {
    k_stamp<<<1, 64, 0, st>>>(t, idx);
    LAUNCH_CHECK();
    return hip_ok();
}

__global__ void k_make_keys(uint32_t m, const int32_t *__restrict__ cI, const int32_t *__restrict__ cJ,
                            const double2 *__restrict__ xg, uint64_t *__restrict__ key,
                            double *__restrict__ gam, const uint32_t *__restrict__ md, uint32_t *__restrict__ mprev_d) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t mm = dcount(m, md);
    if (k == 0 && mprev_d) *mprev_d = mm;
    if (k >= mm) return;
    key[k] = ((uint64_t)(uint32_t)cI[k] << 32) | (uint32_t)cJ[k];
    gam[k] = xg[k].x;
}
HipErr launch_make_keys(hipStream_t st, uint32_t m, const int32_t *cI, const int32_t *cJ,
                        const int32_t *gid, const double2 *xg, uint64_t *key, double *gam, const uint32_t *md,
                        uint32_t *mprev_d) {
    (void)gid;
    if (m == 0) {
        if (mprev_d) {
            k_fill_u32_one<<<1, 64, 0, st>>>(mprev_d, 0u);
            LAUNCH_CHECK();
        }
        return hip_ok();
    }
    k_make_keys<<<nblk(m, 256), 256, 0, st>>>(m, cI, cJ, xg, key, gam, md, mprev_d);
    LAUNCH_CHECK();
    return hip_ok();
}

}  // namespace alens
