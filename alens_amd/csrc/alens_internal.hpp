// alens_internal.hpp -- host-side types shared by the HIP translation units of libalens.so.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "../../include/alens_capi.h"

namespace alens {

#define ALENS_HIP_TRY(call)                                                                     \
    do {                                                                                        \
        hipError_t e_ = (call);                                                                 \
        if (e_ != hipSuccess) return ::alens::HipErr{e_, #call, __FILE__, __LINE__};             \
    } while (0)

struct HipErr {
    hipError_t code = hipSuccess;
    const char *what = nullptr;
    const char *file = nullptr;
    int line = 0;
    bool ok() const { return code == hipSuccess; }
};
inline HipErr hip_ok() { return HipErr{}; }

// growable device buffer (contents not preserved on growth); freed when it goes out of scope
template <typename T>
struct DBuf {
    T *p = nullptr;
    size_t cap = 0;
    DBuf() = default;
    DBuf(const DBuf &) = delete;
    DBuf &operator=(const DBuf &) = delete;
    DBuf(DBuf &&o) noexcept : p(o.p), cap(o.cap) {
        o.p = nullptr;
        o.cap = 0;
    }
    DBuf &operator=(DBuf &&o) noexcept {
        if (this != &o) {
            release();
            p = o.p;
            cap = o.cap;
            o.p = nullptr;
            o.cap = 0;
        }
        return *this;
    }
    ~DBuf() { release(); }
    hipError_t ensure(size_t n) {
        if (n <= cap && p) return hipSuccess;
        if (p) {
            hipError_t e = hipFree(p);
            if (e != hipSuccess) return e;
            p = nullptr;
            cap = 0;
        }
        size_t want = n < 16 ? 16 : n + n / 8;
        hipError_t e = hipMalloc((void **)&p, want * sizeof(T));
        if (e != hipSuccess) {
            p = nullptr;
            return e;
        }
        cap = want;
        return hipSuccess;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
};

struct Grid {
    int nc[3];
    double h[3];
    double len[3];
    int64_t ncell;
    int keybits;
};

// One independent rod system of a context (a context holds 1 or many): its box and cell grid.
struct SysDev {
    double box_lo[3], len[3], h[3];
    double infl;        // AABB / tube inflation: max collision radius + colBuf / 2 + margin
    int32_t pbc[3], nc[3];
    int32_t mono;
    uint32_t cell_off;  // first global cell id of this system
    int32_t rod_off, rod_end;
};
constexpr int kMaxSys = 256;

struct RodsDev {
    int32_t n;
    const int32_t *rsys;  // system of each rod
    const SysDev *sys;
    const int32_t *gid;
    const double *center;  // 3n
    const double *dir;     // 3n
    const double *hcol;    // collision half-length (0 for spheres)
    const double *rcol;    // collision radius
};

struct CfgDev {
    double box_lo[3], len[3];
    int pbc[3];
    int mono;
    double col_buf, dt, kT, mu;
    uint64_t seed;
    int64_t step;
    const int64_t *step_dev;  // when set, the step counter is read here (a replayed graph's kernels)
};

// Counts that live on the device (a step that is replayed from a captured graph never reads them
// back): kernels take the host capacity (grid size, buffer bound) and, when the pointer is set, use
// min(capacity, *count).
__device__ inline uint32_t dcount(uint32_t cap, const uint32_t *d) {
    if (!d) return cap;
    const uint32_t v = *d;
    return v < cap ? v : cap;
}

// ---- scratch for the generic device primitives ----
struct ScanScratch {
    DBuf<uint32_t> lvl[4];
};

// collide kernels (k_collide.hip)
HipErr launch_mob_unc(hipStream_t st, const RodsDev &r, const double *length, const double *radius,
                      const uint8_t *immovable, const double *fext, const CfgDev &c, double *mob, double *unc);
HipErr launch_aabb(hipStream_t st, const RodsDev &r, const CfgDev &c, int32_t *clo, int32_t *chi, double *blo,
                   double *bhi, uint32_t *ecnt);
HipErr exclusive_scan_u32(hipStream_t st, const uint32_t *in, uint32_t *out, int64_t n,
                          ScanScratch &s);
HipErr ensure_scan_scratch(ScanScratch &s, int64_t n);
HipErr ensure_radix_scratch(DBuf<uint32_t> &hist, ScanScratch &s, int64_t n);
// entries beyond ecap are not written (the count eoff[n] still tells how many there were)
HipErr launch_fill_entries(hipStream_t st, const RodsDev &r, const int32_t *clo, const int32_t *chi,
                           const uint32_t *eoff, uint32_t *keys, uint32_t *vals, uint32_t ecap);
// n: capacity (grid); nd: device count or NULL
HipErr radix_sort_pairs(hipStream_t st, uint32_t *keys, uint32_t *vals, uint32_t *keys_alt,
                        uint32_t *vals_alt, int64_t n, int keybits, DBuf<uint32_t> &hist,
                        ScanScratch &s, bool &result_in_alt, const uint32_t *nd = nullptr);
HipErr launch_fill_u32(hipStream_t st, uint32_t *p, uint32_t v, int64_t n);
// several u32 buffers zeroed by one launch
struct ZeroJobs {
    uint32_t *p[8];
    int64_t n[8];
    int count;
};
HipErr launch_zero_many(hipStream_t st, const ZeroJobs &z);
// End of a captured step: the counts, clock readings, barrier abort word and solver reports gathered
// into one buffer (read back with one copy): tail[0..4] = entries, pairs, adjacency slots, valid rows,
// chunks; tail[5..8] = clock readings; tail[9] = abort word; tail[16 ..] = the FP64 reports (nrep
// words), then the FP32 ones when out32 is given.
constexpr int kTailHead = 16;
HipErr launch_step_tail(hipStream_t st, const uint32_t *entries, const uint32_t *pairs, const uint32_t *slots,
                        const uint32_t *rows, const int32_t *chunks, const uint64_t *stamps, const uint32_t *abort_word,
                        const double *out, const double *out32, int nrep, double *tail);
HipErr launch_cell_ranges(hipStream_t st, const uint32_t *keys, int64_t E, uint32_t *cstart,
                          uint32_t *cend, const uint32_t *Ed = nullptr);
HipErr launch_sweep(hipStream_t st, const RodsDev &r, const CfgDev &c, const uint32_t *keys, int64_t E,
                    const uint32_t *cstart, const uint32_t *cend, const uint32_t *vals, const int32_t *clo, const double *blo,
                    const double *bhi, uint32_t *pair_count, uint32_t cap, uint32_t *pI, uint32_t *pJ,
                    uint32_t *pImg, uint32_t *cntI, uint32_t *cntJ, const uint32_t *Ed = nullptr);
HipErr launch_morton_keys(hipStream_t st, const RodsDev &r, int32_t i0, int32_t n, uint32_t *keys,
                          uint32_t *vals);
HipErr launch_permute_rods(hipStream_t st, int32_t n, const uint32_t *perm, const int32_t *gid,
                           const double *center, const double *dir, const double *length, const double *radius,
                           const double *hcol, const double *rcol, const uint8_t *immov, int32_t *gid2,
                           double *center2, double *dir2, double *length2, double *radius2, double *hcol2,
                           double *rcol2, uint8_t *immov2);
HipErr launch_order_I(hipStream_t st, int32_t n, uint32_t m, const uint32_t *pI,
                      const uint32_t *pJ, const uint32_t *pImg, const uint32_t *rowStart,
                      uint32_t *cur, int32_t *cI, int32_t *cJ, uint32_t *cImg, const uint32_t *md = nullptr);
HipErr launch_order_J(hipStream_t st, int32_t n, uint32_t m, const int32_t *cJ,
                      const uint32_t *jStart, uint32_t *cur, uint32_t *jlist, const uint32_t *md = nullptr);
// constraint geometry is stored component-major: geo[q * gs + k], q = 0..8 = {n, rI, rJ}, gs = the
// step's total row count (collision rows first, then the caller's rows)
HipErr launch_rows(hipStream_t st, const RodsDev &r, const CfgDev &c, uint32_t m, uint32_t gs, const int32_t *cI,
                   const int32_t *cJ, const uint32_t *cImg, const double *unc, double *delta0,
                   double *geo, double *posI, double *posJ, double *b, const uint32_t *md = nullptr);
HipErr launch_rows_ext(hipStream_t st, uint32_t me, uint32_t k0, uint32_t gs, const int32_t *cI, const int32_t *cJ,
                       const double *nrm, const double *posI, const double *posJ, const double *delta0,
                       const double *unc, double dt, double *geo, double *b);

// Rod-major adjacency (column blocks of D) in sliced, wave-interleaved form: rods are taken in
// slices of kSlice consecutive indices; entry j of rod i (ascending constraint index, J-side entries
// first) sits at sellStart[i / kSlice] + j * kSlice + i % kSlice, so the kSlice lanes of a wave that
// walk their rods' entries in step read consecutive words.  sellK holds the constraint index (top
// bit set on J-side entries); sellW holds the six weights component-major, sellW[c * epad + idx]:
// (n, rI) on the I side, (-n, -rJ) on the J side.  adjStart keeps the CSR row starts (degrees).
// Caller rows (k = mc + e) follow a rod's collision entries: extStart (n + 1, or NULL) / extEnt
// list, per rod in ascending e, e with the top bit set on the J side.
// Slot order (optional): within each window of `win` consecutive rods, rods take slice positions
// by descending degree (ties by index), so a slice's 64 rods have similar degrees and few padding
// slots.  sord[p] = rod at position p, spos[i] = position of rod i; NULL = identity.  Only which lane
// walks which rod changes: every rod's sums still run over its entries in ascending k.
// Windows are kOrderWin rods when every rod has a solver lane of its own (one pass: the lanes of a
// block hold one window, and the heavy-rod split pairs its heaviest and lightest rods), and
// kOrderWinPasses rods when lanes walk several rods in turn: the rods of one small window, whose
// constraint ranges are adjacent, are then walked at the same time, so their gathers share lines
// (1M-rod box: solver fetches 1150 -> 1065 MB per iteration, 274 -> 268 us).
constexpr int kSlice = 64;
constexpr int kOrderWin = 1024;
constexpr int kOrderWinPasses = 128;
inline int slot_order_window(int64_t n, int ncu) { return n > (int64_t)ncu * kOrderWin ? kOrderWinPasses : kOrderWin; }
HipErr launch_slot_order(hipStream_t st, int32_t n, const uint32_t *rowStart, const uint32_t *jStart,
                         const uint32_t *extStart, uint32_t *sord, uint32_t *spos, int win);
HipErr launch_slice_sizes(hipStream_t st, int32_t n, const uint32_t *rowStart, const uint32_t *jStart,
                          const uint32_t *extStart, uint32_t *slice_len, const uint32_t *sord = nullptr);
HipErr launch_adjacency(hipStream_t st, int32_t n, uint32_t m, const uint32_t *rowStart,
                        const uint32_t *jStart, const uint32_t *jlist, const uint32_t *extStart,
                        const uint32_t *extEnt, uint32_t mc, const double *geo, const uint32_t *sellStart,
                        uint64_t epad, uint32_t *adjStart, uint32_t *sellK, double *sellW,
                        const uint32_t *spos = nullptr);
HipErr launch_guard_entries(hipStream_t st, const uint32_t *eoff_n, uint32_t ecap, uint32_t *ee);
HipErr launch_guard_pairs(hipStream_t st, const uint32_t *count, uint32_t cap, uint32_t *mv, uint32_t *cntI,
                          uint32_t *cntJ, int32_t n);
// Chunk table on the device: con_off[s] = first row of system s (rows grouped by rod I; me caller rows
// close the last system), chunk_off (scan of the systems' 512-row chunk counts), *nchunks, and
// chunk_sys for the first chunk_cap chunks.
HipErr launch_chunk_table(hipStream_t st, int32_t nsys, int32_t n, const SysDev *sys, const uint32_t *rowStart,
                          uint32_t me, uint32_t *con_off, uint32_t *chunk_off, int32_t *nchunks, int32_t *chunk_sys,
                          int32_t chunk_cap, const uint32_t *mv);
HipErr launch_guard_slots(hipStream_t st, const uint32_t *slots, uint32_t cap, uint32_t *mv);

// solve kernels (k_solve.hip)
template <typename T>
struct Vec2;
template <>
struct Vec2<double> {
    using type = double2;
};
template <>
struct Vec2<float> {
    using type = float2;
};

// Arguments of the persistent BBPGD kernel; T is the storage/arithmetic type of the iterate,
// operator data and velocities (double: plain method; float: precision form).  Sums of products are
// accumulated in double in both cases.
template <typename T>
struct SolveArgsT {
    int32_t n;
    uint32_t m;
    const uint32_t *adjStart;    // CSR row starts (rod degrees)
    const uint32_t *sellStart;   // sliced adjacency (see launch_adjacency)
    const uint32_t *sord;        // rod at each slice position (see launch_slot_order), NULL = identity
    const uint32_t *sellK;
    const T *sellW;
    uint64_t epad;               // component stride of sellW
    const T *mob;
    const int32_t *cI;
    const int32_t *cJ;
    const T *geo;                // component-major, stride m
    const T *b;
    const T *lb;                 // NULL = all zero
    const T *invk;               // NULL = no compliance; else A += diag(invk)
    typename Vec2<T>::type *xg;  // (x, g) per constraint; x holds the start point on entry
    T *U;                        // kUStride n scratch (rod velocities, padded rows)
    double *part;                // kRedQ * nchunks partials
    uint32_t *bar;               // barrier words (zeroed before launch)
    double *out;                 // per-system report words (8 per system)
    const int32_t *rsys;         // system of each rod
    const int32_t *smono;        // monolayer flag per system
    const uint32_t *con_off;     // constraint range per system (nsys + 1)
    const uint32_t *chunk_off;   // reduction-chunk range per system (nsys + 1)
    const int32_t *chunk_sys;    // system of each chunk
    double dt, tol;
    int32_t max_ite;
    int32_t nsys;
    int32_t warm;
    int32_t nchunks;             // chunk capacity (stride of part)
    const int32_t *nch_dev;      // chunk count on the device, or NULL (= nchunks)
    const uint32_t *m_dev;       // row count on the device, or NULL (= m, which is also the geo stride)
    int32_t bar_mode;            // grid barrier variant (k_solve.hip)
    int32_t xcd_map;             // 1: blocks of one XCD take consecutive work (k_solve.hip)
    int32_t split_ok;            // 1: heavy rods may take helper lanes in phase A (k_solve.hip)
    int32_t group_width;         // phase-B reduction groups: 128 = 128 lanes, else 256 (k_solve.hip)
    uint32_t gw_rows_lo, gw_rows_hi;  // group_width applies to row counts in [lo, hi] (hi = 0: any)
    int32_t stream_nt;           // 1: phase B loads its row streams non-temporally (k_solve.hip)
};
using SolveArgs = SolveArgsT<double>;
// entries: adjacency entries in flight per phase-A step (2, or 0 for the largest instantiated)
HipErr launch_bbpgd64(hipStream_t st, SolveArgsT<double> a, int max_blocks, int entries, int *grid_used);
HipErr launch_bbpgd32(hipStream_t st, SolveArgsT<float> a, int max_blocks, int entries, int *grid_used);
HipErr launch_to_f32(hipStream_t st, const double *src, float *dst, size_t count);
HipErr launch_xg32_to_x64(hipStream_t st, const float2 *xg32, double *x64, uint32_t m);
HipErr launch_init_xg32(hipStream_t st, float2 *xg, const double2 *xg64, uint32_t m);
HipErr launch_writeback(hipStream_t st, int32_t n, const int32_t *rsys, const SysDev *sys, const uint32_t *adjStart,
                        const uint32_t *sellStart, const uint32_t *sellK, const double *sellW, uint64_t epad,
                        const double2 *xg, const double *mob, const double *unc, double *F, double *U,
                        double *gamma, uint32_t m, const uint32_t *md = nullptr, const uint32_t *spos = nullptr);
HipErr launch_euler(hipStream_t st, int32_t n, const int32_t *rsys, const SysDev *sys, double *center, double *dir,
                    const double *u, const CfgDev &c);
HipErr launch_init_xg(hipStream_t st, double2 *xg, const double *x0, uint32_t m);
HipErr launch_set_x(hipStream_t st, double2 *xg, const double *x0, uint32_t me);
HipErr launch_stamp(hipStream_t st, uint64_t *t, int idx);
HipErr launch_warm_match(hipStream_t st, uint32_t m, const int32_t *cI, const int32_t *cJ,
                         const int32_t *gid, uint32_t mprev, const uint64_t *prev_key,
                         const double *prev_gamma, double2 *xg, const uint32_t *md = nullptr,
                         const uint32_t *mprev_d = nullptr);
// writes the keys of this step's rows; *mprev_d (when set) receives their number
HipErr launch_make_keys(hipStream_t st, uint32_t m, const int32_t *cI, const int32_t *cJ,
                        const int32_t *gid, const double2 *xg, uint64_t *key, double *gam,
                        const uint32_t *md = nullptr, uint32_t *mprev_d = nullptr);

constexpr int kRedQ = 8;  // partial-sum slots per chunk
// values per solver velocity row (padding rows to 8 measured slower than the compact 6)
constexpr int kUStride = 6;
// grid-barrier words of the persistent solver (layout in k_solve.hip)
constexpr int kBarWords = 32;
constexpr int kBarGroups = 8;
constexpr int kBarXcc = 16;  // XCC id slots (the hardware field is 4 bits wide)
// MI355X's last-level (Infinity) cache, which serves data that HBM has already sent once
constexpr double kLastLevelCacheBytes = 256.0 * 1024 * 1024;
constexpr int kBarXccArr = 2 + kBarGroups, kBarXccSize = kBarXccArr + kBarXcc, kBarSetup = kBarXccSize + kBarXcc;
constexpr int kBarTotalWords = (kBarSetup + 1) * kBarWords;

// learned warm start (k_mlp.hip): 12 -> 64 -> 64 -> 64 -> 1; W1 stored 64 x 16 (columns 12..15 zero)
constexpr int kMlpIn = 12;
constexpr int kPW1 = 0, kPB1 = 1024, kPW2 = 1088, kPB2 = 5184, kPW3 = 5248, kPB3 = 9344, kPW4 = 9408,
              kPB4 = 9472;
constexpr int kMlpParams = 9473;
constexpr int kMlpNorm = 2 * kMlpIn + 1;  // mean[12], 1/std[12], gamma scale
struct MlpScratch {
    float *XT, *Z1, *H1, *Z2, *H2, *Z3, *H3, *dO, *dZ1, *dZ2, *dZ3, *loss, *lossv;
    float *kpart;  // K-slice partial products of the weight-gradient GEMMs (kMlpSplitFloats)
};
constexpr size_t kMlpSplitFloats = 256 * 64 * 64;  // at most 256 K-slices of a 64 x 64 product
HipErr launch_features(hipStream_t st, uint32_t m, uint32_t gs, const int32_t *cI, const int32_t *cJ, const double *delta0,
                       const double *geo, const double *posI, const double *posJ, const double *dir,
                       const double *length, const double *unc, const double2 *xprev, double dt, float *X,
                       const uint32_t *md = nullptr);
HipErr launch_mlp_infer_f32(hipStream_t st, uint32_t m, const float *X, const float *P, const float *norm,
                            float *out);
HipErr launch_mlp_infer_xg(hipStream_t st, uint32_t m, const float *X, const float *P, const float *norm,
                           double2 *xg, const uint32_t *md = nullptr);
// part: K-slice partials (ksplit > 1), summed in order into C; NULL = one slice
HipErr launch_gemm_bf16(hipStream_t st, int M, int N, int K, const float *A, int lda, int tA, const float *B,
                        int ldb, int tB, float *C, int ldc, int ksplit, float *part);
HipErr launch_gather_batch(hipStream_t st, int B, const int32_t *idx, const float *X, const float *t, float *XB,
                           float *tB);
HipErr mlp_gradient(hipStream_t st, int B, const float *X, const float *t, const float *P, const float *norm,
                    MlpScratch &s, float *grad);
HipErr mlp_adam_step(hipStream_t st, float *P, const float *grad, float *m1, float *m2, float lr, int step);

}  // namespace alens
