// k_mlp.hip -- learned warm start for the BBPGD solve (gfx950, BF16 MFMA with FP32 accumulation).
//
// Per constraint, 12 inputs (gap delta0, normal n, contact arm positions s_I / (L_I/2),
// s_J / (L_J/2), relative contact-point velocity of U_nc times dt, rod lengths L_I, L_J, previous-step
// gamma) are standardised and fed to 12 -> 64 -> 64 -> 64 -> 1 (SiLU hidden, softplus output);
// gamma0 = gscale * output.  Parameters live in one FP32 master array (layout kP*); the MFMA operands
// are the BF16 roundings of the master values.
//
// MFMA, inference: v_mfma_f32_16x16x16_bf16 (builtin ..._16x16x16bf16_1k).  Lane l holds
// A[l % 16][4 (l / 16) + i], B[4 (l / 16) + i][l % 16] and D[4 (l / 16) + i][l % 16], i = 0..3.
// Constraints run along N, so a layer's output tile (features along rows 4 (l / 16) + i, constraint
// l % 16) is already laid out as the next layer's B operand and stays in registers from layer to layer.
// MFMA, training GEMMs: gfx950's v_mfma_f32_16x16x32_bf16 where K is a multiple of 32 (lane l holds
// A[l % 16][8 (l / 16) + i], B[8 (l / 16) + i][l % 16], i = 0..7, and D as above), else 16x16x16.
#include "alens_dev.hpp"
#include "alens_internal.hpp"

namespace alens {

#define LAUNCH_CHECK()                                                                           \
    do {                                                                                         \
        hipError_t e_ = hipGetLastError();                                                       \
        if (e_ != hipSuccess) return HipErr{e_, "kernel launch", __FILE__, __LINE__};             \
    } while (0)

static inline unsigned nblk(int64_t n, int t) { return (unsigned)((n + t - 1) / t); }

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ inline short f2bf(float f) {
    // round to nearest even
    uint32_t u = __builtin_bit_cast(uint32_t, f);
    if ((u & 0x7f800000u) == 0x7f800000u) return (short)(u >> 16);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (short)(u >> 16);
}

__device__ inline f32x4 mfma16(bf16x4 a, bf16x4 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}
__device__ inline f32x4 mfma32(bf16x8 a, bf16x8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}
__device__ inline __bf16 f2bf16(float f) { return __builtin_bit_cast(__bf16, f2bf(f)); }

__device__ inline float silu(float z) { return z / (1.0f + __expf(-z)); }
__device__ inline float silu_grad(float z) {
    const float s = 1.0f / (1.0f + __expf(-z));
    return s * (1.0f + z * (1.0f - s));
}
__device__ inline float softplus(float o) { return o > 20.0f ? o : log1pf(__expf(o)); }
__device__ inline float sigmoidf(float o) { return 1.0f / (1.0f + __expf(-o)); }

// ------------------------------------------------------------------------------------------------
// raw features, 16 floats per constraint (12 used, 4 zero)
// ------------------------------------------------------------------------------------------------
__global__ void k_features(uint32_t m, uint32_t gs, const int32_t *__restrict__ cI, const int32_t *__restrict__ cJ,
                           const double *__restrict__ delta0, const double *__restrict__ geo,
                           const double *__restrict__ posI, const double *__restrict__ posJ,
                           const double *__restrict__ dir, const double *__restrict__ length,
                           const double *__restrict__ unc, const double2 *__restrict__ xprev, double dt,
                           float *__restrict__ X, const uint32_t *__restrict__ md) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= dcount(m, md)) return;
    const int32_t i = cI[k], j = cJ[k];
    const double n[3] = {geo[k], geo[(size_t)gs + k], geo[2 * (size_t)gs + k]};  // component-major, stride gs
    const double *pi = posI + 3 * (size_t)k, *pj = posJ + 3 * (size_t)k;
    const double *ti = dir + 3 * (size_t)i, *tj = dir + 3 * (size_t)j;
    const double Li = length[i], Lj = length[j];
    const double si = pi[0] * ti[0] + pi[1] * ti[1] + pi[2] * ti[2];
    const double sj = pj[0] * tj[0] + pj[1] * tj[1] + pj[2] * tj[2];
    const double *ui = unc + 6 * (size_t)i, *uj = unc + 6 * (size_t)j;
    double vr[3];
    // v + w x p at each contact point, I minus J
    vr[0] = (ui[0] + (ui[4] * pi[2] - ui[5] * pi[1])) - (uj[0] + (uj[4] * pj[2] - uj[5] * pj[1]));
    vr[1] = (ui[1] + (ui[5] * pi[0] - ui[3] * pi[2])) - (uj[1] + (uj[5] * pj[0] - uj[3] * pj[2]));
    vr[2] = (ui[2] + (ui[3] * pi[1] - ui[4] * pi[0])) - (uj[2] + (uj[3] * pj[1] - uj[4] * pj[0]));
    float *x = X + 16 * (size_t)k;
    x[0] = (float)delta0[k];
    x[1] = (float)n[0];
    x[2] = (float)n[1];
    x[3] = (float)n[2];
    x[4] = (float)(Li > 0.0 ? si / (0.5 * Li) : 0.0);
    x[5] = (float)(Lj > 0.0 ? sj / (0.5 * Lj) : 0.0);
    x[6] = (float)(vr[0] * dt);
    x[7] = (float)(vr[1] * dt);
    x[8] = (float)(vr[2] * dt);
    x[9] = (float)Li;
    x[10] = (float)Lj;
    x[11] = (float)(xprev ? xprev[k].x : 0.0);
    x[12] = 0.0f;
    x[13] = 0.0f;
    x[14] = 0.0f;
    x[15] = 0.0f;
}

HipErr launch_features(hipStream_t st, uint32_t m, uint32_t gs, const int32_t *cI, const int32_t *cJ, const double *delta0,
                       const double *geo, const double *posI, const double *posJ, const double *dir,
                       const double *length, const double *unc, const double2 *xprev, double dt, float *X,
                       const uint32_t *md) {
    if (m == 0) return hip_ok();
    k_features<<<nblk(m, 256), 256, 0, st>>>(m, gs, cI, cJ, delta0, geo, posI, posJ, dir, length, unc, xprev, dt, X,
                                             md);
    LAUNCH_CHECK();
    return hip_ok();
}

// ------------------------------------------------------------------------------------------------
// fused inference: 16 constraints per tile, one wave per tile, all layers in registers
// ------------------------------------------------------------------------------------------------
__device__ inline float norm_feat(const float *X, size_t c, int f, const float *mean, const float *istd) {
    return f < kMlpIn ? (X[16 * c + f] - mean[f]) * istd[f] : 0.0f;
}

template <typename OutT>
__global__ __launch_bounds__(256) void k_mlp_infer(uint32_t m, const float *__restrict__ X,
                                                   const float *__restrict__ P, const float *__restrict__ norm,
                                                   OutT *__restrict__ out, int out_stride,
                                                   const uint32_t *__restrict__ md) {
    m = dcount(m, md);
    const int lane = threadIdx.x & 63;
    const int r16 = lane & 15, q = lane >> 4;
    const float *mean = norm, *istd = norm + kMlpIn;
    const float gscale = norm[2 * kMlpIn];
    // A operands (weights, row = output unit, k = input unit) and biases for this lane
    bf16x4 a1[4], a2[4][4], a3[4][4];
    float b1v[4][4], b2v[4][4], b3v[4][4], w4v[4][4];
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
        const int o = 16 * mt + r16;
#pragma unroll
        for (int i = 0; i < 4; ++i) a1[mt][i] = f2bf(P[kPW1 + o * 16 + 4 * q + i]);
#pragma unroll
        for (int kt = 0; kt < 4; ++kt)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                a2[mt][kt][i] = f2bf(P[kPW2 + o * 64 + 16 * kt + 4 * q + i]);
                a3[mt][kt][i] = f2bf(P[kPW3 + o * 64 + 16 * kt + 4 * q + i]);
            }
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int oo = 16 * mt + 4 * q + i;
            b1v[mt][i] = P[kPB1 + oo];
            b2v[mt][i] = P[kPB2 + oo];
            b3v[mt][i] = P[kPB3 + oo];
            w4v[mt][i] = P[kPW4 + oo];
        }
    }
    const float b4 = P[kPB4];
    const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
    const int64_t ntiles = ((int64_t)m + 15) / 16;
    for (int64_t t = wave; t < ntiles; t += nwaves) {
        const int64_t c = t * 16 + r16;
        const bool ok = c < (int64_t)m;
        bf16x4 bx;
#pragma unroll
        for (int i = 0; i < 4; ++i) bx[i] = f2bf(ok ? norm_feat(X, (size_t)c, 4 * q + i, mean, istd) : 0.0f);
        bf16x4 h[4];
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) {
            f32x4 z = {0.f, 0.f, 0.f, 0.f};
            z = mfma16(a1[mt], bx, z);
#pragma unroll
            for (int i = 0; i < 4; ++i) h[mt][i] = f2bf(silu(z[i] + b1v[mt][i]));
        }
        bf16x4 h2[4];
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) {
            f32x4 z = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int kt = 0; kt < 4; ++kt) z = mfma16(a2[mt][kt], h[kt], z);
#pragma unroll
            for (int i = 0; i < 4; ++i) h2[mt][i] = f2bf(silu(z[i] + b2v[mt][i]));
        }
        float part = 0.0f;
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) {
            f32x4 z = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int kt = 0; kt < 4; ++kt) z = mfma16(a3[mt][kt], h2[kt], z);
#pragma unroll
            for (int i = 0; i < 4; ++i) part += w4v[mt][i] * silu(z[i] + b3v[mt][i]);
        }
        part += __shfl_xor(part, 16, 64);
        part += __shfl_xor(part, 32, 64);
        if (ok && q == 0) out[(size_t)c * out_stride] = (OutT)(gscale * softplus(part + b4));
    }
}

HipErr launch_mlp_infer_f32(hipStream_t st, uint32_t m, const float *X, const float *P, const float *norm,
                            float *out) {
    if (m == 0) return hip_ok();
    const int64_t tiles = ((int64_t)m + 15) / 16;
    int64_t blocks = (tiles + 3) / 4;
    if (blocks > 4096) blocks = 4096;
    k_mlp_infer<float><<<(unsigned)blocks, 256, 0, st>>>(m, X, P, norm, out, 1, nullptr);
    LAUNCH_CHECK();
    return hip_ok();
}

// writes gamma0 into the x component of the solver state (double2 stride 2)
HipErr launch_mlp_infer_xg(hipStream_t st, uint32_t m, const float *X, const float *P, const float *norm,
                           double2 *xg, const uint32_t *md) {
    if (m == 0) return hip_ok();
    const int64_t tiles = ((int64_t)m + 15) / 16;
    int64_t blocks = (tiles + 3) / 4;
    if (blocks > 4096) blocks = 4096;
    k_mlp_infer<double><<<(unsigned)blocks, 256, 0, st>>>(m, X, P, norm, reinterpret_cast<double *>(xg), 2, md);
    LAUNCH_CHECK();
    return hip_ok();
}

// ------------------------------------------------------------------------------------------------
// training: generic BF16 MFMA GEMM  C (M x N) = op(A) (M x K) * op(B) (K x N), FP32 accumulate.
// op(A)(m, k) = tA ? A[k * lda + m] : A[m * lda + k];  op(B)(k, n) = tB ? B[n * ldb + k] : B[k * ldb + n].
// One wave per 16 x 16 tile of C and K-slice.  When the K range is split, slice y writes its tile to
// part[y] and k_sum_slices adds the slices in order (the result does not depend on scheduling).
// M, N, K are multiples of 16; K slices of a multiple of 32 use the 16x16x32 MFMA.
// ------------------------------------------------------------------------------------------------
__global__ __launch_bounds__(64) void k_gemm_bf16(int M, int N, int K, const float *__restrict__ A, int lda, int tA,
                                                  const float *__restrict__ B, int ldb, int tB, float *__restrict__ C,
                                                  int ldc, int kslice, float *__restrict__ part) {
    const int lane = threadIdx.x;
    const int r16 = lane & 15, q = lane >> 4;
    const int tilesN = N / 16;
    const int tm = blockIdx.x / tilesN, tn = blockIdx.x % tilesN;
    const int k0 = blockIdx.y * kslice;
    int k1 = k0 + kslice;
    if (k1 > K) k1 = K;
    const int m = 16 * tm + r16, n = 16 * tn + r16;
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    int k = k0;
    for (; k + 32 <= k1; k += 32) {
        bf16x8 a, b;
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int kk = k + 8 * q + i;
            a[i] = f2bf16(tA ? A[(size_t)kk * lda + m] : A[(size_t)m * lda + kk]);
            b[i] = f2bf16(tB ? B[(size_t)n * ldb + kk] : B[(size_t)kk * ldb + n]);
        }
        acc = mfma32(a, b, acc);
    }
    for (; k < k1; k += 16) {
        bf16x4 a, b;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int kk = k + 4 * q + i;
            a[i] = f2bf(tA ? A[(size_t)kk * lda + m] : A[(size_t)m * lda + kk]);
            b[i] = f2bf(tB ? B[(size_t)n * ldb + kk] : B[(size_t)kk * ldb + n]);
        }
        acc = mfma16(a, b, acc);
    }
    float *dst = part ? part + (size_t)blockIdx.y * M * N : C;
    const int ld = part ? N : ldc;
#pragma unroll
    for (int i = 0; i < 4; ++i) dst[(size_t)(16 * tm + 4 * q + i) * ld + n] = acc[i];
}

// C[r][c] = part[0][r][c] + part[1][r][c] + ... (slices in order)
__global__ void k_sum_slices(int M, int N, int ks, const float *__restrict__ part, float *__restrict__ C, int ldc) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= M * N) return;
    float s = 0.0f;
    for (int y = 0; y < ks; ++y) s += part[(size_t)y * M * N + idx];
    C[(size_t)(idx / N) * ldc + idx % N] = s;
}

// K-slices of a split GEMM: a multiple of 32 at least 32 (the 16x16x32 MFMA), at most kMaxSplit slices
constexpr int kMaxSplit = 256;
static int gemm_kslice(int K, int ksplit) {
    if (ksplit <= 1) return K;
    int kslice = ((K / ksplit + 31) / 32) * 32;
    if (kslice < 32) kslice = 32;
    if ((K + kslice - 1) / kslice > kMaxSplit) kslice = ((K / kMaxSplit + 31) / 32 + 1) * 32;
    return kslice;
}

HipErr launch_gemm_bf16(hipStream_t st, int M, int N, int K, const float *A, int lda, int tA, const float *B,
                        int ldb, int tB, float *C, int ldc, int ksplit, float *part) {
    if (M % 16 || N % 16 || K % 16) return HipErr{hipErrorInvalidValue, "gemm dims", __FILE__, __LINE__};
    const int kslice = part ? gemm_kslice(K, ksplit) : K;
    const int ks = (K + kslice - 1) / kslice;
    dim3 grid((unsigned)((M / 16) * (N / 16)), (unsigned)ks);
    k_gemm_bf16<<<grid, 64, 0, st>>>(M, N, K, A, lda, tA, B, ldb, tB, C, ldc, kslice, ks > 1 ? part : nullptr);
    LAUNCH_CHECK();
    if (ks > 1) {
        k_sum_slices<<<nblk((int64_t)M * N, 256), 256, 0, st>>>(M, N, ks, part, C, ldc);
        LAUNCH_CHECK();
    }
    return hip_ok();
}

// standardised input, feature-major [16][B] (rows >= 12 zero)
__global__ void k_mlp_input(int B, const float *__restrict__ X, const float *__restrict__ norm,
                            float *__restrict__ XT) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= B) return;
#pragma unroll
    for (int f = 0; f < 16; ++f)
        XT[(size_t)f * B + n] = f < kMlpIn ? (X[16 * (size_t)n + f] - norm[f]) * norm[kMlpIn + f] : 0.0f;
}

// Z (rows x B) += bias per row; H = SiLU(Z)
__global__ void k_bias_silu(int rows, int B, float *__restrict__ Z, const float *__restrict__ bias,
                            float *__restrict__ H) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (int64_t)rows * B) return;
    const int r = (int)(idx / B);
    const float z = Z[idx] + bias[r];
    Z[idx] = z;
    H[idx] = silu(z);
}

// output layer: o = w4 . h3 + b4, y = softplus(o); per-sample loss (y - t)^2 / B and
// dO = 2 (y - t) / B * sigmoid(o)
__global__ void k_mlp_head(int B, const float *__restrict__ H3, const float *__restrict__ P,
                           const float *__restrict__ t, float *__restrict__ dO, float *__restrict__ lossv) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= B) return;
    float o = P[kPB4];
    for (int r = 0; r < 64; ++r) o += P[kPW4 + r] * H3[(size_t)r * B + n];
    const float y = softplus(o);
    const float d = y - t[n];
    dO[n] = 2.0f * d / (float)B * sigmoidf(o);
    lossv[n] = d * d / (float)B;
}

// out[r] += sum_n M[r][n] * v[n] (one block per row); v = NULL sums the row
__global__ void k_row_dot(int B, const float *__restrict__ Mx, const float *__restrict__ v, float *__restrict__ out) {
    const int r = blockIdx.x;
    float s = 0.0f;
    for (int n = threadIdx.x; n < B; n += blockDim.x) s += Mx[(size_t)r * B + n] * (v ? v[n] : 1.0f);
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) s += __shfl_xor(s, o, 64);
    __shared__ float sh[4];
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(&out[r], ((sh[0] + sh[1]) + sh[2]) + sh[3]);
}

// dZ3 = (w4 dO) * SiLU'(Z3)
__global__ void k_head_back(int B, const float *__restrict__ P, const float *__restrict__ dO,
                            const float *__restrict__ Z, float *__restrict__ dZ) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (int64_t)64 * B) return;
    const int r = (int)(idx / B), n = (int)(idx % B);
    dZ[idx] = P[kPW4 + r] * dO[n] * silu_grad(Z[idx]);
}

// dZ = dH * SiLU'(Z) in place on dH
__global__ void k_silu_back(int64_t count, const float *__restrict__ Z, float *__restrict__ dH) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx < count) dH[idx] = dH[idx] * silu_grad(Z[idx]);
}

// Adam (Kingma & Ba 2015); W1 columns >= 12 stay zero
__global__ void k_adam(int np, float *__restrict__ P, const float *__restrict__ g, float *__restrict__ m1,
                       float *__restrict__ m2, float lr, float b1, float b2, float eps, float c1, float c2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    if (i < kPB1 && (i % 16) >= kMlpIn) {
        P[i] = 0.0f;
        return;
    }
    const float gi = g[i];
    const float a = b1 * m1[i] + (1.0f - b1) * gi;
    const float b = b2 * m2[i] + (1.0f - b2) * gi * gi;
    m1[i] = a;
    m2[i] = b;
    P[i] -= lr * (a / c1) / (sqrtf(b / c2) + eps);
}

static HipErr fwd_bwd(hipStream_t st, int B, const float *X, const float *t, const float *P, const float *norm,
                      MlpScratch &s, float *grad) {
    const int64_t nb = (int64_t)64 * B;
    k_mlp_input<<<nblk(B, 256), 256, 0, st>>>(B, X, norm, s.XT);
    LAUNCH_CHECK();
    HipErr e;
    // forward (feature-major activations: row = unit, column = sample)
    e = launch_gemm_bf16(st, 64, B, 16, P + kPW1, 16, 0, s.XT, B, 0, s.Z1, B, 1, nullptr);
    if (!e.ok()) return e;
    k_bias_silu<<<nblk(nb, 256), 256, 0, st>>>(64, B, s.Z1, P + kPB1, s.H1);
    LAUNCH_CHECK();
    e = launch_gemm_bf16(st, 64, B, 64, P + kPW2, 64, 0, s.H1, B, 0, s.Z2, B, 1, nullptr);
    if (!e.ok()) return e;
    k_bias_silu<<<nblk(nb, 256), 256, 0, st>>>(64, B, s.Z2, P + kPB2, s.H2);
    LAUNCH_CHECK();
    e = launch_gemm_bf16(st, 64, B, 64, P + kPW3, 64, 0, s.H2, B, 0, s.Z3, B, 1, nullptr);
    if (!e.ok()) return e;
    k_bias_silu<<<nblk(nb, 256), 256, 0, st>>>(64, B, s.Z3, P + kPB3, s.H3);
    LAUNCH_CHECK();
    k_mlp_head<<<nblk(B, 256), 256, 0, st>>>(B, s.H3, P, t, s.dO, s.lossv);
    LAUNCH_CHECK();
    k_row_dot<<<64, 256, 0, st>>>(B, s.H3, s.dO, grad + kPW4);
    LAUNCH_CHECK();
    k_row_dot<<<1, 256, 0, st>>>(B, s.dO, nullptr, grad + kPB4);
    LAUNCH_CHECK();
    k_row_dot<<<1, 256, 0, st>>>(B, s.lossv, nullptr, s.loss);
    LAUNCH_CHECK();
    // backward
    k_head_back<<<nblk(nb, 256), 256, 0, st>>>(B, P, s.dO, s.Z3, s.dZ3);
    LAUNCH_CHECK();
    const int ks = B / 1024 > 1 ? B / 1024 : 1;
    e = launch_gemm_bf16(st, 64, 64, B, s.dZ3, B, 0, s.H2, B, 1, grad + kPW3, 64, ks, s.kpart);
    if (!e.ok()) return e;
    k_row_dot<<<64, 256, 0, st>>>(B, s.dZ3, nullptr, grad + kPB3);
    LAUNCH_CHECK();
    e = launch_gemm_bf16(st, 64, B, 64, P + kPW3, 64, 1, s.dZ3, B, 0, s.dZ2, B, 1, nullptr);
    if (!e.ok()) return e;
    k_silu_back<<<nblk(nb, 256), 256, 0, st>>>(nb, s.Z2, s.dZ2);
    LAUNCH_CHECK();
    e = launch_gemm_bf16(st, 64, 64, B, s.dZ2, B, 0, s.H1, B, 1, grad + kPW2, 64, ks, s.kpart);
    if (!e.ok()) return e;
    k_row_dot<<<64, 256, 0, st>>>(B, s.dZ2, nullptr, grad + kPB2);
    LAUNCH_CHECK();
    e = launch_gemm_bf16(st, 64, B, 64, P + kPW2, 64, 1, s.dZ2, B, 0, s.dZ1, B, 1, nullptr);
    if (!e.ok()) return e;
    k_silu_back<<<nblk(nb, 256), 256, 0, st>>>(nb, s.Z1, s.dZ1);
    LAUNCH_CHECK();
    e = launch_gemm_bf16(st, 64, 16, B, s.dZ1, B, 0, s.XT, B, 1, grad + kPW1, 16, ks, s.kpart);
    if (!e.ok()) return e;
    k_row_dot<<<64, 256, 0, st>>>(B, s.dZ1, nullptr, grad + kPB1);
    LAUNCH_CHECK();
    return hip_ok();
}

// minibatch gather: XB[b] = X[idx[b]], tB[b] = t[idx[b]] (idx < 0 -> zero row, zero weight)
__global__ void k_gather_batch(int B, const int32_t *__restrict__ idx, const float *__restrict__ X,
                               const float *__restrict__ t, float *__restrict__ XB, float *__restrict__ tB) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    const int32_t s = idx[b];
#pragma unroll
    for (int f = 0; f < 16; ++f) XB[16 * (size_t)b + f] = X[16 * (size_t)s + f];
    tB[b] = t[s];
}
HipErr launch_gather_batch(hipStream_t st, int B, const int32_t *idx, const float *X, const float *t, float *XB,
                           float *tB) {
    k_gather_batch<<<nblk(B, 256), 256, 0, st>>>(B, idx, X, t, XB, tB);
    LAUNCH_CHECK();
    return hip_ok();
}

HipErr mlp_gradient(hipStream_t st, int B, const float *X, const float *t, const float *P, const float *norm,
                    MlpScratch &s, float *grad) {
    ALENS_HIP_TRY(hipMemsetAsync(grad, 0, sizeof(float) * kMlpParams, st));
    ALENS_HIP_TRY(hipMemsetAsync(s.loss, 0, sizeof(float), st));
    return fwd_bwd(st, B, X, t, P, norm, s, grad);
}

HipErr mlp_adam_step(hipStream_t st, float *P, const float *grad, float *m1, float *m2, float lr, int step) {
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float c1 = 1.0f - powf(b1, (float)step), c2 = 1.0f - powf(b2, (float)step);
    k_adam<<<nblk(kMlpParams, 256), 256, 0, st>>>(kMlpParams, P, grad, m1, m2, lr, b1, b2, eps, c1, c2);
    LAUNCH_CHECK();
    return hip_ok();
}

}  // namespace alens
