// alens_dev.hpp -- device arithmetic of the collision-constraint path (gfx950).
//
// Every floating-point expression here is evaluated in the order written with no FMA contraction
// (the library is compiled with -ffp-contract=off); sqrt and division are correctly rounded on
// gfx950.  The log / sin / cos used by the drag law and the Brownian draw are polynomial routines of
// this library, so results do not depend on the device math library.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace alens {

constexpr double kPi = 3.14159265358979323846;
constexpr double kEpsPar = 1e-30;      // |tI x tJ|^2 <= kEpsPar * a * c  =>  parallel segments
constexpr double kAabbMargin = 1e-6;   // um added to AABB half-widths (rounding safety)
constexpr int kRedThreads = 256;       // reduction lanes per block (4 waves of 64)
constexpr int kRedChunk = 512;         // elements per reduction chunk (2 passes of 256 lanes)
constexpr int kRedPasses = kRedChunk / kRedThreads;

// ---------------------------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11)
// ---------------------------------------------------------------------------------------------
__host__ __device__ inline void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
    uint32_t c0 = c[0], c1 = c[1], c2 = c[2], c3 = c[3];
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        if (r > 0) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c1 = (uint32_t)p1;
        c3 = (uint32_t)p0;
        c0 = n0;
        c2 = n2;
    }
    c[0] = c0;
    c[1] = c1;
    c[2] = c2;
    c[3] = c3;
}

// natural log via m in [sqrt(1/2), sqrt(2)], 2 atanh((m-1)/(m+1)) series to z^25, ln2 head/tail
__host__ __device__ inline double alog(double x) {
    uint64_t bits = __builtin_bit_cast(uint64_t, x);
    int e = (int)((bits >> 52) & 0x7ff);
    int adj = 0;
    if (e == 0) {
        const double y = x * 18014398509481984.0;
        bits = __builtin_bit_cast(uint64_t, y);
        e = (int)((bits >> 52) & 0x7ff);
        adj = -54;
    }
    e = e - 1023 + adj;
    double m = __builtin_bit_cast(double, (bits & 0x000FFFFFFFFFFFFFull) | 0x3FF0000000000000ull);
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
    const double ln2_hi = 6.93147180369123816490e-01;
    const double ln2_lo = 1.90821492927058770002e-10;
    return de * ln2_hi + (lm + de * ln2_lo);
}

// sin(2 pi u), cos(2 pi u) via quadrant reduction and Taylor polynomials of degree 17 / 18
__host__ __device__ inline void asincos2pi(double u, double &s_out, double &c_out) {
    const double u4 = 4.0 * u;
    const double qf = floor(u4 + 0.5);
    const double f = u4 - qf;
    const double th = f * 1.57079632679489661923;
    const double t2 = th * th;
    double ps = 1.0 / 355687428096000.0;
    ps = -1.0 / 1307674368000.0 + t2 * ps;
    ps = 1.0 / 6227020800.0 + t2 * ps;
    ps = -1.0 / 39916800.0 + t2 * ps;
    ps = 1.0 / 362880.0 + t2 * ps;
    ps = -1.0 / 5040.0 + t2 * ps;
    ps = 1.0 / 120.0 + t2 * ps;
    ps = -1.0 / 6.0 + t2 * ps;
    const double sn = th + th * (t2 * ps);
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
        s_out = sn;
        c_out = cs;
    } else if (q == 1) {
        s_out = cs;
        c_out = -sn;
    } else if (q == 2) {
        s_out = -sn;
        c_out = -cs;
    } else {
        s_out = -cs;
        c_out = sn;
    }
}

__host__ __device__ inline double clampd(double x, double lo, double hi) {
    return x < lo ? lo : (x > hi ? hi : x);
}

// ---------------------------------------------------------------------------------------------
// Segment-segment closest points (clamped alternating solution of the 2x2 normal equations,
// Ericson, Real-Time Collision Detection 5.1.9; det and numerator via the Lagrange identity).
// Segment I: cI + s tI, |s| <= hI; segment J: cJ + u tJ, |u| <= hJ (cJ at the image).
// ---------------------------------------------------------------------------------------------
struct SegRes {
    double s, u, dist, D0, D1, D2;
};

__host__ __device__ inline SegRes segseg(double cI0, double cI1, double cI2, double tI0, double tI1,
                                         double tI2, double hI, double cJ0, double cJ1, double cJ2,
                                         double tJ0, double tJ1, double tJ2, double hJ) {
    const double r0 = cI0 - cJ0, r1 = cI1 - cJ1, r2 = cI2 - cJ2;
    const double a = tI0 * tI0 + tI1 * tI1 + tI2 * tI2;
    const double b = tI0 * tJ0 + tI1 * tJ1 + tI2 * tJ2;
    const double c = tJ0 * tJ0 + tJ1 * tJ1 + tJ2 * tJ2;
    const double d = tI0 * r0 + tI1 * r1 + tI2 * r2;
    const double e = tJ0 * r0 + tJ1 * r1 + tJ2 * r2;
    const double w0 = tI1 * tJ2 - tI2 * tJ1;
    const double w1 = tI2 * tJ0 - tI0 * tJ2;
    const double w2 = tI0 * tJ1 - tI1 * tJ0;
    const double v0 = tJ1 * r2 - tJ2 * r1;
    const double v1 = tJ2 * r0 - tJ0 * r2;
    const double v2 = tJ0 * r1 - tJ1 * r0;
    const double det = w0 * w0 + w1 * w1 + w2 * w2;
    double s;
    if (det > kEpsPar * (a * c)) {
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
    SegRes o;
    o.D0 = (r0 + s * tI0) - u * tJ0;
    o.D1 = (r1 + s * tI1) - u * tJ1;
    o.D2 = (r2 + s * tI2) - u * tJ2;
    o.dist = sqrt(o.D0 * o.D0 + o.D1 * o.D1 + o.D2 * o.D2);
    o.s = s;
    o.u = u;
    return o;
}

// unit normal from J to I; zero distance falls back to tI x tJ, then tI x (least-aligned axis)
__host__ __device__ inline void contact_normal(const SegRes &g, double tI0, double tI1, double tI2,
                                               double tJ0, double tJ1, double tJ2, double n[3]) {
    if (g.dist > 0.0) {
        n[0] = g.D0 / g.dist;
        n[1] = g.D1 / g.dist;
        n[2] = g.D2 / g.dist;
        return;
    }
    double v0 = tI1 * tJ2 - tI2 * tJ1;
    double v1 = tI2 * tJ0 - tI0 * tJ2;
    double v2 = tI0 * tJ1 - tI1 * tJ0;
    double vn = sqrt(v0 * v0 + v1 * v1 + v2 * v2);
    if (!(vn > 0.0)) {
        const double ax = fabs(tI0), ay = fabs(tI1), az = fabs(tI2);
        if (ax <= ay && ax <= az) {
            v0 = 0.0;
            v1 = tI2;
            v2 = -tI1;
        } else if (ay <= az) {
            v0 = -tI2;
            v1 = 0.0;
            v2 = tI0;
        } else {
            v0 = tI1;
            v1 = -tI0;
            v2 = 0.0;
        }
        vn = sqrt(v0 * v0 + v1 * v1 + v2 * v2);
    }
    n[0] = v0 / vn;
    n[1] = v1 / vn;
    n[2] = v2 / vn;
}

// ---------------------------------------------------------------------------------------------
// Mobility: per rod {t0, t1, t2, 1/zperp, 1/zpar - 1/zperp, 1/zrot}
// ---------------------------------------------------------------------------------------------
__host__ __device__ inline void drag_inv(double L, double r, double mu, double &ipar, double &iperp,
                                         double &irot) {
    double zpar, zperp, zrot;
    if (L < r * 2.0) {
        zpar = 6.0 * kPi * mu * r;
        zperp = zpar;
        zrot = 8.0 * kPi * mu * r * r * r;
    } else {
        const double bb = -(1.0 + 2.0 * alog(r / L));
        zpar = 8.0 * kPi * mu * L / (2.0 * bb);
        zperp = 8.0 * kPi * mu * L / (bb + 2.0);
        zrot = 2.0 * kPi * mu * L * L * L / (3.0 * (bb + 2.0));
    }
    ipar = 1.0 / zpar;
    iperp = 1.0 / zperp;
    irot = 1.0 / zrot;
}

template <typename T>
__host__ __device__ inline void apply_M(const T *mb, const T F[6], int mono, T U[6]) {
    T f0 = F[0], f1 = F[1], f2 = F[2], q0 = F[3], q1 = F[4], q2 = F[5];
    if (mono) {
        f2 = T(0);
        q0 = T(0);
        q1 = T(0);
    }
    const T tf = mb[0] * f0 + mb[1] * f1 + mb[2] * f2;
    const T c = mb[4] * tf;
    U[0] = mb[3] * f0 + c * mb[0];
    U[1] = mb[3] * f1 + c * mb[1];
    U[2] = mb[3] * f2 + c * mb[2];
    U[3] = mb[5] * q0;
    U[4] = mb[5] * q1;
    U[5] = mb[5] * q2;
    if (mono) {
        U[2] = T(0);
        U[3] = T(0);
        U[4] = T(0);
    }
}

// (D^T U)_k = (n.vI + rI.wI) - (n.vJ + rJ.wJ); geo = {n[3], rI[3], rJ[3]}
template <typename T>
__host__ __device__ inline T dt_row(const T *geo, const T *uI, const T *uJ) {
    const T tI = geo[0] * uI[0] + geo[1] * uI[1] + geo[2] * uI[2];
    const T aI = geo[3] * uI[3] + geo[4] * uI[4] + geo[5] * uI[5];
    const T tJ = geo[0] * uJ[0] + geo[1] * uJ[1] + geo[2] * uJ[2];
    const T aJ = geo[6] * uJ[3] + geo[7] * uJ[4] + geo[8] * uJ[5];
    return (tI + aI) - (tJ + aJ);
}

template <typename T>
__host__ __device__ inline T proj(T v, T lb) { return v < lb ? lb : v; }

// image shift packing: 3 signed 8-bit periods
__host__ __device__ inline uint32_t img_pack(int mx, int my, int mz) {
    return (uint32_t)(mx + 128) | ((uint32_t)(my + 128) << 8) | ((uint32_t)(mz + 128) << 16);
}
__host__ __device__ inline int img_get(uint32_t p, int k) { return (int)((p >> (8 * k)) & 0xff) - 128; }

}  // namespace alens
