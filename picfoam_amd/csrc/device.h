// Device-side helpers shared by the libpicmi kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "internal.h"

namespace picmi {

template <class T>
struct vec3 {
    T x, y, z;
};
template <class T>
__device__ __forceinline__ vec3<T> mk3(T x, T y, T z) { return {x, y, z}; }
template <class T>
__device__ __forceinline__ vec3<T> operator+(vec3<T> a, vec3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class T>
__device__ __forceinline__ vec3<T> operator-(vec3<T> a, vec3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class T>
__device__ __forceinline__ vec3<T> operator*(T s, vec3<T> a) { return {s * a.x, s * a.y, s * a.z}; }
template <class T>
__device__ __forceinline__ vec3<T> operator/(vec3<T> a, T s) { return {a.x / s, a.y / s, a.z / s}; }
template <class T>
__device__ __forceinline__ T dot(vec3<T> a, vec3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class T>
__device__ __forceinline__ vec3<T> cross(vec3<T> a, vec3<T> b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <class T>
__device__ __forceinline__ T mag(vec3<T> a) { return sqrt(dot(a, a)); }
// 1/sqrt(x) in the precision of x
__device__ __forceinline__ float rsq(float x) { return rsqrtf(x); }
__device__ __forceinline__ double rsq(double x) { return rsqrt(x); }

// FP64 elementary functions for the collision kernels, accurate to a few ulp over the
// arguments those kernels pass, in about a third of the instructions of the general
// library routines (no special-value or denormal-argument handling).  The float overloads
// call the library.
//
// ln 2 and pi/2 split into a head with 32 significant bits and a tail, so n ln2_hi is exact
// for |n| < 2^20
constexpr double kLn2Hi = 0x1.62e42fefp-1, kLn2Lo = 7.440617110012397e-11;
constexpr double kPio2Hi = 0x1.921fb544p+0, kPio2Lo = 6.077100506506192e-11;

// n / d from the hardware reciprocal estimate, two Newton steps and one correction of the
// quotient (d normal, not near the exponent limits)
__device__ __forceinline__ double quot(double n, double d) {
    double r = __builtin_amdgcn_rcp(d);
    r = fma(r, fma(-d, r, 1.0), r);
    r = fma(r, fma(-d, r, 1.0), r);
    const double q = n * r;
    return fma(r, fma(-d, q, n), q);
}
__device__ __forceinline__ float quot(float n, float d) { return n / d; }

// log x, x positive and normal: x = m 2^e with m in [sqrt(1/2), sqrt(2)), log m = 2 atanh z
// with z = (m - 1)/(m + 1), |z| < 0.1716, by the series 2 sum z^(2k+1)/(2k+1) to k = 10
// (next term below 1e-18 relative)
__device__ __forceinline__ double log_pos(double x) {
    double m = __builtin_amdgcn_frexp_mant(x);
    int e = __builtin_amdgcn_frexp_exp(x);
    if (m < 0.70710678118654752) { m *= 2.0; --e; }
    const double z = quot(m - 1.0, m + 1.0);
    const double w = z * z;
    double p = 1.0 / 21.0;
    p = fma(p, w, 1.0 / 19.0); p = fma(p, w, 1.0 / 17.0); p = fma(p, w, 1.0 / 15.0);
    p = fma(p, w, 1.0 / 13.0); p = fma(p, w, 1.0 / 11.0); p = fma(p, w, 1.0 / 9.0);
    p = fma(p, w, 1.0 / 7.0); p = fma(p, w, 1.0 / 5.0); p = fma(p, w, 1.0 / 3.0);
    const double lm = fma(2.0 * z * w, p, 2.0 * z);
    const double de = (double)e;
    return fma(de, kLn2Hi, fma(de, kLn2Lo, lm));
}
__device__ __forceinline__ float log_pos(float x) { return logf(x); }

// exp x for x in [-700, 700]: x = n ln 2 + r, |r| <= ln2 / 2, Taylor series of e^r to r^13
// (remainder below 5e-18 relative), scaled by 2^n
__device__ __forceinline__ double exp_small(double x) {
    const double n = rint(x * 1.4426950408889634);
    const double r = fma(-n, kLn2Lo, fma(-n, kLn2Hi, x));
    double p = 1.0 / 6227020800.0;                        // 1/13!
    p = fma(p, r, 1.0 / 479001600.0); p = fma(p, r, 1.0 / 39916800.0); p = fma(p, r, 1.0 / 3628800.0);
    p = fma(p, r, 1.0 / 362880.0); p = fma(p, r, 1.0 / 40320.0); p = fma(p, r, 1.0 / 5040.0);
    p = fma(p, r, 1.0 / 720.0); p = fma(p, r, 1.0 / 120.0); p = fma(p, r, 1.0 / 24.0);
    p = fma(p, r, 1.0 / 6.0); p = fma(p, r, 0.5); p = fma(p, r, 1.0); p = fma(p, r, 1.0);
    return __builtin_amdgcn_ldexp(p, (int)n);
}
__device__ __forceinline__ float exp_small(float x) { return expf(x); }

// sin and cos of 2 pi u for u in [0, 1]: 4u = q + f with q the nearest integer (the
// quadrant) and |f| <= 1/2 exact, theta = f pi/2 in [-pi/4, pi/4], Taylor series of sin to
// theta^17 and of cos to theta^16 (remainders below 1e-17), rotated by q quarter turns
__device__ __forceinline__ void sincos_turn(double u, double* s, double* c) {
    const double t = 4.0 * u;
    const double q = rint(t);
    const double f = t - q;
    const double th = fma(f, kPio2Hi, f * kPio2Lo);
    const double w = th * th;
    double ps = 1.0 / 355687428096000.0;                  // 1/17!
    ps = fma(ps, -w, 1.0 / 1307674368000.0); ps = fma(ps, -w, 1.0 / 6227020800.0);
    ps = fma(ps, -w, 1.0 / 39916800.0); ps = fma(ps, -w, 1.0 / 362880.0); ps = fma(ps, -w, 1.0 / 5040.0);
    ps = fma(ps, -w, 1.0 / 120.0); ps = fma(ps, -w, 1.0 / 6.0);
    const double sn = fma(th * w, -ps, th);
    double pc = 1.0 / 20922789888000.0;                   // 1/16!
    pc = fma(pc, -w, 1.0 / 87178291200.0); pc = fma(pc, -w, 1.0 / 479001600.0);
    pc = fma(pc, -w, 1.0 / 3628800.0); pc = fma(pc, -w, 1.0 / 40320.0); pc = fma(pc, -w, 1.0 / 720.0);
    pc = fma(pc, -w, 1.0 / 24.0); pc = fma(pc, -w, 0.5);
    const double cs = fma(w, -pc, 1.0);
    const int k = (int)q & 3;
    const double a = (k & 1) ? cs : sn, b = (k & 1) ? sn : cs;
    *s = (k & 2) ? -a : a;
    *c = ((k + 1) & 2) ? -b : b;
}
__device__ __forceinline__ void sincos_turn(float u, float* s, float* c) { sincospif(2.f * u, s, c); }

// Philox4x32-10 (Salmon et al., SC'11): counter (4 x u32), key (2 x u32).
__device__ __forceinline__ uint4 philox4x32(uint4 c, uint2 k) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        uint32_t hi0 = __umulhi(0xD2511F53u, c.x), lo0 = 0xD2511F53u * c.x;
        uint32_t hi1 = __umulhi(0xCD9E8D57u, c.z), lo1 = 0xCD9E8D57u * c.z;
        uint4 n;
        n.x = hi1 ^ c.y ^ k.x;
        n.y = lo1;
        n.z = hi0 ^ c.w ^ k.y;
        n.w = lo0;
        c = n;
        k.x += 0x9E3779B9u;
        k.y += 0xBB67AE85u;
    }
    return c;
}

__device__ __forceinline__ uint4 rng4(uint64_t seed, uint32_t a, uint32_t b, uint32_t c, uint32_t d) {
    return philox4x32(make_uint4(a, b, c, d), make_uint2((uint32_t)seed, (uint32_t)(seed >> 32)));
}

// 53-bit uniform in [0,1) from two words
__device__ __forceinline__ double u01(uint32_t a, uint32_t b) {
    uint64_t x = ((((uint64_t)a) << 32) | b) >> 11;
    return (double)x * (1.0 / 9007199254740992.0);
}

constexpr uint32_t kTagShuffle = 1u, kTagCoulomb = 2u, kTagMcc = 3u, kTagNewId = 4u, kTagInit = 5u, kTagWall = 6u,
                   kTagIonMcc = 7u, kTagIonPair = 8u, kTagDsmc = 9u, kTagDsmcBg = 10u, kTagInject = 11u, kTagReemit = 12u;

// error flags raised by kernels (bitmask in device memory)
enum : uint32_t {
    kFlagSuperluminal = 1u,
    kFlagCellOverflow = 2u,
    kFlagCapacity = 4u,
    kFlagTrackBroken = 8u,
};

// Particle arrays of one species (structure of arrays).
template <class UT>
struct PartSoA {
    double* y1;
    double* y2;
    double* y3;
    UT* ux;
    UT* uy;
    UT* uz;
    int32_t* tet;
    uint32_t* id;
};

// array index of the parcel in sorted slot i: the arrays are in cell order themselves when
// perm is null, otherwise perm holds the order (an index sort between moving sorts)
__device__ __forceinline__ int slot_index(const int* perm, int i) { return perm ? perm[i] : i; }

// Groups the wave's lanes by key (key < 0 lanes take no part) without touching memory:
// each lane learns its group's leader lane and its rank in the group, each leader the group
// size.  Must be called by all lanes of the wave.
__device__ __forceinline__ void wave_group(int key, int& leader, int& rank, int& cnt) {
    const int lane = __lane_id();
    const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    uint64_t active = __ballot(key >= 0);
    leader = lane; rank = 0; cnt = 0;
    while (active) {
        const int l = __ffsll((unsigned long long)active) - 1;
        const int lk = __shfl(key, l);
        const uint64_t mask = __ballot(key == lk) & active;
        if ((mask >> lane) & 1ull) {
            leader = l;
            rank = __popcll(mask & below);
        }
        if (lane == l) cnt = __popcll(mask);
        active &= ~mask;
    }
}

// Wave-aggregated atomic increment: lanes with equal keys (neighbouring lanes usually
// share a cell because parcels are kept sorted by cell) issue one atomic per distinct key.
// Returns this lane's slot (old value + rank among the equal-key lanes).  key < 0 lanes
// take no part.  Must be called by all lanes of the wave.
__device__ __forceinline__ int wave_aggregated_inc(int* counter, int key) {
    const int lane = __lane_id();
    int leader, rank, cnt;
    wave_group(key, leader, rank, cnt);
    // then every leader adds its group size at once: one atomic round trip per wave
    int base = 0;
    if (key >= 0 && lane == leader) base = atomicAdd(counter + key, cnt);
    base = __shfl(base, leader);
    return key >= 0 ? base + rank : -1;
}

// Lorentz transformation of two parcels to their centre-of-momentum frame and back, the
// common part of PerezRelativisticCollision.C:84-196 and
// RelativisticElectronNeutralCollision.C:84-158 (FP64)
struct ComFrame {
    double gP, gQ, msum, gc, gcs, gPs, gQs, pm;
    vec3<double> vc, pPs;
};

__device__ __forceinline__ void com_enter(vec3<double> uP, vec3<double> uQ, double mP, double mQ, ComFrame& f) {
    const double c2 = kC * kC;
    const double bP = mag(uP) / kC, bQ = mag(uQ) / kC;
    f.gP = 1.0 / sqrt(1.0 - bP * bP);
    f.gQ = 1.0 / sqrt(1.0 - bQ * bQ);
    const vec3<double> pP = (mP * f.gP) * uP, pQ = (mQ * f.gQ) * uQ;
    f.msum = mP * f.gP + mQ * f.gQ;
    f.vc = (1.0 / f.msum) * (pP + pQ);
    const double vc2 = dot(f.vc, f.vc);
    if (vc2 == 0.0) {
        f.gc = 1.0; f.gcs = 0.5; f.pPs = pP; f.gPs = f.gP; f.gQs = f.gQ;
    } else {
        f.gc = 1.0 / sqrt(1.0 - vc2 / c2);
        f.gcs = (f.gc - 1.0) / vc2;
        f.pPs = pP + ((f.gcs * dot(f.vc, uP) - f.gc) * mP * f.gP) * f.vc;
        f.gPs = (1.0 - dot(f.vc, uP) / c2) * f.gP * f.gc;
        f.gQs = (1.0 - dot(f.vc, uQ) / c2) * f.gQ * f.gc;
    }
    f.pm = mag(f.pPs);
}

// rotate P's centre-of-momentum momentum by chi (cos cc, sin sc) and azimuth (sin se,
// cos ce), then transform both parcels back
__device__ __forceinline__ void com_leave(const ComFrame& f, double mP, double mQ, double cc, double sc, double se,
                                          double ce, vec3<double>& UP, vec3<double>& UQ) {
    const double sce = sc * ce, scs = sc * se;
    const vec3<double> p = f.pPs;
    const double pm = f.pm;
    vec3<double> np;
    const double pperp = sqrt(p.x * p.x + p.y * p.y);
    if (pperp > 1.e-10 * pm) {
        const double ip = 1.0 / pperp;
        np = mk3<double>((p.x * p.z * sce - p.y * pm * scs) * ip + p.x * cc,
                         (p.y * p.z * sce + p.x * pm * scs) * ip + p.y * cc, -pperp * sce + p.z * cc);
    } else {
        np = mk3<double>(pm * sce, pm * scs, pm * cc);
    }
    const vec3<double> newP = np + (f.gcs * dot(f.vc, np) + mP * f.gPs * f.gc) * f.vc;
    const double aP = mag(newP) / (mP * kC);
    const vec3<double> mnp = -1.0 * np;
    const vec3<double> newQ = mnp + (f.gcs * dot(f.vc, mnp) + mQ * f.gQs * f.gc) * f.vc;
    const double aQ = mag(newQ) / (mQ * kC);
    UP = (1.0 / (mP * sqrt(1.0 + aP * aP))) * newP;
    UQ = (1.0 / (mQ * sqrt(1.0 + aQ * aQ))) * newQ;
}

}  // namespace picmi
