// Helpers shared by the candidate-drawing collision kernels (k_mcc.hip, k_dsmc.hip):
// cross-section look-up, the per-candidate draws, Box-Muller and the id hash that orders
// a cell's parcels.
#pragma once
#include "device.h"
#include "kernels.h"

namespace picmi {

// interpolateXY on an ascending table: clamped linear interpolation between e[lo] <= x <
// e[lo + 1].  With a guide (tables starting above zero) the bracket comes from kXsGuide
// bins equal in log(e), each holding the last entry at or below its lower edge, and a
// step or two along the table; otherwise from a bisection.  Both give the same lo.
__device__ __forceinline__ double xs_lookup(const double* __restrict__ e, const double* __restrict__ s, int n,
                                            const int* __restrict__ guide, double le0, double idl, double x) {
    if (n <= 0) return 0.0;
    if (!(x > e[0])) return s[0];
    if (x >= e[n - 1]) return s[n - 1];
    int lo = 0;
    if (guide) {
        int b = (int)((log_pos(x) - le0) * idl);
        b = b < 0 ? 0 : (b >= kXsGuide ? kXsGuide - 1 : b);
        lo = guide[b];
        while (lo > 0 && e[lo] > x) --lo;
        while (lo + 1 < n && e[lo + 1] <= x) ++lo;
    } else {
        int hi = n - 1;
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (e[mid] <= x) lo = mid;
            else hi = mid;
        }
    }
    if (e[lo] == x) return s[lo];
    return s[lo] + ((x - e[lo]) / (e[lo + 1] - e[lo])) * (s[lo + 1] - s[lo]);
}

// the draws j of candidate k of a cell, keyed by the collision model's tag
__device__ __forceinline__ void mcc_rand(uint32_t tag, uint64_t seed, int64_t step, int cell, int k, int j, double& a,
                                         double& b) {
    uint4 o = rng4(seed, (uint32_t)cell, (uint32_t)step, (uint32_t)k, (tag << 24) | (uint32_t)j);
    a = u01(o.x, o.y);
    b = u01(o.z, o.w);
}

__device__ __forceinline__ void box_muller(double a, double b, double& n1, double& n2) {
    const double r = sqrt(-2.0 * log_pos(1.0 - a));
    double sn, cs;
    sincos_turn(b, &sn, &cs);
    n1 = r * cs;
    n2 = r * sn;
}

// the splitmix64 finaliser of a parcel id: the order of a cell's parcels in the collision
// kernels comes from its top bits
__device__ __forceinline__ uint64_t id_mix(uint32_t id) {
    uint64_t z = (uint64_t)id + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

}  // namespace picmi
