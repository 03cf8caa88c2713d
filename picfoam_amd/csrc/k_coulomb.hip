// Binary Coulomb collisions: Nanbu scattering (NanbuCollision.C:60-159) with
// Takizuka-Abe pairing (TakizukaAbePairing.C:60-338).  One wave per cell; parcels are
// contiguous per species because the arrays are sorted by cell.  The reference's
// shuffle (OpenFOAM Random::shuffle) becomes a sort of the cell's parcels by a
// counter-based random key of (id, step, stage, species), so the pairing does not depend
// on the order the parcels sit in memory; every collision draws its two uniforms from a
// counter (cell, step, stage, collision index).
#include "device.h"
#include "kernels.h"

namespace picmi {

namespace {

// NanbuCollision.C:137-159, in the arithmetic M of the collision (FP64, or FP32 in the
// mixed-precision mode)
template <class M>
__device__ M cos_chi(M s, M U) {
    M A, invA;
    if (s < (M)0.1) {
        if (U < (M)0.0001) U = (M)0.0001;
        return (M)1 + s * log(U);
    }
    if (s < (M)3) {
        invA = (M)0.00569578 +
               ((M)0.95602 + ((M)-0.508139 + ((M)0.479139 + ((M)-0.12789 + (M)0.0238957 * s) * s) * s) * s) * s;
        A = (M)1 / invA;
        return invA * log(exp(-A) + (M)2 * U * sinh(A));
    }
    if (s < (M)6) {
        A = (M)3 * exp(-s);
        return ((M)1 / A) * log(exp(-A) + (M)2 * U * sinh(A));
    }
    return (M)2 * U - (M)1;
}

__device__ __forceinline__ void sincos_m(double x, double* s, double* c) { sincos(x, s, c); }
__device__ __forceinline__ void sincos_m(float x, float* s, float* c) { sincosf(x, s, c); }

template <class UT>
struct Vel {
    UT* x;
    UT* y;
    UT* z;
};

// NanbuCollision::collide (weight correction "none"); arithmetic in M = the velocity type
template <class UT>
__device__ void nanbu_collide(Vel<UT> vp, int64_t ip, Vel<UT> vq, int64_t iq, double mP, double qP, double mQ,
                              double qQ, double lnL_fixed, int calc_bmin, double debye, double nP, double nQ,
                              double nPQ, double dt, double u1d, double u2d) {
    using M = UT;
    const M reMass = (M)(mP * mQ / (mP + mQ));
    vec3<M> uP = mk3<M>(vp.x[ip], vp.y[ip], vp.z[ip]);
    vec3<M> uQ = mk3<M>(vq.x[iq], vq.y[iq], vq.z[iq]);
    vec3<M> g = uP - uQ;
    const M gm = mag(g);
    const M gp = sqrt(g.y * g.y + g.z * g.z);
    if (!(gm > (M)0) || !(gp > (M)0)) return;
    M lnL;
    if (lnL_fixed == 0.0) {
        M b0 = (M)(qP * qQ / (2.0 * kPi * kEps0)) / (reMass * gm * gm);
        if (calc_bmin) {
            M hb = (M)kH / ((M)2 * (gm * reMass));
            b0 = hb > fabs(b0) ? hb : fabs(b0);
        }
        M l = (M)0.5 * log((M)1 + ((M)debye * (M)debye) / (b0 * b0));
        lnL = l > (M)2 ? l : (M)2;
    } else {
        lnL = (M)lnL_fixed;
    }
    const double ad = qP * qQ / (kEps0 * (mP * mQ / (mP + mQ)));
    // s = lnL/(4 pi) (qP qQ / eps0 mu)^2 (nP nQ / nPQ) g^-3 dt, the constant part in FP64
    const M cst = (M)(1.0 / (4.0 * kPi) * (ad * ad) * (nQ * nP / nPQ) * dt);
    const M s = lnL * cst * ((M)1 / (gm * gm * gm));
    const M cc = cos_chi<M>(s, (M)u1d);
    const M sc = sqrt((M)1 - cc * cc);
    const M eps = (M)(2.0 * kPi) * (M)u2d;
    M se, ce;
    sincos_m(eps, &se, &ce);
    const M igp = (M)1 / gp;
    vec3<M> h = mk3<M>(gp * ce, -(g.y * g.x * ce + gm * g.z * se) * igp, -(g.z * g.x * ce - gm * g.y * se) * igp);
    vec3<M> dv = ((M)1 - cc) * g + sc * h;
    uP = uP - (M)(mQ / (mP + mQ)) * dv;
    uQ = uQ + (M)(mP / (mP + mQ)) * dv;
    vp.x[ip] = uP.x; vp.y[ip] = uP.y; vp.z[ip] = uP.z;
    vq.x[iq] = uQ.x; vq.y[iq] = uQ.y; vq.z[iq] = uQ.z;
}

__device__ __forceinline__ void coll_rand(uint64_t seed, int64_t step, int cell, int stage, int k, double& u1,
                                          double& u2) {
    uint4 o = rng4(seed, (uint32_t)cell, (uint32_t)step, ((uint32_t)stage << 24) | (uint32_t)k, kTagCoulomb << 24);
    u1 = u01(o.x, o.y);
    u2 = u01(o.z, o.w);
}

// bitonic sort of npow2 64-bit keys with int payload in LDS by one wave
__device__ void bitonic_sort(uint64_t* key, int* val, int npow2) {
    for (int k = 2; k <= npow2; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = threadIdx.x; i < npow2; i += blockDim.x) {
                const int l = i ^ j;
                if (l > i) {
                    const bool up = (i & k) == 0;
                    const uint64_t a = key[i], b = key[l];
                    if ((a > b) == up) {
                        key[i] = b; key[l] = a;
                        const int t = val[i]; val[i] = val[l]; val[l] = t;
                    }
                }
            }
            __syncthreads();
        }
    }
}

// the cell's parcels of one species ordered by the random key of (id, step, stage,
// species).  Up to 256 parcels the order is found by ranking (each key counts the keys
// below it, no barriers inside); larger cells use a bitonic sort.
__device__ void keyed_perm(uint64_t seed, int64_t step, int stage, int species, const uint32_t* id, int first, int n,
                           uint64_t* key, int* val) {
    if (n <= 256) {
        for (int k = threadIdx.x; k < n; k += blockDim.x) {
            const uint32_t pid = id[first + k];
            uint4 o = rng4(seed, pid, (uint32_t)step, ((uint32_t)stage << 8) | (uint32_t)species, kTagShuffle << 24);
            key[k] = ((uint64_t)o.x << 32) | pid;
        }
        __syncthreads();
        for (int k = threadIdx.x; k < n; k += blockDim.x) {
            const uint64_t kk = key[k];
            int r = 0;
            for (int j = 0; j < n; ++j) {
                const uint64_t kj = key[j];
                r += (kj < kk) || (kj == kk && j < k);
            }
            val[r] = first + k;
        }
        __syncthreads();
        return;
    }
    int np2 = 1;
    while (np2 < n) np2 <<= 1;
    for (int k = threadIdx.x; k < np2; k += blockDim.x) {
        if (k < n) {
            const uint32_t pid = id[first + k];
            uint4 o = rng4(seed, pid, (uint32_t)step, ((uint32_t)stage << 8) | (uint32_t)species, kTagShuffle << 24);
            key[k] = ((uint64_t)o.x << 32) | pid;
            val[k] = first + k;
        } else {
            key[k] = ~0ull;
            val[k] = -1;
        }
    }
    __syncthreads();
    bitonic_sort(key, val, np2);
}

__device__ double wave_sum(double v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

}  // namespace

template <class UT>
__global__ void __launch_bounds__(64) k_coulomb(CoulombArgs a) {
    extern __shared__ uint64_t lds[];
    const int cell = blockIdx.x;
    const int maxp = a.maxp;
    uint64_t* keyA = lds;
    uint64_t* keyB = lds + maxp;
    int* valA = (int*)(lds + 2 * maxp);
    int* valB = valA + maxp;
    const int ns = a.n_species;
    for (int s = 0; s < ns; ++s) {
        if (a.charge[s] == 0.0) continue;
        if (a.start[s][cell + 1] - a.start[s][cell] > maxp) {
            if (threadIdx.x == 0) atomicOr(a.flags, kFlagCellOverflow);
            return;
        }
    }
    const double Vc = a.V[cell];
    // CoulombCollisionModel::debyeLength (CoulombCollisionModel.C:249-310), parcels summed
    // in ascending id order
    double debye = 0.0;
    if (a.calc_debye) {
        double deb = 0.0, nmax = 0.0;
        for (int s = 0; s < ns; ++s) {
            if (a.charge[s] == 0.0) continue;
            const int f0 = a.start[s][cell], n = a.start[s][cell + 1] - f0;
            if (n <= 0) continue;
            int np2 = 1;
            while (np2 < n) np2 <<= 1;
            for (int k = threadIdx.x; k < np2; k += blockDim.x) {
                keyA[k] = (k < n) ? (((uint64_t)a.id[s][f0 + k] << 32) | (uint32_t)k) : ~0ull;
                valA[k] = (k < n) ? f0 + k : -1;
            }
            __syncthreads();
            bitonic_sort(keyA, valA, np2);
            const UT* ux = (const UT*)a.ux[s];
            const UT* uy = (const UT*)a.uy[s];
            const UT* uz = (const UT*)a.uz[s];
            const double w = a.weight[s];
            double vs = 0.0, vx = 0.0, vy = 0.0, vz = 0.0;
            for (int k = threadIdx.x; k < n; k += blockDim.x) {
                const int i = valA[k];
                const double x = ux[i], y = uy[i], z = uz[i];
                vs += (x * x + y * y + z * z) * w;
                vx += w * x; vy += w * y; vz += w * z;
            }
            vs = wave_sum(vs); vx = wave_sum(vx); vy = wave_sum(vy); vz = wave_sum(vz);
            double nn = n * w;
            double ch = a.charge[s];
            double mvx = vx / nn, mvy = vy / nn, mvz = vz / nn;
            double mvs = vs / nn;
            nn /= Vc;
            if (nn > nmax) nmax = nn;
            double T = a.mass[s] / (3.0 * kKB) * (mvs - (mvx * mvx + mvy * mvy + mvz * mvz));
            if (T > 0.0) deb += nn * ch * ch / kEps0 * 1.0 / (T * kKB);
            __syncthreads();
        }
        if (deb > 0.0) {
            double rmin = pow(4.0 * kPi * nmax / 3.0, -1.0 / 3.0);
            double d = sqrt(1.0 / deb);
            debye = d > rmin ? d : rmin;
        }
    }
    // stages in the reference order: species i ascending, partner j descending from the
    // last charged species down to i.  An intra-species stage (i, i) is followed by the
    // intra stage of the last charged species when that is the next charged species; the
    // two touch disjoint parcels and run together on one set of lanes.
    int last = -1;
    for (int s = 0; s < ns; ++s)
        if (a.charge[s] != 0.0) last = s;
    unsigned long long ncoll = 0;
    int st = 0;
    bool skip_last_intra = false;
    for (int i = 0; i < ns; ++i) {
        if (a.charge[i] == 0.0) continue;
        for (int j = last; j >= i; --j) {
            if (a.charge[j] == 0.0) continue;
            const int stage_now = st++;
            if (i == j && i == last && skip_last_intra) continue;
            const int a0 = a.start[i][cell], na = a.start[i][cell + 1] - a0;
            const int b0 = a.start[j][cell], nb = a.start[j][cell + 1] - b0;
            const double mA = a.mass[i], qA = a.charge[i], mB = a.mass[j], qB = a.charge[j];
            const double lnL = a.coulomb_log[i][j];
            const double w = a.weight[i];
            Vel<UT> VA{(UT*)a.ux[i], (UT*)a.uy[i], (UT*)a.uz[i]};
            Vel<UT> VB{(UT*)a.ux[j], (UT*)a.uy[j], (UT*)a.uz[j]};
            double u1, u2;
            if (i == j) {
                // this intra stage, and the last species' intra stage when it comes next
                int nxt = -1;
                for (int s2 = i + 1; s2 < ns; ++s2)
                    if (a.charge[s2] != 0.0) { nxt = s2; break; }
                const bool pair2 = (nxt >= 0) && (nxt == last);
                if (pair2) skip_last_intra = true;
                int sp[2] = {i, pair2 ? last : -1};
                int sg[2] = {stage_now, stage_now + 1};
                int nn[2] = {0, 0}, items[2] = {0, 0};
                for (int u = 0; u < (pair2 ? 2 : 1); ++u) {
                    const int spc = sp[u];
                    const int f0 = a.start[spc][cell];
                    const int cnt = a.start[spc][cell + 1] - f0;
                    if (cnt <= 1 || !a.allow_intra) continue;
                    nn[u] = cnt;
                    keyed_perm(a.seed, a.step, sg[u], spc, a.id[spc], f0, cnt, u == 0 ? keyA : keyB, u == 0 ? valA : valB);
                    items[u] = (cnt % 2) ? 1 + (cnt - 3) / 2 : cnt / 2;
                    ncoll += (cnt % 2) ? 3 + (cnt - 3) / 2 : cnt / 2;
                }
                const int total = items[0] + items[1];
                for (int t = threadIdx.x; t < total; t += blockDim.x) {
                    const int u = t < items[0] ? 0 : 1;
                    const int it = u == 0 ? t : t - items[0];
                    const int spc = sp[u], cnt = nn[u], stg = sg[u];
                    const int* v = u == 0 ? valA : valB;
                    Vel<UT> Vs{(UT*)a.ux[spc], (UT*)a.uy[spc], (UT*)a.uz[spc]};
                    const double ms = a.mass[spc], qs = a.charge[spc];
                    const double lnLs = a.coulomb_log[spc][spc];
                    const double n = cnt * a.weight[spc] / Vc;
                    if (cnt % 2) {
                        const double n12 = (cnt + 3.0) * a.weight[spc] / Vc;
                        if (it == 0) {
                            // odd count: the first three collide cyclically (TakizukaAbePairing.C:115-139)
                            const int P = v[0], Q = v[1], R = v[2];
                            coll_rand(a.seed, a.step, cell, stg, 0, u1, u2);
                            nanbu_collide(Vs, P, Vs, Q, ms, qs, ms, qs, lnLs, a.calculate_bmin, debye, n, n, n12, a.dt, u1, u2);
                            coll_rand(a.seed, a.step, cell, stg, 1, u1, u2);
                            nanbu_collide(Vs, Q, Vs, R, ms, qs, ms, qs, lnLs, a.calculate_bmin, debye, n, n, n12, a.dt, u1, u2);
                            coll_rand(a.seed, a.step, cell, stg, 2, u1, u2);
                            nanbu_collide(Vs, P, Vs, R, ms, qs, ms, qs, lnLs, a.calculate_bmin, debye, n, n, n12, a.dt, u1, u2);
                        } else {
                            const int q = it - 1, x = 3 + 2 * q;
                            coll_rand(a.seed, a.step, cell, stg, 3 + q, u1, u2);
                            nanbu_collide(Vs, v[x], Vs, v[x + 1], ms, qs, ms, qs, lnLs, a.calculate_bmin, debye, n, n,
                                          n12, a.dt, u1, u2);
                        }
                    } else {
                        coll_rand(a.seed, a.step, cell, stg, it, u1, u2);
                        nanbu_collide(Vs, v[2 * it], Vs, v[2 * it + 1], ms, qs, ms, qs, lnLs, a.calculate_bmin, debye, n,
                                      n, n, a.dt, u1, u2);
                    }
                }
            } else {
                if (na == 0 || nb == 0) continue;
                keyed_perm(a.seed, a.step, stage_now, i, a.id[i], a0, na, keyA, valA);
                keyed_perm(a.seed, a.step, stage_now, j, a.id[j], b0, nb, keyB, valB);
                if (na == nb) {
                    const double nP = na * w / Vc, nQ = nb * w / Vc;
                    for (int k = threadIdx.x; k < na; k += blockDim.x) {
                        coll_rand(a.seed, a.step, cell, stage_now, k, u1, u2);
                        nanbu_collide(VA, valA[k], VB, valB[na - k - 1], mA, qA, mB, qB, lnL, a.calculate_bmin, debye,
                                      nP, nQ, nP, a.dt, u1, u2);
                    }
                    ncoll += na;
                } else {
                    const bool minIsA = na < nb;
                    const int nMin = minIsA ? na : nb, nMax = minIsA ? nb : na;
                    const int* pmin = minIsA ? valA : valB;
                    const int* pmax = minIsA ? valB : valA;
                    const int smin = minIsA ? i : j, smax = minIsA ? j : i;
                    Vel<UT> Vmin = minIsA ? VA : VB, Vmax = minIsA ? VB : VA;
                    const double nMinS = nMin * w / Vc, nMaxS = nMax * w / Vc;
                    const int delta = nMax / nMin;
                    const double delta_r = (double)(nMax % nMin) / nMin;
                    const int nMinG1 = (int)(delta_r * nMin);
                    const int nMaxG1 = (int)((delta + 1) * delta_r * nMin);
                    const int nMinG2 = (int)((1.0 - delta_r) * nMin);
                    const int nMaxG2 = (int)(delta * (1.0 - delta_r) * nMin);
                    const double mMin = a.mass[smin], qMin = a.charge[smin], mMax = a.mass[smax], qMax = a.charge[smax];
                    const double lnLp = a.coulomb_log[smax][smin];
                    // group 1: min parcel m collides with max parcels m, m+nMinG1, ... (in order)
                    if (nMinG1 > 0)
                        for (int m = threadIdx.x; m < nMinG1; m += blockDim.x)
                            for (int g1 = m; g1 < nMaxG1; g1 += nMinG1) {
                                coll_rand(a.seed, a.step, cell, stage_now, g1, u1, u2);
                                nanbu_collide(Vmax, pmax[g1], Vmin, pmin[m], mMax, qMax, mMin, qMin, lnLp,
                                              a.calculate_bmin, debye, nMaxS, nMinS, nMaxS, a.dt, u1, u2);
                            }
                    // group 2: min parcels nMinG1.. cycle over max parcels nMaxG1..
                    const int cyc = nMinG2 > 0 ? nMinG2 : 1;
                    for (int m = threadIdx.x; m < cyc; m += blockDim.x)
                        for (int g2 = m; g2 < nMaxG2; g2 += cyc) {
                            coll_rand(a.seed, a.step, cell, stage_now, nMaxG1 + g2, u1, u2);
                            nanbu_collide(Vmax, pmax[nMaxG1 + g2], Vmin, pmin[nMinG1 + m], mMax, qMax, mMin, qMin, lnLp,
                                          a.calculate_bmin, debye, nMaxS, nMinS, nMaxS, a.dt, u1, u2);
                        }
                    ncoll += (unsigned long long)(nMaxG1 > 0 && nMinG1 > 0 ? nMaxG1 : 0) + (unsigned long long)(nMaxG2 > 0 ? nMaxG2 : 0);
                }
            }
            __syncthreads();
        }
    }
    if (threadIdx.x == 0 && ncoll) atomicAdd(a.ncoll, ncoll);
}

hipError_t launch_coulomb(const CoulombArgs& a, bool fp32_velocity, hipStream_t s) {
    if (a.n_cells <= 0) return hipSuccess;
    const size_t lds = (size_t)a.maxp * (2 * sizeof(uint64_t) + 2 * sizeof(int));
    if (fp32_velocity) hipLaunchKernelGGL(k_coulomb<float>, dim3(a.n_cells), dim3(64), lds, s, a);
    else hipLaunchKernelGGL(k_coulomb<double>, dim3(a.n_cells), dim3(64), lds, s, a);
    return hipGetLastError();
}

}  // namespace picmi
