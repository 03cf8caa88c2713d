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

// NanbuCollision.C:137-159
__device__ double cos_chi(double s, double U) {
    double A, invA;
    if (s < 0.1) {
        if (U < 0.0001) U = 0.0001;
        return 1.0 + s * log(U);
    }
    if (s < 3.0) {
        invA = 0.00569578 + (0.95602 + (-0.508139 + (0.479139 + (-0.12789 + 0.0238957 * s) * s) * s) * s) * s;
        A = 1.0 / invA;
        return invA * log(exp(-A) + 2. * U * sinh(A));
    }
    if (s < 6.0) {
        A = 3.0 * exp(-s);
        return (1.0 / A) * log(exp(-A) + 2.0 * U * sinh(A));
    }
    return 2.0 * U - 1.0;
}

template <class UT>
struct Vel {
    UT* x;
    UT* y;
    UT* z;
};

// NanbuCollision::collide (weight correction "none")
template <class UT>
__device__ void nanbu_collide(Vel<UT> vp, int64_t ip, Vel<UT> vq, int64_t iq, double mP, double qP, double mQ,
                              double qQ, double lnL_fixed, int calc_bmin, double debye, double nP, double nQ,
                              double nPQ, double dt, double u1, double u2) {
    const double reMass = mP * mQ / (mP + mQ);
    vec3<double> uP = mk3<double>(vp.x[ip], vp.y[ip], vp.z[ip]);
    vec3<double> uQ = mk3<double>(vq.x[iq], vq.y[iq], vq.z[iq]);
    vec3<double> g = uP - uQ;
    const double gm = mag(g);
    const double gp = sqrt(g.y * g.y + g.z * g.z);
    if (!(gm > 0.0) || !(gp > 0.0)) return;
    double lnL;
    if (lnL_fixed == 0.0) {
        double b0 = qP * qQ / (2.0 * kPi * kEps0 * reMass * gm * gm);
        if (calc_bmin) {
            double hb = kH / (2.0 * (gm * reMass));
            b0 = hb > fabs(b0) ? hb : fabs(b0);
        }
        double l = 0.5 * log(1.0 + (debye * debye) / (b0 * b0));
        lnL = l > 2.0 ? l : 2.0;
    } else {
        lnL = lnL_fixed;
    }
    const double a = qP * qQ / (kEps0 * reMass);
    const double s = lnL / (4.0 * kPi) * (a * a) * (nQ * nP / nPQ) * pow(gm, -3.0) * dt;
    const double cc = cos_chi(s, u1);
    const double sc = sqrt(1.0 - cc * cc);
    const double eps = 2.0 * kPi * u2;
    const double ce = cos(eps), se = sin(eps);
    vec3<double> h = mk3<double>(gp * ce, -1.0 * (g.y * g.x * ce + gm * g.z * se) / gp,
                                 -1.0 * (g.z * g.x * ce - gm * g.y * se) / gp);
    vec3<double> dv = (1.0 - cc) * g + sc * h;
    uP = uP - (mQ / (mP + mQ)) * dv;
    uQ = uQ + (mP / (mP + mQ)) * dv;
    vp.x[ip] = (UT)uP.x; vp.y[ip] = (UT)uP.y; vp.z[ip] = (UT)uP.z;
    vq.x[iq] = (UT)uQ.x; vq.y[iq] = (UT)uQ.y; vq.z[iq] = (UT)uQ.z;
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

// the cell's parcels of one species ordered by the random key of (id, step, stage, species)
__device__ void keyed_perm(uint64_t seed, int64_t step, int stage, int species, const uint32_t* id, int first, int n,
                           uint64_t* key, int* val) {
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
    int stage = 0;
    unsigned long long ncoll = 0;
    for (int i = 0; i < ns; ++i) {
        if (a.charge[i] == 0.0) continue;
        for (int j = ns - 1; j >= i; --j) {
            if (a.charge[j] == 0.0) continue;
            const int st = stage++;
            const int a0 = a.start[i][cell], na = a.start[i][cell + 1] - a0;
            const int b0 = a.start[j][cell], nb = a.start[j][cell + 1] - b0;
            const double mA = a.mass[i], qA = a.charge[i], mB = a.mass[j], qB = a.charge[j];
            const double lnL = a.coulomb_log[i][j];
            const double w = a.weight[i];
            Vel<UT> VA{(UT*)a.ux[i], (UT*)a.uy[i], (UT*)a.uz[i]};
            Vel<UT> VB{(UT*)a.ux[j], (UT*)a.uy[j], (UT*)a.uz[j]};
            double u1, u2;
            if (i == j) {
                if (na <= 1 || !a.allow_intra) continue;
                keyed_perm(a.seed, a.step, st, i, a.id[i], a0, na, keyA, valA);
                const double n = na * w / Vc;
                if (na % 2) {
                    const double n12 = (na + 3.0) * w / Vc;
                    if (threadIdx.x == 0) {
                        const int P = valA[0], Q = valA[1], R = valA[2];
                        coll_rand(a.seed, a.step, cell, st, 0, u1, u2);
                        nanbu_collide(VA, P, VA, Q, mA, qA, mA, qA, lnL, a.calculate_bmin, debye, n, n, n12, a.dt, u1, u2);
                        coll_rand(a.seed, a.step, cell, st, 1, u1, u2);
                        nanbu_collide(VA, Q, VA, R, mA, qA, mA, qA, lnL, a.calculate_bmin, debye, n, n, n12, a.dt, u1, u2);
                        coll_rand(a.seed, a.step, cell, st, 2, u1, u2);
                        nanbu_collide(VA, P, VA, R, mA, qA, mA, qA, lnL, a.calculate_bmin, debye, n, n, n12, a.dt, u1, u2);
                    }
                    __syncthreads();
                    const int npairs = (na - 3) / 2;
                    for (int q = threadIdx.x; q < npairs; q += blockDim.x) {
                        const int x = 3 + 2 * q;
                        coll_rand(a.seed, a.step, cell, st, 3 + q, u1, u2);
                        nanbu_collide(VA, valA[x], VA, valA[x + 1], mA, qA, mA, qA, lnL, a.calculate_bmin, debye, n, n,
                                      n12, a.dt, u1, u2);
                    }
                    ncoll += 3 + npairs;
                } else {
                    const int npairs = na / 2;
                    for (int q = threadIdx.x; q < npairs; q += blockDim.x) {
                        coll_rand(a.seed, a.step, cell, st, q, u1, u2);
                        nanbu_collide(VA, valA[2 * q], VA, valA[2 * q + 1], mA, qA, mA, qA, lnL, a.calculate_bmin,
                                      debye, n, n, n, a.dt, u1, u2);
                    }
                    ncoll += npairs;
                }
            } else {
                if (na == 0 || nb == 0) continue;
                keyed_perm(a.seed, a.step, st, i, a.id[i], a0, na, keyA, valA);
                keyed_perm(a.seed, a.step, st, j, a.id[j], b0, nb, keyB, valB);
                if (na == nb) {
                    const double nP = na * w / Vc, nQ = nb * w / Vc;
                    for (int k = threadIdx.x; k < na; k += blockDim.x) {
                        coll_rand(a.seed, a.step, cell, st, k, u1, u2);
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
                                coll_rand(a.seed, a.step, cell, st, g1, u1, u2);
                                nanbu_collide(Vmax, pmax[g1], Vmin, pmin[m], mMax, qMax, mMin, qMin, lnLp,
                                              a.calculate_bmin, debye, nMaxS, nMinS, nMaxS, a.dt, u1, u2);
                            }
                    // group 2: min parcels nMinG1.. cycle over max parcels nMaxG1..
                    const int cyc = nMinG2 > 0 ? nMinG2 : 1;
                    for (int m = threadIdx.x; m < cyc; m += blockDim.x)
                        for (int g2 = m; g2 < nMaxG2; g2 += cyc) {
                            coll_rand(a.seed, a.step, cell, st, nMaxG1 + g2, u1, u2);
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
