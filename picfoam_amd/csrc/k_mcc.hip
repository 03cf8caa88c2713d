// Electron-neutral null-collision Monte Carlo with a uniform background gas
// (ElectronNeutralCollisionModel::performBackgroundCollisions,
// ElectronNeutralCollisionModel.C:703-829, events :504-700, Nanbu scattering
// NanbuElectronNeutralCollision.C:98-130, ListCrossSection interpolateXY
// ListCrossSection.C:32-92).
//
// One wave per cell.  The cell's electrons are ranked by id in LDS so that a candidate
// index names the same parcel whatever order the sort scatter left them in.
// Candidates are drawn with replacement as in the reference; lanes take one candidate
// each and the lowest lane of a group of candidates that picked the same parcel replays
// the group in candidate order, so every parcel sees its collisions in sequence.
#include <cmath>

#include "device.h"
#include "kernels.h"

namespace picmi {

namespace {

// interpolateXY on an ascending table: clamped linear interpolation between e[lo] <= x <
// e[lo + 1].  With a guide (tables starting above zero) the bracket comes from kXsGuide
// bins equal in log(e), each holding the last entry at or below its lower edge, and a
// step or two along the table; otherwise from a bisection.  Both give the same lo.
__device__ double xs_lookup(const double* __restrict__ e, const double* __restrict__ s, int n,
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

__device__ __forceinline__ void mcc_rand(uint64_t seed, int64_t step, int cell, int k, int j, double& a, double& b) {
    uint4 o = rng4(seed, (uint32_t)cell, (uint32_t)step, (uint32_t)k, (kTagMcc << 24) | (uint32_t)j);
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

// (1 + eV)^u for eV >= 0, u in [0, 1)
__device__ __forceinline__ double pow1p(double eV, double u) { return exp_small(u * log_pos(1.0 + eV)); }

// NanbuElectronNeutralCollision::updateVelocity with a sampled partner velocity
__device__ void nanbu_scatter(double eV, vec3<double>& UE, vec3<double>& Un, double mE, double mN, double uphi,
                              double uchi) {
    double cc;
    if (eV < 1e-30) cc = 1.0;
    else cc = 1.0 / eV * (eV + 2. - 2. * pow1p(eV, uchi));
    double sc;
    if (cc * cc > 1.0) { cc = 1.0; sc = 0.0; }
    else sc = sqrt(1.0 - cc * cc);
    const vec3<double> g = UE - Un;
    const double gm = mag(g), gp = sqrt(g.y * g.y + g.z * g.z);
    if (!(gp > 0.0)) return;
    double sp, cp;
    sincos_turn(uphi, &sp, &cp);
    const vec3<double> h = mk3<double>(gp * cp, -1.0 * (g.y * g.x * cp + gm * g.z * sp) / gp,
                                       -1.0 * (g.z * g.x * cp - gm * g.y * sp) / gp);
    const vec3<double> dv = (1.0 - cc) * g + sc * h;
    UE = UE - (mN / (mN + mE)) * dv;
    Un = Un + (mE / (mN + mE)) * dv;
}

// RelativisticElectronNeutralCollision::updateVelocity with a sampled partner velocity
// (RelativisticElectronNeutralCollision.C:161-240).  cos(chi) is clamped to [-1, 1] as in
// the two-parcel variant (:120-126); the sampled variant would produce a NaN there.
__device__ void rel_scatter(double eV, vec3<double>& UE, vec3<double>& Un, double mE, double mN, double ueps,
                            double uchi) {
    ComFrame f;
    com_enter(UE, Un, mE, mN, f);
    if (f.pm == 0.0) return;
    double cc, sc;
    if (eV < 1e-30) {
        cc = 1.0; sc = 0.0;
    } else {
        cc = 1.0 / eV * (eV + 2. - 2. * pow1p(eV, uchi));
        if (cc * cc > 1.0) { cc = 1.0; sc = 0.0; }
        else sc = sqrt(1.0 - cc * cc);
    }
    double sp, cp;
    sincos_turn(ueps, &sp, &cp);
    com_leave(f, mE, mN, cc, sc, sp, cp, UE, Un);
}

template <class UT>
__device__ __forceinline__ void en_scatter(const MccArgs<UT>& a, double eV, vec3<double>& UE, vec3<double>& Un,
                                           double us, double uc) {
    if (a.relativistic) rel_scatter(eV, UE, Un, a.massE, a.massN, us, uc);
    else nanbu_scatter(eV, UE, Un, a.massE, a.massN, us, uc);
}

template <class UT>
__device__ void deposit_new(const MccArgs<UT>& a, int t, double y1, double y2, double y3, int* count, double s42) {
    const TetGeo& g = a.tets[t];
    atomicAdd(count + g.cell, 1);
    if (a.dep_vw) {
        const double y0 = 1.0 - (y1 + y2 + y3);
        atomicAdd(a.acc_cell + g.cell, (unsigned long long)llrint(y0 * s42));
        atomicAdd(a.acc_point + g.vm[0], (unsigned long long)llrint(y1 * s42));
        atomicAdd(a.acc_point + g.vm[1], (unsigned long long)llrint(y2 * s42));
        atomicAdd(a.acc_point + g.vm[2], (unsigned long long)llrint(y3 * s42));
    }
}

// one candidate (ElectronNeutralCollisionModel.C:747-800 and the event handlers)
template <class UT>
__device__ void candidate(const MccArgs<UT>& a, int cell, int k, int ei, double smax_local, double& qtcr,
                          unsigned& c_el, unsigned& c_ex, unsigned& c_iz) {
    double ua, ub, n1, n2, n3, n4, upick, uacc;
    mcc_rand(a.seed, a.step, cell, k, 0, upick, uacc);
    mcc_rand(a.seed, a.step, cell, k, 1, ua, ub);
    box_muller(ua, ub, n1, n2);
    mcc_rand(a.seed, a.step, cell, k, 2, ua, ub);
    box_muller(ua, ub, n3, n4);
    vec3<double> Un = mk3<double>(a.vthN * n1 + a.drift[0], a.vthN * n2 + a.drift[1], a.vthN * n3 + a.drift[2]);
    vec3<double> Ue = mk3<double>(a.e.ux[ei], a.e.uy[ei], a.e.uz[ei]);
    const double c = kC;
    const double cR = mag(Ue - Un);
    const double gamma = 1.0 / sqrt(1.0 - ((cR * cR) / (c * c)));
    double eV = (gamma - 1.0) * a.massE * c * c / kQe;
    const double sigmaT = smax_local / cR;
    const double Qel = xs_lookup(a.xs_e[0], a.xs_s[0], a.xs_n[0], a.xs_guide[0], a.xs_le0[0], a.xs_idl[0], eV);
    const double Qex = xs_lookup(a.xs_e[1], a.xs_s[1], a.xs_n[1], a.xs_guide[1], a.xs_le0[1], a.xs_idl[1], eV);
    const double Qiz = xs_lookup(a.xs_e[2], a.xs_s[2], a.xs_n[2], a.xs_guide[2], a.xs_le0[2], a.xs_idl[2], eV);
    const double Qt = Qel + Qex + Qiz;
    qtcr = fmax(qtcr, Qt * cR);
    double us, uc;
    mcc_rand(a.seed, a.step, cell, k, 3, us, uc);
    if (uacc <= Qel / sigmaT) {
        en_scatter(a, eV, Ue, Un, us, uc);
        ++c_el;
    } else if (eV >= a.th_ex && uacc <= (Qel + Qex) / sigmaT) {
        Ue = sqrt(1.0 - a.th_ex / eV) * Ue;
        eV -= a.th_ex;
        en_scatter(a, eV, Ue, Un, us, uc);
        ++c_ex;
    } else if (eV >= a.th_iz && uacc <= Qt / sigmaT) {
        double uE, uR, up2, uc2;
        mcc_rand(a.seed, a.step, cell, k, 4, uE, uR);
        mcc_rand(a.seed, a.step, cell, k, 5, up2, uc2);
        const double E0 = 2.0 - 100.0 / (eV + 10.0);
        const double E1 = (eV - a.th_iz) / 2.0 - E0;
        const double alpha = 10.3;
        const double Ep = E0 + alpha * tan(uE * (atan(E1 / alpha) + atan(E0 / alpha)) - atan(E0 / alpha));
        const vec3<double> preUe = Ue, preUn = Un;
        Ue = sqrt(1.0 - (a.th_iz + Ep) / eV) * Ue;
        eV -= (a.th_iz + Ep);
        en_scatter(a, eV, Ue, Un, us, uc);
        const int t = a.e.tet[ei];
        const double y1 = a.e.y1[ei], y2 = a.e.y2[ei], y3 = a.e.y3[ei];
        bool created = false;
        // the ion and the electron of one ionization (ElectronNeutralCollisionModel.C:680-698)
        // are created together or not at all: both slots are reserved first, and when either
        // species is full the event creates neither, raises the capacity flag and leaves any
        // slot it did get as a deleted parcel (tet -1), so the charge stays balanced
        const bool make = uR <= a.wEwN;
        const bool mk_i = make && a.create_ion, mk_e = make && a.create_electron;
        const int64_t slot_i = mk_i ? (int64_t)atomicAdd((unsigned long long*)a.n_i, 1ull) : -1;
        const int64_t slot_e = mk_e ? (int64_t)atomicAdd((unsigned long long*)a.n_e, 1ull) : -1;
        const bool fit = (!mk_i || slot_i < a.cap_i) && (!mk_e || slot_e < a.cap_e);
        if (make && !fit) {
            atomicOr(a.flags, kFlagCapacity);
            if (mk_i && slot_i < a.cap_i) { a.ion.tet[slot_i] = -1; a.ion.id[slot_i] = 0u; }
            if (mk_e && slot_e < a.cap_e) { a.e.tet[slot_e] = -1; a.e.id[slot_e] = 0u; }
        }
        if (mk_i && fit) {
            a.ion.y1[slot_i] = y1; a.ion.y2[slot_i] = y2; a.ion.y3[slot_i] = y3;
            a.ion.tet[slot_i] = t;
            a.ion.ux[slot_i] = (UT)Un.x; a.ion.uy[slot_i] = (UT)Un.y; a.ion.uz[slot_i] = (UT)Un.z;
            uint4 o = rng4(a.seed, a.e.id[ei], (uint32_t)a.step, (uint32_t)k, (kTagNewId << 24) | 1u);
            a.ion.id[slot_i] = o.x;
            deposit_new(a, t, y1, y2, y3, a.count_i, a.s42_i);
            created = true;
        }
        if (uR > a.wNwE) Ue = preUe;
        if (mk_e && fit) {
            const double sp = sqrt(2.0 * Ep * kQe / a.massE);
            vec3<double> Unew = sp * (preUe / mag(preUe));
            vec3<double> Un2 = preUn;
            en_scatter(a, Ep, Unew, Un2, up2, uc2);
            a.e.y1[slot_e] = y1; a.e.y2[slot_e] = y2; a.e.y3[slot_e] = y3;
            a.e.tet[slot_e] = t;
            a.e.ux[slot_e] = (UT)Unew.x; a.e.uy[slot_e] = (UT)Unew.y; a.e.uz[slot_e] = (UT)Unew.z;
            uint4 o = rng4(a.seed, a.e.id[ei], (uint32_t)a.step, (uint32_t)k, (kTagNewId << 24) | 2u);
            a.e.id[slot_e] = o.x;
            deposit_new(a, t, y1, y2, y3, a.count_e, a.s42_e);
        }
        if (created) ++c_iz;
    }
    a.e.ux[ei] = (UT)Ue.x; a.e.uy[ei] = (UT)Ue.y; a.e.uz[ei] = (UT)Ue.z;
}

}  // namespace

// the candidates of one cell by one wave; maxp = LDS slots for the cell's electrons
template <class UT>
__device__ void mcc_cell(const MccArgs<UT>& a, int cell, int maxp, uint64_t* lds, int* picks) {
    uint64_t* key = lds;
    int* val = (int*)(lds + maxp);
    const int a0 = a.start[cell], eC = a.start[cell + 1] - a0;
    const double smax = a.smax[cell];
    const double coll = a.rem[cell] + eC * a.nN * smax * a.dt;
    const int nc = (int)coll;
    __syncthreads();
    if (threadIdx.x == 0) a.rem[cell] = coll - nc;
    unsigned c_el = 0, c_ex = 0, c_iz = 0;
    double qtcr = 0.0;
    if (eC > 0 && nc > 0) {
        // val[r] = the electron of id rank r: each key (id, cell-local index) is distinct and
        // its rank is the number of smaller keys, counted eight keys per LDS read
        const int n8 = (eC + 7) & ~7;
        for (int k = threadIdx.x; k < n8; k += blockDim.x)
            key[k] = (k < eC) ? (((uint64_t)a.id_cell[a0 + k] << 32) | (uint32_t)k) : ~0ull;
        __syncthreads();
        for (int k = threadIdx.x; k < eC; k += blockDim.x) {
            const uint64_t kk = key[k];
            int r = 0;
            for (int j = 0; j < n8; j += 8) {
                const uint64_t* q = key + j;
                r += (int)(q[0] < kk) + (int)(q[1] < kk) + (int)(q[2] < kk) + (int)(q[3] < kk) + (int)(q[4] < kk) +
                     (int)(q[5] < kk) + (int)(q[6] < kk) + (int)(q[7] < kk);
            }
            val[r] = slot_index(a.perm, a0 + k);
        }
        __syncthreads();
        for (int base = 0; base < nc; base += 64) {
            const int k = base + (int)threadIdx.x;
            int pick = -1;
            if (k < nc) {
                double upick, uacc;
                mcc_rand(a.seed, a.step, cell, k, 0, upick, uacc);
                pick = (int)(upick * eC);
                if (pick >= eC) pick = eC - 1;
            }
            picks[threadIdx.x] = pick;
            __syncthreads();
            if (k < nc) {
                bool first = true;
                for (int l = 0; l < (int)threadIdx.x; ++l)
                    if (picks[l] == pick) { first = false; break; }
                if (first) {
                    const int ei = val[pick];
                    for (int l = (int)threadIdx.x; l < 64 && base + l < nc; ++l)
                        if (picks[l] == pick) candidate(a, cell, base + l, ei, smax, qtcr, c_el, c_ex, c_iz);
                }
            }
            __syncthreads();
        }
    }
    for (int off = 32; off > 0; off >>= 1) qtcr = fmax(qtcr, __shfl_xor(qtcr, off));
    if (threadIdx.x == 0 && qtcr > smax) a.smax[cell] = qtcr;
    unsigned v1 = c_el, v2 = c_ex, v3 = c_iz;
    for (int off = 32; off > 0; off >>= 1) {
        v1 += __shfl_xor(v1, off);
        v2 += __shfl_xor(v2, off);
        v3 += __shfl_xor(v3, off);
    }
    // one of kNcollSlots counter groups a cache line apart: one address taking an atomic
    // from every cell would serialise the cells
    if (threadIdx.x == 0) {
        unsigned long long* cnt = a.counts + kNcollStride * (blockIdx.x % kNcollSlots);
        if (nc) atomicAdd(cnt + 0, (unsigned long long)nc);
        if (v1) atomicAdd(cnt + 1, (unsigned long long)v1);
        if (v2) atomicAdd(cnt + 2, (unsigned long long)v2);
        if (v3) atomicAdd(cnt + 3, (unsigned long long)v3);
    }
}

// one wave per cell; cells with more electrons than a.maxp go to the large-cell pass.  Three
// waves per SIMD with a few spilled registers beat two without (measured on the 100^3 argon
// case: 6.7 against 8.4 ms per step)
template <class UT>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3, 8))) k_mcc(MccArgs<UT> a) {
    extern __shared__ uint64_t lds[];
    __shared__ int picks[64];
    const int cell = blockIdx.x;
    if (a.start[cell + 1] - a.start[cell] > a.maxp) {
        if (threadIdx.x == 0) a.big_list[atomicAdd(a.big_n, 1)] = cell;
        return;
    }
    mcc_cell(a, cell, a.maxp, lds, picks);
}

// Several cells per wave: 64 / L cells, L lanes each (L = 16 or 32), for cells of up to a.maxp_group
// electrons (larger ones go to the large-cell pass).  A cell draws a few candidates per
// step (about 5 of 100 electrons at the argon case's dt), so one cell per wave leaves most
// lanes idle while the wave pays the loads, the id ranking and the barriers for that cell
// alone.  Same candidate rule as mcc_cell, in rounds of L candidates per cell: a parcel
// still sees its candidates in candidate order, so the results do not depend on L.
template <class UT, int L>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3, 8))) k_mcc_group(MccArgs<UT> a) {
    constexpr int W = 64 / L;
    extern __shared__ uint64_t lds[];
    __shared__ int picks_all[64];
    const int w = threadIdx.x / L, r = threadIdx.x % L;
    const int maxp = a.maxp_group;
    uint64_t* key = lds + (size_t)w * (maxp + maxp / 2);
    int* val = (int*)(key + maxp);
    int* picks = picks_all + w * L;
    const int cell = blockIdx.x * W + w;
    bool on = cell < a.n_cells;
    const int a0 = on ? a.start[cell] : 0;
    const int eC = on ? a.start[cell + 1] - a0 : 0;
    if (on && eC > maxp) {
        if (r == 0) a.big_list[atomicAdd(a.big_n, 1)] = cell;
        on = false;
    }
    // candidates of this step: the group's first lane reads and rewrites the remainder
    double smax = 0.0;
    int nc = 0;
    if (on && r == 0) {
        smax = a.smax[cell];
        const double coll = a.rem[cell] + eC * a.nN * smax * a.dt;
        nc = (int)coll;
        a.rem[cell] = coll - nc;
    }
    smax = __shfl(smax, w * L);
    nc = __shfl(nc, w * L);
    const int ncw = (eC > 0) ? nc : 0;
    unsigned c_el = 0, c_ex = 0, c_iz = 0;
    double qtcr = 0.0;
    // val[r] = the electron of id rank r (keys (id, cell-local index) are distinct)
    const int n8 = ncw > 0 ? (eC + 7) & ~7 : 0;
    for (int k = r; k < n8; k += L)
        key[k] = (k < eC) ? (((uint64_t)a.id_cell[a0 + k] << 32) | (uint32_t)k) : ~0ull;
    __syncthreads();
    for (int k = r; k < (ncw > 0 ? eC : 0); k += L) {
        const uint64_t kk = key[k];
        int rk = 0;
        for (int j = 0; j < n8; j += 8) {
            const uint64_t* q = key + j;
            rk += (int)(q[0] < kk) + (int)(q[1] < kk) + (int)(q[2] < kk) + (int)(q[3] < kk) + (int)(q[4] < kk) +
                  (int)(q[5] < kk) + (int)(q[6] < kk) + (int)(q[7] < kk);
        }
        val[rk] = slot_index(a.perm, a0 + k);
    }
    __syncthreads();
    // rounds of L candidates; every group takes part in the barriers of the busiest one
    int rounds = (ncw + L - 1) / L;
    for (int off = 32; off > 0; off >>= 1) rounds = max(rounds, __shfl_xor(rounds, off));
    for (int base = 0; base < rounds * L; base += L) {
        const int k = base + r;
        int pick = -1;
        if (k < ncw) {
            double upick, uacc;
            mcc_rand(a.seed, a.step, cell, k, 0, upick, uacc);
            pick = (int)(upick * eC);
            if (pick >= eC) pick = eC - 1;
        }
        picks[r] = pick;
        __syncthreads();
        if (k < ncw) {
            bool first = true;
            for (int l = 0; l < r; ++l)
                if (picks[l] == pick) { first = false; break; }
            if (first) {
                const int ei = val[pick];
                for (int l = r; l < L && base + l < ncw; ++l)
                    if (picks[l] == pick) candidate(a, cell, base + l, ei, smax, qtcr, c_el, c_ex, c_iz);
            }
        }
        __syncthreads();
    }
    for (int off = L / 2; off > 0; off >>= 1) qtcr = fmax(qtcr, __shfl_xor(qtcr, off));
    if (on && r == 0 && qtcr > smax) a.smax[cell] = qtcr;
    unsigned v0 = (on && r == 0) ? (unsigned)nc : 0u, v1 = c_el, v2 = c_ex, v3 = c_iz;
    for (int off = 32; off > 0; off >>= 1) {
        v0 += __shfl_xor(v0, off);
        v1 += __shfl_xor(v1, off);
        v2 += __shfl_xor(v2, off);
        v3 += __shfl_xor(v3, off);
    }
    if (threadIdx.x == 0) {
        unsigned long long* cnt = a.counts + kNcollStride * (blockIdx.x % kNcollSlots);
        if (v0) atomicAdd(cnt + 0, (unsigned long long)v0);
        if (v1) atomicAdd(cnt + 1, (unsigned long long)v1);
        if (v2) atomicAdd(cnt + 2, (unsigned long long)v2);
        if (v3) atomicAdd(cnt + 3, (unsigned long long)v3);
    }
}

template <class UT>
__global__ void __launch_bounds__(64) k_mcc_big(MccArgs<UT> a) {
    extern __shared__ uint64_t lds[];
    __shared__ int picks[64];
    const int nbig = *a.big_n;
    for (int i = blockIdx.x; i < nbig; i += gridDim.x) {
        const int cell = a.big_list[i];
        if (a.start[cell + 1] - a.start[cell] > a.maxp_big) {
            if (threadIdx.x == 0) atomicOr(a.flags, kFlagCellOverflow);
            continue;
        }
        mcc_cell(a, cell, a.maxp_big, lds, picks);
        __syncthreads();
    }
}

template <class UT>
static hipError_t launch_mcc_t(const MccArgs<UT>& a, hipStream_t s) {
    if (a.n_cells <= 0) return hipSuccess;
    if (a.group == 16) {
        const size_t lds = (size_t)4 * (a.maxp_group + a.maxp_group / 2) * sizeof(uint64_t);
        hipLaunchKernelGGL((k_mcc_group<UT, 16>), dim3((a.n_cells + 3) / 4), dim3(64), lds, s, a);
    } else if (a.group == 32) {
        const size_t lds = (size_t)2 * (a.maxp_group + a.maxp_group / 2) * sizeof(uint64_t);
        hipLaunchKernelGGL((k_mcc_group<UT, 32>), dim3((a.n_cells + 1) / 2), dim3(64), lds, s, a);
    } else {
        const size_t lds = (size_t)a.maxp * (sizeof(uint64_t) + sizeof(int));
        hipLaunchKernelGGL((k_mcc<UT>), dim3(a.n_cells), dim3(64), lds, s, a);
    }
    const size_t lds_big = (size_t)a.maxp_big * (sizeof(uint64_t) + sizeof(int));
    const int nb = a.n_cells < 512 ? a.n_cells : 512;
    hipLaunchKernelGGL((k_mcc_big<UT>), dim3(nb), dim3(64), lds_big, s, a);
    return hipGetLastError();
}
hipError_t launch_mcc(const MccArgs<double>& a, hipStream_t s) { return launch_mcc_t(a, s); }
hipError_t launch_mcc(const MccArgs<float>& a, hipStream_t s) { return launch_mcc_t(a, s); }

// the lookup guide of an ascending table with e[0] > 0 and n >= 2 (host)
void xs_guide(const double* e, int n, int* guide, double* le0, double* idl) {
    const double l0 = std::log(e[0]), l1 = std::log(e[n - 1]);
    *le0 = l0;
    *idl = kXsGuide / (l1 - l0);
    int lo = 0;
    for (int b = 0; b < kXsGuide; ++b) {
        const double edge = std::exp(l0 + (l1 - l0) * b / kXsGuide);
        while (lo + 1 < n - 1 && e[lo + 1] <= edge) ++lo;
        guide[b] = lo;
    }
}

}  // namespace picmi
