// Electron-neutral null-collision Monte Carlo with a uniform background gas
// (ElectronNeutralCollisionModel::performBackgroundCollisions,
// ElectronNeutralCollisionModel.C:703-829, events :504-700, Nanbu scattering
// NanbuElectronNeutralCollision.C:98-130, ListCrossSection interpolateXY
// ListCrossSection.C:32-92).
//
// The cell's electrons are ordered by a 52-bit mix of their ids (the cell-local index below
// it), so that a candidate index names the same parcel whatever order the sort scatter left
// them in.  The order is not computed in full: the keys go into LDS buckets by their top
// bits (a histogram, its prefix and one placement per key), and a candidate's rank is
// resolved inside its bucket (about one key per bucket).  Candidates are drawn with
// replacement as in the reference; lanes take one candidate each and the lowest lane of a
// group of candidates that picked the same parcel replays the group in candidate order, so
// every parcel sees its collisions in sequence.
#include <cmath>

#include "collide.h"

namespace picmi {

namespace {

// the tag of a scattering model's draws: kTagMcc for electrons, kTagIonMcc for ions
template <int REL>
constexpr uint32_t mcc_tag() { return REL == 2 ? kTagIonMcc : kTagMcc; }

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

// REL: the scattering model, a template parameter so that each kernel holds the registers
// of one model only: kMccNanbu, kMccRelativistic (RelativisticElectronNeutral) or kMccIon
// (ion-neutral IsotropicScattering with charge exchange)
template <int REL, class UT>
__device__ __forceinline__ void en_scatter(const MccArgs<UT>& a, double eV, vec3<double>& UE, vec3<double>& Un,
                                           double us, double uc) {
    if constexpr (REL == kMccRelativistic) rel_scatter(eV, UE, Un, a.massE, a.massN, us, uc);
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

// An ionization whose pair (ion and secondary electron) is created after the round of
// candidates it belongs to, from the primary's velocity before the event: everything else
// the creation needs is recomputed from the candidate's random numbers.  Creating the pairs
// apart keeps the candidate loop's registers free of the creation's.
struct MccPair {
    double pu[3];
    int ei, k, cell, pad;
};

// the partner of candidate k of a cell, from the background Maxwellian
// (uniformBackground::sampleVelocity, uniformBackground.C:121-127)
template <int REL, class UT>
__device__ __forceinline__ vec3<double> bg_partner(const MccArgs<UT>& a, int cell, int k) {
    double ua, ub, n1, n2, n3, n4;
    mcc_rand(mcc_tag<REL>(), a.seed, a.step, cell, k, 1, ua, ub);
    box_muller(ua, ub, n1, n2);
    mcc_rand(mcc_tag<REL>(), a.seed, a.step, cell, k, 2, ua, ub);
    box_muller(ua, ub, n3, n4);
    return mk3<double>(a.vthN * n1 + a.drift[0], a.vthN * n2 + a.drift[1], a.vthN * n3 + a.drift[2]);
}

// relativistic kinetic energy [eV] of the relative motion (ElectronNeutralCollisionModel.C:755-759)
__device__ __forceinline__ double rel_energy_ev(double cR, double massE) {
    const double c = kC;
    const double gamma = 1.0 / sqrt(1.0 - ((cR * cR) / (c * c)));
    return (gamma - 1.0) * massE * c * c / kQe;
}

// energy of the secondary electron, the Opal-like split (ElectronNeutralCollisionModel.C:656-659)
__device__ __forceinline__ double opal_energy(double eV, double th, double uE) {
    const double E0 = 2.0 - 100.0 / (eV + 10.0);
    const double E1 = (eV - th) / 2.0 - E0;
    const double alpha = 10.3;
    return E0 + alpha * tan(uE * (atan(E1 / alpha) + atan(E0 / alpha)) - atan(E0 / alpha));
}

// one ion-neutral candidate against the background (IonNeutralCollisionModel.C:497-545):
// elastic with probability Qel cR / max, scattered isotropically in the centre of mass
// (IonIsotropicScattering.C:108-140: cos(theta) = 2 u1 - 1 along x, phi = 2 pi u2), else up
// to Qt cR / max above the threshold a charge exchange, the ion taking the partner's
// velocity.  The ion's slot is a.e (the ion species) and its mass a.massE.
template <class UT>
__device__ void ion_candidate(const MccArgs<UT>& a, int cell, int k, int ii, double smax_local, double& qtcr,
                              unsigned& c_el, unsigned& c_cx) {
    constexpr int REL = 2;
    double upick, uacc;
    mcc_rand(mcc_tag<REL>(), a.seed, a.step, cell, k, 0, upick, uacc);
    const vec3<double> Un = bg_partner<REL>(a, cell, k);
    vec3<double> Ui = mk3<double>(a.e.ux[ii], a.e.uy[ii], a.e.uz[ii]);
    const double cR = mag(Ui - Un);
    const double eV = rel_energy_ev(cR, a.massE);
    const double sigmaT = smax_local / cR;
    const double Qel = xs_lookup(a.xs_e[0], a.xs_s[0], a.xs_n[0], a.xs_guide[0], a.xs_le0[0], a.xs_idl[0], eV);
    const double Qcx = xs_lookup(a.xs_e[1], a.xs_s[1], a.xs_n[1], a.xs_guide[1], a.xs_le0[1], a.xs_idl[1], eV);
    const double Qt = Qel + Qcx;
    qtcr = fmax(qtcr, Qt * cR);
    if (uacc <= Qel / sigmaT) {
        double u1, u2;
        mcc_rand(mcc_tag<REL>(), a.seed, a.step, cell, k, 3, u1, u2);
        const double mP = a.massE, mQ = a.massN, ims = 1.0 / (mP + mQ);
        const vec3<double> cm = ims * (mP * Ui + mQ * Un);
        const double ct = 2.0 * u1 - 1.0, st = sqrt(1.0 - ct * ct);
        double sp, cp;
        sincos_turn(u2, &sp, &cp);
        Ui = cm + (mQ * ims * cR) * mk3<double>(ct, st * cp, st * sp);
        ++c_el;
    } else if (eV >= a.th_ex && uacc <= Qt / sigmaT) {
        Ui = Un;
        ++c_cx;
    }
    a.e.ux[ii] = (UT)Ui.x; a.e.uy[ii] = (UT)Ui.y; a.e.uz[ii] = (UT)Ui.z;
}

// one candidate (ElectronNeutralCollisionModel.C:747-800 and the event handlers); an
// ionization that creates a pair appends it to pairs[*npairs] (LDS)
template <int REL, class UT>
__device__ void candidate(const MccArgs<UT>& a, int cell, int k, int ei, double smax_local, double& qtcr,
                          unsigned& c_el, unsigned& c_ex, MccPair* pairs, int* npairs) {
    if constexpr (REL == kMccIon) {
        ion_candidate(a, cell, k, ei, smax_local, qtcr, c_el, c_ex);
        return;
    }
    double upick, uacc;
    mcc_rand(mcc_tag<REL>(), a.seed, a.step, cell, k, 0, upick, uacc);
    vec3<double> Un = bg_partner<REL>(a, cell, k);
    vec3<double> Ue = mk3<double>(a.e.ux[ei], a.e.uy[ei], a.e.uz[ei]);
    const double cR = mag(Ue - Un);
    double eV = rel_energy_ev(cR, a.massE);
    const double sigmaT = smax_local / cR;
    const double Qel = xs_lookup(a.xs_e[0], a.xs_s[0], a.xs_n[0], a.xs_guide[0], a.xs_le0[0], a.xs_idl[0], eV);
    const double Qex = xs_lookup(a.xs_e[1], a.xs_s[1], a.xs_n[1], a.xs_guide[1], a.xs_le0[1], a.xs_idl[1], eV);
    const double Qiz = xs_lookup(a.xs_e[2], a.xs_s[2], a.xs_n[2], a.xs_guide[2], a.xs_le0[2], a.xs_idl[2], eV);
    const double Qt = Qel + Qex + Qiz;
    qtcr = fmax(qtcr, Qt * cR);
    double us, uc;
    mcc_rand(mcc_tag<REL>(), a.seed, a.step, cell, k, 3, us, uc);
    if (uacc <= Qel / sigmaT) {
        en_scatter<REL>(a, eV, Ue, Un, us, uc);
        ++c_el;
    } else if (eV >= a.th_ex && uacc <= (Qel + Qex) / sigmaT) {
        Ue = sqrt(1.0 - a.th_ex / eV) * Ue;
        eV -= a.th_ex;
        en_scatter<REL>(a, eV, Ue, Un, us, uc);
        ++c_ex;
    } else if (eV >= a.th_iz && uacc <= Qt / sigmaT) {
        double uE, uR;
        mcc_rand(mcc_tag<REL>(), a.seed, a.step, cell, k, 4, uE, uR);
        const double Ep = opal_energy(eV, a.th_iz, uE);
        const vec3<double> preUe = Ue;
        Ue = sqrt(1.0 - (a.th_iz + Ep) / eV) * Ue;
        eV -= (a.th_iz + Ep);
        en_scatter<REL>(a, eV, Ue, Un, us, uc);
        if (uR > a.wNwE) Ue = preUe;
        if (uR <= a.wEwN && (a.create_ion || a.create_electron)) {
            const int e = atomicAdd(npairs, 1);
            MccPair& P = pairs[e];
            P.pu[0] = preUe.x; P.pu[1] = preUe.y; P.pu[2] = preUe.z;
            P.ei = ei; P.k = k; P.cell = cell;
        }
    }
    a.e.ux[ei] = (UT)Ue.x; a.e.uy[ei] = (UT)Ue.y; a.e.uz[ei] = (UT)Ue.z;
}

// The ion and the secondary electron of an ionization (ElectronNeutralCollisionModel.C:
// 662-698): the ion carries the partner's velocity after the primary's scattering, the
// electron Ep along the primary's old direction, scattered on the partner's old velocity.
// They are created together or not at all: both slots are reserved first, and when either
// species is full the event creates neither, raises the capacity flag and leaves any slot
// it did get as a deleted parcel (tet -1), so the charge stays balanced.
template <int REL, class UT>
__device__ void create_pair(const MccArgs<UT>& a, const MccPair& P, unsigned& c_iz) {
    const int cell = P.cell, k = P.k, ei = P.ei;
    const vec3<double> preUe = mk3<double>(P.pu[0], P.pu[1], P.pu[2]);
    vec3<double> Un = bg_partner<REL>(a, cell, k);
    const vec3<double> preUn = Un;
    double eV = rel_energy_ev(mag(preUe - Un), a.massE);
    double us, uc, uE, uR, up2, uc2;
    mcc_rand(mcc_tag<REL>(), a.seed, a.step, cell, k, 3, us, uc);
    mcc_rand(mcc_tag<REL>(), a.seed, a.step, cell, k, 4, uE, uR);
    mcc_rand(mcc_tag<REL>(), a.seed, a.step, cell, k, 5, up2, uc2);
    const double Ep = opal_energy(eV, a.th_iz, uE);
    vec3<double> Ue = sqrt(1.0 - (a.th_iz + Ep) / eV) * preUe;
    eV -= (a.th_iz + Ep);
    en_scatter<REL>(a, eV, Ue, Un, us, uc);
    const int t = a.e.tet[ei];
    const double y1 = a.e.y1[ei], y2 = a.e.y2[ei], y3 = a.e.y3[ei];
    const uint32_t pid = a.e.id[ei];
    const bool mk_i = a.create_ion, mk_e = a.create_electron;
    const int64_t slot_i = mk_i ? (int64_t)atomicAdd((unsigned long long*)a.n_i, 1ull) : -1;
    const int64_t slot_e = mk_e ? (int64_t)atomicAdd((unsigned long long*)a.n_e, 1ull) : -1;
    const bool fit = (!mk_i || slot_i < a.cap_i) && (!mk_e || slot_e < a.cap_e);
    if (!fit) {
        atomicOr(a.flags, kFlagCapacity);
        if (mk_i && slot_i < a.cap_i) { a.ion.tet[slot_i] = -1; a.ion.id[slot_i] = 0u; }
        if (mk_e && slot_e < a.cap_e) { a.e.tet[slot_e] = -1; a.e.id[slot_e] = 0u; }
        return;
    }
    if (mk_i) {
        a.ion.y1[slot_i] = y1; a.ion.y2[slot_i] = y2; a.ion.y3[slot_i] = y3;
        a.ion.tet[slot_i] = t;
        a.ion.ux[slot_i] = (UT)Un.x; a.ion.uy[slot_i] = (UT)Un.y; a.ion.uz[slot_i] = (UT)Un.z;
        a.ion.id[slot_i] = rng4(a.seed, pid, (uint32_t)a.step, (uint32_t)k, (kTagNewId << 24) | 1u).x;
        deposit_new(a, t, y1, y2, y3, a.count_i, a.s42_i);
        ++c_iz;
    }
    if (mk_e) {
        const double sp = sqrt(2.0 * Ep * kQe / a.massE);
        vec3<double> Unew = sp * (preUe / mag(preUe));
        vec3<double> Un2 = preUn;
        en_scatter<REL>(a, Ep, Unew, Un2, up2, uc2);
        a.e.y1[slot_e] = y1; a.e.y2[slot_e] = y2; a.e.y3[slot_e] = y3;
        a.e.tet[slot_e] = t;
        a.e.ux[slot_e] = (UT)Unew.x; a.e.uy[slot_e] = (UT)Unew.y; a.e.uz[slot_e] = (UT)Unew.z;
        a.e.id[slot_e] = rng4(a.seed, pid, (uint32_t)a.step, (uint32_t)k, (kTagNewId << 24) | 2u).x;
        deposit_new(a, t, y1, y2, y3, a.count_e, a.s42_e);
    }
}

// Order key of an electron in its cell: 52 bits of the splitmix64 finaliser of its id, the
// cell-local index in the 12 bits below (cells of up to kMccSlotsBig = 4096 electrons), so
// keys are distinct and parcels that share an id keep their list order.
constexpr int kMccIndexBits = 12;
static_assert(kMccSlotsBig <= (1 << kMccIndexBits), "cell-local index must fit below the order key");
__device__ __forceinline__ uint64_t mcc_key(uint32_t id, int k) {
    const uint64_t z = id_mix(id);
    return ((z >> kMccIndexBits) << kMccIndexBits) | (uint64_t)k;
}

// The n electrons of a cell (ids id[first + k]) bucketed by the top log2(NB) bits of their
// keys, by the cell's L lanes (lane r): skey[0..n) holds the keys bucket after bucket (each
// bucket unordered), bins[b] the start of bucket b.  place[] is scratch.  Ends with a barrier.
template <int L, int NB>
__device__ void mcc_bucket(const uint32_t* __restrict__ id, int first, int n, uint64_t* skey, int* place, int* bins,
                           int r) {
    constexpr int PER = NB / L;
    constexpr int SH = 64 - __builtin_ctz(NB);
    for (int b = r; b < NB; b += L) bins[b] = 0;
    __syncthreads();
    for (int k = r; k < n; k += L) place[k] = atomicAdd(&bins[(int)(mcc_key(id[first + k], k) >> SH)], 1);
    __syncthreads();
    // exclusive prefix of the bins: PER consecutive bins per lane, scanned over the lanes
    int loc = 0;
#pragma unroll
    for (int b = 0; b < PER; ++b) loc += bins[r * PER + b];
    int inc = loc;
#pragma unroll
    for (int off = 1; off < L; off <<= 1) {
        const int y = __shfl_up(inc, off, L);
        if (r >= off) inc += y;
    }
    int run = inc - loc;
#pragma unroll
    for (int b = 0; b < PER; ++b) {
        const int c = bins[r * PER + b];
        bins[r * PER + b] = run;
        run += c;
    }
    __syncthreads();
    for (int k = r; k < n; k += L) {
        const uint64_t key = mcc_key(id[first + k], k);
        skey[bins[(int)(key >> SH)] + place[k]] = key;
    }
    __syncthreads();
}

// cell-local index of the electron of rank `rank` (0 <= rank < n) in the key order
template <int NB>
__device__ __forceinline__ int mcc_select(const uint64_t* skey, const int* bins, int n, int rank) {
    // the last bucket starting at or below rank (it holds rank: starts never decrease)
    int lo = 0, hi = NB - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (bins[mid] <= rank) lo = mid;
        else hi = mid - 1;
    }
    const int s0 = bins[lo], s1 = lo + 1 < NB ? bins[lo + 1] : n;
    const int want = rank - s0;
    int pick = s0;
    for (int j = s0; j < s1; ++j) {
        const uint64_t kj = skey[j];
        int c = 0;
        for (int i = s0; i < s1; ++i) c += (int)(skey[i] < kj);
        if (c == want) pick = j;
    }
    return (int)(skey[pick] & ((1u << kMccIndexBits) - 1));
}

}  // namespace

// the candidates of one cell by one wave; maxp = LDS slots for the cell's electrons
template <class UT, int REL, int NB>
__device__ void mcc_cell(const MccArgs<UT>& a, int cell, uint64_t* lds, int* picks, MccPair* pairs, int* npairs) {
    uint64_t* skey = lds;
    const int a0 = a.start[cell], eC = a.start[cell + 1] - a0;
    int* place = (int*)(skey + eC);
    int* bins = place + eC;
    const double smax = a.smax[cell];
    const double coll = a.rem[cell] + eC * a.nN * smax * a.dt;
    const int nc = (int)coll;
    __syncthreads();
    if (threadIdx.x == 0) a.rem[cell] = coll - nc;
    unsigned c_el = 0, c_ex = 0, c_iz = 0;
    double qtcr = 0.0;
    if (eC > 0 && nc > 0) {
        mcc_bucket<64, NB>(a.id_cell, a0, eC, skey, place, bins, (int)threadIdx.x);
        for (int base = 0; base < nc; base += 64) {
            const int k = base + (int)threadIdx.x;
            int pick = -1;
            if (k < nc) {
                double upick, uacc;
                mcc_rand(mcc_tag<REL>(), a.seed, a.step, cell, k, 0, upick, uacc);
                pick = (int)(upick * eC);
                if (pick >= eC) pick = eC - 1;
            }
            picks[threadIdx.x] = pick;
            if (threadIdx.x == 0) *npairs = 0;
            __syncthreads();
            if (k < nc) {
                bool first = true;
                for (int l = 0; l < (int)threadIdx.x; ++l)
                    if (picks[l] == pick) { first = false; break; }
                if (first) {
                    const int ei = slot_index(a.perm, a0 + mcc_select<NB>(skey, bins, eC, pick));
                    for (int l = (int)threadIdx.x; l < 64 && base + l < nc; ++l)
                        if (picks[l] == pick)
                            candidate<REL>(a, cell, base + l, ei, smax, qtcr, c_el, c_ex, pairs, npairs);
                }
            }
            __syncthreads();
            if constexpr (REL != kMccIon) {
                const int np = *npairs;
                if ((int)threadIdx.x < np) create_pair<REL>(a, pairs[threadIdx.x], c_iz);
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
        unsigned long long* cnt = a.counts + kNcollStride * (blockIdx.x % kNcollSlots) + (REL == kMccIon ? 4 : 0);
        if (nc) atomicAdd(cnt + 0, (unsigned long long)nc);
        if (v1) atomicAdd(cnt + 1, (unsigned long long)v1);
        if (v2) atomicAdd(cnt + 2, (unsigned long long)v2);
        if (v3) atomicAdd(cnt + 3, (unsigned long long)v3);
    }
}

// one wave per cell; cells with more electrons than a.maxp go to the large-cell pass.  Three
// waves per SIMD with a few spilled registers beat two without (measured on the 100^3 argon
// case: 6.7 against 8.4 ms per step)
template <class UT, int REL>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3, 8))) k_mcc(MccArgs<UT> a) {
    extern __shared__ uint64_t lds[];
    __shared__ int picks[64];
    __shared__ MccPair pairs[64];
    __shared__ int npairs;
    const int cell = blockIdx.x;
    if (a.start[cell + 1] - a.start[cell] > a.maxp) {
        if (threadIdx.x == 0) a.big_list[atomicAdd(a.big_n, 1)] = cell;
        return;
    }
    mcc_cell<UT, REL, kMccBins>(a, cell, lds, picks, pairs, &npairs);
}

// Several cells per wave: 64 / L cells, L lanes each (L = 8, 16 or 32), for cells of up to a.maxp_group
// electrons (larger ones go to the large-cell pass).  A cell draws a few candidates per
// step (about 5 of 100 electrons at the argon case's dt), so one cell per wave leaves most
// lanes idle while the wave pays the loads, the id ranking and the barriers for that cell
// alone.  Same candidate rule as mcc_cell, in rounds of L candidates per cell: a parcel
// still sees its candidates in candidate order, so the results do not depend on L.
template <class UT, int REL, int L>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3, 8))) k_mcc_group(MccArgs<UT> a) {
    constexpr int W = 64 / L, NB = 8 * L;
    extern __shared__ uint64_t lds[];
    __shared__ int picks_all[64];
    __shared__ MccPair pairs[64];
    __shared__ int npairs;
    const int w = threadIdx.x / L, r = threadIdx.x % L;
    const int maxp = a.maxp_group;
    uint64_t* skey = lds + (size_t)w * (maxp + maxp / 2 + NB / 2);
    int* place = (int*)(skey + maxp);
    int* bins = place + maxp;
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
    const int ncw = (on && eC > 0) ? nc : 0;
    unsigned c_el = 0, c_ex = 0, c_iz = 0;
    double qtcr = 0.0;
    mcc_bucket<L, NB>(a.id_cell, a0, ncw > 0 ? eC : 0, skey, place, bins, r);
    // rounds of L candidates; every group takes part in the barriers of the busiest one
    int rounds = (ncw + L - 1) / L;
    for (int off = 32; off > 0; off >>= 1) rounds = max(rounds, __shfl_xor(rounds, off));
    for (int base = 0; base < rounds * L; base += L) {
        const int k = base + r;
        int pick = -1;
        if (k < ncw) {
            double upick, uacc;
            mcc_rand(mcc_tag<REL>(), a.seed, a.step, cell, k, 0, upick, uacc);
            pick = (int)(upick * eC);
            if (pick >= eC) pick = eC - 1;
        }
        picks[r] = pick;
        if (threadIdx.x == 0) npairs = 0;
        __syncthreads();
        if (k < ncw) {
            bool first = true;
            for (int l = 0; l < r; ++l)
                if (picks[l] == pick) { first = false; break; }
            if (first) {
                const int ei = slot_index(a.perm, a0 + mcc_select<NB>(skey, bins, eC, pick));
                for (int l = r; l < L && base + l < ncw; ++l)
                    if (picks[l] == pick)
                        candidate<REL>(a, cell, base + l, ei, smax, qtcr, c_el, c_ex, pairs, &npairs);
            }
        }
        __syncthreads();
        if constexpr (REL != kMccIon) {
            const int np = npairs;
            if ((int)threadIdx.x < np) create_pair<REL>(a, pairs[threadIdx.x], c_iz);
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
        unsigned long long* cnt = a.counts + kNcollStride * (blockIdx.x % kNcollSlots) + (REL == kMccIon ? 4 : 0);
        if (v0) atomicAdd(cnt + 0, (unsigned long long)v0);
        if (v1) atomicAdd(cnt + 1, (unsigned long long)v1);
        if (v2) atomicAdd(cnt + 2, (unsigned long long)v2);
        if (v3) atomicAdd(cnt + 3, (unsigned long long)v3);
    }
}

template <class UT, int REL>
__global__ void __launch_bounds__(64) k_mcc_big(MccArgs<UT> a) {
    extern __shared__ uint64_t lds[];
    __shared__ int picks[64];
    __shared__ MccPair pairs[64];
    __shared__ int npairs;
    const int nbig = *a.big_n;
    for (int i = blockIdx.x; i < nbig; i += gridDim.x) {
        const int cell = a.big_list[i];
        if (a.start[cell + 1] - a.start[cell] > a.maxp_big) {
            if (threadIdx.x == 0) atomicOr(a.flags, kFlagCellOverflow);
            continue;
        }
        mcc_cell<UT, REL, kMccBinsBig>(a, cell, lds, picks, pairs, &npairs);
        __syncthreads();
    }
}

template <class UT, int REL>
static hipError_t launch_mcc_t(const MccArgs<UT>& a, hipStream_t s) {
    if (a.n_cells <= 0) return hipSuccess;
    // per cell: maxp keys (8 B) and placements (4 B), and the bucket starts (4 B each)
    if (a.group == 8) {
        const size_t lds = (size_t)8 * (a.maxp_group + a.maxp_group / 2 + 8 * 8 / 2) * sizeof(uint64_t);
        hipLaunchKernelGGL((k_mcc_group<UT, REL, 8>), dim3((a.n_cells + 7) / 8), dim3(64), lds, s, a);
    } else if (a.group == 16) {
        const size_t lds = (size_t)4 * (a.maxp_group + a.maxp_group / 2 + 8 * 16 / 2) * sizeof(uint64_t);
        hipLaunchKernelGGL((k_mcc_group<UT, REL, 16>), dim3((a.n_cells + 3) / 4), dim3(64), lds, s, a);
    } else if (a.group == 32) {
        const size_t lds = (size_t)2 * (a.maxp_group + a.maxp_group / 2 + 8 * 32 / 2) * sizeof(uint64_t);
        hipLaunchKernelGGL((k_mcc_group<UT, REL, 32>), dim3((a.n_cells + 1) / 2), dim3(64), lds, s, a);
    } else {
        const size_t lds = (size_t)a.maxp * (sizeof(uint64_t) + sizeof(int)) + kMccBins * sizeof(int);
        hipLaunchKernelGGL((k_mcc<UT, REL>), dim3(a.n_cells), dim3(64), lds, s, a);
    }
    const size_t lds_big = (size_t)a.maxp_big * (sizeof(uint64_t) + sizeof(int)) + kMccBinsBig * sizeof(int);
    const int nb = a.n_cells < 512 ? a.n_cells : 512;
    hipLaunchKernelGGL((k_mcc_big<UT, REL>), dim3(nb), dim3(64), lds_big, s, a);
    return hipGetLastError();
}
template <class UT>
static hipError_t launch_mcc_t(const MccArgs<UT>& a, hipStream_t s) {
    if (a.ion_neutral) return launch_mcc_t<UT, kMccIon>(a, s);
    return a.relativistic ? launch_mcc_t<UT, kMccRelativistic>(a, s) : launch_mcc_t<UT, kMccNanbu>(a, s);
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
