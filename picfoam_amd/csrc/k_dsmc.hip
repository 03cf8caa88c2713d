// Collisions between the parcels of a cell drawn as candidate pairs:
//  * BinaryCollisionModel (neutral DSMC, BinaryCollisionModel.C:229-425): pairs among the
//    neutral parcels with the no-time-counter count 1/2 n (n - 1) w (sigma_T cR)_max dt / V,
//    the second parcel from the first one's octant of the cell when it holds another parcel
//    (8 sub-cells about the cell centre), acceptance sigma_T cR / (sigma_T cR)_max,
//    IsotropicScattering (IsotropicScattering.C:61-105) and the weight correction; then the
//    neutrals with the background gas (:427-525);
//  * IonNeutralCollisionModel with neutral parcels (IonNeutralCollisionModel.C:580-713):
//    n_neutral n_ion w (sigma cR)_max dt / V candidate pairs, elastic (isotropic in the
//    centre of mass) or charge exchange by the cross sections of the ion energy.
//
// One wave per cell.  The cell's parcels are listed species after species, each in the
// order of a hash of its ids (a key sorted in LDS), so that a drawn index names the same
// parcel wherever the sort left it; their velocities are held in LDS as FP64 for the whole
// pass and written back once.  Candidates are drawn with counter-based draws keyed by
// (cell, step, candidate).  Each candidate reads and writes two parcels, so a later
// candidate that meets a parcel of an earlier one must see its result: the candidates of
// a round are given levels (one more than the last earlier candidate sharing a parcel) and
// run level by level, the candidates of one level in parallel.  The result is that of
// running the candidates in order.
//
// Cells are planned one per lane (parcel counts, candidate numbers and remainders); the
// wave then works through the cells that have candidates.  Cells with more parcels than
// the LDS list holds go to a large-cell pass.
#include <cmath>

#include "collide.h"

namespace picmi {

namespace {

constexpr int kPairCand = 256;      // candidates per round
constexpr double kVSmall = 1e-300;  // OpenFOAM's vSmall
constexpr int kKeySpeciesShift = 61, kKeyIndexBits = 12;
static_assert(kPairSlotsBig < (1 << kKeyIndexBits), "cell-local index must fit below the key's hash bits");
static_assert(kMaxSpecies <= 8, "species index must fit in the key's top three bits");

template <int MAXP>
struct PairLds {
    uint64_t key[MAXP];       // (species, id hash, cell-local index), then sorted
    double u[MAXP][3];        // velocities by rank
    int gidx[MAXP];           // array index of the parcel of rank r
    int sl[MAXP];             // sub-cell lists: the ranks of sub-cell 0, then 1, ...
    int last[MAXP];           // level of the round's last candidate that met rank r
    unsigned char sub[MAXP];
    unsigned char tch[MAXP];  // the velocity changed
    int sst[9];               // sub-cell list starts
    short cp[kPairCand], cq[kPairCand];
    unsigned short lv[kPairCand];
    int nlev;
};

struct Plan {
    int n, nA, nB, nc, ncb;
    double wA, wB, smax;
};

// per-lane counters: ion pairs (candidates, elastic, charge exchange) or binary pairs
// (candidates, collisions) and background (candidates, collisions)
struct Cnt {
    unsigned long long v[4];
};

// the parcel counts of a cell, the largest weights present in groups A and B
template <class UT>
__device__ Plan count_cell(const PairArgs<UT>& a, int cell) {
    Plan p{};
    for (int s = 0; s < a.ns; ++s) {
        const int c = a.start[s][cell + 1] - a.start[s][cell];
        if (s < a.na) {
            p.nA += c;
            if (c > 0 && a.w[s] > p.wA) p.wA = a.w[s];
        } else {
            p.nB += c;
            if (c > 0 && a.w[s] > p.wB) p.wB = a.w[s];
        }
    }
    p.n = p.nA + p.nB;
    return p;
}

// the candidate numbers of this step and the remainders they leave (written when `write`)
template <class UT, int MODE>
__device__ void decide(const PairArgs<UT>& a, int cell, Plan& p, bool write, Cnt* cnt) {
    p.smax = a.smax[cell];
    if constexpr (MODE == kPairIon) {
        // IonNeutralCollisionModel.C:640-652
        if (p.nA > 0 && p.nB > 0) {
            const double wm = p.wB > p.wA ? p.wB : p.wA;
            const double sel = a.rem[cell] + (double)(p.nA * p.nB) * wm * p.smax * a.dt / a.vol[cell];
            p.nc = (int)sel;
            if (write) a.rem[cell] = sel - p.nc;
        }
    } else {
        // BinaryCollisionModel.C:320-329
        if (p.nA > 1) {
            const double sel = a.rem[cell] + 0.5 * p.nA * (p.nA - 1) * p.wA * p.smax * a.dt / a.vol[cell];
            p.nc = (int)sel;
            if (write) a.rem[cell] = sel - p.nc;
        }
        // BinaryCollisionModel.C:470-476, with the maximum the pair pass leaves: unchanged
        // when it had no candidates (otherwise the cell's pass computes it)
        if (a.bg && p.nA > 0 && p.nc == 0) {
            const double sel = a.rem_bg[cell] + __dmul_rn(p.nA * a.nN * p.smax, a.dt);
            p.ncb = (int)sel;
            if (write) a.rem_bg[cell] = sel - p.ncb;
        }
    }
    if (cnt) {
        cnt->v[0] += (unsigned long long)p.nc;
        if (MODE == kPairDsmc) cnt->v[2] += (unsigned long long)p.ncb;
    }
}

__device__ __forceinline__ Plan shfl_plan(const Plan& p, int l) {
    Plan q;
    q.n = __shfl(p.n, l); q.nA = __shfl(p.nA, l); q.nB = __shfl(p.nB, l);
    q.nc = __shfl(p.nc, l); q.ncb = __shfl(p.ncb, l);
    q.wA = __shfl(p.wA, l); q.wB = __shfl(p.wB, l); q.smax = __shfl(p.smax, l);
    return q;
}

__device__ __forceinline__ vec3<double> ld_u(const double (&u)[3]) { return mk3<double>(u[0], u[1], u[2]); }
__device__ __forceinline__ void st_u(double (&u)[3], vec3<double> v) { u[0] = v.x; u[1] = v.y; u[2] = v.z; }

// IsotropicScattering::updateVelocity (IsotropicScattering.C:61-105; the ion model's,
// IonIsotropicScattering.C:60-100, is the same): the relative velocity turned to a random
// direction of the centre-of-mass frame, cos(theta) = 2 u1 - 1 along x, phi = 2 pi u2
__device__ __forceinline__ void iso_scatter(vec3<double>& P, vec3<double>& Q, double mP, double mQ, double u1,
                                            double u2) {
    const double ims = 1.0 / (mP + mQ);
    const vec3<double> cm = ims * (mP * P + mQ * Q);
    const double cR = mag(P - Q);
    const double ct = 2.0 * u1 - 1.0, st = sqrt(1.0 - ct * ct);
    double sp, cp;
    sincos_turn(u2, &sp, &cp);
    const vec3<double> rel = cR * mk3<double>(ct, st * cp, st * sp);
    P = cm + (mQ * ims) * rel;
    Q = cm - (mP * ims) * rel;
}

// sigma_T cR of the binary model: HardSphere (HardSphere.C:50-82) or VariableHardSphere
// (VariableHardSphere.C:51-100); zero below vSmall
template <class UT>
__device__ __forceinline__ double sigma_tcr(const PairArgs<UT>& a, const DsmcPair& t, vec3<double> d) {
    const double cR = mag(d);
    if (cR < kVSmall) return 0.0;
    if (!a.vhs) return t.hs * cR;
    return t.hs * pow(t.vn / (t.mr * cR * cR), t.ve) / t.gl * cR;
}

// the partner of candidate k of a cell from the background Maxwellian
// (uniformBackground::sampleVelocity, uniformBackground.C:121-127)
template <class UT>
__device__ __forceinline__ vec3<double> bg_velocity(const PairArgs<UT>& a, int cell, int k) {
    double ua, ub, n1, n2, n3, n4;
    mcc_rand(kTagDsmcBg, a.seed, a.step, cell, k, 1, ua, ub);
    box_muller(ua, ub, n1, n2);
    mcc_rand(kTagDsmcBg, a.seed, a.step, cell, k, 2, ua, ub);
    box_muller(ua, ub, n3, n4);
    return mk3<double>(a.vthN * n1 + a.drift[0], a.vthN * n2 + a.drift[1], a.vthN * n3 + a.drift[2]);
}

__device__ __forceinline__ int species_of(uint64_t key) { return (int)(key >> kKeySpeciesShift); }

// the relativistic kinetic energy [eV] of the relative motion at the ion mass
// (IonNeutralCollisionModel.C:666-670)
__device__ __forceinline__ double ion_energy_ev(double cR, double mI) {
    const double c = kC;
    const double gamma = 1.0 / sqrt(1.0 - ((cR * cR) / (c * c)));
    return (gamma - 1.0) * mI * c * c / kQe;
}

// candidate k of the ion pass between the neutral of rank P and the ion of rank Q
// (IonNeutralCollisionModel.C:655-701, elastic :438-449 with IonIsotropicScattering, charge
// exchange :401-415); draws 1 (acceptance, identity switch), 2 (angles), 3 (weight correction)
template <class UT, int MAXP>
__device__ void ion_pair(const PairArgs<UT>& a, PairLds<MAXP>& L, int cell, int k, int P, int Q, double smax,
                         double& qtcr, Cnt& cnt) {
    const int sP = species_of(L.key[P]), sQ = species_of(L.key[Q]);
    vec3<double> up = ld_u(L.u[P]), uq = ld_u(L.u[Q]);
    const double mN = a.mass[sP], mI = a.mass[sQ];
    const double cR = mag(uq - up);
    const double eV = ion_energy_ev(cR, mI);
    const double Qel = xs_lookup(a.xs_e[0], a.xs_s[0], a.xs_n[0], a.xs_guide[0], a.xs_le0[0], a.xs_idl[0], eV);
    const double Qcx = xs_lookup(a.xs_e[1], a.xs_s[1], a.xs_n[1], a.xs_guide[1], a.xs_le0[1], a.xs_idl[1], eV);
    const double Qt = Qel + Qcx;
    qtcr = fmax(qtcr, Qt * cR);
    double uacc, ucx;
    mcc_rand(kTagIonPair, a.seed, a.step, cell, k, 1, uacc, ucx);
    const double wP = a.w[sP], wQ = a.w[sQ];
    if (uacc <= Qel * cR / smax) {
        double u1, u2;
        mcc_rand(kTagIonPair, a.seed, a.step, cell, k, 2, u1, u2);
        const vec3<double> preP = up, preQ = uq;
        iso_scatter(up, uq, mN, mI, u1, u2);
        if (a.wc) {
            // NanbuWeightCorrectionModel::correctVelocity (NanbuCorrection.C:50-69)
            double uw, unused;
            mcc_rand(kTagIonPair, a.seed, a.step, cell, k, 3, uw, unused);
            if (uw > wP / wQ) uq = preQ;
            if (uw > wQ / wP) up = preP;
        }
        if (a.keep_neutral) up = preP;
        ++cnt.v[1];
    } else if (eV >= a.th_cx && uacc <= Qt * cR / smax) {
        const vec3<double> tmp = up;
        if (ucx <= wQ / wP) up = uq;
        if (ucx <= wP / wQ) uq = tmp;
        if (a.keep_neutral) up = tmp;
        ++cnt.v[2];
    } else {
        return;
    }
    st_u(L.u[P], up);
    st_u(L.u[Q], uq);
    L.tch[P] = 1;
    L.tch[Q] = 1;
}

// candidate k of the binary pass between ranks P and Q (BinaryCollisionModel.C:376-400);
// draws 0 (its acceptance is the second number), 1 (angles), 2 (weight correction)
template <class UT, int MAXP>
__device__ void dsmc_pair(const PairArgs<UT>& a, PairLds<MAXP>& L, int cell, int k, int P, int Q, double smax,
                          double& qtcr, Cnt& cnt) {
    const int sP = species_of(L.key[P]), sQ = species_of(L.key[Q]);
    vec3<double> up = ld_u(L.u[P]), uq = ld_u(L.u[Q]);
    const double s = sigma_tcr(a, a.tab[sP * kMaxSpecies + sQ], up - uq);
    qtcr = fmax(qtcr, s);
    double upick, uacc;
    mcc_rand(kTagDsmc, a.seed, a.step, cell, k, 0, upick, uacc);
    if (!(s / smax > uacc)) return;
    double u1, u2;
    mcc_rand(kTagDsmc, a.seed, a.step, cell, k, 1, u1, u2);
    const vec3<double> preP = up, preQ = uq;
    iso_scatter(up, uq, a.mass[sP], a.mass[sQ], u1, u2);
    if (a.wc) {
        double uw, unused;
        mcc_rand(kTagDsmc, a.seed, a.step, cell, k, 2, uw, unused);
        if (uw > a.w[sP] / a.w[sQ]) uq = preQ;
        if (uw > a.w[sQ] / a.w[sP]) up = preP;
    }
    st_u(L.u[P], up);
    st_u(L.u[Q], uq);
    L.tch[P] = 1;
    L.tch[Q] = 1;
    ++cnt.v[1];
}

// candidate k of the binary background pass on rank P (BinaryCollisionModel.C:478-505);
// draws 0 (pick, acceptance), 1 and 2 (the partner), 3 (angles)
template <class UT, int MAXP>
__device__ void dsmc_bg(const PairArgs<UT>& a, PairLds<MAXP>& L, int cell, int k, int P, double smax, double& qtcr,
                        Cnt& cnt) {
    const int sP = species_of(L.key[P]);
    vec3<double> up = ld_u(L.u[P]);
    vec3<double> un = bg_velocity(a, cell, k);
    const double s = sigma_tcr(a, a.tab[kMaxSpecies * kMaxSpecies + sP], up - un);
    qtcr = fmax(qtcr, s);
    double upick, uacc;
    mcc_rand(kTagDsmcBg, a.seed, a.step, cell, k, 0, upick, uacc);
    if (!(s / smax > uacc)) return;
    double u1, u2;
    mcc_rand(kTagDsmcBg, a.seed, a.step, cell, k, 3, u1, u2);
    iso_scatter(up, un, a.mass[sP], a.massN, u1, u2);
    st_u(L.u[P], up);
    L.tch[P] = 1;
    ++cnt.v[3];
}

// the picks of candidate k: the ion pass draws a neutral and an ion; the binary pass a
// first parcel, then a second one (not the first) from its sub-cell when that holds
// another parcel, else from the cell (BinaryCollisionModel.C:336-367); the background pass
// one parcel.  A redraw uses draw 16 + r; after 64 equal draws the next parcel of the list
// is taken.
template <class UT, int MODE, int MAXP>
__device__ void pick(const PairArgs<UT>& a, const PairLds<MAXP>& L, int cell, int k, int nA, int nB, bool bg, int& P,
                     int& Q) {
    double u0, u1;
    if constexpr (MODE == kPairIon) {
        mcc_rand(kTagIonPair, a.seed, a.step, cell, k, 0, u0, u1);
        P = min((int)(u0 * nA), nA - 1);
        Q = nA + min((int)(u1 * nB), nB - 1);
        return;
    }
    if (bg) {
        mcc_rand(kTagDsmcBg, a.seed, a.step, cell, k, 0, u0, u1);
        P = min((int)(u0 * nA), nA - 1);
        Q = -1;
        return;
    }
    mcc_rand(kTagDsmc, a.seed, a.step, cell, k, 0, u0, u1);
    P = min((int)(u0 * nA), nA - 1);
    Q = -1;
    const int sc = L.sub[P], s0 = L.sst[sc], nsc = L.sst[sc + 1] - s0;
    const int m = nsc > 1 ? nsc : nA;
    for (int r = 0; r < 64 && Q < 0; ++r) {
        double v0, v1;
        mcc_rand(kTagDsmc, a.seed, a.step, cell, k, 16 + r, v0, v1);
        const int j = min((int)(v0 * m), m - 1);
        const int q = nsc > 1 ? L.sl[s0 + j] : j;
        if (q != P) Q = q;
    }
    if (Q < 0) {
        if (nsc > 1) {
            int j = 0;
            while (L.sl[s0 + j] != P) ++j;
            Q = L.sl[s0 + (j + 1) % nsc];
        } else {
            Q = (P + 1) % nA;
        }
    }
}

// nc candidates of one kind (bg: the binary background pass) in rounds of kPairCand: picks,
// levels (one lane, in candidate order), then level by level
template <class UT, int MODE, int MAXP>
__device__ void run_candidates(const PairArgs<UT>& a, PairLds<MAXP>& L, int cell, const Plan& p, int nc, bool bg,
                               double smax, double& qtcr, Cnt& cnt) {
    const int lane = threadIdx.x;
    for (int base = 0; base < nc; base += kPairCand) {
        const int m = min(kPairCand, nc - base);
        for (int j = lane; j < m; j += 64) {
            int P, Q;
            pick<UT, MODE, MAXP>(a, L, cell, base + j, p.nA, p.nB, bg, P, Q);
            L.cp[j] = (short)P;
            L.cq[j] = (short)Q;
        }
        for (int r = lane; r < p.n; r += 64) L.last[r] = 0;
        __syncthreads();
        if (lane == 0) {
            int nl = 0;
            for (int j = 0; j < m; ++j) {
                const int P = L.cp[j], Q = L.cq[j];
                int l = L.last[P];
                if (Q >= 0 && L.last[Q] > l) l = L.last[Q];
                ++l;
                L.last[P] = l;
                if (Q >= 0) L.last[Q] = l;
                L.lv[j] = (unsigned short)l;
                nl = l > nl ? l : nl;
            }
            L.nlev = nl;
        }
        __syncthreads();
        const int nlev = L.nlev;
        for (int lev = 1; lev <= nlev; ++lev) {
            for (int j = lane; j < m; j += 64) {
                if (L.lv[j] != lev) continue;
                const int k = base + j;
                if constexpr (MODE == kPairIon) ion_pair<UT, MAXP>(a, L, cell, k, L.cp[j], L.cq[j], smax, qtcr, cnt);
                else if (bg) dsmc_bg<UT, MAXP>(a, L, cell, k, L.cp[j], smax, qtcr, cnt);
                else dsmc_pair<UT, MAXP>(a, L, cell, k, L.cp[j], L.cq[j], smax, qtcr, cnt);
            }
            __syncthreads();
        }
    }
}

__device__ __forceinline__ double wave_max(double v) {
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off));
    return v;
}

// one cell with candidates, by the whole wave
template <class UT, int MODE, int MAXP>
__device__ void pair_cell(const PairArgs<UT>& a, int cell, const Plan& p, PairLds<MAXP>& L, Cnt& cnt) {
    const int lane = threadIdx.x;
    const int n = p.n;
    int off[kMaxSpecies + 1], st[kMaxSpecies];
    off[0] = 0;
    for (int s = 0; s < a.ns; ++s) {
        st[s] = a.start[s][cell];
        off[s + 1] = off[s] + a.start[s][cell + 1] - st[s];
    }
    // keys: (species, top 49 bits of the id hash, cell-local index), padded to a power of two
    int np = 2;
    while (np < n) np <<= 1;
    for (int j = lane; j < np; j += 64) {
        uint64_t key = ~0ull;
        if (j < n) {
            int s = 0;
            while (j >= off[s + 1]) ++s;
            const int kl = j - off[s];
            const uint32_t id = a.id_cell[s][st[s] + kl];
            key = ((uint64_t)s << kKeySpeciesShift) | ((id_mix(id) >> 15) << kKeyIndexBits) | (uint64_t)kl;
        }
        L.key[j] = key;
    }
    __syncthreads();
    // bitonic sort of the np keys
    for (int size = 2; size <= np; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int t = lane; t < (np >> 1); t += 64) {
                const int lo = 2 * t - (t & (stride - 1)), hi = lo + stride;
                const bool up = (lo & size) == 0;
                const uint64_t x = L.key[lo], y = L.key[hi];
                if ((x > y) == up) { L.key[lo] = y; L.key[hi] = x; }
            }
            __syncthreads();
        }
    }
    // the parcels by rank: array index, velocity, sub-cell (the octant of its position
    // about the cell centre, BinaryCollisionModel.C:296-311)
    const bool subcells = MODE == kPairDsmc && p.nc > 0;
    const double* Cc = a.C + 3 * (size_t)cell;
    for (int r = lane; r < n; r += 64) {
        const uint64_t key = L.key[r];
        const int s = species_of(key);
        const int kl = (int)(key & ((1u << kKeyIndexBits) - 1));
        const int slot = st[s] + kl;
        const int gi = a.perm[s] ? a.perm[s][slot] : slot;
        const PartSoA<UT>& ps = a.p[s];
        L.gidx[r] = gi;
        L.u[r][0] = (double)ps.ux[gi];
        L.u[r][1] = (double)ps.uy[gi];
        L.u[r][2] = (double)ps.uz[gi];
        L.tch[r] = 0;
        if (subcells) {
            const TetGeo& g = a.tets[ps.tet[gi]];
            const double y1 = ps.y1[gi], y2 = ps.y2[gi], y3 = ps.y3[gi];
            const double y0 = 1.0 - (y1 + y2 + y3);
            const double* p1 = a.pts + 3 * (size_t)g.vid[0];
            const double* p2 = a.pts + 3 * (size_t)g.vid[1];
            const double* p3 = a.pts + 3 * (size_t)g.vid[2];
            // an empty direction counts as zero (the upper side): the parcels of a
            // reduced-dimension case sit on the mesh's mid-plane, where the sign of the
            // difference is rounding
            int sc = 0;
            for (int d = 0; d < 3; ++d) {
                const double x = (y0 * Cc[d] + y1 * p1[d]) + (y2 * p2[d] + y3 * p3[d]);
                sc |= (((a.empty_mask >> d) & 1u) || x - Cc[d] >= 0.0 ? 1 : 0) << d;
            }
            L.sub[r] = (unsigned char)sc;
        }
    }
    __syncthreads();
    if (subcells) {
        // the ranks of each sub-cell in rank order
        const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
        int at = 0;
        for (int sc = 0; sc < 8; ++sc) {
            if (lane == 0) L.sst[sc] = at;
            for (int r0 = 0; r0 < n; r0 += 64) {
                const int r = r0 + lane;
                const bool in = r < n && L.sub[r] == sc;
                const uint64_t b = __ballot(in);
                if (in) L.sl[at + __popcll(b & below)] = r;
                at += __popcll(b);
            }
        }
        if (lane == 0) L.sst[8] = at;
        __syncthreads();
    }
    double qtcr = 0.0;
    if (p.nc > 0) run_candidates<UT, MODE, MAXP>(a, L, cell, p, p.nc, false, p.smax, qtcr, cnt);
    double smax = fmax(p.smax, wave_max(qtcr));
    if constexpr (MODE == kPairDsmc) {
        // the background pass, with the maximum the pair pass left
        // (BinaryCollisionModel.C:466-476)
        int ncb = p.ncb;
        if (a.bg && p.nc > 0) {
            const double sel = a.rem_bg[cell] + __dmul_rn(p.nA * a.nN * smax, a.dt);
            ncb = (int)sel;
            if (lane == 0) {
                a.rem_bg[cell] = sel - ncb;
                cnt.v[2] += (unsigned long long)ncb;
            }
        }
        if (ncb > 0) {
            double q2 = 0.0;
            run_candidates<UT, MODE, MAXP>(a, L, cell, p, ncb, true, smax, q2, cnt);
            smax = fmax(smax, wave_max(q2));
        }
    }
    if (lane == 0 && smax > p.smax) a.smax[cell] = smax;
    // velocities back, once
    for (int r = lane; r < n; r += 64) {
        if (!L.tch[r]) continue;
        const PartSoA<UT>& ps = a.p[species_of(L.key[r])];
        const int gi = L.gidx[r];
        ps.ux[gi] = (UT)L.u[r][0];
        ps.uy[gi] = (UT)L.u[r][1];
        ps.uz[gi] = (UT)L.u[r][2];
    }
    __syncthreads();
}

template <int MODE>
__device__ void flush_counts(const unsigned long long* c0, unsigned long long* counts) {
    unsigned long long v[4] = {c0[0], c0[1], c0[2], c0[3]};
    for (int off = 32; off > 0; off >>= 1)
        for (int j = 0; j < 4; ++j) v[j] += __shfl_xor(v[j], off);
    if (threadIdx.x == 0) {
        unsigned long long* c = counts + kNcollStride * (blockIdx.x % kNcollSlots) +
                                (MODE == kPairIon ? kPairWordIon : kPairWordDsmc);
        const int nw = MODE == kPairIon ? 3 : 4;
        for (int j = 0; j < nw; ++j)
            if (v[j]) atomicAdd(c + j, v[j]);
    }
}

}  // namespace

// 64 cells per wave: each lane plans one cell, then the wave runs the cells with
// candidates; cells of more than kPairSlots parcels go to the large-cell pass
template <class UT, int MODE>
__global__ void __launch_bounds__(64) k_pairs(PairArgs<UT> a) {
    __shared__ PairLds<kPairSlots> L;
    const int cell = blockIdx.x * 64 + (int)threadIdx.x;
    Cnt cnt{};
    Plan p{};
    bool work = false;
    if (cell < a.n_cells) {
        p = count_cell(a, cell);
        if (p.n > kPairSlots) {
            a.big_list[atomicAdd(a.big_n, 1)] = cell;
        } else {
            decide<UT, MODE>(a, cell, p, true, &cnt);
            work = p.nc > 0 || p.ncb > 0;
        }
    }
    uint64_t todo = __ballot(work);
    while (todo) {
        const int l = __ffsll((unsigned long long)todo) - 1;
        todo &= todo - 1;
        const Plan q = shfl_plan(p, l);
        pair_cell<UT, MODE, kPairSlots>(a, blockIdx.x * 64 + l, q, L, cnt);
    }
    flush_counts<MODE>(cnt.v, a.counts);
}

// the cells left by k_pairs, one wave each; cells beyond kPairSlotsBig parcels raise the
// cell-overflow flag and are left unchanged
template <class UT, int MODE>
__global__ void __launch_bounds__(64) k_pairs_big(PairArgs<UT> a) {
    __shared__ PairLds<kPairSlotsBig> L;
    Cnt cnt{};
    const int nbig = *a.big_n;
    for (int i = blockIdx.x; i < nbig; i += gridDim.x) {
        const int cell = a.big_list[i];
        Plan p = count_cell(a, cell);
        if (p.n > kPairSlotsBig) {
            if (threadIdx.x == 0) atomicOr(a.flags, kFlagCellOverflow);
            continue;
        }
        decide<UT, MODE>(a, cell, p, threadIdx.x == 0, threadIdx.x == 0 ? &cnt : nullptr);
        if (p.nc > 0 || p.ncb > 0) pair_cell<UT, MODE, kPairSlotsBig>(a, cell, p, L, cnt);
    }
    flush_counts<MODE>(cnt.v, a.counts);
}

template <class UT, int MODE>
static hipError_t launch_pairs_t(const PairArgs<UT>& a, hipStream_t s) {
    if (a.n_cells <= 0) return hipSuccess;
    hipLaunchKernelGGL((k_pairs<UT, MODE>), dim3((a.n_cells + 63) / 64), dim3(64), 0, s, a);
    const int nb = a.n_cells < 256 ? a.n_cells : 256;
    hipLaunchKernelGGL((k_pairs_big<UT, MODE>), dim3(nb), dim3(64), 0, s, a);
    return hipGetLastError();
}
hipError_t launch_pairs(const PairArgs<double>& a, int mode, hipStream_t s) {
    return mode == kPairIon ? launch_pairs_t<double, kPairIon>(a, s) : launch_pairs_t<double, kPairDsmc>(a, s);
}
hipError_t launch_pairs(const PairArgs<float>& a, int mode, hipStream_t s) {
    return mode == kPairIon ? launch_pairs_t<float, kPairIon>(a, s) : launch_pairs_t<float, kPairDsmc>(a, s);
}

}  // namespace picmi
