// Binary Coulomb collisions (CoulombCollisionModel::handleCollisions,
// CoulombCollisionModel.C:241-245): Nanbu (NanbuCollision.C:60-159) or Perez
// (PerezRelativisticCollision.C:62-219) scattering, Takizuka-Abe (TakizukaAbePairing.C:60-338)
// or Nanbu-Yonemura (NanbuYonemuraPairing.C:56-322) pairing, optional Nanbu weight
// correction (NanbuCorrection.C:50-69).
//
// One wave per cell; parcels are contiguous per species because the arrays are sorted by
// cell.  The reference's shuffle (OpenFOAM Random::shuffle) becomes a ranking of the cell's
// parcels by a counter-based random key of (id, step, stage, species), so the pairing does
// not depend on the order the parcels sit in memory; every collision draws its uniforms
// from a counter (cell, step, stage, collision index).
//
// Each pairing stage is described by a small table in LDS and split into work units: a
// unit is a run of collisions that must happen in order (the Takizuka-Abe odd triple, a
// parcel of the smaller species meeting several parcels of the larger one); units touch
// disjoint parcels and run on separate lanes.  All collisions go through one call site.
#include "device.h"
#include "kernels.h"

namespace picmi {

namespace {

// NanbuCollision.C:137-159 / PerezRelativisticCollision.C:198-219 in the arithmetic M.  The
// three regimes are evaluated for every lane with one exp and one log, so lanes of
// different s do not diverge:
//   (1/A) log(exp(-A) + 2 U sinh(A)) = 1 + (1/A) log(U + (1 - U) exp(-2A)),
// with 1/A = invA on the polynomial branch (s < 3) and A = 3 exp(-s) below s = 6.  exp(-s)
// is evaluated only when some lane of the wave needs it.
template <class M>
__device__ __forceinline__ M cos_chi(M s, M U) {
    const bool small = s < (M)0.1, mid = s < (M)3, big = s < (M)6;
    const M invA = (M)0.00569578 +
                   ((M)0.95602 + ((M)-0.508139 + ((M)0.479139 + ((M)-0.12789 + (M)0.0238957 * s) * s) * s) * s) * s;
    M Abig = (M)1;
    if (__any(!mid && big)) Abig = (M)3 * exp_small(-s);
    const M r = quot((M)1, mid ? invA : Abig);
    M A = mid ? r : Abig;
    const M coef = mid ? invA : r;
    if (small || !big) A = (M)0;  // keeps the unused branch finite
    const M e2 = exp_small((M)-2 * A);
    const M Us = U < (M)0.0001 ? (M)0.0001 : U;
    const M arg = small ? Us : U + ((M)1 - U) * e2;
    const M L = log_pos(arg);
    const M cc = small ? (M)1 + s * L : (big ? (M)1 + coef * L : (M)2 * U - (M)1);
    // rounding can put the value an ulp outside [-1, 1] (U near 0 or 1), which would make
    // sin(chi) = sqrt(1 - cos^2) a NaN; the reference formula has the same edge
    return cc > (M)1 ? (M)1 : (cc < (M)-1 ? (M)-1 : cc);
}


__device__ __forceinline__ double wave_sum(double v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

enum : int { kIntraEven = 0, kIntraOddTA = 1, kIntraOddNY = 2, kInterEqual = 3, kCyclic = 4 };

// one pairing stage: P and Q species velocity arrays, the shuffled index lists (LDS) and
// the densities of the collide() call.  kCyclic (unequal inter-species counts): P is the
// larger species; unit m < g1_units is parcel g1_yo + m of Q meeting P parcels
// g1_xo + l, l = m, m + g1_stride, ... < g1_lmax; the remaining units form group 2 alike.
struct StageK {
    void* u[6];
    int cP, cQ;              // cached velocities: x component offsets of P and Q in the cell's LDS cache
    int permP, permQ;        // offsets of the P and Q index lists from the first list in LDS
    double mP, qP, mQ, qQ, wP, wQ;
    double nP, nQ, nPQ, lnL;
    // Nanbu constants: cst = (qP qQ / eps0 mu)^2 (nP nQ / nPQ) dt / (4 pi), fP = mQ/(mP+mQ),
    // fQ = mP/(mP+mQ), b0c = qP qQ / (2 pi eps0 mu), hbc = h / (2 mu)
    double cst, fP, fQ, b0c, hbc;
    int kind, n, units, stage;
    int heavy, hw, lw;       // the first `heavy` units hold hw collisions each, the others lw
    int g1_units, g1_lmax, g1_stride, g1_xo, g1_yo, g1_ko;
    int g2_lmax, g2_stride, g2_xo, g2_yo, g2_ko;
};

// the stage's Nanbu constants from the pair table (computed once on the host)
__device__ __forceinline__ void finish_stage(StageK& d, double dt, const PairConst& pc) {
    d.cst = pc.cst0 * (d.nQ * d.nP / d.nPQ) * dt;
    d.fP = pc.fP;
    d.fQ = pc.fQ;
    d.b0c = pc.b0c;
    d.hbc = pc.hbc;
}

// collision t of unit u; false once the unit is exhausted
__device__ __forceinline__ bool decode(const StageK& d, const int* lists, int u, int t, int& iP, int& iQ, int& k) {
    const int* P = lists + d.permP;
    const int* Q = lists + d.permQ;
    switch (d.kind) {
        case kIntraEven:
            if (t) return false;
            iP = P[2 * u]; iQ = P[2 * u + 1]; k = u;
            return true;
        case kIntraOddTA:  // TakizukaAbePairing.C:115-139: the first three collide cyclically
            if (u == 0) {
                if (t >= 3) return false;
                iP = P[t == 1 ? 1 : 0]; iQ = P[t == 0 ? 1 : 2]; k = t;
                return true;
            }
            if (t) return false;
            iP = P[2 * u + 1]; iQ = P[2 * u + 2]; k = u + 2;
            return true;
        case kIntraOddNY:  // NanbuYonemuraPairing.C:127-152: pairs, then last with first
            if (t == 0) { iP = P[2 * u]; iQ = P[2 * u + 1]; k = u; return true; }
            if (u == 0 && t == 1) { iP = P[d.n - 1]; iQ = P[0]; k = (d.n - 1) >> 1; return true; }
            return false;
        case kInterEqual:
            if (t) return false;
            iP = P[u]; iQ = Q[d.n - 1 - u]; k = u;
            return true;
        default: {
            int l, m, xo, yo, ko;
            if (u < d.g1_units) {
                m = u; l = u + t * d.g1_stride;
                if (l >= d.g1_lmax) return false;
                xo = d.g1_xo; yo = d.g1_yo; ko = d.g1_ko;
            } else {
                m = u - d.g1_units; l = m + t * d.g2_stride;
                if (l >= d.g2_lmax) return false;
                xo = d.g2_xo; yo = d.g2_yo; ko = d.g2_ko;
            }
            iP = P[xo + l]; iQ = Q[yo + m]; k = ko + l;
            return true;
        }
    }
}

// number of collisions of unit u (the t for which decode returns true)
__device__ __forceinline__ int unit_count(const StageK& d, int u) {
    switch (d.kind) {
        case kIntraOddTA: return u == 0 ? 3 : 1;
        case kIntraOddNY: return u == 0 ? 2 : 1;
        case kIntraEven:
        case kInterEqual: return 1;
        default: {
            const bool g1 = u < d.g1_units;
            const int m = g1 ? u : u - d.g1_units;
            const int lmax = g1 ? d.g1_lmax : d.g2_lmax, stride = g1 ? d.g1_stride : d.g2_stride;
            return lmax > m ? (lmax - m + stride - 1) / stride : 0;
        }
    }
}

// NanbuCollision::collide (NanbuCollision.C:60-134) on two velocities, arithmetic M
template <class M>
__device__ __forceinline__ void nanbu_scatter(vec3<M>& uP, vec3<M>& uQ, const StageK& d, int calc_bmin, double debye,
                                              double u1, double u2) {
    const vec3<M> g = uP - uQ;
    const M g2 = dot(g, g), gp2 = g.y * g.y + g.z * g.z;
    if (!(g2 > (M)0) || !(gp2 > (M)0)) return;  // the reference divides by zero here
    const M ig = rsqrt(g2), igp = rsqrt(gp2);
    const M gm = g2 * ig, gp = gp2 * igp;
    M lnL;
    if (d.lnL == 0.0) {
        const M ig2 = ig * ig;
        M b0 = (M)d.b0c * ig2;
        if (calc_bmin) {
            const M hb = (M)d.hbc / gm;
            b0 = hb > fabs(b0) ? hb : fabs(b0);
        }
        const M l = (M)0.5 * log_pos((M)1 + ((M)debye * (M)debye) / (b0 * b0));
        lnL = l > (M)2 ? l : (M)2;
    } else {
        lnL = (M)d.lnL;
    }
    // s = lnL/(4 pi) (qP qQ / eps0 mu)^2 (nP nQ / nPQ) g^-3 dt
    const M s = lnL * (M)d.cst * (ig * ig * ig);
    const M cc = cos_chi<M>(s, (M)u1);
    const M sc2 = (M)1 - cc * cc;
    const M sc = sc2 > (M)0 ? sc2 * rsqrt(sc2) : (M)0;
    M se, ce;
    sincos_turn((M)u2, &se, &ce);
    const vec3<M> h = mk3<M>(gp * ce, -(g.y * g.x * ce + gm * g.z * se) * igp, -(g.z * g.x * ce - gm * g.y * se) * igp);
    const vec3<M> dv = ((M)1 - cc) * g + sc * h;
    uP = uP - (M)d.fP * dv;
    uQ = uQ + (M)d.fQ * dv;
}

// PerezRelativisticCollision::collide (PerezRelativisticCollision.C:62-196), FP64
__device__ __forceinline__ void perez_scatter(vec3<double>& UP, vec3<double>& UQ, const StageK& d, int calc_bmin,
                                              double debye, double dt, double u1, double u2) {
    const double c2 = kC * kC;
    const double mP = d.mP, mQ = d.mQ;
    ComFrame f;
    com_enter(UP, UQ, mP, mQ, f);
    const double pm = f.pm;
    if (pm == 0.0) return;
    const double t2 = 1.0 + mP * f.gPs * mQ * f.gQs * c2 / (pm * pm);
    double lnL;
    if (d.lnL == 0.0) {
        double b0 = d.qP * d.qQ / (4.0 * kPi * kEps0 * c2) * f.gc / f.msum * t2;
        if (calc_bmin) {
            const double hb = kH / (2.0 * pm);
            b0 = hb > fabs(b0) ? hb : fabs(b0);
        }
        const double l = 0.5 * log_pos(1.0 + (debye * debye) / (b0 * b0));
        lnL = l > 2.0 ? l : 2.0;
    } else {
        lnL = d.lnL;
    }
    const double qq = d.qP * d.qQ;
    double s = (d.nP * d.nQ / d.nPQ) * dt * lnL * qq * qq /
               (4.0 * kPi * kEps0 * kEps0 * (c2 * c2) * mP * f.gP * mQ * f.gQ) * (f.gc * pm / f.msum) * (t2 * t2);
    // low-temperature limit of s; 1.61199... = (4 pi / 3)^(1/3)
    const double vrel = pm * f.msum / (mP * f.gPs * mQ * f.gQs * f.gc);
    const double dP = mP * pow(d.nP, 2.0 / 3.0), dQ = mQ * pow(d.nQ, 2.0 / 3.0);
    const double smax = 1.6119919540164696 * (d.nP * d.nQ / d.nPQ) * dt * (mP + mQ) / (dP > dQ ? dP : dQ) * vrel;
    if (smax < s) s = smax;
    const double cc = cos_chi<double>(s, u1);
    const double sc = sqrt(1.0 - cc * cc);
    double se, ce;
    sincos_turn(u2, &se, &ce);
    com_leave(f, mP, mQ, cc, sc, se, ce, UP, UQ);
}

// one collision of parcel iP (species of d's P side) with parcel iQ; MODEL 0 Nanbu, 1 Perez;
// WC: Nanbu weight correction active
// CU: velocities in the LDS cache (component stride cs), else in the species arrays
template <class UT, int MODEL, bool WC, bool CU>
__device__ __forceinline__ void collide(const CoulombArgs& a, const StageK& d, int cell, int iP, int iQ, int k,
                                        double debye, UT* cache, int cs) {
    uint4 o = rng4(a.seed, (uint32_t)cell, (uint32_t)a.step, ((uint32_t)d.stage << 24) | (uint32_t)k,
                   kTagCoulomb << 24);
    const double u1 = u01(o.x, o.y), u2 = u01(o.z, o.w);
    // the table sits in LDS, so the arrays' address space is stated for global loads
    typedef __attribute__((address_space(1))) UT GUT;
    UT* cP = cache + d.cP;
    UT* cQ = cache + d.cQ;
    GUT* gP[3] = {(GUT*)d.u[0], (GUT*)d.u[1], (GUT*)d.u[2]};
    GUT* gQ[3] = {(GUT*)d.u[3], (GUT*)d.u[4], (GUT*)d.u[5]};
    const vec3<UT> preP = CU ? mk3<UT>(cP[iP], cP[cs + iP], cP[2 * cs + iP]) : mk3<UT>(gP[0][iP], gP[1][iP], gP[2][iP]);
    const vec3<UT> preQ = CU ? mk3<UT>(cQ[iQ], cQ[cs + iQ], cQ[2 * cs + iQ]) : mk3<UT>(gQ[0][iQ], gQ[1][iQ], gQ[2][iQ]);
    vec3<UT> nP = preP, nQ = preQ;
    if (MODEL == 1) {
        vec3<double> dP = mk3<double>(preP.x, preP.y, preP.z), dQ = mk3<double>(preQ.x, preQ.y, preQ.z);
        perez_scatter(dP, dQ, d, a.calculate_bmin, debye, a.dt, u1, u2);
        nP = mk3<UT>((UT)dP.x, (UT)dP.y, (UT)dP.z);
        nQ = mk3<UT>((UT)dQ.x, (UT)dQ.y, (UT)dQ.z);
    } else {
        nanbu_scatter<UT>(nP, nQ, d, a.calculate_bmin, debye, u1, u2);
    }
    if (WC) {
        // NanbuWeightCorrectionModel::correctVelocity
        o = rng4(a.seed, (uint32_t)cell, (uint32_t)a.step, ((uint32_t)d.stage << 24) | (uint32_t)k,
                 (kTagCoulomb << 24) | 1u);
        const double r = u01(o.x, o.y);
        if (r > d.wP / d.wQ) nQ = preQ;
        if (r > d.wQ / d.wP) nP = preP;
    }
    if (CU) {
        cP[iP] = nP.x; cP[cs + iP] = nP.y; cP[2 * cs + iP] = nP.z;
        cQ[iQ] = nQ.x; cQ[cs + iQ] = nQ.y; cQ[2 * cs + iQ] = nQ.z;
    } else {
        gP[0][iP] = nP.x; gP[1][iP] = nP.y; gP[2][iP] = nP.z;
        gQ[0][iQ] = nQ.x; gQ[1][iQ] = nQ.y; gQ[2][iQ] = nQ.z;
    }
}

// bitonic sort of npow2 64-bit keys with int payload in LDS by one wave
__device__ void bitonic_sort(uint64_t* key, int* val, int npow2) {
    for (int k = 2; k <= npow2; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = threadIdx.x; i < npow2; i += blockDim.x) {
                const int l = i ^ j;
                if (l > i) {
                    const bool up = (i & k) == 0;
                    const uint64_t x = key[i], y = key[l];
                    if ((x > y) == up) {
                        key[i] = y; key[l] = x;
                        const int t = val[i]; val[i] = val[l]; val[l] = t;
                    }
                }
            }
            __syncthreads();
        }
    }
}

// The shuffle key of a parcel is 53 random bits; the sort keys carry the parcel's cell-local
// index in the 11 bits below it (cells of up to kCoulombSlotsBig = 2048 parcels of a
// species), so sort keys are distinct even when ids repeat, and equal random keys order by
// the cell-local index.
constexpr int kKeyIndexBits = 11;
static_assert(kCoulombSlotsBig <= (1 << kKeyIndexBits), "cell-local index must fit below the shuffle key");

__device__ __forceinline__ uint64_t shuffle_key(uint64_t seed, int64_t step, int stage, int species, uint32_t pid) {
    const uint4 o = rng4(seed, pid, (uint32_t)step, ((uint32_t)stage << 8) | (uint32_t)species, kTagShuffle << 24);
    return ((uint64_t)o.x << 21) | (o.y >> 11);
}

// val[0..n) = the cell's parcels of one species ordered by their 53-bit random key: base +
// cell-local index, or the array index through perm (perm non-null: base is unused); id
// holds the ids in cell order (id[first + k] for the cell's parcel k).  Sort keys are
// distinct (the cell-local index sits below the random key), so up to 256 parcels each
// parcel's place is the number of smaller keys (no barriers in the count; eight keys per
// LDS round trip).  Larger cells use a bitonic sort.
__device__ void keyed_perm(uint64_t seed, int64_t step, int stage, int species, const uint32_t* id, int first, int n,
                           uint64_t* key, int* val, int base, const int* perm) {
    if (n <= 256) {
        const int n8 = (n + 7) & ~7;
        for (int k = threadIdx.x; k < n8; k += blockDim.x)
            key[k] = k < n ? (shuffle_key(seed, step, stage, species, id[first + k]) << kKeyIndexBits) | (uint64_t)k : ~0ull;
        __syncthreads();
        // rounds of up to 64 keys; a partial round of E keys gives each key G = 64 / E lanes
        // (rounded down to a power of two) that count over G slices of the list and add up
        for (int k0 = 0; k0 < n; k0 += 64) {
            const int E = n - k0 < 64 ? n - k0 : 64;
            int lg = 0;
            while ((E << (lg + 1)) <= 64) ++lg;
            const int lane = (int)threadIdx.x;
            const int e = lane >> lg, g = lane & ((1 << lg) - 1);
            const int k = k0 + e;
            const bool act = e < E;
            const uint64_t kk = act ? key[k] : 0ull;
            // slice length: ceil(n8 / G) rounded up to whole 8-key reads, so the G slices
            // cover the list (G * seg >= n8) also when n8 < G
            const int seg = ((((n8 + (1 << lg) - 1) >> lg) + 7) & ~7);
            const int j0 = g * seg;
            const int j1 = j0 + seg < n8 ? j0 + seg : n8;
            int r = 0;
            for (int j = j0; j < j1; j += 8) {
                const uint64_t* q = key + j;
                const uint64_t x0 = q[0], x1 = q[1], x2 = q[2], x3 = q[3], x4 = q[4], x5 = q[5], x6 = q[6], x7 = q[7];
                r += (int)(x0 < kk) + (int)(x1 < kk) + (int)(x2 < kk) + (int)(x3 < kk) + (int)(x4 < kk) +
                     (int)(x5 < kk) + (int)(x6 < kk) + (int)(x7 < kk);
            }
            for (int off = (1 << lg) >> 1; off > 0; off >>= 1) r += __shfl_xor(r, off);
            if (act && g == 0) val[r] = perm ? perm[first + k] : base + k;
        }
        __syncthreads();
        return;
    }
    int np2 = 1;
    while (np2 < n) np2 <<= 1;
    for (int k = threadIdx.x; k < np2; k += blockDim.x) {
        if (k < n) {
            key[k] = (shuffle_key(seed, step, stage, species, id[first + k]) << kKeyIndexBits) | (uint64_t)k;
            val[k] = perm ? perm[first + k] : base + k;
        } else {
            key[k] = ~0ull;
            val[k] = -1;
        }
    }
    __syncthreads();
    bitonic_sort(key, val, np2);
}

__device__ __forceinline__ double dmin(double x, double y) { return x < y ? x : y; }

}  // namespace

// largest charged-species parcel count of the cell
__device__ __forceinline__ int cell_max_count(const CoulombArgs& a, int cell) {
    int m = 0;
    for (int s = 0; s < a.n_species; ++s) {
        if (a.charge[s] == 0.0) continue;
        const int n = a.start[s][cell + 1] - a.start[s][cell];
        m = n > m ? n : m;
    }
    return m;
}

// next charged species after i, or -1
__device__ __forceinline__ int next_charged(const CoulombArgs& a, int i) {
    for (int s2 = i + 1; s2 < a.n_species; ++s2)
        if (a.charge[s2] != 0.0) return s2;
    return -1;
}

// table of the intra-species stage of species spc (cnt >= 2 parcels in a cell of volume Vc);
// its shuffled parcel list sits at offset perm of the stage lists; velocities in the arrays
__device__ __forceinline__ void stage_intra(StageK& d, const CoulombArgs& a, int spc, int cnt, int stage, double Vc,
                                            bool ny, int perm) {
    d.u[0] = d.u[3] = a.ux[spc]; d.u[1] = d.u[4] = a.uy[spc]; d.u[2] = d.u[5] = a.uz[spc];
    d.permP = d.permQ = perm;
    d.mP = d.mQ = a.mass[spc]; d.qP = d.qQ = a.charge[spc]; d.wP = d.wQ = a.weight[spc];
    d.lnL = a.coulomb_log[spc][spc];
    d.n = cnt;
    d.stage = stage;
    const double n = cnt * a.weight[spc] / Vc;
    d.nP = d.nQ = n;
    d.lw = 1;
    if (cnt % 2 == 0) {
        d.kind = kIntraEven; d.units = cnt / 2; d.nPQ = n;
        d.heavy = 0; d.hw = 1;
    } else if (!ny) {
        d.kind = kIntraOddTA; d.units = 1 + (cnt - 3) / 2; d.nPQ = (cnt + 3.0) * a.weight[spc] / Vc;
        d.heavy = 1; d.hw = 3;
    } else {
        d.kind = kIntraOddNY; d.units = (cnt - 1) / 2;
        d.nPQ = 2.0 * a.weight[spc] * ((cnt - 1) / 2 + 1) / Vc;
        d.heavy = 1; d.hw = 2;
    }
    finish_stage(d, a.dt, a.pair[spc * kMaxSpecies + spc]);
}

// table of the stage between species i (na parcels, list at offset permA) and j (nb, permB);
// returns true when j is the P side (the Takizuka-Abe / Nanbu-Yonemura cyclic pairing puts
// the species with more parcels there)
__device__ __forceinline__ bool stage_inter(StageK& d, const CoulombArgs& a, int i, int j, int na, int nb, int stage,
                                            double Vc, bool ny, int permA, int permB) {
    d.stage = stage;
    const double wi = a.weight[i], wj = a.weight[j];
    bool jP = false;
    if (na == nb) {
        d.kind = kInterEqual; d.n = na; d.units = na;
        d.heavy = 0; d.hw = d.lw = 1;
        d.permP = permA; d.permQ = permB;
        d.mP = a.mass[i]; d.qP = a.charge[i]; d.wP = wi;
        d.mQ = a.mass[j]; d.qQ = a.charge[j]; d.wQ = wj;
        d.lnL = a.coulomb_log[i][j];
        d.nP = na * wi / Vc; d.nQ = nb * wj / Vc;
        d.nPQ = ny ? na * dmin(wi, wj) / Vc : d.nP;
    } else {
        const bool minIsA = na < nb;
        jP = minIsA;
        const int nMin = minIsA ? na : nb, nMax = minIsA ? nb : na;
        const int smin = minIsA ? i : j, smax = minIsA ? j : i;
        d.kind = kCyclic; d.n = nMax;
        d.permP = minIsA ? permB : permA;
        d.permQ = minIsA ? permA : permB;
        d.mP = a.mass[smax]; d.qP = a.charge[smax]; d.wP = a.weight[smax];
        d.mQ = a.mass[smin]; d.qQ = a.charge[smin]; d.wQ = a.weight[smin];
        d.lnL = a.coulomb_log[smax][smin];
        d.nP = nMax * a.weight[smax] / Vc;
        d.nQ = nMin * a.weight[smin] / Vc;
        if (ny) {
            // NanbuYonemuraPairing.C:253-300
            d.nPQ = nMax * dmin(wi, wj) / Vc;
            d.g1_units = nMin; d.g1_lmax = nMax; d.g1_stride = nMin;
            d.g1_xo = 0; d.g1_yo = 0; d.g1_ko = 0;
            d.units = nMin;
            d.g2_lmax = 0; d.g2_stride = 1; d.g2_xo = 0; d.g2_yo = 0; d.g2_ko = 0;
            d.heavy = nMax % nMin; d.lw = nMax / nMin; d.hw = d.lw + 1;
        } else {
            // TakizukaAbePairing.C:197-308: two groups
            d.nPQ = d.nP;
            const int delta = nMax / nMin;
            const double delta_r = (double)(nMax % nMin) / nMin;
            const int nMinG1 = (int)(delta_r * nMin);
            const int nMaxG1 = (int)((delta + 1) * delta_r * nMin);
            const int nMinG2 = (int)((1.0 - delta_r) * nMin);
            const int nMaxG2 = (int)(delta * (1.0 - delta_r) * nMin);
            const int cyc = nMinG2 > 0 ? nMinG2 : 1;
            d.g1_units = nMinG1 > 0 ? nMinG1 : 0;
            d.g1_lmax = nMaxG1; d.g1_stride = nMinG1 > 0 ? nMinG1 : 1;
            d.g1_xo = 0; d.g1_yo = 0; d.g1_ko = 0;
            d.g2_lmax = nMaxG2; d.g2_stride = cyc;
            d.g2_xo = nMaxG1; d.g2_yo = nMinG1; d.g2_ko = nMaxG1;
            d.units = d.g1_units + (nMaxG2 > 0 ? cyc : 0);
            d.heavy = d.g1_units; d.lw = delta; d.hw = delta + 1;
        }
    }
    const int sP = jP ? j : i, sQ = jP ? i : j;
    d.u[0] = a.ux[sP]; d.u[1] = a.uy[sP]; d.u[2] = a.uz[sP];
    d.u[3] = a.ux[sQ]; d.u[4] = a.uy[sQ]; d.u[5] = a.uz[sQ];
    finish_stage(d, a.dt, a.pair[sP * kMaxSpecies + sQ]);
    return jP;
}

// The unit run by a lane's slot s of a round of `total` units on L lanes (lane s mod L).
// Units are independent, so any order gives the same result; the units with extra
// collisions (the first `heavy` of a cyclic stage, the odd intra units 0) are moved onto
// lanes that have no unit in the round's last, partial pass.
__device__ __forceinline__ int balance_slot(int s, int total, int L, int nd, const StageK* sk, int units0) {
    const int m = total % L;
    if (m == 0) return s;
    const StageK& d0 = sk[0];
    if (d0.kind == kCyclic) {
        const int H = d0.heavy;
        const int span = m + H;
        return (H > 0 && span <= total && s < span) ? (s + H) % span : s;
    }
    int u = s;
    if (d0.heavy > 0 && L - 1 < total) u = u == 0 ? L - 1 : (u == L - 1 ? 0 : u);
    if (nd > 1 && sk[1].heavy > 0 && L - 2 < total) u = u == units0 ? L - 2 : (u == L - 2 ? units0 : u);
    return u;
}

// every pairing stage of one cell by one wave; maxp = LDS slots per index list.  With CU the
// velocities of the (at most two) species a stage needs are held in LDS for the whole cell:
// read once and written back once, coalesced, instead of gathered per collision.
template <class UT, int MODEL, bool WC, bool CU>
__device__ unsigned long long coulomb_cell(const CoulombArgs& a, int cell, int maxp, uint64_t* lds, StageK* sk) {
    const int kstride = maxp + 8;
    uint64_t* keyA = lds;
    uint64_t* keyB = lds + kstride;
    int* valA = (int*)(lds + 2 * kstride);
    int* valB = valA + maxp;
    UT* cache = (UT*)(lds + 2 * kstride + maxp);   // CU: slot q component c at cache[(3q + c) maxp]
    uint32_t* idc = (uint32_t*)(cache + 6 * maxp);  // CU: slot q parcel ids at idc[q maxp]
    int hold0 = -1, hold1 = -1;                     // species held by the two cache slots
    const int ns = a.n_species;
    const double Vc = a.V[cell];
    // CoulombCollisionModel::debyeLength (CoulombCollisionModel.C:249-310), parcels summed
    // in ascending id order
    double debye = 0.0;
    if (a.calc_debye) {
        double deb = 0.0, nmax = 0.0;
        for (int s = 0; s < ns; ++s) {
            if (a.charge[s] == 0.0 || !((a.debye_mask >> s) & 1u)) continue;
            const int f0 = a.start[s][cell], n = a.start[s][cell + 1] - f0;
            if (n <= 0) continue;
            int np2 = 1;
            while (np2 < n) np2 <<= 1;
            for (int k = threadIdx.x; k < np2; k += blockDim.x) {
                keyA[k] = (k < n) ? (((uint64_t)a.id[s][f0 + k] << 32) | (uint32_t)k) : ~0ull;
                valA[k] = (k < n) ? slot_index(a.perm[s], f0 + k) : -1;
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
            const double ch = a.charge[s];
            const double mvx = vx / nn, mvy = vy / nn, mvz = vz / nn;
            const double mvs = vs / nn;
            nn /= Vc;
            if (nn > nmax) nmax = nn;
            const double vd2 = a.debye_drift ? (mvx * mvx + mvy * mvy + mvz * mvz) : 0.0;
            const double T = a.mass[s] / (3.0 * kKB) * (mvs - vd2);
            if (T > 0.0) deb += nn * ch * ch / kEps0 * 1.0 / (T * kKB);
            __syncthreads();
        }
        if (deb > 0.0) {
            const double rmin = pow(4.0 * kPi * nmax / 3.0, -1.0 / 3.0);
            const double dl = sqrt(1.0 / deb);
            debye = dl > rmin ? dl : rmin;
        }
    }
    // stages in the reference order: species i ascending, partner j descending from the
    // last charged species down to i.  An intra-species stage (i, i) is followed by the
    // intra stage of the last charged species when that is the next charged species; the
    // two touch disjoint parcels and run together.
    int last = -1;
    for (int s = 0; s < ns; ++s)
        if (a.charge[s] != 0.0) last = s;
    const bool ny = a.pairing == 1;
    // cache slots (CU): write slot q back to the arrays of species sp
    auto put_slot = [&](int q, int sp) {
        const int f0 = a.start[sp][cell], n = a.start[sp][cell + 1] - f0;
        const UT* cx = cache + (size_t)(3 * q) * maxp;
        UT* gx = (UT*)a.ux[sp];
        UT* gy = (UT*)a.uy[sp];
        UT* gz = (UT*)a.uz[sp];
        for (int k = threadIdx.x; k < n; k += blockDim.x) {
            const int j = slot_index(a.perm[sp], f0 + k);
            gx[j] = cx[k]; gy[j] = cx[maxp + k]; gz[j] = cx[2 * maxp + k];
        }
    };
    // fill slot qa with species sa and/or slot qb with species sb (a negative slot is left
    // alone); the loads of both are in flight together
    auto get_slots = [&](int qa, int sa, int qb, int sb) {
        const int fa = qa >= 0 ? a.start[sa][cell] : 0, na = qa >= 0 ? a.start[sa][cell + 1] - fa : 0;
        const int fb = qb >= 0 ? a.start[sb][cell] : 0, nb = qb >= 0 ? a.start[sb][cell + 1] - fb : 0;
        const int n = na > nb ? na : nb;
        for (int k = threadIdx.x; k < n; k += blockDim.x) {
            UT ax = 0, ay = 0, az = 0, bx = 0, by = 0, bz = 0;
            uint32_t ai = 0, bi = 0;
            if (k < na) {
                const int j = slot_index(a.perm[sa], fa + k);
                ax = ((const UT*)a.ux[sa])[j]; ay = ((const UT*)a.uy[sa])[j];
                az = ((const UT*)a.uz[sa])[j]; ai = a.id[sa][fa + k];
            }
            if (k < nb) {
                const int j = slot_index(a.perm[sb], fb + k);
                bx = ((const UT*)a.ux[sb])[j]; by = ((const UT*)a.uy[sb])[j];
                bz = ((const UT*)a.uz[sb])[j]; bi = a.id[sb][fb + k];
            }
            if (k < na) {
                UT* cx = cache + (size_t)(3 * qa) * maxp;
                cx[k] = ax; cx[maxp + k] = ay; cx[2 * maxp + k] = az; idc[qa * maxp + k] = ai;
            }
            if (k < nb) {
                UT* cx = cache + (size_t)(3 * qb) * maxp;
                cx[k] = bx; cx[maxp + k] = by; cx[2 * maxp + k] = bz; idc[qb * maxp + k] = bi;
            }
        }
    };
    // bring species sa (and sb >= 0) into the two slots; returns their slots
    auto hold = [&](int sa, int sb, int& qa, int& qb) {
        qa = hold0 == sa ? 0 : (hold1 == sa ? 1 : -1);
        qb = sb < 0 ? -1 : (hold0 == sb ? 0 : (hold1 == sb ? 1 : -1));
        int fa = -1, fb = -1;
        if (qa < 0) fa = qa = qb == 0 ? 1 : 0;
        if (sb >= 0 && qb < 0) fb = qb = 1 - qa;
        if (fa < 0 && fb < 0) return;
        if (fa >= 0 && (fa == 0 ? hold0 : hold1) >= 0) put_slot(fa, fa == 0 ? hold0 : hold1);
        if (fb >= 0 && (fb == 0 ? hold0 : hold1) >= 0) put_slot(fb, fb == 0 ? hold0 : hold1);
        get_slots(fa, sa, fb, sb);
        if (fa == 0 || fb == 0) hold0 = fa == 0 ? sa : sb;
        if (fa == 1 || fb == 1) hold1 = fa == 1 ? sa : sb;
        __syncthreads();
    };
    auto slot_ptrs = [&](StageK& d, int side, int q, int sp) {
        if (CU) {
            if (side == 0) d.cP = 3 * q * maxp; else d.cQ = 3 * q * maxp;
        }
    };
    unsigned long long ncoll = 0;
    int st = 0;
    bool skip_last_intra = false;
    for (int i = 0; i < ns; ++i) {
        if (a.charge[i] == 0.0) continue;
        for (int j = last; j >= i; --j) {
            if (a.charge[j] == 0.0) continue;
            const int stage_now = st++;
            if (i == j && i == last && skip_last_intra) continue;
            int nd = 0;        // stage tables filled
            if (i == j) {
                const bool pair2 = next_charged(a, i) == last;
                if (pair2) skip_last_intra = true;
                int q0 = 0, q1 = 1;
                if (CU && a.allow_intra) hold(i, pair2 ? last : -1, q0, q1);
                for (int v = 0; v < (pair2 ? 2 : 1); ++v) {
                    const int spc = v == 0 ? i : last;
                    const int f0 = a.start[spc][cell];
                    const int cnt = a.start[spc][cell + 1] - f0;
                    if (cnt <= 1 || !a.allow_intra) continue;
                    uint64_t* key = nd == 0 ? keyA : keyB;
                    int* val = nd == 0 ? valA : valB;
                    if (CU)
                        keyed_perm(a.seed, a.step, stage_now + v, spc, idc + (v == 0 ? q0 : q1) * maxp, 0, cnt, key,
                                   val, 0, nullptr);
                    else
                        keyed_perm(a.seed, a.step, stage_now + v, spc, a.id[spc], f0, cnt, key, val, f0, a.perm[spc]);
                    if (threadIdx.x == 0) {
                        StageK& d = sk[nd];
                        stage_intra(d, a, spc, cnt, stage_now + v, Vc, ny, nd == 0 ? 0 : maxp);
                        slot_ptrs(d, 0, v == 0 ? q0 : q1, spc);
                        slot_ptrs(d, 1, v == 0 ? q0 : q1, spc);
                    }
                    ++nd;
                }
            } else {
                const int a0 = a.start[i][cell], na = a.start[i][cell + 1] - a0;
                const int b0 = a.start[j][cell], nb = a.start[j][cell + 1] - b0;
                if (na > 0 && nb > 0) {
                    int qi = 0, qj = 1;
                    if (CU) hold(i, j, qi, qj);
                    if (CU) {
                        keyed_perm(a.seed, a.step, stage_now, i, idc + qi * maxp, 0, na, keyA, valA, 0, nullptr);
                        keyed_perm(a.seed, a.step, stage_now, j, idc + qj * maxp, 0, nb, keyB, valB, 0, nullptr);
                    } else {
                        keyed_perm(a.seed, a.step, stage_now, i, a.id[i], a0, na, keyA, valA, a0, a.perm[i]);
                        keyed_perm(a.seed, a.step, stage_now, j, a.id[j], b0, nb, keyB, valB, b0, a.perm[j]);
                    }
                    if (threadIdx.x == 0) {
                        StageK& d = sk[0];
                        const bool jP = stage_inter(d, a, i, j, na, nb, stage_now, Vc, ny, 0, maxp);
                        slot_ptrs(d, 0, jP ? qj : qi, jP ? j : i);
                        slot_ptrs(d, 1, jP ? qi : qj, jP ? i : j);
                    }
                    nd = 1;
                }
            }
            if (nd == 0) continue;
            __syncthreads();
            // one collision per loop iteration, as in coulomb_group
            const int units0 = sk[0].units;
            const int total = units0 + (nd > 1 ? sk[1].units : 0);
            const int L = blockDim.x;
            int slot = threadIdx.x, t = 0, u = 0, cnt = 0;
            if (slot < total) {
                u = balance_slot(slot, total, L, nd, sk, units0);
                cnt = unit_count(sk[u < units0 ? 0 : 1], u < units0 ? u : u - units0);
            }
            while (slot < total) {
                const StageK& d = sk[u < units0 ? 0 : 1];
                const int uu = u < units0 ? u : u - units0;
                if (t < cnt) {
                    int iP, iQ, k;
                    decode(d, valA, uu, t, iP, iQ, k);
                    collide<UT, MODEL, WC, CU>(a, d, cell, iP, iQ, k, debye, cache, maxp);
                    ++ncoll;
                    if (++t < cnt) continue;
                }
                slot += L;
                t = 0;
                if (slot < total) {
                    u = balance_slot(slot, total, L, nd, sk, units0);
                    cnt = unit_count(sk[u < units0 ? 0 : 1], u < units0 ? u : u - units0);
                }
            }
            __syncthreads();
        }
    }
    if (CU) {
        if (hold0 >= 0) put_slot(0, hold0);
        if (hold1 >= 0) put_slot(1, hold1);
    }
    return ncoll;
}

// the wave's collision count into the step's total, spread over kNcollSlots counters a cache
// line apart (one address taking an atomic from every block serialises the blocks)
__device__ __forceinline__ void add_ncoll(const CoulombArgs& a, unsigned long long ncoll) {
    for (int off = 32; off > 0; off >>= 1) ncoll += __shfl_xor(ncoll, off);
    if (threadIdx.x == 0 && ncoll) atomicAdd(a.ncoll + kNcollStride * (blockIdx.x % kNcollSlots), ncoll);
}

// one wave per cell with small index lists in LDS; cells with more parcels of a species
// than a.maxp go to the list of the large-cell pass
template <class UT, int MODEL, bool WC>
__device__ __forceinline__ void coulomb_pass(const CoulombArgs& a) {
    extern __shared__ uint64_t lds[];
    __shared__ StageK sk[2];
    const int cell = blockIdx.x;
    if (cell_max_count(a, cell) > a.maxp) {
        if (threadIdx.x == 0) a.big_list[atomicAdd(a.big_n, 1)] = cell;
        return;
    }
    add_ncoll(a, coulomb_cell<UT, MODEL, WC, true>(a, cell, a.maxp, lds, sk));
}

// Nanbu scattering fits three waves per SIMD without spilling; Perez keeps the default
template <class UT, bool WC>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3, 8))) k_coulomb_nanbu(CoulombArgs a) {
    coulomb_pass<UT, 0, WC>(a);
}

template <class UT, bool WC>
__global__ void __launch_bounds__(64) k_coulomb_perez(CoulombArgs a) {
    coulomb_pass<UT, 1, WC>(a);
}

// the listed large cells, up to a.maxp_big parcels of a species; larger cells raise the
// overflow flag and are left alone
template <class UT, int MODEL, bool WC>
__global__ void __launch_bounds__(64) k_coulomb_big(CoulombArgs a) {
    extern __shared__ uint64_t lds[];
    __shared__ StageK sk[2];
    const int nbig = *a.big_n;
    unsigned long long ncoll = 0;
    for (int i = blockIdx.x; i < nbig; i += gridDim.x) {
        const int cell = a.big_list[i];
        if (cell_max_count(a, cell) > a.maxp_big) {
            if (threadIdx.x == 0) atomicOr(a.flags, kFlagCellOverflow);
            continue;
        }
        ncoll += coulomb_cell<UT, MODEL, WC, false>(a, cell, a.maxp_big, lds, sk);
        __syncthreads();
    }
    add_ncoll(a, ncoll);
}

// ---------------------------------------------------------------- several cells per wave
// 64 / L cells per wave, L lanes per cell (L = 16, 32 or 64).  A stage's collision units are spread over the
// cell's L lanes, so the short sequential runs that end a stage (a Takizuka-Abe odd triple,
// parcels of the smaller species meeting two partners) share the wave's passes with the
// other cells' collisions instead of holding all 64 lanes, and the stage tables of the
// cells are filled together.  Per cell in LDS: one key list (maxp + 8), 4 L bins and two
// index lists (maxp = 8 L each).  A Coulomb logarithm from the Debye length takes a pre-pass
// over the cell's charged species (ranked by id, summed by a fixed tree).

// The cell's group of L lanes (lane r) ranks its n <= 8 L parcels of one species by their
// random key, the order keyed_perm produces.  The keys are bucketed by their top bits into
// 4 L bins (an LDS histogram, its prefix sum over the group's lanes and a placement in
// bucket order); a key's rank is its bucket's start plus the smaller keys of its own bucket,
// about two comparisons instead of n.  A group with n = 0 only takes part in the barriers.
template <int L>
__device__ void group_perm(uint64_t seed, int64_t step, int stage, int species, const uint32_t* id, const int* perm,
                           int first, int n, uint64_t* key, int* bins, int* val, int r) {
    constexpr int KM = 8, NB = 4 * L, PER = NB / L;
    constexpr int SH = 64 - (L == 16 ? 6 : (L == 32 ? 7 : 8));   // top log2(NB) bits pick the bin
    for (int b = r; b < NB; b += L) bins[b] = 0;
    __syncthreads();
    uint64_t kk[KM];
#pragma unroll
    for (int q = 0; q < KM; ++q) {
        const int k = r + q * L;
        kk[q] = 0;
        if (k < n) {
            kk[q] = (shuffle_key(seed, step, stage, species, id[first + k]) << kKeyIndexBits) | (uint64_t)k;
            val[k] = atomicAdd(&bins[kk[q] >> SH], 1);   // place in the bucket, until the ranks
        }
    }
    __syncthreads();
    // exclusive prefix of the bins: PER consecutive bins per lane, scanned over the group
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
#pragma unroll
    for (int q = 0; q < KM; ++q)
        if (r + q * L < n) key[bins[kk[q] >> SH] + val[r + q * L]] = kk[q];
    __syncthreads();
#pragma unroll
    for (int q = 0; q < KM; ++q) {
        const int k = r + q * L;
        if (k < n) {
            const int b = (int)(kk[q] >> SH);
            const int s0 = bins[b], s1 = b + 1 < NB ? bins[b + 1] : n;
            int c = s0;
            for (int j = s0; j < s1; ++j) c += (int)(key[j] < kk[q]);
            val[c] = slot_index(perm, first + k);
        }
    }
    __syncthreads();
}


template <class UT, int MODEL, bool WC, int L, bool DEB>
__device__ __forceinline__ void coulomb_group(const CoulombArgs& a) {
    constexpr int W = 64 / L;
    extern __shared__ uint64_t lds[];
    __shared__ StageK skw[W][2];
    const int w = threadIdx.x / L, r = threadIdx.x % L;
    const int maxp = a.maxp;                 // 8 L
    uint64_t* key = lds + (size_t)w * (2 * maxp + 8 + 2 * L);
    int* bins = (int*)(key + maxp + 8);      // 4 L
    int* valA = bins + 4 * L;
    int* valB = valA + maxp;
    StageK* sk = skw[w];
    const int cell = blockIdx.x * W + w;
    bool on = cell < a.n_cells;
    if (on && cell_max_count(a, cell) > maxp) {
        if (r == 0) a.big_list[atomicAdd(a.big_n, 1)] = cell;
        on = false;
    }
    const double Vc = on ? a.V[cell] : 1.0;
    const int ns = a.n_species;
    // CoulombCollisionModel::debyeLength (CoulombCollisionModel.C:249-310) as coulomb_cell
    // computes it: each charged species' parcels ranked by id (distinct keys (id, cell-local
    // index), the rank is the count of smaller keys), lane r summing ranks r, r + L, ...,
    // then a fixed shuffle tree over the group, so the sums do not depend on the order the
    // sort left the parcels in
    double debye = 0.0;
    if constexpr (DEB) {
        double deb = 0.0, nmax = 0.0;
        for (int s = 0; s < ns; ++s) {
            if (a.charge[s] == 0.0 || !((a.debye_mask >> s) & 1u)) continue;
            const int f0 = on ? a.start[s][cell] : 0;
            const int n = on ? a.start[s][cell + 1] - f0 : 0;
            const int n8 = (n + 7) & ~7;
            for (int k = r; k < n8; k += L)
                key[k] = k < n ? (((uint64_t)a.id[s][f0 + k] << 32) | (uint32_t)k) : ~0ull;
            __syncthreads();
            for (int k = r; k < n; k += L) {
                const uint64_t kk = key[k];
                int rk = 0;
                for (int j = 0; j < n8; j += 8) {
                    const uint64_t* q = key + j;
                    rk += (int)(q[0] < kk) + (int)(q[1] < kk) + (int)(q[2] < kk) + (int)(q[3] < kk) +
                          (int)(q[4] < kk) + (int)(q[5] < kk) + (int)(q[6] < kk) + (int)(q[7] < kk);
                }
                valA[rk] = slot_index(a.perm[s], f0 + k);
            }
            __syncthreads();
            const UT* ux = (const UT*)a.ux[s];
            const UT* uy = (const UT*)a.uy[s];
            const UT* uz = (const UT*)a.uz[s];
            const double wgt = a.weight[s];
            double vs = 0.0, vx = 0.0, vy = 0.0, vz = 0.0;
            for (int k = r; k < n; k += L) {
                const int i = valA[k];
                const double x = ux[i], y = uy[i], z = uz[i];
                vs += (x * x + y * y + z * z) * wgt;
                vx += wgt * x; vy += wgt * y; vz += wgt * z;
            }
            for (int off = L / 2; off > 0; off >>= 1) {
                vs += __shfl_xor(vs, off); vx += __shfl_xor(vx, off);
                vy += __shfl_xor(vy, off); vz += __shfl_xor(vz, off);
            }
            __syncthreads();   // valA and key are reused by the next species
            if (n <= 0) continue;
            double nn = n * wgt;
            const double mvx = vx / nn, mvy = vy / nn, mvz = vz / nn, mvs = vs / nn;
            nn /= Vc;
            if (nn > nmax) nmax = nn;
            const double vd2 = a.debye_drift ? (mvx * mvx + mvy * mvy + mvz * mvz) : 0.0;
            const double T = a.mass[s] / (3.0 * kKB) * (mvs - vd2);
            if (T > 0.0) deb += nn * a.charge[s] * a.charge[s] / kEps0 * 1.0 / (T * kKB);
        }
        if (deb > 0.0) {
            const double rmin = pow(4.0 * kPi * nmax / 3.0, -1.0 / 3.0);
            const double dl = sqrt(1.0 / deb);
            debye = dl > rmin ? dl : rmin;
        }
    }
    int last = -1;
    for (int s = 0; s < ns; ++s)
        if (a.charge[s] != 0.0) last = s;
    const bool ny = a.pairing == 1;
    unsigned long long ncoll = 0;
    int st = 0;
    bool skip_last_intra = false;
    // the stage order of coulomb_cell
    for (int i = 0; i < ns; ++i) {
        if (a.charge[i] == 0.0) continue;
        for (int j = last; j >= i; --j) {
            if (a.charge[j] == 0.0) continue;
            const int stage_now = st++;
            if (i == j && i == last && skip_last_intra) continue;
            int nd = 0;
            if (i == j) {
                const bool pair2 = next_charged(a, i) == last;
                if (pair2) skip_last_intra = true;
                for (int v = 0; v < (pair2 ? 2 : 1); ++v) {
                    const int spc = v == 0 ? i : last;
                    const int f0 = on ? a.start[spc][cell] : 0;
                    const int cnt = on ? a.start[spc][cell + 1] - f0 : 0;
                    const bool go = cnt > 1 && a.allow_intra;
                    group_perm<L>(a.seed, a.step, stage_now + v, spc, a.id[spc], a.perm[spc], f0, go ? cnt : 0, key,
                                  bins, nd == 0 ? valA : valB, r);
                    if (go && r == 0) stage_intra(sk[nd], a, spc, cnt, stage_now + v, Vc, ny, nd == 0 ? 0 : maxp);
                    if (go) ++nd;
                }
            } else {
                const int a0 = on ? a.start[i][cell] : 0, na = on ? a.start[i][cell + 1] - a0 : 0;
                const int b0 = on ? a.start[j][cell] : 0, nb = on ? a.start[j][cell + 1] - b0 : 0;
                const bool go = na > 0 && nb > 0;
                group_perm<L>(a.seed, a.step, stage_now, i, a.id[i], a.perm[i], a0, go ? na : 0, key, bins, valA, r);
                group_perm<L>(a.seed, a.step, stage_now, j, a.id[j], a.perm[j], b0, go ? nb : 0, key, bins, valB, r);
                if (go && r == 0) stage_inter(sk[0], a, i, j, na, nb, stage_now, Vc, ny, 0, maxp);
                nd = go ? 1 : 0;
            }
            __syncthreads();
            // one collision per iteration: a lane moves on to its next unit (slots r, r + L,
            // ...) inside the same loop, so the wave runs as many iterations as its busiest
            // lane has collisions
            const int units0 = nd > 0 ? sk[0].units : 0;
            const int total = units0 + (nd > 1 ? sk[1].units : 0);
            int slot = r, t = 0, u = 0, cnt = 0;
            if (slot < total) {
                u = balance_slot(slot, total, L, nd, sk, units0);
                cnt = unit_count(sk[u < units0 ? 0 : 1], u < units0 ? u : u - units0);
            }
            while (slot < total) {
                const StageK& d = sk[u < units0 ? 0 : 1];
                const int uu = u < units0 ? u : u - units0;
                if (t < cnt) {
                    int iP, iQ, k;
                    decode(d, valA, uu, t, iP, iQ, k);
                    collide<UT, MODEL, WC, false>(a, d, cell, iP, iQ, k, debye, (UT*)nullptr, 0);
                    ++ncoll;
                    if (++t < cnt) continue;
                }
                slot += L;
                t = 0;
                if (slot < total) {
                    u = balance_slot(slot, total, L, nd, sk, units0);
                    cnt = unit_count(sk[u < units0 ? 0 : 1], u < units0 ? u : u - units0);
                }
            }
            __syncthreads();
        }
    }
    add_ncoll(a, ncoll);
}

template <class UT, bool WC, int L, bool DEB>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3, 8))) k_coulomb_group_nanbu(CoulombArgs a) {
    coulomb_group<UT, 0, WC, L, DEB>(a);
}

template <class UT, bool WC, int L, bool DEB>
__global__ void __launch_bounds__(64) k_coulomb_group_perez(CoulombArgs a) {
    coulomb_group<UT, 1, WC, L, DEB>(a);
}

void pair_constants(const double* mass, const double* charge, int n_species, PairConst* out) {
    for (int p = 0; p < n_species; ++p)
        for (int q = 0; q < n_species; ++q) {
            const double mP = mass[p], mQ = mass[q], qP = charge[p], qQ = charge[q];
            const double mu = mP * mQ / (mP + mQ);
            const double ad = qP * qQ / (kEps0 * mu);
            PairConst& c = out[p * kMaxSpecies + q];
            c.cst0 = 1.0 / (4.0 * kPi) * (ad * ad);
            c.fP = mQ / (mP + mQ);
            c.fQ = mP / (mP + mQ);
            c.b0c = qP * qQ / (2.0 * kPi * kEps0) / mu;
            c.hbc = kH / (2.0 * mu);
        }
}

// key lists, index lists and (cached passes, cache_elem > 0) two slots of velocities and ids
size_t coulomb_lds_bytes(int maxp, size_t cache_elem) {
    return (size_t)(2 * (maxp + 8)) * sizeof(uint64_t) + (size_t)(2 * maxp) * sizeof(int) +
           (cache_elem ? (size_t)(6 * maxp) * cache_elem + (size_t)(2 * maxp) * sizeof(uint32_t) : 0);
}

template <class UT, int MODEL, bool WC, int L>
static void launch_coulomb_group(const CoulombArgs& a, hipStream_t s) {
    constexpr int W = 64 / L;
    const size_t lds = (size_t)W * (2 * a.maxp + 8 + 2 * L) * sizeof(uint64_t);
    const int grid = (a.n_cells + W - 1) / W;
    // the Debye pre-pass is its own instantiation, so fixed logarithms keep their registers
    if (MODEL == 1) {
        if (a.calc_debye) hipLaunchKernelGGL((k_coulomb_group_perez<UT, WC, L, true>), dim3(grid), dim3(64), lds, s, a);
        else hipLaunchKernelGGL((k_coulomb_group_perez<UT, WC, L, false>), dim3(grid), dim3(64), lds, s, a);
    } else {
        if (a.calc_debye) hipLaunchKernelGGL((k_coulomb_group_nanbu<UT, WC, L, true>), dim3(grid), dim3(64), lds, s, a);
        else hipLaunchKernelGGL((k_coulomb_group_nanbu<UT, WC, L, false>), dim3(grid), dim3(64), lds, s, a);
    }
}

template <class UT, int MODEL, bool WC>
static void launch_coulomb_t(const CoulombArgs& a, hipStream_t s) {
    const size_t lds = coulomb_lds_bytes(a.maxp, sizeof(UT));
    if (a.group == 16)
        launch_coulomb_group<UT, MODEL, WC, 16>(a, s);
    else if (a.group == 32)
        launch_coulomb_group<UT, MODEL, WC, 32>(a, s);
    else if (a.group == 64)
        launch_coulomb_group<UT, MODEL, WC, 64>(a, s);
    else if (MODEL == 1)
        hipLaunchKernelGGL((k_coulomb_perez<UT, WC>), dim3(a.n_cells), dim3(64), lds, s, a);
    else
        hipLaunchKernelGGL((k_coulomb_nanbu<UT, WC>), dim3(a.n_cells), dim3(64), lds, s, a);
    const int nb = a.n_cells < 2048 ? a.n_cells : 2048;
    hipLaunchKernelGGL((k_coulomb_big<UT, MODEL, WC>), dim3(nb), dim3(64), coulomb_lds_bytes(a.maxp_big, 0), s, a);
}

template <class UT, int MODEL>
static void launch_coulomb_m(const CoulombArgs& a, hipStream_t s) {
    if (a.weight_correction) launch_coulomb_t<UT, MODEL, true>(a, s);
    else launch_coulomb_t<UT, MODEL, false>(a, s);
}

hipError_t launch_coulomb(const CoulombArgs& a, bool fp32_velocity, hipStream_t s) {
    if (a.n_cells <= 0) return hipSuccess;
    if (a.model == 1) {
        if (fp32_velocity) launch_coulomb_m<float, 1>(a, s);
        else launch_coulomb_m<double, 1>(a, s);
    } else {
        if (fp32_velocity) launch_coulomb_m<float, 0>(a, s);
        else launch_coulomb_m<double, 0>(a, s);
    }
    return hipGetLastError();
}

}  // namespace picmi
