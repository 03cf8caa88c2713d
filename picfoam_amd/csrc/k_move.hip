// Particle kernels of libpicmi: push + E interpolation + barycentric tet tracking +
// charge deposit / cell count (one lane per parcel, parcels kept sorted by cell), push
// only, point location, positions, Maxwellian initialisation and the counting-sort
// scatter.
#include "device.h"
#include "kernels.h"
#include "collide.h"

namespace picmi {

// ---------------------------------------------------------------- pushers
// BorisPusher.C:69-112, BorisNRPusher.C:69-99, VayPusher.C:54-100, HigueraCary.C:69-115.
// k = q*(dt/2)/m and kv = q*dt/m are precomputed on the host in FP64.
// Divisions by a scalar are multiplications by its reciprocal (1/sqrt as rsqrt): a few ulp
// from the reference's operation order, a fraction of the instructions.
template <class T, int P>
__device__ __forceinline__ vec3<T> push_u(vec3<T> U, vec3<T> E, vec3<T> B, T k, T kv, bool rotate, bool hc_ref) {
    const T ic = (T)(1.0 / kC);
    if constexpr (P == kPushBoris) {
        T mu = mag(U) * ic;
        vec3<T> rU = rsq((T)1 - mu * mu) * U;
        vec3<T> oldrU = rU;
        vec3<T> rUsm = k * E;
        rU = rU + rUsm;
        if (rotate) {
            T a = mag(oldrU) * ic, b = mag(rU) * ic;
            T lF = (sqrt((T)1 + a * a) + sqrt((T)1 + b * b)) * (T)0.5;
            vec3<T> t = (k / lF) * B;
            vec3<T> Ud = rU + cross(rU, t);
            rU = rU + cross(Ud, ((T)2 / ((T)1 + dot(t, t))) * t);
        }
        rU = rU + rUsm;
        T g = mag(rU) * ic;
        return rsq((T)1 + g * g) * rU;
    } else if constexpr (P == kPushBorisNR) {
        U = U + k * E;
        if (rotate) {
            vec3<T> t = k * B;
            vec3<T> Ud = U + cross(U, t);
            U = U + cross(Ud, ((T)2 / ((T)1 + dot(t, t))) * t);
        }
        return U + k * E;
    } else if constexpr (P == kPushVay) {
        T mu = mag(U) * ic;
        T lF = rsq((T)1 - mu * mu);
        vec3<T> rU = lF * U;
        vec3<T> Ud = rU + kv * E;
        vec3<T> Tv = k * B;
        Ud = Ud + cross(U, Tv);
        T ud = mag(Ud) * ic;
        T alpha = (T)1 + ud * ud;
        T T2 = dot(Tv, Tv);
        T s = alpha - T2;
        T w = dot(Ud, Tv) * ic;
        T us2 = w * w;
        alpha = rsq((T)0.5 * (s + sqrt(s * s + (T)4 * (T2 + us2))));
        Tv = alpha * Tv;
        s = (T)1 / ((T)1 + dot(Tv, Tv));
        alpha = dot(Ud, Tv);
        rU = s * ((Ud + alpha * Tv) + cross(Ud, Tv));
        T g = mag(rU) * ic;
        return rsq((T)1 + g * g) * rU;
    } else if constexpr (P == kPushHC) {
        T mu = mag(U) * ic;
        vec3<T> rU = rsq((T)1 - mu * mu) * U;
        vec3<T> rUsm = k * E;
        rU = rU + rUsm;
        if (rotate) {
            T g = mag(rU) * ic;
            T gfm2 = (T)1 + g * g;
            vec3<T> t = k * B;
            T beta2 = dot(t, t);
            // published Higuera-Cary: (t.u/c)^2; the reference writes (t.u)^2
            T tu = hc_ref ? dot(t, rU) : dot(t, rU) * ic;
            T sg = gfm2 - beta2;
            T iG = rsq((T)0.5 * (sg + sqrt(sg * sg + (T)4 * (beta2 + tu * tu))));
            t = iG * t;
            vec3<T> Ud = rU + cross(rU, t);
            rU = rU + cross(Ud, ((T)2 / ((T)1 + dot(t, t))) * t);
        }
        rU = rU + rUsm;
        T g = mag(rU) * ic;
        return rsq((T)1 + g * g) * rU;
    } else {
        return U;
    }
}

// CellAverageFW.C:58-65 / VolumeWeightingFW.C:60-67 (interpolationCellPoint)
template <class T, int W>
__device__ __forceinline__ vec3<T> interp_E(const TetGeo& g, double y1, double y2, double y3, const T* Ec,
                                            const T* Ep) {
    const int c = g.cell;
    vec3<T> ec = mk3<T>(Ec[3 * c], Ec[3 * c + 1], Ec[3 * c + 2]);
    if constexpr (W == kCellAverage) {
        return ec;
    } else {
        T b0 = (T)(1.0 - (y1 + y2 + y3));
        vec3<T> e = b0 * ec;
        const int a = g.vm[0], b = g.vm[1], d = g.vm[2];
        e = e + (T)y1 * mk3<T>(Ep[3 * a], Ep[3 * a + 1], Ep[3 * a + 2]);
        e = e + (T)y2 * mk3<T>(Ep[3 * b], Ep[3 * b + 1], Ep[3 * b + 2]);
        e = e + (T)y3 * mk3<T>(Ep[3 * d], Ep[3 * d + 1], Ep[3 * d + 2]);
        return e;
    }
}

// The four barycentric slots of a lane pass through LDS when the parcel changes tet: the
// coordinates at the hit are written to the lane's column (slot k at ys[k * kMoveBlock]), the
// hit slot is overwritten with an exact zero, and the neighbour's slots are read back in
// its order.  Selecting them with v_cndmask instead costs about 40 VALU instructions per
// crossing; the LDS round trip costs nine LDS instructions and their latency.  The column
// layout keeps the lanes of a wave on distinct banks.
constexpr int kMoveBlock = 256;

// Barycentric tracking through the tet decomposition (PICParcel::moveForward,
// PICParcel.C:81-144, with OpenFOAM-9 particle::trackToFace re-derived): per tet the
// remaining displacement (1-sf)*dt*U becomes Tx_i = r_i.d (det-scaled barycentric);
// the first face with Tx_i < 0 and 0 <= -y_i/Tx_i < 1/det is hit and the parcel
// continues in the neighbouring tet with its coordinates permuted by the adjacency table.
// Returns the final tet (-1 = left through a non-coupled patch, deleted as NoInflow does);
// *status = 1 when the iteration guard stopped the track, 2 when the table had no
// neighbour, 3 when the parcel reached a wall face (MoveArgs::patch_kind 1): it then stands
// on the face in its tet, sf holds the step fraction done, and the wall pass (k_wall) goes on
// from there.  The track starts at step fraction sf (0 for a new step); on_wall starts it on
// the tet's face 0 with Y0 exactly 0, as the parcel stood when its wall was hit (recomputing
// 1 - (Y1 + Y2 + Y3) there can give -1e-17, and a parcel still heading out would then miss
// the wall).  U is tracked without its empty components (reduced-dimension cases,
// PICParcel.C:118-123).
template <class UT>
__device__ __forceinline__ int track(const MoveArgs<UT>& a, int t, double& y1, double& y2, double& y3,
                                     const vec3<UT>& U, double& sf, int& status, double* __restrict__ ys,
                                     bool on_wall = false) {
    const TetGeo* __restrict__ tets = a.tets;
    const double dt = a.dt;
    const uint32_t m = a.track_mask;
    const double Ux = (m & 1u) ? 0.0 : (double)U.x, Uy = (m & 2u) ? 0.0 : (double)U.y,
                 Uz = (m & 4u) ? 0.0 : (double)U.z;
    double Y0 = on_wall ? 0.0 : 1.0 - (y1 + y2 + y3), Y1 = y1, Y2 = y2, Y3 = y3;
    int guard = 0;
    status = 0;
    while (sf < 1.0) {
        const double f = 1.0 - sf;
        const double fdt = f * dt;
        const double dx = fdt * Ux, dy = fdt * Uy, dz = fdt * Uz;
        const TetGeo& g = tets[t];
        // the whole record at once: neighbours and permutation travel with the geometry
        // instead of costing two more dependent loads after the face is known
        const int4 nb4 = *reinterpret_cast<const int4*>(g.nbr);
        const uint32_t perm = g.perm;
        const double T1 = g.r[0] * dx + g.r[1] * dy + g.r[2] * dz;
        const double T2 = g.r[3] * dx + g.r[4] * dy + g.r[5] * dz;
        const double T3 = g.r[6] * dx + g.r[7] * dy + g.r[8] * dz;
        const double T0 = -(T1 + T2 + T3);
        const double det = g.det;
        const bool dok = det > 0.0 && __builtin_isnormal(det);
        // first face hit: the smallest mu_i = -Y_i / T_i >= 0 over T_i < 0 below the tet's
        // end 1/det, found by comparing the fractions crosswise (as selects, in face order);
        // one division for the winner
        double bn = 1.0, bd = dok ? det : 1e-300;
        int iH = -1;
        bool h = T0 < 0.0 && Y0 >= 0.0 && Y0 * bd < -T0 * bn;
        bn = h ? Y0 : bn; bd = h ? -T0 : bd; iH = h ? 0 : iH;
        h = T1 < 0.0 && Y1 >= 0.0 && Y1 * bd < -T1 * bn;
        bn = h ? Y1 : bn; bd = h ? -T1 : bd; iH = h ? 1 : iH;
        h = T2 < 0.0 && Y2 >= 0.0 && Y2 * bd < -T2 * bn;
        bn = h ? Y2 : bn; bd = h ? -T2 : bd; iH = h ? 2 : iH;
        h = T3 < 0.0 && Y3 >= 0.0 && Y3 * bd < -T3 * bn;
        bn = h ? Y3 : bn; bd = h ? -T3 : bd; iH = h ? 3 : iH;
        const double muH = (iH >= 0 || dok) ? bn / bd : 1e300;
        const double Z0 = Y0 + muH * T0, Z1 = Y1 + muH * T1, Z2 = Y2 + muH * T2, Z3 = Y3 + muH * T3;
        if (iH < 0) {
            Y0 = Z0; Y1 = Z1; Y2 = Z2; Y3 = Z3;
            break;
        }
        double frac = muH * det;
        frac = frac < 0.0 ? 0.0 : (frac > 1.0 ? 1.0 : frac);
        sf = sf + f * frac;
        const int nb = (iH & 2) ? ((iH & 1) ? nb4.w : nb4.z) : ((iH & 1) ? nb4.y : nb4.x);
        if (nb < 0) {
            // the parcel stops on the face (an empty face, reached only by rounding; a wall,
            // where it waits for the wall pass) or leaves: the hit slot is exactly zero
            Y0 = iH == 0 ? 0.0 : Z0; Y1 = iH == 1 ? 0.0 : Z1;
            Y2 = iH == 2 ? 0.0 : Z2; Y3 = iH == 3 ? 0.0 : Z3;
            if (nb == kTetEmptyFace) break;
            if (nb > -1000000 && a.patch_kind && a.patch_kind[-1 - nb] != 0) {
                status = 3;
                break;
            }
            if (nb <= -1000000) status = 2;
            t = -1;
            break;
        }
        // slot j of the neighbour takes the coordinate of this tet's slot src_j
        ys[0 * kMoveBlock] = Z0;
        ys[1 * kMoveBlock] = Z1;
        ys[2 * kMoveBlock] = Z2;
        ys[3 * kMoveBlock] = Z3;
        ys[iH * kMoveBlock] = 0.0;
        const uint32_t p = (perm >> (8 * iH)) & 0xffu;
        Y0 = ys[(p & 3u) * kMoveBlock];
        Y1 = ys[((p >> 2) & 3u) * kMoveBlock];
        Y2 = ys[((p >> 4) & 3u) * kMoveBlock];
        Y3 = ys[(p >> 6) * kMoveBlock];
        t = nb;
        if (++guard > kMaxTrackSteps) { status = 1; break; }
    }
    y1 = Y1; y2 = Y2; y3 = Y3;
    return t;
}

// A wall hit (PICParcel::hitPatch, PICParcel.C:153-184, with the resync of moveForward,
// :126-137): the velocity is synchronised to the hit time by a push of (sf - 1/2) dt with the
// field at the hit point, reflected by the wall model (SpecularReflection.C:55-68,
// MaxwellianThermal.C:55-118, MixedDiffuseSpecular.C:55-131) about the face triangle's
// outward normal (r1 + r2 + r3 of the tet record), and synchronised back by (1/2 - sf) dt,
// that last push only on the first reflection of the step as in the reference.  The
// diffuse draws are Philox keyed by (seed; id, step, species and hit, wall tag).
// One side of a wall hit's contribution to the wall surface fields (PICParcel.C:266-287 before
// the reflection, :299-320 after it): 1 / max(|U.n| |Sf|, vSmall) weighted by 1, m, m |U|^2 / 2
// and m times the tangential velocity, for the totals and the species' own columns.
template <class UT>
__device__ __forceinline__ void wall_fields_add(const MoveArgs<UT>& a, double* row, double nx, double ny, double nz,
                                                double fA, double ux, double uy, double uz) {
    const double m = a.mass;
    const double un = ux * nx + uy * ny + uz * nz;
    const double tx = ux - un * nx, ty = uy - un * ny, tz = uz - un * nz;
    const double inv = 1.0 / fmax(fabs(un) * fA, 1e-300);
    const double ke = 0.5 * m * (ux * ux + uy * uy + uz * uz) * inv;
    double* sp = row + 10 + 9 * a.species;
    atomicAdd(row + 0, inv); atomicAdd(row + 1, m * inv); atomicAdd(row + 2, ke);
    atomicAdd(row + 3, m * tx * inv); atomicAdd(row + 4, m * ty * inv); atomicAdd(row + 5, m * tz * inv);
    atomicAdd(sp + 0, inv); atomicAdd(sp + 1, m * inv);
    atomicAdd(sp + 2, m * tx * inv); atomicAdd(sp + 3, m * ty * inv); atomicAdd(sp + 4, m * tz * inv);
    atomicAdd(sp + 5, ke);
}

template <class UT, int P, int W>
__device__ __forceinline__ void wall_hit(const MoveArgs<UT>& a, const TetGeo& g, double Y1, double Y2, double Y3,
                                         double sf, int patch, vec3<UT>& U, uint32_t id, int hits, int& resyncs,
                                         double* row = nullptr, double fA = 0.0) {
    double nx = g.r[0] + g.r[3] + g.r[6], ny = g.r[1] + g.r[4] + g.r[7], nz = g.r[2] + g.r[5] + g.r[8];
    const double inv = 1.0 / sqrt(nx * nx + ny * ny + nz * nz);
    nx *= inv; ny *= inv; nz *= inv;
    const bool sync = a.wall_sync && a.do_push;
    vec3<UT> E = mk3<UT>(0, 0, 0), B = mk3<UT>(0, 0, 0);
    if (sync) {
        E = interp_E<UT, W>(g, Y1, Y2, Y3, a.Ecell, a.Epoint);
        if (a.rotate) B = mk3<UT>(a.Bcell[3 * g.cell], a.Bcell[3 * g.cell + 1], a.Bcell[3 * g.cell + 2]);
        const UT f = (UT)(sf - 0.5);
        U = push_u<UT, P>(U, E, B, a.k * f, a.kv * f, a.rotate != 0, a.hc_ref != 0);
    }
    double ux = (double)U.x, uy = (double)U.y, uz = (double)U.z;
    const double px = ux, py = uy, pz = uz;
    if (row) wall_fields_add(a, row, nx, ny, nz, fA, ux, uy, uz);
    const uint32_t c = (a.species << 16) | ((uint32_t)hits & 0xffffu);
    uint4 o = rng4(a.seed, id, (uint32_t)a.step, c, kTagWall << 24);
    const double u_mix = u01(o.x, o.y), u_n = u01(o.z, o.w);
    const bool diffuse = a.wall_model == 1 || (a.wall_model == 2 && a.diffuse_fraction > u_mix);
    double un = ux * nx + uy * ny + uz * nz;
    if (!diffuse) {
        if (un > 0.0) { ux -= 2.0 * un * nx; uy -= 2.0 * un * ny; uz -= 2.0 * un * nz; }
    } else {
        double tx = ux - un * nx, ty = uy - un * ny, tz = uz - un * nz;
        for (uint32_t k = 2; sqrt(tx * tx + ty * ty + tz * tz) < 1e-15 && k < 200; k += 2) {
            // incident along the normal: perturb until a tangent exists
            const uint4 p = rng4(a.seed, id, (uint32_t)a.step, c, (kTagWall << 24) | k);
            const uint4 q = rng4(a.seed, id, (uint32_t)a.step, c, (kTagWall << 24) | (k + 1));
            ux *= 0.8 + 0.2 * u01(p.x, p.y);
            uy *= 0.8 + 0.2 * u01(p.z, p.w);
            uz *= 0.8 + 0.2 * u01(q.x, q.y);
            un = ux * nx + uy * ny + uz * nz;
            tx = ux - un * nx; ty = uy - un * ny; tz = uz - un * nz;
        }
        const double it = 1.0 / sqrt(tx * tx + ty * ty + tz * tz);
        tx *= it; ty *= it; tz *= it;
        const double bx = ny * tz - nz * ty, by = nz * tx - nx * tz, bz = nx * ty - ny * tx;   // n x t1
        o = rng4(a.seed, id, (uint32_t)a.step, c, (kTagWall << 24) | 1u);
        const double uc = u01(o.x, o.y), ud = u01(o.z, o.w);
        const double r = sqrt(-2.0 * log(1.0 - uc));
        double sn, cs;
        sincospi(2.0 * ud, &sn, &cs);
        const double n1 = r * cs, n2 = r * sn;
        const double s = sqrt(kKB * a.wall_T[patch] / a.mass);
        const double vn = sqrt(-2.0 * log(fmax(1.0 - u_n, 1e-300)));
        ux = s * (n1 * tx + n2 * bx - vn * nx) + a.wall_U[3 * patch];
        uy = s * (n1 * ty + n2 * by - vn * ny) + a.wall_U[3 * patch + 1];
        uz = s * (n1 * tz + n2 * bz - vn * nz) + a.wall_U[3 * patch + 2];
    }
    U = mk3<UT>((UT)ux, (UT)uy, (UT)uz);
    if (row) {
        // the reflected velocity as the parcel keeps it, then the energy and momentum given to
        // the wall per unit time and area (PICParcel.C:322-337)
        const double vx = (double)U.x, vy = (double)U.y, vz = (double)U.z;
        wall_fields_add(a, row, nx, ny, nz, fA, vx, vy, vz);
        const double s = a.weight / (a.dt_step * fA), m = a.mass;
        atomicAdd(row + 6, s * (0.5 * m * (px * px + py * py + pz * pz) - 0.5 * m * (vx * vx + vy * vy + vz * vz)));
        const double fx = s * (m * px - m * vx), fy = s * (m * py - m * vy), fz = s * (m * pz - m * vz);
        atomicAdd(row + 7, fx); atomicAdd(row + 8, fy); atomicAdd(row + 9, fz);
        double* sp = row + 10 + 9 * a.species;
        atomicAdd(sp + 6, fx); atomicAdd(sp + 7, fy); atomicAdd(sp + 8, fz);
    }
    if (sync && resyncs++ == 0) {
        const UT f = (UT)(0.5 - sf);
        U = push_u<UT, P>(U, E, B, a.k * f, a.kv * f, a.rotate != 0, a.hc_ref != 0);
    }
}

// VolumeWeighting deposit (VolumeWeighting.C:105-123) of one block's parcels, privatised in
// LDS: the cell shares go to a window of kDepCells cells from the block's lowest cell, the
// vertex shares to an open-addressing table of kDepPoints point ids, and each occupied LDS
// slot is flushed with one global atomic (the parcels are in cell order, so a block's
// 256 parcels share a few cells and their points: a few dozen global atomics instead of
// four per parcel).  Shares that fit neither (a cell outside the window, a probe run past
// kDepProbes) go to the global accumulators directly.  The sums are int64 fixed point, so
// the result does not depend on the order.  Called by every thread of the block.
constexpr int kDepCells = 256, kDepPoints = 512, kDepProbes = 16;
struct DepLds {
    unsigned long long cval[kDepCells];
    unsigned long long pval[kDepPoints];
    int pkey[kDepPoints];
    int wmin[8];
};

__device__ void deposit_vw_block(const TetGeo* __restrict__ tets, int cell, int t, double y1, double y2, double y3,
                                 double s42, unsigned long long* acc_cell, unsigned long long* acc_point, DepLds& L) {
    int cm = cell >= 0 ? cell : 0x7fffffff;
    for (int off = 32; off > 0; off >>= 1) cm = min(cm, __shfl_xor(cm, off));
    if (__lane_id() == 0) L.wmin[threadIdx.x >> 6] = cm;
    for (int k = threadIdx.x; k < kDepCells; k += blockDim.x) L.cval[k] = 0ull;
    for (int k = threadIdx.x; k < kDepPoints; k += blockDim.x) { L.pkey[k] = -1; L.pval[k] = 0ull; }
    __syncthreads();
    cm = L.wmin[0];
    for (int w = 1; w < (int)((blockDim.x + 63) >> 6); ++w) cm = min(cm, L.wmin[w]);
    if (cell >= 0) {
        const TetGeo& g = tets[t];
        const double y0 = 1.0 - (y1 + y2 + y3);
        const unsigned long long vc = (unsigned long long)llrint(y0 * s42);
        const unsigned dc = (unsigned)(cell - cm);
        if (dc < (unsigned)kDepCells) atomicAdd(&L.cval[dc], vc);
        else atomicAdd(acc_cell + cell, vc);
        const double yk[3] = {y1, y2, y3};
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const int p = g.vm[k];
            const unsigned long long v = (unsigned long long)llrint(yk[k] * s42);
            unsigned h = ((unsigned)p * 2654435761u) & (kDepPoints - 1);
            bool done = false;
            for (int probe = 0; probe < kDepProbes && !done; ++probe) {
                const int old = atomicCAS(&L.pkey[h], -1, p);
                if (old == -1 || old == p) {
                    atomicAdd(&L.pval[h], v);
                    done = true;
                }
                h = (h + 1) & (kDepPoints - 1);
            }
            if (!done) atomicAdd(acc_point + p, v);
        }
    }
    __syncthreads();
    if (cm != 0x7fffffff)
        for (int k = threadIdx.x; k < kDepCells; k += blockDim.x)
            if (L.cval[k]) atomicAdd(acc_cell + cm + k, L.cval[k]);
    for (int k = threadIdx.x; k < kDepPoints; k += blockDim.x)
        if (L.pkey[k] >= 0) atomicAdd(acc_point + L.pkey[k], L.pval[k]);
}

// One lane per parcel: push (if charged and moving) + track + deposit/count; `block` is
// the block's index within its species.
template <class UT, int P, int W, int DEP>
__device__ __forceinline__ void move_body(const MoveArgs<UT>& a, int64_t block, int set, DepLds* dl, double* ys) {
    const int64_t i = block * blockDim.x + threadIdx.x;
    const int64_t n = a.n_live ? min(*a.n_live, a.n) : a.n;
    // blocks wholly past the live count (the launch bound can be a species' capacity while
    // created parcels are possible) leave at once
    if (block * blockDim.x >= n) return;
    int newcell = -1;
    double y1 = 0, y2 = 0, y3 = 0;
    int t = -1;
    int status = 0;
    bool deleted = false;
    if (i < n) {
        t = a.p.tet[i];
        if (t >= 0) {
            y1 = a.p.y1[i]; y2 = a.p.y2[i]; y3 = a.p.y3[i];
            vec3<UT> U = mk3<UT>(a.p.ux[i], a.p.uy[i], a.p.uz[i]);
            if (a.do_push) {
                const TetGeo& g = a.tets[t];
                vec3<UT> E = interp_E<UT, W>(g, y1, y2, y3, a.Ecell, a.Epoint);
                vec3<UT> B = mk3<UT>(0, 0, 0);
                if (a.rotate) {
                    const int c = g.cell;
                    B = mk3<UT>(a.Bcell[3 * c], a.Bcell[3 * c + 1], a.Bcell[3 * c + 2]);
                }
                U = push_u<UT, P>(U, E, B, a.k, a.kv, a.rotate != 0, a.hc_ref != 0);
            }
            const double Ux = (double)U.x, Uy = (double)U.y, Uz = (double)U.z;
            if (sqrt(Ux * Ux + Uy * Uy + Uz * Uz) >= kC) atomicOr(a.flags, kFlagSuperluminal);
            if (a.do_push) { a.p.ux[i] = U.x; a.p.uy[i] = U.y; a.p.uz[i] = U.z; }
            double sf = 0.0;
            t = track(a, t, y1, y2, y3, U, sf, status, ys);
            a.p.y1[i] = y1; a.p.y2[i] = y2; a.p.y3[i] = y3;
            a.p.tet[i] = t;
            if (status == 3) {
                // on a wall face: counted, placed and deposited by the wall pass
                const unsigned k = atomicAdd(a.wall_n, 1u);
                if (k < a.wall_cap) a.wall_q[k] = WallHit{i, sf, set, 0};
                else atomicOr(a.flags, kFlagTrackBroken);
            } else if (t >= 0) {
                newcell = a.tets[t].cell;
            } else {
                deleted = (status != 2);
            }
        }
    }
    // statistics: deleted parcels and guard hits (one atomic per wave)
    {
        const uint64_t del = __ballot(deleted);
        const uint64_t gd = __ballot(status == 1);
        const uint64_t br = __ballot(status == 2);
        if (__lane_id() == 0) {
            if (del) atomicAdd(a.stats + 0, (unsigned long long)__popcll(del));
            if (gd) atomicAdd(a.stats + 1, (unsigned long long)__popcll(gd));
            if (br) atomicOr(a.flags, kFlagTrackBroken);
        }
    }
    // cell count (sort histogram; for CellAverage also the deposit) and the parcel's slot
    // in its new cell, which the sort's scatter uses directly
    const int slot = wave_aggregated_inc(a.count, newcell);
    if (a.place && i < n) a.place[i] = make_int2(newcell, slot);
    if constexpr (DEP == kVolumeWeighting)
        deposit_vw_block(a.tets, newcell, t, y1, y2, y3, a.s42, a.acc_cell, a.acc_point, *dl);
}

// every moving species in one launch: blocks [block0[s], block0[s+1]) belong to species s,
// so the short passes of slow species overlap the long tracks of fast ones
template <class UT, int P, int W, int DEP>
__global__ void __launch_bounds__(kMoveBlock) k_move(MoveSet<UT> m) {
    __shared__ double ys_all[4 * kMoveBlock];
    const int64_t b = blockIdx.x;
    int s = 0;
    while (s + 1 < m.ns && b >= m.block0[s + 1]) ++s;
    if constexpr (DEP == kVolumeWeighting) {
        __shared__ DepLds dl;
        move_body<UT, P, W, DEP>(m.a[s], b - m.block0[s], s, &dl, ys_all + threadIdx.x);
    } else {
        move_body<UT, P, W, DEP>(m.a[s], b - m.block0[s], s, nullptr, ys_all + threadIdx.x);
    }
}

// SimpleInjection's reflection on its own patch (SimpleInjection.C:245-252 after the hit's
// sync of a plain patch, PICParcel.C:158-161; specularReflect, PICParcel.C:347-357): the
// velocity synchronised to the hit time and reflected about the face's outward normal, and
// not synchronised back (the reference sets no resync for it).
template <class UT, int P, int W>
__device__ __forceinline__ void patch_reflect(const MoveArgs<UT>& a, const TetGeo& g, double Y1, double Y2, double Y3,
                                              double sf, vec3<UT>& U) {
    double nx = g.r[0] + g.r[3] + g.r[6], ny = g.r[1] + g.r[4] + g.r[7], nz = g.r[2] + g.r[5] + g.r[8];
    const double inv = 1.0 / sqrt(nx * nx + ny * ny + nz * nz);
    nx *= inv; ny *= inv; nz *= inv;
    if (a.wall_sync && a.do_push) {
        const vec3<UT> E = interp_E<UT, W>(g, Y1, Y2, Y3, a.Ecell, a.Epoint);
        vec3<UT> B = mk3<UT>(0, 0, 0);
        if (a.rotate) B = mk3<UT>(a.Bcell[3 * g.cell], a.Bcell[3 * g.cell + 1], a.Bcell[3 * g.cell + 2]);
        const UT f = (UT)(sf - 0.5);
        U = push_u<UT, P>(U, E, B, a.k * f, a.kv * f, a.rotate != 0, a.hc_ref != 0);
    }
    double ux = (double)U.x, uy = (double)U.y, uz = (double)U.z;
    const double un = ux * nx + uy * ny + uz * nz;
    if (un > 0.0) { ux -= 2.0 * un * nx; uy -= 2.0 * un * ny; uz -= 2.0 * un * nz; }
    U = mk3<UT>((UT)ux, (UT)uy, (UT)uz);
}

template <class UT>
__device__ void emit(const MoveArgs<UT>& m, const InjectTri* __restrict__ tri, const double* __restrict__ cum,
                     int ntri, const int32_t* __restrict__ cell_tet0, int vmodel, double vt, bool by_area,
                     uint32_t k, uint32_t c, uint32_t tag, int& t, double& y1, double& y2, double& y3, double U[3]);

// The wall pass: every queued parcel is reflected on its wall (wall_hit), tracks the rest of
// its step (reaching further walls, reflected here in turn: the second and later ones
// without the resync), and is then counted, placed and deposited as the move does for the
// others.  A grid-stride loop over the queue length read on the device; the statistics
// (deleted, guard, wall hits) are summed per wave and added once at the end.
template <class UT, int P, int W, int DEP>
__global__ void __launch_bounds__(kMoveBlock) k_wall(MoveSet<UT> m) {
    __shared__ double ys_all[4 * kMoveBlock];
    double* ys = ys_all + threadIdx.x;
    const MoveArgs<UT>& a0 = m.a[0];
    const unsigned nq = min(*a0.wall_n, a0.wall_cap);
    unsigned long long n_del = 0, n_guard = 0, n_hits = 0;
    for (unsigned k = blockIdx.x * blockDim.x + threadIdx.x; k < nq; k += gridDim.x * blockDim.x) {
        const WallHit h = a0.wall_q[k];
        const MoveArgs<UT>& a = m.a[h.set];
        const int64_t i = h.i;
        int t = a.p.tet[i];
        double y1 = a.p.y1[i], y2 = a.p.y2[i], y3 = a.p.y3[i];
        vec3<UT> U = mk3<UT>(a.p.ux[i], a.p.uy[i], a.p.uz[i]);
        const uint32_t id = a.p.id[i];
        double sf = h.sf;
        int status = 3, hits = 0, resyncs = 0;
        while (status == 3) {
            const TetGeo& g = a.tets[t];
            const int patch = -1 - g.nbr[0];
            const int kind = a.patch_kind[patch];
            bool on_face = true;
            if (kind == 1) {
                double* row = nullptr;
                double fA = 0.0;
                if (a.wall_acc) {
                    const int bf = a.tet_face[t] - a.n_internal;
                    row = a.wall_acc + (size_t)bf * a.wall_stride;
                    fA = a.face_area[bf];
                }
                wall_hit<UT, P, W>(a, g, y1, y2, y3, sf, patch, U, id, hits++, resyncs, row, fA);
            } else if (kind == 2) {
                patch_reflect<UT, P, W>(a, g, y1, y2, y3, sf, U);
                ++hits;
            } else {
                // SimpleInjection.C:253-270: a new parcel from the injector's emitter at the hit's
                // step fraction (ParticleEmitter.C:136-221), synchronised to it, in this slot
                double Un[3];
                const uint32_t c = (a.species << 16) | ((uint32_t)hits & 0xffffu);
                emit(a, a.re_tri, nullptr, a.re_ntri, a.cell_tet0, a.re_vmodel, a.re_vt, false, id, c, kTagReemit, t,
                     y1, y2, y3, Un);
                U = mk3<UT>((UT)Un[0], (UT)Un[1], (UT)Un[2]);
                if (a.wall_sync && a.do_push && t >= 0) {
                    const TetGeo& h = a.tets[t];
                    const vec3<UT> E = interp_E<UT, W>(h, y1, y2, y3, a.Ecell, a.Epoint);
                    vec3<UT> B = mk3<UT>(0, 0, 0);
                    if (a.rotate) B = mk3<UT>(a.Bcell[3 * h.cell], a.Bcell[3 * h.cell + 1], a.Bcell[3 * h.cell + 2]);
                    const UT f = (UT)(sf - 0.5);
                    U = push_u<UT, P>(U, E, B, a.k * f, a.kv * f, a.rotate != 0, a.hc_ref != 0);
                }
                ++hits;
                on_face = false;
                if (t < 0) { status = 2; break; }
            }
            t = track(a, t, y1, y2, y3, U, sf, status, ys, on_face);
            if (status == 3 && hits > kMaxTrackSteps) { status = 1; break; }
        }
        a.p.ux[i] = U.x; a.p.uy[i] = U.y; a.p.uz[i] = U.z;
        a.p.y1[i] = y1; a.p.y2[i] = y2; a.p.y3[i] = y3;
        a.p.tet[i] = t;
        n_hits += (unsigned long long)hits;
        n_guard += status == 1 ? 1ull : 0ull;
        if (status == 2) atomicOr(a.flags, kFlagTrackBroken);
        if (t < 0) {
            n_del += status != 2 ? 1ull : 0ull;
            if (a.place) a.place[i] = make_int2(-1, -1);
            continue;
        }
        const TetGeo& g = a.tets[t];
        const int cell = g.cell;
        const int slot = atomicAdd(a.count + cell, 1);
        if (a.place) a.place[i] = make_int2(cell, slot);
        if constexpr (DEP == kVolumeWeighting) {
            // VolumeWeighting.C:105-123 with the move's fixed-point rounding
            const double y0 = 1.0 - (y1 + y2 + y3);
            atomicAdd(a.acc_cell + cell, (unsigned long long)llrint(y0 * a.s42));
            atomicAdd(a.acc_point + g.vm[0], (unsigned long long)llrint(y1 * a.s42));
            atomicAdd(a.acc_point + g.vm[1], (unsigned long long)llrint(y2 * a.s42));
            atomicAdd(a.acc_point + g.vm[2], (unsigned long long)llrint(y3 * a.s42));
        }
    }
    // every species' MoveArgs carries the same statistics words
    for (int off = 32; off > 0; off >>= 1) {
        n_del += __shfl_xor(n_del, off);
        n_guard += __shfl_xor(n_guard, off);
        n_hits += __shfl_xor(n_hits, off);
    }
    if (__lane_id() == 0) {
        if (n_del) atomicAdd(a0.stats + 0, n_del);
        if (n_guard) atomicAdd(a0.stats + 1, n_guard);
        if (n_hits) atomicAdd(a0.stats + kStatsWallHits, n_hits);
    }
}

template <class UT, int P, int W>
__global__ void __launch_bounds__(256) k_push(MoveArgs<UT> a) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (a.n_live ? min(*a.n_live, a.n) : a.n)) return;
    const int t = a.p.tet[i];
    if (t < 0) return;
    const double y1 = a.p.y1[i], y2 = a.p.y2[i], y3 = a.p.y3[i];
    vec3<UT> U = mk3<UT>(a.p.ux[i], a.p.uy[i], a.p.uz[i]);
    const TetGeo& g = a.tets[t];
    vec3<UT> E = interp_E<UT, W>(g, y1, y2, y3, a.Ecell, a.Epoint);
    vec3<UT> B = mk3<UT>(0, 0, 0);
    if (a.rotate) {
        const int c = g.cell;
        B = mk3<UT>(a.Bcell[3 * c], a.Bcell[3 * c + 1], a.Bcell[3 * c + 2]);
    }
    U = push_u<UT, P>(U, E, B, a.k, a.kv, a.rotate != 0, a.hc_ref != 0);
    a.p.ux[i] = U.x; a.p.uy[i] = U.y; a.p.uz[i] = U.z;
}

// Deposit / count only (calculateFields without a preceding move).
template <class UT, int DEP>
__global__ void __launch_bounds__(256) k_deposit(MoveArgs<UT> a) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t n = a.n_live ? min(*a.n_live, a.n) : a.n;
    int cell = -1, t = -1;
    double y1 = 0, y2 = 0, y3 = 0;
    if (i < n) {
        t = a.p.tet[i];
        if (t >= 0) { cell = a.tets[t].cell; y1 = a.p.y1[i]; y2 = a.p.y2[i]; y3 = a.p.y3[i]; }
    }
    const int slot = wave_aggregated_inc(a.count, cell);
    if (a.place && i < n) a.place[i] = make_int2(cell, slot);
    if constexpr (DEP == kVolumeWeighting) {
        __shared__ DepLds dl;
        deposit_vw_block(a.tets, cell, t, y1, y2, y3, a.s42, a.acc_cell, a.acc_point, dl);
    }
}

// Point location inside the given cell: the tet with the largest minimum barycentric
// coordinate (y_i = r_i.(x - C)/det, the tracking kernel's own transform).
template <class UT>
__global__ void k_locate(LocateArgs<UT> a) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const int c = a.cell[i];
    const double px = a.x[3 * i], py = a.x[3 * i + 1], pz = a.x[3 * i + 2];
    const double cx = a.C[3 * c], cy = a.C[3 * c + 1], cz = a.C[3 * c + 2];
    const double dx = px - cx, dy = py - cy, dz = pz - cz;
    int best = -1;
    double bestmin = -1e300, b1 = 0, b2 = 0, b3 = 0;
    for (int t = a.cell_tet0[c]; t < a.cell_tet0[c + 1]; ++t) {
        const TetGeo& g = a.tets[t];
        double v1 = (g.r[0] * dx + g.r[1] * dy + g.r[2] * dz) / g.det;
        double v2 = (g.r[3] * dx + g.r[4] * dy + g.r[5] * dz) / g.det;
        double v3_ = (g.r[6] * dx + g.r[7] * dy + g.r[8] * dz) / g.det;
        double v0 = 1.0 - (v1 + v2 + v3_);
        double mn = fmin(fmin(v0, v1), fmin(v2, v3_));
        if (mn > bestmin) { bestmin = mn; best = t; b1 = v1; b2 = v2; b3 = v3_; }
    }
    if (bestmin < -1e-9) atomicAdd(a.fails, 1u);
    const int64_t o = a.offset + i;
    a.p.y1[o] = b1; a.p.y2[o] = b2; a.p.y3[o] = b3;
    a.p.tet[o] = best;
    a.p.ux[o] = (UT)a.U[3 * i]; a.p.uy[o] = (UT)a.U[3 * i + 1]; a.p.uz[o] = (UT)a.U[3 * i + 2];
    a.p.id[o] = a.id[i];
}

// ParticleEmitter::emitParticle (ParticleEmitter.C:223-349) for parcel k of one injection: a
// triangle of the patch drawn by area (the cumulative table's last entry below the drawn
// area), a uniform point in it (barycentric (1 - t, (1 - s) t, s t) with t = sqrt(u)), moved
// into the cell by a in [0.1, 0.5) of its normal distance from the cell centre, its empty
// components set to zero (meshTools::constrainDirection), located in the owner cell (the
// tet with the largest minimum barycentric coordinate, as k_locate), and a velocity from
// the model (ParticleEmitter.C:357-535) in the frame (inward normal, t1, t2).  The draws
// are Philox keyed by (seed; k, step, injector and species, tag and draw).
// by_area false: the triangle uniformly over the patch's triangles, as the re-emission of a
// reflected parcel draws it (emitParticleExplicit, ParticleEmitter.C:143-146).
template <class UT>
__device__ void emit(const MoveArgs<UT>& m, const InjectTri* __restrict__ tri, const double* __restrict__ cum,
                     int ntri, const int32_t* __restrict__ cell_tet0, int vmodel, double vt, bool by_area,
                     uint32_t k, uint32_t c, uint32_t tag, int& t, double& y1, double& y2, double& y3, double U[3]) {
    uint4 o = rng4(m.seed, k, (uint32_t)m.step, c, tag << 24);
    const double u = u01(o.x, o.y), s = u01(o.z, o.w);
    int lo = 0;
    if (by_area) {
        const double area = u * cum[ntri];
        int hi = ntri - 1;   // the last triangle whose cumulative start lies below the area
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (cum[mid] < area) lo = mid; else hi = mid - 1;
        }
    } else {
        lo = min((int)(u * (double)ntri), ntri - 1);
    }
    const InjectTri& T = tri[lo];
    o = rng4(m.seed, k, (uint32_t)m.step, c, (tag << 24) | 1u);
    const double tt = sqrt(u01(o.x, o.y)), frac = 0.1 + 0.4 * u01(o.z, o.w);
    const double wa = 1.0 - tt, wb = (1.0 - s) * tt, wc = s * tt;
    double x[3], dn = 0.0;
    for (int d = 0; d < 3; ++d) {
        x[d] = wa * T.a[d] + wb * T.b[d] + wc * T.c[d];
        dn += (x[d] - T.pc[d]) * T.n[d];
    }
    dn = fabs(dn);
    for (int d = 0; d < 3; ++d) {
        x[d] -= frac * (dn * T.n[d]);
        if ((m.track_mask >> d) & 1u) x[d] = 0.0;
    }
    const int cell = T.cell;
    const double dx = x[0] - T.pc[0], dy = x[1] - T.pc[1], dz = x[2] - T.pc[2];
    int best = -1;
    double bestmin = -1e300;
    y1 = y2 = y3 = 0.0;
    for (int q = cell_tet0[cell]; q < cell_tet0[cell + 1]; ++q) {
        const TetGeo& g = m.tets[q];
        const double v1 = (g.r[0] * dx + g.r[1] * dy + g.r[2] * dz) / g.det;
        const double v2 = (g.r[3] * dx + g.r[4] * dy + g.r[5] * dz) / g.det;
        const double v3 = (g.r[6] * dx + g.r[7] * dy + g.r[8] * dz) / g.det;
        const double mn = fmin(fmin(1.0 - (v1 + v2 + v3), v1), fmin(v2, v3));
        if (mn > bestmin) { bestmin = mn; best = q; y1 = v1; y2 = v2; y3 = v3; }
    }
    t = best;
    // velocity in the frame of the face: n inward, t1, t2 tangential
    double vn = 0.0, n1 = 0.0, n2 = 0.0;
    if (vmodel != 1) {
        o = rng4(m.seed, k, (uint32_t)m.step, c, (tag << 24) | 2u);
        box_muller(u01(o.x, o.y), u01(o.z, o.w), n1, n2);
        n1 *= vt; n2 *= vt;
    }
    const double vmp = sqrt(2.0) * vt;   // maxwellianMostProbableSpeed, PICCloudI.H:541-548
    if (vmodel == 0) {
        // Bird eqn 12.5 by acceptance-rejection (ParticleEmitter.C:380-398)
        double un = 0.0;
        for (uint32_t j = 4; j < 4 + 2000; ++j) {
            o = rng4(m.seed, k, (uint32_t)m.step, c, (tag << 24) | j);
            un = 3.0 * (2.0 * u01(o.x, o.y) - 1.0);
            const double P = un < 0.0 ? -1.0 : 2.0 * un / sqrt(2.0) * exp(0.5 - un * un);
            if (!(P < u01(o.z, o.w))) break;
        }
        vn = vmp * un;
    } else if (vmodel == 1) {
        vn = vmp;
    } else {
        o = rng4(m.seed, k, (uint32_t)m.step, c, (tag << 24) | 3u);
        if (vmodel == 2) {
            double g1, g2;
            box_muller(u01(o.x, o.y), u01(o.z, o.w), g1, g2);
            vn = fabs(g1) * vt;
        } else {
            vn = vt * sqrt(-2.0 * log(1.0 - u01(o.x, o.y)));
        }
    }
    for (int d = 0; d < 3; ++d) U[d] = n1 * T.t1[d] + n2 * T.t2[d] - vn * T.n[d];
}

// SimpleInjection::injection for one species of one injector: parcels base + k, k < count,
// after the live ones (base read before any is added; the live count is raised afterwards by
// k_live_add), with new ids; charged species are synchronised back by half a step
// (ParticleEmitter.C:338, PICParcel.C:214-227) with the field at their position.
template <class UT, int P, int W>
__global__ void __launch_bounds__(256) k_inject(InjectArgs<UT> a) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= a.count) return;
    const MoveArgs<UT>& m = a.m;
    const int64_t i = *m.n_live + k;
    if (i >= a.cap) { atomicOr(m.flags, kFlagCapacity); return; }
    int t;
    double y1, y2, y3, U[3];
    emit(m, a.tri, a.cum, a.ntri, a.cell_tet0, a.vmodel, a.vt, true, (uint32_t)k, (a.model << 16) | m.species,
         kTagInject, t, y1, y2, y3, U);
    vec3<UT> u = mk3<UT>((UT)U[0], (UT)U[1], (UT)U[2]);
    if (a.sync && t >= 0) {
        const TetGeo& g = m.tets[t];
        const vec3<UT> E = interp_E<UT, W>(g, y1, y2, y3, m.Ecell, m.Epoint);
        vec3<UT> B = mk3<UT>(0, 0, 0);
        if (m.rotate) B = mk3<UT>(m.Bcell[3 * g.cell], m.Bcell[3 * g.cell + 1], m.Bcell[3 * g.cell + 2]);
        const UT f = (UT)-0.5;
        u = push_u<UT, P>(u, E, B, m.k * f, m.kv * f, m.rotate != 0, m.hc_ref != 0);
    }
    m.p.y1[i] = y1; m.p.y2[i] = y2; m.p.y3[i] = y3;
    m.p.tet[i] = t;
    m.p.ux[i] = u.x; m.p.uy[i] = u.y; m.p.uz[i] = u.z;
    const uint32_t c = (a.model << 16) | m.species;
    m.p.id[i] = rng4(m.seed, (uint32_t)k, (uint32_t)m.step, c, (kTagNewId << 24) | 3u).x;
}

// the live count after an injection: raised by the parcels that fitted
__global__ void k_live_add(int64_t* n, int64_t add, int64_t cap) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *n = min(*n + add, max(cap, *n));
}

template <class UT>
__global__ void k_positions(PartSoA<UT> p, int64_t n, const TetGeo* __restrict__ tets, const double* __restrict__ C,
                            const double* __restrict__ pts, double* __restrict__ x) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int t = p.tet[i];
    if (t < 0) { x[3 * i] = x[3 * i + 1] = x[3 * i + 2] = __builtin_nan(""); return; }
    const TetGeo& g = tets[t];
    const double y1 = p.y1[i], y2 = p.y2[i], y3 = p.y3[i];
    const double y0 = 1.0 - (y1 + y2 + y3);
    for (int d = 0; d < 3; ++d) {
        double v = y0 * C[3 * g.cell + d];
        v = v + y1 * pts[3 * g.vid[0] + d];
        v = v + y2 * pts[3 * g.vid[1] + d];
        v = v + y3 * pts[3 * g.vid[2] + d];
        x[3 * i + d] = v;
    }
}

// picInitialise Uniform + Maxwellian: parcel j of cell c sits in a tet chosen with
// probability proportional to its volume, uniformly inside it (normalised exponential
// spacings), with velocity sqrt(kT/m) N(0,1)^3 + drift.
template <class UT>
__global__ void k_init(InitArgs<UT> a) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const int c = (int)(i / a.nppc);
    const uint32_t j = (uint32_t)(i - (int64_t)c * a.nppc);
    uint4 r0 = rng4(a.seed, (uint32_t)c, j, 0u, (kTagInit << 24));
    uint4 r1 = rng4(a.seed, (uint32_t)c, j, 1u, (kTagInit << 24));
    uint4 r2 = rng4(a.seed, (uint32_t)c, j, 2u, (kTagInit << 24));
    const int t0 = a.cell_tet0[c], t1 = a.cell_tet0[c + 1];
    double vol = 0.0;
    for (int t = t0; t < t1; ++t) vol += a.tets[t].det;
    double pick = u01(r0.x, r0.y) * vol, acc = 0.0;
    int tt = t1 - 1;
    for (int t = t0; t < t1; ++t) { acc += a.tets[t].det; if (pick < acc) { tt = t; break; } }
    double e0 = -log(1.0 - u01(r0.z, r0.w)), e1 = -log(1.0 - u01(r1.x, r1.y));
    double e2 = -log(1.0 - u01(r1.z, r1.w)), e3 = -log(1.0 - u01(r2.x, r2.y));
    double s = (e0 + e1) + (e2 + e3);
    const int64_t o = a.offset + i;
    a.p.y1[o] = e1 / s; a.p.y2[o] = e2 / s; a.p.y3[o] = e3 / s;
    a.p.tet[o] = tt;
    uint4 r3 = rng4(a.seed, (uint32_t)c, j, 3u, (kTagInit << 24));
    double ua = u01(r2.z, r2.w), ub = u01(r3.x, r3.y), uc = u01(r3.z, r3.w);
    uint4 r4 = rng4(a.seed, (uint32_t)c, j, 4u, (kTagInit << 24));
    double ud = u01(r4.x, r4.y);
    double ra = sqrt(-2.0 * log(1.0 - ua)), rb = sqrt(-2.0 * log(1.0 - uc));
    double n1 = ra * cos(2.0 * kPi * ub), n2 = ra * sin(2.0 * kPi * ub), n3 = rb * cos(2.0 * kPi * ud);
    a.p.ux[o] = (UT)(a.vth * n1 + a.drift[0]);
    a.p.uy[o] = (UT)(a.vth * n2 + a.drift[1]);
    a.p.uz[o] = (UT)(a.vth * n3 + a.drift[2]);
    a.p.id[o] = a.id0 + (uint32_t)i;
}

// Counting-sort scatter (buildCellOccupancy, PICCloud.C:327-354): live parcels go to
// start[cell] + slot, with (cell, slot) handed over by the counting pass (the slot comes
// from a wave-aggregated atomic there, so the order inside a cell is not fixed; every
// consumer that depends on it orders the cell's parcels by id or by a random key).  Deleted
// parcels carry cell -1.  A species marked index[s] keeps its parcels in place and gets the
// placement as an index (perm, plus its ids in cell order for the collision shuffle keys).
// Every species in one launch, blocks [block0[s], block0[s+1]) for species s.
template <class UT>
__global__ void __launch_bounds__(256) k_scatter(ScatterSet<UT> m) {
    const int64_t b = blockIdx.x;
    int sp = 0;
    while (sp + 1 < m.ns && b >= m.block0[sp + 1]) ++sp;
    const PartSoA<UT>& src = m.src[sp];
    const PartSoA<UT>& dst = m.dst[sp];
    const int2* __restrict__ place = m.place[sp];
    const int* __restrict__ start = m.start[sp];
    const int64_t i = (b - m.block0[sp]) * blockDim.x + threadIdx.x;
    // the live count before the sort (the scan left the sorted count aside); it can exceed
    // the arrays after a capacity overflow (flagged)
    if (i >= min(*m.n_live[sp], m.n[sp])) return;
    const int2 pc = place[i];
    if (pc.x < 0) return;
    const int64_t o = (int64_t)start[pc.x] + pc.y;
    if (m.index[sp]) {
        // the parcels stay; kernels that walk a cell's parcels read them through perm
        m.perm[sp][o] = (int)i;
        m.id_cell[sp][o] = src.id[i];
        return;
    }
    dst.y1[o] = src.y1[i]; dst.y2[o] = src.y2[i]; dst.y3[o] = src.y3[i];
    dst.ux[o] = src.ux[i]; dst.uy[o] = src.uy[i]; dst.uz[o] = src.uz[i];
    dst.tet[o] = src.tet[i];
    dst.id[o] = src.id[i];
}

// ---------------------------------------------------------------- instantiation / launch

template <class UT>
static hipError_t launch_move_t(MoveSet<UT> m, int pusher, int weighting, int dep, hipStream_t s) {
    int64_t blocks = 0;
    for (int k = 0; k < m.ns; ++k) {
        m.block0[k] = blocks;
        blocks += (m.a[k].n + kMoveBlock - 1) / kMoveBlock;
    }
    m.block0[m.ns] = blocks;
    if (blocks <= 0) return hipSuccess;
    const dim3 blk(kMoveBlock), grd((unsigned)blocks);
#define PICMI_MOVE_CASE(P, W, D)                                                     \
    if (pusher == P && weighting == W && dep == D) {                                 \
        hipLaunchKernelGGL((k_move<UT, P, W, D>), grd, blk, 0, s, m);                \
        return hipGetLastError();                                                    \
    }
#define PICMI_MOVE_P(P)                                                              \
    PICMI_MOVE_CASE(P, kCellAverage, kCellAverage)                                   \
    PICMI_MOVE_CASE(P, kCellAverage, kVolumeWeighting)                               \
    PICMI_MOVE_CASE(P, kVolumeWeighting, kCellAverage)                               \
    PICMI_MOVE_CASE(P, kVolumeWeighting, kVolumeWeighting)
    PICMI_MOVE_P(kPushNone)
    PICMI_MOVE_P(kPushBoris)
    PICMI_MOVE_P(kPushBorisNR)
    PICMI_MOVE_P(kPushVay)
    PICMI_MOVE_P(kPushHC)
#undef PICMI_MOVE_P
#undef PICMI_MOVE_CASE
    return hipErrorInvalidValue;
}

template <class UT>
static hipError_t launch_wall_t(const MoveSet<UT>& m, int pusher, int weighting, int dep, hipStream_t s) {
    if (m.ns <= 0 || !m.a[0].wall_q) return hipSuccess;
    // the queue length stays on the device: a fixed grid strides over it
    const dim3 blk(kMoveBlock), grd(256);
#define PICMI_WALL_CASE(P, W, D)                                                     \
    if (pusher == P && weighting == W && dep == D) {                                 \
        hipLaunchKernelGGL((k_wall<UT, P, W, D>), grd, blk, 0, s, m);                \
        return hipGetLastError();                                                    \
    }
#define PICMI_WALL_P(P)                                                              \
    PICMI_WALL_CASE(P, kCellAverage, kCellAverage)                                   \
    PICMI_WALL_CASE(P, kCellAverage, kVolumeWeighting)                               \
    PICMI_WALL_CASE(P, kVolumeWeighting, kCellAverage)                               \
    PICMI_WALL_CASE(P, kVolumeWeighting, kVolumeWeighting)
    PICMI_WALL_P(kPushNone)
    PICMI_WALL_P(kPushBoris)
    PICMI_WALL_P(kPushBorisNR)
    PICMI_WALL_P(kPushVay)
    PICMI_WALL_P(kPushHC)
#undef PICMI_WALL_P
#undef PICMI_WALL_CASE
    return hipErrorInvalidValue;
}

hipError_t launch_wall(const MoveSet<double>& m, int pusher, int weighting, int dep, hipStream_t s) {
    return launch_wall_t(m, pusher, weighting, dep, s);
}
hipError_t launch_wall(const MoveSet<float>& m, int pusher, int weighting, int dep, hipStream_t s) {
    return launch_wall_t(m, pusher, weighting, dep, s);
}

template <class UT>
static hipError_t launch_push_t(const MoveArgs<UT>& a, int pusher, int weighting, hipStream_t s) {
    if (a.n <= 0) return hipSuccess;
    const dim3 blk(256), grd((unsigned)((a.n + 255) / 256));
#define PICMI_PUSH_CASE(P, W)                                                        \
    if (pusher == P && weighting == W) {                                             \
        hipLaunchKernelGGL((k_push<UT, P, W>), grd, blk, 0, s, a);                   \
        return hipGetLastError();                                                    \
    }
    PICMI_PUSH_CASE(kPushNone, kCellAverage) PICMI_PUSH_CASE(kPushNone, kVolumeWeighting)
    PICMI_PUSH_CASE(kPushBoris, kCellAverage) PICMI_PUSH_CASE(kPushBoris, kVolumeWeighting)
    PICMI_PUSH_CASE(kPushBorisNR, kCellAverage) PICMI_PUSH_CASE(kPushBorisNR, kVolumeWeighting)
    PICMI_PUSH_CASE(kPushVay, kCellAverage) PICMI_PUSH_CASE(kPushVay, kVolumeWeighting)
    PICMI_PUSH_CASE(kPushHC, kCellAverage) PICMI_PUSH_CASE(kPushHC, kVolumeWeighting)
#undef PICMI_PUSH_CASE
    return hipErrorInvalidValue;
}

template <class UT>
static hipError_t launch_inject_t(const InjectArgs<UT>& a, int pusher, int weighting, hipStream_t s) {
    if (a.count <= 0) return hipSuccess;
    if (a.ntri <= 0 || !a.m.n_live) return hipErrorInvalidValue;
    const dim3 blk(256), grd((unsigned)((a.count + 255) / 256));
    bool done = false;
#define PICMI_INJECT_CASE(P, W)                                                      \
    if (!done && pusher == P && weighting == W) {                                    \
        hipLaunchKernelGGL((k_inject<UT, P, W>), grd, blk, 0, s, a);                 \
        done = true;                                                                 \
    }
    PICMI_INJECT_CASE(kPushNone, kCellAverage) PICMI_INJECT_CASE(kPushNone, kVolumeWeighting)
    PICMI_INJECT_CASE(kPushBoris, kCellAverage) PICMI_INJECT_CASE(kPushBoris, kVolumeWeighting)
    PICMI_INJECT_CASE(kPushBorisNR, kCellAverage) PICMI_INJECT_CASE(kPushBorisNR, kVolumeWeighting)
    PICMI_INJECT_CASE(kPushVay, kCellAverage) PICMI_INJECT_CASE(kPushVay, kVolumeWeighting)
    PICMI_INJECT_CASE(kPushHC, kCellAverage) PICMI_INJECT_CASE(kPushHC, kVolumeWeighting)
#undef PICMI_INJECT_CASE
    if (!done) return hipErrorInvalidValue;
    if (hipError_t e = hipGetLastError()) return e;
    hipLaunchKernelGGL(k_live_add, dim3(1), dim3(64), 0, s, const_cast<int64_t*>(a.m.n_live), a.count, a.cap);
    return hipGetLastError();
}
hipError_t launch_inject(const InjectArgs<double>& a, int pusher, int weighting, hipStream_t s) {
    return launch_inject_t(a, pusher, weighting, s);
}
hipError_t launch_inject(const InjectArgs<float>& a, int pusher, int weighting, hipStream_t s) {
    return launch_inject_t(a, pusher, weighting, s);
}

hipError_t launch_move(const MoveSet<double>& m, int pusher, int weighting, int dep, hipStream_t s) {
    return launch_move_t(m, pusher, weighting, dep, s);
}
hipError_t launch_move(const MoveSet<float>& m, int pusher, int weighting, int dep, hipStream_t s) {
    return launch_move_t(m, pusher, weighting, dep, s);
}
hipError_t launch_push(const MoveArgs<double>& a, int pusher, int weighting, hipStream_t s) {
    return launch_push_t(a, pusher, weighting, s);
}
hipError_t launch_push(const MoveArgs<float>& a, int pusher, int weighting, hipStream_t s) {
    return launch_push_t(a, pusher, weighting, s);
}

template <class UT>
static hipError_t launch_deposit_t(const MoveArgs<UT>& a, int dep, hipStream_t s) {
    if (a.n <= 0) return hipSuccess;
    const dim3 blk(256), grd((unsigned)((a.n + 255) / 256));
    if (dep == kVolumeWeighting) hipLaunchKernelGGL((k_deposit<UT, kVolumeWeighting>), grd, blk, 0, s, a);
    else hipLaunchKernelGGL((k_deposit<UT, kCellAverage>), grd, blk, 0, s, a);
    return hipGetLastError();
}
hipError_t launch_deposit(const MoveArgs<double>& a, int dep, hipStream_t s) { return launch_deposit_t(a, dep, s); }
hipError_t launch_deposit(const MoveArgs<float>& a, int dep, hipStream_t s) { return launch_deposit_t(a, dep, s); }

template <class UT>
static hipError_t launch_locate_t(const LocateArgs<UT>& a, hipStream_t s) {
    if (a.n <= 0) return hipSuccess;
    hipLaunchKernelGGL((k_locate<UT>), dim3((unsigned)((a.n + 255) / 256)), dim3(256), 0, s, a);
    return hipGetLastError();
}
hipError_t launch_locate(const LocateArgs<double>& a, hipStream_t s) { return launch_locate_t(a, s); }
hipError_t launch_locate(const LocateArgs<float>& a, hipStream_t s) { return launch_locate_t(a, s); }

template <class UT>
static hipError_t launch_positions_t(const PartSoA<UT>& p, int64_t n, const TetGeo* tets, const double* C,
                                     const double* pts, double* x, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL((k_positions<UT>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n, tets, C, pts, x);
    return hipGetLastError();
}
hipError_t launch_positions(const PartSoA<double>& p, int64_t n, const TetGeo* tets, const double* C,
                            const double* pts, double* x, hipStream_t s) {
    return launch_positions_t(p, n, tets, C, pts, x, s);
}
hipError_t launch_positions(const PartSoA<float>& p, int64_t n, const TetGeo* tets, const double* C,
                            const double* pts, double* x, hipStream_t s) {
    return launch_positions_t(p, n, tets, C, pts, x, s);
}

template <class UT>
static hipError_t launch_init_t(const InitArgs<UT>& a, hipStream_t s) {
    if (a.n <= 0) return hipSuccess;
    hipLaunchKernelGGL((k_init<UT>), dim3((unsigned)((a.n + 255) / 256)), dim3(256), 0, s, a);
    return hipGetLastError();
}
hipError_t launch_init(const InitArgs<double>& a, hipStream_t s) { return launch_init_t(a, s); }
hipError_t launch_init(const InitArgs<float>& a, hipStream_t s) { return launch_init_t(a, s); }

template <class UT>
static hipError_t launch_scatter_t(ScatterSet<UT> m, hipStream_t s) {
    int64_t blocks = 0;
    for (int k = 0; k < m.ns; ++k) {
        m.block0[k] = blocks;
        blocks += (m.n[k] + 255) / 256;
    }
    m.block0[m.ns] = blocks;
    if (blocks <= 0) return hipSuccess;
    hipLaunchKernelGGL((k_scatter<UT>), dim3((unsigned)blocks), dim3(256), 0, s, m);
    return hipGetLastError();
}
hipError_t launch_scatter(const ScatterSet<double>& m, hipStream_t s) { return launch_scatter_t(m, s); }
hipError_t launch_scatter(const ScatterSet<float>& m, hipStream_t s) { return launch_scatter_t(m, s); }

}  // namespace picmi
