// Kernel argument blocks and launch wrappers of libpicmi.
#pragma once
#include "device.h"

namespace picmi {

// a parcel standing on a wall face after the move: its index, its species' slot in the
// move set, and the step fraction done
struct WallHit {
    int64_t i;
    double sf;
    int32_t set;
    int32_t pad;
};

struct InjectTri;

template <class UT>
struct MoveArgs {
    PartSoA<UT> p;
    int64_t n;                       // launch bound (>= the live count)
    const int64_t* n_live;           // live parcel count on the device (null: n is exact)
    const TetGeo* tets;
    const int32_t* cell_tet0;
    const UT* Ecell;
    const UT* Epoint;
    const UT* Bcell;
    UT k, kv;          // q*(dt/2)/m, q*dt/m
    double dt;
    int do_push, rotate, hc_ref;
    uint32_t track_mask;             // bit d: direction d is empty, its velocity component is not tracked
    // non-coupled patches (PICParcel.C:153-184, :209-330; SimpleInjection.C:233-292): kind 0
    // deletes, 1 reflects by the wall model, 2 reflects specularly without the resync, 3
    // re-emits the parcel from the injector's triangles (null: no walls, every kind 0)
    const uint8_t* patch_kind;       // [n_patches] of this species
    const InjectTri* re_tri;         // kind 3: the injector patch's triangles
    int re_ntri, re_vmodel;
    double re_vt;
    // wall surface fields (PICParcel.C:240-339), per boundary face (face - n_internal) a row
    // of wall_stride doubles: rhoN, rhoM, linearKE, momentum[3], q, fD[3], then per species
    // rhoN, rhoM, momentum[3], linearKE, fD[3] (null: not kept)
    double* wall_acc;
    const int32_t* tet_face;
    const double* face_area;         // [n_faces - n_internal] |Sf|
    int n_internal, wall_stride;
    double weight, dt_step;
    const double* wall_T;            // [n_patches] boundaryT
    const double* wall_U;            // [n_patches * 3] boundaryU
    int wall_model;                  // 0 SpecularReflection, 1 MaxwellianThermal, 2 MixedDiffuseSpecular
    int wall_sync;                   // syncVelocityAtBoundary
    double diffuse_fraction;
    double mass;
    uint64_t seed;
    int64_t step;
    uint32_t species;
    // parcels that reached a wall in the move wait in this queue for the wall pass
    struct WallHit* wall_q;
    unsigned* wall_n;
    unsigned wall_cap;
    int* count;                      // per-cell parcel count of this species
    int2* place;                     // per parcel (new cell, slot in the cell) for the sort, or null
    unsigned long long* acc_cell;    // VolumeWeighting fixed-point accumulators
    unsigned long long* acc_point;
    double s42;                      // q*w/Q0 * 2^42
    uint32_t* flags;
    unsigned long long* stats;       // [0] deleted, [1] guard hits
};

template <class UT>
struct LocateArgs {
    PartSoA<UT> p;
    int64_t n, offset;
    const double* x;
    const int32_t* cell;
    const double* U;
    const uint32_t* id;
    const TetGeo* tets;
    const int32_t* cell_tet0;
    const double* C;
    unsigned* fails;
};

template <class UT>
struct InitArgs {
    PartSoA<UT> p;
    int64_t n, offset;
    int nppc;
    const TetGeo* tets;
    const int32_t* cell_tet0;
    double vth;
    double drift[3];
    uint32_t id0;
    uint64_t seed;
};

// push + track + count/deposit of every moving species in one launch (block0 filled by
// the launcher)
template <class UT>
struct MoveSet {
    MoveArgs<UT> a[kMaxSpecies];
    int64_t block0[kMaxSpecies + 1];
    int ns;
};
hipError_t launch_move(const MoveSet<double>& m, int pusher, int weighting, int dep, hipStream_t s);
hipError_t launch_move(const MoveSet<float>& m, int pusher, int weighting, int dep, hipStream_t s);
// the wall pass after a move with wall patches: reflection, resync and the rest of each
// queued parcel's track, then its count, place and deposit
hipError_t launch_wall(const MoveSet<double>& m, int pusher, int weighting, int dep, hipStream_t s);
hipError_t launch_wall(const MoveSet<float>& m, int pusher, int weighting, int dep, hipStream_t s);
// SimpleInjection (SimpleInjection.C:186-229) through its emitter (ParticleEmitter.C): one
// triangle of the injector patch's faces, fanned from each face's first point
struct InjectTri {
    double a[3], b[3], c[3];    // vertices; a is the face's first point
    double n[3];                // unit outward normal of the face
    double t1[3], t2[3];        // wall tangents: t1 from a to the face centre reversed, t2 = -n x t1
    double pc[3];               // owner cell centre
    int32_t cell, pad;
};
template <class UT>
struct InjectArgs {
    MoveArgs<UT> m;             // the species' push constants, fields, tets, seed, step and flags
    const InjectTri* tri;
    const double* cum;          // [ntri + 1] cumulative triangle areas from 0
    const int32_t* cell_tet0;
    int ntri;
    int vmodel;                 // 0 birdMaxwellianFlux, 1 mostProbableSpeed, 2 halfMaxwellianFlux, 3 maxwellianFlux
    double vt;                  // sqrt(k T / m) of the emitted distribution
    uint32_t model;             // injector index (a key of the draws)
    int sync;                   // synchronise the new velocity by -dt/2 (charged species with the boundary sync)
    int64_t count;              // parcels to add after the live ones
    int64_t cap;                // the species' capacity
};
hipError_t launch_inject(const InjectArgs<double>& a, int pusher, int weighting, hipStream_t s);
hipError_t launch_inject(const InjectArgs<float>& a, int pusher, int weighting, hipStream_t s);
hipError_t launch_push(const MoveArgs<double>& a, int pusher, int weighting, hipStream_t s);
hipError_t launch_push(const MoveArgs<float>& a, int pusher, int weighting, hipStream_t s);
hipError_t launch_deposit(const MoveArgs<double>& a, int dep, hipStream_t s);
hipError_t launch_deposit(const MoveArgs<float>& a, int dep, hipStream_t s);
hipError_t launch_locate(const LocateArgs<double>& a, hipStream_t s);
hipError_t launch_locate(const LocateArgs<float>& a, hipStream_t s);
hipError_t launch_positions(const PartSoA<double>& p, int64_t n, const TetGeo* tets, const double* C,
                            const double* pts, double* x, hipStream_t s);
hipError_t launch_positions(const PartSoA<float>& p, int64_t n, const TetGeo* tets, const double* C,
                            const double* pts, double* x, hipStream_t s);
// device scratch of the allocation-free entry points (noalloc.hip): noalloc_init takes the
// array's address on `device` (current) once; noalloc_scratch returns it, or null before
hipError_t noalloc_init(int device);
void* noalloc_scratch(int device);
size_t noalloc_scratch_bytes();
hipError_t launch_init(const InitArgs<double>& a, hipStream_t s);
hipError_t launch_init(const InitArgs<float>& a, hipStream_t s);
// counting-sort scatter of every species in one launch: n = launch bound (>= the live
// source count n_live); place = (cell, slot) per parcel from the counting pass, cell < 0
// for dropped parcels; block0 filled by the launcher
template <class UT>
struct ScatterSet {
    PartSoA<UT> src[kMaxSpecies], dst[kMaxSpecies];
    int64_t n[kMaxSpecies];
    const int64_t* n_live[kMaxSpecies];
    const int2* place[kMaxSpecies];
    const int* start[kMaxSpecies];
    // index[s] != 0: the cell order of species s as an index instead of moved parcels,
    // perm[start[cell] + slot] = i and id_cell[start[cell] + slot] = id[i] for every live
    // parcel i (place[i] = (cell, slot)); its src, dst and live count are left alone
    int index[kMaxSpecies];
    int* perm[kMaxSpecies];
    uint32_t* id_cell[kMaxSpecies];
    int64_t block0[kMaxSpecies + 1];
    int ns;
};
hipError_t launch_scatter(const ScatterSet<double>& m, hipStream_t s);
hipError_t launch_scatter(const ScatterSet<float>& m, hipStream_t s);

// ---------------------------------------------------------------- fields (k_field.hip)
// exclusive scans of n ints (the last count zero, so out[n-1] is the total) for every
// species in the same launches (one for up to 32768 cells, else three), totals stored as
// int64 into a second counter array: the scatter still reads the live counts before the
// sort, and the host makes the second array the live one after it; tmp sized by
// scan_cells_temp_bytes
constexpr int kScanBlock = 1024;
struct ScanSet {
    const int* in[kMaxSpecies];
    int* out[kMaxSpecies];
    int64_t* total[kMaxSpecies];
    int* tmp;
    int n;
};
size_t scan_cells_temp_bytes(int n);
// up to 12 ranges of 4-byte words zeroed by one launch
struct ZeroSet {
    void* p[16];
    int64_t words[16];
    int k;
};
hipError_t launch_zero(const ZeroSet& z, hipStream_t s);
hipError_t launch_scan_cells(const ScanSet& a, int n_species, hipStream_t s);

struct DepositFinishArgs {
    int n_cells;
    int n_species;
    const int* counts[kMaxSpecies];   // per species per cell
    long long whole[kMaxSpecies];     // llrint(q*w/Q0 * 2^42) per species (CellAverage)
    const unsigned long long* acc_cell;
    const unsigned long long* acc_point;
    const int32_t* cpo;               // cell -> cell points
    const int32_t* cp_master;
    const double* vw_w;
    const double* V;
    const double* space_charge;
    double q42;                       // Q0 / 2^42
    int weighting;
    double* rho;
};
hipError_t launch_deposit_finish(const DepositFinishArgs& a, hipStream_t s);

// E = -grad(phi) (Gauss linear) then point interpolation; Ecell_f / Epoint_f are the FP32
// copies used by the mixed-precision push (may be null).  The face table is column-major,
// entry e = k * n_cells + c for slot k < K of cell c (the cell's faces in ascending order):
// src >= 0 is the other cell (phi_f = w phi_c + val phi_src), -1 a fixed value (phi_f =
// val), -2 the cell's own value (zeroGradient), -3 an unused slot (only after the used ones).
// S and gk are three planes of K * n_cells each (x, y, z).
constexpr int32_t kGradFixed = -1, kGradSelf = -2, kGradUnused = -3;
struct GradArgs {
    int n_cells, K;
    const int32_t* src;
    const double* w;
    const double* val;
    const double* S;     // outward area vector
    const double* gk;    // |Sf| x nonOrthCorrectionVector out of the cell (null: orthogonal)
    const double* V;
    const double* phi;
    double* E;
    float* E_f;
};
hipError_t launch_grad(const GradArgs& a, hipStream_t s);
// corr[c] = sum over the cell's internal and coupled faces of |Sf| k . interpolated grad phi,
// with a.E holding -grad phi (k_grad's output)
hipError_t launch_nonorth_corr(const GradArgs& a, double* corr, hipStream_t s);

struct PointInterpArgs {
    int n_points;
    const int32_t* pco;
    const int32_t* pc_cell;
    const double* pc_w;
    const int32_t* pmaster;
    const double* Ecell;
    double* Epoint;
    float* Epoint_f;
};
hipError_t launch_point_interp(const PointInterpArgs& a, hipStream_t s);
hipError_t launch_to_float(const double* src, float* dst, int64_t n, hipStream_t s);

// per-species moment sums for diagnostics, deterministic two-stage reduction:
// out[0]=sum w, [1..3]=sum w U, [4]=sum w U^2, [5]=relativistic KE / (m c^2)
hipError_t launch_diag(const PartSoA<double>& p, int64_t n, double* partial, int nblocks, double* out,
                       hipStream_t s);
hipError_t launch_diag(const PartSoA<float>& p, int64_t n, double* partial, int nblocks, double* out,
                       hipStream_t s);
hipError_t launch_field_energy(const double* E, const double* V, int n, double* partial, int nblocks, double* out,
                               hipStream_t s);

// cell moments of one species (added into out[n_cells][kMomentVals]); cells sorted
constexpr int kMomentVals = 11;
template <class UT>
struct MomentArgs {
    int n_cells;
    PartSoA<UT> p;
    const int* start;
    const double* V;
    double mass, charge, weight;
    int maxp;                 // LDS slots: parcels of the species per cell (power of two)
    uint32_t* flags;
    double* out;
};
hipError_t launch_moments(const MomentArgs<double>& a, hipStream_t s);
hipError_t launch_moments(const MomentArgs<float>& a, hipStream_t s);
// picFields: per cell and species the time sums of rhoN, rhoM, linearKE and momentum
constexpr int kFieldAvgVals = 6;
hipError_t launch_fields_accumulate(const double* moments, double* acc, int n, hipStream_t s);

// ---------------------------------------------------------------- Coulomb (k_coulomb.hip)
struct CoulombArgs {
    int n_cells;
    int n_species;
    double mass[kMaxSpecies], charge[kMaxSpecies], weight[kMaxSpecies];
    double coulomb_log[kMaxSpecies][kMaxSpecies];
    int calc_debye, calculate_bmin, allow_intra;
    uint32_t debye_mask;              // species summed into the Debye length (charged ones only)
    int debye_drift;                  // 1: the species temperature has the mean drift removed
    const int* start[kMaxSpecies];    // cell_start per species [n_cells+1]
    const int* perm[kMaxSpecies];     // sorted slot -> array index, or null (arrays sorted by cell)
    void* ux[kMaxSpecies];
    void* uy[kMaxSpecies];
    void* uz[kMaxSpecies];
    const uint32_t* id[kMaxSpecies];  // parcel ids in cell order (slot start[cell] + k)
    const double* V;
    double dt;
    uint64_t seed;
    int64_t step;
    int maxp;                         // LDS index-list slots of the per-cell pass (power of two)
    int maxp_big;                     // slots of the large-cell pass (power of two)
    int group;                        // lanes per cell of the grouped pass: 16, 32 or 64 (64 / group
                                      // cells per wave); 0 = the one-cell-per-wave pass with the
                                      // velocities in LDS
    int* big_list;                    // cells left to the large-cell pass, [n_cells]
    int* big_n;                       // their number (zero before the launch)
    int model;                        // 0 Nanbu, 1 Perez
    int pairing;                      // 0 TakizukaAbe, 1 NanbuYonemura
    int weight_correction;            // Nanbu weight correction active (unequal weights)
    uint32_t* flags;
    unsigned long long* ncoll;        // kNcollSlots counters kNcollStride apart; the total is their sum
    const struct PairConst* pair;     // [P * kMaxSpecies + Q]
};
// Coulomb constants of an ordered species pair (P, Q), reduced mass mu:
// cst0 = (qP qQ / eps0 mu)^2 / (4 pi), fP = mQ / (mP + mQ), fQ = mP / (mP + mQ),
// b0c = qP qQ / (2 pi eps0 mu), hbc = h / (2 mu)
struct PairConst {
    double cst0, fP, fQ, b0c, hbc, pad[3];
};
void pair_constants(const double* mass, const double* charge, int n_species, PairConst* out);
constexpr int kNcollSlots = 64, kNcollStride = 16;
// step counters on the device: 16 words, then the Coulomb collision counters, then the MCC
// counters (spread over slots for the same reason)
// statistics words: 0 deleted, 1 tracking guard, 2 wall hits, then the collision counters
constexpr int kStatsWallHits = 2;
constexpr int kStatsNcoll = 16, kStatsMcc = kStatsNcoll + kNcollSlots * kNcollStride,
              kStatsWords = kStatsMcc + kNcollSlots * kNcollStride;
size_t coulomb_lds_bytes(int maxp, size_t cache_elem);
hipError_t launch_coulomb(const CoulombArgs& a, bool fp32_velocity, hipStream_t s);

// ---------------------------------------------------------------- MCC (k_mcc.hip)
constexpr int kXsGuide = 1024;
void xs_guide(const double* e, int n, int* guide, double* le0, double* idl);

template <class UT>
struct MccArgs {
    int n_cells;
    PartSoA<UT> e;            // electrons, in cell order directly or through perm
    const int* start;         // electron cell_start
    const int* perm;          // sorted slot -> electron index, or null (electrons sorted by cell)
    const uint32_t* id_cell;  // electron ids in cell order (slot start[cell] + k)
    int64_t* n_e;             // device count of electrons (appends)
    int64_t cap_e;
    PartSoA<UT> ion;          // created ions (may have null pointers when no ion species)
    int64_t* n_i;
    int64_t cap_i;
    double* smax;             // sigma*cR max per cell
    double* rem;              // candidate remainder per cell
    const double* xs_e[3];    // energies per process
    const double* xs_s[3];    // values
    int xs_n[3];
    const int* xs_guide[3];   // kXsGuide bins equal in log(e) -> last entry at or below the bin, or null
    double xs_le0[3], xs_idl[3];  // log(e[0]) and kXsGuide / (log(e[n-1]) - log(e[0]))
    double th_ex, th_iz;
    double massE, massN, nN, vthN, drift[3], wEwN, wNwE;
    int create_ion, create_electron;
    int relativistic;         // RelativisticElectronNeutral scattering instead of Nanbu
    double dt;
    uint64_t seed;
    int64_t step;
    int maxp;                     // LDS slots (electrons of a cell) of the per-cell pass
    int group;                    // lanes per cell of the several-cells-per-wave pass, 0 = one cell per wave
    int maxp_group;               // LDS slots per cell of that pass
    int maxp_big;                 // slots of the large-cell pass
    int* big_list;                // cells left to the large-cell pass, [n_cells]
    int* big_n;                   // their number (zero before the launch)
    uint32_t* flags;
    unsigned long long* counts;   // kNcollSlots slots kNcollStride apart, each: candidates,
                                  // elastic, excitation, ionization; the totals are the sums
    // deposit of created parcels
    const TetGeo* tets;
    int* count_e;                 // per-cell counts (CellAverage / sort histogram)
    int* count_i;
    unsigned long long* acc_cell;
    unsigned long long* acc_point;
    double s42_e, s42_i;
    int dep_vw;
    // the collision model: electron-neutral (draws tagged kTagMcc, counters from word 0 of
    // each slot) or ion-neutral against the background (ion_neutral 1: a.e, start, perm and
    // id_cell are the ion species', xs_*[0] elastic and [1] charge exchange with th_ex its
    // threshold, massE the ion mass; draws tagged kTagIonMcc, counters from word 4:
    // candidates, elastic, charge exchange)
    int ion_neutral;
};
// scattering models of the MCC kernels
constexpr int kMccNanbu = 0, kMccRelativistic = 1, kMccIon = 2;
hipError_t launch_mcc(const MccArgs<double>& a, hipStream_t s);
hipError_t launch_mcc(const MccArgs<float>& a, hipStream_t s);

// ---------------------------------------------------------------- parcel pairs (k_dsmc.hip)
// Collisions between parcels of one cell drawn as candidate pairs: the neutral DSMC of
// BinaryCollisionModel (pairs within the neutral species, then the neutrals with the
// background gas) and the ions with the neutral parcels of IonNeutralCollisionModel.
// A cell's parcels are listed species after species (group A first, then group B), each
// species in the order of a hash of its ids, so that a candidate index names the same
// parcel whatever order the sort left them in.
constexpr int kPairIon = 0, kPairDsmc = 1;
constexpr int kPairSlots = 256, kPairSlotsBig = 1024;   // parcels of a cell per pass
// per ordered species pair (a, b) of the binary model: pi d_ab^2 (hs) with d_ab the mean
// diameter, and for the VariableHardSphere model 2 k Tref (vn), the reduced mass (mr),
// omega_ab - 1/2 (ve) and exp(lgamma(5/2 - omega_ab)) (gl); row kMaxSpecies holds
// (a, background)
struct DsmcPair {
    double hs, vn, mr, ve, gl;
};
template <class UT>
struct PairArgs {
    int n_cells;
    int ns, na;                   // species in the cell list; the first na form group A
    PartSoA<UT> p[kMaxSpecies];   // in list order
    const int* start[kMaxSpecies];
    const int* perm[kMaxSpecies];           // sorted slot -> array index, or null
    const uint32_t* id_cell[kMaxSpecies];   // ids in cell order
    double w[kMaxSpecies], mass[kMaxSpecies];
    const double* vol;            // cell volumes
    double* smax;                 // (sigma cR)_max per cell of the pair pass (and the background pass)
    double* rem;                  // pair candidate remainder
    double* rem_bg;               // background candidate remainder (binary model)
    // binary model
    const DsmcPair* tab;          // [(kMaxSpecies + 1) * kMaxSpecies]: [a * kMaxSpecies + b], background row
    int vhs;
    const TetGeo* tets;           // sub-cells: parcel positions against the cell centre
    const double* C;
    const double* pts;
    uint32_t empty_mask;          // bit d: direction d is empty (its position component is not tested)
    int bg;                       // collisions with the uniformBackground gas
    double nN, vthN, drift[3], massN;
    // ion-neutral: elastic [0] and charge exchange [1] cross sections
    const double* xs_e[2];
    const double* xs_s[2];
    int xs_n[2];
    const int* xs_guide[2];
    double xs_le0[2], xs_idl[2];
    double th_cx;
    int keep_neutral;             // updateNeutralVelocity false
    int wc;                       // Nanbu weight correction
    double dt;
    uint64_t seed;
    int64_t step;
    int* big_list;                // cells left to the large-cell pass, [n_cells]
    int* big_n;
    uint32_t* flags;
    unsigned long long* counts;   // the MCC counter slots; words kPairWordIon.. / kPairWordDsmc..
};
// counter words of an MCC slot: ion pairs (candidates, elastic, charge exchange), binary
// pairs (candidates, collisions), binary background (candidates, collisions)
constexpr int kPairWordIon = 7, kPairWordDsmc = 10;
hipError_t launch_pairs(const PairArgs<double>& a, int mode, hipStream_t s);
hipError_t launch_pairs(const PairArgs<float>& a, int mode, hipStream_t s);

// ---------------------------------------------------------------- Poisson (k_poisson.hip)
struct PoissonDev {
    int n, K, ncol;
    const int32_t* nb;
    const double* a;
    const float* a_f;
    const double* diag;
    const float* diag_f;
    const double* rD;      // DIC diagonal
    const float* rD_f;
    const int32_t* colour;
    const int32_t* corder;
    const int32_t* col_start_h;  // host copy of colour starts
    const double* bsrc;
    const double* V;
    // the same system renumbered by colour (cell new -> corder[new]), coefficients stored
    // column-major [k * n + i]: colour c is the contiguous range col_start[c]..col_start[c+1]
    const int32_t* nbT;
    const double* aT;
    const float* aT_f;
    const double* diag_p;
    const float* diag_pf;
    const double* rD_p;
    const float* rD_pf;
    const int32_t* col_start_d;  // device copy of colour starts
};
// b = rho V / eps0 + Dirichlet sources (+ corr when not null), zero mean when singular
hipError_t poisson_rhs(const PoissonDev& P, const double* rho, const double* corr, double* b, bool singular,
                       double* scratch, hipStream_t s);
// PCG with DIC in FP64 (mode 0) or FP32-inner CG inside FP64 iterative refinement (mode 1)
struct PoissonWork {
    double *r, *z, *p, *q, *x, *b;
    float *rf, *zf, *pf, *qf, *xf;
    double *p2, *xp, *bp;  // second search-direction buffer, x and b in colour order
    float* pf2;
    double* red;          // reduction scratch: (ncol + 3) * 1024 + 64 doubles
    double* scal;         // device scalars
    unsigned* cnt;        // block-completion counters of the fused reductions (zero between uses)
    int nblocks;
};
// phi <- 2 phi - phi_old, phi_old <- the previous phi (linear extrapolation in time)
hipError_t poisson_extrapolate(double* phi, double* phi_old, int n, hipStream_t s);
// r = b - A x (FP64)
hipError_t poisson_residual(const PoissonDev& P, PoissonWork& W, const double* x, const double* b, double* r,
                            hipStream_t s);
int poisson_solve(const PoissonDev& P, PoissonWork& W, double* phi, const double* b, double tol, int max_iter,
                  int mode, bool singular, int* iters, int* refinements, double* rel, hipStream_t s);

}  // namespace picmi
