// C-ABI of libpicmi: the context owns the mesh tables, fields and particle arrays on one
// HIP device and runs the stages of PICCloud::evolve() (PICCloud.C:1426-1471) as batched
// kernels on its stream.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/picmi.h"
#include "internal.h"
#include "gnn.h"
#include "kernels.h"

using namespace picmi;

namespace {

thread_local std::string g_last_error;

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
};

struct Species {
    double mass = 0, charge = 0, weight = 0;
    double diameter = 0, omega = 0;          // moleculeProperties diameter and omega
    int solve_movement = 1;
    int64_t capacity = 0;
    int64_t n = 0;  // host mirror of the parcel count
    int64_t n_sync = 0;  // the count at the last synchronisation (n can be a launch bound)
    // double-buffered SoA
    void* y[2][3] = {{nullptr, nullptr, nullptr}, {nullptr, nullptr, nullptr}};
    void* u[2][3] = {{nullptr, nullptr, nullptr}, {nullptr, nullptr, nullptr}};
    int32_t* tet[2] = {nullptr, nullptr};
    uint32_t* id[2] = {nullptr, nullptr};
    int cur = 0;
    int* count = nullptr;   // [n_cells+1]
    int* start = nullptr;   // [n_cells+1]
    int2* place = nullptr;  // [capacity] (cell, slot) from the counting pass
    int* perm = nullptr;    // [capacity] sorted slot -> array index after an index sort
    uint32_t* id_cell = nullptr;  // [capacity] ids in cell order after an index sort
    bool indexed = false;   // the cell order is perm (and id_cell), not the arrays' order
    int sort_interval = 1;  // moving sorts every k steps (index sorts in between)
    int ncur = 0;           // which of the context's two count rows holds the live count
};

struct StageEvent {
    int stage;
    hipEvent_t a, b;
};

}  // namespace

struct pic_ctx {
    std::string err;
    int device = 0;
    hipStream_t stream = nullptr;
    std::vector<void*> allocs;
    HostMesh hm;
    HostPoisson hp;
    bool has_poisson = false;
    // models
    int pusher = kPushBoris, fw = kCellAverage, cd = kCellAverage;
    bool electrostatic = false, coulomb = false, mcc = false, bg = false;
    bool ion_mcc = false;                    // IonNeutralCollisionModel (neutral parcels, background)
    bool binary = false;                     // BinaryCollisionModel IsotropicScattering
    int vhs = 0;                             // TotalCrossSectionModel VariableHardSphere
    double vhs_tref = 0.0;
    int binary_wc = 0, ion_wc = 0;           // Nanbu weight correction of the binary / ion-neutral models
    int ion_keep_neutral = 0;                // IonNeutralCollisionCoeffs.updateNeutralVelocity false
    int bg_species = -1;                     // the background gas' species
    int coulomb_model = 0, pairing = 0, weight_corr = 0;
    int en_relativistic = 0;
    int precision = PIC_PREC_FP64;
    int hc_ref = 0;
    int allow_intra = 1, calc_bmin = 1;
    uint32_t debye_mask = 0xFFu;             // species in the Debye length (the charged ones of them)
    int debye_drift = 1;                     // DebyeLength.accountForDrift
    int wall_model = 0;                      // WallReflectionModel: Specular, MaxwellianThermal, Mixed
    double diffuse_fraction = 0.0;
    int wall_sync = 1;                       // syncVelocityAtBoundary
    bool walls = false;                      // the mesh has wall patches
    uint8_t* d_patch_kind = nullptr;         // per species and patch: what a parcel reaching it does
    double* d_wall_acc = nullptr;            // wall surface fields of the last step (per boundary face)
    int32_t* d_tet_face = nullptr;
    double* d_face_area = nullptr;           // |Sf| per boundary face
    int wall_stride = 0;
    double *d_wall_T = nullptr, *d_wall_U = nullptr;
    WallHit* d_wall_q = nullptr;             // parcels on a wall after the move (all moving species)
    unsigned* d_wall_n = nullptr;
    unsigned wall_cap = 0;
    double coulomb_log[kMaxSpecies][kMaxSpecies] = {};
    int electron = -1, ion = -1;
    double bg_n = 0, bg_T = 0, bg_m = 0, bg_w = 0, bg_v[3] = {0, 0, 0};
    int create_ion = 1, create_electron = 1;
    // SimpleInjection models (BoundaryModelList of PatchBoundaryModels)
    struct Injector {
        pic_injection cfg{};
        double rem[8] = {};            // particleRemainder_ per injected species
        InjectTri* d_tri = nullptr;
        double* d_cum = nullptr;
        int ntri = 0;
    };
    std::vector<Injector> inj;
    // per species (rows of kMaxSpecies) and patch: what a parcel reaching the patch does
    // (0 delete, 1 wall model, 2 specular without resync, 3 re-emission), and per species
    // the injector that re-emits it
    std::vector<uint8_t> h_kind;
    struct Reemit {
        const InjectTri* tri = nullptr;
        int ntri = 0, vmodel = 0;
        double vt = 0.0;
    } re[kMaxSpecies];
    double inj_time = 0.0;             // time at the start of the next step
    int64_t n_injected = 0;            // parcels injected since the statistics were reset
    uint64_t seed = 0;
    double ptol = 1e-8;
    int pmaxit = 10000, pmode = 0;
    int sort_interval = 1;
    int64_t step = 0;
    bool rotate = false;
    bool e_zero = true;                      // E is zero in every cell (never set, no field solve)
    bool counts_valid = false;
    bool place_valid = false;                // Species::place matches the current layout
    bool big_n_zero = false;                 // d_big_n zeroed since its last use
    // species
    std::vector<Species> sp;
    double Q0 = 1.0;
    // device mesh
    TetGeo* d_tets = nullptr;
    int32_t *d_cell_tet0 = nullptr, *d_cpo = nullptr, *d_cp_master = nullptr, *d_pco = nullptr, *d_pc_cell = nullptr,
            *d_pmaster = nullptr;
    double *d_C = nullptr, *d_V = nullptr, *d_points = nullptr, *d_vw_w = nullptr, *d_pc_w = nullptr;
    // fields
    double *d_E = nullptr, *d_B = nullptr, *d_phi = nullptr, *d_rho = nullptr, *d_space = nullptr,
           *d_Ep = nullptr;
    float *d_Ef = nullptr, *d_Epf = nullptr, *d_Bf = nullptr;
    bool has_space = false;
    unsigned long long *d_acc_cell = nullptr, *d_acc_point = nullptr;
    uint32_t* d_flags = nullptr;
    // 0 deleted, 1 guard; Coulomb collisions in kNcollSlots counters from kStatsNcoll, MCC
    // counts in kNcollSlots groups from kStatsMcc
    unsigned long long* d_stats = nullptr;
    // device parcel counts, two rows of kMaxSpecies: row Species::ncur holds species s's live
    // count (authoritative), the other row receives the sorted count from the scan and
    // becomes the live one once the parcels have moved into cell order
    int64_t* d_nrows = nullptr;
    int* d_big_list = nullptr;               // cells left to the large-cell collision passes
    double* d_moments = nullptr;             // cell moments scratch [n_cells][11]
    double* d_favg = nullptr;                // picFields time sums [species][n_cells][kFieldAvgVals]
    int64_t n_favg = 0;                      // samples in them
    int* d_big_n = nullptr;
    PairConst* d_pair = nullptr;             // Coulomb constants per ordered species pair
    void* d_scan_tmp = nullptr;
    size_t scan_bytes = 0;
    double* d_diag_part = nullptr;
    double* d_diag_out = nullptr;
    // MCC: (sigma cR)max and candidate remainders per cell, electrons and ions; cross-section
    // tables by PIC_XS_* process (electron elastic, excitation, ionization; ion elastic,
    // charge exchange)
    double *d_smax = nullptr, *d_rem = nullptr;
    double *d_smax_i = nullptr, *d_rem_i = nullptr;
    // ion-neutral parcel pairs; binary model pairs and background (one maximum for both)
    double *d_smax_ip = nullptr, *d_rem_ip = nullptr;
    double *d_smax_n = nullptr, *d_rem_n = nullptr, *d_rem_nb = nullptr;
    DsmcPair* d_dsmc_tab = nullptr;
    static constexpr int kXsProcs = 5;
    double* d_xs_e[kXsProcs] = {};
    double* d_xs_s[kXsProcs] = {};
    int xs_n[kXsProcs] = {};
    int* d_xs_guide[kXsProcs] = {};
    double xs_le0[kXsProcs] = {}, xs_idl[kXsProcs] = {};
    double xs_th[kXsProcs] = {};
    std::vector<double> xs_he[kXsProcs], xs_hs[kXsProcs];
    // Poisson
    PoissonDev pd{};
    PoissonWork pw{};
    double* d_b = nullptr;
    // gradient tables
    // Gauss-gradient face table (column-major, see GradArgs) with the non-orthogonal
    // correction vectors (null on an orthogonal mesh); -grad of the potential a solve starts
    // from and the correction per cell (non-orthogonal meshes only)
    GradArgs gtab{};
    double *d_mgrad = nullptr, *d_corr = nullptr;
    // learned initial guess
    std::unique_ptr<Gnn> gnn;
    bool gnn_use = false, gnn_rec = false;
    double* d_phi_prev = nullptr;
    double* d_phi_old = nullptr;             // phi of the solve before the last (extrapolation)
    int poisson_guess = 0;
    int64_t n_solves = 0;
    // profiling
    uint32_t profile = 0;                    // bit s: stage s is timed with HIP events
    std::vector<StageEvent> pending;
    std::vector<hipEvent_t> pool;
    double stage_ms[PIC_STAGE_COUNT] = {};
    int64_t stage_n[PIC_STAGE_COUNT] = {};

    int fail(int code, const std::string& m) {
        err = m;
        g_last_error = m;
        return code;
    }
    template <class T>
    T* alloc(size_t count) {
        if (count == 0) count = 1;
        void* p = nullptr;
        if (hipMalloc(&p, count * sizeof(T)) != hipSuccess) return nullptr;
        hipMemsetAsync(p, 0, count * sizeof(T), stream);
        allocs.push_back(p);
        return (T*)p;
    }
    void release(void* p) {
        if (!p) return;
        for (auto& q : allocs)
            if (q == p) { q = allocs.back(); allocs.pop_back(); break; }
        hipFree(p);
    }
    template <class T>
    T* upload(const std::vector<T>& v) {
        T* p = alloc<T>(v.size());
        if (p && !v.empty()) hipMemcpyAsync(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, stream);
        return p;
    }
    hipEvent_t ev() {
        if (!pool.empty()) { hipEvent_t e = pool.back(); pool.pop_back(); return e; }
        hipEvent_t e;
        hipEventCreate(&e);
        return e;
    }
    struct Scope {
        pic_ctx* c;
        int stage;
        hipEvent_t a = nullptr;
        Scope(pic_ctx* c_, int s) : c(c_), stage(s) {
            if ((c->profile >> s) & 1u) { a = c->ev(); hipEventRecord(a, c->stream); }
        }
        ~Scope() {
            if (a) {
                hipEvent_t b = c->ev();
                hipEventRecord(b, c->stream);
                c->pending.push_back({stage, a, b});
            }
        }
    };
    bool mixed() const { return precision == PIC_PREC_MIXED; }
};

namespace {

int check_hip(pic_ctx* c, hipError_t e, const char* what) {
    if (e == hipSuccess) return PIC_OK;
    return c->fail(PIC_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

template <class UT>
PartSoA<UT> soa(const Species& s, int buf) {
    PartSoA<UT> p;
    p.y1 = (double*)s.y[buf][0]; p.y2 = (double*)s.y[buf][1]; p.y3 = (double*)s.y[buf][2];
    p.ux = (UT*)s.u[buf][0]; p.uy = (UT*)s.u[buf][1]; p.uz = (UT*)s.u[buf][2];
    p.tet = s.tet[buf];
    p.id = s.id[buf];
    return p;
}

int parse_model(pic_ctx* c, const char* name, const char* keyword, const std::vector<std::string>& valid,
                int* out) {
    std::string n = name ? name : "";
    for (size_t i = 0; i < valid.size(); ++i)
        if (valid[i] == n) { *out = (int)i; return PIC_OK; }
    std::string m = "Unknown " + std::string(keyword) + " type " + n + "\n\nValid " + keyword + " types are:\n" +
                    std::to_string(valid.size()) + "\n(\n";
    for (auto& v : valid) m += v + "\n";
    m += ")";
    return c->fail(PIC_ERR_UNKNOWN_MODEL, m);
}

// the live parcel count of species s on the device, and the row the next sort's scan fills
int64_t* live_n(pic_ctx* c, int s) { return c->d_nrows + c->sp[s].ncur * kMaxSpecies + s; }
int64_t* next_n(pic_ctx* c, int s) { return c->d_nrows + (1 - c->sp[s].ncur) * kMaxSpecies + s; }

void set_dev_n(pic_ctx* c, int s);

// error flags + counts after a synchronisation point
int sync_state(pic_ctx* c) {
    uint32_t flags = 0;
    std::vector<int64_t> n(2 * kMaxSpecies);
    hipMemcpyAsync(&flags, c->d_flags, sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream);
    hipMemcpyAsync(n.data(), c->d_nrows, sizeof(int64_t) * 2 * kMaxSpecies, hipMemcpyDeviceToHost, c->stream);
    hipError_t e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return check_hip(c, e, "stream synchronize");
    for (size_t s = 0; s < c->sp.size(); ++s) {
        const int64_t dev = n[c->sp[s].ncur * kMaxSpecies + s];
        c->sp[s].n = std::min<int64_t>(dev, c->sp[s].capacity);
        c->sp[s].n_sync = c->sp[s].n;
        // creations past the capacity (flagged below) left the device count above it: the
        // live row is clamped too, so later steps see the parcels that exist
        if (dev > c->sp[s].capacity) set_dev_n(c, (int)s);
    }
    if (flags) {
        hipMemsetAsync(c->d_flags, 0, sizeof(uint32_t), c->stream);
        if (flags & 1u) return c->fail(PIC_ERR_SUPERLUMINAL, "Particle velocity greater than speed of light (299792458)! Aborting...");
        if (flags & 2u) return c->fail(PIC_ERR_CELL_OVERFLOW, "a cell holds more parcels than the collision kernel capacity");
        if (flags & 4u) return c->fail(PIC_ERR_CAPACITY, "species capacity exceeded by created parcels");
        if (flags & 8u) return c->fail(PIC_ERR_MESH, "tracking reached a tet face without a neighbour");
    }
    return PIC_OK;
}

void set_dev_n(pic_ctx* c, int s) {
    hipMemcpyAsync(live_n(c, s), &c->sp[s].n, sizeof(int64_t), hipMemcpyHostToDevice, c->stream);
    c->sp[s].n_sync = c->sp[s].n;
}

int alloc_species(pic_ctx* c, Species& s) {
    const size_t cap = (size_t)s.capacity;
    const size_t ub = c->mixed() ? sizeof(float) : sizeof(double);
    for (int b = 0; b < 2; ++b) {
        for (int k = 0; k < 3; ++k) {
            s.y[b][k] = c->alloc<double>(cap);
            if (!s.y[b][k]) return c->fail(PIC_ERR_HIP, "out of device memory (particles)");
            void* p = nullptr;
            if (hipMalloc(&p, std::max<size_t>(cap, 1) * ub) != hipSuccess) return c->fail(PIC_ERR_HIP, "out of device memory (particles)");
            c->allocs.push_back(p);
            s.u[b][k] = p;
        }
        s.tet[b] = c->alloc<int32_t>(cap);
        s.id[b] = c->alloc<uint32_t>(cap);
        if (!s.tet[b] || !s.id[b]) return c->fail(PIC_ERR_HIP, "out of device memory (particles)");
    }
    const int nc = c->hm.n_cells;
    s.count = c->alloc<int>(nc + 1);
    s.start = c->alloc<int>(nc + 1);
    s.place = c->alloc<int2>(cap);
    s.perm = c->alloc<int>(cap);
    s.id_cell = c->alloc<uint32_t>(cap);
    if (!s.count || !s.start || !s.place || !s.perm || !s.id_cell)
        return c->fail(PIC_ERR_HIP, "out of device memory (cell counts)");
    return PIC_OK;
}

double fixed_scale(const pic_ctx* c, int s) { return c->sp[s].charge * c->sp[s].weight / c->Q0 * kFixedScale; }

// zero counts and accumulators before a deposit pass, and the large-cell list count of the
// Coulomb pass that follows, in one launch
void zero_deposit(pic_ctx* c) {
    ZeroSet z{};
    for (auto& s : c->sp) { z.p[z.k] = s.count; z.words[z.k++] = c->hm.n_cells + 1; }
    if (c->cd == kVolumeWeighting) {
        z.p[z.k] = c->d_acc_cell; z.words[z.k++] = 2 * (int64_t)c->hm.n_cells;
        z.p[z.k] = c->d_acc_point; z.words[z.k++] = 2 * (int64_t)c->hm.n_points;
    }
    z.p[z.k] = c->d_big_n; z.words[z.k++] = 1;
    if (c->d_wall_n) { z.p[z.k] = c->d_wall_n; z.words[z.k++] = 1; }
    if (c->d_wall_acc) {
        z.p[z.k] = c->d_wall_acc;
        z.words[z.k++] = 2 * (int64_t)std::max(c->hm.n_faces - c->hm.n_internal, 1) * c->wall_stride;
    }
    launch_zero(z, c->stream);
    c->big_n_zero = true;
}

template <class UT>
MoveArgs<UT> move_args(pic_ctx* c, int s, double dt) {
    Species& S = c->sp[s];
    MoveArgs<UT> a{};
    a.p = soa<UT>(S, S.cur);
    a.n = S.n;
    a.n_live = live_n(c, s);
    a.tets = c->d_tets;
    if constexpr (sizeof(UT) == 4) {
        a.Ecell = (const UT*)c->d_Ef; a.Epoint = (const UT*)c->d_Epf; a.Bcell = (const UT*)c->d_Bf;
    } else {
        a.Ecell = (const UT*)c->d_E; a.Epoint = (const UT*)c->d_Ep; a.Bcell = (const UT*)c->d_B;
    }
    a.k = (UT)(S.charge * (0.5 * dt) / S.mass);
    a.kv = (UT)(S.charge * dt / S.mass);
    a.dt = dt;
    // with E and B zero everywhere the Boris push (rotation skipped, BorisPusher.C:42-55) is
    // the identity: the kernel then leaves U untouched instead of rewriting it
    const bool identity_push = c->e_zero && !c->rotate &&
                               (c->pusher == kPushBoris || c->pusher == kPushBorisNR || c->pusher == kPushNone);
    a.do_push = (S.charge != 0.0 && !identity_push) ? 1 : 0;
    a.rotate = c->rotate ? 1 : 0;
    a.hc_ref = c->hc_ref;
    a.track_mask = c->hm.track_mask;
    a.cell_tet0 = c->d_cell_tet0;
    if (c->walls) {
        a.patch_kind = c->d_patch_kind + (size_t)s * c->hm.n_patches;
        a.re_tri = c->re[s].tri;
        a.re_ntri = c->re[s].ntri;
        a.re_vmodel = c->re[s].vmodel;
        a.re_vt = c->re[s].vt;
        a.wall_acc = c->d_wall_acc;
        a.tet_face = c->d_tet_face;
        a.face_area = c->d_face_area;
        a.n_internal = c->hm.n_internal;
        a.wall_stride = c->wall_stride;
        a.weight = S.weight;
        a.dt_step = dt;
        a.wall_T = c->d_wall_T;
        a.wall_U = c->d_wall_U;
        a.wall_model = c->wall_model;
        a.wall_sync = c->wall_sync;
        a.diffuse_fraction = c->diffuse_fraction;
        a.wall_q = c->d_wall_q;
        a.wall_n = c->d_wall_n;
        a.wall_cap = c->wall_cap;
    }
    a.mass = S.mass;
    a.seed = c->seed;
    a.step = c->step;
    a.species = (uint32_t)s;
    a.count = S.count;
    a.place = S.place;
    a.acc_cell = c->d_acc_cell;
    a.acc_point = c->d_acc_point;
    a.s42 = fixed_scale(c, s);
    a.flags = c->d_flags;
    a.stats = c->d_stats;
    return a;
}

template <class UT>
int do_move_t(pic_ctx* c, double dt) {
    pic_ctx::Scope sc(c, PIC_STAGE_MOVE);  // one timed launch for all moving species
    MoveSet<UT> m{};
    for (int s = 0; s < (int)c->sp.size(); ++s) {
        Species& S = c->sp[s];
        if (S.n == 0) continue;
        if (!S.solve_movement) {
            // stationary species are still counted / deposited (PICParcel.C:51)
            if (int r = check_hip(c, launch_deposit(move_args<UT>(c, s, dt), c->cd, c->stream), "deposit")) return r;
        } else {
            m.a[m.ns++] = move_args<UT>(c, s, dt);
        }
    }
    if (m.ns == 0) return PIC_OK;
    if (int r = check_hip(c, launch_move(m, c->pusher, c->fw, c->cd, c->stream), "move kernel")) return r;
    if (c->walls) return check_hip(c, launch_wall(m, c->pusher, c->fw, c->cd, c->stream), "wall pass");
    return PIC_OK;
}

int do_move(pic_ctx* c, double dt) {
    zero_deposit(c);
    if (int r = c->mixed() ? do_move_t<float>(c, dt) : do_move_t<double>(c, dt)) return r;
    c->counts_valid = true;
    c->place_valid = true;
    for (auto& S : c->sp) S.indexed = false;
    return PIC_OK;
}

template <class UT>
ScatterSet<UT> scatter_set(pic_ctx* c) {
    ScatterSet<UT> m{};
    m.ns = (int)c->sp.size();
    for (int s = 0; s < m.ns; ++s) {
        Species& S = c->sp[s];
        m.src[s] = soa<UT>(S, S.cur);
        m.dst[s] = soa<UT>(S, 1 - S.cur);
        m.n[s] = S.n;
        m.n_live[s] = live_n(c, s);
        m.place[s] = S.place;
        m.start[s] = S.start;
        m.perm[s] = S.perm;
        m.id_cell[s] = S.id_cell;
    }
    return m;
}

// buildCellOccupancy (PICCloud.C:327-354): counting sort of every species by cell.  No
// host round trip: the live count after the sort (start[n_cells]) stays on the device and
// the host keeps S.n as an upper bound for launch sizes until the next synchronisation.
// in_step (the sort inside pic_step / pic_run): a species whose moving sort is not due this
// step (Species::sort_interval) keeps its parcels where they are and gets the cell order as
// an index, Species::perm and id_cell (the reference's occupancy lists hold parcel
// pointers, PICCloud.H:166); its deleted parcels stay in the arrays, marked by tet -1,
// until its next moving sort.
int do_sort(pic_ctx* c, bool in_step = false) {
    pic_ctx::Scope sc(c, PIC_STAGE_SORT);
    const int nc = c->hm.n_cells;
    const int ns = (int)c->sp.size();
    if (!c->counts_valid || !c->place_valid) {
        for (auto& s : c->sp) hipMemsetAsync(s.count, 0, sizeof(int) * (nc + 1), c->stream);
        for (int s = 0; s < ns; ++s) {
            hipError_t e = c->mixed() ? launch_deposit(move_args<float>(c, s, 0.0), kCellAverage, c->stream)
                                      : launch_deposit(move_args<double>(c, s, 0.0), kCellAverage, c->stream);
            if (int r = check_hip(c, e, "count kernel")) return r;
        }
        c->counts_valid = true;
        c->place_valid = true;
    }
    // every species' scan in the same launches; the sorted counts go to the species' other
    // count row while the scatter reads the live counts before the sort
    ScanSet ss{};
    ss.n = nc + 1;
    ss.tmp = (int*)c->d_scan_tmp;
    for (int s = 0; s < ns; ++s) {
        ss.in[s] = c->sp[s].count; ss.out[s] = c->sp[s].start; ss.total[s] = next_n(c, s);
    }
    if (int r = check_hip(c, launch_scan_cells(ss, ns, c->stream), "scan")) return r;
    int index[kMaxSpecies] = {};
    bool moved = false;
    for (int s = 0; s < ns; ++s) {
        index[s] = in_step && (c->step + 1) % c->sp[s].sort_interval != 0;
        moved = moved || !index[s];
    }
    hipError_t e;
    if (c->mixed()) {
        ScatterSet<float> m = scatter_set<float>(c);
        for (int s = 0; s < ns; ++s) m.index[s] = index[s];
        e = launch_scatter(m, c->stream);
    } else {
        ScatterSet<double> m = scatter_set<double>(c);
        for (int s = 0; s < ns; ++s) m.index[s] = index[s];
        e = launch_scatter(m, c->stream);
    }
    if (int r = check_hip(c, e, "scatter")) return r;
    for (int s = 0; s < ns; ++s) {
        Species& S = c->sp[s];
        S.indexed = index[s] != 0;
        if (S.indexed) continue;
        S.cur = 1 - S.cur;
        S.ncur = 1 - S.ncur;   // the sorted count is the live one from here on
    }
    if (moved) c->place_valid = false;   // the moved species' slots refer to the layout before this sort
    return PIC_OK;
}

int do_coulomb(pic_ctx* c, double dt) {
    pic_ctx::Scope sc(c, PIC_STAGE_COULOMB);
    const int ns = (int)c->sp.size();
    CoulombArgs a{};
    a.n_cells = c->hm.n_cells;
    a.n_species = ns;
    bool calc_debye = false;
    for (int i = 0; i < ns; ++i) {
        a.mass[i] = c->sp[i].mass;
        a.charge[i] = c->sp[i].charge;
        a.weight[i] = c->sp[i].weight;
        a.start[i] = c->sp[i].start;
        a.perm[i] = c->sp[i].indexed ? c->sp[i].perm : nullptr;
        const int b = c->sp[i].cur;
        a.ux[i] = c->sp[i].u[b][0]; a.uy[i] = c->sp[i].u[b][1]; a.uz[i] = c->sp[i].u[b][2];
        a.id[i] = c->sp[i].indexed ? c->sp[i].id_cell : c->sp[i].id[b];
        for (int j = 0; j < ns; ++j) {
            a.coulomb_log[i][j] = c->coulomb_log[i][j];
            if (j >= i && c->sp[i].charge != 0.0 && c->sp[j].charge != 0.0 && c->coulomb_log[i][j] <= 0.0) calc_debye = true;
        }
    }
    a.calc_debye = calc_debye ? 1 : 0;
    a.debye_mask = c->debye_mask;
    a.debye_drift = c->debye_drift;
    a.calculate_bmin = c->calc_bmin;
    a.allow_intra = c->allow_intra;
    a.V = c->d_V;
    a.dt = dt;
    a.seed = c->seed;
    a.step = c->step;
    // 64 / L cells share a wave in the grouped pass, L lanes and up to 8 L parcels of a
    // species each (L = 16, 32, 64), chosen by the mean population of the fullest species at
    // the last synchronisation; denser plasmas run one cell per wave with the stage
    // velocities in LDS (group 0, 512 parcels of a species).  Larger cells go to the
    // large-cell pass (kCoulombSlotsBig).
    a.maxp = kCoulombSlots;
    a.maxp_big = kCoulombSlotsBig;
    a.group = 0;
    int64_t most = 0;
    for (int i = 0; i < ns; ++i)
        if (c->sp[i].charge != 0.0) most = std::max<int64_t>(most, c->sp[i].n_sync);
    const double ppc = (double)most / std::max(1, a.n_cells);
    if (ppc <= 80.0) {
        a.group = 16; a.maxp = 128;
    } else if (ppc <= 180.0) {
        a.group = 32; a.maxp = 256;
    } else if (ppc <= 420.0) {
        // 512 slots stay 5 standard deviations above a mean of 420 (Poisson counts)
        a.group = 64; a.maxp = 512;
    } else {
        a.maxp = 512;   // LDS lists and velocity slots for the denser cells
    }
    a.big_list = c->d_big_list;
    a.big_n = c->d_big_n;
    if (!c->big_n_zero) hipMemsetAsync(c->d_big_n, 0, sizeof(int), c->stream);
    c->big_n_zero = false;
    a.model = c->coulomb_model;
    a.pairing = c->pairing;
    a.weight_correction = c->weight_corr;
    a.flags = c->d_flags;
    a.ncoll = c->d_stats + kStatsNcoll;
    a.pair = c->d_pair;
    return check_hip(c, launch_coulomb(a, c->mixed(), c->stream), "coulomb kernel");
}

template <class UT>
int do_mcc_t(pic_ctx* c, double dt) {
    Species& E = c->sp[c->electron];
    MccArgs<UT> a{};
    a.n_cells = c->hm.n_cells;
    a.e = soa<UT>(E, E.cur);
    a.start = E.start;
    a.perm = E.indexed ? E.perm : nullptr;
    a.id_cell = E.indexed ? E.id_cell : E.id[E.cur];
    a.n_e = live_n(c, c->electron);
    a.cap_e = E.capacity;
    if (c->ion >= 0) {
        Species& I = c->sp[c->ion];
        a.ion = soa<UT>(I, I.cur);
        a.n_i = live_n(c, c->ion);
        a.cap_i = I.capacity;
        a.count_i = I.count;
        a.s42_i = fixed_scale(c, c->ion);
    } else {
        a.n_i = live_n(c, c->electron);   // unused: no ions are created
        a.cap_i = 0;
        a.count_i = E.count;
    }
    a.smax = c->d_smax;
    a.rem = c->d_rem;
    for (int k = 0; k < 3; ++k) {
        a.xs_e[k] = c->d_xs_e[k]; a.xs_s[k] = c->d_xs_s[k]; a.xs_n[k] = c->xs_n[k];
        a.xs_guide[k] = c->d_xs_guide[k]; a.xs_le0[k] = c->xs_le0[k]; a.xs_idl[k] = c->xs_idl[k];
    }
    a.th_ex = c->xs_th[1];
    a.th_iz = c->xs_th[2];
    a.relativistic = c->en_relativistic;
    a.massE = E.mass;
    a.massN = c->bg_m;
    a.nN = c->bg_n;
    a.vthN = std::sqrt(kKB * c->bg_T / c->bg_m);
    for (int d = 0; d < 3; ++d) a.drift[d] = c->bg_v[d];
    a.wEwN = E.weight / c->bg_w;
    a.wNwE = 1.0 / a.wEwN;
    a.create_ion = (c->create_ion && c->ion >= 0) ? 1 : 0;
    a.create_electron = c->create_electron;
    a.dt = dt;
    a.seed = c->seed;
    a.step = c->step;
    // one key/index list in LDS: 512 electrons per cell in the per-cell pass (6 KiB), 4096
    // in the large-cell pass (48 KiB)
    a.maxp = kMccSlots;
    a.maxp_big = kMccSlotsBig;
    // up to 120 (150, 300) electrons per cell on average at the last synchronisation, eight
    // (four, two) cells share a wave with LDS room for 1.5 times the mean plus 40; fuller
    // cells take the large-cell pass.  A cell draws a few candidates per step, so the
    // candidate rounds run about as fast with 8 lanes per cell as with 16, for twice the cells
    const double ppc = (double)E.n_sync / std::max(1, a.n_cells);
    if (ppc <= 120.0) {
        a.group = 8;
        a.maxp_group = std::min(256, ((int)(1.5 * ppc) + 40 + 7) & ~7);
    } else if (ppc <= 150.0) {
        a.group = 16;
        a.maxp_group = std::min(256, ((int)(1.5 * ppc) + 40 + 7) & ~7);
    } else if (ppc <= 300.0) {
        a.group = 32;
        a.maxp_group = std::min(512, ((int)(1.5 * ppc) + 40 + 7) & ~7);
    }
    a.big_list = c->d_big_list;
    a.big_n = c->d_big_n;
    hipMemsetAsync(c->d_big_n, 0, sizeof(int), c->stream);
    c->big_n_zero = false;
    a.flags = c->d_flags;
    a.counts = c->d_stats + kStatsMcc;
    a.tets = c->d_tets;
    a.count_e = E.count;
    a.acc_cell = c->d_acc_cell;
    a.acc_point = c->d_acc_point;
    a.s42_e = fixed_scale(c, c->electron);
    a.dep_vw = c->cd == kVolumeWeighting ? 1 : 0;
    return check_hip(c, launch_mcc(a, c->stream), "mcc kernel");
}

// IonNeutralCollisionModel::performBackgroundCollisions (IonNeutralCollisionModel.C:447-560)
// on the MCC kernels: the ion species in the electrons' place, no parcels created
template <class UT>
int do_ion_mcc_t(pic_ctx* c, double dt) {
    Species& I = c->sp[c->ion];
    MccArgs<UT> a{};
    a.n_cells = c->hm.n_cells;
    a.e = soa<UT>(I, I.cur);
    a.start = I.start;
    a.perm = I.indexed ? I.perm : nullptr;
    a.id_cell = I.indexed ? I.id_cell : I.id[I.cur];
    a.n_e = live_n(c, c->ion);
    a.cap_e = I.capacity;
    a.n_i = a.n_e;
    a.count_i = I.count;
    a.smax = c->d_smax_i;
    a.rem = c->d_rem_i;
    for (int k = 0; k < 2; ++k) {
        const int p = PIC_XS_ION_ELASTIC + k;
        a.xs_e[k] = c->d_xs_e[p]; a.xs_s[k] = c->d_xs_s[p]; a.xs_n[k] = c->xs_n[p];
        a.xs_guide[k] = c->d_xs_guide[p]; a.xs_le0[k] = c->xs_le0[p]; a.xs_idl[k] = c->xs_idl[p];
    }
    a.th_ex = c->xs_th[PIC_XS_ION_CHARGE_EXCHANGE];
    a.ion_neutral = 1;
    a.massE = I.mass;
    a.massN = c->bg_m;
    a.nN = c->bg_n;
    a.vthN = std::sqrt(kKB * c->bg_T / c->bg_m);
    for (int d = 0; d < 3; ++d) a.drift[d] = c->bg_v[d];
    a.dt = dt;
    a.seed = c->seed;
    a.step = c->step;
    a.maxp = kMccSlots;
    a.maxp_big = kMccSlotsBig;
    const double ppc = (double)I.n_sync / std::max(1, a.n_cells);
    if (ppc <= 120.0) {
        a.group = 8;
        a.maxp_group = std::min(256, ((int)(1.5 * ppc) + 40 + 7) & ~7);
    } else if (ppc <= 150.0) {
        a.group = 16;
        a.maxp_group = std::min(256, ((int)(1.5 * ppc) + 40 + 7) & ~7);
    } else if (ppc <= 300.0) {
        a.group = 32;
        a.maxp_group = std::min(512, ((int)(1.5 * ppc) + 40 + 7) & ~7);
    }
    a.big_list = c->d_big_list;
    a.big_n = c->d_big_n;
    hipMemsetAsync(c->d_big_n, 0, sizeof(int), c->stream);
    c->big_n_zero = false;
    a.flags = c->d_flags;
    a.counts = c->d_stats + kStatsMcc;
    a.tets = c->d_tets;
    a.count_e = I.count;
    return check_hip(c, launch_mcc(a, c->stream), "ion-neutral mcc kernel");
}

// the parcel-pair passes' arguments for a list of species (the first na form group A)
template <class UT>
PairArgs<UT> pair_args(pic_ctx* c, const std::vector<int>& list, int na, double dt) {
    PairArgs<UT> a{};
    a.n_cells = c->hm.n_cells;
    a.ns = (int)list.size();
    a.na = na;
    for (int j = 0; j < a.ns; ++j) {
        Species& S = c->sp[list[j]];
        a.p[j] = soa<UT>(S, S.cur);
        a.start[j] = S.start;
        a.perm[j] = S.indexed ? S.perm : nullptr;
        a.id_cell[j] = S.indexed ? S.id_cell : S.id[S.cur];
        a.w[j] = S.weight;
        a.mass[j] = S.mass;
    }
    a.vol = c->d_V;
    a.tets = c->d_tets;
    a.C = c->d_C;
    a.pts = c->d_points;
    a.empty_mask = c->hm.track_mask;
    a.dt = dt;
    a.seed = c->seed;
    a.step = c->step;
    a.big_list = c->d_big_list;
    a.big_n = c->d_big_n;
    hipMemsetAsync(c->d_big_n, 0, sizeof(int), c->stream);
    c->big_n_zero = false;
    a.flags = c->d_flags;
    a.counts = c->d_stats + kStatsMcc;
    return a;
}

// IonNeutralCollisionModel::performCollisions (IonNeutralCollisionModel.C:580-713): the ion
// species against the neutral species' parcels
template <class UT>
int do_ion_pairs_t(pic_ctx* c, double dt, int neutral, int ion) {
    PairArgs<UT> a = pair_args<UT>(c, {neutral, ion}, 1, dt);
    a.smax = c->d_smax_ip;
    a.rem = c->d_rem_ip;
    for (int k = 0; k < 2; ++k) {
        const int p = PIC_XS_ION_ELASTIC + k;
        a.xs_e[k] = c->d_xs_e[p]; a.xs_s[k] = c->d_xs_s[p]; a.xs_n[k] = c->xs_n[p];
        a.xs_guide[k] = c->d_xs_guide[p]; a.xs_le0[k] = c->xs_le0[p]; a.xs_idl[k] = c->xs_idl[p];
    }
    a.th_cx = c->xs_th[PIC_XS_ION_CHARGE_EXCHANGE];
    a.keep_neutral = c->ion_keep_neutral;
    a.wc = c->ion_wc;
    return check_hip(c, launch_pairs(a, kPairIon, c->stream), "ion-neutral pair kernel");
}

// the ion species of the ion-neutral model: the background's ion, else the one charged
// species that is not the electrons (one at most, checked at creation)
int ion_species_of(const pic_ctx* c) {
    if (c->ion >= 0) return c->ion;
    for (int s = 0; s < (int)c->sp.size(); ++s)
        if (c->sp[s].charge != 0.0 && s != c->electron) return s;
    return -1;
}

// IonNeutralCollisionModel::handleCollisions (IonNeutralCollisionModel.C:246-254): the ions
// with the neutral parcels, then with the background gas
int do_ion_mcc(pic_ctx* c, double dt) {
    pic_ctx::Scope sc(c, PIC_STAGE_MCC);
    if (c->xs_n[PIC_XS_ION_ELASTIC] + c->xs_n[PIC_XS_ION_CHARGE_EXCHANGE] == 0)
        return c->fail(PIC_ERR_INVALID, "no ion-neutral cross sections set");
    int neutral = -1;
    for (int s = 0; s < (int)c->sp.size(); ++s)
        if (c->sp[s].charge == 0.0) neutral = s;
    const int ion = ion_species_of(c);
    if (neutral >= 0 && ion >= 0 && c->sp[neutral].n > 0 && c->sp[ion].n > 0) {
        const int r = c->mixed() ? do_ion_pairs_t<float>(c, dt, neutral, ion) : do_ion_pairs_t<double>(c, dt, neutral, ion);
        if (r) return r;
    }
    if (!c->bg || c->sp[c->ion].n == 0) return PIC_OK;
    return c->mixed() ? do_ion_mcc_t<float>(c, dt) : do_ion_mcc_t<double>(c, dt);
}

// BinaryCollisionModel::handleCollisions (BinaryCollisionModel.C:527-535): the neutral
// species' parcels in pairs, then with the background gas
template <class UT>
int do_binary_t(pic_ctx* c, double dt, const std::vector<int>& nl) {
    PairArgs<UT> a = pair_args<UT>(c, nl, (int)nl.size(), dt);
    a.smax = c->d_smax_n;
    a.rem = c->d_rem_n;
    a.rem_bg = c->d_rem_nb;
    a.tab = c->d_dsmc_tab;
    a.vhs = c->vhs;
    a.wc = c->binary_wc;
    a.bg = c->bg ? 1 : 0;
    a.nN = c->bg_n;
    a.vthN = c->bg ? std::sqrt(kKB * c->bg_T / c->bg_m) : 0.0;
    for (int d = 0; d < 3; ++d) a.drift[d] = c->bg_v[d];
    a.massN = c->bg_m;
    return check_hip(c, launch_pairs(a, kPairDsmc, c->stream), "binary collision kernel");
}

int do_binary(pic_ctx* c, double dt) {
    pic_ctx::Scope sc(c, PIC_STAGE_MCC);
    std::vector<int> nl;
    bool any = false;
    for (int s = 0; s < (int)c->sp.size(); ++s)
        if (c->sp[s].charge == 0.0) {
            nl.push_back(s);
            any = any || c->sp[s].n > 0;
        }
    if (!any) return PIC_OK;
    return c->mixed() ? do_binary_t<float>(c, dt, nl) : do_binary_t<double>(c, dt, nl);
}

int do_mcc(pic_ctx* c, double dt) {
    pic_ctx::Scope sc(c, PIC_STAGE_MCC);
    if (c->xs_n[0] + c->xs_n[1] + c->xs_n[2] == 0) return c->fail(PIC_ERR_INVALID, "no cross sections set");
    const int r = c->mixed() ? do_mcc_t<float>(c, dt) : do_mcc_t<double>(c, dt);
    c->place_valid = false;   // created parcels have no slot; the next sort recounts
    // created parcels are appended on the device: until the next synchronisation the
    // launch bound of the species that can grow is their capacity
    if (c->create_electron) c->sp[c->electron].n = c->sp[c->electron].capacity;
    if (c->create_ion && c->ion >= 0) c->sp[c->ion].n = c->sp[c->ion].capacity;
    return r;
}

int do_deposit_finish(pic_ctx* c) {
    pic_ctx::Scope sc(c, PIC_STAGE_DEPOSIT);
    DepositFinishArgs a{};
    a.n_cells = c->hm.n_cells;
    a.n_species = (int)c->sp.size();
    for (int s = 0; s < a.n_species; ++s) {
        a.counts[s] = c->sp[s].count;
        a.whole[s] = llrint(fixed_scale(c, s));
    }
    a.acc_cell = c->d_acc_cell;
    a.acc_point = c->d_acc_point;
    a.cpo = c->d_cpo;
    a.cp_master = c->d_cp_master;
    a.vw_w = c->d_vw_w;
    a.V = c->d_V;
    a.space_charge = c->has_space ? c->d_space : nullptr;
    a.q42 = c->Q0 / kFixedScale;
    a.weighting = c->cd;
    a.rho = c->d_rho;
    return check_hip(c, launch_deposit_finish(a, c->stream), "deposit finish");
}

int update_field_derived(pic_ctx* c) {
    pic_ctx::Scope sc(c, PIC_STAGE_FIELD);
    const HostMesh& m = c->hm;
    PointInterpArgs pa{};
    pa.n_points = m.n_points;
    pa.pco = c->d_pco;
    pa.pc_cell = c->d_pc_cell;
    pa.pc_w = c->d_pc_w;
    pa.pmaster = c->d_pmaster;
    pa.Ecell = c->d_E;
    pa.Epoint = c->d_Ep;
    pa.Epoint_f = c->d_Epf;
    hipError_t e = hipSuccess;
    if (c->fw == kVolumeWeighting) e = launch_point_interp(pa, c->stream);
    if (e == hipSuccess) e = launch_to_float(c->d_E, c->d_Ef, 3 * (int64_t)m.n_cells, c->stream);
    return check_hip(c, e, "field update");
}

int do_poisson(pic_ctx* c, pic_poisson_stats* st) {
    if (!c->has_poisson) return c->fail(PIC_ERR_INVALID, "MaxwellSolver is not ElectroStatic");
    hipEvent_t e0 = c->ev(), e1 = c->ev();
    hipEventRecord(e0, c->stream);
    int iters = 0, refs = 0, status;
    double rel = 0.0;
    {
        pic_ctx::Scope sc(c, PIC_STAGE_POISSON);
        hipError_t e = hipSuccess;
        if (c->d_corr) {
            // the explicit non-orthogonal correction from the potential the solve starts
            // from (the last solution), before any extrapolated or learned start replaces it
            GradArgs g = c->gtab;
            g.phi = c->d_phi;
            g.E = c->d_mgrad;
            g.E_f = nullptr;
            e = launch_grad(g, c->stream);
            if (e == hipSuccess) e = launch_nonorth_corr(g, c->d_corr, c->stream);
            if (int r = check_hip(c, e, "non-orthogonal correction")) return r;
        }
        e = poisson_rhs(c->pd, c->d_rho, c->d_corr, c->d_b, c->hp.singular, c->pw.red, c->stream);
        if (int r = check_hip(c, e, "poisson rhs")) return r;
        if (c->poisson_guess == 1) {
            // warm start 2 phi(n-1) - phi(n-2) once two solutions exist
            if (c->n_solves >= 2)
                poisson_extrapolate(c->d_phi, c->d_phi_old, c->hm.n_cells, c->stream);
            else
                hipMemcpyAsync(c->d_phi_old, c->d_phi, sizeof(double) * c->hm.n_cells, hipMemcpyDeviceToDevice,
                               c->stream);
        }
        ++c->n_solves;
        const bool learned = c->gnn && (c->gnn_use || c->gnn_rec);
        if (learned) {
            // features of the warm start: r = b - A phi_prev, b, phi_prev
            hipMemcpyAsync(c->d_phi_prev, c->d_phi, sizeof(double) * c->hm.n_cells, hipMemcpyDeviceToDevice, c->stream);
            poisson_residual(c->pd, c->pw, c->d_phi_prev, c->d_b, c->pw.r, c->stream);
            c->gnn->features(c->pw.r, c->d_b, c->d_phi_prev);
            if (c->gnn_use) {
                c->gnn->forward(c->gnn->x, false);
                c->gnn->guess(c->d_phi_prev, c->d_phi);
            }
            if (int r = check_hip(c, hipGetLastError(), "learned guess")) return r;
        }
        status = poisson_solve(c->pd, c->pw, c->d_phi, c->d_b, c->ptol, c->pmaxit, c->pmode, c->hp.singular, &iters,
                               &refs, &rel, c->stream);
        if (c->gnn && c->gnn_rec && status == 0) {
            Gnn& G = *c->gnn;
            const int slot = G.snap_next;
            const size_t n = (size_t)c->hm.n_cells;
            G.target(c->d_phi, c->d_phi_prev, G.snap_t + slot * n);
            hipMemcpyAsync(G.snap_x + slot * n * 3, G.x, sizeof(float) * 3 * n, hipMemcpyDeviceToDevice, c->stream);
            G.snap_norm[slot] = gnn_anorm(G, G.snap_t + slot * n, c->pd.nb, c->pd.a, c->pd.diag, c->pd.K);
            G.snap_rnorm[slot] = gnn_rnorm(G, G.snap_t + slot * n, c->pd.nb, c->pd.a, c->pd.diag, c->pd.K);
            G.snap_next = (slot + 1) % G.snap_cap;
            G.snap_count = std::min(G.snap_count + 1, G.snap_cap);
        }
        GradArgs g = c->gtab;
        g.phi = c->d_phi;
        g.E = c->d_E;
        g.E_f = c->d_Ef;
        if (int r = check_hip(c, launch_grad(g, c->stream), "grad")) return r;
    }
    c->e_zero = false;
    if (int r = update_field_derived(c)) return r;
    hipEventRecord(e1, c->stream);
    hipEventSynchronize(e1);
    float ms = 0.f;
    hipEventElapsedTime(&ms, e0, e1);
    c->pool.push_back(e0);
    c->pool.push_back(e1);
    if (st) { st->iterations = iters; st->refinements = refs; st->rel_residual = rel; st->ms = ms; }
    if (status == -2) return c->fail(PIC_ERR_HIP, "poisson kernels failed");
    if (status == -1) return c->fail(PIC_ERR_NOT_CONVERGED, "Poisson PCG did not reach the tolerance");
    return PIC_OK;
}

}  // namespace

// ====================================================================== C-ABI

// BoundaryModelList::injection (BoundaryModelList.C:193-200) over the SimpleInjection models
// (SimpleInjection.C:186-229): the parcel counts are host arithmetic on the frequencies and
// remainders; the parcels are written on the device after each species' live ones.
template <class UT>
int do_inject_t(pic_ctx* c, double dt) {
    const double t_end = c->inj_time + dt;   // the reference tests the time the step ends at
    for (size_t m = 0; m < c->inj.size(); ++m) {
        auto& I = c->inj[m];
        if (t_end > I.cfg.end_time) continue;
        for (int e = 0; e < I.cfg.n_species; ++e) {
            const int s = I.cfg.species[e];
            Species& S = c->sp[s];
            const double np = I.cfg.frequency[e] / S.weight * dt + I.rem[e];
            const int64_t n = (int64_t)np;
            I.rem[e] = np - (double)n;
            if (n <= 0) continue;
            InjectArgs<UT> a{};
            a.m = move_args<UT>(c, s, dt);
            a.tri = I.d_tri;
            a.cum = I.d_cum;
            a.cell_tet0 = c->d_cell_tet0;
            a.ntri = I.ntri;
            a.vmodel = I.cfg.velocity_model;
            const double T = I.cfg.eV[e] * kQe / kKB;   // SimpleInjection.C:215
            a.vt = std::sqrt(kKB * T / S.mass);
            a.model = (uint32_t)m;
            a.sync = (c->wall_sync && S.charge != 0.0) ? 1 : 0;
            a.count = n;
            a.cap = S.capacity;
            if (int r = check_hip(c, launch_inject(a, c->pusher, c->fw, c->stream), "injection")) return r;
            S.n = std::min<int64_t>(S.n + n, S.capacity);   // an upper bound until the next synchronisation
            c->n_injected += n;
            c->counts_valid = false;
            c->place_valid = false;
        }
    }
    c->inj_time = t_end;
    return PIC_OK;
}

int do_inject(pic_ctx* c, double dt) {
    if (c->inj.empty()) return PIC_OK;
    pic_ctx::Scope sc(c, PIC_STAGE_MOVE);
    return c->mixed() ? do_inject_t<float>(c, dt) : do_inject_t<double>(c, dt);
}

extern "C" {

const char* pic_version(void) { return "picmi 0.1 (gfx950)"; }

int pic_device_count(int32_t* n) {
    int d = 0;
    if (hipGetDeviceCount(&d) != hipSuccess) d = 0;
    if (n) *n = d;
    return PIC_OK;
}

const char* pic_last_error(const pic_ctx* ctx) { return ctx ? ctx->err.c_str() : g_last_error.c_str(); }

void pic_destroy(pic_ctx* c) {
    if (!c) return;
    if (c->stream) hipStreamSynchronize(c->stream);
    if (c->gnn) c->gnn->release();
    for (void* p : c->allocs) hipFree(p);
    for (auto& pe : c->pending) { hipEventDestroy(pe.a); hipEventDestroy(pe.b); }
    for (auto e : c->pool) hipEventDestroy(e);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
}

int pic_create(const pic_mesh* mesh, const pic_species* species, int32_t n_species, const pic_models* models,
               pic_ctx** out) {
    if (!out) { g_last_error = "out is null"; return PIC_ERR_INVALID; }
    *out = nullptr;
    if (!mesh || !species || !models || n_species <= 0 || n_species > kMaxSpecies) {
        g_last_error = "invalid arguments to pic_create";
        return PIC_ERR_INVALID;
    }
    std::unique_ptr<pic_ctx, void (*)(pic_ctx*)> c(new pic_ctx(), pic_destroy);
    int r;
    int ms = 0;
    if ((r = parse_model(c.get(), models->particle_pusher, "ParticlePusher",
                         {"none", "Boris", "BorisNR", "Vay", "HigueraCary"}, &c->pusher)))
        return r;
    if ((r = parse_model(c.get(), models->field_weighting, "FieldWeigthing", {"CellAverage", "VolumeWeighting"}, &c->fw)))
        return r;
    if ((r = parse_model(c.get(), models->charge_distribution, "ChargeDistribution", {"CellAverage", "VolumeWeighting"},
                         &c->cd)))
        return r;
    if ((r = parse_model(c.get(), models->maxwell_solver, "MaxwellSolver", {"none", "ElectroStatic"}, &ms))) return r;
    c->electrostatic = ms == 1;
    int cm = 0;
    if ((r = parse_model(c.get(), models->coulomb_collision, "CoulombCollisionModel", {"none", "Nanbu", "Perez"}, &cm)))
        return r;
    c->coulomb = cm != 0;
    c->coulomb_model = cm == 2 ? 1 : 0;
    if (c->coulomb) {
        int pa = 0, wc = 0;
        if ((r = parse_model(c.get(), models->pairing_algorithm ? models->pairing_algorithm : "TakizukaAbe",
                             "PairingAlgorithm", {"TakizukaAbe", "NanbuYonemura"}, &pa)))
            return r;
        if ((r = parse_model(c.get(), models->weight_correction ? models->weight_correction : "none",
                             "WeightCorrectionModel", {"none", "Nanbu"}, &wc)))
            return r;
        c->pairing = pa;
        c->weight_corr = wc;
    }
    int en = 0, bgm = 0;
    if ((r = parse_model(c.get(), models->electron_neutral, "ElectronNeutralCollisionModel",
                         {"none", "NanbuElectronNeutral", "RelativisticElectronNeutral"}, &en)))
        return r;
    if ((r = parse_model(c.get(), models->background_gas, "BackgroundGasModel", {"none", "uniformBackground"}, &bgm)))
        return r;
    int inm = 0;
    if ((r = parse_model(c.get(), models->ion_neutral_collision ? models->ion_neutral_collision : "none",
                         "IonNeutralCollisionModel", {"none", "IsotropicScattering"}, &inm)))
        return r;
    c->mcc = en != 0;
    c->en_relativistic = en == 2 ? 1 : 0;
    c->bg = bgm == 1;
    c->ion_mcc = inm != 0;
    if (c->ion_mcc) {
        int wc = 0;
        if ((r = parse_model(c.get(), models->ion_neutral_weight_correction ? models->ion_neutral_weight_correction : "none",
                             "WeightCorrectionModel", {"none", "Nanbu"}, &wc)))
            return r;
        c->ion_wc = wc;
        c->ion_keep_neutral = models->ion_neutral_keep_neutral ? 1 : 0;
    }
    // BinaryCollisionModel (BinaryCollisionModelNew.C:39-61) with its total cross section
    // (TotalCrossSectionModelNew.C) and weight correction
    int bin = 0;
    if ((r = parse_model(c.get(), models->binary_collision ? models->binary_collision : "none", "BinaryCollisionModel",
                         {"none", "IsotropicScattering"}, &bin)))
        return r;
    c->binary = bin != 0;
    if (c->binary) {
        int xs = 0, wc = 0;
        if ((r = parse_model(c.get(), models->total_cross_section ? models->total_cross_section : "HardSphere",
                             "TotalCrossSectionModel", {"HardSphere", "VariableHardSphere"}, &xs)))
            return r;
        if ((r = parse_model(c.get(), models->binary_weight_correction ? models->binary_weight_correction : "none",
                             "WeightCorrectionModel", {"none", "Nanbu"}, &wc)))
            return r;
        c->vhs = xs;
        c->binary_wc = wc;
        c->vhs_tref = models->vhs_tref;
        if (c->vhs && !(c->vhs_tref > 0.0))
            return c->fail(PIC_ERR_INVALID, "VariableHardSphereCoeffs.Tref must be positive");
    }
    c->bg_species = models->bg_species;
    if (c->mcc && !c->bg) {
        g_last_error = c->err = "ElectronNeutralCollisionModel needs the uniformBackground gas model (particle neutrals are not implemented)";
        return PIC_ERR_UNSUPPORTED;
    }
    if ((r = parse_model(c.get(), models->wall_reflection ? models->wall_reflection : "SpecularReflection",
                         "WallReflectionModel", {"SpecularReflection", "MaxwellianThermal", "MixedDiffuseSpecular"},
                         &c->wall_model)))
        return r;
    c->diffuse_fraction = models->diffuse_fraction;
    if (c->wall_model == 2 && !(c->diffuse_fraction >= 0.0 && c->diffuse_fraction <= 1.0))
        return c->fail(PIC_ERR_INVALID, "MixedDiffuseSpecularCoeffs.diffuseFraction must lie in [0, 1]");
    c->wall_sync = models->no_boundary_sync ? 0 : 1;
    c->precision = models->precision;
    if (c->precision != PIC_PREC_FP64 && c->precision != PIC_PREC_MIXED) return c->fail(PIC_ERR_INVALID, "precision");
    c->hc_ref = models->hc_reference_form;
    c->allow_intra = models->allow_intra_collision;
    c->calc_bmin = models->calculate_bmin;
    if (models->debye_species) {
        // CoulombCollisionModel.C:106-121: every listed species must exist and carry charge
        for (int k = 0; k < 32; ++k) {
            if (!((models->debye_species >> k) & 1u)) continue;
            if (k >= n_species) return c->fail(PIC_ERR_INVALID, "DebyeLength species " + std::to_string(k) + " does not exist");
            if (species[k].charge == 0.0)
                return c->fail(PIC_ERR_INVALID, "Species " + std::to_string(k) +
                                                    " considered for calculation of the Debye length is not a particle with charge");
        }
        c->debye_mask = models->debye_species;
    }
    c->debye_drift = models->debye_ignore_drift ? 0 : 1;
    std::memcpy(c->coulomb_log, models->coulomb_log, sizeof(c->coulomb_log));
    c->electron = models->electron_species;
    c->ion = models->ion_species;
    c->bg_n = models->bg_number_density;
    c->bg_T = models->bg_temperature;
    c->bg_m = models->bg_mass;
    c->bg_w = models->bg_n_equivalent;
    for (int d = 0; d < 3; ++d) c->bg_v[d] = models->bg_velocity[d];
    c->create_ion = models->create_ion;
    c->create_electron = models->create_electron;
    c->seed = models->seed;
    c->ptol = models->poisson_tolerance > 0 ? models->poisson_tolerance : 1e-8;
    c->pmaxit = models->poisson_max_iter > 0 ? models->poisson_max_iter : 10000;
    c->pmode = models->poisson_mode;
    c->sort_interval = models->sort_interval > 0 ? models->sort_interval : 1;
    c->poisson_guess = models->poisson_guess;
    if (c->poisson_guess < 0 || c->poisson_guess > 1) return c->fail(PIC_ERR_INVALID, "poisson_guess");
    if (c->ion_mcc && c->bg) {
        // IonNeutralCollisionModel.C:181-188: the background neutral needs its ion species
        if (c->ion < 0 || c->ion >= n_species)
            return c->fail(PIC_ERR_INVALID, "IonNeutralCollisionModel: no ion type for the background gas (SpeciesRelations ionTypeId)");
        if (!(c->bg_m > 0)) return c->fail(PIC_ERR_INVALID, "background gas mass");
    }
    if (c->binary && c->bg && !(c->bg_m > 0)) return c->fail(PIC_ERR_INVALID, "background gas mass");
    if (c->mcc) {
        if (c->electron < 0 || c->electron >= n_species) return c->fail(PIC_ERR_INVALID, "electronTypeId not set");
        if (c->ion >= n_species) return c->fail(PIC_ERR_INVALID, "ion species out of range");
        if (!(c->bg_m > 0 && c->bg_w > 0)) return c->fail(PIC_ERR_INVALID, "background gas mass / weight");
        if (c->ion >= 0 && species[c->ion].n_equivalent != c->bg_w)
            return c->fail(PIC_ERR_UNSUPPORTED, "created ions carry the background weight; the ion species weight must equal it");
        if (species[c->electron].n_equivalent != c->bg_w && c->create_electron)
            return c->fail(PIC_ERR_UNSUPPORTED, "created electrons carry the background weight; the electron weight must equal it");
    }
    // nParticleEqWeight (PICCloud.C:86-93, :414-417): zero unless every species and the
    // background gas carry the same weight
    double eqw = species[0].n_equivalent;
    for (int s = 1; s < n_species; ++s)
        if (species[s].n_equivalent != eqw) eqw = 0.0;
    if (c->bg && c->bg_w != eqw) eqw = 0.0;
    if (c->coulomb && c->pairing == 0 && eqw == 0.0)
        return c->fail(PIC_ERR_UNSUPPORTED, "Model TakizukaAbe is only valid for equal weighted particles");
    // WeightCorrectionModel::active (WeightCorrectionModel.C:91-95)
    c->weight_corr = (c->weight_corr && eqw == 0.0) ? 1 : 0;
    // the model selection above is validated before any device is touched
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        g_last_error = "no HIP device available";
        return PIC_ERR_NO_DEVICE;
    }
    hipGetDevice(&c->device);
    // the allocation-free entry points' scratch address on this device, once per device
    if (hipError_t e = noalloc_init(c->device)) {
        g_last_error = std::string("device scratch: ") + hipGetErrorString(e);
        return PIC_ERR_HIP;
    }
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        g_last_error = "hipStreamCreate failed";
        return PIC_ERR_HIP;
    }
    // mesh
    HostMesh& m = c->hm;
    m.n_points = mesh->n_points; m.n_faces = mesh->n_faces; m.n_internal = mesh->n_internal_faces;
    m.n_cells = mesh->n_cells; m.n_patches = mesh->n_patches;
    if (m.n_points <= 0 || m.n_faces <= 0 || m.n_cells <= 0 || m.n_internal < 0 || m.n_internal > m.n_faces)
        return c->fail(PIC_ERR_MESH, "empty or inconsistent mesh sizes");
    m.points.assign(mesh->points, mesh->points + 3 * (size_t)m.n_points);
    m.fo.assign(mesh->face_offsets, mesh->face_offsets + m.n_faces + 1);
    if (m.fo[0] != 0) return c->fail(PIC_ERR_MESH, "face_offsets[0] must be 0");
    m.fp.assign(mesh->face_points, mesh->face_points + m.fo[m.n_faces]);
    m.owner.assign(mesh->owner, mesh->owner + m.n_faces);
    m.nbr.assign(mesh->neighbour, mesh->neighbour + m.n_internal);
    for (int p = 0; p < m.n_patches; ++p) {
        const pic_patch& pp = mesh->patches[p];
        if (pp.type == PIC_PATCH_WALL) {
            c->walls = true;
            if (!(pp.wall_temperature >= 0.0)) return c->fail(PIC_ERR_INVALID, "patch " + std::to_string(p) + ": wall temperature");
        }
        if (pp.type < PIC_PATCH_PATCH || pp.type > PIC_PATCH_EMPTY)
            return c->fail(PIC_ERR_INVALID, "patch " + std::to_string(p) + ": unknown type");
        if (pp.type == PIC_PATCH_EMPTY && pp.phi_bc != PIC_BC_ZERO_GRADIENT)
            return c->fail(PIC_ERR_INVALID, "patch " + std::to_string(p) + ": an empty patch takes no phi value");
        m.ptype.push_back(pp.type); m.pstart.push_back(pp.start); m.psize.push_back(pp.size);
        m.pnbr.push_back(pp.neighbour_patch); m.phi_bc.push_back(pp.phi_bc); m.phi_val.push_back(pp.phi_value);
    }
    std::string merr = build_host_mesh(m);
    if (!merr.empty()) return c->fail(PIC_ERR_MESH, merr);
    if (c->electrostatic) {
        merr = build_host_poisson(m, c->hp);
        if (!merr.empty()) return c->fail(PIC_ERR_MESH, merr);
        c->has_poisson = true;
    }
    // species
    double qmax = 0.0;
    for (int s = 0; s < n_species; ++s) {
        Species S;
        S.mass = species[s].mass;
        S.charge = species[s].charge;
        S.weight = species[s].n_equivalent;
        S.diameter = species[s].diameter;
        S.omega = species[s].omega;
        S.solve_movement = species[s].solve_movement;
        S.capacity = species[s].capacity;
        S.sort_interval = species[s].sort_interval > 0 ? species[s].sort_interval : c->sort_interval;
        if (!(S.mass > 0) || !(S.weight > 0) || S.capacity < 0 || S.capacity > 2000000000LL)
            return c->fail(PIC_ERR_INVALID, "species " + std::to_string(s) + ": mass, weight and capacity must be positive (capacity < 2^31)");
        qmax = std::max(qmax, std::fabs(S.charge * S.weight));
        c->sp.push_back(S);
    }
    c->Q0 = qmax > 0 ? qmax : 1.0;
    if (c->ion_mcc) {
        // IonNeutralCollisionModel.C:170-175: one neutral and one ion species at most
        int nn = 0, ni = 0;
        for (int s = 0; s < n_species; ++s) {
            if (c->sp[s].charge == 0.0) ++nn;
            else if (s != c->electron) ++ni;
        }
        if (nn > 1 || ni > 1)
            return c->fail(PIC_ERR_UNSUPPORTED, "IonNeutralCollisionModel supports one neutral and one ion species only");
    }
    if (c->binary && c->bg && (c->bg_species < 0 || c->bg_species >= n_species || c->sp[c->bg_species].charge != 0.0))
        return c->fail(PIC_ERR_INVALID, "BinaryCollisionModel with a background gas: bg_species must name a neutral species");
    {
        std::vector<double> ms(n_species), qs(n_species);
        for (int s = 0; s < n_species; ++s) { ms[s] = c->sp[s].mass; qs[s] = c->sp[s].charge; }
        std::vector<PairConst> pc(kMaxSpecies * kMaxSpecies);
        pair_constants(ms.data(), qs.data(), n_species, pc.data());
        c->d_pair = c->upload(pc);
    }
    // device mesh
    const int nc = m.n_cells, np = m.n_points;
    c->d_tets = c->upload(m.tets);
    if (c->walls) {
        std::vector<uint8_t> kind((size_t)kMaxSpecies * m.n_patches, 0);
        std::vector<double> wT(m.n_patches, 0.0), wU(3 * (size_t)m.n_patches, 0.0);
        for (int p = 0; p < m.n_patches; ++p) {
            const pic_patch& pp = mesh->patches[p];
            if (pp.type != PIC_PATCH_WALL) continue;
            for (int s = 0; s < kMaxSpecies; ++s) kind[(size_t)s * m.n_patches + p] = 1;
            wT[p] = pp.wall_temperature;
            for (int d = 0; d < 3; ++d) wU[3 * p + d] = pp.wall_velocity[d];
        }
        c->h_kind = kind;
        c->d_patch_kind = c->upload(kind);
        // the wall surface fields: 10 totals and 9 per species for every boundary face
        const int nb = m.n_faces - m.n_internal;
        c->wall_stride = 10 + 9 * (int)c->sp.size();
        std::vector<double> fa(std::max(nb, 1), 0.0);
        for (int f = m.n_internal; f < m.n_faces; ++f)
            fa[f - m.n_internal] = std::sqrt(m.Sf[3 * f] * m.Sf[3 * f] + m.Sf[3 * f + 1] * m.Sf[3 * f + 1] +
                                             m.Sf[3 * f + 2] * m.Sf[3 * f + 2]);
        c->d_face_area = c->upload(fa);
        c->d_tet_face = c->upload(m.tet_face);
        c->d_wall_acc = c->alloc<double>((size_t)std::max(nb, 1) * c->wall_stride);
        if (!c->d_face_area || !c->d_tet_face || !c->d_wall_acc) return c->fail(PIC_ERR_HIP, "out of device memory (wall fields)");
        c->d_wall_T = c->upload(wT);
        c->d_wall_U = c->upload(wU);
        if (!c->d_patch_kind || !c->d_wall_T || !c->d_wall_U) return c->fail(PIC_ERR_HIP, "out of device memory (walls)");
    }
    c->d_cell_tet0 = c->upload(m.cell_tet0);
    c->d_C = c->upload(m.C);
    c->d_V = c->upload(m.V);
    c->d_points = c->upload(m.points);
    c->d_cpo = c->upload(m.cpo);
    {
        std::vector<int32_t> cpm(m.cp.size());
        for (size_t k = 0; k < m.cp.size(); ++k) cpm[k] = m.pmaster[m.cp[k]];
        c->d_cp_master = c->upload(cpm);
    }
    c->d_vw_w = c->upload(m.vw_w);
    c->d_pco = c->upload(m.pco);
    c->d_pc_cell = c->upload(m.pc_cell);
    c->d_pc_w = c->upload(m.pc_w);
    c->d_pmaster = c->upload(m.pmaster);
    c->d_E = c->alloc<double>(3 * (size_t)nc);
    c->d_B = c->alloc<double>(3 * (size_t)nc);
    c->d_phi = c->alloc<double>(nc);
    if (c->electrostatic && c->poisson_guess == 1) c->d_phi_old = c->alloc<double>(nc);
    c->d_rho = c->alloc<double>(nc);
    c->d_space = c->alloc<double>(nc);
    c->d_Ep = c->alloc<double>(3 * (size_t)np);
    c->d_Ef = c->alloc<float>(3 * (size_t)nc);
    c->d_Epf = c->alloc<float>(3 * (size_t)np);
    c->d_Bf = c->alloc<float>(3 * (size_t)nc);
    c->d_acc_cell = c->alloc<unsigned long long>(nc);
    c->d_acc_point = c->alloc<unsigned long long>(np);
    c->d_flags = c->alloc<uint32_t>(1);
    c->d_stats = c->alloc<unsigned long long>(kStatsWords);
    c->d_nrows = c->alloc<int64_t>(2 * kMaxSpecies);
    c->d_big_list = c->alloc<int>(nc);
    c->d_big_n = c->alloc<int>(1);
    c->scan_bytes = scan_cells_temp_bytes(nc + 1);
    c->d_scan_tmp = c->alloc<char>(c->scan_bytes);
    c->d_diag_part = c->alloc<double>(1024 * 12);
    c->d_diag_out = c->alloc<double>(16);
    if (!c->d_tets || !c->d_E || !c->d_diag_out) return c->fail(PIC_ERR_HIP, "out of device memory (mesh)");
    if (c->mcc) {
        c->d_smax = c->alloc<double>(nc);
        c->d_rem = c->alloc<double>(nc);
        // remainders start random in [0,1) (ElectronNeutralCollisionModel.C:286-296)
        std::vector<double> rem(nc);
        for (int i = 0; i < nc; ++i) {
            uint64_t x = (uint64_t)i * 0x9E3779B97F4A7C15ull ^ (c->seed + 0x632BE59BD9B4E019ull);
            x ^= x >> 33; x *= 0xff51afd7ed558ccdull; x ^= x >> 33; x *= 0xc4ceb9fe1a85ec53ull; x ^= x >> 33;
            rem[i] = (double)(x >> 11) * (1.0 / 9007199254740992.0);
        }
        hipMemcpyAsync(c->d_rem, rem.data(), sizeof(double) * nc, hipMemcpyHostToDevice, c->stream);
    }
    // candidate remainders start random in [0,1) (IonNeutralCollisionModel.C:190-194,
    // BinaryCollisionModel.C:155-159), each field from its own salt
    auto random_remainders = [&](double* d, uint64_t mul, uint64_t salt) {
        std::vector<double> rem(nc);
        for (int i = 0; i < nc; ++i) {
            uint64_t x = (uint64_t)i * mul ^ (c->seed + salt);
            x ^= x >> 33; x *= 0xff51afd7ed558ccdull; x ^= x >> 33; x *= 0xc4ceb9fe1a85ec53ull; x ^= x >> 33;
            rem[i] = (double)(x >> 11) * (1.0 / 9007199254740992.0);
        }
        hipMemcpyAsync(d, rem.data(), sizeof(double) * nc, hipMemcpyHostToDevice, c->stream);
        // the host vector dies here: the copy must have read it
        hipStreamSynchronize(c->stream);
    };
    if (c->ion_mcc) {
        c->d_smax_i = c->alloc<double>(nc);
        c->d_rem_i = c->alloc<double>(nc);
        c->d_smax_ip = c->alloc<double>(nc);
        c->d_rem_ip = c->alloc<double>(nc);
        if (!c->d_smax_i || !c->d_rem_i || !c->d_smax_ip || !c->d_rem_ip)
            return c->fail(PIC_ERR_HIP, "out of device memory (ion-neutral)");
        random_remainders(c->d_rem_i, 0xD1B54A32D192ED03ull, 0x8CB92BA72F3D8DD7ull);
        random_remainders(c->d_rem_ip, 0xA24BAED4963EE407ull, 0x9FB21C651E98DF25ull);
    }
    if (c->binary) {
        c->d_smax_n = c->alloc<double>(nc);
        c->d_rem_n = c->alloc<double>(nc);
        c->d_rem_nb = c->alloc<double>(nc);
        // pair constants in the order of the neutral species, background row last
        // (HardSphere.C:64-79, VariableHardSphere.C:64-97)
        std::vector<int> nl;
        for (int s = 0; s < n_species; ++s)
            if (c->sp[s].charge == 0.0) nl.push_back(s);
        std::vector<DsmcPair> tab((kMaxSpecies + 1) * kMaxSpecies);
        auto pair = [&](const Species& P, double dQ, double omQ, double mQ) {
            DsmcPair t{};
            const double d = 0.5 * (P.diameter + dQ);
            const double om = 0.5 * (P.omega + omQ);
            t.hs = kPi * d * d;
            t.vn = 2.0 * kKB * c->vhs_tref;
            t.mr = P.mass * mQ / (P.mass + mQ);
            t.ve = om - 0.5;
            t.gl = std::exp(std::lgamma(2.5 - om));
            return t;
        };
        for (size_t a = 0; a < nl.size(); ++a) {
            const Species& P = c->sp[nl[a]];
            for (size_t b = 0; b < nl.size(); ++b) {
                const Species& Q = c->sp[nl[b]];
                tab[a * kMaxSpecies + b] = pair(P, Q.diameter, Q.omega, Q.mass);
            }
            if (c->bg) {
                const Species& G = c->sp[c->bg_species];
                tab[kMaxSpecies * kMaxSpecies + a] = pair(P, G.diameter, G.omega, c->bg_m);
            }
        }
        c->d_dsmc_tab = c->upload(tab);
        if (!c->d_smax_n || !c->d_rem_n || !c->d_rem_nb || !c->d_dsmc_tab)
            return c->fail(PIC_ERR_HIP, "out of device memory (binary collisions)");
        random_remainders(c->d_rem_n, 0xE7037ED1A0B428DBull, 0x1D8E4E27C47D124Full);
        random_remainders(c->d_rem_nb, 0x8EBC6AF09C88C6E3ull, 0x589965CC75374CC3ull);
    }
    for (auto& S : c->sp)
        if ((r = alloc_species(c.get(), S))) return r;
    if (c->walls) {
        // every parcel of every species may end a move on a wall
        size_t cap = 0;
        for (auto& S : c->sp) cap += (size_t)S.capacity;
        if (cap > 0xFFFFFFFFull) return c->fail(PIC_ERR_CAPACITY, "wall queue: total capacity exceeds 2^32 parcels");
        c->wall_cap = (unsigned)cap;
        c->d_wall_q = c->alloc<WallHit>(std::max<size_t>(cap, 1));
        c->d_wall_n = c->alloc<unsigned>(1);
        if (!c->d_wall_q || !c->d_wall_n) return c->fail(PIC_ERR_HIP, "out of device memory (wall queue)");
    }
    // Poisson device tables
    if (c->has_poisson) {
        HostPoisson& P = c->hp;
        std::vector<float> af(P.a.begin(), P.a.end()), df(P.diag.begin(), P.diag.end()), rdf(P.rD.begin(), P.rD.end());
        c->pd.n = P.n; c->pd.K = P.K; c->pd.ncol = P.ncol;
        c->pd.nb = c->upload(P.nb);
        c->pd.a = c->upload(P.a);
        c->pd.a_f = c->upload(af);
        c->pd.diag = c->upload(P.diag);
        c->pd.diag_f = c->upload(df);
        c->pd.rD = c->upload(P.rD);
        c->pd.rD_f = c->upload(rdf);
        c->pd.colour = c->upload(P.colour);
        c->pd.corder = c->upload(P.corder);
        c->pd.col_start_h = P.col_start.data();
        c->pd.bsrc = c->upload(P.bsrc);
        c->pd.V = c->d_V;
        {
            // the system renumbered by colour, coefficients column-major [k * n + new]
            const int n = P.n, K = P.K;
            std::vector<int32_t> inv(n);
            for (int t = 0; t < n; ++t) inv[P.corder[t]] = t;
            std::vector<int32_t> nbT((size_t)n * K);
            std::vector<double> aT((size_t)n * K), dp(n), rdp(n);
            for (int t = 0; t < n; ++t) {
                const int i = P.corder[t];
                for (int k = 0; k < K; ++k) {
                    const int j = P.nb[(size_t)i * K + k];
                    nbT[(size_t)k * n + t] = j >= 0 ? inv[j] : -1;
                    aT[(size_t)k * n + t] = P.a[(size_t)i * K + k];
                }
                dp[t] = P.diag[i];
                rdp[t] = P.rD[i];
            }
            std::vector<float> aTf(aT.begin(), aT.end()), dpf(dp.begin(), dp.end()), rdpf(rdp.begin(), rdp.end());
            c->pd.nbT = c->upload(nbT);
            c->pd.aT = c->upload(aT);
            c->pd.aT_f = c->upload(aTf);
            c->pd.diag_p = c->upload(dp);
            c->pd.diag_pf = c->upload(dpf);
            c->pd.rD_p = c->upload(rdp);
            c->pd.rD_pf = c->upload(rdpf);
            c->pd.col_start_d = c->upload(P.col_start);
        }
        c->pw.r = c->alloc<double>(nc); c->pw.z = c->alloc<double>(nc); c->pw.p = c->alloc<double>(nc);
        c->pw.q = c->alloc<double>(nc); c->pw.x = c->alloc<double>(nc); c->pw.b = c->alloc<double>(nc);
        c->pw.p2 = c->alloc<double>(nc); c->pw.xp = c->alloc<double>(nc); c->pw.bp = c->alloc<double>(nc);
        c->pw.rf = c->alloc<float>(nc); c->pw.zf = c->alloc<float>(nc); c->pw.pf = c->alloc<float>(nc);
        c->pw.qf = c->alloc<float>(nc); c->pw.xf = c->alloc<float>(nc); c->pw.pf2 = c->alloc<float>(nc);
        // reduction segments: p.q, r.r, spare, one r.z segment per colour (1024 blocks each)
        c->pw.red = c->alloc<double>((size_t)(P.ncol + 3) * 1024 + 64);
        c->pw.scal = c->alloc<double>(64);
        c->pw.cnt = c->alloc<unsigned>(8);
        c->d_b = c->alloc<double>(nc);
        {
            // the per-cell face lists of the gradient, column-major over K slots
            int K = 0;
            for (int i = 0; i < nc; ++i) K = std::max(K, P.gfo[i + 1] - P.gfo[i]);
            const size_t kn = (size_t)K * nc;
            std::vector<int32_t> src(kn, kGradUnused);
            std::vector<double> w(kn, 0.0), val(kn, 0.0), S(3 * kn, 0.0), gk(P.nonorth ? 3 * kn : 0, 0.0);
            for (int i = 0; i < nc; ++i)
                for (int t = P.gfo[i]; t < P.gfo[i + 1]; ++t) {
                    const size_t e = (size_t)(t - P.gfo[i]) * nc + i;
                    src[e] = P.g_kind[t] == 0 ? P.g_other[t] : (P.g_kind[t] == 1 ? kGradFixed : kGradSelf);
                    w[e] = P.g_w[t];
                    val[e] = P.g_val[t];
                    for (int d = 0; d < 3; ++d) {
                        S[d * kn + e] = P.g_S[3 * (size_t)t + d];
                        if (P.nonorth) gk[d * kn + e] = P.g_k[3 * (size_t)t + d];
                    }
                }
            GradArgs& g = c->gtab;
            g.n_cells = nc;
            g.K = K;
            g.src = c->upload(src);
            g.w = c->upload(w);
            g.val = c->upload(val);
            g.S = c->upload(S);
            g.gk = P.nonorth ? c->upload(gk) : nullptr;
            g.V = c->d_V;
            if (!g.src || !g.w || !g.val || !g.S || (P.nonorth && !g.gk))
                return c->fail(PIC_ERR_HIP, "out of device memory (poisson)");
        }
        if (P.nonorth) {
            c->d_mgrad = c->alloc<double>(3 * (size_t)nc);
            c->d_corr = c->alloc<double>(nc);
            if (!c->d_mgrad || !c->d_corr) return c->fail(PIC_ERR_HIP, "out of device memory (poisson)");
        }
        if (!c->pw.scal) return c->fail(PIC_ERR_HIP, "out of device memory (poisson)");
    }
    if ((r = check_hip(c.get(), hipStreamSynchronize(c->stream), "create"))) return r;
    *out = c.release();
    return PIC_OK;
}

int pic_mesh_sizes(const pic_ctx* c, int64_t* nt, int32_t* nc, int32_t* np) {
    if (!c) return PIC_ERR_INVALID;
    if (nt) *nt = (int64_t)c->hm.tet_cell.size();
    if (nc) *nc = c->hm.n_cells;
    if (np) *np = c->hm.n_points;
    return PIC_OK;
}

int pic_mesh_tets(const pic_ctx* c, int32_t* tc, int32_t* tf, int32_t* tp) {
    if (!c) return PIC_ERR_INVALID;
    const size_t n = c->hm.tet_cell.size();
    if (tc) std::memcpy(tc, c->hm.tet_cell.data(), n * sizeof(int32_t));
    if (tf) std::memcpy(tf, c->hm.tet_face.data(), n * sizeof(int32_t));
    if (tp) std::memcpy(tp, c->hm.tet_pt.data(), n * sizeof(int32_t));
    return PIC_OK;
}

int pic_mesh_geometry(const pic_ctx* c, double* cc, double* cv) {
    if (!c) return PIC_ERR_INVALID;
    if (cc) std::memcpy(cc, c->hm.C.data(), c->hm.C.size() * sizeof(double));
    if (cv) std::memcpy(cv, c->hm.V.data(), c->hm.V.size() * sizeof(double));
    return PIC_OK;
}

int pic_set_cell_field(pic_ctx* c, int32_t field, const double* v) {
    if (!c || !v) return PIC_ERR_INVALID;
    const int nc = c->hm.n_cells;
    switch (field) {
        case PIC_FIELD_E: {
            hipMemcpyAsync(c->d_E, v, sizeof(double) * 3 * nc, hipMemcpyHostToDevice, c->stream);
            bool any = false;
            for (size_t i = 0; i < 3 * (size_t)nc; ++i) if (v[i] != 0.0) { any = true; break; }
            c->e_zero = !any;
            if (int r = update_field_derived(c)) return r;
            break;
        }
        case PIC_FIELD_B: {
            hipMemcpyAsync(c->d_B, v, sizeof(double) * 3 * nc, hipMemcpyHostToDevice, c->stream);
            std::vector<float> f(3 * (size_t)nc);
            bool any = false;
            for (size_t i = 0; i < f.size(); ++i) { f[i] = (float)v[i]; if (v[i] != 0.0) any = true; }
            hipMemcpyAsync(c->d_Bf, f.data(), sizeof(float) * f.size(), hipMemcpyHostToDevice, c->stream);
            // BorisPusher.C:42-55: the rotation is skipped when B is zero everywhere
            c->rotate = any || c->pusher == kPushVay;
            hipStreamSynchronize(c->stream);
            break;
        }
        case PIC_FIELD_PHI:
            hipMemcpyAsync(c->d_phi, v, sizeof(double) * nc, hipMemcpyHostToDevice, c->stream);
            break;
        case PIC_FIELD_SPACE_CHARGE:
            hipMemcpyAsync(c->d_space, v, sizeof(double) * nc, hipMemcpyHostToDevice, c->stream);
            c->has_space = true;
            break;
        case PIC_FIELD_MCC_SIGMA_CR_MAX:
        case PIC_FIELD_MCC_REMAINDER:
            if (!c->mcc) return c->fail(PIC_ERR_INVALID, "no electron-neutral collision model");
            hipMemcpyAsync(field == PIC_FIELD_MCC_REMAINDER ? c->d_rem : c->d_smax, v, sizeof(double) * nc,
                           hipMemcpyHostToDevice, c->stream);
            break;
        case PIC_FIELD_ION_SIGMA_CR_MAX:
        case PIC_FIELD_ION_REMAINDER:
            if (!c->ion_mcc) return c->fail(PIC_ERR_INVALID, "no ion-neutral collision model");
            hipMemcpyAsync(field == PIC_FIELD_ION_REMAINDER ? c->d_rem_i : c->d_smax_i, v, sizeof(double) * nc,
                           hipMemcpyHostToDevice, c->stream);
            break;
        case PIC_FIELD_ION_PARCEL_SIGMA_CR_MAX:
        case PIC_FIELD_ION_PARCEL_REMAINDER:
            if (!c->ion_mcc) return c->fail(PIC_ERR_INVALID, "no ion-neutral collision model");
            hipMemcpyAsync(field == PIC_FIELD_ION_PARCEL_REMAINDER ? c->d_rem_ip : c->d_smax_ip, v, sizeof(double) * nc,
                           hipMemcpyHostToDevice, c->stream);
            break;
        case PIC_FIELD_BINARY_SIGMA_CR_MAX:
        case PIC_FIELD_BINARY_REMAINDER:
        case PIC_FIELD_BINARY_BG_REMAINDER:
            if (!c->binary) return c->fail(PIC_ERR_INVALID, "no binary collision model");
            hipMemcpyAsync(field == PIC_FIELD_BINARY_SIGMA_CR_MAX ? c->d_smax_n
                                                                  : (field == PIC_FIELD_BINARY_REMAINDER ? c->d_rem_n : c->d_rem_nb),
                           v, sizeof(double) * nc, hipMemcpyHostToDevice, c->stream);
            break;
        default:
            return c->fail(PIC_ERR_INVALID, "field cannot be set");
    }
    return check_hip(c, hipStreamSynchronize(c->stream), "set field");
}

int pic_get_cell_field(pic_ctx* c, int32_t field, double* v) {
    if (!c || !v) return PIC_ERR_INVALID;
    const int nc = c->hm.n_cells;
    const double* src = nullptr;
    size_t n = 0;
    switch (field) {
        case PIC_FIELD_E: src = c->d_E; n = 3 * (size_t)nc; break;
        case PIC_FIELD_B: src = c->d_B; n = 3 * (size_t)nc; break;
        case PIC_FIELD_PHI: src = c->d_phi; n = nc; break;
        case PIC_FIELD_RHO: src = c->d_rho; n = nc; break;
        case PIC_FIELD_SPACE_CHARGE: src = c->d_space; n = nc; break;
        case PIC_FIELD_E_POINT: src = c->d_Ep; n = 3 * (size_t)c->hm.n_points; break;
        case PIC_FIELD_MCC_SIGMA_CR_MAX:
        case PIC_FIELD_MCC_REMAINDER:
            if (!c->mcc) return c->fail(PIC_ERR_INVALID, "no electron-neutral collision model");
            src = field == PIC_FIELD_MCC_REMAINDER ? c->d_rem : c->d_smax;
            n = nc;
            break;
        case PIC_FIELD_ION_SIGMA_CR_MAX:
        case PIC_FIELD_ION_REMAINDER:
            if (!c->ion_mcc) return c->fail(PIC_ERR_INVALID, "no ion-neutral collision model");
            src = field == PIC_FIELD_ION_REMAINDER ? c->d_rem_i : c->d_smax_i;
            n = nc;
            break;
        case PIC_FIELD_ION_PARCEL_SIGMA_CR_MAX:
        case PIC_FIELD_ION_PARCEL_REMAINDER:
            if (!c->ion_mcc) return c->fail(PIC_ERR_INVALID, "no ion-neutral collision model");
            src = field == PIC_FIELD_ION_PARCEL_REMAINDER ? c->d_rem_ip : c->d_smax_ip;
            n = nc;
            break;
        case PIC_FIELD_BINARY_SIGMA_CR_MAX:
        case PIC_FIELD_BINARY_REMAINDER:
        case PIC_FIELD_BINARY_BG_REMAINDER:
            if (!c->binary) return c->fail(PIC_ERR_INVALID, "no binary collision model");
            src = field == PIC_FIELD_BINARY_SIGMA_CR_MAX ? c->d_smax_n
                                                         : (field == PIC_FIELD_BINARY_REMAINDER ? c->d_rem_n : c->d_rem_nb);
            n = nc;
            break;
        default: return c->fail(PIC_ERR_INVALID, "unknown field");
    }
    hipMemcpyAsync(v, src, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream);
    return check_hip(c, hipStreamSynchronize(c->stream), "get field");
}

int pic_set_cross_section(pic_ctx* c, int32_t process, int32_t n, const double* e, const double* s, double th) {
    if (!c || process < 0 || process >= pic_ctx::kXsProcs || n < 0 || (n > 0 && (!e || !s))) return PIC_ERR_INVALID;
    for (int i = 1; i < n; ++i)
        if (!(e[i] > e[i - 1])) return c->fail(PIC_ERR_INVALID, "cross-section energies must be strictly ascending");
    c->xs_he[process].assign(e, e + n);
    c->xs_hs[process].assign(s, s + n);
    c->xs_n[process] = n;
    c->xs_th[process] = th;
    c->d_xs_e[process] = c->upload(c->xs_he[process]);
    c->d_xs_s[process] = c->upload(c->xs_hs[process]);
    c->d_xs_guide[process] = nullptr;
    if (n >= 2 && e[0] > 0.0) {
        std::vector<int> g(kXsGuide);
        xs_guide(e, n, g.data(), &c->xs_le0[process], &c->xs_idl[process]);
        c->d_xs_guide[process] = c->upload(g);
    }
    return check_hip(c, hipStreamSynchronize(c->stream), "set cross section");
}

namespace {
// ListCrossSection interpolateXY on the host tables (clamped at both ends)
double xs_host(const pic_ctx* c, int k, double x) {
    const int n = c->xs_n[k];
    if (n == 0) return 0.0;
    const double* xs = c->xs_he[k].data();
    const double* ys = c->xs_hs[k].data();
    if (x <= xs[0]) return ys[0];
    if (x >= xs[n - 1]) return ys[n - 1];
    int lo = (int)(std::upper_bound(xs, xs + n, x) - xs) - 1;
    if (xs[lo] == x) return ys[lo];
    return ys[lo] + ((x - xs[lo]) / (xs[lo + 1] - xs[lo])) * (ys[lo + 1] - ys[lo]);
}

// (sigma cR)max of a collision model's processes [p0, p1) for a parcel of mass m:
// FindMaximum (the largest sqrt(2 e E / m) Qt(E) on E = 0, step, ... below emax) or
// Temperature (Qt(kT/e) times the most probable speed sqrt(2 k T / m), 1 eV when T is zero)
// (ElectronNeutralCollisionModel.C:370-482, IonNeutralCollisionModel.C:256-379,
// PICCloudI.H:541-548)
double smax_find(const pic_ctx* c, int p0, int p1, double m, double emax, double estep) {
    double smax = 0.0;
    for (double eV = 0.0; eV < emax; eV += estep) {
        double q = 0.0;
        for (int k = p0; k < p1; ++k) q += xs_host(c, k, eV);
        smax = std::max(smax, std::sqrt(2.0 * kQe * eV / m) * q);
    }
    return smax;
}
double smax_temperature(const pic_ctx* c, int p0, int p1, double m, double T) {
    if (T < 1e-300) T = 11604.52;
    const double eV = T * kKB / kQe;
    double q = 0.0;
    for (int k = p0; k < p1; ++k) q += xs_host(c, k, eV);
    return q * std::sqrt(2.0 * kKB * T / m);
}
int fill_cells(pic_ctx* c, double* d, double v, const char* what) {
    std::vector<double> h(c->hm.n_cells, v);
    hipMemcpyAsync(d, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice, c->stream);
    return check_hip(c, hipStreamSynchronize(c->stream), what);
}
}  // namespace

int pic_mcc_initialize(pic_ctx* c, double emax, double estep) {
    if (!c || !c->mcc) return c ? c->fail(PIC_ERR_INVALID, "no electron-neutral model") : PIC_ERR_INVALID;
    if (!(estep > 0)) return c->fail(PIC_ERR_INVALID, "energy step");
    const double smax = smax_find(c, PIC_XS_ELASTIC, PIC_XS_IONIZATION + 1, c->sp[c->electron].mass, emax, estep);
    return fill_cells(c, c->d_smax, smax, "mcc initialize");
}

// the parcel field and the background field from one value: the model has one neutral and
// one ion species, so both searches run over the same tables with the same ion mass
// (IonNeutralCollisionModel.C:283-330, :331-379)
int pic_ion_mcc_initialize(pic_ctx* c, double emax, double estep) {
    if (!c || !c->ion_mcc) return c ? c->fail(PIC_ERR_INVALID, "no ion-neutral model") : PIC_ERR_INVALID;
    if (!(estep > 0)) return c->fail(PIC_ERR_INVALID, "energy step");
    const int ion = ion_species_of(c);
    if (ion < 0) return c->fail(PIC_ERR_INVALID, "no ion species");
    const double smax = smax_find(c, PIC_XS_ION_ELASTIC, PIC_XS_ION_CHARGE_EXCHANGE + 1, c->sp[ion].mass, emax, estep);
    if (int r = fill_cells(c, c->d_smax_ip, smax, "ion-neutral initialize")) return r;
    return fill_cells(c, c->d_smax_i, smax, "ion-neutral initialize");
}

int pic_mcc_initialize_temperature(pic_ctx* c, int32_t model, double temperature) {
    if (!c || (model != 0 && model != 1)) return PIC_ERR_INVALID;
    if (!(temperature >= 0.0)) return c->fail(PIC_ERR_INVALID, "temperature");
    if (model == 0) {
        if (!c->mcc) return c->fail(PIC_ERR_INVALID, "no electron-neutral model");
        const double smax = smax_temperature(c, PIC_XS_ELASTIC, PIC_XS_IONIZATION + 1, c->sp[c->electron].mass,
                                             temperature);
        return fill_cells(c, c->d_smax, smax, "mcc initialize");
    }
    if (!c->ion_mcc) return c->fail(PIC_ERR_INVALID, "no ion-neutral model");
    const int ion = ion_species_of(c);
    if (ion < 0) return c->fail(PIC_ERR_INVALID, "no ion species");
    const double smax = smax_temperature(c, PIC_XS_ION_ELASTIC, PIC_XS_ION_CHARGE_EXCHANGE + 1, c->sp[ion].mass,
                                         temperature);
    if (int r = fill_cells(c, c->d_smax_ip, smax, "ion-neutral initialize")) return r;
    return fill_cells(c, c->d_smax_i, smax, "ion-neutral initialize");
}

// BinaryCollisionModel::initialize (BinaryCollisionModel.C:540-600)
int pic_binary_initialize(pic_ctx* c, const double* T, const double* n) {
    if (!c || !T || !n) return PIC_ERR_INVALID;
    if (!c->binary) return c->fail(PIC_ERR_INVALID, "no binary collision model");
    int best = -1;
    double dens = 0.0;
    const double bgn = c->bg ? c->bg_n : 0.0;
    for (int s = 0; s < (int)c->sp.size(); ++s) {
        if (c->sp[s].charge != 0.0) continue;
        if (best < 0) best = s;   // the first neutral species when none is denser than zero
        double d = n[s];
        if (c->bg && c->bg_species == s) d += bgn;
        if (d > dens) { dens = d; best = s; }
    }
    if (best < 0) return PIC_OK;   // no neutral species: the reference warns and leaves the field
    double t = T[best];
    if (c->bg && best == c->bg_species && c->bg_T > t) t = c->bg_T;
    if (t < 1e-300) t = 293.15;
    const Species& S = c->sp[best];
    const double smax = kPi * S.diameter * S.diameter * std::sqrt(2.0 * kKB * t / S.mass);
    return fill_cells(c, c->d_smax_n, smax, "binary initialize");
}

int pic_particles_set(pic_ctx* c, int32_t s, int64_t n, const double* bary, const int32_t* tet, const double* U,
                      const uint32_t* id) {
    if (!c || s < 0 || s >= (int)c->sp.size() || n < 0 || (n > 0 && (!bary || !tet || !U || !id))) return PIC_ERR_INVALID;
    Species& S = c->sp[s];
    if (n > S.capacity) return c->fail(PIC_ERR_CAPACITY, "more parcels than the species capacity");
    const int64_t nt = (int64_t)c->hm.tet_cell.size();
    for (int64_t i = 0; i < n; ++i)
        if (tet[i] < -1 || tet[i] >= nt) return c->fail(PIC_ERR_INVALID, "tet index out of range");
    std::vector<double> comp(n);
    for (int k = 0; k < 3; ++k) {
        for (int64_t i = 0; i < n; ++i) comp[i] = bary[3 * i + k];
        hipMemcpy(S.y[S.cur][k], comp.data(), sizeof(double) * n, hipMemcpyHostToDevice);
    }
    for (int k = 0; k < 3; ++k) {
        if (c->mixed()) {
            std::vector<float> f(n);
            for (int64_t i = 0; i < n; ++i) f[i] = (float)U[3 * i + k];
            hipMemcpy(S.u[S.cur][k], f.data(), sizeof(float) * n, hipMemcpyHostToDevice);
        } else {
            for (int64_t i = 0; i < n; ++i) comp[i] = U[3 * i + k];
            hipMemcpy(S.u[S.cur][k], comp.data(), sizeof(double) * n, hipMemcpyHostToDevice);
        }
    }
    hipMemcpy(S.tet[S.cur], tet, sizeof(int32_t) * n, hipMemcpyHostToDevice);
    hipMemcpy(S.id[S.cur], id, sizeof(uint32_t) * n, hipMemcpyHostToDevice);
    S.n = n;
    set_dev_n(c, s);
    for (auto& S : c->sp) S.indexed = false;
    c->counts_valid = false;
    return check_hip(c, hipStreamSynchronize(c->stream), "particles set");
}

int pic_particles_set_positions(pic_ctx* c, int32_t s, int64_t n, const double* x, const int32_t* cell,
                                const double* U, const uint32_t* id) {
    if (!c || s < 0 || s >= (int)c->sp.size() || n < 0 || (n > 0 && (!x || !cell || !U || !id))) return PIC_ERR_INVALID;
    Species& S = c->sp[s];
    if (n > S.capacity) return c->fail(PIC_ERR_CAPACITY, "more parcels than the species capacity");
    for (int64_t i = 0; i < n; ++i)
        if (cell[i] < 0 || cell[i] >= c->hm.n_cells) return c->fail(PIC_ERR_INVALID, "cell index out of range");
    double* dx = nullptr; int32_t* dc = nullptr; double* dU = nullptr; uint32_t* did = nullptr; unsigned* dfail = nullptr;
    const size_t nn = std::max<int64_t>(n, 1);
    if (hipMalloc(&dx, nn * 3 * sizeof(double)) || hipMalloc(&dc, nn * sizeof(int32_t)) ||
        hipMalloc(&dU, nn * 3 * sizeof(double)) || hipMalloc(&did, nn * sizeof(uint32_t)) ||
        hipMalloc(&dfail, sizeof(unsigned)))
        return c->fail(PIC_ERR_HIP, "out of device memory (locate)");
    hipMemcpy(dx, x, n * 3 * sizeof(double), hipMemcpyHostToDevice);
    hipMemcpy(dc, cell, n * sizeof(int32_t), hipMemcpyHostToDevice);
    hipMemcpy(dU, U, n * 3 * sizeof(double), hipMemcpyHostToDevice);
    hipMemcpy(did, id, n * sizeof(uint32_t), hipMemcpyHostToDevice);
    hipMemset(dfail, 0, sizeof(unsigned));
    hipError_t e;
    if (c->mixed()) {
        LocateArgs<float> a{soa<float>(S, S.cur), n, 0, dx, dc, dU, did, c->d_tets, c->d_cell_tet0, c->d_C, dfail};
        e = launch_locate(a, c->stream);
    } else {
        LocateArgs<double> a{soa<double>(S, S.cur), n, 0, dx, dc, dU, did, c->d_tets, c->d_cell_tet0, c->d_C, dfail};
        e = launch_locate(a, c->stream);
    }
    unsigned fails = 0;
    hipStreamSynchronize(c->stream);
    hipMemcpy(&fails, dfail, sizeof(unsigned), hipMemcpyDeviceToHost);
    hipFree(dx); hipFree(dc); hipFree(dU); hipFree(did); hipFree(dfail);
    if (int r = check_hip(c, e, "locate")) return r;
    S.n = n;
    set_dev_n(c, s);
    for (auto& S : c->sp) S.indexed = false;
    c->counts_valid = false;
    if (fails) return c->fail(PIC_ERR_INVALID, std::to_string(fails) + " positions lie outside their given cell");
    return check_hip(c, hipStreamSynchronize(c->stream), "particles set positions");
}

int pic_particles_count(pic_ctx* c, int32_t s, int64_t* n) {
    if (!c || s < 0 || s >= (int)c->sp.size() || !n) return PIC_ERR_INVALID;
    *n = c->sp[s].n;
    return PIC_OK;
}

int pic_particles_get(pic_ctx* c, int32_t s, int64_t cap, int64_t* nout, double* bary, int32_t* tet, double* U,
                      uint32_t* id) {
    if (!c || s < 0 || s >= (int)c->sp.size()) return PIC_ERR_INVALID;
    if (int r = check_hip(c, hipStreamSynchronize(c->stream), "particles get")) return r;
    Species& S = c->sp[s];
    const int64_t n = S.n;
    if (nout) *nout = n;
    if (n > cap) return c->fail(PIC_ERR_CAPACITY, "output buffer too small");
    std::vector<double> comp(n);
    if (bary)
        for (int k = 0; k < 3; ++k) {
            hipMemcpy(comp.data(), S.y[S.cur][k], sizeof(double) * n, hipMemcpyDeviceToHost);
            for (int64_t i = 0; i < n; ++i) bary[3 * i + k] = comp[i];
        }
    if (U)
        for (int k = 0; k < 3; ++k) {
            if (c->mixed()) {
                std::vector<float> f(n);
                hipMemcpy(f.data(), S.u[S.cur][k], sizeof(float) * n, hipMemcpyDeviceToHost);
                for (int64_t i = 0; i < n; ++i) U[3 * i + k] = f[i];
            } else {
                hipMemcpy(comp.data(), S.u[S.cur][k], sizeof(double) * n, hipMemcpyDeviceToHost);
                for (int64_t i = 0; i < n; ++i) U[3 * i + k] = comp[i];
            }
        }
    if (tet) hipMemcpy(tet, S.tet[S.cur], sizeof(int32_t) * n, hipMemcpyDeviceToHost);
    if (id) hipMemcpy(id, S.id[S.cur], sizeof(uint32_t) * n, hipMemcpyDeviceToHost);
    return check_hip(c, hipGetLastError(), "particles get");
}

int pic_particles_positions(pic_ctx* c, int32_t s, int64_t cap, double* x) {
    if (!c || s < 0 || s >= (int)c->sp.size() || !x) return PIC_ERR_INVALID;
    Species& S = c->sp[s];
    if (S.n > cap) return c->fail(PIC_ERR_CAPACITY, "output buffer too small");
    double* dx = nullptr;
    if (hipMalloc(&dx, std::max<int64_t>(S.n, 1) * 3 * sizeof(double))) return c->fail(PIC_ERR_HIP, "alloc");
    hipError_t e = c->mixed() ? launch_positions(soa<float>(S, S.cur), S.n, c->d_tets, c->d_C, c->d_points, dx, c->stream)
                              : launch_positions(soa<double>(S, S.cur), S.n, c->d_tets, c->d_C, c->d_points, dx, c->stream);
    hipStreamSynchronize(c->stream);
    hipMemcpy(x, dx, S.n * 3 * sizeof(double), hipMemcpyDeviceToHost);
    hipFree(dx);
    return check_hip(c, e, "positions");
}

int pic_particles_positions_noalloc(pic_ctx* c, int32_t s, int64_t cap, double* x) {
    if (!c || s < 0 || s >= (int)c->sp.size() || !x) return PIC_ERR_INVALID;
    Species& S = c->sp[s];
    if (S.n > cap) return c->fail(PIC_ERR_CAPACITY, "output buffer too small");
    const size_t bytes = (size_t)S.n * 3 * sizeof(double);
    const size_t most = noalloc_scratch_bytes() / (3 * sizeof(double));
    if (bytes > noalloc_scratch_bytes())
        return c->fail(PIC_ERR_CAPACITY, std::to_string(S.n) + " parcels exceed the " + std::to_string(most) +
                                             " the allocation-free positions call stages at once");
    void* d = noalloc_scratch(c->device);
    if (!d) return c->fail(PIC_ERR_HIP, "device scratch was not set up for this device");
    if (S.n == 0) return PIC_OK;
    hipError_t e = c->mixed() ? launch_positions(soa<float>(S, S.cur), S.n, c->d_tets, c->d_C, c->d_points, (double*)d, c->stream)
                              : launch_positions(soa<double>(S, S.cur), S.n, c->d_tets, c->d_C, c->d_points, (double*)d, c->stream);
    if (int r = check_hip(c, e, "positions")) return r;
    if (int r = check_hip(c, hipMemcpyAsync(x, d, bytes, hipMemcpyDeviceToHost, c->stream), "positions copy")) return r;
    return check_hip(c, hipStreamSynchronize(c->stream), "positions");
}

int pic_init_maxwellian(pic_ctx* c, int32_t s, int32_t nppc, double T, const double* drift, uint32_t id0,
                        uint64_t seed) {
    if (!c || s < 0 || s >= (int)c->sp.size() || nppc < 0) return PIC_ERR_INVALID;
    Species& S = c->sp[s];
    const int64_t add = (int64_t)nppc * c->hm.n_cells;
    if (S.n + add > S.capacity) return c->fail(PIC_ERR_CAPACITY, "more parcels than the species capacity");
    hipError_t e;
    if (c->mixed()) {
        InitArgs<float> a{soa<float>(S, S.cur), add, S.n, nppc, c->d_tets, c->d_cell_tet0, std::sqrt(kKB * T / S.mass),
                          {drift ? drift[0] : 0.0, drift ? drift[1] : 0.0, drift ? drift[2] : 0.0}, id0, seed};
        e = launch_init(a, c->stream);
    } else {
        InitArgs<double> a{soa<double>(S, S.cur), add, S.n, nppc, c->d_tets, c->d_cell_tet0, std::sqrt(kKB * T / S.mass),
                           {drift ? drift[0] : 0.0, drift ? drift[1] : 0.0, drift ? drift[2] : 0.0}, id0, seed};
        e = launch_init(a, c->stream);
    }
    if (int r = check_hip(c, e, "init")) return r;
    S.n += add;
    set_dev_n(c, s);
    for (auto& S : c->sp) S.indexed = false;
    c->counts_valid = false;
    return check_hip(c, hipStreamSynchronize(c->stream), "init");
}

int pic_push(pic_ctx* c, double dt) {
    if (!c) return PIC_ERR_INVALID;
    for (int s = 0; s < (int)c->sp.size(); ++s) {
        if (!c->sp[s].solve_movement || c->sp[s].charge == 0.0) continue;
        hipError_t e = c->mixed() ? launch_push(move_args<float>(c, s, dt), c->pusher, c->fw, c->stream)
                                  : launch_push(move_args<double>(c, s, dt), c->pusher, c->fw, c->stream);
        if (int r = check_hip(c, e, "push")) return r;
    }
    return sync_state(c);
}

int pic_move(pic_ctx* c, double dt) {
    if (!c) return PIC_ERR_INVALID;
    if (int r = do_move(c, dt)) return r;
    return sync_state(c);
}

int pic_sort(pic_ctx* c) {
    if (!c) return PIC_ERR_INVALID;
    if (int r = do_sort(c)) return r;
    return sync_state(c);
}

int pic_collide_coulomb(pic_ctx* c, double dt) {
    if (!c || !c->coulomb) return c ? c->fail(PIC_ERR_INVALID, "CoulombCollisionModel is none") : PIC_ERR_INVALID;
    if (int r = do_sort(c)) return r;
    if (int r = do_coulomb(c, dt)) return r;
    return sync_state(c);
}

int pic_collide_ion_neutral(pic_ctx* c, double dt) {
    if (!c || !c->ion_mcc) return c ? c->fail(PIC_ERR_INVALID, "IonNeutralCollisionModel is none") : PIC_ERR_INVALID;
    if (int r = do_sort(c)) return r;
    if (int r = do_ion_mcc(c, dt)) return r;
    return sync_state(c);
}

int pic_collide_binary(pic_ctx* c, double dt) {
    if (!c || !c->binary) return c ? c->fail(PIC_ERR_INVALID, "BinaryCollisionModel is none") : PIC_ERR_INVALID;
    if (int r = do_sort(c)) return r;
    if (int r = do_binary(c, dt)) return r;
    return sync_state(c);
}

int pic_set_injection(pic_ctx* c, const pic_injection* models, int32_t n) {
    if (!c || n < 0 || n > 4 || (n > 0 && !models)) return c ? c->fail(PIC_ERR_INVALID, "pic_set_injection: 0-4 models") : PIC_ERR_INVALID;
    const HostMesh& m = c->hm;
    const int ns = (int)c->sp.size();
    std::vector<pic_ctx::Injector> out(n);
    for (int k = 0; k < n; ++k) {
        const pic_injection& q = models[k];
        const std::string tag = "SimpleInjection " + std::to_string(k) + ": ";
        if (q.patch < 0 || q.patch >= m.n_patches) return c->fail(PIC_ERR_INVALID, tag + "no such patch");
        // SimpleInjection.C:99-104: the injector must be a plain `patch`
        if (m.ptype[q.patch] != PIC_PATCH_PATCH) return c->fail(PIC_ERR_INVALID, tag + "the patch is not of type patch");
        if (m.psize[q.patch] <= 0) return c->fail(PIC_ERR_INVALID, tag + "the patch has no faces");
        if (q.velocity_model < 0 || q.velocity_model > 3)
            return c->fail(PIC_ERR_UNKNOWN_MODEL, tag + "valid injection velocity models are: birdMaxwellianFlux mostProbableSpeed maxwellianFlux halfMaxwellianFlux");
        if (q.n_species < 1 || q.n_species > 8) return c->fail(PIC_ERR_INVALID, tag + "1-8 injectionTypeIds");
        for (int e = 0; e < q.n_species; ++e) {
            if (q.species[e] < 0 || q.species[e] >= ns) return c->fail(PIC_ERR_INVALID, tag + "species not defined in cloud");
            if (!(q.frequency[e] >= 0.0) || !(q.eV[e] >= 0.0)) return c->fail(PIC_ERR_INVALID, tag + "frequencies and eV must be >= 0");
        }
        for (int s = 0; s < ns; ++s) {
            if (q.interaction[s] < 0 || q.interaction[s] > 3) return c->fail(PIC_ERR_INVALID, tag + "interaction 0-3");
            if (q.interaction[s] == 3)
                return c->fail(PIC_ERR_UNSUPPORTED, tag + "the floating interaction needs the circuit boundary condition, which is not built");
        }
        out[k].cfg = q;
    }
    // the patch's faces fanned from their first point (ParticleEmitter.C:78-131 triangulates
    // them; a fan has the same areas for the planar convex faces of a mesh)
    hipStreamSynchronize(c->stream);
    for (auto& I : c->inj) { c->release(I.d_tri); c->release(I.d_cum); }
    c->inj.clear();
    for (int k = 0; k < n; ++k) {
        auto& I = out[k];
        const int p = I.cfg.patch;
        std::vector<InjectTri> tri;
        std::vector<double> cum(1, 0.0);
        for (int f = m.pstart[p]; f < m.pstart[p] + m.psize[p]; ++f) {
            const int f0 = m.fo[f], nfp = m.fo[f + 1] - f0;
            const double* Sf = &m.Sf[3 * f];
            const double mS = std::sqrt(Sf[0] * Sf[0] + Sf[1] * Sf[1] + Sf[2] * Sf[2]);
            const int own = m.owner[f];
            const double* P0 = &m.points[3 * m.fp[f0]];
            double t1[3], t2[3], nin[3];
            for (int d = 0; d < 3; ++d) { t1[d] = m.Cf[3 * f + d] - P0[d]; nin[d] = -Sf[d] / mS; }
            const double it1 = 1.0 / std::sqrt(t1[0] * t1[0] + t1[1] * t1[1] + t1[2] * t1[2]);
            for (double& v : t1) v *= it1;
            t2[0] = nin[1] * t1[2] - nin[2] * t1[1];
            t2[1] = nin[2] * t1[0] - nin[0] * t1[2];
            t2[2] = nin[0] * t1[1] - nin[1] * t1[0];
            const double it2 = 1.0 / std::sqrt(t2[0] * t2[0] + t2[1] * t2[1] + t2[2] * t2[2]);
            for (double& v : t2) v *= it2;
            for (int j = 1; j + 1 < nfp; ++j) {
                InjectTri T{};
                const double* B = &m.points[3 * m.fp[f0 + j]];
                const double* Cc = &m.points[3 * m.fp[f0 + j + 1]];
                double e1[3], e2[3];
                for (int d = 0; d < 3; ++d) {
                    T.a[d] = P0[d]; T.b[d] = B[d]; T.c[d] = Cc[d];
                    T.n[d] = Sf[d] / mS; T.t1[d] = t1[d]; T.t2[d] = t2[d];
                    T.pc[d] = m.C[3 * own + d];
                    e1[d] = B[d] - P0[d]; e2[d] = Cc[d] - P0[d];
                }
                T.cell = own;
                const double cx = e1[1] * e2[2] - e1[2] * e2[1], cy = e1[2] * e2[0] - e1[0] * e2[2],
                             cz = e1[0] * e2[1] - e1[1] * e2[0];
                cum.push_back(cum.back() + 0.5 * std::sqrt(cx * cx + cy * cy + cz * cz));
                tri.push_back(T);
            }
        }
        if (tri.empty()) return c->fail(PIC_ERR_MESH, "SimpleInjection: the patch has no triangles");
        I.ntri = (int)tri.size();
        I.d_tri = c->upload(tri);
        I.d_cum = c->upload(cum);
        if (!I.d_tri || !I.d_cum) return c->fail(PIC_ERR_HIP, "out of device memory (injection)");
        c->inj.push_back(I);
    }
    if (n > 0) c->inj_time = models[0].time;
    // the interactions on the injector patches (SimpleInjection.C:233-292): 1 reflection
    // (kind 2), 2 reflection with useInitalVelocity (kind 3, or 2 for a species the model
    // does not inject); the wall queue and tables are made when the mesh has no walls
    bool any = false;
    for (auto& I : c->inj)
        for (int s = 0; s < ns; ++s) any = any || I.cfg.interaction[s] != 0;
    if (any && !c->walls) {
        size_t cap = 0;
        for (auto& S : c->sp) cap += (size_t)S.capacity;
        if (cap > 0xFFFFFFFFull) return c->fail(PIC_ERR_CAPACITY, "wall queue: total capacity exceeds 2^32 parcels");
        c->wall_cap = (unsigned)cap;
        c->d_wall_q = c->alloc<WallHit>(std::max<size_t>(cap, 1));
        c->d_wall_n = c->alloc<unsigned>(1);
        c->h_kind.assign((size_t)kMaxSpecies * m.n_patches, 0);
        c->d_patch_kind = c->upload(c->h_kind);
        c->d_wall_T = c->alloc<double>(m.n_patches);
        c->d_wall_U = c->alloc<double>(3 * (size_t)m.n_patches);
        if (!c->d_wall_q || !c->d_wall_n || !c->d_patch_kind || !c->d_wall_T || !c->d_wall_U)
            return c->fail(PIC_ERR_HIP, "out of device memory (wall queue)");
        c->walls = true;
    }
    if (c->walls) {
        for (int s = 0; s < kMaxSpecies; ++s) c->re[s] = pic_ctx::Reemit{};
        for (int p = 0; p < m.n_patches; ++p)
            if (m.ptype[p] == PIC_PATCH_PATCH)
                for (int s = 0; s < kMaxSpecies; ++s) c->h_kind[(size_t)s * m.n_patches + p] = 0;
        for (auto& I : c->inj)
            for (int s = 0; s < ns; ++s) {
                const int it = I.cfg.interaction[s];
                if (it == 0) continue;
                int e = -1;
                for (int k = 0; k < I.cfg.n_species; ++k)
                    if (I.cfg.species[k] == s) e = k;
                uint8_t kind = (it == 2 && e >= 0) ? 3 : 2;
                if (kind == 3) {
                    if (c->re[s].tri) return c->fail(PIC_ERR_UNSUPPORTED, "SimpleInjection: a species re-emitted by two models");
                    c->re[s].tri = I.d_tri;
                    c->re[s].ntri = I.ntri;
                    c->re[s].vmodel = I.cfg.velocity_model;
                    c->re[s].vt = std::sqrt(kKB * (I.cfg.eV[e] * kQe / kKB) / c->sp[s].mass);
                }
                c->h_kind[(size_t)s * m.n_patches + I.cfg.patch] = kind;
            }
        hipMemcpyAsync(c->d_patch_kind, c->h_kind.data(), c->h_kind.size(), hipMemcpyHostToDevice, c->stream);
    }
    return PIC_OK;
}

int pic_wall_fields(pic_ctx* c, int32_t species, double* out) {
    if (!c || !out || species < -1 || species >= (int32_t)c->sp.size()) return c ? c->fail(PIC_ERR_INVALID, "pic_wall_fields: species") : PIC_ERR_INVALID;
    const int nb = c->hm.n_faces - c->hm.n_internal;
    const int k = species < 0 ? 10 : 9;
    if (!c->d_wall_acc) {
        std::memset(out, 0, sizeof(double) * (size_t)nb * k);
        return PIC_OK;
    }
    std::vector<double> h((size_t)std::max(nb, 1) * c->wall_stride);
    hipMemcpyAsync(h.data(), c->d_wall_acc, sizeof(double) * h.size(), hipMemcpyDeviceToHost, c->stream);
    if (int r = check_hip(c, hipStreamSynchronize(c->stream), "wall fields")) return r;
    const int off = species < 0 ? 0 : 10 + 9 * species;
    for (int f = 0; f < nb; ++f)
        for (int j = 0; j < k; ++j) out[(size_t)f * k + j] = h[(size_t)f * c->wall_stride + off + j];
    return PIC_OK;
}

int pic_inject(pic_ctx* c, double dt) {
    if (!c || c->inj.empty()) return c ? c->fail(PIC_ERR_INVALID, "no SimpleInjection model set") : PIC_ERR_INVALID;
    if (int r = do_inject(c, dt)) return r;
    return sync_state(c);
}

int pic_collide_mcc(pic_ctx* c, double dt) {
    if (!c || !c->mcc) return c ? c->fail(PIC_ERR_INVALID, "ElectronNeutralCollisionModel is none") : PIC_ERR_INVALID;
    if (int r = do_sort(c)) return r;
    if (int r = do_mcc(c, dt)) return r;
    return sync_state(c);
}

int pic_deposit(pic_ctx* c) {
    if (!c) return PIC_ERR_INVALID;
    zero_deposit(c);
    for (int s = 0; s < (int)c->sp.size(); ++s) {
        hipError_t e = c->mixed() ? launch_deposit(move_args<float>(c, s, 0.0), c->cd, c->stream)
                                  : launch_deposit(move_args<double>(c, s, 0.0), c->cd, c->stream);
        if (int r = check_hip(c, e, "deposit")) return r;
    }
    c->counts_valid = true;
    c->place_valid = true;
    if (int r = do_deposit_finish(c)) return r;
    return sync_state(c);
}

int pic_poisson_solve(pic_ctx* c, pic_poisson_stats* st) {
    if (!c) return PIC_ERR_INVALID;
    return do_poisson(c, st);
}

// PICCloud::evolve (PICCloud.C:1426-1471) enqueued on the context stream without a host
// round trip, except inside the Poisson solve (convergence checks) and GNN recording
int step_body(pic_ctx* c, double dt, pic_poisson_stats* ps) {
    if (int r = do_inject(c, dt)) return r;
    if (int r = do_move(c, dt)) return r;
    // a species moves into cell order every Species::sort_interval steps; collision models
    // need the cell order every step, and in between get it as an index
    bool due = c->coulomb || c->mcc || c->ion_mcc || c->binary;
    for (const auto& S : c->sp) due = due || (c->step + 1) % S.sort_interval == 0;
    if (due)
        if (int r = do_sort(c, true)) return r;
    if (c->coulomb)
        if (int r = do_coulomb(c, dt)) return r;
    if (c->mcc)
        if (int r = do_mcc(c, dt)) return r;
    if (c->ion_mcc)
        if (int r = do_ion_mcc(c, dt)) return r;
    if (c->binary)
        if (int r = do_binary(c, dt)) return r;
    if (int r = do_deposit_finish(c)) return r;
    if (c->electrostatic)
        if (int r = do_poisson(c, ps)) return r;
    c->step++;
    return PIC_OK;
}

void read_step_stats(pic_ctx* c, const pic_poisson_stats& ps, pic_step_stats* st) {
    std::vector<unsigned long long> h(kStatsWords);
    hipMemcpy(h.data(), c->d_stats, sizeof(unsigned long long) * kStatsWords, hipMemcpyDeviceToHost);
    unsigned long long ncoll = 0;
    for (int k = 0; k < kNcollSlots; ++k) ncoll += h[kStatsNcoll + k * kNcollStride];
    std::memset(st, 0, sizeof(*st));
    for (size_t s = 0; s < c->sp.size(); ++s) st->n_parcels[s] = c->sp[s].n;
    st->n_deleted = (int64_t)h[0];
    st->n_track_guard = (int64_t)h[1];
    st->n_wall_hits = (int64_t)h[kStatsWallHits];
    st->coulomb_collisions = (int64_t)ncoll;
    unsigned long long mc[4] = {0, 0, 0, 0};
    for (int k = 0; k < kNcollSlots; ++k)
        for (int j = 0; j < 4; ++j) mc[j] += h[kStatsMcc + k * kNcollStride + j];
    st->mcc_candidates = (int64_t)mc[0];
    st->mcc_elastic = (int64_t)mc[1];
    st->mcc_excitation = (int64_t)mc[2];
    st->mcc_ionization = (int64_t)mc[3];
    unsigned long long ic[3] = {0, 0, 0};
    for (int k = 0; k < kNcollSlots; ++k)
        for (int j = 0; j < 3; ++j) ic[j] += h[kStatsMcc + k * kNcollStride + 4 + j];
    st->ion_neutral_candidates = (int64_t)ic[0];
    st->ion_neutral_elastic = (int64_t)ic[1];
    st->ion_neutral_charge_exchange = (int64_t)ic[2];
    unsigned long long pw[7] = {0, 0, 0, 0, 0, 0, 0};
    for (int k = 0; k < kNcollSlots; ++k)
        for (int j = 0; j < 7; ++j) pw[j] += h[kStatsMcc + k * kNcollStride + kPairWordIon + j];
    st->ion_parcel_candidates = (int64_t)pw[0];
    st->ion_parcel_elastic = (int64_t)pw[1];
    st->ion_parcel_charge_exchange = (int64_t)pw[2];
    st->binary_candidates = (int64_t)pw[3];
    st->binary_collisions = (int64_t)pw[4];
    st->binary_bg_candidates = (int64_t)pw[5];
    st->binary_bg_collisions = (int64_t)pw[6];
    st->n_injected = c->n_injected;
    st->poisson = ps;
}

int pic_step(pic_ctx* c, double dt, pic_step_stats* st) {
    if (!c) return PIC_ERR_INVALID;
    c->n_injected = 0;
    hipMemsetAsync(c->d_stats, 0, sizeof(unsigned long long) * kStatsWords, c->stream);
    pic_poisson_stats ps{};
    if (int r = step_body(c, dt, &ps)) return r;
    if (int r = sync_state(c)) return r;
    if (st) read_step_stats(c, ps, st);
    return PIC_OK;
}

int pic_run(pic_ctx* c, double dt, int64_t n_steps, pic_step_stats* st) {
    if (!c || n_steps < 0) return PIC_ERR_INVALID;
    c->n_injected = 0;
    hipMemsetAsync(c->d_stats, 0, sizeof(unsigned long long) * kStatsWords, c->stream);
    pic_poisson_stats ps{};
    for (int64_t k = 0; k < n_steps; ++k)
        if (int r = step_body(c, dt, &ps)) return r;
    if (int r = sync_state(c)) return r;
    if (st) read_step_stats(c, ps, st);
    return PIC_OK;
}

int pic_diag_reduce(pic_ctx* c, pic_diag* out) {
    if (!c || !out) return PIC_ERR_INVALID;
    std::memset(out, 0, sizeof(*out));
    for (int s = 0; s < (int)c->sp.size(); ++s) {
        Species& S = c->sp[s];
        int nb = (int)std::min<int64_t>(1024, std::max<int64_t>(1, (S.n + 255) / 256));
        hipError_t e = c->mixed() ? launch_diag(soa<float>(S, S.cur), S.n, c->d_diag_part, nb, c->d_diag_out, c->stream)
                                  : launch_diag(soa<double>(S, S.cur), S.n, c->d_diag_part, nb, c->d_diag_out, c->stream);
        if (int r = check_hip(c, e, "diag")) return r;
        double h[12];
        hipMemcpyAsync(h, c->d_diag_out, sizeof(h), hipMemcpyDeviceToHost, c->stream);
        hipStreamSynchronize(c->stream);
        const double w = S.weight;
        out->n_parcels[s] = (int64_t)h[0];
        out->sum_w[s] = h[0] * w;
        for (int d = 0; d < 3; ++d) out->sum_wU[s][d] = h[1 + d] * w;
        out->sum_wU2[s] = h[4] * w;
        out->kinetic_energy[s] = 0.5 * S.mass * h[4] * w;
        out->kinetic_energy_relativistic[s] = h[5] * S.mass * kC * kC;
        for (int d = 0; d < 3; ++d) {
            out->sum_wUU[s][d] = h[6 + d] * w;
            out->momentum_relativistic[s][d] = S.mass * h[9 + d] * w;
        }
        out->charge[s] = S.charge * h[0] * w;
    }
    {
        int nb = (int)std::min<int64_t>(1024, std::max<int64_t>(1, (c->hm.n_cells + 255) / 256));
        hipError_t e = launch_field_energy(c->d_E, c->d_V, c->hm.n_cells, c->d_diag_part, nb, c->d_diag_out, c->stream);
        if (int r = check_hip(c, e, "field energy")) return r;
        double fe = 0;
        hipMemcpyAsync(&fe, c->d_diag_out, sizeof(double), hipMemcpyDeviceToHost, c->stream);
        hipStreamSynchronize(c->stream);
        out->field_energy = 0.5 * kEps0 * fe;
    }
    return PIC_OK;
}

// the cell moments of species s into c->d_moments (the arrays in cell order first)
int species_moments(pic_ctx* c, int s) {
    const int nc = c->hm.n_cells;
    Species& S = c->sp[s];
    hipError_t e;
    if (c->mixed()) {
        MomentArgs<float> a{nc, soa<float>(S, S.cur), S.start, c->d_V, S.mass, S.charge, S.weight, 4096, c->d_flags,
                            c->d_moments};
        e = launch_moments(a, c->stream);
    } else {
        MomentArgs<double> a{nc, soa<double>(S, S.cur), S.start, c->d_V, S.mass, S.charge, S.weight, 4096,
                             c->d_flags, c->d_moments};
        e = launch_moments(a, c->stream);
    }
    return check_hip(c, e, "moments");
}

int pic_cell_moments(pic_ctx* c, int32_t species, double* out) {
    if (!c || !out || species < -1 || species >= (int)c->sp.size()) return c ? c->fail(PIC_ERR_INVALID, "species") : PIC_ERR_INVALID;
    const int nc = c->hm.n_cells;
    if (!c->d_moments) c->d_moments = c->alloc<double>((size_t)nc * kMomentVals);
    if (!c->d_moments) return c->fail(PIC_ERR_HIP, "out of device memory (moments)");
    if (int r = do_sort(c)) return r;
    hipMemsetAsync(c->d_moments, 0, sizeof(double) * (size_t)nc * kMomentVals, c->stream);
    for (int s = 0; s < (int)c->sp.size(); ++s) {
        if (species >= 0 && s != species) continue;
        if (int r = species_moments(c, s)) return r;
    }
    hipMemcpyAsync(out, c->d_moments, sizeof(double) * (size_t)nc * kMomentVals, hipMemcpyDeviceToHost, c->stream);
    return sync_state(c);
}

int pic_fields_average(pic_ctx* c, int32_t op) {
    if (!c || (op != 0 && op != 1)) return c ? c->fail(PIC_ERR_INVALID, "op") : PIC_ERR_INVALID;
    const int nc = c->hm.n_cells, ns = (int)c->sp.size();
    const size_t per = (size_t)nc * kFieldAvgVals;
    if (!c->d_favg) {
        c->d_favg = c->alloc<double>(per * ns);
        if (!c->d_favg) return c->fail(PIC_ERR_HIP, "out of device memory (field averages)");
    }
    if (op == 0) {
        hipMemsetAsync(c->d_favg, 0, sizeof(double) * per * ns, c->stream);
        c->n_favg = 0;
        return check_hip(c, hipStreamSynchronize(c->stream), "field averages");
    }
    if (!c->d_moments) c->d_moments = c->alloc<double>((size_t)nc * kMomentVals);
    if (!c->d_moments) return c->fail(PIC_ERR_HIP, "out of device memory (moments)");
    if (int r = do_sort(c)) return r;
    for (int s = 0; s < ns; ++s) {
        hipMemsetAsync(c->d_moments, 0, sizeof(double) * (size_t)nc * kMomentVals, c->stream);
        if (int r = species_moments(c, s)) return r;
        if (int r = check_hip(c, launch_fields_accumulate(c->d_moments, c->d_favg + per * s, nc, c->stream),
                              "field averages"))
            return r;
    }
    c->n_favg++;
    return sync_state(c);
}

// picFields::write (picFields.C:103-211) from the averages: UMean = momentumMean / rhoMMean,
// translationalT = 2 / (3 k rhoNMean) (linearKEMean - 1/2 rhoMMean UMean.UMean), p = k
// rhoNMean translationalT
int pic_fields_mean(pic_ctx* c, int32_t species, double* out, int32_t* derived) {
    if (!c || !out || species < 0 || species >= (int)c->sp.size()) return c ? c->fail(PIC_ERR_INVALID, "species") : PIC_ERR_INVALID;
    if (!c->d_favg || c->n_favg == 0) return c->fail(PIC_ERR_INVALID, "no field averages accumulated");
    const int nc = c->hm.n_cells;
    std::vector<double> h((size_t)nc * kFieldAvgVals);
    hipMemcpyAsync(h.data(), c->d_favg + (size_t)nc * kFieldAvgVals * species, sizeof(double) * h.size(),
                   hipMemcpyDeviceToHost, c->stream);
    if (int r = check_hip(c, hipStreamSynchronize(c->stream), "field averages")) return r;
    const double inv = 1.0 / (double)c->n_favg;
    bool ok = true;
    for (int i = 0; i < nc; ++i) {
        const double* a = &h[(size_t)i * kFieldAvgVals];
        const double rn = a[0] * inv, rm = a[1] * inv, ke = a[2] * inv;
        // picFields.C:132: the derived fields are written only when every cell has both
        ok = ok && std::fabs(rn) > 1e-300 && std::fabs(ke) > 1e-300;
        double* o = out + (size_t)i * 11;
        o[0] = rn; o[1] = rm; o[2] = ke;
        for (int d = 0; d < 3; ++d) o[3 + d] = a[3 + d] * inv;
    }
    for (int i = 0; i < nc; ++i) {
        double* o = out + (size_t)i * 11;
        double uu = 0.0;
        for (int d = 0; d < 3; ++d) {
            o[6 + d] = ok ? o[3 + d] / o[1] : 0.0;
            uu += o[6 + d] * o[6 + d];
        }
        o[9] = ok ? 2.0 / (3.0 * kKB * o[0]) * (o[2] - 0.5 * o[1] * uu) : 0.0;
        o[10] = ok ? kKB * o[0] * o[9] : 0.0;
    }
    if (derived) *derived = ok ? 1 : 0;
    return PIC_OK;
}

int pic_sync(pic_ctx* c) {
    if (!c) return PIC_ERR_INVALID;
    return check_hip(c, hipStreamSynchronize(c->stream), "sync");
}

int64_t pic_step_index(const pic_ctx* c) { return c ? c->step : -1; }
int pic_set_step_index(pic_ctx* c, int64_t s) {
    if (!c) return PIC_ERR_INVALID;
    c->step = s;
    return PIC_OK;
}

int pic_profile_enable(pic_ctx* c, int32_t stages) {
    if (!c) return PIC_ERR_INVALID;
    c->profile = (uint32_t)stages;
    return PIC_OK;
}

int pic_profile_reset(pic_ctx* c) {
    if (!c) return PIC_ERR_INVALID;
    hipStreamSynchronize(c->stream);
    for (auto& pe : c->pending) { c->pool.push_back(pe.a); c->pool.push_back(pe.b); }
    c->pending.clear();
    std::memset(c->stage_ms, 0, sizeof(c->stage_ms));
    std::memset(c->stage_n, 0, sizeof(c->stage_n));
    return PIC_OK;
}

int pic_stage_times(pic_ctx* c, double* ms, int64_t* n) {
    if (!c) return PIC_ERR_INVALID;
    hipStreamSynchronize(c->stream);
    for (auto& pe : c->pending) {
        float t = 0.f;
        hipEventElapsedTime(&t, pe.a, pe.b);
        c->stage_ms[pe.stage] += t;
        c->stage_n[pe.stage] += 1;
        c->pool.push_back(pe.a);
        c->pool.push_back(pe.b);
    }
    c->pending.clear();
    for (int k = 0; k < PIC_STAGE_COUNT; ++k) {
        if (ms) ms[k] = c->stage_ms[k];
        if (n) n[k] = c->stage_n[k];
    }
    return PIC_OK;
}

// ---------------------------------------------------------------- learned initial guess

int pic_gnn_init(pic_ctx* c, int32_t layers, int32_t width, uint64_t seed) {
    if (!c) return PIC_ERR_INVALID;
    if (!c->has_poisson) return c->fail(PIC_ERR_INVALID, "the learned guess needs MaxwellSolver ElectroStatic");
    if (width != 64) return c->fail(PIC_ERR_UNSUPPORTED, "hidden width must be 64 (MFMA tiles)");
    if (layers < 1 || layers > 8) return c->fail(PIC_ERR_INVALID, "layers must be 1..8");
    if (c->gnn) { c->gnn->release(); c->gnn.reset(); }
    c->gnn.reset(new Gnn());
    Gnn& G = *c->gnn;
    G.nb = c->pd.nb;
    if (int r = check_hip(c, G.alloc(c->hm.n_cells, c->pd.K, layers, c->stream), "gnn alloc")) return r;
    double ds = 0.0;
    for (double d : c->hp.diag) ds += d;
    G.dbar = ds / c->hp.diag.size();
    if (!c->d_phi_prev) c->d_phi_prev = c->alloc<double>(c->hm.n_cells);
    // He-style initialisation from a splitmix stream; residual layers start small
    std::vector<float> th(G.np, 0.f);
    uint64_t st = seed * 0x9E3779B97F4A7C15ull + 1;
    auto normal = [&]() {
        auto u = [&]() {
            st += 0x9E3779B97F4A7C15ull;
            uint64_t z = st;
            z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
            z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
            z ^= z >> 31;
            return ((z >> 11) + 0.5) * (1.0 / 9007199254740992.0);
        };
        double a = u(), b = u();
        return std::sqrt(-2.0 * std::log(a)) * std::cos(2.0 * kPi * b);
    };
    size_t o = 0;
    for (int k = 0; k < 192; ++k) th[o++] = (float)(normal() * std::sqrt(2.0 / 3.0));
    o += 64;
    for (int l = 0; l < layers; ++l) {
        for (int k = 0; k < 8192; ++k) th[o++] = (float)(normal() * 0.5 * std::sqrt(2.0 / 128.0));
        o += 64;
    }
    for (int k = 0; k < 64; ++k) th[o++] = (float)(normal() * 0.1 * std::sqrt(1.0 / 64.0));
    hipMemcpyAsync(G.theta, th.data(), sizeof(float) * G.np, hipMemcpyHostToDevice, c->stream);
    G.refresh_weights();
    return check_hip(c, hipStreamSynchronize(c->stream), "gnn init");
}

int pic_gnn_n_params(pic_ctx* c, int64_t* n) {
    if (!c || !c->gnn || !n) return c ? c->fail(PIC_ERR_INVALID, "pic_gnn_init first") : PIC_ERR_INVALID;
    *n = c->gnn->np;
    return PIC_OK;
}

int pic_gnn_params(pic_ctx* c, int64_t n, float* get, const float* set) {
    if (!c || !c->gnn) return c ? c->fail(PIC_ERR_INVALID, "pic_gnn_init first") : PIC_ERR_INVALID;
    Gnn& G = *c->gnn;
    if (n != G.np) return c->fail(PIC_ERR_INVALID, "parameter count mismatch");
    if (set) {
        hipMemcpyAsync(G.theta, set, sizeof(float) * n, hipMemcpyHostToDevice, c->stream);
        G.refresh_weights();
    }
    if (get) hipMemcpyAsync(get, G.theta, sizeof(float) * n, hipMemcpyDeviceToHost, c->stream);
    return check_hip(c, hipStreamSynchronize(c->stream), "gnn params");
}

int pic_gnn_record(pic_ctx* c, int32_t capacity) {
    if (!c || !c->gnn) return c ? c->fail(PIC_ERR_INVALID, "pic_gnn_init first") : PIC_ERR_INVALID;
    Gnn& G = *c->gnn;
    if (capacity <= 0) { c->gnn_rec = false; return PIC_OK; }
    if (capacity != G.snap_cap) {
        const size_t n = (size_t)c->hm.n_cells;
        G.snap_x = c->alloc<float>(3 * n * capacity);
        G.snap_t = c->alloc<float>(n * capacity);
        if (!G.snap_x || !G.snap_t) return c->fail(PIC_ERR_HIP, "out of device memory (snapshots)");
        G.snap_cap = capacity;
        G.snap_count = 0;
        G.snap_next = 0;
        G.snap_norm.assign(capacity, 1.0);
        G.snap_rnorm.assign(capacity, 1.0);
    }
    c->gnn_rec = true;
    return PIC_OK;
}

int pic_gnn_snapshot_count(pic_ctx* c, int32_t* n) {
    if (!c || !c->gnn || !n) return PIC_ERR_INVALID;
    *n = c->gnn->snap_count;
    return PIC_OK;
}

int pic_gnn_snapshot(pic_ctx* c, int32_t idx, float* x, float* t, double* norm) {
    if (!c || !c->gnn || idx < 0 || idx >= c->gnn->snap_count) return c ? c->fail(PIC_ERR_INVALID, "snapshot index") : PIC_ERR_INVALID;
    Gnn& G = *c->gnn;
    const size_t n = (size_t)c->hm.n_cells;
    if (x) hipMemcpyAsync(x, G.snap_x + idx * n * 3, sizeof(float) * 3 * n, hipMemcpyDeviceToHost, c->stream);
    if (t) hipMemcpyAsync(t, G.snap_t + idx * n, sizeof(float) * n, hipMemcpyDeviceToHost, c->stream);
    if (norm) *norm = G.snap_norm[idx];
    return check_hip(c, hipStreamSynchronize(c->stream), "gnn snapshot");
}

int pic_gnn_grad(pic_ctx* c, int32_t idx, float* grad, double* loss) {
    if (!c || !c->gnn || idx < 0 || idx >= c->gnn->snap_count) return c ? c->fail(PIC_ERR_INVALID, "snapshot index") : PIC_ERR_INVALID;
    Gnn& G = *c->gnn;
    const size_t n = (size_t)c->hm.n_cells;
    G.forward(G.snap_x + idx * n * 3, true);
    double l = 0.0;
    G.loss_backward2(G.snap_t + idx * n, G.snap_norm[idx], G.snap_rnorm[idx], c->pd.nb, c->pd.a, c->pd.diag, c->pd.K,
                     &l);
    if (grad) hipMemcpyAsync(grad, G.grad, sizeof(float) * G.np, hipMemcpyDeviceToHost, c->stream);
    if (loss) *loss = l;
    return check_hip(c, hipStreamSynchronize(c->stream), "gnn grad");
}

int pic_gnn_train(pic_ctx* c, int32_t steps, double lr, uint64_t seed, double* loss_first, double* loss_last) {
    if (!c || !c->gnn) return c ? c->fail(PIC_ERR_INVALID, "pic_gnn_init first") : PIC_ERR_INVALID;
    Gnn& G = *c->gnn;
    if (G.snap_count == 0) return c->fail(PIC_ERR_INVALID, "no snapshots recorded");
    const size_t n = (size_t)c->hm.n_cells;
    double first = 0.0, last = 0.0;
    int nlast = 0;
    const int tail = std::max(1, steps / 10);
    for (int s = 0; s < steps; ++s) {
        uint64_t z = (seed + 1) * 0x9E3779B97F4A7C15ull + (uint64_t)s * 0xBF58476D1CE4E5B9ull;
        z ^= z >> 31; z *= 0x94D049BB133111EBull; z ^= z >> 29;
        const int idx = (int)(z % (uint64_t)G.snap_count);
        G.forward(G.snap_x + idx * n * 3, true);
        const bool want = (s == 0) || (s >= steps - tail);
        double l = 0.0;
        G.loss_backward2(G.snap_t + idx * n, G.snap_norm[idx], G.snap_rnorm[idx], c->pd.nb, c->pd.a, c->pd.diag,
                         c->pd.K, want ? &l : nullptr);
        if (s == 0) first = l;
        if (s >= steps - tail) { last += l; ++nlast; }
        G.adam((float)lr, ++G.adam_t);
    }
    if (loss_first) *loss_first = first;
    if (loss_last) *loss_last = nlast ? last / nlast : 0.0;
    return check_hip(c, hipStreamSynchronize(c->stream), "gnn train");
}

int pic_gnn_eval(pic_ctx* c, int32_t idx, double* loss) {
    return pic_gnn_grad(c, idx, nullptr, loss);
}

int pic_gnn_loss(pic_ctx* c, int32_t kind) {
    if (!c || !c->gnn) return c ? c->fail(PIC_ERR_INVALID, "pic_gnn_init first") : PIC_ERR_INVALID;
    if (kind < 0 || kind > 2) return c->fail(PIC_ERR_INVALID, "loss kind 0 (A-norm), 1 (residual) or 2 (both)");
    c->gnn->loss_kind = kind;
    return PIC_OK;
}

int pic_gnn_use(pic_ctx* c, int32_t on) {
    if (!c || !c->gnn) return c ? c->fail(PIC_ERR_INVALID, "pic_gnn_init first") : PIC_ERR_INVALID;
    c->gnn_use = on != 0;
    return PIC_OK;
}

}  // extern "C"
