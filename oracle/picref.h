/*
 * picref — FP64 CPU restatement of picFoam's per-timestep particle loop.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the oracle the HIP path is checked against.  Only
 * tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg may load it; the
 * product library never links or calls it.
 *
 * Every function names the reference file:line it restates (paths relative to the
 * picFoam source tree).  Where the arithmetic lives in OpenFOAM-9 (not vendored in the
 * reference: tet tracking, primitiveMesh geometry, PCG/DIC, Gauss grad,
 * volPointInterpolation, Random) the restatement re-derives the published algorithm and
 * says so; those parts are pinned by closed forms and brute force in tests/, not by the
 * reference's own numbers (the reference ships no tests and no outputs).
 */
#ifndef PICREF_H
#define PICREF_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ref_mesh ref_mesh;

/* patch types (constant/polyMesh/boundary `type`) */
enum { REF_PATCH_PATCH = 0, REF_PATCH_WALL = 1, REF_PATCH_CYCLIC = 2, REF_PATCH_EMPTY = 3 };
/* pushers (ParticlePusherNew.C:39) */
enum { REF_PUSH_NONE = 0, REF_PUSH_BORIS = 1, REF_PUSH_BORISNR = 2, REF_PUSH_VAY = 3, REF_PUSH_HC = 4 };
/* FieldWeigthing / ChargeDistribution (FieldWeigthing.C:77, ChargeDistributionNew.C:44) */
enum { REF_W_CELLAVERAGE = 0, REF_W_VOLUMEWEIGHTING = 1 };

#define REF_FIXED_SHIFT 42 /* int64 fixed-point deposit: 1 charge quantum = 2^42 units */

ref_mesh* ref_mesh_create(int n_points, const double* points, int n_faces, const int* face_offsets,
                          const int* face_points, const int* owner, int n_internal, const int* neighbour,
                          int n_cells, int n_patches, const int* patch_type, const int* patch_start,
                          const int* patch_size, const int* patch_nbr);
void ref_mesh_destroy(ref_mesh* m);
int ref_mesh_n_tets(const ref_mesh* m);
int ref_mesh_n_points(const ref_mesh* m);
void ref_mesh_geometry(const ref_mesh* m, double* cell_centres, double* cell_volumes, double* face_centres,
                       double* face_areas);
void ref_mesh_tets(const ref_mesh* m, int* tet_cell, int* tet_face, int* tet_pt);
void ref_mesh_point_master(const ref_mesh* m, int* pmaster);
/* face base point per face (an index into its point list) */
void ref_mesh_face_base(const ref_mesh* m, int* base);

/* positions <-> (tet, barycentric y1..y3) */
void ref_positions(const ref_mesh* m, int n, const double* y, const int* tet, double* x);
int ref_locate(const ref_mesh* m, int n, const double* x, const int* cell, double* y, int* tet);

/* single-particle pusher: U (velocity) updated in place; E, B at the particle */
void ref_push_one(int pusher, double* U, const double* E, const double* B, double q, double mass, double dt,
                  int hc_reference_form, int rotate);

/* E at particles (FieldWeigthing::getFieldVector) */
void ref_interp_E(const ref_mesh* m, int weighting, int n, const double* y, const int* tet, const double* Ecell,
                  const double* Epoint, double* Ep);

/* one call of the move loop for one species: push (if charged & movable) + track.
 * y[n*3], U[n*3], tet[n] updated in place; tet=-1 marks a particle that left through a
 * non-coupled patch.  Returns number of particles that hit the tracking guard. */
int ref_move(const ref_mesh* m, int pusher, int weighting, int n, double* y, double* U, int* tet, double q,
             double mass, double dt, const double* Ecell, const double* Epoint, const double* Bcell,
             int do_push, int hc_reference_form, int rotate, int* n_superluminal);

/* Wall patches (PICParcel.C:153-184, :209-330; the WallReflectionModel models): a parcel that
 * reaches a wall face has its velocity synchronised to the hit time, reflected by the
 * model, and synchronised back (the second and later resyncs of a step skipped), then tracks
 * on.  Without walls (NULL) a wall face deletes the parcel like any non-coupled patch. */
enum { REF_WALL_SPECULAR = 0, REF_WALL_MAXWELLIAN_THERMAL = 1, REF_WALL_MIXED = 2 };
typedef struct {
    int model;                 /* REF_WALL_* */
    double diffuse_fraction;   /* MixedDiffuseSpecularCoeffs.diffuseFraction */
    const double* T;           /* [n_patches] wall temperature (boundaryT) */
    const double* U;           /* [n_patches*3] wall velocity (boundaryU) */
    int sync;                  /* SolverSettings.syncVelocityAtBoundary */
    uint64_t seed;
    int64_t step;
    uint32_t species;
    const uint32_t* id;        /* [n] parcel ids: the keys of the reflection draws */
    int64_t* hits;             /* optional [1]: wall hits counted */
    /* optional [n_patches]: what a face of each patch does to this species (NULL: walls
     * reflect by the model, other non-coupled patches delete): 0 delete, 1 the wall model,
     * 2 SimpleInjection's reflection (hit sync, specular, no resync; SimpleInjection.C:
     * 245-252, PICParcel.C:347-357), 3 its re-emission with useInitalVelocity (:253-270,
     * ParticleEmitter.C:136-221) by the emitter `reemit` describes */
    const int* patch_kind;
    const struct ref_inject_params_s* reemit;
    /* optional: the wall surface fields (PICParcel.C:240-339) added per boundary face
     * (face - n_internal) in rows of acc_stride doubles: rhoN, rhoM, linearKE, momentum[3],
     * q, fD[3], then at 10 + 9 * species: rhoN, rhoM, momentum[3], linearKE, fD[3] */
    double* face_acc;
    int acc_stride;
    double weight;             /* nEquivalentParticles */
} ref_wall_params;
int ref_move_walls(const ref_mesh* m, int pusher, int weighting, int n, double* y, double* U, int* tet, double q,
                   double mass, double dt, const double* Ecell, const double* Epoint, const double* Bcell,
                   int do_push, int hc_reference_form, int rotate, int* n_superluminal,
                   const ref_wall_params* walls);

/* SimpleInjection (SimpleInjection.C:186-229) through ParticleEmitter::emitParticle
 * (ParticleEmitter.C:223-349, velocity models :357-535): `count` new parcels of one species
 * from one injector patch.  velocity_model: 0 birdMaxwellianFlux, 1 mostProbableSpeed,
 * 2 halfMaxwellianFlux, 3 maxwellianFlux.  Writes y[3*count], tet[count], U[3*count] and,
 * when id is not NULL, id[count]. */
typedef struct ref_inject_params_s {
    int patch;
    int velocity_model;
    double T;                  /* K: eV * e / k */
    double mass, q;
    int sync;                  /* push charged parcels by -dt/2 with the field at their position */
    int pusher, weighting, hc_reference_form, rotate;
    double dt;
    const double *Ecell, *Epoint, *Bcell;
    uint64_t seed;
    int64_t step;
    uint32_t species, model;
} ref_inject_params;
void ref_inject(const ref_mesh* m, const ref_inject_params* p, int count, double* y, int* tet, double* U,
                uint32_t* id);

/* push only (leapfrog initialisation, picInitialise.C:228-237) */
void ref_push(const ref_mesh* m, int pusher, int weighting, int n, const double* y, double* U, const int* tet,
              double q, double mass, double dt, const double* Ecell, const double* Epoint, const double* Bcell,
              int hc_reference_form, int rotate);

/* int64 fixed-point charge deposit (ChargeDistribution::add) for one species;
 * scale = q*w / Q0 where Q0 is the charge quantum.  acc_cell[n_cells], acc_point[n_points] */
void ref_deposit(const ref_mesh* m, int weighting, int n, const double* y, const int* tet, double scale,
                 int64_t* acc_cell, int64_t* acc_point);
/* ChargeDistribution::update: rho[n_cells] from the accumulators */
void ref_deposit_finish(const ref_mesh* m, int weighting, const int64_t* acc_cell, const int64_t* acc_point,
                        double Q0, double* rho);

/* volPointInterpolation (inverse distance, cyclic-coupled): Ecell[n_cells*3] -> Epoint[n_points*3] */
void ref_cell_to_point(const ref_mesh* m, const double* Ecell, double* Epoint);

/* Philox4x32-10 */
void ref_philox4x32(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* counting sort by cell (buildCellOccupancy, PICCloud.C:327-354): returns order[] (stable)
 * and cell_start[n_cells+1]; dead particles (tet<0) are dropped. */
int ref_sort_by_cell(const ref_mesh* m, int n, const int* tet, int* order, int* cell_start);

/* Coulomb collisions on sorted species arrays: Nanbu (model 0) or Perez (model 1)
 * scattering, Takizuka-Abe (pairing 0) or Nanbu-Yonemura (pairing 1) pairing, optional
 * Nanbu weight correction.  stats: [0] collisions, [1] sum lnL, [2] sum s, [3] sum Debye. */
typedef struct {
    int n_species;
    double mass[8], charge[8], weight[8];
    double coulomb_log[8][8]; /* >0 fixed per species pair, 0 computed from the Debye length */
    int calculate_bmin;
    int allow_intra;
    uint64_t seed;
    int64_t step;
    int model;
    int pairing;
    int weight_correction;
    unsigned debye_mask;   /* species in the Debye length (charged ones among them) */
    int debye_drift;       /* 1: drift removed from the species temperature */
    int* const* hits;      /* optional per species, per parcel: collisions taken part in */
} ref_coulomb_params;
void ref_set_threads(int n);
void ref_coulomb(const ref_mesh* m, const ref_coulomb_params* p, double dt, double* const* U,
                 const uint32_t* const* id, const int* const* cell_start, double* stats);

/* electron-neutral null-collision MCC with a uniform background */
typedef struct {
    double massE, massN, nN, TN, wE, wN, driftN[3];
    int n_el, n_ex, n_iz;
    const double *e_el, *s_el, *e_ex, *s_ex, *e_iz, *s_iz;
    double th_ex, th_iz;
    int create_ion, create_electron, relativistic_scatter;
    uint64_t seed;
    int64_t step;
} ref_mcc_params;
/* electrons sorted by cell; sigma_cr_max[n_cells], remainder[n_cells] updated in place;
 * new electrons / ions appended after n_e / n_i (capacity checked); returns 0 or -1 */
int ref_mcc(const ref_mesh* m, const ref_mcc_params* p, double dt, double* yE, double* UE, int* tetE,
            uint32_t* idE, const int* cell_start, int* n_e, int cap_e, double* yI, double* UI, int* tetI,
            uint32_t* idI, int* n_i, int cap_i, double* sigma_cr_max, double* remainder, int64_t* counts);

/* ion-neutral null-collision MCC with a uniform background (IsotropicScattering) */
typedef struct {
    double massI, massN, nN, TN, driftN[3];
    int n_el, n_cx;
    const double *e_el, *s_el, *e_cx, *s_cx;
    double th_cx;
    uint64_t seed;
    int64_t step;
} ref_ion_mcc_params;
/* ions sorted by cell (cell_start); sigma_cr_max and remainder updated in place;
 * counts[0..2] += candidates, elastic, charge exchange */
int ref_ion_mcc(const ref_mesh* m, const ref_ion_mcc_params* p, double dt, double* UI, const uint32_t* idI,
                const int* cell_start, double* sigma_cr_max, double* remainder, int64_t* counts);
/* IonIsotropicScattering::updateVelocity on (UP, UQ) in place */
void ref_ion_isotropic(double* UP, double* UQ, double mP, double mQ, double u1, double u2);

/* parcel pairs: the binary model's neutral DSMC and the ion-neutral model's parcel
 * collisions.  Species [0, ns) in list order, each sorted by cell (cell_start); the first
 * na form the first group (the neutrals). */
typedef struct {
    int ns, na;
    double mass[8], w[8], d[8], omega[8];
    double* U[8];
    const double* y[8];
    const int* tet[8];
    const uint32_t* id[8];
    const int* cell_start[8];
    int vhs;
    double tref;
    int weight_correction, keep_neutral;
    int bg;
    double nN, TN, massN, dN, omegaN, driftN[3];
    int n_el, n_cx;
    const double *e_el, *s_el, *e_cx, *s_cx;
    double th_cx;
    uint64_t seed;
    int64_t step;
} ref_pair_params;
int ref_binary(const ref_mesh* m, const ref_pair_params* p, double dt, double* sigma_cr_max, double* remainder,
               double* remainder_bg, int64_t* counts);
int ref_ion_pairs(const ref_mesh* m, const ref_pair_params* p, double dt, double* sigma_cr_max, double* remainder,
                  int64_t* counts);
double ref_sigma_tcr(int vhs, double tref, double dP, double dQ, double omP, double omQ, double mP, double mQ,
                     const double* UP, const double* UQ);

/* electron-neutral scattering with a sampled partner velocity, in place:
 * NanbuElectronNeutralCollision.C:98-130 (relativistic 0) or
 * RelativisticElectronNeutralCollision.C:161-240 (relativistic 1) */
void ref_en_scatter(int relativistic, double eV, double* UE, double* Un, double mE, double mN, double ueps,
                    double uchi);

/* ListCrossSection interpolateXY (ListCrossSection.C:32-92) */
double ref_interpolate_xy(double x, int n, const double* xs, const double* ys);

/* Poisson: FV Laplacian (Gauss linear corrected: implicit orthogonal part, explicit
 * non-orthogonal correction of the starting potential) + multicolour DIC PCG */
typedef struct ref_poisson ref_poisson;
ref_poisson* ref_poisson_create(const ref_mesh* m, const int* phi_bc_type, const double* phi_bc_value);
void ref_poisson_destroy(ref_poisson* p);
int ref_poisson_n_colours(const ref_poisson* p);
void ref_poisson_colours(const ref_poisson* p, int* colour);
/* solve A phi = rho V / eps0 (+BC) + correction(phi in) to ||r||/||b|| <= tol; phi warm start
 * in/out; returns iters */
int ref_poisson_solve(const ref_poisson* p, const double* rho, double* phi, double tol, int max_iter,
                      int precond, double* final_rel_res);
void ref_poisson_apply(const ref_poisson* p, const double* x, double* y);
void ref_poisson_rhs(const ref_poisson* p, const double* rho, double* b);
void ref_grad(const ref_poisson* p, const double* phi, double* gradphi);
/* 1 when some face's non-orthogonal correction vector exceeds 1e-10 |Sf| */
int ref_poisson_nonorth(const ref_poisson* p);
/* the explicit non-orthogonal correction of phi per cell (adds to the right-hand side) */
void ref_poisson_correction(const ref_poisson* p, const double* phi, double* corr);

#ifdef __cplusplus
}
#endif
#endif
