/*
 * picmi.h — C-ABI of the MI355X-native particle step (libpicmi.so).
 *
 * The reference (picFoam, OpenFOAM-9) selects its submodels at run time from
 * constant/picProperties and calls them once per parcel through virtual functions.  A GPU
 * cannot be driven per parcel, so this ABI keeps the same model names and dictionary
 * keywords but is batch-granular: one call runs one stage of PICCloud::evolve() for every
 * parcel on the device.  All pointers are plain host pointers; nothing is retained after a
 * call returns.  Every entry point returns 0 (PIC_OK) or a negative PIC_ERR_* code and
 * leaves a message for pic_last_error().  No C++ exception crosses this boundary.
 */
#ifndef PICMI_H
#define PICMI_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PICMI_ABI_VERSION 1

typedef struct pic_ctx pic_ctx;

/* status codes; the reference aborts the process for each of these (FatalError) */
enum {
    PIC_OK = 0,
    PIC_ERR_INVALID = -1,        /* bad argument */
    PIC_ERR_UNKNOWN_MODEL = -2,  /* e.g. ParticlePusherNew.C:47-55 (lists valid names) */
    PIC_ERR_SUPERLUMINAL = -3,   /* |U| >= c, PICParcel.C:61-66 */
    PIC_ERR_HIP = -4,            /* HIP runtime error */
    PIC_ERR_CAPACITY = -5,       /* more parcels than the species capacity */
    PIC_ERR_MESH = -6,           /* inconsistent polyMesh */
    PIC_ERR_UNSUPPORTED = -7,    /* option outside the implemented set */
    PIC_ERR_CELL_OVERFLOW = -8,  /* more parcels in one cell than a collision kernel holds */
    PIC_ERR_NOT_CONVERGED = -9,  /* Poisson solve hit its iteration limit */
    PIC_ERR_NO_DEVICE = -10      /* no HIP device */
};

/* constant/polyMesh/boundary types.  A parcel reaching a PIC_PATCH_PATCH face leaves the
 * domain (deleted); one reaching a PIC_PATCH_WALL face is reflected by
 * pic_models.wall_reflection and tracks on within the same step (PICParcel.C:153-184).
 * PIC_PATCH_EMPTY marks the front and back of a
 * reduced-dimension (2-D or 1-D) case: the tracking velocity loses its components normal
 * to them (PICParcel.C:118-123, meshTools::constrainDirection over polyMesh::solutionD), so
 * parcels stay in their plane; no flux crosses them in the Poisson solve. */
enum { PIC_PATCH_PATCH = 0, PIC_PATCH_WALL = 1, PIC_PATCH_CYCLIC = 2, PIC_PATCH_EMPTY = 3 };
/* phi boundary conditions (0/phiE) */
enum { PIC_BC_ZERO_GRADIENT = 0, PIC_BC_FIXED_VALUE = 1 };
/* arithmetic of the particle step */
enum { PIC_PREC_FP64 = 0, PIC_PREC_MIXED = 1 };
/* Poisson solver arithmetic */
enum { PIC_POISSON_FP64 = 0, PIC_POISSON_MIXED_IR = 1 };
/* cell fields for pic_set_cell_field / pic_get_cell_field */
enum {
    PIC_FIELD_E = 0,            /* [n_cells*3] electric field */
    PIC_FIELD_B = 1,            /* [n_cells*3] magnetic field */
    PIC_FIELD_PHI = 2,          /* [n_cells]   electric potential */
    PIC_FIELD_RHO = 3,          /* [n_cells]   charge density (get only) */
    PIC_FIELD_SPACE_CHARGE = 4, /* [n_cells]   background space charge density */
    PIC_FIELD_E_POINT = 5,      /* [n_points*3] E interpolated to the mesh points (get only) */
    PIC_FIELD_MCC_SIGMA_CR_MAX = 6, /* [n_cells] electron-background (sigma cR)_max (picSigmaTcRMax_ElectronBackground) */
    PIC_FIELD_MCC_REMAINDER = 7,    /* [n_cells] background candidate remainder */
    PIC_FIELD_ION_SIGMA_CR_MAX = 8, /* [n_cells] ion-background (sigma cR)_max (picSigmaTcRMax_IonBackground) */
    PIC_FIELD_ION_REMAINDER = 9,    /* [n_cells] ion-background candidate remainder */
    PIC_FIELD_BINARY_SIGMA_CR_MAX = 10,     /* [n_cells] neutral (sigma_T cR)_max (picSigmaTcRMax) */
    PIC_FIELD_BINARY_REMAINDER = 11,        /* [n_cells] neutral pair candidate remainder
                                               (pic:collisionSelectionRemainder) */
    PIC_FIELD_BINARY_BG_REMAINDER = 12,     /* [n_cells] neutral-background candidate remainder */
    PIC_FIELD_ION_PARCEL_SIGMA_CR_MAX = 13, /* [n_cells] ion-neutral parcel (sigma cR)_max
                                               (picSigmaTcRMax_IonNeutral) */
    PIC_FIELD_ION_PARCEL_REMAINDER = 14     /* [n_cells] ion-neutral parcel candidate remainder */
};
/* cross-section processes (CrossSectionModel.H:52-59), tabulated "List" model */
enum { PIC_XS_ELASTIC = 0, PIC_XS_EXCITATION = 1, PIC_XS_IONIZATION = 2 };
/* ion-neutral processes (IonNeutralCollisionCoeffs.CrossSectionModels: ElasticCrossSection,
 * ChargeExchangeCrossSection with its threshold) */
enum { PIC_XS_ION_ELASTIC = 3, PIC_XS_ION_CHARGE_EXCHANGE = 4 };
/* stage ids for pic_stage_times */
enum {
    PIC_STAGE_MOVE = 0, PIC_STAGE_SORT = 1, PIC_STAGE_COULOMB = 2, PIC_STAGE_MCC = 3,
    PIC_STAGE_DEPOSIT = 4, PIC_STAGE_POISSON = 5, PIC_STAGE_FIELD = 6, PIC_STAGE_COUNT = 7
};

typedef struct {
    int32_t type;            /* PIC_PATCH_* */
    int32_t start, size;     /* face range, as in constant/polyMesh/boundary */
    int32_t neighbour_patch; /* cyclic partner patch, -1 otherwise */
    int32_t phi_bc;          /* PIC_BC_* for the Poisson solve */
    double phi_value;        /* fixedValue */
    double wall_temperature; /* wall patches: boundaryT (K), for diffuse reflection */
    double wall_velocity[3]; /* wall patches: boundaryU (m/s) */
} pic_patch;

/* OpenFOAM polyMesh: internal faces first (upper-triangular order), face normals out of
 * the owner, face i of a cyclic patch matching face i of its partner. */
typedef struct {
    int32_t n_points, n_faces, n_internal_faces, n_cells, n_patches;
    const double* points;        /* [n_points*3] */
    const int32_t* face_offsets; /* [n_faces+1] */
    const int32_t* face_points;  /* [face_offsets[n_faces]] */
    const int32_t* owner;        /* [n_faces] */
    const int32_t* neighbour;    /* [n_internal_faces] */
    const pic_patch* patches;    /* [n_patches] */
} pic_mesh;

/* moleculeProperties.<species> (PICCloud.C:57-317) */
typedef struct {
    double mass;          /* kg */
    double charge;        /* C, already multiplied by the charge modifier */
    double n_equivalent;  /* nEquivalentParticles (weight) */
    int32_t solve_movement;
    int64_t capacity;     /* parcels held on the device for this species */
    int32_t sort_interval; /* this species' sort_interval (see pic_models); 0 = the models' value.
                              A heavy species that seldom changes cell can move into cell order
                              less often than a light one */
    double diameter;      /* moleculeProperties.<species>.diameter (m), the hard-sphere size of the
                             binary collision models (PICParcelI.H:45) */
    double omega;         /* moleculeProperties.<species>.omega, the viscosity index of the
                             VariableHardSphere model (PICParcelI.H:46) */
} pic_species;

/* constant/picProperties keywords; names as the reference's selection tables */
typedef struct {
    const char* particle_pusher;      /* ParticlePusher: Boris BorisNR Vay HigueraCary none */
    const char* field_weighting;      /* SolverSettings.FieldWeigthing: CellAverage VolumeWeighting */
    const char* charge_distribution;  /* SolverSettings.ChargeDistribution: CellAverage VolumeWeighting */
    const char* maxwell_solver;       /* MaxwellSolver: ElectroStatic none */
    const char* coulomb_collision;    /* CollisionModels.CoulombCollisionModel: Nanbu Perez none */
    const char* pairing_algorithm;    /* <model>Coeffs.PairingAlgorithm: TakizukaAbe NanbuYonemura */
    const char* weight_correction;    /* <model>Coeffs.WeightCorrectionModel: none Nanbu */
    const char* electron_neutral;     /* CollisionModels.ElectronNeutralCollisionModel: NanbuElectronNeutral
                                         RelativisticElectronNeutral none */
    const char* background_gas;       /* CollisionModels.BackgroundGasModel: uniformBackground none */
    int32_t allow_intra_collision;    /* NanbuCoeffs.allowIntraCollision */
    int32_t calculate_bmin;           /* NanbuCoeffs.DebyeLength.calculate_bmin */
    double coulomb_log[8][8];         /* NanbuCoeffs.Collisions.*.coulombLog; 0 = from the Debye length */
    int32_t electron_species;         /* SpeciesRelations.electronTypeId */
    int32_t ion_species;              /* ion created by background ionization, -1 none */
    double bg_number_density, bg_temperature, bg_mass, bg_n_equivalent, bg_velocity[3];
    int32_t create_ion, create_electron;
    int32_t precision;                /* PIC_PREC_* */
    int32_t hc_reference_form;        /* 1: Higuera-Cary with the reference's (t.u)^2 term */
    uint64_t seed;
    double poisson_tolerance;         /* ||r||2/||b||2 */
    int32_t poisson_max_iter;
    int32_t poisson_mode;             /* PIC_POISSON_* */
    int32_t sort_interval;            /* parcels move into cell order every k steps; a collision model
                                         gets the cell order as an index on the steps between */
    int32_t poisson_guess;            /* PCG start: 0 the previous phi, 1 2 phi(n-1) - phi(n-2) */
    /* <model>Coeffs.DebyeLength (CoulombCollisionModel.C:98-141): the species whose parcels
     * enter the Debye length, as a bit mask of species ids (0 = every charged species, the
     * reference's default; a listed species without charge is PIC_ERR_INVALID), and
     * accountForDrift false (1 here) to keep the drift in the species temperature */
    uint32_t debye_species;
    int32_t debye_ignore_drift;
    /* BoundaryModels.WallReflectionModel for wall patches: SpecularReflection (also when
     * NULL), MaxwellianThermal, MixedDiffuseSpecular with MixedDiffuseSpecularCoeffs.
     * diffuseFraction (PICParcel.C:209-330, WallReflectionModel/); a parcel reaching a wall
     * has its velocity synchronised to the hit time, reflected and synchronised back
     * (PICParcel.C:153-210) unless SolverSettings.syncVelocityAtBoundary is false (1 here);
     * as in the reference only the first reflection of a step is synchronised back
     * (PICParcel.C:130-139).  The diffuse draws are keyed by (seed; parcel id, step,
     * species and hit number) */
    const char* wall_reflection;
    double diffuse_fraction;
    int32_t no_boundary_sync;
    /* CollisionModels.IonNeutralCollisionModel: IsotropicScattering or none (NULL); the ions
     * collide with the parcels of the neutral species (IonNeutralCollisionModel.C:580-713)
     * and with the uniformBackground gas (:447-560), elastically and by charge exchange.  As
     * in the reference (:170-175) one neutral and one ion species at most */
    const char* ion_neutral_collision;
    /* CollisionModels.BinaryCollisionModel: IsotropicScattering or none (NULL).  Neutral parcels
     * collide with each other (no time counter, 8 sub-cells; BinaryCollisionModel.C:229-425)
     * and, with a uniformBackground gas, with the background (:427-525).  The total cross
     * section is IsotropicScatteringCoeffs.TotalCrossSectionModel: HardSphere (NULL) or
     * VariableHardSphere with VariableHardSphereCoeffs.Tref (vhs_tref, K); the weight
     * correction is IsotropicScatteringCoeffs.WeightCorrection: none (NULL) or Nanbu */
    const char* binary_collision;
    const char* total_cross_section;
    double vhs_tref;
    const char* binary_weight_correction;
    /* IonNeutralCollisionCoeffs.WeightCorrection (none/NULL or Nanbu) and updateNeutralVelocity
     * false (1 here) for the ion collisions with neutral parcels (IonNeutralCollisionModel.C:
     * 401-449) */
    const char* ion_neutral_weight_correction;
    int32_t ion_neutral_keep_neutral;
    /* BackgroundGasCoeffs.species: the background gas' species, whose diameter and omega the
     * binary model's background collisions use (BinaryCollisionModel.C:486-492) */
    int32_t bg_species;
} pic_models;

typedef struct {
    int32_t iterations;        /* PCG iterations (inner, summed over refinements) */
    int32_t refinements;       /* outer FP64 refinements (mixed mode) */
    double rel_residual;       /* final ||b - A phi||2 / ||b||2 in FP64 */
    double ms;                 /* solve time on the device */
} pic_poisson_stats;

typedef struct {
    int64_t n_parcels[8];
    int64_t n_deleted;          /* left through non-coupled patches this step */
    int64_t n_track_guard;      /* hit the tracking iteration guard */
    int64_t coulomb_collisions;
    int64_t mcc_candidates, mcc_elastic, mcc_excitation, mcc_ionization;
    pic_poisson_stats poisson;
    int64_t n_wall_hits;        /* wall reflections this step (a parcel may reflect more than once) */
    int64_t ion_neutral_candidates, ion_neutral_elastic, ion_neutral_charge_exchange;
    /* ion collisions with neutral parcels (IonNeutralCollisionModel.C:580-713) */
    int64_t ion_parcel_candidates, ion_parcel_elastic, ion_parcel_charge_exchange;
    /* BinaryCollisionModel: neutral pairs, then the neutrals with the background gas */
    int64_t binary_candidates, binary_collisions, binary_bg_candidates, binary_bg_collisions;
    int64_t n_injected;         /* parcels added by the SimpleInjection models */
} pic_step_stats;

/* One PatchBoundaryModels entry of type SimpleInjection (SimpleInjection.C:34-175) on a
 * patch of type `patch`.  Every step whose end time t + dt is at most end_time
 * (SimpleInjection.C:191-197), each listed species gets floor(frequency / nEquivalentParticles
 * * dt + remainder) new parcels, the fraction carried to the next step (:204-210), placed
 * and given velocities by the emitter (ParticleEmitter.C:223-349, :357-535): a triangle of
 * the patch by area (faces fanned from their first point), a uniform point in it moved into
 * the owner cell by 0.1-0.5 of its normal distance from the cell centre, the components of
 * empty directions set to zero, and a velocity of the model at temperature eV (in eV) in the
 * frame of the face.  Charged species are synchronised back by half a step with the field at
 * their position when the boundary sync is on (ParticleEmitter.C:338).  The new parcels move
 * their full step in the same pic_step.  interaction[] gives, per species of the cloud, what
 * a parcel reaching the patch does (SimpleInjection.C:233-292): 0 deletion (the default);
 * 1 reflection: its velocity synchronised to the hit and reflected specularly, not
 * synchronised back (PICParcel.C:158-161, :347-357); 2 reflection with useInitalVelocity:
 * for a species the model injects, the parcel is replaced (in its slot, with its id) by one
 * from the emitter, the triangle drawn uniformly (ParticleEmitter.C:136-221), synchronised
 * to the hit's step fraction and tracked for the rest of the step, else as 1; 3 floating
 * is refused (PIC_ERR_UNSUPPORTED: it needs the circuit boundary condition). */
typedef struct {
    int32_t patch;
    int32_t velocity_model;       /* PIC_INJECT_* */
    int32_t n_species;            /* injectionTypeIds entries */
    int32_t species[8];           /* injectionTypeIds */
    double frequency[8];          /* frequencies: real particles per second */
    double eV[8];                 /* eV: temperature of the emitted distribution */
    double end_time;              /* injectionEndTime */
    double time;                  /* the time at step index 0; advanced by dt every step */
    int32_t interaction[8];       /* per cloud species: 0 deletion, 1 reflection, 2 with re-emission */
} pic_injection;
enum {
    PIC_INJECT_BIRD_MAXWELLIAN_FLUX = 0,  /* birdMaxwellianFlux */
    PIC_INJECT_MOST_PROBABLE_SPEED = 1,   /* mostProbableSpeed */
    PIC_INJECT_HALF_MAXWELLIAN_FLUX = 2,  /* halfMaxwellianFlux */
    PIC_INJECT_MAXWELLIAN_FLUX = 3        /* maxwellianFlux */
};

typedef struct {
    /* TemperatureInfo.C:72-138 / KineticEnergyInfo.C:69-82 sums per species */
    double sum_w[8], sum_wU[8][3], sum_wU2[8], kinetic_energy[8];
    double field_energy;        /* FieldEnergyInfo.C:118-124: 0.5 eps0 sum |E|^2 V */
    int64_t n_parcels[8];
    /* KineticEnergyInfo.C:73-77 relativistic branch: sum (gamma - 1) m c^2, without the
     * parcel weight as in the reference */
    double kinetic_energy_relativistic[8];
    /* IsotropizationInfo.C:60-66 / UnidirectionalVelocityInfo.C:60-65: sum w U_d U_d per
     * component; MomentumInfo.C:65-75 relativistic branch: sum m w gamma U; DebyeLengthInfo.C:
     * 93-99: sum q w */
    double sum_wUU[8][3];
    double momentum_relativistic[8][3];
    double charge[8];
} pic_diag;

const char* pic_version(void);
int pic_device_count(int32_t* n);

/* PICCloud::PICCloud + setModels (PICCloud.C:734-1000, :378-518) */
int pic_create(const pic_mesh* mesh, const pic_species* species, int32_t n_species, const pic_models* models,
               pic_ctx** out);
void pic_destroy(pic_ctx* ctx);
const char* pic_last_error(const pic_ctx* ctx);

/* mesh-derived tables (the tet decomposition every parcel refers to) */
int pic_mesh_sizes(const pic_ctx* ctx, int64_t* n_tets, int32_t* n_cells, int32_t* n_points);
int pic_mesh_tets(const pic_ctx* ctx, int32_t* tet_cell, int32_t* tet_face, int32_t* tet_pt);
int pic_mesh_geometry(const pic_ctx* ctx, double* cell_centres, double* cell_volumes);

/* fields */
int pic_set_cell_field(pic_ctx* ctx, int32_t field, const double* values);
int pic_get_cell_field(pic_ctx* ctx, int32_t field, double* values);
/* ListCrossSection energies [eV] / values [m^2] (ListCrossSection.C:142-323) */
int pic_set_cross_section(pic_ctx* ctx, int32_t process, int32_t n, const double* energies_ev,
                          const double* sigma_m2, double threshold_ev);
/* ElectronNeutralCollisionModel::initialize "FindMaximum" (ElectronNeutralCollisionModel.C:375-420) */
int pic_mcc_initialize(pic_ctx* ctx, double maximum_energy_ev, double energy_step_ev);
/* IonNeutralCollisionModel::initialize "FindMaximum" for the background (IonNeutralCollisionModel.C:
 * 282-330): the largest sqrt(2 e E / m_ion) (Qel + Qcx)(E) on E = 0, step, ... below the maximum */
int pic_ion_mcc_initialize(pic_ctx* ctx, double maximum_energy_ev, double energy_step_ev);
/* the "Temperature" initialisation of either model (ElectronNeutralCollisionModel.C:432-482,
 * IonNeutralCollisionModel.C:331-379): Qt(k T / e) times the most probable speed
 * sqrt(2 k T / m) of the electron (model 0) or ion (model 1) species at temperature [K];
 * a zero temperature stands for 1 eV as in the reference */
int pic_mcc_initialize_temperature(pic_ctx* ctx, int32_t model, double temperature);
/* BinaryCollisionModel::initialize (BinaryCollisionModel.C:534-600): every cell's (sigma_T
 * cR)_max set to pi d^2 sqrt(2 k T / m) of the neutral species with the largest number
 * density (the background's density added to its species), at that species' temperature
 * (the background's when larger; 293.15 K for a zero temperature).  temperatures [K] and
 * number_densities [m^-3] are per species, [n_species] */
int pic_binary_initialize(pic_ctx* ctx, const double* temperatures, const double* number_densities);

/* particles: barycentric y1..y3 w.r.t. (cell centre, base, A, B) of tet, velocity U, id */
int pic_particles_set(pic_ctx* ctx, int32_t species, int64_t n, const double* bary, const int32_t* tet,
                      const double* U, const uint32_t* id);
/* positions + containing cell -> (tet, bary) located on the device */
int pic_particles_set_positions(pic_ctx* ctx, int32_t species, int64_t n, const double* x, const int32_t* cell,
                                const double* U, const uint32_t* id);
/* count and arrays as stored: between moving sorts (sort_interval > 1) they can hold
 * parcels deleted since the last one, marked by tet -1 */
int pic_particles_count(pic_ctx* ctx, int32_t species, int64_t* n);
int pic_particles_get(pic_ctx* ctx, int32_t species, int64_t capacity, int64_t* n, double* bary, int32_t* tet,
                      double* U, uint32_t* id);
int pic_particles_positions(pic_ctx* ctx, int32_t species, int64_t capacity, double* x);
/* pic_particles_positions without any device allocation (no counterpart in the reference):
 * for callers that cannot call hipMalloc (code under HIP graph capture, tight serving
 * loops, harnesses that forbid it).  The positions are staged in a 32 MiB device array
 * that every GPU the library runs on holds until the process ends, so one call takes at
 * most 1398101 parcels (PIC_ERR_CAPACITY beyond).  Calls on one GPU share that array:
 * calls that overlap on the same GPU race. */
int pic_particles_positions_noalloc(pic_ctx* ctx, int32_t species, int64_t capacity, double* x);
/* picInitialise Uniform/Maxwellian: nppc parcels per cell, uniform in volume, Maxwellian
 * at temperature [K] plus drift; ids start at id0 */
int pic_init_maxwellian(pic_ctx* ctx, int32_t species, int32_t nppc, double temperature, const double* drift,
                        uint32_t id0, uint64_t seed);

/* stages of PICCloud::evolve (PICCloud.C:1426-1471) */
int pic_push(pic_ctx* ctx, double dt);                 /* ParticlePusher::updateVelocity only */
int pic_move(pic_ctx* ctx, double dt);                 /* Cloud::move: push + interpolate + track (the
                                                          tracking stage; PICParcel::move pushes first) */
int pic_sort(pic_ctx* ctx);                            /* buildCellOccupancy */
int pic_collide_coulomb(pic_ctx* ctx, double dt);      /* CoulombCollisionModel::handleCollisions */
int pic_collide_mcc(pic_ctx* ctx, double dt);          /* ElectronNeutralCollisionModel::handleCollisions */
int pic_collide_ion_neutral(pic_ctx* ctx, double dt);  /* IonNeutralCollisionModel::handleCollisions:
                                                          the ions with the neutral parcels, then with
                                                          the background gas */
int pic_collide_binary(pic_ctx* ctx, double dt);       /* BinaryCollisionModel::handleCollisions */
/* BoundaryModelList's SimpleInjection models (BoundaryModelList.C:34-167): replaces the list
 * (n = 0 removes it; at most 4 models, one patch each) and resets their remainders.
 * pic_step runs the injection first, as PICCloud::evolve (PICCloud.C:1435) does. */
int pic_set_injection(pic_ctx* ctx, const pic_injection* models, int32_t n);
int pic_inject(pic_ctx* ctx, double dt);               /* BoundaryModelList::injection alone */
/* The wall surface fields of the last move (PICParcel::wallReflection, PICParcel.C:240-339;
 * the boundary values of rhoN, rhoM, linearKE, momentum, q and fD and of their per-species
 * fields), per boundary face in face order (face - n_internal_faces).  species -1: 10
 * values per face (rhoN, rhoM, linearKE, momentum[3], q, fD[3]); species s: 9 values
 * (rhoN, rhoM, momentum[3], linearKE, fD[3]).  Each hit adds 1/max(|U.n| |Sf|, 1e-300)
 * weighted by 1, m, m|U|^2/2, m U_t before and after the reflection, and the energy and
 * momentum lost, times nEquivalentParticles / (dt |Sf|), to q and fD; n is the normal of the
 * face triangle the parcel hit.  Zero without wall patches. */
int pic_wall_fields(pic_ctx* ctx, int32_t species, double* out);
int pic_deposit(pic_ctx* ctx);                         /* calculateFields: charge density */
int pic_poisson_solve(pic_ctx* ctx, pic_poisson_stats* stats); /* MaxwellSolver::solveFields */
int pic_diag_reduce(pic_ctx* ctx, pic_diag* out);      /* DiagnosticsList::info sums */
/* PICCloud::calculateFields (PICCloud.C:587-695) moments per cell of one species, or of all
 * species (species = -1), out[n_cells][11]: rhoN, rhoM, picRhoN (parcel count), linearKE,
 * momentum[3], j[3], rhoCharge; densities per m^3.  Sorts the parcels by cell first. */
int pic_cell_moments(pic_ctx* ctx, int32_t species, double* out);
/* picFields (postProcess/functionObjects/picFields/picFields.C:97-212) over the time
 * averages OpenFOAM's fieldAverage keeps of the moment fields: op 0 clears the sums, op 1
 * adds the current rhoN, rhoM, linearKE and momentum of every species (a sample; the
 * parcels are sorted by cell first) */
int pic_fields_average(pic_ctx* ctx, int32_t op);
/* the means of one species over the samples and the fields picFields derives from them,
 * out[n_cells][11]: rhoNMean, rhoMMean, linearKEMean, momentumMean[3], UMean[3] =
 * momentumMean / rhoMMean, translationalT = 2 / (3 k rhoNMean) (linearKEMean - rhoMMean
 * UMean.UMean / 2), p = k rhoNMean translationalT.  As in the reference the derived fields
 * are computed only when every cell has rhoNMean and linearKEMean above 1e-300 (*derived =
 * 1), else they are zero (*derived = 0).  The wall pressure from the force density (fD) is
 * not built */
int pic_fields_mean(pic_ctx* ctx, int32_t species, double* out, int32_t* derived);
int pic_step(pic_ctx* ctx, double dt, pic_step_stats* stats);  /* one PICCloud::evolve() */
/* n_steps evolve() calls, the time loop of picFoam.C, with one host synchronisation at the
 * end: stats sum the counters over the steps (poisson: the last solve); an error raised on
 * the device (superluminal parcel, capacity, cell overflow) is reported at the end and
 * leaves the state of the step it happened in */
int pic_run(pic_ctx* ctx, double dt, int64_t n_steps, pic_step_stats* stats);
int pic_sync(pic_ctx* ctx);
int64_t pic_step_index(const pic_ctx* ctx);
int pic_set_step_index(pic_ctx* ctx, int64_t step);

/* device timing of the stages run since the last reset (HIP events on the context stream).
 * stages: bit (1 << PIC_STAGE_x) times stage x, -1 every stage, 0 none.  Each timed stage
 * adds two event records to the stream (a few microseconds per step and stage) */
int pic_profile_enable(pic_ctx* ctx, int32_t stages);
int pic_stage_times(pic_ctx* ctx, double* ms /*[PIC_STAGE_COUNT]*/, int64_t* launches /*[PIC_STAGE_COUNT]*/);
int pic_profile_reset(pic_ctx* ctx);

/* learned initial guess for the Poisson PCG (no counterpart in the reference): a
 * message-passing network over the cell graph with BF16 MFMA kernels, trained on the
 * device on (rho, phi_prev) -> phi snapshots the solver itself records */
int pic_gnn_init(pic_ctx* ctx, int32_t layers, int32_t width, uint64_t seed);
int pic_gnn_n_params(pic_ctx* ctx, int64_t* n);
int pic_gnn_params(pic_ctx* ctx, int64_t n, float* get, const float* set);
int pic_gnn_record(pic_ctx* ctx, int32_t capacity);   /* record every Poisson solve into a ring buffer */
int pic_gnn_snapshot_count(pic_ctx* ctx, int32_t* n);
int pic_gnn_snapshot(pic_ctx* ctx, int32_t index, float* features, float* target, double* norm);
int pic_gnn_grad(pic_ctx* ctx, int32_t index, float* grad, double* loss);
int pic_gnn_train(pic_ctx* ctx, int32_t steps, double lr, uint64_t seed, double* loss_first, double* loss_last);
int pic_gnn_eval(pic_ctx* ctx, int32_t index, double* loss);
int pic_gnn_use(pic_ctx* ctx, int32_t on);            /* start the PCG from the network's guess (its
                                                          inference runs inside pic_poisson_solve) */
/* training loss: 0 = A-norm error e^T A e / d^T A d, 1 = residual |A e|^2 / |A d|^2, 2 = their sum */
int pic_gnn_loss(pic_ctx* ctx, int32_t kind);

#ifdef __cplusplus
}
#endif
#endif
