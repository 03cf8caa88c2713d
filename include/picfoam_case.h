/*
 * picfoam_case.h — the case layer of libpicmi: a picFoam case directory read into the
 * structures of picmi.h, the picInitialise models, and the lagrangian and field files
 * written back in OpenFOAM's ASCII format.
 *
 * The reference builds these from OpenFOAM-9's Time, fvMesh and IOdictionary classes
 * (solver/picFoam/createFields.H, preProcess/picInitialise/picInitialise.C); here they are
 * plain C++ inside the library, behind the same kind of C-ABI as the particle step.  A case
 * holds host data only: pic_case_open, pic_case_initialise and pic_case_read_parcels run
 * without a GPU; pic_case_create hands everything to pic_create and the device.
 *
 * Files read (ASCII):
 *   system/controlDict       startTime, endTime, deltaT, writeInterval (writeControl
 *                            timeStep or runTime/adjustableRunTime)
 *   constant/picProperties   SolverSettings, moleculeProperties, MaxwellSolver,
 *                            ParticlePusher, CollisionModels (PICCloud.C:57-317, 380-520)
 *   constant/polyMesh/       points, faces, owner, neighbour, boundary; without them,
 *                            system/blockMeshDict with one hex block (blockMesh)
 *   <startTime>/phiE, E, B   patch conditions of phiE (fixedValue / zeroGradient / cyclic /
 *                            empty) and the uniform or per-cell E and B
 *   system/picInitialiseDict picInitialise: temperature, numberDensities, electronVolt,
 *                            InitializationModel (UniformDisplacement, Maxwellian,
 *                            QuasineutralPlasma, Equipartition), solveMaxwellEquations,
 *                            initalizeLeapFrog
 *   <time>/lagrangian/<cloud>/ positions or coordinates, U, typeId, nParticle, chargeModifier
 *                            (PICParcelIO.C:83-162)
 *
 * Status codes are those of picmi.h; pic_case_error() holds the message of the last failure.
 */
#ifndef PICFOAM_CASE_H
#define PICFOAM_CASE_H
#include <stdint.h>

#include "picmi.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pic_case pic_case;

/* Reads the case directory (createTime + createMesh + the PICCloud constructor's
 * dictionaries).  On failure *out is still set when the case object exists, so that
 * pic_case_error can report; pic_case_close frees it. */
int pic_case_open(const char* case_dir, pic_case** out);
void pic_case_close(pic_case* cs);
const char* pic_case_error(const pic_case* cs);

/* Sizes and times of the case. */
int pic_case_sizes(const pic_case* cs, int32_t* n_species, int32_t* n_cells, int32_t* n_patches);
int pic_case_times(const pic_case* cs, double* start_time, double* end_time, double* delta_t,
                   double* write_interval);
/* SolverSettings.typeIdList[species] */
const char* pic_case_species_name(const pic_case* cs, int32_t species);
/* constant/polyMesh/boundary name of a patch */
const char* pic_case_patch_name(const pic_case* cs, int32_t patch);

/* The case's PatchBoundaryModels of type SimpleInjection (BoundaryModelList.C:104-167,
 * SimpleInjection.C:34-175) as pic_create's context gets them through pic_set_injection:
 * up to `capacity` entries into out, their number into *n.  Other patch boundary models are
 * refused when the case is opened. */
int pic_case_injection(const pic_case* cs, int32_t capacity, pic_injection* out, int32_t* n);

/* The structures pic_create takes, as the case defines them.  The pointers inside *mesh
 * and the model names inside *models stay valid until pic_case_close.  species[] holds
 * n_species entries; its capacities are the parcel counts the case holds so far. */
int pic_case_describe(const pic_case* cs, pic_mesh* mesh, pic_species* species, pic_models* models);

/* picInitialise (picInitialise.C:45-316): runs the InitializationModel of
 * system/<dict_name> (NULL: picInitialiseDict) into the case's host parcels, replacing any
 * read before.  seed keys the counter-based draws (the reference's serial Random stream
 * cannot be reproduced). */
int pic_case_initialise(pic_case* cs, const char* dict_name, uint64_t seed);

/* PICParcel::readFields: the parcels of <time_name>/lagrangian/<cloud>/. */
int pic_case_read_parcels(pic_case* cs, const char* time_name);

/* Host parcels of one species: count, then positions [n*3], cells [n] and U [n*3]
 * (capacity = the room in each array, in parcels). */
int pic_case_parcel_count(const pic_case* cs, int32_t species, int64_t* n);
int pic_case_parcels(const pic_case* cs, int32_t species, int64_t capacity, double* x, int32_t* cell, double* U);

/* pic_create with the case's mesh, species and models (precision and the parcel capacity
 * factor chosen here), the fields of the start time, and the host parcels uploaded.  After
 * pic_case_initialise it also does picInitialise's last steps: the charge density, the
 * Poisson solve when solveMaxwellEquations is set, and the half step back of the velocities
 * when initalizeLeapFrog is set (picInitialise.C:213-237). */
int pic_case_create(pic_case* cs, int32_t precision, double capacity_factor, pic_ctx** out);

/* Writes <time_name>/lagrangian/<cloud>/{positions,coordinates,U,typeId,nParticle,
 * chargeModifier} and <time_name>/{phiE,E,rhoCharge} (PICParcel::writeFields and the
 * cloud's fields) from the device state. */
int pic_case_write(pic_case* cs, pic_ctx* ctx, const char* time_name);

/* The picFoam time loop (picFoam.C): steps of deltaT from the case's current time up to
 * until_time (<= 0: endTime), writing at every writeInterval; *steps counts the steps run. */
int pic_case_run(pic_case* cs, pic_ctx* ctx, double until_time, int64_t* steps);

/* The case's current time and its OpenFOAM directory name (timeFormat general,
 * timePrecision 6). */
double pic_case_time(const pic_case* cs);
const char* pic_case_time_name(const pic_case* cs);

#ifdef __cplusplus
}
#endif
#endif
