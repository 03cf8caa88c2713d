/*
 * picref.c — FP64 CPU oracle for picFoam's per-timestep particle loop.
 *
 * TEST INFRASTRUCTURE ONLY (see picref.h).  Plain C, compiled with -ffp-contract=off so
 * that every product and sum is rounded as written.
 *
 * Parity status per function (the tests that pin each one are in tests/test_oracle.py and
 * tests/test_oracle_pins.py):
 *   geometry, tets, locate, positions ............ pinned: closed-form volumes/centres,
 *                                                    barycentric round trips
 *   face base points ............................. pinned: point 0 on convex meshes; a folded
 *                                                    face against a brute-force positive-volume search
 *   pushers ...................................... pinned: |u| conservation, E-only kick,
 *                                                    gyro period/radius, ExB drift, Boris=HC=Vay (v<<c)
 *   tracking ..................................... pinned: straight-line endpoint x+U*dt with
 *                                                    periodic wrap, brute-force point location
 *   interpolation / deposit ...................... pinned: linear field reproduced exactly,
 *                                                    total charge, hand-computed single particles,
 *                                                    VolumeWeighting distance weights by hand
 *   cell->point interpolation .................... pinned: constant/linear fields on interior points
 *   Poisson PCG/DIC + grad ....................... pinned: dense solve, manufactured sin(kx) at
 *                                                    second order; the non-orthogonal correction by
 *                                                    second-order convergence on a sheared lattice
 *                                                    (the implicit part alone does not converge)
 *   Debye length ................................. pinned: closed form sqrt(eps0 kT/(n q^2)) summed
 *                                                    over species, with drift and the Wigner-Seitz floor
 *   Coulomb (Nanbu + TA pairing) ................. pinned: momentum/energy conservation per pair,
 *                                                    <cos chi> = exp(-s) in every branch of the
 *                                                    sampler, TA collisions per parcel and odd triples
 *   Perez Coulomb ................................ pinned: relativistic 4-momentum conservation,
 *                                                    <cos chi> = exp(-s) and Nanbu's angle at v<<c
 *   Nanbu-Yonemura pairing, Nanbu weight
 *   correction ................................... pinned: the reference's collision counts;
 *                                                    keep probability w_P/w_Q (binomial test)
 *   MCC (null collision) ......................... pinned: collision frequency n*sigma*v,
 *                                                    candidate count, elastic energy loss 2m/M,
 *                                                    excitation/ionization energy bookkeeping, the
 *                                                    Opal split, pair probability w_E/w_N,
 *                                                    background Maxwellian width kT/M
 *   relativistic e-n scattering .................. pinned: 4-momentum conservation, Nanbu's
 *                                                    cos chi at v<<c
 *   ion-neutral MCC (background) ................. pinned: candidate count and remainder, full
 *                                                    acceptance, elastic share, charge exchange,
 *                                                    isotropic scattering's conservation laws
 *   binary DSMC (pairs, sub-cells, background) ... pinned: HardSphere / VariableHardSphere closed
 *                                                    forms, candidate count, sub-cell partners,
 *                                                    the closed-form hard-sphere two-temperature
 *                                                    rate, background rates for HS and VHS
 *   ion-neutral parcel pairs ..................... pinned: candidate count, acceptance 1/2,
 *                                                    charge-exchange multiset, conservation,
 *                                                    updateNeutralVelocity false
 *   Philox4x32-10 ................................ pinned: Random123 known-answer vectors
 *   OpenFOAM-9 Random/shuffle streams, tracking tie-breaks, volPointInterpolation at
 *   non-coupled boundary points
 *                                                  parity unpinned (not reproducible offline)
 */
#ifdef _OPENMP
#include <omp.h>
#endif
#include "picref.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define C_LIGHT 299792458.0
#define EPS0 8.8541878128e-12
#define QE 1.602176634e-19
#define KB 1.380649e-23
#define HPLANCK 6.62607015e-34
#define PI_D 3.14159265358979323846
#define MAX_TRACK_STEPS 1000
#define REF_EMPTY_FACE (-900000)

typedef struct { double x, y, z; } v3;
static inline v3 V(double x, double y, double z) { v3 r = {x, y, z}; return r; }
static inline v3 vadd(v3 a, v3 b) { return V(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 vsub(v3 a, v3 b) { return V(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 vscale(double s, v3 a) { return V(s * a.x, s * a.y, s * a.z); }
static inline v3 vdiv(v3 a, double s) { return V(a.x / s, a.y / s, a.z / s); }
static inline double vdot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline v3 vcross(v3 a, v3 b) { return V(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
static inline double vmag(v3 a) { return sqrt(vdot(a, a)); }
static inline v3 ld3(const double* p) { return V(p[0], p[1], p[2]); }
static inline void st3(double* p, v3 a) { p[0] = a.x; p[1] = a.y; p[2] = a.z; }

struct ref_mesh {
    int n_points, n_faces, n_internal, n_cells, n_patches;
    double* points;
    int *fo, *fp, *owner, *nbr;
    int *ptype, *pstart, *psize, *pnbr;
    double *Cf, *Sf, *C, *V;
    int *cfo, *cf;               /* cell -> faces, OpenFOAM primitiveMesh::calcCells order */
    int n_tets;
    int *tet_cell, *tet_face, *tet_pt, *cell_tet0;
    int* fbase;                  /* face base point, an index into the face's point list */
    int* pmaster;                /* cyclic-coupled point groups -> smallest point id */
    int* face_patch;             /* boundary face -> patch index (-1 for internal) */
    double* psep;                /* per patch: x_partner = x + sep */
    int track_mask;              /* polyMesh::solutionD: bit d set for an empty direction */
    /* VolumeWeighting (ChargeDistribution) weights, per cell-point (ascending point id) */
    int *cpo, *cp;
    double* vw_w;
    /* volPointInterpolation inverse-distance weights, same cell-point layout */
    double* ip_w;
};

/* ------------------------------------------------------------------ geometry */

/* OpenFOAM-9 primitiveMeshFaceCentresAndAreas: triangles direct, polygons by a fan around
 * the point average with |n|-weighted triangle centroids. */
static void face_geometry(ref_mesh* m) {
    for (int f = 0; f < m->n_faces; ++f) {
        int s = m->fo[f], n = m->fo[f + 1] - s;
        const int* fp = m->fp + s;
        if (n == 3) {
            v3 p0 = ld3(m->points + 3 * fp[0]), p1 = ld3(m->points + 3 * fp[1]), p2 = ld3(m->points + 3 * fp[2]);
            st3(m->Cf + 3 * f, vscale(1.0 / 3.0, vadd(vadd(p0, p1), p2)));
            st3(m->Sf + 3 * f, vscale(0.5, vcross(vsub(p1, p0), vsub(p2, p0))));
        } else {
            v3 est = V(0, 0, 0);
            for (int i = 0; i < n; ++i) est = vadd(est, ld3(m->points + 3 * fp[i]));
            est = vdiv(est, (double)n);
            v3 sumN = V(0, 0, 0), sumAc = V(0, 0, 0);
            double sumA = 0.0;
            for (int i = 0; i < n; ++i) {
                v3 p = ld3(m->points + 3 * fp[i]), q = ld3(m->points + 3 * fp[(i + 1) % n]);
                v3 c = vadd(vadd(p, q), est);
                v3 nn = vcross(vsub(q, p), vsub(est, p));
                double a = vmag(nn);
                sumN = vadd(sumN, nn);
                sumA += a;
                sumAc = vadd(sumAc, vscale(a, c));
            }
            if (sumA < 1e-300) {
                st3(m->Cf + 3 * f, est);
                st3(m->Sf + 3 * f, V(0, 0, 0));
            } else {
                st3(m->Cf + 3 * f, vscale(1.0 / 3.0, vdiv(sumAc, sumA)));
                st3(m->Sf + 3 * f, vscale(0.5, sumN));
            }
        }
    }
}

/* OpenFOAM-9 primitiveMeshCellCentresAndVols: pyramid decomposition about the face-centre
 * average. */
static void cell_geometry(ref_mesh* m) {
    int nc = m->n_cells;
    v3* est = (v3*)calloc(nc, sizeof(v3));
    int* nf = (int*)calloc(nc, sizeof(int));
    for (int f = 0; f < m->n_faces; ++f) {
        est[m->owner[f]] = vadd(est[m->owner[f]], ld3(m->Cf + 3 * f));
        nf[m->owner[f]]++;
    }
    for (int f = 0; f < m->n_internal; ++f) {
        est[m->nbr[f]] = vadd(est[m->nbr[f]], ld3(m->Cf + 3 * f));
        nf[m->nbr[f]]++;
    }
    for (int c = 0; c < nc; ++c) est[c] = vdiv(est[c], (double)nf[c]);
    v3* ctr = (v3*)calloc(nc, sizeof(v3));
    double* vol = (double*)calloc(nc, sizeof(double));
    for (int f = 0; f < m->n_faces; ++f) {
        int o = m->owner[f];
        v3 cf = ld3(m->Cf + 3 * f), sf = ld3(m->Sf + 3 * f);
        double pv = vdot(sf, vsub(cf, est[o]));
        if (pv < 1e-300) pv = 1e-300;
        v3 pc = vadd(vscale(0.75, cf), vscale(0.25, est[o]));
        ctr[o] = vadd(ctr[o], vscale(pv, pc));
        vol[o] += pv;
    }
    for (int f = 0; f < m->n_internal; ++f) {
        int o = m->nbr[f];
        v3 cf = ld3(m->Cf + 3 * f), sf = ld3(m->Sf + 3 * f);
        double pv = vdot(sf, vsub(est[o], cf));
        if (pv < 1e-300) pv = 1e-300;
        v3 pc = vadd(vscale(0.75, cf), vscale(0.25, est[o]));
        ctr[o] = vadd(ctr[o], vscale(pv, pc));
        vol[o] += pv;
    }
    for (int c = 0; c < nc; ++c) {
        st3(m->C + 3 * c, vdiv(ctr[c], vol[c]));
        m->V[c] = vol[c] / 3.0;
    }
    free(est); free(nf); free(ctr); free(vol);
}

/* tetrahedron::quality (OpenFOAM-9): signed volume over (8/(9 sqrt 3)) R^3, R the
 * circumradius from the circumcentre offset (|u|^2 v x w + |v|^2 w x u + |w|^2 u x v) /
 * (2 u . v x w) of the edges u, v, w from a */
static double tet_quality(v3 a, v3 b, v3 c, v3 d) {
    v3 u = vsub(b, a), v = vsub(c, a), w = vsub(d, a);
    double t6 = vdot(u, vcross(v, w));
    if (t6 == 0.0) return 0.0;
    v3 o = vadd(vadd(vscale(vdot(u, u), vcross(v, w)), vscale(vdot(v, v), vcross(w, u))), vscale(vdot(w, w), vcross(u, v)));
    double R = vmag(vscale(1.0 / (2.0 * t6), o));
    return (t6 / 6.0) / (8.0 / (9.0 * sqrt(3.0)) * R * R * R + 1e-150);
}

/* polyMeshTetDecomposition::findSharedBasePoint / findBasePoint (OpenFOAM-9): the first
 * face point from which the minimum fan-tet quality with the owner centre (and the
 * neighbour centre nc, A and B swapped, when given) exceeds minTetQuality = 1e-15 */
static int face_base_point(const ref_mesh* m, int f, v3 oc, const v3* nc) {
    const int* fp = m->fp + m->fo[f];
    int n = m->fo[f + 1] - m->fo[f];
    for (int b = 0; b < n; ++b) {
        v3 pb = ld3(m->points + 3 * fp[b]);
        double q = 1e300;
        for (int i = 1; i <= n - 2; ++i) {
            v3 pA = ld3(m->points + 3 * fp[(b + i) % n]), pB = ld3(m->points + 3 * fp[(b + i + 1) % n]);
            double qo = tet_quality(oc, pb, pA, pB);
            if (qo < q) q = qo;
            if (nc) { double qn = tet_quality(*nc, pb, pB, pA); if (qn < q) q = qn; }
        }
        if (q > 1e-15) return b;
    }
    return -1;
}

static int uf_find(int* p, int a) {
    while (p[a] != a) { p[a] = p[p[a]]; a = p[a]; }
    return a;
}

ref_mesh* ref_mesh_create(int n_points, const double* points, int n_faces, const int* face_offsets,
                          const int* face_points, const int* owner, int n_internal, const int* neighbour,
                          int n_cells, int n_patches, const int* patch_type, const int* patch_start,
                          const int* patch_size, const int* patch_nbr) {
    ref_mesh* m = (ref_mesh*)calloc(1, sizeof(ref_mesh));
    m->n_points = n_points; m->n_faces = n_faces; m->n_internal = n_internal; m->n_cells = n_cells;
    m->n_patches = n_patches;
    m->points = (double*)malloc(sizeof(double) * 3 * n_points);
    memcpy(m->points, points, sizeof(double) * 3 * n_points);
    m->fo = (int*)malloc(sizeof(int) * (n_faces + 1));
    memcpy(m->fo, face_offsets, sizeof(int) * (n_faces + 1));
    m->fp = (int*)malloc(sizeof(int) * face_offsets[n_faces]);
    memcpy(m->fp, face_points, sizeof(int) * face_offsets[n_faces]);
    m->owner = (int*)malloc(sizeof(int) * n_faces);
    memcpy(m->owner, owner, sizeof(int) * n_faces);
    m->nbr = (int*)malloc(sizeof(int) * (n_internal > 0 ? n_internal : 1));
    if (n_internal) memcpy(m->nbr, neighbour, sizeof(int) * n_internal);
    m->ptype = (int*)malloc(sizeof(int) * (n_patches + 1));
    m->pstart = (int*)malloc(sizeof(int) * (n_patches + 1));
    m->psize = (int*)malloc(sizeof(int) * (n_patches + 1));
    m->pnbr = (int*)malloc(sizeof(int) * (n_patches + 1));
    memcpy(m->ptype, patch_type, sizeof(int) * n_patches);
    memcpy(m->pstart, patch_start, sizeof(int) * n_patches);
    memcpy(m->psize, patch_size, sizeof(int) * n_patches);
    memcpy(m->pnbr, patch_nbr, sizeof(int) * n_patches);
    m->Cf = (double*)calloc(3 * n_faces, sizeof(double));
    m->Sf = (double*)calloc(3 * n_faces, sizeof(double));
    m->C = (double*)calloc(3 * n_cells, sizeof(double));
    m->V = (double*)calloc(n_cells, sizeof(double));
    face_geometry(m);
    cell_geometry(m);

    /* cells(): owned faces in face order, then neighboured faces in face order */
    int* cnt = (int*)calloc(n_cells + 1, sizeof(int));
    for (int f = 0; f < n_faces; ++f) cnt[owner[f]]++;
    for (int f = 0; f < n_internal; ++f) cnt[neighbour[f]]++;
    m->cfo = (int*)malloc(sizeof(int) * (n_cells + 1));
    m->cfo[0] = 0;
    for (int c = 0; c < n_cells; ++c) m->cfo[c + 1] = m->cfo[c] + cnt[c];
    m->cf = (int*)malloc(sizeof(int) * m->cfo[n_cells]);
    memset(cnt, 0, sizeof(int) * (n_cells + 1));
    for (int f = 0; f < n_faces; ++f) { int c = owner[f]; m->cf[m->cfo[c] + cnt[c]++] = f; }
    for (int f = 0; f < n_internal; ++f) { int c = neighbour[f]; m->cf[m->cfo[c] + cnt[c]++] = f; }
    free(cnt);

    /* tets: (cell, face, pt), pt = 1..n-2 (face base point 0) */
    int nt = 0;
    m->cell_tet0 = (int*)malloc(sizeof(int) * (n_cells + 1));
    for (int c = 0; c < n_cells; ++c) {
        m->cell_tet0[c] = nt;
        for (int k = m->cfo[c]; k < m->cfo[c + 1]; ++k) { int f = m->cf[k]; nt += m->fo[f + 1] - m->fo[f] - 2; }
    }
    m->cell_tet0[n_cells] = nt;
    m->n_tets = nt;
    m->tet_cell = (int*)malloc(sizeof(int) * nt);
    m->tet_face = (int*)malloc(sizeof(int) * nt);
    m->tet_pt = (int*)malloc(sizeof(int) * nt);
    nt = 0;
    for (int c = 0; c < n_cells; ++c)
        for (int k = m->cfo[c]; k < m->cfo[c + 1]; ++k) {
            int f = m->cf[k], n = m->fo[f + 1] - m->fo[f];
            for (int pt = 1; pt <= n - 2; ++pt) { m->tet_cell[nt] = c; m->tet_face[nt] = f; m->tet_pt[nt] = pt; ++nt; }
        }

    /* boundary face -> patch, cyclic separations, coupled point groups */
    m->face_patch = (int*)malloc(sizeof(int) * n_faces);
    for (int f = 0; f < n_faces; ++f) m->face_patch[f] = -1;
    m->psep = (double*)calloc(3 * (n_patches + 1), sizeof(double));
    m->pmaster = (int*)malloc(sizeof(int) * n_points);
    for (int p = 0; p < n_points; ++p) m->pmaster[p] = p;
    /* polyMesh::calcDirections: summed |area components| of the empty patches, normalised;
     * a component above 1e-6 marks an empty direction */
    {
        double ed[3] = {0.0, 0.0, 0.0};
        int any = 0;
        for (int pi = 0; pi < n_patches; ++pi) {
            if (m->ptype[pi] != REF_PATCH_EMPTY) continue;
            for (int f = m->pstart[pi]; f < m->pstart[pi] + m->psize[pi]; ++f)
                for (int d = 0; d < 3; ++d) ed[d] += fabs(m->Sf[3 * f + d]);
            if (m->psize[pi] > 0) any = 1;
        }
        double e = sqrt(ed[0] * ed[0] + ed[1] * ed[1] + ed[2] * ed[2]);
        m->track_mask = 0;
        if (any && e > 0.0)
            for (int d = 0; d < 3; ++d) if (ed[d] / e > 1e-6) m->track_mask |= 1 << d;
    }
    for (int pi = 0; pi < n_patches; ++pi) {
        for (int f = m->pstart[pi]; f < m->pstart[pi] + m->psize[pi]; ++f) m->face_patch[f] = pi;
        if (m->ptype[pi] == REF_PATCH_CYCLIC && m->psize[pi] > 0) {
            int q = m->pnbr[pi];
            int f0 = m->pstart[pi], g0 = m->pstart[q];
            st3(m->psep + 3 * pi, vsub(ld3(m->Cf + 3 * g0), ld3(m->Cf + 3 * f0)));
            for (int k = 0; k < m->psize[pi]; ++k) {
                int f = f0 + k, g = g0 + k, n = m->fo[f + 1] - m->fo[f];
                /* cyclic ordering: g[0] ~ f[0], g[j] ~ f[n-j]; confirmed by position */
                for (int j = 0; j < n; ++j) {
                    int a = m->fp[m->fo[f] + j];
                    v3 img = vadd(ld3(m->points + 3 * a), ld3(m->psep + 3 * pi));
                    int best = -1; double bd = 1e300;
                    for (int i = 0; i < n; ++i) {
                        int b = m->fp[m->fo[g] + i];
                        double d = vmag(vsub(ld3(m->points + 3 * b), img));
                        if (d < bd) { bd = d; best = b; }
                    }
                    int ra = uf_find(m->pmaster, a), rb = uf_find(m->pmaster, best);
                    if (ra != rb) { if (ra < rb) m->pmaster[rb] = ra; else m->pmaster[ra] = rb; }
                }
            }
        }
    }
    for (int p = 0; p < n_points; ++p) m->pmaster[p] = uf_find(m->pmaster, p);

    /* face base points (polyMeshTetDecomposition::findFaceBasePts, OpenFOAM-9): internal
     * faces with both cell centres; on a cyclic pair the lower-numbered patch with the
     * partner cell's centre carried across by the separation, and the partner face with the
     * matching point (point j of one face is point n - j of the other); other boundary faces
     * with the owner alone.  No qualifying point: point 0 */
    m->fbase = (int*)malloc(sizeof(int) * n_faces);
    for (int f = 0; f < n_faces; ++f) {
        v3 oc = ld3(m->C + 3 * owner[f]);
        m->fbase[f] = -1;
        if (f < n_internal) {
            v3 nc = ld3(m->C + 3 * neighbour[f]);
            m->fbase[f] = face_base_point(m, f, oc, &nc);
        } else {
            int pi = m->face_patch[f];
            if (m->ptype[pi] == REF_PATCH_CYCLIC) {
                int q = m->pnbr[pi];
                if (pi < q) {
                    int g = m->pstart[q] + (f - m->pstart[pi]);
                    v3 nc = vsub(ld3(m->C + 3 * owner[g]), ld3(m->psep + 3 * pi));
                    m->fbase[f] = face_base_point(m, f, oc, &nc);
                }
            } else {
                m->fbase[f] = face_base_point(m, f, oc, NULL);
            }
        }
    }
    for (int f = n_internal; f < n_faces; ++f) {
        int pi = m->face_patch[f];
        if (m->ptype[pi] != REF_PATCH_CYCLIC || pi < m->pnbr[pi]) continue;
        int q = m->pnbr[pi], g = m->pstart[q] + (f - m->pstart[pi]), n = m->fo[f + 1] - m->fo[f];
        m->fbase[f] = m->fbase[g] < 0 ? -1 : (n - m->fbase[g]) % n;
    }
    for (int f = 0; f < n_faces; ++f) if (m->fbase[f] < 0) m->fbase[f] = 0;

    /* cellPoints (ascending point id) + weights */
    m->cpo = (int*)malloc(sizeof(int) * (n_cells + 1));
    int cap = 0;
    for (int c = 0; c < n_cells; ++c)
        for (int k = m->cfo[c]; k < m->cfo[c + 1]; ++k) { int f = m->cf[k]; cap += m->fo[f + 1] - m->fo[f]; }
    m->cp = (int*)malloc(sizeof(int) * (cap + 1));
    int ncp = 0;
    int* tmp = (int*)malloc(sizeof(int) * 512);
    for (int c = 0; c < n_cells; ++c) {
        m->cpo[c] = ncp;
        int nt2 = 0;
        for (int k = m->cfo[c]; k < m->cfo[c + 1]; ++k) {
            int f = m->cf[k];
            for (int j = m->fo[f]; j < m->fo[f + 1]; ++j) {
                int p = m->fp[j], dup = 0;
                for (int i = 0; i < nt2; ++i) if (tmp[i] == p) { dup = 1; break; }
                if (!dup) tmp[nt2++] = p;
            }
        }
        /* insertion sort ascending */
        for (int i = 1; i < nt2; ++i) { int v = tmp[i], j = i - 1; while (j >= 0 && tmp[j] > v) { tmp[j + 1] = tmp[j]; --j; } tmp[j + 1] = v; }
        for (int i = 0; i < nt2; ++i) m->cp[ncp++] = tmp[i];
    }
    m->cpo[n_cells] = ncp;
    free(tmp);
    m->vw_w = (double*)malloc(sizeof(double) * (ncp + 1));
    m->ip_w = (double*)malloc(sizeof(double) * (ncp + 1));
    double* sumd = (double*)calloc(n_points, sizeof(double));
    double* sumi = (double*)calloc(n_points, sizeof(double));
    for (int c = 0; c < n_cells; ++c)
        for (int k = m->cpo[c]; k < m->cpo[c + 1]; ++k) {
            int p = m->cp[k];
            double d = vmag(vsub(ld3(m->points + 3 * p), ld3(m->C + 3 * c)));
            m->vw_w[k] = d;          /* VolumeWeighting.C:66 distance weight */
            m->ip_w[k] = 1.0 / d;    /* volPointInterpolation inverse distance */
            sumd[m->pmaster[p]] += d;
            sumi[m->pmaster[p]] += 1.0 / d;
        }
    for (int c = 0; c < n_cells; ++c)
        for (int k = m->cpo[c]; k < m->cpo[c + 1]; ++k) {
            int p = m->pmaster[m->cp[k]];
            m->vw_w[k] /= sumd[p];
            m->ip_w[k] /= sumi[p];
        }
    free(sumd); free(sumi);
    return m;
}

void ref_mesh_destroy(ref_mesh* m) {
    if (!m) return;
    free(m->points); free(m->fo); free(m->fp); free(m->owner); free(m->nbr);
    free(m->ptype); free(m->pstart); free(m->psize); free(m->pnbr);
    free(m->Cf); free(m->Sf); free(m->C); free(m->V); free(m->cfo); free(m->cf);
    free(m->tet_cell); free(m->tet_face); free(m->tet_pt); free(m->cell_tet0);
    free(m->pmaster); free(m->face_patch); free(m->psep); free(m->fbase);
    free(m->cpo); free(m->cp); free(m->vw_w); free(m->ip_w);
    free(m);
}

int ref_mesh_n_tets(const ref_mesh* m) { return m->n_tets; }
int ref_mesh_n_points(const ref_mesh* m) { return m->n_points; }

void ref_mesh_geometry(const ref_mesh* m, double* cc, double* cv, double* fc, double* fa) {
    if (cc) memcpy(cc, m->C, sizeof(double) * 3 * m->n_cells);
    if (cv) memcpy(cv, m->V, sizeof(double) * m->n_cells);
    if (fc) memcpy(fc, m->Cf, sizeof(double) * 3 * m->n_faces);
    if (fa) memcpy(fa, m->Sf, sizeof(double) * 3 * m->n_faces);
}

void ref_mesh_tets(const ref_mesh* m, int* tc, int* tf, int* tp) {
    memcpy(tc, m->tet_cell, sizeof(int) * m->n_tets);
    memcpy(tf, m->tet_face, sizeof(int) * m->n_tets);
    memcpy(tp, m->tet_pt, sizeof(int) * m->n_tets);
}

void ref_mesh_point_master(const ref_mesh* m, int* pm) { memcpy(pm, m->pmaster, sizeof(int) * m->n_points); }

void ref_mesh_face_base(const ref_mesh* m, int* base) { memcpy(base, m->fbase, sizeof(int) * m->n_faces); }

/* tetIndices::faceTriIs (OpenFOAM-9): (base, A, B) = (f[b], f[b+pt], f[b+pt+1]) around the
 * face from its base point b, A and B swapped when the cell is the face's neighbour. */
static void tet_points(const ref_mesh* m, int t, int vid[3]) {
    int f = m->tet_face[t], pt = m->tet_pt[t], c = m->tet_cell[t];
    const int* fp = m->fp + m->fo[f];
    int n = m->fo[f + 1] - m->fo[f], b = m->fbase[f];
    vid[0] = fp[b];
    if (m->owner[f] == c) { vid[1] = fp[(b + pt) % n]; vid[2] = fp[(b + pt + 1) % n]; }
    else { vid[1] = fp[(b + pt + 1) % n]; vid[2] = fp[(b + pt) % n]; }
}

static void tet_vertices(const ref_mesh* m, int t, v3 x[4], int vid[3]) {
    tet_points(m, t, vid);
    x[0] = ld3(m->C + 3 * m->tet_cell[t]);
    for (int i = 0; i < 3; ++i) x[i + 1] = ld3(m->points + 3 * vid[i]);
}

void ref_positions(const ref_mesh* m, int n, const double* y, const int* tet, double* x) {
    for (int i = 0; i < n; ++i) {
        if (tet[i] < 0) { x[3 * i] = x[3 * i + 1] = x[3 * i + 2] = NAN; continue; }
        v3 X[4]; int vid[3];
        tet_vertices(m, tet[i], X, vid);
        double y1 = y[3 * i], y2 = y[3 * i + 1], y3 = y[3 * i + 2];
        double y0 = 1.0 - (y1 + y2 + y3);
        v3 p = vadd(vadd(vscale(y0, X[0]), vscale(y1, X[1])), vadd(vscale(y2, X[2]), vscale(y3, X[3])));
        st3(x + 3 * i, p);
    }
}

/* barycentric coordinates of p in tet X: y_i = r_i.(p - x0)/det, r1 = e2 x e3, ... */
static void bary_of(const v3 X[4], v3 p, double y[4]) {
    v3 e1 = vsub(X[1], X[0]), e2 = vsub(X[2], X[0]), e3 = vsub(X[3], X[0]);
    v3 r1 = vcross(e2, e3), r2 = vcross(e3, e1), r3 = vcross(e1, e2);
    double det = vdot(e1, r1);
    v3 d = vsub(p, X[0]);
    y[1] = vdot(r1, d) / det; y[2] = vdot(r2, d) / det; y[3] = vdot(r3, d) / det;
    y[0] = 1.0 - (y[1] + y[2] + y[3]);
}

int ref_locate(const ref_mesh* m, int n, const double* x, const int* cell, double* y, int* tet) {
    int fails = 0;
    for (int i = 0; i < n; ++i) {
        int c = cell[i], best = -1;
        double bestmin = -1e300, by[4] = {0, 0, 0, 0};
        v3 p = ld3(x + 3 * i);
        for (int t = m->cell_tet0[c]; t < m->cell_tet0[c + 1]; ++t) {
            v3 X[4]; int vid[3]; double yy[4];
            tet_vertices(m, t, X, vid);
            bary_of(X, p, yy);
            double mn = yy[0];
            for (int k = 1; k < 4; ++k) if (yy[k] < mn) mn = yy[k];
            if (mn > bestmin) { bestmin = mn; best = t; memcpy(by, yy, sizeof by); }
        }
        if (bestmin < -1e-9) ++fails;
        tet[i] = best;
        y[3 * i] = by[1]; y[3 * i + 1] = by[2]; y[3 * i + 2] = by[3];
    }
    return fails;
}

/* ------------------------------------------------------------------ pushers */

/* BorisPusher.C:69-112 (relativistic; rotation Lorentz factor = mean of the old and the
 * half-kicked gamma, :98), BorisNRPusher.C:69-99, VayPusher.C:54-100, HigueraCary.C:69-115. */
void ref_push_one(int pusher, double* Up, const double* Ep, const double* Bp, double q, double mass, double dt,
                  int hc_reference_form, int rotate) {
    v3 U = ld3(Up), E = ld3(Ep), B = ld3(Bp);
    const double c = C_LIGHT;
    double dt2 = 0.5 * dt;
    if (pusher == REF_PUSH_NONE) return;
    if (pusher == REF_PUSH_BORIS) {
        double mu = vmag(U) / c;
        v3 rU = vdiv(U, sqrt(1.0 - mu * mu));
        v3 oldrU = rU;
        double k = q * dt2 / mass;
        v3 rUsm = vscale(k, E);
        rU = vadd(rU, rUsm);
        if (rotate) {
            double a = vmag(oldrU) / c, b = vmag(rU) / c;
            double lF = (sqrt(1.0 + a * a) + sqrt(1.0 + b * b)) / 2.0;
            v3 t = vdiv(vscale(k, B), lF);
            v3 Ud = vadd(rU, vcross(rU, t));
            rU = vadd(rU, vcross(Ud, vdiv(vscale(2.0, t), 1.0 + vdot(t, t))));
        }
        rU = vadd(rU, rUsm);
        double g = vmag(rU) / c;
        U = vdiv(rU, sqrt(1.0 + g * g));
    } else if (pusher == REF_PUSH_BORISNR) {
        double alpha = q * dt2 / mass;
        U = vadd(U, vscale(alpha, E));
        if (rotate) {
            v3 t = vscale(alpha, B);
            v3 Ud = vadd(U, vcross(U, t));
            U = vadd(U, vcross(Ud, vdiv(vscale(2.0, t), 1.0 + vdot(t, t))));
        }
        U = vadd(U, vscale(alpha, E));
    } else if (pusher == REF_PUSH_VAY) {
        double mu = vmag(U) / c;
        double lF = 1.0 / sqrt(1.0 - mu * mu);
        v3 rU = vscale(lF, U);
        v3 Ud = vadd(rU, vscale(q * dt / mass, E));
        v3 T = vscale(q * dt2 / mass, B);
        Ud = vadd(Ud, vcross(U, T));
        double ud = vmag(Ud) / c;
        double alpha = 1.0 + ud * ud;
        double T2 = vdot(T, T);
        double s = alpha - T2;
        double w = vdot(Ud, T) / c;
        double us2 = w * w;
        alpha = 1.0 / sqrt(0.5 * (s + sqrt(s * s + 4.0 * (T2 + us2))));
        T = vscale(alpha, T);
        s = 1.0 / (1.0 + vdot(T, T));
        alpha = vdot(Ud, T);
        rU = vscale(s, vadd(vadd(Ud, vscale(alpha, T)), vcross(Ud, T)));
        double g = vmag(rU) / c;
        U = vdiv(rU, sqrt(1.0 + g * g));
    } else if (pusher == REF_PUSH_HC) {
        double mu = vmag(U) / c;
        v3 rU = vdiv(U, sqrt(1.0 - mu * mu));
        double k = q * dt2 / mass;
        v3 rUsm = vscale(k, E);
        rU = vadd(rU, rUsm);
        if (rotate) {
            double g = vmag(rU) / c;
            double gfm2 = 1.0 + g * g;
            v3 t = vscale(k, B);
            double beta2 = vdot(t, t);
            /* published Higuera-Cary form uses (t.u/c)^2; the reference writes (t.u)^2 */
            double tu = hc_reference_form ? vdot(t, rU) : vdot(t, rU) / c;
            double sg = gfm2 - beta2;
            double iG = 1.0 / sqrt(0.5 * (sg + sqrt(sg * sg + 4.0 * (beta2 + tu * tu))));
            t = vscale(iG, t);
            v3 Ud = vadd(rU, vcross(rU, t));
            rU = vadd(rU, vcross(Ud, vdiv(vscale(2.0, t), 1.0 + vdot(t, t))));
        }
        rU = vadd(rU, rUsm);
        double g = vmag(rU) / c;
        U = vdiv(rU, sqrt(1.0 + g * g));
    }
    st3(Up, U);
}

/* CellAverageFW.C:58-65 / VolumeWeightingFW.C:60-67 (interpolationCellPoint): cell value
 * weighted by y0 plus the face-triangle point values weighted by y1..y3. */
static v3 interp_one(const ref_mesh* m, int weighting, int t, const double* yy, const double* Ecell,
                     const double* Epoint) {
    int c = m->tet_cell[t];
    if (weighting == REF_W_CELLAVERAGE) return ld3(Ecell + 3 * c);
    int vid[3];
    tet_points(m, t, vid);
    double y1 = yy[0], y2 = yy[1], y3 = yy[2];
    double y0 = 1.0 - (y1 + y2 + y3);
    v3 e = vscale(y0, ld3(Ecell + 3 * c));
    e = vadd(e, vscale(y1, ld3(Epoint + 3 * m->pmaster[vid[0]])));
    e = vadd(e, vscale(y2, ld3(Epoint + 3 * m->pmaster[vid[1]])));
    e = vadd(e, vscale(y3, ld3(Epoint + 3 * m->pmaster[vid[2]])));
    return e;
}

void ref_interp_E(const ref_mesh* m, int weighting, int n, const double* y, const int* tet, const double* Ecell,
                  const double* Epoint, double* Ep) {
    for (int i = 0; i < n; ++i) st3(Ep + 3 * i, interp_one(m, weighting, tet[i], y + 3 * i, Ecell, Epoint));
}

void ref_push(const ref_mesh* m, int pusher, int weighting, int n, const double* y, double* U, const int* tet,
              double q, double mass, double dt, const double* Ecell, const double* Epoint, const double* Bcell,
              int hc_reference_form, int rotate) {
    for (int i = 0; i < n; ++i) {
        if (tet[i] < 0) continue;
        v3 E = interp_one(m, weighting, tet[i], y + 3 * i, Ecell, Epoint);
        double e[3] = {E.x, E.y, E.z};
        ref_push_one(pusher, U + 3 * i, e, Bcell + 3 * m->tet_cell[tet[i]], q, mass, dt, hc_reference_form, rotate);
    }
}

/* ------------------------------------------------------------------ tracking */

/* Find the tet of cell c (other than t) whose point set contains points P and Q. */
static int find_tet_with_edge(const ref_mesh* m, int c, int t, int P, int Q, int vid_out[3]) {
    for (int s = m->cell_tet0[c]; s < m->cell_tet0[c + 1]; ++s) {
        if (s == t) continue;
        int v[3];
        tet_points(m, s, v);
        int hp = (v[0] == P || v[1] == P || v[2] == P), hq = (v[0] == Q || v[1] == Q || v[2] == Q);
        if (hp && hq) { memcpy(vid_out, v, sizeof v); return s; }
    }
    return -1;
}

/* Cross tet face iH (opposite vertex iH) of tet t; y[4] is remapped in place.
 * Returns the new tet, or -1 - patch for a particle leaving through a non-coupled patch. */
static int cross_face(const ref_mesh* m, int t, int iH, double y[4]) {
    int c = m->tet_cell[t];
    int vid[3];
    tet_points(m, t, vid);
    if (iH != 0) {
        /* internal tet face (cc, P, Q): the two face points other than vertex iH */
        int P = -1, Q = -1, k = 0;
        int idx[2];
        for (int i = 1; i <= 3; ++i) if (i != iH) { idx[k] = i; if (k == 0) P = vid[i - 1]; else Q = vid[i - 1]; ++k; }
        int nv[3];
        int s = find_tet_with_edge(m, c, t, P, Q, nv);
        if (s < 0) return -1000000;
        double ny[4] = {y[0], 0.0, 0.0, 0.0};
        for (int i = 0; i < 3; ++i) {
            if (nv[i] == P) ny[i + 1] = y[idx[0]];
            else if (nv[i] == Q) ny[i + 1] = y[idx[1]];
        }
        memcpy(y, ny, sizeof ny);
        return s;
    }
    int f = m->tet_face[t], pt = m->tet_pt[t];
    if (f < m->n_internal) {
        int c2 = (m->owner[f] == c) ? m->nbr[f] : m->owner[f];
        for (int s = m->cell_tet0[c2]; s < m->cell_tet0[c2 + 1]; ++s)
            if (m->tet_face[s] == f && m->tet_pt[s] == pt) {
                int nv[3];
                tet_points(m, s, nv);
                double ny[4] = {0.0, 0.0, 0.0, 0.0};
                for (int i = 0; i < 3; ++i)
                    for (int j = 0; j < 3; ++j) if (nv[j] == vid[i]) ny[j + 1] = y[i + 1];
                memcpy(y, ny, sizeof ny);
                return s;
            }
        return -1000000;
    }
    int pi = m->face_patch[f];
    if (m->ptype[pi] == REF_PATCH_EMPTY) return REF_EMPTY_FACE;
    if (m->ptype[pi] != REF_PATCH_CYCLIC) return -1 - pi;
    int q = m->pnbr[pi];
    int g = m->pstart[q] + (f - m->pstart[pi]);
    int c2 = m->owner[g];
    v3 sep = ld3(m->psep + 3 * pi);
    v3 img[3];
    for (int i = 0; i < 3; ++i) img[i] = vadd(ld3(m->points + 3 * vid[i]), sep);
    double scale = 0.0;
    for (int i = 0; i < 3; ++i) scale += vmag(vsub(img[i], img[(i + 1) % 3]));
    for (int s = m->cell_tet0[c2]; s < m->cell_tet0[c2 + 1]; ++s) {
        if (m->tet_face[s] != g) continue;
        int nv[3];
        tet_points(m, s, nv);
        int map[3] = {-1, -1, -1}, ok = 1;
        for (int i = 0; i < 3 && ok; ++i) {
            for (int j = 0; j < 3; ++j)
                if (vmag(vsub(ld3(m->points + 3 * nv[j]), img[i])) < 1e-9 * scale) { map[i] = j; break; }
            if (map[i] < 0) ok = 0;
        }
        if (!ok) continue;
        double ny[4] = {0.0, 0.0, 0.0, 0.0};
        for (int i = 0; i < 3; ++i) ny[map[i] + 1] = y[i + 1];
        memcpy(y, ny, sizeof ny);
        return s;
    }
    return -1000000;
}

/* what a wall hit needs besides the mesh: the push of the velocity synchronisation and
 * the reflection model */
struct track_ctx {
    int pusher, weighting, hc_reference_form, rotate, do_push;
    double q, mass, dt;
    const double *Ecell, *Epoint, *Bcell;
    const ref_wall_params* walls;
    uint32_t id;
    int hits;      /* wall hits of this parcel in this step */
    int resyncs;   /* velocity resyncs after a reflection (PICParcel.C:126-137) */
    int64_t total_hits;
    /* SimpleInjection re-emission: the injector patch's triangles */
    const int *r_tface, *r_tj;
    const double* r_cum;
    int r_ntri;
};

#define TAG_WALL 6u
#define TAG_REEMIT 12u
static void emit_one(const ref_mesh* m, const ref_inject_params* p, const int* tface, const int* tj,
                     const double* cum, int ntri, int by_area, uint32_t ka, uint32_t kc, uint32_t tag, double* y,
                     int* tet, double* Uo);
static int inject_table(const ref_mesh* m, int pt, int** tface, int** tj, double** cum);
static inline double u01(uint32_t a, uint32_t b);
static inline void rng4(uint64_t seed, uint32_t a, uint32_t b, uint32_t c, uint32_t d, uint32_t out[4]);

/* syncVelocityAtBoundary (PICParcel.C:201-210): a push of the velocity by fraction * dt
 * with the field at the parcel's place on the face */
static void sync_velocity(const ref_mesh* m, const struct track_ctx* x, int t, const double y[4], double* U,
                          double fraction) {
    if (!x->walls->sync || x->q == 0.0 || !x->do_push) return;
    const double yy[3] = {y[1], y[2], y[3]};
    v3 E = interp_one(m, x->weighting, t, yy, x->Ecell, x->Epoint);
    double e[3] = {E.x, E.y, E.z};
    ref_push_one(x->pusher, U, e, x->Bcell + 3 * m->tet_cell[t], x->q, x->mass, fraction * x->dt,
                 x->hc_reference_form, x->rotate);
}

/* The wall's reflection (SpecularReflection.C:55-68, MaxwellianThermal.C:55-118,
 * MixedDiffuseSpecular.C:55-131): n is the outward unit normal of the face triangle. */
static void reflect(const ref_mesh* m, struct track_ctx* x, int patch, v3 n, double* Up) {
    const ref_wall_params* w = x->walls;
    v3 U = ld3(Up);
    uint32_t o[4];
    const uint32_t c = (x->walls->species << 16) | (uint32_t)(x->hits & 0xffff);
    rng4(w->seed, x->id, (uint32_t)w->step, c, TAG_WALL << 24, o);
    const double u_mix = u01(o[0], o[1]), u_n = u01(o[2], o[3]);
    int diffuse = w->model == REF_WALL_MAXWELLIAN_THERMAL ||
                  (w->model == REF_WALL_MIXED && w->diffuse_fraction > u_mix);
    double un = vdot(U, n);
    if (!diffuse) {
        if (un > 0.0) U = vsub(U, vscale(2.0 * un, n));
        st3(Up, U);
        return;
    }
    v3 Ut = vsub(U, vscale(un, n));
    for (uint32_t k = 2; vmag(Ut) < 1e-15 && k < 200; k += 2) {
        /* incident along the normal: perturb the velocity until a tangent exists */
        uint32_t a[4], b[4];
        rng4(w->seed, x->id, (uint32_t)w->step, c, (TAG_WALL << 24) | k, a);
        rng4(w->seed, x->id, (uint32_t)w->step, c, (TAG_WALL << 24) | (k + 1), b);
        U = V(U.x * (0.8 + 0.2 * u01(a[0], a[1])), U.y * (0.8 + 0.2 * u01(a[2], a[3])),
              U.z * (0.8 + 0.2 * u01(b[0], b[1])));
        un = vdot(U, n);
        Ut = vsub(U, vscale(un, n));
    }
    const v3 t1 = vdiv(Ut, vmag(Ut));
    const v3 t2 = vcross(n, t1);
    uint32_t g[4];
    rng4(w->seed, x->id, (uint32_t)w->step, c, (TAG_WALL << 24) | 1u, g);
    const double uc = u01(g[0], g[1]), ud = u01(g[2], g[3]);
    const double r = sqrt(-2.0 * log(1.0 - uc));
    const double n1 = r * cos(2.0 * PI_D * ud), n2 = r * sin(2.0 * PI_D * ud);
    const double s = sqrt(KB * w->T[patch] / x->mass);
    const double vn = sqrt(-2.0 * log(fmax(1.0 - u_n, 1e-300)));
    U = vscale(s, vsub(vadd(vscale(n1, t1), vscale(n2, t2)), vscale(vn, n)));
    U = vadd(U, ld3(w->U + 3 * patch));
    st3(Up, U);
}

/* one side of a wall hit in the wall surface fields (PICParcel.C:266-287, :299-320) */
static void wall_fields_add(const struct track_ctx* x, double* row, v3 n, double fA, v3 U) {
    const double m = x->mass, un = vdot(U, n);
    const v3 Ut = vsub(U, vscale(un, n));
    const double inv = 1.0 / fmax(fabs(un) * fA, 1e-300);
    const double ke = 0.5 * m * vdot(U, U) * inv;
    double* sp = row + 10 + 9 * (int)x->walls->species;
    const double v[6] = {inv, m * inv, ke, m * Ut.x * inv, m * Ut.y * inv, m * Ut.z * inv};
    const double w[6] = {inv, m * inv, m * Ut.x * inv, m * Ut.y * inv, m * Ut.z * inv, ke};
    for (int k = 0; k < 6; ++k) {
#pragma omp atomic
        row[k] += v[k];
#pragma omp atomic
        sp[k] += w[k];
    }
}

/* PICParcel::moveForward (PICParcel.C:81-144) with OpenFOAM-9 barycentric tracking
 * (particle::trackToFace / trackToStationaryTet re-derived): per tet, the displacement
 * (1-f)*dt*U is taken to barycentric form Tx_i = r_i.d (det-scaled); the first face with
 * Tx_i < 0 and 0 <= -y_i/Tx_i < 1/det is hit. */
static int track_one(const ref_mesh* m, double* yy, double* Up, int* tetp, double dt, struct track_ctx* x) {
    int t = *tetp;
    double y[4];
    y[1] = yy[0]; y[2] = yy[1]; y[3] = yy[2];
    y[0] = 1.0 - (y[1] + y[2] + y[3]);
    v3 U = ld3(Up);
    /* reduced-dimension cases track with the empty components removed (PICParcel.C:118-123:
     * meshTools::constrainDirection over solutionD); the parcel keeps its full U */
    if (m->track_mask & 1) U.x = 0.0;
    if (m->track_mask & 2) U.y = 0.0;
    if (m->track_mask & 4) U.z = 0.0;
    double sf = 0.0;
    int guard = 0, status = 0;
    while (sf < 1.0) {
        double f = 1.0 - sf;
        v3 d = vscale(f * dt, U);
        v3 X[4]; int vid[3];
        tet_vertices(m, t, X, vid);
        v3 e1 = vsub(X[1], X[0]), e2 = vsub(X[2], X[0]), e3 = vsub(X[3], X[0]);
        v3 r1 = vcross(e2, e3), r2 = vcross(e3, e1), r3 = vcross(e1, e2);
        double det = vdot(e1, r1);
        double T[4];
        T[1] = vdot(r1, d); T[2] = vdot(r2, d); T[3] = vdot(r3, d);
        T[0] = -(T[1] + T[2] + T[3]);
        double muH = (det > 0.0 && isnormal(det)) ? 1.0 / det : 1e300;
        int iH = -1;
        for (int i = 0; i < 4; ++i)
            if (T[i] < 0.0) {
                double mu = -y[i] / T[i];
                if (mu >= 0.0 && mu < muH) { muH = mu; iH = i; }
            }
        for (int i = 0; i < 4; ++i) y[i] = y[i] + muH * T[i];
        if (iH < 0) { sf = 1.0; break; }
        y[iH] = 0.0;
        double frac = muH * det;
        if (frac < 0.0) frac = 0.0;
        if (frac > 1.0) frac = 1.0;
        sf = sf + f * frac;
        int s = cross_face(m, t, iH, y);
        if (s == REF_EMPTY_FACE) break; /* on an empty face (rounding only): stays */
        if (s <= -1000000) { status = 2; break; }
        const int kind = (s < 0 && s > -1000000 && x && x->walls)
                             ? (x->walls->patch_kind ? x->walls->patch_kind[-1 - s]
                                                     : (m->ptype[-1 - s] == REF_PATCH_WALL ? 1 : 0))
                             : 0;
        if (kind == 2 || kind == 3) {
            /* SimpleInjection::particleBC on its patch (SimpleInjection.C:241-272) after the
             * hit's sync of a plain patch (PICParcel.C:158-161): reflection without a resync,
             * or the parcel replaced by a new one from the emitter at the same step fraction,
             * synchronised to it (ParticleEmitter.C:136-221) */
            const int patch = -1 - s;
            if (kind == 2) {
                v3 r1 = vcross(e2, e3), r2 = vcross(e3, e1), r3 = vcross(e1, e2);
                v3 nn = vadd(vadd(r1, r2), r3);
                nn = vdiv(nn, vmag(nn));
                sync_velocity(m, x, t, y, Up, sf - 0.5);
                v3 Uh = ld3(Up);
                const double un = vdot(Uh, nn);
                if (un > 0.0) st3(Up, vsub(Uh, vscale(2.0 * un, nn)));
            } else {
                const ref_inject_params* rp = x->walls->reemit;
                double yy3[3];
                const uint32_t kc = (x->walls->species << 16) | (uint32_t)(x->hits & 0xffff);
                emit_one(m, rp, x->r_tface, x->r_tj, x->r_cum, x->r_ntri, 0, x->id, kc, TAG_REEMIT, yy3, &t, Up);
                y[1] = yy3[0]; y[2] = yy3[1]; y[3] = yy3[2];
                y[0] = 1.0 - (y[1] + y[2] + y[3]);
                sync_velocity(m, x, t, y, Up, sf - 0.5);
            }
            (void)patch;
            ++x->hits;
            ++x->total_hits;
            U = ld3(Up);
            if (m->track_mask & 1) U.x = 0.0;
            if (m->track_mask & 2) U.y = 0.0;
            if (m->track_mask & 4) U.z = 0.0;
            if (++guard > MAX_TRACK_STEPS) { status = 1; break; }
            continue;
        }
        if (kind == 1) {
            /* PICParcel::hitPatch then the resync of moveForward: the parcel stays on the
             * face in its tet and tracks the rest of the step with the new velocity */
            const int patch = -1 - s;
            v3 r1 = vcross(e2, e3), r2 = vcross(e3, e1), r3 = vcross(e1, e2);
            v3 nn = vadd(vadd(r1, r2), r3);
            nn = vdiv(nn, vmag(nn));
            sync_velocity(m, x, t, y, Up, sf - 0.5);
            const v3 pre = ld3(Up);
            double* row = NULL;
            double fA = 0.0;
            if (x->walls->face_acc) {
                const int f = m->tet_face[t];
                row = x->walls->face_acc + (size_t)(f - m->n_internal) * x->walls->acc_stride;
                fA = vmag(ld3(m->Sf + 3 * f));
                wall_fields_add(x, row, nn, fA, pre);
            }
            reflect(m, x, patch, nn, Up);
            if (row) {
                const v3 post = ld3(Up);
                wall_fields_add(x, row, nn, fA, post);
                const double sc = x->walls->weight / (x->dt * fA), ms = x->mass;
                const double dq = sc * (0.5 * ms * vdot(pre, pre) - 0.5 * ms * vdot(post, post));
                const v3 fd = vscale(sc, vsub(vscale(ms, pre), vscale(ms, post)));
                double* sp = row + 10 + 9 * (int)x->walls->species;
#pragma omp atomic
                row[6] += dq;
#pragma omp atomic
                row[7] += fd.x;
#pragma omp atomic
                row[8] += fd.y;
#pragma omp atomic
                row[9] += fd.z;
#pragma omp atomic
                sp[6] += fd.x;
#pragma omp atomic
                sp[7] += fd.y;
#pragma omp atomic
                sp[8] += fd.z;
            }
            if (x->resyncs++ == 0) sync_velocity(m, x, t, y, Up, 0.5 - sf);
            ++x->hits;
            ++x->total_hits;
            U = ld3(Up);
            if (m->track_mask & 1) U.x = 0.0;
            if (m->track_mask & 2) U.y = 0.0;
            if (m->track_mask & 4) U.z = 0.0;
            if (++guard > MAX_TRACK_STEPS) { status = 1; break; }
            continue;
        }
        if (s < 0) { t = -1; break; } /* left through a non-coupled patch: deleted (NoInflow) */
        t = s;
        if (++guard > MAX_TRACK_STEPS) { status = 1; break; }
    }
    *tetp = t;
    yy[0] = y[1]; yy[1] = y[2]; yy[2] = y[3];
    return status;
}

int ref_move_walls(const ref_mesh* m, int pusher, int weighting, int n, double* y, double* U, int* tet, double q,
                   double mass, double dt, const double* Ecell, const double* Epoint, const double* Bcell,
                   int do_push, int hc_reference_form, int rotate, int* n_superluminal,
                   const ref_wall_params* walls) {
    int guards = 0, sl = 0;
    int64_t hits = 0;
    int *r_tface = NULL, *r_tj = NULL, r_ntri = 0;
    double* r_cum = NULL;
    if (walls && walls->reemit) r_ntri = inject_table(m, walls->reemit->patch, &r_tface, &r_tj, &r_cum);
    /* parcels are independent: spread over the host threads for the CPU baseline */
#pragma omp parallel for schedule(static) reduction(+ : guards, sl, hits)
    for (int i = 0; i < n; ++i) {
        if (tet[i] < 0) continue;
        if (do_push && q != 0.0) {
            v3 E = interp_one(m, weighting, tet[i], y + 3 * i, Ecell, Epoint);
            double e[3] = {E.x, E.y, E.z};
            ref_push_one(pusher, U + 3 * i, e, Bcell + 3 * m->tet_cell[tet[i]], q, mass, dt, hc_reference_form,
                         rotate);
        }
        if (vmag(ld3(U + 3 * i)) >= C_LIGHT) ++sl; /* PICParcel.C:61-66 */
        struct track_ctx x = {pusher, weighting, hc_reference_form, rotate, do_push, q, mass, dt,
                              Ecell, Epoint, Bcell, walls, walls ? walls->id[i] : 0u, 0, 0, 0,
                              r_tface, r_tj, r_cum, r_ntri};
        if (track_one(m, y + 3 * i, U + 3 * i, tet + i, dt, &x)) ++guards;
        hits += x.total_hits;
    }
    free(r_tface); free(r_tj); free(r_cum);
    if (n_superluminal) *n_superluminal = sl;
    if (walls && walls->hits) *walls->hits += hits;
    return guards;
}

int ref_move(const ref_mesh* m, int pusher, int weighting, int n, double* y, double* U, int* tet, double q,
             double mass, double dt, const double* Ecell, const double* Epoint, const double* Bcell, int do_push,
             int hc_reference_form, int rotate, int* n_superluminal) {
    return ref_move_walls(m, pusher, weighting, n, y, U, tet, q, mass, dt, Ecell, Epoint, Bcell, do_push,
                          hc_reference_form, rotate, n_superluminal, NULL);
}

/* ------------------------------------------------------------------ deposit */

/* CellAverage.C:60-63 (q*w into the cell) and VolumeWeighting.C:105-124 (y0 to the cell
 * centre, y1..y3 to the face-triangle points), in int64 fixed point. */
void ref_deposit(const ref_mesh* m, int weighting, int n, const double* y, const int* tet, double scale,
                 int64_t* acc_cell, int64_t* acc_point) {
    const double s = scale * 4398046511104.0; /* 2^42 */
    const int64_t whole = llrint(s);
    for (int i = 0; i < n; ++i) {
        int t = tet[i];
        if (t < 0) continue;
        int c = m->tet_cell[t];
        if (weighting == REF_W_CELLAVERAGE) { acc_cell[c] += whole; continue; }
        int vid[3];
        tet_points(m, t, vid);
        double y1 = y[3 * i], y2 = y[3 * i + 1], y3 = y[3 * i + 2];
        double y0 = 1.0 - (y1 + y2 + y3);
        acc_cell[c] += llrint(y0 * s);
        acc_point[m->pmaster[vid[0]]] += llrint(y1 * s);
        acc_point[m->pmaster[vid[1]]] += llrint(y2 * s);
        acc_point[m->pmaster[vid[2]]] += llrint(y3 * s);
    }
}

/* CellAverage.C:66-68 / VolumeWeighting.C:127-164 */
void ref_deposit_finish(const ref_mesh* m, int weighting, const int64_t* acc_cell, const int64_t* acc_point,
                        double Q0, double* rho) {
    const double q = Q0 / 4398046511104.0;
    for (int c = 0; c < m->n_cells; ++c) {
        double V = m->V[c];
        double r = ((double)acc_cell[c] * q) / V;
        if (weighting == REF_W_VOLUMEWEIGHTING)
            for (int k = m->cpo[c]; k < m->cpo[c + 1]; ++k)
                r += m->vw_w[k] * ((double)acc_point[m->pmaster[m->cp[k]]] * q) / V;
        rho[c] = r;
    }
}

void ref_cell_to_point(const ref_mesh* m, const double* Ecell, double* Epoint) {
    memset(Epoint, 0, sizeof(double) * 3 * m->n_points);
    for (int c = 0; c < m->n_cells; ++c)
        for (int k = m->cpo[c]; k < m->cpo[c + 1]; ++k) {
            int p = m->pmaster[m->cp[k]];
            double w = m->ip_w[k];
            for (int d = 0; d < 3; ++d) Epoint[3 * p + d] += w * Ecell[3 * c + d];
        }
    /* non-master points carry the value of their master */
    for (int p = 0; p < m->n_points; ++p)
        if (m->pmaster[p] != p) for (int d = 0; d < 3; ++d) Epoint[3 * p + d] = Epoint[3 * m->pmaster[p] + d];
}

/* ------------------------------------------------------------------ Philox4x32-10 */

/* Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as 1, 2, 3", SC'11. */
void ref_philox4x32(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4]) {
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0, n1 = lo1, n2 = hi0 ^ c3 ^ k1, n3 = lo0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

static inline double u01(uint32_t a, uint32_t b) {
    uint64_t x = (((uint64_t)a << 32) | b) >> 11;
    return (double)x * (1.0 / 9007199254740992.0);
}

static inline void rng4(uint64_t seed, uint32_t a, uint32_t b, uint32_t c, uint32_t d, uint32_t out[4]) {
    uint32_t ctr[4] = {a, b, c, d}, key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    ref_philox4x32(ctr, key, out);
}

/* stream tags (top byte of counter word 3) */
#define TAG_SHUFFLE 1u
#define TAG_COULOMB 2u
#define TAG_MCC 3u
#define TAG_NEWID 4u

/* ------------------------------------------------------------------ sort */

int ref_sort_by_cell(const ref_mesh* m, int n, const int* tet, int* order, int* cell_start) {
    int nc = m->n_cells;
    memset(cell_start, 0, sizeof(int) * (nc + 1));
    for (int i = 0; i < n; ++i) if (tet[i] >= 0) cell_start[m->tet_cell[tet[i]] + 1]++;
    for (int c = 0; c < nc; ++c) cell_start[c + 1] += cell_start[c];
    int* cur = (int*)malloc(sizeof(int) * (nc + 1));
    memcpy(cur, cell_start, sizeof(int) * (nc + 1));
    for (int i = 0; i < n; ++i) if (tet[i] >= 0) order[cur[m->tet_cell[tet[i]]]++] = i;
    int live = cell_start[nc];
    free(cur);
    return live;
}

/* ------------------------------------------------------------------ Coulomb */

typedef struct { uint64_t key; int idx; } keyrec;
static int keyrec_cmp(const void* a, const void* b) {
    const keyrec* x = (const keyrec*)a; const keyrec* y = (const keyrec*)b;
    if (x->key != y->key) return x->key < y->key ? -1 : 1;
    return x->idx < y->idx ? -1 : (x->idx > y->idx);
}

/* shuffle (OpenFOAM Random::shuffle, TakizukaAbePairing.C:110, NanbuYonemuraPairing.C:124)
 * restated as a sort by a counter-based 53-bit random key of (particle id, step, stage,
 * species); equal keys (parcels sharing an id) keep their order in the cell's list */
static void keyed_perm(uint64_t seed, int64_t step, int stage, int species, const uint32_t* id, int first, int n,
                       int* perm) {
    keyrec* r = (keyrec*)malloc(sizeof(keyrec) * (n > 0 ? n : 1));
    for (int k = 0; k < n; ++k) {
        uint32_t o[4];
        rng4(seed, id[first + k], (uint32_t)step, (uint32_t)((stage << 8) | species), TAG_SHUFFLE << 24, o);
        r[k].key = ((uint64_t)o[0] << 21) | (o[1] >> 11);
        r[k].idx = first + k;
    }
    qsort(r, n, sizeof(keyrec), keyrec_cmp);
    for (int k = 0; k < n; ++k) perm[k] = r[k].idx;
    free(r);
}

/* NanbuCollision.C:137-159 (PerezRelativisticCollision.C:198-219 is the same function) */
static double cos_chi(double s, double U) {
    double A, invA;
    if (s < 0.1) {
        if (U < 0.0001) U = 0.0001;
        return 1.0 + s * log(U);
    }
    if (s < 3.0) {
        invA = 0.00569578 + (0.95602 + (-0.508139 + (0.479139 + (-0.12789 + 0.0238957 * s) * s) * s) * s) * s;
        A = 1.0 / invA;
        return invA * log(exp(-A) + 2. * U * sinh(A));
    }
    if (s < 6.0) {
        A = 3.0 * exp(-s);
        return (1.0 / A) * log(exp(-A) + 2.0 * U * sinh(A));
    }
    return 2.0 * U - 1.0;
}

/* NanbuCollision::collide, NanbuCollision.C:60-134, on the two velocities */
static void nanbu_collide(v3* UP, v3* UQ, double mP, double qP, double mQ, double qQ, double lnL_fixed,
                          int calc_bmin, double debye, double nP, double nQ, double nPQ, double dt, double u1,
                          double u2, double* stats) {
    double reMass = mP * mQ / (mP + mQ);
    v3 uP = *UP, uQ = *UQ;
    v3 g = vsub(uP, uQ);
    double gm = vmag(g);
    double gp = sqrt(g.y * g.y + g.z * g.z);
    if (!(gm > 0.0) || !(gp > 0.0)) return; /* degenerate relative velocity: the reference divides by zero */
    double lnL;
    if (lnL_fixed == 0.0) {
        double b0 = qP * qQ / (2.0 * PI_D * EPS0 * reMass * gm * gm);
        if (calc_bmin) { double hb = HPLANCK / (2.0 * (gm * reMass)); b0 = hb > fabs(b0) ? hb : fabs(b0); }
        double l = 0.5 * log(1.0 + (debye * debye) / (b0 * b0));
        lnL = l > 2.0 ? l : 2.0;
    } else lnL = lnL_fixed;
    double a = qP * qQ / (EPS0 * reMass);
    double s = lnL / (4.0 * PI_D) * (a * a) * (nQ * nP / nPQ) * pow(gm, -3.0) * dt;
    double cc = cos_chi(s, u1);
    double sc = sqrt(1.0 - cc * cc);
    double eps = 2.0 * PI_D * u2;
    double ce = cos(eps), se = sin(eps);
    v3 h = V(gp * ce, -1.0 * (g.y * g.x * ce + gm * g.z * se) / gp, -1.0 * (g.z * g.x * ce - gm * g.y * se) / gp);
    v3 dv = vadd(vscale(1.0 - cc, g), vscale(sc, h));
    *UP = vsub(uP, vscale(mQ / (mP + mQ), dv));
    *UQ = vadd(uQ, vscale(mP / (mP + mQ), dv));
    if (stats) { stats[1] += lnL; stats[2] += s; }
}

/* Lorentz transformation to the centre-of-momentum frame of two parcels and back, the
 * common part of PerezRelativisticCollision.C:84-196 and
 * RelativisticElectronNeutralCollision.C:84-158 */
typedef struct {
    double gP, gQ, msum, gc, gcs, gPs, gQs, pm;
    v3 vc, pPs;
} com_frame;

static void com_enter(v3 uP, v3 uQ, double mP, double mQ, com_frame* f) {
    const double c = C_LIGHT, c2 = C_LIGHT * C_LIGHT;
    double bP = vmag(uP) / c, bQ = vmag(uQ) / c;
    f->gP = 1.0 / sqrt(1.0 - bP * bP);
    f->gQ = 1.0 / sqrt(1.0 - bQ * bQ);
    v3 pP = vscale(mP * f->gP, uP), pQ = vscale(mQ * f->gQ, uQ);
    f->msum = mP * f->gP + mQ * f->gQ;
    f->vc = vdiv(vadd(pP, pQ), f->msum);
    double vc2 = vdot(f->vc, f->vc);
    if (vc2 == 0.0) {
        f->gc = 1.0; f->gcs = 0.5; f->pPs = pP; f->gPs = f->gP; f->gQs = f->gQ;
    } else {
        f->gc = 1.0 / sqrt(1.0 - vc2 / c2);
        f->gcs = (f->gc - 1.0) / vc2;
        f->pPs = vadd(pP, vscale((f->gcs * vdot(f->vc, uP) - f->gc) * mP * f->gP, f->vc));
        f->gPs = (1.0 - vdot(f->vc, uP) / c2) * f->gP * f->gc;
        f->gQs = (1.0 - vdot(f->vc, uQ) / c2) * f->gQ * f->gc;
    }
    f->pm = vmag(f->pPs);
}

/* rotate P's centre-of-momentum momentum by (chi, eps) and transform both parcels back */
static void com_leave(const com_frame* f, double mP, double mQ, double cc, double sc, double eps, v3* UP, v3* UQ) {
    const double c = C_LIGHT;
    double sce = sc * cos(eps), scs = sc * sin(eps);
    v3 p = f->pPs;
    double pm = f->pm;
    v3 np;
    double pperp = sqrt(p.x * p.x + p.y * p.y);
    if (pperp > 1.e-10 * pm) {
        double ip = 1.0 / pperp;
        np = V((p.x * p.z * sce - p.y * pm * scs) * ip + p.x * cc, (p.y * p.z * sce + p.x * pm * scs) * ip + p.y * cc,
               -pperp * sce + p.z * cc);
    } else {
        np = V(pm * sce, pm * scs, pm * cc);
    }
    v3 newP = vadd(np, vscale(f->gcs * vdot(f->vc, np) + mP * f->gPs * f->gc, f->vc));
    double aP = vmag(newP) / (mP * c);
    double ngP = sqrt(1.0 + aP * aP);
    v3 mnp = vscale(-1.0, np);
    v3 newQ = vadd(mnp, vscale(f->gcs * vdot(f->vc, mnp) + mQ * f->gQs * f->gc, f->vc));
    double aQ = vmag(newQ) / (mQ * c);
    double ngQ = sqrt(1.0 + aQ * aQ);
    *UP = vdiv(newP, mP * ngP);
    *UQ = vdiv(newQ, mQ * ngQ);
}

/* PerezRelativisticCollision::collide, PerezRelativisticCollision.C:62-196: the scattering
 * is done on the momentum of P in the centre-of-momentum frame */
static void perez_collide(v3* UP, v3* UQ, double mP, double qP, double mQ, double qQ, double lnL_fixed,
                          int calc_bmin, double debye, double nP, double nQ, double nPQ, double dt, double u1,
                          double u2, double* stats) {
    const double c2 = C_LIGHT * C_LIGHT;
    com_frame f;
    com_enter(*UP, *UQ, mP, mQ, &f);
    double pm = f.pm;
    if (pm == 0.0) return;
    double lnL;
    double t2 = 1.0 + mP * f.gPs * mQ * f.gQs * c2 / (pm * pm);
    if (lnL_fixed == 0.0) {
        double b0 = qP * qQ / (4.0 * PI_D * EPS0 * c2) * f.gc / f.msum * t2;
        if (calc_bmin) { double hb = HPLANCK / (2.0 * pm); b0 = hb > fabs(b0) ? hb : fabs(b0); }
        double l = 0.5 * log(1.0 + (debye * debye) / (b0 * b0));
        lnL = l > 2.0 ? l : 2.0;
    } else lnL = lnL_fixed;
    double s1 = f.gc * pm / f.msum;
    double s2 = t2 * t2;
    double s = (nP * nQ / nPQ) * dt * lnL * (qP * qQ) * (qP * qQ) /
               (4.0 * PI_D * EPS0 * EPS0 * (c2 * c2) * mP * f.gP * mQ * f.gQ) * s1 * s2;
    double vrel = pm * f.msum / (mP * f.gPs * mQ * f.gQs * f.gc);
    double dmax = mP * pow(nP, 2.0 / 3.0), dq = mQ * pow(nQ, 2.0 / 3.0);
    if (dq > dmax) dmax = dq;
    double smax = pow(4.0 * PI_D / 3.0, 1.0 / 3.0) * (nP * nQ / nPQ) * dt * (mP + mQ) / dmax * vrel;
    if (smax < s) s = smax;
    double cc = cos_chi(s, u1);
    double sc = sqrt(1.0 - cc * cc);
    com_leave(&f, mP, mQ, cc, sc, 2.0 * PI_D * u2, UP, UQ);
    if (stats) { stats[1] += lnL; stats[2] += s; }
}

/* CoulombCollisionModel::debyeLength, CoulombCollisionModel.C:249-310: the charged species of
 * debyeLength_Species_ (ctor :98-122), the drift removed from the temperature when
 * debyeLength_accountForDrift_ (:136-141, :288-292) */
static double debye_length(const ref_coulomb_params* p, double* const* U, const int* const* cs, int c, double Vc) {
    double debye = 0.0, nmax = 0.0;
    for (int s = 0; s < p->n_species; ++s) {
        if (p->charge[s] == 0.0 || !((p->debye_mask >> s) & 1u)) continue;
        int a = cs[s][c], b = cs[s][c + 1];
        double n = 0.0, ch = 0.0, vs = 0.0;
        v3 v = V(0, 0, 0);
        for (int k = a; k < b; ++k) {
            v3 u = ld3(U[s] + 3 * k);
            vs += vdot(u, u) * p->weight[s];
            ch += p->charge[s] * p->weight[s];
            v = vadd(v, vscale(p->weight[s], u));
            n += p->weight[s];
        }
        if (n <= 0.0) continue;
        v = vdiv(v, n); ch /= n; vs /= n; n /= Vc;
        if (n > nmax) nmax = n;
        double T = p->mass[s] / (3.0 * KB) * (vs - (p->debye_drift ? vdot(v, v) : 0.0));
        if (T <= 0.0) continue;
        debye += n * ch * ch / EPS0 * 1.0 / (T * KB);
    }
    if (debye <= 0.0) return 0.0;
    double rmin = pow(4.0 * PI_D * nmax / 3.0, -1.0 / 3.0);
    double d = sqrt(1.0 / debye);
    return d > rmin ? d : rmin;
}

typedef struct {
    const ref_coulomb_params* p;
    double* const* U;
    int cell, stage;
    double debye, dt;
    double* stats;
} coll_ctx;

/* one collision of parcel iP of species sP with parcel iQ of species sQ: two uniforms
 * (cell, step, stage, k) for cos(chi) and the azimuth, a third for the weight correction */
static void collide(const coll_ctx* x, int sP, int iP, int sQ, int iQ, double lnL, double nP, double nQ, double nPQ,
                    int k) {
    const ref_coulomb_params* p = x->p;
    uint32_t o[4];
    rng4(p->seed, (uint32_t)x->cell, (uint32_t)p->step, (uint32_t)((x->stage << 24) | k), TAG_COULOMB << 24, o);
    double u1 = u01(o[0], o[1]), u2 = u01(o[2], o[3]);
    v3 uP = ld3(x->U[sP] + 3 * iP), uQ = ld3(x->U[sQ] + 3 * iQ);
    v3 preP = uP, preQ = uQ;
    if (p->model == 1)
        perez_collide(&uP, &uQ, p->mass[sP], p->charge[sP], p->mass[sQ], p->charge[sQ], lnL, p->calculate_bmin,
                      x->debye, nP, nQ, nPQ, x->dt, u1, u2, x->stats);
    else
        nanbu_collide(&uP, &uQ, p->mass[sP], p->charge[sP], p->mass[sQ], p->charge[sQ], lnL, p->calculate_bmin,
                      x->debye, nP, nQ, nPQ, x->dt, u1, u2, x->stats);
    if (p->weight_correction) {
        /* NanbuWeightCorrectionModel::correctVelocity, NanbuCorrection.C:50-69 */
        rng4(p->seed, (uint32_t)x->cell, (uint32_t)p->step, (uint32_t)((x->stage << 24) | k),
             (TAG_COULOMB << 24) | 1u, o);
        double r = u01(o[0], o[1]);
        double wP = p->weight[sP], wQ = p->weight[sQ];
        if (r > wP / wQ) uQ = preQ;
        if (r > wQ / wP) uP = preP;
    }
    st3(x->U[sP] + 3 * iP, uP);
    st3(x->U[sQ] + 3 * iQ, uQ);
    if (x->stats) x->stats[0] += 1.0;
    if (p->hits) { p->hits[sP][iP] += 1; p->hits[sQ][iQ] += 1; }
}

static double dmin(double a, double b) { return a < b ? a : b; }

/* TakizukaAbePairing::pairANDcollide (TakizukaAbePairing.C:60-338) and
 * NanbuYonemuraPairing::pairANDcollide (NanbuYonemuraPairing.C:56-322): per cell, species
 * pairs (i, j>=i) visited as i ascending, j descending, one stage each.
 *   intra, even count: consecutive pairs of the shuffled list, n12 = n.
 *   intra, odd count: Takizuka-Abe collides the first three cyclically (n12 = (n+3)w/V);
 *     Nanbu-Yonemura collides the last with the first after the pairs (n12 = (n+1)w/V).
 *   inter, equal counts: shuffled k with shuffled n-1-k.
 *   inter, unequal counts: the two Takizuka-Abe groups, or Nanbu-Yonemura's cycling of
 *     the smaller species over the larger one. */
void ref_coulomb(const ref_mesh* m, const ref_coulomb_params* p, double dt, double* const* U,
                 const uint32_t* const* id, const int* const* cs, double* stats) {
    int ns = p->n_species;
    int maxn = 1;
    for (int s = 0; s < ns; ++s)
        for (int c = 0; c < m->n_cells; ++c) if (cs[s][c + 1] - cs[s][c] > maxn) maxn = cs[s][c + 1] - cs[s][c];
    int calc_debye = 0;
    for (int i = 0; i < ns; ++i)
        for (int j = i; j < ns; ++j)
            if (p->charge[i] != 0.0 && p->charge[j] != 0.0 && p->coulomb_log[i][j] <= 0.0) calc_debye = 1;
    const int ny = p->pairing == 1;
    /* cells are independent: spread over the host threads for the CPU baseline, each
     * thread with its own lists and counters (the counters are summed at the end) */
#pragma omp parallel
    {
    int* pa = (int*)malloc(sizeof(int) * maxn);
    int* pb = (int*)malloc(sizeof(int) * maxn);
    double lst[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    coll_ctx x;
    x.p = p; x.U = U; x.dt = dt; x.stats = stats ? lst : NULL;
#pragma omp for schedule(dynamic, 16)
    for (int c = 0; c < m->n_cells; ++c) {
        double Vc = m->V[c];
        x.cell = c;
        x.debye = calc_debye ? debye_length(p, U, cs, c, Vc) : 0.0;
        if (stats) lst[3] += x.debye;
        int stage = 0;
        for (int i = 0; i < ns; ++i) {
            if (p->charge[i] == 0.0) continue;
            for (int j = ns - 1; j >= i; --j) {
                if (p->charge[j] == 0.0) continue;
                int st = stage++;
                x.stage = st;
                int a0 = cs[i][c], na = cs[i][c + 1] - a0;
                int b0 = cs[j][c], nb = cs[j][c + 1] - b0;
                double lnL = p->coulomb_log[i][j];
                double wi = p->weight[i], wj = p->weight[j];
                if (i == j) {
                    if (na <= 1 || !p->allow_intra) continue;
                    keyed_perm(p->seed, p->step, st, i, id[i], a0, na, pa);
                    double n = na * wi / Vc;
                    if (na % 2 == 0) {
                        for (int x2 = 0; x2 < na - 1; x2 += 2)
                            collide(&x, i, pa[x2], i, pa[x2 + 1], lnL, n, n, n, x2 / 2);
                    } else if (!ny) {
                        double n12 = (na + 3.0) * wi / Vc;
                        collide(&x, i, pa[0], i, pa[1], lnL, n, n, n12, 0);
                        collide(&x, i, pa[1], i, pa[2], lnL, n, n, n12, 1);
                        collide(&x, i, pa[0], i, pa[2], lnL, n, n, n12, 2);
                        for (int x2 = 3; x2 < na - 1; x2 += 2)
                            collide(&x, i, pa[x2], i, pa[x2 + 1], lnL, n, n, n12, 3 + (x2 - 3) / 2);
                    } else {
                        double n12 = 2.0 * wi * ((na - 1) / 2 + 1) / Vc;
                        for (int x2 = 0; x2 < na - 2; x2 += 2)
                            collide(&x, i, pa[x2], i, pa[x2 + 1], lnL, n, n, n12, x2 / 2);
                        collide(&x, i, pa[na - 1], i, pa[0], lnL, n, n, n12, (na - 1) / 2);
                    }
                } else {
                    if (na == 0 || nb == 0) continue;
                    keyed_perm(p->seed, p->step, st, i, id[i], a0, na, pa);
                    keyed_perm(p->seed, p->step, st, j, id[j], b0, nb, pb);
                    if (na == nb) {
                        double nP = na * wi / Vc, nQ = nb * wj / Vc;
                        double nPQ = ny ? na * dmin(wi, wj) / Vc : nP;
                        for (int k = 0; k < na; ++k)
                            collide(&x, i, pa[k], j, pb[na - k - 1], lnL, nP, nQ, nPQ, k);
                    } else {
                        int minIsA = na < nb;
                        int nMin = minIsA ? na : nb, nMax = minIsA ? nb : na;
                        int* pmin = minIsA ? pa : pb;
                        int* pmax = minIsA ? pb : pa;
                        int smin = minIsA ? i : j, smax = minIsA ? j : i;
                        double nMinS = nMin * p->weight[smin] / Vc, nMaxS = nMax * p->weight[smax] / Vc;
                        double lnLp = p->coulomb_log[smax][smin];
                        if (ny) {
                            /* NanbuYonemuraPairing.C:253-300: min parcel l mod nMin with max parcel l */
                            double n12 = nMax * dmin(wi, wj) / Vc;
                            for (int l = 0; l < nMax; ++l)
                                collide(&x, smax, pmax[l], smin, pmin[l % nMin], lnLp, nMaxS, nMinS, n12, l);
                        } else {
                            /* TakizukaAbePairing.C:197-308 */
                            int delta = nMax / nMin;
                            double delta_r = (double)(nMax % nMin) / nMin;
                            int nMinG1 = (int)(delta_r * nMin);
                            int nMaxG1 = (int)((delta + 1) * delta_r * nMin);
                            int nMinG2 = (int)((1.0 - delta_r) * nMin);
                            int nMaxG2 = (int)(delta * (1.0 - delta_r) * nMin);
                            if (nMinG1 > 0) {
                                int s1 = 0, s2 = 0;
                                for (int g1 = 0; g1 < nMaxG1; ++g1) {
                                    if (s1 >= nMinG1) s1 = 0;
                                    collide(&x, smax, pmax[s2], smin, pmin[s1], lnLp, nMaxS, nMinS, nMaxS, g1);
                                    ++s1; ++s2;
                                }
                            }
                            int s1 = nMinG1, s2 = nMaxG1;
                            for (int g2 = 0; g2 < nMaxG2; ++g2) {
                                if (s1 >= nMinG2 + nMinG1) s1 = nMinG1;
                                collide(&x, smax, pmax[s2], smin, pmin[s1], lnLp, nMaxS, nMinS, nMaxS, nMaxG1 + g2);
                                ++s1; ++s2;
                            }
                        }
                    }
                }
            }
        }
    }
    free(pa); free(pb);
    if (stats) {
#pragma omp critical
        for (int k = 0; k < 8; ++k) stats[k] += lst[k];
    }
    }
}

/* number of host threads the oracle's parallel loops use (CPU baseline timing) */
void ref_set_threads(int n) {
#ifdef _OPENMP
    omp_set_num_threads(n > 0 ? n : 1);
#else
    (void)n;
#endif
}

/* ------------------------------------------------------------------ MCC */

double ref_interpolate_xy(double x, int n, const double* xs, const double* ys) {
    int lo = 0;
    for (lo = 0; lo < n && xs[lo] > x; ++lo) {}
    int low = lo;
    if (low < n)
        for (int i = low; i < n; ++i) if (xs[i] > xs[lo] && xs[i] <= x) lo = i;
    int hi = 0;
    for (hi = 0; hi < n && xs[hi] < x; ++hi) {}
    int high = hi;
    if (high < n)
        for (int i = high; i < n; ++i) if (xs[i] < xs[hi] && xs[i] >= x) hi = i;
    if (lo < n && hi < n && lo != hi) return ys[lo] + ((x - xs[lo]) / (xs[hi] - xs[lo])) * (ys[hi] - ys[lo]);
    if (lo == hi) return ys[lo];
    if (lo == n) return ys[hi];
    return ys[lo];
}

static double sigma_of(int n, const double* e, const double* s, double eV) {
    if (n <= 0) return 0.0;
    return ref_interpolate_xy(eV, n, e, s);
}

/* NanbuElectronNeutralCollision::updateVelocity (NanbuElectronNeutralCollision.C:98-130) */
static void nanbu_en_scatter(double eV, v3* UE, v3* Un, double mE, double mN, double uphi, double uchi) {
    double phi = 2.0 * PI_D * uphi;
    double cc;
    if (eV < 1e-30) cc = 1.0;
    else cc = 1.0 / eV * (eV + 2. - 2. * pow(1. + eV, uchi));
    double sc;
    if (cc * cc > 1.0) { cc = 1.0; sc = 0.0; }
    else sc = sqrt(1.0 - cc * cc);
    v3 g = vsub(*UE, *Un);
    double gm = vmag(g), gp = sqrt(g.y * g.y + g.z * g.z);
    if (!(gp > 0.0)) return;
    double cp = cos(phi), sp = sin(phi);
    v3 h = V(gp * cp, -1.0 * (g.y * g.x * cp + gm * g.z * sp) / gp, -1.0 * (g.z * g.x * cp - gm * g.y * sp) / gp);
    v3 dv = vadd(vscale(1.0 - cc, g), vscale(sc, h));
    *UE = vsub(*UE, vscale(mN / (mN + mE), dv));
    *Un = vadd(*Un, vscale(mE / (mN + mE), dv));
}

/* RelativisticElectronNeutralCollision::updateVelocity with a sampled partner velocity
 * (RelativisticElectronNeutralCollision.C:161-240).  cos(chi) is clamped to [-1, 1] as in
 * the two-parcel variant (:120-126); the sampled variant would produce a NaN there. */
static void rel_en_scatter(double eV, v3* UE, v3* Un, double mE, double mN, double ueps, double uchi) {
    com_frame f;
    com_enter(*UE, *Un, mE, mN, &f);
    if (f.pm == 0.0) return;
    double cc, sc;
    if (eV < 1e-30) { cc = 1.0; sc = 0.0; }
    else {
        cc = 1.0 / eV * (eV + 2. - 2. * pow(1. + eV, uchi));
        if (cc * cc > 1.0) { cc = 1.0; sc = 0.0; }
        else sc = sqrt(1.0 - cc * cc);
    }
    com_leave(&f, mE, mN, cc, sc, 2.0 * PI_D * ueps, UE, Un);
}

static void en_scatter(const ref_mcc_params* p, double eV, v3* UE, v3* Un, double us, double uc) {
    if (p->relativistic_scatter) rel_en_scatter(eV, UE, Un, p->massE, p->massN, us, uc);
    else nanbu_en_scatter(eV, UE, Un, p->massE, p->massN, us, uc);
}

/* one electron-neutral scattering on (UE, Un) in place: Nanbu (relativistic = 0) or the
 * relativistic centre-of-momentum variant (1) */
void ref_en_scatter(int relativistic, double eV, double* UE, double* Un, double mE, double mN, double ueps,
                    double uchi) {
    v3 e = ld3(UE), n = ld3(Un);
    if (relativistic) rel_en_scatter(eV, &e, &n, mE, mN, ueps, uchi);
    else nanbu_en_scatter(eV, &e, &n, mE, mN, ueps, uchi);
    st3(UE, e); st3(Un, n);
}

static inline void mcc_rand(uint64_t seed, int64_t step, int cell, int k, int j, double* a, double* b) {
    uint32_t o[4];
    rng4(seed, (uint32_t)cell, (uint32_t)step, (uint32_t)k, (TAG_MCC << 24) | (uint32_t)j, o);
    *a = u01(o[0], o[1]);
    *b = u01(o[2], o[3]);
}

static inline void box_muller(double a, double b, double* n1, double* n2) {
    double r = sqrt(-2.0 * log(1.0 - a));
    *n1 = r * cos(2.0 * PI_D * b);
    *n2 = r * sin(2.0 * PI_D * b);
}

/* order of a cell's electrons for the candidate draw: 52 bits of the splitmix64 finaliser
 * of the id, then the position in the cell's list (a fixed, id-determined order that does
 * not depend on where the sort left the parcels) */
typedef struct { uint64_t key; int idx; } idrec;
static uint64_t mcc_order_key(uint32_t id) {
    uint64_t z = (uint64_t)id + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return z >> 12;
}
static int idrec_cmp(const void* a, const void* b) {
    const idrec* x = (const idrec*)a; const idrec* y = (const idrec*)b;
    if (x->key != y->key) return x->key < y->key ? -1 : 1;
    return x->idx < y->idx ? -1 : (x->idx > y->idx);
}

/* ElectronNeutralCollisionModel::performBackgroundCollisions (ElectronNeutralCollisionModel.C:703-829)
 * with the elastic/excitation/ionization events (:504-700); candidates are drawn with
 * replacement from the cell's electrons taken in mcc_order_key order. */
int ref_mcc(const ref_mesh* m, const ref_mcc_params* p, double dt, double* yE, double* UE, int* tetE,
            uint32_t* idE, const int* cs, int* n_e, int cap_e, double* yI, double* UI, int* tetI, uint32_t* idI,
            int* n_i, int cap_i, double* scrmax, double* rem, int64_t* counts) {
    const double c = C_LIGHT;
    int maxn = 1;
    for (int cc = 0; cc < m->n_cells; ++cc) if (cs[cc + 1] - cs[cc] > maxn) maxn = cs[cc + 1] - cs[cc];
    idrec* L = (idrec*)malloc(sizeof(idrec) * maxn);
    double vth = sqrt(KB * p->TN / p->massN);
    double wEwN = p->wE / p->wN, wNwE = 1.0 / wEwN;
    int err = 0;
    for (int cell = 0; cell < m->n_cells; ++cell) {
        int a0 = cs[cell], eC = cs[cell + 1] - a0;
        double smax = scrmax[cell];
        double coll = rem[cell] + eC * p->nN * smax * dt;
        int nc = (int)coll;
        rem[cell] = coll - nc;
        if (counts) counts[0] += nc;
        if (eC == 0 || nc == 0) continue;
        for (int k = 0; k < eC; ++k) { L[k].key = mcc_order_key(idE[a0 + k]); L[k].idx = a0 + k; }
        qsort(L, eC, sizeof(idrec), idrec_cmp);
        for (int k = 0; k < nc; ++k) {
            double upick, uacc, ua, ub, n1, n2, n3, n4;
            mcc_rand(p->seed, p->step, cell, k, 0, &upick, &uacc);
            int pick = (int)(upick * eC);
            if (pick >= eC) pick = eC - 1;
            int ei = L[pick].idx;
            mcc_rand(p->seed, p->step, cell, k, 1, &ua, &ub);
            box_muller(ua, ub, &n1, &n2);
            mcc_rand(p->seed, p->step, cell, k, 2, &ua, &ub);
            box_muller(ua, ub, &n3, &n4);
            v3 Un = vadd(vscale(vth, V(n1, n2, n3)), V(p->driftN[0], p->driftN[1], p->driftN[2]));
            v3 Ue = ld3(UE + 3 * ei);
            double cR = vmag(vsub(Ue, Un));
            double gamma = 1.0 / sqrt(1.0 - ((cR * cR) / (c * c)));
            double eV = (gamma - 1.0) * p->massE * c * c / QE;
            double sigmaT = smax / cR;
            double Qel = sigma_of(p->n_el, p->e_el, p->s_el, eV);
            double Qex = sigma_of(p->n_ex, p->e_ex, p->s_ex, eV);
            double Qiz = sigma_of(p->n_iz, p->e_iz, p->s_iz, eV);
            double Qt = Qel + Qex + Qiz;
            if (Qt * cR > scrmax[cell]) scrmax[cell] = Qt * cR;
            double us, uc;
            mcc_rand(p->seed, p->step, cell, k, 3, &us, &uc);
            if (uacc <= Qel / sigmaT) {
                en_scatter(p, eV, &Ue, &Un, us, uc);
                if (counts) counts[1]++;
            } else if (eV >= p->th_ex && uacc <= (Qel + Qex) / sigmaT) {
                Ue = vscale(sqrt(1.0 - p->th_ex / eV), Ue);
                eV -= p->th_ex;
                en_scatter(p, eV, &Ue, &Un, us, uc);
                if (counts) counts[2]++;
            } else if (eV >= p->th_iz && uacc <= Qt / sigmaT) {
                double uE, uR, up2, uc2;
                mcc_rand(p->seed, p->step, cell, k, 4, &uE, &uR);
                mcc_rand(p->seed, p->step, cell, k, 5, &up2, &uc2);
                double E0 = 2.0 - 100.0 / (eV + 10.0);
                double E1 = (eV - p->th_iz) / 2.0 - E0;
                double alpha = 10.3;
                double Ep = E0 + alpha * tan(uE * (atan(E1 / alpha) + atan(E0 / alpha)) - atan(E0 / alpha));
                v3 preUe = Ue, preUn = Un;
                Ue = vscale(sqrt(1.0 - (p->th_iz + Ep) / eV), Ue);
                eV -= (p->th_iz + Ep);
                en_scatter(p, eV, &Ue, &Un, us, uc);
                int created = 0;
                /* the ion and the electron are created together or not at all (a full
                 * species is an error; the reference's lists have no capacity) */
                int make = uR <= wEwN;
                int mk_i = make && p->create_ion, mk_e = make && p->create_electron;
                int fit = (!mk_i || *n_i < cap_i) && (!mk_e || *n_e < cap_e);
                if (make && !fit) err = -1;
                if (mk_i && fit) {
                    int ni = (*n_i)++;
                    memcpy(yI + 3 * ni, yE + 3 * ei, sizeof(double) * 3);
                    tetI[ni] = tetE[ei];
                    st3(UI + 3 * ni, Un);
                    uint32_t o[4];
                    rng4(p->seed, idE[ei], (uint32_t)p->step, (uint32_t)k, (TAG_NEWID << 24) | 1u, o);
                    idI[ni] = o[0];
                    created = 1;
                }
                if (uR > wNwE) Ue = preUe;
                if (mk_e && fit) {
                    int ne = (*n_e)++;
                    memcpy(yE + 3 * ne, yE + 3 * ei, sizeof(double) * 3);
                    tetE[ne] = tetE[ei];
                    v3 Unew = vscale(sqrt(2.0 * Ep * QE / p->massE), vdiv(preUe, vmag(preUe)));
                    v3 Un2 = preUn;
                    en_scatter(p, Ep, &Unew, &Un2, up2, uc2);
                    st3(UE + 3 * ne, Unew);
                    uint32_t o[4];
                    rng4(p->seed, idE[ei], (uint32_t)p->step, (uint32_t)k, (TAG_NEWID << 24) | 2u, o);
                    idE[ne] = o[0];
                }
                if (counts && created) counts[3]++;
            }
            st3(UE + 3 * ei, Ue);
        }
    }
    free(L);
    return err;
}

/* ------------------------------------------------------------------ ion-neutral MCC */

#define TAG_ION_MCC 7u

static inline void ion_rand(uint64_t seed, int64_t step, int cell, int k, int j, double* a, double* b) {
    uint32_t o[4];
    rng4(seed, (uint32_t)cell, (uint32_t)step, (uint32_t)k, (TAG_ION_MCC << 24) | (uint32_t)j, o);
    *a = u01(o[0], o[1]);
    *b = u01(o[2], o[3]);
}

/* IonIsotropicScattering::updateVelocity with a sampled partner (IonIsotropicScattering.C:
 * 108-140): the relative velocity turned to a uniformly random direction in the centre of
 * mass, cos(theta) = 2 u1 - 1, phi = 2 pi u2 */
void ref_ion_isotropic(double* UP, double* UQ, double mP, double mQ, double u1, double u2) {
    v3 P = ld3(UP), Q = ld3(UQ);
    v3 Ucm = vscale(1.0 / (mP + mQ), vadd(vscale(mP, P), vscale(mQ, Q)));
    double cR = vmag(vsub(P, Q));
    double ct = 2.0 * u1 - 1.0;
    double st = sqrt(1.0 - ct * ct);
    double phi = 2.0 * PI_D * u2;
    v3 rel = vscale(cR, V(ct, st * cos(phi), st * sin(phi)));
    st3(UP, vadd(Ucm, vscale(mQ / (mP + mQ), rel)));
    st3(UQ, vsub(Ucm, vscale(mP / (mP + mQ), rel)));
}

/* IonNeutralCollisionModel::performBackgroundCollisions (IonNeutralCollisionModel.C:
 * 447-560): per cell nI n sigma_cR_max dt candidates plus the remainder, each an ion drawn
 * with replacement (the cell's ions in mcc_order_key order) against a partner sampled
 * from the background Maxwellian; elastic (isotropic) with probability Qel cR / max, else
 * charge exchange (the ion takes the partner's velocity) up to Qt cR / max above the
 * threshold.  sigma_cR_max grows to the largest Qt cR met, while the step's acceptance uses
 * the value it started with (as the reference's local copy does). */
int ref_ion_mcc(const ref_mesh* m, const ref_ion_mcc_params* p, double dt, double* UI, const uint32_t* idI,
                const int* cs, double* scrmax, double* rem, int64_t* counts) {
    const double c = C_LIGHT;
    int maxn = 1;
    for (int cc = 0; cc < m->n_cells; ++cc) if (cs[cc + 1] - cs[cc] > maxn) maxn = cs[cc + 1] - cs[cc];
    idrec* L = (idrec*)malloc(sizeof(idrec) * maxn);
    double vth = sqrt(KB * p->TN / p->massN);
    for (int cell = 0; cell < m->n_cells; ++cell) {
        int a0 = cs[cell], nI = cs[cell + 1] - a0;
        double smax = scrmax[cell];
        double coll = rem[cell] + nI * p->nN * smax * dt;
        int nc = (int)coll;
        rem[cell] = coll - nc;
        if (counts) counts[0] += nc;
        if (nI == 0 || nc == 0) continue;
        for (int k = 0; k < nI; ++k) { L[k].key = mcc_order_key(idI[a0 + k]); L[k].idx = a0 + k; }
        qsort(L, nI, sizeof(idrec), idrec_cmp);
        for (int k = 0; k < nc; ++k) {
            double upick, uacc, ua, ub, n1, n2, n3, n4;
            ion_rand(p->seed, p->step, cell, k, 0, &upick, &uacc);
            int pick = (int)(upick * nI);
            if (pick >= nI) pick = nI - 1;
            int ii = L[pick].idx;
            ion_rand(p->seed, p->step, cell, k, 1, &ua, &ub);
            box_muller(ua, ub, &n1, &n2);
            ion_rand(p->seed, p->step, cell, k, 2, &ua, &ub);
            box_muller(ua, ub, &n3, &n4);
            v3 Un = vadd(vscale(vth, V(n1, n2, n3)), V(p->driftN[0], p->driftN[1], p->driftN[2]));
            v3 Ui = ld3(UI + 3 * ii);
            double cR = vmag(vsub(Ui, Un));
            double gamma = 1.0 / sqrt(1.0 - ((cR * cR) / (c * c)));
            double eV = (gamma - 1.0) * p->massI * c * c / QE;
            double sigmaT = smax / cR;
            double Qel = sigma_of(p->n_el, p->e_el, p->s_el, eV);
            double Qcx = sigma_of(p->n_cx, p->e_cx, p->s_cx, eV);
            double Qt = Qel + Qcx;
            if (Qt * cR > scrmax[cell]) scrmax[cell] = Qt * cR;
            if (uacc <= Qel / sigmaT) {
                double u1, u2, q[3] = {Un.x, Un.y, Un.z}, P[3] = {Ui.x, Ui.y, Ui.z};
                ion_rand(p->seed, p->step, cell, k, 3, &u1, &u2);
                ref_ion_isotropic(P, q, p->massI, p->massN, u1, u2);
                Ui = ld3(P);
                if (counts) counts[1]++;
            } else if (eV >= p->th_cx && uacc <= Qt / sigmaT) {
                Ui = Un;
                if (counts) counts[2]++;
            }
            st3(UI + 3 * ii, Ui);
        }
    }
    free(L);
    return 0;
}

/* ------------------------------------------------------------------ parcel pairs */

#define TAG_ION_PAIR 8u
#define TAG_DSMC 9u
#define TAG_DSMC_BG 10u
#define VSMALL 1e-300

static inline void pair_rand(uint32_t tag, uint64_t seed, int64_t step, int cell, int k, int j, double* a, double* b) {
    uint32_t o[4];
    rng4(seed, (uint32_t)cell, (uint32_t)step, (uint32_t)k, (tag << 24) | (uint32_t)j, o);
    *a = u01(o[0], o[1]);
    *b = u01(o[2], o[3]);
}

/* the order of a cell's parcels in the pair passes: species after species (list order),
 * each by the top 49 bits of the splitmix64 finaliser of its id, then by its position in
 * the cell's list */
typedef struct { uint64_t key; int s, idx; } prec;
static int prec_cmp(const void* a, const void* b) {
    const prec* x = (const prec*)a; const prec* y = (const prec*)b;
    return x->key < y->key ? -1 : (x->key > y->key);
}
static uint64_t pair_key(int s, uint32_t id, int k) {
    uint64_t z = (uint64_t)id + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return ((uint64_t)s << 61) | ((z >> 15) << 12) | (uint64_t)k;
}
/* the cell's parcels of species [0, ns) in that order; returns their number */
static int cell_list(const ref_pair_params* p, int cell, prec* L, int* nA, double* wA, double* wB) {
    int n = 0;
    *nA = 0; *wA = 0.0; *wB = 0.0;
    for (int s = 0; s < p->ns; ++s) {
        int a0 = p->cell_start[s][cell], c = p->cell_start[s][cell + 1] - a0;
        for (int k = 0; k < c; ++k) {
            L[n].key = pair_key(s, p->id[s][a0 + k], k);
            L[n].s = s;
            L[n].idx = a0 + k;
            ++n;
        }
        if (s < p->na) {
            *nA += c;
            if (c > 0 && p->w[s] > *wA) *wA = p->w[s];
        } else if (c > 0 && p->w[s] > *wB) {
            *wB = p->w[s];
        }
    }
    qsort(L, n, sizeof(prec), prec_cmp);
    return n;
}

/* HardSphere::sigmaTcR (HardSphere.C:50-82) and VariableHardSphere::sigmaTcR
 * (VariableHardSphere.C:51-100): d_PQ and omega_PQ the means of the two species' */
double ref_sigma_tcr(int vhs, double tref, double dP, double dQ, double omP, double omQ, double mP, double mQ,
                     const double* UP, const double* UQ) {
    double dPQ = 0.5 * (dP + dQ);
    double cR = vmag(vsub(ld3(UP), ld3(UQ)));
    if (cR < VSMALL) return 0.0;
    if (!vhs) {
        double sigmaTPQ = PI_D * dPQ * dPQ;
        return sigmaTPQ * cR;
    }
    double omegaPQ = 0.5 * (omP + omQ);
    double mR = mP * mQ / (mP + mQ);
    double sigmaTPQ = PI_D * dPQ * dPQ * pow(2.0 * KB * tref / (mR * cR * cR), omegaPQ - 0.5) / exp(lgamma(2.5 - omegaPQ));
    return sigmaTPQ * cR;
}

/* NanbuWeightCorrectionModel::correctVelocity (NanbuCorrection.C:50-69) */
static void nanbu_weight_correct(double uw, double wP, double wQ, double* UP, double* UQ, const double* preP,
                                 const double* preQ) {
    if (uw > wP / wQ) { UQ[0] = preQ[0]; UQ[1] = preQ[1]; UQ[2] = preQ[2]; }
    if (uw > wQ / wP) { UP[0] = preP[0]; UP[1] = preP[1]; UP[2] = preP[2]; }
}

/* BinaryCollisionModel::performCollisions (BinaryCollisionModel.C:229-425) and, with a
 * background, performBackgroundCollisions (:427-525), over the na neutral species.  Per cell:
 * the no-time-counter candidates rem + 1/2 n (n - 1) w_max (sigma_T cR)_max dt / V; the first
 * parcel drawn from the cell, the second from the first one's octant about the cell centre
 * when that holds another parcel (else from the cell), redrawn while it is the first (draw
 * 16 + r; after 64 equal draws the next parcel of the list); accepted when sigma_T cR /
 * (sigma_T cR)_max exceeds a uniform; IsotropicScattering (IsotropicScattering.C:61-105) and
 * the weight correction.  (sigma_T cR)_max grows to the largest value met while the step
 * accepts against the value it started with.  The background pass draws n n_g max dt
 * candidates, with the maximum the pair pass left, against Maxwellian partners.
 * counts[0..3] += pair candidates, pair collisions, background candidates, collisions. */
int ref_binary(const ref_mesh* m, const ref_pair_params* p, double dt, double* scrmax, double* rem, double* rem_bg,
               int64_t* counts) {
    int maxn = 2;
    for (int cell = 0; cell < m->n_cells; ++cell) {
        int n = 0;
        for (int s = 0; s < p->ns; ++s) n += p->cell_start[s][cell + 1] - p->cell_start[s][cell];
        if (n > maxn) maxn = n;
    }
    prec* L = (prec*)malloc(sizeof(prec) * maxn);
    int* sub = (int*)malloc(sizeof(int) * maxn);
    int* sl = (int*)malloc(sizeof(int) * maxn);
    double vth = p->bg ? sqrt(KB * p->TN / p->massN) : 0.0;
    for (int cell = 0; cell < m->n_cells; ++cell) {
        int nA;
        double wA, wB;
        int n = cell_list(p, cell, L, &nA, &wA, &wB);
        if (n > 1) {
            double smax0 = scrmax[cell];
            double sel = rem[cell] + 0.5 * n * (n - 1) * wA * smax0 * dt / m->V[cell];
            int nc = (int)sel;
            rem[cell] = sel - nc;
            if (counts) counts[0] += nc;
            if (nc > 0) {
                /* sub-cells: the octant of the position about the cell centre */
                v3 cC = ld3(m->C + 3 * cell);
                for (int r = 0; r < n; ++r) {
                    int s = L[r].s, i = L[r].idx;
                    v3 X[4]; int vid[3];
                    tet_vertices(m, p->tet[s][i], X, vid);
                    double y1 = p->y[s][3 * i], y2 = p->y[s][3 * i + 1], y3 = p->y[s][3 * i + 2];
                    double y0 = 1.0 - (y1 + y2 + y3);
                    v3 pos = vadd(vadd(vscale(y0, X[0]), vscale(y1, X[1])), vadd(vscale(y2, X[2]), vscale(y3, X[3])));
                    v3 rel = vsub(pos, cC);
                    /* an empty direction counts as zero: reduced-dimension parcels sit on the
                     * mid-plane, where the sign of the difference is rounding */
                    int em = m->track_mask;
                    sub[r] = ((em & 1) || rel.x >= 0.0) + 2 * ((em & 2) || rel.y >= 0.0) + 4 * ((em & 4) || rel.z >= 0.0);
                }
                int sst[9], at = 0;
                for (int sc = 0; sc < 8; ++sc) {
                    sst[sc] = at;
                    for (int r = 0; r < n; ++r) if (sub[r] == sc) sl[at++] = r;
                }
                sst[8] = at;
                for (int k = 0; k < nc; ++k) {
                    double upick, uacc;
                    pair_rand(TAG_DSMC, p->seed, p->step, cell, k, 0, &upick, &uacc);
                    int P = (int)(upick * n);
                    if (P >= n) P = n - 1;
                    int sc = sub[P], s0 = sst[sc], nsc = sst[sc + 1] - s0;
                    int mm = nsc > 1 ? nsc : n;
                    int Q = -1;
                    for (int r = 0; r < 64 && Q < 0; ++r) {
                        double v0, v1;
                        pair_rand(TAG_DSMC, p->seed, p->step, cell, k, 16 + r, &v0, &v1);
                        int j = (int)(v0 * mm);
                        if (j >= mm) j = mm - 1;
                        int q = nsc > 1 ? sl[s0 + j] : j;
                        if (q != P) Q = q;
                    }
                    if (Q < 0) {
                        if (nsc > 1) {
                            int j = 0;
                            while (sl[s0 + j] != P) ++j;
                            Q = sl[s0 + (j + 1) % nsc];
                        } else {
                            Q = (P + 1) % n;
                        }
                    }
                    int sP = L[P].s, sQ = L[Q].s;
                    double* UP = p->U[sP] + 3 * L[P].idx;
                    double* UQ = p->U[sQ] + 3 * L[Q].idx;
                    double sT = ref_sigma_tcr(p->vhs, p->tref, p->d[sP], p->d[sQ], p->omega[sP], p->omega[sQ],
                                              p->mass[sP], p->mass[sQ], UP, UQ);
                    if (sT > scrmax[cell]) scrmax[cell] = sT;
                    if (sT / smax0 > uacc) {
                        double u1, u2, preP[3] = {UP[0], UP[1], UP[2]}, preQ[3] = {UQ[0], UQ[1], UQ[2]};
                        pair_rand(TAG_DSMC, p->seed, p->step, cell, k, 1, &u1, &u2);
                        ref_ion_isotropic(UP, UQ, p->mass[sP], p->mass[sQ], u1, u2);
                        if (p->weight_correction) {
                            double uw, unused;
                            pair_rand(TAG_DSMC, p->seed, p->step, cell, k, 2, &uw, &unused);
                            nanbu_weight_correct(uw, p->w[sP], p->w[sQ], UP, UQ, preP, preQ);
                        }
                        if (counts) counts[1]++;
                    }
                }
            }
        }
        if (p->bg && n > 0) {
            double smaxb = scrmax[cell];
            double sel = rem_bg[cell] + n * p->nN * smaxb * dt;
            int ncb = (int)sel;
            rem_bg[cell] = sel - ncb;
            if (counts) counts[2] += ncb;
            for (int k = 0; k < ncb; ++k) {
                double upick, uacc, ua, ub, n1, n2, n3, n4;
                pair_rand(TAG_DSMC_BG, p->seed, p->step, cell, k, 0, &upick, &uacc);
                int P = (int)(upick * n);
                if (P >= n) P = n - 1;
                pair_rand(TAG_DSMC_BG, p->seed, p->step, cell, k, 1, &ua, &ub);
                box_muller(ua, ub, &n1, &n2);
                pair_rand(TAG_DSMC_BG, p->seed, p->step, cell, k, 2, &ua, &ub);
                box_muller(ua, ub, &n3, &n4);
                double Un[3] = {vth * n1 + p->driftN[0], vth * n2 + p->driftN[1], vth * n3 + p->driftN[2]};
                int sP = L[P].s;
                double* UP = p->U[sP] + 3 * L[P].idx;
                double sT = ref_sigma_tcr(p->vhs, p->tref, p->d[sP], p->dN, p->omega[sP], p->omegaN, p->mass[sP],
                                          p->massN, UP, Un);
                if (sT > scrmax[cell]) scrmax[cell] = sT;
                if (sT / smaxb > uacc) {
                    double u1, u2;
                    pair_rand(TAG_DSMC_BG, p->seed, p->step, cell, k, 3, &u1, &u2);
                    ref_ion_isotropic(UP, Un, p->mass[sP], p->massN, u1, u2);
                    if (counts) counts[3]++;
                }
            }
        }
    }
    free(L); free(sub); free(sl);
    return 0;
}

/* IonNeutralCollisionModel::performCollisions (IonNeutralCollisionModel.C:580-713): species
 * [0, na) the neutrals, [na, ns) the ions.  Per cell rem + n_N n_I w_max (sigma cR)_max dt / V
 * candidates, each a neutral and an ion drawn with replacement; the ion energy
 * (gamma - 1) m_I c^2 / e of their relative speed; elastic (IonIsotropicScattering with the
 * weight correction, the neutral reset when updateNeutralVelocity is false) when a uniform
 * is at most Qel cR / max, else a charge exchange (the identity switch with probabilities
 * w_Q / w_P and w_P / w_Q) up to Qt cR / max above the threshold.  counts[0..2] +=
 * candidates, elastic, charge exchange. */
int ref_ion_pairs(const ref_mesh* m, const ref_pair_params* p, double dt, double* scrmax, double* rem, int64_t* counts) {
    const double c = C_LIGHT;
    int maxn = 2;
    for (int cell = 0; cell < m->n_cells; ++cell) {
        int n = 0;
        for (int s = 0; s < p->ns; ++s) n += p->cell_start[s][cell + 1] - p->cell_start[s][cell];
        if (n > maxn) maxn = n;
    }
    prec* L = (prec*)malloc(sizeof(prec) * maxn);
    for (int cell = 0; cell < m->n_cells; ++cell) {
        int nA;
        double wA, wB;
        int n = cell_list(p, cell, L, &nA, &wA, &wB);
        int nB = n - nA;
        if (nA == 0 || nB == 0) continue;
        double wm = wB > wA ? wB : wA;
        double smax0 = scrmax[cell];
        double sel = rem[cell] + (double)(nA * nB) * wm * smax0 * dt / m->V[cell];
        int nc = (int)sel;
        rem[cell] = sel - nc;
        if (counts) counts[0] += nc;
        for (int k = 0; k < nc; ++k) {
            double uP, uQ;
            pair_rand(TAG_ION_PAIR, p->seed, p->step, cell, k, 0, &uP, &uQ);
            int P = (int)(uP * nA);
            if (P >= nA) P = nA - 1;
            int Q = (int)(uQ * nB);
            if (Q >= nB) Q = nB - 1;
            Q += nA;
            int sP = L[P].s, sQ = L[Q].s;
            double* UP = p->U[sP] + 3 * L[P].idx;
            double* UQ = p->U[sQ] + 3 * L[Q].idx;
            double massI = p->mass[sQ];
            double cR = vmag(vsub(ld3(UQ), ld3(UP)));
            double gamma = 1.0 / sqrt(1.0 - ((cR * cR) / (c * c)));
            double eV = (gamma - 1.0) * massI * c * c / QE;
            double Qel = sigma_of(p->n_el, p->e_el, p->s_el, eV);
            double Qcx = sigma_of(p->n_cx, p->e_cx, p->s_cx, eV);
            double Qt = Qel + Qcx;
            if (Qt * cR > scrmax[cell]) scrmax[cell] = Qt * cR;
            double uacc, ucx;
            pair_rand(TAG_ION_PAIR, p->seed, p->step, cell, k, 1, &uacc, &ucx);
            double wP = p->w[sP], wQ = p->w[sQ];
            if (uacc <= Qel * cR / smax0) {
                double u1, u2, preP[3] = {UP[0], UP[1], UP[2]}, preQ[3] = {UQ[0], UQ[1], UQ[2]};
                pair_rand(TAG_ION_PAIR, p->seed, p->step, cell, k, 2, &u1, &u2);
                ref_ion_isotropic(UP, UQ, p->mass[sP], massI, u1, u2);
                if (p->weight_correction) {
                    double uw, unused;
                    pair_rand(TAG_ION_PAIR, p->seed, p->step, cell, k, 3, &uw, &unused);
                    nanbu_weight_correct(uw, wP, wQ, UP, UQ, preP, preQ);
                }
                if (p->keep_neutral) { UP[0] = preP[0]; UP[1] = preP[1]; UP[2] = preP[2]; }
                if (counts) counts[1]++;
            } else if (eV >= p->th_cx && uacc <= Qt * cR / smax0) {
                double tmp[3] = {UP[0], UP[1], UP[2]};
                if (ucx <= wQ / wP) { UP[0] = UQ[0]; UP[1] = UQ[1]; UP[2] = UQ[2]; }
                if (ucx <= wP / wQ) { UQ[0] = tmp[0]; UQ[1] = tmp[1]; UQ[2] = tmp[2]; }
                if (p->keep_neutral) { UP[0] = tmp[0]; UP[1] = tmp[1]; UP[2] = tmp[2]; }
                if (counts) counts[2]++;
            }
        }
    }
    free(L);
    return 0;
}

/* ------------------------------------------------------------------ Poisson */

struct ref_poisson {
    const ref_mesh* m;
    int n, K;               /* cells, max couplings per cell */
    int* nb;                /* [n*K] neighbour cell (-1 unused) */
    double* a;              /* [n*K] coupling coefficient |Sf|*deltaCoeff (>0) */
    double* diag;
    double* bsrc;           /* Dirichlet source per cell (a_b * value) */
    int singular;
    int* colour;
    int ncol;
    int* corder;            /* cells sorted by (colour, index) */
    double* rD;             /* DIC reciprocal diagonal source (not inverted) */
    int* bctype;
    double* bcval;
    /* face interpolation weights for Gauss grad */
    double* w;              /* per face owner weight */
    /* |Sf| x nonOrthCorrectionVector per face, oriented out of the face's owner (zero on
     * non-coupled boundary faces); nonorth = 0 when every one is below 1e-10 |Sf| */
    double* kc;
    int nonorth;
};

static void add_coupling(struct ref_poisson* P, int i, int j, double a) {
    for (int k = 0; k < P->K; ++k) {
        if (P->nb[i * P->K + k] == j) { P->a[i * P->K + k] += a; return; }
        if (P->nb[i * P->K + k] < 0) { P->nb[i * P->K + k] = j; P->a[i * P->K + k] = a; return; }
    }
}

/* OpenFOAM's surfaceInterpolation::makeNonOrthCorrectionVectors: k = nf - d*deltaCoeff on
 * internal and coupled faces (boundary faces of other patches keep k = 0), stored times
 * |Sf| */
static void nonorth_vector(struct ref_poisson* P, int f, v3 Sf, v3 nf, v3 d, double delta_coeff) {
    double mS = vmag(Sf);
    P->kc[3 * f + 0] = mS * (nf.x - d.x * delta_coeff);
    P->kc[3 * f + 1] = mS * (nf.y - d.y * delta_coeff);
    P->kc[3 * f + 2] = mS * (nf.z - d.z * delta_coeff);
}

/* fvm::laplacian with "Gauss linear corrected": the implicit part with nonOrthDeltaCoeffs =
 * 1/max(nf.d, 0.05|d|) (boundary faces: 1/(nf.(Cf - C))), cyclic coupling, fixedValue /
 * zeroGradient patches, and the explicit non-orthogonal correction (ref_poisson_correction)
 * from the potential the solve starts from; DIC is IC(0) in a greedy multicolour ordering
 * (cells coloured in index order). */
ref_poisson* ref_poisson_create(const ref_mesh* m, const int* bct, const double* bcv) {
    ref_poisson* P = (ref_poisson*)calloc(1, sizeof(ref_poisson));
    P->m = m;
    int n = m->n_cells;
    P->n = n;
    int K = 0;
    for (int c = 0; c < n; ++c) if (m->cfo[c + 1] - m->cfo[c] > K) K = m->cfo[c + 1] - m->cfo[c];
    P->K = K;
    P->nb = (int*)malloc(sizeof(int) * n * K);
    for (int i = 0; i < n * K; ++i) P->nb[i] = -1;
    P->a = (double*)calloc(n * K, sizeof(double));
    P->diag = (double*)calloc(n, sizeof(double));
    P->bsrc = (double*)calloc(n, sizeof(double));
    P->w = (double*)calloc(m->n_faces, sizeof(double));
    P->kc = (double*)calloc(3 * (size_t)m->n_faces, sizeof(double));
    P->bctype = (int*)malloc(sizeof(int) * (m->n_patches + 1));
    P->bcval = (double*)malloc(sizeof(double) * (m->n_patches + 1));
    memcpy(P->bctype, bct, sizeof(int) * m->n_patches);
    memcpy(P->bcval, bcv, sizeof(double) * m->n_patches);
    int has_dirichlet = 0;
    for (int f = 0; f < m->n_faces; ++f) {
        v3 Sf = ld3(m->Sf + 3 * f), Cf = ld3(m->Cf + 3 * f);
        double mS = vmag(Sf);
        v3 nf = vdiv(Sf, mS);
        int o = m->owner[f];
        v3 Co = ld3(m->C + 3 * o);
        if (f < m->n_internal) {
            int nb = m->nbr[f];
            v3 Cn = ld3(m->C + 3 * nb);
            v3 d = vsub(Cn, Co);
            double nd = vdot(nf, d), md = 0.05 * vmag(d);
            double a = mS / (nd > md ? nd : md);
            add_coupling(P, o, nb, a);
            add_coupling(P, nb, o, a);
            P->diag[o] += a; P->diag[nb] += a;
            double so = fabs(vdot(Sf, vsub(Cf, Co))), sn = fabs(vdot(Sf, vsub(Cn, Cf)));
            P->w[f] = sn / (so + sn);
            nonorth_vector(P, f, Sf, nf, d, a / mS);
            continue;
        }
        int pi = m->face_patch[f];
        if (m->ptype[pi] == REF_PATCH_CYCLIC) {
            int q = m->pnbr[pi];
            int g = m->pstart[q] + (f - m->pstart[pi]);
            int nb = m->owner[g];
            v3 Cn = vsub(ld3(m->C + 3 * nb), ld3(m->psep + 3 * pi));
            v3 Cfn = Cf; /* this side's face centre */
            double so = fabs(vdot(Sf, vsub(Cfn, Co))), sn = fabs(vdot(Sf, vsub(Cn, Cfn)));
            P->w[f] = sn / (so + sn);
            v3 d = vsub(Cn, Co);
            double nd = vdot(nf, d), md = 0.05 * vmag(d);
            double a = mS / (nd > md ? nd : md);
            nonorth_vector(P, f, Sf, nf, d, a / mS);
            if (nb == o) continue; /* self-coupled: no net contribution */
            add_coupling(P, o, nb, a);
            P->diag[o] += a;
        } else if (P->bctype[pi] == 1) {
            v3 d = vsub(Cf, Co);
            double a = mS / vdot(nf, d);
            P->diag[o] += a;
            P->bsrc[o] += a * P->bcval[pi];
            has_dirichlet = 1;
            P->w[f] = 1.0;
        } else {
            P->w[f] = 1.0;
        }
    }
    P->singular = !has_dirichlet;
    P->nonorth = 0;
    for (int f = 0; f < m->n_faces; ++f) {
        double mS = vmag(ld3(m->Sf + 3 * f));
        if (vmag(ld3(P->kc + 3 * f)) > 1e-10 * mS) P->nonorth = 1;
    }
    if (!P->nonorth) memset(P->kc, 0, 3 * (size_t)m->n_faces * sizeof(double));
    /* greedy colouring in cell order */
    P->colour = (int*)malloc(sizeof(int) * n);
    P->ncol = 0;
    for (int i = 0; i < n; ++i) {
        unsigned long long used = 0;
        for (int k = 0; k < K; ++k) {
            int j = P->nb[i * K + k];
            if (j >= 0 && j < i) used |= 1ull << P->colour[j];
        }
        int col = 0;
        while (used & (1ull << col)) ++col;
        P->colour[i] = col;
        if (col + 1 > P->ncol) P->ncol = col + 1;
    }
    P->corder = (int*)malloc(sizeof(int) * n);
    int t = 0;
    for (int col = 0; col < P->ncol; ++col)
        for (int i = 0; i < n; ++i) if (P->colour[i] == col) P->corder[t++] = i;
    P->rD = (double*)malloc(sizeof(double) * n);
    for (int t2 = 0; t2 < n; ++t2) {
        int i = P->corder[t2];
        double r = P->diag[i];
        for (int k = 0; k < K; ++k) {
            int j = P->nb[i * K + k];
            if (j >= 0 && P->colour[j] < P->colour[i]) r -= P->a[i * K + k] * P->a[i * K + k] / P->rD[j];
        }
        P->rD[i] = r;
    }
    return P;
}

void ref_poisson_destroy(ref_poisson* P) {
    if (!P) return;
    free(P->nb); free(P->a); free(P->diag); free(P->bsrc); free(P->colour); free(P->corder); free(P->rD);
    free(P->bctype); free(P->bcval); free(P->w); free(P->kc);
    free(P);
}

int ref_poisson_n_colours(const ref_poisson* P) { return P->ncol; }
void ref_poisson_colours(const ref_poisson* P, int* col) { memcpy(col, P->colour, sizeof(int) * P->n); }

void ref_poisson_apply(const ref_poisson* P, const double* x, double* y) {
    for (int i = 0; i < P->n; ++i) {
        double s = P->diag[i] * x[i];
        for (int k = 0; k < P->K; ++k) {
            int j = P->nb[i * P->K + k];
            if (j >= 0) s -= P->a[i * P->K + k] * x[j];
        }
        y[i] = s;
    }
}

void ref_poisson_rhs(const ref_poisson* P, const double* rho, double* b) {
    for (int i = 0; i < P->n; ++i) b[i] = rho[i] * P->m->V[i] / EPS0 + P->bsrc[i];
    if (P->singular) {
        double s = 0.0;
        for (int i = 0; i < P->n; ++i) s += b[i];
        s /= P->n;
        for (int i = 0; i < P->n; ++i) b[i] -= s;
    }
}

static void dic_apply(const ref_poisson* P, const double* r, double* z) {
    int n = P->n, K = P->K;
    for (int t = 0; t < n; ++t) {
        int i = P->corder[t];
        double s = r[i];
        for (int k = 0; k < K; ++k) {
            int j = P->nb[i * K + k];
            if (j >= 0 && P->colour[j] < P->colour[i]) s += P->a[i * K + k] * z[j];
        }
        z[i] = s / P->rD[i];
    }
    for (int t = n - 1; t >= 0; --t) {
        int i = P->corder[t];
        double s = 0.0;
        for (int k = 0; k < K; ++k) {
            int j = P->nb[i * K + k];
            if (j >= 0 && P->colour[j] > P->colour[i]) s += P->a[i * K + k] * z[j];
        }
        z[i] = z[i] + s / P->rD[i];
    }
}

static double dotp(int n, const double* a, const double* b) {
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}

/* The explicit part of fvm::laplacian(phiE) with "Gauss linear corrected"
 * (gaussLaplacianScheme: the source takes -V div(|Sf| k . interpolate(grad phi)), with
 * correctedSnGrad::correction = nonOrthCorrectionVectors & linear-interpolated Gauss-linear
 * grad), evaluated on phi; ElectroStaticSolver.C:61-64 assembles it from the potential of
 * the previous step, the field the solve starts from.  Returned as the term that adds to
 * the right-hand side V rho/eps0 of the SPD form used here; zero on an orthogonal mesh. */
int ref_poisson_nonorth(const ref_poisson* P) { return P->nonorth; }

void ref_poisson_correction(const ref_poisson* P, const double* phi, double* corr) {
    const ref_mesh* m = P->m;
    memset(corr, 0, sizeof(double) * m->n_cells);
    if (!P->nonorth) return;
    double* g = (double*)malloc(sizeof(double) * 3 * m->n_cells);
    ref_grad(P, phi, g);
    for (int f = 0; f < m->n_faces; ++f) {
        int o = m->owner[f], nb;
        if (f < m->n_internal) {
            nb = m->nbr[f];
        } else {
            int pi = m->face_patch[f];
            if (m->ptype[pi] != REF_PATCH_CYCLIC) continue;
            int q = m->pnbr[pi];
            nb = m->owner[m->pstart[q] + (f - m->pstart[pi])];
        }
        double w = P->w[f], s = 0.0;
        for (int d = 0; d < 3; ++d) s += P->kc[3 * f + d] * (w * g[3 * o + d] + (1.0 - w) * g[3 * nb + d]);
        corr[o] += s;
        if (f < m->n_internal) corr[nb] -= s;
    }
    free(g);
}

int ref_poisson_solve(const ref_poisson* P, const double* rho, double* phi, double tol, int max_iter, int precond,
                      double* final_rel) {
    int n = P->n;
    double* b = (double*)malloc(sizeof(double) * n);
    double* r = (double*)malloc(sizeof(double) * n);
    double* z = (double*)malloc(sizeof(double) * n);
    double* pp = (double*)malloc(sizeof(double) * n);
    double* q = (double*)malloc(sizeof(double) * n);
    if (P->nonorth) {
        /* right-hand side with the correction of the starting potential, then the zero
         * mean of a singular system */
        ref_poisson_correction(P, phi, q);
        for (int i = 0; i < n; ++i) b[i] = rho[i] * P->m->V[i] / EPS0 + P->bsrc[i] + q[i];
        if (P->singular) {
            double s = 0.0;
            for (int i = 0; i < n; ++i) s += b[i];
            s /= n;
            for (int i = 0; i < n; ++i) b[i] -= s;
        }
    } else {
        ref_poisson_rhs(P, rho, b);
    }
    ref_poisson_apply(P, phi, q);
    for (int i = 0; i < n; ++i) r[i] = b[i] - q[i];
    double bn = sqrt(dotp(n, b, b));
    if (bn == 0.0) bn = 1.0;
    double rn = sqrt(dotp(n, r, r));
    int it = 0;
    if (rn / bn > tol) {
        if (precond) dic_apply(P, r, z); else memcpy(z, r, sizeof(double) * n);
        memcpy(pp, z, sizeof(double) * n);
        double rz = dotp(n, r, z);
        for (it = 1; it <= max_iter; ++it) {
            ref_poisson_apply(P, pp, q);
            double alpha = rz / dotp(n, pp, q);
            for (int i = 0; i < n; ++i) { phi[i] += alpha * pp[i]; r[i] -= alpha * q[i]; }
            rn = sqrt(dotp(n, r, r));
            if (rn / bn <= tol) break;
            if (precond) dic_apply(P, r, z); else memcpy(z, r, sizeof(double) * n);
            double rz2 = dotp(n, r, z);
            double beta = rz2 / rz;
            rz = rz2;
            for (int i = 0; i < n; ++i) pp[i] = z[i] + beta * pp[i];
        }
        if (it > max_iter) it = max_iter;
    }
    if (P->singular) {
        double s = 0.0;
        for (int i = 0; i < n; ++i) s += phi[i];
        s /= n;
        for (int i = 0; i < n; ++i) phi[i] -= s;
    }
    if (final_rel) *final_rel = rn / bn;
    free(b); free(r); free(z); free(pp); free(q);
    return it;
}

/* fvc::grad, Gauss linear (ElectroStaticSolver.C:68) */
void ref_grad(const ref_poisson* P, const double* phi, double* gr) {
    const ref_mesh* m = P->m;
    memset(gr, 0, sizeof(double) * 3 * m->n_cells);
    for (int f = 0; f < m->n_faces; ++f) {
        v3 Sf = ld3(m->Sf + 3 * f);
        int o = m->owner[f];
        double pf;
        if (f < m->n_internal) {
            int nb = m->nbr[f];
            pf = P->w[f] * phi[o] + (1.0 - P->w[f]) * phi[nb];
            gr[3 * nb + 0] -= Sf.x * pf; gr[3 * nb + 1] -= Sf.y * pf; gr[3 * nb + 2] -= Sf.z * pf;
        } else {
            int pi = m->face_patch[f];
            if (m->ptype[pi] == REF_PATCH_EMPTY) continue; /* an empty fvPatch has no faces */
            if (m->ptype[pi] == REF_PATCH_CYCLIC) {
                int q = m->pnbr[pi];
                int g = m->pstart[q] + (f - m->pstart[pi]);
                pf = P->w[f] * phi[o] + (1.0 - P->w[f]) * phi[m->owner[g]];
            } else if (P->bctype[pi] == 1) pf = P->bcval[pi];
            else pf = phi[o];
        }
        gr[3 * o + 0] += Sf.x * pf; gr[3 * o + 1] += Sf.y * pf; gr[3 * o + 2] += Sf.z * pf;
    }
    for (int c = 0; c < m->n_cells; ++c)
        for (int d = 0; d < 3; ++d) gr[3 * c + d] /= m->V[c];
}

/* ------------------------------------------------------------------ SimpleInjection */

#define TAG_INJECT 11u

/* SimpleInjection::injection (SimpleInjection.C:186-229) for one species through
 * ParticleEmitter::emitParticle (ParticleEmitter.C:223-349) and its velocity models
 * (:357-535): `count` new parcels of one injector.  The patch's faces are fanned from
 * their first point into triangles; a triangle is drawn by area (the last whose
 * cumulative start lies below the drawn area), a point in it by the barycentric
 * (1 - t, (1 - s) t, s t) with t = sqrt(u) (OpenFOAM barycentric2D01), moved into the
 * cell by a in [0.1, 0.5) of its normal distance from the owner's centre, its empty
 * components zeroed (meshTools::constrainDirection) and located in the owner cell.  The
 * velocity frame: n the inward normal, t1 from the face's first point towards the face
 * centre, t2 = n x t1.  Charged parcels are pushed by -dt/2 when p->sync is set
 * (ParticleEmitter.C:338, PICParcel.C:214-227).  The draws are Philox keyed by
 * (seed; k, step, injector << 16 | species, TAG_INJECT << 24 | draw). */
/* the patch's faces fanned from their first point: triangle -> face, its second point's
 * index in the face, and the cumulative areas from 0 (ParticleEmitter.C:78-131) */
static int inject_table(const ref_mesh* m, int pt, int** tface, int** tj, double** cum) {
    int ntri = 0;
    for (int f = m->pstart[pt]; f < m->pstart[pt] + m->psize[pt]; ++f) ntri += m->fo[f + 1] - m->fo[f] - 2;
    *tface = (int*)malloc(sizeof(int) * (ntri > 0 ? ntri : 1));
    *tj = (int*)malloc(sizeof(int) * (ntri > 0 ? ntri : 1));
    *cum = (double*)malloc(sizeof(double) * (ntri + 1));
    (*cum)[0] = 0.0;
    int k = 0;
    for (int f = m->pstart[pt]; f < m->pstart[pt] + m->psize[pt]; ++f) {
        const int* fp = m->fp + m->fo[f];
        const int n = m->fo[f + 1] - m->fo[f];
        v3 a = ld3(m->points + 3 * fp[0]);
        for (int j = 1; j + 1 < n; ++j) {
            v3 b = ld3(m->points + 3 * fp[j]), c = ld3(m->points + 3 * fp[j + 1]);
            (*cum)[k + 1] = (*cum)[k] + 0.5 * vmag(vcross(vsub(b, a), vsub(c, a)));
            (*tface)[k] = f; (*tj)[k] = j;
            ++k;
        }
    }
    return ntri;
}

/* One emitted parcel: position, tet and velocity (no synchronisation).  by_area picks the
 * triangle by area (emitParticle, ParticleEmitter.C:232-262), else uniformly over the
 * triangles (emitParticleExplicit, :143-146).  Draws keyed (seed; ka, step, kc, tag << 24 |
 * draw). */
static void emit_one(const ref_mesh* m, const ref_inject_params* p, const int* tface, const int* tj,
                     const double* cum, int ntri, int by_area, uint32_t ka, uint32_t kc, uint32_t tag, double* y,
                     int* tet, double* Uo) {
    const double vt = sqrt(KB * p->T / p->mass);
    const double vmp = sqrt(2.0 * KB * p->T / p->mass);
    uint32_t o[4];
    rng4(p->seed, ka, (uint32_t)p->step, kc, tag << 24, o);
    const double u = u01(o[0], o[1]), s = u01(o[2], o[3]);
    int tri = 0;
    if (by_area) {
        const double area = u * cum[ntri];
        for (int q = ntri - 1; q >= 0; --q)
            if (area > cum[q]) { tri = q; break; }
    } else {
        tri = (int)(u * (double)ntri);
        if (tri > ntri - 1) tri = ntri - 1;
    }
    const int f = tface[tri];
    const int* fp = m->fp + m->fo[f];
    v3 a = ld3(m->points + 3 * fp[0]), b = ld3(m->points + 3 * fp[tj[tri]]), c = ld3(m->points + 3 * fp[tj[tri] + 1]);
    rng4(p->seed, ka, (uint32_t)p->step, kc, (tag << 24) | 1u, o);
    const double t = sqrt(u01(o[0], o[1])), frac = 0.1 + 0.4 * u01(o[2], o[3]);
    v3 pf = vadd(vadd(vscale(1.0 - t, a), vscale((1.0 - s) * t, b)), vscale(s * t, c));
    v3 Sf = ld3(m->Sf + 3 * f);
    v3 nout = vdiv(Sf, vmag(Sf));
    const int own = m->owner[f];
    v3 pc = ld3(m->C + 3 * own);
    v3 d = vscale(fabs(vdot(vsub(pf, pc), nout)), nout);
    v3 x = vsub(pf, vscale(frac, d));
    if (m->track_mask & 1) x.x = 0.0;
    if (m->track_mask & 2) x.y = 0.0;
    if (m->track_mask & 4) x.z = 0.0;
    double xx[3] = {x.x, x.y, x.z};
    ref_locate(m, 1, xx, &own, y, tet);
    /* the frame of the face (ParticleEmitter.C:308-322) */
    v3 nin = vscale(-1.0, nout);
    v3 t1 = vsub(ld3(m->Cf + 3 * f), a);
    t1 = vdiv(t1, vmag(t1));
    v3 t2 = vcross(nin, t1);
    t2 = vdiv(t2, vmag(t2));
    double n1 = 0.0, n2 = 0.0, vn;
    if (p->velocity_model != 1) {
        rng4(p->seed, ka, (uint32_t)p->step, kc, (tag << 24) | 2u, o);
        box_muller(u01(o[0], o[1]), u01(o[2], o[3]), &n1, &n2);
    }
    if (p->velocity_model == 0) {
        /* Bird eqn 12.5: uNormal = 3 (2u - 1) accepted with P = 2 u / sqrt(2) exp(1/2 - u^2) */
        double un = 0.0;
        for (uint32_t j = 4; j < 4 + 2000; ++j) {
            rng4(p->seed, ka, (uint32_t)p->step, kc, (tag << 24) | j, o);
            un = 3.0 * (2.0 * u01(o[0], o[1]) - 1.0);
            double P = un < 0.0 ? -1.0 : 2.0 * un / sqrt(2.0) * exp(0.5 - un * un);
            if (!(P < u01(o[2], o[3]))) break;
        }
        vn = vmp * un;
    } else if (p->velocity_model == 1) {
        vn = vmp;
    } else {
        rng4(p->seed, ka, (uint32_t)p->step, kc, (tag << 24) | 3u, o);
        if (p->velocity_model == 2) {
            double g1, g2;
            box_muller(u01(o[0], o[1]), u01(o[2], o[3]), &g1, &g2);
            vn = fabs(g1) * vt;
        } else {
            vn = vt * sqrt(-2.0 * log(1.0 - u01(o[0], o[1])));
        }
    }
    v3 uu = vadd(vadd(vscale(vt * n1, t1), vscale(vt * n2, t2)), vscale(vn, nin));
    st3(Uo, uu);
}

void ref_inject(const ref_mesh* m, const ref_inject_params* p, int count, double* y, int* tet, double* U,
                uint32_t* id) {
    int *tface, *tj;
    double* cum;
    const int ntri = inject_table(m, p->patch, &tface, &tj, &cum);
    const uint32_t ckey = (p->model << 16) | p->species;
    for (int i = 0; i < count; ++i) {
        uint32_t o[4];
        emit_one(m, p, tface, tj, cum, ntri, 1, (uint32_t)i, ckey, TAG_INJECT, y + 3 * i, tet + i, U + 3 * i);
        if (p->sync && p->q != 0.0 && tet[i] >= 0) {
            v3 E = interp_one(m, p->weighting, tet[i], y + 3 * i, p->Ecell, p->Epoint);
            double e[3] = {E.x, E.y, E.z};
            double zero[3] = {0.0, 0.0, 0.0};
            const double* B = p->Bcell ? p->Bcell + 3 * m->tet_cell[tet[i]] : zero;
            ref_push_one(p->pusher, U + 3 * i, e, B, p->q, p->mass, -0.5 * p->dt, p->hc_reference_form, p->rotate);
        }
        rng4(p->seed, (uint32_t)i, (uint32_t)p->step, ckey, (TAG_NEWID << 24) | 3u, o);
        if (id) id[i] = o[0];
    }
    free(tface); free(tj); free(cum);
}
