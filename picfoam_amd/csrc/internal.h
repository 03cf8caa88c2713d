// Shared host/device declarations of libpicmi (not part of the public ABI).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

namespace picmi {

constexpr double kC = 299792458.0;
constexpr double kEps0 = 8.8541878128e-12;
constexpr double kQe = 1.602176634e-19;
constexpr double kKB = 1.380649e-23;
constexpr double kH = 6.62607015e-34;
constexpr double kPi = 3.14159265358979323846;
constexpr double kFixedScale = 4398046511104.0;  // 2^42 units per charge quantum
constexpr int kMaxTrackSteps = 1000;
// tet-face neighbour code of a face on an empty patch (a reduced-dimension case): a parcel
// that reaches one stops there and keeps its place (other non-coupled patches: -1 - patch)
constexpr int kTetEmptyFace = -900000;
constexpr int kMaxSpecies = 8;
// per-cell index-list slots in LDS of the collision kernels (per-cell pass / large cells)
constexpr int kCoulombSlots = 128, kCoulombSlotsBig = 2048;
constexpr int kMccSlots = 512, kMccSlotsBig = 4096;
// order-key buckets of the one-cell-per-wave MCC passes
constexpr int kMccBins = 512, kMccBinsBig = 1024;

enum Pusher { kPushNone = 0, kPushBoris = 1, kPushBorisNR = 2, kPushVay = 3, kPushHC = 4 };
enum Weighting { kCellAverage = 0, kVolumeWeighting = 1 };

// One tet of the cell decomposition, one 128-byte line: the det-scaled inverse of the tet
// edge matrix (rows r1 = e2 x e3, r2 = e3 x e1, r3 = e1 x e2, det = e1 . r1), the owning
// cell, the coupled (master) point ids of (base, A, B), the neighbouring tet across each
// tet face (face i is opposite vertex i; <0 encodes -1-patch for a non-coupled patch) and,
// per face (byte i), for each slot j of the neighbour the slot of this tet whose barycentric
// coordinate moves there (2 bits each; the dropped vertex's slot, zero on the face, goes to
// the neighbour's new vertex).
struct alignas(128) TetGeo {
    double r[9];
    double det;
    int32_t cell;
    int32_t vm[3];
    int32_t nbr[4];
    uint32_t perm;
    int32_t vid[3];  // raw point ids of (base, A, B) for positions
};
static_assert(sizeof(TetGeo) == 128, "TetGeo must be one cache line");

// Host-side mesh tables built once in pic_create.
struct HostMesh {
    int n_points = 0, n_faces = 0, n_internal = 0, n_cells = 0, n_patches = 0;
    std::vector<double> points;
    std::vector<int32_t> fo, fp, owner, nbr;
    std::vector<int32_t> ptype, pstart, psize, pnbr, phi_bc;
    std::vector<double> phi_val;
    std::vector<double> Cf, Sf, C, V;
    std::vector<int32_t> cfo, cf;              // cell -> faces (OpenFOAM cells() order)
    std::vector<int32_t> cell_tet0;            // [n_cells+1]
    std::vector<int32_t> tet_cell, tet_face, tet_pt;
    std::vector<int32_t> fbase;                // face base point (index into the face's points)
    std::vector<int32_t> pmaster;              // coupled point groups
    std::vector<int32_t> face_patch;
    std::vector<double> psep;                  // per patch separation
    // polyMesh::solutionD: bit d set when direction d is empty (reduced-dimension case);
    // the tracking velocity drops those components
    uint32_t track_mask = 0;
    std::vector<TetGeo> tets;
    // cell points (ascending id) with VolumeWeighting distance weights and
    // volPointInterpolation inverse-distance weights
    std::vector<int32_t> cpo, cp;
    std::vector<double> vw_w, ip_w;
    // master point -> (cell, weight) lists for cell->point interpolation
    std::vector<int32_t> pco, pc_cell;
    std::vector<double> pc_w;
    int n_masters = 0;
};

// Returns empty string on success, otherwise an error message.
std::string build_host_mesh(HostMesh& m);
// point ids (base, A, B) of tet t (after build_host_mesh); with the cell centre they are its
// four vertices
void tet_points(const HostMesh& m, int t, int vid[3]);

// Poisson operator in ELL form (K slots per cell) with multicolour DIC.
struct HostPoisson {
    int n = 0, K = 0, ncol = 0;
    bool singular = false;
    std::vector<int32_t> nb;   // [n*K], -1 unused
    std::vector<double> a;     // [n*K]
    std::vector<double> diag, bsrc, rD;
    std::vector<int32_t> colour, corder, col_start;
    std::vector<double> w;     // per face owner interpolation weight
    // grad: per cell list of (face-centre value source) terms
    std::vector<int32_t> gfo;  // [n+1] cell -> entries
    std::vector<int32_t> g_kind;   // 0: interp between cells (w*self + (1-w)*other), 1: fixed value, 2: self
    std::vector<int32_t> g_other;
    std::vector<double> g_w, g_val, g_S;  // weight of self, value, signed area vector (3 each)
    // |Sf| x nonOrthCorrectionVector per grad entry, pointing out of the cell (3 each; zero
    // on non-coupled boundary faces); nonorth false when every one is below 1e-10 |Sf|
    std::vector<double> g_k;
    bool nonorth = false;
};
std::string build_host_poisson(const HostMesh& m, HostPoisson& p);

}  // namespace picmi
