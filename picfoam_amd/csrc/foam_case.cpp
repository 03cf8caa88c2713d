// The case layer (include/picfoam_case.h): a picFoam case directory read into pic_mesh,
// pic_species and pic_models; the picInitialise models; lagrangian and field files written
// back in OpenFOAM's ASCII format.  Host code; the device is reached through picmi.h only.
#include <dirent.h>
#include <sys/stat.h>
#include <sys/types.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <functional>
#include <memory>
#include <regex>
#include <set>
#include <sstream>
#include <string>
#include <vector>

#include "foam_dict.h"
#include "internal.h"
#include "picfoam_case.h"
#include "picmi.h"

using picmi::foam::Dict;
using picmi::foam::FoamError;
using picmi::foam::Reader;
using picmi::foam::Token;

namespace {

constexpr double kBoltzmann = 1.380649e-23;
constexpr double kElementary = 1.602176634e-19;
constexpr double kPiC = 3.14159265358979323846;

struct V3 {
    double x, y, z;
};
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator*(double s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline V3 at(const std::vector<double>& v, size_t i) { return {v[3 * i], v[3 * i + 1], v[3 * i + 2]}; }
inline V3 arr(const std::array<double, 3>& a) { return {a[0], a[1], a[2]}; }

// ---------------------------------------------------------------- counter-based draws
// Philox4x32-10 (Salmon, Moraes, Dror and Shaw, SC'11), the generator the device kernels
// use, here on the host: counter (a, b, c, d), key = seed.
struct Draw {
    uint32_t w[4];
};
Draw philox(uint64_t seed, uint32_t a, uint32_t b, uint32_t c, uint32_t d) {
    uint32_t x0 = a, x1 = b, x2 = c, x3 = d;
    uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = 0xD2511F53ull * x0, p1 = 0xCD9E8D57ull * x2;
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ x1 ^ k0, n1 = (uint32_t)p1;
        const uint32_t n2 = (uint32_t)(p0 >> 32) ^ x3 ^ k1, n3 = (uint32_t)p0;
        x0 = n0; x1 = n1; x2 = n2; x3 = n3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return {{x0, x1, x2, x3}};
}
inline double u01(uint32_t hi, uint32_t lo) { return (double)((((uint64_t)hi << 32) | lo) >> 11) * (1.0 / 9007199254740992.0); }

// A stream of uniforms for one parcel (or one tet): successive Philox blocks keyed by
// (seed; stream id, model tag, block index).
class Stream {
public:
    Stream(uint64_t seed, uint64_t id, uint32_t tag) : seed_(seed), id_(id), tag_(tag) {}
    double uniform() {
        if (k_ == 2) refill();
        const double u = u01(d_.w[2 * k_], d_.w[2 * k_ + 1]);
        ++k_;
        return u;
    }
    // Box-Muller, one normal per call (the second is discarded to keep the stream simple)
    double normal() {
        const double u1 = uniform(), u2 = uniform();
        return std::sqrt(-2.0 * std::log(1.0 - u1)) * std::cos(2.0 * kPiC * u2);
    }

private:
    void refill() {
        d_ = philox(seed_, (uint32_t)id_, (uint32_t)(id_ >> 32), tag_, block_++);
        k_ = 0;
    }
    uint64_t seed_, id_;
    uint32_t tag_, block_ = 0;
    Draw d_{};
    int k_ = 2;
};

// ---------------------------------------------------------------- the case object

struct HostParcels {
    std::vector<double> x;      // positions [n*3] (from a positions file or an initialiser)
    std::vector<int32_t> cell;  // [n]
    std::vector<double> bary;   // barycentric (y1, y2, y3) [n*3] when tet is set
    std::vector<int32_t> tet;   // [n], or empty: locate x in cell
    std::vector<double> U;      // [n*3]
    size_t size() const { return cell.size(); }
};

struct InitPost {
    bool active = false;
    bool solve_fields = false, leap_frog = false;
};

std::string time_name(double t) {
    char b[32];
    std::snprintf(b, sizeof b, "%.6g", t);
    return b;
}

}  // namespace

struct pic_case {
    std::string dir, error, cloud = "pic";
    // mesh
    std::vector<double> points;
    std::vector<int32_t> fo, fp, owner, nbr;
    int n_cells = 0;
    std::vector<pic_patch> patches;
    std::vector<std::string> patch_names;
    picmi::HostMesh hm;  // geometry and tets (build_host_mesh) for the initialisers
    double mid[3] = {0.0, 0.0, 0.0};  // centre of the mesh's bounding box
    // species and models
    std::vector<std::string> species_names;
    std::vector<pic_species> species;
    pic_models models{};
    std::vector<std::string> model_words;  // storage behind the model name pointers
    double space_charge = 0.0;             // SolverSettings.spaceCharge [A s]
    struct CrossSection {
        int process;
        std::vector<double> e, s;
        double threshold;
    };
    std::vector<CrossSection> xs;
    // (sigma cR)_max of the start time per collision model (0 electron-neutral:
    // picSigmaTcRMax; 1 ion-neutral: picSigmaTcRMax_IonBackground), or how picInitialise's
    // CrossSectionInitialization sets it: FindMaximum over [0, emax) in steps of estep (the
    // default when nothing says otherwise: up to the largest tabulated energy, 1000 eV
    // without one) or Temperature at T [K]
    struct XsInit {
        bool temperature = false;
        double emax = 0.0, estep = 0.0, T = 0.0;
    };
    std::vector<double> smax[2];
    XsInit xs_init[2];
    // the binary model's picSigmaTcRMax and the ion-neutral parcel field
    // picSigmaTcRMax_IonNeutral as read from the start time (READ_IF_PRESENT), and the
    // picInitialise temperatures and number densities BinaryCollisionModel::initialize takes
    std::vector<double> smax_bin, smax_ip;
    // PatchBoundaryModels of type SimpleInjection (time set at creation: the start time)
    std::vector<pic_injection> injection;
    bool binary_init = false;
    std::vector<double> init_T, init_n;
    // time
    double start = 0.0, end = 0.0, dt = 0.0, write_interval = 0.0, now = 0.0;
    bool write_steps = false;  // writeControl timeStep
    std::string now_name;
    // fields of the start time
    std::vector<double> E, B;  // [n_cells*3], empty = zero
    // boundaryT and boundaryU as read (rewritten into every time written, as AUTO_WRITE does)
    std::vector<std::pair<std::string, std::string>> wall_files;
    // parcels
    std::vector<HostParcels> parcels;
    InitPost post;
    bool located = false;

    int fail(int code, const std::string& msg) {
        error = msg;
        return code;
    }
};

namespace {

using Case = pic_case;

// ---------------------------------------------------------------- polyMesh

int patch_type(const std::string& t, const std::string& where) {
    if (t == "patch") return PIC_PATCH_PATCH;
    if (t == "wall") return PIC_PATCH_WALL;
    if (t == "cyclic") return PIC_PATCH_CYCLIC;
    if (t == "empty") return PIC_PATCH_EMPTY;
    throw FoamError(where + ": patch type '" + t + "' is not supported (patch, wall, cyclic, empty)");
}

std::vector<long long> read_labels(const std::string& path) {
    picmi::foam::ListFile lf = picmi::foam::read_list_file(path);
    Reader r(lf.tokens, lf.where, lf.body);
    const long long n = r.list_open();
    std::vector<long long> v;
    if (n > 0) v.reserve((size_t)n);
    while (!r.accept(')')) v.push_back(r.label());
    if (n >= 0 && (long long)v.size() != n) r.fail("list size differs from its count");
    return v;
}

void read_polymesh(Case& c, const std::string& pm) {
    {
        picmi::foam::ListFile lf = picmi::foam::read_list_file(pm + "/points");
        Reader r(lf.tokens, lf.where, lf.body);
        const long long n = r.list_open();
        while (!r.accept(')')) {
            auto p = r.vector3();
            c.points.insert(c.points.end(), p.begin(), p.end());
        }
        if (n >= 0 && (long long)c.points.size() != 3 * n) r.fail("list size differs from its count");
    }
    {
        picmi::foam::ListFile lf = picmi::foam::read_list_file(pm + "/faces");
        Reader r(lf.tokens, lf.where, lf.body);
        const long long n = r.list_open();
        c.fo.assign(1, 0);
        while (!r.accept(')')) {
            const long long k = r.list_open();
            long long got = 0;
            while (!r.accept(')')) { c.fp.push_back((int32_t)r.label()); ++got; }
            if (k >= 0 && got != k) r.fail("face size differs from its count");
            c.fo.push_back((int32_t)c.fp.size());
        }
        if (n >= 0 && (long long)c.fo.size() - 1 != n) r.fail("list size differs from its count");
    }
    for (long long v : read_labels(pm + "/owner")) c.owner.push_back((int32_t)v);
    for (long long v : read_labels(pm + "/neighbour")) c.nbr.push_back((int32_t)v);
    const int nf = (int)c.fo.size() - 1;
    if ((int)c.owner.size() != nf) throw FoamError(pm + "/owner: " + std::to_string(c.owner.size()) +
                                                   " entries for " + std::to_string(nf) + " faces");
    int nc = 0;
    for (int32_t o : c.owner) nc = std::max(nc, o + 1);
    for (int32_t o : c.nbr) nc = std::max(nc, o + 1);
    c.n_cells = nc;
    // boundary: N ( name { type; nFaces; startFace; neighbourPatch; } ... )
    picmi::foam::ListFile lf = picmi::foam::read_list_file(pm + "/boundary");
    const std::vector<Token>& t = lf.tokens;
    size_t pos = lf.body;
    if (pos < t.size() && t[pos].kind == Token::Number) ++pos;
    if (pos >= t.size() || !t[pos].is('(')) throw FoamError(lf.where + ": expected the patch list");
    ++pos;
    std::vector<std::string> nbr_names;
    while (pos < t.size() && !t[pos].is(')')) {
        const std::string name = t[pos].text;
        ++pos;
        if (pos >= t.size() || !t[pos].is('{')) throw FoamError(lf.where + ": patch " + name + ": expected '{'");
        ++pos;
        auto d = picmi::foam::parse_entries(t, pos, lf.where + "/" + name, true);
        pic_patch p{};
        p.type = patch_type(d->word("type"), d->where);
        p.size = (int32_t)d->label("nFaces");
        p.start = (int32_t)d->label("startFace");
        p.neighbour_patch = -1;
        if (p.type == PIC_PATCH_CYCLIC) {
            const std::string tr = d->word_or("transform", "unspecified");
            if (tr == "rotational") throw FoamError(d->where + ": rotational cyclics are not supported");
        }
        nbr_names.push_back(p.type == PIC_PATCH_CYCLIC ? d->word("neighbourPatch") : "");
        c.patches.push_back(p);
        c.patch_names.push_back(name);
    }
    for (size_t i = 0; i < c.patches.size(); ++i) {
        if (nbr_names[i].empty()) continue;
        auto it = std::find(c.patch_names.begin(), c.patch_names.end(), nbr_names[i]);
        if (it == c.patch_names.end()) throw FoamError(lf.where + ": neighbourPatch " + nbr_names[i] + " not found");
        c.patches[i].neighbour_patch = (int32_t)(it - c.patch_names.begin());
    }
}

// ---------------------------------------------------------------- blockMesh (one hex block)

// blockMesh's cell spacing along an edge of n cells with expansion ratio r (last / first):
// geometric with ratio r^(1/(n-1)) per cell
std::vector<double> grading(int n, double r) {
    std::vector<double> lam(n + 1);
    if (n <= 1 || std::fabs(r - 1.0) < 1e-12) {
        for (int i = 0; i <= n; ++i) lam[i] = (double)i / n;
        return lam;
    }
    const double q = std::pow(r, 1.0 / (n - 1));
    for (int i = 0; i <= n; ++i) lam[i] = (1.0 - std::pow(q, i)) / (1.0 - std::pow(q, n));
    return lam;
}

void block_mesh(Case& c, const std::string& path) {
    auto dp = picmi::foam::read_dict_file(path);
    const Dict& d = *dp;
    const double scale = d.has("convertToMeters") ? d.scalar("convertToMeters") : d.scalar_or("scale", 1.0);
    std::vector<V3> verts;
    {
        Reader r = d.value("vertices");
        r.list_open();
        while (!r.accept(')')) verts.push_back(scale * arr(r.vector3()));
    }
    int hv[8], nn[3];
    double gr[3] = {1.0, 1.0, 1.0};
    {
        Reader r = d.value("blocks");
        r.list_open();
        int nblocks = 0;
        while (!r.accept(')')) {
            if (++nblocks > 1) r.fail("only one hex block is supported");
            if (r.word() != "hex") r.fail("only hex blocks are supported");
            r.expect('(');
            for (int i = 0; i < 8; ++i) {
                hv[i] = (int)r.label();
                if (hv[i] < 0 || hv[i] >= (int)verts.size()) r.fail("block vertex out of range");
            }
            r.expect(')');
            if (r.peek().kind == Token::Word && r.peek().text != "simpleGrading" && r.peek().text != "edgeGrading")
                r.word();  // cell zone name
            r.expect('(');
            for (int a = 0; a < 3; ++a) {
                nn[a] = (int)r.label();
                if (nn[a] < 1) r.fail("block cell counts must be positive");
            }
            r.expect(')');
            const std::string g = r.word();
            if (g == "simpleGrading") {
                r.expect('(');
                for (int a = 0; a < 3; ++a) gr[a] = r.number();
                r.expect(')');
            } else if (g == "edgeGrading") {
                r.expect('(');
                double e[12];
                for (int i = 0; i < 12; ++i) e[i] = r.number();
                r.expect(')');
                for (int a = 0; a < 3; ++a) {
                    for (int i = 1; i < 4; ++i)
                        if (e[4 * a + i] != e[4 * a]) r.fail("edgeGrading with unequal edges is not supported");
                    gr[a] = e[4 * a];
                }
            } else {
                r.fail("expected simpleGrading or edgeGrading");
            }
        }
        if (nblocks == 0) r.fail("no block");
    }
    if (d.has("edges")) {
        Reader r = d.value("edges");
        r.list_open();
        if (!r.accept(')')) r.fail("curved edges are not supported");
    }
    if (d.has("mergePatchPairs")) {
        Reader r = d.value("mergePatchPairs");
        r.list_open();
        if (!r.accept(')')) r.fail("mergePatchPairs are not supported");
    }
    const int nx = nn[0], ny = nn[1], nz = nn[2];
    const std::vector<double> lx = grading(nx, gr[0]), ly = grading(ny, gr[1]), lz = grading(nz, gr[2]);
    V3 X[8];
    for (int i = 0; i < 8; ++i) X[i] = verts[hv[i]];
    // points (i, j, k) by trilinear interpolation of the block corners (OpenFOAM hex order:
    // 0 (000) 1 (100) 2 (110) 3 (010) 4 (001) 5 (101) 6 (111) 7 (011))
    auto pid = [&](int i, int j, int k) { return i + (nx + 1) * (j + (ny + 1) * k); };
    c.points.assign(3 * (size_t)(nx + 1) * (ny + 1) * (nz + 1), 0.0);
    for (int k = 0; k <= nz; ++k)
        for (int j = 0; j <= ny; ++j)
            for (int i = 0; i <= nx; ++i) {
                const double u = lx[i], v = ly[j], w = lz[k];
                const V3 p = (1 - u) * (1 - v) * (1 - w) * X[0] + u * (1 - v) * (1 - w) * X[1] + u * v * (1 - w) * X[2] +
                             (1 - u) * v * (1 - w) * X[3] + (1 - u) * (1 - v) * w * X[4] + u * (1 - v) * w * X[5] +
                             u * v * w * X[6] + (1 - u) * v * w * X[7];
                const size_t q = (size_t)pid(i, j, k);
                c.points[3 * q] = p.x; c.points[3 * q + 1] = p.y; c.points[3 * q + 2] = p.z;
            }
    auto cid = [&](int i, int j, int k) { return i + nx * (j + ny * k); };
    // faces normal to the block's local axis a at index s, in the local +a orientation
    auto face = [&](int a, int s, int p, int q) -> std::array<int, 4> {
        if (a == 0) return {pid(s, p, q), pid(s, p + 1, q), pid(s, p + 1, q + 1), pid(s, p, q + 1)};
        if (a == 1) return {pid(p, s, q), pid(p, s, q + 1), pid(p + 1, s, q + 1), pid(p + 1, s, q)};
        return {pid(p, q, s), pid(p + 1, q, s), pid(p + 1, q + 1, s), pid(p, q + 1, s)};
    };
    c.fo.assign(1, 0);
    c.fp.clear(); c.owner.clear(); c.nbr.clear();
    auto push_face = [&](const std::array<int, 4>& f, int own) {
        for (int v : f) c.fp.push_back(v);
        c.fo.push_back((int32_t)c.fp.size());
        c.owner.push_back(own);
    };
    // internal faces in upper-triangular order: by owner, then neighbour (+i, +j, +k)
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                const int o = cid(i, j, k);
                if (i + 1 < nx) { push_face(face(0, i + 1, j, k), o); c.nbr.push_back(cid(i + 1, j, k)); }
                if (j + 1 < ny) { push_face(face(1, j + 1, i, k), o); c.nbr.push_back(cid(i, j + 1, k)); }
                if (k + 1 < nz) { push_face(face(2, k + 1, i, j), o); c.nbr.push_back(cid(i, j, k + 1)); }
            }
    c.n_cells = nx * ny * nz;
    // the six block sides by their corner sets: (axis, max side)
    const int side_corners[6][4] = {{0, 3, 7, 4}, {1, 2, 6, 5}, {0, 1, 5, 4}, {3, 2, 6, 7}, {0, 1, 2, 3}, {4, 5, 6, 7}};
    auto side_of = [&](const std::vector<int>& q, const std::string& where) {
        std::set<int> s(q.begin(), q.end());
        for (int sd = 0; sd < 6; ++sd) {
            std::set<int> b;
            for (int i = 0; i < 4; ++i) b.insert(hv[side_corners[sd][i]]);
            if (b == s) return sd;
        }
        throw FoamError(where + ": face is not a side of the block");
    };
    // faces of a side in a common transverse order, so that face m of one side matches face
    // m of the opposite side; the min side lists its points reversed from point 0, which
    // makes point 0 the image of the max side's point 0 (OpenFOAM's cyclic ordering)
    auto side_faces = [&](int sd, std::vector<std::array<int, 4>>& fs, std::vector<int>& ow) {
        const int a = sd / 2, mx = sd % 2;
        const int n1 = a == 0 ? ny : nx, n2 = a == 2 ? ny : nz;
        const int s = mx ? nn[a] : 0;
        for (int q = 0; q < n2; ++q)
            for (int p = 0; p < n1; ++p) {
                std::array<int, 4> f = face(a, s, p, q);
                if (!mx) f = {f[0], f[3], f[2], f[1]};
                fs.push_back(f);
                const int ci = a == 0 ? (mx ? nx - 1 : 0) : p, cj = a == 0 ? p : (a == 1 ? (mx ? ny - 1 : 0) : q);
                const int ck = a == 2 ? (mx ? nz - 1 : 0) : q;
                ow.push_back(cid(ci, cj, ck));
            }
    };
    struct BPatch {
        std::string name, type, nbr;
        std::vector<int> sides;
    };
    std::vector<BPatch> bp;
    std::vector<int> used(6, 0);
    if (d.has("boundary")) {
        const picmi::foam::Entry* e = d.find("boundary");
        if (e->dict) throw FoamError(d.where + ": boundary must be a list");
        const std::vector<Token>& t = e->value;
        size_t pos = 0;
        if (pos >= t.size() || !t[pos].is('(')) throw FoamError(d.where + ": boundary: expected '('");
        ++pos;
        while (pos < t.size() && !t[pos].is(')')) {
            BPatch P;
            P.name = t[pos++].text;
            if (pos >= t.size() || !t[pos].is('{')) throw FoamError(d.where + ": boundary " + P.name + ": expected '{'");
            ++pos;
            auto pd = picmi::foam::parse_entries(t, pos, d.where + "/boundary/" + P.name, true);
            P.type = pd->word("type");
            if (P.type == "cyclic") P.nbr = pd->word("neighbourPatch");
            Reader r = pd->value("faces");
            r.list_open();
            while (!r.accept(')')) {
                r.expect('(');
                std::vector<int> q;
                while (!r.accept(')')) q.push_back((int)r.label());
                if (q.size() != 4) r.fail("a block face has 4 vertices");
                const int sd = side_of(q, pd->where);
                if (used[sd]++) r.fail("block side listed twice");
                P.sides.push_back(sd);
            }
            bp.push_back(P);
        }
    }
    std::vector<int> rest;
    for (int sd = 0; sd < 6; ++sd) if (!used[sd]) rest.push_back(sd);
    if (!rest.empty()) {
        BPatch P;
        P.name = "defaultFaces";
        P.type = "empty";
        if (const Dict* dd = d.sub("defaultPatch")) {
            P.name = dd->word_or("name", P.name);
            P.type = dd->word_or("type", P.type);
        }
        P.sides = rest;
        bp.push_back(P);
    }
    for (const BPatch& P : bp) {
        pic_patch pp{};
        pp.type = patch_type(P.type, path + ": patch " + P.name);
        pp.start = (int32_t)(c.fo.size() - 1);
        pp.neighbour_patch = -1;
        std::vector<std::array<int, 4>> fs;
        std::vector<int> ow;
        for (int sd : P.sides) side_faces(sd, fs, ow);
        for (size_t i = 0; i < fs.size(); ++i) push_face(fs[i], ow[i]);
        pp.size = (int32_t)fs.size();
        c.patches.push_back(pp);
        c.patch_names.push_back(P.name);
    }
    for (size_t i = 0; i < bp.size(); ++i) {
        if (bp[i].type != "cyclic") continue;
        auto it = std::find(c.patch_names.begin(), c.patch_names.end(), bp[i].nbr);
        if (it == c.patch_names.end()) throw FoamError(path + ": neighbourPatch " + bp[i].nbr + " not found");
        const size_t j = (size_t)(it - c.patch_names.begin());
        if (bp[i].sides.size() != 1 || bp[j].sides.size() != 1 || bp[i].sides[0] / 2 != bp[j].sides[0] / 2 ||
            bp[i].sides[0] == bp[j].sides[0])
            throw FoamError(path + ": cyclic " + bp[i].name + " must pair one block side with the opposite side");
        c.patches[i].neighbour_patch = (int32_t)j;
    }
}

// ---------------------------------------------------------------- fields of a time directory

// internalField of a volVectorField: uniform (x y z) or nonuniform List<vector> N (...)
// The boundaryField entry of a patch: its own name, else the first key that matches it as a
// regular expression (OpenFOAM's quoted patch-name patterns)
const Dict* patch_entry(const Dict& bf, const std::string& name) {
    if (const Dict* pd = bf.sub(name)) return pd;
    for (const picmi::foam::Entry& e : bf.entries) {
        if (!e.dict) continue;
        try {
            if (std::regex_match(name, std::regex(e.key))) return e.dict.get();
        } catch (const std::regex_error&) {
        }
    }
    return nullptr;
}

std::vector<double> read_vector_field(const std::string& path, int n_cells) {
    auto d = picmi::foam::read_dict_file(path);
    Reader r = d->value("internalField");
    const std::string kind = r.word();
    std::vector<double> out;
    if (kind == "uniform") {
        auto v = r.vector3();
        out.resize(3 * (size_t)n_cells);
        for (int i = 0; i < n_cells; ++i) { out[3 * i] = v[0]; out[3 * i + 1] = v[1]; out[3 * i + 2] = v[2]; }
    } else if (kind == "nonuniform") {
        r.word();  // List<vector>
        const long long n = r.list_open();
        while (!r.accept(')')) { auto v = r.vector3(); out.insert(out.end(), v.begin(), v.end()); }
        if ((n >= 0 && (long long)out.size() != 3 * n) || (long long)out.size() != 3LL * n_cells)
            r.fail("internalField size differs from the cell count");
    } else {
        r.fail("internalField must be uniform or nonuniform");
    }
    return out;
}

// the phiE patch conditions: fixedValue (uniform value), zeroGradient, cyclic, empty
void read_phi_bc(Case& c, const std::string& path) {
    auto d = picmi::foam::read_dict_file(path);
    const Dict& bf = d->sub_required("boundaryField");
    for (size_t p = 0; p < c.patches.size(); ++p) {
        const Dict* pd = patch_entry(bf, c.patch_names[p]);
        if (!pd) throw FoamError(path + ": boundaryField has no entry for patch " + c.patch_names[p]);
        const std::string t = pd->word("type");
        pic_patch& pp = c.patches[p];
        pp.phi_bc = PIC_BC_ZERO_GRADIENT;
        pp.phi_value = 0.0;
        if (t == "fixedValue") {
            Reader r = pd->value("value");
            if (r.word() != "uniform") r.fail("only a uniform fixedValue is supported");
            pp.phi_bc = PIC_BC_FIXED_VALUE;
            pp.phi_value = r.number();
        } else if (t == "zeroGradient") {
        } else if (t == "cyclic" || t == "empty") {
            const int want = t == "cyclic" ? PIC_PATCH_CYCLIC : PIC_PATCH_EMPTY;
            if (pp.type != want) throw FoamError(pd->where + ": type " + t + " on a patch of another type");
        } else {
            throw FoamError(pd->where + ": phiE condition '" + t + "' is not supported");
        }
    }
}

// The wall temperatures and velocities of 0/boundaryT and 0/boundaryU (PICCloud.C:932-955,
// read at the wall by MaxwellianThermal.C:105-117 and MixedDiffuseSpecular.C:107-119): one
// uniform value per wall patch (the patch's value entry, else a uniform internalField)
void read_wall_fields(Case& c, const std::string& t0) {
    bool any = false;
    for (const pic_patch& p : c.patches) any = any || p.type == PIC_PATCH_WALL;
    c.wall_files.clear();
    if (!any) return;
    for (int f = 0; f < 2; ++f) {
        const std::string name = f == 0 ? "boundaryT" : "boundaryU";
        const std::string path = t0 + "/" + name;
        if (!picmi::foam::file_exists(path)) throw FoamError(path + ": missing (the mesh has wall patches)");
        c.wall_files.push_back({name, picmi::foam::read_text(path)});
        auto d = picmi::foam::read_dict_file(path);
        const Dict& bf = d->sub_required("boundaryField");
        for (size_t p = 0; p < c.patches.size(); ++p) {
            pic_patch& pp = c.patches[p];
            if (pp.type != PIC_PATCH_WALL) continue;
            const Dict* pd = patch_entry(bf, c.patch_names[p]);
            if (!pd) throw FoamError(path + ": boundaryField has no entry for wall patch " + c.patch_names[p]);
            Reader r = pd->has("value") ? pd->value("value") : d->value("internalField");
            const std::string kind = r.word();
            if (kind != "uniform")
                r.fail("per-face values on wall patch " + c.patch_names[p] + " are not supported; give a uniform value");
            if (f == 0) {
                pp.wall_temperature = r.number();
                if (!(pp.wall_temperature >= 0.0)) r.fail("negative wall temperature");
            } else {
                const auto v = r.vector3();
                for (int k = 0; k < 3; ++k) pp.wall_velocity[k] = v[k];
            }
        }
    }
}

// ---------------------------------------------------------------- picProperties

int species_index(const Case& c, const std::string& name, const std::string& where) {
    auto it = std::find(c.species_names.begin(), c.species_names.end(), name);
    if (it == c.species_names.end()) throw FoamError(where + ": species '" + name + "' is not in typeIdList");
    return (int)(it - c.species_names.begin());
}

void require_none(const Dict& d, const std::string& key, const std::string& why) {
    if (d.has(key) && d.word(key) != "none") throw FoamError(d.where + ": " + key + " " + d.word(key) + ": " + why);
}

// a modelList naming a model other than none: "( patch { type <model>; ... } ... )"
// (BoundaryModelList.C, BoundaryEventModelList.C: an entry of type none adds nothing)
bool list_has_models(const picmi::foam::Entry* e) {
    if (!e || e->dict) return false;
    const std::vector<picmi::foam::Token>& t = e->value;
    size_t k = 0;
    if (k < t.size() && t[k].kind == picmi::foam::Token::Number) ++k;
    if (t.size() <= k + 2) return false;   // "()" and "0()"
    int depth = 0, entries = 0;
    for (size_t i = k; i < t.size(); ++i) {
        if (t[i].is('{')) { ++depth; if (depth == 1) ++entries; continue; }
        if (t[i].is('}')) { --depth; continue; }
        if (depth == 1 && t[i].kind == picmi::foam::Token::Word && t[i].text == "type" && i + 1 < t.size() &&
            t[i + 1].text != "none")
            return true;
    }
    return entries == 0;   // elements that are not dictionaries: not understood, refused
}

// PatchBoundaryModels.modelList (BoundaryModelList.C:104-167): "( <patch> { type <model>;
// <model>Coeffs { ... } } ... )".  SimpleInjection (SimpleInjection.C:34-175) becomes a
// pic_injection; an entry of type none adds nothing; other models are refused.
void read_patch_boundary_models(Case& c, const Dict& l) {
    const picmi::foam::Entry* e = l.find("modelList");
    if (!e || e->dict) return;
    Reader r(e->value, l.where + "/modelList");
    r.list_open();
    while (!r.accept(')')) {
        const std::string patch = r.word();
        r.expect('{');
        size_t pos = r.pos();
        auto md = picmi::foam::parse_entries(e->value, pos, l.where + "/modelList/" + patch, true);
        while (r.pos() < pos) r.next();
        const std::string type = md->word("type");
        if (type == "none") continue;
        if (type != "SimpleInjection")
            throw FoamError(md->where + ": patch boundary model " + type + " is not implemented (SimpleInjection is)");
        auto it = std::find(c.patch_names.begin(), c.patch_names.end(), patch);
        if (it == c.patch_names.end()) throw FoamError(md->where + ": no patch " + patch);
        const Dict& cf = md->sub_required("SimpleInjectionCoeffs");
        pic_injection q{};
        q.patch = (int32_t)(it - c.patch_names.begin());
        const std::string vm = cf.word("velocityModel");
        const char* vms[] = {"birdMaxwellianFlux", "mostProbableSpeed", "halfMaxwellianFlux", "maxwellianFlux"};
        q.velocity_model = -1;
        for (int k = 0; k < 4; ++k)
            if (vm == vms[k]) q.velocity_model = k;
        if (q.velocity_model < 0)
            cf.fail("velocityModel", "valid injection velocity models are: birdMaxwellianFlux mostProbableSpeed "
                                     "maxwellianFlux halfMaxwellianFlux");
        const std::vector<std::string> ids = cf.words("injectionTypeIds");
        const std::vector<double> fr = cf.scalars("frequencies"), ev = cf.scalars("eV");
        if (ids.size() != fr.size() || fr.size() != ev.size())
            throw FoamError(cf.where + ": not the same number of entries in frequencies, eV, injectionTypeIds");
        if (ids.empty() || ids.size() > 8) throw FoamError(cf.where + ": 1 to 8 injectionTypeIds");
        q.n_species = (int32_t)ids.size();
        for (size_t k = 0; k < ids.size(); ++k) {
            auto sp = std::find(c.species_names.begin(), c.species_names.end(), ids[k]);
            if (sp == c.species_names.end()) throw FoamError(cf.where + ": species " + ids[k] + " not defined in cloud.");
            q.species[k] = (int32_t)(sp - c.species_names.begin());
            q.frequency[k] = fr[k];
            q.eV[k] = ev[k];
        }
        q.end_time = cf.scalar("injectionEndTime");
        // interactions { <species> { type deletion|reflection|floating; useInitalVelocity b; } }
        if (const Dict* in = cf.sub("interactions")) {
            for (const picmi::foam::Entry& ie : in->entries) {
                if (!ie.dict) continue;
                auto sp = std::find(c.species_names.begin(), c.species_names.end(), ie.key);
                if (sp == c.species_names.end()) throw FoamError(in->where + ": Species " + ie.key + " is not defined in cloud");
                const std::string it2 = ie.dict->word("type");
                int kind = -1;
                if (it2 == "deletion") kind = 0;
                else if (it2 == "reflection") kind = ie.dict->boolean_or("useInitalVelocity", false) ? 2 : 1;
                else if (it2 == "floating") kind = 3;
                else throw FoamError(ie.dict->where + ": interaction type " + it2 + " not defined. valid options [deletion/reflection/floating]");
                q.interaction[sp - c.species_names.begin()] = kind;
            }
        }
        c.injection.push_back(q);
    }
}

void read_pic_properties(Case& c, const std::string& path) {
    auto dp = picmi::foam::read_dict_file(path);
    const Dict& d = *dp;
    const Dict& ss = d.sub_required("SolverSettings");
    c.species_names = ss.words("typeIdList");
    if (c.species_names.empty() || c.species_names.size() > 8)
        throw FoamError(ss.where + ": typeIdList must name 1 to 8 species");
    c.space_charge = ss.scalar_or("spaceCharge", 0.0);
    const Dict& mp = d.sub_required("moleculeProperties");
    c.species.clear();
    for (const std::string& s : c.species_names) {
        const Dict& sd = mp.sub_required(s);
        pic_species sp{};
        sp.mass = sd.scalar("mass");
        sp.charge = sd.scalar("charge");
        sp.n_equivalent = sd.scalar("nEquivalentParticles");
        sp.solve_movement = sd.boolean_or("solveMovement", true) ? 1 : 0;
        // the hard-sphere diameter and viscosity index of the binary collision models
        // (PICParcelI.H:45-46)
        sp.diameter = sd.scalar_or("diameter", 0.0);
        sp.omega = sd.scalar_or("omega", 0.0);
        sp.capacity = 0;
        sp.sort_interval = 0;
        c.species.push_back(sp);
    }
    // model words live in c.model_words; the pointers are set once the vector stops growing
    std::vector<std::pair<const char**, std::string>> names;
    pic_models& m = c.models;
    m = pic_models{};
    m.bg_species = -1;
    names.push_back({&m.particle_pusher, d.word("ParticlePusher")});
    names.push_back({&m.field_weighting, ss.word("FieldWeigthing")});
    names.push_back({&m.charge_distribution, ss.word("ChargeDistribution")});
    names.push_back({&m.maxwell_solver, d.word("MaxwellSolver")});
    require_none(d, "ParticleMerging", "particle merging is not implemented");
    // BoundaryModels (PICCloud.C:493-506): the wall model (WallReflectionModelNew.C:40, its
    // coefficients in <model>Coeffs, WallReflectionModel.C:52); patch boundary and event
    // models are not implemented
    std::string wall = "SpecularReflection";
    m.diffuse_fraction = 0.0;
    c.injection.clear();
    if (const Dict* bm = d.sub("BoundaryModels")) {
        if (const Dict* l = bm->sub("PatchEventModels"))
            if (list_has_models(l->find("modelList")))
                throw FoamError(l->where + ": patch-event models are not implemented");
        if (const Dict* l = bm->sub("PatchBoundaryModels")) read_patch_boundary_models(c, *l);
        wall = bm->word_or("WallReflectionModel", wall);
        if (wall == "MixedDiffuseSpecular")
            m.diffuse_fraction = bm->sub_required("MixedDiffuseSpecularCoeffs").scalar("diffuseFraction");
    }
    names.push_back({&m.wall_reflection, wall});
    // SolverSettings.syncVelocityAtBoundary, true by default (PICCloud.C:926)
    m.no_boundary_sync = ss.boolean_or("syncVelocityAtBoundary", true) ? 0 : 1;
    const Dict& cm = d.sub_required("CollisionModels");
    // BinaryCollisionModel (BinaryCollisionModelNew.C:39-61) with <model>Coeffs: the
    // TotalCrossSectionModel (TotalCrossSectionModelNew.C:39, VariableHardSphereCoeffs.Tref,
    // VariableHardSphere.C:39-45) and the WeightCorrection (WeightCorrectionModelNew.C:42)
    const std::string bin = cm.word_or("BinaryCollisionModel", "none");
    names.push_back({&m.binary_collision, bin});
    std::string bxs = "HardSphere", bwc = "none";
    if (bin != "none") {
        const Dict& bc = cm.sub_required(bin + "Coeffs");
        bxs = bc.word("TotalCrossSectionModel");
        bwc = bc.word("WeightCorrection");
        if (bxs == "VariableHardSphere") m.vhs_tref = bc.sub_required("VariableHardSphereCoeffs").scalar("Tref");
    }
    names.push_back({&m.total_cross_section, bxs});
    names.push_back({&m.binary_weight_correction, bwc});
    // SpeciesRelations (PICCloud.C:96-130)
    m.electron_species = -1;
    m.ion_species = -1;
    int neutral = -1;
    if (const Dict* rel = mp.sub("SpeciesRelations")) {
        const std::string e = rel->word_or("electronTypeId", "none");
        if (e != "none") m.electron_species = species_index(c, e, rel->where);
        for (const picmi::foam::Entry& en : rel->entries) {
            if (!en.dict) continue;
            const std::string ion = en.dict->word_or("ionTypeId", "none");
            const std::string neu = en.dict->word_or("neutralTypeId", "none");
            if (m.ion_species < 0 && ion != "none") m.ion_species = species_index(c, ion, en.dict->where);
            if (neutral < 0 && neu != "none") neutral = species_index(c, neu, en.dict->where);
        }
    }
    // Coulomb collisions (CoulombCollisionModel.C:55-190)
    const std::string coul = cm.word_or("CoulombCollisionModel", "none");
    names.push_back({&m.coulomb_collision, coul});
    m.allow_intra_collision = 1;
    m.calculate_bmin = 1;
    std::string pairing = "TakizukaAbe", wcorr = "none";
    if (coul != "none") {
        const Dict& cc = cm.sub_required("CoulombCollisionCoeffs");
        pairing = cc.word("pairingAlgorithm");
        wcorr = cc.word_or("WeightCorrection", "none");
        m.allow_intra_collision = cc.boolean("allowIntraCollision") ? 1 : 0;
        require_none(cc, "IonizationModel", "Perez ionization is not implemented");
        if (const Dict* dl = cc.sub("DebyeLength")) {
            m.calculate_bmin = dl->boolean_or("calculate_bmin", true) ? 1 : 0;
            m.debye_ignore_drift = dl->boolean_or("accountForDrift", true) ? 0 : 1;
            if (dl->has("species"))
                for (const std::string& s : dl->words("species")) m.debye_species |= 1u << species_index(c, s, dl->where);
        }
        if (const Dict* col = cc.sub("Collisions")) {
            for (const picmi::foam::Entry& en : col->entries) {
                if (!en.dict) continue;
                const std::vector<std::string> pr = en.dict->words("collisionPartner");
                if (pr.size() != 2) throw FoamError(en.dict->where + ": collisionPartner names two species");
                const int a = species_index(c, pr[0], en.dict->where), b = species_index(c, pr[1], en.dict->where);
                const double l = en.dict->scalar("coulombLog");
                if (m.coulomb_log[a][b] != 0.0 && m.coulomb_log[a][b] != l)
                    throw FoamError(en.dict->where + ": ambiguous coulombLog for " + pr[0] + " " + pr[1]);
                m.coulomb_log[a][b] = m.coulomb_log[b][a] = l;
            }
        }
    }
    names.push_back({&m.pairing_algorithm, pairing});
    names.push_back({&m.weight_correction, wcorr});
    // the uniform background gas (uniformBackground.C:40-80), for electron-neutral
    // (ElectronNeutralCollisionModel.C:120-230) and ion-neutral (IonNeutralCollisionModel.C:
    // 93-197) collisions with their cross-section models
    const std::string en = cm.word_or("ElectronNeutralCollisionModel", "none");
    const std::string in = cm.word_or("IonNeutralCollisionModel", "none");
    const std::string bg = cm.word_or("BackgroundGasModel", "none");
    names.push_back({&m.electron_neutral, en});
    names.push_back({&m.ion_neutral_collision, in});
    names.push_back({&m.background_gas, bg});
    m.create_ion = m.create_electron = 1;
    int gas = -1;
    if (bg == "uniformBackground") {
        const Dict& bc = cm.sub_required("BackgroundGasCoeffs");
        m.bg_temperature = bc.scalar("temperature");
        m.bg_number_density = bc.scalar_or("numberDensity", -1.0);
        if (m.bg_number_density < 0.0) {
            const double p = bc.scalar("pressure");
            m.bg_number_density = p / (kBoltzmann * m.bg_temperature);
        }
        m.bg_n_equivalent = bc.scalar("nEquivalentParticles");
        const auto v = bc.vector3("velocity");
        for (int k = 0; k < 3; ++k) m.bg_velocity[k] = v[k];
        gas = species_index(c, bc.word("species"), bc.where);
        m.bg_mass = c.species[gas].mass;
        m.bg_species = gas;
        // the background's ion: the SpeciesRelations entry naming the gas as its neutral
        // (PICCloud::ionTypeId)
        if (const Dict* rel = mp.sub("SpeciesRelations"))
            for (const picmi::foam::Entry& e : rel->entries)
                if (e.dict && e.dict->word_or("neutralTypeId", "none") == c.species_names[gas]) {
                    const std::string ion = e.dict->word_or("ionTypeId", "none");
                    if (ion != "none") m.ion_species = species_index(c, ion, e.dict->where);
                    break;
                }
    } else if (bg != "none") {
        throw FoamError(cm.where + ": BackgroundGasModel " + bg + " is not supported (uniformBackground, none)");
    }
    // a model's cross sections for the background gas: <Kind>CrossSection <model> with
    // <Tag><model>Coeffs (CrossSectionModelList.C:48-88; FixedValueCrossSection.C, value and,
    // but for elastic scattering, threshold; ListCrossSection.C:118-310, energies and values,
    // the ion charge exchange's list under the keyword "value")
    auto read_xs = [&](const Dict& coeffs, const char* what, int first, const char* const* kinds,
                       const char* const* tags, int nproc, int neutral_sp) {
        const Dict& xm = coeffs.sub_required("CrossSectionModels");
        const Dict& xg = xm.sub_required(c.species_names[neutral_sp]);
        for (int p = 0; p < nproc; ++p) {
            const std::string model = xg.word_or(kinds[p], "none");
            if (model == "none") continue;
            Case::CrossSection x;
            x.process = first + p;
            const bool elastic = std::string(tags[p]) == "Elastic";
            const Dict& co = xg.sub_required(std::string(tags[p]) + model + "Coeffs");
            x.threshold = elastic ? 0.0 : co.scalar("threshold");
            if (model == "FixedValue") {
                const double v = co.scalar("value");
                x.e = {0.0, 1e30};
                x.s = {v, v};
            } else if (model == "List") {
                x.e = co.scalars("energies");
                x.s = co.scalars(first + p == PIC_XS_ION_CHARGE_EXCHANGE ? "value" : "values");
                if (x.e.size() != x.s.size() || x.e.size() < 2)
                    throw FoamError(co.where + ": energies and values differ in length");
            } else {
                throw FoamError(xg.where + ": " + what + " cross-section model " + model +
                                " is not supported (List, FixedValue)");
            }
            c.xs.push_back(x);
        }
    };
    if (en != "none") {
        if (gas < 0)
            throw FoamError(cm.where + ": electron-neutral collisions with neutral parcels are not implemented; "
                            "use BackgroundGasModel uniformBackground");
        const Dict& ec = cm.sub_required("ElectronNeutralCollisionCoeffs");
        require_none(ec, "WeightCorrection", "weight correction of electron-neutral collisions is not implemented");
        m.create_ion = ec.boolean_or("createIon", true) ? 1 : 0;
        m.create_electron = ec.boolean_or("createElectron", true) ? 1 : 0;
        const char* kinds[3] = {"ElasticCrossSection", "ExcitationCrossSection", "IonizationCrossSection"};
        const char* tags[3] = {"Elastic", "Excitation", "Ionization"};
        read_xs(ec, "electron-neutral", PIC_XS_ELASTIC, kinds, tags, 3, gas);
    }
    std::string iwc = "none";
    if (in != "none") {
        // one neutral species (IonNeutralCollisionModel.C:170-175): the background's, else the
        // species without charge; its cross sections serve the parcels and the background
        int nsp = gas;
        if (nsp < 0)
            for (size_t s = 0; s < c.species.size(); ++s)
                if (c.species[s].charge == 0.0) { nsp = (int)s; break; }
        if (nsp < 0) throw FoamError(cm.where + ": IonNeutralCollisionModel needs a neutral species or a background gas");
        const Dict& ic = cm.sub_required("IonNeutralCollisionCoeffs");
        iwc = ic.word_or("WeightCorrection", "none");
        m.ion_neutral_keep_neutral = ic.boolean_or("updateNeutralVelocity", true) ? 0 : 1;
        const char* kinds[2] = {"ElasticCrossSection", "ChargeExchangeCrossSection"};
        const char* tags[2] = {"Elastic", "ChargeExchange"};
        read_xs(ic, "ion-neutral", PIC_XS_ION_ELASTIC, kinds, tags, 2, nsp);
    }
    names.push_back({&m.ion_neutral_weight_correction, iwc});
    (void)neutral;
    m.precision = PIC_PREC_FP64;
    m.seed = 12345;
    m.poisson_tolerance = 1e-8;
    m.poisson_max_iter = 20000;
    m.sort_interval = 1;
    // the words are stored first, then the pointers taken (the vector does not move after)
    c.model_words.clear();
    c.model_words.reserve(names.size());
    for (auto& nw : names) c.model_words.push_back(nw.second);
    for (size_t i = 0; i < names.size(); ++i) *names[i].first = c.model_words[i].c_str();
}

// PCG settings of system/fvSolution solvers.phiE (tolerance is absolute in OpenFOAM; here the
// relative residual of the same solve is used, relTol must be 0)
void read_fv_solution(Case& c, const std::string& path) {
    if (!picmi::foam::file_exists(path)) return;
    auto d = picmi::foam::read_dict_file(path);
    const Dict* s = d->sub("solvers");
    const Dict* p = s ? s->sub("phiE") : nullptr;
    if (!p) return;
    c.models.poisson_tolerance = p->scalar_or("tolerance", c.models.poisson_tolerance);
    c.models.poisson_max_iter = (int32_t)p->scalar_or("maxIter", c.models.poisson_max_iter);
}

void read_control(Case& c, const std::string& path) {
    auto d = picmi::foam::read_dict_file(path);
    const std::string from = d->word_or("startFrom", "startTime");
    c.start = d->scalar_or("startTime", 0.0);
    if (from == "latestTime" || from == "firstTime") {
        // the numeric time directories of the case
        std::vector<double> times;
        if (DIR* dh = opendir(c.dir.c_str())) {
            while (dirent* e = readdir(dh)) {
                char* end = nullptr;
                const double t = std::strtod(e->d_name, &end);
                if (end != e->d_name && *end == '\0') times.push_back(t);
            }
            closedir(dh);
        }
        if (!times.empty())
            c.start = from == "latestTime" ? *std::max_element(times.begin(), times.end())
                                           : *std::min_element(times.begin(), times.end());
    }
    c.end = d->scalar("endTime");
    c.dt = d->scalar("deltaT");
    if (!(c.dt > 0.0)) throw FoamError(path + ": deltaT must be positive");
    const std::string wc = d->word_or("writeControl", "timeStep");
    if (wc == "timeStep") c.write_steps = true;
    else if (wc == "runTime" || wc == "adjustableRunTime") c.write_steps = false;
    else throw FoamError(path + ": writeControl " + wc + " is not supported (timeStep, runTime)");
    c.write_interval = d->scalar("writeInterval");
    c.now = c.start;
    c.now_name = time_name(c.start);
}

// the time directory of the start time, as OpenFOAM names it or as written in the case
std::string start_dir(const Case& c) {
    const std::string a = c.dir + "/" + time_name(c.start);
    struct stat st;
    if (stat(a.c_str(), &st) == 0 && S_ISDIR(st.st_mode)) return a;
    if (DIR* dh = opendir(c.dir.c_str())) {
        while (dirent* e = readdir(dh)) {
            char* end = nullptr;
            const double t = std::strtod(e->d_name, &end);
            if (end != e->d_name && *end == '\0' && t == c.start) {
                std::string r = c.dir + "/" + e->d_name;
                closedir(dh);
                return r;
            }
        }
        closedir(dh);
    }
    return a;
}

void open_case(Case& c) {
    read_control(c, c.dir + "/system/controlDict");
    const std::string pm = c.dir + "/constant/polyMesh";
    if (picmi::foam::file_exists(pm + "/faces")) read_polymesh(c, pm);
    else if (picmi::foam::file_exists(c.dir + "/system/blockMeshDict")) block_mesh(c, c.dir + "/system/blockMeshDict");
    else throw FoamError(c.dir + ": neither constant/polyMesh nor system/blockMeshDict");
    read_pic_properties(c, c.dir + "/constant/picProperties");
    read_fv_solution(c, c.dir + "/system/fvSolution");
    const std::string t0 = start_dir(c);
    for (pic_patch& p : c.patches) { p.phi_bc = PIC_BC_ZERO_GRADIENT; p.phi_value = 0.0; }
    if (std::string(c.models.maxwell_solver) != "none") {
        if (!picmi::foam::file_exists(t0 + "/phiE")) throw FoamError(t0 + "/phiE: missing (MaxwellSolver needs it)");
        read_phi_bc(c, t0 + "/phiE");
    }
    read_wall_fields(c, t0);
    if (picmi::foam::file_exists(t0 + "/E")) c.E = read_vector_field(t0 + "/E", c.n_cells);
    // a (sigma cR)_max field of the start time, if present and positive
    auto read_smax = [&](const std::string& f, std::vector<double>& out) {
        out.clear();
        if (!picmi::foam::file_exists(f)) return;
        auto d = picmi::foam::read_dict_file(f);
        Reader r = d->value("internalField");
        const std::string kind = r.word();
        if (kind == "uniform") {
            const double v = r.number();
            if (v > 0.0) out.assign(c.n_cells, v);
        } else {
            r.word();
            r.list_open();
            while (!r.accept(')')) out.push_back(r.number());
            if ((int)out.size() != c.n_cells) r.fail("internalField size differs from the cell count");
        }
    };
    c.smax_bin.clear();
    c.smax_ip.clear();
    if (std::string(c.models.binary_collision) != "none") read_smax(t0 + "/picSigmaTcRMax", c.smax_bin);
    if (std::string(c.models.ion_neutral_collision) != "none") read_smax(t0 + "/picSigmaTcRMax_IonNeutral", c.smax_ip);
    for (int k = 0; k < 2; ++k) {
        const bool on = std::string(k == 0 ? c.models.electron_neutral : c.models.ion_neutral_collision) != "none";
        c.smax[k].clear();
        if (!on) continue;
        read_smax(t0 + (k == 0 ? "/picSigmaTcRMax_ElectronBackground" : "/picSigmaTcRMax_IonBackground"), c.smax[k]);
        double emax = 0.0;
        for (const Case::CrossSection& x : c.xs)
            if ((x.process >= PIC_XS_ION_ELASTIC) == (k == 1) && x.e.back() < 1e29) emax = std::max(emax, x.e.back());
        c.xs_init[k] = Case::XsInit{};
        c.xs_init[k].emax = emax > 0.0 ? emax : 1000.0;
        c.xs_init[k].estep = c.xs_init[k].emax * 1e-5;
    }
    if (picmi::foam::file_exists(t0 + "/B")) c.B = read_vector_field(t0 + "/B", c.n_cells);
    // host geometry and tets (the same tables pic_create builds)
    picmi::HostMesh& m = c.hm;
    m = picmi::HostMesh{};
    m.n_points = (int)(c.points.size() / 3);
    m.n_faces = (int)c.fo.size() - 1;
    m.n_internal = (int)c.nbr.size();
    m.n_cells = c.n_cells;
    m.n_patches = (int)c.patches.size();
    m.points = c.points;
    m.fo = c.fo; m.fp = c.fp; m.owner = c.owner; m.nbr = c.nbr;
    for (const pic_patch& p : c.patches) {
        m.ptype.push_back(p.type); m.pstart.push_back(p.start); m.psize.push_back(p.size);
        m.pnbr.push_back(p.neighbour_patch); m.phi_bc.push_back(p.phi_bc); m.phi_val.push_back(p.phi_value);
    }
    const std::string err = picmi::build_host_mesh(m);
    if (!err.empty()) throw FoamError(c.dir + ": mesh: " + err);
    for (int d = 0; d < 3; ++d) {
        double lo = 1e300, hi = -1e300;
        for (size_t i = d; i < c.points.size(); i += 3) { lo = std::min(lo, c.points[i]); hi = std::max(hi, c.points[i]); }
        c.mid[d] = 0.5 * (lo + hi);
    }
    c.parcels.assign(c.species.size(), HostParcels{});
}

// ---------------------------------------------------------------- the picInitialise models

// the four vertices (cell centre, base, A, B) of tet t
void tet_vertices(const picmi::HostMesh& m, int t, V3 X[4]) {
    int vid[3];
    picmi::tet_points(m, t, vid);
    X[0] = at(m.C, (size_t)m.tet_cell[t]);
    for (int i = 0; i < 3; ++i) X[i + 1] = at(m.points, (size_t)vid[i]);
}

double tet_volume(const V3 X[4]) { return dot(X[1] - X[0], cross(X[2] - X[0], X[3] - X[0])) / 6.0; }

// a uniform point of the unit tetrahedron as barycentric coordinates: three uniforms in the
// unit cube folded into the tet (Rocchini and Cignoni, "Generating random points in a
// tetrahedron", J. Graphics Tools 5(4), 2000), the construction tetrahedron::randomPoint uses
void random_bary(Stream& r, double b[4]) {
    double s = r.uniform(), t = r.uniform(), u = r.uniform();
    if (s + t > 1.0) { s = 1.0 - s; t = 1.0 - t; }
    if (t + u > 1.0) {
        const double tmp = u;
        u = 1.0 - s - t;
        t = 1.0 - tmp;
    } else if (s + t + u > 1.0) {
        const double tmp = u;
        u = s + t + u - 1.0;
        s = 1.0 - t - tmp;
    }
    b[0] = 1.0 - s - t - u; b[1] = s; b[2] = t; b[3] = u;
}

void add_in_tet(HostParcels& P, int t, int cell, const double b[4], V3 U) {
    P.tet.push_back(t);
    P.cell.push_back(cell);
    P.bary.push_back(b[1]); P.bary.push_back(b[2]); P.bary.push_back(b[3]);
    P.U.push_back(U.x); P.U.push_back(U.y); P.U.push_back(U.z);
}

// meshTools::constrainDirection on a position: the empty directions take the mesh centre
V3 constrain(const picmi::HostMesh& m, V3 p, V3 centre) {
    if (m.track_mask & 1u) p.x = centre.x;
    if (m.track_mask & 2u) p.y = centre.y;
    if (m.track_mask & 4u) p.z = centre.z;
    return p;
}

// the tet of cell containing x and its barycentric (y1, y2, y3), or -1
int locate_in_cell(const picmi::HostMesh& m, int cell, V3 x, double y[3]) {
    for (int t = m.cell_tet0[cell]; t < m.cell_tet0[cell + 1]; ++t) {
        V3 X[4];
        tet_vertices(m, t, X);
        const V3 e1 = X[1] - X[0], e2 = X[2] - X[0], e3 = X[3] - X[0], d = x - X[0];
        const double det = dot(e1, cross(e2, e3));
        if (!(det > 0.0)) continue;
        const double y1 = dot(d, cross(e2, e3)) / det, y2 = dot(e1, cross(d, e3)) / det, y3 = dot(e1, cross(e2, d)) / det;
        const double tol = -1e-12;
        if (1.0 - y1 - y2 - y3 >= tol && y1 >= tol && y2 >= tol && y3 >= tol) {
            y[0] = y1; y[1] = y2; y[2] = y3;
            return t;
        }
    }
    return -1;
}

// a parcel at barycentric b of tet t; in a reduced-dimension case its position is first
// moved to the mesh's mid-plane (meshTools::constrainDirection, as the tet-based
// picInitialise models do) and located again in its cell
void add_constrained(Case& c, HostParcels& P, int t, int cell, const double b[4], V3 U) {
    const picmi::HostMesh& m = c.hm;
    if (!m.track_mask) { add_in_tet(P, t, cell, b, U); return; }
    V3 X[4];
    tet_vertices(m, t, X);
    V3 p = b[0] * X[0] + b[1] * X[1] + b[2] * X[2] + b[3] * X[3];
    p = constrain(m, p, V3{c.mid[0], c.mid[1], c.mid[2]});
    double y[3];
    const int t2 = locate_in_cell(m, cell, p, y);
    if (t2 < 0) throw FoamError("a parcel moved to the mid-plane left its cell " + std::to_string(cell));
    const double bb[4] = {1.0 - y[0] - y[1] - y[2], y[0], y[1], y[2]};
    add_in_tet(P, t2, cell, bb, U);
}

// the cell containing x: a uniform grid over the mesh's bounding box lists the cells whose
// bounding boxes overlap each bin; a cell contains x when one of its tets does (barycentric
// coordinates >= -1e-12); the lowest such cell id is taken (polyMesh::findCell)
class CellFinder {
public:
    explicit CellFinder(const picmi::HostMesh& m) : m_(m) {
        const int nc = m.n_cells;
        lo_ = {1e300, 1e300, 1e300}; hi_ = {-1e300, -1e300, -1e300};
        std::vector<V3> clo(nc, V3{1e300, 1e300, 1e300}), chi(nc, V3{-1e300, -1e300, -1e300});
        for (int t = 0; t < (int)m.tet_cell.size(); ++t) {
            V3 X[4];
            tet_vertices(m, t, X);
            const int c = m.tet_cell[t];
            for (int i = 0; i < 4; ++i) {
                clo[c] = {std::min(clo[c].x, X[i].x), std::min(clo[c].y, X[i].y), std::min(clo[c].z, X[i].z)};
                chi[c] = {std::max(chi[c].x, X[i].x), std::max(chi[c].y, X[i].y), std::max(chi[c].z, X[i].z)};
            }
        }
        for (int c = 0; c < nc; ++c) {
            lo_ = {std::min(lo_.x, clo[c].x), std::min(lo_.y, clo[c].y), std::min(lo_.z, clo[c].z)};
            hi_ = {std::max(hi_.x, chi[c].x), std::max(hi_.y, chi[c].y), std::max(hi_.z, chi[c].z)};
        }
        n_ = std::max(1, (int)std::cbrt((double)nc));
        bins_.assign((size_t)n_ * n_ * n_, {});
        for (int c = 0; c < nc; ++c) {
            int a0[3], a1[3];
            bin(clo[c], a0);
            bin(chi[c], a1);
            for (int k = a0[2]; k <= a1[2]; ++k)
                for (int j = a0[1]; j <= a1[1]; ++j)
                    for (int i = a0[0]; i <= a1[0]; ++i) bins_[((size_t)k * n_ + j) * n_ + i].push_back(c);
        }
    }
    // returns the cell, and the tet and barycentric (y1, y2, y3) of x in it
    int find(V3 x, int* tet, double y[3]) const {
        int b[3];
        bin(x, b);
        for (int c : bins_[((size_t)b[2] * n_ + b[1]) * n_ + b[0]]) {
            for (int t = m_.cell_tet0[c]; t < m_.cell_tet0[c + 1]; ++t) {
                V3 X[4];
                tet_vertices(m_, t, X);
                const V3 e1 = X[1] - X[0], e2 = X[2] - X[0], e3 = X[3] - X[0], d = x - X[0];
                const double det = dot(e1, cross(e2, e3));
                if (!(det > 0.0)) continue;
                const double y1 = dot(d, cross(e2, e3)) / det, y2 = dot(e1, cross(d, e3)) / det,
                             y3 = dot(e1, cross(e2, d)) / det;
                const double y0 = 1.0 - y1 - y2 - y3, tol = -1e-12;
                if (y0 >= tol && y1 >= tol && y2 >= tol && y3 >= tol) {
                    *tet = t;
                    y[0] = y1; y[1] = y2; y[2] = y3;
                    return c;
                }
            }
        }
        return -1;
    }
    V3 lo() const { return lo_; }
    V3 hi() const { return hi_; }

private:
    void bin(V3 p, int b[3]) const {
        const double q[3] = {p.x, p.y, p.z}, l[3] = {lo_.x, lo_.y, lo_.z}, h[3] = {hi_.x, hi_.y, hi_.z};
        for (int d = 0; d < 3; ++d) {
            const double s = h[d] > l[d] ? (q[d] - l[d]) / (h[d] - l[d]) : 0.0;
            b[d] = std::min(n_ - 1, std::max(0, (int)std::floor(s * n_)));
        }
    }
    const picmi::HostMesh& m_;
    V3 lo_, hi_;
    int n_ = 1;
    std::vector<std::vector<int>> bins_;
};

// InitializationModel::readDefaultSettings: temperatures [K] and number densities [m^-3]
// per species (zero when not given), electronVolt converting the temperatures
void read_defaults(const Case& c, const Dict& d, std::vector<double>& T, std::vector<double>& n) {
    T.assign(c.species.size(), 0.0);
    n.assign(c.species.size(), 0.0);
    const bool ev = d.boolean_or("electronVolt", false);
    if (const Dict* td = d.sub("temperature"))
        for (const picmi::foam::Entry& e : td->entries) {
            const int s = species_index(c, e.key, td->where);
            T[s] = td->scalar(e.key) * (ev ? kElementary / kBoltzmann : 1.0);
        }
    if (const Dict* nd = d.sub("numberDensities"))
        for (const picmi::foam::Entry& e : nd->entries) n[species_index(c, e.key, nd->where)] = nd->scalar(e.key);
}

inline V3 maxwell(Stream& r, double T, double mass) {
    const double s = std::sqrt(kBoltzmann * T / mass);
    const double a = r.normal(), b = r.normal(), cc = r.normal();
    return {s * a, s * b, s * cc};
}

constexpr uint32_t kTagCount = 0x11, kTagParcel = 0x12;

// UniformInitialization.C:55-126: particleNumbers on a lattice from min to max with spacing
// (max - min) / max(n - 1, 2) per axis, each species displaced along displacementDir by
// displacement[s] sin(distance / spanLength * cycles * pi)
void init_uniform(Case& c, const Dict& d, const Dict& co, uint64_t seed, const std::vector<double>& T) {
    const picmi::HostMesh& m = c.hm;
    CellFinder cf(m);
    const std::vector<std::string> sp = co.words("species");
    const V3 mins = co.has("min") ? arr(co.vector3("min")) : cf.lo();
    const V3 maxs = co.has("max") ? arr(co.vector3("max")) : cf.hi();
    const auto pn = co.vector3("particleNumbers");
    V3 dir = arr(co.vector3("displacementDir"));
    const std::vector<double> disp = co.scalars("displacement");
    const double cycles = co.scalar("cycles");
    const double dm = std::sqrt(dot(dir, dir));
    if (!(dm > 0.0)) co.fail("displacementDir", "must not be zero");
    dir = (1.0 / dm) * dir;
    if (disp.size() < sp.size()) co.fail("displacement", "needs one value per species");
    const V3 span = maxs - mins;
    const double spanLength = dot(span, dir);
    const long long n[3] = {(long long)pn[0], (long long)pn[1], (long long)pn[2]};
    const V3 delta = {span.x / (double)std::max(n[0] - 1, 2LL), span.y / (double)std::max(n[1] - 1, 2LL),
                      span.z / (double)std::max(n[2] - 1, 2LL)};
    (void)d;
    for (size_t si = 0; si < sp.size(); ++si) {
        const int s = species_index(c, sp[si], co.where);
        HostParcels& P = c.parcels[s];
        uint64_t idx = 0;
        for (long long i = 0; i < n[0]; ++i)
            for (long long j = 0; j < n[1]; ++j)
                for (long long k = 0; k < n[2]; ++k, ++idx) {
                    const V3 p = mins + V3{delta.x * i, delta.y * j, delta.z * k};
                    const double dist = dot(p - mins, dir);
                    const double cur = disp[si] * std::sin(dist / spanLength * cycles * kPiC);
                    const V3 q = p + cur * dir;
                    Stream r(seed, ((uint64_t)s << 40) | idx, kTagParcel);
                    const V3 U = maxwell(r, T[s], c.species[s].mass);
                    int t;
                    double y[3];
                    const int cell = cf.find(q, &t, y);
                    if (cell < 0) {
                        char b[160];
                        std::snprintf(b, sizeof b, "UniformDisplacement: parcel outside the mesh at (%g %g %g)", q.x, q.y, q.z);
                        throw FoamError(co.where + ": " + b);
                    }
                    const double bb[4] = {1.0 - y[0] - y[1] - y[2], y[0], y[1], y[2]};
                    add_in_tet(P, t, cell, bb, U);
                }
    }
}

// QuasineutralPlasmaInitialization.C:83-174: cellCount / nParticle parcels of the ion
// species and of the electrons in every cell, each in the tet label((nTets - 1) * u) of its
// cell (the reference's draw, which never picks the cell's last tet), at a uniform point of
// that tet, Maxwellian (useMaxwellian) or with the speed sqrt(3kT/m) in a direction drawn from
// the unit cube's corner octant (rndGen.sample01<vector>() normalised)
void init_quasineutral(Case& c, const Dict& co, uint64_t seed, const std::vector<double>& T) {
    const picmi::HostMesh& m = c.hm;
    const int ion = species_index(c, co.word("ionSpecies"), co.where);
    const int el = c.models.electron_species;
    if (el < 0) co.fail("ionSpecies", "the electron species is not defined (SpeciesRelations.electronTypeId)");
    if (c.species[ion].charge + c.species[el].charge != 0.0)
        co.fail("ionSpecies", "ion and electron charges must be of the same magnitude and opposite sign");
    const double cc = co.scalar("cellCount");
    const bool maxw = co.boolean("useMaxwellian");
    for (int s : {ion, el}) {
        const long long per = (long long)(cc / c.species[s].n_equivalent);
        HostParcels& P = c.parcels[s];
        for (int cell = 0; cell < m.n_cells; ++cell) {
            const int t0 = m.cell_tet0[cell], nt = m.cell_tet0[cell + 1] - t0;
            for (long long i = 0; i < per; ++i) {
                Stream r(seed, ((uint64_t)s << 56) | ((uint64_t)cell << 24) | (uint64_t)i, kTagParcel);
                const int t = t0 + (int)((double)(nt - 1) * r.uniform());
                double b[4];
                random_bary(r, b);
                V3 U;
                if (maxw) {
                    U = maxwell(r, T[s], c.species[s].mass);
                } else {
                    V3 dvec = {r.uniform(), r.uniform(), r.uniform()};
                    const double l = std::sqrt(dot(dvec, dvec));
                    U = (std::sqrt(T[s] * kBoltzmann * 3.0 / c.species[s].mass) / l) * dvec;
                }
                add_constrained(c, P, t, cell, b, U);
            }
        }
    }
}

// MaxwellianInitialization.C:66-126 and EquipartitionInitialization.C:66-148: per tet and
// species, label(n V_tet / w) parcels plus one with the probability of the remainder, at
// uniform points of the tet; Maxwellian: U = thermalVelocity * (normal draws) + driftVelocity;
// Equipartition: U = sqrt(kT/m) (normal draws) + velocity
void init_per_tet(Case& c, const Dict& co, uint64_t seed, const std::vector<double>& T, const std::vector<double>& n,
                  bool maxwellian) {
    const picmi::HostMesh& m = c.hm;
    const size_t ns = c.species.size();
    std::vector<V3> drift(ns, V3{0, 0, 0}), therm(ns, V3{0, 0, 0});
    V3 vel{0, 0, 0};
    if (maxwellian) {
        for (const picmi::foam::Entry& e : co.entries) {
            if (!e.dict) continue;
            const int s = species_index(c, e.key, co.where);
            drift[s] = arr(e.dict->vector3("driftVelocity"));
            therm[s] = arr(e.dict->vector3("thermalVelocity"));
        }
    } else {
        vel = arr(co.vector3("velocity"));
        if (co.word_or("cellZone", "none") != "none") co.fail("cellZone", "cell zones are not supported");
    }
    for (int t = 0; t < (int)m.tet_cell.size(); ++t) {
        V3 X[4];
        tet_vertices(m, t, X);
        const double vol = tet_volume(X);
        const int cell = m.tet_cell[t];
        for (size_t s = 0; s < ns; ++s) {
            const double req = n[s] * vol / c.species[s].n_equivalent;
            long long k = (long long)req;
            Stream rc(seed, ((uint64_t)s << 40) | (uint64_t)t, kTagCount);
            if (req - (double)k > rc.uniform()) ++k;
            for (long long i = 0; i < k; ++i) {
                Stream r(seed, ((uint64_t)s << 56) | ((uint64_t)t << 20) | (uint64_t)i, kTagParcel);
                double b[4];
                random_bary(r, b);
                V3 U;
                if (maxwellian) {
                    const double a = r.normal(), bb = r.normal(), cz = r.normal();
                    U = V3{therm[s].x * a, therm[s].y * bb, therm[s].z * cz} + drift[s];
                } else {
                    U = maxwell(r, T[s], c.species[s].mass) + vel;
                }
                add_constrained(c, c.parcels[s], t, cell, b, U);
            }
        }
    }
}

// ParticleListInitialization.C:55-108: one parcel per sub-dictionary, at position (or the
// centre of cellId), velocity plus a Maxwellian draw at its temperature
void init_list(Case& c, const Dict& d, const Dict& co, uint64_t seed) {
    const picmi::HostMesh& m = c.hm;
    CellFinder cf(m);
    const bool ev = d.boolean_or("electronVolt", false);
    uint64_t k = 0;
    for (const picmi::foam::Entry& e : co.entries) {
        if (!e.dict) continue;
        const Dict& pd = *e.dict;
        const int s = species_index(c, pd.word("species"), pd.where);
        const V3 vel = arr(pd.vector3("velocity"));
        const long long cid = pd.has("cellId") ? pd.label("cellId") : -1;
        V3 pos;
        if (cid > -1) {
            if (cid >= m.n_cells) pd.fail("cellId", "out of range");
            pos = at(m.C, (size_t)cid);
        } else {
            pos = arr(pd.vector3("position"));
        }
        int t;
        double y[3];
        const int cell = cf.find(pos, &t, y);
        if (cell < 0) pd.fail("position", "outside of the mesh");
        double T = pd.scalar("temperature");
        if (ev) T *= kElementary / kBoltzmann;
        Stream r(seed, ((uint64_t)s << 56) | k++, kTagParcel);
        const V3 U = maxwell(r, T, c.species[s].mass) + vel;
        const double b[4] = {1.0 - y[0] - y[1] - y[2], y[0], y[1], y[2]};
        add_in_tet(c.parcels[s], t, cell, b, U);
    }
}

void initialise(Case& c, const std::string& name, uint64_t seed) {
    auto dp = picmi::foam::read_dict_file(c.dir + "/system/" + name);
    const Dict& d = *dp;
    const std::string model = d.word("InitializationModel");
    std::vector<double> T, n;
    read_defaults(c, d, T, n);
    c.parcels.assign(c.species.size(), HostParcels{});
    if (model == "NoParticle") {
    } else if (model == "UniformDisplacement") {
        init_uniform(c, d, d.sub_required("UniformDisplacementCoeffs"), seed, T);
    } else if (model == "QuasineutralPlasma") {
        init_quasineutral(c, d.sub_required("QuasineutralPlasmaCoeffs"), seed, T);
    } else if (model == "Maxwellian") {
        init_per_tet(c, d.sub_required("MaxwellianCoeffs"), seed, T, n, true);
    } else if (model == "Equipartition") {
        init_per_tet(c, d.sub_required("EquipartitionCoeffs"), seed, T, n, false);
    } else if (model == "ParticleList") {
        init_list(c, d, d.sub_required("ParticleListCoeffs"), seed);
    } else {
        throw FoamError(d.where + ": InitializationModel " + model +
                        " is not supported (UniformDisplacement, QuasineutralPlasma, Maxwellian, Equipartition, "
                        "ParticleList, NoParticle)");
    }
    // the collision models' (sigma cR)_max (picInitialise.C:204-213; the models' initialize,
    // ElectronNeutralCollisionModel.C:370-482, IonNeutralCollisionModel.C:256-379) at the
    // species temperatures: the dictionary's, replaced by the parcels' own for the models that
    // compute them (Maxwellian, ParticleList, NoParticle; picInitialise.C:129-170)
    if (d.boolean_or("initalizeCollisionModels", false)) {
        if (model == "Maxwellian" || model == "ParticleList" || model == "NoParticle")
            for (size_t s = 0; s < c.species.size(); ++s) {
                const HostParcels& P = c.parcels[s];
                const size_t np = P.size();
                if (np == 0) continue;
                double m1[3] = {0, 0, 0}, m2 = 0.0;
                for (size_t i = 0; i < np; ++i) {
                    for (int k = 0; k < 3; ++k) m1[k] += P.U[3 * i + k];
                    m2 += P.U[3 * i] * P.U[3 * i] + P.U[3 * i + 1] * P.U[3 * i + 1] + P.U[3 * i + 2] * P.U[3 * i + 2];
                }
                for (int k = 0; k < 3; ++k) m1[k] /= (double)np;
                T[s] = c.species[s].mass / (3.0 * kBoltzmann) *
                       (m2 / (double)np - (m1[0] * m1[0] + m1[1] * m1[1] + m1[2] * m1[2]));
            }
        // number densities from the parcels for the models that compute them
        // (UniformDisplacement, QuasineutralPlasma, NoParticle, ParticleList; picInitialise.C:
        // 172-200): the species' real particles over the mesh volume
        if (model == "UniformDisplacement" || model == "QuasineutralPlasma" || model == "NoParticle" ||
            model == "ParticleList") {
            double vol = 0.0;
            for (int k = 0; k < c.hm.n_cells; ++k) vol += c.hm.V[k];
            for (size_t s = 0; s < c.species.size(); ++s)
                n[s] = (double)c.parcels[s].size() * c.species[s].n_equivalent / vol;
        }
        // BinaryCollisionModel::initialize (BinaryCollisionModel.C:540-600)
        c.binary_init = std::string(c.models.binary_collision) != "none";
        c.init_T = T;
        c.init_n = n;
        const char* key[2] = {"electronNeutral", "ionNeutral"};
        const int sp[2] = {c.models.electron_species, c.models.ion_species};
        for (int k = 0; k < 2; ++k) {
            // an inactive model returns before it reads the dictionary
            const bool on = std::string(k == 0 ? c.models.electron_neutral : c.models.ion_neutral_collision) != "none";
            if (!on) continue;
            const Dict& xi = d.sub_required("CrossSectionInitialization");
            const std::string mode = xi.word(key[k]);
            Case::XsInit& xs = c.xs_init[k];
            if (mode == "FindMaximum") {
                const Dict& sd = xi.sub_required(std::string(key[k]) + "_searchParameter");
                xs.temperature = false;
                xs.emax = sd.scalar("maximumEnergy");
                xs.estep = sd.scalar("energyStep");
                if (!(xs.estep > 0.0)) sd.fail("energyStep", "must be positive");
            } else if (mode == "Temperature") {
                xs.temperature = true;
                xs.T = sp[k] >= 0 ? T[sp[k]] : 0.0;
            } else {
                xi.fail(key[k], "Invalid initialization mode, valid options are: (FindMaximum, Temperature)");
            }
            c.smax[k].clear();   // initialised here rather than read from the start time
        }
    }
    c.post.active = true;
    c.post.solve_fields = d.boolean_or("solveMaxwellEquations", false);
    c.post.leap_frog = d.boolean_or("initalizeLeapFrog", false);
}

// ---------------------------------------------------------------- lagrangian files

std::string cloud_dir(const Case& c, const std::string& tn) { return c.dir + "/" + tn + "/lagrangian/" + c.cloud; }

void mkdirs(const std::string& path) {
    std::string cur;
    for (size_t i = 0; i <= path.size(); ++i) {
        if (i == path.size() || path[i] == '/') {
            if (!cur.empty()) mkdir(cur.c_str(), 0755);
        }
        if (i < path.size()) cur.push_back(path[i]);
    }
    struct stat st;
    if (stat(path.c_str(), &st) != 0 || !S_ISDIR(st.st_mode)) throw FoamError(path + ": cannot create the directory");
}

void write_file(const std::string& path, const std::string& text) {
    std::ofstream f(path, std::ios::binary);
    if (!f) throw FoamError(path + ": cannot write");
    f << text;
    if (!f) throw FoamError(path + ": write failed");
}

void read_parcels(Case& c, const std::string& tn) {
    const std::string cd = cloud_dir(c, tn);
    const bool coords = picmi::foam::file_exists(cd + "/coordinates");
    if (!coords && !picmi::foam::file_exists(cd + "/positions")) throw FoamError(cd + ": no positions or coordinates");
    std::vector<double> bary, x;
    std::vector<int32_t> cell, tetf, tetp;
    {
        picmi::foam::ListFile lf = picmi::foam::read_list_file(cd + (coords ? "/coordinates" : "/positions"));
        Reader r(lf.tokens, lf.where, lf.body);
        const long long n = r.list_open();
        while (!r.accept(')')) {
            r.expect('(');
            if (coords) {
                double b[4];
                for (double& v : b) v = r.number();
                r.expect(')');
                bary.insert(bary.end(), {b[1], b[2], b[3]});
                cell.push_back((int32_t)r.label());
                tetf.push_back((int32_t)r.label());
                tetp.push_back((int32_t)r.label());
            } else {
                for (int k = 0; k < 3; ++k) x.push_back(r.number());
                r.expect(')');
                const int ln = r.peek().line;
                cell.push_back((int32_t)r.label());
                // an older positions line may carry tetFace and tetPt too; the parcel is
                // located again from its position and cell
                while (!r.done() && r.peek().kind == Token::Number && r.peek().integral && r.peek().line == ln) r.label();
            }
        }
        if (n >= 0 && (long long)cell.size() != n) r.fail("list size differs from its count");
    }
    const size_t np = cell.size();
    auto read_vec = [&](const std::string& f) {
        picmi::foam::ListFile lf = picmi::foam::read_list_file(cd + "/" + f);
        Reader r(lf.tokens, lf.where, lf.body);
        r.list_open();
        std::vector<double> v;
        while (!r.accept(')')) { auto a = r.vector3(); v.insert(v.end(), a.begin(), a.end()); }
        if (v.size() != 3 * np) r.fail("field size differs from the parcel count");
        return v;
    };
    auto read_num = [&](const std::string& f, bool required, double def) {
        std::vector<double> v;
        if (!picmi::foam::file_exists(cd + "/" + f)) {
            if (required) throw FoamError(cd + "/" + f + ": missing");
            v.assign(np, def);
            return v;
        }
        picmi::foam::ListFile lf = picmi::foam::read_list_file(cd + "/" + f);
        Reader r(lf.tokens, lf.where, lf.body);
        // a uniform list N{value}
        if (r.peek().kind == Token::Number && lf.tokens.size() > r.pos() + 1 && lf.tokens[r.pos() + 1].is('{')) {
            const long long n = r.label();
            r.expect('{');
            const double val = r.number();
            r.expect('}');
            v.assign((size_t)n, val);
        } else {
            r.list_open();
            while (!r.accept(')')) v.push_back(r.number());
        }
        if (v.size() != np) r.fail("field size differs from the parcel count");
        return v;
    };
    const std::vector<double> U = read_vec("U");
    const std::vector<double> tid = read_num("typeId", true, 0.0);
    const std::vector<double> nP = read_num("nParticle", false, -1.0);
    const std::vector<double> cm = read_num("chargeModifier", false, 1.0);
    c.parcels.assign(c.species.size(), HostParcels{});
    for (size_t i = 0; i < np; ++i) {
        const int s = (int)std::llround(tid[i]);
        if (s < 0 || s >= (int)c.species.size()) throw FoamError(cd + "/typeId: species " + std::to_string(s) + " undefined");
        if (cm[i] != 1.0) throw FoamError(cd + "/chargeModifier: only 1 is supported (the charge is per species)");
        if (nP[i] >= 0.0 && nP[i] != c.species[s].n_equivalent)
            throw FoamError(cd + "/nParticle: parcels of a species must carry its nEquivalentParticles");
        if (cell[i] < 0 || cell[i] >= c.n_cells) throw FoamError(cd + ": cell out of range");
        HostParcels& P = c.parcels[s];
        P.cell.push_back(cell[i]);
        P.U.insert(P.U.end(), {U[3 * i], U[3 * i + 1], U[3 * i + 2]});
        if (coords) {
            // (cell, tetFace, tetPt) -> the tet of the decomposition
            int found = -1;
            for (int t = c.hm.cell_tet0[cell[i]]; t < c.hm.cell_tet0[cell[i] + 1]; ++t)
                if (c.hm.tet_face[t] == tetf[i] && c.hm.tet_pt[t] == tetp[i]) { found = t; break; }
            if (found < 0)
                throw FoamError(cd + "/coordinates: no tet (cell " + std::to_string(cell[i]) + ", face " +
                                std::to_string(tetf[i]) + ", point " + std::to_string(tetp[i]) + ")");
            P.tet.push_back(found);
            P.bary.insert(P.bary.end(), {bary[3 * i], bary[3 * i + 1], bary[3 * i + 2]});
        } else {
            P.x.insert(P.x.end(), {x[3 * i], x[3 * i + 1], x[3 * i + 2]});
        }
    }
    c.post = InitPost{};
}

template <class T>
std::string list_text(const std::string& head, const std::vector<T>& v, int per, int prec) {
    std::string s = head;
    const size_t n = per ? v.size() / per : 0;
    s += std::to_string(n) + "\n(\n";
    for (size_t i = 0; i < n; ++i) {
        if (per == 1) {
            s += picmi::foam::fmt((double)v[i], prec);
        } else {
            s += "(";
            for (int k = 0; k < per; ++k) {
                if (k) s += " ";
                s += picmi::foam::fmt((double)v[(size_t)per * i + k], prec);
            }
            s += ")";
        }
        s += "\n";
    }
    s += ")\n";
    return s;
}

std::string vol_field(const Case& c, const std::string& cls, const std::string& loc, const std::string& name,
                      const std::string& dims, const std::vector<double>& v, int per, bool phi) {
    std::string s = picmi::foam::header(cls, loc, name);
    s += "dimensions      " + dims + ";\n\n";
    s += "internalField   nonuniform List<" + std::string(per == 1 ? "scalar" : "vector") + "> ";
    s += list_text("", v, per, 17);
    s += ";\n\nboundaryField\n{\n";
    for (size_t p = 0; p < c.patches.size(); ++p) {
        const pic_patch& pp = c.patches[p];
        s += "    " + c.patch_names[p] + "\n    {\n";
        if (pp.type == PIC_PATCH_CYCLIC) s += "        type            cyclic;\n";
        else if (pp.type == PIC_PATCH_EMPTY) s += "        type            empty;\n";
        else if (phi && pp.phi_bc == PIC_BC_FIXED_VALUE)
            s += "        type            fixedValue;\n        value           uniform " + picmi::foam::fmt(pp.phi_value) + ";\n";
        else s += "        type            zeroGradient;\n";
        s += "    }\n";
    }
    s += "}\n";
    return s;
}

void check(int rc, pic_ctx* ctx, const char* what) {
    if (rc != PIC_OK) throw FoamError(std::string(what) + ": " + (pic_last_error(ctx) ? pic_last_error(ctx) : "failed"));
}

void write_time(Case& c, pic_ctx* ctx, const std::string& tn) {
    const std::string cd = cloud_dir(c, tn);
    mkdirs(cd);
    const std::string loc = tn + "/lagrangian/" + c.cloud;
    std::vector<double> pos, bary4, U;
    std::vector<int32_t> cell, tface, tpt, tid, cmod;
    std::vector<double> nP;
    int64_t nt = 0;
    int32_t nc = 0, np = 0;
    check(pic_mesh_sizes(ctx, &nt, &nc, &np), ctx, "pic_mesh_sizes");
    std::vector<int32_t> tcell(nt), tfc(nt), tp(nt);
    check(pic_mesh_tets(ctx, tcell.data(), tfc.data(), tp.data()), ctx, "pic_mesh_tets");
    for (int s = 0; s < (int)c.species.size(); ++s) {
        int64_t n = 0;
        check(pic_particles_count(ctx, s, &n), ctx, "pic_particles_count");
        std::vector<double> y(3 * (size_t)n), u(3 * (size_t)n), x(3 * (size_t)n);
        std::vector<int32_t> t(n);
        std::vector<uint32_t> id(n);
        int64_t got = 0;
        if (n) {
            check(pic_particles_get(ctx, s, n, &got, y.data(), t.data(), u.data(), id.data()), ctx, "pic_particles_get");
            check(pic_particles_positions(ctx, s, n, x.data()), ctx, "pic_particles_positions");
        }
        for (int64_t i = 0; i < got; ++i) {
            const int tt = t[i];
            pos.insert(pos.end(), {x[3 * i], x[3 * i + 1], x[3 * i + 2]});
            bary4.insert(bary4.end(), {1.0 - (y[3 * i] + y[3 * i + 1] + y[3 * i + 2]), y[3 * i], y[3 * i + 1], y[3 * i + 2]});
            cell.push_back(tcell[tt]);
            tface.push_back(tfc[tt]);
            tpt.push_back(tp[tt]);
            U.insert(U.end(), {u[3 * i], u[3 * i + 1], u[3 * i + 2]});
            tid.push_back(s);
            cmod.push_back(1);
            nP.push_back(c.species[s].n_equivalent);
        }
    }
    const size_t n = cell.size();
    {
        std::string s = picmi::foam::header("Cloud<PICParcel>", loc, "positions");
        s += std::to_string(n) + "\n(\n";
        for (size_t i = 0; i < n; ++i)
            s += "(" + picmi::foam::fmt(pos[3 * i]) + " " + picmi::foam::fmt(pos[3 * i + 1]) + " " +
                 picmi::foam::fmt(pos[3 * i + 2]) + ") " + std::to_string(cell[i]) + "\n";
        s += ")\n";
        write_file(cd + "/positions", s);
    }
    {
        std::string s = picmi::foam::header("Cloud<PICParcel>", loc, "coordinates");
        s += std::to_string(n) + "\n(\n";
        for (size_t i = 0; i < n; ++i)
            s += "(" + picmi::foam::fmt(bary4[4 * i]) + " " + picmi::foam::fmt(bary4[4 * i + 1]) + " " +
                 picmi::foam::fmt(bary4[4 * i + 2]) + " " + picmi::foam::fmt(bary4[4 * i + 3]) + ") " +
                 std::to_string(cell[i]) + " " + std::to_string(tface[i]) + " " + std::to_string(tpt[i]) + "\n";
        s += ")\n";
        write_file(cd + "/coordinates", s);
    }
    write_file(cd + "/U", list_text(picmi::foam::header("vectorField", loc, "U"), U, 3, 17));
    write_file(cd + "/typeId", list_text(picmi::foam::header("labelField", loc, "typeId"), tid, 1, 17));
    write_file(cd + "/chargeModifier", list_text(picmi::foam::header("labelField", loc, "chargeModifier"), cmod, 1, 17));
    write_file(cd + "/nParticle", list_text(picmi::foam::header("scalarField", loc, "nParticle"), nP, 1, 17));
    // fields
    std::vector<double> rho(nc), E(3 * (size_t)nc);
    check(pic_get_cell_field(ctx, PIC_FIELD_RHO, rho.data()), ctx, "rhoCharge");
    check(pic_get_cell_field(ctx, PIC_FIELD_E, E.data()), ctx, "E");
    const std::string td = c.dir + "/" + tn;
    for (const auto& wf : c.wall_files)
        write_file(td + "/" + wf.first,
                   std::regex_replace(wf.second, std::regex("location\\s+\"[^\"]*\";"), "location    \"" + tn + "\";",
                                      std::regex_constants::format_first_only));
    write_file(td + "/E", vol_field(c, "volVectorField", tn, "E", "[1 1 -3 0 0 -1 0]", E, 3, false));
    write_file(td + "/rhoCharge", vol_field(c, "volScalarField", tn, "rhoCharge", "[0 -3 1 0 0 1 0]", rho, 1, false));
    if (std::string(c.models.maxwell_solver) != "none") {
        std::vector<double> phi(nc);
        check(pic_get_cell_field(ctx, PIC_FIELD_PHI, phi.data()), ctx, "phiE");
        write_file(td + "/phiE", vol_field(c, "volScalarField", tn, "phiE", "[1 2 -3 0 0 -1 0]", phi, 1, true));
    }
}

// positions of host parcels (initialised ones carry a tet and barycentric coordinates)
void host_positions(const Case& c, const HostParcels& P, std::vector<double>& x) {
    const size_t n = P.size();
    x.resize(3 * n);
    if (P.tet.empty()) { x = P.x; return; }
    for (size_t i = 0; i < n; ++i) {
        V3 X[4];
        tet_vertices(c.hm, P.tet[i], X);
        const double y1 = P.bary[3 * i], y2 = P.bary[3 * i + 1], y3 = P.bary[3 * i + 2], y0 = 1.0 - (y1 + y2 + y3);
        const V3 p = y0 * X[0] + y1 * X[1] + y2 * X[2] + y3 * X[3];
        x[3 * i] = p.x; x[3 * i + 1] = p.y; x[3 * i + 2] = p.z;
    }
}

}  // namespace

// ====================================================================== C-ABI

namespace {
int run_guarded(pic_case* c, const std::function<int()>& f) {
    try {
        return f();
    } catch (const FoamError& e) {
        return c->fail(PIC_ERR_INVALID, e.what());
    } catch (const std::bad_alloc&) {
        return c->fail(PIC_ERR_INVALID, "out of host memory");
    } catch (const std::exception& e) {
        return c->fail(PIC_ERR_INVALID, e.what());
    }
}
}  // namespace

extern "C" {

int pic_case_open(const char* case_dir, pic_case** out) {
    if (!out) return PIC_ERR_INVALID;
    *out = nullptr;
    if (!case_dir) return PIC_ERR_INVALID;
    pic_case* c = new (std::nothrow) pic_case();
    if (!c) return PIC_ERR_INVALID;
    *out = c;
    c->dir = case_dir;
    while (c->dir.size() > 1 && c->dir.back() == '/') c->dir.pop_back();
    return run_guarded(c, [&]() -> int {
        open_case(*c);
        return PIC_OK;
    });
}

void pic_case_close(pic_case* c) { delete c; }

const char* pic_case_error(const pic_case* c) { return c ? c->error.c_str() : "no case"; }

int pic_case_sizes(const pic_case* c, int32_t* ns, int32_t* nc, int32_t* np) {
    if (!c) return PIC_ERR_INVALID;
    if (ns) *ns = (int32_t)c->species.size();
    if (nc) *nc = c->n_cells;
    if (np) *np = (int32_t)c->patches.size();
    return PIC_OK;
}

int pic_case_times(const pic_case* c, double* start, double* end, double* dt, double* wi) {
    if (!c) return PIC_ERR_INVALID;
    if (start) *start = c->start;
    if (end) *end = c->end;
    if (dt) *dt = c->dt;
    if (wi) *wi = c->write_steps ? c->write_interval * c->dt : c->write_interval;
    return PIC_OK;
}

const char* pic_case_species_name(const pic_case* c, int32_t s) {
    if (!c || s < 0 || s >= (int32_t)c->species_names.size()) return nullptr;
    return c->species_names[s].c_str();
}

const char* pic_case_patch_name(const pic_case* c, int32_t p) {
    if (!c || p < 0 || p >= (int32_t)c->patch_names.size()) return nullptr;
    return c->patch_names[p].c_str();
}

int pic_case_injection(const pic_case* c, int32_t capacity, pic_injection* out, int32_t* n) {
    if (!c || !n || capacity < 0 || (capacity > 0 && !out)) return PIC_ERR_INVALID;
    *n = (int32_t)c->injection.size();
    for (int32_t k = 0; k < *n && k < capacity; ++k) {
        out[k] = c->injection[k];
        out[k].time = c->now;
    }
    return PIC_OK;
}

int pic_case_describe(const pic_case* c, pic_mesh* mesh, pic_species* species, pic_models* models) {
    if (!c) return PIC_ERR_INVALID;
    if (mesh) {
        mesh->n_points = (int32_t)(c->points.size() / 3);
        mesh->n_faces = (int32_t)c->fo.size() - 1;
        mesh->n_internal_faces = (int32_t)c->nbr.size();
        mesh->n_cells = c->n_cells;
        mesh->n_patches = (int32_t)c->patches.size();
        mesh->points = c->points.data();
        mesh->face_offsets = c->fo.data();
        mesh->face_points = c->fp.data();
        mesh->owner = c->owner.data();
        mesh->neighbour = c->nbr.data();
        mesh->patches = c->patches.data();
    }
    if (species)
        for (size_t s = 0; s < c->species.size(); ++s) {
            species[s] = c->species[s];
            species[s].capacity = (int64_t)c->parcels[s].size();
        }
    if (models) *models = c->models;
    return PIC_OK;
}

int pic_case_initialise(pic_case* c, const char* dict_name, uint64_t seed) {
    if (!c) return PIC_ERR_INVALID;
    return run_guarded(c, [&]() -> int {
        initialise(*c, dict_name ? dict_name : "picInitialiseDict", seed);
        return PIC_OK;
    });
}

int pic_case_read_parcels(pic_case* c, const char* tn) {
    if (!c || !tn) return PIC_ERR_INVALID;
    return run_guarded(c, [&]() -> int {
        read_parcels(*c, tn);
        char* end = nullptr;
        const double t = std::strtod(tn, &end);
        if (end != tn && *end == '\0') { c->now = t; c->now_name = tn; }
        return PIC_OK;
    });
}

int pic_case_parcel_count(const pic_case* c, int32_t s, int64_t* n) {
    if (!c || !n || s < 0 || s >= (int32_t)c->species.size()) return PIC_ERR_INVALID;
    *n = (int64_t)c->parcels[s].size();
    return PIC_OK;
}

int pic_case_parcels(const pic_case* c, int32_t s, int64_t cap, double* x, int32_t* cell, double* U) {
    if (!c || s < 0 || s >= (int32_t)c->species.size()) return PIC_ERR_INVALID;
    const HostParcels& P = c->parcels[s];
    if ((int64_t)P.size() > cap) return PIC_ERR_CAPACITY;
    if (x) {
        std::vector<double> xx;
        host_positions(*c, P, xx);
        std::memcpy(x, xx.data(), sizeof(double) * xx.size());
    }
    if (cell) std::memcpy(cell, P.cell.data(), sizeof(int32_t) * P.size());
    if (U) std::memcpy(U, P.U.data(), sizeof(double) * P.U.size());
    return PIC_OK;
}

int pic_case_create(pic_case* c, int32_t precision, double factor, pic_ctx** out) {
    if (!c || !out) return PIC_ERR_INVALID;
    *out = nullptr;
    if (!(factor >= 1.0)) return c->fail(PIC_ERR_INVALID, "capacity_factor must be >= 1");
    pic_ctx* ctx = nullptr;
    const int rc = run_guarded(c, [&]() -> int {
        pic_mesh mesh{};
        std::vector<pic_species> sp(c->species.size());
        pic_models mo{};
        pic_case_describe(c, &mesh, sp.data(), &mo);
        for (pic_species& s : sp) s.capacity = std::max<int64_t>(1024, (int64_t)std::ceil((double)s.capacity * factor));
        mo.precision = precision;
        int r = pic_create(&mesh, sp.data(), (int32_t)sp.size(), &mo, &ctx);
        if (r != PIC_OK) return c->fail(r, std::string("pic_create: ") + pic_last_error(nullptr));
        auto fail = [&](int code, const char* what) {
            c->fail(code, std::string(what) + ": " + pic_last_error(ctx));
            pic_destroy(ctx);
            ctx = nullptr;
            return code;
        };
        if (!c->E.empty() && (r = pic_set_cell_field(ctx, PIC_FIELD_E, c->E.data())) != PIC_OK) return fail(r, "E");
        if (!c->B.empty() && (r = pic_set_cell_field(ctx, PIC_FIELD_B, c->B.data())) != PIC_OK) return fail(r, "B");
        if (c->space_charge != 0.0) {
            std::vector<double> sc(c->n_cells);
            for (int i = 0; i < c->n_cells; ++i) sc[i] = c->space_charge / (double)c->n_cells / c->hm.V[i];
            if ((r = pic_set_cell_field(ctx, PIC_FIELD_SPACE_CHARGE, sc.data())) != PIC_OK) return fail(r, "space charge");
        }
        for (const pic_case::CrossSection& x : c->xs)
            if ((r = pic_set_cross_section(ctx, x.process, (int32_t)x.e.size(), x.e.data(), x.s.data(), x.threshold)) !=
                PIC_OK)
                return fail(r, "cross section");
        for (int s = 0; s < (int)c->species.size(); ++s) {
            const HostParcels& P = c->parcels[s];
            const int64_t n = (int64_t)P.size();
            if (!n) continue;
            std::vector<uint32_t> ids(n);
            for (int64_t i = 0; i < n; ++i) ids[i] = (uint32_t)i;
            if (!P.tet.empty())
                r = pic_particles_set(ctx, s, n, P.bary.data(), P.tet.data(), P.U.data(), ids.data());
            else
                r = pic_particles_set_positions(ctx, s, n, P.x.data(), P.cell.data(), P.U.data(), ids.data());
            if (r != PIC_OK) return fail(r, "parcels");
        }
        for (int k = 0; k < 2 && r == PIC_OK; ++k) {
            if (std::string(k == 0 ? c->models.electron_neutral : c->models.ion_neutral_collision) == "none") continue;
            const pic_case::XsInit& xs = c->xs_init[k];
            if (!c->smax[k].empty())
                r = pic_set_cell_field(ctx, k == 0 ? PIC_FIELD_MCC_SIGMA_CR_MAX : PIC_FIELD_ION_SIGMA_CR_MAX,
                                       c->smax[k].data());
            else if (xs.temperature)
                r = pic_mcc_initialize_temperature(ctx, k, xs.T);
            else
                r = k == 0 ? pic_mcc_initialize(ctx, xs.emax, xs.estep) : pic_ion_mcc_initialize(ctx, xs.emax, xs.estep);
        }
        if (r != PIC_OK) return fail(r, "MCC (sigma cR)max");
        // the ion-neutral parcel field as read, and the binary model's field: initialised by
        // picInitialise, else as read (else zero, as the reference leaves it)
        if (!c->smax_ip.empty() && std::string(c->models.ion_neutral_collision) != "none")
            r = pic_set_cell_field(ctx, PIC_FIELD_ION_PARCEL_SIGMA_CR_MAX, c->smax_ip.data());
        if (r == PIC_OK && std::string(c->models.binary_collision) != "none") {
            if (c->binary_init)
                r = pic_binary_initialize(ctx, c->init_T.data(), c->init_n.data());
            else if (!c->smax_bin.empty())
                r = pic_set_cell_field(ctx, PIC_FIELD_BINARY_SIGMA_CR_MAX, c->smax_bin.data());
        }
        if (r != PIC_OK) return fail(r, "binary / ion-neutral (sigma cR)max");
        if (!c->injection.empty()) {
            std::vector<pic_injection> inj = c->injection;
            for (pic_injection& q : inj) q.time = c->now;
            if ((r = pic_set_injection(ctx, inj.data(), (int32_t)inj.size())) != PIC_OK) return fail(r, "SimpleInjection");
        }
        if ((r = pic_deposit(ctx)) != PIC_OK) return fail(r, "charge density");
        if (c->post.active) {
            // picInitialise.C:213-237
            if (c->post.solve_fields && std::string(c->models.maxwell_solver) != "none" &&
                (r = pic_poisson_solve(ctx, nullptr)) != PIC_OK)
                return fail(r, "solveMaxwellEquations");
            if (c->post.leap_frog && (r = pic_push(ctx, -0.5 * c->dt)) != PIC_OK) return fail(r, "initalizeLeapFrog");
        }
        return PIC_OK;
    });
    *out = ctx;
    return rc;
}

int pic_case_write(pic_case* c, pic_ctx* ctx, const char* tn) {
    if (!c || !ctx) return PIC_ERR_INVALID;
    return run_guarded(c, [&]() -> int {
        write_time(*c, ctx, tn ? tn : c->now_name);
        return PIC_OK;
    });
}

int pic_case_run(pic_case* c, pic_ctx* ctx, double until, int64_t* steps) {
    if (!c || !ctx) return PIC_ERR_INVALID;
    if (steps) *steps = 0;
    const double stop = until > 0.0 ? until : c->end;
    return run_guarded(c, [&]() -> int {
        // Time::run/write: steps of deltaT; a write every writeInterval steps (timeStep) or
        // when the time reaches the next multiple of writeInterval (runTime)
        int64_t k0 = std::llround((c->now - c->start) / c->dt);
        const int64_t k1 = std::llround((stop - c->start) / c->dt);
        while (k0 < k1) {
            int64_t next = k1;
            if (c->write_interval > 0.0) {
                if (c->write_steps) {
                    const int64_t wi = std::max<int64_t>(1, std::llround(c->write_interval));
                    next = std::min(next, (k0 / wi + 1) * wi);
                } else {
                    const double tw = (std::floor((c->start + k0 * c->dt) / c->write_interval + 1e-9) + 1.0) * c->write_interval;
                    next = std::min(next, (int64_t)std::llround((tw - c->start) / c->dt));
                    if (next <= k0) next = k0 + 1;
                }
            }
            pic_step_stats st{};
            const int r = pic_run(ctx, c->dt, next - k0, &st);
            if (r != PIC_OK) return c->fail(r, std::string("pic_run: ") + pic_last_error(ctx));
            if (steps) *steps += next - k0;
            k0 = next;
            c->now = c->start + (double)k0 * c->dt;
            c->now_name = time_name(c->now);
            bool write = false;
            if (c->write_interval > 0.0) {
                if (c->write_steps) write = k0 % std::max<int64_t>(1, std::llround(c->write_interval)) == 0;
                else {
                    const double q = c->now / c->write_interval;
                    write = std::fabs(q - std::round(q)) < 1e-6;
                }
            }
            if (write || k0 == k1) write_time(*c, ctx, c->now_name);
        }
        return PIC_OK;
    });
}

double pic_case_time(const pic_case* c) { return c ? c->now : 0.0; }

const char* pic_case_time_name(const pic_case* c) { return c ? c->now_name.c_str() : nullptr; }

}  // extern "C"
