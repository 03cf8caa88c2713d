// Host-side mesh preparation for libpicmi: OpenFOAM-9 primitiveMesh geometry, the cell
// tet decomposition every parcel refers to, the tet adjacency/permutation table used by
// the tracking kernel, cyclic point coupling, and the Poisson operator.
// Compiled with -ffp-contract=off so the geometry rounds as written.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>

#include "internal.h"

namespace picmi {

namespace {
struct v3 {
    double x, y, z;
};
inline v3 mk(double x, double y, double z) { return {x, y, z}; }
inline v3 operator+(v3 a, v3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
inline v3 operator-(v3 a, v3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
inline v3 operator*(double s, v3 a) { return mk(s * a.x, s * a.y, s * a.z); }
inline v3 operator/(v3 a, double s) { return mk(a.x / s, a.y / s, a.z / s); }
inline double dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline v3 cross(v3 a, v3 b) { return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
inline double mag(v3 a) { return std::sqrt(dot(a, a)); }
inline v3 ld(const double* p) { return mk(p[0], p[1], p[2]); }
inline void st(double* p, v3 a) { p[0] = a.x; p[1] = a.y; p[2] = a.z; }

int uf_find(std::vector<int32_t>& p, int a) {
    while (p[a] != a) { p[a] = p[p[a]]; a = p[a]; }
    return a;
}

}  // namespace

// tetIndices::faceTriIs: (f[b], f[b+pt], f[b+pt+1]) around the face from its base point b,
// A/B swapped for the neighbour
void tet_points(const HostMesh& m, int t, int vid[3]) {
    int f = m.tet_face[t], pt = m.tet_pt[t], c = m.tet_cell[t];
    const int32_t* fp = m.fp.data() + m.fo[f];
    const int n = m.fo[f + 1] - m.fo[f], b = m.fbase.empty() ? 0 : m.fbase[f];
    vid[0] = fp[b];
    const int pa = fp[(b + pt) % n], pb = fp[(b + pt + 1) % n];
    if (m.owner[f] == c) { vid[1] = pa; vid[2] = pb; }
    else { vid[1] = pb; vid[2] = pa; }
}

namespace {

// quality of the tet (a, b, c, d): its signed volume over the volume of the regular tet of
// the same circumradius R, (8 / (9 sqrt 3)) R^3 (1 for a regular tet, <= 0 when inverted or
// flat), the measure OpenFOAM's tetrahedron::quality takes
double tet_quality(v3 a, v3 b, v3 c, v3 d) {
    const v3 u = b - a, v = c - a, w = d - a;
    const double t6 = dot(u, cross(v, w));
    if (t6 == 0.0) return 0.0;
    const v3 o = (1.0 / (2.0 * t6)) * ((dot(u, u) * cross(v, w) + dot(v, v) * cross(w, u)) + dot(w, w) * cross(u, v));
    const double R = mag(o);
    return (t6 / 6.0) / (8.0 / (9.0 * std::sqrt(3.0)) * R * R * R + 1e-150);
}

// the first point of face f from which every fan tet has quality above minTetQuality
// (1e-15) with the owner cell centre oc and, if given, the neighbour cell centre nc (its
// tets with A and B swapped); -1 when no point qualifies
int face_base_point(const HostMesh& m, int f, v3 oc, const v3* nc) {
    const int32_t* fp = m.fp.data() + m.fo[f];
    const int n = m.fo[f + 1] - m.fo[f];
    for (int b = 0; b < n; ++b) {
        const v3 pb = ld(&m.points[3 * fp[b]]);
        double q = 1e300;
        for (int i = 1; i <= n - 2; ++i) {
            const v3 pA = ld(&m.points[3 * fp[(b + i) % n]]), pB = ld(&m.points[3 * fp[(b + i + 1) % n]]);
            q = std::min(q, tet_quality(oc, pb, pA, pB));
            if (nc) q = std::min(q, tet_quality(*nc, pb, pB, pA));
        }
        if (q > 1e-15) return b;
    }
    return -1;
}
}  // namespace

std::string build_host_mesh(HostMesh& m) {
    const int nf = m.n_faces, nc = m.n_cells, np = m.n_points, ni = m.n_internal;
    if ((int)m.fo.size() != nf + 1) return "face_offsets size";
    for (int f = 0; f < nf; ++f) {
        if (m.owner[f] < 0 || m.owner[f] >= nc) return "owner out of range at face " + std::to_string(f);
        if (m.fo[f + 1] - m.fo[f] < 3) return "face with fewer than 3 points: " + std::to_string(f);
    }
    for (int f = 0; f < ni; ++f)
        if (m.nbr[f] < 0 || m.nbr[f] >= nc) return "neighbour out of range at face " + std::to_string(f);
    for (int i = 0; i < m.fo[nf]; ++i)
        if (m.fp[i] < 0 || m.fp[i] >= np) return "face point out of range";

    // face centres and area vectors (primitiveMeshFaceCentresAndAreas)
    m.Cf.assign(3 * nf, 0.0);
    m.Sf.assign(3 * nf, 0.0);
    for (int f = 0; f < nf; ++f) {
        int s = m.fo[f], n = m.fo[f + 1] - s;
        const int32_t* fp = m.fp.data() + s;
        if (n == 3) {
            v3 p0 = ld(&m.points[3 * fp[0]]), p1 = ld(&m.points[3 * fp[1]]), p2 = ld(&m.points[3 * fp[2]]);
            st(&m.Cf[3 * f], (1.0 / 3.0) * ((p0 + p1) + p2));
            st(&m.Sf[3 * f], 0.5 * cross(p1 - p0, p2 - p0));
        } else {
            v3 est = mk(0, 0, 0);
            for (int i = 0; i < n; ++i) est = est + ld(&m.points[3 * fp[i]]);
            est = est / (double)n;
            v3 sumN = mk(0, 0, 0), sumAc = mk(0, 0, 0);
            double sumA = 0.0;
            for (int i = 0; i < n; ++i) {
                v3 p = ld(&m.points[3 * fp[i]]), q = ld(&m.points[3 * fp[(i + 1) % n]]);
                v3 c = (p + q) + est;
                v3 nn = cross(q - p, est - p);
                double a = mag(nn);
                sumN = sumN + nn;
                sumA += a;
                sumAc = sumAc + a * c;
            }
            if (sumA < 1e-300) { st(&m.Cf[3 * f], est); st(&m.Sf[3 * f], mk(0, 0, 0)); }
            else { st(&m.Cf[3 * f], (1.0 / 3.0) * (sumAc / sumA)); st(&m.Sf[3 * f], 0.5 * sumN); }
        }
    }
    // cell centres and volumes (primitiveMeshCellCentresAndVols)
    {
        std::vector<v3> est(nc, mk(0, 0, 0)), ctr(nc, mk(0, 0, 0));
        std::vector<int> cnt(nc, 0);
        std::vector<double> vol(nc, 0.0);
        for (int f = 0; f < nf; ++f) { est[m.owner[f]] = est[m.owner[f]] + ld(&m.Cf[3 * f]); cnt[m.owner[f]]++; }
        for (int f = 0; f < ni; ++f) { est[m.nbr[f]] = est[m.nbr[f]] + ld(&m.Cf[3 * f]); cnt[m.nbr[f]]++; }
        for (int c = 0; c < nc; ++c) {
            if (cnt[c] < 4) return "cell with fewer than 4 faces: " + std::to_string(c);
            est[c] = est[c] / (double)cnt[c];
        }
        for (int f = 0; f < nf; ++f) {
            int o = m.owner[f];
            v3 cf = ld(&m.Cf[3 * f]), sf = ld(&m.Sf[3 * f]);
            double pv = dot(sf, cf - est[o]);
            if (pv < 1e-300) pv = 1e-300;
            v3 pc = 0.75 * cf + 0.25 * est[o];
            ctr[o] = ctr[o] + pv * pc;
            vol[o] += pv;
        }
        for (int f = 0; f < ni; ++f) {
            int o = m.nbr[f];
            v3 cf = ld(&m.Cf[3 * f]), sf = ld(&m.Sf[3 * f]);
            double pv = dot(sf, est[o] - cf);
            if (pv < 1e-300) pv = 1e-300;
            v3 pc = 0.75 * cf + 0.25 * est[o];
            ctr[o] = ctr[o] + pv * pc;
            vol[o] += pv;
        }
        m.C.assign(3 * nc, 0.0);
        m.V.assign(nc, 0.0);
        for (int c = 0; c < nc; ++c) { st(&m.C[3 * c], ctr[c] / vol[c]); m.V[c] = vol[c] / 3.0; }
    }
    // cells(): owned faces ascending, then neighboured faces ascending
    {
        std::vector<int> cnt(nc + 1, 0);
        for (int f = 0; f < nf; ++f) cnt[m.owner[f]]++;
        for (int f = 0; f < ni; ++f) cnt[m.nbr[f]]++;
        m.cfo.assign(nc + 1, 0);
        for (int c = 0; c < nc; ++c) m.cfo[c + 1] = m.cfo[c] + cnt[c];
        m.cf.assign(m.cfo[nc], 0);
        std::fill(cnt.begin(), cnt.end(), 0);
        for (int f = 0; f < nf; ++f) { int c = m.owner[f]; m.cf[m.cfo[c] + cnt[c]++] = f; }
        for (int f = 0; f < ni; ++f) { int c = m.nbr[f]; m.cf[m.cfo[c] + cnt[c]++] = f; }
    }
    // tets (cell, face, pt), face base point 0
    {
        m.cell_tet0.assign(nc + 1, 0);
        int64_t nt = 0;
        for (int c = 0; c < nc; ++c) {
            m.cell_tet0[c] = (int32_t)nt;
            for (int k = m.cfo[c]; k < m.cfo[c + 1]; ++k) { int f = m.cf[k]; nt += m.fo[f + 1] - m.fo[f] - 2; }
        }
        if (nt > 2000000000LL) return "too many tets";
        m.cell_tet0[nc] = (int32_t)nt;
        m.tet_cell.resize(nt); m.tet_face.resize(nt); m.tet_pt.resize(nt);
        nt = 0;
        for (int c = 0; c < nc; ++c)
            for (int k = m.cfo[c]; k < m.cfo[c + 1]; ++k) {
                int f = m.cf[k], n = m.fo[f + 1] - m.fo[f];
                for (int pt = 1; pt <= n - 2; ++pt) { m.tet_cell[nt] = c; m.tet_face[nt] = f; m.tet_pt[nt] = pt; ++nt; }
            }
    }
    // patches: face -> patch, separations, coupled points
    m.face_patch.assign(nf, -1);
    m.psep.assign(3 * (m.n_patches + 1), 0.0);
    m.pmaster.resize(np);
    std::iota(m.pmaster.begin(), m.pmaster.end(), 0);
    for (int pi = 0; pi < m.n_patches; ++pi) {
        if (m.pstart[pi] < ni || m.pstart[pi] + m.psize[pi] > nf) return "patch range invalid";
        for (int f = m.pstart[pi]; f < m.pstart[pi] + m.psize[pi]; ++f) m.face_patch[f] = pi;
    }
    for (int f = ni; f < nf; ++f) if (m.face_patch[f] < 0) return "boundary face not in any patch";
    // polyMesh::calcDirections: the summed |area components| of the empty patches, normalised;
    // a component above 1e-6 is an empty direction
    {
        double ed[3] = {0.0, 0.0, 0.0};
        bool any = false;
        for (int pi = 0; pi < m.n_patches; ++pi) {
            if (m.ptype[pi] != 3) continue;
            for (int f = m.pstart[pi]; f < m.pstart[pi] + m.psize[pi]; ++f)
                for (int d = 0; d < 3; ++d) ed[d] += std::fabs(m.Sf[3 * (size_t)f + d]);
            any = any || m.psize[pi] > 0;
        }
        m.track_mask = 0;
        const double e = std::sqrt(ed[0] * ed[0] + ed[1] * ed[1] + ed[2] * ed[2]);
        if (any && e > 0.0)
            for (int d = 0; d < 3; ++d) if (ed[d] / e > 1e-6) m.track_mask |= 1u << d;
    }
    for (int pi = 0; pi < m.n_patches; ++pi) {
        if (m.ptype[pi] != 2 || m.psize[pi] == 0) continue;
        int q = m.pnbr[pi];
        if (q < 0 || q >= m.n_patches || m.psize[q] != m.psize[pi] || m.ptype[q] != 2) return "cyclic partner mismatch";
        int f0 = m.pstart[pi], g0 = m.pstart[q];
        st(&m.psep[3 * pi], ld(&m.Cf[3 * g0]) - ld(&m.Cf[3 * f0]));
        v3 sep = ld(&m.psep[3 * pi]);
        for (int k = 0; k < m.psize[pi]; ++k) {
            int f = f0 + k, g = g0 + k, n = m.fo[f + 1] - m.fo[f];
            if (m.fo[g + 1] - m.fo[g] != n) return "cyclic face size mismatch";
            double scale = mag(ld(&m.Sf[3 * f]));
            scale = std::sqrt(scale);
            for (int j = 0; j < n; ++j) {
                int a = m.fp[m.fo[f] + j];
                v3 img = ld(&m.points[3 * a]) + sep;
                int best = -1;
                double bd = 1e300;
                for (int i = 0; i < n; ++i) {
                    int b = m.fp[m.fo[g] + i];
                    double d = mag(ld(&m.points[3 * b]) - img);
                    if (d < bd) { bd = d; best = b; }
                }
                if (bd > 1e-6 * scale) return "cyclic points do not match under translation";
                int ra = uf_find(m.pmaster, a), rb = uf_find(m.pmaster, best);
                if (ra != rb) { if (ra < rb) m.pmaster[rb] = ra; else m.pmaster[ra] = rb; }
            }
        }
    }
    for (int p = 0; p < np; ++p) m.pmaster[p] = uf_find(m.pmaster, p);

    // face base points (polyMeshTetDecomposition::findFaceBasePts of OpenFOAM-9, restated):
    // internal faces test both cell centres, a cyclic face on the lower-numbered patch of its
    // pair tests its own cell centre and the partner cell's centre carried across, and the
    // partner face then takes the matching point (the two faces list their points in
    // opposite order from a common point 0, so point j there is point n - j here); other
    // boundary faces test the owner alone.  A face without a qualifying point keeps point 0.
    m.fbase.assign(nf, 0);
    {
        std::vector<int> base(nf, -1);
        for (int f = 0; f < nf; ++f) {
            const v3 oc = ld(&m.C[3 * m.owner[f]]);
            if (f < ni) {
                const v3 nc = ld(&m.C[3 * m.nbr[f]]);
                base[f] = face_base_point(m, f, oc, &nc);
                continue;
            }
            const int pi = m.face_patch[f];
            if (m.ptype[pi] == 2) {
                const int q = m.pnbr[pi];
                if (pi > q) continue;   // the partner side follows below
                const int g = m.pstart[q] + (f - m.pstart[pi]);
                const v3 nc = ld(&m.C[3 * m.owner[g]]) - ld(&m.psep[3 * pi]);
                base[f] = face_base_point(m, f, oc, &nc);
            } else {
                base[f] = face_base_point(m, f, oc, nullptr);
            }
        }
        for (int f = ni; f < nf; ++f) {
            const int pi = m.face_patch[f];
            if (m.ptype[pi] != 2 || pi < m.pnbr[pi]) continue;
            const int q = m.pnbr[pi];
            const int g = m.pstart[q] + (f - m.pstart[pi]);
            const int n = m.fo[f + 1] - m.fo[f];
            base[f] = base[g] < 0 ? -1 : (n - base[g]) % n;
        }
        for (int f = 0; f < nf; ++f) m.fbase[f] = base[f] < 0 ? 0 : base[f];
    }

    // tet geometry + adjacency
    const int64_t nt = m.tet_cell.size();
    m.tets.assign(nt, TetGeo{});
    for (int64_t t = 0; t < nt; ++t) {
        int vid[3];
        tet_points(m, (int)t, vid);
        TetGeo& g = m.tets[t];
        v3 x0 = ld(&m.C[3 * m.tet_cell[t]]);
        v3 x1 = ld(&m.points[3 * vid[0]]), x2 = ld(&m.points[3 * vid[1]]), x3 = ld(&m.points[3 * vid[2]]);
        v3 e1 = x1 - x0, e2 = x2 - x0, e3 = x3 - x0;
        v3 r1 = cross(e2, e3), r2 = cross(e3, e1), r3 = cross(e1, e2);
        g.r[0] = r1.x; g.r[1] = r1.y; g.r[2] = r1.z;
        g.r[3] = r2.x; g.r[4] = r2.y; g.r[5] = r2.z;
        g.r[6] = r3.x; g.r[7] = r3.y; g.r[8] = r3.z;
        g.det = dot(e1, r1);
        g.cell = m.tet_cell[t];
        for (int i = 0; i < 3; ++i) { g.vm[i] = m.pmaster[vid[i]]; g.vid[i] = vid[i]; }
        for (int i = 0; i < 4; ++i) g.nbr[i] = -1000000;
        g.perm = 0;
    }
    // packed per tet face: bits 2j..2j+1 hold the slot of this tet whose coordinate moves to
    // slot j of the neighbour (map[k] is the neighbour slot of this tet's slot k; the hit
    // face's own slot, whose coordinate is zero on the face, goes to the new vertex)
    auto source_code = [](const int map[4]) {
        uint32_t code = 0;
        for (int k = 0; k < 4; ++k) code |= (uint32_t)k << (2 * map[k]);
        return code;
    };
    for (int c = 0; c < nc; ++c) {
        int t0 = m.cell_tet0[c], t1 = m.cell_tet0[c + 1];
        for (int t = t0; t < t1; ++t) {
            int vid[3];
            tet_points(m, t, vid);
            TetGeo& g = m.tets[t];
            // internal tet faces i = 1..3: (cc, P, Q)
            for (int i = 1; i <= 3; ++i) {
                int slots[2], k = 0;
                for (int j = 1; j <= 3; ++j) if (j != i) slots[k++] = j;
                int P = vid[slots[0] - 1], Q = vid[slots[1] - 1];
                int found = -1, sv[3] = {0, 0, 0};
                for (int s = t0; s < t1; ++s) {
                    if (s == t) continue;
                    int w[3];
                    tet_points(m, s, w);
                    bool hp = (w[0] == P || w[1] == P || w[2] == P), hq = (w[0] == Q || w[1] == Q || w[2] == Q);
                    if (hp && hq) { found = s; sv[0] = w[0]; sv[1] = w[1]; sv[2] = w[2]; break; }
                }
                if (found < 0) return "open cell (tet face without partner) in cell " + std::to_string(c);
                int map[4] = {0, -1, -1, -1};
                bool used[4] = {true, false, false, false};
                for (int j = 0; j < 3; ++j) {
                    if (sv[j] == P) { map[slots[0]] = j + 1; used[j + 1] = true; }
                    if (sv[j] == Q) { map[slots[1]] = j + 1; used[j + 1] = true; }
                }
                for (int j = 1; j <= 3; ++j) if (!used[j]) map[i] = j;
                g.nbr[i] = found;
                g.perm |= source_code(map) << (8 * i);
            }
            // face triangle i = 0
            int f = m.tet_face[t], pt = m.tet_pt[t];
            if (f < ni) {
                int c2 = (m.owner[f] == c) ? m.nbr[f] : m.owner[f];
                int found = -1;
                for (int s = m.cell_tet0[c2]; s < m.cell_tet0[c2 + 1]; ++s)
                    if (m.tet_face[s] == f && m.tet_pt[s] == pt) { found = s; break; }
                if (found < 0) return "internal face tet partner missing";
                int w[3];
                tet_points(m, found, w);
                int map[4] = {0, -1, -1, -1};
                for (int a = 0; a < 3; ++a)
                    for (int b = 0; b < 3; ++b) if (w[b] == vid[a]) map[a + 1] = b + 1;
                g.nbr[0] = found;
                g.perm |= source_code(map);
            } else {
                int pi = m.face_patch[f];
                if (m.ptype[pi] != 2) { g.nbr[0] = m.ptype[pi] == 3 ? kTetEmptyFace : -1 - pi; continue; }
                int q = m.pnbr[pi];
                int gface = m.pstart[q] + (f - m.pstart[pi]);
                int c2 = m.owner[gface];
                v3 sep = ld(&m.psep[3 * pi]);
                v3 img[3];
                for (int a = 0; a < 3; ++a) img[a] = ld(&m.points[3 * vid[a]]) + sep;
                double scale = mag(img[0] - img[1]) + mag(img[1] - img[2]) + mag(img[2] - img[0]);
                int found = -1, map[4] = {0, -1, -1, -1};
                for (int s = m.cell_tet0[c2]; s < m.cell_tet0[c2 + 1] && found < 0; ++s) {
                    if (m.tet_face[s] != gface) continue;
                    int w[3];
                    tet_points(m, s, w);
                    int mp[4] = {0, -1, -1, -1};
                    bool ok = true;
                    for (int a = 0; a < 3 && ok; ++a) {
                        for (int b = 0; b < 3; ++b)
                            if (mag(ld(&m.points[3 * w[b]]) - img[a]) < 1e-9 * scale) { mp[a + 1] = b + 1; break; }
                        if (mp[a + 1] < 0) ok = false;
                    }
                    if (ok) { found = s; std::memcpy(map, mp, sizeof map); }
                }
                if (found < 0) return "cyclic tet partner missing (face triangulations differ)";
                g.nbr[0] = found;
                g.perm |= source_code(map);
            }
        }
    }
    // cell points + weights
    m.cpo.assign(nc + 1, 0);
    m.cp.clear();
    std::vector<int> tmp;
    for (int c = 0; c < nc; ++c) {
        m.cpo[c] = (int32_t)m.cp.size();
        tmp.clear();
        for (int k = m.cfo[c]; k < m.cfo[c + 1]; ++k) {
            int f = m.cf[k];
            for (int j = m.fo[f]; j < m.fo[f + 1]; ++j) tmp.push_back(m.fp[j]);
        }
        std::sort(tmp.begin(), tmp.end());
        tmp.erase(std::unique(tmp.begin(), tmp.end()), tmp.end());
        for (int p : tmp) m.cp.push_back(p);
    }
    m.cpo[nc] = (int32_t)m.cp.size();
    const int ncp = (int)m.cp.size();
    m.vw_w.assign(ncp, 0.0);
    m.ip_w.assign(ncp, 0.0);
    {
        std::vector<double> sumd(np, 0.0), sumi(np, 0.0);
        for (int c = 0; c < nc; ++c)
            for (int k = m.cpo[c]; k < m.cpo[c + 1]; ++k) {
                int p = m.cp[k];
                double d = mag(ld(&m.points[3 * p]) - ld(&m.C[3 * c]));
                m.vw_w[k] = d;
                m.ip_w[k] = 1.0 / d;
                sumd[m.pmaster[p]] += d;
                sumi[m.pmaster[p]] += 1.0 / d;
            }
        for (int c = 0; c < nc; ++c)
            for (int k = m.cpo[c]; k < m.cpo[c + 1]; ++k) {
                int p = m.pmaster[m.cp[k]];
                m.vw_w[k] /= sumd[p];
                m.ip_w[k] /= sumi[p];
            }
    }
    // master point -> (cell, weight), in cell order (the order ref_cell_to_point sums in)
    {
        std::vector<int> cnt(np + 1, 0);
        for (int k = 0; k < ncp; ++k) cnt[m.pmaster[m.cp[k]]]++;
        m.pco.assign(np + 1, 0);
        for (int p = 0; p < np; ++p) m.pco[p + 1] = m.pco[p] + cnt[p];
        m.pc_cell.assign(ncp, 0);
        m.pc_w.assign(ncp, 0.0);
        std::fill(cnt.begin(), cnt.end(), 0);
        for (int c = 0; c < nc; ++c)
            for (int k = m.cpo[c]; k < m.cpo[c + 1]; ++k) {
                int p = m.pmaster[m.cp[k]];
                int slot = m.pco[p] + cnt[p]++;
                m.pc_cell[slot] = c;
                m.pc_w[slot] = m.ip_w[k];
            }
        m.n_masters = 0;
        for (int p = 0; p < np; ++p) if (m.pmaster[p] == p) ++m.n_masters;
    }
    return "";
}

namespace {
void add_coupling(HostPoisson& P, int i, int j, double a) {
    for (int k = 0; k < P.K; ++k) {
        if (P.nb[i * P.K + k] == j) { P.a[i * P.K + k] += a; return; }
        if (P.nb[i * P.K + k] < 0) { P.nb[i * P.K + k] = j; P.a[i * P.K + k] = a; return; }
    }
}
}  // namespace

// fvm::laplacian (Gauss linear corrected): the implicit part with nonOrthDeltaCoeffs
// 1/max(nf.d, 0.05|d|) (boundary faces 1/(nf.(Cf - C))), cyclic coupling, fixedValue /
// zeroGradient patches, and the vectors of the explicit non-orthogonal correction; IC(0)
// "DIC" factor in the greedy multicolour ordering.
std::string build_host_poisson(const HostMesh& m, HostPoisson& P) {
    const int n = m.n_cells;
    P.n = n;
    int K = 0;
    for (int c = 0; c < n; ++c) K = std::max(K, m.cfo[c + 1] - m.cfo[c]);
    P.K = K;
    P.nb.assign((size_t)n * K, -1);
    P.a.assign((size_t)n * K, 0.0);
    P.diag.assign(n, 0.0);
    P.bsrc.assign(n, 0.0);
    P.w.assign(m.n_faces, 0.0);
    // surfaceInterpolation's nonOrthCorrectionVectors nf - d*deltaCoeff, times |Sf|, out of
    // the face's owner; internal and coupled faces only
    std::vector<double> kf(3 * (size_t)m.n_faces, 0.0);
    auto nonorth_vector = [&](int f, v3 Sf, v3 nf, v3 d, double delta_coeff) {
        const double mS = mag(Sf);
        st(&kf[3 * (size_t)f], mS * (nf - delta_coeff * d));
    };
    bool dir = false;
    for (int f = 0; f < m.n_faces; ++f) {
        v3 Sf = ld(&m.Sf[3 * f]), Cf = ld(&m.Cf[3 * f]);
        double mS = mag(Sf);
        v3 nf = Sf / mS;
        int o = m.owner[f];
        v3 Co = ld(&m.C[3 * o]);
        if (f < m.n_internal) {
            int nb = m.nbr[f];
            v3 Cn = ld(&m.C[3 * nb]);
            v3 d = Cn - Co;
            double nd = dot(nf, d), md = 0.05 * mag(d);
            double a = mS / (nd > md ? nd : md);
            add_coupling(P, o, nb, a);
            add_coupling(P, nb, o, a);
            P.diag[o] += a;
            P.diag[nb] += a;
            double so = std::fabs(dot(Sf, Cf - Co)), sn = std::fabs(dot(Sf, Cn - Cf));
            P.w[f] = sn / (so + sn);
            nonorth_vector(f, Sf, nf, d, a / mS);
            continue;
        }
        int pi = m.face_patch[f];
        if (m.ptype[pi] == 2) {
            int q = m.pnbr[pi];
            int g = m.pstart[q] + (f - m.pstart[pi]);
            int nb = m.owner[g];
            v3 Cn = ld(&m.C[3 * nb]) - ld(&m.psep[3 * pi]);
            double so = std::fabs(dot(Sf, Cf - Co)), sn = std::fabs(dot(Sf, Cn - Cf));
            P.w[f] = sn / (so + sn);
            v3 d = Cn - Co;
            double nd = dot(nf, d), md = 0.05 * mag(d);
            double a = mS / (nd > md ? nd : md);
            nonorth_vector(f, Sf, nf, d, a / mS);
            if (nb == o) continue;
            add_coupling(P, o, nb, a);
            P.diag[o] += a;
        } else if (m.phi_bc[pi] == 1) {
            v3 d = Cf - Co;
            double a = mS / dot(nf, d);
            P.diag[o] += a;
            P.bsrc[o] += a * m.phi_val[pi];
            dir = true;
            P.w[f] = 1.0;
        } else {
            P.w[f] = 1.0;
        }
    }
    P.singular = !dir;
    P.nonorth = false;
    for (int f = 0; f < m.n_faces; ++f)
        if (mag(ld(&kf[3 * (size_t)f])) > 1e-10 * mag(ld(&m.Sf[3 * f]))) P.nonorth = true;
    P.colour.assign(n, 0);
    P.ncol = 0;
    for (int i = 0; i < n; ++i) {
        uint64_t used = 0;
        for (int k = 0; k < K; ++k) {
            int j = P.nb[(size_t)i * K + k];
            if (j >= 0 && j < i) used |= 1ull << P.colour[j];
        }
        int col = 0;
        while (used & (1ull << col)) ++col;
        if (col >= 63) return "colouring needs too many colours";
        P.colour[i] = col;
        P.ncol = std::max(P.ncol, col + 1);
    }
    P.corder.resize(n);
    P.col_start.assign(P.ncol + 1, 0);
    {
        int t = 0;
        for (int col = 0; col < P.ncol; ++col) {
            P.col_start[col] = t;
            for (int i = 0; i < n; ++i) if (P.colour[i] == col) P.corder[t++] = i;
        }
        P.col_start[P.ncol] = t;
    }
    P.rD.assign(n, 0.0);
    for (int t = 0; t < n; ++t) {
        int i = P.corder[t];
        double r = P.diag[i];
        for (int k = 0; k < K; ++k) {
            int j = P.nb[(size_t)i * K + k];
            if (j >= 0 && P.colour[j] < P.colour[i]) r -= P.a[(size_t)i * K + k] * P.a[(size_t)i * K + k] / P.rD[j];
        }
        if (!(r > 0.0)) return "DIC factorisation broke down";
        P.rD[i] = r;
    }
    // Gauss grad entries per cell in ascending face order
    std::vector<std::vector<int>> cf(n);
    for (int c = 0; c < n; ++c) {
        for (int k = m.cfo[c]; k < m.cfo[c + 1]; ++k) cf[c].push_back(m.cf[k]);
        std::sort(cf[c].begin(), cf[c].end());
    }
    P.gfo.assign(n + 1, 0);
    P.g_kind.clear(); P.g_other.clear(); P.g_w.clear(); P.g_val.clear(); P.g_S.clear(); P.g_k.clear();
    for (int c = 0; c < n; ++c) {
        P.gfo[c] = (int32_t)P.g_kind.size();
        for (int f : cf[c]) {
            if (f >= m.n_internal && m.ptype[m.face_patch[f]] == 3) continue;
            v3 Sf = ld(&m.Sf[3 * f]);
            bool own = m.owner[f] == c;
            // an internal face whose owner and neighbour coincide cannot occur; a cell can
            // own a face and be its neighbour only through two different faces
            double sgn = own ? 1.0 : -1.0;
            if (f < m.n_internal) {
                P.g_kind.push_back(0);
                P.g_other.push_back(own ? m.nbr[f] : m.owner[f]);
                // phi_f = w*phi_owner + (1-w)*phi_neighbour
                P.g_w.push_back(own ? P.w[f] : 1.0 - P.w[f]);
                P.g_val.push_back(own ? 1.0 - P.w[f] : P.w[f]);
            } else {
                int pi = m.face_patch[f];
                if (m.ptype[pi] == 2) {
                    int q = m.pnbr[pi];
                    int g = m.pstart[q] + (f - m.pstart[pi]);
                    P.g_kind.push_back(0);
                    P.g_other.push_back(m.owner[g]);
                    P.g_w.push_back(P.w[f]);
                    P.g_val.push_back(1.0 - P.w[f]);
                } else if (m.phi_bc[pi] == 1) {
                    P.g_kind.push_back(1);
                    P.g_other.push_back(-1);
                    P.g_w.push_back(0.0);
                    P.g_val.push_back(m.phi_val[pi]);
                } else {
                    P.g_kind.push_back(2);
                    P.g_other.push_back(-1);
                    P.g_w.push_back(1.0);
                    P.g_val.push_back(0.0);
                }
            }
            P.g_S.push_back(sgn * Sf.x);
            P.g_S.push_back(sgn * Sf.y);
            P.g_S.push_back(sgn * Sf.z);
            // a coupled face's vector already points out of this side's owner
            const double ks = P.nonorth ? sgn : 0.0;
            for (int d = 0; d < 3; ++d) P.g_k.push_back(ks * kf[3 * (size_t)f + d]);
        }
    }
    P.gfo[n] = (int32_t)P.g_kind.size();
    return "";
}

}  // namespace picmi
