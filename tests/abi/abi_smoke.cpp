// A C++ program that includes include/picmi.h, links libpicmi.so and drives one context
// through the C-ABI alone: pic_create -> pic_particles_set_positions -> pic_move ->
// pic_particles_positions -> pic_step -> pic_run -> pic_diag_reduce -> pic_destroy.  The
// mesh (an all-cyclic hex box) comes from a small binary file; with E = B = 0 every
// parcel moves on a straight line, so its end position is x + U dt folded into the box.
// Exit status 0 and "abi smoke ok" on success.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "picmi.h"

#define CHECK(call)                                                                         \
    do {                                                                                    \
        int rc_ = (call);                                                                   \
        if (rc_ != PIC_OK) {                                                                \
            std::fprintf(stderr, "%s failed (%d): %s\n", #call, rc_, pic_last_error(ctx)); \
            return 1;                                                                       \
        }                                                                                   \
    } while (0)

template <class T>
static bool rd(std::FILE* f, std::vector<T>& v, size_t n) {
    v.resize(n);
    return std::fread(v.data(), sizeof(T), n, f) == n;
}

int main(int argc, char** argv) {
    if (argc < 2) { std::fprintf(stderr, "usage: abi_smoke mesh.bin\n"); return 2; }
    std::FILE* f = std::fopen(argv[1], "rb");
    if (!f) { std::perror(argv[1]); return 2; }
    int32_t hdr[5];
    double box[3];
    if (std::fread(hdr, sizeof(int32_t), 5, f) != 5 || std::fread(box, sizeof(double), 3, f) != 3) return 2;
    const int np = hdr[0], nf = hdr[1], ni = hdr[2], nc = hdr[3], npatch = hdr[4];
    std::vector<double> points;
    std::vector<int32_t> fo, fp, owner, nbr, pat;
    if (!rd(f, points, 3 * (size_t)np) || !rd(f, fo, nf + 1)) return 2;
    if (!rd(f, fp, fo[nf]) || !rd(f, owner, nf) || !rd(f, nbr, ni) || !rd(f, pat, 4 * (size_t)npatch)) return 2;
    std::fclose(f);
    std::vector<pic_patch> patches(npatch);
    for (int p = 0; p < npatch; ++p) {
        patches[p].type = pat[4 * p];
        patches[p].start = pat[4 * p + 1];
        patches[p].size = pat[4 * p + 2];
        patches[p].neighbour_patch = pat[4 * p + 3];
        patches[p].phi_bc = PIC_BC_ZERO_GRADIENT;
        patches[p].phi_value = 0.0;
    }
    pic_mesh mesh{np, nf, ni, nc, npatch, points.data(), fo.data(), fp.data(), owner.data(), nbr.data(),
                  patches.data()};
    const int n = 4 * nc;
    pic_species sp{};
    sp.mass = 9.1093837015e-31;
    sp.charge = -1.602176634e-19;
    sp.n_equivalent = 1e5;
    sp.solve_movement = 1;
    sp.capacity = n;
    pic_models m{};
    m.particle_pusher = "Boris";
    m.field_weighting = "CellAverage";
    m.charge_distribution = "CellAverage";
    m.maxwell_solver = "none";
    m.coulomb_collision = "none";
    m.electron_neutral = "none";
    m.background_gas = "none";
    m.electron_species = -1;
    m.ion_species = -1;
    m.precision = PIC_PREC_FP64;
    m.seed = 1;
    pic_ctx* ctx = nullptr;
    int rc = pic_create(&mesh, &sp, 1, &m, &ctx);
    if (rc != PIC_OK) { std::fprintf(stderr, "pic_create failed (%d): %s\n", rc, pic_last_error(nullptr)); return 1; }
    std::vector<double> cc(3 * (size_t)nc), cv(nc);
    CHECK(pic_mesh_geometry(ctx, cc.data(), cv.data()));
    std::vector<double> x(3 * (size_t)n), U(3 * (size_t)n);
    std::vector<int32_t> cell(n);
    std::vector<uint32_t> id(n);
    const double h = box[0] / std::cbrt((double)nc);
    for (int i = 0; i < n; ++i) {
        const int c = i % nc;
        cell[i] = c;
        id[i] = (uint32_t)i;
        for (int d = 0; d < 3; ++d) {
            x[3 * i + d] = cc[3 * c + d] + 0.2 * h * std::sin(1.3 * i + 2.1 * d);
            U[3 * i + d] = 3.0e5 * std::cos(0.7 * i + 1.9 * d + 0.3);
        }
    }
    CHECK(pic_particles_set_positions(ctx, 0, n, x.data(), cell.data(), U.data(), id.data()));
    const double dt = 1.5 * h / 3.0e5;
    CHECK(pic_move(ctx, dt));
    std::vector<double> xe(3 * (size_t)n);
    CHECK(pic_particles_positions(ctx, 0, n, xe.data()));
    double worst = 0.0;
    for (int i = 0; i < n; ++i)
        for (int d = 0; d < 3; ++d) {
            double e = xe[3 * i + d] - (x[3 * i + d] + U[3 * i + d] * dt);
            e -= box[d] * std::round(e / box[d]);
            worst = std::fmax(worst, std::fabs(e) / box[d]);
        }
    if (!(worst < 1e-12)) { std::fprintf(stderr, "straight-line end positions off by %.3e\n", worst); return 1; }
    pic_step_stats st{};
    CHECK(pic_step(ctx, dt, &st));
    CHECK(pic_run(ctx, dt, 3, &st));
    if (st.n_parcels[0] != n || st.n_deleted != 0) {
        std::fprintf(stderr, "parcel count %lld deleted %lld\n", (long long)st.n_parcels[0], (long long)st.n_deleted);
        return 1;
    }
    pic_diag dg{};
    CHECK(pic_diag_reduce(ctx, &dg));
    if (dg.n_parcels[0] != n || std::fabs(dg.sum_w[0] - n * sp.n_equivalent) > 1e-9 * n * sp.n_equivalent) {
        std::fprintf(stderr, "diagnostics: %lld parcels, sum w %.6e\n", (long long)dg.n_parcels[0], dg.sum_w[0]);
        return 1;
    }
    std::vector<double> rho(nc);
    CHECK(pic_get_cell_field(ctx, PIC_FIELD_RHO, rho.data()));
    double q = 0.0;
    for (int c = 0; c < nc; ++c) q += rho[c] * cv[c];
    if (std::fabs(q - n * sp.charge * sp.n_equivalent) > 1e-9 * std::fabs(n * sp.charge * sp.n_equivalent)) {
        std::fprintf(stderr, "deposited charge %.6e\n", q);
        return 1;
    }
    std::printf("abi smoke ok: %d parcels, %d cells, straight-line error %.2e\n", n, nc, worst);
    pic_destroy(ctx);
    return 0;
}
