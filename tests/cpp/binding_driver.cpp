// A C++ caller of the C ABI in the pattern a maintainer's binding of the
// reference uses (member function -> one ABI call): the descriptor built
// from plain arrays in the reference's (i*ny + j)*nz + k layout (the
// material_cell.matrix hand-off), m uploaded from three component arrays,
// then main.cpp:243-483's loop body as member calls --
// time_marching_device_serial -> gf_time_advance, get_dm_RKs -> gf_mag_stage,
// update_m_RKs (no call: folded into the stage), get_H_static -> gf_demag --
// and get_averagem / copym_from_device at the end.  Errors are reported
// through gf_last_error, the create error through gf_last_error(nullptr).
//
// Usage:
//   binding_driver probe
//       creates a context and expects success (GPU) or a clean failure with
//       a message (no GPU); exit 0 either way unless the error path breaks.
//   binding_driver run PARAMS M_IN M_OUT
//       PARAMS: text "nx ny nz steps dx dt Ms gyro alpha K1 Aex Hz"; M_IN and
//       M_OUT: raw little-endian doubles mx|my|mz (3n).  Prints the average m.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <vector>

#include "goferro.h"

namespace {

struct Params {
    long nx = 4, ny = 4, nz = 4, steps = 1;
    double dx = 2e-9, dt = 1e-14, Ms = 1.4e5, gyro = 1.76e11, alpha = 0.01, K1 = -610., Aex = 3.7e-12, Hz = 0.;
};

gf_ctx* create(const Params& p, std::vector<uint32_t>& cell, std::vector<gf_material>& mats) {
    mats.assign(1, gf_material{});
    gf_material& g = mats[0];
    g.if_FM = 1;
    g.anisotropy_type = 1;
    g.Ms = p.Ms; g.gyro = p.gyro; g.FM_damping = p.alpha;
    g.K1 = p.K1; g.K2 = 0.; g.Aex = p.Aex; g.iDMI = 0.; g.B1 = 0.; g.B2 = 0.;
    const long n = p.nx * p.ny * p.nz;
    cell.assign(static_cast<size_t>(n), 1u);
    gf_desc d;
    std::memset(&d, 0, sizeof d);
    d.struct_size = sizeof(gf_desc);
    d.abi_version = GF_ABI_VERSION;
    d.nx = p.nx; d.ny = p.ny; d.nz = p.nz;
    d.dx = d.dy = d.dz = p.dx;
    d.periodic[0] = 1; d.periodic[1] = 1; d.periodic[2] = 0;
    d.flags = GF_FLAG_MAGNETOSTATICS;
    d.dt = p.dt;
    d.n_materials = 1;
    d.materials = mats.data();
    d.material_cell = cell.data();
    for (int a = 0; a < 3; a++) d.Tg2c[a][a] = 1.;
    d.Hext_stat[2] = p.Hz;
    gf_ctx* ctx = nullptr;
    if (gf_create(&d, &ctx) != GF_OK) {
        std::fprintf(stderr, "gf_create: %s\n", gf_last_error(nullptr));
        return nullptr;
    }
    return ctx;
}

int check(gf_ctx* ctx, gf_status st, const char* what) {
    if (st == GF_OK) return 0;
    std::fprintf(stderr, "%s failed (%d): %s\n", what, static_cast<int>(st), gf_last_error(ctx));
    return 1;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc >= 2 && std::strcmp(argv[1], "probe") == 0) {
        Params p;
        std::vector<uint32_t> cell;
        std::vector<gf_material> mats;
        gf_ctx* ctx = create(p, cell, mats);
        if (ctx) {
            std::printf("probe: context created\n");
            return gf_destroy(ctx) == GF_OK ? 0 : 3;
        }
        const char* msg = gf_last_error(nullptr);
        if (!msg || !*msg) {
            std::fprintf(stderr, "probe: create failed without a message\n");
            return 2;
        }
        std::printf("probe: create refused: %s\n", msg);
        return 0;
    }
    if (argc != 5 || std::strcmp(argv[1], "run") != 0) {
        std::fprintf(stderr, "usage: binding_driver probe | run PARAMS M_IN M_OUT\n");
        return 2;
    }
    Params p;
    {
        std::ifstream in(argv[2]);
        if (!(in >> p.nx >> p.ny >> p.nz >> p.steps >> p.dx >> p.dt >> p.Ms >> p.gyro >> p.alpha >> p.K1 >> p.Aex >>
              p.Hz)) {
            std::fprintf(stderr, "cannot read %s\n", argv[2]);
            return 2;
        }
    }
    const long n = p.nx * p.ny * p.nz;
    std::vector<double> mx(static_cast<size_t>(n)), my(mx), mz(mx);
    {
        std::ifstream in(argv[3], std::ios::binary);
        in.read(reinterpret_cast<char*>(mx.data()), static_cast<std::streamsize>(sizeof(double) * n));
        in.read(reinterpret_cast<char*>(my.data()), static_cast<std::streamsize>(sizeof(double) * n));
        in.read(reinterpret_cast<char*>(mz.data()), static_cast<std::streamsize>(sizeof(double) * n));
        if (!in) {
            std::fprintf(stderr, "cannot read %s\n", argv[3]);
            return 2;
        }
    }
    std::vector<uint32_t> cell;
    std::vector<gf_material> mats;
    gf_ctx* ctx = create(p, cell, mats);
    if (!ctx) return 1;
    // initialize_device: the three component arrays into one upload
    std::vector<double> buf(static_cast<size_t>(3 * n));
    std::copy(mx.begin(), mx.end(), buf.begin());
    std::copy(my.begin(), my.end(), buf.begin() + n);
    std::copy(mz.begin(), mz.end(), buf.begin() + 2 * n);
    int bad = check(ctx, gf_upload(ctx, GF_FIELD_M, buf.data(), buf.size()), "gf_upload");
    // the initial static field (main.cpp:113-143)
    if (!bad) bad = check(ctx, gf_demag(ctx), "get_H_static");
    for (long s = 0; s < p.steps && !bad; s++) {
        bad = check(ctx, gf_time_advance(ctx), "time_marching_device_serial");
        for (int stage = 1; stage <= 4 && !bad; stage++) {
            bad = check(ctx, gf_mag_stage(ctx, stage), "get_dm_RK");
            if (!bad) bad = check(ctx, gf_demag(ctx), "get_H_static");
        }
    }
    double avg[3] = {0., 0., 0.};
    if (!bad) bad = check(ctx, gf_average(ctx, GF_FIELD_M, avg), "get_averagem");
    if (!bad) bad = check(ctx, gf_download(ctx, GF_FIELD_M, buf.data(), buf.size()), "copym_from_device");
    // a stage out of order is refused with the expected order in the message
    if (!bad && gf_mag_stage(ctx, 3) != GF_ERR_STATE) {
        std::fprintf(stderr, "a stage out of order was accepted\n");
        bad = 1;
    }
    gf_destroy(ctx);
    if (bad) return 1;
    std::ofstream out(argv[4], std::ios::binary);
    out.write(reinterpret_cast<const char*>(buf.data()), static_cast<std::streamsize>(sizeof(double) * 3 * n));
    std::printf("average m %.17g %.17g %.17g\n", avg[0], avg[1], avg[2]);
    return out ? 0 : 1;
}
