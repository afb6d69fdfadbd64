// gf_api.hip — the C ABI (include/goferro.h): context set-up, transfers,
// the RK4 protocol and the drives.  Host side of the gfx950 library.
//
// Set-up mirrors magnetic_system::initialize_host / initialize_host_supple /
// initialize_device (magnetic_initialize.cpp:4-1129): interface masks become
// bits of a per-cell code, per-material constants are hoisted into a device
// table, the Newell tensor spectra are built once.  The step protocol
// mirrors main.cpp:243-483 for the magnetic physics.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "gf_internal.h"

namespace gf {

gf_status set_error(gf_ctx* c, gf_status code, const std::string& msg) {
    if (c) c->err = msg;
    return code;
}

gf_status hip_check(gf_ctx* c, hipError_t e, const char* what) {
    if (e == hipSuccess) return GF_OK;
    return set_error(c, GF_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

static hipEvent_t take_event(gf_ctx* c) {
    if (!c->event_pool.empty()) {
        hipEvent_t e = c->event_pool.back();
        c->event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

void timer_begin(gf_ctx* c, int cls, hipEvent_t* ev) {
    (void)cls;
    *ev = nullptr;
    if (!c->timing) return;
    *ev = take_event(c);
    if (*ev) hipEventRecord(*ev, c->stream);
}

void timer_end(gf_ctx* c, int cls, hipEvent_t start) {
    if (!c->timing || !start) return;
    hipEvent_t stop = take_event(c);
    if (!stop) return;
    hipEventRecord(stop, c->stream);
    c->timers.push_back(TimerSlot{start, stop, cls});
}

static void drain_timers(gf_ctx* c) {
    if (c->timers.empty()) return;
    hipStreamSynchronize(c->stream);
    for (auto& t : c->timers) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, t.start, t.stop) == hipSuccess) {
            c->kernel_ms[t.cls] += ms;
            c->kernel_launches[t.cls] += 1;
        }
        c->event_pool.push_back(t.start);
        c->event_pool.push_back(t.stop);
    }
    c->timers.clear();
}

}  // namespace gf

using namespace gf;

namespace {

thread_local std::string g_create_error;

bool is_fm(const gf_ctx* c, int64_t id) {
    const uint32_t t = c->host_code[id] & kMatMask;
    return t && c->host_mats[t - 1].if_FM;
}
bool is_fe(const gf_ctx* c, int64_t id) {
    const uint32_t t = c->host_code[id] & kMatMask;
    return t && c->host_mats[t - 1].if_FE;
}
bool is_afm(const gf_ctx* c, int64_t id) {
    const uint32_t t = c->host_code[id] & kMatMask;
    return t && c->host_mats[t - 1].if_AFM;
}

// Face masks of one magnetic kind as code bits (magnetic_initialize.cpp:
// 136-379 for FM, 635-~870 for AFM): a face is a surface when exactly one
// side is of the kind, or, at a non-periodic wall, when the cell is.
void build_surface_bits(gf_ctx* c, bool (*kind)(const gf_ctx*, int64_t), int shift) {
    const int64_t nx = c->nx, ny = c->ny, nz = c->nz;
    auto id = [&](int64_t i, int64_t j, int64_t k) { return (i * ny + j) * nz + k; };
    for (int64_t i = 0; i < nx; i++)
        for (int64_t j = 0; j < ny; j++)
            for (int64_t k = 0; k < nz; k++) {
                const int64_t me = id(i, j, k);
                if (!kind(c, me)) continue;
                uint32_t bits = 0;
                // neighbour on each side, or the wall rule
                const bool per[3] = {c->per[0] != 0, c->per[1] != 0, c->per[2] != 0};
                const int64_t lo[3] = {i, j, k};
                const int64_t n3[3] = {nx, ny, nz};
                for (int ax = 0; ax < 3; ax++) {
                    for (int side = 0; side < 2; side++) {   // 0 = backward, 1 = forward
                        int64_t q[3] = {lo[0], lo[1], lo[2]};
                        const bool at_wall = side == 0 ? lo[ax] == 0 : lo[ax] == n3[ax] - 1;
                        bool surf;
                        if (at_wall && !per[ax]) {
                            surf = true;   // wall face next to a cell of the kind
                        } else {
                            q[ax] = side == 0 ? (lo[ax] == 0 ? n3[ax] - 1 : lo[ax] - 1)
                                              : (lo[ax] == n3[ax] - 1 ? 0 : lo[ax] + 1);
                            surf = !kind(c, id(q[0], q[1], q[2]));
                        }
                        if (surf) bits |= 1u << (shift + 2 * ax + side);
                    }
                }
                c->host_code[me] |= bits;
            }
}

DevMaterial make_dev_material(const gf_material& m, double dt) {
    DevMaterial d{};
    d.is_fm = m.if_FM ? 1 : 0;
    d.is_afm = m.if_AFM ? 1 : 0;
    d.anis_type = m.anisotropy_type;
    d.Ms = m.Ms;
    if (m.if_FM) {
        const double Ms = m.Ms;
        d.K1n = m.K1 / kMu0 / Ms;
        d.K2n = m.K2 / kMu0 / Ms;
        d.B1n = m.B1 / kMu0 / Ms;
        d.B2n = m.B2 / kMu0 / Ms;
        d.Aex = m.Aex;
        d.Aex_self = 2. * m.Aex * m.Aex / (m.Aex + m.Aex);
        d.dmi_slope = m.iDMI / 2. / m.Aex;
        d.exch = 2. / kMu0 / Ms;
        d.dmi_pos = 2. * m.iDMI / kMu0 / Ms;
        d.dmi_neg = -2. * m.iDMI / kMu0 / Ms;
        d.precess = -m.gyro / (1 + m.FM_damping * m.FM_damping) * dt;
        d.alpha_precess = m.FM_damping * d.precess;
    } else {
        d.Aex = m.Aex;
    }
    const double ms[2] = {m.Ms_AFM1, m.Ms_AFM2};
    const double k1[2] = {m.K1_AFM1, m.K1_AFM2};
    const double b1[2] = {m.B1_AFM1, m.B1_AFM2};
    const double b2[2] = {m.B2_AFM1, m.B2_AFM2};
    const double ae[2] = {m.Aex_AFM1, m.Aex_AFM2};
    const double gy[2] = {m.gyro_AFM1, m.gyro_AFM2};
    const double al[2] = {m.damping_AFM1, m.damping_AFM2};
    for (int s = 0; s < 2; s++) {
        d.aMs[s] = ms[s];
        d.aAex[s] = ae[s];
        if (!m.if_AFM) continue;
        d.aK1n[s] = k1[s] / kMu0 / ms[s];
        d.aB1n[s] = b1[s] / kMu0 / ms[s];
        d.aB2n[s] = b2[s] / kMu0 / ms[s];
        d.aJ[s] = -1. * m.J_AFM / kMu0 / ms[s];
        d.aAex_self[s] = 2. * ae[s] * ae[s] / (ae[s] + ae[s]);
        d.aExch[s] = 2. / kMu0 / ms[s];
        d.aPrecess[s] = -gy[s] / (1. + al[s] * al[s]) * dt;
        d.aAlphaPrecess[s] = al[s] * d.aPrecess[s];
    }
    for (int a = 0; a < 3; a++) d.u[a] = m.uniaxis_AFM[a];
    d.is_fe = m.if_FE ? 1 : 0;
    if (m.if_FE) {
        d.fe_mass = -1. * dt / m.FE_mass;
        d.fe_damping = d.fe_mass * m.FE_damping;
        d.a1 = m.a1; d.a11 = m.a11; d.a12 = m.a12; d.a111 = m.a111; d.a112 = m.a112; d.a123 = m.a123;
        d.a1111 = m.a1111; d.a1112 = m.a1112; d.a1122 = m.a1122; d.a1123 = m.a1123;
        d.G11 = m.G11; d.tv1 = m.tv1; d.tv2 = m.tv2; d.tv3 = m.tv3;
    }
    // elastic_dynamic_solver.cpp:1306-1309 and the eigenstrain constants (:4-104)
    std::memcpy(d.cij, m.cijkl_glb, sizeof(d.cij));
    d.el_rho = dt / m.density;
    d.el_mass = -dt * m.elast_mass_damping;
    d.el_stiff = m.elast_stiff_damping / m.density;
    d.Q11 = m.Q11; d.Q12 = m.Q12; d.Q44 = m.Q44;
    d.l100 = m.lamda100; d.l111 = m.lamda111;
    d.l100a = m.lamda100_AFM1; d.l111a = m.lamda111_AFM1;
    d.l100b = m.lamda100_AFM2; d.l111b = m.lamda111_AFM2;
    // EMdynamic_solver.cpp:392-499 and get_dJp_RK* (:4318-4451)
    d.em_er = m.r_permittivity;
    d.em_cond = m.conductivity;
    d.cn[0] = m.comp_n1; d.cn[1] = m.comp_n2; d.cn[2] = m.comp_n3;
    d.wp[0] = m.omega_plasma_n1; d.wp[1] = m.omega_plasma_n2; d.wp[2] = m.omega_plasma_n3;
    d.re[0] = 1. / m.tao_e_n1; d.re[1] = 1. / m.tao_e_n2; d.re[2] = 1. / m.tao_e_n3;
    return d;
}

gf_status alloc_n(gf_ctx* c, double** p, int64_t count, const char* what) {
    gf_status st = hip_check(c, hipMalloc(p, sizeof(double) * static_cast<size_t>(count)), what);
    if (st) return st;
    return hip_check(c, hipMemsetAsync(*p, 0, sizeof(double) * static_cast<size_t>(count), c->stream), what);
}

// Elastic set-up checks: the grids on which the reference's indexing
// differs from its intent in a way that cannot be reproduced, and the
// planes the absorbing boundaries read (elastic_dynamic_solver.cpp:1247,
// elastodynamics_AbsorbingBoundary.cpp:92-230).
std::string elastic_unsupported(const gf_ctx* c) {
    if (c->per[2] && c->ny != c->nz)
        return "elastodynamics with periodic z needs ny == nz (the reference wraps the forward z neighbour at "
               "k == ny-1)";
    const bool mur = c->flags & GF_FLAG_1D_ABC;
    if (!mur && c->nz < 5) return "the Liao boundary needs nz >= 5 (it reads planes k = 1..4)";
    if (mur && c->nz < 2) return "the Mur boundary needs nz >= 2";
    if (c->ny < c->nz) {
        // at k = ny-1 the forward z stress reads cell (i, j, 0) instead of
        // (i, j, ny); that cell must hold a material
        const int64_t k = c->ny - 1;
        for (int64_t i = 0; i < c->nx; i++)
            for (int64_t j = 0; j < c->ny; j++) {
                const int64_t col = (i * c->ny + j) * c->nz;
                const bool joined = (c->host_code[col + k] & kMatMask) && (c->host_code[col + k + 1] & kMatMask);
                if (joined && !(c->host_code[col] & kMatMask))
                    return "elastodynamics: the reference reads the material of cell (i, j, 0) at k = ny-1, which "
                           "is vacuum at i=" + std::to_string(i) + " j=" + std::to_string(j);
            }
    }
    return "";
}

// Sound speeds of the first material met on the planes k = 0 and k = nz-1
// (elastic_initialize.cpp:4-83), scanning i-major.
void elastic_speeds(gf_ctx* c) {
    auto scan = [&](int64_t k, double& vl, double& vt) {
        double c11 = 0., c44 = 0., rho = 1.e-8;
        for (int64_t i = 0; i < c->nx; i++)
            for (int64_t j = 0; j < c->ny; j++) {
                const uint32_t t = c->host_code[(i * c->ny + j) * c->nz + k] & kMatMask;
                if (t) {
                    const gf_material& m = c->host_mats[t - 1];
                    c11 = m.c11; c44 = m.c44; rho = m.density;
                    vl = std::sqrt((c11 / rho));
                    vt = std::sqrt((c44 / rho));
                    return;
                }
            }
        vl = std::sqrt((c11 / rho));
        vt = std::sqrt((c44 / rho));
    };
    scan(0, c->vl_first, c->vt_first);
    scan(c->nz - 1, c->vl_last, c->vt_last);
}

gf_status alloc3(gf_ctx* c, double* p[3], const char* what) {
    for (int a = 0; a < 3; a++) {
        gf_status st = hip_check(c, hipMalloc(&p[a], sizeof(double) * c->n), what);
        if (st) return st;
        st = hip_check(c, hipMemsetAsync(p[a], 0, sizeof(double) * c->n, c->stream), what);
        if (st) return st;
    }
    return GF_OK;
}

void free3(double* p[3]) {
    for (int a = 0; a < 3; a++) {
        if (p[a]) hipFree(p[a]);
        p[a] = nullptr;
    }
}

// magnetic_solver.cpp:1834-1861 (drive time per stage)
void stage_drive(gf_ctx* c, int stage, double out[3]) {
    if (c->flags & GF_FLAG_PRESCRIBE_HEXT) {
        for (int a = 0; a < 3; a++) out[a] = c->hext[a];
        return;
    }
    const double t = stage == 1 ? c->time - c->dt : (stage == 4 ? c->time : c->time - c->dt * 0.5);
    for (int a = 0; a < 3; a++) out[a] = c->hext_stat[a] + c->hext_altr[a] * std::sin(2. * kPi * c->h_altr_freq[a] * t);
}

double** fm_store(gf_ctx* c, int sel) { return sel == 0 ? c->m : (sel == 1 ? c->st_a : c->st_b); }
double** afm1_store(gf_ctx* c, int sel) { return sel == 0 ? c->m1 : (sel == 1 ? c->st1_a : c->st1_b); }
double** afm2_store(gf_ctx* c, int sel) { return sel == 0 ? c->m2 : (sel == 1 ? c->st2_a : c->st2_b); }
double** fe_pstore(gf_ctx* c, int sel) { return sel == 0 ? c->p : (sel == 1 ? c->p_a : c->p_b); }
double** fe_qstore(gf_ctx* c, int sel) { return sel == 0 ? c->q : (sel == 1 ? c->q_a : c->q_b); }

// The static strain of the current magnet and ferroelectric stores.
gf_status elastostatic_static(gf_ctx* c) {
    double** fm = c->any_fm ? fm_store(c, c->mag_store_sel) : nullptr;
    double** s1 = c->any_afm ? afm1_store(c, c->afm_store_sel) : nullptr;
    double** s2 = c->any_afm ? afm2_store(c, c->afm_store_sel) : nullptr;
    double** ps = c->any_fe ? fe_pstore(c, c->fe_store_sel) : nullptr;
    return elastostatic_run(c, fm, s1, s2, ps);
}

// The magnetostatic part of gf_demag: H of the current magnet stores.
gf_status mag_statics(gf_ctx* c) {
    if (!(c->any_fm || c->any_afm)) return GF_OK;
    double* zero[3] = {nullptr, nullptr, nullptr};
    double** fm = c->any_fm ? fm_store(c, c->mag_store_sel) : zero;
    double** s1 = c->any_afm ? afm1_store(c, c->afm_store_sel) : zero;
    double** s2 = c->any_afm ? afm2_store(c, c->afm_store_sel) : zero;
    if (c->flags & GF_FLAG_MAG_1D) return hip_check(c, launch_demag_local(c, fm, s1, s2, 1), "demag 1d");
    if (c->flags & GF_FLAG_DEMAG_FACTOR) return hip_check(c, launch_demag_local(c, fm, s1, s2, 0), "demag factor");
    // no magnetostatics asked for (the static fields of other physics may be)
    if (!(c->flags & GF_FLAG_MAGNETOSTATICS)) return GF_OK;
    if (!c->demag.ready) return set_error(c, GF_ERR_STATE, "FFT magnetostatics was not set up (flags)");
    if (c->demag.fused) return demag_run_fused(c, fm);
    return demag_run(c, fm, s1, s2);
}

// Flexoelectric constants of one material (read.cpp:181-199), zero for a
// non-ferroelectric as check_material leaves them (global.cpp:339-358).
FlexoMat make_flexo(const gf_material& m) {
    FlexoMat f{};
    if (!m.if_FE) return f;
    const double c11 = m.c11, c12 = m.c12, c44 = m.c44, Q11 = m.Q11, Q12 = m.Q12, Q44 = m.Q44;
    f.F11 = ((c11 + c12) * m.f11 - 2. * c12 * m.f12) / (c11 + 2. * c12) / (c11 - c12);
    f.F12 = (c11 * m.f12 - c12 * m.f11) / (c11 + 2. * c12) / (c11 - c12);
    f.F44 = m.f44 / c44;
    f.tv4 = 2. * (c11 * (Q11 * f.F11 + 2. * Q12 * f.F12) + 2. * c12 * (Q11 * f.F12 + Q12 * f.F11 + Q12 * f.F12));
    f.tv5 = 2. * (c11 * (Q11 * f.F12 + Q12 * f.F11 + Q12 * f.F12) +
                  c12 * (Q11 * f.F11 + Q11 * f.F12 + Q12 * f.F11 + 3. * Q12 * f.F12));
    f.tv6 = 2. * c44 * Q44 * f.F44;
    f.tv7 = c11 * (f.F11 * f.F11 + 2. * f.F12 * f.F12) + 2. * c12 * (f.F12 * f.F12 + 2. * f.F11 * f.F12);
    f.tv8 = c44 * f.F44 * f.F44;
    f.f11 = m.f11; f.f12 = m.f12; f.f44 = m.f44;
    return f;
}

gf_status flexo_setup(gf_ctx* c) {
    if (!c->any_fe) return set_error(c, GF_ERR_UNSUPPORTED, "flexoelectricity needs a ferroelectric material");
    std::vector<FlexoMat> tab;
    for (const auto& m : c->host_mats) {
        if (m.if_FE && (m.c11 == m.c12 || m.c11 + 2. * m.c12 == 0. || m.c44 == 0.))
            return set_error(c, GF_ERR_ARG, "flexoelectricity needs c11 != c12, c11 + 2 c12 != 0 and c44 != 0");
        tab.push_back(make_flexo(m));
    }
    gf_status st = hip_check(c, hipMalloc(&c->d_flexo, sizeof(FlexoMat) * tab.size()), "hipMalloc flexo table");
    if (!st)
        st = hip_check(c, hipMemcpy(c->d_flexo, tab.data(), sizeof(FlexoMat) * tab.size(), hipMemcpyHostToDevice),
                       "copy flexo table");
    if (!st) st = alloc3(c, c->flexo_force, "hipMalloc flexo force");
    for (int a = 0; a < 3 && !st; a++)
        st = hip_check(c, hipMemset(c->flexo_force[a], 0, sizeof(double) * c->n), "clear flexo force");
    return st;
}

// Spin-pumping planes (global.cpp:169-235) and their coefficients
// (magnetic_solver.cpp:245-259).  The layers are walked bottom to top from
// k = 0 (gf_desc.nz_layer, else the runs of equal ids along z of column
// (0, 0)); a layer whose material at column (0, 0) pumps, with a magnet on
// exactly one side, carries J_ISHE on each plane k, with the source plane
// next to it and the layer thickness t: coefficient theta_SH e / (2 pi)
// g_mix times +-sinh((t - d) / l_sd) / sinh(t / l_sd), d = (|k - sz| -
// 1/2) dz, + when the source lies below.  The magnet stages then also write
// their increments, which the current reads.
gf_status spin_setup(gf_ctx* c, const gf_desc* d) {
    const int64_t nz = c->nz;
    std::vector<int64_t> lay;
    if (d->n_layers > 0) {
        if (!d->nz_layer) return set_error(c, GF_ERR_ARG, "nz_layer is null");
        int64_t tot = 0;
        for (int i = 0; i < d->n_layers; i++) {
            if (d->nz_layer[i] < 1) return set_error(c, GF_ERR_ARG, "nz_layer entries must be positive");
            tot += d->nz_layer[i];
            lay.push_back(d->nz_layer[i]);
        }
        if (tot > nz) return set_error(c, GF_ERR_ARG, "the layers hold more planes than nz");
    } else {
        for (int64_t k = 0; k < nz;) {
            int64_t e = k + 1;
            while (e < nz && c->host_code[e] == c->host_code[k]) e++;
            lay.push_back(e - k);
            k = e;
        }
    }
    auto id_of = [&](int64_t k) { return c->host_code[k] & kMatMask; };   // column (0, 0)
    auto magnet = [&](int64_t k) {
        const uint32_t t = id_of(k);
        return t && (c->host_mats[t - 1].if_FM || c->host_mats[t - 1].if_AFM);
    };
    std::vector<int32_t> src(static_cast<size_t>(nz), -1);
    std::vector<double> thick(static_cast<size_t>(nz), 0.), coef(static_cast<size_t>(nz), 0.);
    int64_t zf = -1;
    for (int64_t len : lay) {
        const int64_t zi = zf + 1;
        zf = zi + len - 1;
        const uint32_t t = id_of(zi);
        if (!t || !c->host_mats[t - 1].if_spin_pump) continue;
        const bool si = zi != 0 && magnet(zi - 1);
        const bool sf = zf != nz - 1 && magnet(zf + 1);
        int64_t sz = -1;
        if (si && !sf) sz = zi - 1;
        else if (!si && sf) sz = zf + 1;
        if (sz < 0) continue;
        for (int64_t k = zi; k <= zf; k++) {
            src[static_cast<size_t>(k)] = static_cast<int32_t>(sz);
            thick[static_cast<size_t>(k)] = static_cast<double>(len) * c->dz;
        }
    }
    for (int64_t k = 0; k < nz; k++) {
        const size_t kk = static_cast<size_t>(k);
        if (src[kk] < 0 || !id_of(k)) { src[kk] = -1; continue; }
        const gf_material& m = c->host_mats[id_of(k) - 1];
        const double cjs = m.spin_hall_angle * kUnitE / 2. / kPi * m.spin_mix_cond;
        const double dist = (std::fabs(static_cast<double>(k - src[kk])) - 0.5) * c->dz;
        double cl = std::sinh((thick[kk] - dist) / m.spin_diffus_length) / std::sinh(thick[kk] / m.spin_diffus_length);
        if (!(k > src[kk])) cl = -1. * cl;
        coef[kk] = cjs * cl;
    }
    c->spin_on = true;
    gf_status st = alloc3(c, c->jis, "hipMalloc J_ISHE");
    for (int a = 0; a < 3 && !st; a++)
        st = hip_check(c, hipMemset(c->jis[a], 0, sizeof(double) * c->n), "clear J_ISHE");
    if (!st) st = hip_check(c, hipMalloc(&c->spin_src, sizeof(int32_t) * src.size()), "hipMalloc spin planes");
    if (!st) st = hip_check(c, hipMalloc(&c->spin_coef, sizeof(double) * coef.size()), "hipMalloc spin planes");
    if (!st)
        st = hip_check(c, hipMemcpy(c->spin_src, src.data(), sizeof(int32_t) * src.size(), hipMemcpyHostToDevice),
                       "copy spin planes");
    if (!st)
        st = hip_check(c, hipMemcpy(c->spin_coef, coef.data(), sizeof(double) * coef.size(), hipMemcpyHostToDevice),
                       "copy spin planes");
    if (!st && c->any_fm && !c->em_dm[0]) st = alloc3(c, c->em_dm, "hipMalloc dm");
    if (!st && c->any_afm && !c->em_dm1[0]) st = alloc3(c, c->em_dm1, "hipMalloc dm");
    if (!st && c->any_afm && !c->em_dm2[0]) st = alloc3(c, c->em_dm2, "hipMalloc dm");
    // the step-boundary current reads the source planes of the state before
    // the stage-4 update: the FM one only without an AFM (with both magnets
    // the AFM update's call sees the FM state already updated)
    std::vector<int32_t> is_src(static_cast<size_t>(nz), 0);
    for (int64_t k = 0; k < nz; k++)
        if (src[static_cast<size_t>(k)] >= 0) is_src[static_cast<size_t>(src[static_cast<size_t>(k)])] = 1;
    if (!st) st = hip_check(c, hipMalloc(&c->spin_is_src, sizeof(int32_t) * is_src.size()), "hipMalloc spin planes");
    if (!st)
        st = hip_check(c, hipMemcpy(c->spin_is_src, is_src.data(), sizeof(int32_t) * is_src.size(),
                                    hipMemcpyHostToDevice), "copy spin planes");
    if (!st && c->any_fm && !c->any_afm) st = alloc3(c, c->spin_m0, "hipMalloc spin state");
    if (!st && c->any_afm) st = alloc3(c, c->spin_m10, "hipMalloc spin state");
    if (!st && c->any_afm) st = alloc3(c, c->spin_m20, "hipMalloc spin state");
    return st;
}

// The source planes of a magnet's state before its stage-4 kernel updates
// it in place (FM: which = 0, AFM: which = 1).
gf_status spin_save(gf_ctx* c, int which) {
    SpinSaveArgs a{};
    if (which == 0) {
        for (int i = 0; i < 3; i++) { a.src[i] = c->m[i]; a.dst[i] = c->spin_m0[i]; }
        a.ncomp = 3;
    } else {
        for (int i = 0; i < 3; i++) {
            a.src[i] = c->m1[i]; a.dst[i] = c->spin_m10[i];
            a.src[3 + i] = c->m2[i]; a.dst[3 + i] = c->spin_m20[i];
        }
        a.ncomp = 6;
    }
    a.is_src = c->spin_is_src;
    a.nx = c->nx; a.ny = c->ny; a.nz = c->nz;
    hipEvent_t ev = nullptr;
    timer_begin(c, GF_K_OTHER, &ev);
    gf_status st = hip_check(c, launch_spin_save(a, c->stream), "spin pumping");
    timer_end(c, GF_K_OTHER, ev);
    return st;
}

// J_ISHE of the stage that just ran its magnet stages (get_J_ISHE_RK*).
gf_status spin_pump(gf_ctx* c) {
    SpinArgs a{};
    double** fs = c->any_fm ? fm_store(c, c->mag_in_sel) : nullptr;
    double** s1 = c->any_afm ? afm1_store(c, c->afm_in_sel) : nullptr;
    double** s2 = c->any_afm ? afm2_store(c, c->afm_in_sel) : nullptr;
    for (int i = 0; i < 3; i++) {
        a.fs[i] = fs ? fs[i] : nullptr;
        a.fd[i] = fs ? c->em_dm[i] : nullptr;
        a.s1[i] = s1 ? s1[i] : nullptr;
        a.d1[i] = s1 ? c->em_dm1[i] : nullptr;
        a.s2[i] = s2 ? s2[i] : nullptr;
        a.d2[i] = s2 ? c->em_dm2[i] : nullptr;
    }
    a.src = c->spin_src;
    a.coef = c->spin_coef;
    a.j[0] = c->jis[0];
    a.j[1] = c->jis[1];
    a.nx = c->nx; a.ny = c->ny; a.nz = c->nz; a.dt = c->dt;
    hipEvent_t ev = nullptr;
    timer_begin(c, GF_K_OTHER, &ev);
    gf_status st = hip_check(c, launch_spin_pump(a, c->stream), "spin pumping");
    timer_end(c, GF_K_OTHER, ev);
    return st;
}

// get_J_ISHE at the step boundary (magnetic_solver.cpp:1460-1511, called by
// update_m and update_m_AFM before m += dm, :1593-1597 and
// magnetic_AFM_solver.cpp:1801-1803), after every stage-4 call that reads
// the stage-4 current: the states before the update (the saved source
// planes; the updated FM state when an AFM is present, whose call comes
// last) and the combined increments acc + k4/6.
gf_status spin_pump_step(gf_ctx* c) {
    SpinArgs a{};
    for (int i = 0; i < 3; i++) {
        if (c->any_fm) {
            a.fs[i] = c->any_afm ? c->m[i] : c->spin_m0[i];
            a.fd[i] = c->em_dm[i];
            a.fa[i] = c->acc[i];
        }
        if (c->any_afm) {
            a.s1[i] = c->spin_m10[i]; a.d1[i] = c->em_dm1[i]; a.a1[i] = c->acc1[i];
            a.s2[i] = c->spin_m20[i]; a.d2[i] = c->em_dm2[i]; a.a2[i] = c->acc2[i];
        }
    }
    a.src = c->spin_src;
    a.coef = c->spin_coef;
    a.j[0] = c->jis[0];
    a.j[1] = c->jis[1];
    a.nx = c->nx; a.ny = c->ny; a.nz = c->nz; a.dt = c->dt;
    hipEvent_t ev = nullptr;
    timer_begin(c, GF_K_OTHER, &ev);
    gf_status st = hip_check(c, launch_spin_pump(a, c->stream), "spin pumping");
    timer_end(c, GF_K_OTHER, ev);
    return st;
}

// statics after an update phase (main.cpp:299-306 and the 1D field inside
// update_m*, magnetic_solver.cpp:1533-1538, magnetic_AFM_solver.cpp:1746-1751)
gf_status statics_after_update(gf_ctx* c) {
    const bool em = c->flags & GF_FLAG_EMDYNAMIC;
    const bool ms = c->flags & GF_FLAG_MAGNETOSTATICS;
    const bool one_d = c->flags & GF_FLAG_MAG_1D;
    if (em) return GF_OK;
    if ((c->flags & GF_FLAG_ELECTROSTATICS) && c->any_fe) {
        // get_E_static after the update phase (main.cpp:299-302)
        if (gf_status st = elec_run(c, fe_pstore(c, c->fe_store_sel))) return st;
    }
    if (one_d) {
        if ((c->any_fm && ms) || c->any_afm) return mag_statics(c);
        return GF_OK;
    }
    if ((c->any_fm || c->any_afm) && ms) return mag_statics(c);
    return GF_OK;
}

// Face gradients and strain of the current u store.
gf_status elastic_strain(gf_ctx* c) {
    ElStrainArgs a{};
    a.code = c->d_code;
    for (int i = 0; i < 3; i++) {
        a.us[i] = c->us[i]; a.gyz[i] = c->gyz[i]; a.gxz[i] = c->gxz[i]; a.gxy[i] = c->gxy[i];
    }
    for (int i = 0; i < 6; i++) { a.eg[i] = c->eg[i]; a.ec[i] = c->ec[i]; }
    a.nx = c->nx; a.ny = c->ny; a.nz = c->nz; a.dx = c->dx; a.dy = c->dy; a.dz = c->dz;
    a.idx = 1. / c->dx; a.idy = 1. / c->dy; a.idz = 1. / c->dz;
    for (int i = 0; i < 3; i++) a.per[i] = c->per[i];
    a.rotate = (c->flags & GF_FLAG_ORIENT_ROTATE) ? 1 : 0;
    std::memcpy(a.Tg2c, c->Tg2c, sizeof(a.Tg2c));
    hipEvent_t ev = nullptr;
    timer_begin(c, GF_K_ELASTIC_STRAIN, &ev);
    gf_status st = hip_check(c, launch_el_strain(a, c->stream), "elastic strain");
    timer_end(c, GF_K_ELASTIC_STRAIN, ev);
    return st;
}

// Electrodynamics set-up checks that need no device: the source plane and
// the current box must lie inside the grid.
std::string em_unsupported(const gf_ctx* c, const gf_desc* d) {
    if ((d->flags & GF_FLAG_PLANE_EM) && (d->planeEM_source_z < 0 || d->planeEM_source_z >= c->nz))
        return "planeEM_source_z outside the grid";
    if (d->flags & GF_FLAG_JF_INPUT) {
        const int64_t dims[3] = {c->nx, c->ny, c->nz};
        for (int a = 0; a < 3; a++)
            if (d->Jfin_lo[a] < 0 || d->Jfin_hi[a] < d->Jfin_lo[a] || d->Jfin_hi[a] >= dims[a])
                return "the Jf input box lies outside the grid";
    }
    return "";
}

// EMdynamic_initialize.cpp:3-156: Mur coefficients from the permittivity of
// the first material on the bottom and top planes, the PEC edges
// (initialize_PEC_edges, :158-366), the source tables, and whether the Jp
// currents can leave zero (a plasma frequency, a non-finite 1/tau, or an
// uploaded Jp).
gf_status em_setup(gf_ctx* c, const gf_desc* d) {
    const int64_t nx = c->nx, ny = c->ny, nz = c->nz, n = c->n;
    c->em_on = true;
    c->em_periodic_all = c->per[0] && c->per[1] && c->per[2];
    c->nE[0] = nx * (ny + 1) * (nz + 1); c->nE[1] = (nx + 1) * ny * (nz + 1); c->nE[2] = (nx + 1) * (ny + 1) * nz;
    c->nH[0] = (nx + 1) * ny * nz; c->nH[1] = nx * (ny + 1) * nz; c->nH[2] = nx * ny * (nz + 1);
    gf_status st = GF_OK;
    for (int a = 0; a < 3 && !st; a++) {
        double** eb[] = {&c->Eb[a], &c->Ea[a], &c->Ebb[a], &c->Eacc[a], &c->Et1[a]};
        for (double** p : eb)
            if (!st) st = alloc_n(c, p, c->nE[a], "hipMalloc E");
        double** hb[] = {&c->Hb[a], &c->Ha[a], &c->Hbb[a], &c->Hacc[a]};
        for (double** p : hb)
            if (!st) st = alloc_n(c, p, c->nH[a], "hipMalloc H");
    }
    if (!st) st = alloc3(c, c->Ecell, "hipMalloc E cell");
    if (!st) st = alloc3(c, c->Hcell, "hipMalloc H cell");
    for (int q = 0; q < 9 && !st; q++) {
        st = alloc_n(c, &c->Jp[q], n, "hipMalloc Jp");
        if (!st) st = alloc_n(c, &c->Jps[q], n, "hipMalloc Jp");
        if (!st) st = alloc_n(c, &c->Jacc[q], n, "hipMalloc Jp");
    }
    if (!st && (c->flags & GF_FLAG_EM_FROM_M)) {
        if (c->any_fm) st = alloc3(c, c->em_dm, "hipMalloc dm");
        if (!st && c->any_afm) st = alloc3(c, c->em_dm1, "hipMalloc dm");
        if (!st && c->any_afm) st = alloc3(c, c->em_dm2, "hipMalloc dm");
    }
    if (st) return st;
    double er_bot = 1., er_top = 1.;
    auto first = [&](int64_t k, double& er) {
        for (int64_t i = 0; i < nx; i++)
            for (int64_t j = 0; j < ny; j++) {
                const uint32_t t = c->host_code[(i * ny + j) * nz + k] & kMatMask;
                if (t) { er = std::sqrt(c->host_mats[t - 1].r_permittivity); return; }
            }
    };
    first(0, er_bot);
    first(nz - 1, er_top);
    const double dt = c->dt;
    c->cf_top_half = (2. * c->dz - kCLight / er_top * dt) / (2. * c->dz + kCLight / er_top * dt);
    c->cf_bot_half = (2. * c->dz - kCLight / er_bot * dt) / (2. * c->dz + kCLight / er_bot * dt);
    c->cf_x_half = (2. * c->dx - kCLight * dt) / (2. * c->dx + kCLight * dt);
    c->cf_y_half = (2. * c->dy - kCLight * dt) / (2. * c->dy + kCLight * dt);
    c->cf_top = (c->dz - kCLight / er_top * dt) / (c->dz + kCLight / er_top * dt);
    c->cf_bot = (c->dz - kCLight / er_bot * dt) / (c->dz + kCLight / er_bot * dt);
    c->cf_x = (c->dx - kCLight * dt) / (c->dx + kCLight * dt);
    c->cf_y = (c->dy - kCLight * dt) / (c->dy + kCLight * dt);
    for (int a = 0; a < 3; a++) { c->plane_E[a] = d->planeEM_E[a]; c->plane_f[a] = d->planeEM_E_freq[a]; }
    c->plane_z = d->planeEM_source_z;
    c->jf_type = d->Jf_input_type;
    const int ch = d->Jf_input_component;
    c->jf_comp = ch == 'x' ? 0 : (ch == 'y' ? 1 : (ch == 'z' ? 2 : -1));
    c->jf_cycles = d->num_Jf_cycles;
    for (int a = 0; a < 3; a++) { c->jf_lo[a] = d->Jfin_lo[a]; c->jf_hi[a] = d->Jfin_hi[a]; }
    c->jf_amp = d->Jf_input_amp; c->jf_freq = d->Jf_input_freq; c->jf_sigma = d->Jf_input_sigma;
    if ((c->flags & GF_FLAG_JF_INPUT) && (c->jf_type == 1 || c->jf_type == 2)) {
        // sin / cos of the spatial phase per plane of the box (:3902-3903)
        const int64_t jnz = static_cast<int64_t>(c->jf_hi[2]) - c->jf_lo[2] + 1;
        std::vector<double> sp(static_cast<size_t>(jnz));
        for (int64_t k = 0; k < jnz; k++) {
            const double sv = kPi / (static_cast<double>(jnz) * c->dz) * (static_cast<double>(k) + 0.5) * c->dz;
            sp[static_cast<size_t>(k)] = c->jf_type == 1 ? std::sin(sv) : std::cos(sv);
        }
        st = hip_check(c, hipMalloc(&c->jf_spatial, sizeof(double) * sp.size()), "hipMalloc Jf table");
        if (!st)
            st = hip_check(c, hipMemcpy(c->jf_spatial, sp.data(), sizeof(double) * sp.size(), hipMemcpyHostToDevice),
                           "copy Jf table");
        if (st) return st;
    }
    for (const auto& m : c->host_mats) if (m.if_PEC) c->em_pec_all = true;
    if (c->em_pec_all) {
        auto pec = [&](int64_t i, int64_t j, int64_t k) {
            const uint32_t t = c->host_code[(i * ny + j) * nz + k] & kMatMask;
            return t && c->host_mats[t - 1].if_PEC;
        };
        std::vector<uint8_t> f[3];
        for (int a = 0; a < 3; a++) f[a].assign(static_cast<size_t>(c->nE[a]), 0);
        for (int64_t i = 0; i < nx; i++)
            for (int64_t j = 0; j <= ny; j++)
                for (int64_t k = 0; k <= nz; k++) {
                    const bool wj = j == 0 || j == ny, wk = k == 0 || k == nz;
                    if ((wj && !c->per[1]) || (wk && !c->per[2])) continue;
                    const int64_t y1 = wj ? ny - 1 : j - 1, y3 = wj ? 0 : j, z1 = wk ? nz - 1 : k - 1, z2 = wk ? 0 : k;
                    f[0][static_cast<size_t>((i * (ny + 1) + j) * (nz + 1) + k)] =
                        pec(i, y1, z1) || pec(i, y1, z2) || pec(i, y3, z1) || pec(i, y3, z2);
                }
        for (int64_t i = 0; i <= nx; i++)
            for (int64_t j = 0; j < ny; j++)
                for (int64_t k = 0; k <= nz; k++) {
                    const bool wi = i == 0 || i == nx, wk = k == 0 || k == nz;
                    if ((wi && !c->per[0]) || (wk && !c->per[2])) continue;
                    const int64_t x1 = wi ? nx - 1 : i - 1, x3 = wi ? 0 : i, z1 = wk ? nz - 1 : k - 1, z2 = wk ? 0 : k;
                    f[1][static_cast<size_t>((i * ny + j) * (nz + 1) + k)] =
                        pec(x1, j, z1) || pec(x1, j, z2) || pec(x3, j, z1) || pec(x3, j, z2);
                }
        for (int64_t i = 0; i <= nx; i++)
            for (int64_t j = 0; j <= ny; j++)
                for (int64_t k = 0; k < nz; k++) {
                    const bool wi = i == 0 || i == nx, wj = j == 0 || j == ny;
                    if ((wi && !c->per[0]) || (wj && !c->per[1])) continue;
                    const int64_t x1 = wi ? nx - 1 : i - 1, x3 = wi ? 0 : i, y1 = wj ? ny - 1 : j - 1, y2 = wj ? 0 : j;
                    f[2][static_cast<size_t>((i * (ny + 1) + j) * nz + k)] =
                        pec(x1, y1, k) || pec(x1, y2, k) || pec(x3, y1, k) || pec(x3, y2, k);
                }
        for (int a = 0; a < 3 && !st; a++) {
            st = hip_check(c, hipMalloc(&c->pec[a], f[a].size()), "hipMalloc PEC");
            if (!st)
                st = hip_check(c, hipMemcpy(c->pec[a], f[a].data(), f[a].size(), hipMemcpyHostToDevice), "copy PEC");
        }
        if (st) return st;
    }
    for (int64_t id = 0; id < n && !c->jp_on; id++) {
        const uint32_t t = c->host_code[static_cast<size_t>(id)] & kMatMask;
        if (!t) continue;
        const gf_material& m = c->host_mats[t - 1];
        const double wp[3] = {m.omega_plasma_n1, m.omega_plasma_n2, m.omega_plasma_n3};
        const double tau[3] = {m.tao_e_n1, m.tao_e_n2, m.tao_e_n3};
        const double cn[3] = {m.comp_n1, m.comp_n2, m.comp_n3};
        for (int p = 0; p < 3; p++)
            if (wp[p] != 0. || !std::isfinite(1. / tau[p]) || !std::isfinite(cn[p])) c->jp_on = true;
    }
    return GF_OK;
}

double** e_store(gf_ctx* c, int sel) { return sel == 0 ? c->Eb : (sel == 1 ? c->Ea : c->Ebb); }
double** h_store(gf_ctx* c, int sel) { return sel == 0 ? c->Hb : (sel == 1 ? c->Ha : c->Hbb); }

// Cell-centred E and/or H of the current stores (update_DE_cell /
// update_DH_cell, EMdynamic_solver.cpp:3780-3826).
gf_status em_cells(gf_ctx* c, bool do_e, bool do_h, const EmWallArgs* zwall = nullptr) {
    EmCellArgs a{};
    if (zwall) {
        a.zwall = 1;
        a.w = *zwall;
    }
    a.code = c->d_code; a.mats = c->d_mats;
    double** es = e_store(c, c->e_sel);
    double** hs = h_store(c, c->h_sel);
    for (int i = 0; i < 3; i++) { a.es[i] = es[i]; a.hs[i] = hs[i]; a.ecell[i] = c->Ecell[i]; a.hcell[i] = c->Hcell[i]; }
    a.one_d = (c->flags & GF_FLAG_MAG_1D) && (c->flags & GF_FLAG_EM_FROM_M);
    if (a.one_d) {
        a.mz = c->any_fm ? fm_store(c, c->mag_store_sel)[2] : nullptr;
        a.mz1 = c->any_afm ? afm1_store(c, c->afm_store_sel)[2] : nullptr;
        a.mz2 = c->any_afm ? afm2_store(c, c->afm_store_sel)[2] : nullptr;
        a.hzstat = c->hstat[2];
    }
    a.nx = c->nx; a.ny = c->ny; a.nz = c->nz;
    a.do_e = do_e; a.do_h = do_h;
    hipEvent_t ev = nullptr;
    timer_begin(c, GF_K_OTHER, &ev);
    gf_status st = hip_check(c, launch_em_cells(a, c->stream), "EM cells");
    timer_end(c, GF_K_OTHER, ev);
    return st;
}

// With the electrodynamics on, each stage ends with it: its get phase reads
// dp and dm of the stage, and the other stages read the cell fields of the
// stage-s stores (main.cpp:265-297).
gf_status check_before_em(gf_ctx* c, int stage) {
    if (c->em_on && c->em_next_stage != stage)
        return set_error(c, GF_ERR_STATE, "the electrodynamics expects stage " + std::to_string(c->em_next_stage) +
                                              "; within a stage gf_em_stage comes after the other stage calls");
    return GF_OK;
}

// Strain arrays handed to the magnetic and ferroelectric stages: the
// crystal-frame dynamic strain when the elastic solver runs and acts on
// m and p, else none.
bool strain_coupled(const gf_ctx* c) { return c->el_on && (c->flags & GF_FLAG_ELASTO_ON_PM); }

// The magnets read a static strain (the reference's ex*t0_crt) when strain
// acts on them and no elastodynamics supplies one.
bool static_strain(const gf_ctx* c) {
    return (c->flags & GF_FLAG_ELASTO_ON_PM) && !(c->flags & GF_FLAG_ELASTODYNAMICS);
}

// Strain arrays the magnet stages read (crystal frame), or nulls: the
// dynamic or the static strain, and a static strain uploaded beside the
// dynamic one (exx = exxt0_crt + Dexx_crt, magnetic_solver.cpp:89-96).
void magnet_strain(gf_ctx* c, const double* out[6], const double* out0[6]) {
    for (int i = 0; i < 6; i++) {
        out[i] = strain_coupled(c) ? c->ec[i] : (static_strain(c) ? c->e_t0[i] : nullptr);
        out0[i] = strain_coupled(c) && c->has_et0_dyn ? c->e_t0[i] : nullptr;
    }
}

// sin(2 pi f t) of the shaped drive at the stage time of the uniform drive
// (magnetic_solver.cpp:1834-1861).
double shaped_drive(const gf_ctx* c, int stage) {
    const double t = stage == 1 ? c->time - c->dt : (stage == 4 ? c->time : c->time - c->dt * 0.5);
    return std::sin(2. * kPi * c->hdrive_freq * t);
}

// With the elastic solver on, each stage starts with it: its get phase
// reads every physics' store of the stage (main.cpp:265-297).
gf_status check_after_elastic(gf_ctx* c, int stage) {
    if (!c->el_on) return GF_OK;
    if (c->el_next_stage != (stage == 4 ? 1 : stage + 1))
        return set_error(c, GF_ERR_STATE, "call gf_elastic_stage(" + std::to_string(stage) +
                                              ") before the other stage calls of that stage");
    return GF_OK;
}

}  // namespace

namespace gf {
gf_status set_thread_error(gf_status code, const std::string& msg) {
    g_create_error = msg;
    return code;
}
}  // namespace gf

extern "C" {

const char* gf_last_error(const gf_ctx* ctx) {
    if (!ctx) return g_create_error.c_str();
    return ctx->err.c_str();
}

gf_status gf_create(const gf_desc* d, gf_ctx** out) {
    g_create_error.clear();
    if (!out) { g_create_error = "out is null"; return GF_ERR_ARG; }
    *out = nullptr;
    if (!d) { g_create_error = "desc is null"; return GF_ERR_ARG; }
    if (d->struct_size != sizeof(gf_desc) || d->abi_version != GF_ABI_VERSION) {
        g_create_error = "gf_desc size/version mismatch (header and library differ)";
        return GF_ERR_ARG;
    }
    if (d->nx < 1 || d->ny < 1 || d->nz < 1 || !(d->dx > 0) || !(d->dy > 0) || !(d->dz > 0)) {
        g_create_error = "grid sizes and spacings must be positive";
        return GF_ERR_ARG;
    }
    if (d->n_materials < 0 || (d->n_materials > 0 && !d->materials) || !d->material_cell) {
        g_create_error = "materials / material_cell missing";
        return GF_ERR_ARG;
    }
    if (d->n_materials > static_cast<int32_t>(kMatMask)) {
        g_create_error = "too many materials";
        return GF_ERR_ARG;
    }
    if ((d->flags & GF_FLAG_PRESCRIBE_HEXT) && d->n_prescribe_Hext > 0 && !d->prescribe_Hext) {
        g_create_error = "prescribed drive table missing";
        return GF_ERR_ARG;
    }
    gf_ctx* c = nullptr;
    try {
        c = new gf_ctx();
    } catch (const std::bad_alloc&) {
        g_create_error = "host allocation failed";
        return GF_ERR_ALLOC;
    }
    gf_status st = GF_OK;
    try {
        c->nx = d->nx; c->ny = d->ny; c->nz = d->nz; c->n = d->nx * d->ny * d->nz;
        c->dx = d->dx; c->dy = d->dy; c->dz = d->dz;
        for (int a = 0; a < 3; a++) c->per[a] = d->periodic[a] != 0;
        c->flags = d->flags; c->dt = d->dt; c->time = 0.;
        c->hdrive_freq = d->Hdrive_freq;
        c->precess_angle = d->precess_angle;
        c->precess_freq = d->precess_frequency;
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++) { c->Tg2c[a][b] = d->Tg2c[a][b]; c->Tc2g[b][a] = d->Tg2c[a][b]; }
        for (int a = 0; a < 3; a++) {
            c->hext_stat[a] = d->Hext_stat[a]; c->hext_altr[a] = d->Hext_altr[a];
            c->h_altr_freq[a] = d->H_altr_freq[a]; c->demag_fac[a] = d->demag_fac[a]; c->hext[a] = 0.;
        }
        if ((d->flags & GF_FLAG_PRESCRIBE_HEXT) && d->n_prescribe_Hext > 0)
            c->prescribe.assign(d->prescribe_Hext, d->prescribe_Hext + 4 * d->n_prescribe_Hext);
        for (int a = 0; a < 3; a++) {
            c->eext_stat[a] = d->Eext_stat[a]; c->eext_altr[a] = d->Eext_altr[a];
            c->e_altr_freq[a] = d->E_altr_freq[a]; c->eext[a] = 0.;
        }
        if ((d->flags & GF_FLAG_PRESCRIBE_EEXT) && d->n_prescribe_Eext > 0 && d->prescribe_Eext)
            c->prescribe_e.assign(d->prescribe_Eext, d->prescribe_Eext + 4 * d->n_prescribe_Eext);
        c->host_mats.assign(d->materials, d->materials + d->n_materials);
        c->host_code.resize(static_cast<size_t>(c->n));
        for (int64_t i = 0; i < c->n; i++) {
            const uint32_t t = d->material_cell[i];
            if (t > static_cast<uint32_t>(d->n_materials)) {
                c->err = "material_cell id out of range at cell " + std::to_string(i);
                g_create_error = c->err;
                delete c;
                return GF_ERR_ARG;
            }
            uint32_t code = t;
            if (t && d->materials[t - 1].if_FM) { code |= kIsFM; c->nfm++; }
            if (t && d->materials[t - 1].if_AFM) { code |= kIsAFM; c->nafm++; }
            if (t && d->materials[t - 1].if_FE) { code |= kIsFE; c->nfe++; }
            c->host_code[static_cast<size_t>(i)] = code;
        }
        for (const auto& m : c->host_mats) {
            if (m.if_FM) c->any_fm = true;
            if (m.if_AFM) c->any_afm = true;
            if (m.if_FE) c->any_fe = true;
        }
        if (c->any_fe) build_surface_bits(c, is_fe, kFeSurfShift);
        if (c->any_fm) build_surface_bits(c, is_fm, kFmSurfShift);
        if (c->any_afm) build_surface_bits(c, is_afm, kAfmSurfShift);
        if (c->flags & GF_FLAG_EMDYNAMIC) {
            const std::string why = em_unsupported(c, d);
            if (!why.empty()) {
                g_create_error = why;
                delete c;
                return GF_ERR_UNSUPPORTED;
            }
        }
        if (c->flags & GF_FLAG_ELASTODYNAMICS) {
            const std::string why = elastic_unsupported(c);
            if (!why.empty()) {
                g_create_error = why;
                delete c;
                return GF_ERR_UNSUPPORTED;
            }
        }
    } catch (const std::bad_alloc&) {
        g_create_error = "host allocation failed";
        delete c;
        return GF_ERR_ALLOC;
    }

    st = hip_check(c, hipGetDevice(&c->device), "hipGetDevice");
    if (!st) st = hip_check(c, hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking), "hipStreamCreate");
    if (!st) st = hip_check(c, hipMalloc(&c->d_code, sizeof(uint32_t) * c->n), "hipMalloc code");
    if (!st)
        st = hip_check(c, hipMemcpy(c->d_code, c->host_code.data(), sizeof(uint32_t) * c->n, hipMemcpyHostToDevice),
                       "copy code");
    if (!st && c->any_fm && c->nfm < c->n && c->n <= INT64_C(0xffffffff)) {
        // the FM cells in index order: the unfused FM stage runs over them
        // alone (material kinds cannot change after set-up)
        std::vector<uint32_t> ids;
        ids.reserve(static_cast<size_t>(c->nfm));
        for (int64_t i = 0; i < c->n; i++)
            if (c->host_code[static_cast<size_t>(i)] & kIsFM) ids.push_back(static_cast<uint32_t>(i));
        c->n_fm_ids = static_cast<int64_t>(ids.size());
        if (!ids.empty()) {
            st = hip_check(c, hipMalloc(&c->d_fm_ids, sizeof(uint32_t) * ids.size()), "hipMalloc FM cell list");
            if (!st)
                st = hip_check(c, hipMemcpy(c->d_fm_ids, ids.data(), sizeof(uint32_t) * ids.size(), hipMemcpyHostToDevice),
                               "copy FM cell list");
        }
    }
    if (!st) {
        std::vector<DevMaterial> dm;
        for (const auto& m : c->host_mats) dm.push_back(make_dev_material(m, c->dt));
        if (dm.empty()) dm.push_back(DevMaterial{});
        st = hip_check(c, hipMalloc(&c->d_mats, sizeof(DevMaterial) * dm.size()), "hipMalloc materials");
        if (!st)
            st = hip_check(c, hipMemcpy(c->d_mats, dm.data(), sizeof(DevMaterial) * dm.size(), hipMemcpyHostToDevice),
                           "copy materials");
    }
    if (!st) st = alloc3(c, c->hstat, "hipMalloc hstat");
    if (!st && c->any_fm) {
        st = alloc3(c, c->m, "hipMalloc m");
        if (!st) st = alloc3(c, c->acc, "hipMalloc acc");
        if (!st) st = alloc3(c, c->st_a, "hipMalloc store a");
        if (!st) st = alloc3(c, c->st_b, "hipMalloc store b");
    }
    if (!st && c->any_afm) {
        st = alloc3(c, c->m1, "hipMalloc m1");
        if (!st) st = alloc3(c, c->m2, "hipMalloc m2");
        if (!st) st = alloc3(c, c->acc1, "hipMalloc acc1");
        if (!st) st = alloc3(c, c->acc2, "hipMalloc acc2");
        if (!st) st = alloc3(c, c->st1_a, "hipMalloc store1 a");
        if (!st) st = alloc3(c, c->st1_b, "hipMalloc store1 b");
        if (!st) st = alloc3(c, c->st2_a, "hipMalloc store2 a");
        if (!st) st = alloc3(c, c->st2_b, "hipMalloc store2 b");
    }
    if (!st && c->any_fe) {
        double** fe_bufs[] = {c->p, c->q, c->pacc, c->qacc, c->p_a, c->p_b, c->q_a, c->q_b, c->estat};
        for (double** b : fe_bufs)
            if (!st) st = alloc3(c, b, "hipMalloc ferroelectric state");
    }
    if (!st && (c->flags & GF_FLAG_ELASTODYNAMICS)) {
        c->el_on = true;
        double** el_bufs[] = {c->u, c->v, c->us, c->us_alt, c->vs, c->accu, c->accv, c->fstore, c->epx, c->epy,
                              c->epz};
        for (double** b : el_bufs)
            if (!st) st = alloc3(c, b, "hipMalloc elastic state");
        const int64_t nx = c->nx, ny = c->ny, nz = c->nz;
        for (int a = 0; a < 3 && !st; a++) {
            st = alloc_n(c, &c->gyz[a], (nx + 1) * ny * nz, "hipMalloc gradients");
            if (!st) st = alloc_n(c, &c->gxz[a], nx * (ny + 1) * nz, "hipMalloc gradients");
            if (!st) st = alloc_n(c, &c->gxy[a], nx * ny * (nz + 1), "hipMalloc gradients");
        }
        const bool rotate = c->flags & GF_FLAG_ORIENT_ROTATE;
        for (int a = 0; a < 6 && !st; a++) {
            st = alloc_n(c, &c->eg[a], c->n, "hipMalloc strain");
            if (!st) st = alloc_n(c, &c->e0g[a], c->n, "hipMalloc eigenstrain");
            if (!st && rotate) st = alloc_n(c, &c->ec[a], c->n, "hipMalloc strain");
            if (!st && rotate) st = alloc_n(c, &c->e0c[a], c->n, "hipMalloc eigenstrain");
            if (!rotate) { c->ec[a] = c->eg[a]; c->e0c[a] = c->e0g[a]; }
        }
        c->hist_slots = (c->flags & GF_FLAG_1D_ABC) ? 1 : 4;
        c->hist_head = 0;
        if (!st)
            st = alloc_n(c, &c->hist, static_cast<int64_t>(c->hist_slots) * kHistPlanes * 3 * nx * ny,
                         "hipMalloc boundary history");
        c->src_type = d->input_strain_type;
        c->src_k = d->input_strain_sourcez;
        const int ch = d->input_strain_component;
        c->src_comp = ch == 'x' ? 0 : (ch == 'y' ? 1 : (ch == 'z' ? 2 : -1));
        c->src_cycles = d->num_strain_cycles;
        c->src_sigma = d->sigma_gauss;
        c->src_amp = d->amplitude_gauss;
        c->w_liao = d->weighting_factor_fourth_Liao;
        elastic_speeds(c);
    }
    if (!st && (c->flags & GF_FLAG_HDRIVE)) st = alloc3(c, c->hdrive, "hipMalloc drive profile");
    if (!st && static_strain(c)) {
        st = alloc3(c, c->e_t0, "hipMalloc static strain");
        if (!st) st = alloc3(c, c->e_t0 + 3, "hipMalloc static strain");
    }
    if (!st && (c->flags & GF_FLAG_EMDYNAMIC)) st = em_setup(c, d);
    if (!st && (c->flags & GF_FLAG_SPIN_PUMPING)) st = spin_setup(c, d);
    if (!st && (c->flags & GF_FLAG_FLEXO)) st = flexo_setup(c);
    if (!st && (c->flags & GF_FLAG_ELASTOSTATIC)) st = elastostatic_setup(c, d);
    if (!st) {
        c->red_blocks = std::min<int64_t>(1024, (c->n + 255) / 256);
        st = hip_check(c, hipMalloc(&c->red_partial, sizeof(double) * 3 * c->red_blocks), "hipMalloc reduction");
    }
    if (!st && (c->flags & GF_FLAG_MAGNETOSTATICS) && !(c->flags & GF_FLAG_MAG_1D) &&
        !(c->flags & GF_FLAG_DEMAG_FACTOR) && (c->any_fm || c->any_afm))
        st = demag_setup(c);
    if (!st && (c->flags & GF_FLAG_ELECTROSTATICS)) {
        if (!c->any_fe) {
            st = set_error(c, GF_ERR_UNSUPPORTED, "FFT electrostatics needs a ferroelectric material");
        } else {
            // er_homo: the largest relative permittivity times scale_elec
            // (ferroelectric_initialize.cpp:311-318)
            double er = 0.;
            for (const auto& m : c->host_mats) er = std::max(er, m.r_permittivity);
            st = elec_setup(c, er * (d->scale_elec > 0. ? d->scale_elec : 1.));
        }
    }
    if (!st) st = hip_check(c, hipStreamSynchronize(c->stream), "create sync");
    if (st) {
        g_create_error = c->err;
        gf_destroy(c);
        return st;
    }
    *out = c;
    return GF_OK;
}

gf_status gf_destroy(gf_ctx* c) {
    if (!c) return GF_ERR_ARG;
    if (c->stream) hipStreamSynchronize(c->stream);
    for (auto& t : c->timers) { hipEventDestroy(t.start); hipEventDestroy(t.stop); }
    for (auto e : c->event_pool) hipEventDestroy(e);
    free3(c->m); free3(c->acc); free3(c->st_a); free3(c->st_b); free3(c->dm_last);
    free3(c->m1); free3(c->m2); free3(c->acc1); free3(c->acc2);
    free3(c->st1_a); free3(c->st1_b); free3(c->st2_a); free3(c->st2_b);
    free3(c->hstat); free3(c->hnu); free3(c->hdrive); free3(c->e_t0); free3(c->e_t0 + 3);
    free3(c->e0t0); free3(c->e0t0 + 3);
    free3(c->p); free3(c->q); free3(c->pacc); free3(c->qacc);
    free3(c->p_a); free3(c->p_b); free3(c->q_a); free3(c->q_b); free3(c->estat);
    free3(c->u); free3(c->v); free3(c->us); free3(c->vs); free3(c->accu); free3(c->accv); free3(c->fstore);
    free3(c->gyz); free3(c->gxz); free3(c->gxy); free3(c->us_alt); free3(c->epx); free3(c->epy); free3(c->epz);
    for (int a = 0; a < 6; a++) {
        if (c->ec[a] && c->ec[a] != c->eg[a]) hipFree(c->ec[a]);
        if (c->e0c[a] && c->e0c[a] != c->e0g[a]) hipFree(c->e0c[a]);
        if (c->eg[a]) hipFree(c->eg[a]);
        if (c->e0g[a]) hipFree(c->e0g[a]);
        c->ec[a] = c->eg[a] = c->e0c[a] = c->e0g[a] = nullptr;
    }
    if (c->hist) hipFree(c->hist);
    free3(c->Eb); free3(c->Ea); free3(c->Ebb); free3(c->Eacc); free3(c->Et1);
    free3(c->Hb); free3(c->Ha); free3(c->Hbb); free3(c->Hacc);
    free3(c->Ecell); free3(c->Hcell);
    free3(c->em_dm); free3(c->em_dm1); free3(c->em_dm2);
    free3(c->jis);
    free3(c->flexo_force);
    if (c->d_flexo) hipFree(c->d_flexo);
    if (c->es_buf) hipFree(c->es_buf);
    if (c->spin_src) hipFree(c->spin_src);
    if (c->spin_coef) hipFree(c->spin_coef);
    if (c->spin_is_src) hipFree(c->spin_is_src);
    free3(c->spin_m0); free3(c->spin_m10); free3(c->spin_m20);
    for (int q = 0; q < 9; q++) {
        if (c->Jp[q]) hipFree(c->Jp[q]);
        if (c->Jps[q]) hipFree(c->Jps[q]);
        if (c->Jacc[q]) hipFree(c->Jacc[q]);
    }
    for (int a = 0; a < 3; a++)
        if (c->pec[a]) hipFree(c->pec[a]);
    if (c->jf_spatial) hipFree(c->jf_spatial);
    if (c->red_partial) hipFree(c->red_partial);
    if (c->d_code) hipFree(c->d_code);
    if (c->d_fm_ids) hipFree(c->d_fm_ids);
    if (c->d_mats) hipFree(c->d_mats);
    adjoint_free(c);
    adjoint32_free(c);
    demag_free(c);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
    return GF_OK;
}

static double** field_for(gf_ctx* c, int id, bool for_write, gf_status* st) {
    *st = GF_OK;
    switch (id) {
    case GF_FIELD_M:
        if (!c->any_fm) break;
        return c->m;
    case GF_FIELD_M_STORE:
        if (!c->any_fm) break;
        if (for_write && c->mag_store_sel == 0) {
            // a store that differs from m lives in a free buffer
            c->mag_store_sel = 2;
        }
        return fm_store(c, c->mag_store_sel);
    case GF_FIELD_HSTAT:
        if (for_write) c->hstat_written = true;
        return c->hstat;
    case GF_FIELD_HEXT_NONUNIF:
        if (for_write && !c->hnu[0]) {
            if ((*st = alloc3(c, c->hnu, "hipMalloc hext_nonunif"))) return nullptr;
        }
        if (!c->hnu[0]) break;
        return c->hnu;
    case GF_FIELD_M_AFM1:
        if (!c->any_afm) break;
        return c->m1;
    case GF_FIELD_M_AFM2:
        if (!c->any_afm) break;
        return c->m2;
    case GF_FIELD_P:
        if (!c->any_fe) break;
        return c->p;
    case GF_FIELD_Q:
        if (!c->any_fe) break;
        return c->q;
    case GF_FIELD_P_STORE:
        if (!c->any_fe) break;
        if (for_write && c->fe_store_sel == 0) c->fe_store_sel = 2;
        return fe_pstore(c, c->fe_store_sel);
    case GF_FIELD_ESTAT:
        if (!c->any_fe) break;
        return c->estat;
    case GF_FIELD_U:
        if (!c->el_on) break;
        return c->u;
    case GF_FIELD_V:
        if (!c->el_on) break;
        return c->v;
    case GF_FIELD_U_STORE:
        if (!c->el_on) break;
        return c->us;
    case GF_FIELD_ELASTOFORCE:
        if (!c->el_on) break;
        return c->fstore;
    case GF_FIELD_EIGENSTRAIN_T0_DIAG:
    case GF_FIELD_EIGENSTRAIN_T0_OFF:
        if (!c->el_on && !(c->flags & GF_FLAG_ELASTOSTATIC)) break;
        if (!c->e0t0[0]) {
            // six zeroed arrays on first use; the stage reads them only once written
            if ((*st = alloc3(c, c->e0t0, "hipMalloc static eigenstrain"))) return nullptr;
            if ((*st = alloc3(c, c->e0t0 + 3, "hipMalloc static eigenstrain"))) return nullptr;
        }
        return id == GF_FIELD_EIGENSTRAIN_T0_DIAG ? c->e0t0 : c->e0t0 + 3;
    case GF_FIELD_STRAIN_DIAG:
        if (!c->el_on || for_write) break;
        return c->eg;
    case GF_FIELD_STRAIN_OFF:
        if (!c->el_on || for_write) break;
        return c->eg + 3;
    case GF_FIELD_HDRIVE:
        if (!c->hdrive[0]) break;
        return c->hdrive;
    case GF_FIELD_STRAIN_T0_DIAG:
    case GF_FIELD_STRAIN_T0_OFF:
        if (!c->e_t0[0] && c->el_on && for_write) {
            // a static strain beside the elastodynamics, on first upload
            if ((*st = alloc3(c, c->e_t0, "hipMalloc static strain"))) return nullptr;
            if ((*st = alloc3(c, c->e_t0 + 3, "hipMalloc static strain"))) return nullptr;
        }
        if (!c->e_t0[0]) break;
        return id == GF_FIELD_STRAIN_T0_DIAG ? c->e_t0 : c->e_t0 + 3;
    case GF_FIELD_J_ISHE:
        if (!c->spin_on || for_write) break;
        return c->jis;
    case GF_FIELD_E_CELL:
        if (!c->em_on || for_write) break;
        return c->Ecell;
    case GF_FIELD_H_CELL:
        if (!c->em_on || for_write) break;
        return c->Hcell;
    default:
        break;
    }
    *st = set_error(c, GF_ERR_ARG, "field " + std::to_string(id) + " is not available in this context");
    return nullptr;
}

// The Yee fields (three staggered arrays, concatenated x|y|z) and Jp (nine
// cell arrays): the device arrays and their sizes, or 0 when id is none of
// them or the electrodynamics is off.
static int em_parts(gf_ctx* c, int id, bool for_write, double* parts[9], int64_t sizes[9]) {
    if (!c->em_on) return 0;
    double** arr = nullptr;
    const int64_t* sz = nullptr;
    if (id == GF_FIELD_E_EM) { arr = c->Eb; sz = c->nE; }
    else if (id == GF_FIELD_H_EM) { arr = c->Hb; sz = c->nH; }
    else if (id == GF_FIELD_E_EM_STORE) {
        if (for_write && c->e_sel == 0) c->e_sel = 2;   // a store that differs from E lives in a free buffer
        arr = e_store(c, c->e_sel); sz = c->nE;
    } else if (id == GF_FIELD_H_EM_STORE) {
        if (for_write && c->h_sel == 0) c->h_sel = 2;
        arr = h_store(c, c->h_sel); sz = c->nH;
    }
    if (arr) {
        for (int a = 0; a < 3; a++) { parts[a] = arr[a]; sizes[a] = sz[a]; }
        return 3;
    }
    if (id == GF_FIELD_JP) {
        for (int q = 0; q < 9; q++) { parts[q] = c->Jp[q]; sizes[q] = c->n; }
        return 9;
    }
    return 0;
}

static bool is_em_field(int id) {
    return id == GF_FIELD_E_EM || id == GF_FIELD_H_EM || id == GF_FIELD_E_EM_STORE || id == GF_FIELD_H_EM_STORE ||
           id == GF_FIELD_JP;
}

static gf_status em_transfer(gf_ctx* c, int id, double* host, size_t count, bool up) {
    double* parts[9];
    int64_t sizes[9];
    const int np = em_parts(c, id, up, parts, sizes);
    if (!np) return set_error(c, GF_ERR_ARG, "field " + std::to_string(id) + " is not available in this context");
    size_t tot = 0;
    for (int q = 0; q < np; q++) tot += static_cast<size_t>(sizes[q]);
    if (count != tot) return set_error(c, GF_ERR_ARG, "count must be " + std::to_string(tot) + " for this field");
    size_t o = 0;
    gf_status st = GF_OK;
    for (int q = 0; q < np && !st; q++) {
        const size_t b = sizeof(double) * static_cast<size_t>(sizes[q]);
        st = up ? hip_check(c, hipMemcpyAsync(parts[q], host + o, b, hipMemcpyHostToDevice, c->stream), "upload")
                : hip_check(c, hipMemcpyAsync(host + o, parts[q], b, hipMemcpyDeviceToHost, c->stream), "download");
        o += static_cast<size_t>(sizes[q]);
    }
    if (!st && up) {
        // EMdynamic_initialize.cpp:376-404: the stores start equal to the
        // fields and the cell fields follow them
        if (id == GF_FIELD_E_EM) { c->e_sel = 0; c->em_next_stage = 1; st = em_cells(c, true, false); }
        if (id == GF_FIELD_H_EM) { c->h_sel = 0; c->em_next_stage = 1; st = em_cells(c, false, true); }
        if (id == GF_FIELD_JP) {
            for (int q = 0; q < 9 && !st; q++)
                st = hip_check(c, hipMemcpyAsync(c->Jps[q], c->Jp[q], sizeof(double) * c->n, hipMemcpyDeviceToDevice,
                                                 c->stream), "Jp store");
            c->jp_on = true;
            c->jp_uploaded = true;
            c->em_next_stage = 1;
        }
    }
    if (st) return st;
    return hip_check(c, hipStreamSynchronize(c->stream), up ? "upload sync" : "download sync");
}

gf_status gf_set_material(gf_ctx* c, int index, const gf_material* m) {
    if (!c || !m) return set_error(c, GF_ERR_ARG, "null argument");
    if (index < 1 || index > static_cast<int>(c->host_mats.size()))
        return set_error(c, GF_ERR_ARG, "material index out of range");
    if (c->el_on || c->em_on)
        return set_error(c, GF_ERR_UNSUPPORTED,
                         "material updates need a context without elastodynamics or electrodynamics "
                         "(their set-up derives per-cell constants from the materials)");
    const gf_material& o = c->host_mats[static_cast<size_t>(index - 1)];
    if (o.if_FM != m->if_FM || o.if_AFM != m->if_AFM || o.if_FE != m->if_FE || o.if_PEC != m->if_PEC)
        return set_error(c, GF_ERR_ARG, "the kind flags of a material cannot change");
    gf_status st = hip_check(c, hipStreamSynchronize(c->stream), "material sync");
    if (st) return st;
    c->host_mats[static_cast<size_t>(index - 1)] = *m;
    const DevMaterial d = make_dev_material(*m, c->dt);
    st = hip_check(c, hipMemcpy(c->d_mats + (index - 1), &d, sizeof(DevMaterial), hipMemcpyHostToDevice),
                   "copy material");
    if (!st && c->d_flexo) {
        const FlexoMat f = make_flexo(*m);
        st = hip_check(c, hipMemcpy(c->d_flexo + (index - 1), &f, sizeof(FlexoMat), hipMemcpyHostToDevice),
                       "copy flexo constants");
    }
    return st;
}

gf_status gf_field_size(const gf_ctx* c, int id, size_t* count) {
    if (!c || !count) return GF_ERR_ARG;
    if (is_em_field(id)) {
        if (!c->em_on) return GF_ERR_ARG;
        const int64_t* sz = (id == GF_FIELD_H_EM || id == GF_FIELD_H_EM_STORE) ? c->nH : c->nE;
        *count = id == GF_FIELD_JP ? static_cast<size_t>(9 * c->n) : static_cast<size_t>(sz[0] + sz[1] + sz[2]);
        return GF_OK;
    }
    if (id < 0 || id >= GF_FIELD_COUNT) return GF_ERR_ARG;
    *count = static_cast<size_t>(3 * c->n);
    return GF_OK;
}

gf_status gf_upload(gf_ctx* c, int id, const double* host, size_t count) {
    if (!c || !host) return set_error(c, GF_ERR_ARG, "null argument");
    if (is_em_field(id)) return em_transfer(c, id, const_cast<double*>(host), count, true);
    if (count != static_cast<size_t>(3 * c->n)) return set_error(c, GF_ERR_ARG, "count must be 3*nx*ny*nz");
    gf_status st;
    double** f = field_for(c, id, true, &st);
    if (!f) return st;
    const size_t bytes = sizeof(double) * c->n;
    for (int a = 0; a < 3; a++) {
        st = hip_check(c, hipMemcpyAsync(f[a], host + a * c->n, bytes, hipMemcpyHostToDevice, c->stream), "upload");
        if (st) return st;
    }
    // initialize_device: the stores start equal to m (magnetic_initialize.cpp:1081-1097)
    auto reset = [&](double** m, double** sa, double** sb) -> gf_status {
        for (int a = 0; a < 3; a++) {
            gf_status s2 = hip_check(c, hipMemcpyAsync(sa[a], m[a], bytes, hipMemcpyDeviceToDevice, c->stream), "store");
            if (s2) return s2;
            s2 = hip_check(c, hipMemcpyAsync(sb[a], m[a], bytes, hipMemcpyDeviceToDevice, c->stream), "store");
            if (s2) return s2;
        }
        return GF_OK;
    };
    if (id == GF_FIELD_M) { st = reset(c->m, c->st_a, c->st_b); c->mag_store_sel = 0; c->mag_next_stage = 1; }
    if (id == GF_FIELD_M_AFM1) { st = reset(c->m1, c->st1_a, c->st1_b); c->afm_store_sel = 0; c->afm_next_stage = 1; }
    if (id == GF_FIELD_M_AFM2) { st = reset(c->m2, c->st2_a, c->st2_b); c->afm_store_sel = 0; c->afm_next_stage = 1; }
    // ferroelectric_initialize.cpp:355-366: the stores start equal to p and q
    if (id == GF_FIELD_P || id == GF_FIELD_Q) {
        const bool isp = id == GF_FIELD_P;
        st = reset(isp ? c->p : c->q, isp ? c->p_a : c->q_a, isp ? c->p_b : c->q_b);
        if (c->fe_store_sel != 0) {
            // a stage store that differs from the state was uploaded; the
            // other one of the pair restarts from its state as well
            st = st ? st : reset(isp ? c->q : c->p, isp ? c->q_a : c->p_a, isp ? c->q_b : c->p_b);
        }
        c->fe_store_sel = 0;
        c->fe_next_stage = 1;
    }
    if (id == GF_FIELD_HEXT_NONUNIF) c->has_hnu = true;
    if (id == GF_FIELD_EIGENSTRAIN_T0_DIAG || id == GF_FIELD_EIGENSTRAIN_T0_OFF) c->has_e0t0 = true;
    if ((id == GF_FIELD_STRAIN_T0_DIAG || id == GF_FIELD_STRAIN_T0_OFF) && c->el_on) c->has_et0_dyn = true;
    // an uploaded static strain or eigenstrain replaces the start-up solve
    // of the elastodynamics (main.cpp:148)
    if (id == GF_FIELD_STRAIN_T0_DIAG || id == GF_FIELD_STRAIN_T0_OFF || id == GF_FIELD_EIGENSTRAIN_T0_DIAG ||
        id == GF_FIELD_EIGENSTRAIN_T0_OFF)
        c->t0_fixed = true;
    if (id == GF_FIELD_M || id == GF_FIELD_M_STORE) c->demag.m_in_spec = false;
    if (id == GF_FIELD_HSTAT) { c->demag.h_in_spec = false; c->demag.h_real_valid = true; }
    // elastic_initialize.cpp:386-410: the stores start equal to u and v
    if (id == GF_FIELD_U || id == GF_FIELD_V) {
        const bool isu = id == GF_FIELD_U;
        for (int a = 0; a < 3 && !st; a++)
            st = hip_check(c, hipMemcpyAsync(isu ? c->us[a] : c->vs[a], isu ? c->u[a] : c->v[a], bytes,
                                             hipMemcpyDeviceToDevice, c->stream), "store");
        c->el_next_stage = 1;
    }
    if (st) return st;
    return hip_check(c, hipStreamSynchronize(c->stream), "upload sync");
}

gf_status gf_download(gf_ctx* c, int id, double* host, size_t count) {
    if (!c || !host) return set_error(c, GF_ERR_ARG, "null argument");
    if (is_em_field(id)) return em_transfer(c, id, host, count, false);
    if (count != static_cast<size_t>(3 * c->n)) return set_error(c, GF_ERR_ARG, "count must be 3*nx*ny*nz");
    gf_status st;
    double** f = field_for(c, id, false, &st);
    if (!f) return st;
    // the strain arrays hold the strain of the store the last elastic stage
    // started from; refresh them from the current store
    if (id == GF_FIELD_STRAIN_DIAG || id == GF_FIELD_STRAIN_OFF) {
        st = elastic_strain(c);
        if (st) return st;
    }
    if (id == GF_FIELD_HSTAT && (st = demag_materialize_h(c))) return st;
    for (int a = 0; a < 3; a++) {
        st = hip_check(c, hipMemcpyAsync(host + a * c->n, f[a], sizeof(double) * c->n, hipMemcpyDeviceToHost, c->stream),
                       "download");
        if (st) return st;
    }
    return hip_check(c, hipStreamSynchronize(c->stream), "download sync");
}

// global.cpp:513-549 (H drive table; t advances every call)
gf_status gf_time_advance(gf_ctx* c) {
    if (!c) return GF_ERR_ARG;
    c->time = c->time + c->dt;
    if (c->flags & GF_FLAG_PRESCRIBE_EEXT) {
        const int np = static_cast<int>(c->prescribe_e.size() / 4);
        if (c->prescribe_e_idx < np && (c->time - c->dt) > c->prescribe_e[4 * c->prescribe_e_idx]) c->prescribe_e_idx++;
        if (c->prescribe_e_idx < np) {
            for (int a = 0; a < 3; a++) c->eext[a] = c->prescribe_e[4 * c->prescribe_e_idx + 1 + a];
        } else {
            for (int a = 0; a < 3; a++) c->eext[a] = 0.;
        }
    }
    if (c->flags & GF_FLAG_PRESCRIBE_HEXT) {
        const int np = static_cast<int>(c->prescribe.size() / 4);
        if (c->prescribe_idx < np && (c->time - c->dt) > c->prescribe[4 * c->prescribe_idx]) c->prescribe_idx++;
        if (c->prescribe_idx < np) {
            for (int a = 0; a < 3; a++) c->hext[a] = c->prescribe[4 * c->prescribe_idx + 1 + a];
        } else {
            for (int a = 0; a < 3; a++) c->hext[a] = 0.;
        }
    }
    return GF_OK;
}

}  // extern "C"

namespace gf {

// Arguments of one FM stage reading store in_sel and writing store out_sel
// (0 = m, 1 = st_a, 2 = st_b); the drive is left at zero.
// Hstat can be non-zero only when statics run or it was uploaded; the
// stage kernels skip reading it otherwise (it is zero from allocation).
static bool hstat_live(const gf_ctx* c) {
    return c->hstat_written || (c->flags & (GF_FLAG_MAGNETOSTATICS | GF_FLAG_MAG_1D | GF_FLAG_DEMAG_FACTOR));
}

StageArgs fm_stage_args(gf_ctx* c, int stage, int in_sel, int out_sel) {
    StageArgs a{};
    magnet_strain(c, a.strain, a.strain0);
    const bool em_m = c->em_on && (c->flags & GF_FLAG_EM_BACKACTION_M);
    a.code = c->d_code; a.mats = c->d_mats;
    double** s = fm_store(c, in_sel);
    double** o = fm_store(c, out_sel);
    for (int i = 0; i < 3; i++) {
        a.store[i] = s[i]; a.base[i] = c->m[i]; a.acc[i] = c->acc[i]; a.out[i] = o[i];
        a.hstat[i] = hstat_live(c) ? c->hstat[i] : nullptr;
        a.hnu[i] = c->has_hnu ? c->hnu[i] : nullptr; a.dm_out[i] = c->em_dm[i];
        a.hdrive[i] = c->hdrive[i];
        a.hcell[i] = em_m ? c->Hcell[i] : nullptr;
    }
    a.nx = c->nx; a.ny = c->ny; a.nz = c->nz; a.dx = c->dx; a.dy = c->dy; a.dz = c->dz;
    std::memcpy(a.Tg2c, c->Tg2c, sizeof(a.Tg2c));
    std::memcpy(a.Tc2g, c->Tc2g, sizeof(a.Tc2g));
    a.rotate = (c->flags & GF_FLAG_ORIENT_ROTATE) ? 1 : 0;
    a.stage = stage;
    a.nmat = static_cast<int32_t>(c->host_mats.size());
    return a;
}

// Uniform and shaped drives of FM stage s at the context's current time.
// The prescribed cone of the bottom layer at time t
// (magnetic_solver.cpp:1891-1957, the same expressions).
void prescribed_cone(const gf_ctx* c, double t, double out[3]) {
    const double m0 = std::sin(c->precess_angle / 180. * kPi);
    const double omega = 2. * kPi * c->precess_freq;
    out[0] = std::sqrt((1. - m0 * m0));
    out[1] = m0 * std::sin(omega * t);
    out[2] = m0 * std::cos(omega * t);
}

void fm_stage_drives(gf_ctx* c, int stage, StageArgs& a) {
    stage_drive(c, stage, a.hext);
    a.drive_s = a.hdrive[0] ? shaped_drive(c, stage) : 0.;
    // update_m_RK1/RK2 prescribe the half-step cone, update_m_RK3 and
    // update_m the full-step one (magnetic_solver.cpp:1526-1609)
    a.prescribe = (c->flags & GF_FLAG_PRESCRIBED_M) ? 1 : 0;
    if (a.prescribe) prescribed_cone(c, stage <= 2 ? c->time - c->dt * 0.5 : c->time, a.presc_m);
}

// The same for the two AFM sublattices.
AfmStageArgs afm_stage_args(gf_ctx* c, int stage, int in_sel, int out_sel) {
    AfmStageArgs a{};
    magnet_strain(c, a.strain, a.strain0);
    const bool em_m = c->em_on && (c->flags & GF_FLAG_EM_BACKACTION_M);
    for (int i = 0; i < 3; i++) {
        a.hcell[i] = em_m ? c->Hcell[i] : nullptr;
        a.dm_out1[i] = c->em_dm1[i];
        a.dm_out2[i] = c->em_dm2[i];
    }
    a.code = c->d_code; a.mats = c->d_mats;
    double** s1 = afm1_store(c, in_sel);
    double** s2 = afm2_store(c, in_sel);
    double** o1 = afm1_store(c, out_sel);
    double** o2 = afm2_store(c, out_sel);
    for (int i = 0; i < 3; i++) {
        a.store1[i] = s1[i]; a.store2[i] = s2[i];
        a.base1[i] = c->m1[i]; a.base2[i] = c->m2[i];
        a.acc1[i] = c->acc1[i]; a.acc2[i] = c->acc2[i];
        a.out1[i] = o1[i]; a.out2[i] = o2[i];
        a.hstat[i] = hstat_live(c) ? c->hstat[i] : nullptr;
    }
    a.nx = c->nx; a.ny = c->ny; a.nz = c->nz; a.dx = c->dx; a.dy = c->dy; a.dz = c->dz;
    std::memcpy(a.Tg2c, c->Tg2c, sizeof(a.Tg2c));
    std::memcpy(a.Tc2g, c->Tc2g, sizeof(a.Tc2g));
    a.rotate = (c->flags & GF_FLAG_ORIENT_ROTATE) ? 1 : 0;
    a.stage = stage;
    return a;
}

// ---- argument aliasing ----------------------------------------------------
// A stage launch reads some arrays at neighbouring cells (the stencil's
// store, the cell codes) and the rest at its own cell, most of them before
// the cell's code is known (all of a cell's loads are issued together);
// it writes its outputs at its own cell.  The launch is well defined only
// if every array it writes holds one quantity for the launch (one role) and
// every array it reads is either not written by it, or read only at the
// cell and as the same quantity (the in-place RK update of the state and
// the accumulator).  stage_io_conflict names the first pair that breaks it.
struct IoRef {
    const void* p;
    const char* role;
    int comp;
    bool gather;      // read at neighbouring cells
};

struct IoList {
    IoRef r[80];
    int n = 0;
    void add(const void* p, const char* role, int comp, bool gather = false) {
        if (p && n < 80) r[n++] = IoRef{p, role, comp, gather};
    }
};

static bool same_role(const IoRef& a, const IoRef& b) {
    return a.comp == b.comp && std::strcmp(a.role, b.role) == 0;
}

static std::string io_conflict(const IoList& rd, const IoList& wr) {
    auto name = [](const IoRef& x) { return std::string(x.role) + "[" + std::to_string(x.comp) + "]"; };
    for (int i = 0; i < wr.n; i++) {
        for (int j = 0; j < i; j++)
            if (wr.r[j].p == wr.r[i].p) return "it writes " + name(wr.r[j]) + " and " + name(wr.r[i]) + " to one array";
        for (int j = 0; j < rd.n; j++) {
            const IoRef& r = rd.r[j];
            if (r.p != wr.r[i].p) continue;
            if (r.gather) return "it reads " + name(r) + " at neighbouring cells from the array it writes as " +
                                 name(wr.r[i]);
            if (!same_role(r, wr.r[i])) return "it reads " + name(r) + " from the array it writes as " + name(wr.r[i]);
        }
    }
    return "";
}

std::string stage_io_conflict(const StageArgs& a, const FusedZArgs* f) {
    IoList rd, wr;
    rd.add(a.code, "code", 0, true);
    for (int c = 0; c < 3; c++) {
        rd.add(a.store[c], "m store", c, true);
        rd.add(a.base[c], "m", c);
        if (a.stage > 1) rd.add(a.acc[c], "acc", c);
        rd.add(a.hstat[c], "Hstat", c);
        rd.add(a.hnu[c], "Hext_nonunif", c);
        rd.add(a.hdrive[c], "Hdrive", c);
        rd.add(a.hcell[c], "H cell", c);
        wr.add(a.out[c], a.stage == 4 ? "m" : "next m store", c);
        if (a.stage < 4) wr.add(a.acc[c], "acc", c);
        wr.add(a.dm_out[c], "k", c);
    }
    for (int c = 0; c < 6; c++) rd.add(a.strain[c], "strain", c);
    if (f) {
        rd.add(f->hspec, "H z-spectrum", 0, true);
        rd.add(f->tw, "twiddles", 0, true);
        wr.add(f->mspec, "M z-spectrum", 0);
    }
    return io_conflict(rd, wr);
}

std::string stage_io_conflict(const AfmStageArgs& a) {
    IoList rd, wr;
    rd.add(a.code, "code", 0, true);
    for (int c = 0; c < 3; c++) {
        rd.add(a.store1[c], "m1 store", c, true);
        rd.add(a.store2[c], "m2 store", c, true);
        rd.add(a.base1[c], "m1", c);
        rd.add(a.base2[c], "m2", c);
        if (a.stage > 1) {
            rd.add(a.acc1[c], "acc1", c);
            rd.add(a.acc2[c], "acc2", c);
        }
        rd.add(a.hstat[c], "Hstat", c);
        rd.add(a.hcell[c], "H cell", c);
        wr.add(a.out1[c], a.stage == 4 ? "m1" : "next m1 store", c);
        wr.add(a.out2[c], a.stage == 4 ? "m2" : "next m2 store", c);
        if (a.stage < 4) {
            wr.add(a.acc1[c], "acc1", c);
            wr.add(a.acc2[c], "acc2", c);
        }
        wr.add(a.dm_out1[c], "k1", c);
        wr.add(a.dm_out2[c], "k2", c);
    }
    for (int c = 0; c < 6; c++) rd.add(a.strain[c], "strain", c);
    return io_conflict(rd, wr);
}

std::string stage_io_conflict(const FeStageArgs& a) {
    IoList rd, wr;
    rd.add(a.code, "code", 0, true);
    for (int c = 0; c < 3; c++) {
        rd.add(a.pstore[c], "p store", c, true);
        rd.add(a.qstore[c], "q store", c);
        rd.add(a.pbase[c], "p", c);
        rd.add(a.qbase[c], "q", c);
        if (a.stage > 1) {
            rd.add(a.pacc[c], "p acc", c);
            rd.add(a.qacc[c], "q acc", c);
        }
        rd.add(a.estat[c], "Estat", c);
        rd.add(a.ecell[c], "E cell", c);
        rd.add(a.flexo[c], "flexo force", c);
        wr.add(a.pout[c], a.stage == 4 ? "p" : "next p store", c);
        wr.add(a.qout[c], a.stage == 4 ? "q" : "next q store", c);
        wr.add(a.pacc[c], "p acc", c);
        wr.add(a.qacc[c], "q acc", c);
    }
    for (int c = 0; c < 6; c++) {
        rd.add(a.strain[c], "strain", c);
        rd.add(a.eigen[c], "eigenstrain", c);
        rd.add(a.strain0[c], "static strain", c);
        rd.add(a.eigen0[c], "static eigenstrain", c);
    }
    return io_conflict(rd, wr);
}

// Arguments of one FE stage reading store in_sel and writing out_sel
// (0 = p/q, 1 = the a stores, 2 = the b stores); the drive is left at zero.
FeStageArgs fe_stage_args(gf_ctx* c, int stage, int in_sel, int out_sel) {
    FeStageArgs a{};
    // the strain and eigenstrain of the elastodynamics, or, with the
    // elastostatics instead, the static ones (ex*t0_crt, ex*0t0_crt:
    // ferroelectric_solver.cpp:624-653 with the dynamic parts zero); with
    // both, the static ones added to the dynamic ones (:636-650)
    const bool stat = !c->el_on && (c->flags & GF_FLAG_ELASTOSTATIC) && (c->flags & GF_FLAG_ELASTO_ON_PM);
    const bool both = strain_coupled(c) && (c->flags & GF_FLAG_ELASTOSTATIC);
    for (int i = 0; i < 6; i++) {
        a.strain[i] = strain_coupled(c) ? c->ec[i] : (stat ? c->e_t0[i] : nullptr);
        a.eigen[i] = strain_coupled(c) ? c->e0c[i] : (stat ? c->e0t0[i] : nullptr);
        a.strain0[i] = both ? c->e_t0[i] : nullptr;
        a.eigen0[i] = both ? c->e0t0[i] : nullptr;
    }
    const bool em_p = c->em_on && (c->flags & GF_FLAG_EM_BACKACTION_P);
    for (int i = 0; i < 3; i++) a.ecell[i] = em_p ? c->Ecell[i] : nullptr;
    a.code = c->d_code; a.mats = c->d_mats;
    double** ps = fe_pstore(c, in_sel);
    double** qs = fe_qstore(c, in_sel);
    double** po = fe_pstore(c, out_sel);
    double** qo = fe_qstore(c, out_sel);
    for (int i = 0; i < 3; i++) {
        a.pstore[i] = ps[i]; a.qstore[i] = qs[i]; a.pbase[i] = c->p[i]; a.qbase[i] = c->q[i];
        a.pacc[i] = c->pacc[i]; a.qacc[i] = c->qacc[i]; a.pout[i] = po[i]; a.qout[i] = qo[i];
        a.estat[i] = c->estat[i];
        a.flexo[i] = c->d_flexo ? c->flexo_force[i] : nullptr;
    }
    a.nx = c->nx; a.ny = c->ny; a.nz = c->nz; a.dx = c->dx; a.dy = c->dy; a.dz = c->dz; a.dt = c->dt;
    std::memcpy(a.Tg2c, c->Tg2c, sizeof(a.Tg2c));
    std::memcpy(a.Tc2g, c->Tc2g, sizeof(a.Tc2g));
    a.rotate = (c->flags & GF_FLAG_ORIENT_ROTATE) ? 1 : 0;
    a.stage = stage;
    return a;
}

}  // namespace gf

extern "C" {

gf_status gf_mag_stage(gf_ctx* c, int stage) {
    if (!c) return GF_ERR_ARG;
    if (stage < 1 || stage > 4) return set_error(c, GF_ERR_ARG, "stage must be 1..4");
    if (!c->any_fm) return set_error(c, GF_ERR_STATE, "no ferromagnetic material in this context");
    if (stage != c->mag_next_stage)
        return set_error(c, GF_ERR_STATE, "FM stages must run in the order 1,2,3,4 (expected " +
                                              std::to_string(c->mag_next_stage) + ")");
    if (gf_status e = check_after_elastic(c, stage)) return e;
    if (gf_status e = check_before_em(c, stage)) return e;
    const int out_sel = stage == 4 ? 0 : (c->mag_store_sel == 1 ? 2 : 1);
    StageArgs a = fm_stage_args(c, stage, c->mag_store_sel, out_sel);
    fm_stage_drives(c, stage, a);
    FusedZArgs f{};
    if (c->demag.fused) {
        f.hspec = c->demag.spec_a; f.mspec = c->demag.spec_m; f.tw = c->demag.twz.dev;
        f.nh = c->nx * c->ny * c->demag.nzh;
        f.nzh = static_cast<int>(c->demag.nzh);
        int lz = 0;
        while ((int64_t{1} << lz) < c->nz) lz++;
        f.lognz = lz;
        f.P = c->demag.P;
    }
    {
        const std::string why = stage_io_conflict(a, c->demag.fused ? &f : nullptr);
        if (!why.empty()) return set_error(c, GF_ERR_STATE, "FM stage arguments alias: " + why);
    }
    if (stage == 4 && c->spin_on && !c->any_afm)
        if (gf_status e = spin_save(c, 0)) return e;
    hipEvent_t ev = nullptr;
    timer_begin(c, GF_K_MAG_STAGE, &ev);
    gf_status st;
    if (c->demag.fused) {
        st = hip_check(c, launch_mag_stage_fused(a, f, c->demag.h_in_spec, c->stream), "mag stage");
        if (!st) c->demag.m_in_spec = true;
    } else {
        st = hip_check(c, launch_mag_stage(a, c->n, c->stream, c->d_fm_ids, c->n_fm_ids), "mag stage");
    }
    timer_end(c, GF_K_MAG_STAGE, ev);
    if (st) return st;
    c->mag_in_sel = c->mag_store_sel;
    c->mag_store_sel = out_sel;
    c->mag_next_stage = stage == 4 ? 1 : stage + 1;
    // get_J_ISHE_RK* at the end of get_dm_RK* (magnetic_solver.cpp:226-228);
    // with an AFM it is recomputed after the AFM stage, whose value stands
    // at stage 4 without electrodynamics nothing reads the stage current:
    // the step-boundary one replaces it at once
    if (c->spin_on && !c->any_afm) return stage == 4 && !c->em_on ? spin_pump_step(c) : spin_pump(c);
    return GF_OK;
}

gf_status gf_afm_stage(gf_ctx* c, int stage) {
    if (!c) return GF_ERR_ARG;
    if (stage < 1 || stage > 4) return set_error(c, GF_ERR_ARG, "stage must be 1..4");
    if (!c->any_afm) return set_error(c, GF_ERR_STATE, "no antiferromagnetic material in this context");
    if (stage != c->afm_next_stage)
        return set_error(c, GF_ERR_STATE, "AFM stages must run in the order 1,2,3,4");
    if (gf_status e = check_after_elastic(c, stage)) return e;
    if (gf_status e = check_before_em(c, stage)) return e;
    const int out_sel = stage == 4 ? 0 : (c->afm_store_sel == 1 ? 2 : 1);
    AfmStageArgs a = afm_stage_args(c, stage, c->afm_store_sel, out_sel);
    // the AFM path refreshes the sinusoidal drive unconditionally
    // (magnetic_AFM_solver.cpp:61-64)
    const double t = stage == 1 ? c->time - c->dt : (stage == 4 ? c->time : c->time - c->dt * 0.5);
    for (int i = 0; i < 3; i++) {
        c->hext[i] = c->hext_stat[i] + c->hext_altr[i] * std::sin(2. * kPi * c->h_altr_freq[i] * t);
        a.hext[i] = c->hext[i];
    }
    {
        const std::string why = stage_io_conflict(a);
        if (!why.empty()) return set_error(c, GF_ERR_STATE, "AFM stage arguments alias: " + why);
    }
    if (stage == 4 && c->spin_on)
        if (gf_status e = spin_save(c, 1)) return e;
    hipEvent_t ev = nullptr;
    timer_begin(c, GF_K_AFM_STAGE, &ev);
    gf_status st = hip_check(c, launch_afm_stage(a, c->n, c->stream), "afm stage");
    timer_end(c, GF_K_AFM_STAGE, ev);
    if (st) return st;
    c->afm_in_sel = c->afm_store_sel;
    c->afm_store_sel = out_sel;
    c->afm_next_stage = stage == 4 ? 1 : stage + 1;
    // magnetic_AFM_solver.cpp:426-428: after the FM stage of the same stage
    if (c->spin_on) return stage == 4 && !c->em_on ? spin_pump_step(c) : spin_pump(c);
    return GF_OK;
}

gf_status gf_fe_stage(gf_ctx* c, int stage) {
    if (!c) return GF_ERR_ARG;
    if (stage < 1 || stage > 4) return set_error(c, GF_ERR_ARG, "stage must be 1..4");
    if (!c->any_fe) return set_error(c, GF_ERR_STATE, "no ferroelectric material in this context");
    if (stage != c->fe_next_stage) return set_error(c, GF_ERR_STATE, "FE stages must run in the order 1,2,3,4");
    if (gf_status e = check_after_elastic(c, stage)) return e;
    if (gf_status e = check_before_em(c, stage)) return e;
    const int out_sel = stage == 4 ? 0 : (c->fe_store_sel == 1 ? 2 : 1);
    FeStageArgs a = fe_stage_args(c, stage, c->fe_store_sel, out_sel);
    double** ps = fe_pstore(c, c->fe_store_sel);
    double** qs = fe_qstore(c, c->fe_store_sel);
    {
        const std::string why = stage_io_conflict(a);
        if (!why.empty()) return set_error(c, GF_ERR_STATE, "FE stage arguments alias: " + why);
    }
    // ferroelectric_solver.cpp:2002-2029 (drive time per stage)
    if (!(c->flags & GF_FLAG_PRESCRIBE_EEXT)) {
        const double t = stage == 1 ? c->time - c->dt : (stage == 4 ? c->time : c->time - c->dt * 0.5);
        for (int i = 0; i < 3; i++)
            c->eext[i] = c->eext_stat[i] + c->eext_altr[i] * std::sin(2. * kPi * c->e_altr_freq[i] * t);
    }
    for (int i = 0; i < 3; i++) a.eext[i] = c->eext[i];
    hipEvent_t ev = nullptr;
    gf_status st = GF_OK;
    if (c->d_flexo) {
        // the flexoelectric force of the stage store and of the strain the
        // elastic stage found for it (get_p_gradient, get_Dstrain_gradient)
        FeFlexoArgs f{};
        f.code = c->d_code; f.fm = c->d_flexo;
        for (int i = 0; i < 3; i++) {
            f.p[i] = ps[i]; f.out[i] = c->flexo_force[i];
            f.gyz[i] = c->el_on ? c->gyz[i] : nullptr;
            f.gxz[i] = c->el_on ? c->gxz[i] : nullptr;
            f.gxy[i] = c->el_on ? c->gxy[i] : nullptr;
            f.per[i] = c->per[i];
        }
        for (int i = 0; i < 6; i++) f.eg[i] = c->el_on ? c->eg[i] : nullptr;
        f.nx = c->nx; f.ny = c->ny; f.nz = c->nz; f.dx = c->dx; f.dy = c->dy; f.dz = c->dz;
        timer_begin(c, GF_K_OTHER, &ev);
        st = hip_check(c, launch_fe_flexo(f, c->stream), "flexo force");
        timer_end(c, GF_K_OTHER, ev);
    }
    if (!st && c->fe_defer && c->em_on) {
        // runs in the EM stage's launch, which comes next in this stage
        c->fe_args = a;
        c->fe_pending = true;
    } else {
        timer_begin(c, GF_K_FE_STAGE, &ev);
        if (!st) st = hip_check(c, launch_fe_stage(a, c->n, c->stream), "fe stage");
        timer_end(c, GF_K_FE_STAGE, ev);
    }
    if (st) return st;
    // dp = q_s dt of this stage feeds Ampere's law (get_dp_RK*, :1786-1832)
    for (int i = 0; i < 3; i++) c->fe_q_used[i] = qs[i];
    c->fe_store_sel = out_sel;
    c->fe_next_stage = stage == 4 ? 1 : stage + 1;
    return GF_OK;
}

gf_status gf_elastic_stage(gf_ctx* c, int stage) {
    if (!c) return GF_ERR_ARG;
    if (stage < 1 || stage > 4) return set_error(c, GF_ERR_ARG, "stage must be 1..4");
    if (!c->el_on) return set_error(c, GF_ERR_STATE, "elastodynamics is off in this context");
    if (stage != c->el_next_stage) return set_error(c, GF_ERR_STATE, "elastic stages must run in the order 1,2,3,4");
    if ((c->any_fm && c->mag_next_stage != stage) || (c->any_afm && c->afm_next_stage != stage) ||
        (c->any_fe && c->fe_next_stage != stage))
        return set_error(c, GF_ERR_STATE, "gf_elastic_stage must come first within its stage");
    const int rotate = (c->flags & GF_FLAG_ORIENT_ROTATE) ? 1 : 0;
    const bool mur = c->flags & GF_FLAG_1D_ABC;
    const bool gauss = c->flags & GF_FLAG_GAUSSIAN_PULSE;
    const int64_t nxy = c->nx * c->ny;
    const int64_t slot_size = static_cast<int64_t>(kHistPlanes) * 3 * nxy;

    gf_status st = GF_OK;
    // stage 4: u enters the boundary history before the final combine
    // (transfer_pointer, elastodynamics_AbsorbingBoundary.cpp:3-89)
    if (stage == 4) {
        c->hist_head = (c->hist_head + 1) % c->hist_slots;
        ElHistArgs h{};
        for (int i = 0; i < 3; i++) h.u[i] = c->u[i];
        h.slot = c->hist + c->hist_head * slot_size;
        h.nx = c->nx; h.ny = c->ny; h.nz = c->nz;
        const int32_t lp[kHistPlanes] = {1, 2, 3, 4};
        const int32_t mp[kHistPlanes] = {0, 1, static_cast<int32_t>(c->nz - 2), static_cast<int32_t>(c->nz - 1)};
        for (int p = 0; p < kHistPlanes; p++) h.planes[p] = mur ? mp[p] : lp[p];
        hipEvent_t ev = nullptr;
        timer_begin(c, GF_K_OTHER, &ev);
        st = hip_check(c, launch_el_history(h, c->stream), "boundary history");
        timer_end(c, GF_K_OTHER, ev);
    }
    const double t_stage = stage == 1 ? c->time - c->dt : (stage == 4 ? c->time : c->time - 0.5 * c->dt);
    // strain of the stage store (elastic_dynamic_solver.cpp:107-227), the
    // eigenstrain from the stores of this stage (:4-104) and the cells'
    // parts of their face stresses; then the stress divergence and the v, u
    // updates (:841-3057)
    if (!st) {
        ElStageArgs sa{};
        ElStrainArgs& a = sa.s;
        a.code = c->d_code;
        // the face gradients are kept only for the flexoelectric force
        const bool faces = c->d_flexo != nullptr;
        for (int i = 0; i < 3; i++) {
            a.us[i] = c->us[i];
            a.gyz[i] = faces ? c->gyz[i] : nullptr;
            a.gxz[i] = faces ? c->gxz[i] : nullptr;
            a.gxy[i] = faces ? c->gxy[i] : nullptr;
        }
        for (int i = 0; i < 6; i++) { a.eg[i] = c->eg[i]; a.ec[i] = c->ec[i]; }
        a.nx = c->nx; a.ny = c->ny; a.nz = c->nz; a.dx = c->dx; a.dy = c->dy; a.dz = c->dz;
        a.idx = 1. / c->dx; a.idy = 1. / c->dy; a.idz = 1. / c->dz;
        for (int i = 0; i < 3; i++) a.per[i] = c->per[i];
        a.rotate = rotate;
        std::memcpy(a.Tg2c, c->Tg2c, sizeof(a.Tg2c));
        ElEigenArgs& e = sa.e;
        e.code = c->d_code; e.mats = c->d_mats;
        double** fm = c->any_fm ? fm_store(c, c->mag_store_sel) : nullptr;
        double** s1 = c->any_afm ? afm1_store(c, c->afm_store_sel) : nullptr;
        double** s2 = c->any_afm ? afm2_store(c, c->afm_store_sel) : nullptr;
        double** ps = c->any_fe ? fe_pstore(c, c->fe_store_sel) : nullptr;
        for (int i = 0; i < 3; i++) {
            e.m[i] = fm ? fm[i] : nullptr;
            e.m1[i] = s1 ? s1[i] : nullptr;
            e.m2[i] = s2 ? s2[i] : nullptr;
            e.p[i] = ps ? ps[i] : nullptr;
        }
        for (int i = 0; i < 6; i++) {
            e.e0g[i] = c->e0g[i]; e.e0c[i] = c->e0c[i];
            e.e0t0[i] = c->has_e0t0 ? c->e0t0[i] : nullptr;
        }
        e.flexo = c->any_fe ? c->d_flexo : nullptr;
        e.n = c->n;
        e.nx = c->nx; e.ny = c->ny; e.nz = c->nz; e.dx = c->dx; e.dy = c->dy; e.dz = c->dz;
        e.rotate = rotate;
        std::memcpy(e.Tg2c, c->Tg2c, sizeof(e.Tg2c));
        std::memcpy(e.Tc2g, c->Tc2g, sizeof(e.Tc2g));
        sa.eig = (c->flags & GF_FLAG_ELASTO_BACKACTION) ? 1 : 0;
        // only a ferroelectric stage reads the eigenstrain back (the next
        // elastic stage recomputes it before any other use)
        sa.e0_out = (c->any_fe && strain_coupled(c)) ? 1 : 0;
        for (int i = 0; i < 3; i++) { sa.px[i] = c->epx[i]; sa.py[i] = c->epy[i]; sa.pz[i] = c->epz[i]; }
        ElForceArgs& f = sa.f;
        f.code = c->d_code; f.mats = c->d_mats;
        f.nmat = static_cast<int32_t>(c->host_mats.size());
        for (int i = 0; i < 3; i++) {
            f.u[i] = c->u[i]; f.v[i] = c->v[i]; f.us[i] = c->us_alt[i]; f.vs[i] = c->vs[i];
            f.accu[i] = c->accu[i]; f.accv[i] = c->accv[i]; f.fstore[i] = c->fstore[i];
        }
        f.nx = c->nx; f.ny = c->ny; f.nz = c->nz; f.dx = c->dx; f.dy = c->dy; f.dz = c->dz; f.dt = c->dt;
        f.idx = 1. / c->dx; f.idy = 1. / c->dy; f.idz = 1. / c->dz;
        for (int i = 0; i < 3; i++) f.per[i] = c->per[i];
        f.stage = stage;
        f.trac_k = -1;
        if (gauss && c->src_type == 3 && c->src_comp >= 0) {
            // Gaussian traction at the stage time (:1190-1224)
            const double sg = c->src_sigma;
            const double tv = t_stage - 5.0 * sg;
            f.trac_k = c->src_k;
            f.trac_row = c->src_comp;
            f.trac_amp = c->src_amp * std::exp(-1. * std::pow(tv, 2.) / (2. * std::pow(sg, 2.)));
        }
        hipEvent_t ev = nullptr;
        timer_begin(c, GF_K_ELASTIC_STAGE, &ev);
        st = hip_check(c, launch_el_stage(sa, c->stream), "elastic stage");
        timer_end(c, GF_K_ELASTIC_STAGE, ev);
        // the store just written is the stage's u store from here on
        if (!st) std::swap(c->us, c->us_alt);
    }
    // absorbing boundary and displacement source
    // (elastodynamics_AbsorbingBoundary.cpp:92-230, elastic_dynamic_solver.cpp:3060-3284)
    if (!st) {
        const int kind = stage < 3 ? 0 : (stage == 3 ? 1 : 2);
        ElBoundaryArgs b{};
        b.code = c->d_code;
        for (int i = 0; i < 3; i++) { b.u[i] = c->u[i]; b.us[i] = c->us[i]; }
        for (int j = 0; j < 4; j++) {
            const int slot = ((c->hist_head - j) % c->hist_slots + c->hist_slots) % c->hist_slots;
            b.t[j] = c->hist + slot * slot_size;
        }
        b.nx = c->nx; b.ny = c->ny; b.nz = c->nz;
        b.kind = kind;
        b.liao = mur ? 0 : 1;
        b.mur_bottom = (c->flags & GF_FLAG_1D_ABC_ELAST_ONLYTOP) ? 0 : 1;
        b.w = c->w_liao;
        const double num = kind == 0 ? 2. * c->dz : c->dz;
        const double dt = c->dt;
        b.cb_l = (num - c->vl_first * dt) / (num + c->vl_first * dt);
        b.cb_t = (num - c->vt_first * dt) / (num + c->vt_first * dt);
        b.ct_l = (num - c->vl_last * dt) / (num + c->vl_last * dt);
        b.ct_t = (num - c->vt_last * dt) / (num + c->vt_last * dt);
        b.src_k = -1;
        if (gauss && c->src_comp >= 0 && c->src_k >= 0 && c->src_k < c->nz) {
            const double tsrc = kind == 0 ? c->time - 0.5 * c->dt : c->time;
            const double sg = c->src_sigma, A = c->src_amp;
            const double vsp = c->src_comp == 2 ? c->vl_first : c->vt_first;
            if (c->src_type == 1 && c->time < 12.5 * sg) {
                b.src_k = c->src_k;
                b.src_val = A * 1.6487210 * sg * vsp * std::exp(-1. * std::pow((tsrc - 5.0 * sg), 2.) / (2. * sg * sg));
            } else if (c->src_type == 2) {
                b.src_k = c->src_k;
                if (c->src_cycles == 0 || tsrc <= static_cast<double>(c->src_cycles) / sg)
                    b.src_val = A * vsp / 2. / kPi / sg * (std::cos(2. * kPi * sg * tsrc) - 1);
                else
                    b.src_val = 0.;
            }
            b.src_comp = c->src_comp;
        }
        hipEvent_t ev = nullptr;
        timer_begin(c, GF_K_OTHER, &ev);
        st = hip_check(c, launch_el_boundary(b, c->stream), "elastic boundary");
        timer_end(c, GF_K_OTHER, ev);
    }
    if (st) return st;
    c->el_next_stage = stage == 4 ? 1 : stage + 1;
    return GF_OK;
}

gf_status gf_em_stage(gf_ctx* c, int stage) {
    if (!c) return GF_ERR_ARG;
    if (stage < 1 || stage > 4) return set_error(c, GF_ERR_ARG, "stage must be 1..4");
    if (!c->em_on) return set_error(c, GF_ERR_STATE, "the electrodynamics is off in this context");
    if (stage != c->em_next_stage) return set_error(c, GF_ERR_STATE, "EM stages must run in the order 1,2,3,4");
    const int done = stage == 4 ? 1 : stage + 1;
    if ((c->any_fm && c->mag_next_stage != done) || (c->any_afm && c->afm_next_stage != done) ||
        (c->any_fe && c->fe_next_stage != done) || (c->el_on && c->el_next_stage != done))
        return set_error(c, GF_ERR_STATE, "gf_em_stage comes after the other stage calls of its stage");
    const int64_t nx = c->nx, ny = c->ny, nz = c->nz;
    const double t_stage = stage == 1 ? c->time - c->dt : (stage == 4 ? c->time : c->time - 0.5 * c->dt);
    const int out_e = stage == 4 ? 0 : (c->e_sel == 1 ? 2 : 1);
    const int out_h = stage == 4 ? 0 : (c->h_sel == 1 ? 2 : 1);
    double** es = e_store(c, c->e_sel);
    double** eo = e_store(c, out_e);
    double** hs = h_store(c, c->h_sel);
    double** ho = h_store(c, out_h);
    gf_status st = GF_OK;
    hipEvent_t ev = nullptr;

    // free-current box at the stage time (update_Jf_input*, :3885-4315): its
    // value, which the Ampere launch evaluates per cell
    int32_t jf_comp = -1;
    double jf_value = 0.;
    const double* jf_spatial = nullptr;
    if (c->flags & GF_FLAG_JF_INPUT) {
        const double t = t_stage, f = c->jf_freq, amp = c->jf_amp, sg = c->jf_sigma;
        bool write = true;
        if (c->jf_type < 3) {
            const double tv = std::sin(2. * kPi * f * t);
            jf_value = amp * tv;
            jf_spatial = c->jf_type == 0 ? nullptr : c->jf_spatial;
        } else if (c->jf_type == 3) {
            const double tv = t <= static_cast<double>(c->jf_cycles) / f ? std::sin(2. * kPi * f * t) : 0.;
            jf_value = amp * tv;
        } else if (c->jf_type == 4) {
            const double tv = t - 5. * f;
            jf_value = amp * std::exp(-1. * std::pow(tv, 2.) / (2. * std::pow(f, 2.)));
        } else if (c->jf_type == 5) {
            const double tv = t - 5. * f;
            jf_value = amp * std::sqrt(std::exp(1.)) * tv / f * std::exp(-1. * std::pow(tv, 2.) / (2. * std::pow(f, 2.)));
        } else if (c->jf_type == 6) {
            const double tv = t - 5. * sg;
            jf_value = amp * std::exp(-1. * std::pow(tv, 2.) / (2. * std::pow(sg, 2.))) * std::sin(2. * kPi * f * t);
        } else {
            write = false;
        }
        if (write) jf_comp = c->jf_comp;
    }
    EmArgs a{};
    // stage 4: the E update keeps the planes the Mur walls read before the
    // combine (transfer_pointer, EMdynamic_AbsorbingBoundary.cpp:3-56)
    const bool hist = stage == 4 && !c->em_periodic_all;
    for (int i = 0; i < 3; i++) a.t1[i] = hist ? c->Et1[i] : nullptr;
    a.code = c->d_code; a.mats = c->d_mats;
    for (int i = 0; i < 3; i++) {
        a.es[i] = es[i]; a.hs[i] = hs[i];
        a.eb[i] = c->Eb[i]; a.eo[i] = eo[i]; a.eacc[i] = c->Eacc[i];
        a.hb[i] = c->Hb[i]; a.ho[i] = ho[i]; a.hacc[i] = c->Hacc[i];
        a.q[i] = ((c->flags & GF_FLAG_EM_FROM_P) && c->any_fe) ? c->fe_q_used[i] : nullptr;
        const bool from_m = c->flags & GF_FLAG_EM_FROM_M;
        a.dm[i] = from_m ? c->em_dm[i] : nullptr;
        a.dma1[i] = from_m ? c->em_dm1[i] : nullptr;
        a.dma2[i] = from_m ? c->em_dm2[i] : nullptr;
    }
    a.jishe[0] = c->spin_on ? c->jis[0] : nullptr;
    a.jishe[1] = c->spin_on ? c->jis[1] : nullptr;
    for (int q = 0; q < 9; q++) a.jps[q] = c->jp_on ? c->Jps[q] : nullptr;
    a.jf_comp = jf_comp; a.jf_value = jf_value; a.jf_spatial = jf_spatial;
    for (int i = 0; i < 3; i++) { a.jf_lo[i] = c->jf_lo[i]; a.jf_hi[i] = c->jf_hi[i]; }
    a.nx = nx; a.ny = ny; a.nz = nz; a.dx = c->dx; a.dy = c->dy; a.dz = c->dz; a.dt = c->dt;
    for (int i = 0; i < 3; i++) a.per[i] = c->per[i];
    a.stage = stage;
    // Ampere and Faraday with the RK updates (both read the stage-s stores),
    // with a deferred ferroelectric stage in the same launch
    if (!st) {
        timer_begin(c, GF_K_EM_E, &ev);
        if (c->fe_pending) st = hip_check(c, launch_fe_em(c->fe_args, a, c->stream), "FE stage, EM E and H");
        else st = hip_check(c, launch_em_eh(a, c->stream), "EM E and H");
        timer_end(c, GF_K_EM_E, ev);
        c->fe_pending = false;
    }
    // Jp from the cell E of the stage store (get_dJp_RK*, update_Jp*)
    if (!st && c->jp_on) {
        EmJpArgs j{};
        j.code = c->d_code; j.mats = c->d_mats;
        for (int i = 0; i < 3; i++) j.ecell[i] = c->Ecell[i];
        for (int q = 0; q < 9; q++) { j.jp[q] = c->Jp[q]; j.jps[q] = c->Jps[q]; j.jacc[q] = c->Jacc[q]; }
        j.n = c->n; j.dt = c->dt; j.stage = stage;
        timer_begin(c, GF_K_OTHER, &ev);
        st = hip_check(c, launch_em_jp(j, c->stream), "EM Jp");
        timer_end(c, GF_K_OTHER, ev);
    }
    const int kind = stage < 3 ? 0 : (stage == 3 ? 1 : 2);
    // hard plane source (update_planeEM*, :3829-3883)
    if (!st && (c->flags & GF_FLAG_PLANE_EM)) {
        EmMiscArgs p{};
        const double t = kind == 0 ? c->time - 0.5 * c->dt : c->time;
        for (int i = 0; i < 3; i++) {
            p.b[i] = c->Eb[i]; p.s[i] = eo[i];
            p.plane[i] = c->plane_E[i] * std::sin(2. * kPi * c->plane_f[i] * t);
        }
        p.nx = nx; p.ny = ny; p.nz = nz; p.plane_z = c->plane_z; p.kind = kind;
        timer_begin(c, GF_K_OTHER, &ev);
        st = hip_check(c, launch_em_plane(p, c->stream), "EM plane source");
        timer_end(c, GF_K_OTHER, ev);
    }
    // Mur or PEC walls, x then y then z (EMdynamic_AbsorbingBoundary.cpp:58-874);
    // with only z walls (and no PEC edges after them) the cells launch
    // applies them to the edges it reads
    const bool zwall_in_cells = !c->em_periodic_all && c->per[0] && c->per[1] && !c->per[2] && !c->em_pec_all &&
                                c->nz >= 2;
    EmWallArgs zw{};
    if (!c->em_periodic_all) {
        for (int axis = 0; axis < 3 && !st; axis++) {
            if (c->per[axis]) continue;
            EmWallArgs w{};
            for (int i = 0; i < 3; i++) { w.b[i] = c->Eb[i]; w.s[i] = eo[i]; w.t1[i] = c->Et1[i]; }
            w.nx = nx; w.ny = ny; w.nz = nz; w.axis = axis; w.kind = kind;
            w.pec = axis < 2 ? ((c->flags & GF_FLAG_PEC_XY) ? 1 : 0) : ((c->flags & GF_FLAG_PEC_Z) ? 1 : 0);
            w.onlytop = ((c->flags & GF_FLAG_1D_ABC) && (c->flags & GF_FLAG_1D_ABC_EM_ONLYTOP)) ? 1 : 0;
            if (axis == 0) w.cf = kind == 0 ? c->cf_x_half : c->cf_x;
            else if (axis == 1) w.cf = kind == 0 ? c->cf_y_half : c->cf_y;
            else { w.cf = kind == 0 ? c->cf_top_half : c->cf_top; w.cb = kind == 0 ? c->cf_bot_half : c->cf_bot; }
            if (axis == 2 && zwall_in_cells) {
                zw = w;
                continue;
            }
            timer_begin(c, GF_K_OTHER, &ev);
            st = hip_check(c, launch_em_wall(w, c->stream), "EM walls");
            timer_end(c, GF_K_OTHER, ev);
        }
    }
    // PEC edges (update_DE_PEC, :4526-4556)
    if (!st && c->em_pec_all) {
        EmMiscArgs p{};
        for (int i = 0; i < 3; i++) { p.b[i] = c->Eb[i]; p.s[i] = eo[i]; p.pec[i] = c->pec[i]; }
        p.nx = nx; p.ny = ny; p.nz = nz;
        timer_begin(c, GF_K_OTHER, &ev);
        st = hip_check(c, launch_em_pec(p, c->stream), "EM PEC");
        timer_end(c, GF_K_OTHER, ev);
    }
    if (st) return st;
    c->e_sel = out_e;
    c->h_sel = out_h;
    st = em_cells(c, true, true, zwall_in_cells ? &zw : nullptr);
    // Ampere's law of stage 4 has read the stage current: now the step one
    if (!st && stage == 4 && c->spin_on) st = spin_pump_step(c);
    if (st) return st;
    c->em_next_stage = stage == 4 ? 1 : stage + 1;
    return GF_OK;
}

gf_status gf_demag(gf_ctx* c) {
    if (!c) return GF_ERR_ARG;
    if ((c->flags & GF_FLAG_ELASTOSTATIC) && !(c->el_on && c->t0_fixed)) {
        // the static strain of the current magnet and ferroelectric stores:
        // before the loop without elastodynamics (main.cpp:214), once at
        // start-up with it unless a static strain or eigenstrain was read
        // (main.cpp:148-158)
        if (gf_status st = elastostatic_static(c)) return st;
        c->t0_fixed = true;
    }
    if ((c->flags & GF_FLAG_ELECTROSTATICS) && c->any_fe) {
        // initial get_E_static (main.cpp:120-131)
        if (!c->eplan.ready) return set_error(c, GF_ERR_STATE, "FFT electrostatics was not set up");
        if (gf_status st = elec_run(c, fe_pstore(c, c->fe_store_sel))) return st;
    }
    return mag_statics(c);
}

gf_status gf_prescribe_m(gf_ctx* c) {
    if (!c) return GF_ERR_ARG;
    if (!(c->flags & GF_FLAG_PRESCRIBED_M)) return set_error(c, GF_ERR_STATE, "GF_FLAG_PRESCRIBED_M is not set");
    if (!c->any_fm) return set_error(c, GF_ERR_STATE, "no ferromagnetic material in this context");
    if (c->mag_next_stage != 1) return set_error(c, GF_ERR_STATE, "gf_prescribe_m needs a step boundary");
    double v[3];
    prescribed_cone(c, c->time, v);
    gf_status st = hip_check(c, launch_prescribe_plane(c->m, c->d_code, c->nx, c->ny, c->nz, v, c->stream),
                             "prescribe m");
    if (st) return st;
    // initialize_device then copies m into the store (magnetic_initialize.cpp:1081-1086)
    c->mag_store_sel = 0;
    c->demag.m_in_spec = false;
    return hip_check(c, hipStreamSynchronize(c->stream), "prescribe m sync");
}

// main.cpp:243-483 for the magnetic physics
gf_status gf_step(gf_ctx* c, int64_t nsteps) {
    if (!c) return GF_ERR_ARG;
    if (nsteps < 0) return set_error(c, GF_ERR_ARG, "nsteps must be >= 0");
    if ((c->any_fm && c->mag_next_stage != 1) || (c->any_afm && c->afm_next_stage != 1) ||
        (c->any_fe && c->fe_next_stage != 1) || (c->el_on && c->el_next_stage != 1) ||
        (c->em_on && c->em_next_stage != 1))
        return set_error(c, GF_ERR_STATE, "gf_step called in the middle of a step");
    // the FE stage runs inside the EM stage's launch (they are independent)
    c->fe_defer = c->any_fe && c->em_on;
    for (int64_t s = 0; s < nsteps; s++) {
        gf_status st = gf_time_advance(c);
        for (int stage = 1; stage <= 4 && !st; stage++) {
            if (c->el_on) st = gf_elastic_stage(c, stage);
            if (!st && c->any_fm) st = gf_mag_stage(c, stage);
            if (!st && c->any_afm) st = gf_afm_stage(c, stage);
            if (!st && c->any_fe) st = gf_fe_stage(c, stage);
            if (!st && c->em_on) st = gf_em_stage(c, stage);
            if (!st) st = statics_after_update(c);
        }
        // the static strain after the step without elastodynamics (main.cpp:472-474)
        if (!st && (c->flags & GF_FLAG_ELASTOSTATIC) && !c->el_on) st = elastostatic_static(c);
        if (st) {
            c->fe_defer = false;
            c->fe_pending = false;
            return st;
        }
    }
    c->fe_defer = false;
    return GF_OK;
}

gf_status gf_average(gf_ctx* c, int id, double out[3]) {
    if (!c || !out) return GF_ERR_ARG;
    gf_status st;
    if (id != GF_FIELD_M && id != GF_FIELD_M_AFM1 && id != GF_FIELD_M_AFM2 && id != GF_FIELD_P)
        return set_error(c, GF_ERR_ARG, "average is defined for m, m_AFM1, m_AFM2, p");
    double** f = field_for(c, id, false, &st);
    if (!f) return st;
    const double* ff[3] = {f[0], f[1], f[2]};
    st = hip_check(c, launch_sum3(ff, c->n, c->red_partial, c->red_blocks, c->stream), "sum");
    if (st) return st;
    std::vector<double> part(static_cast<size_t>(3 * c->red_blocks));
    st = hip_check(c, hipMemcpyAsync(part.data(), c->red_partial, sizeof(double) * part.size(), hipMemcpyDeviceToHost,
                                     c->stream), "sum copy");
    if (!st) st = hip_check(c, hipStreamSynchronize(c->stream), "sum sync");
    if (st) return st;
    const double cnt = static_cast<double>(id == GF_FIELD_M ? c->nfm : (id == GF_FIELD_P ? c->nfe : c->nafm));
    for (int a = 0; a < 3; a++) {
        double s = 0.;
        for (int64_t b = 0; b < c->red_blocks; b++) s += part[static_cast<size_t>(3 * b + a)];
        out[a] = s / cnt;
    }
    return GF_OK;
}

gf_status gf_get_time(const gf_ctx* c, double* t) {
    if (!c || !t) return GF_ERR_ARG;
    *t = c->time;
    return GF_OK;
}

gf_status gf_set_time(gf_ctx* c, double t) {
    if (!c) return GF_ERR_ARG;
    if ((c->any_fm && c->mag_next_stage != 1) || (c->any_afm && c->afm_next_stage != 1) ||
        (c->any_fe && c->fe_next_stage != 1) || (c->el_on && c->el_next_stage != 1) ||
        (c->em_on && c->em_next_stage != 1))
        return set_error(c, GF_ERR_STATE, "gf_set_time needs a step boundary");
    c->time = t;
    c->prescribe_idx = 0;
    c->prescribe_e_idx = 0;
    return GF_OK;
}

gf_status gf_sync(gf_ctx* c) {
    if (!c) return GF_ERR_ARG;
    return hip_check(c, hipStreamSynchronize(c->stream), "sync");
}

gf_status gf_get_stream(gf_ctx* c, void** stream) {
    if (!c || !stream) return GF_ERR_ARG;
    *stream = c->stream;
    return GF_OK;
}

gf_status gf_set_kernel_timing(gf_ctx* c, int enable) {
    if (!c) return GF_ERR_ARG;
    drain_timers(c);
    c->timing = enable != 0;
    for (int k = 0; k < GF_K_COUNT; k++) { c->kernel_ms[k] = 0.; c->kernel_launches[k] = 0; }
    return GF_OK;
}

gf_status gf_kernel_times(gf_ctx* c, double ms[GF_K_COUNT], int64_t launches[GF_K_COUNT]) {
    if (!c || !ms || !launches) return GF_ERR_ARG;
    drain_timers(c);
    for (int k = 0; k < GF_K_COUNT; k++) { ms[k] = c->kernel_ms[k]; launches[k] = c->kernel_launches[k]; }
    return GF_OK;
}

}  // extern "C"

// ---- checkpoint / restart --------------------------------------------------
//
// One host image of everything a step boundary carries into the next step:
// a header (grid, flags, dt, a hash of the material ids, t, the drive-table
// cursors, the store selectors and the boundary-history head), the material
// table, every persistent device array under a tag, and a checksum.  The
// reference restarts from its text inputs (inoutput.cpp:11-327: m, m_AFM,
// Hstat, p, q, Estat, E and H on the Yee grid, u, v, the elastic force
// store, the strains, Jp) at seven digits and t = 0; the image holds them at
// full precision together with t, the Mur and Liao histories and the H
// z-spectrum of the fused magnetostatics, so a restarted run continues bit
// for bit.

namespace {

constexpr char kStateMagic[8] = {'G', 'F', 'S', 'T', 'A', 'T', 'E', '1'};
constexpr uint32_t kStateVersion = 1;
constexpr uint64_t kHashSeed = 14695981039346656037ull;

struct StateHeader {
    char magic[8];
    uint32_t version;
    uint32_t header_bytes;
    int64_t nx, ny, nz;
    double dx, dy, dz, dt;
    int32_t per[3];
    uint32_t flags;
    uint64_t cell_hash;
    double time;
    int64_t prescribe_idx, prescribe_e_idx;
    int32_t hist_head, hist_slots;
    int32_t mag_sel, afm_sel, fe_sel, e_sel, h_sel;
    uint32_t bits;
    int32_t n_materials;
    int32_t n_arrays;
    uint64_t material_bytes;
};

struct StateArrayHeader {
    uint32_t tag;
    uint32_t reserved;
    int64_t count;   // doubles
};

enum : uint32_t {
    kBitHstatWritten = 1u, kBitJpOn = 2u, kBitJpUploaded = 4u, kBitHasHnu = 8u,
    kBitHInSpec = 16u, kBitHRealValid = 32u, kBitMInSpec = 64u, kBitHasE0t0 = 128u, kBitEt0Dyn = 256u,
    kBitT0Fixed = 512u
};

// FNV-1a over 64-bit words, then over the tail bytes.
uint64_t hash_bytes(uint64_t h, const void* p, size_t n) {
    constexpr uint64_t kPrime = 1099511628211ull;
    const unsigned char* b = static_cast<const unsigned char*>(p);
    size_t i = 0;
    for (; i + 8 <= n; i += 8) {
        uint64_t w;
        std::memcpy(&w, b + i, 8);
        h = (h ^ w) * kPrime;
    }
    for (; i < n; i++) h = (h ^ b[i]) * kPrime;
    return h;
}

struct StateArray {
    uint32_t tag;
    double* p;
    int64_t count;
};

// The persistent device arrays for the given selectors and bits, tagged
// group << 8 | component.  Stage stores appear only when a selector points
// away from the state (an uploaded store); spectra only while they hold the
// current field.
struct StateSel {
    int mag, afm, fe, e, h;
};

StateSel current_sel(const gf_ctx* c) {
    return {c->mag_store_sel, c->afm_store_sel, c->fe_store_sel, c->e_sel, c->h_sel};
}

std::vector<StateArray> state_arrays(gf_ctx* c, const StateSel& sel, uint32_t bits) {
    std::vector<StateArray> v;
    const int64_t n = c->n;
    auto add = [&](uint32_t group, double* const* p, int k, int64_t count) {
        for (int i = 0; i < k; i++)
            if (p[i]) v.push_back({group << 8 | static_cast<uint32_t>(i), p[i], count});
    };
    if (c->any_fm) {
        add(1, c->m, 3, n);
        if (sel.mag != 0) add(2, fm_store(c, sel.mag), 3, n);
    }
    if (c->any_afm) {
        add(3, c->m1, 3, n);
        add(4, c->m2, 3, n);
        if (sel.afm != 0) { add(5, afm1_store(c, sel.afm), 3, n); add(6, afm2_store(c, sel.afm), 3, n); }
    }
    if (c->any_fe) {
        add(7, c->p, 3, n);
        add(8, c->q, 3, n);
        if (sel.fe != 0) { add(9, fe_pstore(c, sel.fe), 3, n); add(10, fe_qstore(c, sel.fe), 3, n); }
        add(11, c->estat, 3, n);
    }
    if (c->el_on) {
        add(12, c->u, 3, n);
        add(13, c->v, 3, n);
        add(14, c->us, 3, n);
        add(15, c->vs, 3, n);
        add(16, c->fstore, 3, n);
        add(17, c->eg, 6, n);
        add(18, c->e0g, 6, n);
        if (c->flags & GF_FLAG_ORIENT_ROTATE) { add(19, c->ec, 6, n); add(20, c->e0c, 6, n); }
        if (c->hist)
            v.push_back({21u << 8, c->hist, static_cast<int64_t>(c->hist_slots) * kHistPlanes * 3 * c->nx * c->ny});
    }
    if (c->em_on) {
        for (int a = 0; a < 3; a++) {
            v.push_back({22u << 8 | static_cast<uint32_t>(a), c->Eb[a], c->nE[a]});
            v.push_back({23u << 8 | static_cast<uint32_t>(a), c->Hb[a], c->nH[a]});
            v.push_back({24u << 8 | static_cast<uint32_t>(a), c->Et1[a], c->nE[a]});
        }
        for (int a = 0; a < 3 && sel.e != 0; a++)
            v.push_back({33u << 8 | static_cast<uint32_t>(a), e_store(c, sel.e)[a], c->nE[a]});
        for (int a = 0; a < 3 && sel.h != 0; a++)
            v.push_back({34u << 8 | static_cast<uint32_t>(a), h_store(c, sel.h)[a], c->nH[a]});
        add(25, c->Ecell, 3, n);
        add(26, c->Hcell, 3, n);
        add(28, c->Jp, 9, n);
        add(35, c->Jps, 9, n);
    }
    add(29, c->hstat, 3, n);
    if (bits & kBitHasHnu) add(30, c->hnu, 3, n);
    add(31, c->hdrive, 3, n);
    add(32, c->e_t0, 6, n);
    if (bits & kBitHasE0t0) add(37, c->e0t0, 6, n);
    add(36, c->jis, 2, n);
    const DemagPlan& d = c->demag;
    if (d.ready) {
        const int64_t spec = 2 * 3 * c->nx * c->ny * d.nzh;
        if (bits & kBitHInSpec) v.push_back({40u << 8, reinterpret_cast<double*>(d.spec_a), spec});
        if (bits & kBitMInSpec) v.push_back({41u << 8, reinterpret_cast<double*>(d.spec_m), spec});
    }
    return v;
}

bool at_step_boundary(const gf_ctx* c) {
    return !((c->any_fm && c->mag_next_stage != 1) || (c->any_afm && c->afm_next_stage != 1) ||
             (c->any_fe && c->fe_next_stage != 1) || (c->el_on && c->el_next_stage != 1) ||
             (c->em_on && c->em_next_stage != 1));
}

uint32_t state_bits(const gf_ctx* c) {
    uint32_t b = 0;
    if (c->hstat_written) b |= kBitHstatWritten;
    if (c->jp_on) b |= kBitJpOn;
    if (c->jp_uploaded) b |= kBitJpUploaded;
    if (c->has_hnu) b |= kBitHasHnu;
    if (c->has_e0t0) b |= kBitHasE0t0;
    if (c->has_et0_dyn) b |= kBitEt0Dyn;
    if (c->t0_fixed) b |= kBitT0Fixed;
    if (c->demag.h_in_spec) b |= kBitHInSpec;
    if (c->demag.h_real_valid) b |= kBitHRealValid;
    if (c->demag.m_in_spec) b |= kBitMInSpec;
    return b;
}

size_t state_bytes(const gf_ctx* c, const std::vector<StateArray>& arr) {
    size_t b = sizeof(StateHeader) + sizeof(gf_material) * c->host_mats.size() + sizeof(uint64_t);
    for (const auto& a : arr) b += sizeof(StateArrayHeader) + sizeof(double) * static_cast<size_t>(a.count);
    return b;
}

uint64_t cell_hash(const gf_ctx* c) {
    return hash_bytes(kHashSeed, c->host_code.data(), sizeof(uint32_t) * c->host_code.size());
}

}  // namespace

extern "C" {

gf_status gf_state_size(gf_ctx* c, size_t* bytes) {
    if (!c || !bytes) return set_error(c, GF_ERR_ARG, "null argument");
    const auto arr = state_arrays(c, current_sel(c), state_bits(c));
    *bytes = state_bytes(c, arr);
    return GF_OK;
}

gf_status gf_state_save(gf_ctx* c, void* host, size_t bytes) {
    if (!c || !host) return set_error(c, GF_ERR_ARG, "null argument");
    if (!at_step_boundary(c)) return set_error(c, GF_ERR_STATE, "the state can be saved only at a step boundary");
    const uint32_t bits = state_bits(c);
    const auto arr = state_arrays(c, current_sel(c), bits);
    const size_t need = state_bytes(c, arr);
    if (bytes != need) return set_error(c, GF_ERR_ARG, "the state image takes " + std::to_string(need) + " bytes");
    unsigned char* out = static_cast<unsigned char*>(host);
    StateHeader h{};
    std::memcpy(h.magic, kStateMagic, sizeof(kStateMagic));
    h.version = kStateVersion;
    h.header_bytes = sizeof(StateHeader);
    h.nx = c->nx; h.ny = c->ny; h.nz = c->nz;
    h.dx = c->dx; h.dy = c->dy; h.dz = c->dz; h.dt = c->dt;
    for (int a = 0; a < 3; a++) h.per[a] = c->per[a];
    h.flags = c->flags;
    h.cell_hash = cell_hash(c);
    h.time = c->time;
    h.prescribe_idx = c->prescribe_idx;
    h.prescribe_e_idx = c->prescribe_e_idx;
    h.hist_head = c->hist_head;
    h.hist_slots = c->hist_slots;
    h.mag_sel = c->mag_store_sel; h.afm_sel = c->afm_store_sel; h.fe_sel = c->fe_store_sel;
    h.e_sel = c->e_sel; h.h_sel = c->h_sel;
    h.bits = bits;
    h.n_materials = static_cast<int32_t>(c->host_mats.size());
    h.n_arrays = static_cast<int32_t>(arr.size());
    h.material_bytes = sizeof(gf_material) * c->host_mats.size();
    size_t o = 0;
    std::memcpy(out + o, &h, sizeof(h));
    o += sizeof(h);
    std::memcpy(out + o, c->host_mats.data(), h.material_bytes);
    o += h.material_bytes;
    gf_status st = GF_OK;
    for (const auto& a : arr) {
        StateArrayHeader ah{a.tag, 0u, a.count};
        std::memcpy(out + o, &ah, sizeof(ah));
        o += sizeof(ah);
        const size_t b = sizeof(double) * static_cast<size_t>(a.count);
        if (!st) st = hip_check(c, hipMemcpyAsync(out + o, a.p, b, hipMemcpyDeviceToHost, c->stream), "state save");
        o += b;
    }
    if (!st) st = hip_check(c, hipStreamSynchronize(c->stream), "state save sync");
    if (st) return st;
    const uint64_t sum = hash_bytes(kHashSeed, out, o);
    std::memcpy(out + o, &sum, sizeof(sum));
    return GF_OK;
}

gf_status gf_state_load(gf_ctx* c, const void* host, size_t bytes) {
    if (!c || !host) return set_error(c, GF_ERR_ARG, "null argument");
    if (!at_step_boundary(c)) return set_error(c, GF_ERR_STATE, "the state can be loaded only at a step boundary");
    const unsigned char* in = static_cast<const unsigned char*>(host);
    StateHeader h{};
    if (bytes < sizeof(h) + sizeof(uint64_t)) return set_error(c, GF_ERR_ARG, "the state image is truncated");
    std::memcpy(&h, in, sizeof(h));
    if (std::memcmp(h.magic, kStateMagic, sizeof(kStateMagic)) != 0)
        return set_error(c, GF_ERR_ARG, "not a goferro state image");
    if (h.version != kStateVersion || h.header_bytes != sizeof(StateHeader))
        return set_error(c, GF_ERR_ARG, "state image version " + std::to_string(h.version) + " is not supported");
    uint64_t sum = 0;
    std::memcpy(&sum, in + bytes - sizeof(sum), sizeof(sum));
    if (hash_bytes(kHashSeed, in, bytes - sizeof(sum)) != sum)
        return set_error(c, GF_ERR_ARG, "the state image checksum does not match (corrupt or truncated)");
    if (h.nx != c->nx || h.ny != c->ny || h.nz != c->nz)
        return set_error(c, GF_ERR_ARG, "the state image is of another grid");
    if (std::memcmp(&h.dx, &c->dx, sizeof(double)) || std::memcmp(&h.dy, &c->dy, sizeof(double)) ||
        std::memcmp(&h.dz, &c->dz, sizeof(double)) || std::memcmp(&h.dt, &c->dt, sizeof(double)))
        return set_error(c, GF_ERR_ARG, "the state image has other cell sizes or another dt");
    if (h.per[0] != c->per[0] || h.per[1] != c->per[1] || h.per[2] != c->per[2] || h.flags != c->flags)
        return set_error(c, GF_ERR_ARG, "the state image has other boundary conditions or flags");
    if (h.cell_hash != cell_hash(c)) return set_error(c, GF_ERR_ARG, "the state image has other material ids");
    if (h.n_materials != static_cast<int32_t>(c->host_mats.size()) ||
        h.material_bytes != sizeof(gf_material) * c->host_mats.size())
        return set_error(c, GF_ERR_ARG, "the state image has another material table");
    if (bytes < sizeof(h) + h.material_bytes + sizeof(uint64_t))
        return set_error(c, GF_ERR_ARG, "the state image is truncated");
    if (h.hist_slots != c->hist_slots || h.hist_head < 0 || (c->hist_slots > 0 && h.hist_head >= c->hist_slots))
        return set_error(c, GF_ERR_ARG, "the state image has another boundary history");
    auto sel_ok = [](int32_t s) { return s >= 0 && s <= 2; };
    if (!sel_ok(h.mag_sel) || !sel_ok(h.afm_sel) || !sel_ok(h.fe_sel) || !sel_ok(h.e_sel) || !sel_ok(h.h_sel))
        return set_error(c, GF_ERR_ARG, "the state image has invalid store selectors");
    if ((h.bits & (kBitHInSpec | kBitMInSpec)) && !c->demag.ready)
        return set_error(c, GF_ERR_ARG, "the state image holds spectra this context has no plan for");
    // the material table: equal, or applied record by record where allowed
    size_t o = sizeof(h);
    std::vector<gf_material> mats(c->host_mats.size());
    std::memcpy(mats.data(), in + o, h.material_bytes);
    o += h.material_bytes;
    // the array list this context expects for the image's selectors
    gf_status st = GF_OK;
    if ((h.bits & kBitHasHnu) && !c->hnu[0]) {
        st = alloc3(c, c->hnu, "hipMalloc hext_nonunif");
        if (st) return st;
    }
    if ((h.bits & kBitEt0Dyn) && !c->el_on)
        return set_error(c, GF_ERR_ARG, "the state image holds a static strain beside elastodynamics this context lacks");
    if ((h.bits & kBitEt0Dyn) && !c->e_t0[0]) {
        st = alloc3(c, c->e_t0, "hipMalloc static strain");
        if (!st) st = alloc3(c, c->e_t0 + 3, "hipMalloc static strain");
        if (st) return st;
    }
    if ((h.bits & kBitHasE0t0) && !c->el_on && !(c->flags & GF_FLAG_ELASTOSTATIC))
        return set_error(c, GF_ERR_ARG,
                         "the state image holds a static eigenstrain and this context has no elastodynamics or elastostatics");
    if ((h.bits & kBitHasE0t0) && !c->e0t0[0]) {
        st = alloc3(c, c->e0t0, "hipMalloc static eigenstrain");
        if (!st) st = alloc3(c, c->e0t0 + 3, "hipMalloc static eigenstrain");
        if (st) return st;
    }
    const auto arr = state_arrays(c, StateSel{h.mag_sel, h.afm_sel, h.fe_sel, h.e_sel, h.h_sel}, h.bits);
    if (h.n_arrays != static_cast<int32_t>(arr.size()) || bytes != state_bytes(c, arr))
        return set_error(c, GF_ERR_ARG, "the state image does not hold the arrays of this context");
    size_t scan = o;
    for (const auto& a : arr) {
        StateArrayHeader ah{};
        std::memcpy(&ah, in + scan, sizeof(ah));
        if (ah.tag != a.tag || ah.count != a.count)
            return set_error(c, GF_ERR_ARG, "the state image does not hold the arrays of this context");
        scan += sizeof(ah) + sizeof(double) * static_cast<size_t>(a.count);
    }
    for (size_t i = 0; i < mats.size() && !st; i++) {
        if (std::memcmp(&mats[i], &c->host_mats[i], sizeof(gf_material)) == 0) continue;
        st = gf_set_material(c, static_cast<int>(i + 1), &mats[i]);
    }
    if (st) return st;
    for (const auto& a : arr) {
        o += sizeof(StateArrayHeader);
        const size_t b = sizeof(double) * static_cast<size_t>(a.count);
        if (!st) st = hip_check(c, hipMemcpyAsync(a.p, in + o, b, hipMemcpyHostToDevice, c->stream), "state load");
        o += b;
    }
    if (!st) st = hip_check(c, hipStreamSynchronize(c->stream), "state load sync");
    if (st) return st;
    c->time = h.time;
    c->prescribe_idx = static_cast<int>(h.prescribe_idx);
    c->prescribe_e_idx = static_cast<int>(h.prescribe_e_idx);
    c->hist_head = h.hist_head;
    c->mag_store_sel = h.mag_sel; c->afm_store_sel = h.afm_sel; c->fe_store_sel = h.fe_sel;
    c->e_sel = h.e_sel; c->h_sel = h.h_sel;
    c->hstat_written = h.bits & kBitHstatWritten;
    c->jp_on = h.bits & kBitJpOn;
    c->jp_uploaded = h.bits & kBitJpUploaded;
    c->has_hnu = h.bits & kBitHasHnu;
    c->has_e0t0 = h.bits & kBitHasE0t0;
    c->has_et0_dyn = h.bits & kBitEt0Dyn;
    c->t0_fixed = h.bits & kBitT0Fixed;
    c->demag.h_in_spec = h.bits & kBitHInSpec;
    c->demag.h_real_valid = h.bits & kBitHRealValid;
    c->demag.m_in_spec = h.bits & kBitMInSpec;
    return GF_OK;
}

gf_status gf_state_write(gf_ctx* c, const char* path) {
    if (!c || !path) return set_error(c, GF_ERR_ARG, "null argument");
    size_t bytes = 0;
    gf_status st = gf_state_size(c, &bytes);
    if (st) return st;
    std::vector<unsigned char> buf;
    try {
        buf.resize(bytes);
    } catch (const std::bad_alloc&) {
        return set_error(c, GF_ERR_ALLOC, "host allocation of the state image failed");
    }
    if ((st = gf_state_save(c, buf.data(), bytes))) return st;
    const std::string tmp = std::string(path) + ".tmp";
    FILE* f = std::fopen(tmp.c_str(), "wb");
    if (!f) return set_error(c, GF_ERR_ARG, std::string("cannot open ") + tmp + " for writing");
    const bool ok = std::fwrite(buf.data(), 1, bytes, f) == bytes;
    const bool closed = std::fclose(f) == 0;
    if (!ok || !closed || std::rename(tmp.c_str(), path) != 0) {
        std::remove(tmp.c_str());
        return set_error(c, GF_ERR_ARG, std::string("writing ") + path + " failed");
    }
    return GF_OK;
}

gf_status gf_state_read(gf_ctx* c, const char* path) {
    if (!c || !path) return set_error(c, GF_ERR_ARG, "null argument");
    FILE* f = std::fopen(path, "rb");
    if (!f) return set_error(c, GF_ERR_ARG, std::string("cannot open ") + path);
    std::vector<unsigned char> buf;
    bool ok = std::fseek(f, 0, SEEK_END) == 0;
    const long len = ok ? std::ftell(f) : -1;
    ok = ok && len >= 0 && std::fseek(f, 0, SEEK_SET) == 0;
    if (ok) {
        try {
            buf.resize(static_cast<size_t>(len));
        } catch (const std::bad_alloc&) {
            std::fclose(f);
            return set_error(c, GF_ERR_ALLOC, "host allocation of the state image failed");
        }
        ok = std::fread(buf.data(), 1, buf.size(), f) == buf.size();
    }
    std::fclose(f);
    if (!ok) return set_error(c, GF_ERR_ARG, std::string("reading ") + path + " failed");
    return gf_state_load(c, buf.data(), buf.size());
}


// The aliasing check over the stage arguments the builders produce for
// every combination of physics, couplings and optional fields, every stage
// and both stage-1 stores (m, or an uploaded store), on host contexts whose
// arrays are distinct placeholder addresses (nothing touches the device).
gf_status gf_selftest_stage_io(int plant, char* msg, size_t cap) {
    auto put = [&](const std::string& s) {
        if (msg && cap) {
            std::strncpy(msg, s.c_str(), cap - 1);
            msg[cap - 1] = 0;
        }
    };
    if (plant < 0 || plant > 4) { put("plant must be 0..4"); return GF_ERR_ARG; }
    int64_t checked = 0, found = 0;
    std::string first;
    for (uint32_t combo = 0; combo < (1u << 12); combo++) {
        gf_ctx* c = new gf_ctx();
        uintptr_t next = 0x100000;
        auto fake = [&]() { next += 0x1000; return reinterpret_cast<double*>(next); };
        auto fake3 = [&](double** a) { for (int i = 0; i < 3; i++) a[i] = fake(); };
        auto bit = [&](int k) { return ((combo >> k) & 1u) != 0; };
        c->any_fm = bit(0);
        c->any_afm = bit(1);
        c->any_fe = bit(2);
        c->el_on = bit(3);
        c->em_on = bit(4);
        c->flags = (bit(5) ? GF_FLAG_ELASTO_ON_PM : 0u) | (c->el_on ? GF_FLAG_ELASTODYNAMICS : 0u) |
                   (bit(6) ? GF_FLAG_EM_BACKACTION_M | GF_FLAG_EM_BACKACTION_P : 0u) |
                   (bit(7) ? GF_FLAG_MAGNETOSTATICS : 0u) | (bit(8) ? GF_FLAG_ORIENT_ROTATE : 0u);
        c->has_hnu = bit(9);
        c->hstat_written = bit(10);
        c->d_code = reinterpret_cast<uint32_t*>(fake());
        c->d_mats = reinterpret_cast<DevMaterial*>(fake());
        fake3(c->m); fake3(c->st_a); fake3(c->st_b); fake3(c->acc);
        fake3(c->m1); fake3(c->m2); fake3(c->st1_a); fake3(c->st1_b); fake3(c->st2_a); fake3(c->st2_b);
        fake3(c->acc1); fake3(c->acc2);
        fake3(c->p); fake3(c->q); fake3(c->p_a); fake3(c->p_b); fake3(c->q_a); fake3(c->q_b);
        fake3(c->pacc); fake3(c->qacc);
        fake3(c->hstat); fake3(c->hnu); fake3(c->Hcell); fake3(c->Ecell); fake3(c->estat);
        if (bit(11)) {   // optional fields: k outputs (Faraday, spin pumping), shaped drive, flexo force
            fake3(c->em_dm); fake3(c->em_dm1); fake3(c->em_dm2); fake3(c->hdrive); fake3(c->flexo_force);
            c->d_flexo = reinterpret_cast<FlexoMat*>(fake());
        }
        for (int i = 0; i < 6; i++) { c->ec[i] = fake(); c->e0c[i] = fake(); c->e_t0[i] = fake(); }
        FusedZArgs f{};
        f.hspec = reinterpret_cast<double2*>(fake());
        f.mspec = reinterpret_cast<double2*>(fake());
        f.tw = reinterpret_cast<double2*>(fake());
        for (int s0 : {0, 2}) {
            int sel = s0;
            for (int stage = 1; stage <= 4; stage++) {
                const int out = stage == 4 ? 0 : (sel == 1 ? 2 : 1);
                std::string why;
                if (c->any_fm) {
                    StageArgs a = fm_stage_args(c, stage, sel, out);
                    if (plant == 2) a.out[1] = const_cast<double*>(a.store[1]);
                    why = stage_io_conflict(a, bit(7) ? &f : nullptr);
                    checked++;
                    if (!why.empty()) { found++; if (first.empty()) first = "FM stage: " + why; }
                }
                if (c->any_afm) {
                    AfmStageArgs a = afm_stage_args(c, stage, sel, out);
                    if (plant == 3) a.acc1[0] = const_cast<double*>(a.base2[0]);
                    why = stage_io_conflict(a);
                    checked++;
                    if (!why.empty()) { found++; if (first.empty()) first = "AFM stage: " + why; }
                }
                if (c->any_fe) {
                    FeStageArgs a = fe_stage_args(c, stage, sel, out);
                    if (plant == 1) a.pout[0] = const_cast<double*>(a.qstore[1]);
                    if (plant == 4) a.qacc[2] = a.pacc[2];
                    why = stage_io_conflict(a);
                    checked++;
                    if (!why.empty()) { found++; if (first.empty()) first = "FE stage: " + why; }
                }
                sel = out;
            }
        }
        delete c;
    }
    if (plant == 0) {
        if (found) { put(first); return GF_ERR_STATE; }
        put("no aliasing in " + std::to_string(checked) + " stage argument sets");
        return GF_OK;
    }
    // a planted alias must be reported
    put(first.empty() ? std::string("planted alias not reported") : first);
    return found ? GF_OK : GF_ERR_STATE;
}

}  // extern "C"
