// gf_em_dev.h — device code of the electrodynamics stage shared by the EM
// launches and the launch that runs a ferroelectric stage beside them.
// Ampere per edge with the four-cell means of permittivity, conductivity,
// Jf, Jp and dp (get_dE_RK*, EMdynamic_solver.cpp:337-2237) fused with the
// RK update (update_DE_RK*, :3588-3777); Faraday per face with dM of the two
// cells (get_dH_RK*, :2239-3475) fused with update_DH_RK* (:3477-3586).
// RK4 keeps a state accumulator, acc = y + k1/6, acc += k2/3, acc += k3/3,
// y = acc + k4/6, the reference's left-to-right sum.
#pragma once

#include "gf_internal.h"

namespace gf {
namespace em_dev {

__device__ __forceinline__ void rk_state(int stage, int64_t i, double k, double* base, double* acc, double* out) {
    const double y = base[i];
    if (stage == 1) {
        out[i] = y + k * 0.5;
        acc[i] = y + k * kSixth;
    } else if (stage == 2) {
        out[i] = y + k * 0.5;
        acc[i] = acc[i] + k * kThird;
    } else if (stage == 3) {
        out[i] = y + k;
        acc[i] = acc[i] + k * kThird;
    } else {
        base[i] = acc[i] + k * kSixth;
    }
}

// The free current of cell (i, j, k), component comp, at the stage time:
// the value update_Jf_input* writes in its box (EMdynamic_solver.cpp:
// 3885-4315), zero elsewhere.
__device__ __forceinline__ double jf_at(const EmArgs& a, int comp, int64_t i, int64_t j, int64_t k) {
    if (comp != a.jf_comp || i < a.jf_lo[0] || i > a.jf_hi[0] || j < a.jf_lo[1] || j > a.jf_hi[1] || k < a.jf_lo[2] ||
        k > a.jf_hi[2])
        return 0.;
    return a.jf_spatial ? a.jf_value * a.jf_spatial[k - a.jf_lo[2]] : a.jf_value;
}

// Four-cell means around an edge and the Ampere increment
// (EMdynamic_solver.cpp:392-499), with the spin-pumping current on x and y.
// The cells are (ci, cj, ck)[q], index cell[q].
__device__ __forceinline__ double edge_increment(const EmArgs& a, int comp, const int64_t cell[4], const int64_t ci[4],
                                                 const int64_t cj[4], const int64_t ck[4], double curl,
                                                 double e_store) {
    double Jf = 0., Jp = 0., Ji = 0., dP = 0., er = 0., cond = 0., pc = 0.;
    // the codes and the q store of the four cells are read together, before
    // any code is known (every index is in bounds)
    uint32_t tq[4];
    double qv[4] = {0., 0., 0., 0.};
    for (int q = 0; q < 4; q++) {
        tq[q] = a.code[cell[q]] & kMatMask;
        if (a.q[0]) qv[q] = a.q[comp][cell[q]];
    }
    for (int q = 0; q < 4; q++) {
        const int64_t id = cell[q];
        const uint32_t t = tq[q];
        if (!t) {
            er = er + 1.;
        } else {
            const DevMaterial& M = a.mats[t - 1];
            if (a.q[0]) dP = dP + qv[q] * a.dt;
            if (M.is_fe) pc = pc + 1.;
            er = er + M.em_er;
            cond = cond + M.em_cond;
        }
        if (a.jf_comp >= 0) Jf = Jf + jf_at(a, comp, ci[q], cj[q], ck[q]);
        if (a.jps[0]) Jp = Jp + a.jps[comp][id] + a.jps[3 + comp][id] + a.jps[6 + comp][id];
        if (a.jishe[0] && comp < 2) Ji = Ji + a.jishe[comp][id];
    }
    if (pc < 0.5) pc = 1.;
    Jf = Jf / 4.;
    Jp = Jp / 4.;
    Ji = Ji / 4.;
    dP = dP / pc;
    er = er / 4.;
    cond = cond / 4.;
    return a.dt / kEps0 / er * (curl - Jf - Jp - Ji - cond * e_store - dP / a.dt);
}

// Ampere on edge e of the concatenated Ex|Ey|Ez grids, with its RK update.
__device__ __forceinline__ void em_edge(const EmArgs& a, int64_t e) {
    const int64_t nx = a.nx, ny = a.ny, nz = a.nz;
    const int64_t nEx = nx * (ny + 1) * (nz + 1), nEy = (nx + 1) * ny * (nz + 1);
    auto C = [&](int64_t i, int64_t j, int64_t k) { return (i * ny + j) * nz + k; };
    auto HX = [&](int64_t i, int64_t j, int64_t k) { return (i * ny + j) * nz + k; };
    auto HY = [&](int64_t i, int64_t j, int64_t k) { return (i * (ny + 1) + j) * nz + k; };
    auto HZ = [&](int64_t i, int64_t j, int64_t k) { return (i * ny + j) * (nz + 1) + k; };
    int comp;
    double k_ = 0.;
    bool hist = false;   // the edge lies on a plane a Mur wall reads (two next to each non-periodic wall)
    auto near = [](int64_t v, int64_t n) { return v <= 1 || v >= n - 1; };
    if (e < nEx) {
        comp = 0;
        const int64_t i = e / ((ny + 1) * (nz + 1)), j = (e / (nz + 1)) % (ny + 1), k = e % (nz + 1);
        hist = (!a.per[1] && near(j, ny)) || (!a.per[2] && near(k, nz));
        const bool wj = j == 0 || j == ny, wk = k == 0 || k == nz;
        if (!(wj && !a.per[1]) && !(wk && !a.per[2])) {
            const int64_t y1 = wj ? ny - 1 : j - 1, y3 = wj ? 0 : j, z1 = wk ? nz - 1 : k - 1, z2 = wk ? 0 : k;
            const double dHzdy = (a.hs[2][HZ(i, y3, z2)] - a.hs[2][HZ(i, y1, z2)]) / a.dy;
            const double dHydz = (a.hs[1][HY(i, y3, z2)] - a.hs[1][HY(i, y3, z1)]) / a.dz;
            const int64_t cell[4] = {C(i, y1, z1), C(i, y1, z2), C(i, y3, z1), C(i, y3, z2)};
            const int64_t ci[4] = {i, i, i, i}, cj[4] = {y1, y1, y3, y3}, ck[4] = {z1, z2, z1, z2};
            k_ = edge_increment(a, 0, cell, ci, cj, ck, dHzdy - dHydz, a.es[0][e]);
        }
    } else if (e < nEx + nEy) {
        comp = 1;
        e -= nEx;
        const int64_t i = e / (ny * (nz + 1)), j = (e / (nz + 1)) % ny, k = e % (nz + 1);
        hist = (!a.per[0] && near(i, nx)) || (!a.per[2] && near(k, nz));
        const bool wi = i == 0 || i == nx, wk = k == 0 || k == nz;
        if (!(wi && !a.per[0]) && !(wk && !a.per[2])) {
            const int64_t x1 = wi ? nx - 1 : i - 1, x3 = wi ? 0 : i, z1 = wk ? nz - 1 : k - 1, z2 = wk ? 0 : k;
            const double dHxdz = (a.hs[0][HX(x3, j, z2)] - a.hs[0][HX(x3, j, z1)]) / a.dz;
            const double dHzdx = (a.hs[2][HZ(x3, j, z2)] - a.hs[2][HZ(x1, j, z2)]) / a.dx;
            const int64_t cell[4] = {C(x1, j, z1), C(x1, j, z2), C(x3, j, z1), C(x3, j, z2)};
            const int64_t ci[4] = {x1, x1, x3, x3}, cj[4] = {j, j, j, j}, ck[4] = {z1, z2, z1, z2};
            k_ = edge_increment(a, 1, cell, ci, cj, ck, dHxdz - dHzdx, a.es[1][e]);
        }
    } else {
        comp = 2;
        e -= nEx + nEy;
        const int64_t i = e / ((ny + 1) * nz), j = (e / nz) % (ny + 1), k = e % nz;
        hist = (!a.per[0] && near(i, nx)) || (!a.per[1] && near(j, ny));
        const bool wi = i == 0 || i == nx, wj = j == 0 || j == ny;
        if (!(wi && !a.per[0]) && !(wj && !a.per[1])) {
            const int64_t x1 = wi ? nx - 1 : i - 1, x3 = wi ? 0 : i, y1 = wj ? ny - 1 : j - 1, y2 = wj ? 0 : j;
            const double dHydx = (a.hs[1][HY(x3, y2, k)] - a.hs[1][HY(x1, y2, k)]) / a.dx;
            const double dHxdy = (a.hs[0][HX(x3, y2, k)] - a.hs[0][HX(x3, y1, k)]) / a.dy;
            const int64_t cell[4] = {C(x1, y1, k), C(x1, y2, k), C(x3, y1, k), C(x3, y2, k)};
            const int64_t ci[4] = {x1, x1, x3, x3}, cj[4] = {y1, y2, y1, y2}, ck[4] = {k, k, k, k};
            k_ = edge_increment(a, 2, cell, ci, cj, ck, dHydx - dHxdy, a.es[2][e]);
        }
    }
    // stage 4: the base value before the combine overwrites it, for the
    // Mur walls (transfer_pointer, EMdynamic_AbsorbingBoundary.cpp:3-56);
    // each edge's own thread reads and then writes its base
    if (a.t1[0] && hist) a.t1[comp][e] = a.eb[comp][e];
    rk_state(a.stage, e, k_, a.eb[comp], a.eacc[comp], a.eo[comp]);
}

// dM of one cell next to a face (EMdynamic_solver.cpp:2290-2340)
__device__ __forceinline__ double cell_dm(const EmArgs& a, int comp, int64_t id, double& count) {
    const uint32_t t = a.code[id] & kMatMask;
    if (!t) return 0.;
    const DevMaterial& M = a.mats[t - 1];
    double d = 0.;
    if (a.dm[0] || a.dma1[0]) {
        const double f = a.dm[0] ? a.dm[comp][id] : 0.;
        const double f1 = a.dma1[0] ? a.dma1[comp][id] : 0.;
        const double f2 = a.dma2[0] ? a.dma2[comp][id] : 0.;
        d = f * M.Ms + f1 * M.aMs[0] + f2 * M.aMs[1];
    }
    if (M.is_fm || M.is_afm) count = count + 1.;
    return d;
}

// Faraday on face f of the concatenated Hx|Hy|Hz grids, with its RK update.
__device__ __forceinline__ void em_face(const EmArgs& a, int64_t f) {
    const int64_t nx = a.nx, ny = a.ny, nz = a.nz;
    const int64_t nHx = (nx + 1) * ny * nz, nHy = nx * (ny + 1) * nz;
    auto C = [&](int64_t i, int64_t j, int64_t k) { return (i * ny + j) * nz + k; };
    auto EX = [&](int64_t i, int64_t j, int64_t k) { return (i * (ny + 1) + j) * (nz + 1) + k; };
    auto EY = [&](int64_t i, int64_t j, int64_t k) { return (i * ny + j) * (nz + 1) + k; };
    auto EZ = [&](int64_t i, int64_t j, int64_t k) { return (i * (ny + 1) + j) * nz + k; };
    const double pre = -a.dt / kMu0;
    int comp;
    double k_, cnt = 0., d1, d2;
    if (f < nHx) {
        comp = 0;
        const int64_t i = f / (ny * nz), j = (f / nz) % ny, k = f % nz;
        const int64_t x1 = i == 0 ? (a.per[0] ? nx - 1 : 0) : i - 1, x2 = i == nx ? (a.per[0] ? 0 : nx - 1) : i;
        const double dEzdy = (a.es[2][EZ(i, j + 1, k)] - a.es[2][EZ(i, j, k)]) / a.dy;
        const double dEydz = (a.es[1][EY(i, j, k + 1)] - a.es[1][EY(i, j, k)]) / a.dz;
        d1 = cell_dm(a, 0, C(x1, j, k), cnt);
        d2 = cell_dm(a, 0, C(x2, j, k), cnt);
        if (cnt < 0.5) cnt = 1.;
        k_ = pre * (dEzdy - dEydz) - (d1 + d2) / cnt;
    } else if (f < nHx + nHy) {
        comp = 1;
        f -= nHx;
        const int64_t i = f / ((ny + 1) * nz), j = (f / nz) % (ny + 1), k = f % nz;
        const int64_t y1 = j == 0 ? (a.per[1] ? ny - 1 : 0) : j - 1, y2 = j == ny ? (a.per[1] ? 0 : ny - 1) : j;
        const double dExdz = (a.es[0][EX(i, j, k + 1)] - a.es[0][EX(i, j, k)]) / a.dz;
        const double dEzdx = (a.es[2][EZ(i + 1, j, k)] - a.es[2][EZ(i, j, k)]) / a.dx;
        d1 = cell_dm(a, 1, C(i, y1, k), cnt);
        d2 = cell_dm(a, 1, C(i, y2, k), cnt);
        if (cnt < 0.5) cnt = 1.;
        k_ = pre * (dExdz - dEzdx) - (d1 + d2) / cnt;
    } else {
        comp = 2;
        f -= nHx + nHy;
        const int64_t i = f / (ny * (nz + 1)), j = (f / (nz + 1)) % ny, k = f % (nz + 1);
        const int64_t z1 = k == 0 ? (a.per[2] ? nz - 1 : 0) : k - 1, z2 = k == nz ? (a.per[2] ? 0 : nz - 1) : k;
        const double dEydx = (a.es[1][EY(i + 1, j, k)] - a.es[1][EY(i, j, k)]) / a.dx;
        const double dExdy = (a.es[0][EX(i, j + 1, k)] - a.es[0][EX(i, j, k)]) / a.dy;
        d1 = cell_dm(a, 2, C(i, j, z1), cnt);
        d2 = cell_dm(a, 2, C(i, j, z2), cnt);
        if (cnt < 0.5) cnt = 1.;
        k_ = pre * (dEydx - dExdy) - (d1 + d2) / cnt;
    }
    rk_state(a.stage, f, k_, a.hb[comp], a.hacc[comp], a.ho[comp]);
}

}  // namespace em_dev
}  // namespace gf
