// gf_fe.hip — fused ferroelectric RK4 stage for gfx950 (FP64).
//
// One launch per stage does get_dq_RKs, get_dp_RKs, update_q_RKs and
// update_p_RKs (ferroelectric_solver.cpp:534-1832, 1835-1999), alone or, in
// a context with electrodynamics stepped by gf_step, together with the EM
// stage's Ampere and Faraday updates:
//   k_q = -dt/m [ dF_Landau/dp - G11 lap(p) - E_stat - E_ext ] - dt*gamma/m q_s   (FE cells)
//   k_p = q_s dt                                                                 (all cells)
//   stage 1..3: next = y + c_s k, acc = y + k/6 | acc + k/3
//   stage 4:    y = acc + k/6
// which is the reference's left-to-right y + k1/6 + k2/3 + k3/3 + k4/6.
// Powers are products instead of pow(), and the RK weights products with the
// rounded reciprocals instead of quotients; the difference is a few ulp.
#include "gf_em_dev.h"
#include "gf_internal.h"

namespace gf {

namespace {

__device__ __forceinline__ void rot(const double T[3][3], double ix, double iy, double iz, double& ox, double& oy,
                                    double& oz) {
    ox = T[0][0] * ix + T[0][1] * iy + T[0][2] * iz;
    oy = T[1][0] * ix + T[1][1] * iy + T[1][2] * iz;
    oz = T[2][0] * ix + T[2][1] * iy + T[2][2] * iz;
}

// Derivative of the 8th-order Landau energy along component a, with b and c
// the other two (ferroelectric_solver.cpp:740-786, strain terms zero).
__device__ __forceinline__ double landau(const DevMaterial& M, double pa, double pb, double pc) {
    const double a2 = pa * pa, b2 = pb * pb, c2 = pc * pc;
    const double a3 = a2 * pa, a4 = a2 * a2, a5 = a4 * pa, a7 = a5 * a2;
    const double b4 = b2 * b2, c4 = c2 * c2, b6 = b4 * b2, c6 = c4 * c2;
    const double bc2 = b2 + c2;
    return 2. * M.a1 * pa + 4. * M.a11 * a3 + 2. * M.a12 * bc2 * pa + 6. * M.a111 * a5 +
           4. * M.a112 * a3 * bc2 + 2. * M.a112 * (b4 + c4) * pa + 2. * M.a123 * pa * b2 * c2 +
           8. * M.a1111 * a7 + 2. * M.a1112 * pa * (b6 + c6 + 3. * a4 * bc2) + 4. * M.a1122 * a3 * (b4 + c4) +
           2. * M.a1123 * pa * b2 * c2 * (2. * a2 + b2 + c2);
}

__device__ __forceinline__ void rk(int stage, int64_t id, double k, const double* base, double* acc, double* out) {
    const double y = base[id];
    if (stage == 1) {
        out[id] = y + k * 0.5;
        acc[id] = y + k * kSixth;
    } else if (stage == 2) {
        out[id] = y + k * 0.5;
        acc[id] = acc[id] + k * kThird;
    } else if (stage == 3) {
        out[id] = y + k;
        acc[id] = acc[id] + k * kThird;
    } else {
        out[id] = acc[id] + k * kSixth;
    }
}

// The stage of cell id.
__device__ __forceinline__ void fe_cell(const FeStageArgs& a, int64_t id) {
    const uint32_t own = a.code[id];
    double kq[3] = {0., 0., 0.};
    const double q0 = a.qstore[0][id], q1 = a.qstore[1][id], q2 = a.qstore[2][id];
    if (own & kIsFE) {
        const DevMaterial& M = a.mats[(own & kMatMask) - 1];
        const int64_t nyz = a.ny * a.nz;
        const int64_t i = id / nyz, j = (id - i * nyz) / a.nz, k = id - i * nyz - j * a.nz;
        const double c[3] = {a.pstore[0][id], a.pstore[1][id], a.pstore[2][id]};
        // Laplacian, mirror value across a surface face (ferroelectric_solver.cpp:4-131)
        double lap[3] = {0., 0., 0.};
        const int64_t nb[6] = {((i == 0 ? a.nx - 1 : i - 1) * a.ny + j) * a.nz + k,
                               ((i == a.nx - 1 ? 0 : i + 1) * a.ny + j) * a.nz + k,
                               (i * a.ny + (j == 0 ? a.ny - 1 : j - 1)) * a.nz + k,
                               (i * a.ny + (j == a.ny - 1 ? 0 : j + 1)) * a.nz + k,
                               (i * a.ny + j) * a.nz + (k == 0 ? a.nz - 1 : k - 1),
                               (i * a.ny + j) * a.nz + (k == a.nz - 1 ? 0 : k + 1)};
        const double h[3] = {a.dx, a.dy, a.dz};
        for (int ax = 0; ax < 3; ax++) {
            const bool sb = (own >> (kFeSurfShift + 2 * ax)) & 1u;
            const bool sf = (own >> (kFeSurfShift + 2 * ax + 1)) & 1u;
            for (int cc = 0; cc < 3; cc++) {
                const double* s = cc == 0 ? a.pstore[0] : (cc == 1 ? a.pstore[1] : a.pstore[2]);
                const double bw = sb ? c[cc] : s[nb[2 * ax]];
                const double fw = sf ? c[cc] : s[nb[2 * ax + 1]];
                lap[cc] = lap[cc] + (fw - 2. * c[cc] + bw) / h[ax] / h[ax];
            }
        }
        double px = c[0], py = c[1], pz = c[2];
        if (a.rotate) rot(a.Tg2c, c[0], c[1], c[2], px, py, pz);
        double lx = landau(M, px, py, pz);
        double ly = landau(M, py, pz, px);
        double lz = landau(M, pz, px, py);
        if (a.strain[0]) {
            // electrostrictive coupling to strain minus eigenstrain
            // (ferroelectric_solver.cpp:624-653, 776-786), same sum order
            double exx = a.strain[0][id], eyy = a.strain[1][id], ezz = a.strain[2][id];
            double eyz = a.strain[3][id], exz = a.strain[4][id], exy = a.strain[5][id];
            double exx0 = a.eigen[0][id], eyy0 = a.eigen[1][id], ezz0 = a.eigen[2][id];
            double eyz0 = a.eigen[3][id], exz0 = a.eigen[4][id], exy0 = a.eigen[5][id];
            if (a.strain0[0]) {
                // the static strain and eigenstrain added to the dynamic
                // ones (ferroelectric_solver.cpp:636-650)
                exx = exx + a.strain0[0][id]; eyy = eyy + a.strain0[1][id]; ezz = ezz + a.strain0[2][id];
                eyz = eyz + a.strain0[3][id]; exz = exz + a.strain0[4][id]; exy = exy + a.strain0[5][id];
                exx0 = exx0 + a.eigen0[0][id]; eyy0 = eyy0 + a.eigen0[1][id]; ezz0 = ezz0 + a.eigen0[2][id];
                eyz0 = eyz0 + a.eigen0[3][id]; exz0 = exz0 + a.eigen0[4][id]; exy0 = exy0 + a.eigen0[5][id];
            }
            lx = lx + (M.tv1 * (exx0 - exx) + M.tv2 * (eyy0 - eyy + ezz0 - ezz)) * px;
            lx = lx + M.tv3 * ((exy0 - exy) * py + (exz0 - exz) * pz);
            ly = ly + (M.tv1 * (eyy0 - eyy) + M.tv2 * (exx0 - exx + ezz0 - ezz)) * py;
            ly = ly + M.tv3 * ((exy0 - exy) * px + (eyz0 - eyz) * pz);
            lz = lz + (M.tv1 * (ezz0 - ezz) + M.tv2 * (exx0 - exx + eyy0 - eyy)) * pz;
            lz = lz + M.tv3 * ((exz0 - exz) * px + (eyz0 - eyz) * py);
        }
        const double fx = M.fe_mass * lx;
        const double fy = M.fe_mass * ly;
        const double fz = M.fe_mass * lz;
        double gx = fx, gy = fy, gz = fz;
        if (a.rotate) rot(a.Tc2g, fx, fy, fz, gx, gy, gz);
        kq[0] = gx + M.fe_mass * (-1. * M.G11 * lap[0] - a.estat[0][id] - a.eext[0]) + M.fe_damping * q0;
        kq[1] = gy + M.fe_mass * (-1. * M.G11 * lap[1] - a.estat[1][id] - a.eext[1]) + M.fe_damping * q1;
        kq[2] = gz + M.fe_mass * (-1. * M.G11 * lap[2] - a.estat[2][id] - a.eext[2]) + M.fe_damping * q2;
        if (a.ecell[0]) {
            // electrodynamic back-action (ferroelectric_solver.cpp:794-798)
            kq[0] = kq[0] - M.fe_mass * a.ecell[0][id];
            kq[1] = kq[1] - M.fe_mass * a.ecell[1][id];
            kq[2] = kq[2] - M.fe_mass * a.ecell[2][id];
        }
        if (a.flexo[0]) {
            // flexoelectric force, computed by fe_flexo_kernel (:800-832)
            kq[0] = kq[0] + M.fe_mass * a.flexo[0][id];
            kq[1] = kq[1] + M.fe_mass * a.flexo[1][id];
            kq[2] = kq[2] + M.fe_mass * a.flexo[2][id];
        }
    }
    rk(a.stage, id, q0 * a.dt, a.pbase[0], a.pacc[0], a.pout[0]);
    rk(a.stage, id, q1 * a.dt, a.pbase[1], a.pacc[1], a.pout[1]);
    rk(a.stage, id, q2 * a.dt, a.pbase[2], a.pacc[2], a.pout[2]);
    rk(a.stage, id, kq[0], a.qbase[0], a.qacc[0], a.qout[0]);
    rk(a.stage, id, kq[1], a.qbase[1], a.qacc[1], a.qout[1]);
    rk(a.stage, id, kq[2], a.qbase[2], a.qacc[2], a.qout[2]);
}

__global__ void __launch_bounds__(256) fe_stage_kernel(FeStageArgs a) {
    const int64_t id = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (id < a.nx * a.ny * a.nz) fe_cell(a, id);
}

// The ferroelectric stage and the electrodynamics stage's Ampere and
// Faraday updates in one launch (cells, then edges, then faces).  They are
// independent: the EM update reads the q store the FE stage reads (dp =
// q_s dt) and neither reads what the other writes; the FE stage reads the
// cell-centred E of the stage store, which the cells launch after the EM
// update writes.
__global__ void __launch_bounds__(256) fe_em_kernel(FeStageArgs f, EmArgs e) {
    const int64_t n = f.nx * f.ny * f.nz, nx = e.nx, ny = e.ny, nz = e.nz;
    const int64_t nE = nx * (ny + 1) * (nz + 1) + (nx + 1) * ny * (nz + 1) + (nx + 1) * (ny + 1) * nz;
    const int64_t nH = (nx + 1) * ny * nz + nx * (ny + 1) * nz + nx * ny * (nz + 1);
    const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t < n) fe_cell(f, t);
    else if (t < n + nE) em_dev::em_edge(e, t - n);
    else if (t < n + nE + nH) em_dev::em_face(e, t - n - nE);
}

// The bracket of the flexoelectric term of get_dq_RK* at each FE cell
// (ferroelectric_solver.cpp:800-832), the same sum order:
//   + tv4 p_a dp_a/da + tv5 (...) + tv6 (...) - tv7 d2p_a/da2 - tv8 (...)
//   - f11 de_aa/da - f12 (...) - 2 f44 (...)
// with the p gradients of get_p_gradient, the second derivatives of the
// Laplacian (FE surfaces mirror the centre) and the gradients of the
// dynamic strain of get_Dstrain_gradient (elastic_dynamic_solver.cpp:
// 230-399: the normal strain from the two face gradients, the others
// central between faces that are not stress free, one-sided against a free
// face, zero between two free faces).
__global__ void __launch_bounds__(256) fe_flexo_kernel(FeFlexoArgs a) {
    const int64_t nx = a.nx, ny = a.ny, nz = a.nz;
    const int64_t id = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (id >= nx * ny * nz) return;
    const uint32_t own = a.code[id];
    if (!(own & kIsFE)) return;
    const FlexoMat& F = a.fm[(own & kMatMask) - 1];
    const int64_t nyz = ny * nz;
    const int64_t i = id / nyz, j = (id - i * nyz) / nz, k = id - i * nyz - j * nz;
    const double h[3] = {a.dx, a.dy, a.dz};
    double g[3][3];
    fe_p_gradient(own, a.p, id, i, j, k, nx, ny, nz, h, g);
    const int64_t nb[6] = {((i == 0 ? nx - 1 : i - 1) * ny + j) * nz + k, ((i == nx - 1 ? 0 : i + 1) * ny + j) * nz + k,
                           (i * ny + (j == 0 ? ny - 1 : j - 1)) * nz + k, (i * ny + (j == ny - 1 ? 0 : j + 1)) * nz + k,
                           (i * ny + j) * nz + (k == 0 ? nz - 1 : k - 1), (i * ny + j) * nz + (k == nz - 1 ? 0 : k + 1)};
    const double c[3] = {a.p[0][id], a.p[1][id], a.p[2][id]};
    double d2[3][3];
    for (int ax = 0; ax < 3; ax++) {
        const bool sb = (own >> (kFeSurfShift + 2 * ax)) & 1u;
        const bool sf = (own >> (kFeSurfShift + 2 * ax + 1)) & 1u;
        for (int cc = 0; cc < 3; cc++) {
            const double bw = sb ? c[cc] : a.p[cc][nb[2 * ax]];
            const double fw = sf ? c[cc] : a.p[cc][nb[2 * ax + 1]];
            d2[cc][ax] = (fw - 2. * c[cc] + bw) / h[ax] / h[ax];
        }
    }
    double e[6][3] = {{0., 0., 0.}, {0., 0., 0.}, {0., 0., 0.}, {0., 0., 0.}, {0., 0., 0.}, {0., 0., 0.}};
    if (a.eg[0]) {
        // stress-free faces as the strain kernel finds them: a wall that is
        // not periodic, or a vacuum neighbour
        const bool wall[6] = {i == 0 && !a.per[0], i == nx - 1 && !a.per[0], j == 0 && !a.per[1],
                              j == ny - 1 && !a.per[1], k == 0 && !a.per[2], k == nz - 1 && !a.per[2]};
        for (int ax = 0; ax < 3; ax++) {
            const bool sb = wall[2 * ax] || (a.code[nb[2 * ax]] & kMatMask) == 0u;
            const bool sf = wall[2 * ax + 1] || (a.code[nb[2 * ax + 1]] & kMatMask) == 0u;
            int64_t bwd = nb[2 * ax], fwd = nb[2 * ax + 1];
            double scale = 2.;
            if (sb && sf) { bwd = id; fwd = id; scale = 1.; }
            else if (sb) { bwd = id; scale = 1.; }
            else if (sf) { fwd = id; scale = 1.; }
            for (int q = 0; q < 6; q++) e[q][ax] = (a.eg[q][fwd] - a.eg[q][bwd]) / h[ax] / scale;
        }
        e[0][0] = (a.gyz[0][((i + 1) * ny + j) * nz + k] - a.gyz[0][(i * ny + j) * nz + k]) / a.dx;
        e[1][1] = (a.gxz[1][(i * (ny + 1) + j + 1) * nz + k] - a.gxz[1][(i * (ny + 1) + j) * nz + k]) / a.dy;
        e[2][2] = (a.gxy[2][(i * ny + j) * (nz + 1) + k + 1] - a.gxy[2][(i * ny + j) * (nz + 1) + k]) / a.dz;
    }
    const double px = c[0], py = c[1], pz = c[2];
    a.out[0][id] = +F.tv4 * px * g[0][0] + F.tv5 * (py * g[1][0] + pz * g[2][0]) +
                   F.tv6 * (px * g[1][1] + px * g[2][2] + py * g[0][1] + pz * g[0][2]) - F.tv7 * d2[0][0] -
                   F.tv8 * (d2[0][2] + d2[0][1]) - F.f11 * e[0][0] - F.f12 * (e[1][0] + e[2][0]) -
                   2. * F.f44 * (e[5][1] + e[4][2]);
    a.out[1][id] = +F.tv4 * py * g[1][1] + F.tv5 * (px * g[0][1] + pz * g[2][1]) +
                   F.tv6 * (py * g[0][0] + py * g[2][2] + px * g[1][0] + pz * g[1][2]) - F.tv7 * d2[1][1] -
                   F.tv8 * (d2[1][0] + d2[1][2]) - F.f11 * e[1][1] - F.f12 * (e[0][1] + e[2][1]) -
                   2. * F.f44 * (e[5][0] + e[3][2]);
    a.out[2][id] = +F.tv4 * pz * g[2][2] + F.tv5 * (px * g[0][2] + py * g[1][2]) +
                   F.tv6 * (pz * g[0][0] + pz * g[1][1] + px * g[2][0] + py * g[2][1]) - F.tv7 * d2[2][2] -
                   F.tv8 * (d2[2][0] + d2[2][1]) - F.f11 * e[2][2] - F.f12 * (e[0][2] + e[1][2]) -
                   2. * F.f44 * (e[3][1] + e[4][0]);
}

}  // namespace

hipError_t launch_fe_flexo(const FeFlexoArgs& a, hipStream_t s) {
    const int64_t blocks = (a.nx * a.ny * a.nz + 255) / 256;
    hipLaunchKernelGGL(fe_flexo_kernel, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_fe_em(const FeStageArgs& f, const EmArgs& e, hipStream_t s) {
    const int64_t n = f.nx * f.ny * f.nz, nx = e.nx, ny = e.ny, nz = e.nz;
    const int64_t nE = nx * (ny + 1) * (nz + 1) + (nx + 1) * ny * (nz + 1) + (nx + 1) * (ny + 1) * nz;
    const int64_t nH = (nx + 1) * ny * nz + nx * (ny + 1) * nz + nx * ny * (nz + 1);
    const int64_t blocks = (n + nE + nH + 255) / 256;
    hipLaunchKernelGGL(fe_em_kernel, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, s, f, e);
    return hipGetLastError();
}

hipError_t launch_fe_stage(const FeStageArgs& a, int64_t n, hipStream_t s) {
    const int64_t blocks = (n + 255) / 256;
    hipLaunchKernelGGL(fe_stage_kernel, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, s, a);
    return hipGetLastError();
}

}  // namespace gf
