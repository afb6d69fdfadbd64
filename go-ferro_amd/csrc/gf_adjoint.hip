// gf_adjoint.hip — discrete adjoint of the FM RK4 trajectory on gfx950.
//
// The forward map of one stage is k = f(y) = P (y x H) + Q (y x (y x H))
// with H = Hd(y) + Ha(y) + Hme(y) + exch S(y) + Hdmi(y) + Hext + Hnu + Hdrive s(t)
// (magnetic_solver.cpp:3-229): Hd the FFT convolution with the Newell
// tensor, Ha the anisotropy, Hme the magnetoelastic field of a static
// strain, S the exchange stencil and P, Q = -gyro dt / (1 + alpha^2),
// alpha P.  Its vector-Jacobian product for a cotangent kbar is
//   qbar = Q kbar,  pbar = P kbar + qbar x y,
//   ybar = p x qbar + H x pbar + (dHa/dy)^T Hbar + (dHme/dy)^T Hbar
//          + S(exch Hbar) + Ms N^T (*) Hbar,       Hbar = pbar x y,
// (+ the transposed DMI stencil, which is not symmetric), where S is
// symmetric (each face carries one harmonic-mean coefficient),
// exch = 2/(mu0 Ms) of each cell's material, and N^T (*) is the convolution
// with the transposed, conjugated tensor spectra.  The parameter gradients
// are those of one material: alpha and B1 at its cells, Aex through the
// coefficient of each face it takes part in (its own faces and, through the
// harmonic mean, its interfaces with other ferromagnets).  Two kernels per stage: a pointwise one (cross products, local
// Jacobians, parameter gradients) writing ybar_local and Hbar, and after the
// adjoint convolution of Hbar a stencil one that adds exch S(Hbar) and
// Ms N^T (*) Hbar and does the RK4 bookkeeping
//   y1 = y0 + k1/6 + k2/3 + k3/3 + k4/6,  ya = y0 + k1/2, yb = y0 + k2/2, yc = y0 + k3
// in reverse: kbar4 = lam/6; acc = lam + ybar_c, kbar3 = lam/3 + ybar_c;
// acc += ybar_b, kbar2 = lam/3 + ybar_b/2; acc += ybar_a, kbar1 = lam/6 +
// ybar_a/2; lam = acc + ybar_0 (+ the loss gradient of y0).
//
// Checkpointing: the forward pass keeps m every K steps (K = ceil(sqrt N)
// by default); the reverse pass recomputes one segment with the same
// kernels (bit-identical trajectory), keeping its stage stores and fields,
// then walks it backwards.  Given a memory budget, the forward pass keeps
// the stage stores and fields of the last M steps that fit it (all N when
// they do), the reverse sweep walks those without recomputing them, and
// only the first N - M steps are checkpointed.  The trajectory and the
// gradients are the same bit for bit.  Per-cell gradient accumulators are summed once
// at the end in a fixed order, so the gradients are deterministic.
#include "gf_llg_dev.h"

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <string>
#include <vector>

namespace gf {

struct AdjState {
    gf_adjoint_desc d{};
    bool ready = false;            // a forward pass whose reverse sweep can run
    int64_t N = 0, K = 0, nseg = 0;
    int64_t M = 0;                 // the last M steps are resident; the first N - M are checkpointed
    double* pool = nullptr;        // (nseg + 1) checkpoints of 3n: segment starts, then the final state
    double* RY = nullptr;          // resident steps: 4M stage stores of 3n (store s of step j at 4j + s - 1)
    double* RH = nullptr;          // and their 4M magnetostatic fields, with magnetostatics
    std::vector<double> ck_time;   // time at each segment start
    double t_final = 0.;
    std::vector<double> tstep;     // time after each step
    std::vector<double> g;         // dL/dX_n
    double* X = nullptr;           // probe means, device
    double loss = 0.;
    // reverse-sweep buffers
    int64_t kbuf = 0;              // segment length the buffers hold
    double* Y = nullptr;           // (4K + 1) stage stores of 3n
    double* HD = nullptr;          // 4K stage fields of 3n
    double* work = nullptr;        // lam, acc, kbar, mbar, hbar, dadj, ghdrive: 7 x 3n
    double* gpar = nullptr;        // galpha, gaex, gb1: 3 x n
};

namespace {

struct AdjPointArgs {
    const uint32_t* code;
    const DevMaterial* mats;
    const double* y[3];            // stage store
    const double* hd[3];           // its magnetostatic field, null when off
    const double* hnu[3];          // may be null
    const double* hdrive[3];       // may be null
    const double* strain[6];       // may be null
    const double* kbar[3];
    double* mbar[3];               // out: local part of the store's cotangent
    double* hbar[3];               // out: cotangent of H
    double* galpha;                // per-cell gradient accumulators
    double* gaex;
    double* gb1;
    double* ghdrive[3];            // may be null
    int64_t nx, ny, nz;
    double dx, dy, dz;
    double hext[3];
    double drive_s;
    double Tg2c[3][3];
    double Tc2g[3][3];
    int32_t rotate;
    int32_t mat;                   // the material whose alpha, Aex, B1 are differentiated (1-based)
    double alpha;                  // its damping
};

// d(face coefficient)/dA_mat of the face between a cell of material mi and
// a neighbour of material mj: the coefficient is A_self (= A) within a
// material and the harmonic mean 2 Ai Aj / (Ai + Aj) across an interface
// (magnetic_initialize.cpp:381-524).
__device__ __forceinline__ double dcoef_dA(const DevMaterial* mats, uint32_t mi, uint32_t mj, uint32_t mat) {
    if (mi == mj) return mi == mat ? 1. : 0.;
    if (mi != mat && mj != mat) return 0.;
    const double Ai = mats[mi - 1].Aex, Aj = mats[mj - 1].Aex;
    const double other = mi == mat ? Aj : Ai;
    return 2. * other * other / ((Ai + Aj) * (Ai + Aj));
}

__device__ __forceinline__ void cross(double ax, double ay, double az, double bx, double by, double bz, double& cx,
                                      double& cy, double& cz) {
    cx = ay * bz - az * by;
    cy = az * bx - ax * bz;
    cz = ax * by - ay * bx;
}

// v -> T^T v
__device__ __forceinline__ void rot_t(const double T[3][3], double ix, double iy, double iz, double& ox, double& oy,
                                      double& oz) {
    ox = T[0][0] * ix + T[1][0] * iy + T[2][0] * iz;
    oy = T[0][1] * ix + T[1][1] * iy + T[2][1] * iz;
    oz = T[0][2] * ix + T[1][2] * iy + T[2][2] * iz;
}

template <class I>
__global__ void __launch_bounds__(256) adj_point_kernel(AdjPointArgs a) {
    const I n = static_cast<I>(a.nx * a.ny * a.nz);
    const I id = static_cast<I>(blockIdx.x) * static_cast<I>(blockDim.x) + static_cast<I>(threadIdx.x);
    if (id >= n) return;
    const uint32_t own = ldu(a.code, id);
    if (!(own & kIsFM)) {
        for (int c = 0; c < 3; c++) { std_(a.mbar[c], id, 0.); std_(a.hbar[c], id, 0.); }
        return;
    }
    GMat& M = mat_of(a.mats, own);
    const I ny = static_cast<I>(a.ny), nz = static_cast<I>(a.nz);
    const I col = id / nz, z = id - col * nz;
    const I x = col / ny, y = col - x * ny;

    const double mgx = ldd(a.y[0], id), mgy = ldd(a.y[1], id), mgz = ldd(a.y[2], id);
    double mx = mgx, my = mgy, mz = mgz;
    if (a.rotate) rot(a.Tg2c, mgx, mgy, mgz, mx, my, mz);
    const P3 sp = {{a.y[0], a.y[1], a.y[2]}};
    double nv[6][3];
    uint32_t ncd[6];
    gather6<I>(sp, a.code, x, y, z, static_cast<I>(a.nx), ny, nz, nv, ncd);
    const Stencil st = stencil_core<true>(mgx, mgy, mgz, nv, ncd, a.mats, own, kFmSurfShift, -1, a.dx, a.dy, a.dz,
                                          M.Aex, M.Aex_self, M.dmi_slope);
    // exchange and interfacial DMI (zero without it) of the forward stage
    const double hex0 = M.exch * st.ex[0] + M.dmi_pos * st.d[2], hex1 = M.exch * st.ex[1] + M.dmi_pos * st.d[3];
    const double hex2 = M.exch * st.ex[2] + M.dmi_neg * (st.d[0] + st.d[1]);
    // the exchange field's derivative in Aex of material a.mat: each
    // non-surface face's difference weighted by d(coefficient)/dAex
    double exd0 = 0., exd1 = 0., exd2 = 0.;
    {
        const double ih2[3] = {1. / (a.dx * a.dx), 1. / (a.dy * a.dy), 1. / (a.dz * a.dz)};
        const uint32_t mi = own & kMatMask;
        for (int f = 0; f < 6; f++) {
            if ((own >> (kFmSurfShift + f)) & 1u) continue;
            const double w = dcoef_dA(a.mats, mi, ncd[f] & kMatMask, static_cast<uint32_t>(a.mat)) * ih2[f >> 1];
            exd0 += w * (nv[f][0] - mgx);
            exd1 += w * (nv[f][1] - mgy);
            exd2 += w * (nv[f][2] - mgz);
        }
    }

    // anisotropy field and Jacobian in the crystal frame (symmetric)
    double hac0 = 0., hac1 = 0., hac2 = 0.;
    double J00 = 0., J11 = 0., J22 = 0., J01 = 0., J02 = 0., J12 = 0.;
    const double x2 = mx * mx, y2 = my * my, z2 = mz * mz;
    if (M.anis_type == 1) {
        hac0 = -2. * (M.K1n * (y2 + z2) + M.K2n * y2 * z2) * mx;
        hac1 = -2. * (M.K1n * (x2 + z2) + M.K2n * x2 * z2) * my;
        hac2 = -2. * (M.K1n * (x2 + y2) + M.K2n * x2 * y2) * mz;
        J00 = -2. * (M.K1n * (y2 + z2) + M.K2n * y2 * z2);
        J11 = -2. * (M.K1n * (x2 + z2) + M.K2n * x2 * z2);
        J22 = -2. * (M.K1n * (x2 + y2) + M.K2n * x2 * y2);
        J01 = -4. * mx * my * (M.K1n + M.K2n * z2);
        J02 = -4. * mx * mz * (M.K1n + M.K2n * y2);
        J12 = -4. * my * mz * (M.K1n + M.K2n * x2);
    } else if (M.anis_type == 2) {
        hac2 = 2. * M.K1n * mz;
        J22 = 2. * M.K1n;
    }
    double ha0 = hac0, ha1 = hac1, ha2 = hac2;
    if (a.rotate) rot(a.Tc2g, hac0, hac1, hac2, ha0, ha1, ha2);
    double H0 = ha0 + hex0 + a.hext[0], H1 = ha1 + hex1 + a.hext[1], H2 = ha2 + hex2 + a.hext[2];
    if (a.hd[0]) { H0 += ldd(a.hd[0], id); H1 += ldd(a.hd[1], id); H2 += ldd(a.hd[2], id); }
    double exx = 0., eyy = 0., ezz = 0., eyz = 0., exz = 0., exy = 0.;
    if (a.strain[0]) {
        exx = ldd(a.strain[0], id); eyy = ldd(a.strain[1], id); ezz = ldd(a.strain[2], id);
        eyz = ldd(a.strain[3], id); exz = ldd(a.strain[4], id); exy = ldd(a.strain[5], id);
        double he0, he1, he2;
        const double* none[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
        magnetoelastic(a.strain, none, id, -2., M.B1n, M.B2n, mx, my, mz, a.rotate, a.Tc2g, he0, he1, he2);
        H0 += he0; H1 += he1; H2 += he2;
    }
    if (a.hnu[0]) { H0 += ldd(a.hnu[0], id); H1 += ldd(a.hnu[1], id); H2 += ldd(a.hnu[2], id); }
    if (a.hdrive[0]) {
        H0 += ldd(a.hdrive[0], id) * a.drive_s;
        H1 += ldd(a.hdrive[1], id) * a.drive_s;
        H2 += ldd(a.hdrive[2], id) * a.drive_s;
    }

    double px, py, pz, qx, qy, qz;
    cross(mgx, mgy, mgz, H0, H1, H2, px, py, pz);
    cross(mgx, mgy, mgz, px, py, pz, qx, qy, qz);
    const double kb0 = ldd(a.kbar[0], id), kb1 = ldd(a.kbar[1], id), kb2 = ldd(a.kbar[2], id);
    const double P = M.precess, Q = M.alpha_precess;
    const double qb0 = Q * kb0, qb1 = Q * kb1, qb2 = Q * kb2;
    double t0, t1, t2;
    cross(qb0, qb1, qb2, mgx, mgy, mgz, t0, t1, t2);
    const double pb0 = P * kb0 + t0, pb1 = P * kb1 + t1, pb2 = P * kb2 + t2;
    double mb0, mb1, mb2, u0, u1, u2;
    cross(px, py, pz, qb0, qb1, qb2, mb0, mb1, mb2);
    cross(H0, H1, H2, pb0, pb1, pb2, u0, u1, u2);
    mb0 += u0; mb1 += u1; mb2 += u2;
    double hb0, hb1, hb2;
    cross(pb0, pb1, pb2, mgx, mgy, mgz, hb0, hb1, hb2);

    // local Jacobians of the crystal-frame fields
    double hc0 = hb0, hc1 = hb1, hc2 = hb2;
    if (a.rotate) rot_t(a.Tc2g, hb0, hb1, hb2, hc0, hc1, hc2);
    double mc0 = J00 * hc0 + J01 * hc1 + J02 * hc2;
    double mc1 = J01 * hc0 + J11 * hc1 + J12 * hc2;
    double mc2 = J02 * hc0 + J12 * hc1 + J22 * hc2;
    if (a.strain[0]) {
        mc0 += -2. * (M.B1n * exx * hc0 + M.B2n * (exy * hc1 + exz * hc2));
        mc1 += -2. * (M.B1n * eyy * hc1 + M.B2n * (exy * hc0 + eyz * hc2));
        mc2 += -2. * (M.B1n * ezz * hc2 + M.B2n * (exz * hc0 + eyz * hc1));
    }
    double mg0 = mc0, mg1 = mc1, mg2 = mc2;
    if (a.rotate) rot_t(a.Tg2c, mc0, mc1, mc2, mg0, mg1, mg2);
    mb0 += mg0; mb1 += mg1; mb2 += mg2;

    // the transposed DMI terms of the cell's own field: the DMI field is
    // (P d_zx, P d_zy, -P (d_xx + d_yy)) with each d the mean of a backward
    // and a forward slope of one x or y face: the difference to the
    // neighbour, or at a surface the slope s = D/(2A) times a component of
    // the cell itself (magnetic_solver.cpp:1639-1831); the exchange carries
    // A s times that component at a surface.  The neighbours' parts of the
    // differences are gathered by the combine kernel.
    int nxs = 0, nys = 0;
    {
        const double s = M.dmi_slope;
        const double ga = M.dmi_neg * hb2 * 0.5, al = M.dmi_pos * hb0 * 0.5, be = M.dmi_pos * hb1 * 0.5;
        const double ix = 1. / a.dx, iy = 1. / a.dy;
        const double kx = M.exch * M.Aex * s * ix, ky = M.exch * M.Aex * s * iy;
        const bool sf[4] = {((own >> (kFmSurfShift + kXB)) & 1u) != 0, ((own >> (kFmSurfShift + kXF)) & 1u) != 0,
                            ((own >> (kFmSurfShift + kYB)) & 1u) != 0, ((own >> (kFmSurfShift + kYF)) & 1u) != 0};
        // x faces: d_xx from the x component, d_zx from the z component
        for (int side = 0; side < 2; side++) {
            const double sg = side ? -1. : 1.;   // backward slope c - L, forward R - c
            if (sf[side]) {
                mb2 += -ga * s; mb0 += al * s;
                mb2 += sg * kx * hb0; mb0 += -sg * kx * hb2;
                nxs++;
            } else {
                mb0 += sg * ga * ix; mb2 += sg * al * ix;
            }
        }
        // y faces: d_yy from the y component, d_zy from the z component
        for (int side = 0; side < 2; side++) {
            const double sg = side ? -1. : 1.;
            if (sf[2 + side]) {
                mb2 += -ga * s; mb1 += be * s;
                mb2 += sg * ky * hb1; mb1 += -sg * ky * hb2;
                nys++;
            } else {
                mb1 += sg * ga * iy; mb2 += sg * be * iy;
            }
        }
    }

    // parameter gradients of material a.mat: P, Q in alpha and Hme linear in
    // B1 at its own cells; the exchange field through the face coefficients
    // (its own faces and its interfaces)
    const bool mine = (own & kMatMask) == static_cast<uint32_t>(a.mat);
    if (mine) {
        const double dP = -2. * a.alpha * P / (1. + a.alpha * a.alpha);
        const double dQ = P + a.alpha * dP;
        std_(a.galpha, id,
             ldd(a.galpha, id) + dP * (px * kb0 + py * kb1 + pz * kb2) + dQ * (qx * kb0 + qy * kb1 + qz * kb2));
    }
    double gx = M.exch * (hb0 * exd0 + hb1 * exd1 + hb2 * exd2);
    if (mine && nxs + nys > 0) {
        // the DMI field's surface slopes, s = D/(2A) with ds/dA = -s/A (the
        // exchange's A s at a surface does not depend on A)
        const double r = -M.dmi_slope / M.Aex * 0.5;
        gx += r * (M.dmi_pos * (hb0 * nxs * mgx + hb1 * nys * mgy) - M.dmi_neg * hb2 * (nxs + nys) * mgz);
    }
    std_(a.gaex, id, ldd(a.gaex, id) + gx);
    if (mine && a.strain[0])
        std_(a.gb1, id, ldd(a.gb1, id) + -2. / kMu0 / M.Ms * (hc0 * exx * mx + hc1 * eyy * my + hc2 * ezz * mz));
    if (a.ghdrive[0]) {
        std_(a.ghdrive[0], id, ldd(a.ghdrive[0], id) + a.drive_s * hb0);
        std_(a.ghdrive[1], id, ldd(a.ghdrive[1], id) + a.drive_s * hb1);
        std_(a.ghdrive[2], id, ldd(a.ghdrive[2], id) + a.drive_s * hb2);
    }
    std_(a.mbar[0], id, mb0); std_(a.mbar[1], id, mb1); std_(a.mbar[2], id, mb2);
    std_(a.hbar[0], id, hb0); std_(a.hbar[1], id, hb1); std_(a.hbar[2], id, hb2);
}

struct AdjCombineArgs {
    const uint32_t* code;
    const DevMaterial* mats;
    const double* mbar[3];
    const double* hbar[3];
    const double* dadj[3];         // N^T (*) Hbar, null without magnetostatics
    double* lam[3];
    double* acc[3];
    double* kbar[3];
    double* ybar[3];               // mode 0 output
    int64_t nx, ny, nz;
    double dx, dy, dz;
    int32_t mode;                  // 0: ybar only; 1..4: RK4 bookkeeping after stage `mode`
    int32_t lo[3], hi[3];          // probe box
    double g;                      // loss gradient per probe cell (mode 1)
};

template <class I>
__global__ void __launch_bounds__(256) adj_combine_kernel(AdjCombineArgs a) {
    const I n = static_cast<I>(a.nx * a.ny * a.nz);
    const I id = static_cast<I>(blockIdx.x) * static_cast<I>(blockDim.x) + static_cast<I>(threadIdx.x);
    if (id >= n) return;
    const uint32_t own = ldu(a.code, id);
    if (!(own & kIsFM)) {
        if (a.mode == 0) for (int c = 0; c < 3; c++) std_(a.ybar[c], id, 0.);
        return;
    }
    GMat& M = mat_of(a.mats, own);
    const I ny = static_cast<I>(a.ny), nz = static_cast<I>(a.nz);
    const I col = id / nz, z = id - col * nz;
    const I x = col / ny, y = col - x * ny;
    // the exchange part of the transposed Jacobian: the field at cell i is
    // exch_i sum_j A_ij (y_j - y_i)/h^2, so cell j receives
    // sum_i A_ij (exch_i Hbar_i - exch_j Hbar_j)/h^2, the stencil of
    // exch Hbar (A_ij symmetric; a surface face carries nothing)
    const P3 sp = {{a.hbar[0], a.hbar[1], a.hbar[2]}};
    double nv[6][3];
    uint32_t ncd[6];
    gather6<I>(sp, a.code, x, y, z, static_cast<I>(a.nx), ny, nz, nv, ncd);
    // the DMI differences of the x and y neighbours that read this cell (a
    // face that is not a surface): a neighbour below reads it in its forward
    // slope (+), one above in its backward slope (-)
    double yd[3] = {0., 0., 0.};
    {
        const double ih[2] = {1. / a.dx, 1. / a.dy};
        for (int f = 0; f < 4; f++) {
            if (((own >> (kFmSurfShift + f)) & 1u) || !(ncd[f] & kIsFM)) continue;
            const DevMaterial& Mn = a.mats[(ncd[f] & kMatMask) - 1];
            const double w = ((f & 1) ? -1. : 1.) * ih[f >> 1];
            const double ga = Mn.dmi_neg * nv[f][2] * 0.5, td = Mn.dmi_pos * nv[f][f >> 1] * 0.5;
            yd[f >> 1] += w * ga;   // d_xx (x faces) or d_yy (y faces)
            yd[2] += w * td;        // d_zx or d_zy
        }
    }
    for (int f = 0; f < 6; f++) {
        const double e = (ncd[f] & kIsFM) ? a.mats[(ncd[f] & kMatMask) - 1].exch : 0.;
        for (int c = 0; c < 3; c++) nv[f][c] = e * nv[f][c];
    }
    const Stencil st = stencil_core<false>(M.exch * ldd(a.hbar[0], id), M.exch * ldd(a.hbar[1], id),
                                           M.exch * ldd(a.hbar[2], id), nv, ncd, a.mats, own, kFmSurfShift, -1, a.dx,
                                           a.dy, a.dz, M.Aex, M.Aex_self, 0.);
    const bool probe = x >= static_cast<I>(a.lo[0]) && x <= static_cast<I>(a.hi[0]) && y >= static_cast<I>(a.lo[1]) &&
                       y <= static_cast<I>(a.hi[1]) && z >= static_cast<I>(a.lo[2]) && z <= static_cast<I>(a.hi[2]);
    for (int c = 0; c < 3; c++) {
        double yb = ldd(a.mbar[c], id) + st.ex[c] + yd[c];
        if (a.dadj[0]) yb += M.Ms * ldd(a.dadj[c], id);
        if (a.mode == 0) {
            std_(a.ybar[c], id, yb);
            continue;
        }
        const double lam = ldd(a.lam[c], id);
        if (a.mode == 4) {
            std_(a.acc[c], id, lam + yb);
            std_(a.kbar[c], id, lam / 3. + yb);
        } else if (a.mode == 3) {
            std_(a.acc[c], id, ldd(a.acc[c], id) + yb);
            std_(a.kbar[c], id, lam / 3. + 0.5 * yb);
        } else if (a.mode == 2) {
            std_(a.acc[c], id, ldd(a.acc[c], id) + yb);
            std_(a.kbar[c], id, lam / 6. + 0.5 * yb);
        } else {
            double l = ldd(a.acc[c], id) + yb;
            if (c == 0 && probe) l += a.g;
            std_(a.lam[c], id, l);
            std_(a.kbar[c], id, l / 6.);
        }
    }
}

// Start of the reverse sweep: lam = dL/dy_N (x component on the probe),
// kbar = lam / 6; the rest of lam, acc and kbar were cleared.
__global__ void __launch_bounds__(256) adj_init_kernel(const uint32_t* code, double* lamx, double* kbarx, int64_t nx,
                                                        int64_t ny, int64_t nz, int lo0, int lo1, int lo2, int hi0,
                                                        int hi1, int hi2, double g) {
    const int64_t id = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (id >= nx * ny * nz) return;
    if (!(code[id] & kIsFM)) return;
    const int64_t z = id % nz, y = (id / nz) % ny, x = id / (ny * nz);
    if (x < lo0 || x > hi0 || y < lo1 || y > hi1 || z < lo2 || z > hi2) return;
    lamx[id] = g;
    kbarx[id] = g / 6.;
}

// Mean of m_x over the probe box, one workgroup, fixed summation order.
__global__ void __launch_bounds__(256) probe_mean_kernel(const double* mx, int64_t ny, int64_t nz, int lo0, int lo1,
                                                          int lo2, int hi0, int hi1, int hi2, double* out) {
    __shared__ double red[256];
    const int64_t bx = hi0 - lo0 + 1, by = hi1 - lo1 + 1, bz = hi2 - lo2 + 1;
    const int64_t cells = bx * by * bz;
    double s = 0.;
    for (int64_t t = threadIdx.x; t < cells; t += blockDim.x) {
        const int64_t k = t % bz, j = (t / bz) % by, i = t / (by * bz);
        s += mx[((lo0 + i) * ny + (lo1 + j)) * nz + (lo2 + k)];
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (static_cast<int>(threadIdx.x) < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = red[0] / static_cast<double>(cells);
}

unsigned blocks256(int64_t n) { return static_cast<unsigned>((n + 255) / 256); }

// --- host helpers ---------------------------------------------------------

gf_status check_context(gf_ctx* c, int material) {
    if (!c->any_fm || c->any_afm || c->any_fe || c->el_on || c->em_on)
        return set_error(c, GF_ERR_UNSUPPORTED, "the adjoint sweep needs a ferromagnet-only context");
    const uint32_t bad = GF_FLAG_MAG_1D | GF_FLAG_DEMAG_FACTOR | GF_FLAG_PRESCRIBE_HEXT | GF_FLAG_EMDYNAMIC |
                         GF_FLAG_PRESCRIBE_EEXT | GF_FLAG_ELASTODYNAMICS | GF_FLAG_PRESCRIBED_M;
    if (c->flags & bad)
        return set_error(c, GF_ERR_UNSUPPORTED,
                         "the adjoint sweep supports FFT or no magnetostatics and sinusoidal drives only "
                         "(no prescribed layer)");
    // gf_step recomputes the static strain from m after every step; the
    // sweep's forward would keep it fixed and its gradient would miss
    // d(strain)/dm, so a context that solves the elastostatics is refused
    if (c->flags & GF_FLAG_ELASTOSTATIC)
        return set_error(c, GF_ERR_UNSUPPORTED,
                         "the adjoint sweep does not support the elastostatic solver (static strain from m)");
    if (material < 1 || material > static_cast<int>(c->host_mats.size()) || !c->host_mats[material - 1].if_FM)
        return set_error(c, GF_ERR_ARG, "material must name an FM material");
    const gf_material& m = c->host_mats[material - 1];
    if (m.Aex == 0.) return set_error(c, GF_ERR_UNSUPPORTED, "the adjoint sweep needs Aex != 0");
    if (c->mag_next_stage != 1) return set_error(c, GF_ERR_STATE, "the context is in the middle of a step");
    return GF_OK;
}

bool demag_on(const gf_ctx* c) { return (c->flags & GF_FLAG_MAGNETOSTATICS) && c->demag.ready; }

gf_status dalloc(gf_ctx* c, double** p, int64_t count, const char* what) {
    if (*p) { hipFree(*p); *p = nullptr; }
    return hip_check(c, hipMalloc(p, sizeof(double) * static_cast<size_t>(count)), what);
}

void slot3(double* base, int64_t slot, int64_t n, double* out[3]) {
    for (int a = 0; a < 3; a++) out[a] = base + (slot * 3 + a) * n;
}

gf_status copy3(gf_ctx* c, double* const* dst, const double* const* src) {
    for (int a = 0; a < 3; a++) {
        gf_status st = hip_check(c, hipMemcpyAsync(dst[a], src[a], sizeof(double) * static_cast<size_t>(c->n),
                                                   hipMemcpyDeviceToDevice, c->stream), "adjoint copy");
        if (st) return st;
    }
    return GF_OK;
}

// One FM stage with explicit stores: reads `in`, base `base`, writes `out`,
// field `h` (the magnetostatic field of `in`; the context's Hstat when
// magnetostatics is off, so that an uploaded static field still acts).
gf_status stage_to(gf_ctx* c, int stage, double* const* base, double* const* in, double* const* out,
                   double* const* h) {
    StageArgs a = fm_stage_args(c, stage, 0, 1);
    for (int i = 0; i < 3; i++) {
        a.store[i] = in[i]; a.base[i] = base[i]; a.out[i] = out[i]; a.hstat[i] = h[i];
    }
    fm_stage_drives(c, stage, a);
    return hip_check(c, launch_mag_stage(a, c->n, c->stream), "adjoint forward stage");
}

// One RK4 step from y0 (statics between stages as gf_step does), stage
// stores to st[0..2], the new state to y1, fields to h[0..3] (h[s] of the
// store stage s+1 reads; h[4] = field of y1, may equal h[0] storage of the
// next step).  Without magnetostatics every h[s] is the context's Hstat.
gf_status step_to(gf_ctx* c, double* const* y0, double* const* const st[3], double* const* y1,
                  double* const* const h[5]) {
    gf_status e = gf_time_advance(c);
    if (e) return e;
    const bool ms = demag_on(c);
    double* const* ins[4] = {y0, st[0], st[1], st[2]};
    double* const* outs[4] = {st[0], st[1], st[2], y1};
    for (int s = 1; s <= 4; s++) {
        if ((e = stage_to(c, s, y0, ins[s - 1], outs[s - 1], h[s - 1]))) return e;
        if (ms && h[s] && (e = demag_run_to(c, outs[s - 1], h[s]))) return e;
    }
    return GF_OK;
}

void free_state(AdjState* s) {
    if (!s) return;
    hipFree(s->pool); hipFree(s->X); hipFree(s->Y); hipFree(s->HD); hipFree(s->work); hipFree(s->gpar);
    hipFree(s->RY); hipFree(s->RH);
    delete s;
}

void stage_point_args(gf_ctx* c, int material, AdjPointArgs& p) {
    p = AdjPointArgs{};
    p.code = c->d_code; p.mats = c->d_mats;
    const double* strain[6];
    StageArgs sa = fm_stage_args(c, 1, 0, 1);
    for (int i = 0; i < 6; i++) strain[i] = sa.strain[i];
    for (int i = 0; i < 6; i++) p.strain[i] = strain[i];
    for (int i = 0; i < 3; i++) { p.hnu[i] = sa.hnu[i]; p.hdrive[i] = sa.hdrive[i]; }
    p.nx = c->nx; p.ny = c->ny; p.nz = c->nz; p.dx = c->dx; p.dy = c->dy; p.dz = c->dz;
    std::memcpy(p.Tg2c, c->Tg2c, sizeof(p.Tg2c));
    std::memcpy(p.Tc2g, c->Tc2g, sizeof(p.Tc2g));
    p.rotate = sa.rotate;
    const gf_material& m = c->host_mats[material - 1];
    p.mat = material;
    p.alpha = m.FM_damping;
}

// Stage VJP on device buffers: y (store), h (its magnetostatic field or
// null), kbar -> mbar, hbar, dadj; gradient accumulators updated.
gf_status stage_vjp(gf_ctx* c, int material, int stage, double* const* y, double* const* h, double* const* kbar,
                    double* const* mbar, double* const* hbar, double* const* dadj, double* gpar, double* const* ghd) {
    AdjPointArgs p;
    stage_point_args(c, material, p);
    StageArgs sa = fm_stage_args(c, stage, 0, 1);
    fm_stage_drives(c, stage, sa);
    for (int i = 0; i < 3; i++) {
        p.y[i] = y[i]; p.hd[i] = h ? h[i] : nullptr; p.kbar[i] = kbar[i];
        p.mbar[i] = mbar[i]; p.hbar[i] = hbar[i];
        p.ghdrive[i] = p.hdrive[0] ? ghd[i] : nullptr;
        p.hext[i] = sa.hext[i];
    }
    if (!h) for (int i = 0; i < 3; i++) p.hd[i] = c->hstat[i];   // uploaded static field, no convolution
    p.drive_s = sa.drive_s;
    p.galpha = gpar; p.gaex = gpar + c->n; p.gb1 = gpar + 2 * c->n;
    if (narrow(c->n)) hipLaunchKernelGGL(adj_point_kernel<uint32_t>, dim3(blocks256(c->n)), dim3(256), 0, c->stream, p);
    else hipLaunchKernelGGL(adj_point_kernel<int64_t>, dim3(blocks256(c->n)), dim3(256), 0, c->stream, p);
    gf_status st = hip_check(c, hipGetLastError(), "adjoint point kernel");
    if (st) return st;
    if (h && dadj) return demag_adjoint(c, hbar, dadj);
    return GF_OK;
}

gf_status combine(gf_ctx* c, int mode, double* const* mbar, double* const* hbar, double* const* dadj,
                  double* const* lam, double* const* acc, double* const* kbar, double* const* ybar,
                  const int lo[3], const int hi[3], double g) {
    AdjCombineArgs a{};
    a.code = c->d_code; a.mats = c->d_mats;
    for (int i = 0; i < 3; i++) {
        a.mbar[i] = mbar[i]; a.hbar[i] = hbar[i]; a.dadj[i] = dadj ? dadj[i] : nullptr;
        a.lam[i] = lam ? lam[i] : nullptr; a.acc[i] = acc ? acc[i] : nullptr;
        a.kbar[i] = kbar ? kbar[i] : nullptr; a.ybar[i] = ybar ? ybar[i] : nullptr;
        a.lo[i] = lo[i]; a.hi[i] = hi[i];
    }
    a.nx = c->nx; a.ny = c->ny; a.nz = c->nz; a.dx = c->dx; a.dy = c->dy; a.dz = c->dz;
    a.mode = mode; a.g = g;
    if (narrow(c->n)) hipLaunchKernelGGL(adj_combine_kernel<uint32_t>, dim3(blocks256(c->n)), dim3(256), 0, c->stream, a);
    else hipLaunchKernelGGL(adj_combine_kernel<int64_t>, dim3(blocks256(c->n)), dim3(256), 0, c->stream, a);
    return hip_check(c, hipGetLastError(), "adjoint combine kernel");
}

gf_status sum_grads(gf_ctx* c, double* gpar, double grad[3]) {
    const double* f[3] = {gpar, gpar + c->n, gpar + 2 * c->n};
    gf_status st = hip_check(c, launch_sum3(f, c->n, c->red_partial, c->red_blocks, c->stream), "gradient sum");
    if (st) return st;
    std::vector<double> part(static_cast<size_t>(3 * c->red_blocks));
    st = hip_check(c, hipMemcpyAsync(part.data(), c->red_partial, sizeof(double) * part.size(), hipMemcpyDeviceToHost,
                                     c->stream), "gradient copy");
    if (!st) st = hip_check(c, hipStreamSynchronize(c->stream), "gradient sync");
    if (st) return st;
    for (int a = 0; a < 3; a++) {
        double s = 0.;
        for (int64_t b = 0; b < c->red_blocks; b++) s += part[static_cast<size_t>(3 * b + a)];
        grad[a] += s;
    }
    return GF_OK;
}

gf_status ensure_work(gf_ctx* c, AdjState* s) {
    gf_status st = GF_OK;
    if (!s->work) st = hip_check(c, hipMalloc(&s->work, sizeof(double) * 21 * static_cast<size_t>(c->n)), "hipMalloc adjoint work");
    if (!st && !s->gpar) st = hip_check(c, hipMalloc(&s->gpar, sizeof(double) * 3 * static_cast<size_t>(c->n)), "hipMalloc adjoint gradients");
    return st;
}

}  // namespace

void adjoint_free(gf_ctx* c) {
    free_state(c->adj);
    c->adj = nullptr;
}

}  // namespace gf

using namespace gf;

extern "C" {

gf_status gf_adjoint_forward(gf_ctx* c, const gf_adjoint_desc* d, double* loss) {
    if (!c || !d || !loss) return set_error(c, GF_ERR_ARG, "null argument");
    if (d->struct_size != sizeof(gf_adjoint_desc)) return set_error(c, GF_ERR_ARG, "gf_adjoint_desc size mismatch");
    if (d->nsteps < 1) return set_error(c, GF_ERR_ARG, "nsteps must be >= 1");
    if (gf_status e = check_context(c, d->material)) return e;
    const int64_t dims[3] = {c->nx, c->ny, c->nz};
    for (int a = 0; a < 3; a++)
        if (d->probe_lo[a] < 0 || d->probe_hi[a] < d->probe_lo[a] || d->probe_hi[a] >= dims[a])
            return set_error(c, GF_ERR_ARG, "probe box outside the grid");
    if (d->precision != 0 && d->precision != 64 && d->precision != 32)
        return set_error(c, GF_ERR_ARG, "precision must be 64 or 32");
    if (d->resident_mib < 0) return set_error(c, GF_ERR_ARG, "resident_mib must be >= 0");
    adjoint_free(c);
    adjoint32_free(c);
    if (d->precision == 32) return adjoint32_forward(c, d, loss);
    AdjState* s = new AdjState();
    c->adj = s;
    s->d = *d;
    s->N = d->nsteps;
    const int64_t n = c->n;
    const bool ms = demag_on(c);
    // resident steps: the stage stores and fields of each, within the budget
    const double per_step = static_cast<double>(ms ? 8 : 4) * 3. * static_cast<double>(n) * sizeof(double);
    const double budget = static_cast<double>(d->resident_mib) * 1048576.;
    s->M = std::min<int64_t>(s->N, static_cast<int64_t>(std::floor(budget / per_step)));
    const int64_t P = s->N - s->M;   // checkpointed steps
    s->K = d->checkpoint_every > 0 ? d->checkpoint_every
                                   : static_cast<int64_t>(std::ceil(std::sqrt(static_cast<double>(P))));
    s->K = std::max<int64_t>(1, std::min(s->K, P));
    s->nseg = (P + s->K - 1) / s->K;
    gf_status st = dalloc(c, &s->pool, (s->nseg + 1) * 3 * n, "hipMalloc adjoint checkpoints");
    if (!st) st = dalloc(c, &s->X, s->N, "hipMalloc probe means");
    if (!st && s->M > 0) st = dalloc(c, &s->RY, 4 * s->M * 3 * n, "hipMalloc adjoint resident stores");
    if (!st && s->M > 0 && ms) st = dalloc(c, &s->RH, 4 * s->M * 3 * n, "hipMalloc adjoint resident fields");
    if (st) { adjoint_free(c); return st; }
    s->ck_time.assign(static_cast<size_t>(s->nseg), 0.);
    s->tstep.assign(static_cast<size_t>(s->N), 0.);
    double* const* h_all[5] = {c->hstat, c->hstat, c->hstat, c->hstat, c->hstat};
    double* const* st3[3] = {c->st_a, c->st_b, c->st_a};
    // stage stores alternate a, b, a as in gf_step; the new state goes to m
    // (base m is read per cell before the cell's own stage-4 write)
    if (ms) st = demag_run_to(c, c->m, c->hstat);
    for (int64_t k = 0; k < s->N && !st; k++) {
        if (k >= P) {
            // a resident step: stage stores and fields kept where the
            // reverse sweep reads them; the new state into m and its field
            // into Hstat after the last one
            const int64_t j = k - P;
            double *ys[4][3], *hs[5][3];
            for (int q = 0; q < 4; q++) {
                slot3(s->RY, 4 * j + q, n, ys[q]);
                if (ms) slot3(s->RH, 4 * j + q, n, hs[q]);
                else for (int a = 0; a < 3; a++) hs[q][a] = c->hstat[a];
            }
            const bool last = k + 1 == s->N;
            if (ms && !last) slot3(s->RH, 4 * (j + 1), n, hs[4]);
            else for (int a = 0; a < 3; a++) hs[4][a] = c->hstat[a];
            double* yn[3];
            if (last) for (int a = 0; a < 3; a++) yn[a] = c->m[a];
            else slot3(s->RY, 4 * (j + 1), n, yn);
            if (j == 0) {
                st = copy3(c, ys[0], c->m);
                if (!st && ms) st = copy3(c, hs[0], c->hstat);
            }
            double* const* sts[3] = {ys[1], ys[2], ys[3]};
            double* const* hp[5] = {hs[0], hs[1], hs[2], hs[3], ms ? hs[4] : nullptr};
            if (!st) st = step_to(c, ys[0], sts, yn, hp);
            if (!st) {
                hipLaunchKernelGGL(probe_mean_kernel, dim3(1), dim3(256), 0, c->stream, yn[0], c->ny, c->nz,
                                   d->probe_lo[0], d->probe_lo[1], d->probe_lo[2], d->probe_hi[0], d->probe_hi[1],
                                   d->probe_hi[2], s->X + k);
                st = hip_check(c, hipGetLastError(), "probe mean");
            }
            s->tstep[static_cast<size_t>(k)] = c->time;
            continue;
        }
        if (k % s->K == 0) {
            double* ck[3];
            slot3(s->pool, k / s->K, n, ck);
            st = copy3(c, ck, c->m);
            s->ck_time[static_cast<size_t>(k / s->K)] = c->time;
        }
        if (!st) st = step_to(c, c->m, st3, c->m, h_all);
        if (!st) {
            hipLaunchKernelGGL(probe_mean_kernel, dim3(1), dim3(256), 0, c->stream, c->m[0], c->ny, c->nz,
                               d->probe_lo[0], d->probe_lo[1], d->probe_lo[2], d->probe_hi[0], d->probe_hi[1],
                               d->probe_hi[2], s->X + k);
            st = hip_check(c, hipGetLastError(), "probe mean");
        }
        s->tstep[static_cast<size_t>(k)] = c->time;
    }
    if (!st) {
        double* fin[3];
        slot3(s->pool, s->nseg, n, fin);
        st = copy3(c, fin, c->m);
        s->t_final = c->time;
    }
    std::vector<double> X(static_cast<size_t>(s->N));
    if (!st) st = hip_check(c, hipMemcpyAsync(X.data(), s->X, sizeof(double) * X.size(), hipMemcpyDeviceToHost,
                                              c->stream), "probe copy");
    if (!st) st = hip_check(c, hipStreamSynchronize(c->stream), "adjoint forward sync");
    c->mag_store_sel = 0;
    c->mag_next_stage = 1;
    c->demag.h_in_spec = false;
    c->demag.h_real_valid = true;
    c->demag.m_in_spec = false;
    if (st) { adjoint_free(c); return st; }
    const double w = 2. * kPi * d->probe_freq;
    std::complex<double> S(0., 0.);
    for (int64_t k = 0; k < s->N; k++)
        S += X[static_cast<size_t>(k)] * std::polar(c->dt, -w * s->tstep[static_cast<size_t>(k)]);
    s->loss = std::norm(S);
    const double np = static_cast<double>(d->probe_hi[0] - d->probe_lo[0] + 1) *
                      static_cast<double>(d->probe_hi[1] - d->probe_lo[1] + 1) *
                      static_cast<double>(d->probe_hi[2] - d->probe_lo[2] + 1);
    s->g.resize(static_cast<size_t>(s->N));
    for (int64_t k = 0; k < s->N; k++) {
        const std::complex<double> wk = std::polar(c->dt, -w * s->tstep[static_cast<size_t>(k)]);
        s->g[static_cast<size_t>(k)] = 2. * (std::conj(S) * wk).real() / np;
    }
    s->ready = true;
    *loss = s->loss;
    return GF_OK;
}

gf_status gf_adjoint_backward(gf_ctx* c, double grad[3], double* grad_hdrive, size_t count) {
    if (!c || !grad) return set_error(c, GF_ERR_ARG, "null argument");
    if (c->adj32) {
        if (gf_status e = check_context(c, adjoint32_material(c))) return e;
        return adjoint32_backward(c, grad, grad_hdrive, count);
    }
    AdjState* s = c->adj;
    if (!s || !s->ready) return set_error(c, GF_ERR_STATE, "gf_adjoint_backward needs a gf_adjoint_forward first");
    if (grad_hdrive && !(c->flags & GF_FLAG_HDRIVE))
        return set_error(c, GF_ERR_ARG, "grad_hdrive needs the shaped drive (GF_FLAG_HDRIVE)");
    if (grad_hdrive && count != static_cast<size_t>(3 * c->n)) return set_error(c, GF_ERR_ARG, "count must be 3n");
    if (gf_status e = check_context(c, s->d.material)) return e;
    const int64_t n = c->n, K = s->K;
    const bool ms = demag_on(c);
    gf_status st = ensure_work(c, s);
    if (!st && s->nseg > 0 && s->kbuf != K) {
        st = dalloc(c, &s->Y, (4 * K + 1) * 3 * n, "hipMalloc adjoint stage stores");
        if (!st && ms) st = dalloc(c, &s->HD, 4 * K * 3 * n, "hipMalloc adjoint stage fields");
        s->kbuf = st ? 0 : K;
    }
    if (st) return st;
    double *lam[3], *acc[3], *kbar[3], *mbar[3], *hbar[3], *dadj[3], *ghd[3];
    slot3(s->work, 0, n, lam); slot3(s->work, 1, n, acc); slot3(s->work, 2, n, kbar);
    slot3(s->work, 3, n, mbar); slot3(s->work, 4, n, hbar); slot3(s->work, 5, n, dadj); slot3(s->work, 6, n, ghd);
    st = hip_check(c, hipMemsetAsync(s->work, 0, sizeof(double) * 21 * static_cast<size_t>(n), c->stream), "clear");
    if (!st) st = hip_check(c, hipMemsetAsync(s->gpar, 0, sizeof(double) * 3 * static_cast<size_t>(n), c->stream), "clear");
    const int* lo = s->d.probe_lo;
    const int* hi = s->d.probe_hi;
    if (!st) {
        hipLaunchKernelGGL(adj_init_kernel, dim3(blocks256(n)), dim3(256), 0, c->stream, c->d_code, lam[0], kbar[0],
                           c->nx, c->ny, c->nz, lo[0], lo[1], lo[2], hi[0], hi[1], hi[2], s->g.back());
        st = hip_check(c, hipGetLastError(), "adjoint init");
    }
    // one step of the reverse sweep over its four stage stores and fields
    auto walk_step = [&](int64_t step, double* ybuf, double* hbuf, int64_t j) {
        gf_status e = GF_OK;
        for (int stg = 4; stg >= 1 && !e; stg--) {
            double *ys[3], *hs[3];
            slot3(ybuf, 4 * j + (stg - 1), n, ys);
            if (ms) slot3(hbuf, 4 * j + (stg - 1), n, hs);
            e = stage_vjp(c, s->d.material, stg, ys, ms ? hs : nullptr, kbar, mbar, hbar, ms ? dadj : nullptr,
                          s->gpar, ghd);
            const double g = (stg == 1 && step >= 1) ? s->g[static_cast<size_t>(step - 1)] : 0.;
            if (!e) e = combine(c, stg, mbar, hbar, ms ? dadj : nullptr, lam, acc, kbar, nullptr, lo, hi, g);
        }
        return e;
    };
    const int64_t P = s->N - s->M;
    for (int64_t j = s->M - 1; j >= 0 && !st; j--) {
        c->time = s->tstep[static_cast<size_t>(P + j)];
        st = walk_step(P + j, s->RY, s->RH, j);
    }
    std::vector<double> tseg(static_cast<size_t>(K));
    for (int64_t seg = s->nseg - 1; seg >= 0 && !st; seg--) {
        const int64_t n0 = seg * K, n1 = std::min(P, n0 + K), J = n1 - n0;
        // recompute the segment, keeping every stage store and field
        double* y0[3];
        slot3(s->Y, 0, n, y0);
        double* ck[3];
        slot3(s->pool, seg, n, ck);
        st = copy3(c, y0, ck);
        c->time = s->ck_time[static_cast<size_t>(seg)];
        if (!st && ms) {
            double* h0[3];
            slot3(s->HD, 0, n, h0);
            st = demag_run_to(c, y0, h0);
        }
        for (int64_t j = 0; j < J && !st; j++) {
            double *ys[4][3], *yn[3], *hs[5][3];
            for (int q = 0; q < 4; q++) slot3(s->Y, 4 * j + q, n, ys[q]);
            slot3(s->Y, 4 * (j + 1), n, yn);
            for (int q = 0; q < 4; q++) {
                if (ms) slot3(s->HD, 4 * j + q, n, hs[q]);
                else for (int a = 0; a < 3; a++) hs[q][a] = c->hstat[a];
            }
            const bool more = j + 1 < J;
            if (ms && more) slot3(s->HD, 4 * (j + 1), n, hs[4]);
            double* const* sts[3] = {ys[1], ys[2], ys[3]};
            double* const* hp[5] = {hs[0], hs[1], hs[2], hs[3], (ms && more) ? hs[4] : nullptr};
            if (!ms) hp[4] = nullptr;
            st = step_to(c, ys[0], sts, yn, hp);
            tseg[static_cast<size_t>(j)] = c->time;
        }
        // walk the segment backwards
        for (int64_t j = J - 1; j >= 0 && !st; j--) {
            c->time = tseg[static_cast<size_t>(j)];
            st = walk_step(n0 + j, s->Y, s->HD, j);   // step n0 + j maps y_step -> y_step+1
        }
    }
    // back to the forward's final state with its field
    if (!st) {
        double* fin[3];
        slot3(s->pool, s->nseg, n, fin);
        st = copy3(c, c->m, fin);
        c->time = s->t_final;
        if (!st && ms) st = demag_run_to(c, c->m, c->hstat);
        c->mag_store_sel = 0;
        c->mag_next_stage = 1;
        c->demag.h_in_spec = false;
        c->demag.h_real_valid = true;
        c->demag.m_in_spec = false;
    }
    grad[0] = grad[1] = grad[2] = 0.;
    if (!st) st = sum_grads(c, s->gpar, grad);
    if (!st && grad_hdrive)
        for (int a = 0; a < 3 && !st; a++)
            st = hip_check(c, hipMemcpyAsync(grad_hdrive + a * n, ghd[a], sizeof(double) * static_cast<size_t>(n),
                                             hipMemcpyDeviceToHost, c->stream), "gradient copy");
    if (!st) st = hip_check(c, hipStreamSynchronize(c->stream), "adjoint backward sync");
    return st;
}

gf_status gf_adjoint_stage(gf_ctx* c, int material, int stage, const double* store, const double* kbar_h,
                           double* ybar_h, double grad[3], double* grad_hdrive, size_t count) {
    if (!c || !store || !kbar_h || !ybar_h || !grad) return set_error(c, GF_ERR_ARG, "null argument");
    if (stage < 1 || stage > 4) return set_error(c, GF_ERR_ARG, "stage must be 1..4");
    if (count != static_cast<size_t>(3 * c->n)) return set_error(c, GF_ERR_ARG, "count must be 3n");
    if (grad_hdrive && !(c->flags & GF_FLAG_HDRIVE))
        return set_error(c, GF_ERR_ARG, "grad_hdrive needs the shaped drive (GF_FLAG_HDRIVE)");
    if (gf_status e = check_context(c, material)) return e;
    // the convolution buffers are reused below: a field kept only as a
    // spectrum is written out first
    if (c->demag.h_in_spec && !c->demag.h_real_valid) {
        if (gf_status e = demag_materialize_h(c)) return e;
    }
    if (!c->adj) c->adj = new AdjState();
    AdjState* s = c->adj;
    const int64_t n = c->n;
    gf_status st = ensure_work(c, s);
    if (st) return st;
    double *y[3], *h[3], *kb[3], *mbar[3], *hbar[3], *dadj[3], *ghd[3];
    slot3(s->work, 0, n, y); slot3(s->work, 1, n, h); slot3(s->work, 2, n, kb);
    slot3(s->work, 3, n, mbar); slot3(s->work, 4, n, hbar); slot3(s->work, 5, n, dadj); slot3(s->work, 6, n, ghd);
    for (int a = 0; a < 3 && !st; a++) {
        st = hip_check(c, hipMemcpyAsync(y[a], store + a * n, sizeof(double) * static_cast<size_t>(n),
                                         hipMemcpyHostToDevice, c->stream), "store copy");
        if (!st) st = hip_check(c, hipMemcpyAsync(kb[a], kbar_h + a * n, sizeof(double) * static_cast<size_t>(n),
                                                  hipMemcpyHostToDevice, c->stream), "kbar copy");
    }
    if (!st) st = hip_check(c, hipMemsetAsync(s->gpar, 0, sizeof(double) * 3 * static_cast<size_t>(n), c->stream), "clear");
    if (!st) st = hip_check(c, hipMemsetAsync(ghd[0], 0, sizeof(double) * 3 * static_cast<size_t>(n), c->stream), "clear");
    const bool ms = demag_on(c);
    if (!st && ms) st = demag_run_to(c, y, h);
    if (!st) st = stage_vjp(c, material, stage, y, ms ? h : nullptr, kb, mbar, hbar, ms ? dadj : nullptr, s->gpar, ghd);
    const int lo[3] = {0, 0, 0}, hi[3] = {-1, -1, -1};
    double* ybar[3];
    slot3(s->work, 1, n, ybar);   // the field slot is free again
    if (!st) st = combine(c, 0, mbar, hbar, ms ? dadj : nullptr, nullptr, nullptr, nullptr, ybar, lo, hi, 0.);
    for (int a = 0; a < 3 && !st; a++)
        st = hip_check(c, hipMemcpyAsync(ybar_h + a * n, ybar[a], sizeof(double) * static_cast<size_t>(n),
                                         hipMemcpyDeviceToHost, c->stream), "ybar copy");
    if (!st && grad_hdrive)
        for (int a = 0; a < 3 && !st; a++)
            st = hip_check(c, hipMemcpyAsync(grad_hdrive + a * n, ghd[a], sizeof(double) * static_cast<size_t>(n),
                                             hipMemcpyDeviceToHost, c->stream), "gradient copy");
    if (!st) {
        double g3[3] = {0., 0., 0.};
        st = sum_grads(c, s->gpar, g3);
        for (int a = 0; a < 3; a++) grad[a] += g3[a];
    }
    if (!st) st = hip_check(c, hipStreamSynchronize(c->stream), "adjoint stage sync");
    // the convolution buffers were used: the context's field is rebuilt by
    // its next statics; keep the flags honest
    c->demag.h_in_spec = false;
    c->demag.h_real_valid = true;
    c->demag.m_in_spec = false;
    return st;
}

}  // extern "C"
