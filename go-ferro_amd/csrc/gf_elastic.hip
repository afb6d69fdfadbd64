// gf_elastic.hip — elastodynamics RK4 stage for gfx950 (FP64).
//
// One stage of the reference's elastic solver (elastic_dynamic_solver.cpp,
// elastodynamics_AbsorbingBoundary.cpp) is up to four launches:
//   history   stage 4 only: the boundary planes of u into the history ring
//             (transfer_pointer, elastodynamics_AbsorbingBoundary.cpp:3-89)
//   prep      face gradients of the u store and the cell strain (:107-227),
//             the eigenstrain from the m, m_AFM and p stores (:4-104) with
//             back-action, and each cell's part of its face stresses
//   force     stress divergence, dv and du = v_s dt, the v and u RK updates
//             (get_dv_RKs / get_du_RKs / update_v_RKs / update_u_RKs, :841-3057)
//   boundary  Liao or Mur planes and the displacement source
//             (elastodynamics_AbsorbingBoundary.cpp:92-230, :3060-3284)
// The force launch reads the u store around each cell (its face gradients)
// and writes the next store into a second array (the two are swapped after
// it); v and its store are read and written by their own cell's thread only.
//
// RK4 bookkeeping keeps a state accumulator per field, acc = y + k1/6,
// acc = acc + k2/3, acc = acc + k3/3, y = acc + k4/6, which is the
// reference's left-to-right y + k1/6 + k2/3 + k3/3 + k4/6 (the weights applied
// as products with the rounded reciprocals, within about an ulp of the quotients).
#include "gf_internal.h"

namespace gf {

namespace {

__device__ __forceinline__ bool solid(const uint32_t* code, int64_t id) { return (code[id] & kMatMask) != 0u; }

__device__ __forceinline__ void rot_vec(const double T[3][3], double ix, double iy, double iz, double& ox,
                                        double& oy, double& oz) {
    ox = T[0][0] * ix + T[0][1] * iy + T[0][2] * iz;
    oy = T[1][0] * ix + T[1][1] * iy + T[1][2] * iz;
    oz = T[2][0] * ix + T[2][1] * iy + T[2][2] * iz;
}

// o_ij = sum_kl T_ik T_jl s_kl of a symmetric tensor given in Voigt order,
// result in Voigt order (mathlib.cpp:78-125, summed k-major as there).
__device__ __forceinline__ void rot_voigt(const double T[3][3], const double v[6], double o[6]) {
    const double s[3][3] = {{v[0], v[5], v[4]}, {v[5], v[1], v[3]}, {v[4], v[3], v[2]}};
    const int oi[6] = {0, 1, 2, 1, 0, 0};
    const int oj[6] = {0, 1, 2, 2, 2, 1};
    for (int c = 0; c < 6; c++) {
        const int i = oi[c], j = oj[c];
        double r = 0.;
        for (int k = 0; k < 3; k++)
            for (int l = 0; l < 3; l++) r = r + T[i][k] * T[j][l] * s[k][l];
        o[c] = r;
    }
}

// ---------------------------------------------------------------- strain
// (i, j, k) of cell id = (i ny + j) nz + k, by 32-bit divisions when the
// index and ny nz fit 32 bits (64-bit divisions are emulated and slow).
__device__ __forceinline__ void cell_ijk(int64_t id, int64_t ny, int64_t nz, int64_t& i, int64_t& j, int64_t& k) {
    const int64_t nyz = ny * nz;
    if (((id | nyz) >> 32) == 0) {
        const uint32_t u = static_cast<uint32_t>(id), q = static_cast<uint32_t>(nyz), z = static_cast<uint32_t>(nz);
        const uint32_t ii = u / q, r = u - ii * q, jj = r / z;
        i = ii; j = jj; k = r - jj * z;
    } else {
        i = id / nyz; j = (id - i * nyz) / nz; k = id - i * nyz - j * nz;
    }
}

// Face gradients g[face][comp] of cell id = (i, j, k) of the u store,
// faces ordered x-, x+, y-, y+, z-, z+.  A stress-free face (vacuum on a
// side, or a non-periodic wall) has zero gradient; across a periodic wall
// the difference wraps (elastic_dynamic_solver.cpp:107-187).  The
// differences are multiplied by the reciprocal spacings (within an ulp of
// the reference's quotients); the upper face of a cell and the lower face
// of the cell above it are the same expression, hence the same value.
// The cell's face gradients from the u store.  Every load is issued before
// any code is known (all indices are in bounds), so that one memory round
// trip feeds the cell; the launches that call it run consecutive cells in
// consecutive lanes (one cell per thread, none past the grid's last cell),
// so the z neighbours' values and codes come from the adjacent lanes (DPP
// wave shifts) except at a column end or a wave edge, where they are
// loaded.
__device__ __forceinline__ void cell_grad(const ElStrainArgs& a, int64_t id, int64_t i, int64_t j, int64_t k,
                                          double (&g)[6][3]) {
    const int64_t nx = a.nx, ny = a.ny, nz = a.nz;
    const int64_t nb[6] = {((i == 0 ? nx - 1 : i - 1) * ny + j) * nz + k, ((i == nx - 1 ? 0 : i + 1) * ny + j) * nz + k,
                           (i * ny + (j == 0 ? ny - 1 : j - 1)) * nz + k, (i * ny + (j == ny - 1 ? 0 : j + 1)) * nz + k,
                           (i * ny + j) * nz + (k == 0 ? nz - 1 : k - 1), (i * ny + j) * nz + (k == nz - 1 ? 0 : k + 1)};
    const uint32_t own = a.code[id];
    double u[3], un[6][3];
    uint32_t cn[6];
    for (int c = 0; c < 3; c++) u[c] = a.us[c][id];
    for (int f = 0; f < 4; f++) {
        for (int c = 0; c < 3; c++) un[f][c] = a.us[c][nb[f]];
        cn[f] = a.code[nb[f]];
    }
    const int lane = static_cast<int>(threadIdx.x & 63u);
    const bool gm = lane == 0 || k == 0, gp = lane == 63 || k == nz - 1;
    if (gm) {
        for (int c = 0; c < 3; c++) un[4][c] = a.us[c][nb[4]];
        cn[4] = a.code[nb[4]];
    }
    if (gp) {
        for (int c = 0; c < 3; c++) un[5][c] = a.us[c][nb[5]];
        cn[5] = a.code[nb[5]];
    }
    for (int c = 0; c < 3; c++) {
        const double dn = lane_up(u[c]), up = lane_down(u[c]);
        if (!gm) un[4][c] = dn;
        if (!gp) un[5][c] = up;
    }
    const uint32_t cdn = lane_up(own), cup = lane_down(own);
    if (!gm) cn[4] = cdn;
    if (!gp) cn[5] = cup;
    const bool me = (own & kMatMask) != 0u;
    const bool wall[6] = {i == 0 && !a.per[0], i == nx - 1 && !a.per[0], j == 0 && !a.per[1],
                          j == ny - 1 && !a.per[1], k == 0 && !a.per[2], k == nz - 1 && !a.per[2]};
    const double ih[3] = {a.idx, a.idy, a.idz};
    for (int f = 0; f < 6; f++) {
        const bool fr = wall[f] || !(me && (cn[f] & kMatMask) != 0u);
        for (int c = 0; c < 3; c++) {
            if (fr) g[f][c] = 0.;
            else if (f & 1) g[f][c] = (un[f][c] - u[c]) * ih[f >> 1];
            else g[f][c] = (u[c] - un[f][c]) * ih[f >> 1];
        }
    }
}

// The face arrays: a cell writes its lower faces, the last cell along an
// axis its upper face too.
__device__ __forceinline__ void grad_out(const ElStrainArgs& a, int64_t i, int64_t j, int64_t k,
                                         const double (&g)[6][3]) {
    const int64_t nx = a.nx, ny = a.ny, nz = a.nz;
    for (int c = 0; c < 3; c++) {
        a.gyz[c][(i * ny + j) * nz + k] = g[0][c];
        if (i == nx - 1) a.gyz[c][(nx * ny + j) * nz + k] = g[1][c];
        a.gxz[c][(i * (ny + 1) + j) * nz + k] = g[2][c];
        if (j == ny - 1) a.gxz[c][(i * (ny + 1) + ny) * nz + k] = g[3][c];
        a.gxy[c][(i * ny + j) * (nz + 1) + k] = g[4][c];
        if (k == nz - 1) a.gxy[c][(i * ny + j) * (nz + 1) + nz] = g[5][c];
    }
}

// Cell strain as the average of its face gradients (:190-227).
__device__ __forceinline__ void strain_of(const double (&g)[6][3], double (&e)[6]) {
    e[0] = (g[0][0] + g[1][0]) / 2.;
    e[1] = (g[2][1] + g[3][1]) / 2.;
    e[2] = (g[4][2] + g[5][2]) / 2.;
    e[3] = (g[4][1] + g[5][1] + g[2][2] + g[3][2]) / 4.;
    e[4] = (g[4][0] + g[5][0] + g[0][2] + g[1][2]) / 4.;
    e[5] = (g[2][0] + g[3][0] + g[0][1] + g[1][1]) / 4.;
}

// The strain out, global frame and (rotated) crystal frame.
__device__ __forceinline__ void strain_out(const ElStrainArgs& a, int64_t id, const double (&e)[6]) {
    for (int c = 0; c < 6; c++) a.eg[c][id] = e[c];
    if (a.ec[0] != a.eg[0]) {
        double o[6];
        if (a.rotate) rot_voigt(a.Tg2c, e, o);
        else for (int c = 0; c < 6; c++) o[c] = e[c];
        for (int c = 0; c < 6; c++) a.ec[c][id] = o[c];
    }
}

__global__ void __launch_bounds__(256) el_strain_kernel(ElStrainArgs a) {
    const int64_t id = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (id >= a.nx * a.ny * a.nz) return;
    int64_t i, j, k;
    cell_ijk(id, a.ny, a.nz, i, j, k);
    double g[6][3], e[6];
    cell_grad(a, id, i, j, k, g);
    grad_out(a, i, j, k, g);
    strain_of(g, e);
    strain_out(a, id, e);
}

// ----------------------------------------------------------- eigenstrain
// Eigenstrain of cell id from the stage stores (elastic_dynamic_solver.cpp:
// 4-104) less the static eigenstrain at t0 when there is one: s in the
// crystal frame, o in the global frame, zero for vacuum.
// kGeneral: with rotation and the flexoelectric terms; the lean form is
// used only when the context has neither.  code is the cell's code; lc, when
// set, the workgroup's LDS copy of every material's constants (kEigC each:
// l100 l111 l100a l111a l100b l111b Q11 Q12 Q44).  The stores are read
// before the code is known (every index is in bounds), so one memory round
// trip feeds the cell.
constexpr int kEigC = 9;

template <bool kGeneral = true>
__device__ __forceinline__ void eigenstrain(const ElEigenArgs& a, const double* lc, int64_t id, uint32_t code,
                                            double s[6], double o[6]) {
    const uint32_t t = code & kMatMask;
    double v[4][3];
    const double* const* src[4] = {a.p, a.m, a.m1, a.m2};
    for (int f = 0; f < 4; f++) {
        v[f][0] = 0.; v[f][1] = 0.; v[f][2] = 0.;
        if (src[f][0]) { v[f][0] = src[f][0][id]; v[f][1] = src[f][1][id]; v[f][2] = src[f][2][id]; }
    }
    double t0[6] = {0., 0., 0., 0., 0., 0.};
    if (a.e0t0[0])
        for (int c = 0; c < 6; c++) t0[c] = a.e0t0[c][id];
    for (int c = 0; c < 6; c++) { s[c] = 0.; o[c] = 0.; }
    if (t) {
        double C[kEigC];
        if (lc) {
            for (int q = 0; q < kEigC; q++) C[q] = lc[(t - 1) * kEigC + q];
        } else {
            const DevMaterial& M = a.mats[t - 1];
            C[0] = M.l100; C[1] = M.l111; C[2] = M.l100a; C[3] = M.l111a; C[4] = M.l100b; C[5] = M.l111b;
            C[6] = M.Q11; C[7] = M.Q12; C[8] = M.Q44;
        }
        for (int f = 0; f < 4; f++) {
            const double x = v[f][0], y = v[f][1], z = v[f][2];
            if (kGeneral && a.rotate) rot_vec(a.Tg2c, x, y, z, v[f][0], v[f][1], v[f][2]);
        }
        const double px = v[0][0], py = v[0][1], pz = v[0][2];
        const double mx = v[1][0], my = v[1][1], mz = v[1][2];
        const double ax = v[2][0], ay = v[2][1], az = v[2][2];
        const double bx = v[3][0], by = v[3][1], bz = v[3][2];
        s[0] = 1.5 * C[0] * (mx * mx - 1. / 3.) + C[6] * (px * px) + C[7] * (py * py + pz * pz) +
               0.5 * 1.5 * C[2] * (ax * ax - 1. / 3.) + 0.5 * 1.5 * C[4] * (bx * bx - 1. / 3.) - t0[0];
        s[1] = 1.5 * C[0] * (my * my - 1. / 3.) + C[6] * (py * py) + C[7] * (px * px + pz * pz) +
               0.5 * 1.5 * C[2] * (ay * ay - 1. / 3.) + 0.5 * 1.5 * C[4] * (by * by - 1. / 3.) - t0[1];
        s[2] = 1.5 * C[0] * (mz * mz - 1. / 3.) + C[6] * (pz * pz) + C[7] * (py * py + px * px) +
               0.5 * 1.5 * C[2] * (az * az - 1. / 3.) + 0.5 * 1.5 * C[4] * (bz * bz - 1. / 3.) - t0[2];
        s[3] = 1.5 * C[1] * my * mz + C[8] * py * pz + 0.5 * 1.5 * C[3] * ay * az + 0.5 * 1.5 * C[5] * by * bz - t0[3];
        s[4] = 1.5 * C[1] * mx * mz + C[8] * px * pz + 0.5 * 1.5 * C[3] * ax * az + 0.5 * 1.5 * C[5] * bx * bz - t0[4];
        s[5] = 1.5 * C[1] * my * mx + C[8] * py * px + 0.5 * 1.5 * C[3] * ay * ax + 0.5 * 1.5 * C[5] * by * bx - t0[5];
        if (kGeneral && a.flexo) {
            // -F grad p (:82-89): the global-frame gradient of FE cells,
            // zero elsewhere, enters the crystal-frame tensor
            const FlexoMat& F = a.flexo[t - 1];
            double g[3][3] = {{0., 0., 0.}, {0., 0., 0.}, {0., 0., 0.}};
            const uint32_t own = a.code[id];
            if (own & kIsFE) {
                const int64_t nyz = a.ny * a.nz;
                const int64_t i = id / nyz, j = (id - i * nyz) / a.nz, k = id - i * nyz - j * a.nz;
                const double h[3] = {a.dx, a.dy, a.dz};
                fe_p_gradient(own, a.p, id, i, j, k, a.nx, a.ny, a.nz, h, g);
            }
            s[0] = s[0] - F.F11 * g[0][0] - F.F12 * (g[2][2] + g[1][1]);
            s[1] = s[1] - F.F11 * g[1][1] - F.F12 * (g[2][2] + g[0][0]);
            s[2] = s[2] - F.F11 * g[2][2] - F.F12 * (g[0][0] + g[1][1]);
            s[3] = s[3] - F.F44 / 2. * (g[1][2] + g[2][1]);
            s[4] = s[4] - F.F44 / 2. * (g[0][2] + g[2][0]);
            s[5] = s[5] - F.F44 / 2. * (g[1][0] + g[0][1]);
        }
        if (kGeneral && a.rotate) rot_voigt(a.Tc2g, s, o);
        else for (int c = 0; c < 6; c++) o[c] = s[c];
    }
}

// ----------------------------------------------------------------- force
__device__ __forceinline__ double crow(const double (*C)[6], int r, const double e[6]) {
    double s = 0.;
    for (int j = 0; j < 6; j++) s = s + C[r][j] * e[j];
    return s;
}

// The stress the reference evaluates on a face from the strain of one cell
// beside it is C e with (elastic_dynamic_solver.cpp:876-933, 994-1050,
// 1112-1166, 1229-1290)
//   YZ face (rows xx, xy, xz): e = [gx_f - e0xx, eyy - e0yy, ezz - e0zz,
//       2(eyz - e0yz), (gx_zb + gx_zt + 2 gz_f)/2 - 2 e0xz, (gx_yb + gx_yt + 2 gy_f)/2 - 2 e0xy]
//   XZ face (rows xy, yy, yz) and XY face (rows xz, yz, zz) likewise,
// where g?_f is the face's own gradient and the others belong to the cell
// (its strain e, eigenstrain e0, the gradients on its faces across the
// axis).  C e is linear in e, so it splits into a part of the cell, P =
// C (e with the face gradients zero), and C fp with fp holding the face
// gradients: the face gradient of component c of the face's axis a enters
// strain entry (a, c).  The face stress is the mean of its two cells'
// evaluations, (P_q + C_q fp + P_n + C_n fp) / 2: the prep launch writes
// P^x, P^y, P^z of every cell, the force launch forms the face stresses.

// Rows of C per face axis, and the strain entries the face gradient
// components (x, y, z) enter.
__constant__ int kFaceRows[3][3] = {{0, 5, 4}, {5, 1, 3}, {4, 3, 2}};
__constant__ int kFaceCols[3][3] = {{0, 5, 4}, {5, 1, 3}, {4, 3, 2}};

// P^ax of a cell (its strain e, eigenstrain e0, face gradients g; C its
// stiffness): the strain of its faces along ax with their own gradients
// zero, times the rows of C those faces carry.
__device__ __forceinline__ void p_axis(const double (*C)[6], int ax, const double (&g)[6][3], const double (&e)[6],
                                       const double (&e0)[6], double (&P)[3]) {
    double b[6];
    if (ax == 0) {
        b[0] = -e0[0]; b[1] = e[1] - e0[1]; b[2] = e[2] - e0[2]; b[3] = (e[3] - e0[3]) * 2.;
        b[4] = ((g[4][0] + g[5][0]) / 4. - e0[4]) * 2.; b[5] = ((g[2][0] + g[3][0]) / 4. - e0[5]) * 2.;
    } else if (ax == 1) {
        b[0] = e[0] - e0[0]; b[1] = -e0[1]; b[2] = e[2] - e0[2]; b[3] = ((g[4][1] + g[5][1]) / 4. - e0[3]) * 2.;
        b[4] = (e[4] - e0[4]) * 2.; b[5] = ((g[0][1] + g[1][1]) / 4. - e0[5]) * 2.;
    } else {
        b[0] = e[0] - e0[0]; b[1] = e[1] - e0[1]; b[2] = -e0[2]; b[3] = ((g[2][2] + g[3][2]) / 4. - e0[3]) * 2.;
        b[4] = ((g[0][2] + g[1][2]) / 4. - e0[4]) * 2.; b[5] = (e[5] - e0[5]) * 2.;
    }
    for (int o = 0; o < 3; o++) P[o] = crow(C, kFaceRows[ax][o], b);
}

__device__ __forceinline__ const double (*stiffness(const double* ctab, uint32_t code))[6] {
    return reinterpret_cast<const double (*)[6]>(ctab + ((code & kMatMask) - 1u) * 36u);
}

// The workgroup's LDS copy of every material's constants: the stiffness
// matrices (36 each), then the eigenstrain constants (kEigC each), then
// el_rho, el_mass, el_stiff (kElC each).
constexpr int kElC = 3;
constexpr int kLdsPerMat = 36 + kEigC + kElC;

__device__ __forceinline__ void load_ctab(double* ctab, const DevMaterial* mats, int nmat) {
    for (int t = threadIdx.x; t < 36 * nmat; t += blockDim.x) ctab[t] = mats[t / 36].cij[(t % 36) / 6][t % 6];
    double* ec = ctab + 36 * nmat;
    for (int t = threadIdx.x; t < kEigC * nmat; t += blockDim.x) {
        const DevMaterial& M = mats[t / kEigC];
        const int q = t % kEigC;
        ec[t] = q == 0 ? M.l100 : q == 1 ? M.l111 : q == 2 ? M.l100a : q == 3 ? M.l111a : q == 4 ? M.l100b
              : q == 5 ? M.l111b : q == 6 ? M.Q11 : q == 7 ? M.Q12 : M.Q44;
    }
    double* lc = ec + kEigC * nmat;
    for (int t = threadIdx.x; t < kElC * nmat; t += blockDim.x) {
        const DevMaterial& M = mats[t / kElC];
        const int q = t % kElC;
        lc[t] = q == 0 ? M.el_rho : q == 1 ? M.el_mass : M.el_stiff;
    }
    __syncthreads();
}

// ---------------------------------------------------------------- prep
// The strain of the u store and the eigenstrain of the stage stores, both
// written out, and the cell's P^x, P^y, P^z.  kGeneral: rotation and the
// flexoelectric terms (the face gradients are written out for the latter).
template <bool kGeneral>
__global__ void __launch_bounds__(256) el_prep_kernel(ElStageArgs a) {
    extern __shared__ double ctab[];   // the materials' constants (load_ctab)
    load_ctab(ctab, a.e.mats, a.f.nmat);
    const int64_t id = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (id >= a.s.nx * a.s.ny * a.s.nz) return;
    int64_t i, j, k;
    cell_ijk(id, a.s.ny, a.s.nz, i, j, k);
    double g[6][3], e[6], e0[6] = {0., 0., 0., 0., 0., 0.}, sc[6];
    const uint32_t code = a.s.code[id];
    // the eigenstrain first: its loads then go out with the store's, before
    // any of the launch's writes
    if (a.eig) eigenstrain<kGeneral>(a.e, ctab + 36 * a.f.nmat, id, code, sc, e0);
    cell_grad(a.s, id, i, j, k, g);
    if (kGeneral && a.s.gyz[0]) grad_out(a.s, i, j, k, g);
    strain_of(g, e);
    // the strain array is read by the magnets and ferroelectrics at their own
    // cells (and, for the flexoelectric force, around them: the general form);
    // a download recomputes it everywhere
    if (kGeneral || (code & (kIsFM | kIsAFM | kIsFE))) strain_out(a.s, id, e);
    if (a.eig && a.e0_out) {
        for (int c = 0; c < 6; c++) a.e.e0g[c][id] = e0[c];
        if (a.e.e0c[0] != a.e.e0g[0])
            for (int c = 0; c < 6; c++) a.e.e0c[c][id] = sc[c];
    }
    double P[3][3] = {{0., 0., 0.}, {0., 0., 0.}, {0., 0., 0.}};
    if (code & kMatMask) {
        const double(*C)[6] = stiffness(ctab, code);
        p_axis(C, 0, g, e, e0, P[0]);
        p_axis(C, 1, g, e, e0, P[1]);
        p_axis(C, 2, g, e, e0, P[2]);
    }
    for (int o = 0; o < 3; o++) {
        a.px[o][id] = P[0][o];
        a.py[o][id] = P[1][o];
        a.pz[o][id] = P[2][o];
    }
}

// ----------------------------------------------------------------- force
// The stress of each face from the P's of its two cells, its gradient (from
// the u store the prep launch read) and the two stiffnesses; dv; the v and
// u RK updates, the next u store into f.us.
__global__ void __launch_bounds__(256) el_force_kernel(ElStageArgs a) {
    extern __shared__ double ctab[];   // the materials' constants (load_ctab)
    const ElForceArgs& f = a.f;
    load_ctab(ctab, f.mats, f.nmat);
    const int64_t nx = f.nx, ny = f.ny, nz = f.nz;
    const int64_t id = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (id >= nx * ny * nz) return;
    int64_t i, j, k;
    cell_ijk(id, ny, nz, i, j, k);
    const uint32_t code = f.code[id];
    const int stage = f.stage;
    double dv[3] = {0., 0., 0.};
    if (code & kMatMask) {
        const double(*Cq)[6] = stiffness(ctab, code);
        const int64_t ib = i == 0 ? nx - 1 : i - 1, iff = i == nx - 1 ? 0 : i + 1;
        const int64_t jb = j == 0 ? ny - 1 : j - 1, jf = j == ny - 1 ? 0 : j + 1;
        const int64_t kb = k == 0 ? nz - 1 : k - 1;
        // the reference wraps the forward z neighbour on k == ny - 1
        // (elastic_dynamic_solver.cpp:1247); set-up rejects the grids where
        // that differs from the periodic wrap and reads outside a column
        const int64_t kf = k == ny - 1 ? 0 : k + 1;
        const int64_t kt = k == nz - 1 ? 0 : k + 1;   // the true forward z neighbour
        const int64_t nb[6] = {(ib * ny + j) * nz + k, (iff * ny + j) * nz + k, (i * ny + jb) * nz + k,
                               (i * ny + jf) * nz + k, (i * ny + j) * nz + kb, (i * ny + j) * nz + kf};
        const int64_t nbt = (i * ny + j) * nz + kt;
        const bool fr[6] = {(i == 0 && !f.per[0]) || !solid(f.code, nb[0]),
                            (i == nx - 1 && !f.per[0]) || !solid(f.code, nb[1]),
                            (j == 0 && !f.per[1]) || !solid(f.code, nb[2]),
                            (j == ny - 1 && !f.per[1]) || !solid(f.code, nb[3]),
                            (k == 0 && !f.per[2]) || !solid(f.code, nb[4]),
                            (k == nz - 1 && !f.per[2]) || !solid(f.code, nbt)};
        const double ih[3] = {f.idx, f.idy, f.idz};
        const double* const* Pa[3] = {a.px, a.py, a.pz};
        double d[3][3];
        for (int ax = 0; ax < 3; ax++) {
            double s1[3] = {0., 0., 0.}, s2[3] = {0., 0., 0.};
            for (int side = 0; side < 2; side++) {
                const int fc = 2 * ax + side;
                double* sf = side ? s2 : s1;
                if (fr[fc]) {
                    // Gaussian traction on a stress-free top face (input_strain_type 3)
                    if (fc == 5 && k == f.trac_k) sf[f.trac_row] = f.trac_amp;
                    continue;
                }
                // the face gradient, across to the face's true neighbour (the
                // value the cell on either side computes for it)
                const int64_t n = fc == 5 ? nbt : nb[fc];
                double fg[3];
                for (int c = 0; c < 3; c++)
                    fg[c] = side ? (a.s.us[c][n] - a.s.us[c][id]) * ih[ax] : (a.s.us[c][id] - a.s.us[c][n]) * ih[ax];
                const double(*Cn)[6] = stiffness(ctab, f.code[nb[fc]]);
                for (int o = 0; o < 3; o++) {
                    const int r = kFaceRows[ax][o];
                    double tq = Pa[ax][o][id], tn = Pa[ax][o][nb[fc]];
                    for (int c = 0; c < 3; c++) {
                        const int col = kFaceCols[ax][c];
                        tq = tq + Cq[r][col] * fg[c];
                        tn = tn + Cn[r][col] * fg[c];
                    }
                    sf[o] = (tq + tn) / 2.;
                }
            }
            for (int c = 0; c < 3; c++) d[ax][c] = s2[c] - s1[c];
        }
        const double* lc = ctab + (36 + kEigC) * f.nmat + ((code & kMatMask) - 1) * kElC;
        const double el_rho = lc[0], el_mass = lc[1], el_stiff = lc[2];
        for (int c = 0; c < 3; c++) {
            const double fo = d[0][c] * f.idx + d[1][c] * f.idy + d[2][c] * f.idz;
            // the stored force only matters with stiffness damping
            const double fs = el_stiff != 0. ? f.fstore[c][id] : 0.;
            if (stage == 2 || stage == 3)
                dv[c] = fo * el_rho + el_mass * f.v[c][id] + el_stiff * (fo - fs) * 2.;
            else
                dv[c] = fo * el_rho + el_mass * f.v[c][id] + el_stiff * (fo - fs);
            if (stage == 1) f.fstore[c][id] = fo;
        }
    }
    for (int c = 0; c < 3; c++) {
        const double du = f.vs[c][id] * f.dt;
        const double v = f.v[c][id], u = f.u[c][id];
        if (stage == 1) {
            f.vs[c][id] = v + dv[c] * 0.5;
            f.accv[c][id] = v + dv[c] * kSixth;
            f.us[c][id] = u + du * 0.5;
            f.accu[c][id] = u + du * kSixth;
        } else if (stage == 2) {
            f.vs[c][id] = v + dv[c] * 0.5;
            f.accv[c][id] = f.accv[c][id] + dv[c] * kThird;
            f.us[c][id] = u + du * 0.5;
            f.accu[c][id] = f.accu[c][id] + du * kThird;
        } else if (stage == 3) {
            f.vs[c][id] = v + dv[c];
            f.accv[c][id] = f.accv[c][id] + dv[c] * kThird;
            f.us[c][id] = u + du;
            f.accu[c][id] = f.accu[c][id] + du * kThird;
        } else {
            const double vn = f.accv[c][id] + dv[c] * kSixth;
            const double un = f.accu[c][id] + du * kSixth;
            f.v[c][id] = vn;
            f.vs[c][id] = vn;
            f.u[c][id] = un;
            f.us[c][id] = un;
        }
    }
}

// --------------------------------------------------------------- history
__global__ void __launch_bounds__(256) el_history_kernel(ElHistArgs a) {
    const int64_t nxy = a.nx * a.ny;
    const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= kHistPlanes * 3 * nxy) return;
    const int64_t ij = t % nxy;
    const int pc = static_cast<int>(t / nxy);   // plane * 3 + comp
    const int p = pc / 3, c = pc % 3;
    a.slot[t] = a.u[c][ij * a.nz + a.planes[p]];
}

// -------------------------------------------------------------- boundary
__global__ void __launch_bounds__(256) el_boundary_kernel(ElBoundaryArgs a) {
    const int64_t nxy = a.nx * a.ny, nz = a.nz;
    const int64_t ij = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (ij >= nxy) return;
    const int64_t col = ij * nz;
    auto hist = [&](int slot, int plane, int c) { return a.t[slot][(plane * 3 + c) * nxy + ij]; };
    if (a.liao) {
        // fourth-order Liao extrapolation from planes 1..4 of t1..t4
        const double w = a.w;
        for (int c = 0; c < 3; c++) {
            const double val = (4. * w + 2. * (1. - w)) * hist(0, 0, c) - (6. * w + (1. - w)) * hist(1, 1, c) +
                               (4. * w) * hist(2, 2, c) - w * hist(3, 3, c);
            a.u[c][col] = val;
            if (a.kind == 2) a.us[c][col] = val;
        }
    } else {
        if (a.mur_bottom) {
            for (int c = 0; c < 3; c++) {
                const double cc = c == 2 ? a.cb_l : a.cb_t;
                if (a.kind < 2) {
                    a.us[c][col] = a.u[c][col + 1] + cc * (a.u[c][col] - a.us[c][col + 1]);
                } else {
                    const double val = hist(0, 1, c) + cc * (hist(0, 0, c) - a.u[c][col + 1]);
                    a.u[c][col] = val;
                    a.us[c][col] = val;
                }
            }
        }
        for (int c = 0; c < 3; c++) {
            const double cc = c == 2 ? a.ct_l : a.ct_t;
            const int64_t top = col + nz - 1, below = col + nz - 2;
            if (a.kind < 2) {
                a.us[c][top] = a.u[c][below] + cc * (a.u[c][top] - a.us[c][below]);
            } else {
                const double val = hist(0, 2, c) + cc * (hist(0, 3, c) - a.u[c][below]);
                a.u[c][top] = val;
                a.us[c][top] = val;
            }
        }
    }
    if (a.src_k >= 0 && solid(a.code, col + a.src_k)) {
        const int64_t s = col + a.src_k;
        if (a.kind < 2) {
            a.us[a.src_comp][s] = a.src_val;
        } else {
            a.u[a.src_comp][s] = a.src_val;
            a.us[a.src_comp][s] = a.src_val;
        }
    }
}

unsigned blocks_for(int64_t n) { return static_cast<unsigned>((n + 255) / 256); }

}  // namespace

hipError_t launch_el_strain(const ElStrainArgs& a, hipStream_t s) {
    hipLaunchKernelGGL(el_strain_kernel, dim3(blocks_for(a.nx * a.ny * a.nz)), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_el_stage(const ElStageArgs& a, hipStream_t s) {
    const bool general = a.s.rotate || a.e.rotate || a.e.flexo;
    const void* prep = general ? reinterpret_cast<const void*>(el_prep_kernel<true>)
                               : reinterpret_cast<const void*>(el_prep_kernel<false>);
    const size_t lds = sizeof(double) * kLdsPerMat * static_cast<size_t>(a.f.nmat > 0 ? a.f.nmat : 1);
    if (lds > 65536) {   // up to 255 materials: 96 KiB of the CU's 160 KiB
        hipError_t e = hipFuncSetAttribute(prep, hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds));
        if (e == hipSuccess)
            e = hipFuncSetAttribute(reinterpret_cast<const void*>(el_force_kernel),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds));
        if (e != hipSuccess) return e;
    }
    const dim3 grid(blocks_for(a.f.nx * a.f.ny * a.f.nz));
    if (general) hipLaunchKernelGGL(el_prep_kernel<true>, grid, dim3(256), lds, s, a);
    else hipLaunchKernelGGL(el_prep_kernel<false>, grid, dim3(256), lds, s, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(el_force_kernel, grid, dim3(256), lds, s, a);
    return hipGetLastError();
}

hipError_t launch_el_history(const ElHistArgs& a, hipStream_t s) {
    hipLaunchKernelGGL(el_history_kernel, dim3(blocks_for(kHistPlanes * 3 * a.nx * a.ny)), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_el_boundary(const ElBoundaryArgs& a, hipStream_t s) {
    hipLaunchKernelGGL(el_boundary_kernel, dim3(blocks_for(a.nx * a.ny)), dim3(256), 0, s, a);
    return hipGetLastError();
}

}  // namespace gf
