// gf_em.hip — electrodynamics (Yee FDTD with RK4 in time) for gfx950, FP64.
//
// One stage of EMdynamic_system (EMdynamic_solver.cpp,
// EMdynamic_AbsorbingBoundary.cpp) as a handful of launches:
//   eh        Ampere per edge with the four-cell means of permittivity,
//             conductivity, Jf, Jp and dp (get_dE_RK*, :337-2237), fused
//             with the RK update (update_DE_RK*, :3588-3777); and in the
//             same launch Faraday per face with dM of the two cells
//             (get_dH_RK*, :2239-3475), fused with update_DH_RK*
//             (:3477-3586).  The free current of the box is evaluated where
//             an edge reads it (update_Jf_input*, :3885-4315).  At stage 4
//             each edge on a plane a Mur wall reads keeps its base value
//             before the combine (transfer_pointer, EMdynamic_Absorbing
//             Boundary.cpp:3-56).
//   jp        Drude/Lorentz currents (get_dJp_RK*, update_Jp*, :4318-4524)
//   plane     hard plane source (update_planeEM*, :3829-3883)
//   wall x3   first-order Mur or PEC walls, x then y then z (:58-874)
//   pec       PEC edges (update_DE_PEC, :4526-4556)
//   cells     cell-centred E and H of the new stores (:3780-3826)
// E and H keep two stage stores each: the eh launch reads the stage-s
// stores of both fields and writes stage s+1 into the other buffer.
// RK4 keeps a state accumulator, acc = y + k1/6, acc += k2/3, acc += k3/3,
// y = acc + k4/6, the reference's left-to-right sum.
#include <algorithm>

#include "gf_em_dev.h"
#include "gf_internal.h"

namespace gf {

namespace {

using em_dev::em_edge;
using em_dev::em_face;
using em_dev::rk_state;

// Ampere and Faraday of a stage in one launch: both read only the stage
// stores and write other arrays, so the edges (first nE threads) and the
// faces (the rest) are independent.
__global__ void __launch_bounds__(256) em_eh_kernel(EmArgs a) {
    const int64_t nx = a.nx, ny = a.ny, nz = a.nz;
    const int64_t nE = nx * (ny + 1) * (nz + 1) + (nx + 1) * ny * (nz + 1) + (nx + 1) * (ny + 1) * nz;
    const int64_t nH = (nx + 1) * ny * nz + nx * (ny + 1) * nz + nx * ny * (nz + 1);
    const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t < nE) em_edge(a, t);
    else if (t < nE + nH) em_face(a, t - nE);
}

__global__ void __launch_bounds__(256) em_jp_kernel(EmJpArgs a) {
    const int64_t id = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (id >= a.n) return;
    const uint32_t t = a.code[id] & kMatMask;
    for (int p = 0; p < 3; p++)
        for (int c = 0; c < 3; c++) {
            double k_ = 0.;
            if (t) {
                const DevMaterial& M = a.mats[t - 1];
                k_ = (kEps0 * M.wp[p] * M.wp[p] * a.ecell[c][id] * M.cn[p] - M.re[p] * a.jps[3 * p + c][id]) * a.dt;
            }
            // the store is updated in place: only this thread reads it now
            rk_state(a.stage, id, k_, a.jp[3 * p + c], a.jacc[3 * p + c], a.jps[3 * p + c]);
            if (a.stage == 4) a.jps[3 * p + c][id] = a.jp[3 * p + c][id];
        }
}

// The new value of the z-wall edge of line col (x or y component comp;
// both layouts end in nz + 1 along z) at the top (k = nz) or the bottom
// (k = 0), as em_wall_kernel writes it.
__device__ __forceinline__ double z_wall_value(const EmWallArgs& w, int comp, int64_t col, bool top) {
    const int64_t nz = w.nz;
    if (w.pec || (!top && w.onlytop)) return 0.;
    const int64_t wall = top ? col + nz : col, inner = top ? col + nz - 1 : col + 1;
    const double cf = top ? w.cf : w.cb;
    if (w.kind < 2) return w.b[comp][inner] + cf * (w.b[comp][wall] - w.s[comp][inner]);
    return w.t1[comp][inner] + cf * (w.t1[comp][wall] - w.b[comp][inner]);
}

__global__ void __launch_bounds__(256) em_cells_kernel(EmCellArgs a) {
    const int64_t nx = a.nx, ny = a.ny, nz = a.nz;
    const int64_t id = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (id >= nx * ny * nz) return;
    const int64_t i = id / (ny * nz), j = (id / nz) % ny, k = id % nz;
    if (a.do_e) {
        const double* ex = a.es[0];
        const double* ey = a.es[1];
        const double* ez = a.es[2];
        const int64_t x00 = (i * (ny + 1) + j) * (nz + 1) + k, x10 = (i * (ny + 1) + j + 1) * (nz + 1) + k;
        const int64_t y00 = (i * ny + j) * (nz + 1) + k, y10 = ((i + 1) * ny + j) * (nz + 1) + k;
        double vx[4] = {ex[x00], ex[x10], ex[x00 + 1], ex[x10 + 1]};
        double vy[4] = {ey[y00], ey[y10], ey[y00 + 1], ey[y10 + 1]};
        if (a.zwall && (k == 0 || k == nz - 1)) {
            // the wall edges this cell reads, from the wall condition; the
            // cell of a line's own (i, j) writes it (and the last cell the
            // line past the grid)
            double* wx = a.w.kind < 2 ? a.w.s[0] : a.w.b[0];
            double* wy = a.w.kind < 2 ? a.w.s[1] : a.w.b[1];
            for (int top = 0; top < 2; top++) {
                if (top ? k != nz - 1 : k != 0) continue;
                const int slot = top ? 2 : 0;
                const int64_t cx0 = x00 - k, cx1 = x10 - k, cy0 = y00 - k, cy1 = y10 - k;   // line starts
                vx[slot] = z_wall_value(a.w, 0, cx0, top);
                vx[slot + 1] = z_wall_value(a.w, 0, cx1, top);
                vy[slot] = z_wall_value(a.w, 1, cy0, top);
                vy[slot + 1] = z_wall_value(a.w, 1, cy1, top);
                const int64_t off = top ? nz : 0;
                wx[cx0 + off] = vx[slot];
                if (j == ny - 1) wx[cx1 + off] = vx[slot + 1];
                wy[cy0 + off] = vy[slot];
                if (i == nx - 1) wy[cy1 + off] = vy[slot + 1];
            }
        }
        a.ecell[0][id] = (vx[0] + vx[1] + vx[2] + vx[3]) / 4.;
        a.ecell[1][id] = (vy[0] + vy[1] + vy[2] + vy[3]) / 4.;
        const int64_t z00 = (i * (ny + 1) + j) * nz + k, z10 = ((i + 1) * (ny + 1) + j) * nz + k;
        const int64_t z01 = (i * (ny + 1) + j + 1) * nz + k, z11 = ((i + 1) * (ny + 1) + j + 1) * nz + k;
        a.ecell[2][id] = (ez[z00] + ez[z10] + ez[z01] + ez[z11]) / 4.;
    }
    if (a.do_h) {
        a.hcell[0][id] = (a.hs[0][(i * ny + j) * nz + k] + a.hs[0][((i + 1) * ny + j) * nz + k]) / 2.;
        a.hcell[1][id] = (a.hs[1][(i * (ny + 1) + j) * nz + k] + a.hs[1][(i * (ny + 1) + j + 1) * nz + k]) / 2.;
        if (a.one_d) {
            // 1D magnetostatics through the cell field (EMdynamic_solver.cpp:3793-3804)
            const uint32_t t = a.code[id] & kMatMask;
            double h = 0.;
            if (t) {
                const DevMaterial& M = a.mats[t - 1];
                const double mz = a.mz ? a.mz[id] : 0., mz1 = a.mz1 ? a.mz1[id] : 0., mz2 = a.mz2 ? a.mz2[id] : 0.;
                h = -1. * M.Ms * mz - 1. * M.aMs[0] * mz1 - 1. * M.aMs[1] * mz2 - a.hzstat[id];
            }
            a.hcell[2][id] = h;
        } else {
            a.hcell[2][id] = (a.hs[2][(i * ny + j) * (nz + 1) + k] + a.hs[2][(i * ny + j) * (nz + 1) + k + 1]) / 2.;
        }
    }
}

// One wall pair of the first-order Mur condition for one line of edges:
// kind < 2 writes the out store from the base and the new store, kind 2
// writes the base from the history (EMdynamic_AbsorbingBoundary.cpp:58-874).
__device__ __forceinline__ void mur(const EmWallArgs& a, int comp, int64_t wall, int64_t inner, double cf) {
    if (a.kind < 2) a.s[comp][wall] = a.b[comp][inner] + cf * (a.b[comp][wall] - a.s[comp][inner]);
    else a.b[comp][wall] = a.t1[comp][inner] + cf * (a.t1[comp][wall] - a.b[comp][inner]);
}

__device__ __forceinline__ void zero(const EmWallArgs& a, int comp, int64_t at) {
    if (a.kind < 2) a.s[comp][at] = 0.;
    else a.b[comp][at] = 0.;
}

__global__ void __launch_bounds__(256) em_wall_kernel(EmWallArgs a) {
    const int64_t nx = a.nx, ny = a.ny, nz = a.nz;
    const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (a.axis == 0) {
        // Ey lines (j < ny, k <= nz) and Ez lines (j <= ny, k < nz)
        if (t < ny * (nz + 1)) {
            const int64_t j = t / (nz + 1), k = t % (nz + 1);
            auto EY = [&](int64_t i) { return (i * ny + j) * (nz + 1) + k; };
            if (!a.pec) { mur(a, 1, EY(0), EY(1), a.cf); mur(a, 1, EY(nx), EY(nx - 1), a.cf); }
            else { zero(a, 1, EY(0)); zero(a, 1, EY(nx)); }
        }
        if (t < (ny + 1) * nz) {
            const int64_t j = t / nz, k = t % nz;
            auto EZ = [&](int64_t i) { return (i * (ny + 1) + j) * nz + k; };
            if (!a.pec) { mur(a, 2, EZ(0), EZ(1), a.cf); mur(a, 2, EZ(nx), EZ(nx - 1), a.cf); }
            else { zero(a, 2, EZ(0)); zero(a, 2, EZ(nx)); }
        }
    } else if (a.axis == 1) {
        // Ex lines (i < nx, k <= nz) and Ez lines (i <= nx, k < nz)
        if (t < nx * (nz + 1)) {
            const int64_t i = t / (nz + 1), k = t % (nz + 1);
            auto EX = [&](int64_t j) { return (i * (ny + 1) + j) * (nz + 1) + k; };
            if (!a.pec) { mur(a, 0, EX(0), EX(1), a.cf); mur(a, 0, EX(ny), EX(ny - 1), a.cf); }
            else { zero(a, 0, EX(0)); zero(a, 0, EX(ny)); }
        }
        if (t < (nx + 1) * nz) {
            const int64_t i = t / nz, k = t % nz;
            auto EZ = [&](int64_t j) { return (i * (ny + 1) + j) * nz + k; };
            if (!a.pec) { mur(a, 2, EZ(0), EZ(1), a.cf); mur(a, 2, EZ(ny), EZ(ny - 1), a.cf); }
            else { zero(a, 2, EZ(0)); zero(a, 2, EZ(ny)); }
        }
    } else {
        // Ex lines (i < nx, j <= ny) and Ey lines (i <= nx, j < ny); top first
        for (int comp = 0; comp < 2; comp++) {
            const int64_t lines = comp == 0 ? nx * (ny + 1) : (nx + 1) * ny;
            if (t >= lines) continue;
            const int64_t col = t * (nz + 1);   // both layouts end in (nz + 1) along z
            if (!a.pec) {
                mur(a, comp, col + nz, col + nz - 1, a.cf);
                if (a.onlytop) zero(a, comp, col);
                else mur(a, comp, col, col + 1, a.cb);
            } else {
                zero(a, comp, col);
                zero(a, comp, col + nz);
            }
        }
    }
}

__global__ void __launch_bounds__(256) em_plane_kernel(EmMiscArgs a) {
    const int64_t nx = a.nx, ny = a.ny, nz = a.nz;
    const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= nx * ny) return;
    const int64_t i = t / ny, j = t % ny, z = a.plane_z;
    double* const* d = a.kind < 2 ? a.s : a.b;
    d[0][(i * (ny + 1) + j) * (nz + 1) + z] = a.plane[0];
    d[1][(i * ny + j) * (nz + 1) + z] = a.plane[1];
    d[2][(i * (ny + 1) + j) * nz + z] = a.plane[2];
}

__global__ void __launch_bounds__(256) em_pec_kernel(EmMiscArgs a) {
    const int64_t nx = a.nx, ny = a.ny, nz = a.nz;
    const int64_t nEx = nx * (ny + 1) * (nz + 1), nEy = (nx + 1) * ny * (nz + 1), nEz = (nx + 1) * (ny + 1) * nz;
    int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= nEx + nEy + nEz) return;
    int comp = 0;
    if (e >= nEx + nEy) { comp = 2; e -= nEx + nEy; }
    else if (e >= nEx) { comp = 1; e -= nEx; }
    if (a.pec[comp][e]) {
        a.b[comp][e] = 0.;
        a.s[comp][e] = 0.;
    }
}

unsigned blocks_for(int64_t n) { return static_cast<unsigned>((n + 255) / 256); }

}  // namespace

hipError_t launch_em_eh(const EmArgs& a, hipStream_t s) {
    const int64_t nE = a.nx * (a.ny + 1) * (a.nz + 1) + (a.nx + 1) * a.ny * (a.nz + 1) + (a.nx + 1) * (a.ny + 1) * a.nz;
    const int64_t nH = (a.nx + 1) * a.ny * a.nz + a.nx * (a.ny + 1) * a.nz + a.nx * a.ny * (a.nz + 1);
    hipLaunchKernelGGL(em_eh_kernel, dim3(blocks_for(nE + nH)), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_em_jp(const EmJpArgs& a, hipStream_t s) {
    hipLaunchKernelGGL(em_jp_kernel, dim3(blocks_for(a.n)), dim3(256), 0, s, a);
    return hipGetLastError();
}

// One thread per cell; cells of planes without a source keep their zero.
// The sums run FM, AFM1, AFM2 as the reference writes them.
__global__ void __launch_bounds__(256) spin_pump_kernel(SpinArgs a) {
    const int64_t n = a.nx * a.ny * a.nz;
    const int64_t id = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (id >= n) return;
    const int64_t k = id % a.nz;
    const int32_t sz = a.src[k];
    if (sz < 0) return;
    const int64_t is = id - k + sz;
    // the increment: the stage's k, or at the step boundary acc + k4/6 (the
    // RK combination in the order the stage-4 update forms it)
    auto inc = [&](const double* const* d, const double* const* acc, int c) {
        return acc[0] ? acc[c][is] + d[c][is] * kSixth : d[c][is];
    };
    double sx = 0., sy = 0.;
    if (a.fs[0]) {
        const double d0 = inc(a.fd, a.fa, 0), d1 = inc(a.fd, a.fa, 1), d2 = inc(a.fd, a.fa, 2);
        sx = a.fs[0][is] * d2 - a.fs[2][is] * d0;
        sy = a.fs[1][is] * d2 - a.fs[2][is] * d1;
    }
    if (a.s1[0]) {
        const double p0 = inc(a.d1, a.a1, 0), p1 = inc(a.d1, a.a1, 1), p2 = inc(a.d1, a.a1, 2);
        const double q0 = inc(a.d2, a.a2, 0), q1 = inc(a.d2, a.a2, 1), q2 = inc(a.d2, a.a2, 2);
        sx = sx + a.s1[0][is] * p2 - a.s1[2][is] * p0 + a.s2[0][is] * q2 - a.s2[2][is] * q0;
        sy = sy + a.s1[1][is] * p2 - a.s1[2][is] * p1 + a.s2[1][is] * q2 - a.s2[2][is] * q1;
    }
    a.j[0][id] = a.coef[k] * sx / a.dt;
    a.j[1][id] = a.coef[k] * sy / a.dt;
}

hipError_t launch_spin_pump(const SpinArgs& a, hipStream_t s) {
    hipLaunchKernelGGL(spin_pump_kernel, dim3(blocks_for(a.nx * a.ny * a.nz)), dim3(256), 0, s, a);
    return hipGetLastError();
}

__global__ void __launch_bounds__(256) spin_save_kernel(SpinSaveArgs a) {
    const int64_t n = a.nx * a.ny * a.nz;
    const int64_t id = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (id >= n || !a.is_src[id % a.nz]) return;
    for (int c = 0; c < a.ncomp; c++) a.dst[c][id] = a.src[c][id];
}

hipError_t launch_spin_save(const SpinSaveArgs& a, hipStream_t s) {
    hipLaunchKernelGGL(spin_save_kernel, dim3(blocks_for(a.nx * a.ny * a.nz)), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_em_cells(const EmCellArgs& a, hipStream_t s) {
    hipLaunchKernelGGL(em_cells_kernel, dim3(blocks_for(a.nx * a.ny * a.nz)), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_em_wall(const EmWallArgs& a, hipStream_t s) {
    int64_t n;
    if (a.axis == 0) n = std::max(a.ny * (a.nz + 1), (a.ny + 1) * a.nz);
    else if (a.axis == 1) n = std::max(a.nx * (a.nz + 1), (a.nx + 1) * a.nz);
    else n = std::max(a.nx * (a.ny + 1), (a.nx + 1) * a.ny);
    hipLaunchKernelGGL(em_wall_kernel, dim3(blocks_for(n)), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_em_plane(const EmMiscArgs& a, hipStream_t s) {
    hipLaunchKernelGGL(em_plane_kernel, dim3(blocks_for(a.nx * a.ny)), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_em_pec(const EmMiscArgs& a, hipStream_t s) {
    const int64_t n = a.nx * (a.ny + 1) * (a.nz + 1) + (a.nx + 1) * a.ny * (a.nz + 1) + (a.nx + 1) * (a.ny + 1) * a.nz;
    hipLaunchKernelGGL(em_pec_kernel, dim3(blocks_for(n)), dim3(256), 0, s, a);
    return hipGetLastError();
}


}  // namespace gf
