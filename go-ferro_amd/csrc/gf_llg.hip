// gf_llg.hip — fused LLG RK4 stage kernels for gfx950 (FP64).
//
// One launch per RK stage does what the reference spreads over get_dm_RKs,
// get_dm (at stage 4) and update_m_RKs / update_m
// (magnetic_solver.cpp:3-229, 1449-1457, 1515-1636;
//  magnetic_AFM_solver.cpp:3-430, 1715-1834):
//   stage 1:  k = f(m);      next = m + k/2;  acc = k/6
//   stage 2:  k = f(s2);     next = m + k/2;  acc = acc + k/3
//   stage 3:  k = f(s3);     next = m + k;    acc = acc + k/3
//   stage 4:  k = f(s4);     m = m + (acc + k/6)
// The accumulator is summed in the reference's order
// (((k1/6 + k2/3) + k3/3) + k4/6), the weights applied as products with the
// rounded reciprocals (within about an ulp of the reference's quotients);
// the four k-arrays and the separate combine/update/copy launches disappear.  The
// stencil store is read-only within a stage and the next store goes to the
// other buffer, so physics that share a stage still see the old store.
#include "gf_llg_dev.h"

namespace gf {

namespace {

// Material records the lean fused stage copies into LDS (about 0.9 KB each),
// and the set-up loads per thread of a 256-thread workgroup: up to
// kMatWordsPerThread words of those records, and the field spectrum in
// chunks of kSetupChunk complex values.
constexpr int kLeanMats = 8;
constexpr int kMatWordsPerThread = 4;
constexpr int kSetupChunk = 4;
static_assert(kLeanMats * sizeof(DevMaterial) / sizeof(double) <= kMatWordsPerThread * 256,
              "the lean stage's material records must fit the set-up loads of 256 threads");

// gather6 for launches whose active lanes hold consecutive cells along z
// (lane i of a wave at cell id0 + i, and a lane without a cell only past
// the grid's last cell): a cell's z neighbours are then the store values
// and codes of the lanes next to it, moved with DPP wave shifts, except at
// a column end (k = 0 or nz - 1, where the neighbour wraps or the face is a
// surface) or a wave edge, where they are loaded.  A lane next to one
// without a cell is at a column end.  Two parts: the loads, and after the
// cell's other loads have been issued, the lane shifts of its own store c
// and code own.
template <class I>
__device__ __forceinline__ void gather6_lanes_load(const P3 s3, const uint32_t* code, I x, I y, I z, I nx, I ny,
                                                   I nz, double (&v)[6][3], uint32_t (&cd)[6]) {
    const I nb[4] = {((x == 0 ? nx - 1 : x - 1) * ny + y) * nz + z, ((x == nx - 1 ? 0 : x + 1) * ny + y) * nz + z,
                     (x * ny + (y == 0 ? ny - 1 : y - 1)) * nz + z, (x * ny + (y == ny - 1 ? 0 : y + 1)) * nz + z};
#pragma unroll
    for (int f = 0; f < 4; f++) {
        for (int q = 0; q < 3; q++) v[f][q] = ldd(s3.p[q], nb[f]);
        cd[f] = ldu(code, nb[f]);
    }
    const int lane = static_cast<int>(threadIdx.x & 63u);
    const bool gm = lane == 0 || z == 0, gp = lane == 63 || z == nz - 1;
    if (gm) {
        const I zm = (x * ny + y) * nz + (z == 0 ? nz - 1 : z - 1);
        for (int q = 0; q < 3; q++) v[4][q] = ldd(s3.p[q], zm);
        cd[4] = ldu(code, zm);
    }
    if (gp) {
        const I zp = (x * ny + y) * nz + (z == nz - 1 ? 0 : z + 1);
        for (int q = 0; q < 3; q++) v[5][q] = ldd(s3.p[q], zp);
        cd[5] = ldu(code, zp);
    }
}

template <class I>
__device__ __forceinline__ void gather6_lanes_shift(I z, I nz, const double (&c)[3], uint32_t own, double (&v)[6][3],
                                                    uint32_t (&cd)[6]) {
    const int lane = static_cast<int>(threadIdx.x & 63u);
    const bool gm = lane == 0 || z == 0, gp = lane == 63 || z == nz - 1;
    for (int q = 0; q < 3; q++) {
        const double dn = lane_up(c[q]), up = lane_down(c[q]);
        if (!gm) v[4][q] = dn;
        if (!gp) v[5][q] = up;
    }
    const uint32_t cdn = lane_up(own), cup = lane_down(own);
    if (!gm) cd[4] = cdn;
    if (!gp) cd[5] = cup;
}

// Every global input of one FM cell of a stage except the static field:
// loaded together, before the cell's own code is known, so that one
// memory round trip feeds the cell (all indices are in bounds whatever the
// code says).
struct FmIn {
    uint32_t own;
    double c[3];        // stage store at the cell
    double v[6][3];     // stage store at the six face neighbours (gather6)
    uint32_t cd[6];     // their codes
    double m[3];        // m at the start of the step
    double acc[3];      // RK accumulator (read at stages 2-4)
};

// kLanes: the launch's lanes hold consecutive cells (see gather6_lanes).
template <class I, bool kLanes = false>
__device__ __forceinline__ void fm_load(const StageArgs& a, I id, I x, I y, I z, FmIn& in) {
    in.own = ldu(a.code, id);
    const P3 sp = {{a.store[0], a.store[1], a.store[2]}};
    for (int c = 0; c < 3; c++) in.c[c] = ldd(a.store[c], id);
    const I nx = static_cast<I>(a.nx), ny = static_cast<I>(a.ny), nz = static_cast<I>(a.nz);
    if (kLanes) gather6_lanes_load<I>(sp, a.code, x, y, z, nx, ny, nz, in.v, in.cd);
    else gather6<I>(sp, a.code, x, y, z, nx, ny, nz, in.v, in.cd);
    // at stage 1 the stage store is m itself: read it once
    const bool same = a.base[0] == a.store[0] && a.base[1] == a.store[1] && a.base[2] == a.store[2];
    for (int c = 0; c < 3; c++) in.m[c] = same ? in.c[c] : ldd_nt(a.base[c], id);
    for (int c = 0; c < 3; c++) in.acc[c] = a.stage > 1 ? ldd_nt(a.acc[c], id) : 0.;
    if (kLanes) gather6_lanes_shift<I>(z, nz, in.c, in.own, in.v, in.cd);
}

// RK update of one component (see the file comment) from its preloaded m
// and accumulator; returns the value written to the next store (m itself
// at stage 4): the RK value, or pv where the cell's value is prescribed.
template <class I>
__device__ __forceinline__ double rk_store(int stage, I id, int c, double k, double m, double acc,
                                           double* const* accp, double* const* out, bool prescribed = false,
                                           double pv = 0.) {
    double v;
    if (stage == 1) {
        v = m + k * 0.5;
        std_nt(accp[c], id, k * kSixth);
    } else if (stage == 2) {
        v = m + k * 0.5;
        std_nt(accp[c], id, acc + k * kThird);
    } else if (stage == 3) {
        v = m + k;
        std_nt(accp[c], id, acc + k * kThird);
    } else {
        v = m + (acc + k * kSixth);
    }
    if (prescribed) v = pv;
    std_nt(out[c], id, v);
    return v;
}

// One FM cell of a stage: H_eff with the static field hs, the LLG
// increment and the RK update; n receives the next store value.  kLean
// compiles out the frame rotation and the optional fields (strain, non-
// uniform and shaped drives, electrodynamic back-action, k output, the
// prescribed bottom layer) for contexts that have none of them: their
// pointers and the rotation matrices then occupy no scalar registers
// across the cell loop.  mats is the material table, in global memory
// (GMat) or a workgroup's LDS copy (LMat).  bottom: the cell is at k = 0.
template <bool kLean, class I, class MT>
__device__ __forceinline__ void fm_cell(const StageArgs& a, MT* mats, const FmIn& in, I id, double hs0, double hs1,
                                        double hs2, double& n0, double& n1, double& n2, bool bottom) {
    MT& M = mats[(in.own & kMatMask) - 1];

    const double mgx = in.c[0], mgy = in.c[1], mgz = in.c[2];
    double mx = mgx, my = mgy, mz = mgz;
    if (!kLean && a.rotate) rot(a.Tg2c, mgx, mgy, mgz, mx, my, mz);

    const Stencil st = stencil_core<true>(mgx, mgy, mgz, in.v, in.cd, mats, in.own, kFmSurfShift, -1, a.dx, a.dy,
                                          a.dz, M.Aex, M.Aex_self, M.dmi_slope);

    // magnetocrystalline anisotropy (magnetic_solver.cpp:124-141)
    double hac0 = 0., hac1 = 0., hac2 = 0.;
    if (M.anis_type == 1) {
        hac0 = -2. * (M.K1n * (my * my + mz * mz) + M.K2n * my * my * mz * mz) * mx;
        hac1 = -2. * (M.K1n * (mx * mx + mz * mz) + M.K2n * mx * mx * mz * mz) * my;
        hac2 = -2. * (M.K1n * (mx * mx + my * my) + M.K2n * mx * mx * my * my) * mz;
    } else if (M.anis_type == 2) {
        hac2 = 2. * M.K1n * mz;
    }
    double ha0 = hac0, ha1 = hac1, ha2 = hac2;
    if (!kLean && a.rotate) rot(a.Tc2g, hac0, hac1, hac2, ha0, ha1, ha2);

    // H_eff (magnetic_solver.cpp:164-198); the magnetoelastic term
    // (:89-96, 143-150) enters only when a strain is coupled, since it is
    // zero otherwise.
    double H0 = hs0 + ha0;
    double H1 = hs1 + ha1;
    double H2 = hs2 + ha2;
    if (!kLean && a.strain[0]) {
        double he0, he1, he2;
        magnetoelastic(a.strain, a.strain0, id, -2., M.B1n, M.B2n, mx, my, mz, a.rotate, a.Tc2g, he0, he1, he2);
        H0 = H0 + he0;
        H1 = H1 + he1;
        H2 = H2 + he2;
    }
    H0 = H0 + M.exch * st.ex[0] + M.dmi_pos * st.d[2] + a.hext[0];
    H1 = H1 + M.exch * st.ex[1] + M.dmi_pos * st.d[3] + a.hext[1];
    H2 = H2 + M.exch * st.ex[2] + M.dmi_neg * (st.d[0] + st.d[1]) + a.hext[2];
    if (!kLean && a.hnu[0]) {
        H0 = H0 + ldd(a.hnu[0], id);
        H1 = H1 + ldd(a.hnu[1], id);
        H2 = H2 + ldd(a.hnu[2], id);
    }
    if (!kLean && a.hdrive[0]) {
        H0 = H0 + ldd(a.hdrive[0], id) * a.drive_s;
        H1 = H1 + ldd(a.hdrive[1], id) * a.drive_s;
        H2 = H2 + ldd(a.hdrive[2], id) * a.drive_s;
    }
    if (!kLean && a.hcell[0]) {
        // electrodynamic back-action (magnetic_solver.cpp:200-204)
        H0 = H0 + ldd(a.hcell[0], id);
        H1 = H1 + ldd(a.hcell[1], id);
        H2 = H2 + ldd(a.hcell[2], id);
    }

    // LLG right-hand side times dt (magnetic_solver.cpp:206-220)
    const double px = mgy * H2 - mgz * H1, py = mgz * H0 - mgx * H2, pz = mgx * H1 - mgy * H0;
    const double qx = mgy * pz - mgz * py, qy = mgz * px - mgx * pz, qz = mgx * py - mgy * px;
    const double k0 = M.precess * px + M.alpha_precess * qx;
    const double k1 = M.precess * py + M.alpha_precess * qy;
    const double k2 = M.precess * pz + M.alpha_precess * qz;
    if (!kLean && a.dm_out[0]) {
        std_(a.dm_out[0], id, k0); std_(a.dm_out[1], id, k1); std_(a.dm_out[2], id, k2);
    }
    // update_prescribed_m{_half,_full,} over the update (magnetic_solver.cpp:
    // 1526-1609, 1891-1957); the accumulator keeps the LLG increment
    const bool pres = !kLean && a.prescribe && bottom;
    n0 = rk_store(a.stage, id, 0, k0, in.m[0], in.acc[0], a.acc, a.out, pres, a.presc_m[0]);
    n1 = rk_store(a.stage, id, 1, k1, in.m[1], in.acc[1], a.acc, a.out, pres, a.presc_m[1]);
    n2 = rk_store(a.stage, id, 2, k2, in.m[2], in.acc[2], a.acc, a.out, pres, a.presc_m[2]);
}

// One FM cell of the unfused stage: coordinates from the linear index,
// magnetostatic field from Hstat.
template <class I, bool kLanes = false>
__device__ __forceinline__ void fm_point(const StageArgs& a, I id) {
    const I ny = static_cast<I>(a.ny), nz = static_cast<I>(a.nz);
    const I col = id / nz, z = id - col * nz;
    const I x = col / ny, y = col - x * ny;
    FmIn in;
    fm_load<I, kLanes>(a, id, x, y, z, in);
    // a null Hstat is zero (no statics, nothing uploaded)
    const bool hs = a.hstat[0] != nullptr;
    const double hs0 = hs ? ldd(a.hstat[0], id) : 0., hs1 = hs ? ldd(a.hstat[1], id) : 0.;
    const double hs2 = hs ? ldd(a.hstat[2], id) : 0.;
    if (!(in.own & kIsFM)) return;
    double n0, n1, n2;
    fm_cell<false, I>(a, (GMat*)a.mats, in, id, hs0, hs1, hs2, n0, n1, n2, z == 0);
}

template <class I>
__global__ void __launch_bounds__(256) mag_stage_kernel(StageArgs a) {
    const I n = static_cast<I>(a.nx * a.ny * a.nz);
    const I id = static_cast<I>(blockIdx.x) * static_cast<I>(blockDim.x) + static_cast<I>(threadIdx.x);
    if (id >= n) return;
    fm_point<I, true>(a, id);
}

// The same over a list of the FM cells: a layered magnet fills whole
// waves instead of loading the inputs of every non-magnetic cell first.
__global__ void __launch_bounds__(256) mag_stage_list_kernel(StageArgs a, const uint32_t* ids, uint32_t count) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    fm_point<uint32_t>(a, ldu(ids, t));
}

// The FM stage fused with the z transforms of the FFT magnetostatics, for
// contexts whose magnetisation is FM only.  A workgroup owns P whole z
// columns (fixed x, consecutive y): it rebuilds their static field from
// the H z-spectrum (inverse real transform in LDS, two real columns per
// complex transform, the imaginary parts of the self-conjugate bins
// ignored as a complex-to-real transform does), runs the stage on every
// cell, and forward-transforms Ms m of the new store into the M z-spectrum.
// The three components are transformed together.  Column p < P/2 lives in
// the real parts and column p + P/2 in the imaginary parts of transform p;
// the inverse leaves sample k at bit-reversed position, which is where the
// forward transform wants its input, so each cell's H is read and its M
// written in the same LDS slot.  Both spectra use the layout [x][kz][y] of
// the separate z passes (gf_fft.hip), so the y and x passes are unchanged.
// Workgroups are ordered so that each XCD takes a contiguous range of x
// planes, keeping stencil neighbours in its L2.  The lean form also copies
// the material records into LDS (at most kLeanMats of them).
template <bool kHspec, bool kLean, class I>
__global__ void __launch_bounds__(256) mag_stage_fused_kernel(StageArgs a, FusedZArgs f) {
    extern __shared__ double2 fsm[];
    const int nz = static_cast<int>(a.nz), ny = static_cast<int>(a.ny);
    const int P = f.P, lognz = f.lognz, nzh = f.nzh;
    const int pairs = P > 1 ? P >> 1 : 1;
    double2* tw = fsm;
    double2* cbuf = tw + (nz >> 1);          // [3][pairs][nz], element i at slot fpad(i)
    DevMaterial* ltab = reinterpret_cast<DevMaterial*>(cbuf + fft_lds_elems(3 * pairs * nz));
    const int tid = threadIdx.x, bd = blockDim.x;
    const unsigned b = blockIdx.x, G = gridDim.x;
    const unsigned q = G / 8u, r = G % 8u, xcd = b % 8u, slot = b / 8u;
    const unsigned L = xcd * q + (xcd < r ? xcd : r) + slot;
    const int gpx = ny / P;
    const int x = static_cast<int>(L) / gpx, y0 = (static_cast<int>(L) % gpx) * P;
    const int cells = P * nz;
    // The set-up loads (twiddles, material records, the field spectrum in
    // chunks of kSetupChunk elements per thread) are issued before any of
    // them is written to LDS, so that together they cost one memory round
    // trip per chunk rather than one each.
    const int ntw = nz >> 1;
    const int words = kLean ? a.nmat * static_cast<int>(sizeof(DevMaterial) / sizeof(double)) : 0;
    const double2 twv = tid < ntw ? f.tw[tid] : make_double2(0., 0.);
    double mw[kMatWordsPerThread];
#pragma unroll
    for (int u = 0; u < kMatWordsPerThread; u++) {
        const int t = tid + u * bd;
        mw[u] = t < words ? ldd(reinterpret_cast<const double*>(a.mats), static_cast<uint32_t>(t)) : 0.;
    }
    bool setup_stored = false;
    auto store_setup = [&]() {
        if (tid < ntw) tw[tid] = twv;
        for (int j = tid + bd; j < ntw; j += bd) tw[j] = f.tw[j];
#pragma unroll
        for (int u = 0; u < kMatWordsPerThread; u++) {
            const int t = tid + u * bd;
            if (t < words) reinterpret_cast<double*>(ltab)[t] = mw[u];
        }
        setup_stored = true;
    };
    if (kHspec) {
        const int per = pairs * nzh, tot = 3 * per;
        for (int t0 = 0; t0 < tot; t0 += kSetupChunk * bd) {
            double2 A[kSetupChunk], B[kSetupChunk];
#pragma unroll
            for (int u = 0; u < kSetupChunk; u++) {
                const int t = t0 + u * bd + tid;
                A[u] = make_double2(0., 0.);
                B[u] = make_double2(0., 0.);
                if (t < tot) {
                    const int c = t / per, w = t - c * per;
                    const int p = w % pairs, kz = w / pairs;
                    const int64_t row = (static_cast<int64_t>(x) * nzh + kz) * ny + y0;
                    const double2* in = f.hspec + c * f.nh;
                    A[u] = in[row + p];
                    if (P > 1) B[u] = in[row + p + pairs];
                }
            }
            if (!setup_stored) store_setup();
#pragma unroll
            for (int u = 0; u < kSetupChunk; u++) {
                const int t = t0 + u * bd + tid;
                if (t < tot) {
                    const int c = t / per, w = t - c * per;
                    const int p = w % pairs, kz = w / pairs;
                    double2 Av = A[u], Bv = B[u];
                    if (kz == 0 || 2 * kz == nz) { Av.y = 0.; Bv.y = 0.; }
                    const int cb = (c * pairs + p) * nz;
                    cbuf[fpad(cb + kz)] = make_double2(Av.x - Bv.y, Av.y + Bv.x);
                    if (kz > 0 && 2 * kz < nz) cbuf[fpad(cb + nz - kz)] = make_double2(Av.x + Bv.y, Bv.x - Av.y);
                }
            }
        }
        __syncthreads();
        ifft_dif(cbuf, 3 * pairs, lognz, tw);
    } else {
        store_setup();
        for (int t = threadIdx.x; t < 3 * cells; t += blockDim.x) {
            const int c = t / cells, w = t - c * cells;
            const int col = w >> lognz, k = w & (nz - 1);
            const double h = a.hstat[c][(static_cast<int64_t>(x) * ny + y0 + col) * nz + k];
            const int p = col < pairs ? col : col - pairs;
            double* slot2 = reinterpret_cast<double*>(cbuf + fpad((c * pairs + p) * nz + fft_bitrev(k, lognz)));
            slot2[col < pairs ? 0 : 1] = h;
        }
        __syncthreads();
    }
    for (int w = threadIdx.x; w < cells; w += blockDim.x) {
        const int col = w >> lognz, k = w & (nz - 1);
        const int y = y0 + col;
        const int p = col < pairs ? col : col - pairs;
        const int part = col < pairs ? 0 : 1;
        const int pos = p * nz + fft_bitrev(k, lognz);
        double* h0 = reinterpret_cast<double*>(cbuf + fpad(pos)) + part;
        double* h1 = reinterpret_cast<double*>(cbuf + fpad(pairs * nz + pos)) + part;
        double* h2 = reinterpret_cast<double*>(cbuf + fpad(2 * pairs * nz + pos)) + part;
        const I id = (static_cast<I>(x) * static_cast<I>(ny) + static_cast<I>(y)) * static_cast<I>(nz) + static_cast<I>(k);
        FmIn in;
        fm_load<I, true>(a, id, static_cast<I>(x), static_cast<I>(y), static_cast<I>(k), in);
        double mx = 0., my = 0., mz = 0.;
        if (in.own & kIsFM) {
            double n0, n1, n2;
            double Ms;
            if (kLean) {
                LMat* tab = (LMat*)ltab;
                fm_cell<kLean, I>(a, tab, in, id, *h0, *h1, *h2, n0, n1, n2, k == 0);
                Ms = tab[(in.own & kMatMask) - 1].Ms;
            } else {
                fm_cell<kLean, I>(a, (GMat*)a.mats, in, id, *h0, *h1, *h2, n0, n1, n2, k == 0);
                Ms = mat_of(a.mats, in.own).Ms;
            }
            mx = Ms * n0; my = Ms * n1; mz = Ms * n2;
        }
        *h0 = mx; *h1 = my; *h2 = mz;
        if (P == 1) {
            // a lone column: the imaginary parts of the forward input are zero
            h0[1] = 0.; h1[1] = 0.; h2[1] = 0.;
        }
    }
    __syncthreads();
    fft_dit(cbuf, 3 * pairs, lognz, tw);
    const int per = pairs * nzh;
    for (int t = threadIdx.x; t < 3 * per; t += blockDim.x) {
        const int c = t / per, u = t - c * per;
        const int p = u % pairs, kz = u / pairs;
        const int cb = (c * pairs + p) * nz;
        const double2 Z = cbuf[fpad(cb + kz)];
        const double2 Zm = cbuf[fpad(cb + ((nz - kz) & (nz - 1)))];
        const int64_t row = (static_cast<int64_t>(x) * nzh + kz) * ny + y0;
        double2* out = f.mspec + c * f.nh;
        if (P > 1) {
            out[row + p] = make_double2((Z.x + Zm.x) * 0.5, (Z.y - Zm.y) * 0.5);
            out[row + p + pairs] = make_double2((Z.y + Zm.y) * 0.5, (Zm.x - Z.x) * 0.5);
        } else {
            out[row + p] = Z;
        }
    }
}

// One sublattice of one AFM cell: the two sublattices of a stage read only
// stage-(s-1) stores, so they run as independent threads.  Every global
// input of the sublattice is loaded before the cell's code is known (all
// indices are in bounds), as for the FM cell, so one memory round trip
// feeds the cell.
template <class I, bool kLanes = false>
__device__ __forceinline__ void afm_point(const AfmStageArgs& a, I id, int sub) {
    const I ny = static_cast<I>(a.ny), nz = static_cast<I>(a.nz);
    const I col = id / nz, z = id - col * nz;
    const I x = col / ny, y = col - x * ny;
    const P3 s1 = {{a.store1[0], a.store1[1], a.store1[2]}}, s2 = {{a.store2[0], a.store2[1], a.store2[2]}};
    const P3 st = sub == 0 ? s1 : s2;
    const P3 other = sub == 0 ? s2 : s1;
    const double* const* base = sub == 0 ? a.base1 : a.base2;
    double* const* acc = sub == 0 ? a.acc1 : a.acc2;
    double* const* out = sub == 0 ? a.out1 : a.out2;
    const uint32_t own = ldu(a.code, id);
    const double mgx = ldd(st.p[0], id), mgy = ldd(st.p[1], id), mgz = ldd(st.p[2], id);
    double v[6][3];
    uint32_t cd[6];
    if (kLanes) gather6_lanes_load<I>(st, a.code, x, y, z, static_cast<I>(a.nx), ny, nz, v, cd);
    else gather6<I>(st, a.code, x, y, z, static_cast<I>(a.nx), ny, nz, v, cd);
    const double o0 = ldd(other.p[0], id), o1 = ldd(other.p[1], id), o2 = ldd(other.p[2], id);
    // a null Hstat is zero (no statics, nothing uploaded)
    const bool hs = a.hstat[0] != nullptr;
    const double hs0 = hs ? ldd(a.hstat[0], id) : 0., hs1 = hs ? ldd(a.hstat[1], id) : 0.;
    const double hs2 = hs ? ldd(a.hstat[2], id) : 0.;
    // at stage 1 the stage store is m itself: read it once
    const bool same = base[0] == st.p[0] && base[1] == st.p[1] && base[2] == st.p[2];
    double mb[3], ac[3];
    mb[0] = same ? mgx : ldd_nt(base[0], id);
    mb[1] = same ? mgy : ldd_nt(base[1], id);
    mb[2] = same ? mgz : ldd_nt(base[2], id);
    for (int c = 0; c < 3; c++) ac[c] = a.stage > 1 ? ldd_nt(acc[c], id) : 0.;
    if (kLanes) {
        const double c3[3] = {mgx, mgy, mgz};
        gather6_lanes_shift<I>(z, nz, c3, own, v, cd);
    }
    if (!(own & kIsAFM)) return;
    GMat& M = mat_of(a.mats, own);
    double mx = mgx, my = mgy, mz = mgz;
    if (a.rotate) rot(a.Tg2c, mgx, mgy, mgz, mx, my, mz);
    const Stencil s = stencil_core<false>(mgx, mgy, mgz, v, cd, (GMat*)a.mats, own, kAfmSurfShift, sub, a.dx, a.dy,
                                          a.dz, M.aAex[sub], M.aAex_self[sub], 0.);
    // uniaxial anisotropy 2K1 (u.m) u (magnetic_AFM_solver.cpp:148-154)
    const double udotm = mx * M.u[0] + my * M.u[1] + mz * M.u[2];
    const double hac0 = 2. * M.aK1n[sub] * udotm * M.u[0];
    const double hac1 = 2. * M.aK1n[sub] * udotm * M.u[1];
    const double hac2 = 2. * M.aK1n[sub] * udotm * M.u[2];
    double ha0 = hac0, ha1 = hac1, ha2 = hac2;
    if (a.rotate) rot(a.Tc2g, hac0, hac1, hac2, ha0, ha1, ha2);
    // magnetoelastic (magnetic_AFM_solver.cpp:89-96, 156-162), exchange,
    // inter-sublattice coupling and drive (:170-200)
    double H0 = hs0 + ha0;
    double H1 = hs1 + ha1;
    double H2 = hs2 + ha2;
    if (a.strain[0]) {
        double he0, he1, he2;
        magnetoelastic(a.strain, a.strain0, id, -1., M.aB1n[sub], M.aB2n[sub], mx, my, mz, a.rotate, a.Tc2g, he0, he1, he2);
        H0 = H0 + he0;
        H1 = H1 + he1;
        H2 = H2 + he2;
    }
    H0 = H0 + M.aExch[sub] * s.ex[0] + M.aJ[sub] * o0 + a.hext[0];
    H1 = H1 + M.aExch[sub] * s.ex[1] + M.aJ[sub] * o1 + a.hext[1];
    H2 = H2 + M.aExch[sub] * s.ex[2] + M.aJ[sub] * o2 + a.hext[2];
    if (a.hcell[0]) {
        // electrodynamic back-action (magnetic_AFM_solver.cpp:222-226)
        H0 = H0 + ldd(a.hcell[0], id);
        H1 = H1 + ldd(a.hcell[1], id);
        H2 = H2 + ldd(a.hcell[2], id);
    }
    const double px = mgy * H2 - mgz * H1, py = mgz * H0 - mgx * H2, pz = mgx * H1 - mgy * H0;
    const double qx = mgy * pz - mgz * py, qy = mgz * px - mgx * pz, qz = mgx * py - mgy * px;
    const double k0 = M.aPrecess[sub] * px + M.aAlphaPrecess[sub] * qx;
    const double k1 = M.aPrecess[sub] * py + M.aAlphaPrecess[sub] * qy;
    const double k2 = M.aPrecess[sub] * pz + M.aAlphaPrecess[sub] * qz;
    double* const* dmo = sub == 0 ? a.dm_out1 : a.dm_out2;
    if (dmo[0]) {
        std_(dmo[0], id, k0); std_(dmo[1], id, k1); std_(dmo[2], id, k2);
    }
    rk_store(a.stage, id, 0, k0, mb[0], ac[0], acc, out);
    rk_store(a.stage, id, 1, k1, mb[1], ac[1], acc, out);
    rk_store(a.stage, id, 2, k2, mb[2], ac[2], acc, out);
}

// Workgroup b serves cell block (b / 16) * 8 + b % 8 and sublattice
// (b / 8) % 2: the two sublattices of a cell block land on the same XCD
// (workgroups are dealt to the 8 XCDs round-robin) one dispatch round
// apart, so the second finds the shared reads (code, Hstat, the other
// sublattice's store) in that XCD's L2.
template <class I>
__global__ void __launch_bounds__(256) afm_stage_kernel(AfmStageArgs a) {
    const unsigned b = blockIdx.x;
    const unsigned blk = (b >> 4) * 8u + (b & 7u);
    const int sub = static_cast<int>((b >> 3) & 1u);
    const I n = static_cast<I>(a.nx * a.ny * a.nz);
    const I id = static_cast<I>(blk) * static_cast<I>(blockDim.x) + static_cast<I>(threadIdx.x);
    if (id >= n) return;
    afm_point<I, true>(a, id, sub);
}

// Batched stages: one launch covers every instance of a batch.  The stage
// arguments of instance i and stage s live in device memory at
// items[4*i + s - 1]; workgroup slot b (of L.total slots) serves instance
// L.item[b], the chunk of L.cpt * blockDim.x cells starting at
// (b - L.block0[instance]) * L.cpt * blockDim.x, one cell per thread per
// pass.  A grid smaller than L.total walks the slots with a stride of
// gridDim.x.
template <class I>
__global__ void __launch_bounds__(256) mag_batch_kernel(const StageArgs* __restrict__ items, BatchLaunch L, int stage) {
    for (uint32_t b = blockIdx.x; b < L.total; b += gridDim.x) {
        const uint32_t inst = L.item[b];
        const StageArgs& a = items[4 * inst + stage - 1];
        const I n = static_cast<I>(a.nx * a.ny * a.nz);
        const I chunk = static_cast<I>(L.cpt) * static_cast<I>(blockDim.x);
        const I first = static_cast<I>(b - L.block0[inst]) * chunk + static_cast<I>(threadIdx.x);
        for (int j = 0; j < L.cpt; j++) {
            const I id = first + static_cast<I>(j) * static_cast<I>(blockDim.x);
            if (id < n) fm_point<I, true>(a, id);
        }
    }
}

// Work item w of 2 * ceil(total / 8) * 8: slot (w / 16) * 8 + w % 8 and
// sublattice (w / 8) % 2, so that the two sublattices of a slot run on one
// XCD one dispatch round apart (see afm_stage_kernel).
// kOne: one work item per workgroup and one cell per thread (no loops).
template <class I, bool kOne>
__global__ void __launch_bounds__(256) afm_batch_kernel(const AfmStageArgs* __restrict__ items, BatchLaunch L,
                                                         int stage) {
    const uint32_t work = 2u * ((L.total + 7u) / 8u * 8u);
    if (kOne) {
        const uint32_t w = blockIdx.x;
        const uint32_t b = (w >> 4) * 8u + (w & 7u);
        if (w >= work || b >= L.total) return;
        const uint32_t inst = L.item[b];
        const AfmStageArgs& a = items[4 * inst + stage - 1];
        const I n = static_cast<I>(a.nx * a.ny * a.nz);
        const I id = static_cast<I>(b - L.block0[inst]) * static_cast<I>(blockDim.x) + static_cast<I>(threadIdx.x);
        if (id < n) afm_point<I, true>(a, id, static_cast<int>((w >> 3) & 1u));
        return;
    }
    for (uint32_t w = blockIdx.x; w < work; w += gridDim.x) {
        const uint32_t b = (w >> 4) * 8u + (w & 7u);
        const int sub = static_cast<int>((w >> 3) & 1u);
        if (b >= L.total) continue;
        const uint32_t inst = L.item[b];
        const AfmStageArgs& a = items[4 * inst + stage - 1];
        const I n = static_cast<I>(a.nx * a.ny * a.nz);
        const I chunk = static_cast<I>(L.cpt) * static_cast<I>(blockDim.x);
        const I first = static_cast<I>(b - L.block0[inst]) * chunk + static_cast<I>(threadIdx.x);
        for (int j = 0; j < L.cpt; j++) {
            const I id = first + static_cast<I>(j) * static_cast<I>(blockDim.x);
            if (id < n) afm_point<I>(a, id, sub);   // lane gathers cost this loop a wave of occupancy
        }
    }
}

// Per-instance clock of a batch: t += dt (global.cpp:513-549), then the
// sinusoidal drive at each stage time t - dt, t - dt/2, t - dt/2, t
// (magnetic_solver.cpp:1834-1861, magnetic_AFM_solver.cpp:61-64) written
// into the stage arguments of both magnet kinds.
// A batch without FM (or AFM) instances has no table of that kind: its
// pointer is null and is not written.
__global__ void __launch_bounds__(256) batch_clock_kernel(BatchClock* clk, StageArgs* fm, AfmStageArgs* afm,
                                                           int count) {
    const int i = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= count) return;
    const BatchClock c = clk[i];
    const double t = c.t + c.dt;
    clk[i].t = t;
    for (int s = 1; s <= 4; s++) {
        const double ts = s == 1 ? t - c.dt : (s == 4 ? t : t - c.dt * 0.5);
        for (int k = 0; k < 3; k++) {
            const double h = c.stat[k] + c.altr[k] * sin(2. * kPi * c.freq[k] * ts);
            if (fm) fm[4 * i + s - 1].hext[k] = h;
            if (afm) afm[4 * i + s - 1].hext[k] = h;
        }
    }
}

// update_prescribed_m (magnetic_solver.cpp:1937-1957): the FM cells of the
// bottom layer k = 0 take the cone value v.
__global__ void __launch_bounds__(256) prescribe_plane_kernel(double* m0, double* m1, double* m2,
                                                              const uint32_t* code, int64_t nxy, int64_t nz,
                                                              double v0, double v1, double v2) {
    const int64_t col = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (col >= nxy) return;
    const int64_t id = col * nz;
    if (!(code[id] & kIsFM)) return;
    m0[id] = v0; m1[id] = v1; m2[id] = v2;
}

// Per-block partial sums of three fields; block b writes partial[3*b + c].
__global__ void __launch_bounds__(256) sum3_kernel(const double* f0, const double* f1, const double* f2, int64_t n,
                                                    double* partial) {
    __shared__ double red[3][256];
    double s0 = 0., s1 = 0., s2 = 0.;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        s0 += f0[i]; s1 += f1[i]; s2 += f2[i];
    }
    red[0][threadIdx.x] = s0; red[1][threadIdx.x] = s1; red[2][threadIdx.x] = s2;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (static_cast<int>(threadIdx.x) < w) {
            red[0][threadIdx.x] += red[0][threadIdx.x + w];
            red[1][threadIdx.x] += red[1][threadIdx.x + w];
            red[2][threadIdx.x] += red[2][threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partial[3 * blockIdx.x + 0] = red[0][0];
        partial[3 * blockIdx.x + 1] = red[1][0];
        partial[3 * blockIdx.x + 2] = red[2][0];
    }
}

}  // namespace

// 32-bit element offsets when every cell array holds fewer than 2^28 doubles.
bool narrow(int64_t n) { return n < (int64_t{1} << 28); }

hipError_t launch_mag_stage(const StageArgs& a, int64_t n, hipStream_t s, const uint32_t* ids, int64_t count) {
    if (ids) {
        if (count > 0)
            hipLaunchKernelGGL(mag_stage_list_kernel, dim3(static_cast<unsigned>((count + 255) / 256)), dim3(256), 0, s,
                               a, ids, static_cast<uint32_t>(count));
        return hipGetLastError();
    }
    const int64_t blocks = (n + 255) / 256;
    if (narrow(n)) hipLaunchKernelGGL(mag_stage_kernel<uint32_t>, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, s, a);
    else hipLaunchKernelGGL(mag_stage_kernel<int64_t>, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_mag_stage_fused(const StageArgs& a, const FusedZArgs& f, bool hspec, hipStream_t s) {
    const int nz = static_cast<int>(a.nz);
    const int pairs = f.P > 1 ? f.P / 2 : 1;
    const bool lean = !a.rotate && !a.strain[0] && !a.hnu[0] && !a.hdrive[0] && !a.hcell[0] && !a.dm_out[0] &&
                      !a.prescribe && a.nmat >= 1 && a.nmat <= kLeanMats;
    const size_t lds = sizeof(double2) * (nz / 2 + fft_lds_elems(3 * pairs * nz)) +
                       (lean ? sizeof(DevMaterial) * static_cast<size_t>(a.nmat) : 0);
    const unsigned groups = static_cast<unsigned>(a.nx * a.ny / f.P);
    const bool n32 = narrow(a.nx * a.ny * a.nz);
    const dim3 g(groups), t(256);
    if (hspec && lean && n32) hipLaunchKernelGGL((mag_stage_fused_kernel<true, true, uint32_t>), g, t, lds, s, a, f);
    else if (hspec && lean) hipLaunchKernelGGL((mag_stage_fused_kernel<true, true, int64_t>), g, t, lds, s, a, f);
    else if (hspec && n32) hipLaunchKernelGGL((mag_stage_fused_kernel<true, false, uint32_t>), g, t, lds, s, a, f);
    else if (hspec) hipLaunchKernelGGL((mag_stage_fused_kernel<true, false, int64_t>), g, t, lds, s, a, f);
    else if (n32) hipLaunchKernelGGL((mag_stage_fused_kernel<false, false, uint32_t>), g, t, lds, s, a, f);
    else hipLaunchKernelGGL((mag_stage_fused_kernel<false, false, int64_t>), g, t, lds, s, a, f);
    return hipGetLastError();
}

hipError_t launch_afm_stage(const AfmStageArgs& a, int64_t n, hipStream_t s) {
    const int64_t blocks = (n + 255) / 256;
    const dim3 grid(static_cast<unsigned>((blocks + 7) / 8 * 16));
    if (narrow(n)) hipLaunchKernelGGL(afm_stage_kernel<uint32_t>, grid, dim3(256), 0, s, a);
    else hipLaunchKernelGGL(afm_stage_kernel<int64_t>, grid, dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_mag_batch(const StageArgs* items, const BatchLaunch& L, int stage, hipStream_t s) {
    if (L.narrow) hipLaunchKernelGGL(mag_batch_kernel<uint32_t>, dim3(L.grid), dim3(L.threads), 0, s, items, L, stage);
    else hipLaunchKernelGGL(mag_batch_kernel<int64_t>, dim3(L.grid), dim3(L.threads), 0, s, items, L, stage);
    return hipGetLastError();
}

hipError_t launch_afm_batch(const AfmStageArgs* items, const BatchLaunch& L, int stage, hipStream_t s) {
    // L.grid counts slots; the grid covers both sublattices of each
    const unsigned grid = 2u * ((L.grid + 7u) / 8u * 8u);
    const bool one = L.cpt == 1 && L.grid == L.total;
    if (L.narrow && one)
        hipLaunchKernelGGL((afm_batch_kernel<uint32_t, true>), dim3(grid), dim3(L.threads), 0, s, items, L, stage);
    else if (L.narrow)
        hipLaunchKernelGGL((afm_batch_kernel<uint32_t, false>), dim3(grid), dim3(L.threads), 0, s, items, L, stage);
    else if (one)
        hipLaunchKernelGGL((afm_batch_kernel<int64_t, true>), dim3(grid), dim3(L.threads), 0, s, items, L, stage);
    else
        hipLaunchKernelGGL((afm_batch_kernel<int64_t, false>), dim3(grid), dim3(L.threads), 0, s, items, L, stage);
    return hipGetLastError();
}

hipError_t launch_batch_clock(BatchClock* clk, StageArgs* fm, AfmStageArgs* afm, int count, hipStream_t s) {
    hipLaunchKernelGGL(batch_clock_kernel, dim3(static_cast<unsigned>((count + 255) / 256)), dim3(256), 0, s, clk, fm,
                       afm, count);
    return hipGetLastError();
}

hipError_t launch_prescribe_plane(double* const* m, const uint32_t* code, int64_t nx, int64_t ny, int64_t nz,
                                  const double v[3], hipStream_t s) {
    const int64_t nxy = nx * ny;
    hipLaunchKernelGGL(prescribe_plane_kernel, dim3(static_cast<unsigned>((nxy + 255) / 256)), dim3(256), 0, s,
                       m[0], m[1], m[2], code, nxy, nz, v[0], v[1], v[2]);
    return hipGetLastError();
}

hipError_t launch_sum3(const double* const* f, int64_t n, double* partial, int64_t blocks, hipStream_t s) {
    hipLaunchKernelGGL(sum3_kernel, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, s, f[0], f[1], f[2], n,
                       partial);
    return hipGetLastError();
}

}  // namespace gf
