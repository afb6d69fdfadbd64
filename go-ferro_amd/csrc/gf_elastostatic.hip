// gf_elastostatic.hip — static strain of the eigenstrain of the magnets and
// ferroelectrics (GF_FLAG_ELASTOSTATIC) for gfx950, FP64.
//
// The reference declares get_strain_static (elastic_system.h:190) without a
// definition; its set-up fixes a homogeneous cubic stiffness (the
// component-wise largest c11, c12, c44 times scale_elasto,
// elastic_initialize.cpp:22-40) and the Green tensor G_ik of that medium
// over crystal-frame wave vectors (:340-389).  Here, in a periodic body:
//   sigma0 = C e0                    e0 the eigenstrain (elastic_dynamic_solver.cpp:21-73:
//                                    magnetostriction and electrostriction)
//   w(k)   = G(k) (sigma0(k) k)      u = -i w
//   e(k)   = (k (x) w + w (x) k) / 2 for k != 0,   e(0) = C^-1 sigma0(0)
// with a Nyquist component averaged over its two signs so that e is real.
// e0 is written out as the static eigenstrain ex*0t0_crt, which a
// ferroelectric reads beside the static strain (ferroelectric_solver.cpp:
// 636-650) and which main.cpp copies out with it (:148-156).
// This is not a hot path (once per step at most): six complex fields are
// transformed by one workgroup per line with the line in LDS.
#include <algorithm>
#include <cstring>

#include "gf_internal.h"

namespace gf {

namespace {

constexpr int kLineMax = 2048;

// Radix-2 complex FFT of `lines` lines of length n = 2^logn (1 <= logn <= 11),
// line q starting at (q / m) * A + q % m with stride S; sign -1 forward, +1
// inverse, unnormalised.
__global__ void __launch_bounds__(256) es_fft_lines(double2* data, int n, int logn, int64_t m, int64_t A, int64_t S,
                                                     double sign) {
    __shared__ double2 buf[kLineMax];
    const int64_t q = blockIdx.x;
    const int64_t base = (q / m) * A + q % m;
    for (int t = threadIdx.x; t < n; t += blockDim.x) {
        const unsigned r = __brev(static_cast<unsigned>(t)) >> (32 - logn);
        buf[r] = data[base + static_cast<int64_t>(t) * S];
    }
    __syncthreads();
    for (int len = 2; len <= n; len <<= 1) {
        const int half = len >> 1;
        for (int t = threadIdx.x; t < n / 2; t += blockDim.x) {
            const int g = t / half, pos = t - g * half;
            const int a = g * len + pos, b = a + half;
            double sn, cs;
            sincospi(sign * 2.0 * pos / len, &sn, &cs);
            const double2 u = buf[a], v = buf[b];
            const double vr = v.x * cs - v.y * sn, vi = v.x * sn + v.y * cs;
            buf[a] = make_double2(u.x + vr, u.y + vi);
            buf[b] = make_double2(u.x - vr, u.y - vi);
        }
        __syncthreads();
    }
    for (int t = threadIdx.x; t < n; t += blockDim.x) data[base + static_cast<int64_t>(t) * S] = buf[t];
}

struct EsArgs {
    const uint32_t* code;
    const DevMaterial* mats;
    const double* m[3];     // FM store, null when no FM
    const double* m1[3];    // AFM stores, null when no AFM
    const double* m2[3];
    const double* p[3];     // FE p store, null when no FE
    double2* f[6];          // sigma0, then the strain spectra, then the strain
    double* out[6];
    double* e0[6];          // the eigenstrain e0 (crystal frame), null when not kept
    int64_t nx, ny, nz;
    double dx, dy, dz;
    double c11, c12, c44;
    int32_t rotate;
    double Tg2c[3][3];
};

__device__ __forceinline__ void to_crystal(const EsArgs& a, double x, double y, double z, double& ox, double& oy,
                                           double& oz) {
    if (a.rotate) {
        ox = a.Tg2c[0][0] * x + a.Tg2c[0][1] * y + a.Tg2c[0][2] * z;
        oy = a.Tg2c[1][0] * x + a.Tg2c[1][1] * y + a.Tg2c[1][2] * z;
        oz = a.Tg2c[2][0] * x + a.Tg2c[2][1] * y + a.Tg2c[2][2] * z;
    } else {
        ox = x; oy = y; oz = z;
    }
}

// sigma0 = C e0 of every cell, e0 the magnetostrictive and electrostrictive
// eigenstrain in the crystal frame (zero in vacuum), in the reference's sum
// order; e0 itself written out when kept.
__global__ void __launch_bounds__(256) es_sigma_kernel(EsArgs a) {
    const int64_t n = a.nx * a.ny * a.nz;
    const int64_t id = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (id >= n) return;
    const uint32_t t = a.code[id] & kMatMask;
    double e[6] = {0., 0., 0., 0., 0., 0.};
    if (t) {
        const DevMaterial& M = a.mats[t - 1];
        double mx = 0., my = 0., mz = 0., ax = 0., ay = 0., az = 0., bx = 0., by = 0., bz = 0.;
        double px = 0., py = 0., pz = 0.;
        if (a.m[0]) to_crystal(a, a.m[0][id], a.m[1][id], a.m[2][id], mx, my, mz);
        if (a.m1[0]) {
            to_crystal(a, a.m1[0][id], a.m1[1][id], a.m1[2][id], ax, ay, az);
            to_crystal(a, a.m2[0][id], a.m2[1][id], a.m2[2][id], bx, by, bz);
        }
        if (a.p[0]) to_crystal(a, a.p[0][id], a.p[1][id], a.p[2][id], px, py, pz);
        e[0] = 1.5 * M.l100 * (mx * mx - 1. / 3.) + M.Q11 * (px * px) + M.Q12 * (py * py + pz * pz) +
               0.5 * 1.5 * M.l100a * (ax * ax - 1. / 3.) + 0.5 * 1.5 * M.l100b * (bx * bx - 1. / 3.);
        e[1] = 1.5 * M.l100 * (my * my - 1. / 3.) + M.Q11 * (py * py) + M.Q12 * (px * px + pz * pz) +
               0.5 * 1.5 * M.l100a * (ay * ay - 1. / 3.) + 0.5 * 1.5 * M.l100b * (by * by - 1. / 3.);
        e[2] = 1.5 * M.l100 * (mz * mz - 1. / 3.) + M.Q11 * (pz * pz) + M.Q12 * (py * py + px * px) +
               0.5 * 1.5 * M.l100a * (az * az - 1. / 3.) + 0.5 * 1.5 * M.l100b * (bz * bz - 1. / 3.);
        e[3] = 1.5 * M.l111 * my * mz + M.Q44 * py * pz + 0.5 * 1.5 * M.l111a * ay * az +
               0.5 * 1.5 * M.l111b * by * bz;
        e[4] = 1.5 * M.l111 * mx * mz + M.Q44 * px * pz + 0.5 * 1.5 * M.l111a * ax * az +
               0.5 * 1.5 * M.l111b * bx * bz;
        e[5] = 1.5 * M.l111 * my * mx + M.Q44 * py * px + 0.5 * 1.5 * M.l111a * ay * ax +
               0.5 * 1.5 * M.l111b * by * bx;
    }
    if (a.e0[0])
        for (int c = 0; c < 6; c++) a.e0[c][id] = e[c];
    const double s[6] = {a.c11 * e[0] + a.c12 * (e[1] + e[2]), a.c11 * e[1] + a.c12 * (e[0] + e[2]),
                         a.c11 * e[2] + a.c12 * (e[0] + e[1]), 2. * a.c44 * e[3],
                         2. * a.c44 * e[4], 2. * a.c44 * e[5]};
    for (int c = 0; c < 6; c++) a.f[c][id] = make_double2(s[c], 0.);
}

// The strain of sigma0 at one global wave vector (elastic_initialize.cpp:
// 340-389 for G), one real/imaginary part.
__device__ void es_operator(const EsArgs& a, const double kg[3], const double sig[6][2], double eps[6][2]) {
    const double c11 = a.c11, c12 = a.c12, c44 = a.c44;
    double kq[3];
    to_crystal(a, kg[0], kg[1], kg[2], kq[0], kq[1], kq[2]);
    const double kn = kg[0] * kg[0] + kg[1] * kg[1] + kg[2] * kg[2];
    if (kn == 0.) {
        const double den = (c11 - c12) * (c11 + 2. * c12);
        for (int r = 0; r < 2; r++) {
            eps[0][r] = ((c11 + c12) * sig[0][r] - c12 * (sig[1][r] + sig[2][r])) / den;
            eps[1][r] = ((c11 + c12) * sig[1][r] - c12 * (sig[0][r] + sig[2][r])) / den;
            eps[2][r] = ((c11 + c12) * sig[2][r] - c12 * (sig[0][r] + sig[1][r])) / den;
            eps[3][r] = sig[3][r] / (2. * c44);
            eps[4][r] = sig[4][r] / (2. * c44);
            eps[5][r] = sig[5][r] / (2. * c44);
        }
        return;
    }
    const double chi = (c11 - c12 - 2. * c44) / c44;
    const double x2 = kq[0] * kq[0], y2 = kq[1] * kq[1], z2 = kq[2] * kq[2];
    const double d0 = c11 * kn * kn * kn + chi * (c11 + c12) * kn * (x2 * y2 + y2 * z2 + z2 * x2) +
                      chi * chi * (c11 + 2. * c12 + c44) * x2 * y2 * z2;
    double G[3][3];
    G[0][0] = (c44 * kn * kn + (c11 + c12) * chi * y2 * z2 + (c11 - c44) * kn * (y2 + z2)) / (c44 * d0);
    G[1][1] = (c44 * kn * kn + (c11 + c12) * chi * z2 * x2 + (c11 - c44) * kn * (z2 + x2)) / (c44 * d0);
    G[2][2] = (c44 * kn * kn + (c11 + c12) * chi * x2 * y2 + (c11 - c44) * kn * (x2 + y2)) / (c44 * d0);
    G[0][1] = G[1][0] = -(c12 + c44) * (kn + chi * z2) * kq[0] * kq[1] / (c44 * d0);
    G[1][2] = G[2][1] = -(c12 + c44) * (kn + chi * x2) * kq[1] * kq[2] / (c44 * d0);
    G[0][2] = G[2][0] = -(c12 + c44) * (kn + chi * y2) * kq[0] * kq[2] / (c44 * d0);
    for (int r = 0; r < 2; r++) {
        const double S[3][3] = {{sig[0][r], sig[5][r], sig[4][r]},
                                {sig[5][r], sig[1][r], sig[3][r]},
                                {sig[4][r], sig[3][r], sig[2][r]}};
        double t[3], w[3];
        for (int m = 0; m < 3; m++) t[m] = S[m][0] * kq[0] + S[m][1] * kq[1] + S[m][2] * kq[2];
        for (int i = 0; i < 3; i++) w[i] = G[i][0] * t[0] + G[i][1] * t[1] + G[i][2] * t[2];
        eps[0][r] = kq[0] * w[0];
        eps[1][r] = kq[1] * w[1];
        eps[2][r] = kq[2] * w[2];
        eps[3][r] = (kq[2] * w[1] + kq[1] * w[2]) / 2.;
        eps[4][r] = (kq[2] * w[0] + kq[0] * w[2]) / 2.;
        eps[5][r] = (kq[1] * w[0] + kq[0] * w[1]) / 2.;
    }
}

__device__ __forceinline__ double es_wave(int64_t i, int64_t n, double h, bool& nyq) {
    const int64_t t = i <= n / 2 ? i : i - n;
    nyq = n >= 2 && 2 * i == n;
    return static_cast<double>(t) * (2. * kPi / (static_cast<double>(n) * h));
}

__global__ void __launch_bounds__(128) es_gamma_kernel(EsArgs a) {
    const int64_t nx = a.nx, ny = a.ny, nz = a.nz;
    const int64_t id = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (id >= nx * ny * nz) return;
    const int64_t i = id / (ny * nz), j = (id / nz) % ny, k = id % nz;
    bool nyq[3];
    const double kg[3] = {es_wave(i, nx, a.dx, nyq[0]), es_wave(j, ny, a.dy, nyq[1]), es_wave(k, nz, a.dz, nyq[2])};
    double sig[6][2], acc[6][2], eps[6][2];
    for (int c = 0; c < 6; c++) {
        const double2 v = a.f[c][id];
        sig[c][0] = v.x; sig[c][1] = v.y;
        acc[c][0] = 0.; acc[c][1] = 0.;
    }
    int nv = 0;
    for (int f = 0; f < 8; f++) {
        if (((f & 1) && !nyq[0]) || ((f & 2) && !nyq[1]) || ((f & 4) && !nyq[2])) continue;
        const double kf[3] = {(f & 1) ? -kg[0] : kg[0], (f & 2) ? -kg[1] : kg[1], (f & 4) ? -kg[2] : kg[2]};
        es_operator(a, kf, sig, eps);
        for (int c = 0; c < 6; c++) { acc[c][0] += eps[c][0]; acc[c][1] += eps[c][1]; }
        nv++;
    }
    for (int c = 0; c < 6; c++) a.f[c][id] = make_double2(acc[c][0] / nv, acc[c][1] / nv);
}

__global__ void __launch_bounds__(256) es_out_kernel(EsArgs a) {
    const int64_t n = a.nx * a.ny * a.nz;
    const int64_t id = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (id >= n) return;
    for (int c = 0; c < 6; c++) a.out[c][id] = a.f[c][id].x / static_cast<double>(n);
}

unsigned es_blocks(int64_t n, int threads) { return static_cast<unsigned>((n + threads - 1) / threads); }

int log2_of(int64_t n) {
    int l = 0;
    while ((int64_t{1} << l) < n) l++;
    return l;
}

// FFTs of one field along the three axes (z, y, x forward; x, y, z inverse).
hipError_t es_fft3(const gf_ctx* c, double2* f, double sign, hipStream_t s) {
    const int64_t nx = c->nx, ny = c->ny, nz = c->nz;
    struct Axis { int64_t n, lines, m, A, S; };
    const Axis z{nz, nx * ny, 1, nz, 1}, y{ny, nx * nz, nz, ny * nz, nz}, x{nx, ny * nz, ny * nz, 0, ny * nz};
    const Axis order[3] = {sign < 0 ? z : x, y, sign < 0 ? x : z};
    for (const Axis& ax : order) {
        if (ax.n < 2) continue;
        hipLaunchKernelGGL(es_fft_lines, dim3(static_cast<unsigned>(ax.lines)), dim3(256), 0, s, f,
                           static_cast<int>(ax.n), log2_of(ax.n), ax.m, ax.A, ax.S, sign);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

}  // namespace

gf_status elastostatic_setup(gf_ctx* c, const gf_desc* d) {
    auto pow2 = [](int64_t v) { return v >= 1 && (v & (v - 1)) == 0; };
    if (!pow2(c->nx) || !pow2(c->ny) || !pow2(c->nz) || c->nx > kLineMax || c->ny > kLineMax || c->nz > kLineMax)
        return set_error(c, GF_ERR_UNSUPPORTED, "the elastostatics needs power-of-two nx, ny, nz up to 2048");
    if (!(c->any_fm || c->any_afm || c->any_fe))
        return set_error(c, GF_ERR_UNSUPPORTED, "the elastostatics needs a magnetic or ferroelectric material");
    if (c->flags & GF_FLAG_FLEXO)
        return set_error(c, GF_ERR_UNSUPPORTED, "the elastostatics runs without the flexoelectric coupling");
    const double sc = d->scale_elasto > 0. ? d->scale_elasto : 1.;
    double c11 = c->host_mats[0].c11, c12 = c->host_mats[0].c12, c44 = c->host_mats[0].c44;
    for (const auto& m : c->host_mats) {
        c11 = std::max(c11, m.c11);
        c12 = std::max(c12, m.c12);
        c44 = std::max(c44, m.c44);
    }
    c->es_c[0] = c11 * sc; c->es_c[1] = c12 * sc; c->es_c[2] = c44 * sc;
    if (!(c->es_c[0] > c->es_c[1]) || !(c->es_c[0] + 2. * c->es_c[1] > 0.) || !(c->es_c[2] > 0.))
        return set_error(c, GF_ERR_ARG, "the elastostatics needs c11 > c12, c11 + 2 c12 > 0 and c44 > 0");
    gf_status st = hip_check(c, hipMalloc(&c->es_buf, sizeof(double2) * 6 * static_cast<size_t>(c->n)),
                             "hipMalloc elastostatic spectra");
    for (int a = 0; a < 6 && !st; a++) {
        if (c->e_t0[a]) continue;
        st = hip_check(c, hipMalloc(&c->e_t0[a], sizeof(double) * static_cast<size_t>(c->n)), "hipMalloc static strain");
        if (!st) st = hip_check(c, hipMemset(c->e_t0[a], 0, sizeof(double) * static_cast<size_t>(c->n)), "clear static strain");
    }
    // the static eigenstrain of each solve, kept (and carried by the state
    // image) for the ferroelectric stage and the eigenstrain output
    for (int a = 0; a < 6 && !st; a++) {
        if (c->e0t0[a]) continue;
        st = hip_check(c, hipMalloc(&c->e0t0[a], sizeof(double) * static_cast<size_t>(c->n)), "hipMalloc static eigenstrain");
        if (!st) st = hip_check(c, hipMemset(c->e0t0[a], 0, sizeof(double) * static_cast<size_t>(c->n)), "clear static eigenstrain");
    }
    if (!st) c->has_e0t0 = true;
    // with elastodynamics the magnets read the static strain beside the
    // dynamic one, and the prep launch subtracts the static eigenstrain
    if (!st && c->el_on) c->has_et0_dyn = true;
    return st;
}

gf_status elastostatic_run(gf_ctx* c, double* const* fm, double* const* a1, double* const* a2, double* const* p) {
    EsArgs a{};
    a.code = c->d_code; a.mats = c->d_mats;
    for (int i = 0; i < 3; i++) {
        a.m[i] = fm ? fm[i] : nullptr;
        a.m1[i] = a1 ? a1[i] : nullptr;
        a.m2[i] = a2 ? a2[i] : nullptr;
        a.p[i] = p ? p[i] : nullptr;
    }
    for (int q = 0; q < 6; q++) { a.f[q] = c->es_buf + q * c->n; a.out[q] = c->e_t0[q]; a.e0[q] = c->e0t0[q]; }
    a.nx = c->nx; a.ny = c->ny; a.nz = c->nz; a.dx = c->dx; a.dy = c->dy; a.dz = c->dz;
    a.c11 = c->es_c[0]; a.c12 = c->es_c[1]; a.c44 = c->es_c[2];
    a.rotate = (c->flags & GF_FLAG_ORIENT_ROTATE) ? 1 : 0;
    std::memcpy(a.Tg2c, c->Tg2c, sizeof(a.Tg2c));
    hipEvent_t ev = nullptr;
    timer_begin(c, GF_K_OTHER, &ev);
    hipLaunchKernelGGL(es_sigma_kernel, dim3(es_blocks(c->n, 256)), dim3(256), 0, c->stream, a);
    hipError_t e = hipGetLastError();
    for (int q = 0; q < 6 && e == hipSuccess; q++) e = es_fft3(c, a.f[q], -1., c->stream);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(es_gamma_kernel, dim3(es_blocks(c->n, 128)), dim3(128), 0, c->stream, a);
        e = hipGetLastError();
    }
    for (int q = 0; q < 6 && e == hipSuccess; q++) e = es_fft3(c, a.f[q], 1., c->stream);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(es_out_kernel, dim3(es_blocks(c->n, 256)), dim3(256), 0, c->stream, a);
        e = hipGetLastError();
    }
    timer_end(c, GF_K_OTHER, ev);
    return hip_check(c, e, "elastostatics");
}

}  // namespace gf
