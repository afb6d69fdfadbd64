// gf_fp32.hip — FP32 variant of the adjoint sweep (forward trajectory,
// checkpointed reverse sweep) for the ferromagnet physics of the FP64 sweep
// in gf_adjoint.hip, restricted to contexts without frame rotation and
// without a non-uniform static field: exchange, cubic or uniaxial
// anisotropy, uniform and shaped drives, the magnetoelastic field of a
// static strain, and FFT or no magnetostatics (magnetic_solver.cpp:3-229,
// magnetostatic_solver.cpp:114-249, differentiated as in gf_adjoint.hip).
//
// Every state array, field, spectrum, tensor spectrum, checkpoint and
// cotangent is FP32 and the kernels compute in FP32; only the probe
// spectral loss (a sum over N steps) and the final gradient sums are formed
// in FP64.  The context's own FP64 state is not touched: the sweep starts
// from its m rounded to FP32 and its time, and leaves both as they were.
// The algorithms are those of the FP64 path: the same stencil, anisotropy,
// LLG increment, RK4 accumulator order, five-pass FFT convolution with the
// context's Newell spectra (rounded to FP32) and the same discrete adjoint.
#include "gf_llg_dev.h"

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <vector>

namespace gf {

// The FM material's coefficients in FP32 (the FP64 record rounded).
struct Coef32 {
    float Ms, K1n, K2n, B1n, B2n, Aself, exch, precess, alpha_precess;
    int anis;
};

struct Adj32 {
    gf_adjoint_desc d{};
    bool ready = false;
    int64_t N = 0, K = 0, nseg = 0;
    int material = 0;
    Coef32 co{};
    float alpha = 0.f, aex = 0.f;
    double t0 = 0.;                // context time at the start of the sweep
    std::vector<double> ck_time, tstep, g;
    double loss = 0.;
    bool ms = false;               // FFT magnetostatics
    // state and fields
    float *m = nullptr, *sa = nullptr, *sb = nullptr, *acc = nullptr, *hd = nullptr;
    float *hdrive = nullptr, *strain = nullptr;
    float* pool = nullptr;         // (nseg + 1) checkpoints of 3n
    float* X = nullptr;            // probe means
    // FFT plan
    int64_t nzh = 0, nh = 0;
    float2 *twx = nullptr, *twy = nullptr, *twz = nullptr;
    float2* kt = nullptr;          // 9 tensor spectra, layout [kz][y][x]
    float2 *spa = nullptr, *spb = nullptr;
    int pz = 1, py = 1;
    // reverse sweep
    int64_t kbuf = 0;
    float *Y = nullptr, *HD = nullptr, *work = nullptr, *gpar = nullptr;
    double* partial = nullptr;
};

namespace {

constexpr int kT = 256;
constexpr int kBatch32 = 8;

unsigned blocks_of(int64_t n) { return static_cast<unsigned>((n + kT - 1) / kT); }

__device__ __forceinline__ void cross3(float ax, float ay, float az, float bx, float by, float bz, float& cx,
                                       float& cy, float& cz) {
    cx = ay * bz - az * by;
    cy = az * bx - ax * bz;
    cz = ax * by - ay * bx;
}

// The exchange sum of get_m_spatial_derivative_local (magnetic_solver.cpp:
// 1639-1831) in FP32 for one FM material: a face toward a cell that is not
// FM, or a non-periodic wall, is a surface with zero slope; every other
// face carries the material's self mean.
struct Ex3 {
    float v[3];
};

__device__ __forceinline__ Ex3 exch32(const float* const* f, const uint32_t* code, uint32_t own, int64_t id, int x,
                                      int y, int z, int nx, int ny, int nz, float ix, float iy, float iz, float A) {
    const int64_t nb[6] = {
        ((static_cast<int64_t>(x == 0 ? nx - 1 : x - 1) * ny + y) * nz + z),
        ((static_cast<int64_t>(x == nx - 1 ? 0 : x + 1) * ny + y) * nz + z),
        ((static_cast<int64_t>(x) * ny + (y == 0 ? ny - 1 : y - 1)) * nz + z),
        ((static_cast<int64_t>(x) * ny + (y == ny - 1 ? 0 : y + 1)) * nz + z),
        ((static_cast<int64_t>(x) * ny + y) * nz + (z == 0 ? nz - 1 : z - 1)),
        ((static_cast<int64_t>(x) * ny + y) * nz + (z == nz - 1 ? 0 : z + 1))};
    float c[3], v[6][3];
    for (int a = 0; a < 3; a++) c[a] = f[a][id];
    for (int q = 0; q < 6; q++)
        for (int a = 0; a < 3; a++) v[q][a] = f[a][nb[q]];
    const float inv[3] = {ix, iy, iz};
    Ex3 r = {{0.f, 0.f, 0.f}};
    for (int ax = 0; ax < 3; ax++) {
        const bool sb = (own >> (kFmSurfShift + 2 * ax)) & 1u;
        const bool sf = (own >> (kFmSurfShift + 2 * ax + 1)) & 1u;
        const float d = inv[ax];
        for (int a = 0; a < 3; a++) {
            const float bw = sb ? 0.f : (c[a] - v[2 * ax][a]) * d;
            const float fw = sf ? 0.f : (v[2 * ax + 1][a] - c[a]) * d;
            r.v[a] = r.v[a] + A * (fw * d) + A * (-bw * d);
        }
    }
    (void)code;
    return r;
}

struct S32Args {
    const uint32_t* code;
    const float* store[3];
    const float* base[3];
    float* acc[3];
    float* out[3];
    const float* hd[3];            // field of the store (or the static field), may be null
    const float* hdrive[3];        // may be null
    const float* strain[6];        // may be null
    int nx, ny, nz;
    float ix, iy, iz;
    float hext[3];
    float drive_s;
    Coef32 m;
    int stage;
};

// H_eff of one FM cell (magnetic_solver.cpp:124-198) in FP32, the
// anisotropy Jacobian J (symmetric) when asked for.
struct H32 {
    float H[3];
    float hex[3];
    float J00, J11, J22, J01, J02, J12;
    float e[6];
};

__device__ __forceinline__ H32 field32(const Coef32& M, const float* const* hd, const float* const* hdrive,
                                       const float* const* strain, const float hext[3], float drive_s, int64_t id,
                                       float mx, float my, float mz, const Ex3& ex) {
    H32 r{};
    for (int a = 0; a < 3; a++) r.hex[a] = M.exch * ex.v[a];
    float ha0 = 0.f, ha1 = 0.f, ha2 = 0.f;
    const float x2 = mx * mx, y2 = my * my, z2 = mz * mz;
    if (M.anis == 1) {
        r.J00 = -2.f * (M.K1n * (y2 + z2) + M.K2n * y2 * z2);
        r.J11 = -2.f * (M.K1n * (x2 + z2) + M.K2n * x2 * z2);
        r.J22 = -2.f * (M.K1n * (x2 + y2) + M.K2n * x2 * y2);
        r.J01 = -4.f * mx * my * (M.K1n + M.K2n * z2);
        r.J02 = -4.f * mx * mz * (M.K1n + M.K2n * y2);
        r.J12 = -4.f * my * mz * (M.K1n + M.K2n * x2);
        ha0 = r.J00 * mx;
        ha1 = r.J11 * my;
        ha2 = r.J22 * mz;
    } else if (M.anis == 2) {
        r.J22 = 2.f * M.K1n;
        ha2 = r.J22 * mz;
    }
    float H0 = ha0, H1 = ha1, H2 = ha2;
    if (hd[0]) { H0 += hd[0][id]; H1 += hd[1][id]; H2 += hd[2][id]; }
    if (strain[0]) {
        for (int q = 0; q < 6; q++) r.e[q] = strain[q][id];
        // -2 (B1 e_aa m_a + B2 (e_ab m_b + e_ac m_c)), Voigt xx yy zz yz xz xy
        H0 += -2.f * (M.B1n * r.e[0] * mx + M.B2n * (r.e[5] * my + r.e[4] * mz));
        H1 += -2.f * (M.B1n * r.e[1] * my + M.B2n * (r.e[5] * mx + r.e[3] * mz));
        H2 += -2.f * (M.B1n * r.e[2] * mz + M.B2n * (r.e[4] * mx + r.e[3] * my));
    }
    H0 += r.hex[0] + hext[0];
    H1 += r.hex[1] + hext[1];
    H2 += r.hex[2] + hext[2];
    if (hdrive[0]) {
        H0 += hdrive[0][id] * drive_s;
        H1 += hdrive[1][id] * drive_s;
        H2 += hdrive[2][id] * drive_s;
    }
    r.H[0] = H0; r.H[1] = H1; r.H[2] = H2;
    return r;
}

// One FM stage in FP32 with the accumulator form of the FP64 stage kernels.
__global__ void __launch_bounds__(kT) stage32_kernel(S32Args a) {
    const int64_t n = static_cast<int64_t>(a.nx) * a.ny * a.nz;
    const int64_t id = static_cast<int64_t>(blockIdx.x) * kT + threadIdx.x;
    if (id >= n) return;
    const uint32_t own = a.code[id];
    if (!(own & kIsFM)) return;
    const int z = static_cast<int>(id % a.nz), y = static_cast<int>((id / a.nz) % a.ny);
    const int x = static_cast<int>(id / (static_cast<int64_t>(a.ny) * a.nz));
    const float mx = a.store[0][id], my = a.store[1][id], mz = a.store[2][id];
    const Ex3 ex = exch32(a.store, a.code, own, id, x, y, z, a.nx, a.ny, a.nz, a.ix, a.iy, a.iz, a.m.Aself);
    const H32 h = field32(a.m, a.hd, a.hdrive, a.strain, a.hext, a.drive_s, id, mx, my, mz, ex);
    float px, py, pz, qx, qy, qz;
    cross3(mx, my, mz, h.H[0], h.H[1], h.H[2], px, py, pz);
    cross3(mx, my, mz, px, py, pz, qx, qy, qz);
    const float k[3] = {a.m.precess * px + a.m.alpha_precess * qx, a.m.precess * py + a.m.alpha_precess * qy,
                        a.m.precess * pz + a.m.alpha_precess * qz};
    for (int c = 0; c < 3; c++) {
        const float m0 = a.base[c][id];
        float v;
        if (a.stage == 1) {
            v = m0 + k[c] * 0.5f;
            a.acc[c][id] = k[c] / 6.f;
        } else if (a.stage == 2) {
            v = m0 + k[c] * 0.5f;
            a.acc[c][id] = a.acc[c][id] + k[c] / 3.f;
        } else if (a.stage == 3) {
            v = m0 + k[c];
            a.acc[c][id] = a.acc[c][id] + k[c] / 3.f;
        } else {
            v = m0 + (a.acc[c][id] + k[c] / 6.f);
        }
        a.out[c][id] = v;
    }
}

// --- FFT convolution in FP32: the five passes of gf_fft.hip --------------

struct Z32Args {
    const uint32_t* code;
    const float* in[3];
    float ms;                      // > 0: transform Ms*in on FM cells (zero elsewhere); 0: raw input
    float2* out;                   // [x][kz][y]
    const float2* tw;
    int64_t nh;
    int nx, ny, nz, nzh, lognz, pencils, logp;
};

__global__ void __launch_bounds__(kT) zfwd32_kernel(Z32Args a) {
    extern __shared__ float2 sm32[];
    const int c = blockIdx.z, x = blockIdx.y, y0 = blockIdx.x * a.pencils;
    float2* tw = sm32;
    float2* buf = sm32 + (a.nz >> 1);
    const float2 tw0 = fft_twiddle_prefetch(a.tw, a.nz);
    const float* in = a.in[c];
    fft_staged_load<kBatch32, float>(
        a.pencils * a.nz,
        [&](int idx) {
            const int y = y0 + (idx >> a.lognz);
            if (y >= a.ny) return 0.f;
            const int64_t cell = (static_cast<int64_t>(x) * a.ny + y) * a.nz + (idx & (a.nz - 1));
            if (a.ms > 0.f) return (a.code[cell] & kIsFM) ? a.ms * in[cell] : 0.f;
            return in[cell];
        },
        [&](int idx, float v) {
            buf[fpad((idx >> a.lognz) * a.nz + fft_bitrev(idx & (a.nz - 1), a.lognz))] = make_float2(v, 0.f);
        });
    fft_twiddle_store(tw, a.tw, a.nz, tw0);
    __syncthreads();
    fft_dit(buf, a.pencils, a.lognz, tw);
    float2* out = a.out + c * a.nh;
    for (int idx = threadIdx.x; idx < a.pencils * a.nzh; idx += blockDim.x) {
        const int r = idx & (a.pencils - 1), kz = idx >> a.logp, y = y0 + r;
        if (y < a.ny) out[(static_cast<int64_t>(x) * a.nzh + kz) * a.ny + y] = buf[fpad(r * a.nz + kz)];
    }
}

struct Y32Args {
    const float2* in;
    float2* out;
    const float2* tw;
    int64_t nh;
    int nx, ny, nzh, logny, pencils, logp;
};

// y forward: [x][kz][y] -> [kz][y][x]
__global__ void __launch_bounds__(kT) yfwd32_kernel(Y32Args a) {
    extern __shared__ float2 sm32[];
    const int c = blockIdx.z, kz = blockIdx.y, x0 = blockIdx.x * a.pencils;
    float2* tw = sm32;
    float2* buf = sm32 + (a.ny >> 1);
    const float2 tw0 = fft_twiddle_prefetch(a.tw, a.ny);
    const float2* in = a.in + c * a.nh;
    fft_staged_load<kBatch32, float2>(
        a.pencils * a.ny,
        [&](int idx) {
            const int x = x0 + (idx >> a.logny);
            return x < a.nx ? in[(static_cast<int64_t>(x) * a.nzh + kz) * a.ny + (idx & (a.ny - 1))]
                            : make_float2(0.f, 0.f);
        },
        [&](int idx, float2 v) { buf[fpad((idx >> a.logny) * a.ny + fft_bitrev(idx & (a.ny - 1), a.logny))] = v; });
    fft_twiddle_store(tw, a.tw, a.ny, tw0);
    __syncthreads();
    fft_dit(buf, a.pencils, a.logny, tw);
    float2* out = a.out + c * a.nh;
    for (int idx = threadIdx.x; idx < a.pencils * a.ny; idx += blockDim.x) {
        const int r = idx & (a.pencils - 1), y = idx >> a.logp, x = x0 + r;
        if (x < a.nx) out[(static_cast<int64_t>(kz) * a.ny + y) * a.nx + x] = buf[fpad(r * a.ny + y)];
    }
}

// y inverse: [kz][y][x] -> [x][kz][y]
__global__ void __launch_bounds__(kT) yinv32_kernel(Y32Args a) {
    extern __shared__ float2 sm32[];
    const int c = blockIdx.z, kz = blockIdx.y, x0 = blockIdx.x * a.pencils;
    float2* tw = sm32;
    float2* buf = sm32 + (a.ny >> 1);
    const float2 tw0 = fft_twiddle_prefetch(a.tw, a.ny);
    const float2* in = a.in + c * a.nh;
    fft_staged_load<kBatch32, float2>(
        a.pencils * a.ny,
        [&](int idx) {
            const int x = x0 + (idx & (a.pencils - 1));
            return x < a.nx ? in[(static_cast<int64_t>(kz) * a.ny + (idx >> a.logp)) * a.nx + x]
                            : make_float2(0.f, 0.f);
        },
        [&](int idx, float2 v) { buf[fpad((idx & (a.pencils - 1)) * a.ny + (idx >> a.logp))] = v; });
    fft_twiddle_store(tw, a.tw, a.ny, tw0);
    __syncthreads();
    ifft_dif(buf, a.pencils, a.logny, tw);
    float2* out = a.out + c * a.nh;
    for (int idx = threadIdx.x; idx < a.pencils * a.ny; idx += blockDim.x) {
        const int r = idx >> a.logny, y = idx & (a.ny - 1), x = x0 + r;
        if (x < a.nx)
            out[(static_cast<int64_t>(x) * a.nzh + kz) * a.ny + y] = buf[fpad(r * a.ny + fft_bitrev(y, a.logny))];
    }
}

struct X32Args {
    float2* data;                  // [kz][y][x], component c at data + c*nh
    const float2* kt;              // 9 tensor spectra, spectrum s at kt + s*nh
    const float2* tw;
    int64_t nh;
    float den;                     // -dx*dy*dz
    int nx, ny, lognx;
    int adjoint;                   // conjugate transpose of the tensor
};

// x forward transform, tensor product (magnetostatic_solver.cpp:163-177,
// spectra xx, yy, zz, xy, yz, zx, yx, zy, xz), x inverse, one y row of the
// three components per workgroup.
__global__ void __launch_bounds__(kT) xconv32_kernel(X32Args a) {
    extern __shared__ float2 sm32[];
    const int kz = blockIdx.y, y = blockIdx.x;
    float2* tw = sm32;
    float2* buf = sm32 + (a.nx >> 1);
    const float2 tw0 = fft_twiddle_prefetch(a.tw, a.nx);
    const int64_t row = (static_cast<int64_t>(kz) * a.ny + y) * a.nx;
    fft_staged_load<kBatch32, float2>(
        3 * a.nx, [&](int idx) { return a.data[(idx >> a.lognx) * a.nh + row + (idx & (a.nx - 1))]; },
        [&](int idx, float2 v) { buf[fpad((idx >> a.lognx) * a.nx + fft_bitrev(idx & (a.nx - 1), a.lognx))] = v; });
    fft_twiddle_store(tw, a.tw, a.nx, tw0);
    __syncthreads();
    fft_dit<4>(buf, 3, a.lognx, tw);
    for (int x = threadIdx.x; x < a.nx; x += blockDim.x) {
        const int64_t g = row + x;
        float2 K[9];
        for (int q = 0; q < 9; q++) {
            const int t = a.adjoint ? (q < 3 ? q : (q < 6 ? q + 3 : q - 3)) : q;
            const float2 v = a.kt[t * a.nh + g];
            K[q] = make_float2(v.x, a.adjoint ? -v.y : v.y);
        }
        const float2 mx = buf[fpad(x)], my = buf[fpad(a.nx + x)], mz = buf[fpad(2 * a.nx + x)];
        auto cm = [](float2 p, float2 q) { return make_float2(p.x * q.x - p.y * q.y, p.x * q.y + p.y * q.x); };
        float2 t0 = cm(mx, K[0]), t1 = cm(my, K[3]), t2 = cm(mz, K[8]);
        const float2 hx = make_float2((t0.x + t1.x + t2.x) / a.den, (t0.y + t1.y + t2.y) / a.den);
        t0 = cm(my, K[1]); t1 = cm(mz, K[4]); t2 = cm(mx, K[6]);
        const float2 hy = make_float2((t0.x + t1.x + t2.x) / a.den, (t0.y + t1.y + t2.y) / a.den);
        t0 = cm(mz, K[2]); t1 = cm(mx, K[5]); t2 = cm(my, K[7]);
        const float2 hz = make_float2((t0.x + t1.x + t2.x) / a.den, (t0.y + t1.y + t2.y) / a.den);
        buf[fpad(x)] = hx; buf[fpad(a.nx + x)] = hy; buf[fpad(2 * a.nx + x)] = hz;
    }
    __syncthreads();
    ifft_dif<4>(buf, 3, a.lognx, tw);
    for (int idx = threadIdx.x; idx < 3 * a.nx; idx += blockDim.x) {
        const int c = idx >> a.lognx, x = idx & (a.nx - 1);
        a.data[c * a.nh + row + x] = buf[fpad(c * a.nx + fft_bitrev(x, a.lognx))];
    }
}

struct ZI32Args {
    const float2* in;              // [x][kz][y]
    float* out[3];
    const float2* tw;
    int64_t nh;
    int nx, ny, nz, nzh, lognz, pencils, logp;
};

// z complex-to-real inverse (the Hermitian extension of gf_fft.hip's zinv).
__global__ void __launch_bounds__(kT) zinv32_kernel(ZI32Args a) {
    extern __shared__ float2 sm32[];
    const int c = blockIdx.z, x = blockIdx.y, y0 = blockIdx.x * a.pencils;
    float2* tw = sm32;
    float2* buf = sm32 + (a.nz >> 1);
    const float2 tw0 = fft_twiddle_prefetch(a.tw, a.nz);
    const float2* in = a.in + c * a.nh;
    fft_staged_load<kBatch32, float2>(
        a.pencils * a.nzh,
        [&](int idx) {
            const int y = y0 + (idx & (a.pencils - 1));
            return y < a.ny ? in[(static_cast<int64_t>(x) * a.nzh + (idx >> a.logp)) * a.ny + y]
                            : make_float2(0.f, 0.f);
        },
        [&](int idx, float2 v) {
            const int r = idx & (a.pencils - 1), kz = idx >> a.logp;
            buf[fpad(r * a.nz + kz)] = v;
            const int mirror = a.nz - kz;
            if (kz >= 1 && mirror > kz && mirror < a.nz) buf[fpad(r * a.nz + mirror)] = make_float2(v.x, -v.y);
        });
    fft_twiddle_store(tw, a.tw, a.nz, tw0);
    __syncthreads();
    ifft_dif(buf, a.pencils, a.lognz, tw);
    float* out = a.out[c];
    for (int idx = threadIdx.x; idx < a.pencils * a.nz; idx += blockDim.x) {
        const int r = idx >> a.lognz, k = idx & (a.nz - 1), y = y0 + r;
        if (y < a.ny) out[(static_cast<int64_t>(x) * a.ny + y) * a.nz + k] = buf[fpad(r * a.nz + fft_bitrev(k, a.lognz))].x;
    }
}

// --- adjoint kernels in FP32 (gf_adjoint.hip's adj_point / adj_combine) ---

struct P32Args {
    const uint32_t* code;
    const float* y[3];
    const float* hd[3];
    const float* hdrive[3];
    const float* strain[6];
    const float* kbar[3];
    float* mbar[3];
    float* hbar[3];
    float *galpha, *gaex, *gb1;
    float* ghdrive[3];
    int nx, ny, nz;
    float ix, iy, iz;
    float hext[3];
    float drive_s;
    Coef32 m;
    float alpha, aex;
};

__global__ void __launch_bounds__(kT) adj_point32_kernel(P32Args a) {
    const int64_t n = static_cast<int64_t>(a.nx) * a.ny * a.nz;
    const int64_t id = static_cast<int64_t>(blockIdx.x) * kT + threadIdx.x;
    if (id >= n) return;
    const uint32_t own = a.code[id];
    if (!(own & kIsFM)) {
        for (int c = 0; c < 3; c++) { a.mbar[c][id] = 0.f; a.hbar[c][id] = 0.f; }
        return;
    }
    const Coef32& M = a.m;
    const int z = static_cast<int>(id % a.nz), y = static_cast<int>((id / a.nz) % a.ny);
    const int x = static_cast<int>(id / (static_cast<int64_t>(a.ny) * a.nz));
    const float mx = a.y[0][id], my = a.y[1][id], mz = a.y[2][id];
    const Ex3 ex = exch32(a.y, a.code, own, id, x, y, z, a.nx, a.ny, a.nz, a.ix, a.iy, a.iz, M.Aself);
    const H32 h = field32(M, a.hd, a.hdrive, a.strain, a.hext, a.drive_s, id, mx, my, mz, ex);
    float px, py, pz, qx, qy, qz;
    cross3(mx, my, mz, h.H[0], h.H[1], h.H[2], px, py, pz);
    cross3(mx, my, mz, px, py, pz, qx, qy, qz);
    const float kb0 = a.kbar[0][id], kb1 = a.kbar[1][id], kb2 = a.kbar[2][id];
    const float P = M.precess, Q = M.alpha_precess;
    const float qb0 = Q * kb0, qb1 = Q * kb1, qb2 = Q * kb2;
    float t0, t1, t2;
    cross3(qb0, qb1, qb2, mx, my, mz, t0, t1, t2);
    const float pb0 = P * kb0 + t0, pb1 = P * kb1 + t1, pb2 = P * kb2 + t2;
    float mb0, mb1, mb2, u0, u1, u2;
    cross3(px, py, pz, qb0, qb1, qb2, mb0, mb1, mb2);
    cross3(h.H[0], h.H[1], h.H[2], pb0, pb1, pb2, u0, u1, u2);
    mb0 += u0; mb1 += u1; mb2 += u2;
    float hb0, hb1, hb2;
    cross3(pb0, pb1, pb2, mx, my, mz, hb0, hb1, hb2);
    // local Jacobians (anisotropy, magnetoelastic), both symmetric
    mb0 += h.J00 * hb0 + h.J01 * hb1 + h.J02 * hb2;
    mb1 += h.J01 * hb0 + h.J11 * hb1 + h.J12 * hb2;
    mb2 += h.J02 * hb0 + h.J12 * hb1 + h.J22 * hb2;
    if (a.strain[0]) {
        const float* e = h.e;
        mb0 += -2.f * (M.B1n * e[0] * hb0 + M.B2n * (e[5] * hb1 + e[4] * hb2));
        mb1 += -2.f * (M.B1n * e[1] * hb1 + M.B2n * (e[5] * hb0 + e[3] * hb2));
        mb2 += -2.f * (M.B1n * e[2] * hb2 + M.B2n * (e[4] * hb0 + e[3] * hb1));
    }
    // parameter gradients: P, Q in alpha; exchange linear in Aex; Hme linear in B1
    const float dP = -2.f * a.alpha * P / (1.f + a.alpha * a.alpha);
    const float dQ = P + a.alpha * dP;
    a.galpha[id] += dP * (px * kb0 + py * kb1 + pz * kb2) + dQ * (qx * kb0 + qy * kb1 + qz * kb2);
    a.gaex[id] += (hb0 * h.hex[0] + hb1 * h.hex[1] + hb2 * h.hex[2]) / a.aex;
    if (a.strain[0])
        a.gb1[id] += -2.f / static_cast<float>(kMu0) / M.Ms * (hb0 * h.e[0] * mx + hb1 * h.e[1] * my + hb2 * h.e[2] * mz);
    if (a.ghdrive[0]) {
        a.ghdrive[0][id] += a.drive_s * hb0;
        a.ghdrive[1][id] += a.drive_s * hb1;
        a.ghdrive[2][id] += a.drive_s * hb2;
    }
    a.mbar[0][id] = mb0; a.mbar[1][id] = mb1; a.mbar[2][id] = mb2;
    a.hbar[0][id] = hb0; a.hbar[1][id] = hb1; a.hbar[2][id] = hb2;
}

struct C32Args {
    const uint32_t* code;
    const float* mbar[3];
    const float* hbar[3];
    const float* dadj[3];          // null without magnetostatics
    float* lam[3];
    float* acc[3];
    float* kbar[3];
    int nx, ny, nz;
    float ix, iy, iz;
    float Aself, exch, Ms;
    int mode;                      // RK4 bookkeeping after stage `mode`
    int lo[3], hi[3];
    float g;
};

__global__ void __launch_bounds__(kT) adj_combine32_kernel(C32Args a) {
    const int64_t n = static_cast<int64_t>(a.nx) * a.ny * a.nz;
    const int64_t id = static_cast<int64_t>(blockIdx.x) * kT + threadIdx.x;
    if (id >= n) return;
    const uint32_t own = a.code[id];
    if (!(own & kIsFM)) return;
    const int z = static_cast<int>(id % a.nz), y = static_cast<int>((id / a.nz) % a.ny);
    const int x = static_cast<int>(id / (static_cast<int64_t>(a.ny) * a.nz));
    const Ex3 st = exch32(a.hbar, a.code, own, id, x, y, z, a.nx, a.ny, a.nz, a.ix, a.iy, a.iz, a.Aself);
    const bool probe = x >= a.lo[0] && x <= a.hi[0] && y >= a.lo[1] && y <= a.hi[1] && z >= a.lo[2] && z <= a.hi[2];
    for (int c = 0; c < 3; c++) {
        float yb = a.mbar[c][id] + a.exch * st.v[c];
        if (a.dadj[0]) yb += a.Ms * a.dadj[c][id];
        const float lam = a.lam[c][id];
        if (a.mode == 4) {
            a.acc[c][id] = lam + yb;
            a.kbar[c][id] = lam / 3.f + yb;
        } else if (a.mode == 3) {
            a.acc[c][id] = a.acc[c][id] + yb;
            a.kbar[c][id] = lam / 3.f + 0.5f * yb;
        } else if (a.mode == 2) {
            a.acc[c][id] = a.acc[c][id] + yb;
            a.kbar[c][id] = lam / 6.f + 0.5f * yb;
        } else {
            float l = a.acc[c][id] + yb;
            if (c == 0 && probe) l += a.g;
            a.lam[c][id] = l;
            a.kbar[c][id] = l / 6.f;
        }
    }
}

__global__ void __launch_bounds__(kT) adj_init32_kernel(const uint32_t* code, float* lamx, float* kbarx, int nx,
                                                         int ny, int nz, int lo0, int lo1, int lo2, int hi0, int hi1,
                                                         int hi2, float g) {
    const int64_t id = static_cast<int64_t>(blockIdx.x) * kT + threadIdx.x;
    if (id >= static_cast<int64_t>(nx) * ny * nz) return;
    if (!(code[id] & kIsFM)) return;
    const int z = static_cast<int>(id % nz), y = static_cast<int>((id / nz) % ny);
    const int x = static_cast<int>(id / (static_cast<int64_t>(ny) * nz));
    if (x < lo0 || x > hi0 || y < lo1 || y > hi1 || z < lo2 || z > hi2) return;
    lamx[id] = g;
    kbarx[id] = g / 6.f;
}

// Mean of m_x over the probe box in FP32, one workgroup, fixed order.
__global__ void __launch_bounds__(kT) probe32_kernel(const float* mx, int ny, int nz, int lo0, int lo1, int lo2,
                                                     int hi0, int hi1, int hi2, float* out) {
    __shared__ float red[kT];
    const int bx = hi0 - lo0 + 1, by = hi1 - lo1 + 1, bz = hi2 - lo2 + 1;
    const int cells = bx * by * bz;
    float s = 0.f;
    for (int t = threadIdx.x; t < cells; t += kT) {
        const int k = t % bz, j = (t / bz) % by, i = t / (by * bz);
        s += mx[(static_cast<int64_t>(lo0 + i) * ny + (lo1 + j)) * nz + (lo2 + k)];
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = kT / 2; w > 0; w >>= 1) {
        if (static_cast<int>(threadIdx.x) < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = red[0] / static_cast<float>(cells);
}

__global__ void __launch_bounds__(kT) to_f32_kernel(const double* in, float* out, int64_t n) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kT + threadIdx.x;
    if (i < n) out[i] = static_cast<float>(in[i]);
}

// Tensor spectra rounded to FP32, complex: real part from the FP64 array,
// imaginary part from the FP32 residue.
__global__ void __launch_bounds__(kT) tensor32_kernel(const double* re, const float* im, float2* out, int64_t n) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kT + threadIdx.x;
    if (i < n) out[i] = make_float2(static_cast<float>(re[i]), im[i]);
}

// Per-block FP64 partial sums of three FP32 arrays; block b writes
// partial[3*b + c].
__global__ void __launch_bounds__(kT) sum3_32_kernel(const float* f0, const float* f1, const float* f2, int64_t n,
                                                      double* partial) {
    __shared__ double red[3][kT];
    double s0 = 0., s1 = 0., s2 = 0.;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * kT + threadIdx.x; i < n;
         i += static_cast<int64_t>(gridDim.x) * kT) {
        s0 += f0[i]; s1 += f1[i]; s2 += f2[i];
    }
    red[0][threadIdx.x] = s0; red[1][threadIdx.x] = s1; red[2][threadIdx.x] = s2;
    __syncthreads();
    for (int w = kT / 2; w > 0; w >>= 1) {
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

constexpr int kSumBlocks = 256;

// --- host side ----------------------------------------------------------

int ilog2i(int64_t v) {
    int l = 0;
    while ((int64_t{1} << l) < v) l++;
    return l;
}

int pencils_for(int64_t len, int64_t budget, int64_t cap, int64_t other) {
    int64_t p = std::max<int64_t>(1, budget / len);
    p = std::min(p, cap);
    int64_t pow = 1;
    while (pow * 2 <= p) pow *= 2;
    while (pow > 1 && pow > other) pow /= 2;
    return static_cast<int>(pow);
}

template <class T>
gf_status alloc(gf_ctx* c, T** p, int64_t count, const char* what) {
    if (*p) { hipFree(*p); *p = nullptr; }
    return hip_check(c, hipMalloc(p, sizeof(T) * static_cast<size_t>(std::max<int64_t>(count, 1))), what);
}

gf_status twiddles32(gf_ctx* c, float2** dst, int64_t n) {
    const int64_t cnt = n / 2 > 0 ? n / 2 : 1;
    std::vector<float2> h(static_cast<size_t>(cnt));
    for (int64_t j = 0; j < cnt; j++) {
        const double ang = 2.0 * kPi * static_cast<double>(j) / static_cast<double>(n);
        h[static_cast<size_t>(j)] = make_float2(static_cast<float>(std::cos(ang)), static_cast<float>(-std::sin(ang)));
    }
    gf_status st = alloc(c, dst, cnt, "hipMalloc FP32 twiddles");
    if (!st) st = hip_check(c, hipMemcpy(*dst, h.data(), sizeof(float2) * h.size(), hipMemcpyHostToDevice), "twiddles");
    return st;
}

void slot3f(float* base, int64_t slot, int64_t n, float* out[3]) {
    for (int a = 0; a < 3; a++) out[a] = base + (slot * 3 + a) * n;
}

gf_status copy3f(gf_ctx* c, float* const* dst, const float* const* src) {
    for (int a = 0; a < 3; a++) {
        gf_status st = hip_check(c, hipMemcpyAsync(dst[a], src[a], sizeof(float) * static_cast<size_t>(c->n),
                                                   hipMemcpyDeviceToDevice, c->stream), "FP32 copy");
        if (st) return st;
    }
    return GF_OK;
}

gf_status to_f32(gf_ctx* c, const double* in, float* out, int64_t n) {
    hipLaunchKernelGGL(to_f32_kernel, dim3(blocks_of(n)), dim3(kT), 0, c->stream, in, out, n);
    return hip_check(c, hipGetLastError(), "FP32 conversion");
}

// FM-only convolution out = N (*) (Ms in) (ms > 0), or the adjoint
// convolution out = N^T (*) in of raw arrays (adjoint, ms = 0).
gf_status conv32(gf_ctx* c, Adj32* s, const float* const* in, float* const* out, bool adjoint) {
    const int nx = static_cast<int>(c->nx), ny = static_cast<int>(c->ny), nz = static_cast<int>(c->nz);
    const int nzh = static_cast<int>(s->nzh);
    Z32Args z{};
    z.code = c->d_code;
    for (int i = 0; i < 3; i++) z.in[i] = in[i];
    z.ms = adjoint ? 0.f : s->co.Ms;
    z.out = s->spa; z.tw = s->twz; z.nh = s->nh;
    z.nx = nx; z.ny = ny; z.nz = nz; z.nzh = nzh; z.lognz = ilog2i(nz); z.pencils = s->pz; z.logp = ilog2i(s->pz);
    const size_t lz = sizeof(float2) * static_cast<size_t>(nz / 2 + fft_lds_elems(static_cast<int64_t>(s->pz) * nz));
    hipLaunchKernelGGL(zfwd32_kernel, dim3((ny + s->pz - 1) / s->pz, nx, 3), dim3(kT), lz, c->stream, z);
    Y32Args y{};
    y.in = s->spa; y.out = s->spb; y.tw = s->twy; y.nh = s->nh;
    y.nx = nx; y.ny = ny; y.nzh = nzh; y.logny = ilog2i(ny); y.pencils = s->py; y.logp = ilog2i(s->py);
    const size_t ly = sizeof(float2) * static_cast<size_t>(ny / 2 + fft_lds_elems(static_cast<int64_t>(s->py) * ny));
    const dim3 gy((nx + s->py - 1) / s->py, nzh, 3);
    hipLaunchKernelGGL(yfwd32_kernel, gy, dim3(kT), ly, c->stream, y);
    X32Args x{};
    x.data = s->spb; x.kt = s->kt; x.tw = s->twx; x.nh = s->nh;
    x.den = static_cast<float>(-c->dx * c->dy * c->dz);
    x.nx = nx; x.ny = ny; x.lognx = ilog2i(nx); x.adjoint = adjoint ? 1 : 0;
    const size_t lx = sizeof(float2) * static_cast<size_t>(nx / 2 + fft_lds_elems(3 * static_cast<int64_t>(nx)));
    hipLaunchKernelGGL(xconv32_kernel, dim3(ny, nzh), dim3(kT), lx, c->stream, x);
    y.in = s->spb; y.out = s->spa;
    hipLaunchKernelGGL(yinv32_kernel, gy, dim3(kT), ly, c->stream, y);
    ZI32Args zi{};
    zi.in = s->spa;
    for (int i = 0; i < 3; i++) zi.out[i] = out[i];
    zi.tw = s->twz; zi.nh = s->nh;
    zi.nx = nx; zi.ny = ny; zi.nz = nz; zi.nzh = nzh; zi.lognz = ilog2i(nz); zi.pencils = s->pz;
    zi.logp = ilog2i(s->pz);
    hipLaunchKernelGGL(zinv32_kernel, dim3((ny + s->pz - 1) / s->pz, nx, 3), dim3(kT), lz, c->stream, zi);
    return hip_check(c, hipGetLastError(), "FP32 convolution");
}

void fill_common(gf_ctx* c, Adj32* s, int stage, float hext[3], float& drive_s, const float* hdr[3],
                 const float* str[6]) {
    StageArgs sa = fm_stage_args(c, stage, 0, 1);
    fm_stage_drives(c, stage, sa);
    for (int i = 0; i < 3; i++) {
        hext[i] = static_cast<float>(sa.hext[i]);
        hdr[i] = s->hdrive ? s->hdrive + i * c->n : nullptr;
    }
    drive_s = static_cast<float>(sa.drive_s);
    for (int i = 0; i < 6; i++) str[i] = s->strain ? s->strain + i * c->n : nullptr;
}

gf_status stage32(gf_ctx* c, Adj32* s, int stage, float* const* base, float* const* in, float* const* out,
                  float* const* h) {
    S32Args a{};
    a.code = c->d_code;
    for (int i = 0; i < 3; i++) {
        a.store[i] = in[i]; a.base[i] = base[i]; a.acc[i] = s->acc + i * c->n; a.out[i] = out[i]; a.hd[i] = h[i];
    }
    fill_common(c, s, stage, a.hext, a.drive_s, a.hdrive, a.strain);
    a.nx = static_cast<int>(c->nx); a.ny = static_cast<int>(c->ny); a.nz = static_cast<int>(c->nz);
    a.ix = static_cast<float>(1. / c->dx); a.iy = static_cast<float>(1. / c->dy); a.iz = static_cast<float>(1. / c->dz);
    a.m = s->co;
    a.stage = stage;
    hipLaunchKernelGGL(stage32_kernel, dim3(blocks_of(c->n)), dim3(kT), 0, c->stream, a);
    return hip_check(c, hipGetLastError(), "FP32 stage");
}

// One RK4 step from y0 (the field of y0 in h[0]); stage stores st[0..2],
// the new state y1; fields h[1..4] of the stores and of y1 (h[4] may be
// null).  Without magnetostatics every h[s] is the static field.
gf_status step32(gf_ctx* c, Adj32* s, float* const* y0, float* const* const st[3], float* const* y1,
                 float* const* const h[5]) {
    gf_status e = gf_time_advance(c);
    if (e) return e;
    float* const* ins[4] = {y0, st[0], st[1], st[2]};
    float* const* outs[4] = {st[0], st[1], st[2], y1};
    for (int q = 1; q <= 4; q++) {
        if ((e = stage32(c, s, q, y0, ins[q - 1], outs[q - 1], h[q - 1]))) return e;
        if (s->ms && h[q] && (e = conv32(c, s, outs[q - 1], h[q], false))) return e;
    }
    return GF_OK;
}

gf_status sum32(gf_ctx* c, Adj32* s, const float* f0, const float* f1, const float* f2, double out[3]) {
    hipLaunchKernelGGL(sum3_32_kernel, dim3(kSumBlocks), dim3(kT), 0, c->stream, f0, f1, f2, c->n, s->partial);
    std::vector<double> part(3 * kSumBlocks);
    gf_status st = hip_check(c, hipGetLastError(), "FP32 gradient sum");
    if (!st) st = hip_check(c, hipMemcpyAsync(part.data(), s->partial, sizeof(double) * part.size(),
                                              hipMemcpyDeviceToHost, c->stream), "gradient copy");
    if (!st) st = hip_check(c, hipStreamSynchronize(c->stream), "gradient sync");
    if (st) return st;
    for (int a = 0; a < 3; a++) {
        double t = 0.;
        for (int b = 0; b < kSumBlocks; b++) t += part[static_cast<size_t>(3 * b + a)];
        out[a] = t;
    }
    return GF_OK;
}

void free32(Adj32* s) {
    if (!s) return;
    float* fb[] = {s->m, s->sa, s->sb, s->acc, s->hd, s->hdrive, s->strain, s->pool, s->X, s->Y, s->HD, s->work,
                   s->gpar};
    for (float* p : fb) hipFree(p);
    float2* cb[] = {s->twx, s->twy, s->twz, s->kt, s->spa, s->spb};
    for (float2* p : cb) hipFree(p);
    hipFree(s->partial);
    delete s;
}

}  // namespace

void adjoint32_free(gf_ctx* c) {
    free32(c->adj32);
    c->adj32 = nullptr;
}

int adjoint32_material(const gf_ctx* c) { return c->adj32 ? c->adj32->material : 0; }

gf_status adjoint32_forward(gf_ctx* c, const gf_adjoint_desc* d, double* loss) {
    if (c->flags & GF_FLAG_ORIENT_ROTATE)
        return set_error(c, GF_ERR_UNSUPPORTED, "the FP32 adjoint sweep does not support frame rotation");
    if (c->has_hnu)
        return set_error(c, GF_ERR_UNSUPPORTED, "the FP32 adjoint sweep does not support Hext_nonunif");
    int fm_mats = 0;
    for (const auto& m : c->host_mats) fm_mats += m.if_FM ? 1 : 0;
    if (fm_mats != 1)
        return set_error(c, GF_ERR_UNSUPPORTED, "the FP32 adjoint sweep supports one FM material (the FP64 sweep several)");
    for (const auto& m : c->host_mats)
        if (m.if_FM && m.iDMI != 0.)
            return set_error(c, GF_ERR_UNSUPPORTED, "the FP32 adjoint sweep does not support DMI (the FP64 sweep does)");
    if ((c->flags & GF_FLAG_MAGNETOSTATICS) && c->demag.ready && c->demag.general)
        return set_error(c, GF_ERR_UNSUPPORTED,
                         "the FP32 adjoint sweep needs power-of-two nx, ny, nz with FFT magnetostatics");
    adjoint32_free(c);
    Adj32* s = new Adj32();
    c->adj32 = s;
    s->d = *d;
    s->material = d->material;
    s->N = d->nsteps;
    s->K = d->checkpoint_every > 0 ? d->checkpoint_every
                                   : static_cast<int64_t>(std::ceil(std::sqrt(static_cast<double>(d->nsteps))));
    s->K = std::min(s->K, s->N);
    s->nseg = (s->N + s->K - 1) / s->K;
    s->ms = (c->flags & GF_FLAG_MAGNETOSTATICS) && c->demag.ready;
    s->t0 = c->time;
    const int64_t n = c->n;
    // material coefficients rounded to FP32
    DevMaterial dm{};
    gf_status st = hip_check(c, hipMemcpy(&dm, c->d_mats + (d->material - 1), sizeof(DevMaterial),
                                          hipMemcpyDeviceToHost), "material copy");
    if (st) { adjoint32_free(c); return st; }
    s->co = Coef32{static_cast<float>(dm.Ms), static_cast<float>(dm.K1n), static_cast<float>(dm.K2n),
                   static_cast<float>(dm.B1n), static_cast<float>(dm.B2n), static_cast<float>(dm.Aex_self),
                   static_cast<float>(dm.exch), static_cast<float>(dm.precess), static_cast<float>(dm.alpha_precess),
                   dm.anis_type};
    const gf_material& hm = c->host_mats[static_cast<size_t>(d->material - 1)];
    s->alpha = static_cast<float>(hm.FM_damping);
    s->aex = static_cast<float>(hm.Aex);
    // state, fields and the inputs in FP32
    if (!st) st = alloc(c, &s->m, 3 * n, "hipMalloc FP32 m");
    if (!st) st = alloc(c, &s->sa, 3 * n, "hipMalloc FP32 store");
    if (!st) st = alloc(c, &s->sb, 3 * n, "hipMalloc FP32 store");
    if (!st) st = alloc(c, &s->acc, 3 * n, "hipMalloc FP32 accumulator");
    if (!st) st = alloc(c, &s->hd, 3 * n, "hipMalloc FP32 field");
    if (!st) st = alloc(c, &s->pool, (s->nseg + 1) * 3 * n, "hipMalloc FP32 checkpoints");
    if (!st) st = alloc(c, &s->X, s->N, "hipMalloc FP32 probe means");
    if (!st) st = alloc(c, &s->partial, 3 * kSumBlocks, "hipMalloc FP32 partial sums");
    for (int a = 0; a < 3 && !st; a++) st = to_f32(c, c->m[a], s->m + a * n, n);
    if (!st) st = hip_check(c, hipMemsetAsync(s->sa, 0, sizeof(float) * 3 * n, c->stream), "clear");
    if (!st) st = hip_check(c, hipMemsetAsync(s->sb, 0, sizeof(float) * 3 * n, c->stream), "clear");
    if (!st) st = hip_check(c, hipMemsetAsync(s->acc, 0, sizeof(float) * 3 * n, c->stream), "clear");
    StageArgs sa = fm_stage_args(c, 1, 0, 1);
    if (!st && sa.hdrive[0]) {
        st = alloc(c, &s->hdrive, 3 * n, "hipMalloc FP32 drive");
        for (int a = 0; a < 3 && !st; a++) st = to_f32(c, sa.hdrive[a], s->hdrive + a * n, n);
    }
    if (!st && sa.strain[0]) {
        st = alloc(c, &s->strain, 6 * n, "hipMalloc FP32 strain");
        for (int a = 0; a < 6 && !st; a++) st = to_f32(c, sa.strain[a], s->strain + a * n, n);
    }
    if (!st && s->ms) {
        s->nzh = c->nz / 2 + 1;
        s->nh = c->nx * c->ny * s->nzh;
        s->pz = pencils_for(c->nz, 2048, 16, c->ny);
        s->py = pencils_for(c->ny, 2048, 16, c->nx);
        st = twiddles32(c, &s->twx, c->nx);
        if (!st) st = twiddles32(c, &s->twy, c->ny);
        if (!st) st = twiddles32(c, &s->twz, c->nz);
        if (!st) st = alloc(c, &s->kt, 9 * s->nh, "hipMalloc FP32 tensor");
        if (!st) st = alloc(c, &s->spa, 3 * s->nh, "hipMalloc FP32 spectra");
        if (!st) st = alloc(c, &s->spb, 3 * s->nh, "hipMalloc FP32 spectra");
        if (!st) {
            hipLaunchKernelGGL(tensor32_kernel, dim3(blocks_of(9 * s->nh)), dim3(kT), 0, c->stream, c->demag.kre,
                               c->demag.kim, s->kt, 9 * s->nh);
            st = hip_check(c, hipGetLastError(), "FP32 tensor");
        }
    } else if (!st) {
        // the static field (uploaded, or zero) acts in every stage
        for (int a = 0; a < 3 && !st; a++) st = to_f32(c, c->hstat[a], s->hd + a * n, n);
    }
    if (st) { adjoint32_free(c); return st; }
    s->ck_time.assign(static_cast<size_t>(s->nseg), 0.);
    s->tstep.assign(static_cast<size_t>(s->N), 0.);
    float *m3[3], *a3[3], *b3[3], *h3[3];
    slot3f(s->m, 0, n, m3); slot3f(s->sa, 0, n, a3); slot3f(s->sb, 0, n, b3); slot3f(s->hd, 0, n, h3);
    float* const* h_all[5] = {h3, h3, h3, h3, h3};
    float* const* st3[3] = {a3, b3, a3};
    if (s->ms) st = conv32(c, s, m3, h3, false);
    for (int64_t k = 0; k < s->N && !st; k++) {
        if (k % s->K == 0) {
            float* ck[3];
            slot3f(s->pool, k / s->K, n, ck);
            st = copy3f(c, ck, m3);
            s->ck_time[static_cast<size_t>(k / s->K)] = c->time;
        }
        if (!st) st = step32(c, s, m3, st3, m3, h_all);
        if (!st) {
            hipLaunchKernelGGL(probe32_kernel, dim3(1), dim3(kT), 0, c->stream, m3[0], static_cast<int>(c->ny),
                               static_cast<int>(c->nz), d->probe_lo[0], d->probe_lo[1], d->probe_lo[2],
                               d->probe_hi[0], d->probe_hi[1], d->probe_hi[2], s->X + k);
            st = hip_check(c, hipGetLastError(), "FP32 probe mean");
        }
        s->tstep[static_cast<size_t>(k)] = c->time;
    }
    if (!st) {
        float* fin[3];
        slot3f(s->pool, s->nseg, n, fin);
        st = copy3f(c, fin, m3);
    }
    std::vector<float> X(static_cast<size_t>(s->N));
    if (!st) st = hip_check(c, hipMemcpyAsync(X.data(), s->X, sizeof(float) * X.size(), hipMemcpyDeviceToHost,
                                              c->stream), "probe copy");
    if (!st) st = hip_check(c, hipStreamSynchronize(c->stream), "FP32 forward sync");
    c->time = s->t0;
    if (st) { adjoint32_free(c); return st; }
    const double w = 2. * kPi * d->probe_freq;
    std::complex<double> S(0., 0.);
    for (int64_t k = 0; k < s->N; k++)
        S += static_cast<double>(X[static_cast<size_t>(k)]) * std::polar(c->dt, -w * s->tstep[static_cast<size_t>(k)]);
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

gf_status adjoint32_backward(gf_ctx* c, double grad[3], double* grad_hdrive, size_t count) {
    Adj32* s = c->adj32;
    if (!s || !s->ready) return set_error(c, GF_ERR_STATE, "gf_adjoint_backward needs a gf_adjoint_forward first");
    if (grad_hdrive && !s->hdrive) return set_error(c, GF_ERR_ARG, "grad_hdrive needs the shaped drive (GF_FLAG_HDRIVE)");
    if (grad_hdrive && count != static_cast<size_t>(3 * c->n)) return set_error(c, GF_ERR_ARG, "count must be 3n");
    const int64_t n = c->n, K = s->K;
    const double t_save = c->time;
    // The loss gradients dL/dX_n are of the size of the loss over the probe
    // signal (~1e-30 for the film workloads), so FP32 cotangents times the
    // LLG prefactors (~1e-9) would fall into the subnormal range.  The seeds
    // are scaled by a power of two to order one (exact in FP32) and the
    // gradients unscaled in FP64.
    double gmax = 0.;
    for (double v : s->g) gmax = std::max(gmax, std::fabs(v));
    const int gexp = gmax > 0. ? std::ilogb(gmax) : 0;
    const double sig = std::ldexp(1.0, -gexp);
    gf_status st = GF_OK;
    if (!s->work) st = alloc(c, &s->work, 21 * n, "hipMalloc FP32 adjoint work");
    if (!st && !s->gpar) st = alloc(c, &s->gpar, 3 * n, "hipMalloc FP32 gradients");
    if (!st && s->kbuf != K) {
        st = alloc(c, &s->Y, (4 * K + 1) * 3 * n, "hipMalloc FP32 stage stores");
        if (!st && s->ms) st = alloc(c, &s->HD, 4 * K * 3 * n, "hipMalloc FP32 stage fields");
        s->kbuf = st ? 0 : K;
    }
    if (st) return st;
    float *lam[3], *acc[3], *kbar[3], *mbar[3], *hbar[3], *dadj[3], *ghd[3];
    slot3f(s->work, 0, n, lam); slot3f(s->work, 1, n, acc); slot3f(s->work, 2, n, kbar);
    slot3f(s->work, 3, n, mbar); slot3f(s->work, 4, n, hbar); slot3f(s->work, 5, n, dadj); slot3f(s->work, 6, n, ghd);
    st = hip_check(c, hipMemsetAsync(s->work, 0, sizeof(float) * 21 * static_cast<size_t>(n), c->stream), "clear");
    if (!st) st = hip_check(c, hipMemsetAsync(s->gpar, 0, sizeof(float) * 3 * static_cast<size_t>(n), c->stream), "clear");
    const int* lo = s->d.probe_lo;
    const int* hi = s->d.probe_hi;
    const int nx = static_cast<int>(c->nx), ny = static_cast<int>(c->ny), nz = static_cast<int>(c->nz);
    if (!st) {
        hipLaunchKernelGGL(adj_init32_kernel, dim3(blocks_of(n)), dim3(kT), 0, c->stream, c->d_code, lam[0], kbar[0],
                           nx, ny, nz, lo[0], lo[1], lo[2], hi[0], hi[1], hi[2],
                           static_cast<float>(s->g.back() * sig));
        st = hip_check(c, hipGetLastError(), "FP32 adjoint init");
    }
    float* h3[3];
    slot3f(s->hd, 0, n, h3);       // the static field when magnetostatics is off
    std::vector<double> tseg(static_cast<size_t>(K));
    for (int64_t seg = s->nseg - 1; seg >= 0 && !st; seg--) {
        const int64_t n0 = seg * K, n1 = std::min(s->N, n0 + K), J = n1 - n0;
        float *y0[3], *ck[3];
        slot3f(s->Y, 0, n, y0);
        slot3f(s->pool, seg, n, ck);
        st = copy3f(c, y0, ck);
        c->time = s->ck_time[static_cast<size_t>(seg)];
        if (!st && s->ms) {
            float* h0[3];
            slot3f(s->HD, 0, n, h0);
            st = conv32(c, s, y0, h0, false);
        }
        for (int64_t j = 0; j < J && !st; j++) {
            float *ys[4][3], *yn[3], *hs[5][3];
            for (int q = 0; q < 4; q++) slot3f(s->Y, 4 * j + q, n, ys[q]);
            slot3f(s->Y, 4 * (j + 1), n, yn);
            for (int q = 0; q < 4; q++) {
                if (s->ms) slot3f(s->HD, 4 * j + q, n, hs[q]);
                else for (int a = 0; a < 3; a++) hs[q][a] = h3[a];
            }
            const bool more = j + 1 < J;
            if (s->ms && more) slot3f(s->HD, 4 * (j + 1), n, hs[4]);
            float* const* sts[3] = {ys[1], ys[2], ys[3]};
            float* const* hp[5] = {hs[0], hs[1], hs[2], hs[3], (s->ms && more) ? hs[4] : nullptr};
            st = step32(c, s, ys[0], sts, yn, hp);
            tseg[static_cast<size_t>(j)] = c->time;
        }
        for (int64_t j = J - 1; j >= 0 && !st; j--) {
            c->time = tseg[static_cast<size_t>(j)];
            const int64_t step = n0 + j;
            for (int stg = 4; stg >= 1 && !st; stg--) {
                float *ys[3], *hs[3];
                slot3f(s->Y, 4 * j + (stg - 1), n, ys);
                if (s->ms) slot3f(s->HD, 4 * j + (stg - 1), n, hs);
                else for (int a = 0; a < 3; a++) hs[a] = h3[a];
                P32Args p{};
                p.code = c->d_code;
                for (int i = 0; i < 3; i++) {
                    p.y[i] = ys[i]; p.hd[i] = hs[i]; p.kbar[i] = kbar[i]; p.mbar[i] = mbar[i]; p.hbar[i] = hbar[i];
                    p.ghdrive[i] = s->hdrive ? ghd[i] : nullptr;
                }
                fill_common(c, s, stg, p.hext, p.drive_s, p.hdrive, p.strain);
                p.galpha = s->gpar; p.gaex = s->gpar + n; p.gb1 = s->gpar + 2 * n;
                p.nx = nx; p.ny = ny; p.nz = nz;
                p.ix = static_cast<float>(1. / c->dx); p.iy = static_cast<float>(1. / c->dy);
                p.iz = static_cast<float>(1. / c->dz);
                p.m = s->co; p.alpha = s->alpha; p.aex = s->aex;
                hipLaunchKernelGGL(adj_point32_kernel, dim3(blocks_of(n)), dim3(kT), 0, c->stream, p);
                st = hip_check(c, hipGetLastError(), "FP32 adjoint point");
                if (!st && s->ms) st = conv32(c, s, hbar, dadj, true);
                if (st) break;
                C32Args a{};
                a.code = c->d_code;
                for (int i = 0; i < 3; i++) {
                    a.mbar[i] = mbar[i]; a.hbar[i] = hbar[i]; a.dadj[i] = s->ms ? dadj[i] : nullptr;
                    a.lam[i] = lam[i]; a.acc[i] = acc[i]; a.kbar[i] = kbar[i]; a.lo[i] = lo[i]; a.hi[i] = hi[i];
                }
                a.nx = nx; a.ny = ny; a.nz = nz;
                a.ix = p.ix; a.iy = p.iy; a.iz = p.iz;
                a.Aself = s->co.Aself; a.exch = s->co.exch; a.Ms = s->co.Ms;
                a.mode = stg;
                a.g = (stg == 1 && step >= 1) ? static_cast<float>(s->g[static_cast<size_t>(step - 1)] * sig) : 0.f;
                hipLaunchKernelGGL(adj_combine32_kernel, dim3(blocks_of(n)), dim3(kT), 0, c->stream, a);
                st = hip_check(c, hipGetLastError(), "FP32 adjoint combine");
            }
        }
    }
    c->time = t_save;
    grad[0] = grad[1] = grad[2] = 0.;
    if (!st) st = sum32(c, s, s->gpar, s->gpar + n, s->gpar + 2 * n, grad);
    for (int a = 0; a < 3; a++) grad[a] /= sig;
    if (!st && grad_hdrive) {
        std::vector<float> gh(static_cast<size_t>(3 * n));
        st = hip_check(c, hipMemcpyAsync(gh.data(), ghd[0], sizeof(float) * gh.size(), hipMemcpyDeviceToHost,
                                         c->stream), "gradient copy");
        if (!st) st = hip_check(c, hipStreamSynchronize(c->stream), "FP32 backward sync");
        if (!st)
            for (size_t i = 0; i < gh.size(); i++) grad_hdrive[i] = static_cast<double>(gh[i]) / sig;
    }
    if (!st) st = hip_check(c, hipStreamSynchronize(c->stream), "FP32 backward sync");
    return st;
}

}  // namespace gf
