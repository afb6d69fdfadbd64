// gf_fft.hip — FP64 FFT magnetostatic convolution for gfx950.
//
// Replaces magnetic_system::get_H_static (magnetostatic_solver.cpp:114-219):
// M = Ms m (FM) or Ms1 m1 + Ms2 m2 (AFM) -> 3 real-to-complex 3D transforms
// -> H_k = (sum of three M_k * N_k) / (-dx dy dz) with the nine tensor
// spectra -> 3 complex-to-real transforms.  Transforms are unnormalised; the
// kernel is pre-divided by n (as cuFFT D2Z/Z2D with the reference's kernel).
//
// Five launches, each a set of 1D transforms held in LDS, each reading and
// writing global memory in runs of >= 16 consecutive complex values:
//   zfwd   [x][y][z] reals      -> z transform -> [x][kz][y]     (spec_a)
//   yfwd   [x][kz][y]           -> y transform -> [kz][y][x]     (spec_b)
//   xconv  [kz][y][x]           -> x fwd, tensor product, x inv  (spec_b, in place)
//   yinv   [kz][y][x]           -> y inverse   -> [x][kz][y]     (spec_a)
//   zinv   [x][kz][y]           -> z inverse (Hermitian) -> H [x][y][z]
// Each layout puts the next transform's axis or its store run innermost,
// so the transposes ride on the LDS staging the transforms need anyway.
// Transforms are radix-2: forward decimation in time (bit-reversed load,
// natural result), inverse decimation in frequency (natural load,
// bit-reversed result read back at the store), twiddles staged in LDS.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <thread>
#include <vector>

#include "gf_fft_dev.h"

namespace gf {

namespace {

constexpr int kThreads = 256;
// Global loads per thread issued together when a pass fills its LDS buffer.
constexpr int kBatch = 8;

__device__ __forceinline__ int bitrev(int j, int logn) { return fft_bitrev(j, logn); }
__device__ __forceinline__ double2 cmul(double2 a, double2 b) { return fft_cmul(a, b); }
__device__ __forceinline__ void load_twiddles(double2* dst, const double2* src, int n) { fft_load_twiddles(dst, src, n); }

struct ZArgs {
    const uint32_t* code;
    const DevMaterial* mats;
    const double* fm[3];
    const double* a1[3];
    const double* a2[3];
    const double* raw[3];       // kRaw input, component c at raw[c]
    double2* out;               // component c at out + c*nh, layout [x][kz][y]
    const double2* tw;
    int64_t nh;
    int nx, ny, nz, nzh;
    int logny, lognz, pencils, logp;
};

// Magnetisation assembly (magnetostatic_solver.cpp:121-145) fused with the
// z real-to-complex transform.  kRaw = true transforms plain real arrays
// (the tensor kernels at set-up, the adjoint field of the adjoint sweep).
template <bool kRaw>
__global__ void __launch_bounds__(kThreads) zfwd_kernel(ZArgs a) {
    extern __shared__ double2 smem[];
    const int c = blockIdx.z;
    const int x = blockIdx.y;
    const int y0 = blockIdx.x * a.pencils;
    const double* fmc = c == 0 ? a.fm[0] : (c == 1 ? a.fm[1] : a.fm[2]);
    const double* a1c = c == 0 ? a.a1[0] : (c == 1 ? a.a1[1] : a.a1[2]);
    const double* a2c = c == 0 ? a.a2[0] : (c == 1 ? a.a2[1] : a.a2[2]);
    double2* tw = smem;
    double2* buf = smem + (a.nz >> 1);
    load_twiddles(tw, a.tw, a.nz);
    const int total = a.pencils * a.nz;
    for (int idx = threadIdx.x; idx < total; idx += blockDim.x) {
        const int r = idx >> a.lognz;
        const int k = idx & (a.nz - 1);
        const int y = y0 + r;
        double v = 0.;
        if (y < a.ny) {
            const int64_t cell = (static_cast<int64_t>(x) * a.ny + y) * a.nz + k;
            if (kRaw) {
                v = (c == 0 ? a.raw[0] : (c == 1 ? a.raw[1] : a.raw[2]))[cell];
            } else {
                const uint32_t code = a.code[cell];
                if (code & kIsFM) {
                    v = a.mats[(code & kMatMask) - 1].Ms * fmc[cell];
                } else if (code & kIsAFM) {
                    const DevMaterial& M = a.mats[(code & kMatMask) - 1];
                    v = M.aMs[0] * a1c[cell] + M.aMs[1] * a2c[cell];
                }
            }
        }
        buf[fpad(r * a.nz + bitrev(k, a.lognz))] = make_double2(v, 0.);
    }
    __syncthreads();
    fft_dit(buf, a.pencils, a.lognz, tw);
    double2* out = a.out + c * a.nh;
    const int ztotal = a.pencils * a.nzh;
    for (int idx = threadIdx.x; idx < ztotal; idx += blockDim.x) {
        const int r = idx & (a.pencils - 1);
        const int kz = idx >> a.logp;
        const int y = y0 + r;
        if (y < a.ny) out[(static_cast<int64_t>(x) * a.nzh + kz) * a.ny + y] = buf[fpad(r * a.nz + kz)];
    }
}

// The y and x passes of a convolution may run over a range of kz planes
// at a time (kz = kz0 + blockIdx.y); the [kz][y][x] buffer between them
// then holds plane kz at kz - kzb0, component c at c*nhb, so that a chunk
// of planes can reuse the start of the buffer while it is still in cache.
struct YArgs {
    const double2* in;
    double2* out;
    const double2* tw;
    int64_t nh, nhb;
    int nx, ny, nzh;
    int logny, pencils, logp;
    int kz0, kzb0;
};

// y forward transform: [x][kz][y] -> [kz][y][x]
__global__ void __launch_bounds__(kThreads) yfwd_kernel(YArgs a) {
    extern __shared__ double2 smem[];
    const int c = blockIdx.z;
    const int kz = a.kz0 + static_cast<int>(blockIdx.y);
    const int x0 = blockIdx.x * a.pencils;
    double2* tw = smem;
    double2* buf = smem + (a.ny >> 1);
    const double2 tw0 = fft_twiddle_prefetch(a.tw, a.ny);
    const double2* in = a.in + c * a.nh;
    const int total = a.pencils * a.ny;
    fft_staged_load<kBatch, double2>(
        total,
        [&](int idx) {
            const int x = x0 + (idx >> a.logny);
            return x < a.nx ? in[(static_cast<int64_t>(x) * a.nzh + kz) * a.ny + (idx & (a.ny - 1))]
                            : make_double2(0., 0.);
        },
        [&](int idx, double2 v) { buf[fpad((idx >> a.logny) * a.ny + bitrev(idx & (a.ny - 1), a.logny))] = v; });
    fft_twiddle_store(tw, a.tw, a.ny, tw0);
    __syncthreads();
    fft_dit(buf, a.pencils, a.logny, tw);
    double2* out = a.out + c * a.nhb;
    const int kb = kz - a.kzb0;
    for (int idx = threadIdx.x; idx < total; idx += blockDim.x) {
        const int r = idx & (a.pencils - 1);
        const int y = idx >> a.logp;
        const int x = x0 + r;
        if (x < a.nx) out[(static_cast<int64_t>(kb) * a.ny + y) * a.nx + x] = buf[fpad(r * a.ny + y)];
    }
}

// y inverse transform: [kz][y][x] -> [x][kz][y]
__global__ void __launch_bounds__(kThreads) yinv_kernel(YArgs a) {
    extern __shared__ double2 smem[];
    const int c = blockIdx.z;
    const int kz = a.kz0 + static_cast<int>(blockIdx.y);
    const int kb = kz - a.kzb0;
    const int x0 = blockIdx.x * a.pencils;
    double2* tw = smem;
    double2* buf = smem + (a.ny >> 1);
    const double2 tw0 = fft_twiddle_prefetch(a.tw, a.ny);
    const double2* in = a.in + c * a.nhb;
    const int total = a.pencils * a.ny;
    fft_staged_load<kBatch, double2>(
        total,
        [&](int idx) {
            const int x = x0 + (idx & (a.pencils - 1));
            return x < a.nx ? in[(static_cast<int64_t>(kb) * a.ny + (idx >> a.logp)) * a.nx + x] : make_double2(0., 0.);
        },
        [&](int idx, double2 v) { buf[fpad((idx & (a.pencils - 1)) * a.ny + (idx >> a.logp))] = v; });
    fft_twiddle_store(tw, a.tw, a.ny, tw0);
    __syncthreads();
    ifft_dif(buf, a.pencils, a.logny, tw);
    double2* out = a.out + c * a.nh;
    for (int idx = threadIdx.x; idx < total; idx += blockDim.x) {
        const int r = idx >> a.logny;
        const int y = idx & (a.ny - 1);
        const int x = x0 + r;
        if (x < a.nx) out[(static_cast<int64_t>(x) * a.nzh + kz) * a.ny + y] = buf[fpad(r * a.ny + bitrev(y, a.logny))];
    }
}

struct XArgs {
    double2* data;              // component c at data + c*nhb, plane kz at kz - kzb0, layout [kz][y][x]
    const double* kre;          // real part of tensor spectrum s at kre + s*nh
    const float* kim;           // its imaginary part (round-off of an even/odd tensor) at kim + s*nh
    const double2* tw;
    int64_t nh, nhb;
    int kz0, kzb0;
    double den;                 // -dx*dy*dz
    int nx, ny, nzh;
    int lognx, pencils, logp;
    int adjoint;                // 1: conjugate transpose of the tensor (adjoint convolution)
};

// x forward transform, tensor product and x inverse transform, in place.
// kConv = false: forward transform of component 0 only (tensor set-up).
template <bool kConv>
__global__ void __launch_bounds__(kThreads) xconv_kernel(XArgs a) {
    extern __shared__ double2 smem[];
    const int kz = a.kz0 + static_cast<int>(blockIdx.y);
    const int kb = kz - a.kzb0;
    const int y0 = blockIdx.x * a.pencils;
    const int comps = kConv ? 3 : 1;
    double2* tw = smem;
    double2* buf = smem + (a.nx >> 1);
    const double2 tw0 = fft_twiddle_prefetch(a.tw, a.nx);
    const int total = comps * a.pencils * a.nx;
    const int pts = a.pencils * a.nx;
    const int64_t s = a.nh;
    // Tensor spectra of one k-point in the order xx, yy, zz, xy, yz, zx, yx,
    // zy, xz; the adjoint convolution uses N^T with conjugated spectra:
    // xy<->yx, yz<->zy, zx<->xz (the real-space correlation with the kernel).
    auto load_tensor = [&](int64_t g, double (&kr)[9], double (&ki)[9]) {
#pragma unroll
        for (int q = 0; q < 9; q++) {
            const int t = a.adjoint ? (q < 3 ? q : (q < 6 ? q + 3 : q - 3)) : q;
            // the tensor is streamed once per convolution
            kr[q] = __builtin_nontemporal_load(&a.kre[t * s + g]);
            const double im = static_cast<double>(__builtin_nontemporal_load(&a.kim[t * s + g]));
            ki[q] = a.adjoint ? -im : im;
        }
    };
    auto point_of = [&](int idx) {
        return (static_cast<int64_t>(kz) * a.ny + y0 + (idx >> a.lognx)) * a.nx + (idx & (a.nx - 1));
    };
    // this thread's first k-point's tensor is loaded with the magnetisation
    // spectra, so the product does not wait for a second round trip
    double kr0[9], ki0[9];
    const int first = threadIdx.x;
    const bool first_ok = kConv && first < pts && y0 + (first >> a.lognx) < a.ny;
    if (first_ok) load_tensor(point_of(first), kr0, ki0);
    fft_staged_load<kBatch, double2>(
        total,
        [&](int idx) {
            const int t = idx >> a.lognx;
            const int y = y0 + (t & (a.pencils - 1));
            const int64_t g = (static_cast<int64_t>(kb) * a.ny + y) * a.nx + (idx & (a.nx - 1));
            return y < a.ny ? a.data[(t >> a.logp) * a.nhb + g] : make_double2(0., 0.);
        },
        [&](int idx, double2 v) { buf[fpad((idx >> a.lognx) * a.nx + bitrev(idx & (a.nx - 1), a.lognx))] = v; });
    fft_twiddle_store(tw, a.tw, a.nx, tw0);
    __syncthreads();
    fft_dit<4>(buf, comps * a.pencils, a.lognx, tw);
    if (kConv) {
        // magnetostatic_solver.cpp:163-177
        auto product = [&](int idx, const double (&kr)[9], const double (&ki)[9]) {
            const int x = idx & (a.nx - 1);
            const int r = idx >> a.lognx;
            const int ix = (0 * a.pencils + r) * a.nx + x;
            const int iy = (1 * a.pencils + r) * a.nx + x;
            const int iz = (2 * a.pencils + r) * a.nx + x;
            const double2 mx = buf[fpad(ix)], my = buf[fpad(iy)], mz = buf[fpad(iz)];
            auto K = [&](int q) { return make_double2(kr[q], ki[q]); };
            double2 t0 = cmul(mx, K(0)), t1 = cmul(my, K(3)), t2 = cmul(mz, K(8));
            const double2 hx = make_double2((t0.x + t1.x + t2.x) / a.den, (t0.y + t1.y + t2.y) / a.den);
            t0 = cmul(my, K(1)); t1 = cmul(mz, K(4)); t2 = cmul(mx, K(6));
            const double2 hy = make_double2((t0.x + t1.x + t2.x) / a.den, (t0.y + t1.y + t2.y) / a.den);
            t0 = cmul(mz, K(2)); t1 = cmul(mx, K(5)); t2 = cmul(my, K(7));
            const double2 hz = make_double2((t0.x + t1.x + t2.x) / a.den, (t0.y + t1.y + t2.y) / a.den);
            buf[fpad(ix)] = hx; buf[fpad(iy)] = hy; buf[fpad(iz)] = hz;
        };
        if (first_ok) product(first, kr0, ki0);
        for (int idx = first + static_cast<int>(blockDim.x); idx < pts; idx += blockDim.x) {
            if (y0 + (idx >> a.lognx) >= a.ny) continue;
            double kr[9], ki[9];
            load_tensor(point_of(idx), kr, ki);
            product(idx, kr, ki);
        }
        __syncthreads();
        ifft_dif<4>(buf, comps * a.pencils, a.lognx, tw);
    }
    for (int idx = threadIdx.x; idx < total; idx += blockDim.x) {
        const int x = idx & (a.nx - 1);
        const int t = idx >> a.lognx;
        const int r = t & (a.pencils - 1);
        const int c = t >> a.logp;
        const int y = y0 + r;
        if (y < a.ny) {
            const int src = kConv ? bitrev(x, a.lognx) : x;
            a.data[c * a.nhb + (static_cast<int64_t>(kb) * a.ny + y) * a.nx + x] = buf[fpad(t * a.nx + src)];
        }
    }
}

struct ZInvArgs {
    const double2* in;          // layout [x][kz][y]
    double* out[3];
    const double2* tw;
    int64_t nh;
    int nx, ny, nz, nzh;
    int lognz, pencils, logp;
};

// z complex-to-real inverse: Hermitian extension of the half spectrum, the
// imaginary parts of the self-conjugate bins left as they are (they only
// reach the discarded imaginary output).
__global__ void __launch_bounds__(kThreads) zinv_kernel(ZInvArgs a) {
    extern __shared__ double2 smem[];
    const int c = blockIdx.z;
    const int x = blockIdx.y;
    const int y0 = blockIdx.x * a.pencils;
    double2* tw = smem;
    double2* buf = smem + (a.nz >> 1);
    const double2 tw0 = fft_twiddle_prefetch(a.tw, a.nz);
    const double2* in = a.in + c * a.nh;
    const int ztotal = a.pencils * a.nzh;
    fft_staged_load<kBatch, double2>(
        ztotal,
        [&](int idx) {
            const int y = y0 + (idx & (a.pencils - 1));
            return y < a.ny ? in[(static_cast<int64_t>(x) * a.nzh + (idx >> a.logp)) * a.ny + y] : make_double2(0., 0.);
        },
        [&](int idx, double2 v) {
            const int r = idx & (a.pencils - 1);
            const int kz = idx >> a.logp;
            buf[fpad(r * a.nz + kz)] = v;
            const int mirror = a.nz - kz;
            if (kz >= 1 && mirror > kz && mirror < a.nz) buf[fpad(r * a.nz + mirror)] = make_double2(v.x, -v.y);
        });
    fft_twiddle_store(tw, a.tw, a.nz, tw0);
    __syncthreads();
    ifft_dif(buf, a.pencils, a.lognz, tw);
    double* out = c == 0 ? a.out[0] : (c == 1 ? a.out[1] : a.out[2]);
    const int total = a.pencils * a.nz;
    for (int idx = threadIdx.x; idx < total; idx += blockDim.x) {
        const int r = idx >> a.lognz;
        const int k = idx & (a.nz - 1);
        const int y = y0 + r;
        if (y < a.ny) out[(static_cast<int64_t>(x) * a.ny + y) * a.nz + k] = buf[fpad(r * a.nz + bitrev(k, a.lognz))].x;
    }
}

struct LocalArgs {
    const uint32_t* code;
    const DevMaterial* mats;
    const double* fm[3];
    const double* a1[3];
    const double* a2[3];
    double* h[3];
    double fac[3];
    int64_t n;
    int one_d;
};

// Demag-factor field (magnetostatic_solver.cpp:192-218) or the 1D model
// Hz = -Ms mz (:222-249); cells that are neither FM nor AFM keep H.
__global__ void __launch_bounds__(kThreads) demag_local_kernel(LocalArgs a) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const uint32_t code = a.code[i];
    if (code & kIsFM) {
        const double Ms = a.mats[(code & kMatMask) - 1].Ms;
        if (a.one_d) {
            a.h[0][i] = 0.; a.h[1][i] = 0.; a.h[2][i] = -Ms * a.fm[2][i];
        } else {
            for (int c = 0; c < 3; c++) a.h[c][i] = Ms * a.fm[c][i] * a.fac[c];
        }
    } else if (code & kIsAFM) {
        const DevMaterial& M = a.mats[(code & kMatMask) - 1];
        if (a.one_d) {
            a.h[0][i] = 0.; a.h[1][i] = 0.; a.h[2][i] = -M.aMs[0] * a.a1[2][i] - M.aMs[1] * a.a2[2][i];
        } else {
            for (int c = 0; c < 3; c++) a.h[c][i] = (M.aMs[0] * a.a1[c][i] + M.aMs[1] * a.a2[c][i]) * a.fac[c];
        }
    }
}

// Tensor spectra as an FP64 real part and an FP32 imaginary part: the
// tensor components are even or odd along every axis, so their spectra are
// real up to the round-off of the reference's kernel evaluation (~1e-9 of
// the peak); FP32 keeps that residue to ~1e-16 of the peak.
__global__ void __launch_bounds__(kThreads) split_spectrum_kernel(const double2* in, double* re, float* im, int64_t n) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double2 v = in[i];
    re[i] = v.x;
    im[i] = static_cast<float>(v.y);
}

int ilog2(int64_t n) {
    int l = 0;
    while ((int64_t{1} << l) < n) l++;
    return l;
}

int pick_pencils(int64_t len, int64_t budget, int64_t cap, int64_t other) {
    int64_t p = budget / len;
    if (p < 1) p = 1;
    if (p > cap) p = cap;
    int64_t pow = 1;
    while (pow * 2 <= p) pow *= 2;
    while (pow > 1 && pow > other) pow /= 2;
    return static_cast<int>(pow);
}

gf_status make_twiddles(gf_ctx* c, Twiddles& t, int64_t n) {
    t.n = n;
    const int64_t cnt = n / 2 > 0 ? n / 2 : 1;
    std::vector<double2> h(static_cast<size_t>(cnt));
    for (int64_t j = 0; j < cnt; j++) {
        // same angle expression as a radix-2 stage of span n
        const double ang = 2.0 * kPi * static_cast<double>(j) / static_cast<double>(n);
        h[static_cast<size_t>(j)] = make_double2(std::cos(ang), -std::sin(ang));
    }
    gf_status st = hip_check(c, hipMalloc(&t.dev, sizeof(double2) * cnt), "hipMalloc twiddles");
    if (st) return st;
    return hip_check(c, hipMemcpy(t.dev, h.data(), sizeof(double2) * cnt, hipMemcpyHostToDevice),
                     "copy twiddles");
}

struct Shapes {
    int pz, py, px;
    size_t lds_z, lds_y, lds_x;
};

// Pencils per workgroup.  z and y passes write their output transposed,
// so their pencil count sets the length of the store runs (16 B each);
// the x pass reads and writes whole rows, so it keeps one pencil per
// component and more workgroups per CU (at 256x256x128 one pencil beat two
// by 7% and eight y pencils matched four).  GOFERRO_FFT_PENCILS="z,y,x"
// overrides the defaults (powers of two) for tuning sweeps.
Shapes shapes(const gf_ctx* c) {
    Shapes s;
    int64_t bz = 2048, by = 2048, bx = 512, cz = 16, cy = 16, cx = 1;
    if (const char* env = std::getenv("GOFERRO_FFT_PENCILS")) {
        int z = 0, y = 0, x = 0;
        if (std::sscanf(env, "%d,%d,%d", &z, &y, &x) == 3 && z > 0 && y > 0 && x > 0) {
            bz = int64_t{z} * c->nz; by = int64_t{y} * c->ny; bx = int64_t{x} * c->nx;
            cz = z; cy = y; cx = x;
        }
    }
    s.pz = pick_pencils(c->nz, bz, cz, c->ny);
    s.py = pick_pencils(c->ny, by, cy, c->nx);
    s.px = pick_pencils(c->nx, bx, cx, c->ny);
    s.lds_z = sizeof(double2) * (c->nz / 2 + fft_lds_elems(s.pz * c->nz));
    s.lds_y = sizeof(double2) * (c->ny / 2 + fft_lds_elems(s.py * c->ny));
    s.lds_x = sizeof(double2) * (c->nx / 2 + fft_lds_elems(3 * s.px * c->nx));
    return s;
}

unsigned blocks_for(int64_t n, int p) { return static_cast<unsigned>((n + p - 1) / p); }

}  // namespace

// Twiddle table of n entries w[t] = exp(-2 pi i t / n) for the mixed-radix
// passes.
static gf_status make_table(gf_ctx* c, Twiddles& t, int64_t n) {
    t.n = n;
    std::vector<double2> h(static_cast<size_t>(n));
    for (int64_t j = 0; j < n; j++) {
        const double ang = 2.0 * kPi * static_cast<double>(j) / static_cast<double>(n);
        h[static_cast<size_t>(j)] = make_double2(std::cos(ang), -std::sin(ang));
    }
    gf_status st = hip_check(c, hipMalloc(&t.dev, sizeof(double2) * n), "hipMalloc twiddles");
    if (st) return st;
    return hip_check(c, hipMemcpy(t.dev, h.data(), sizeof(double2) * n, hipMemcpyHostToDevice), "copy twiddles");
}

// Radices of n from {4, 2, 3, 5, 7}, or false when n has another prime
// factor.
static bool factor(int64_t n, FftFactors& f) {
    f = FftFactors{};
    f.n = static_cast<int>(n);
    int64_t m = n;
    auto push = [&](int r) {
        if (f.count >= kMaxRadices) return false;
        f.r[f.count++] = r;
        m /= r;
        return true;
    };
    while (m % 4 == 0) if (!push(4)) return false;
    while (m % 2 == 0) if (!push(2)) return false;
    for (int r : {3, 5, 7})
        while (m % r == 0) if (!push(r)) return false;
    return m == 1;
}

// Pencils per workgroup of the mixed-radix passes: as many lines as keep
// the two LDS buffers within 2048 elements (z and y) or 1024 (x, three
// components), a power of two, at most 16.
static int any_pencils(int64_t len, int64_t budget, int64_t other) {
    int64_t p = 1;
    while (p * 2 <= 16 && p * 2 * len <= budget && p * 2 <= other) p *= 2;
    return static_cast<int>(p);
}

static AnyPassArgs any_args(const gf_ctx* c, const DemagPlan& d) {
    AnyPassArgs a{};
    a.code = c->d_code; a.mats = c->d_mats;
    a.spec = d.spec_m; a.spec2 = d.spec_b;
    a.kre = d.kre; a.kim = d.kim;
    a.nh = c->nx * c->ny * d.nzh;
    a.nhb = a.nh;
    a.den = -c->dx * c->dy * c->dz;
    a.nx = static_cast<int>(c->nx); a.ny = static_cast<int>(c->ny); a.nzh = static_cast<int>(d.nzh);
    return a;
}

// Forward pipeline helpers shared by the tensor set-up and the convolution.
static hipError_t run_zfwd(gf_ctx* c, DemagPlan& d, bool raw, const double* const* rawp, const double* const* fm,
                           const double* const* s1, const double* const* s2, int comps) {
    hipEvent_t ev = nullptr;
    timer_begin(c, GF_K_FFT_Z_FWD, &ev);
    if (d.general) {
        AnyPassArgs a = any_args(c, d);
        for (int i = 0; i < 3; i++) {
            a.fm[i] = fm ? fm[i] : nullptr;
            a.a1[i] = s1 ? s1[i] : nullptr;
            a.a2[i] = s2 ? s2[i] : nullptr;
            a.raw[i] = rawp ? rawp[i] : nullptr;
        }
        a.raw_in = raw ? 1 : 0;
        a.f = d.fz; a.tw = d.gtz.dev;
        a.pencils = any_pencils(c->nz, 2048, c->ny);
        dim3 grid(blocks_for(c->ny, a.pencils), static_cast<unsigned>(c->nx), static_cast<unsigned>(comps));
        const hipError_t e = launch_any_pass(AnyPass::kZFwd, a, grid, c->stream);
        timer_end(c, GF_K_FFT_Z_FWD, ev);
        return e;
    }
    const Shapes sh = shapes(c);
    ZArgs a{};
    a.code = c->d_code; a.mats = c->d_mats;
    for (int i = 0; i < 3; i++) {
        a.fm[i] = fm ? fm[i] : nullptr;
        a.a1[i] = s1 ? s1[i] : nullptr;
        a.a2[i] = s2 ? s2[i] : nullptr;
    }
    for (int i = 0; i < 3; i++) a.raw[i] = rawp ? rawp[i] : nullptr;
    a.out = d.spec_m; a.tw = d.twz.dev;
    a.nh = c->nx * c->ny * d.nzh;
    a.nx = static_cast<int>(c->nx); a.ny = static_cast<int>(c->ny); a.nz = static_cast<int>(c->nz);
    a.nzh = static_cast<int>(d.nzh);
    a.logny = ilog2(c->ny); a.lognz = ilog2(c->nz); a.pencils = sh.pz; a.logp = ilog2(sh.pz);
    dim3 grid(blocks_for(c->ny, sh.pz), static_cast<unsigned>(c->nx), static_cast<unsigned>(comps));
    if (raw) hipLaunchKernelGGL(zfwd_kernel<true>, grid, dim3(kThreads), sh.lds_z, c->stream, a);
    else hipLaunchKernelGGL(zfwd_kernel<false>, grid, dim3(kThreads), sh.lds_z, c->stream, a);
    timer_end(c, GF_K_FFT_Z_FWD, ev);
    return hipGetLastError();
}

// A range of kz planes of the y and x passes: planes k0 .. k0+cnt-1, and
// where the [kz][y][x] buffer holds them (plane kz at kz - kb0, component
// stride nhb).  The whole spectrum is {0, nzh, 0, nh}.
struct KzRange {
    int k0, cnt, kb0;
    int64_t nhb;
};

static KzRange whole_range(const gf_ctx* c, const DemagPlan& d) {
    return KzRange{0, static_cast<int>(d.nzh), 0, c->nx * c->ny * d.nzh};
}

static hipError_t run_y(gf_ctx* c, DemagPlan& d, bool inverse, int comps, KzRange kr) {
    hipEvent_t ev = nullptr;
    const int cls = inverse ? GF_K_FFT_Y_INV : GF_K_FFT_Y_FWD;
    timer_begin(c, cls, &ev);
    if (d.general) {
        AnyPassArgs a = any_args(c, d);
        a.spec = inverse ? d.spec_a : d.spec_m;
        a.nhb = kr.nhb; a.kz0 = kr.k0; a.kzb0 = kr.kb0;
        a.f = d.fy; a.tw = d.gty.dev;
        a.pencils = any_pencils(c->ny, 2048, c->nx);
        dim3 grid(blocks_for(c->nx, a.pencils), static_cast<unsigned>(kr.cnt), static_cast<unsigned>(comps));
        const hipError_t e = launch_any_pass(inverse ? AnyPass::kYInv : AnyPass::kYFwd, a, grid, c->stream);
        timer_end(c, cls, ev);
        return e;
    }
    const Shapes sh = shapes(c);
    YArgs a{};
    a.in = inverse ? d.spec_b : d.spec_m;
    a.out = inverse ? d.spec_a : d.spec_b;
    a.tw = d.twy.dev;
    a.nh = c->nx * c->ny * d.nzh;
    a.nhb = kr.nhb; a.kz0 = kr.k0; a.kzb0 = kr.kb0;
    a.nx = static_cast<int>(c->nx); a.ny = static_cast<int>(c->ny); a.nzh = static_cast<int>(d.nzh);
    a.logny = ilog2(c->ny); a.pencils = sh.py; a.logp = ilog2(sh.py);
    dim3 grid(blocks_for(c->nx, sh.py), static_cast<unsigned>(kr.cnt), static_cast<unsigned>(comps));
    if (inverse) hipLaunchKernelGGL(yinv_kernel, grid, dim3(kThreads), sh.lds_y, c->stream, a);
    else hipLaunchKernelGGL(yfwd_kernel, grid, dim3(kThreads), sh.lds_y, c->stream, a);
    timer_end(c, cls, ev);
    return hipGetLastError();
}

static hipError_t run_x(gf_ctx* c, DemagPlan& d, bool conv, KzRange kr, bool adjoint = false) {
    hipEvent_t ev = nullptr;
    timer_begin(c, GF_K_FFT_X_CONV, &ev);
    if (d.general) {
        AnyPassArgs a = any_args(c, d);
        a.nhb = kr.nhb; a.kz0 = kr.k0; a.kzb0 = kr.kb0;
        a.conv = conv ? 1 : 0;
        a.adjoint = adjoint ? 1 : 0;
        a.f = d.fx; a.tw = d.gtx.dev;
        a.pencils = any_pencils(c->nx, conv ? 1024 / 3 : 1024, c->ny);
        dim3 grid(blocks_for(c->ny, a.pencils), static_cast<unsigned>(kr.cnt));
        const hipError_t e = launch_any_pass(AnyPass::kXConv, a, grid, c->stream);
        timer_end(c, GF_K_FFT_X_CONV, ev);
        return e;
    }
    const Shapes sh = shapes(c);
    XArgs a{};
    a.adjoint = adjoint ? 1 : 0;
    a.data = d.spec_b; a.kre = d.kre; a.kim = d.kim; a.tw = d.twx.dev;
    a.nh = c->nx * c->ny * d.nzh;
    a.nhb = kr.nhb; a.kz0 = kr.k0; a.kzb0 = kr.kb0;
    a.den = -c->dx * c->dy * c->dz;
    a.nx = static_cast<int>(c->nx); a.ny = static_cast<int>(c->ny); a.nzh = static_cast<int>(d.nzh);
    a.lognx = ilog2(c->nx); a.pencils = sh.px; a.logp = ilog2(sh.px);
    dim3 grid(blocks_for(c->ny, sh.px), static_cast<unsigned>(kr.cnt));
    if (conv) hipLaunchKernelGGL(xconv_kernel<true>, grid, dim3(kThreads), sh.lds_x, c->stream, a);
    else hipLaunchKernelGGL(xconv_kernel<false>, grid, dim3(kThreads), sh.lds_x, c->stream, a);
    timer_end(c, GF_K_FFT_X_CONV, ev);
    return hipGetLastError();
}

static hipError_t run_zinv(gf_ctx* c, DemagPlan& d, double* const* out = nullptr) {
    hipEvent_t ev = nullptr;
    timer_begin(c, GF_K_FFT_Z_INV, &ev);
    if (d.general) {
        AnyPassArgs a = any_args(c, d);
        a.spec = d.spec_a;
        for (int i = 0; i < 3; i++) a.out[i] = out ? out[i] : c->hstat[i];
        a.f = d.fz; a.tw = d.gtz.dev;
        a.pencils = any_pencils(c->nz, 2048, c->ny);
        dim3 grid(blocks_for(c->ny, a.pencils), static_cast<unsigned>(c->nx), 3);
        const hipError_t e = launch_any_pass(AnyPass::kZInv, a, grid, c->stream);
        timer_end(c, GF_K_FFT_Z_INV, ev);
        return e;
    }
    const Shapes sh = shapes(c);
    ZInvArgs a{};
    a.in = d.spec_a;
    for (int i = 0; i < 3; i++) a.out[i] = out ? out[i] : c->hstat[i];
    a.tw = d.twz.dev;
    a.nh = c->nx * c->ny * d.nzh;
    a.nx = static_cast<int>(c->nx); a.ny = static_cast<int>(c->ny); a.nz = static_cast<int>(c->nz);
    a.nzh = static_cast<int>(d.nzh);
    a.lognz = ilog2(c->nz); a.pencils = sh.pz; a.logp = ilog2(sh.pz);
    dim3 grid(blocks_for(c->ny, sh.pz), static_cast<unsigned>(c->nx), 3);
    hipLaunchKernelGGL(zinv_kernel, grid, dim3(kThreads), sh.lds_z, c->stream, a);
    timer_end(c, GF_K_FFT_Z_INV, ev);
    return hipGetLastError();
}

static bool pow2(int64_t n) { return n >= 1 && (n & (n - 1)) == 0; }

// The work buffers, the tensor spectra storage and the twiddles of a plan:
// radix-2 tables for power-of-two grids, or the mixed-radix factors and
// full tables when a size is not a power of two (prime factors 2, 3, 5, 7;
// the LDS of the x pass, three components of one line through two
// buffers, bounds nx).
static gf_status plan_alloc(gf_ctx* c, DemagPlan& d, const char* what) {
    const std::string name(what);
    d.general = !pow2(c->nx) || !pow2(c->ny) || !pow2(c->nz);
    if (d.general) {
        if (!factor(c->nx, d.fx) || !factor(c->ny, d.fy) || !factor(c->nz, d.fz))
            return set_error(c, GF_ERR_UNSUPPORTED,
                             name + " needs nx, ny, nz whose prime factors are 2, 3, 5 and 7");
        if (c->nx > 1024 || c->ny > 2048 || c->nz > 2048)
            return set_error(c, GF_ERR_UNSUPPORTED, name + " supports nx <= 1024, ny, nz <= 2048");
    } else if (c->nx > 1024 || c->ny > 2048 || c->nz > 2048) {
        return set_error(c, GF_ERR_UNSUPPORTED, name + " supports nx <= 1024, ny, nz <= 2048");
    }
    d.nzh = c->nz / 2 + 1;
    const int64_t nh = c->nx * c->ny * d.nzh;
    gf_status st;
    if ((st = hip_check(c, hipMalloc(&d.spec_m, sizeof(double2) * 3 * nh), "hipMalloc spec_m"))) return st;
    if ((st = hip_check(c, hipMalloc(&d.spec_a, sizeof(double2) * 3 * nh), "hipMalloc spec_a"))) return st;
    if ((st = hip_check(c, hipMalloc(&d.spec_b, sizeof(double2) * 3 * nh), "hipMalloc spec_b"))) return st;
    if ((st = hip_check(c, hipMalloc(&d.kre, sizeof(double) * 9 * nh), "hipMalloc kre"))) return st;
    if ((st = hip_check(c, hipMalloc(&d.kim, sizeof(float) * 9 * nh), "hipMalloc kim"))) return st;
    if (d.general) {
        if ((st = make_table(c, d.gtx, c->nx))) return st;
        if ((st = make_table(c, d.gty, c->ny))) return st;
        if ((st = make_table(c, d.gtz, c->nz))) return st;
        // the fused stage's radix-2 z transforms (power-of-two nz)
        return pow2(c->nz) ? make_twiddles(c, d.twz, c->nz) : GF_OK;
    }
    if ((st = make_twiddles(c, d.twx, c->nx))) return st;
    if ((st = make_twiddles(c, d.twy, c->ny))) return st;
    return make_twiddles(c, d.twz, c->nz);
}

gf_status demag_setup(gf_ctx* c) {
    DemagPlan& d = c->demag;
    gf_status st;
    if ((st = plan_alloc(c, d, "FFT magnetostatics"))) return st;
    const int64_t nh = c->nx * c->ny * d.nzh;

    // Real-space tensor kernels on the host (bit-exact with the reference's
    // evaluation), then forward transforms on the device.
    std::vector<double> real9(static_cast<size_t>(9 * c->n));
    unsigned hw = std::thread::hardware_concurrency();
    int threads = hw ? static_cast<int>(hw) : 1;
    if (threads > 64) threads = 64;
    newell_real_kernels(c->nx, c->ny, c->nz, c->dx, c->dy, c->dz, real9.data(), threads);
    double* draw = nullptr;
    if ((st = hip_check(c, hipMalloc(&draw, sizeof(double) * c->n), "hipMalloc kernel staging"))) return st;
    for (int s = 0; s < 9; s++) {
        st = hip_check(c, hipMemcpyAsync(draw, real9.data() + s * c->n, sizeof(double) * c->n, hipMemcpyHostToDevice,
                                         c->stream), "copy kernel");
        const double* raw3[3] = {draw, nullptr, nullptr};
        if (!st) st = hip_check(c, run_zfwd(c, d, true, raw3, nullptr, nullptr, nullptr, 1), "zfwd kernel");
        if (!st) st = hip_check(c, run_y(c, d, false, 1, whole_range(c, d)), "yfwd kernel");
        if (!st) st = hip_check(c, run_x(c, d, false, whole_range(c, d)), "xfwd kernel");
        if (!st) {
            hipLaunchKernelGGL(split_spectrum_kernel, dim3(blocks_for(nh, kThreads)), dim3(kThreads), 0, c->stream,
                               d.spec_b, d.kre + s * nh, d.kim + s * nh, nh);
            st = hip_check(c, hipGetLastError(), "split spectrum");
        }
        if (!st) st = hip_check(c, hipStreamSynchronize(c->stream), "kernel spectra");
        if (st) { hipFree(draw); return st; }
    }
    hipFree(draw);
    d.ready = true;
    // Fused z transforms when the magnetisation is FM only, nz is a power of
    // two and P whole z columns (P dividing ny) fit the LDS budget of the
    // fused stage (gf_llg.hip); the y and x passes may be mixed-radix.
    d.fused = false;
    if (c->any_fm && !c->any_afm && !(c->flags & GF_FLAG_EMDYNAMIC) && c->nz >= 2 && pow2(c->nz) && d.twz.dev) {
        int P = 1, pmax = 8;
        if (const char* env = std::getenv("GOFERRO_FUSED_P")) {
            const int v = std::atoi(env);
            if (v >= 1 && v <= 64) pmax = v;
        }
        auto lds = [&](int p) {
            const int pairs = p > 1 ? p / 2 : 1;
            return sizeof(double2) * (c->nz / 2 + fft_lds_elems(3 * pairs * c->nz));
        };
        while (P * 2 <= pmax && P * 2 <= c->ny && c->ny % (P * 2) == 0 && lds(P * 2) <= 64 * 1024) P *= 2;
        if (lds(P) <= 48 * 1024) { d.fused = true; d.P = P; }
    }
    d.h_in_spec = false;
    d.h_real_valid = true;
    d.m_in_spec = false;
    return GF_OK;
}

// FFT electrostatics of the polarisation on the same five passes (the
// reference declares get_E_static, ferroelectric_system.h:103, without a
// definition; its set-up, ferroelectric_initialize.cpp:310-338 and
// global.cpp:270-318, fixes a homogeneous permittivity er and continuous
// wave vectors k_a = t 2 pi / (n_a d_a), t = i for i <= n/2 else i - n).  The
// nine "tensor" spectra are E(k) = G(k) P(k) with G_ab = -k_a k_b /
// (eps0 er |k|^2), G(0) = 0, stored real (imaginary parts zero) and scaled
// by den/n for the passes' unnormalised transforms and their division by
// den = -dx dy dz.  An off-diagonal product with a component on its Nyquist
// plane (t = n/2, its own mirror) is set to zero, so that the spectrum is
// Hermitian and the field real.  The plan is the context's second one, so
// the electrostatics runs beside the FFT magnetostatics of a multiferroic.
gf_status elec_setup(gf_ctx* c, double er) {
    if (!(er > 0.)) return set_error(c, GF_ERR_ARG, "FFT electrostatics needs a positive r_permittivity");
    DemagPlan& d = c->eplan;
    gf_status st;
    if ((st = plan_alloc(c, d, "FFT electrostatics"))) return st;
    const int64_t nx = c->nx, ny = c->ny, nzh = d.nzh, nh = nx * ny * nzh;
    auto wave = [](int64_t i, int64_t n, double h, bool& nyq) {
        const int64_t t = i <= n / 2 ? i : i - n;
        nyq = n >= 2 && 2 * i == n;
        return static_cast<double>(t) * (2. * kPi / (static_cast<double>(n) * h));
    };
    const double scale = (-c->dx * c->dy * c->dz) / static_cast<double>(c->n);
    std::vector<double> g(static_cast<size_t>(9 * nh));
    // spectra order xx, yy, zz, xy, yz, zx, yx, zy, xz; layout [kz][y][x]
    const int pa[9] = {0, 1, 2, 0, 1, 2, 1, 2, 0}, pb[9] = {0, 1, 2, 1, 2, 0, 0, 1, 2};
    for (int64_t k = 0; k < nzh; k++)
        for (int64_t j = 0; j < ny; j++)
            for (int64_t i = 0; i < nx; i++) {
                bool ny3[3];
                const double kv[3] = {wave(i, nx, c->dx, ny3[0]), wave(j, ny, c->dy, ny3[1]),
                                      wave(k, c->nz, c->dz, ny3[2])};
                const double k2 = kv[0] * kv[0] + kv[1] * kv[1] + kv[2] * kv[2];
                const int64_t q = (k * ny + j) * nx + i;
                for (int s = 0; s < 9; s++) {
                    const int a = pa[s], b = pb[s];
                    double v = 0.;
                    if (k2 > 0. && (a == b || (!ny3[a] && !ny3[b])))
                        v = -kv[a] * kv[b] / (kEps0 * er * k2) * scale;
                    g[static_cast<size_t>(s * nh + q)] = v;
                }
            }
    if ((st = hip_check(c, hipMemcpy(d.kre, g.data(), sizeof(double) * g.size(), hipMemcpyHostToDevice),
                        "copy electrostatic spectra")))
        return st;
    if ((st = hip_check(c, hipMemset(d.kim, 0, sizeof(float) * 9 * nh), "clear electrostatic spectra"))) return st;
    d.ready = true;
    d.fused = false;
    d.h_in_spec = false;
    d.h_real_valid = true;
    d.m_in_spec = false;
    return GF_OK;
}

// The y forward, x convolution and y inverse passes of one convolution
// (spec_m -> spec_a through spec_b).  With GOFERRO_FFT_KZCHUNKS = C > 1 the
// three passes run over C ranges of kz planes in turn, each range's
// intermediate at the start of spec_b (measured slower on the MI355X at
// 256x256x128 for every C > 1; kept for tuning other grids).
static gf_status conv_passes(gf_ctx* c, DemagPlan& d, bool adjoint, const char* what) {
    int chunks = 1;
    if (const char* env = std::getenv("GOFERRO_FFT_KZCHUNKS")) {
        const int v = std::atoi(env);
        if (v >= 1 && v <= 1024) chunks = v;
    }
    const int nzh = static_cast<int>(d.nzh);
    if (chunks > nzh) chunks = nzh;
    if (chunks <= 1) {
        const KzRange w = whole_range(c, d);
        gf_status st;
        if ((st = hip_check(c, run_y(c, d, false, 3, w), what))) return st;
        if ((st = hip_check(c, run_x(c, d, true, w, adjoint), what))) return st;
        return hip_check(c, run_y(c, d, true, 3, w), what);
    }
    const int per = (nzh + chunks - 1) / chunks;
    const int64_t nhb = static_cast<int64_t>(per) * c->nx * c->ny;
    for (int k0 = 0; k0 < nzh; k0 += per) {
        const KzRange r{k0, k0 + per <= nzh ? per : nzh - k0, k0, nhb};
        gf_status st;
        if ((st = hip_check(c, run_y(c, d, false, 3, r), what))) return st;
        if ((st = hip_check(c, run_x(c, d, true, r, adjoint), what))) return st;
        if ((st = hip_check(c, run_y(c, d, true, 3, r), what))) return st;
    }
    return GF_OK;
}

// The static field of the current polarisation store into Estat.
gf_status elec_run(gf_ctx* c, const double* const* p) {
    DemagPlan& d = c->eplan;
    gf_status st;
    if ((st = hip_check(c, run_zfwd(c, d, true, p, nullptr, nullptr, nullptr, 3), "zfwd electrostatics"))) return st;
    if ((st = conv_passes(c, d, false, "y/x/y passes electrostatics"))) return st;
    return hip_check(c, run_zinv(c, d, c->estat), "zinv electrostatics");
}

// Fused mode: M z-spectrum (from the stage, or a z pass when the store
// changed otherwise) -> y, x with the tensor product, y inverse -> H
// z-spectrum in spec_a; the real field is built only when asked for.
gf_status demag_run_fused(gf_ctx* c, const double* const* fm) {
    DemagPlan& d = c->demag;
    gf_status st;
    if (!d.m_in_spec && (st = hip_check(c, run_zfwd(c, d, false, nullptr, fm, nullptr, nullptr, 3), "zfwd")))
        return st;
    if ((st = conv_passes(c, d, false, "y/x/y passes"))) return st;
    d.h_in_spec = true;
    d.h_real_valid = false;
    d.m_in_spec = false;
    return GF_OK;
}

gf_status demag_materialize_h(gf_ctx* c) {
    DemagPlan& d = c->demag;
    if (!d.ready || !d.fused || d.h_real_valid) return GF_OK;
    gf_status st = hip_check(c, run_zinv(c, d), "zinv");
    if (!st) d.h_real_valid = true;
    return st;
}

// FM-only convolution H = N * (Ms m) of a given store into out (3 arrays).
gf_status demag_run_to(gf_ctx* c, const double* const* fm, double* const* out) {
    DemagPlan& d = c->demag;
    gf_status st;
    if ((st = hip_check(c, run_zfwd(c, d, false, nullptr, fm, nullptr, nullptr, 3), "zfwd"))) return st;
    if ((st = conv_passes(c, d, false, "y/x/y passes"))) return st;
    return hip_check(c, run_zinv(c, d, out), "zinv");
}

// Adjoint convolution out = N^T (*) in of plain real arrays (no Ms): the
// transpose of demag_run_to up to the Ms factor the caller applies.
gf_status demag_adjoint(gf_ctx* c, const double* const* in, double* const* out) {
    DemagPlan& d = c->demag;
    gf_status st;
    if ((st = hip_check(c, run_zfwd(c, d, true, in, nullptr, nullptr, nullptr, 3), "zfwd adjoint"))) return st;
    if ((st = conv_passes(c, d, true, "y/x/y passes adjoint"))) return st;
    return hip_check(c, run_zinv(c, d, out), "zinv adjoint");
}

gf_status demag_run(gf_ctx* c, const double* const* fm, const double* const* s1, const double* const* s2) {
    DemagPlan& d = c->demag;
    gf_status st;
    if ((st = hip_check(c, run_zfwd(c, d, false, nullptr, fm, s1, s2, 3), "zfwd"))) return st;
    if ((st = conv_passes(c, d, false, "y/x/y passes"))) return st;
    return hip_check(c, run_zinv(c, d), "zinv");
}

hipError_t launch_demag_local(gf_ctx* c, const double* const* fm, const double* const* s1, const double* const* s2,
                              int one_d) {
    LocalArgs a{};
    a.code = c->d_code; a.mats = c->d_mats;
    for (int i = 0; i < 3; i++) {
        a.fm[i] = fm[i]; a.a1[i] = s1[i]; a.a2[i] = s2[i]; a.h[i] = c->hstat[i]; a.fac[i] = c->demag_fac[i];
    }
    a.n = c->n; a.one_d = one_d;
    hipEvent_t ev = nullptr;
    timer_begin(c, GF_K_OTHER, &ev);
    hipLaunchKernelGGL(demag_local_kernel, dim3(blocks_for(c->n, kThreads)), dim3(kThreads), 0, c->stream, a);
    timer_end(c, GF_K_OTHER, ev);
    return hipGetLastError();
}

static void plan_free(DemagPlan& d) {
    hipFree(d.spec_m); hipFree(d.spec_a); hipFree(d.spec_b); hipFree(d.kre); hipFree(d.kim);
    hipFree(d.twx.dev); hipFree(d.twy.dev); hipFree(d.twz.dev);
    hipFree(d.gtx.dev); hipFree(d.gty.dev); hipFree(d.gtz.dev);
    d = DemagPlan{};
}

void demag_free(gf_ctx* c) {
    plan_free(c->demag);
    plan_free(c->eplan);
}

}  // namespace gf
