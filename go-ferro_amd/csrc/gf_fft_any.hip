// gf_fft_any.hip — FP64 magnetostatic and electrostatic FFT passes for
// grids whose sizes are not all powers of two.
//
// The reference plans cufftPlan3d(nx, ny, nz) for any grid
// (magnetic_initialize.cpp:114-115, ferroelectric_initialize.cpp:351-352).
// These kernels are the five passes of gf_fft.hip (same layouts, same
// launches, same tensor product) with a mixed-radix line transform instead
// of the radix-2 one: lengths whose prime factors are 2, 3, 5 and 7.
//
// The line transform is a Stockham autosort FFT in LDS: natural order in,
// natural order out, one radix per pass, ping-ponging between two LDS
// buffers.  Pass s of radix R over sub-transforms of size L (the product
// of the radices before it) takes, for each j < n/R, the elements
// j + r n/R (r < R), multiplies element r by w^(r k n/(L R)) with
// k = j mod L, applies the R-point DFT and stores the results at
// (j div L) L R + k + q L.  Twiddles come from one table of n entries,
// w[t] = exp(-2 pi i t / n); the small DFTs use w[m n/R].
#include "gf_internal.h"

namespace gf {

namespace {

constexpr int kThreads = 256;

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 conj_if(double2 a, bool c) { return c ? make_double2(a.x, -a.y) : a; }

// R-point DFT of v in place, forward (exp(-2 pi i q r / R)) or inverse.
template <int R, bool kInv>
__device__ __forceinline__ void small_dft(double2 (&v)[8], const double2* tw, int n) {
    if (R == 2) {
        const double2 a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    } else if (R == 4) {
        // forward: -i for the odd quarter turn, inverse: +i
        const double2 s0 = cadd(v[0], v[2]), d0 = csub(v[0], v[2]);
        const double2 s1 = cadd(v[1], v[3]), d1 = csub(v[1], v[3]);
        const double2 jd1 = kInv ? make_double2(-d1.y, d1.x) : make_double2(d1.y, -d1.x);
        v[0] = cadd(s0, s1);
        v[2] = csub(s0, s1);
        v[1] = cadd(d0, jd1);
        v[3] = csub(d0, jd1);
    } else {
        double2 in[R];
#pragma unroll
        for (int r = 0; r < R; r++) in[r] = v[r];
        const int unit = n / R;
#pragma unroll
        for (int q = 0; q < R; q++) {
            double2 acc = in[0];
#pragma unroll
            for (int r = 1; r < R; r++) acc = cadd(acc, cmul(in[r], conj_if(tw[((q * r) % R) * unit], kInv)));
            v[q] = acc;
        }
    }
}

template <int R, bool kInv>
__device__ __forceinline__ void stockham_pass(const double2* src, double2* dst, int pencils, int n, int L,
                                              const double2* tw) {
    const int per = n / R;
    const int unit = n / (L * R);
    const int total = pencils * per;
    for (int idx = threadIdx.x; idx < total; idx += blockDim.x) {
        const int p = idx / per, j = idx - p * per;
        const double2* in = src + p * n;
        double2* out = dst + p * n;
        const int k = j % L;
        double2 v[8];
#pragma unroll
        for (int r = 0; r < R; r++) v[r] = in[j + r * per];
#pragma unroll
        for (int r = 1; r < R; r++) v[r] = cmul(v[r], conj_if(tw[r * k * unit], kInv));
        small_dft<R, kInv>(v, tw, n);
        const int base = (j / L) * L * R + k;
#pragma unroll
        for (int q = 0; q < R; q++) out[base + q * L] = v[q];
    }
    __syncthreads();
}

// P pencils of length n at a[p*n] (natural order), transformed through b;
// returns the buffer that holds the result.  The caller has synchronised
// the workgroup after filling a and the twiddle table.
template <bool kInv>
__device__ double2* fft_any(double2* a, double2* b, int pencils, const FftFactors& f, const double2* tw) {
    int L = 1;
    for (int s = 0; s < f.count; s++) {
        const int R = f.r[s];
        if (R == 4) stockham_pass<4, kInv>(a, b, pencils, f.n, L, tw);
        else if (R == 2) stockham_pass<2, kInv>(a, b, pencils, f.n, L, tw);
        else if (R == 3) stockham_pass<3, kInv>(a, b, pencils, f.n, L, tw);
        else if (R == 5) stockham_pass<5, kInv>(a, b, pencils, f.n, L, tw);
        else stockham_pass<7, kInv>(a, b, pencils, f.n, L, tw);
        double2* t = a; a = b; b = t;
        L *= R;
    }
    return a;
}

__device__ __forceinline__ void load_table(double2* dst, const double2* src, int n) {
    for (int t = threadIdx.x; t < n; t += blockDim.x) dst[t] = src[t];
}

// zfwd: magnetisation (or raw reals) of pencils of z columns -> z
// transform -> half spectrum in [x][kz][y] (magnetostatic_solver.cpp:121-151).
__global__ void __launch_bounds__(kThreads) zfwd_any_kernel(AnyPassArgs a) {
    extern __shared__ double2 smem[];
    const int c = blockIdx.z, x = blockIdx.y, y0 = blockIdx.x * a.pencils;
    const int nz = a.f.n;
    double2* tw = smem;
    double2* b0 = tw + nz;
    double2* b1 = b0 + a.pencils * nz;
    load_table(tw, a.tw, nz);
    const double* fmc = a.fm[c];
    const double* a1c = a.a1[c];
    const double* a2c = a.a2[c];
    const int total = a.pencils * nz;
    for (int idx = threadIdx.x; idx < total; idx += blockDim.x) {
        const int r = idx / nz, k = idx - r * nz;
        const int y = y0 + r;
        double v = 0.;
        if (y < a.ny) {
            const int64_t cell = (static_cast<int64_t>(x) * a.ny + y) * nz + k;
            if (a.raw_in) {
                v = a.raw[c][cell];
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
        b0[idx] = make_double2(v, 0.);
    }
    __syncthreads();
    const double2* res = fft_any<false>(b0, b1, a.pencils, a.f, tw);
    double2* out = a.spec + c * a.nh;
    const int ztotal = a.pencils * a.nzh;
    for (int idx = threadIdx.x; idx < ztotal; idx += blockDim.x) {
        const int r = idx % a.pencils, kz = idx / a.pencils;
        const int y = y0 + r;
        if (y < a.ny) out[(static_cast<int64_t>(x) * a.nzh + kz) * a.ny + y] = res[r * nz + kz];
    }
}

// y forward: [x][kz][y] (spec) -> [kz][y][x] (spec2, planes from kz0,
// component stride nhb); y inverse the other way.
template <bool kInv>
__global__ void __launch_bounds__(kThreads) y_any_kernel(AnyPassArgs a) {
    extern __shared__ double2 smem[];
    const int c = blockIdx.z, kz = a.kz0 + static_cast<int>(blockIdx.y), x0 = blockIdx.x * a.pencils;
    const int kb = kz - a.kzb0;
    const int ny = a.f.n;
    double2* tw = smem;
    double2* b0 = tw + ny;
    double2* b1 = b0 + a.pencils * ny;
    load_table(tw, a.tw, ny);
    const int total = a.pencils * ny;
    if (!kInv) {
        const double2* in = a.spec + c * a.nh;
        for (int idx = threadIdx.x; idx < total; idx += blockDim.x) {
            const int r = idx / ny, y = idx - r * ny;
            const int x = x0 + r;
            b0[idx] = x < a.nx ? in[(static_cast<int64_t>(x) * a.nzh + kz) * ny + y] : make_double2(0., 0.);
        }
    } else {
        const double2* in = a.spec2 + c * a.nhb;
        for (int idx = threadIdx.x; idx < total; idx += blockDim.x) {
            const int y = idx / a.pencils, r = idx - y * a.pencils;
            const int x = x0 + r;
            b0[r * ny + y] = x < a.nx ? in[(static_cast<int64_t>(kb) * ny + y) * a.nx + x] : make_double2(0., 0.);
        }
    }
    __syncthreads();
    const double2* res = fft_any<kInv>(b0, b1, a.pencils, a.f, tw);
    if (!kInv) {
        double2* out = a.spec2 + c * a.nhb;
        for (int idx = threadIdx.x; idx < total; idx += blockDim.x) {
            const int y = idx / a.pencils, r = idx - y * a.pencils;
            const int x = x0 + r;
            if (x < a.nx) out[(static_cast<int64_t>(kb) * ny + y) * a.nx + x] = res[r * ny + y];
        }
    } else {
        double2* out = a.spec + c * a.nh;
        for (int idx = threadIdx.x; idx < total; idx += blockDim.x) {
            const int r = idx / ny, y = idx - r * ny;
            const int x = x0 + r;
            if (x < a.nx) out[(static_cast<int64_t>(x) * a.nzh + kz) * ny + y] = res[idx];
        }
    }
}

// x forward, the tensor product (magnetostatic_solver.cpp:163-177) and x
// inverse in place on [kz][y][x] (spec2); conv = 0: forward transform of
// component 0 only (tensor set-up).
__global__ void __launch_bounds__(kThreads) xconv_any_kernel(AnyPassArgs a) {
    extern __shared__ double2 smem[];
    const int kz = a.kz0 + static_cast<int>(blockIdx.y), y0 = blockIdx.x * a.pencils;
    const int kb = kz - a.kzb0;
    const int nx = a.f.n;
    const int comps = a.conv ? 3 : 1;
    const int lines = comps * a.pencils;
    double2* tw = smem;
    double2* b0 = tw + nx;
    double2* b1 = b0 + lines * nx;
    load_table(tw, a.tw, nx);
    const int total = lines * nx;
    for (int idx = threadIdx.x; idx < total; idx += blockDim.x) {
        const int t = idx / nx, x = idx - t * nx;
        const int cc = t / a.pencils, r = t - cc * a.pencils;
        const int y = y0 + r;
        b0[idx] = y < a.ny ? a.spec2[cc * a.nhb + (static_cast<int64_t>(kb) * a.ny + y) * nx + x]
                           : make_double2(0., 0.);
    }
    __syncthreads();
    double2* res = fft_any<false>(b0, b1, lines, a.f, tw);
    if (a.conv) {
        const int64_t s = a.nh;
        const int pts = a.pencils * nx;
        for (int idx = threadIdx.x; idx < pts; idx += blockDim.x) {
            const int r = idx / nx, x = idx - r * nx;
            const int y = y0 + r;
            if (y >= a.ny) continue;
            const int64_t g = (static_cast<int64_t>(kz) * a.ny + y) * nx + x;
            double kr[9], ki[9];
#pragma unroll
            for (int q = 0; q < 9; q++) {
                // the adjoint convolution uses N^T with conjugated spectra
                const int tq = a.adjoint ? (q < 3 ? q : (q < 6 ? q + 3 : q - 3)) : q;
                kr[q] = a.kre[tq * s + g];
                const double im = static_cast<double>(a.kim[tq * s + g]);
                ki[q] = a.adjoint ? -im : im;
            }
            const int ix = (0 * a.pencils + r) * nx + x;
            const int iy = (1 * a.pencils + r) * nx + x;
            const int iz = (2 * a.pencils + r) * nx + x;
            const double2 mx = res[ix], my = res[iy], mz = res[iz];
            auto K = [&](int q) { return make_double2(kr[q], ki[q]); };
            double2 t0 = cmul(mx, K(0)), t1 = cmul(my, K(3)), t2 = cmul(mz, K(8));
            const double2 hx = make_double2((t0.x + t1.x + t2.x) / a.den, (t0.y + t1.y + t2.y) / a.den);
            t0 = cmul(my, K(1)); t1 = cmul(mz, K(4)); t2 = cmul(mx, K(6));
            const double2 hy = make_double2((t0.x + t1.x + t2.x) / a.den, (t0.y + t1.y + t2.y) / a.den);
            t0 = cmul(mz, K(2)); t1 = cmul(mx, K(5)); t2 = cmul(my, K(7));
            const double2 hz = make_double2((t0.x + t1.x + t2.x) / a.den, (t0.y + t1.y + t2.y) / a.den);
            res[ix] = hx; res[iy] = hy; res[iz] = hz;
        }
        __syncthreads();
        double2* other = res == b0 ? b1 : b0;
        res = fft_any<true>(res, other, lines, a.f, tw);
    }
    for (int idx = threadIdx.x; idx < total; idx += blockDim.x) {
        const int t = idx / nx, x = idx - t * nx;
        const int cc = t / a.pencils, r = t - cc * a.pencils;
        const int y = y0 + r;
        if (y < a.ny) a.spec2[cc * a.nhb + (static_cast<int64_t>(kb) * a.ny + y) * nx + x] = res[idx];
    }
}

// z complex-to-real inverse: Hermitian extension of the half spectrum
// [x][kz][y], inverse transform, real parts (cuFFT Z2D semantics).
__global__ void __launch_bounds__(kThreads) zinv_any_kernel(AnyPassArgs a) {
    extern __shared__ double2 smem[];
    const int c = blockIdx.z, x = blockIdx.y, y0 = blockIdx.x * a.pencils;
    const int nz = a.f.n;
    double2* tw = smem;
    double2* b0 = tw + nz;
    double2* b1 = b0 + a.pencils * nz;
    load_table(tw, a.tw, nz);
    const double2* in = a.spec + c * a.nh;
    const int ztotal = a.pencils * a.nzh;
    for (int idx = threadIdx.x; idx < ztotal; idx += blockDim.x) {
        const int r = idx % a.pencils, kz = idx / a.pencils;
        const int y = y0 + r;
        const double2 v = y < a.ny ? in[(static_cast<int64_t>(x) * a.nzh + kz) * a.ny + y] : make_double2(0., 0.);
        b0[r * nz + kz] = v;
        const int mirror = nz - kz;
        if (kz >= 1 && mirror > kz && mirror < nz) b0[r * nz + mirror] = make_double2(v.x, -v.y);
    }
    __syncthreads();
    const double2* res = fft_any<true>(b0, b1, a.pencils, a.f, tw);
    double* out = a.out[c];
    const int total = a.pencils * nz;
    for (int idx = threadIdx.x; idx < total; idx += blockDim.x) {
        const int r = idx / nz, k = idx - r * nz;
        const int y = y0 + r;
        if (y < a.ny) out[(static_cast<int64_t>(x) * a.ny + y) * nz + k] = res[idx].x;
    }
}

}  // namespace

// LDS bytes of a generic pass over `lines` pencils of length n: the
// twiddle table and the two Stockham buffers.
size_t any_pass_lds(int lines, int n) { return sizeof(double2) * (static_cast<size_t>(n) + 2u * lines * n); }

// Dynamic LDS beyond 64 KB (up to the 160 KB of a gfx950 CU) is allowed
// per kernel; a refusal of the attribute is left to the launch to report.
static void allow_lds(const void* fn, size_t lds) {
    if (lds <= 64 * 1024) return;
    if (hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds)) != hipSuccess)
        (void)hipGetLastError();
}

hipError_t launch_any_pass(AnyPass kind, const AnyPassArgs& a, dim3 grid, hipStream_t s) {
    const int lines = kind == AnyPass::kXConv ? (a.conv ? 3 : 1) * a.pencils : a.pencils;
    const size_t lds = any_pass_lds(lines, a.f.n);
    if (lds > 160 * 1024) return hipErrorInvalidValue;
    switch (kind) {
    case AnyPass::kZFwd: allow_lds(reinterpret_cast<const void*>(zfwd_any_kernel), lds); break;
    case AnyPass::kYFwd: allow_lds(reinterpret_cast<const void*>(y_any_kernel<false>), lds); break;
    case AnyPass::kYInv: allow_lds(reinterpret_cast<const void*>(y_any_kernel<true>), lds); break;
    case AnyPass::kXConv: allow_lds(reinterpret_cast<const void*>(xconv_any_kernel), lds); break;
    case AnyPass::kZInv: allow_lds(reinterpret_cast<const void*>(zinv_any_kernel), lds); break;
    }
    switch (kind) {
    case AnyPass::kZFwd: hipLaunchKernelGGL(zfwd_any_kernel, grid, dim3(kThreads), lds, s, a); break;
    case AnyPass::kYFwd: hipLaunchKernelGGL(y_any_kernel<false>, grid, dim3(kThreads), lds, s, a); break;
    case AnyPass::kYInv: hipLaunchKernelGGL(y_any_kernel<true>, grid, dim3(kThreads), lds, s, a); break;
    case AnyPass::kXConv: hipLaunchKernelGGL(xconv_any_kernel, grid, dim3(kThreads), lds, s, a); break;
    case AnyPass::kZInv: hipLaunchKernelGGL(zinv_any_kernel, grid, dim3(kThreads), lds, s, a); break;
    }
    return hipGetLastError();
}

}  // namespace gf
