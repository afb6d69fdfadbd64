// gf_fft_dev.h — FFT building blocks on LDS, shared by the
// magnetostatic passes (gf_fft.hip) and the fused magnetic stage
// (gf_llg.hip).  Transforms are unnormalised; twiddles w[j] =
// exp(-2 pi i j / N), j < N/2, are staged in LDS by the caller.
#pragma once

#include "gf_internal.h"

namespace gf {

__device__ __forceinline__ int fft_bitrev(int j, int logn) {
    return logn ? static_cast<int>(__brev(static_cast<unsigned>(j)) >> (32 - logn)) : 0;
}

// LDS slot of element i of a transform buffer: one spare slot after every
// 16 elements, so that the power-of-two strides of the butterflies and of
// the bit-reversed loads do not pile onto the same LDS banks.  A buffer of
// m elements takes fft_lds_elems(m) slots.
__device__ __forceinline__ int fpad(int i) { return i + (i >> 4); }
__host__ __device__ __forceinline__ int64_t fft_lds_elems(int64_t m) { return m + (m >> 4) + 1; }

// Complex values of the LDS transforms: double2 for the FP64 path, float2
// for the FP32 variant of the adjoint sweep.
template <class C> struct FftComplex;
template <> struct FftComplex<double2> {
    static __device__ __forceinline__ double2 make(double re, double im) { return make_double2(re, im); }
};
template <> struct FftComplex<float2> {
    static __device__ __forceinline__ float2 make(float re, float im) { return make_float2(re, im); }
};

__device__ __forceinline__ double2 fft_cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// One pass of R = 2, 4 or 8 elements per thread over log2(R) consecutive
// radix-2 stages s, s+1, ... of P pencils of length N = 2^logn at buf[p*N]:
// the elements g + m*2^s (m < R) of each group are read once, go through
// the stages' butterflies in registers, and are written back once.  The
// butterflies, their twiddles tw[(i mod 2^t) * N/2^(t+1)] (stage t, pair
// (i, i + 2^t)) and their arithmetic are those of the plain radix-2
// transform, so the grouping changes no rounding.  Decimation in time
// (forward, kInverse = false) runs the stages upwards with
// (u, v) -> (u + v w, u - v w); decimation in frequency (inverse,
// conjugate twiddles) runs them downwards with (u, v) -> (u + v, (u - v) w*).
template <int R, bool kInverse, class C>
__device__ __forceinline__ void fft_pass(C* buf, int pencils, int logn, int s, const C* tw) {
    using F = FftComplex<C>;
    constexpr int LR = R == 2 ? 1 : (R == 4 ? 2 : 3);
    const int N = 1 << logn, h = 1 << s;
    const int groups = N >> LR;
    const int total = pencils * groups;
    for (int b = threadIdx.x; b < total; b += blockDim.x) {
        const int p = b >> (logn - LR);
        const int q = b & (groups - 1);
        const int j = q & (h - 1);
        const int g = p * N + ((q >> s) << (s + LR)) + j;
        C a[R];
#pragma unroll
        for (int m = 0; m < R; m++) a[m] = buf[fpad(g + m * h)];
#pragma unroll
        for (int rr = 0; rr < LR; rr++) {
            const int r = kInverse ? LR - 1 - rr : rr;
            const int span = 1 << r;
            const int ts = N >> (s + r + 1);
#pragma unroll
            for (int m = 0; m < R; m++) {
                if (m & span) continue;
                const C tv = tw[(j + (m & (span - 1)) * h) * ts];
                const C u = a[m], v = a[m + span];
                if (!kInverse) {
                    const C w = tv;
                    const C t = F::make(v.x * w.x - v.y * w.y, v.x * w.y + v.y * w.x);
                    a[m] = F::make(u.x + t.x, u.y + t.y);
                    a[m + span] = F::make(u.x - t.x, u.y - t.y);
                } else {
                    const C w = F::make(tv.x, -tv.y);
                    const C d = F::make(u.x - v.x, u.y - v.y);
                    a[m] = F::make(u.x + v.x, u.y + v.y);
                    a[m + span] = F::make(d.x * w.x - d.y * w.y, d.x * w.y + d.y * w.x);
                }
            }
        }
#pragma unroll
        for (int m = 0; m < R; m++) buf[fpad(g + m * h)] = a[m];
    }
    __syncthreads();
}

// P pencils of length N = 2^logn at buf[p*N], bit-reversed in, natural out:
// radix-2 decimation in time, log2(kRadix) stages per LDS pass (the
// leftover stages first).  Radix 8 takes fewer passes and barriers; radix 4
// needs fewer registers and keeps more threads busy on short transforms.
template <int kRadix = 8, class C>
__device__ __forceinline__ void fft_dit(C* buf, int pencils, int logn, const C* tw) {
    constexpr int LR = kRadix == 8 ? 3 : 2;
    int s = logn % LR;
    if (s == 1) fft_pass<2, false>(buf, pencils, logn, 0, tw);
    else if (s == 2) fft_pass<4, false>(buf, pencils, logn, 0, tw);
    for (; s < logn; s += LR) fft_pass<kRadix, false>(buf, pencils, logn, s, tw);
}

// Inverse transform (conjugate twiddles), natural in, bit-reversed out:
// radix-2 decimation in frequency, log2(kRadix) stages per pass from the
// top, the leftover stages last.
template <int kRadix = 8, class C>
__device__ __forceinline__ void ifft_dif(C* buf, int pencils, int logn, const C* tw) {
    constexpr int LR = kRadix == 8 ? 3 : 2;
    int s = logn - LR;
    for (; s >= 0; s -= LR) fft_pass<kRadix, true>(buf, pencils, logn, s, tw);
    if (s + LR == 2) fft_pass<4, true>(buf, pencils, logn, 0, tw);
    else if (s + LR == 1) fft_pass<2, true>(buf, pencils, logn, 0, tw);
}

__device__ __forceinline__ void fft_load_twiddles(double2* dst, const double2* src, int n) {
    for (int j = threadIdx.x; j < (n >> 1); j += blockDim.x) dst[j] = src[j];
}

// The same in two halves, so that the first load can be issued together
// with a kernel's other set-up loads: fft_twiddle_prefetch returns this
// thread's first twiddle, fft_twiddle_store writes it and loads the rest.
template <class C>
__device__ __forceinline__ C fft_twiddle_prefetch(const C* src, int n) {
    return static_cast<int>(threadIdx.x) < (n >> 1) ? src[threadIdx.x] : FftComplex<C>::make(0, 0);
}
template <class C>
__device__ __forceinline__ void fft_twiddle_store(C* dst, const C* src, int n, C first) {
    if (static_cast<int>(threadIdx.x) < (n >> 1)) dst[threadIdx.x] = first;
    for (int j = threadIdx.x + blockDim.x; j < (n >> 1); j += blockDim.x) dst[j] = src[j];
}

// Copies elements 0..total-1 from global memory into LDS through
// registers, kBatch elements per thread at a time: ld(idx) loads element
// idx and st(idx, v) writes it.  Every load of a batch is issued before the
// first LDS write waits for one, so a workgroup pays one memory round trip
// per batch rather than one per element.
template <int kBatch, class T, class Ld, class St>
__device__ __forceinline__ void fft_staged_load(int total, Ld ld, St st) {
    for (int b0 = threadIdx.x; b0 < total; b0 += kBatch * blockDim.x) {
        T v[kBatch];
#pragma unroll
        for (int u = 0; u < kBatch; u++) {
            const int idx = b0 + u * static_cast<int>(blockDim.x);
            if (idx < total) v[u] = ld(idx);
        }
#pragma unroll
        for (int u = 0; u < kBatch; u++) {
            const int idx = b0 + u * static_cast<int>(blockDim.x);
            if (idx < total) st(idx, v[u]);
        }
    }
}

}  // namespace gf
