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

__device__ __forceinline__ double2 fft_cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// P pencils of length N = 2^logn at buf[p*N], bit-reversed in, natural out.
// Radix-2 decimation in time; two consecutive stages run in one pass per
// thread (four elements, the same butterflies in the same order), which
// halves the LDS round trips and barriers without changing a rounding.
__device__ __forceinline__ void fft_dit(double2* buf, int pencils, int logn, const double2* tw) {
    const int N = 1 << logn;
    int s = 0;
    if (logn & 1) {
        // lone first stage (span 2, twiddle 1)
        const int total = pencils * (N >> 1);
        for (int b = threadIdx.x; b < total; b += blockDim.x) {
            const int i0 = 2 * b, i1 = i0 + 1;
            const double2 w = tw[0];
            const double2 u = buf[fpad(i0)];
            const double2 v = buf[fpad(i1)];
            const double2 t = make_double2(v.x * w.x - v.y * w.y, v.x * w.y + v.y * w.x);
            buf[fpad(i0)] = make_double2(u.x + t.x, u.y + t.y);
            buf[fpad(i1)] = make_double2(u.x - t.x, u.y - t.y);
        }
        __syncthreads();
        s = 1;
    }
    const int total4 = pencils * (N >> 2);
    for (; s < logn; s += 2) {
        const int h = 1 << s;
        const int ts0 = N >> (s + 1), ts1 = N >> (s + 2);
        for (int b = threadIdx.x; b < total4; b += blockDim.x) {
            const int p = b >> (logn - 2);
            const int q = b & ((N >> 2) - 1);
            const int j = q & (h - 1);
            const int g = p * N + ((q >> s) << (s + 2)) + j;
            double2 a0 = buf[fpad(g)], a1 = buf[fpad(g + h)], a2 = buf[fpad(g + 2 * h)], a3 = buf[fpad(g + 3 * h)];
            // stage s: (a0, a1) and (a2, a3), twiddle j at span 2h
            double2 w = tw[j * ts0];
            double2 t = make_double2(a1.x * w.x - a1.y * w.y, a1.x * w.y + a1.y * w.x);
            double2 b0 = make_double2(a0.x + t.x, a0.y + t.y), b1 = make_double2(a0.x - t.x, a0.y - t.y);
            t = make_double2(a3.x * w.x - a3.y * w.y, a3.x * w.y + a3.y * w.x);
            double2 b2 = make_double2(a2.x + t.x, a2.y + t.y), b3 = make_double2(a2.x - t.x, a2.y - t.y);
            // stage s+1: (b0, b2) twiddle j, (b1, b3) twiddle j + h, at span 4h
            w = tw[j * ts1];
            t = make_double2(b2.x * w.x - b2.y * w.y, b2.x * w.y + b2.y * w.x);
            buf[fpad(g)] = make_double2(b0.x + t.x, b0.y + t.y);
            buf[fpad(g + 2 * h)] = make_double2(b0.x - t.x, b0.y - t.y);
            w = tw[(j + h) * ts1];
            t = make_double2(b3.x * w.x - b3.y * w.y, b3.x * w.y + b3.y * w.x);
            buf[fpad(g + h)] = make_double2(b1.x + t.x, b1.y + t.y);
            buf[fpad(g + 3 * h)] = make_double2(b1.x - t.x, b1.y - t.y);
        }
        __syncthreads();
    }
}

// Inverse transform (conjugate twiddles), natural in, bit-reversed out;
// radix-2 decimation in frequency with two stages per pass as above.
__device__ __forceinline__ void ifft_dif(double2* buf, int pencils, int logn, const double2* tw) {
    const int N = 1 << logn;
    const int total4 = pencils * (N >> 2);
    int s = logn - 1;
    for (; s >= 1; s -= 2) {
        const int h = 1 << s, h2 = h >> 1;
        const int ts0 = N >> (s + 1), ts1 = N >> s;
        for (int b = threadIdx.x; b < total4; b += blockDim.x) {
            const int p = b >> (logn - 2);
            const int q = b & ((N >> 2) - 1);
            const int j = q & (h2 - 1);
            const int g = p * N + ((q >> (s - 1)) << (s + 1)) + j;
            const double2 a0 = buf[fpad(g)], a1 = buf[fpad(g + h2)], a2 = buf[fpad(g + h)], a3 = buf[fpad(g + h + h2)];
            // stage s: (a0, a2) twiddle j, (a1, a3) twiddle j + h/2, at span 2h
            double2 w = make_double2(tw[j * ts0].x, -tw[j * ts0].y);
            double2 d = make_double2(a0.x - a2.x, a0.y - a2.y);
            const double2 b0 = make_double2(a0.x + a2.x, a0.y + a2.y);
            const double2 b2 = make_double2(d.x * w.x - d.y * w.y, d.x * w.y + d.y * w.x);
            w = make_double2(tw[(j + h2) * ts0].x, -tw[(j + h2) * ts0].y);
            d = make_double2(a1.x - a3.x, a1.y - a3.y);
            const double2 b1 = make_double2(a1.x + a3.x, a1.y + a3.y);
            const double2 b3 = make_double2(d.x * w.x - d.y * w.y, d.x * w.y + d.y * w.x);
            // stage s-1: (b0, b1) and (b2, b3), twiddle j at span h
            w = make_double2(tw[j * ts1].x, -tw[j * ts1].y);
            d = make_double2(b0.x - b1.x, b0.y - b1.y);
            buf[fpad(g)] = make_double2(b0.x + b1.x, b0.y + b1.y);
            buf[fpad(g + h2)] = make_double2(d.x * w.x - d.y * w.y, d.x * w.y + d.y * w.x);
            d = make_double2(b2.x - b3.x, b2.y - b3.y);
            buf[fpad(g + h)] = make_double2(b2.x + b3.x, b2.y + b3.y);
            buf[fpad(g + h + h2)] = make_double2(d.x * w.x - d.y * w.y, d.x * w.y + d.y * w.x);
        }
        __syncthreads();
    }
    if (s == 0) {
        // lone last stage (span 2, twiddle 1)
        const int total = pencils * (N >> 1);
        for (int b = threadIdx.x; b < total; b += blockDim.x) {
            const int i0 = 2 * b, i1 = i0 + 1;
            const double2 w = make_double2(tw[0].x, -tw[0].y);
            const double2 u = buf[fpad(i0)];
            const double2 v = buf[fpad(i1)];
            const double2 d = make_double2(u.x - v.x, u.y - v.y);
            buf[fpad(i0)] = make_double2(u.x + v.x, u.y + v.y);
            buf[fpad(i1)] = make_double2(d.x * w.x - d.y * w.y, d.x * w.y + d.y * w.x);
        }
        __syncthreads();
    }
}

__device__ __forceinline__ void fft_load_twiddles(double2* dst, const double2* src, int n) {
    for (int j = threadIdx.x; j < (n >> 1); j += blockDim.x) dst[j] = src[j];
}

// The same in two halves, so that the first load can be issued together
// with a kernel's other set-up loads: fft_twiddle_prefetch returns this
// thread's first twiddle, fft_twiddle_store writes it and loads the rest.
__device__ __forceinline__ double2 fft_twiddle_prefetch(const double2* src, int n) {
    return static_cast<int>(threadIdx.x) < (n >> 1) ? src[threadIdx.x] : make_double2(0., 0.);
}
__device__ __forceinline__ void fft_twiddle_store(double2* dst, const double2* src, int n, double2 first) {
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
