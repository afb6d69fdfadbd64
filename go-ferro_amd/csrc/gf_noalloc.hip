// gf_noalloc.hip — an averaging entry point that allocates nothing.
//
// gf_average reduces a field with per-block partial sums in a buffer the
// context allocates, copies the partials to the host and waits.  Callers
// that cannot allocate or wait (code under HIP graph capture, serving loops
// that forbid hipMalloc) use gf_average_noalloc instead: the same sum3
// kernel writes its partials into `scratch`, a file-scope device array of
// fixed size that every GPU holds until the process ends, and a one-block
// kernel finishes the sums in the host loop's order (bit-identical to
// gf_average) into the caller's device output.  Nothing is allocated and
// nothing waits.
//
// The array's host-side name is not a device address: each device's copy
// is resolved once with hipGetSymbolAddress (with that device current) and
// only that address is handed to kernels.  Every call on one GPU shares that
// GPU's array, so calls that overlap on one GPU (two streams, two threads)
// race on it; calls ordered on one stream do not.  sum3 writes every
// partial that the finishing kernel reads, so nothing needs clearing.
#include "gf_internal.h"

#include <mutex>

using namespace gf;

namespace {

constexpr int kMaxBlocks = 1024;                      // sum3 blocks a context uses at most
constexpr int kScratchDoubles = 3 * kMaxBlocks;       // 24 KiB per device
constexpr int kMaxDevices = 64;

static __device__ __align__(256) double scratch[kScratchDoubles];

std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

// This device's copy of `scratch`, resolved on the first call on it.
gf_status scratch_for(gf_ctx* c, double** out) {
    const int dev = c->device;
    if (dev < 0 || dev >= kMaxDevices) return set_error(c, GF_ERR_UNSUPPORTED, "device index beyond 63");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        int cur = -1;
        gf_status st = hip_check(c, hipGetDevice(&cur), "hipGetDevice");
        if (st) return st;
        if (cur != dev && (st = hip_check(c, hipSetDevice(dev), "hipSetDevice"))) return st;
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != dev) hipSetDevice(cur);
        if ((st = hip_check(c, e, "hipGetSymbolAddress scratch"))) return st;
        g_scratch_ptr[dev] = static_cast<double*>(p);
    }
    *out = g_scratch_ptr[dev];
    return GF_OK;
}

// out[a] = (sum over b of partial[3b + a], left to right) / cnt: the host
// loop of gf_average, one thread per component.
__global__ void __launch_bounds__(64) average_finish_kernel(const double* partial, int64_t blocks, double cnt,
                                                            double* out) {
    const int a = static_cast<int>(threadIdx.x);
    if (a >= 3) return;
    double s = 0.;
    for (int64_t b = 0; b < blocks; b++) s += partial[3 * b + a];
    out[a] = s / cnt;
}

}  // namespace

extern "C" gf_status gf_average_noalloc(gf_ctx* c, int id, double* out) {
    if (!c || !out) return GF_ERR_ARG;
    const double* f[3];
    int64_t cnt;
    if (id == GF_FIELD_M && c->any_fm) {
        for (int a = 0; a < 3; a++) f[a] = c->m[a];
        cnt = c->nfm;
    } else if ((id == GF_FIELD_M_AFM1 || id == GF_FIELD_M_AFM2) && c->any_afm) {
        for (int a = 0; a < 3; a++) f[a] = id == GF_FIELD_M_AFM1 ? c->m1[a] : c->m2[a];
        cnt = c->nafm;
    } else if (id == GF_FIELD_P && c->any_fe) {
        for (int a = 0; a < 3; a++) f[a] = c->p[a];
        cnt = c->nfe;
    } else {
        return set_error(c, GF_ERR_ARG, "average is defined for m, m_AFM1, m_AFM2, p of a context that has them");
    }
    if (c->red_blocks < 1 || 3 * c->red_blocks > kScratchDoubles)
        return set_error(c, GF_ERR_UNSUPPORTED,
                         "gf_average_noalloc holds at most 1024 partial sums per component (a context has "
                         "min(1024, ceil(n / 256)))");
    double* part = nullptr;
    gf_status st = scratch_for(c, &part);
    if (st) return st;
    if ((st = hip_check(c, launch_sum3(f, c->n, part, c->red_blocks, c->stream), "sum"))) return st;
    hipLaunchKernelGGL(average_finish_kernel, dim3(1), dim3(64), 0, c->stream, part, c->red_blocks,
                       static_cast<double>(cnt), out);
    return hip_check(c, hipGetLastError(), "average finish");
}
