// gf_llg_dev.h — device building blocks of the magnet stages, shared by
// the forward stage kernels (gf_llg.hip) and their adjoints
// (gf_adjoint.hip): frame rotation, indexed access with 32-bit offsets, the
// exchange/DMI stencil and the magnetoelastic field.
#pragma once

#include "gf_fft_dev.h"

namespace gf {

namespace {

__device__ __forceinline__ void rot(const double T[3][3], double ix, double iy, double iz, double& ox, double& oy,
                                    double& oz) {
    ox = T[0][0] * ix + T[0][1] * iy + T[0][2] * iz;
    oy = T[1][0] * ix + T[1][1] * iy + T[1][2] * iz;
    oz = T[2][0] * ix + T[2][1] * iy + T[2][2] * iz;
}

// Element access by a 64-bit index, or by a 32-bit index turned into a
// 32-bit byte offset from the (uniform) array base: global loads and stores
// then take the base from scalar registers and one 32-bit offset from
// vector registers, which saves registers and address arithmetic.  The
// 32-bit form is used when every array holds fewer than 2^28 elements.
__device__ __forceinline__ double ldd(const double* p, int64_t i) { return p[i]; }
__device__ __forceinline__ double ldd(const double* p, uint32_t i) {
    return *reinterpret_cast<const double*>(reinterpret_cast<const char*>(p) + (i << 3));
}
__device__ __forceinline__ void std_(double* p, int64_t i, double v) { p[i] = v; }
__device__ __forceinline__ void std_(double* p, uint32_t i, double v) {
    *reinterpret_cast<double*>(reinterpret_cast<char*>(p) + (i << 3)) = v;
}
__device__ __forceinline__ uint32_t ldu(const uint32_t* p, int64_t i) { return p[i]; }
__device__ __forceinline__ uint32_t ldu(const uint32_t* p, uint32_t i) {
    return *reinterpret_cast<const uint32_t*>(reinterpret_cast<const char*>(p) + (i << 2));
}

// Exchange-weighted neighbour differences of one store for one cell
// (get_m_spatial_derivative_local, magnetic_solver.cpp:1639-1831).
// Returns hx[3] = sum_faces Aij * d2m (before the 2/mu0/Ms factor) and the
// DMI first derivatives d[4] = dmxdx, dmydy, dmzdx, dmzdy.
struct Stencil {
    double ex[3];
    double d[4];
};

struct P3 {
    const double* p[3];
};

template <bool kDmi, class I>
__device__ __forceinline__ Stencil stencil(const P3 s3, const uint32_t* code, const DevMaterial* mats,
                                           uint32_t own, int surf_shift, int sub, I i, I j, I k,
                                           I nx, I ny, I nz, double dx, double dy, double dz,
                                           double A, double A_self, double slope) {
    const I id = (i * ny + j) * nz + k;
    const double* s[3] = {s3.p[0], s3.p[1], s3.p[2]};
    const double c0 = ldd(s[0], id), c1 = ldd(s[1], id), c2 = ldd(s[2], id);
    const uint32_t mat = own & kMatMask;
    // one reciprocal per spacing instead of a division per difference: each
    // quotient moves by at most an ulp against the reference's c/dx
    const double ix = 1. / dx, iy = 1. / dy, iz = 1. / dz;
    Stencil r;
    double bw[3], fw[3], ab, af;
    // x faces
    {
        I nb;
        if ((own >> (surf_shift + kXB)) & 1u) {
            if (kDmi) { bw[0] = -slope * c2; bw[1] = 0.; bw[2] = slope * c0; }
            else { bw[0] = 0.; bw[1] = 0.; bw[2] = 0.; }
            ab = A;
        } else {
            nb = ((i == 0 ? nx - 1 : i - 1) * ny + j) * nz + k;
            bw[0] = (c0 - ldd(s[0], nb)) * ix; bw[1] = (c1 - ldd(s[1], nb)) * ix; bw[2] = (c2 - ldd(s[2], nb)) * ix;
            const uint32_t nm = ldu(code, nb) & kMatMask;
            if (nm == mat) ab = A_self;
            else {
                const double a2 = sub < 0 ? mats[nm - 1].Aex : mats[nm - 1].aAex[sub];
                ab = 2. * A * a2 / (A + a2);
            }
        }
        if ((own >> (surf_shift + kXF)) & 1u) {
            if (kDmi) { fw[0] = -slope * c2; fw[1] = 0.; fw[2] = slope * c0; }
            else { fw[0] = 0.; fw[1] = 0.; fw[2] = 0.; }
            af = A;
        } else {
            nb = ((i == nx - 1 ? 0 : i + 1) * ny + j) * nz + k;
            fw[0] = (ldd(s[0], nb) - c0) * ix; fw[1] = (ldd(s[1], nb) - c1) * ix; fw[2] = (ldd(s[2], nb) - c2) * ix;
            const uint32_t nm = ldu(code, nb) & kMatMask;
            if (nm == mat) af = A_self;
            else {
                const double a2 = sub < 0 ? mats[nm - 1].Aex : mats[nm - 1].aAex[sub];
                af = 2. * A * a2 / (A + a2);
            }
        }
        r.d[0] = (bw[0] + fw[0]) * 0.5;
        r.d[2] = (bw[2] + fw[2]) * 0.5;
        for (int a = 0; a < 3; a++) r.ex[a] = af * (fw[a] * ix) + ab * (-bw[a] * ix);
    }
    // y faces
    {
        I nb;
        if ((own >> (surf_shift + kYB)) & 1u) {
            if (kDmi) { bw[0] = 0.; bw[1] = -slope * c2; bw[2] = slope * c1; }
            else { bw[0] = 0.; bw[1] = 0.; bw[2] = 0.; }
            ab = A;
        } else {
            nb = (i * ny + (j == 0 ? ny - 1 : j - 1)) * nz + k;
            bw[0] = (c0 - ldd(s[0], nb)) * iy; bw[1] = (c1 - ldd(s[1], nb)) * iy; bw[2] = (c2 - ldd(s[2], nb)) * iy;
            const uint32_t nm = ldu(code, nb) & kMatMask;
            if (nm == mat) ab = A_self;
            else {
                const double a2 = sub < 0 ? mats[nm - 1].Aex : mats[nm - 1].aAex[sub];
                ab = 2. * A * a2 / (A + a2);
            }
        }
        if ((own >> (surf_shift + kYF)) & 1u) {
            if (kDmi) { fw[0] = 0.; fw[1] = -slope * c2; fw[2] = slope * c1; }
            else { fw[0] = 0.; fw[1] = 0.; fw[2] = 0.; }
            af = A;
        } else {
            nb = (i * ny + (j == ny - 1 ? 0 : j + 1)) * nz + k;
            fw[0] = (ldd(s[0], nb) - c0) * iy; fw[1] = (ldd(s[1], nb) - c1) * iy; fw[2] = (ldd(s[2], nb) - c2) * iy;
            const uint32_t nm = ldu(code, nb) & kMatMask;
            if (nm == mat) af = A_self;
            else {
                const double a2 = sub < 0 ? mats[nm - 1].Aex : mats[nm - 1].aAex[sub];
                af = 2. * A * a2 / (A + a2);
            }
        }
        r.d[1] = (bw[1] + fw[1]) * 0.5;
        r.d[3] = (bw[2] + fw[2]) * 0.5;
        for (int a = 0; a < 3; a++) r.ex[a] = r.ex[a] + af * (fw[a] * iy) + ab * (-bw[a] * iy);
    }
    // z faces: zero slope at surfaces for both kinds
    {
        I nb;
        if ((own >> (surf_shift + kZB)) & 1u) {
            bw[0] = 0.; bw[1] = 0.; bw[2] = 0.;
            ab = A;
        } else {
            nb = (i * ny + j) * nz + (k == 0 ? nz - 1 : k - 1);
            bw[0] = (c0 - ldd(s[0], nb)) * iz; bw[1] = (c1 - ldd(s[1], nb)) * iz; bw[2] = (c2 - ldd(s[2], nb)) * iz;
            const uint32_t nm = ldu(code, nb) & kMatMask;
            if (nm == mat) ab = A_self;
            else {
                const double a2 = sub < 0 ? mats[nm - 1].Aex : mats[nm - 1].aAex[sub];
                ab = 2. * A * a2 / (A + a2);
            }
        }
        if ((own >> (surf_shift + kZF)) & 1u) {
            fw[0] = 0.; fw[1] = 0.; fw[2] = 0.;
            af = A;
        } else {
            nb = (i * ny + j) * nz + (k == nz - 1 ? 0 : k + 1);
            fw[0] = (ldd(s[0], nb) - c0) * iz; fw[1] = (ldd(s[1], nb) - c1) * iz; fw[2] = (ldd(s[2], nb) - c2) * iz;
            const uint32_t nm = ldu(code, nb) & kMatMask;
            if (nm == mat) af = A_self;
            else {
                const double a2 = sub < 0 ? mats[nm - 1].Aex : mats[nm - 1].aAex[sub];
                af = 2. * A * a2 / (A + a2);
            }
        }
        for (int a = 0; a < 3; a++) r.ex[a] = r.ex[a] + af * (fw[a] * iz) + ab * (-bw[a] * iz);
    }
    return r;
}

// Magnetoelastic field of a crystal-frame strain, in the global frame:
// f (B1 e_aa m_a + B2 (e_ab m_b + e_ac m_c)) with f = -2 (FM) or -1 (per
// AFM sublattice), B already divided by mu0 Ms (magnetic_solver.cpp:143-150,
// magnetic_AFM_solver.cpp:156-162).  Strain is Voigt xx yy zz yz xz xy.
template <class I>
__device__ __forceinline__ void magnetoelastic(const double* const* e, I id, double f, double B1, double B2,
                                               double mx, double my, double mz, int rotate, const double T[3][3],
                                               double& h0, double& h1, double& h2) {
    const double exx = 0. + ldd(e[0], id), eyy = 0. + ldd(e[1], id), ezz = 0. + ldd(e[2], id);
    const double eyz = 0. + ldd(e[3], id), exz = 0. + ldd(e[4], id), exy = 0. + ldd(e[5], id);
    const double c0 = f * (B1 * (exx) * mx + B2 * ((exy) * my + (exz) * mz));
    const double c1 = f * (B1 * (eyy) * my + B2 * ((exy) * mx + (eyz) * mz));
    const double c2 = f * (B1 * (ezz) * mz + B2 * ((exz) * mx + (eyz) * my));
    if (rotate) rot(T, c0, c1, c2, h0, h1, h2);
    else { h0 = c0; h1 = c1; h2 = c2; }
}

}  // namespace

}  // namespace gf
