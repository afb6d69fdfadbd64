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
// Every access goes through the global address space: pointers read from
// memory (the per-instance arguments of a batch) would otherwise be generic
// and compile to slower flat accesses with 64-bit addresses.
typedef const __attribute__((address_space(1))) double GConstD;
typedef __attribute__((address_space(1))) double GD;
typedef const __attribute__((address_space(1))) uint32_t GConstU;
typedef const __attribute__((address_space(1))) char GConstC;
typedef __attribute__((address_space(1))) char GC;

__device__ __forceinline__ double ldd(const double* p, int64_t i) { return ((GConstD*)p)[i]; }
__device__ __forceinline__ double ldd(const double* p, uint32_t i) {
    return *(GConstD*)((GConstC*)p + (i << 3));
}
__device__ __forceinline__ void std_(double* p, int64_t i, double v) { ((GD*)p)[i] = v; }
__device__ __forceinline__ void std_(double* p, uint32_t i, double v) { *(GD*)((GC*)p + (i << 3)) = v; }
// The material record of a cell, read through the global address space,
// or from a workgroup's LDS copy of the records (LMat).
typedef const __attribute__((address_space(1))) DevMaterial GMat;
typedef const __attribute__((address_space(3))) DevMaterial LMat;
__device__ __forceinline__ GMat& mat_of(const DevMaterial* mats, uint32_t code) {
    return ((GMat*)mats)[(code & kMatMask) - 1];
}

// Nontemporal forms for arrays a stage streams through once (m, the RK
// accumulator, the next store): they then displace less of the stencil
// stores and cell codes that neighbouring workgroups re-read from L2.
__device__ __forceinline__ double ldd_nt(const double* p, int64_t i) {
    return __builtin_nontemporal_load(&((GConstD*)p)[i]);
}
__device__ __forceinline__ double ldd_nt(const double* p, uint32_t i) {
    return __builtin_nontemporal_load((GConstD*)((GConstC*)p + (i << 3)));
}
__device__ __forceinline__ void std_nt(double* p, int64_t i, double v) { __builtin_nontemporal_store(v, &((GD*)p)[i]); }
__device__ __forceinline__ void std_nt(double* p, uint32_t i, double v) {
    __builtin_nontemporal_store(v, (GD*)((GC*)p + (i << 3)));
}

__device__ __forceinline__ uint32_t ldu(const uint32_t* p, int64_t i) { return ((GConstU*)p)[i]; }
__device__ __forceinline__ uint32_t ldu(const uint32_t* p, uint32_t i) {
    return *(GConstU*)((GConstC*)p + (i << 2));
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

// The six face neighbours of cell (i, j, k) in the order x-, x+, y-, y+,
// z-, z+, with the reference's wrap (used or not, every index is in
// bounds), and their cell codes.  Loaded whether or not a face is a
// surface, so that the loads do not wait for the cell's own code.
template <class I>
__device__ __forceinline__ void gather6(const P3 s3, const uint32_t* code, I i, I j, I k, I nx, I ny, I nz,
                                        double (&v)[6][3], uint32_t (&cd)[6]) {
    I nb[6];
    nb[0] = ((i == 0 ? nx - 1 : i - 1) * ny + j) * nz + k;
    nb[1] = ((i == nx - 1 ? 0 : i + 1) * ny + j) * nz + k;
    nb[2] = (i * ny + (j == 0 ? ny - 1 : j - 1)) * nz + k;
    nb[3] = (i * ny + (j == ny - 1 ? 0 : j + 1)) * nz + k;
    nb[4] = (i * ny + j) * nz + (k == 0 ? nz - 1 : k - 1);
    nb[5] = (i * ny + j) * nz + (k == nz - 1 ? 0 : k + 1);
#pragma unroll
    for (int f = 0; f < 6; f++) {
        v[f][0] = ldd(s3.p[0], nb[f]);
        v[f][1] = ldd(s3.p[1], nb[f]);
        v[f][2] = ldd(s3.p[2], nb[f]);
        cd[f] = ldu(code, nb[f]);
    }
}

// Face coefficient between a cell of material `mat` (constant A, self mean
// A_self) and a neighbour of code nc: the harmonic mean 2 A1 A2 / (A1 + A2)
// of magnetic_initialize.cpp:381-524, A_self for the same material.
template <class MP>
__device__ __forceinline__ double face_coef(MP mats, uint32_t mat, uint32_t nc, int sub, double A, double A_self) {
    const uint32_t nm = nc & kMatMask;
    if (nm == mat) return A_self;
    const double a2 = sub < 0 ? mats[nm - 1].Aex : mats[nm - 1].aAex[sub];
    return 2. * A * a2 / (A + a2);
}

// The stencil of get_m_spatial_derivative_local from the cell's store
// (c0, c1, c2) and its gathered neighbours; see gather6.
template <bool kDmi, class MP>
__device__ __forceinline__ Stencil stencil_core(double c0, double c1, double c2, const double (&v)[6][3],
                                                const uint32_t (&cd)[6], MP mats, uint32_t own,
                                                int surf_shift, int sub, double dx, double dy, double dz, double A,
                                                double A_self, double slope) {
    const uint32_t mat = own & kMatMask;
    // one reciprocal per spacing instead of a division per difference: each
    // quotient moves by at most an ulp against the reference's c/dx
    const double ix = 1. / dx, iy = 1. / dy, iz = 1. / dz;
    Stencil r;
    double bw[3], fw[3], ab, af;
    // x faces
    if ((own >> (surf_shift + kXB)) & 1u) {
        if (kDmi) { bw[0] = -slope * c2; bw[1] = 0.; bw[2] = slope * c0; }
        else { bw[0] = 0.; bw[1] = 0.; bw[2] = 0.; }
        ab = A;
    } else {
        bw[0] = (c0 - v[0][0]) * ix; bw[1] = (c1 - v[0][1]) * ix; bw[2] = (c2 - v[0][2]) * ix;
        ab = face_coef(mats, mat, cd[0], sub, A, A_self);
    }
    if ((own >> (surf_shift + kXF)) & 1u) {
        if (kDmi) { fw[0] = -slope * c2; fw[1] = 0.; fw[2] = slope * c0; }
        else { fw[0] = 0.; fw[1] = 0.; fw[2] = 0.; }
        af = A;
    } else {
        fw[0] = (v[1][0] - c0) * ix; fw[1] = (v[1][1] - c1) * ix; fw[2] = (v[1][2] - c2) * ix;
        af = face_coef(mats, mat, cd[1], sub, A, A_self);
    }
    r.d[0] = (bw[0] + fw[0]) * 0.5;
    r.d[2] = (bw[2] + fw[2]) * 0.5;
    for (int a = 0; a < 3; a++) r.ex[a] = af * (fw[a] * ix) + ab * (-bw[a] * ix);
    // y faces
    if ((own >> (surf_shift + kYB)) & 1u) {
        if (kDmi) { bw[0] = 0.; bw[1] = -slope * c2; bw[2] = slope * c1; }
        else { bw[0] = 0.; bw[1] = 0.; bw[2] = 0.; }
        ab = A;
    } else {
        bw[0] = (c0 - v[2][0]) * iy; bw[1] = (c1 - v[2][1]) * iy; bw[2] = (c2 - v[2][2]) * iy;
        ab = face_coef(mats, mat, cd[2], sub, A, A_self);
    }
    if ((own >> (surf_shift + kYF)) & 1u) {
        if (kDmi) { fw[0] = 0.; fw[1] = -slope * c2; fw[2] = slope * c1; }
        else { fw[0] = 0.; fw[1] = 0.; fw[2] = 0.; }
        af = A;
    } else {
        fw[0] = (v[3][0] - c0) * iy; fw[1] = (v[3][1] - c1) * iy; fw[2] = (v[3][2] - c2) * iy;
        af = face_coef(mats, mat, cd[3], sub, A, A_self);
    }
    r.d[1] = (bw[1] + fw[1]) * 0.5;
    r.d[3] = (bw[2] + fw[2]) * 0.5;
    for (int a = 0; a < 3; a++) r.ex[a] = r.ex[a] + af * (fw[a] * iy) + ab * (-bw[a] * iy);
    // z faces: zero slope at surfaces for both kinds
    if ((own >> (surf_shift + kZB)) & 1u) {
        bw[0] = 0.; bw[1] = 0.; bw[2] = 0.;
        ab = A;
    } else {
        bw[0] = (c0 - v[4][0]) * iz; bw[1] = (c1 - v[4][1]) * iz; bw[2] = (c2 - v[4][2]) * iz;
        ab = face_coef(mats, mat, cd[4], sub, A, A_self);
    }
    if ((own >> (surf_shift + kZF)) & 1u) {
        fw[0] = 0.; fw[1] = 0.; fw[2] = 0.;
        af = A;
    } else {
        fw[0] = (v[5][0] - c0) * iz; fw[1] = (v[5][1] - c1) * iz; fw[2] = (v[5][2] - c2) * iz;
        af = face_coef(mats, mat, cd[5], sub, A, A_self);
    }
    for (int a = 0; a < 3; a++) r.ex[a] = r.ex[a] + af * (fw[a] * iz) + ab * (-bw[a] * iz);
    return r;
}

template <bool kDmi, class I>
__device__ __forceinline__ Stencil stencil(const P3 s3, const uint32_t* code, const DevMaterial* mats,
                                           uint32_t own, int surf_shift, int sub, I i, I j, I k,
                                           I nx, I ny, I nz, double dx, double dy, double dz,
                                           double A, double A_self, double slope) {
    const I id = (i * ny + j) * nz + k;
    const double c0 = ldd(s3.p[0], id), c1 = ldd(s3.p[1], id), c2 = ldd(s3.p[2], id);
    double v[6][3];
    uint32_t cd[6];
    gather6<I>(s3, code, i, j, k, nx, ny, nz, v, cd);
    return stencil_core<kDmi>(c0, c1, c2, v, cd, mats, own, surf_shift, sub, dx, dy, dz, A, A_self, slope);
}

// Magnetoelastic field of a crystal-frame strain, in the global frame:
// f (B1 e_aa m_a + B2 (e_ab m_b + e_ac m_c)) with f = -2 (FM) or -1 (per
// AFM sublattice), B already divided by mu0 Ms (magnetic_solver.cpp:143-150,
// magnetic_AFM_solver.cpp:156-162).  Strain is Voigt xx yy zz yz xz xy.
// e0, when set, is a static strain added to e (exx = exxt0 + Dexx, the
// uploaded static strain beside the elastodynamics).
template <class I>
__device__ __forceinline__ void magnetoelastic(const double* const* e, const double* const* e0, I id, double f,
                                               double B1, double B2, double mx, double my, double mz, int rotate,
                                               const double T[3][3], double& h0, double& h1, double& h2) {
    double s[6] = {0., 0., 0., 0., 0., 0.};
    if (e0[0])
        for (int c = 0; c < 6; c++) s[c] = ldd(e0[c], id);
    const double exx = s[0] + ldd(e[0], id), eyy = s[1] + ldd(e[1], id), ezz = s[2] + ldd(e[2], id);
    const double eyz = s[3] + ldd(e[3], id), exz = s[4] + ldd(e[4], id), exy = s[5] + ldd(e[5], id);
    const double c0 = f * (B1 * (exx) * mx + B2 * ((exy) * my + (exz) * mz));
    const double c1 = f * (B1 * (eyy) * my + B2 * ((exy) * mx + (eyz) * mz));
    const double c2 = f * (B1 * (ezz) * mz + B2 * ((exz) * mx + (eyz) * my));
    if (rotate) rot(T, c0, c1, c2, h0, h1, h2);
    else { h0 = c0; h1 = c1; h2 = c2; }
}

}  // namespace

}  // namespace gf
