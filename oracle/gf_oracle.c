/*
 * gf_oracle.c — CPU restatement of the reference's FP64 RK4 hot path
 * (ferromagnet LLG, two-sublattice antiferromagnet LLG, FFT magnetostatics).
 *
 * TEST INFRASTRUCTURE ONLY: this is the checker.  Only tests/,
 * __graft_entry__.smoke() and the CPU-baseline leg of bench.py may load it;
 * the shipped library never links it.
 *
 * Each function names the reference file:line it restates.  Arithmetic is
 * written in the reference's own operation order (the same parenthesisation,
 * the same divisions) so that with -ffp-contract=off on x86-64 the values
 * are those the reference's serial build would produce.  The FFT is this
 * file's own radix-2 transform (cuFFT, which the reference calls, is absent
 * here; its unnormalised D2Z/Z2D semantics are restated).
 *
 * Pinning status of each piece is recorded in the test files and in the
 * oracle section of the design notes:
 *   - FFT: pinned against a naive DFT and numpy.fft (tests/test_oracle_fft.py)
 *   - Newell kernel: self-term sum rule, thin-film and cube limits
 *   - FM LLG stage: macrospin closed form, exchange spin-wave dispersion
 *   - AFM LLG stage: uniaxial AFMR frequency, exchange-free macrospin limit
 *   - RK4 protocol: fourth-order convergence
 */
#include "gf_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* Loops over cells, edges, faces and FFT lines run on all host cores
 * (OpenMP) once a grid holds GFO_OMP_MIN cells.  Every parallel loop writes
 * each element from one iteration only and keeps the per-element operation
 * order, so the results are bit-identical for any thread count; sums over
 * cells (averages, loss) stay serial. */
#ifndef GFO_OMP_MIN
#define GFO_OMP_MIN 32768
#endif

/* GFO_PLANT selects one deliberate mistake for the pin tests, which build
 * a variant of this file with it to show that their check fails on it:
 *   1 DMI surface slope with the wrong sign,  2 DMI slope D/A for D/(2A),
 *   3 Liao reading the history one step too old, 4 Liao coefficients of
 *   t1 and t2 swapped, 5 Mur half- and full-stage coefficients swapped,
 *   6 Mur half stages reading the base state for the stage store,
 *   7 elastic Mur half- and full-stage coefficients swapped, 8 the
 *   prescribed layer of the stage-2 update at t instead of t - dt/2, 9 the
 *   tangent-linear reference without the DMI slope's derivative in Aex.
 * The default build (0) plants nothing. */
#ifndef GFO_PLANT
#define GFO_PLANT 0
#endif
#if GFO_PLANT == 1
#define DMI_SLOPE(D, A) (-(D) / 2. / (A))
#elif GFO_PLANT == 2
#define DMI_SLOPE(D, A) ((D) / (A))
#else
#define DMI_SLOPE(D, A) ((D) / 2. / (A))
#endif

static const double MU0 = 1.256637061e-6;          /* magnetic_system.h:14 */
static const double PI = 3.14159265358979323846;   /* magnetic_system.h:16 */
static const double UNIT_E = 1.602176634e-19;      /* magnetic_system.h:15 */

struct gfo_ctx {
    int64_t nx, ny, nz, n;
    double dx, dy, dz;
    int per[3];
    uint32_t flags;
    double dt, time;
    int nmat;
    gf_material* mat;
    uint32_t* cell;
    double Tg2c[3][3], Tc2g[3][3];
    double Hext_stat[3], Hext_altr[3], H_altr_freq[3], Hext[3];
    double demag_fac[3];
    int npre; double* pre; int pre_idx;
    int if_FM_all, if_AFM_all;
    int64_t NFM, NAFM;
    /* FM (magnetic_system.h:109-133) */
    double *m[3], *ms[3], *k[4][3], *dm[3];
    unsigned char *surfYZ, *surfXZ, *surfXY;
    double *Aij[6];   /* xf, xb, yf, yb, zf, zb */
    /* AFM (magnetic_system.h:116-147) */
    double *m1[3], *m2[3], *ms1[3], *ms2[3], *k1[4][3], *k2[4][3], *dm1[3], *dm2[3];
    unsigned char *asurfYZ, *asurfXZ, *asurfXY;
    double *Aij1[6], *Aij2[6];
    /* FE (ferroelectric_system.h) */
    double *p[3], *q[3], *ps[3], *qs[3], *kp[4][3], *kq[4][3], *Es[3];
    unsigned char *fsurfYZ, *fsurfXZ, *fsurfXY;
    int if_FE_all; int64_t NFE;
    double Eext_stat[3], Eext_altr[3], E_altr_freq[3], Eext[3];
    int npreE; double* preE; int preE_idx;
    /* elastodynamics (elastic_system.h) */
    int if_elasto;
    double *u[3], *v[3], *us[3], *vs[3], *ku[4][3], *kv[4][3];
    double *force[3], *force_store[3];
    double *gYZ[3], *gXZ[3], *gXY[3];          /* du{x,y,z}/d{x,y,z} on faces */
    double *e_glb[6], *e_crt[6];               /* dynamic strain xx yy zz yz xz xy */
    double *e0_glb[6], *e0_crt[6];             /* dynamic eigenstrain */
    double *e0t0[6];                           /* static crystal-frame eigenstrain at t0 (input) */
    double *t1[3], *t2[3], *t3[3], *t4[3];     /* displacement history (Liao / Mur) */
    unsigned char *sfYZ, *sfXZ, *sfXY;         /* stress-free faces */
    double vs_long_1st, vs_tran_1st, vs_long_last, vs_tran_last;
    int source_type, source_z, source_comp, num_cycles;
    double sigma_gauss, amplitude_gauss, w_liao;
    /* statics */
    double *H[3], *Hnu[3];
    double *Hdrive[3], hdrive_freq;            /* shaped drive (GF_FLAG_HDRIVE), no reference counterpart */
    double precess_angle, precess_freq;        /* prescribed bottom layer (GF_FLAG_PRESCRIBED_M, read.cpp:101) */
    double er_homo;                            /* electrostatics (GF_FLAG_ELECTROSTATICS) */
    double *e_t0[6];                           /* static crystal-frame strain, ex*t0_crt */
    int t0_fixed;                              /* the static strain at t0 was solved for or uploaded */
    /* electrodynamics (EMdynamic_system.h): E on edges, H on faces */
    int if_em, periodic_all, pec_all, jp_on;
    int64_t nE[3], nH[3];
    double *Em[3], *Ems[3], *kE[4][3], *Et1[3];
    double *Hm[3], *Hms[3], *kH[4][3];
    double *Ecell[3], *Hcell[3], *Jf[3];
    double *Jp[3][3], *Jps[3][3], *kJ[4][3][3];   /* [pole][comp] */
    unsigned char *pec[3];
    gf_desc desc;                              /* scalars of the descriptor */
    double cf_top_half, cf_bot_half, cf_x_half, cf_y_half, cf_top, cf_bot, cf_x, cf_y;
    double *kspec[9];   /* complex spectra, interleaved, nx*ny*(nz/2+1) each */
    /* spin pumping (GF_FLAG_SPIN_PUMPING, global.cpp:169-235) */
    int spin_on;
    int64_t* spin_src;  /* source_z per plane, -1 = no J_ISHE */
    double* spin_coef;  /* coeff_Js * coeff_L per plane */
    double* Jis[3];     /* Jx_ISHE, Jy_ISHE, zero */
    /* flexoelectricity (GF_FLAG_FLEXO): per material F11, F12, F44,
     * tv4..tv8, f11, f12, f44 (read.cpp:181-199, zero unless FE) */
    int flexo_on;
    double (*flx)[11];
    /* elastostatics (GF_FLAG_ELASTOSTATIC): homogeneous c11, c12, c44 */
    double ch11, ch12, ch44;
};

static void elastic_gradients(gfo_ctx* c);
static void elastic_strain(gfo_ctx* c);
static int em_setup(gfo_ctx* c, const gf_desc* d);
static void em_cells(gfo_ctx* c, int which);

/* ------------------------------------------------------------------ FFT */
/* Mixed-radix decimation in time for lengths that are not powers of two
 * (cuFFT plans any size, magnetic_initialize.cpp:114-115): the smallest
 * prime factor p splits x (stride `stride`) into p interleaved sequences
 * of length m = n/p, each transformed recursively into y[q*m ..], then
 * X[k + s m] = sum_q exp(-+2 pi i q (k + s m) / n) Y_q[k], the angle taken
 * from the index reduced mod n. */
static void dft_rec(const double* xr, const double* xi, int64_t stride, int64_t n, double* yr, double* yi,
                    int inverse) {
    if (n == 1) { yr[0] = xr[0]; yi[0] = xi[0]; return; }
    int64_t p = 2;
    while (n % p) p++;
    const int64_t m = n / p;
    double* tr = (double*)malloc(sizeof(double) * (size_t)n);
    double* ti = (double*)malloc(sizeof(double) * (size_t)n);
    for (int64_t q = 0; q < p; q++) dft_rec(xr + q * stride, xi + q * stride, stride * p, m, tr + q * m, ti + q * m, inverse);
    for (int64_t k = 0; k < m; k++)
        for (int64_t s = 0; s < p; s++) {
            double ar = 0.0, ai = 0.0;
            for (int64_t q = 0; q < p; q++) {
                const int64_t t = (q * (k + s * m)) % n;
                const double ang = 2.0 * PI * (double)t / (double)n;
                const double wr = cos(ang), wi = inverse ? sin(ang) : -sin(ang);
                ar += tr[q * m + k] * wr - ti[q * m + k] * wi;
                ai += tr[q * m + k] * wi + ti[q * m + k] * wr;
            }
            yr[k + s * m] = ar;
            yi[k + s * m] = ai;
        }
    free(tr);
    free(ti);
}

static int is_pow2(int64_t n) { return n >= 1 && (n & (n - 1)) == 0; }

/* In-place complex FFT, unnormalised, sign -1 forward: iterative radix 2
 * for powers of two, dft_rec otherwise. */
static void fft1(double* re, double* im, int64_t n, int inverse) {
    if (n <= 1) return;
    if (!is_pow2(n)) {
        double* yr = (double*)malloc(sizeof(double) * (size_t)n);
        double* yi = (double*)malloc(sizeof(double) * (size_t)n);
        dft_rec(re, im, 1, n, yr, yi, inverse);
        memcpy(re, yr, sizeof(double) * (size_t)n);
        memcpy(im, yi, sizeof(double) * (size_t)n);
        free(yr);
        free(yi);
        return;
    }
    for (int64_t i = 1, j = 0; i < n; i++) {
        int64_t bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) {
            double t = re[i]; re[i] = re[j]; re[j] = t;
            t = im[i]; im[i] = im[j]; im[j] = t;
        }
    }
    for (int64_t len = 2; len <= n; len <<= 1) {
        int64_t h = len >> 1;
        for (int64_t j = 0; j < h; j++) {
            /* twiddle exp(-+2 pi i j / len) from the table angle directly */
            double ang = 2.0 * PI * (double)j / (double)len;
            double wr = cos(ang), wi = inverse ? sin(ang) : -sin(ang);
            for (int64_t s = 0; s < n; s += len) {
                int64_t a = s + j, b = s + j + h;
                double vr = re[b] * wr - im[b] * wi;
                double vi = re[b] * wi + im[b] * wr;
                re[b] = re[a] - vr; im[b] = im[a] - vi;
                re[a] = re[a] + vr; im[a] = im[a] + vi;
            }
        }
    }
}

/* cuFFT D2Z semantics (magnetostatic_solver.cpp:149-151): unnormalised
 * forward transform over (nx,ny,nz), row-major, half spectrum along z. */
void gfo_rfft3(int64_t nx, int64_t ny, int64_t nz, const double* in, double* out) {
    int64_t nzh = nz / 2 + 1;
    int64_t mx = nx > ny ? nx : ny; if (nz > mx) mx = nz;
    /* lines are independent: each thread transforms whole lines with its
     * own scratch, so the result does not depend on the thread count */
    #pragma omp parallel if (nx * ny * nz >= GFO_OMP_MIN)
    {
    double* re = (double*)malloc(sizeof(double) * mx);
    double* im = (double*)malloc(sizeof(double) * mx);
    #pragma omp for collapse(2) schedule(static)
    for (int64_t i = 0; i < nx; i++)
        for (int64_t j = 0; j < ny; j++) {
            for (int64_t k = 0; k < nz; k++) { re[k] = in[(i * ny + j) * nz + k]; im[k] = 0.0; }
            fft1(re, im, nz, 0);
            for (int64_t k = 0; k < nzh; k++) {
                out[2 * ((i * ny + j) * nzh + k)] = re[k];
                out[2 * ((i * ny + j) * nzh + k) + 1] = im[k];
            }
        }
    #pragma omp for collapse(2) schedule(static)
    for (int64_t i = 0; i < nx; i++)
        for (int64_t k = 0; k < nzh; k++) {
            for (int64_t j = 0; j < ny; j++) { re[j] = out[2 * ((i * ny + j) * nzh + k)]; im[j] = out[2 * ((i * ny + j) * nzh + k) + 1]; }
            fft1(re, im, ny, 0);
            for (int64_t j = 0; j < ny; j++) { out[2 * ((i * ny + j) * nzh + k)] = re[j]; out[2 * ((i * ny + j) * nzh + k) + 1] = im[j]; }
        }
    #pragma omp for collapse(2) schedule(static)
    for (int64_t j = 0; j < ny; j++)
        for (int64_t k = 0; k < nzh; k++) {
            for (int64_t i = 0; i < nx; i++) { re[i] = out[2 * ((i * ny + j) * nzh + k)]; im[i] = out[2 * ((i * ny + j) * nzh + k) + 1]; }
            fft1(re, im, nx, 0);
            for (int64_t i = 0; i < nx; i++) { out[2 * ((i * ny + j) * nzh + k)] = re[i]; out[2 * ((i * ny + j) * nzh + k) + 1] = im[i]; }
        }
    free(re); free(im);
    }
}

/* cuFFT Z2D semantics (magnetostatic_solver.cpp:183-185): unnormalised
 * inverse of the half spectrum; x and y inverted first, then z by
 * Hermitian extension, imaginary parts of the self-conjugate bins unused. */
void gfo_irfft3(int64_t nx, int64_t ny, int64_t nz, const double* in, double* out) {
    int64_t nzh = nz / 2 + 1;
    int64_t mx = nx > ny ? nx : ny; if (nz > mx) mx = nz;
    double* tmp = (double*)malloc(sizeof(double) * 2 * nx * ny * nzh);
    memcpy(tmp, in, sizeof(double) * 2 * nx * ny * nzh);
    #pragma omp parallel if (nx * ny * nz >= GFO_OMP_MIN)
    {
    double* re = (double*)malloc(sizeof(double) * mx);
    double* im = (double*)malloc(sizeof(double) * mx);
    #pragma omp for collapse(2) schedule(static)
    for (int64_t j = 0; j < ny; j++)
        for (int64_t k = 0; k < nzh; k++) {
            for (int64_t i = 0; i < nx; i++) { re[i] = tmp[2 * ((i * ny + j) * nzh + k)]; im[i] = tmp[2 * ((i * ny + j) * nzh + k) + 1]; }
            fft1(re, im, nx, 1);
            for (int64_t i = 0; i < nx; i++) { tmp[2 * ((i * ny + j) * nzh + k)] = re[i]; tmp[2 * ((i * ny + j) * nzh + k) + 1] = im[i]; }
        }
    #pragma omp for collapse(2) schedule(static)
    for (int64_t i = 0; i < nx; i++)
        for (int64_t k = 0; k < nzh; k++) {
            for (int64_t j = 0; j < ny; j++) { re[j] = tmp[2 * ((i * ny + j) * nzh + k)]; im[j] = tmp[2 * ((i * ny + j) * nzh + k) + 1]; }
            fft1(re, im, ny, 1);
            for (int64_t j = 0; j < ny; j++) { tmp[2 * ((i * ny + j) * nzh + k)] = re[j]; tmp[2 * ((i * ny + j) * nzh + k) + 1] = im[j]; }
        }
    #pragma omp for collapse(2) schedule(static)
    for (int64_t i = 0; i < nx; i++)
        for (int64_t j = 0; j < ny; j++) {
            const double* row = tmp + 2 * ((i * ny + j) * nzh);
            for (int64_t k = 0; k < nz; k++) {
                if (k < nzh) { re[k] = row[2 * k]; im[k] = row[2 * k + 1]; }
                else { re[k] = row[2 * (nz - k)]; im[k] = -row[2 * (nz - k) + 1]; }
            }
            fft1(re, im, nz, 1);
            for (int64_t k = 0; k < nz; k++) out[(i * ny + j) * nz + k] = re[k];
        }
    free(re); free(im);
    }
    free(tmp);
}

void gfo_rdft3_naive(int64_t nx, int64_t ny, int64_t nz, const double* in, double* out) {
    int64_t nzh = nz / 2 + 1;
    for (int64_t a = 0; a < nx; a++)
        for (int64_t b = 0; b < ny; b++)
            for (int64_t c = 0; c < nzh; c++) {
                long double sr = 0.0L, si = 0.0L;
                for (int64_t i = 0; i < nx; i++)
                    for (int64_t j = 0; j < ny; j++)
                        for (int64_t k = 0; k < nz; k++) {
                            long double ph = -2.0L * 3.14159265358979323846264338327950288L *
                                ((long double)(a * i % nx) / nx + (long double)(b * j % ny) / ny +
                                 (long double)(c * k % nz) / nz);
                            long double v = in[(i * ny + j) * nz + k];
                            sr += v * cosl(ph); si += v * sinl(ph);
                        }
                out[2 * ((a * ny + b) * nzh + c)] = (double)sr;
                out[2 * ((a * ny + b) * nzh + c) + 1] = (double)si;
            }
}

/* --------------------------------------------------------- Newell kernel */
/* magnetic_initialize.cpp:1427-1476 */
double gfo_newell_F2(double x, double y, double z) {
    double f2 = 0.;
    double sigma = 1.0e-33;
    double u = x * x, v = y * y, w = z * z;
    double r = sqrt(u + v + w);
    if (fabs(x) > sigma && fabs(y) > sigma && fabs(z) > sigma) {
        f2 = z * (u - v) * log(r - z) / 2.
           + y * (u - w) * log(r - y) / 2.
           - x * y * z * atan(y * z / (x * r))
           + (2. * u - v - w) * r / 6.;
    } else if (fabs(x) < sigma && fabs(y) > sigma && fabs(z) > sigma) {
        f2 = -(z * v / 2.) * log(r - z) - (y * w / 2.) * log(r - y) - r * r * r / 6.;
    } else if (fabs(x) > sigma && fabs(y) < sigma && fabs(z) > sigma) {
        f2 = (z * u / 2.) * log(r - z) + (2. * u - w) * r / 6.;
    } else if (fabs(x) > sigma && fabs(y) > sigma && fabs(z) < sigma) {
        f2 = (y * u / 2.) * log(r - y) + (2. * u - v) * r / 6.;
    } else if (fabs(x) < sigma && fabs(y) < sigma && fabs(z) > sigma) {
        f2 = -fabs(z) * fabs(z) * fabs(z) / 6.;
    } else if (fabs(x) < sigma && fabs(y) > sigma && fabs(z) < sigma) {
        f2 = -fabs(y) * fabs(y) * fabs(y) / 6.;
    } else if (fabs(x) > sigma && fabs(y) < sigma && fabs(z) < sigma) {
        f2 = fabs(x) * fabs(x) * fabs(x) / 3.;
    } else if (fabs(x) < sigma && fabs(y) < sigma && fabs(z) < sigma) {
        f2 = 0.;
    }
    return f2;
}

/* magnetic_initialize.cpp:1419-1425 */
static double newell_F1(double x, double y, double z, double dx, double dy, double dz) {
    (void)dx; (void)dz;
    double ym = y - dy, yp = y + dy;
    return 2. * gfo_newell_F2(x, y, z) - gfo_newell_F2(x, ym, z) - gfo_newell_F2(x, yp, z);
}

/* magnetic_initialize.cpp:1404-1417 */
double gfo_newell_A(double x, double y, double z, double dx, double dy, double dz) {
    double xp = x + dx, xm = x - dx, zp = z + dz, zm = z - dz;
    return 4. * newell_F1(x, y, z, dx, dy, dz) - 2. * newell_F1(x, y, zp, dx, dy, dz)
         - 2. * newell_F1(x, y, zm, dx, dy, dz) - 2. * newell_F1(xp, y, z, dx, dy, dz)
         + newell_F1(xp, y, zp, dx, dy, dz)
         + newell_F1(xp, y, zm, dx, dy, dz)
         - 2. * newell_F1(xm, y, z, dx, dy, dz) + newell_F1(xm, y, zp, dx, dy, dz) + newell_F1(xm, y, zm, dx, dy, dz);
}

/* magnetic_initialize.cpp:1507-1557 */
double gfo_newell_G2(double x, double y, double z) {
    double g2 = 0.;
    double sigma = 1.0e-33;
    double u = x * x, v = y * y, w = z * z;
    double r = sqrt(u + v + w);
    if (fabs(x * y * z) > sigma) {
        g2 = (y * y * y / 6. - y * w / 2.) * log(r + x)
           + (x / 6.) * (3. * w - u) * log(r - y)
           + x * y * r / 3. + x * y * z * log(r - z)
           + (v * z / 2.) * atan(x * z / (r * y))
           + (z * z * z / 6.) * atan(x * y / (r * z))
           + (u * z / 2.) * atan(y * z / (r * x));
    } else if (fabs(x) < sigma && fabs(y * z) > sigma) {
        g2 = (y * y * y / 6. - y * w / 2.) * log(r);
    } else if (fabs(x * z) > sigma && fabs(y) < sigma) {
        g2 = (x * w / 2. - x * x * x / 6.) * log(r);
    } else if (fabs(x * y) > sigma && fabs(z) < sigma) {
        g2 = ((y * y * y) / 6.) * log(r + x) - (x * x * x / 6.) * log(r - y) + x * y * r / 3.;
    } else if ((fabs(x) + fabs(y)) < sigma && fabs(z) > sigma) {
        g2 = 0.;
    } else if ((fabs(x) + fabs(z)) < sigma && fabs(y) > sigma) {
        g2 = (y * y * y) / 12. * log(v);
    } else if (fabs(x) > sigma && (fabs(y) + fabs(z)) < sigma) {
        g2 = -((x * x * x) / 12.) * log(u);
    } else if (fabs(x) < sigma && fabs(y) < sigma && fabs(z) < sigma) {
        g2 = 0.;
    }
    return g2;
}

/* magnetic_initialize.cpp:1486-1505 */
static double newell_G1(double x, double y, double z, double dx, double dy, double dz) {
    double xp = x + dx, xm = x - dx, yp = y + dy / 2., ym = y - dy / 2., zp = z + dz, zm = z - dz;
    return 4. * gfo_newell_G2(x, yp, z) - 4. * gfo_newell_G2(x, ym, z)
         - 2. * gfo_newell_G2(x, yp, zp) + 2. * gfo_newell_G2(x, ym, zp)
         - 2. * gfo_newell_G2(x, yp, zm) + 2. * gfo_newell_G2(x, ym, zm)
         - 2. * gfo_newell_G2(xm, yp, z) + 2. * gfo_newell_G2(xm, ym, z)
         + gfo_newell_G2(xm, yp, zp) - gfo_newell_G2(xm, ym, zp)
         + gfo_newell_G2(xm, yp, zm) - gfo_newell_G2(xm, ym, zm)
         - 2. * gfo_newell_G2(xp, yp, z) + 2. * gfo_newell_G2(xp, ym, z)
         + gfo_newell_G2(xp, yp, zp) - gfo_newell_G2(xp, ym, zp)
         + gfo_newell_G2(xp, yp, zm) - gfo_newell_G2(xp, ym, zm);
}

/* magnetic_initialize.cpp:1478-1484 */
double gfo_newell_B(double x, double y, double z, double dx, double dy, double dz) {
    double yp = y + dy / 2., ym = y - dy / 2.;
    return newell_G1(x, yp, z, dx, dy, dz) - newell_G1(x, ym, z, dx, dy, dz);
}

/* magnetic_initialize.cpp:1176-1373: the nine tensor kernels evaluated on
 * offsets |i| <= n/2-1 per axis, placed circularly (offset d at d mod n,
 * index n/2 left zero) and divided by n.  Order of out9:
 * Axyz, Ayzx, Azxy, Bxyz, Byzx, Bzxy, Byxz, Bzyx, Bxzy. */
void gfo_newell_real(int64_t nx, int64_t ny, int64_t nz, double dx, double dy, double dz, double* out9) {
    int64_t n = nx * ny * nz;
    double twopi = 2. * PI;
    memset(out9, 0, sizeof(double) * 9 * n);
    #pragma omp parallel for collapse(3) schedule(dynamic, 256) if (n >= GFO_OMP_MIN)
    for (int64_t i = 1 - nx / 2; i < nx / 2; i++)
        for (int64_t j = 1 - ny / 2; j < ny / 2; j++)
            for (int64_t k = 1 - nz / 2; k < nz / 2; k++) {
                double x = (double)i * dx, y = (double)j * dy, z = (double)k * dz;
                int64_t pi_ = (i + nx) % nx, pj = (j + ny) % ny, pk = (k + nz) % nz;
                int64_t id = (pi_ * ny + pj) * nz + pk;
                out9[0 * n + id] = gfo_newell_A(x, y, z, dx, dy, dz) / (2. * twopi);
                out9[1 * n + id] = gfo_newell_A(y, z, x, dy, dz, dx) / (2. * twopi);
                out9[2 * n + id] = gfo_newell_A(z, x, y, dz, dx, dy) / (2. * twopi);
                out9[3 * n + id] = gfo_newell_B(x, y, z, dx, dy, dz) / (2. * twopi);
                out9[4 * n + id] = gfo_newell_B(y, z, x, dy, dz, dx) / (2. * twopi);
                out9[5 * n + id] = gfo_newell_B(z, x, y, dz, dx, dy) / (2. * twopi);
                out9[6 * n + id] = gfo_newell_B(y, x, z, dy, dx, dz) / (2. * twopi);
                out9[7 * n + id] = gfo_newell_B(z, y, x, dz, dy, dx) / (2. * twopi);
                out9[8 * n + id] = gfo_newell_B(x, z, y, dx, dz, dy) / (2. * twopi);
            }
    for (int64_t s = 0; s < 9; s++)
        for (int64_t id = 0; id < n; id++) out9[s * n + id] = out9[s * n + id] / (double)n;
}

/* ------------------------------------------------------------ helpers */
static double* dalloc(int64_t n) { return (double*)calloc((size_t)(n > 0 ? n : 1), sizeof(double)); }
static int64_t ID(const gfo_ctx* c, int64_t i, int64_t j, int64_t k) { return (i * c->ny + j) * c->nz + k; }
static const gf_material* matof(const gfo_ctx* c, int64_t id) {
    uint32_t t = c->cell[id];
    return t ? &c->mat[t - 1] : NULL;
}
static int isFM(const gfo_ctx* c, int64_t id) { const gf_material* m = matof(c, id); return m && m->if_FM; }
static int isAFM(const gfo_ctx* c, int64_t id) { const gf_material* m = matof(c, id); return m && m->if_AFM; }
static int isFE(const gfo_ctx* c, int64_t id) { const gf_material* m = matof(c, id); return m && m->if_FE; }

/* mathlib.cpp:46-75 */
static void glb2crt(const gfo_ctx* c, double ix, double iy, double iz, double* ox, double* oy, double* oz) {
    if (c->flags & GF_FLAG_ORIENT_ROTATE) {
        *ox = c->Tg2c[0][0] * ix + c->Tg2c[0][1] * iy + c->Tg2c[0][2] * iz;
        *oy = c->Tg2c[1][0] * ix + c->Tg2c[1][1] * iy + c->Tg2c[1][2] * iz;
        *oz = c->Tg2c[2][0] * ix + c->Tg2c[2][1] * iy + c->Tg2c[2][2] * iz;
    } else { *ox = ix; *oy = iy; *oz = iz; }
}
static void crt2glb(const gfo_ctx* c, double ix, double iy, double iz, double* ox, double* oy, double* oz) {
    if (c->flags & GF_FLAG_ORIENT_ROTATE) {
        *ox = c->Tc2g[0][0] * ix + c->Tc2g[0][1] * iy + c->Tc2g[0][2] * iz;
        *oy = c->Tc2g[1][0] * ix + c->Tc2g[1][1] * iy + c->Tc2g[1][2] * iz;
        *oz = c->Tc2g[2][0] * ix + c->Tc2g[2][1] * iy + c->Tc2g[2][2] * iz;
    } else { *ox = ix; *oy = iy; *oz = iz; }
}

/* Interface masks, magnetic_initialize.cpp:136-379 (FM) and :635-~870
 * (AFM): a face is a surface when exactly one side is magnetic of the
 * kind, or at a non-periodic wall next to a magnetic cell. */
static void build_masks(gfo_ctx* c, int (*pred)(const gfo_ctx*, int64_t),
                        unsigned char* sYZ, unsigned char* sXZ, unsigned char* sXY) {
    int64_t nx = c->nx, ny = c->ny, nz = c->nz;
    for (int64_t j = 0; j < ny; j++)
        for (int64_t k = 0; k < nz; k++) {
            int f1 = pred(c, ID(c, 0, j, k)), f2 = pred(c, ID(c, nx - 1, j, k));
            if (c->per[0]) { sYZ[(0 * ny + j) * nz + k] = sYZ[(nx * ny + j) * nz + k] = (unsigned char)(f1 ^ f2); }
            else { sYZ[(0 * ny + j) * nz + k] = (unsigned char)f1; sYZ[(nx * ny + j) * nz + k] = (unsigned char)f2; }
        }
    for (int64_t i = 1; i < nx; i++)
        for (int64_t j = 0; j < ny; j++)
            for (int64_t k = 0; k < nz; k++)
                sYZ[(i * ny + j) * nz + k] = (unsigned char)(pred(c, ID(c, i - 1, j, k)) ^ pred(c, ID(c, i, j, k)));
    for (int64_t i = 0; i < nx; i++)
        for (int64_t k = 0; k < nz; k++) {
            int f1 = pred(c, ID(c, i, 0, k)), f2 = pred(c, ID(c, i, ny - 1, k));
            if (c->per[1]) { sXZ[(i * (ny + 1) + 0) * nz + k] = sXZ[(i * (ny + 1) + ny) * nz + k] = (unsigned char)(f1 ^ f2); }
            else { sXZ[(i * (ny + 1) + 0) * nz + k] = (unsigned char)f1; sXZ[(i * (ny + 1) + ny) * nz + k] = (unsigned char)f2; }
        }
    for (int64_t i = 0; i < nx; i++)
        for (int64_t j = 1; j < ny; j++)
            for (int64_t k = 0; k < nz; k++)
                sXZ[(i * (ny + 1) + j) * nz + k] = (unsigned char)(pred(c, ID(c, i, j - 1, k)) ^ pred(c, ID(c, i, j, k)));
    for (int64_t i = 0; i < nx; i++)
        for (int64_t j = 0; j < ny; j++) {
            int f1 = pred(c, ID(c, i, j, 0)), f2 = pred(c, ID(c, i, j, nz - 1));
            if (c->per[2]) { sXY[(i * ny + j) * (nz + 1) + 0] = sXY[(i * ny + j) * (nz + 1) + nz] = (unsigned char)(f1 ^ f2); }
            else { sXY[(i * ny + j) * (nz + 1) + 0] = (unsigned char)f1; sXY[(i * ny + j) * (nz + 1) + nz] = (unsigned char)f2; }
        }
    for (int64_t i = 0; i < nx; i++)
        for (int64_t j = 0; j < ny; j++)
            for (int64_t k = 1; k < nz; k++)
                sXY[(i * ny + j) * (nz + 1) + k] = (unsigned char)(pred(c, ID(c, i, j, k - 1)) ^ pred(c, ID(c, i, j, k)));
}

static int sYZ(const unsigned char* s, const gfo_ctx* c, int64_t i, int64_t j, int64_t k) { return s[(i * c->ny + j) * c->nz + k]; }
static int sXZ(const unsigned char* s, const gfo_ctx* c, int64_t i, int64_t j, int64_t k) { return s[(i * (c->ny + 1) + j) * c->nz + k]; }
static int sXY(const unsigned char* s, const gfo_ctx* c, int64_t i, int64_t j, int64_t k) { return s[(i * c->ny + j) * (c->nz + 1) + k]; }

static double hmean(double a1, double a2) { return 2. * a1 * a2 / (a1 + a2); }

/* Face exchange constants, magnetic_initialize.cpp:381-524 (FM) and
 * :870-~1040 (AFM); sel: 0 = FM Aex, 1 = Aex_AFM1, 2 = Aex_AFM2. */
static double aex_of(const gf_material* m, int sel) { return sel == 0 ? m->Aex : (sel == 1 ? m->Aex_AFM1 : m->Aex_AFM2); }
static void build_aij(gfo_ctx* c, int (*pred)(const gfo_ctx*, int64_t), const unsigned char* mYZ,
                      const unsigned char* mXZ, const unsigned char* mXY, int sel, double** A) {
    int64_t nx = c->nx, ny = c->ny, nz = c->nz;
    for (int64_t id = 0; id < c->n; id++) {
        if (!pred(c, id)) continue;
        const gf_material* m = matof(c, id);
        int64_t x = id / (ny * nz), y = (id - x * (ny * nz)) / nz, z = id - x * (ny * nz) - y * nz;
        double a = aex_of(m, sel);
        A[0][id] = sYZ(mYZ, c, x + 1, y, z) ? a : hmean(a, aex_of(matof(c, ID(c, x == nx - 1 ? 0 : x + 1, y, z)), sel));
        A[1][id] = sYZ(mYZ, c, x, y, z) ? a : hmean(a, aex_of(matof(c, ID(c, x == 0 ? nx - 1 : x - 1, y, z)), sel));
        A[2][id] = sXZ(mXZ, c, x, y + 1, z) ? a : hmean(a, aex_of(matof(c, ID(c, x, y == ny - 1 ? 0 : y + 1, z)), sel));
        A[3][id] = sXZ(mXZ, c, x, y, z) ? a : hmean(a, aex_of(matof(c, ID(c, x, y == 0 ? ny - 1 : y - 1, z)), sel));
        A[4][id] = sXY(mXY, c, x, y, z + 1) ? a : hmean(a, aex_of(matof(c, ID(c, x, y, z == nz - 1 ? 0 : z + 1)), sel));
        A[5][id] = sXY(mXY, c, x, y, z) ? a : hmean(a, aex_of(matof(c, ID(c, x, y, z == 0 ? nz - 1 : z - 1)), sel));
    }
}

/* ----------------------------------------------------------- create */
/* The spin-pumping planes (global.cpp:169-235): the layers bottom to top
 * (d->nz_layer, else the runs of equal ids along z of column (0, 0)); a
 * layer whose material (at column (0, 0)) pumps, with a magnet (FM or AFM)
 * on exactly one side, carries J_ISHE on each of its planes k, with the
 * source plane sz next to it and the layer thickness t.  Per plane, the
 * coefficient of get_J_ISHE_RK1 (magnetic_solver.cpp:245-259):
 * theta_SH e / (2 pi) g_mix times +-sinh((t - d) / l_sd) / sinh(t / l_sd),
 * d = (|k - sz| - 1/2) dz, + when the source lies below. */
static int spin_setup(gfo_ctx* c, const gf_desc* d) {
    int64_t nz = c->nz;
    c->spin_on = 1;
    c->spin_src = (int64_t*)malloc(sizeof(int64_t) * (size_t)nz);
    c->spin_coef = (double*)calloc((size_t)nz, sizeof(double));
    double* thick = (double*)calloc((size_t)nz, sizeof(double));
    for (int64_t k = 0; k < nz; k++) c->spin_src[k] = -1;
    for (int a = 0; a < 3; a++) c->Jis[a] = dalloc(c->n);
    int64_t nl = 0;
    int64_t* lay = (int64_t*)malloc(sizeof(int64_t) * (size_t)nz);
    if (d->n_layers > 0) {
        int64_t tot = 0;
        for (int i = 0; i < d->n_layers; i++) {
            if (d->nz_layer[i] < 1 || nl >= nz) { free(lay); free(thick); return 1; }
            lay[nl++] = d->nz_layer[i];
            tot += d->nz_layer[i];
        }
        if (tot > nz) { free(lay); free(thick); return 1; }
    } else {
        for (int64_t k = 0; k < nz;) {
            int64_t e = k + 1;
            while (e < nz && c->cell[e] == c->cell[k]) e++;
            lay[nl++] = e - k;
            k = e;
        }
    }
    int64_t zf = -1;
    for (int64_t l = 0; l < nl; l++) {
        int64_t zi = zf + 1;
        zf = zi + lay[l] - 1;
        uint32_t t = c->cell[zi];          /* column (0, 0): index k */
        if (!t || !c->mat[t - 1].if_spin_pump) continue;
        int src_i = 0, src_f = 0;
        if (zi != 0) {
            uint32_t ti = c->cell[zi - 1];
            src_i = ti && (c->mat[ti - 1].if_FM || c->mat[ti - 1].if_AFM);
        }
        if (zf != nz - 1) {
            uint32_t tf = c->cell[zf + 1];
            src_f = tf && (c->mat[tf - 1].if_FM || c->mat[tf - 1].if_AFM);
        }
        int64_t sz = -1;
        if (src_i && !src_f) sz = zi - 1;
        else if (!src_i && src_f) sz = zf + 1;
        if (sz < 0) continue;
        for (int64_t k = zi; k <= zf; k++) {
            c->spin_src[k] = sz;
            thick[k] = (double)lay[l] * c->dz;
        }
    }
    for (int64_t k = 0; k < nz; k++) {
        if (c->spin_src[k] < 0 || !c->cell[k]) continue;
        const gf_material* mat = &c->mat[c->cell[k] - 1];
        double cjs = mat->spin_hall_angle * UNIT_E / 2. / PI * mat->spin_mix_cond;
        double dist = (fabs((double)(k - c->spin_src[k])) - 0.5) * c->dz;
        double cl = sinh((thick[k] - dist) / mat->spin_diffus_length) / sinh(thick[k] / mat->spin_diffus_length);
        if (!(k > c->spin_src[k])) cl = -1. * cl;
        c->spin_coef[k] = cjs * cl;
    }
    free(lay);
    free(thick);
    return 0;
}

/* get_J_ISHE_RK* (magnetic_solver.cpp:231-283 and its stage copies): the
 * current of each pumping plane from the stage store and the stage
 * increment of its source plane, FM then both AFM sublattices. */
/* J_ISHE at the source planes from the given magnetisations and
 * increments (FM, AFM1, AFM2; NULL when absent), the bracket of
 * magnetic_solver.cpp:245-259 evaluated left to right. */
static void spin_current(gfo_ctx* c, double* const* fm, double* const* fd, double* const* s1, double* const* d1,
                         double* const* s2, double* const* d2) {
    int64_t ny = c->ny, nz = c->nz;
    for (int64_t k = 0; k < nz; k++) {
        int64_t sz = c->spin_src[k];
        if (sz < 0) continue;
        for (int64_t i = 0; i < c->nx; i++)
            for (int64_t j = 0; j < ny; j++) {
                int64_t id = (i * ny + j) * nz + k, is = (i * ny + j) * nz + sz;
                double sx = 0., sy = 0.;
                if (fm) {
                    sx = fm[0][is] * fd[2][is] - fm[2][is] * fd[0][is];
                    sy = fm[1][is] * fd[2][is] - fm[2][is] * fd[1][is];
                }
                if (s1) {
                    sx = sx + s1[0][is] * d1[2][is] - s1[2][is] * d1[0][is] + s2[0][is] * d2[2][is]
                         - s2[2][is] * d2[0][is];
                    sy = sy + s1[1][is] * d1[2][is] - s1[2][is] * d1[1][is] + s2[1][is] * d2[2][is]
                         - s2[2][is] * d2[1][is];
                }
                c->Jis[0][id] = c->spin_coef[k] * sx / c->dt;
                c->Jis[1][id] = c->spin_coef[k] * sy / c->dt;
            }
    }
}

/* get_J_ISHE_RKs (magnetic_solver.cpp:230-283 and its RK2-4 copies): the
 * stage stores and the stage increments */
static void spin_pump(gfo_ctx* c, int stage) {
    spin_current(c, c->if_FM_all ? c->ms : NULL, c->k[stage - 1], c->if_AFM_all ? c->ms1 : NULL,
                 c->k1[stage - 1], c->ms2, c->k2[stage - 1]);
}

/* get_J_ISHE (magnetic_solver.cpp:1460-1511), called by update_m and
 * update_m_AFM before their m += dm (:1593-1597, magnetic_AFM_solver.cpp:
 * 1801-1803): the states before the update and the RK-combined increments.
 * With both magnets the AFM call comes second, after the FM state has been
 * updated, and its value stands. */
static void spin_pump_step(gfo_ctx* c) {
    spin_current(c, c->if_FM_all ? c->m : NULL, c->dm, c->if_AFM_all ? c->m1 : NULL, c->dm1, c->m2, c->dm2);
}

gfo_ctx* gfo_create(const gf_desc* d) {
    if (!d || d->nx < 1 || d->ny < 1 || d->nz < 1) return NULL;
    gfo_ctx* c = (gfo_ctx*)calloc(1, sizeof(gfo_ctx));
    c->nx = d->nx; c->ny = d->ny; c->nz = d->nz; c->n = d->nx * d->ny * d->nz;
    c->dx = d->dx; c->dy = d->dy; c->dz = d->dz;
    for (int a = 0; a < 3; a++) c->per[a] = d->periodic[a] != 0;
    c->flags = d->flags; c->dt = d->dt; c->time = 0.;
    c->hdrive_freq = d->Hdrive_freq;
    c->precess_angle = d->precess_angle;
    c->precess_freq = d->precess_frequency;
    /* er_homo = max r_permittivity * scale_elec (ferroelectric_initialize.cpp:311-318) */
    c->er_homo = 0.;
    for (int t = 0; t < d->n_materials; t++)
        if (d->materials[t].r_permittivity > c->er_homo) c->er_homo = d->materials[t].r_permittivity;
    c->er_homo = c->er_homo * (d->scale_elec > 0. ? d->scale_elec : 1.);
    c->nmat = d->n_materials;
    c->mat = (gf_material*)malloc(sizeof(gf_material) * (size_t)(c->nmat > 0 ? c->nmat : 1));
    memcpy(c->mat, d->materials, sizeof(gf_material) * (size_t)c->nmat);
    c->cell = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)c->n);
    memcpy(c->cell, d->material_cell, sizeof(uint32_t) * (size_t)c->n);
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) { c->Tg2c[a][b] = d->Tg2c[a][b]; c->Tc2g[b][a] = d->Tg2c[a][b]; }  /* read.cpp:105-107 */
    for (int a = 0; a < 3; a++) {
        c->Hext_stat[a] = d->Hext_stat[a]; c->Hext_altr[a] = d->Hext_altr[a];
        c->H_altr_freq[a] = d->H_altr_freq[a]; c->demag_fac[a] = d->demag_fac[a]; c->Hext[a] = 0.;
    }
    c->npre = d->n_prescribe_Hext;
    if (c->npre > 0) {
        c->pre = (double*)malloc(sizeof(double) * 4 * (size_t)c->npre);
        memcpy(c->pre, d->prescribe_Hext, sizeof(double) * 4 * (size_t)c->npre);
    }
    c->pre_idx = 0;
    /* global.cpp:79-116 */
    for (int i = 0; i < c->nmat; i++) { if (c->mat[i].if_FM) c->if_FM_all = 1; if (c->mat[i].if_AFM) c->if_AFM_all = 1; }
    for (int i = 0; i < c->nmat; i++) if (c->mat[i].if_FE) c->if_FE_all = 1;
    for (int64_t id = 0; id < c->n; id++) {
        if (isFM(c, id)) c->NFM++;
        if (isAFM(c, id)) c->NAFM++;
        if (isFE(c, id)) c->NFE++;
    }
    for (int a = 0; a < 3; a++) {
        c->Eext_stat[a] = d->Eext_stat[a]; c->Eext_altr[a] = d->Eext_altr[a];
        c->E_altr_freq[a] = d->E_altr_freq[a]; c->Eext[a] = 0.;
    }
    c->npreE = d->n_prescribe_Eext;
    if (c->npreE > 0) {
        c->preE = (double*)malloc(sizeof(double) * 4 * (size_t)c->npreE);
        memcpy(c->preE, d->prescribe_Eext, sizeof(double) * 4 * (size_t)c->npreE);
    }
    c->preE_idx = 0;
    /* elastic_initialize.cpp:4-83, 245-336 */
    c->if_elasto = (c->flags & GF_FLAG_ELASTODYNAMICS) != 0;
    /* periodic z with ny != nz makes the reference read past its arrays
     * (the k == ny-1 test of elastic_dynamic_solver.cpp:1247) */
    if (c->if_elasto && c->per[2] && c->ny != c->nz) { free(c->mat); free(c->cell); free(c->pre); free(c); return NULL; }
    if (c->if_elasto && !(c->flags & GF_FLAG_1D_ABC) && c->nz < 5) { free(c->mat); free(c->cell); free(c->pre); free(c); return NULL; }
    if (c->if_elasto && (c->flags & GF_FLAG_1D_ABC) && c->nz < 2) { free(c->mat); free(c->cell); free(c->pre); free(c); return NULL; }
    /* at k = ny-1 the forward z stress takes the material of cell (i,j,0),
     * undefined in the reference when that cell is vacuum */
    if (c->if_elasto && c->ny < c->nz)
        for (int64_t i = 0; i < c->nx; i++)
            for (int64_t j = 0; j < c->ny; j++) {
                int64_t k = c->ny - 1;
                if (c->cell[ID(c, i, j, k)] && c->cell[ID(c, i, j, k + 1)] && !c->cell[ID(c, i, j, 0)]) {
                    free(c->mat); free(c->cell); free(c->pre); free(c); return NULL;
                }
            }
    c->source_type = d->input_strain_type; c->source_z = d->input_strain_sourcez;
    c->source_comp = d->input_strain_component; c->num_cycles = d->num_strain_cycles;
    c->sigma_gauss = d->sigma_gauss; c->amplitude_gauss = d->amplitude_gauss;
    c->w_liao = d->weighting_factor_fourth_Liao;
    {
        double c11_1st = 0., c44_1st = 0., rho_1st = 1.e-8, c11_last = 0., c44_last = 0., rho_last = 1.e-8;
        int found = 0;
        for (int64_t i = 0; i < c->nx && !found; i++)
            for (int64_t j = 0; j < c->ny && !found; j++) {
                const gf_material* m = matof(c, ID(c, i, j, 0));
                if (m) { c11_1st = m->c11; c44_1st = m->c44; rho_1st = m->density; found = 1; }
            }
        found = 0;
        for (int64_t i = 0; i < c->nx && !found; i++)
            for (int64_t j = 0; j < c->ny && !found; j++) {
                const gf_material* m = matof(c, ID(c, i, j, c->nz - 1));
                if (m) { c11_last = m->c11; c44_last = m->c44; rho_last = m->density; found = 1; }
            }
        c->vs_long_1st = sqrt((c11_1st / rho_1st)); c->vs_tran_1st = sqrt((c44_1st / rho_1st));
        c->vs_long_last = sqrt((c11_last / rho_last)); c->vs_tran_last = sqrt((c44_last / rho_last));
    }
    int64_t n = c->n;
    for (int a = 0; a < 3; a++) {
        c->m[a] = dalloc(n); c->ms[a] = dalloc(n); c->dm[a] = dalloc(n);
        for (int s = 0; s < 4; s++) { c->k[s][a] = dalloc(n); c->k1[s][a] = dalloc(n); c->k2[s][a] = dalloc(n); }
        c->m1[a] = dalloc(n); c->m2[a] = dalloc(n); c->ms1[a] = dalloc(n); c->ms2[a] = dalloc(n);
        c->dm1[a] = dalloc(n); c->dm2[a] = dalloc(n);
        c->H[a] = dalloc(n); c->Hnu[a] = dalloc(n); c->Hdrive[a] = dalloc(n);
        c->p[a] = dalloc(n); c->q[a] = dalloc(n); c->ps[a] = dalloc(n); c->qs[a] = dalloc(n); c->Es[a] = dalloc(n);
        for (int s = 0; s < 4; s++) { c->kp[s][a] = dalloc(n); c->kq[s][a] = dalloc(n); }
    }
    int64_t nyz = c->ny * c->nz;
    c->surfYZ = (unsigned char*)calloc((size_t)((c->nx + 1) * nyz), 1);
    c->surfXZ = (unsigned char*)calloc((size_t)(c->nx * (c->ny + 1) * c->nz), 1);
    c->surfXY = (unsigned char*)calloc((size_t)(c->nx * c->ny * (c->nz + 1)), 1);
    c->asurfYZ = (unsigned char*)calloc((size_t)((c->nx + 1) * nyz), 1);
    c->asurfXZ = (unsigned char*)calloc((size_t)(c->nx * (c->ny + 1) * c->nz), 1);
    c->asurfXY = (unsigned char*)calloc((size_t)(c->nx * c->ny * (c->nz + 1)), 1);
    c->fsurfYZ = (unsigned char*)calloc((size_t)((c->nx + 1) * nyz), 1);
    c->fsurfXZ = (unsigned char*)calloc((size_t)(c->nx * (c->ny + 1) * c->nz), 1);
    c->fsurfXY = (unsigned char*)calloc((size_t)(c->nx * c->ny * (c->nz + 1)), 1);
    if (c->if_FE_all) build_masks(c, isFE, c->fsurfYZ, c->fsurfXZ, c->fsurfXY);   /* ferroelectric_initialize.cpp:74-308 */
    for (int a = 0; a < 3; a++) {
        c->u[a] = dalloc(n); c->v[a] = dalloc(n); c->us[a] = dalloc(n); c->vs[a] = dalloc(n);
        for (int s = 0; s < 4; s++) { c->ku[s][a] = dalloc(n); c->kv[s][a] = dalloc(n); }
        c->force[a] = dalloc(n); c->force_store[a] = dalloc(n);
        c->gYZ[a] = dalloc((c->nx + 1) * nyz); c->gXZ[a] = dalloc(c->nx * (c->ny + 1) * c->nz);
        c->gXY[a] = dalloc(c->nx * c->ny * (c->nz + 1));
        c->t1[a] = dalloc(n); c->t2[a] = dalloc(n); c->t3[a] = dalloc(n); c->t4[a] = dalloc(n);
    }
    for (int a = 0; a < 6; a++) { c->e_glb[a] = dalloc(n); c->e_crt[a] = dalloc(n); c->e0_glb[a] = dalloc(n); c->e0_crt[a] = dalloc(n); c->e0t0[a] = dalloc(n); }
    for (int a = 0; a < 6; a++) c->e_t0[a] = dalloc(n);
    c->sfYZ = (unsigned char*)calloc((size_t)((c->nx + 1) * nyz), 1);
    c->sfXZ = (unsigned char*)calloc((size_t)(c->nx * (c->ny + 1) * c->nz), 1);
    c->sfXY = (unsigned char*)calloc((size_t)(c->nx * c->ny * (c->nz + 1)), 1);
    /* stress-free faces: a vacuum side or a non-periodic wall (elastic_initialize.cpp:245-336) */
    for (int64_t i = 0; i <= c->nx; i++)
        for (int64_t j = 0; j < c->ny; j++)
            for (int64_t k = 0; k < c->nz; k++) {
                int f;
                if (i == 0 || i == c->nx) f = c->per[0] ? !(c->cell[ID(c, 0, j, k)] && c->cell[ID(c, c->nx - 1, j, k)]) : 1;
                else f = !(c->cell[ID(c, i - 1, j, k)] && c->cell[ID(c, i, j, k)]);
                c->sfYZ[(i * c->ny + j) * c->nz + k] = (unsigned char)f;
            }
    for (int64_t i = 0; i < c->nx; i++)
        for (int64_t j = 0; j <= c->ny; j++)
            for (int64_t k = 0; k < c->nz; k++) {
                int f;
                if (j == 0 || j == c->ny) f = c->per[1] ? !(c->cell[ID(c, i, 0, k)] && c->cell[ID(c, i, c->ny - 1, k)]) : 1;
                else f = !(c->cell[ID(c, i, j - 1, k)] && c->cell[ID(c, i, j, k)]);
                c->sfXZ[(i * (c->ny + 1) + j) * c->nz + k] = (unsigned char)f;
            }
    for (int64_t i = 0; i < c->nx; i++)
        for (int64_t j = 0; j < c->ny; j++)
            for (int64_t k = 0; k <= c->nz; k++) {
                int f;
                if (k == 0 || k == c->nz) f = c->per[2] ? !(c->cell[ID(c, i, j, 0)] && c->cell[ID(c, i, j, c->nz - 1)]) : 1;
                else f = !(c->cell[ID(c, i, j, k - 1)] && c->cell[ID(c, i, j, k)]);
                c->sfXY[(i * c->ny + j) * (c->nz + 1) + k] = (unsigned char)f;
            }
    for (int f = 0; f < 6; f++) { c->Aij[f] = dalloc(n); c->Aij1[f] = dalloc(n); c->Aij2[f] = dalloc(n); }
    if (c->if_FM_all) {
        build_masks(c, isFM, c->surfYZ, c->surfXZ, c->surfXY);
        build_aij(c, isFM, c->surfYZ, c->surfXZ, c->surfXY, 0, c->Aij);
    }
    if (c->if_AFM_all) {
        build_masks(c, isAFM, c->asurfYZ, c->asurfXZ, c->asurfXY);
        build_aij(c, isAFM, c->asurfYZ, c->asurfXZ, c->asurfXY, 1, c->Aij1);
        build_aij(c, isAFM, c->asurfYZ, c->asurfXZ, c->asurfXY, 2, c->Aij2);
    }
    /* demag spectra: magnetic_initialize.cpp:1099-1112 */
    if ((c->flags & GF_FLAG_MAGNETOSTATICS) && !(c->flags & GF_FLAG_MAG_1D) && !(c->flags & GF_FLAG_DEMAG_FACTOR)) {
        int64_t nh = c->nx * c->ny * (c->nz / 2 + 1);
        double* real9 = dalloc(9 * n);
        gfo_newell_real(c->nx, c->ny, c->nz, c->dx, c->dy, c->dz, real9);
        for (int s = 0; s < 9; s++) {
            c->kspec[s] = dalloc(2 * nh);
            gfo_rfft3(c->nx, c->ny, c->nz, real9 + s * n, c->kspec[s]);
        }
        free(real9);
    }
    if (em_setup(c, d)) { gfo_destroy(c); return NULL; }
    if ((c->flags & GF_FLAG_SPIN_PUMPING) && spin_setup(c, d)) { gfo_destroy(c); return NULL; }
    if (c->flags & GF_FLAG_ELASTOSTATIC) {
        /* elastic_initialize.cpp:22-40: the component-wise largest
         * constants, scaled */
        double sc = d->scale_elasto > 0. ? d->scale_elasto : 1.;
        c->ch11 = c->mat[0].c11; c->ch12 = c->mat[0].c12; c->ch44 = c->mat[0].c44;
        for (int t = 1; t < c->nmat; t++) {
            if (c->mat[t].c11 > c->ch11) c->ch11 = c->mat[t].c11;
            if (c->mat[t].c12 > c->ch12) c->ch12 = c->mat[t].c12;
            if (c->mat[t].c44 > c->ch44) c->ch44 = c->mat[t].c44;
        }
        c->ch11 = c->ch11 * sc; c->ch12 = c->ch12 * sc; c->ch44 = c->ch44 * sc;
    }
    if (c->flags & GF_FLAG_FLEXO) {
        /* read.cpp:181-199 and check_material's zeroing (global.cpp:339-358) */
        c->flexo_on = 1;
        c->flx = (double (*)[11])calloc((size_t)(c->nmat > 0 ? c->nmat : 1), sizeof(double[11]));
        for (int t = 0; t < c->nmat; t++) {
            const gf_material* m = &c->mat[t];
            if (!m->if_FE) continue;
            double c11 = m->c11, c12 = m->c12, c44 = m->c44, Q11 = m->Q11, Q12 = m->Q12, Q44 = m->Q44;
            double F11 = ((c11 + c12) * m->f11 - 2. * c12 * m->f12) / (c11 + 2. * c12) / (c11 - c12);
            double F12 = (c11 * m->f12 - c12 * m->f11) / (c11 + 2. * c12) / (c11 - c12);
            double F44 = m->f44 / c44;
            double* f = c->flx[t];
            f[0] = F11; f[1] = F12; f[2] = F44;
            f[3] = 2. * (c11 * (Q11 * F11 + 2. * Q12 * F12) + 2. * c12 * (Q11 * F12 + Q12 * F11 + Q12 * F12));
            f[4] = 2. * (c11 * (Q11 * F12 + Q12 * F11 + Q12 * F12) + c12 * (Q11 * F11 + Q11 * F12 + Q12 * F11 + 3. * Q12 * F12));
            f[5] = 2. * c44 * Q44 * F44;
            f[6] = c11 * (F11 * F11 + 2. * F12 * F12) + 2. * c12 * (F12 * F12 + 2. * F11 * F12);
            f[7] = c44 * F44 * F44;
            f[8] = m->f11; f[9] = m->f12; f[10] = m->f44;
        }
    }
    return c;
}

void gfo_destroy(gfo_ctx* c) {
    if (!c) return;
    free(c->mat); free(c->cell); free(c->pre);
    free(c->spin_src); free(c->spin_coef); free(c->flx);
    for (int a = 0; a < 3; a++) free(c->Jis[a]);
    for (int a = 0; a < 3; a++) {
        free(c->m[a]); free(c->ms[a]); free(c->dm[a]);
        for (int s = 0; s < 4; s++) { free(c->k[s][a]); free(c->k1[s][a]); free(c->k2[s][a]); }
        free(c->m1[a]); free(c->m2[a]); free(c->ms1[a]); free(c->ms2[a]); free(c->dm1[a]); free(c->dm2[a]);
        free(c->H[a]); free(c->Hnu[a]); free(c->Hdrive[a]);
        free(c->p[a]); free(c->q[a]); free(c->ps[a]); free(c->qs[a]); free(c->Es[a]);
        for (int s = 0; s < 4; s++) { free(c->kp[s][a]); free(c->kq[s][a]); }
    }
    free(c->fsurfYZ); free(c->fsurfXZ); free(c->fsurfXY); free(c->preE);
    for (int a = 0; a < 3; a++) {
        free(c->u[a]); free(c->v[a]); free(c->us[a]); free(c->vs[a]);
        for (int s = 0; s < 4; s++) { free(c->ku[s][a]); free(c->kv[s][a]); }
        free(c->force[a]); free(c->force_store[a]); free(c->gYZ[a]); free(c->gXZ[a]); free(c->gXY[a]);
        free(c->t1[a]); free(c->t2[a]); free(c->t3[a]); free(c->t4[a]);
    }
    for (int a = 0; a < 6; a++) { free(c->e_glb[a]); free(c->e_crt[a]); free(c->e0_glb[a]); free(c->e0_crt[a]); free(c->e0t0[a]); }
    for (int a = 0; a < 6; a++) free(c->e_t0[a]);
    free(c->sfYZ); free(c->sfXZ); free(c->sfXY);
    free(c->surfYZ); free(c->surfXZ); free(c->surfXY); free(c->asurfYZ); free(c->asurfXZ); free(c->asurfXY);
    for (int f = 0; f < 6; f++) { free(c->Aij[f]); free(c->Aij1[f]); free(c->Aij2[f]); }
    for (int s = 0; s < 9; s++) free(c->kspec[s]);
    for (int a = 0; a < 3; a++) {
        free(c->Em[a]); free(c->Ems[a]); free(c->Et1[a]); free(c->Hm[a]); free(c->Hms[a]);
        for (int s = 0; s < 4; s++) { free(c->kE[s][a]); free(c->kH[s][a]); }
        free(c->Ecell[a]); free(c->Hcell[a]); free(c->Jf[a]); free(c->pec[a]);
        for (int p = 0; p < 3; p++) {
            free(c->Jp[p][a]); free(c->Jps[p][a]);
            for (int s = 0; s < 4; s++) free(c->kJ[s][p][a]);
        }
    }
    free(c);
}

/* The magnets read the static strain ex*t0_crt (magnetic_solver.cpp:89-96:
 * exx = exxt0_crt + Dexx_crt) when strain acts on them without
 * elastodynamics, whose dynamic part is then zero. */
static int static_strain(const gfo_ctx* c) {
    return (c->flags & GF_FLAG_ELASTO_ON_PM) && !(c->flags & GF_FLAG_ELASTODYNAMICS);
}

static double** field_ptrs(gfo_ctx* c, int id) {
    switch (id) {
    case GF_FIELD_M: return c->m;
    case GF_FIELD_M_STORE: return c->ms;
    case GF_FIELD_HSTAT: return c->H;
    case GF_FIELD_HEXT_NONUNIF: return c->Hnu;
    case GF_FIELD_M_AFM1: return c->m1;
    case GF_FIELD_M_AFM2: return c->m2;
    case GF_FIELD_P: return c->p;
    case GF_FIELD_Q: return c->q;
    case GF_FIELD_P_STORE: return c->ps;
    case GF_FIELD_ESTAT: return c->Es;
    case GFO_FIELD_Q_STORE: return c->qs;
    case GFO_FIELD_DV1: return c->kv[0];
    case GFO_FIELD_K1: return c->k[0];
    case GFO_FIELD_KA1: return c->k1[0];
    case GFO_FIELD_KA2: return c->k2[0];
    case GFO_FIELD_KQ1: return c->kq[0];
    case GFO_FIELD_E0_DIAG: return c->if_elasto ? c->e0_glb : NULL;
    case GFO_FIELD_E0_OFF: return c->if_elasto ? c->e0_glb + 3 : NULL;
    case GF_FIELD_U: return c->u;
    case GF_FIELD_V: return c->v;
    case GF_FIELD_U_STORE: return c->us;
    case GF_FIELD_ELASTOFORCE: return (c->flags & GF_FLAG_ELASTODYNAMICS) ? c->force_store : NULL;
    case GF_FIELD_EIGENSTRAIN_T0_DIAG: return (c->flags & (GF_FLAG_ELASTODYNAMICS | GF_FLAG_ELASTOSTATIC)) ? c->e0t0 : NULL;
    case GF_FIELD_EIGENSTRAIN_T0_OFF: return (c->flags & (GF_FLAG_ELASTODYNAMICS | GF_FLAG_ELASTOSTATIC)) ? c->e0t0 + 3 : NULL;
    case GF_FIELD_STRAIN_DIAG: return c->e_glb;
    case GF_FIELD_STRAIN_OFF: return c->e_glb + 3;
    case GF_FIELD_HDRIVE: return (c->flags & GF_FLAG_HDRIVE) ? c->Hdrive : NULL;
    case GF_FIELD_STRAIN_T0_DIAG:
        return (static_strain(c) || (c->flags & (GF_FLAG_ELASTOSTATIC | GF_FLAG_ELASTODYNAMICS))) ? c->e_t0 : NULL;
    case GF_FIELD_STRAIN_T0_OFF:
        return (static_strain(c) || (c->flags & (GF_FLAG_ELASTOSTATIC | GF_FLAG_ELASTODYNAMICS))) ? c->e_t0 + 3 : NULL;
    case GF_FIELD_E_CELL: return c->if_em ? c->Ecell : NULL;
    case GF_FIELD_H_CELL: return c->if_em ? c->Hcell : NULL;
    case GF_FIELD_J_ISHE: return c->spin_on ? c->Jis : NULL;
    default: return NULL;
    }
}

/* Uploading m also resets the store (magnetic_initialize.cpp:1081-1097). */
/* The Yee fields (three staggered arrays, concatenated) and Jp (nine cell
 * arrays); returns the number of arrays, 0 for other ids. */
static int em_parts(gfo_ctx* c, int id, double** parts[9], int64_t sizes[9]) {
    if (!c->if_em) return 0;
    double** src = NULL;
    const int64_t* sz = NULL;
    if (id == GF_FIELD_E_EM) { src = c->Em; sz = c->nE; }
    else if (id == GF_FIELD_E_EM_STORE) { src = c->Ems; sz = c->nE; }
    else if (id == GF_FIELD_H_EM) { src = c->Hm; sz = c->nH; }
    else if (id == GF_FIELD_H_EM_STORE) { src = c->Hms; sz = c->nH; }
    if (src) {
        for (int a = 0; a < 3; a++) { parts[a] = &src[a]; sizes[a] = sz[a]; }
        return 3;
    }
    if (id == GF_FIELD_JP) {
        for (int p = 0; p < 3; p++)
            for (int a = 0; a < 3; a++) { parts[3 * p + a] = &c->Jp[p][a]; sizes[3 * p + a] = c->n; }
        return 9;
    }
    return 0;
}

static int em_copy(gfo_ctx* c, int id, double* host, size_t count, int to_ctx) {
    double** parts[9];
    int64_t sizes[9];
    int np = em_parts(c, id, parts, sizes);
    if (!np) return -1;
    size_t tot = 0;
    for (int q = 0; q < np; q++) tot += (size_t)sizes[q];
    if (count != tot) return -1;
    size_t o = 0;
    for (int q = 0; q < np; q++) {
        if (to_ctx) memcpy(*parts[q], host + o, sizeof(double) * (size_t)sizes[q]);
        else memcpy(host + o, *parts[q], sizeof(double) * (size_t)sizes[q]);
        o += (size_t)sizes[q];
    }
    if (!to_ctx) return 0;
    /* EMdynamic_initialize.cpp:376-404: the stores start equal to the
     * fields and the cell fields follow them */
    if (id == GF_FIELD_E_EM) {
        for (int a = 0; a < 3; a++) memcpy(c->Ems[a], c->Em[a], sizeof(double) * (size_t)c->nE[a]);
        em_cells(c, 1);
    }
    if (id == GF_FIELD_H_EM) {
        for (int a = 0; a < 3; a++) memcpy(c->Hms[a], c->Hm[a], sizeof(double) * (size_t)c->nH[a]);
        em_cells(c, 2);
    }
    if (id == GF_FIELD_JP)
        for (int p = 0; p < 3; p++)
            for (int a = 0; a < 3; a++) memcpy(c->Jps[p][a], c->Jp[p][a], sizeof(double) * (size_t)c->n);
    return 0;
}

int gfo_set_field(gfo_ctx* c, int id, const double* src, size_t count) {
    if (id == GF_FIELD_E_EM || id == GF_FIELD_E_EM_STORE || id == GF_FIELD_H_EM || id == GF_FIELD_H_EM_STORE ||
        id == GF_FIELD_JP)
        return em_copy(c, id, (double*)src, count, 1);
    if (id == GF_FIELD_E_CELL || id == GF_FIELD_H_CELL) return -1;
    double** f = field_ptrs(c, id);
    if (!f || count != (size_t)(3 * c->n)) return -1;
    for (int a = 0; a < 3; a++) memcpy(f[a], src + a * c->n, sizeof(double) * (size_t)c->n);
    if (id == GF_FIELD_M) for (int a = 0; a < 3; a++) memcpy(c->ms[a], c->m[a], sizeof(double) * (size_t)c->n);
    /* main.cpp:148: read static strain or eigenstrain replaces the start-up solve */
    if (id == GF_FIELD_STRAIN_T0_DIAG || id == GF_FIELD_STRAIN_T0_OFF || id == GF_FIELD_EIGENSTRAIN_T0_DIAG ||
        id == GF_FIELD_EIGENSTRAIN_T0_OFF)
        c->t0_fixed = 1;
    if (id == GF_FIELD_M_AFM1) for (int a = 0; a < 3; a++) memcpy(c->ms1[a], c->m1[a], sizeof(double) * (size_t)c->n);
    if (id == GF_FIELD_M_AFM2) for (int a = 0; a < 3; a++) memcpy(c->ms2[a], c->m2[a], sizeof(double) * (size_t)c->n);
    /* ferroelectric_initialize.cpp:355-366 */
    if (id == GF_FIELD_P) for (int a = 0; a < 3; a++) memcpy(c->ps[a], c->p[a], sizeof(double) * (size_t)c->n);
    if (id == GF_FIELD_Q) for (int a = 0; a < 3; a++) memcpy(c->qs[a], c->q[a], sizeof(double) * (size_t)c->n);
    /* elastic_initialize.cpp:386-410: stores equal the state, strain from u */
    if (id == GF_FIELD_V) for (int a = 0; a < 3; a++) memcpy(c->vs[a], c->v[a], sizeof(double) * (size_t)c->n);
    if (id == GF_FIELD_U) {
        for (int a = 0; a < 3; a++) memcpy(c->us[a], c->u[a], sizeof(double) * (size_t)c->n);
        elastic_gradients(c);
        elastic_strain(c);
    }
    return 0;
}

int gfo_get_field(const gfo_ctx* c, int id, double* dst, size_t count) {
    if (id == GF_FIELD_E_EM || id == GF_FIELD_E_EM_STORE || id == GF_FIELD_H_EM || id == GF_FIELD_H_EM_STORE ||
        id == GF_FIELD_JP)
        return em_copy((gfo_ctx*)c, id, dst, count, 0);
    double** f = field_ptrs((gfo_ctx*)c, id);
    if (id == GFO_FIELD_DM) f = ((gfo_ctx*)c)->dm;
    if (id == GFO_FIELD_DM1) f = ((gfo_ctx*)c)->dm1;
    if (id == GFO_FIELD_DM2) f = ((gfo_ctx*)c)->dm2;
    if (!f || count != (size_t)(3 * c->n)) return -1;
    for (int a = 0; a < 3; a++) memcpy(dst + a * c->n, f[a], sizeof(double) * (size_t)c->n);
    return 0;
}

/* ----------------------------------------------------------- drives */
/* global.cpp:513-549 (the H part; the E table belongs to the FE path) */
void gfo_time_advance(gfo_ctx* c) {
    c->time = c->time + c->dt;
    if (c->flags & GF_FLAG_PRESCRIBE_EEXT) {
        if (c->preE_idx < c->npreE && (c->time - c->dt) > c->preE[4 * c->preE_idx]) c->preE_idx = c->preE_idx + 1;
        if (c->preE_idx < c->npreE) {
            c->Eext[0] = c->preE[4 * c->preE_idx + 1];
            c->Eext[1] = c->preE[4 * c->preE_idx + 2];
            c->Eext[2] = c->preE[4 * c->preE_idx + 3];
        } else { c->Eext[0] = 0.; c->Eext[1] = 0.; c->Eext[2] = 0.; }
    }
    if (c->flags & GF_FLAG_PRESCRIBE_HEXT) {
        if (c->pre_idx < c->npre && (c->time - c->dt) > c->pre[4 * c->pre_idx]) c->pre_idx = c->pre_idx + 1;
        if (c->pre_idx < c->npre) {
            c->Hext[0] = c->pre[4 * c->pre_idx + 1];
            c->Hext[1] = c->pre[4 * c->pre_idx + 2];
            c->Hext[2] = c->pre[4 * c->pre_idx + 3];
        } else { c->Hext[0] = 0.; c->Hext[1] = 0.; c->Hext[2] = 0.; }
    }
}

/* magnetic_solver.cpp:1834-1861: drive at t-dt (stage 1), t-dt/2
 * (stages 2, 3) and t (stage 4). */
static void update_external_H(gfo_ctx* c, int stage) {
    double local_time = stage == 1 ? c->time - c->dt : (stage == 4 ? c->time : c->time - c->dt * 0.5);
    for (int a = 0; a < 3; a++)
        c->Hext[a] = c->Hext_stat[a] + c->Hext_altr[a] * sin(2. * PI * c->H_altr_freq[a] * local_time);
}

/* ------------------------------------------------------------ stencil */
/* magnetic_solver.cpp:1639-1831 (FM, with DMI surface slope) and
 * magnetic_AFM_solver.cpp:1837-2219 (AFM, slope zero): d[0..3] =
 * dmxdx, dmydy, dmzdx, dmzdy; d2[axis][side][comp] with side 0 = f. */
static void stencil(const gfo_ctx* c, double* const* s, const unsigned char* mYZ, const unsigned char* mXZ,
                    const unsigned char* mXY, int dmi, double A, double D, int64_t i, int64_t j, int64_t k,
                    double d[4], double d2[3][2][3]) {
    int64_t nx = c->nx, ny = c->ny, nz = c->nz;
    double dx = c->dx, dy = c->dy, dz = c->dz;
    int64_t id = ID(c, i, j, k);
    double mx_c = s[0][id], my_c = s[1][id], mz_c = s[2][id];
    double bw[3], fw[3];
    int64_t nb;
    /* x */
    if (sYZ(mYZ, c, i, j, k)) {
        if (dmi) { bw[0] = -DMI_SLOPE(D, A) * mz_c; bw[1] = 0.; bw[2] = DMI_SLOPE(D, A) * mx_c; }
        else { bw[0] = 0.; bw[1] = 0.; bw[2] = 0.; }
    } else {
        nb = ID(c, i == 0 ? nx - 1 : i - 1, j, k);
        bw[0] = (mx_c - s[0][nb]) / dx; bw[1] = (my_c - s[1][nb]) / dx; bw[2] = (mz_c - s[2][nb]) / dx;
    }
    if (sYZ(mYZ, c, i + 1, j, k)) {
        if (dmi) { fw[0] = -DMI_SLOPE(D, A) * mz_c; fw[1] = 0.; fw[2] = DMI_SLOPE(D, A) * mx_c; }
        else { fw[0] = 0.; fw[1] = 0.; fw[2] = 0.; }
    } else {
        nb = ID(c, i == nx - 1 ? 0 : i + 1, j, k);
        fw[0] = (s[0][nb] - mx_c) / dx; fw[1] = (s[1][nb] - my_c) / dx; fw[2] = (s[2][nb] - mz_c) / dx;
    }
    d[0] = (bw[0] + fw[0]) * 0.5; d[2] = (bw[2] + fw[2]) * 0.5;
    for (int a = 0; a < 3; a++) { d2[0][0][a] = fw[a] / dx; d2[0][1][a] = -bw[a] / dx; }
    /* y */
    if (sXZ(mXZ, c, i, j, k)) {
        if (dmi) { bw[0] = 0.; bw[1] = -DMI_SLOPE(D, A) * mz_c; bw[2] = DMI_SLOPE(D, A) * my_c; }
        else { bw[0] = 0.; bw[1] = 0.; bw[2] = 0.; }
    } else {
        nb = ID(c, i, j == 0 ? ny - 1 : j - 1, k);
        bw[0] = (mx_c - s[0][nb]) / dy; bw[1] = (my_c - s[1][nb]) / dy; bw[2] = (mz_c - s[2][nb]) / dy;
    }
    if (sXZ(mXZ, c, i, j + 1, k)) {
        if (dmi) { fw[0] = 0.; fw[1] = -DMI_SLOPE(D, A) * mz_c; fw[2] = DMI_SLOPE(D, A) * my_c; }
        else { fw[0] = 0.; fw[1] = 0.; fw[2] = 0.; }
    } else {
        nb = ID(c, i, j == ny - 1 ? 0 : j + 1, k);
        fw[0] = (s[0][nb] - mx_c) / dy; fw[1] = (s[1][nb] - my_c) / dy; fw[2] = (s[2][nb] - mz_c) / dy;
    }
    d[1] = (bw[1] + fw[1]) * 0.5; d[3] = (bw[2] + fw[2]) * 0.5;
    for (int a = 0; a < 3; a++) { d2[1][0][a] = fw[a] / dy; d2[1][1][a] = -bw[a] / dy; }
    /* z: slope zero at a surface for both kinds */
    if (sXY(mXY, c, i, j, k)) { bw[0] = 0.; bw[1] = 0.; bw[2] = 0.; }
    else {
        nb = ID(c, i, j, k == 0 ? nz - 1 : k - 1);
        bw[0] = (mx_c - s[0][nb]) / dz; bw[1] = (my_c - s[1][nb]) / dz; bw[2] = (mz_c - s[2][nb]) / dz;
    }
    if (sXY(mXY, c, i, j, k + 1)) { fw[0] = 0.; fw[1] = 0.; fw[2] = 0.; }
    else {
        nb = ID(c, i, j, k == nz - 1 ? 0 : k + 1);
        fw[0] = (s[0][nb] - mx_c) / dz; fw[1] = (s[1][nb] - my_c) / dz; fw[2] = (s[2][nb] - mz_c) / dz;
    }
    for (int a = 0; a < 3; a++) { d2[2][0][a] = fw[a] / dz; d2[2][1][a] = -bw[a] / dz; }
}

/* ------------------------------------------------------- FM stage */
/* magnetic_solver.cpp:3-229 (RK1; RK2-4 at :285, :568, :849 differ only in
 * the drive time and the output array), plus get_dm :1449-1457 at s=4. */
void gfo_mag_get_dm(gfo_ctx* c, int stage) {
    if (!(c->flags & GF_FLAG_PRESCRIBE_HEXT)) update_external_H(c, stage);
    const double drive_t = stage == 1 ? c->time - c->dt : (stage == 4 ? c->time : c->time - c->dt * 0.5);
    const double drive_s = sin(2. * PI * c->hdrive_freq * drive_t);
    int64_t ny = c->ny, nz = c->nz;
    double** out = c->k[stage - 1];
    #pragma omp parallel for schedule(static) if (c->n >= GFO_OMP_MIN)
    for (int64_t i = 0; i < c->n; i++) {
        const gf_material* mat = matof(c, i);
        if (!mat || !mat->if_FM) continue;
        int64_t x = i / (ny * nz), y = (i - x * (ny * nz)) / nz, z = i - x * (ny * nz) - y * nz;
        double mgx = c->ms[0][i], mgy = c->ms[1][i], mgz = c->ms[2][i];
        double mx, my, mz;
        glb2crt(c, mgx, mgy, mgz, &mx, &my, &mz);
        double Ms = mat->Ms;
        double K1 = mat->K1 / MU0 / Ms, K2 = mat->K2 / MU0 / Ms;
        double B1 = mat->B1 / MU0 / Ms, B2 = mat->B2 / MU0 / Ms;
        double Aexi = mat->Aex, iDMIi = mat->iDMI;
        double exx = 0., eyy = 0., ezz = 0., eyz = 0., exz = 0., exy = 0.;
        if (static_strain(c)) {                  /* :89-96, dynamic strain zero */
            exx = c->e_t0[0][i] + 0.; eyy = c->e_t0[1][i] + 0.; ezz = c->e_t0[2][i] + 0.;
            eyz = c->e_t0[3][i] + 0.; exz = c->e_t0[4][i] + 0.; exy = c->e_t0[5][i] + 0.;
        } else if (c->flags & GF_FLAG_ELASTO_ON_PM) {   /* :89-96: the static strain (zero unless uploaded) plus the dynamic one */
            exx = c->e_t0[0][i] + c->e_crt[0][i]; eyy = c->e_t0[1][i] + c->e_crt[1][i]; ezz = c->e_t0[2][i] + c->e_crt[2][i];
            eyz = c->e_t0[3][i] + c->e_crt[3][i]; exz = c->e_t0[4][i] + c->e_crt[4][i]; exy = c->e_t0[5][i] + c->e_crt[5][i];
        }
        double d[4], d2[3][2][3];
        stencil(c, c->ms, c->surfYZ, c->surfXZ, c->surfXY, 1, Aexi, iDMIi, x, y, z, d, d2);
        double hac[3] = {0., 0., 0.}, ha[3], hec[3], he[3], hx[3], hd[3], H[3];
        if (mat->anisotropy_type == 1) {
            hac[0] = -2. * (K1 * (my * my + mz * mz) + K2 * my * my * mz * mz) * mx;
            hac[1] = -2. * (K1 * (mx * mx + mz * mz) + K2 * mx * mx * mz * mz) * my;
            hac[2] = -2. * (K1 * (mx * mx + my * my) + K2 * mx * mx * my * my) * mz;
        } else if (mat->anisotropy_type == 2) {
            hac[0] = 0.; hac[1] = 0.; hac[2] = 2. * K1 * mz;
        }
        crt2glb(c, hac[0], hac[1], hac[2], &ha[0], &ha[1], &ha[2]);
        hec[0] = -2. * (B1 * (exx)*mx + B2 * ((exy)*my + (exz)*mz));
        hec[1] = -2. * (B1 * (eyy)*my + B2 * ((exy)*mx + (eyz)*mz));
        hec[2] = -2. * (B1 * (ezz)*mz + B2 * ((exz)*mx + (eyz)*my));
        crt2glb(c, hec[0], hec[1], hec[2], &he[0], &he[1], &he[2]);
        for (int a = 0; a < 3; a++)
            hx[a] = 2. / MU0 / Ms * (c->Aij[0][i] * d2[0][0][a] + c->Aij[1][i] * d2[0][1][a] + c->Aij[2][i] * d2[1][0][a]
                                    + c->Aij[3][i] * d2[1][1][a] + c->Aij[4][i] * d2[2][0][a] + c->Aij[5][i] * d2[2][1][a]);
        hd[0] = 2. * iDMIi / MU0 / Ms * d[2];
        hd[1] = 2. * iDMIi / MU0 / Ms * d[3];
        hd[2] = -2. * iDMIi / MU0 / Ms * (d[0] + d[1]);
        for (int a = 0; a < 3; a++)
            H[a] = c->H[a][i] + ha[a] + he[a] + hx[a] + hd[a] + c->Hext[a] + c->Hnu[a][i];
        if (c->flags & GF_FLAG_HDRIVE)   /* shaped drive at the stage time of the uniform one */
            for (int a = 0; a < 3; a++) H[a] = H[a] + c->Hdrive[a][i] * drive_s;
        if (c->if_em && (c->flags & GF_FLAG_EM_BACKACTION_M))   /* magnetic_solver.cpp:200-204 */
            for (int a = 0; a < 3; a++) H[a] = H[a] + c->Hcell[a][i];
        double alpha = mat->FM_damping, gyro = mat->gyro;
        double precess = -gyro / (1 + alpha * alpha) * c->dt;
        double px = mgy * H[2] - mgz * H[1], py = mgz * H[0] - mgx * H[2], pz = mgx * H[1] - mgy * H[0];
        double qx = mgy * pz - mgz * py, qy = mgz * px - mgx * pz, qz = mgx * py - mgy * px;
        out[0][i] = precess * px + alpha * precess * qx;
        out[1][i] = precess * py + alpha * precess * qy;
        out[2][i] = precess * pz + alpha * precess * qz;
    }
    if (stage == 4)
        for (int a = 0; a < 3; a++)
            #pragma omp parallel for schedule(static) if (c->n >= GFO_OMP_MIN)
            for (int64_t i = 0; i < c->n; i++)
                c->dm[a][i] = c->k[0][a][i] / 6. + c->k[1][a][i] / 3. + c->k[2][a][i] / 3. + c->k[3][a][i] / 6.;
    if (c->spin_on) spin_pump(c, stage);     /* magnetic_solver.cpp:226-228 */
}

/* update_prescribed_m_half / _full / update_prescribed_m
 * (magnetic_solver.cpp:1891-1957): every FM cell of the bottom layer k = 0
 * takes the cone (sqrt(1 - m0^2), m0 sin wt, m0 cos wt) at time t. */
static void prescribe_m(gfo_ctx* c, double* const* dst, double t) {
    double m0 = sin(c->precess_angle / 180. * PI);
    double omega = 2. * PI * c->precess_freq;
    for (int64_t i = 0; i < c->nx; i++)
        for (int64_t j = 0; j < c->ny; j++) {
            int64_t id = ID(c, i, j, 0);
            if (!isFM(c, id)) continue;
            dst[0][id] = sqrt((1. - m0 * m0));
            dst[1][id] = m0 * sin(omega * t);
            dst[2][id] = m0 * cos(omega * t);
        }
}

/* The time of the prescribed cone written by the update of a stage:
 * update_m_RK1 and _RK2 call the half-step form (t - dt/2), update_m_RK3
 * and update_m the full-step one (t) (magnetic_solver.cpp:1526-1609). */
static double prescribe_time(const gfo_ctx* c, int stage) {
#if GFO_PLANT == 8
    return stage == 1 ? c->time - c->dt * 0.5 : c->time;
#else
    return stage <= 2 ? c->time - c->dt * 0.5 : c->time;
#endif
}

/* magnetic_solver.cpp:1515-1636 */
void gfo_mag_update(gfo_ctx* c, int stage) {
    int presc = (c->flags & GF_FLAG_PRESCRIBED_M) != 0;
    if (stage == 4 && c->spin_on) spin_pump_step(c);
    for (int a = 0; a < 3; a++) {
        double *m = c->m[a], *s = c->ms[a];
        if (stage == 1 || stage == 2) { const double* k = c->k[stage - 1][a]; for (int64_t i = 0; i < c->n; i++) s[i] = m[i] + k[i] * 0.5; }
        else if (stage == 3) { const double* k = c->k[2][a]; for (int64_t i = 0; i < c->n; i++) s[i] = m[i] + k[i]; }
        else { const double* dm = c->dm[a]; for (int64_t i = 0; i < c->n; i++) m[i] = m[i] + dm[i]; }
    }
    if (stage < 4) {
        if (presc) prescribe_m(c, c->ms, prescribe_time(c, stage));
        return;
    }
    /* update_m: m += dm, the prescribed layer, then the store copies m */
    if (presc) prescribe_m(c, c->m, prescribe_time(c, 4));
    for (int a = 0; a < 3; a++) memcpy(c->ms[a], c->m[a], sizeof(double) * (size_t)c->n);
}

/* initialize_device (magnetic_initialize.cpp:1074-1086): the prescribed
 * layer at the current time, then the store copies m */
void gfo_prescribe_m(gfo_ctx* c) {
    if (!(c->flags & GF_FLAG_PRESCRIBED_M) || !c->if_FM_all) return;
    prescribe_m(c, c->m, c->time);
    for (int a = 0; a < 3; a++) memcpy(c->ms[a], c->m[a], sizeof(double) * (size_t)c->n);
}

/* ------------------------------------------------------ AFM stage */
/* magnetic_AFM_solver.cpp:3-430 (RK1; RK2-4 :431, :859, :1287), get_dm_AFM
 * :1715-1729.  Sublattice loops read only stores, so their order is free. */
void gfo_afm_get_dm(gfo_ctx* c, int stage) {
    update_external_H(c, stage);   /* unconditional in the AFM path (:61-64) */
    int64_t ny = c->ny, nz = c->nz;
    for (int sub = 0; sub < 2; sub++) {
        double** st = sub == 0 ? c->ms1 : c->ms2;
        double** other = sub == 0 ? c->ms2 : c->ms1;
        double** A = sub == 0 ? c->Aij1 : c->Aij2;
        double** out = sub == 0 ? c->k1[stage - 1] : c->k2[stage - 1];
        #pragma omp parallel for schedule(static) if (c->n >= GFO_OMP_MIN)
        for (int64_t i = 0; i < c->n; i++) {
            const gf_material* mat = matof(c, i);
            if (!mat || !mat->if_AFM) continue;
            int64_t x = i / (ny * nz), y = (i - x * (ny * nz)) / nz, z = i - x * (ny * nz) - y * nz;
            double ux = mat->uniaxis_AFM[0], uy = mat->uniaxis_AFM[1], uz = mat->uniaxis_AFM[2];
            double exx = 0., eyy = 0., ezz = 0., eyz = 0., exz = 0., exy = 0.;
            if (static_strain(c)) {                  /* magnetic_AFM_solver.cpp:89-96 */
                exx = c->e_t0[0][i] + 0.; eyy = c->e_t0[1][i] + 0.; ezz = c->e_t0[2][i] + 0.;
                eyz = c->e_t0[3][i] + 0.; exz = c->e_t0[4][i] + 0.; exy = c->e_t0[5][i] + 0.;
            } else if (c->flags & GF_FLAG_ELASTO_ON_PM) {   /* magnetic_AFM_solver.cpp:89-96: static plus dynamic */
                exx = c->e_t0[0][i] + c->e_crt[0][i]; eyy = c->e_t0[1][i] + c->e_crt[1][i]; ezz = c->e_t0[2][i] + c->e_crt[2][i];
                eyz = c->e_t0[3][i] + c->e_crt[3][i]; exz = c->e_t0[4][i] + c->e_crt[4][i]; exy = c->e_t0[5][i] + c->e_crt[5][i];
            }
            double mgx = st[0][i], mgy = st[1][i], mgz = st[2][i];
            double mx, my, mz;
            glb2crt(c, mgx, mgy, mgz, &mx, &my, &mz);
            double Ms = sub == 0 ? mat->Ms_AFM1 : mat->Ms_AFM2;
            double K1 = (sub == 0 ? mat->K1_AFM1 : mat->K1_AFM2) / MU0 / Ms;
            double udotm = mx * ux + my * uy + mz * uz;
            double B1 = (sub == 0 ? mat->B1_AFM1 : mat->B1_AFM2) / MU0 / Ms;
            double B2 = (sub == 0 ? mat->B2_AFM1 : mat->B2_AFM2) / MU0 / Ms;
            double JAFM = -1. * mat->J_AFM / MU0 / Ms;
            double d[4], d2[3][2][3];
            stencil(c, st, c->asurfYZ, c->asurfXZ, c->asurfXY, 0, 0., 0., x, y, z, d, d2);
            double hac[3], ha[3], hec[3], he[3], hx[3], hi[3], H[3];
            hac[0] = 2. * K1 * udotm * ux; hac[1] = 2. * K1 * udotm * uy; hac[2] = 2. * K1 * udotm * uz;
            crt2glb(c, hac[0], hac[1], hac[2], &ha[0], &ha[1], &ha[2]);
            hec[0] = -1. * (B1 * (exx)*mx + B2 * ((exy)*my + (exz)*mz));
            hec[1] = -1. * (B1 * (eyy)*my + B2 * ((exy)*mx + (eyz)*mz));
            hec[2] = -1. * (B1 * (ezz)*mz + B2 * ((exz)*mx + (eyz)*my));
            crt2glb(c, hec[0], hec[1], hec[2], &he[0], &he[1], &he[2]);
            for (int a = 0; a < 3; a++)
                hx[a] = 2. / MU0 / Ms * (A[0][i] * d2[0][0][a] + A[1][i] * d2[0][1][a] + A[2][i] * d2[1][0][a]
                                        + A[3][i] * d2[1][1][a] + A[4][i] * d2[2][0][a] + A[5][i] * d2[2][1][a]);
            for (int a = 0; a < 3; a++) hi[a] = JAFM * other[a][i];
            for (int a = 0; a < 3; a++) H[a] = c->H[a][i] + ha[a] + he[a] + hx[a] + hi[a] + c->Hext[a];
            if (c->if_em && (c->flags & GF_FLAG_EM_BACKACTION_M))   /* magnetic_AFM_solver.cpp:222-226 */
                for (int a = 0; a < 3; a++) H[a] = H[a] + c->Hcell[a][i];
            double alpha = sub == 0 ? mat->damping_AFM1 : mat->damping_AFM2;
            double gyro = sub == 0 ? mat->gyro_AFM1 : mat->gyro_AFM2;
            double precess = -gyro / (1. + alpha * alpha) * c->dt;
            double px = mgy * H[2] - mgz * H[1], py = mgz * H[0] - mgx * H[2], pz = mgx * H[1] - mgy * H[0];
            double qx = mgy * pz - mgz * py, qy = mgz * px - mgx * pz, qz = mgx * py - mgy * px;
            out[0][i] = precess * px + alpha * precess * qx;
            out[1][i] = precess * py + alpha * precess * qy;
            out[2][i] = precess * pz + alpha * precess * qz;
        }
    }
    if (stage == 4)
        for (int a = 0; a < 3; a++)
            #pragma omp parallel for schedule(static) if (c->n >= GFO_OMP_MIN)
            for (int64_t i = 0; i < c->n; i++) {
                c->dm1[a][i] = c->k1[0][a][i] / 6. + c->k1[1][a][i] / 3. + c->k1[2][a][i] / 3. + c->k1[3][a][i] / 6.;
                c->dm2[a][i] = c->k2[0][a][i] / 6. + c->k2[1][a][i] / 3. + c->k2[2][a][i] / 3. + c->k2[3][a][i] / 6.;
            }
    if (c->spin_on) spin_pump(c, stage);     /* magnetic_AFM_solver.cpp:426-428 */
}

/* magnetic_AFM_solver.cpp:1731-1834 */
void gfo_afm_update(gfo_ctx* c, int stage) {
    if (stage == 4 && c->spin_on) spin_pump_step(c);
    for (int sub = 0; sub < 2; sub++)
        for (int a = 0; a < 3; a++) {
            double* m = sub == 0 ? c->m1[a] : c->m2[a];
            double* s = sub == 0 ? c->ms1[a] : c->ms2[a];
            if (stage == 1 || stage == 2) {
                const double* k = (sub == 0 ? c->k1 : c->k2)[stage - 1][a];
                for (int64_t i = 0; i < c->n; i++) s[i] = m[i] + k[i] * 0.5;
            } else if (stage == 3) {
                const double* k = (sub == 0 ? c->k1 : c->k2)[2][a];
                for (int64_t i = 0; i < c->n; i++) s[i] = m[i] + k[i];
            } else {
                const double* dm = sub == 0 ? c->dm1[a] : c->dm2[a];
                for (int64_t i = 0; i < c->n; i++) { m[i] = m[i] + dm[i]; s[i] = m[i]; }
            }
        }
}

/* ------------------------------------------------------- FE stage */
/* ferroelectric_solver.cpp:2002-2029 */
static void update_external_E(gfo_ctx* c, int stage) {
    double local_time = stage == 1 ? c->time - c->dt : (stage == 4 ? c->time : c->time - c->dt * 0.5);
    for (int a = 0; a < 3; a++)
        c->Eext[a] = c->Eext_stat[a] + c->Eext_altr[a] * sin(2. * PI * c->E_altr_freq[a] * local_time);
}

/* ferroelectric_solver.cpp:4-131: 7-point Laplacian of the p store, a
 * surface face mirrors the centre value. d2[comp][axis]. */
static void fe_laplacian(const gfo_ctx* c, int64_t i, int64_t j, int64_t k, double d2[3][3]) {
    int64_t nx = c->nx, ny = c->ny, nz = c->nz;
    double cen[3], bw[3], fw[3];
    int64_t id = ID(c, i, j, k);
    for (int a = 0; a < 3; a++) cen[a] = c->ps[a][id];
    double dd[3] = {c->dx, c->dy, c->dz};
    for (int ax = 0; ax < 3; ax++) {
        int sb, sf;
        int64_t nb, nf;
        if (ax == 0) {
            sb = sYZ(c->fsurfYZ, c, i, j, k); sf = sYZ(c->fsurfYZ, c, i + 1, j, k);
            nb = ID(c, i == 0 ? nx - 1 : i - 1, j, k); nf = ID(c, i == nx - 1 ? 0 : i + 1, j, k);
        } else if (ax == 1) {
            sb = sXZ(c->fsurfXZ, c, i, j, k); sf = sXZ(c->fsurfXZ, c, i, j + 1, k);
            nb = ID(c, i, j == 0 ? ny - 1 : j - 1, k); nf = ID(c, i, j == ny - 1 ? 0 : j + 1, k);
        } else {
            sb = sXY(c->fsurfXY, c, i, j, k); sf = sXY(c->fsurfXY, c, i, j, k + 1);
            nb = ID(c, i, j, k == 0 ? nz - 1 : k - 1); nf = ID(c, i, j, k == nz - 1 ? 0 : k + 1);
        }
        for (int a = 0; a < 3; a++) {
            bw[a] = sb ? cen[a] : c->ps[a][nb];
            fw[a] = sf ? cen[a] : c->ps[a][nf];
            d2[a][ax] = (fw[a] - 2. * cen[a] + bw[a]) / dd[ax] / dd[ax];
        }
    }
}

/* get_p_gradient (ferroelectric_solver.cpp:134-281) of the p store at an FE
 * cell: (forward - backward) / h / 2 with the cell's own value across an
 * FE surface face, halved at the surface too.  g[comp][axis]. */
static void fe_p_gradient(const gfo_ctx* c, int64_t i, int64_t j, int64_t k, double g[3][3]) {
    int64_t nx = c->nx, ny = c->ny, nz = c->nz;
    int64_t id = ID(c, i, j, k);
    double dd[3] = {c->dx, c->dy, c->dz};
    for (int ax = 0; ax < 3; ax++) {
        int sb, sf;
        int64_t nb, nf;
        if (ax == 0) {
            sb = sYZ(c->fsurfYZ, c, i, j, k); sf = sYZ(c->fsurfYZ, c, i + 1, j, k);
            nb = ID(c, i == 0 ? nx - 1 : i - 1, j, k); nf = ID(c, i == nx - 1 ? 0 : i + 1, j, k);
        } else if (ax == 1) {
            sb = sXZ(c->fsurfXZ, c, i, j, k); sf = sXZ(c->fsurfXZ, c, i, j + 1, k);
            nb = ID(c, i, j == 0 ? ny - 1 : j - 1, k); nf = ID(c, i, j == ny - 1 ? 0 : j + 1, k);
        } else {
            sb = sXY(c->fsurfXY, c, i, j, k); sf = sXY(c->fsurfXY, c, i, j, k + 1);
            nb = ID(c, i, j, k == 0 ? nz - 1 : k - 1); nf = ID(c, i, j, k == nz - 1 ? 0 : k + 1);
        }
        for (int a = 0; a < 3; a++) {
            double bw = sb ? c->ps[a][id] : c->ps[a][nb];
            double fw = sf ? c->ps[a][id] : c->ps[a][nf];
            g[a][ax] = (fw - bw) / dd[ax] / 2.;
        }
    }
}

static void strain_gradients(const gfo_ctx* c, int64_t i, int64_t j, int64_t k, double d[6][3]);

/* get_dq_RKs (ferroelectric_solver.cpp:534-842, RK2-4 at :844, :1154,
 * :1464 differ only in the drive time) and get_dp_RKs (:1786-1832).
 * Strain terms are zero without an elastic solver; no flexo, no EM. */
void gfo_fe_get(gfo_ctx* c, int stage) {
    if (!(c->flags & GF_FLAG_PRESCRIBE_EEXT)) update_external_E(c, stage);
    int64_t ny = c->ny, nz = c->nz;
    double** dq = c->kq[stage - 1];
    double** dp = c->kp[stage - 1];
    #pragma omp parallel for schedule(static) if (c->n >= GFO_OMP_MIN)
    for (int64_t id = 0; id < c->n; id++) {
        int64_t i = id / (ny * nz), j = (id - i * (ny * nz)) / nz, k = id - i * (ny * nz) - j * nz;
        const gf_material* mat = matof(c, id);
        if (!mat || !mat->if_FE) continue;
        double pgx = c->ps[0][id], pgy = c->ps[1][id], pgz = c->ps[2][id];
        double d2[3][3];
        fe_laplacian(c, i, j, k, d2);
        double px, py, pz;
        glb2crt(c, pgx, pgy, pgz, &px, &py, &pz);
        double exx = 0., eyy = 0., ezz = 0., eyz = 0., exz = 0., exy = 0.;
        double exx0 = 0, eyy0 = 0, ezz0 = 0, eyz0 = 0, exz0 = 0, exy0 = 0;
        if (c->flags & GF_FLAG_ELASTO_ON_PM) {   /* ferroelectric_solver.cpp:624-653 */
            exx = c->e_crt[0][id]; eyy = c->e_crt[1][id]; ezz = c->e_crt[2][id];
            eyz = c->e_crt[3][id]; exz = c->e_crt[4][id]; exy = c->e_crt[5][id];
            exx0 = c->e0_crt[0][id]; eyy0 = c->e0_crt[1][id]; ezz0 = c->e0_crt[2][id];
            eyz0 = c->e0_crt[3][id]; exz0 = c->e0_crt[4][id]; exy0 = c->e0_crt[5][id];
            if (c->flags & GF_FLAG_ELASTOSTATIC) {
                /* :636-650: the static strain and eigenstrain added (the
                 * dynamic ones are zero without elastodynamics) */
                exx = exx + c->e_t0[0][id]; eyy = eyy + c->e_t0[1][id]; ezz = ezz + c->e_t0[2][id];
                eyz = eyz + c->e_t0[3][id]; exz = exz + c->e_t0[4][id]; exy = exy + c->e_t0[5][id];
                exx0 = exx0 + c->e0t0[0][id]; eyy0 = eyy0 + c->e0t0[1][id]; ezz0 = ezz0 + c->e0t0[2][id];
                eyz0 = eyz0 + c->e0t0[3][id]; exz0 = exz0 + c->e0t0[4][id]; exy0 = exy0 + c->e0t0[5][id];
            }
        }
        double a1 = mat->a1, a11 = mat->a11, a12 = mat->a12, a111 = mat->a111, a112 = mat->a112, a123 = mat->a123;
        double a1111 = mat->a1111, a1112 = mat->a1112, a1122 = mat->a1122, a1123 = mat->a1123;
        double tv1 = mat->tv1, tv2 = mat->tv2, tv3 = mat->tv3, G11 = mat->G11;
        double FE_mass = -1. * c->dt / mat->FE_mass;
        double FE_damping = FE_mass * mat->FE_damping;
        double qx_c = FE_mass * (
            2. * a1 * px + 4. * a11 * pow(px, 3.)
            + 2. * a12 * (pow(py, 2.) + pow(pz, 2.)) * px
            + 6. * a111 * pow(px, 5.)
            + 4. * a112 * pow(px, 3.) * (pow(py, 2.) + pow(pz, 2.))
            + 2. * a112 * (pow(py, 4.) + pow(pz, 4.)) * px
            + 2. * a123 * px * pow(py, 2.) * pow(pz, 2.)
            + 8. * a1111 * pow(px, 7.)
            + 2. * a1112 * px * (pow(py, 6.) + pow(pz, 6.) + 3. * pow(px, 4.) * (pow(py, 2.) + pow(pz, 2.)))
            + 4. * a1122 * pow(px, 3.) * (pow(py, 4.) + pow(pz, 4.))
            + 2. * a1123 * px * pow(py, 2.) * pow(pz, 2.) * (2. * pow(px, 2.) + pow(py, 2.) + pow(pz, 2.))
            + (tv1 * (exx0 - exx) + tv2 * (eyy0 - eyy + ezz0 - ezz)) * px
            + tv3 * ((exy0 - exy) * py + (exz0 - exz) * pz));
        double qy_c = FE_mass * (
            2. * a1 * py + 4. * a11 * pow(py, 3.)
            + 2. * a12 * (pow(pz, 2.) + pow(px, 2.)) * py
            + 6. * a111 * pow(py, 5.)
            + 4. * a112 * pow(py, 3.) * (pow(px, 2.) + pow(pz, 2.))
            + 2. * a112 * (pow(px, 4.) + pow(pz, 4.)) * py
            + 2. * a123 * py * pow(px, 2.) * pow(pz, 2.)
            + 8. * a1111 * pow(py, 7)
            + 2. * a1112 * py * (pow(px, 6.) + pow(pz, 6.) + 3. * pow(py, 4.) * (pow(px, 2.) + pow(pz, 2.)))
            + 4. * a1122 * pow(py, 3.) * (pow(px, 4.) + pow(pz, 4.))
            + 2. * a1123 * py * pow(px, 2.) * pow(pz, 2.) * (pow(px, 2.) + 2 * pow(py, 2.) + pow(pz, 2.))
            + (tv1 * (eyy0 - eyy) + tv2 * (exx0 - exx + ezz0 - ezz)) * py
            + tv3 * ((exy0 - exy) * px + (eyz0 - eyz) * pz));
        double qz_c = FE_mass * (
            2. * a1 * pz + 4. * a11 * pow(pz, 3.)
            + 2. * a12 * (pow(py, 2.) + pow(px, 2.)) * pz
            + 6. * a111 * pow(pz, 5.)
            + 4. * a112 * pow(pz, 3.) * (pow(py, 2.) + pow(px, 2.))
            + 2. * a112 * (pow(py, 4.) + pow(px, 4.)) * pz
            + 2. * a123 * pz * pow(py, 2.) * pow(px, 2.)
            + 8. * a1111 * pow(pz, 7.)
            + 2. * a1112 * pz * (pow(py, 6.) + pow(px, 6.) + 3. * pow(pz, 4.) * (pow(py, 2.) + pow(px, 2.)))
            + 4. * a1122 * pow(pz, 3.) * (pow(py, 4.) + pow(px, 4.))
            + 2. * a1123 * pz * pow(px, 2.) * pow(py, 2.) * (pow(px, 2.) + pow(py, 2.) + 2. * pow(pz, 2.))
            + (tv1 * (ezz0 - ezz) + tv2 * (exx0 - exx + eyy0 - eyy)) * pz
            + tv3 * ((exz0 - exz) * px + (eyz0 - eyz) * py));
        double qg[3];
        crt2glb(c, qx_c, qy_c, qz_c, &qg[0], &qg[1], &qg[2]);
        for (int a = 0; a < 3; a++)
            dq[a][id] = qg[a] + FE_mass * (-1. * G11 * (d2[a][0] + d2[a][1] + d2[a][2]) - c->Es[a][id] - c->Eext[a])
                      + FE_damping * c->qs[a][id];
        if (c->if_em && (c->flags & GF_FLAG_EM_BACKACTION_P))   /* ferroelectric_solver.cpp:794-798 */
            for (int a = 0; a < 3; a++) dq[a][id] = dq[a][id] - FE_mass * c->Ecell[a][id];
        if (c->flexo_on) {
            /* ferroelectric_solver.cpp:800-832; the strain gradients are
             * those of the dynamic strain (zero without elastodynamics) */
            const double* f = c->flx[c->cell[id] - 1];
            double tv4 = f[3], tv5 = f[4], tv6 = f[5], tv7 = f[6], tv8 = f[7], f11 = f[8], f12 = f[9], f44 = f[10];
            double g[3][3], e[6][3] = {{0.}};
            fe_p_gradient(c, i, j, k, g);
            if (c->if_elasto) strain_gradients(c, i, j, k, e);
            /* e[comp][axis], comp xx yy zz yz xz xy */
            dq[0][id] = dq[0][id] + FE_mass * (
                + tv4 * pgx * g[0][0]
                + tv5 * (pgy * g[1][0] + pgz * g[2][0])
                + tv6 * (pgx * g[1][1] + pgx * g[2][2] + pgy * g[0][1] + pgz * g[0][2])
                - tv7 * d2[0][0]
                - tv8 * (d2[0][2] + d2[0][1])
                - f11 * e[0][0] - f12 * (e[1][0] + e[2][0])
                - 2. * f44 * (e[5][1] + e[4][2]));
            dq[1][id] = dq[1][id] + FE_mass * (
                + tv4 * pgy * g[1][1]
                + tv5 * (pgx * g[0][1] + pgz * g[2][1])
                + tv6 * (pgy * g[0][0] + pgy * g[2][2] + pgx * g[1][0] + pgz * g[1][2])
                - tv7 * d2[1][1]
                - tv8 * (d2[1][0] + d2[1][2])
                - f11 * e[1][1] - f12 * (e[0][1] + e[2][1])
                - 2. * f44 * (e[5][0] + e[3][2]));
            dq[2][id] = dq[2][id] + FE_mass * (
                + tv4 * pgz * g[2][2]
                + tv5 * (pgx * g[0][2] + pgy * g[1][2])
                + tv6 * (pgz * g[0][0] + pgz * g[1][1] + pgx * g[2][0] + pgy * g[2][1])
                - tv7 * d2[2][2]
                - tv8 * (d2[2][0] + d2[2][1])
                - f11 * e[2][2] - f12 * (e[0][2] + e[1][2])
                - 2. * f44 * (e[3][1] + e[4][0]));
        }
    }
    for (int a = 0; a < 3; a++)
        for (int64_t id = 0; id < c->n; id++) dp[a][id] = c->qs[a][id] * c->dt;
}

/* update_q_RKs / update_p_RKs / update_q / update_p
 * (ferroelectric_solver.cpp:1835-1999): the state accumulates term by term,
 * q = q + k1/6 + k2/3 + k3/3 + k4/6, evaluated left to right. */
void gfo_fe_update(gfo_ctx* c, int stage) {
    for (int w = 0; w < 2; w++)
        for (int a = 0; a < 3; a++) {
            double* y = w == 0 ? c->q[a] : c->p[a];
            double* s = w == 0 ? c->qs[a] : c->ps[a];
            const double* k1 = (w == 0 ? c->kq[0] : c->kp[0])[a];
            const double* k2 = (w == 0 ? c->kq[1] : c->kp[1])[a];
            const double* k3 = (w == 0 ? c->kq[2] : c->kp[2])[a];
            const double* kl = (w == 0 ? c->kq[3] : c->kp[3])[a];
            #pragma omp parallel for schedule(static) if (c->n >= GFO_OMP_MIN)
            for (int64_t i = 0; i < c->n; i++) {
                if (stage == 1) s[i] = y[i] + k1[i] * 0.5;
                else if (stage == 2) s[i] = y[i] + k2[i] * 0.5;
                else if (stage == 3) s[i] = y[i] + k3[i];
                else { y[i] = y[i] + k1[i] / 6. + k2[i] / 3. + k3[i] / 3. + kl[i] / 6.; s[i] = y[i]; }
            }
        }
}

/* --------------------------------------------------- elastodynamics */
static int sfx(const gfo_ctx* c, int64_t i, int64_t j, int64_t k) { return c->sfYZ[(i * c->ny + j) * c->nz + k]; }
static int sfy(const gfo_ctx* c, int64_t i, int64_t j, int64_t k) { return c->sfXZ[(i * (c->ny + 1) + j) * c->nz + k]; }
static int sfz(const gfo_ctx* c, int64_t i, int64_t j, int64_t k) { return c->sfXY[(i * c->ny + j) * (c->nz + 1) + k]; }
#define GYZ(a, i, j, k) c->gYZ[a][((i) * c->ny + (j)) * c->nz + (k)]
#define GXZ(a, i, j, k) c->gXZ[a][((i) * (c->ny + 1) + (j)) * c->nz + (k)]
#define GXY(a, i, j, k) c->gXY[a][((i) * c->ny + (j)) * (c->nz + 1) + (k)]

/* get_Dstrain_gradient (elastic_dynamic_solver.cpp:230-399) at a solid cell:
 * along each axis the normal strain from its two face gradients, the other
 * components as (forward - backward) / h / scale, central (scale 2) between
 * two faces that are not stress free, one-sided (scale 1) against a free
 * face, zero between two free faces.  d[comp][axis]. */
static void strain_gradients(const gfo_ctx* c, int64_t i, int64_t j, int64_t k, double d[6][3]) {
    int64_t nx = c->nx, ny = c->ny, nz = c->nz;
    double dd[3] = {c->dx, c->dy, c->dz};
    for (int ax = 0; ax < 3; ax++) {
        int sb, sf;
        int64_t pos, n;
        if (ax == 0) { sb = sfx(c, i, j, k); sf = sfx(c, i + 1, j, k); pos = i; n = nx; }
        else if (ax == 1) { sb = sfy(c, i, j, k); sf = sfy(c, i, j + 1, k); pos = j; n = ny; }
        else { sb = sfz(c, i, j, k); sf = sfz(c, i, j, k + 1); pos = k; n = nz; }
        int64_t bwd, fwd;
        double scale;
        int64_t prev = pos == 0 ? n - 1 : pos - 1, next = pos == n - 1 ? 0 : pos + 1;
        if (sb == sf) {
            if (sb) { bwd = pos; fwd = pos; scale = 1.; }
            else { scale = 2.; bwd = prev; fwd = next; }
        } else if (sb) {
            bwd = pos; scale = 1.; fwd = next;
        } else {
            fwd = pos; scale = 1.; bwd = prev;
        }
        int64_t ib = ax == 0 ? ID(c, bwd, j, k) : (ax == 1 ? ID(c, i, bwd, k) : ID(c, i, j, bwd));
        int64_t ifw = ax == 0 ? ID(c, fwd, j, k) : (ax == 1 ? ID(c, i, fwd, k) : ID(c, i, j, fwd));
        for (int e = 0; e < 6; e++) d[e][ax] = (c->e_glb[e][ifw] - c->e_glb[e][ib]) / dd[ax] / scale;
        if (ax == 0) d[0][0] = (GYZ(0, i + 1, j, k) - GYZ(0, i, j, k)) / c->dx;
        else if (ax == 1) d[1][1] = (GXZ(1, i, j + 1, k) - GXZ(1, i, j, k)) / c->dy;
        else d[2][2] = (GXY(2, i, j, k + 1) - GXY(2, i, j, k)) / c->dz;
    }
}

/* mathlib.cpp:78-125 */
static void mat_rot(const double T[3][3], const double in[3][3], double out[3][3]) {
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            out[i][j] = 0.;
            for (int k = 0; k < 3; k++)
                for (int l = 0; l < 3; l++) out[i][j] = out[i][j] + T[i][k] * T[j][l] * in[k][l];
        }
}

/* elastic_dynamic_solver.cpp:4-104 (static eigenstrain and flexo terms zero) */
static void elastic_eigenstrain(gfo_ctx* c) {
    #pragma omp parallel for schedule(static) if (c->n >= GFO_OMP_MIN)
    for (int64_t id = 0; id < c->n; id++) {
        const gf_material* mat = matof(c, id);
        if (!mat) {
            for (int a = 0; a < 6; a++) { c->e0_crt[a][id] = 0.; c->e0_glb[a][id] = 0.; }
            continue;
        }
        double Q11 = mat->Q11, Q12 = mat->Q12, Q44 = mat->Q44;
        double l100 = mat->lamda100, l111 = mat->lamda111;
        double l100a = mat->lamda100_AFM1, l111a = mat->lamda111_AFM1, l100b = mat->lamda100_AFM2, l111b = mat->lamda111_AFM2;
        double px, py, pz, mx, my, mz, ax, ay, az, bx, by, bz;
        glb2crt(c, c->ps[0][id], c->ps[1][id], c->ps[2][id], &px, &py, &pz);
        glb2crt(c, c->ms[0][id], c->ms[1][id], c->ms[2][id], &mx, &my, &mz);
        glb2crt(c, c->ms1[0][id], c->ms1[1][id], c->ms1[2][id], &ax, &ay, &az);
        glb2crt(c, c->ms2[0][id], c->ms2[1][id], c->ms2[2][id], &bx, &by, &bz);
        double s[3][3], o[3][3];
        s[0][0] = 1.5 * l100 * (pow(mx, 2.) - 1. / 3.) + Q11 * pow(px, 2.) + Q12 * (pow(py, 2.) + pow(pz, 2.))
                + 0.5 * 1.5 * l100a * (pow(ax, 2.) - 1. / 3.) + 0.5 * 1.5 * l100b * (pow(bx, 2.) - 1. / 3.) - c->e0t0[0][id];
        s[1][1] = 1.5 * l100 * (pow(my, 2.) - 1. / 3.) + Q11 * pow(py, 2.) + Q12 * (pow(px, 2.) + pow(pz, 2.))
                + 0.5 * 1.5 * l100a * (pow(ay, 2.) - 1. / 3.) + 0.5 * 1.5 * l100b * (pow(by, 2.) - 1. / 3.) - c->e0t0[1][id];
        s[2][2] = 1.5 * l100 * (pow(mz, 2.) - 1. / 3.) + Q11 * pow(pz, 2.) + Q12 * (pow(py, 2.) + pow(px, 2.))
                + 0.5 * 1.5 * l100a * (pow(az, 2.) - 1. / 3.) + 0.5 * 1.5 * l100b * (pow(bz, 2.) - 1. / 3.) - c->e0t0[2][id];
        s[1][2] = 1.5 * l111 * my * mz + Q44 * py * pz + 0.5 * 1.5 * l111a * ay * az + 0.5 * 1.5 * l111b * by * bz - c->e0t0[3][id];
        s[0][2] = 1.5 * l111 * mx * mz + Q44 * px * pz + 0.5 * 1.5 * l111a * ax * az + 0.5 * 1.5 * l111b * bx * bz - c->e0t0[4][id];
        s[0][1] = 1.5 * l111 * my * mx + Q44 * py * px + 0.5 * 1.5 * l111a * ay * ax + 0.5 * 1.5 * l111b * by * bx - c->e0t0[5][id];
        if (c->flexo_on && c->if_FE_all) {
            /* elastic_dynamic_solver.cpp:82-89: the global-frame p gradient
             * of FE cells (zero elsewhere) enters the crystal-frame tensor */
            const double* f = c->flx[c->cell[id] - 1];
            double F11 = f[0], F12 = f[1], F44 = f[2], g[3][3] = {{0.}};
            if (mat->if_FE) {
                int64_t i = id / (c->ny * c->nz), j = (id / c->nz) % c->ny, k = id % c->nz;
                fe_p_gradient(c, i, j, k, g);
            }
            s[0][0] = s[0][0] - F11 * g[0][0] - F12 * (g[2][2] + g[1][1]);
            s[1][1] = s[1][1] - F11 * g[1][1] - F12 * (g[2][2] + g[0][0]);
            s[2][2] = s[2][2] - F11 * g[2][2] - F12 * (g[0][0] + g[1][1]);
            s[1][2] = s[1][2] - F44 / 2. * (g[1][2] + g[2][1]);
            s[0][2] = s[0][2] - F44 / 2. * (g[0][2] + g[2][0]);
            s[0][1] = s[0][1] - F44 / 2. * (g[1][0] + g[0][1]);
        }
        c->e0_crt[0][id] = s[0][0]; c->e0_crt[1][id] = s[1][1]; c->e0_crt[2][id] = s[2][2];
        c->e0_crt[3][id] = s[1][2]; s[2][1] = s[1][2];
        c->e0_crt[4][id] = s[0][2]; s[2][0] = s[0][2];
        c->e0_crt[5][id] = s[0][1]; s[1][0] = s[0][1];
        if (c->flags & GF_FLAG_ORIENT_ROTATE) mat_rot(c->Tc2g, s, o);
        else memcpy(o, s, sizeof o);
        c->e0_glb[0][id] = o[0][0]; c->e0_glb[1][id] = o[1][1]; c->e0_glb[2][id] = o[2][2];
        c->e0_glb[3][id] = o[1][2]; c->e0_glb[4][id] = o[0][2]; c->e0_glb[5][id] = o[0][1];
    }
}

/* elastic_dynamic_solver.cpp:107-187 */
static void elastic_gradients(gfo_ctx* c) {
    int64_t nx = c->nx, ny = c->ny, nz = c->nz;
    double** s = c->us;
    #pragma omp parallel for collapse(3) schedule(static) if (c->n >= GFO_OMP_MIN)
    for (int64_t i = 0; i <= nx; i++)
        for (int64_t j = 0; j < ny; j++)
            for (int64_t k = 0; k < nz; k++)
                for (int a = 0; a < 3; a++) {
                    double g;
                    if (sfx(c, i, j, k)) g = 0.;
                    else if (i == 0 || i == nx) g = (s[a][ID(c, 0, j, k)] - s[a][ID(c, nx - 1, j, k)]) / c->dx;
                    else g = (s[a][ID(c, i, j, k)] - s[a][ID(c, i - 1, j, k)]) / c->dx;
                    GYZ(a, i, j, k) = g;
                }
    #pragma omp parallel for collapse(3) schedule(static) if (c->n >= GFO_OMP_MIN)
    for (int64_t i = 0; i < nx; i++)
        for (int64_t j = 0; j <= ny; j++)
            for (int64_t k = 0; k < nz; k++)
                for (int a = 0; a < 3; a++) {
                    double g;
                    if (sfy(c, i, j, k)) g = 0.;
                    else if (j == 0 || j == ny) g = (s[a][ID(c, i, 0, k)] - s[a][ID(c, i, ny - 1, k)]) / c->dy;
                    else g = (s[a][ID(c, i, j, k)] - s[a][ID(c, i, j - 1, k)]) / c->dy;
                    GXZ(a, i, j, k) = g;
                }
    #pragma omp parallel for collapse(3) schedule(static) if (c->n >= GFO_OMP_MIN)
    for (int64_t i = 0; i < nx; i++)
        for (int64_t j = 0; j < ny; j++)
            for (int64_t k = 0; k <= nz; k++)
                for (int a = 0; a < 3; a++) {
                    double g;
                    if (sfz(c, i, j, k)) g = 0.;
                    else if (k == 0 || k == nz) g = (s[a][ID(c, i, j, 0)] - s[a][ID(c, i, j, nz - 1)]) / c->dz;
                    else g = (s[a][ID(c, i, j, k)] - s[a][ID(c, i, j, k - 1)]) / c->dz;
                    GXY(a, i, j, k) = g;
                }
}

/* elastic_dynamic_solver.cpp:190-227 */
static void elastic_strain(gfo_ctx* c) {
    #pragma omp parallel for collapse(3) schedule(static) if (c->n >= GFO_OMP_MIN)
    for (int64_t i = 0; i < c->nx; i++)
        for (int64_t j = 0; j < c->ny; j++)
            for (int64_t k = 0; k < c->nz; k++) {
                int64_t id = ID(c, i, j, k);
                double s[3][3], o[3][3];
                s[0][0] = (GYZ(0, i, j, k) + GYZ(0, i + 1, j, k)) / 2.;
                s[1][1] = (GXZ(1, i, j, k) + GXZ(1, i, j + 1, k)) / 2.;
                s[2][2] = (GXY(2, i, j, k) + GXY(2, i, j, k + 1)) / 2.;
                s[1][2] = (GXY(1, i, j, k) + GXY(1, i, j, k + 1) + GXZ(2, i, j, k) + GXZ(2, i, j + 1, k)) / 4.;
                s[2][1] = s[1][2];
                s[0][2] = (GXY(0, i, j, k) + GXY(0, i, j, k + 1) + GYZ(2, i, j, k) + GYZ(2, i + 1, j, k)) / 4.;
                s[2][0] = s[0][2];
                s[0][1] = (GXZ(0, i, j, k) + GXZ(0, i, j + 1, k) + GYZ(1, i, j, k) + GYZ(1, i + 1, j, k)) / 4.;
                s[1][0] = s[0][1];
                if (c->flags & GF_FLAG_ORIENT_ROTATE) mat_rot(c->Tg2c, s, o);
                else memcpy(o, s, sizeof o);
                c->e_glb[0][id] = s[0][0]; c->e_crt[0][id] = o[0][0];
                c->e_glb[1][id] = s[1][1]; c->e_crt[1][id] = o[1][1];
                c->e_glb[2][id] = s[2][2]; c->e_crt[2][id] = o[2][2];
                c->e_glb[3][id] = s[1][2]; c->e_crt[3][id] = o[1][2];
                c->e_glb[4][id] = s[0][2]; c->e_crt[4][id] = o[0][2];
                c->e_glb[5][id] = s[0][1]; c->e_crt[5][id] = o[0][1];
            }
}

/* mathlib.cpp:128-135 */
static void c_dot(const double C[6][6], const double e[6], double s[6]) {
    for (int i = 0; i < 6; i++) {
        s[i] = 0.;
        for (int j = 0; j < 6; j++) s[i] = s[i] + C[i][j] * e[j];
    }
}

#define EG(a, id) c->e_glb[a][id]
#define E0(a, id) c->e0_glb[a][id]

/* stress on a YZ face fi seen from cell (qi,j,k) (elastic_dynamic_solver.cpp:876-933) */
static void stress_yz(const gfo_ctx* c, int64_t fi, int64_t qi, int64_t j, int64_t k, double st[6]) {
    int64_t q = ID(c, qi, j, k);
    double e[6];
    e[0] = GYZ(0, fi, j, k) - E0(0, q);
    e[1] = EG(1, q) - E0(1, q);
    e[2] = EG(2, q) - E0(2, q);
    e[3] = (EG(3, q) - E0(3, q)) * 2.;
    e[4] = ((GXY(0, qi, j, k) + GXY(0, qi, j, k + 1) + GYZ(2, fi, j, k) * 2.) / 4. - E0(4, q)) * 2.;
    e[5] = ((GXZ(0, qi, j, k) + GXZ(0, qi, j + 1, k) + GYZ(1, fi, j, k) * 2.) / 4. - E0(5, q)) * 2.;
    c_dot(matof(c, q)->cijkl_glb, e, st);
}

/* stress on an XZ face fj seen from cell (i,qj,k) (:994-1050) */
static void stress_xz(const gfo_ctx* c, int64_t i, int64_t fj, int64_t qj, int64_t k, double st[6]) {
    int64_t q = ID(c, i, qj, k);
    double e[6];
    e[0] = EG(0, q) - E0(0, q);
    e[1] = GXZ(1, i, fj, k) - E0(1, q);
    e[2] = EG(2, q) - E0(2, q);
    e[3] = ((GXY(1, i, qj, k) + GXY(1, i, qj, k + 1) + GXZ(2, i, fj, k) * 2.) / 4. - E0(3, q)) * 2.;
    e[4] = (EG(4, q) - E0(4, q)) * 2.;
    e[5] = ((GXZ(0, i, fj, k) * 2. + GYZ(1, i, qj, k) + GYZ(1, i + 1, qj, k)) / 4. - E0(5, q)) * 2.;
    c_dot(matof(c, q)->cijkl_glb, e, st);
}

/* stress on an XY face fk seen from cell with flat index q and z index qk
 * (:1112-1166 and :1229-1290; q may alias past the column, see below) */
static void stress_xy(const gfo_ctx* c, int64_t i, int64_t j, int64_t fk, int64_t qk, double st[6]) {
    int64_t q = (i * c->ny + j) * c->nz + qk;
    double e[6];
    e[0] = EG(0, q) - E0(0, q);
    e[1] = EG(1, q) - E0(1, q);
    e[2] = GXY(2, i, j, fk) - E0(2, q);
    e[3] = ((GXY(1, i, j, fk) * 2. + GXZ(2, i, j, qk) + GXZ(2, i, j + 1, qk)) / 4. - E0(3, q)) * 2.;
    e[4] = ((GXY(0, i, j, fk) * 2. + GYZ(2, i, j, qk) + GYZ(2, i + 1, j, qk)) / 4. - E0(4, q)) * 2.;
    e[5] = (EG(5, q) - E0(5, q)) * 2.;
    c_dot(matof(c, q)->cijkl_glb, e, st);
}

/* get_dv_RKs (elastic_dynamic_solver.cpp:841-1329; RK2-4 :1331, :1817,
 * :2303): stiffness damping (F - F_store) in RK1 (which then stores F),
 * 2(F - F_store) in RK2 and RK3, (F - F_store) in RK4; mass damping on
 * the base velocity in every stage. */
void gfo_elastic_get(gfo_ctx* c, int stage) {
    if (c->flags & GF_FLAG_ELASTO_BACKACTION) elastic_eigenstrain(c);
    int64_t nx = c->nx, ny = c->ny, nz = c->nz;
    double** dv = c->kv[stage - 1];
    double** du = c->ku[stage - 1];
    double ts = stage == 1 ? c->time - c->dt : (stage == 4 ? c->time : c->time - 0.5 * c->dt);
    #pragma omp parallel for collapse(3) schedule(static) if (c->n >= GFO_OMP_MIN)
    for (int64_t i = 0; i < nx; i++)
        for (int64_t j = 0; j < ny; j++)
            for (int64_t k = 0; k < nz; k++) {
                int64_t id = ID(c, i, j, k);
                const gf_material* mat = matof(c, id);
                if (!mat) { dv[0][id] = 0.; dv[1][id] = 0.; dv[2][id] = 0.; continue; }
                double s1x[6] = {0}, s2x[6] = {0}, s1y[6] = {0}, s2y[6] = {0}, s1z[6] = {0}, s2z[6] = {0}, t[6];
                if (!sfx(c, i, j, k)) {
                    stress_yz(c, i, i, j, k, s1x);
                    stress_yz(c, i, i == 0 ? nx - 1 : i - 1, j, k, t);
                    for (int a = 0; a < 6; a++) s1x[a] = (s1x[a] + t[a]) / 2.;
                }
                if (!sfx(c, i + 1, j, k)) {
                    stress_yz(c, i + 1, i, j, k, s2x);
                    stress_yz(c, i + 1, i == nx - 1 ? 0 : i + 1, j, k, t);
                    for (int a = 0; a < 6; a++) s2x[a] = (s2x[a] + t[a]) / 2.;
                }
                if (!sfy(c, i, j, k)) {
                    stress_xz(c, i, j, j, k, s1y);
                    stress_xz(c, i, j, j == 0 ? ny - 1 : j - 1, k, t);
                    for (int a = 0; a < 6; a++) s1y[a] = (s1y[a] + t[a]) / 2.;
                }
                if (!sfy(c, i, j + 1, k)) {
                    stress_xz(c, i, j + 1, j, k, s2y);
                    stress_xz(c, i, j + 1, j == ny - 1 ? 0 : j + 1, k, t);
                    for (int a = 0; a < 6; a++) s2y[a] = (s2y[a] + t[a]) / 2.;
                }
                if (!sfz(c, i, j, k)) {
                    stress_xy(c, i, j, k, k, s1z);
                    stress_xy(c, i, j, k, k == 0 ? nz - 1 : k - 1, t);
                    for (int a = 0; a < 6; a++) s1z[a] = (s1z[a] + t[a]) / 2.;
                }
                if (!sfz(c, i, j, k + 1)) {
                    stress_xy(c, i, j, k + 1, k, s2z);
                    /* the reference tests k == ny-1 for the periodic wrap (:1247) */
                    stress_xy(c, i, j, k + 1, k == ny - 1 ? 0 : k + 1, t);
                    for (int a = 0; a < 6; a++) s2z[a] = (s2z[a] + t[a]) / 2.;
                } else if ((c->flags & GF_FLAG_GAUSSIAN_PULSE) && c->source_type == 3 && k == c->source_z) {
                    double tv = ts - 5.0 * c->sigma_gauss;
                    double amp = c->amplitude_gauss * exp(-1. * pow(tv, 2.) / (2. * pow(c->sigma_gauss, 2.)));
                    for (int a = 0; a < 6; a++) s2z[a] = 0.;
                    if (c->source_comp == 'x') s2z[4] = amp;
                    else if (c->source_comp == 'y') s2z[3] = amp;
                    else if (c->source_comp == 'z') s2z[2] = amp;
                }
                double density = c->dt / mat->density;
                double mass_damping = -c->dt * mat->elast_mass_damping;
                double stiff_damping = mat->elast_stiff_damping / mat->density;
                double f[3];
                f[0] = (s2x[0] - s1x[0]) / c->dx + (s2y[5] - s1y[5]) / c->dy + (s2z[4] - s1z[4]) / c->dz;
                f[1] = (s2x[5] - s1x[5]) / c->dx + (s2y[1] - s1y[1]) / c->dy + (s2z[3] - s1z[3]) / c->dz;
                f[2] = (s2x[4] - s1x[4]) / c->dx + (s2y[3] - s1y[3]) / c->dy + (s2z[2] - s1z[2]) / c->dz;
                for (int a = 0; a < 3; a++) {
                    c->force[a][id] = f[a];
                    if (stage == 2 || stage == 3)
                        dv[a][id] = f[a] * density + mass_damping * c->v[a][id] + stiff_damping * (f[a] - c->force_store[a][id]) * 2.;
                    else
                        dv[a][id] = f[a] * density + mass_damping * c->v[a][id] + stiff_damping * (f[a] - c->force_store[a][id]);
                    if (stage == 1) c->force_store[a][id] = f[a];
                }
            }
    /* get_du_RKs (:2801-2847) */
    for (int a = 0; a < 3; a++)
        for (int64_t id = 0; id < c->n; id++) du[a][id] = c->vs[a][id] * c->dt;
}

/* transfer_pointer (elastodynamics_AbsorbingBoundary.cpp:3-89) */
static void elastic_shift_history(gfo_ctx* c) {
    for (int a = 0; a < 3; a++) {
        if (!(c->flags & GF_FLAG_1D_ABC)) {
            memcpy(c->t4[a], c->t3[a], sizeof(double) * (size_t)c->n);
            memcpy(c->t3[a], c->t2[a], sizeof(double) * (size_t)c->n);
            memcpy(c->t2[a], c->t1[a], sizeof(double) * (size_t)c->n);
        }
        memcpy(c->t1[a], c->u[a], sizeof(double) * (size_t)c->n);
    }
}

/* update_Du_Boundary{_half,_full,} (elastodynamics_AbsorbingBoundary.cpp:92-230):
 * kind 0 = half (stages 1, 2), 1 = full (stage 3), 2 = after the step. */
static void elastic_boundary(gfo_ctx* c, int kind) {
    int64_t nz = c->nz;
    double dz = c->dz, dt = c->dt;
    for (int64_t i = 0; i < c->nx; i++)
        for (int64_t j = 0; j < c->ny; j++) {
            if (!(c->flags & GF_FLAG_1D_ABC)) {
                double w = c->w_liao;
#if GFO_PLANT == 3
                double **h1 = c->t2, **h2 = c->t3, **h3 = c->t4, **h4 = c->t4;
#else
                double **h1 = c->t1, **h2 = c->t2, **h3 = c->t3, **h4 = c->t4;
#endif
#if GFO_PLANT == 4
                double c1 = 6. * w + (1. - w), c2 = 4. * w + 2. * (1. - w);
#else
                double c1 = 4. * w + 2. * (1. - w), c2 = 6. * w + (1. - w);
#endif
                for (int a = 0; a < 3; a++)
                    c->u[a][ID(c, i, j, 0)] = c1 * h1[a][ID(c, i, j, 1)]
                        - c2 * h2[a][ID(c, i, j, 2)]
                        + (4. * w) * h3[a][ID(c, i, j, 3)]
                        - w * h4[a][ID(c, i, j, 4)];
                continue;
            }
#if GFO_PLANT == 7
            double num = kind == 0 ? dz : 2. * dz;
#else
            double num = kind == 0 ? 2. * dz : dz;
#endif
            if (!(c->flags & GF_FLAG_1D_ABC_ELAST_ONLYTOP)) {
                double cl = (num - c->vs_long_1st * dt) / (num + c->vs_long_1st * dt);
                double ct = (num - c->vs_tran_1st * dt) / (num + c->vs_tran_1st * dt);
                for (int a = 0; a < 3; a++) {
                    double cc = a == 2 ? cl : ct;
                    if (kind < 2)
                        c->us[a][ID(c, i, j, 0)] = c->u[a][ID(c, i, j, 1)] + cc * (c->u[a][ID(c, i, j, 0)] - c->us[a][ID(c, i, j, 1)]);
                    else
                        c->u[a][ID(c, i, j, 0)] = c->t1[a][ID(c, i, j, 1)] + cc * (c->t1[a][ID(c, i, j, 0)] - c->u[a][ID(c, i, j, 1)]);
                }
            }
            double cl = (num - c->vs_long_last * dt) / (num + c->vs_long_last * dt);
            double ct = (num - c->vs_tran_last * dt) / (num + c->vs_tran_last * dt);
            for (int a = 0; a < 3; a++) {
                double cc = a == 2 ? cl : ct;
                if (kind < 2)
                    c->us[a][ID(c, i, j, nz - 1)] = c->u[a][ID(c, i, j, nz - 2)] + cc * (c->u[a][ID(c, i, j, nz - 1)] - c->us[a][ID(c, i, j, nz - 2)]);
                else
                    c->u[a][ID(c, i, j, nz - 1)] = c->t1[a][ID(c, i, j, nz - 2)] + cc * (c->t1[a][ID(c, i, j, nz - 1)] - c->u[a][ID(c, i, j, nz - 2)]);
            }
        }
}

/* update_Du_Gaussian{_half,_full,} (elastic_dynamic_solver.cpp:3060-3284):
 * displacement sources on the plane k = source_z; kind as above; the
 * store is written during the stages, the state after the step. */
static void elastic_source(gfo_ctx* c, int kind) {
    double tsrc = kind == 0 ? c->time - 0.5 * c->dt : c->time;
    double sg = c->sigma_gauss, A = c->amplitude_gauss;
    int a = c->source_comp == 'x' ? 0 : (c->source_comp == 'y' ? 1 : (c->source_comp == 'z' ? 2 : -1));
    if (a < 0 || c->source_z < 0 || c->source_z >= c->nz) return;
    double vsp = a == 2 ? c->vs_long_1st : c->vs_tran_1st;
    double** dst = kind < 2 ? c->us : c->u;
    for (int64_t i = 0; i < c->nx; i++)
        for (int64_t j = 0; j < c->ny; j++) {
            int64_t id = ID(c, i, j, c->source_z);
            if (!c->cell[id]) continue;
            if (c->source_type == 1 && c->time < 12.5 * sg) {
                dst[a][id] = A * 1.6487210 * sg * vsp * exp(-1. * pow((tsrc - 5.0 * sg), 2.) / (2. * sg * sg));
            } else if (c->source_type == 2) {
                if (c->num_cycles == 0 || tsrc <= (double)c->num_cycles / sg)
                    dst[a][id] = A * vsp / 2. / PI / sg * (cos(2. * PI * sg * tsrc) - 1);
                else
                    dst[a][id] = 0.;
            }
        }
}

/* update_v_RKs / update_v and update_u_RKs / update_u
 * (elastic_dynamic_solver.cpp:2849-3057) */
void gfo_elastic_update(gfo_ctx* c, int stage) {
    int64_t n = c->n;
    for (int a = 0; a < 3; a++) {
        double *v = c->v[a], *vs = c->vs[a];
        if (stage == 1 || stage == 2) { const double* k = c->kv[stage - 1][a]; for (int64_t i = 0; i < n; i++) vs[i] = v[i] + k[i] * 0.5; }
        else if (stage == 3) { const double* k = c->kv[2][a]; for (int64_t i = 0; i < n; i++) vs[i] = v[i] + k[i]; }
        else {
            for (int64_t i = 0; i < n; i++) {
                v[i] = v[i] + c->kv[0][a][i] / 6. + c->kv[1][a][i] / 3. + c->kv[2][a][i] / 3. + c->kv[3][a][i] / 6.;
                vs[i] = v[i];
            }
        }
    }
    int gauss = (c->flags & GF_FLAG_GAUSSIAN_PULSE) != 0;
    if (stage < 4) {
        for (int a = 0; a < 3; a++) {
            double *u = c->u[a], *us = c->us[a];
            const double* k = c->ku[stage - 1][a];
            if (stage < 3) for (int64_t i = 0; i < n; i++) us[i] = u[i] + k[i] * 0.5;
            else for (int64_t i = 0; i < n; i++) us[i] = u[i] + k[i];
        }
        elastic_boundary(c, stage < 3 ? 0 : 1);
        if (gauss) elastic_source(c, stage < 3 ? 0 : 1);
    } else {
        elastic_shift_history(c);
        for (int a = 0; a < 3; a++)
            for (int64_t i = 0; i < n; i++)
                c->u[a][i] = c->u[a][i] + c->ku[0][a][i] / 6. + c->ku[1][a][i] / 3. + c->ku[2][a][i] / 3. + c->ku[3][a][i] / 6.;
        elastic_boundary(c, 2);
        if (gauss) elastic_source(c, 2);
        for (int a = 0; a < 3; a++) memcpy(c->us[a], c->u[a], sizeof(double) * (size_t)n);
    }
    elastic_gradients(c);
    elastic_strain(c);
}

/* -------------------------------------------------- electrodynamics */
static const double EPS0 = 8.854187817e-12;    /* EMdynamic_system.h:13 */
static const double C_LIGHT = 2.99792458e8;    /* EMdynamic_system.h:14 */

/* staggered indices: E on edges, H on faces (EMdynamic_initialize.cpp:68-95) */
static int64_t EXI(const gfo_ctx* c, int64_t i, int64_t j, int64_t k) { return (i * (c->ny + 1) + j) * (c->nz + 1) + k; }
static int64_t EYI(const gfo_ctx* c, int64_t i, int64_t j, int64_t k) { return (i * c->ny + j) * (c->nz + 1) + k; }
static int64_t EZI(const gfo_ctx* c, int64_t i, int64_t j, int64_t k) { return (i * (c->ny + 1) + j) * c->nz + k; }
static int64_t HXI(const gfo_ctx* c, int64_t i, int64_t j, int64_t k) { return (i * c->ny + j) * c->nz + k; }
static int64_t HYI(const gfo_ctx* c, int64_t i, int64_t j, int64_t k) { return (i * (c->ny + 1) + j) * c->nz + k; }
static int64_t HZI(const gfo_ctx* c, int64_t i, int64_t j, int64_t k) { return (i * c->ny + j) * (c->nz + 1) + k; }

static int pec_cell(const gfo_ctx* c, int64_t id) { const gf_material* m = matof(c, id); return m && m->if_PEC; }

/* EMdynamic_initialize.cpp:3-156 and initialize_PEC_edges (:158-366) */
static int em_setup(gfo_ctx* c, const gf_desc* d) {
    c->desc = *d;
    c->desc.materials = NULL; c->desc.material_cell = NULL;
    c->desc.prescribe_Hext = NULL; c->desc.prescribe_Eext = NULL;
    c->if_em = (c->flags & GF_FLAG_EMDYNAMIC) != 0;
    if (!c->if_em) return 0;
    int64_t nx = c->nx, ny = c->ny, nz = c->nz, n = c->n;
    if ((c->flags & GF_FLAG_PLANE_EM) && (d->planeEM_source_z < 0 || d->planeEM_source_z >= nz)) return 1;
    if (c->flags & GF_FLAG_JF_INPUT) {
        int64_t dims[3] = {nx, ny, nz};
        for (int a = 0; a < 3; a++)
            if (d->Jfin_lo[a] < 0 || d->Jfin_hi[a] < d->Jfin_lo[a] || d->Jfin_hi[a] >= dims[a]) return 1;
    }
    c->periodic_all = c->per[0] && c->per[1] && c->per[2];   /* global.cpp:136-139 */
    c->nE[0] = nx * (ny + 1) * (nz + 1); c->nE[1] = (nx + 1) * ny * (nz + 1); c->nE[2] = (nx + 1) * (ny + 1) * nz;
    c->nH[0] = (nx + 1) * ny * nz; c->nH[1] = nx * (ny + 1) * nz; c->nH[2] = nx * ny * (nz + 1);
    for (int a = 0; a < 3; a++) {
        c->Em[a] = dalloc(c->nE[a]); c->Ems[a] = dalloc(c->nE[a]); c->Et1[a] = dalloc(c->nE[a]);
        c->Hm[a] = dalloc(c->nH[a]); c->Hms[a] = dalloc(c->nH[a]);
        for (int s = 0; s < 4; s++) { c->kE[s][a] = dalloc(c->nE[a]); c->kH[s][a] = dalloc(c->nH[a]); }
        c->Ecell[a] = dalloc(n); c->Hcell[a] = dalloc(n); c->Jf[a] = dalloc(n);
        for (int p = 0; p < 3; p++) {
            c->Jp[p][a] = dalloc(n); c->Jps[p][a] = dalloc(n);
            for (int s = 0; s < 4; s++) c->kJ[s][p][a] = dalloc(n);
        }
        c->pec[a] = (unsigned char*)calloc((size_t)c->nE[a], 1);
    }
    /* Mur coefficients from the permittivity of the first material met on
     * the bottom and top planes */
    double er_bot = 1., er_top = 1.;
    int found = 0;
    for (int64_t i = 0; i < nx && !found; i++)
        for (int64_t j = 0; j < ny && !found; j++) {
            const gf_material* m = matof(c, ID(c, i, j, 0));
            if (m) { er_bot = sqrt(m->r_permittivity); found = 1; }
        }
    found = 0;
    for (int64_t i = 0; i < nx && !found; i++)
        for (int64_t j = 0; j < ny && !found; j++) {
            const gf_material* m = matof(c, ID(c, i, j, nz - 1));
            if (m) { er_top = sqrt(m->r_permittivity); found = 1; }
        }
    double dt = c->dt;
    c->cf_top_half = (2. * c->dz - C_LIGHT / er_top * dt) / (2. * c->dz + C_LIGHT / er_top * dt);
    c->cf_bot_half = (2. * c->dz - C_LIGHT / er_bot * dt) / (2. * c->dz + C_LIGHT / er_bot * dt);
    c->cf_x_half = (2. * c->dx - C_LIGHT * dt) / (2. * c->dx + C_LIGHT * dt);
    c->cf_y_half = (2. * c->dy - C_LIGHT * dt) / (2. * c->dy + C_LIGHT * dt);
    c->cf_top = (c->dz - C_LIGHT / er_top * dt) / (c->dz + C_LIGHT / er_top * dt);
    c->cf_bot = (c->dz - C_LIGHT / er_bot * dt) / (c->dz + C_LIGHT / er_bot * dt);
    c->cf_x = (c->dx - C_LIGHT * dt) / (c->dx + C_LIGHT * dt);
    c->cf_y = (c->dy - C_LIGHT * dt) / (c->dy + C_LIGHT * dt);
    for (int i = 0; i < c->nmat; i++) if (c->mat[i].if_PEC) c->pec_all = 1;
    if (c->pec_all) {
        /* an edge is PEC when one of its four cells is, walls that are not
         * periodic excluded */
        for (int64_t i = 0; i < nx; i++)
            for (int64_t j = 0; j <= ny; j++)
                for (int64_t k = 0; k <= nz; k++) {
                    if ((j == 0 || j == ny) && !c->per[1]) continue;
                    if ((k == 0 || k == nz) && !c->per[2]) continue;
                    int64_t y1 = (j == 0 || j == ny) ? ny - 1 : j - 1, y3 = (j == 0 || j == ny) ? 0 : j;
                    int64_t z1 = (k == 0 || k == nz) ? nz - 1 : k - 1, z2 = (k == 0 || k == nz) ? 0 : k;
                    c->pec[0][EXI(c, i, j, k)] = (unsigned char)(pec_cell(c, ID(c, i, y1, z1)) || pec_cell(c, ID(c, i, y1, z2)) ||
                                                                 pec_cell(c, ID(c, i, y3, z1)) || pec_cell(c, ID(c, i, y3, z2)));
                }
        for (int64_t i = 0; i <= nx; i++)
            for (int64_t j = 0; j < ny; j++)
                for (int64_t k = 0; k <= nz; k++) {
                    if ((i == 0 || i == nx) && !c->per[0]) continue;
                    if ((k == 0 || k == nz) && !c->per[2]) continue;
                    int64_t x1 = (i == 0 || i == nx) ? nx - 1 : i - 1, x3 = (i == 0 || i == nx) ? 0 : i;
                    int64_t z1 = (k == 0 || k == nz) ? nz - 1 : k - 1, z2 = (k == 0 || k == nz) ? 0 : k;
                    c->pec[1][EYI(c, i, j, k)] = (unsigned char)(pec_cell(c, ID(c, x1, j, z1)) || pec_cell(c, ID(c, x1, j, z2)) ||
                                                                 pec_cell(c, ID(c, x3, j, z1)) || pec_cell(c, ID(c, x3, j, z2)));
                }
        for (int64_t i = 0; i <= nx; i++)
            for (int64_t j = 0; j <= ny; j++)
                for (int64_t k = 0; k < nz; k++) {
                    if ((i == 0 || i == nx) && !c->per[0]) continue;
                    if ((j == 0 || j == ny) && !c->per[1]) continue;
                    int64_t x1 = (i == 0 || i == nx) ? nx - 1 : i - 1, x3 = (i == 0 || i == nx) ? 0 : i;
                    int64_t y1 = (j == 0 || j == ny) ? ny - 1 : j - 1, y2 = (j == 0 || j == ny) ? 0 : j;
                    c->pec[2][EZI(c, i, j, k)] = (unsigned char)(pec_cell(c, ID(c, x1, y1, k)) || pec_cell(c, ID(c, x1, y2, k)) ||
                                                                 pec_cell(c, ID(c, x3, y1, k)) || pec_cell(c, ID(c, x3, y2, k)));
                }
    }
    return 0;
}

/* update_DE_cell / update_DH_cell (EMdynamic_solver.cpp:3780-3826); which:
 * 1 = E, 2 = H, 3 = both */
static void em_cells(gfo_ctx* c, int which) {
    int one_d = (c->flags & GF_FLAG_MAG_1D) && (c->flags & GF_FLAG_EM_FROM_M);
    #pragma omp parallel for collapse(3) schedule(static) if (c->n >= GFO_OMP_MIN)
    for (int64_t i = 0; i < c->nx; i++)
        for (int64_t j = 0; j < c->ny; j++)
            for (int64_t k = 0; k < c->nz; k++) {
                int64_t id = ID(c, i, j, k);
                if (which & 1) {
                    c->Ecell[0][id] = (c->Ems[0][EXI(c, i, j, k)] + c->Ems[0][EXI(c, i, j + 1, k)] + c->Ems[0][EXI(c, i, j, k + 1)]
                                       + c->Ems[0][EXI(c, i, j + 1, k + 1)]) / 4.;
                    c->Ecell[1][id] = (c->Ems[1][EYI(c, i, j, k)] + c->Ems[1][EYI(c, i + 1, j, k)] + c->Ems[1][EYI(c, i, j, k + 1)]
                                       + c->Ems[1][EYI(c, i + 1, j, k + 1)]) / 4.;
                    c->Ecell[2][id] = (c->Ems[2][EZI(c, i, j, k)] + c->Ems[2][EZI(c, i + 1, j, k)] + c->Ems[2][EZI(c, i, j + 1, k)]
                                       + c->Ems[2][EZI(c, i + 1, j + 1, k)]) / 4.;
                }
                if (which & 2) {
                    c->Hcell[0][id] = (c->Hms[0][HXI(c, i, j, k)] + c->Hms[0][HXI(c, i + 1, j, k)]) / 2.;
                    c->Hcell[1][id] = (c->Hms[1][HYI(c, i, j, k)] + c->Hms[1][HYI(c, i, j + 1, k)]) / 2.;
                    if (one_d) {
                        const gf_material* m = matof(c, id);
                        c->Hcell[2][id] = !m ? 0. : -1. * m->Ms * c->ms[2][id] - 1. * m->Ms_AFM1 * c->ms1[2][id]
                                                    - 1. * m->Ms_AFM2 * c->ms2[2][id] - c->H[2][id];
                    } else {
                        c->Hcell[2][id] = (c->Hms[2][HZI(c, i, j, k)] + c->Hms[2][HZI(c, i, j, k + 1)]) / 2.;
                    }
                }
            }
}

/* update_Jf_input{,_half,_full} (EMdynamic_solver.cpp:3885-4315) */
static void em_jf(gfo_ctx* c, int stage) {
    const gf_desc* d = &c->desc;
    double t = stage == 1 ? c->time - c->dt : (stage == 4 ? c->time : c->time - 0.5 * c->dt);
    int a = d->Jf_input_component == 'x' ? 0 : (d->Jf_input_component == 'y' ? 1 : (d->Jf_input_component == 'z' ? 2 : -1));
    int64_t Jnz = (int64_t)d->Jfin_hi[2] - d->Jfin_lo[2] + 1;
    double amp = d->Jf_input_amp, f = d->Jf_input_freq, sg = d->Jf_input_sigma;
    for (int64_t i = d->Jfin_lo[0]; i <= d->Jfin_hi[0]; i++)
        for (int64_t j = d->Jfin_lo[1]; j <= d->Jfin_hi[1]; j++)
            for (int64_t k = d->Jfin_lo[2]; k <= d->Jfin_hi[2]; k++) {
                double val, tv, sv;
                int type = d->Jf_input_type;
                if (type < 3) {
                    tv = sin(2. * PI * f * t);
                    sv = PI / ((double)Jnz * c->dz) * ((double)(k - d->Jfin_lo[2]) + 0.5) * c->dz;
                    if (type == 0) val = amp * tv;
                    else if (type == 1) val = amp * tv * sin(sv);
                    else val = amp * tv * cos(sv);
                } else if (type == 3) {
                    tv = t <= (double)d->num_Jf_cycles / f ? sin(2. * PI * f * t) : 0.;
                    val = amp * tv;
                } else if (type == 4) {
                    tv = t - 5. * f;
                    val = amp * exp(-1. * pow(tv, 2.) / (2. * pow(f, 2.)));
                } else if (type == 5) {
                    tv = t - 5. * f;
                    val = amp * sqrt(exp(1.)) * tv / f * exp(-1. * pow(tv, 2.) / (2. * pow(f, 2.)));
                } else if (type == 6) {
                    tv = t - 5. * sg;
                    val = amp * exp(-1. * pow(tv, 2.) / (2. * pow(sg, 2.))) * sin(2. * PI * f * t);
                } else {
                    continue;
                }
                if (a >= 0) c->Jf[a][ID(c, i, j, k)] = val;
            }
}

/* The four-cell means of get_dE_RK* (EMdynamic_solver.cpp:392-499) and the
 * Ampere increment, with the spin-pumping current on x and y edges. */
static double edge_dE(const gfo_ctx* c, int comp, const int64_t cell[4], double curl, double e_store, int stage) {
    double Jf = 0., Jp = 0., Ji = 0., dP = 0., er = 0., cond = 0., Pc = 0.;
    for (int q = 0; q < 4; q++) {
        int64_t id = cell[q];
        const gf_material* mat = matof(c, id);
        if (!mat) {
            er = er + 1.;
        } else {
            if (c->flags & GF_FLAG_EM_FROM_P) dP = dP + c->kp[stage - 1][comp][id];
            if (mat->if_FE) Pc = Pc + 1.;
            er = er + mat->r_permittivity;
            cond = cond + mat->conductivity;
        }
        Jf = Jf + c->Jf[comp][id];
        Jp = Jp + c->Jps[0][comp][id] + c->Jps[1][comp][id] + c->Jps[2][comp][id];
        if (c->spin_on && comp < 2) Ji = Ji + c->Jis[comp][id];
    }
    if (Pc < 0.5) Pc = 1.;
    Jf = Jf / 4.;
    Jp = Jp / 4.;
    Ji = Ji / 4.;
    dP = dP / Pc; er = er / 4.; cond = cond / 4.;
    return c->dt / EPS0 / er * (curl - Jf - Jp - Ji - cond * e_store - dP / c->dt);
}

/* dM of one cell next to an H face (EMdynamic_solver.cpp:2290-2340) */
static double face_dM(const gfo_ctx* c, int comp, int64_t id, int stage, double* count) {
    const gf_material* mat = matof(c, id);
    if (!mat) return 0.;
    double d = 0.;
    if (c->flags & GF_FLAG_EM_FROM_M)
        d = c->k[stage - 1][comp][id] * mat->Ms + c->k1[stage - 1][comp][id] * mat->Ms_AFM1
          + c->k2[stage - 1][comp][id] * mat->Ms_AFM2;
    if (mat->if_FM || mat->if_AFM) *count = *count + 1.;
    return d;
}

/* get_dE_RK*, get_dH_RK*, get_dJp_RK* with the Jf source
 * (EMdynamic_solver.cpp:337-3476, 4318-4451) */
void gfo_em_get(gfo_ctx* c, int stage) {
    int64_t nx = c->nx, ny = c->ny, nz = c->nz;
    int px = c->per[0], py = c->per[1], pz = c->per[2];
    double dt = c->dt;
    if (c->flags & GF_FLAG_JF_INPUT) em_jf(c, stage);
    #pragma omp parallel for schedule(static) if (c->n >= GFO_OMP_MIN)
    for (int64_t id = 0; id < c->n; id++) {
        const gf_material* mat = matof(c, id);
        if (!mat) continue;
        double cn[3] = {mat->comp_n1, mat->comp_n2, mat->comp_n3};
        double wp[3] = {mat->omega_plasma_n1, mat->omega_plasma_n2, mat->omega_plasma_n3};
        double re[3] = {1. / mat->tao_e_n1, 1. / mat->tao_e_n2, 1. / mat->tao_e_n3};
        for (int p = 0; p < 3; p++)
            for (int a = 0; a < 3; a++)
                c->kJ[stage - 1][p][a][id] = (EPS0 * wp[p] * wp[p] * c->Ecell[a][id] * cn[p] - re[p] * c->Jps[p][a][id]) * dt;
    }
    double** Hs = c->Hms;
    double** Es = c->Ems;
    /* Ex edges */
    #pragma omp parallel for collapse(3) schedule(static) if (c->n >= GFO_OMP_MIN)
    for (int64_t i = 0; i < nx; i++)
        for (int64_t j = 0; j <= ny; j++)
            for (int64_t k = 0; k <= nz; k++) {
                if ((j == 0 || j == ny) && !py) continue;
                if ((k == 0 || k == nz) && !pz) continue;
                int64_t y1 = (j == 0 || j == ny) ? ny - 1 : j - 1, y3 = (j == 0 || j == ny) ? 0 : j;
                int64_t z1 = (k == 0 || k == nz) ? nz - 1 : k - 1, z2 = (k == 0 || k == nz) ? 0 : k;
                double dHzdy = (Hs[2][HZI(c, i, y3, z2)] - Hs[2][HZI(c, i, y1, z2)]) / c->dy;
                double dHydz = (Hs[1][HYI(c, i, y3, z2)] - Hs[1][HYI(c, i, y3, z1)]) / c->dz;
                int64_t cell[4] = {ID(c, i, y1, z1), ID(c, i, y1, z2), ID(c, i, y3, z1), ID(c, i, y3, z2)};
                int64_t e = EXI(c, i, j, k);
                c->kE[stage - 1][0][e] = edge_dE(c, 0, cell, dHzdy - dHydz, Es[0][e], stage);
            }
    /* Ey edges */
    #pragma omp parallel for collapse(3) schedule(static) if (c->n >= GFO_OMP_MIN)
    for (int64_t i = 0; i <= nx; i++)
        for (int64_t j = 0; j < ny; j++)
            for (int64_t k = 0; k <= nz; k++) {
                if ((i == 0 || i == nx) && !px) continue;
                if ((k == 0 || k == nz) && !pz) continue;
                int64_t x1 = (i == 0 || i == nx) ? nx - 1 : i - 1, x3 = (i == 0 || i == nx) ? 0 : i;
                int64_t z1 = (k == 0 || k == nz) ? nz - 1 : k - 1, z2 = (k == 0 || k == nz) ? 0 : k;
                double dHxdz = (Hs[0][HXI(c, x3, j, z2)] - Hs[0][HXI(c, x3, j, z1)]) / c->dz;
                double dHzdx = (Hs[2][HZI(c, x3, j, z2)] - Hs[2][HZI(c, x1, j, z2)]) / c->dx;
                int64_t cell[4] = {ID(c, x1, j, z1), ID(c, x1, j, z2), ID(c, x3, j, z1), ID(c, x3, j, z2)};
                int64_t e = EYI(c, i, j, k);
                c->kE[stage - 1][1][e] = edge_dE(c, 1, cell, dHxdz - dHzdx, Es[1][e], stage);
            }
    /* Ez edges */
    #pragma omp parallel for collapse(3) schedule(static) if (c->n >= GFO_OMP_MIN)
    for (int64_t i = 0; i <= nx; i++)
        for (int64_t j = 0; j <= ny; j++)
            for (int64_t k = 0; k < nz; k++) {
                if ((i == 0 || i == nx) && !px) continue;
                if ((j == 0 || j == ny) && !py) continue;
                int64_t x1 = (i == 0 || i == nx) ? nx - 1 : i - 1, x3 = (i == 0 || i == nx) ? 0 : i;
                int64_t y1 = (j == 0 || j == ny) ? ny - 1 : j - 1, y2 = (j == 0 || j == ny) ? 0 : j;
                double dHydx = (Hs[1][HYI(c, x3, y2, k)] - Hs[1][HYI(c, x1, y2, k)]) / c->dx;
                double dHxdy = (Hs[0][HXI(c, x3, y2, k)] - Hs[0][HXI(c, x3, y1, k)]) / c->dy;
                int64_t cell[4] = {ID(c, x1, y1, k), ID(c, x1, y2, k), ID(c, x3, y1, k), ID(c, x3, y2, k)};
                int64_t e = EZI(c, i, j, k);
                c->kE[stage - 1][2][e] = edge_dE(c, 2, cell, dHydx - dHxdy, Es[2][e], stage);
            }
    /* H faces (EMdynamic_solver.cpp:2239-2547) */
    #pragma omp parallel for collapse(3) schedule(static) if (c->n >= GFO_OMP_MIN)
    for (int64_t i = 0; i <= nx; i++)
        for (int64_t j = 0; j < ny; j++)
            for (int64_t k = 0; k < nz; k++) {
                int64_t x1 = i == 0 ? (px ? nx - 1 : 0) : i - 1, x2 = i == nx ? (px ? 0 : nx - 1) : i;
                double dEzdy = (Es[2][EZI(c, i, j + 1, k)] - Es[2][EZI(c, i, j, k)]) / c->dy;
                double dEydz = (Es[1][EYI(c, i, j, k + 1)] - Es[1][EYI(c, i, j, k)]) / c->dz;
                double cnt = 0.;
                double d1 = face_dM(c, 0, ID(c, x1, j, k), stage, &cnt), d2 = face_dM(c, 0, ID(c, x2, j, k), stage, &cnt);
                if (cnt < 0.5) cnt = 1.;
                c->kH[stage - 1][0][HXI(c, i, j, k)] = -dt / MU0 * (dEzdy - dEydz) - (d1 + d2) / cnt;
            }
    #pragma omp parallel for collapse(3) schedule(static) if (c->n >= GFO_OMP_MIN)
    for (int64_t i = 0; i < nx; i++)
        for (int64_t j = 0; j <= ny; j++)
            for (int64_t k = 0; k < nz; k++) {
                int64_t y1 = j == 0 ? (py ? ny - 1 : 0) : j - 1, y2 = j == ny ? (py ? 0 : ny - 1) : j;
                double dExdz = (Es[0][EXI(c, i, j, k + 1)] - Es[0][EXI(c, i, j, k)]) / c->dz;
                double dEzdx = (Es[2][EZI(c, i + 1, j, k)] - Es[2][EZI(c, i, j, k)]) / c->dx;
                double cnt = 0.;
                double d1 = face_dM(c, 1, ID(c, i, y1, k), stage, &cnt), d2 = face_dM(c, 1, ID(c, i, y2, k), stage, &cnt);
                if (cnt < 0.5) cnt = 1.;
                c->kH[stage - 1][1][HYI(c, i, j, k)] = -dt / MU0 * (dExdz - dEzdx) - (d1 + d2) / cnt;
            }
    #pragma omp parallel for collapse(3) schedule(static) if (c->n >= GFO_OMP_MIN)
    for (int64_t i = 0; i < nx; i++)
        for (int64_t j = 0; j < ny; j++)
            for (int64_t k = 0; k <= nz; k++) {
                int64_t z1 = k == 0 ? (pz ? nz - 1 : 0) : k - 1, z2 = k == nz ? (pz ? 0 : nz - 1) : k;
                double dEydx = (Es[1][EYI(c, i + 1, j, k)] - Es[1][EYI(c, i, j, k)]) / c->dx;
                double dExdy = (Es[0][EXI(c, i, j + 1, k)] - Es[0][EXI(c, i, j, k)]) / c->dy;
                double cnt = 0.;
                double d1 = face_dM(c, 2, ID(c, i, j, z1), stage, &cnt), d2 = face_dM(c, 2, ID(c, i, j, z2), stage, &cnt);
                if (cnt < 0.5) cnt = 1.;
                c->kH[stage - 1][2][HZI(c, i, j, k)] = -dt / MU0 * (dEydx - dExdy) - (d1 + d2) / cnt;
            }
}

/* update_planeEM{_half,_full,} (EMdynamic_solver.cpp:3829-3883); kind as
 * for the boundary */
static void em_plane(gfo_ctx* c, int kind) {
    const gf_desc* d = &c->desc;
    double t = kind == 0 ? c->time - 0.5 * c->dt : c->time;
    double** dst = kind < 2 ? c->Ems : c->Em;
    int64_t z = d->planeEM_source_z;
    for (int64_t i = 0; i < c->nx; i++)
        for (int64_t j = 0; j < c->ny; j++) {
            dst[0][EXI(c, i, j, z)] = d->planeEM_E[0] * sin(2. * PI * d->planeEM_E_freq[0] * t);
            dst[1][EYI(c, i, j, z)] = d->planeEM_E[1] * sin(2. * PI * d->planeEM_E_freq[1] * t);
            dst[2][EZI(c, i, j, z)] = d->planeEM_E[2] * sin(2. * PI * d->planeEM_E_freq[2] * t);
        }
}

/* first-order Mur walls (EMdynamic_AbsorbingBoundary.cpp:58-874): kind 0 =
 * stages 1-2 (coefficients with 2d), 1 = stage 3, 2 = after the step
 * (from the history t1); x walls, then y walls, then z walls */
static void em_mur(double* dst, const double* b, const double* s, const double* t1, int kind, double cf,
                   int64_t wall, int64_t inner) {
#if GFO_PLANT == 6
    s = b;
#endif
    if (kind < 2) dst[wall] = b[inner] + cf * (b[wall] - s[inner]);
    else dst[wall] = t1[inner] + cf * (t1[wall] - b[inner]);
}

static void em_boundary(gfo_ctx* c, int kind) {
    int64_t nx = c->nx, ny = c->ny, nz = c->nz;
#if GFO_PLANT == 5
    kind = kind == 0 ? 1 : (kind == 1 ? 0 : kind);
#endif
    double cx = kind == 0 ? c->cf_x_half : c->cf_x, cy = kind == 0 ? c->cf_y_half : c->cf_y;
    double ct = kind == 0 ? c->cf_top_half : c->cf_top, cb = kind == 0 ? c->cf_bot_half : c->cf_bot;
    int pec_xy = (c->flags & GF_FLAG_PEC_XY) != 0, pec_z = (c->flags & GF_FLAG_PEC_Z) != 0;
    int onlytop = (c->flags & GF_FLAG_1D_ABC) && (c->flags & GF_FLAG_1D_ABC_EM_ONLYTOP);
    double** D = kind < 2 ? c->Ems : c->Em;
    if (!c->per[0]) {
        for (int64_t j = 0; j < ny; j++)
            for (int64_t k = 0; k <= nz; k++) {
                int64_t w0 = EYI(c, 0, j, k), w1 = EYI(c, 1, j, k), wn = EYI(c, nx, j, k), wm = EYI(c, nx - 1, j, k);
                if (!pec_xy) {
                    em_mur(D[1], c->Em[1], c->Ems[1], c->Et1[1], kind, cx, w0, w1);
                    em_mur(D[1], c->Em[1], c->Ems[1], c->Et1[1], kind, cx, wn, wm);
                } else { D[1][w0] = 0.; D[1][wn] = 0.; }
            }
        for (int64_t j = 0; j <= ny; j++)
            for (int64_t k = 0; k < nz; k++) {
                int64_t w0 = EZI(c, 0, j, k), w1 = EZI(c, 1, j, k), wn = EZI(c, nx, j, k), wm = EZI(c, nx - 1, j, k);
                if (!pec_xy) {
                    em_mur(D[2], c->Em[2], c->Ems[2], c->Et1[2], kind, cx, w0, w1);
                    em_mur(D[2], c->Em[2], c->Ems[2], c->Et1[2], kind, cx, wn, wm);
                } else { D[2][w0] = 0.; D[2][wn] = 0.; }
            }
    }
    if (!c->per[1]) {
        for (int64_t i = 0; i < nx; i++)
            for (int64_t k = 0; k <= nz; k++) {
                int64_t w0 = EXI(c, i, 0, k), w1 = EXI(c, i, 1, k), wn = EXI(c, i, ny, k), wm = EXI(c, i, ny - 1, k);
                if (!pec_xy) {
                    em_mur(D[0], c->Em[0], c->Ems[0], c->Et1[0], kind, cy, w0, w1);
                    em_mur(D[0], c->Em[0], c->Ems[0], c->Et1[0], kind, cy, wn, wm);
                } else { D[0][w0] = 0.; D[0][wn] = 0.; }
            }
        for (int64_t i = 0; i <= nx; i++)
            for (int64_t k = 0; k < nz; k++) {
                int64_t w0 = EZI(c, i, 0, k), w1 = EZI(c, i, 1, k), wn = EZI(c, i, ny, k), wm = EZI(c, i, ny - 1, k);
                if (!pec_xy) {
                    em_mur(D[2], c->Em[2], c->Ems[2], c->Et1[2], kind, cy, w0, w1);
                    em_mur(D[2], c->Em[2], c->Ems[2], c->Et1[2], kind, cy, wn, wm);
                } else { D[2][w0] = 0.; D[2][wn] = 0.; }
            }
    }
    if (!c->per[2]) {
        for (int comp = 0; comp < 2; comp++) {
            int64_t ni = comp == 0 ? nx : nx + 1, nj = comp == 0 ? ny + 1 : ny;
            for (int64_t i = 0; i < ni; i++)
                for (int64_t j = 0; j < nj; j++) {
                    int64_t w0, w1, wn, wm;
                    if (comp == 0) { w0 = EXI(c, i, j, 0); w1 = EXI(c, i, j, 1); wn = EXI(c, i, j, nz); wm = EXI(c, i, j, nz - 1); }
                    else { w0 = EYI(c, i, j, 0); w1 = EYI(c, i, j, 1); wn = EYI(c, i, j, nz); wm = EYI(c, i, j, nz - 1); }
                    if (!pec_z) {
                        em_mur(D[comp], c->Em[comp], c->Ems[comp], c->Et1[comp], kind, ct, wn, wm);
                        if (onlytop) D[comp][w0] = 0.;
                        else em_mur(D[comp], c->Em[comp], c->Ems[comp], c->Et1[comp], kind, cb, w0, w1);
                    } else { D[comp][w0] = 0.; D[comp][wn] = 0.; }
                }
        }
    }
}

/* update_DE_PEC (EMdynamic_solver.cpp:4526-4556) */
static void em_pec(gfo_ctx* c) {
    for (int a = 0; a < 3; a++)
        for (int64_t e = 0; e < c->nE[a]; e++)
            if (c->pec[a][e]) { c->Em[a][e] = 0.; c->Ems[a][e] = 0.; }
}

/* update_DE_RK* / update_DE, update_DH_RK* / update_DH, update_Jp_RK* /
 * update_Jp (EMdynamic_solver.cpp:3477-3777, 4453-4524) */
void gfo_em_update(gfo_ctx* c, int stage) {
    int plane = (c->flags & GF_FLAG_PLANE_EM) != 0;
    if (stage < 4) {
        double w = stage == 3 ? 1. : 0.5;
        int kind = stage < 3 ? 0 : 1;
        for (int a = 0; a < 3; a++)
            #pragma omp parallel for schedule(static) if (c->n >= GFO_OMP_MIN)
            for (int64_t e = 0; e < c->nE[a]; e++)
                c->Ems[a][e] = stage == 3 ? c->Em[a][e] + c->kE[2][a][e] : c->Em[a][e] + c->kE[stage - 1][a][e] * 0.5;
        if (plane) em_plane(c, kind);
        if (!c->periodic_all) em_boundary(c, kind);
        if (c->pec_all) em_pec(c);
        em_cells(c, 1);
        for (int p = 0; p < 3; p++)
            for (int a = 0; a < 3; a++)
                #pragma omp parallel for schedule(static) if (c->n >= GFO_OMP_MIN)
                for (int64_t id = 0; id < c->n; id++)
                    c->Jps[p][a][id] = stage == 3 ? c->Jp[p][a][id] + c->kJ[2][p][a][id]
                                                  : c->Jp[p][a][id] + c->kJ[stage - 1][p][a][id] * 0.5;
        for (int a = 0; a < 3; a++)
            #pragma omp parallel for schedule(static) if (c->n >= GFO_OMP_MIN)
            for (int64_t f = 0; f < c->nH[a]; f++)
                c->Hms[a][f] = stage == 3 ? c->Hm[a][f] + c->kH[2][a][f] : c->Hm[a][f] + c->kH[stage - 1][a][f] * 0.5;
        (void)w;
        em_cells(c, 2);
        return;
    }
    if (!c->periodic_all)
        for (int a = 0; a < 3; a++) memcpy(c->Et1[a], c->Em[a], sizeof(double) * (size_t)c->nE[a]);
    for (int a = 0; a < 3; a++)
        #pragma omp parallel for schedule(static) if (c->n >= GFO_OMP_MIN)
        for (int64_t e = 0; e < c->nE[a]; e++)
            c->Em[a][e] = c->Em[a][e] + c->kE[0][a][e] / 6. + c->kE[1][a][e] / 3. + c->kE[2][a][e] / 3. + c->kE[3][a][e] / 6.;
    if (plane) em_plane(c, 2);
    if (!c->periodic_all) em_boundary(c, 2);
    if (c->pec_all) em_pec(c);
    for (int a = 0; a < 3; a++) memcpy(c->Ems[a], c->Em[a], sizeof(double) * (size_t)c->nE[a]);
    em_cells(c, 1);
    for (int p = 0; p < 3; p++)
        for (int a = 0; a < 3; a++)
            #pragma omp parallel for schedule(static) if (c->n >= GFO_OMP_MIN)
            for (int64_t id = 0; id < c->n; id++) {
                c->Jp[p][a][id] = c->Jp[p][a][id] + c->kJ[0][p][a][id] / 6. + c->kJ[1][p][a][id] / 3.
                                + c->kJ[2][p][a][id] / 3. + c->kJ[3][p][a][id] / 6.;
                c->Jps[p][a][id] = c->Jp[p][a][id];
            }
    for (int a = 0; a < 3; a++)
        #pragma omp parallel for schedule(static) if (c->n >= GFO_OMP_MIN)
        for (int64_t f = 0; f < c->nH[a]; f++) {
            c->Hm[a][f] = c->Hm[a][f] + c->kH[0][a][f] / 6. + c->kH[1][a][f] / 3. + c->kH[2][a][f] / 3. + c->kH[3][a][f] / 6.;
            c->Hms[a][f] = c->Hm[a][f];
        }
    em_cells(c, 2);
}

/* ------------------------------------------------------ statics */
/* magnetostatic_solver.cpp:222-249 */
static void demag_1d(gfo_ctx* c) {
    for (int64_t i = 0; i < c->n; i++) {
        const gf_material* mat = matof(c, i);
        if (!mat) continue;
        if (mat->if_FM) { c->H[0][i] = 0.; c->H[1][i] = 0.; c->H[2][i] = -mat->Ms * c->ms[2][i]; }
        else if (mat->if_AFM) {
            c->H[0][i] = 0.; c->H[1][i] = 0.;
            c->H[2][i] = -mat->Ms_AFM1 * c->ms1[2][i] - mat->Ms_AFM2 * c->ms2[2][i];
        }
    }
}

/* magnetostatic_solver.cpp:114-219 */
static void conv_3d(gfo_ctx* c, const double* M, double* const* out);

static void demag_3d(gfo_ctx* c) {
    int64_t n = c->n;
    if (c->flags & GF_FLAG_DEMAG_FACTOR) {
        for (int64_t i = 0; i < n; i++) {
            const gf_material* mat = matof(c, i);
            if (!mat) continue;
            if (mat->if_FM) {
                for (int a = 0; a < 3; a++) c->H[a][i] = mat->Ms * c->ms[a][i] * c->demag_fac[a];
            } else if (mat->if_AFM) {
                for (int a = 0; a < 3; a++)
                    c->H[a][i] = (mat->Ms_AFM1 * c->ms1[a][i] + mat->Ms_AFM2 * c->ms2[a][i]) * c->demag_fac[a];
            }
        }
        return;
    }
    double* M = dalloc(3 * n);
    #pragma omp parallel for schedule(static) if (c->n >= GFO_OMP_MIN)
    for (int64_t i = 0; i < n; i++) {
        const gf_material* mat = matof(c, i);
        if (!mat) continue;
        if (mat->if_FM) {
            for (int a = 0; a < 3; a++) M[a * n + i] = mat->Ms * c->ms[a][i];
        } else if (mat->if_AFM) {
            for (int a = 0; a < 3; a++) M[a * n + i] = mat->Ms_AFM1 * c->ms1[a][i] + mat->Ms_AFM2 * c->ms2[a][i];
        }
    }
    conv_3d(c, M, c->H);
    free(M);
}

/* magnetostatic_solver.cpp:114-219: H = N (*) M for a magnetisation M of
 * 3n doubles (component a at M + a n) into out. */
static void conv_3d(gfo_ctx* c, const double* M, double* const* out) {
    int64_t n = c->n;
    int64_t nh = c->nx * c->ny * (c->nz / 2 + 1);
    double* Mk = dalloc(3 * 2 * nh);
    double* Hk = dalloc(2 * nh);
    for (int a = 0; a < 3; a++) gfo_rfft3(c->nx, c->ny, c->nz, M + a * n, Mk + a * 2 * nh);
    /* H_k = (M_a K_a + M_b K_b + M_c K_c) / (-dx dy dz), complex products
     * in the order of :163-177; spectra order Axyz,Ayzx,Azxy,Bxyz,Byzx,Bzxy,Byxz,Bzyx,Bxzy */
    static const int comp[3][3] = {{0, 1, 2}, {1, 2, 0}, {2, 0, 1}};
    static const int spec[3][3] = {{0, 3, 8}, {1, 4, 6}, {2, 5, 7}};
    double den = -c->dx * c->dy * c->dz;
    for (int a = 0; a < 3; a++) {
        #pragma omp parallel for schedule(static) if (c->n >= GFO_OMP_MIN)
        for (int64_t q = 0; q < nh; q++) {
            double sr = 0., si = 0.;
            for (int t = 0; t < 3; t++) {
                const double* mk = Mk + comp[a][t] * 2 * nh;
                const double* kk = c->kspec[spec[a][t]];
                double ar = mk[2 * q], ai = mk[2 * q + 1], br = kk[2 * q], bi = kk[2 * q + 1];
                double pr = ar * br - ai * bi, pi_ = ar * bi + ai * br;
                if (t == 0) { sr = pr; si = pi_; } else { sr = sr + pr; si = si + pi_; }
            }
            Hk[2 * q] = sr / den; Hk[2 * q + 1] = si / den;
        }
        gfo_irfft3(c->nx, c->ny, c->nz, Hk, out[a]);
    }
    free(Mk); free(Hk);
}

/* FFT electrostatics (GF_FLAG_ELECTROSTATICS; no counterpart in the
 * reference: get_E_static is declared at ferroelectric_system.h:103 but not
 * defined).  From the reference's set-up (ferroelectric_initialize.cpp:
 * 310-338, global.cpp:270-318): homogeneous permittivity er_homo and
 * continuous wave vectors k_a = t 2 pi / (n_a d_a), t = i for i <= n/2, else
 * i - n.  E(k) = -k (k.P(k)) / (eps0 er |k|^2) from the current P store,
 * E(0) = 0; an off-diagonal product with a component on its Nyquist plane
 * (t = n/2) is zero so that the field is real. */
static double kwave(int64_t i, int64_t n, double h, int* nyq) {
    int64_t t = i <= n / 2 ? i : i - n;
    *nyq = n >= 2 && 2 * i == n;
    return (double)t * (2. * PI / ((double)n * h));
}

void gfo_estatic(gfo_ctx* c) {
    const double EPS0_ = 8.854187817e-12;   /* EMdynamic_system.h:13 */
    int64_t nx = c->nx, ny = c->ny, nz = c->nz, nzh = nz / 2 + 1, nh = nx * ny * nzh;
    double* Pk = (double*)malloc(sizeof(double) * 2 * nh * 3);
    double* Ek = (double*)malloc(sizeof(double) * 2 * nh);
    for (int a = 0; a < 3; a++) gfo_rfft3(nx, ny, nz, c->ps[a], Pk + a * 2 * nh);
    for (int a = 0; a < 3; a++) {
        for (int64_t i = 0; i < nx; i++)
            for (int64_t j = 0; j < ny; j++)
                for (int64_t k = 0; k < nzh; k++) {
                    int nyq[3];
                    double kv[3];
                    kv[0] = kwave(i, nx, c->dx, &nyq[0]);
                    kv[1] = kwave(j, ny, c->dy, &nyq[1]);
                    kv[2] = kwave(k, nz, c->dz, &nyq[2]);
                    double k2 = kv[0] * kv[0] + kv[1] * kv[1] + kv[2] * kv[2];
                    int64_t q = (i * ny + j) * nzh + k;
                    double er = 0., ei = 0.;
                    if (k2 > 0.)
                        for (int b = 0; b < 3; b++) {
                            if (a != b && (nyq[a] || nyq[b])) continue;
                            double g = -kv[a] * kv[b] / (EPS0_ * c->er_homo * k2);
                            er += g * Pk[b * 2 * nh + 2 * q];
                            ei += g * Pk[b * 2 * nh + 2 * q + 1];
                        }
                    Ek[2 * q] = er;
                    Ek[2 * q + 1] = ei;
                }
        gfo_irfft3(nx, ny, nz, Ek, c->Es[a]);
        for (int64_t q = 0; q < c->n; q++) c->Es[a][q] = c->Es[a][q] / (double)c->n;
    }
    free(Pk); free(Ek);
}

/* Static strain of the eigenstrain of the magnets and ferroelectrics (GF_FLAG_ELASTOSTATIC;
 * get_strain_static is declared at elastic_system.h:190 without a
 * definition).  Periodic body of the homogeneous cubic stiffness
 * (elastic_initialize.cpp:22-40), crystal frame.  With sigma0 = C e0 and
 * the Green tensor G = (C_ijkl k_j k_l)^-1 of elastic_initialize.cpp:
 * 340-389 (k in the crystal frame, |k|^2 of the global grid),
 * u = -i G (sigma0 k) and e_ij = (k_j w_i + k_i w_j) / 2, w = G (sigma0 k);
 * e(0) = e0(0) (zero mean stress).  A Nyquist component is averaged over
 * its two signs. */
static void es_operator(const gfo_ctx* c, const double kg[3], const double sig[6][2], double eps[6][2]) {
    double c11 = c->ch11, c12 = c->ch12, c44 = c->ch44;
    double kq[3];
    glb2crt(c, kg[0], kg[1], kg[2], &kq[0], &kq[1], &kq[2]);
    double kn = kg[0] * kg[0] + kg[1] * kg[1] + kg[2] * kg[2];
    if (kn == 0.) {
        double den = (c11 - c12) * (c11 + 2. * c12);
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
    double chi = (c11 - c12 - 2. * c44) / c44;
    double x2 = kq[0] * kq[0], y2 = kq[1] * kq[1], z2 = kq[2] * kq[2];
    double d0 = c11 * kn * kn * kn + chi * (c11 + c12) * kn * (x2 * y2 + y2 * z2 + z2 * x2)
              + chi * chi * (c11 + 2. * c12 + c44) * x2 * y2 * z2;
    double G[3][3];
    G[0][0] = (c44 * kn * kn + (c11 + c12) * chi * y2 * z2 + (c11 - c44) * kn * (y2 + z2)) / (c44 * d0);
    G[1][1] = (c44 * kn * kn + (c11 + c12) * chi * z2 * x2 + (c11 - c44) * kn * (z2 + x2)) / (c44 * d0);
    G[2][2] = (c44 * kn * kn + (c11 + c12) * chi * x2 * y2 + (c11 - c44) * kn * (x2 + y2)) / (c44 * d0);
    G[0][1] = G[1][0] = -(c12 + c44) * (kn + chi * z2) * kq[0] * kq[1] / (c44 * d0);
    G[1][2] = G[2][1] = -(c12 + c44) * (kn + chi * x2) * kq[1] * kq[2] / (c44 * d0);
    G[0][2] = G[2][0] = -(c12 + c44) * (kn + chi * y2) * kq[0] * kq[2] / (c44 * d0);
    for (int r = 0; r < 2; r++) {
        double S[3][3] = {{sig[0][r], sig[5][r], sig[4][r]}, {sig[5][r], sig[1][r], sig[3][r]},
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

/* The eigenstrain of the magnets and ferroelectrics in the crystal frame
 * (elastic_dynamic_solver.cpp:21-73 in its sum order, without the flexo
 * terms), Voigt order xx yy zz yz xz xy. */
static void es_eigen(const gfo_ctx* c, int64_t id, double e[6]) {
    const gf_material* mat = matof(c, id);
    for (int a = 0; a < 6; a++) e[a] = 0.;
    if (!mat) return;
    double mx = 0, my = 0, mz = 0, ax = 0, ay = 0, az = 0, bx = 0, by = 0, bz = 0, px = 0, py = 0, pz = 0;
    if (c->if_FM_all) glb2crt(c, c->ms[0][id], c->ms[1][id], c->ms[2][id], &mx, &my, &mz);
    if (c->if_AFM_all) {
        glb2crt(c, c->ms1[0][id], c->ms1[1][id], c->ms1[2][id], &ax, &ay, &az);
        glb2crt(c, c->ms2[0][id], c->ms2[1][id], c->ms2[2][id], &bx, &by, &bz);
    }
    if (c->if_FE_all) glb2crt(c, c->ps[0][id], c->ps[1][id], c->ps[2][id], &px, &py, &pz);
    double l100 = mat->lamda100, l111 = mat->lamda111, l100a = mat->lamda100_AFM1, l111a = mat->lamda111_AFM1;
    double l100b = mat->lamda100_AFM2, l111b = mat->lamda111_AFM2;
    double Q11 = mat->Q11, Q12 = mat->Q12, Q44 = mat->Q44;
    e[0] = 1.5 * l100 * (mx * mx - 1. / 3.) + Q11 * (px * px) + Q12 * (py * py + pz * pz)
         + 0.5 * 1.5 * l100a * (ax * ax - 1. / 3.) + 0.5 * 1.5 * l100b * (bx * bx - 1. / 3.);
    e[1] = 1.5 * l100 * (my * my - 1. / 3.) + Q11 * (py * py) + Q12 * (px * px + pz * pz)
         + 0.5 * 1.5 * l100a * (ay * ay - 1. / 3.) + 0.5 * 1.5 * l100b * (by * by - 1. / 3.);
    e[2] = 1.5 * l100 * (mz * mz - 1. / 3.) + Q11 * (pz * pz) + Q12 * (py * py + px * px)
         + 0.5 * 1.5 * l100a * (az * az - 1. / 3.) + 0.5 * 1.5 * l100b * (bz * bz - 1. / 3.);
    e[3] = 1.5 * l111 * my * mz + Q44 * py * pz + 0.5 * 1.5 * l111a * ay * az + 0.5 * 1.5 * l111b * by * bz;
    e[4] = 1.5 * l111 * mx * mz + Q44 * px * pz + 0.5 * 1.5 * l111a * ax * az + 0.5 * 1.5 * l111b * bx * bz;
    e[5] = 1.5 * l111 * my * mx + Q44 * py * px + 0.5 * 1.5 * l111a * ay * ax + 0.5 * 1.5 * l111b * by * bx;
}

void gfo_elastostatic(gfo_ctx* c) {
    int64_t nx = c->nx, ny = c->ny, nz = c->nz, nzh = nz / 2 + 1, nh = nx * ny * nzh, n = c->n;
    double c11 = c->ch11, c12 = c->ch12, c44 = c->ch44;
    double* sg = (double*)malloc(sizeof(double) * 6 * n);
    double* S = (double*)malloc(sizeof(double) * 2 * nh * 6);
    double* E = (double*)malloc(sizeof(double) * 2 * nh * 6);
    for (int64_t id = 0; id < n; id++) {
        double e[6];
        es_eigen(c, id, e);
        /* the static eigenstrain ex*0t0_crt, kept beside the static strain */
        for (int a = 0; a < 6; a++) c->e0t0[a][id] = e[a];
        sg[0 * n + id] = c11 * e[0] + c12 * (e[1] + e[2]);
        sg[1 * n + id] = c11 * e[1] + c12 * (e[0] + e[2]);
        sg[2 * n + id] = c11 * e[2] + c12 * (e[0] + e[1]);
        sg[3 * n + id] = 2. * c44 * e[3];
        sg[4 * n + id] = 2. * c44 * e[4];
        sg[5 * n + id] = 2. * c44 * e[5];
    }
    for (int a = 0; a < 6; a++) gfo_rfft3(nx, ny, nz, sg + a * n, S + a * 2 * nh);
    for (int64_t i = 0; i < nx; i++)
        for (int64_t j = 0; j < ny; j++)
            for (int64_t k = 0; k < nzh; k++) {
                int nyq[3];
                double kg[3];
                kg[0] = kwave(i, nx, c->dx, &nyq[0]);
                kg[1] = kwave(j, ny, c->dy, &nyq[1]);
                kg[2] = kwave(k, nz, c->dz, &nyq[2]);
                int64_t q = (i * ny + j) * nzh + k;
                double sig[6][2], acc[6][2] = {{0.}}, eps[6][2];
                for (int a = 0; a < 6; a++) { sig[a][0] = S[a * 2 * nh + 2 * q]; sig[a][1] = S[a * 2 * nh + 2 * q + 1]; }
                int nv = 0;
                for (int f = 0; f < 8; f++) {
                    if (((f & 1) && !nyq[0]) || ((f & 2) && !nyq[1]) || ((f & 4) && !nyq[2])) continue;
                    double kf[3] = {(f & 1) ? -kg[0] : kg[0], (f & 2) ? -kg[1] : kg[1], (f & 4) ? -kg[2] : kg[2]};
                    es_operator(c, kf, sig, eps);
                    for (int a = 0; a < 6; a++) { acc[a][0] += eps[a][0]; acc[a][1] += eps[a][1]; }
                    nv++;
                }
                for (int a = 0; a < 6; a++) {
                    E[a * 2 * nh + 2 * q] = acc[a][0] / (double)nv;
                    E[a * 2 * nh + 2 * q + 1] = acc[a][1] / (double)nv;
                }
            }
    for (int a = 0; a < 6; a++) {
        gfo_irfft3(nx, ny, nz, E + a * 2 * nh, c->e_t0[a]);
        for (int64_t q = 0; q < n; q++) c->e_t0[a][q] = c->e_t0[a][q] / (double)n;
    }
    free(sg); free(S); free(E);
}

void gfo_demag(gfo_ctx* c) {
    /* main.cpp:214 without elastodynamics; with it once at start-up unless
     * a static strain or eigenstrain was read (main.cpp:148-158) */
    if ((c->flags & GF_FLAG_ELASTOSTATIC) && !(c->if_elasto && c->t0_fixed)) {
        gfo_elastostatic(c);
        c->t0_fixed = 1;
    }
    if ((c->flags & GF_FLAG_ELECTROSTATICS) && c->if_FE_all) gfo_estatic(c);
    if (!(c->if_FM_all || c->if_AFM_all)) return;
    if (!(c->flags & (GF_FLAG_MAGNETOSTATICS | GF_FLAG_MAG_1D | GF_FLAG_DEMAG_FACTOR))) return;
    if (c->flags & GF_FLAG_MAG_1D) demag_1d(c);
    else demag_3d(c);
}

/* --------------------------------------------------------- protocol */
/* main.cpp:243-483 restricted to the magnetic physics: time march; for
 * each stage all get_d*, then all updates; the 1D stray field inside the
 * updates (magnetic_solver.cpp:1533-1538, magnetic_AFM_solver.cpp:1746-1751);
 * the 3D/demag-factor static between stages and after the step unless the
 * EM solver is on. */
void gfo_step(gfo_ctx* c, int64_t nsteps) {
    int mag = c->if_FM_all || c->if_AFM_all;
    int ms = (c->flags & GF_FLAG_MAGNETOSTATICS) != 0;
    int em = (c->flags & GF_FLAG_EMDYNAMIC) != 0;
    int one_d = (c->flags & GF_FLAG_MAG_1D) != 0;
    for (int64_t st = 0; st < nsteps; st++) {
        gfo_time_advance(c);
        for (int s = 1; s <= 4; s++) {
            if (c->if_elasto) gfo_elastic_get(c, s);
            if (c->if_FM_all) gfo_mag_get_dm(c, s);
            if (c->if_AFM_all) gfo_afm_get_dm(c, s);
            if (c->if_FE_all) gfo_fe_get(c, s);
            if (c->if_em) gfo_em_get(c, s);
            if (c->if_elasto) gfo_elastic_update(c, s);
            if (c->if_FM_all) gfo_mag_update(c, s);
            if (c->if_AFM_all) gfo_afm_update(c, s);
            if (c->if_FE_all) gfo_fe_update(c, s);
            if (c->if_em) gfo_em_update(c, s);
            if (!em && c->if_FE_all && (c->flags & GF_FLAG_ELECTROSTATICS)) gfo_estatic(c);
            if (!em && one_d && ((c->if_FM_all && ms) || c->if_AFM_all)) demag_1d(c);
            if (!em && mag && !one_d && ms) demag_3d(c);
        }
        /* main.cpp:472-474: the static strain after the step without elastodynamics */
        if ((c->flags & GF_FLAG_ELASTOSTATIC) && !c->if_elasto) gfo_elastostatic(c);
    }
}

/* magnetic_solver.cpp:1863-1888, magnetic_AFM_solver.cpp:2221-2266,
 * ferroelectric_solver.cpp:2031-2055 */
void gfo_average(const gfo_ctx* c, int id, double out[3]) {
    double** f = field_ptrs((gfo_ctx*)c, id);
    int64_t cnt = id == GF_FIELD_M ? c->NFM : (id == GF_FIELD_P ? c->NFE : c->NAFM);
    for (int a = 0; a < 3; a++) {
        double s = 0.;
        for (int64_t i = 0; i < c->n; i++) s = s + f[a][i];
        out[a] = s / (double)cnt;
    }
}

double gfo_time(const gfo_ctx* c) { return c->time; }

/* Host threads the parallel loops use (1 in a build without OpenMP). */
int gfo_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ------------------------------------------- tangent-linear reference */
/* Forward-mode derivative of the probe spectral loss of the adjoint sweep
 * (no counterpart in the reference; test infrastructure): the primal runs
 * through the oracle's own stages (gfo_mag_get_dm / gfo_mag_update /
 * demag_3d), the tangent through the linearisation of
 * magnetic_solver.cpp:3-229 written out term by term:
 *   dH = N (*) (Ms dm) + 2/(mu0 Ms) sum Aij d2(dm) + Ja(m) dm + Me dm
 *        + dHdrive s(t) + (dAex/Aex) Hex + (dB1/(mu0 Ms)) f (e_aa m_a)
 *        + 2D/(mu0 Ms) (d_zx, d_zy, -(d_xx + d_yy))(dm) - (dAex/Aex) (its
 *          surface-slope part, linear in D/(2A))
 *   dk = P (dm x H + m x dH) + Q (dm x p + m x dp) + dP p + dQ q
 * and the RK4 combination of the tangents (Hex without the surface slope,
 * which enters the exchange as A D/(2A)).  FM only, one FM material. */
static void tl_stage(gfo_ctx* c, int stage, const gf_material* mat, double* const* dms, double* const* dk,
                     const double dpar[3], double* const* dhdrive, double* const* dHd, double* const* dhscratch) {
    int64_t n = c->n, ny = c->ny, nz = c->nz;
    const double drive_t = stage == 1 ? c->time - c->dt : (stage == 4 ? c->time : c->time - c->dt * 0.5);
    const double drive_s = sin(2. * PI * c->hdrive_freq * drive_t);
    double Hext[3];
    for (int a = 0; a < 3; a++)
        Hext[a] = c->Hext_stat[a] + c->Hext_altr[a] * sin(2. * PI * c->H_altr_freq[a] * drive_t);
    /* tangent of the magnetostatic field */
    int ms_on = (c->flags & GF_FLAG_MAGNETOSTATICS) != 0;
    if (ms_on) {
        double* M = dalloc(3 * n);
        for (int64_t i = 0; i < n; i++) {
            const gf_material* m = matof(c, i);
            if (m && m->if_FM) for (int a = 0; a < 3; a++) M[a * n + i] = m->Ms * dms[a][i];
        }
        conv_3d(c, M, dHd);
        free(M);
    }
    (void)dhscratch;
    for (int64_t i = 0; i < n; i++) {
        const gf_material* m = matof(c, i);
        if (!m || !m->if_FM) { for (int a = 0; a < 3; a++) dk[a][i] = 0.; continue; }
        int64_t x = i / (ny * nz), y = (i - x * (ny * nz)) / nz, z = i - x * (ny * nz) - y * nz;
        double Ms = m->Ms;
        double K1 = m->K1 / MU0 / Ms, K2 = m->K2 / MU0 / Ms, B1 = m->B1 / MU0 / Ms, B2 = m->B2 / MU0 / Ms;
        double mg[3] = {c->ms[0][i], c->ms[1][i], c->ms[2][i]}, dg[3] = {dms[0][i], dms[1][i], dms[2][i]};
        double mc[3], dc[3];
        glb2crt(c, mg[0], mg[1], mg[2], &mc[0], &mc[1], &mc[2]);
        glb2crt(c, dg[0], dg[1], dg[2], &dc[0], &dc[1], &dc[2]);
        /* primal field (same terms as gfo_mag_get_dm).  The stencil is
         * linear in the field for given A and D, so its tangent is the
         * stencil of dm; its surface part (the DMI slope s = D/(2A)) is
         * linear in s, ds/dA = -s/A, and the exchange carries it as A s,
         * which does not depend on A */
        double d[4], d2[3][2][3], dd[4], dd2[3][2][3], dn[4], d2n[3][2][3];
        stencil(c, c->ms, c->surfYZ, c->surfXZ, c->surfXY, 1, m->Aex, m->iDMI, x, y, z, d, d2);
        stencil(c, dms, c->surfYZ, c->surfXZ, c->surfXY, 1, m->Aex, m->iDMI, x, y, z, dd, dd2);
        stencil(c, c->ms, c->surfYZ, c->surfXZ, c->surfXY, 1, m->Aex, 0., x, y, z, dn, d2n);
        double hx[3], dhx[3];
        for (int a = 0; a < 3; a++) {
            hx[a] = 2. / MU0 / Ms * (c->Aij[0][i] * d2[0][0][a] + c->Aij[1][i] * d2[0][1][a] + c->Aij[2][i] * d2[1][0][a]
                                    + c->Aij[3][i] * d2[1][1][a] + c->Aij[4][i] * d2[2][0][a] + c->Aij[5][i] * d2[2][1][a]);
            dhx[a] = 2. / MU0 / Ms * (c->Aij[0][i] * dd2[0][0][a] + c->Aij[1][i] * dd2[0][1][a] + c->Aij[2][i] * dd2[1][0][a]
                                     + c->Aij[3][i] * dd2[1][1][a] + c->Aij[4][i] * dd2[2][0][a] + c->Aij[5][i] * dd2[2][1][a]);
            /* the faces between cells, linear in A */
            const double hxn = 2. / MU0 / Ms * (c->Aij[0][i] * d2n[0][0][a] + c->Aij[1][i] * d2n[0][1][a]
                                               + c->Aij[2][i] * d2n[1][0][a] + c->Aij[3][i] * d2n[1][1][a]
                                               + c->Aij[4][i] * d2n[2][0][a] + c->Aij[5][i] * d2n[2][1][a]);
            dhx[a] += hxn / m->Aex * dpar[1];
        }
        /* the DMI field (magnetic_solver.cpp, as gfo_mag_get_dm) and its tangent */
        const double kd = 2. * m->iDMI / MU0 / Ms;
        double ddA[4];
#if GFO_PLANT == 9
        for (int q = 0; q < 4; q++) ddA[q] = dd[q];
#else
        for (int q = 0; q < 4; q++) ddA[q] = dd[q] - (d[q] - dn[q]) / m->Aex * dpar[1];
#endif
        const double hdm[3] = {kd * d[2], kd * d[3], -kd * (d[0] + d[1])};
        const double dhdm[3] = {kd * ddA[2], kd * ddA[3], -kd * (ddA[0] + ddA[1])};
        double hac[3] = {0., 0., 0.}, dhac[3] = {0., 0., 0.};
        double X = mc[0], Y = mc[1], Z = mc[2], dX = dc[0], dY = dc[1], dZ = dc[2];
        if (mat->anisotropy_type == 1) {
            hac[0] = -2. * (K1 * (Y * Y + Z * Z) + K2 * Y * Y * Z * Z) * X;
            hac[1] = -2. * (K1 * (X * X + Z * Z) + K2 * X * X * Z * Z) * Y;
            hac[2] = -2. * (K1 * (X * X + Y * Y) + K2 * X * X * Y * Y) * Z;
            /* product rule on each component */
            dhac[0] = -2. * (K1 * (2. * Y * dY + 2. * Z * dZ) + K2 * (2. * Y * dY * Z * Z + Y * Y * 2. * Z * dZ)) * X
                      - 2. * (K1 * (Y * Y + Z * Z) + K2 * Y * Y * Z * Z) * dX;
            dhac[1] = -2. * (K1 * (2. * X * dX + 2. * Z * dZ) + K2 * (2. * X * dX * Z * Z + X * X * 2. * Z * dZ)) * Y
                      - 2. * (K1 * (X * X + Z * Z) + K2 * X * X * Z * Z) * dY;
            dhac[2] = -2. * (K1 * (2. * X * dX + 2. * Y * dY) + K2 * (2. * X * dX * Y * Y + X * X * 2. * Y * dY)) * Z
                      - 2. * (K1 * (X * X + Y * Y) + K2 * X * X * Y * Y) * dZ;
        } else if (mat->anisotropy_type == 2) {
            hac[2] = 2. * K1 * Z;
            dhac[2] = 2. * K1 * dZ;
        }
        double exx = 0., eyy = 0., ezz = 0., eyz = 0., exz = 0., exy = 0.;
        if (static_strain(c)) {
            exx = c->e_t0[0][i]; eyy = c->e_t0[1][i]; ezz = c->e_t0[2][i];
            eyz = c->e_t0[3][i]; exz = c->e_t0[4][i]; exy = c->e_t0[5][i];
        }
        double hec[3], dhec[3];
        hec[0] = -2. * (B1 * exx * X + B2 * (exy * Y + exz * Z));
        hec[1] = -2. * (B1 * eyy * Y + B2 * (exy * X + eyz * Z));
        hec[2] = -2. * (B1 * ezz * Z + B2 * (exz * X + eyz * Y));
        dhec[0] = -2. * (B1 * exx * dX + B2 * (exy * dY + exz * dZ)) - 2. * dpar[2] / MU0 / Ms * exx * X;
        dhec[1] = -2. * (B1 * eyy * dY + B2 * (exy * dX + eyz * dZ)) - 2. * dpar[2] / MU0 / Ms * eyy * Y;
        dhec[2] = -2. * (B1 * ezz * dZ + B2 * (exz * dX + eyz * dY)) - 2. * dpar[2] / MU0 / Ms * ezz * Z;
        double ha[3], dha[3], he[3], dhe[3];
        crt2glb(c, hac[0], hac[1], hac[2], &ha[0], &ha[1], &ha[2]);
        crt2glb(c, dhac[0], dhac[1], dhac[2], &dha[0], &dha[1], &dha[2]);
        crt2glb(c, hec[0], hec[1], hec[2], &he[0], &he[1], &he[2]);
        crt2glb(c, dhec[0], dhec[1], dhec[2], &dhe[0], &dhe[1], &dhe[2]);
        double H[3], dH[3];
        for (int a = 0; a < 3; a++) {
            H[a] = c->H[a][i] + ha[a] + he[a] + hx[a] + hdm[a] + Hext[a] + c->Hnu[a][i];
            dH[a] = dha[a] + dhe[a] + dhx[a] + dhdm[a];
            if (ms_on) dH[a] += dHd[a][i];
            if (c->flags & GF_FLAG_HDRIVE) {
                H[a] += c->Hdrive[a][i] * drive_s;
                if (dhdrive) dH[a] += dhdrive[a][i] * drive_s;
            }
        }
        double alpha = mat->FM_damping, gyro = mat->gyro;
        double P = -gyro / (1 + alpha * alpha) * c->dt, Q = alpha * P;
        double dP = -2. * alpha * P / (1. + alpha * alpha) * dpar[0];
        double dQ = P * dpar[0] + alpha * dP;
        double p[3] = {mg[1] * H[2] - mg[2] * H[1], mg[2] * H[0] - mg[0] * H[2], mg[0] * H[1] - mg[1] * H[0]};
        double q[3] = {mg[1] * p[2] - mg[2] * p[1], mg[2] * p[0] - mg[0] * p[2], mg[0] * p[1] - mg[1] * p[0]};
        double dp[3] = {dg[1] * H[2] - dg[2] * H[1] + mg[1] * dH[2] - mg[2] * dH[1],
                        dg[2] * H[0] - dg[0] * H[2] + mg[2] * dH[0] - mg[0] * dH[2],
                        dg[0] * H[1] - dg[1] * H[0] + mg[0] * dH[1] - mg[1] * dH[0]};
        double dq[3] = {dg[1] * p[2] - dg[2] * p[1] + mg[1] * dp[2] - mg[2] * dp[1],
                        dg[2] * p[0] - dg[0] * p[2] + mg[2] * dp[0] - mg[0] * dp[2],
                        dg[0] * p[1] - dg[1] * p[0] + mg[0] * dp[1] - mg[1] * dp[0]};
        for (int a = 0; a < 3; a++) dk[a][i] = P * dp[a] + Q * dq[a] + dP * p[a] + dQ * q[a];
    }
}

int gfo_tl_probe_loss(gfo_ctx* c, int material, const int lo[3], const int hi[3], double freq, int64_t nsteps,
                      const double dpar[3], const double* dhdrive_flat, const double* dm0_flat, double* loss,
                      double* dloss) {
    if (material < 1 || material > c->nmat || !c->mat[material - 1].if_FM) return -1;
    if (c->if_AFM_all || c->if_FE_all || c->if_elasto || c->if_em) return -1;
    const gf_material* mat = &c->mat[material - 1];
    int64_t n = c->n;
    double *dm[3], *dms[3], *dk[4][3], *dHd[3], *dhd[3];
    for (int a = 0; a < 3; a++) {
        dm[a] = dalloc(n); dms[a] = dalloc(n); dHd[a] = dalloc(n);
        for (int s = 0; s < 4; s++) dk[s][a] = dalloc(n);
        if (dm0_flat) memcpy(dm[a], dm0_flat + a * n, sizeof(double) * (size_t)n);
        memcpy(dms[a], dm[a], sizeof(double) * (size_t)n);
        dhd[a] = (double*)(dhdrive_flat ? dhdrive_flat + a * n : NULL);
    }
    int ms_on = (c->flags & GF_FLAG_MAGNETOSTATICS) != 0;
    if (ms_on) demag_3d(c);
    double Sr = 0., Si = 0., dSr = 0., dSi = 0.;
    double w = 2. * PI * freq;
    int64_t np = (int64_t)(hi[0] - lo[0] + 1) * (hi[1] - lo[1] + 1) * (hi[2] - lo[2] + 1);
    for (int64_t st = 0; st < nsteps; st++) {
        gfo_time_advance(c);
        for (int s = 1; s <= 4; s++) {
            tl_stage(c, s, mat, dms, dk[s - 1], dpar, dhdrive_flat ? dhd : NULL, dHd, NULL);
            gfo_mag_get_dm(c, s);
            gfo_mag_update(c, s);
            for (int a = 0; a < 3; a++)
                for (int64_t i = 0; i < n; i++) {
                    if (s == 1 || s == 2) dms[a][i] = dm[a][i] + dk[s - 1][a][i] * 0.5;
                    else if (s == 3) dms[a][i] = dm[a][i] + dk[2][a][i];
                    else {
                        dm[a][i] = dm[a][i] + (dk[0][a][i] / 6. + dk[1][a][i] / 3. + dk[2][a][i] / 3. + dk[3][a][i] / 6.);
                        dms[a][i] = dm[a][i];
                    }
                }
            if (ms_on) demag_3d(c);
        }
        double X = 0., dX = 0.;
        for (int i = lo[0]; i <= hi[0]; i++)
            for (int j = lo[1]; j <= hi[1]; j++)
                for (int k = lo[2]; k <= hi[2]; k++) {
                    int64_t id = ((int64_t)i * c->ny + j) * c->nz + k;
                    X += c->m[0][id];
                    dX += dm[0][id];
                }
        X /= (double)np;
        dX /= (double)np;
        double wr = c->dt * cos(-w * c->time), wi = c->dt * sin(-w * c->time);
        Sr += X * wr; Si += X * wi;
        dSr += dX * wr; dSi += dX * wi;
    }
    *loss = Sr * Sr + Si * Si;
    *dloss = 2. * (Sr * dSr + Si * dSi);
    for (int a = 0; a < 3; a++) {
        free(dm[a]); free(dms[a]); free(dHd[a]);
        for (int s = 0; s < 4; s++) free(dk[s][a]);
    }
    return 0;
}
