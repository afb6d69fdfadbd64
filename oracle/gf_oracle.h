/*
 * gf_oracle.h — CPU restatement of the reference's FP64 RK4 hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker the HIP product is
 * compared against; only tests/, __graft_entry__.smoke() and the CPU-baseline
 * leg of bench.py may load it.  The product never links or calls it.
 *
 * It follows the reference's own data layout and protocol (separate k-arrays
 * per stage, one store buffer, the get -> wait -> update order of
 * main.cpp:243-483), not the fused design of the GPU path, so that a fusion
 * mistake on the GPU cannot be mirrored here.
 */
#ifndef GF_ORACLE_H
#define GF_ORACLE_H

#include <stddef.h>
#include <stdint.h>
#include "../include/goferro.h"   /* gf_desc / gf_material / field ids (types only) */

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gfo_ctx gfo_ctx;

/* oracle-only field: the combined RK4 increment dm of the last step */
#define GFO_FIELD_DM 100
/* oracle-only field: the FE q store (the ABI exposes only the p store) */
#define GFO_FIELD_Q_STORE 101
/* oracle-only field: the elastic velocity increment of stage 1 */
#define GFO_FIELD_DV1 102
/* oracle-only field: the FM increment k of the last stage-1 get_dm */
#define GFO_FIELD_K1 103
#define GFO_FIELD_KA1 104   /* stage-1 increments of the AFM sublattices */
#define GFO_FIELD_KA2 105
#define GFO_FIELD_KQ1 106   /* stage-1 increment of q */
#define GFO_FIELD_E0_DIAG 107   /* dynamic eigenstrain (global frame) */
#define GFO_FIELD_E0_OFF 108
#define GFO_FIELD_DM1 109   /* combined RK4 increments of the AFM sublattices, last step */
#define GFO_FIELD_DM2 110

gfo_ctx* gfo_create(const gf_desc* desc);
void gfo_destroy(gfo_ctx* c);
int gfo_set_field(gfo_ctx* c, int field_id, const double* src, size_t count);
int gfo_get_field(const gfo_ctx* c, int field_id, double* dst, size_t count);

/* protocol pieces (reference names in the .c file) */
void gfo_time_advance(gfo_ctx* c);
void gfo_mag_get_dm(gfo_ctx* c, int stage);   /* get_dm_RKs (+ get_dm at s=4) */
void gfo_mag_update(gfo_ctx* c, int stage);   /* update_m_RKs / update_m      */
void gfo_afm_get_dm(gfo_ctx* c, int stage);
void gfo_afm_update(gfo_ctx* c, int stage);
void gfo_fe_get(gfo_ctx* c, int stage);       /* get_dq_RKs + get_dp_RKs      */
void gfo_fe_update(gfo_ctx* c, int stage);    /* update_q_RKs + update_p_RKs  */
void gfo_elastic_get(gfo_ctx* c, int stage);  /* get_dv_RKs + get_du_RKs      */
void gfo_elastic_update(gfo_ctx* c, int stage); /* update_v_RKs + update_u_RKs */
void gfo_em_get(gfo_ctx* c, int stage);       /* get_dE_RK + get_dH_RK + Jp/Jf */
void gfo_em_update(gfo_ctx* c, int stage);    /* update_DE_RK + update_DH_RK  */
void gfo_demag(gfo_ctx* c);
void gfo_prescribe_m(gfo_ctx* c);              /* update_prescribed_m at start-up */
void gfo_estatic(gfo_ctx* c);
void gfo_elastostatic(gfo_ctx* c);             /* static strain of the magnets' eigenstrain */                   /* get_H_static / _1D           */
void gfo_step(gfo_ctx* c, int64_t nsteps);    /* main.cpp:243-483             */
void gfo_average(const gfo_ctx* c, int field_id, double out[3]);
double gfo_time(const gfo_ctx* c);

/* Tangent-linear reference of the adjoint sweep's probe spectral loss:
 * nsteps FM steps from the current state; loss = |sum_n X_n e^{-i w t_n} dt|^2
 * with X_n the probe mean of m_x; dloss = its derivative along
 * (dpar = dalpha, dAex, dB1 of `material`, dhdrive 3n or NULL, dm0 3n or
 * NULL).  Returns 0, or -1 for an unsupported context. */
int gfo_tl_probe_loss(gfo_ctx* c, int material, const int lo[3], const int hi[3], double freq, int64_t nsteps,
                      const double dpar[3], const double* dhdrive, const double* dm0, double* loss, double* dloss);

/* building blocks, exposed so tests can pin them one by one */
void gfo_newell_real(int64_t nx, int64_t ny, int64_t nz, double dx, double dy, double dz,
                     double* out9);   /* 9 placed, /n real kernels, each n doubles */
double gfo_newell_F2(double x, double y, double z);
double gfo_newell_G2(double x, double y, double z);
double gfo_newell_A(double x, double y, double z, double dx, double dy, double dz);
double gfo_newell_B(double x, double y, double z, double dx, double dy, double dz);
/* unnormalised real 3D transforms with the half spectrum along z:
 * in: nx*ny*nz reals; out: nx*ny*(nz/2+1) complex as interleaved doubles */
void gfo_rfft3(int64_t nx, int64_t ny, int64_t nz, const double* in, double* out);
void gfo_irfft3(int64_t nx, int64_t ny, int64_t nz, const double* in, double* out);
/* naive O(N^2) DFT of the same transform, for pinning gfo_rfft3 */
void gfo_rdft3_naive(int64_t nx, int64_t ny, int64_t nz, const double* in, double* out);
/* host threads the oracle's parallel loops use */
int gfo_threads(void);

#ifdef __cplusplus
}
#endif
#endif
