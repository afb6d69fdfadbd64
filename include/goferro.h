/*
 * goferro.h — C ABI of the MI355X (gfx950) coupled RK4 hot path.
 *
 * The reference drives its physics through argument-less member functions
 * that act on OpenACC-resident arrays (magnetic_system.h:160-210,
 * EMdynamic_system.h:132-179, ferroelectric_system.h:96-131,
 * elastic_system.h:172-224) in the fixed protocol of main.cpp:243-483.
 * This header replaces that surface with plain C: a context owns the device
 * state, the caller owns host buffers, every call is ordered on the context's
 * HIP stream, and every call returns a status (0 = OK, < 0 = error).
 *
 * Layout contract (matrix.h:46-52): every cell field is an SoA array of
 * nx*ny*nz doubles indexed (i*ny + j)*nz + k, z fastest.  material_cell is
 * uint32 with 0 = vacuum and m >= 1 selecting materials[m-1] (global.h:91-92).
 *
 * No torch types appear here; the PyTorch extension and the ctypes wrapper
 * call the same entry points.
 */
#ifndef GOFERRO_H
#define GOFERRO_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GF_ABI_VERSION 4

/* ---- status codes ------------------------------------------------------ */
enum gf_status_code {
    GF_OK = 0,
    GF_ERR_ARG = -1,          /* invalid argument (null pointer, bad id, size) */
    GF_ERR_UNSUPPORTED = -2,  /* valid input the build does not implement */
    GF_ERR_HIP = -3,          /* HIP runtime error; text in gf_last_error */
    GF_ERR_ALLOC = -4,        /* host or device allocation failed */
    GF_ERR_STATE = -5         /* call out of protocol order */
};
typedef int gf_status;

/* ---- physics flags (gf_desc.flags) -------------------------------------- */
#define GF_FLAG_MAGNETOSTATICS   (1u << 0)  /* if_magnetostatics */
#define GF_FLAG_MAG_1D           (1u << 1)  /* if_mag_1Dmodel: Hz = -Ms mz */
#define GF_FLAG_DEMAG_FACTOR     (1u << 2)  /* if_demag_factor */
#define GF_FLAG_PRESCRIBE_HEXT   (1u << 3)  /* if_prescribe_Hext (table drive) */
#define GF_FLAG_ORIENT_ROTATE    (1u << 4)  /* if_OrientRotate */
#define GF_FLAG_HEXT_NONUNIF     (1u << 5)  /* Hext_nonunif field uploaded */
#define GF_FLAG_EMDYNAMIC        (1u << 6)  /* if_EMdynamic (statics frozen) */
#define GF_FLAG_PRESCRIBE_EEXT   (1u << 7)  /* if_prescribe_Eext (table drive) */
#define GF_FLAG_ELASTODYNAMICS   (1u << 8)  /* if_elastodynamics */
#define GF_FLAG_ELASTO_BACKACTION (1u << 9) /* if_elasto_backaction_from_pORm: eigenstrain from m, p */
#define GF_FLAG_ELASTO_ON_PM     (1u << 10) /* if_elasto_on_pORm: strain acts on m and p */
#define GF_FLAG_GAUSSIAN_PULSE   (1u << 11) /* if_gaussian_strain_pulse */
#define GF_FLAG_1D_ABC           (1u << 12) /* if_1D_ABC: Mur instead of Liao */
#define GF_FLAG_1D_ABC_ELAST_ONLYTOP (1u << 13) /* if_1D_ABC_ELAST_onlytop */
#define GF_FLAG_EM_FROM_P        (1u << 14) /* if_EM_fromP: dP/dt is a source current */
#define GF_FLAG_EM_FROM_M        (1u << 15) /* if_EM_fromM: dM/dt enters Faraday's law */
#define GF_FLAG_EM_BACKACTION_P  (1u << 16) /* if_EM_backaction_on_P: E_em drives p */
#define GF_FLAG_EM_BACKACTION_M  (1u << 17) /* if_EM_backaction_on_M: H_em drives m */
#define GF_FLAG_PLANE_EM         (1u << 18) /* if_input_planeEM_E: hard plane source */
#define GF_FLAG_JF_INPUT         (1u << 19) /* if_Jf_input: free-current source box */
#define GF_FLAG_PEC_XY           (1u << 20) /* if_PEC_XY: x and y walls are PEC */
#define GF_FLAG_PEC_Z            (1u << 21) /* if_PEC_Z: z walls are PEC */
#define GF_FLAG_1D_ABC_EM_ONLYTOP (1u << 22) /* if_1D_ABC_EM_onlytop (with GF_FLAG_1D_ABC) */
/* Spatially shaped AC drive of the ferromagnet: H += Hdrive(r) sin(2 pi
 * Hdrive_freq t) at the stage times of the uniform drive, Hdrive uploaded as
 * GF_FIELD_HDRIVE.  No counterpart in the reference (its non-uniform field
 * Hext_nonunif is static); it carries the per-cell drive amplitudes whose
 * gradients the adjoint returns. */
#define GF_FLAG_HDRIVE           (1u << 23)
/* FFT electrostatics of the ferroelectric (get_E_static, declared at
 * ferroelectric_system.h:103 but not defined in the reference): the static
 * field of the polarisation, E(k) = -k (k.P(k)) / (eps0 er |k|^2), E(0) = 0,
 * with er = max r_permittivity * scale_elec and the continuous wave vectors
 * of global.cpp:270-318, recomputed into GF_FIELD_ESTAT by the initial
 * statics (gf_demag) and after every update phase unless the
 * electrodynamics is on (main.cpp:120-167, 299-306).  Not together with
 * the FFT magnetostatics (the two would share one FFT plan); the 1D and
 * demag-factor magnetostatics combine with it. */
#define GF_FLAG_ELECTROSTATICS   (1u << 24)
/* if_spin_pumping: the inverse spin Hall current J_ISHE of the normal-metal
 * layers next to a magnet (get_J_ISHE_RK*, magnetic_solver.cpp:231-283 and
 * its per-stage copies; set-up global.cpp:169-235), recomputed after the
 * magnet stages of every RK stage and added to Ampere's law of the x and y
 * edges as a four-cell mean (EMdynamic_solver.cpp:422-482, 554-614 and the
 * later stages).  The layers come from gf_desc.nz_layer; GF_FIELD_J_ISHE
 * downloads it. */
#define GF_FLAG_SPIN_PUMPING     (1u << 25)
/* if_flexo: flexoelectric coupling of the ferroelectric.  The driving
 * force gains tv4..tv8 times p and its gradients and f11, f12, f44 times
 * the dynamic strain gradients (ferroelectric_solver.cpp:800-832 and the
 * later stages; the gradients of get_p_gradient, :134-281, and
 * get_Dstrain_gradient, elastic_dynamic_solver.cpp:230-399); the dynamic
 * eigenstrain gains -F grad p (elastic_dynamic_solver.cpp:82-89).  F11,
 * F12, F44 and tv4..tv8 are derived from c11, c12, c44, Q11, Q12, Q44 and
 * f11, f12, f44 as read.cpp:181-199 does (zero for a non-ferroelectric). */
#define GF_FLAG_FLEXO            (1u << 26)
/* if_elastostatic: the static strain of the eigenstrain of the magnets
 * and ferroelectrics (get_strain_static, declared at elastic_system.h:190
 * but not defined in the reference), in a periodic body of the homogeneous
 * cubic stiffness of
 * elastic_initialize.cpp:22-40 (component-wise largest c11, c12, c44 times
 * scale_elasto) through its Green tensor G_ik (:340-389):
 * e(k) = sym(k (x) G(k) sigma0(k) k) for k != 0, with sigma0 = C e0 and
 * e0 the eigenstrain of elastic_dynamic_solver.cpp:4-104, and the mean
 * strain equal to the mean eigenstrain (zero mean stress).  A Nyquist
 * component enters symmetrised over its sign, so the strain is real.
 * Written into GF_FIELD_STRAIN_T0_* (crystal frame), with e0 into
 * GF_FIELD_EIGENSTRAIN_T0_*: without elastodynamics by gf_demag and after
 * every gf_step step (main.cpp:214, 473); with it once, by the first
 * gf_demag, unless a static strain or eigenstrain was uploaded first
 * (main.cpp:148-158), and the elastodynamics then carries the change from
 * t0.  With GF_FLAG_ELASTO_ON_PM the magnets read the static strain (plus
 * the dynamic one) and the ferroelectrics both, added to the dynamic ones
 * (ferroelectric_solver.cpp:636-650).  Without the flexoelectric coupling;
 * power-of-two nx, ny, nz up to 2048. */
#define GF_FLAG_ELASTOSTATIC     (1u << 27)
/* if_prescribed_m: the ferromagnetic cells of the bottom layer (k = 0)
 * precess on a prescribed cone, m = (sqrt(1 - m0^2), m0 sin wt, m0 cos wt)
 * with m0 = sin(precess_angle / 180 pi) and w = 2 pi precess_frequency
 * (magnetic_solver.cpp:1891-1957), written over the RK update of every
 * stage: the stage-2 and stage-3 stores at t - dt/2, the stage-4 store at
 * t and m itself after the step at t (:1526-1609).  gf_prescribe_m sets
 * the layer at the current time, as initialize_device does at start-up
 * (magnetic_initialize.cpp:1074-1077). */
#define GF_FLAG_PRESCRIBED_M     (1u << 28)

/*
 * One material record.  Field meaning and units follow the reference's
 * material class (material.h:3-58) as filled by read_materials
 * (read.cpp:113-215), including the derived B1/B2 (read.cpp:141-142,
 * 162-165).  Booleans are int32 (0/1).
 */
typedef struct gf_material {
    /* ferromagnet (material.h:12-18) */
    int32_t if_FM;
    int32_t anisotropy_type;   /* 1 = cubic, 2 = uniaxial along crystal z */
    double Ms, gyro, FM_damping;
    double K1, K2, Aex, iDMI;
    double B1, B2;
    /* two-sublattice antiferromagnet (material.h:24-35) */
    int32_t if_AFM;
    int32_t reserved0;
    double Ms_AFM1, Ms_AFM2;
    double gyro_AFM1, gyro_AFM2;
    double damping_AFM1, damping_AFM2;
    double K1_AFM1, K1_AFM2;
    double uniaxis_AFM[3];     /* already normalised, read.cpp:151-155 */
    double Aex_AFM1, Aex_AFM2;
    double J_AFM;
    double B1_AFM1, B1_AFM2, B2_AFM1, B2_AFM2;
    /* ferroelectric (material.h:41-49), tv1..tv3 derived as read.cpp:188-190 */
    int32_t if_FE;
    int32_t reserved1;
    double FE_mass, FE_damping;
    double a1, a11, a12, a111, a112, a123, a1111, a1112, a1122, a1123;
    double G11;
    double Q11, Q12, Q44;
    double tv1, tv2, tv3;
    /* elastic (material.h:6-9) and magnetostriction constants (material.h:17, 32-33) */
    double c11, c12, c44, density;
    double elast_mass_damping, elast_stiff_damping;
    double cijkl_glb[6][6];    /* global-frame stiffness, global.cpp:365-441 */
    double lamda100, lamda111;
    double lamda100_AFM1, lamda100_AFM2, lamda111_AFM1, lamda111_AFM2;
    /* electromagnetic (material.h:52-57), comp_n* already normalised as
     * global.cpp:360-362 does */
    int32_t if_PEC;
    int32_t reserved2;
    double r_permittivity, conductivity;
    double comp_n1, omega_plasma_n1, tao_e_n1;
    double comp_n2, omega_plasma_n2, tao_e_n2;
    double comp_n3, omega_plasma_n3, tao_e_n3;
    /* spin pumping into a normal metal (material.h, read.cpp:170-171) */
    int32_t if_spin_pump;
    int32_t reserved3;
    double spin_hall_angle, spin_mix_cond, spin_diffus_length;
    /* flexoelectric coefficients (material.h, read.cpp:179) */
    double f11, f12, f44;
} gf_material;

/*
 * Simulation descriptor: the subset of geometry_parameters and
 * global_parameters (read.cpp:4-111, 217-237) that the hot path reads.
 */
typedef struct gf_desc {
    uint32_t struct_size;      /* sizeof(gf_desc), ABI check */
    uint32_t abi_version;      /* GF_ABI_VERSION */
    int64_t nx, ny, nz;        /* nx_system, ny_system, nz_system */
    double dx, dy, dz;
    int32_t periodic[3];       /* periodicX/Y/Z */
    uint32_t flags;            /* GF_FLAG_* */
    double dt;
    int32_t n_materials;
    int32_t reserved0;
    const gf_material* materials;   /* n_materials records, copied at create */
    const uint32_t* material_cell;  /* nx*ny*nz ids, copied at create */
    double Tg2c[3][3];         /* global -> crystal rotation (read.cpp:27-29) */
    double Hext_stat[3];       /* A/m (read.cpp:35) */
    double Hext_altr[3];       /* A/m (read.cpp:36) */
    double H_altr_freq[3];     /* Hz  (read.cpp:37) */
    double demag_fac[3];       /* read.cpp:46 */
    int32_t n_prescribe_Hext;  /* rows of the prescribed drive table */
    int32_t reserved1;
    const double* prescribe_Hext;   /* rows of (t, Hx, Hy, Hz), copied at create */
    /* electric drive of the ferroelectric (read.cpp:40-43) */
    double Eext_stat[3];       /* V/m */
    double Eext_altr[3];
    double E_altr_freq[3];     /* Hz */
    int32_t n_prescribe_Eext;
    int32_t reserved2;
    const double* prescribe_Eext;   /* rows of (t, Ex, Ey, Ez), copied at create */
    /* elastodynamic sources and boundaries (read.cpp:58-66) */
    int32_t input_strain_type;      /* 1 Gaussian displacement, 2 cosine, 3 Gaussian traction */
    int32_t input_strain_sourcez;   /* source plane k */
    int32_t input_strain_component; /* 'x', 'y' or 'z' */
    int32_t num_strain_cycles;
    double sigma_gauss;             /* s (or Hz for type 2) */
    double amplitude_gauss;
    double weighting_factor_fourth_Liao;
    /* electrodynamic sources (read.cpp:48-53, 97-100) */
    double planeEM_E[3];            /* V/m */
    double planeEM_E_freq[3];       /* Hz */
    int32_t planeEM_source_z;       /* plane k of the hard source (EMdynamic_initialize.cpp:18) */
    int32_t Jf_input_type;          /* 0..6 (EMdynamic_solver.cpp:3885-4025) */
    int32_t Jf_input_component;     /* 'x', 'y' or 'z' */
    int32_t num_Jf_cycles;
    int32_t Jfin_lo[3];             /* Jfin_xi, Jfin_yi, Jfin_zi */
    int32_t Jfin_hi[3];             /* Jfin_xf, Jfin_yf, Jfin_zf (inclusive) */
    double Jf_input_amp;            /* A/m^2 */
    double Jf_input_freq;           /* Hz (sigma in s for types 4, 5) */
    double Jf_input_sigma;          /* s (type 6) */
    /* shaped drive (GF_FLAG_HDRIVE) */
    double Hdrive_freq;             /* Hz */
    /* permittivity scale of the electrostatics (read.cpp:56); 0 = 1 */
    double scale_elec;
    /* the layer stack, bottom to top from k = 0 (pt_nz_layer after
     * set_geometry, geometry.cpp:3-26): the spin-pumping set-up walks it
     * (global.cpp:180-235).  n_layers = 0 takes the runs of equal material
     * ids along z of column (0, 0) as the layers. */
    int32_t n_layers;
    int32_t reserved3;
    const int32_t* nz_layer;
    /* stiffness scale of the elastostatics (read.cpp:65); 0 = 1 */
    double scale_elasto;
    /* prescribed precession of the bottom layer (GF_FLAG_PRESCRIBED_M,
     * read.cpp:101) */
    double precess_angle;           /* degrees */
    double precess_frequency;       /* Hz */
} gf_desc;

/* ---- field identifiers for upload/download ------------------------------ */
enum gf_field_id {
    GF_FIELD_M = 0,         /* FM m: 3 arrays (mx,my,mz), 3n doubles       */
    GF_FIELD_M_STORE = 1,   /* FM m at the current RK stage, 3n            */
    GF_FIELD_HSTAT = 2,     /* magnetostatic field Hx,Hy,Hz, 3n            */
    GF_FIELD_HEXT_NONUNIF = 3, /* non-uniform external field, 3n          */
    GF_FIELD_M_AFM1 = 4,    /* AFM sublattice 1, 3n                        */
    GF_FIELD_M_AFM2 = 5,    /* AFM sublattice 2, 3n                        */
    GF_FIELD_P = 6,         /* polarisation p, 3n (C/m^2)                  */
    GF_FIELD_Q = 7,         /* dp/dt, 3n                                   */
    GF_FIELD_P_STORE = 8,   /* p at the current RK stage, 3n               */
    GF_FIELD_ESTAT = 9,     /* static electric field, 3n (input)           */
    GF_FIELD_U = 10,        /* displacement u, 3n (m)                      */
    GF_FIELD_V = 11,        /* velocity v, 3n (m/s)                        */
    GF_FIELD_U_STORE = 12,  /* u at the current RK stage, 3n               */
    GF_FIELD_STRAIN_DIAG = 13, /* dynamic strain exx, eyy, ezz (global), 3n, read only */
    GF_FIELD_STRAIN_OFF = 14,  /* dynamic strain eyz, exz, exy (global), 3n, read only */
    /* Yee fields on their staggered grids, concatenated x|y|z:
     *   E edges: Ex nx(ny+1)(nz+1), Ey (nx+1)ny(nz+1), Ez (nx+1)(ny+1)nz
     *   H faces: Hx (nx+1)ny nz,    Hy nx(ny+1)nz,     Hz nx ny(nz+1)
     * (EMdynamic_initialize.cpp:68-95); gf_field_size gives the count. */
    GF_FIELD_E_EM = 15,
    GF_FIELD_H_EM = 16,
    GF_FIELD_E_EM_STORE = 17,  /* E at the current RK stage              */
    GF_FIELD_H_EM_STORE = 18,
    GF_FIELD_E_CELL = 19,      /* cell-centred E of the stage store, 3n, read only */
    GF_FIELD_H_CELL = 20,      /* cell-centred H of the stage store, 3n, read only */
    GF_FIELD_JP = 21,          /* polarisation currents, 9n: poles 1..3, each x,y,z */
    GF_FIELD_HDRIVE = 22,      /* profile of the shaped drive, 3n (GF_FLAG_HDRIVE) */
    /* static strain in the crystal frame, exx eyy ezz | eyz exz exy, 3n each:
     * the reference's ex*t0_crt, which the magnets read under
     * if_elasto_on_pORm (magnetic_solver.cpp:89-96, magnetic_AFM_solver.cpp:
     * 89-96).  Used with GF_FLAG_ELASTO_ON_PM: alone when elastodynamics
     * is off, added to the dynamic strain when it is on (exx = exxt0_crt +
     * Dexx_crt); zero until uploaded. */
    GF_FIELD_STRAIN_T0_DIAG = 23,
    GF_FIELD_STRAIN_T0_OFF = 24,
    /* inverse spin Hall current Jx, Jy (and a zero z), 3n, read only
     * (GF_FLAG_SPIN_PUMPING; output_Jishe, inoutput.cpp:846-864) */
    GF_FIELD_J_ISHE = 25,
    /* elastic force per unit volume stored at the first RK stage, 3n (N/m^3):
     * the reference's force_*_store that the stiffness damping differences
     * against (elastic_dynamic_solver.cpp:1313-1323); read and written by
     * input_elastoforce / output_elastoforce (inoutput.cpp:220-237,
     * 742-761).  With GF_FLAG_ELASTODYNAMICS. */
    GF_FIELD_ELASTOFORCE = 26,
    /* static eigenstrain at t0 in the crystal frame, exx eyy ezz | eyz exz
     * exy, 3n each: the reference's ex*0t0_crt (elastic_system.h:154), read
     * by input_eigenstraint0 (inoutput.cpp:259-280).  The elastic stage
     * subtracts it from the eigenstrain of the m, m_AFM and p stores
     * (elastic_dynamic_solver.cpp:39-65), so the elastodynamics carries
     * the change from t0.  With GF_FLAG_ELASTODYNAMICS; zero until uploaded.
     * With GF_FLAG_ELASTOSTATIC each static solve writes it (the e0 of the
     * solve). */
    GF_FIELD_EIGENSTRAIN_T0_DIAG = 27,
    GF_FIELD_EIGENSTRAIN_T0_OFF = 28,
    GF_FIELD_COUNT
};

typedef struct gf_ctx gf_ctx;

/* Create a context on the current HIP device.  Copies the material table,
 * material_cell and drive table, builds interface masks and (when the FFT
 * demag is on) the Newell kernel spectra.  Replaces initialize_host +
 * copy_to_device + initialize_device of the magnetic system
 * (magnetic_initialize.cpp:4-116, 1072-1129, 1559-1788). */
gf_status gf_create(const gf_desc* desc, gf_ctx** out);
gf_status gf_destroy(gf_ctx* ctx);

/* Last error message of a context (or, when ctx is NULL, of the calling
 * thread's last failed gf_create or context-free call such as
 * gf_read_cells).  Never NULL. */
const char* gf_last_error(const gf_ctx* ctx);

/* Host <-> device copies of one field.  count is the number of doubles
 * (3*n for vector fields, see gf_field_size).  Uploading GF_FIELD_M also sets the stage store
 * to m, as initialize_device does (magnetic_initialize.cpp:1081-1097).
 * Both calls synchronise the context stream before returning. */
gf_status gf_upload(gf_ctx* ctx, int field_id, const double* host, size_t count);
gf_status gf_download(gf_ctx* ctx, int field_id, double* host, size_t count);

/* global_parameters::time_marching_device_serial (global.cpp:513-549):
 * t += dt and step the prescribed drive table. */
gf_status gf_time_advance(gf_ctx* ctx);

/* One fused FM RK4 stage s in 1..4: get_dm_RKs followed by update_m_RKs
 * (magnetic_solver.cpp:3-229, 285-1076 with 1449-1457 and 1515-1636).
 * Reads the stage-(s-1) store, writes the stage-s store (double-buffered,
 * so other physics of the same stage still see the old store). */
gf_status gf_mag_stage(gf_ctx* ctx, int stage);

/* One fused AFM RK4 stage (magnetic_AFM_solver.cpp:3-1834). */
gf_status gf_afm_stage(gf_ctx* ctx, int stage);

/* One fused ferroelectric RK4 stage: get_dq_RKs + get_dp_RKs followed by
 * update_q_RKs + update_p_RKs (ferroelectric_solver.cpp:4-131, 534-1784,
 * 1786-2029), with E_stat taken from the uploaded GF_FIELD_ESTAT. */
gf_status gf_fe_stage(gf_ctx* ctx, int stage);

/* One elastodynamics RK4 stage: get_dv_RKs + get_du_RKs followed by
 * update_v_RKs + update_u_RKs with the absorbing boundary and the strain
 * source (elastic_dynamic_solver.cpp:4-227, 841-3284;
 * elastodynamics_AbsorbingBoundary.cpp:3-230).  Its get phase reads the
 * stage-s stores of m, m_AFM and p (eigenstrain back-action), so within a
 * stage it must run before gf_mag_stage / gf_afm_stage / gf_fe_stage, which
 * then read the strain of the stage-s u store.  Grids where the reference's
 * forward-z wrap (k == ny-1, elastic_dynamic_solver.cpp:1247) reads past a
 * column or differs from a periodic wrap are rejected by gf_create. */
gf_status gf_elastic_stage(gf_ctx* ctx, int stage);

/* One electrodynamics RK4 stage: get_dE_RK + get_dH_RK (with the Jf source
 * and the Jp currents) followed by update_DE_RK + update_DH_RK with the
 * plane source, Mur walls, PEC edges and the cell-centred fields
 * (EMdynamic_solver.cpp:337-3476, 3477-4556; EMdynamic_AbsorbingBoundary.cpp:
 * 3-874).  Its get phase reads dp and dm of the same stage (ferroelectric
 * and magnetic get phases), so within a stage it runs after gf_fe_stage,
 * gf_mag_stage and gf_afm_stage, which read the cell fields of the
 * stage-s E and H stores. */
gf_status gf_em_stage(gf_ctx* ctx, int stage);

/* Replace the numeric parameters of material `index` (1-based, as in
 * material_cell) and rebuild its derived constants; the kind flags (if_FM,
 * if_AFM, if_FE, elastic, ...) must stay as they were.  Takes effect from
 * the next stage.  No counterpart in the reference (its materials are fixed
 * at read time); parameter sweeps and the adjoint's parameter updates use it. */
gf_status gf_set_material(gf_ctx* ctx, int index, const gf_material* material);

/* Number of doubles gf_upload / gf_download move for a field. */
gf_status gf_field_size(const gf_ctx* ctx, int field_id, size_t* count);

/* magnetic_system::get_H_static (magnetostatic_solver.cpp:114-219) or the
 * 1D model get_H_static_1D (:222-249) per the descriptor flags. */
gf_status gf_demag(gf_ctx* ctx);

/* update_prescribed_m at start-up (magnetic_initialize.cpp:1074-1077, the
 * body at magnetic_solver.cpp:1937-1957): the ferromagnetic cells of the
 * bottom layer take the prescribed cone at the context's time, in m and
 * in its stage store.  Needs GF_FLAG_PRESCRIBED_M and a step boundary;
 * call it after uploading m and before the initial gf_demag, as the
 * reference's start-up order does (main.cpp:113-143). */
gf_status gf_prescribe_m(gf_ctx* ctx);

/* Whole RK4 steps in the protocol of main.cpp:243-483 (time march, four
 * stages with statics between them, statics after the step). */
gf_status gf_step(gf_ctx* ctx, int64_t nsteps);

/* get_averagem / get_averagem_AFM (magnetic_solver.cpp:1863-1888) and
 * get_averagep (ferroelectric_solver.cpp:2031-2055): sums of the field over
 * all cells divided by the FM, AFM or FE cell count. */
gf_status gf_average(gf_ctx* ctx, int field_id, double out[3]);

/* No counterpart in the reference: for callers that cannot allocate device
 * memory or wait on the host (code under HIP graph capture, serving loops
 * that forbid hipMalloc).  The average of gf_average, bit for bit, written
 * to out_dev (device memory, 3 doubles) on the context stream; it returns
 * without waiting (gf_sync or later work on that stream orders the read).
 * The partial sums live in a fixed 24 KiB device array per GPU, resolved
 * on the first call on that GPU and held until the process ends, so
 * nothing is allocated; at most 1024 partial sums per component (every
 * context fits).  Calls that overlap on one GPU (different streams or
 * threads) race on that array; calls ordered on one stream do not. */
gf_status gf_average_noalloc(gf_ctx* ctx, int field_id, double* out_dev);

/* Current simulated time (global_parameters::time_device). */
gf_status gf_get_time(const gf_ctx* ctx, double* t);

/* Set the simulated time, e.g. to restart from a saved state at its time
 * (the drive tables restart from their first row).  At a step boundary. */
gf_status gf_set_time(gf_ctx* ctx, double t);

/* Wait for all work queued on the context stream. */
gf_status gf_sync(gf_ctx* ctx);

/* The context's HIP stream (a hipStream_t), for callers that order their
 * own work (events, copies) with the context's kernels. */
gf_status gf_get_stream(gf_ctx* ctx, void** stream);

/* Per-kernel timing.  When enabled, every launch of the listed kernel
 * classes is bracketed by HIP events on the context stream; gf_kernel_times
 * returns accumulated milliseconds and launch counts per class. */
enum gf_kernel_class {
    GF_K_MAG_STAGE = 0,
    GF_K_AFM_STAGE = 1,
    GF_K_FFT_Z_FWD = 2,
    GF_K_FFT_Y_FWD = 3,
    GF_K_FFT_X_CONV = 4,
    GF_K_FFT_Y_INV = 5,
    GF_K_FFT_Z_INV = 6,
    GF_K_OTHER = 7,
    GF_K_FE_STAGE = 8,
    GF_K_ELASTIC_STAGE = 9,   /* stress divergence + v, u updates */
    GF_K_ELASTIC_STRAIN = 10, /* face gradients, strain, eigenstrain */
    GF_K_EM_E = 11,           /* Ampere update of the E edges and Faraday update of the H faces (one launch) */
    GF_K_EM_H = 12,           /* no longer timed apart: the Faraday update runs in the GF_K_EM_E launch */
    GF_K_COUNT = 13
};
gf_status gf_set_kernel_timing(gf_ctx* ctx, int enable);
gf_status gf_kernel_times(gf_ctx* ctx, double ms[GF_K_COUNT], int64_t launches[GF_K_COUNT]);

/* ---- Adjoint sweep ----------------------------------------------------
 * No counterpart in the reference.  Gradients of a probe-region spectral
 * loss through the discrete RK4 trajectory of an FM context:
 *   X_n = mean of m_x over the probe box after step n (n = 1..N),
 *   S   = sum_n X_n exp(-i 2 pi f t_n) dt,   L = |S|^2,
 * with respect to the damping alpha, the exchange constant Aex and B1 of
 * one FM material and to the shaped-drive profile (GF_FIELD_HDRIVE).  The
 * forward pass keeps a checkpoint of m every checkpoint_every steps; the
 * backward pass recomputes one segment at a time (keeping its stage stores
 * and fields) and runs the adjoint of each stage: transposed exchange
 * stencil, adjoint FFT convolution (conjugate-transposed tensor spectra),
 * adjoint LLG cross products, anisotropy and magnetoelastic Jacobians.
 * Supported contexts: FM only (several FM materials; the gradients are
 * one material's), interfacial DMI with its surface slopes, FFT or no
 * magnetostatics, uniform/shaped sinusoidal drives, optional static strain
 * and rotation; no coupled physics or drive tables. */
typedef struct gf_adjoint_desc {
    uint32_t struct_size;      /* sizeof(gf_adjoint_desc) */
    int32_t material;          /* 1-based FM material whose alpha, Aex, B1 are differentiated */
    int32_t probe_lo[3];       /* probe box, inclusive cell indices */
    int32_t probe_hi[3];
    double probe_freq;         /* Hz */
    int64_t nsteps;            /* N */
    int64_t checkpoint_every;  /* steps per checkpoint of the checkpointed steps; 0 = ceil(sqrt) of their count */
    int32_t precision;         /* 64 (or 0): FP64 sweep; 32: FP32 variant (see gf_adjoint_forward) */
    int32_t resident_mib;      /* FP64 sweep: device memory (MiB) for the stage stores and fields of the
                                * last steps, kept by the forward pass so that the reverse sweep does
                                * not recompute them (8 x 3n doubles a step with magnetostatics, 4
                                * without); 0 = checkpoint every step */
} gf_adjoint_desc;

/* Run N steps from the current state (the magnetostatic field is rebuilt
 * from it first), keep the checkpoints and return L.  Leaves the context at
 * the final state, at a step boundary.  With precision 32 the trajectory,
 * checkpoints and the reverse sweep run in FP32 on copies of the state
 * (frame rotation and Hext_nonunif unsupported), and the context's own
 * state and time are left unchanged. */
gf_status gf_adjoint_forward(gf_ctx* ctx, const gf_adjoint_desc* desc, double* loss);

/* Reverse sweep of the last gf_adjoint_forward: grad = dL/d(alpha, Aex, B1)
 * of the chosen material; grad_hdrive (3n doubles, may be NULL) = dL/dHdrive.
 * Leaves the context at the forward's final state. */
gf_status gf_adjoint_backward(gf_ctx* ctx, double grad[3], double* grad_hdrive, size_t count);

/* The adjoint of one stage increment k = f(store) (magnetic_solver.cpp:
 * 3-229 differentiated) at the stage time of `stage` for the context's
 * current time: ybar = (dk/dstore)^T kbar including the magnetostatic field
 * of the store, grad += (dk/d(alpha, Aex, B1))^T kbar, grad_hdrive (may be
 * NULL) += (dk/dHdrive)^T kbar.  Host arrays of 3n doubles; the context's
 * state is not changed. */
gf_status gf_adjoint_stage(gf_ctx* ctx, int material, int stage, const double* store, const double* kbar,
                           double* ybar, double grad[3], double* grad_hdrive, size_t count);

/* ---- Batched sweeps ---------------------------------------------------
 * No counterpart in the reference, which runs one simulation per process.
 * A batch steps many independent contexts of uneven grid sizes together:
 * one launch per RK4 stage and magnet kind covers every instance, and a
 * whole RK4 step (per-instance clock and drive, four stages) is replayed
 * from a HIP graph.  Contexts keep their own state and stay usable through
 * the single-context calls between batch steps (gf_download, gf_average,
 * gf_step).  Batchable contexts hold FM and/or AFM material only, with
 * sinusoidal or no drive, optional Hext_nonunif and rotation, and no
 * magnetostatics, 1D model, drive tables or coupled physics; their Hstat
 * stays as uploaded. */
typedef struct gf_batch gf_batch;

/* Build a batch over count contexts created on the current device.  The
 * contexts are borrowed: destroy the batch before them. */
gf_status gf_batch_create(gf_ctx* const* ctxs, int count, gf_batch** out);
gf_status gf_batch_destroy(gf_batch* batch);

/* Last error of a batch (or of the last failed gf_batch_create when batch
 * is NULL).  Never NULL. */
const char* gf_batch_last_error(const gf_batch* batch);

/* nsteps RK4 steps of every instance (gf_time_advance and the four magnet
 * stages, per instance).  Ordered after work queued on every context
 * stream; every context stream is ordered after it.  The stage arguments
 * are rebuilt from the contexts at each call, so uploads made between
 * calls (Hstat, Hext_nonunif, a stage store, a loaded state image) take
 * effect as they would for gf_step.  Waits for the previous call's queued
 * work before it starts; returns after queueing the steps. */
gf_status gf_batch_step(gf_batch* batch, int64_t nsteps);

/* Wait for the batch's queued work. */
gf_status gf_batch_sync(gf_batch* batch);

/* The batch's HIP stream (a hipStream_t). */
gf_status gf_batch_get_stream(gf_batch* batch, void** stream);

/* Launch shape of the batched stage kernels: workgroup size (64, 128 or
 * 256), cells per thread (1..16: a workgroup slot covers threads *
 * cells_per_thread cells of one instance) and grid cap (0 = one workgroup
 * per slot; otherwise at most grid_cap workgroups, each looping over the
 * slots).  Changing it re-captures the step graph. */
gf_status gf_batch_config(const gf_batch* batch, int* threads, int* cells_per_thread, int* grid_cap);
gf_status gf_batch_set_config(gf_batch* batch, int threads, int cells_per_thread, int grid_cap);

/* Autotuner: times stage 1 of every (workgroup size, cells per thread,
 * grid cap) candidate reps times on the live instances (leaving the state
 * as it was) and keeps the fastest.  best_ms (may be NULL) receives its
 * time per stage. */
gf_status gf_batch_tune(gf_batch* batch, int reps, double* best_ms);

/* ---- restart and the reference's text files ------------------------------
 *
 * State image: everything a step boundary carries into the next step, at
 * full precision -- t, the drive-table cursors, m / m_AFM / p / q / u / v /
 * E / H / Jp and any uploaded stage stores, the elastic force store and
 * strains, the Liao or Mur boundary history, the Mur planes of the Yee
 * field, Hstat (or the H z-spectrum of the fused magnetostatics), the
 * external-field and strain inputs, and the material table.  The reference
 * restarts from its text inputs instead (input_m ... input_Jp,
 * inoutput.cpp:11-327, called at main.cpp:44-89): seven digits, t = 0, no
 * force store or boundary history.  An image loads only into a context
 * created from the same grid, cell sizes, dt, flags and material ids, at a
 * step boundary; a material record that differs is applied as
 * gf_set_material would.  Layout: a header, the material records, then per
 * array a (tag, count) pair and the doubles, then a 64-bit FNV-1a checksum
 * of everything before it (little-endian, this library's build only). */
gf_status gf_state_size(gf_ctx* ctx, size_t* bytes);
gf_status gf_state_save(gf_ctx* ctx, void* host, size_t bytes);      /* bytes must equal gf_state_size */
gf_status gf_state_load(gf_ctx* ctx, const void* host, size_t bytes);
gf_status gf_state_write(gf_ctx* ctx, const char* path);             /* the image as a file */
gf_status gf_state_read(gf_ctx* ctx, const char* path);

/* Per-cell text files of the reference (output_m and its kind,
 * inoutput.cpp:330-864; input_m and its kind, :11-327): one line per cell,
 * "i  j  k" 1-based with k fastest, then ncomp values "%.7e" two spaces
 * apart.  data holds ncomp arrays of dims[0]*dims[1]*dims[2] doubles, cell
 * (i, j, k) at (i*dims[1] + j)*dims[2] + k; a Yee component passes its own
 * staggered dims.  Reading accepts the records in any order but, unlike the
 * reference (which reads n records blindly and keeps the old value of a
 * cell it never saw), requires every cell exactly once and nothing after
 * them.  No context: gf_last_error(NULL) holds the reason of a failure. */
gf_status gf_write_cells(const char* path, const int64_t dims[3], int32_t ncomp, const double* data);
gf_status gf_read_cells(const char* path, const int64_t dims[3], int32_t ncomp, double* data);

/* Diagnostic, with no counterpart in the reference.  Every stage kernel
 * issues all of a cell's loads before the cell's code is known, so no array
 * a launch writes may be read by it at neighbouring cells or as another
 * quantity; gf_mag_stage, gf_afm_stage and gf_fe_stage refuse arguments
 * that break this (GF_ERR_STATE).  This call runs that check over the
 * stage arguments of every combination of physics, couplings, optional
 * fields, stage and stage-1 store, on host placeholder contexts (no device
 * work).  plant = 0 returns GF_OK when nothing aliases; plant = 1..4
 * plants one alias (an FE output on a q store, an FM output on its stencil
 * store, an AFM accumulator on the other sublattice's state, two FE
 * accumulators on one array) and returns GF_OK when it is reported.  msg
 * (cap bytes) receives a summary or the first conflict. */
gf_status gf_selftest_stage_io(int plant, char* msg, size_t cap);

#ifdef __cplusplus
}
#endif
#endif /* GOFERRO_H */
