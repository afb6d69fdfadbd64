// gf_internal.h — shared definitions of the gfx950 hot-path library.
//
// Device data layout (all FP64, SoA, z fastest, index (i*ny + j)*nz + k):
//   m[3], acc[3], store_a[3], store_b[3]   FM state, RK accumulator, stage stores
//   m1/m2 + their acc/stores                AFM sublattices
//   hstat[3]                                magnetostatic field
//   code[n] (uint32)                        material id + interface bits (below)
//   spec_a/spec_b                           complex FFT work buffers, 3 components
//   kre[9], kim[9]                          demag tensor spectra (FP64 real, FP32 imaginary)
// A cell code packs what the reference keeps in material_cell plus the six
// FM_surf* / AFM_surf* face masks (magnetic_initialize.cpp:136-379, 635-870).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/goferro.h"

namespace gf {

constexpr double kMu0 = 1.256637061e-6;               // magnetic_system.h:14
constexpr double kPi = 3.14159265358979323846;        // magnetic_system.h:16
// RK4 weights 1/6 and 1/3 as rounded reciprocals: the stage updates multiply
// by them instead of dividing (an FP64 division is about ten instructions);
// each product is within about an ulp of the quotient.
constexpr double kSixth = 1. / 6.;
constexpr double kThird = 1. / 3.;
constexpr double kUnitE = 1.602176634e-19;          // magnetic_system.h:15

// cell code bits
constexpr uint32_t kMatMask = 0xFFu;   // material id, up to 255 materials
constexpr uint32_t kIsFM = 1u << 8;
constexpr uint32_t kIsAFM = 1u << 9;
constexpr uint32_t kIsFE = 1u << 10;
constexpr int kFmSurfShift = 11;    // 6 bits: xb, xf, yb, yf, zb, zf
constexpr int kAfmSurfShift = 17;   // 6 bits, same order
constexpr int kFeSurfShift = 23;    // 6 bits, same order (FE_surf*, ferroelectric_initialize.cpp:74-308)
enum Face { kXB = 0, kXF = 1, kYB = 2, kYF = 3, kZB = 4, kZF = 5 };

// Per-material constants as the stage kernels need them.  Every entry is
// the value the reference computes per cell from the material record, in
// the same operation order, so hoisting it to set-up changes no bit.
struct DevMaterial {
    // FM (magnetic_solver.cpp:84-87, 164-220)
    double Ms;
    double K1n, K2n;        // K / mu0 / Ms
    double B1n, B2n;        // B / mu0 / Ms
    double Aex;
    double Aex_self;        // 2*Aex*Aex/(Aex+Aex): harmonic mean with itself
    double dmi_slope;       // iDMI / 2. / Aex
    double exch;            // 2. / mu0 / Ms
    double dmi_pos;         // 2. * iDMI / mu0 / Ms
    double dmi_neg;         // -2. * iDMI / mu0 / Ms
    double precess;         // -gyro / (1 + alpha^2) * dt
    double alpha_precess;   // alpha * precess
    int32_t anis_type;
    int32_t is_fm;
    // AFM (magnetic_AFM_solver.cpp:108-200), index 0/1 = sublattice 1/2
    double aMs[2];
    double aK1n[2];
    double aB1n[2], aB2n[2];
    double aJ[2];           // -1. * J_AFM / mu0 / Ms_sub
    double aAex[2];
    double aAex_self[2];
    double aExch[2];        // 2. / mu0 / Ms_sub
    double aPrecess[2];
    double aAlphaPrecess[2];
    double u[3];
    int32_t is_afm;
    int32_t is_fe;
    // FE (ferroelectric_solver.cpp:740-792)
    double fe_mass;         // -1. * dt / FE_mass
    double fe_damping;      // fe_mass * FE_damping
    double a1, a11, a12, a111, a112, a123, a1111, a1112, a1122, a1123;
    double G11;
    double tv1, tv2, tv3;
    // elastodynamics (elastic_dynamic_solver.cpp:1306-1329) and the
    // eigenstrain sources (:4-104)
    double cij[6][6];       // cijkl_glb, Voigt xx yy zz yz xz xy
    double el_rho;          // dt / density
    double el_mass;         // -dt * elast_mass_damping
    double el_stiff;        // elast_stiff_damping / density
    double Q11, Q12, Q44;
    double l100, l111, l100a, l111a, l100b, l111b;
    // electrodynamics (EMdynamic_solver.cpp:392-499, 4318-4451)
    double em_er;           // r_permittivity
    double em_cond;         // conductivity
    double cn[3];           // comp_n1..3
    double wp[3];           // omega_plasma_n1..3
    double re[3];           // 1 / tao_e_n1..3
};

constexpr double kEps0 = 8.854187817e-12;     // EMdynamic_system.h:13
constexpr double kCLight = 2.99792458e8;      // EMdynamic_system.h:14

// Yee grids: E on edges, Ex nx(ny+1)(nz+1), Ey (nx+1)ny(nz+1), Ez
// (nx+1)(ny+1)nz; H on faces, Hx (nx+1)ny nz, Hy nx(ny+1)nz, Hz nx ny(nz+1);
// each z fastest in its own dimensions.
// A double moved one lane up (lane i gets lane i - 1) or down (lane i
// gets lane i + 1) across the 64-lane wave, with DPP wave shifts.
__device__ __forceinline__ double lane_up(double v) {
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), 0x138, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), 0x138, 0xf, 0xf, false);
    return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double lane_down(double v) {
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), 0x130, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), 0x130, 0xf, 0xf, false);
    return __hiloint2double(hi, lo);
}
__device__ __forceinline__ uint32_t lane_up(uint32_t v) {
    return static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0x138, 0xf, 0xf, false));
}
__device__ __forceinline__ uint32_t lane_down(uint32_t v) {
    return static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0x130, 0xf, 0xf, false));
}

struct EmArgs {
    const uint32_t* code;
    const DevMaterial* mats;
    const double* es[3];        // E store of the stage
    const double* hs[3];        // H store of the stage
    double* eb[3];              // E base, E out store, E accumulator
    double* eo[3];
    double* eacc[3];
    double* hb[3];
    double* ho[3];
    double* hacc[3];
    const double* jf_spatial;   // free current in its box (update_Jf_input*): the per-plane factor
    double jf_value;            //   (types 1, 2) or null, and the value at the stage time
    int32_t jf_comp;            //   of the driven component, -1 when there is none
    int32_t jf_lo[3], jf_hi[3];
    const double* jps[9];       // Jp stores (pole-major), null when identically zero
    const double* q[3];         // ferroelectric q store of the stage (dp = q dt), null when uncoupled
    const double* dm[3];        // magnetic increments of the stage, null when absent
    const double* dma1[3];
    const double* dma2[3];
    const double* jishe[2];     // inverse spin Hall current Jx, Jy, null when off
    double* t1[3];              // stage 4 with Mur walls: E of the planes the walls read, kept before the combine
    int64_t nx, ny, nz;
    double dx, dy, dz, dt;
    int32_t per[3];
    int32_t stage;
};

// get_J_ISHE_RK* (magnetic_solver.cpp:231-283): per pumping plane k, the
// current from the stage store and increment of its source plane src[k].
// With the accumulators set it is get_J_ISHE (:1460-1511) at the step
// boundary instead: the increment is the RK combination acc + k4/6 and the
// states are those before the update.
struct SpinArgs {
    const double* fs[3];        // FM store and increment of the stage, null when no FM
    const double* fd[3];
    const double* s1[3];        // AFM sublattices, null when no AFM
    const double* d1[3];
    const double* s2[3];
    const double* d2[3];
    const double* fa[3];        // RK accumulators after stage 3, null for a stage current
    const double* a1[3];
    const double* a2[3];
    const int32_t* src;         // source plane per plane, -1 = none
    const double* coef;         // theta_SH e / (2 pi) g_mix * profile, per plane
    double* j[2];
    int64_t nx, ny, nz;
    double dt;
};

struct EmJpArgs {
    const uint32_t* code;
    const DevMaterial* mats;
    const double* ecell[3];
    double* jp[9];
    double* jps[9];
    double* jacc[9];
    int64_t n;
    double dt;
    int32_t stage;
};

struct EmWallArgs {
    double* b[3];               // E base
    double* s[3];               // E out store
    const double* t1[3];        // E base before the final combine
    int64_t nx, ny, nz;
    int32_t axis;               // walls normal to x, y or z
    int32_t kind;               // 0 = stages 1-2, 1 = stage 3, 2 = after the step
    int32_t pec;                // the walls are PEC
    int32_t onlytop;            // bottom z wall held at zero
    double cf;                  // x / y coefficient, or the top one for z
    double cb;                  // bottom coefficient for z
};

struct EmCellArgs {
    const uint32_t* code;
    const DevMaterial* mats;
    const double* es[3];
    const double* hs[3];
    double* ecell[3];
    double* hcell[3];
    const double* mz;           // 1D model: FM and AFM mz stores and Hz_stat, else null
    const double* mz1;
    const double* mz2;
    const double* hzstat;
    int64_t nx, ny, nz;
    int32_t do_e, do_h, one_d;
    // the z walls (axis 2) applied by the cells launch to the edges it
    // reads, when the x and y walls are periodic: set zwall
    int32_t zwall;
    EmWallArgs w;
};


struct EmMiscArgs {
    double* b[3];
    double* s[3];
    const double* src[3];       // history copy source (E base)
    double* t1[3];
    const uint8_t* pec[3];
    double plane[3];
    int64_t nx, ny, nz;
    int32_t per[3];
    int32_t plane_z;
    int32_t kind;
};

// Elastodynamics.  Face gradients live on the staggered face grids
// (gyz: (nx+1) x ny x nz, gxz: nx x (ny+1) x nz, gxy: nx x ny x (nz+1));
// strain and eigenstrain are six cell arrays in Voigt order.
struct ElStrainArgs {
    const uint32_t* code;
    const double* us[3];
    double* gyz[3];
    double* gxz[3];
    double* gxy[3];
    double* eg[6];
    double* ec[6];          // may equal eg (no rotation)
    int64_t nx, ny, nz;
    double dx, dy, dz;
    double idx, idy, idz;   // 1/dx, 1/dy, 1/dz
    int32_t per[3];
    int32_t rotate;
    double Tg2c[3][3];
};

// Flexoelectric constants of one material (read.cpp:181-199; zero unless FE).
struct FlexoMat {
    double F11, F12, F44, tv4, tv5, tv6, tv7, tv8, f11, f12, f44;
};

// get_p_gradient (ferroelectric_solver.cpp:134-281) at an FE cell of code
// `own`: (forward - backward) / h / 2 of the p store, the cell's own value
// standing in across an FE surface face (also halved there).  g[comp][axis].
__device__ __forceinline__ void fe_p_gradient(uint32_t own, const double* const* p, int64_t id, int64_t i, int64_t j,
                                              int64_t k, int64_t nx, int64_t ny, int64_t nz, const double h[3],
                                              double g[3][3]) {
    const int64_t nb[6] = {((i == 0 ? nx - 1 : i - 1) * ny + j) * nz + k, ((i == nx - 1 ? 0 : i + 1) * ny + j) * nz + k,
                           (i * ny + (j == 0 ? ny - 1 : j - 1)) * nz + k, (i * ny + (j == ny - 1 ? 0 : j + 1)) * nz + k,
                           (i * ny + j) * nz + (k == 0 ? nz - 1 : k - 1), (i * ny + j) * nz + (k == nz - 1 ? 0 : k + 1)};
    for (int ax = 0; ax < 3; ax++) {
        const bool sb = (own >> (kFeSurfShift + 2 * ax)) & 1u;
        const bool sf = (own >> (kFeSurfShift + 2 * ax + 1)) & 1u;
        for (int c = 0; c < 3; c++) {
            const double bw = sb ? p[c][id] : p[c][nb[2 * ax]];
            const double fw = sf ? p[c][id] : p[c][nb[2 * ax + 1]];
            g[c][ax] = (fw - bw) / h[ax] / 2.;
        }
    }
}

struct ElEigenArgs {
    const uint32_t* code;
    const DevMaterial* mats;
    const double* m[3];     // FM store, null when there is no FM
    const double* m1[3];    // AFM stores, null when there is no AFM
    const double* m2[3];
    const double* p[3];     // FE p store, null when there is no FE
    double* e0g[6];
    double* e0c[6];         // may equal e0g (no rotation)
    const double* e0t0[6];  // static eigenstrain at t0 (crystal frame), null when not in use
    const FlexoMat* flexo;  // flexoelectric -F grad p of the FE cells, null when off
    int64_t n;
    int64_t nx, ny, nz;
    double dx, dy, dz;
    int32_t rotate;
    double Tg2c[3][3];
    double Tc2g[3][3];
};

// The v and u RK updates of an elastic stage.  us is the u store the stage
// writes: a different array from the store its strain is read from, since
// a cell's stress reads the store around it up to two cells away.
struct ElForceArgs {
    const uint32_t* code;
    const DevMaterial* mats;
    double* u[3];
    double* v[3];
    double* us[3];
    double* vs[3];
    double* accu[3];
    double* accv[3];
    double* fstore[3];
    int64_t nx, ny, nz;
    double dx, dy, dz, dt;
    double idx, idy, idz;   // 1/dx, 1/dy, 1/dz
    int32_t per[3];
    int32_t stage;
    int32_t trac_k;         // plane of a traction pulse, -1 when off
    int32_t trac_row;       // 0, 1, 2 for a traction along x, y, z (stress xz, yz, zz)
    double trac_amp;
    int32_t nmat;           // material records (their stiffness matrices are staged in LDS)
};

// One elastic stage (two launches): the face gradients and strain of the
// u store s.us (written out: the strain for the magnets and ferroelectrics,
// the face gradients only when s.gyz is set, for the flexoelectric force),
// the eigenstrain of the stage stores (e, when eig is set; zero otherwise;
// written out when e0_out is set) and each cell's part of the stress its
// faces carry, P^x, P^y, P^z: the rows of C (the cell's stiffness) applied
// to the strain terms that do not depend on which of its two faces along
// the axis is evaluated; then the face stresses and the updates (f).
struct ElStageArgs {
    ElStrainArgs s;
    ElEigenArgs e;
    ElForceArgs f;
    double* px[3];
    double* py[3];
    double* pz[3];
    int32_t eig;
    int32_t e0_out;
};

// Displacement history of the absorbing boundary: slot = [plane][comp][nx*ny]
// with the four planes k = 1..4 (Liao) or k = 0, 1, nz-2, nz-1 (Mur).
constexpr int kHistPlanes = 4;

struct ElHistArgs {
    const double* u[3];
    double* slot;
    int64_t nx, ny, nz;
    int32_t planes[kHistPlanes];
};

struct ElBoundaryArgs {
    const uint32_t* code;
    double* u[3];
    double* us[3];
    const double* t[4];     // slots of t1..t4 (Mur reads t1 only)
    int64_t nx, ny, nz;
    int32_t kind;           // 0 = stages 1-2, 1 = stage 3, 2 = after the step
    int32_t liao;           // 1 = Liao at k = 0, 0 = Mur
    int32_t mur_bottom;     // Mur also at k = 0
    double w;               // Liao weight of the fourth-order term
    double cb_l, cb_t, ct_l, ct_t;   // Mur coefficients, bottom/top, long/trans
    int32_t src_k;          // source plane written this call, -1 when none
    int32_t src_comp;
    double src_val;
};

// The flexoelectric force of one stage at the FE cells
// (ferroelectric_solver.cpp:800-832): p store, its gradients and second
// derivatives, and the gradients of the dynamic strain
// (elastic_dynamic_solver.cpp:230-399: strain eg and face gradients gyz,
// gxz, gxy of the stage, null without elastodynamics).
struct FeFlexoArgs {
    const uint32_t* code;
    const FlexoMat* fm;
    const double* p[3];
    const double* eg[6];
    const double* gyz[3];
    const double* gxz[3];
    const double* gxy[3];
    double* out[3];
    int64_t nx, ny, nz;
    double dx, dy, dz;
    int32_t per[3];
};

struct FeStageArgs {
    const uint32_t* code;
    const DevMaterial* mats;
    const double* pstore[3];
    const double* qstore[3];
    const double* pbase[3];
    const double* qbase[3];
    double* pacc[3];
    double* qacc[3];
    double* pout[3];
    double* qout[3];
    const double* estat[3];
    const double* strain[6];    // dynamic strain (crystal frame), null when uncoupled
    const double* eigen[6];     // eigenstrain (crystal frame), with strain
    const double* strain0[6];   // static strain added to strain, or null
    const double* eigen0[6];    // static eigenstrain added to eigen, with strain0
    const double* ecell[3];     // electrodynamic E at the cells, null when uncoupled
    const double* flexo[3];     // flexoelectric force of the stage, null when off
    int64_t nx, ny, nz;
    double dx, dy, dz, dt;
    double eext[3];
    double Tg2c[3][3];
    double Tc2g[3][3];
    int32_t rotate;
    int32_t stage;
};

// Everything a stage kernel reads, passed by value as kernel arguments.
struct StageArgs {
    const uint32_t* code;
    const DevMaterial* mats;
    const double* store[3];     // stage store read by the stencil
    const double* base[3];      // m at the start of the step
    double* acc[3];             // RK accumulator
    double* out[3];             // next store, or m at stage 4
    const double* hstat[3];
    const double* hnu[3];       // may be null
    const double* hdrive[3];    // shaped drive profile, null when off
    const double* strain[6];    // dynamic strain (crystal frame), null when uncoupled
    const double* strain0[6];   // static strain added to it (crystal frame), null when none
    const double* hcell[3];     // electrodynamic H at the cells, null when uncoupled
    double* dm_out[3];          // k of this stage, null unless Faraday's law needs it
    int64_t nx, ny, nz;
    double dx, dy, dz;
    double hext[3];
    double drive_s;             // sin(2 pi f t) of the shaped drive at the stage time
    double Tg2c[3][3];
    double Tc2g[3][3];
    int32_t rotate;
    int32_t stage;
    int32_t nmat;               // material records at mats
    int32_t prescribe;          // the FM cells at k = 0 take presc_m as their next store
    double presc_m[3];
};

struct AfmStageArgs {
    const uint32_t* code;
    const DevMaterial* mats;
    const double* store1[3];
    const double* store2[3];
    const double* base1[3];
    const double* base2[3];
    double* acc1[3];
    double* acc2[3];
    double* out1[3];
    double* out2[3];
    const double* hstat[3];
    const double* strain[6];    // dynamic strain (crystal frame), null when uncoupled
    const double* strain0[6];   // static strain added to it (crystal frame), null when none
    const double* hcell[3];     // electrodynamic H at the cells, null when uncoupled
    double* dm_out1[3];         // k of this stage per sublattice, null unless needed
    double* dm_out2[3];
    int64_t nx, ny, nz;
    double dx, dy, dz;
    double hext[3];
    double Tg2c[3][3];
    double Tc2g[3][3];
    int32_t rotate;
    int32_t stage;
};

// Launch shape of a batched stage: slot -> instance table, first slot of
// each instance, slot count, grid, workgroup size, cells per thread.
struct BatchLaunch {
    const uint32_t* item;
    const uint32_t* block0;
    uint32_t total;
    unsigned grid, threads;
    int cpt;
    bool narrow;
};

struct AdjState;
struct Adj32;

// Clock and sinusoidal drive of one instance of a batch.
struct BatchClock {
    double t, dt;
    double stat[3], altr[3], freq[3];
};

// Twiddle table of one transform length: w[j] = exp(-2 pi i j / N), j < N/2.
struct Twiddles {
    int64_t n = 0;
    double2* dev = nullptr;
};

// Radices of a mixed-radix line transform of length n (their product), for
// the passes of grids that are not powers of two.
constexpr int kMaxRadices = 16;
struct FftFactors {
    int n = 0;
    int count = 0;
    int r[kMaxRadices] = {};
};

struct DemagPlan {
    bool ready = false;
    // Any of nx, ny, nz not a power of two: the passes use the mixed-radix
    // kernels (gf_fft_any.hip) with natural-order twiddle tables of n
    // entries, w[t] = exp(-2 pi i t / n).
    bool general = false;
    FftFactors fx, fy, fz;
    Twiddles gtx, gty, gtz;
    int64_t nzh = 0;                 // nz/2 + 1
    double2* spec_m = nullptr;       // M z-spectrum, 3 components, layout [x][kz][y]
    double2* spec_a = nullptr;       // H z-spectrum, 3 components, layout [x][kz][y]
    double2* spec_b = nullptr;       // 3 components, layout [kz][y][x]
    double* kre = nullptr;           // 9 tensor spectra, real parts, layout [kz][y][x]
    float* kim = nullptr;            // their imaginary parts (round-off residue)
    Twiddles twx, twy, twz;
    // Fused mode (FM-only magnetisation): the FM stage consumes the H
    // z-spectrum and produces the M z-spectrum itself.
    bool fused = false;
    int P = 0;                       // z columns per fused workgroup
    bool h_in_spec = false;          // spec_a holds the current H
    bool h_real_valid = true;        // hstat holds the current H
    bool m_in_spec = false;          // spec_m holds Ms m of the current FM store
};

struct FusedZArgs {
    const double2* hspec;
    double2* mspec;
    const double2* tw;
    int64_t nh;
    int nzh, lognz, P;
};

// Arguments of the mixed-radix passes (gf_fft_any.hip); each pass reads
// the fields it needs.
enum class AnyPass { kZFwd, kYFwd, kYInv, kXConv, kZInv };
struct AnyPassArgs {
    const uint32_t* code;
    const DevMaterial* mats;
    const double* fm[3];
    const double* a1[3];
    const double* a2[3];
    const double* raw[3];       // raw_in: plain real input arrays
    double* out[3];             // z inverse output
    double2* spec;              // [x][kz][y], component c at spec + c*nh
    double2* spec2;             // [kz][y][x], plane kz at kz - kzb0, component c at spec2 + c*nhb
    const double* kre;          // tensor spectra (real, FP32 imaginary), [kz][y][x]
    const float* kim;
    const double2* tw;          // twiddle table of the pass's length
    FftFactors f;               // the pass's transform length and radices
    int64_t nh, nhb;
    double den;                 // -dx dy dz
    int nx, ny, nzh;
    int pencils;
    int kz0, kzb0;
    int raw_in, conv, adjoint;
};
size_t any_pass_lds(int lines, int n);
hipError_t launch_any_pass(AnyPass kind, const AnyPassArgs& a, dim3 grid, hipStream_t s);

struct TimerSlot {
    hipEvent_t start, stop;
    int cls;
};

}  // namespace gf

struct gf_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;

    int64_t nx = 0, ny = 0, nz = 0, n = 0;
    double dx = 0, dy = 0, dz = 0;
    int per[3] = {0, 0, 0};
    uint32_t flags = 0;
    double dt = 0, time = 0;
    double Tg2c[3][3] = {}, Tc2g[3][3] = {};
    double hext_stat[3] = {}, hext_altr[3] = {}, h_altr_freq[3] = {}, hext[3] = {};
    double demag_fac[3] = {};
    std::vector<double> prescribe;   // rows of (t, Hx, Hy, Hz)
    int prescribe_idx = 0;
    double eext_stat[3] = {}, eext_altr[3] = {}, e_altr_freq[3] = {}, eext[3] = {};
    std::vector<double> prescribe_e; // rows of (t, Ex, Ey, Ez)
    int prescribe_e_idx = 0;

    std::vector<gf_material> host_mats;
    std::vector<uint32_t> host_code;
    bool any_fm = false, any_afm = false, any_fe = false;
    int64_t nfm = 0, nafm = 0, nfe = 0;
    bool has_hnu = false;
    bool has_e0t0 = false;          // the static eigenstrain was written: the elastic stage subtracts it
    bool has_et0_dyn = false;       // a static strain was uploaded beside the elastodynamics: the magnets add it
    bool t0_fixed = false;          // the static strain at t0 was solved for, uploaded or loaded

    uint32_t* d_code = nullptr;
    uint32_t* d_fm_ids = nullptr;   // FM cells in index order, when not every cell is FM
    int64_t n_fm_ids = 0;
    gf::DevMaterial* d_mats = nullptr;

    // FM
    double* m[3] = {};
    double* acc[3] = {};
    double* st_a[3] = {};
    double* st_b[3] = {};
    double* dm_last[3] = {};
    // AFM
    double* m1[3] = {};
    double* m2[3] = {};
    double* acc1[3] = {};
    double* acc2[3] = {};
    double* st1_a[3] = {};
    double* st1_b[3] = {};
    double* st2_a[3] = {};
    double* st2_b[3] = {};
    // FE: p, q, their accumulators and two stage stores each
    double* p[3] = {};
    double* q[3] = {};
    double* pacc[3] = {};
    double* qacc[3] = {};
    double* p_a[3] = {};
    double* p_b[3] = {};
    double* q_a[3] = {};
    double* q_b[3] = {};
    double* estat[3] = {};
    // elastodynamics: state, single stage stores (each cell's store is read
    // only by its own thread within a stage), RK accumulators, the force of
    // stage 1, face gradients, strain and eigenstrain (crystal-frame arrays
    // alias the global ones without a rotation), boundary history ring
    bool el_on = false;
    double* u[3] = {};
    double* v[3] = {};
    double* us[3] = {};
    double* vs[3] = {};
    double* accu[3] = {};
    double* accv[3] = {};
    double* fstore[3] = {};
    double* gyz[3] = {};
    double* gxz[3] = {};
    double* gxy[3] = {};
    double* eg[6] = {};
    double* ec[6] = {};
    double* e0g[6] = {};
    double* e0c[6] = {};
    double* us_alt[3] = {};         // the u store an elastic stage writes (swapped with us after it)
    double* epx[3] = {};            // face-stress parts of the cells (ElStageArgs)
    double* epy[3] = {};
    double* epz[3] = {};
    double* hist = nullptr;
    int hist_slots = 0;
    int hist_head = 0;
    int el_next_stage = 1;
    double vl_first = 0, vt_first = 0, vl_last = 0, vt_last = 0;   // sound speeds at k = 0 and nz-1
    int src_type = 0, src_k = -1, src_comp = -1, src_cycles = 0;
    double src_sigma = 0, src_amp = 0, w_liao = 0;
    // electrodynamics: Yee E/H with double-buffered stage stores (0 = base,
    // 1 = a, 2 = b), accumulators, boundary history, cell fields, sources
    bool em_on = false, em_periodic_all = false, em_pec_all = false, jp_on = false, jp_uploaded = false;
    int64_t nE[3] = {}, nH[3] = {};
    double* Eb[3] = {};
    double* Ea[3] = {};
    double* Ebb[3] = {};
    double* Eacc[3] = {};
    double* Hb[3] = {};
    double* Ha[3] = {};
    double* Hbb[3] = {};
    double* Hacc[3] = {};
    double* Et1[3] = {};
    double* Ecell[3] = {};
    double* Hcell[3] = {};
    double* Jp[9] = {};
    double* Jps[9] = {};
    double* Jacc[9] = {};
    uint8_t* pec[3] = {};
    double* jf_spatial = nullptr;
    double* em_dm[3] = {};          // FM increments of the stage (Faraday's law)
    double* em_dm1[3] = {};
    double* em_dm2[3] = {};
    const double* fe_q_used[3] = {}; // q store read by the last ferroelectric stage
    // gf_step with a ferroelectric and electrodynamics: the FE stage's
    // launch is deferred and runs in the EM stage's Ampere/Faraday launch
    bool fe_defer = false;
    bool fe_pending = false;
    gf::FeStageArgs fe_args{};
    int e_sel = 0, h_sel = 0, em_next_stage = 1;
    double cf_top_half = 0, cf_bot_half = 0, cf_x_half = 0, cf_y_half = 0;
    double cf_top = 0, cf_bot = 0, cf_x = 0, cf_y = 0;
    double plane_E[3] = {}, plane_f[3] = {};
    int plane_z = 0;
    int jf_type = 0, jf_comp = -1, jf_cycles = 0;
    int jf_lo[3] = {}, jf_hi[3] = {};
    double jf_amp = 0, jf_freq = 0, jf_sigma = 0;
    // spin pumping (GF_FLAG_SPIN_PUMPING): J_ISHE (x, y, zero z), the
    // per-plane source and coefficient, the stores the last magnet stages read
    bool spin_on = false;
    double* jis[3] = {};
    int32_t* spin_src = nullptr;
    double* spin_coef = nullptr;
    int32_t* spin_is_src = nullptr;  // per plane: a source plane of some pumping plane
    double* spin_m0[3] = {};         // source planes of the state before the stage-4 update
    double* spin_m10[3] = {};
    double* spin_m20[3] = {};
    int mag_in_sel = 0, afm_in_sel = 0;
    // flexoelectricity (GF_FLAG_FLEXO): per-material constants, the force
    // of the current stage
    gf::FlexoMat* d_flexo = nullptr;
    double* flexo_force[3] = {};
    // elastostatics (GF_FLAG_ELASTOSTATIC): six complex work fields and the
    // homogeneous c11, c12, c44
    double2* es_buf = nullptr;
    double es_c[3] = {};
    // statics
    double* hstat[3] = {};
    bool hstat_written = false;     // Hstat uploaded (it is zero otherwise unless statics run)
    double* hnu[3] = {};
    double* hdrive[3] = {};         // shaped drive profile (GF_FLAG_HDRIVE)
    double hdrive_freq = 0;
    double precess_angle = 0;       // prescribed bottom layer (GF_FLAG_PRESCRIBED_M), degrees
    double precess_freq = 0;        // its frequency, Hz
    double* e_t0[6] = {};           // static crystal-frame strain (magnets, no elastodynamics)
    double* e0t0[6] = {};           // static crystal-frame eigenstrain at t0 (elastodynamics), on first use
    // reduction scratch
    double* red_partial = nullptr;
    int64_t red_blocks = 0;

    // Which buffer holds the store the next stage reads: 0 = m, 1 = a, 2 = b.
    int mag_store_sel = 0;
    int afm_store_sel = 0;
    int mag_next_stage = 1;
    int afm_next_stage = 1;
    int fe_store_sel = 0;
    int fe_next_stage = 1;

    gf::DemagPlan demag;
    gf::DemagPlan eplan;            // FFT electrostatics (GF_FLAG_ELECTROSTATICS)
    gf::AdjState* adj = nullptr;    // checkpoints and buffers of the adjoint sweep
    gf::Adj32* adj32 = nullptr;     // the same for its FP32 variant

    bool timing = false;
    std::vector<gf::TimerSlot> timers;
    std::vector<hipEvent_t> event_pool;
    double kernel_ms[GF_K_COUNT] = {};
    int64_t kernel_launches[GF_K_COUNT] = {};
};

namespace gf {

// error helpers (gf_api.hip)
gf_status set_error(gf_ctx* c, gf_status code, const std::string& msg);
gf_status hip_check(gf_ctx* c, hipError_t e, const char* what);
// elastostatics (gf_elastostatic.hip): set-up, and the static strain of the
// given magnet stores into e_t0
gf_status elastostatic_setup(gf_ctx* c, const gf_desc* d);
gf_status elastostatic_run(gf_ctx* c, double* const* fm, double* const* a1, double* const* a2, double* const* p);
// the thread's error of a call without a context (gf_last_error(NULL))
gf_status set_thread_error(gf_status code, const std::string& msg);
void timer_begin(gf_ctx* c, int cls, hipEvent_t* ev);
void timer_end(gf_ctx* c, int cls, hipEvent_t start);

// stage arguments of a context (gf_api.hip): stage store in_sel is read and
// out_sel written (0 = m, 1 = first, 2 = second double-buffered store);
// the drive is left at zero
StageArgs fm_stage_args(gf_ctx* c, int stage, int in_sel, int out_sel);
void fm_stage_drives(gf_ctx* c, int stage, StageArgs& a);
// adjoint sweep state (gf_adjoint.hip) and its FP32 variant (gf_fp32.hip)
void adjoint_free(gf_ctx* c);
void adjoint32_free(gf_ctx* c);
int adjoint32_material(const gf_ctx* c);
gf_status adjoint32_forward(gf_ctx* c, const gf_adjoint_desc* d, double* loss);
gf_status adjoint32_backward(gf_ctx* c, double grad[3], double* grad_hdrive, size_t count);
AfmStageArgs afm_stage_args(gf_ctx* c, int stage, int in_sel, int out_sel);
FeStageArgs fe_stage_args(gf_ctx* c, int stage, int in_sel, int out_sel);
// "" when no array a stage launch writes is read by it at neighbouring
// cells or as another quantity, else what aliases (gf_api.hip)
std::string stage_io_conflict(const StageArgs& a, const FusedZArgs* f);
std::string stage_io_conflict(const AfmStageArgs& a);
std::string stage_io_conflict(const FeStageArgs& a);
// true when cell indices of an n-cell grid fit 32-bit byte offsets (gf_llg.hip)
bool narrow(int64_t n);

// launchers (gf_llg.hip)
// ids: when set, the count FM cells the stage runs over (else all n cells)
hipError_t launch_mag_stage(const StageArgs& a, int64_t n, hipStream_t s, const uint32_t* ids = nullptr,
                            int64_t count = 0);
hipError_t launch_afm_stage(const AfmStageArgs& a, int64_t n, hipStream_t s);
hipError_t launch_mag_batch(const StageArgs* items, const BatchLaunch& L, int stage, hipStream_t s);
hipError_t launch_afm_batch(const AfmStageArgs* items, const BatchLaunch& L, int stage, hipStream_t s);
hipError_t launch_batch_clock(BatchClock* clk, StageArgs* fm, AfmStageArgs* afm, int count, hipStream_t s);
hipError_t launch_mag_stage_fused(const StageArgs& a, const FusedZArgs& f, bool hspec, hipStream_t s);

// ferroelectric (gf_fe.hip)
hipError_t launch_fe_stage(const FeStageArgs& a, int64_t n, hipStream_t s);

// elastodynamics (gf_elastic.hip)
hipError_t launch_el_strain(const ElStrainArgs& a, hipStream_t s);
hipError_t launch_el_stage(const ElStageArgs& a, hipStream_t s);
hipError_t launch_el_history(const ElHistArgs& a, hipStream_t s);
hipError_t launch_el_boundary(const ElBoundaryArgs& a, hipStream_t s);

// electrodynamics (gf_em.hip)
hipError_t launch_em_eh(const EmArgs& a, hipStream_t s);
hipError_t launch_fe_em(const FeStageArgs& f, const EmArgs& e, hipStream_t s);
hipError_t launch_em_jp(const EmJpArgs& a, hipStream_t s);
hipError_t launch_em_cells(const EmCellArgs& a, hipStream_t s);
hipError_t launch_spin_pump(const SpinArgs& a, hipStream_t s);
// copies the source planes (is_src[k] != 0) of ncomp arrays
struct SpinSaveArgs {
    const double* src[6];
    double* dst[6];
    int32_t ncomp;
    const int32_t* is_src;
    int64_t nx, ny, nz;
};
hipError_t launch_spin_save(const SpinSaveArgs& a, hipStream_t s);
hipError_t launch_fe_flexo(const FeFlexoArgs& a, hipStream_t s);
hipError_t launch_em_wall(const EmWallArgs& a, hipStream_t s);
hipError_t launch_em_plane(const EmMiscArgs& a, hipStream_t s);
hipError_t launch_em_pec(const EmMiscArgs& a, hipStream_t s);

// demag (gf_fft.hip)
gf_status demag_setup(gf_ctx* c);
gf_status demag_run(gf_ctx* c, const double* const* fm_store, const double* const* s1,
                    const double* const* s2);
gf_status demag_materialize_h(gf_ctx* c);
gf_status demag_run_fused(gf_ctx* c, const double* const* fm_store);
gf_status demag_run_to(gf_ctx* c, const double* const* fm_store, double* const* out);
gf_status demag_adjoint(gf_ctx* c, const double* const* in, double* const* out);
// FFT electrostatics on the same plan (gf_fft.hip)
gf_status elec_setup(gf_ctx* c, double er);
gf_status elec_run(gf_ctx* c, const double* const* p);
void demag_free(gf_ctx* c);
hipError_t launch_demag_local(gf_ctx* c, const double* const* fm_store, const double* const* s1,
                              const double* const* s2, int one_d);

// reductions (gf_llg.hip)
hipError_t launch_prescribe_plane(double* const* m, const uint32_t* code, int64_t nx, int64_t ny, int64_t nz,
                                  const double v[3], hipStream_t s);
hipError_t launch_sum3(const double* const* f, int64_t n, double* partial, int64_t blocks, hipStream_t s);

// host Newell kernel (gf_newell.cpp)
void newell_real_kernels(int64_t nx, int64_t ny, int64_t nz, double dx, double dy, double dz,
                         double* out9, int threads);

}  // namespace gf
