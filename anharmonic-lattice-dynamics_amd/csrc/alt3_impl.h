// Internal declarations of the alt3 library (HIP/CDNA4 three-phonon path).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/alt3.h"

struct alt3_ctx;

namespace alt3 {

// ---------------------------------------------------------------- errors
void set_error(const std::string &msg);
int fail(int code, const std::string &msg);

#define ALT3_HIP(call)                                                                      \
    do {                                                                                    \
        hipError_t e_ = (call);                                                             \
        if (e_ != hipSuccess)                                                               \
            return ::alt3::fail(ALT3_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define ALT3_TRY(call)              \
    do {                            \
        int rc_ = (call);           \
        if (rc_ != ALT3_OK) return rc_; \
    } while (0)

// ---------------------------------------------------------------- physical constants
// THERMACOND/modules.f90:4-24; cnst is a REAL*4 literal there, so it carries float rounding.
constexpr double PI = 3.14159265358979323846;
constexpr double H_PLANK = 6.62606896e-34;
constexpr double K_B = 1.3806504e-23;
constexpr double C_LIGHT = 2.99792458e+8;
constexpr double HBAR = 1.054571628e-34;
constexpr double EE = 1.60217653e-19;
constexpr double UMA = 1.66053878e-27;
inline double cnst() { return (double)521.1098918f; }
double const33();

// ---------------------------------------------------------------- device buffer
// owns its allocation: freed by release() or when the buffer goes out of scope (hipFree
// waits for the device, so a scoped temporary outlives the kernels that use it)
template <typename T>
struct DBuf {
    T *p = nullptr;
    size_t n = 0;
    DBuf() = default;
    DBuf(const DBuf &) = delete;
    DBuf &operator=(const DBuf &) = delete;
    ~DBuf() { release(); }
    int alloc(size_t count) {
        if (count == n && p) return ALT3_OK;
        release();
        if (count == 0) return ALT3_OK;
        if (hipMalloc(&p, count * sizeof(T)) != hipSuccess) {
            p = nullptr;
            return fail(ALT3_E_NOMEM, "hipMalloc of " + std::to_string(count * sizeof(T)) + " bytes failed");
        }
        n = count;
        return ALT3_OK;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    // grow-only: keeps a larger allocation (scratch reused across calls without hipMalloc)
    int reserve(size_t count) { return (p && count <= n) ? ALT3_OK : alloc(count); }
    size_t bytes() const { return n * sizeof(T); }
};

// Mesh geometry shared by the kernels: mesh point index  idx = c2 + n2*(c1 + n1*c0)
// with c in [0,n) (THERMACOND's i-outer / k-inner loop), reduced coordinate c/n - 1/2.
struct MeshDims {
    int n0, n1, n2;
};

// ---------------------------------------------------------------- Born charges
// constants of nonanal (rest.f90:3517-3578) for the eigen kernels; on == 0: no term
struct BornDev {
    int on;
    double epsil[9];     // row-major epsil(i,j)
    double eps_scale;    // 8.8541878176e-12 / 1e10 / ee
    double om0;          // |r3 . (r1 x r2)|
    double rho2;         // ((|g1||g2||g3|)^0.33 / 4)^2
};

// ---------------------------------------------------------------- host tables
struct HostTables {
    int nk = 0, nibz = 0, nops = 0, npos = 0;
    double g[9] = {0};           // rows g1,g2,g3 (2 pi included)
    double volume_r = 0;
    std::vector<double> kpc_eig; // [nk][3] REAL*4-fraction mesh (eigen solves)
    std::vector<double> kpc;     // [nk][3] FP64 mesh
    std::vector<int> mapibz, mapinv, mappos;
    std::vector<double> kibz, wibz;
    std::vector<double> opk;     // [nops][9] column-major
    std::vector<double> opk_inv; // [nops][9] column-major inverse
    // stars on the mesh (kstarsinfbz1) per IBZ point
    std::vector<int> star_n, star_op, star_n0;        // [nibz], [nibz][48], [nibz][48]
    // full stars (getstar) folded onto the mesh (veloc_avergeIBZ_invsyFBZ)
    std::vector<int> fstar_n, fstar_op, fstar_n0;
    std::vector<int> tet;        // [6nk][4]
};

int build_tables(const alt3_system *sys, HostTables &t);

// ---------------------------------------------------------------- FC3 preparation
struct V3Tables {
    // reference-order term list (home-cell first index), for the per-element kernel
    int nterm = 0;
    std::vector<int> t_abc;        // [nterm][3] row indices a,b,c (alpha + 3*(atom0-1))
    std::vector<double> t_psi;     // [nterm] fc3/sqrt(mi mj mk)
    std::vector<double> t_rj, t_rk;// [nterm][3]
    // phase-GEMM layout: slots (dR, b, c) with term lists
    int ndr = 0, nslot = 0;
    std::vector<double> dr;        // [ndr][3]
    std::vector<int> slot_dr, slot_bc;   // [nslot]
    std::vector<int> slot_ptr;     // [nslot+1]
    std::vector<int> sterm_a;      // [nterm] a of each term in slot order
    std::vector<double> sterm_psi; // [nterm]
    std::vector<double> sterm_rk;  // [nterm][3]
};
int build_v3_tables(const alt3_system *sys, V3Tables &v);
int v3_prep_launch(alt3_ctx *ctx, int ksub);
int v3_prep_points(alt3_ctx *ctx, int nq, const double *q1, const double *e1, double *gp);
int eig_launch(alt3_ctx *ctx, int np, int nvar, const double *kq, double dq, double *evals, double *evec);
int v3_rows_points(alt3_ctx *ctx, int nq, const double *gp, const double *ev3, const double *w3, const double *w1,
                   double *out, int *d_err);
bool v3_rows_supported(const alt3_ctx *ctx, bool fp32);
bool v3_chain_supported(const alt3_ctx *ctx);
int v3_rows_inputs(alt3_ctx *ctx, int ksub, bool fp32, bool want_gpf);
bool v3_batch_supported(const alt3_ctx *ctx);
int v3_tile_for(alt3_ctx *ctx, bool fp32, int *tq);
int v3_batch_launch(alt3_ctx *const *ctxs, int n, bool fp32, int tq);
int v3_chain_run(alt3_ctx *ctx, int ksub, const float *evf, const float *ephf, const double *wq, float *out,
                 int *d_err);
bool v3_chain16_supported(const alt3_ctx *ctx);
size_t v3_chain16_pack_floats(int nk, int ksub);
int v3_chain16_run(alt3_ctx *ctx, int ksub, const float *evf, const float *ephf, const double *wq, float *out,
                   int *d_err);
int v3_rows_run(alt3_ctx *ctx, int ksub, bool fp32, int *d_err);
int prow_all(alt3_ctx *ctx);   // mark every P row as stored (writers that fill all rows)
int piso_run(alt3_ctx *ctx);   // P_iso_split into d_piso when giso is set and piso_ok is not
int cg_start_launch(alt3_ctx *ctx, double *x, double *te);
int cg_device_ws(alt3_ctx *ctx, int n, const double *dA, const double *db, double *dx, int max_iter, int ndig,
                 double *funcmin, int *iters, bool rowmajor, const float *dAf, double *work, double *red, int *state);
int scratch_setup();   // the static scratch of alt3_cg_static on the current device (cg_static.hip)
// bracket the launch of kernel class k (0: V3 contraction, 1: FGR) with the context's events
inline void kernel_mark(alt3_ctx *ctx, int k, bool end);

}  // namespace alt3

// ---------------------------------------------------------------- the context
struct alt3_ctx {
    int natoms0 = 0, nb = 0, nk = 0, nibz = 0, nops = 0, classical = 0;
    bool host_only = false;
    int nc[3] = {0, 0, 0};
    alt3::HostTables T;
    alt3::V3Tables V;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // HIP events around the last |V3|^2 contraction kernel and the last FGR kernel (alt3_kernel_ms)
    hipEvent_t ev_k[4] = {nullptr, nullptr, nullptr, nullptr};
    bool ev_k_set[2] = {false, false};
    double last_ms = 0;
    long long peak_bytes = 0;
    size_t lds_max = 65536;                        // LDS per workgroup on this device
    int device = 0;                                // the HIP device of stream and buffers

    // harmonic
    int nt2h = 0;                                 // FC2 terms with a home-cell first atom
    alt3::DBuf<int> d_fc2_rc;                      // [nt2h][2] row, col
    alt3::DBuf<double> d_fc2_amp, d_fc2_rr;        // [nt2h], [nt2h][3]
    alt3::DBuf<double> d_kpc_eig, d_kpc;           // [nk][3]
    alt3::DBuf<double> d_eival, d_freq, d_vel;     // [nk][nb], [nk][nb], [nk][nb][3]
    alt3::DBuf<double> d_evec;                     // [nk][nb][nb] complex (interleaved)
    alt3::DBuf<int> d_mappos;
    alt3::BornDev born{};                          // non-analytic term (born_flag != 0)
    alt3::DBuf<double> d_born;                     // [natoms0][10]: charge(i,j) row-major, mass
    bool have_eig = false;
    bool have_veloci = false;                      // d_veloci holds the star averages of d_vel

    // V3
    alt3::DBuf<int> d_t_abc;
    alt3::DBuf<double> d_t_psi, d_t_rj, d_t_rk;
    alt3::DBuf<double> d_dr;
    alt3::DBuf<int> d_slot_dr, d_slot_bc, d_slot_ptr, d_sterm_a;
    alt3::DBuf<double> d_sterm_psi, d_sterm_rk;
    alt3::DBuf<double> d_eph;                      // [nk][ndr] complex e^{i q2.dR}
    alt3::DBuf<double> d_gp;                       // per q1: [ndr][nb][nb*nb] complex (l1,b,c)
    alt3::DBuf<float> d_gpf;                       // packed FP32 B operand of the mixed path
    alt3::DBuf<int> d_mapinv;
    int ibz_lo = -1, ibz_hi = -1, v3_prec = -1;    // v3_prec: what d_v33/d_v33f hold
    int v3_req = -1;                               // precision the caller asked for
    alt3::DBuf<double> d_v33;                      // [ksub][nb][nb][nb][nk] FP64
    alt3::DBuf<float> d_v33f;                      // same, FP32 (mixed path)
    bool have_v3 = false;

    // FGR
    int delta_kind = ALT3_DELTA_TETRA;             // used by alt3_solve
    int v3_mma = ALT3_MMA_F16X3;                   // matrix-core arithmetic of the 24-branch mixed V3
    double delta_sigma = 0;                        // Gaussian width, cm^-1
    int fgr_kind = ALT3_DELTA_TETRA;               // what d_p1/d_p2 hold
    double fgr_sigma = 0;
    alt3::DBuf<double> d_p1, d_p2;                 // [ksub][nb][nb][nb][nk]
    alt3::DBuf<float> d_p1f, d_p2f;                // same in FP32 (mixed path: P = W |V3|^2 with FP32 |V3|^2)
    bool p_fp32 = false;                           // which pair holds the current P
    // P rows held: [process][ksub][nb][nb][nb] (l1, l2, l3 of one (s, process)); 0 = the row is
    // zero and not stored (its memory is stale), 1 = stored.  The tetrahedron FGR leaves rows
    // whose l1 no tetrahedron reaches unstored; every reader of P goes through this map.
    alt3::DBuf<unsigned char> d_prow;
    alt3::DBuf<double> d_pzero;                    // a line of zeros the tau loads of unstored rows read
    alt3::DBuf<unsigned long long> d_prowb;        // the same packed per (s, l1, l2): P1 bits 0-31, P2 bits 32-63
    bool have_fgr = false;

    // per temperature: one tau pass fills slot t of each array for every temperature it
    // was given (tau_temps); the consumers read the slot of the temperature they are asked for
    double tempk = -1;
    std::vector<double> tau_temps;
    int tau_cur = 0;
    alt3::DBuf<double> d_dist;                     // [t][nk][nb]
    alt3::DBuf<double> d_q, d_qn, d_qu;            // [t][ksub][nb]
    alt3::DBuf<double> d_r1, d_r2;                 // [t][ksub][nb][nb][nk]  sum_l3 P1, P2
    alt3::DBuf<float> d_r1f, d_r2f;                // same in FP32 (mixed path, R has the precision of P)
    bool r_fp32 = false;
    size_t ksub_nb() const { return (size_t)(ibz_hi - ibz_lo + 1) * nb; }
    double *dist() const { return d_dist.p + (size_t)tau_cur * nk * nb; }
    double *q() const { return d_q.p + (size_t)tau_cur * ksub_nb(); }
    double *qn() const { return d_qn.p + (size_t)tau_cur * ksub_nb(); }
    double *qu() const { return d_qu.p + (size_t)tau_cur * ksub_nb(); }
    double *r1() const { return d_r1.p + (size_t)tau_cur * ksub_nb() * nb * nk; }
    double *r2() const { return d_r2.p + (size_t)tau_cur * ksub_nb() * nb * nk; }
    float *r1f() const { return d_r1f.p + (size_t)tau_cur * ksub_nb() * nb * nk; }
    float *r2f() const { return d_r2f.p + (size_t)tau_cur * ksub_nb() * nb * nk; }
    bool have_tau = false;
    alt3::DBuf<int> d_star_n, d_star_op, d_star_n0, d_fstar_n, d_fstar_op, d_fstar_n0;
    alt3::DBuf<int> d_cm_slot_n2, d_cm_tab, d_cm_split_x;   // collision star table, see star_sums
    alt3::DBuf<int4> d_cm_task;
    int cm_ntask = 0, cm_nsplit = 0;
    alt3::DBuf<double> d_opk, d_opk_inv, d_kibz;
    alt3::DBuf<double> d_frta, d_kpart, d_kappa, d_taueff, d_veloci;
    alt3::DBuf<double> d_cm, d_rhs;
    alt3::DBuf<float> d_cmf;                       // the collision matrix in FP32 (mixed path)
    bool cm_fp32 = false;                          // which of d_cm / d_cmf holds it
    bool have_cm = false;
    alt3::DBuf<double> d_cg;                       // CG work vectors
    // iterative BTE: S rows (sum over l2 of P with occupations, [ksub][nb][nb][nk], in the
    // precision of R) and the update matrix U (rows (s,l1,alpha), columns (i,l,beta)) of upd_temp
    alt3::DBuf<double> d_s1, d_s2, d_upd, d_iter;
    alt3::DBuf<float> d_s1f, d_s2f;
    bool have_upd = false;
    double upd_temp = -1;
    // isotope scattering (iso = 1): per-atom g, Piso [nibz][nb] of the current eigenvectors
    std::vector<double> giso;                      // empty: off
    alt3::DBuf<double> d_giso, d_piso;
    bool piso_ok = false;                          // d_piso matches the eigen data
    // per-call scratch kept across calls (no hipMalloc/hipFree on the solve path)
    alt3::DBuf<int> d_flags;                       // [0] V3 den=0, [1] tau, [2] kappa_rta, [3] kappa_bte, [4,5] CG,
                                                   // [6,7,8] iterative BTE (converged, iterations, N_BE error)
    alt3::DBuf<int> d_negs;                        // subst_eivecs pairs
    alt3::DBuf<double> d_qpart, d_wq, d_ev7, d_evec0;
    alt3::DBuf<float> d_ephf, d_evf, d_gpf32;
    alt3::DBuf<float> d_chain_e, d_chain_g;        // operand packs of k_v3_chain (eigenvectors, G')
    alt3::DBuf<unsigned char> d_batch;             // alt3_v3_batch: work counter and job table
    alt3::DBuf<float> d_fgr_big;                   // k_fgr_big: per-workgroup energies, ranges, weight sums
    alt3::DBuf<double> d_red;                      // reduction scratch
    alt3::DBuf<double> d_freqT;                    // frequencies branch-major [nb][nk] (FGR)
    alt3::DBuf<double> d_solve_x, d_solve_te;      // alt3_solve: CG vector, tauinv_eff
    alt3::DBuf<double> d_solve_k;                  // alt3_solve: kappa_RTA, kappa_BTE per temperature
    std::vector<hipEvent_t> evpool;                // alt3_solve: stage events
};

inline void alt3::kernel_mark(alt3_ctx *ctx, int k, bool end)
{
    (void)hipEventRecord(ctx->ev_k[2 * k + (end ? 1 : 0)], ctx->stream);
    if (end) ctx->ev_k_set[k] = true;
}
