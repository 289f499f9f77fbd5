// C ABI of the alt3 library (declared in include/alt3.h).
#include <algorithm>
#include <chrono>
#include <cstring>

#include "alt3_impl.h"

namespace alt3 {

thread_local std::string g_last_error;
void set_error(const std::string &msg) { g_last_error = msg; }
int fail(int code, const std::string &msg)
{
    set_error(msg);
    return code;
}

int eig_prepare(alt3_ctx *ctx, const alt3_system *sys);
int eig_run(alt3_ctx *ctx);
int v3_prepare(alt3_ctx *ctx);
int v3_run(alt3_ctx *ctx, int ibz_lo, int ibz_hi, int prec, bool sync);
int v3_copy_out(alt3_ctx *ctx, double *host);
bool v3_mixed_fits(const alt3_ctx *ctx);   // v3_mixed.hip
template <typename T>
int copy_fortran5(alt3_ctx *ctx, const T *dev, double *host, const unsigned char *rmap = nullptr);
int fgr_run(alt3_ctx *ctx, bool sync);
int fgr_gauss_run(alt3_ctx *ctx, double sigma);
int upload_bte_tables(alt3_ctx *ctx);
int tau_run(alt3_ctx *ctx, double tempk, double *Q, double *QN, double *QU, double *tN, double *tU, double *tT);
int tau_multi_run(alt3_ctx *ctx, const double *temps, int nt, bool sync);
int kappa_rta_run(alt3_ctx *ctx, double tempk, double *kappa, double *FRTA, double *taueff, bool sync);
int collision_run(alt3_ctx *ctx, double tempk, double *CM, double *RHS, bool sync);
int cg_device(alt3_ctx *ctx, int n, const double *dA, const double *db, double *dx, int max_iter, int ndig,
              double *funcmin, int *iters, bool rowmajor, const float *dAf);
int kappa_bte_run(alt3_ctx *ctx, double tempk, const double *dx, double *kappa, double *taueff_dev, bool sync);
int update_run(alt3_ctx *ctx, double tempk, const double *FFF, const double *FRTA, double *F2);
int load_fortran5(alt3_ctx *ctx, int ksub, const double *host, double *dev);
int self_run(alt3_ctx *ctx, int nq, const double *q, double tempk, double eta, double v3thr, int kind,
             double *omega0,
             double *nself, double *uself);
int kappa_iter_run(alt3_ctx *ctx, double tempk, int max_iter, double mix, double tol, double *kappa, double *khist,
                   double *tehist, double *FFF, int *iters);

namespace {
int copy_p(alt3_ctx *ctx, double *P1sm, double *P2sm)
{
    const size_t p2 = ctx->d_prow.n / 2;   // rows not stored come out as zeros
    const unsigned char *m1 = ctx->d_prow.p, *m2 = m1 ? m1 + p2 : nullptr;
    if (ctx->p_fp32) {
        if (P1sm) ALT3_TRY(copy_fortran5(ctx, ctx->d_p1f.p, P1sm, m1));
        if (P2sm) ALT3_TRY(copy_fortran5(ctx, ctx->d_p2f.p, P2sm, m2));
    } else {
        if (P1sm) ALT3_TRY(copy_fortran5(ctx, ctx->d_p1.p, P1sm, m1));
        if (P2sm) ALT3_TRY(copy_fortran5(ctx, ctx->d_p2.p, P2sm, m2));
    }
    return ALT3_OK;
}

struct Timer {
    alt3_ctx *c;
    explicit Timer(alt3_ctx *ctx) : c(ctx) { (void)hipEventRecord(c->ev0, c->stream); }
    void stop()
    {
        float ms = 0;
        (void)hipEventRecord(c->ev1, c->stream);
        (void)hipEventSynchronize(c->ev1);
        (void)hipEventElapsedTime(&ms, c->ev0, c->ev1);
        c->last_ms = ms;
    }
};
}  // namespace

}  // namespace alt3

using namespace alt3;

extern "C" {

const char *alt3_last_error(void) { return g_last_error.c_str(); }

int alt3_create(const alt3_system *sys, alt3_ctx **out)
{
    if (!sys || !out) return fail(ALT3_E_ARG, "null argument");
    *out = nullptr;
    if (sys->natoms0 < 1 || sys->natoms < sys->natoms0 || !sys->mass || !sys->r123 || !sys->atompos || !sys->iatomcell0)
        return fail(ALT3_E_ARG, "incomplete crystal description");
    if (3 * sys->natoms0 > 64) return fail(ALT3_E_ARG, "more than 64 branches per cell are not supported");
    for (int i = 0; i < sys->natoms; i++)
        if (sys->iatomcell0[i] < 1 || sys->iatomcell0[i] > sys->natoms0) return fail(ALT3_E_ARG, "iatomcell0 out of range");
    for (int i = 0; i < sys->natoms0; i++)
        if (sys->iatomcell0[i] != i + 1) return fail(ALT3_E_ARG, "the first natoms0 listed atoms must be the primitive cell");
    int ndev = 0;
    if (!sys->host_only && (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1))
        return fail(ALT3_E_HIP, "no HIP device available");
    alt3_ctx *ctx = new alt3_ctx();
    ctx->host_only = sys->host_only != 0;
    ctx->natoms0 = sys->natoms0;
    ctx->nb = 3 * sys->natoms0;
    ctx->classical = sys->classical;
    for (int a = 0; a < 3; a++) ctx->nc[a] = sys->nc[a];
    int rc = build_tables(sys, ctx->T);
    if (rc == ALT3_OK) rc = build_v3_tables(sys, ctx->V);
    if (rc != ALT3_OK) {
        delete ctx;
        return rc;
    }
    ctx->nk = ctx->T.nk;
    ctx->nibz = ctx->T.nibz;
    ctx->nops = ctx->T.nops;
    if (ctx->host_only) {
        *out = ctx;
        return ALT3_OK;
    }
    auto init = [&]() -> int {
        ALT3_HIP(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
        ALT3_TRY(ctx->d_flags.alloc(16));
        int dev = 0, lds = 0;
        ALT3_HIP(hipGetDevice(&dev));
        ctx->device = dev;
        ALT3_TRY(scratch_setup());
        ALT3_HIP(hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
        if (lds > 0) ctx->lds_max = (size_t)lds;
        ALT3_HIP(hipEventCreate(&ctx->ev0));
        ALT3_HIP(hipEventCreate(&ctx->ev1));
        for (hipEvent_t &e : ctx->ev_k) ALT3_HIP(hipEventCreate(&e));
        ALT3_TRY(ctx->d_kpc_eig.alloc(3 * ctx->nk));
        ALT3_TRY(ctx->d_kpc.alloc(3 * ctx->nk));
        ALT3_TRY(ctx->d_mapinv.alloc(ctx->nibz));
        ALT3_HIP(hipMemcpy(ctx->d_kpc_eig.p, ctx->T.kpc_eig.data(), sizeof(double) * 3 * ctx->nk, hipMemcpyHostToDevice));
        ALT3_HIP(hipMemcpy(ctx->d_kpc.p, ctx->T.kpc.data(), sizeof(double) * 3 * ctx->nk, hipMemcpyHostToDevice));
        ALT3_HIP(hipMemcpy(ctx->d_mapinv.p, ctx->T.mapinv.data(), sizeof(int) * ctx->nibz, hipMemcpyHostToDevice));
        ALT3_TRY(eig_prepare(ctx, sys));
        ALT3_TRY(v3_prepare(ctx));
        ALT3_TRY(upload_bte_tables(ctx));
        return ALT3_OK;
    };
    rc = init();
    if (rc != ALT3_OK) {
        alt3_destroy(ctx);
        return rc;
    }
    *out = ctx;
    return ALT3_OK;
}

int alt3_destroy(alt3_ctx *ctx)
{
    if (!ctx) return ALT3_OK;
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    // device buffers go first, while the context's stream and device state still exist
    ctx->d_fc2_rc.release(); ctx->d_fc2_amp.release(); ctx->d_fc2_rr.release();
    ctx->d_kpc_eig.release(); ctx->d_kpc.release(); ctx->d_eival.release(); ctx->d_freq.release();
    ctx->d_vel.release(); ctx->d_evec.release(); ctx->d_mappos.release(); ctx->d_born.release();
    ctx->d_t_abc.release(); ctx->d_t_psi.release(); ctx->d_t_rj.release(); ctx->d_t_rk.release();
    ctx->d_dr.release(); ctx->d_slot_dr.release(); ctx->d_slot_bc.release(); ctx->d_slot_ptr.release();
    ctx->d_sterm_a.release(); ctx->d_sterm_psi.release(); ctx->d_sterm_rk.release();
    ctx->d_eph.release(); ctx->d_gp.release(); ctx->d_gpf.release(); ctx->d_mapinv.release(); ctx->d_v33.release(); ctx->d_v33f.release();
    ctx->d_p1.release(); ctx->d_p2.release(); ctx->d_p1f.release(); ctx->d_p2f.release(); ctx->d_dist.release(); ctx->d_q.release(); ctx->d_qn.release();
    ctx->d_qu.release(); ctx->d_r1.release(); ctx->d_r2.release(); ctx->d_r1f.release(); ctx->d_r2f.release();
    ctx->d_star_n.release(); ctx->d_star_op.release(); ctx->d_star_n0.release();
    ctx->d_cm_slot_n2.release(); ctx->d_cm_tab.release(); ctx->d_cm_task.release(); ctx->d_cm_split_x.release();
    ctx->d_fstar_n.release(); ctx->d_fstar_op.release(); ctx->d_fstar_n0.release();
    ctx->d_opk.release(); ctx->d_opk_inv.release(); ctx->d_kibz.release(); ctx->d_frta.release();
    ctx->d_kpart.release(); ctx->d_kappa.release(); ctx->d_taueff.release(); ctx->d_veloci.release();
    ctx->d_cm.release(); ctx->d_cmf.release(); ctx->d_rhs.release(); ctx->d_cg.release(); ctx->d_red.release();
    ctx->d_flags.release(); ctx->d_negs.release(); ctx->d_qpart.release(); ctx->d_wq.release();
    ctx->d_ev7.release(); ctx->d_evec0.release(); ctx->d_ephf.release(); ctx->d_evf.release(); ctx->d_gpf32.release();
    ctx->d_chain_e.release(); ctx->d_chain_g.release(); ctx->d_batch.release(); ctx->d_fgr_big.release(); ctx->d_prow.release(); ctx->d_prowb.release(); ctx->d_pzero.release();
    ctx->d_s1.release(); ctx->d_s2.release(); ctx->d_s1f.release(); ctx->d_s2f.release(); ctx->d_upd.release(); ctx->d_iter.release();
    for (hipEvent_t e : ctx->evpool) (void)hipEventDestroy(e);
    for (hipEvent_t e : ctx->ev_k)
        if (e) (void)hipEventDestroy(e);
    if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
    if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return ALT3_OK;
}

int alt3_device_free(long long *free_bytes, long long *total_bytes)
{
    if (!free_bytes) return fail(ALT3_E_ARG, "null argument");
    size_t f = 0, t = 0;
    ALT3_HIP(hipMemGetInfo(&f, &t));
    *free_bytes = (long long)f;
    if (total_bytes) *total_bytes = (long long)t;
    return ALT3_OK;
}

int alt3_device_bytes(const alt3_ctx *ctx, long long *bytes)
{
    if (!ctx || !bytes) return fail(ALT3_E_ARG, "null argument");
    long long b = 0;
    b += ctx->d_evec.bytes() + ctx->d_eival.bytes() + ctx->d_freq.bytes() + ctx->d_vel.bytes();
    b += ctx->d_v33.bytes() + ctx->d_v33f.bytes() + ctx->d_gp.bytes() + ctx->d_gpf.bytes();
    b += ctx->d_p1.bytes() + ctx->d_p2.bytes() + ctx->d_p1f.bytes() + ctx->d_p2f.bytes() + ctx->d_r1.bytes() + ctx->d_r2.bytes() + ctx->d_cm.bytes() + ctx->d_cmf.bytes();
    b += ctx->d_r1f.bytes() + ctx->d_r2f.bytes();
    b += ctx->d_s1.bytes() + ctx->d_s2.bytes() + ctx->d_s1f.bytes() + ctx->d_s2f.bytes() + ctx->d_upd.bytes();
    b += ctx->d_qpart.bytes() + ctx->d_wq.bytes() + ctx->d_ev7.bytes() + ctx->d_evec0.bytes();
    b += ctx->d_ephf.bytes() + ctx->d_evf.bytes() + ctx->d_gpf32.bytes();
    b += ctx->d_chain_e.bytes() + ctx->d_chain_g.bytes() + ctx->d_prow.bytes() + ctx->d_prowb.bytes();
    *bytes = b;
    return ALT3_OK;
}

int alt3_footprint(const alt3_ctx *ctx, int ntemps, int precision, long long *bytes)
{
    if (!ctx || !bytes || ntemps < 1) return fail(ALT3_E_ARG, "bad argument");
    if (precision < ALT3_V3_FP64_REF || precision > ALT3_V3_MIXED) return fail(ALT3_E_ARG, "unknown precision");
    // cells whose mixed tile does not fit in LDS run the FP64 path (v3_run): their footprint too
    if (precision == ALT3_V3_MIXED && !v3_mixed_fits(ctx)) precision = ALT3_V3_FP64;
    const long long nk = ctx->nk, nb = ctx->nb, ni = ctx->nibz, ndr = ctx->V.ndr, nt = std::min(ntemps, 4);
    const long long nel = ni * nb * nb * nb * nk, n = ni * nb * 3;
    long long b = 0;
    b += nk * nb * nb * 16 + nk * nb * 8 * 3 + nk * nb * 3 * 8 + nk * 3 * 8 * 2 + nk * 4 * 2;   // eigen data, meshes
    b += (long long)ctx->nt2h * 40 + (long long)ctx->V.nterm * 80 + nk * ndr * 16;             // FC tables, phases
    b += ni * ndr * nb * nb * nb * 16;                                                          // G'
    if (precision == ALT3_V3_MIXED) {
        const long long Kp = ((2 * ndr + 3) / 4) * 4, Np = ((2 * nb * nb + 15) / 16) * 16;
        b += nel * 4 + ni * nb * Kp * Np * 4 + nk * nb * nb * 8 + nk * ndr * 8;
        if (nb == 24) {                          // k_v3_chain packs (FP32) or k_v3_chain16's, the larger
            const long long k0 = Kp / 4, ks = k0 <= 8 ? 8 : k0 <= 12 ? 12 : k0 <= 14 ? 14 : 16;
            b += std::max(nk * 60 * 64 * 4 + ni * nb * ks * 36 * 64 * 4,
                          (long long)alt3::v3_chain16_pack_floats((int)nk, (int)ni) * 4);
        }
    } else {
        b += nel * 8;
    }
    b += 2 * nel * (precision == ALT3_V3_MIXED ? 4 : 8) + 2 * nel / nk;                        // P1sm, P2sm, row map
    if (nk > 8192) b += 1024 * 20 * nk;          // k_fgr_big scratch: 5 floats per mesh point per workgroup
    const long long rb = precision == ALT3_V3_MIXED ? 4 : 8;                                  // R in the precision of P
    b += nt * (nk * nb * 8 + 3 * ni * nb * 8 + 2 * ni * nb * nb * nk * rb + 3 * ni * nb * nb * 8);  // tau pass
    b += n * n * 8 + n * 8 * 8 + ni * nb * 9 * 8 * 2;                                           // CM, RHS, CG, kappa
    b += (long long)ni * 48 * 4 * 7 + 9 * 48 * 8 * 2;                                           // star tables
    b += 16ll << 20;                                                                            // small buffers
    *bytes = b;
    return ALT3_OK;
}

int alt3_sizes(const alt3_ctx *ctx, int *nk, int *nibz, int *nb, int *nops)
{
    if (!ctx) return fail(ALT3_E_ARG, "null context");
    if (nk) *nk = ctx->nk;
    if (nibz) *nibz = ctx->nibz;
    if (nb) *nb = ctx->nb;
    if (nops) *nops = ctx->nops;
    return ALT3_OK;
}

int alt3_get_mesh(const alt3_ctx *ctx, double *kpc_eig, double *kpc, int *mapibz, int *mapinv, double *kibz,
                  double *wibz, int *tet, double *opk, int *mappos, int *npos)
{
    if (!ctx) return fail(ALT3_E_ARG, "null context");
    const HostTables &T = ctx->T;
    if (kpc_eig) std::memcpy(kpc_eig, T.kpc_eig.data(), sizeof(double) * T.kpc_eig.size());
    if (kpc) std::memcpy(kpc, T.kpc.data(), sizeof(double) * T.kpc.size());
    if (mapibz) std::memcpy(mapibz, T.mapibz.data(), sizeof(int) * T.mapibz.size());
    if (mapinv) std::memcpy(mapinv, T.mapinv.data(), sizeof(int) * T.mapinv.size());
    if (kibz) std::memcpy(kibz, T.kibz.data(), sizeof(double) * T.kibz.size());
    if (wibz) std::memcpy(wibz, T.wibz.data(), sizeof(double) * T.wibz.size());
    if (tet) std::memcpy(tet, T.tet.data(), sizeof(int) * T.tet.size());
    if (opk) std::memcpy(opk, T.opk.data(), sizeof(double) * T.opk.size());
    if (mappos) std::memcpy(mappos, T.mappos.data(), sizeof(int) * T.mappos.size());
    if (npos) *npos = T.npos;
    return ALT3_OK;
}

int alt3_kstars(const alt3_ctx *ctx, int ibz, int *narms, int *ops, int *n0)
{
    if (!ctx || ibz < 0 || ibz >= ctx->nibz) return fail(ALT3_E_ARG, "bad ibz index");
    const HostTables &T = ctx->T;
    int n = T.star_n[ibz];
    if (narms) *narms = n;
    for (int s = 0; s < n; s++) {
        if (ops) ops[s] = T.star_op[48 * ibz + s];
        if (n0) n0[s] = T.star_n0[48 * ibz + s];
    }
    return ALT3_OK;
}

int alt3_eig(alt3_ctx *ctx, double *eigenval, double *eigenvec, double *veloc)
{
    if (!ctx) return fail(ALT3_E_ARG, "null context");
    if (ctx->host_only) return fail(ALT3_E_STATE, "host-only context has no device state");
    Timer tm(ctx);
    ALT3_TRY(eig_run(ctx));
    tm.stop();
    const size_t nk = ctx->nk, nb = ctx->nb;
    if (eigenval) ALT3_HIP(hipMemcpy(eigenval, ctx->d_eival.p, sizeof(double) * nk * nb, hipMemcpyDeviceToHost));
    if (eigenvec) ALT3_HIP(hipMemcpy(eigenvec, ctx->d_evec.p, sizeof(double) * nk * nb * nb * 2, hipMemcpyDeviceToHost));
    if (veloc) ALT3_HIP(hipMemcpy(veloc, ctx->d_vel.p, sizeof(double) * nk * nb * 3, hipMemcpyDeviceToHost));
    ctx->have_v3 = ctx->have_fgr = ctx->have_tau = ctx->have_cm = false;
    return ALT3_OK;
}

int alt3_set_eig(alt3_ctx *ctx, const double *eigenval, const double *eigenvec, const double *veloc)
{
    if (!ctx || !eigenval || !eigenvec || !veloc) return fail(ALT3_E_ARG, "null argument");
    if (ctx->host_only) return fail(ALT3_E_STATE, "host-only context has no device state");
    const size_t nk = ctx->nk, nb = ctx->nb;
    ALT3_TRY(ctx->d_eival.alloc(nk * nb));
    ALT3_TRY(ctx->d_freq.alloc(nk * nb));
    ALT3_TRY(ctx->d_vel.alloc(nk * nb * 3));
    ALT3_TRY(ctx->d_evec.alloc(nk * nb * nb * 2));
    std::vector<double> fr(nk * nb);
    for (size_t i = 0; i < nk * nb; i++) fr[i] = std::sqrt(eigenval[i]) * cnst();  // cal_freq
    ALT3_HIP(hipMemcpy(ctx->d_eival.p, eigenval, sizeof(double) * nk * nb, hipMemcpyHostToDevice));
    ALT3_HIP(hipMemcpy(ctx->d_freq.p, fr.data(), sizeof(double) * nk * nb, hipMemcpyHostToDevice));
    ALT3_HIP(hipMemcpy(ctx->d_evec.p, eigenvec, sizeof(double) * nk * nb * nb * 2, hipMemcpyHostToDevice));
    ALT3_HIP(hipMemcpy(ctx->d_vel.p, veloc, sizeof(double) * nk * nb * 3, hipMemcpyHostToDevice));
    ctx->have_eig = true;
    ctx->piso_ok = false;
    ctx->have_veloci = false;
    ctx->have_v3 = ctx->have_fgr = ctx->have_tau = ctx->have_cm = false;
    return ALT3_OK;
}

int alt3_v3sq_ibz(alt3_ctx *ctx, int ibz_lo, int ibz_hi, int precision, double *v33s8)
{
    if (!ctx) return fail(ALT3_E_ARG, "null context");
    if (ctx->host_only) return fail(ALT3_E_STATE, "host-only context has no device state");
    Timer tm(ctx);
    ALT3_TRY(v3_run(ctx, ibz_lo, ibz_hi, precision, true));
    tm.stop();
    if (v33s8) ALT3_TRY(v3_copy_out(ctx, v33s8));
    return ALT3_OK;
}

int alt3_get_v3sq(alt3_ctx *ctx, double *v33s8)
{
    if (!ctx || !v33s8) return fail(ALT3_E_ARG, "null argument");
    if (ctx->host_only) return fail(ALT3_E_STATE, "host-only context has no device state");
    if (!ctx->have_v3) return fail(ALT3_E_STATE, "no |V3|^2 held: call alt3_v3sq_ibz or alt3_v3_batch first");
    return v3_copy_out(ctx, v33s8);
}

int alt3_v3_batch(alt3_ctx *const *ctxs, int n, int precision, int tile, int *tile_used)
{
    if (!ctxs || n <= 0) return fail(ALT3_E_ARG, "bad argument");
    if (precision != ALT3_V3_FP64 && precision != ALT3_V3_MIXED)
        return fail(ALT3_E_ARG, "alt3_v3_batch: precision must be ALT3_V3_FP64 or ALT3_V3_MIXED");
    for (int i = 0; i < n; i++) {
        alt3_ctx *c = ctxs[i];
        if (!c) return fail(ALT3_E_ARG, "null context");
        if (c->host_only) return fail(ALT3_E_STATE, "host-only context has no device state");
        if (!c->have_eig) return fail(ALT3_E_STATE, "eigenvectors missing: call alt3_eig or alt3_set_eig first");
        if (c->nb != ctxs[0]->nb) return fail(ALT3_E_ARG, "alt3_v3_batch: every context must have the same branch count");
        if (c->device != ctxs[0]->device) return fail(ALT3_E_ARG, "alt3_v3_batch: every context must be on one device");
        for (int k = 0; k < i; k++)
            if (ctxs[k] == c) return fail(ALT3_E_ARG, "alt3_v3_batch: a context appears twice");
    }
    if (tile_used) *tile_used = 0;
    const bool fp32 = precision == ALT3_V3_MIXED;
    bool packed = v3_batch_supported(ctxs[0]);
    for (int i = 0; i < n && packed; i++) packed = v3_batch_supported(ctxs[i]);
    if (!packed) {
        // 24 branches (their own MFMA kernels fill the device per context) and cells without
        // FC3 pair vectors: one launch per context, each on its own stream, then one wait each
        for (int i = 0; i < n; i++) ALT3_TRY(v3_run(ctxs[i], 0, ctxs[i]->nibz - 1, precision, false));
        int bad = 0;
        for (int i = 0; i < n; i++) {
            int herr = 0;
            ALT3_HIP(hipMemcpyAsync(&herr, ctxs[i]->d_flags.p, sizeof(int), hipMemcpyDeviceToHost, ctxs[i]->stream));
            ALT3_HIP(hipStreamSynchronize(ctxs[i]->stream));
            if (herr) {
                ctxs[i]->have_v3 = false;
                bad++;
            }
        }
        if (bad) return fail(ALT3_E_DEN0, "MATRIX_ELT: den=0 (a zero eigenvalue in a V3 triple)");
        return ALT3_OK;
    }
    for (int i = 0; i < n; i++) {   // G', the per-mode scale and the FP32 copies, per context stream
        alt3_ctx *c = ctxs[i];
        const int ksub = c->nibz;
        const size_t nel = (size_t)ksub * c->nb * c->nb * c->nb * c->nk;
        c->have_v3 = c->have_fgr = c->have_tau = c->have_cm = false;
        c->ibz_lo = 0;
        c->ibz_hi = c->nibz - 1;
        c->v3_prec = c->v3_req = precision;
        c->ev_k_set[0] = false;
        if (fp32) {
            c->d_v33.release();
            ALT3_TRY(c->d_v33f.alloc(nel));
        } else {
            c->d_v33f.release();
            ALT3_TRY(c->d_v33.alloc(nel));
        }
        ALT3_HIP(hipMemsetAsync(c->d_flags.p, 0, sizeof(int), c->stream));
        ALT3_TRY(v3_prep_launch(c, ksub));
        ALT3_TRY(v3_rows_inputs(c, ksub, fp32, fp32));
    }
    for (int i = 0; i < n; i++) ALT3_HIP(hipStreamSynchronize(ctxs[i]->stream));
    int tq = tile;
    if (tq <= 0) ALT3_TRY(v3_tile_for(ctxs[0], fp32, &tq));
    ALT3_TRY(v3_batch_launch(ctxs, n, fp32, tq));   // on the first context's stream, waited for
    if (tile_used) *tile_used = tq;
    int bad = 0;
    for (int i = 0; i < n; i++) {
        int herr = 0;
        ALT3_HIP(hipMemcpy(&herr, ctxs[i]->d_flags.p, sizeof(int), hipMemcpyDeviceToHost));
        if (herr) bad++;
        else ctxs[i]->have_v3 = true;
    }
    if (bad) return fail(ALT3_E_DEN0, "MATRIX_ELT: den=0 (a zero eigenvalue in a V3 triple)");
    return ALT3_OK;
}

int alt3_fgr(alt3_ctx *ctx, double *P1sm, double *P2sm)
{
    if (!ctx) return fail(ALT3_E_ARG, "null context");
    if (ctx->host_only) return fail(ALT3_E_STATE, "host-only context has no device state");
    Timer tm(ctx);
    ALT3_TRY(fgr_run(ctx, true));
    tm.stop();
    ALT3_TRY(copy_p(ctx, P1sm, P2sm));
    return ALT3_OK;
}

int alt3_set_fgr(alt3_ctx *ctx, int ibz_lo, int ibz_hi, const double *P1sm, const double *P2sm)
{
    if (!ctx || !P1sm || !P2sm) return fail(ALT3_E_ARG, "null argument");
    if (ctx->host_only) return fail(ALT3_E_STATE, "host-only context has no device state");
    if (!ctx->have_eig) return fail(ALT3_E_STATE, "eigen data missing: call alt3_eig or alt3_set_eig first");
    if (ibz_lo < 0 || ibz_hi >= ctx->nibz || ibz_lo > ibz_hi) return fail(ALT3_E_ARG, "bad ibz range");
    const int ksub = ibz_hi - ibz_lo + 1;
    const size_t nel = (size_t)ksub * ctx->nb * ctx->nb * ctx->nb * ctx->nk;
    ctx->have_v3 = ctx->have_fgr = ctx->have_tau = ctx->have_cm = false;
    ctx->ibz_lo = ibz_lo;
    ctx->ibz_hi = ibz_hi;
    ctx->v3_prec = ctx->v3_req = ALT3_V3_FP64;   // P held in FP64; no |V3|^2 in the context
    ctx->p_fp32 = false;
    ctx->d_p1f.release();
    ctx->d_p2f.release();
    ALT3_TRY(ctx->d_p1.alloc(nel));
    ALT3_TRY(ctx->d_p2.alloc(nel));
    ALT3_TRY(ctx->d_prow.alloc(2 * nel / ctx->nk));
    ALT3_TRY(load_fortran5(ctx, ksub, P1sm, ctx->d_p1.p));
    ALT3_TRY(load_fortran5(ctx, ksub, P2sm, ctx->d_p2.p));
    ALT3_TRY(prow_all(ctx));
    ctx->have_fgr = true;
    ctx->fgr_kind = ALT3_DELTA_TETRA;
    return ALT3_OK;
}

int alt3_fgr_gauss(alt3_ctx *ctx, double sigma, double *P1sm, double *P2sm)
{
    if (!ctx) return fail(ALT3_E_ARG, "null context");
    if (ctx->host_only) return fail(ALT3_E_STATE, "host-only context has no device state");
    Timer tm(ctx);
    ALT3_TRY(fgr_gauss_run(ctx, sigma));
    tm.stop();
    ALT3_TRY(copy_p(ctx, P1sm, P2sm));
    return ALT3_OK;
}

int alt3_set_delta(alt3_ctx *ctx, int kind, double sigma)
{
    if (!ctx) return fail(ALT3_E_ARG, "null context");
    if (kind != ALT3_DELTA_TETRA && kind != ALT3_DELTA_GAUSS) return fail(ALT3_E_ARG, "unknown delta kind");
    if (kind == ALT3_DELTA_GAUSS && !(sigma > 0)) return fail(ALT3_E_ARG, "Gaussian width must be positive");
    ctx->delta_kind = kind;
    ctx->delta_sigma = sigma;
    return ALT3_OK;
}

int alt3_set_v3_mma(alt3_ctx *ctx, int kind)
{
    if (!ctx) return fail(ALT3_E_ARG, "null context");
    if (kind != ALT3_MMA_F16X3 && kind != ALT3_MMA_F32) return fail(ALT3_E_ARG, "unknown matrix-core kind");
    if (kind != ctx->v3_mma) ctx->have_v3 = ctx->have_fgr = ctx->have_tau = ctx->have_cm = false;
    ctx->v3_mma = kind;
    return ALT3_OK;
}

int alt3_tau(alt3_ctx *ctx, double tempk, double *Q, double *QN, double *QU, double *tauN, double *tauU, double *tautot)
{
    if (!ctx) return fail(ALT3_E_ARG, "null context");
    if (ctx->host_only) return fail(ALT3_E_STATE, "host-only context has no device state");
    Timer tm(ctx);
    ALT3_TRY(tau_run(ctx, tempk, Q, QN, QU, tauN, tauU, tautot));
    tm.stop();
    return ALT3_OK;
}

int alt3_kappa_rta(alt3_ctx *ctx, double tempk, double *kappa, double *FRTA, double *taueff)
{
    if (!ctx) return fail(ALT3_E_ARG, "null context");
    if (ctx->host_only) return fail(ALT3_E_STATE, "host-only context has no device state");
    Timer tm(ctx);
    ALT3_TRY(kappa_rta_run(ctx, tempk, kappa, FRTA, taueff, true));
    tm.stop();
    return ALT3_OK;
}

int alt3_collision(alt3_ctx *ctx, double tempk, double *CM, double *RHS)
{
    if (!ctx) return fail(ALT3_E_ARG, "null context");
    if (ctx->host_only) return fail(ALT3_E_STATE, "host-only context has no device state");
    Timer tm(ctx);
    ALT3_TRY(collision_run(ctx, tempk, CM, RHS, true));
    tm.stop();
    return ALT3_OK;
}

int alt3_cg(alt3_ctx *ctx, int n, const double *A, const double *b, double *x, int max_iter, int ndig, double *funcmin,
            int *iters)
{
    if (!ctx || !x || n <= 0) return fail(ALT3_E_ARG, "bad argument");
    if (ctx->host_only) return fail(ALT3_E_STATE, "host-only context has no device state");
    if ((A == nullptr) != (b == nullptr)) return fail(ALT3_E_ARG, "A and b must both be given or both be NULL");
    DBuf<double> dA, db, dx;
    const double *pA, *pb;
    if (!A) {
        if (!ctx->have_cm) return fail(ALT3_E_STATE, "alt3_cg(A=NULL) needs alt3_collision first");
        if (ctx->ibz_lo != 0 || ctx->ibz_hi != ctx->nibz - 1) return fail(ALT3_E_STATE, "collision matrix does not cover the IBZ");
        if (n != ctx->nibz * ctx->nb * 3) return fail(ALT3_E_ARG, "n must be nibz*nb*3");
        pA = ctx->d_cm.p;
        pb = ctx->d_rhs.p;
    } else {
        ALT3_TRY(dA.alloc((size_t)n * n));
        ALT3_TRY(db.alloc(n));
        ALT3_HIP(hipMemcpy(dA.p, A, sizeof(double) * n * n, hipMemcpyHostToDevice));
        ALT3_HIP(hipMemcpy(db.p, b, sizeof(double) * n, hipMemcpyHostToDevice));
        pA = dA.p;
        pb = db.p;
    }
    ALT3_TRY(dx.alloc(n));
    ALT3_HIP(hipMemcpy(dx.p, x, sizeof(double) * n, hipMemcpyHostToDevice));
    Timer tm(ctx);
    ALT3_TRY(cg_device(ctx, n, pA, pb, dx.p, max_iter, ndig, funcmin, iters, A == nullptr,
                       A == nullptr && ctx->cm_fp32 ? ctx->d_cmf.p : nullptr));
    tm.stop();
    ALT3_HIP(hipMemcpy(x, dx.p, sizeof(double) * n, hipMemcpyDeviceToHost));
    return ALT3_OK;
}

int alt3_kappa_bte(alt3_ctx *ctx, double tempk, const double *x, double *kappa, double *taueff)
{
    if (!ctx || !x) return fail(ALT3_E_ARG, "null argument");
    if (ctx->host_only) return fail(ALT3_E_STATE, "host-only context has no device state");
    const int n = ctx->nibz * ctx->nb * 3;
    DBuf<double> dx, dte;
    ALT3_TRY(dx.alloc(n));
    ALT3_TRY(dte.alloc((size_t)ctx->nibz * ctx->nb));
    ALT3_HIP(hipMemcpy(dx.p, x, sizeof(double) * n, hipMemcpyHostToDevice));
    if (taueff)
        ALT3_HIP(hipMemcpy(dte.p, taueff, sizeof(double) * ctx->nibz * ctx->nb, hipMemcpyHostToDevice));
    else
        ALT3_HIP(hipMemset(dte.p, 0, sizeof(double) * ctx->nibz * ctx->nb));
    Timer tm(ctx);
    ALT3_TRY(kappa_bte_run(ctx, tempk, dx.p, kappa, dte.p, true));
    tm.stop();
    if (taueff) ALT3_HIP(hipMemcpy(taueff, dte.p, sizeof(double) * ctx->nibz * ctx->nb, hipMemcpyDeviceToHost));
    return ALT3_OK;
}

int alt3_update(alt3_ctx *ctx, double tempk, const double *FFF, const double *FRTA, double *F2)
{
    if (!ctx || !FFF || !FRTA || !F2) return fail(ALT3_E_ARG, "null argument");
    if (ctx->host_only) return fail(ALT3_E_STATE, "host-only context has no device state");
    Timer tm(ctx);
    ALT3_TRY(update_run(ctx, tempk, FFF, FRTA, F2));
    tm.stop();
    return ALT3_OK;
}

int alt3_set_iso(alt3_ctx *ctx, const double *giso, double *Piso)
{
    if (!ctx) return fail(ALT3_E_ARG, "null context");
    if (ctx->host_only) return fail(ALT3_E_STATE, "host-only context has no device state");
    ctx->have_tau = ctx->have_cm = ctx->have_upd = false;
    ctx->piso_ok = false;
    if (!giso) {
        ctx->giso.clear();
        if (Piso) return fail(ALT3_E_ARG, "Piso needs giso");
        return ALT3_OK;
    }
    for (int a = 0; a < ctx->natoms0; a++)
        if (!(giso[a] >= 0)) return fail(ALT3_E_ARG, "giso must be >= 0");
    ctx->giso.assign(giso, giso + ctx->natoms0);
    if (!Piso) return ALT3_OK;
    Timer tm(ctx);
    ALT3_TRY(piso_run(ctx));
    tm.stop();
    const size_t nibz = ctx->nibz, nb = ctx->nb;
    std::vector<double> h(nibz * nb);
    ALT3_HIP(hipMemcpy(h.data(), ctx->d_piso.p, sizeof(double) * nibz * nb, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < nibz; i++)
        for (size_t l = 0; l < nb; l++) Piso[i + nibz * l] = h[i * nb + l];
    return ALT3_OK;
}

int alt3_self_w3(alt3_ctx *ctx, int nq, const double *q, double tempk, double eta, double v3_threshold, double *omega0,
                 double *nself, double *uself)
{
    if (!ctx) return fail(ALT3_E_ARG, "null context");
    if (ctx->host_only) return fail(ALT3_E_STATE, "host-only context has no device state");
    Timer tm(ctx);
    ALT3_TRY(self_run(ctx, nq, q, tempk, eta, v3_threshold, 0, omega0, nself, uself));
    tm.stop();
    return ALT3_OK;
}

int alt3_self_w4(alt3_ctx *ctx, int nq, const double *q, double tempk, double eta, double v3_threshold, double *omega0,
                 double *nself, double *uself)
{
    if (!ctx) return fail(ALT3_E_ARG, "null context");
    if (ctx->host_only) return fail(ALT3_E_STATE, "host-only context has no device state");
    if (!(eta > 0)) return fail(ALT3_E_ARG, "the Gaussian width eta must be positive");
    Timer tm(ctx);
    ALT3_TRY(self_run(ctx, nq, q, tempk, eta, v3_threshold, 1, omega0, nself, uself));
    tm.stop();
    return ALT3_OK;
}

int alt3_kappa_iter(alt3_ctx *ctx, double tempk, int max_iter, double update_mix, double tol, double *kappa,
                    double *kappa_hist, double *taueff_hist, double *FFF, int *iters)
{
    if (!ctx) return fail(ALT3_E_ARG, "null context");
    if (ctx->host_only) return fail(ALT3_E_STATE, "host-only context has no device state");
    Timer tm(ctx);
    ALT3_TRY(kappa_iter_run(ctx, tempk, max_iter, update_mix, tol, kappa, kappa_hist, taueff_hist, FFF, iters));
    tm.stop();
    return ALT3_OK;
}

int alt3_solve(alt3_ctx *ctx, int ntemps, const double *temps, int precision, double *kappa_rta, double *kappa_bte,
               int *cg_iters, double *stage_ms)
{
    if (!ctx || ntemps < 1 || !temps) return fail(ALT3_E_ARG, "bad argument");
    if (ctx->host_only) return fail(ALT3_E_STATE, "host-only context has no device state");
    double ms[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    const int nb = ctx->nb, nibz = ctx->nibz, n = nibz * nb * 3;
    // Stage times come from events recorded on the stream and read once the last stage is
    // done, so the temperature loop runs without waiting on the host (only CG looks at its
    // convergence flag); the N_BE flags of tau, kappa_RTA and kappa_BTE are read at the end.
    struct Mark {
        int slot;
        hipEvent_t a, b;
    };
    std::vector<Mark> marks;
    auto event = [&](size_t i, hipEvent_t *e) -> int {
        while (ctx->evpool.size() <= i) {
            hipEvent_t x;
            ALT3_HIP(hipEventCreate(&x));
            ctx->evpool.push_back(x);
        }
        *e = ctx->evpool[i];
        return ALT3_OK;
    };
    auto stage = [&](int slot, auto &&fn) -> int {
        Mark m{slot, nullptr, nullptr};
        ALT3_TRY(event(2 * marks.size(), &m.a));
        ALT3_TRY(event(2 * marks.size() + 1, &m.b));
        ALT3_HIP(hipEventRecord(m.a, ctx->stream));
        const int rc = fn();
        ALT3_HIP(hipEventRecord(m.b, ctx->stream));
        marks.push_back(m);
        return rc;
    };
    if (!ctx->have_eig) ALT3_TRY(stage(0, [&] { return eig_run(ctx); }));
    // |V3|^2 and the scattering weights do not depend on temperature: a context that still
    // holds them for the whole IBZ at this precision and delta integration reuses them
    const bool v3_held = ctx->have_v3 && ctx->v3_req == precision && ctx->ibz_lo == 0 && ctx->ibz_hi == nibz - 1;
    if (!v3_held) {
        ALT3_HIP(hipMemsetAsync(ctx->d_flags.p, 0, sizeof(int), ctx->stream));   // den = 0 flag
        ALT3_TRY(stage(1, [&] { return v3_run(ctx, 0, nibz - 1, precision, false); }));
    }
    const bool fgr_held = v3_held && ctx->have_fgr && ctx->fgr_kind == ctx->delta_kind &&
                          (ctx->delta_kind == ALT3_DELTA_TETRA || ctx->fgr_sigma == ctx->delta_sigma);
    if (!fgr_held)
        ALT3_TRY(stage(2, [&] {
            return ctx->delta_kind == ALT3_DELTA_GAUSS ? fgr_gauss_run(ctx, ctx->delta_sigma) : fgr_run(ctx, false);
        }));
    DBuf<double> &dx = ctx->d_solve_x, &dte = ctx->d_solve_te, &dk = ctx->d_solve_k;
    ALT3_TRY(dx.reserve(n));
    ALT3_TRY(dte.reserve((size_t)nibz * nb));
    ALT3_TRY(dk.reserve((size_t)2 * ntemps * nb * 9));
    ALT3_HIP(hipMemsetAsync(ctx->d_flags.p + 1, 0, 3 * sizeof(int), ctx->stream));   // N_BE flags
    const size_t kb = (size_t)nb * 9 * sizeof(double);
    for (int it = 0; it < ntemps; it++) {
        double T = temps[it];
        // one tau pass (one read of P1sm/P2sm) per group of four temperatures
        if (it % 4 == 0) {
            const int g = std::min(4, ntemps - it);
            ALT3_TRY(stage(3, [&] { return tau_multi_run(ctx, temps + it, g, false); }));
        }
        ALT3_TRY(stage(4, [&] { return kappa_rta_run(ctx, T, nullptr, nullptr, nullptr, false); }));
        ALT3_HIP(hipMemcpyAsync(dk.p + (size_t)2 * it * nb * 9, ctx->d_kappa.p, kb, hipMemcpyDeviceToDevice, ctx->stream));
        ALT3_TRY(stage(5, [&] { return collision_run(ctx, T, nullptr, nullptr, false); }));
        // x0 = F_RTA flattened (i, l, alpha), alpha fastest (kappa.f90:835-843), on the device
        ALT3_TRY(cg_start_launch(ctx, dx.p, dte.p));
        int its = 0;
        ALT3_TRY(stage(6, [&] {
            return cg_device(ctx, n, ctx->d_cm.p, ctx->d_rhs.p, dx.p, 100, 2, nullptr, &its, true,
                             ctx->cm_fp32 ? ctx->d_cmf.p : nullptr);
        }));
        if (cg_iters) cg_iters[it] = its;
        ALT3_TRY(stage(7, [&] { return kappa_bte_run(ctx, T, dx.p, nullptr, dte.p, false); }));
        ALT3_HIP(hipMemcpyAsync(dk.p + (size_t)(2 * it + 1) * nb * 9, ctx->d_kappa.p, kb, hipMemcpyDeviceToDevice,
                                ctx->stream));
    }
    int flags[4] = {0, 0, 0, 0};
    ALT3_HIP(hipMemcpyAsync(flags, ctx->d_flags.p, sizeof(flags), hipMemcpyDeviceToHost, ctx->stream));
    ALT3_HIP(hipStreamSynchronize(ctx->stream));
    if (flags[0] && !v3_held) {   // what was built from a den = 0 triple is not kept
        ctx->have_v3 = ctx->have_fgr = ctx->have_tau = ctx->have_cm = false;
        return fail(ALT3_E_DEN0, "MATRIX_ELT: den=0 (a zero eigenvalue in a V3 triple)");
    }
    if (flags[1]) return fail(ALT3_E_NBE, "ERROR in N_BE: w=<0 (a zero or imaginary mode frequency)");
    if (flags[2] || flags[3]) return fail(ALT3_E_NBE, "ERROR in N_BE: w=<0 (a zero frequency on a star member)");
    std::vector<double> hk((size_t)2 * ntemps * nb * 9);
    ALT3_HIP(hipMemcpy(hk.data(), dk.p, sizeof(double) * hk.size(), hipMemcpyDeviceToHost));
    for (int it = 0; it < ntemps; it++) {
        if (kappa_rta) std::memcpy(kappa_rta + (size_t)it * nb * 9, hk.data() + (size_t)2 * it * nb * 9, kb);
        if (kappa_bte) std::memcpy(kappa_bte + (size_t)it * nb * 9, hk.data() + (size_t)(2 * it + 1) * nb * 9, kb);
    }
    double total = 0;
    for (const Mark &m : marks) {
        float t = 0;
        ALT3_HIP(hipEventElapsedTime(&t, m.a, m.b));
        ms[m.slot] += t;
        total += t;
    }
    ctx->last_ms = total;
    if (stage_ms) std::memcpy(stage_ms, ms, sizeof(ms));
    return ALT3_OK;
}

int alt3_kernel_ms(const alt3_ctx *ctx, double *v3_ms, double *fgr_ms)
{
    if (!ctx) return fail(ALT3_E_ARG, "null context");
    if (ctx->host_only) return fail(ALT3_E_STATE, "host-only context has no device state");
    double *out[2] = {v3_ms, fgr_ms};
    for (int k = 0; k < 2; k++) {
        if (!out[k]) continue;
        *out[k] = -1;
        if (!ctx->ev_k_set[k]) continue;
        ALT3_HIP(hipEventSynchronize(ctx->ev_k[2 * k + 1]));
        float ms = 0;
        ALT3_HIP(hipEventElapsedTime(&ms, ctx->ev_k[2 * k], ctx->ev_k[2 * k + 1]));
        *out[k] = ms;
    }
    return ALT3_OK;
}

int alt3_last_kernel_ms(const alt3_ctx *ctx, double *ms)
{
    if (!ctx || !ms) return fail(ALT3_E_ARG, "null argument");
    *ms = ctx->last_ms;
    return ALT3_OK;
}

}  // extern "C"
