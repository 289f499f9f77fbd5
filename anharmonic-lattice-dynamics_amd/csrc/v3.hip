// |V3(q1 l1, q2 l2, q3 l3)|^2 on the GPU, q1 in the irreducible zone, q2 on the mesh,
// q3 = -q1-q2.  Replaces calculate_w3_ibz_split_sq (THERMACOND/sub.f90:1050-1233) and
// matrix_elt_simplified (sub.f90:5888-6011):
//
//   V3 = const33 / sqrt(8 sqrt|l1 l2 l3|) * sum_t psi_t e^{i q2.Rj_t} e^{i q3.Rk_t} e1[a_t] e2[b_t] e3[c_t]
//
// Three arithmetic paths (ALT3_V3_*):
//  FP64_REF  the reference algorithm: every element sums all FC3 terms (one workgroup per
//            (q1,q2), one thread per branch triple).  Kept as the drop-in baseline.
//  FP64      batched form.  Because q3 = -q1-q2 the phase factorises,
//              e^{i q2.Rj} e^{i q3.Rk} = e^{i q2.(Rj-Rk)} e^{-i q1.Rk},
//            so per q1 a small tensor G'[dR][l1][b,c] = sum_t psi e^{-i q1.Rk} e1[a,l1]
//            (k_v3_prep) turns the term sum into a phase GEMM over dR for a whole tile of
//            q2, followed by two nb x nb contractions with e3 and e2 (k_v3_main).
//  MIXED     the same with the phase GEMM on FP32 MFMA (see v3_mixed.hip).
#include "alt3_impl.h"

namespace {

struct zd {
    double r, i;
};

__device__ __forceinline__ int mesh_index(const alt3::MeshDims &d, int c0, int c1, int c2)
{
    return c2 + d.n2 * (c1 + d.n1 * c0);
}
__device__ __forceinline__ void mesh_coords(const alt3::MeshDims &d, int k, int &c0, int &c1, int &c2)
{
    c2 = k % d.n2;
    c1 = (k / d.n2) % d.n1;
    c0 = k / (d.n1 * d.n2);
}
__device__ __forceinline__ int wrap(int x, int n)
{
    x %= n;
    return x < 0 ? x + n : x;
}
// mesh index of -q1-q2 (reduced coordinates c/n - 1/2 on even meshes)
__device__ __forceinline__ int mesh_neg_sum(const alt3::MeshDims &d, int k1, int k2)
{
    int a0, a1, a2, b0, b1, b2;
    mesh_coords(d, k1, a0, a1, a2);
    mesh_coords(d, k2, b0, b1, b2);
    // -(u1+u2) = -(c1+c2)/n + 1  ->  c = (3n/2 - c1 - c2) mod n
    return mesh_index(d, wrap(3 * d.n0 / 2 - a0 - b0, d.n0), wrap(3 * d.n1 / 2 - a1 - b1, d.n1),
                      wrap(3 * d.n2 / 2 - a2 - b2, d.n2));
}

// ------------------------------------------------------------------ reference algorithm
// one workgroup per (q1, q2); threads over (l1,l2,l3); shared per-term phase factors
__global__ void __launch_bounds__(256) k_v3_ref(alt3::MeshDims md, int nk, int nb, int ibz_lo,
                                                const int *__restrict__ mapinv, const double *__restrict__ kpc,
                                                const double *__restrict__ evec, const double *__restrict__ eival,
                                                int nterm, const int *__restrict__ t_abc,
                                                const double *__restrict__ t_psi, const double *__restrict__ t_rj,
                                                const double *__restrict__ t_rk, double c33,
                                                double *__restrict__ out, int *__restrict__ err)
{
    const int s = blockIdx.y, n2 = blockIdx.x;
    const int k1 = mapinv[ibz_lo + s];
    const int k3 = mesh_neg_sum(md, k1, n2);
    constexpr int TT = 256;
    __shared__ double ph[2 * TT];
    __shared__ int abc[3 * TT];
    const int nel = nb * nb * nb;
    const int e = threadIdx.x;
    int l1 = 0, l2 = 0, l3 = 0;
    if (e < nel) {
        l1 = e / (nb * nb);
        l2 = (e / nb) % nb;
        l3 = e % nb;
    }
    const double *e1 = evec + 2 * ((size_t)k1 * nb + l1) * nb;
    const double *e2 = evec + 2 * ((size_t)n2 * nb + l2) * nb;
    const double *e3 = evec + 2 * ((size_t)k3 * nb + l3) * nb;
    const double *q2 = kpc + 3 * n2, *q3 = kpc + 3 * k3;
    double xr = 0, xi = 0;
    for (int t0 = 0; t0 < nterm; t0 += TT) {
        int t = t0 + threadIdx.x;
        if (t < nterm) {
            // psi * eiqr2(q2,j) * eiqr2(q3,k) (cal_eiqr2, sub.f90:6015-6067)
            double a2 = q2[0] * t_rj[3 * t] + q2[1] * t_rj[3 * t + 1] + q2[2] * t_rj[3 * t + 2];
            double a3 = q3[0] * t_rk[3 * t] + q3[1] * t_rk[3 * t + 1] + q3[2] * t_rk[3 * t + 2];
            double s2, c2, s3, c3;
            sincos(a2, &s2, &c2);
            sincos(a3, &s3, &c3);
            double pr = t_psi[t] * c2, pi = t_psi[t] * s2;
            ph[2 * threadIdx.x] = pr * c3 - pi * s3;
            ph[2 * threadIdx.x + 1] = pr * s3 + pi * c3;
            abc[3 * threadIdx.x] = t_abc[3 * t];
            abc[3 * threadIdx.x + 1] = t_abc[3 * t + 1];
            abc[3 * threadIdx.x + 2] = t_abc[3 * t + 2];
        }
        __syncthreads();
        int tn = min(TT, nterm - t0);
        if (e < nel)
            for (int u = 0; u < tn; u++) {
                int a = abc[3 * u], b = abc[3 * u + 1], c = abc[3 * u + 2];
                double ar = e1[2 * a], ai = e1[2 * a + 1], br = e2[2 * b], bi = e2[2 * b + 1];
                double cr = e3[2 * c], ci = e3[2 * c + 1];
                double abr = ar * br - ai * bi, abi = ar * bi + ai * br;
                double er = abr * cr - abi * ci, ei = abr * ci + abi * cr;
                double pr = ph[2 * u], pi = ph[2 * u + 1];
                xr += pr * er - pi * ei;
                xi += pr * ei + pi * er;
            }
        __syncthreads();
    }
    if (e < nel) {
        double den = sqrt(8 * sqrt(fabs(eival[(size_t)k1 * nb + l1] * eival[(size_t)n2 * nb + l2] * eival[(size_t)k3 * nb + l3])));
        if (den == 0) {
            atomicExch(err, 1);
            den = 1;
        }
        double wr = xr / den * c33, wi = xi / den * c33;
        out[(((size_t)s * nb + l1) * nb + l2) * (size_t)nb * nk + (size_t)l3 * nk + n2] = wr * wr + wi * wi;
    }
}

// ------------------------------------------------------------------ batched path
// G'[s][dr][l1][bc] = sum_{t in slot(dr,bc)} psi_t e^{-i q1.Rk_t} e1[a_t][l1]
__global__ void k_v3_prep(int nb, int ibz_lo, int nslot, int ndr, const int *__restrict__ mapinv,
                          const double *__restrict__ kpc, const double *__restrict__ evec,
                          const int *__restrict__ slot_dr, const int *__restrict__ slot_bc,
                          const int *__restrict__ slot_ptr, const int *__restrict__ sterm_a,
                          const double *__restrict__ sterm_psi, const double *__restrict__ sterm_rk,
                          const double *__restrict__ q1pts, const double *__restrict__ e1pts,
                          double *__restrict__ gp)
{
    const int s = blockIdx.y;
    const int id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= nslot * nb) return;
    const int slot = id / nb, l1 = id % nb;
    // q1 = the irreducible mesh point s, or (q1pts given) the point s of a list with its own
    // eigenvectors e1pts [s][nb col][nb]
    const int k1 = q1pts ? 0 : mapinv[ibz_lo + s];
    const double *q1 = q1pts ? q1pts + 3 * s : kpc + 3 * k1;
    const double *e1 = q1pts ? e1pts + 2 * ((size_t)s * nb + l1) * nb : evec + 2 * ((size_t)k1 * nb + l1) * nb;
    double gr = 0, gi = 0;
    for (int t = slot_ptr[slot]; t < slot_ptr[slot + 1]; t++) {
        double a = -(q1[0] * sterm_rk[3 * t] + q1[1] * sterm_rk[3 * t + 1] + q1[2] * sterm_rk[3 * t + 2]);
        double sn, cs;
        sincos(a, &sn, &cs);
        double pr = sterm_psi[t] * cs, pi = sterm_psi[t] * sn;
        int ai = sterm_a[t];
        double er = e1[2 * ai], ei = e1[2 * ai + 1];
        gr += pr * er - pi * ei;
        gi += pr * ei + pi * er;
    }
    size_t o = ((((size_t)s * ndr + slot_dr[slot]) * nb + l1) * nb * nb + slot_bc[slot]);
    gp[2 * o] = gr;
    gp[2 * o + 1] = gi;
}

// e^{i q2.dR_r} for every mesh point and cell-pair difference (independent of q1)
__global__ void k_phase_table(int nk, int ndr, const double *__restrict__ kpc, const double *__restrict__ dr,
                              double *__restrict__ eph)
{
    int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= nk * ndr) return;
    int n2 = x / ndr, r = x % ndr;
    double a = kpc[3 * n2] * dr[3 * r] + kpc[3 * n2 + 1] * dr[3 * r + 1] + kpc[3 * n2 + 2] * dr[3 * r + 2];
    double sn, cs;
    sincos(a, &sn, &cs);
    eph[2 * (size_t)x] = cs;
    eph[2 * (size_t)x + 1] = sn;
}

// one workgroup per (q1 s, l1, tile of TQ q2); 256 threads
template <int TQ>
__global__ void __launch_bounds__(256) k_v3_main(alt3::MeshDims md, int nk, int nb, int ibz_lo, int ndr,
                                                 const int *__restrict__ mapinv, const double *__restrict__ eph,
                                                 const double *__restrict__ evec,
                                                 const double *__restrict__ eival, const double *__restrict__ gp,
                                                 double c33, double *__restrict__ out, int *__restrict__ err)
{
    extern __shared__ double lds[];
    const int s = blockIdx.z, l1 = blockIdx.y, q0 = blockIdx.x * TQ;
    const int nb2 = nb * nb;
    const int k1 = mapinv[ibz_lo + s];
    zd *E = reinterpret_cast<zd *>(lds);  // [TQ][ndr]
    zd *Phi = E + TQ * ndr;               // [TQ][nb2]   Phi'[b][c]
    zd *U = Phi + TQ * nb2;               // [TQ][nb2]   U[b][l3]
    __shared__ int k3s[TQ];
    const int tid = threadIdx.x;
    if (tid < TQ) k3s[tid] = (q0 + tid < nk) ? mesh_neg_sum(md, k1, q0 + tid) : 0;
    for (int x = tid; x < TQ * ndr; x += 256) {
        int qq = x / ndr, r = x % ndr;
        int n2 = min(q0 + qq, nk - 1);
        E[x] = {eph[2 * ((size_t)n2 * ndr + r)], eph[2 * ((size_t)n2 * ndr + r) + 1]};
    }
    __syncthreads();
    // Phi'[qq][bc] = sum_r E[qq][r] G'[r][l1][bc]
    const double *g = gp + 2 * (((size_t)s * ndr) * nb + l1) * nb2;  // + r*nb*nb2
    for (int x = tid; x < TQ * nb2; x += 256) {
        int qq = x / nb2, bc = x % nb2;
        double pr = 0, pi = 0;
        for (int r = 0; r < ndr; r++) {
            const double *gg = g + 2 * ((size_t)r * nb * nb2 + bc);
            double gr = gg[0], gi = gg[1];
            zd e = E[qq * ndr + r];
            pr += e.r * gr - e.i * gi;
            pi += e.r * gi + e.i * gr;
        }
        Phi[x] = {pr, pi};
    }
    __syncthreads();
    // U[qq][b][l3] = sum_c Phi'[qq][b][c] e3[c][l3]
    for (int x = tid; x < TQ * nb2; x += 256) {
        int qq = x / nb2, b = (x / nb) % nb, l3 = x % nb;
        const double *e3 = evec + 2 * ((size_t)k3s[qq] * nb + l3) * nb;
        const zd *ph = Phi + qq * nb2 + b * nb;
        double ur = 0, ui = 0;
        for (int c = 0; c < nb; c++) {
            double er = e3[2 * c], ei = e3[2 * c + 1];
            ur += ph[c].r * er - ph[c].i * ei;
            ui += ph[c].r * ei + ph[c].i * er;
        }
        U[x] = {ur, ui};
    }
    __syncthreads();
    // V[qq][l2][l3] = sum_b e2[b][l2] U[qq][b][l3]; out[s][l1][l2][l3][n2]
    const double lam1 = eival[(size_t)k1 * nb + l1];
    for (int x = tid; x < TQ * nb2; x += 256) {
        int l2 = x / (TQ * nb), rem = x % (TQ * nb), l3 = rem / TQ, qq = rem % TQ;
        int n2 = q0 + qq;
        if (n2 >= nk) continue;
        const double *e2 = evec + 2 * ((size_t)n2 * nb + l2) * nb;
        double vr = 0, vi = 0;
        for (int b = 0; b < nb; b++) {
            zd u = U[qq * nb2 + b * nb + l3];
            double er = e2[2 * b], ei = e2[2 * b + 1];
            vr += er * u.r - ei * u.i;
            vi += er * u.i + ei * u.r;
        }
        double den = sqrt(8 * sqrt(fabs(lam1 * eival[(size_t)n2 * nb + l2] * eival[(size_t)k3s[qq] * nb + l3])));
        if (den == 0) {
            atomicExch(err, 1);
            den = 1;
        }
        double wr = vr / den * c33, wi = vi / den * c33;
        out[(((size_t)s * nb + l1) * nb + l2) * (size_t)nb * nk + (size_t)l3 * nk + n2] = wr * wr + wi * wi;
    }
}

// [s][l1][l2][l3][n2] -> Fortran v33s8(ksub,nk,nb,nb,nb)
template <typename T>
// rmap (or NULL): the P row map, 0 for rows that are zero and not stored
__global__ void k_to_fortran5(int ksub, int nk, int nb, const T *__restrict__ in, double *__restrict__ out,
                              const unsigned char *__restrict__ rmap = nullptr)
{
    size_t tot = (size_t)ksub * nk * nb * nb * nb;
    size_t o = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= tot) return;
    size_t x = o;
    int s = x % ksub; x /= ksub;
    int n2 = x % nk; x /= nk;
    int l1 = x % nb; x /= nb;
    int l2 = x % nb; x /= nb;
    int l3 = (int)x;
    const size_t row = (((size_t)s * nb + l1) * nb + l2) * nb + l3;
    out[o] = (rmap && !rmap[row]) ? 0.0 : (double)in[row * nk + n2];
}

// Fortran (ksub,nk,nb,nb,nb) -> [s][l1][l2][l3][n2] (the inverse of k_to_fortran5)
__global__ void k_from_fortran5(int ksub, int nk, int nb, const double *__restrict__ in, double *__restrict__ out)
{
    size_t tot = (size_t)ksub * nk * nb * nb * nb;
    size_t o = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= tot) return;
    size_t x = o;
    int s = x % ksub; x /= ksub;
    int n2 = x % nk; x /= nk;
    int l1 = x % nb; x /= nb;
    int l2 = x % nb; x /= nb;
    int l3 = (int)x;
    out[((((size_t)s * nb + l1) * nb + l2) * nb + l3) * nk + n2] = in[o];
}

}  // namespace

namespace alt3 {

int v3_mixed_run(alt3_ctx *ctx, int ksub, int *d_err);  // v3_mixed.hip
bool v3_mixed_fits(const alt3_ctx *ctx);                 // v3_mixed.hip

int v3_prepare(alt3_ctx *ctx)
{
    V3Tables &v = ctx->V;
    ALT3_TRY(ctx->d_t_abc.alloc(std::max(1, 3 * v.nterm)));
    ALT3_TRY(ctx->d_t_psi.alloc(std::max(1, v.nterm)));
    ALT3_TRY(ctx->d_t_rj.alloc(std::max(1, 3 * v.nterm)));
    ALT3_TRY(ctx->d_t_rk.alloc(std::max(1, 3 * v.nterm)));
    ALT3_TRY(ctx->d_dr.alloc(std::max(1, 3 * v.ndr)));
    ALT3_TRY(ctx->d_slot_dr.alloc(std::max(1, v.nslot)));
    ALT3_TRY(ctx->d_slot_bc.alloc(std::max(1, v.nslot)));
    ALT3_TRY(ctx->d_slot_ptr.alloc(v.nslot + 1));
    ALT3_TRY(ctx->d_sterm_a.alloc(std::max(1, v.nterm)));
    ALT3_TRY(ctx->d_sterm_psi.alloc(std::max(1, v.nterm)));
    ALT3_TRY(ctx->d_sterm_rk.alloc(std::max(1, 3 * v.nterm)));
    auto up = [&](void *d, const void *h, size_t bytes) -> int {
        if (bytes) ALT3_HIP(hipMemcpy(d, h, bytes, hipMemcpyHostToDevice));
        return ALT3_OK;
    };
    ALT3_TRY(up(ctx->d_t_abc.p, v.t_abc.data(), sizeof(int) * v.t_abc.size()));
    ALT3_TRY(up(ctx->d_t_psi.p, v.t_psi.data(), sizeof(double) * v.t_psi.size()));
    ALT3_TRY(up(ctx->d_t_rj.p, v.t_rj.data(), sizeof(double) * v.t_rj.size()));
    ALT3_TRY(up(ctx->d_t_rk.p, v.t_rk.data(), sizeof(double) * v.t_rk.size()));
    ALT3_TRY(up(ctx->d_dr.p, v.dr.data(), sizeof(double) * v.dr.size()));
    ALT3_TRY(up(ctx->d_slot_dr.p, v.slot_dr.data(), sizeof(int) * v.slot_dr.size()));
    ALT3_TRY(up(ctx->d_slot_bc.p, v.slot_bc.data(), sizeof(int) * v.slot_bc.size()));
    ALT3_TRY(up(ctx->d_slot_ptr.p, v.slot_ptr.data(), sizeof(int) * v.slot_ptr.size()));
    ALT3_TRY(up(ctx->d_sterm_a.p, v.sterm_a.data(), sizeof(int) * v.sterm_a.size()));
    ALT3_TRY(up(ctx->d_sterm_psi.p, v.sterm_psi.data(), sizeof(double) * v.sterm_psi.size()));
    ALT3_TRY(up(ctx->d_sterm_rk.p, v.sterm_rk.data(), sizeof(double) * v.sterm_rk.size()));
    if (v.ndr > 0) {
        ALT3_TRY(ctx->d_eph.alloc((size_t)ctx->nk * v.ndr * 2));
        int n = ctx->nk * v.ndr;
        hipLaunchKernelGGL(k_phase_table, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, ctx->nk, v.ndr, ctx->d_kpc.p,
                           ctx->d_dr.p, ctx->d_eph.p);
        ALT3_HIP(hipGetLastError());
        ALT3_HIP(hipStreamSynchronize(ctx->stream));
    }
    return ALT3_OK;
}

// G'[s][dr][l1][bc] for the current ibz range (zeroed first: empty slots stay 0)
int v3_prep_launch(alt3_ctx *ctx, int ksub)
{
    const int nb = ctx->nb, ndr = ctx->V.ndr, nb2 = nb * nb;
    ALT3_TRY(ctx->d_gp.alloc((size_t)ksub * ndr * nb * nb2 * 2));
    ALT3_HIP(hipMemsetAsync(ctx->d_gp.p, 0, ctx->d_gp.bytes(), ctx->stream));
    int nth = ctx->V.nslot * nb;
    if (nth > 0) {
        hipLaunchKernelGGL(k_v3_prep, dim3((nth + 127) / 128, ksub), dim3(128), 0, ctx->stream, nb, ctx->ibz_lo,
                           ctx->V.nslot, ndr, ctx->d_mapinv.p, ctx->d_kpc.p, ctx->d_evec.p, ctx->d_slot_dr.p,
                           ctx->d_slot_bc.p, ctx->d_slot_ptr.p, ctx->d_sterm_a.p, ctx->d_sterm_psi.p,
                           ctx->d_sterm_rk.p, nullptr, nullptr, ctx->d_gp.p);
        ALT3_HIP(hipGetLastError());
    }
    return ALT3_OK;
}

// G' of nq arbitrary points q1 (Cartesian) with eigenvectors e1 [q][nb col][nb] into gp
// [q][ndr][nb][nb*nb] complex
int v3_prep_points(alt3_ctx *ctx, int nq, const double *q1, const double *e1, double *gp)
{
    const int nb = ctx->nb, ndr = ctx->V.ndr;
    ALT3_HIP(hipMemsetAsync(gp, 0, sizeof(double) * 2 * (size_t)nq * ndr * nb * nb * nb, ctx->stream));
    const int nth = ctx->V.nslot * nb;
    if (nth > 0) {
        hipLaunchKernelGGL(k_v3_prep, dim3((nth + 127) / 128, nq), dim3(128), 0, ctx->stream, nb, 0, ctx->V.nslot, ndr,
                           ctx->d_mapinv.p, ctx->d_kpc.p, ctx->d_evec.p, ctx->d_slot_dr.p, ctx->d_slot_bc.p,
                           ctx->d_slot_ptr.p, ctx->d_sterm_a.p, ctx->d_sterm_psi.p, ctx->d_sterm_rk.p, q1, e1, gp);
        ALT3_HIP(hipGetLastError());
    }
    return ALT3_OK;
}

// sync = false (alt3_solve): no wait at the end; a den = 0 stays in the device flag
int v3_run(alt3_ctx *ctx, int ibz_lo, int ibz_hi, int prec, bool sync)
{
    if (!ctx->have_eig) return fail(ALT3_E_STATE, "eigenvectors missing: call alt3_eig or alt3_set_eig first");
    if (ibz_lo < 0 || ibz_hi >= ctx->nibz || ibz_lo > ibz_hi) return fail(ALT3_E_ARG, "bad ibz range");
    if (prec < ALT3_V3_FP64_REF || prec > ALT3_V3_MIXED) return fail(ALT3_E_ARG, "bad precision");
    const int nk = ctx->nk, nb = ctx->nb, ksub = ibz_hi - ibz_lo + 1;
    const size_t nel = (size_t)ksub * nb * nb * nb * nk;
    MeshDims md{ctx->nc[0], ctx->nc[1], ctx->nc[2]};
    int *derr = ctx->d_flags.p + 0;
    ALT3_HIP(hipMemsetAsync(derr, 0, sizeof(int), ctx->stream));
    ctx->have_v3 = ctx->have_fgr = ctx->have_tau = ctx->have_cm = false;
    ctx->ibz_lo = ibz_lo;
    ctx->ibz_hi = ibz_hi;
    ctx->v3_prec = prec;
    ctx->v3_req = prec;
    const double c33 = const33();
    if (prec == ALT3_V3_MIXED && !v3_mixed_fits(ctx)) {
        // the 16-point MFMA tile of the mixed path does not fit in LDS for this many
        // branches: the FP64 path computes |V3|^2 instead (and consumers read FP64)
        prec = ALT3_V3_FP64;
        ctx->v3_prec = prec;
    }
    if (prec != ALT3_V3_FP64_REF && v3_rows_supported(ctx, prec == ALT3_V3_MIXED)) {
        // 6, 12 or 24 branches: one workgroup per q2 tile handles every l1 (v3_rows.hip)
        const bool fp32 = prec == ALT3_V3_MIXED;
        if (fp32) {
            ctx->d_v33.release();
            ALT3_TRY(ctx->d_v33f.alloc(nel));
        } else {
            ctx->d_v33f.release();
            ALT3_TRY(ctx->d_v33.alloc(nel));
        }
        ALT3_TRY(v3_prep_launch(ctx, ksub));
        ALT3_TRY(v3_rows_run(ctx, ksub, fp32, derr));
    } else if (prec == ALT3_V3_MIXED) {
        ctx->d_v33.release();
        ALT3_TRY(ctx->d_v33f.alloc(nel));
        ALT3_TRY(v3_mixed_run(ctx, ksub, derr));
    } else {
        ctx->d_v33f.release();
        ALT3_TRY(ctx->d_v33.alloc(nel));
        if (prec == ALT3_V3_FP64_REF) {
            if (nb * nb * nb > 256) return fail(ALT3_E_ARG, "FP64_REF path supports nb^3 <= 256");
            hipLaunchKernelGGL(k_v3_ref, dim3(nk, ksub), dim3(256), 0, ctx->stream, md, nk, nb, ibz_lo,
                               ctx->d_mapinv.p, ctx->d_kpc.p, ctx->d_evec.p, ctx->d_eival.p, ctx->V.nterm,
                               ctx->d_t_abc.p, ctx->d_t_psi.p, ctx->d_t_rj.p, ctx->d_t_rk.p, c33, ctx->d_v33.p, derr);
            ALT3_HIP(hipGetLastError());
        } else {
            const int ndr = ctx->V.ndr, nb2 = nb * nb;
            ALT3_TRY(v3_prep_launch(ctx, ksub));
            // tile of q2 sized so the three LDS arrays stay within 64 KB (two groups per CU);
            // large cells take one q2 per group and up to the device's LDS per group
            auto need = [&](int t) { return (size_t)t * (ndr + 2 * nb2) * sizeof(zd); };
            int tq = 64;
            while (tq > 1 && need(tq) > 65536) tq /= 2;
            size_t lds = need(tq);
            if (lds > ctx->lds_max) return fail(ALT3_E_ARG, "V3 tile does not fit in LDS");
            dim3 grid((nk + tq - 1) / tq, nb, ksub);
#define LAUNCH(T)                                                                                              \
    if (lds > 65536)                                                                                           \
        ALT3_HIP(hipFuncSetAttribute((const void *)k_v3_main<T>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds)); \
    hipLaunchKernelGGL(k_v3_main<T>, grid, dim3(256), lds, ctx->stream, md, nk, nb, ibz_lo, ndr, ctx->d_mapinv.p, \
                       ctx->d_eph.p, ctx->d_evec.p, ctx->d_eival.p, ctx->d_gp.p, c33, ctx->d_v33.p, derr)
            kernel_mark(ctx, 0, false);
            switch (tq) {
                case 64: LAUNCH(64); break;
                case 32: LAUNCH(32); break;
                case 16: LAUNCH(16); break;
                case 8: LAUNCH(8); break;
                case 4: LAUNCH(4); break;
                case 2: LAUNCH(2); break;
                default: LAUNCH(1); break;
            }
#undef LAUNCH
            ALT3_HIP(hipGetLastError());
            kernel_mark(ctx, 0, true);
        }
    }
    if (sync) {
        int herr = 0;
        ALT3_HIP(hipMemcpyAsync(&herr, derr, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
        ALT3_HIP(hipStreamSynchronize(ctx->stream));
        if (herr) return fail(ALT3_E_DEN0, "MATRIX_ELT: den=0 (a zero eigenvalue in a V3 triple)");
    }
    ctx->have_v3 = true;
    return ALT3_OK;
}

int v3_copy_out(alt3_ctx *ctx, double *host)
{
    const int nk = ctx->nk, nb = ctx->nb, ksub = ctx->ibz_hi - ctx->ibz_lo + 1;
    size_t tot = (size_t)ksub * nk * nb * nb * nb;
    DBuf<double> tmp;
    ALT3_TRY(tmp.alloc(tot));
    if (ctx->v3_prec == ALT3_V3_MIXED)
        hipLaunchKernelGGL(k_to_fortran5<float>, dim3((tot + 255) / 256), dim3(256), 0, ctx->stream, ksub, nk, nb,
                           ctx->d_v33f.p, tmp.p);
    else
        hipLaunchKernelGGL(k_to_fortran5<double>, dim3((tot + 255) / 256), dim3(256), 0, ctx->stream, ksub, nk, nb,
                           ctx->d_v33.p, tmp.p);
    ALT3_HIP(hipGetLastError());
    ALT3_HIP(hipMemcpyAsync(host, tmp.p, tot * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    ALT3_HIP(hipStreamSynchronize(ctx->stream));
    return ALT3_OK;
}

template <typename T>
int copy_fortran5(alt3_ctx *ctx, const T *dev, double *host, const unsigned char *rmap)
{
    const int nk = ctx->nk, nb = ctx->nb, ksub = ctx->ibz_hi - ctx->ibz_lo + 1;
    size_t tot = (size_t)ksub * nk * nb * nb * nb;
    DBuf<double> tmp;
    ALT3_TRY(tmp.alloc(tot));
    hipLaunchKernelGGL(k_to_fortran5<T>, dim3((tot + 255) / 256), dim3(256), 0, ctx->stream, ksub, nk, nb, dev, tmp.p, rmap);
    ALT3_HIP(hipGetLastError());
    ALT3_HIP(hipMemcpyAsync(host, tmp.p, tot * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    ALT3_HIP(hipStreamSynchronize(ctx->stream));
    return ALT3_OK;
}

template int copy_fortran5<double>(alt3_ctx *, const double *, double *, const unsigned char *);

// host Fortran (ksub,nk,nb,nb,nb) into a device array of the kernels' layout
int load_fortran5(alt3_ctx *ctx, int ksub, const double *host, double *dev)
{
    const int nk = ctx->nk, nb = ctx->nb;
    size_t tot = (size_t)ksub * nk * nb * nb * nb;
    DBuf<double> tmp;
    ALT3_TRY(tmp.alloc(tot));
    ALT3_HIP(hipMemcpyAsync(tmp.p, host, tot * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(k_from_fortran5, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, ctx->stream, ksub, nk, nb,
                       tmp.p, dev);
    ALT3_HIP(hipGetLastError());
    ALT3_HIP(hipStreamSynchronize(ctx->stream));
    return ALT3_OK;
}
template int copy_fortran5<float>(alt3_ctx *, const float *, double *, const unsigned char *);

}  // namespace alt3
