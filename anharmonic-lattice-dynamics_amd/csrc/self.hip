// Three-phonon self-energy at arbitrary q on the GPU (FP64): jobs 2/6/7 of kappa.f90
// (kappa.f90:900-925) evaluate, for every branch la of a point q off the mesh,
//   function_self_w3 (rest.f90:4012-4124)  at omega = omega(q, la),
// a sum over the mesh k2 with k3 = -q-k2 (unfolded) of |V3(q la, k2 l2, k3 l3)|^2 times the
// Lorentzian response of the four processes, split into normal and umklapp parts by
// check_inside_bz (extratools.f90:325-341).  V3 is matrix_elt_full (rest.f90:3664-3719); the
// eigen data at q and at every k3 come from get_freq (rest.f90:4319-4445).
//
// Per batch of points: the eigensolves at q and at the nk points k3 (k_eig_group, one
// variant), G'(q) (k_v3_prep on the point list), |V3|^2 [s][l1][l2][l3][k2] by the batched
// kernel with e3 and w3 taken per (s, k2) (k_v3_rows<EXT>), then the Lorentzian sums
// (k_self_sum) and their ordered reduction over the k2 chunks (k_self_final).  The phase of a
// term, e^{i(k2.rr2 + k3.rr3)} with k3 = -q-k2, factors exactly into e^{i k2.(rr2-rr3)} of the
// mesh phase table and e^{-i q.rr3} of G'(q).
#include <algorithm>
#include <cmath>

#include "alt3_impl.h"

namespace {

__device__ __forceinline__ double nbe_self(double w, double t, int m)
{
    double z = w / t;   // nbe, extratools.f90:257-288
    if (m == 0) {
        if (z > 60) return 0;
        if (z < (double)0.0010f) return 1 / (z * (1 + z / 2 * (1 + z / 3)));
        return 1 / (exp(z) - 1);
    }
    return 1 / z;
}

// the eigensolve points of a batch: q_s (s < nq), then -q_s - kpc[k2] at nq + s*nk + k2
__global__ void k_self_points(int nq, int nk, const double *__restrict__ q, const double *__restrict__ kpc,
                              double *__restrict__ pts)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)nq * (nk + 1)) return;
    double *o = pts + 3 * i;
    if (i < (size_t)nq) {
        for (int c = 0; c < 3; c++) o[c] = q[3 * i + c];
    } else {
        const size_t r = i - nq;
        const int s = (int)(r / nk), k = (int)(r % nk);
        for (int c = 0; c < 3; c++) o[c] = -q[3 * s + c] - kpc[3 * (size_t)k + c];
    }
}

// |lambda|^(-1/4): +inf for a zero eigenvalue, which the V3 kernel reports (the reference's
// den = 0 stop)
__global__ void k_self_w(size_t n, const double *__restrict__ lam, double *__restrict__ w)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) w[i] = 1.0 / sqrt(sqrt(fabs(lam[i])));
}

// Gaussian delta of width w (delta_g, extratools.f90:239-246)
__device__ __forceinline__ double delta_gauss(double x, double w)
{
    return exp(-x * x / 2.0 / w / w) / sqrt(2 * alt3::PI) / w;
}

// Sums of the self-energy for point s and 256 mesh points k2 per group:
// part[s][chunk][l1][4] = (Re, Im) of the normal and of the umklapp sums (before / nk * pi).
// kind 0: function_self_w3 (rest.f90:4012-4124), four Lorentzian processes cut at 3000 eta;
// kind 1: function_self_w4 (rest.f90:4126-4240), the imaginary part alone from the Gaussian
// deltas of three processes cut at 6 eta (the real part stays zero)
__global__ void __launch_bounds__(256) k_self_sum(int nk, int nb, const double *__restrict__ v33,
                                                  const double *__restrict__ lamq, const double *__restrict__ lam3,
                                                  const double *__restrict__ freq, const double *__restrict__ q,
                                                  const double *__restrict__ kpc, double g0, double g1, double g2,
                                                  double rr0, double rr1, double rr2, double rr3, double rr4,
                                                  double rr5, double rr6, double rr7, double rr8, double temp,
                                                  int classical, double eta, double v3thr, double cnst,
                                                  int kind, double *__restrict__ part)
{
    __shared__ double red[4][4];   // [sum][wavefront]
    const int s = blockIdx.y, nch = gridDim.x, tid = threadIdx.x;
    const int k2 = blockIdx.x * 256 + tid;
    const bool valid = k2 < nk;
    const int kc = valid ? k2 : 0;
    // check_inside_bz(k3): reduced coordinates of k3 + (g1+g2+g3)/2 in [0,1), epsl 5e-10
    double k3[3];
    for (int c = 0; c < 3; c++) k3[c] = -q[3 * s + c] - kpc[3 * (size_t)kc + c];
    const double sh[3] = {k3[0] + 0.5 * g0, k3[1] + 0.5 * g1, k3[2] + 0.5 * g2};
    const double rr[9] = {rr0, rr1, rr2, rr3, rr4, rr5, rr6, rr7, rr8};
    bool inside = true;
    for (int i = 0; i < 3; i++) {
        const double x = (sh[0] * rr[3 * i] + sh[1] * rr[3 * i + 1] + sh[2] * rr[3 * i + 2]) + 5e-10;
        if (x < 0 || x >= 1) inside = false;
    }
    const double etacut = kind == 1 ? eta * 6 : eta * 3000;
    for (int l1 = 0; l1 < nb; l1++) {
        double nr = 0, ni = 0, ur = 0, ui = 0;
        if (valid) {
            const double omega = sqrt(fabs(lamq[(size_t)s * nb + l1])) * cnst;
            for (int l2 = 0; l2 < nb; l2++) {
                const double om2 = freq[(size_t)k2 * nb + l2], nb2 = nbe_self(om2, temp, classical);
                for (int l3 = 0; l3 < nb; l3++) {
                    const double term = v33[((((size_t)s * nb + l1) * nb + l2) * nb + l3) * (size_t)nk + k2];
                    if (sqrt(term) < v3thr) continue;   // |w33| < v3_threshold
                    const double om3 = sqrt(fabs(lam3[((size_t)s * nk + k2) * nb + l3])) * cnst;
                    const double nb3 = nbe_self(om3, temp, classical);
                    double ires = 0, rres = 0, arg;
                    if (kind == 1) {
                        arg = omega - om3 - om2;
                        if (fabs(arg) < etacut) ires = (nb3 + nb2 + 1) * delta_gauss(arg, eta);
                        arg = omega - om3 + om2;
                        if (fabs(arg) < etacut) ires = ires + (nb2 - nb3) * delta_gauss(arg, eta);
                        arg = omega - om2 + om3;
                        if (fabs(arg) < etacut) ires = ires + (nb3 - nb2) * delta_gauss(arg, eta);
                        goto summed;
                    }
                    arg = omega - om3 - om2;
                    if (fabs(arg) < etacut) {
                        ires = (nb3 + nb2 + 1) * (eta / (arg * arg + eta * eta) / alt3::PI);
                        rres = (nb3 + nb2 + 1) * arg / (arg * arg + eta * eta);
                    }
                    arg = omega + om3 + om2;
                    if (fabs(arg) < etacut) {
                        ires = ires - (nb3 + nb2 + 1) * (eta / (arg * arg + eta * eta) / alt3::PI);
                        rres = rres - (nb3 + nb2 + 1) * arg / (arg * arg + eta * eta);
                    }
                    arg = omega - om3 + om2;
                    if (fabs(arg) < etacut) {
                        ires = ires + (nb2 - nb3) * (eta / (arg * arg + eta * eta) / alt3::PI);
                        rres = rres + (nb2 - nb3) * arg / (arg * arg + eta * eta);
                    }
                    arg = omega - om2 + om3;
                    if (fabs(arg) < etacut) {
                        ires = ires + (nb3 - nb2) * (eta / (arg * arg + eta * eta) / alt3::PI);
                        rres = rres + (nb3 - nb2) * arg / (arg * arg + eta * eta);
                    }
                summed:
                    if (inside) {
                        nr += term * rres;
                        ni += term * (ires * alt3::PI);
                    } else {
                        ur += term * rres;
                        ui += term * (ires * alt3::PI);
                    }
                }
            }
        }
        // butterflies inside each of the four wavefronts, then the wavefront partials in LDS
        for (int off = 32; off > 0; off >>= 1) {
            nr += __shfl_xor(nr, off, 64);
            ni += __shfl_xor(ni, off, 64);
            ur += __shfl_xor(ur, off, 64);
            ui += __shfl_xor(ui, off, 64);
        }
        if ((tid & 63) == 0) {
            red[0][tid >> 6] = nr;
            red[1][tid >> 6] = ni;
            red[2][tid >> 6] = ur;
            red[3][tid >> 6] = ui;
        }
        __syncthreads();
        if (tid < 4)
            part[(((size_t)s * nch + blockIdx.x) * nb + l1) * 4 + tid] =
                ((red[tid][0] + red[tid][1]) + red[tid][2]) + red[tid][3];
        __syncthreads();
    }
}

// nself, uself (complex, Fortran (nb, nq)) = sum over the chunks / nk * pi
__global__ void k_self_final(int nq, int nb, int nch, int nk, const double *__restrict__ part,
                             double *__restrict__ nself, double *__restrict__ uself)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x;   // l1 + nb * s
    if (x >= nq * nb) return;
    const int s = x / nb, l1 = x % nb;
    double a[4] = {0, 0, 0, 0};
    for (int c = 0; c < nch; c++)
        for (int i = 0; i < 4; i++) a[i] += part[(((size_t)s * nch + c) * nb + l1) * 4 + i];
    nself[2 * x] = a[0] / nk * alt3::PI;
    nself[2 * x + 1] = a[1] / nk * alt3::PI;
    uself[2 * x] = a[2] / nk * alt3::PI;
    uself[2 * x + 1] = a[3] / nk * alt3::PI;
}

}  // namespace

namespace alt3 {

int self_run(alt3_ctx *ctx, int nq, const double *q, double tempk, double eta, double v3thr, int kind,
             double *omega0, double *nself, double *uself)
{
    if (nq < 1 || !q) return fail(ALT3_E_ARG, "no q point");
    if (!(tempk > 0)) return fail(ALT3_E_ARG, "ERROR in N_BE: t=<0");
    if (!ctx->have_eig) return fail(ALT3_E_STATE, "eigenvectors missing: call alt3_eig or alt3_set_eig first");
    const int nk = ctx->nk, nb = ctx->nb, ndr = ctx->V.ndr;
    if (!(ndr > 0 && (nb == 6 || nb == 12 || nb == 24)))
        return fail(ALT3_E_ARG, "the self-energy handles cells of 6, 12 or 24 branches");
    // batch of points: the k3 eigenvectors and |V3|^2 of a point take nk*nb^2*16 + nk*nb^3*8 B
    const size_t per = (size_t)nk * nb * nb * 16 + (size_t)nk * nb * nb * nb * 8;
    const int B = (int)std::max<size_t>(1, std::min<size_t>(64, ((size_t)1 << 30) / per));
    const int nch = (nk + 255) / 256;
    const double temp = tempk * K_B / (100 * H_PLANK * C_LIGHT);
    // make_reciprocal_lattice without 2 pi (rr_i . g_j = delta_ij) for check_inside_bz
    const double *g = ctx->T.g;
    double rr[9];
    {
        auto cross = [](const double *a, const double *b, double *c) {
            c[0] = a[1] * b[2] - a[2] * b[1];
            c[1] = a[2] * b[0] - a[0] * b[2];
            c[2] = a[0] * b[1] - a[1] * b[0];
        };
        double c[3];
        cross(g + 3, g + 6, c);
        const double om = g[0] * c[0] + g[1] * c[1] + g[2] * c[2];
        for (int i = 0; i < 3; i++) rr[i] = c[i] / om;
        cross(g + 6, g, c);
        for (int i = 0; i < 3; i++) rr[3 + i] = c[i] / om;
        cross(g, g + 3, c);
        for (int i = 0; i < 3; i++) rr[6 + i] = c[i] / om;
    }
    const double gs[3] = {g[0] + g[3] + g[6], g[1] + g[4] + g[7], g[2] + g[5] + g[8]};
    const size_t npts = (size_t)B * (nk + 1);
    DBuf<double> dq, pts, lam, evec, w, gp, v33, part, outs;
    ALT3_TRY(dq.alloc((size_t)3 * B));
    ALT3_TRY(pts.alloc(3 * npts));
    ALT3_TRY(lam.alloc(npts * nb));
    ALT3_TRY(evec.alloc(npts * nb * nb * 2));
    ALT3_TRY(w.alloc(npts * nb));
    ALT3_TRY(gp.alloc((size_t)B * ndr * nb * nb * nb * 2));
    ALT3_TRY(v33.alloc((size_t)B * nb * nb * nb * nk));
    ALT3_TRY(part.alloc((size_t)B * nch * nb * 4));
    ALT3_TRY(outs.alloc((size_t)4 * B * nb));
    int *derr = ctx->d_flags.p + 0;
    std::vector<double> hl(nb);
    for (int q0 = 0; q0 < nq; q0 += B) {
        const int b = std::min(B, nq - q0);
        const size_t np = (size_t)b * (nk + 1);
        ALT3_HIP(hipMemsetAsync(derr, 0, sizeof(int), ctx->stream));
        ALT3_HIP(hipMemcpyAsync(dq.p, q + 3 * (size_t)q0, sizeof(double) * 3 * b, hipMemcpyHostToDevice, ctx->stream));
        hipLaunchKernelGGL(k_self_points, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, ctx->stream, b, nk, dq.p,
                           ctx->d_kpc.p, pts.p);
        ALT3_HIP(hipGetLastError());
        // get_freq at q and at every k3 (zheev order, |lambda|, x i fix)
        ALT3_TRY(eig_launch(ctx, (int)np, 1, pts.p, 0.0, lam.p, evec.p));
        hipLaunchKernelGGL(k_self_w, dim3((unsigned)((np * nb + 255) / 256)), dim3(256), 0, ctx->stream, np * nb, lam.p,
                           w.p);
        ALT3_HIP(hipGetLastError());
        ALT3_TRY(v3_prep_points(ctx, b, dq.p, evec.p, gp.p));
        const double *lam3 = lam.p + (size_t)b * nb, *ev3 = evec.p + (size_t)b * nb * nb * 2, *w3 = w.p + (size_t)b * nb;
        ALT3_TRY(v3_rows_points(ctx, b, gp.p, ev3, w3, w.p, v33.p, derr));
        hipLaunchKernelGGL(k_self_sum, dim3((unsigned)nch, (unsigned)b), dim3(256), 0, ctx->stream, nk, nb, v33.p, lam.p,
                           lam3, ctx->d_freq.p, dq.p, ctx->d_kpc.p, gs[0], gs[1], gs[2], rr[0], rr[1], rr[2], rr[3],
                           rr[4], rr[5], rr[6], rr[7], rr[8], temp, ctx->classical, eta, v3thr, cnst(), kind, part.p);
        hipLaunchKernelGGL(k_self_final, dim3((b * nb + 63) / 64), dim3(64), 0, ctx->stream, b, nb, nch, nk, part.p,
                           outs.p, outs.p + 2 * (size_t)b * nb);
        ALT3_HIP(hipGetLastError());
        int herr = 0;
        ALT3_HIP(hipMemcpyAsync(&herr, derr, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
        if (nself)
            ALT3_HIP(hipMemcpyAsync(nself + 2 * (size_t)q0 * nb, outs.p, sizeof(double) * 2 * b * nb,
                                    hipMemcpyDeviceToHost, ctx->stream));
        if (uself)
            ALT3_HIP(hipMemcpyAsync(uself + 2 * (size_t)q0 * nb, outs.p + 2 * (size_t)b * nb, sizeof(double) * 2 * b * nb,
                                    hipMemcpyDeviceToHost, ctx->stream));
        if (omega0) {
            for (int s = 0; s < b; s++) {
                ALT3_HIP(hipMemcpyAsync(hl.data(), lam.p + (size_t)s * nb, sizeof(double) * nb, hipMemcpyDeviceToHost,
                                        ctx->stream));
                ALT3_HIP(hipStreamSynchronize(ctx->stream));
                for (int l = 0; l < nb; l++) omega0[(size_t)(q0 + s) * nb + l] = sqrt(fabs(hl[l])) * cnst();
            }
        }
        ALT3_HIP(hipStreamSynchronize(ctx->stream));
        if (herr) return fail(ALT3_E_DEN0, "MATRIX_ELT: den=0 (a zero frequency at q or at -q-k)");
    }
    return ALT3_OK;
}

}  // namespace alt3
