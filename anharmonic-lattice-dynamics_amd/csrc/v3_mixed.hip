// Mixed-precision |V3|^2: the phase GEMM on FP32 matrix cores (v_mfma_f32_16x16x4f32).
//
// For one q1 (s) and branch l1 the batched form of matrix_elt_simplified
// (THERMACOND/sub.f90:5888-6011) is
//   Phi'(q2)[b,c] = sum_r e^{i q2.dR_r} G'[r][l1][b,c]           (complex GEMM, K = ndr)
//   V[l2,l3]      = sum_b e2[b,l2] sum_c Phi'[b,c] e3[c,l3]      (two nb^3 contractions)
// The GEMM is run as the real product
//   [Re Phi' | Im Phi'] = [Re E | Im E] . [[Re G', Im G'], [-Im G', Re G']]
// with M = a tile of TQ q2 points, N = 2 nb^2, K = 2 ndr.  A (the phase tile, from a
// per-mesh table) sits in LDS; B (G' packed per (q1,l1) by k_pack_b) is read from L2,
// where all nk/TQ workgroups of the same (q1,l1) share it.  Inputs are FP32 with FP32
// accumulation; the contractions and |V|^2 are FP32, the 1/sqrt(8 sqrt|l1 l2 l3|) scale
// and const33 are applied in FP64.
#include "alt3_impl.h"

namespace {

typedef float f4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ void coords(const alt3::MeshDims &d, int k, int &c0, int &c1, int &c2)
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
__device__ __forceinline__ int mesh_neg_sum(const alt3::MeshDims &d, int k1, int k2)
{
    int a0, a1, a2, b0, b1, b2;
    coords(d, k1, a0, a1, a2);
    coords(d, k2, b0, b1, b2);
    return wrap(3 * d.n2 / 2 - a2 - b2, d.n2) +
           d.n2 * (wrap(3 * d.n1 / 2 - a1 - b1, d.n1) + d.n1 * wrap(3 * d.n0 / 2 - a0 - b0, d.n0));
}

// B[s][l1][k][n] (FP32) from G'[s][r][l1][bc] (FP64 complex)
__global__ void k_pack_b(int ksub, int nb, int ndr, int Kp, int Np, const double *__restrict__ gp, float *__restrict__ B)
{
    const int nb2 = nb * nb;
    size_t tot = (size_t)ksub * nb * Kp * Np;
    size_t o = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= tot) return;
    size_t x = o;
    int n = x % Np; x /= Np;
    int k = x % Kp; x /= Kp;
    int l1 = x % nb; x /= nb;
    int s = (int)x;
    float v = 0.f;
    if (k < 2 * ndr && n < 2 * nb2) {
        int r = k < ndr ? k : k - ndr;
        int bc = n < nb2 ? n : n - nb2;
        const double *g = gp + 2 * ((((size_t)s * ndr + r) * nb + l1) * nb2 + bc);
        double gr = g[0], gi = g[1];
        if (k < ndr) v = (float)(n < nb2 ? gr : gi);
        else v = (float)(n < nb2 ? -gi : gr);
    }
    B[o] = v;
}

__global__ void k_to_float2(size_t n, const double *__restrict__ in, float *__restrict__ out)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (float)in[i];
}

// one workgroup (4 waves) per (s, l1, tile of TQ q2)
template <int TQ>
__global__ void __launch_bounds__(256) k_v3_mfma(alt3::MeshDims md, int nk, int nb, int ibz_lo, int ndr, int Kp, int Np,
                                                 const int *__restrict__ mapinv, const float *__restrict__ ephf,
                                                 const float *__restrict__ evecf, const double *__restrict__ eival,
                                                 const float *__restrict__ B, double c33, float *__restrict__ out,
                                                 int *__restrict__ err)
{
    extern __shared__ float lds[];
    const int s = blockIdx.z, l1 = blockIdx.y, q0 = blockIdx.x * TQ;
    const int nb2 = nb * nb;
    float *A = lds;                  // [TQ][Kp]
    float *Phi = A + TQ * Kp;        // [TQ][Np]   Re in [0,nb2), Im in [nb2, 2 nb2)
    float *U = Phi + TQ * Np;        // [TQ][2*nb2] (b, l3) complex interleaved
    __shared__ int k3s[TQ];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int k1 = mapinv[ibz_lo + s];
    if (tid < TQ) k3s[tid] = (q0 + tid < nk) ? mesh_neg_sum(md, k1, q0 + tid) : k1;
    for (int x = tid; x < TQ * Kp; x += 256) {
        int qq = x / Kp, k = x % Kp;
        int n2 = q0 + qq;
        float v = 0.f;
        if (n2 < nk && k < 2 * ndr) {
            int r = k < ndr ? k : k - ndr;
            v = ephf[2 * ((size_t)n2 * ndr + r) + (k < ndr ? 0 : 1)];
        }
        A[x] = v;
    }
    __syncthreads();
    // GEMM tiles (mt, nt) spread over the 4 waves
    const float *Bs = B + ((size_t)s * nb + l1) * (size_t)Kp * Np;
    const int MT = TQ / 16, NT = Np / 16;
    for (int p = wave; p < MT * NT; p += 4) {
        const int mt = p % MT, nt = p / MT;
        f4 acc = {0.f, 0.f, 0.f, 0.f};
        const int row = mt * 16 + (lane & 15), kq = lane >> 4, col = nt * 16 + (lane & 15);
        for (int kk = 0; kk < Kp; kk += 4) {
            float a = A[row * Kp + kk + kq];
            float b = Bs[(size_t)(kk + kq) * Np + col];
            acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc, 0, 0, 0);
        }
        // D layout: lane l, element i -> row 4*(l/16)+i, column l%16 of the 16x16 tile
#pragma unroll
        for (int i = 0; i < 4; i++) Phi[(mt * 16 + 4 * kq + i) * Np + col] = acc[i];
    }
    __syncthreads();
    // U[qq][b][l3] = sum_c Phi'[qq][b][c] e3[c][l3]
    for (int x = tid; x < TQ * nb2; x += 256) {
        int qq = x / nb2, b = (x / nb) % nb, l3 = x % nb;
        const float *e3 = evecf + 2 * ((size_t)k3s[qq] * nb + l3) * nb;
        const float *pr = Phi + qq * Np + b * nb, *pi = pr + nb2;
        float ur = 0.f, ui = 0.f;
        for (int c = 0; c < nb; c++) {
            float er = e3[2 * c], ei = e3[2 * c + 1];
            ur += pr[c] * er - pi[c] * ei;
            ui += pr[c] * ei + pi[c] * er;
        }
        U[2 * x] = ur;
        U[2 * x + 1] = ui;
    }
    __syncthreads();
    const double lam1 = eival[(size_t)k1 * nb + l1];
    for (int x = tid; x < TQ * nb2; x += 256) {
        int l2 = x / (TQ * nb), rem = x % (TQ * nb), l3 = rem / TQ, qq = rem % TQ;
        int n2 = q0 + qq;
        if (n2 >= nk) continue;
        const float *e2 = evecf + 2 * ((size_t)n2 * nb + l2) * nb;
        float vr = 0.f, vi = 0.f;
        for (int b = 0; b < nb; b++) {
            float ur = U[2 * (qq * nb2 + b * nb + l3)], ui = U[2 * (qq * nb2 + b * nb + l3) + 1];
            float er = e2[2 * b], ei = e2[2 * b + 1];
            vr += er * ur - ei * ui;
            vi += er * ui + ei * ur;
        }
        double den = sqrt(8 * sqrt(fabs(lam1 * eival[(size_t)n2 * nb + l2] * eival[(size_t)k3s[qq] * nb + l3])));
        if (den == 0) {
            atomicExch(err, 1);
            den = 1;
        }
        double sc = c33 / den;
        float f = (float)(sc * sc);
        out[(((size_t)s * nb + l1) * nb + l2) * (size_t)nb * nk + (size_t)l3 * nk + n2] = (vr * vr + vi * vi) * f;
    }
}

}  // namespace

namespace alt3 {

static size_t mixed_lds(const alt3_ctx *ctx, int t)
{
    const int nb2 = ctx->nb * ctx->nb, ndr = ctx->V.ndr;
    const int Kp = ((2 * ndr + 3) / 4) * 4, Np = ((2 * nb2 + 15) / 16) * 16;
    return (size_t)t * (Kp + Np + 2 * nb2) * sizeof(float);
}

// host-only contexts have no device to ask: they assume the 160 KB of a gfx950 workgroup
bool v3_mixed_fits(const alt3_ctx *ctx)
{
    return mixed_lds(ctx, 16) <= (ctx->host_only ? (size_t)163840 : ctx->lds_max);
}

int v3_mixed_run(alt3_ctx *ctx, int ksub, int *d_err)
{
    const int nk = ctx->nk, nb = ctx->nb, nb2 = nb * nb, ndr = ctx->V.ndr;
    const int Kp = ((2 * ndr + 3) / 4) * 4, Np = ((2 * nb2 + 15) / 16) * 16;
    int tq = 64;
    auto lds_of = [&](int t) { return mixed_lds(ctx, t); };
    while (tq > 16 && lds_of(tq) > 65536) tq /= 2;
    if (lds_of(tq) > ctx->lds_max) return fail(ALT3_E_ARG, "ALT3_V3_MIXED: a 16-point q2 tile does not fit in LDS (nb too large)");
    MeshDims md{ctx->nc[0], ctx->nc[1], ctx->nc[2]};
    // G' (FP64) then packed FP32 B
    ALT3_TRY(v3_prep_launch(ctx, ksub));
    size_t nB = (size_t)ksub * nb * Kp * Np;
    ALT3_TRY(ctx->d_gpf.alloc(nB));
    hipLaunchKernelGGL(k_pack_b, dim3((unsigned)((nB + 255) / 256)), dim3(256), 0, ctx->stream, ksub, nb, ndr, Kp, Np,
                       ctx->d_gp.p, ctx->d_gpf.p);
    ALT3_HIP(hipGetLastError());
    // FP32 copies of the eigenvectors and the phase table
    DBuf<float> evf, ephf;
    size_t ne = (size_t)nk * nb * nb * 2, np = (size_t)nk * ndr * 2;
    ALT3_TRY(evf.alloc(ne));
    ALT3_TRY(ephf.alloc(np));
    hipLaunchKernelGGL(k_to_float2, dim3((unsigned)((ne + 255) / 256)), dim3(256), 0, ctx->stream, ne, ctx->d_evec.p, evf.p);
    hipLaunchKernelGGL(k_to_float2, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, ctx->stream, np, ctx->d_eph.p, ephf.p);
    ALT3_HIP(hipGetLastError());
    dim3 grid((nk + tq - 1) / tq, nb, ksub);
    size_t lds = lds_of(tq);
    const double c33 = const33();
#define LAUNCH(T)                                                                                                     \
    if (lds > 65536)                                                                                                  \
        ALT3_HIP(hipFuncSetAttribute((const void *)k_v3_mfma<T>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds)); \
    hipLaunchKernelGGL(k_v3_mfma<T>, grid, dim3(256), lds, ctx->stream, md, nk, nb, ctx->ibz_lo, ndr, Kp, Np,         \
                       ctx->d_mapinv.p, ephf.p, evf.p, ctx->d_eival.p, ctx->d_gpf.p, c33, ctx->d_v33f.p, d_err)
    kernel_mark(ctx, 0, false);
    switch (tq) {
        case 64: LAUNCH(64); break;
        case 32: LAUNCH(32); break;
        default: LAUNCH(16); break;
    }
#undef LAUNCH
    ALT3_HIP(hipGetLastError());
    kernel_mark(ctx, 0, true);
    ALT3_HIP(hipStreamSynchronize(ctx->stream));
    return ALT3_OK;
}

}  // namespace alt3
