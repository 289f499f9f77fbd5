// Batched |V3|^2 with every l1 of a q2 tile in one workgroup (cells of 6, 12 or 24 branches).
//
// Same algebra as k_v3_main (THERMACOND/sub.f90:5888-6011 in batched form):
//   Phi'(q2)[b,c] = sum_r e^{i q2.dR_r} G'[r][l1][b,c]
//   U[b,l3]       = sum_c Phi'[b,c] e3[c,l3]            e3 = e(q3 = -q1-q2)
//   V[l2,l3]      = sum_b e2[b,l2] U[b,l3]
//   |V3|^2        = |V|^2 const33^2 / (8 sqrt|l1 l2 l3|)
// but organised for the hardware: one workgroup holds a tile of TQ q2 points and loops over
// l1, so the phase rows, the e3 blocks and the per-mode scales are staged once per tile
// instead of once per (tile, l1).  Thread (q, j) owns row b = j of Phi'/U (the phase sum
// over r streams into the U row one column c at a time, so Phi' never materialises) and,
// after a barrier, column l2 = j of V.  LDS arrays are stored q-fastest so that the lanes of
// a wave (consecutive q) hit consecutive banks, and the output rows [s][l1][l2][l3][q2] are
// written with consecutive q2 per lane.  T = double (ALT3_V3_FP64) or float (ALT3_V3_MIXED:
// FP32 phase sum and contractions, FP64 scale); the scale uses the per-mode table
// |lambda|^(-1/4), so 1/den = w1 w2 w3 / sqrt 8.
#include <map>
#include <mutex>
#include <tuple>

#include "alt3_impl.h"
#include "mesh_dev.h"

namespace {

// a*b + c as one fused multiply-add: each complex multiply-accumulate below is four of them
__device__ __forceinline__ float fmx(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fmx(double a, double b, double c) { return __builtin_fma(a, b, c); }

// |lambda(k,l)|^(-1/4); +inf for a zero eigenvalue (the reference's den = 0 stop)
__global__ void k_lam_quarter(int n, const double *__restrict__ eival, double *__restrict__ w)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) w[i] = 1.0 / sqrt(sqrt(fabs(eival[i])));
}

template <typename T>
__global__ void k_copy_real(size_t n, const double *__restrict__ in, T *__restrict__ out)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (T)in[i];
}

// EXT: q1 is point s of a list (G' of that point in gp, w1x [s][NB]) and e3 / w3 at
// q3 = -q1-q2 come per (s, q2) from ev3x [s][q2][NB col][NB] and w3x [s][q2][NB] (the
// self-energy's q3 is not a mesh point); otherwise q1 is the irreducible point s and e3, w3
// are read at the mesh point -q1-q2.
// one tile: q1 = irreducible point s (of the block ibz_lo...), q2 = q0 .. q0 + TQ - 1
template <typename T, int NB, int TQ, bool EXT>
__device__ __forceinline__ void rows_tile(const alt3::MeshDev &m, int nk, int ibz_lo, int ndr,
                                          const int *__restrict__ mapinv, const T *__restrict__ eph,
                                          const T *__restrict__ evec, const double *__restrict__ wq,
                                          const T *__restrict__ gp, double c33sq8, T *__restrict__ out,
                                          int *__restrict__ err, const T *__restrict__ ev3x,
                                          const double *__restrict__ w3x, const double *__restrict__ w1x, int s, int q0)
{
    extern __shared__ __align__(16) unsigned char smem[];
    T *E = reinterpret_cast<T *>(smem);     // [r][q] re, then im        (2 ndr TQ)
    T *E3 = E + 2 * ndr * TQ;                // [c][l3][q] re, then im    (2 NB NB TQ)
    T *U = E3 + 2 * NB * NB * TQ;            // [b][l3][q] re, then im    (2 NB NB TQ)
    // G' rows of the current l1, [r][b][c] (re, im), staged in LDS for NB <= 12 so that the
    // phase loop reads them as LDS broadcasts (two rows b per wave) instead of vector loads
    constexpr bool kStageG = NB <= 12;
    constexpr int kGr = kStageG ? NB * NB : NB * NB * NB;   // stride of r in the G' rows read
    T *Gs = U + 2 * NB * NB * TQ;            // (2 ndr NB NB) when kStageG
    __shared__ double w3s[NB][TQ];
    __shared__ int k3s[TQ];
    constexpr int NT = TQ * NB, NN = NB * NB;
    const int tid = threadIdx.x, q = tid % TQ, j = tid / TQ;
    const int k1 = EXT ? 0 : mapinv[ibz_lo + s];
    const int n2 = q0 + q;
    const bool valid = n2 < nk;
    const int n2c = valid ? n2 : nk - 1;
    if (!EXT && j == 0) {
        int a0, a1, a2, b0, b1, b2;
        m.coords(k1, a0, a1, a2);
        m.coords(n2c, b0, b1, b2);
        k3s[q] = m.neg_sum(a0, a1, a2, b0, b1, b2);
    }
    for (int x = tid; x < ndr * TQ; x += NT) {
        const int r = x / TQ, qq = x - r * TQ;
        const int nn = min(q0 + qq, nk - 1);
        E[x] = eph[2 * ((size_t)nn * ndr + r)];
        E[ndr * TQ + x] = eph[2 * ((size_t)nn * ndr + r) + 1];
    }
    __syncthreads();
    for (int x = tid; x < NN * TQ; x += NT) {      // x = (c*NB + l3)*TQ + qq
        const int cl = x / TQ, qq = x - cl * TQ, c = cl / NB, l3 = cl - c * NB;
        const T *e;
        if constexpr (EXT)
            e = ev3x + 2 * ((((size_t)s * nk + min(q0 + qq, nk - 1)) * NB + l3) * NB + c);
        else
            e = evec + 2 * (((size_t)k3s[qq] * NB + l3) * NB + c);
        E3[x] = e[0];
        E3[NN * TQ + x] = e[1];
    }
    for (int x = tid; x < NB * TQ; x += NT) {
        const int l3 = x / TQ, qq = x - l3 * TQ;
        if constexpr (EXT)
            w3s[l3][qq] = w3x[((size_t)s * nk + min(q0 + qq, nk - 1)) * NB + l3];
        else
            w3s[l3][qq] = wq[(size_t)k3s[qq] * NB + l3];
    }
    const double w2 = wq[(size_t)n2c * NB + j];
    const T *e2 = evec + 2 * ((size_t)n2c * NB + j) * NB;   // e(q2) column l2 = j
    __syncthreads();
    for (int l1 = 0; l1 < NB; l1++) {
        if constexpr (kStageG) {
            // (r, b, c) of this l1; the previous l1's reads ended at the barrier after V
            for (int x = tid; x < ndr * NN; x += NT) {
                const int r = x / NN, bc = x - r * NN;
                const T *src = gp + 2 * ((((size_t)s * ndr + r) * NB + l1) * NN + bc);
                Gs[2 * x] = src[0];
                Gs[2 * x + 1] = src[1];
            }
            __syncthreads();
        }
        // row b = j of U = Phi' e3, Phi' streamed column by column
        T ur[NB], ui[NB];
#pragma unroll
        for (int l3 = 0; l3 < NB; l3++) ur[l3] = ui[l3] = 0;
        const T *g = kStageG ? Gs + 2 * (j * NB) : gp + 2 * (((size_t)s * ndr * NB + l1) * NN + j * NB);   // + r*NB*NN + c
#pragma unroll 1
        for (int c = 0; c < NB; c++) {
            T pr = 0, pi = 0;
#pragma unroll 4
            for (int r = 0; r < ndr; r++) {
                const T gr = g[2 * (r * kGr + c)], gi = g[2 * (r * kGr + c) + 1];
                const T er = E[r * TQ + q], ei = E[ndr * TQ + r * TQ + q];
                pr = fmx(er, gr, pr);
                pr = fmx(-ei, gi, pr);
                pi = fmx(er, gi, pi);
                pi = fmx(ei, gr, pi);
            }
#pragma unroll
            for (int l3 = 0; l3 < NB; l3++) {
                const T xr = E3[(c * NB + l3) * TQ + q], xi = E3[NN * TQ + (c * NB + l3) * TQ + q];
                ur[l3] = fmx(pr, xr, ur[l3]);
                ur[l3] = fmx(-pi, xi, ur[l3]);
                ui[l3] = fmx(pr, xi, ui[l3]);
                ui[l3] = fmx(pi, xr, ui[l3]);
            }
        }
#pragma unroll
        for (int l3 = 0; l3 < NB; l3++) {
            U[(j * NB + l3) * TQ + q] = ur[l3];
            U[NN * TQ + (j * NB + l3) * TQ + q] = ui[l3];
        }
        __syncthreads();
        // column l2 = j of V = e2^T U, scaled and squared
        if (valid) {
            T vr[NB], vi[NB];
#pragma unroll
            for (int l3 = 0; l3 < NB; l3++) vr[l3] = vi[l3] = 0;
#pragma unroll 2
            for (int b = 0; b < NB; b++) {
                const T er = e2[2 * b], ei = e2[2 * b + 1];
#pragma unroll
                for (int l3 = 0; l3 < NB; l3++) {
                    const T xr = U[(b * NB + l3) * TQ + q], xi = U[NN * TQ + (b * NB + l3) * TQ + q];
                    vr[l3] = fmx(er, xr, vr[l3]);
                    vr[l3] = fmx(-ei, xi, vr[l3]);
                    vi[l3] = fmx(er, xi, vi[l3]);
                    vi[l3] = fmx(ei, xr, vi[l3]);
                }
            }
            const double w12 = (EXT ? w1x[(size_t)s * NB + l1] : wq[(size_t)k1 * NB + l1]) * w2;
            T *o = out + ((((size_t)s * NB + l1) * NB + j) * NB) * (size_t)nk + n2;
#pragma unroll
            for (int l3 = 0; l3 < NB; l3++) {
                const double w = w12 * w3s[l3][q];
                if (!(w <= 1.7e308)) atomicExch(err, 1);   // a zero eigenvalue in the triple
                o[(size_t)l3 * nk] = (vr[l3] * vr[l3] + vi[l3] * vi[l3]) * (T)(c33sq8 * w * w);
            }
        }
        __syncthreads();
    }
}

template <typename T, int NB, int TQ, bool EXT>
__global__ void __launch_bounds__(TQ * NB) k_v3_rows(alt3::MeshDev m, int nk, int ibz_lo, int ndr,
                                                     const int *__restrict__ mapinv, const T *__restrict__ eph,
                                                     const T *__restrict__ evec, const double *__restrict__ wq,
                                                     const T *__restrict__ gp, double c33sq8, T *__restrict__ out,
                                                     int *__restrict__ err, const T *__restrict__ ev3x,
                                                     const double *__restrict__ w3x, const double *__restrict__ w1x)
{
    rows_tile<T, NB, TQ, EXT>(m, nk, ibz_lo, ndr, mapinv, eph, evec, wq, gp, c33sq8, out, err, ev3x, w3x, w1x,
                              blockIdx.y, blockIdx.x * TQ);
}

// the arrays of one context in a batched launch (alt3_v3_batch)
template <typename T>
struct RowsJob {
    alt3::MeshDev m;
    int nk, ndr, tiles;       // tiles of q2 per irreducible point
    long long first;          // first work item of this context in the batch
    const int *mapinv;
    const T *eph, *evec, *gp;
    const double *wq;
    T *out;
    int *err;
};

// Shared persistent launch over the irreducible zones of several contexts of NB branches:
// a grid sized to the device pops work items (context, q1, tile of q2) from one counter;
// every wavefront reads the same item and leaves the loop together once the counter passes
// the last item.  Each tile is computed exactly as k_v3_rows computes it.
template <typename T, int NB, int TQ>
__global__ void __launch_bounds__(TQ * NB) k_v3_rows_batch(const RowsJob<T> *__restrict__ jobs, int njobs,
                                                           long long nitems, unsigned long long *__restrict__ counter,
                                                           double c33sq8)
{
    __shared__ long long s_item;
    __shared__ int s_job;
    for (;;) {
        __syncthreads();   // the previous item is done with s_item and the LDS tiles
        if (threadIdx.x == 0) {
            const long long it = (long long)atomicAdd(counter, 1ull);
            int j = 0;
            if (it < nitems) {   // last context whose first item is <= it (jobs sorted by first)
                int lo = 0, hi = njobs - 1;
                while (lo < hi) {
                    const int mid = (lo + hi + 1) / 2;
                    if (jobs[mid].first <= it) lo = mid; else hi = mid - 1;
                }
                j = lo;
            }
            s_item = it;
            s_job = j;
        }
        __syncthreads();
        const long long it = s_item;
        if (it >= nitems) break;
        const RowsJob<T> &J = jobs[s_job];
        const long long local = it - J.first;
        const int sq = (int)(local / J.tiles), tile = (int)(local - (long long)sq * J.tiles);
        rows_tile<T, NB, TQ, false>(J.m, J.nk, 0, J.ndr, J.mapinv, J.eph, J.evec, J.wq, J.gp, c33sq8, J.out, J.err,
                                    nullptr, nullptr, nullptr, sq, tile * TQ);
    }
}

__global__ void k_zero_u64(unsigned long long *p)
{
    if (threadIdx.x == 0) *p = 0ull;
}

template <typename T, int NB, int TQ, bool EXT>
int launch_rows(alt3_ctx *ctx, int ksub, const T *eph, const T *evec, const double *wq, const T *gp, T *out,
                int *d_err, const T *ev3x, const double *w3x, const double *w1x)
{
    const int ndr = ctx->V.ndr;
    const size_t lds = (size_t)2 * ((ndr + 2 * NB * NB) * TQ + (NB <= 12 ? ndr * NB * NB : 0)) * sizeof(T);
    if (lds > ctx->lds_max) return alt3::fail(ALT3_E_ARG, "V3 tile does not fit in LDS");
    if (lds > 65536)
        ALT3_HIP(hipFuncSetAttribute((const void *)k_v3_rows<T, NB, TQ, EXT>,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    const double c33 = alt3::const33();
    dim3 grid((unsigned)((ctx->nk + TQ - 1) / TQ), (unsigned)ksub);
    if (!EXT) alt3::kernel_mark(ctx, 0, false);
    hipLaunchKernelGGL((k_v3_rows<T, NB, TQ, EXT>), grid, dim3(TQ * NB), lds, ctx->stream,
                       alt3::MeshDev::make(ctx->nc), ctx->nk, ctx->ibz_lo, ndr, ctx->d_mapinv.p, eph, evec, wq, gp,
                       c33 * c33 / 8, out, d_err, ev3x, w3x, w1x);
    ALT3_HIP(hipGetLastError());
    if (!EXT) alt3::kernel_mark(ctx, 0, true);
    return ALT3_OK;
}

template <typename T, bool EXT = false>
int dispatch_rows(alt3_ctx *ctx, int ksub, const T *eph, const T *evec, const double *wq, const T *gp, T *out,
                  int *d_err, const T *ev3x = nullptr, const double *w3x = nullptr, const double *w1x = nullptr)
{
    switch (ctx->nb) {
        case 6: return launch_rows<T, 6, 32, EXT>(ctx, ksub, eph, evec, wq, gp, out, d_err, ev3x, w3x, w1x);
        case 12: return launch_rows<T, 12, 16, EXT>(ctx, ksub, eph, evec, wq, gp, out, d_err, ev3x, w3x, w1x);
        default: return launch_rows<T, 24, 4, EXT>(ctx, ksub, eph, evec, wq, gp, out, d_err, ev3x, w3x, w1x);
    }
}

}  // namespace

namespace alt3 {

// 24 branches only in FP32: the FP64 tile holds 4 q2 points, too few to amortise the G' rows
// (k_v3_main is faster there)
bool v3_rows_supported(const alt3_ctx *ctx, bool fp32)
{
    return ctx->V.ndr > 0 && (ctx->nb == 6 || ctx->nb == 12 || (ctx->nb == 24 && fp32));
}

// wq = |lambda|^(-1/4) and, for the FP32 kernels, FP32 copies of the phase table, the
// eigenvectors and (rows kernel) G' of ksub irreducible points, on the context stream
int v3_rows_inputs(alt3_ctx *ctx, int ksub, bool fp32, bool want_gpf)
{
    const int nk = ctx->nk, nb = ctx->nb, ndr = ctx->V.ndr;
    DBuf<double> &wq = ctx->d_wq;
    ALT3_TRY(wq.reserve((size_t)nk * nb));
    hipLaunchKernelGGL(k_lam_quarter, dim3((nk * nb + 255) / 256), dim3(256), 0, ctx->stream, nk * nb, ctx->d_eival.p,
                       wq.p);
    ALT3_HIP(hipGetLastError());
    if (!fp32) return ALT3_OK;
    DBuf<float> &ephf = ctx->d_ephf, &evf = ctx->d_evf, &gpf = ctx->d_gpf32;
    const size_t ne = (size_t)nk * nb * nb * 2, np = (size_t)nk * ndr * 2, ng = (size_t)ksub * ndr * nb * nb * nb * 2;
    ALT3_TRY(ephf.reserve(np));
    ALT3_TRY(evf.reserve(ne));
    hipLaunchKernelGGL(k_copy_real<float>, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, ctx->stream, np,
                       ctx->d_eph.p, ephf.p);
    hipLaunchKernelGGL(k_copy_real<float>, dim3((unsigned)((ne + 255) / 256)), dim3(256), 0, ctx->stream, ne,
                       ctx->d_evec.p, evf.p);
    if (want_gpf) {
        ALT3_TRY(gpf.reserve(ng));
        hipLaunchKernelGGL(k_copy_real<float>, dim3((unsigned)((ng + 255) / 256)), dim3(256), 0, ctx->stream, ng,
                           ctx->d_gp.p, gpf.p);
    }
    ALT3_HIP(hipGetLastError());
    return ALT3_OK;
}

// G' must be prepared (v3_prep_launch); out holds ksub*nb^3*nk values of the precision's type
int v3_rows_run(alt3_ctx *ctx, int ksub, bool fp32, int *d_err)
{
    const bool chain = fp32 && v3_chain_supported(ctx);
    ALT3_TRY(v3_rows_inputs(ctx, ksub, fp32, fp32 && !chain));
    DBuf<double> &wq = ctx->d_wq;
    if (!fp32) {
        ALT3_TRY(dispatch_rows<double>(ctx, ksub, ctx->d_eph.p, ctx->d_evec.p, wq.p, ctx->d_gp.p, ctx->d_v33.p, d_err));
    } else if (chain) {   // 24 branches: every contraction on the matrix cores
        if (ctx->v3_mma == ALT3_MMA_F16X3 && v3_chain16_supported(ctx))
            ALT3_TRY(v3_chain16_run(ctx, ksub, ctx->d_evf.p, ctx->d_ephf.p, wq.p, ctx->d_v33f.p, d_err));
        else
            ALT3_TRY(v3_chain_run(ctx, ksub, ctx->d_evf.p, ctx->d_ephf.p, wq.p, ctx->d_v33f.p, d_err));
    } else {
        ALT3_TRY(dispatch_rows<float>(ctx, ksub, ctx->d_ephf.p, ctx->d_evf.p, wq.p, ctx->d_gpf32.p, ctx->d_v33f.p, d_err));
    }
    // no wait: every buffer used here belongs to the context and outlives the kernels
    return ALT3_OK;
}

// ------------------------------------------------------------------ batched launches
namespace {

// the tile sizes the autotuner tries for a branch count (TQ * NB threads per group)
const int *tile_candidates(int nb, int &n)
{
    static const int c6[] = {16, 32, 64}, c12[] = {8, 16, 32};
    n = 3;
    return nb == 6 ? c6 : c12;
}

size_t rows_lds(int nb, int ndr, int tq, size_t tsize)
{
    return (size_t)2 * ((ndr + 2 * nb * nb) * tq + (nb <= 12 ? ndr * nb * nb : 0)) * tsize;
}

std::mutex g_tune_mu;
std::map<std::tuple<int, int, int>, int> g_tune;   // (fp32, nb, ndr) -> tile size

template <typename T, int NB, int TQ>
int batch_launch_t(alt3_ctx *lead, const RowsJob<T> *d_jobs, int njobs, long long nitems, int ndr_max,
                   unsigned long long *counter)
{
    const void *kern = (const void *)k_v3_rows_batch<T, NB, TQ>;
    const size_t lds = rows_lds(NB, ndr_max, TQ, sizeof(T));
    if (lds > lead->lds_max) return fail(ALT3_E_ARG, "batched V3 tile does not fit in LDS");
    if (lds > 65536) ALT3_HIP(hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    int per_cu = 0, cus = 0;
    ALT3_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, TQ * NB, lds));
    ALT3_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, lead->device));
    const long long want = (long long)std::max(1, per_cu) * std::max(1, cus);
    const unsigned grid = (unsigned)std::max(1ll, std::min(want, nitems));
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, lead->stream, counter);
    const double c33 = const33();
    kernel_mark(lead, 0, false);
    hipLaunchKernelGGL((k_v3_rows_batch<T, NB, TQ>), dim3(grid), dim3(TQ * NB), lds, lead->stream, d_jobs, njobs, nitems,
                       counter, c33 * c33 / 8);
    ALT3_HIP(hipGetLastError());
    kernel_mark(lead, 0, true);
    return ALT3_OK;
}

template <typename T>
int batch_launch(alt3_ctx *lead, int nb, int tq, const RowsJob<T> *d_jobs, int njobs, long long nitems, int ndr_max,
                 unsigned long long *counter)
{
#define ALT3_BATCH(NB, TQ) \
    if (nb == NB && tq == TQ) return batch_launch_t<T, NB, TQ>(lead, d_jobs, njobs, nitems, ndr_max, counter);
    ALT3_BATCH(6, 16) ALT3_BATCH(6, 32) ALT3_BATCH(6, 64)
    ALT3_BATCH(12, 8) ALT3_BATCH(12, 16) ALT3_BATCH(12, 32)
#undef ALT3_BATCH
    return fail(ALT3_E_ARG, "no batched V3 kernel for this branch count and tile");
}

template <typename T>
RowsJob<T> job_of(alt3_ctx *c, int tq, long long first, bool fp32)
{
    RowsJob<T> j;
    j.m = MeshDev::make(c->nc);
    j.nk = c->nk;
    j.ndr = c->V.ndr;
    j.tiles = (c->nk + tq - 1) / tq;
    j.first = first;
    j.mapinv = c->d_mapinv.p;
    if constexpr (sizeof(T) == sizeof(float)) {
        j.eph = c->d_ephf.p;
        j.evec = c->d_evf.p;
        j.gp = c->d_gpf32.p;
        j.out = c->d_v33f.p;
    } else {
        j.eph = c->d_eph.p;
        j.evec = c->d_evec.p;
        j.gp = c->d_gp.p;
        j.out = c->d_v33.p;
    }
    (void)fp32;
    j.wq = c->d_wq.p;
    j.err = c->d_flags.p + 0;
    return j;
}

// one batched launch over ctxs with tile tq; the job table and the counter live in the lead
// context's batch scratch
template <typename T>
int batch_run(alt3_ctx *const *ctxs, int n, int tq, bool fp32)
{
    alt3_ctx *lead = ctxs[0];
    std::vector<RowsJob<T>> jobs;
    long long first = 0;
    int ndr_max = 0;
    for (int i = 0; i < n; i++) {
        jobs.push_back(job_of<T>(ctxs[i], tq, first, fp32));
        first += (long long)ctxs[i]->nibz * jobs.back().tiles;
        ndr_max = std::max(ndr_max, ctxs[i]->V.ndr);
    }
    const size_t bytes = sizeof(RowsJob<T>) * jobs.size() + 64;
    ALT3_TRY(lead->d_batch.reserve(bytes));
    unsigned long long *counter = reinterpret_cast<unsigned long long *>(lead->d_batch.p);
    RowsJob<T> *d_jobs = reinterpret_cast<RowsJob<T> *>(lead->d_batch.p + 64);
    ALT3_HIP(hipMemcpyAsync(d_jobs, jobs.data(), sizeof(RowsJob<T>) * jobs.size(), hipMemcpyHostToDevice, lead->stream));
    ALT3_TRY(batch_launch<T>(lead, lead->nb, tq, d_jobs, n, first, ndr_max, counter));
    // the host copy of the job table must outlive the copy
    ALT3_HIP(hipStreamSynchronize(lead->stream));
    return ALT3_OK;
}

}  // namespace

bool v3_batch_supported(const alt3_ctx *ctx) { return ctx->V.ndr > 0 && (ctx->nb == 6 || ctx->nb == 12); }

// the tile size of the batched kernel for the context's (precision, nb, ndr): the cached
// choice, or the fastest candidate timed on this context's own zone (its inputs prepared)
int v3_tile_for(alt3_ctx *ctx, bool fp32, int *tq)
{
    const auto key = std::make_tuple(fp32 ? 1 : 0, ctx->nb, ctx->V.ndr);
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto it = g_tune.find(key);
        if (it != g_tune.end()) {
            *tq = it->second;
            return ALT3_OK;
        }
    }
    int nc = 0;
    const int *cand = tile_candidates(ctx->nb, nc);
    float best = 3.4e38f;
    int best_tq = 0;
    alt3_ctx *one[1] = {ctx};
    for (int c = 0; c < nc; c++) {
        if (rows_lds(ctx->nb, ctx->V.ndr, cand[c], fp32 ? 4 : 8) > ctx->lds_max) continue;
        for (int rep = 0; rep < 2; rep++) {   // the first launch pays the kernel load
            ALT3_TRY(fp32 ? batch_run<float>(one, 1, cand[c], true) : batch_run<double>(one, 1, cand[c], false));
            float ms = 0;
            ALT3_HIP(hipEventElapsedTime(&ms, ctx->ev_k[0], ctx->ev_k[1]));
            if (rep == 1 && ms < best) {
                best = ms;
                best_tq = cand[c];
            }
        }
    }
    if (!best_tq) return fail(ALT3_E_ARG, "no V3 tile fits in LDS");
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune[key] = best_tq;
    *tq = best_tq;
    return ALT3_OK;
}

int v3_batch_launch(alt3_ctx *const *ctxs, int n, bool fp32, int tq)
{
    return fp32 ? batch_run<float>(ctxs, n, tq, true) : batch_run<double>(ctxs, n, tq, false);
}

// |V3|^2 [s][l1][l2][l3][q2] (FP64) for nq list points q1 with G' in gp, e3 / w3 per (s, q2)
// and w1 per s (the self-energy); cells of 6, 12 or 24 branches
int v3_rows_points(alt3_ctx *ctx, int nq, const double *gp, const double *ev3, const double *w3, const double *w1,
                   double *out, int *d_err)
{
    const int nk = ctx->nk, nb = ctx->nb;
    if (!(ctx->V.ndr > 0 && (nb == 6 || nb == 12 || nb == 24)))
        return fail(ALT3_E_ARG, "the self-energy kernel handles cells of 6, 12 or 24 branches");
    DBuf<double> &wq = ctx->d_wq;
    ALT3_TRY(wq.reserve((size_t)nk * nb));
    hipLaunchKernelGGL(k_lam_quarter, dim3((nk * nb + 255) / 256), dim3(256), 0, ctx->stream, nk * nb, ctx->d_eival.p,
                       wq.p);
    ALT3_HIP(hipGetLastError());
    return dispatch_rows<double, true>(ctx, nq, ctx->d_eph.p, ctx->d_evec.p, wq.p, gp, out, d_err, ev3, w3, w1);
}

}  // namespace alt3
