// Mixed-precision |V3|^2 for 24-branch cells with every contraction on FP32 matrix cores
// (v_mfma_f32_16x16x4f32), the eigenvector products chained through the MFMA registers.
//
// The algebra is the batched form of matrix_elt_simplified (THERMACOND/sub.f90:5888-6011)
// used by k_v3_rows: for q1 = s, one branch l1 and a tile of 16 mesh points q2,
//   Phi'(q2)[b,c] = sum_r e^{i q2.dR_r} G'[r][l1][b,c]          (phase GEMM over the tile)
//   U[b,l3]       = sum_c Phi'[b,c] e3[c,l3]                    (per q2, e3 = e(-q1-q2))
//   V[l2,l3]      = sum_b e2[b,l2] U[b,l3]                      (per q2, e2 = e(q2))
//   |V3|^2        = |V|^2 const33^2 / 8 |l1 l2 l3|^(-1/2)
// Each complex product is a real GEMM on its real blocks:
//   phase  [Re Phi'; Im Phi'] = [[Re E, -Im E], [Im E, Re E]] . [Re G'; Im G']
//          M = 32 (re/im rows of the 16 points), K = 2 ndr (padded to 4), N = 576 (b,c)
//   U      [Ur; Ui] = [[Re Phi', -Im Phi'], [Im Phi', Re Phi']] . [Re e3; Im e3]
//          M = 48 (p,b), K = 48 (p',c), N = 24 (l3, padded to 32)
//   V      [Vr | Vi](l3, l2) = [Ur; Ui]^T . [[Re e2, Im e2], [-Im e2, Re e2]]
//          M = l3, K = 48 (p,b), N = 48 (p'',l2)
// The U accumulators are the A operands of V without leaving the registers: lane l, element i
// of the U tile (mt, nt) holds U[(p,b) = 16 mt + 4(l/16) + i][l3 = 16 nt + l%16], which is
// exactly the A operand A[m = l3][k] of a 16x16x4 step whose four k lie at rows 16 mt + 4k' + i;
// the e2 operand is packed per mesh point in that permuted k order.  The phase output goes
// through LDS (the points of the tile are spread over the lanes there, one point per wave
// after), the |V3|^2 of the tile through an LDS staging area so that every row of the output
// [s][l1][l2][l3][q2] is written as 16 consecutive q2.
//
// Operand layouts of v_mfma_f32_16x16x4f32: A[m][k] at lane m + 16 k, B[k][n] at lane n + 16 k,
// D[m][n] at lane n + 16 (m / 4), element m % 4.
#include "alt3_impl.h"
#include "mesh_dev.h"

namespace {

typedef float f4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ f4 mma(float a, float b, f4 c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }

constexpr int NB = 24;           // branches
constexpr int NN = NB * NB;      // (b,c) columns of the phase GEMM
constexpr int NTP = NN / 16;     // its N-tiles
constexpr int MU = 2 * NB / 16;  // M-tiles of U, rows (p,b)
constexpr int NU = 2;            // N-tiles of U, columns l3 (24 of 32 used)
constexpr int KU = 2 * NB / 4;   // k-steps of U, (p',c)
constexpr int NV = 2 * NB / 16;  // N-tiles of V, columns (p'',l2)
constexpr int FU = KU * NU;      // e3 operand fragments per mesh point
constexpr int FV = MU * 4 * NV;  // e2 operand fragments per mesh point
constexpr int QT = 16;           // mesh points q2 per workgroup
constexpr int NW = 8;            // waves per workgroup
constexpr int QW = QT / NW;      // points per wave in the U and V products
constexpr int KSMAX = 16;        // phase k-steps held in registers: ndr <= 32
// LDS layouts, padded against bank conflicts (32 banks of 4 bytes):
//   Phi'[q][part][b][c]: b stride 26 (the 16 rows b of a U operand read fall on distinct even
//   banks, the next k-group on the odd ones), q stride 4 mod 8 (the four lane groups of a phase
//   write on distinct banks)
//   Ost[q][l2][l3]: l2 stride 25 (the 16 columns l2 of a V tile on distinct banks), q stride odd
constexpr int BS = 26;
constexpr int PP = NB * BS;          // one part (re or im) of one point
constexpr int PS = 2 * PP + 4;       // one point
constexpr int OS = NB * 25 + 1;      // one point of the staging area

// e3 and e2 operand fragments of every mesh point from the FP32 eigenvectors evf
// [k][l col][b comp] (complex interleaved):
//   PU[k][t][nt][lane] = B_U[(p',c) = 4t + lane/16][l3 = 16 nt + lane%16]  (0 for l3 >= 24)
//   PV[k][mt][i][nv][lane] = B_V[(p,b) = 16 mt + 4(lane/16) + i][(p'',l2) = 16 nv + lane%16]
__global__ void k_chain_pack_e(int nk, const float *__restrict__ evf, float *__restrict__ PU, float *__restrict__ PV)
{
    const size_t x = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    constexpr int per = (FU + FV) * 64;
    if (x >= (size_t)nk * per) return;
    const int k = (int)(x / per), f = (int)(x % per);
    const int lane = f & 63, frag = f >> 6;
    const float *e = evf + (size_t)k * NB * NB * 2;   // [l][b] complex
    if (frag < FU) {
        const int t = frag / NU, nt = frag % NU;
        const int kk = 4 * t + (lane >> 4), l3 = 16 * nt + (lane & 15);
        const int pp = kk / NB, c = kk % NB;
        PU[(size_t)k * FU * 64 + f] = l3 < NB ? e[2 * (l3 * NB + c) + pp] : 0.0f;
    } else {
        const int g = frag - FU, nv = g % NV, i = (g / NV) % 4, mt = g / (NV * 4);
        const int kk = 16 * mt + 4 * (lane >> 4) + i, n = 16 * nv + (lane & 15);
        const int p = kk / NB, b = kk % NB, p2 = n / NB, l2 = n % NB;
        const float er = e[2 * (l2 * NB + b)], ei = e[2 * (l2 * NB + b) + 1];
        const float v = (p == p2) ? er : (p == 0 ? ei : -ei);
        PV[(size_t)k * FV * 64 + (f - FU * 64)] = v;
    }
}

// phase B operands from G' [s][r][l1][(b,c)] (FP64 complex):
//   Gpk[s][l1][t][nt][lane] = [Re G'; Im G'][k = 4t + lane/16][(b,c) = 16 nt + lane%16]
__global__ void k_chain_pack_g(int ksub, int ndr, int ks, const double *__restrict__ gp, float *__restrict__ Gpk)
{
    const size_t x = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t per = (size_t)ks * NTP * 64;
    if (x >= (size_t)ksub * NB * per) return;
    const int sl = (int)(x / per), f = (int)(x % per);
    const int s = sl / NB, l1 = sl % NB;
    const int lane = f & 63, frag = f >> 6, t = frag / NTP, nt = frag % NTP;
    const int kk = 4 * t + (lane >> 4), n = 16 * nt + (lane & 15);
    float v = 0.0f;
    if (kk < 2 * ndr) {
        const int r = kk < ndr ? kk : kk - ndr;
        v = (float)gp[2 * ((((size_t)s * ndr + r) * NB + l1) * NN + n) + (kk < ndr ? 0 : 1)];
    }
    Gpk[x] = v;
}

constexpr int CH = 4;                // chunks of QT points per workgroup
constexpr int TPW = (NTP + NW - 1) / NW;   // phase N-tiles per wave (5: waves 0-3, 4: waves 4-7)

// One workgroup of 8 waves per (q1 = s, tile of CH x 16 mesh points q2), looping over l1.  The
// phase B operands of (s, l1) -- G', the largest operand -- are read from memory once per l1
// into the registers of the waves that use them and serve all CH chunks of 16 points (with one
// chunk per workgroup, as before, the phase GEMM waited on these loads for every 16 points).
// Per chunk: phase GEMM into LDS, barrier, U and V per point with |V3|^2 into the staging
// area, barrier, rows out (overlapping the next chunk's phase GEMM, which writes only Phi).
template <int KS>
__global__ void __launch_bounds__(NW * 64) k_v3_chain(alt3::MeshDev m, int nk, int ibz_lo, int ndr, int ks,
                                                     const int *__restrict__ mapinv, const float *__restrict__ ephf,
                                                     const float *__restrict__ Gpk, const float *__restrict__ PU,
                                                     const float *__restrict__ PV, const double *__restrict__ wq,
                                                     double c33sq8, float *__restrict__ out, int *__restrict__ err)
{
    extern __shared__ __align__(16) float lds[];
    float *Phi = lds;                  // [QT][2][NB][BS]: Re Phi'[b][c], then Im
    float *Ost = Phi + QT * PS;        // [QT][NB][25] scaled |V3|^2 (l2, l3) of the current l1
    float *Ao = Ost + QT * OS;         // [CH][KS][64] phase A operands (rows of Re Phi'), per lane
    float *Ai = Ao + CH * KS * 64;     // [CH][KS][64] (rows of Im Phi')
    __shared__ double w2s[QT][NB], w3s[QT][NB];
    __shared__ int k3s[CH * QT];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int s = blockIdx.y, q00 = blockIdx.x * (CH * QT);
    const int nch = min(CH, (nk - q00 + QT - 1) / QT);   // chunks with at least one mesh point
    const int k1 = mapinv[ibz_lo + s];
    int a0, a1, a2;
    m.coords(k1, a0, a1, a2);
    if (tid < CH * QT) {
        int b0, b1, b2;
        m.coords(min(q00 + tid, nk - 1), b0, b1, b2);
        k3s[tid] = m.neg_sum(a0, a1, a2, b0, b1, b2);
    }
    // phase A operands of every chunk in the lane order of the MFMA: lane (col, grp) of k-step t
    // holds row q = col, k = 4t + grp of [Re E, -Im E] (Ao) and [Im E, Re E] (Ai)
    for (int x = tid; x < CH * KS * 64; x += NW * 64) {
        const int ln = x & 63, ct = x >> 6, t = ct % KS, ch = ct / KS;
        const int kk = 4 * t + (ln >> 4);
        float vr = 0.0f, vi = 0.0f;
        if (kk < 2 * ndr) {
            const int r = kk < ndr ? kk : kk - ndr;
            const float *e = ephf + 2 * ((size_t)min(q00 + ch * QT + (ln & 15), nk - 1) * ndr + r);
            vr = kk < ndr ? e[0] : -e[1];
            vi = kk < ndr ? e[1] : e[0];
        }
        Ao[x] = vr;
        Ai[x] = vi;
    }
    const int col = lane & 15, grp = lane >> 4;
    // phase B operands of this wave's N-tiles nt = wave + 8 j
    float gb[TPW][KS];
    // (the packs hold KS k-steps, zero beyond 2 ndr: no guards, so that the MFMA chains below
    // are straight-line code; a wave without a fifth tile reads the last tile again, unused)
    auto load_g = [&](int l1) {
        const float *G = Gpk + (size_t)(s * NB + l1) * KS * NTP * 64 + lane;
#pragma unroll
        for (int j = 0; j < TPW; j++)
#pragma unroll
            for (int t = 0; t < KS; t++) {
                const int nt = min(wave + NW * j, NTP - 1);
                gb[j][t] = G[((size_t)t * NTP + nt) * 64];
            }
    };
    __syncthreads();
    for (int l1 = 0; l1 < NB; l1++) {
        const double w1 = wq[(size_t)k1 * NB + l1];
        load_g(l1);
        for (int ch = 0; ch < nch; ch++) {
            const int q0 = q00 + ch * QT;
            // ---- phase GEMM of the chunk: rows of point q = lane%16, k = 4t + lane/16;
            // ar: rows of Re Phi' ([Re E, -Im E]), ai: rows of Im Phi' ([Im E, Re E])
            float ar[KS], ai[KS];
#pragma unroll
            for (int t = 0; t < KS; t++) {
                ar[t] = Ao[(ch * KS + t) * 64 + lane];
                ai[t] = Ai[(ch * KS + t) * 64 + lane];
            }
#pragma unroll
            for (int j = 0; j < TPW; j++) {
                const int nt = wave + NW * j;
                if (j < TPW - 1 || nt < NTP) {   // the same for every lane of the wave
                    f4 cr = {0.f, 0.f, 0.f, 0.f}, ci = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
                    for (int t = 0; t < KS; t++) {
                        cr = mma(ar[t], gb[j][t], cr);
                        ci = mma(ai[t], gb[j][t], ci);
                    }
                    const int n = 16 * nt + col, bc = (n / NB) * BS + n % NB;
#pragma unroll
                    for (int i = 0; i < 4; i++) {
                        Phi[(4 * grp + i) * PS + bc] = cr[i];
                        Phi[(4 * grp + i) * PS + PP + bc] = ci[i];
                    }
                }
            }
            for (int x = tid; x < QT * NB; x += NW * 64) {
                const int q = x / NB, l = x % NB;
                w2s[q][l] = wq[(size_t)min(q0 + q, nk - 1) * NB + l];
                w3s[q][l] = wq[(size_t)k3s[ch * QT + q] * NB + l];
            }
            __syncthreads();
            // ---- U and V, one point at a time per wave
            for (int jq = 0; jq < QW; jq++) {
                const int q = wave * QW + jq;
                const int n2c = min(q0 + q, nk - 1);
                const float *ph = Phi + q * PS;
                const float *pu = PU + (size_t)k3s[ch * QT + q] * FU * 64 + lane;
                const float *pv = PV + (size_t)n2c * FV * 64 + lane;
                f4 u[MU][NU];
#pragma unroll
                for (int a = 0; a < MU; a++)
#pragma unroll
                    for (int b = 0; b < NU; b++) u[a][b] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int t = 0; t < KU; t++) {
                    float bu[NU];
#pragma unroll
                    for (int nt = 0; nt < NU; nt++) bu[nt] = pu[(t * NU + nt) * 64];
                    const int kk = 4 * t + grp, pp = kk >= NB ? 1 : 0, c = kk - pp * NB;
#pragma unroll
                    for (int mt = 0; mt < MU; mt++) {
                        const int row = 16 * mt + col, p = row >= NB ? 1 : 0, b = row - p * NB;
                        float a = ph[(p ^ pp) * PP + b * BS + c];
                        if (p == 0 && pp == 1) a = -a;
#pragma unroll
                        for (int nt = 0; nt < NU; nt++) u[mt][nt] = mma(a, bu[nt], u[mt][nt]);
                    }
                }
                f4 v[NU][NV];
#pragma unroll
                for (int a = 0; a < NU; a++)
#pragma unroll
                    for (int b = 0; b < NV; b++) v[a][b] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int mt = 0; mt < MU; mt++)
#pragma unroll
                    for (int i = 0; i < 4; i++) {
                        float bv[NV];
#pragma unroll
                        for (int nv = 0; nv < NV; nv++) bv[nv] = pv[((mt * 4 + i) * NV + nv) * 64];
#pragma unroll
                        for (int mv = 0; mv < NU; mv++)
#pragma unroll
                            for (int nv = 0; nv < NV; nv++) v[mv][nv] = mma(u[mt][mv][i], bv[nv], v[mv][nv]);
                    }
                // |V|^2: element i of V tile (mv, nv) at lane l is column n = 16 nv + l%16 of row
                // l3 = 16 mv + 4(l/16) + i; n < 24 is Re V(l2 = n), n >= 24 is Im V(l2 = n - 24),
                // found at column (n + 8) % 16 of the next tile(s): lane l ^ 8
#pragma unroll
                for (int mv = 0; mv < NU; mv++)
#pragma unroll
                    for (int i = 0; i < 4; i++) {
                        const int l3 = 16 * mv + 4 * grp + i;
                        const float s0 = v[mv][0][i] * v[mv][0][i], s1 = v[mv][1][i] * v[mv][1][i],
                                    s2 = v[mv][2][i] * v[mv][2][i];
                        const float y1 = __shfl_xor(s1, 8, 64), y2 = __shfl_xor(s2, 8, 64);
                        if (l3 < NB) {
                            const double w13 = w1 * w3s[q][l3];
                            {   // l2 = col
                                const float v2 = s0 + (col < 8 ? y1 : y2);
                                const double w = w13 * w2s[q][col];
                                if (!(w <= 1.7e308)) atomicExch(err, 1);   // a zero eigenvalue in the triple
                                Ost[q * OS + col * 25 + l3] = v2 * (float)(c33sq8 * w * w);
                            }
                            if (col < 8) {   // l2 = 16 + col
                                const float v2 = s1 + y2;
                                const double w = w13 * w2s[q][16 + col];
                                if (!(w <= 1.7e308)) atomicExch(err, 1);
                                Ost[q * OS + (16 + col) * 25 + l3] = v2 * (float)(c33sq8 * w * w);
                            }
                        }
                    }
            }
            __syncthreads();
            // ---- rows (l1, l2, l3) of the chunk: 16 consecutive q2 each, as float4
            float *o = out + (size_t)(s * NB + l1) * NN * nk + q0;
            for (int x = tid; x < NN * (QT / 4); x += NW * 64) {
                const int row = x / (QT / 4), c4 = x % (QT / 4);
                if (q0 + 4 * c4 < nk) {
                    const float *src = Ost + 4 * c4 * OS + (row / NB) * 25 + row % NB;
                    const float4 val = {src[0], src[OS], src[2 * OS], src[3 * OS]};
                    *reinterpret_cast<float4 *>(o + (size_t)row * nk + 4 * c4) = val;
                }
            }
            // the next chunk's phase GEMM writes only Phi (read by nobody now) and the w tables
            // (read by nobody now); its U and V write Ost after the barrier that follows it,
            // when these reads are done
        }
    }
}

}  // namespace

namespace alt3 {

// 24-branch cells with at most 32 distinct cell differences on a mesh of a multiple of 4
// points (so that the float4 rows stay aligned)
bool v3_chain_supported(const alt3_ctx *ctx)
{
    return ctx->nb == NB && ctx->V.ndr > 0 && (2 * ctx->V.ndr + 3) / 4 <= KSMAX && ctx->nk % 4 == 0;
}

// G' must be prepared (v3_prep_launch) and wq (|lambda|^(-1/4)) filled; out: ksub*nb^3*nk floats
int v3_chain_run(alt3_ctx *ctx, int ksub, const float *evf, const float *ephf, const double *wq, float *out,
                 int *d_err)
{
    const int nk = ctx->nk, ndr = ctx->V.ndr, ks0 = (2 * ndr + 3) / 4;
    // k-steps of the phase GEMM rounded up to a kernel instance; the packs are zero beyond 2 ndr
    const int ks = ks0 <= 8 ? 8 : ks0 <= 12 ? 12 : ks0 <= 14 ? 14 : 16;
    DBuf<float> &pe = ctx->d_chain_e, &pg = ctx->d_chain_g;
    const size_t ne = (size_t)nk * (FU + FV) * 64, ng = (size_t)ksub * NB * ks * NTP * 64;
    ALT3_TRY(pe.reserve(ne));
    ALT3_TRY(pg.reserve(ng));
    float *PU = pe.p, *PV = pe.p + (size_t)nk * FU * 64;
    hipLaunchKernelGGL(k_chain_pack_e, dim3((unsigned)((ne + 255) / 256)), dim3(256), 0, ctx->stream, nk, evf, PU, PV);
    hipLaunchKernelGGL(k_chain_pack_g, dim3((unsigned)((ng + 255) / 256)), dim3(256), 0, ctx->stream, ksub, ndr, ks,
                       ctx->d_gp.p, pg.p);
    ALT3_HIP(hipGetLastError());
    const size_t lds = (size_t)(QT * PS + QT * OS + 2 * CH * ks * 64) * sizeof(float);
    if (lds + 2 * QT * NB * sizeof(double) + CH * QT * sizeof(int) + 64 > ctx->lds_max)
        return fail(ALT3_E_ARG, "V3 chain tile does not fit in LDS");
    const double c33 = const33();
    dim3 grid((unsigned)((nk + CH * QT - 1) / (CH * QT)), (unsigned)ksub);
#define ALT3_CHAIN(KS)                                                                                                \
    {                                                                                                                 \
        ALT3_HIP(hipFuncSetAttribute((const void *)k_v3_chain<KS>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds)); \
        kernel_mark(ctx, 0, false);                                                                                   \
        hipLaunchKernelGGL(k_v3_chain<KS>, grid, dim3(NW * 64), lds, ctx->stream, MeshDev::make(ctx->nc), nk,          \
                           ctx->ibz_lo, ndr, ks, ctx->d_mapinv.p, ephf, pg.p, PU, PV, wq, c33 * c33 / 8, out, d_err); \
    }
    if (ks == 8) ALT3_CHAIN(8)
    else if (ks == 12) ALT3_CHAIN(12)
    else if (ks == 14) ALT3_CHAIN(14)
    else ALT3_CHAIN(16)
#undef ALT3_CHAIN
    ALT3_HIP(hipGetLastError());
    kernel_mark(ctx, 0, true);
    return ALT3_OK;
}

}  // namespace alt3
