// Mixed-precision |V3|^2 for 24-branch cells on FP16 matrix cores with split (hi + lo)
// operands: every FP32 operand x is carried as two halves, hi = fp16(x) and lo = fp16(x - hi),
// and each product A.B is formed as Ahi.Bhi + Ahi.Blo + Alo.Bhi on v_mfma_f32_16x16x32_f16
// with FP32 accumulation.  The dropped Alo.Blo term and the rounding of lo leave a relative
// error of about 2^-21 per product, close to FP32's 2^-24, at 3 FP16 products where the FP32
// chain (v3_chain.hip) issues one FP32 product of an eighth of the rate.
//
// The algebra is that of k_v3_chain (the batched form of matrix_elt_simplified,
// THERMACOND/sub.f90:5888-6011): for q1 = s, one branch l1 and a tile of 16 mesh points q2,
//   Phi'(q2)[b,c] = sum_r e^{i q2.dR_r} G'[r][l1][b,c]          (phase GEMM over the tile)
//   U[b,l3]       = sum_c Phi'[b,c] e3[c,l3]                    (per q2, e3 = e(-q1-q2))
//   V[l2,l3]      = sum_b e2[b,l2] U[b,l3]                      (per q2, e2 = e(q2))
//   |V3|^2        = |V|^2 const33^2 / 8 |l1 l2 l3|^(-1/2)
// as real-block GEMMs:
//   phase  [Re Phi'^T | Im Phi'^T] = [Re G'; Im G']^T . [[Re E, Im E], [-Im E, Re E]]
//          M = 576 (b,c), K = 2 ndr (padded to 64), N = 16 points (one tile for Re, one for Im)
//   U      [Ur; Ui] = [[Re Phi', -Im Phi'], [Im Phi', Re Phi']] . [Re e3; Im e3]
//          M = 48 (p,b), K = 48 (p',c) (padded to 64), N = 24 (l3, padded to 32)
//   V      [Vr | Vi](l3, l2) = [Ur; Ui]^T . [[Re e2, Im e2], [-Im e2, Re e2]]
//          M = l3, K = 48 (p,b) (padded to 64), N = 48 (p'',l2)
// Operand layouts of v_mfma_f32_16x16x32_f16: A[m][k] at lane m + 16 g, B[k][n] at lane n + 16 g,
// with the eight k of lane group g = lane/16 in the eight halves of the lane; D[m][n] at lane
// n + 16 (m / 4), element m % 4.  A and B see the same k order, so each operand below is packed
// in one order of its own choosing (stated with it) and its partner in the same order.
//
// Range.  FP16 holds magnitudes up to 65504, normal down to 2^-14, subnormal in steps of 2^-24.
// The eigenvector operands e2, e3 are packed times 2^4 (components up to 16, so that the halves
// of small components stay well above the subnormal steps); G' of each (q1, l1) is scaled by its
// own sg = 2^(7 - e), where 2^e bounds max_{b,c} sum_r (|Re G'_r[b,c]| + |Im G'_r[b,c]|) over
// that (q1, l1) (k_c16_bound), so |Phi'| <= 2^7 (phases are at most 1), and the U products,
// bounded by |Phi' row|_2 |16 e3 col|_2 <= sqrt(24) 2^11, are split as they come.  The scales
// are powers of two and come out of |V|^2 exactly (V carries 2^8 sg).  The scale is per (q1, l1) because a row can
// be far below the others: at q1 = Gamma the acoustic l1 give G' at the round-off of the sum rule
// (1e-16 of the optical rows), which one global scale would flush to zero, leaving Q = 0 for
// those modes where the FP64 and FP32 paths keep a round-off value.
#include "alt3_impl.h"
#include "mesh_dev.h"

namespace {

typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef unsigned int u4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ f4 mma(h8 a, h8 b, f4 c) { return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0); }
// the three products of the split operands into one accumulator
__device__ __forceinline__ f4 mma3(h8 ah, h8 al, h8 bh, h8 bl, f4 c)
{
    c = mma(ah, bh, c);
    c = mma(ah, bl, c);
    return mma(al, bh, c);
}
__device__ __forceinline__ void split(float x, _Float16 &hi, _Float16 &lo)
{
    hi = (_Float16)x;
    lo = (_Float16)(x - (float)hi);
}
typedef _Float16 h2 __attribute__((ext_vector_type(2)));
typedef _Float16 h4 __attribute__((ext_vector_type(4)));
typedef float f2 __attribute__((ext_vector_type(2)));
// two values at once: one packed conversion for the two hi halves, one for the two lo halves
__device__ __forceinline__ void split2(float a, float b, h2 &hi, h2 &lo)
{
    const f2 x = {a, b};
    hi = __builtin_convertvector(x, h2);
    lo = __builtin_convertvector(x - __builtin_convertvector(hi, f2), h2);
}
__device__ __forceinline__ h4 cat(h2 a, h2 b) { return __builtin_shufflevector(a, b, 0, 1, 2, 3); }
__device__ __forceinline__ h8 cat(h4 a, h4 b) { return __builtin_shufflevector(a, b, 0, 1, 2, 3, 4, 5, 6, 7); }
__device__ __forceinline__ h8 flip_sign(h8 x, unsigned mask)
{
    u4 u = __builtin_bit_cast(u4, x);
    u ^= mask;
    return __builtin_bit_cast(h8, u);
}

constexpr int NB = 24;           // branches
constexpr int NN = NB * NB;      // (b,c) rows of the phase GEMM
constexpr int NTP = NN / 16;     // its M-tiles
constexpr int KT = 2;            // k-steps of 32 in every product (K <= 64)
constexpr int QT = 16;           // mesh points q2 per chunk
constexpr int NW = 8;            // waves per workgroup
constexpr int QW = QT / NW;      // points per wave in the U and V products
constexpr int CH = 4;            // chunks of QT points per workgroup
static_assert(QT * NB <= NW * 64, "one weight pair per thread");
constexpr int TPW = (NTP + NW - 1) / NW;   // phase M-tiles per wave (5: waves 0-3, 4: waves 4-7)
constexpr int KSPL = 32;         // 2 ndr <= 64: ndr <= 32
// LDS: Phi'[q][plane][b][c] halves, planes Re hi, Im hi, Re lo, Im lo (the Im plane 32 banks from
// the Re one: a U operand tile that reads Re rows b = 16..23 and Im rows b = 0..7 hits distinct
// banks); rows of 24 halves, a point every 1156 dwords (4 mod 32, 16-byte aligned for the U
// operand reads: the 16 points of a phase write, two dwords each, fall on two lanes per bank
// pair at most)
constexpr int PLANE = NB * NB;               // halves per plane
constexpr int PSTR = 4 * PLANE + 8;          // halves per point
// FP32 staging area of |V3|^2 [q][l2][l3]: l2 stride 27, a point every 650 floats (the V-tile
// writes and the row reads of the output hit at most 3 and 2 lanes per bank with 32 banks, 2
// with 64)
constexpr int LS = 27;
constexpr int OS = NB * LS + 2;

// e3 and e2 operands of every mesh point from the FP32 eigenvectors evf [k][l col][b comp]
// (complex interleaved), as halves hi then lo, 16 bytes per lane:
//   EU[k][t][nt][hl][lane]: 16 B_U[k = (p',c)][l3 = 16 nt + lane%16] (zero for l3 >= 24), row
//     k = 24 p' + c; half j of lane group g holds row 8g + j in step t = 0 and row 32 + 4g + j
//     (p' = 1, c = 8 + 4g + j) in step t = 1, j < 4; the halves j >= 4 of step 1 are zero and
//     not stored: that step's slots hold 8 bytes per lane, after the four of step 0
//   EV[k][t][nv][hl][lane]: 16 B_V[k = (p,b)][(p'',l2)], column lane%16 of tile nv: Re V(l2 = 0..15)
//     (nv = 0), Im V(l2 = 0..15) (nv = 1), Re V(16..23) | Im V(16..23) (nv = 2); half j of lane group g
//     holds row (p,b) = 4g + j (t = 0, j < 4), 16 + 4g + j - 4 (t = 0, j >= 4), 32 + 4g + j
//     (t = 1, j < 4) -- the rows the U accumulators hold; the halves j >= 4 of step t = 1 are
//     zero and not stored: that step's slots hold 8 bytes per lane, after the six of step 0
constexpr int EUF = 2 * 2 + 2;       // 16-byte operand slots per lane and point (EU: four of
                                     // step 0, four 8-byte slots of step 1)
constexpr int EVF = 3 * 2 + 3;       // (EV: six slots of step 0, six 8-byte slots of step 1)

__device__ __forceinline__ int ev_row(int t, int g, int j)
{
    return t == 0 ? (j < 4 ? 4 * g + j : 16 + 4 * g + j - 4) : (j < 4 ? 32 + 4 * g + j : -1);
}

__global__ void k_c16_pack_e(int nk, const float *__restrict__ evf, h8 *__restrict__ EU, h8 *__restrict__ EV)
{
    const size_t x = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    constexpr int per = (KT * 2 + KT * 3) * 64;   // (operand, lane) tasks per point
    if (x >= (size_t)nk * per) return;
    const int k = (int)(x / per), f = (int)(x % per);
    const int lane = f & 63, frag = f >> 6, g = lane >> 4;
    const float *e = evf + (size_t)k * NB * NB * 2;   // [l][b] complex
    h8 hi, lo;
    if (frag < KT * 2) {
        const int t = frag / 2, nt = frag % 2, l3 = 16 * nt + (lane & 15);
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const int row = t == 0 ? 8 * g + j : (j < 4 ? 32 + 4 * g + j : -1);
            float v = 0.0f;
            if (row >= 0 && l3 < NB) v = e[2 * (l3 * NB + row % NB) + row / NB] * 16.0f;
            _Float16 a, b;
            split(v, a, b);
            hi[j] = a;
            lo[j] = b;
        }
        if (t == 0) {
            h8 *o = EU + ((size_t)k * EUF + nt * 2) * 64 + lane;
            o[0] = hi;
            o[64] = lo;
        } else {
            h4 *o = reinterpret_cast<h4 *>(EU + ((size_t)k * EUF + 4) * 64) + nt * 2 * 64 + lane;
            o[0] = __builtin_shufflevector(hi, hi, 0, 1, 2, 3);
            o[64] = __builtin_shufflevector(lo, lo, 0, 1, 2, 3);
        }
    } else {
        const int fr = frag - KT * 2, t = fr / 3, nv = fr % 3;
        // columns: Re V(l2 = 0..15), Im V(l2 = 0..15), then Re V(16..23) | Im V(16..23)
        const int cl = lane & 15, p2 = nv < 2 ? nv : (cl >= 8 ? 1 : 0), l2 = nv < 2 ? cl : 16 + (cl & 7);
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const int row = ev_row(t, g, j);
            float v = 0.0f;
            if (row >= 0) {
                const int p = row / NB, b = row % NB;
                const float er = e[2 * (l2 * NB + b)], ei = e[2 * (l2 * NB + b) + 1];
                v = ((p == p2) ? er : (p == 0 ? ei : -ei)) * 16.0f;
            }
            _Float16 a, b;
            split(v, a, b);
            hi[j] = a;
            lo[j] = b;
        }
        if (t == 0) {
            h8 *o = EV + ((size_t)k * EVF + nv * 2) * 64 + lane;
            o[0] = hi;
            o[64] = lo;
        } else {
            h4 *o = reinterpret_cast<h4 *>(EV + ((size_t)k * EVF + 6) * 64) + nv * 2 * 64 + lane;
            o[0] = __builtin_shufflevector(hi, hi, 0, 1, 2, 3);
            o[64] = __builtin_shufflevector(lo, lo, 0, 1, 2, 3);
        }
    }
}

// per (s, l1): the largest sum_r (|Re G'_r[b,c]| + |Im G'_r[b,c]|) over (b,c), as float bits
// (positive floats order as unsigned integers), one atomic per wavefront (the 576 (b,c) of one
// (s, l1) are nine whole wavefronts)
__global__ void k_c16_bound(int ksub, int ndr, const double *__restrict__ gp, unsigned *__restrict__ bound)
{
    const size_t x = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    float m = 0.0f;
    if (x < (size_t)ksub * NB * NN) {
        const int n = (int)(x % NN), l1 = (int)((x / NN) % NB), s = (int)(x / ((size_t)NN * NB));
        double acc = 0.0;
        for (int r = 0; r < ndr; r++) {
            const double *g = gp + 2 * ((((size_t)s * ndr + r) * NB + l1) * NN + n);
            acc += fabs(g[0]) + fabs(g[1]);
        }
        m = (float)acc * 1.0000002f;   // round the bound up
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
    if ((threadIdx.x & 63) == 0 && x < (size_t)ksub * NB * NN) atomicMax(bound + x / NN, __float_as_uint(m));
}

__device__ __forceinline__ float g_scale(const unsigned *bound)
{
    const float b = __uint_as_float(*bound);
    if (!(b > 0.0f)) return 1.0f;
    int e;
    frexpf(b, &e);   // b < 2^e
    return ldexpf(1.0f, 7 - e);
}

// phase A operands from G' [s][r][l1][(b,c)] (FP64 complex), scaled by sg of (s, l1):
//   Gh[s][l1][t][nt][hl][lane]: [Re G'; Im G'][k][(b,c) = 16 nt + lane%16], half j of lane group
//   g of step t at k = 32 t + 8 g + j (Re G'_k for k < ndr, Im G'_{k-ndr} below 2 ndr, else 0)
__global__ void k_c16_pack_g(int ksub, int ndr, const double *__restrict__ gp, const unsigned *__restrict__ bound,
                             h8 *__restrict__ Gh)
{
    const size_t x = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t per = (size_t)KT * NTP * 64;
    if (x >= (size_t)ksub * NB * per) return;
    const int sl = (int)(x / per), f = (int)(x % per);
    const float sg = g_scale(bound + sl);
    const int s = sl / NB, l1 = sl % NB;
    const int lane = f & 63, frag = f >> 6, t = frag / NTP, nt = frag % NTP, g = lane >> 4;
    const int n = 16 * nt + (lane & 15);
    h8 hi, lo;
#pragma unroll
    for (int j = 0; j < 8; j++) {
        const int kk = 32 * t + 8 * g + j;
        float v = 0.0f;
        if (kk < 2 * ndr) {
            const int r = kk < ndr ? kk : kk - ndr;
            v = (float)gp[2 * ((((size_t)s * ndr + r) * NB + l1) * NN + n) + (kk < ndr ? 0 : 1)] * sg;
        }
        _Float16 a, b;
        split(v, a, b);
        hi[j] = a;
        lo[j] = b;
    }
    h8 *o = Gh + ((size_t)sl * KT * NTP * 2 + (t * NTP + nt) * 2) * 64 + lane;
    o[0] = hi;
    o[64] = lo;
}

// lane l ^ 8 of the row of 16 lanes (DPP row_ror:8), for the Im V columns of the V tiles
__device__ __forceinline__ float xor8(float x)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x128, 0xf, 0xf, false));
}

// One workgroup of 8 waves per (q1 = s, tile of CH x 16 mesh points q2), looping over l1, as
// k_v3_chain: the G' operands of (s, l1) in the registers of the waves that use them for all CH
// chunks; per chunk the phase GEMM into LDS (as split halves), barrier, U and V per point with
// |V3|^2 into the staging area, barrier, rows out.  Work items (tile, s) are dealt to the XCDs in
// contiguous ranges, tile-major: the workgroups resident on one XCD at a time hold one or two
// tiles of points and read their eigenvector operands (15 KB a point) from that XCD's L2.
__global__ void __launch_bounds__(NW * 64) k_v3_chain16(alt3::MeshDev m, int nk, int ibz_lo, int ksub, int ndr,
                                                       const int *__restrict__ mapinv, const float *__restrict__ ephf,
                                                       const h8 *__restrict__ Gh, const h8 *__restrict__ EU,
                                                       const h8 *__restrict__ EV, const double *__restrict__ wq,
                                                       double c33sq8, const unsigned *__restrict__ bound,
                                                       float *__restrict__ out, int *__restrict__ err)
{
    extern __shared__ __align__(16) unsigned char lds_raw[];
    _Float16 *Phi = reinterpret_cast<_Float16 *>(lds_raw);              // [QT][PSTR]
    float *Ost = reinterpret_cast<float *>(Phi + QT * PSTR);             // [QT][OS]
    h8 *Aph = reinterpret_cast<h8 *>(Ost + QT * OS);                     // [CH][KT][4][64]
    __shared__ float w2s[QT][NB], w3s[QT][NB];     // |lambda|^(-1/2) at q2, and at q3 times cs (FP32)
    __shared__ int k3s[CH * QT];
    __shared__ int bad;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    // blocks go round-robin to the 8 XCDs: XCD x runs items x*per .. x*per + per - 1 in order
    const int ntile = (nk + CH * QT - 1) / (CH * QT), nitem = ntile * ksub, per = (nitem + 7) / 8;
    const int item = (int)(blockIdx.x % 8) * per + (int)(blockIdx.x / 8);
    if (item >= nitem) return;
    const int s = item % ksub, q00 = (item / ksub) * (CH * QT);
    const int nch = min(CH, (nk - q00 + QT - 1) / QT);
    const int k1 = mapinv[ibz_lo + s];
    int a0, a1, a2;
    m.coords(k1, a0, a1, a2);
    if (tid == 0) bad = 0;
    if (tid < CH * QT) {
        int b0, b1, b2;
        m.coords(min(q00 + tid, nk - 1), b0, b1, b2);
        k3s[tid] = m.neg_sum(a0, a1, a2, b0, b1, b2);
    }
    const int col = lane & 15, grp = lane >> 4;
    // phase B operands of every chunk: lane (q, g) of step t holds k = 32 t + 8 g + j of
    // [Re E, Im E] rows for Re Phi' (kind 0, 1: hi, lo) and [Im E, Re E]... as below
    //   Re Phi' uses k < ndr: Re E_k, k < 2 ndr: -Im E_{k-ndr}   (kinds 0, 1)
    //   Im Phi' uses k < ndr: Im E_k, k < 2 ndr:  Re E_{k-ndr}   (kinds 2, 3)
    for (int x = tid; x < CH * KT * 64; x += NW * 64) {
        const int ln = x & 63, ct = x >> 6, t = ct % KT, ch = ct / KT;
        const float *e = ephf + 2 * (size_t)min(q00 + ch * QT + (ln & 15), nk - 1) * ndr;
        h8 rh, rl, ih, il;
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const int kk = 32 * t + 8 * (ln >> 4) + j;
            float vr = 0.0f, vi = 0.0f;
            if (kk < 2 * ndr) {
                const int r = kk < ndr ? kk : kk - ndr;
                vr = kk < ndr ? e[2 * r] : -e[2 * r + 1];
                vi = kk < ndr ? e[2 * r + 1] : e[2 * r];
            }
            _Float16 a, b;
            split(vr, a, b);
            rh[j] = a;
            rl[j] = b;
            split(vi, a, b);
            ih[j] = a;
            il[j] = b;
        }
        h8 *o = Aph + ((ch * KT + t) * 4) * 64 + ln;
        o[0] = rh;
        o[64] = rl;
        o[128] = ih;
        o[192] = il;
    }
    h8 gh[TPW][KT], gl[TPW][KT];
    auto load_g = [&](int l1) {
        const h8 *G = Gh + (size_t)(s * NB + l1) * KT * NTP * 2 * 64 + lane;
#pragma unroll
        for (int j = 0; j < TPW; j++)
#pragma unroll
            for (int t = 0; t < KT; t++) {
                const int nt = min(wave + NW * j, NTP - 1);
                gh[j][t] = G[((size_t)(t * NTP + nt) * 2) * 64];
                gl[j][t] = G[((size_t)(t * NTP + nt) * 2 + 1) * 64];
            }
    };
    // e3 operands of one point loaded one step ahead of their use (before the barrier for the
    // wave's first point of a chunk, during V for the next); e2 operands where V starts (the
    // compiler issues them among U's products)
    h8 puh[KT][2], pul[KT][2], pvh[KT][3], pvl[KT][3];
    auto load_u = [&](int ch, int q) {
        const h8 *eu = EU + (size_t)k3s[ch * QT + q] * EUF * 64;
        const h4 *eu1 = reinterpret_cast<const h4 *>(eu + 4 * 64) + lane;
#pragma unroll
        for (int nt = 0; nt < 2; nt++) {
            puh[0][nt] = eu[(nt * 2) * 64 + lane];
            pul[0][nt] = eu[(nt * 2 + 1) * 64 + lane];
            puh[1][nt] = cat(eu1[(nt * 2) * 64], h4{});
            pul[1][nt] = cat(eu1[(nt * 2 + 1) * 64], h4{});
        }
    };
    // wq at q2 and q3 of this thread's (q, l) of the next chunk, loaded during U and V of the
    // current one and stored as the weights once its epilogue has read them
    double wn2 = 0.0, wn3 = 0.0;
    auto load_w = [&](int ch) {
        if (tid < QT * NB) {
            const int q = tid / NB, l = tid % NB;
            wn2 = wq[(size_t)min(q00 + ch * QT + q, nk - 1) * NB + l];
            wn3 = wq[(size_t)k3s[ch * QT + q] * NB + l];
        }
    };
    auto store_w = [&](float cs) {
        if (tid < QT * NB) {
            const int q = tid / NB, l = tid % NB;
            if (!(wn2 <= 1.7e308) || !(wn3 <= 1.7e308)) bad = 1;
            w2s[q][l] = (float)(wn2 * wn2);
            w3s[q][l] = cs * (float)(wn3 * wn3);   // with the l1 factor
        }
    };
    // const33^2/8 |l1|^(-1/2) (2^8 sg)^-2 (the scales come out); a zero eigenvalue in a triple
    // (w = inf) is flagged from the tables, every triple of the tile being computed
    auto cs_of = [&](int l1) {
        const double w1 = wq[(size_t)k1 * NB + l1];
        const float sg = g_scale(bound + s * NB + l1);
        if (!(w1 <= 1.7e308)) bad = 1;
        return (float)(c33sq8 / (65536.0 * (double)sg * (double)sg) * w1 * w1);
    };
    auto load_v = [&](int ch, int q) {
        const h8 *ev = EV + (size_t)min(q00 + ch * QT + q, nk - 1) * EVF * 64;
        const h4 *ev1 = reinterpret_cast<const h4 *>(ev + 6 * 64) + lane;
#pragma unroll
        for (int nv = 0; nv < 3; nv++) {
            pvh[0][nv] = ev[(nv * 2) * 64 + lane];
            pvl[0][nv] = ev[(nv * 2 + 1) * 64 + lane];
            pvh[1][nv] = cat(ev1[(nv * 2) * 64], h4{});
            pvl[1][nv] = cat(ev1[(nv * 2 + 1) * 64], h4{});
        }
    };
    __syncthreads();
    load_w(0);
    store_w(cs_of(0));
    for (int l1 = 0; l1 < NB; l1++) {
        load_g(l1);
        for (int ch = 0; ch < nch; ch++) {
            // the chunk after this one
            const int nl1 = ch + 1 < nch ? l1 : l1 + 1, nc = ch + 1 < nch ? ch + 1 : 0;
            const int q0 = q00 + ch * QT;
            // ---- phase GEMM of the chunk: D[(b,c)][q] at lane q + 16 (m/4), m = 4 grp + i
            h8 erh[KT], erl[KT], eih[KT], eil[KT];
#pragma unroll
            for (int t = 0; t < KT; t++) {
                const h8 *a = Aph + ((ch * KT + t) * 4) * 64 + lane;
                erh[t] = a[0];
                erl[t] = a[64];
                eih[t] = a[128];
                eil[t] = a[192];
            }
#pragma unroll
            for (int j = 0; j < TPW; j++) {
                const int nt = wave + NW * j;
                if (j < TPW - 1 || nt < NTP) {   // the same for every lane of the wave
                    f4 cr = {0.f, 0.f, 0.f, 0.f}, ci = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
                    for (int t = 0; t < KT; t++) {
                        cr = mma3(gh[j][t], gl[j][t], erh[t], erl[t], cr);
                        ci = mma3(gh[j][t], gl[j][t], eih[t], eil[t], ci);
                    }
                    // (b,c) = 16 nt + 4 grp + i: four consecutive c of one row b
                    _Float16 *dst = Phi + col * PSTR + 16 * nt + 4 * grp;
                    h2 rh0, rl0, rh1, rl1, ih0, il0, ih1, il1;
                    split2(cr[0], cr[1], rh0, rl0);
                    split2(cr[2], cr[3], rh1, rl1);
                    split2(ci[0], ci[1], ih0, il0);
                    split2(ci[2], ci[3], ih1, il1);
                    *reinterpret_cast<h4 *>(dst) = cat(rh0, rh1);
                    *reinterpret_cast<h4 *>(dst + PLANE) = cat(ih0, ih1);
                    *reinterpret_cast<h4 *>(dst + 2 * PLANE) = cat(rl0, rl1);
                    *reinterpret_cast<h4 *>(dst + 3 * PLANE) = cat(il0, il1);
                }
            }
            load_u(ch, wave * QW);
            __syncthreads();
            if (nl1 < NB) load_w(nc);
            // ---- U and V, one point at a time per wave
            for (int jq = 0; jq < QW; jq++) {
                const int q = wave * QW + jq;
                const _Float16 *ph = Phi + q * PSTR;
                f4 u[3][2];
#pragma unroll
                for (int a = 0; a < 3; a++)
#pragma unroll
                    for (int b = 0; b < 2; b++) u[a][b] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int t = 0; t < KT; t++) {
                    // k = (p',c) of lane group grp: 8 grp .. + 7 (t = 0: p' = grp / 3, c0 = 8 (grp % 3)),
                    // 32 + 4 grp .. + 3 (t = 1: p' = 1, c0 = 8 + 4 grp; the upper four halves meet
                    // zero e3 operands and are zero here too)
                    const int pp = t == 0 ? grp / 3 : 1, c0 = t == 0 ? 8 * (grp % 3) : 8 + 4 * grp;
#pragma unroll
                    for (int mt = 0; mt < 3; mt++) {
                        const int row = 16 * mt + col, p = row >= NB ? 1 : 0, b = row - p * NB;
                        // Re Phi' on the diagonal blocks, -Im Phi' above, Im Phi' below
                        const _Float16 *src = ph + (p == pp ? 0 : PLANE) + b * NB + c0;
                        const unsigned mask = (p == 0 && pp == 1) ? 0x80008000u : 0u;
                        h8 ah, al;
                        if (t == 0) {
                            ah = *reinterpret_cast<const h8 *>(src);
                            al = *reinterpret_cast<const h8 *>(src + 2 * PLANE);
                        } else {
                            ah = cat(*reinterpret_cast<const h4 *>(src), h4{});
                            al = cat(*reinterpret_cast<const h4 *>(src + 2 * PLANE), h4{});
                        }
                        ah = flip_sign(ah, mask);
                        al = flip_sign(al, mask);
#pragma unroll
                        for (int nt = 0; nt < 2; nt++) u[mt][nt] = mma3(ah, al, puh[t][nt], pul[t][nt], u[mt][nt]);
                    }
                }
                if (jq + 1 < QW) load_u(ch, q + 1);
                // V: the U accumulators, split, are the A operands
                load_v(ch, q);
                f4 v[2][3];
#pragma unroll
                for (int a = 0; a < 2; a++)
#pragma unroll
                    for (int b = 0; b < 3; b++) v[a][b] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int t = 0; t < KT; t++) {
#pragma unroll
                    for (int mv = 0; mv < 2; mv++) {
                        const f4 x0 = t == 0 ? u[0][mv] : u[2][mv];
                        h2 h0, l0, h1, l1, h2_, l2_, h3, l3_;
                        split2(x0[0], x0[1], h0, l0);
                        split2(x0[2], x0[3], h1, l1);
                        if (t == 0) {
                            split2(u[1][mv][0], u[1][mv][1], h2_, l2_);
                            split2(u[1][mv][2], u[1][mv][3], h3, l3_);
                        } else {
                            h2_ = l2_ = h3 = l3_ = h2{};
                        }
                        const h8 ah = cat(cat(h0, h1), cat(h2_, h3)), al = cat(cat(l0, l1), cat(l2_, l3_));
#pragma unroll
                        for (int nv = 0; nv < 3; nv++) v[mv][nv] = mma3(ah, al, pvh[t][nv], pvl[t][nv], v[mv][nv]);
                    }
                }
                // |V|^2: element i of V tile (mv, nv) at lane l is row l3 = 16 mv + 4(l/16) + i; tiles
                // 0 and 1 hold Re and Im V of l2 = l%16, tile 2 Re V(16 + l%16) in lanes l%16 < 8
                // and the matching Im V in lane l ^ 8
                const float wa = w2s[q][col], wb = w2s[q][16 + (col & 7)];
                float w3r[2][4];
#pragma unroll
                for (int mv = 0; mv < 2; mv++)
#pragma unroll
                    for (int i = 0; i < 4; i++) w3r[mv][i] = w3s[q][min(16 * mv + 4 * grp + i, NB - 1)];
#pragma unroll
                for (int mv = 0; mv < 2; mv++)
#pragma unroll
                    for (int i = 0; i < 4; i++) {
                        const int l3 = 16 * mv + 4 * grp + i;
                        const float sa = fmaf(v[mv][1][i], v[mv][1][i], v[mv][0][i] * v[mv][0][i]);
                        const float s2 = v[mv][2][i] * v[mv][2][i];
                        const float sb = s2 + xor8(s2);
                        if (l3 < NB) {
                            Ost[q * OS + col * LS + l3] = sa * (w3r[mv][i] * wa);   // l2 = col
                            if (col < 8)                                               // l2 = 16 + col
                                Ost[q * OS + (16 + col) * LS + l3] = sb * (w3r[mv][i] * wb);
                        }
                    }
            }
            __syncthreads();
            if (nl1 < NB) store_w(cs_of(nl1));
            // ---- rows (l1, l2, l3) of the chunk: 16 consecutive q2 each, as float4, stored
            // non-temporal: the rows are not read again by this kernel, and cached they would
            // push the eigenvector operands the workgroups of the XCD share out of its L2
            float *o = out + (size_t)(s * NB + l1) * NN * nk + q0;
            for (int x = tid; x < NN * (QT / 4); x += NW * 64) {
                const int row = x / (QT / 4), c4 = x % (QT / 4);
                if (q0 + 4 * c4 < nk) {
                    const float *src = Ost + 4 * c4 * OS + (row / NB) * LS + row % NB;
                    const f4 val = {src[0], src[OS], src[2 * OS], src[3 * OS]};
                    __builtin_nontemporal_store(val, reinterpret_cast<f4 *>(o + (size_t)row * nk + 4 * c4));
                }
            }
        }
    }
    __syncthreads();
    if (tid == 0 && bad) atomicExch(err, 1);   // a zero eigenvalue in a triple (den = 0)
}

}  // namespace

namespace alt3 {

bool v3_chain16_supported(const alt3_ctx *ctx)
{
    return ctx->nb == NB && ctx->V.ndr > 0 && ctx->V.ndr <= KSPL && ctx->nk % 4 == 0;
}

size_t v3_chain16_pack_floats(int nk, int ksub)
{
    // EU + EV per mesh point, G' per (s, l1); counted in floats (16 bytes = 4 floats per slot)
    return ((size_t)nk * (EUF + EVF) * 64 + (size_t)ksub * NB * KT * NTP * 2 * 64) * 4 + (size_t)ksub * NB + 64;
}

// G' must be prepared (v3_prep_launch) and wq (|lambda|^(-1/4)) filled; out: ksub*nb^3*nk floats
int v3_chain16_run(alt3_ctx *ctx, int ksub, const float *evf, const float *ephf, const double *wq, float *out,
                   int *d_err)
{
    const int nk = ctx->nk, ndr = ctx->V.ndr;
    if (!v3_chain16_supported(ctx)) return fail(ALT3_E_ARG, "FP16 split chain: unsupported cell");
    DBuf<float> &pk = ctx->d_chain_e;
    ctx->d_chain_g.release();
    const size_t neu = (size_t)nk * EUF * 64, nev = (size_t)nk * EVF * 64, ng = (size_t)ksub * NB * KT * NTP * 2 * 64;
    ALT3_TRY(pk.reserve(v3_chain16_pack_floats(nk, ksub)));
    h8 *EU = reinterpret_cast<h8 *>(pk.p), *EV = EU + neu, *Gh = EV + nev;
    unsigned *bound = reinterpret_cast<unsigned *>(Gh + ng);
    ALT3_HIP(hipMemsetAsync(bound, 0, sizeof(unsigned) * ksub * NB, ctx->stream));
    const size_t ne_tasks = (size_t)nk * (KT * 2 + KT * 3) * 64;
    hipLaunchKernelGGL(k_c16_pack_e, dim3((unsigned)((ne_tasks + 255) / 256)), dim3(256), 0, ctx->stream, nk, evf, EU,
                       EV);
    const size_t nb_tasks = (size_t)ksub * NB * NN;
    hipLaunchKernelGGL(k_c16_bound, dim3((unsigned)((nb_tasks + 255) / 256)), dim3(256), 0, ctx->stream, ksub, ndr,
                       ctx->d_gp.p, bound);
    const size_t ng_tasks = (size_t)ksub * NB * KT * NTP * 64;
    hipLaunchKernelGGL(k_c16_pack_g, dim3((unsigned)((ng_tasks + 255) / 256)), dim3(256), 0, ctx->stream, ksub, ndr,
                       ctx->d_gp.p, bound, Gh);
    ALT3_HIP(hipGetLastError());
    const size_t lds = (size_t)QT * PSTR * sizeof(_Float16) + (size_t)QT * OS * sizeof(float) +
                       (size_t)CH * KT * 4 * 64 * sizeof(h8);
    if (lds + 2 * QT * NB * sizeof(double) + CH * QT * sizeof(int) + 64 > ctx->lds_max)
        return fail(ALT3_E_ARG, "FP16 split chain tile does not fit in LDS");
    const double c33 = const33();
    const int nitem = ((nk + CH * QT - 1) / (CH * QT)) * ksub;
    dim3 grid((unsigned)(8 * ((nitem + 7) / 8)));
    ALT3_HIP(hipFuncSetAttribute((const void *)k_v3_chain16, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    kernel_mark(ctx, 0, false);
    hipLaunchKernelGGL(k_v3_chain16, grid, dim3(NW * 64), lds, ctx->stream, MeshDev::make(ctx->nc), nk, ctx->ibz_lo,
                       ksub, ndr, ctx->d_mapinv.p, ephf, Gh, EU, EV, wq, c33 * c33 / 8, bound, out, d_err);
    ALT3_HIP(hipGetLastError());
    kernel_mark(ctx, 0, true);
    return ALT3_OK;
}

}  // namespace alt3
