// Tetrahedron-integrated three-phonon scattering weights on the GPU.
//
// Replaces calculate_FGR_mpi2 (THERMACOND/sub.f90:7199-7499) and weight_tet
// (tetrahedron.f90:188-384).  For every (q1 in the irreducible zone, l1, l2, l3) and process
// (1: coalescence, corner energy -w(q2,l2)+w(q1+q2,l3); 2: decay, w(q2,l2)+w(q1-q2,l3)) the
// linear-tetrahedron weights of delta(w(q1,l1) - e) over the 6*nk tetrahedra are
// accumulated per mesh point in LDS and multiplied by F(q2) = |V3|^2 (2pi)^2/nk with the
// acoustic-Gamma mask:
//   P_sm(q1,q2) = F(q2) * sum_{tetra containing q2} c/6.
// The corner energies and weights keep the reference's REAL*4 types (energies rounded to
// float, differences in float, mixed float/double formulae) so ties and the degenerate-gap
// clamp behave identically; FMA contraction is disabled in this file for the same reason.
#include <algorithm>

#include "alt3_impl.h"
#include "mesh_dev.h"

#pragma clang fp contract(off)

namespace {

// SSORT of the corner energies (descending, first maximum wins) and the reference's map
// uns from sorted position to corner (tetrahedron.f90:248-300).  The sorted values do not
// depend on how they are sorted (the energies are never NaN or -0), so a compare-exchange
// network on registers replaces the selection sort; uns is then built from the sorted values
// exactly as the reference builds it.  Every index below is a compile-time constant once the
// loops are unrolled, so nothing goes to per-lane scratch memory.
__device__ __forceinline__ void cx_desc(float &a, float &b)
{
    const float hi = fmaxf(a, b), lo = fminf(a, b);
    a = hi;
    b = lo;
}

__device__ __forceinline__ void tet_sort(const float *w, float *en, int *uns)
{
    en[0] = w[0];
    en[1] = w[1];
    en[2] = w[2];
    en[3] = w[3];
    cx_desc(en[0], en[1]);
    cx_desc(en[2], en[3]);
    cx_desc(en[0], en[2]);
    cx_desc(en[1], en[3]);
    cx_desc(en[1], en[2]);
    if (en[0] != en[1] && en[1] != en[2] && en[2] != en[3]) {
#pragma unroll
        for (int k = 0; k < 4; k++) uns[k] = (en[k] == w[0]) ? 1 : (en[k] == w[1]) ? 2 : (en[k] == w[2]) ? 3 : 4;
    } else {
#pragma unroll
        for (int k = 0; k < 4; k++) {
            uns[k] = (en[k] == w[0]) ? 1 : (en[k] == w[1]) ? 2 : (en[k] == w[2]) ? 3 : 4;
#pragma unroll
            for (int l = 0; l < k; l++) {
                if (uns[l] == uns[k] && uns[k] == 1) {
                    if (en[k] == w[1]) uns[k] = 2;
                    else if (en[k] == w[2]) uns[k] = 3;
                    else if (en[k] == w[3]) uns[k] = 4;
                } else if (uns[l] == uns[k] && uns[k] == 2) {
                    if (en[k] == w[2]) uns[k] = 3;
                    else if (en[k] == w[3]) uns[k] = 4;
                } else if (uns[l] == uns[k] && uns[k] == 3) {
                    uns[k] = 4;
                }
            }
        }
    }
}

// c[uns[k] - 1] = cc[k] for k = 3, 2, 1, 0 in that order (a later k overwrites), as selects
template <typename T>
__device__ __forceinline__ void unsort(const int *uns, const T *cc, T *c)
{
#pragma unroll
    for (int i = 0; i < 4; i++) {
        T v = 0;
#pragma unroll
        for (int k = 3; k >= 0; k--)
            if (uns[k] - 1 == i) v = cc[k];
        c[i] = v;
    }
}

// weight_tet for one tetrahedron: corner energies w[4] (already rounded to REAL*4) at e
__device__ void weight_tet1(const float *w, double e, double *c)
{
    float en[4];
    int uns[4];
    tet_sort(w, en, uns);
    float e1 = en[3], e2 = en[2], e3 = en[1], e4 = en[0];
    float e21 = e2 - e1, e31 = e3 - e1, e41 = e4 - e1, e32 = e3 - e2, e42 = e4 - e2, e43 = e4 - e3;
    const double small1 = 1e-10;
    const float small2 = (float)1e-10;
#define CLAMP(x)                                              \
    {                                                         \
        if (fabs((double)(x)) < small1 && (x) > 0.0f) (x) = small2;  \
        if (fabs((double)(x)) < small1 && (x) < 0.0f) (x) = -small2; \
    }
    CLAMP(e21) CLAMP(e31) CLAMP(e41) CLAMP(e42) CLAMP(e43)
#undef CLAMP
    const double E1 = e1, E2 = e2, E3 = e3, E4 = e4;
    double cc[4];
    if (e < E1 || e > E4) {
        c[0] = c[1] = c[2] = c[3] = 0;
        return;
    } else if (E1 <= e && e <= E2) {
        double d1 = e - E1;
        cc[3] = ((E2 - e) / (double)e21 + (E3 - e) / (double)e31 + (E4 - e) / (double)e41) * (d1 * d1) / (double)(e41 * e31 * e21);
        cc[2] = (d1 * d1 * d1) / (double)(e21 * e21 * e31 * e41);
        cc[1] = (d1 * d1 * d1) / (double)(e21 * (e31 * e31) * e41);
        cc[0] = (d1 * d1 * d1) / (double)(e21 * e31 * (e41 * e41));
    } else if (E2 <= e && e <= E3) {
        // the reference's quotients as divisions, rounded to REAL*4 (a reciprocal times the
        // numerator differs by an ulp of FP64 now and then, and near a degenerate corner pair that
        // ulp can move the REAL*4 rounding of a quotient whose terms nearly cancel in P)
        float c11 = (float)((E3 - e) / (double)(e31 * e31));
        float c12 = (float)((E4 - e) / (double)(e41 * e41));
        float c21 = (float)((E3 - e) / (double)(e32 * e32));
        float c22 = (float)((E4 - e) / (double)(e42 * e42));
        float c31 = (float)((e - E2) / (double)(e32 * e32));
        float c32 = (float)((e - E1) / (double)(e31 * e31));
        float c41 = (float)((e - E2) / (double)(e42 * e42));
        float c42 = (float)((e - E1) / (double)(e41 * e41));
        const double a32 = (E3 - e) * (e - E2), a41 = (E4 - e) * (e - E1), a31 = (E3 - e) * (e - E1), a42 = (E4 - e) * (e - E2);
        float b11 = (float)(a32 / (double)(e42 * e32) + a41 / (double)(e41 * e42) + a31 / (double)(e32 * e41));
        float b12 = (float)(a41 / (double)(e42 * e31) + a42 / (double)(e42 * e32) + a31 / (double)(e31 * e32));
        float b21 = (float)(a32 / (double)(e42 * e31) + a42 / (double)(e42 * e41) + a31 / (double)(e31 * e41));
        float b22 = (float)(a32 / (double)(e32 * e31) + a41 / (double)(e41 * e31) + a42 / (double)(e32 * e41));
        float b31 = b21;
        float b32 = b11;
        float b41 = b22;
        float b42 = b12;
        cc[3] = 0.5f * (c11 * b11 + c12 * b12);
        cc[2] = 0.5f * (c21 * b21 + c22 * b22);
        cc[1] = 0.5f * (c31 * b31 + c32 * b32);
        cc[0] = 0.5f * (c41 * b41 + c42 * b42);
    } else {
        double d4 = E4 - e;
        cc[3] = (d4 * d4 * d4) / (double)(e41 * e41 * e42 * e43);
        cc[2] = (d4 * d4 * d4) / (double)(e41 * (e42 * e42) * e43);
        cc[1] = (d4 * d4 * d4) / (double)(e41 * e42 * (e43 * e43));
        cc[0] = ((e - E3) / (double)e43 + (e - E2) / (double)e42 + (e - E1) / (double)e41) * (d4 * d4) / (double)(e41 * e42 * e43);
    }
    unsort(uns, cc, c);
}

// The same weights in FP32 arithmetic, for the mixed path (whose |V3|^2 and P are FP32): the
// same sort, case split and 1e-10 clamp.  The quotients are rewritten over the six hardware
// reciprocals 1/e_ij (v_rcp_f32, 1 ulp), grouped so that every partial product stays bounded
// (d/e_ij <= 1 inside its case), instead of up to twenty IEEE divisions.  e stays REAL*8 as in
// the reference: the case tests and the differences e - e_i are taken in FP64 and only then
// rounded.  (Rounding e itself to REAL*4 makes it tie exactly with corner energies -- the
// energy -w(Gamma, acoustic) + w(q1, l1) of a Gamma corner rounds to the same float as w(q1, l1)
// -- and a tie inside a degenerate pair e_i = e_j gives 0 * (1/0).)
__device__ __forceinline__ void weight_tet1f(const float *w, double e, float *c)
{
    float en[4];
    int uns[4];
    tet_sort(w, en, uns);
    const float e1 = en[3], e2 = en[2], e3 = en[1], e4 = en[0];
    float e21 = e2 - e1, e31 = e3 - e1, e41 = e4 - e1, e32 = e3 - e2, e42 = e4 - e2, e43 = e4 - e3;
    const float small = 1e-10f;
#define CLAMPF(x)                                \
    {                                            \
        if (fabsf(x) < small && (x) > 0.0f) (x) = small;  \
        if (fabsf(x) < small && (x) < 0.0f) (x) = -small; \
    }
    CLAMPF(e21) CLAMPF(e31) CLAMPF(e41) CLAMPF(e42) CLAMPF(e43)
#undef CLAMPF
    const double E1 = e1, E2 = e2, E3 = e3, E4 = e4;
    float cc[4];
    if (e < E1 || e > E4) {
        c[0] = c[1] = c[2] = c[3] = 0;
        return;
    }
    // e - e_i (>= 0 for the corners below e) and e_i - e (above), each rounded once
    const float d1 = (float)(e - E1), d2 = (float)(e - E2), u3 = (float)(E3 - e), u4 = (float)(E4 - e);
    if (e <= E2) {
        const float q21 = __builtin_amdgcn_rcpf(e21), q31 = __builtin_amdgcn_rcpf(e31), q41 = __builtin_amdgcn_rcpf(e41);
        const float g = (d1 * q21) * (d1 * q31) * q41;   // d1^2 / (e21 e31 e41)
        cc[3] = (-d2 * q21 + u3 * q31 + u4 * q41) * g;
        cc[2] = g * (d1 * q21);
        cc[1] = g * (d1 * q31);
        cc[0] = g * (d1 * q41);
    } else if (e <= E3) {
        const float q31 = __builtin_amdgcn_rcpf(e31), q41 = __builtin_amdgcn_rcpf(e41);
        const float q32 = __builtin_amdgcn_rcpf(e32), q42 = __builtin_amdgcn_rcpf(e42);
        const float a32 = u3 * d2, a41 = u4 * d1, a31 = u3 * d1, a42 = u4 * d2;
        const float i1 = q42 * q32, i2 = q41 * q42, i3 = q32 * q41, i4 = q42 * q31, i5 = q31 * q32, i6 = q31 * q41;
        const float b11 = a32 * i1 + a41 * i2 + a31 * i3, b12 = a41 * i4 + a42 * i1 + a31 * i5;
        const float b21 = a32 * i4 + a42 * i2 + a31 * i6, b22 = a32 * i5 + a41 * i6 + a42 * i3;
        cc[3] = 0.5f * ((u3 * q31) * q31 * b11 + (u4 * q41) * q41 * b12);
        cc[2] = 0.5f * ((u3 * q32) * q32 * b21 + (u4 * q42) * q42 * b22);
        cc[1] = 0.5f * ((d2 * q32) * q32 * b21 + (d1 * q31) * q31 * b11);
        cc[0] = 0.5f * ((d2 * q42) * q42 * b22 + (d1 * q41) * q41 * b12);
    } else {
        const float q41 = __builtin_amdgcn_rcpf(e41), q42 = __builtin_amdgcn_rcpf(e42), q43 = __builtin_amdgcn_rcpf(e43);
        const float d3 = (float)(e - E3);
        const float h = (u4 * q41) * (u4 * q42) * q43;   // d4^2 / (e41 e42 e43)
        cc[3] = h * (u4 * q41);
        cc[2] = h * (u4 * q42);
        cc[1] = h * (u4 * q43);
        cc[0] = (d3 * q43 + d2 * q42 + d1 * q41) * h;
    }
    unsort(uns, cc, c);
}

// corners (0..7 = bits g1,g2,g3 of the cube) of the 6 tetrahedra of make_kp_reg_tet:
// {0,1,2,5} {1,2,3,5} {0,2,4,5} {2,3,5,7} {2,4,5,6} {2,5,6,7}; column y of the table packed
// as 3 bits per tetrahedron, so that a runtime tetrahedron index needs no memory load
constexpr unsigned tc_col(int y)
{
    constexpr int t[6][4] = {{0, 1, 2, 5}, {1, 2, 3, 5}, {0, 2, 4, 5}, {2, 3, 5, 7}, {2, 4, 5, 6}, {2, 5, 6, 7}};
    unsigned v = 0;
    for (int l = 0; l < 6; l++) v |= (unsigned)t[l][y] << (3 * l);
    return v;
}
__device__ __forceinline__ int tet_corner(int l, int y)
{
    constexpr unsigned c0 = tc_col(0), c1 = tc_col(1), c2 = tc_col(2), c3 = tc_col(3);
    const unsigned col = y == 0 ? c0 : y == 1 ? c1 : y == 2 ? c2 : c3;
    return (int)((col >> (3 * l)) & 7u);
}

// One workgroup per (s, l2, l3, process).  The corner energies e(q2) depend on l2, l3 and
// the process but not on l1, so they are built once in LDS (with the energy range of every
// cube kept in registers) and reused for the nb values omega(q1, l1); per l1 the LDS weight
// sums W are accumulated (FP64 on both paths: LDS ds_add_f32 measured slower than ds_add_f64
// here) and then written out as P = W * F.
constexpr int kFgrThreads = 512;   // 8 waves; LDS (12 B per mesh point) allows 2 groups per CU
// the frequencies branch-major [l][q], so that the energy gathers of a wave are contiguous
__global__ void k_freq_t(int nk, int nb, const double *__restrict__ freq, double *__restrict__ freqT)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x;   // (l, q)
    if (x >= nk * nb) return;
    const int l = x / nk, q = x - l * nk;
    freqT[x] = freq[(size_t)q * nb + l];
}

constexpr int kFgrCase = 4096;     // LDS list per weight_tet case; one chunk of cubes (<= 6*512) always fits
constexpr int kFgrMaxNb = 64;      // branches (bits of the per-cube l1 masks)
constexpr int kFgrList = 3 * kFgrCase;

// e inside [lo, hi] as the reference tests it (no weight outside, tetrahedron.f90:343)
__device__ __forceinline__ bool omega_in(double e, float lo, float hi) { return !(e < (double)lo || e > (double)hi); }

// kFgrPer: mesh points (and cubes) per thread; MT: the per-cube l1 masks (32 bits for up to 32
// branches, 64 for up to 64)
template <typename VT, int kFgrPer, typename MT>
__global__ void __launch_bounds__(kFgrThreads, (sizeof(VT) == 4 && kFgrPer <= 4) ? 6 : 4) k_fgr(alt3::MeshDev m, int nk, int nb, int ibz_lo,
                                                     const int *__restrict__ mapinv, const double *__restrict__ freq,
                                                     const double *__restrict__ freqT,
                                                     const VT *__restrict__ v33, VT *__restrict__ P1,
                                                     VT *__restrict__ P2, unsigned char *__restrict__ prow,
                                                     size_t prow2)
{
    extern __shared__ double lds[];
    double *W = lds;               // [nk]
    float *en = reinterpret_cast<float *>(W + nk);     // [nk]
    unsigned short *s_cube = reinterpret_cast<unsigned short *>(en + nk);   // [nk] cubes active for this l1
    __shared__ float s_lo[kFgrThreads / 64], s_hi[kFgrThreads / 64];
    __shared__ int s_cnt[3], s_ncube;
    __shared__ double s_om[kFgrMaxNb];           // omega(q1, l1)
    __shared__ unsigned short s_list[kFgrList];  // [case][kFgrCase] of (cube << 3 | tetrahedron), nk <= 8192
    // The groups of one q1 read the same |V3|^2 rows (both processes, and the decay mirror the
    // (l3, l2) rows of the coalescence groups).  Workgroups are dealt to the eight XCDs round
    // robin, so work item (s, l2, l3, proc) is taken from a contiguous eighth of the items per
    // XCD: the groups of one q1 then run together on one XCD and share its L2.
    const int G = (int)gridDim.x, b = blockIdx.x;
    const int item = (G & 7) == 0 ? (b & 7) * (G >> 3) + (b >> 3) : b;
    const int proc = item & 1;
    int r = item >> 1;
    const int l3 = r % nb;
    r /= nb;
    const int l2 = r % nb;
    const int s = r / nb;
    // Mixed path: the decay energies of (l2, l3) at q2 are those of (l3, l2) at q1-q2, and the
    // mesh map q2 -> q1-q2 carries the tetrahedra of make_kp_reg_tet onto themselves (it
    // inverts each cube through its centre, which keeps the diagonal all six share), so the
    // decay group with l2 < l3 also writes the (l3, l2) rows from W(q1-q2) and the l2 > l3
    // groups have nothing to do.  (Tied corner energies may then assign the tie's weights to
    // the other corner, which the FP64 path, compared with the reference at 1e-8, avoids.)
    constexpr bool kSym = sizeof(VT) == sizeof(float);
    if (kSym && proc == 1 && l2 > l3) return;   // whole group
    const bool mirror = kSym && proc == 1 && l2 < l3;
    const int k1 = mapinv[ibz_lo + s];
    int a0, a1, a2;
    m.coords(k1, a0, a1, a2);
    const int tid = threadIdx.x, gam = m.gamma();
    float lo = 3.4e38f, hi = -3.4e38f;
    unsigned vzero = 0;  // bit j: acoustic branch at Gamma in q2 or q3
    unsigned vzerom = 0; // the same for the mirrored (l3, l2) row
#pragma unroll
    for (int j = 0; j < kFgrPer; j++) {
        const int w = tid + j * kFgrThreads;
        if (w < nk) {
            int c0, c1, c2;
            m.coords(w, c0, c1, c2);
            int p;
            double e;
            if (proc == 0) {  // -w(q2,l2) + w(q1+q2,l3)
                p = m.sum(a0, a1, a2, c0, c1, c2);
                e = -freqT[(size_t)l2 * nk + w] + freqT[(size_t)l3 * nk + p];
            } else {          // w(q2,l2) + w(q1-q2,l3)
                p = m.diff(a0, a1, a2, c0, c1, c2);
                e = freqT[(size_t)l2 * nk + w] + freqT[(size_t)l3 * nk + p];
            }
            if ((l2 < 3 && w == gam) || (l3 < 3 && p == gam)) vzero |= 1u << j;
            if ((l3 < 3 && w == gam) || (l2 < 3 && p == gam)) vzerom |= 1u << j;
            const float ef = (float)e;
            en[w] = ef;
            W[w] = 0.0;
            lo = fminf(lo, ef);
            hi = fmaxf(hi, ef);
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        lo = fminf(lo, __shfl_xor(lo, off, 64));
        hi = fmaxf(hi, __shfl_xor(hi, off, 64));
    }
    if ((tid & 63) == 0) {
        s_lo[tid >> 6] = lo;
        s_hi[tid >> 6] = hi;
    }
    if (tid < nb) s_om[tid] = freq[(size_t)k1 * nb + tid];
    // list counters: cleared here and at the end of every l1 pass that uses them
    if (tid < 3) s_cnt[tid] = 0;
    if (tid == 0) s_ncube = 0;
    __syncthreads();
    for (int w = 0; w < kFgrThreads / 64; w++) {
        lo = fminf(lo, s_lo[w]);
        hi = fmaxf(hi, s_hi[w]);
    }
    // most items reach no omega(q1, l1) at all (their rows are all zero): they skip the cube
    // ranges and masks below, and every l1 pass takes its zero-row path (uniform over the group)
    bool any = false;
    for (int l = 0; l < nb; l++) any |= omega_in(s_om[l], lo, hi);
    // energy range of cube k = tid + j*T; corners c + {0,e1,e2,e1+e2,e3,e3+e1,e3+e2,e3+e1+e2}
    // (pt1..pt8 of make_kp_reg_tet), tetrahedra (1,2,3,6) (2,3,4,6) (1,3,5,6) (3,4,6,8)
    // (3,5,6,7) (3,6,7,8)
    float cmn[kFgrPer], cmx[kFgrPer];
#pragma unroll
    for (int j = 0; j < kFgrPer; j++) {
        const int k = tid + j * kFgrThreads;
        cmn[j] = 3.4e38f;
        cmx[j] = -3.4e38f;
        if (any && k < nk) {
            int c0, c1, c2;
            m.coords(k, c0, c1, c2);
            const int d0 = (c0 + 1 == m.n0) ? 0 : c0 + 1, d1 = (c1 + 1 == m.n1) ? 0 : c1 + 1,
                      d2 = (c2 + 1 == m.n2) ? 0 : c2 + 1;
            const float ce[8] = {en[m.index(c0, c1, c2)], en[m.index(d0, c1, c2)], en[m.index(c0, d1, c2)],
                                 en[m.index(d0, d1, c2)], en[m.index(c0, c1, d2)], en[m.index(d0, c1, d2)],
                                 en[m.index(c0, d1, d2)], en[m.index(d0, d1, d2)]};
            for (int p = 0; p < 8; p++) {
                cmn[j] = fminf(cmn[j], ce[p]);
                cmx[j] = fmaxf(cmx[j], ce[p]);
            }
        }
    }
    const double tpi2 = (2 * alt3::PI) * (2 * alt3::PI);
    VT *P = proc == 0 ? P1 : P2;
    // the branches l1 whose omega(q1, l1) falls in each cube's energy range, once per item
    // (bit l1 of cm[j] for cube k = tid + j*T), so that each l1 pass visits only its cubes
    MT cm[kFgrPer];
#pragma unroll
    for (int j = 0; j < kFgrPer; j++) {
        cm[j] = 0;
        if (any && tid + j * kFgrThreads < nk)
            for (int l = 0; l < nb; l++)
                if (omega_in(s_om[l], cmn[j], cmx[j])) cm[j] |= (MT)1 << l;
    }
    // tetrahedra of cube k whose energy range contains e go to the LDS list of their
    // weight_tet case (e below the second-lowest corner, below the third, or above), so that
    // the lanes of a wave take the same branch of the weight formula.  Called by every lane of
    // the group (cube_in false where a lane has no cube); the list slots are claimed once per
    // wave and case (one LDS atomic per wave instead of one per tetrahedron).
    auto push = [&](int k, bool cube_in, double e) {
        float ce[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        if (cube_in) {
            int c0, c1, c2;
            m.coords(k, c0, c1, c2);
            const int d0 = (c0 + 1 == m.n0) ? 0 : c0 + 1, d1 = (c1 + 1 == m.n1) ? 0 : c1 + 1,
                      d2 = (c2 + 1 == m.n2) ? 0 : c2 + 1;
            ce[0] = en[m.index(c0, c1, c2)];
            ce[1] = en[m.index(d0, c1, c2)];
            ce[2] = en[m.index(c0, d1, c2)];
            ce[3] = en[m.index(d0, d1, c2)];
            ce[4] = en[m.index(c0, c1, d2)];
            ce[5] = en[m.index(d0, c1, d2)];
            ce[6] = en[m.index(c0, d1, d2)];
            ce[7] = en[m.index(d0, d1, d2)];
        }
#pragma unroll
        for (int l = 0; l < 6; l++) {
            const float a = ce[tet_corner(l, 0)], b = ce[tet_corner(l, 1)], c = ce[tet_corner(l, 2)],
                        d = ce[tet_corner(l, 3)];
            const float s1 = fminf(a, b), s2 = fmaxf(a, b), s3 = fminf(c, d), s4 = fmaxf(c, d);
            const bool act = cube_in && omega_in(e, fminf(s1, s3), fmaxf(s2, s4));
            const float m1 = fmaxf(s1, s3), m2 = fminf(s2, s4);
            const int cs = (e <= (double)fminf(m1, m2)) ? 0 : (e <= (double)fmaxf(m1, m2)) ? 1 : 2;
#pragma unroll
            for (int q = 0; q < 3; q++) {
                const bool mine = act && cs == q;
                const unsigned long long mask = __ballot(mine);
                if (mask) {   // the same for every lane of the wave
                    const int leader = __ffsll((long long)mask) - 1;
                    int base = 0;
                    if ((int)__lane_id() == leader) base = atomicAdd(&s_cnt[q], __popcll(mask));
                    base = __shfl(base, leader, 64);
                    const int x = base + __popcll(mask & ((1ull << __lane_id()) - 1ull));
                    if (mine && x < kFgrCase) s_list[q * kFgrCase + x] = (unsigned short)((k << 3) | l);
                }
            }
        }
    };
    // weights of the listed tetrahedra, added to W at their corners
    auto process = [&](int n0, int n1, int n2, double e) {
        for (int x = tid; x < n0 + n1 + n2; x += kFgrThreads) {
            const int at = x < n0 ? x : x < n0 + n1 ? kFgrCase + x - n0 : 2 * kFgrCase + x - n0 - n1;
            const int code = (int)s_list[at], kk = code >> 3, l = code & 7;
            int c0, c1, c2;
            m.coords(kk, c0, c1, c2);
            int v[4];
#pragma unroll
            for (int y = 0; y < 4; y++) {
                const int p = tet_corner(l, y);
                int e0 = c0 + (p & 1), e1 = c1 + ((p >> 1) & 1), e2 = c2 + ((p >> 2) & 1);
                if (e0 == m.n0) e0 = 0;
                if (e1 == m.n1) e1 = 0;
                if (e2 == m.n2) e2 = 0;
                v[y] = m.index(e0, e1, e2);
            }
            const float w4[4] = {en[v[0]], en[v[1]], en[v[2]], en[v[3]]};
            if constexpr (sizeof(VT) == sizeof(float)) {   // W in units of 1/6 (the 1/6 is in F)
                float c[4];
                weight_tet1f(w4, e, c);
                // a zero weight adds +-0, which leaves W as it is: no test before the atomics
                for (int y = 0; y < 4; y++) atomicAdd(&W[v[y]], (double)c[y]);
            } else {
                double c[4];
                weight_tet1(w4, e, c);
                for (int y = 0; y < 4; y++)
                    if (c[y] != 0.0) atomicAdd(&W[v[y]], c[y] / 6.0);
            }
        }
    };
    // P = W F with F = |V3|^2 (2 pi)^2 / nk
    const double fscale = tpi2 / nk;
    for (int l1 = 0; l1 < nb; l1++) {
        const double omega1 = s_om[l1];
        const size_t base = ((((size_t)s * nb + l1) * nb + l2) * nb + l3) * (size_t)nk;
        const size_t basem = ((((size_t)s * nb + l1) * nb + l3) * nb + l2) * (size_t)nk;   // mirrored row
        unsigned char *rmap = prow + (proc ? prow2 : 0);
        if (!omega_in(omega1, lo, hi)) {
            // no tetrahedron of this item reaches omega1: W = 0, so P = 0 without reading
            // |V3|^2 (uniform over the group).  FP32 P: the row is marked zero in the row map
            // and not written; FP64 P: written as zeros (its readers take every row)
            if constexpr (kSym) {
                if (tid == 0) {
                    rmap[base / nk] = 0;
                    if (mirror) rmap[basem / nk] = 0;
                }
            } else {
                if (tid == 0) rmap[base / nk] = 1;
#pragma unroll
                for (int j = 0; j < kFgrPer; j++) {
                    const int w = tid + j * kFgrThreads;
                    if (w < nk) P[base + w] = (VT)0;
                }
            }
            continue;
        }
        if (tid == 0) {
            rmap[base / nk] = 1;
            if (mirror) rmap[basem / nk] = 1;
        }
        // |V3|^2 of this l1 (at q2 = q_w for coalescence, q2 = -q_w for decay), requested
        // before the weight phase so that the load latency overlaps it
        VT fv[kFgrPer], fvm[kFgrPer];
#pragma unroll
        for (int j = 0; j < kFgrPer; j++) {
            const int w = tid + j * kFgrThreads;
            fv[j] = fvm[j] = 0;
            if (w < nk) {
                int vo = w;
                if (proc) {
                    int c0, c1, c2;
                    m.coords(w, c0, c1, c2);
                    vo = m.neg(c0, c1, c2);
                }
                fv[j] = v33[base + vo];
                if (mirror) fvm[j] = v33[basem + vo];
            }
        }
        {
            // Only tetrahedra whose energy range contains omega1 have weights, and they are
            // few: they are compacted into LDS lists so that the weight formula runs with
            // every lane busy.  The lists normally hold all of them; if one overflows, the
            // cubes are taken again chunk by chunk (at most 6 entries per cube and thread).
            // the cubes of this l1, compacted (wave-aggregated appends)
#pragma unroll
            for (int j = 0; j < kFgrPer; j++) {
                const int k = tid + j * kFgrThreads;
                const bool act = k < nk && (cm[j] >> l1 & 1);
                const unsigned long long mask = __ballot(act);
                if (mask) {   // the same for every lane of the wave
                    const int leader = __ffsll((long long)mask) - 1;
                    int base = 0;
                    if ((int)__lane_id() == leader) base = atomicAdd(&s_ncube, __popcll(mask));
                    base = __shfl(base, leader, 64);
                    if (act) s_cube[base + __popcll(mask & ((1ull << __lane_id()) - 1ull))] = (unsigned short)k;
                }
            }
            __syncthreads();
            const int ncube = s_ncube;
            // the waves whose lanes hold no cube skip the pushes (their ballots would be empty)
#pragma unroll 1
            for (int x0 = 0; x0 < ncube; x0 += kFgrThreads)
                if (x0 + (tid & ~63) < ncube) push(x0 + tid < ncube ? s_cube[x0 + tid] : 0, x0 + tid < ncube, omega1);
            __syncthreads();
            const int n0 = s_cnt[0], n1 = s_cnt[1], n2 = s_cnt[2];
            if (n0 <= kFgrCase && n1 <= kFgrCase && n2 <= kFgrCase) {
                process(n0, n1, n2, omega1);
            } else {   // a list overflowed: the cubes again, T at a time (at most 6T entries a case)
#pragma unroll 1
                for (int x0 = 0; x0 < ncube; x0 += kFgrThreads) {
                    __syncthreads();
                    if (tid < 3) s_cnt[tid] = 0;
                    __syncthreads();
                    if (x0 + (tid & ~63) < ncube) push(x0 + tid < ncube ? s_cube[x0 + tid] : 0, x0 + tid < ncube, omega1);
                    __syncthreads();
                    process(s_cnt[0], s_cnt[1], s_cnt[2], omega1);
                }
            }
            __syncthreads();
        }
        // P(q_w) = W(q_w) F(q_w) (and the mirrored row W(q1-q_w) F'(q_w)); the weight sums are
        // cleared once every thread has read them (at once where no thread reads another's)
        const float fs6 = (float)(fscale / 6.0);
#pragma unroll
        for (int j = 0; j < kFgrPer; j++) {
            const int w = tid + j * kFgrThreads;
            if (w < nk) {
                if constexpr (sizeof(VT) == sizeof(float)) {   // FP32 P: the product in FP32
                    const float F = (vzero >> j & 1u) ? 0.0f : fv[j] * fs6;
                    P[base + w] = (float)W[w] * F;
                } else {
                    double F = (double)fv[j] * tpi2 / nk;
                    if (vzero >> j & 1u) F = 0.0;
                    P[base + w] = (VT)(W[w] * F);
                }
                if (mirror) {
                    int c0, c1, c2;
                    m.coords(w, c0, c1, c2);
                    const int p = m.diff(a0, a1, a2, c0, c1, c2);
                    const float Fm = (vzerom >> j & 1u) ? 0.0f : (float)fvm[j] * fs6;
                    P[basem + w] = (VT)((float)W[p] * Fm);
                } else {
                    W[w] = 0.0;
                }
            }
        }
        if (mirror) {   // uniform over the group
            __syncthreads();
#pragma unroll
            for (int j = 0; j < kFgrPer; j++) {
                const int w = tid + j * kFgrThreads;
                if (w < nk) W[w] = 0.0;
            }
        }
        if (tid < 3) s_cnt[tid] = 0;
        if (tid == 0) s_ncube = 0;
        __syncthreads();
    }
}

// Meshes whose corner energies and weight sums do not fit in LDS (nk > 8192, e.g. 30^3): the
// same work items and arithmetic as k_fgr, with the per-item arrays -- corner energies,
// cube energy ranges and weight sums -- in a slice of global scratch owned by the
// workgroup.  The grid is persistent (one workgroup per scratch slice), each workgroup
// taking items blockIdx.x, blockIdx.x + gridDim.x, ...; the weight sums are global FP64
// atomics (L2), the tetrahedron lists stay in LDS with 32-bit codes and are filled chunk by
// chunk when one pass over the mesh overflows them.
constexpr int kBigCase = 4096;
template <typename VT>
__global__ void __launch_bounds__(kFgrThreads) k_fgr_big(alt3::MeshDev m, int nk, int nb, int ibz_lo, int nitems,
                                                         const int *__restrict__ mapinv,
                                                         const double *__restrict__ freq,
                                                         const double *__restrict__ freqT,
                                                         const VT *__restrict__ v33, VT *__restrict__ P1,
                                                         VT *__restrict__ P2, float *__restrict__ scratch)
{
    __shared__ float s_lo[kFgrThreads / 64], s_hi[kFgrThreads / 64];
    __shared__ int s_cnt[3];
    __shared__ unsigned s_list[3 * kBigCase];
    float *en = scratch + (size_t)blockIdx.x * 5 * nk;            // [nk]   (5 nk floats per slice)
    float2 *cr = reinterpret_cast<float2 *>(en + nk);              // [nk] cube (min, max)
    double *W = reinterpret_cast<double *>(en + 3 * nk);           // [nk]  (nk even: 8-byte aligned)
    constexpr bool kSym = sizeof(VT) == sizeof(float);
    const int tid = threadIdx.x, gam = m.gamma();
    const double tpi2 = (2 * alt3::PI) * (2 * alt3::PI);
    for (int item = blockIdx.x; item < nitems; item += gridDim.x) {
        const int proc = item & 1;
        int r = item >> 1;
        const int l3 = r % nb;
        r /= nb;
        const int l2 = r % nb;
        const int s = r / nb;
        if (kSym && proc == 1 && l2 > l3) continue;   // the same for the whole group
        const bool mirror = kSym && proc == 1 && l2 < l3;
        const int k1 = mapinv[ibz_lo + s];
        int a0, a1, a2;
        m.coords(k1, a0, a1, a2);
        float lo = 3.4e38f, hi = -3.4e38f;
        for (int w = tid; w < nk; w += kFgrThreads) {
            int c0, c1, c2;
            m.coords(w, c0, c1, c2);
            const int p = proc == 0 ? m.sum(a0, a1, a2, c0, c1, c2) : m.diff(a0, a1, a2, c0, c1, c2);
            const double e = proc == 0 ? -freqT[(size_t)l2 * nk + w] + freqT[(size_t)l3 * nk + p]
                                       : freqT[(size_t)l2 * nk + w] + freqT[(size_t)l3 * nk + p];
            const float ef = (float)e;
            en[w] = ef;
            W[w] = 0.0;
            lo = fminf(lo, ef);
            hi = fmaxf(hi, ef);
        }
        for (int off = 32; off > 0; off >>= 1) {
            lo = fminf(lo, __shfl_xor(lo, off, 64));
            hi = fmaxf(hi, __shfl_xor(hi, off, 64));
        }
        if ((tid & 63) == 0) {
            s_lo[tid >> 6] = lo;
            s_hi[tid >> 6] = hi;
        }
        __syncthreads();   // en written by every thread; the range partials
        for (int w = 0; w < kFgrThreads / 64; w++) {
            lo = fminf(lo, s_lo[w]);
            hi = fmaxf(hi, s_hi[w]);
        }
        for (int k = tid; k < nk; k += kFgrThreads) {
            int c0, c1, c2;
            m.coords(k, c0, c1, c2);
            const int d0 = (c0 + 1 == m.n0) ? 0 : c0 + 1, d1 = (c1 + 1 == m.n1) ? 0 : c1 + 1,
                      d2 = (c2 + 1 == m.n2) ? 0 : c2 + 1;
            const float ce[8] = {en[m.index(c0, c1, c2)], en[m.index(d0, c1, c2)], en[m.index(c0, d1, c2)],
                                 en[m.index(d0, d1, c2)], en[m.index(c0, c1, d2)], en[m.index(d0, c1, d2)],
                                 en[m.index(c0, d1, d2)], en[m.index(d0, d1, d2)]};
            float mn = ce[0], mx = ce[0];
#pragma unroll
            for (int q = 1; q < 8; q++) {
                mn = fminf(mn, ce[q]);
                mx = fmaxf(mx, ce[q]);
            }
            cr[k] = make_float2(mn, mx);
        }
        __syncthreads();   // cube ranges complete
        VT *P = proc == 0 ? P1 : P2;
        for (int l1 = 0; l1 < nb; l1++) {
            const double omega1 = freq[(size_t)k1 * nb + l1];
            const size_t base = ((((size_t)s * nb + l1) * nb + l2) * nb + l3) * (size_t)nk;
            const size_t basem = ((((size_t)s * nb + l1) * nb + l3) * nb + l2) * (size_t)nk;
            if (omega_in(omega1, lo, hi)) {
                // chunks of cubes whose tetrahedra always fit the lists (6 per cube)
                constexpr int kChunk = (kBigCase / 6) / kFgrThreads * kFgrThreads;
                for (int k0 = 0; k0 < nk; k0 += kChunk) {
                    if (tid < 3) s_cnt[tid] = 0;
                    __syncthreads();
                    for (int k = k0 + tid; k < min(nk, k0 + kChunk); k += kFgrThreads) {
                        const float2 rg = cr[k];
                        if (!omega_in(omega1, rg.x, rg.y)) continue;
                        int c0, c1, c2;
                        m.coords(k, c0, c1, c2);
                        const int d0 = (c0 + 1 == m.n0) ? 0 : c0 + 1, d1 = (c1 + 1 == m.n1) ? 0 : c1 + 1,
                                  d2 = (c2 + 1 == m.n2) ? 0 : c2 + 1;
                        const float ce[8] = {en[m.index(c0, c1, c2)], en[m.index(d0, c1, c2)], en[m.index(c0, d1, c2)],
                                             en[m.index(d0, d1, c2)], en[m.index(c0, c1, d2)], en[m.index(d0, c1, d2)],
                                             en[m.index(c0, d1, d2)], en[m.index(d0, d1, d2)]};
#pragma unroll
                        for (int l = 0; l < 6; l++) {
                            const float a = ce[tet_corner(l, 0)], b = ce[tet_corner(l, 1)], c = ce[tet_corner(l, 2)],
                                        d = ce[tet_corner(l, 3)];
                            const float s1 = fminf(a, b), s2 = fmaxf(a, b), s3 = fminf(c, d), s4 = fmaxf(c, d);
                            if (omega_in(omega1, fminf(s1, s3), fmaxf(s2, s4))) {
                                const float m1 = fmaxf(s1, s3), m2 = fminf(s2, s4);
                                const int cs = (omega1 <= (double)fminf(m1, m2)) ? 0
                                               : (omega1 <= (double)fmaxf(m1, m2)) ? 1 : 2;
                                const int x = atomicAdd(&s_cnt[cs], 1);
                                s_list[cs * kBigCase + x] = ((unsigned)k << 3) | (unsigned)l;
                            }
                        }
                    }
                    __syncthreads();
                    const int n0 = s_cnt[0], n1 = s_cnt[1], n2 = s_cnt[2];
                    for (int x = tid; x < n0 + n1 + n2; x += kFgrThreads) {
                        const int at = x < n0 ? x : x < n0 + n1 ? kBigCase + x - n0 : 2 * kBigCase + x - n0 - n1;
                        const unsigned code = s_list[at];
                        const int kk = (int)(code >> 3), l = (int)(code & 7u);
                        int c0, c1, c2;
                        m.coords(kk, c0, c1, c2);
                        int v[4];
#pragma unroll
                        for (int y = 0; y < 4; y++) {
                            const int pc = tet_corner(l, y);
                            int e0 = c0 + (pc & 1), e1 = c1 + ((pc >> 1) & 1), e2 = c2 + ((pc >> 2) & 1);
                            if (e0 == m.n0) e0 = 0;
                            if (e1 == m.n1) e1 = 0;
                            if (e2 == m.n2) e2 = 0;
                            v[y] = m.index(e0, e1, e2);
                        }
                        const float w4[4] = {en[v[0]], en[v[1]], en[v[2]], en[v[3]]};
                        if constexpr (sizeof(VT) == sizeof(float)) {
                            float c[4];
                            weight_tet1f(w4, omega1, c);
                            for (int y = 0; y < 4; y++)
                                if (c[y] != 0.0f) atomicAdd(&W[v[y]], (double)(c[y] / 6.0f));
                        } else {
                            double c[4];
                            weight_tet1(w4, omega1, c);
                            for (int y = 0; y < 4; y++)
                                if (c[y] != 0.0) atomicAdd(&W[v[y]], c[y] / 6.0);
                        }
                    }
                    __syncthreads();   // the lists are reused by the next chunk
                }
            }
            // P(q_w) = W(q_w) F(q_w) (and the mirrored row W(q1-q_w) F'(q_w)), W cleared after
            // every thread has read it
            for (int w = tid; w < nk; w += kFgrThreads) {
                int c0, c1, c2;
                m.coords(w, c0, c1, c2);
                const int vo = proc ? m.neg(c0, c1, c2) : w;
                const int pq = proc == 0 ? m.sum(a0, a1, a2, c0, c1, c2) : m.diff(a0, a1, a2, c0, c1, c2);
                const bool vz = (l2 < 3 && w == gam) || (l3 < 3 && pq == gam);
                double F = (double)v33[base + vo] * tpi2 / nk;
                if (vz) F = 0.0;
                P[base + w] = (VT)(W[w] * F);
                if (mirror) {
                    const int pm = m.diff(a0, a1, a2, c0, c1, c2);
                    const bool vzm = (l3 < 3 && w == gam) || (l2 < 3 && pm == gam);
                    double Fm = (double)v33[basem + vo] * tpi2 / nk;
                    if (vzm) Fm = 0.0;
                    P[basem + w] = (VT)(W[pm] * Fm);
                }
            }
            __syncthreads();
            for (int w = tid; w < nk; w += kFgrThreads) W[w] = 0.0;
            __syncthreads();
        }
    }
}

// Gaussian-smeared variant (delta_g, extratools.f90:239-246): the same corner energies and
// masks as calculate_FGR_mpi2, but delta(w1 - e) -> exp(-x^2/2s^2)/(sqrt(2pi) s) at the mesh
// point itself:  P_sm(q1,q2) = F(q2) delta_g(w(q1,l1) - e(q2), sigma).  FP64 throughout.
// One thread per (s, l1, l2, l3, q2), both processes.
template <typename VT>
__global__ void __launch_bounds__(256) k_fgr_gauss(alt3::MeshDev m, int nk, int nb, int ibz_lo, double sigma,
                                                   const int *__restrict__ mapinv, const double *__restrict__ freq,
                                                   const VT *__restrict__ v33, VT *__restrict__ P1,
                                                   VT *__restrict__ P2)
{
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= nk) return;
    int r = blockIdx.y;                  // (l1, l2, l3)
    const int l3 = r % nb;
    r /= nb;
    const int l2 = r % nb, l1 = r / nb;
    const int s = blockIdx.z;
    const int k1 = mapinv[ibz_lo + s];
    int a0, a1, a2, c0, c1, c2;
    m.coords(k1, a0, a1, a2);
    m.coords(w, c0, c1, c2);
    const int gam = m.gamma();
    const double omega1 = freq[(size_t)k1 * nb + l1];
    const double tpi2 = (2 * alt3::PI) * (2 * alt3::PI);
    const double norm = 1.0 / (sqrt(2 * alt3::PI) * sigma);
    const size_t base = ((((size_t)s * nb + l1) * nb + l2) * nb + l3) * (size_t)nk;
    const double f2 = freq[(size_t)w * nb + l2];
    {   // coalescence: -w(q2,l2) + w(q1+q2,l3), F at q2
        const int p = m.sum(a0, a1, a2, c0, c1, c2);
        const double x = omega1 - (-f2 + freq[(size_t)p * nb + l3]);
        double F = (double)v33[base + w] * tpi2 / nk;
        if ((l2 < 3 && w == gam) || (l3 < 3 && p == gam)) F = 0.0;
        P1[base + w] = (VT)(F * (exp(-x * x / 2 / sigma / sigma) * norm));
    }
    {   // decay: w(q2,l2) + w(q1-q2,l3), F at -q2
        const int p = m.diff(a0, a1, a2, c0, c1, c2);
        const double x = omega1 - (f2 + freq[(size_t)p * nb + l3]);
        double F = (double)v33[base + m.neg(c0, c1, c2)] * tpi2 / nk;
        if ((l2 < 3 && w == gam) || (l3 < 3 && p == gam)) F = 0.0;
        P2[base + w] = (VT)(F * (exp(-x * x / 2 / sigma / sigma) * norm));
    }
}

}  // namespace

namespace alt3 {

// P in the precision of |V3|^2: FP32 for the mixed path, FP64 otherwise
__global__ void k_fill_u8(size_t n, unsigned char v, unsigned char *__restrict__ a)
{
    const size_t x = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x < n) a[x] = v;
}

static int alloc_p(alt3_ctx *ctx, size_t nel)
{
    ALT3_TRY(ctx->d_prow.alloc(2 * nel / ctx->nk));
    ctx->p_fp32 = ctx->v3_prec == ALT3_V3_MIXED;
    if (ctx->p_fp32) {
        ctx->d_p1.release();
        ctx->d_p2.release();
        ALT3_TRY(ctx->d_p1f.alloc(nel));
        ALT3_TRY(ctx->d_p2f.alloc(nel));
    } else {
        ctx->d_p1f.release();
        ctx->d_p2f.release();
        ALT3_TRY(ctx->d_p1.alloc(nel));
        ALT3_TRY(ctx->d_p2.alloc(nel));
    }
    return ALT3_OK;
}

// every P row stored (writers that fill whole arrays: the Gaussian FGR, the global-scratch
// tetrahedron FGR, P loaded from the caller)
int prow_all(alt3_ctx *ctx)
{
    const size_t n = ctx->d_prow.n;
    if (n == 0) return ALT3_OK;
    hipLaunchKernelGGL(k_fill_u8, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, n, (unsigned char)1,
                       ctx->d_prow.p);
    ALT3_HIP(hipGetLastError());
    return ALT3_OK;
}

int fgr_gauss_run(alt3_ctx *ctx, double sigma)
{
    if (!ctx->have_v3) return fail(ALT3_E_STATE, "alt3_fgr_gauss needs alt3_v3sq_ibz first");
    if (!(sigma > 0)) return fail(ALT3_E_ARG, "Gaussian width must be positive");
    const int nk = ctx->nk, nb = ctx->nb, ksub = ctx->ibz_hi - ctx->ibz_lo + 1;
    const size_t nel = (size_t)ksub * nb * nb * nb * nk;
    ALT3_TRY(alloc_p(ctx, nel));
    ALT3_TRY(prow_all(ctx));
    const MeshDev md = MeshDev::make(ctx->nc);
    dim3 grid((unsigned)((nk + 255) / 256), (unsigned)(nb * nb * nb), (unsigned)ksub);
    if (ctx->p_fp32)
        hipLaunchKernelGGL(k_fgr_gauss<float>, grid, dim3(256), 0, ctx->stream, md, nk, nb, ctx->ibz_lo, sigma,
                           ctx->d_mapinv.p, ctx->d_freq.p, ctx->d_v33f.p, ctx->d_p1f.p, ctx->d_p2f.p);
    else
        hipLaunchKernelGGL(k_fgr_gauss<double>, grid, dim3(256), 0, ctx->stream, md, nk, nb, ctx->ibz_lo, sigma,
                           ctx->d_mapinv.p, ctx->d_freq.p, ctx->d_v33.p, ctx->d_p1.p, ctx->d_p2.p);
    ALT3_HIP(hipGetLastError());
    ALT3_HIP(hipStreamSynchronize(ctx->stream));
    ctx->have_fgr = true;
    ctx->fgr_kind = ALT3_DELTA_GAUSS;
    ctx->fgr_sigma = sigma;
    ctx->have_tau = ctx->have_cm = false;
    return ALT3_OK;
}

int fgr_run(alt3_ctx *ctx, bool sync)
{
    if (!ctx->have_v3) return fail(ALT3_E_STATE, "alt3_fgr needs alt3_v3sq_ibz first");
    const int nk = ctx->nk, nb = ctx->nb, ksub = ctx->ibz_hi - ctx->ibz_lo + 1;
    const size_t nel = (size_t)ksub * nb * nb * nb * nk;
    const bool mixed = ctx->v3_prec == ALT3_V3_MIXED;
    // W, corner energies, active-cube list
    size_t lds = (size_t)nk * (sizeof(double) + sizeof(float) + sizeof(unsigned short));
    ALT3_TRY(alloc_p(ctx, nel));
    const MeshDev md = MeshDev::make(ctx->nc);
    ALT3_TRY(ctx->d_freqT.reserve((size_t)nk * nb));
    hipLaunchKernelGGL(k_freq_t, dim3((unsigned)((nk * nb + 255) / 256)), dim3(256), 0, ctx->stream, nk, nb, ctx->d_freq.p,
                       ctx->d_freqT.p);
    // the LDS kernel's list codes (cube << 3 | tetrahedron) are 16 bits wide: nk <= 8192 (20^3);
    // larger meshes take the global-scratch kernel
    if (lds + kFgrList * sizeof(unsigned short) + 1024 > ctx->lds_max || nk > 16 * kFgrThreads || nb > kFgrMaxNb) {
        if ((unsigned long long)nk * 8 > (1ull << 31)) return fail(ALT3_E_ARG, "mesh too large for the FGR codes");
        const long long nitems = 2ll * ksub * nb * nb;
        const void *kern = mixed ? (const void *)k_fgr_big<float> : (const void *)k_fgr_big<double>;
        int per_cu = 0, cus = 0;
        ALT3_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, kFgrThreads, 0));
        ALT3_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device));
        const long long grid = std::max(1ll, std::min(nitems, (long long)std::max(1, per_cu) * std::max(1, cus)));
        ALT3_TRY(ctx->d_fgr_big.reserve((size_t)grid * 5 * nk));
        ALT3_TRY(prow_all(ctx));   // this kernel writes every row
        kernel_mark(ctx, 1, false);
        if (mixed)
            hipLaunchKernelGGL(k_fgr_big<float>, dim3((unsigned)grid), dim3(kFgrThreads), 0, ctx->stream, md, nk, nb,
                               ctx->ibz_lo, (int)nitems, ctx->d_mapinv.p, ctx->d_freq.p, ctx->d_freqT.p,
                               ctx->d_v33f.p, ctx->d_p1f.p, ctx->d_p2f.p, ctx->d_fgr_big.p);
        else
            hipLaunchKernelGGL(k_fgr_big<double>, dim3((unsigned)grid), dim3(kFgrThreads), 0, ctx->stream, md, nk, nb,
                               ctx->ibz_lo, (int)nitems, ctx->d_mapinv.p, ctx->d_freq.p, ctx->d_freqT.p,
                               ctx->d_v33.p, ctx->d_p1.p, ctx->d_p2.p, ctx->d_fgr_big.p);
        ALT3_HIP(hipGetLastError());
        kernel_mark(ctx, 1, true);
        if (sync) ALT3_HIP(hipStreamSynchronize(ctx->stream));
        ctx->have_fgr = true;
        ctx->fgr_kind = ALT3_DELTA_TETRA;
        ctx->have_tau = ctx->have_cm = false;
        return ALT3_OK;
    }
    dim3 grid((unsigned)(2 * (size_t)ksub * nb * nb));
    const int per = (nk + kFgrThreads - 1) / kFgrThreads;
    const size_t prow2 = (size_t)ksub * nb * nb * nb;   // the P2 rows of the row map
    const void *v = mixed ? (const void *)ctx->d_v33f.p : (const void *)ctx->d_v33.p;
    const bool big = lds + kFgrList * sizeof(unsigned short) + 1024 > 65536;   // static + dynamic LDS
#define ALT3_FGR_LAUNCH_MT(PER, MT)                                                                              \
    if (big) {                                                                                                   \
        ALT3_HIP(hipFuncSetAttribute((const void *)k_fgr<float, PER, MT>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds)); \
        ALT3_HIP(hipFuncSetAttribute((const void *)k_fgr<double, PER, MT>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds)); \
    }                                                                                                            \
    if (mixed)                                                                                                   \
        hipLaunchKernelGGL((k_fgr<float, PER, MT>), grid, dim3(kFgrThreads), lds, ctx->stream, md, nk, nb, ctx->ibz_lo, \
                           ctx->d_mapinv.p, ctx->d_freq.p, ctx->d_freqT.p, (const float *)v, ctx->d_p1f.p,         \
                           ctx->d_p2f.p, ctx->d_prow.p, prow2);                                                  \
    else                                                                                                         \
        hipLaunchKernelGGL((k_fgr<double, PER, MT>), grid, dim3(kFgrThreads), lds, ctx->stream, md, nk, nb,        \
                           ctx->ibz_lo, ctx->d_mapinv.p, ctx->d_freq.p, ctx->d_freqT.p, (const double *)v, ctx->d_p1.p, \
                           ctx->d_p2.p, ctx->d_prow.p, prow2);
#define ALT3_FGR_LAUNCH(PER)                                                                                     \
    if (nb <= 32) {                                                                                              \
        ALT3_FGR_LAUNCH_MT(PER, unsigned)                                                                        \
    } else {                                                                                                     \
        ALT3_FGR_LAUNCH_MT(PER, unsigned long long)                                                              \
    }
    kernel_mark(ctx, 1, false);
    if (per <= 2) {
        ALT3_FGR_LAUNCH(2)
    } else if (per <= 4) {
        ALT3_FGR_LAUNCH(4)
    } else if (per <= 6) {
        ALT3_FGR_LAUNCH(6)
    } else if (per <= 8) {
        ALT3_FGR_LAUNCH(8)
    } else if (per <= 9) {
        ALT3_FGR_LAUNCH(9)
    } else if (per <= 12) {
        ALT3_FGR_LAUNCH(12)
    } else {
        ALT3_FGR_LAUNCH(16)
    }
#undef ALT3_FGR_LAUNCH
#undef ALT3_FGR_LAUNCH_MT
    ALT3_HIP(hipGetLastError());
    kernel_mark(ctx, 1, true);
    if (sync) ALT3_HIP(hipStreamSynchronize(ctx->stream));
    ctx->have_fgr = true;
    ctx->fgr_kind = ALT3_DELTA_TETRA;
    ctx->have_tau = ctx->have_cm = false;
    return ALT3_OK;
}

// Isotope term of P_iso_split (sub.f90:4825-4961).  One workgroup per (irreducible q1, l1):
//   Piso(q1,l1) = w1^2 pi/2 sum_l2 sum_q W(q,l2) gs(q,l2),
//   W(q,l2)     = sum over the tetrahedra at q of c/6,   c the weight_tet weights of
//                 delta(w1 - w(q,l2)) (corner energies REAL*4, as there),
//   gs(q,l2)    = sum_atoms g(a) |sum_{x of atom a} conj(e(x,l1,q1)) e(x,l2,q)|^2,
// which is the reference's sum over tetrahedra and corners of gs c/6, grouped by mesh point.
// Per branch l2 the mesh energies go to LDS; the threads scan a chunk of cubes and push the
// tetrahedra whose energy range contains w1 onto an LDS list, which all threads then work
// through (few tetrahedra are active, so the FP64 weight formula runs with every lane busy),
// adding c/6 to W in LDS; gs is then formed once per mesh point with a weight.  The reference
// repeats the whole sum inside a loop over n2 = 1..nk that changes nothing in it and divides by
// nk at the end; the sum is formed once.
constexpr int kIsoList = 2048;   // tetrahedra per case and chunk of cubes (kIsoList / 6 cubes)
constexpr int kIsoThreads = 512;

// the mesh energies as REAL*4, branch-major [l][q], for coalesced per-branch loads
__global__ void k_freq_f32t(int nk, int nb, const double *__restrict__ freq, float *__restrict__ enT)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x;   // (l, q)
    if (x >= nk * nb) return;
    const int l = x / nk, q = x - l * nk;
    enT[x] = (float)freq[(size_t)q * nb + l];
}

__global__ void __launch_bounds__(kIsoThreads) k_piso(MeshDev m, int nk, int nb, int natoms0, const int *__restrict__ mapinv,
                                              const double *__restrict__ freq, const float *__restrict__ enT,
                                              const double *__restrict__ evec,
                                              const double *__restrict__ giso, double *__restrict__ piso)
{
    extern __shared__ double Wl[];                          // [nk] weight sums of branch l2
    float *ens = reinterpret_cast<float *>(Wl + nk);        // [nk] energies of branch l2
    __shared__ double e1[2 * 64];
    __shared__ double ga[32];
    __shared__ double red[kIsoThreads / 64];
    __shared__ int s_cnt[3];
    __shared__ unsigned s_list[3 * kIsoList];   // [case][kIsoList] of cube << 3 | tetrahedron
    const int s = blockIdx.x / nb, l1 = blockIdx.x % nb, tid = threadIdx.x;
    const int k1 = mapinv[s];
    const double omega1 = freq[(size_t)k1 * nb + l1];
    for (int x = tid; x < 2 * nb; x += blockDim.x) e1[x] = evec[2 * ((size_t)k1 * nb + l1) * nb + x];
    for (int x = tid; x < natoms0; x += blockDim.x) ga[x] = giso[x];
    for (int w = tid; w < nk; w += blockDim.x) Wl[w] = 0.0;
    double acc = 0;
    constexpr int kChunk = kIsoList / 6;
    // tetrahedra of cubes [c0, c1) whose energy range contains w1 onto the list of their
    // weight_tet case (w1 below the second-lowest corner, below the third, or above), so that
    // the lanes of a wave take the same branch of the weight formula (entries past the end of
    // a list are counted but not stored)
    auto scan = [&](int c0, int c1) {
        for (int k = c0 + tid; k < c1; k += blockDim.x) {
            int a0, a1, a2;
            m.coords(k, a0, a1, a2);
            const int d0 = (a0 + 1 == m.n0) ? 0 : a0 + 1, d1 = (a1 + 1 == m.n1) ? 0 : a1 + 1,
                      d2 = (a2 + 1 == m.n2) ? 0 : a2 + 1;
            const float ce[8] = {ens[m.index(a0, a1, a2)], ens[m.index(d0, a1, a2)], ens[m.index(a0, d1, a2)],
                                 ens[m.index(d0, d1, a2)], ens[m.index(a0, a1, d2)], ens[m.index(d0, a1, d2)],
                                 ens[m.index(a0, d1, d2)], ens[m.index(d0, d1, d2)]};
            float lo = ce[0], hi = ce[0];
            for (int p = 1; p < 8; p++) {
                lo = fminf(lo, ce[p]);
                hi = fmaxf(hi, ce[p]);
            }
            if (!omega_in(omega1, lo, hi)) continue;
            for (int l = 0; l < 6; l++) {
                const float a = ce[tet_corner(l, 0)], b = ce[tet_corner(l, 1)], c = ce[tet_corner(l, 2)],
                            d = ce[tet_corner(l, 3)];
                const float s1 = fminf(a, b), s2 = fmaxf(a, b), s3 = fminf(c, d), s4 = fmaxf(c, d);
                if (omega_in(omega1, fminf(s1, s3), fmaxf(s2, s4))) {
                    const float m1 = fmaxf(s1, s3), m2 = fminf(s2, s4);
                    const int cs = (omega1 <= (double)fminf(m1, m2)) ? 0 : (omega1 <= (double)fmaxf(m1, m2)) ? 1 : 2;
                    const int x = atomicAdd(&s_cnt[cs], 1);
                    if (x < kIsoList) s_list[cs * kIsoList + x] = ((unsigned)k << 3) | (unsigned)l;
                }
            }
        }
    };
    // weight_tet of the listed tetrahedra (n0, n1, n2 of the three cases), c/6 added to W at
    // their corners
    auto process = [&](int n0, int n1, int n2) {
        for (int x = tid; x < n0 + n1 + n2; x += blockDim.x) {
            const int at = x < n0 ? x : x < n0 + n1 ? kIsoList + x - n0 : 2 * kIsoList + x - n0 - n1;
            const int k = (int)(s_list[at] >> 3), l = (int)(s_list[at] & 7u);
            int a0, a1, a2;
            m.coords(k, a0, a1, a2);
            int v[4];
            for (int y = 0; y < 4; y++) {
                const int p = tet_corner(l, y);
                int b0 = a0 + (p & 1), b1 = a1 + ((p >> 1) & 1), b2 = a2 + ((p >> 2) & 1);
                if (b0 == m.n0) b0 = 0;
                if (b1 == m.n1) b1 = 0;
                if (b2 == m.n2) b2 = 0;
                v[y] = m.index(b0, b1, b2);
            }
            const float w4[4] = {ens[v[0]], ens[v[1]], ens[v[2]], ens[v[3]]};
            double c[4];
            weight_tet1(w4, omega1, c);
            for (int y = 0; y < 4; y++)
                if (c[y] != 0.0) atomicAdd(&Wl[v[y]], c[y] / 6.0);
        }
    };
    for (int l2 = 0; l2 < nb; l2++) {
        __syncthreads();   // the previous branch's energies and weights are no longer read
        for (int w = tid; w < nk; w += blockDim.x) ens[w] = enT[(size_t)l2 * nk + w];
        // all cubes at once when their active tetrahedra fit the list (the usual case), else
        // chunk by chunk (kIsoList / 6 cubes cannot overflow it); the branch is uniform
        if (tid < 3) s_cnt[tid] = 0;
        __syncthreads();
        scan(0, nk);
        __syncthreads();
        const int n0 = s_cnt[0], n1 = s_cnt[1], n2 = s_cnt[2];
        if (n0 <= kIsoList && n1 <= kIsoList && n2 <= kIsoList) {
            process(n0, n1, n2);
        } else {
            for (int c0 = 0; c0 < nk; c0 += kChunk) {
                __syncthreads();
                if (tid < 3) s_cnt[tid] = 0;
                __syncthreads();
                scan(c0, min(nk, c0 + kChunk));
                __syncthreads();
                process(s_cnt[0], s_cnt[1], s_cnt[2]);
            }
        }
        __syncthreads();   // W is complete
        // gs at the mesh points with a weight; the weights are cleared for the next branch
        for (int w = tid; w < nk; w += blockDim.x) {
            const double W = Wl[w];
            if (W == 0.0) continue;
            Wl[w] = 0.0;
            const double *e2 = evec + 2 * ((size_t)w * nb + l2) * nb;
            double gs = 0;
            for (int at = 0; at < natoms0; at++) {
                double re = 0, im = 0;
                for (int x3 = 3 * at; x3 < 3 * at + 3; x3++) {
                    const double ar = e1[2 * x3], ai = e1[2 * x3 + 1], br = e2[2 * x3], bi = e2[2 * x3 + 1];
                    re = re + (ar * br + ai * bi);
                    im = im + (ar * bi - ai * br);
                }
                gs = gs + ga[at] * (re * re + im * im);
            }
            acc = acc + gs * W;
        }
    }
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
    if ((tid & 63) == 0) red[tid >> 6] = acc;
    __syncthreads();
    if (tid == 0) {
        double t = 0;
        for (int w = 0; w < (int)(blockDim.x >> 6); w++) t += red[w];
        piso[(size_t)s * nb + l1] = t * omega1 * omega1 * PI / 2;
    }
}

int piso_run(alt3_ctx *ctx)
{
    if (ctx->giso.empty() || ctx->piso_ok) return ALT3_OK;
    if (!ctx->have_eig) return fail(ALT3_E_STATE, "isotope scattering needs the eigen data (alt3_eig or alt3_set_eig)");
    if (ctx->nb > 64) return fail(ALT3_E_ARG, "isotope term: at most 64 branches");
    const int nb = ctx->nb, nibz = ctx->nibz;
    ALT3_TRY(ctx->d_giso.alloc(ctx->giso.size()));
    ALT3_TRY(ctx->d_piso.alloc((size_t)nibz * nb));
    ALT3_HIP(hipMemcpyAsync(ctx->d_giso.p, ctx->giso.data(), sizeof(double) * ctx->giso.size(), hipMemcpyHostToDevice,
                            ctx->stream));
    const size_t lds = (size_t)ctx->nk * (sizeof(double) + sizeof(float));
    if (lds + 3 * kIsoList * sizeof(unsigned) + 2048 > ctx->lds_max)
        return fail(ALT3_E_ARG, "isotope term: the mesh energies do not fit in LDS (nk too large)");
    if (lds + 3 * kIsoList * sizeof(unsigned) + 2048 > 65536)
        ALT3_HIP(hipFuncSetAttribute((const void *)k_piso, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    DBuf<double> &enT = ctx->d_red;   // reduction scratch, nk*nb floats here
    ALT3_TRY(enT.reserve(((size_t)ctx->nk * nb + 1) / 2));
    hipLaunchKernelGGL(k_freq_f32t, dim3((unsigned)((ctx->nk * nb + 255) / 256)), dim3(256), 0, ctx->stream, ctx->nk, nb,
                       ctx->d_freq.p, reinterpret_cast<float *>(enT.p));
    hipLaunchKernelGGL(k_piso, dim3((unsigned)(nibz * nb)), dim3(kIsoThreads), lds, ctx->stream, MeshDev::make(ctx->nc),
                       ctx->nk, nb, ctx->natoms0, ctx->d_mapinv.p, ctx->d_freq.p,
                       reinterpret_cast<const float *>(enT.p), ctx->d_evec.p, ctx->d_giso.p, ctx->d_piso.p);
    ALT3_HIP(hipGetLastError());
    ALT3_HIP(hipStreamSynchronize(ctx->stream));
    ctx->piso_ok = true;
    return ALT3_OK;
}

}  // namespace alt3
