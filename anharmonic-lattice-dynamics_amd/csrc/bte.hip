// Temperature-dependent part of the three-phonon path on the GPU (FP64):
//   distribution           sub.f90:1955-1980, nbe extratools.f90:257-288
//   calculate_tauRTA_mpi2  sub.f90:7953-8194 (Q, N/U split) + Qandtau :8198-8305
//   F_RTA + mode_kappa     sub.f90:1238-1289, :1292-1401
//   CollisionMatrix_mpi    sub.f90:3034-3210, cal_RHS_IBZ :2618-2691
//   cg_quadratic           cg.f90:19-69
//   mode_kappa_d           sub.f90:1407-1530 with veloc_avergeIBZ_invsyFBZ :5479-5537
//   update_mpi + job 3     sub.f90:2295-2449, kappa.f90:1373-1487 (iterative BTE)
//
// Layouts on the device: P1sm/P2sm [s][l1][l2][l3][q2]; the tau kernel also writes
// R1/R2[s][l1][l2][q2] = sum_l3 P1, P2 (occupations applied), which is all the collision
// matrix reads: C = -R1(n0) + R1(q2-q1) + R2(n0)/2 + R2(q1-q2)/2 off the diagonal and -Q on it.
#include <algorithm>

#include "alt3_impl.h"

#include <type_traits>
#include "mesh_dev.h"

namespace {

__device__ __forceinline__ double nbe(double w, double t, int m)
{
    double z = w / t;
    if (m == 0) {
        if (z > 60) return 0;
        if (z < (double)0.0010f) return 1 / (z * (1 + z / 2 * (1 + z / 3)));
        return 1 / (exp(z) - 1);
    }
    return 1 / z;
}

__global__ void k_dist(int n, const double *__restrict__ freq, double temp, int classical, double *__restrict__ dist,
                       int *__restrict__ err)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double w = freq[i];
    if (!(w > 0)) {
        atomicExch(err, 1);
        dist[i] = 0;
        return;
    }
    dist[i] = nbe(w, temp, classical);
}

__device__ double block_sum(double v, double *sh)
{
    // blockDim.x a multiple of 64: butterflies inside each wavefront, then the wavefront
    // partials through LDS, added by every thread in wavefront order
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    double r = 0;
    for (int i = 0; i < (int)(blockDim.x >> 6); i++) r += sh[i];
    __syncthreads();
    return r;
}

// (a, b) summed over the workgroup (blockDim.x a multiple of 64, at most 1024): butterflies
// inside each wavefront, then the wavefront partials through LDS (sh: 2 doubles per
// wavefront), added by every thread in the same order, so all threads hold the same totals
__device__ void block_sum2(double &a, double &b, double *sh)
{
    for (int off = 32; off > 0; off >>= 1) {
        a += __shfl_xor(a, off, 64);
        b += __shfl_xor(b, off, 64);
    }
    const int w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        sh[2 * w] = a;
        sh[2 * w + 1] = b;
    }
    __syncthreads();
    double x = 0, y = 0;
    for (int i = 0; i < nw; i++) {
        x += sh[2 * i];
        y += sh[2 * i + 1];
    }
    __syncthreads();
    a = x;
    b = y;
}

// the star table of the collision and update kernels (see star_sums)
struct StarTasks {
    const int *slot_n2;   // [nibz] irreducible point of slot x (slots by decreasing star size)
    const int4 *task;     // [ntask] (slot, first member, end, part): part -1 = the whole star
    const int *split_x;   // [nsplit] slot of split star p (parts 2p and 2p+1)
    const int *tab;       // [48][nibz] member-major q0 << 6 | op
    int ntask, nsplit;
};

StarTasks star_tasks(const alt3_ctx *ctx)
{
    return StarTasks{ctx->d_cm_slot_n2.p, ctx->d_cm_task.p, ctx->d_cm_split_x.p, ctx->d_cm_tab.p, ctx->cm_ntask,
                     ctx->cm_nsplit};
}

// one workgroup per (s, l1, l2): the l3-reduced R1/R2 rows and partial Q, QN, QU of NT
// temperatures at once, so that P1sm/P2sm are read once for all of them.  R is stored in the
// precision of P (FP32 on the mixed path); the sums are FP64.
template <int NT, typename PT>
__global__ void __launch_bounds__(256) k_tau(alt3::MeshDev m, int nk, int nb, int ibz_lo,
                                             const int *__restrict__ mapinv, const double *__restrict__ dist,
                                             size_t dstr, const PT *__restrict__ P1sm,
                                             const PT *__restrict__ P2sm, double *__restrict__ qpart, size_t qstr,
                                             PT *__restrict__ R1, PT *__restrict__ R2, size_t rstr,
                                             const unsigned char *__restrict__ rm1,
                                             const unsigned char *__restrict__ rm2)
{
    __shared__ double sh[256];
    const int l2 = blockIdx.x % nb, sl = blockIdx.x / nb, s = sl / nb, l1 = sl % nb;
    const int k1 = mapinv[ibz_lo + s];
    int a0, a1, a2;
    m.coords(k1, a0, a1, a2);
    double d1[NT], q[NT], qn[NT], qu[NT];
#pragma unroll
    for (int t = 0; t < NT; t++) {
        d1[t] = dist[t * dstr + (size_t)k1 * nb + l1];
        q[t] = qn[t] = qu[t] = 0;
    }
    const size_t row = ((size_t)sl * nb + l2) * (size_t)nb * nk;   // (s,l1,l2,l3=0,j=0)
    for (int j = threadIdx.x; j < nk; j += blockDim.x) {
        int c0, c1, c2;
        m.coords(j, c0, c1, c2);
        const bool in1 = m.sum_inside(a0, a1, a2, c0, c1, c2);    // q1+q2 inside the first zone
        const int np = m.sum(a0, a1, a2, c0, c1, c2);
        const bool in2 = m.diff_inside(a0, a1, a2, c0, c1, c2);   // q1-q2
        const int nm = m.diff(a0, a1, a2, c0, c1, c2);
        double f1[NT], f2[NT], r1[NT], r2[NT];
#pragma unroll
        for (int t = 0; t < NT; t++) {
            const double d2 = dist[t * dstr + (size_t)j * nb + l2];
            f1[t] = d1[t] * d2;
            f2[t] = d1[t] * (d2 + 1);
            r1[t] = r2[t] = 0;
        }
#pragma unroll 3
        for (int l3 = 0; l3 < nb; l3++) {
            const size_t x = row + (size_t)l3 * nk + j, rid = ((size_t)sl * nb + l2) * nb + l3;
            // rows the row map marks zero are not stored (rm NULL: all stored)
            const double P1 = (!rm1 || rm1[rid]) ? (double)P1sm[x] : 0.0;
            const double P2 = (!rm2 || rm2[rid]) ? (double)P2sm[x] : 0.0;
#pragma unroll
            for (int t = 0; t < NT; t++) {
                const double p1 = P1 * (f1[t] * (dist[t * dstr + (size_t)np * nb + l3] + 1));
                const double p2 = P2 * (f2[t] * (dist[t * dstr + (size_t)nm * nb + l3] + 1));
                q[t] = q[t] + (p1 + 0.5 * p2);
                if (in1) qn[t] += p1; else qu[t] += p1;
                if (in2) qn[t] += 0.5 * p2; else qu[t] += 0.5 * p2;
                r1[t] += p1;
                r2[t] += p2;
            }
        }
        const size_t o = ((size_t)sl * nb + l2) * (size_t)nk + j;
#pragma unroll
        for (int t = 0; t < NT; t++) {
            R1[t * rstr + o] = (PT)r1[t];
            R2[t * rstr + o] = (PT)r2[t];
        }
    }
#pragma unroll
    for (int t = 0; t < NT; t++) {
        const double a = block_sum(q[t], sh), b = block_sum(qn[t], sh), c = block_sum(qu[t], sh);
        if (threadIdx.x == 0) {
            qpart[t * qstr + 3 * (size_t)blockIdx.x] = a;
            qpart[t * qstr + 3 * (size_t)blockIdx.x + 1] = b;
            qpart[t * qstr + 3 * (size_t)blockIdx.x + 2] = c;
        }
    }
}

__global__ void k_fill_zero(int n, double *__restrict__ a)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x < n) a[x] = 0.0;
}

// the P row map packed per (s, l1, l2): bit l3 for the P1 row, bit 32 + l3 for the P2 row
// (cells of up to 32 branches)
__global__ void k_prow_pack(size_t n, int nb, size_t p2, const unsigned char *__restrict__ rm,
                            unsigned long long *__restrict__ out)
{
    const size_t x = (size_t)blockIdx.x * blockDim.x + threadIdx.x;   // (s, l1, l2)
    if (x >= n) return;
    unsigned long long w = 0;
    for (int l3 = 0; l3 < nb; l3++) {
        if (rm[x * nb + l3]) w |= 1ull << l3;
        if (rm[p2 + x * nb + l3]) w |= 1ull << (32 + l3);
    }
    out[x] = w;
}

// one workgroup per (s, chunk of 256 mesh points q2), every (l1, l2, l3) inside, for cells of
// NB branches: the occupations at q1+q2 and q1-q2 (the gathers of k_tau) are read once per
// thread into registers, so the inner loops only stream the P rows (12 independent loads per
// (l1, l2)) and write the R rows.  Partial Q sums per (s, l1, chunk).
template <int NT, typename PT, int NB>
__global__ void __launch_bounds__(256) k_tau_q2(alt3::MeshDev m, int nk, int ibz_lo, const int *__restrict__ mapinv,
                                                const double *__restrict__ dist, size_t dstr,
                                                const PT *__restrict__ P1sm, const PT *__restrict__ P2sm,
                                                double *__restrict__ qpart, size_t qstr, PT *__restrict__ R1,
                                                PT *__restrict__ R2, size_t rstr,
                                                const unsigned long long *__restrict__ rmb,
                                                const PT *__restrict__ pzero)
{
    __shared__ double sh[256];
    const int s = blockIdx.y, nch = gridDim.x, ch = blockIdx.x;
    const int j = ch * 256 + threadIdx.x;
    const bool valid = j < nk;
    const int jc = valid ? j : nk - 1;
    const int k1 = mapinv[ibz_lo + s];
    int a0, a1, a2, c0, c1, c2;
    m.coords(k1, a0, a1, a2);
    m.coords(jc, c0, c1, c2);
    const bool in1 = m.sum_inside(a0, a1, a2, c0, c1, c2), in2 = m.diff_inside(a0, a1, a2, c0, c1, c2);
    const int np = m.sum(a0, a1, a2, c0, c1, c2), nm = m.diff(a0, a1, a2, c0, c1, c2);
    // the occupations at q2 in registers too for small cells; for more branches they are
    // read per (l1, l2) (cached) instead of picked out of a register array by selects
    constexpr bool kRegD2 = NB <= 12;
    double n3p[NT][NB], n3m[NT][NB], d2[NT][kRegD2 ? NB : 1];
#pragma unroll
    for (int t = 0; t < NT; t++)
#pragma unroll
        for (int l = 0; l < NB; l++) {
            n3p[t][l] = dist[t * dstr + (size_t)np * NB + l] + 1;
            n3m[t][l] = dist[t * dstr + (size_t)nm * NB + l] + 1;
            if constexpr (kRegD2) d2[t][l] = dist[t * dstr + (size_t)jc * NB + l];
        }
    for (int l1 = 0; l1 < NB; l1++) {
        double d1[NT], q[NT], qn[NT], qu[NT];
#pragma unroll
        for (int t = 0; t < NT; t++) {
            d1[t] = dist[t * dstr + (size_t)k1 * NB + l1];
            q[t] = qn[t] = qu[t] = 0;
        }
        // the branches l2 with a stored P row (FP32 path with a row map; every l2 otherwise):
        // the others (44 % of the (q1, l1, l2) blocks on rock-salt 12^3) have R1 = R2 = 0 and add
        // nothing, so the loop below visits only these (the sums skip only additions of zero)
        unsigned live = (1u << NB) - 1;
        if constexpr (sizeof(PT) == sizeof(float)) {
            if (rmb) {
                live = 0;
                for (int l2 = 0; l2 < NB; l2++)
                    if (rmb[((size_t)s * NB + l1) * NB + l2]) live |= 1u << l2;
                if (valid) {
                    for (unsigned z = ~live & ((1u << NB) - 1); z; z &= z - 1) {
                        const size_t o = (((size_t)s * NB + l1) * NB + __builtin_ctz(z)) * (size_t)nk + j;
#pragma unroll
                        for (int t = 0; t < NT; t++) {
                            R1[t * rstr + o] = (PT)0;
                            R2[t * rstr + o] = (PT)0;
                        }
                    }
                }
            }
        }
#pragma unroll 1
        for (unsigned lv = live; lv; lv &= lv - 1) {
            const int l2 = __builtin_ctz(lv);
            const size_t rid = (((size_t)s * NB + l1) * NB + l2) * NB;                   // row map (s,l1,l2,l3=0)
            const size_t row = rid * (size_t)nk + jc;                                       // (s,l1,l2,l3=0)
            // rows the row map marks zero are not stored (rm NULL: all stored): bit l3 of b1, b2,
            // the same for the whole group; their loads go to one cached line of zeros (pzero),
            // so they cost no HBM traffic and no branch
            PT P1[NB], P2[NB];
            unsigned bb = (1u << NB) - 1;   // l3 with a stored P1 or P2 row
            if constexpr (sizeof(PT) == sizeof(float)) {   // the FP64 FGR stores every row
                unsigned b1 = (1u << NB) - 1, b2 = b1;
                if (rmb) {   // the packed map: P1 rows in the low word, P2 rows in the high one
                    const unsigned long long w = rmb[((size_t)s * NB + l1) * NB + l2];
                    b1 = (unsigned)w;
                    b2 = (unsigned)(w >> 32);
                }
                bb = b1 | b2;
#pragma unroll
                for (int l3 = 0; l3 < NB; l3++) {   // an unstored row reads the zero line instead
                    P1[l3] = *((b1 >> l3 & 1u) ? P1sm + row + (size_t)l3 * nk : pzero);
                    P2[l3] = *((b2 >> l3 & 1u) ? P2sm + row + (size_t)l3 * nk : pzero);
                }
            } else {
#pragma unroll
                for (int l3 = 0; l3 < NB; l3++) {
                    P1[l3] = P1sm[row + (size_t)l3 * nk];
                    P2[l3] = P2sm[row + (size_t)l3 * nk];
                }
            }
            // select the occupations of branch l2 without dynamic register indexing
            double dd[NT];
#pragma unroll
            for (int t = 0; t < NT; t++) {
                if constexpr (kRegD2) {
                    dd[t] = d2[t][0];
#pragma unroll
                    for (int l = 1; l < NB; l++) dd[t] = (l2 == l) ? d2[t][l] : dd[t];
                } else {
                    dd[t] = dist[t * dstr + (size_t)jc * NB + l2];
                }
            }
#pragma unroll
            for (int t = 0; t < NT; t++) {
                const double f1 = d1[t] * dd[t], f2 = d1[t] * (dd[t] + 1);
                double r1 = 0, r2 = 0;
                // chunks of eight l3 without a stored row add only zeros and are skipped (uniform)
#pragma unroll
                for (int c = 0; c < NB; c += 8) {
                    if (bb >> c & 0xffu) {
#pragma unroll
                        for (int l3 = c; l3 < NB && l3 < c + 8; l3++) {
                            const double p1 = (double)P1[l3] * (f1 * n3p[t][l3]);
                            const double p2 = (double)P2[l3] * (f2 * n3m[t][l3]);
                            q[t] = q[t] + (p1 + 0.5 * p2);
                            if (in1) qn[t] += p1; else qu[t] += p1;
                            if (in2) qn[t] += 0.5 * p2; else qu[t] += 0.5 * p2;
                            r1 += p1;
                            r2 += p2;
                        }
                    }
                }
                if (valid) {
                    const size_t o = (((size_t)s * NB + l1) * NB + l2) * (size_t)nk + j;
                    R1[t * rstr + o] = (PT)r1;
                    R2[t * rstr + o] = (PT)r2;
                }
            }
        }
        // the group's 3*NT partial sums: butterflies inside each of the four wavefronts, then the
        // wavefront partials through LDS, added in wavefront order by thread x < 3*NT
#pragma unroll
        for (int t = 0; t < NT; t++) {
            if (!valid) q[t] = qn[t] = qu[t] = 0;
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) {
                q[t] += __shfl_xor(q[t], off, 64);
                qn[t] += __shfl_xor(qn[t], off, 64);
                qu[t] += __shfl_xor(qu[t], off, 64);
            }
        }
        const int wv = threadIdx.x >> 6;
        if ((threadIdx.x & 63) == 0) {
#pragma unroll
            for (int t = 0; t < NT; t++) {
                sh[(3 * t) * 4 + wv] = q[t];
                sh[(3 * t + 1) * 4 + wv] = qn[t];
                sh[(3 * t + 2) * 4 + wv] = qu[t];
            }
        }
        __syncthreads();
        if (threadIdx.x < 3 * NT) {
            const int x = threadIdx.x, t = x / 3;
            const double tot = ((sh[4 * x] + sh[4 * x + 1]) + sh[4 * x + 2]) + sh[4 * x + 3];
            qpart[t * qstr + 3 * (((size_t)s * NB + l1) * nch + ch) + x % 3] = tot;
        }
        __syncthreads();
    }
}

// Q, QN, QU of (s, l1) = sum of the parts partial sums of that row; with the isotope term
// (piso [nibz][nb], else NULL) Piso of the irreducible point ibz_lo + s is added to all three, as
// calculate_tauRTA_iso does (sub.f90:8502-8504)
__global__ void k_tau_sum(int n, int parts, int nb, int ibz_lo, const double *__restrict__ qpart,
                          const double *__restrict__ piso, double *__restrict__ Q, double *__restrict__ QN,
                          double *__restrict__ QU)
{
    int x = blockIdx.x * blockDim.x + threadIdx.x;   // (s, l1)
    if (x >= n) return;
    double q = 0, qn = 0, qu = 0;
    for (int l2 = 0; l2 < parts; l2++) {
        const double *p = qpart + 3 * ((size_t)x * parts + l2);
        q += p[0];
        qn += p[1];
        qu += p[2];
    }
    if (piso) {
        const double pi = piso[(size_t)ibz_lo * nb + x];
        q = q + pi;
        qn = qn + pi;
        qu = qu + pi;
    }
    Q[x] = q;
    QN[x] = qn;
    QU[x] = qu;
}

// tauinv = Q/(n(n+1)) c 100 ; outputs Fortran (ksub,nb)
__global__ void k_qandtau(int ksub, int nb, int ibz_lo, const int *__restrict__ mapinv, const double *__restrict__ dist,
                          const double *__restrict__ Q, const double *__restrict__ QN, const double *__restrict__ QU,
                          double *__restrict__ fQ, double *__restrict__ fQN, double *__restrict__ fQU,
                          double *__restrict__ tN, double *__restrict__ tU, double *__restrict__ tT)
{
    int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= ksub * nb) return;
    int s = x / nb, l = x % nb;
    double nb1 = dist[(size_t)mapinv[ibz_lo + s] * nb + l];
    size_t f = s + (size_t)ksub * l;
    fQ[f] = Q[x];
    fQN[f] = QN[x];
    fQU[f] = QU[x];
    double a = QN[x] / (nb1 * (nb1 + 1)) * alt3::C_LIGHT * 100;
    double b = QU[x] / (nb1 * (nb1 + 1)) * alt3::C_LIGHT * 100;
    tN[f] = a;
    tU[f] = b;
    tT[f] = a + b;
}

__device__ __forceinline__ void opmul(const double *op, const double *f, double *o)
{
    for (int r = 0; r < 3; r++) {
        double acc = 0;
        for (int c = 0; c < 3; c++) acc = acc + op[3 * c + r] * f[c];
        o[r] = acc;
    }
}

// The per-mode star sums below run one wavefront per irreducible mode: lane s evaluates the
// terms of star member s, and lane 0 adds them in member order, which is the reference's
// summation order.  kStarWave modes per 256-thread block; sh holds 12 doubles per lane.
constexpr int kStarWave = 4;

// mode_kappa (sub.f90:1292-1401) for irreducible mode (i, ii) with distribution F and velocity
// v at the IBZ point: the partial kappa over the mesh star and the tauinv_eff sum (lane 0)
__device__ void mode_kappa_wave(int i, int ii, int nb, const double *F, const double *v,
                                const double *__restrict__ eival, double temp, int classical, double const1,
                                double const2, double cnst, const int *__restrict__ star_n,
                                const int *__restrict__ star_op, const int *__restrict__ star_n0,
                                const double *__restrict__ opk, double *sh, double *kp, double &te,
                                int *__restrict__ err, bool valid)
{
    const int lane = threadIdx.x & 63, narms = star_n[i];
    double *my = sh + 12 * lane;
    if (valid && lane < narms) {
        const double *op = opk + 9 * star_op[48 * i + lane];
        double F12[3], vc[3];
        opmul(op, F, F12);
        opmul(op, v, vc);
        const int n0 = star_n0[48 * i + lane];
        const double om = sqrt(eival[(size_t)n0 * nb + ii]) * cnst;
        if (!(om > 0)) atomicExch(err, 1);
        const double nb0 = nbe(om, temp, classical);
        for (int j = 0; j < 3; j++) {
            my[j] = const1 * om * vc[j] * nb0 * (nb0 + 1);
            my[3 + j] = F12[j];
            my[6 + j] = vc[j] == 0.0 ? const2 * F12[j] / om / (double)0.01f   // REAL*4 literal (sub.f90:1362)
                                     : const2 * F12[j] / om / vc[j] / (narms * 3.0);
        }
    }
    __syncthreads();   // every thread of the block reaches it (invalid modes carry valid = false)
    if (valid && lane == 0) {
        for (int e = 0; e < 9; e++) kp[e] = 0;
        te = 0;
        for (int m = 0; m < narms; m++) {
            const double *t = sh + 12 * m;
            for (int j = 0; j < 3; j++) {
                for (int jj = 0; jj < 3; jj++) kp[j + 3 * jj] = kp[j + 3 * jj] - t[j] * t[3 + jj];
                te = te - t[6 + j];
            }
        }
    }
}

// F_RTA (nibz,nb,3) Fortran layout, then per (i,ii) partial kappa (9) over the mesh star; one
// wavefront per mode
__global__ void __launch_bounds__(64 * kStarWave) k_kappa_rta(int nibz, int nb, int nk, const int *__restrict__ mapinv,
                            const double *__restrict__ freq, const double *__restrict__ dist,
                            const double *__restrict__ eival, const double *__restrict__ vel,
                            const double *__restrict__ Q /*[i][nb]*/, double tempk, double temp, int classical,
                            double const1, double const2, double cnst, const int *__restrict__ star_n,
                            const int *__restrict__ star_op, const int *__restrict__ star_n0,
                            const double *__restrict__ opk, double *__restrict__ FRTA, double *__restrict__ kpart,
                            double *__restrict__ taueff, int *__restrict__ err)
{
    __shared__ double sh[kStarWave][64 * 12];
    const int w = threadIdx.x >> 6, x0 = blockIdx.x * kStarWave + w;
    const bool valid = x0 < nibz * nb;
    const int x = valid ? x0 : 0, i = x / nb, ii = x % nb, m = mapinv[i], lane = threadIdx.x & 63;
    const double omega = freq[(size_t)m * nb + ii], nbv = dist[(size_t)m * nb + ii];
    double F[3], v[3];
    for (int a = 0; a < 3; a++) {
        v[a] = vel[((size_t)m * nb + ii) * 3 + a];
        F[a] = -(alt3::C_LIGHT * v[a]) * alt3::H_PLANK * (omega * alt3::C_LIGHT * 100) * nbv * (nbv + 1) /
               (alt3::K_B * tempk * tempk) / (Q[(size_t)i * nb + ii] * alt3::C_LIGHT * 100);
    }
    double kp[9], te;
    mode_kappa_wave(i, ii, nb, F, v, eival, temp, classical, const1, const2, cnst, star_n, star_op, star_n0, opk, sh[w],
                    kp, te, err, valid);
    if (valid && lane == 0) {
        for (int a = 0; a < 3; a++) FRTA[i + (size_t)nibz * (ii + (size_t)nb * a)] = F[a];
        for (int e = 0; e < 9; e++) kpart[(size_t)x * 9 + e] = kp[e];
        // mode_kappa zeroes the whole tauinv_eff array for every (i,ii): only the last survives
        taueff[i + (size_t)nibz * ii] = (i == nibz - 1 && ii == nb - 1) ? te : 0.0;
    }
}

// kappa(ii, j, jj) Fortran (nb,3,3) = sum over i of the partials: one wavefront per output,
// lane partial sums over i = lane, lane+64, ... then a shuffle tree
__global__ void __launch_bounds__(64) k_kappa_sum(int nibz, int nb, const double *__restrict__ kpart,
                                                  double *__restrict__ kappa, const int *__restrict__ skip)
{
    if (skip && *skip) return;
    const int x = blockIdx.x, ii = x / 9, e = x % 9;
    double acc = 0;
    for (int i = threadIdx.x; i < nibz; i += 64) acc += kpart[((size_t)i * nb + ii) * 9 + e];
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
    if (threadIdx.x == 0) kappa[ii + (size_t)nb * e] = acc;
}

// collision matrix, general fallback.  One workgroup per row block (s,l1); threads over
// column blocks (n2,l2), whose star gathers stay inside the (s,l1) slab of R1/R2.  The
// device matrix is row-major (CMr[row*ncol + col]) so the three rows of a block are written
// contiguously.
template <typename RT, typename MT>
__global__ void __launch_bounds__(256) k_collision(alt3::MeshDev m, int nk, int nb, int ibz_lo, int nibz,
                                                   const int *__restrict__ mapinv, const double *__restrict__ Q,
                                                   const RT *__restrict__ R1, const RT *__restrict__ R2,
                                                   const int *__restrict__ star_n, const int *__restrict__ star_op,
                                                   const int *__restrict__ star_n0, const double *__restrict__ opk,
                                                   const double *__restrict__ piso, MT *__restrict__ CMr)
{
    const int rb = blockIdx.x;  // (s, l1)
    const int s = rb / nb, l1 = rb % nb;
    const int k1 = mapinv[ibz_lo + s];
    int a0, a1, a2;
    m.coords(k1, a0, a1, a2);
    const size_t ncol = (size_t)nibz * nb * 3;
    const double qd = Q[(size_t)s * nb + l1];
    const double pis = piso ? piso[(size_t)(ibz_lo + s) * nb + l1] : 0.0;
    for (int cb = threadIdx.x; cb < nibz * nb; cb += blockDim.x) {
        const int n2 = cb / nb, l2 = cb % nb;
        const size_t rbase = ((size_t)rb * nb + l2) * (size_t)nk;
        double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        for (int ns = 0; ns < star_n[n2]; ns++) {
            int n0 = star_n0[48 * n2 + ns];
            double C;
            if (n0 == k1 && l1 == l2) {
                C = -qd;
            } else if (piso) {
                int c0, c1, c2;
                m.coords(n0, c0, c1, c2);
                const int p3 = m.sum(a0, a1, a2, c0, c1, c2);   // q1+q2
                C = -(double)R1[rbase + n0] + (double)R1[rbase + p3] + (double)R2[rbase + n0] + pis;
            } else {
                int c0, c1, c2;
                m.coords(n0, c0, c1, c2);
                const int p1 = m.diff(a0, a1, a2, c0, c1, c2);  // q1-q2
                const int p2 = m.diff(c0, c1, c2, a0, a1, a2);  // q2-q1
                C = -(double)R1[rbase + n0] + (double)R1[rbase + p2] + 0.5 * (double)R2[rbase + n0] +
                    0.5 * (double)R2[rbase + p1];
            }
            const double *op = opk + 9 * star_op[48 * n2 + ns];
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++) acc[i + 3 * j] = acc[i + 3 * j] + (C * op[3 * j + i]);
        }
        const size_t d1 = 3 * (size_t)rb, d2 = 3 * (size_t)cb;
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) CMr[(d1 + i) * ncol + d2 + j] = (MT)acc[i + 3 * j];
    }
}

// Phase 2 of the LDS collision kernels: the sums of C(q0) op over the mesh points q0 of each
// irreducible q2's star, into row block rb, column branch l2 of the row-major matrix M
// (ncol = nibz*nb*3).  Each thread takes tasks of the star table (longest first): a task is a
// whole star, or one half of a star longer than the split threshold, so that no thread walks
// a 48-member star while the others wait; the two halves meet in LDS (part, 18 doubles per
// split star) after a barrier.  Members are summed in order within a task.  Every thread of
// the block calls this.
template <typename MT>
__device__ __forceinline__ void star_sums(int nibz, int nb, int rb, int l2, const double *C, const double *ops,
                                          double *part, const StarTasks &st, MT *__restrict__ M)
{
    const size_t ncol = (size_t)nibz * nb * 3, d1 = 3 * (size_t)rb;
    for (int t = threadIdx.x; t < st.ntask; t += blockDim.x) {
        const int4 tk = st.task[t];   // slot, first member, end, part (-1: whole star)
        double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll 4   // independent table loads in flight; the sums keep member order
        for (int ns = tk.y; ns < tk.z; ns++) {
            const int e = st.tab[(size_t)ns * nibz + tk.x];
            const double c = C[e >> 6];
            const double *op = ops + 9 * (e & 63);
#pragma unroll
            for (int i = 0; i < 3; i++)
#pragma unroll
                for (int j = 0; j < 3; j++) acc[i + 3 * j] = acc[i + 3 * j] + (c * op[3 * j + i]);
        }
        if (tk.w < 0) {
            const size_t d2 = 3 * ((size_t)st.slot_n2[tk.x] * nb + l2);
#pragma unroll
            for (int i = 0; i < 3; i++)
#pragma unroll
                for (int j = 0; j < 3; j++) M[(d1 + i) * ncol + d2 + j] = (MT)acc[i + 3 * j];
        } else {
#pragma unroll
            for (int k = 0; k < 9; k++) part[9 * tk.w + k] = acc[k];
        }
    }
    if (st.nsplit > 0) {   // uniform over the block
        __syncthreads();
        for (int p = threadIdx.x; p < st.nsplit; p += blockDim.x) {
            const size_t d2 = 3 * ((size_t)st.slot_n2[st.split_x[p]] * nb + l2);
#pragma unroll
            for (int i = 0; i < 3; i++)
#pragma unroll
                for (int j = 0; j < 3; j++)
                    M[(d1 + i) * ncol + d2 + j] = (MT)(part[9 * (2 * p) + i + 3 * j] + part[9 * (2 * p + 1) + i + 3 * j]);
        }
    }
}

// Collision matrix with the per-point collision term in LDS.  One workgroup per (row block
// (s,l1), l2).  Phase 1 builds C(q0) = -R1(q0) + R1(q0-q1) + R2(q0)/2 + R2(q1-q0)/2 (or -Q
// on the diagonal) for every mesh point q0 from row-contiguous reads of R1/R2; phase 2 gives
// each thread one irreducible q2 (slots ordered by decreasing star size, so that the lanes of
// a wave run the same number of star members) and sums C(q0) op over its star from LDS.  The
// star table is stored member-major (tab[ns][slot] = q0 << 6 | op) for coalesced reads.
template <typename RT, typename MT>
__global__ void __launch_bounds__(256) k_collision_c(alt3::MeshDev m, int nk, int nb, int ibz_lo, int nibz, int nops,
                                                     const int *__restrict__ mapinv, const double *__restrict__ Q,
                                                     const RT *__restrict__ R1, const RT *__restrict__ R2,
                                                     const StarTasks st, const double *__restrict__ opk,
                                                     const double *__restrict__ piso, MT *__restrict__ CMr)
{
    extern __shared__ double lds[];
    double *C = lds;            // [nk]
    double *ops = lds + nk;     // [nops][9]
    double *part = ops + 9 * nops;   // [2 nsplit][9]
    const int l2 = blockIdx.x % nb, rb = blockIdx.x / nb;  // rb = (s, l1)
    const int s = rb / nb, l1 = rb % nb;
    const int k1 = mapinv[ibz_lo + s];
    int a0, a1, a2;
    m.coords(k1, a0, a1, a2);
    const size_t rbase = ((size_t)rb * nb + l2) * (size_t)nk;
    for (int x = threadIdx.x; x < 9 * nops; x += blockDim.x) ops[x] = opk[x];
    if (piso) {   // CollisionMatrix_iso: -R1(q0) + R1(q1+q0) + R2(q0) + Piso
        const double pis = piso[(size_t)(ibz_lo + s) * nb + l1];
        for (int n0 = threadIdx.x; n0 < nk; n0 += blockDim.x) {
            int c0, c1, c2;
            m.coords(n0, c0, c1, c2);
            const int p3 = m.sum(a0, a1, a2, c0, c1, c2);   // q1+q0
            C[n0] = -(double)R1[rbase + n0] + (double)R1[rbase + p3] + (double)R2[rbase + n0] + pis;
        }
    } else {
        for (int n0 = threadIdx.x; n0 < nk; n0 += blockDim.x) {
            int c0, c1, c2;
            m.coords(n0, c0, c1, c2);
            const int p1 = m.diff(a0, a1, a2, c0, c1, c2);  // q1-q0
            const int p2 = m.diff(c0, c1, c2, a0, a1, a2);  // q0-q1
            C[n0] = -(double)R1[rbase + n0] + (double)R1[rbase + p2] + 0.5 * (double)R2[rbase + n0] +
                    0.5 * (double)R2[rbase + p1];
        }
    }
    __syncthreads();
    // the diagonal: -Q, which holds Piso when the isotope term is on (-Q3 - Piso, as there)
    if (threadIdx.x == 0 && l1 == l2) C[k1] = -Q[(size_t)s * nb + l1];
    __syncthreads();
    star_sums(nibz, nb, rb, l2, C, ops, part, st, CMr);
}

// update_mpi (sub.f90:2295-2449) as a matrix on the irreducible F, built like the collision
// matrix: one workgroup per (row block (s,l1), branch l), with the coefficient of F(q0, l)
//   U(q0) = -R1(q0) + R2(q0)/2 + S1(q0-q1) + S2(q1-q0)/2
// (R: the -P1 F(q2,l2) and P2/2 F(q2,l2) terms; S: the P1 F(q1+q2,l3) and P2/2 F(q1-q2,l3)
// terms, re-indexed by q0) and no diagonal replacement.  A(F) = U F.
template <typename RT>
__global__ void __launch_bounds__(256) k_update_c(alt3::MeshDev m, int nk, int nb, int ibz_lo, int nibz, int nops,
                                                  const int *__restrict__ mapinv, const RT *__restrict__ R1,
                                                  const RT *__restrict__ R2, const RT *__restrict__ S1,
                                                  const RT *__restrict__ S2, const StarTasks st,
                                                  const double *__restrict__ opk, double *__restrict__ U)
{
    extern __shared__ double lds[];
    double *C = lds;            // [nk]
    double *ops = lds + nk;     // [nops][9]
    double *part = ops + 9 * nops;   // [2 nsplit][9]
    const int l = blockIdx.x % nb, rb = blockIdx.x / nb;  // rb = (s, l1)
    const int s = rb / nb;
    const int k1 = mapinv[ibz_lo + s];
    int a0, a1, a2;
    m.coords(k1, a0, a1, a2);
    const size_t rbase = ((size_t)rb * nb + l) * (size_t)nk;
    for (int x = threadIdx.x; x < 9 * nops; x += blockDim.x) ops[x] = opk[x];
    for (int n0 = threadIdx.x; n0 < nk; n0 += blockDim.x) {
        int c0, c1, c2;
        m.coords(n0, c0, c1, c2);
        const int p1 = m.diff(a0, a1, a2, c0, c1, c2);  // q1-q0
        const int p2 = m.diff(c0, c1, c2, a0, a1, a2);  // q0-q1
        C[n0] = -(double)R1[rbase + n0] + (double)S1[rbase + p2] + 0.5 * (double)R2[rbase + n0] +
                0.5 * (double)S2[rbase + p1];
    }
    __syncthreads();
    star_sums(nibz, nb, rb, l, C, ops, part, st, U);
}

// update_iso (sub.f90:2456-2610) adds Piso(n1,l1) Fsq2 to A after its loops, Fsq2 being what
// the loops left there: F of the last branch at the last irreducible point, rotated by the op
// of the last member of that point's star.  In the matrix U: Piso(n1,l1) op(i,j) in column
// (nibz-1, nb-1, j) of every row (s, l1, i).  One thread per row.
__global__ void k_update_iso(int ksub, int nb, int nibz, int ibz_lo, const double *__restrict__ piso,
                             const double *__restrict__ op, double *__restrict__ U)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;   // (s, l1, i)
    if (r >= ksub * nb * 3) return;
    const int i = r % 3, sl = r / 3;
    const size_t ncol = (size_t)nibz * nb * 3, d2 = 3 * ((size_t)(nibz - 1) * nb + nb - 1);
    const double p = piso[(size_t)ibz_lo * nb + sl];
    for (int j = 0; j < 3; j++) U[(size_t)r * ncol + d2 + j] = U[(size_t)r * ncol + d2 + j] + p * op[3 * j + i];
}

// S1/S2[s][l1][l3][q2] = sum over l2 of P1sm, P2sm times the occupations of
// calculate_tauRTA_mpi2 (sub.f90:8054-8062): the weights of F(q1+q2, l3) and F(q1-q2, l3) in
// update_mpi.  One thread per (s, q2).
template <typename PT>
__global__ void __launch_bounds__(256) k_tau_s(alt3::MeshDev m, int nk, int nb, int ibz_lo,
                                               const int *__restrict__ mapinv, const double *__restrict__ dist,
                                               const PT *__restrict__ P1sm, const PT *__restrict__ P2sm,
                                               PT *__restrict__ S1, PT *__restrict__ S2,
                                               const unsigned char *__restrict__ rm1,
                                               const unsigned char *__restrict__ rm2)
{
    const int s = blockIdx.y, j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nk) return;
    const int k1 = mapinv[ibz_lo + s];
    int a0, a1, a2, c0, c1, c2;
    m.coords(k1, a0, a1, a2);
    m.coords(j, c0, c1, c2);
    const int np = m.sum(a0, a1, a2, c0, c1, c2), nm = m.diff(a0, a1, a2, c0, c1, c2);
    for (int l1 = 0; l1 < nb; l1++) {
        const double d1 = dist[(size_t)k1 * nb + l1];
        for (int l3 = 0; l3 < nb; l3++) {
            const double n3p = dist[(size_t)np * nb + l3] + 1, n3m = dist[(size_t)nm * nb + l3] + 1;
            double s1 = 0, s2 = 0;
            for (int l2 = 0; l2 < nb; l2++) {
                const double d2 = dist[(size_t)j * nb + l2];
                const size_t rid = (((size_t)s * nb + l1) * nb + l2) * nb + l3, x = rid * (size_t)nk + j;
                // rows the row map marks zero are not stored (rm NULL: all stored)
                if (!rm1 || rm1[rid]) s1 += (double)P1sm[x] * ((d1 * d2) * n3p);
                if (!rm2 || rm2[rid]) s2 += (double)P2sm[x] * ((d1 * (d2 + 1)) * n3m);
            }
            const size_t o = (((size_t)s * nb + l1) * nb + l3) * (size_t)nk + j;
            S1[o] = (PT)s1;
            S2[o] = (PT)s2;
        }
    }
}

// F2 = F_RTA + A/Q (update_mpi, sub.f90:2437) and, with mix > 0, F <- F + mix (F2 - F)
// (kappa.f90:1456).  A and F in the vector layout 3*(i*nb+l)+alpha over rows i = ibz_lo..;
// F_RTA Fortran (nibz,nb,3); Q [s][nb].  F2 (optional) Fortran (ksub,nb,3).
__global__ void k_iter_mix(int ksub, int nibz, int nb, int ibz_lo, double mix, const double *__restrict__ FRTA,
                           const double *__restrict__ Q, const double *__restrict__ A, double *__restrict__ F,
                           double *__restrict__ F2out, const int *__restrict__ skip)
{
    if (skip && *skip) return;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ksub * nb * 3) return;
    const int a = e % 3, sl = e / 3, s = sl / nb, ii = sl % nb, i = ibz_lo + s;
    const double F2 = FRTA[i + (size_t)nibz * (ii + (size_t)nb * a)] + A[e] / Q[sl];
    if (F2out) F2out[s + (size_t)ksub * (ii + (size_t)nb * a)] = F2;
    if (F) F[e] = F[e] + mix * (F2 - F[e]);
}

// mode_kappa on F in the vector layout (the iterated FFF); one wavefront per mode
__global__ void __launch_bounds__(64 * kStarWave) k_mode_kappa(int nibz, int nb, const int *__restrict__ mapinv,
                             const double *__restrict__ x, const double *__restrict__ eival,
                             const double *__restrict__ vel, double temp, int classical, double const1, double const2,
                             double cnst, const int *__restrict__ star_n, const int *__restrict__ star_op,
                             const int *__restrict__ star_n0, const double *__restrict__ opk,
                             double *__restrict__ kpart, double *__restrict__ te_last, int *__restrict__ err,
                             const int *__restrict__ skip)
{
    if (skip && *skip) return;
    __shared__ double sh[kStarWave][64 * 12];
    const int w = threadIdx.x >> 6, t0 = blockIdx.x * kStarWave + w;
    const bool valid = t0 < nibz * nb;
    const int t = valid ? t0 : 0, i = t / nb, ii = t % nb, m = mapinv[i];
    const double F[3] = {x[3 * (size_t)t], x[3 * (size_t)t + 1], x[3 * (size_t)t + 2]};
    const double v[3] = {vel[((size_t)m * nb + ii) * 3], vel[((size_t)m * nb + ii) * 3 + 1],
                         vel[((size_t)m * nb + ii) * 3 + 2]};
    double kp[9], te;
    mode_kappa_wave(i, ii, nb, F, v, eival, temp, classical, const1, const2, cnst, star_n, star_op, star_n0, opk, sh[w],
                    kp, te, err, valid);
    if (valid && (threadIdx.x & 63) == 0) {
        for (int e = 0; e < 9; e++) kpart[(size_t)t * 9 + e] = kp[e];
        // mode_kappa clears tauinv_eff for every mode: only the last mode's value survives
        if (t == nibz * nb - 1) *te_last = te;
    }
}

// convergence test of kappa.f90:1466-1476: normk = sqrt(sum((kappa - kappa_old)^2 / (9 nb)));
// kappa (and the surviving tauinv_eff value) is appended to the history and kappa becomes
// kappa_old; state[0] = converged, state[1] = iterations done
__global__ void k_iter_check(int nb, int it, double tol, const double *__restrict__ kappa, double *__restrict__ kold,
                             double *__restrict__ hist, const double *__restrict__ te_last,
                             double *__restrict__ te_hist, int *__restrict__ state)
{
    if (state[0] || threadIdx.x != 0) return;
    te_hist[it] = *te_last;
    const int n = 9 * nb;
    double e = 0;
    for (int x = 0; x < n; x++) {
        const double d = kappa[x] - kold[x];
        e = e + d * d / n;
        hist[(size_t)it * n + x] = kappa[x];
        kold[x] = kappa[x];
    }
    state[1] = it + 1;
    if (sqrt(e) < tol) state[0] = 1;
}

// F_RTA Fortran (nibz,nb,3) -> vector layout 3*(i*nb+l)+alpha
__global__ void k_frta_to_vec(int nibz, int nb, const double *__restrict__ FRTA, double *__restrict__ x)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nibz * nb * 3) return;
    const int a = e % 3, il = e / 3, i = il / nb, ii = il % nb;
    x[e] = FRTA[i + (size_t)nibz * (ii + (size_t)nb * a)];
}

// row-major -> Fortran column-major copy (nrow x ncol)
template <typename MT>
__global__ void k_rows_to_cols(size_t nrow, size_t ncol, const MT *__restrict__ in, double *__restrict__ out)
{
    size_t o = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= nrow * ncol) return;
    size_t r = o % nrow, c = o / nrow;
    out[o] = in[r * ncol + c];
}

// y = A x for a row-major nrow x ncol matrix: one wavefront per row, columns in order per lane
template <typename AT>
__global__ void __launch_bounds__(256) k_matvec_rows(int nrow, int ncol, const AT *__restrict__ A,
                                                     const double *__restrict__ x, double *__restrict__ y,
                                                     const int *__restrict__ skip)
{
    if (skip && *skip) return;   // iteration already converged
    const int lane = threadIdx.x & 63;
    const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (row >= nrow) return;
    const AT *a = A + (size_t)row * ncol;
    double acc = 0;
    for (int j = lane; j < ncol; j += 64) acc += (double)a[j] * x[j];
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
    if (lane == 0) y[row] = acc;
}

__global__ void k_rhs(int nibz, int nb, double tempk, const int *__restrict__ mapinv, const double *__restrict__ freq,
                      const double *__restrict__ dist, const double *__restrict__ vel, double *__restrict__ RHS)
{
    int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= nibz * nb) return;
    int i = x / nb, ii = x % nb, n0 = mapinv[i];
    double omega = freq[(size_t)n0 * nb + ii], nb0 = dist[(size_t)n0 * nb + ii];
    for (int a = 0; a < 3; a++)
        RHS[3 * (size_t)x + a] = (alt3::C_LIGHT * vel[((size_t)n0 * nb + ii) * 3 + a]) * alt3::H_PLANK * (omega * alt3::C_LIGHT * 100) *
                                 nb0 * (nb0 + 1) / (alt3::K_B * tempk * tempk) / (alt3::C_LIGHT * 100);
}

// ------------------------------------------------------------------ CG (HBM-bound)
// y = A x, A column-major n x n: one thread per row, sequential over columns
__global__ void __launch_bounds__(256) k_matvec(int n, const double *__restrict__ A, const double *__restrict__ x,
                                                double *__restrict__ y, const int *__restrict__ skip)
{
    if (skip && *skip) return;   // CG already converged
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double acc = 0;
    const double *a = A + i;
    int j = 0;
    for (; j + 4 <= n; j += 4) {
        double v0 = a[(size_t)n * j] * x[j];
        double v1 = a[(size_t)n * (j + 1)] * x[j + 1];
        double v2 = a[(size_t)n * (j + 2)] * x[j + 2];
        double v3 = a[(size_t)n * (j + 3)] * x[j + 3];
        acc = acc + v0;
        acc = acc + v1;
        acc = acc + v2;
        acc = acc + v3;
    }
    for (; j < n; j++) acc = acc + a[(size_t)n * j] * x[j];
    y[i] = acc;
}

// partial dot products: out[blockIdx.x] for pairs (x1,y1), (x2,y2)
__global__ void k_dot2(int n, const double *__restrict__ x1, const double *__restrict__ y1, const double *__restrict__ x2,
                       const double *__restrict__ y2, double *__restrict__ out)
{
    __shared__ double sh[256];
    double a = 0, b = 0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        a += x1[i] * y1[i];
        if (x2) b += x2[i] * y2[i];
    }
    a = block_sum(a, sh);
    b = block_sum(b, sh);
    if (threadIdx.x == 0) {
        out[2 * blockIdx.x] = a;
        out[2 * blockIdx.x + 1] = b;
    }
}

// final reduction of k_dot2 partials into sc[slot], sc[slot+1]
__global__ void k_reduce2(int nblk, const double *__restrict__ part, double *__restrict__ sc, int slot)
{
    __shared__ double sh[256];
    double a = 0, b = 0;
    for (int i = threadIdx.x; i < nblk; i += blockDim.x) {
        a += part[2 * i];
        b += part[2 * i + 1];
    }
    a = block_sum(a, sh);
    b = block_sum(b, sh);
    if (threadIdx.x == 0) {
        sc[slot] = a;
        sc[slot + 1] = b;
    }
}

// v0 = grad; vmin -= num/den conj; grad -= num/den aux    (sc[0]=num, sc[1]=den)
// conj = -grad + (num/den) conj       (sc[4]=grad.grad, sc[5]=v0.v0)
// One CG iteration of cg_quadratic (cg.f90:33-50) after aux = A conj, in a single workgroup
// (the vectors have nibz*nb*3 entries): num, den, the steps of vmin and grad, the error
// (num/den)^2 |conj|^2 / |vmin|^2 against the tolerance, and the new direction.  state[0]
// turns 1 at the converged iteration (later launches return at once); state[1] = iteration.
__global__ void __launch_bounds__(1024) k_cg_update(int n, int it, double tol, double *__restrict__ vmin,
                                                    double *__restrict__ grad, double *__restrict__ conj,
                                                    const double *__restrict__ aux, double *__restrict__ v0,
                                                    int *__restrict__ state)
{
    __shared__ double sh[32];
    if (state[0]) return;
    const int tid = threadIdx.x, nt = blockDim.x;
    double a = 0, b = 0;
    for (int i = tid; i < n; i += nt) {
        a += grad[i] * conj[i];
        b += aux[i] * conj[i];
    }
    block_sum2(a, b, sh);
    const double num = a, den = b;
    const double r = num / den;
    a = b = 0;
    for (int i = tid; i < n; i += nt) {
        v0[i] = grad[i];
        vmin[i] = vmin[i] - r * conj[i];
        grad[i] = grad[i] - r * aux[i];
        a += vmin[i] * vmin[i];
        b += conj[i] * conj[i];
    }
    block_sum2(a, b, sh);
    const double v2 = a, cc = b;
    const double err = (num * num / den / den) * cc / v2;
    if (err < tol) {
        if (tid == 0) {
            state[0] = 1;
            state[1] = it;
        }
        return;
    }
    a = b = 0;
    for (int i = tid; i < n; i += nt) {
        a += grad[i] * grad[i];
        b += v0[i] * v0[i];
    }
    block_sum2(a, b, sh);
    const double gg = a, vv = b;
    for (int i = tid; i < n; i += nt) conj[i] = -grad[i] + gg / vv * conj[i];
    if (tid == 0) state[1] = it;
}

__global__ void k_axpy_init(int n, const double *__restrict__ ax, const double *__restrict__ b, double *__restrict__ grad,
                            double *__restrict__ conj)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double g = ax[i] - b[i];
    grad[i] = g;
    conj[i] = -g;
}

// star-averaged velocities with inverse operations (veloc_avergeIBZ_invsyFBZ); one wavefront per
// mode, the members' v/narms added in member order by lane 0
__global__ void __launch_bounds__(64 * kStarWave) k_veloci(int nibz, int nb, const double *__restrict__ vel,
                         const int *__restrict__ fstar_n, const int *__restrict__ fstar_op,
                         const int *__restrict__ fstar_n0, const double *__restrict__ opk_inv,
                         double *__restrict__ veloci)
{
    __shared__ double sh[kStarWave][64 * 3];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63, x0 = blockIdx.x * kStarWave + w;
    const bool valid = x0 < nibz * nb;
    const int x = valid ? x0 : 0, i = x / nb, j = x % nb, narms = fstar_n[i];
    if (valid && lane < narms) {
        const int n0 = fstar_n0[48 * i + lane];
        double vi[3];
        opmul(opk_inv + 9 * fstar_op[48 * i + lane], vel + ((size_t)n0 * nb + j) * 3, vi);
        for (int r = 0; r < 3; r++) sh[w][3 * lane + r] = vi[r] / narms;
    }
    __syncthreads();
    if (valid && lane == 0) {
        double va[3] = {0, 0, 0};
        for (int n = 0; n < narms; n++)
            for (int r = 0; r < 3; r++) va[r] = va[r] + sh[w][3 * n + r];
        for (int r = 0; r < 3; r++) veloci[(size_t)x * 3 + r] = va[r];
    }
}

// mode_kappa_d per mode, one wavefront per mode (terms per star member, ordered sums by lane 0)
__global__ void __launch_bounds__(64 * kStarWave) k_kappa_bte(int nibz, int nb, const double *__restrict__ x,
                            const double *__restrict__ veloci, const double *__restrict__ eival, double temp,
                            int classical, double const1, double const2, double cnst,
                            const int *__restrict__ star_n, const int *__restrict__ star_op,
                            const int *__restrict__ star_n0, const double *__restrict__ opk,
                            double *__restrict__ kpart, double *__restrict__ taueff, int *__restrict__ err)
{
    __shared__ double sh[kStarWave][64 * 10];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63, t0 = blockIdx.x * kStarWave + w;
    const bool valid = t0 < nibz * nb;
    const int t = valid ? t0 : 0, i = t / nb, ii = t % nb, narms = star_n[i];
    if (valid && lane < narms) {
        const double f[3] = {x[3 * (size_t)t], x[3 * (size_t)t + 1], x[3 * (size_t)t + 2]};
        const double *vi = veloci + (size_t)t * 3;
        const double *op = opk + 9 * star_op[48 * i + lane];
        double F12[3], vc[3];
        opmul(op, f, F12);
        opmul(op, vi, vc);
        const int n0 = star_n0[48 * i + lane];
        const double om = sqrt(eival[(size_t)n0 * nb + ii]) * cnst;
        if (!(om > 0)) atomicExch(err, 1);
        const double nb0 = nbe(om, temp, classical);
        double *my = sh[w] + 10 * lane;
        for (int j = 0; j < 3; j++) {
            my[j] = const1 * om * vc[j] * nb0 * (nb0 + 1);
            my[3 + j] = F12[j];
        }
        const double fv = F12[0] * vc[0] + F12[1] * vc[1] + F12[2] * vc[2];
        const double vv = vc[0] * vc[0] + vc[1] * vc[1] + vc[2] * vc[2];
        my[9] = const2 * fv / om / vv / (narms * 3.0);
    }
    __syncthreads();
    if (valid && lane == 0) {
        double kp[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        double te = taueff[i + (size_t)nibz * ii];
        for (int m = 0; m < narms; m++) {
            const double *tm = sh[w] + 10 * m;
            for (int j = 0; j < 3; j++) {   // the reference subtracts the same term once per j
                for (int jj = 0; jj < 3; jj++) kp[j + 3 * jj] = kp[j + 3 * jj] - tm[j] * tm[3 + jj];
                te = te - tm[9];
            }
        }
        for (int e = 0; e < 9; e++) kpart[(size_t)t * 9 + e] = kp[e];
        taueff[i + (size_t)nibz * ii] = te;
    }
}

}  // namespace

namespace alt3 {

static double temp_cm(double tempk) { return tempk * K_B / (100 * H_PLANK * C_LIGHT); }

int upload_bte_tables(alt3_ctx *ctx)
{
    HostTables &T = ctx->T;
    ALT3_TRY(ctx->d_star_n.alloc(T.nibz));
    ALT3_TRY(ctx->d_star_op.alloc(48 * T.nibz));
    ALT3_TRY(ctx->d_star_n0.alloc(48 * T.nibz));
    ALT3_TRY(ctx->d_fstar_n.alloc(T.nibz));
    ALT3_TRY(ctx->d_fstar_op.alloc(48 * T.nibz));
    ALT3_TRY(ctx->d_fstar_n0.alloc(48 * T.nibz));
    ALT3_TRY(ctx->d_opk.alloc(9 * T.nops));
    ALT3_TRY(ctx->d_opk_inv.alloc(9 * T.nops));
    ALT3_HIP(hipMemcpy(ctx->d_star_n.p, T.star_n.data(), sizeof(int) * T.nibz, hipMemcpyHostToDevice));
    ALT3_HIP(hipMemcpy(ctx->d_star_op.p, T.star_op.data(), sizeof(int) * 48 * T.nibz, hipMemcpyHostToDevice));
    ALT3_HIP(hipMemcpy(ctx->d_star_n0.p, T.star_n0.data(), sizeof(int) * 48 * T.nibz, hipMemcpyHostToDevice));
    ALT3_HIP(hipMemcpy(ctx->d_fstar_n.p, T.fstar_n.data(), sizeof(int) * T.nibz, hipMemcpyHostToDevice));
    ALT3_HIP(hipMemcpy(ctx->d_fstar_op.p, T.fstar_op.data(), sizeof(int) * 48 * T.nibz, hipMemcpyHostToDevice));
    ALT3_HIP(hipMemcpy(ctx->d_fstar_n0.p, T.fstar_n0.data(), sizeof(int) * 48 * T.nibz, hipMemcpyHostToDevice));
    ALT3_HIP(hipMemcpy(ctx->d_opk.p, T.opk.data(), sizeof(double) * 9 * T.nops, hipMemcpyHostToDevice));
    ALT3_HIP(hipMemcpy(ctx->d_opk_inv.p, T.opk_inv.data(), sizeof(double) * 9 * T.nops, hipMemcpyHostToDevice));
    // collision star table: irreducible points ordered by decreasing star size (stable),
    // members stored member-major as q0 << 6 | op
    std::vector<int> order(T.nibz);
    for (int i = 0; i < T.nibz; i++) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return T.star_n[a] > T.star_n[b]; });
    std::vector<int> cnt(T.nibz), tab((size_t)48 * T.nibz, 0);
    for (int x = 0; x < T.nibz; x++) {
        const int n2 = order[x];
        cnt[x] = T.star_n[n2];
        if (cnt[x] < 0 || cnt[x] > 48) return fail(ALT3_E_MESH, "star larger than 48 members");
        for (int ns = 0; ns < cnt[x]; ns++) {
            const int n0 = T.star_n0[48 * n2 + ns], op = T.star_op[48 * n2 + ns];
            if (n0 < 0 || n0 >= T.nk || op < 0 || op >= 48 || op >= T.nops)
                return fail(ALT3_E_MESH, "star member outside the mesh or operation table");
            tab[(size_t)ns * T.nibz + x] = (n0 << 6) | op;
        }
    }
    // tasks of the star sums: stars longer than kSplitAt members (up to kMaxSplit of them, the
    // longest) are split in two halves; longest tasks first
    constexpr int kSplitAt = 24, kMaxSplit = 128;
    std::vector<int4> tasks;
    std::vector<int> split_x;
    for (int x = 0; x < T.nibz; x++) {
        if (cnt[x] > kSplitAt && (int)split_x.size() < kMaxSplit) {
            const int p = (int)split_x.size(), h = cnt[x] / 2;
            split_x.push_back(x);
            tasks.push_back(make_int4(x, 0, h, 2 * p));
            tasks.push_back(make_int4(x, h, cnt[x], 2 * p + 1));
        } else {
            tasks.push_back(make_int4(x, 0, cnt[x], -1));
        }
    }
    std::stable_sort(tasks.begin(), tasks.end(), [](const int4 &a, const int4 &b) { return a.z - a.y > b.z - b.y; });
    ALT3_TRY(ctx->d_cm_slot_n2.alloc(T.nibz));
    ALT3_TRY(ctx->d_cm_tab.alloc((size_t)48 * T.nibz));
    ALT3_TRY(ctx->d_cm_task.alloc(tasks.size()));
    ALT3_TRY(ctx->d_cm_split_x.alloc(std::max<size_t>(1, split_x.size())));
    ALT3_HIP(hipMemcpy(ctx->d_cm_slot_n2.p, order.data(), sizeof(int) * T.nibz, hipMemcpyHostToDevice));
    ALT3_HIP(hipMemcpy(ctx->d_cm_tab.p, tab.data(), sizeof(int) * 48 * T.nibz, hipMemcpyHostToDevice));
    ALT3_HIP(hipMemcpy(ctx->d_cm_task.p, tasks.data(), sizeof(int4) * tasks.size(), hipMemcpyHostToDevice));
    if (!split_x.empty())
        ALT3_HIP(hipMemcpy(ctx->d_cm_split_x.p, split_x.data(), sizeof(int) * split_x.size(), hipMemcpyHostToDevice));
    ctx->cm_ntask = (int)tasks.size();
    ctx->cm_nsplit = (int)split_x.size();
    return ALT3_OK;
}

// Q, QN, QU and the l3-reduced R1/R2 for each of nt temperatures, P1sm/P2sm read once per
// group of up to four temperatures.  Temperature t lands in slot t of the per-temperature
// arrays; slot 0 is selected afterwards.
// sync = false (alt3_solve): no wait at the end; an N_BE error stays in the device flag, which
// the caller cleared before and reads after its last stage
int tau_multi_run(alt3_ctx *ctx, const double *temps, int nt, bool sync)
{
    if (!ctx->have_fgr) return fail(ALT3_E_STATE, "alt3_tau needs alt3_fgr first");
    if (nt < 1) return fail(ALT3_E_ARG, "no temperature");
    for (int t = 0; t < nt; t++)
        if (!(temps[t] > 0)) return fail(ALT3_E_ARG, "ERROR in N_BE: t=<0");
    const int nk = ctx->nk, nb = ctx->nb, ksub = ctx->ibz_hi - ctx->ibz_lo + 1;
    const MeshDev md = MeshDev::make(ctx->nc);
    ctx->have_tau = ctx->have_cm = ctx->have_upd = false;
    ctx->tau_temps.clear();
    ALT3_TRY(piso_run(ctx));
    int *derr = ctx->d_flags.p + 1;
    if (sync) ALT3_HIP(hipMemsetAsync(derr, 0, sizeof(int), ctx->stream));
    const size_t dstr = (size_t)nk * nb, qstr = (size_t)ksub * nb, rstr = (size_t)ksub * nb * nb * nk;
    ALT3_TRY(ctx->d_dist.alloc(nt * dstr));
    ALT3_TRY(ctx->d_q.alloc(nt * qstr));
    ALT3_TRY(ctx->d_qn.alloc(nt * qstr));
    ALT3_TRY(ctx->d_qu.alloc(nt * qstr));
    ctx->r_fp32 = ctx->p_fp32;
    if (ctx->r_fp32) {
        ctx->d_r1.release();
        ctx->d_r2.release();
        ALT3_TRY(ctx->d_r1f.alloc(nt * rstr));
        ALT3_TRY(ctx->d_r2f.alloc(nt * rstr));
    } else {
        ctx->d_r1f.release();
        ctx->d_r2f.release();
        ALT3_TRY(ctx->d_r1.alloc(nt * rstr));
        ALT3_TRY(ctx->d_r2.alloc(nt * rstr));
    }
    for (int t = 0; t < nt; t++)
        hipLaunchKernelGGL(k_dist, dim3((nk * nb + 255) / 256), dim3(256), 0, ctx->stream, nk * nb, ctx->d_freq.p,
                           temp_cm(temps[t]), ctx->classical, ctx->d_dist.p + t * dstr, derr);
    ALT3_HIP(hipGetLastError());
    const int nch = (nk + 255) / 256;
    // the P row map (rows of l1 no tetrahedron reaches are zero and not stored); packed per
    // (s, l1, l2) for k_tau_q2
    const unsigned char *rm1 = ctx->d_prow.p, *rm2 = rm1 ? rm1 + ctx->d_prow.n / 2 : nullptr;
    const unsigned long long *rmb = nullptr;
    if (ctx->d_pzero.n == 0) {   // one line of zeros for the loads of unstored rows
        ALT3_TRY(ctx->d_pzero.alloc(64));
        hipLaunchKernelGGL(k_fill_zero, dim3(1), dim3(64), 0, ctx->stream, 64, ctx->d_pzero.p);
    }
    if (rm1 && nb <= 32) {
        const size_t n = (size_t)ksub * nb * nb;
        ALT3_TRY(ctx->d_prowb.reserve(n));
        hipLaunchKernelGGL(k_prow_pack, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, n, nb,
                           ctx->d_prow.n / 2, rm1, ctx->d_prowb.p);
        rmb = ctx->d_prowb.p;
    }
    // k_tau_q2 (occupations at q1 +- q2 in registers): 6 branches with up to 4 temperatures per
    // pass, 12 with 2, 24 with 1; other cells k_tau
    const bool rows = nb == 6 || nb == 12 || nb == 24;
    const int tpp = nb == 6 ? 4 : nb == 12 ? 2 : nb == 24 ? 1 : 4;   // temperatures per pass
    const int parts = rows ? nch : nb;   // partial sums per (s, l1)
    const size_t nblk = (size_t)ksub * nb * nb, pstr = (size_t)3 * ksub * nb * parts;
    DBuf<double> &qpart = ctx->d_qpart;
    ALT3_TRY(qpart.reserve((size_t)nt * pstr));
    for (int t0 = 0; t0 < nt; t0 += tpp) {
        const int g = std::min(tpp, nt - t0);
        if (rows) {
            const dim3 grid((unsigned)nch, (unsigned)ksub);
#define ALT3_TAU_Q2(G, PT, NB, P1, P2, R1, R2)                                                                     \
    hipLaunchKernelGGL((k_tau_q2<G, PT, NB>), grid, dim3(256), 0, ctx->stream, md, nk, ctx->ibz_lo, ctx->d_mapinv.p, \
                       ctx->d_dist.p + t0 * dstr, dstr, P1, P2, qpart.p + t0 * pstr, pstr, (R1) + t0 * rstr,         \
                       (R2) + t0 * rstr, rstr, rmb, (const PT *)ctx->d_pzero.p)
#define ALT3_TAU_Q2_ALL(PT, P1, P2, R1, R2)                                 \
    if (nb == 24) ALT3_TAU_Q2(1, PT, 24, P1, P2, R1, R2);                    \
    else if (nb == 12 && g == 1) ALT3_TAU_Q2(1, PT, 12, P1, P2, R1, R2);     \
    else if (nb == 12) ALT3_TAU_Q2(2, PT, 12, P1, P2, R1, R2);               \
    else if (g == 1) ALT3_TAU_Q2(1, PT, 6, P1, P2, R1, R2);                  \
    else if (g == 2) ALT3_TAU_Q2(2, PT, 6, P1, P2, R1, R2);                  \
    else if (g == 3) ALT3_TAU_Q2(3, PT, 6, P1, P2, R1, R2);                  \
    else ALT3_TAU_Q2(4, PT, 6, P1, P2, R1, R2);
            if (ctx->p_fp32) {
                float *r1 = ctx->d_r1f.p, *r2 = ctx->d_r2f.p;
                ALT3_TAU_Q2_ALL(float, ctx->d_p1f.p, ctx->d_p2f.p, r1, r2)
            } else {
                double *r1 = ctx->d_r1.p, *r2 = ctx->d_r2.p;
                ALT3_TAU_Q2_ALL(double, ctx->d_p1.p, ctx->d_p2.p, r1, r2)
            }
#undef ALT3_TAU_Q2_ALL
#undef ALT3_TAU_Q2
        } else if (ctx->p_fp32) {
            auto *kern = g == 1 ? k_tau<1, float> : g == 2 ? k_tau<2, float> : g == 3 ? k_tau<3, float> : k_tau<4, float>;
            hipLaunchKernelGGL(kern, dim3((unsigned)nblk), dim3(256), 0, ctx->stream, md, nk, nb, ctx->ibz_lo,
                               ctx->d_mapinv.p, ctx->d_dist.p + t0 * dstr, dstr, ctx->d_p1f.p, ctx->d_p2f.p,
                               qpart.p + t0 * pstr, pstr, ctx->d_r1f.p + t0 * rstr, ctx->d_r2f.p + t0 * rstr, rstr,
                               rm1, rm2);
        } else {
            auto *kern = g == 1 ? k_tau<1, double> : g == 2 ? k_tau<2, double> : g == 3 ? k_tau<3, double> : k_tau<4, double>;
            hipLaunchKernelGGL(kern, dim3((unsigned)nblk), dim3(256), 0, ctx->stream, md, nk, nb, ctx->ibz_lo,
                               ctx->d_mapinv.p, ctx->d_dist.p + t0 * dstr, dstr, ctx->d_p1.p, ctx->d_p2.p,
                               qpart.p + t0 * pstr, pstr, ctx->d_r1.p + t0 * rstr, ctx->d_r2.p + t0 * rstr, rstr,
                               rm1, rm2);
        }
        ALT3_HIP(hipGetLastError());
    }
    const double *piso = ctx->giso.empty() ? nullptr : ctx->d_piso.p;
    for (int t = 0; t < nt; t++)
        hipLaunchKernelGGL(k_tau_sum, dim3((ksub * nb + 255) / 256), dim3(256), 0, ctx->stream, ksub * nb, parts, nb,
                           ctx->ibz_lo, qpart.p + t * pstr, piso, ctx->d_q.p + t * qstr, ctx->d_qn.p + t * qstr,
                           ctx->d_qu.p + t * qstr);
    ALT3_HIP(hipGetLastError());
    if (sync) {
        int herr = 0;
        ALT3_HIP(hipMemcpyAsync(&herr, derr, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
        ALT3_HIP(hipStreamSynchronize(ctx->stream));
        if (herr) return fail(ALT3_E_NBE, "ERROR in N_BE: w=<0 (a zero or imaginary mode frequency)");
    }
    ctx->tau_temps.assign(temps, temps + nt);
    ctx->tau_cur = 0;
    ctx->tempk = temps[0];
    ctx->have_tau = true;
    return ALT3_OK;
}

// make the slot of temperature tempk (from the last tau pass) current
int tau_select(alt3_ctx *ctx, double tempk, const char *who)
{
    if (!ctx->have_tau) return fail(ALT3_E_STATE, std::string(who) + " needs alt3_tau first");
    for (size_t t = 0; t < ctx->tau_temps.size(); t++)
        if (ctx->tau_temps[t] == tempk) {
            if ((int)t != ctx->tau_cur || ctx->tempk != tempk) ctx->have_cm = false;
            ctx->tau_cur = (int)t;
            ctx->tempk = tempk;
            return ALT3_OK;
        }
    return fail(ALT3_E_STATE, std::string(who) + " temperature differs from the last alt3_tau");
}

int tau_run(alt3_ctx *ctx, double tempk, double *Q, double *QN, double *QU, double *tN, double *tU, double *tT)
{
    ALT3_TRY(tau_multi_run(ctx, &tempk, 1, true));
    const int nb = ctx->nb, ksub = ctx->ibz_hi - ctx->ibz_lo + 1;
    if (Q || QN || QU || tN || tU || tT) {
        DBuf<double> o;
        size_t n = (size_t)ksub * nb;
        ALT3_TRY(o.alloc(6 * n));
        hipLaunchKernelGGL(k_qandtau, dim3((ksub * nb + 255) / 256), dim3(256), 0, ctx->stream, ksub, nb, ctx->ibz_lo,
                           ctx->d_mapinv.p, ctx->dist(), ctx->q(), ctx->qn(), ctx->qu(), o.p, o.p + n,
                           o.p + 2 * n, o.p + 3 * n, o.p + 4 * n, o.p + 5 * n);
        ALT3_HIP(hipGetLastError());
        double *dst[6] = {Q, QN, QU, tN, tU, tT};
        for (int k = 0; k < 6; k++)
            if (dst[k]) ALT3_HIP(hipMemcpyAsync(dst[k], o.p + k * n, n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        ALT3_HIP(hipStreamSynchronize(ctx->stream));
    }
    return ALT3_OK;
}

static int need_full_ibz(alt3_ctx *ctx, double tempk, const char *who)
{
    ALT3_TRY(tau_select(ctx, tempk, who));
    if (ctx->ibz_lo != 0 || ctx->ibz_hi != ctx->nibz - 1)
        return fail(ALT3_E_STATE, std::string(who) + " needs Q over the whole irreducible zone (ibz 0..nibz-1)");
    return ALT3_OK;
}

int kappa_rta_run(alt3_ctx *ctx, double tempk, double *kappa, double *FRTA, double *taueff, bool sync)
{
    ALT3_TRY(need_full_ibz(ctx, tempk, "alt3_kappa_rta"));
    const int nb = ctx->nb, nibz = ctx->nibz, nk = ctx->nk;
    ALT3_TRY(ctx->d_frta.alloc((size_t)nibz * nb * 3));
    ALT3_TRY(ctx->d_kpart.alloc((size_t)nibz * nb * 9));
    ALT3_TRY(ctx->d_kappa.alloc((size_t)nb * 9));
    ALT3_TRY(ctx->d_taueff.alloc((size_t)nibz * nb));
    int *derr = ctx->d_flags.p + 2;
    if (sync) ALT3_HIP(hipMemsetAsync(derr, 0, sizeof(int), ctx->stream));
    double const1 = H_PLANK * C_LIGHT * 100 * C_LIGHT / (nk * ctx->T.volume_r / 1e30);
    double const2 = K_B * tempk * tempk / (H_PLANK * C_LIGHT * 100) / C_LIGHT;
    hipLaunchKernelGGL(k_kappa_rta, dim3((nibz * nb + kStarWave - 1) / kStarWave), dim3(64 * kStarWave), 0, ctx->stream, nibz, nb, nk,
                       ctx->d_mapinv.p, ctx->d_freq.p, ctx->dist(), ctx->d_eival.p, ctx->d_vel.p, ctx->q(), tempk,
                       temp_cm(tempk), ctx->classical, const1, const2, cnst(), ctx->d_star_n.p, ctx->d_star_op.p,
                       ctx->d_star_n0.p, ctx->d_opk.p, ctx->d_frta.p, ctx->d_kpart.p, ctx->d_taueff.p, derr);
    ALT3_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_kappa_sum, dim3(nb * 9), dim3(64), 0, ctx->stream, nibz, nb, ctx->d_kpart.p, ctx->d_kappa.p, nullptr);
    ALT3_HIP(hipGetLastError());
    if (!sync) return ALT3_OK;   // outputs stay on the device (d_kappa, d_frta, d_taueff)
    int herr = 0;
    ALT3_HIP(hipMemcpyAsync(&herr, derr, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    if (kappa) ALT3_HIP(hipMemcpyAsync(kappa, ctx->d_kappa.p, sizeof(double) * nb * 9, hipMemcpyDeviceToHost, ctx->stream));
    if (FRTA) ALT3_HIP(hipMemcpyAsync(FRTA, ctx->d_frta.p, sizeof(double) * nibz * nb * 3, hipMemcpyDeviceToHost, ctx->stream));
    if (taueff) ALT3_HIP(hipMemcpyAsync(taueff, ctx->d_taueff.p, sizeof(double) * nibz * nb, hipMemcpyDeviceToHost, ctx->stream));
    ALT3_HIP(hipStreamSynchronize(ctx->stream));
    if (herr) return fail(ALT3_E_NBE, "ERROR in N_BE: w=<0 (a zero frequency on a star member)");
    return ALT3_OK;
}

int collision_run(alt3_ctx *ctx, double tempk, double *CM, double *RHS, bool sync)
{
    ALT3_TRY(tau_select(ctx, tempk, "alt3_collision"));
    const int nb = ctx->nb, nibz = ctx->nibz, nk = ctx->nk, ksub = ctx->ibz_hi - ctx->ibz_lo + 1;
    const MeshDev md = MeshDev::make(ctx->nc);
    const size_t nrow = (size_t)ksub * nb * 3, ncol = (size_t)nibz * nb * 3;
    // the mixed path (FP32 P and R) keeps the matrix in FP32: half the bytes for the kernel's
    // stores and for every CG product; the products and CG vectors stay FP64
    ctx->cm_fp32 = ctx->r_fp32;
    if (ctx->cm_fp32) {
        ctx->d_cm.release();
        ALT3_TRY(ctx->d_cmf.alloc(nrow * ncol));
    } else {
        ctx->d_cmf.release();
        ALT3_TRY(ctx->d_cm.alloc(nrow * ncol));
    }
    ALT3_TRY(ctx->d_rhs.alloc((size_t)nibz * nb * 3));
    const size_t lds = ((size_t)nk + 9 * ctx->nops + 18 * ctx->cm_nsplit) * sizeof(double);
    const double *piso = ctx->giso.empty() ? nullptr : ctx->d_piso.p;   // CollisionMatrix_iso
    auto launch = [&](auto *R1, auto *R2, auto *M) {
        using RT = std::remove_const_t<std::remove_pointer_t<decltype(R1)>>;
        using MT = std::remove_pointer_t<decltype(M)>;
        if (lds <= 65536)
            hipLaunchKernelGGL((k_collision_c<RT, MT>), dim3(ksub * nb * nb), dim3(256), lds, ctx->stream, md, nk, nb,
                               ctx->ibz_lo, nibz, ctx->nops, ctx->d_mapinv.p, ctx->q(), R1, R2, star_tasks(ctx),
                               ctx->d_opk.p, piso, M);
        else
            hipLaunchKernelGGL((k_collision<RT, MT>), dim3(ksub * nb), dim3(256), 0, ctx->stream, md, nk, nb,
                               ctx->ibz_lo, nibz, ctx->d_mapinv.p, ctx->q(), R1, R2, ctx->d_star_n.p, ctx->d_star_op.p,
                               ctx->d_star_n0.p, ctx->d_opk.p, piso, M);
    };
    if (ctx->r_fp32)
        launch((const float *)ctx->r1f(), (const float *)ctx->r2f(), ctx->d_cmf.p);
    else
        launch((const double *)ctx->r1(), (const double *)ctx->r2(), ctx->d_cm.p);
    ALT3_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_rhs, dim3((nibz * nb + 127) / 128), dim3(128), 0, ctx->stream, nibz, nb, tempk, ctx->d_mapinv.p,
                       ctx->d_freq.p, ctx->dist(), ctx->d_vel.p, ctx->d_rhs.p);
    ALT3_HIP(hipGetLastError());
    if (CM) {  // Fortran Collision_Matrixmpi(nrow, ncol) is column-major
        DBuf<double> t;
        ALT3_TRY(t.alloc(nrow * ncol));
        if (ctx->cm_fp32)
            hipLaunchKernelGGL(k_rows_to_cols<float>, dim3((unsigned)((nrow * ncol + 255) / 256)), dim3(256), 0,
                               ctx->stream, nrow, ncol, ctx->d_cmf.p, t.p);
        else
            hipLaunchKernelGGL(k_rows_to_cols<double>, dim3((unsigned)((nrow * ncol + 255) / 256)), dim3(256), 0,
                               ctx->stream, nrow, ncol, ctx->d_cm.p, t.p);
        ALT3_HIP(hipGetLastError());
        ALT3_HIP(hipMemcpyAsync(CM, t.p, sizeof(double) * nrow * ncol, hipMemcpyDeviceToHost, ctx->stream));
        ALT3_HIP(hipStreamSynchronize(ctx->stream));
    }
    if (RHS) ALT3_HIP(hipMemcpyAsync(RHS, ctx->d_rhs.p, sizeof(double) * nibz * nb * 3, hipMemcpyDeviceToHost, ctx->stream));
    if (sync || RHS) ALT3_HIP(hipStreamSynchronize(ctx->stream));
    ctx->have_cm = true;
    return ALT3_OK;
}

// CG on a device matrix (dA, db); x is a device vector, overwritten.  rowmajor selects the
// layout of dA (the collision matrix is kept row-major on the device).
// dAf (FP32, row-major) replaces dA when given
__global__ void k_zero_int(int n, int *__restrict__ p)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// cg_quadratic on device buffers with the work space given: work [4n] (grad, conj, aux, v0),
// red [2*256 + 16] (dot-product partials and scalars), state [2] (converged flag, iteration;
// cleared here by a kernel, not a memset)
int cg_device_ws(alt3_ctx *ctx, int n, const double *dA, const double *db, double *dx, int max_iter, int ndig,
                 double *funcmin, int *iters, bool rowmajor, const float *dAf, double *work, double *red, int *state)
{
    auto matvec = [&](const double *xin, double *yout, const int *skip) {
        if (dAf)
            hipLaunchKernelGGL(k_matvec_rows<float>, dim3((n + 3) / 4), dim3(256), 0, ctx->stream, n, n, dAf, xin, yout,
                               skip);
        else if (rowmajor)
            hipLaunchKernelGGL(k_matvec_rows<double>, dim3((n + 3) / 4), dim3(256), 0, ctx->stream, n, n, dA, xin, yout,
                               skip);
        else
            hipLaunchKernelGGL(k_matvec, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, n, dA, xin, yout, skip);
    };
    double *grad = work, *conj = grad + n, *aux = conj + n, *v0 = aux + n;
    const int nblk = 256;
    double *part = red, *sc = part + 2 * nblk;
    hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(64), 0, ctx->stream, 2, state);
    const int gb = (n + 255) / 256;
    const double tol = 1.0 / std::pow(10.0, 2 * ndig);
    matvec(dx, aux, nullptr);
    hipLaunchKernelGGL(k_axpy_init, dim3(gb), dim3(256), 0, ctx->stream, n, aux, db, grad, conj);
    ALT3_HIP(hipGetLastError());
    // iterations run on the device; the host looks at the converged flag every 8 of them
    int iter = max_iter;
    for (int it = 1; it <= max_iter; it++) {
        matvec(conj, aux, state);
        hipLaunchKernelGGL(k_cg_update, dim3(1), dim3(1024), 0, ctx->stream, n, it, tol, dx, grad, conj, aux, v0,
                           state);
        ALT3_HIP(hipGetLastError());
        if (it % 8 == 0 || it == max_iter) {
            int st[2];
            ALT3_HIP(hipMemcpyAsync(st, state, sizeof(st), hipMemcpyDeviceToHost, ctx->stream));
            ALT3_HIP(hipStreamSynchronize(ctx->stream));
            if (st[0]) {
                iter = st[1];
                break;
            }
        }
    }
    double hs[2];
    if (funcmin) {
        matvec(dx, v0, nullptr);
        hipLaunchKernelGGL(k_dot2, dim3(nblk), dim3(256), 0, ctx->stream, n, dx, v0, dx, db, part);
        hipLaunchKernelGGL(k_reduce2, dim3(1), dim3(256), 0, ctx->stream, nblk, part, sc, 0);
        ALT3_HIP(hipMemcpyAsync(hs, sc, 2 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        ALT3_HIP(hipStreamSynchronize(ctx->stream));
        *funcmin = 0.5 * hs[0] - hs[1];
    }
    if (iters) *iters = iter;
    return ALT3_OK;
}

int cg_device(alt3_ctx *ctx, int n, const double *dA, const double *db, double *dx, int max_iter, int ndig,
              double *funcmin, int *iters, bool rowmajor, const float *dAf)
{
    ALT3_TRY(ctx->d_cg.alloc((size_t)4 * n));
    ALT3_TRY(ctx->d_red.alloc(2 * 256 + 16));
    return cg_device_ws(ctx, n, dA, db, dx, max_iter, ndig, funcmin, iters, rowmajor, dAf, ctx->d_cg.p, ctx->d_red.p,
                        ctx->d_flags.p + 4);
}

// x (device, vector layout) = F_RTA of the last alt3_kappa_rta, the CG start of job 1
// (kappa.f90:835-843), and te (device) = its tauinv_eff, both on the context stream
int cg_start_launch(alt3_ctx *ctx, double *x, double *te)
{
    const int nibz = ctx->nibz, nb = ctx->nb, N = nibz * nb * 3;
    hipLaunchKernelGGL(k_frta_to_vec, dim3((N + 255) / 256), dim3(256), 0, ctx->stream, nibz, nb, ctx->d_frta.p, x);
    ALT3_HIP(hipGetLastError());
    ALT3_HIP(hipMemcpyAsync(te, ctx->d_taueff.p, sizeof(double) * nibz * nb, hipMemcpyDeviceToDevice, ctx->stream));
    return ALT3_OK;
}

int kappa_bte_run(alt3_ctx *ctx, double tempk, const double *dx, double *kappa, double *taueff_dev, bool sync)
{
    const int nb = ctx->nb, nibz = ctx->nibz, nk = ctx->nk;
    for (int j = 0; j < nibz; j++)
        for (int s = 0; s < ctx->T.fstar_n[j]; s++)
            if (ctx->T.fstar_n0[48 * j + s] < 0)
                return fail(ALT3_E_MESH, "ERROR: a star vector of an IBZ point is not a kpoint (mesh breaks the lattice symmetry)");
    ALT3_TRY(ctx->d_veloci.alloc((size_t)nibz * nb * 3));
    ALT3_TRY(ctx->d_kpart.alloc((size_t)nibz * nb * 9));
    ALT3_TRY(ctx->d_kappa.alloc((size_t)nb * 9));
    int *derr = ctx->d_flags.p + 3;
    if (sync) ALT3_HIP(hipMemsetAsync(derr, 0, sizeof(int), ctx->stream));
    if (!ctx->have_veloci) {   // star-averaged IBZ velocities depend on the eigen data only
        hipLaunchKernelGGL(k_veloci, dim3((nibz * nb + kStarWave - 1) / kStarWave), dim3(64 * kStarWave), 0, ctx->stream, nibz, nb, ctx->d_vel.p,
                           ctx->d_fstar_n.p, ctx->d_fstar_op.p, ctx->d_fstar_n0.p, ctx->d_opk_inv.p, ctx->d_veloci.p);
        ctx->have_veloci = true;
    }
    double const1 = H_PLANK * C_LIGHT * 100 * C_LIGHT / (nk * ctx->T.volume_r / 1e30);
    double const2 = K_B * tempk * tempk / (H_PLANK * C_LIGHT * 100) / C_LIGHT;
    hipLaunchKernelGGL(k_kappa_bte, dim3((nibz * nb + kStarWave - 1) / kStarWave), dim3(64 * kStarWave), 0, ctx->stream, nibz, nb, dx, ctx->d_veloci.p,
                       ctx->d_eival.p, temp_cm(tempk), ctx->classical, const1, const2, cnst(), ctx->d_star_n.p,
                       ctx->d_star_op.p, ctx->d_star_n0.p, ctx->d_opk.p, ctx->d_kpart.p, taueff_dev, derr);
    hipLaunchKernelGGL(k_kappa_sum, dim3(nb * 9), dim3(64), 0, ctx->stream, nibz, nb, ctx->d_kpart.p, ctx->d_kappa.p, nullptr);
    ALT3_HIP(hipGetLastError());
    if (!sync) return ALT3_OK;   // kappa stays in d_kappa
    int herr = 0;
    ALT3_HIP(hipMemcpyAsync(&herr, derr, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    if (kappa) ALT3_HIP(hipMemcpyAsync(kappa, ctx->d_kappa.p, sizeof(double) * nb * 9, hipMemcpyDeviceToHost, ctx->stream));
    ALT3_HIP(hipStreamSynchronize(ctx->stream));
    if (herr) return fail(ALT3_E_NBE, "ERROR in N_BE: w=<0 (a zero frequency on a star member)");
    return ALT3_OK;
}

// ------------------------------------------------------------------ iterative BTE (job 3)
// The update of update_mpi is linear in F: A(F) = U F with U built once per temperature from
// the R rows of the tau pass and the S rows (k_tau_s), in the collision matrix's layout.  It
// relies on the stars of the irreducible points (kstarsinfbz1) covering every mesh point once,
// with mapibz naming the star a point belongs to: then the op update_mpi applies to F at
// q1 +- q2 (get_kvecopn2) is the op of that point in the star table.
static int update_build(alt3_ctx *ctx, double tempk, const char *who)
{
    ALT3_TRY(tau_select(ctx, tempk, who));
    if (ctx->have_upd && ctx->upd_temp == tempk) return ALT3_OK;
    const int nb = ctx->nb, nibz = ctx->nibz, nk = ctx->nk, ksub = ctx->ibz_hi - ctx->ibz_lo + 1;
    {
        std::vector<int> seen(nk, 0);
        for (int n = 0; n < nibz; n++)
            for (int s = 0; s < ctx->T.star_n[n]; s++) {
                const int q = ctx->T.star_n0[48 * n + s];
                if (q < 0 || q >= nk || ctx->T.mapibz[q] != n || seen[q]++)
                    return fail(ALT3_E_MESH, std::string(who) + ": the stars of the irreducible points do not partition the mesh");
            }
        for (int q = 0; q < nk; q++)
            if (!seen[q]) return fail(ALT3_E_MESH, std::string(who) + ": a mesh point is in no star of an irreducible point");
    }
    const size_t lds = ((size_t)nk + 9 * ctx->nops + 18 * ctx->cm_nsplit) * sizeof(double);
    if (lds > ctx->lds_max) return fail(ALT3_E_ARG, std::string(who) + ": the mesh does not fit in LDS (nk too large)");
    const size_t rstr = (size_t)ksub * nb * nb * nk, nrow = (size_t)ksub * nb * 3, ncol = (size_t)nibz * nb * 3;
    ALT3_TRY(ctx->d_upd.alloc(nrow * ncol));
    const MeshDev md = MeshDev::make(ctx->nc);
    const dim3 gs((unsigned)((nk + 255) / 256), (unsigned)ksub);
    auto launch = [&](auto *P1, auto *P2, auto *R1, auto *R2, auto &S1, auto &S2) {
        using PT = std::remove_const_t<std::remove_pointer_t<decltype(P1)>>;
        int rc = S1.alloc(rstr);
        if (rc == ALT3_OK) rc = S2.alloc(rstr);
        if (rc != ALT3_OK) return rc;
        const unsigned char *rm1 = ctx->d_prow.p, *rm2 = rm1 ? rm1 + ctx->d_prow.n / 2 : nullptr;
        hipLaunchKernelGGL(k_tau_s<PT>, gs, dim3(256), 0, ctx->stream, md, nk, nb, ctx->ibz_lo, ctx->d_mapinv.p,
                           ctx->dist(), P1, P2, S1.p, S2.p, rm1, rm2);
        if (lds > 65536) {
            hipError_t e = hipFuncSetAttribute((const void *)k_update_c<PT>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
            if (e != hipSuccess) return fail(ALT3_E_HIP, std::string("hipFuncSetAttribute: ") + hipGetErrorString(e));
        }
        hipLaunchKernelGGL(k_update_c<PT>, dim3(ksub * nb * nb), dim3(256), lds, ctx->stream, md, nk, nb, ctx->ibz_lo,
                           nibz, ctx->nops, ctx->d_mapinv.p, (const PT *)R1, (const PT *)R2, (const PT *)S1.p,
                           (const PT *)S2.p, star_tasks(ctx), ctx->d_opk.p, ctx->d_upd.p);
        return ALT3_OK;
    };
    if (ctx->r_fp32) {
        ctx->d_s1.release();
        ctx->d_s2.release();
        ALT3_TRY(launch((const float *)ctx->d_p1f.p, (const float *)ctx->d_p2f.p, ctx->r1f(), ctx->r2f(), ctx->d_s1f, ctx->d_s2f));
    } else {
        ctx->d_s1f.release();
        ctx->d_s2f.release();
        ALT3_TRY(launch((const double *)ctx->d_p1.p, (const double *)ctx->d_p2.p, ctx->r1(), ctx->r2(), ctx->d_s1, ctx->d_s2));
    }
    ALT3_HIP(hipGetLastError());
    if (!ctx->giso.empty()) {
        const int nl = nibz - 1, op = ctx->T.star_op[48 * nl + ctx->T.star_n[nl] - 1];
        hipLaunchKernelGGL(k_update_iso, dim3((unsigned)((nrow + 255) / 256)), dim3(256), 0, ctx->stream, ksub, nb,
                           nibz, ctx->ibz_lo, ctx->d_piso.p, ctx->d_opk.p + 9 * op, ctx->d_upd.p);
        ALT3_HIP(hipGetLastError());
    }
    ALT3_HIP(hipStreamSynchronize(ctx->stream));
    ctx->have_upd = true;
    ctx->upd_temp = tempk;
    return ALT3_OK;
}

int update_run(alt3_ctx *ctx, double tempk, const double *FFF, const double *FRTA, double *F2)
{
    ALT3_TRY(update_build(ctx, tempk, "alt3_update"));
    const int nb = ctx->nb, nibz = ctx->nibz, ksub = ctx->ibz_hi - ctx->ibz_lo + 1;
    const size_t N = (size_t)nibz * nb * 3, nrow = (size_t)ksub * nb * 3;
    std::vector<double> x(N);
    for (int i = 0; i < nibz; i++)   // Fortran (nibz,nb,3) -> vector layout
        for (int l = 0; l < nb; l++)
            for (int a = 0; a < 3; a++) x[3 * ((size_t)i * nb + l) + a] = FFF[i + (size_t)nibz * (l + (size_t)nb * a)];
    ALT3_TRY(ctx->d_iter.reserve(2 * N + nrow + N));
    double *dx = ctx->d_iter.p, *dfr = dx + N, *dA = dfr + N, *dF2 = dA + nrow;
    ALT3_HIP(hipMemcpyAsync(dx, x.data(), sizeof(double) * N, hipMemcpyHostToDevice, ctx->stream));
    ALT3_HIP(hipMemcpyAsync(dfr, FRTA, sizeof(double) * N, hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(k_matvec_rows<double>, dim3((unsigned)((nrow + 3) / 4)), dim3(256), 0, ctx->stream, (int)nrow, (int)N,
                       ctx->d_upd.p, dx, dA, nullptr);
    hipLaunchKernelGGL(k_iter_mix, dim3((unsigned)((nrow + 255) / 256)), dim3(256), 0, ctx->stream, ksub, nibz, nb,
                       ctx->ibz_lo, 0.0, dfr, ctx->q(), dA, nullptr, dF2, nullptr);
    ALT3_HIP(hipGetLastError());
    ALT3_HIP(hipMemcpyAsync(F2, dF2, sizeof(double) * nrow, hipMemcpyDeviceToHost, ctx->stream));
    ALT3_HIP(hipStreamSynchronize(ctx->stream));
    return ALT3_OK;
}

int kappa_iter_run(alt3_ctx *ctx, double tempk, int max_iter, double mix, double tol, double *kappa, double *khist,
                   double *tehist, double *FFF, int *iters)
{
    if (max_iter < 1) return fail(ALT3_E_ARG, "max_iter must be at least 1");
    ALT3_TRY(need_full_ibz(ctx, tempk, "alt3_kappa_iter"));
    ALT3_TRY(kappa_rta_run(ctx, tempk, nullptr, nullptr, nullptr, true));   // F_RTA, kappa_RTA = first kappa_old
    ALT3_TRY(update_build(ctx, tempk, "alt3_kappa_iter"));
    const int nb = ctx->nb, nibz = ctx->nibz, nk = ctx->nk, N = nibz * nb * 3;
    ALT3_TRY(ctx->d_iter.reserve(2 * (size_t)N + (9 * (size_t)nb + 1) * (1 + (size_t)max_iter)));
    double *x = ctx->d_iter.p, *y = x + N, *kold = y + N, *hist = kold + 9 * nb;
    double *te_hist = hist + 9 * (size_t)nb * max_iter, *te_last = te_hist + max_iter;
    int *state = ctx->d_flags.p + 6, *derr = ctx->d_flags.p + 8;   // converged, iterations; N_BE error
    ALT3_HIP(hipMemsetAsync(state, 0, 3 * sizeof(int), ctx->stream));
    hipLaunchKernelGGL(k_frta_to_vec, dim3((N + 255) / 256), dim3(256), 0, ctx->stream, nibz, nb, ctx->d_frta.p, x);
    ALT3_HIP(hipMemcpyAsync(kold, ctx->d_kappa.p, sizeof(double) * 9 * nb, hipMemcpyDeviceToDevice, ctx->stream));
    const double const1 = H_PLANK * C_LIGHT * 100 * C_LIGHT / (nk * ctx->T.volume_r / 1e30);
    const double const2 = K_B * tempk * tempk / (H_PLANK * C_LIGHT * 100) / C_LIGHT;
    // iterations run on the device; the host looks at the converged flag every 8 of them
    for (int it = 0; it < max_iter; it++) {
        hipLaunchKernelGGL(k_matvec_rows<double>, dim3((N + 3) / 4), dim3(256), 0, ctx->stream, N, N, ctx->d_upd.p, x, y, state);
        hipLaunchKernelGGL(k_iter_mix, dim3((N + 255) / 256), dim3(256), 0, ctx->stream, nibz, nibz, nb, 0, mix,
                           ctx->d_frta.p, ctx->q(), y, x, nullptr, state);
        hipLaunchKernelGGL(k_mode_kappa, dim3((nibz * nb + kStarWave - 1) / kStarWave), dim3(64 * kStarWave), 0, ctx->stream, nibz, nb,
                           ctx->d_mapinv.p, x, ctx->d_eival.p, ctx->d_vel.p, temp_cm(tempk), ctx->classical, const1,
                           const2, cnst(), ctx->d_star_n.p, ctx->d_star_op.p, ctx->d_star_n0.p, ctx->d_opk.p,
                           ctx->d_kpart.p, te_last, derr, state);
        hipLaunchKernelGGL(k_kappa_sum, dim3(nb * 9), dim3(64), 0, ctx->stream, nibz, nb, ctx->d_kpart.p, ctx->d_kappa.p,
                           state);
        hipLaunchKernelGGL(k_iter_check, dim3(1), dim3(64), 0, ctx->stream, nb, it, tol, ctx->d_kappa.p, kold, hist,
                           te_last, te_hist, state);
        ALT3_HIP(hipGetLastError());
        if (it % 8 == 7 || it == max_iter - 1) {
            int st[3];
            ALT3_HIP(hipMemcpyAsync(st, state, sizeof(st), hipMemcpyDeviceToHost, ctx->stream));
            ALT3_HIP(hipStreamSynchronize(ctx->stream));
            if (st[2]) return fail(ALT3_E_NBE, "ERROR in N_BE: w=<0 (a zero frequency on a star member)");
            if (st[0]) break;
        }
    }
    int st[3];
    ALT3_HIP(hipMemcpyAsync(st, state, sizeof(st), hipMemcpyDeviceToHost, ctx->stream));
    ALT3_HIP(hipStreamSynchronize(ctx->stream));
    if (st[2]) return fail(ALT3_E_NBE, "ERROR in N_BE: w=<0 (a zero frequency on a star member)");
    const int done = st[1];
    if (kappa) ALT3_HIP(hipMemcpyAsync(kappa, ctx->d_kappa.p, sizeof(double) * 9 * nb, hipMemcpyDeviceToHost, ctx->stream));
    if (khist && done > 0)
        ALT3_HIP(hipMemcpyAsync(khist, hist, sizeof(double) * 9 * nb * done, hipMemcpyDeviceToHost, ctx->stream));
    if (tehist && done > 0)
        ALT3_HIP(hipMemcpyAsync(tehist, te_hist, sizeof(double) * done, hipMemcpyDeviceToHost, ctx->stream));
    std::vector<double> xv;
    if (FFF) {
        xv.resize(N);
        ALT3_HIP(hipMemcpyAsync(xv.data(), x, sizeof(double) * N, hipMemcpyDeviceToHost, ctx->stream));
    }
    ALT3_HIP(hipStreamSynchronize(ctx->stream));
    if (FFF)
        for (int i = 0; i < nibz; i++)
            for (int l = 0; l < nb; l++)
                for (int a = 0; a < 3; a++) FFF[i + (size_t)nibz * (l + (size_t)nb * a)] = xv[3 * ((size_t)i * nb + l) + a];
    if (iters) *iters = done;
    return ALT3_OK;
}

}  // namespace alt3
