// Harmonic phonons on the q-mesh, FP64 on the GPU.
//
// Replaces get_frequencies (THERMACOND/rest.f90:2230-2405): for every mesh point the
// dynamical matrix (set_dynamical_matrix, rest.f90:445-556) is built and diagonalised at q
// and at q +- dq e_a (finitedif_vel, rest.f90:2587-2631, dq = REAL*4 0.001), i.e. 7
// Hermitian eigenproblems per q.  A group of lanes owns one matrix in LDS and runs cyclic
// complex Jacobi rotations (lane-parallel row/column updates); Jacobi converges to full
// FP64 accuracy on the small (3*natoms)^2 matrices, including degenerate branches.
// Eigenvalue ordering follows zheev (ascending), then the reference's post-processing:
// |lambda|, eigenvectors multiplied by i when "pure imaginary", the second sort of
// get_frequencies that permutes eigenvalues/eigenvectors but not velocities, and
// subst_eivecs (rest.f90:3262-3293).
#include <cmath>

#include "alt3_impl.h"

namespace {

struct alignas(16) zd {
    double r, i;
};
__device__ __forceinline__ zd zmul(zd a, zd b) { return {a.r * b.r - a.i * b.i, a.r * b.i + a.i * b.r}; }

constexpr int kDq = 7;  // q, q+dx, q-dx, q+dy, q-dy, q+dz, q-dz

// nonanal (rest.f90:3517-3578), entry e = (row, col) of D at kq: the real term
//   zag(i) zbg(j) / (qeq eps_scale) / om0 / sqrt(ma mb) exp(-q^2/rho^2)
// with zag(i) = sum_k q_k charge_a(k,i), qeq = q.epsil.q, evaluated in the reference's order;
// get_freq shifts q = 0 to (1e-10, 1e-10, 1e-10) first (rest.f90:4342-4346).  zm [natoms0][10]
// holds charge(i,j) row-major and the mass of each atom.
__device__ __forceinline__ double nonanal_term(const alt3::BornDev &b, const double *__restrict__ zm, const double *q,
                                               double qeq, int nb, int e)
{
    const int r = e / nb, c = e - r * nb, na = r / 3, i = r - 3 * na, nbb = c / 3, j = c - 3 * nbb;
    const double *za = zm + 10 * na, *zb = zm + 10 * nbb;
    const double zag = q[0] * za[i] + q[1] * za[3 + i] + q[2] * za[6 + i];
    const double zbg = q[0] * zb[j] + q[1] * zb[3 + j] + q[2] * zb[6 + j];
    const double q2 = q[0] * q[0] + q[1] * q[1] + q[2] * q[2];
    return zag * zbg / (qeq * b.eps_scale) / b.om0 / sqrt(za[9] * zb[9]) * exp(-q2 / b.rho2);
}

__device__ __forceinline__ void born_q(const double *kq, double *q, const alt3::BornDev &b, double &qeq)
{
    const bool z = kq[0] == 0.0 && kq[1] == 0.0 && kq[2] == 0.0;
    for (int a = 0; a < 3; a++) q[a] = z ? kq[a] + 1.0e-10 : kq[a];
    const double *E = b.epsil;
    qeq = q[0] * (E[0] * q[0] + E[1] * q[1] + E[2] * q[2]) + q[1] * (E[3] * q[0] + E[4] * q[1] + E[5] * q[2]) +
          q[2] * (E[6] * q[0] + E[7] * q[1] + E[8] * q[2]);
}

// entries: e = row*nb + col; per entry a CSR range of FC2 terms (reference term order)
__global__ void __launch_bounds__(64) k_eig_solve(int nk, int nvar, int nb, const double *__restrict__ kpc,
                                                  const int *__restrict__ ent_ptr,
                                                  const double *__restrict__ term_amp,
                                                  const double *__restrict__ term_rr, double dq,
                                                  double *__restrict__ evals7,   // [nk][7][nb]
                                                  double *__restrict__ evec0,    // [nk][nb col][nb comp] complex
                                                  alt3::BornDev born, const double *__restrict__ bzm)
{
    extern __shared__ double lds[];
    const int m = blockIdx.x;  // matrix id = q*nvar + variant (nvar 7: q and q +- dq e_a; 1: q only)
    const int q = m / nvar, var = m % nvar;
    if (q >= nk) return;
    const int lane = threadIdx.x;
    zd *A = reinterpret_cast<zd *>(lds);  // [nb][nb] row-major
    zd *V = A + nb * nb;                  // [nb][nb] row-major, V[k][col]
    double *red = reinterpret_cast<double *>(V + nb * nb);  // [64]
    __shared__ double s_rot[5];           // c, s, cos(phi), sin(phi), skip
    __shared__ int s_ord[64];

    double kq[3] = {kpc[3 * q], kpc[3 * q + 1], kpc[3 * q + 2]};
    if (var > 0) {
        int a = (var - 1) / 2;
        kq[a] = (var & 1) ? kq[a] + dq : kq[a] - dq;
    }
    const int nn = nb * nb;
    // 1. D(q): each entry sums its own terms in reference order
    for (int e = lane; e < nn; e += 64) {
        double re = 0, im = 0;
        for (int t = ent_ptr[e]; t < ent_ptr[e + 1]; t++) {
            double ph = kq[0] * term_rr[3 * t] + kq[1] * term_rr[3 * t + 1] + kq[2] * term_rr[3 * t + 2];
            double sn, cs;
            sincos(ph, &sn, &cs);
            re = re + term_amp[t] * cs;
            im = im + term_amp[t] * sn;
        }
        A[e] = {re, im};
        V[e] = {(e / nb == e % nb) ? 1.0 : 0.0, 0.0};
    }
    __syncthreads();
    // 2. Hermitisation exactly as rest.f90:514-543 (note: column nb is left untouched)
    for (int e = lane; e < nn; e += 64) {
        int t = e / nb, j = e % nb;
        if (t == j) {
            zd d = A[e];
            if (fabs(d.i) > 9e-4 * fabs(d.r)) A[e] = {d.r, 0.0};
        }
    }
    __syncthreads();
    for (int e = lane; e < nn; e += 64) {
        int t = e / nb, j = e % nb;
        if (j > t && j < nb - 1) {
            zd a = A[t * nb + j], b = A[j * nb + t];
            double mi = (a.i - b.i) / 2, mj = (a.r + b.r) / 2;
            A[t * nb + j] = {mj, mi};
            A[j * nb + t] = {mj, -mi};
        }
    }
    __syncthreads();
    {
        double si = 0, sr = 0;
        for (int e = lane; e < nn; e += 64) {
            si += fabs(A[e].i);
            sr += fabs(A[e].r);
        }
        red[lane] = si;
        __syncthreads();
        if (lane == 0) {
            double s = 0;
            for (int l = 0; l < 64; l++) s += red[l];
            s_rot[0] = s;
        }
        __syncthreads();
        red[lane] = sr;
        __syncthreads();
        if (lane == 0) {
            double s = 0;
            for (int l = 0; l < 64; l++) s += red[l];
            s_rot[1] = s;
        }
        __syncthreads();
        bool realify = (s_rot[0] / nn) < (s_rot[1] / nn) * 1e-10;
        __syncthreads();
        if (realify)
            for (int e = lane; e < nn; e += 64) A[e] = {A[e].r, 0.0};
        __syncthreads();
    }
    if (born.on) {   // nonanal after set_dynamical_matrix (rest.f90:4348)
        double qb[3], qeq;
        born_q(kq, qb, born, qeq);
        for (int e = lane; e < nn; e += 64) A[e].r = A[e].r + nonanal_term(born, bzm, qb, qeq, nb, e);
        __syncthreads();
    }
    // zheev('U') reads the upper triangle and a real diagonal: make A exactly Hermitian
    for (int e = lane; e < nn; e += 64) {
        int r = e / nb, c = e % nb;
        if (r > c) A[e] = {A[c * nb + r].r, -A[c * nb + r].i};
        if (r == c) A[e] = {A[e].r, 0.0};
    }
    __syncthreads();
    // 3. cyclic complex Jacobi
    for (int sweep = 0; sweep < 40; sweep++) {
        double off = 0, dia = 0;
        for (int e = lane; e < nn; e += 64) {
            int r = e / nb, c = e % nb;
            double v2 = A[e].r * A[e].r + A[e].i * A[e].i;
            if (r != c) off += v2; else dia += v2;
        }
        red[lane] = off;
        __syncthreads();
        if (lane == 0) {
            double s = 0;
            for (int l = 0; l < 64; l++) s += red[l];
            s_rot[0] = s;
        }
        __syncthreads();
        red[lane] = dia;
        __syncthreads();
        if (lane == 0) {
            double s = 0;
            for (int l = 0; l < 64; l++) s += red[l];
            s_rot[1] = s;
        }
        __syncthreads();
        bool done = !(s_rot[0] > 1e-32 * s_rot[1]) || s_rot[0] == 0.0;
        __syncthreads();
        if (done) break;
        for (int p = 0; p < nb - 1; p++)
            for (int qq = p + 1; qq < nb; qq++) {
                if (lane == 0) {
                    zd apq = A[p * nb + qq];
                    double mag = sqrt(apq.r * apq.r + apq.i * apq.i);
                    double app = A[p * nb + p].r, aqq = A[qq * nb + qq].r;
                    if (mag <= 1e-300 || mag * mag <= 1e-36 * (app * app + aqq * aqq)) {
                        s_rot[4] = 1.0;
                        A[p * nb + qq] = {0.0, 0.0};
                        A[qq * nb + p] = {0.0, 0.0};
                    } else {
                        double tau = (aqq - app) / (2.0 * mag);
                        double t = (tau >= 0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
                        double c = 1.0 / sqrt(1.0 + t * t);
                        s_rot[0] = c;
                        s_rot[1] = t * c;
                        s_rot[2] = apq.r / mag;
                        s_rot[3] = apq.i / mag;
                        s_rot[4] = 0.0;
                    }
                }
                __syncthreads();
                if (s_rot[4] == 0.0) {
                    const double c = s_rot[0], s = s_rot[1];
                    const zd em = {s_rot[2], -s_rot[3]};  // e^{-i phi}
                    const zd ep = {s_rot[2], s_rot[3]};   // e^{+i phi}
                    if (lane < nb) {  // columns of A and V
                        int k = lane;
                        zd x = A[k * nb + p], y = zmul(em, A[k * nb + qq]);
                        A[k * nb + p] = {c * x.r - s * y.r, c * x.i - s * y.i};
                        A[k * nb + qq] = {s * x.r + c * y.r, s * x.i + c * y.i};
                        if (var == 0) {   // the displaced matrices give eigenvalues only
                            zd vx = V[k * nb + p], vy = zmul(em, V[k * nb + qq]);
                            V[k * nb + p] = {c * vx.r - s * vy.r, c * vx.i - s * vy.i};
                            V[k * nb + qq] = {s * vx.r + c * vy.r, s * vx.i + c * vy.i};
                        }
                    }
                    __syncthreads();
                    if (lane < nb) {  // rows of A
                        int k = lane;
                        zd x = A[p * nb + k], y = zmul(ep, A[qq * nb + k]);
                        A[p * nb + k] = {c * x.r - s * y.r, c * x.i - s * y.i};
                        A[qq * nb + k] = {s * x.r + c * y.r, s * x.i + c * y.i};
                    }
                    __syncthreads();
                    if (lane == 0) {
                        A[p * nb + qq] = {0.0, 0.0};
                        A[qq * nb + p] = {0.0, 0.0};
                        A[p * nb + p].i = 0.0;
                        A[qq * nb + qq].i = 0.0;
                    }
                }
                __syncthreads();
            }
    }
    // 4. ascending order of the signed eigenvalues (zheev), ties keep index order
    if (lane == 0) {
        for (int k = 0; k < nb; k++) s_ord[k] = k;
        for (int a = 1; a < nb; a++) {
            int x = s_ord[a];
            double vx = A[x * nb + x].r;
            int b = a - 1;
            while (b >= 0 && A[s_ord[b] * nb + s_ord[b]].r > vx) {
                s_ord[b + 1] = s_ord[b];
                b--;
            }
            s_ord[b + 1] = x;
        }
    }
    __syncthreads();
    if (lane < nb) {
        int src = s_ord[lane];
        evals7[((size_t)q * nvar + var) * nb + lane] = fabs(A[src * nb + src].r);
    }
    if (var == 0) {
        // column j (sorted) of V; "pure imaginary" columns (sum |Re| < 1e-3) times i
        for (int j = lane; j < nb; j += 64) {
            int src = s_ord[j];
            double sre = 0;
            for (int k = 0; k < nb; k++) sre += fabs(V[k * nb + src].r);
            bool rot = sre < 1e-3;
            for (int k = 0; k < nb; k++) {
                zd v = V[k * nb + src];
                if (rot) v = {-v.i, v.r};
                size_t o = 2 * (((size_t)q * nb + j) * nb + k);
                evec0[o] = v.r;
                evec0[o + 1] = v.i;
            }
        }
    }
}

// The same solve with G lanes per matrix and 64/G matrices per wavefront, for cells of up to
// 32 branches (G >= nb).  Every matrix goes through the same operations in the same order as
// in k_eig_solve (the norm sums are taken by the group's first lane in k_eig_solve's
// summation order), so the results are bit-identical; a wavefront now carries up to eight
// rotation chains at once instead of one.
// PAR (G = 64, one matrix per wavefront, cells of more than 8 branches): the sweeps in the
// round-robin (circle) order of parallel Jacobi instead of the cyclic row order: each of the
// m - 1 rounds of a sweep (m = nb rounded up to even) rotates m/2 disjoint pairs (p, q) at once
// -- their angles from the same A, then A J for all of them, then J^H (A J) -- so a sweep takes
// m - 1 dependent steps where the cyclic order takes nb (nb - 1)/2.  The rotation formulas are
// those of the cyclic order; the eigenvalues agree to round-off, the eigenvectors inside a
// degenerate subspace are another orthonormal basis of it.
// Variants var0 .. var0 + nv - 1 of every point (nvar: the layout of evals7).  warm (PAR, the
// displaced variants only): the eigenvectors of the central solve, evec0 as written by the
// variant-0 launch; the displaced matrix is first turned into W^H D(q +- dq) W with W those
// eigenvectors (nearly diagonal: its off-diagonal entries are O(dq)), so that the Jacobi
// sweeps start close to convergence.  A unitary similarity leaves the eigenvalues, which is
// all a displaced solve gives.
template <int G, bool PAR = false>
__global__ void __launch_bounds__(64) k_eig_group(int nk, int nvar, int nb, const double *__restrict__ kpc,
                                                  const int *__restrict__ ent_ptr,
                                                  const double *__restrict__ term_amp,
                                                  const double *__restrict__ term_rr, double dq,
                                                  double *__restrict__ evals7, double *__restrict__ evec0,
                                                  alt3::BornDev born, const double *__restrict__ bzm, int var0 = 0,
                                                  int nv = 0, const double *__restrict__ warm = nullptr)
{
    constexpr int M = 64 / G;
    extern __shared__ double lds[];
    const int lane = threadIdx.x, g = lane / G, sub = lane % G;
    const int nn = nb * nb;
    // row stride: nb, or nb + 1 in the parallel order, whose column updates have consecutive
    // lanes on consecutive rows (a stride of nb complex doubles puts them on two LDS banks)
    const int ld = PAR ? nb + 1 : nb;
    auto EI = [&](int e) { return PAR ? e + e / nb : e; };       // linear index e = r nb + c -> r ld + c
    zd *A = reinterpret_cast<zd *>(lds) + (size_t)g * 2 * nb * ld;   // [nb][ld] row-major
    zd *V = A + nb * ld;                                             // [nb][ld], V[k][col]
    __shared__ double s_rot[M][5];                               // c, s, cos(phi), sin(phi), skip
    __shared__ double s_sum[M][2];
    __shared__ int s_ord[M][G];
    if (nv <= 0) nv = nvar;
    const int mat = blockIdx.x * M + g;
    const bool valid = mat < nk * nv;
    const int q = valid ? mat / nv : 0, var = valid ? var0 + mat % nv : 0;
    // norm sum of k_eig_solve: lane l of 64 sums entries l, l+64, ..., then the lane sums in order
    auto sum2 = [&](auto f, double &s0, double &s1) {
        s0 = 0;
        s1 = 0;
        for (int l = 0; l < 64 && l < nn; l++) {
            double p0 = 0, p1 = 0;
            for (int e = l; e < nn; e += 64) f(e, p0, p1);
            s0 += p0;
            s1 += p1;
        }
    };

    double kq[3] = {kpc[3 * q], kpc[3 * q + 1], kpc[3 * q + 2]};
    if (var > 0) {
        int a = (var - 1) / 2;
        kq[a] = (var & 1) ? kq[a] + dq : kq[a] - dq;
    }
    // 1. D(q)
    for (int e = sub; e < nn; e += G) {
        double re = 0, im = 0;
        for (int t = ent_ptr[e]; t < ent_ptr[e + 1]; t++) {
            double ph = kq[0] * term_rr[3 * t] + kq[1] * term_rr[3 * t + 1] + kq[2] * term_rr[3 * t + 2];
            double sn, cs;
            sincos(ph, &sn, &cs);
            re = re + term_amp[t] * cs;
            im = im + term_amp[t] * sn;
        }
        A[EI(e)] = {re, im};
        V[EI(e)] = {(e / nb == e % nb) ? 1.0 : 0.0, 0.0};
    }
    __syncthreads();
    // 2. Hermitisation as rest.f90:514-543 (column nb untouched)
    for (int e = sub; e < nn; e += G) {
        int t = e / nb, j = e % nb;
        if (t == j) {
            zd d = A[EI(e)];
            if (fabs(d.i) > 9e-4 * fabs(d.r)) A[EI(e)] = {d.r, 0.0};
        }
    }
    __syncthreads();
    for (int e = sub; e < nn; e += G) {
        int t = e / nb, j = e % nb;
        if (j > t && j < nb - 1) {
            zd a = A[t * ld + j], b = A[j * ld + t];
            double mi = (a.i - b.i) / 2, mj = (a.r + b.r) / 2;
            A[t * ld + j] = {mj, mi};
            A[j * ld + t] = {mj, -mi};
        }
    }
    __syncthreads();
    if (sub == 0) {
        double si, sr;
        sum2([&](int e, double &a, double &b) { a += fabs(A[EI(e)].i); b += fabs(A[EI(e)].r); }, si, sr);
        s_sum[g][0] = si;
        s_sum[g][1] = sr;
    }
    __syncthreads();
    {
        bool realify = (s_sum[g][0] / nn) < (s_sum[g][1] / nn) * 1e-10;
        if (realify)
            for (int e = sub; e < nn; e += G) A[EI(e)] = {A[EI(e)].r, 0.0};
    }
    __syncthreads();
    if (born.on) {   // nonanal after set_dynamical_matrix (rest.f90:4348)
        double qb[3], qeq;
        born_q(kq, qb, born, qeq);
        for (int e = sub; e < nn; e += G) A[EI(e)].r = A[EI(e)].r + nonanal_term(born, bzm, qb, qeq, nb, e);
        __syncthreads();
    }
    for (int e = sub; e < nn; e += G) {
        int r = e / nb, c = e % nb;
        if (r > c) A[EI(e)] = {A[c * ld + r].r, -A[c * ld + r].i};
        if (r == c) A[EI(e)] = {A[EI(e)].r, 0.0};
    }
    __syncthreads();
    if constexpr (PAR) {
        static_assert(G == 64, "the parallel order runs one matrix per wavefront");
        if (warm != nullptr && var > 0) {   // A <- W^H A W (uniform over the wavefront)
            constexpr int EPL = 16;         // entries per lane, nb <= 32
            for (int e = sub; e < nn; e += 64) {   // W[k][c] = component k of eigenvector c, into V
                const int r = e / nb, c = e % nb;
                const double *w = warm + 2 * (((size_t)q * nb + c) * nb + r);
                V[r * ld + c] = {w[0], w[1]};
            }
            __syncthreads();
            zd t[EPL];
#pragma unroll
            for (int i = 0; i < EPL; i++) {   // T = A W
                const int e = sub + 64 * i;
                t[i] = {0.0, 0.0};
                if (e < nn) {
                    const int r = e / nb, c = e % nb;
                    double ar = 0, ai = 0;
                    for (int k = 0; k < nb; k++) {
                        const zd a = A[r * ld + k], w = V[k * ld + c];
                        ar += a.r * w.r - a.i * w.i;
                        ai += a.r * w.i + a.i * w.r;
                    }
                    t[i] = {ar, ai};
                }
            }
            __syncthreads();
#pragma unroll
            for (int i = 0; i < EPL; i++) {
                const int e = sub + 64 * i;
                if (e < nn) A[EI(e)] = t[i];
            }
            __syncthreads();
#pragma unroll
            for (int i = 0; i < EPL; i++) {   // W^H T
                const int e = sub + 64 * i;
                if (e < nn) {
                    const int r = e / nb, c = e % nb;
                    double ar = 0, ai = 0;
                    for (int k = 0; k < nb; k++) {
                        const zd w = V[k * ld + r], a = A[k * ld + c];
                        ar += w.r * a.r + w.i * a.i;
                        ai += w.r * a.i - w.i * a.r;
                    }
                    t[i] = {ar, ai};
                }
            }
            __syncthreads();
#pragma unroll
            for (int i = 0; i < EPL; i++) {
                const int e = sub + 64 * i;
                if (e < nn) A[EI(e)] = t[i];
            }
            __syncthreads();
            for (int e = sub; e < nn; e += 64) {   // Hermitian to the last bit: the upper triangle's mean
                const int r = e / nb, c = e % nb;
                if (r < c) {
                    const zd a = A[r * ld + c], b = A[c * ld + r];
                    const zd mid = {0.5 * (a.r + b.r), 0.5 * (a.i - b.i)};
                    A[r * ld + c] = mid;
                    A[c * ld + r] = {mid.r, -mid.i};
                } else if (r == c) {
                    A[r * ld + c].i = 0.0;
                }
            }
            __syncthreads();
        }
        const int m = (nb + 1) & ~1, np = m / 2;
        __shared__ double s_par[32][5];   // per pair: c, s, cos(phi), sin(phi), skip
        __shared__ int s_pq[32][2];
        for (int sweep = 0; valid && sweep < 40; sweep++) {
            double off = 0, dia = 0;
            for (int e = sub; e < nn; e += 64) {
                const int r = e / nb, c = e % nb;
                const double v2 = A[EI(e)].r * A[EI(e)].r + A[EI(e)].i * A[EI(e)].i;
                if (r != c) off += v2; else dia += v2;
            }
            for (int o = 32; o > 0; o >>= 1) {
                off += __shfl_xor(off, o, 64);
                dia += __shfl_xor(dia, o, 64);
            }
            off = __shfl(off, 0, 64);   // lane 0's sums: one decision for the wavefront
            dia = __shfl(dia, 0, 64);
            // the cyclic order's 1e-32 is below the round-off floor of a matrix with degenerate
            // eigenvalues (a 45-degree rotation of a degenerate pair leaves |a_pq| ~ eps |a_pp|),
            // which then runs to the sweep limit: stop at 1e-30, and flush pairs below it
            if (!(off > 1e-30 * dia) || off == 0.0) break;
            for (int rd = 0; rd < m - 1; rd++) {
                if (sub < np) {
                    const int a = sub == 0 ? rd : (rd + sub) % (m - 1);
                    const int b = sub == 0 ? m - 1 : (rd - sub + m - 1) % (m - 1);
                    const int p = min(a, b), qq = max(a, b);
                    s_pq[sub][0] = p;
                    s_pq[sub][1] = qq;
                    double cc = 1, ss = 0, cr = 1, ci = 0, skip = 1;
                    if (qq < nb) {   // with nb odd, the pair of the phantom index nb is idle
                        const zd apq = A[p * ld + qq];
                        const double mag = sqrt(apq.r * apq.r + apq.i * apq.i);
                        const double app = A[p * ld + p].r, aqq = A[qq * ld + qq].r;
                        if (mag <= 1e-300 || mag * mag <= 1e-30 * (app * app + aqq * aqq)) {
                            A[p * ld + qq] = {0.0, 0.0};
                            A[qq * ld + p] = {0.0, 0.0};
                        } else {
                            const double tau = (aqq - app) / (2.0 * mag);
                            const double t = (tau >= 0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
                            cc = 1.0 / sqrt(1.0 + t * t);
                            ss = t * cc;
                            cr = apq.r / mag;
                            ci = apq.i / mag;
                            skip = 0;
                        }
                    }
                    s_par[sub][0] = cc;
                    s_par[sub][1] = ss;
                    s_par[sub][2] = cr;
                    s_par[sub][3] = ci;
                    s_par[sub][4] = skip;
                }
                __syncthreads();
                for (int x = sub; x < np * nb; x += 64) {   // columns p, q of A and V: A <- A J
                    const int kp = x / nb, k = x - kp * nb;
                    if (s_par[kp][4] != 0.0) continue;
                    const int p = s_pq[kp][0], qq = s_pq[kp][1];
                    const double c = s_par[kp][0], sn = s_par[kp][1];
                    const zd em = {s_par[kp][2], -s_par[kp][3]};
                    zd x0 = A[k * ld + p], y0 = zmul(em, A[k * ld + qq]);
                    A[k * ld + p] = {c * x0.r - sn * y0.r, c * x0.i - sn * y0.i};
                    A[k * ld + qq] = {sn * x0.r + c * y0.r, sn * x0.i + c * y0.i};
                    if (var == 0) {   // the displaced matrices give eigenvalues only
                        zd vx = V[k * ld + p], vy = zmul(em, V[k * ld + qq]);
                        V[k * ld + p] = {c * vx.r - sn * vy.r, c * vx.i - sn * vy.i};
                        V[k * ld + qq] = {sn * vx.r + c * vy.r, sn * vx.i + c * vy.i};
                    }
                }
                __syncthreads();
                for (int x = sub; x < np * nb; x += 64) {   // rows p, q of A: A <- J^H A
                    const int kp = x / nb, k = x - kp * nb;
                    if (s_par[kp][4] != 0.0) continue;
                    const int p = s_pq[kp][0], qq = s_pq[kp][1];
                    const double c = s_par[kp][0], sn = s_par[kp][1];
                    const zd ep = {s_par[kp][2], s_par[kp][3]};
                    zd x0 = A[p * ld + k], y0 = zmul(ep, A[qq * ld + k]);
                    A[p * ld + k] = {c * x0.r - sn * y0.r, c * x0.i - sn * y0.i};
                    A[qq * ld + k] = {sn * x0.r + c * y0.r, sn * x0.i + c * y0.i};
                }
                __syncthreads();
                // the rotated entries are zero in exact arithmetic; the next round's angles read
                // other entries (each pair comes once per sweep), and its updates follow a barrier
                if (sub < np && s_par[sub][4] == 0.0) {
                    const int p = s_pq[sub][0], qq = s_pq[sub][1];
                    A[p * ld + qq] = {0.0, 0.0};
                    A[qq * ld + p] = {0.0, 0.0};
                    A[p * ld + p].i = 0.0;
                    A[qq * ld + qq].i = 0.0;
                }
            }
            __syncthreads();
        }
    } else {
    // 3. cyclic complex Jacobi; a finished matrix idles while the others of its wavefront go on
    bool done = !valid;
    for (int sweep = 0; sweep < 40; sweep++) {
        if (sub == 0 && !done) {
            double off, dia;
            sum2([&](int e, double &o, double &d) {
                int r = e / nb, c = e % nb;
                double v2 = A[EI(e)].r * A[EI(e)].r + A[EI(e)].i * A[EI(e)].i;
                if (r != c) o += v2; else d += v2;
            }, off, dia);
            s_sum[g][0] = (!(off > 1e-32 * dia) || off == 0.0) ? 1.0 : 0.0;
        }
        __syncthreads();
        done = done || s_sum[g][0] != 0.0;
        if (__ballot(!done) == 0) break;   // one wavefront per block: uniform
        for (int p = 0; p < nb - 1; p++)
            for (int qq = p + 1; qq < nb; qq++) {
                if (sub == 0 && !done) {
                    zd apq = A[p * ld + qq];
                    double mag = sqrt(apq.r * apq.r + apq.i * apq.i);
                    double app = A[p * ld + p].r, aqq = A[qq * ld + qq].r;
                    if (mag <= 1e-300 || mag * mag <= 1e-36 * (app * app + aqq * aqq)) {
                        s_rot[g][4] = 1.0;
                        A[p * ld + qq] = {0.0, 0.0};
                        A[qq * ld + p] = {0.0, 0.0};
                    } else {
                        double tau = (aqq - app) / (2.0 * mag);
                        double t = (tau >= 0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
                        double c = 1.0 / sqrt(1.0 + t * t);
                        s_rot[g][0] = c;
                        s_rot[g][1] = t * c;
                        s_rot[g][2] = apq.r / mag;
                        s_rot[g][3] = apq.i / mag;
                        s_rot[g][4] = 0.0;
                    }
                }
                __syncthreads();
                const bool act = !done && s_rot[g][4] == 0.0;
                const double c = s_rot[g][0], sn = s_rot[g][1];
                const zd em = {s_rot[g][2], -s_rot[g][3]};  // e^{-i phi}
                const zd ep = {s_rot[g][2], s_rot[g][3]};   // e^{+i phi}
                if (act && sub < nb) {  // columns of A and V
                    const int k = sub;
                    zd x = A[k * ld + p], y = zmul(em, A[k * ld + qq]);
                    A[k * ld + p] = {c * x.r - sn * y.r, c * x.i - sn * y.i};
                    A[k * ld + qq] = {sn * x.r + c * y.r, sn * x.i + c * y.i};
                    if (var == 0) {   // the displaced matrices give eigenvalues only
                        zd vx = V[k * ld + p], vy = zmul(em, V[k * ld + qq]);
                        V[k * ld + p] = {c * vx.r - sn * vy.r, c * vx.i - sn * vy.i};
                        V[k * ld + qq] = {sn * vx.r + c * vy.r, sn * vx.i + c * vy.i};
                    }
                }
                __syncthreads();
                if (act && sub < nb) {  // rows of A
                    const int k = sub;
                    zd x = A[p * ld + k], y = zmul(ep, A[qq * ld + k]);
                    A[p * ld + k] = {c * x.r - sn * y.r, c * x.i - sn * y.i};
                    A[qq * ld + k] = {sn * x.r + c * y.r, sn * x.i + c * y.i};
                }
                __syncthreads();
                if (act && sub == 0) {
                    A[p * ld + qq] = {0.0, 0.0};
                    A[qq * ld + p] = {0.0, 0.0};
                    A[p * ld + p].i = 0.0;
                    A[qq * ld + qq].i = 0.0;
                }
                __syncthreads();
            }
    }
    }
    // 4. ascending order of the signed eigenvalues (zheev), ties keep index order
    if (sub == 0) {
        int *ord = s_ord[g];
        for (int k = 0; k < nb; k++) ord[k] = k;
        for (int a = 1; a < nb; a++) {
            int x = ord[a];
            double vx = A[x * ld + x].r;
            int b = a - 1;
            while (b >= 0 && A[ord[b] * ld + ord[b]].r > vx) {
                ord[b + 1] = ord[b];
                b--;
            }
            ord[b + 1] = x;
        }
    }
    __syncthreads();
    if (!valid || sub >= nb) return;
    const int src = s_ord[g][sub];
    evals7[((size_t)q * nvar + var) * nb + sub] = fabs(A[src * ld + src].r);
    if (var == 0) {
        // column j = sub (sorted) of V; "pure imaginary" columns (sum |Re| < 1e-3) times i
        double sre = 0;
        for (int k = 0; k < nb; k++) sre += fabs(V[k * ld + src].r);
        const bool rot = sre < 1e-3;
        for (int k = 0; k < nb; k++) {
            zd v = V[k * ld + src];
            if (rot) v = {-v.i, v.r};
            size_t o = 2 * (((size_t)q * nb + sub) * nb + k);
            evec0[o] = v.r;
            evec0[o + 1] = v.i;
        }
    }
}

// finitedif_vel velocities + get_frequencies' second sort
// One wavefront per mesh point q, lane j holding mode j (nb <= 64): velocities by the central
// differences of the displaced eigenvalues, the eigenvalues sorted ascending as the reference's
// exchange sort orders them (extratools.f90:154-175), and the eigenvectors permuted to match.
// The exchange sort's step i moves the running minimum of positions i..nb-1 into position i:
// each position j > i holding a value strictly below everything in i..j-1 receives that
// running minimum (the first position of it), and position i receives the minimum of the
// whole range (first position).  Each step is one prefix-minimum scan across the lanes, so
// the permutation, ties included, is the serial sort's.
__global__ void __launch_bounds__(256) k_eig_finish(int nk, int nb, double dq, double cnst,
                                                    const double *__restrict__ evals7,
                                                    const double *__restrict__ evec0, double *__restrict__ eival,
                                                    double *__restrict__ evec, double *__restrict__ vel,
                                                    double *__restrict__ freq)
{
    const int lane = threadIdx.x & 63;
    const int q = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (q >= nk) return;   // uniform over the wavefront
    const double *e = evals7 + (size_t)q * kDq * nb;
    const bool act = lane < nb;
    const double e0 = act ? e[lane] : 0.0;
    if (act)
        for (int a = 0; a < 3; a++) {
            const double ep = e[(1 + 2 * a) * nb + lane], em = e[(2 + 2 * a) * nb + lane];
            vel[((size_t)q * nb + lane) * 3 + a] = (ep - em) / 2 / dq / 2 / sqrt(e0) * cnst * 1e-8 * 2 * alt3::PI;
        }
    double v = e0;   // value at position lane
    int id = lane;   // its mode before sorting
    for (int i = 0; i < nb - 1; i++) {
        // inclusive scan over positions i..lane of (value, id), the earlier one kept on ties
        const bool in = lane >= i && act;
        double sv = in ? v : INFINITY;
        int sid = in ? id : -1;
        for (int off = 1; off < 64; off <<= 1) {
            const double ov = __shfl_up(sv, off, 64);
            const int oid = __shfl_up(sid, off, 64);
            if (lane >= off && !(sv < ov)) {
                sv = ov;
                sid = oid;
            }
        }
        const double xv = __shfl_up(sv, 1, 64);   // exclusive: positions i..lane-1
        const int xid = __shfl_up(sid, 1, 64);
        const double mv = __shfl(sv, nb - 1, 64);   // the whole range
        const int mid = __shfl(sid, nb - 1, 64);
        if (act && lane > i && v < xv) {
            v = xv;
            id = xid;
        } else if (lane == i) {
            v = mv;
            id = mid;
        }
    }
    if (act) {
        eival[(size_t)q * nb + lane] = v;
        freq[(size_t)q * nb + lane] = sqrt(v) * cnst;
    }
    for (int j = 0; j < nb; j++) {
        const int src = __shfl(id, j, 64);
        const double *s = evec0 + 2 * ((size_t)q * nb + src) * nb;
        double *d = evec + 2 * ((size_t)q * nb + j) * nb;
        for (int k = lane; k < 2 * nb; k += 64) d[k] = s[k];
    }
}

// subst_eivecs: for k in the negative half, e(-k) <- conj(e(k))
__global__ void k_subst(int nneg, int nb, const int *__restrict__ negs, const int *__restrict__ negidx,
                        double *__restrict__ evec)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    int per = 2 * nb * nb;
    if (t >= nneg * nb * nb) return;
    int n = t / (nb * nb), e = t % (nb * nb);
    const double *src = evec + (size_t)negs[n] * per;
    double *dst = evec + (size_t)negidx[n] * per;
    dst[2 * e] = src[2 * e];
    dst[2 * e + 1] = -src[2 * e + 1];
}

}  // namespace

namespace alt3 {

// read_born's data and the constants nonanal derives from the lattice (rest.f90:3528-3538)
int born_prepare(alt3_ctx *ctx, const alt3_system *sys)
{
    ctx->born = BornDev{};
    if (!sys->born_flag) return ALT3_OK;
    if (!sys->epsil || !sys->zeu) return fail(ALT3_E_ARG, "born_flag set without epsil and charges");
    BornDev &b = ctx->born;
    b.on = 1;
    for (int k = 0; k < 9; k++) b.epsil[k] = sys->epsil[k];
    b.eps_scale = 8.8541878176e-12 / 1e10 / EE;
    const double *g = ctx->T.g, *r = sys->r123;
    auto len = [](const double *v) { return std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); };
    const double gg = std::pow(len(g) * len(g + 3) * len(g + 6), (double)0.33f);   // **0.33, a REAL*4 literal
    const double rho = gg / 4.0;
    b.rho2 = rho * rho;
    const double cx = r[1] * r[5] - r[2] * r[4], cy = r[2] * r[3] - r[0] * r[5], cz = r[0] * r[4] - r[1] * r[3];
    b.om0 = std::fabs(r[6] * cx + r[7] * cy + r[8] * cz);   // calculate_volume: |r3 . (r1 x r2)|
    std::vector<double> zm((size_t)10 * sys->natoms0);
    for (int a = 0; a < sys->natoms0; a++) {
        for (int k = 0; k < 9; k++) zm[10 * a + k] = sys->zeu[9 * a + k];
        zm[10 * a + 9] = sys->mass[a];
    }
    ALT3_TRY(ctx->d_born.alloc(zm.size()));
    ALT3_HIP(hipMemcpy(ctx->d_born.p, zm.data(), sizeof(double) * zm.size(), hipMemcpyHostToDevice));
    return ALT3_OK;
}

int eig_prepare(alt3_ctx *ctx, const alt3_system *sys)
{
    ALT3_TRY(born_prepare(ctx, sys));
    const int nb = ctx->nb, nn = nb * nb;
    std::vector<std::vector<int>> per(nn);
    for (int t = 0; t < sys->nt2; t++) {
        int i0 = sys->fc2_atoms[2 * t], j = sys->fc2_atoms[2 * t + 1];
        if (i0 < 1 || i0 > sys->natoms0) continue;  // only home-cell rows (rest.f90:481)
        if (j < 1 || j > sys->natoms) return fail(ALT3_E_ARG, "fc2 atom index out of range");
        if (sys->fc2_xyz[2 * t] < 1 || sys->fc2_xyz[2 * t] > 3 || sys->fc2_xyz[2 * t + 1] < 1 || sys->fc2_xyz[2 * t + 1] > 3)
            return fail(ALT3_E_ARG, "fc2 xyz index out of range");
        int j0 = sys->iatomcell0[j - 1];
        int row = sys->fc2_xyz[2 * t] - 1 + 3 * (i0 - 1), col = sys->fc2_xyz[2 * t + 1] - 1 + 3 * (j0 - 1);
        per[row * nb + col].push_back(t);
    }
    std::vector<int> ptr(nn + 1, 0);
    std::vector<double> amp, rr;
    for (int e = 0; e < nn; e++) {
        for (int t : per[e]) {
            int i0 = sys->fc2_atoms[2 * t], j = sys->fc2_atoms[2 * t + 1], j0 = sys->iatomcell0[j - 1];
            amp.push_back(sys->fc2[t] / std::sqrt(sys->mass[i0 - 1] * sys->mass[j0 - 1]));
            for (int a = 0; a < 3; a++) rr.push_back(sys->atompos[3 * (j - 1) + a] - sys->atompos[3 * (j0 - 1) + a]);
        }
        ptr[e + 1] = (int)amp.size();
    }
    ctx->nt2h = (int)amp.size();
    ALT3_TRY(ctx->d_fc2_rc.alloc(nn + 1));
    ALT3_TRY(ctx->d_fc2_amp.alloc(std::max<size_t>(1, amp.size())));
    ALT3_TRY(ctx->d_fc2_rr.alloc(std::max<size_t>(1, rr.size())));
    ALT3_HIP(hipMemcpy(ctx->d_fc2_rc.p, ptr.data(), sizeof(int) * (nn + 1), hipMemcpyHostToDevice));
    if (!amp.empty()) {
        ALT3_HIP(hipMemcpy(ctx->d_fc2_amp.p, amp.data(), sizeof(double) * amp.size(), hipMemcpyHostToDevice));
        ALT3_HIP(hipMemcpy(ctx->d_fc2_rr.p, rr.data(), sizeof(double) * rr.size(), hipMemcpyHostToDevice));
    }
    return ALT3_OK;
}

// Hermitian eigensolves of D(q) for np points kq (Cartesian) and nvar variants each (7: q and
// q +- dq e_a; 1: q only): evals [np][nvar][nb] = |lambda| in zheev order, evec [np][nb col][nb]
// (variant 0, with the x i fix of pure-imaginary columns)
int eig_launch(alt3_ctx *ctx, int np, int nvar, const double *kq, double dq, double *evals, double *evec)
{
    const int nb = ctx->nb;
    if (nb > 8) {   // one matrix per wavefront, parallel (round-robin) Jacobi order
        const size_t nmat = (size_t)np * nvar;
        const size_t lds = (size_t)2 * nb * (nb + 1) * sizeof(zd);
        if (lds + 4096 > ctx->lds_max) return fail(ALT3_E_ARG, "eigen solve: matrix pair does not fit in LDS");
        const void *kern = (const void *)k_eig_group<64, true>;
        if (lds + 4096 > 65536) ALT3_HIP(hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        if (nvar == 1 || nb > 32) {
            hipLaunchKernelGGL((k_eig_group<64, true>), dim3((unsigned)nmat), dim3(64), lds, ctx->stream, np, nvar, nb,
                               kq, ctx->d_fc2_rc.p, ctx->d_fc2_amp.p, ctx->d_fc2_rr.p, dq, evals, evec, ctx->born,
                               ctx->d_born.p, 0, nvar, nullptr);
        } else {   // the central solves, then the displaced ones started from their eigenvectors
            hipLaunchKernelGGL((k_eig_group<64, true>), dim3((unsigned)np), dim3(64), lds, ctx->stream, np, nvar, nb,
                               kq, ctx->d_fc2_rc.p, ctx->d_fc2_amp.p, ctx->d_fc2_rr.p, dq, evals, evec, ctx->born,
                               ctx->d_born.p, 0, 1, nullptr);
            hipLaunchKernelGGL((k_eig_group<64, true>), dim3((unsigned)(np * (size_t)(nvar - 1))), dim3(64), lds,
                               ctx->stream, np, nvar, nb, kq, ctx->d_fc2_rc.p, ctx->d_fc2_amp.p, ctx->d_fc2_rr.p, dq,
                               evals, evec, ctx->born, ctx->d_born.p, 1, nvar - 1, evec);
        }
        ALT3_HIP(hipGetLastError());
        return ALT3_OK;
    }
    const int G = nb <= 8 ? 8 : nb <= 16 ? 16 : nb <= 32 ? 32 : 64;
    if (G < 64) {   // 64/G matrices per wavefront
        const int M = 64 / G;
        const size_t nmat = (size_t)np * nvar;
        const size_t lds = (size_t)M * 2 * nb * nb * sizeof(zd);
        const void *kern = G == 8 ? (const void *)k_eig_group<8> : G == 16 ? (const void *)k_eig_group<16>
                                                                           : (const void *)k_eig_group<32>;
        if (lds + 4096 > 65536) ALT3_HIP(hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        dim3 grid((unsigned)((nmat + M - 1) / M));
#define ALT3_EIG_GROUP(GG)                                                                                     \
        hipLaunchKernelGGL(k_eig_group<GG>, grid, dim3(64), lds, ctx->stream, np, nvar, nb, kq, ctx->d_fc2_rc.p, \
                           ctx->d_fc2_amp.p, ctx->d_fc2_rr.p, dq, evals, evec, ctx->born, ctx->d_born.p);
        if (G == 8) {
            ALT3_EIG_GROUP(8)
        } else if (G == 16) {
            ALT3_EIG_GROUP(16)
        } else {
            ALT3_EIG_GROUP(32)
        }
#undef ALT3_EIG_GROUP
    } else {
        size_t lds = (size_t)2 * nb * nb * sizeof(zd) + 64 * sizeof(double);
        if (lds > ctx->lds_max) return fail(ALT3_E_ARG, "eigen solve: matrix pair does not fit in LDS");
        if (lds > 65536)   // cells with more than 44 branches
            ALT3_HIP(hipFuncSetAttribute((const void *)k_eig_solve, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        hipLaunchKernelGGL(k_eig_solve, dim3((unsigned)((size_t)np * nvar)), dim3(64), lds, ctx->stream, np, nvar, nb,
                           kq, ctx->d_fc2_rc.p, ctx->d_fc2_amp.p, ctx->d_fc2_rr.p, dq, evals, evec, ctx->born,
                           ctx->d_born.p);
    }
    ALT3_HIP(hipGetLastError());
    return ALT3_OK;
}

int eig_run(alt3_ctx *ctx)
{
    const int nk = ctx->nk, nb = ctx->nb;
    if (nb > 64) return fail(ALT3_E_ARG, "more than 64 branches per cell");
    DBuf<double> &ev7 = ctx->d_ev7, &evec0 = ctx->d_evec0;
    ALT3_TRY(ev7.reserve((size_t)nk * kDq * nb));
    ALT3_TRY(evec0.reserve((size_t)nk * nb * nb * 2));
    ALT3_TRY(ctx->d_eival.alloc((size_t)nk * nb));
    ALT3_TRY(ctx->d_freq.alloc((size_t)nk * nb));
    ALT3_TRY(ctx->d_vel.alloc((size_t)nk * nb * 3));
    ALT3_TRY(ctx->d_evec.alloc((size_t)nk * nb * nb * 2));
    const double dq = (double)0.001f;
    ALT3_TRY(eig_launch(ctx, nk, kDq, ctx->d_kpc_eig.p, dq, ev7.p, evec0.p));
    ALT3_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_eig_finish, dim3((nk + 3) / 4), dim3(256), 0, ctx->stream, nk, nb, dq, cnst(), ev7.p,
                       evec0.p, ctx->d_eival.p, ctx->d_evec.p, ctx->d_vel.p, ctx->d_freq.p);
    ALT3_HIP(hipGetLastError());
    // subst_eivecs
    const int npos = ctx->T.npos, nneg = nk - npos;
    if (nneg > 0) {
        std::vector<int> negs(nneg), negidx(nneg);
        for (int n = 0; n < nneg; n++) {
            int k = ctx->T.mappos[npos + n];
            negs[n] = k;
            int c2 = k % ctx->nc[2], c1 = (k / ctx->nc[2]) % ctx->nc[1], c0 = k / (ctx->nc[1] * ctx->nc[2]);
            int m0 = (ctx->nc[0] - c0) % ctx->nc[0], m1 = (ctx->nc[1] - c1) % ctx->nc[1], m2 = (ctx->nc[2] - c2) % ctx->nc[2];
            negidx[n] = m2 + ctx->nc[2] * (m1 + ctx->nc[1] * m0);
        }
        ALT3_TRY(ctx->d_negs.reserve(2 * (size_t)nneg));
        int *dn = ctx->d_negs.p, *di = dn + nneg;
        ALT3_HIP(hipMemcpyAsync(dn, negs.data(), sizeof(int) * nneg, hipMemcpyHostToDevice, ctx->stream));
        ALT3_HIP(hipMemcpyAsync(di, negidx.data(), sizeof(int) * nneg, hipMemcpyHostToDevice, ctx->stream));
        int tot = nneg * nb * nb;
        hipLaunchKernelGGL(k_subst, dim3((tot + 255) / 256), dim3(256), 0, ctx->stream, nneg, nb, dn, di, ctx->d_evec.p);
        ALT3_HIP(hipGetLastError());
        ALT3_HIP(hipStreamSynchronize(ctx->stream));
    }
    ALT3_HIP(hipStreamSynchronize(ctx->stream));
    ctx->have_eig = true;
    ctx->piso_ok = false;
    ctx->have_veloci = false;
    return ALT3_OK;
}

}  // namespace alt3
