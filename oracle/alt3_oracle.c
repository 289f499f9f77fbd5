/*
 * ORACLE -- test infrastructure only.
 *
 * CPU restatement (plain C, FP64, OpenMP) of THERMACOND's three-phonon path, used by the
 * tests as the checker and by bench.py as the CPU baseline.  Nothing in the product links,
 * loads or calls this file.  Arrays use the reference's Fortran column-major layouts with
 * 0-based indices; each function cites the reference routine it restates.
 *
 * Pinning (what fixes each function independently of this file):
 *   indexing (get_k_info, make_kp_reg[_tet], get_negatives, get_weights, getstar,
 *   kstarsinfbz1, lattice ops)  -> integer dumps of the reference build (oracle/_ref)
 *   v3sq                       -> reference v33s8 dumps; closed-form invariances in tests
 *   tetra weights / fgr         -> reference P1sm/P2sm dumps; DOS sum rule of a parabolic band
 *   tau, kappa_rta, cm, rhs, cg, kappa_bte -> reference Q/tau/kappa/CM/RHS/x dumps;
 *                                 CG against a dense direct solve
 *
 * Build: gcc -O2 -fopenmp -ffp-contract=off -shared -fPIC oracle/alt3_oracle.c -lm
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <complex.h>

typedef double complex zc;

/* ---- constants (THERMACOND/modules.f90:4-24); cnst is a REAL*4 literal ---- */
static const double PI = 3.14159265358979323846;
static const double H_PLANK = 6.62606896e-34;
static const double K_B = 1.3806504e-23;
static const double C_LIGHT = 2.99792458e+8;
static const double HBAR = 1.054571628e-34;
static const double EE = 1.60217653e-19;
static const double UMA = 1.66053878e-27;
static double cnst_(void) { return (double)521.1098918f; }
double or_const33(void)
{ /* kappa.f90:182 */
    return 1e30 * pow(HBAR / UMA, 1.5) / pow(sqrt(EE * 1e20 / UMA), 1.5) * EE / H_PLANK / 100 / C_LIGHT;
}
double or_cnst(void) { return cnst_(); }

/* ---------------------------------------------------------------------------- */
/* geometry helpers                                                              */
/* ---------------------------------------------------------------------------- */
static void cross(const double *a, const double *b, double *c)
{
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}
static double dot3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

/* make_reciprocal_lattice_2pi (extratools.f90:209-228): rows r[0..2] -> rows g[0..2] */
void or_reciprocal_2pi(const double *r, double *g, double *volume_r)
{
    double c[3];
    cross(r + 3, r + 6, c);
    double om = dot3(r, c);
    for (int i = 0; i < 3; i++) g[i] = 2 * PI * c[i] / om;
    cross(r + 6, r, c);
    for (int i = 0; i < 3; i++) g[3 + i] = 2 * PI * c[i] / om;
    cross(r, r + 3, c);
    for (int i = 0; i < 3; i++) g[6 + i] = 2 * PI * c[i] / om;
    if (volume_r) *volume_r = fabs(om);
}

/* make_reciprocal_lattice (no 2 pi; extratools.f90:185-207): rr_i . g_j = delta_ij */
static void recip_nopi(const double *g, double *rr)
{
    double c[3];
    cross(g + 3, g + 6, c);
    double om = dot3(g, c);
    for (int i = 0; i < 3; i++) rr[i] = c[i] / om;
    cross(g + 6, g, c);
    for (int i = 0; i < 3; i++) rr[3 + i] = c[i] / om;
    cross(g, g + 3, c);
    for (int i = 0; i < 3; i++) rr[6 + i] = c[i] / om;
}

/* comp1 (extratools.f90:578-595) */
static void comp1(const double *q, const double *rr, int n, int *i, int *insid)
{
    const double epsl = 5e-10;
    double qdr = dot3(q, rr) + epsl;
    if (qdr < 0 || qdr >= 1) *insid = 0;
    double aux = qdr - floor(qdr);
    double y = 1 + n * aux;
    *i = (int)lround(y);
    if (fabs(*i - y) > (double)0.01f) *insid = 2;   /* REAL*4 literal, extratools.f90:609 */
    if (*i == n + 1) *i = 1;
}

/* get_k_info (rest.f90:3024-3058) -> 0-based mesh index; ijk are 1-based as in Fortran.
 * Returns inside (1 in box, 0 outside, 2 not a mesh point). */
int or_get_k_info(const double *q, const int *nc, const double *g, int *nk, int *ijk)
{
    double rr[9], w[3];
    recip_nopi(g, rr);
    int even = (nc[0] % 2 == 0) && (nc[1] % 2 == 0) && (nc[2] % 2 == 0);
    for (int a = 0; a < 3; a++) w[a] = q[a];
    if (even)
        for (int a = 0; a < 3; a++) w[a] = q[a] + 0.5 * (g[a] + g[3 + a] + g[6 + a]);
    int inside = 1, i, j, k;
    comp1(w, rr, nc[0], &i, &inside);
    comp1(w, rr + 3, nc[1], &j, &inside);
    comp1(w, rr + 6, nc[2], &k, &inside);
    if (ijk) { ijk[0] = i; ijk[1] = j; ijk[2] = k; }
    *nk = (k - 1) + (j - 1) * nc[2] + (i - 1) * nc[1] * nc[2];
    return inside;
}

/* make_kp_reg (modules.f90:1721-1784): REAL*4 fractions, even meshes shifted by -G/2 */
void or_make_kp_reg(const int *nc, const double *g, double *kpt)
{
    int nk = 0;
    double sh[3];
    for (int a = 0; a < 3; a++) sh[a] = (-0.5) * (g[a] + g[3 + a] + g[6 + a]);
    for (int i = 1; i <= nc[0]; i++)
        for (int j = 1; j <= nc[1]; j++)
            for (int k = 1; k <= nc[2]; k++) {
                double f1 = (double)((float)(i - 1) / (float)nc[0]);
                double f2 = (double)((float)(j - 1) / (float)nc[1]);
                double f3 = (double)((float)(k - 1) / (float)nc[2]);
                for (int a = 0; a < 3; a++) kpt[3 * nk + a] = f1 * g[a] + f2 * g[3 + a] + f3 * g[6 + a] + 0.0;
                nk++;
            }
    if (nc[0] % 2 == 0 && nc[1] % 2 == 0 && nc[2] % 2 == 0)
        for (int i = 0; i < nk; i++)
            for (int a = 0; a < 3; a++) kpt[3 * i + a] = kpt[3 * i + a] + sh[a];
}

/* make_kp_reg_tet (tetrahedron.f90:2-142): FP64 mesh (shift sx=sy=sz=0) + 6 tetrahedra
 * per cube; tet[4*t + c] = 0-based mesh index of corner c of tetrahedron t. */
int or_make_kp_reg_tet(const int *nc, const double *g, double *kpt, int *tet)
{
    int nx = nc[0], ny = nc[1], nz = nc[2], nk = 0, nt = 0;
    double shft[3];
    for (int a = 0; a < 3; a++) shft[a] = (-0.5) * (g[a] + g[3 + a] + g[6 + a]);
    for (int i = 1; i <= nx; i++)
        for (int j = 1; j <= ny; j++)
            for (int k = 1; k <= nz; k++) {
                for (int a = 0; a < 3; a++)
                    kpt[3 * nk + a] = ((i - 1 + 0.0) / nx) * g[a] + ((j - 1 + 0.0) / ny) * g[3 + a] +
                                      ((k - 1 + 0.0) / nz) * g[6 + a] + 0.0;
                nk++;
            }
    static const int corner[6][4] = {{1, 2, 3, 6}, {2, 3, 4, 6}, {1, 3, 5, 6}, {3, 4, 6, 8}, {3, 5, 6, 7}, {3, 6, 7, 8}};
    for (int i = 1; i <= nx; i++)
        for (int j = 1; j <= ny; j++)
            for (int k = 1; k <= nz; k++) {
                double q[8][3];
                for (int a = 0; a < 3; a++) {
                    double q1 = ((i - 1 + 0.0) / nx) * g[a] + ((j - 1 + 0.0) / ny) * g[3 + a] + ((k - 1 + 0.0) / nz) * g[6 + a] + shft[a];
                    double d1 = (1.0 / nx) * g[a], d2 = (1.0 / ny) * g[3 + a], d3 = (1.0 / nz) * g[6 + a];
                    q[0][a] = q1;
                    q[1][a] = q1 + d1;
                    q[2][a] = q1 + d2;
                    q[3][a] = q1 + d1 + d2;
                    q[4][a] = q1 + d3;
                    q[5][a] = q1 + d3 + d1;
                    q[6][a] = q1 + d3 + d2;
                    q[7][a] = q1 + d3 + d1 + d2;
                }
                int pt[8];
                for (int c = 0; c < 8; c++) {
                    if (or_get_k_info(q[c], nc, g, &pt[c], NULL) == 2) return -1;
                }
                for (int l = 0; l < 6; l++) {
                    for (int c = 0; c < 4; c++) tet[4 * nt + c] = pt[corner[l][c] - 1];
                    nt++;
                }
            }
    if (nx % 2 == 0 && ny % 2 == 0 && nz % 2 == 0)
        for (int i = 0; i < nk; i++)
            for (int a = 0; a < 3; a++) kpt[3 * i + a] = kpt[3 * i + a] + shft[a];
    return 0;
}

/* get_negatives (rest.f90:3208-3260): map (0-based), returns npos */
int or_get_negatives(int nk, const double *kp, const int *nc, const double *g, int *map)
{
    int nps = 0;
    char *inc = calloc(nk, 1);
    for (int i = 0; i < nk; i++) {
        double mq[3] = {-kp[3 * i], -kp[3 * i + 1], -kp[3 * i + 2]};
        int mk;
        or_get_k_info(mq, nc, g, &mk, NULL);
        if (mk < i) continue;
        map[nps++] = i;
        inc[i] = 1;
    }
    int cnt = nps;
    for (int i = 0; i < nk; i++)
        if (!inc[i]) map[cnt++] = i;
    free(inc);
    return nps;
}

/* ---------------------------------------------------------------------------- */
/* lattice point group (force_constants.f90:1379-1525 dlatmat2; op_kmatrix :154-186)   */
/* ---------------------------------------------------------------------------- */
static int ndet3(const int *m) /* m column-major 3x3 */
{
    return m[0] * (m[4] * m[8] - m[7] * m[5]) - m[3] * (m[1] * m[8] - m[7] * m[2]) + m[6] * (m[1] * m[5] - m[4] * m[2]);
}
static double vlen(const double *v) { return sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); }

/* cart: column-major 3x3, column j = j-th primitive vector. matrices: column-major, [48][9] */
static int dlatmat2(const double *cart, double eps, int *matrices)
{
    enum { NMAX = 400 };
    static int ishort[3][NMAX][3];
    static double vshort[3][NMAX][3];
    int nshort[3] = {0, 0, 0};
    double abc[3][3];
    for (int i = 0; i < 3; i++) abc[i][i] = vlen(cart + 3 * i);
    for (int i = 1; i < 3; i++)
        for (int j = 0; j < i; j++) {
            double v[3];
            for (int k = 0; k < 3; k++) v[k] = cart[3 * i + k] - cart[3 * j + k];
            abc[i][j] = abc[j][i] = vlen(v);
        }
    double xmax = fmax(fmax(abc[0][0], abc[1][1]), abc[2][2]) + eps;
    int foundone = 1, i = 0;
    while (foundone) {
        i++;
        foundone = 0;
        for (int j1 = -i; j1 <= i; j1++)
            for (int j2 = -i; j2 <= i; j2++)
                for (int j3 = -i; j3 <= i; j3++) {
                    if (!(abs(j1) == i || abs(j2) == i || abs(j3) == i)) continue;
                    double v[3] = {0, 0, 0};
                    for (int k = 0; k < 3; k++) {
                        v[k] = v[k] + j1 * cart[k];
                        v[k] = v[k] + j2 * cart[3 + k];
                        v[k] = v[k] + j3 * cart[6 + k];
                    }
                    double d = vlen(v);
                    if (d < xmax) foundone = 1;
                    for (int k = 0; k < 3; k++)
                        if (fabs(d - abc[k][k]) < eps) {
                            if (nshort[k] >= NMAX) return -1;
                            int n = nshort[k]++;
                            ishort[k][n][0] = j1; ishort[k][n][1] = j2; ishort[k][n][2] = j3;
                            for (int c = 0; c < 3; c++) vshort[k][n][c] = v[c];
                        }
                }
    }
    int nm = 1;
    int itrans[9];
    for (int i1 = 0; i1 < nshort[0]; i1++)
        for (int i2 = 0; i2 < nshort[1]; i2++)
            for (int i3 = 0; i3 < nshort[2]; i3++) {
                int ch[3] = {i1, i2, i3};
                for (int c = 0; c < 3; c++) {
                    itrans[c] = ishort[0][i1][c];
                    itrans[3 + c] = ishort[1][i2][c];
                    itrans[6 + c] = ishort[2][i3][c];
                }
                if (abs(ndet3(itrans)) != 1) continue;
                int ok = 1;
                for (int a = 1; a < 3 && ok; a++)
                    for (int b = 0; b < a; b++) {
                        double v[3];
                        for (int c = 0; c < 3; c++) v[c] = vshort[a][ch[a]][c] - vshort[b][ch[b]][c];
                        if (fabs(vlen(v) - abc[a][b]) > eps) { ok = 0; break; }
                    }
                if (!ok) continue;
                if (itrans[0] + itrans[4] + itrans[8] == 3) {
                    memcpy(matrices, itrans, sizeof(itrans));
                } else {
                    if (nm >= 48) return -2;
                    memcpy(matrices + 9 * nm, itrans, sizeof(itrans));
                    nm++;
                }
            }
    /* closure by multiplication */
    static char tried[48][48];
    memset(tried, 0, sizeof(tried));
    foundone = 1;
    while (foundone) {
        foundone = 0;
        for (int a = 0; a < nm; a++)
            for (int b = 0; b < nm; b++) {
                if (tried[a][b]) continue;
                tried[a][b] = 1;
                int t[9];
                for (int r = 0; r < 3; r++)
                    for (int c = 0; c < 3; c++) {
                        int s = 0;
                        for (int k = 0; k < 3; k++) s += matrices[9 * a + 3 * k + r] * matrices[9 * b + 3 * c + k];
                        t[3 * c + r] = s;
                    }
                int found = 0;
                for (int k = 0; k < nm && !found; k++) {
                    int same = 1;
                    for (int e = 0; e < 9; e++)
                        if (matrices[9 * k + e] != t[e]) { same = 0; break; }
                    if (same) found = 1;
                }
                if (!found) {
                    if (nm >= 48) return -2;
                    memcpy(matrices + 9 * nm, t, sizeof(t));
                    nm++;
                    foundone = 1;
                }
            }
    }
    return nm;
}

static void mm3(const double *a, const double *b, double *c) /* column-major c = a*b */
{
    double t[9];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            double s = 0;
            for (int k = 0; k < 3; k++) s = s + a[3 * k + i] * b[3 * j + k];
            t[3 * j + i] = s;
        }
    memcpy(c, t, sizeof(t));
}

/* LU inverse with partial pivoting (xmatinv via ludcmp/lubksb, force_constants.f90:1165-1289) */
static int inv3(const double *ain, double *aout)
{
    double a[3][3], b[3][3];
    int idx[3];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) { a[i][j] = ain[3 * j + i]; b[i][j] = (i == j); }
    for (int i = 0; i < 3; i++) idx[i] = i;
    for (int k = 0; k < 3; k++) {
        int p = k;
        for (int i = k + 1; i < 3; i++)
            if (fabs(a[i][k]) > fabs(a[p][k])) p = i;
        if (a[p][k] == 0) return 1;
        if (p != k)
            for (int j = 0; j < 3; j++) {
                double t = a[p][j]; a[p][j] = a[k][j]; a[k][j] = t;
                t = b[p][j]; b[p][j] = b[k][j]; b[k][j] = t;
            }
        for (int i = k + 1; i < 3; i++) {
            double f = a[i][k] / a[k][k];
            for (int j = 0; j < 3; j++) { a[i][j] -= f * a[k][j]; b[i][j] -= f * b[k][j]; }
        }
    }
    for (int c = 0; c < 3; c++)
        for (int i = 2; i >= 0; i--) {
            double s = b[i][c];
            for (int j = i + 1; j < 3; j++) s -= a[i][j] * b[j][c];
            b[i][c] = s / a[i][i];
        }
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) aout[3 * j + i] = b[i][j];
    (void)idx;
    return 0;
}

/* lattice point group in k-space: latparams (6), primlatt column-major (3x3, columns =
 * primitive vectors in conventional units).  opk: [48][9] column-major op_kmatrix. */
int or_lattice_ops(const double *latparams, const double *primlatt, double *opk)
{
    double latp[6];
    memcpy(latp, latparams, sizeof(latp));
    double d2r = 4.0 * atan(1.0) / 180.0;
    for (int i = 3; i < 6; i++) latp[i] *= d2r;
    double c2c[9] = {0};
    c2c[0] = latp[0];
    c2c[3] = latp[1] * cos(latp[5]);
    c2c[4] = latp[1] * sin(latp[5]);
    c2c[6] = latp[2] * cos(latp[4]);
    c2c[7] = latp[2] * (cos(latp[3]) - cos(latp[5]) * cos(latp[4])) / sin(latp[5]);
    c2c[8] = sqrt(latp[2] * latp[2] - c2c[6] * c2c[6] - c2c[7] * c2c[7]);
    double p2c[9], c2p[9];
    mm3(c2c, primlatt, p2c);
    if (inv3(primlatt, c2p)) return -1;
    int mats[48 * 9];
    int n = dlatmat2(p2c, 1e-5, mats);
    if (n <= 0) return n;
    for (int m = 0; m < n; m++) {
        double t[9], ti[9];
        for (int e = 0; e < 9; e++) t[e] = mats[9 * m + e];
        mm3(t, c2p, t);
        mm3(primlatt, t, t);
        if (inv3(t, ti)) return -3;
        for (int j = 0; j < 3; j++)
            for (int k = 0; k < 3; k++) opk[9 * m + 3 * j + k] = (double)lround(ti[3 * k + j]);
    }
    return n;
}

/* getstar (force_constants.f90:1748-1785) */
int or_getstar(const double *kvec, const double *primlatt, int nops, const double *opk, double *star, int *ops)
{
    int narms = 0;
    double starp[48][3];
    for (int i = 0; i < nops; i++) {
        double v[3], v2[3];
        for (int r = 0; r < 3; r++) {
            double s = 0;
            for (int c = 0; c < 3; c++) s = s + opk[9 * i + 3 * c + r] * kvec[c];
            v[r] = s;
        }
        for (int c = 0; c < 3; c++) {
            double s = 0;
            for (int k = 0; k < 3; k++) s = s + v[k] * primlatt[3 * c + k];
            v2[c] = s;
        }
        int dup = 0;
        for (int j = 0; j < narms && !dup; j++) {
            int allint = 1;
            for (int k = 0; k < 3; k++) {
                double d = v2[k] - starp[j][k];
                double n = (double)lround(d);
                if (fabs(d - n) > 1e-6) { allint = 0; break; }
            }
            if (allint) dup = 1;
        }
        if (dup) continue;
        for (int c = 0; c < 3; c++) { star[3 * narms + c] = v[c]; starp[narms][c] = v2[c]; }
        ops[narms] = i;
        narms++;
    }
    return narms;
}

/* get_weights (modules.f90:2006-2162): returns nibz; kibz (3*nibz), wibz, mapibz (nk),
 * mapinv (nibz), all indices 0-based. */
int or_get_weights(int nk, const double *kp, const double *primlatt, int nops, const double *opk,
                   double *kibz, double *wibz, int *mapibz, int *mapinv)
{
    int nibz = 0;
    double *w2 = malloc(sizeof(double) * nk);
    for (int i = 0; i < nk; i++) w2[i] = 1;
    for (int i = 0; i < nk; i++) {
        double star[48 * 3];
        int ops[48];
        int narms = or_getstar(kp + 3 * i, primlatt, nops, opk, star, ops);
        int exists = 0;
        for (int l = 0; l < narms && !exists; l++)
            for (int j = 0; j < nibz; j++) {
                double d[3] = {kibz[3 * j] - star[3 * l], kibz[3 * j + 1] - star[3 * l + 1], kibz[3 * j + 2] - star[3 * l + 2]};
                if (vlen(d) < 1e-4) {
                    exists = 1;
                    w2[j] += 1;
                    mapibz[i] = j;
                    break;
                }
            }
        if (exists) continue;
        if (nibz == 0) w2[0] = 1;
        mapibz[i] = nibz;
        mapinv[nibz] = i;
        for (int a = 0; a < 3; a++) kibz[3 * nibz + a] = kp[3 * i + a];
        nibz++;
    }
    double sw = 0;
    for (int j = 0; j < nibz; j++) sw += w2[j];
    for (int j = 0; j < nibz; j++) wibz[j] = w2[j] / sw;
    free(w2);
    return nibz;
}

/* kstarsinfbz1 (sub.f90:5551-5590): star members that coincide with a mesh point
 * (component-wise |d| < tolerance, tolerance = REAL*4 1e-4, rest.f90:35). */
int or_kstarsinfbz1(const double *q, int nk, const double *kp, const double *primlatt, int nops,
                    const double *opk, double *kstars, int *kops, int *n0)
{
    double star[48 * 3];
    int ops[48];
    int narms = or_getstar(q, primlatt, nops, opk, star, ops), ds = 0;
    const double tol = (double)0.0001f;
    for (int s = 0; s < narms; s++)
        for (int j = 0; j < nk; j++) {
            if (fabs(kp[3 * j] - star[3 * s]) < tol && fabs(kp[3 * j + 1] - star[3 * s + 1]) < tol &&
                fabs(kp[3 * j + 2] - star[3 * s + 2]) < tol) {
                for (int c = 0; c < 3; c++) kstars[3 * ds + c] = star[3 * s + c];
                kops[ds] = ops[s];
                if (n0) n0[ds] = j;
                ds++;
            }
        }
    return ds;
}

/* ---------------------------------------------------------------------------- */
/* dynamical matrix (rest.f90:445-556), FP64                                     */
/* ---------------------------------------------------------------------------- */
/* dynmat: ndim x ndim column-major complex (interleaved re,im). Terms 1-based atoms. */
void or_dynmat(const double *kpt, int natoms0, const double *mass0, const double *atompos,
               const int *iatomcell0, int nt2, const int *at2, const int *xyz2, const double *fc2,
               double *dynmat_out)
{
    int ndim = 3 * natoms0;
    zc *d = calloc((size_t)ndim * ndim, sizeof(zc));
    for (int i0 = 1; i0 <= natoms0; i0++)
        for (int al = 1; al <= 3; al++) {
            int i3 = al + 3 * (i0 - 1);
            double mi = mass0[i0 - 1];
            for (int t = 0; t < nt2; t++) {
                if (i0 == at2[2 * t] && al == xyz2[2 * t]) {
                    int be = xyz2[2 * t + 1], j = at2[2 * t + 1], j0 = iatomcell0[j - 1];
                    double mj = mass0[j0 - 1];
                    int j3 = be + 3 * (j0 - 1);
                    double rr[3];
                    for (int c = 0; c < 3; c++) rr[c] = atompos[3 * (j - 1) + c] - atompos[3 * (j0 - 1) + c];
                    double ph = kpt[0] * rr[0] + kpt[1] * rr[1] + kpt[2] * rr[2];
                    zc junk = fc2[t] * cexp(I * ph) / sqrt(mi * mj);
                    d[(i3 - 1) + (size_t)ndim * (j3 - 1)] += junk;
                }
            }
        }
    double all = 0;
    for (int e = 0; e < ndim * ndim; e++) all += cabs(d[e]);
    all /= (ndim * ndim);
    for (int t = 1; t <= ndim; t++) {
        zc dt = d[(t - 1) * (ndim + 1)];
        if (fabs(cimag(dt)) > 9e-4 * fabs(creal(dt))) d[(t - 1) * (ndim + 1)] = creal(dt);
        for (int j = t + 1; j <= ndim - 1; j++) {
            zc a = d[(t - 1) + (size_t)ndim * (j - 1)], b = d[(j - 1) + (size_t)ndim * (t - 1)];
            double mi = (cimag(a) - cimag(b)) / 2, mj = (creal(a) + creal(b)) / 2;
            d[(t - 1) + (size_t)ndim * (j - 1)] = mj + I * mi;
            d[(j - 1) + (size_t)ndim * (t - 1)] = mj - I * mi;
        }
    }
    double si = 0, sr = 0;
    for (int e = 0; e < ndim * ndim; e++) { si += fabs(cimag(d[e])); sr += fabs(creal(d[e])); }
    if (si / (ndim * ndim) < sr / (ndim * ndim) * 1e-10)
        for (int e = 0; e < ndim * ndim; e++) d[e] = (d[e] + conj(d[e])) / 2.0;
    memcpy(dynmat_out, d, sizeof(zc) * ndim * ndim);
    free(d);
}

/* ---------------------------------------------------------------------------- */
/* V3 (sub.f90:1050-1233 + matrix_elt_simplified :5888-6011 + cal_eiqr2 :6015-6067) */
/* ---------------------------------------------------------------------------- */
/* eigenval (nb,nk), eigenvec (nb,nb,nk) complex interleaved; kp (3,nk) = FP64 mesh.
 * v33 out: (ksub, nk, nb, nb, nb) column-major.  ibz_lo..ibz_hi 0-based inclusive. */
int or_v3sq(int ibz_lo, int ibz_hi, const int *mapinv, int nk, const double *kp, const int *nc,
            const double *g, int nb, const double *eigenval, const double *eigenvec_ri,
            const double *mass0, const double *atompos, const int *iatomcell0, int nt3,
            const int *at3, const int *xyz3, const double *fc3, double *v33)
{
    const zc *ev = (const zc *)eigenvec_ri;
    int ksub = ibz_hi - ibz_lo + 1;
    /* cal_eiqr2 (sub.f90:6035) sizes the table by the largest second atom index and reads it
     * with the third as well; the two agree for lists closed under the (j,k) swap (FOCEX's and
     * the generators'), and sizing by both keeps every read in bounds */
    int maxat = 0;
    for (int t = 0; t < nt3; t++) {
        if (at3[3 * t + 1] > maxat) maxat = at3[3 * t + 1];
        if (at3[3 * t + 2] > maxat) maxat = at3[3 * t + 2];
    }
    zc *eiqr2 = malloc(sizeof(zc) * (size_t)nk * maxat);
    for (int i = 0; i < nk; i++)
        for (int j = 1; j <= maxat; j++) {
            int j0 = iatomcell0[j - 1];
            double rr[3];
            for (int c = 0; c < 3; c++) rr[c] = atompos[3 * (j - 1) + c] - atompos[3 * (j0 - 1) + c];
            double ph = kp[3 * i] * rr[0] + kp[3 * i + 1] * rr[1] + kp[3 * i + 2] * rr[2];
            eiqr2[i + (size_t)nk * (j - 1)] = cexp(I * ph);
        }
    const double c33 = or_const33();
    int err = 0;
#pragma omp parallel for collapse(2) schedule(dynamic)
    for (int s = 0; s < ksub; s++)
        for (int n2 = 0; n2 < nk; n2++) {
            int n1 = ibz_lo + s;
            const double *q1 = kp + 3 * mapinv[n1];
            const double *q2 = kp + 3 * n2;
            double q3[3] = {-q2[0] - q1[0], -q2[1] - q1[1], -q2[2] - q1[2]};
            int nk1, nk2, nk3;
            or_get_k_info(q1, nc, g, &nk1, NULL);
            if (nk1 != mapinv[n1]) { err = 1; continue; }
            or_get_k_info(q2, nc, g, &nk2, NULL);
            or_get_k_info(q3, nc, g, &nk3, NULL);
            for (int l1 = 0; l1 < nb; l1++)
                for (int l2 = 0; l2 < nb; l2++)
                    for (int l3 = 0; l3 < nb; l3++) {
                        zc xx = 0;
                        for (int t = 0; t < nt3; t++) {
                            int i3 = at3[3 * t], j = at3[3 * t + 1], k = at3[3 * t + 2];
                            if (iatomcell0[i3 - 1] != i3) continue;
                            int i0 = i3, j0 = iatomcell0[j - 1], k0 = iatomcell0[k - 1];
                            double mi = mass0[i0 - 1], mj = mass0[j0 - 1], mk = mass0[k0 - 1];
                            zc e3 = ev[(xyz3[3 * t] - 1 + 3 * (i0 - 1)) + (size_t)nb * (l1 + (size_t)nb * nk1)] *
                                    ev[(xyz3[3 * t + 1] - 1 + 3 * (j0 - 1)) + (size_t)nb * (l2 + (size_t)nb * nk2)] *
                                    ev[(xyz3[3 * t + 2] - 1 + 3 * (k0 - 1)) + (size_t)nb * (l3 + (size_t)nb * nk3)];
                            xx = xx + fc3[t] / sqrt(mi * mj * mk) * eiqr2[nk2 + (size_t)nk * (j - 1)] *
                                          eiqr2[nk3 + (size_t)nk * (k - 1)] * e3;
                        }
                        double den = sqrt(8 * sqrt(fabs(eigenval[l1 + nb * nk1] * eigenval[l2 + nb * nk2] * eigenval[l3 + nb * nk3])));
                        if (den == 0) { err = 2; continue; }
                        zc w = xx / den * c33;
                        v33[s + (size_t)ksub * (n2 + (size_t)nk * (l1 + (size_t)nb * (l2 + (size_t)nb * l3)))] = creal(w * conj(w));
                    }
        }
    free(eiqr2);
    return err;
}

/* ---------------------------------------------------------------------------- */
/* tetrahedron weights (tetrahedron.f90:188-384) with the reference's REAL*4 types  */
/* ---------------------------------------------------------------------------- */
static void ssort4(float *x) /* SSORT (tetrahedron.f90:716-738): descending via MAXLOC */
{
    for (int i = 0; i < 3; i++) {
        int m = i;
        for (int k = i + 1; k < 4; k++)
            if (x[k] > x[m]) m = k;
        if (m != i) { float t = x[i]; x[i] = x[m]; x[m] = t; }
    }
}

/* weights c[4] for one tetrahedron with corner energies w[4] at energy e */
void or_weight_tet1(const double *w, double e, double *c)
{
    float en[4], en1[4];
    int uns[4] = {0, 0, 0, 0};
    for (int k = 0; k < 4; k++) en[k] = (float)w[k];
    memcpy(en1, en, sizeof(en));
    ssort4(en);
    if (en[0] != en[1] && en[1] != en[2] && en[2] != en[3]) {
        for (int k = 0; k < 4; k++) {
            if (en[k] == en1[0]) uns[k] = 1;
            else if (en[k] == en1[1]) uns[k] = 2;
            else if (en[k] == en1[2]) uns[k] = 3;
            else if (en[k] == en1[3]) uns[k] = 4;
        }
    } else {
        for (int k = 0; k < 4; k++) {
            if (en[k] == en1[0]) uns[k] = 1;
            else if (en[k] == en1[1]) uns[k] = 2;
            else if (en[k] == en1[2]) uns[k] = 3;
            else if (en[k] == en1[3]) uns[k] = 4;
            for (int l = 0; l < k; l++) {
                if (uns[l] == uns[k] && uns[k] == 1) {
                    if (en[k] == en1[1]) uns[k] = 2;
                    else if (en[k] == en1[2]) uns[k] = 3;
                    else if (en[k] == en1[3]) uns[k] = 4;
                } else if (uns[l] == uns[k] && uns[k] == 2) {
                    if (en[k] == en1[2]) uns[k] = 3;
                    else if (en[k] == en1[3]) uns[k] = 4;
                } else if (uns[l] == uns[k] && uns[k] == 3) {
                    uns[k] = 4;
                }
            }
        }
    }
    float e1 = en[3], e2 = en[2], e3 = en[1], e4 = en[0];
    float e21 = e2 - e1, e31 = e3 - e1, e41 = e4 - e1, e32 = e3 - e2, e42 = e4 - e2, e43 = e4 - e3;
    const double small1 = 1e-10, small2 = 1e-10;
#define CLAMP(x)                                                   \
    do {                                                           \
        if (fabs((double)(x)) < small1 && (x) > 0.0f) (x) = (float)small2;  \
        if (fabs((double)(x)) < small1 && (x) < 0.0f) (x) = (float)(-small2); \
    } while (0)
    CLAMP(e21); CLAMP(e31); CLAMP(e41); CLAMP(e42); CLAMP(e43);
#undef CLAMP
    double cc[4];
    double E1 = e1, E2 = e2, E3 = e3, E4 = e4;
    if (e < E1 || e > E4) {
        cc[3] = cc[2] = cc[1] = cc[0] = 0;
    } else if (E1 <= e && e <= E2) {
        cc[3] = ((E2 - e) / (double)e21 + (E3 - e) / (double)e31 + (E4 - e) / (double)e41) * ((e - E1) * (e - E1)) /
                (double)(e41 * e31 * e21);
        cc[2] = ((e - E1) * (e - E1) * (e - E1)) / (double)(e21 * e21 * e31 * e41);
        cc[1] = ((e - E1) * (e - E1) * (e - E1)) / (double)(e21 * (e31 * e31) * e41);
        cc[0] = ((e - E1) * (e - E1) * (e - E1)) / (double)(e21 * e31 * (e41 * e41));
    } else if (E2 <= e && e <= E3) {
        float c11 = (float)((E3 - e) / (double)(e31 * e31));
        float c12 = (float)((E4 - e) / (double)(e41 * e41));
        float c21 = (float)((E3 - e) / (double)(e32 * e32));
        float c22 = (float)((E4 - e) / (double)(e42 * e42));
        float c31 = (float)((e - E2) / (double)(e32 * e32));
        float c32 = (float)((e - E1) / (double)(e31 * e31));
        float c41 = (float)((e - E2) / (double)(e42 * e42));
        float c42 = (float)((e - E1) / (double)(e41 * e41));
        float b11 = (float)((E3 - e) * (e - E2) / (double)(e42 * e32) + (E4 - e) * (e - E1) / (double)(e41 * e42) + (E3 - e) * (e - E1) / (double)(e32 * e41));
        float b12 = (float)((E4 - e) * (e - E1) / (double)(e42 * e31) + (E4 - e) * (e - E2) / (double)(e42 * e32) + (E3 - e) * (e - E1) / (double)(e31 * e32));
        float b21 = (float)((E3 - e) * (e - E2) / (double)(e42 * e31) + (E4 - e) * (e - E2) / (double)(e42 * e41) + (E3 - e) * (e - E1) / (double)(e31 * e41));
        float b22 = (float)((E3 - e) * (e - E2) / (double)(e32 * e31) + (E4 - e) * (e - E1) / (double)(e41 * e31) + (E4 - e) * (e - E2) / (double)(e32 * e41));
        float b31 = (float)((E3 - e) * (e - E2) / (double)(e42 * e31) + (E4 - e) * (e - E2) / (double)(e42 * e41) + (E3 - e) * (e - E1) / (double)(e31 * e41));
        float b32 = (float)((E3 - e) * (e - E2) / (double)(e42 * e32) + (E4 - e) * (e - E1) / (double)(e41 * e42) + (E3 - e) * (e - E1) / (double)(e32 * e41));
        float b41 = (float)((E3 - e) * (e - E2) / (double)(e32 * e31) + (E4 - e) * (e - E1) / (double)(e41 * e31) + (E4 - e) * (e - E2) / (double)(e32 * e41));
        float b42 = (float)((E4 - e) * (e - E1) / (double)(e42 * e31) + (E4 - e) * (e - E2) / (double)(e42 * e32) + (E3 - e) * (e - E1) / (double)(e31 * e32));
        cc[3] = 0.5f * (c11 * b11 + c12 * b12);
        cc[2] = 0.5f * (c21 * b21 + c22 * b22);
        cc[1] = 0.5f * (c31 * b31 + c32 * b32);
        cc[0] = 0.5f * (c41 * b41 + c42 * b42);
    } else {
        cc[3] = ((E4 - e) * (E4 - e) * (E4 - e)) / (double)(e41 * e41 * e42 * e43);
        cc[2] = ((E4 - e) * (E4 - e) * (E4 - e)) / (double)(e41 * (e42 * e42) * e43);
        cc[1] = ((E4 - e) * (E4 - e) * (E4 - e)) / (double)(e41 * e42 * (e43 * e43));
        cc[0] = ((e - E3) / (double)e43 + (e - E2) / (double)e42 + (e - E1) / (double)e41) * ((E4 - e) * (E4 - e)) /
                (double)(e41 * e42 * e43);
    }
    /* tet(i)%p(uns(k))%c = cc[k]; later assignments win */
    double out[4] = {0, 0, 0, 0};
    int set[4] = {0, 0, 0, 0};
    for (int k = 3; k >= 0; k--) {
        int u = uns[k];
        if (u >= 1 && u <= 4) { out[u - 1] = cc[k]; set[u - 1] = 1; }
    }
    for (int k = 0; k < 4; k++) c[k] = set[k] ? out[k] : 0.0;
}

/* calculate_FGR_mpi2 (sub.f90:7199-7499).  v33 and P1sm/P2sm: (ksub,nk,nb,nb,nb).
 * freq (nk,nb) = sqrt(eigenval)*cnst (cal_freq, sub.f90:1984-2006). */
int or_fgr(int ibz_lo, int ibz_hi, const int *mapinv, int nk, const double *kp, const int *nc,
           const double *g, int nb, const double *freq, const int *tet, const double *v33,
           double *P1, double *P2)
{
    int ksub = ibz_hi - ibz_lo + 1;
    int ntet = 6 * nk;
    /* mesh-index tables (one get_k_info each, as the reference's inner loops) */
    int *nkq2 = malloc(sizeof(int) * nk), *ijk2 = malloc(sizeof(int) * 3 * nk);
    int *nneg = malloc(sizeof(int) * nk), *ijkn = malloc(sizeof(int) * 3 * nk);
    for (int j = 0; j < nk; j++) {
        or_get_k_info(kp + 3 * j, nc, g, &nkq2[j], ijk2 + 3 * j);
        double m[3] = {-kp[3 * j], -kp[3 * j + 1], -kp[3 * j + 2]};
        or_get_k_info(m, nc, g, &nneg[j], ijkn + 3 * j);
    }
    const double tpi2 = (2 * PI) * (2 * PI);
    int h1 = nc[0] / 2 + 1, h2 = nc[1] / 2 + 1, h3 = nc[2] / 2 + 1;
#define ISG(p) ((p)[0] == h1 && (p)[1] == h2 && (p)[2] == h3)
    memset(P1, 0, sizeof(double) * (size_t)ksub * nk * nb * nb * nb);
    memset(P2, 0, sizeof(double) * (size_t)ksub * nk * nb * nb * nb);
#pragma omp parallel
    {
        int *np = malloc(sizeof(int) * nk), *ijkp = malloc(sizeof(int) * 3 * nk);
        int *nm = malloc(sizeof(int) * nk);
        double *F = malloc(sizeof(double) * nk);
#pragma omp for collapse(2) schedule(dynamic)
        for (int s = 0; s < ksub; s++)
            for (int ii = 0; ii < nb; ii++) {
                int i = mapinv[ibz_lo + s];
                const double *q1 = kp + 3 * i;
                for (int j = 0; j < nk; j++) {
                    double a[3] = {q1[0] + kp[3 * j], q1[1] + kp[3 * j + 1], q1[2] + kp[3 * j + 2]};
                    or_get_k_info(a, nc, g, &np[j], ijkp + 3 * j);
                    double b[3] = {q1[0] - kp[3 * j], q1[1] - kp[3 * j + 1], q1[2] - kp[3 * j + 2]};
                    or_get_k_info(b, nc, g, &nm[j], NULL);
                }
                double omega1 = freq[i + (size_t)nk * ii];
                for (int jj = 0; jj < nb; jj++)
                    for (int kk = 0; kk < nb; kk++) {
                        size_t base = (size_t)s + (size_t)ksub * ((size_t)nk * (ii + (size_t)nb * (jj + (size_t)nb * kk)));
                        /* process 1 (coalescence) */
                        for (int j = 0; j < nk; j++) {
                            double v = v33[base + (size_t)ksub * j] * tpi2 / nk;
                            if (jj < 3 && ISG(ijk2 + 3 * j)) v = 0.0;
                            if (kk < 3 && ISG(ijkp + 3 * j)) v = 0.0;
                            F[j] = v;
                        }
                        for (int t = 0; t < ntet; t++) {
                            double w[4], c[4];
                            for (int k = 0; k < 4; k++) {
                                int ww = tet[4 * t + k];
                                w[k] = -freq[ww + (size_t)nk * jj] + freq[np[ww] + (size_t)nk * kk];
                            }
                            or_weight_tet1(w, omega1, c);
                            for (int k = 0; k < 4; k++) {
                                int ww = tet[4 * t + k];
                                P1[base + (size_t)ksub * ww] = P1[base + (size_t)ksub * ww] + c[k] / 6.0 * F[ww];
                            }
                        }
                        /* process 2 (decay) */
                        for (int j = 0; j < nk; j++) {
                            double v = v33[base + (size_t)ksub * j] * tpi2 / nk;
                            if (jj < 3 && ISG(ijkn + 3 * j)) v = 0.0;
                            if (kk < 3 && ISG(ijkp + 3 * j)) v = 0.0;
                            F[nneg[j]] = v;
                        }
                        for (int t = 0; t < ntet; t++) {
                            double w[4], c[4];
                            for (int k = 0; k < 4; k++) {
                                int ww = tet[4 * t + k];
                                w[k] = freq[ww + (size_t)nk * jj] + freq[nm[ww] + (size_t)nk * kk];
                            }
                            or_weight_tet1(w, omega1, c);
                            for (int k = 0; k < 4; k++) {
                                int ww = tet[4 * t + k];
                                P2[base + (size_t)ksub * ww] = P2[base + (size_t)ksub * ww] + c[k] / 6.0 * F[ww];
                            }
                        }
                    }
            }
        free(np); free(ijkp); free(nm); free(F);
    }
#undef ISG
    free(nkq2); free(ijk2); free(nneg); free(ijkn);
    return 0;
}

/* P_iso_split (sub.f90:4825-4961): the isotope term of irreducible point n1, branch l1,
 *   Piso(n1,l1) = w1^2 pi/2 sum_l2 sum_tetra sum_corners gs(q,l2) c/6,
 *   gs(q,l2) = sum_atoms g(a) |dot_product(e(3a-2:3a, l1, q1), e(3a-2:3a, l2, q))|^2
 * with c the weight_tet weights of delta(w1 - w(q,l2)) over all 6 nk tetrahedra (corner
 * energies frequency(q,l2)).  The reference wraps this in a loop over n2 = 1..nk that changes
 * nothing inside, then divides by nk: nk identical copies and the division cancel, so the sum
 * is formed once.  evec: complex (nb,nb,nk) as interleaved doubles; g (natoms0); tet (4,6nk)
 * 0-based; Piso (ksub,nb). */
void or_piso(int ibz_lo, int ibz_hi, const int *mapinv, int nk, int nb, int natoms0, const double *freq,
             const int *tet, const double *evec, const double *gat, double *Piso)
{
    int ksub = ibz_hi - ibz_lo + 1;
#pragma omp parallel for collapse(2) schedule(dynamic)
    for (int s = 0; s < ksub; s++)
        for (int l1 = 0; l1 < nb; l1++) {
            const int i = mapinv[ibz_lo + s];
            const double omega1 = freq[i + (size_t)nk * l1];
            const double *e1 = evec + 2 * ((size_t)i * nb + l1) * nb;
            double S = 0;
            for (int l2 = 0; l2 < nb; l2++)
                for (int t = 0; t < 6 * nk; t++) {
                    double w[4], c[4];
                    for (int k = 0; k < 4; k++) w[k] = freq[tet[4 * t + k] + (size_t)nk * l2];
                    or_weight_tet1(w, omega1, c);
                    for (int k = 0; k < 4; k++) {
                        const double *e2 = evec + 2 * ((size_t)tet[4 * t + k] * nb + l2) * nb;
                        double gs = 0;
                        for (int a = 0; a < natoms0; a++) {
                            double re = 0, im = 0;
                            for (int x = 3 * a; x < 3 * a + 3; x++) {
                                re += e1[2 * x] * e2[2 * x] + e1[2 * x + 1] * e2[2 * x + 1];
                                im += e1[2 * x] * e2[2 * x + 1] - e1[2 * x + 1] * e2[2 * x];
                            }
                            double m = sqrt(re * re + im * im);
                            gs = gs + gat[a] * (m * m);
                        }
                        S = S + gs * c[k] / 6.0;
                    }
                }
            Piso[s + (size_t)ksub * l1] = S * omega1 * omega1 * PI / 2;
        }
}

/* Gaussian-smeared form of or_fgr: delta_g(x, sigma) = exp(-x^2/2/sigma^2)/sqrt(2 pi)/sigma
 * (extratools.f90:239-246) at each mesh point instead of the tetrahedron weights; same
 * processes, energies and acoustic masks as calculate_FGR_mpi2 (sub.f90:7199-7499). */
double or_delta_g(double x, double w) { return exp(-x * x / 2.0 / w / w) / sqrt(2 * PI) / w; }

int or_fgr_gauss(int ibz_lo, int ibz_hi, const int *mapinv, int nk, const double *kp, const int *nc,
                 const double *g, int nb, const double *freq, double sigma, const double *v33,
                 double *P1, double *P2)
{
    int ksub = ibz_hi - ibz_lo + 1;
    int *ijk2 = malloc(sizeof(int) * 3 * nk), *nneg = malloc(sizeof(int) * nk), *ijkn = malloc(sizeof(int) * 3 * nk);
    for (int j = 0; j < nk; j++) {
        int nn;
        or_get_k_info(kp + 3 * j, nc, g, &nn, ijk2 + 3 * j);
        double m[3] = {-kp[3 * j], -kp[3 * j + 1], -kp[3 * j + 2]};
        or_get_k_info(m, nc, g, &nneg[j], ijkn + 3 * j);
    }
    const double tpi2 = (2 * PI) * (2 * PI);
    int h1 = nc[0] / 2 + 1, h2 = nc[1] / 2 + 1, h3 = nc[2] / 2 + 1;
#define ISG(p) ((p)[0] == h1 && (p)[1] == h2 && (p)[2] == h3)
#pragma omp parallel
    {
        int *np = malloc(sizeof(int) * nk), *ijkp = malloc(sizeof(int) * 3 * nk), *nm = malloc(sizeof(int) * nk);
        double *F = malloc(sizeof(double) * nk);
#pragma omp for collapse(2) schedule(dynamic)
        for (int s = 0; s < ksub; s++)
            for (int ii = 0; ii < nb; ii++) {
                int i = mapinv[ibz_lo + s];
                const double *q1 = kp + 3 * i;
                for (int j = 0; j < nk; j++) {
                    double a[3] = {q1[0] + kp[3 * j], q1[1] + kp[3 * j + 1], q1[2] + kp[3 * j + 2]};
                    or_get_k_info(a, nc, g, &np[j], ijkp + 3 * j);
                    double b[3] = {q1[0] - kp[3 * j], q1[1] - kp[3 * j + 1], q1[2] - kp[3 * j + 2]};
                    or_get_k_info(b, nc, g, &nm[j], NULL);
                }
                double omega1 = freq[i + (size_t)nk * ii];
                for (int jj = 0; jj < nb; jj++)
                    for (int kk = 0; kk < nb; kk++) {
                        size_t base = (size_t)s + (size_t)ksub * ((size_t)nk * (ii + (size_t)nb * (jj + (size_t)nb * kk)));
                        for (int j = 0; j < nk; j++) {   /* coalescence */
                            double v = v33[base + (size_t)ksub * j] * tpi2 / nk;
                            if (jj < 3 && ISG(ijk2 + 3 * j)) v = 0.0;
                            if (kk < 3 && ISG(ijkp + 3 * j)) v = 0.0;
                            double e = -freq[j + (size_t)nk * jj] + freq[np[j] + (size_t)nk * kk];
                            P1[base + (size_t)ksub * j] = v * or_delta_g(omega1 - e, sigma);
                        }
                        for (int j = 0; j < nk; j++) {   /* decay, F at -q */
                            double v = v33[base + (size_t)ksub * j] * tpi2 / nk;
                            if (jj < 3 && ISG(ijkn + 3 * j)) v = 0.0;
                            if (kk < 3 && ISG(ijkp + 3 * j)) v = 0.0;
                            F[nneg[j]] = v;
                        }
                        for (int j = 0; j < nk; j++) {
                            double e = freq[j + (size_t)nk * jj] + freq[nm[j] + (size_t)nk * kk];
                            P2[base + (size_t)ksub * j] = F[j] * or_delta_g(omega1 - e, sigma);
                        }
                    }
            }
        free(np); free(ijkp); free(nm); free(F);
    }
#undef ISG
    free(ijk2); free(nneg); free(ijkn);
    return 0;
}

/* ---------------------------------------------------------------------------- */
/* occupations, Q and tau (sub.f90:1955-1980, 7953-8194, 8198-8305; extratools:257-288) */
/* ---------------------------------------------------------------------------- */
double or_nbe(double w, double t, int m)
{
    double z = w / t;
    if (m == 0) {
        if (z > 60) return 0;
        if (z < (double)0.0010f) return 1 / (z * (1 + z / 2 * (1 + z / 3)));
        return 1 / (exp(z) - 1);
    }
    return 1 / z;
}

double or_temp_cm(double tempk) { return tempk * K_B / (100 * H_PLANK * C_LIGHT); }

void or_distribution(int nk, int nb, const double *eigenval, double tempk, int classical, double *freq, double *dist)
{
    double temp = or_temp_cm(tempk);
    for (int i = 0; i < nk; i++)
        for (int ii = 0; ii < nb; ii++) {
            freq[i + (size_t)nk * ii] = sqrt(eigenval[ii + (size_t)nb * i]) * cnst_();
            dist[i + (size_t)nk * ii] = or_nbe(freq[i + (size_t)nk * ii], temp, classical);
        }
}

/* calculate_tauRTA_mpi2 for ibz_lo..hi: Q, QN, QU (ksub,nb); P1, P2 (ksub,nk,nb,nb,nb) */
void or_tau(int ibz_lo, int ibz_hi, const int *mapinv, int nk, const double *kp, const int *nc,
            const double *g, int nb, const double *dist, const double *P1sm, const double *P2sm,
            double *P1, double *P2, double *Q, double *QN, double *QU)
{
    int ksub = ibz_hi - ibz_lo + 1;
#pragma omp parallel
    {
        int *np = malloc(sizeof(int) * nk), *nm = malloc(sizeof(int) * nk);
        int *in1 = malloc(sizeof(int) * nk), *in2 = malloc(sizeof(int) * nk);
#pragma omp for schedule(dynamic)
        for (int s = 0; s < ksub; s++) {
            int i = mapinv[ibz_lo + s];
            const double *q1 = kp + 3 * i;
            for (int j = 0; j < nk; j++) {
                double a[3] = {q1[0] + kp[3 * j], q1[1] + kp[3 * j + 1], q1[2] + kp[3 * j + 2]};
                in1[j] = or_get_k_info(a, nc, g, &np[j], NULL);
                double b[3] = {q1[0] - kp[3 * j], q1[1] - kp[3 * j + 1], q1[2] - kp[3 * j + 2]};
                in2[j] = or_get_k_info(b, nc, g, &nm[j], NULL);
            }
            for (int ii = 0; ii < nb; ii++) {
                double q = 0, qn = 0, qu = 0;
                for (int jj = 0; jj < nb; jj++)
                    for (int kk = 0; kk < nb; kk++)
                        for (int j = 0; j < nk; j++) {
                            size_t x = (size_t)s + (size_t)ksub * (j + (size_t)nk * (ii + (size_t)nb * (jj + (size_t)nb * kk)));
                            double d1 = dist[i + (size_t)nk * ii], d2 = dist[j + (size_t)nk * jj];
                            double f1 = d1 * d2 * (dist[np[j] + (size_t)nk * kk] + 1);
                            double p1 = P1sm[x] * f1;
                            double f2 = d1 * (d2 + 1) * (dist[nm[j] + (size_t)nk * kk] + 1);
                            double p2 = P2sm[x] * f2;
                            if (P1) P1[x] = p1;
                            if (P2) P2[x] = p2;
                            q = q + (p1 + 0.5 * p2);
                            if (in1[j] == 1) qn = qn + p1; else qu = qu + p1;
                            if (in2[j] == 1) qn = qn + 0.5 * p2; else qu = qu + 0.5 * p2;
                        }
                Q[s + (size_t)ksub * ii] = q;
                QN[s + (size_t)ksub * ii] = qn;
                QU[s + (size_t)ksub * ii] = qu;
            }
        }
        free(np); free(nm); free(in1); free(in2);
    }
}

/* Qandtau (sub.f90:8198-8305) */
void or_qandtau(int nibz, int nb, int nk, const int *mapinv, const double *dist, const double *QN,
                const double *QU, double *tauN, double *tauU, double *tautot)
{
    for (int ibz = 0; ibz < nibz; ibz++)
        for (int ii = 0; ii < nb; ii++) {
            double nb1 = dist[mapinv[ibz] + (size_t)nk * ii];
            size_t x = ibz + (size_t)nibz * ii;
            tauN[x] = QN[x] / (nb1 * (nb1 + 1)) * C_LIGHT * 100;
            tauU[x] = QU[x] / (nb1 * (nb1 + 1)) * C_LIGHT * 100;
            tautot[x] = tauN[x] + tauU[x];
        }
}

/* calculate_RTA_distributionfunction (sub.f90:1238-1289): FFF (nibz,nb,3) */
void or_frta(int nibz, int nb, int nk, const int *mapinv, const double *freq, const double *dist,
             const double *veloc, const double *Q, double tempk, double *FFF)
{
    for (int i = 0; i < nibz; i++)
        for (int ii = 0; ii < nb; ii++) {
            int m = mapinv[i];
            double omega = freq[m + (size_t)nk * ii], nb_ = dist[m + (size_t)nk * ii];
            for (int a = 0; a < 3; a++)
                FFF[i + (size_t)nibz * (ii + (size_t)nb * a)] =
                    -(C_LIGHT * veloc[a + 3 * (ii + (size_t)nb * m)]) * H_PLANK * (omega * C_LIGHT * 100) * nb_ * (nb_ + 1) /
                    (K_B * tempk * tempk) / (Q[i + (size_t)nibz * ii] * C_LIGHT * 100);
        }
}

static void opmul(const double *opk, int s, const double *f, double *o)
{
    for (int r = 0; r < 3; r++) {
        double acc = 0;
        for (int c = 0; c < 3; c++) acc = acc + opk[9 * s + 3 * c + r] * f[c];
        o[r] = acc;
    }
}

/* mode_kappa (sub.f90:1292-1401): kappa (nb,3,3), tauinv_eff (nibz,nb) */
void or_mode_kappa(int nibz, int nb, int nk, const double *kibz, const int *mapinv, const double *kp,
                   const int *nc, const double *g, const double *primlatt, int nops, const double *opk,
                   const double *eigenval, const double *veloc, const double *FFF, double tempk,
                   int classical, double volume_r, double *kappa, double *taueff)
{
    double temp = or_temp_cm(tempk);
    double const1 = H_PLANK * C_LIGHT * 100 * C_LIGHT / (nk * volume_r / 1e30);
    double const2 = K_B * tempk * tempk / (H_PLANK * C_LIGHT * 100) / C_LIGHT;
    memset(kappa, 0, sizeof(double) * nb * 9);
    for (int i = 0; i < nibz; i++) {
        double stars[144];
        int ops[48], n0s[48];
        int narms = or_kstarsinfbz1(kibz + 3 * i, nk, kp, primlatt, nops, opk, stars, ops, n0s);
        for (int ii = 0; ii < nb; ii++) {
            memset(taueff, 0, sizeof(double) * nibz * nb);
            for (int s = 0; s < narms; s++) {
                double f[3], v[3], F12[3], velocc[3];
                for (int a = 0; a < 3; a++) {
                    f[a] = FFF[i + (size_t)nibz * (ii + (size_t)nb * a)];
                    v[a] = veloc[a + 3 * (ii + (size_t)nb * mapinv[i])];
                }
                opmul(opk, ops[s], f, F12);
                opmul(opk, ops[s], v, velocc);
                int n0;
                or_get_k_info(stars + 3 * s, nc, g, &n0, NULL);
                double omega = sqrt(eigenval[ii + (size_t)nb * n0]) * cnst_();
                double nb0 = or_nbe(omega, temp, classical);
                for (int j = 0; j < 3; j++) {
                    for (int jj = 0; jj < 3; jj++)
                        kappa[ii + nb * (j + 3 * jj)] = kappa[ii + nb * (j + 3 * jj)] - const1 * omega * velocc[j] * nb0 * (nb0 + 1) * F12[jj];
                    size_t x = i + (size_t)nibz * ii;
                    if (velocc[j] == 0.0)
                        taueff[x] = taueff[x] - const2 * F12[j] / omega / (double)0.01f;   /* REAL*4 literal */
                    else
                        taueff[x] = taueff[x] - const2 * F12[j] / omega / velocc[j] / (narms * 3.0);
                }
            }
        }
    }
}

/* CollisionMatrix_mpi (sub.f90:3034-3210): CM (ksub*nb*3, nibz*nb*3) column-major;
 * P1,P2 (ksub,nk,nb,nb,nb) with occupations applied. */
/* CollisionMatrix_mpi (sub.f90:3034-3210), or with Piso != NULL CollisionMatrix_iso
 * (sub.f90:3387-3570), the older form used when iso = 1: off the diagonal
 *   C = sum_l3 -P1(q0) + P1(q1+q0) + P2(q0)   plus Piso(n1,l1),
 * on it (q0 = q1, l2 = l1) the same diagonal sum minus Piso(n1,l1).  Piso (nibz,nb). */
void or_collision(int ibz_lo, int ibz_hi, int nibz, const double *kibz, const int *mapinv, int nk,
                  const double *kp, const int *nc, const double *g, const double *primlatt, int nops,
                  const double *opk, int nb, const double *P1, const double *P2, const double *Piso, double *CM)
{
    int ksub = ibz_hi - ibz_lo + 1;
    size_t nrow = (size_t)ksub * nb * 3, ncol = (size_t)nibz * nb * 3;
    memset(CM, 0, sizeof(double) * nrow * ncol);
    int *narm = malloc(sizeof(int) * nibz), *sop = malloc(sizeof(int) * 48 * nibz), *sn0 = malloc(sizeof(int) * 48 * nibz);
    double *sq = malloc(sizeof(double) * 144 * nibz);
    for (int n2 = 0; n2 < nibz; n2++) {
        narm[n2] = or_kstarsinfbz1(kibz + 3 * n2, nk, kp, primlatt, nops, opk, sq + 144 * n2, sop + 48 * n2, NULL);
        for (int s = 0; s < narm[n2]; s++) or_get_k_info(sq + 144 * n2 + 3 * s, nc, g, &sn0[48 * n2 + s], NULL);
    }
#define PX(A, n1, n2, l1, l2, l3) A[(size_t)(n1) + (size_t)ksub * ((size_t)(n2) + (size_t)nk * ((l1) + (size_t)nb * ((l2) + (size_t)nb * (l3))))]
#pragma omp parallel for collapse(2) schedule(dynamic)
    for (int n1 = 0; n1 < ksub; n1++)
        for (int l1 = 0; l1 < nb; l1++) {
            int n1i = mapinv[n1 + ibz_lo];
            size_t d1 = 3 * ((size_t)n1 * nb + l1);
            for (int n2 = 0; n2 < nibz; n2++)
                for (int l2 = 0; l2 < nb; l2++) {
                    size_t d2 = 3 * ((size_t)n2 * nb + l2);
                    for (int ns = 0; ns < narm[n2]; ns++) {
                        int n0 = sn0[48 * n2 + ns];
                        double C = 0;
                        if (n1i == n0 && l1 == l2) {
                            for (int nl2 = 0; nl2 < nb; nl2++)
                                for (int l3 = 0; l3 < nb; l3++)
                                    for (int n22 = 0; n22 < nk; n22++)
                                        C = C + (-PX(P1, n1, n22, l1, nl2, l3) - 0.5 * PX(P2, n1, n22, l1, nl2, l3));
                            if (Piso) C = C - Piso[n1 + ibz_lo + (size_t)nibz * l1];
                        } else if (Piso) {
                            const double *q1 = kibz + 3 * (n1 + ibz_lo), *q2 = sq + 144 * n2 + 3 * ns;
                            double a[3] = {q1[0] + q2[0], q1[1] + q2[1], q1[2] + q2[2]};
                            int p1;
                            or_get_k_info(a, nc, g, &p1, NULL);
                            for (int nl3 = 0; nl3 < nb; nl3++)
                                C = C + (-PX(P1, n1, n0, l1, l2, nl3) + PX(P1, n1, p1, l1, l2, nl3) + PX(P2, n1, n0, l1, l2, nl3));
                            C = C + Piso[n1 + ibz_lo + (size_t)nibz * l1];
                        } else {
                            const double *q1 = kibz + 3 * (n1 + ibz_lo), *q2 = sq + 144 * n2 + 3 * ns;
                            double a[3] = {q1[0] - q2[0], q1[1] - q2[1], q1[2] - q2[2]};
                            double b[3] = {q2[0] - q1[0], q2[1] - q1[1], q2[2] - q1[2]};
                            int p1, p2;
                            or_get_k_info(a, nc, g, &p1, NULL);
                            or_get_k_info(b, nc, g, &p2, NULL);
                            for (int nl3 = 0; nl3 < nb; nl3++)
                                C = C + (-PX(P1, n1, n0, l1, l2, nl3) + PX(P1, n1, p2, l1, l2, nl3) + 0.5 * PX(P2, n1, n0, l1, l2, nl3) +
                                         0.5 * PX(P2, n1, p1, l1, l2, nl3));
                        }
                        int op = sop[48 * n2 + ns];
                        for (int i = 0; i < 3; i++)
                            for (int j = 0; j < 3; j++)
                                CM[(d1 + i) + nrow * (d2 + j)] = CM[(d1 + i) + nrow * (d2 + j)] + (C * opk[9 * op + 3 * j + i]);
                    }
                }
        }
#undef PX
    free(narm); free(sop); free(sn0); free(sq);
}

/* update_mpi (sub.f90:2295-2449): one update of the iterative BTE for the irreducible points
 * ibz_lo..hi,  F2(ibz,l1,x) = F_RTA(n1,l1,x) + A(ibz,l1,x) / Q(n1,l1)  with
 *   A = sum_{n2 in IBZ} sum_{q2 in star(n2) on the mesh} sum_{l2,l3}
 *         P1(ibz,q2,l1,l2,l3) (F(q1+q2,l3) - F(q2,l2))_x + 0.5 P2(ibz,q2,l1,l2,l3) (F(q2,l2) + F(q1-q2,l3))_x
 * where F at a mesh point q is op_kmatrix(s) F(IBZ point of q): s = kvecop of q2 in the star of
 * n2 (kstarsinfbz1), and for q3 = q1 +- q2 the op get_kvecopn2 (sub.f90:2045-2111) finds, the
 * first member of the star of mapibz(q3) that is the mesh point q3.  P1, P2 (ksub,nk,nb,nb,nb)
 * carry the occupations (calculate_tauRTA_mpi2); FFF, FRTA (nibz,nb,3); Q (nibz,nb);
 * F2 (ksub,nb,3); mapibz 0-based.  Returns -1 if some q3 is not a member of its star (the
 * reference would then reuse the op of an earlier q3), else 0. */
int or_update(int ibz_lo, int ibz_hi, int nibz, const double *kibz, const int *mapibz, int nk,
              const double *kp, const int *nc, const double *g, const double *primlatt, int nops,
              const double *opk, int nb, const double *P1, const double *P2, const double *FFF,
              const double *FRTA, const double *Q, const double *Piso, double *F2)
{
    int ksub = ibz_hi - ibz_lo + 1, bad = 0;
    int *narm = malloc(sizeof(int) * nibz), *sop = malloc(sizeof(int) * 48 * nibz), *sn0 = malloc(sizeof(int) * 48 * nibz);
    double *sq = malloc(sizeof(double) * 144 * nibz);
    for (int n2 = 0; n2 < nibz; n2++)
        narm[n2] = or_kstarsinfbz1(kibz + 3 * n2, nk, kp, primlatt, nops, opk, sq + 144 * n2, sop + 48 * n2, sn0 + 48 * n2);
    /* get_kvecopn2 for every mesh point: first star member of its IBZ point that lands on it */
    int *opq = malloc(sizeof(int) * nk);
    for (int q = 0; q < nk; q++) {
        int n = mapibz[q];
        opq[q] = -1;
        for (int s = 0; s < narm[n]; s++)
            if (sn0[48 * n + s] == q) {
                opq[q] = sop[48 * n + s];
                break;
            }
    }
#define PX(A, n1, n2, l1, l2, l3) A[(size_t)(n1) + (size_t)ksub * ((size_t)(n2) + (size_t)nk * ((l1) + (size_t)nb * ((l2) + (size_t)nb * (l3))))]
#define FX(n, l, x) FFF[(size_t)(n) + (size_t)nibz * ((l) + (size_t)nb * (x))]
#pragma omp parallel for collapse(2) schedule(dynamic) reduction(| : bad)
    for (int ibz = 0; ibz < ksub; ibz++)
        for (int l1 = 0; l1 < nb; l1++) {
            const int n1 = ibz + ibz_lo;
            const double *q1 = kibz + 3 * n1;
            for (int xyz = 0; xyz < 3; xyz++) {
                double A = 0;
                for (int n2 = 0; n2 < nibz; n2++)
                    for (int s = 0; s < narm[n2]; s++) {
                        const double *q2 = sq + 144 * n2 + 3 * s;
                        double a[3] = {q1[0] + q2[0], q1[1] + q2[1], q1[2] + q2[2]};
                        double b[3] = {q1[0] - q2[0], q1[1] - q2[1], q1[2] - q2[2]};
                        int ns, p1, p2;
                        or_get_k_info(q2, nc, g, &ns, NULL);
                        or_get_k_info(a, nc, g, &p1, NULL);
                        or_get_k_info(b, nc, g, &p2, NULL);
                        const int i1 = mapibz[p1], i2 = mapibz[p2], sq2 = sop[48 * n2 + s];
                        const int sp1f = opq[p1], sp2f = opq[p2];
                        if (sp1f < 0 || sp2f < 0) {
                            bad = 1;
                            continue;
                        }
                        for (int l2 = 0; l2 < nb; l2++)
                            for (int l3 = 0; l3 < nb; l3++) {
                                double f[3], fq2[3], fp1[3], fp2[3];
                                for (int c = 0; c < 3; c++) f[c] = FX(n2, l2, c);
                                opmul(opk, sq2, f, fq2);
                                for (int c = 0; c < 3; c++) f[c] = FX(i1, l3, c);
                                opmul(opk, sp1f, f, fp1);
                                for (int c = 0; c < 3; c++) f[c] = FX(i2, l3, c);
                                opmul(opk, sp2f, f, fp2);
                                A = A + PX(P1, ibz, ns, l1, l2, l3) * (fp1[xyz] - fq2[xyz]) +
                                    0.5 * PX(P2, ibz, ns, l1, l2, l3) * (fq2[xyz] + fp2[xyz]);
                            }
                    }
                if (Piso) {
                    /* update_iso (sub.f90:2456-2610) adds Piso(n1,l1) Fsq2(xyz) after the loops,
                     * Fsq2 as the loops left it: F(nibz, ndn) rotated by the op of the last
                     * star member of the last irreducible point */
                    const int nl = nibz - 1;
                    double f[3], fq2[3];
                    for (int c = 0; c < 3; c++) f[c] = FX(nl, nb - 1, c);
                    opmul(opk, sop[48 * nl + narm[nl] - 1], f, fq2);
                    A = A + Piso[n1 + (size_t)nibz * l1] * fq2[xyz];
                }
                F2[ibz + (size_t)ksub * (l1 + (size_t)nb * xyz)] =
                    FRTA[n1 + (size_t)nibz * (l1 + (size_t)nb * xyz)] + A / Q[n1 + (size_t)nibz * l1];
            }
        }
#undef PX
#undef FX
    free(narm); free(sop); free(sn0); free(sq); free(opq);
    return bad ? -1 : 0;
}

/* check_inside_bz (extratools.f90:325-341): 1 if q + (g1+g2+g3)/2 has reduced coordinates in
 * [0,1) (get_direct_components, :561-592, epsl = 5e-10), i.e. a normal process */
static int check_inside_bz(const double *q, const double *g)
{
    double rr[9], q2[3];
    recip_nopi(g, rr);
    for (int c = 0; c < 3; c++) q2[c] = q[c] + 0.5 * (g[c] + g[3 + c] + g[6 + c]);
    int inside = 1;
    for (int i = 0; i < 3; i++) {
        double x = dot3(q2, rr + 3 * i) + 5e-10;
        if (x < 0 || x >= 1) inside = 0;
    }
    return inside;
}

/* function_self_w3 (rest.f90:4012-4124) with matrix_elt_full (:3664-3719): the three-phonon
 * self-energy of the mode (q, branch with eigenvector evq and frequency omq) at omega and at
 * temp (both cm^-1), summed over the mesh k2 with k3 = -q-k2 unfolded.  Mesh: kp (nk,3),
 * eigenval (nb,nk), eigenvec (nb,nb,nk) [component, branch, k]; at every k3 the get_freq data
 * eival3 (nb,nk) and eivec3 (nb,nb,nk).  eta = etaz; v3thr = v3_threshold.  nself, uself:
 * complex (re, im) of the normal and umklapp parts. */
void or_self_w3(const double *q, int nb, double omq, const double *evq_ri, double omega, double temp,
                int classical, double eta, double v3thr, int nk, const double *kp, const double *g,
                const double *eigenval, const double *eigenvec_ri, const double *eival3, const double *eivec3_ri,
                const double *mass0, const double *atompos, const int *iatomcell0, int nt3, const int *at3,
                const int *xyz3, const double *fc3, int kind, double *nself, double *uself)
{
    /* kind 1: function_self_w4 (rest.f90:4126-4240), the imaginary part alone with Gaussian
     * deltas delta_g(arg, eta) (extratools.f90:239-246) of the three processes, cut at 6 eta */
    const zc *evq = (const zc *)evq_ri, *ev = (const zc *)eigenvec_ri, *ev3a = (const zc *)eivec3_ri;
    const double etacut = kind == 1 ? eta * 6 : eta * 3000, c33 = or_const33();
    const float cf = 521.1098918f;
    const double cnst3 = (double)(cf * cf * cf);   /* cnst*cnst*cnst in REAL*4 (rest.f90:3713) */
    zc ns = 0, us = 0;
    for (int ik = 0; ik < nk; ik++) {
        const double *k2 = kp + 3 * ik;
        double k3[3] = {-q[0] - k2[0], -q[1] - k2[1], -q[2] - k2[2]};
        const int inside = check_inside_bz(k3, g);
        for (int l2 = 0; l2 < nb; l2++) {
            const double om2 = sqrt(fabs(eigenval[l2 + (size_t)nb * ik])) * cnst_();
            const double nb2 = or_nbe(om2, temp, classical);
            for (int l3 = 0; l3 < nb; l3++) {
                const double om3 = sqrt(fabs(eival3[l3 + (size_t)nb * ik])) * cnst_();
                const double nb3 = or_nbe(om3, temp, classical);
                double ires = 0, rres = 0, arg;
                if (kind == 1) {
                    arg = omega - om3 - om2;
                    if (fabs(arg) < etacut) ires = (nb3 + nb2 + 1) * or_delta_g(arg, eta);
                    arg = omega - om3 + om2;
                    if (fabs(arg) < etacut) ires = ires + (nb2 - nb3) * or_delta_g(arg, eta);
                    arg = omega - om2 + om3;
                    if (fabs(arg) < etacut) ires = ires + (nb3 - nb2) * or_delta_g(arg, eta);
                    goto selfdone;
                }
                arg = omega - om3 - om2;
                if (fabs(arg) < etacut) {
                    ires = (nb3 + nb2 + 1) * (eta / (arg * arg + eta * eta) / PI);
                    rres = (nb3 + nb2 + 1) * arg / (arg * arg + eta * eta);
                }
                arg = omega + om3 + om2;
                if (fabs(arg) < etacut) {
                    ires = ires - (nb3 + nb2 + 1) * (eta / (arg * arg + eta * eta) / PI);
                    rres = rres - (nb3 + nb2 + 1) * arg / (arg * arg + eta * eta);
                }
                arg = omega - om3 + om2;
                if (fabs(arg) < etacut) {
                    ires = ires + (nb2 - nb3) * (eta / (arg * arg + eta * eta) / PI);
                    rres = rres + (nb2 - nb3) * arg / (arg * arg + eta * eta);
                }
                arg = omega - om2 + om3;
                if (fabs(arg) < etacut) {
                    ires = ires + (nb3 - nb2) * (eta / (arg * arg + eta * eta) / PI);
                    rres = rres + (nb3 - nb2) * arg / (arg * arg + eta * eta);
                }
            selfdone:;
                const zc self = rres + I * (ires * PI);
                /* matrix_elt_full */
                zc xx = 0;
                for (int t = 0; t < nt3; t++) {
                    int i3 = at3[3 * t], j1 = at3[3 * t + 1], k1 = at3[3 * t + 2];
                    int i0 = iatomcell0[i3 - 1];
                    if (i0 != i3) continue;
                    int j0 = iatomcell0[j1 - 1], k0 = iatomcell0[k1 - 1];
                    double mi = mass0[i0 - 1], mj = mass0[j0 - 1], mk = mass0[k0 - 1];
                    int ta1 = xyz3[3 * t] - 1 + 3 * (i0 - 1), ta2 = xyz3[3 * t + 1] - 1 + 3 * (j0 - 1),
                        ta3 = xyz3[3 * t + 2] - 1 + 3 * (k0 - 1);
                    double ph = 0, ph3 = 0;
                    for (int c = 0; c < 3; c++) {
                        ph += k2[c] * (atompos[3 * (j1 - 1) + c] - atompos[3 * (j0 - 1) + c]);
                        ph3 += k3[c] * (atompos[3 * (k1 - 1) + c] - atompos[3 * (k0 - 1) + c]);
                    }
                    zc eiqr = cexp(I * (ph + ph3));
                    xx = xx + fc3[t] / sqrt(mi * mj * mk) * eiqr * evq[ta1] * ev[ta2 + (size_t)nb * (l2 + (size_t)nb * ik)] *
                                  ev3a[ta3 + (size_t)nb * (l3 + (size_t)nb * ik)];
                }
                const double den = sqrt(8 * omq * om2 * om3 / cnst3);
                if (den == 0) continue;   /* the reference stops here (MATRIX_ELT: den=0) */
                const zc w33 = xx / den * c33;
                if (cabs(w33) < v3thr) continue;
                const zc term = w33 * conj(w33);
                if (inside == 1) ns = ns + term * self;
                else us = us + term * self;
            }
        }
    }
    ns = ns / nk * PI;
    us = us / nk * PI;
    nself[0] = creal(ns); nself[1] = cimag(ns);
    uself[0] = creal(us); uself[1] = cimag(us);
}

/* cal_RHS_IBZ (sub.f90:2618-2691) */
void or_rhs(int nibz, int nb, int nk, const int *mapinv, const double *freq, const double *dist,
            const double *veloc, double tempk, double *RHS)
{
    size_t s = 0;
    for (int i = 0; i < nibz; i++) {
        int n0 = mapinv[i];
        for (int ii = 0; ii < nb; ii++) {
            double omega = freq[n0 + (size_t)nk * ii], nb0 = dist[n0 + (size_t)nk * ii];
            for (int a = 0; a < 3; a++)
                RHS[s++] = (C_LIGHT * veloc[a + 3 * (ii + (size_t)nb * n0)]) * H_PLANK * (omega * C_LIGHT * 100) * nb0 * (nb0 + 1) /
                           (K_B * tempk * tempk) / (C_LIGHT * 100);
        }
    }
}

/* cg_quadratic (cg.f90:19-69); returns iterations */
int or_cg(int n, const double *a, const double *b, double *vmin, int max_iter, int ndig, double *funcmin)
{
    double *conj = malloc(sizeof(double) * n), *grad = malloc(sizeof(double) * n);
    double *aux = malloc(sizeof(double) * n), *v0 = malloc(sizeof(double) * n);
    double tol = 1.0 / pow(10.0, 2 * ndig);
    int iter = 0;
#define MATVEC(out, vec)                                                  \
    do {                                                                  \
        for (int i_ = 0; i_ < n; i_++) out[i_] = 0;                       \
        for (int j_ = 0; j_ < n; j_++) {                                  \
            double vj_ = vec[j_];                                         \
            const double *col_ = a + (size_t)n * j_;                      \
            for (int i_ = 0; i_ < n; i_++) out[i_] = out[i_] + col_[i_] * vj_; \
        }                                                                 \
    } while (0)
#define DOT(x, y, r) do { double s_ = 0; for (int i_ = 0; i_ < n; i_++) s_ = s_ + x[i_] * y[i_]; r = s_; } while (0)
    for (int i = 0; i < n; i++) v0[i] = vmin[i];
    MATVEC(grad, v0);
    for (int i = 0; i < n; i++) { grad[i] = grad[i] - b[i]; conj[i] = -grad[i]; }
    while (iter < max_iter) {
        iter++;
        double num, den, v2, cc, err;
        MATVEC(aux, conj);
        DOT(grad, conj, num);
        DOT(aux, conj, den);
        for (int i = 0; i < n; i++) v0[i] = grad[i];
        for (int i = 0; i < n; i++) vmin[i] = vmin[i] - num / den * conj[i];
        for (int i = 0; i < n; i++) grad[i] = grad[i] - num / den * aux[i];
        DOT(vmin, vmin, v2);
        DOT(conj, conj, cc);
        err = (num * num / den / den) * cc / v2;
        if (err < tol) break;
        DOT(grad, grad, num);
        DOT(v0, v0, den);
        for (int i = 0; i < n; i++) conj[i] = -grad[i] + num / den * conj[i];
    }
    MATVEC(v0, vmin);
    double t1, t2;
    DOT(vmin, v0, t1);
    DOT(vmin, b, t2);
    if (funcmin) *funcmin = 0.5 * t1 - t2;
#undef MATVEC
#undef DOT
    free(conj); free(grad); free(aux); free(v0);
    return iter;
}

/* veloc_avergeIBZ_invsyFBZ (sub.f90:5479-5537) + mode_kappa_d (sub.f90:1407-1530).
 * taueff accumulates onto its input, as the reference's module array does. */
void or_mode_kappa_d(int nibz, int nb, int nk, const double *kibz, const int *mapinv, const double *kp,
                     const int *nc, const double *g, const double *primlatt, int nops, const double *opk,
                     const double *eigenval, const double *veloc, const double *x, double tempk,
                     int classical, double volume_r, double *kappa, double *taueff)
{
    double temp = or_temp_cm(tempk);
    double const1 = H_PLANK * C_LIGHT * 100 * C_LIGHT / (nk * volume_r / 1e30);
    double const2 = K_B * tempk * tempk / (H_PLANK * C_LIGHT * 100) / C_LIGHT;
    double *veloci = calloc((size_t)3 * nb * nibz, sizeof(double));
    for (int i = 0; i < nibz; i++) {
        double st[144];
        int ops[48];
        int narms = or_getstar(kibz + 3 * i, primlatt, nops, opk, st, ops);
        for (int j = 0; j < nb; j++) {
            double va[3] = {0, 0, 0};
            for (int n = 0; n < narms; n++) {
                int n0;
                or_get_k_info(st + 3 * n, nc, g, &n0, NULL);
                double inv[9], vi[3];
                inv3(opk + 9 * ops[n], inv);
                const double *v = veloc + 3 * (j + (size_t)nb * n0);
                for (int r = 0; r < 3; r++) {
                    double acc = 0;
                    for (int c = 0; c < 3; c++) acc = acc + inv[3 * c + r] * v[c];
                    vi[r] = acc;
                }
                for (int r = 0; r < 3; r++) va[r] = va[r] + (vi[r] / narms);
            }
            for (int r = 0; r < 3; r++) veloci[r + 3 * (j + (size_t)nb * i)] = va[r];
        }
    }
    memset(kappa, 0, sizeof(double) * nb * 9);
    for (int i = 0; i < nibz; i++) {
        double stars[144];
        int ops[48];
        int narms = or_kstarsinfbz1(kibz + 3 * i, nk, kp, primlatt, nops, opk, stars, ops, NULL);
        for (int ii = 0; ii < nb; ii++)
            for (int s = 0; s < narms; s++) {
                double f[3], F12[3], velocc[3];
                for (int a = 0; a < 3; a++) f[a] = x[3 * ((size_t)i * nb + ii) + a];
                opmul(opk, ops[s], f, F12);
                opmul(opk, ops[s], veloci + 3 * (ii + (size_t)nb * i), velocc);
                int n0;
                or_get_k_info(stars + 3 * s, nc, g, &n0, NULL);
                double omega = sqrt(eigenval[ii + (size_t)nb * n0]) * cnst_();
                double nb0 = or_nbe(omega, temp, classical);
                for (int j = 0; j < 3; j++) {
                    for (int jj = 0; jj < 3; jj++)
                        kappa[ii + nb * (j + 3 * jj)] = kappa[ii + nb * (j + 3 * jj)] - const1 * omega * velocc[j] * nb0 * (nb0 + 1) * F12[jj];
                    double fv = F12[0] * velocc[0] + F12[1] * velocc[1] + F12[2] * velocc[2];
                    double vv = velocc[0] * velocc[0] + velocc[1] * velocc[1] + velocc[2] * velocc[2];
                    size_t xi = i + (size_t)nibz * ii;
                    taueff[xi] = taueff[xi] - const2 * fv / omega / vv / (narms * 3.0);
                }
            }
    }
    free(veloci);
}
