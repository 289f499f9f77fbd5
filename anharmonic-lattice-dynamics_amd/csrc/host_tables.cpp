// Host-side tables of the three-phonon path: lattice point group, q-mesh, irreducible
// zone, stars and the FC3 term layouts the kernels read.
//
// The q-mesh work is exact integer arithmetic: every q the path touches is a mesh point,
// so a point is represented by integer coordinates c in [0,n) (reduced coordinate
// c/n - 1/2 on the even, Gamma-containing meshes THERMACOND uses) and symmetry operations
// act as integer matrices on 2*lcm(n)-scaled reduced coordinates.  This reproduces the
// reference's floating-point get_k_info/getstar/get_weights decisions (tolerances 1e-4..
// 1e-6 against mesh spacings >= 1e-2) without their cost; tests pin every table against
// integer dumps of the reference.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <numeric>
#include <tuple>

#include "alt3_impl.h"

namespace alt3 {

double const33()
{  // kappa.f90:182
    return 1e30 * std::pow(HBAR / UMA, 1.5) / std::pow(std::sqrt(EE * 1e20 / UMA), 1.5) * EE / H_PLANK / 100 / C_LIGHT;
}

namespace {

void cross3(const double *a, const double *b, double *c)
{
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}
double dot3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
double len3(const double *v) { return std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); }

// column-major 3x3 helpers
void mat_mul(const double *a, const double *b, double *c)
{
    double t[9];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            double s = 0;
            for (int k = 0; k < 3; k++) s = s + a[3 * k + i] * b[3 * j + k];
            t[3 * j + i] = s;
        }
    std::memcpy(c, t, sizeof(t));
}
bool mat_inv(const double *ain, double *aout)
{  // Gauss-Jordan with partial pivoting
    double a[3][6];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            a[i][j] = ain[3 * j + i];
            a[i][3 + j] = (i == j) ? 1.0 : 0.0;
        }
    for (int k = 0; k < 3; k++) {
        int p = k;
        for (int i = k + 1; i < 3; i++)
            if (std::fabs(a[i][k]) > std::fabs(a[p][k])) p = i;
        if (a[p][k] == 0) return false;
        if (p != k)
            for (int j = 0; j < 6; j++) std::swap(a[p][j], a[k][j]);
        double d = a[k][k];
        for (int j = 0; j < 6; j++) a[k][j] /= d;
        for (int i = 0; i < 3; i++)
            if (i != k) {
                double f = a[i][k];
                for (int j = 0; j < 6; j++) a[i][j] -= f * a[k][j];
            }
    }
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) aout[3 * j + i] = a[i][3 + j];
    return true;
}

// Point group of a lattice (force_constants.f90:1379-1525, dlatmat2): integer matrices in
// the primitive basis, found by mapping the basis onto lattice vectors of equal length and
// closed under multiplication; the identity is kept first.  The enumeration order fixes
// the op numbering that op_kmatrix, stars and the collision matrix use.
int lattice_point_group(const double *cart /*col j = prim vector j*/, double eps, std::vector<std::array<int, 9>> &mats)
{
    struct Short { int i[3]; double v[3]; };
    std::vector<Short> sh[3];
    double abc[3][3];
    for (int i = 0; i < 3; i++) abc[i][i] = len3(cart + 3 * i);
    for (int i = 1; i < 3; i++)
        for (int j = 0; j < i; j++) {
            double v[3];
            for (int k = 0; k < 3; k++) v[k] = cart[3 * i + k] - cart[3 * j + k];
            abc[i][j] = abc[j][i] = len3(v);
        }
    double xmax = std::max(std::max(abc[0][0], abc[1][1]), abc[2][2]) + eps;
    bool more = true;
    for (int shell = 1; more; shell++) {
        more = false;
        for (int j1 = -shell; j1 <= shell; j1++)
            for (int j2 = -shell; j2 <= shell; j2++)
                for (int j3 = -shell; j3 <= shell; j3++) {
                    if (std::abs(j1) != shell && std::abs(j2) != shell && std::abs(j3) != shell) continue;
                    double v[3] = {0, 0, 0};
                    for (int k = 0; k < 3; k++) {
                        v[k] = v[k] + j1 * cart[k];
                        v[k] = v[k] + j2 * cart[3 + k];
                        v[k] = v[k] + j3 * cart[6 + k];
                    }
                    double d = len3(v);
                    if (d < xmax) more = true;
                    for (int k = 0; k < 3; k++)
                        if (std::fabs(d - abc[k][k]) < eps) sh[k].push_back(Short{{j1, j2, j3}, {v[0], v[1], v[2]}});
                }
    }
    mats.clear();
    mats.push_back({1, 0, 0, 0, 1, 0, 0, 0, 1});
    for (auto &a : sh[0])
        for (auto &b : sh[1])
            for (auto &c : sh[2]) {
                std::array<int, 9> m = {a.i[0], a.i[1], a.i[2], b.i[0], b.i[1], b.i[2], c.i[0], c.i[1], c.i[2]};
                int det = m[0] * (m[4] * m[8] - m[7] * m[5]) - m[3] * (m[1] * m[8] - m[7] * m[2]) + m[6] * (m[1] * m[5] - m[4] * m[2]);
                if (std::abs(det) != 1) continue;
                const Short *s[3] = {&a, &b, &c};
                bool ok = true;
                for (int x = 1; x < 3 && ok; x++)
                    for (int y = 0; y < x; y++) {
                        double v[3];
                        for (int k = 0; k < 3; k++) v[k] = s[x]->v[k] - s[y]->v[k];
                        if (std::fabs(len3(v) - abc[x][y]) > eps) { ok = false; break; }
                    }
                if (!ok) continue;
                if (m[0] + m[4] + m[8] == 3) { mats[0] = m; continue; }
                if (mats.size() >= 48) return -1;
                mats.push_back(m);
            }
    std::vector<std::vector<char>> tried(48, std::vector<char>(48, 0));
    bool found = true;
    while (found) {
        found = false;
        for (size_t x = 0; x < mats.size(); x++)
            for (size_t y = 0; y < mats.size(); y++) {
                if (tried[x][y]) continue;
                tried[x][y] = 1;
                std::array<int, 9> t{};
                for (int r = 0; r < 3; r++)
                    for (int c = 0; c < 3; c++) {
                        int s = 0;
                        for (int k = 0; k < 3; k++) s += mats[x][3 * k + r] * mats[y][3 * c + k];
                        t[3 * c + r] = s;
                    }
                if (std::find(mats.begin(), mats.end(), t) == mats.end()) {
                    if (mats.size() >= 48) return -1;
                    mats.push_back(t);
                    found = true;
                }
            }
    }
    return (int)mats.size();
}

long long lcm_ll(long long a, long long b) { return a / std::gcd(a, b) * b; }

}  // namespace

int build_tables(const alt3_system *sys, HostTables &t)
{
    const int *nc = sys->nc;
    for (int a = 0; a < 3; a++)
        if (nc[a] <= 0 || nc[a] % 2 != 0)
            return fail(ALT3_E_MESH, "mesh must be even in every direction (make_kp_reg_tet corners are mesh points only then)");
    // the device index arithmetic divides by float reciprocals, exact for n <= 256 and nk < 4e6
    if (nc[0] > 256 || nc[1] > 256 || nc[2] > 256 || (long long)nc[0] * nc[1] * nc[2] >= 4000000)
        return fail(ALT3_E_MESH, "mesh too large: at most 256 points per direction and 4e6 in total");
    const int nk = nc[0] * nc[1] * nc[2];
    t.nk = nk;
    // reciprocal lattice, make_reciprocal_lattice_2pi (extratools.f90:209-228)
    const double *r = sys->r123;
    double c[3];
    cross3(r + 3, r + 6, c);
    double om = dot3(r, c);
    for (int i = 0; i < 3; i++) t.g[i] = 2 * PI * c[i] / om;
    cross3(r + 6, r, c);
    for (int i = 0; i < 3; i++) t.g[3 + i] = 2 * PI * c[i] / om;
    cross3(r, r + 3, c);
    for (int i = 0; i < 3; i++) t.g[6 + i] = 2 * PI * c[i] / om;
    t.volume_r = std::fabs(om);
    const double *g = t.g;

    // lattice point group -> op_kmatrix (force_constants.f90:50-186)
    {
        double latp[6];
        std::memcpy(latp, sys->latparams, sizeof(latp));
        const double d2r = 4.0 * std::atan(1.0) / 180.0;
        for (int i = 3; i < 6; i++) latp[i] *= d2r;
        double c2c[9] = {0};
        c2c[0] = latp[0];
        c2c[3] = latp[1] * std::cos(latp[5]);
        c2c[4] = latp[1] * std::sin(latp[5]);
        c2c[6] = latp[2] * std::cos(latp[4]);
        c2c[7] = latp[2] * (std::cos(latp[3]) - std::cos(latp[5]) * std::cos(latp[4])) / std::sin(latp[5]);
        c2c[8] = std::sqrt(latp[2] * latp[2] - c2c[6] * c2c[6] - c2c[7] * c2c[7]);
        double p2c[9], c2p[9];
        mat_mul(c2c, sys->primlatt, p2c);
        if (!mat_inv(sys->primlatt, c2p)) return fail(ALT3_E_ARG, "primitive lattice is singular");
        std::vector<std::array<int, 9>> mats;
        int n = lattice_point_group(p2c, 1e-5, mats);
        if (n <= 0) return fail(ALT3_E_ARG, "lattice point group: more than 48 operations");
        t.nops = n;
        t.opk.assign(9 * n, 0.0);
        t.opk_inv.assign(9 * n, 0.0);
        for (int m = 0; m < n; m++) {
            double tm[9], ti[9];
            for (int e = 0; e < 9; e++) tm[e] = mats[m][e];
            mat_mul(tm, c2p, tm);
            mat_mul(sys->primlatt, tm, tm);
            if (!mat_inv(tm, ti)) return fail(ALT3_E_ARG, "singular point operation");
            for (int j = 0; j < 3; j++)
                for (int k = 0; k < 3; k++) t.opk[9 * m + 3 * j + k] = (double)std::lround(ti[3 * k + j]);
            if (!mat_inv(&t.opk[9 * m], &t.opk_inv[9 * m])) return fail(ALT3_E_ARG, "singular op_kmatrix");
        }
    }

    // meshes: REAL*4 fractions (make_kp_reg) and FP64 (make_kp_reg_tet), -G/2 shift
    t.kpc_eig.resize(3 * nk);
    t.kpc.resize(3 * nk);
    double sh[3];
    for (int a = 0; a < 3; a++) sh[a] = (-0.5) * (g[a] + g[3 + a] + g[6 + a]);
    for (int i = 0, k = 0; i < nc[0]; i++)
        for (int j = 0; j < nc[1]; j++)
            for (int l = 0; l < nc[2]; l++, k++) {
                double f1 = (double)((float)i / (float)nc[0]);
                double f2 = (double)((float)j / (float)nc[1]);
                double f3 = (double)((float)l / (float)nc[2]);
                for (int a = 0; a < 3; a++) {
                    t.kpc_eig[3 * k + a] = (f1 * g[a] + f2 * g[3 + a] + f3 * g[6 + a] + 0.0) + sh[a];
                    t.kpc[3 * k + a] = (((i + 0.0) / nc[0]) * g[a] + ((j + 0.0) / nc[1]) * g[3 + a] + ((l + 0.0) / nc[2]) * g[6 + a] + 0.0) + sh[a];
                }
            }

    // integer representation: U_a = (2 c_a - n_a) * s_a, s_a = L2 / (2 n_a); u_a = U_a / L2
    long long L2 = 2 * lcm_ll(lcm_ll(nc[0], nc[1]), nc[2]);
    long long s_a[3] = {L2 / (2 * nc[0]), L2 / (2 * nc[1]), L2 / (2 * nc[2])};
    auto coords = [&](int k, int *cc) {
        cc[2] = k % nc[2];
        cc[1] = (k / nc[2]) % nc[1];
        cc[0] = k / (nc[1] * nc[2]);
    };
    auto index_of = [&](const int *cc) { return cc[2] + nc[2] * (cc[1] + nc[1] * cc[0]); };
    // reduced integer op matrices: u' = M u with M = Gmat^-1 S Gmat (Gmat columns = g_i)
    std::vector<std::array<long long, 9>> M(t.nops);
    {
        double G[9], Gi[9];
        for (int i = 0; i < 3; i++)
            for (int a = 0; a < 3; a++) G[3 * i + a] = g[3 * i + a];  // column i = g_i
        if (!mat_inv(G, Gi)) return fail(ALT3_E_ARG, "singular reciprocal lattice");
        for (int m = 0; m < t.nops; m++) {
            double tmp[9], R[9];
            mat_mul(&t.opk[9 * m], G, tmp);
            mat_mul(Gi, tmp, R);
            for (int e = 0; e < 9; e++) {
                long long v = std::llround(R[e]);
                if (std::fabs(R[e] - (double)v) > 1e-6)
                    return fail(ALT3_E_ARG, "point operation is not integral in the reciprocal basis");
                M[m][e] = v;
            }
        }
    }
    using Key = std::array<long long, 3>;
    auto red = [&](int k) {
        int cc[3];
        coords(k, cc);
        return Key{(2LL * cc[0] - nc[0]) * s_a[0], (2LL * cc[1] - nc[1]) * s_a[1], (2LL * cc[2] - nc[2]) * s_a[2]};
    };
    auto apply = [&](int m, const Key &u) {
        Key o;
        for (int r0 = 0; r0 < 3; r0++) {
            long long s = 0;
            for (int cidx = 0; cidx < 3; cidx++) s += M[m][3 * cidx + r0] * u[cidx];
            o[r0] = s;
        }
        return o;
    };
    // exact mesh point (no folding) of a reduced integer vector, or -1
    auto on_mesh = [&](const Key &u) {
        int cc[3];
        for (int a = 0; a < 3; a++) {
            if (u[a] % s_a[a] != 0) return -1;
            long long tt = u[a] / s_a[a] + nc[a];  // = 2 c
            if (tt < 0 || tt >= 2LL * nc[a] || (tt & 1)) return -1;
            cc[a] = (int)(tt / 2);
        }
        return index_of(cc);
    };
    // folded mesh point (get_k_info on an arbitrary mesh vector)
    auto folded = [&](const Key &u) {
        int cc[3];
        for (int a = 0; a < 3; a++) {
            if (u[a] % s_a[a] != 0) return -1;  // not a mesh vector (get_k_info inside=2)
            long long q = u[a] / s_a[a];
            long long tt = q + nc[a];     // 2c, possibly out of range
            if (tt % 2 != 0) return -1;
            long long cval = tt / 2;
            cval %= nc[a];
            if (cval < 0) cval += nc[a];
            cc[a] = (int)cval;
        }
        return index_of(cc);
    };
    // getstar (force_constants.f90:1748-1785): distinct images in op order
    auto star = [&](int k, std::vector<Key> &sv, std::vector<int> &so) {
        sv.clear();
        so.clear();
        Key u = red(k);
        for (int m = 0; m < t.nops; m++) {
            Key v = apply(m, u);
            if (std::find(sv.begin(), sv.end(), v) != sv.end()) continue;
            sv.push_back(v);
            so.push_back(m);
        }
    };

    // get_negatives (rest.f90:3208-3260)
    t.mappos.assign(nk, 0);
    {
        std::vector<char> inc(nk, 0);
        int nps = 0;
        for (int i = 0; i < nk; i++) {
            int cc[3], mc[3];
            coords(i, cc);
            for (int a = 0; a < 3; a++) mc[a] = (nc[a] - cc[a]) % nc[a];
            if (index_of(mc) < i) continue;
            t.mappos[nps++] = i;
            inc[i] = 1;
        }
        t.npos = nps;
        int cnt = nps;
        for (int i = 0; i < nk; i++)
            if (!inc[i]) t.mappos[cnt++] = i;
    }

    // get_weights (modules.f90:2006-2162): first-come IBZ, exact (unfolded) star matching
    t.mapibz.assign(nk, 0);
    t.mapinv.clear();
    std::vector<double> w2;
    std::map<Key, int> ibz_of;
    std::vector<Key> sv;
    std::vector<int> so;
    for (int i = 0; i < nk; i++) {
        star(i, sv, so);
        int hit = -1;
        for (auto &v : sv) {
            auto it = ibz_of.find(v);
            if (it != ibz_of.end()) { hit = it->second; break; }
        }
        if (hit >= 0) {
            w2[hit] += 1;
            t.mapibz[i] = hit;
            continue;
        }
        int j = (int)t.mapinv.size();
        t.mapinv.push_back(i);
        w2.push_back(1.0);
        ibz_of[red(i)] = j;
        t.mapibz[i] = j;
    }
    t.nibz = (int)t.mapinv.size();
    double sw = 0;
    for (double x : w2) sw += x;
    t.wibz.resize(t.nibz);
    t.kibz.resize(3 * t.nibz);
    for (int j = 0; j < t.nibz; j++) {
        t.wibz[j] = w2[j] / sw;
        for (int a = 0; a < 3; a++) t.kibz[3 * j + a] = t.kpc_eig[3 * t.mapinv[j] + a];
    }

    // stars: on-mesh members (kstarsinfbz1) and full folded stars (getstar + get_k_info)
    t.star_n.assign(t.nibz, 0);
    t.star_op.assign(48 * t.nibz, 0);
    t.star_n0.assign(48 * t.nibz, 0);
    t.fstar_n.assign(t.nibz, 0);
    t.fstar_op.assign(48 * t.nibz, 0);
    t.fstar_n0.assign(48 * t.nibz, 0);
    for (int j = 0; j < t.nibz; j++) {
        star(t.mapinv[j], sv, so);
        int ds = 0;
        for (size_t s = 0; s < sv.size(); s++) {
            int m = on_mesh(sv[s]);
            t.fstar_op[48 * j + s] = so[s];
            t.fstar_n0[48 * j + s] = folded(sv[s]);
            if (m < 0) continue;
            t.star_op[48 * j + ds] = so[s];
            t.star_n0[48 * j + ds] = m;
            ds++;
        }
        t.star_n[j] = ds;
        t.fstar_n[j] = (int)sv.size();
    }

    // tetrahedra (make_kp_reg_tet): cube corners pt1..pt8 = c + {0,e1,e2,e1+e2,e3,e3+e1,e3+e2,e3+e1+e2}
    static const int corner[6][4] = {{1, 2, 3, 6}, {2, 3, 4, 6}, {1, 3, 5, 6}, {3, 4, 6, 8}, {3, 5, 6, 7}, {3, 6, 7, 8}};
    static const int off[8][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {1, 1, 0}, {0, 0, 1}, {1, 0, 1}, {0, 1, 1}, {1, 1, 1}};
    t.tet.resize(4 * 6 * (size_t)nk);
    for (int k = 0, nt = 0; k < nk; k++) {
        int cc[3], pt[8];
        coords(k, cc);
        for (int p = 0; p < 8; p++) {
            int q[3];
            for (int a = 0; a < 3; a++) q[a] = (cc[a] + off[p][a]) % nc[a];
            pt[p] = index_of(q);
        }
        for (int l = 0; l < 6; l++, nt++)
            for (int x = 0; x < 4; x++) t.tet[4 * nt + x] = pt[corner[l][x] - 1];
    }
    return ALT3_OK;
}

// FC3 terms in the two layouts the V3 kernels read.
//   V3(q1 l1, q2 l2, q3 l3) * den/const33 = sum_t psi_t e^{i q2.Rj} e^{i q3.Rk} e1[a_t] e2[b_t] e3[c_t]
// (matrix_elt_simplified, sub.f90:5888-6011).  With q3 = -q1-q2 the phase factorises as
// e^{i q2.(Rj-Rk)} e^{-i q1.Rk}, so terms are grouped into slots (dR = Rj-Rk, b, c).
int build_v3_tables(const alt3_system *sys, V3Tables &v)
{
    const int n0 = sys->natoms0;
    std::vector<int> keep;
    for (int t = 0; t < sys->nt3; t++) {
        int i3 = sys->fc3_atoms[3 * t];
        if (i3 < 1 || i3 > sys->natoms) return fail(ALT3_E_ARG, "fc3 atom index out of range");
        if (sys->iatomcell0[i3 - 1] != i3) continue;  // first index must be in the primitive cell
        for (int r = 1; r < 3; r++) {
            int a = sys->fc3_atoms[3 * t + r];
            if (a < 1 || a > sys->natoms) return fail(ALT3_E_ARG, "fc3 atom index out of range");
        }
        for (int r = 0; r < 3; r++)
            if (sys->fc3_xyz[3 * t + r] < 1 || sys->fc3_xyz[3 * t + r] > 3) return fail(ALT3_E_ARG, "fc3 xyz index out of range");
        keep.push_back(t);
    }
    v.nterm = (int)keep.size();
    v.t_abc.resize(3 * v.nterm);
    v.t_psi.resize(v.nterm);
    v.t_rj.resize(3 * v.nterm);
    v.t_rk.resize(3 * v.nterm);
    std::vector<std::array<long long, 3>> dn(v.nterm);
    for (int n = 0; n < v.nterm; n++) {
        int t = keep[n];
        int i3 = sys->fc3_atoms[3 * t], j = sys->fc3_atoms[3 * t + 1], k = sys->fc3_atoms[3 * t + 2];
        int i0 = sys->iatomcell0[i3 - 1], j0 = sys->iatomcell0[j - 1], k0 = sys->iatomcell0[k - 1];
        if (i0 < 1 || i0 > n0 || j0 < 1 || j0 > n0 || k0 < 1 || k0 > n0) return fail(ALT3_E_ARG, "iatomcell0 out of range");
        v.t_abc[3 * n] = sys->fc3_xyz[3 * t] - 1 + 3 * (i0 - 1);
        v.t_abc[3 * n + 1] = sys->fc3_xyz[3 * t + 1] - 1 + 3 * (j0 - 1);
        v.t_abc[3 * n + 2] = sys->fc3_xyz[3 * t + 2] - 1 + 3 * (k0 - 1);
        v.t_psi[n] = sys->fc3[t] / std::sqrt(sys->mass[i0 - 1] * sys->mass[j0 - 1] * sys->mass[k0 - 1]);
        for (int a = 0; a < 3; a++) {
            v.t_rj[3 * n + a] = sys->atompos[3 * (j - 1) + a] - sys->atompos[3 * (j0 - 1) + a];
            v.t_rk[3 * n + a] = sys->atompos[3 * (k - 1) + a] - sys->atompos[3 * (k0 - 1) + a];
        }
        // group key: the lattice vector Rj-Rk rounded to 1e-6 A
        for (int a = 0; a < 3; a++) dn[n][a] = std::llround((v.t_rj[3 * n + a] - v.t_rk[3 * n + a]) * 1e6);
    }
    const int nb = 3 * n0;
    std::map<std::array<long long, 3>, int> drid;
    std::vector<int> term_dr(v.nterm);
    for (int n = 0; n < v.nterm; n++) {
        auto it = drid.find(dn[n]);
        if (it == drid.end()) {
            int id = (int)drid.size();
            drid[dn[n]] = id;
            term_dr[n] = id;
            for (int a = 0; a < 3; a++) v.dr.push_back(v.t_rj[3 * n + a] - v.t_rk[3 * n + a]);
        } else {
            term_dr[n] = it->second;
        }
    }
    v.ndr = (int)drid.size();
    // slots (dr, b, c) in dr-major order; terms in reference order within a slot
    std::map<std::pair<int, int>, std::vector<int>> slots;
    for (int n = 0; n < v.nterm; n++) slots[{term_dr[n], v.t_abc[3 * n + 1] * nb + v.t_abc[3 * n + 2]}].push_back(n);
    v.nslot = (int)slots.size();
    v.slot_ptr.push_back(0);
    for (auto &kv : slots) {
        v.slot_dr.push_back(kv.first.first);
        v.slot_bc.push_back(kv.first.second);
        for (int n : kv.second) {
            v.sterm_a.push_back(v.t_abc[3 * n]);
            v.sterm_psi.push_back(v.t_psi[n]);
            for (int a = 0; a < 3; a++) v.sterm_rk.push_back(v.t_rk[3 * n + a]);
        }
        v.slot_ptr.push_back((int)v.sterm_a.size());
    }
    return ALT3_OK;
}

}  // namespace alt3
