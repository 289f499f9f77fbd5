// Host-only driver of libalt3's C ABI for an AddressSanitizer / UndefinedBehaviorSanitizer build
// of the host C++ (mesh, symmetry, star and FC3-table construction in host_tables.cpp, the
// argument handling and footprint arithmetic in capi.cpp).  No device is touched: the context
// is created with host_only = 1.
//
//   host_check <input.bin> <output.bin>
//
// input.bin (little endian): int32 natoms0, natoms, nt2, nt3, nc[3], classical, born_flag;
// float64 mass[natoms0], r123[9], latparams[6], primlatt[9], atompos[3 natoms];
// int32 iatomcell0[natoms], fc2_atoms[2 nt2], fc2_xyz[2 nt2]; float64 fc2[nt2];
// int32 fc3_atoms[3 nt3], fc3_xyz[3 nt3]; float64 fc3[nt3]; if born_flag: float64 epsil[9],
// zeu[9 natoms0].
// output.bin: int32 nk, nibz, nb, nops, npos; float64 kpc_eig[3nk], kpc[3nk]; int32 mapibz[nk],
// mapinv[nibz]; float64 kibz[3nibz], wibz[nibz]; int32 tet[24nk]; float64 opk[9nops];
// int32 mappos[nk]; per irreducible point int32 narms, ops[narms], n0[narms]; int64 footprint
// of (1, 3, 7 temperatures) x (FP64, mixed).
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "alt3.h"

namespace {

FILE *in_f;

template <typename T>
std::vector<T> rd(size_t n)
{
    std::vector<T> v(n);
    if (n && std::fread(v.data(), sizeof(T), n, in_f) != n) {
        std::fprintf(stderr, "short read\n");
        std::exit(2);
    }
    return v;
}

template <typename T>
void wr(FILE *f, const T *p, size_t n)
{
    if (n) std::fwrite(p, sizeof(T), n, f);
}

int check(int rc, const char *what)
{
    if (rc != ALT3_OK) {
        std::fprintf(stderr, "%s failed (%d): %s\n", what, rc, alt3_last_error());
        std::exit(3);
    }
    return rc;
}

}  // namespace

int main(int argc, char **argv)
{
    if (argc != 3) {
        std::fprintf(stderr, "usage: host_check input.bin output.bin\n");
        return 1;
    }
    in_f = std::fopen(argv[1], "rb");
    if (!in_f) return 1;
    const auto hdr = rd<int>(9);
    const int natoms0 = hdr[0], natoms = hdr[1], nt2 = hdr[2], nt3 = hdr[3];
    const auto mass = rd<double>(natoms0), r123 = rd<double>(9), latp = rd<double>(6), prim = rd<double>(9);
    const auto pos = rd<double>(3 * (size_t)natoms);
    const auto iat0 = rd<int>(natoms), f2a = rd<int>(2 * (size_t)nt2), f2x = rd<int>(2 * (size_t)nt2);
    const auto f2 = rd<double>(nt2);
    const auto f3a = rd<int>(3 * (size_t)nt3), f3x = rd<int>(3 * (size_t)nt3);
    const auto f3 = rd<double>(nt3);
    std::vector<double> eps, zeu;
    if (hdr[8]) {
        eps = rd<double>(9);
        zeu = rd<double>(9 * (size_t)natoms0);
    }
    std::fclose(in_f);

    alt3_system s{};
    s.natoms0 = natoms0;
    s.mass = mass.data();
    s.r123 = r123.data();
    s.latparams = latp.data();
    s.primlatt = prim.data();
    s.natoms = natoms;
    s.atompos = pos.data();
    s.iatomcell0 = iat0.data();
    s.nt2 = nt2;
    s.fc2_atoms = f2a.data();
    s.fc2_xyz = f2x.data();
    s.fc2 = f2.data();
    s.nt3 = nt3;
    s.fc3_atoms = f3a.data();
    s.fc3_xyz = f3x.data();
    s.fc3 = f3.data();
    s.nc[0] = hdr[4];
    s.nc[1] = hdr[5];
    s.nc[2] = hdr[6];
    s.classical = hdr[7];
    s.host_only = 1;
    s.born_flag = hdr[8];
    s.epsil = hdr[8] ? eps.data() : nullptr;
    s.zeu = hdr[8] ? zeu.data() : nullptr;

    alt3_ctx *ctx = nullptr;
    check(alt3_create(&s, &ctx), "alt3_create");
    int nk = 0, nibz = 0, nb = 0, nops = 0, npos = 0;
    check(alt3_sizes(ctx, &nk, &nibz, &nb, &nops), "alt3_sizes");
    std::vector<double> kpc_eig(3 * (size_t)nk), kpc(3 * (size_t)nk), kibz(3 * (size_t)nibz), wibz(nibz), opk(9 * (size_t)nops);
    std::vector<int> mapibz(nk), mapinv(nibz), tet(24 * (size_t)nk), mappos(nk);
    check(alt3_get_mesh(ctx, kpc_eig.data(), kpc.data(), mapibz.data(), mapinv.data(), kibz.data(), wibz.data(),
                        tet.data(), opk.data(), mappos.data(), &npos),
          "alt3_get_mesh");
    FILE *o = std::fopen(argv[2], "wb");
    if (!o) return 1;
    const int sizes[5] = {nk, nibz, nb, nops, npos};
    wr(o, sizes, 5);
    wr(o, kpc_eig.data(), kpc_eig.size());
    wr(o, kpc.data(), kpc.size());
    wr(o, mapibz.data(), mapibz.size());
    wr(o, mapinv.data(), mapinv.size());
    wr(o, kibz.data(), kibz.size());
    wr(o, wibz.data(), wibz.size());
    wr(o, tet.data(), tet.size());
    wr(o, opk.data(), opk.size());
    wr(o, mappos.data(), mappos.size());
    for (int i = 0; i < nibz; i++) {
        int narms = 0, ops[48], n0[48];
        check(alt3_kstars(ctx, i, &narms, ops, n0), "alt3_kstars");
        wr(o, &narms, 1);
        wr(o, ops, narms);
        wr(o, n0, narms);
    }
    for (int nt : {1, 3, 7})
        for (int p : {ALT3_V3_FP64, ALT3_V3_MIXED}) {
            long long b = 0;
            check(alt3_footprint(ctx, nt, p, &b), "alt3_footprint");
            wr(o, &b, 1);
        }
    std::fclose(o);
    // argument checks: errors, not crashes
    int bad = 0;
    if (alt3_kstars(ctx, nibz, &npos, nullptr, nullptr) == ALT3_OK) bad++;
    if (alt3_footprint(ctx, 0, ALT3_V3_FP64, nullptr) == ALT3_OK) bad++;
    if (alt3_sizes(nullptr, &nk, &nibz, &nb, &nops) == ALT3_OK) bad++;
    check(alt3_destroy(ctx), "alt3_destroy");
    if (bad) {
        std::fprintf(stderr, "%d bad arguments accepted\n", bad);
        return 4;
    }
    std::printf("host_check ok nk=%d nibz=%d nb=%d nops=%d\n", nk, nibz, nb, nops);
    return 0;
}
