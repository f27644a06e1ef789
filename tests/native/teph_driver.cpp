// Runs electronphonon_mi355x<complex> the way dm_dynamics drives its Teph
// (src_v4.5.7/DensityMatrix_Dynamics.h:59-63,70,266-319): the operator from an
// ldbd_data folder, the state as the reference's row-pointer tables over one
// contiguous pool (myarray.cpp:145-155), evolve_driver without and with an
// explicit 1 - dm, then compute_ddmdt_eq and evolve_driver again.
//
//   teph_driver <ldbd_dir> <scale_eph> <t> <dm.bin> <f.bin> <out_prefix>
// dm.bin: nk x nb_dm^2 complex128; f.bin: nk x nb_dm float64; writes
// <out_prefix>_{plain,dm1,eq}.bin (nk x nb_dm^2 complex128 each).
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include <vector>

#include "electronphonon_mi355x.h"

struct complex {
    double x, y;
};

static std::vector<char> slurp(const char *path, size_t bytes)
{
    std::vector<char> b(bytes);
    FILE *f = std::fopen(path, "rb");
    if (!f || std::fread(b.data(), 1, bytes, f) != bytes) throw std::runtime_error(std::string("cannot read ") + path);
    std::fclose(f);
    return b;
}
static void spill(const std::string &path, const void *p, size_t bytes)
{
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f || std::fwrite(p, 1, bytes, f) != bytes) throw std::runtime_error("cannot write " + path);
    std::fclose(f);
}

int main(int argc, char **argv)
{
    if (argc != 7) {
        std::fprintf(stderr, "usage: %s ldbd_dir scale_eph t dm.bin f.bin out_prefix\n", argv[0]);
        return 2;
    }
    try {
        dmdrun::electronphonon_mi355x<complex> eph(argv[1], "lindblad", std::atof(argv[2]));
        eph.set_algorithm(true, false);
        const long nk = eph.nk_glob;
        const int nbd = eph.nb_expand, n2 = nbd * nbd;
        const double t = std::atof(argv[3]);
        std::vector<char> dmb = slurp(argv[4], (size_t)nk * n2 * sizeof(complex));
        std::vector<char> fb = slurp(argv[5], (size_t)nk * nbd * sizeof(double));
        // row-pointer tables over contiguous pools, as alloc_array builds them
        std::vector<complex> dm_pool((complex *)dmb.data(), (complex *)dmb.data() + (size_t)nk * n2);
        std::vector<complex> dm1_pool((size_t)nk * n2), out_pool((size_t)nk * n2);
        std::vector<double> f_pool((double *)fb.data(), (double *)fb.data() + (size_t)nk * nbd);
        std::vector<complex *> dm(nk), dm1(nk), ddmdt(nk);
        std::vector<double *> f(nk);
        for (long k = 0; k < nk; k++) {
            dm[k] = &dm_pool[(size_t)k * n2];
            dm1[k] = &dm1_pool[(size_t)k * n2];
            ddmdt[k] = &out_pool[(size_t)k * n2];
            f[k] = &f_pool[(size_t)k * nbd];
            // set_oneminusdm (DenMat.cpp:249-261): 1 - dm
            for (int i = 0; i < nbd; i++)
                for (int j = 0; j < nbd; j++) {
                    const complex d = dm[k][i * nbd + j];
                    dm1[k][i * nbd + j] = {(i == j ? 1.0 : 0.0) - d.x, -d.y};
                }
        }
        const std::string out = argv[6];
        eph.evolve_driver(t, dm.data(), nullptr, ddmdt.data());
        spill(out + "_plain.bin", out_pool.data(), out_pool.size() * sizeof(complex));
        eph.evolve_driver(t, dm.data(), dm1.data(), ddmdt.data());
        spill(out + "_dm1.bin", out_pool.data(), out_pool.size() * sizeof(complex));
        eph.compute_ddmdt_eq(f.data());
        eph.evolve_driver(t, dm.data(), dm1.data(), ddmdt.data());
        spill(out + "_eq.bin", out_pool.data(), out_pool.size() * sizeof(complex));
        std::printf("teph_driver: nk = %ld nb_dm = %d window [%d, %d) prefac %.12le\n", nk, nbd, eph.bStart, eph.bEnd,
                    eph.prefac_eph);
    } catch (const std::exception &e) {
        std::fprintf(stderr, "teph_driver: %s\n", e.what());
        return 1;
    }
    return 0;
}
