// Seeded initial-condition writer in the reference's IC layout
// (input.cpp:121-310: /header@dimension, /seedpos[N,d], /rho[N], /vel[N,d],
// /Energy[N]), in place of the reference's unseeded Python scripts
// (create.py, create_sod.py).
//
//   make_ic <out.hdf5> <dim 2|3> <n> <seed> <uniform|sod> [gamma]
//
// Generators: uniform points in the unit box from SplitMix64 (Steele,
// Lea, Flood 2014) of the given seed, 53-bit doubles.  States: "uniform"
// rho = 1, P = 1, v = 0; "sod" rho, P = (1, 1) for x < 0.5 and
// (0.125, 0.1) otherwise, v = 0.  Energy is the total energy per volume
// P / (gamma - 1) + rho v^2 / 2.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "hdf5.h"

namespace {

struct SplitMix64 {
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9e3779b97f4a7c15ull);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

bool write(hid_t f, const char* name, const double* data, hsize_t n0, hsize_t n1) {
    hsize_t dims[2] = {n0, n1};
    hid_t sp = H5Screate_simple(n1 ? 2 : 1, dims, nullptr);
    hid_t ds = H5Dcreate(f, name, H5T_NATIVE_DOUBLE, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    const bool ok = ds >= 0 && H5Dwrite(ds, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, data) >= 0;
    if (ds >= 0) H5Dclose(ds);
    H5Sclose(sp);
    return ok;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 6) {
        std::fprintf(stderr, "usage: %s out.hdf5 dim n seed uniform|sod [gamma]\n", argv[0]);
        return EXIT_FAILURE;
    }
    const std::string out = argv[1], kind = argv[5];
    const int dim = std::atoi(argv[2]);
    const long n = std::atol(argv[3]);
    const uint64_t seed = std::strtoull(argv[4], nullptr, 10);
    const double gamma = argc > 6 ? std::atof(argv[6]) : 5.0 / 3.0;
    if ((dim != 2 && dim != 3) || n < 1 || (kind != "uniform" && kind != "sod") || !(gamma > 1.0)) {
        std::fprintf(stderr, "make_ic: bad arguments\n");
        return EXIT_FAILURE;
    }
    SplitMix64 rng{seed};
    std::vector<double> pos((size_t)n * dim), rho(n), vel((size_t)n * dim, 0.0), E(n);
    for (long i = 0; i < n; i++) {
        for (int d = 0; d < dim; d++) pos[(size_t)i * dim + d] = rng.uniform();
        const bool left = kind == "uniform" || pos[(size_t)i * dim] < 0.5;
        rho[i] = left ? 1.0 : 0.125;
        E[i] = (left ? 1.0 : 0.1) / (gamma - 1.0);
    }
    hid_t f = H5Fcreate(out.c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
    if (f < 0) {
        std::fprintf(stderr, "make_ic: cannot create %s\n", out.c_str());
        return EXIT_FAILURE;
    }
    hid_t h = H5Gcreate(f, "header", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    hid_t sp = H5Screate(H5S_SCALAR);
    hid_t a = H5Acreate(h, "dimension", H5T_NATIVE_INT, sp, H5P_DEFAULT, H5P_DEFAULT);
    H5Awrite(a, H5T_NATIVE_INT, &dim);
    H5Aclose(a);
    H5Sclose(sp);
    H5Gclose(h);
    const bool ok = write(f, "seedpos", pos.data(), n, dim) && write(f, "rho", rho.data(), n, 0) &&
                    write(f, "vel", vel.data(), n, dim) && write(f, "Energy", E.data(), n, 0);
    H5Fclose(f);
    if (!ok) {
        std::fprintf(stderr, "make_ic: write failed\n");
        return EXIT_FAILURE;
    }
    std::printf("make_ic: wrote %ld %dD seeds (%s, seed %llu) to %s\n", n, dim, kind.c_str(),
                (unsigned long long)seed, out.c_str());
    return EXIT_SUCCESS;
}
