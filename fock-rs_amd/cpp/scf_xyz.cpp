// scf_xyz -- the reference's consumer flow (pointcloud/src/main.rs:44-53) without rendering:
// read an xyz file in a basis set, run try_scf, print the result.
//   scf_xyz <basis.json> <molecule.xyz> [iters=100] [epsilon=1e-6]
#include <cstdio>
#include <cstdlib>

#include "fock_scf.hpp"

int main(int argc, char **argv) {
    if (argc < 3) {
        std::fprintf(stderr, "usage: %s <basis.json> <molecule.xyz> [iters] [epsilon]\n", argv[0]);
        return 2;
    }
    const size_t iters = argc > 3 ? std::strtoul(argv[3], nullptr, 10) : 100;
    const double eps = argc > 4 ? std::strtod(argv[4], nullptr) : 1e-6;
    try {
        const auto basis = fock_rs::BasisSet::load(argv[1]);
        const auto molecule = fock_rs::read_xyz_file(argv[2], basis);
        fock_rs::Context ctx(0);
        auto result = fock_rs::try_scf(molecule, iters, eps, 0, ctx);
        if (!result) {
            std::printf("diverged\n");
            return 1;
        }
        std::printf("n_basis %zu\nn_electrons %zu\niterations %zu\ntotal_energy %.12f\nelectronic_energy %.12f\n",
                    result->n_basis, result->n_electrons, result->iterations, result->total_energy,
                    result->electronic_energy);
        std::printf("orbital_energies");
        for (double e : result->orbital_energies) std::printf(" %.10f", e);
        std::printf("\nhomo_at_origin %.10f\n",
                    result->orbitals[result->n_electrons / 2 - 1].evaluate(fock_rs::Vector3{0.0, 0.0, 0.0}));
    } catch (const std::exception &e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 3;
    }
    return 0;
}
