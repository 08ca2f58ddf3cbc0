// Reference-style usage through the drop-in header (README.md:63-71 of the reference):
// load a basis, build a molecule, Init + Calculate.  Prints the RHF energy.
#include <cstdio>

#include "hartreefock_amd.hpp"

int main(int argc, char** argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s basis_file\n", argv[0]);
        return 2;
    }
    Chemistry::Basis basis;
    basis.Load(argv[1]);
    Systems::Molecule molecule;
    Systems::AtomWithShells O(8), H1(1), H2(1);
    O.position = Vector3D<double>(0, 0, 0.24026010);
    H1.position = Vector3D<double>(0, 1.43233673, -0.96104039);
    H2.position = Vector3D<double>(0, -1.43233673, -0.96104039);
    molecule.atoms = {O, H1, H2};
    molecule.basisFile = basis.fileName;
    molecule.Init();
    HartreeFock::RestrictedHartreeFock hf;
    hf.alpha = 0.5;
    hf.initGuess = 0;
    hf.normalIterAfterDIIS = 0;
    hf.Init(&molecule);
    const double E = hf.Calculate();
    std::printf("E = %.12f converged = %d iterations = %d status = %d\n", E, hf.converged ? 1 : 0, hf.iterations,
                hf.lastStatus);
    const Vector3D<double> mu = hf.GetMoment();
    std::printf("dipole = %.9f %.9f %.9f  q(O) = %.9f\n", mu.X, mu.Y, mu.Z, hf.CalculateAtomicCharge(0));
    std::printf("E(MP2) = %.12f\n", hf.CalculateMp2Energy());
    hf.integralsRepository.CalculateElectronElectronIntegrals();
    std::printf("(11|11) = %.12f\n", hf.integralsRepository.getElectronElectron(0, 0, 0, 0));
    return hf.converged ? 0 : 1;
}
