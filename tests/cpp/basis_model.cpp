// The drop-in header's data model without the GPU: Chemistry::Basis::Load of a basis file, atoms picked from
// basis.atoms, Molecule::Init.  Prints one line per basis function ("bf ID atom L l m n alpha0 coef0 norm0") for
// the tests to compare with the oracle's molecule of the same atoms.
#include <cstdio>
#include <cstdlib>

#include "hartreefock_amd.hpp"

int main(int argc, char** argv) {
    if (argc < 2) return 2;
    Chemistry::Basis basis;
    basis.Load(argv[1]);
    Systems::Molecule m;
    for (int a = 2; a + 3 < argc; a += 4) {
        const unsigned int Z = (unsigned)std::atoi(argv[a]);
        for (const auto& at : basis.atoms)
            if (at.Z == Z) {
                m.atoms.push_back(at);
                m.atoms.back().position = Vector3D<double>(std::atof(argv[a + 1]), std::atof(argv[a + 2]),
                                                           std::atof(argv[a + 3]));
                break;
            }
    }
    m.Init();
    std::printf("nbf %u electrons %u %u maxL %u\n", m.CountNumberOfContractedGaussians(), m.alphaElectrons,
                m.betaElectrons, m.GetMaxAngularMomentum());
    for (const auto& at : m.atoms)
        for (const auto& sh : at.shells)
            for (const auto& o : sh.basisFunctions) {
                const auto& g = o.gaussianOrbitals.front();
                std::printf("bf %u %u %u %u %u %u %.17g %.17g %.17g %.17g\n", o.ID, o.centerID,
                            (unsigned)o.angularMomentum, o.angularMomentum.l, o.angularMomentum.m, o.angularMomentum.n,
                            g.alpha, g.coefficient, g.normalizationFactor, o.center.Z);
            }
    return 0;
}
