// Reference-style callers of the drop-in header (hartreefock_amd.hpp over libhfgpu.so), exercising the reference's
// public API the way its own callers use it:
//   * atoms taken from Chemistry::Basis::atoms by Z, placed, pushed into a Systems::Molecule, Init(), then
//     RestrictedHartreeFock alpha / initGuess, Init(&molecule), Calculate() (the reference README's usage);
//   * the repository's integrals through getElectronElectron(int x 4) and (orbital pointers x 4), getBoysFunctions,
//     the public core Hamiltonian h;
//   * the SCF driven by Step(iter) as HartreeFockAlgorithm::Calculate drives it;
//   * a bond-length scan that moves atoms and calls SetCenterForShells() between Init/Calculate pairs
//     (HartreeFock/HartreeFockThread.cpp:79-121), reading converged and HOMOEnergy;
//   * Molecule::LoadXYZ, electron counts (UHF cation) and an electric field.
// Prints "key = value" lines that tests/test_gpu_parity.py compares with the oracle.
#include <cmath>
#include <cstdio>
#include <fstream>

#include "hartreefock_amd.hpp"

int main(int argc, char** argv) {
    if (argc < 3) {
        std::fprintf(stderr, "usage: %s basis_file scratch_xyz_path\n", argv[0]);
        return 2;
    }
    Chemistry::Basis basis;
    basis.Load(argv[1]);
    Systems::AtomWithShells H1, H2, O;
    for (auto& atom : basis.atoms) {
        if (atom.Z == 1) H1 = H2 = atom;
        else if (atom.Z == 8) O = atom;
    }
    O.position = Vector3D<double>(0, 0, 0.24026010);
    H1.position = Vector3D<double>(0, 1.43233673, -0.96104039);
    H2.position = Vector3D<double>(0, -1.43233673, -0.96104039);
    Systems::Molecule water;
    water.atoms.push_back(H1);
    water.atoms.push_back(H2);
    water.atoms.push_back(O);
    water.Init();
    std::printf("basis_atoms = %zu\nnbf = %u\nelectrons = %u %u\n", basis.atoms.size(),
                water.CountNumberOfContractedGaussians(), water.alphaElectrons, water.betaElectrons);

    HartreeFock::RestrictedHartreeFock hf;
    hf.alpha = 0.5;
    hf.initGuess = 0;
    hf.normalIterAfterDIIS = 0;
    hf.Init(&water);
    const double E = hf.Calculate();
    std::printf("E = %.12f\nconverged = %d\niterations = %d\nstatus = %d\nHOMO = %.10f\n", E, hf.converged ? 1 : 0,
                hf.iterations, hf.lastStatus, hf.HOMOEnergy);
    std::printf("h00 = %.12f\n", hf.h.empty() ? 0.0 : hf.h[0]);
    const Vector3D<double> mu = hf.GetMoment();
    std::printf("dipole = %.9f %.9f %.9f\nq0 = %.9f\n", mu.X, mu.Y, mu.Z, hf.CalculateAtomicCharge(0));
    std::printf("E(MP2) = %.12f\n", hf.CalculateMp2Energy());
    auto& repo = hf.integralsRepository;
    repo.CalculateElectronElectronIntegrals();
    const auto& o = water.atoms[2].shells;   // oxygen's shells: 1s, then the 2sp shell (s, px, py, pz)
    std::printf("eri_int = %.12f\n", repo.getElectronElectron(2, 3, 5, 6));
    std::printf("eri_orb = %.12f\n", repo.getElectronElectron(&o[0].basisFunctions[0], &o[1].basisFunctions[0],
                                                                &o[1].basisFunctions[2], &o[1].basisFunctions[3]));
    const auto& F = repo.getBoysFunctions(0, 2.5);
    std::printf("boys_n = %zu\nboys0 = %.15f\n", F.functions.size(), F.functions[0]);

    // the SCF driven one Step at a time, with the reference's convergence test on the energy change
    HartreeFock::RestrictedHartreeFock st;
    st.alpha = 0.5;
    st.initGuess = 0;
    st.Init(&water);
    double prev = 0, rms = 1;
    int it = 0;
    for (; it < 200; ++it) {
        rms = st.Step(it);
        const double e = st.GetTotalEnergy();
        if (it > 0 && std::fabs(e - prev) <= 1e-11 && rms < 1e-10 && st.lastErrorEst < 1e-10) break;
        prev = e;
    }
    std::printf("E_steps = %.12f\nsteps = %d\n", st.GetTotalEnergy(), it + 1);

    // bond-length scan of H2, moving the atoms and their shells as HartreeFockThread::Calculate does
    Systems::Molecule h2;
    h2.atoms.push_back(H1);
    h2.atoms.push_back(H2);
    h2.Init();
    HartreeFock::RestrictedHartreeFock scan;
    scan.alpha = 0.5;
    scan.initGuess = 0;
    scan.UseDIIS = false;
    scan.normalIterAfterDIIS = 0;
    for (double r : {1.2, 1.4, 1.6}) {
        h2.atoms[0].position = Vector3D<double>(-r / 2, 0, 0);
        h2.atoms[0].SetCenterForShells();
        h2.atoms[1].position = Vector3D<double>(r / 2, 0, 0);
        h2.atoms[1].SetCenterForShells();
        scan.Init(&h2);
        const double e = scan.Calculate();
        std::printf("scan %.1f = %.12f %d %.10f\n", r, e, scan.converged ? 1 : 0, scan.HOMOEnergy);
    }

    // Molecule::LoadXYZ (bohr) with a cation in UHF and an electric field
    {
        std::ofstream xyz(argv[2]);
        xyz << "3\nwater, bohr\nO 0 0 0.24026010\nH 0 1.43233673 -0.96104039\nH 0 -1.43233673 -0.96104039\n";
    }
    Systems::Molecule cation;
    cation.alphaElectrons = 5;
    cation.betaElectrons = 4;
    const bool loaded = cation.LoadXYZ(argv[2], basis);
    cation.ElectricField = Vector3D<double>(0, 0, 0.01);
    HartreeFock::UnrestrictedHartreeFock u;
    u.alpha = 0.5;
    u.initGuess = 0;
    u.normalIterAfterDIIS = 0;
    u.Init(&cation);
    const double Eu = u.Calculate();
    std::printf("loaded = %d\nE_uhf_field = %.12f\nuhf_converged = %d\nE_field_nuc = %.12f\n", loaded ? 1 : 0, Eu,
                u.converged ? 1 : 0, u.GetNuclearElectricFieldEnergy());
    return hf.converged && st.lastStatus == HF_OK && u.converged ? 0 : 1;
}
