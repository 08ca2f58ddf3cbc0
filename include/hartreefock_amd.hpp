// hartreefock_amd.hpp — C++ drop-in for the reference's hot-path API, over the C ABI (hfgpu.h).
//
// Same class names, members and call sequence as the reference so a caller switches by
// changing includes and linking libhfgpu.so:
//   Chemistry::Basis                 HartreeFock/Basis.h           (Load from an NWChem file)
//   Systems::Molecule                HartreeFock/Molecule.h:11-33  (atoms, alpha/betaElectrons, Init, LoadXYZ)
//   GaussianIntegrals::IntegralsRepository
//                                    HartreeFock/IntegralsRepository.h:46-185
//                                    (Reset, CalculateElectronElectronIntegrals, getElectronElectron)
//   HartreeFock::RestrictedHartreeFock / UnrestrictedHartreeFock
//                                    HartreeFock/RestrictedHartreeFock.h:15-41, UnrestrictedHartreeFock.h:13-44
//                                    (Init, Calculate, GetTotalEnergy and the public tuning fields)
// Matrices are std::vector<double>, row-major n x n, instead of Eigen::MatrixXd.
// Errors: like the reference, Calculate() returns +inf when not initialised and sets
// `converged`; ABI failures are reported through lastError() (no exceptions cross the ABI).
#pragma once

#include <cmath>
#include <cstdint>
#include <fstream>
#include <limits>
#include <sstream>
#include <string>
#include <vector>

#include "hfgpu.h"

template <typename T>
struct Vector3D {
    T X = 0, Y = 0, Z = 0;
    Vector3D() = default;
    Vector3D(T x, T y, T z) : X(x), Y(y), Z(z) {}
};

namespace Chemistry {

class Basis {
public:
    std::string fileName;      // the NWChem file; parsed by libhfgpu (HartreeFock/Basis.cpp:15-158 semantics)
    void Load(const std::string& name) { fileName = name; }
};

}  // namespace Chemistry

namespace Systems {

class Atom {
public:
    Vector3D<double> position;
    unsigned int Z = 0;
    unsigned int electronsNumber = 0;
    unsigned int ID = 0;
    explicit Atom(unsigned int nrZ = 0, int nrElectrons = -1)
        : Z(nrZ), electronsNumber(nrElectrons < 0 ? nrZ : (unsigned)nrElectrons) {}
};

class AtomWithShells : public Atom {
public:
    using Atom::Atom;
};

class Molecule {
public:
    std::vector<AtomWithShells> atoms;
    unsigned int alphaElectrons = 0;
    unsigned int betaElectrons = 0;
    Vector3D<double> ElectricField;
    std::string basisFile;

    unsigned int ElectronsNumber() const {
        if (alphaElectrons > 0 || betaElectrons > 0) return alphaElectrons + betaElectrons;
        unsigned int e = 0;
        for (const auto& a : atoms) e += a.electronsNumber;
        return e;
    }
    // HartreeFock/Molecule.cpp:127-140
    void Init() {
        for (unsigned int i = 0; i < atoms.size(); ++i) atoms[i].ID = i;
        const unsigned int n = ElectronsNumber();
        if (alphaElectrons == 0 && betaElectrons == 0) {
            alphaElectrons = n / 2;
            betaElectrons = n - alphaElectrons;
        }
    }
    // HartreeFock/Molecule.cpp:143-199: count line, comment line, then "Sym x y z" in bohr.
    bool LoadXYZ(const std::string& fileName, const Chemistry::Basis& basis) {
        static const char* sym[] = {"H", "He", "Li", "Be", "B", "C", "N", "O", "F", "Ne"};
        std::ifstream f(fileName);
        if (!f) return false;
        std::string line;
        if (!std::getline(f, line)) return false;
        const int nat = std::stoi(line);
        if (!std::getline(f, line)) return false;
        for (int i = 0; i < nat; ++i) {
            if (!std::getline(f, line)) return false;
            std::istringstream ls(line);
            std::string name;
            double x, y, z;
            ls >> name >> x >> y >> z;
            unsigned int Z = 0;
            for (unsigned int k = 0; k < 10; ++k)
                if (name == sym[k]) Z = k + 1;
            if (!Z) return false;
            AtomWithShells a(Z);
            a.position = Vector3D<double>(x, y, z);
            atoms.push_back(a);
        }
        basisFile = basis.fileName;
        Init();
        return true;
    }
};

}  // namespace Systems

namespace GaussianIntegrals {

class IntegralsRepository {
public:
    explicit IntegralsRepository(Systems::Molecule* molecule = nullptr, const hf_options* opt = nullptr) {
        hf_options o;
        hf_default_options(&o);
        if (opt) o = *opt;
        status_ = hf_create(&o, &ctx_);
        if (molecule) Reset(molecule);
    }
    ~IntegralsRepository() { hf_destroy(ctx_); }
    IntegralsRepository(const IntegralsRepository&) = delete;
    IntegralsRepository& operator=(const IntegralsRepository&) = delete;

    // HartreeFock/IntegralsRepository.cpp:19-27
    void Reset(Systems::Molecule* molecule = nullptr) {
        eri_.clear();
        eri_.shrink_to_fit();
        m_Molecule = molecule;
        if (!molecule || !ctx_) return;
        std::vector<int> Z;
        std::vector<double> xyz;
        for (const auto& a : molecule->atoms) {
            Z.push_back((int)a.Z);
            xyz.push_back(a.position.X);
            xyz.push_back(a.position.Y);
            xyz.push_back(a.position.Z);
        }
        status_ = hf_set_molecule(ctx_, molecule->basisFile.c_str(), (int)Z.size(), Z.data(), xyz.data());
    }
    // HartreeFock/IntegralsRepository.cpp:566-584: every unique ERI into the packed store.
    void CalculateElectronElectronIntegrals() {
        const int n = hf_nbf(ctx_);
        eri_.assign(hf_packed_len(n), 0.0);
        status_ = hf_eri_packed(ctx_, eri_.data(), eri_.size());
    }
    // HartreeFock/IntegralsRepository.h:138-141
    double getElectronElectron(int i, int j, int k, int l) const {
        return eri_[hf_packed_index((uint32_t)i, (uint32_t)j, (uint32_t)k, (uint32_t)l)];
    }
    Systems::Molecule* getMolecule() const { return m_Molecule; }
    void SetUseLotsOfMemory(bool) {}   // no host caches to trade against memory here
    void ClearAllMaps() {}
    hf_ctx* context() const { return ctx_; }
    int status() const { return status_; }
    std::string lastError() const { return ctx_ ? hf_last_error(ctx_) : "no context"; }

private:
    hf_ctx* ctx_ = nullptr;
    int status_ = HF_OK;
    Systems::Molecule* m_Molecule = nullptr;
    std::vector<double> eri_;
};

}  // namespace GaussianIntegrals

namespace HartreeFock {

// HartreeFock/HartreeFockAlgorithm.h:46-101 (public surface)
class HartreeFockAlgorithm {
public:
    int numberOfOrbitals = 0;
    GaussianIntegrals::IntegralsRepository integralsRepository;
    double alpha = 0.75;
    double initGuess = 0.75;
    volatile int terminate = 0;
    bool converged = false;
    double HOMOEnergy = 0;
    double lastErrorEst = 0;
    bool UseDIIS = true;
    int maxDIISiterations = 1000;
    int normalIterAfterDIIS = 500;
    int iterations = 0;

    explicit HartreeFockAlgorithm(int iterations_ = 3000) : maxIterations(iterations_) {}
    virtual ~HartreeFockAlgorithm() = default;

    virtual void Init(Systems::Molecule* molecule) {
        converged = false;
        integralsRepository.Reset(molecule);
        numberOfOrbitals = hf_nbf(integralsRepository.context());
        inited = integralsRepository.status() == HF_OK;
    }
    // HartreeFock/HartreeFockAlgorithm.cpp:105-127
    double Calculate() {
        if (!inited) return std::numeric_limits<double>::infinity();
        hf_scf_params p;
        hf_scf_defaults(&p);
        Fill(p);
        p.max_iterations = maxIterations;
        p.alpha = alpha;
        p.init_guess = initGuess;
        p.use_diis = UseDIIS ? 1 : 0;
        p.max_diis_iterations = maxDIISiterations;
        p.normal_iter_after_diis = normalIterAfterDIIS;
        p.terminate = &terminate;
        p.mp2 = wantMp2 ? 1 : 0;
        hf_scf_result r;
        const int n = numberOfOrbitals;
        density_.assign((Unrestricted() ? 2 : 1) * (size_t)n * n, 0.0);
        lastStatus = hf_scf(integralsRepository.context(), &p, &r, density_.data());
        totalEnergy = r.energy;
        HOMOEnergy = r.homo;
        converged = r.converged != 0;
        iterations = r.iterations;
        s2 = r.s2;
        if (wantMp2) mp2Energy = r.mp2_energy;
        Publish();
        return totalEnergy;
    }
    double GetTotalEnergy() const { return totalEnergy; }
    int lastStatus = HF_OK;
    double s2 = 0;
    double mp2Energy = 0;

    // HartreeFock/RestrictedHartreeFock.cpp:251-288, UnrestrictedHartreeFock.cpp:295-375.  The
    // library computes MP2 inside the SCF call (from the orbitals of its last step and the in-core
    // ERI store), so this re-runs Calculate()'s SCF with MP2 requested and returns the correlation
    // energy; the SCF members are refreshed with the identical result.
    double CalculateMp2Energy() {
        if (!inited) return 0;
        wantMp2 = true;
        Calculate();
        wantMp2 = false;
        return mp2Energy;
    }
    // HartreeFock/RestrictedHartreeFock.cpp:291-317 (UHF: UnrestrictedHartreeFock.cpp:381-406)
    double CalculateAtomicCharge(int atom) const {
        std::vector<double> q(integralsRepository.getMolecule() ? integralsRepository.getMolecule()->atoms.size() : 0);
        if (atom < 0 || (size_t)atom >= q.size() || density_.empty()) return 0;
        double mu[3];
        if (hf_properties(integralsRepository.context(), Unrestricted() ? 2 : 1, density_.data(), q.data(), mu) != HF_OK)
            return 0;
        return q[atom];
    }
    // HartreeFock/RestrictedHartreeFock.cpp:319-335 (UHF: UnrestrictedHartreeFock.cpp:408-420)
    Vector3D<double> GetMoment() const {
        double mu[3] = {0, 0, 0};
        if (!density_.empty())
            (void)hf_properties(integralsRepository.context(), Unrestricted() ? 2 : 1, density_.data(), nullptr, mu);
        return Vector3D<double>(mu[0], mu[1], mu[2]);
    }

protected:
    virtual bool Unrestricted() const = 0;
    virtual void Fill(hf_scf_params& p) const = 0;
    virtual void Publish() = 0;
    int maxIterations;
    bool inited = false;
    bool wantMp2 = false;
    double totalEnergy = std::numeric_limits<double>::infinity();
    std::vector<double> density_;
};

class RestrictedHartreeFock : public HartreeFockAlgorithm {
public:
    std::vector<double> DensityMatrix;   // total density (includes the factor 2), n x n
    explicit RestrictedHartreeFock(int iterations_ = 3000) : HartreeFockAlgorithm(iterations_) {}

protected:
    bool Unrestricted() const override { return false; }
    void Fill(hf_scf_params& p) const override { p.unrestricted = 0; }
    void Publish() override { DensityMatrix = density_; }
};

class UnrestrictedHartreeFock : public HartreeFockAlgorithm {
public:
    std::vector<double> DensityMatrixPlus, DensityMatrixMinus;
    double asymmetry = 0.1;
    bool addAsymmetry = true;
    explicit UnrestrictedHartreeFock(int iterations_ = 3000) : HartreeFockAlgorithm(iterations_) {}

protected:
    bool Unrestricted() const override { return true; }
    void Fill(hf_scf_params& p) const override {
        p.unrestricted = 1;
        p.asymmetry = asymmetry;
        p.add_asymmetry = addAsymmetry ? 1 : 0;
    }
    void Publish() override {
        const size_t nn = density_.size() / 2;
        DensityMatrixPlus.assign(density_.begin(), density_.begin() + nn);
        DensityMatrixMinus.assign(density_.begin() + nn, density_.end());
    }
};

}  // namespace HartreeFock
