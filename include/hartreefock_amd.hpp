// hartreefock_amd.hpp — C++ drop-in for the reference's hot-path API, over the C ABI (hfgpu.h).
//
// Same class names, public members and call sequence as the reference, so a caller switches by changing
// includes and linking libhfgpu.so:
//   Orbitals::QuantumNumbers::QuantumNumbers, GaussianOrbital, ContractedGaussianOrbital, ContractedGaussianShell
//                                    HartreeFock/QuantumNumbers.h, GaussianOrbital.h, ContractedGaussianOrbital.h,
//                                    PrimitiveShell.h (the data model: shells, orbitals, exponents, coefficients)
//   Chemistry::Basis                 HartreeFock/Basis.h           (Load from an NWChem file into `atoms`)
//   Systems::Atom / AtomWithShells / Molecule
//                                    HartreeFock/Atom.h, Molecule.h:11-33 (atoms, alpha/betaElectrons,
//                                    ElectricField, Init, SetCenterForShells, LoadXYZ)
//   GaussianIntegrals::IntegralsRepository
//                                    HartreeFock/IntegralsRepository.h:46-185
//                                    (Reset, CalculateElectronElectronIntegrals, getElectronElectron, getBoysFunctions)
//   HartreeFock::RestrictedHartreeFock / UnrestrictedHartreeFock
//                                    HartreeFock/HartreeFockAlgorithm.h:46-101, RestrictedHartreeFock.h:15-41,
//                                    UnrestrictedHartreeFock.h:13-44 (Init, Calculate, Step, h, the tuning fields)
// The molecule reaches the library as shell arrays (hf_set_basis) built from the molecule's own atoms and
// shells, so whatever the caller puts in `atoms` (from a Basis, from LoadXYZ, or built by hand) is what is
// computed.  Matrices are std::vector<double>, row-major n x n, instead of Eigen::MatrixXd.
// Errors: like the reference, Calculate() returns +inf when not initialised and sets `converged`; ABI failures
// are reported through lastStatus / lastError() (no exception crosses the ABI).
#pragma once

#include <algorithm>
#include <atomic>
#include <cctype>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <fstream>
#include <limits>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "hfgpu.h"

template <typename T>
struct Vector3D {
    T X = 0, Y = 0, Z = 0;
    Vector3D() = default;
    Vector3D(T x, T y, T z) : X(x), Y(y), Z(z) {}
    Vector3D operator+(const Vector3D& o) const { return Vector3D(X + o.X, Y + o.Y, Z + o.Z); }
    Vector3D operator-(const Vector3D& o) const { return Vector3D(X - o.X, Y - o.Y, Z - o.Z); }
    Vector3D operator*(T s) const { return Vector3D(X * s, Y * s, Z * s); }
    T operator*(const Vector3D& o) const { return X * o.X + Y * o.Y + Z * o.Z; }   // scalar product
    T Length() const { return std::sqrt(X * X + Y * Y + Z * Z); }
};

namespace Orbitals {

namespace QuantumNumbers {

// Cartesian exponents (l, m, n) of x^l y^m z^n (HartreeFock/QuantumNumbers.h:12-104).
class QuantumNumbers {
public:
    unsigned int l = 0, m = 0, n = 0;
    QuantumNumbers(unsigned int L = 0, unsigned int M = 0, unsigned int N = 0) : l(L), m(M), n(N) {}
    unsigned int AngularMomentum() const { return l + m + n; }
    operator unsigned int() const { return AngularMomentum(); }
    char AtomicOrbital() const { return "spdfgh"[std::min(5u, AngularMomentum())]; }
    static unsigned int NumOrbitals(unsigned int L) { return (L + 1) * (L + 2) / 2; }
    // position within the shell in the order xx, xy, xz, yy, yz, zz (:46-49)
    unsigned int GetCanonicalIndex() const { return (m + n) * (m + n + 1) / 2 + n; }
};

}  // namespace QuantumNumbers

class Orbital {
public:
    Vector3D<double> center;
    QuantumNumbers::QuantumNumbers angularMomentum;
    unsigned int ID = 0;
    unsigned int centerID = 0;
    unsigned int shellID = 0;
    const Vector3D<double>& getCenter() const { return center; }
    char AtomicOrbital() const { return angularMomentum.AtomicOrbital(); }
};

// One primitive: alpha, contraction coefficient and the per-component normalisation factor
// (HartreeFock/GaussianOrbital.cpp:25-30, set by Normalize() as Basis::Load does).
class GaussianOrbital : public Orbital {
public:
    double coefficient = 1;
    double alpha = 1;
    double normalizationFactor = 1;
    double getCoefficient() const { return coefficient; }
    double getAlpha() const { return alpha; }
    double getNormalizationFactor() const {
        auto dfact = [](int k) {
            double r = 1;
            for (int x = k; x > 1; x -= 2) r *= x;
            return r;
        };
        const unsigned int L = angularMomentum;
        return std::pow(2. * alpha / M_PI, 3. / 4.) * std::pow(4. * alpha, L / 2.) /
               std::sqrt(dfact(2 * (int)angularMomentum.l - 1) * dfact(2 * (int)angularMomentum.m - 1) *
                         dfact(2 * (int)angularMomentum.n - 1));
    }
    void Normalize() { normalizationFactor = getNormalizationFactor(); }
};

class ContractedGaussianOrbital : public Orbital {
public:
    std::vector<GaussianOrbital> gaussianOrbitals;
    void AddGaussian(double exponent) {
        GaussianOrbital g;
        g.center = center;
        g.angularMomentum = angularMomentum;
        g.alpha = exponent;
        gaussianOrbitals.push_back(g);
    }
    void Normalize() {
        for (auto& g : gaussianOrbitals) g.Normalize();
    }
};

// A shell of the basis file: one letter group ("S", "SP", "D", ...), each letter contributing all Cartesian
// components of its L in canonical order (HartreeFock/PrimitiveShell.cpp:181-245).
class ContractedGaussianShell {
public:
    unsigned int ID = 0;
    unsigned int centerID = 0;
    std::vector<ContractedGaussianOrbital> basisFunctions;

    void AddOrbitals(char type) {
        const char* order = "spdfgh";
        unsigned int L = 0;
        while (order[L] && order[L] != std::tolower(type)) ++L;
        if (!order[L]) return;
        for (unsigned int x = L + 1; x-- > 0;)              // l descending, then m descending: xx, xy, xz, yy, ...
            for (unsigned int y = L - x + 1; y-- > 0;) {
                ContractedGaussianOrbital o;
                o.angularMomentum = QuantumNumbers::QuantumNumbers(x, y, L - x - y);
                basisFunctions.push_back(o);
            }
    }
    void AddGaussians(double exponent) {
        for (auto& o : basisFunctions) o.AddGaussian(exponent);
    }
    void SetCenters(const Vector3D<double>& c) {
        for (auto& o : basisFunctions) {
            o.center = c;
            for (auto& g : o.gaussianOrbitals) g.center = c;
        }
    }
    const Vector3D<double>& getCenter() const {
        static const Vector3D<double> origin;
        return basisFunctions.empty() ? origin : basisFunctions.front().center;
    }
    void Normalize() {
        for (auto& o : basisFunctions) o.Normalize();
    }
};

}  // namespace Orbitals

namespace Systems {

class Atom {
public:
    Vector3D<double> position;
    unsigned int Z = 0;
    unsigned int electronsNumber = 0;
    unsigned int ID = 0;
    explicit Atom(unsigned int nrZ = 0, int nrElectrons = -1)
        : Z(nrZ), electronsNumber(nrElectrons < 0 ? nrZ : (unsigned)nrElectrons) {}
    virtual ~Atom() = default;
};

class AtomWithShells : public Atom {
public:
    std::vector<Orbitals::ContractedGaussianShell> shells;
    AtomWithShells(unsigned int nrZ = 0, int nrElectrons = -1) : Atom(nrZ, nrElectrons) {}
    void AddShell(const std::string& name) {
        Orbitals::ContractedGaussianShell shell;
        for (char c : name) shell.AddOrbitals(c);
        shells.push_back(shell);
    }
    void SetCenterForShells() {
        for (auto& s : shells) s.SetCenters(position);
    }
    void SetPosition(const Vector3D<double>& pos) {
        position = pos;
        SetCenterForShells();
    }
    unsigned int CountNumberOfContractedGaussians() const {
        unsigned int c = 0;
        for (const auto& s : shells) c += (unsigned int)s.basisFunctions.size();
        return c;
    }
    unsigned int GetMaxAngularMomentum() const {
        unsigned int L = 0;
        for (const auto& s : shells)
            for (const auto& o : s.basisFunctions) L = std::max(L, (unsigned int)o.angularMomentum);
        return L;
    }
    void Normalize() {
        for (auto& s : shells) s.Normalize();
    }
};

namespace detail {
inline unsigned int z_of(const std::string& s) {
    static const char* sym[] = {"H",  "He", "Li", "Be", "B",  "C",  "N", "O",  "F",  "Ne",
                                "Na", "Mg", "Al", "Si", "P",  "S",  "Cl", "Ar"};
    for (unsigned int k = 0; k < sizeof(sym) / sizeof(sym[0]); ++k)
        if (s == sym[k]) return k + 1;
    return 0;
}
}  // namespace detail

}  // namespace Systems

namespace Chemistry {

// The NWChem basis-file reader with the reference's semantics (HartreeFock/Basis.cpp:15-158): lines blank or
// starting with '#', "BASIS" or "END" are skipped; "Elem LETTERS" starts a shell (a new atom when the element
// changes); each numeric line is an exponent followed by one coefficient per letter; every primitive is
// normalised per component and the contraction is not renormalised.
class Basis {
public:
    std::vector<Systems::AtomWithShells> atoms;

    void Load(const std::string& fileName) {
        std::ifstream file(fileName);
        if (!file) return;
        atoms.clear();
        std::string line, shellName;
        auto word = [](const std::string& s) {
            if (s.empty()) return false;
            for (char c : s)
                if (!(std::isalnum((unsigned char)c) || c == '_')) return false;
            return true;
        };
        while (std::getline(file, line)) {
            if (!line.empty() && line.back() == '\r') line.pop_back();
            std::istringstream ls(line);
            std::string a, b, c;
            ls >> a;
            if (a.empty() || a[0] == '#' || a.rfind("BASIS", 0) == 0 || a.rfind("END", 0) == 0) continue;
            ls >> b;
            if (!b.empty() && !(ls >> c) && word(a) && word(b) && Systems::detail::z_of(a) > 0) {
                const unsigned int Z = Systems::detail::z_of(a);
                if (atoms.empty() || atoms.back().Z != Z) atoms.emplace_back(Z);
                shellName = b;
                atoms.back().AddShell(shellName);
                continue;
            }
            if (atoms.empty() || atoms.back().shells.empty()) continue;
            std::istringstream ns(line);
            std::string tok;
            std::vector<double> v;
            while (ns >> tok) {
                for (char& ch : tok)
                    if (ch == 'D' || ch == 'd') ch = 'E';
                char* end = nullptr;
                const double x = std::strtod(tok.c_str(), &end);
                if (end != tok.c_str()) v.push_back(x);
            }
            if (v.empty()) continue;
            Orbitals::ContractedGaussianShell& shell = atoms.back().shells.back();
            shell.AddGaussians(v[0]);
            // the coefficients: one per letter, for every component of that letter
            size_t at = 0;
            for (size_t li = 0; li < shellName.size() && li + 1 < v.size(); ++li) {
                const char* order = "spdfgh";
                unsigned int L = 0;
                while (order[L] && order[L] != std::tolower(shellName[li])) ++L;
                const size_t nc = Orbitals::QuantumNumbers::QuantumNumbers::NumOrbitals(L);
                for (size_t k = 0; k < nc && at + k < shell.basisFunctions.size(); ++k) {
                    auto& g = shell.basisFunctions[at + k].gaussianOrbitals.back();
                    g.coefficient = v[li + 1];
                    g.Normalize();
                }
                at += nc;
            }
        }
    }
};

}  // namespace Chemistry

namespace Systems {

class Molecule {
public:
    std::vector<AtomWithShells> atoms;
    unsigned int alphaElectrons = 0;
    unsigned int betaElectrons = 0;
    Vector3D<double> ElectricField;

    unsigned int ElectronsNumber() const {
        if (alphaElectrons > 0 || betaElectrons > 0) return alphaElectrons + betaElectrons;
        unsigned int e = 0;
        for (const auto& a : atoms) e += a.electronsNumber;
        return e;
    }
    unsigned int CountNumberOfContractedGaussians() const {
        unsigned int c = 0;
        for (const auto& a : atoms) c += a.CountNumberOfContractedGaussians();
        return c;
    }
    unsigned int GetMaxAngularMomentum() const {
        unsigned int L = 0;
        for (const auto& a : atoms) L = std::max(L, a.GetMaxAngularMomentum());
        return L;
    }
    void SetCenterForShells() {
        for (auto& a : atoms) a.SetCenterForShells();
    }
    // IDs in the basis-function order atom -> shell -> component (HartreeFock/Molecule.cpp:41-72)
    void SetIDs() {
        unsigned int id = 0, sid = 0, aid = 0;
        for (auto& a : atoms) {
            a.ID = aid++;
            for (auto& s : a.shells) {
                s.ID = sid++;
                s.centerID = a.ID;
                for (auto& o : s.basisFunctions) {
                    o.ID = id++;
                    o.centerID = a.ID;
                    o.shellID = s.ID;
                }
            }
        }
    }
    double NuclearRepulsionEnergy() const {
        double e = 0;
        for (size_t a = 0; a < atoms.size(); ++a)
            for (size_t b = a + 1; b < atoms.size(); ++b)
                e += (double)atoms[a].Z * atoms[b].Z / (atoms[a].position - atoms[b].position).Length();
        return e;
    }
    double NuclearElectricFieldEnergy() const {
        double e = 0;
        for (const auto& a : atoms) e -= (double)a.Z * (a.position * ElectricField);
        return e;
    }
    void Normalize() {
        for (auto& a : atoms) a.Normalize();
    }
    // HartreeFock/Molecule.cpp:127-140
    void Init() {
        SetCenterForShells();
        SetIDs();
        const unsigned int n = ElectronsNumber();
        if (alphaElectrons == 0 && betaElectrons == 0) {
            alphaElectrons = n / 2;
            betaElectrons = n - alphaElectrons;
        }
    }
    // HartreeFock/Molecule.cpp:143-199: count line, comment line, then "Sym x y z" in bohr; each atom is a copy
    // of the basis's atom of that element
    bool LoadXYZ(const std::string& fileName, const Chemistry::Basis& basis) {
        std::ifstream f(fileName);
        if (!f) return false;
        std::string line;
        if (!std::getline(f, line)) return false;
        const int nat = std::atoi(line.c_str());
        if (!std::getline(f, line)) return false;
        for (int i = 0; i < nat; ++i) {
            if (!std::getline(f, line)) return false;
            std::istringstream ls(line);
            std::string name;
            double x = 0, y = 0, z = 0;
            ls >> name >> x >> y >> z;
            const unsigned int Z = detail::z_of(name);
            if (!Z) return false;
            bool found = false;
            for (const auto& a : basis.atoms)
                if (a.Z == Z) {
                    atoms.push_back(a);
                    atoms.back().position = Vector3D<double>(x, y, z);
                    found = true;
                    break;
                }
            if (!found) return false;
        }
        Init();
        return true;
    }
};

}  // namespace Systems

namespace GaussianIntegrals {

// F_0..F_maxM(T) (HartreeFock/BoysFunctions.h)
class BoysFunctions {
public:
    std::vector<double> functions;
    void GenerateBoysFunctions(int maxM, double T) {
        functions.assign(maxM + 1, 0.0);
        (void)hf_boys(maxM, T, functions.data());
    }
};

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

    // HartreeFock/IntegralsRepository.cpp:19-27: installs the molecule's atoms and shells (SP shells split into
    // S then P, which keeps the basis-function order), its electron counts and its electric field
    void Reset(Systems::Molecule* molecule = nullptr) {
        eri_.clear();
        eri_.shrink_to_fit();
        boys_.clear();
        m_Molecule = molecule;
        if (!molecule || !ctx_) return;
        std::vector<int> L, nprim, poff, bfo, atom, Z;
        std::vector<double> xyz, alpha, coef, axyz;
        int bf = 0;
        for (size_t a = 0; a < molecule->atoms.size(); ++a) {
            const auto& at = molecule->atoms[a];
            Z.push_back((int)at.Z);
            axyz.insert(axyz.end(), {at.position.X, at.position.Y, at.position.Z});
            for (const auto& sh : at.shells) {
                size_t k = 0;
                while (k < sh.basisFunctions.size()) {          // one group per letter: its components are consecutive
                    const auto& o = sh.basisFunctions[k];
                    const unsigned int l = o.angularMomentum;
                    L.push_back((int)l);
                    nprim.push_back((int)o.gaussianOrbitals.size());
                    poff.push_back((int)alpha.size());
                    bfo.push_back(bf);
                    atom.push_back((int)a);
                    xyz.insert(xyz.end(), {o.center.X, o.center.Y, o.center.Z});
                    for (const auto& g : o.gaussianOrbitals) {
                        alpha.push_back(g.alpha);
                        coef.push_back(g.coefficient * std::pow(2. * g.alpha / M_PI, 3. / 4.) *
                                       std::pow(4. * g.alpha, l / 2.));
                    }
                    const size_t nc = Orbitals::QuantumNumbers::QuantumNumbers::NumOrbitals(l);
                    bf += (int)nc;
                    k += nc;
                }
            }
        }
        hf_basis_soa b{};
        b.nshell = (int)L.size();
        b.nbf = bf;
        b.L = L.data();
        b.nprim = nprim.data();
        b.prim_offset = poff.data();
        b.bf_offset = bfo.data();
        b.atom = atom.data();
        b.xyz = xyz.data();
        b.alpha = alpha.data();
        b.coef = coef.data();
        b.natoms = (int)Z.size();
        b.Z = Z.data();
        b.atom_xyz = axyz.data();
        status_ = hf_set_basis(ctx_, &b);
        if (status_ == HF_OK) status_ = hf_set_electrons(ctx_, (int)molecule->alphaElectrons, (int)molecule->betaElectrons);
        const double E[3] = {molecule->ElectricField.X, molecule->ElectricField.Y, molecule->ElectricField.Z};
        if (status_ == HF_OK) status_ = hf_set_electric_field(ctx_, E);
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
    // HartreeFock/IntegralsRepository.cpp:372-433, by the orbitals' IDs (Molecule::SetIDs); the packed store is
    // filled on first use
    double getElectronElectron(const Orbitals::ContractedGaussianOrbital* o1, const Orbitals::ContractedGaussianOrbital* o2,
                               const Orbitals::ContractedGaussianOrbital* o3, const Orbitals::ContractedGaussianOrbital* o4) {
        if (eri_.empty()) CalculateElectronElectronIntegrals();
        return getElectronElectron((int)o1->ID, (int)o2->ID, (int)o3->ID, (int)o4->ID);
    }
    // HartreeFock/IntegralsRepository.cpp:326-342: F_0..F_maxM(T), maxM = 4 Lmax(molecule) + 1 (L + 1 without a
    // molecule), memoised by T
    const BoysFunctions& getBoysFunctions(unsigned int L, double T) {
        auto it = boys_.find(T);
        if (it != boys_.end()) return it->second;
        const unsigned int maxM = m_Molecule ? 4 * m_Molecule->GetMaxAngularMomentum() + 1 : L + 1;
        BoysFunctions& f = boys_[T];
        f.GenerateBoysFunctions((int)maxM, T);
        return f;
    }
    Systems::Molecule* getMolecule() const { return m_Molecule; }
    void SetUseLotsOfMemory(bool) {}   // no host caches to trade against memory here
    void ClearAllMaps() { boys_.clear(); }
    void ClearMatricesMaps() {}
    hf_ctx* context() const { return ctx_; }
    int status() const { return status_; }
    std::string lastError() const { return ctx_ ? hf_last_error(ctx_) : "no context"; }

private:
    hf_ctx* ctx_ = nullptr;
    int status_ = HF_OK;
    Systems::Molecule* m_Molecule = nullptr;
    std::vector<double> eri_;
    std::map<double, BoysFunctions> boys_;
};

}  // namespace GaussianIntegrals

namespace HartreeFock {

// HartreeFock/HartreeFockAlgorithm.h:46-101 (public surface)
class HartreeFockAlgorithm {
public:
    std::vector<double> h;              // core Hamiltonian T + V + E . M after Init (row-major n x n)
    int numberOfOrbitals = 0;
    GaussianIntegrals::IntegralsRepository integralsRepository;
    double alpha = 0.75;
    double initGuess = 0.75;
    std::atomic<int> terminate{0};      // set (to true / non-zero) from another thread to stop Calculate()
    bool converged = false;
    double HOMOEnergy = 0;
    double lastErrorEst = 0;
    bool UseDIIS = true;
    int maxDIISiterations = 1000;
    int normalIterAfterDIIS = 500;
    int iterations = 0;

    explicit HartreeFockAlgorithm(int iterations_ = 3000) : maxIterations(iterations_) {}
    virtual ~HartreeFockAlgorithm() = default;

    // HartreeFock/HartreeFockAlgorithm.cpp:21-101
    virtual void Init(Systems::Molecule* molecule) {
        converged = false;
        stepping_ = false;
        integralsRepository.Reset(molecule);
        numberOfOrbitals = hf_nbf(integralsRepository.context());
        inited = integralsRepository.status() == HF_OK;
        nuclearRepulsionEnergy = molecule ? molecule->NuclearRepulsionEnergy() : 0;
        nuclearElectricFieldEnergy = molecule ? molecule->NuclearElectricFieldEnergy() : 0;
        h.assign((size_t)numberOfOrbitals * numberOfOrbitals, 0.0);
        if (inited) inited = hf_core_hamiltonian(integralsRepository.context(), h.data()) == HF_OK;
    }
    // HartreeFock/HartreeFockAlgorithm.cpp:105-127
    double Calculate() {
        if (!inited) return std::numeric_limits<double>::infinity();
        EndSteps();
        hf_scf_params p = Params();
        static_assert(sizeof(std::atomic<int>) == sizeof(int), "terminate flag layout");
        p.terminate = reinterpret_cast<volatile int*>(&terminate);
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
    // HartreeFock/HartreeFockAlgorithm.h:77: one SCF iteration (the Fock build of the current density, DIIS, the new
    // density, the energy); returns ||D_new - D||_F as the reference's Step does.  The first call after Init starts
    // the iteration sequence from the initial guess.
    virtual double Step(int iter) {
        if (!inited) return std::numeric_limits<double>::infinity();
        hf_ctx* c = integralsRepository.context();
        if (!stepping_) {
            hf_scf_params p = Params();
            lastStatus = hf_scf_begin(c, &p);
            if (lastStatus != HF_OK) return std::numeric_limits<double>::infinity();
            stepping_ = true;
        }
        double e = 0, rms = 0, err = 0;
        lastStatus = hf_scf_step(c, iter, &e, &rms, &err);
        totalEnergy = e;
        lastErrorEst = err;
        return lastStatus == HF_OK ? rms : std::numeric_limits<double>::infinity();
    }
    double GetTotalEnergy() const { return totalEnergy; }
    double GetNuclearElectricFieldEnergy() const { return nuclearElectricFieldEnergy; }
    double GetMP2Energy() const { return mp2Energy; }
    int lastStatus = HF_OK;
    double s2 = 0;
    double mp2Energy = 0;

    // HartreeFock/RestrictedHartreeFock.cpp:251-288, UnrestrictedHartreeFock.cpp:295-375.  The library computes
    // MP2 inside the SCF call (from the orbitals of its last step and the in-core ERI store), so this re-runs
    // Calculate()'s SCF with MP2 requested and returns the correlation energy; the SCF members are refreshed with
    // the identical result.
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
    // HartreeFock/HartreeFockAlgorithm.cpp:234-242
    Vector3D<double> GetNuclearMoment() const {
        Vector3D<double> m;
        if (const Systems::Molecule* mol = integralsRepository.getMolecule())
            for (const auto& a : mol->atoms) m = m + a.position * (double)a.Z;
        return m;
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
    hf_scf_params Params() const {
        hf_scf_params p;
        hf_scf_defaults(&p);
        Fill(p);
        p.max_iterations = maxIterations;
        p.alpha = alpha;
        p.init_guess = initGuess;
        p.use_diis = UseDIIS ? 1 : 0;
        p.max_diis_iterations = maxDIISiterations;
        p.normal_iter_after_diis = normalIterAfterDIIS;
        return p;
    }
    void EndSteps() {
        if (!stepping_) return;
        hf_scf_end(integralsRepository.context(), nullptr, nullptr);
        stepping_ = false;
    }
    int maxIterations;
    bool inited = false;
    bool wantMp2 = false;
    bool stepping_ = false;
    double totalEnergy = std::numeric_limits<double>::infinity();
    double nuclearRepulsionEnergy = 0;
    double nuclearElectricFieldEnergy = 0;
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
