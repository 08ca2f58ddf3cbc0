// Basis-set reader and molecule shell list.
//
// Semantics of the reference's reader (HartreeFock/Basis.cpp:15-158): lines that are
// blank or start with '#', "BASIS" or "END" are skipped; "Elem TYPE" starts a shell
// whose letters (S, P, SP, D, ...) each contribute all Cartesian components of that L
// (HartreeFock/PrimitiveShell.cpp:181-245); each following numeric line gives one
// exponent and one coefficient per letter.  A new atom block starts whenever the
// element changes.  Basis functions are ordered atom -> shell -> letter -> component
// (HartreeFock/Molecule.cpp:41-72), so splitting SP into an S shell followed by a P
// shell keeps the reference's order.
#include <cctype>
#include <cmath>
#include <cstdlib>
#include <fstream>
#include <mutex>
#include <sstream>

#include "hf_internal.h"

namespace hfhost {

namespace {

const char* kSymbols[] = {"H", "He", "Li", "Be", "B", "C", "N", "O", "F", "Ne",
                          "Na", "Mg", "Al", "Si", "P", "S", "Cl", "Ar"};

int z_of(const std::string& s) {
    for (int i = 0; i < (int)(sizeof(kSymbols) / sizeof(kSymbols[0])); ++i)
        if (s == kSymbols[i]) return i + 1;
    return 0;
}

int l_of(char c) {
    const char* order = "spdfgh";
    for (int i = 0; order[i]; ++i)
        if (std::tolower(c) == order[i]) return i;
    return -1;
}

struct FileShell {
    std::string letters;
    std::vector<double> exps;
    std::vector<std::vector<double>> coefs;   // [primitive][letter]
};

bool is_word(const std::string& s) {
    if (s.empty()) return false;
    for (char c : s)
        if (!(std::isalnum((unsigned char)c) || c == '_')) return false;
    return true;
}

// Numeric line: whitespace-separated numbers, Fortran 'D' exponents accepted.
std::vector<double> numbers_in(const std::string& line) {
    std::vector<double> out;
    std::istringstream ls(line);
    std::string tok;
    while (ls >> tok) {
        for (char& c : tok)
            if (c == 'D' || c == 'd') c = 'E';
        char* end = nullptr;
        const double v = std::strtod(tok.c_str(), &end);
        if (end != tok.c_str()) out.push_back(v);
    }
    return out;
}

bool parse_file(const char* path, std::map<int, std::vector<FileShell>>& atoms, std::string& err) {
    std::ifstream in(path);
    if (!in) { err = std::string("cannot open basis file ") + path; return false; }
    std::string line;
    int cur_z = 0;
    FileShell* cur = nullptr;
    std::map<int, bool> closed;   // element blocks already finished
    FileShell ignored;            // a repeated element block: the reference's loader uses the first one
    while (std::getline(in, line)) {
        if (!line.empty() && line.back() == '\r') line.pop_back();
        std::istringstream ls(line);
        std::string first, second, third;
        ls >> first;
        if (first.empty() || first[0] == '#' || first.rfind("BASIS", 0) == 0 || first.rfind("END", 0) == 0)
            continue;
        ls >> second;
        bool shell_line = false;
        if (!second.empty() && !(ls >> third) && is_word(first) && is_word(second)) shell_line = true;
        if (shell_line) {
            const int z = z_of(first);
            if (!z) { err = "unknown element " + first; return false; }
            if (z != cur_z) {
                if (cur_z) closed[cur_z] = true;
                cur_z = z;
            }
            FileShell fs;
            fs.letters = second;
            for (char c : second)
                if (l_of(c) < 0) { err = "unknown orbital type " + second; return false; }
            if (closed.count(z)) {
                ignored = fs;
                cur = &ignored;
            } else {
                atoms[z].push_back(fs);
                cur = &atoms[z].back();
            }
            continue;
        }
        if (!cur) { err = "numbers before any shell in basis file"; return false; }
        std::vector<double> v = numbers_in(line);
        if (v.empty()) continue;
        cur->exps.push_back(v[0]);
        std::vector<double> c(v.begin() + 1, v.end());
        c.resize(cur->letters.size(), 0.0);
        cur->coefs.push_back(c);
    }
    return true;
}

}  // namespace

double radial_norm(double alpha, int L) {
    // (2a/pi)^{3/4} (4a)^{L/2}: the L-dependent part of HartreeFock/GaussianOrbital.cpp:25-30
    return std::pow(2. * alpha / M_PI, 3. / 4.) * std::pow(4. * alpha, L / 2.);
}

// Parsed basis files, by path: parsed once per process (the serving loop loads the same few files
// for every request).  Guarded for concurrent contexts on different host threads.
static const std::map<int, std::vector<FileShell>>* cached_basis(const char* path, std::string& err) {
    static std::mutex mu;
    static std::map<std::string, std::map<int, std::vector<FileShell>>> cache;
    std::lock_guard<std::mutex> lock(mu);
    auto it = cache.find(path);
    if (it != cache.end()) return &it->second;
    std::map<int, std::vector<FileShell>> atoms;
    if (!parse_file(path, atoms, err)) return nullptr;
    return &cache.emplace(path, std::move(atoms)).first->second;
}

bool load_molecule(const char* basis_file, int natoms, const int* Z, const double* xyz, Molecule& mol,
                   std::string& err) {
    const std::map<int, std::vector<FileShell>>* cached = cached_basis(basis_file, err);
    if (!cached) return false;
    const std::map<int, std::vector<FileShell>>& atoms = *cached;
    mol = Molecule();
    int bf = 0;
    for (int a = 0; a < natoms; ++a) {
        auto it = atoms.find(Z[a]);
        if (it == atoms.end()) { err = "element Z=" + std::to_string(Z[a]) + " not in basis file"; return false; }
        mol.Z.push_back(Z[a]);
        for (int k = 0; k < 3; ++k) mol.xyz.push_back(xyz[3 * a + k]);
        for (const FileShell& fs : it->second) {
            for (size_t li = 0; li < fs.letters.size(); ++li) {
                Shell s;
                s.L = l_of(fs.letters[li]);
                s.atom = a;
                s.bf = bf;
                for (int k = 0; k < 3; ++k) s.xyz[k] = xyz[3 * a + k];
                for (size_t p = 0; p < fs.exps.size(); ++p) {
                    s.alpha.push_back(fs.exps[p]);
                    s.coef.push_back(fs.coefs[p][li] * radial_norm(fs.exps[p], s.L));
                }
                bf += ncart(s.L);
                mol.maxL = std::max(mol.maxL, s.L);
                mol.shells.push_back(s);
            }
        }
    }
    mol.nbf = bf;
    return true;
}

bool molecule_from_soa(const hf_basis_soa* b, Molecule& mol, std::string& err) {
    mol = Molecule();
    if (b->nshell <= 0 || b->natoms <= 0) { err = "empty basis"; return false; }
    for (int a = 0; a < b->natoms; ++a) {
        mol.Z.push_back(b->Z[a]);
        for (int k = 0; k < 3; ++k) mol.xyz.push_back(b->atom_xyz[3 * a + k]);
    }
    int bf = 0;
    for (int k = 0; k < 3 * b->natoms; ++k)
        if (!std::isfinite(b->atom_xyz[k])) { err = "atom coordinate is not finite"; return false; }
    for (int s = 0; s < b->nshell; ++s) {
        Shell sh;
        sh.L = b->L[s];
        sh.atom = b->atom[s];
        sh.bf = b->bf_offset[s];
        if (sh.L < 0 || sh.L > 2) { err = "shell L outside 0..2 is not supported"; return false; }
        if (sh.atom < 0 || sh.atom >= b->natoms) { err = "shell atom index out of range"; return false; }
        for (int k = 0; k < 3; ++k) sh.xyz[k] = b->xyz[3 * s + k];
        for (int p = 0; p < b->nprim[s]; ++p) {
            sh.alpha.push_back(b->alpha[b->prim_offset[s] + p]);
            sh.coef.push_back(b->coef[b->prim_offset[s] + p]);
        }
        bf = std::max(bf, sh.bf + ncart(sh.L));
        mol.maxL = std::max(mol.maxL, sh.L);
        mol.shells.push_back(sh);
    }
    mol.nbf = b->nbf > 0 ? b->nbf : bf;
    return true;
}

double nuclear_repulsion(const Molecule& mol) {
    // HartreeFock/Molecule.cpp:75-84
    double e = 0;
    const int na = (int)mol.Z.size();
    for (int a = 0; a < na; ++a)
        for (int b = a + 1; b < na; ++b) {
            const double dx = mol.xyz[3 * a] - mol.xyz[3 * b], dy = mol.xyz[3 * a + 1] - mol.xyz[3 * b + 1],
                         dz = mol.xyz[3 * a + 2] - mol.xyz[3 * b + 2];
            e += static_cast<double>(mol.Z[a]) * mol.Z[b] / std::sqrt(dx * dx + dy * dy + dz * dz);
        }
    return e;
}

}  // namespace hfhost
