/*
 * refcpu.cpp — CPU oracle (TEST INFRASTRUCTURE ONLY; see refcpu.h).
 *
 * A restatement, in this project's own code, of the reference program's
 * algorithm for the two-electron integrals, the Fock build and the SCF loop.
 * Each function names the reference file:line it follows.  Eigen (which the
 * reference uses for dense algebra) is replaced by a cyclic Jacobi eigensolver
 * and a column-pivoted Householder QR written here.
 */
#include "refcpu.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <limits>
#include <list>
#include <map>
#include <regex>
#include <string>
#include <tuple>
#include <vector>

namespace {

const double kPi = 3.14159265358979323846;

/* ------------------------------------------------------------------ */
/* Cartesian quantum numbers: HartreeFock/QuantumNumbers.h:37-104       */
/* ------------------------------------------------------------------ */
struct QN {
    unsigned l = 0, m = 0, n = 0;
    QN() = default;
    QN(unsigned a, unsigned b, unsigned c) : l(a), m(b), n(c) {}
    unsigned L() const { return l + m + n; }
    unsigned maxc() const { return std::max(std::max(l, m), n); }
    unsigned canon() const { return (m + n + 1) * (m + n + 2) / 2 - m - 1; }
    unsigned total() const { unsigned L_ = L(); return L_ * (L_ * (L_ + 3) + 2) / 6 + canon(); }
    static unsigned count(unsigned L_) { return (L_ + 1) * (L_ + 2) / 2; }
    void inc() {  // canonical increment xx,xy,xz,yy,yz,zz then next shell
        unsigned L_ = L();
        if (n == L_) { l = L_ + 1; m = n = 0; }
        else if (m > 0) { --m; ++n; }
        else if (l > 0) { --l; m = L_ - l; n = 0; }
    }
    unsigned& comp(int d) { return d == 0 ? l : (d == 1 ? m : n); }
    unsigned get(int d) const { return d == 0 ? l : (d == 1 ? m : n); }
    int dir() const {   // recursion direction: largest component, ties x > y > z
        unsigned mx = maxc();
        if (l == mx) return 0;
        if (m == mx) return 1;
        return 2;
    }
};

struct V3 {
    double x = 0, y = 0, z = 0;
    V3() = default;
    V3(double a, double b, double c) : x(a), y(b), z(c) {}
    double operator[](int d) const { return d == 0 ? x : (d == 1 ? y : z); }
};
inline V3 operator-(const V3& a, const V3& b) { return V3(a.x - b.x, a.y - b.y, a.z - b.z); }
inline V3 operator+(const V3& a, const V3& b) { return V3(a.x + b.x, a.y + b.y, a.z + b.z); }
inline V3 operator*(double s, const V3& a) { return V3(s * a.x, s * a.y, s * a.z); }
inline V3 operator/(const V3& a, double s) { return V3(a.x / s, a.y / s, a.z / s); }
inline double dot(const V3& a, const V3& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

/* HartreeFock/MathUtils.h:35-45 */
double DoubleFactorial(long n) {
    static const unsigned long long tab[21] = {1, 1, 2, 3, 8, 15, 48, 105, 384, 945, 3840, 10395,
                                               46080, 135135, 645120, 2027025, 10321920, 34459425,
                                               185794560, 654729075, 3715891200ULL};
    if (n <= 1) return 1;
    if (n < 21) return static_cast<double>(tab[n]);
    double res = 1;
    for (; n > 20; n -= 2) res *= n;
    return res * static_cast<double>(tab[n]);
}

/* ------------------------------------------------------------------ */
/* Data model: HartreeFock/GaussianOrbital.*, PrimitiveShell.*, Atom.h  */
/* ------------------------------------------------------------------ */
struct Prim {
    double alpha = 0, coef = 0, norm = 0;
    QN qn;
    V3 c;
    unsigned ID = 0, shellID = 0;
    /* HartreeFock/GaussianOrbital.cpp:25-30 */
    void Normalize() {
        norm = std::pow(2. * alpha / kPi, 3. / 4.) * std::pow(4. * alpha, qn.L() / 2.) /
               std::sqrt(DoubleFactorial(2 * (long)qn.l - 1) * DoubleFactorial(2 * (long)qn.m - 1) *
                         DoubleFactorial(2 * (long)qn.n - 1));
    }
};
struct CGO {
    QN qn;
    V3 c;
    unsigned ID = 0, shellID = 0, atomID = 0;
    std::vector<Prim> g;
};
struct Shell {
    std::vector<CGO> bf;
    unsigned ID = 0;
    /* HartreeFock/PrimitiveShell.cpp:232-245 */
    void AddL(unsigned L) {
        CGO o;
        o.qn.l = L;
        do { bf.push_back(o); o.qn.inc(); } while (o.qn.L() == L);
    }
};
struct Atom {
    unsigned Z = 0, ID = 0;
    V3 pos;
    std::vector<Shell> shells;
};

const char* kElements[] = {"H", "He", "Li", "Be", "B", "C", "N", "O", "F", "Ne", "Na", "Mg",
                           "Al", "Si", "P", "S", "Cl", "Ar"};
unsigned ZForName(const std::string& s) {
    for (unsigned i = 0; i < sizeof(kElements) / sizeof(kElements[0]); ++i)
        if (s == kElements[i]) return i + 1;
    return 0;
}
int LForLetter(char c) {
    switch (std::tolower(c)) {
        case 's': return 0; case 'p': return 1; case 'd': return 2;
        case 'f': return 3; case 'g': return 4; case 'h': return 5;
    }
    return -1;
}

/* NWChem basis reader with the semantics of HartreeFock/Basis.cpp:15-158 */
bool LoadBasis(const std::string& file, std::vector<Atom>& atoms) {
    std::ifstream in(file);
    if (!in) return false;
    std::regex ignore("^\\s*(#|BASIS|END).*$|^\\s*$");
    std::regex number("[+-]?[0-9]*\\.?[0-9]+((E|D)[+-][0-9]+)?");
    std::regex shellre("^(\\w+)\\s+(\\w+)\\s*$");
    std::string line, shellName;
    while (std::getline(in, line)) {
        if (!line.empty() && line.back() == '\r') line.pop_back();
        if (std::regex_match(line, ignore)) continue;
        std::smatch mm;
        if (std::regex_match(line, mm, shellre)) {
            unsigned Z = ZForName(mm[1].str());
            if (!Z) return false;
            if (atoms.empty() || atoms.back().Z != Z) { Atom a; a.Z = Z; atoms.push_back(a); }
            shellName = mm[2].str();
            Shell sh;
            for (char c : shellName) {
                int L = LForLetter(c);
                if (L < 0) return false;
                sh.AddL((unsigned)L);
            }
            atoms.back().shells.push_back(sh);
            continue;
        }
        if (atoms.empty() || atoms.back().shells.empty()) return false;
        Shell& sh = atoms.back().shells.back();
        bool first = true;
        int orbital = 0, total = 0;
        for (auto it = std::sregex_iterator(line.begin(), line.end(), number); it != std::sregex_iterator(); ++it) {
            std::string s = it->str();
            std::replace(s.begin(), s.end(), 'D', 'E');
            double v = std::stod(s);
            if (first) {
                first = false;
                for (CGO& o : sh.bf) { Prim p; p.qn = o.qn; p.alpha = v; o.g.push_back(p); }
            } else {
                int L = LForLetter(shellName.at(orbital));
                int cnt = (int)QN::count((unsigned)L);
                for (int i = 0; i < cnt; ++i) {
                    CGO& o = sh.bf[total + i];
                    o.g.back().coef = v;
                    o.g.back().Normalize();
                }
                total += cnt;
                ++orbital;
            }
        }
    }
    return true;
}

/* ------------------------------------------------------------------ */
/* Boys: HartreeFock/MathUtils.h:163-200, BoysFunction.cpp:7-25,        */
/* BoysFunctions.cpp:10-25                                              */
/* ------------------------------------------------------------------ */
bool IncompleteGamma(double a, double x, double& gin) {
    gin = 0;
    if (a < 0.0 || x < 0) return false;
    if (x == 0.0) return true;
    double xam = -x + a * std::log(x);
    if (xam > 600 || a > 160.0) return false;
    if (x <= 1.0 + a) {
        double sum = 1.0 / a, r = sum;
        for (int k = 1; k <= 80; ++k) {
            r *= x / (a + k);
            sum += r;
            if (std::fabs(r / sum) <= std::numeric_limits<double>::epsilon()) break;
        }
        gin = std::exp(xam) * sum;
    } else {
        double t0 = 0.0;
        for (int k = 80; k >= 1; --k) t0 = (k - a) / (1.0 + k / (x + t0));
        double gim = std::exp(xam) / (x + t0);
        gin = std::tgamma(a) - gim;
    }
    return true;
}

double BoysFunction(double m, double x) {
    if (0. == m && 0. == x) return 1.;
    double res = 0;
    if (std::fabs(x) < 1e-10) x = 1e-10;
    if (x > 160)
        return DoubleFactorial(static_cast<long>(2 * m - 1)) / std::pow(2, m + 1) * std::sqrt(kPi / std::pow(x, 2 * m + 1));
    IncompleteGamma(m + 0.5, x, res);
    return res / (2. * std::pow(x, m + 0.5));
}

void GenerateBoys(int maxM, double T, std::vector<double>& f) {
    f.resize(maxM + 1);
    f[maxM] = BoysFunction(maxM, T);
    for (int m = maxM - 1; m >= 0; --m) f[m] = (2. * T * f[m + 1] + std::exp(-T)) / (2. * m + 1.);
}

/* Small dense helpers */
struct Mat {
    int r = 0, c = 0;
    std::vector<double> v;
    Mat() = default;
    Mat(int rr, int cc) : r(rr), c(cc), v((size_t)rr * cc, 0.0) {}
    double& operator()(int i, int j) { return v[(size_t)i * c + j]; }
    double operator()(int i, int j) const { return v[(size_t)i * c + j]; }
};

}  // namespace

/* ------------------------------------------------------------------ */
/* The molecule + its integral repository state                         */
/* ------------------------------------------------------------------ */
struct ref_mol {
    std::vector<Atom> atoms;
    std::vector<const CGO*> bfs;
    unsigned alphaE = 0, betaE = 0;
    double field[3] = {0, 0, 0};                     /* Molecule::ElectricField, HartreeFock/Molecule.h:19 */
    unsigned maxL = 0;
    std::map<double, std::vector<double>> boys;     /* HartreeFock/IntegralsRepository.h:168 */
    typedef std::tuple<unsigned, unsigned, unsigned, unsigned, unsigned, unsigned, unsigned, unsigned,
                       double, double, double, double> PrimKey;
    std::map<PrimKey, Mat> prim;                     /* HartreeFock/IntegralsRepository.h:180 */

    const std::vector<double>& Boys(double T) {      /* HartreeFock/IntegralsRepository.cpp:326-342 */
        auto it = boys.find(T);
        if (it != boys.end()) return it->second;
        auto& f = boys[T];
        GenerateBoys((int)(4 * maxL + 1), T, f);
        return f;
    }
};

namespace {

/* HartreeFock/GaussianTwoElectrons.cpp:98-133 */
void VerticalRecursion(Mat& M, double alpha, double alpha12, const V3& Rpa, const V3& Rwp, unsigned maxL) {
    const double oneDiv2alpha12 = 1. / (2. * alpha12);
    const double alphaDivAlpha12 = alpha / alpha12;
    const unsigned size = maxL + 1;
    for (QN cur(1, 0, 0); cur.L() < size; cur.inc()) {
        int d = cur.dir();
        QN prev = cur, pp = cur;
        double N = 0;
        bool addpp = false;
        prev.comp(d) -= 1;
        if (prev.get(d) > 0) { N = prev.get(d); pp.comp(d) -= 2; addpp = true; }
        const unsigned ci = cur.total(), pi = prev.total();
        for (unsigned m = 0; m < size - cur.L(); ++m) {
            M(ci, m) = Rpa[d] * M(pi, m) + Rwp[d] * M(pi, m + 1);
            if (addpp) {
                const unsigned ppi = pp.total();
                M(ci, m) += N * oneDiv2alpha12 * (M(ppi, m) - alphaDivAlpha12 * M(ppi, m + 1));
            }
        }
    }
}

/* When set, the transfer relation fills every bra row e (|e| from 0) instead of the
 * reference's |e| >= |f|-1; see ref_set_exact_transfer. */
bool g_exact_transfer = false;

/* HartreeFock/GaussianTwoElectrons.cpp:137-192 (the e-range starts at L(f)-1, as there) */
void ElectronTransfer(Mat& E, double alpha12, double alpha34, const V3& delta, unsigned maxL, unsigned maxL34) {
    const double a12d34 = alpha12 / alpha34;
    const double oneDiv2alpha34 = 1. / (2. * alpha34);
    for (QN q2(1, 0, 0); q2.L() <= maxL34; q2.inc()) {
        const unsigned curL = q2.L();
        const unsigned ci2 = q2.total();
        int d = q2.dir();
        QN p2 = q2, pp2 = q2;
        double Ny = 0;
        bool addpp = false;
        p2.comp(d) -= 1;
        if (p2.get(d) > 0) { Ny = p2.get(d); pp2.comp(d) -= 2; addpp = true; }
        const unsigned pi2 = p2.total();
        const unsigned ppi2 = addpp ? pp2.total() : 0;
        for (QN q1(g_exact_transfer ? 0 : curL - 1, 0, 0); q1.L() <= maxL - curL; q1.inc()) {
            const unsigned ci1 = q1.total();
            QN nx = q1, pv = q1;
            nx.comp(d) += 1;
            double Nx = 0;
            bool addp = false;
            if (pv.get(d) > 0) { Nx = pv.get(d); pv.comp(d) -= 1; addp = true; }
            E(ci1, ci2) = delta[d] * E(ci1, pi2) - a12d34 * E(nx.total(), pi2);
            if (addp) E(ci1, ci2) += Nx * oneDiv2alpha34 * E(pv.total(), pi2);
            if (addpp) E(ci1, ci2) += Ny * oneDiv2alpha34 * E(ci1, ppi2);
        }
    }
}

/* HartreeFock/GaussianTwoElectrons.cpp:15-96 — one primitive quartet, returns the block
 * rows [L1 .. L1+L2] x cols [L3 .. L3+L4] (total canonical indices). */
Mat PrimitiveQuartet(ref_mol* mol, double a1, double a2, double a3, double a4, const V3& c1, const V3& c2,
                     const V3& c3, const V3& c4, unsigned L1, unsigned L2, unsigned L3, unsigned L4) {
    const double a12 = a1 + a2, a34 = a3 + a4, prod = a12 * a34, sum = a12 + a34, alpha = prod / sum;
    const unsigned L12 = L1 + L2, L34 = L3 + L4, L = L12 + L34, size = L + 1;
    const unsigned maxIndex = QN(0, 0, L).total();
    const V3 R12 = c1 - c2, R34 = c3 - c4;
    const V3 Rp = (a1 * c1 + a2 * c2) / a12;
    const V3 Rq = (a3 * c3 + a4 * c4) / a34;
    const V3 Rpq = Rp - Rq;
    const double exponent = -a1 * a2 / a12 * dot(R12, R12) - a3 * a4 / a34 * dot(R34, R34);
    const double factor = 2. * std::pow(kPi, 5. / 2.) / (prod * std::sqrt(sum)) * std::exp(exponent);
    const double T = alpha * dot(Rpq, Rpq);
    Mat M(maxIndex + 1, size);
    const std::vector<double>& f = mol->Boys(T);
    for (unsigned i = 0; i < size; ++i) M(0, i) = factor * f[i];
    VerticalRecursion(M, alpha, a12, Rp - c1, (-alpha / a12) * Rpq, L);
    const unsigned maxIndex12 = QN(0, 0, L12).total(), maxIndex34 = QN(0, 0, L34).total();
    const unsigned iL1 = QN(L1, 0, 0).total(), iL3 = QN(L3, 0, 0).total();
    Mat E(maxIndex + 1, maxIndex34 + 1);
    for (unsigned i = 0; i <= maxIndex; ++i) E(i, 0) = M(i, 0);
    const V3 Delta = (-1.0 / a34) * (a2 * R12 + a4 * R34);
    ElectronTransfer(E, a12, a34, Delta, L, L34);
    Mat B(maxIndex12 - iL1 + 1, maxIndex34 - iL3 + 1);
    for (int i = 0; i < B.r; ++i)
        for (int j = 0; j < B.c; ++j) B(i, j) = E(iL1 + i, iL3 + j);
    return B;
}

/* HartreeFock/IntegralsRepository.cpp:344-370 */
template <class O>
void SwapOrbitals(O*& o1, O*& o2, O*& o3, O*& o4) {
    if (o1->qn.L() < o2->qn.L() || (o1->qn.L() == o2->qn.L() && o1->ID < o2->ID)) std::swap(o1, o2);
    if (o3->qn.L() < o4->qn.L() || (o3->qn.L() == o4->qn.L() && o3->ID < o4->ID)) std::swap(o3, o4);
    const unsigned L12 = o1->qn.L() + o2->qn.L(), L34 = o3->qn.L() + o4->qn.L();
    if (L12 < L34 || (L12 == L34 && (o1->ID < o3->ID || (o1->ID == o3->ID && o2->ID < o4->ID)))) {
        std::swap(o1, o3);
        std::swap(o2, o4);
    }
}

/* HartreeFock/IntegralsRepository.cpp:436-475 */
const Mat& VerticalAndTransfer(ref_mol* mol, const Prim* g1, const Prim* g2, const Prim* g3, const Prim* g4, bool& swapped) {
    if (g1->qn.L() == g2->qn.L() && g1->ID < g2->ID) std::swap(g1, g2);
    if (g3->qn.L() == g4->qn.L() && g3->ID < g4->ID) std::swap(g3, g4);
    if (g1->qn.L() + g2->qn.L() == g3->qn.L() + g4->qn.L() &&
        (g1->ID < g3->ID || (g1->ID == g3->ID && g2->ID < g4->ID))) {
        std::swap(g1, g3);
        std::swap(g2, g4);
        swapped = true;
    } else
        swapped = false;
    ref_mol::PrimKey key(g1->shellID, g2->shellID, g3->shellID, g4->shellID, g1->qn.L(), g2->qn.L(), g3->qn.L(),
                         g4->qn.L(), g1->alpha, g2->alpha, g3->alpha, g4->alpha);
    auto it = mol->prim.find(key);
    if (it != mol->prim.end()) return it->second;
    Mat B = PrimitiveQuartet(mol, g1->alpha, g2->alpha, g3->alpha, g4->alpha, g1->c, g2->c, g3->c, g4->c,
                             g1->qn.L(), g2->qn.L(), g3->qn.L(), g4->qn.L());
    return mol->prim.emplace(key, std::move(B)).first->second;
}

/* HartreeFock/IntegralsRepository.cpp:372-433 + GaussianTwoElectrons.cpp:198-360 */
double ContractedERI(ref_mol* mol, const CGO* o1, const CGO* o2, const CGO* o3, const CGO* o4) {
    SwapOrbitals(o1, o2, o3, o4);
    const unsigned L1 = o1->qn.L(), L2 = o2->qn.L(), L3 = o3->qn.L(), L4 = o4->qn.L();
    const unsigned L12 = L1 + L2, L34 = L3 + L4;
    const unsigned iL1 = QN(L1, 0, 0).total(), iL3 = QN(L3, 0, 0).total();
    const int rows = (int)(QN(0, 0, L12).total() - iL1 + 1), cols = (int)(QN(0, 0, L34).total() - iL3 + 1);
    Mat acc(rows, cols);
    for (const Prim& g1 : o1->g)
        for (const Prim& g2 : o2->g)
            for (const Prim& g3 : o3->g)
                for (const Prim& g4 : o4->g) {
                    const double factor = g1.norm * g2.norm * g3.norm * g4.norm * g1.coef * g2.coef * g3.coef * g4.coef;
                    bool swapped;
                    const Mat& B = VerticalAndTransfer(mol, &g1, &g2, &g3, &g4, swapped);
                    if (swapped) {
                        for (int i = 0; i < rows; ++i)
                            for (int j = 0; j < cols; ++j) acc(i, j) += factor * B(j, i);
                    } else {
                        for (int i = 0; i < rows; ++i)
                            for (int j = 0; j < cols; ++j) acc(i, j) += factor * B(i, j);
                    }
                }
    /* HorizontalRecursion1: (e0|f0) -> (ab|f0); tensor w3[e][b][f], e relative to L1 */
    const V3 AB = o1->c - o2->c;
    const unsigned nb = QN(0, 0, L2).total() + 1;
    std::vector<double> w3((size_t)rows * nb * cols, 0.0);
    auto W3 = [&](unsigned e, unsigned b, unsigned f) -> double& { return w3[((size_t)e * nb + b) * cols + f]; };
    for (int i = 0; i < rows; ++i)
        for (int j = 0; j < cols; ++j) W3(i, 0, j) = acc(i, j);
    {
        unsigned lim = L12;
        for (QN q2(1, 0, 0); q2.L() <= L2;) {
            for (QN q1(L1, 0, 0); q1.L() < lim; q1.inc()) {
                int d = q2.dir();
                QN n1 = q1, p2 = q2;
                n1.comp(d) += 1;
                p2.comp(d) -= 1;
                const unsigned c1 = q1.total() - iL1, nx1 = n1.total() - iL1;
                const unsigned c2 = q2.total(), pv2 = p2.total();
                for (int f = 0; f < cols; ++f) W3(c1, c2, f) = W3(nx1, pv2, f) + AB[d] * W3(c1, pv2, f);
            }
            unsigned old = q2.L();
            q2.inc();
            if (q2.L() != old) --lim;
        }
    }
    const unsigned na = QN::count(L1), nbb = QN::count(L2), b0 = QN(L2, 0, 0).total();
    /* HorizontalRecursion2: (ab|f0) -> (ab|cd); tensor w4[a][b][f][d] */
    const V3 CD = o3->c - o4->c;
    const unsigned nd = QN(0, 0, L4).total() + 1;
    std::vector<double> w4((size_t)na * nbb * cols * nd, 0.0);
    auto W4 = [&](unsigned a, unsigned b, unsigned f, unsigned d) -> double& {
        return w4[(((size_t)a * nbb + b) * cols + f) * nd + d];
    };
    for (unsigned a = 0; a < na; ++a)
        for (unsigned b = 0; b < nbb; ++b)
            for (int f = 0; f < cols; ++f) W4(a, b, f, 0) = W3(a, b0 + b, f);
    {
        unsigned lim = L34;
        for (QN q4(1, 0, 0); q4.L() <= L4;) {
            for (QN q3(L3, 0, 0); q3.L() < lim; q3.inc()) {
                int d = q4.dir();
                QN n3 = q3, p4 = q4;
                n3.comp(d) += 1;
                p4.comp(d) -= 1;
                const unsigned c3 = q3.total() - iL3, nx3 = n3.total() - iL3;
                const unsigned c4 = q4.total(), pv4 = p4.total();
                for (unsigned a = 0; a < na; ++a)
                    for (unsigned b = 0; b < nbb; ++b) W4(a, b, c3, c4) = W4(a, b, nx3, pv4) + CD[d] * W4(a, b, c3, pv4);
            }
            unsigned old = q4.L();
            q4.inc();
            if (q4.L() != old) --lim;
        }
    }
    const unsigned d0 = QN(L4, 0, 0).total();
    return W4(o1->qn.canon(), o2->qn.canon(), o3->qn.canon(), d0 + o4->qn.canon());
}

inline uint64_t TwoIndex(uint64_t i, uint64_t j) { return i < j ? j * (j + 1ULL) / 2 + i : i * (i + 1ULL) / 2 + j; }

/* ------------------------------------------------------------------ */
/* One-electron integrals (HartreeFock/GaussianOverlap.cpp:14-65,       */
/* GaussianKinetic.cpp:23-84, GaussianNuclear.cpp:11-108)               */
/* ------------------------------------------------------------------ */
/* 1-D overlap table s[i][j] for i<=li+1, j<=lj+1 (the +1 serves the kinetic terms) */
void Overlap1D(double a1, double a2, double c1, double c2, unsigned li, unsigned lj, Mat& s) {
    const double alpha = a1 + a2, P = (a1 * c1 + a2 * c2) / alpha, dif = c1 - c2, PA = P - c1, o2a = 1. / (2. * alpha);
    const unsigned I = li + 1, J = lj + 1, lim = I + J;
    Mat t(lim + 1, J + 1);
    t(0, 0) = 1;
    if (lim >= 1) t(1, 0) = PA;
    for (unsigned i = 2; i <= lim; ++i) t(i, 0) = PA * t(i - 1, 0) + (i - 1) * o2a * t(i - 2, 0);
    for (unsigned j = 1; j <= J; ++j)
        for (unsigned i = 0; i + j <= lim; ++i) t(i, j) = t(i + 1, j - 1) + dif * t(i, j - 1);
    s = Mat(I + 1, J + 1);
    for (unsigned i = 0; i <= I; ++i)
        for (unsigned j = 0; j <= J; ++j) s(i, j) = t(i, j);
}
double Kinetic1D(const Mat& s, double a1, double a2, unsigned i, unsigned j) {
    const double t2 = 2 * a1 * a2;
    double v = t2 * s(i + 1, j + 1);
    if (i > 0) v -= i * a2 * s(i - 1, j + 1);
    if (j > 0) v -= j * a1 * s(i + 1, j - 1);
    if (i > 0 && j > 0) v += 0.5 * (double)i * j * s(i - 1, j - 1);
    return v;
}

void PrimOverlapKinetic(const Prim& g1, const Prim& g2, double& S, double& T) {
    Mat sx, sy, sz;
    Overlap1D(g1.alpha, g2.alpha, g1.c.x, g2.c.x, g1.qn.l, g2.qn.l, sx);
    Overlap1D(g1.alpha, g2.alpha, g1.c.y, g2.c.y, g1.qn.m, g2.qn.m, sy);
    Overlap1D(g1.alpha, g2.alpha, g1.c.z, g2.c.z, g1.qn.n, g2.qn.n, sz);
    const V3 d = g1.c - g2.c;
    const double o = 1. / (g1.alpha + g2.alpha);
    const double factor = std::exp(-g1.alpha * g2.alpha * o * dot(d, d)) * std::pow(kPi * o, 1.5);
    const double X = sx(g1.qn.l, g2.qn.l), Y = sy(g1.qn.m, g2.qn.m), Z = sz(g1.qn.n, g2.qn.n);
    S = factor * X * Y * Z;
    T = factor * (Kinetic1D(sx, g1.alpha, g2.alpha, g1.qn.l, g2.qn.l) * Y * Z +
                  X * Kinetic1D(sy, g1.alpha, g2.alpha, g1.qn.m, g2.qn.m) * Z +
                  X * Y * Kinetic1D(sz, g1.alpha, g2.alpha, g1.qn.n, g2.qn.n));
}

/* Nuclear attraction of a contracted pair to one nucleus: VRR on centre 1 per primitive
 * pair (GaussianNuclear::VerticalRecursion), contraction, then HRR to centre 2. */
double ContractedNuclear(ref_mol* mol, const V3& C, const CGO* o1, const CGO* o2) {
    if (o1->qn.L() < o2->qn.L() || (o1->qn.L() == o2->qn.L() && o1->ID > o2->ID)) std::swap(o1, o2);
    const unsigned L1 = o1->qn.L(), L2 = o2->qn.L(), L = L1 + L2;
    const unsigned rows = QN(0, 0, L).total() + 1;
    std::vector<double> col(rows, 0.0);
    for (const Prim& g1 : o1->g)
        for (const Prim& g2 : o2->g) {
            const double a = g1.alpha + g2.alpha;
            const V3 P = (g1.alpha * g1.c + g2.alpha * g2.c) / a;
            const V3 PC = C - P, dif = g1.c - g2.c, PA = P - g1.c;
            const std::vector<double>& f = mol->Boys(a * dot(PC, PC));
            const double factor = 2. * kPi / a * std::exp(-g1.alpha * g2.alpha / a * dot(dif, dif));
            Mat M(rows, L + 1);
            for (unsigned i = 0; i <= L; ++i) M(0, i) = factor * f[i];
            const double o2a = 1. / (2. * a);
            for (QN cur(1, 0, 0); cur.L() < L + 1; cur.inc()) {
                int d = cur.dir();
                QN prev = cur, pp = cur;
                prev.comp(d) -= 1;
                double N = 0;
                bool addpp = false;
                if (prev.get(d) > 0) { N = prev.get(d); pp.comp(d) -= 2; addpp = true; }
                for (unsigned m = 0; m < L + 1 - cur.L(); ++m) {
                    M(cur.total(), m) = PA[d] * M(prev.total(), m) + PC[d] * M(prev.total(), m + 1);
                    if (addpp) M(cur.total(), m) += N * o2a * (M(pp.total(), m) - M(pp.total(), m + 1));
                }
            }
            const double w = g1.norm * g2.norm * g1.coef * g2.coef;
            for (unsigned r = 0; r < rows; ++r) col[r] += w * M(r, 0);
        }
    /* HRR: (a, b+1_i) = (a+1_i, b) + AB_i (a, b) */
    const V3 AB = o1->c - o2->c;
    const unsigned ncol = QN(0, 0, L2).total() + 1;
    Mat H(rows, ncol);
    for (unsigned r = 0; r < rows; ++r) H(r, 0) = col[r];
    unsigned lim = L;
    for (QN qj(1, 0, 0); qj.L() <= L2;) {
        for (QN qi(0, 0, 0); qi.L() < lim; qi.inc()) {
            int d = qj.dir();
            QN nx = qi, pj = qj;
            nx.comp(d) += 1;
            pj.comp(d) -= 1;
            H(qi.total(), qj.total()) = H(nx.total(), pj.total()) + AB[d] * H(qi.total(), pj.total());
        }
        unsigned old = qj.L();
        qj.inc();
        if (qj.L() != old) --lim;
    }
    return H(o1->qn.total(), o2->qn.total());
}

/* ------------------------------------------------------------------ */
/* Dense algebra (replaces Eigen)                                       */
/* ------------------------------------------------------------------ */
/* Cyclic Jacobi: A (row-major n x n, symmetric) -> ascending eigenvalues w, eigenvectors in
 * columns of V. */
void JacobiEigen(int n, std::vector<double> A, std::vector<double>& w, std::vector<double>& V) {
    V.assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) V[(size_t)i * n + i] = 1.0;
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0, diag = 0;
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) {
                double a = A[(size_t)i * n + j];
                if (i != j) off += a * a; else diag += a * a;
            }
        if (off <= 1e-30 * std::max(diag, 1e-300) || off == 0.0) break;
        for (int p = 0; p < n - 1; ++p)
            for (int q = p + 1; q < n; ++q) {
                const double apq = A[(size_t)p * n + q];
                if (apq == 0.0) continue;
                const double app = A[(size_t)p * n + p], aqq = A[(size_t)q * n + q];
                const double theta = (aqq - app) / (2.0 * apq);
                const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
                const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < n; ++k) {
                    const double akp = A[(size_t)k * n + p], akq = A[(size_t)k * n + q];
                    A[(size_t)k * n + p] = c * akp - s * akq;
                    A[(size_t)k * n + q] = s * akp + c * akq;
                }
                for (int k = 0; k < n; ++k) {
                    const double apk = A[(size_t)p * n + k], aqk = A[(size_t)q * n + k];
                    A[(size_t)p * n + k] = c * apk - s * aqk;
                    A[(size_t)q * n + k] = s * apk + c * aqk;
                }
                for (int k = 0; k < n; ++k) {
                    const double vkp = V[(size_t)k * n + p], vkq = V[(size_t)k * n + q];
                    V[(size_t)k * n + p] = c * vkp - s * vkq;
                    V[(size_t)k * n + q] = s * vkp + c * vkq;
                }
            }
    }
    std::vector<int> idx(n);
    for (int i = 0; i < n; ++i) idx[i] = i;
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return A[(size_t)a * n + a] < A[(size_t)b * n + b]; });
    w.resize(n);
    std::vector<double> V2((size_t)n * n);
    for (int j = 0; j < n; ++j) {
        w[j] = A[(size_t)idx[j] * n + idx[j]];
        for (int i = 0; i < n; ++i) V2[(size_t)i * n + j] = V[(size_t)i * n + idx[j]];
    }
    V.swap(V2);
}

std::vector<double> MatMul(int n, const std::vector<double>& A, const std::vector<double>& B) {
    std::vector<double> C((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < n; ++k) {
            const double a = A[(size_t)i * n + k];
            for (int j = 0; j < n; ++j) C[(size_t)i * n + j] += a * B[(size_t)k * n + j];
        }
    return C;
}

/* Least-squares / square solve by column-pivoted Householder QR (the solver the
 * reference's DIIS uses, HartreeFock/DIIS.h:72). */
std::vector<double> ColPivQRSolve(int n, std::vector<double> A, std::vector<double> b) {
    std::vector<int> perm(n);
    for (int i = 0; i < n; ++i) perm[i] = i;
    std::vector<double> norms(n);
    for (int j = 0; j < n; ++j) {
        double s = 0;
        for (int i = 0; i < n; ++i) s += A[(size_t)i * n + j] * A[(size_t)i * n + j];
        norms[j] = s;
    }
    int rank = n;
    double maxnorm0 = 0;
    for (int j = 0; j < n; ++j) maxnorm0 = std::max(maxnorm0, std::sqrt(norms[j]));
    for (int k = 0; k < n; ++k) {
        int piv = k;
        double best = -1;
        for (int j = k; j < n; ++j) {
            double s = 0;
            for (int i = k; i < n; ++i) s += A[(size_t)i * n + j] * A[(size_t)i * n + j];
            if (s > best) { best = s; piv = j; }
        }
        if (piv != k) {
            for (int i = 0; i < n; ++i) std::swap(A[(size_t)i * n + k], A[(size_t)i * n + piv]);
            std::swap(perm[k], perm[piv]);
        }
        double alpha = 0;
        for (int i = k; i < n; ++i) alpha += A[(size_t)i * n + k] * A[(size_t)i * n + k];
        alpha = std::sqrt(alpha);
        if (alpha <= std::numeric_limits<double>::epsilon() * maxnorm0 * n) { rank = k; break; }
        if (A[(size_t)k * n + k] > 0) alpha = -alpha;
        std::vector<double> v(n, 0.0);
        for (int i = k; i < n; ++i) v[i] = A[(size_t)i * n + k];
        v[k] -= alpha;
        double vn = 0;
        for (int i = k; i < n; ++i) vn += v[i] * v[i];
        if (vn == 0) continue;
        for (int j = k; j < n; ++j) {
            double s = 0;
            for (int i = k; i < n; ++i) s += v[i] * A[(size_t)i * n + j];
            s = 2 * s / vn;
            for (int i = k; i < n; ++i) A[(size_t)i * n + j] -= s * v[i];
        }
        double s = 0;
        for (int i = k; i < n; ++i) s += v[i] * b[i];
        s = 2 * s / vn;
        for (int i = k; i < n; ++i) b[i] -= s * v[i];
    }
    std::vector<double> y(n, 0.0);
    for (int i = rank - 1; i >= 0; --i) {
        double s = b[i];
        for (int j = i + 1; j < rank; ++j) s -= A[(size_t)i * n + j] * y[j];
        y[i] = s / A[(size_t)i * n + i];
    }
    std::vector<double> x(n, 0.0);
    for (int i = 0; i < n; ++i) x[perm[i]] = y[i];
    return x;
}

/* HartreeFock/DIIS.h:9-96 (6 kept, estimate from 5) */
struct DIIS {
    std::list<std::vector<double>> errors, values;
    bool Add(const std::vector<double>& v, const std::vector<double>& e) {
        errors.push_back(e);
        values.push_back(v);
        if (errors.size() > 6) { errors.pop_front(); values.pop_front(); }
        return errors.size() >= 5;
    }
    double Estimate(std::vector<double>& value) const {
        const int nm = (int)errors.size(), np = nm + 1;
        std::vector<double> B((size_t)np * np, 0.0);
        double est = 0;
        auto e1 = errors.cbegin();
        for (int i = 0; i < nm; ++i, ++e1) {
            auto e2 = errors.cbegin();
            for (int j = 0; j < i; ++j, ++e2) {
                double s = 0;
                for (size_t k = 0; k < e1->size(); ++k) s += (*e1)[k] * (*e2)[k];
                B[(size_t)i * np + j] = B[(size_t)j * np + i] = s;
            }
            double s = 0;
            for (size_t k = 0; k < e1->size(); ++k) s += (*e1)[k] * (*e1)[k];
            B[(size_t)i * np + i] = s;
            if (i == nm - 1 || i == nm - 2) est += s;
            B[(size_t)nm * np + i] = B[(size_t)i * np + nm] = 1;
        }
        est = std::sqrt(est);
        std::vector<double> rhs(np, 0.0);
        rhs[nm] = 1;
        std::vector<double> c = ColPivQRSolve(np, B, rhs);
        std::fill(value.begin(), value.end(), 0.0);
        auto it = values.cbegin();
        for (int i = 0; i < nm; ++i, ++it)
            for (size_t k = 0; k < value.size(); ++k) value[k] += c[i] * (*it)[k];
        return est;
    }
};

}  // namespace

/* ================================================================== */
/* C API                                                               */
/* ================================================================== */
extern "C" {

int ref_mol_create(const char* basis_file, int natoms, const int* Z, const double* xyz, ref_mol** out) {
    if (!basis_file || natoms <= 0 || !Z || !xyz || !out) return -1;
    std::vector<Atom> basis;
    if (!LoadBasis(basis_file, basis)) return -6;
    ref_mol* m = new ref_mol();
    for (int a = 0; a < natoms; ++a) {
        const Atom* found = nullptr;
        for (const Atom& b : basis)
            if ((int)b.Z == Z[a]) { found = &b; break; }
        if (!found) { delete m; return -1; }
        Atom at = *found;
        at.pos = V3(xyz[3 * a], xyz[3 * a + 1], xyz[3 * a + 2]);
        m->atoms.push_back(at);
    }
    /* HartreeFock/Molecule.cpp:41-72 (SetIDs) and :127-140 (Init) */
    unsigned ID = 0, cID = 0, sID = 0, aID = 0, electrons = 0;
    for (Atom& at : m->atoms) {
        at.ID = aID++;
        electrons += at.Z;
        for (Shell& sh : at.shells) {
            sh.ID = sID++;
            for (CGO& o : sh.bf) {
                o.c = at.pos;
                o.ID = cID++;
                o.atomID = at.ID;
                o.shellID = sh.ID;
                m->maxL = std::max(m->maxL, o.qn.L());
                for (Prim& g : o.g) { g.c = at.pos; g.ID = ID++; g.shellID = sh.ID; }
            }
        }
    }
    for (Atom& at : m->atoms)
        for (Shell& sh : at.shells)
            for (CGO& o : sh.bf) m->bfs.push_back(&o);
    m->alphaE = electrons / 2;
    m->betaE = electrons - m->alphaE;
    *out = m;
    return 0;
}

void ref_mol_destroy(ref_mol* m) { delete m; }
int ref_mol_nbf(const ref_mol* m) { return (int)m->bfs.size(); }
int ref_mol_max_l(const ref_mol* m) { return (int)m->maxL; }
int ref_mol_electrons(const ref_mol* m, int* a, int* b) {
    if (a) *a = (int)m->alphaE;
    if (b) *b = (int)m->betaE;
    return (int)(m->alphaE + m->betaE);
}

/* Molecule::alphaElectrons / betaElectrons set by the caller before the SCF (HartreeFock/Molecule.h:16-17;
   RHF reads alpha + beta, HartreeFock/Molecule.cpp:96-101, UHF each, HartreeFock/UnrestrictedHartreeFock.cpp:33-37) */
int ref_mol_set_electrons(ref_mol* m, int a, int b) {
    if (a < 0 || b < 0) return -1;
    m->alphaE = (unsigned)a;
    m->betaE = (unsigned)b;
    return 0;
}

int ref_mol_set_field(ref_mol* m, const double* E) {
    for (int k = 0; k < 3; ++k) m->field[k] = E[k];
    return 0;
}

/* HartreeFock/Molecule.cpp:75-84 */
double ref_mol_nuclear_repulsion(const ref_mol* m) {
    double e = 0;
    for (size_t a = 0; a < m->atoms.size(); ++a)
        for (size_t b = a + 1; b < m->atoms.size(); ++b) {
            const V3 d = m->atoms[a].pos - m->atoms[b].pos;
            e += static_cast<double>(m->atoms[a].Z) * m->atoms[b].Z / std::sqrt(dot(d, d));
        }
    return e;
}

int ref_mol_bf_info(const ref_mol* m, int* L, int* lmn, int* atom, int* shell) {
    for (size_t i = 0; i < m->bfs.size(); ++i) {
        const CGO* o = m->bfs[i];
        if (L) L[i] = (int)o->qn.L();
        if (lmn) { lmn[3 * i] = o->qn.l; lmn[3 * i + 1] = o->qn.m; lmn[3 * i + 2] = o->qn.n; }
        if (atom) atom[i] = (int)o->atomID;
        if (shell) shell[i] = (int)o->shellID;
    }
    return 0;
}

void ref_set_exact_transfer(int on) { g_exact_transfer = on != 0; }

int ref_boys(int maxm, double T, double* out) {
    std::vector<double> f;
    GenerateBoys(maxm, T, f);
    for (int i = 0; i <= maxm; ++i) out[i] = f[i];
    return 0;
}

double ref_eri(ref_mol* m, int i, int j, int k, int l) {
    const int n = (int)m->bfs.size();
    if (i < 0 || j < 0 || k < 0 || l < 0 || i >= n || j >= n || k >= n || l >= n)
        return std::numeric_limits<double>::quiet_NaN();
    double v = ContractedERI(m, m->bfs[i], m->bfs[j], m->bfs[k], m->bfs[l]);
    m->prim.clear();
    return v;
}

uint64_t ref_packed_index(uint32_t i, uint32_t j, uint32_t k, uint32_t l) {
    return TwoIndex(TwoIndex(i, j), TwoIndex(k, l));
}
uint64_t ref_packed_len(int n) { return ref_packed_index(n, n, n, n) + 1; }

/* HartreeFock/IntegralsRepository.cpp:484-528, with the useLotsOfMemory=false cache
 * lifetime (primitive cache cleared at every shell-4 boundary, :499,505). */
int ref_eri_packed_rows(ref_mol* m, double* out, uint64_t len, int i0, int i1) {
    const int n = (int)m->bfs.size();
    if (len < ref_packed_len(n) || i0 < 0 || i1 > n || i0 > i1) return -1;
    for (int i = i0; i < i1; ++i)
        for (int j = 0; j <= i; ++j) {
            const uint64_t ij = TwoIndex(i, j);
            for (int k = 0; k < n; ++k) {
                for (int l = 0; l <= k; ++l) {
                    const uint64_t kl = TwoIndex(k, l);
                    if (ij <= kl) out[TwoIndex(ij, kl)] = ContractedERI(m, m->bfs[i], m->bfs[j], m->bfs[k], m->bfs[l]);
                    if (l == k || m->bfs[l]->shellID != m->bfs[l + 1]->shellID) m->prim.clear();
                }
            }
        }
    m->prim.clear();
    m->boys.clear();
    return 0;
}

/* Every stride-th (i, j, k) triple of the loop above (offset selects which), with its
 * whole l loop and the same cache lifetime: a bounded, unbiased sample of the ERI
 * phase for timing.  *count receives the number of ERIs evaluated. */
int ref_eri_packed_sample(ref_mol* m, double* out, uint64_t len, uint64_t stride, uint64_t offset, uint64_t* count) {
    const int n = (int)m->bfs.size();
    if (len < ref_packed_len(n) || stride == 0) return -1;
    uint64_t t = 0, cnt = 0;
    for (int i = 0; i < n; ++i)
        for (int j = 0; j <= i; ++j) {
            const uint64_t ij = TwoIndex(i, j);
            for (int k = 0; k < n; ++k, ++t) {
                if (t % stride != offset) continue;
                for (int l = 0; l <= k; ++l) {
                    const uint64_t kl = TwoIndex(k, l);
                    if (ij <= kl) {
                        out[TwoIndex(ij, kl)] = ContractedERI(m, m->bfs[i], m->bfs[j], m->bfs[k], m->bfs[l]);
                        ++cnt;
                    }
                    if (l == k || m->bfs[l]->shellID != m->bfs[l + 1]->shellID) m->prim.clear();
                }
            }
        }
    m->prim.clear();
    m->boys.clear();
    if (count) *count = cnt;
    return 0;
}

/* The RHF two-electron matrix exactly as HartreeFock/RestrictedHartreeFock.cpp:189-202:
 * G_ij = sum_kl D_kl [(ij|kl) - 1/2 (il|kj)], one N^4 loop (timed as the CPU baseline);
 * rows i in [i0, i1) only. */
int ref_rhf_g(int n, const double* eri, const double* D, double* G, int i0, int i1) {
    if (i0 < 0 || i1 > n || i0 > i1) return -1;
    for (int i = i0; i < i1; ++i)
        for (int j = 0; j < n; ++j) {
            double g = 0;
            for (int k = 0; k < n; ++k)
                for (int l = 0; l < n; ++l) {
                    const double coulomb = eri[ref_packed_index(i, j, k, l)];
                    const double exchange = eri[ref_packed_index(i, l, k, j)];
                    g += D[(size_t)k * n + l] * (coulomb - 0.5 * exchange);
                }
            G[(size_t)i * n + j] = g;
        }
    return 0;
}

int ref_eri_packed(ref_mol* m, double* out, uint64_t len) {
    return ref_eri_packed_rows(m, out, len, 0, (int)m->bfs.size());
}

int ref_one_electron(ref_mol* m, double* S, double* T, double* V) {
    const int n = (int)m->bfs.size();
    for (int i = 0; i < n; ++i)
        for (int j = i; j < n; ++j) {
            const CGO* a = m->bfs[i];
            const CGO* b = m->bfs[j];
            double s = 0, t = 0;
            for (const Prim& g1 : a->g)
                for (const Prim& g2 : b->g) {
                    double ps, pt;
                    PrimOverlapKinetic(g1, g2, ps, pt);
                    const double w = g1.norm * g2.norm * g1.coef * g2.coef;
                    s += w * ps;
                    t += w * pt;
                }
            double v = 0;
            for (const Atom& at : m->atoms) v -= at.Z * ContractedNuclear(m, at.pos, a, b);
            if (S) S[(size_t)i * n + j] = S[(size_t)j * n + i] = s;
            if (T) T[(size_t)i * n + j] = T[(size_t)j * n + i] = t;
            if (V) V[(size_t)i * n + j] = V[(size_t)j * n + i] = v;
        }
    return 0;
}

/* Dipole moment integrals about the origin, <a|x|b>, <a|y|b>, <a|z|b> (MomentMatrix::Calculate,
 * HartreeFock/QuantumMatrix.cpp:61-90, GaussianMoment.cpp:15-90 with centre3 = 0 from
 * GaussianMoment.h:38-49): the factor of the moment direction is s(i+1, j) + A s(i, j) (x = x_A + A),
 * the other two directions are the 1-D overlaps. */
int ref_moment(ref_mol* m, double* Mx, double* My, double* Mz) {
    const int n = (int)m->bfs.size();
    for (int i = 0; i < n; ++i)
        for (int j = i; j < n; ++j) {
            const CGO* a = m->bfs[i];
            const CGO* b = m->bfs[j];
            double mx = 0, my = 0, mz = 0;
            for (const Prim& g1 : a->g)
                for (const Prim& g2 : b->g) {
                    Mat sx, sy, sz;
                    Overlap1D(g1.alpha, g2.alpha, g1.c.x, g2.c.x, g1.qn.l, g2.qn.l, sx);
                    Overlap1D(g1.alpha, g2.alpha, g1.c.y, g2.c.y, g1.qn.m, g2.qn.m, sy);
                    Overlap1D(g1.alpha, g2.alpha, g1.c.z, g2.c.z, g1.qn.n, g2.qn.n, sz);
                    const V3 d = g1.c - g2.c;
                    const double o = 1. / (g1.alpha + g2.alpha);
                    const double f = g1.norm * g2.norm * g1.coef * g2.coef *
                                     std::exp(-g1.alpha * g2.alpha * o * dot(d, d)) * std::pow(kPi * o, 1.5);
                    const unsigned l1 = g1.qn.l, l2 = g2.qn.l, m1 = g1.qn.m, m2 = g2.qn.m, n1 = g1.qn.n, n2 = g2.qn.n;
                    const double X = sx(l1, l2), Y = sy(m1, m2), Z = sz(n1, n2);
                    mx += f * (sx(l1 + 1, l2) + g1.c.x * X) * Y * Z;
                    my += f * X * (sy(m1 + 1, m2) + g1.c.y * Y) * Z;
                    mz += f * X * Y * (sz(n1 + 1, n2) + g1.c.z * Z);
                }
            Mx[(size_t)i * n + j] = Mx[(size_t)j * n + i] = mx;
            My[(size_t)i * n + j] = My[(size_t)j * n + i] = my;
            Mz[(size_t)i * n + j] = Mz[(size_t)j * n + i] = mz;
        }
    return 0;
}

int ref_jk_from_packed(int n, const double* eri, const double* D, double* J, double* K) {
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            double jv = 0, kv = 0;
            for (int k = 0; k < n; ++k)
                for (int l = 0; l < n; ++l) {
                    const double d = D[(size_t)k * n + l];
                    jv += d * eri[ref_packed_index(i, j, k, l)];
                    kv += d * eri[ref_packed_index(i, l, k, j)];
                }
            if (J) J[(size_t)i * n + j] = jv;
            if (K) K[(size_t)i * n + j] = kv;
        }
    return 0;
}

void ref_scf_defaults(ref_scf_params* p) {
    /* HartreeFock/HartreeFockAlgorithm.cpp:10-11 and UnrestrictedHartreeFock.cpp:9 */
    p->unrestricted = 0;
    p->max_iterations = 3000;
    p->alpha = 0.75;
    p->init_guess = 0.75;
    p->use_diis = 1;
    p->max_diis_iterations = 1000;
    p->normal_iter_after_diis = 500;
    p->add_asymmetry = 1;
    p->asymmetry = 0.1;
}

/* HartreeFockAlgorithm::Init/Calculate + RHF/UHF::Step (see file header of each block) */
int ref_scf(ref_mol* m, const ref_scf_params* p, ref_scf_result* r, double* D_out) {
    const int n = (int)m->bfs.size();
    const size_t nn = (size_t)n * n;
    std::vector<double> S(nn), T(nn), Vn(nn), h(nn);
    ref_one_electron(m, S.data(), T.data(), Vn.data());
    for (size_t i = 0; i < nn; ++i) h[i] = T[i] + Vn[i];
    if (m->field[0] != 0 || m->field[1] != 0 || m->field[2] != 0) {
        /* h = T + V + Ex Mx + Ey My + Ez Mz (HartreeFock/HartreeFockAlgorithm.cpp:38) */
        std::vector<double> M(3 * nn);
        ref_moment(m, M.data(), M.data() + nn, M.data() + 2 * nn);
        for (size_t i = 0; i < nn; ++i)
            h[i] += m->field[0] * M[i] + m->field[1] * M[nn + i] + m->field[2] * M[2 * nn + i];
    }
    const double Enuc = ref_mol_nuclear_repulsion(m);
    std::vector<double> eri(ref_packed_len(n));
    ref_eri_packed(m, eri.data(), eri.size());
    /* S^-1/2 = U s^-1/2 U^T, symmetrised (HartreeFock/HartreeFockAlgorithm.cpp:48-59) */
    std::vector<double> sw, U;
    JacobiEigen(n, S, sw, U);
    std::vector<double> X(nn, 0.0);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            double s = 0;
            for (int k = 0; k < n; ++k) s += U[(size_t)i * n + k] * (1.0 / std::sqrt(sw[k])) * U[(size_t)j * n + k];
            X[(size_t)i * n + j] = s;
        }
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < i; ++j) {
            const double v = 0.5 * (X[(size_t)i * n + j] + X[(size_t)j * n + i]);
            X[(size_t)i * n + j] = X[(size_t)j * n + i] = v;
        }
    const bool uhf = p->unrestricted != 0;
    /* occupations: RHF HartreeFock/RestrictedHartreeFock.cpp:18-32, UHF UnrestrictedHartreeFock.cpp:20-60 */
    int occA, occB;
    const int ne = (int)(m->alphaE + m->betaE);
    if (!uhf) {
        occA = ne / 2;
        if (!occA) occA = 1;
        if (occA > n) occA = n;
        occB = occA;
    } else {
        occB = ne / 2;
        occA = ne - occB;
        if (m->alphaE > 0 || m->betaE > 0) { occA = (int)m->alphaE; occB = (int)m->betaE; }
        if (occA + occB > 2 * n) { occA = occB = n; }
        if (occA > n) { occB += occA - n; occA = n; }
        if (occB > n) occB = n;
    }
    std::vector<double> DA(nn, 0.0), DB(nn, 0.0);
    DIIS diisA, diisB;
    double energy = 0, homo = 0, lastErr = 0;
    bool useDIIS = p->use_diis != 0;
    std::vector<double> epsA, epsB, CA, CB;

    auto diag = [&](const std::vector<double>& F, std::vector<double>& eps, std::vector<double>& C) {
        std::vector<double> Ft = MatMul(n, MatMul(n, X, F), X);
        std::vector<double> Cp;
        JacobiEigen(n, Ft, eps, Cp);
        C = MatMul(n, X, Cp);
    };
    auto density = [&](const std::vector<double>& C, int occ, double w) {
        std::vector<double> D(nn, 0.0);
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) {
                double s = 0;
                for (int v = 0; v < occ; ++v) s += w * C[(size_t)i * n + v] * C[(size_t)j * n + v];
                D[(size_t)i * n + j] = s;
            }
        return D;
    };
    auto errmat = [&](const std::vector<double>& D, const std::vector<double>& F) {
        std::vector<double> a = MatMul(n, MatMul(n, S, D), F), b = MatMul(n, MatMul(n, F, D), S);
        for (size_t i = 0; i < nn; ++i) a[i] -= b[i];
        return a;
    };

    auto step = [&](int iter) -> double {
        std::vector<double> FA(nn), FB(nn);
        if (iter == 0) {
            for (int i = 0; i < n; ++i)
                for (int j = 0; j < n; ++j)
                    FA[(size_t)i * n + j] = p->init_guess > 0
                        ? p->init_guess * S[(size_t)i * n + j] * (h[(size_t)i * n + i] + h[(size_t)j * n + j]) * 0.5
                        : h[(size_t)i * n + j];
            FB = FA;
            if (uhf && p->add_asymmetry && n > 1) {
                FA[1] += p->asymmetry;
                FA[n] = FA[1];
            }
        } else if (!uhf) {
            /* HartreeFock/RestrictedHartreeFock.cpp:189-202 */
            std::vector<double> G(nn, 0.0);
            for (int i = 0; i < n; ++i)
                for (int j = 0; j < n; ++j)
                    for (int k = 0; k < n; ++k)
                        for (int l = 0; l < n; ++l) {
                            const double coulomb = eri[ref_packed_index(i, j, k, l)];
                            const double exchange = eri[ref_packed_index(i, l, k, j)];
                            G[(size_t)i * n + j] += DA[(size_t)k * n + l] * (coulomb - 0.5 * exchange);
                        }
            for (size_t i = 0; i < nn; ++i) FA[i] = h[i] + G[i];
        } else {
            /* HartreeFock/UnrestrictedHartreeFock.cpp:228-238 */
            std::vector<double> GA(nn, 0.0), GB(nn, 0.0);
            for (int i = 0; i < n; ++i)
                for (int j = 0; j < n; ++j)
                    for (int k = 0; k < n; ++k)
                        for (int l = 0; l < n; ++l) {
                            const double coulomb = eri[ref_packed_index(i, j, k, l)];
                            const double exchange = eri[ref_packed_index(i, l, k, j)];
                            const size_t kl = (size_t)k * n + l;
                            GA[(size_t)i * n + j] += DA[kl] * (coulomb - exchange) + DB[kl] * coulomb;
                            GB[(size_t)i * n + j] += DB[kl] * (coulomb - exchange) + DA[kl] * coulomb;
                        }
            for (size_t i = 0; i < nn; ++i) { FA[i] = h[i] + GA[i]; FB[i] = h[i] + GB[i]; }
        }
        if (!uhf) {  /* RHF symmetrises F (HartreeFock/RestrictedHartreeFock.cpp:206) */
            for (int i = 0; i < n; ++i)
                for (int j = 0; j < i; ++j) {
                    const double v = (FA[(size_t)i * n + j] + FA[(size_t)j * n + i]) * 0.5;
                    FA[(size_t)i * n + j] = FA[(size_t)j * n + i] = v;
                }
        }
        /* DIISStep */
        bool used = false;
        if (useDIIS && iter && iter < p->max_diis_iterations) {
            if (!uhf) {
                if (diisA.Add(FA, errmat(DA, FA))) { lastErr = diisA.Estimate(FA); used = true; }
            } else {
                diisA.Add(FA, errmat(DA, FA));
                if (diisB.Add(FB, errmat(DB, FB))) {
                    lastErr = diisB.Estimate(FB) + diisA.Estimate(FA);
                    used = true;
                }
            }
        } else
            lastErr = 0;
        diag(FA, epsA, CA);
        if (uhf) diag(FB, epsB, CB);
        std::vector<double> nDA = density(CA, occA, uhf ? 1.0 : 2.0);
        std::vector<double> nDB;
        if (uhf) nDB = density(CB, occB, 1.0);
        /* CalculateEnergy: RHF HartreeFock/RestrictedHartreeFock.cpp:209-249, UHF UnrestrictedHartreeFock.cpp:246-293 */
        double e = 0;
        for (int i = 1; i < n; ++i)
            for (int j = 0; j < i; ++j)
                e += (uhf ? nDA[(size_t)i * n + j] + nDB[(size_t)i * n + j] : nDA[(size_t)i * n + j]) * h[(size_t)j * n + i];
        e *= 2.;
        for (int i = 0; i < n; ++i)
            e += (uhf ? nDA[(size_t)i * n + i] + nDB[(size_t)i * n + i] : nDA[(size_t)i * n + i]) * h[(size_t)i * n + i];
        if (!uhf) {
            e *= 0.5;
            for (int v = 0; v < occA; ++v) e += epsA[v];
            homo = epsA[occA - 1];
        } else {
            for (int v = 0; v < occA; ++v) e += epsA[v];
            for (int v = 0; v < occB; ++v) e += epsB[v];
            homo = std::max(occA ? epsA[occA - 1] : 0, occB ? epsB[occB - 1] : 0);
            e /= 2.;
        }
        energy = e + Enuc;
        double rms = 0;
        for (size_t i = 0; i < nn; ++i) { const double d = nDA[i] - DA[i]; rms += d * d; }
        if (uhf) for (size_t i = 0; i < nn; ++i) { const double d = nDB[i] - DB[i]; rms += d * d; }
        rms = std::sqrt(rms);
        if (used) { DA = nDA; if (uhf) DB = nDB; }
        else {
            for (size_t i = 0; i < nn; ++i) DA[i] = p->alpha * nDA[i] + (1. - p->alpha) * DA[i];
            if (uhf) for (size_t i = 0; i < nn; ++i) DB[i] = p->alpha * nDB[i] + (1. - p->alpha) * DB[i];
        }
        return rms;
    };

    /* HartreeFock/HartreeFockAlgorithm.cpp:105-192 */
    double cur = 0, prev = std::numeric_limits<double>::infinity();
    int iter = 0;
    bool converged = false;
    for (; iter < p->max_iterations; ++iter) {
        const double rms = step(iter);
        cur = energy;
        if (std::fabs(prev - cur) <= (useDIIS ? 1e-11 : 1e-13) && rms < 1e-10 && lastErr < 1e-10) { converged = true; break; }
        prev = cur;
    }
    if (useDIIS && iter < p->max_diis_iterations && converged && p->normal_iter_after_diis) {
        useDIIS = false;
        for (int i = 0; i < p->normal_iter_after_diis; ++i) { step(iter + i); cur = energy; prev = cur; }
        iter += p->normal_iter_after_diis;
        useDIIS = true;
    }
    if (!converged) {
        for (; iter < p->max_iterations; ++iter) {
            const double rms = step(iter);
            cur = energy;
            if (std::fabs(prev - cur) <= 1e-13 && rms < 1e-10) { converged = true; break; }
            prev = cur;
        }
    }
    r->energy = cur;
    r->homo = homo;
    r->iterations = iter + 1;
    r->converged = converged ? 1 : 0;
    r->s2 = 0;
    if (uhf) {
        /* <S^2> = ((Na-Nb)/2)^2 + (Na+Nb)/2 - sum_ij |(Ca_occ^T S Cb_occ)_ij|^2 (not in the reference) */
        double ov = 0;
        std::vector<double> SC = MatMul(n, S, CB);
        for (int i = 0; i < occA; ++i)
            for (int j = 0; j < occB; ++j) {
                double s = 0;
                for (int k = 0; k < n; ++k) s += CA[(size_t)k * n + i] * SC[(size_t)k * n + j];
                ov += s * s;
            }
        const double sz = 0.5 * (occA - occB);
        r->s2 = sz * sz + 0.5 * (occA + occB) - ov;
    }
    if (D_out) {
        std::memcpy(D_out, DA.data(), nn * sizeof(double));
        if (uhf) std::memcpy(D_out + nn, DB.data(), nn * sizeof(double));
    }
    return 0;
}

}  // extern "C"
