// One-electron matrices S, T, V on the host (the reference builds them in
// HartreeFock/QuantumMatrix.cpp:34-155 from Obara-Saika recursions,
// GaussianOverlap/GaussianKinetic/GaussianNuclear).  This file uses the
// McMurchie-Davidson Hermite expansion instead, shell pair by shell pair:
//   S_ab = E^x_0 E^y_0 E^z_0 (pi/p)^{3/2}
//   T    = sum over directions of  b(2j+1) S_ij - 2b^2 S_i,j+2 - j(j-1)/2 S_i,j-2  (times the other two S)
//   V_ab = -sum_C Z_C 2pi/p sum_tuv E^x_t E^y_u E^z_v R_tuv(p, P - C)
// with the same basis-function normalisation as the ERI path.
#include <cmath>
#include <cstring>

#include "hf_internal.h"

namespace hfhost {

namespace {

constexpr int kMaxL = 2;
constexpr int kI = kMaxL + 1;       // i <= LA
constexpr int kJ = kMaxL + 3;       // j <= LB + 2
constexpr int kT = kI + kJ;

double dfact(int n) {
    double r = 1;
    for (; n > 1; n -= 2) r *= n;
    return r;
}

struct Comp {
    int l, m, n;
};

std::vector<Comp> comps(int L) {
    std::vector<Comp> out;
    for (int l = L; l >= 0; --l)
        for (int m = L - l; m >= 0; --m) out.push_back({l, m, L - l - m});
    return out;
}

double comp_norm(const Comp& c) {
    return 1.0 / std::sqrt(dfact(2 * c.l - 1) * dfact(2 * c.m - 1) * dfact(2 * c.n - 1));
}

// Hermite expansion coefficients E[i][j][t] for one direction.
void hermite_e(double a, double b, double X, int imax, int jmax, double E[kI][kJ][kT]) {
    const double p = a + b, q = a * b / p, o2p = 0.5 / p;
    std::memset(E, 0, sizeof(double) * kI * kJ * kT);
    E[0][0][0] = std::exp(-q * X * X);
    for (int i = 0; i <= imax; ++i) {
        for (int j = 0; j <= jmax; ++j) {
            if (i == 0 && j == 0) continue;
            for (int t = 0; t <= i + j; ++t) {
                double v;
                if (j == 0) {
                    const int ii = i - 1;
                    v = -(q * X / a) * E[ii][0][t] + (t + 1 <= ii ? (t + 1) * E[ii][0][t + 1] : 0.0);
                    if (t > 0) v += o2p * E[ii][0][t - 1];
                } else {
                    const int jj = j - 1;
                    v = (q * X / b) * E[i][jj][t] + (t + 1 <= i + jj ? (t + 1) * E[i][jj][t + 1] : 0.0);
                    if (t > 0) v += o2p * E[i][jj][t - 1];
                }
                E[i][j][t] = v;
            }
        }
    }
}

}  // namespace

void one_electron(const Molecule& mol, double* S, double* T, double* V) {
    const int n = mol.nbf;
    const size_t nn = (size_t)n * n;
    std::memset(S, 0, nn * sizeof(double));
    std::memset(T, 0, nn * sizeof(double));
    std::memset(V, 0, nn * sizeof(double));
    const int natoms = (int)mol.Z.size();
    for (size_t sa = 0; sa < mol.shells.size(); ++sa) {
        for (size_t sb = 0; sb <= sa; ++sb) {
            const Shell& A = mol.shells[sa];
            const Shell& B = mol.shells[sb];
            const std::vector<Comp> ca = comps(A.L), cb = comps(B.L);
            const int na = (int)ca.size(), nb = (int)cb.size();
            std::vector<double> s(na * nb, 0.0), t(na * nb, 0.0), v(na * nb, 0.0);
            const double AB[3] = {A.xyz[0] - B.xyz[0], A.xyz[1] - B.xyz[1], A.xyz[2] - B.xyz[2]};
            const int L = A.L + B.L;
            for (size_t pa = 0; pa < A.alpha.size(); ++pa)
                for (size_t pb = 0; pb < B.alpha.size(); ++pb) {
                    const double a = A.alpha[pa], b = B.alpha[pb], p = a + b;
                    const double w = A.coef[pa] * B.coef[pb];
                    double E[3][kI][kJ][kT];
                    for (int d = 0; d < 3; ++d) hermite_e(a, b, AB[d], A.L, B.L + 2, E[d]);
                    const double sq = std::sqrt(M_PI / p);
                    const double P[3] = {(a * A.xyz[0] + b * B.xyz[0]) / p, (a * A.xyz[1] + b * B.xyz[1]) / p,
                                         (a * A.xyz[2] + b * B.xyz[2]) / p};
                    // Hermite Coulomb integrals R_tuv for each nucleus
                    std::vector<double> Rall((size_t)natoms * (L + 1) * (L + 1) * (L + 1));
                    for (int c = 0; c < natoms; ++c) {
                        const double PC[3] = {P[0] - mol.xyz[3 * c], P[1] - mol.xyz[3 * c + 1], P[2] - mol.xyz[3 * c + 2]};
                        double F[2 * kMaxL + 1];
                        boys_host(L, p * (PC[0] * PC[0] + PC[1] * PC[1] + PC[2] * PC[2]), F);
                        // R[n][t][u][v] for t+u+v <= L-n
                        const int D = L + 1;
                        std::vector<double> R((size_t)D * D * D * D, 0.0);
                        auto idx = [D](int nn_, int t_, int u_, int v_) { return ((nn_ * D + t_) * D + u_) * D + v_; };
                        double fac = 1.0;
                        for (int k = 0; k <= L; ++k) {
                            R[idx(k, 0, 0, 0)] = fac * F[k];
                            fac *= -2.0 * p;
                        }
                        for (int k = L - 1; k >= 0; --k)
                            for (int tot = 1; tot <= L - k; ++tot)
                                for (int t_ = 0; t_ <= tot; ++t_)
                                    for (int u_ = 0; u_ <= tot - t_; ++u_) {
                                        const int v_ = tot - t_ - u_;
                                        double val;
                                        if (t_ > 0) {
                                            val = PC[0] * R[idx(k + 1, t_ - 1, u_, v_)];
                                            if (t_ > 1) val += (t_ - 1) * R[idx(k + 1, t_ - 2, u_, v_)];
                                        } else if (u_ > 0) {
                                            val = PC[1] * R[idx(k + 1, t_, u_ - 1, v_)];
                                            if (u_ > 1) val += (u_ - 1) * R[idx(k + 1, t_, u_ - 2, v_)];
                                        } else {
                                            val = PC[2] * R[idx(k + 1, t_, u_, v_ - 1)];
                                            if (v_ > 1) val += (v_ - 1) * R[idx(k + 1, t_, u_, v_ - 2)];
                                        }
                                        R[idx(k, t_, u_, v_)] = val;
                                    }
                        for (int t_ = 0; t_ <= L; ++t_)
                            for (int u_ = 0; u_ <= L - t_; ++u_)
                                for (int v_ = 0; v_ <= L - t_ - u_; ++v_)
                                    Rall[(((size_t)c * (L + 1) + t_) * (L + 1) + u_) * (L + 1) + v_] = R[idx(0, t_, u_, v_)];
                    }
                    for (int i = 0; i < na; ++i)
                        for (int j = 0; j < nb; ++j) {
                            const int li[3] = {ca[i].l, ca[i].m, ca[i].n}, lj[3] = {cb[j].l, cb[j].m, cb[j].n};
                            double s1[3], t1[3];
                            for (int d = 0; d < 3; ++d) {
                                auto S1 = [&](int ii, int jj) { return jj < 0 ? 0.0 : E[d][ii][jj][0] * sq; };
                                s1[d] = S1(li[d], lj[d]);
                                t1[d] = b * (2 * lj[d] + 1) * S1(li[d], lj[d]) - 2 * b * b * S1(li[d], lj[d] + 2) -
                                        0.5 * lj[d] * (lj[d] - 1) * S1(li[d], lj[d] - 2);
                            }
                            s[i * nb + j] += w * s1[0] * s1[1] * s1[2];
                            t[i * nb + j] += w * (t1[0] * s1[1] * s1[2] + s1[0] * t1[1] * s1[2] + s1[0] * s1[1] * t1[2]);
                            double vv = 0;
                            for (int c = 0; c < natoms; ++c) {
                                double acc = 0;
                                for (int t_ = 0; t_ <= li[0] + lj[0]; ++t_)
                                    for (int u_ = 0; u_ <= li[1] + lj[1]; ++u_)
                                        for (int v_ = 0; v_ <= li[2] + lj[2]; ++v_)
                                            acc += E[0][li[0]][lj[0]][t_] * E[1][li[1]][lj[1]][u_] *
                                                   E[2][li[2]][lj[2]][v_] *
                                                   Rall[(((size_t)c * (L + 1) + t_) * (L + 1) + u_) * (L + 1) + v_];
                                vv -= mol.Z[c] * acc;
                            }
                            v[i * nb + j] += w * 2.0 * M_PI / p * vv;
                        }
                }
            for (int i = 0; i < na; ++i)
                for (int j = 0; j < nb; ++j) {
                    const double f = comp_norm(ca[i]) * comp_norm(cb[j]);
                    const size_t r = (size_t)(A.bf + i), c = (size_t)(B.bf + j);
                    S[r * n + c] = S[c * n + r] = f * s[i * nb + j];
                    T[r * n + c] = T[c * n + r] = f * t[i * nb + j];
                    V[r * n + c] = V[c * n + r] = f * v[i * nb + j];
                }
        }
    }
}

}  // namespace hfhost
