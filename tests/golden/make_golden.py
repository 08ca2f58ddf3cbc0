"""Write the golden fixtures in tests/golden/ from the independent McMurchie-Davidson
code (md_eri.py, mpmath Boys).  Nothing here runs or imports the reference.

Formats follow the reference's test harness (HartreeFock/Test.cpp:497-507, 545-555):
matrices as `i j value`, ERIs as `i j k l value`, 1-based indices.

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, HERE)

import md_eri  # noqa: E402
from hartreefock_amd import geometry as g  # noqa: E402

BASIS = os.path.join(ROOT, "hartreefock_amd", "basis")


def unique_quartets(n):
    for i in range(n):
        for j in range(i + 1):
            ij = i * (i + 1) // 2 + j
            for k in range(n):
                for l in range(k + 1):
                    kl = k * (k + 1) // 2 + l
                    if ij <= kl:
                        yield i, j, k, l


def write_eris(path, bfs, quartets, header):
    with open(path, "w") as fh:
        fh.write("# %s\n# i j k l (ij|kl)   1-based; McMurchie-Davidson with mpmath Boys\n" % header)
        for (i, j, k, l) in quartets:
            v = md_eri.contracted_eri(bfs[i], bfs[j], bfs[k], bfs[l])
            fh.write("%d %d %d %d % .17e\n" % (i + 1, j + 1, k + 1, l + 1, v))


def write_matrix(path, M, header):
    n = M.shape[0]
    with open(path, "w") as fh:
        fh.write("# %s\n# i j value   1-based\n" % header)
        for i in range(n):
            for j in range(i + 1):
                fh.write("%d %d % .17e\n" % (i + 1, j + 1, M[i, j]))


def main():
    # 1. every unique ERI of H2O / STO-3G at the README geometry (406 values)
    atoms = g.water_cfg1()
    bfs = md_eri.build_basis(os.path.join(BASIS, "sto3g.txt"), atoms)
    write_eris(os.path.join(HERE, "eri_h2o_sto3g.txt"), bfs, list(unique_quartets(len(bfs))),
               "H2O STO-3G, O (0,0,0.24026010) H (0,+-1.43233673,-0.96104039) bohr")
    S, T, V = md_eri.one_electron(bfs, atoms)
    for name, M in (("S", S), ("T", T), ("V", V)):
        write_matrix(os.path.join(HERE, "%s_h2o_sto3g.txt" % name.lower()), M, "H2O STO-3G %s" % name)
    # 2. H2O / 6-31G*: one-electron matrices and a seeded sample of unique ERIs (with d)
    bfs = md_eri.build_basis(os.path.join(BASIS, "6-31g_st_.1.nw"), atoms)
    S, T, V = md_eri.one_electron(bfs, atoms)
    for name, M in (("S", S), ("T", T), ("V", V)):
        write_matrix(os.path.join(HERE, "%s_h2o_631gs.txt" % name.lower()), M, "H2O 6-31G* %s" % name)
    rng = np.random.default_rng(11)
    allq = list(unique_quartets(len(bfs)))
    pick = sorted(rng.choice(len(allq), size=150, replace=False))
    write_eris(os.path.join(HERE, "eri_h2o_631gs_sample.txt"), bfs, [allq[p] for p in pick],
               "H2O 6-31G* (README geometry), 150 seeded unique quartets")
    # 3. two carbons in 6-31G*: (dd|dd) quartets with one- and two-centre bra pairs
    catoms = [(6, (0.0, 0.0, 0.0)), (6, (0.3, 1.2, 2.5))]
    bfs = md_eri.build_basis(os.path.join(BASIS, "6-31g_st_.1.nw"), catoms)
    d0 = [i for i, f in enumerate(bfs) if sum(f.lmn) == 2 and f.atom == 0]
    d1 = [i for i, f in enumerate(bfs) if sum(f.lmn) == 2 and f.atom == 1]
    qs = []
    for a in (0, 3, 5):
        for b in (1, 4):
            qs.append((d1[a], d0[b], d1[b], d0[a]))      # two-centre bra and ket
            qs.append((d1[a], d1[b], d0[a], d0[b]))      # one-centre bra and ket
            qs.append((d1[a], d1[b], d1[b], d0[a]))      # one-centre bra, two-centre ket
    write_eris(os.path.join(HERE, "eri_c2_631gs_dddd.txt"), bfs, qs,
               "C2 6-31G*, C (0,0,0) and (0.3,1.2,2.5) bohr: (dd|dd) quartets, exact values")
    # 4. Boys functions F_m(T), m = 0..16, at 40 digits
    Ts = [0.0, 1e-12, 1e-6, 0.05, 0.1, 0.37, 1.0, 2.5, 5.0, 9.99, 15.0, 20.0, 29.9, 35.95, 36.0, 36.07,
          40.0, 50.0, 80.0, 120.0, 159.9, 160.5, 250.0, 1000.0]
    with open(os.path.join(HERE, "boys_mpmath.txt"), "w") as fh:
        fh.write("# T m F_m(T) from mpmath hyp1f1(m+1/2; m+3/2; -T)/(2m+1) at 40 digits\n")
        for T in Ts:
            for m in range(17):
                fh.write("%.17g %d %.17e\n" % (T, m, md_eri.boys_mp(m, T)))


if __name__ == "__main__":
    main()
