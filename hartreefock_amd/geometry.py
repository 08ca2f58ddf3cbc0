"""Seeded molecule generators for the benchmark and test workloads.

Every generator returns a list of (Z, (x, y, z)) with coordinates in bohr, the
unit the reference's Molecule works in (reference: HartreeFock/Molecule.cpp:143-199
reads XYZ files as bohr; only the GUI scan converts from angstrom, with
Bohr = 0.529177210903 from HartreeFock/Constants.h:4).
"""
import math

import numpy as np

BOHR_ANGSTROM = 0.529177210903          # reference HartreeFock/Constants.h:4
SYMBOLS = {1: "H", 2: "He", 3: "Li", 4: "Be", 5: "B", 6: "C", 7: "N", 8: "O",
           9: "F", 10: "Ne"}
Z_OF = {v: k for k, v in SYMBOLS.items()}

TETRAHEDRAL = math.acos(-1.0 / 3.0)     # 109.4712...°


def _to_bohr(atoms_ang):
    return [(z, tuple(float(c) / BOHR_ANGSTROM for c in xyz)) for z, xyz in atoms_ang]


def water_cfg1():
    """H2O at the geometry of the reference README (bohr)."""
    return [(8, (0.0, 0.0, 0.24026010)),
            (1, (0.0, 1.43233673, -0.96104039)),
            (1, (0.0, -1.43233673, -0.96104039))]


def water_crawford():
    """H2O at the geometry of reference HartreeFock/Test.cpp:586-596 (bohr)."""
    return [(8, (0.0, -0.143225816552, 0.0)),
            (1, (1.638036840407, 1.136548822547, 0.0)),
            (1, (-1.638036840407, 1.136548822547, 0.0))]


def methane_crawford():
    """CH4 at the geometry of reference HartreeFock/Test.cpp:752-770 (bohr)."""
    a = 1.183771681898
    return [(6, (0.0, 0.0, 0.0)),
            (1, (a, -a, -a)), (1, (a, a, a)), (1, (-a, a, -a)), (1, (-a, -a, a))]


def _unit(v):
    return v / np.linalg.norm(v)


def alkane_zigzag(n, seed=0, jitter=0.05, cc=1.54, ch=1.09, drop_last_h=False):
    """n-alkane C_nH_{2n+2} as a planar all-trans zigzag chain.

    Carbons lie in the xy plane (C-C `cc` angstrom, C-C-C tetrahedral); each
    carbon's hydrogens sit at tetrahedral angles (C-H `ch` angstrom).  Every
    coordinate of every atom is then moved by U(-jitter, +jitter) angstrom from
    numpy's default_rng(seed).  Atom order: carbons along the chain, then the
    hydrogens carbon by carbon.  drop_last_h removes the final hydrogen, giving
    the C_nH_{2n+1} alkyl radical.
    """
    half = TETRAHEDRAL / 2.0
    step = cc * math.sin(half)              # advance along x per bond
    rise = cc * math.cos(half)              # zigzag offset in y
    carbons = [np.array([i * step, (i % 2) * rise, 0.0]) for i in range(n)]

    def virtual(i):                          # zigzag continued one site past either end
        return np.array([i * step, (i % 2) * rise, 0.0])

    zaxis = np.array([0.0, 0.0, 1.0])
    hydrogens = []
    phi = TETRAHEDRAL / 2.0
    for i in range(n):
        c = carbons[i]
        left = carbons[i - 1] if i > 0 else virtual(-1)
        right = carbons[i + 1] if i < n - 1 else virtual(n)
        b1, b2 = _unit(left - c), _unit(right - c)
        u = _unit(-(b1 + b2))
        hs = [c + ch * (u * math.cos(phi) + zaxis * math.sin(phi)),
              c + ch * (u * math.cos(phi) - zaxis * math.sin(phi))]
        if n == 1:                           # methane: H at both virtual sites
            hs += [c + ch * b1, c + ch * b2]
        elif i == 0:
            hs.append(c + ch * b1)
        elif i == n - 1:
            hs.append(c + ch * b2)
        hydrogens.extend(hs)
    if drop_last_h:
        hydrogens = hydrogens[:-1]
    rng = np.random.default_rng(seed)
    atoms = [(6, tuple(p)) for p in carbons] + [(1, tuple(p)) for p in hydrogens]
    out = []
    for z, p in atoms:
        d = rng.uniform(-jitter, jitter, size=3)
        out.append((z, tuple(np.asarray(p) + d)))
    return _to_bohr(out)


def _random_rotation(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
        [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
        [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def water_cluster(nwater, seed=0, box=None, min_oo=2.6, roh=0.9572, angle_deg=104.52):
    """(H2O)_n: rigid waters, random orientations, O centres by seeded rejection
    sampling in a cube of side `box` angstrom (default 10 angstrom scaled to the
    density of 32 waters in 10 angstrom) with every O-O distance >= min_oo."""
    if box is None:
        box = max(3.0, 10.0 * (nwater / 32.0) ** (1.0 / 3.0))
    rng = np.random.default_rng(seed)
    centres = []
    tries = 0
    while len(centres) < nwater:
        tries += 1
        if tries > 1000000:
            raise RuntimeError("water_cluster: packing failed")
        p = rng.uniform(0.0, box, size=3)
        if all(np.linalg.norm(p - c) >= min_oo for c in centres):
            centres.append(p)
    half = math.radians(angle_deg) / 2.0
    local = [np.zeros(3),
             np.array([roh * math.sin(half), 0.0, roh * math.cos(half)]),
             np.array([-roh * math.sin(half), 0.0, roh * math.cos(half)])]
    out = []
    for c in centres:
        r = _random_rotation(rng)
        for z, v in zip((8, 1, 1), local):
            out.append((z, tuple(c + r @ v)))
    return _to_bohr(out)


def decane_cfg2(seed=2024):
    """n-decane C10H22 (194 functions in 6-31G*)."""
    return alkane_zigzag(10, seed=seed)


def water32_cfg3(seed=2024):
    """(H2O)32 (800 functions in 6-31G**)."""
    return water_cluster(32, seed=seed, box=10.0)


def eicosyl_cfg5(seed=2024):
    """Linear C20H41 alkyl radical (462 functions in 6-31+G*)."""
    return alkane_zigzag(20, seed=seed, drop_last_h=True)


def serving_mix(nreq, seed=2024):
    """Request molecules: water clusters (1-12) and alkanes (n=1-16), uniform."""
    rng = np.random.default_rng(seed)
    mols = []
    for i in range(nreq):
        if rng.integers(0, 2) == 0:
            k = int(rng.integers(1, 13))
            mols.append(("water%d" % k, water_cluster(k, seed=seed + 7919 * i)))
        else:
            k = int(rng.integers(1, 17))
            mols.append(("alkane%d" % k, alkane_zigzag(k, seed=seed + 7919 * i)))
    return mols


def electrons(atoms):
    return sum(z for z, _ in atoms)


def to_arrays(atoms):
    z = np.array([a[0] for a in atoms], dtype=np.int32)
    xyz = np.array([a[1] for a in atoms], dtype=np.float64).reshape(-1, 3)
    return z, xyz
