"""Write the trimmed basis-set data files this package ships (H-Ne only).

The basis sets are Basis Set Exchange data in NWChem format, the same data the
reference program loads (reference: HartreeFock/sto3g.txt, HartreeFock/*.nw,
parsed by HartreeFock/Basis.cpp:15-158).  Only the first-row elements are kept,
because every molecule this package benchmarks or tests is built from H, C and O.

6-31G** is not among the reference's files.  It is composed here from two of
them: heavy atoms from 6-31G* (6-31G** differs from 6-31G* only on H and He)
and H/He from 6-31+G** (the '+' adds diffuse functions to heavy atoms only, so
the H/He blocks of 6-31+G** are exactly those of 6-31G**).

Usage:  python extract_basis.py [/path/to/reference/HartreeFock]
"""
import os
import re
import sys

ELEMENTS = ["H", "He", "Li", "Be", "B", "C", "N", "O", "F", "Ne"]
FILES = ["sto3g.txt", "6-31g.1.nw", "6-31g_st_.1.nw", "6-31+g_st_.1.nw",
         "6-31+g_st__st_.1.nw"]
SHELL_RE = re.compile(r"^(\w+)\s+(\w+)\s*$")


def read_blocks(path):
    """Return (header_lines, {element: [lines...]}) keeping file order."""
    header, blocks, order = [], {}, []
    current = None
    seen_basis = False
    with open(path) as fh:
        for raw in fh:
            line = raw.rstrip("\n")
            stripped = line.strip()
            if not seen_basis:
                header.append(line)
                if stripped.startswith("BASIS"):
                    seen_basis = True
                continue
            if stripped.startswith("END"):
                break
            m = SHELL_RE.match(line)
            if m:
                current = m.group(1)
                if current not in blocks:
                    blocks[current] = []
                    order.append(current)
                blocks[current].append(line)
            elif stripped.startswith("#") or not stripped:
                continue
            elif current is not None:
                blocks[current].append(line)
    return header, blocks, order


def write(path, header, blocks):
    with open(path, "w") as fh:
        for line in header:
            fh.write(line + "\n")
        for el in ELEMENTS:
            if el in blocks:
                for line in blocks[el]:
                    fh.write(line + "\n")
        fh.write("END\n")


def main():
    src = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/HartreeFock"
    # the package's data directory (this script is the recipe for those files and is not shipped)
    here = os.path.join(os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__)))),
                        "hartreefock_amd", "basis")
    for name in FILES:
        header, blocks, _ = read_blocks(os.path.join(src, name))
        write(os.path.join(here, name), header, blocks)
    # 6-31G** = heavy atoms of 6-31G*, H/He of 6-31+G**
    h1, heavy, _ = read_blocks(os.path.join(src, "6-31g_st_.1.nw"))
    _, light, _ = read_blocks(os.path.join(src, "6-31+g_st__st_.1.nw"))
    merged = dict(heavy)
    for el in ("H", "He"):
        merged[el] = light[el]
    hdr = ["# 6-31G** composed from 6-31G* (Li-Ne) and 6-31+G** (H, He);",
           "# see extract_basis.py for why this equals 6-31G**.",
           'BASIS "ao basis" PRINT']
    write(os.path.join(here, "6-31g_st__st_.nw"), hdr, merged)


if __name__ == "__main__":
    main()
