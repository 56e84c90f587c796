"""Extract the basis-set blocks the known-answer tests need into small fixtures.

The reference ships its basis sets as plain text (`element=Z`, `l=<l> n=<K>`, K lines
`exp coeff`, `end=Z`; parser `src/lible/ints/basis_sets.cpp:214-278`).  The known-answer tests
(`test/tests_ints.cpp:694-1148`) use O and H from def2-svp, def2-tzvp, def2-qzvp-rifit and
def2-universal-jkfit.  This script copies exactly those element blocks, unchanged, into
`tests/golden/basis/<name>.txt` so that the tests run without the reference tree (plus
C and H of cc-pVDZ and C and O of 6-31+G for the one-electron known answers, and H and O of
def2-QZVP, a main basis with g shells).

Run once in the build container:  python tests/golden/make_basis_fixtures.py
"""
import os

REF = "/root/reference/src/lible/ints"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "basis")

WANTED = {
    ("lible_basis/ahlrichs/def2-svp.txt", "def2-svp"): [1, 8],
    ("lible_basis/ahlrichs/def2-tzvp.txt", "def2-tzvp"): [1, 8],
    ("lible_aux_basis/ahlrichs/def2-qzvp-rifit.txt", "def2-qzvp-rifit"): [1, 8],
    ("lible_aux_basis/ahlrichs/def2-universal-jkfit.txt", "def2-universal-jkfit"): [1, 8],
    # the one-electron known answers: overlap / kinetic on C2H4 in cc-pVDZ, nuclear attraction
    # on CO2 in 6-31+G (tests_ints.cpp:167-318)
    ("lible_basis/dunning/cc-pvdz.txt", "cc-pvdz"): [1, 6],
    ("lible_basis/pople/6-31+g.txt", "6-31+g"): [6, 8],
    # a main basis with g functions (O 7s4p3d2f1g, H 4s3p2d1f) for the g-shell parity tests
    ("lible_basis/ahlrichs/def2-qzvp.txt", "def2-qzvp"): [1, 8],
}


def extract(path, elements):
    lines = open(path).read().splitlines()
    out = []
    for z in elements:
        start = lines.index(f"element={z}")
        end = lines.index(f"end={z}", start)
        out.extend(lines[start:end + 1])
    return out


def main():
    os.makedirs(OUT, exist_ok=True)
    for (rel, name), elements in WANTED.items():
        block = extract(os.path.join(REF, rel), elements)
        with open(os.path.join(OUT, name + ".txt"), "w") as f:
            f.write("\n".join(block) + "\n")
        print(name, len(block), "lines")


if __name__ == "__main__":
    main()
