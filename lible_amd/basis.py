"""Basis-set input for the ERI path: the reference's text format and the seeded synthetic shells.

A basis is handed to the library as flat shell arrays, atoms in input order and shells in basis
order within each atom (the ordering of `constructShells`, `src/lible/ints/shell.cpp:89-135`):

    shell_atom[nsh]  int32   atom index of each shell
    shell_l[nsh]     int32   angular momentum
    shell_nprim[nsh] int32   contraction depth K
    exps[sum K]      float64 primitive exponents
    coeffs[sum K]    float64 raw contraction coefficients (normalisation is applied inside)

Coordinates are in bohr.  `ANG_TO_BOHR` is the reference's conversion (`defs.hpp:9-12`).
"""
from __future__ import annotations

import dataclasses
import os

import numpy as np

BOHR_TO_ANG = 0.5291772083
ANG_TO_BOHR = 1.0 / BOHR_TO_ANG


@dataclasses.dataclass
class Shells:
    atom: np.ndarray
    l: np.ndarray
    nprim: np.ndarray
    exps: np.ndarray
    coeffs: np.ndarray

    @property
    def nsh(self) -> int:
        return int(self.l.size)

    @property
    def nbf(self) -> int:
        return int(np.sum(2 * self.l + 1))

    @staticmethod
    def from_lists(per_atom: list[list[tuple[int, list[float], list[float]]]]) -> "Shells":
        atom, l, nprim, exps, coeffs = [], [], [], [], []
        for ia, shells in enumerate(per_atom):
            for lv, e, c in shells:
                if len(e) != len(c):
                    raise ValueError("number of exponents/coefficients doesn't match")
                atom.append(ia)
                l.append(lv)
                nprim.append(len(e))
                exps.extend(e)
                coeffs.extend(c)
        return Shells(np.asarray(atom, np.int32), np.asarray(l, np.int32),
                      np.asarray(nprim, np.int32), np.asarray(exps, np.float64),
                      np.asarray(coeffs, np.float64))


@dataclasses.dataclass
class Molecule:
    coords: np.ndarray          # (natoms, 3) bohr
    shells: Shells
    aux: Shells | None = None

    @property
    def natoms(self) -> int:
        return int(self.coords.shape[0])


def read_basis_file(path: str, atomic_nr: int) -> list[tuple[int, list[float], list[float]]]:
    """Shells of one element from a basis file in the reference's text format
    (`basis_sets.cpp:214-278`): `element=Z`, then `l=<l> n=<K>` headers with K `exp coeff` lines,
    closed by `end=Z`."""
    lines = open(path).read().splitlines()
    key = f"element={atomic_nr}"
    if key not in lines:
        raise RuntimeError(f"Basis set {os.path.basename(path)} not found for element {atomic_nr}!")
    i = lines.index(key) + 1
    shells = []
    cur = None
    while i < len(lines) and lines[i] != f"end={atomic_nr}":
        line = lines[i]
        if "l=" in line:
            if cur is not None:
                shells.append(cur)
            cur = (int(line[2]), [], [])
        else:
            e, c = line.split()[:2]
            cur[1].append(float(e))
            cur[2].append(float(c))
        i += 1
    if cur is not None:
        shells.append(cur)
    return shells


def molecule_from_basis_files(main_path: str, atomic_nrs, coords_angstrom,
                              aux_path: str | None = None) -> Molecule:
    """A molecule with a named basis, coordinates given in angstrom as the reference's
    `Structure(basis, atomic_nrs, coords)` constructor takes them (`structure.cpp:12-40`)."""
    coords = np.asarray(coords_angstrom, np.float64) * ANG_TO_BOHR
    main = Shells.from_lists([read_basis_file(main_path, z) for z in atomic_nrs])
    aux = None
    if aux_path is not None:
        aux = Shells.from_lists([read_basis_file(aux_path, z) for z in atomic_nrs])
    return Molecule(coords, main, aux)


# ---------------------------------------------------------------------------------------------
# Seeded synthetic workloads.  Per-atom shell recipe (l, K, count); exponents log-uniform in a
# range, contraction coefficients uniform in [0.1, 1] (raw; the library normalises them).

RECIPE_SMALL = [(0, 3, 3), (1, 2, 2), (2, 1, 1)]              # 14 AO per atom
RECIPE_MAIN = [(0, 4, 4), (1, 3, 3), (2, 1, 2), (3, 1, 1)]     # 30 AO per atom
RECIPE_AUX = [(0, 1, 6), (1, 1, 4), (2, 1, 3), (3, 1, 2), (4, 1, 1)]   # 56 AO per atom


def random_atoms(rng: np.random.Generator, natoms: int, box: float, min_dist: float = 1.2):
    """Rejection-sample atom positions uniformly in a cube of side `box` (bohr) until every pair
    is at least `min_dist` apart."""
    pts = []
    while len(pts) < natoms:
        p = rng.uniform(0.0, box, 3)
        if all(np.linalg.norm(p - q) >= min_dist for q in pts):
            pts.append(p)
    return np.asarray(pts, np.float64)


def random_shells(rng: np.random.Generator, natoms: int, recipe, exp_lo: float = 0.1,
                  exp_hi: float = 1e3) -> Shells:
    per_atom = []
    for _ in range(natoms):
        shells = []
        for l, K, count in recipe:
            for _ in range(count):
                e = np.exp(rng.uniform(np.log(exp_lo), np.log(exp_hi), K))
                c = rng.uniform(0.1, 1.0, K)
                shells.append((l, list(e), list(c)))
        per_atom.append(shells)
    return Shells.from_lists(per_atom)


def synthetic_molecule(seed: int, natoms: int, box: float, recipe=RECIPE_MAIN,
                       aux_recipe=None) -> Molecule:
    rng = np.random.default_rng(seed)
    coords = random_atoms(rng, natoms, box)
    shells = random_shells(rng, natoms, recipe)
    aux = random_shells(rng, natoms, aux_recipe, 0.05, 500.0) if aux_recipe else None
    return Molecule(coords, shells, aux)


def density_from_orbitals(seed: int, nbf: int, nocc: int, sigma: float) -> np.ndarray:
    """D = C C^T with C an nbf x nocc matrix of N(0, sigma) entries."""
    rng = np.random.default_rng(seed)
    C = rng.normal(0.0, sigma, (nbf, nocc))
    return C @ C.T


# The workloads named by the benchmark: (seed, natoms, box bohr, recipe, nocc, sigma).
CONFIGS = {
    "c1_eri4_dense": dict(seed=1001, natoms=6, box=6.0, recipe=RECIPE_SMALL),
    "c2_jk_fp64": dict(seed=1002, natoms=40, box=14.0, recipe=RECIPE_MAIN, nocc=100, sigma=0.1),
    "c3_jk_mixed": dict(seed=1003, natoms=120, box=22.0, recipe=RECIPE_MAIN, nocc=300,
                        sigma=0.05),
    "c5_gradient": dict(seed=1005, natoms=60, box=16.0, recipe=RECIPE_MAIN, nocc=150,
                        sigma=0.07),
}


def config_molecule(name: str) -> Molecule:
    c = CONFIGS[name]
    return synthetic_molecule(c["seed"], c["natoms"], c["box"], c["recipe"])


# The batched RI workload: 256 independent molecules of 3..30 atoms (count drawn per molecule),
# main basis 30 functions per atom, auxiliary 6s 4p 3d 2f 1g = 56 per atom, at the atom density
# of the 40-atom / 14-bohr workload (68.6 bohr^3 per atom).
RI_BATCH = dict(seed=1004, nmol=256, natoms_lo=3, natoms_hi=30, bohr3_per_atom=68.6)


def batch_molecules(nmol: int | None = None, seed: int | None = None) -> list[Molecule]:
    c = RI_BATCH
    seed = c["seed"] if seed is None else seed
    nmol = c["nmol"] if nmol is None else nmol
    out = []
    for k in range(nmol):
        rng = np.random.default_rng(seed * 100000 + k)
        n = int(rng.integers(c["natoms_lo"], c["natoms_hi"] + 1))
        box = (c["bohr3_per_atom"] * n) ** (1.0 / 3.0)
        out.append(synthetic_molecule(seed * 100000 + k, n, box, RECIPE_MAIN, RECIPE_AUX))
    return out
