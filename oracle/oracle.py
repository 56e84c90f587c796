"""ctypes binding of the CPU FP64 oracle (oracle/liblbx_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, `__graft_entry__.smoke()` and bench.py's
cpu_baseline leg, always as the results compared against (or the timed CPU baseline).  The
product never imports this module.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_d = ctypes.c_double
_i = ctypes.c_int
_l = ctypes.c_long
_p = ctypes.c_void_p
_dp = ctypes.POINTER(ctypes.c_double)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liblbx_oracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"oracle library missing: {path} (run `make -C oracle`)")
        L = ctypes.CDLL(path)
        L.orc_last_error.restype = ctypes.c_char_p
        L.orc_pure_primitive_norm.restype = _d
        L.orc_pure_primitive_norm.argtypes = [_i, _d]
        L.orc_ecoeffs_shell_abs_sum.restype = _d
        L.orc_ecoeffs_shell_abs_sum.argtypes = [_i, _i, _p]
        L.orc_structure_create.restype = _p
        L.orc_structure_create.argtypes = [_i, _p, _i, _p, _p, _p, _p, _p, _i, _p, _p, _p, _p, _p]
        L.orc_structure_destroy.argtypes = [_p]
        L.orc_structure_dims.argtypes = [_p, _p, _p, _p]
        L.orc_structure_norms.argtypes = [_p, _p]
        for name in ("orc_eri4", "orc_eri4_diag", "orc_eri3", "orc_eri2", "orc_eri2_diag"):
            getattr(L, name).argtypes = [_p, _p]
        L.orc_ri_sums.argtypes = [_p, _p]
        for name in ("orc_eri4d1", "orc_eri3d1", "orc_eri2d1"):
            getattr(L, name).argtypes = [_p, _p]
        L.orc_deploy_sum.argtypes = [_p, _i, _i, _p]
        L.orc_deploy_sum2.argtypes = [_p, _i, _p]
        L.orc_blocks.argtypes = [_p, _i, _l, _p, _p]
        L.orc_oneel.argtypes = [_p, _i, _l, _p, _p]
        L.orc_jk.argtypes = [_p, _p, _p, _p, _d, _i, _p]
        L.orc_jk_grad.argtypes = [_p, _p, _p, _d]
        L.orc_boys.argtypes = [_i, _d, _p]
        L.orc_boys_many.argtypes = [_i, _l, _p, _p]
        L.orc_rints3d.argtypes = [_i, _d, _p, _p, _p]
        L.orc_sph_trafo.argtypes = [_i, _i, _p, _p, _p]
        L.orc_ecoeffs_recurrence1.argtypes = [_d, _i, _p]
        L.orc_ecoeffs_recurrence2.argtypes = [_d, _d, _i, _i, _d, _d, _d, _p]
        L.orc_ecoeffs_recurrence2_n1.argtypes = [_d, _d, _i, _i, _d, _d, _d, _d, _d, _p]
        L.orc_ecoeffs_primitive_pair.argtypes = [_d, _d, _i, _i, _p, _p, _p, _p, _i, _p]
        _LIB = L
    return _LIB


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def _check(rc):
    if rc != 0:
        raise RuntimeError(lib().orc_last_error().decode())


def boys(n: int, x: float) -> np.ndarray:
    out = np.zeros(n + 1)
    lib().orc_boys(n, x, _ptr(out))
    return out


def boys_many(n: int, x: np.ndarray) -> np.ndarray:
    x = np.ascontiguousarray(x, np.float64)
    out = np.zeros((x.size, n + 1))
    lib().orc_boys_many(n, x.size, _ptr(x), _ptr(out))
    return out


def rints3d(l: int, p: float, xyz, fnx) -> np.ndarray:
    xyz = np.ascontiguousarray(xyz, np.float64)
    fnx = np.ascontiguousarray(fnx, np.float64)
    out = np.zeros((l + 1, l + 1, l + 1))
    lib().orc_rints3d(l, p, _ptr(xyz), _ptr(fnx), _ptr(out))
    return out


def sph_trafo(l: int):
    cap = 2000
    s = np.zeros(cap, np.int32)
    c = np.zeros(cap, np.int32)
    v = np.zeros(cap)
    n = lib().orc_sph_trafo(l, cap, _ptr(s), _ptr(c), _ptr(v))
    if n < 0:
        raise RuntimeError(lib().orc_last_error().decode())
    return list(zip(s[:n].tolist(), c[:n].tolist(), v[:n].tolist()))


def pure_primitive_norm(l: int, e: float) -> float:
    return lib().orc_pure_primitive_norm(l, e)


def ecoeffs_recurrence1(one_o_2a: float, l: int) -> np.ndarray:
    out = np.zeros((l + 1, l + 1))
    lib().orc_ecoeffs_recurrence1(one_o_2a, l, _ptr(out))
    return out


def ecoeffs_recurrence2(a, b, la, lb, PA, PB, Kab) -> np.ndarray:
    out = np.zeros((la + 1, lb + 1, la + lb + 1))
    lib().orc_ecoeffs_recurrence2(a, b, la, lb, PA, PB, Kab, _ptr(out))
    return out


def ecoeffs_recurrence2_n1(a, b, la, lb, PA, PB, Kab, A, B) -> np.ndarray:
    out = np.zeros((la + 1, lb + 1, la + lb + 1))
    lib().orc_ecoeffs_recurrence2_n1(a, b, la, lb, PA, PB, Kab, A, B, _ptr(out))
    return out


def ecoeffs_primitive_pair(a, b, la, lb, xa, xb, A=None, B=None, deriv=0) -> np.ndarray:
    xa = np.ascontiguousarray(xa, np.float64)
    xb = np.ascontiguousarray(xb, np.float64)
    A = np.ascontiguousarray(A if A is not None else np.zeros(3), np.float64)
    B = np.ascontiguousarray(B if B is not None else np.zeros(3), np.float64)
    out = np.zeros((3, la + 1, lb + 1, la + lb + 1))
    lib().orc_ecoeffs_primitive_pair(a, b, la, lb, _ptr(xa), _ptr(xb), _ptr(A), _ptr(B),
                                     deriv, _ptr(out))
    return out


def ecoeffs_shell_abs_sum(l: int, exps) -> float:
    e = np.ascontiguousarray(exps, np.float64)
    return lib().orc_ecoeffs_shell_abs_sum(l, e.size, _ptr(e))


class Structure:
    """Oracle structure built from a lible_amd.basis.Molecule (coordinates in bohr)."""

    def __init__(self, mol):
        L = lib()
        self._keep = []

        def arrs(sh):
            a = [np.ascontiguousarray(sh.atom, np.int32), np.ascontiguousarray(sh.l, np.int32),
                 np.ascontiguousarray(sh.nprim, np.int32),
                 np.ascontiguousarray(sh.exps, np.float64),
                 np.ascontiguousarray(sh.coeffs, np.float64)]
            self._keep.extend(a)
            return a

        coords = np.ascontiguousarray(mol.coords, np.float64)
        self._keep.append(coords)
        m = arrs(mol.shells)
        if mol.aux is not None:
            x = arrs(mol.aux)
            nx = mol.aux.nsh
        else:
            x = [np.zeros(1, np.int32)] * 3 + [np.zeros(1)] * 2
            nx = 0
        self.h = L.orc_structure_create(mol.natoms, _ptr(coords), mol.shells.nsh,
                                        *[_ptr(a) for a in m], nx, *[_ptr(a) for a in x])
        if not self.h:
            raise RuntimeError(L.orc_last_error().decode())
        nbf, nbf_aux, nat = _l(), _l(), _l()
        L.orc_structure_dims(self.h, ctypes.byref(nbf), ctypes.byref(nbf_aux), ctypes.byref(nat))
        self.nbf, self.nbf_aux, self.natoms = nbf.value, nbf_aux.value, nat.value

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_structure_destroy(self.h)
            self.h = None

    def norms(self) -> np.ndarray:
        out = np.zeros(self.nbf)
        lib().orc_structure_norms(self.h, _ptr(out))
        return out

    def eri4(self) -> np.ndarray:
        n = self.nbf
        out = np.zeros((n, n, n, n))
        _check(lib().orc_eri4(self.h, _ptr(out)))
        return out

    def eri4_diag(self) -> np.ndarray:
        out = np.zeros((self.nbf, self.nbf))
        _check(lib().orc_eri4_diag(self.h, _ptr(out)))
        return out

    def eri3(self) -> np.ndarray:
        out = np.zeros((self.nbf, self.nbf, self.nbf_aux))
        _check(lib().orc_eri3(self.h, _ptr(out)))
        return out

    def eri2(self) -> np.ndarray:
        out = np.zeros((self.nbf_aux, self.nbf_aux))
        _check(lib().orc_eri2(self.h, _ptr(out)))
        return out

    def eri2_diag(self) -> np.ndarray:
        out = np.zeros(self.nbf_aux)
        _check(lib().orc_eri2_diag(self.h, _ptr(out)))
        return out

    def ri_sums(self) -> np.ndarray:
        """Sums of |x| and x over the dense eri3 and eri2 tensors: [|eri3|, eri3, |eri2|, eri2]."""
        out = np.zeros(4)
        _check(lib().orc_ri_sums(self.h, _ptr(out)))
        return out

    def eri4d1(self) -> np.ndarray:
        """(12, nbf, nbf, nbf, nbf): d/dR of the centre of each index's shell, x y z."""
        n = self.nbf
        out = np.zeros((12, n, n, n, n))
        _check(lib().orc_eri4d1(self.h, _ptr(out)))
        return out

    def eri3d1(self) -> np.ndarray:
        """(9, nbf, nbf, naux)."""
        out = np.zeros((9, self.nbf, self.nbf, self.nbf_aux))
        _check(lib().orc_eri3d1(self.h, _ptr(out)))
        return out

    def eri2d1(self) -> np.ndarray:
        """(6, naux, naux)."""
        out = np.zeros((6, self.nbf_aux, self.nbf_aux))
        _check(lib().orc_eri2d1(self.h, _ptr(out)))
        return out

    def deploy_sum(self, which: str, use_symm: bool) -> float:
        code = {"eri4": 0, "eri3": 1, "eri2": 2, "eri4d1": 3, "eri3d1": 4, "eri2d1": 5}[which]
        s = _d()
        _check(lib().orc_deploy_sum(self.h, code, int(use_symm), ctypes.byref(s)))
        return s.value

    def deploy_sum2(self, which: str) -> float:
        """Sum of |x| over the blocks of the second-derivative / spin-orbit kernels for all
        ordered pairs (deployERI2D2Kernel / deployERI4SOCKernel / deployERI3SOCKernel)."""
        code = {"eri2d2": 6, "eri4soc": 7, "eri3soc": 8}[which]
        s = _d()
        _check(lib().orc_deploy_sum2(self.h, code, ctypes.byref(s)))
        return s.value

    BLOCK_KINDS = {"eri4": (0, 1), "eri3": (1, 1), "eri2": (2, 1), "eri4d1": (3, 12),
                   "eri3d1": (4, 9), "eri2d1": (5, 6), "eri2d2": (6, 36), "eri4soc": (7, 3),
                   "eri3soc": (8, 3)}

    def blocks(self, which: str, shells: np.ndarray, dims) -> np.ndarray:
        """Blocks of the listed entries (n, 4 / 3 / 2 shell indices) in the caller's shell order
        -> (n, ncomp, *dims) with dims the entries' spherical sizes."""
        code, ncomp = self.BLOCK_KINDS[which]
        sh = np.ascontiguousarray(shells, dtype=np.int32)
        out = np.zeros((sh.shape[0], ncomp, *dims))
        _check(lib().orc_blocks(self.h, code, sh.shape[0], _ptr(sh), _ptr(out)))
        return out

    def overlap(self) -> np.ndarray:
        out = np.zeros((self.nbf, self.nbf))
        _check(lib().orc_oneel(self.h, 0, 0, None, _ptr(out)))
        return out

    def kinetic(self) -> np.ndarray:
        out = np.zeros((self.nbf, self.nbf))
        _check(lib().orc_oneel(self.h, 1, 0, None, _ptr(out)))
        return out

    def external_charges(self, charges) -> np.ndarray:
        """Attraction to point charges, rows (x, y, z bohr, charge); the nuclei give the
        nuclear attraction."""
        c = np.ascontiguousarray(charges, np.float64).reshape(-1, 4)
        out = np.zeros((self.nbf, self.nbf))
        _check(lib().orc_oneel(self.h, 2, c.shape[0], _ptr(c), _ptr(out)))
        return out

    def jk(self, D: np.ndarray, thresh: float = 1e-12, stride: int = 1):
        D = np.ascontiguousarray(D, np.float64)
        J = np.zeros_like(D)
        K = np.zeros_like(D)
        stats = np.zeros(4)
        _check(lib().orc_jk(self.h, _ptr(D), _ptr(J), _ptr(K), thresh, stride, _ptr(stats)))
        return J, K, {"quartets": stats[0], "flops": stats[1], "seconds_setup": stats[2],
                      "seconds_jk": stats[3]}

    def jk_grad(self, D: np.ndarray, thresh: float = 1e-12) -> np.ndarray:
        D = np.ascontiguousarray(D, np.float64)
        g = np.zeros((self.natoms, 3))
        _check(lib().orc_jk_grad(self.h, _ptr(D), _ptr(g), thresh))
        return g
