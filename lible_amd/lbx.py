"""ctypes binding of liblbx.so, the MI355X SHARK ERI library (C-ABI in include/lbx.h).

Device memory comes from torch (CUDA tensors on the current HIP device); the integrals are
computed by the library's HIP kernels.  There is no CPU fallback: without the built library or
without a GPU every entry point raises.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblbx.so")

LBX_FP64 = 0
LBX_MIXED = 1

# every symbol include/lbx.h declares
SYMBOLS = ["lbx_create", "lbx_destroy", "lbx_last_error", "lbx_dims", "lbx_set_stream",
           "lbx_synchronize", "lbx_boys", "lbx_eri4", "lbx_eri4_diag", "lbx_eri3", "lbx_eri2",
           "lbx_eri2_diag", "lbx_eri4d1", "lbx_eri3d1", "lbx_eri2d1", "lbx_eri4_batch",
           "lbx_eri3_batch", "lbx_eri2_batch", "lbx_eri4d1_batch", "lbx_eri3d1_batch",
           "lbx_eri2d1_batch", "lbx_eri4_batch_static", "lbx_eri4soc_batch", "lbx_eri3soc_batch", "lbx_eri2d2_batch", "lbx_jk", "lbx_create_batch", "lbx_ri_sums", "lbx_jk_grad",
           "lbx_tune", "lbx_set_variant", "lbx_profile_enable",
           "lbx_profile_read", "lbx_set_class_filter", "lbx_set_bound_window", "lbx_ri_jk",
           "lbx_one_electron", "lbx_one_electron_grad", "lbx_set_graphs"]
VARIANTS = ("valu-bra", "valu-ket", "mfma", "valu-bra-f32", "valu-ket-f32", "mfma-f32", "block",
            "block-f32")


class LbxError(RuntimeError):
    pass


class _Basis(ctypes.Structure):
    _fields_ = [("natoms", ctypes.c_int), ("coords", ctypes.c_void_p), ("nshells", ctypes.c_int),
                ("shell_atom", ctypes.c_void_p), ("shell_l", ctypes.c_void_p),
                ("shell_nprim", ctypes.c_void_p), ("exps", ctypes.c_void_p),
                ("coeffs", ctypes.c_void_p)]


class JKStats(ctypes.Structure):
    _fields_ = [("quartets", ctypes.c_double), ("flops", ctypes.c_double),
                ("quartets_fp32", ctypes.c_double)]


class RIStats(ctypes.Structure):
    _fields_ = [("triplets", ctypes.c_double), ("doublets", ctypes.c_double),
                ("flops", ctypes.c_double), ("triplets_fp32", ctypes.c_double),
                ("doublets_fp32", ctypes.c_double)]


class ClassProfile(ctypes.Structure):
    _fields_ = [("la", ctypes.c_int), ("lb", ctypes.c_int), ("lc", ctypes.c_int),
                ("ld", ctypes.c_int), ("kernel", ctypes.c_int), ("calls", ctypes.c_int),
                ("ms", ctypes.c_double), ("quartets", ctypes.c_double),
                ("flops", ctypes.c_double)]


_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise LbxError(f"{LIB_PATH} is not built (run `python -m lible_amd.build`)")
        # One HIP runtime per process: torch carries its own libamdhip64 / HSA runtime.  Loading
        # torch first makes liblbx.so bind to that runtime (same sonames) instead of mapping a
        # second one from /opt/rocm, which could not open the device.
        import torch  # noqa: F401
        L = ctypes.CDLL(LIB_PATH)
        p, i, d, l = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_long
        L.lbx_last_error.restype = ctypes.c_char_p
        L.lbx_create.argtypes = [ctypes.POINTER(p), ctypes.POINTER(_Basis), ctypes.POINTER(_Basis)]
        L.lbx_destroy.argtypes = [p]
        L.lbx_dims.argtypes = [p, ctypes.POINTER(l), ctypes.POINTER(l), ctypes.POINTER(l)]
        L.lbx_set_stream.argtypes = [p, p]
        L.lbx_synchronize.argtypes = [p]
        L.lbx_boys.argtypes = [i, l, p, p, p]
        for name in ("lbx_eri4", "lbx_eri4_diag", "lbx_eri3", "lbx_eri2", "lbx_eri2_diag",
                     "lbx_eri4d1", "lbx_eri3d1", "lbx_eri2d1"):
            getattr(L, name).argtypes = [p, p]
        for name in ("lbx_eri4_batch", "lbx_eri3_batch", "lbx_eri2_batch", "lbx_eri4d1_batch",
                     "lbx_eri3d1_batch", "lbx_eri2d1_batch", "lbx_eri4_batch_static", "lbx_eri4soc_batch",
                     "lbx_eri3soc_batch", "lbx_eri2d2_batch"):
            getattr(L, name).argtypes = [p, l, p, p]
        L.lbx_jk.argtypes = [p, p, p, p, d, i, d, ctypes.POINTER(JKStats)]
        L.lbx_jk_grad.argtypes = [p, p, p, d, i, d, ctypes.POINTER(JKStats)]
        L.lbx_create_batch.argtypes = [ctypes.POINTER(p), i, ctypes.POINTER(_Basis),
                                       ctypes.POINTER(_Basis)]
        L.lbx_ri_sums.argtypes = [p, p, i, d, ctypes.POINTER(RIStats)]
        L.lbx_tune.argtypes = [p, p, d]
        L.lbx_set_variant.argtypes = [p, i, i, i, i, i]
        L.lbx_set_class_filter.argtypes = [p, i, i, i, i]
        L.lbx_set_bound_window.argtypes = [p, d, d]
        L.lbx_ri_jk.argtypes = [p, p, p, p]
        L.lbx_one_electron.argtypes = [p, p, p, i, p, p]
        L.lbx_one_electron_grad.argtypes = [p, p, p, i, p, p, p]
        L.lbx_set_graphs.argtypes = [p, i]
        L.lbx_profile_enable.argtypes = [p, i]
        L.lbx_profile_read.argtypes = [p, ctypes.POINTER(ClassProfile), i, ctypes.POINTER(i)]
        _LIB = L
    return _LIB


def _check(rc: int):
    if rc != 0:
        raise LbxError(f"lbx error {rc}: {lib().lbx_last_error().decode()}")


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise LbxError("no GPU available: the lbx library runs on the MI355X only")
    return torch


def boys(n: int, x):
    """F_0..F_n at the points of the CUDA tensor x (float64) -> tensor (len(x), n+1)."""
    torch = _torch()
    x = x.contiguous().to(torch.float64)
    out = torch.empty((x.numel(), n + 1), dtype=torch.float64, device=x.device)
    _check(lib().lbx_boys(n, x.numel(), x.data_ptr(), out.data_ptr(),
                          torch.cuda.current_stream().cuda_stream))
    return out


class Context:
    """A molecule (lible_amd.basis.Molecule) resident on the current device."""

    def __init__(self, mol):
        _torch()
        self._keep = []
        main = self._basis(mol.coords, mol.shells, mol.natoms)
        aux = self._basis(mol.coords, mol.aux, mol.natoms) if mol.aux is not None else None
        self._ls = (list(mol.shells.l), list(mol.aux.l) if mol.aux is not None else None)
        h = ctypes.c_void_p()
        _check(lib().lbx_create(ctypes.byref(h), ctypes.byref(main),
                                ctypes.byref(aux) if aux is not None else None))
        self._attach(h, 1)

    @classmethod
    def batch(cls, mols):
        """One context over many independent molecules (lbx_create_batch); each needs an
        auxiliary basis.  Only ri_sums() is available on it."""
        _torch()
        self = cls.__new__(cls)
        self._keep = []
        self._ls = (None, None)
        n = len(mols)
        mains, auxs = (_Basis * n)(), (_Basis * n)()
        for k, m in enumerate(mols):
            if m.aux is None:
                raise LbxError("every molecule of a batch needs an auxiliary basis")
            mains[k] = self._basis(m.coords, m.shells, m.natoms)
            auxs[k] = self._basis(m.coords, m.aux, m.natoms)
        h = ctypes.c_void_p()
        _check(lib().lbx_create_batch(ctypes.byref(h), n, mains, auxs))
        self._attach(h, n)
        return self

    def _attach(self, h, nmol: int):
        torch = _torch()
        self.h = h
        self.nmol = nmol
        _check(lib().lbx_set_stream(self.h, torch.cuda.current_stream().cuda_stream))
        nbf, nbf_aux, nat = ctypes.c_long(), ctypes.c_long(), ctypes.c_long()
        _check(lib().lbx_dims(self.h, ctypes.byref(nbf), ctypes.byref(nbf_aux), ctypes.byref(nat)))
        self.nbf, self.nbf_aux, self.natoms = nbf.value, nbf_aux.value, nat.value
        self.device = torch.device("cuda", torch.cuda.current_device())

    def ri_sums(self, out=None, precision: int = LBX_FP64, mixed_thresh: float = 1e-5):
        """Per molecule: sum |eri3|, sum eri3, sum |eri2|, sum eri2 over the dense tensors
        (lbx_ri_sums) -> ((nmol, 4) tensor, stats); LBX_MIXED evaluates the triplets and
        doublets with Schwarz bound below mixed_thresh in FP32."""
        o = self._out((self.nmol, 4)) if out is None else out
        st = RIStats()
        _check(lib().lbx_ri_sums(self.h, o.data_ptr(), int(precision), float(mixed_thresh),
                                 ctypes.byref(st)))
        return o, {"triplets": st.triplets, "doublets": st.doublets, "flops": st.flops,
                   "triplets_fp32": st.triplets_fp32, "doublets_fp32": st.doublets_fp32}

    def _basis(self, coords, sh, natoms) -> _Basis:
        arrs = [np.ascontiguousarray(coords, np.float64), np.ascontiguousarray(sh.atom, np.int32),
                np.ascontiguousarray(sh.l, np.int32), np.ascontiguousarray(sh.nprim, np.int32),
                np.ascontiguousarray(sh.exps, np.float64),
                np.ascontiguousarray(sh.coeffs, np.float64)]
        self._keep.extend(arrs)
        ptr = [a.ctypes.data for a in arrs]
        return _Basis(natoms, ptr[0], sh.nsh, ptr[1], ptr[2], ptr[3], ptr[4], ptr[5])

    def close(self):
        if getattr(self, "h", None):
            lib().lbx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _out(self, shape):
        torch = _torch()
        return torch.empty(shape, dtype=torch.float64, device=self.device)

    def eri4(self, out=None):
        n = self.nbf
        out = self._out((n, n, n, n)) if out is None else out
        _check(lib().lbx_eri4(self.h, out.data_ptr()))
        return out

    def eri4_diag(self):
        out = self._out((self.nbf, self.nbf))
        _check(lib().lbx_eri4_diag(self.h, out.data_ptr()))
        return out

    def eri3(self):
        out = self._out((self.nbf, self.nbf, self.nbf_aux))
        _check(lib().lbx_eri3(self.h, out.data_ptr()))
        return out

    def eri2(self):
        out = self._out((self.nbf_aux, self.nbf_aux))
        _check(lib().lbx_eri2(self.h, out.data_ptr()))
        return out

    def eri2_diag(self):
        out = self._out((self.nbf_aux,))
        _check(lib().lbx_eri2_diag(self.h, out.data_ptr()))
        return out

    def eri4d1(self):
        """(12, nbf, nbf, nbf, nbf): derivatives of (mn|kl) with respect to the centres of the
        shells of m, n, k, l, x y z each (lbx_eri4d1)."""
        n = self.nbf
        out = self._out((12, n, n, n, n))
        _check(lib().lbx_eri4d1(self.h, out.data_ptr()))
        return out

    def eri3d1(self):
        """(9, nbf, nbf, naux) (lbx_eri3d1)."""
        out = self._out((9, self.nbf, self.nbf, self.nbf_aux))
        _check(lib().lbx_eri3d1(self.h, out.data_ptr()))
        return out

    def eri2d1(self):
        """(6, naux, naux) (lbx_eri2d1)."""
        out = self._out((6, self.nbf_aux, self.nbf_aux))
        _check(lib().lbx_eri2d1(self.h, out.data_ptr()))
        return out

    def blocks(self, shells, deriv: bool = False, op: str | None = None):
        """Blocks of listed shell quartets (n, 4), triplets (n, 3: main, main, auxiliary) or
        doublets (n, 2: auxiliary) -> tensor (n, ncomp, S_a, S_b[, S_c[, S_d]]) with ncomp 1 for
        the integrals, 12 / 9 / 6 for their derivatives (lbx_eri*_batch); op "soc" (quartets,
        triplets: 3 components) or "d2" (doublets: 36 components) for the spin-orbit-type
        integrals and the second derivatives."""
        sh = np.ascontiguousarray(shells, dtype=np.int32)
        if sh.ndim != 2 or sh.shape[1] not in (2, 3, 4):
            raise LbxError("shells must be (n, 4), (n, 3) or (n, 2)")
        per = sh.shape[1]
        ls_main = np.asarray(self._ls[0]) if self._ls[0] is not None else None
        ls_aux = np.asarray(self._ls[1]) if self._ls[1] is not None else None
        dims = []
        for k in range(per):
            aux = per == 2 or (per == 3 and k == 2)
            ls = ls_aux if aux else ls_main
            if ls is None or sh.shape[0] == 0:
                dims.append(1)
            else:
                idx = int(sh[0, k])
                if not 0 <= idx < len(ls):
                    raise LbxError("shell index out of range")
                dims.append(2 * int(ls[idx]) + 1)
        base = {4: "lbx_eri4", 3: "lbx_eri3", 2: "lbx_eri2"}[per]
        if op == "soc" and per in (3, 4):
            ncomp, name = 3, base + "soc_batch"
        elif op == "d2" and per == 2:
            ncomp, name = 36, base + "d2_batch"
        elif op is None:
            ncomp = (12, 9, 6)[4 - per] if deriv else 1
            name = base + ("d1_batch" if deriv else "_batch")
        else:
            raise LbxError(f"no {op} blocks for {per}-index entries")
        out = self._out((sh.shape[0], ncomp, *dims))
        _check(getattr(lib(), name)(self.h, sh.shape[0], sh.ctypes.data, out.data_ptr()))
        return out

    def jk_grad(self, D, schwarz: float = 1e-12, grad=None, precision: int = LBX_FP64,
                mixed_thresh: float = 1e-5):
        """Nuclear gradient (natoms, 3) of E = 1/2 tr(DJ) - 1/4 tr(DK) at fixed D (lbx_jk_grad);
        LBX_MIXED evaluates the quartets with Schwarz bound below mixed_thresh in FP32."""
        torch = _torch()
        D = D.to(device=self.device, dtype=torch.float64).contiguous()
        if D.shape != (self.nbf, self.nbf):
            raise LbxError(f"density must be ({self.nbf}, {self.nbf})")
        g = self._out((self.natoms, 3)) if grad is None else grad
        st = JKStats()
        _check(lib().lbx_jk_grad(self.h, D.data_ptr(), g.data_ptr(), schwarz, int(precision),
                                 float(mixed_thresh), ctypes.byref(st)))
        return g, {"quartets": st.quartets, "flops": st.flops, "quartets_fp32": st.quartets_fp32}

    def jk(self, D, schwarz: float = 1e-12, precision: int = LBX_FP64, mixed_thresh: float = 1e-5,
           J=None, K=None):
        torch = _torch()
        D = D.to(device=self.device, dtype=torch.float64).contiguous()
        n = self.nbf
        if D.shape != (n, n):
            raise LbxError(f"density must be ({n}, {n})")
        J = self._out((n, n)) if J is None else J
        K = self._out((n, n)) if K is None else K
        st = JKStats()
        _check(lib().lbx_jk(self.h, D.data_ptr(), J.data_ptr(), K.data_ptr(), schwarz, precision,
                            mixed_thresh, ctypes.byref(st)))
        return J, K, {"quartets": st.quartets, "flops": st.flops, "quartets_fp32": st.quartets_fp32}

    def ri_jk(self, D, exchange: bool = True):
        """Density-fitted J and K (K None when exchange is False) over the context's
        auxiliary basis (lbx_ri_jk)."""
        torch = _torch()
        D = D.to(device=self.device, dtype=torch.float64).contiguous()
        n = self.nbf
        if D.shape != (n, n):
            raise LbxError(f"density must be ({n}, {n})")
        J = self._out((n, n))
        K = self._out((n, n)) if exchange else None
        _check(lib().lbx_ri_jk(self.h, D.data_ptr(), J.data_ptr(), K.data_ptr() if K is not None else None))
        return J, K

    def one_electron(self, overlap: bool = True, kinetic: bool = True, charges=None):
        """Overlap S, kinetic energy T and the attraction V to point charges ((k, 4) rows
        x, y, z, q in bohr; None: no V) over the main basis (lbx_one_electron) -> (S, T, V),
        None for what was not asked."""
        n = self.nbf
        S = self._out((n, n)) if overlap else None
        T = self._out((n, n)) if kinetic else None
        V, q = None, None
        if charges is not None:
            q = np.ascontiguousarray(charges, np.float64).reshape(-1, 4)
            V = self._out((n, n))
        if S is None and T is None and V is None:
            raise LbxError("one_electron: nothing to compute")
        _check(lib().lbx_one_electron(self.h, S.data_ptr() if S is not None else None,
                                      T.data_ptr() if T is not None else None,
                                      0 if q is None else q.shape[0],
                                      q.ctypes.data if q is not None and q.size else None,
                                      V.data_ptr() if V is not None else None))
        return S, T, V

    def one_electron_grad(self, D, W=None, charges=None):
        """Nuclear gradient of tr(D (T + V)) - tr(W S) (lbx_one_electron_grad) -> (grad
        (natoms, 3) over the basis-function centres, charge_grad (k, 3) over the charge
        positions, None without charges)."""
        torch = _torch()
        n = self.nbf
        D = D.to(device=self.device, dtype=torch.float64).contiguous()
        if D.shape != (n, n):
            raise LbxError(f"density must be ({n}, {n})")
        if W is not None:
            W = W.to(device=self.device, dtype=torch.float64).contiguous()
            if W.shape != (n, n):
                raise LbxError(f"energy-weighted density must be ({n}, {n})")
        q = None
        if charges is not None:
            q = np.ascontiguousarray(charges, np.float64).reshape(-1, 4)
        nq = 0 if q is None else q.shape[0]
        g = self._out((self.natoms, 3))
        cg = self._out((nq, 3)) if nq else None
        _check(lib().lbx_one_electron_grad(self.h, D.data_ptr(), W.data_ptr() if W is not None else None, nq,
                                           q.ctypes.data if nq else None, g.data_ptr(),
                                           cg.data_ptr() if cg is not None else None))
        return g, cg

    def tune(self, D=None, schwarz: float = 1e-12):
        """Times every kernel variant of every class pair on this density and keeps the
        fastest (lbx_tune); on a batched context the RI sums' kernels (D not used)."""
        torch = _torch()
        if D is None:
            _check(lib().lbx_tune(self.h, None, schwarz))
            return
        D = D.to(device=self.device, dtype=torch.float64).contiguous()
        _check(lib().lbx_tune(self.h, D.data_ptr(), schwarz))

    def set_class_filter(self, cls: tuple | None):
        """Restricts the following jk() calls to one class pair (la, lb, lc, ld); None: all."""
        _check(lib().lbx_set_class_filter(self.h, *(cls if cls is not None else (-1, -1, -1, -1))))

    def set_graphs(self, on: bool):
        """Capture repeated jk() / ri_sums() launch sequences into HIP graphs and replay them
        (lbx_set_graphs; off by default, keep off when several threads drive the GPU)."""
        _check(lib().lbx_set_graphs(self.h, int(bool(on))))

    def set_bound_window(self, lo: float = 0.0, hi: float = float("inf")):
        """Restricts the following jk() calls to quartets with lo <= Q_ab Q_cd < hi."""
        _check(lib().lbx_set_bound_window(self.h, float(lo), float(hi)))

    def set_variant(self, cls: tuple, variant: int) -> bool:
        """Fix the kernel variant of a class pair; False if it has no such variant."""
        return lib().lbx_set_variant(self.h, *cls, variant) == 0

    def profile(self, on: bool):
        _check(lib().lbx_profile_enable(self.h, int(on)))

    def profile_read(self) -> list[dict]:
        n = ctypes.c_int()
        _check(lib().lbx_profile_read(self.h, None, 0, ctypes.byref(n)))
        buf = (ClassProfile * max(n.value, 1))()
        _check(lib().lbx_profile_read(self.h, buf, n.value, ctypes.byref(n)))
        return [{"class": (r.la, r.lb, r.lc, r.ld), "kernel": VARIANTS[r.kernel],
                 "calls": r.calls, "ms": r.ms, "quartets": r.quartets, "flops": r.flops}
                for r in buf[:n.value]]
