"""Restricted Hartree-Fock around the integral engine: the caller step of the J/K path.

The reference library stops at the integrals; an SCF program turns them into an energy and a
nuclear gradient.  This module is that caller, on the device end to end: one-electron integrals
(lbx_one_electron), J and K of the current density (lbx_jk, or lbx_ri_jk with density fitting),
Fock assembly, orthogonalisation and diagonalisation (torch.linalg on the GPU, FP64), Pulay DIIS,
and the analytic gradient

    dE/dR = sum_mn D_mn d(T + V)_mn/dR - sum_mn W_mn dS_mn/dR + dE_2e/dR + dE_nn/dR

with D = 2 C_occ C_occ^T, W = 2 C_occ eps_occ C_occ^T, the one-electron part from
lbx_one_electron_grad (nuclei as point charges, orbital + operator parts) and the two-electron
part E_2e = 1/2 tr(DJ) - 1/4 tr(DK) at fixed D from lbx_jk_grad.
"""
from __future__ import annotations

import dataclasses

import numpy as np

from . import lbx as L


def nuclear_repulsion(coords: np.ndarray, Z) -> float:
    """sum_{A<B} Z_A Z_B / |R_A - R_B| (bohr)."""
    Z = np.asarray(Z, np.float64)
    d = coords[:, None, :] - coords[None, :, :]
    r = np.sqrt((d * d).sum(-1))
    iu = np.triu_indices(len(Z), 1)
    return float((Z[iu[0]] * Z[iu[1]] / r[iu]).sum())


def nuclear_repulsion_grad(coords: np.ndarray, Z) -> np.ndarray:
    """d E_nn / d R_A = -sum_{B != A} Z_A Z_B (R_A - R_B) / |R_A - R_B|^3."""
    Z = np.asarray(Z, np.float64)
    d = coords[:, None, :] - coords[None, :, :]
    r = np.sqrt((d * d).sum(-1))
    np.fill_diagonal(r, np.inf)
    return -((Z[:, None] * Z[None, :] / r ** 3)[:, :, None] * d).sum(1)


@dataclasses.dataclass
class SCFResult:
    energy: float
    converged: bool
    iterations: int
    e_nuc: float
    mo_energy: object      # (nbf,) torch tensor
    mo_coeff: object       # (nbf, nmo) torch tensor
    density: object        # (nbf, nbf) total density D = 2 C_occ C_occ^T
    history: list


class RHF:
    """Closed-shell Hartree-Fock on one molecule.

    mol: lible_amd.basis.Molecule (coordinates in bohr; an auxiliary basis is needed only with
    density_fit); atomic_nrs: nuclear charges per atom; charge: molecular charge."""

    def __init__(self, mol, atomic_nrs, charge: int = 0, density_fit: bool = False,
                 schwarz: float = 1e-12, precision: int = L.LBX_FP64, mixed_thresh: float = 1e-5):
        import torch
        self.torch = torch
        self.mol = mol
        self.Z = np.asarray(atomic_nrs, np.float64)
        if len(self.Z) != mol.natoms:
            raise ValueError("one nuclear charge per atom")
        nelec = int(round(self.Z.sum())) - charge
        if nelec <= 0 or nelec % 2:
            raise ValueError(f"closed-shell RHF needs a positive even electron count, got {nelec}")
        self.nocc = nelec // 2
        self.density_fit = density_fit
        if density_fit and mol.aux is None:
            raise ValueError("density fitting needs an auxiliary basis")
        self.schwarz, self.precision, self.mixed_thresh = schwarz, precision, mixed_thresh
        self.ctx = L.Context(mol)
        self.ctx.set_graphs(True)   # the J/K buffers are reused every iteration: replays
        self.nuclei = np.column_stack([mol.coords, self.Z])
        S, T, V = self.ctx.one_electron(charges=self.nuclei)
        self.S, self.H = S, T + V
        self.e_nuc = nuclear_repulsion(mol.coords, self.Z)
        s, U = torch.linalg.eigh(S)
        keep = s > 1e-9 * s.max()
        self.X = U[:, keep] / torch.sqrt(s[keep])          # canonical orthogonalisation
        self.result: SCFResult | None = None

    def jk(self, D):
        if self.density_fit:
            return self.ctx.ri_jk(D)
        # the same D, J and K buffers every iteration: lbx_jk captures its launch sequence into a
        # HIP graph on the first call and replays it on the later ones
        if not hasattr(self, "_bufs"):
            self._bufs = tuple(self.torch.empty_like(D) for _ in range(3))
        Db, Jb, Kb = self._bufs
        Db.copy_(D)
        self.ctx.jk(Db, schwarz=self.schwarz, precision=self.precision, mixed_thresh=self.mixed_thresh,
                    J=Jb, K=Kb)
        return Jb.clone(), Kb.clone()

    def _density(self, F):
        torch = self.torch
        e, Co = torch.linalg.eigh(self.X.T @ F @ self.X)
        C = self.X @ Co
        Cocc = C[:, :self.nocc]
        return e, C, 2.0 * Cocc @ Cocc.T

    def kernel(self, max_iter: int = 100, e_tol: float = 1e-10, g_tol: float = 1e-8,
               diis_space: int = 8) -> SCFResult:
        """SCF iterations from the core-Hamiltonian guess with Pulay DIIS on the orthogonalised
        commutator FDS - SDF; converged when the energy change is below e_tol and the largest
        commutator element below g_tol."""
        torch = self.torch
        H, S, X = self.H, self.S, self.X
        e, C, D = self._density(H)
        focks, errs, hist = [], [], []
        E_old, converged, it = 0.0, False, 0
        for it in range(1, max_iter + 1):
            J, K = self.jk(D)
            F = H + J - 0.5 * K
            E = 0.5 * float((D * (H + F)).sum()) + self.e_nuc
            err = X.T @ (F @ D @ S - S @ D @ F) @ X
            gnorm = float(err.abs().max())
            hist.append((E, gnorm))
            if abs(E - E_old) < e_tol and gnorm < g_tol:
                converged = True
                break
            E_old = E
            focks.append(F)
            errs.append(err)
            if len(focks) > diis_space:
                focks.pop(0)
                errs.pop(0)
            if len(focks) >= 2:
                m = len(focks)
                B = torch.empty((m + 1, m + 1), dtype=torch.float64, device=F.device)
                for i in range(m):
                    for j in range(i + 1):
                        B[i, j] = B[j, i] = (errs[i] * errs[j]).sum()
                B[m, :m] = B[:m, m] = -1.0
                B[m, m] = 0.0
                rhs = torch.zeros(m + 1, dtype=torch.float64, device=F.device)
                rhs[m] = -1.0
                c = torch.linalg.lstsq(B.cpu(), rhs.cpu().unsqueeze(1)).solution.squeeze(1).to(F.device)
                F = sum(c[i] * focks[i] for i in range(m))
            e, C, D = self._density(F)
        # orbitals of the final Fock matrix (D was built from them before the last J/K)
        self.result = SCFResult(E, converged, it, self.e_nuc, e, C, D, hist)
        self._F = F
        return self.result

    def energy_weighted_density(self):
        """W = 2 sum_occ eps_i C_i C_i^T of the converged orbitals."""
        eF, C, _ = self._density(self._F)
        Cocc = C[:, :self.nocc]
        return 2.0 * (Cocc * eF[:self.nocc]) @ Cocc.T

    def gradient(self) -> np.ndarray:
        """Analytic nuclear gradient (natoms, 3) in hartree/bohr at the converged density (exact
        J/K only: the density-fitted energy has extra metric terms)."""
        if self.result is None:
            raise RuntimeError("run kernel() first")
        if self.density_fit:
            raise NotImplementedError("analytic gradient with density fitting")
        D = self.result.density
        W = self.energy_weighted_density()
        g1, gq = self.ctx.one_electron_grad(D, W, self.nuclei)
        g2, _ = self.ctx.jk_grad(D, schwarz=self.schwarz, precision=self.precision,
                                 mixed_thresh=self.mixed_thresh)
        g = (g1 + gq + g2).cpu().numpy()
        return g + nuclear_repulsion_grad(self.mol.coords, self.Z)

    def close(self):
        self.ctx.close()
