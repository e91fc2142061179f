"""Full-space numpy brute force for tiny systems -- TEST INFRASTRUCTURE ONLY.

Independent of pomerol's sector machinery and of its TermList merging: the Hamiltonian is
diagonalised in the whole Fock space (``numpy.linalg.eigh``) and the Lehmann sums of
``doc/gamma4.tex:18-70`` are evaluated term by term.  Used to pin the C++ oracle
(``ref_cpu.cpp``) on systems of up to ~4 sites.

Convention: Fock state bit ``i`` is mode ``i``; ``c_i``/``c_i^dag`` carry the sign
``(-1)^(occupied modes j < i)``.
"""
from __future__ import annotations

import itertools

import numpy as np


def op_matrix(n_modes: int, mono) -> np.ndarray:
    """Dense matrix of a monomial ``[(dag, mode), ...]`` (rightmost acts first)."""
    dim = 1 << n_modes
    M = np.zeros((dim, dim))
    for s in range(dim):
        t, sign = s, 1
        ok = True
        for dag, i in reversed(mono):
            bit = 1 << i
            occ = bool(t & bit)
            if occ == dag:
                ok = False
                break
            if bin(t & (bit - 1)).count("1") & 1:
                sign = -sign
            t ^= bit
        if ok:
            M[t, s] += sign
    return M


def hamiltonian(n_modes: int, terms) -> np.ndarray:
    return sum(c * op_matrix(n_modes, m) for c, m in terms)


class FullED:
    def __init__(self, n_modes: int, terms, beta: float):
        self.n = n_modes
        self.beta = beta
        self.H = hamiltonian(n_modes, terms)
        self.E, self.U = np.linalg.eigh(self.H)
        self.E0 = self.E.min()
        w = np.exp(-beta * (self.E - self.E0))
        self.w = w / w.sum()

    def rot(self, mono) -> np.ndarray:
        return self.U.T @ op_matrix(self.n, mono) @ self.U

    def gf(self, i: int, j: int, z: np.ndarray) -> np.ndarray:
        """G_ij(z) = sum_{mn} <n|c_i|m><m|c_j^dag|n>(w_m+w_n)/(z-(E_m-E_n))  (gamma4.tex:22-26)."""
        A = self.rot([(False, i)])
        B = self.rot([(True, j)])
        R = A * B.T * (self.w[None, :] + self.w[:, None])  # R[n, m]
        P = self.E[None, :] - self.E[:, None]
        mask = np.abs(R) > 0
        return np.array([np.sum(R[mask] / (zz - P[mask])) for zz in z])

    def chi(self, A_mono, B_mono, z: np.ndarray) -> np.ndarray:
        """<A;B>(z) = -sum R/(z-P), R = A_nm B_mn (w_n - w_m), plus beta*sum A_nm B_mn w_n at P=0, z=0."""
        A = self.rot(A_mono)
        B = self.rot(B_mono)
        P = self.E[None, :] - self.E[:, None]
        AB = A * B.T
        zero = np.abs(P) < 1e-8
        out = []
        for zz in z:
            R = AB * (self.w[:, None] - self.w[None, :])
            v = -np.sum(R[~zero] / (zz - P[~zero]))
            if abs(zz) < 1e-14:
                v += self.beta * np.sum((AB * self.w[:, None])[zero])
            out.append(v)
        return np.array(out)

    def g2(self, i: int, j: int, k: int, l: int, triplets: np.ndarray, res_tol: float = 1e-8, mask=None) -> np.ndarray:
        """chi_{ijkl}(w1,w2;w3) of gamma4.tex:38-70 by direct quadruple sums (over the eigenstate
        quadruples where mask, if given, is true)."""
        ops = [self.rot([(False, i)]), self.rot([(False, j)]), self.rot([(True, k)])]
        CX4 = self.rot([(True, l)])
        E, w, beta = self.E, self.w, self.beta
        Ei, Ej, Ek, El = np.ix_(E, E, E, E)
        wi, wj, wk, wl = np.ix_(w, w, w, w)
        out = np.zeros(len(triplets), dtype=complex)
        for perm in itertools.permutations(range(3)):
            sign = np.linalg.det(np.eye(3)[list(perm)])
            O1, O2, O3 = (ops[p] for p in perm)
            M = np.einsum("ij,jk,kl,li->ijkl", O1, O2, O3, CX4) * sign
            if mask is not None:
                M = M * mask
            nz = M != 0
            for t, (z1, z2, z3) in enumerate(triplets):
                Z = [z1, z2, -z3]
                a, b, c = (Z[p] for p in perm)
                d1 = a + Ei - Ej
                d3 = c + Ek - El
                val = (wi + wl) / (d1 * (a + b + c + Ei - El) * d3) - (wj + wk) / (d1 * (b + Ej - Ek) * d3)
                D12 = a + b + Ei - Ek
                res12 = np.abs(D12) <= res_tol
                r12 = np.where(res12, beta * wi, (wk - wi) / np.where(res12, 1.0, D12))
                D23 = b + c + Ej - El
                res23 = np.abs(D23) <= res_tol
                r23 = np.where(res23, beta * wj, (wl - wj) / np.where(res23, 1.0, D23))
                val = val + (r12 - r23) / (d1 * d3)
                out[t] += np.sum(M[nz] * np.broadcast_to(val, M.shape)[nz])
        return out
