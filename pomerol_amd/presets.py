"""Second-quantised model expressions (fermionic subset of pomerol's operator DSL).

Mirrors ``include/pomerol/Operators.hpp:52-59`` (``c``, ``c_dag``, ``n``, ``hc``) and the
fermionic ``LatticePresets`` (``src/pomerol/LatticePresets.cpp:46-126,140-147,402-409``):
``Level``, ``Hopping``, ``Magnetization``, ``NupNdown`` and ``CoulombS``.  An expression is a
list of ``(coefficient, monomial)`` pairs; a monomial is a tuple of ``(dagger, index)`` with
``index = (label, orbital, spin)`` and ``spin`` 0 (down) or 1 (up), the numbering of
``LatticePresets.hpp:31-35``.  The rightmost operator of a monomial acts first.

Index tuples are ordered like ``IndexClassification`` orders them (``IndexClassification.hpp:
133-140``): lexicographically on (label, orbital, spin).
"""
from __future__ import annotations

from typing import Iterable

DOWN = 0
UP = 1


class Expr:
    """Sum of fermionic monomials with real coefficients."""

    def __init__(self, terms: Iterable | None = None):
        self.terms = []
        for coeff, mono in terms or []:
            self.terms.append((float(coeff), tuple(mono)))

    # arithmetic -----------------------------------------------------------
    def __add__(self, other):
        if isinstance(other, (int, float)):
            other = Expr([(other, ())]) if other else Expr()
        return Expr(self.terms + other.terms)

    __radd__ = __add__

    def __iadd__(self, other):
        return self.__add__(other)

    def __neg__(self):
        return Expr([(-c, m) for c, m in self.terms])

    def __sub__(self, other):
        return self + (-other)

    def __mul__(self, other):
        if isinstance(other, (int, float)):
            return Expr([(c * other, m) for c, m in self.terms])
        return Expr([(c1 * c2, m1 + m2) for c1, m1 in self.terms for c2, m2 in other.terms])

    def __rmul__(self, other):
        if isinstance(other, (int, float)):
            return Expr([(other * c, m) for c, m in self.terms])
        return NotImplemented

    def hc(self) -> "Expr":
        """Hermitian conjugate (real coefficients)."""
        return Expr([(c, tuple((not d, i) for d, i in reversed(m))) for c, m in self.terms])

    # queries --------------------------------------------------------------
    def indices(self) -> list:
        return sorted({i for _, m in self.terms for _, i in m})

    def mode_terms(self, indices: list | None = None) -> list:
        """Terms with integer mode numbers, ``[(coeff, [(dag, mode), ...]), ...]``."""
        idx = {t: k for k, t in enumerate(indices if indices is not None else self.indices())}
        return [(c, [(bool(d), idx[i]) for d, i in m]) for c, m in self.terms if c != 0.0]


def c(label: str, orbital: int = 0, spin: int = UP) -> Expr:
    return Expr([(1.0, ((False, (label, orbital, spin)),))])


def c_dag(label: str, orbital: int = 0, spin: int = UP) -> Expr:
    return Expr([(1.0, ((True, (label, orbital, spin)),))])


def n(label: str, orbital: int = 0, spin: int = UP) -> Expr:
    return c_dag(label, orbital, spin) * c(label, orbital, spin)


def Level(label: str, eps: float, orbital: int = 0, spin: int | None = None) -> Expr:
    if spin is None:
        return Level(label, eps, orbital, UP) + Level(label, eps, orbital, DOWN)
    return eps * n(label, orbital, spin)


def Hopping(label1: str, label2: str, t: float, orbital: int = 0, spin: int | None = None) -> Expr:
    if spin is None:
        return Hopping(label1, label2, t, orbital, UP) + Hopping(label1, label2, t, orbital, DOWN)
    e = t * c_dag(label1, orbital, spin) * c(label2, orbital, spin)
    return e + e.hc()


def Magnetization(label: str, h: float, orbital: int = 0) -> Expr:
    return Level(label, h, orbital, UP) + Level(label, -h, orbital, DOWN)


def NupNdown(label: str, U: float, orbital: int = 0) -> Expr:
    return U * n(label, orbital, UP) * n(label, orbital, DOWN)


def CoulombS(label: str, U: float, eps: float, orbital: int = 0) -> Expr:
    return Level(label, eps, orbital, UP) + Level(label, eps, orbital, DOWN) + NupNdown(label, U, orbital)


def anderson(eps_bath, V_bath, U: float, mu: float, h: float = 0.0, imp: str = "C", bath_prefix: str = "b") -> Expr:
    """Single-orbital Anderson impurity: CoulombS(imp, U, -mu) + sum_k Level(b_k) + Hopping(imp, b_k)."""
    H = CoulombS(imp, U, -mu)
    if h:
        H += Magnetization(imp, h)
    for k, (e, v) in enumerate(zip(eps_bath, V_bath)):
        name = f"{bath_prefix}{k}"
        H += Level(name, e)
        H += Hopping(imp, name, v)
    return H


def hubbard_cluster(Lx: int, Ly: int, t: float, tp: float, U: float, mu: float) -> Expr:
    """Periodic Lx x Ly Hubbard cluster with nearest (-t) and diagonal (-t') hopping."""
    def name(x, y):
        return f"s{y * Lx + x:02d}"

    H = Expr()
    for y in range(Ly):
        for x in range(Lx):
            H += CoulombS(name(x, y), U, -mu)
    bonds = {}
    for y in range(Ly):
        for x in range(Lx):
            a = name(x, y)
            for dx, dy, amp in ((1, 0, -t), (0, 1, -t), (1, 1, -tp), (1, -1, -tp)):
                b = name((x + dx) % Lx, (y + dy) % Ly)
                if a == b or amp == 0.0:
                    continue
                key = (min(a, b), max(a, b))
                bonds[key] = bonds.get(key, 0.0) + amp
    for (a, b), amp in sorted(bonds.items()):
        H += Hopping(a, b, amp)
    return H
