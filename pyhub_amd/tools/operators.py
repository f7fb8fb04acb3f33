"""Operator constructors with PyHub's names (pyhub/tools/operators.py).

Each class is a one-factor Operators whose factor string selects the elementary
routine of the device program (the class names are the API).  `sisj` composes
S_i . S_j from `c_dagger_c`, `n`, `_n` and `sz`; `opeexp` sums exp(op) psi.
"""
from __future__ import annotations

import numpy as np

from pyhub_amd.core._operators import Operators


class c_dagger_c(Operators):
    def __init__(self, *args):
        Operators.__init__(self, *args)
        self.str_index = [["c_dagger_c"]]


class c_dagger(Operators):
    def __init__(self, *args):
        Operators.__init__(self, *args)
        self.str_index = [["c_dagger"]]


class _c(Operators):
    def __init__(self, *args):
        Operators.__init__(self, *args)
        self.str_index = [["c"]]


class n(Operators):
    def __init__(self, *args):
        Operators.__init__(self, *args)
        self.str_index = [["n"]]


class _n(Operators):
    def __init__(self, *args):
        Operators.__init__(self, *args)
        self.str_index = [["1_n"]]


class sz(Operators):
    def __init__(self, *args):
        Operators.__init__(self, *args)
        self.str_index = [["sz"]]


class splus(Operators):
    def __init__(self, *args):
        Operators.__init__(self, *args)
        self.str_index = [["splus"]]


class sminus(Operators):
    def __init__(self, *args):
        Operators.__init__(self, *args)
        self.str_index = [["sminus"]]


class idt(Operators):
    def __init__(self):
        Operators.__init__(self)
        self.coeff = [1.0]
        self.elem_index = [[()]]
        self.str_index = [["id."]]


class empty_operator(Operators):
    def __init__(self):
        Operators.__init__(self)
        self.coeff = []
        self.elem_index = []
        self.str_index = []


def sisj(*args):
    """The S_i . S_j operator of PyHub's helper (tools/operators.py:51-60), as a sum
    of products in this order: the transverse part, then S^z_i S^z_j.

    The transverse part (S+_i S-_j + S-_i S+_j) / 2 is written with the hopping
    pairs that move one spin of each species between the two sites; bringing
    c^dag_iu c_iu-bar c^dag_ju-bar c_ju into that order flips its sign, so each of
    the two exchange products enters with -1/2.  On one site the helper keeps its
    own form n_up (1 - n_dn), as two halves (it is the reference's operator, so
    <sisj(i, i)> is what PyHub reports)."""
    i, j = args[0][0], args[1][0]
    if i == j:
        half = n((i, "up")) * _n((i, "down")) / 2
        parts = [half, half]
    else:
        exchange = (c_dagger_c((i, a), (j, a)) * c_dagger_c((j, b), (i, b))
                    for a, b in (("u", "d"), ("d", "u")))
        parts = [-0.5 * term for term in exchange]
    return opesum(parts + [sz((i,)) * sz((j,))])


def _taylor_terms(op: Operators, v, count):
    """Successive terms op^k v / k! for k = 1 .. count, each from the previous."""
    term = v
    for k in range(1, count + 1):
        term = op @ term / k
        yield term


def opeexp(op: Operators, psi, ftol=1.0e-6, unitary=False):
    """exp(op) psi by its Taylor series (PyHub's helper, tools/operators.py:96-125).

    The partial sums stop when their squared norm settles: it moves by less than
    ftol from one term to the next (the first comparison is against 1e3), or, with
    unitary=True, it is within ftol of the norm of psi (the helper compares the norm
    of psi with the squared norm of the sum).  At most 99 terms are summed, and
    reaching the 99th raises ValueError, as in the helper.  When psi is not a list or
    array (e.g. None): the matrix whose row i is exp(op) e_i over the selected basis
    elements, each row summed with the default ftol and unitary."""
    if not isinstance(psi, (list, np.ndarray)):
        rows = [opeexp(op, e) for e in np.eye(op.nb_selected)]
        return np.array(rows, dtype=float).reshape(op.nb_selected, op.nb_selected)
    total = np.copy(psi)
    ref = np.linalg.norm(psi) if unitary else None
    last = 1.0e3
    max_terms = 99
    for k, term in enumerate(_taylor_terms(op, np.copy(psi), max_terms), start=1):
        total = total + term
        sq = total.conj() @ total
        settled = abs(ref - sq) < ftol if unitary else abs(last - sq) < ftol
        last = sq
        if settled and k < max_terms:
            return total
    raise ValueError("exponential operator does not converge")


def opesum(lst):
    H = empty_operator()
    for op in lst:
        H += op
    return H
