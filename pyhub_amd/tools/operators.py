"""Operator constructors with PyHub's names (mirror of pyhub/tools/operators.py).

Each class is a one-factor Operators whose factor string selects the elementary
routine of the device program; `sisj` is built from `c_dagger_c`, `n`, `_n` and
`sz` as in the reference (operators.py:36-48).
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
    """S_i . S_j as a sum of products (operators.py:36-48): on one site
    n_up (1 - n_dn) / 2 twice, else minus half the two spin-exchange products; then
    sz * sz in both cases."""
    op = empty_operator()
    i, j = args[0][0], args[1][0]
    if i == j:
        op += n((i, "up")) * _n((i, "down")) / 2 + n((i, "up")) * _n((i, "down")) / 2
    else:
        op -= (c_dagger_c((i, "u"), (j, "u")) * c_dagger_c((j, "d"), (i, "d"))
               + c_dagger_c((i, "d"), (j, "d")) * c_dagger_c((j, "u"), (i, "u"))) / 2
    op += sz((i,)) * sz((j,))
    return op


def opeexp(op: Operators, psi, ftol=1.0e-6, unitary=False):
    """exp(op) psi by its series (operators.py:96-125); without psi, the matrix
    whose row i is exp(op) e_i."""
    if isinstance(psi, (list, np.ndarray)):
        conv = False
        new_psi, full_psi, i, val = np.copy(psi), np.copy(psi), 1, 1.0e3
        nmr0 = np.linalg.norm(psi)
        while not conv and i < 100:
            new_psi = op @ new_psi / i
            full_psi = full_psi + new_psi
            i += 1
            nmr = full_psi.conj() @ full_psi
            if unitary:
                if np.abs(nmr0 - nmr) < ftol:
                    conv = True
            else:
                if np.abs(val - nmr) < ftol:
                    conv = True
                else:
                    val = nmr
        if i == 100:
            raise ValueError("exponential operator does not converge")
        return full_psi
    matrix = np.zeros((op.nb_selected, op.nb_selected))
    psi0 = np.zeros(op.nb_selected)
    for i in range(op.nb_selected):
        psi0[i] = 1.0
        matrix[i, :] = opeexp(op, psi0)
        psi0[i] = 0.0
    return matrix


def opesum(lst):
    H = empty_operator()
    for op in lst:
        H += op
    return H
