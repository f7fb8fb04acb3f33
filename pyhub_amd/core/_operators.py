"""Generic operator strings with PyHub's Operators API (mirror of pyhub/core/_operators.py).

The reference keeps an operator as a sum of products (`coeff`, `elem_index`,
`str_index`), writes it to operators.h5 and runs the Fortran operators.x for every
application (`_write_operator` :396-471, `exec_operators` :208-219).  Here
`set_basis` compiles the sum into a device program (phub_opprog_create) on a device
product basis of the listed sectors (phub_fock_create), and every application is one
pull-form kernel over the basis (phub_operator_apply / _diag / _matrix).  The
factors of a product act in list order, the first one first, as the Fortran driver
applies them (compute_operators.f90:33-67).

Kept from the reference on purpose (drop-in behaviour):
- `sz` on a vector is SZ_OPE's (1 - n_up)/2 - (1 - n_dn)/2; on a basis element
  (to_matrix, the comparisons, bit_repr) it is the proper Sz (operators_basis.f90);
- `c_dagger_c` picks the spin-flip routine when the two spin labels differ as given
  ('u' and 'up' count as different), as `_write_operator` does (:450-451);
- the comparisons list the basis elements up-index-major (:569-575);
- a complex vector is applied to its real and imaginary parts (:148-152).
Not kept: the HDF5 file, the debug print of a restricted vector (:156) and the
`beta` dataset written with alpha's values (:361).
"""
from __future__ import annotations

import ctypes as C
import re
from itertools import product

import numpy as np
import torch

from pyhub_amd.core.basis import Basis

_CODES = {"n": 2, "id.": 3, "1_n": 4, "sz": 5, "splus": 6, "sminus": 7, "sisj": 8,
          "c_dagger": 10, "c": 11}
_UP_LABELS = (1, "u", "up", "+")


def _spin_code(label):
    return 1 if label in _UP_LABELS else 2


def _factor_code(string, args):
    """(code, site1, spin1, site2, spin2) of one factor (_operators.py:426-469)."""
    site1 = spin1 = site2 = spin2 = 0
    try:
        site1 = int(args[0][0]) + 1
        site2 = int(args[1][0]) + 1
    except (IndexError, TypeError):
        pass
    try:
        spin1 = _spin_code(args[0][1])
        spin2 = _spin_code(args[1][1])
    except (IndexError, TypeError):
        pass
    if string == "c_dagger_c":
        code = 1 if args[0][1] == args[1][1] else 9
    elif string in _CODES:
        code = _CODES[string]
    else:
        raise NotImplementedError(string)
    return code, site1, spin1, site2, spin2


class _Compiled:
    """Device basis + program shared by an operator and its restrictions."""

    def __init__(self, ctx, nb_sites, nup, ndown, coeff, factors):
        self.ctx = ctx
        lib = ctx.lib
        nu = (C.c_int * len(nup))(*[int(k) for k in nup])
        nd = (C.c_int * len(ndown))(*[int(k) for k in ndown])
        self.fock = C.c_void_p()
        ctx.check(lib.phub_fock_create(ctx.h, int(nb_sites), nu, len(nup), nd, len(ndown),
                                       C.byref(self.fock)))
        a, b = C.c_int64(0), C.c_int64(0)
        lib.phub_fock_dims(self.fock, C.byref(a), C.byref(b))
        self.nsup, self.nsdown = int(a.value), int(b.value)
        self.nstates = self.nsup * self.nsdown
        cf = np.ascontiguousarray(coeff, dtype=np.float64)
        lens = np.ascontiguousarray([len(f) for f in factors], dtype=np.int32)
        flat = [x for f in factors for fac in f for x in fac]
        el = np.ascontiguousarray(flat if flat else [0], dtype=np.int32)
        self.prog = C.c_void_p()
        ctx.check(lib.phub_opprog_create(ctx.h, len(cf), cf.ctypes.data_as(C.POINTER(C.c_double)),
                                         lens.ctypes.data_as(C.POINTER(C.c_int)),
                                         el.ctypes.data_as(C.POINTER(C.c_int)), C.byref(self.prog)))

    def __del__(self):
        try:
            lib = self.ctx.lib
            if getattr(self, "prog", None):
                lib.phub_opprog_destroy(self.prog)
            if getattr(self, "fock", None):
                lib.phub_fock_destroy(self.fock)
        except Exception:
            pass

    def apply(self, x: torch.Tensor) -> torch.Tensor:
        """O x for a device FP64 vector (or (k, nstates) block) of the full basis."""
        x = x.to(self.ctx.device, torch.float64).contiguous()
        out = torch.empty_like(x)
        nvec = 1 if x.dim() == 1 else x.shape[0]
        self.ctx.check(self.ctx.lib.phub_operator_apply(self.ctx.h, self.fock, self.prog,
                                                        C.c_void_p(x.data_ptr()),
                                                        C.c_void_p(out.data_ptr()), int(nvec)))
        return out

    def diag(self) -> np.ndarray:
        d = torch.empty(self.nstates, dtype=torch.float64, device=self.ctx.device)
        self.ctx.check(self.ctx.lib.phub_operator_diag(self.ctx.h, self.fock, self.prog,
                                                       C.c_void_p(d.data_ptr())))
        return d.cpu().numpy()

    def matrix(self) -> np.ndarray:
        m = torch.empty((self.nstates, self.nstates), dtype=torch.float64, device=self.ctx.device)
        self.ctx.check(self.ctx.lib.phub_operator_matrix(self.ctx.h, self.fock, self.prog,
                                                         C.c_void_p(m.data_ptr())))
        return m.cpu().numpy()


class Operators(Basis):
    _next_index = 1

    def __init__(self, *args):
        self.coeff = [1.0]
        self.elem_index = [[args]]
        self.op2write = True

    # -- algebra (_operators.py:17-116) -------------------------------------------
    def __add__(self, other):
        if isinstance(other, Operators):
            out = Operators.empty_operator()
            out.coeff = self.coeff + other.coeff
            out.elem_index = self.elem_index + other.elem_index
            out.str_index = self.str_index + other.str_index
            return out
        if isinstance(other, (float, int)):
            return self + other * Operators._idt()
        return NotImplemented

    def __radd__(self, other):
        return self.__add__(other)

    def __sub__(self, other):
        if isinstance(other, Operators):
            out = Operators.empty_operator()
            out.coeff = self.coeff + [-1.0 * c for c in other.coeff]
            out.elem_index = self.elem_index + other.elem_index
            out.str_index = self.str_index + other.str_index
            return out
        if isinstance(other, (float, int)):
            return self - other * Operators._idt()
        return NotImplemented

    def __rsub__(self, other):
        return -1 * self.__sub__(other)

    def _with_same_basis(self, out):
        if not self.op2write:
            out.set_basis(Basis(self.nb_sites, self.hilbert, self.order))
            out.selected = self.selected
            out.nb_selected = self.nb_selected
        return out

    def __mul__(self, other):
        out = Operators.empty_operator()
        if isinstance(other, Operators):
            out.coeff = [a * b for a, b in product(self.coeff, other.coeff)]
            out.elem_index = [[*a, *b] for a, b in product(self.elem_index, other.elem_index)]
            out.str_index = [[*a, *b] for a, b in product(self.str_index, other.str_index)]
            return out
        if isinstance(other, (float, int)):
            out.coeff = [other * c for c in self.coeff]
            out.elem_index = self.elem_index
            out.str_index = self.str_index
            return self._with_same_basis(out)
        return NotImplemented

    def __neg__(self):
        out = Operators.empty_operator()
        out.coeff = [-1.0 * c for c in self.coeff]
        out.elem_index = self.elem_index
        out.str_index = self.str_index
        return self._with_same_basis(out)

    def __rmul__(self, other):
        return self.__mul__(other)

    def __truediv__(self, other):
        if isinstance(other, (float, int)):
            return self.__mul__(1 / float(other))
        return NotImplemented

    def __matmul__(self, other):
        return self.__call__(other)

    def __rmatmul__(self, other):
        return self.__call__(np.conj(other))

    def __getitem__(self, index):
        """The operator restricted to the basis elements `index` (list or array)."""
        if isinstance(index, (list, np.ndarray)):
            index = np.asarray(index).astype(int)
            other = Operators.empty_operator()
            Basis.__init__(other, self.nb_sites, hilbert=self.hilbert, order=self.order)
            other.coeff = self.coeff
            other.elem_index = self.elem_index
            other.str_index = self.str_index
            other.op2write = False
            other.stable = True
            other._index = self._index
            other._prog = self._prog
            other.selected = index
            other.nb_selected = len(index)
            other.max_len_ope = self.max_len_ope
            return other
        raise TypeError("index must be a list or an array")

    # -- application (_operators.py:145-206) ----------------------------------------
    def _require_basis(self):
        if self.op2write:
            raise AttributeError("Set the basis set using self.set_basis(YourBasis:Basis)")

    def _full(self, psi):
        ns = self._prog.nstates
        x = torch.as_tensor(np.asarray(psi, dtype=np.float64)) if not isinstance(psi, torch.Tensor) \
            else psi.to(torch.float64)
        if self.selected is None:
            if x.shape[-1] != ns:
                raise ValueError(f"vector of {x.shape[-1]} elements for {ns} basis states")
            return x
        full = torch.zeros(ns, dtype=torch.float64)
        full[torch.as_tensor(self.selected)] = x.cpu()
        return full

    def __call__(self, psi, avg=False):
        self._require_basis()
        arr = psi.cpu().numpy() if isinstance(psi, torch.Tensor) else np.asarray(psi)
        if np.iscomplexobj(arr):
            if not avg:
                return self.__call__(arr.real, avg=False) + 1j * self.__call__(arr.imag, avg=False)
            return self.__call__(arr.real, avg=True) + self.__call__(arr.imag, avg=True)
        x = self._full(arr)
        y = self._prog.apply(x)
        if avg:
            return float(torch.dot(x.to(y.device), y))
        out = y.cpu().numpy()
        return out if self.selected is None else out[self.selected]

    def avg(self, psi):
        return self.__call__(psi, avg=True)

    def exec_operators(self):
        """Nothing to run: the program is resident on the device."""
        self._require_basis()

    # -- basis and compilation (_operators.py:250-256, 396-471) ---------------------
    def set_basis(self, B: Basis, selected: np.ndarray = None, memspace="10Go"):
        Basis.__init__(self, B.nb_sites, hilbert=B.hilbert, order=B.order)
        self.nb_selected = len(selected) if isinstance(selected, np.ndarray) else self.nstates
        self.selected = selected if isinstance(selected, np.ndarray) else None
        self._set_basis = True
        if self.op2write:
            self._write_operator(max_len_ope=40)
        else:
            # the program is compiled against a basis: follow the new one
            self._compile()

    def _write_operator(self, max_len_ope=40):
        self.op2write = False
        self.stable = True
        self.max_len_ope = max_len_ope
        self._index = Operators._next_index
        Operators._next_index += 1
        self._compile()

    def _terms(self, max_len_ope=None):
        """[(coeff, [(code, site1, spin1, site2, spin2), ...]), ...]: the encoding
        `_write_operator` writes for operators.x (:422-471), one entry per product."""
        out = []
        for c, term in zip(self.coeff, self):
            f = [_factor_code(s, e) for s, e in zip(term.str_index[0], term.elem_index[0])]
            if max_len_ope is not None and len(f) > max_len_ope:
                raise ValueError(f"a product of {len(f)} factors exceeds max_len_ope={max_len_ope}")
            out.append((float(c), f))
        return out

    def _compile(self):
        from pyhub_amd.engine import Context
        terms = self._terms(self.max_len_ope)
        for _, fs in terms:
            for code, s1, _, s2, _ in fs:
                if code != 3 and max(s1, s2 if code in (1, 8, 9) else 0) > self.nb_sites:
                    raise ValueError(f"operator site beyond the {self.nb_sites}-site basis")
        if not terms:   # the empty operator: one zero term
            terms = [(0.0, [(3, 0, 0, 0, 0)])]
        self._prog = _Compiled(Context.default(), self.nb_sites, list(self.nup), list(self.ndown),
                               [c for c, _ in terms], [f for _, f in terms])

    # -- matrices and per-element values ----------------------------------------------
    def _basis_values(self):
        """<e_i|O|e_i> for every basis element (the basis-element path)."""
        self._require_basis()
        return self._prog.diag()

    @property
    def to_matrix(self):
        """H[i, :] = O e_i over the basis (or its selected elements), as the
        reference builds it from the basis-element path (:296-323)."""
        self._require_basis()
        M = self._prog.matrix()
        if self.selected is not None:
            sel = np.asarray(self.selected, dtype=int)
            return M[np.ix_(sel, sel)]
        return M

    def _element_order(self):
        """State indices in the order the comparisons visit them (:557-575)."""
        if self.selected is not None:
            return np.asarray(self.selected, dtype=int)
        nsup, nsd = self.nsup, self.nsdown
        return np.array([d * nsup + u for u in range(nsup) for d in range(nsd)], dtype=int)

    def _compare(self, other, fn):
        if not isinstance(other, (float, int)):
            return NotImplemented
        vals = self._basis_values()[self._element_order()]
        return [fn(v) for v in vals]

    def __eq__(self, other):
        return self._compare(other, lambda v: bool(np.isclose(v, other)))

    def __ne__(self, other):
        eq = self.__eq__(other)
        return eq if eq is NotImplemented else [not e for e in eq]

    def __ge__(self, other):
        return self._compare(other, lambda v: bool(v >= other))

    def __gt__(self, other):
        return self._compare(other, lambda v: bool(v > other))

    def __le__(self, other):
        return self._compare(other, lambda v: bool(v <= other))

    def __lt__(self, other):
        return self._compare(other, lambda v: bool(v < other))

    __hash__ = object.__hash__

    def __req__(self, other):
        return self.__eq__(other)

    def bit_repr(self):
        """Every selected basis element with its occupation string and <e|O|e> (:474-507)."""
        self._require_basis()
        vals = self._basis_values()
        select = self.selected if self.selected is not None else list(range(self.nstates))
        bup, bdn = np.asarray(self.basis_up), np.asarray(self.basis_down)
        s = "-" * (3 * (self.nb_sites + 4)) + "\n"
        s += "empty        : 0\nuparrow      : 1\ndownarrow    :-1\nupdoawnarrow : 2\n"
        s += "-" * (3 * (self.nb_sites + 4)) + "\n"
        s += "index | bit" + " " * (3 * (self.nb_sites - 1) - 1) + " | <H> " + "\n"
        s += "------|----" + "-" * (3 * (self.nb_sites - 1) - 1) + "-|-----" + "\n"
        for index in self.basis[select]:
            up, down = self.decompose(int(index))
            posu = int(np.where(up == bup)[0][0])
            posd = int(np.where(down == bdn)[0][0])
            avg = vals[posd * self.nsup + posu]
            s += "{:4d}  |".format(int(index))
            for i in range(self.nb_sites):
                u, d = (up >> i) & 1, (down >> i) & 1
                s += " 0 " if (not u and not d) else (" 1 " if u and not d else ("-1 " if d and not u else " 2 "))
            s += f" | {np.around(avg, 4)}" + "\n"
        return s

    def find_stable_space(self):
        raise NotImplementedError("find_stable_space is experimental in the reference (:510-553)")

    # -- Lanczos over the operator (_operators.py:325-366) -----------------------------
    def lanczos(self, maxstep=200, acc_lcz=1.0e-8, neigen=1, init_states=None):
        """The reference's three-term Lanczos on O (NADD = 5, stop when the mean
        change of the neigen lowest Ritz values drops below acc_lcz); O v runs on the
        device, the recurrence in FP64 on the device too."""
        from scipy.linalg import eigh_tridiagonal
        self._require_basis()
        dev = self._prog.ctx.device
        nsel = self.nb_selected
        v = (np.random.random(nsel) - 0.5) if not isinstance(init_states, np.ndarray) else init_states
        v = torch.as_tensor(np.asarray(v, dtype=np.float64), device=dev)
        alpha, beta = [], []
        beta.append(float(torch.linalg.norm(v)))
        v = v / beta[0]
        v0 = torch.zeros_like(v)
        sel = None if self.selected is None else torch.as_tensor(self.selected, device=dev)
        vecs = []
        nstep, ek, ekm1 = 0, np.full(neigen, 100.0), np.full(neigen, 0.0)
        nadd = 5
        Vk = np.ones((1, 1))
        while nstep < maxstep - nadd and not np.linalg.norm(ek[:neigen] - ekm1) / neigen < acc_lcz:
            ekm1 = np.copy(ek[:neigen])
            for _ in range(nadd):
                vecs.append(v)
                if sel is None:
                    v2 = self._prog.apply(v)
                else:
                    full = torch.zeros(self._prog.nstates, dtype=torch.float64, device=dev)
                    full[sel] = v
                    v2 = self._prog.apply(full)[sel]
                alpha.append(float(torch.dot(v2, v)))
                v2 = v2 - alpha[-1] * v - beta[-1] * v0
                beta.append(float(torch.linalg.norm(v2)))
                v0 = v
                v = v2 / beta[-1]
                nstep += 1
            ek, Vk = eigh_tridiagonal(np.array(alpha), np.array(beta[1:nstep]), eigvals_only=False,
                                      select="a")
        self._alpha, self._beta = np.array(alpha), np.array(beta)
        self._ek, self._Vk = ek[:neigen], Vk[:, :neigen]
        basis = torch.stack(vecs[: len(alpha)])
        psi = (basis.T @ torch.as_tensor(Vk[:, :neigen], device=dev)).cpu().numpy()
        return ek[:neigen], psi

    @property
    def alpha(self):
        return self._alpha

    @property
    def beta(self):
        return self._beta

    # -- bookkeeping ----------------------------------------------------------------
    @property
    def filename_operators(self):
        return None

    @property
    def file_size(self):
        return 0.0

    def __delitem__(self, *args):
        pass

    @property
    def index_str(self):
        return str(self._index).zfill(6)

    @property
    def nb_ope(self):
        return len(self.coeff)

    @property
    def T(self):
        return None

    @property
    def dagger(self):
        return self.T

    @staticmethod
    def empty_operator():
        empty = Operators()
        empty.coeff = []
        empty.elem_index = []
        empty.str_index = []
        return empty

    @staticmethod
    def _idt():
        idt = Operators()
        idt.coeff = [1.0]
        idt.elem_index = [[()]]
        idt.str_index = [["id."]]
        return idt

    def __iter__(self):
        terms = []
        for i, c in enumerate(self.coeff):
            prod_ = Operators._idt()
            for s, e in zip(self.str_index[i], self.elem_index[i]):
                prod_ = prod_ * _operator_from_string(s, e)
            terms.append(c * prod_)
        return iter(terms)

    def __multiplemul__(self, other):
        out = Operators._idt()
        for o in other:
            out = out * o
        return out

    def __repr__(self):
        return self.__str__()

    def __str__(self):
        full = ""
        for i, c in enumerate(self.coeff):
            s = f"{np.around(c, 3)}*"
            for j, e in enumerate(self.elem_index[i]):
                s += _print_from_string(self.str_index[i][j], e)
            full += s + " + "
        full = re.sub(r"\*\s", r" ", full)
        full = re.sub(r"(.*) \+ $", r"\1", full)
        full = re.sub(r"\+\s-", r"- ", full)
        return full


def _operator_from_string(string, args):
    if string not in ("c_dagger_c", "c_dagger", "c", "n", "1_n", "sz", "sisj", "splus", "sminus",
                      "id."):
        raise ValueError(f"{string} is not an implemented operator")
    op = Operators(*args)
    op.str_index = [[string]]
    return op


def _print_from_string(string, args):
    if string == "c_dagger_c":
        return f"c^\\dagger_{{{args[0][0]},{args[0][1]}}}c_{{{args[1][0]},{args[1][1]}}}"
    if string == "c_dagger":
        return f"c^\\dagger_{{{args[0][0]},{args[0][1]}}}"
    if string == "c":
        return f"c_{{{args[0][0]},{args[0][1]}}}"
    if string == "n":
        return f"n_{{{args[0][0]},{args[0][1]}}}"
    if string == "1_n":
        return f"(1 - n_{{{args[0][0]},{args[0][1]}}})"
    if string == "sz":
        return f"S^z_{{{args[0][0]}}}"
    if string == "sisj":
        return f"S_{{{args[0][0]}}}S_{{{args[1][0]}}}"
    if string == "splus":
        return f"S^+_{{{args[0][0]}}}"
    if string == "sminus":
        return f"S^-_{{{args[0][0]}}}"
    if string == "id.":
        return ""
    raise ValueError(f"{string} is not an implemented operator")
