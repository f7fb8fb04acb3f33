"""CPU oracle for the exact-diagonalization hot path (TEST INFRASTRUCTURE ONLY).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import
this module, and only as the checker or the timed CPU baseline; the product
package (pyhub_amd) never imports it.

Contents
- ctypes bindings to liboracle.so (ed_oracle.c: a C restatement of the
  reference's Fortran, every function citing the file:line it follows);
- ctypes bindings to the reference's own HPROD compiled from its Fortran
  sources into oracle/_ref (when that build exists);
- numpy dense ED (the DIAGMAT/EXACT_DIAG pattern, exact_diag.f90:2-33),
  closed forms and a finite-ring Lieb-Wu Bethe-ansatz solver.

Parity status of each function is listed in ed_oracle.c's header; the numpy
helpers here are pinned by the closed forms in tests/test_oracle.py.
"""
from __future__ import annotations

import ctypes as C
import math
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_lib = None
_ref_hub = None
_ref_heis = None

u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


def lib():
    global _lib
    if _lib is None:
        # PHUB_ORACLE_SAN=1 loads the ASan/UBSan build (oracle/Makefile `sanitize`)
        name = "liboracle_san.so" if os.environ.get("PHUB_ORACLE_SAN") == "1" else "liboracle.so"
        path = os.path.join(_HERE, name)
        if not os.path.exists(path):
            raise RuntimeError(f"oracle library missing: {path} (run `make -C oracle`)")
        L = C.CDLL(path)
        L.orc_binom.restype = C.c_uint64
        L.orc_binom.argtypes = [C.c_int, C.c_int]
        L.orc_hilbert.argtypes = [C.c_int, C.c_int, C.c_longlong, u64p]
        L.orc_rank_ref.restype = C.c_longlong
        L.orc_rank_ref.argtypes = [C.c_int, C.c_int, C.c_uint64]
        hub_args = [C.c_int, C.c_int, C.c_int, C.c_longlong, C.c_longlong, u64p, u64p,
                    f64p, f64p, f64p, f64p]
        L.orc_hprod_hubbard.argtypes = hub_args
        L.orc_hprod_hubbard_fast.argtypes = hub_args + [C.c_int]
        L.orc_hprod_hubbard_ul.argtypes = [C.c_int, C.c_int, C.c_int, C.c_longlong, C.c_longlong,
                                           u64p, u64p, f64p, f64p, f64p, f64p, f64p]
        L.orc_two_rdm.argtypes = [C.c_int, C.c_longlong, C.c_longlong, u64p, u64p, f64p, f64p, f64p]
        L.orc_hprod_heis.argtypes = [C.c_int, C.c_longlong, u64p, f64p, f64p, f64p]
        L.orc_hprod_heis_fast.argtypes = [C.c_int, C.c_longlong, u64p, f64p, f64p, f64p]
        L.orc_tridiag_eig.restype = C.c_int
        L.orc_tridiag_eig.argtypes = [C.c_int, f64p, f64p, C.c_void_p]
        L.orc_lanczos.restype = C.c_int
        L.orc_lanczos.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_longlong, C.c_longlong,
                                  u64p, u64p, f64p, f64p, f64p, C.c_int, C.c_int,
                                  C.c_int, C.c_double, f64p, f64p, f64p, f64p, f64p,
                                  C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.orc_one_rdm.argtypes = [C.c_int, C.c_int, C.c_int, C.c_longlong, C.c_longlong,
                                  u64p, u64p, f64p, f64p, f64p]
        L.orc_ni.argtypes = [C.c_int, C.c_longlong, C.c_longlong, u64p, u64p, f64p, f64p, f64p]
        L.orc_j_spin.argtypes = [C.c_int, C.c_longlong, C.c_longlong, u64p, u64p, f64p, f64p, f64p]
        L.orc_j_spin_heis.argtypes = [C.c_int, C.c_longlong, u64p, f64p, f64p, f64p]
        L.orc_excite.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                 C.c_longlong, C.c_longlong, u64p, u64p,
                                 C.c_longlong, C.c_longlong, u64p, u64p, f64p, f64p]
        L.orc_cf_chain.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_longlong, C.c_longlong,
                                   u64p, u64p, f64p, f64p, f64p, C.c_int, C.c_int, f64p, C.c_int,
                                   f64p, f64p, C.POINTER(C.c_double)]
        L.orc_unrank_ref.restype = C.c_uint64
        L.orc_unrank_ref.argtypes = [C.c_int, C.c_int, C.c_uint64]
        L.orc_heis_column.restype = C.c_longlong
        L.orc_heis_column.argtypes = [C.c_int, f64p, C.c_int, C.c_uint64, u64p, f64p]
        _lib = L
    return _lib


def ref_available() -> bool:
    return (os.path.exists(os.path.join(_HERE, "_ref", "libref_hub.so"))
            and os.path.exists(os.path.join(_HERE, "_ref", "libref_heis.so")))


def _ref_libs():
    global _ref_hub, _ref_heis
    if _ref_hub is None:
        _ref_hub = C.CDLL(os.path.join(_HERE, "_ref", "libref_hub.so"))
        _ref_hub.phref_hub_setup.argtypes = [C.c_int] * 5 + [i32p, i32p, f64p, f64p]
        _ref_hub.phref_hub_hprod.argtypes = [f64p, f64p]
        _ref_hub.phref_hub_setup_ul.argtypes = [C.c_int] * 5 + [i32p, i32p, f64p, f64p, f64p]
        _ref_hub.phref_hub_one_rdm.argtypes = [f64p, f64p, f64p]
        _ref_hub.phref_hub_two_rdm.argtypes = [f64p, f64p, C.c_int, f64p, f64p]
        _ref_hub.phref_hub_j_spin.argtypes = [f64p, f64p, f64p]
        _ref_hub.phref_hub_gf.argtypes = ([C.c_int] * 6 + [i32p, i32p, C.c_int, C.c_int, i32p, i32p,
                                           f64p, C.c_int, f64p])
        _ref_heis = C.CDLL(os.path.join(_HERE, "_ref", "libref_heis.so"))
        _ref_heis.phref_heis_setup.argtypes = [C.c_int, C.c_int, i32p, f64p]
        _ref_heis.phref_heis_hprod.argtypes = [f64p, f64p]
        _ref_heis.phref_heis_j_spin.argtypes = [f64p, f64p, f64p]
    return _ref_hub, _ref_heis


# ---------------------------------------------------------------------------
# basis
# ---------------------------------------------------------------------------
def binom(n: int, k: int) -> int:
    return math.comb(n, k) if 0 <= k <= n else 0


def hilbert(n: int, k: int) -> np.ndarray:
    """Reference basis order (restates HILBERT, compute_basis.f90:67-96)."""
    ns = binom(n, k)
    out = np.zeros(ns, dtype=np.uint64)
    lib().orc_hilbert(n, k, ns, out)
    return out


def rank_ref(n: int, k: int, x: int) -> int:
    return int(lib().orc_rank_ref(n, k, int(x)))


def unrank_ref(n: int, k: int, idx: int) -> int:
    """Mask of reference-order index idx (inverse of rank_ref; any n <= 63)."""
    return int(lib().orc_unrank_ref(n, k, int(idx)))


def heis_column(n: int, Jm, idx: int, sign_convention: int = 0) -> dict:
    """Column idx of the Heisenberg HPROD matrix (orc_heis_column) as {row: value}
    with repeated rows summed; works for n = 36 (indices past 2^32)."""
    Jm = np.ascontiguousarray(Jm, dtype=np.float64)
    rows = np.zeros(n * n + 1, dtype=np.uint64)
    vals = np.zeros(n * n + 1)
    m = int(lib().orc_heis_column(n, Jm, int(sign_convention), int(idx), rows, vals))
    out: dict = {}
    for r, v in zip(rows[:m].tolist(), vals[:m].tolist()):
        out[r] = out.get(r, 0.0) + v
    return out


def heis_dense(n: int, Jm, sign_convention: int = 0) -> np.ndarray:
    """Dense S_z = 0 Heisenberg matrix from orc_heis_column (reference order), in
    either flip-sign convention; small n only."""
    D = binom(n, n // 2)
    H = np.zeros((D, D))
    for c in range(D):
        for r, v in heis_column(n, Jm, c, sign_convention).items():
            H[r, c] += v
    return H


def brute_basis(n: int, k: int) -> np.ndarray:
    """All n-bit masks with k set bits, sorted lexicographically by their
    increasing list of set-bit positions (bit 0 first).  This is the order
    HILBERT produces (compute_basis.f90:67-96), written as a brute-force sort."""
    masks = [x for x in range(1 << n) if bin(x).count("1") == k]
    masks.sort(key=lambda x: tuple(b for b in range(n) if x >> b & 1))
    return np.array(masks, dtype=np.uint64)


# ---------------------------------------------------------------------------
# models
# ---------------------------------------------------------------------------
class HubbardModel:
    """Hubbard cluster in the reference conventions: T[J-1][K-1] = T(J,K)
    (hopping -t on bonds, eps_i on the diagonal), U per site."""

    def __init__(self, n, nup, ndn, T, U):
        self.n, self.nup, self.ndn = int(n), int(nup), int(ndn)
        self.T = np.ascontiguousarray(T, dtype=np.float64)
        U = np.asarray(U, dtype=np.float64)
        # a 4-index U (n, n, n, n) is the is_Uloc = 3 interaction (core_solver.py:142-148)
        self.UL = None
        if U.ndim == 4:
            self.UL = np.ascontiguousarray(U)
            self.UP = np.ascontiguousarray(np.einsum("ijkl,ijkl->ij", U, U))
            U = np.zeros(n)
        self.U = np.ascontiguousarray(np.full(n, float(U)) if U.ndim == 0 else U)
        self.bup = hilbert(n, nup)
        self.bdn = hilbert(n, ndn)
        self.nsup, self.nsdn = len(self.bup), len(self.bdn)
        self.dim = self.nsup * self.nsdn

    def hv(self, v, fast=False, nthreads=1):
        v = np.ascontiguousarray(v, dtype=np.float64)
        w = np.empty_like(v)
        if self.UL is not None:
            lib().orc_hprod_hubbard_ul(self.n, self.nup, self.ndn, self.nsup, self.nsdn,
                                       self.bup, self.bdn, self.T, self.UL, self.UP, v, w)
        elif fast:
            lib().orc_hprod_hubbard_fast(self.n, self.nup, self.ndn, self.nsup, self.nsdn,
                                         self.bup, self.bdn, self.T, self.U, v, w, nthreads)
        else:
            lib().orc_hprod_hubbard(self.n, self.nup, self.ndn, self.nsup, self.nsdn,
                                    self.bup, self.bdn, self.T, self.U, v, w)
        return w

    def hv_ref(self, v):
        """The reference's own HPROD (Fortran, compiled from its sources)."""
        hub, _ = _ref_libs()
        if self.UL is not None:
            hub.phref_hub_setup_ul(self.n, self.nup, self.ndn, self.nsup, self.nsdn,
                                   self.bup.astype(np.int32), self.bdn.astype(np.int32),
                                   self.T, self.UL, self.UP)
        else:
            hub.phref_hub_setup(self.n, self.nup, self.ndn, self.nsup, self.nsdn,
                                self.bup.astype(np.int32), self.bdn.astype(np.int32),
                                self.T, self.U)
        v = np.ascontiguousarray(v, dtype=np.float64)
        w = np.empty_like(v)
        hub.phref_hub_hprod(v, w)
        return w

    def _ref_setup(self):
        hub, _ = _ref_libs()
        hub.phref_hub_setup(self.n, self.nup, self.ndn, self.nsup, self.nsdn,
                            self.bup.astype(np.int32), self.bdn.astype(np.int32),
                            self.T, self.U if self.UL is None else np.zeros(self.n))
        return hub

    def one_rdm_ref(self, bra, ket=None):
        """DENS_MAT of the reference's own ONE_RDM (static_rq.f90:153-244), as
        one_rdm(): [s][J-1][K-1]."""
        hub = self._ref_setup()
        ket = bra if ket is None else ket
        out = np.zeros(2 * self.n * self.n)
        hub.phref_hub_one_rdm(np.ascontiguousarray(bra, dtype=np.float64),
                              np.ascontiguousarray(ket, dtype=np.float64), out)
        return out.reshape(2, self.n, self.n).transpose(0, 2, 1).copy()

    def two_rdm_ref(self, bra, ket=None, full=False):
        """The reference's own TWO_RDM (static_rq.f90:247-320): (NI, DBL_OCC or
        None); DBL_OCC as two_rdm(): [J-1][K-1][M-1][N-1]."""
        hub = self._ref_setup()
        ket = bra if ket is None else ket
        db = np.zeros(self.n ** 4)
        ni = np.zeros(self.n)
        hub.phref_hub_two_rdm(np.ascontiguousarray(bra, dtype=np.float64),
                              np.ascontiguousarray(ket, dtype=np.float64), int(full), db, ni)
        d4 = db.reshape(self.n, self.n, self.n, self.n).transpose(3, 2, 1, 0).copy() if full else None
        return ni, d4

    def sisj_ref(self, bra, ket=None):
        """SMATRIX of the reference's own J_SPIN (static_rq.f90:323-395)."""
        hub = self._ref_setup()
        ket = bra if ket is None else ket
        out = np.zeros(self.n * self.n)
        hub.phref_hub_j_spin(np.ascontiguousarray(bra, dtype=np.float64),
                             np.ascontiguousarray(ket, dtype=np.float64), out)
        return out.reshape(self.n, self.n).T.copy()

    def excite_ref(self, psi, spin, create, ncomp=None):
        """Excitation vectors of the reference's own GF_E_UP / GF_E_DOWN / GF_H_UP /
        GF_H_DOWN (spgf_ed.f90:304-429) for sites 1..ncomp: (target model,
        array [site-1][target state])."""
        hub, _ = _ref_libs()
        ncomp = self.n if ncomp is None else int(ncomp)
        dn = 1 if create else -1
        tgt = HubbardModel(self.n, self.nup + (dn if spin == 0 else 0),
                           self.ndn + (dn if spin == 1 else 0), self.T, self.U)
        kind = (0 if create else 2) + (1 if spin == 1 else 0)
        q = np.zeros(tgt.dim * ncomp)
        hub.phref_hub_gf(kind, self.n, self.nup, self.ndn, self.nsup, self.nsdn,
                         self.bup.astype(np.int32), self.bdn.astype(np.int32),
                         tgt.nsup, tgt.nsdn, tgt.bup.astype(np.int32), tgt.bdn.astype(np.int32),
                         np.ascontiguousarray(psi, dtype=np.float64), ncomp, q)
        return tgt, q.reshape(ncomp, tgt.dim)

    def dense(self, fast=True):
        H = np.zeros((self.dim, self.dim))
        e = np.zeros(self.dim)
        for i in range(self.dim):
            e[i] = 1.0
            H[:, i] = self.hv(e, fast=fast)
            e[i] = 0.0
        return H

    def _args(self, fast, nthreads):
        return (0, self.n, self.nup, self.ndn, self.nsup, self.nsdn, self.bup, self.bdn,
                self.T, self.U, np.zeros(1), int(fast), int(nthreads))

    def lanczos(self, v0, maxlcz=600, acc=1e-8, fast=True, nthreads=1):
        if self.UL is not None:
            raise NotImplementedError("the oracle Lanczos covers the per-site U only")
        return _lanczos(self._args(fast, nthreads), v0, maxlcz, acc, self.dim)

    def two_rdm(self, bra, ket=None):
        """DBL_OCC(J,K,M,N) of TWO_RDM (static_rq.f90:247-306), shape (n, n, n, n)."""
        ket = bra if ket is None else ket
        out = np.zeros(self.n ** 4)
        lib().orc_two_rdm(self.n, self.nsup, self.nsdn, self.bup, self.bdn,
                          np.ascontiguousarray(bra, dtype=np.float64),
                          np.ascontiguousarray(ket, dtype=np.float64), out)
        return out.reshape(self.n, self.n, self.n, self.n)

    def one_rdm(self, bra, ket=None):
        ket = bra if ket is None else ket
        out = np.zeros(2 * self.n * self.n)
        lib().orc_one_rdm(self.n, self.nup, self.ndn, self.nsup, self.nsdn, self.bup, self.bdn,
                          np.ascontiguousarray(bra, dtype=np.float64),
                          np.ascontiguousarray(ket, dtype=np.float64), out)
        return out.reshape(2, self.n, self.n)

    def ni(self, bra, ket=None):
        ket = bra if ket is None else ket
        out = np.zeros(self.n)
        lib().orc_ni(self.n, self.nsup, self.nsdn, self.bup, self.bdn,
                     np.ascontiguousarray(bra, dtype=np.float64),
                     np.ascontiguousarray(ket, dtype=np.float64), out)
        return out

    def sisj(self, bra, ket=None):
        ket = bra if ket is None else ket
        out = np.zeros(self.n * self.n)
        lib().orc_j_spin(self.n, self.nsup, self.nsdn, self.bup, self.bdn,
                         np.ascontiguousarray(bra, dtype=np.float64),
                         np.ascontiguousarray(ket, dtype=np.float64), out)
        return out.reshape(self.n, self.n)

    def excite(self, psi, site, spin, create):
        """c^dag_{site,spin} psi (create=1) or c_{site,spin} psi (create=0),
        site 1-based; returns (target model, vector)."""
        dn = (1 if create else -1)
        tgt = HubbardModel(self.n, self.nup + (dn if spin == 0 else 0),
                           self.ndn + (dn if spin == 1 else 0), self.T, self.U)
        out = np.zeros(tgt.dim)
        lib().orc_excite(self.n, site, spin, int(create), self.ndn,
                         self.nsup, self.nsdn, self.bup, self.bdn,
                         tgt.nsup, tgt.nsdn, tgt.bup, tgt.bdn,
                         np.ascontiguousarray(psi, dtype=np.float64), out)
        return tgt, out

    def cf_chain(self, start, nsteps, fast=True, nthreads=1):
        a = np.zeros(nsteps)
        b = np.zeros(nsteps)
        n2 = C.c_double(0)
        lib().orc_cf_chain(*self._args(fast, nthreads), np.ascontiguousarray(start, dtype=np.float64),
                           nsteps, a, b, C.byref(n2))
        return a, b, n2.value


class HeisenbergModel:
    """Heisenberg cluster in the reference convention (hprod_heis.f90:1-61):
    H = sum over ORDERED pairs J(J,K) S_J.S_K, S_z = 0 sector."""

    def __init__(self, n, Jm):
        self.n = int(n)
        self.Jm = np.ascontiguousarray(Jm, dtype=np.float64)
        self.bup = hilbert(n, n // 2)
        self.nsup = len(self.bup)
        self.dim = self.nsup

    def hv(self, v, fast=False):
        v = np.ascontiguousarray(v, dtype=np.float64)
        w = np.empty_like(v)
        f = lib().orc_hprod_heis_fast if fast else lib().orc_hprod_heis
        f(self.n, self.nsup, self.bup, self.Jm, v, w)
        return w

    def hv_ref(self, v):
        _, heis = _ref_libs()
        heis.phref_heis_setup(self.n, self.nsup, self.bup.astype(np.int32), self.Jm)
        v = np.ascontiguousarray(v, dtype=np.float64)
        w = np.empty_like(v)
        heis.phref_heis_hprod(v, w)
        return w

    def dense(self):
        H = np.zeros((self.dim, self.dim))
        e = np.zeros(self.dim)
        for i in range(self.dim):
            e[i] = 1.0
            H[:, i] = self.hv(e)
            e[i] = 0.0
        return H

    def lanczos(self, v0, maxlcz=600, acc=1e-8):
        args = (1, self.n, self.n // 2, 0, self.nsup, 1, self.bup, self.bup,
                np.zeros(1), np.zeros(1), self.Jm, 0, 1)
        return _lanczos(args, v0, maxlcz, acc, self.dim)

    def sisj_ref(self, bra, ket=None):
        """SMATRIX of the reference's own Heisenberg J_SPIN (static_rq_heis.f90:58-116)."""
        _, heis = _ref_libs()
        heis.phref_heis_setup(self.n, self.nsup, self.bup.astype(np.int32), self.Jm)
        ket = bra if ket is None else ket
        out = np.zeros(self.n * self.n)
        heis.phref_heis_j_spin(np.ascontiguousarray(bra, dtype=np.float64),
                               np.ascontiguousarray(ket, dtype=np.float64), out)
        return out.reshape(self.n, self.n).T.copy()

    def sisj(self, bra, ket=None):
        ket = bra if ket is None else ket
        out = np.zeros(self.n * self.n)
        lib().orc_j_spin_heis(self.n, self.nsup, self.bup,
                              np.ascontiguousarray(bra, dtype=np.float64),
                              np.ascontiguousarray(ket, dtype=np.float64), out)
        return out.reshape(self.n, self.n)


def _lanczos(args, v0, maxlcz, acc, dim):
    a = np.zeros(maxlcz + 1)
    b = np.zeros(maxlcz + 2)
    ev = np.zeros(maxlcz + 1)
    psi = np.zeros(dim)
    ndeg = C.c_int(0)
    lost = C.c_int(0)
    nb = lib().orc_lanczos(*args, maxlcz, acc, np.ascontiguousarray(v0, dtype=np.float64),
                           a, b, ev, psi, C.byref(ndeg), C.byref(lost))
    return {"e0": ev[0], "eig": ev[:nb].copy(), "alpha": a[:nb].copy(), "beta": b[:nb].copy(),
            "nblcz": nb, "psi": psi, "ndeg": ndeg.value, "norm_lost": bool(lost.value)}


def tridiag_eig(alpha, beta_off):
    """Eigen-decomposition of the symmetric tridiagonal (alpha, beta_off)."""
    m = len(alpha)
    d = np.ascontiguousarray(alpha, dtype=np.float64).copy()
    e = np.ascontiguousarray(np.append(beta_off, 0.0)[: max(m - 1, 1)], dtype=np.float64)
    z = np.zeros(m * m)
    rc = lib().orc_tridiag_eig(m, d, e, z.ctypes.data_as(C.c_void_p))
    if rc:
        raise RuntimeError("tridiagonal QL did not converge")
    return d, z.reshape(m, m).T  # columns are eigenvectors


# ---------------------------------------------------------------------------
# lattices and closed forms
# ---------------------------------------------------------------------------
def ring_T(L, t=1.0, eps=None, periodic=True):
    T = np.zeros((L, L))
    for i in range(L - (0 if periodic else 1)):
        j = (i + 1) % L
        T[i, j] = T[j, i] = -t
    if eps is not None:
        T[np.arange(L), np.arange(L)] = eps
    return T


def disordered_ring_T(L, seed=11, scale=0.6):
    """Ring hoppings plus a random symmetric matrix (all pairs, on-site terms
    included): no lattice symmetry is left, so low spectra are non-degenerate."""
    rng = np.random.default_rng(seed)
    T = rng.normal(size=(L, L)) * scale
    return (T + T.T) / 2 + ring_T(L)


def square_T(Lx, Ly, t=1.0, tp=0.0, periodic=True):
    """Site s = x + Lx*y; T = -t on NN bonds, -tp on NNN bonds."""
    n = Lx * Ly
    T = np.zeros((n, n))

    def s(x, y):
        return (x % Lx) + Lx * (y % Ly)

    for y in range(Ly):
        for x in range(Lx):
            i = s(x, y)
            for dx, dy, amp in ((1, 0, -t), (0, 1, -t), (1, 1, -tp), (1, -1, -tp)):
                xx, yy = x + dx, y + dy
                if not periodic and not (0 <= xx < Lx and 0 <= yy < Ly):
                    continue
                j = s(xx, yy)
                if i == j or amp == 0.0:
                    continue
                T[i, j] += amp if T[i, j] == 0.0 else 0.0
                T[j, i] = T[i, j]
    return T


def free_fermion_e0(T, nup, ndn):
    ek = np.linalg.eigvalsh(T)
    return float(np.sum(ek[:nup]) + np.sum(ek[:ndn]))


def dimer_e0(t, U):
    return U / 2 - math.sqrt(U * U / 4 + 4 * t * t)


def heisenberg_standard_dense(n, Jstd):
    """Standard H = sum_{i<j} J_ij S_i.S_j on n spins-1/2, dense (S_z=0 block
    is not extracted: full 2^n space).  For small n only."""
    sx = np.array([[0, 0.5], [0.5, 0]])
    sy = np.array([[0, -0.5j], [0.5j, 0]])
    sz = np.array([[0.5, 0], [0, -0.5]])
    I = np.eye(2)

    def op(o, i):
        m = np.array([[1.0]])
        for k in range(n):
            m = np.kron(m, o if k == i else I)
        return m

    H = np.zeros((2 ** n, 2 ** n), dtype=complex)
    for i in range(n):
        for j in range(i + 1, n):
            if Jstd[i, j] != 0:
                for o in (sx, sy, sz):
                    H += Jstd[i, j] * op(o, i) @ op(o, j)
    return H.real


def heisenberg_standard_sz0(n, Jstd):
    """Standard H = sum_{i<j} J_ij S_i.S_j restricted to S_z = 0 (n even), as a
    scipy sparse matrix over the masks with n/2 set bits (set bit = spin up,
    numerical order).  S_i.S_j = S^z_i S^z_j + (S^+_i S^-_j + S^-_i S^+_j)/2:
    +-1/4 on the diagonal, 1/2 between states that differ by swapping i and j.
    Independent of any Jordan-Wigner gauge; checked against the full-space
    Kronecker construction above."""
    import itertools
    import scipy.sparse as sp
    k = n // 2
    states = np.array(sorted(sum(1 << b for b in c) for c in itertools.combinations(range(n), k)),
                      dtype=np.int64)
    D = len(states)
    diag = np.zeros(D)
    rows, cols, vals = [], [], []
    idx = np.arange(D)
    for i in range(n):
        for j in range(i + 1, n):
            Jij = Jstd[i, j]
            if Jij == 0:
                continue
            bi = (states >> i) & 1
            bj = (states >> j) & 1
            diag += np.where(bi == bj, 0.25, -0.25) * Jij
            sel = bi != bj
            tgt = states[sel] ^ ((1 << i) | (1 << j))
            rows.append(np.searchsorted(states, tgt))
            cols.append(idx[sel])
            vals.append(np.full(int(sel.sum()), 0.5 * Jij))
    rows.append(idx)
    cols.append(idx)
    vals.append(diag)
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                         shape=(D, D))


def j1j2_square(Lx, Ly, J1, J2):
    """Standard-convention couplings Jstd[i, j] (i<j) of the J1-J2 model on an
    Lx x Ly torus, site s = x + Lx*y; duplicate bonds of small tori add up."""
    n = Lx * Ly
    Jstd = np.zeros((n, n))
    for y in range(Ly):
        for x in range(Lx):
            s = x + Lx * y
            for dx, dy, Jb in ((1, 0, J1), (0, 1, J1), (1, 1, J2), (1, -1, J2)):
                t = (x + dx) % Lx + Lx * ((y + dy) % Ly)
                if t == s:
                    continue
                i, j = min(s, t), max(s, t)
                Jstd[i, j] += Jb
    return Jstd


def lieb_wu_ring(L, nup, ndn, U, t=1.0, I=None, Jq=None, tol=1e-14):
    """Finite periodic ring, Lieb-Wu Bethe equations for H = -t sum c^dag c + U sum n n,
    N = nup + ndn electrons, M = ndn down spins (M <= N/2):
        L k_j = 2 pi I_j - sum_b 2 atan((sin k_j - Lam_b)/u),           u = U/(4t)
        sum_j 2 atan((Lam_a - sin k_j)/u) = 2 pi J_a + sum_b 2 atan((Lam_a - Lam_b)/(2u))
        E = -2 t sum_j cos k_j.
    Quantum numbers default to the ground-state choice (I_j integer for M even,
    half-odd for M odd, J_a integer for N-M odd); when N is even and I_j are
    integers, the set is shifted to {-(N/2-1)..N/2}."""
    from scipy.optimize import fsolve
    N, M = nup + ndn, min(nup, ndn)
    u = U / (4.0 * t)
    if I is None:
        off = 0.0 if M % 2 == 0 else 0.5
        I = np.arange(N) - (N - 1) / 2.0
        if (I[0] - off) % 1.0 != 0.0:
            I = I + 0.5
    if Jq is None:
        Jq = np.arange(M) - (M - 1) / 2.0
    I = np.asarray(I, float)
    Jq = np.asarray(Jq, float)

    def eqs(x):
        k, lam = x[:N], x[N:]
        r1 = L * k - 2 * np.pi * I + np.sum(2 * np.arctan((np.sin(k)[:, None] - lam[None, :]) / u), axis=1)
        r2 = (np.sum(2 * np.arctan((lam[:, None] - np.sin(k)[None, :]) / u), axis=1)
              - 2 * np.pi * Jq - np.sum(2 * np.arctan((lam[:, None] - lam[None, :]) / (2 * u)), axis=1))
        return np.concatenate([r1, r2])

    k0 = 2 * np.pi * I / L
    lam0 = np.sinh(np.pi * Jq / max(M, 1)) * 0.5 if M else np.zeros(0)
    x, info, ier, msg = fsolve(eqs, np.concatenate([k0, lam0]), xtol=tol, full_output=True)
    res = np.max(np.abs(eqs(x))) if len(x) else 0.0
    return float(-2 * t * np.sum(np.cos(x[:N]))), float(res)


# ---------------------------------------------------------------------------
# generic operator strings (operators.x), restated in Python loops
# ---------------------------------------------------------------------------
def _anticom(n, pos, m):
    """ANTICOM (funcmod.f90:29-41): (-1)^(occupied sites 1..pos-1), site J = bit n-J."""
    s = 1.0
    for i2 in range(1, pos):
        if (m >> (n - i2)) & 1:
            s = -s
    return s


def _bit(n, pos):
    return 1 << (n - pos)


class OperatorOracle:
    """Restates operators.x on a product basis of fixed-N sectors.

    bup / bdn are the concatenated sector lists in the given order
    (compute_operators.f90:10-26); the state index is d * nsup + u.  Terms are
    (coeff, [(code, site1, spin1, site2, spin2), ...]) with the OPE codes of
    compute_operators.f90:36-63; factors act in list order (first one first).

    apply() follows the vector path: C_DAGGER_C (operators.f90:1-55),
    C_DAGGER_C_UNRES (:58-144), NI (:148-172), NIM1 (:196-220), SZ_OPE (:224-245),
    SISJ (:303-394), C_DAGGER (:397-446), C_ (:450-499).  matrix() follows the
    basis-element path OPERATOR_BASIS (operators_basis.f90:1-56) with
    C_DAGGER_C_BASIS (:59-100), NI/NIM1_BASIS (:103-130), SZ_BASIS (:134-143),
    SISJ_BASIS (:147-186), C_DAGGER_C_UNRES_BASIS, C_DAGGER_BASIS, C_BASIS.
    Three places differ from the Fortran text, all stated: SISJ with K = J reads
    the current up state (the Fortran reads BUP(K), the K-th up state); a
    basis-element path whose next state is missing from the lists is dropped (the
    Fortran keeps the previous position); and SISJ_BASIS's second try restarts
    from the state the factor started on (the Fortran restarts from the product's
    first basis element, the same state when S.S is the first factor)."""

    def __init__(self, n, nup, ndn):
        self.n = n
        self.bup = np.concatenate([hilbert(n, k) for k in dict.fromkeys(nup)]).astype(np.int64)
        self.bdn = np.concatenate([hilbert(n, k) for k in dict.fromkeys(ndn)]).astype(np.int64)
        self.nsup, self.nsdn = len(self.bup), len(self.bdn)
        self.dim = self.nsup * self.nsdn
        self.pu = {int(m): i for i, m in enumerate(self.bup)}
        self.pd = {int(m): i for i, m in enumerate(self.bdn)}

    # -- vector path ---------------------------------------------------------------
    def _cdc(self, J, K, spin, v):
        n, out = self.n, np.zeros_like(v)
        if spin == 1:
            for i, x in enumerate(self.bup):
                x = int(x)
                a, b = _bit(n, K), _bit(n, J)
                new = 0
                if x & a and not x & b:
                    new = x - a + b
                if x & a and J == K:
                    new = x
                if new > 0 and new in self.pu:
                    pos = self.pu[new]
                    ac = _anticom(n, K, x) * _anticom(n, J, new)
                    for l1 in range(self.nsdn):
                        out[l1 * self.nsup + pos] += ac * v[l1 * self.nsup + i]
        else:
            for l1, x in enumerate(self.bdn):
                x = int(x)
                a, b = _bit(n, K), _bit(n, J)
                new = 0
                if x & a and not x & b:
                    new = x - a + b
                if x & a and J == K:
                    new = x
                if new > 0 and new in self.pd:
                    pos = self.pd[new]
                    ac = _anticom(n, K, x) * _anticom(n, J, new)
                    for i in range(self.nsup):
                        out[pos * self.nsup + i] += ac * v[l1 * self.nsup + i]
        return out

    def _unres(self, J, K, spink, v):
        n, out = self.n, np.zeros_like(v)
        if spink == 1:   # remove up at K, add down at J
            for i, x in enumerate(self.bup):
                x = int(x)
                if not x & _bit(n, K):
                    continue
                new = x - _bit(n, K)
                if new not in self.pu:
                    continue
                pos, ac = self.pu[new], _anticom(n, K, new)
                for l1, y in enumerate(self.bdn):
                    y = int(y)
                    if y & _bit(n, J):
                        continue
                    new2 = y + _bit(n, J)
                    if new2 in self.pd:
                        ac2 = _anticom(n, J, new2)
                        out[self.pd[new2] * self.nsup + pos] += ac * ac2 * v[l1 * self.nsup + i]
        else:            # remove down at K, add up at J
            for i, x in enumerate(self.bdn):
                x = int(x)
                if not x & _bit(n, K):
                    continue
                new = x - _bit(n, K)
                if new not in self.pd:
                    continue
                pos, ac = self.pd[new], _anticom(n, K, new)
                for l1, y in enumerate(self.bup):
                    y = int(y)
                    if y & _bit(n, J):
                        continue
                    new2 = y + _bit(n, J)
                    if new2 in self.pu:
                        ac2 = _anticom(n, J, new2)
                        out[pos * self.nsup + self.pu[new2]] += ac * ac2 * v[i * self.nsup + l1]
        return out

    def _diag(self, fn, v):
        out = np.zeros_like(v)
        for d, y in enumerate(self.bdn):
            for u, x in enumerate(self.bup):
                out[d * self.nsup + u] = fn(int(x), int(y)) * v[d * self.nsup + u]
        return out

    def _sz1(self, J, x, y):
        n = self.n
        return (0.5 if x & _bit(n, J) else 0.0) - (0.5 if y & _bit(n, J) else 0.0)

    def _create_annihilate(self, J, spin, v, create):
        n, out = self.n, np.zeros_like(v)
        b = _bit(n, J)
        if spin == 1:
            for i, x in enumerate(self.bup):
                x = int(x)
                if bool(x & b) == create:
                    continue
                new = x + b if create else x - b
                if new in self.pu:
                    ac = _anticom(n, J, x)   # ANTICOM(NORB, 0, .) = 1 for the down part
                    for l2 in range(self.nsdn):
                        out[l2 * self.nsup + self.pu[new]] += ac * v[l2 * self.nsup + i]
        else:
            for i, x in enumerate(self.bdn):
                x = int(x)
                if bool(x & b) == create:
                    continue
                new = x + b if create else x - b
                if new in self.pd:
                    ac = _anticom(n, J, x)
                    for l1 in range(self.nsup):
                        out[self.pd[new] * self.nsup + l1] += ac * v[i * self.nsup + l1]
        return out

    def _sisj(self, K, J, v):
        n, out = self.n, np.zeros_like(v)
        if K == J:
            b = _bit(n, J)
            out += self._diag(lambda x, y: 0.5 if bool(x & b) != bool(y & b) else 0.0, v)
        else:
            # up J -> K with down K -> J, then down J -> K with up K -> J, each times -1/2
            out -= 0.5 * self._cdc(J, K, 2, self._cdc(K, J, 1, v))
            out -= 0.5 * self._cdc(J, K, 1, self._cdc(K, J, 2, v))
        out += self._diag(lambda x, y: self._sz1(J, x, y) * self._sz1(K, x, y), v)
        return out

    def _factor(self, f, v):
        code, s1, p1, s2, p2 = f
        n = self.n
        if code == 1:
            return self._cdc(s1, s2, p1, v)
        if code == 2:
            return self._diag(lambda x, y: 1.0 if ((x if p1 == 1 else y) & _bit(n, s1)) else 0.0, v)
        if code == 3:
            return v.copy()
        if code == 4:
            return self._diag(lambda x, y: 0.0 if ((x if p1 == 1 else y) & _bit(n, s1)) else 1.0, v)
        if code == 5:   # SZ_OPE: (1 - n_up)/2 - (1 - n_dn)/2
            return self._diag(lambda x, y: (0.0 if x & _bit(n, s1) else 0.5)
                              - (0.0 if y & _bit(n, s1) else 0.5), v)
        if code == 6:
            return self._unres(s1, s1, 2, v)
        if code == 7:
            return self._unres(s1, s1, 1, v)
        if code == 8:
            return self._sisj(s1, s2, v)
        if code == 9:
            return self._unres(s1, s2, p2, v)
        if code == 10:
            return self._create_annihilate(s1, p1, v, True)
        if code == 11:
            return self._create_annihilate(s1, p1, v, False)
        raise ValueError(code)

    def apply(self, terms, v):
        v = np.asarray(v, dtype=np.float64)
        out = np.zeros(self.dim)
        for coeff, factors in terms:
            if not abs(coeff) > 1e-14:
                continue
            cur = v.copy()
            for f in factors:
                cur = self._factor(f, cur)
            out += coeff * cur
        return out

    # -- basis-element path ----------------------------------------------------------
    def _walk(self, f, x, y, val):
        """One factor on the single state (x, y) with value val (operators_basis.f90)."""
        code, s1, p1, s2, p2 = f
        n = self.n
        if code == 1:
            m = x if p1 == 1 else y
            a, b = _bit(n, s2), _bit(n, s1)
            new = 0
            if m & a and not m & b:
                new = m - a + b
            if m & a and s1 == s2:
                new = m
            if new == 0:
                return None
            val *= _anticom(n, s2, m) * _anticom(n, s1, new)
            return (new, y, val) if p1 == 1 else (x, new, val)
        if code == 2:
            return (x, y, val) if ((x if p1 == 1 else y) & _bit(n, s1)) else None
        if code == 3:
            return x, y, val
        if code == 4:
            return None if ((x if p1 == 1 else y) & _bit(n, s1)) else (x, y, val)
        if code == 5:   # SZ_BASIS: funcmod's SZ
            z = self._sz1(s1, x, y)
            return (x, y, val * z) if z != 0.0 else None
        if code in (6, 7, 9):
            J, K = (s1, s1) if code in (6, 7) else (s1, s2)
            spk = 2 if code == 6 else (1 if code == 7 else p2)
            rem, add = (x, y) if spk == 1 else (y, x)
            if not rem & _bit(n, K) or add & _bit(n, J):
                return None
            val *= _anticom(n, K, rem) * _anticom(n, J, add)
            rem, add = rem - _bit(n, K), add + _bit(n, J)
            return (rem, add, val) if spk == 1 else (add, rem, val)
        if code == 8:
            K, J = s1, s2
            if K == J:
                b = _bit(n, J)
                return (x, y, val / 2) if bool(x & b) != bool(y & b) else None
            r = self._walk((1, K, 1, J, 1), x, y, val)
            r = self._walk((1, J, 2, K, 2), *r) if r else None
            if r is None:
                r = self._walk((1, K, 2, J, 2), x, y, val)
                r = self._walk((1, J, 1, K, 1), *r) if r else None
            return (r[0], r[1], -r[2] / 2) if r else None
        if code in (10, 11):
            m = x if p1 == 1 else y
            b = _bit(n, s1)
            if bool(m & b) == (code == 10):
                return None
            val *= _anticom(n, s1, m)
            m = m + b if code == 10 else m - b
            return (m, y, val) if p1 == 1 else (x, m, val)
        raise ValueError(code)

    def matrix(self, terms):
        """Row i = O e_i through the basis-element path (to_matrix)."""
        M = np.zeros((self.dim, self.dim))
        for d, y0 in enumerate(self.bdn):
            for u, x0 in enumerate(self.bup):
                i = d * self.nsup + u
                for coeff, factors in terms:
                    if not abs(coeff) > 1e-14:
                        continue
                    st = (int(x0), int(y0), 1.0)
                    for f in factors:
                        st = self._walk(f, *st)
                        if st is None or st[0] not in self.pu or st[1] not in self.pd:
                            st = None
                            break
                    if st is not None:
                        M[i, self.pd[st[1]] * self.nsup + self.pu[st[0]]] += coeff * st[2]
        return M
