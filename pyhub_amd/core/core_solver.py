"""CoreSolver with PyHub's API (mirror of pyhub/core/core_solver.py).

The reference writes solver.h5, spawns hubbard.x / heisenberg.x with os.system
and reads every result back from HDF5 (core_solver.py:31-249).  Here the same
stages are in-process calls into libphub.so: the operator, the Krylov vectors
and the ground vector stay on the GPU; results are kept on the solver object.

Selection as in SOLVE (solve.f90:15-19): dense diagonalisation when
nstates < max_lcz (EXACT_DIAG, exact_diag.f90:2-33: H from H*v on unit vectors,
then a symmetric eigensolver), Lanczos otherwise.
"""
from __future__ import annotations

from time import perf_counter as pc

import numpy as np
import torch

from pyhub_amd import _lib as L
from pyhub_amd.core.basis import Basis
from pyhub_amd.core.reduced_quantities import GreenFunction, StaticQuantities
from pyhub_amd.engine import Context, HeisenbergOp, HubbardOp

_MODES = {"plain64": L.LCZ_PLAIN64, "full64": L.LCZ_FULL64, "mixed": L.LCZ_MIXED,
          "plain32": L.LCZ_PLAIN32}
# relative distance within which a Ritz value of the three-term recurrence is a copy of
# a lower one (FP64 / FP32 vectors); the full-reorthogonalisation modes keep repeats
_GHOST_TOL = {"plain64": 1e-8, "plain32": 1e-5}


def _genuine_ritz(val, u1, tol):
    """Positions of the genuine values among the ascending Ritz values of the three-term
    recurrence (u1: first components of their tridiagonal eigenvectors).  Without
    reorthogonalisation, rounding adds copies of converged values (Paige's ghosts) and,
    while a copy forms, values with no weight on the start vector (Cullum and
    Willoughby's spurious values); either would enter the spectrum as an excited state.
    Values within tol (relative) of a cluster's first member form one cluster, kept once
    through its member of largest |u1|; a value alone with |u1| < 1e-10 is dropped.  The
    lowest value is always kept as the ground state.  The same rule as the engine's
    excited-state selection."""
    keep, i, m = [], 0, len(val)
    while i < m:
        j, best = i + 1, i
        while j < m and val[j] - val[i] <= tol * max(1.0, abs(val[j])):
            if abs(u1[j]) > abs(u1[best]):
                best = j
            j += 1
        if i == 0:
            best = 0       # the ground state: the lowest value, as the engine reports it
        if i == 0 or j - i > 1 or abs(u1[i]) >= 1e-10:
            keep.append(best)
        i = j
    return np.asarray(keep, dtype=int)


def boltzmann_weights(val, T, nb_comp_states):
    """BOLTZMANN (boltzmann.f90:1-29): weights exp((E1-Ei)/T) above 1e-8; T=0 -> (1,0,...)."""
    zero = 1.0e-8
    nval = len(val)
    nw = nval
    for i in range(1, nval):
        with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
            x = np.exp((val[0] - val[i]) / T) if T != 0 else (0.0 if val[i] != val[0] else np.nan)
        if x <= zero:
            nw = i
            break
    if nw < nb_comp_states:
        nw = nb_comp_states
    nw = min(nw, nval)
    if T >= zero:
        w = np.exp((val[0] - val[:nw]) / T)
    else:
        w = np.zeros(nw)
        w[0] = 1.0
    return w


class CoreSolver(Basis, GreenFunction, StaticQuantities):
    def __init__(self, nb_sites, n_up: int, n_down: int, exec_name, T=0.0, order="spin"):
        Basis.__init__(self, nb_sites, hilbert=(n_up, n_down), order=order)
        self.T = T
        self.exec_name = exec_name
        self.t_matrix = np.zeros((nb_sites, nb_sites))
        self.J_matrix = np.zeros((nb_sites, nb_sites))
        self.U = 0.0
        self.is_solve = False
        self.is_reduced_quantities = False
        self.is_spgf = False
        self.is_tprf = False
        self.excited_state = 0
        self.lanczos_mode = "plain64"
        self.start_vector = None
        self._op = None
        self._res = {}

    # ---- operator -----------------------------------------------------------
    def _model_op(self, n_up=None, n_down=None):
        ctx = Context.default()
        nu = self.n_up[0] if n_up is None else n_up
        nd = self.n_down[0] if n_down is None else n_down
        if self.exec_name == "heisenberg.x":
            return HeisenbergOp(ctx, self.nb_sites, self.J_matrix,
                                sign_convention=getattr(self, "sign_convention", 0))
        return HubbardOp(ctx, self.nb_sites, nu, nd, self.t_matrix, self.U_matrix)

    def _validate(self):
        if isinstance(self.U, (float, int, np.floating, np.integer)):
            pass
        elif isinstance(self.U, (np.ndarray, list)):
            U = np.asarray(self.U, dtype=float)
            if U.ndim == 4:
                if not np.isclose(U, np.einsum("ijkl->jilk", U)).all():
                    raise ValueError("U tensor must be symmetric")
            elif U.ndim != 1:
                raise ValueError("U must be a float or a np.ndarray of size (nb_sites,) or "
                                 f"(nb_sites,)*4, not {type(self.U)}")
        else:
            raise ValueError("U must be a float or a np.ndarray of size (nb_sites,) or "
                             f"(nb_sites,)*4, not {type(self.U)}")

    @staticmethod
    def _dense_h(op):
        """EXACT_DIAG's H: one H*v per unit vector (exact_diag.f90:14-19)."""
        D = op.dim
        I = torch.eye(D, dtype=torch.float64, device=op.ctx.device)
        H = torch.empty((D, D), dtype=torch.float64, device=op.ctx.device)
        for i in range(D):
            H[:, i] = op.hv(I[:, i].contiguous())
        return H

    # ---- stages ---------------------------------------------------------------
    def kernel(self, max_lcz: int = 600, acc_lcz: float = 1.0e-8, nb_comp_states=1,
               compute_rq=True, compute_spgf=False, nb_sites_comp=None, store_H=False,
               verbose=False):
        self.max_lcz = max_lcz
        self.acc_lcz = acc_lcz
        self.compute_rq = compute_rq
        self.nb_comp_states_ = nb_comp_states
        self.compute_spgf = compute_spgf
        self.compute_tprf = False
        self.store_H = store_H
        self.verbose = verbose
        self.nb_sites_comp = self.nb_sites if nb_sites_comp is None else nb_sites_comp
        self._validate()
        tall = pc()
        self.printv("start calculation")
        self._solve()
        self.wgt = self.boltzmann.copy()
        if compute_rq:
            self.reduced_quantities()
        if compute_spgf:
            self.spgf(nb_sites_comp=self.nb_sites_comp)
        self.time = pc() - tall
        self.printv(f"total elapsed time {np.around(self.time, 4)} s\n")

    def _solve(self):
        op = self._model_op()
        self._op = op
        D = op.dim
        r = {}
        if D < self.max_lcz:
            H = self._dense_h(op)
            val, vec = torch.linalg.eigh(H)
            val = val.cpu().numpy()
            ndeg = 1
            for i in range(1, D):
                if abs(val[i] - val[0]) < 1.0e-14:
                    ndeg += 1
                else:
                    break
            r.update(is_lanczos=False, e=val, vectors=vec, ndeg=ndeg,
                     H=H.cpu().numpy() if self.store_H else None)
        else:
            v0 = None
            if self.start_vector is not None:
                v0 = torch.as_tensor(np.asarray(self.start_vector, float), device=op.ctx.device)
            from scipy.linalg import eigh_tridiagonal
            # The Ritz vectors of every state with a Boltzmann weight are rebuilt from
            # the Lanczos vectors (READ_PSI_HUBBARD over NW states).  NW is known only
            # after the run: when T > 0 lifts it above the states rebuilt, the
            # recurrence is repeated (same start vector, same alpha/beta) for NW states.
            nst = max(1, self.nb_comp_states_)
            while True:
                if nst > L.MAX_STATES:
                    raise NotImplementedError(
                        f"{nst} Lanczos states carry a Boltzmann weight; at most "
                        f"{L.MAX_STATES} are rebuilt (lower T or use the dense path)")
                out = op.lanczos(mode=_MODES[self.lanczos_mode], max_steps=self.max_lcz,
                                 tol=self.acc_lcz, v0=v0, refine_steps=2, want_psi=True,
                                 n_states=nst)
                if out.norm_lost:
                    # the reference prints this and keeps its results (lanczos.f90:74-76)
                    print(" FORTRAN LANCZOS ERROR: NORM LOST", flush=True)
                m = len(out.alpha)
                val, vec = eigh_tridiagonal(out.alpha, out.beta[1:m]) if m > 1 else (out.alpha.copy(), np.ones((1, 1)))
                if self.lanczos_mode in _GHOST_TOL:
                    keep = _genuine_ritz(val, vec[0], _GHOST_TOL[self.lanczos_mode])
                    val, vec = val[keep], vec[:, keep]
                nw = len(boltzmann_weights(val, self.T, self.nb_comp_states_))
                if nw <= nst or nst >= m:
                    break
                nst = nw
            if self.lanczos_mode == "mixed":
                val = val.copy()
                val[0] = out.e0
            r.update(is_lanczos=True, e=val, tri_vectors=vec, psi0_dev=out.psi, ndeg=out.ndeg,
                     psis_dev=out.psis,
                     alpha=out.alpha, beta=out.beta, steps=out.steps, converged=out.converged,
                     conv=out.e0 - out.e0_ritz, residual=out.residual, lanczos=out,
                     H=(np.diag(out.alpha) + np.diag(out.beta[1:m], 1) + np.diag(out.beta[1:m], -1))
                     if self.store_H else None)
        r["boltzmann"] = boltzmann_weights(r["e"], self.T, self.nb_comp_states_)
        self._res = r
        self.is_solve = True

    def set_state(self, excited_state=-1):
        self.excited_state = excited_state
        self.wgt = self.boltzmann.copy()
        if excited_state != -1:
            self.wgt[:] = 0.0
            self.wgt[excited_state] = 1.0

    def _state_vectors(self):
        """Device vectors of the Boltzmann-weighted states (READ_PSI_HUBBARD)."""
        r = self._res
        nw = len(self.wgt)
        if not r["is_lanczos"]:
            return [r["vectors"][:, b].contiguous() for b in range(nw)]
        # Ritz vectors of the lowest states (READ_PSI_HUBBARD, compute_solver.f90:271-285)
        if r.get("psis_dev") is not None:
            ps = r["psis_dev"]
            return [ps[b] if b < ps.shape[0] else None for b in range(nw)]
        return [r["psi0_dev"]] + [None] * (nw - 1)

    def reduced_quantities(self):
        t0 = pc()
        self.printv("start reduced quantities calculation")
        op = self._op
        n = self.nb_sites
        vecs = self._state_vectors()
        nw = len(vecs)
        rq = {"sisj": np.zeros((nw, n, n))}
        # the 4-index DBL_OCC is computed when U is a 4-index tensor (do_rq_two_body,
        # core_solver.py:147)
        two_body = self.exec_name != "heisenberg.x" and self.U_matrix.ndim == 4
        if self.exec_name != "heisenberg.x":
            rq.update(density_matrix_up=np.zeros((nw, n, n)), density_matrix_down=np.zeros((nw, n, n)),
                      ni=np.zeros((nw, n)))
        if two_body:
            rq["dbl_occ"] = np.zeros((nw, n, n, n, n))
        for b, v in enumerate(vecs):
            if v is None:
                continue
            if self.exec_name == "heisenberg.x":
                rq["sisj"][b] = op.sisj(v)
            else:
                up, dn, ni, ss = op.rdm(v)
                if self.n_up[0] == self.n_down[0]:
                    dn = up.copy()      # static_rq.f90:212-241 copies the up matrix
                rq["density_matrix_up"][b] = up
                rq["density_matrix_down"][b] = dn
                rq["ni"][b] = ni
                rq["sisj"][b] = ss
                if two_body:
                    rq["dbl_occ"][b] = op.two_rdm(v)
        self._rq = rq
        self.is_reduced_quantities = True
        self.printv(f"end calculation | elapsed time {np.around(pc() - t0, 4)} s")

    def spgf(self, nb_sites_comp=None, method=None):
        """Green's functions of the first nb_sites_comp sites (core_solver.py:81-101).
        Lanczos path: method "band" (default, SPGF_BANDLANCZOS: full G_ij) or "chains"
        (one continued fraction per site: diagonal G_ii only, cheaper)."""
        self.nb_sites_comp = self.nb_sites if nb_sites_comp is None else nb_sites_comp
        t0 = pc()
        self.printv("start spgf calculation")
        if self.exec_name == "heisenberg.x":
            raise NotImplementedError("Green's functions are defined for the Hubbard model")
        if self._res["is_lanczos"]:
            if (method or "band") == "band":
                self._spgf_band()
            elif method == "chains":
                self._spgf_chains()
            else:
                raise ValueError("method must be 'band' or 'chains'")
        else:
            self._spgf_ed()
        self.is_spgf = True
        self.printv(f"end calculation | elapsed time {np.around(pc() - t0, 4)} s")
        self.gf = self.compute_gf()

    def _spgf_ed(self):
        """SPGF_ED (spgf_ed.f90:1-298): full ED of the N+-1 sectors, Q = VEC^T c psi."""
        op = self._op
        vecs = self._state_vectors()
        nw = len(vecs)
        ncomp = self.nb_sites_comp
        out = {}
        nb_poles = np.zeros((2, 2), dtype=int)
        for si, spin in enumerate(("up", "down")):
            if spin == "down" and self.n_up[0] == self.n_down[0]:
                for kind in ("lesser", "greater"):
                    out[f"positions_{kind}_down"] = out[f"positions_{kind}_up"]
                    out[f"Q_{kind}_down"] = out[f"Q_{kind}_up"]
                nb_poles[:, 1] = nb_poles[:, 0]
                break
            for ki, (kind, create) in enumerate((("lesser", False), ("greater", True))):
                nu = self.n_up[0] + ((1 if create else -1) if si == 0 else 0)
                nd = self.n_down[0] + ((1 if create else -1) if si == 1 else 0)
                if not (0 <= nu <= self.nb_sites and 0 <= nd <= self.nb_sites):
                    out[f"positions_{kind}_{spin}"] = np.zeros(0)
                    out[f"Q_{kind}_{spin}"] = np.zeros((ncomp, 0, nw))
                    continue
                tgt = self._model_op(nu, nd)
                H = self._dense_h(tgt)
                val, vec = torch.linalg.eigh(H)
                Q = np.zeros((ncomp, tgt.dim, nw))
                for b, v in enumerate(vecs):
                    for i in range(ncomp):
                        _, x = op.excite(v, i + 1, si, create, target=tgt)
                        Q[i, :, b] = (vec.T @ x).cpu().numpy()
                out[f"positions_{kind}_{spin}"] = val.cpu().numpy()
                out[f"Q_{kind}_{spin}"] = Q
                nb_poles[ki, si] = tgt.dim
        out["nb_poles"] = nb_poles
        self._gfdata = out

    def _spgf_band(self):
        """SPGF_BANDLANCZOS (spgf_bandlanczos.f90:1-300): for each spin and kind the
        nb_sites_comp excitations c_i psi0 (lesser) / c^dag_i psi0 (greater) start a band
        Lanczos in the N-1 / N+1 sector; positions are the band matrix eigenvalues and
        Q[i, k] the weights of pole k for site i.  The down spin reuses the up results
        when N_up = N_dn, as the reference does."""
        if np.count_nonzero(self.wgt) > 1:
            raise NotImplementedError("band Lanczos Green's functions are ground-state only "
                                      "(spgf_bandlanczos.f90:63)")
        op = self._op
        psi = self._state_vectors()[0]
        ncomp = self.nb_sites_comp
        out = {}
        nb_poles = np.zeros((2, 2), dtype=int)
        for si, spin in enumerate(("up", "down")):
            if spin == "down" and self.n_up[0] == self.n_down[0]:
                for kind in ("lesser", "greater"):
                    out[f"positions_{kind}_down"] = out[f"positions_{kind}_up"]
                    out[f"Q_{kind}_down"] = out[f"Q_{kind}_up"]
                nb_poles[:, 1] = nb_poles[:, 0]
                break
            for ki, (kind, create) in enumerate((("lesser", False), ("greater", True))):
                nu = self.n_up[0] + ((1 if create else -1) if si == 0 else 0)
                nd = self.n_down[0] + ((1 if create else -1) if si == 1 else 0)
                if not (0 <= nu <= self.nb_sites and 0 <= nd <= self.nb_sites):
                    out[f"positions_{kind}_{spin}"] = np.zeros(0)
                    out[f"Q_{kind}_{spin}"] = np.zeros((ncomp, 0, 1))
                    continue
                tgt = self._model_op(nu, nd)
                V = torch.stack([op.excite(psi, i + 1, si, create, target=tgt)[1]
                                 for i in range(ncomp)])
                pos, Q = tgt.band_lanczos(V, max_steps=self.max_lcz, tol=self.acc_lcz)
                out[f"positions_{kind}_{spin}"] = pos
                out[f"Q_{kind}_{spin}"] = Q.T[:, :, None]
                nb_poles[ki, si] = len(pos)
        out["nb_poles"] = nb_poles
        self._gfdata = out

    def _spgf_chains(self, nsteps=None):
        """Diagonal G_ii from continued-fraction Lanczos chains (the single-vector
        case of GF_BANDLANCZOS, spgf_bandlanczos.f90:309-448).  Each chain gives its
        own poles (Ritz values) and weights Q_ik = |c psi| y_k[0]."""
        from scipy.linalg import eigh_tridiagonal
        if np.count_nonzero(self.wgt) > 1:
            raise NotImplementedError("Lanczos-chain Green's functions are ground-state only "
                                      "(spgf_bandlanczos.f90:63); use the dense path at T > 0")
        op = self._op
        psi = self._state_vectors()[0]
        nsteps = nsteps or min(self.max_lcz, 200)
        out = {"chains": {}}
        for si, spin in enumerate(("up", "down")):
            for kind, create in (("lesser", False), ("greater", True)):
                nu = self.n_up[0] + ((1 if create else -1) if si == 0 else 0)
                nd = self.n_down[0] + ((1 if create else -1) if si == 1 else 0)
                chains = []
                if 0 <= nu <= self.nb_sites and 0 <= nd <= self.nb_sites:
                    tgt = self._model_op(nu, nd)
                    for i in range(self.nb_sites_comp):
                        _, x = op.excite(psi, i + 1, si, create, target=tgt)
                        a, b, n2 = tgt.cf_chain(x, min(nsteps, tgt.dim), precision=64)
                        m = int(np.count_nonzero(b)) + 1 if n2 > 0 else 0
                        m = min(m, len(a))
                        if m == 0:
                            chains.append((np.zeros(0), np.zeros(0), a, b, n2))
                            continue
                        th, y = eigh_tridiagonal(a[:m], b[: m - 1]) if m > 1 else (a[:1], np.ones((1, 1)))
                        chains.append((th, np.sqrt(n2) * y[0, :], a, b, n2))
                out["chains"][f"{kind}_{spin}"] = chains
        self._gfdata = out

    def printv(self, *args):
        if getattr(self, "verbose", False):
            print(*args)

    # ---- results (core_solver.py:165-249) ---------------------------------------
    @property
    def filename(self):
        return None

    @property
    def boltzmann(self):
        return np.array(self._res["boltzmann"], dtype=float)

    @property
    def Z(self):
        return np.sum(self.wgt)

    @property
    def nb_deg(self):
        return self._res["ndeg"]

    @property
    def H(self):
        if self._res.get("H") is None:
            raise KeyError("H was not stored (kernel(store_H=True))")
        return self._res["H"]

    @property
    def e(self):
        return np.array(self._res["e"])

    @property
    def etot(self):
        return self.e0

    @property
    def e0(self):
        return self.e[: self.nb_comp_states] @ self.wgt / self.Z

    @property
    def lcz_results(self):
        r = self._res
        if not r.get("is_lanczos"):
            raise KeyError("the dense path does not run Lanczos")
        lcz = {"alpha": r["alpha"], "beta": r["beta"][: r["steps"]], "conv": r["conv"],
               "step": r["steps"]}
        if not r["converged"]:
            raise StopIteration("Lanczos not converged, step,", lcz["step"],
                                f" greater than the maximum wanted {self.max_lcz}")
        return lcz

    @property
    def nb_lcz(self):
        if not self._res.get("is_lanczos"):
            raise KeyError("the dense path does not run Lanczos")
        return self._res["steps"]

    @property
    def lcz_vectors(self):
        raise NotImplementedError("Krylov vectors stay on the GPU and are not stored; use psi0")

    @property
    def nb_comp_states(self):
        return len(self.boltzmann)

    @property
    def psi0(self):
        r = self._res
        w = self.boltzmann / np.sum(self.boltzmann)
        if not r["is_lanczos"]:
            V = r["vectors"][:, : len(w)].cpu().numpy()
            return V @ w
        ps = r.get("psis_dev")
        if ps is not None and len(w) > 1:
            return w @ ps[: len(w)].cpu().numpy()
        return r["psi0_dev"].cpu().numpy()

    def psi(self, vector: int):
        r = self._res
        if not r["is_lanczos"]:
            return r["vectors"][:, vector].cpu().numpy()
        ps = r.get("psis_dev")
        if ps is None or not (0 <= vector < ps.shape[0]):
            raise ValueError(f"Lanczos state {vector} was not reconstructed "
                             f"(nb_comp_states = {self.nb_comp_states_})")
        return ps[vector].cpu().numpy()

    def psi_as_dict(self, vector: int):
        psi = self.psi(vector)
        index = np.where(abs(psi) > 1.0e-12)[0]
        return {"basis": self.basis[index], "coeff": psi[index]}

    @property
    def is_lanczos(self):
        return bool(self._res.get("is_lanczos", False))

    @property
    def U_matrix(self):
        if isinstance(self.U, np.ndarray):
            if len(self.U.shape) in (1, 4):
                return self.U
            raise ValueError
        if isinstance(self.U, (list, tuple)):
            return np.asarray(self.U, dtype=float)
        return np.full(self.nb_sites, float(self.U), dtype=float)
