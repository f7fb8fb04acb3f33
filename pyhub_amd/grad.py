"""Differentiable ground-state quantities of Hubbard clusters.

No counterpart in the reference (it has no autograd); the derivatives are pinned
by the reference's own contractions (ONE_RDM / TWO_RDM, static_rq.f90) and checked
against central finite differences of the engine's energies and RDMs.

* Energy: Hellmann-Feynman, one pass of libphub.so over psi (phub_hf_grad_hubbard)
      dE0/dT(J,K) = sum_s <c^dag_Ks c_Js>,   dE0/dU_J = <n_Ju n_Jd>,
  mapped to bond amplitudes t_b (T = -t_b on both ordered pairs of bond b) and
  on-site energies eps_J = T(J,J).
* 1-RDM loss L = sum_s ||gamma_s - target_s||_F^2, gamma_s[J][K] = <c^dag_Js c_Ks>:
  adjoint method.  With G_s = dL/dgamma_s = 2 (gamma_s - target_s) (symmetric),
      g = dL/dpsi = 2 sum_s sum_JK G_s[J][K] c^dag_Js c_Ks psi
  is one H*v of an operator with T_up = G_up, T_dn = G_dn, U = 0 (same kernels);
      (H - E0) lam = P_perp g      (MINRES on the complement of psi, 7 work vectors),
      dL/dtheta = -<lam| dH/dtheta |psi>      (phub_grad_hubbard).
  No Krylov basis is stored for the adjoint.

`E0Function` / `RDMLossFunction` wrap both as torch.autograd.Functions over the
parameter tensors (t, eps, U).

* Heisenberg energy (hprod_heis.f90 operator): Hellmann-Feynman over psi,
      dE0/dJ(J,K) = <dH/dJ(J,K)>  (phub_hf_grad_heisenberg, ordered pairs),
  mapped to bond couplings J_b with J(i,j) = J(j,i) = scale * J_b (scale 1/2 in the
  standard convention H = sum_b J_b S_i.S_j, 1 in the reference's ordered-pair
  one).  In the physical sign convention dE0/dJ_b = <S_i.S_j> of the reference's
  J_SPIN (static_rq_heis.f90:58-116); in the reference's flip sign it differs on
  frustrated bond graphs.  `HeisenbergE0Function` wraps it.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch

from pyhub_amd import _lib as L
from pyhub_amd.engine import Context, HeisenbergOp, HubbardOp


@dataclass
class ClusterSpec:
    """Cluster geometry: n sites, bonds (i, j) 0-based, particle numbers."""
    n: int
    bonds: list
    n_up: int
    n_dn: int
    lanczos_tol: float = 1e-13
    max_steps: int = 400
    minres_tol: float = 1e-12
    # the reference's three-term recurrence, psi combined from the kept vectors; at
    # these tolerances its ground state and RDMs equal the fully reorthogonalised
    # run's (LCZ_FULL64) to rounding, at a fraction of the cost
    lanczos_mode: int = L.LCZ_PLAIN64

    def T(self, t, eps):
        T = np.zeros((self.n, self.n))
        for b, (i, j) in enumerate(self.bonds):
            T[i, j] = T[j, i] = -float(t[b])
        T[np.arange(self.n), np.arange(self.n)] = np.asarray(eps, float)
        return T

    def U_vec(self, U):
        U = np.asarray(U, float)
        return np.full(self.n, float(U)) if U.ndim == 0 else U


def solve(spec: ClusterSpec, t, eps, U, ctx: Context | None = None):
    """FP64 Lanczos (spec.lanczos_mode); returns (op, E0, psi) with E0 the Rayleigh
    quotient of the reconstructed psi."""
    ctx = ctx or Context.default()
    op = HubbardOp(ctx, spec.n, spec.n_up, spec.n_dn, spec.T(t, eps), spec.U_vec(U))
    r = op.lanczos(mode=spec.lanczos_mode, max_steps=min(spec.max_steps, op.dim),
                   tol=spec.lanczos_tol, check_every=5, seed=12345)
    w = op.hv(r.psi)
    e0 = float(torch.dot(r.psi, w))
    return op, e0, r.psi


def _param_grads(spec: ClusterSpec, gT: np.ndarray, gU: np.ndarray, scalar_u: bool):
    """Map d/dT(J,K) and d/dU_J to the parameters (t_b, eps_J, U)."""
    gt = np.array([-(gT[i, j] + gT[j, i]) for (i, j) in spec.bonds])
    ge = np.diag(gT).copy()
    gu = float(gU.sum()) if scalar_u else gU.copy()
    return gt, ge, gu


def hellmann_feynman(spec: ClusterSpec, t, eps, U, ctx=None):
    """E0 and (dE/dt_b, dE/deps_J, dE/dU) from one pass over psi."""
    op, e0, psi = solve(spec, t, eps, U, ctx)
    gT, gU = op.hf_grad(psi)
    gt, ge, gu = _param_grads(spec, gT, gU, np.ndim(U) == 0)
    return e0, gt, ge, gu


def rdm_loss_and_grad(spec: ClusterSpec, t, eps, U, target_up, target_dn, ctx=None,
                      want_energy_grad=False):
    """L = sum_s ||gamma_s - target_s||_F^2 and dL/d(t, eps, U) by the adjoint solve.
    With want_energy_grad, info["energy_grad"] holds (dE/dt, dE/deps, dE/dU) of the same
    ground state (Hellmann-Feynman, one more pass over psi)."""
    ctx = ctx or Context.default()
    op, e0, psi = solve(spec, t, eps, U, ctx)
    n = spec.n
    gam_up, gam_dn = op.one_rdm_spin(psi)
    tu, td = np.asarray(target_up, float), np.asarray(target_dn, float)
    L_val = float(np.sum((gam_up - tu) ** 2) + np.sum((gam_dn - td) ** 2))
    Gu = 2.0 * (gam_up - tu)
    Gd = 2.0 * (gam_dn - td)
    Gu = 0.5 * (Gu + Gu.T)
    Gd = 0.5 * (Gd + Gd.T)
    gop = HubbardOp(ctx, n, spec.n_up, spec.n_dn, Gu, np.zeros(n), T_dn=Gd)
    g = 2.0 * gop.hv(psi)
    lam, iters, rr = op.minres(psi, e0, g, max_iter=3000, tol=spec.minres_tol)
    gT, gU = op.grad(lam, psi)
    gt, ge, gu = _param_grads(spec, -gT, -gU, np.ndim(U) == 0)
    info = {"minres_iters": iters, "minres_rel_res": rr, "e0": e0,
            "gamma_up": gam_up, "gamma_dn": gam_dn}
    if want_energy_grad:
        hT, hU = op.hf_grad(psi)
        info["energy_grad"] = _param_grads(spec, hT, hU, np.ndim(U) == 0)
    return L_val, gt, ge, gu, info


def run_sets(fn, items, nthreads: int = 8, device: int = 0):
    """fn(ctx, item) for every item; returns the results in item order.

    The items are dealt round-robin to `nthreads` host threads. Each thread makes
    a torch stream and a Context on it, and runs its items inside that stream.
    Parameter sets are independent, and one set is a chain of small launches with
    a host round trip per Krylov or MINRES step, so several sets in flight keep
    the GPU busy where one leaves it idle between steps. ctypes releases the GIL
    during every library call. After every thread has ended, the first exception
    a thread raised is raised here."""
    import threading
    items = list(items)
    nth = max(1, min(int(nthreads), len(items)))
    L.load()   # load the library once, before the threads
    out = [None] * len(items)
    errs = []

    def work(k):
        try:
            s = torch.cuda.Stream(device=device)
            ctx = Context(device, stream=s)
            with torch.cuda.stream(s):
                for i in range(k, len(items), nth):
                    out[i] = fn(ctx, items[i])
            s.synchronize()
        except BaseException as e:   # re-raised by the caller's thread
            errs.append(e)

    if nth == 1:
        work(0)
    else:
        th = [threading.Thread(target=work, args=(k,)) for k in range(nth)]
        for t in th:
            t.start()
        for t in th:
            t.join()
    if errs:
        raise errs[0]
    return out


class E0Function(torch.autograd.Function):
    """E0(t, eps, U) with the Hellmann-Feynman backward."""

    @staticmethod
    def forward(ctx, t, eps, U, spec):
        e0, gt, ge, gu = hellmann_feynman(spec, t.detach().cpu().numpy(),
                                          eps.detach().cpu().numpy(), U.detach().cpu().numpy())
        ctx.save_for_backward(torch.as_tensor(gt), torch.as_tensor(ge),
                              torch.as_tensor(np.asarray(gu, float)))
        return torch.tensor(e0, dtype=torch.float64)

    @staticmethod
    def backward(ctx, go):
        gt, ge, gu = ctx.saved_tensors
        return go * gt, go * ge, go * gu, None


class RDMLossFunction(torch.autograd.Function):
    """L(t, eps, U) = sum_s ||gamma_s - target_s||^2 with the adjoint backward."""

    @staticmethod
    def forward(ctx, t, eps, U, spec, target_up, target_dn):
        L_val, gt, ge, gu, _ = rdm_loss_and_grad(spec, t.detach().cpu().numpy(),
                                                 eps.detach().cpu().numpy(),
                                                 U.detach().cpu().numpy(), target_up, target_dn)
        ctx.save_for_backward(torch.as_tensor(gt), torch.as_tensor(ge),
                              torch.as_tensor(np.asarray(gu, float)))
        return torch.tensor(L_val, dtype=torch.float64)

    @staticmethod
    def backward(ctx, go):
        gt, ge, gu = ctx.saved_tensors
        return go * gt, go * ge, go * gu, None, None, None


def ground_energy(t, eps, U, spec: ClusterSpec):
    return E0Function.apply(t, eps, U, spec)


def rdm_loss(t, eps, U, spec: ClusterSpec, target_up, target_dn):
    return RDMLossFunction.apply(t, eps, U, spec, target_up, target_dn)


def lattice_bonds(Lx, Ly, periodic=False):
    """Nearest-neighbour bonds of an Lx x Ly lattice, site s = x + Lx*y.  Each
    unordered bond appears once: on a periodic side of length 2 the wrap-around
    bond is the bond already listed, and on a side of length 1 it would join a
    site to itself; neither is added again (every listed bond owns its own
    matrix entries, so its gradient is the whole derivative of that entry)."""
    b, seen = [], set()

    def add(i, j):
        key = (min(i, j), max(i, j))
        if i != j and key not in seen:
            seen.add(key)
            b.append((i, j))

    for y in range(Ly):
        for x in range(Lx):
            s = x + Lx * y
            if x + 1 < Lx or periodic:
                add(s, (x + 1) % Lx + Lx * y)
            if y + 1 < Ly or periodic:
                add(s, x + Lx * ((y + 1) % Ly))
    return b


@dataclass
class HeisenbergSpec:
    """Heisenberg cluster: n sites (even), bonds (i, j) 0-based.  convention
    "standard": H = sum_b J_b S_i.S_j, the physical flip sign (J(i,j) = J(j,i) =
    J_b / 2, sign_convention 1); "reference": the reference's ordered-pair matrix
    and flip sign (J(i,j) = J(j,i) = J_b, sign_convention 0)."""
    n: int
    bonds: list
    convention: str = "standard"
    lanczos_tol: float = 1e-13
    max_steps: int = 400

    @property
    def scale(self):
        return 0.5 if self.convention == "standard" else 1.0

    @property
    def sign_convention(self):
        return 1 if self.convention == "standard" else 0

    def Jm(self, Jb):
        J = np.zeros((self.n, self.n))
        for b, (i, j) in enumerate(self.bonds):
            J[i, j] += self.scale * float(Jb[b])
            J[j, i] += self.scale * float(Jb[b])
        return J


def heisenberg_hellmann_feynman(spec: HeisenbergSpec, Jb, ctx: Context | None = None):
    """E0 and dE0/dJ_b from one pass over psi (FP64 Lanczos, full reorthogonalisation)."""
    ctx = ctx or Context.default()
    op = HeisenbergOp(ctx, spec.n, spec.Jm(Jb), sign_convention=spec.sign_convention)
    r = op.lanczos(mode=L.LCZ_FULL64, max_steps=min(spec.max_steps, op.dim),
                   tol=spec.lanczos_tol, check_every=5, seed=12345)
    e0 = float(torch.dot(r.psi, op.hv(r.psi)))
    g = op.hf_grad(r.psi)
    gb = np.array([spec.scale * (g[i, j] + g[j, i]) for (i, j) in spec.bonds])
    return e0, gb


class HeisenbergE0Function(torch.autograd.Function):
    """E0(J_b) of a Heisenberg cluster with the Hellmann-Feynman backward."""

    @staticmethod
    def forward(ctx, Jb, spec):
        e0, gb = heisenberg_hellmann_feynman(spec, Jb.detach().cpu().numpy())
        ctx.save_for_backward(torch.as_tensor(gb))
        return torch.tensor(e0, dtype=torch.float64)

    @staticmethod
    def backward(ctx, go):
        (gb,) = ctx.saved_tensors
        return go * gb, None


def heisenberg_energy(Jb, spec: HeisenbergSpec):
    return HeisenbergE0Function.apply(Jb, spec)
