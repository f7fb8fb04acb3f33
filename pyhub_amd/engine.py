"""Device-side engine objects over the C ABI (include/phub.h).

`Context` owns a phub_ctx bound to a torch device and its current stream;
`HubbardOp` / `HeisenbergOp` own a phub_op (basis, rank and hop tables on the
device).  Vectors are torch tensors on the context's device; the C ABI only
sees their data pointers.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib as L


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Context:
    """A phub_ctx on one GPU (one process per GPU; default stream of torch)."""

    _default = None

    def __init__(self, device: int = 0, stream=None):
        if not torch.cuda.is_available():
            raise RuntimeError("pyhub_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
        self.lib = L.load()
        self.device = torch.device("cuda", device)
        torch.cuda.set_device(self.device)
        s = stream if stream is not None else torch.cuda.current_stream(self.device)
        self.stream = s
        h = C.c_void_p()
        rc = self.lib.phub_ctx_create(device, C.c_void_p(s.cuda_stream), C.byref(h))
        self.h = h
        L.check(h, rc)

    @classmethod
    def default(cls) -> "Context":
        if cls._default is None:
            cls._default = Context(0)
        return cls._default

    def check(self, rc):
        L.check(self.h, rc)

    def sync(self):
        self.check(self.lib.phub_sync(self.h))

    def last_kernel_ms(self) -> float:
        return float(self.lib.phub_last_kernel_ms(self.h))

    def __del__(self):
        try:
            if getattr(self, "h", None) and self.h.value:
                self.lib.phub_ctx_destroy(self.h)
        except Exception:
            pass

    # -- basis ---------------------------------------------------------------
    def basis(self, n: int, k: int) -> torch.Tensor:
        """Masks of the n-site, k-particle basis in the reference order (uint64 as int64)."""
        cnt = int(self.lib.phub_binom(n, k))
        out = torch.empty(cnt, dtype=torch.int64, device=self.device)
        if cnt:
            self.check(self.lib.phub_basis_build(self.h, n, k, C.c_void_p(out.data_ptr())))
        return out

    def rank(self, n: int, k: int, masks: torch.Tensor) -> torch.Tensor:
        masks = masks.to(self.device, torch.int64).contiguous()
        out = torch.empty_like(masks)
        self.check(self.lib.phub_basis_rank(self.h, n, k, C.c_void_p(masks.data_ptr()),
                                            masks.numel(), C.c_void_p(out.data_ptr())))
        return out


@dataclass
class LanczosOutput:
    e0: float
    e0_ritz: float
    residual: float
    steps: int
    converged: bool
    ndeg: int
    seconds: float
    seconds_lanczos: float
    hv_ms_avg: float
    alpha: np.ndarray
    beta: np.ndarray
    psi: torch.Tensor | None
    psis: torch.Tensor | None = None        # (n_states, dim) when n_states > 1
    e_states: np.ndarray | None = None      # lowest Ritz values
    norm_lost: bool = False                 # lanczos.f90:74-76 NORM LOST (psi renormalised)
    corrections: int = 0                    # FP64 corrections of psi (MIXED / PLAIN32)
    correction_iters: int = 0               # their inner MINRES iterations


class _Op:
    ctx: Context
    h: C.c_void_p

    def _info(self):
        info = L.OpInfo()
        self.ctx.check(self.ctx.lib.phub_op_get_info(self.h, C.byref(info)))
        self.info = info
        self.dim = int(info.dim)
        self.dim_up = int(info.dim_up)
        self.dim_dn = int(info.dim_dn)

    def __del__(self):
        try:
            if getattr(self, "h", None) and self.h.value:
                self.ctx.lib.phub_op_destroy(self.h)
        except Exception:
            pass

    def vec(self, dtype=torch.float64) -> torch.Tensor:
        return torch.empty(self.dim, dtype=dtype, device=self.ctx.device)

    def hv(self, v: torch.Tensor, out: torch.Tensor | None = None) -> torch.Tensor:
        """w = H v on the device (FP64 or FP32 by v's dtype)."""
        if v.device != self.ctx.device or v.numel() != self.dim or not v.is_contiguous():
            raise ValueError("v must be a contiguous device vector of the operator dimension")
        w = torch.empty_like(v) if out is None else out
        if v.dtype == torch.float64:
            rc = self.ctx.lib.phub_hv_f64(self.ctx.h, self.h, C.c_void_p(v.data_ptr()),
                                          C.c_void_p(w.data_ptr()))
        elif v.dtype == torch.float32:
            rc = self.ctx.lib.phub_hv_f32(self.ctx.h, self.h, C.c_void_p(v.data_ptr()),
                                          C.c_void_p(w.data_ptr()))
        else:
            raise TypeError("v must be float64 or float32")
        self.ctx.check(rc)
        return w

    def masks(self):
        up = torch.empty(self.dim_up, dtype=torch.int64, device=self.ctx.device)
        dn = torch.empty(max(self.dim_dn, 1), dtype=torch.int64, device=self.ctx.device)
        self.ctx.check(self.ctx.lib.phub_op_masks(self.ctx.h, self.h, C.c_void_p(up.data_ptr()),
                                                  C.c_void_p(dn.data_ptr())))
        return up, dn[: self.dim_dn]

    def lanczos_workspace_bytes(self, mode=L.LCZ_PLAIN64, max_steps=600, fixed_steps=0,
                                refine_steps=0, want_psi=True) -> int:
        """Device bytes the Krylov vectors of such a run need (phub_lanczos_workspace_bytes)."""
        o = L.LanczosOpts(mode, max_steps, 0.0, 5, fixed_steps, refine_steps, 1,
                          1 if want_psi else 0, 1, None, 0)
        nb = C.c_size_t(0)
        rc = self.ctx.lib.phub_lanczos_workspace_bytes(self.h, C.byref(o), C.byref(nb))
        if rc:
            raise ValueError("bad Lanczos options")
        return int(nb.value)

    def _workspace(self, mode, max_steps, fixed_steps, refine_steps, want_psi):
        """A torch-owned workspace for the Krylov vectors (torch's caching allocator
        reuses it across calls instead of returning it to the driver).  The plain
        modes take what fits in free memory; they keep fewer vectors and rebuild psi
        by a second pass when it is smaller than the full need."""
        need = self.lanczos_workspace_bytes(mode, max_steps, fixed_steps, refine_steps, want_psi)
        dev = self.ctx.device
        if mode in (L.LCZ_PLAIN64, L.LCZ_PLAIN32):
            free, _ = torch.cuda.mem_get_info(dev)
            cached = torch.cuda.memory_reserved(dev) - torch.cuda.memory_allocated(dev)
            need = min(need, max(int(0.85 * (free + cached)), 0))
        try:
            return torch.empty(max(need, 1), dtype=torch.uint8, device=dev)
        except torch.cuda.OutOfMemoryError:
            return None

    def lanczos(self, mode=L.LCZ_PLAIN64, max_steps=600, tol=1e-8, check_every=5,
                fixed_steps=0, refine_steps=0, seed=1, v0: torch.Tensor | None = None,
                want_psi=True, raise_noconv=False, n_states=1,
                workspace: torch.Tensor | None = None, psi_tol: float = 0.0) -> LanczosOutput:
        """workspace: device memory for the Krylov vectors (any dtype, contiguous);
        by default one is taken from torch's caching allocator for the call.
        psi_tol > 0: residual target ||H psi - E psi|| (FP64 vectors: a stopping
        condition besides tol; MIXED / PLAIN32: the FP64 corrections of psi, at most
        refine_steps of them, stop once it is met)."""
        ws = workspace if workspace is not None else self._workspace(
            mode, max_steps, fixed_steps, refine_steps, want_psi)
        if ws is not None and not ws.is_contiguous():
            raise ValueError("workspace must be contiguous")
        o = L.LanczosOpts(mode, max_steps, tol, check_every, fixed_steps, refine_steps, seed,
                          1 if want_psi else 0, int(n_states),
                          C.c_void_p(ws.data_ptr()) if ws is not None else None,
                          ws.numel() * ws.element_size() if ws is not None else 0,
                          float(psi_tol))
        res = L.LanczosResult()
        nal = max(max_steps, fixed_steps) + 2
        alpha = np.zeros(nal)
        beta = np.zeros(nal + 1)
        psis = None
        if n_states > 1 and want_psi:
            psis = torch.zeros((int(n_states), self.dim), dtype=torch.float64, device=self.ctx.device)
            psi = psis[0]
        else:
            psi = self.vec(torch.float64) if (want_psi or mode == L.LCZ_MIXED) else None
        v0p = None
        if v0 is not None:
            v0 = v0.to(self.ctx.device, torch.float64).contiguous()
            v0p = C.c_void_p(v0.data_ptr())
        pbuf = psis if psis is not None else psi
        rc = self.ctx.lib.phub_lanczos(self.ctx.h, self.h, C.byref(o), v0p, _dptr(alpha),
                                       _dptr(beta), C.byref(res),
                                       C.c_void_p(pbuf.data_ptr()) if pbuf is not None else None)
        if rc == L.ERR_NOCONV and not raise_noconv:
            pass
        else:
            self.ctx.check(rc)
        nst = min(res.steps, L.MAX_STATES)
        if psis is None and psi is not None and want_psi:
            psis = psi.unsqueeze(0)
        return LanczosOutput(res.e0, res.e0_ritz, res.residual, res.steps, bool(res.converged),
                             res.ndeg, res.seconds, res.seconds_lanczos, res.hv_ms_avg,
                             alpha[: res.steps].copy(), beta[: res.steps + 1].copy(), psi,
                             psis, np.array(res.e_states[:nst]), bool(res.norm_lost),
                             int(res.corrections), int(res.correction_iters))

    def _ws(self, query, *args) -> torch.Tensor:
        """A device workspace of the size an entry point's *_workspace_bytes query gives."""
        nb = C.c_size_t(0)
        rc = query(self.h, *args, C.byref(nb))
        if rc:
            raise ValueError("workspace size query failed")
        return torch.empty(max(int(nb.value), 1), dtype=torch.uint8, device=self.ctx.device)

    def cf_chain_workspace(self, precision: int = 64) -> torch.Tensor:
        return self._ws(self.ctx.lib.phub_cf_chain_workspace_bytes, precision)

    def cf_chain(self, start: torch.Tensor, nsteps: int, precision: int = 64,
                 workspace: torch.Tensor | None = None):
        """Single-vector chain (alpha, beta, <start|start>); workspace: device memory of
        cf_chain_workspace() size for the three chain vectors (else allocated per call)."""
        a = np.zeros(nsteps)
        b = np.zeros(nsteps)
        n2 = np.zeros(1)
        start = start.to(self.ctx.device, torch.float64).contiguous()
        ws = workspace if workspace is not None else self.cf_chain_workspace(precision)
        self.ctx.check(self.ctx.lib.phub_cf_chain(self.ctx.h, self.h, C.c_void_p(start.data_ptr()),
                                                  nsteps, precision, _dptr(a), _dptr(b), _dptr(n2),
                                                  C.c_void_p(ws.data_ptr()),
                                                  ws.numel() * ws.element_size()))
        return a, b, float(n2[0])

    def band_workspace(self, p: int) -> torch.Tensor:
        return self._ws(self.ctx.lib.phub_band_lanczos_workspace_bytes, int(p))

    def band_lanczos(self, V: torch.Tensor, max_steps: int = 600, tol: float = 1e-8,
                     raise_noconv: bool = False, workspace: torch.Tensor | None = None):
        """Band Lanczos from the p start vectors V (p, dim) (GF_BANDLANCZOS): returns the
        poles (ascending eigenvalues of the band matrix) and Q (npoles, p) with
        Q[k, i] = sum_{j<=i} R[i, j] VEC[j, k]."""
        V = V.to(self.ctx.device, torch.float64).contiguous()
        p = V.shape[0]
        if V.shape[1] != self.dim:
            raise ValueError("start vectors must have the operator dimension")
        ws = workspace if workspace is not None else self.band_workspace(p)
        o = L.BandOpts(int(max_steps), float(tol), C.c_void_p(ws.data_ptr()),
                       ws.numel() * ws.element_size())
        pos = np.zeros(max_steps + p)
        Q = np.zeros((max_steps + p) * p)
        npl = C.c_int(0)
        rc = self.ctx.lib.phub_band_lanczos(self.ctx.h, self.h, C.c_void_p(V.data_ptr()), p,
                                            C.byref(o), _dptr(pos), _dptr(Q), C.byref(npl))
        if not (rc == L.ERR_NOCONV and not raise_noconv):
            self.ctx.check(rc)
        m = npl.value
        return pos[:m].copy(), Q[: m * p].reshape(m, p).copy()

    def minres_workspace(self) -> torch.Tensor:
        return self._ws(self.ctx.lib.phub_adjoint_minres_workspace_bytes)

    def minres(self, psi: torch.Tensor, e0: float, b: torch.Tensor, max_iter=500, tol=1e-10,
               workspace: torch.Tensor | None = None, out: torch.Tensor | None = None):
        """Adjoint MINRES (H - e0) x = b on the complement of psi; workspace of
        minres_workspace() size for the eight work vectors (torch's caching allocator
        reuses it across calls when none is passed)."""
        x = self.vec(torch.float64) if out is None else out
        ws = workspace if workspace is not None else self.minres_workspace()
        it = C.c_int(0)
        rr = C.c_double(0)
        self.ctx.check(self.ctx.lib.phub_adjoint_minres(
            self.ctx.h, self.h, C.c_void_p(psi.data_ptr()), e0, C.c_void_p(b.contiguous().data_ptr()),
            C.c_void_p(x.data_ptr()), max_iter, tol, C.byref(it), C.byref(rr),
            C.c_void_p(ws.data_ptr()), ws.numel() * ws.element_size()))
        return x, it.value, rr.value


class HubbardOp(_Op):
    """Hubbard operator of HPROD (hprod_hubbard.f90:17-118) on one (N_up, N_dn) sector."""

    def __init__(self, ctx: Context, n: int, nup: int, ndn: int, T, U, T_dn=None):
        self.ctx = ctx
        self.n, self.nup, self.ndn = int(n), int(nup), int(ndn)
        self.T = np.ascontiguousarray(np.asarray(T, dtype=np.float64).reshape(n, n))
        U = np.asarray(U, dtype=np.float64)
        # a (n, n, n, n) U is the 4-index interaction (is_Uloc = 3); per-site U is then 0
        self.U4 = np.ascontiguousarray(U) if U.ndim == 4 else None
        if self.U4 is not None:
            U = np.zeros(n)
        self.U = np.ascontiguousarray(np.full(n, float(U)) if U.ndim == 0 else U.reshape(n))
        self.T_dn = None if T_dn is None else np.ascontiguousarray(
            np.asarray(T_dn, dtype=np.float64).reshape(n, n))
        d = L.HubbardDesc(self.n, self.nup, self.ndn, _dptr(self.T), _dptr(self.U),
                          _dptr(self.T_dn) if self.T_dn is not None else None,
                          _dptr(self.U4) if self.U4 is not None else None)
        h = C.c_void_p()
        rc = ctx.lib.phub_op_hubbard(ctx.h, C.byref(d), C.byref(h))
        self.h = h
        ctx.check(rc)
        self._info()

    def rdm(self, bra: torch.Tensor, ket: torch.Tensor | None = None):
        ket = bra if ket is None else ket
        n = self.n
        up, dn, ni, ss = (np.zeros(n * n), np.zeros(n * n), np.zeros(n), np.zeros(n * n))
        self.ctx.check(self.ctx.lib.phub_rdm_hubbard(self.ctx.h, self.h, C.c_void_p(bra.data_ptr()),
                                                     C.c_void_p(ket.data_ptr()), _dptr(up),
                                                     _dptr(dn), _dptr(ni), _dptr(ss)))
        return up.reshape(n, n), dn.reshape(n, n), ni, ss.reshape(n, n)

    def two_rdm(self, bra: torch.Tensor, ket: torch.Tensor | None = None):
        """DBL_OCC(J,K,M,N) (TWO_RDM, static_rq.f90:247-306), shape (n, n, n, n)."""
        ket = bra if ket is None else ket
        n = self.n
        out = np.zeros(n ** 4)
        self.ctx.check(self.ctx.lib.phub_two_rdm_hubbard(self.ctx.h, self.h,
                                                         C.c_void_p(bra.data_ptr()),
                                                         C.c_void_p(ket.data_ptr()), _dptr(out)))
        return out.reshape(n, n, n, n)

    def hf_grad(self, psi: torch.Tensor):
        n = self.n
        gt, gu = np.zeros(n * n), np.zeros(n)
        self.ctx.check(self.ctx.lib.phub_hf_grad_hubbard(self.ctx.h, self.h,
                                                         C.c_void_p(psi.data_ptr()), _dptr(gt),
                                                         _dptr(gu)))
        return gt.reshape(n, n), gu

    def one_rdm_spin(self, bra: torch.Tensor, ket: torch.Tensor | None = None):
        """Proper spin-resolved gamma_s[J][K] = <bra| c^dag_J c_K |ket> (0-based J, K)."""
        ket = bra if ket is None else ket
        n = self.n
        gu, gd = np.zeros(n * n), np.zeros(n * n)
        self.ctx.check(self.ctx.lib.phub_one_rdm_spin(self.ctx.h, self.h, C.c_void_p(bra.data_ptr()),
                                                      C.c_void_p(ket.data_ptr()), _dptr(gu),
                                                      _dptr(gd)))
        return gu.reshape(n, n), gd.reshape(n, n)

    def grad(self, bra: torch.Tensor, ket: torch.Tensor):
        """<bra| dH/dT(J,K) |ket> (n x n) and <bra| dH/dU_J |ket> (n)."""
        n = self.n
        gt, gu = np.zeros(n * n), np.zeros(n)
        self.ctx.check(self.ctx.lib.phub_grad_hubbard(self.ctx.h, self.h, C.c_void_p(bra.data_ptr()),
                                                      C.c_void_p(ket.data_ptr()), _dptr(gt),
                                                      _dptr(gu)))
        return gt.reshape(n, n), gu

    def excite(self, psi: torch.Tensor, site: int, spin: int, create: bool, target=None):
        """c^dag_{site,spin} psi or c_{site,spin} psi (site 1-based); returns (op, vector)."""
        dn = 1 if create else -1
        if target is None:
            target = HubbardOp(self.ctx, self.n, self.nup + (dn if spin == 0 else 0),
                               self.ndn + (dn if spin == 1 else 0), self.T, self.U)
        out = target.vec(torch.float64)
        if target.dim:
            self.ctx.check(self.ctx.lib.phub_excite(self.ctx.h, self.h, target.h, site, spin,
                                                    1 if create else 0, C.c_void_p(psi.data_ptr()),
                                                    C.c_void_p(out.data_ptr())))
        return target, out


class HeisenbergOp(_Op):
    """Heisenberg operator of HPROD (hprod_heis.f90:1-61), S_z = 0 sector."""

    def __init__(self, ctx: Context, n: int, J, sign_convention: int = 0):
        self.ctx = ctx
        self.n = int(n)
        self.J = np.ascontiguousarray(np.asarray(J, dtype=np.float64).reshape(n, n))
        self.sign_convention = int(sign_convention)
        d = L.HeisDesc(self.n, _dptr(self.J), self.sign_convention)
        h = C.c_void_p()
        rc = ctx.lib.phub_op_heisenberg(ctx.h, C.byref(d), C.byref(h))
        self.h = h
        ctx.check(rc)
        self._info()

    def sisj(self, bra: torch.Tensor, ket: torch.Tensor | None = None):
        ket = bra if ket is None else ket
        n = self.n
        ss = np.zeros(n * n)
        self.ctx.check(self.ctx.lib.phub_rdm_heisenberg(self.ctx.h, self.h,
                                                        C.c_void_p(bra.data_ptr()),
                                                        C.c_void_p(ket.data_ptr()), _dptr(ss)))
        return ss.reshape(n, n)

    def hf_grad(self, psi: torch.Tensor):
        """dE0/dJ(J,K) over ordered pairs: <psi| dH/dJ(J,K) |psi> (the diagonal is
        0.75 <psi|psi>)."""
        n = self.n
        g = np.zeros(n * n)
        self.ctx.check(self.ctx.lib.phub_hf_grad_heisenberg(self.ctx.h, self.h,
                                                            C.c_void_p(psi.data_ptr()), _dptr(g)))
        return g.reshape(n, n)


def cf_eval(ctx: Context, alpha, beta, norm2, shift, sign, omega, eta):
    """G(omega) for a set of chains on the device; returns complex128 (nch, nw)."""
    alpha = np.ascontiguousarray(np.atleast_2d(alpha), dtype=np.float64)
    beta = np.ascontiguousarray(np.atleast_2d(beta), dtype=np.float64)
    nch, m = alpha.shape
    norm2 = np.ascontiguousarray(np.broadcast_to(np.asarray(norm2, float), (nch,)))
    shift = np.ascontiguousarray(np.broadcast_to(np.asarray(shift, float), (nch,)))
    sign = np.ascontiguousarray(np.broadcast_to(np.asarray(sign, float), (nch,)))
    omega = np.ascontiguousarray(omega, dtype=np.float64)
    out = np.zeros(2 * nch * len(omega))
    ctx.check(ctx.lib.phub_cf_eval(ctx.h, nch, m, _dptr(alpha), _dptr(beta), _dptr(norm2),
                                   _dptr(shift), _dptr(sign), len(omega), _dptr(omega), float(eta),
                                   _dptr(out)))
    return out.view(np.complex128).reshape(nch, len(omega))
