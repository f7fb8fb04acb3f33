"""Batched clusters: ground states and diagonal Green's functions of many small
Hubbard clusters in one call (phub_batch_hubbard: two persistent launches, ground
states in FP64 vectors then all chains in FP32 vectors with FP64 scalars), and the
continued-fraction evaluation of G_ii(omega) on a grid (phub_cf_eval).

The random-cluster workload follows the batched configuration: L sites uniform in
{6..12} on a random connected graph (a ring over a random permutation plus random
chords, node degree 2..4), t_ij ~ U(0.5, 1.5) (T = -t), eps_i ~ N(0, 0.5),
U ~ U(2, 10), N_up, N_dn uniform in [1, L-1]; generator default_rng([3, c]).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from pyhub_amd import _lib as L
from pyhub_amd.engine import Context


@dataclass
class Cluster:
    n: int
    n_up: int
    n_dn: int
    T: np.ndarray
    U: np.ndarray


def random_cluster(c: int) -> Cluster:
    rng = np.random.default_rng([3, c])
    L_ = int(rng.integers(6, 13))
    perm = rng.permutation(L_)
    adj = np.zeros((L_, L_), dtype=bool)
    for a in range(L_):
        i, j = perm[a], perm[(a + 1) % L_]
        adj[i, j] = adj[j, i] = True
    target = rng.integers(2, 5, L_)
    for _ in range(4 * L_):
        i, j = rng.integers(0, L_, 2)
        if i == j or adj[i, j]:
            continue
        if adj[i].sum() < min(target[i], 4) and adj[j].sum() < min(target[j], 4):
            adj[i, j] = adj[j, i] = True
    T = np.zeros((L_, L_))
    iu, ju = np.nonzero(np.triu(adj, 1))
    t = rng.uniform(0.5, 1.5, len(iu))
    T[iu, ju] = -t
    T[ju, iu] = -t
    T[np.arange(L_), np.arange(L_)] = rng.normal(0.0, 0.5, L_)
    U = np.full(L_, rng.uniform(2.0, 10.0))
    nup = int(rng.integers(1, L_))
    ndn = int(rng.integers(1, L_))
    return Cluster(L_, nup, ndn, T, U)


def run_batch(clusters, gs_steps=256, gf_steps=120, seed=7, gs_tol=1e-9,
              ctx: Context | None = None):
    """One persistent launch; returns e0 (ncl,), alpha/beta (ncl, 32, 2, 2, steps) and
    norm2 (ncl, 32, 2, 2): index [c, site-1, spin, kind] with kind 0 = particle
    (c^dag psi0), 1 = hole (c psi0)."""
    ctx = ctx or Context.default()
    ncl = len(clusters)
    keep = []
    arr = (L.Cluster * ncl)()
    for i, cl in enumerate(clusters):
        T = np.ascontiguousarray(cl.T, dtype=np.float64)
        U = np.ascontiguousarray(cl.U, dtype=np.float64)
        keep += [T, U]
        arr[i] = L.Cluster(cl.n, cl.n_up, cl.n_dn, T.ctypes.data_as(C.POINTER(C.c_double)),
                           U.ctypes.data_as(C.POINTER(C.c_double)))
    opts = L.BatchOpts(gs_steps, gf_steps, seed, gs_tol)
    e0 = np.zeros(ncl)
    nslot = ncl * 32 * 2 * 2
    ga = np.zeros(nslot * gf_steps)
    gb = np.zeros(nslot * gf_steps)
    gn = np.zeros(nslot)
    dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))  # noqa: E731
    ctx.check(ctx.lib.phub_batch_hubbard(ctx.h, arr, ncl, C.byref(opts), dp(e0), dp(ga), dp(gb),
                                         dp(gn)))
    shp = (ncl, 32, 2, 2)
    return e0, ga.reshape(*shp, gf_steps), gb.reshape(*shp, gf_steps), gn.reshape(shp)


def gii_omega(e0, ga, gb, gn, clusters, omega, eta=0.05, ctx: Context | None = None):
    """G_ii,s(omega) = particle + hole continued fractions, evaluated on the device.
    Returns complex (ncl, 32, 2, nw); entries of absent sites are zero."""
    from pyhub_amd.engine import cf_eval
    ctx = ctx or Context.default()
    ncl = len(clusters)
    m = ga.shape[-1]
    nw = len(omega)
    # only the chains of sites a cluster has (slot [c, site-1, spin, kind]); the other
    # slots of the (ncl, 32, 2) result stay zero
    nsite = np.array([cl.n for cl in clusters])
    site_ok = np.arange(32)[None, :] < nsite[:, None]                    # (ncl, 32)
    sel = np.broadcast_to(site_ok[:, :, None, None], (ncl, 32, 2, 2)).reshape(-1)
    a = ga.reshape(-1, m)[sel]
    b = gb.reshape(-1, m)[sel]
    n2 = gn.reshape(-1)[sel]
    shift = np.repeat(np.asarray(e0), 32 * 2 * 2)[sel]
    sign = np.tile(np.array([1.0, -1.0]), ncl * 32 * 2)[sel]           # particle +, hole -
    Gc = cf_eval(ctx, a, b, n2, shift, sign, np.asarray(omega, float), eta)
    G = np.zeros((ncl, 32, 2, nw), dtype=np.complex128)
    G[np.broadcast_to(site_ok[:, :, None], (ncl, 32, 2))] = Gc[0::2] + Gc[1::2]
    return G
