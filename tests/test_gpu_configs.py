"""GPU tests of the benchmark configurations at their full sizes, against the oracle.

- 4x4 t-t' Hubbard (t = 1, t' = -0.2, U = 8, 8 + 8, D = 165,636,900): one H*v in
  FP64 against the oracle's rank-table HPROD (OpenMP) within 1e-13 relative,
  and in FP32 within 2e-5; the mixed-precision Lanczos against the FP64 one
  (E0, 1-RDM, double occupancy and S.S at residual 1e-10).
- 256 batched clusters (120-step chains, 2,048 omega points, eta = 0.05): the
  per-chain sum rule on every cluster, E0 against dense FP64 ED of the oracle's
  matrix for every cluster up to 2,000 states, and the chains of the first and
  last site against dense ED, the oracle's moments and its FP64 chain (the FP32
  error printed).
- 6x6 J1-J2 Heisenberg (S_z = 0, D = 9,075,135,300): H applied to a sparse vector
  of 1,000 unit vectors spread over the index range (many past 2^32) equals the
  oracle's columns exactly (dyadic values make every FP32 sum exact), nothing
  else is touched, in both sign conventions; <u, H v> = <H u, v> at full size.
"""
import math
import os

import numpy as np
import pytest
import torch

from oracle import oracle as O

pytestmark = pytest.mark.gpu

NCPU = max(1, min(16, os.cpu_count() or 1))


def _free():
    torch.cuda.synchronize()
    torch.cuda.empty_cache()


# ---------------------------------------------------------------------------
# 4x4 t-t' Hubbard
# ---------------------------------------------------------------------------
def _hub4x4():
    T = O.square_T(4, 4, t=1.0, tp=-0.2, periodic=True)
    return T, 8.0


def test_config2_hv_full_size_vs_oracle(ctx):
    from pyhub_amd.engine import HubbardOp
    T, U = _hub4x4()
    op = HubbardOp(ctx, 16, 8, 8, T, U)
    m = O.HubbardModel(16, 8, 8, T, U)
    assert op.dim == m.dim == 12870 ** 2
    v = np.random.default_rng(2).uniform(-0.5, 0.5, m.dim)
    ref = m.hv(v, fast=True, nthreads=NCPU)
    sc = float(np.max(np.abs(ref)))
    vd = torch.from_numpy(v).cuda()
    w = op.hv(vd).cpu().numpy()
    e64 = float(np.max(np.abs(w - ref))) / sc
    del w
    w32 = op.hv(vd.float()).double().cpu().numpy()
    e32 = float(np.max(np.abs(w32 - ref))) / sc
    print(f"4x4 H*v: FP64 max rel err {e64:.2e}, FP32 {e32:.2e}", flush=True)
    del vd, w32
    _free()
    assert e64 <= 1e-13
    assert e32 <= 2e-5


def test_config2_mixed_lanczos_vs_fp64(ctx):
    """Lanczos of the 4x4 model from the same start vector, both to the residual
    ||H psi - E psi|| <= 1e-10: the mixed form (FP32 basis with full
    reorthogonalisation, FP64 scalars, FP64 corrections of psi) and the reference's
    FP64 three-term form agree on E0 to 1e-12 relative and on the 1-RDM, the double
    occupancy and <S_i.S_j> to 1e-8 (the reduced-quantity tolerance)."""
    from pyhub_amd import _lib as L
    from pyhub_amd.engine import HubbardOp
    T, U = _hub4x4()
    op = HubbardOp(ctx, 16, 8, 8, T, U)
    v0 = torch.from_numpy(np.random.default_rng(2).uniform(-0.5, 0.5, op.dim)).cuda()
    out = {}
    for name, mode in (("fp64", L.LCZ_PLAIN64), ("mixed", L.LCZ_MIXED)):
        r = op.lanczos(mode=mode, max_steps=300, tol=8.5e-9, v0=v0, refine_steps=4, psi_tol=1e-10)
        out[name] = (r.e0, r.residual, r.steps, op.rdm(r.psi))
        print(f"4x4 {name}: E0 {r.e0:.13f}, {r.steps} steps, residual {r.residual:.1e}, "
              f"{r.corrections} FP64 corrections", flush=True)
        del r
        _free()
    e64, res64, _, q64 = out["fp64"]
    emx, resmx, _, qmx = out["mixed"]
    dq = max(float(np.max(np.abs(a - b))) for a, b in zip(q64, qmx))
    print(f"4x4 mixed vs fp64: |dE0|/|E0| {abs(emx - e64) / abs(e64):.1e}, max |d(1-RDM, NI, S.S)| "
          f"{dq:.1e}", flush=True)
    assert res64 <= 1e-9 and resmx <= 1e-9
    assert abs(emx - e64) <= 1e-12 * abs(e64)
    assert dq <= 1e-8
    del v0
    _free()


# ---------------------------------------------------------------------------
# 256 batched clusters
# ---------------------------------------------------------------------------
def _dense_ed(m):
    H = m.dense(fast=True)
    return np.linalg.eigh(H)


def _cf(a, b, n2, zz):
    """n2 / (zz - a0 - b0^2 / (zz - a1 - ...)), trailing zero levels dropped."""
    f = np.zeros_like(zz)
    ln = len(a)
    while ln > 1 and a[ln - 1] == 0.0 and b[ln - 2] == 0.0:
        ln -= 1
    for k in range(ln - 1, -1, -1):
        bb = b[k] ** 2 if k < ln - 1 else 0.0
        f = 1.0 / (zz - a[k] - bb * f)
    return n2 * f


def _chain_moments(a, b, n2, J):
    """n2 (T^j)_00, j < J: the spectral moments the chain reproduces."""
    m = len(a)
    T = np.diag(a) + np.diag(b[:m - 1], 1) + np.diag(b[:m - 1], -1)
    e = np.zeros(m)
    e[0] = 1.0
    out = []
    for _ in range(J):
        out.append(n2 * e[0])
        e = T @ e
    return np.array(out)


def test_config3_all_clusters_vs_oracle(ctx):
    """All 256 clusters in one batched call (FP32 chain vectors, FP64 scalars):

    - sum rule n_particle + n_hole = 1 for every site and spin of every cluster;
    - E0 against dense FP64 ED for every cluster up to 2,000 states;
    - for those with a non-degenerate ground state, every chain of the first and
      last site: on N+-1 sectors of at most 120 states (exhausted by the chain)
      G(omega) at eta = 0.05 against dense ED; on larger sectors the spectral
      moments 0..7 against <x|H^j|x> of the oracle's HPROD and G at eta = 0.5
      (resolved by 120 steps) against the oracle's FP64 chain.  The eta = 0.05
      difference to the FP64 chain on the larger sectors is printed: there 120
      steps do not resolve the spectrum, and FP32 and FP64 chains place the
      unconverged poles differently (the FP64 chain itself is off the exact
      G by about the same amount)."""
    from pyhub_amd.batch import gii_omega, random_cluster, run_batch
    clusters = [random_cluster(c) for c in range(256)]
    nsteps = 120
    e0, ga, gb, gn = run_batch(clusters, gs_steps=256, gf_steps=nsteps, seed=3, gs_tol=1e-9)
    omega = np.linspace(-15, 15, 2048)
    G = gii_omega(e0, ga, gb, gn, clusters, omega, eta=0.05)
    worst_sum = 0.0
    for ci, cl in enumerate(clusters):
        for i in range(cl.n):
            for s in (0, 1):
                worst_sum = max(worst_sum, abs(gn[ci, i, s].sum() - 1.0))
    print(f"sum rule: worst |n_p + n_h - 1| = {worst_sum:.2e} over all 256 clusters", flush=True)
    assert worst_sum < 1e-6
    e_err, ex, mom, g05, g005 = [], [], [], [], []
    for ci, cl in enumerate(clusters):
        D = math.comb(cl.n, cl.n_up) * math.comb(cl.n, cl.n_dn)
        if D > 2000:
            continue
        m = O.HubbardModel(cl.n, cl.n_up, cl.n_dn, cl.T, cl.U)
        ev, V = _dense_ed(m)
        e_err.append((abs(e0[ci] - ev[0]) / max(1.0, abs(ev[0])), ci))
        if ev[1] - ev[0] < 1e-6:
            continue                       # degenerate ground state: psi0 not unique
        psi0 = V[:, 0]
        for site in sorted({1, cl.n}):
            for spin in (0, 1):
                gsum = np.zeros(len(omega), dtype=complex)
                gref = np.zeros(len(omega), dtype=complex)
                exact = True
                for kind, create, sgn in ((0, 1, 1.0), (1, 0, -1.0)):
                    nsp = (cl.n_up if spin == 0 else cl.n_dn) + (1 if create else -1)
                    if not 0 <= nsp <= cl.n:
                        continue
                    tgt, x = m.excite(psi0, site, spin, create)
                    if x @ x == 0.0:
                        continue
                    a_g, b_g, n_g = ga[ci, site - 1, spin, kind], gb[ci, site - 1, spin, kind], \
                        gn[ci, site - 1, spin, kind]
                    if tgt.dim <= nsteps:
                        zz = sgn * (omega + 0.05j) + ev[0]
                        evt, Vt = np.linalg.eigh(tgt.dense(fast=True))
                        w = (Vt.T @ x) ** 2
                        gref += sgn * (w[None, :] / (zz[:, None] - evt[None, :])).sum(axis=1)
                        continue
                    exact = False
                    y, mu = x.copy(), []
                    for _ in range(8):
                        mu.append(x @ y)
                        y = tgt.hv(y, fast=True)
                    R = abs(ev[0]) + 20.0
                    mom.append(np.max(np.abs(_chain_moments(a_g, b_g, n_g, 8) - np.array(mu))
                                      / (n_g * R ** np.arange(8))))
                    a, b, n2 = tgt.cf_chain(x, nsteps, fast=True)
                    for eta, lst in ((0.5, g05), (0.05, g005)):
                        zz = sgn * (omega + 1j * eta) + ev[0]
                        go = _cf(a, b, n2, zz)
                        lst.append(np.max(np.abs(_cf(a_g, b_g, n_g, zz) - go)) / np.max(np.abs(go)))
                if exact:
                    ex.append(np.max(np.abs(G[ci, site - 1, spin] - gref)) / np.max(np.abs(gref)))
    e_err.sort(reverse=True)

    def st(v):
        v = np.sort(v)
        return f"n={len(v)} median {v[len(v) // 2]:.1e} p90 {v[int(0.9 * len(v))]:.1e} max {v[-1]:.1e}"

    print(f"E0 vs dense ED: {len(e_err)} clusters, worst rel err {e_err[0][0]:.2e}", flush=True)
    print(f"G_ii (eta 0.05) vs dense ED, exhausted sectors: {st(ex)}", flush=True)
    print(f"chain moments 0..7 vs <x|H^j|x>: {st(mom)}", flush=True)
    print(f"G (eta 0.5) vs oracle FP64 chain: {st(g05)}", flush=True)
    print(f"G (eta 0.05) vs oracle FP64 chain (unresolved, reported): {st(g005)}", flush=True)
    assert e_err[0][0] < 1e-9
    assert np.median(ex) < 1e-4 and max(ex) < 3e-2
    assert max(mom) < 1e-6
    assert max(g05) < 1e-3


# ---------------------------------------------------------------------------
# 6x6 J1-J2 Heisenberg
# ---------------------------------------------------------------------------
def _j1j2_6x6():
    Jstd = O.j1j2_square(6, 6, 1.0, 0.5)
    return (Jstd + Jstd.T) / 2.0


def _spread_indices(D, rng, count):
    idx = {0, 1, D - 1, D - 2, 2 ** 31 - 1, 2 ** 31, 2 ** 32 - 1, 2 ** 32, 2 ** 32 + 1,
           2 ** 33 + 12345}
    while len(idx) < count:
        idx.add(int(rng.integers(0, D)))
    return sorted(i for i in idx if i < D)


@pytest.mark.parametrize("conv,dtype", [(1, torch.float32), (0, torch.float32), (1, torch.float64)])
def test_config5_heis6x6_sparse_columns_exact(ctx, conv, dtype):
    """H applied to 1,000 scaled unit vectors spread over the 6x6 index range (indices past
    2^32 included) equals the oracle's columns exactly (dyadic coefficients), FP32 and FP64
    (145 GB of vectors), both sign conventions; no other row is touched."""
    from pyhub_amd.engine import HeisenbergOp
    n = 36
    Jm = _j1j2_6x6()
    op = HeisenbergOp(ctx, n, Jm, sign_convention=conv)
    D = O.binom(36, 18)
    assert op.dim == D
    rng = np.random.default_rng(50 + conv)
    cols = _spread_indices(D, rng, 1000)
    coef = rng.integers(1, 9, len(cols)) / 8.0 * rng.choice([-1.0, 1.0], len(cols))
    expect: dict = {}
    for c, a in zip(cols, coef):
        for r, val in O.heis_column(n, Jm, c, conv).items():
            expect[r] = expect.get(r, 0.0) + a * val
    rows = np.array(sorted(expect), dtype=np.int64)
    want = np.array([expect[r] for r in rows])
    v = torch.zeros(D, dtype=dtype, device="cuda")
    v[torch.as_tensor(np.array(cols, dtype=np.int64), device="cuda")] = torch.as_tensor(
        coef, dtype=dtype, device="cuda")
    w = op.hv(v)
    del v
    got = w[torch.as_tensor(rows, device="cuda")].double().cpu().numpy()
    nz_expect = int(np.count_nonzero(want))
    nz = int(torch.count_nonzero(w).item())
    del w
    _free()
    print(f"6x6 conv {conv} {dtype}: {len(cols)} columns, {len(rows)} rows reached, max |diff| "
          f"{np.max(np.abs(got - want)):.1e}, nonzeros {nz} (expected {nz_expect})", flush=True)
    assert np.array_equal(got, want)
    assert nz == nz_expect


def test_config5_heis6x6_symmetric_full_size(ctx):
    """<u, H v> = <H u, v> for random FP32 u, v of 9.1e9 states (FP64 dots)."""
    from pyhub_amd.engine import HeisenbergOp
    op = HeisenbergOp(ctx, 36, _j1j2_6x6(), sign_convention=1)
    g = torch.Generator(device="cuda")
    g.manual_seed(7)
    u = torch.rand(op.dim, dtype=torch.float32, device="cuda", generator=g) - 0.5
    v = torch.rand(op.dim, dtype=torch.float32, device="cuda", generator=g) - 0.5
    hv = op.hv(v)
    hu = op.hv(u)

    def dot(a, b):
        s = 0.0
        ch = 1 << 28
        for i in range(0, a.numel(), ch):
            s += float(torch.dot(a[i:i + ch].double(), b[i:i + ch].double()).item())
        return s

    a, b = dot(u, hv), dot(hu, v)
    scale = math.sqrt(dot(u, u) * dot(hv, hv))
    del u, v, hu, hv
    _free()
    print(f"6x6 <u,Hv> = {a:.10e}, <Hu,v> = {b:.10e}, rel diff {abs(a - b) / scale:.1e}", flush=True)
    assert abs(a - b) <= 1e-6 * scale
