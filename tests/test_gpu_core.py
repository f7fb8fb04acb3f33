"""GPU parity tests for the core path (basis, H*v, Lanczos, reduced quantities),
calling the HIP library through its C ABI and checking against the oracle."""
import numpy as np
import pytest
import torch

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def rand_cluster(rng, n, eps_sigma=0.3, dyadic=False):
    if dyadic:
        T = rng.integers(-4, 5, size=(n, n)) / 4.0
        T = np.triu(T, 1)
        T = T + T.T
        T[np.arange(n), np.arange(n)] = rng.integers(-2, 3, n) / 2.0
        U = rng.integers(0, 9, n) / 2.0
    else:
        T = rng.normal(size=(n, n))
        T = (T + T.T) / 2
        T[rng.random((n, n)) < 0.3] = 0.0
        T = np.triu(T, 1)
        T = T + T.T
        T[np.arange(n), np.arange(n)] = rng.normal(0, eps_sigma, n)
        U = rng.uniform(1, 6, n)
    return T, U


def dev(x, dtype=torch.float64):
    return torch.as_tensor(np.asarray(x), dtype=dtype, device="cuda")


@pytest.mark.parametrize("n,k", [(1, 0), (1, 1), (4, 2), (8, 4), (10, 3), (12, 6), (16, 8), (16, 1), (13, 13)])
def test_basis_masks_and_ranks_bit_exact(ctx, n, k):
    m = ctx.basis(n, k).cpu().numpy().astype(np.uint64)
    assert np.array_equal(m, O.hilbert(n, k))
    r = ctx.rank(n, k, torch.as_tensor(m.astype(np.int64), device="cuda")).cpu().numpy()
    assert np.array_equal(r, np.arange(len(m)))


@pytest.mark.parametrize("n,nu,nd,seed", [(4, 2, 2, 0), (5, 2, 3, 1), (6, 3, 2, 2), (6, 1, 4, 3),
                                          (7, 3, 3, 4), (8, 4, 4, 5), (6, 0, 3, 6), (5, 5, 2, 7),
                                          (9, 4, 5, 8), (10, 5, 5, 9)])
def test_hubbard_hv_f64_vs_oracle(ctx, n, nu, nd, seed):
    from pyhub_amd.engine import HubbardOp
    rng = np.random.default_rng(seed)
    T, U = rand_cluster(rng, n)
    op = HubbardOp(ctx, n, nu, nd, T, U)
    m = O.HubbardModel(n, nu, nd, T, U)
    assert op.dim == m.dim
    up, dn = op.masks()
    assert np.array_equal(up.cpu().numpy().astype(np.uint64), m.bup)
    assert np.array_equal(dn.cpu().numpy().astype(np.uint64), m.bdn)
    v = rng.normal(size=m.dim)
    w = op.hv(dev(v)).cpu().numpy()
    wref = m.hv(v)
    assert np.max(np.abs(w - wref)) <= 1e-13 * max(1.0, np.max(np.abs(wref)))
    # FP32 path: same operator, FP32 vectors and coefficients
    w32 = op.hv(dev(v, torch.float32)).cpu().numpy().astype(np.float64)
    assert np.max(np.abs(w32 - wref)) <= 2e-5 * max(1.0, np.max(np.abs(wref)))


@pytest.mark.parametrize("n,nu,nd,seed", [(4, 2, 2, 10), (5, 2, 3, 11), (6, 3, 3, 12), (6, 2, 4, 13)])
def test_hubbard_dense_matrix_bit_exact(ctx, n, nu, nd, seed):
    """With dyadic T and U every matrix element is exact: the GPU H built column by
    column must equal the reference-faithful oracle's bit for bit (signs, order)."""
    from pyhub_amd.engine import HubbardOp
    rng = np.random.default_rng(seed)
    T, U = rand_cluster(rng, n, dyadic=True)
    op = HubbardOp(ctx, n, nu, nd, T, U)
    m = O.HubbardModel(n, nu, nd, T, U)
    I = torch.eye(m.dim, dtype=torch.float64, device="cuda")
    Hg = np.stack([op.hv(I[:, i].contiguous()).cpu().numpy() for i in range(m.dim)], axis=1)
    Ho = m.dense(fast=False)
    assert np.array_equal(Hg, Ho)


N_DN_TILES = 25   # the down-hop tuner's candidates: 17 of k_hv_dn, 8 of k_hv_dn2


@pytest.mark.parametrize("tile", list(range(N_DN_TILES)))
def test_hubbard_down_tiles_vs_oracle(ctx, tile, monkeypatch):
    """Every down-hop tile the H*v tuner can pick gives the oracle's H*v (FP64 to
    1e-13, FP32 to 2e-5 relative), on a sector whose dimensions are not
    multiples of any tile."""
    from pyhub_amd.engine import HubbardOp
    monkeypatch.setenv("PHUB_DN_TILE", str(tile))
    rng = np.random.default_rng(70 + tile)
    n, nu, nd = 10, 4, 3
    T, U = rand_cluster(rng, n)
    op = HubbardOp(ctx, n, nu, nd, T, U)
    m = O.HubbardModel(n, nu, nd, T, U)
    v = rng.normal(size=m.dim)
    ref = m.hv(v)
    w64 = op.hv(dev(v)).cpu().numpy()
    assert np.max(np.abs(w64 - ref)) <= 1e-13 * max(1.0, np.max(np.abs(ref)))
    w32 = op.hv(dev(v, torch.float32)).double().cpu().numpy()
    assert np.max(np.abs(w32 - ref)) <= 2e-5 * max(1.0, np.max(np.abs(ref)))


@pytest.mark.parametrize("tile", list(range(N_DN_TILES)))
def test_hubbard_down_tiles_lanczos_vs_dense(ctx, tile, monkeypatch):
    """The fused Lanczos epilogue of every down-hop tile (scale, beta * prev and the
    dot partials): FP64 three-term Lanczos E0 against the dense eigenvalue of the
    oracle's matrix to 1e-10 relative, on a sector of 1,568 states."""
    from pyhub_amd import _lib as L
    from pyhub_amd.engine import HubbardOp
    monkeypatch.setenv("PHUB_DN_TILE", str(tile))
    rng = np.random.default_rng(170 + tile)
    n, nu, nd = 8, 3, 2
    T, U = rand_cluster(rng, n)
    op = HubbardOp(ctx, n, nu, nd, T, U)
    m = O.HubbardModel(n, nu, nd, T, U)
    e_dense = float(np.linalg.eigvalsh(m.dense())[0])
    v0 = rng.uniform(-0.5, 0.5, m.dim)
    r = op.lanczos(mode=L.LCZ_PLAIN64, max_steps=400, tol=1e-12, v0=dev(v0))
    assert abs(r.e0 - e_dense) <= 1e-10 * abs(e_dense)


@pytest.mark.parametrize("n,conv,seed", [(4, 0, 0), (6, 0, 1), (8, 0, 2), (10, 0, 3), (6, 1, 4), (8, 1, 5)])
def test_heisenberg_hv_vs_oracle(ctx, n, conv, seed):
    from pyhub_amd.engine import HeisenbergOp
    rng = np.random.default_rng(seed)
    J = rng.normal(size=(n, n))
    J = (J + J.T) / 2
    op = HeisenbergOp(ctx, n, J, sign_convention=conv)
    m = O.HeisenbergModel(n, J)
    v = rng.normal(size=m.dim)
    w = op.hv(dev(v)).cpu().numpy()
    if conv == 0:
        assert np.max(np.abs(w - m.hv(v))) <= 1e-12 * max(1.0, np.max(np.abs(w)))
    else:
        # physical convention: spectrum = 2x the standard sum_{i<j} J_ij S_i.S_j
        I = torch.eye(m.dim, dtype=torch.float64, device="cuda")
        H = np.stack([op.hv(I[:, i].contiguous()).cpu().numpy() for i in range(m.dim)], axis=1)
        e = np.linalg.eigvalsh(H)
        es = np.linalg.eigvalsh(O.heisenberg_standard_dense(n, np.triu(J, 1)))
        # S_z = 0 block holds a member of every multiplet; the ordered-pair sum
        # doubles the off-diagonal bonds, the diagonal J(J,J) adds 0.75 J(J,J)
        assert abs(e[0] - (2 * es[0] + 0.75 * np.trace(J))) < 1e-10


@pytest.mark.parametrize("n,seed,dtype", [(5, 40, torch.float64), (12, 41, torch.float64),
                                          (14, 42, torch.float64), (16, 43, torch.float64),
                                          (20, 44, torch.float64), (16, 45, torch.float32)])
def test_heisenberg_block_hv_vs_oracle(ctx, n, seed, dtype):
    """The block H*v at sizes where blocks of every popcount and all three bond
    kinds (low-low, high-high, cross) occur, against the oracle's HPROD."""
    from pyhub_amd.engine import HeisenbergOp
    rng = np.random.default_rng(seed)
    J = rng.normal(size=(n, n))
    J[rng.random((n, n)) < 0.5] = 0.0
    J = (J + J.T) / 2
    op = HeisenbergOp(ctx, n, J)
    m = O.HeisenbergModel(n, J)
    v = rng.normal(size=m.dim)
    ref = m.hv(v, fast=True)
    w = op.hv(dev(v, dtype)).double().cpu().numpy()
    tol = 1e-12 if dtype == torch.float64 else 2e-5
    print(f"n={n} dim={m.dim} max|w-ref|={np.max(np.abs(w - ref)):.3e}", flush=True)
    assert np.max(np.abs(w - ref)) <= tol * max(1.0, np.max(np.abs(ref)))


@pytest.mark.parametrize("n,nh,seed", [(14, 3, 47), (16, 5, 48), (18, 7, 49), (20, 6, 50)])
def test_heisenberg_hv_high_half_sizes(ctx, monkeypatch, n, nh, seed):
    """The H*v with fewer high sites than n - n/2 (PHUB_HEIS_NH): more low-low and
    cross bonds, more and smaller blocks, popcounts with no neighbouring block; FP64
    and FP32 against the oracle's HPROD, both sign conventions' flip signs."""
    from pyhub_amd.engine import HeisenbergOp
    monkeypatch.setenv("PHUB_HEIS_NH", str(nh))
    rng = np.random.default_rng(seed)
    J = rng.normal(size=(n, n))
    J[rng.random((n, n)) < 0.4] = 0.0
    J = (J + J.T) / 2
    op = HeisenbergOp(ctx, n, J)
    m = O.HeisenbergModel(n, J)
    v = rng.normal(size=m.dim)
    ref = m.hv(v, fast=True)
    w64 = op.hv(dev(v)).cpu().numpy()
    w32 = op.hv(dev(v, torch.float32)).double().cpu().numpy()
    scale = max(1.0, np.max(np.abs(ref)))
    print(f"n={n} nh={nh} fp64 {np.max(np.abs(w64 - ref)):.2e} fp32 {np.max(np.abs(w32 - ref)):.2e}",
          flush=True)
    assert np.max(np.abs(w64 - ref)) <= 1e-12 * scale
    assert np.max(np.abs(w32 - ref)) <= 2e-5 * scale


@pytest.mark.parametrize("code", range(8))
def test_heisenberg_block_hv_item_orders(ctx, monkeypatch, code):
    """Every block order the tuner can pick (PHUB_HEIS_ORDER forces one: workgroup to
    XCD map and memory or popcount-major block list) gives the oracle's HPROD."""
    from pyhub_amd.engine import HeisenbergOp
    monkeypatch.setenv("PHUB_HEIS_ORDER", str(code))
    n = 20
    rng = np.random.default_rng(46)
    J = rng.normal(size=(n, n))
    J[rng.random((n, n)) < 0.5] = 0.0
    J = (J + J.T) / 2
    op = HeisenbergOp(ctx, n, J)
    m = O.HeisenbergModel(n, J)
    v = rng.normal(size=m.dim)
    ref = m.hv(v, fast=True)
    w = op.hv(dev(v)).cpu().numpy()
    assert np.max(np.abs(w - ref)) <= 1e-12 * max(1.0, np.max(np.abs(ref)))


def test_j1j2_4x4_standard_convention_e0(ctx):
    """J1-J2 on the 4x4 torus (J2/J1 = 0.5) in the standard convention: the
    ordered-pair matrix carries J/2 per pair (sign_convention=1); E0 against the
    sparse S_z=0 matrix of sum_{i<j} J_ij S_i.S_j (scipy eigsh), FP64 plain
    recurrence and FP32 vectors with FP64 scalars."""
    import scipy.sparse.linalg as sla
    from pyhub_amd import _lib as L
    from pyhub_amd.engine import HeisenbergOp
    Jstd = O.j1j2_square(4, 4, 1.0, 0.5)
    e_ref = sla.eigsh(O.heisenberg_standard_sz0(16, Jstd), k=1, which="SA")[0][0]
    Jm = (Jstd + Jstd.T) / 2.0
    op = HeisenbergOp(ctx, 16, Jm, sign_convention=1)
    r64 = op.lanczos(mode=L.LCZ_PLAIN64, max_steps=300, tol=1e-12, want_psi=True)
    r32 = op.lanczos(mode=L.LCZ_PLAIN32, fixed_steps=150, max_steps=150, want_psi=False)
    print(f"E0 ref {e_ref:.12f} plain64 {r64.e0:.12f} plain32 {r32.e0:.12f}", flush=True)
    assert abs(r64.e0 - e_ref) < 1e-10 * abs(e_ref)
    assert r64.residual < 1e-5
    assert abs(r32.e0 - e_ref) < 1e-5 * abs(e_ref)


@pytest.mark.parametrize("Lx,Ly", [(6, 4), (8, 4)])
def test_heisenberg_large_sector_properties(ctx, Lx, Ly):
    """J1-J2 tori of 24 and 32 sites (dim 2,704,156 and 601,080,390, where the
    state index passes 2^29): symmetry <u, H v> = <H u, v>, FP32 vs FP64 H*v,
    and the reduced-quantity kernels refuse a sector without a mask table."""
    from pyhub_amd.engine import HeisenbergOp
    Jstd = O.j1j2_square(Lx, Ly, 1.0, 0.5)
    n = Lx * Ly
    print(f"n={n}", flush=True)
    op = HeisenbergOp(ctx, n, (Jstd + Jstd.T) / 2.0, sign_convention=1)
    g = torch.Generator(device="cuda").manual_seed(5)
    u = torch.randn(op.dim, dtype=torch.float64, device="cuda", generator=g)
    v = torch.randn(op.dim, dtype=torch.float64, device="cuda", generator=g)
    hu, hv = op.hv(u), op.hv(v)
    a, b = float(torch.dot(u, hv)), float(torch.dot(hu, v))
    assert abs(a - b) < 1e-10 * max(1.0, abs(a))
    h32 = op.hv(v.float()).double()
    assert float(torch.max(torch.abs(h32 - hv))) < 2e-5 * float(torch.max(torch.abs(hv)))
    if op.dim > (1 << 28):
        del hu, h32
        with pytest.raises(Exception):
            op.sisj(u, v)


CFG1_E0 = -4.603526299989137


@pytest.mark.parametrize("mode,keep", [(0, None), (0, "4"), (1, None), (2, None), (3, None),
                                       (3, "6")])
def test_config1_lanczos_modes(ctx, mode, keep, monkeypatch):
    """Every Lanczos mode on config 1.  The plain modes keep their vectors and
    combine psi from them; with PHUB_LANCZOS_KEEP capping the vectors kept, psi
    comes from a second pass of the recurrence instead.  PLAIN32 refines E0 in
    FP64 (refine_steps = 2), as MIXED does."""
    from pyhub_amd.engine import HubbardOp
    if keep:
        monkeypatch.setenv("PHUB_LANCZOS_KEEP", keep)
    op = HubbardOp(ctx, 8, 4, 4, O.ring_T(8), 4.0)
    v0 = np.random.default_rng(1).uniform(-0.5, 0.5, op.dim)
    # the FP32 recurrence's Ritz values settle at ~1e-9, not 1e-12
    r = op.lanczos(mode=mode, max_steps=300, tol=1e-12 if mode != 3 else 1e-9, v0=dev(v0),
                   refine_steps=2)
    assert r.converged
    tol = 1e-10 if mode in (0, 1) else 1e-9
    assert abs(r.e0 - CFG1_E0) / abs(CFG1_E0) < tol, (r.e0, r.steps)
    assert r.residual < (1e-6 if mode == 0 else 1e-5)
    m = O.HubbardModel(8, 4, 4, O.ring_T(8), 4.0)
    psi = r.psi.cpu().numpy()
    assert abs(np.linalg.norm(psi) - 1) < 1e-12
    # 1-RDM agrees with the oracle contraction of the same psi and with dense ED
    up, dn, ni, ss = op.rdm(r.psi)
    ro = m.one_rdm(psi)
    assert np.max(np.abs(up - ro[0])) < 1e-12
    ev, V = np.linalg.eigh(m.dense())
    rd = m.one_rdm(V[:, 0])
    # psi error ~ residual / gap (gap 0.3035 above E0)
    assert np.max(np.abs(up - rd[0])) < 1e-10 + 20 * r.residual / 0.3


@pytest.mark.parametrize("mode", [0, 1, 2, 3])
def test_lanczos_workspace(ctx, mode):
    """A caller workspace of phub_lanczos_workspace_bytes holds every Krylov vector;
    a smaller one is accepted (the plain modes keep fewer vectors and rebuild psi by
    a second pass, the stored-basis modes allocate the rest), with the same result."""
    from pyhub_amd.engine import HubbardOp
    op = HubbardOp(ctx, 8, 4, 4, O.ring_T(8), 4.0)
    v0 = dev(np.random.default_rng(1).uniform(-0.5, 0.5, op.dim))
    kw = dict(mode=mode, max_steps=200, tol=1e-12 if mode != 3 else 1e-9, v0=v0, refine_steps=2)
    need = op.lanczos_workspace_bytes(mode, 200, 0, 2, True)
    ld = (op.dim + 63) // 64 * 64
    s = 4 if mode in (2, 3) else 8
    assert need >= 201 * ld * s
    full = op.lanczos(workspace=torch.empty(need, dtype=torch.uint8, device=ctx.device), **kw)
    small = op.lanczos(workspace=torch.empty(5 * ld * 8 + 256, dtype=torch.uint8,
                                             device=ctx.device), **kw)
    auto = op.lanczos(**kw)
    for r in (full, small, auto):
        assert r.converged
    assert full.steps == small.steps == auto.steps
    tol = 1e-12 if mode != 3 else 1e-9
    assert abs(full.e0 - small.e0) < tol and abs(full.e0 - auto.e0) < tol
    ov = abs(float(torch.dot(full.psi, small.psi)))
    assert abs(ov - 1.0) < (1e-10 if mode in (0, 1) else 1e-6)


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_lanczos_excited_states(ctx, mode):
    """n_states = 4 Ritz pairs, the reference's psi of every weighted state
    (Lanczos vectors combined with the tridiagonal eigenvectors,
    compute_solver.f90:271-285):
    - e_states are the lowest eigenvalues of the returned tridiagonal matrix
      (scipy's eigh_tridiagonal; MIXED's e_states[0] is the FP64-refined E0);
    - each psi_k is a unit vector, and the pairs equal dense ED's lowest four
      (eigenvalue and vector up to sign) -- the run continues until all four
      Ritz values are stable, so none of them is left unconverged."""
    from scipy.linalg import eigh_tridiagonal
    from pyhub_amd.engine import HubbardOp
    T = O.disordered_ring_T(6)      # (3 up, 2 down): 300 states, non-degenerate
    op = HubbardOp(ctx, 6, 3, 2, T, 2.5)
    m = O.HubbardModel(6, 3, 2, T, 2.5)
    ev, V = np.linalg.eigh(m.dense())
    r = op.lanczos(mode=mode, max_steps=250, tol=1e-12 if mode != 2 else 1e-9, n_states=4,
                   refine_steps=2)
    assert r.converged
    k = r.steps
    th = eigh_tridiagonal(r.alpha, r.beta[1:k], eigvals_only=True)
    assert r.psis.shape == (4, op.dim) and len(r.e_states) >= 4
    lo = 1 if mode == 2 else 0
    assert np.max(np.abs(r.e_states[lo:4] - th[lo:4])) < 1e-10
    etol, vtol = (1e-9, 1e-9) if mode != 2 else (1e-5, 1e-5)
    for q in range(4):
        psi = r.psis[q].cpu().numpy()
        assert abs(np.linalg.norm(psi) - 1.0) < 1e-12
        assert abs(r.e_states[q] - ev[q]) < etol, (q, r.e_states[q], ev[q])
        assert abs(abs(psi @ V[:, q]) - 1.0) < vtol, (q, psi @ V[:, q])
    # the n_states = 1 path is unchanged: same ground state as the 4-state run
    r1 = op.lanczos(mode=mode, max_steps=250, tol=1e-12 if mode != 2 else 1e-9, refine_steps=2)
    assert abs(r1.e0 - ev[0]) < etol
    assert r1.psis.shape == (1, op.dim)


@pytest.mark.parametrize("pairmap", ["1", "0"])
@pytest.mark.parametrize("n,nu,nd,seed", [(6, 3, 2, 20), (7, 3, 3, 21), (8, 4, 3, 22), (6, 3, 3, 23),
                                          (10, 5, 4, 24), (18, 2, 1, 25)])
def test_rdm_transition_vs_oracle(ctx, monkeypatch, n, nu, nd, seed, pairmap):
    """ONE_RDM, NI and J_SPIN for bra != ket (all loops of static_rq.f90); the
    spin-flip term through the down-spin pair map and through per-row ranks
    (PHUB_RDM_PAIRMAP=0, the path for maps beyond 1 GiB).  The on-device
    assembly of the diagonal sums runs over chunks of 256 up-then-down states:
    10 sites (252 + 210 states: the up/down boundary inside the first chunk, two
    chunks) and 18 sites (the branch that keeps its tables in global memory)."""
    from pyhub_amd.engine import HubbardOp
    monkeypatch.setenv("PHUB_RDM_PAIRMAP", pairmap)
    rng = np.random.default_rng(seed)
    T, U = rand_cluster(rng, n)
    op = HubbardOp(ctx, n, nu, nd, T, U)
    m = O.HubbardModel(n, nu, nd, T, U)
    a = rng.normal(size=m.dim)
    b = rng.normal(size=m.dim)
    up, dn, ni, ss = op.rdm(dev(a), dev(b))
    ro = m.one_rdm(a, b)
    assert np.max(np.abs(up - ro[0])) < 1e-10
    assert np.max(np.abs(dn - ro[1])) < 1e-10
    assert np.max(np.abs(ni - m.ni(a, b))) < 1e-10
    assert np.max(np.abs(ss - m.sisj(a, b))) < 1e-10


@pytest.mark.parametrize("mode", [0, 3])
def test_plain_recurrence_excited_states_skip_ghosts(ctx, mode):
    """The three-term recurrence run far past convergence (200 fixed steps on 300
    states) makes copies of its converged Ritz values; the excited states skip them:
    the three lowest energies match the oracle's dense spectrum, and each returned
    state is an eigenvector of the oracle's matrix (FP64: residual 1e-6; FP32 vectors
    with FP64 scalars: 1e-3)."""
    from pyhub_amd.engine import HubbardOp
    T = O.disordered_ring_T(6)
    op = HubbardOp(ctx, 6, 3, 2, T, 2.5)
    m = O.HubbardModel(6, 3, 2, T, 2.5)
    H = m.dense()
    e = np.linalg.eigvalsh(H)
    r = op.lanczos(mode=mode, fixed_steps=200, max_steps=200, n_states=3, want_psi=True, seed=3)
    print(f"mode {mode}: e_states {r.e_states} dense {e[:3]}", flush=True)
    tol_e, tol_r = (1e-9, 1e-6) if mode == 0 else (1e-5, 1e-3)
    assert np.max(np.abs(np.asarray(r.e_states[:3]) - e[:3])) < tol_e
    for k in range(3):
        p = r.psis[k].cpu().numpy()
        p = p / np.linalg.norm(p)
        assert np.linalg.norm(H @ p - e[k] * p) < tol_r, k


def test_jd_correction_preconditioned(ctx, monkeypatch):
    """FP64 corrections of the FP32 ground state with the Jacobi-preconditioned inner
    MINRES (the default) and without it (PHUB_JD_PRECOND=0), on the 4x3 t-t' torus at
    half filling (853,776 states): both reach the residual target, agree with the FP64
    three-term recurrence on E0 to 1e-12 relative and on the 1-RDM, double occupancy
    and <S.S> to 1e-8, and the preconditioned solve takes fewer inner iterations."""
    from pyhub_amd import _lib as L
    from pyhub_amd.engine import HubbardOp
    T = O.square_T(4, 3, t=1.0, tp=-0.2, periodic=True)
    op = HubbardOp(ctx, 12, 6, 6, T, 8.0)
    v0 = dev(np.random.default_rng(9).uniform(-0.5, 0.5, op.dim))
    out = {}
    for name, mode, pc in (("fp64", L.LCZ_PLAIN64, "1"), ("pc", L.LCZ_PLAIN32, "1"),
                           ("nopc", L.LCZ_PLAIN32, "0")):
        monkeypatch.setenv("PHUB_JD_PRECOND", pc)
        r = op.lanczos(mode=mode, max_steps=400, tol=1e-9, v0=v0, refine_steps=4, psi_tol=1e-10)
        out[name] = (r.e0, r.residual, r.correction_iters, op.rdm(r.psi))
        print(f"4x3 {name}: E0 {r.e0:.13f} residual {r.residual:.1e} {r.corrections} corrections "
              f"({r.correction_iters} inner)", flush=True)
        del r
    e64 = out["fp64"][0]
    for name in ("pc", "nopc"):
        e, res, _, q = out[name]
        assert res <= 1e-9
        assert abs(e - e64) <= 1e-12 * abs(e64)
        assert max(float(np.max(np.abs(a - b))) for a, b in zip(q, out["fp64"][3])) <= 1e-8
    assert 0 < out["pc"][2] < out["nopc"][2]


@pytest.mark.parametrize("n,seed", [(6, 30), (8, 31)])
def test_heisenberg_sisj_vs_oracle(ctx, n, seed):
    from pyhub_amd.engine import HeisenbergOp
    rng = np.random.default_rng(seed)
    J = np.zeros((n, n))
    for i in range(n):
        J[i, (i + 1) % n] = J[(i + 1) % n, i] = 1.0
    op = HeisenbergOp(ctx, n, J)
    m = O.HeisenbergModel(n, J)
    a = rng.normal(size=m.dim)
    b = rng.normal(size=m.dim)
    assert np.max(np.abs(op.sisj(dev(a), dev(b)) - m.sisj(a, b))) < 1e-10


def test_band_lanczos_exact_and_moments(ctx):
    """Band Lanczos (GF_BANDLANCZOS) from the p excitations c_i psi0:
    - run past the size of a 225-state sector (the recurrence deflates), every
      G_ij(omega) equals dense ED;
    - on a 52,920-state sector the zeroth and first moments are exact at any step
      count: sum_k Q_ik Q_jk = <v_i|v_j>, sum_k Q_ik Q_jk E_k = <v_i|H|v_j>."""
    from pyhub_amd.engine import HubbardOp
    T = O.ring_T(6)
    T[0, 3] = T[3, 0] = -0.3
    T[2, 2] = 0.1
    op = HubbardOp(ctx, 6, 3, 2, T, 3.0)
    m = O.HubbardModel(6, 3, 2, T, 3.0)
    ev, V = np.linalg.eigh(m.dense())
    psi = dev(V[:, 0])
    xs, tgt = [], None
    for site in range(1, 7):
        tgt, x = op.excite(psi, site, 0, False)
        xs.append(x)
    Vs = torch.stack(xs)
    # tol 0: run to max_steps - 10 = 390 steps, past the 225 states (deflated tail)
    pos, Q = tgt.band_lanczos(Vs, max_steps=400, tol=0.0)
    mt = O.HubbardModel(6, 2, 2, T, 3.0)
    et, Ut = np.linalg.eigh(mt.dense())
    Qd = Ut.T @ Vs.cpu().numpy().T          # (poles, sites)
    z = np.linspace(-8, 8, 81) + 0.3j
    Gb = np.einsum("ki,kj,wk->wij", Q, Q, 1.0 / (z[:, None] - pos[None, :]))
    Gd = np.einsum("ki,kj,wk->wij", Qd, Qd, 1.0 / (z[:, None] - et[None, :]))
    # no reorthogonalisation in the recurrence (as the reference's): ~1e-9 drift
    assert np.max(np.abs(Gb - Gd)) < 1e-8
    # moments on a larger sector
    op = HubbardOp(ctx, 10, 5, 5, O.ring_T(10), 4.0)
    r = op.lanczos(mode=1, max_steps=300, tol=1e-12)
    xs = []
    for site in range(1, 11):
        tgt, x = op.excite(r.psi, site, 1, False)
        xs.append(x)
    Vs = torch.stack(xs)
    pos, Q = tgt.band_lanczos(Vs, max_steps=120, tol=1e-12)
    S0 = (Vs @ Vs.T).cpu().numpy()
    HV = torch.stack([tgt.hv(v.contiguous()) for v in Vs])
    S1 = (Vs @ HV.T).cpu().numpy()
    assert np.max(np.abs(Q.T @ Q - S0)) < 1e-10
    assert np.max(np.abs((Q.T * pos) @ Q - S1)) < 1e-9 * max(1.0, np.max(np.abs(S1)))


def _rand_u4(rng, n, density=0.15):
    U = rng.normal(size=(n, n, n, n))
    U[rng.random((n, n, n, n)) > density] = 0.0
    U = 0.5 * (U + np.einsum("ijkl->jilk", U))
    for J in range(n):
        U[J, J, J, J] = rng.uniform(1, 4)
    return U


@pytest.mark.parametrize("n,nu,nd,seed,pull", [(6, 3, 2, 0, False), (6, 3, 3, 1, False),
                                               (7, 2, 4, 2, False), (6, 3, 2, 0, True),
                                               (7, 2, 4, 2, True)])
def test_hubbard_4index_u_vs_oracle(ctx, n, nu, nd, seed, pull, monkeypatch):
    """is_Uloc = 3: H*v (FP64 and FP32) and DBL_OCC against the oracle's HPROD /
    TWO_RDM restatements, and the Lanczos ground energy against dense ED; both the
    row-staged kernel and the one-lane-per-target kernel (PHUB_UL_PULL=1, the form
    used when the rows do not fit in LDS)."""
    from pyhub_amd import _lib as L
    from pyhub_amd.engine import HubbardOp
    if pull:
        monkeypatch.setenv("PHUB_UL_PULL", "1")
    rng = np.random.default_rng(80 + seed)
    T, _ = rand_cluster(rng, n)
    U4 = _rand_u4(rng, n)
    op = HubbardOp(ctx, n, nu, nd, T, U4)
    m = O.HubbardModel(n, nu, nd, T, U4)
    v = rng.normal(size=m.dim)
    ref = m.hv(v)
    w = op.hv(dev(v)).cpu().numpy()
    assert np.max(np.abs(w - ref)) <= 1e-12 * max(1.0, np.max(np.abs(ref)))
    w32 = op.hv(dev(v, torch.float32)).double().cpu().numpy()
    assert np.max(np.abs(w32 - ref)) <= 2e-5 * max(1.0, np.max(np.abs(ref)))
    a, b = rng.normal(size=m.dim), rng.normal(size=m.dim)
    D = op.two_rdm(dev(a), dev(b))
    assert np.max(np.abs(D - m.two_rdm(a, b))) < 1e-11
    e_ref = np.linalg.eigvalsh(m.dense())[0]
    r = op.lanczos(mode=L.LCZ_FULL64, max_steps=min(300, op.dim), tol=1e-13)
    assert abs(r.e0 - e_ref) < 1e-10 * max(1.0, abs(e_ref))


@pytest.mark.parametrize("site,spin,create", [(1, 0, 1), (3, 0, 0), (2, 1, 1), (6, 1, 0)])
def test_excitation_and_cf_chain_vs_oracle(ctx, site, spin, create):
    from pyhub_amd.engine import HubbardOp
    T = O.ring_T(6)
    T[0, 2] = T[2, 0] = 0.3
    op = HubbardOp(ctx, 6, 3, 2, T, 3.0)
    m = O.HubbardModel(6, 3, 2, T, 3.0)
    ev, V = np.linalg.eigh(m.dense())
    psi = V[:, 0]
    tgt, out = op.excite(dev(psi), site, spin, bool(create))
    mt, ref = m.excite(psi, site, spin, create)
    assert np.array_equal(out.cpu().numpy(), ref)
    a, b, n2 = tgt.cf_chain(out, 12, precision=64)
    ao, bo, n2o = mt.cf_chain(ref, 12)
    assert abs(n2 - n2o) < 1e-13
    k = 8   # before loss of orthogonality matters
    assert np.max(np.abs(a[:k] - ao[:k])) < 1e-9 and np.max(np.abs(b[:k] - bo[:k])) < 1e-9


@pytest.mark.parametrize("n,nu,nd", [(4, 0, 2), (4, 4, 1), (1, 1, 1), (5, 0, 0), (3, 3, 3),
                                     (6, 1, 0), (2, 1, 1)])
def test_edge_sectors(ctx, n, nu, nd):
    """Empty and full species, one-site and one-state sectors: H*v (FP64, FP32) against
    the oracle, and the plain / full Lanczos ground energy against dense ED (a
    one-state sector breaks down after one step with E0 = its diagonal)."""
    from pyhub_amd.engine import HubbardOp
    T = O.ring_T(n) if n > 2 else -np.ones((n, n)) + np.eye(n)
    T[0, 0] = 0.3
    op = HubbardOp(ctx, n, nu, nd, T, 2.0)
    m = O.HubbardModel(n, nu, nd, T, 2.0)
    assert op.dim == m.dim
    v = np.random.default_rng(n + 10 * nu + 100 * nd).normal(size=m.dim)
    ref = m.hv(v)
    scale = max(1.0, np.max(np.abs(ref)))
    assert np.max(np.abs(op.hv(dev(v)).cpu().numpy() - ref)) <= 1e-13 * scale
    assert np.max(np.abs(op.hv(dev(v, torch.float32)).double().cpu().numpy() - ref)) <= 2e-5 * scale
    e_ref = np.linalg.eigvalsh(m.dense())[0]
    for mode in (0, 1):
        # the run ends when the space is exhausted (dim steps); the plain recurrence
        # may get there before its every-5-steps test has seen two equal Ritz values
        # (the reference solver takes the dense path on such sectors)
        r = op.lanczos(mode=mode, max_steps=50, tol=1e-12)
        assert r.converged or (mode == 0 and r.steps == op.dim)
        assert abs(r.e0 - e_ref) < 1e-10 * max(1.0, abs(e_ref)), (mode, r.e0, e_ref)
        assert abs(np.linalg.norm(r.psi.cpu().numpy()) - 1.0) < 1e-12
