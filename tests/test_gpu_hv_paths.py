"""GPU tests of the up-hop kernel paths of the two-kernel Hubbard H*v against the
oracle (hprod_hubbard.f90:17-118 restated, bit-exact against the reference):

- lattice hoppings (few signed values) and random hoppings (many); a sector whose
  rows do not fit LDS (the fallback kernel);
- FP64 with two rows per workgroup where they fit (dim_up <= 10,048), and with one
  row: forced by PHUB_UP_ROWS=1 on small sectors, and on a sector whose up
  dimension (11,440) is past the two-row limit, over every down-hop tile;
- FP32 with three rows per workgroup where they fit (dim_up <= 13,397; every FP32
  H*v of the other tests), and with two: forced by PHUB_UP_ROWS=2;
- 31 and 32 sites (every bit of the 32-bit state masks).

Tolerances: FP64 within 1e-13 of the oracle relative to max|w|; FP32 within 2e-5."""
import numpy as np
import pytest
import torch

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def dev(x, dtype=torch.float64):
    return torch.as_tensor(np.asarray(x), dtype=dtype, device="cuda")


def _check(op, m, rng):
    v = rng.normal(size=m.dim)
    ref = m.hv(v, fast=True, nthreads=8)
    sc = max(1.0, np.max(np.abs(ref)))
    w64 = op.hv(dev(v)).cpu().numpy()
    assert np.max(np.abs(w64 - ref)) <= 1e-13 * sc
    w32 = op.hv(dev(v, torch.float32)).double().cpu().numpy()
    assert np.max(np.abs(w32 - ref)) <= 2e-5 * sc


def _lattice(n, rng, tp=0.2, eps=True):
    """A ring (n not a square) or torus with t = 1, t' on the next-nearest bonds
    and random on-site energies: four signed hopping values at most."""
    if n in (9, 16):
        L = int(round(np.sqrt(n)))
        T = O.square_T(L, L, t=1.0, tp=-tp, periodic=True)
    else:
        T = O.ring_T(n)
        for i in range(n):
            j = (i + 2) % n
            T[i, j] = T[j, i] = tp
    if eps:
        T[np.arange(n), np.arange(n)] = rng.normal(0, 0.4, n)
    return T


@pytest.mark.parametrize("n,nu,nd,seed", [(8, 4, 4, 0), (9, 4, 5, 1), (10, 3, 6, 2), (6, 0, 3, 3),
                                          (12, 6, 2, 4), (16, 8, 1, 5)])
def test_lattice_hoppings(ctx, n, nu, nd, seed):
    from pyhub_amd.engine import HubbardOp
    rng = np.random.default_rng(seed)
    T = _lattice(n, rng)
    U = rng.uniform(2, 6, n)
    op = HubbardOp(ctx, n, nu, nd, T, U)
    _check(op, O.HubbardModel(n, nu, nd, T, U), rng)


def test_lattice_hoppings_large_up_sector(ctx):
    """18 sites, 8 up (43,758 states): one row (171 KB in FP32) does not fit LDS, so
    the H*v runs the one-lane-per-state fallback kernel (k_hv_hub)."""
    from pyhub_amd.engine import HubbardOp
    rng = np.random.default_rng(11)
    n, nu, nd = 18, 8, 1
    T = _lattice(n, rng)
    op = HubbardOp(ctx, n, nu, nd, T, 4.0)
    _check(op, O.HubbardModel(n, nu, nd, T, 4.0), rng)


@pytest.mark.parametrize("lattice", [True, False])
def test_fp64_one_row_forced(ctx, monkeypatch, lattice):
    """PHUB_UP_ROWS=1: the FP64 one-row up kernel on sectors where two rows fit."""
    from pyhub_amd.engine import HubbardOp
    monkeypatch.setenv("PHUB_UP_ROWS", "1")
    rng = np.random.default_rng(21)
    n, nu, nd = 10, 5, 4
    if lattice:
        T = _lattice(n, rng)
    else:
        T = rng.normal(size=(n, n))
        T = (T + T.T) / 2
    U = rng.uniform(2, 6, n)
    op = HubbardOp(ctx, n, nu, nd, T, U)
    _check(op, O.HubbardModel(n, nu, nd, T, U), rng)


@pytest.mark.parametrize("n,nu,nd,seed", [(7, 3, 3, 31), (13, 6, 1, 32)])
def test_fp32_two_rows_forced(ctx, monkeypatch, n, nu, nd, seed):
    """PHUB_UP_ROWS=2: FP32 on two rows per up-kernel workgroup where three fit (the
    path of up sectors between 13,398 and 18,944 states), random hoppings, odd down
    dimensions (35 and 13) that leave the last workgroup one row."""
    from pyhub_amd.engine import HubbardOp
    monkeypatch.setenv("PHUB_UP_ROWS", "2")
    rng = np.random.default_rng(seed)
    T = rng.normal(size=(n, n))
    T = (T + T.T) / 2
    U = rng.uniform(2, 6, n)
    op = HubbardOp(ctx, n, nu, nd, T, U)
    _check(op, O.HubbardModel(n, nu, nd, T, U), rng)


@pytest.mark.parametrize("tile", list(range(25)))
@pytest.mark.parametrize("lattice", [True, False])
def test_fp64_one_row_large_up_sector_all_tiles(ctx, monkeypatch, tile, lattice):
    """n = 16, 7 up and 1 down (dim_up = 11,440: two FP64 rows do not fit LDS, so
    FP64 runs the one-row kernel, the config-2 FP64 path), every down tile."""
    from pyhub_amd.engine import HubbardOp
    monkeypatch.setenv("PHUB_DN_TILE", str(tile))
    rng = np.random.default_rng(300 + tile)
    n, nu, nd = 16, 7, 1
    if lattice:
        T = _lattice(n, rng)
    else:
        T = rng.normal(size=(n, n)) * 0.5
        T = (T + T.T) / 2
    U = rng.uniform(2, 6, n)
    op = HubbardOp(ctx, n, nu, nd, T, U)
    _check(op, O.HubbardModel(n, nu, nd, T, U), rng)



@pytest.mark.parametrize("n,nu,nd,seed", [(31, 2, 1, 41), (32, 2, 1, 42), (32, 4, 1, 43)])
def test_32_bit_masks(ctx, n, nu, nd, seed):
    """Up to 32 sites, every bit of the 32-bit state masks: 2 up and 1 down electron
    (the two-kernel path) and 4 up and 1 down (35,960 up states, rows too long for LDS
    in either precision: the one-lane-per-state kernel), random hoppings, against the
    oracle (pinned at 32 sites by second-quantised spectra)."""
    from pyhub_amd.engine import HubbardOp
    rng = np.random.default_rng(seed)
    T = rng.normal(size=(n, n)) * 0.3
    T = (T + T.T) / 2
    U = rng.uniform(2, 6, n)
    op = HubbardOp(ctx, n, nu, nd, T, U)
    _check(op, O.HubbardModel(n, nu, nd, T, U), rng)


def test_32_sites_ground_state(ctx):
    """E0 of 1 up and 1 down electron on 32 sites: the lowest eigenvalue of
    T (x) 1 + 1 (x) T + U_i delta_ij, through the FP64 Lanczos."""
    from pyhub_amd import _lib as L
    from pyhub_amd.engine import HubbardOp
    rng = np.random.default_rng(44)
    n = 32
    T = rng.normal(size=(n, n)) * 0.3
    T = (T + T.T) / 2
    U = rng.uniform(2, 6, n)
    Hk = np.kron(T, np.eye(n)) + np.kron(np.eye(n), T) + np.diag(
        [U[i] if i == j else 0.0 for i in range(n) for j in range(n)])
    e_ref = np.linalg.eigvalsh(Hk)[0]
    op = HubbardOp(ctx, n, 1, 1, T, U)
    r = op.lanczos(mode=L.LCZ_FULL64, max_steps=300, tol=1e-12)
    print(f"32 sites (1,1): E0 {r.e0:.13f} exact {e_ref:.13f}", flush=True)
    assert abs(r.e0 - e_ref) < 1e-10 * max(1.0, abs(e_ref))
