"""GPU tests of the generic operator engine (Operators, tools/operators.py,
tools/models.py) against the operators.x restatement in the oracle, modelled on the
reference's examples 02 and 07."""
import numpy as np
import pytest
import torch

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _random_terms(rng, n, nterms, maxlen):
    """Random products of every factor code (sites 1..n, both spins)."""
    terms = []
    for _ in range(nterms):
        fs = []
        for _ in range(rng.integers(1, maxlen + 1)):
            code = int(rng.choice([1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11]))
            s1, s2 = (int(x) for x in rng.integers(1, n + 1, size=2))
            p1, p2 = (int(x) for x in rng.integers(1, 3, size=2))
            if code == 9 and p1 == p2:
                p1 = 3 - p2
            fs.append((code, s1, p1, s2, p2))
        terms.append((float(rng.normal()), fs))
    return terms


def _program(ctx, n, nup, ndn, terms):
    from pyhub_amd.core._operators import _Compiled
    return _Compiled(ctx, n, nup, ndn, [c for c, _ in terms], [f for _, f in terms])


@pytest.mark.parametrize("n,nup,ndn,seed", [(3, [1], [2], 0), (4, [1, 2], [2, 1], 1),
                                            (4, [0, 1, 2, 3, 4], [0, 1, 2, 3, 4], 2),
                                            (5, [2], [2, 3], 3)])
def test_operator_programs_vs_oracle(ctx, n, nup, ndn, seed):
    """Random sums of products of all eleven factor codes on single- and
    multi-sector bases: O v (vector path) and the to_matrix rows / diagonal
    (basis-element path) equal the oracle's restatement to 1e-12."""
    rng = np.random.default_rng(seed)
    terms = _random_terms(rng, n, 24, 4)
    # one product of S.S with itself and a zero-coefficient term (skipped)
    terms.append((0.7, [(8, 1, 0, 2, 0), (8, 2, 0, 2, 0)]))
    terms.append((0.0, [(2, 1, 1, 0, 0)]))
    oc = O.OperatorOracle(n, nup, ndn)
    prog = _program(ctx, n, nup, ndn, terms)
    assert prog.nstates == oc.dim
    V = rng.normal(size=(3, oc.dim))
    out = prog.apply(torch.as_tensor(V)).cpu().numpy()
    for k in range(3):
        ref = oc.apply(terms, V[k])
        assert np.max(np.abs(out[k] - ref)) < 1e-12 * max(1.0, np.max(np.abs(ref)))
    M = oc.matrix(terms)
    assert np.max(np.abs(prog.matrix() - M)) < 1e-12 * max(1.0, np.max(np.abs(M)))
    assert np.max(np.abs(prog.diag() - np.diag(M))) < 1e-12 * max(1.0, np.max(np.abs(M)))


@pytest.mark.parametrize("maps", ["1", "0"])
def test_operator_walk_and_map_paths(ctx, monkeypatch, maps):
    """O v through the per-species source maps (the default) and through the
    in-kernel factor walk (PHUB_OP_MAPS=0, the path for lists beyond 2^30 elements or
    maps beyond 1 GiB) on a multi-sector basis: hops, densities, c / c^dag and S.S
    products (mixed-species factors included), against the oracle to 1e-12; both
    paths on an odd number of single-species products and a restricted basis."""
    monkeypatch.setenv("PHUB_OP_MAPS", maps)
    rng = np.random.default_rng(7)
    n, nup, ndn = 5, [1, 2, 3], [2, 3]
    terms = _random_terms(rng, n, 31, 3)
    terms += [(0.3, [(1, 2, 1, 4, 1)]), (-0.8, [(1, 5, 2, 1, 2), (2, 3, 1, 0, 0)]),
              (1.1, [(10, 2, 2, 0, 0), (11, 4, 2, 0, 0)]), (0.4, [(8, 1, 0, 3, 0)])]
    oc = O.OperatorOracle(n, nup, ndn)
    prog = _program(ctx, n, nup, ndn, terms)
    V = rng.normal(size=(2, oc.dim))
    out = prog.apply(torch.as_tensor(V)).cpu().numpy()
    for k in range(2):
        ref = oc.apply(terms, V[k])
        assert np.max(np.abs(out[k] - ref)) < 1e-12 * max(1.0, np.max(np.abs(ref)))
    single = [(0.5, [(1, 1, 1, 2, 1)]), (0.25, [(2, 3, 2, 0, 0)]), (-1.5, [(1, 4, 2, 2, 2)])]
    oc1 = O.OperatorOracle(4, [2], [1, 2])
    prog1 = _program(ctx, 4, [2], [1, 2], single)
    v = rng.normal(size=oc1.dim)
    ref = oc1.apply(single, v)
    assert np.max(np.abs(prog1.apply(torch.as_tensor(v)).cpu().numpy() - ref)) < 1e-12


def test_hubbard_operator_api_vs_solver(ctx):
    """fermi_hubbard(t, U) as an operator sum (tools/models.py:5-11) on Basis(6,
    (3, 2)): to_matrix equals the pinned Hubbard matrix, O @ v equals the H*v
    kernel, the operator's own Lanczos (_operators.py:325-366) reaches the dense
    ground energy, avg and the complex call follow the reference."""
    from pyhub_amd.core.basis import Basis
    from pyhub_amd.engine import HubbardOp
    from pyhub_amd.tools.models import fermi_hubbard
    T = O.disordered_ring_T(6)
    U = 2.5
    H = fermi_hubbard(T, U)
    H.set_basis(Basis(6, hilbert=(3, 2)))
    m = O.HubbardModel(6, 3, 2, T, np.full(6, U))
    Hd = m.dense()
    assert np.max(np.abs(H.to_matrix - Hd)) < 1e-12
    rng = np.random.default_rng(0)
    v = rng.normal(size=H.nstates)
    hv = HubbardOp(ctx, 6, 3, 2, T, U).hv(torch.as_tensor(v, device=ctx.device)).cpu().numpy()
    assert np.max(np.abs(H @ v - hv)) < 1e-12
    assert abs(H.avg(v) - v @ Hd @ v) < 1e-9
    w = v + 1j * rng.normal(size=H.nstates)
    assert np.max(np.abs(H(w) - (Hd @ w.real + 1j * (Hd @ w.imag)))) < 1e-12
    np.random.seed(4)
    ek, psi = H.lanczos(maxstep=200, acc_lcz=1e-10)
    e_ref = np.linalg.eigvalsh(Hd)[0]
    assert abs(ek[0] - e_ref) < 1e-8
    assert abs(abs(np.linalg.norm(psi[:, 0])) - 1.0) < 1e-6
    assert np.max(np.abs(Hd @ psi[:, 0] - ek[0] * psi[:, 0])) < 1e-4


def test_counting_operator_and_restriction(ctx):
    """Example 02 (N on a 2-site (1,1) basis: avg = 2) and the example 07 flow:
    the comparisons list basis elements up-index-major, and H[idx] acts on the
    selected elements (its to_matrix is the submatrix)."""
    from itertools import product
    from pyhub_amd.core.basis import Basis
    from pyhub_amd.tools.models import heisenberg
    from pyhub_amd.tools.operators import n, opesum
    N = opesum([n((i, s)) for i, s in product(range(2), ("up", "down"))])
    N.set_basis(Basis(2, hilbert=(1, 1)))
    v = np.random.default_rng(10).random(N.nstates) - 0.5
    v /= np.linalg.norm(v)
    assert abs(N.avg(v) - 2.0) < 1e-12
    assert np.max(np.abs(N.to_matrix - 2 * np.eye(4))) < 1e-14
    nb = 4
    J = np.diag(np.ones(nb - 1), 1) + np.diag(np.ones(nb - 1), -1)
    J[0, -1] = J[-1, 0] = 1.0
    B = Basis(nb, hilbert=(2, 2))
    H = heisenberg(J)
    H.set_basis(B)
    N2 = opesum([n((i, "u")) * n((i, "d")) for i in range(nb)])
    N2.set_basis(B)
    eq = N2 == 0
    order = [d * B.nsup + u for u in range(B.nsup) for d in range(B.nsdown)]
    diag = np.diag(N2.to_matrix)
    assert eq == [bool(np.isclose(diag[s], 0.0)) for s in order]
    idx = np.where(eq)[0]
    Hs = H[idx]
    assert np.max(np.abs(Hs.to_matrix - H.to_matrix[np.ix_(idx, idx)])) < 1e-14
    x = np.random.default_rng(1).normal(size=len(idx))
    full = np.zeros(H.nstates)
    full[idx] = x
    assert np.max(np.abs(Hs @ x - (H @ full)[idx])) < 1e-12
    assert isinstance(H.bit_repr(), str) and H.bit_repr().count("\n") == 8 + H.nstates


def test_operator_errors(ctx):
    """An operator that names a site beyond the basis, an unknown factor string, an
    application before set_basis and overlapping in / out blocks are rejected."""
    from pyhub_amd.core.basis import Basis
    from pyhub_amd.core._operators import Operators
    from pyhub_amd.tools.operators import n
    op = n((5, "up"))
    with pytest.raises(AttributeError):
        op(np.zeros(4))
    with pytest.raises(ValueError):
        op.set_basis(Basis(3, hilbert=(1, 1)))
    bad = Operators((0, "u"))
    bad.str_index = [["foo"]]
    with pytest.raises((ValueError, NotImplementedError)):
        bad.set_basis(Basis(2, hilbert=(1, 1)))
    # overlapping input and output blocks are rejected before any launch
    import ctypes as C
    prog = _program(ctx, 3, [1], [1], [(1.0, [(1, 1, 1, 2, 1)])])
    buf = torch.zeros(3 * prog.nstates, dtype=torch.float64, device=ctx.device)
    with pytest.raises(ValueError):
        ctx.check(ctx.lib.phub_operator_apply(ctx.h, prog.fock, prog.prog, C.c_void_p(buf.data_ptr()),
                                              C.c_void_p(buf.data_ptr() + 8 * prog.nstates), 2))


def test_opeexp_and_sisj_helpers(ctx):
    """opeexp: exp(op) psi of the Taylor partial sums against scipy's expm of the
    operator matrix (series converged to ~ftol), the matrix form row by row, the
    unitary stopping rule on an anti-Hermitian exponent, and ValueError when 99
    terms do not settle.  sisj: the operator's expectation equals the oracle's J_SPIN
    (pinned to the reference) for i != j, and n_up (1 - n_dn) + Sz Sz on one site."""
    import scipy.linalg as sla
    from pyhub_amd.core.basis import Basis
    from pyhub_amd.tools.models import fermi_hubbard
    from pyhub_amd.tools.operators import opeexp, sisj
    n_s = 4
    T = O.ring_T(n_s)
    H = fermi_hubbard(T, 3.0)
    H.set_basis(Basis(n_s, hilbert=(2, 1)))
    A = -0.2 * H
    A.set_basis(Basis(n_s, hilbert=(2, 1)))
    rng = np.random.default_rng(4)
    psi = rng.normal(size=A.nstates)
    psi /= np.linalg.norm(psi)
    M = A.to_matrix
    want = sla.expm(M) @ psi
    got = opeexp(A, psi, ftol=1e-14)
    assert np.max(np.abs(got - want)) < 1e-10
    E = opeexp(A, None)
    assert E.shape == (A.nstates, A.nstates)
    assert np.max(np.abs(E - sla.expm(M).T)) < 1e-5          # rows exp(op) e_i, default ftol
    B = 20.0 * H
    B.set_basis(Basis(n_s, hilbert=(2, 1)))
    with pytest.raises(ValueError):
        opeexp(B, psi, ftol=1e-300)
    m = O.HubbardModel(n_s, 2, 1, T, 3.0)
    ev, V = np.linalg.eigh(m.dense())
    g = V[:, 0]
    ss = m.sisj(g)
    for i, j in ((0, 1), (1, 3), (2, 2)):
        op = sisj((i,), (j,))
        op.set_basis(Basis(n_s, hilbert=(2, 1)))
        val = op.avg(g)
        if i != j:
            assert abs(val - ss[i, j]) < 1e-12
        else:
            nu = m.one_rdm(g)[0][i, i]
            assert abs(val - (nu - m.ni(g)[i]) - 0.25 * (nu + m.one_rdm(g)[1][i, i] - 2 * m.ni(g)[i])) < 1e-12
