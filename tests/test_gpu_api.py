"""GPU tests of the PyHub-compatible API (FermiHubbard / Heisenberg / CoreSolver),
modelled on the reference's examples (examples/04,05,07,08,09)."""
import numpy as np
import pytest
import torch

from oracle import oracle as O

pytestmark = pytest.mark.gpu

CFG1_E0 = -4.603526299989137


def test_fermi_hubbard_config1_lanczos_path(ctx):
    from pyhub_amd.solver.fermi_hubbard import FermiHubbard
    T = O.ring_T(8)
    fh = FermiHubbard(8, 4, 4, T, U=4.0)
    fh.start_vector = np.random.default_rng(1).uniform(-0.5, 0.5, 4900)
    fh.kernel(max_lcz=600, acc_lcz=1e-12)
    assert fh.is_lanczos
    assert abs(fh.e0 - CFG1_E0) < 1e-10 * abs(CFG1_E0)
    m = O.HubbardModel(8, 4, 4, T, 4.0)
    # the GPU quantities equal the oracle's contractions of the same psi ...
    psi = fh.psi0
    ro = m.one_rdm(psi)
    assert np.max(np.abs(fh.one_rdm["up"] - ro[0])) < 1e-12
    assert np.max(np.abs(fh.one_rdm["down"] - ro[1])) < 1e-12
    assert np.max(np.abs(fh.ni - m.ni(psi))) < 1e-12
    assert np.max(np.abs(fh.spin_cor_matrix - m.sisj(psi))) < 1e-12
    # ... and the dense-ED ground state's to the accuracy of psi
    ev, V = np.linalg.eigh(m.dense())
    rd = m.one_rdm(V[:, 0])
    assert np.max(np.abs(fh.one_rdm["up"] - rd[0])) < 1e-6
    assert abs(fh.s2) < 1e-6     # singlet
    # Hellmann-Feynman identity E0 = <T> + <U n n>
    assert abs(fh.e_one_body + np.sum(fh.ni * 4.0) - fh.e0) < 1e-9
    assert fh.lcz_results["step"] == fh.nb_lcz


def test_fermi_hubbard_dense_path_and_green_functions(ctx):
    from pyhub_amd.solver.fermi_hubbard import FermiHubbard
    T = O.ring_T(6)
    T[0, 3] = T[3, 0] = -0.3
    fh = FermiHubbard(6, 3, 2, T, U=3.0)
    fh.kernel(max_lcz=600, compute_spgf=True)      # dim 300 < 600 -> EXACT_DIAG path
    assert not fh.is_lanczos
    m = O.HubbardModel(6, 3, 2, T, 3.0)
    ev, V = np.linalg.eigh(m.dense())
    assert abs(fh.e0 - ev[0]) < 1e-12
    assert np.max(np.abs(fh.e - ev)) < 1e-10
    # G_ii sum rule: total weight of the diagonal = 1
    for spin in ("up", "down"):
        for i in range(6):
            assert abs(np.sum(fh.gf[spin][i, i].pole["weights"]) - 1.0) < 1e-6
    # lesser weights reproduce the 1-RDM (static_rq vs spgf consistency)
    A = fh.one_rdm_from_gf["up"]
    assert np.max(np.abs(A - fh.one_rdm["up"])) < 1e-10
    # poles vs oracle: lowest N+1 (up) eigenvalue
    mt, _ = m.excite(V[:, 0], 1, 0, 1)
    e_np1 = np.linalg.eigvalsh(mt.dense())[0]
    assert abs(fh._gfdata["positions_greater_up"][0] - e_np1) < 1e-10
    assert abs(fh.ae["up"] - (fh.e0 - e_np1)) < 1e-10


def test_lanczos_path_chains_sum_rule(ctx):
    from pyhub_amd.solver.fermi_hubbard import FermiHubbard
    T = O.ring_T(8)
    fh = FermiHubbard(8, 4, 4, T, U=4.0)
    fh.kernel(max_lcz=600, acc_lcz=1e-12)
    fh.spgf(nb_sites_comp=2, method="chains")
    for spin in ("up", "down"):
        for i in range(2):
            w = fh.gf[spin][i, i].pole["weights"]
            assert abs(np.sum(w) - 1.0) < 1e-6
    # device continued fraction == partial fractions of the same chains
    omega = np.linspace(-8, 8, 64)
    G = fh.gf_diag(omega, eta=0.2)
    e0 = fh.e[0]
    thg, qg = fh._gfdata["chains"]["greater_up"][0][:2]
    thl, ql = fh._gfdata["chains"]["lesser_up"][0][:2]
    z = omega + 0.2j
    Gp = (np.sum(qg ** 2 / (z[:, None] - (thg - e0)), axis=1)
          + np.sum(ql ** 2 / (z[:, None] - (e0 - thl)), axis=1))
    assert np.max(np.abs(G[0, 0] - Gp)) < 1e-9


def test_lanczos_path_band_green_function_vs_dense_path(ctx):
    """Lanczos path (max_lcz below the dimension) with the band Lanczos G_ij against
    the dense path (SPGF_ED) of the same model.  The band recurrence stops when its
    lowest pole has converged (as the reference's), so the lowest poles (ae) and the
    zeroth moment (one_rdm_from_gf) are exact and G_ij(omega) is close (the
    unconverged upper poles move G by a few 1e-3 of its peak at eta = 0.3)."""
    from pyhub_amd.solver.fermi_hubbard import FermiHubbard
    T = O.ring_T(6)
    T[0, 3] = T[3, 0] = -0.3
    T[1, 1] = 0.2
    fl = FermiHubbard(6, 3, 2, T, U=3.0)
    fl.kernel(max_lcz=250, acc_lcz=1e-13, compute_spgf=True)   # dim 300 > 250: Lanczos
    assert fl.is_lanczos
    fd = FermiHubbard(6, 3, 2, T, U=3.0)
    fd.kernel(max_lcz=600, compute_spgf=True)
    omega = np.linspace(-10, 10, 101)
    for spin in ("up", "down"):
        for i in range(6):
            for j in range(6):
                gl = fl.gf[spin][i, j].retarded(omega, 0.3)
                gd = fd.gf[spin][i, j].retarded(omega, 0.3)
                assert np.max(np.abs(gl - gd)) < 1e-2 * np.max(np.abs(gd)) + 1e-6, (spin, i, j)
    assert abs(fl.ae["up"] - fd.ae["up"]) < 1e-9
    assert abs(fl.ip["down"] - fd.ip["down"]) < 1e-9
    assert np.max(np.abs(fl.one_rdm_from_gf["down"] - fd.one_rdm["down"])) < 1e-8


def test_lanczos_path_temperature_and_excited_states(ctx):
    """T > 0 ensembles and excited states on the Lanczos path against the dense
    path (EXACT_DIAG) of the same model: BOLTZMANN over the Ritz values
    (boltzmann.f90:1-29) weights NW = 7 states at T = 0.1 here, each rebuilt from
    the Lanczos vectors; the weighted 1-RDM, spin correlations and the ensemble
    energy match, and set_state(b) selects state b on both paths."""
    from pyhub_amd.solver.fermi_hubbard import FermiHubbard
    T = O.disordered_ring_T(6)
    fl = FermiHubbard(6, 3, 2, T, U=2.5, T=0.1, lanczos_mode="full64")
    fl.kernel(max_lcz=250, acc_lcz=1e-12, nb_comp_states=3)    # dim 300 > 250
    assert fl.is_lanczos
    fd = FermiHubbard(6, 3, 2, T, U=2.5, T=0.1)
    fd.kernel(max_lcz=600, nb_comp_states=3)
    assert not fd.is_lanczos
    nw = len(fd.boltzmann)
    assert nw == 7 and len(fl.boltzmann) == nw
    assert np.max(np.abs(fl.e[:nw] - fd.e[:nw])) < 1e-9
    assert np.max(np.abs(fl.boltzmann - fd.boltzmann) / fd.boltzmann) < 1e-7
    for spin in ("up", "down"):
        assert np.max(np.abs(fl.one_rdm[spin] - fd.one_rdm[spin])) < 1e-7
    assert np.max(np.abs(fl.spin_cor_matrix - fd.spin_cor_matrix)) < 1e-7
    assert abs(fl.e_one_body - fd.e_one_body) < 1e-7
    for b in range(nw):
        assert abs(abs(fl.psi(b) @ fd.psi(b)) - 1.0) < 1e-8, b
    # a single excited state (excited_state in the reference's input)
    fl.set_state(2)
    fd.set_state(2)
    fl.reduced_quantities()
    fd.reduced_quantities()
    assert np.max(np.abs(fl.one_rdm["up"] - fd.one_rdm["up"])) < 1e-7
    # T = 0 with three computed states on the plain (reference) recurrence
    fp = FermiHubbard(6, 3, 2, T, U=2.5)
    fp.kernel(max_lcz=250, acc_lcz=1e-12, nb_comp_states=3)
    assert fp.is_lanczos and len(fp.boltzmann) == 3
    fp.set_state(1)
    fp.reduced_quantities()
    fd.set_state(1)
    fd.reduced_quantities()
    assert abs(fp.e[1] - fd.e[1]) < 1e-9
    assert np.max(np.abs(fp.one_rdm["up"] - fd.one_rdm["up"])) < 1e-7
    # Green's functions of a T > 0 ensemble need the dense path
    with pytest.raises(NotImplementedError):
        fl.set_state()
        fl.spgf()


@pytest.mark.parametrize("temp,path", [(0.1, "lanczos"), (0.1, "dense"), (0.6, "dense")])
def test_thermal_ensemble_vs_canonical_average(ctx, temp, path):
    """BOLTZMANN (boltzmann.f90:1-29) weights against the canonical ensemble of the
    oracle's full spectrum: Z, the ensemble energy e0 = sum_k w_k E_k / Z and the
    ensemble 1-RDM, densities and spin correlations equal the exact thermal averages
    sum_k exp(-E_k/T) <k|A|k> / sum_k exp(-E_k/T) over all states (dense ED of the
    pinned HPROD, expectations from the pinned ONE_RDM / NI / J_SPIN restatements),
    to the 1e-8 weight cut of BOLTZMANN.  A wrong sign in the exponent, a missing
    1/Z or weights paired with the wrong states fails these comparisons."""
    from pyhub_amd.solver.fermi_hubbard import FermiHubbard
    T = O.disordered_ring_T(6)
    U = 2.5
    m = O.HubbardModel(6, 3, 2, T, U)                      # dim 300, N_up != N_dn
    ev, V = np.linalg.eigh(m.dense())
    p = np.exp((ev[0] - ev) / temp)
    Zx = p.sum()
    p /= Zx
    E_x = p @ ev
    rho = sum(pk * m.one_rdm(V[:, k]) for k, pk in enumerate(p) if pk > 1e-14)
    ni = sum(pk * m.ni(V[:, k]) for k, pk in enumerate(p) if pk > 1e-14)
    ss = sum(pk * m.sisj(V[:, k]) for k, pk in enumerate(p) if pk > 1e-14)
    if path == "lanczos":
        f = FermiHubbard(6, 3, 2, T, U=U, T=temp, lanczos_mode="full64")
        f.kernel(max_lcz=250, acc_lcz=1e-12, nb_comp_states=3)
        assert f.is_lanczos
    else:
        f = FermiHubbard(6, 3, 2, T, U=U, T=temp)
        f.kernel(max_lcz=600, nb_comp_states=3)
        assert not f.is_lanczos
    nw = len(f.boltzmann)
    # NW: the states whose weight exceeds 1e-8 (at least nb_comp_states)
    assert nw == max(3, int(np.sum(np.exp((ev[0] - ev) / temp) > 1e-8)))
    cut = np.exp((ev[0] - ev[nw:]) / temp).sum() / Zx      # weight BOLTZMANN leaves out
    print(f"T={temp} {path}: NW={nw}, weight left out {cut:.1e}", flush=True)
    tol = 1e-8 + 4 * cut
    assert abs(f.Z - Zx) <= tol * Zx
    assert abs(f.e0 - E_x) <= tol * max(1.0, abs(E_x))
    assert np.max(np.abs(f.one_rdm["up"] - rho[0])) <= tol
    assert np.max(np.abs(f.one_rdm["down"] - rho[1])) <= tol
    assert np.max(np.abs(f.ni - ni)) <= tol
    assert np.max(np.abs(f.spin_cor_matrix - ss)) <= tol


def test_self_energy_dimer_closed_form(ctx):
    """self_energy / qpw (reduced_quantities.py:134-169) on the half-filled Hubbard
    dimer, whose self-energy is known in closed form: in the bonding (e = -t) and
    antibonding (e = +t) orbitals Sigma(z) = U/2 + U^2 / (4 (z - U/2 -+ 3t)), so the
    quasiparticle weight at mu = U/2 is 1 / (1 + U^2 / (36 t^2)); U = 0 gives 0."""
    from pyhub_amd.solver.fermi_hubbard import FermiHubbard
    t, U = 1.0, 3.0
    T = np.array([[0.0, -t], [-t, 0.0]])
    fh = FermiHubbard(2, 1, 1, T, U=U)
    fh.kernel(compute_spgf=True)
    assert not fh.is_lanczos
    ek, Vk = np.linalg.eigh(T)
    for w, eta in ((0.37, 0.2), (-2.5, 0.05), (4.0, 0.5)):
        Sk = Vk.T @ fh.self_energy(w, eta=eta)["up"] @ Vk
        z = w + 1j * eta
        exact = U / 2 + U * U / (4 * (z - U / 2 + np.array([-3 * t, 3 * t])))
        assert np.max(np.abs(np.diag(Sk) - exact)) < 1e-10
        assert abs(Sk[0, 1]) < 1e-10
    ws = np.array([0.1, 0.7, 1.3])
    S = fh.self_energy(ws, eta=0.1)
    assert S["up"].shape == (2, 2, 3)
    for k, w in enumerate(ws):
        assert np.max(np.abs(S["down"][:, :, k].T - fh.self_energy(w, eta=0.1)["down"])) < 1e-12
    assert abs(fh.mu["up"] - U / 2) < 1e-10
    assert np.max(np.abs(fh.qpw - np.identity(2) / (1 + U * U / (36 * t * t)))) < 1e-5
    assert fh.galitskii_migdal is None
    f0 = FermiHubbard(2, 1, 1, T, U=0.0)
    f0.kernel(compute_spgf=True)
    assert np.max(np.abs(f0.self_energy(0.3, eta=0.1)["up"])) < 1e-10


def test_empty_species_and_green_functions(ctx):
    """No up electron (N_up = 0): the lesser up Green's function has no pole (an
    empty N-1 sector), the greater one carries the whole sum rule; E0 and the
    1-RDM match dense ED of the oracle's matrix."""
    from pyhub_amd.solver.fermi_hubbard import FermiHubbard
    T = O.ring_T(5)
    fh = FermiHubbard(5, 0, 2, T, U=3.0)
    fh.kernel(compute_spgf=True)
    m = O.HubbardModel(5, 0, 2, T, 3.0)
    ev, V = np.linalg.eigh(m.dense())
    assert abs(fh.e0 - ev[0]) < 1e-12
    assert np.max(np.abs(fh.one_rdm["up"])) < 1e-14
    assert abs(np.trace(fh.one_rdm["down"]) - 2.0) < 1e-12
    assert fh.nb_poles[0, 0] == 0
    for i in range(5):
        w = fh.gf["up"][i, i].pole["weights"]
        assert abs(np.sum(w) - 1.0) < 1e-10


def test_fermi_hubbard_4index_u(ctx):
    """FermiHubbard with a 4-index U (is_Uloc = 3): energies against dense ED of the
    oracle operator, two_rdm against the oracle's DBL_OCC of the same psi, ni equal to
    the diagonal of two_rdm, and e_two_body as the reference evaluates it
    (reduced_quantities.py:67-71: is_Uloc = 3 is truthy, so the sum(ni * U) branch)."""
    from pyhub_amd.solver.fermi_hubbard import FermiHubbard
    n = 6
    rng = np.random.default_rng(11)
    U4 = rng.normal(size=(n, n, n, n))
    U4[rng.random((n, n, n, n)) > 0.1] = 0.0
    U4 = 0.5 * (U4 + np.einsum("ijkl->jilk", U4))
    for J in range(n):
        U4[J, J, J, J] = 3.0
    T = O.ring_T(n)
    fh = FermiHubbard(n, 3, 2, T, U=U4)
    fh.kernel(max_lcz=600, compute_rq=True)
    m = O.HubbardModel(n, 3, 2, T, U4)
    ev, V = np.linalg.eigh(m.dense())
    assert abs(fh.e0 - ev[0]) < 1e-10
    psi = fh.psi0.cpu().numpy() if hasattr(fh.psi0, "cpu") else np.asarray(fh.psi0)
    D = fh.two_rdm
    assert np.max(np.abs(D - m.two_rdm(psi))) < 1e-10
    assert np.max(np.abs(np.array([D[j, j, j, j] for j in range(n)]) - fh.ni)) < 1e-10
    # the reference's expression: the per-site sum broadcast against the 4-index U
    assert abs(fh.e_two_body - (np.sum(fh.ni * U4)
                                + np.einsum("ij,ij", fh.spin_cor_matrix, fh.J_matrix))) < 1e-9


def test_heisenberg_ring(ctx):
    from pyhub_amd.solver.heisenberg import Heisenberg
    n = 8
    J = np.zeros((n, n))
    for i in range(n):
        J[i, (i + 1) % n] = J[(i + 1) % n, i] = 1.0
    h = Heisenberg(n, J)
    h.kernel(max_lcz=60, acc_lcz=1e-12)              # dim 70 >= 60 -> Lanczos path
    m = O.HeisenbergModel(n, J)
    ev, V = np.linalg.eigh(m.dense())
    assert abs(h.e0 - ev[0]) < 1e-9
    assert np.max(np.abs(h.spin_cor_matrix - m.sisj(V[:, 0]))) < 1e-6
    # reference convention on a bipartite ring: 2x the standard energy
    es = np.linalg.eigvalsh(O.heisenberg_standard_dense(n, np.triu(J, 1)))[0]
    assert abs(h.e0 - 2 * es) < 1e-9


def test_norm_lost_is_reported_not_raised(ctx):
    """Plain FP64 Lanczos run far past convergence (60 fixed steps on the 70-state
    ring) loses orthogonality; as in the reference (lanczos.f90:74-76, a printed
    NORM LOST) the call succeeds, psi comes back normalised and is still the ground
    state."""
    from pyhub_amd import _lib as L
    from pyhub_amd.engine import HeisenbergOp
    n = 8
    J = np.zeros((n, n))
    for i in range(n):
        J[i, (i + 1) % n] = J[(i + 1) % n, i] = 1.0
    op = HeisenbergOp(ctx, n, J)
    m = O.HeisenbergModel(n, J)
    ev, V = np.linalg.eigh(m.dense())
    v0 = torch.as_tensor(np.random.default_rng(3).uniform(-0.5, 0.5, m.dim))
    out = op.lanczos(mode=L.LCZ_PLAIN64, fixed_steps=60, v0=v0, want_psi=True)
    psi = out.psi.cpu().numpy()
    assert abs(np.linalg.norm(psi) - 1.0) < 1e-12
    assert abs(out.e0 - ev[0]) < 1e-9
    assert abs(abs(psi @ V[:, 0]) - 1.0) < 1e-8
    assert isinstance(out.norm_lost, bool)


def test_dimer_closed_form(ctx):
    from pyhub_amd.solver.fermi_hubbard import FermiHubbard
    for U in (0.0, 2.0, 8.0):
        fh = FermiHubbard(2, 1, 1, np.array([[0.0, -1.0], [-1.0, 0.0]]), U=U)
        fh.kernel()
        assert abs(fh.e0 - O.dimer_e0(1.0, U)) < 1e-12


def test_errors(ctx):
    from pyhub_amd.solver.fermi_hubbard import FermiHubbard
    with pytest.raises(ValueError):
        FermiHubbard(3, 1, 1, np.array([[0, 1, 0], [0, 0, 1], [0, 0, 0.0]]), U=1.0)
    fh = FermiHubbard(8, 4, 4, O.ring_T(8), U=4.0)
    fh.kernel(max_lcz=12, acc_lcz=1e-14, compute_rq=False)   # step cap reached
    with pytest.raises(StopIteration):
        fh.lcz_results
    T = O.ring_T(4)
    T[0, 0] = np.nan
    with pytest.raises(ValueError):
        FermiHubbard(4, 2, 2, T, U=1.0).kernel()
