"""CPU tests of host-side logic of the package (no GPU calls)."""
from pyhub_amd.grad import HeisenbergSpec, lattice_bonds


def test_lattice_bonds_small_periodic_sides():
    """Periodic sides of length 2 and 1: every unordered bond once, no self bond
    (each listed bond owns its matrix entries, so its gradient is the whole
    derivative of those entries)."""
    b = lattice_bonds(2, 3, periodic=True)
    keys = [tuple(sorted(x)) for x in b]
    assert len(keys) == len(set(keys)) and all(i != j for i, j in b)
    assert len(b) == 3 + 6          # one horizontal bond per row, a 3-ring per column
    b1 = lattice_bonds(1, 4, periodic=True)
    assert len(b1) == 4 and all(i != j for i, j in b1)
    assert len(lattice_bonds(4, 3)) == 17     # the 3x4 open cluster of the gradient workload


def test_heisenberg_spec_coupling_matrix():
    """Standard convention: J(i,j) = J(j,i) = J_b / 2 (the ordered-pair sum then
    gives J_b S_i.S_j); reference convention: J_b on both ordered pairs."""
    s = HeisenbergSpec(n=4, bonds=[(0, 1), (1, 2)], convention="standard")
    J = s.Jm([2.0, 4.0])
    assert J[0, 1] == J[1, 0] == 1.0 and J[1, 2] == J[2, 1] == 2.0 and s.sign_convention == 1
    r = HeisenbergSpec(n=4, bonds=[(0, 1)], convention="reference")
    assert r.Jm([3.0])[0, 1] == 3.0 and r.sign_convention == 0


def test_pyhub_import_paths_alias_the_engine():
    """The reference's module paths resolve to this engine's modules (same objects)."""
    import importlib
    import pyhub_amd.solver.fermi_hubbard as ours
    ref = importlib.import_module("pyhub.solver.fermi_hubbard")
    assert ref is ours and ref.FermiHubbard is ours.FermiHubbard
    from pyhub.core.basis import Basis
    from pyhub_amd.core.basis import Basis as B2
    assert Basis is B2
    from pyhub.tools.operators import c_dagger_c, opeexp, sisj  # noqa: F401
    from pyhub.tools.pole import Pole  # noqa: F401
    from pyhub.solver.heisenberg import Heisenberg  # noqa: F401


def test_genuine_ritz_skips_ghosts_and_spurious_values():
    """Ritz values of the three-term recurrence with FP32 vectors (from a 200-step run
    on 300 states, below a ground level added): the copies of a converged level (1e-8
    apart) count once, through the member of largest first component; a value whose
    eigenvector has no weight on the start vector (a copy still forming) is dropped;
    the lowest value is always the ground state; the FP64 tolerance keeps levels 1e-6
    apart distinct."""
    import numpy as np
    from pyhub_amd.core.core_solver import _GHOST_TOL, _genuine_ritz
    val = np.array([-7.5, -7.2696915827, -7.2696915683, -7.2696915654, -6.5907965758,
                    -6.5906695780, -6.5906695309, -6.3282888371, -6.3269646192])
    u1 = np.array([0.3, 3.5e-5, 5.7e-3, 5.8e-3, 8.6e-19, 2.9e-4, 4.9e-2, 1.1e-19, 1.0e-3])
    assert list(_genuine_ritz(val, u1, _GHOST_TOL["plain32"])) == [0, 3, 6, 8]
    # a ghost cluster at the bottom stays the ground state through its lowest member
    assert list(_genuine_ritz(val[1:], u1[1:], _GHOST_TOL["plain32"])) == [0, 5, 7]
    v64 = np.array([-7.2696915682, -7.2696905682, -6.5906695363])
    assert list(_genuine_ritz(v64, np.array([1e-2, 1e-3, 1e-12]), _GHOST_TOL["plain64"])) == [0, 1]
    assert list(_genuine_ritz(np.array([1.0]), np.array([1.0]), 1e-8)) == [0]
    # the lowest value stays the ground state, even alone with no start-vector weight
    assert list(_genuine_ritz(np.array([-2.0, -1.0]), np.array([1e-16, 0.5]), 1e-8)) == [0, 1]
