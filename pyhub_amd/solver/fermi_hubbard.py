"""FermiHubbard with PyHub's API (mirror of pyhub/solver/fermi_hubbard.py)."""
from __future__ import annotations

import numpy as np

from pyhub_amd.core.core_solver import CoreSolver


class FermiHubbard(CoreSolver):
    r"""Solution of a general Hubbard cluster by exact diagonalisation or Lanczos
    on the GPU.  H = \sum_{ij\sigma} t_{ij} c^\dagger_{j\sigma} c_{i\sigma}
    + \sum_i U_i n_{i\uparrow} n_{i\downarrow} (hprod_hubbard.f90:17-118).

    Extra keyword (not in the reference): lanczos_mode in {'plain64' (the
    reference's recurrence, default), 'full64', 'mixed'}."""

    def __init__(self, nb_sites, n_up: int, n_down: int, t_matrix: np.ndarray, U: float = 0,
                 T=0.0, order="spin", lanczos_mode="plain64"):
        CoreSolver.__init__(self, nb_sites, n_up, n_down, "hubbard.x", T, order=order)
        self.t_matrix = np.asarray(t_matrix, dtype=float)
        if not np.isclose(self.t_matrix, self.t_matrix.T).all():
            raise ValueError("t_matrix must be symmetric")
        self.ek, self.Vk = np.linalg.eigh(self.t_matrix)
        self.U = U
        self.lanczos_mode = lanczos_mode

    def __fileremove__(self):
        pass
