"""MI355X-native exact-diagonalization engine with PyHub's solver API.

The Python classes mirror pyhub's (Basis, CoreSolver, FermiHubbard, Heisenberg,
Pole); the work runs in libphub.so (hand-written HIP for gfx950) through the C
ABI of include/phub.h.
"""
__version__ = "0.1.0"


def __getattr__(name):
    # lazy: importing the package must not require a GPU (build / CPU tests)
    if name == "FermiHubbard":
        from pyhub_amd.solver.fermi_hubbard import FermiHubbard
        return FermiHubbard
    if name == "Heisenberg":
        from pyhub_amd.solver.heisenberg import Heisenberg
        return Heisenberg
    if name == "Basis":
        from pyhub_amd.core.basis import Basis
        return Basis
    if name == "Pole":
        from pyhub_amd.tools.pole import Pole
        return Pole
    raise AttributeError(name)
