"""PyHub's own import paths over pyhub_amd, so that code written against the
reference package (pyhub/solver/fermi_hubbard.py, pyhub/core/basis.py, ...) runs
unchanged: `from pyhub.solver.fermi_hubbard import FermiHubbard` resolves to
pyhub_amd.solver.fermi_hubbard.  The submodules load on first import (importing
pyhub itself touches neither torch nor the GPU).  The reference's analytic
solvers (bethe_anzatz, gutzwiller) and tools.dos are not part of this engine.
"""
import importlib
import importlib.abc
import importlib.util
import sys

__version__ = "1.0.0"

_SUBMODULES = {
    "core", "core.basis", "core.core_solver", "core.reduced_quantities", "core._operators",
    "solver", "solver.fermi_hubbard", "solver.heisenberg",
    "tools", "tools.models", "tools.operators", "tools.pole", "tools.tools",
}


class _AliasLoader(importlib.abc.Loader):
    def __init__(self, target):
        self.target = target

    def create_module(self, spec):
        # the pyhub_amd module itself becomes pyhub.<sub> (one module object, so
        # classes compare equal whichever path imported them)
        return importlib.import_module(self.target)

    def exec_module(self, module):
        pass


class _AliasFinder(importlib.abc.MetaPathFinder):
    def find_spec(self, fullname, path, target=None):
        if fullname.startswith("pyhub.") and fullname[len("pyhub."):] in _SUBMODULES:
            sub = fullname[len("pyhub."):]
            return importlib.util.spec_from_loader(fullname, _AliasLoader("pyhub_amd." + sub))
        return None


if not any(isinstance(f, _AliasFinder) for f in sys.meta_path):
    sys.meta_path.insert(0, _AliasFinder())
