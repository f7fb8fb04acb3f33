"""Pole representation of a Green's function (mirror of pyhub/tools/pole.py).

G(w) = sum_i weights_i / (w - positions_i + i*eta).  Host-side numpy: this is
bookkeeping on the poles the engine produces, not part of the device path.
Differences from the reference: vectorised evaluation; greater/lesser/causal on
arrays use each grid point (pole.py:105,114 evaluate at the whole array `w`);
pole_research and raise_deg are not provided (fitting / plotting helpers).
"""
from __future__ import annotations

import numpy as np


class Pole:
    def __init__(self, pole: dict, atol: float = 1.0e-10):
        if set(pole.keys()) != {"positions", "weights"}:
            raise ValueError(
                "Pole asked is not correct, use correct notation -> "
                f"{ {'positions': [], 'weights': []} } not {pole.keys()}")
        self.pole = {"positions": np.array(pole["positions"], copy=True, dtype=float).reshape(-1),
                     "weights": np.array(pole["weights"], copy=True, dtype=float).reshape(-1)}
        self.atol = atol
        if len(self.pole["weights"]) != len(self.pole["positions"]):
            raise ValueError(f"number of poles {len(self.pole['positions'])} should be equal as "
                             f"len of weights {len(self.pole['weights'])}")

    def __call__(self, w, eta: float = 0.05):
        return self.retarded(w, eta=eta)

    def __str__(self):
        return f"{self.pole}"

    def __add__(self, other):
        return Pole({"positions": np.concatenate((self.pole["positions"], other.pole["positions"])),
                     "weights": np.concatenate((self.pole["weights"], other.pole["weights"]))})

    def __sub__(self, other):
        return Pole({"positions": np.concatenate((self.pole["positions"], other.pole["positions"])),
                     "weights": np.concatenate((self.pole["weights"], -other.pole["weights"]))})

    def __mul__(self, other):
        if isinstance(other, (float, int, complex, np.complex128, np.floating)):
            return Pole({"positions": self.pole["positions"], "weights": self.pole["weights"] * other})
        raise TypeError(f"Other must be float complex or int, not {type(other)}")

    def __truediv__(self, other):
        if isinstance(other, (float, int, np.floating)):
            return Pole({"positions": self.pole["positions"], "weights": self.pole["weights"] / other})
        raise TypeError(f"Other must be float or int, not {type(other)}")

    def shift(self, s: float):
        return Pole({"positions": s + self.pole["positions"], "weights": self.pole["weights"]})

    def clean(self, ptol: float = 1.0e-5, wtol: float = 1.0e-5):
        """Drop poles with |weight| < wtol, merge positions equal after rounding to
        -log10(ptol) decimals (pole.py:69-78)."""
        keep = np.abs(self.pole["weights"]) >= wtol
        w = self.pole["weights"][keep]
        p = self.pole["positions"][keep]
        pos, inv = np.unique(np.around(p, int(-np.log10(ptol))), return_inverse=True)
        nw = np.zeros(len(pos))
        np.add.at(nw, inv, w)
        self.pole["positions"], self.pole["weights"] = pos, nw
        return self

    def _eval(self, w, shift_eta):
        w = np.asarray(w)
        z = w[..., None] + shift_eta
        return np.sum(self.pole["weights"] / (z - self.pole["positions"]), axis=-1)

    def retarded(self, w, eta: float = 0.05):
        return self._eval(w, 1j * eta)

    def advanced(self, w, eta: float = 0.05):
        return self._eval(w, -1j * eta)

    def greater(self, w, ef, eta: float = 0.05):
        sel = self.pole["positions"] > ef
        return Pole({"positions": self.pole["positions"][sel],
                     "weights": self.pole["weights"][sel]})._eval(w, -1j * eta)

    def lesser(self, w, ef, eta: float = 0.05):
        sel = self.pole["positions"] < ef
        return Pole({"positions": self.pole["positions"][sel],
                     "weights": self.pole["weights"][sel]})._eval(w, 1j * eta)

    def causal(self, w, ef, eta: float = 0.05):
        w = np.asarray(w)
        z = w[..., None] + np.sign(self.pole["positions"] - ef) * 1j * eta
        return np.sum(self.pole["weights"] / (z - self.pole["positions"]), axis=-1)

    def spectral_density(self, w, eta: float = 0.05):
        return -1.0 * np.imag(self.retarded(w, eta)) / np.pi

    @classmethod
    def empty(cls):
        return cls({"positions": np.array([]), "weights": np.array([])})

    @classmethod
    def zero(cls):
        return cls({"positions": np.array([0.0]), "weights": np.array([0.0])})

    @staticmethod
    def sum(lst):
        p = Pole.empty()
        for q in lst:
            p = p + q
        return p

    @property
    def nb_poles(self):
        return len(self.pole["positions"])
