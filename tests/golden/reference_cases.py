"""Loader of the reference's golden vectors (``reference_residuals.npz`` /
``reference_residuals.json``, written by ``make_reference_residuals.py``
from the reference's own residual and RK steps).  Pure data access: no
reference code is read or run here."""
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
NPZ = os.path.join(HERE, "reference_residuals.npz")
META = os.path.join(HERE, "reference_residuals.json")

_meta = None
_npz = None


def names():
    return sorted(_load_meta())


def _load_meta():
    global _meta
    if _meta is None:
        with open(META) as fh:
            _meta = json.load(fh)["cases"]
    return _meta


def _arrays():
    global _npz
    if _npz is None:
        _npz = np.load(NPZ, allow_pickle=False)
    return _npz


class GoldenCase:
    """One fixture case: mesh, free parameters, scheme and the reference's
    outputs (R, f at U0; UN and per-step norms after nsteps)."""

    def __init__(self, name):
        m = _load_meta()[name]
        z = _arrays()
        self.name = name
        self.p, self.q = m["p"], m["q"]
        self.nx, self.ny, self.nz = m["nx"], m["ny"], m["nz"]
        self.bounds = list(m["bounds"])
        self.params = m["params"]
        self.scheme, self.dt, self.nsteps = m["scheme"], m["dt"], m["nsteps"]
        for k in ("nodes", "U0", "R", "f", "UN", "norms"):
            setattr(self, k, np.array(z[name + "/" + k]))

    @property
    def nelem(self):
        return self.nx * self.ny * self.nz

    def mesh_args(self):
        return (self.p, self.q, self.nx, self.ny, self.nz, self.bounds, self.nodes)

    def param_kwargs(self):
        fp = self.params
        return dict(gamma=fp["gamma"], mu=fp["mu"], c=fp["c"], rho=fp["rho"], u=fp["u"],
                    v=fp["v"], w=fp["w"], eta=fp["eta"], source=fp["source"])


def load(name):
    return GoldenCase(name)
