"""Shared fixtures: the libpf binding, the oracle, and generated cases."""
from __future__ import annotations

import importlib.util
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG_DIR = os.path.join(ROOT, "experience-of-using-gpu-in-power-flow-computation_amd")
GOLDEN = os.path.join(ROOT, "tests", "golden")
IEEE = os.path.join(GOLDEN, "ieee")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built libpf.so")


def _load_pkg():
    if "pfamd" in sys.modules:
        return sys.modules["pfamd"]
    spec = importlib.util.spec_from_file_location("pfamd", os.path.join(PKG_DIR, "__init__.py"),
                                                  submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["pfamd"] = mod
    spec.loader.exec_module(mod)
    return mod


def _load_synth():
    _load_pkg()
    import importlib
    return importlib.import_module("pfamd.synth")


@pytest.fixture(scope="session")
def pf():
    lib = os.path.join(PKG_DIR, "libpf.so")
    if not os.path.exists(lib):
        subprocess.run(["make", "-s", "-C", PKG_DIR, "-j8", "libpf.so"], check=True)
    m = _load_pkg()
    m.lib()
    return m


@pytest.fixture(scope="session")
def oracle():
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    O.build()
    return O


@pytest.fixture(scope="session")
def synth():
    return _load_synth()


class Cases:
    """Synthetic cases written once per session (the CDF file is canonical)."""

    def __init__(self, root, synth):
        self.root = root
        self.synth = synth
        self._paths = {}

    def path(self, name: str) -> str:
        if name in self._paths:
            return self._paths[name]
        s = self.synth
        if name.startswith("ieee"):
            p = os.path.join(IEEE, name + ".cdf")
        else:
            makers = {"c14": (s.case14, "cdf"), "c118": (s.case118, "cdf"), "c944": (s.case944, "cdf"),
                      "c1888": (s.case1888, "cdf"), "c15104": (s.case15104, "cdfx")}
            fn, ext = makers[name]
            p = os.path.join(self.root, f"{name}.{ext}")
            s.write_cdf(fn(), p)
        self._paths[name] = p
        return p


@pytest.fixture(scope="session")
def cases(tmp_path_factory, synth):
    return Cases(str(tmp_path_factory.mktemp("cases")), synth)


def ybus_dense(net) -> np.ndarray:
    """Independent dense Ybus from the branch list (pi model, MATPOWER form)."""
    n = net.nB
    Y = np.zeros((n, n), complex)
    for l in range(net.nL):
        f, t = net.frm[l], net.to[l]
        y = 1.0 / complex(net.r[l], net.x[l])
        tap = complex(net.tap_re[l], net.tap_im[l])
        b = net.bch[l]
        Y[f, f] += (y + 0.5j * b) / abs(tap) ** 2
        Y[f, t] += -y / tap.conjugate()
        Y[t, f] += -y / tap
        Y[t, t] += y + 0.5j * b
    Y[np.arange(n), np.arange(n)] += net.gsh + 1j * net.bsh
    return Y


def power(Y, V, theta):
    E = V * np.exp(1j * theta)
    S = E * np.conj(Y @ E)
    return S.real, S.imag


def csr_to_dense(rp, ci, v, n):
    A = np.zeros((n, n), dtype=v.dtype)
    for i in range(n):
        A[i, ci[rp[i]:rp[i + 1]]] = v[rp[i]:rp[i + 1]]
    return A
