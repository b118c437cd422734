import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def lib():
    from derek_amd import lib as L
    L.load()
    return L


def oracle_system(spec):
    """oracle.scf.System built from a derek_amd.config.SystemSpec."""
    from oracle import scf
    return scf.System(spec.lengths, spec.ecut, spec.g_fine_scale, spec.mp, spec.nelec, spec.nbands, spec.kT,
                      spec.centres, spec.heights, spec.widths, spec.cos_amp, spec.seed)


def rel_err(a, b):
    import numpy as np
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(1e-300, float(np.max(np.abs(b)))))


def assert_close(got, ref, tol, what=""):
    e = rel_err(got, ref)
    assert e < tol, f"{what} relative error {e:.3e} > {tol:.1e}"
