import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C-ABI)")


def load_golden(name, shape=None, dtype=np.float64):
    a = np.fromfile(os.path.join(GOLDEN, name), dtype=dtype)
    if shape is not None:
        a = a.reshape(shape, order="F")
    return a


@pytest.fixture(scope="session")
def h2o():
    """The reference's h2o fixtures (test_occ_ri_k_h2o.sh:2 argv)."""
    return {
        "n_bf": 24, "n_aux": 113, "n_occ": 5, "n_vec2": 322, "bs": 3, "thresh": 1e-20,
        "ints3c": load_golden("3c_ints_h2o.data", (322, 113)),
        "pqm12": load_golden("3c_ints_PQm12_h2o.data", (322, 113)),
        "inv05": load_golden("2c_ints_inv05_h2o.data", (113, 113)),
        "mo": load_golden("occ_mo_h2o.data", (24, 5)),
        "keys": load_golden("v2m_key_h2o.data", None, np.uint64),
    }


def rel_fro(x, ref):
    n = np.linalg.norm(ref)
    return np.linalg.norm(x - ref) / (n if n > 0 else 1.0)
