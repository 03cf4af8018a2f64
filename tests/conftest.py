import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built libsemigpu.so")


def golden_names(full_only=False):
    out = []
    for f in sorted(os.listdir(GOLDEN)):
        if f.endswith(".npz"):
            if full_only:
                with np.load(os.path.join(GOLDEN, f)) as d:
                    if not int(d["full"]):
                        continue
            out.append(f[:-4])
    return out


def load_golden(name):
    with np.load(os.path.join(GOLDEN, name + ".npz")) as d:
        return {k: d[k] for k in d.files}


def atoms_of(g):
    return [(int(z), float(x[0]), float(x[1]), float(x[2])) for z, x in zip(g["Z"], g["xyz_ang"])]


@pytest.fixture(scope="session")
def engine():
    from semigpu_amd import api
    eng = api.Engine()
    yield eng
    eng.close()
