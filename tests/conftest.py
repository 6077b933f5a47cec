import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")
FIXTURES = ("seeded_11", "seeded_12", "seeded_13_cic", "edges", "seeded_14_fdm", "seeded_15_128")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


def load_fixture(name):
    """(Snapshot, dict of the reference's outputs) for one golden fixture."""
    from asohf_amd.geometry import Geometry
    from asohf_amd.synth import Snapshot
    d = np.load(os.path.join(GOLDEN, name + ".npz"))
    geom = Geometry(int(d["nx"]), int(d["ny"]), int(d["nz"]), d["lado0"], d["rote"], d["rete"],
                    d["npatch"], d["pnx"], d["pny"], d["pnz"], d["prx"], d["pry"], d["prz"],
                    fdm=d["fdm"], interp_degree=int(d["interp_degree"]))
    snap = Snapshot(d["x"], d["y"], d["z"], d["m"], d["npart_esp"], geom)
    return snap, {k: d[k] for k in d.files}


@pytest.fixture(params=FIXTURES)
def golden(request):
    return load_fixture(request.param)
