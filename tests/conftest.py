import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from nqpu_amd.engine import Gate  # noqa: E402

GOLDEN = os.path.join(ROOT, "tests", "golden", "golden.npz")
KINDS = ["h", "x", "y", "z", "s", "t", "rx", "ry", "rz", "u", "cnot", "cz", "swap",
         "toffoli", "crx", "cry", "crz", "cr", "sx", "phase", "iswap", "ccz", "rxx",
         "ryy", "rzz", "cswap", "cu", "custom1", "custom2"]


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def decode_gates(d, prefix):
    gates = []
    for i in range(len(d[prefix + "_kind"])):
        k = KINDS[int(d[prefix + "_kind"][i])]
        nt, nc = int(d[prefix + "_nt"][i]), int(d[prefix + "_nc"][i])
        m = None
        if k == "custom1":
            m = d[prefix + "_matrix"][i][:4].reshape(2, 2)
        elif k == "custom2":
            m = d[prefix + "_matrix"][i][:16].reshape(4, 4)
        gates.append(Gate(k, tuple(int(q) for q in d[prefix + "_targets"][i][:nt]),
                          tuple(int(q) for q in d[prefix + "_controls"][i][:nc]),
                          tuple(float(p) for p in d[prefix + "_params"][i]), m))
    return gates


@pytest.fixture(scope="session")
def golden():
    return dict(np.load(GOLDEN, allow_pickle=False))


@pytest.fixture(scope="session")
def oracle():
    from oracle import cpu_ref
    cpu_ref.lib()
    return cpu_ref
