import gzip
import json
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
DATA = os.path.join(GOLDEN, "data")
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def golden(name):
    with gzip.open(os.path.join(GOLDEN, name + ".json.gz"), "rt") as f:
        return json.load(f)


def data_path(name, tmp_path_factory=None):
    """Path of an input in tests/golden/data (gz inputs are unpacked to a temp dir)."""
    p = os.path.join(DATA, name)
    if os.path.exists(p):
        return p
    gz = p + ".gz"
    if os.path.exists(gz):
        import tempfile
        d = tempfile.mkdtemp(prefix="rmx_golden_")
        out = os.path.join(d, name)
        with gzip.open(gz, "rb") as fi, open(out, "wb") as fo:
            fo.write(fi.read())
        return out
    raise FileNotFoundError(name)


INPUTS = {
    "flow_net": "flow.net",
    "flow_txt": "flow.txt",
    "lfr1k_mu03": "directNet_1k_sm_mu_03.txt",
    "lfr5k_mu05": "directNet_5k_sm_mu_05.txt",
    "sbm1k": "sbm_1k_c25_s1.txt",
}
