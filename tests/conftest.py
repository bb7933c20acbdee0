import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")
    config.addinivalue_line("markers", "slow: longer CPU checks")


@pytest.fixture(scope="session")
def plain_case(tmp_path_factory):
    """The plain Pg case (401x61 grid, 20 OBS, 399 shots) written to disk."""
    from fmt2d_amd import synth
    d = str(tmp_path_factory.mktemp("plain"))
    m, g = synth.plain_pg_case()
    synth.write_fmt2d_case(d, m, g)
    return d


def load_case(d):
    from fmt2d_amd import formats
    sm = formats.read_smesh(os.path.join(d, "model.smesh"))
    sx, sz = formats.read_refl(os.path.join(d, "sed.dat"))
    mx, mz = formats.read_refl(os.path.join(d, "moho.dat"))
    return sm, sx, sz, mx, mz
