import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU; runs the HIP engine")


@pytest.fixture(scope="session")
def pkg():
    import __graft_entry__ as g
    return g.load_package()


@pytest.fixture(scope="session")
def orc():
    import oracle
    oracle.lib()
    return oracle


def new_env(pkg, preset=0, plasticities=0, **overrides):
    """An environment holding the synthetic scaffold `preset` with config overrides
    (room for `plasticities` plasticity rules)."""
    env = pkg.NetworkEnvironment(1, 7, 16, 2, plasticities)
    cfg = pkg.synth_preset(preset)
    for k, v in overrides.items():
        setattr(cfg, k, v)
    env.describe_synthetic(cfg)
    return env
