import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built libkf.so")


@pytest.fixture(scope="session")
def ctx():
    """One HIP context (device 0) shared by the GPU tests of the session."""
    from killingfusion_amd import Context
    c = Context(0)
    yield c
    c.close()


@pytest.fixture
def loop_capture(ctx):
    """The shared context with its active-set iteration loops captured as hipGraphs and
    replayed (kf_context_set_loop_capture) for one test; direct launches again afterwards."""
    ctx.set_loop_capture(True)
    yield ctx
    ctx.set_loop_capture(False)
