"""Shared fixtures.  ``gpu`` marks tests that need a HIP device; everything
else runs on the CPU.  The oracle (oracle/) is the checker only."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)

from g3lhalo_amd import tables as T  # noqa: E402

ARCMIN = T.ARCMIN_IN_RAD


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")


@pytest.fixture(scope="session")
def tables():
    return T.make_tables()


@pytest.fixture(scope="session")
def oracle_model(tables):
    import oracle as O
    O.build()
    m = O.Model(tables)
    m.set_hods(T.FIDUCIAL_HOD, T.FIDUCIAL_HOD)
    return m


def say(msg):
    """Names the case before a GPU step (flushed, so it survives a crash)."""
    print(msg, flush=True)
    sys.stdout.flush()
