"""Shared fixtures.  GPU tests are marked ``gpu``; CPU tests run with
``-m "not gpu"``.  Nothing here reads /root/reference at run time: the
reference's data files are in tests/golden/."""
from __future__ import annotations

import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and libclf.so")


@pytest.fixture(scope="session")
def mcdata_raw():
    """The reference's 52 mcdata events as stored (hits not x-sorted)."""
    z = np.load(os.path.join(GOLDEN, "mcdata_events.npz"))
    return {k: z[k].tobytes() for k in sorted(z.files, key=int)}


@pytest.fixture(scope="session")
def mcdata_sorted():
    """The same events after the reference's own preorder_by_x."""
    z = np.load(os.path.join(GOLDEN, "mcdata_presort_ref.npz"))
    return {k: z[k].tobytes() for k in sorted(z.files, key=int)}


@pytest.fixture(scope="session")
def oracle_digests():
    with open(os.path.join(GOLDEN, "oracle_digests.json")) as f:
        return json.load(f)


@pytest.fixture(scope="session")
def oracle_mod():
    from oracle import oracle

    oracle.build()
    return oracle


@pytest.fixture(scope="session")
def single_blob():
    from cl_forward_amd import generator

    return generator.single_event().pack()


@pytest.fixture(scope="session")
def plain51_blobs():
    from cl_forward_amd import generator

    return [generator.plain_event(i).pack() for i in range(51)]


@pytest.fixture(scope="session")
def gpu_ctx():
    from cl_forward_amd import search

    ctx = search.Context(0)
    yield ctx
    ctx.close()
