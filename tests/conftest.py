"""Shared test setup: the `gpu` marker, import paths, the C oracle build and
golden-fixture loading."""
from __future__ import annotations

import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
GOLDEN = os.path.join(ROOT, "tests", "golden")
for p in (ROOT, ORACLE_DIR):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def _ensure_oracle():
    so = os.path.join(ORACLE_DIR, "liboracle.so")
    src = os.path.join(ORACLE_DIR, "v2d_oracle.c")
    if not os.path.exists(so) or os.path.getmtime(so) < os.path.getmtime(src):
        subprocess.run(["make", "-s", "-C", ORACLE_DIR, "all"], check=True)


_ensure_oracle()


def golden(name: str) -> dict:
    """Load a fixture of tests/golden (plain arrays, no pickle)."""
    with np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False) as d:
        return {k: d[k] for k in d.files}


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle
    return oracle


@pytest.fixture(scope="session")
def C(oracle_mod):
    """The C restatement of the reference core."""
    return oracle_mod.load_c()


@pytest.fixture(scope="session")
def REF(oracle_mod):
    """The reference Fortran built from its sources (skips where not built)."""
    r = oracle_mod.load_ref()
    if r is None:
        pytest.skip("oracle/_ref not built (make -C oracle ref needs /root/reference)")
    return r
