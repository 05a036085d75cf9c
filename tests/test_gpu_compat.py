"""Runs the C++ driver of the reference-shaped API (overlap_amd/compat/ovl_reference_api.hpp):
WilsonEvoOperator::Compute, OverlapOperator::Compute, Dynamics::Fermions::Update, NCsDot and
ZolotarevCoefficients on reference-shaped objects, each checked against the oracle."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu
BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "compat", "test_compat")


def test_reference_shaped_api(gpu):
    assert os.path.exists(BIN), "tests/compat/test_compat not built (run __graft_entry__.build())"
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0 and "PASS" in r.stdout
