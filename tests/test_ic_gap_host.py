"""Host check of the IC gap arithmetic the samplers run (ripples_amd/csrc/
ic_gap.h): the reciprocal division and the block skip thresholds against
exact 128-bit division and the plain gap formula (ic_gap_check.cpp)."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc not available")
    exe = str(tmp_path_factory.mktemp("ic_gap") / "ic_gap_check")
    subprocess.run([HIPCC, "-x", "hip", "--offload-host-only", "-O2", "-std=c++17",
                    os.path.join(HERE, "ic_gap_check.cpp"), "-o", exe],
                   check=True, capture_output=True, timeout=300)
    return exe


@pytest.mark.parametrize("seed", [1, 2])
def test_fast_gap_and_skip_thresholds(checker, seed):
    r = subprocess.run([checker, str(seed), "3000"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.startswith("ok")
