"""Host check of the LT row breakpoints (ripples_amd/csrc/lt_scan.h).

lt_scan_check.cpp compiles the header for the host and checks that the
breakpoint search equals plain bisection over the threshold with the float
chain, and that the crossing edge found from the breakpoints equals the
subtraction loop of include/ripples/add_rrrset.h:136-150, on rows built to hit
exact grid values, halfway ties, subnormals, zeros, -0 and wide ranges."""
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
    exe = str(tmp_path_factory.mktemp("lt_scan") / "lt_scan_check")
    subprocess.run([HIPCC, "-x", "hip", "--offload-host-only", "-O2", "-std=c++17",
                    os.path.join(HERE, "lt_scan_check.cpp"), "-o", exe],
                   check=True, capture_output=True, timeout=300)
    return exe


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_breakpoints_match_float_chain(checker, seed):
    r = subprocess.run([checker, str(seed), "3000"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.startswith("ok")
