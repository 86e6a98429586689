"""The oracle built with AddressSanitizer and UBSan (oracle/Makefile target
`asan`) runs a short case of every path -- RKC, CN, BE and the exact tree
solve for the Purkinje, olive, granule and Golgi models, a Purkinje pair
with its gap junctions, and the fixed in-degree network over the three
synaptic populations -- on the reference's morphology files, in bounds,
with no undefined behaviour and finite voltages."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_oracle_runs_clean_under_asan_and_ubsan(tmp_path, ref_data):
    if not shutil.which("gcc"):
        pytest.skip("no host C compiler")
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "asan", f"ASAN_OUT={tmp_path}"], check=True,
                   capture_output=True, text=True)
    r = subprocess.run([str(tmp_path / "oracle_asan"), ref_data], capture_output=True, text=True,
                       env=dict(os.environ, OMP_NUM_THREADS="1", ASAN_OPTIONS="detect_leaks=1"))
    print(r.stdout)
    assert r.returncode == 0, r.stderr[-4000:]
    assert r.stdout.count(": ok") == 20
