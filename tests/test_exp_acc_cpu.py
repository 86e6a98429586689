"""exp_acc (csrc/exp_acc.hpp), the exponential the inferior-olive kernels use
on the device, is correctly rounded: built for the host with hipcc, its
results equal exp rounded to nearest from a 60-digit Decimal evaluation on
random and edge arguments (the host C library's exp, by comparison, rounds
some 0.1 % of random arguments the other way)."""
import os
import random
import subprocess
from decimal import Decimal, getcontext

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "cerebellum-neuron-model-simulation-using-rkc_amd", "csrc")
HIPCC = "/opt/rocm/bin/hipcc"

PROGRAM = r"""
#include "exp_acc.hpp"
#include <cstdio>
int main() {
  double x;
  while (std::scanf("%la", &x) == 1) std::printf("%a\n", cbrkc::exp_acc(x));
}
"""


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not available")
def test_exp_acc_is_correctly_rounded(tmp_path):
    src = tmp_path / "t.cpp"
    src.write_text(PROGRAM)
    exe = tmp_path / "t"
    subprocess.run([HIPCC, "-O2", "-ffp-contract=off", "-x", "hip", "--offload-arch=gfx950", "-I", CSRC, str(src),
                    "-o", str(exe)], check=True, capture_output=True)
    rng = random.Random(5)
    xs = [rng.uniform(-2.0, 2.0) for _ in range(3000)] + [rng.uniform(-700.0, 700.0) for _ in range(3000)]
    xs += [0.0, -0.0, 1e-300, -1e-300, 0.5, 1.0, -1.0, 0.6931471805599453, 700.0, -700.0, 708.9, -707.9]
    out = subprocess.run([str(exe)], input="\n".join(x.hex() for x in xs), capture_output=True, text=True,
                         check=True).stdout.split()
    getcontext().prec = 60
    # float(Decimal) rounds the 60-digit value to the nearest double
    bad = [(x, y) for x, y in zip(xs, out) if float.fromhex(y) != float(Decimal(x).exp())]
    assert not bad, bad[:5]
