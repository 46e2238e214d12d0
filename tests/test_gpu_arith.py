"""The step code's device arithmetic (csrc/dev_common.h) over 4M operands: random exponents over the
specified range and operands shaped like the step's (tests/cpp/arith_check.hip, built by
`make -C low-fidelity-dynamic-models_amd adjcheck`).

- qdiv, qsqrt (Newton-Raphson refinement of the hardware reciprocal and reciprocal square root, the
  strict form's quotient and square root) return the same bits as the compiler's IEEE `/` and sqrt;
- the fast form's fdiv, frcp, fsqrt and fsqrt_rcp are within one ulp of the correctly rounded results;
- its sincos_fast (angles up to 1e5 rad) and sin_2pi_unit (turns in [0, 1]) within two ulps of the host's
  long double sine and cosine rounded to double."""
import os
import subprocess
import sys

import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu

BIN = os.path.join(ROOT, "low-fidelity-dynamic-models_amd", "lfdm_amd", "arith_check")


def test_device_arithmetic_matches_ieee_and_long_double_references():
    assert os.path.exists(BIN), "run make -C low-fidelity-dynamic-models_amd adjcheck"
    print("[arith] device quotients, square roots and sines vs IEEE and long double, 4,194,304 operands", flush=True)
    sys.stdout.flush()
    r = subprocess.run([BIN, str(1 << 22)], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ALL OK" in r.stdout
