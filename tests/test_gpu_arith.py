"""The step code's device quotient and square root (qdiv, qsqrt: Newton-Raphson refinement of the
hardware reciprocal and reciprocal square root, csrc/dev_common.h) return the same bits as the
compiler's IEEE `/` and sqrt over 4M operand pairs: random exponents over the specified range and
operands shaped like the step's (tests/cpp/arith_check.hip, built by
`make -C low-fidelity-dynamic-models_amd adjcheck`)."""
import os
import subprocess
import sys

import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu

BIN = os.path.join(ROOT, "low-fidelity-dynamic-models_amd", "lfdm_amd", "arith_check")


def test_quotient_and_square_root_match_ieee_bit_for_bit():
    assert os.path.exists(BIN), "run make -C low-fidelity-dynamic-models_amd adjcheck"
    print("[arith] qdiv / qsqrt vs device '/' and sqrt, 4,194,304 operands", flush=True)
    sys.stdout.flush()
    r = subprocess.run([BIN, str(1 << 22)], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ALL OK" in r.stdout
