"""The reference's GPU known-answer tests and SolveStep(t, functor) demo, compiled as C++ against the
drop-in classes of include/lfdm_amd.hpp (tests/cpp/dropin_test.cpp, built by
`make -C low-fidelity-dynamic-models_amd dropin`)."""
import os
import subprocess

import pytest

from conftest import DATA, ROOT

BIN = os.path.join(ROOT, "low-fidelity-dynamic-models_amd", "lfdm_amd", "dropin_test")


def test_dropin_program_is_built():
    assert os.path.exists(BIN), "run make -C low-fidelity-dynamic-models_amd dropin"


@pytest.mark.gpu
def test_dropin_program_passes():
    r = subprocess.run([BIN, DATA], capture_output=True, text=True, timeout=600)
    print(r.stdout, flush=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ALL PASSED" in r.stdout
