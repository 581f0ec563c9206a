"""The oracle still reproduces the committed golden fixtures (tests/golden/make_golden.py) bit for bit."""
from pathlib import Path

import numpy as np
import pytest

from golden.make_golden import golden_arrays

HERE = Path(__file__).resolve().parent / "golden"


@pytest.mark.parametrize("dtype,name", [(np.float32, "tiny_fp32.npz"), (np.float64, "tiny_fp64.npz")])
def test_oracle_reproduces_golden(dtype, name):
    want = np.load(HERE / name)
    got = golden_arrays(dtype)
    assert set(got) == set(want.files)
    for k in want.files:
        np.testing.assert_array_equal(got[k], want[k], err_msg=k)
