"""The oracle against the committed golden trajectories (tests/golden): five
small cases, three steps each, every field with its ghost layer, bit for bit.
The fixtures were written by tests/golden/make_golden.py."""
import json
import os
import sys

import numpy as np
import pytest

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
sys.path.insert(0, HERE)

CASES = json.load(open(os.path.join(HERE, "cases.json")))


def expected(name):
    with np.load(os.path.join(HERE, f"{name}.npz")) as z:
        return {k: z[k] for k in z.files}


@pytest.mark.parametrize("name", sorted(CASES))
def test_oracle_reproduces_golden(vofmod, orc, name):
    import make_golden

    kw, init, first, steps = make_golden.cases()[name]
    got = make_golden.run_oracle(orc, vofmod, kw, init, first, steps)
    want = expected(name)
    for k in make_golden.FIELDS:
        np.testing.assert_array_equal(got[k], want[k], err_msg=f"{name} {k}")


def test_golden_cases_match_their_description():
    import make_golden

    cs = make_golden.cases()
    assert set(cs) == set(CASES)
    for name, (kw, _, first, steps) in cs.items():
        assert CASES[name]["first_istep"] == first and CASES[name]["steps"] == steps
        want = expected(name)
        nz = kw.get("nz", kw.get("n"))
        assert want["F"].shape[0] == nz + 2
        assert want["F"].dtype == (np.float64 if kw["precision"] == 8 else np.float32)
