"""The committed per-step digests of the 32^3 workload are the oracle's: its
first three steps, rerun here from the same seeded host sources, reproduce
them bit for bit (the GPU test compares all 100 steps)."""
import json
import os
import sys

HERE = os.path.dirname(__file__)
sys.path.insert(0, os.path.join(HERE, "golden"))

import make_config1_digests  # noqa: E402


def test_oracle_reproduces_first_digests(orc):
    with open(os.path.join(HERE, "golden", "config1_digests.json")) as f:
        want = json.load(f)["steps"]
    _, got = make_config1_digests.run(3)
    for g, w in zip(got, want[:3]):
        # U's 2-norm is informational: its last bit follows the BLAS summation order
        assert abs(g.pop("U_2") - w["U_2"]) <= 1e-12 * abs(w["U_2"])
        assert g == {k: v for k, v in w.items() if k != "U_2"}
    assert [w["step"] for w in want] == list(range(1, 101))
