"""CPU checks of the batched workload's problem generator."""
import numpy as np

from classic_amd import problems as P


def test_specs_reproducible_and_in_range():
    a = P.batch_problem_specs(256, seed=0)
    b = P.batch_problem_specs(256, seed=0)
    c = P.batch_problem_specs(256, seed=1)
    assert a == b and a != c
    sides = np.array([[s["mx"], s["my"]] for s in a])
    assert sides.min() >= 128 and sides.max() <= 1024
    # log-uniform: the median side is near sqrt(128 * 1024) = 362
    assert 250 < np.median(sides) < 520
    kinds = [s["solver"] for s in a]
    assert 80 < kinds.count("euler") < 176
    for s in a:
        if s["solver"] == "acoustics":
            rho, bulk = s["params"]
            assert 0.5 <= rho <= 4.0 and 1.0 <= bulk <= 8.0
            assert s["mthbc"] == (3, 3, 3, 3)
        else:
            assert s["params"] == (1.4, 1.0) and s["mthbc"] == (1, 1, 1, 1)


def test_problem_initial_data_shapes():
    for s in P.batch_problem_specs(6, seed=3, side_lo=8, side_hi=40):
        q, dx, dy = P.batch_problem_q(s)
        meqn = 4 if s["solver"] == "euler" else 3
        assert q.shape == (s["my"] + 4, s["mx"] + 4, meqn)
        assert dx > 0 and dy > 0
        inner = q[2:-2, 2:-2]
        assert np.all(inner[..., 0] > 0)
