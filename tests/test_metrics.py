"""The host-side arithmetic behind the learned part's reported figures, on the CPU: the plain resolution
that matches a learned error and its time (log-log interpolation along the plain ladder,
npde.learned.pipeline.matched_speedup), checked against closed forms of exact power laws, and the
Pareto front of the accuracy-versus-cost table (npde.learned.pareto.pareto_flags), checked against hand
cases.  Power laws make log-log interpolation exact, so interpolating in n or t instead of log n and
log t, swapping the bracketing levels, or taking the time of the wrong segment fails them."""
import json
import math
from pathlib import Path

import numpy as np
import pytest

from npde.learned.pareto import pareto_flags
from npde.learned.pipeline import matched_speedup

LADDER = np.array([64.0, 128.0, 256.0])


def power_ladder(c, p, d, q, ns=LADDER):
    """err = c n^-p, time = d n^q on the ladder."""
    return c * ns ** -p, d * ns ** q


@pytest.mark.parametrize("p,q", [(0.83, 2.0), (1.2, 3.0), (0.5, 2.4)])
@pytest.mark.parametrize("n_true", [70.0, 105.66, 128.0, 200.0, 255.0])
def test_matched_resolution_is_exact_on_power_laws(p, q, n_true):
    errs, times = power_ladder(5.2, p, 1.5e-6, q)
    e = 5.2 * n_true ** -p
    t_learned = 0.034
    n_star, t_star, sp = matched_speedup(LADDER, errs, times, e, t_learned)
    assert n_star == pytest.approx(n_true, rel=1e-12)
    assert t_star == pytest.approx(1.5e-6 * n_true ** q, rel=1e-12)
    assert sp == pytest.approx(t_star / t_learned, rel=1e-15)


def test_matched_resolution_at_a_ladder_level_is_that_level():
    errs = np.array([0.0814, 0.0457, 0.0204])
    times = np.array([0.025, 0.100, 0.531])
    for k in range(3):
        n_star, t_star, _ = matched_speedup(LADDER, errs, times, errs[k], 1.0)
        assert n_star == pytest.approx(LADDER[k], rel=1e-12)
        assert t_star == pytest.approx(times[k], rel=1e-12)


def test_errors_beyond_the_ladder():
    errs, times = power_ladder(3.0, 0.9, 2e-6, 2.5)
    # finer than the finest level: the last segment's power laws, extended
    n_true = 400.0
    n_star, t_star, _ = matched_speedup(LADDER, errs, times, 3.0 * n_true ** -0.9, 1.0)
    assert n_star == pytest.approx(n_true, rel=1e-12)
    assert t_star == pytest.approx(2e-6 * n_true ** 2.5, rel=1e-12)
    # no better than the coarsest level: that level's resolution and time
    for e in (errs[0], 2 * errs[0]):
        n_star, t_star, sp = matched_speedup(LADDER, errs, times, e, 0.5)
        assert (n_star, t_star) == (LADDER[0], times[0]) and sp == times[0] / 0.5


def test_bent_ladder_uses_the_bracketing_segment():
    """Different exponents on the two segments: a target between 128 and 256 takes the second segment's
    laws only."""
    errs = np.array([0.08, 0.04, 0.025])      # p = 1 on the first segment, log2(0.04/0.025) on the second
    times = np.array([0.02, 0.09, 0.6])
    e = 0.03
    n_star, t_star, _ = matched_speedup(LADDER, errs, times, e, 1.0)
    p2 = math.log(0.04 / 0.025) / math.log(2.0)
    q2 = math.log(0.6 / 0.09) / math.log(2.0)
    n_exp = 128.0 * (0.04 / e) ** (1.0 / p2)
    assert n_star == pytest.approx(n_exp, rel=1e-12)
    assert t_star == pytest.approx(0.09 * (n_exp / 128.0) ** q2, rel=1e-12)


def test_committed_evaluation_is_internally_consistent():
    """The committed evaluation's matched resolution, time and ratio follow from its own ladder and
    learned row."""
    path = Path(__file__).resolve().parents[1] / "neural-pde-surrogate_amd" / "npde" / "learned" / "results" / \
        "riemann64_eval.json"
    r = json.loads(path.read_text())
    ns = np.array(sorted(int(k) for k in r["plain"]), dtype=float)
    errs = np.array([r["plain"][str(int(n))]["l1_rho"] for n in ns])
    times = np.array([r["plain"][str(int(n))]["seconds"] for n in ns])
    learned = r["learned64"]
    n_star, t_star, sp = matched_speedup(ns, errs, times, learned["l1_rho"], learned["seconds"])
    assert n_star == pytest.approx(r["matched"]["plain_resolution"], rel=1e-12)
    assert t_star == pytest.approx(r["matched"]["plain_seconds"], rel=1e-12)
    assert sp == pytest.approx(r["matched"]["speedup"], rel=1e-12)
    assert 64 < n_star < 128 and learned["l1_rho"] < r["plain"]["64"]["l1_rho"]


def row(cost, rmse):
    return {"cost": cost, "rmse_last": rmse}


def test_pareto_front_hand_cases():
    rows = [row(1.0, 0.30), row(2.0, 0.20), row(2.5, 0.25), row(4.0, 0.10), row(4.0, 0.12), row(8.0, 0.10)]
    # (2.5, 0.25) is beaten by (2.0, 0.20); (4.0, 0.12) by (4.0, 0.10) at equal cost; (8.0, 0.10) by
    # (4.0, 0.10) at equal accuracy
    assert pareto_flags(rows) == [True, True, False, True, False, False]
    # identical points do not remove each other
    assert pareto_flags([row(1.0, 0.2), row(1.0, 0.2)]) == [True, True]
    assert pareto_flags([row(1.0, 0.2)]) == [True]


def test_pareto_front_of_the_committed_table():
    """The committed accuracy-cost table marks exactly the rows the rule selects."""
    path = Path(__file__).resolve().parents[1] / "neural-pde-surrogate_amd" / "npde" / "learned" / "results" / \
        "pareto.json"
    rows = json.loads(path.read_text())["rows"]
    assert [r["pareto"] for r in rows] == pareto_flags(rows)
