"""CPU checks of trace_bayes_minimize, the host search behind
trace_phr_optimize_bayes (bayesianOptimizer, optimizer.go:96-155).

The reference's search is the third-party module changkun.de/x/bo v0.0.1,
which is absent here, and it draws from Go's math/rand: the sequence of points
it would try is parity unpinned.  What is pinned here: every point lies in the
box, the values are the objective's at those points, the result is the first
evaluated minimum, the same seed gives the same sequence, and on smooth
functions the model rounds find the minimum far better than the same number of
uniform samples."""
import numpy as np
import pytest

from trace_amd import _lib
from trace_amd.device import bayes_minimize


def bowl(c):
    c = np.asarray(c, np.float64)
    return lambda x: float(np.sum((np.asarray(x) - c) ** 2))


def test_points_in_box_values_and_first_minimum():
    lo, hi = [0.4, 6.0], [0.55, 9.0]
    f = bowl([0.47, 7.3])
    best, xs, ys = bayes_minimize(f, lo, hi, rounds=12, random_rounds=4, seed=3)
    assert xs.shape == (12, 2) and ys.shape == (12,)
    assert np.all(xs >= np.array(lo)) and np.all(xs <= np.array(hi))
    assert np.array_equal(ys, np.array([f(x) for x in xs]))
    assert np.array_equal(best, xs[int(np.argmin(ys))])  # argmin is the first minimum


def test_first_minimum_wins_on_ties():
    # a constant objective: every value ties, the first point is the result
    best, xs, ys = bayes_minimize(lambda x: 1.0, [0.0], [1.0], rounds=6, random_rounds=2, seed=9)
    assert np.all(ys == 1.0)
    assert np.array_equal(best, xs[0])


def test_seeded_and_deterministic():
    f = bowl([0.2, 0.8, 0.5])
    a = bayes_minimize(f, [0, 0, 0], [1, 1, 1], rounds=10, random_rounds=3, seed=5)
    b = bayes_minimize(f, [0, 0, 0], [1, 1, 1], rounds=10, random_rounds=3, seed=5)
    c = bayes_minimize(f, [0, 0, 0], [1, 1, 1], rounds=10, random_rounds=3, seed=6)
    assert np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2])
    assert not np.array_equal(a[1][0], c[1][0])


def test_random_rounds_are_uniform_samples():
    # with rounds == random_rounds no model is fitted: the points are the
    # seeded uniform stream, so their mean over many rounds is the box centre
    _, xs, _ = bayes_minimize(lambda x: float(x[0]), [2.0, -1.0], [4.0, 1.0], rounds=2000, random_rounds=2000,
                              seed=11)
    assert abs(xs[:, 0].mean() - 3.0) < 0.05 and abs(xs[:, 1].mean()) < 0.05
    assert xs[:, 0].min() < 2.01 and xs[:, 0].max() > 3.99


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_model_rounds_beat_uniform_sampling(seed):
    lo, hi = np.array([0.4, 6.0]), np.array([0.55, 9.0])
    c = np.array([0.512, 6.4])
    scale = hi - lo
    f = lambda x: float(np.sum(((np.asarray(x) - c) / scale) ** 2))  # noqa: E731
    best, xs, ys = bayes_minimize(f, lo, hi, rounds=20, random_rounds=5, seed=seed)
    # the model rounds reach the minimum to within 2 % of the box
    assert np.all(np.abs((best - c) / scale) < 0.02), best
    # twenty uniform samples of the same stream land much further away
    _, _, yr = bayes_minimize(f, lo, hi, rounds=20, random_rounds=20, seed=seed)
    assert ys.min() < 0.1 * yr.min()


def test_one_dimension_with_a_kink():
    f = lambda x: abs(float(x[0]) - 0.3)  # noqa: E731
    best, _, ys = bayes_minimize(f, [0.0], [1.0], rounds=16, random_rounds=3, seed=2)
    assert abs(best[0] - 0.3) < 0.01


def test_errors():
    with pytest.raises(_lib.TraceError):
        bayes_minimize(lambda x: 0.0, [1.0], [0.0], rounds=3)  # lo > hi
    with pytest.raises(_lib.TraceError):
        bayes_minimize(lambda x: float("nan"), [0.0], [1.0], rounds=3)
    with pytest.raises(_lib.TraceError):
        bayes_minimize(lambda x: 0.0, np.zeros(9), np.ones(9), rounds=3)  # more than 8 dims

    def boom(x):
        raise KeyError("objective failed")

    with pytest.raises(KeyError):
        bayes_minimize(boom, [0.0], [1.0], rounds=3)
    lib = _lib.load()
    assert lib.trace_bayes_minimize(None, None, 1, 1, 0, 0, _lib.OBJECTIVE_FN(lambda u, x, y: 0), None, None, None,
                                    None) == _lib.TRACE_E_INVALID
