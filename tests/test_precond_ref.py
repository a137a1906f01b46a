"""The numpy restatement of the preconditioner's training pass
(tests/precond_ref.py) against central finite differences, and planted
mistakes in its backward pass that the same check must catch (a missing tanh
derivative, a transposed weight)."""
import numpy as np
import pytest

from . import precond_ref as R


def _pairs(args, w, idx, bug=None, h=1e-6):
    g = R.loss_and_grad(*args, w, bug=bug)[1]
    out = []
    for i in idx:
        wp, wm = w.copy(), w.copy()
        wp[i] += h
        wm[i] -= h
        fd = (R.loss_only(*args, wp) - R.loss_only(*args, wm)) / (2 * h)
        out.append((g[i], fd))
    return out


def _idx(seed):
    rng = np.random.default_rng(seed)
    return np.concatenate([rng.integers(R.W1, R.B1, 6), rng.integers(R.B1, R.W2, 3), rng.integers(R.W2, R.B2, 6),
                           rng.integers(R.B2, R.W3, 3), rng.integers(R.W3, R.B3, 6), [R.B3, R.B3 + 1]])


@pytest.mark.parametrize("seed", [1, 2])
def test_analytic_gradient_matches_finite_differences(seed):
    p = R.random_problem(seed)
    for g, fd in _pairs(p[:-1], p[-1], _idx(seed)):
        assert abs(g - fd) <= 1e-6 * max(1.0, abs(fd)) + 1e-8, (g, fd)


@pytest.mark.parametrize("bug", ["no_tanh_grad", "w2_transposed"])
def test_finite_differences_catch_planted_mistakes(bug):
    p = R.random_problem(3)
    # first-layer weights see both mistakes
    idx = np.arange(R.W1, R.W1 + 40)
    bad = _pairs(p[:-1], p[-1], idx, bug=bug)
    assert max(abs(g - fd) / max(abs(fd), 1e-9) for g, fd in bad) > 1e-2
