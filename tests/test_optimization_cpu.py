"""The one-dimensional Newton solver of the branch-length sweep (CPU)."""
import math

import numpy as np
import pytest

from fixtures import fixed_model, load_sample
from netrax_amd.optimization import newton_maximise
from oracle import oracle as O


@pytest.mark.parametrize("x0", [1e-6, 0.01, 1.0, 50.0])
def test_newton_finds_interior_maximum(x0):
    # f = 5 log x - 2 x, maximum at 2.5
    f = lambda x: (5.0 / x - 2.0, -5.0 / x**2)
    x = newton_maximise(f, x0, 1e-6, 100.0, 1e-7, 64)
    assert abs(x - 2.5) < 1e-6


def test_newton_respects_bounds():
    assert newton_maximise(lambda x: (-1.0, 0.0), 3.0, 1e-6, 100.0, 1e-7, 200) == pytest.approx(1e-6, abs=1e-6)
    assert newton_maximise(lambda x: (1.0, 0.0), 3.0, 1e-6, 100.0, 1e-7, 200) == pytest.approx(100.0, rel=1e-6)


def test_newton_on_oracle_branch_derivatives():
    """Driven by the oracle's reference-normalised derivatives, the solver
    lands where the branch derivative vanishes (or on a bound)."""
    net, parts = load_sample("two_reticulations", fixed_model())
    for e in (0, 3, net.n_edges - 1):
        def deriv(x):
            b = net.brlen.copy()
            b[e] = x
            d1, d2, _, _ = O.brlen_derivatives(net, parts, e, O.AVERAGE, mode=0, brlen=b)
            return d1, d2
        x = newton_maximise(deriv, float(net.brlen[e]), 1e-6, 100.0, 1e-7, 32)
        d1, _ = deriv(x)
        # converged (tolerance 1e-7) at a stationary point or at a bound
        assert abs(d1) < 1e-4 or x <= 1e-6 + 2e-7 or x >= 100.0 - 2e-7, (e, x, d1)
