"""CPU checks of the Strouhal-measurement helper used by the unsteady tests."""
import numpy as np

from lco import saturated_start, strouhal


def test_strouhal_of_saturating_sinusoid():
    """A growing-then-saturated sine of known frequency sampled at dt = 0.2: the
    helper finds the saturated window and recovers the frequency."""
    t = np.arange(0.0, 80.0 + 1e-9, 0.2)
    f0 = 0.155
    amp = 0.35 * np.tanh(np.exp(0.25 * (t - 30.0)))
    y = 0.01 + amp * np.sin(2 * np.pi * f0 * t + 0.3)
    r = strouhal(t, y)
    assert t[r["i0"]] > 25.0
    assert abs(r["st"] - f0) < 2e-3
    assert abs(r["amp"] - 0.35) < 0.01


def test_saturated_start_of_pure_cycle_is_early():
    t = np.arange(0.0, 60.0, 0.2)
    y = np.sin(2 * np.pi * 0.2 * t)
    assert t[saturated_start(y)] < 5.0
