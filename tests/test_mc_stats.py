"""Host-side statistics of the samplers' runMC protocol (correlation time,
variance of a correlated mean, convergence-check schedule), pinned by series
whose answers are known in closed form."""
import numpy as np

from cemc_amd.mcmc import correlation_time, prec_check_steps, variance_of_mean


def _ar1(rho, n, seed):
    rng = np.random.Generator(np.random.Philox(seed))
    noise = rng.standard_normal(n)
    x = np.empty(n)
    x[0] = noise[0] / np.sqrt(1.0 - rho * rho)
    for i in range(1, n):
        x[i] = rho * x[i - 1] + noise[i]
    return x


def test_correlation_time_of_ar1_series():
    """AR(1) with rho = 0.9 has autocorrelation 0.9^k: 0.531 at lag 6 and
    0.478 at lag 7, so the first lag below 1/2 is 7 and the estimate is
    tau = -1/ln(2^(-1/7)) = 7/ln 2 (montecarlo.py:505-515)."""
    info = correlation_time(_ar1(0.9, 200000, 3))
    assert info["correlation_time_found"]
    assert abs(info["correlation_time"] - 7.0 / np.log(2.0)) < 1e-12
    # rho = 0.6: 0.6 at lag 1 and 0.36 at lag 2, so tau = 2 / ln 2
    info = correlation_time(_ar1(0.6, 200000, 4))
    assert abs(info["correlation_time"] - 2.0 / np.log(2.0)) < 1e-12


def test_correlation_time_edge_cases():
    """Zero variance: 'infinite' correlation, reported as the window length;
    an AR(1) window much shorter than its correlation time never drops below
    1/2 and is reported as too short."""
    info = correlation_time(np.full(100, -3.5))
    assert info["correlation_time_found"] and info["correlation_time"] == 100
    x = _ar1(0.9999, 20, 5)
    d = x - x.mean()
    ac = np.correlate(d, d, mode="full")[19:] / (20 * np.var(d))
    info = correlation_time(x)
    assert info["correlation_time_found"] == bool(np.min(ac) <= 0.5)
    assert info["correlation_time_found"] or info["msg"].startswith("Window is too short")


def test_variance_of_mean():
    assert variance_of_mean(2.0, 10, None) == 0.2
    found = {"correlation_time_found": True, "correlation_time": 4.0}
    assert variance_of_mean(2.0, 10, found) == 2.0 * 2.0 * 4.0 / 10
    short = {"correlation_time_found": True, "correlation_time": 0.3}
    assert variance_of_mean(-2.0, 10, short) == 2.0 * 2.0 * 1.0 / 10
    assert variance_of_mean(2.0, 10, {"correlation_time_found": False, "correlation_time": 9.0}) == 0.2


def test_prec_check_schedule():
    """The reference tests after the first step past each multiple of the
    check interval: steps 3, 6, 8 for an interval of 2.5 in 10 steps."""
    assert prec_check_steps(2.5, 10) == [3, 6, 8]
    assert prec_check_steps(4.0, 13) == [5, 9]
    assert prec_check_steps(20.0, 10) == []
