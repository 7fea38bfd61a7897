"""CPU checks of the committed golden fixtures (tests/golden/, written by
tests/golden/make_fixtures.py with the FP64 oracle): the cheap ones are
recomputed by the oracle, the expensive ones are checked against laws that fix
them independently of the Hamiltonian."""
import json
import math
import os

import numpy as np
import pytest

from molspin_amd import abi, workloads
from oracle import binding as ref

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    with open(os.path.join(HERE, name + ".json")) as f:
        return json.load(f)


def trapz_decay(k, total, dt):
    """Trapezoid over the free-evolution grid of k e^{-k t}: the sum of the
    yields of a complete set of equal-rate sinks (Tr rho(t) = e^{-k t})."""
    t = np.array(ref.time_grid(total, dt))
    v = k * np.exp(-k * t)
    return float(np.sum(0.5 * np.diff(t) * (v[1:] + v[:-1])))


def test_fibonacci_fixture_matches_generator():
    fx = load("fib1024")
    _, cfg = workloads.config2(n_orientations=1024)
    assert np.array_equal(np.array(fx["points"]), np.array(cfg["fields"]))
    for i in (0, 1, 511, 1023):
        assert np.allclose(fx["points"][i], ref.fibonacci_point(1024, i, 0.05), rtol=0, atol=1e-15)


def test_config1_fixture_recomputed():
    fx = load("cfg1")
    s, cfg = workloads.config1()
    p = s.pack()
    prm = abi.run_params(total_time=fx["total_time"], time_step=fx["time_step"], method=abi.MSG_METHOD_KRYLOV)
    y, _ = ref.run_hsgeneral(p, prm, threads=8)
    assert np.abs(y[0] - np.array(fx["yields"])).max() <= 1e-12
    tr = np.array(fx["trace"])
    assert tr.shape == (4, 5001)
    # populations of S, T0, T+, T- add up to Tr rho(t) = e^{-k t} (k = 1e-3)
    t = np.array(ref.time_grid(fx["total_time"], fx["time_step"]))
    assert np.abs(tr.sum(axis=0) - np.exp(-1e-3 * t)).max() <= 1e-10


@pytest.mark.parametrize("name,k", [("cfg2_subset", 1e-3), ("cfg3_subset", 1e-3)])
def test_fixture_yields_sum_to_decay_law(name, k):
    fx = load(name)
    y = np.array(fx["yields"])
    want = trapz_decay(k, fx["total_time"], fx["time_step"])
    sums = y.reshape(-1, 4).sum(axis=1)
    assert np.abs(sums - want).max() <= 1e-10 * max(1.0, want) + 1e-13
    assert (y > 0).all()


def test_symdecay_fixture_completeness():
    fx = load("symdecay_d512")
    for row in fx["yields"]:
        assert abs(sum(row) - 1.0) <= 1e-10
        assert all(v > 0 for v in row)
