"""Time-dependent Hamiltonians (dynamics=dynamic): Zeeman fields with
linearpolarized / circularpolarized time dependence.

CPU: the oracle's dynamic loop (midpoint Hamiltonian per interval, RK4 stages at
t, t+dt/2, t+dt; TaskHSGeneral.cpp:487-537) is pinned by the Rabi closed form of a
spin in a rotating field (with the expected convergence orders) and by the static
limit of a non-oscillating field.
GPU: msg_propagate with dynamics=dynamic (matrix-free Lanczos with per-interval
electron fields; dense path with per-interval propagators) against that oracle.
"""
import math

import numpy as np
import pytest

from molspin_amd import abi
from molspin_amd import workloads
from molspin_amd.system import SpinSystem
from oracle import binding as ref

YIELD_TOL = 1e-5
TRACE_TOL = 1e-4


def rabi_system(b0, b1, freq):
    """One electron (g = 2.0023) in B0 along z and B1 rotating in the xy plane at
    `freq` rad/ns (circularpolarized about z), starting spin up."""
    s = SpinSystem("rabi")
    s.add_spin("e", "1/2", electron=True, tensor=2.0023)
    s.add_zeeman(["e"], (0.0, 0.0, b0))
    s.add_zeeman(["e"], (b1, 0.0, 0.0), time_dependence="circularpolarized", frequency=freq)
    s.add_state("up", ["e"], [1, 0])
    s.set_initial(["up"])
    return s


def rabi_exact(b0, b1, freq, times):
    """P_up(t) = 1 - (w1/W)^2 sin^2(W t / 2), W^2 = (w0 - w)^2 + w1^2, in the frame
    rotating with the field (w0, w1: Larmor frequencies of B0, B1)."""
    g = 2.0023 * 8.79410005e1  # rad/ns/T (common prefactor x g)
    w0, w1 = g * b0, g * b1
    W = math.sqrt((w0 - freq) ** 2 + w1 ** 2)
    return 1.0 - (w1 / W) ** 2 * np.sin(0.5 * W * np.asarray(times)) ** 2


def test_rabi_closed_form_and_orders():
    """Midpoint exponential: second order in dt; RK4: fourth order (1e-3 T fields)."""
    b0, b1, freq = 1e-3, 2e-4, 0.17
    s = rabi_system(b0, b1, freq)
    p = s.pack()
    errs = {}
    for method in (abi.MSG_METHOD_EXPONENTIAL, abi.MSG_METHOD_RK4):
        for dt in (0.02, 0.01):
            prm = abi.run_params(calculation=abi.MSG_CALC_TIMEEVOLUTION, observables=abi.MSG_OBS_STATE_POPULATIONS,
                                 method=method, total_time=40.0, time_step=dt, dynamics=True)
            _, tr = ref.run_hsgeneral(p, prm, trace=True)
            exact = rabi_exact(b0, b1, freq, ref.time_grid(40.0, dt))
            errs[(method, dt)] = np.abs(tr[0, :, 0] - exact).max()
    e1, e2 = errs[(abi.MSG_METHOD_EXPONENTIAL, 0.02)], errs[(abi.MSG_METHOD_EXPONENTIAL, 0.01)]
    assert e2 < 1e-4 and 3.0 < e1 / e2 < 5.0
    r1, r2 = errs[(abi.MSG_METHOD_RK4, 0.02)], errs[(abi.MSG_METHOD_RK4, 0.01)]
    assert r2 < 1e-7 and 12.0 < r1 / r2 < 20.0
    # the static plan keeps the field as written (B1 fixed along x): the w = 0 closed form
    prm = abi.run_params(calculation=abi.MSG_CALC_TIMEEVOLUTION, observables=abi.MSG_OBS_STATE_POPULATIONS,
                         method=abi.MSG_METHOD_EXPONENTIAL, total_time=40.0, time_step=0.01)
    _, trs = ref.run_hsgeneral(p, prm, trace=True)
    assert np.abs(trs[0, :, 0] - rabi_exact(b0, b1, 0.0, ref.time_grid(40.0, 0.01))).max() < 1e-10


def rp_system(n_nuc=3, td="linearpolarized", rf=(0.0, 0.0, 0.0), freq=0.3, phase=0.4, axis=(0.0, 0.0, 1.0),
              perpendicular=False, rates=None, nuclear_rf=False):
    """Radical pair with anisotropic hyperfine, a static 50 uT field and a
    time-dependent Zeeman field on both electrons (or on nucleus N0)."""
    rng = np.random.default_rng(7)
    s = SpinSystem("rydmr")
    s.add_spin("E1", "1/2", electron=True, tensor=2.0023)
    s.add_spin("E2", "1/2", electron=True, tensor=2.0023)
    for i in range(n_nuc):
        s.add_spin("N%d" % i, "1/2", tensor=1.0)
        A = rng.normal(size=(3, 3))
        A = 0.5 * (A + A.T)
        A = A - np.trace(A) / 3 * np.eye(3) + (0.3 + 0.8 * rng.random()) * np.eye(3)
        s.add_hyperfine(["E1" if i % 2 == 0 else "E2"], ["N%d" % i], tensor=A.tolist(), prefactor=0.001)
    s.add_zeeman(["E1", "E2"], (0.0, 0.0, 5e-5), swept=True)
    targets = ["N0"] if nuclear_rf else ["E1", "E2"]
    s.add_zeeman(targets, rf, time_dependence=td, frequency=freq, phase=phase, axis=axis,
                 perpendicular=perpendicular, prefactor=1.0)
    s.add_radical_pair_states("E1", "E2")
    for name, k in zip(("S", "T0", "Tp", "Tm"), rates or (1e-3,) * 4):
        s.add_transition("k" + name, name, k)
    s.set_initial(["S"])
    return s


def test_non_oscillating_field_equals_the_static_run():
    """frequency 0, phase phi: the linearly polarized field is constant at B cos(phi),
    so the dynamic run equals the static run of a field scaled by cos(phi)."""
    rf = (3e-4, -1e-4, 2e-4)
    dyn = rp_system(td="linearpolarized", rf=rf, freq=0.0, phase=0.7).pack()
    sta = rp_system(td="linearpolarized", rf=[x * math.cos(0.7) for x in rf], freq=0.0, phase=0.0).pack()
    prm = abi.run_params(method=abi.MSG_METHOD_EXPONENTIAL, total_time=60.0, time_step=1.0, dynamics=True)
    yd, _ = ref.run_hsgeneral(dyn, prm)
    prm_s = abi.run_params(method=abi.MSG_METHOD_EXPONENTIAL, total_time=60.0, time_step=1.0)
    ys, _ = ref.run_hsgeneral(sta, prm_s)
    assert np.abs(yd - ys).max() < 1e-13
    # ... and an oscillating field changes the yields
    osc = rp_system(td="linearpolarized", rf=rf, freq=0.3, phase=0.7).pack()
    yo, _ = ref.run_hsgeneral(osc, prm)
    assert np.abs(yo - ys).max() > 1e-6


def test_circular_field_along_its_own_axis():
    """A circularly polarized field parallel to its axis does not rotate (the static
    field as written); with perpendicularoscillations its perpendicular part, zero,
    is all that acts (the run without that field)."""
    ax = (1 / math.sqrt(2), 1 / math.sqrt(2), 0.0)
    v = tuple(3e-4 * x for x in ax)
    prm_d = abi.run_params(method=abi.MSG_METHOD_EXPONENTIAL, total_time=40.0, time_step=1.0, dynamics=True)
    prm_s = abi.run_params(method=abi.MSG_METHOD_EXPONENTIAL, total_time=40.0, time_step=1.0)
    yd, _ = ref.run_hsgeneral(rp_system(td="circularpolarized", rf=v, freq=0.5, axis=ax).pack(), prm_d)
    ys, _ = ref.run_hsgeneral(rp_system(td="circularpolarized", rf=v, freq=0.5, axis=ax).pack(), prm_s)
    assert np.abs(yd - ys).max() < 1e-12
    yp, _ = ref.run_hsgeneral(rp_system(td="circularpolarized", rf=v, freq=0.5, axis=ax, perpendicular=True).pack(),
                              prm_d)
    y0, _ = ref.run_hsgeneral(rp_system(td="circularpolarized", rf=(0.0, 0.0, 0.0)).pack(), prm_s)
    assert np.abs(yp - y0).max() < 1e-12


# ---------------------------------------------------------------------------
# Device against the oracle
# ---------------------------------------------------------------------------

@pytest.fixture(scope="module")
def ctx():
    from molspin_amd import backend
    c = backend.Context(0)
    yield c
    c.close()


def trace_err(dev, ora):
    return np.abs(dev - ora).max() / np.abs(ora).max()


FIELDS = [(0.0, 0.0, 5e-5), (2e-5, 0.0, 4e-5), (0.0, 1e-4, 1e-4)]


@pytest.mark.gpu
@pytest.mark.parametrize("td", ["linearpolarized", "circularpolarized"])
def test_device_lanczos_dynamic(ctx, td):
    """D = 256 (6 nuclei) on the matrix-free Lanczos kernel with per-interval electron
    fields, three field points, yields and traces against the oracle (Krylov m = 16)."""
    s = rp_system(n_nuc=6, td=td, rf=(2e-4, 1e-4, 0.0), freq=0.25, phase=0.3, axis=(0.0, 0.6, 0.8),
                  perpendicular=(td == "circularpolarized"))
    p = s.pack()
    sysd = ctx.upload(p)
    prm = abi.run_params(calculation=abi.MSG_CALC_TIMEEVOLUTION, method=abi.MSG_METHOD_KRYLOV,
                         total_time=120.0, time_step=2.0, dynamics=True)
    _, tr = sysd.propagate(prm, fields=FIELDS, trace=True)
    _, tro = ref.run_hsgeneral(p, prm, fields=FIELDS, trace=True, threads=8)
    print(td, "trace err", trace_err(tr, tro))
    assert trace_err(tr, tro) <= TRACE_TOL
    prm_y = abi.run_params(method=abi.MSG_METHOD_KRYLOV, total_time=120.0, time_step=2.0, dynamics=True)
    y, _ = sysd.propagate(prm_y, fields=FIELDS)
    yo, _ = ref.run_hsgeneral(p, prm_y, fields=FIELDS, threads=8)
    print(td, "yields", y[:, 0], yo[:, 0])
    assert np.abs(y - yo).max() <= YIELD_TOL


@pytest.mark.gpu
def test_device_lanczos_dynamic_large(ctx):
    """D = 16384 (12 nuclei, LDS-mirrored streaming kernel) with 4 trace-sample columns
    and a circularly polarized field, 5 intervals, state populations per sample."""
    s = rp_system(n_nuc=12, td="circularpolarized", rf=(3e-4, 0.0, 0.0), freq=0.4)
    p = s.pack()
    sysd = ctx.upload(p)
    assert sysd.dim == 16384
    M = 4
    cols = ref.coherent_samples(p, s.state_index("S"), 5, M) / math.sqrt(M)
    prm = abi.run_params(calculation=abi.MSG_CALC_TIMEEVOLUTION, observables=abi.MSG_OBS_STATE_POPULATIONS,
                         method=abi.MSG_METHOD_KRYLOV, sampling=abi.MSG_SAMPLING_GIVEN, total_time=10.0,
                         time_step=2.0, dynamics=True)
    _, tr = sysd.propagate(prm, columns=cols, trace=True)
    _, tro = ref.run_hsgeneral(p, prm, columns=cols, trace=True, threads=8)
    print("D=16384 dynamic trace err", trace_err(tr, tro))
    assert trace_err(tr, tro) <= TRACE_TOL


@pytest.mark.gpu
@pytest.mark.parametrize("method", [abi.MSG_METHOD_EXPONENTIAL, abi.MSG_METHOD_RK4])
def test_device_dense_dynamic(ctx, method):
    """Dense path, per-interval propagators: midpoint exponential and RK4 stages, unequal
    sinks (K inside the propagator), against the oracle; a run split into 7-interval
    propagator chunks (state saved and resumed) gives the same numbers."""
    import os
    s = rp_system(n_nuc=3, td="linearpolarized", rf=(2e-4, 0.0, 1e-4), freq=0.2, rates=(3e-3, 1e-3, 2e-3, 2e-3))
    p = s.pack()
    sysd = ctx.upload(p)
    prm = abi.run_params(calculation=abi.MSG_CALC_TIMEEVOLUTION, method=method, total_time=100.0, time_step=0.5,
                         dynamics=True)
    _, tr = sysd.propagate(prm, fields=FIELDS, trace=True)
    _, tro = ref.run_hsgeneral(p, prm, fields=FIELDS, trace=True, threads=8)
    print("dense dynamic method", method, "trace err", trace_err(tr, tro))
    assert trace_err(tr, tro) <= TRACE_TOL
    os.environ["MSG_DYN_CHUNK"] = "7"
    try:
        _, tr7 = sysd.propagate(prm, fields=FIELDS, trace=True)
    finally:
        del os.environ["MSG_DYN_CHUNK"]
    assert np.abs(tr7 - tr).max() <= 1e-12


@pytest.mark.gpu
def test_device_dense_dynamic_nuclear_field(ctx):
    """A time-dependent field on a nucleus (dense path, AUTO) against the oracle."""
    s = rp_system(n_nuc=3, td="linearpolarized", rf=(0.0, 0.0, 0.02), freq=0.05, nuclear_rf=True)
    p = s.pack()
    sysd = ctx.upload(p)
    prm = abi.run_params(method=abi.MSG_METHOD_AUTO, total_time=200.0, time_step=1.0, dynamics=True)
    y, _ = sysd.propagate(prm, fields=FIELDS)
    prm_o = abi.run_params(method=abi.MSG_METHOD_EXPONENTIAL, total_time=200.0, time_step=1.0, dynamics=True)
    yo, _ = ref.run_hsgeneral(p, prm_o, fields=FIELDS, threads=8)
    print("nuclear rf", y[:, 0], yo[:, 0])
    assert np.abs(y - yo).max() <= YIELD_TOL


@pytest.mark.gpu
def test_device_dynamic_rejections(ctx):
    """msg_step takes a static Hamiltonian; time infinity is static-only (the
    reference's message); the server does not take dense dynamic requests."""
    from molspin_amd import backend
    s = rp_system(n_nuc=3, td="linearpolarized", rf=(2e-4, 0.0, 0.0))
    p = s.pack()
    sysd = ctx.upload(p)
    prm = abi.run_params(dynamics=True)
    with pytest.raises(backend.BackendError) as e:
        sysd.step(prm, ref.direct_factors(p), 1.0)
    assert e.value.status == abi.MSG_UNSUPPORTED
    with pytest.raises(backend.BackendError) as e:
        sysd.time_infinity(prm)
    assert e.value.status == abi.MSG_INVALID_ARG and "finite-time" in str(e.value)
    srv = backend.Server(ctx, 2, 1 << 30)
    try:
        with pytest.raises(backend.BackendError) as e:
            srv.submit(sysd, abi.run_params(method=abi.MSG_METHOD_AUTO, total_time=10.0, dynamics=True), [FIELDS[0]])
        assert e.value.status == abi.MSG_UNSUPPORTED
    finally:
        srv.close()
