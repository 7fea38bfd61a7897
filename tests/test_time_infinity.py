"""Time-infinity yields (HSGeneral method=timeinf).

CPU: the oracle's D^2 x D^2 time-integral solve is pinned by the reference's own
closed forms (tests_HSGeneral.cpp:1410-1431 and :1704-1730), by scipy's
Bartels-Stewart Lyapunov solver, by the long-time limit of the finite-time loop
and by the eigen closed form for K = kappa 1.
GPU: msg_time_infinity (both device routes) against that oracle; FP64 on both
sides, so the bar is 1e-9 absolute on the yields.
"""
import math

import numpy as np
import pytest

from molspin_amd import abi
from molspin_amd import workloads
from molspin_amd.system import SpinSystem
from oracle import binding as ref

TI_TOL = 1e-9


def powder_yield_system():
    """BuildHSGeneralPowderYieldSystem (tests_HSGeneral.cpp:99-121): two electrons,
    a decoupled nucleus, Zeeman 1 mT without tensors or common prefactor, a singlet
    sink at 0.02 /ns, singlet initial state."""
    s = SpinSystem("powder")
    s.add_spin("E1", "1/2", electron=True)
    s.add_spin("E2", "1/2", electron=True)
    s.add_spin("N", "1/2")
    s.add_zeeman(["E1", "E2"], (0.0, 0.0, 0.001), common_prefactor=False, ignore_tensors=True)
    r = 1.0 / math.sqrt(2.0)
    s.add_state("Singlet", ["E1", "E2"], [0, r, -r, 0])
    s.add_transition("sink", "Singlet", 0.02)
    s.set_initial(["Singlet"])
    return s


def unequal_sink_system(n_nuc=2, rates=(0.05, 0.02, 0.03, 0.03), field=(0.0, 0.02, 0.05)):
    """Radical pair with hyperfine on both electrons and four sinks at unequal rates
    (K not proportional to 1, every state decays)."""
    s = SpinSystem("unequal")
    s.add_spin("E1", "1/2", electron=True, tensor=2.0023)
    s.add_spin("E2", "1/2", electron=True, tensor=2.0023)
    for i, a in enumerate([0.9, 1.3, 0.4, 0.7][:n_nuc]):
        s.add_spin("N%d" % i, "1/2", tensor=1.0)
        s.add_hyperfine(["E1" if i % 2 == 0 else "E2"], ["N%d" % i], tensor=a, prefactor=0.001)
    s.add_zeeman(["E1", "E2"], field, prefactor=0.001)
    s.add_radical_pair_states("E1", "E2")
    for name, k in zip(("S", "T0", "Tp", "Tm"), rates):
        s.add_transition("k" + name, name, k)
    s.set_initial(["S"])
    return s


# ---------------------------------------------------------------------------
# Oracle pins (CPU)
# ---------------------------------------------------------------------------

def test_time_integral_reference_known_answer():
    """tests_HSGeneral.cpp:1704-1730: H = 0, K = 0.1 I, rho0 = e_0 e_0^T gives X = 5 rho0."""
    D = 8
    rho0 = np.zeros((D, D), complex)
    rho0[0, 0] = 1.0
    X = ref.lyapunov_integral(-0.1 * np.eye(D, dtype=complex), rho0)
    assert np.abs(X - 5.0 * rho0).max() < 1e-11


@pytest.mark.parametrize("D", [4, 9, 16])
def test_time_integral_matches_bartels_stewart(D):
    """The Kronecker-form solve equals scipy's Bartels-Stewart solver (a library
    routine) on random stable generators."""
    import scipy.linalg as sl
    rng = np.random.default_rng(D)
    A = rng.normal(size=(D, D)) + 1j * rng.normal(size=(D, D))
    B = rng.normal(size=(D, D)) + 1j * rng.normal(size=(D, D))
    G = -1j * (A + A.conj().T) / 2 - (B @ B.conj().T / D + 0.05 * np.eye(D))
    C = rng.normal(size=(D, 3)) + 1j * rng.normal(size=(D, 3))
    rho0 = C @ C.conj().T
    X = ref.lyapunov_integral(G, rho0)
    Xs = sl.solve_continuous_lyapunov(G, -rho0)
    assert np.abs(X - Xs).max() <= 1e-12 * np.abs(Xs).max()


def test_timeinf_singlet_sink_yield_is_one():
    """tests_HSGeneral.cpp:1410-1431: starting in the singlet with a single singlet
    sink and no singlet-triplet mixing, the infinite-time product yield is exactly 1."""
    p = powder_yield_system().pack()
    y = ref.time_infinity(p, abi.run_params())
    assert y.shape == (1,) and abs(y[0] - 1.0) < 1e-10


def test_singular_time_integral_takes_the_minimum_norm_solution():
    """The powder fixture's triplets never decay and are never populated, so the
    D^2 x D^2 system is exactly singular; like arma::solve the oracle then returns
    the minimum-norm least-squares solution (numpy's lstsq, a library routine)."""
    p = powder_yield_system().pack()
    H = ref.hamiltonian(p)
    K = ref.reaction(p)
    B = ref.direct_factors(p)
    rho0 = B @ B.conj().T
    D = H.shape[0]
    G = -1j * H - K
    L = np.kron(G, np.eye(D)) + np.kron(np.eye(D), G.conj())
    assert np.linalg.matrix_rank(L) < D * D
    x = np.linalg.lstsq(L, -rho0.reshape(-1), rcond=1e-10)[0]
    X = ref.lyapunov_integral(G, rho0)
    assert np.abs(X - x.reshape(D, D)).max() < 1e-12
    # populations: the singlet decays at 0.02 /ns (integral 50), the triplets stay empty
    y = ref.time_infinity(p, abi.run_params(observables=abi.MSG_OBS_STATE_POPULATIONS))
    assert abs(y[0] - 50.0) < 1e-9


def test_timeinf_is_the_long_time_limit_of_the_finite_loop():
    """Unequal sinks (K not proportional to 1): the time-infinity yields equal the
    finite-time trapezoid yields of the oracle's own loop over [0, 2500] ns,
    Richardson-extrapolated in dt (the tail e^{-0.01 t} beyond 2500 ns is < 1e-10)."""
    p = unequal_sink_system().pack()
    yt = ref.time_infinity(p, abi.run_params())
    ys = []
    for dt in (0.5, 0.25):
        prm = abi.run_params(total_time=2500.0, time_step=dt, method=abi.MSG_METHOD_EXPONENTIAL)
        ys.append(ref.run_hsgeneral(p, prm)[0][0])
    extrap = (4.0 * ys[1] - ys[0]) / 3.0
    assert np.abs(extrap - yt).max() < 1e-7
    assert abs(yt.sum() - 1.0) < 1e-10  # every state decays into one of the four sinks


def test_timeinf_symmetric_matches_eigen_closed_form():
    """K = kappa 1 (config 1, four equal sinks): the solve equals the eigenbasis sum
    Re sum_gh conj(P'_gh) rho'_gh / (2 kappa + i (l_g - l_h)) computed with numpy;
    the singlet yield (rho0 proportional to P_S) also equals the symmetric-decay
    yield at rate 2 kappa, whose oracle the reference's golden values pin."""
    s, cfg = workloads.config1()
    p = s.pack()
    prm = abi.run_params()
    yt = ref.time_infinity(p, prm)
    H = ref.hamiltonian(p)
    B = ref.direct_factors(p)
    rho0 = B @ B.conj().T
    lam, V = np.linalg.eigh(H)
    rp = V.conj().T @ rho0 @ V
    k = 1e-3
    w = lam[:, None] - lam[None, :]
    for i, t in enumerate(ref.packed_transitions(p)):
        P = ref.projector(p, p.desc.transitions[t].source_state)
        pp = V.conj().T @ P @ V
        val = (np.conj(pp) * rp / (k + 1j * w)).sum().real * k
        assert abs(val - yt[i]) < 1e-10
    ysd = ref.symmetric_decay(p, k)
    assert abs(yt[0] - ysd[s.state_index("S")]) < 1e-10


# ---------------------------------------------------------------------------
# Device (msg_time_infinity) against the oracle
# ---------------------------------------------------------------------------

@pytest.fixture(scope="module")
def ctx():
    from molspin_amd import backend
    c = backend.Context(0)
    yield c
    c.close()


@pytest.mark.gpu
@pytest.mark.parametrize("method", [abi.MSG_METHOD_AUTO, abi.MSG_METHOD_EXPONENTIAL])
@pytest.mark.parametrize("obs", [abi.MSG_OBS_TRANSITION_YIELDS, abi.MSG_OBS_STATE_POPULATIONS])
def test_device_config1(ctx, method, obs):
    """Config 1 (D = 48, K = kappa 1): eigen route (AUTO) and doubling route
    (EXPONENTIAL) against the oracle's solve, transition yields and populations."""
    s, cfg = workloads.config1()
    p = s.pack()
    sysd = ctx.upload(p)
    prm = abi.run_params(method=method, observables=obs)
    y = sysd.time_infinity(prm)
    yo = ref.time_infinity(p, prm)
    print("config1 method", method, "obs", obs, "dev", y[0], "oracle", yo)
    scale = max(1.0, np.abs(yo).max())
    assert np.abs(y[0] - yo).max() <= TI_TOL * scale


@pytest.mark.gpu
def test_device_unequal_sinks_fields(ctx):
    """Unequal sinks (doubling route), three field vectors in one call, against the oracle."""
    s = unequal_sink_system(n_nuc=3)
    p = s.pack()
    sysd = ctx.upload(p)
    fields = [(0.0, 0.0, 5e-5), (3e-5, 0.0, 4e-5), (0.0, 1e-3, 2e-3)]
    prm = abi.run_params()
    y = sysd.time_infinity(prm, fields=fields)
    for i, f in enumerate(fields):
        yo = ref.time_infinity(p, prm, field=f)
        print("unequal field", f, "dev", y[i], "oracle", yo)
        assert np.abs(y[i] - yo).max() <= TI_TOL
        assert abs(y[i].sum() - 1.0) < 1e-9


@pytest.mark.gpu
@pytest.mark.parametrize("method", [abi.MSG_METHOD_AUTO, abi.MSG_METHOD_EXPONENTIAL])
def test_device_given_columns(ctx, method):
    """sampling=given: rho0 = C C^dagger from caller columns (trace samples)."""
    s, cfg = workloads.config1()
    p = s.pack()
    sysd = ctx.upload(p)
    M = 6
    cols = ref.coherent_samples(p, s.state_index("S"), 7, M) / math.sqrt(M)
    prm = abi.run_params(method=method, sampling=abi.MSG_SAMPLING_GIVEN)
    y = sysd.time_infinity(prm, columns=cols)
    yo = ref.time_infinity(p, prm, columns=cols)
    print("given columns method", method, "dev", y[0], "oracle", yo)
    assert np.abs(y[0] - yo).max() <= TI_TOL


@pytest.mark.gpu
@pytest.mark.parametrize("n", [7, 8])
def test_device_routes_agree_with_bartels_stewart(ctx, n):
    """D = 512 and 1024 (7 and 8 nuclei, equal sinks): the eigen route and the
    doubling route agree, and both equal scipy's Bartels-Stewart solution on the
    oracle's H, K and rho0."""
    import scipy.linalg as sl
    idx = next(i for i in range(1000) if workloads.config4_pair(i)[1]["n"] == n)
    s, cfg = workloads.config4_pair(idx)
    p = s.pack()
    sysd = ctx.upload(p)
    f = cfg["fields"][40]
    ye = sysd.time_infinity(abi.run_params(method=abi.MSG_METHOD_AUTO), fields=[f])[0]
    yd = sysd.time_infinity(abi.run_params(method=abi.MSG_METHOD_EXPONENTIAL), fields=[f])[0]
    H = ref.hamiltonian(p, f)
    K = ref.reaction(p)
    B = ref.direct_factors(p)
    X = sl.solve_continuous_lyapunov(-1j * H - K, -(B @ B.conj().T))
    yo = np.array([p.desc.transitions[t].rate * np.trace(ref.projector(p, p.desc.transitions[t].source_state) @ X).real
                   for t in ref.packed_transitions(p)])
    print("D", sysd.dim, "eigen", ye, "doubling", yd, "scipy", yo)
    assert np.abs(ye - yo).max() <= TI_TOL
    assert np.abs(yd - yo).max() <= TI_TOL


@pytest.mark.gpu
def test_device_d4096_completeness_and_singlet(ctx):
    """Config 2 (D = 4096, eigen route), two orientations: the four equal-rate sink
    yields sum to 1, and the singlet yield equals the device symmetric-decay yield at
    rate k (rho0 proportional to P_S, so the dispersive term vanishes)."""
    s, cfg = workloads.config2(n_orientations=1024)
    p = s.pack()
    sysd = ctx.upload(p)
    fields = cfg["fields"][:2]
    y = sysd.time_infinity(abi.run_params(), fields=fields)
    ysd = sysd.symmetric_decay(fields=fields, rate=1e-3)
    print("D=4096 timeinf", y, "symmetric decay S", ysd[:, s.state_index("S")])
    assert np.abs(y.sum(axis=1) - 1.0).max() < 1e-9
    assert np.abs(y[:, 0] - ysd[:, s.state_index("S")]).max() < 1e-9


@pytest.mark.gpu
def test_device_reference_fixture_singular_system(ctx):
    """tests_HSGeneral.cpp:1410-1431 on the device (doubling route): the triplets
    never decay but are never populated, so they add nothing; the singlet-sink yield
    is exactly 1 and the singlet population integrates to 1/0.02 = 50."""
    p = powder_yield_system().pack()
    sysd = ctx.upload(p)
    y = sysd.time_infinity(abi.run_params())
    pop = sysd.time_infinity(abi.run_params(observables=abi.MSG_OBS_STATE_POPULATIONS))
    print("powder fixture yield", y, "population", pop)
    assert abs(y[0, 0] - 1.0) < 1e-10
    assert abs(pop[0, 0] - 50.0) < 1e-8


@pytest.mark.gpu
def test_device_rejections(ctx):
    """A populated part that never decays (T+ initial state, singlet sink, no
    mixing: doubling route), no sinks at all (eigen route) and relaxation operators."""
    from molspin_amd import backend
    s = powder_yield_system()
    s.add_state("Tp", ["E1", "E2"], [1, 0, 0, 0])
    s.set_initial(["Tp"])
    sysd = ctx.upload(s.pack())
    with pytest.raises(backend.BackendError) as e:
        sysd.time_infinity(abi.run_params())
    assert e.value.status == abi.MSG_NUMERIC_BREAKDOWN and "never decays" in str(e.value)
    s = unequal_sink_system(rates=(0.0, 0.0, 0.0, 0.0))
    sysd0 = ctx.upload(s.pack())
    with pytest.raises(backend.BackendError) as e:
        sysd0.time_infinity(abi.run_params(observables=abi.MSG_OBS_STATE_POPULATIONS))
    assert e.value.status == abi.MSG_NUMERIC_BREAKDOWN
    s = unequal_sink_system()
    s.add_st_dephasing("E1", "E2", 1e-3)
    sysd1 = ctx.upload(s.pack())
    with pytest.raises(backend.BackendError) as e:
        sysd1.time_infinity(abi.run_params())
    assert e.value.status == abi.MSG_UNSUPPORTED
