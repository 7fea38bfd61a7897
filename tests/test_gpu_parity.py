"""Device (HIP, through the C ABI) against the FP64 oracle on the same seeded
inputs.  Tolerances follow the specification of the path: singlet yields to
|dPhi| <= 1e-5 and P(t) traces to 1e-4 max-relative (against the trace maximum)."""
import math

import numpy as np
import pytest

from molspin_amd import abi
from molspin_amd import workloads
from molspin_amd.system import SpinSystem
from oracle import binding as ref

pytestmark = pytest.mark.gpu

YIELD_TOL = 1e-5
TRACE_TOL = 1e-4


@pytest.fixture(scope="module")
def ctx():
    from molspin_amd import backend
    c = backend.Context(0)
    yield c
    c.close()


def load_fixture(name):
    import json
    import os
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", name + ".json")) as f:
        return json.load(f)


def trace_err(dev, ora):
    return np.abs(dev - ora).max() / np.abs(ora).max()


@pytest.mark.parametrize("method", [abi.MSG_METHOD_AUTO, abi.MSG_METHOD_EXPONENTIAL])
def test_config1_yields_and_trace(ctx, method):
    """Config 1: D = 48 (spin-1 nucleus), 0-5000 ns at dt = 1 ns, dense path."""
    s, cfg = workloads.config1()
    p = s.pack()
    sysd = ctx.upload(p)
    prm = abi.run_params(total_time=cfg["total_time"], time_step=cfg["time_step"], method=method)
    y, _ = sysd.propagate(prm)
    yo, _ = ref.run_hsgeneral(p, abi.run_params(total_time=cfg["total_time"], time_step=cfg["time_step"],
                                                method=abi.MSG_METHOD_KRYLOV), threads=8)
    print("cfg1 yields dev", y[0], "oracle", yo[0])
    assert np.abs(y - yo).max() <= YIELD_TOL
    prm_t = abi.run_params(calculation=abi.MSG_CALC_TIMEEVOLUTION, observables=abi.MSG_OBS_STATE_POPULATIONS,
                           total_time=cfg["total_time"], time_step=cfg["time_step"], method=method)
    _, tr = sysd.propagate(prm_t, trace=True)
    _, tro = ref.run_hsgeneral(p, prm_t, trace=True, threads=8)
    assert tr.shape == tro.shape
    assert trace_err(tr[0, :, 0], tro[0, :, 0]) <= TRACE_TOL


def test_golden_system_dense(ctx):
    """The reference's golden system (tests_TaskStaticHSSymmetricDecay.cpp:110-174,
    first row: field along z, k = 1e-3 /ns) through the HSGeneral loop with S, T0,
    T+, T- sinks over 0-2e4 ns (dt = 1 ns, 2e4 steps) and the finite-time
    correction: the yields reach the infinite-time golden values 0.327212 and
    0.174110, and match the oracle on the same grid."""
    from tests.test_oracle import golden_system
    s = golden_system()
    r = 1 / math.sqrt(2)
    s.add_state("T0", ["electron1", "electron2"], [0, r, r, 0])
    s.add_state("Tm", ["electron1", "electron2"], [0, 0, 0, 1])
    for name in ("state1", "state2", "T0", "Tm"):
        s.add_transition("k" + name, name, 1e-3)
    p = s.pack()
    sysd = ctx.upload(p)
    prm = abi.run_params(total_time=2e4, time_step=1.0, yield_corrections=True, method=abi.MSG_METHOD_EXPONENTIAL)
    y, _ = sysd.propagate(prm)
    yo, _ = ref.run_hsgeneral(p, prm)
    print("golden dev", y[0], "oracle", yo[0])
    assert np.abs(y - yo).max() <= YIELD_TOL
    assert abs(yo[0, 0] - 0.327212) < 1e-6 and abs(yo[0, 1] - 0.174110) < 1e-6
    assert abs(y[0, 0] - 0.327212) < 5e-6 and abs(y[0, 1] - 0.174110) < 5e-6


def test_config2_orientations_short(ctx):
    """Config 2 (D = 4096, anisotropic hyperfine) on 2 Fibonacci orientations,
    all 1024 direct columns, 0-20 ns: matrix-free Lanczos path."""
    s, cfg = workloads.config2(n_orientations=1024)
    p = s.pack()
    sysd = ctx.upload(p)
    fields = [cfg["fields"][0], cfg["fields"][517]]
    prm = abi.run_params(total_time=20.0, time_step=2.0, method=abi.MSG_METHOD_KRYLOV)
    y, _ = sysd.propagate(prm, fields=fields)
    yo, _ = ref.run_hsgeneral(p, prm, fields=fields, threads=8)
    print("cfg2 short dev", y, "oracle", yo)
    assert np.abs(y - yo).max() <= YIELD_TOL


def test_config2_long_given_columns(ctx):
    """Config 2 over the full 0-3000 ns grid (1500 steps) on 16 direct columns:
    FP32 drift of the matrix-free path stays within the yield tolerance."""
    s, cfg = workloads.config2(n_orientations=1024)
    p = s.pack()
    sysd = ctx.upload(p)
    B = ref.direct_factors(p)
    cols = B[:, ::64] * math.sqrt(B.shape[1] / 16.0)  # 16 columns, rho renormalised to trace 1
    field = cfg["fields"][300]
    prm = abi.run_params(total_time=3000.0, time_step=2.0, sampling=abi.MSG_SAMPLING_GIVEN)
    y, _ = sysd.propagate(prm, fields=[field], columns=cols)
    yo, _ = ref.run_hsgeneral(p, prm, fields=[field], columns=cols, threads=8)
    print("cfg2 long dev", y, "oracle", yo, "diff", y - yo)
    assert np.abs(y - yo).max() <= YIELD_TOL


def st_system(n_a, n_b, rate, seed):
    rng = workloads.SplitMix64(77, seed)
    hf = [math.exp(rng.uniform(math.log(0.05), math.log(1.5))) for _ in range(n_a + n_b)]
    s = workloads._radical_pair("st", ["1/2"] * n_a, ["1/2"] * n_b, hf, 1e-3)
    s.add_st_dephasing("E1", "E2", rate)
    return s


@pytest.mark.parametrize("na,nb", [(2, 2), (5, 3)])
def test_stochastic_trajectories_st_dephasing(ctx, na, nb):
    """Config-3-style trajectories (coherent nuclear states, singlet electrons,
    ST dephasing) with the counter-based flips shared by device and oracle:
    per-trajectory parity, so yields agree deterministically."""
    s = st_system(na, nb, 0.02, na * 10 + nb)
    p = s.pack()
    sysd = ctx.upload(p)
    M = 64
    cols = ref.coherent_samples(p, s.state_index("S"), 4242, M) / math.sqrt(M)
    prm = abi.run_params(total_time=400.0, time_step=1.0, sampling=abi.MSG_SAMPLING_GIVEN, seed=99,
                         method=abi.MSG_METHOD_KRYLOV)
    y, _ = sysd.propagate(prm, columns=cols)
    yo, _ = ref.run_hsgeneral(p, prm, columns=cols, threads=8)
    print("st dev", y, "oracle", yo)
    assert np.abs(y - yo).max() <= YIELD_TOL


def test_singlet_only_recombination_dense(ctx):
    """Unequal Haberkorn rates (singlet and T0 sinks only): the dense path
    exponentiates -iH - K exactly."""
    s = SpinSystem("asym")
    s.add_spin("E1", "1/2", electron=True, tensor=2.0023)
    s.add_spin("E2", "1/2", electron=True, tensor=2.0023)
    for i, a in enumerate([0.9, 0.4, 1.3]):
        s.add_spin("N%d" % i, "1/2", tensor=1.0)
        s.add_hyperfine(["E1" if i < 2 else "E2"], ["N%d" % i], tensor=a, prefactor=0.001)
    s.add_zeeman(["E1", "E2"], (0.0, 0.02, 0.05), prefactor=0.001)
    s.add_radical_pair_states("E1", "E2")
    s.add_transition("kS", "S", 3e-3)
    s.add_transition("kT0", "T0", 1e-3)
    s.set_initial(["S"])
    p = s.pack()
    sysd = ctx.upload(p)
    prm = abi.run_params(total_time=800.0, time_step=1.0, method=abi.MSG_METHOD_EXPONENTIAL)
    y, _ = sysd.propagate(prm)
    yo, _ = ref.run_hsgeneral(p, prm)
    assert np.abs(y - yo).max() <= YIELD_TOL


def test_step_entry_matches_oracle(ctx):
    """msg_step (HSPropagator::StepStochastic) on one interval."""
    s = st_system(4, 3, 0.05, 5)
    p = s.pack()
    sysd = ctx.upload(p)
    M = 8
    cols = ref.coherent_samples(p, s.state_index("S"), 11, M) / math.sqrt(M)
    prm = abi.run_params(sampling=abi.MSG_SAMPLING_GIVEN, method=abi.MSG_METHOD_KRYLOV, seed=5)
    new = sysd.step(prm, cols, 2.0, step_index=0)
    # oracle: one interval of the same loop, observed through state populations
    prm_t = abi.run_params(calculation=abi.MSG_CALC_TIMEEVOLUTION, observables=abi.MSG_OBS_STATE_POPULATIONS,
                           total_time=2.0, time_step=2.0, sampling=abi.MSG_SAMPLING_GIVEN, seed=5)
    _, tro = ref.run_hsgeneral(p, prm_t, columns=cols, trace=True)
    pops = [np.real(np.trace(new.conj().T @ ref.projector(p, st) @ new)) for st in range(4)]
    assert np.abs(np.array(pops) - tro[0, 1]).max() < 1e-6


def force_shape(monkeypatch, shape):
    """Launch shape override (MSG_LANCZOS_G); "-3" is the split-column shape with
    512-thread workgroups (four groups per thread) and "-3/1024" the one with
    1024-thread workgroups (two groups per thread, MSG_SPLIT_T)."""
    if shape is not None:
        g, _, t = shape.partition("/")
        monkeypatch.setenv("MSG_LANCZOS_G", g)
        if t:
            monkeypatch.setenv("MSG_SPLIT_T", t)


@pytest.mark.parametrize("shape", [None, "-1", "0", "-3", "-3/1024"])
def test_large_dimension_streaming_path(ctx, shape, monkeypatch):
    """D = 16384 (7 + 5 anisotropic spin-1/2 nuclei): the large-D matrix-free
    variants on 6 direct columns and 2 orientations over 0-60 ns: the default shape,
    the LDS-mirrored form, the plain streaming form and the column split over two
    workgroups (all tuning candidates)."""
    force_shape(monkeypatch, shape)
    rng = workloads.SplitMix64(41, 3)
    hf = [workloads._aniso_hf(rng) for _ in range(12)]
    s = workloads._radical_pair("big", ["1/2"] * 7, ["1/2"] * 5, hf, 1e-3, swept=True)
    p = s.pack()
    sysd = ctx.upload(p)
    assert sysd.dim == 16384
    B = ref.direct_factors(p)
    cols = B[:, [0, 77, 1023, 2048, 3001, 4095]] * math.sqrt(B.shape[1] / 6.0)
    fields = [workloads.fibonacci_point(64, i, 0.05) for i in (5, 40)]
    prm = abi.run_params(total_time=60.0, time_step=2.0, sampling=abi.MSG_SAMPLING_GIVEN)
    y, _ = sysd.propagate(prm, fields=fields, columns=cols)
    yo, _ = ref.run_hsgeneral(p, prm, fields=fields, columns=cols, threads=8)
    print("D=16384 dev", y, "oracle", yo)
    assert np.abs(y - yo).max() <= YIELD_TOL


@pytest.mark.parametrize("shape", [None, "-2", "-3", "-3/1024"])
def test_config3_trajectories(ctx, shape, monkeypatch):
    """Config 3 (D = 65536, 14 isotropic nuclei, ST dephasing 1e-4 /ns): 4 spin-
    coherent trajectories drawn by the backend's sampler, 0-30 ns, against the
    oracle on the same columns and the same counter-based flips; the default shape,
    the tiled streaming form and the column split over eight workgroups."""
    force_shape(monkeypatch, shape)
    s, cfg = workloads.config3(n_traj=4)
    p = s.pack()
    sysd = ctx.upload(p)
    M = 4
    cols = sysd.trace_samples(s.state_index("S"), M, cfg["sample_seed"]) / math.sqrt(M)
    prm = abi.run_params(total_time=30.0, time_step=1.0, sampling=abi.MSG_SAMPLING_GIVEN, seed=2026)
    y, _ = sysd.propagate(prm, columns=cols)
    yo, _ = ref.run_hsgeneral(p, prm, columns=cols, threads=8)
    print("cfg3 dev", y, "oracle", yo)
    assert np.abs(y - yo).max() <= YIELD_TOL


@pytest.mark.parametrize("shape", ["-3", "-3/1024"])
def test_split_column_d32768(ctx, shape, monkeypatch):
    """D = 32768 (13 anisotropic spin-1/2 nuclei) with each column split over four
    workgroups: 4 direct columns, 2 orientations, 0-20 ns, against the oracle."""
    rng = workloads.SplitMix64(43, 1)
    hf = [workloads._aniso_hf(rng) for _ in range(13)]
    s = workloads._radical_pair("split", ["1/2"] * 7, ["1/2"] * 6, hf, 1e-3, swept=True)
    p = s.pack()
    sysd = ctx.upload(p)
    assert sysd.dim == 32768
    B = ref.direct_factors(p)
    cols = B[:, [3, 2050, 5000, 8191]] * math.sqrt(B.shape[1] / 4.0)
    fields = [workloads.fibonacci_point(64, i, 0.05) for i in (9, 33)]
    prm = abi.run_params(total_time=20.0, time_step=2.0, sampling=abi.MSG_SAMPLING_GIVEN)
    force_shape(monkeypatch, shape)
    y, _ = sysd.propagate(prm, fields=fields, columns=cols)
    yo, _ = ref.run_hsgeneral(p, prm, fields=fields, columns=cols, threads=8)
    print("D=32768 split dev", y, "oracle", yo)
    assert np.abs(y - yo).max() <= YIELD_TOL


def test_legacy_direct_yields(ctx):
    """statichs-direct-yields (endpoint excluded, Lanczos with basis reuse in the
    reference): device task mode against the oracle's legacy loop."""
    s, _ = workloads.config1()
    p = s.pack()
    sysd = ctx.upload(p)
    T, dt = 600.0, 1.0
    prm = abi.run_params(total_time=T, time_step=dt, task=abi.MSG_TASK_LEGACY_DIRECT_YIELDS, yield_corrections=True)
    y, _ = sysd.propagate(prm)
    yo = ref.legacy_direct_yields(p, T, dt, correction=True)
    print("legacy direct dev", y[0], "oracle", yo)
    assert np.abs(y[0] - yo).max() <= YIELD_TOL


def test_legacy_stoch_yields_singlet_only_sink(ctx):
    """statichs-stoch-yields with a single singlet sink: the legacy rule treats
    equal rates as symmetric and applies e^{-kt} analytically."""
    s = SpinSystem("leg")
    s.add_spin("E1", "1/2", electron=True, tensor=2.0023)
    s.add_spin("E2", "1/2", electron=True, tensor=2.0023)
    for i, a in enumerate([0.9, 0.4, 1.3, 0.7]):
        s.add_spin("N%d" % i, "1/2", tensor=1.0)
        s.add_hyperfine(["E1" if i < 2 else "E2"], ["N%d" % i], tensor=a, prefactor=0.001)
    s.add_zeeman(["E1", "E2"], (0.0, 0.0, 0.05), prefactor=0.001)
    s.add_radical_pair_states("E1", "E2")
    s.add_transition("kS", "S", 2e-3)
    s.set_initial(["S"])
    p = s.pack()
    sysd = ctx.upload(p)
    M = 32
    cols = sysd.trace_samples(s.state_index("S"), M, 31337)
    T, dt = 500.0, 1.0
    prm = abi.run_params(total_time=T, time_step=dt, sampling=abi.MSG_SAMPLING_GIVEN,
                         task=abi.MSG_TASK_LEGACY_STOCH_YIELDS)
    y, _ = sysd.propagate(prm, columns=cols / math.sqrt(M))
    yo = ref.legacy_stoch_yields(p, T, dt, cols)
    print("legacy stoch dev", y[0], "oracle", yo)
    assert np.abs(y[0] - yo).max() <= YIELD_TOL


@pytest.mark.parametrize("index", [12, 0, 20])
def test_dense_sweep_matches_oracle(ctx, index):
    """Dense MFMA path (three-product complex GEMM) on uneven-batch pairs of D = 256,
    128 and 512 (16, 8 and 32 row tiles: eight waves from 12 row tiles up, four
    below), 4 fields over 0-40 ns, against the oracle."""
    s, cfg = workloads.config4_pair(index)
    p = s.pack()
    sysd = ctx.upload(p)
    assert sysd.dim == 4 << cfg["n"] and sysd.dim <= 512
    fields = cfg["fields"][::16]
    prm = abi.run_params(total_time=40.0, time_step=2.0, method=abi.MSG_METHOD_EXPONENTIAL)
    y, _ = sysd.propagate(prm, fields=fields)
    yo, _ = ref.run_hsgeneral(p, abi.run_params(total_time=40.0, time_step=2.0, method=abi.MSG_METHOD_KRYLOV),
                              fields=fields, threads=8)
    print("D", sysd.dim, "max |dPhi|", np.abs(y - yo).max())
    assert np.abs(y - yo).max() <= YIELD_TOL


def test_server_alone_equals_mixed(ctx):
    """Continuous batching: requests of different radical pairs (dense and
    matrix-free paths) admitted while others run give the same yields as each
    request propagated alone, and match the oracle."""
    from molspin_amd import backend
    reqs = []
    for idx in (4, 3, 20):  # D = 1024 and 2048 (matrix-free), 512 (dense)
        s, cfg = workloads.config4_pair(idx)
        p = s.pack()
        reqs.append((s, p, ctx.upload(p), cfg["fields"][::32]))
    prm = abi.run_params(total_time=40.0, time_step=2.0)
    alone = [sd.propagate(prm, fields=f)[0] for _, _, sd, f in reqs]
    srv = backend.Server(ctx, max_slots=2, hbm_budget_bytes=8 << 30)
    ids = [srv.submit(sd, prm, f) for _, _, sd, f in reqs[:2]]
    srv.pump(5)  # first two resident and part-way
    ids.append(srv.submit(reqs[2][2], prm, reqs[2][3]))
    active = 3
    while active:
        active = srv.pump(5)
    mixed = [srv.collect(i) for i in ids]
    srv.close()
    for (s, p, sd, f), a, m in zip(reqs, alone, mixed):
        print("D", sd.dim, "alone", a[0], "mixed", m[0])
        assert np.abs(a - m).max() <= 1e-12
        yo, _ = ref.run_hsgeneral(p, prm, fields=f, threads=8)
        assert np.abs(m - yo).max() <= YIELD_TOL


def test_server_whole_mode_field_chunks(ctx):
    """Whole-mode serving (pump(0)): a D = 4096 request of 40 fields runs as
    persistent launches of a few fields each (so it cannot hold the device for
    long), beside a D = 512 request admitted after it; both equal their
    msg_propagate results."""
    import time
    from molspin_amd import backend
    s4, cfg4 = workloads.config4_pair(1)
    s5, cfg5 = workloads.config4_pair(20)
    sd4, sd5 = ctx.upload(s4.pack()), ctx.upload(s5.pack())
    assert sd4.dim == 4096 and sd5.dim == 512
    prm = abi.run_params(total_time=40.0, time_step=2.0)
    f4, f5 = cfg4["fields"][:40], cfg5["fields"][::8]
    a4, a5 = sd4.propagate(prm, fields=f4)[0], sd5.propagate(prm, fields=f5)[0]
    srv = backend.Server(ctx, max_slots=2, hbm_budget_bytes=8 << 30)
    i4 = srv.submit(sd4, prm, f4)
    srv.pump(0)
    i5 = srv.submit(sd5, prm, f5)
    while srv.pump(0):
        time.sleep(0.001)
    m4, m5 = srv.collect(i4), srv.collect(i5)
    srv.close()
    assert m4.shape == a4.shape and m5.shape == a5.shape
    assert np.abs(m4 - a4).max() <= 1e-12 and np.abs(m5 - a5).max() <= 1e-12


def test_symmetric_decay_golden_values(ctx):
    """Device eigenbasis yields (TaskStaticHSSymmetricDecay) reproduce the
    reference's golden values at k = 1e-4 (tests_TaskStaticHSSymmetricDecay.cpp:64-69)
    and the oracle to 1e-10; rate 0 takes the default 1e-3 (no transitions)."""
    from tests.test_oracle import golden_system
    p = golden_system().pack()
    sysd = ctx.upload(p)
    y = sysd.symmetric_decay(rate=1e-4)[0]
    print("golden", y)
    assert "%g %g" % (y[0], y[1]) == "0.325757 0.174256"
    assert np.abs(y - ref.symmetric_decay(p, 1e-4)).max() <= 1e-10
    y0 = sysd.symmetric_decay(rate=0.0)[0]
    assert np.abs(y0 - ref.symmetric_decay(p, 1e-3)).max() <= 1e-10


def test_symmetric_decay_golden_sweep(ctx):
    """The 13-row rotated-field sweep (tests_TaskStaticHSSymmetricDecay.cpp:110-174)
    as one batch of fields."""
    from tests.test_oracle import GOLDEN_SWEEP, golden_system
    p = golden_system().pack()
    sysd = ctx.upload(p)
    fields = []
    for step in range(len(GOLDEN_SWEEP)):
        th = math.radians(15.0 * step)
        fields.append((5e-5 * math.sin(th), 0.0, 5e-5 * math.cos(th)))
    y = sysd.symmetric_decay(fields=fields, rate=1e-3)
    for step, (gs, gt) in enumerate(GOLDEN_SWEEP):
        assert "%g %g" % (y[step, 0], y[step, 1]) == "%g %g" % (gs, gt), step


def test_symmetric_decay_dense_pair_and_completeness(ctx):
    """A 7-nucleus anisotropic pair (D = 512) against the oracle's eigensolver, and
    at D = 4096 (config 2 system) the completeness law: the yields of S, T0, T+, T-
    (projectors summing to the identity) add up to Tr rho0 = 1."""
    fx = load_fixture("symdecay_d512")
    s, cfg = workloads.config4_pair(fx["pair"])
    sysd = ctx.upload(s.pack())
    y = sysd.symmetric_decay(fields=fx["fields"], rate=fx["rate"])
    print("D=512 dev", y, "oracle", fx["yields"])
    assert np.abs(y - np.array(fx["yields"])).max() <= 1e-9
    s2, cfg2 = workloads.config2(n_orientations=1024)
    sysd2 = ctx.upload(s2.pack())
    y2 = sysd2.symmetric_decay(fields=[cfg2["fields"][0]], rate=1e-3)[0]
    print("D=4096 yields", y2, "sum", y2.sum())
    assert abs(y2.sum() - 1.0) <= 1e-9
    assert (y2 > 0).all()


def test_config2_against_fixture(ctx):
    """Config 2 (D = 4096): 16 Fibonacci orientations, 0-50 ns, all 1024 direct
    columns, against the committed FP64 fixture."""
    fx = load_fixture("cfg2_subset")
    s, cfg = workloads.config2(n_orientations=1024)
    sysd = ctx.upload(s.pack())
    prm = abi.run_params(total_time=fx["total_time"], time_step=fx["time_step"], method=abi.MSG_METHOD_KRYLOV,
                         krylov_size=fx["krylov_size"])
    y, _ = sysd.propagate(prm, fields=[cfg["fields"][i] for i in fx["orientations"]])
    err = np.abs(y - np.array(fx["yields"])).max()
    print("cfg2 fixture max |dPhi|", err)
    assert err <= YIELD_TOL


@pytest.mark.parametrize("shape", [None, "-3"])
def test_config3_against_fixture(ctx, shape, monkeypatch):
    """Config 3 (D = 65536): 32 spin-coherent trajectories drawn by the backend's
    sampler (same columns as the oracle's up to a phase per column), ST dephasing
    with flip seed 2026, 0-200 ns, against the committed FP64 fixture; the default
    shape and the column split over eight workgroups."""
    force_shape(monkeypatch, shape)
    fx = load_fixture("cfg3_subset")
    M = fx["trajectories"]
    s, cfg = workloads.config3(n_traj=M)
    sysd = ctx.upload(s.pack())
    cols = sysd.trace_samples(s.state_index("S"), M, fx["sample_seed"]) / math.sqrt(M)
    prm = abi.run_params(total_time=fx["total_time"], time_step=fx["time_step"], method=abi.MSG_METHOD_KRYLOV,
                         sampling=abi.MSG_SAMPLING_GIVEN, seed=fx["flip_seed"], krylov_size=fx["krylov_size"])
    y, _ = sysd.propagate(prm, columns=cols)
    err = np.abs(y[0] - np.array(fx["yields"])).max()
    print("cfg3 fixture max |dPhi|", err, y[0])
    assert err <= YIELD_TOL


def test_config1_trace_against_fixture(ctx):
    """Config 1 P(t) traces of S, T0, T+, T- at 5001 points against the fixture."""
    fx = load_fixture("cfg1")
    s, cfg = workloads.config1()
    sysd = ctx.upload(s.pack())
    prm = abi.run_params(calculation=abi.MSG_CALC_TIMEEVOLUTION, observables=abi.MSG_OBS_STATE_POPULATIONS,
                         total_time=fx["total_time"], time_step=fx["time_step"])
    _, tr = sysd.propagate(prm, trace=True)
    want = np.array(fx["trace"]).T
    for c in range(4):
        assert trace_err(tr[0, :, c], want[:, c]) <= TRACE_TOL, c


def test_autotune_keeps_results(ctx):
    """msg_autotune picks a launch shape per Hilbert dimension; yields computed
    with the tuned shape agree with the default shape's to FP32 rounding, for the
    LDS path (D = 4096) and the large-D path (D = 65536: tiled, plain streaming or
    split columns)."""
    from molspin_amd import backend
    s, cfg = workloads.config2(n_orientations=1024)
    fields = cfg["fields"][:4]
    prm = abi.run_params(total_time=20.0, time_step=2.0, method=abi.MSG_METHOD_KRYLOV)
    c2 = backend.Context(0)
    try:
        sysd = c2.upload(s.pack())
        y0, _ = sysd.propagate(prm, fields=fields)
        shape = sysd.autotune(prm, fields=fields, sample_steps=3)
        print("D=4096 tuned", shape.groups_per_thread, shape.threads, shape.slots, shape.kernel_ms)
        # candidates <2, 512>, <1, 1024> and <4, 256>: G x T covers the D / 4 = 1024 groups
        assert shape.path == 1 and shape.groups_per_thread * shape.threads == 1024 and shape.slots >= 1
        y1, _ = sysd.propagate(prm, fields=fields)
        assert np.abs(y1 - y0).max() <= 1e-7
        s3, cfg3 = workloads.config3(n_traj=8)
        sys3 = c2.upload(s3.pack())
        cols = sys3.trace_samples(s3.state_index("S"), 8, cfg3["sample_seed"]) / math.sqrt(8)
        p3 = abi.run_params(total_time=5.0, time_step=1.0, method=abi.MSG_METHOD_KRYLOV,
                            sampling=abi.MSG_SAMPLING_GIVEN, seed=2026)
        z0, _ = sys3.propagate(p3, columns=cols)
        sh3 = sys3.autotune(p3, columns=cols, sample_steps=2)
        print("D=65536 tuned", sh3.groups_per_thread, sh3.threads, sh3.slots, sh3.kernel_ms)
        assert sh3.groups_per_thread in (-3, -2, 0)
        z1, _ = sys3.propagate(p3, columns=cols)
        assert np.abs(z1 - z0).max() <= 1e-7
    finally:
        c2.close()


@pytest.mark.parametrize("index,shapes", [(4, ["1", "2"]), (3, ["1", "2"]), (1, ["2", "1", "4"])])
def test_launch_shapes_match_oracle(ctx, index, shapes, monkeypatch):
    """Every shape of the LDS Lanczos kernel that msg_autotune may pick (G groups
    per thread, D / 4G threads; the operator forms the nucleus blocks of a
    thread's shared index bits once for its G groups): D = 1024 and 2048 with
    G = 1, 2 and D = 4096 with G = 1, 2, 4, on 3 fields of an uneven-batch pair,
    each against the oracle."""
    s, cfg = workloads.config4_pair(index)
    p = s.pack()
    sysd = ctx.upload(p)
    assert sysd.dim == 4 << cfg["n"] and sysd.dim in (1024, 2048, 4096)
    fields = cfg["fields"][::24]
    prm = abi.run_params(total_time=20.0, time_step=2.0, method=abi.MSG_METHOD_KRYLOV)
    yo, _ = ref.run_hsgeneral(p, prm, fields=fields, threads=8)
    for g in shapes:
        monkeypatch.setenv("MSG_LANCZOS_G", g)
        y, _ = sysd.propagate(prm, fields=fields)
        print("D", sysd.dim, "G", g, "max |dPhi|", np.abs(y - yo).max())
        assert np.abs(y - yo).max() <= YIELD_TOL, g


@pytest.mark.parametrize("singlet_only", [False, True])
def test_rk4_matches_oracle(ctx, singlet_only):
    """propagationmethod=rk4 (HSPropagator.cpp:80-90): config 1 (D = 48) with the
    four equal sinks, and with a singlet-only sink (K not a multiple of the
    identity), yields and P(t) traces against the oracle's RK4 on the same grid."""
    s, cfg = workloads.config1()
    if singlet_only:
        s.transitions = [t for t in s.transitions if t["name"] == "kS"] or s.transitions[:1]
    p = s.pack()
    sysd = ctx.upload(p)
    prm = abi.run_params(total_time=400.0, time_step=1.0, method=abi.MSG_METHOD_RK4)
    y, _ = sysd.propagate(prm)
    yo, _ = ref.run_hsgeneral(p, prm, threads=8)
    print("rk4 yields dev", y[0], "oracle", yo[0])
    assert np.abs(y - yo).max() <= YIELD_TOL
    prm_t = abi.run_params(calculation=abi.MSG_CALC_TIMEEVOLUTION, observables=abi.MSG_OBS_STATE_POPULATIONS,
                           total_time=400.0, time_step=1.0, method=abi.MSG_METHOD_RK4)
    _, tr = sysd.propagate(prm_t, trace=True)
    _, tro = ref.run_hsgeneral(p, prm_t, trace=True, threads=8)
    for c in range(tr.shape[2]):
        assert trace_err(tr[0, :, c], tro[0, :, c]) <= TRACE_TOL, c


def test_uncoupled_config1_matches_oracle(ctx):
    """propagationmethod=uncoupled (H = H_A (x) 1 + 1 (x) H_B, correlation functions
    from the radicals' eigenpairs): config 1 (radical B holds a spin-1 nucleus)
    yields over 0-5000 ns and the four State traces against the oracle's exact
    exponential propagation."""
    s, cfg = workloads.config1()
    p = s.pack()
    sysd = ctx.upload(p)
    prm = abi.run_params(total_time=cfg["total_time"], time_step=cfg["time_step"], method=abi.MSG_METHOD_UNCOUPLED)
    y, _ = sysd.propagate(prm)
    yo, _ = ref.run_hsgeneral(p, abi.run_params(total_time=cfg["total_time"], time_step=cfg["time_step"],
                                                method=abi.MSG_METHOD_KRYLOV), threads=8)
    print("uncoupled cfg1", y[0], "oracle", yo[0])
    assert np.abs(y - yo).max() <= YIELD_TOL
    fx = load_fixture("cfg1")
    prm_t = abi.run_params(calculation=abi.MSG_CALC_TIMEEVOLUTION, observables=abi.MSG_OBS_STATE_POPULATIONS,
                           total_time=fx["total_time"], time_step=fx["time_step"], method=abi.MSG_METHOD_UNCOUPLED)
    _, tr = sysd.propagate(prm_t, trace=True)
    want = np.array(fx["trace"]).T
    for c in range(4):
        assert trace_err(tr[0, :, c], want[:, c]) <= TRACE_TOL, c


def test_uncoupled_bare_radical(ctx):
    """Radical B without nuclei (dimension 2: its 3 eigenvalue pairs leave one of the
    correlation kernel's 4 pair ranges empty) beside 3 nuclei on A, at 3 field
    orientations, with traces: the uncoupled method against the oracle's propagation."""
    rng = workloads.SplitMix64(79, 0)
    s = workloads._radical_pair("bare", ["1/2"] * 3, [], [workloads._aniso_hf(rng) for _ in range(3)], 2e-3,
                                swept=True)
    p = s.pack()
    sysd = ctx.upload(p)
    fields = [workloads.fibonacci_point(16, i, 0.05) for i in (1, 6, 11)]
    prm = abi.run_params(total_time=600.0, time_step=2.0, method=abi.MSG_METHOD_UNCOUPLED)
    y, tr = sysd.propagate(prm, fields=fields, trace=True)
    yo, tro = ref.run_hsgeneral(p, abi.run_params(total_time=600.0, time_step=2.0, method=abi.MSG_METHOD_KRYLOV),
                                fields=fields, trace=True)
    print("uncoupled bare", y[:, 0], "oracle", yo[:, 0])
    assert np.abs(y - yo).max() <= YIELD_TOL
    for k in range(len(fields)):
        assert trace_err(tr[k], tro[k]) <= TRACE_TOL, k


def test_uncoupled_config2_against_fixture(ctx):
    """Config 2 (D = 4096 as radicals of 128 and 32): the 16 fixture orientations."""
    fx = load_fixture("cfg2_subset")
    s, cfg = workloads.config2(n_orientations=1024)
    sysd = ctx.upload(s.pack())
    prm = abi.run_params(total_time=fx["total_time"], time_step=fx["time_step"], method=abi.MSG_METHOD_UNCOUPLED)
    y, _ = sysd.propagate(prm, fields=[cfg["fields"][i] for i in fx["orientations"]])
    err = np.abs(y - np.array(fx["yields"])).max()
    print("uncoupled cfg2 max |dPhi|", err)
    assert err <= YIELD_TOL


def test_uncoupled_rejects_coupled_electrons(ctx):
    from molspin_amd import backend
    s = SpinSystem("ex")
    s.add_spin("E1", "1/2", electron=True, tensor=2.0023)
    s.add_spin("E2", "1/2", electron=True, tensor=2.0023)
    s.add_spin("N", "1/2", tensor=1.0)
    s.add_hyperfine(["E1"], ["N"], tensor=1.0, prefactor=0.001)
    s.add_exchange(["E1"], ["E2"], 0.2, prefactor=0.001)
    s.add_zeeman(["E1", "E2"], (0.0, 0.0, 0.05), prefactor=0.001)
    s.add_radical_pair_states("E1", "E2")
    for st in ("S", "T0", "Tp", "Tm"):
        s.add_transition("k" + st, st, 1e-3)
    s.set_initial(["S"])
    sysd = ctx.upload(s.pack())
    with pytest.raises(backend.BackendError) as e:
        sysd.propagate(abi.run_params(total_time=10.0, time_step=1.0, method=abi.MSG_METHOD_UNCOUPLED))
    assert e.value.status == abi.MSG_UNSUPPORTED


def test_server_uncoupled_requests(ctx):
    """Server requests with propagationmethod=uncoupled, mixed with Krylov-path
    requests: each result equals the same request run alone, and the uncoupled
    yields match the oracle."""
    from molspin_amd import backend
    reqs = []
    for idx in (4, 20):
        s, cfg = workloads.config4_pair(idx)
        p = s.pack()
        reqs.append((p, ctx.upload(p), cfg["fields"][::32]))
    pu = abi.run_params(total_time=40.0, time_step=2.0, method=abi.MSG_METHOD_UNCOUPLED)
    pk = abi.run_params(total_time=40.0, time_step=2.0)
    alone = [sd.propagate(pu, fields=f)[0] for _, sd, f in reqs]
    srv = backend.Server(ctx, max_slots=2, hbm_budget_bytes=8 << 30)
    ids = [srv.submit(reqs[0][1], pk, reqs[0][2]), srv.submit(reqs[0][1], pu, reqs[0][2]),
           srv.submit(reqs[1][1], pu, reqs[1][2])]
    active = 3
    while active:
        active = srv.pump(0)
    got = [srv.collect(i) for i in ids]
    srv.close()
    assert np.abs(got[1] - alone[0]).max() <= 1e-12
    assert np.abs(got[2] - alone[1]).max() <= 1e-12
    assert np.abs(got[0] - got[1]).max() <= YIELD_TOL  # Krylov path and exact uncoupled agree
    for (p, sd, f), a in zip(reqs, alone):
        yo, _ = ref.run_hsgeneral(p, pk, fields=f, threads=8)
        assert np.abs(a - yo).max() <= YIELD_TOL


@pytest.mark.parametrize("shape", [None, "-3", "-3/1024"])
def test_config3_full_horizon(ctx, shape, monkeypatch):
    """Config 3 (D = 65536) over its whole 0-2000 ns grid (2000 steps) on 8
    spin-coherent trajectories with the ST flips of seed 2026, against the committed
    FP64 fixture: the FP32 drift of the matrix-free path over the full horizon stays
    inside the yield tolerance, and the singlet/triplet traces (every 10th sample)
    inside the trace tolerance.  Default (tuner-free) shape and split columns."""
    force_shape(monkeypatch, shape)
    fx = load_fixture("cfg3_full")
    M = fx["trajectories"]
    s, cfg = workloads.config3(n_traj=M)
    sysd = ctx.upload(s.pack())
    cols = sysd.trace_samples(s.state_index("S"), M, fx["sample_seed"]) / math.sqrt(M)
    prm = abi.run_params(total_time=fx["total_time"], time_step=fx["time_step"], method=abi.MSG_METHOD_KRYLOV,
                         sampling=abi.MSG_SAMPLING_GIVEN, seed=fx["flip_seed"], krylov_size=fx["krylov_size"])
    y, tr = sysd.propagate(prm, columns=cols, trace=True)
    err = np.abs(y[0] - np.array(fx["yields"])).max()
    terr = trace_err(tr[0, ::10], np.array(fx["trace_every_10"]))
    print("cfg3 full horizon max |dPhi|", err, "trace max-rel", terr)
    assert err <= YIELD_TOL
    assert terr <= TRACE_TOL


def test_config2_bench_fixture_traces(ctx):
    """Config 2 (D = 4096) on the bench's 100 ns horizon: 8 orientations, all 1024
    direct columns, through the static k_lanczos kernel at the default shape; yields
    and per-sample traces of every transition observable against the fixture."""
    fx = load_fixture("cfg2_bench")
    s, cfg = workloads.config2(n_orientations=1024)
    sysd = ctx.upload(s.pack())
    prm = abi.run_params(total_time=fx["total_time"], time_step=fx["time_step"], method=abi.MSG_METHOD_KRYLOV,
                         krylov_size=fx["krylov_size"])
    y, tr = sysd.propagate(prm, fields=[cfg["fields"][i] for i in fx["orientations"]], trace=True)
    ftr = np.array(fx["trace"])
    err = np.abs(y - np.array(fx["yields"])).max()
    terr = max(trace_err(tr[k], ftr[k]) for k in range(len(ftr)))
    print("cfg2 bench fixture max |dPhi|", err, "trace max-rel", terr)
    assert err <= YIELD_TOL
    assert terr <= TRACE_TOL


def test_legacy_direct_yields_lanczos_d1024(ctx):
    """statichs-direct-yields at D = 1024 (config-4 pair, 8 anisotropic nuclei): the
    task mode on the matrix-free Lanczos kernel (a fresh Krylov basis every step,
    the reference's rule at krylovtol = 0) against the oracle's legacy loop."""
    idx = next(i for i in range(1000) if workloads.config4_pair(i)[1]["n"] == 8)
    s, cfg = workloads.config4_pair(idx)
    p = s.pack()
    sysd = ctx.upload(p)
    assert sysd.dim == 1024
    T, dt = 200.0, 2.0
    field = cfg["fields"][30]
    prm = abi.run_params(total_time=T, time_step=dt, task=abi.MSG_TASK_LEGACY_DIRECT_YIELDS, yield_corrections=True,
                         method=abi.MSG_METHOD_KRYLOV)
    y, _ = sysd.propagate(prm, fields=[field])
    yo = ref.legacy_direct_yields(p, T, dt, correction=True, field=field)
    print("legacy direct D=1024 dev", y[0], "oracle", yo)
    assert np.abs(y[0] - yo).max() <= YIELD_TOL


def test_legacy_direct_yields_basis_reuse(ctx):
    """statichs-direct-yields with krylovtol at D = 1024: each column keeps its Krylov
    basis while h_{m+1,m} |c_{m-1}| <= krylovtol and advances its coefficients by
    exp(-i dt T) (TaskStaticHSDirectYields.cpp:569-652), against the oracle's loop with
    the same tolerance.  At 1e-4 the reuse moves the yields by about 4e-6 from the
    fresh-basis result (on both sides), so the device follows the oracle's restarts."""
    idx = next(i for i in range(1000) if workloads.config4_pair(i)[1]["n"] == 8)
    s, cfg = workloads.config4_pair(idx)
    p = s.pack()
    sysd = ctx.upload(p)
    T, dt = 200.0, 2.0
    field = cfg["fields"][30]
    out = {}
    for tol in (0.0, 1e-6, 1e-4):
        prm = abi.run_params(total_time=T, time_step=dt, task=abi.MSG_TASK_LEGACY_DIRECT_YIELDS,
                             yield_corrections=True, method=abi.MSG_METHOD_KRYLOV, krylov_tol=tol)
        y, _ = sysd.propagate(prm, fields=[field])
        yo = ref.legacy_direct_yields(p, T, dt, tol=tol, correction=True, field=field)
        print("krylovtol", tol, "dev", y[0], "oracle", yo, "diff", np.abs(y[0] - yo).max())
        assert np.abs(y[0] - yo).max() <= YIELD_TOL
        out[tol] = (y[0], yo)
    # the reuse changes the result on both sides by the same amount
    dev_shift, ora_shift = out[1e-4][0] - out[0.0][0], out[1e-4][1] - out[0.0][1]
    assert np.abs(ora_shift).max() > 1e-6
    assert np.abs(dev_shift - ora_shift).max() <= 0.5 * np.abs(ora_shift).max()


@pytest.mark.parametrize("na,nb", [(2, 2), (5, 3)])
def test_stochastic_trajectories_two_channels(ctx, na, nb):
    """Two relaxationdephasing channels on the electron pair (rates 0.015 and 0.01 /ns):
    the superposed Poisson count (SpinSpace_relaxation.cpp:1237-1252), drawn by
    inversion of the same Philox word on the device and in the oracle, so trajectories
    agree one by one; dense path (D = 64) and matrix-free Lanczos (D = 1024)."""
    s = st_system(na, nb, 0.015, na * 10 + nb + 1)
    s.add_st_dephasing("E1", "E2", 0.01)
    p = s.pack()
    sysd = ctx.upload(p)
    M = 64
    cols = ref.coherent_samples(p, s.state_index("S"), 4243, M) / math.sqrt(M)
    prm = abi.run_params(total_time=400.0, time_step=1.0, sampling=abi.MSG_SAMPLING_GIVEN, seed=101,
                         method=abi.MSG_METHOD_KRYLOV)
    y, _ = sysd.propagate(prm, columns=cols)
    yo, _ = ref.run_hsgeneral(p, prm, columns=cols, threads=8)
    # the same columns without relaxation: the channels change the yields
    s0 = st_system(na, nb, 0.0, na * 10 + nb + 1)
    y0, _ = ref.run_hsgeneral(s0.pack(), prm, columns=cols, threads=8)
    print("two channels dev", y, "oracle", yo, "no relaxation", y0)
    assert np.abs(y - yo).max() <= YIELD_TOL
    assert np.abs(yo - y0).max() > 1e-3


@pytest.mark.parametrize("n,cols", [(8, None), (10, 8), (12, 4)])
def test_rk4_matrix_free(ctx, n, cols):
    """propagationmethod=rk4 above D = 512 (HSPropagator.cpp:80-90): the matrix-free
    RK4 stages of -iH - K with K inside (singlet sink faster than the triplet ones),
    D = 1024 on all direct columns, D = 4096 and 16384 on trace-sample columns,
    against the oracle's RK4 loop: yields and traces."""
    idx = next(i for i in range(1000) if workloads.config4_pair(i)[1]["n"] == n)
    s, cfg = workloads.config4_pair(idx)
    for t in s.transitions:
        t["rate"] = 4e-3 if s.states[t["source"]]["name"] == "S" else 1e-3
    p = s.pack()
    sysd = ctx.upload(p)
    kw = {}
    sampling = abi.MSG_SAMPLING_DIRECT
    if cols is not None:
        kw["columns"] = ref.coherent_samples(p, s.state_index("S"), 21, cols) / math.sqrt(cols)
        sampling = abi.MSG_SAMPLING_GIVEN
    prm = abi.run_params(calculation=abi.MSG_CALC_TIMEEVOLUTION, method=abi.MSG_METHOD_RK4, total_time=40.0,
                         time_step=0.25, sampling=sampling)
    fields = [cfg["fields"][40]]
    y, tr = sysd.propagate(prm, fields=fields, trace=True, **kw)
    yo, tro = ref.run_hsgeneral(p, prm, fields=fields, trace=True, threads=8, **kw)
    print("rk4 D", sysd.dim, "dev", y[0], "oracle", yo[0], "trace err", trace_err(tr, tro))
    assert np.abs(y - yo).max() <= YIELD_TOL
    assert trace_err(tr, tro) <= TRACE_TOL


def mixed_system(n_half, mult, seed):
    """Radical pair with n_half spin-1/2 nuclei and one nucleus of multiplicity `mult`
    (spin 1 or 3/2, like FAD's N5, example_2_5_ps.msd) with anisotropic hyperfine."""
    rng = workloads.SplitMix64(91, seed)
    spin = {3: "1", 4: "3/2"}[mult]
    hf = [workloads._aniso_hf(rng) for _ in range(n_half + 1)]
    na = (n_half + 1) // 2
    s = workloads._radical_pair("mix", [spin] + ["1/2"] * (na - 1), ["1/2"] * (n_half + 1 - na), hf, 1e-3,
                                swept=True)
    return s


@pytest.mark.parametrize("n_half,mult,cols", [(7, 3, None), (6, 4, None), (10, 3, 4)])
def test_matrix_free_spin1_nucleus(ctx, n_half, mult, cols):
    """A nucleus of spin 1 or 3/2 beside 6-10 spin-1/2 nuclei on the matrix-free
    Lanczos path (Spin.cpp:330-410: I+ elements sqrt(I(I+1) - m(m-1))): D = 1536 and
    1024 on all direct columns at two field orientations, D = 12288 on trace-sample
    columns, against the oracle."""
    s = mixed_system(n_half, mult, n_half * 10 + mult)
    p = s.pack()
    sysd = ctx.upload(p)
    assert sysd.dim == 4 * mult * 2 ** n_half
    fields = [workloads.fibonacci_point(64, i, 0.05) for i in (5, 40)]
    kw = {}
    sampling = abi.MSG_SAMPLING_DIRECT
    if cols is not None:
        kw["columns"] = ref.coherent_samples(p, s.state_index("S"), 3, cols) / math.sqrt(cols)
        sampling = abi.MSG_SAMPLING_GIVEN
        fields = fields[:1]
    prm = abi.run_params(total_time=40.0, time_step=2.0, method=abi.MSG_METHOD_KRYLOV, sampling=sampling)
    y, _ = sysd.propagate(prm, fields=fields, **kw)
    yo, _ = ref.run_hsgeneral(p, prm, fields=fields, threads=8, **kw)
    print("mixed D", sysd.dim, "dev", y[:, 0], "oracle", yo[:, 0], "max diff", np.abs(y - yo).max())
    assert np.abs(y - yo).max() <= YIELD_TOL
