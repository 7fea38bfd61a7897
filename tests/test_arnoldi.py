"""Matrix-free Arnoldi on G = -iH - K (unequal recombination rates, D up to 16384)
against the oracle's KrylovExpmGeneral restatement (Arnoldi, modified Gram-Schmidt,
exp(dt H_m) e1), same Krylov order.  Tolerances as for the Lanczos path:
|dPhi| <= 1e-5 on yields, 1e-4 max-relative on traces."""
import math

import numpy as np
import pytest

from molspin_amd import abi, backend
from molspin_amd.system import SpinSystem
from oracle import binding as ref

pytestmark = pytest.mark.gpu

YIELD_TOL = 1e-5
TRACE_TOL = 1e-4


def pair(n_a, n_b, rates, seed=3, rf=None):
    """Radical pair with anisotropic hyperfine on n_a + n_b spin-1/2 nuclei, a swept
    Zeeman field and sinks S, T0, T+, T- at `rates` (unequal: K not proportional to 1)."""
    rng = np.random.default_rng(seed)
    s = SpinSystem("arn")
    s.add_spin("E1", "1/2", electron=True, tensor=2.0023)
    s.add_spin("E2", "1/2", electron=True, tensor=2.0023)
    for i in range(n_a + n_b):
        s.add_spin("N%d" % i, "1/2", tensor=1.0)
        A = rng.normal(size=(3, 3))
        A = 0.4 * (A + A.T) / 2 + (0.2 + rng.random()) * np.eye(3)
        s.add_hyperfine(["E1" if i < n_a else "E2"], ["N%d" % i], tensor=A.tolist(), prefactor=0.001)
    s.add_zeeman(["E1", "E2"], (0.0, 0.0, 5e-5), prefactor=1.0, swept=True)
    if rf is not None:
        s.add_zeeman(["E1", "E2"], rf, time_dependence="linearpolarized", frequency=0.3)
    s.add_radical_pair_states("E1", "E2")
    for name, k in zip(("S", "T0", "Tp", "Tm"), rates):
        s.add_transition("k" + name, name, k)
    s.set_initial(["S"])
    return s


FIELDS = [(0.0, 0.0, 5e-5), (3e-5, 0.0, 4e-5), (0.0, 2e-4, 1e-4)]


@pytest.fixture(scope="module")
def ctx():
    c = backend.Context(0)
    yield c
    c.close()


def trace_err(dev, ora):
    return np.abs(dev - ora).max() / np.abs(ora).max()


@pytest.mark.parametrize("groups", ["1", "2"])
def test_arnoldi_d1024_singlet_only_sink(ctx, groups, monkeypatch):
    """D = 1024: singlet-only recombination (the Haberkorn singlet sink), three fields,
    yields and traces against the oracle, with one and two groups per thread."""
    s = pair(4, 4, (3e-3, 0.0, 0.0, 0.0))
    p = s.pack()
    sysd = ctx.upload(p)
    monkeypatch.setenv("MSG_LANCZOS_G", groups)
    prm = abi.run_params(calculation=abi.MSG_CALC_TIMEEVOLUTION, method=abi.MSG_METHOD_KRYLOV,
                         total_time=200.0, time_step=2.0)
    _, tr = sysd.propagate(prm, fields=FIELDS, trace=True)
    _, tro = ref.run_hsgeneral(p, prm, fields=FIELDS, trace=True, threads=8)
    print("D=1024 arnoldi trace err", trace_err(tr, tro))
    assert trace_err(tr, tro) <= TRACE_TOL
    prm = abi.run_params(method=abi.MSG_METHOD_KRYLOV, total_time=200.0, time_step=2.0)
    y, _ = sysd.propagate(prm, fields=FIELDS)
    yo, _ = ref.run_hsgeneral(p, prm, fields=FIELDS, threads=8)
    print("yields", y[:, 0], yo[:, 0])
    assert np.abs(y - yo).max() <= YIELD_TOL


@pytest.mark.parametrize("groups", ["1", "2"])
def test_arnoldi_d2048_singlet_only_sink(ctx, groups, monkeypatch):
    """D = 2048 (5 + 4 nuclei), singlet-only sink, two fields, 10 intervals, with one
    and two groups per thread."""
    s = pair(5, 4, (2e-3, 0.0, 0.0, 0.0), seed=5)
    p = s.pack()
    sysd = ctx.upload(p)
    assert sysd.dim == 2048
    prm = abi.run_params(method=abi.MSG_METHOD_KRYLOV, total_time=20.0, time_step=2.0)
    monkeypatch.setenv("MSG_LANCZOS_G", groups)
    y, _ = sysd.propagate(prm, fields=FIELDS[1:])
    yo, _ = ref.run_hsgeneral(p, prm, fields=FIELDS[1:], threads=16)
    print("D=2048 arnoldi G", groups, y[:, 0], yo[:, 0])
    assert np.abs(y - yo).max() <= YIELD_TOL


@pytest.mark.parametrize("groups", ["1", "2"])
def test_arnoldi_d4096_unequal_rates(ctx, groups, monkeypatch):
    """D = 4096 (6 + 4 nuclei), k_S != k_T, one orientation, 20 intervals, in both
    launch shapes the tuner may pick (1024 threads with one group each, 512 with two)."""
    s = pair(6, 4, (2e-3, 5e-4, 5e-4, 5e-4))
    p = s.pack()
    sysd = ctx.upload(p)
    prm = abi.run_params(method=abi.MSG_METHOD_KRYLOV, total_time=40.0, time_step=2.0)
    monkeypatch.setenv("MSG_LANCZOS_G", groups)
    y, _ = sysd.propagate(prm, fields=FIELDS[2:])
    yo, _ = ref.run_hsgeneral(p, prm, fields=FIELDS[2:], threads=16)
    print("D=4096 arnoldi G", groups, y[:, 0], yo[:, 0])
    assert np.abs(y - yo).max() <= YIELD_TOL


def test_arnoldi_d8192_given_columns_with_dephasing(ctx):
    """D = 8192 (two groups per thread): trace-sample columns with singlet-triplet
    dephasing (Philox flips shared with the oracle), state populations per sample."""
    s = pair(7, 4, (2e-3, 1e-3, 0.0, 0.0))
    s.add_st_dephasing("E1", "E2", 0.02)
    p = s.pack()
    sysd = ctx.upload(p)
    assert sysd.dim == 8192
    M = 6
    cols = ref.coherent_samples(p, s.state_index("S"), 9, M) / math.sqrt(M)
    prm = abi.run_params(calculation=abi.MSG_CALC_TIMEEVOLUTION, observables=abi.MSG_OBS_STATE_POPULATIONS,
                         method=abi.MSG_METHOD_KRYLOV, sampling=abi.MSG_SAMPLING_GIVEN, total_time=30.0,
                         time_step=2.0, seed=77)
    _, tr = sysd.propagate(prm, columns=cols, trace=True)
    _, tro = ref.run_hsgeneral(p, prm, columns=cols, trace=True, threads=8)
    print("D=8192 arnoldi trace err", trace_err(tr, tro))
    assert trace_err(tr, tro) <= TRACE_TOL


def test_arnoldi_dynamic_field(ctx):
    """Unequal rates with a linearly polarized field (dynamics=dynamic): the Arnoldi
    kernel reads the per-interval electron fields."""
    s = pair(4, 4, (3e-3, 1e-3, 1e-3, 1e-3), rf=(2e-4, 0.0, 1e-4))
    p = s.pack()
    sysd = ctx.upload(p)
    prm = abi.run_params(method=abi.MSG_METHOD_KRYLOV, total_time=100.0, time_step=2.0, dynamics=True)
    y, _ = sysd.propagate(prm, fields=FIELDS[:2])
    yo, _ = ref.run_hsgeneral(p, prm, fields=FIELDS[:2], threads=8)
    print("dynamic arnoldi", y[:, 0], yo[:, 0])
    assert np.abs(y - yo).max() <= YIELD_TOL


def test_arnoldi_d16384_given_columns(ctx):
    """D = 16384 (four groups per thread, the Krylov vector in 128 KiB of LDS):
    singlet-only recombination on trace-sample columns, yields and traces against
    the oracle."""
    s = pair(8, 4, (2e-3, 0.0, 0.0, 0.0))
    p = s.pack()
    sysd = ctx.upload(p)
    assert sysd.dim == 16384
    M = 4
    cols = ref.coherent_samples(p, s.state_index("S"), 13, M) / math.sqrt(M)
    prm = abi.run_params(calculation=abi.MSG_CALC_TIMEEVOLUTION, method=abi.MSG_METHOD_KRYLOV,
                         sampling=abi.MSG_SAMPLING_GIVEN, total_time=30.0, time_step=2.0)
    y, tr = sysd.propagate(prm, fields=FIELDS[1:2], columns=cols, trace=True)
    yo, tro = ref.run_hsgeneral(p, prm, fields=FIELDS[1:2], columns=cols, trace=True, threads=8)
    print("D=16384 arnoldi", y[0], yo[0], "trace err", trace_err(tr, tro))
    assert np.abs(y - yo).max() <= YIELD_TOL
    assert trace_err(tr, tro) <= TRACE_TOL


def test_arnoldi_limit(ctx):
    """Unequal rates above D = 16384 stay refused, naming the limit."""
    s = pair(9, 4, (2e-3, 0.0, 0.0, 0.0))
    sysd = ctx.upload(s.pack())
    with pytest.raises(backend.BackendError) as e:
        sysd.propagate(abi.run_params(method=abi.MSG_METHOD_KRYLOV, total_time=4.0, time_step=2.0),
                       fields=FIELDS[:1])
    assert e.value.status == abi.MSG_UNSUPPORTED and "16384" in str(e.value)
