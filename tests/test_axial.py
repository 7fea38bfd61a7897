"""Axial (isotropic among others) electron-nucleus couplings on the matrix-free path:
when every such coupling of a system is axial about z, the kernels apply each flip block
as its two real entries (SpinSpace_hamiltonians.cpp:438-535 builds the same operator from
the full tensors).  Isotropic radical pairs, some with an axial tensor diag(a, a, c) and a
nuclear Zeeman term in an off-axis field, on every kernel that takes the short form:
Lanczos in every launch shape (also beside couplings between nuclei), the LDS-mirrored
and plain streaming kernels at D = 16384, Arnoldi (unequal rates), RK4, the mixed-spin
kernel and the uneven launch, against the FP64 oracle.  |dPhi| <= 1e-5 on yields, 1e-4 max-relative on traces."""
import math

import numpy as np
import pytest

from molspin_amd import abi, backend, workloads
from oracle import binding as ref

pytestmark = pytest.mark.gpu

YIELD_TOL = 1e-5
TRACE_TOL = 1e-4
FIELDS = [workloads.fibonacci_point(64, i, 0.05) for i in (11, 37)]
AXIAL = [[0.5, 0.0, 0.0], [0.0, 0.5, 0.0], [0.0, 0.0, 1.1]]


@pytest.fixture(scope="module")
def ctx():
    c = backend.Context(0)
    yield c
    c.close()


def iso_pair(na, nb, seed, rates=None, nuc_b=None, zeeman=True):
    """Isotropic hyperfine in [0.05, 1.5] mT, the second nucleus of A axial; a nuclear
    Zeeman term on A0 in an off-axis 1 mT field; swept electron Zeeman field."""
    rng = workloads.SplitMix64(71, seed)
    nuc_b = nuc_b or ["1/2"] * nb
    hf = [math.exp(rng.uniform(math.log(0.05), math.log(1.5))) for _ in range(na + len(nuc_b))]
    if na >= 2:
        hf[1] = AXIAL
    s = workloads._radical_pair("iso", ["1/2"] * na, nuc_b, hf, 1e-3, swept=True)
    if zeeman:
        s.add_zeeman(["A0"], (0.3, -0.2, 1.0), prefactor=0.001, swept=False)
    if rates is not None:
        for t in s.transitions:
            t["rate"] = rates[s.states[t["source"]]["name"]]
    return s


def trace_err(dev, ora):
    return float(np.abs(dev - ora).max() / np.abs(ora).max())


def check(ctx, s, prm, columns=None, fields=FIELDS, trace=True):
    p = s.pack()
    sysd = ctx.upload(p)
    y, tr = sysd.propagate(prm, fields=fields, columns=columns, trace=trace)
    yo, tro = ref.run_hsgeneral(p, prm, fields=fields, columns=columns, threads=8, trace=trace)
    dy = float(np.abs(y - yo).max())
    dt = max(trace_err(tr[k], tro[k]) for k in range(len(fields))) if trace else 0.0
    print("D = %d: max |dPhi| %.2e, trace max-rel %.2e" % (sysd.dim, dy, dt))
    assert dy <= YIELD_TOL and dt <= TRACE_TOL
    return sysd


@pytest.mark.parametrize("groups", [None, "1", "2", "4"])
def test_lanczos_d1024(ctx, groups, monkeypatch):
    """D = 1024 (5 + 3 nuclei), all 256 direct columns, 0-40 ns: the default shape and
    one, two and four groups per thread."""
    if groups is not None:
        monkeypatch.setenv("MSG_LANCZOS_G", groups)
    prm = abi.run_params(total_time=40.0, time_step=2.0, method=abi.MSG_METHOD_KRYLOV)
    check(ctx, iso_pair(5, 3, 1), prm)


def test_lanczos_d4096_trace_samples(ctx):
    """D = 4096 (6 + 4 nuclei) on 24 spin-coherent trace samples, 0-30 ns."""
    s = iso_pair(6, 4, 2)
    p = s.pack()
    cols = ref.coherent_samples(p, s.state_index("S"), 13, 24) / math.sqrt(24)
    prm = abi.run_params(total_time=30.0, time_step=2.0, method=abi.MSG_METHOD_KRYLOV,
                         sampling=abi.MSG_SAMPLING_GIVEN)
    check(ctx, s, prm, columns=cols)


@pytest.mark.parametrize("shape", [None, "-1", "0"])
def test_streaming_d16384(ctx, shape, monkeypatch):
    """D = 16384 (7 + 5 nuclei): the default shape, the LDS-mirrored and the plain
    streaming kernels on 6 direct columns and 2 fields, 0-40 ns."""
    if shape is not None:
        monkeypatch.setenv("MSG_LANCZOS_G", shape)
    s = iso_pair(7, 5, 3)
    p = s.pack()
    B = ref.direct_factors(p)
    cols = B[:, [1, 500, 1777, 2222, 3333, 4095]] * math.sqrt(B.shape[1] / 6.0)
    prm = abi.run_params(total_time=40.0, time_step=2.0, method=abi.MSG_METHOD_KRYLOV,
                         sampling=abi.MSG_SAMPLING_GIVEN)
    sysd = check(ctx, s, prm, columns=cols, trace=False)
    assert sysd.dim == 16384


@pytest.mark.parametrize("groups", ["1", "2"])
def test_with_nuclear_couplings(ctx, groups, monkeypatch):
    """The short form beside couplings between nuclei (isotropic and anisotropic, within
    a radical and across the two): D = 1024, one and two groups per thread."""
    monkeypatch.setenv("MSG_LANCZOS_G", groups)
    s = workloads.add_nuclear_couplings(iso_pair(5, 3, 9), 5, 3, 9)
    prm = abi.run_params(total_time=40.0, time_step=2.0, method=abi.MSG_METHOD_KRYLOV)
    check(ctx, s, prm)


def test_arnoldi_unequal_rates(ctx):
    """Unequal recombination (singlet sink faster): matrix-free Arnoldi at D = 1024."""
    s = iso_pair(4, 4, 4, rates={"S": 3e-3, "T0": 1e-3, "Tp": 1e-3, "Tm": 1e-3})
    prm = abi.run_params(total_time=40.0, time_step=2.0, method=abi.MSG_METHOD_KRYLOV)
    check(ctx, s, prm)


def test_rk4_matrix_free(ctx):
    """propagationmethod=rk4 above D = 512: the matrix-free RK4 stages at D = 1024."""
    prm = abi.run_params(total_time=20.0, time_step=0.5, method=abi.MSG_METHOD_RK4)
    check(ctx, iso_pair(4, 4, 5), prm)


def test_mixed_spin_kernel(ctx):
    """An isotropic spin-1 nucleus on B beside 6 + 1 spin-1/2 nuclei (k_lanczos_mix,
    D = 1536)."""
    s = iso_pair(6, 1, 6, nuc_b=["1/2", "1"])
    prm = abi.run_params(total_time=30.0, time_step=2.0, method=abi.MSG_METHOD_KRYLOV)
    sysd = check(ctx, s, prm)
    assert sysd.dim == 1536


def test_uneven_launch(ctx):
    """The uneven launch with isotropic pairs at D = 1024 and 4096 beside an
    anisotropic one at D = 2048: each equals its alone run, the D = 1024 pair the
    oracle."""
    prm = abi.run_params(total_time=20.0, time_step=2.0, method=abi.MSG_METHOD_KRYLOV)
    s1, s2 = iso_pair(5, 3, 7), iso_pair(6, 4, 8)
    rng = workloads.SplitMix64(73, 0)
    s3 = workloads._radical_pair("an", ["1/2"] * 5, ["1/2"] * 4, [workloads._aniso_hf(rng) for _ in range(9)],
                                 1e-3, swept=True)
    ds = [ctx.upload(s.pack()) for s in (s1, s2, s3)]
    out = ctx.propagate_uneven([dict(system=d, params=prm, fields=FIELDS) for d in ds])
    for d, o in zip(ds, out):
        alone, _ = d.propagate(prm, fields=FIELDS)
        assert np.abs(o[0] - alone).max() <= 1e-12
    yo, _ = ref.run_hsgeneral(s1.pack(), prm, fields=FIELDS, threads=8)
    assert np.abs(out[0][0] - yo).max() <= YIELD_TOL
