"""Couplings between nuclei on the matrix-free path (SpinSpace_hamiltonians.cpp:481-535:
a DoubleSpin interaction between any two spins, here two spin-1/2 nuclei): the operator
adds, per coupling, row 2 b1 + b2 of its 4x4 block on the group and on its three
partners across the two nuclei's index bits.  Radical pairs with an anisotropic coupling
inside radical A, an isotropic one between the most significant nuclei of A and B and an
anisotropic one across the radicals (`workloads.add_nuclear_couplings`), on every kernel
that holds the whole Krylov vector in LDS: Lanczos (all launch shapes), Arnoldi (unequal
rates), RK4, the mixed-spin kernel and the uneven launch, against the FP64 oracle.
|dPhi| <= 1e-5 on yields, 1e-4 max-relative on traces."""
import math

import numpy as np
import pytest

from molspin_amd import abi, backend, workloads
from oracle import binding as ref

pytestmark = pytest.mark.gpu

YIELD_TOL = 1e-5
TRACE_TOL = 1e-4
FIELDS = [workloads.fibonacci_point(64, i, 0.05) for i in (7, 40)]


@pytest.fixture(scope="module")
def ctx():
    c = backend.Context(0)
    yield c
    c.close()


def trace_err(dev, ora):
    return float(np.abs(dev - ora).max() / np.abs(ora).max())


def coupled_pair(na, nb, seed, rates=None, nuc_b=None):
    rng = workloads.SplitMix64(61, seed)
    nuc_b = nuc_b or ["1/2"] * nb
    hf = [workloads._aniso_hf(rng) for _ in range(na + len(nuc_b))]
    s = workloads._radical_pair("nn", ["1/2"] * na, nuc_b, hf, 1e-3, swept=True)
    if rates is not None:
        for t in s.transitions:
            t["rate"] = rates[s.states[t["source"]]["name"]]
    return workloads.add_nuclear_couplings(s, na, nb, seed)


def check(ctx, s, prm, columns=None, fields=FIELDS):
    p = s.pack()
    sysd = ctx.upload(p)
    y, tr = sysd.propagate(prm, fields=fields, columns=columns, trace=True)
    yo, tro = ref.run_hsgeneral(p, prm, fields=fields, columns=columns, threads=8, trace=True)
    dy = float(np.abs(y - yo).max())
    dt = max(trace_err(tr[k], tro[k]) for k in range(len(fields)))
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
    check(ctx, coupled_pair(5, 3, 1), prm)


def test_lanczos_d4096_trace_samples(ctx):
    """D = 4096 (6 + 4 nuclei; the coupling across A and B involves the group-index
    bit of the two-groups-per-thread shape) on 24 spin-coherent trace samples."""
    s = coupled_pair(6, 4, 2)
    p = s.pack()
    cols = ref.coherent_samples(p, s.state_index("S"), 11, 24) / math.sqrt(24)
    prm = abi.run_params(total_time=30.0, time_step=2.0, method=abi.MSG_METHOD_KRYLOV,
                         sampling=abi.MSG_SAMPLING_GIVEN)
    check(ctx, s, prm, columns=cols)


def test_arnoldi_unequal_rates(ctx):
    """Unequal recombination (singlet sink faster): matrix-free Arnoldi at D = 1024."""
    s = coupled_pair(4, 4, 3, rates={"S": 3e-3, "T0": 1e-3, "Tp": 1e-3, "Tm": 1e-3})
    prm = abi.run_params(total_time=40.0, time_step=2.0, method=abi.MSG_METHOD_KRYLOV)
    check(ctx, s, prm)


def test_rk4_matrix_free(ctx):
    """propagationmethod=rk4 above D = 512: the matrix-free RK4 stages at D = 1024."""
    prm = abi.run_params(total_time=20.0, time_step=0.5, method=abi.MSG_METHOD_RK4)
    check(ctx, coupled_pair(4, 4, 4), prm)


def test_mixed_spin_kernel(ctx):
    """A spin-1 nucleus on B beside 6 + 1 spin-1/2 nuclei (k_lanczos_mix, D = 1536) with
    the couplings between the spin-1/2 nuclei."""
    s = coupled_pair(6, 1, 5, nuc_b=["1/2", "1"])
    prm = abi.run_params(total_time=30.0, time_step=2.0, method=abi.MSG_METHOD_KRYLOV)
    sysd = check(ctx, s, prm)
    assert sysd.dim == 1536


def test_uneven_launch_and_limits(ctx):
    """The uneven launch takes a coupled pair at D = 1024 beside an uncoupled one; at
    D = 16384 the couplings are refused with a reason (the streaming kernels keep one
    Krylov vector's slab on chip)."""
    prm = abi.run_params(total_time=20.0, time_step=2.0, method=abi.MSG_METHOD_KRYLOV)
    s1 = coupled_pair(5, 3, 6)
    p1 = s1.pack()
    s2, _ = workloads.config4_pair(next(i for i in range(1000) if workloads.config4_pair(i)[1]["n"] == 9))
    d1, d2 = ctx.upload(p1), ctx.upload(s2.pack())
    out = ctx.propagate_uneven([dict(system=d1, params=prm, fields=FIELDS), dict(system=d2, params=prm, fields=FIELDS)])
    yo, _ = ref.run_hsgeneral(p1, prm, fields=FIELDS, threads=8)
    assert np.abs(out[0][0] - yo).max() <= YIELD_TOL
    alone, _ = d2.propagate(prm, fields=FIELDS)
    assert np.abs(out[1][0] - alone).max() <= 1e-12
    big = ctx.upload(coupled_pair(7, 5, 7).pack())
    assert big.dim == 16384
    with pytest.raises(backend.BackendError) as exc:
        big.propagate(prm, fields=FIELDS[:1])
    assert exc.value.status == abi.MSG_UNSUPPORTED and "nucleus-nucleus" in str(exc.value)
