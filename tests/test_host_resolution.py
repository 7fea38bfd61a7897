"""CPU checks of the backend's host resolution (molspin_amd/csrc/msg_resolve.cpp)
against the oracle: the device-order dense Hamiltonian, the basis permutation,
and the matrix-free operator restated on the host from the same per-item
coefficient blocks the kernel reads."""
import ctypes as C
import hashlib
import math
import os
import subprocess

import numpy as np
import pytest

from molspin_amd import abi, workloads
from molspin_amd.system import SpinSystem
from oracle import binding as ref

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = [os.path.join(ROOT, "tests", "host_shim.cpp"), os.path.join(ROOT, "molspin_amd", "csrc", "msg_resolve.cpp")]
HEADERS = [os.path.join(ROOT, "include", "molspin_gpu.h"), os.path.join(ROOT, "molspin_amd", "csrc", "msg_resolve.hpp"),
           os.path.join(ROOT, "molspin_amd", "csrc", "msg_internal.hpp")]


@pytest.fixture(scope="module")
def shim():
    h = hashlib.sha1(b"".join(open(p, "rb").read() for p in SRC + HEADERS)).hexdigest()[:12]
    out = "/tmp/msg_host_shim_%s.so" % h
    if not os.path.exists(out):
        subprocess.run(["g++", "-O1", "-std=c++17", "-fPIC", "-shared", "-o", out + ".tmp"] + SRC, check=True)
        os.replace(out + ".tmp", out)
    L = C.CDLL(out)
    P = C.POINTER
    L.shim_resolve.argtypes = [P(abi.SystemDesc), P(C.c_int64), P(C.c_int), P(C.c_int), P(C.c_int)]
    L.shim_perm.argtypes = [P(abi.SystemDesc), C.c_void_p]
    L.shim_dense_hamiltonian.argtypes = [P(abi.SystemDesc), C.c_void_p, C.c_void_p]
    L.shim_fast_apply.argtypes = [P(abi.SystemDesc), C.c_void_p, C.c_void_p, C.c_void_p]
    L.shim_trace_samples.argtypes = [P(abi.SystemDesc), C.c_int, C.c_int, C.c_uint64, C.c_int, C.c_void_p]
    L.shim_nuc_flags.argtypes = [P(abi.SystemDesc), C.c_void_p, C.c_void_p]
    return L


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def perm_of(shim, packed):
    perm = np.zeros(packed.dim, dtype=np.int64)
    assert shim.shim_perm(packed.ref(), _p(perm)) == 0
    assert sorted(perm) == list(range(packed.dim))
    return perm


def device_dense(shim, packed, field=None):
    D = packed.dim
    out = np.zeros(2 * D * D)
    f = None if field is None else np.array(field, dtype=np.float64)
    assert shim.shim_dense_hamiltonian(packed.ref(), _p(f), _p(out)) == 0
    return out.view(np.complex128).reshape(D, D)  # row-major


def small_aniso_pair(na, nb, seed=0, exchange=None, nuclear_zeeman=False, nuclear_pairs=False):
    rng = workloads.SplitMix64(31, seed)
    hf = [workloads._aniso_hf(rng) for _ in range(na + nb)]
    s = workloads._radical_pair("t", ["1/2"] * na, ["1/2"] * nb, hf, 1e-3, field=(0.01, -0.02, 0.05), swept=True)
    if exchange is not None:
        s.add_exchange(["E1"], ["E2"], exchange, prefactor=0.001)
    if nuclear_zeeman:
        s.add_zeeman(["A0", "B0"], (0.0, 0.0, 50.0), prefactor=0.001, swept=False)
    if nuclear_pairs:
        workloads.add_nuclear_couplings(s, na, nb, seed)
    return s


@pytest.mark.parametrize("name", ["config1", "golden", "aniso"])
def test_device_order_dense_hamiltonian_matches_oracle(shim, name):
    if name == "config1":
        s, _ = workloads.config1()
    elif name == "golden":
        from tests.test_oracle import golden_system
        s = golden_system()
    else:
        s = small_aniso_pair(3, 2, 1, exchange=0.3)
    p = s.pack()
    Hr = ref.hamiltonian(p, (0.0, 0.01, 0.05))
    Hd = device_dense(shim, p, (0.0, 0.01, 0.05))
    perm = perm_of(shim, p)
    assert np.abs(Hd[np.ix_(perm, perm)] - Hr).max() < 1e-14


@pytest.mark.parametrize("variant", [dict(), dict(exchange=0.4), dict(nuclear_zeeman=True), dict(nuclear_pairs=True)])
def test_matrix_free_operator_matches_oracle(shim, variant):
    """The per-item Kronecker factors reproduce H x of the oracle's CSR
    Hamiltonian (float accuracy) for anisotropic hyperfine on both radicals,
    off-axis fields, exchange, nuclear Zeeman terms and couplings between nuclei
    (isotropic and anisotropic, within a radical and across the two)."""
    s = small_aniso_pair(4, 3, 7, **variant)
    p = s.pack()
    field = (0.03, -0.01, 0.04)
    Hr = ref.hamiltonian(p, field)
    perm = perm_of(shim, p)
    rng = np.random.default_rng(5)
    x = rng.normal(size=p.dim) + 1j * rng.normal(size=p.dim)
    xd = np.zeros(p.dim, dtype=np.complex64)
    xd[perm] = x
    yd = np.zeros(p.dim, dtype=np.complex64)
    f = np.array(field)
    assert shim.shim_fast_apply(p.ref(), _p(f), _p(xd.view(np.float32)), _p(yd.view(np.float32))) == 0
    want = Hr @ x
    got = yd[perm].astype(np.complex128)
    assert np.abs(got - want).max() <= 2e-6 * np.abs(want).max()


NUC_E1, NUC_E2, NUC_A1, NUC_A2 = 1, 2, 8, 16


def test_axial_couplings_take_the_short_form(shim):
    """Isotropic and axial (diag(a, a, c)) hyperfine couplings get the short flip block
    (two real entries), the anisotropic ones keep the full 2x2 block; with the short form
    on, the matrix-free operator still reproduces the oracle's H x, in an off-axis field
    with a nuclear Zeeman term on an isotropic nucleus."""
    rng = workloads.SplitMix64(37, 0)
    axial = [[0.6, 0.0, 0.0], [0.0, 0.6, 0.0], [0.0, 0.0, -0.9]]
    hf = [0.8, workloads._aniso_hf(rng), axial, 1.3, 0.4, workloads._aniso_hf(rng), 0.07]
    s = workloads._radical_pair("ax", ["1/2"] * 4, ["1/2"] * 3, hf, 1e-3, field=(0.01, -0.02, 0.05), swept=True)
    s.add_zeeman(["A0"], (0.0, 0.0, 50.0), prefactor=0.001, swept=False)
    p = s.pack()
    field = np.array([0.03, -0.01, 0.04])
    fl = np.zeros(16, dtype=np.int32)
    assert shim.shim_nuc_flags(p.ref(), _p(field), _p(fl)) == 0
    perm = perm_of(shim, p)
    Hr = ref.hamiltonian(p, tuple(field))
    # the short form only beside the coupling it describes; 5 of the 7 couplings are axial
    assert all((not f & NUC_A1 or f & NUC_E1) and (not f & NUC_A2 or f & NUC_E2) for f in fl)
    assert sum(bool(f & (NUC_A1 | NUC_A2)) for f in fl) == 5
    rng2 = np.random.default_rng(9)
    x = rng2.normal(size=p.dim) + 1j * rng2.normal(size=p.dim)
    xd = np.zeros(p.dim, dtype=np.complex64)
    xd[perm] = x
    yd = np.zeros(p.dim, dtype=np.complex64)
    assert shim.shim_fast_apply(p.ref(), _p(field), _p(xd.view(np.float32)), _p(yd.view(np.float32))) == 0
    want = Hr @ x
    got = yd[perm].astype(np.complex128)
    assert np.abs(got - want).max() <= 2e-6 * np.abs(want).max()


def test_resolution_flags(shim):
    dim, fast, pair, sym = C.c_int64(), C.c_int(), C.c_int(), C.c_int()
    s, _ = workloads.config1()
    assert shim.shim_resolve(s.pack().ref(), C.byref(dim), C.byref(fast), C.byref(pair), C.byref(sym)) == 0
    assert dim.value == 48 and pair.value == 1 and fast.value == 0 and sym.value == 1  # spin-1 nucleus
    s2, _ = workloads.config2(n_orientations=4)
    assert shim.shim_resolve(s2.pack().ref(), C.byref(dim), C.byref(fast), C.byref(pair), C.byref(sym)) == 0
    assert dim.value == 4096 and fast.value == 1 and sym.value == 1


@pytest.mark.parametrize("method", [0, 1])
def test_trace_samples_match_oracle_sampler(shim, method):
    """The backend's host sampler (BuildTraceSamples restated, coherent and SU(Z))
    draws the same columns as the oracle's from the same counter-based stream,
    for spin-1/2 and spin-1 nuclei and a singlet support."""
    s, _ = workloads.config1()
    p = s.pack()
    M = 16
    out = np.zeros(2 * p.dim * M)
    assert shim.shim_trace_samples(p.ref(), s.state_index("S"), method, 777, M, _p(out)) == 0
    got = out.view(np.complex128).reshape((p.dim, M), order="F")
    want = (ref.coherent_samples if method == 0 else ref.suz_samples)(p, s.state_index("S"), 777, M)
    # columns agree up to a global phase of each column
    for j in range(M):
        ph = np.vdot(want[:, j], got[:, j])
        assert abs(abs(ph) - 1.0) < 1e-12
        assert np.abs(got[:, j] - ph * want[:, j]).max() < 1e-12
