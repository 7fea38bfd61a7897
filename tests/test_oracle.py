"""Pins the FP64 oracle (oracle/refcpu.cpp) against the reference's own golden
values, known-answer vectors and closed forms, so a plausible mistake (a dropped
term, a wrong sign, a transposed operand) fails one of them."""
import math

import numpy as np
import pytest

from molspin_amd import abi
from molspin_amd.system import SpinSystem
from molspin_amd import workloads
from oracle import binding as ref

GAMMA = 87.9410005


def golden_system():
    """Tests/tests_TaskStaticHSSymmetricDecay.cpp:14-46."""
    s = SpinSystem("System")
    s.add_spin("electron1", "1/2", electron=True, tensor=2.0)
    s.add_spin("electron2", "1/2", electron=True, tensor=2.0)
    s.add_spin("nucleus1", "1/2", tensor=1.0)
    s.add_spin("nucleus2", "1/2", tensor=1.0)
    s.add_hyperfine(["electron1"], ["nucleus1"], tensor=5e-4)
    s.add_hyperfine(["electron2"], ["nucleus2"], tensor=(1e-4, 1e-4, 1e-3))
    s.add_zeeman(["electron1", "electron2"], (0.0, 0.0, 5e-5), swept=True)
    r = 1 / math.sqrt(2)
    s.add_state("state1", ["electron1", "electron2"], [0, r, -r, 0])
    s.add_state("state2", ["electron1", "electron2"], [1, 0, 0, 0])
    s.set_initial(["state1"])
    return s


def test_symmetric_decay_golden_values():
    """Phi_S = 0.325757, Phi_T+ = 0.174256 at k = 1e-4 (tests_TaskStaticHSSymmetricDecay.cpp:64-69)."""
    p = golden_system().pack()
    y = ref.symmetric_decay(p, 1e-4)
    assert "%g %g" % (y[0], y[1]) == "0.325757 0.174256"


GOLDEN_SWEEP = [
    (0.327212, 0.174110), (0.321633, 0.179687), (0.307003, 0.194457), (0.288503, 0.213453),
    (0.271831, 0.230932), (0.260868, 0.242648), (0.257143, 0.246686), (0.260868, 0.242648),
    (0.271831, 0.230932), (0.288503, 0.213453), (0.307003, 0.194457), (0.321633, 0.179687),
    (0.327212, 0.174110)]


def test_symmetric_decay_golden_sweep():
    """Field rotated 15 degrees per step about y, k = 1e-3
    (tests_TaskStaticHSSymmetricDecay.cpp:110-174)."""
    p = golden_system().pack()
    for step, (gs, gt) in enumerate(GOLDEN_SWEEP):
        th = math.radians(15.0 * step)
        field = (5e-5 * math.sin(th), 0.0, 5e-5 * math.cos(th))
        y = ref.symmetric_decay(p, 1e-3, field)
        assert "%g %g" % (y[0], y[1]) == "%g %g" % (gs, gt), step


def test_philox_known_answers():
    """Random123 known-answer vectors for philox4x32-10."""
    assert ref.philox([0, 0, 0, 0], [0, 0]) == [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]
    assert ref.philox([0xffffffff] * 4, [0xffffffff] * 2) == [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]
    assert ref.philox([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0]) == \
        [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]


def test_fibonacci_points_match_reference_vectors():
    """Tests/tests_actions.cpp:266-307 (n = 200; points 50, 100, 198)."""
    want = {50: (0.707203326, 0.4974874372, 0.5023641165),
            100: (0.3292513656, -0.005025125628, 0.9442289375),
            198: (-0.09728003857, -0.9899497487, -0.1026454532)}
    for i, v in want.items():
        got = ref.fibonacci_point(200, i, 1.0)
        assert np.allclose(got, v, atol=1e-4), (i, got)
        # the product's workload generator restates the same float-truncated point
        assert np.allclose(workloads.fibonacci_point(200, i, 1.0), got, rtol=0, atol=2e-7)


def test_logspace():
    """Tests/tests_actions.cpp:309-335: logspace(0, 6, 50)."""
    v = ref.logspace(0.0, 6.0, 50)
    assert v[0] == 1.0 and abs(v[-1] - 1e6) < 1e-6
    assert abs(v[10] - 10 ** (60 / 49)) < 1e-9 * v[10]
    assert np.allclose(workloads.logspace(0.0, 6.0, 50), v, rtol=1e-15)


def test_time_grid_exact_endpoint():
    """tests_HSGeneral.cpp:1753-1769: T = 0.25, dt = 0.2 gives 0, 0.2, 0.25."""
    assert list(ref.time_grid(0.25, 0.2)) == [0.0, 0.2, 0.25]
    t = ref.time_grid(0.5, 0.05)
    assert len(t) == 11 and t[-1] == 0.5
    assert list(ref.time_grid(0.0, 1.0)) == [0.0]


def test_trapezoid():
    """tests_HSGeneral.cpp:1732-1751."""
    assert abs(ref.trapezoid([0.0, 0.5, 1.25], [1.0, 1.0, 1.0]) - 1.25) < 1e-12


def single_sink_system(rate=0.02):
    s = SpinSystem("decay")
    s.add_spin("e", "1/2", electron=True)
    s.add_state("up", ["e"], [1, 0])
    s.add_transition("k", "up", rate)
    s.set_initial(["up"])
    return s


@pytest.mark.parametrize("method", [abi.MSG_METHOD_KRYLOV, abi.MSG_METHOD_EXPONENTIAL])
def test_yield_correction_finite_decay_law(method):
    """tests_HSGeneral.cpp:1085-1112: k e^{-kt} trapezoid / (1 - e^{-kT}) to 1e-10."""
    p = single_sink_system().pack()
    prm = abi.run_params(total_time=0.5, time_step=0.05, method=method, yield_corrections=True)
    y, _ = ref.run_hsgeneral(p, prm)
    rate, total, dt = 0.02, 0.5, 0.05
    trap = sum(0.5 * dt * rate * (math.exp(-rate * (s - 1) * dt) + math.exp(-rate * s * dt)) for s in range(1, 11))
    assert abs(y[0, 0] - trap / (1 - math.exp(-rate * total))) < 1e-10


def breit_rabi_system(a, b):
    s = SpinSystem("br")
    s.add_spin("e", "1/2", electron=True, tensor=2.0023)
    s.add_spin("n", "1/2", tensor=1.0)
    s.add_hyperfine(["e"], ["n"], tensor=a)
    s.add_zeeman(["e"], (0.0, 0.0, b))
    return s


def test_hamiltonian_breit_rabi_and_units():
    """One electron + one spin-1/2 nucleus: H = w Sz + A S.I has eigenvalues
    w/2 + A/4, -w/2 + A/4, -A/4 +- sqrt(w^2 + A^2)/2, with w = 87.94 g B and
    A = 87.94 g_e g_n a (T operators carry the spin tensors)."""
    a, b = 1.3e-3, 0.7e-3
    H = ref.hamiltonian(breit_rabi_system(a, b).pack())
    w = GAMMA * 2.0023 * b
    A = GAMMA * 2.0023 * 1.0 * a
    want = sorted([w / 2 + A / 4, -w / 2 + A / 4, -A / 4 + math.hypot(w, A) / 2, -A / 4 - math.hypot(w, A) / 2])
    got = np.linalg.eigvalsh(H)
    assert np.allclose(got, want, rtol=0, atol=1e-12)
    # 0.001 * 87.9410005 * 2.0023 rad/ns per mT agrees with the free-electron
    # gamma_e = 176.0859644 rad/mT/us quoted in TaskStaticHSStochYields.cpp:71 to 1e-5
    assert abs(0.001 * GAMMA * 2.0023 / 0.1760859644 - 1) < 1e-5


def test_hamiltonian_invariants():
    s, _ = workloads.config1()
    p = s.pack()
    H = ref.hamiltonian(p)
    assert H.shape == (48, 48)
    assert np.abs(H - H.conj().T).max() < 1e-15
    assert abs(np.trace(H)) < 1e-12  # no exchange, traceless spin operators
    # field || z, isotropic couplings: H commutes with total Fz
    mults = [2, 2, 2, 2, 3]
    fz = np.zeros(48)
    for idx in range(48):
        rem, tot = idx, 0.0
        for m in reversed(mults):
            d = rem % m
            rem //= m
            tot += (m - 1) / 2 - d
        fz[idx] = tot
    Fz = np.diag(fz)
    assert np.abs(H @ Fz - Fz @ H).max() < 1e-14


def test_anisotropic_hyperfine_term_orientation():
    """S1 A S2 with a non-symmetric matrix: the Hamiltonian's xz element couples
    S1x with S2z (not S1z S2x), pinning operand order."""
    s = SpinSystem("t")
    s.add_spin("e", "1/2", electron=True, tensor=1.0)
    s.add_spin("n", "1/2", tensor=1.0)
    A = [[0, 0, 1.0], [0, 0, 0], [0, 0, 0]]
    s.add_hyperfine(["e"], ["n"], tensor=A, common_prefactor=False)
    H = ref.hamiltonian(s.pack())
    sx = np.array([[0, 0.5], [0.5, 0]])
    sz = np.array([[0.5, 0], [0, -0.5]])
    assert np.abs(H - np.kron(sx, sz)).max() < 1e-15


def test_exchange_convention():
    """Exchange J(2 S1.S2 + 1/2) (SpinSpace_hamiltonians.cpp:619-628): singlet at
    -J, triplets at +J."""
    s = SpinSystem("x")
    s.add_spin("a", "1/2", electron=True, tensor=1.0)
    s.add_spin("b", "1/2", electron=True, tensor=1.0)
    s.add_exchange(["a"], ["b"], 0.3, common_prefactor=False)
    w = np.linalg.eigvalsh(ref.hamiltonian(s.pack()))
    assert np.allclose(w, [-0.3, 0.3, 0.3, 0.3], atol=1e-14)


def test_direct_factors_reproduce_rho0():
    s, _ = workloads.config1()
    p = s.pack()
    B = ref.direct_factors(p)
    P = ref.projector(p, s.state_index("S"))
    assert B.shape == (48, 12)
    assert np.abs(B @ B.conj().T - P / np.trace(P).real).max() < 1e-15


def closed_form_trace(H, rho0, P, k, times):
    """k-free population trace Tr[P e^{-iHt} rho0 e^{iHt}] e^{-kt} by
    eigen-decomposition (numpy, a textbook routine)."""
    lam, V = np.linalg.eigh(H)
    rp = V.conj().T @ rho0 @ V
    pp = V.conj().T @ P @ V
    M = pp * rp.T  # M[g,h] = P'_{gh} rho'_{hg}
    out = []
    for t in times:
        ph = np.exp(-1j * (lam[None, :] - lam[:, None]) * t)  # e^{-i(l_h - l_g)t}
        out.append((M * ph).sum().real * math.exp(-k * t))
    return np.array(out)


@pytest.mark.parametrize("method", [abi.MSG_METHOD_KRYLOV, abi.MSG_METHOD_EXPONENTIAL])
def test_hsgeneral_matches_eigen_closed_form(method):
    """Finite-time trapezoid yield and P_S(t) of config 1 (shortened to 400 ns)
    against the eigen-decomposition closed form."""
    s, cfg = workloads.config1()
    p = s.pack()
    T, dt, k = 400.0, 1.0, 1e-3
    prm = abi.run_params(total_time=T, time_step=dt, method=method, krylov_size=16)
    y, tr = ref.run_hsgeneral(p, prm, trace=True)
    times = ref.time_grid(T, dt)
    H = ref.hamiltonian(p)
    P = ref.projector(p, s.state_index("S"))
    rho0 = P / np.trace(P).real
    pop = closed_form_trace(H, rho0, P, k, times)
    assert np.abs(tr[0, :, 0] - k * pop).max() < 1e-12
    phi = np.trapezoid(k * pop, times)
    assert abs(y[0, 0] - phi) < 1e-11
    # the four sinks together integrate the trapezoid of k e^{-kt}
    assert abs(y[0].sum() - np.trapezoid(k * np.exp(-k * times), times)) < 1e-12


def test_rk4_converges_to_exponential():
    s, _ = workloads.config1()
    p = s.pack()
    ye, _ = ref.run_hsgeneral(p, abi.run_params(total_time=20.0, time_step=0.5, method=abi.MSG_METHOD_EXPONENTIAL))
    yr, _ = ref.run_hsgeneral(p, abi.run_params(total_time=20.0, time_step=0.5, method=abi.MSG_METHOD_RK4))
    assert np.abs(ye - yr).max() < 1e-7


def test_time_evolution_state_populations():
    """observables=states: Tr[P rho(t)] without the rate, endpoint included."""
    s, _ = workloads.config1()
    p = s.pack()
    prm = abi.run_params(calculation=abi.MSG_CALC_TIMEEVOLUTION, observables=abi.MSG_OBS_STATE_POPULATIONS,
                         total_time=10.0, time_step=0.7)
    _, tr = ref.run_hsgeneral(p, prm, trace=True)
    times = ref.time_grid(10.0, 0.7)
    assert tr.shape == (1, len(times), 4) and times[-1] == 10.0
    # populations of S, T0, T+, T- sum to the surviving norm e^{-kt}
    assert np.allclose(tr[0].sum(axis=1), np.exp(-1e-3 * times), atol=1e-13)
    assert abs(tr[0, 0, 0] - 1.0) < 1e-14


def test_legacy_direct_yields_matches_hsgeneral():
    """Legacy Lanczos loop with basis reuse (endpoint excluded) at T + dt equals
    HSGeneral at T (tests_HSGeneral.cpp:729-759 use the same shift)."""
    s, _ = workloads.config1()
    p = s.pack()
    T, dt = 300.0, 1.0
    yg, _ = ref.run_hsgeneral(p, abi.run_params(total_time=T, time_step=dt))
    yl = ref.legacy_direct_yields(p, T + dt, dt)
    assert np.abs(yg[0] - yl).max() < 1e-10


def electron_pair(rate):
    s = SpinSystem("pair")
    s.add_spin("e1", "1/2", electron=True)
    s.add_spin("e2", "1/2", electron=True)
    s.add_radical_pair_states("e1", "e2")
    s.add_state("ud", ["e1", "e2"], [0, 1, 0, 0])
    s.add_st_dephasing("e1", "e2", rate)
    return s


def test_st_dephasing_statistics():
    """tests_StochasticRelaxation.cpp:75-93: flips with the exact parity
    probability, S and T populations conserved, ST coherence e^{-kt}."""
    k, t, N = 0.7, 1.3, 32768
    s = electron_pair(k)
    p = s.pack()
    cols = np.zeros((4, N), dtype=complex)
    cols[1, :] = 0.6
    prm = abi.run_params(calculation=abi.MSG_CALC_TIMEEVOLUTION, observables=abi.MSG_OBS_STATE_POPULATIONS,
                         sampling=abi.MSG_SAMPLING_GIVEN, total_time=t, time_step=t, seed=519)
    _, tr = ref.run_hsgeneral(p, prm, columns=cols, trace=True)
    pops = tr[0, 1] / N  # S, T0, T+, T-, ud
    ptot = -math.expm1(-k * t) / 2.0
    sigma = 0.36 * math.sqrt(ptot * (1 - ptot) / N)
    assert abs(pops[0] - 0.18) < 1e-12 and abs(pops[1] - 0.18) < 1e-12
    assert abs(0.36 * (1 - ptot) - pops[4]) < 6 * sigma
    # coherence <S|rho|T0> = (P_ud - P_du)/2 = 0.18 e^{-kt}
    assert abs((2 * pops[4] - 0.36) / 2 - 0.18 * math.exp(-k * t)) < 6 * sigma


def test_st_dephasing_channels_superpose():
    """Several relaxationdephasing channels on the electron pair
    (SpinSpace_relaxation.cpp:1237-1252): the superposed Poisson process with the same
    U = P_S - P_T in every channel flips a column with the single-channel probability of
    the summed rate."""
    k1, k2, t, N = 0.4, 0.3, 1.3, 32768
    s = electron_pair(k1)
    s.add_st_dephasing("e1", "e2", k2)
    p = s.pack()
    cols = np.zeros((4, N), dtype=complex)
    cols[1, :] = 0.6
    prm = abi.run_params(calculation=abi.MSG_CALC_TIMEEVOLUTION, observables=abi.MSG_OBS_STATE_POPULATIONS,
                         sampling=abi.MSG_SAMPLING_GIVEN, total_time=t, time_step=t / 2, seed=77)
    _, tr = ref.run_hsgeneral(p, prm, columns=cols, trace=True)
    pops = tr[0, -1] / N
    k = k1 + k2
    ptot = -math.expm1(-k * t) / 2.0
    sigma = 0.36 * math.sqrt(ptot * (1 - ptot) / N)
    assert abs(pops[0] - 0.18) < 1e-12 and abs(pops[1] - 0.18) < 1e-12
    assert abs(0.36 * (1 - ptot) - pops[4]) < 6 * sigma


def test_relaxation_channels_on_different_pairs():
    """Channels on two spin pairs that share a spin, A = (e1, e2) and B = (e1, n):
    U = -SWAP of the pair, so a basis state hops udd <-> dud at rate lambda_A = k_A / 2
    and udd <-> ddu at rate lambda_B = k_B / 2.  The trajectory average of the udd
    population follows that three-state Markov chain, which pins the event count and
    the channel draw (probabilities lambda_j / sum lambda) of the superposed process."""
    from scipy.linalg import expm
    kA, kB, t, N = 0.5, 0.2, 2.0, 32768
    s = SpinSystem("three")
    s.add_spin("e1", "1/2", electron=True)
    s.add_spin("e2", "1/2", electron=True)
    s.add_spin("n", "1/2", tensor=1.0)
    amp = lambda i: [1.0 if j == i else 0.0 for j in range(8)]
    for name, i in (("udd", 3), ("dud", 5), ("ddu", 6)):
        s.add_state(name, ["e1", "e2", "n"], amp(i))
    s.add_st_dephasing("e1", "e2", kA)
    s.add_st_dephasing("e1", "n", kB)
    p = s.pack()
    cols = np.zeros((8, N), dtype=complex)
    cols[3, :] = 1.0
    prm = abi.run_params(calculation=abi.MSG_CALC_TIMEEVOLUTION, observables=abi.MSG_OBS_STATE_POPULATIONS,
                         sampling=abi.MSG_SAMPLING_GIVEN, total_time=t, time_step=0.25, seed=5)
    _, tr = ref.run_hsgeneral(p, prm, columns=cols, trace=True)
    pops = tr[0, -1] / N
    lA, lB = kA / 2, kB / 2
    Q = np.array([[-(lA + lB), lA, lB], [lA, -lA, 0.0], [lB, 0.0, -lB]])
    want = expm(Q * t)[0]
    sigma = math.sqrt(0.25 / N)
    assert np.abs(pops - want).max() < 6 * sigma
    assert abs(pops.sum() - 1.0) < 1e-12


def test_coherent_samples_statistics():
    """Spin-coherent nuclear states projected on the singlet: unit norm, singlet
    electronic part, and E|Omega><Omega| = I/(2I+1) per nucleus."""
    s = SpinSystem("c")
    s.add_spin("e1", "1/2", electron=True)
    s.add_spin("e2", "1/2", electron=True)
    s.add_spin("n1", "1/2", tensor=1.0)
    s.add_spin("n2", "1", tensor=1.0)
    s.add_radical_pair_states("e1", "e2")
    p = s.pack()
    M = 8192
    B = ref.coherent_samples(p, 0, 12345, M)
    assert np.allclose(np.linalg.norm(B, axis=0), 1.0, atol=1e-14)
    P = ref.projector(p, 0)
    assert np.abs(P @ B - B).max() < 1e-14
    rho = B @ B.conj().T / M
    want = P / np.trace(P).real
    assert np.abs(rho - want).max() < 6.0 / math.sqrt(M)


def test_stoch_legacy_matches_hsgeneral_given_columns():
    """Legacy stoch-yields (dense propagator, e^{-kt} applied analytically,
    endpoint excluded) equals HSGeneral on the same columns at T - dt."""
    s, _ = workloads.config1()
    p = s.pack()
    M = 64
    cols = ref.coherent_samples(p, s.state_index("S"), 99, M)
    T, dt = 200.0, 1.0
    yl = ref.legacy_stoch_yields(p, T, dt, cols)
    yg, _ = ref.run_hsgeneral(p, abi.run_params(total_time=T - dt, time_step=dt, sampling=abi.MSG_SAMPLING_GIVEN),
                              columns=cols / math.sqrt(M))
    assert np.abs(yl - yg[0]).max() < 1e-10


def test_expm_and_eig_pins():
    rng = np.random.default_rng(3)
    A = rng.normal(size=(9, 9)) + 1j * rng.normal(size=(9, 9))
    H = A + A.conj().T
    w, V = ref.eig_hermitian(H)
    assert np.allclose(w, np.linalg.eigvalsh(H), atol=1e-12)
    assert np.abs(V @ np.diag(w) @ V.conj().T - H).max() < 1e-12
    U = ref.expm(-1j * H * 0.7)
    lam, Q = np.linalg.eigh(H)
    assert np.abs(U - Q @ np.diag(np.exp(-0.7j * lam)) @ Q.conj().T).max() < 1e-12


def test_suz_samples_statistics():
    """SU(Z) trace samples: unit norm, fixed singlet part and E|chi><chi| = 1/Z
    on the open spins (STOCHASTIC_RELAXATION.md:322-337)."""
    s = SpinSystem("c")
    s.add_spin("e1", "1/2", electron=True)
    s.add_spin("e2", "1/2", electron=True)
    s.add_spin("n1", "1/2", tensor=1.0)
    s.add_spin("n2", "1", tensor=1.0)
    s.add_radical_pair_states("e1", "e2")
    p = s.pack()
    M = 8192
    B = ref.suz_samples(p, 0, 4321, M)
    assert np.allclose(np.linalg.norm(B, axis=0), 1.0, atol=1e-14)
    P = ref.projector(p, 0)
    assert np.abs(P @ B - B).max() < 1e-14
    rho = B @ B.conj().T / M
    assert np.abs(rho - P / np.trace(P).real).max() < 6.0 / math.sqrt(M)
