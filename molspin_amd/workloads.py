"""Seeded synthetic radical-pair workloads.

Every object is drawn from a SplitMix64 stream keyed by (master seed, workload,
index), so a workload is reproducible from its name alone.  Units follow MolSpin
input files: hyperfine tensors and fields in mT written with prefactor=0.001
(Example/magnetoreception/timeevolution/example_2.msd:38-53), rates in 1/ns,
times in ns.  Electrons have tensor isotropic(2.0023); nuclei tensor
isotropic(1) (the Spin default would be 2.0, SpinAPI/Spin.cpp:29).
"""
import math

from .system import SpinSystem

MASTER_SEED = 20261203
G_ELECTRON = 2.0023
MASK64 = (1 << 64) - 1


class SplitMix64:
    """SplitMix64 (Steele, Lea, Flood, OOPSLA'14) with uniform/normal helpers."""

    def __init__(self, *key):
        s = MASTER_SEED
        for k in key:
            s = (s * 0x9E3779B97F4A7C15 + (k & MASK64) + 0x632BE59BD9B4E019) & MASK64
        self.state = s

    def next(self):
        self.state = (self.state + 0x9E3779B97F4A7C15) & MASK64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def uniform(self, a=0.0, b=1.0):
        return a + (b - a) * ((self.next() >> 11) + 0.5) * 2.0 ** -53

    def normal(self):
        u1, u2 = self.uniform(), self.uniform()
        return math.sqrt(-2.0 * math.log(u1)) * math.cos(2.0 * math.pi * u2)

    def randint(self, lo, hi):
        """Uniform integer in [lo, hi]."""
        return lo + int(self.uniform() * (hi - lo + 1))


def _radical_pair(name, nuc_a, nuc_b, hyperfine, rate, field=(0.0, 0.0, 0.05), swept=False):
    """Two electrons E1, E2, nuclei on each; hyperfine(list) gives per-nucleus
    tensor specs in mT; four sinks S, T0, T+, T- at `rate` (symmetric recombination)."""
    s = SpinSystem(name)
    s.add_spin("E1", "1/2", electron=True, tensor=G_ELECTRON)
    s.add_spin("E2", "1/2", electron=True, tensor=G_ELECTRON)
    names = []
    for i, spin in enumerate(nuc_a):
        names.append(("A%d" % i, "E1"))
        s.add_spin("A%d" % i, spin, tensor=1.0)
    for i, spin in enumerate(nuc_b):
        names.append(("B%d" % i, "E2"))
        s.add_spin("B%d" % i, spin, tensor=1.0)
    for (nuc, el), hf in zip(names, hyperfine):
        s.add_hyperfine([el], [nuc], tensor=hf, prefactor=0.001, name="hf_" + nuc)
    s.add_zeeman(["E1", "E2"], field, prefactor=0.001, swept=swept, name="zee")
    s.add_radical_pair_states("E1", "E2")
    for st in ("S", "T0", "Tp", "Tm"):
        s.add_transition("k" + st, st, rate)
    s.set_initial(["S"])
    return s


def add_nuclear_couplings(s, na, nb, seed=0):
    """Couplings between nuclei of a `_radical_pair` system with na + nb spin-1/2
    nuclei (mT, hyperfine-type two-spin interactions): an anisotropic one between the
    first two nuclei of A, an isotropic one between the last nuclei of A and B (the
    most significant index bits), and an anisotropic one across the radicals between A0
    and the last B nucleus (needs na >= 2, nb >= 1)."""
    rng = SplitMix64(53, seed)

    def tensor(scale):
        m = _aniso_hf(rng)
        return [[scale * m[i][j] for j in range(3)] for i in range(3)]
    s.add_hyperfine(["A0"], ["A1"], tensor=tensor(0.2), prefactor=0.001, name="nn_a01")
    s.add_hyperfine(["A%d" % (na - 1)], ["B%d" % (nb - 1)], tensor=0.15, prefactor=0.001, name="nn_ab")
    s.add_hyperfine(["B%d" % (nb - 1)], ["A0"], tensor=tensor(0.1), prefactor=0.001, name="nn_ba")
    return s


def config1():
    """2 x spin-1/2 on A (0.5, 1.2 mT), 1 x spin-1 on B (0.8 mT); D = 48;
    B = 0.05 mT along z; k = 1e-3 /ns; 0-5000 ns at dt = 1 ns."""
    s = _radical_pair("cfg1", ["1/2", "1/2"], ["1"], [0.5, 1.2, 0.8], 1e-3)
    return s, dict(total_time=5000.0, time_step=1.0, krylov_size=16)


def _aniso_hf(rng):
    """Isotropic U(0.1, 1.5) mT plus a random symmetric traceless part of
    Frobenius norm u * 1 mT, u ~ U(0, 1)."""
    a = rng.uniform(0.1, 1.5)
    g = [[rng.normal() for _ in range(3)] for _ in range(3)]
    sym = [[0.5 * (g[i][j] + g[j][i]) for j in range(3)] for i in range(3)]
    tr = (sym[0][0] + sym[1][1] + sym[2][2]) / 3.0
    for i in range(3):
        sym[i][i] -= tr
    nrm = math.sqrt(sum(sym[i][j] ** 2 for i in range(3) for j in range(3)))
    u = rng.uniform()
    m = [[a * (i == j) + sym[i][j] * u / nrm for j in range(3)] for i in range(3)]
    return m


def config2(n_orientations=1024, isotropic=False):
    """6 + 4 spin-1/2 nuclei with anisotropic hyperfine, D = 4096; field
    directions from the Fibonacci sphere (|B| = 0.05 mT); k = 1e-3; 0-3000 ns at
    dt = 2 ns; Lanczos order 16.  isotropic: each tensor replaced by its isotropic
    part (a third of its trace)."""
    rng = SplitMix64(2, 0)
    hf = [_aniso_hf(rng) for _ in range(10)]
    if isotropic:
        hf = [(t[0][0] + t[1][1] + t[2][2]) / 3.0 for t in hf]
    s = _radical_pair("cfg2", ["1/2"] * 6, ["1/2"] * 4, hf, 1e-3, swept=True)
    fields = [fibonacci_point(n_orientations, i, 0.05) for i in range(n_orientations)]
    return s, dict(total_time=3000.0, time_step=2.0, krylov_size=16, fields=fields)


def config3(n_traj=2048):
    """8 + 6 spin-1/2 nuclei, isotropic hyperfine log-uniform in [0.05, 1.5] mT,
    D = 65536; spin-coherent nuclear states (x) singlet, n_traj trajectories;
    ST dephasing 1e-4 /ns; k = 1e-3; 0-2000 ns at dt = 1 ns."""
    rng = SplitMix64(3, 0)
    lo, hi = math.log(0.05), math.log(1.5)
    hf = [math.exp(rng.uniform(lo, hi)) for _ in range(14)]
    s = _radical_pair("cfg3", ["1/2"] * 8, ["1/2"] * 6, hf, 1e-3)
    s.add_st_dephasing("E1", "E2", 1e-4)
    return s, dict(total_time=2000.0, time_step=1.0, krylov_size=16, n_traj=n_traj,
                   sample_seed=(MASTER_SEED << 8) | 3)


def config4_pair(index):
    """Radical pair `index` of the uneven batch: n ~ U{2..12} nuclei split 60/40
    between the radicals, hyperfine as config2, k ~ U(5e-4, 5e-3) /ns; fields
    logspace(-3, 1) mT along z (64 points); 0-3000 ns at dt = 2 ns."""
    rng = SplitMix64(4, index)
    n = rng.randint(2, 12)
    na = max(1, int(math.ceil(0.6 * n)))
    nb = n - na
    hf = [_aniso_hf(rng) for _ in range(n)]
    k = rng.uniform(5e-4, 5e-3)
    s = _radical_pair("cfg4_%d" % index, ["1/2"] * na, ["1/2"] * nb, hf, k, field=(0.0, 0.0, 1.0), swept=True)
    fields = [(0.0, 0.0, b) for b in logspace(-3.0, 1.0, 64)]
    return s, dict(total_time=3000.0, time_step=2.0, krylov_size=16, fields=fields, n=n, k=k)


def serving_request(index):
    """Request `index` of the serving stream: a config-4-style pair with 16-128
    fields and an exponential inter-arrival gap (mean 1/200 s)."""
    rng = SplitMix64(5, index)
    pair = rng.randint(0, 1 << 20)
    sys_, p = config4_pair(pair)
    nf = rng.randint(16, 128)
    fields = [(0.0, 0.0, b) for b in logspace(-3.0, 1.0, nf)]
    gap = -math.log(rng.uniform()) / 200.0
    p["fields"] = fields
    return sys_, p, gap


def fibonacci_point(n, index, magnitude):
    """ActionFibonacciSphere point (RunSection/Utility.cpp:44-79): y and theta
    pass through float before the trigonometry."""
    import numpy as np
    phi = math.pi * (3.0 - math.sqrt(5.0))
    y = np.float32(1.0 - (index / (n - 1)) * 2)
    th = np.float32(phi * index)
    r = math.sqrt(1.0 - float(y * y))  # float product, as in the reference
    x = float(np.cos(th)) * r
    z = float(np.sin(th)) * r
    return (magnitude * x, magnitude * float(y), magnitude * z)


def logspace(a, b, n):
    """arma::logspace(a, b, n) (ActionLogSpace.cpp:16)."""
    if n == 1:
        return [10.0 ** b]
    return [10.0 ** (a + (b - a) * i / (n - 1)) for i in range(n)]
