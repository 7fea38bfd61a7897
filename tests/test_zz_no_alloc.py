"""msg_propagate_noalloc (every device buffer in the library's fixed scratch)
against msg_propagate on the same inputs: two calls in a row, the second smaller,
to the bit (each column runs the same arithmetic whichever workgroup takes it;
only the number of resident columns may differ).  Runs that do not fit the
scratch are refused with MSG_OOM_ADMISSION, and the dense path with
MSG_UNSUPPORTED."""
import math

import numpy as np
import pytest

from molspin_amd import abi, workloads

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from molspin_amd import backend
    c = backend.Context(0)
    yield c
    c.close()


def test_noalloc_equals_propagate_twice(ctx):
    s, cfg = workloads.config2(n_orientations=1024)
    sysd = ctx.upload(s.pack())
    # first call: 3 orientations x 1024 direct columns x 12 steps, with traces
    prm = abi.run_params(total_time=24.0, time_step=2.0, method=abi.MSG_METHOD_KRYLOV)
    fields = [cfg["fields"][i] for i in (5, 400, 901)]
    y0, t0 = sysd.propagate(prm, fields=fields, trace=True)
    y1, t1 = sysd.propagate_noalloc(prm, fields=fields, trace=True)
    assert np.abs(y1 - y0).max() <= 1e-12 and np.abs(t1 - t0).max() <= 1e-12
    # second call, smaller: 1 orientation x 4 steps, no trace
    prm2 = abi.run_params(total_time=8.0, time_step=2.0, method=abi.MSG_METHOD_KRYLOV)
    y0, _ = sysd.propagate(prm2, fields=[cfg["fields"][77]])
    y1, _ = sysd.propagate_noalloc(prm2, fields=[cfg["fields"][77]])
    assert np.abs(y1 - y0).max() <= 1e-12


def test_noalloc_given_columns_large_dimension(ctx):
    """D = 65536 trajectories (given columns, ST flips): the scratch holds only a
    few resident columns at this dimension, so the run takes fewer workgroups."""
    s, cfg = workloads.config3(n_traj=12)
    sysd = ctx.upload(s.pack())
    cols = sysd.trace_samples(s.state_index("S"), 12, cfg["sample_seed"]) / math.sqrt(12)
    prm = abi.run_params(total_time=6.0, time_step=1.0, method=abi.MSG_METHOD_KRYLOV,
                         sampling=abi.MSG_SAMPLING_GIVEN, seed=2026)
    y0, _ = sysd.propagate(prm, columns=cols)
    y1, _ = sysd.propagate_noalloc(prm, columns=cols)
    assert np.abs(y1 - y0).max() <= 1e-12
    y0, _ = sysd.propagate(prm, columns=cols[:, :3])
    y1, _ = sysd.propagate_noalloc(prm, columns=cols[:, :3])
    assert np.abs(y1 - y0).max() <= 1e-12


def test_noalloc_refusals(ctx):
    from molspin_amd import backend
    s, cfg = workloads.config1()
    small = ctx.upload(s.pack())
    prm = abi.run_params(total_time=10.0, time_step=1.0)
    with pytest.raises(backend.BackendError) as e:
        small.propagate_noalloc(prm)  # D = 48: dense path
    assert e.value.status == abi.MSG_UNSUPPORTED
    s2, cfg2 = workloads.config2(n_orientations=1024)
    big = ctx.upload(s2.pack())
    prm2 = abi.run_params(total_time=20.0, time_step=2.0, method=abi.MSG_METHOD_KRYLOV)
    with pytest.raises(backend.BackendError) as e:  # 4096 orientations x 1024 columns of partial sums
        big.propagate_noalloc(prm2, fields=[cfg2["fields"][i % 1024] for i in range(4096)])
    assert e.value.status == abi.MSG_OOM_ADMISSION and "scratch" in str(e.value)
    # the scratch is usable again after a refusal
    y0, _ = big.propagate(prm2, fields=[cfg2["fields"][3]])
    y1, _ = big.propagate_noalloc(prm2, fields=[cfg2["fields"][3]])
    assert np.abs(y1 - y0).max() <= 1e-12
