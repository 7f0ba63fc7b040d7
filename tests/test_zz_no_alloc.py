"""fwiflow_cufd_noalloc (device buffers carved from the library's fixed
per-GPU scratch array) against fwiflow_cufd (hipMalloc'd buffers), through
the C ABI.

Both entry points run the same kernels on the same plan, so wherever the
scratch array holds the full boundary ring their results are equal bit for
bit.  Calls run back to back, the later ones smaller, so that stale contents
a call leaves in the shared array would show in the next call's results
if anything read memory it had not written."""
import numpy as np
import pytest

import fwiflow
from fwiflow import _lib, models
from fwiflow.ops import cufd

pytestmark = pytest.mark.gpu

SCRATCH_BYTES = 128 << 20  # FWIFLOW_SCRATCH_BYTES in include/fwiflow.h
GRAD_TOL = 1e-3


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def _case(ws, **kw):
    c = models.case_small(**kw)
    para, stf = models.write_case(c, ws)
    ids = list(range(len(c.src_x)))
    lam, mu, den = models.moduli(c, "true")
    fwiflow.fwi_obs_op(lam * 1.02, mu, den, stf, 0, ids, para)
    return para, stf, ids, models.moduli(c, "start")


def _both(para, stf, ids, m0, calc_id):
    a = cufd(*m0, stf, calc_id, 0, ids, para, noalloc=True)
    b = cufd(*m0, stf, calc_id, 0, ids, para)
    return a, b


def test_noalloc_equals_cufd_on_two_calls_in_a_row(tmp_path):
    """A 3-shot gradient on a 64x84 grid, then a 1-shot gradient and misfit
    on a 54x64 grid: every output equal to fwiflow_cufd's, bit for bit."""
    big = _case(str(tmp_path / "a"), nz=40, nx=60, nSteps=300, nshots=3)
    small = _case(str(tmp_path / "b"), nz=30, nx=40, nSteps=200, nshots=1, seed=5)
    for name, (para, stf, ids, m0) in (("3 shots", big), ("1 shot", small)):
        print("noalloc gradient", name, flush=True)
        a, b = _both(para, stf, ids, m0, 1)
        for x, y in zip(a[1:], b[1:]):
            assert np.array_equal(x, y)
        a, b = _both(para, stf, ids, m0, 0)
        assert a[0] == b[0] and a[0] > 0


def test_noalloc_forward_data_equal_cufd(tmp_path):
    """calc_id 2 writes the same Shot files."""
    import os

    para, stf, ids, m0 = _case(str(tmp_path), nz=40, nx=60, nSteps=250, nshots=2)
    data = os.path.join(str(tmp_path), "Data")
    out = {}
    for noalloc in (False, True):
        cufd(*m0, stf, 2, 0, ids, para, noalloc=noalloc)
        out[noalloc] = [np.fromfile(os.path.join(data, f"Shot{i}.bin"), dtype=np.float32)
                        for i in ids]
    for x, y in zip(out[True], out[False]):
        assert np.array_equal(x, y)


def test_noalloc_keeps_checkpoints_when_the_ring_does_not_fit(tmp_path):
    """100x200 (+32-cell C-PML), 1000 steps, 3 shots: the full ring of every
    step (about 60 MB a shot) does not fit in the scratch array, so the call
    keeps checkpoints (so its gradients differ from fwiflow_cufd's in the
    last bits), within the contract tolerance of them; misfit equal (the
    forward pass is the same)."""
    import dataclasses

    ws = str(tmp_path)
    c = dataclasses.replace(models.case1(), src_z=np.array([2, 2, 2]), src_x=np.array([40, 100, 160]))
    para, stf = models.write_case(c, ws)
    ids = [0, 1, 2]
    lam, mu, den = models.moduli(c, "true")
    fwiflow.fwi_obs_op(lam, mu, den, stf, 0, ids, para)
    m0 = models.moduli(c, "start")
    a, b = _both(para, stf, ids, m0, 1)
    errs = [rel(x, y) for x, y in zip(a[1:], b[1:])]
    print("checkpointed noalloc vs cufd", errs, flush=True)
    assert 0 < max(errs) < GRAD_TOL
    a, b = _both(para, stf, ids, m0, 0)
    assert a[0] == b[0]


def test_noalloc_rejects_a_grid_beyond_the_scratch_array(tmp_path):
    """Three shots on the 500x1700 grid (+40-cell C-PML): their forward state
    alone (150 MB) exceeds the array; the call fails with FWIFLOW_ERR_MEMORY
    before any kernel runs, and the next call works."""
    c = models.case3(nSteps=20, nshots=64)
    para, stf = models.write_case(c, str(tmp_path / "big"))
    m = models.moduli(c, "start")
    with pytest.raises(_lib.FwiflowError) as e:
        cufd(*m, stf, 2, 0, [0, 1, 2], para, noalloc=True)
    assert e.value.status == "FWIFLOW_ERR_MEMORY"
    assert str(SCRATCH_BYTES) in str(e.value)
    para, stf, ids, m0 = _case(str(tmp_path / "s"), nz=30, nx=40, nSteps=150, nshots=1)
    a, b = _both(para, stf, ids, m0, 0)
    assert a[0] == b[0]
