"""C-ABI library: loads, exports every declared symbol, and turns each of the
reference's exit(1) paths into a status code before any GPU work.  CPU only:
no compute entry point is exercised on a device here."""
import json
import os
import re

import numpy as np
import pytest

import fwiflow
from fwiflow import _lib, models
from fwiflow.utils import paraGen, surveyGen

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_symbols():
    text = open(os.path.join(ROOT, "include", "fwiflow.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fwiflow_[a-z_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = _lib.load()
    decl = _declared_symbols()
    assert len(decl) >= 20
    assert sorted(_lib.EXPORTS) == decl
    for name in decl:
        assert hasattr(L, name), name
    assert "gfx950" in fwiflow.version()


def _small(ws, **kw):
    c = models.case_small(nSteps=100)
    para, stf = models.write_case(c, ws)
    lam, mu, den = models.moduli(c, "true")
    return c, para, stf, lam, mu, den


def _code(exc):
    return exc.value.status


def test_bad_calc_id(workspace):
    c, para, stf, lam, mu, den = _small(workspace)
    with pytest.raises(fwiflow.FwiflowError) as e:
        fwiflow.check_call(3, [0], para, lam, mu, den)
    assert _code(e) == "FWIFLOW_ERR_CALC_ID"


def test_missing_parameter_file(workspace):
    c, para, stf, lam, mu, den = _small(workspace)
    with pytest.raises(fwiflow.FwiflowError) as e:
        fwiflow.check_call(2, [0], os.path.join(workspace, "nope.json"), lam, mu, den)
    assert _code(e) == "FWIFLOW_ERR_IO"


def test_garbled_and_incomplete_json(workspace):
    c, para, stf, lam, mu, den = _small(workspace)
    good = open(para).read()
    open(para, "w").write(good[:-5])
    with pytest.raises(fwiflow.FwiflowError) as e:
        fwiflow.check_call(2, [0], para, lam, mu, den)
    assert _code(e) == "FWIFLOW_ERR_JSON"
    d = json.loads(good)
    del d["nSteps"]
    open(para, "w").write(json.dumps(d))
    with pytest.raises(fwiflow.FwiflowError) as e:
        fwiflow.check_call(2, [0], para, lam, mu, den)
    assert _code(e) == "FWIFLOW_ERR_JSON" and "nSteps" in str(e.value)


def test_conditioning_options(workspace):
    """if_win / filter / if_src_update (Parameter.cpp:146-174, Src_Rec.cu:156-199)
    are parsed and validated; what the transforms cannot hold is refused."""
    c, para, stf, lam, mu, den = _small(workspace)
    d = json.loads(open(para).read())
    survey = d["survey_fname"]
    sv = json.loads(open(survey).read())

    def try_para(dd, code=None):
        open(para, "w").write(json.dumps(dd))
        if code is None:
            fwiflow.check_call(2, [0], para, lam, mu, den)
            return
        with pytest.raises(fwiflow.FwiflowError) as e:
            fwiflow.check_call(2, [0], para, lam, mu, den)
        assert _code(e) == code

    try_para(dict(d, filter=[0, 0.1, 100, 200], if_src_update=True))
    try_para(dict(d, filter=[0, 0.1, 100]), "FWIFLOW_ERR_JSON")
    try_para(dict(d, if_src_update=1), "FWIFLOW_ERR_JSON")
    try_para(dict(d, if_win=False, if_src_update=False))
    # per-trace windows need win_start / win_end / weights of nrec values
    try_para(dict(d, if_win=True), "FWIFLOW_ERR_JSON")
    nrec = sv["shot0"]["nrec"]
    sv2 = json.loads(json.dumps(sv))
    for k in range(sv2["nShots"]):
        sv2[f"shot{k}"].update(win_start=[0.0] * nrec, win_end=[0.1] * nrec, weights=[1.0] * nrec)
    open(survey, "w").write(json.dumps(sv2))
    try_para(dict(d, if_win=True))
    sv2["shot0"]["weights"] = [1.0] * (nrec - 1)
    open(survey, "w").write(json.dumps(sv2))
    try_para(dict(d, if_win=True), "FWIFLOW_ERR_SHAPE")
    open(survey, "w").write(json.dumps(sv))
    # 2*nSteps beyond one workgroup's LDS, or with a prime factor > 61
    try_para(dict(d, filter=[0, 1, 2, 3], nSteps=11000), "FWIFLOW_ERR_UNSUPPORTED")
    try_para(dict(d, if_src_update=True, nSteps=1009), "FWIFLOW_ERR_UNSUPPORTED")
    try_para(dict(d, nSteps=1009))  # no transform needed: accepted


def test_shot_id_out_of_range(workspace):
    c, para, stf, lam, mu, den = _small(workspace)
    with pytest.raises(fwiflow.FwiflowError) as e:
        fwiflow.check_call(2, [0, 7], para, lam, mu, den)
    assert _code(e) == "FWIFLOW_ERR_SHAPE"


def test_empty_shot_list(workspace):
    """An empty group is refused (the reference's loops would run no shot and
    return uninitialised outputs)."""
    c, para, stf, lam, mu, den = _small(workspace)
    with pytest.raises(fwiflow.FwiflowError) as e:
        fwiflow.check_call(2, [], para, lam, mu, den)
    assert _code(e) == "FWIFLOW_ERR_ARG"


def test_source_outside_grid(workspace):
    c, para, stf, lam, mu, den = _small(workspace)
    surveyGen([2], [c.nx + 40], c.rec_z, c.rec_x, json.loads(open(para).read())["survey_fname"])
    with pytest.raises(fwiflow.FwiflowError) as e:
        fwiflow.check_call(2, [0], para, lam, mu, den)
    assert _code(e) == "FWIFLOW_ERR_SHAPE"


def test_courant_violation(workspace):
    c, para, stf, lam, mu, den = _small(workspace)
    with pytest.raises(fwiflow.FwiflowError) as e:
        fwiflow.check_call(2, [0], para, lam * 10, mu * 10, den)
    assert _code(e) == "FWIFLOW_ERR_COURANT"


def test_missing_observed_data(workspace):
    c, para, stf, lam, mu, den = _small(workspace)
    fwiflow.check_call(2, [0, 1], para, lam, mu, den)  # calc 2 writes, needs none
    with pytest.raises(fwiflow.FwiflowError) as e:
        fwiflow.check_call(0, [0], para, lam, mu, den)
    assert _code(e) == "FWIFLOW_ERR_IO"
    # truncated Shot file
    os.makedirs(os.path.join(workspace, "Data"), exist_ok=True)
    np.zeros(10, np.float32).tofile(os.path.join(workspace, "Data", "Shot0.bin"))
    with pytest.raises(fwiflow.FwiflowError) as e:
        fwiflow.check_call(1, [0], para, lam, mu, den)
    assert _code(e) == "FWIFLOW_ERR_IO"


def test_host_tables_match_oracle(workspace, oracle_mod):
    """C-PML profiles and window tables of the GPU path equal the oracle's
    reference-arithmetic FP32 values (up to 2 ulp: host pow/exp overloads)."""
    c, para, stf, lam, mu, den = _small(workspace)
    nzA = c.NZ - c.nPad
    cz, cx, wd, ws = fwiflow.host_tables(para, nzA, c.NX, c.nSteps)
    oz = oracle_mod.cpml(nzA, c.nPml, c.dz, c.f0, c.dt, precision=32)
    ox = oracle_mod.cpml(c.NX, c.nPml, c.dx, c.f0, c.dt, precision=32)
    np.testing.assert_allclose(cz, oz, rtol=3e-7, atol=1e-12)
    np.testing.assert_allclose(cx, ox, rtol=3e-7, atol=1e-12)
    w = oracle_mod.window(np.ones((1, c.nSteps)), c.dt, 0.005, precision=32)[0]
    np.testing.assert_allclose(wd, w, rtol=3e-7, atol=1e-12)
    w = oracle_mod.window(np.ones((1, c.nSteps)), c.dt, 0.001, precision=32)[0]
    np.testing.assert_allclose(ws, w, rtol=3e-7, atol=1e-12)


def test_compute_without_gpu_fails_loudly(workspace):
    """No CPU fallback: with no device the op returns FWIFLOW_ERR_HIP."""
    try:
        import torch

        if torch.cuda.is_available():
            pytest.skip("a GPU is present")
    except ImportError:
        pass
    c, para, stf, lam, mu, den = _small(workspace)
    with pytest.raises(fwiflow.FwiflowError) as e:
        fwiflow.fwi_obs_op(lam, mu, den, stf, 0, [0], para)
    assert _code(e) == "FWIFLOW_ERR_HIP"


def test_paragen_surveygen_format(workspace):
    """The files are one-line JSON with the keys Parameter.cpp / Src_Rec.cu read."""
    p = os.path.join(workspace, "p.json")
    s = os.path.join(workspace, "s.json")
    paraGen(100, 200, 20, 20, 1000, 0.002, 10, 32, 28, p, s, os.path.join(workspace, "D"))
    surveyGen([2], [100], [1, 1], [3, 4], s)
    lines = open(p).read().splitlines()
    assert len(lines) == 1
    d = json.loads(lines[0])
    assert list(d)[:9] == ["nz", "nx", "dz", "dx", "nSteps", "dt", "f0", "nPoints_pml", "nPad"]
    assert isinstance(d["nz"], int) and isinstance(d["dt"], float)
    sv = json.loads(open(s).read())
    assert sv["nShots"] == 1 and sv["shot0"]["nrec"] == 2 and sv["shot0"]["x_rec"] == [3, 4]
    assert os.path.isdir(os.path.join(workspace, "D"))


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU: --gpus 2 exits with status 2 and says why,
    before any GPU or library work."""
    import subprocess
    import sys
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


def test_bench_byte_model_matches_the_stated_per_unit_figure():
    """766.3 GB per shot-gradient at the 500x1700 workload: 168 B per active
    cell per step plus 96 B per C-PML cell-direction (bench.py's roofline
    denominator), and the per-launch bytes scale with the shots of a launch."""
    import importlib.util
    spec = importlib.util.spec_from_file_location("bench", os.path.join(ROOT, "bench.py"))
    bench = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(bench)
    c = models.case3(nSteps=4000, nshots=2)
    assert bench.algorithmic_bytes_per_shot_gradient(c, 4000) == pytest.approx(766.272e9, rel=1e-6)
    one = bench.kernel_launch_bytes("bwd_pair_a", c, 1.0)
    assert bench.kernel_launch_bytes("bwd_pair_a", c, 64 / 22) == pytest.approx(one * 64 / 22)


def test_launch_tuning_cache_round_trip(workspace):
    """The autotuner's cache: key per (arch, grid, C-PML, shot count), entries
    kept across stores, and apply_tuned sets the cached shape on a plan."""
    import json

    from fwiflow import tune

    c = models.case_small(nz=40, nx=60, nSteps=100, nshots=3)
    para, _ = models.write_case(c, workspace)
    path = os.path.join(workspace, "tune", "launch.json")
    key = tune.cache_key(para, 3)
    assert key == f"gfx950/{c.NZ}x{c.NX}/pml{c.nPml}/shots3"
    assert tune.lookup(para, 3, path) is None
    tune.store(key, {"tile_x": 2, "shots_fwd": 3, "shots_bwd": 1}, path)
    tune.store("other", {"tile_x": 4, "shots_fwd": 1, "shots_bwd": 1}, path)
    assert json.load(open(path))[key]["shots_bwd"] == 1

    class FakePlan:
        def set_launch(self, *a):
            self.shape = a

    fp = FakePlan()
    assert tune.apply_tuned(fp, para, 3, path)["tile_x"] == 2 and fp.shape == (2, 3, 1)
    assert tune.apply_tuned(FakePlan(), para, 4, path) is None
