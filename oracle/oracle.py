"""ctypes front end of the CPU oracle (``oracle/cufd_oracle.cpp``).

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, ``__graft_entry__.smoke()``
and the CPU-baseline leg of ``bench.py``, and there only as the checker.  The
product library never imports, links or calls anything in this directory.

``cufd()`` mirrors the reference driver ``cufd`` (deps/CustomOps/FWI/Src/
libCUFD.cu:34-580) including its file side effects: it reads the one-line
parameter and survey JSON files (Parameter.cpp:16-178, Src_Rec.cu:19-245),
reads ``{data_dir}/Shot{id}.bin`` for calc_id 0/1 and writes it for calc_id 2
(utilities.cu:10-31, libCUFD.cu:514-521).
"""
from __future__ import annotations

import ctypes as C
import json
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None


class Cond(C.Structure):
    _fields_ = [("if_win", C.c_int), ("if_filter", C.c_int), ("if_src_update", C.c_int),
                ("filter", C.c_double * 4), ("win_start", C.POINTER(C.c_double)),
                ("win_end", C.POINTER(C.c_double)), ("weights", C.POINTER(C.c_double)),
                ("cond_obs", C.POINTER(C.c_float)), ("cond_syn", C.POINTER(C.c_float)),
                ("src_out", C.POINTER(C.c_float))]


class Geom(C.Structure):
    _fields_ = [("nz", C.c_int), ("nx", C.c_int), ("nPml", C.c_int), ("nPad", C.c_int),
                ("nSteps", C.c_int), ("dz", C.c_double), ("dx", C.c_double),
                ("dt", C.c_double), ("f0", C.c_double)]


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc, OpenMP)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE, "-B" if force else "liboracle.so"],
                       check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        build()
        L = C.CDLL(_LIB_PATH)
        dp, fp, ip = C.POINTER(C.c_double), C.POINTER(C.c_float), C.POINTER(C.c_int)
        L.oracle_cufd.restype = C.c_int
        L.oracle_cufd.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(Geom), C.c_int,
                                  ip, ip, ip, ip, ip, dp, dp, dp, dp, fp, fp, fp,
                                  dp, dp, dp, dp, dp, C.c_int, dp, C.POINTER(Cond)]
        L.oracle_dft.argtypes = [C.c_int, C.c_int, dp]
        L.oracle_condition.restype = C.c_int
        L.oracle_condition.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, dp, dp,
                                       dp, dp, dp, dp, dp, dp]
        L.oracle_cpml.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                                  C.c_double, dp]
        L.oracle_model.argtypes = [C.c_int, C.POINTER(Geom), dp, dp, dp, dp]
        L.oracle_window.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, dp]
        L.oracle_set_mutation.argtypes = [C.c_int]
        L.oracle_ring.restype = C.c_int
        L.oracle_ring.argtypes = [C.POINTER(Geom), C.POINTER(C.c_longlong)]
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t)) if a is not None else None


MUTATIONS = {"velocity_gate_of_stress": 1, "adj_vel_a_half": 2, "adj_vel_K_half": 3,
             "adj_vel_drop_psi": 4, "adj_stress_drop_psi": 5, "adj_stress_a_half": 6,
             "adj_stress_gated": 0x100}


class mutation:
    """Context manager planting named mistakes in the oracle (for the
    oracle's own tests: checks that must fail under them), optionally with
    the one named variant, ``adj_stress_gated`` (the reference's
    commented-out adjoint memory-variable gates restored)."""

    def __init__(self, *names):
        self.code = 0
        for n in names:
            if n is not None:
                self.code |= MUTATIONS[n]

    def __enter__(self):
        lib().oracle_set_mutation(self.code)
        return self

    def __exit__(self, *a):
        lib().oracle_set_mutation(0)


def geom(nz, nx, nPml, nPad, nSteps, dz, dx, dt, f0) -> Geom:
    return Geom(int(nz), int(nx), int(nPml), int(nPad), int(nSteps), float(dz), float(dx),
                float(dt), float(f0))


def cpml(N, nPml, dh, f0, dt, precision=32):
    """(K, a, b, K_half, a_half, b_half) of one axis, utilities.cu:242-358."""
    out = np.zeros(6 * N)
    lib().oracle_cpml(precision, N, nPml, dh, f0, dt, _p(out, C.c_double))
    return out.reshape(6, N)


def model_arrays(g: Geom, lam, mu, den, precision=32):
    """ave_mu, byc_a, byc_b, cp on the padded grid, returned as [NZ][NX]."""
    n = g.nz * g.nx
    out = np.zeros(4 * n)
    L, M, D = (np.ascontiguousarray(a, dtype=np.float64) for a in (lam, mu, den))
    lib().oracle_model(precision, C.byref(g), _p(L, C.c_double), _p(M, C.c_double),
                       _p(D, C.c_double), _p(out, C.c_double))
    return [out[k * n:(k + 1) * n].reshape(g.nx, g.nz).T.copy() for k in range(4)]


def window(data, dt, ratio, precision=32):
    """Apply the default sin^2/cos^2 time window (utilities.cu:707-747) to [ntr][nt]."""
    d = np.ascontiguousarray(np.atleast_2d(data), dtype=np.float64).copy()
    lib().oracle_window(precision, d.shape[1], d.shape[0], dt, ratio, _p(d, C.c_double))
    return d


def dft(x, sign=-1):
    """Exact DFT (sign -1 forward, +1 inverse unnormalised) of a complex vector."""
    a = np.ascontiguousarray(np.asarray(x, dtype=np.complex128)).copy()
    lib().oracle_dft(a.size, int(sign), a.view(np.float64).ctypes.data_as(C.POINTER(C.c_double)))
    return a


def _cond_call(op, data, dt, precision, params=(0, 0, 0, 0), wstart=None, wend=None,
               weight=None, aux=None, src=None, coef=None):
    d = np.ascontiguousarray(np.atleast_2d(data), dtype=np.float64).copy()
    ntr, nt = d.shape
    prm = np.ascontiguousarray(params, dtype=np.float64)
    arr = lambda a: None if a is None else np.ascontiguousarray(a, dtype=np.float64)  # noqa: E731
    ws, we, wt, ax = arr(wstart), arr(wend), arr(weight), arr(aux)
    sr = None if src is None else np.ascontiguousarray(src, dtype=np.float64).copy()
    cf = np.zeros(2 * (nt + 1)) if coef is None else np.ascontiguousarray(
        np.column_stack([np.real(coef), np.imag(coef)]).ravel(), dtype=np.float64)
    rc = lib().oracle_condition(precision, op, nt, ntr, float(dt), _p(prm, C.c_double),
                                _p(ws, C.c_double), _p(we, C.c_double), _p(wt, C.c_double),
                                _p(d, C.c_double), _p(ax, C.c_double), _p(sr, C.c_double),
                                _p(cf, C.c_double))
    if rc != 0:
        raise RuntimeError(f"oracle_condition({op}) failed")
    return d, sr, cf[0::2] + 1j * cf[1::2]


def window_traces(data, dt, wstart, wend, weights, ratio, precision=32):
    """Per-trace window and weights (utilities.cu:653-704) on [ntr][nt]."""
    return _cond_call(0, data, dt, precision, (ratio, 0, 0, 0), wstart, wend, weights)[0]


def bandpass(data, dt, f, precision=32):
    """bp_filter1d (utilities.cu:872-925) on [ntr][nt]."""
    return _cond_call(1, data, dt, precision, tuple(f))[0]


def source_update(obs, cal, src, dt, precision=32):
    """source_update (utilities.cu:927-1035): returns (cal', src', coef[nt+1])."""
    d, s, cf = _cond_call(2, cal, dt, precision, aux=obs, src=src)
    return d, s, cf


def source_update_adj(data, coef, dt, precision=32):
    """source_update_adj (utilities.cu:1037-1082) with amp_ratio 1."""
    return _cond_call(3, data, dt, precision, coef=coef)[0]


def ring_indices(g: Geom):
    n = lib().oracle_ring(C.byref(g), None)
    out = np.zeros(n, dtype=np.int64)
    lib().oracle_ring(C.byref(g), out.ctypes.data_as(C.POINTER(C.c_longlong)))
    return out


def run(g: Geom, calc_id, lam, mu, den, stf_rows, src, recs, obs=None, precision=32,
        scheme=0, ckpt_every=0, nthreads=None, illumination=False, cond=None):
    """Run the restated operator on explicit inputs.

    ``src``: list of (z, x) grid indices per shot (already + nPml);
    ``recs``: list of (z_array, x_array) per shot (already + nPml);
    ``stf_rows``: [group][nSteps]; ``obs``: list of [nrec][nSteps] float32 (calc 0/1).
    ``cond``: optional dict of the conditioning stages: ``if_win`` with
    ``win_start``/``win_end``/``weights`` (lists of per-shot arrays),
    ``filter`` (four frequencies) and ``if_src_update``.
    Returns dict with syn, res (lists), misfit, grad_lambda/mu/den [NZ][NX],
    grad_stf [group][nSteps], with ``illumination`` illum [NZ][NX], and with
    ``cond`` the conditioned obs and syn (lists) and the updated sources.
    """
    G = len(src)
    nrec = np.array([len(r[0]) for r in recs], dtype=np.int32)
    zr = np.ascontiguousarray(np.concatenate([np.asarray(r[0]) for r in recs]), dtype=np.int32)
    xr = np.ascontiguousarray(np.concatenate([np.asarray(r[1]) for r in recs]), dtype=np.int32)
    zs = np.array([s[0] for s in src], dtype=np.int32)
    xs = np.array([s[1] for s in src], dtype=np.int32)
    L, M, D = (np.ascontiguousarray(a, dtype=np.float64) for a in (lam, mu, den))
    S = np.ascontiguousarray(stf_rows, dtype=np.float64).reshape(G, g.nSteps)
    tot = int(nrec.sum()) * g.nSteps
    syn = np.zeros(tot, dtype=np.float32)
    res = np.zeros(tot, dtype=np.float32)
    ob = None
    if calc_id in (0, 1):
        ob = np.ascontiguousarray(np.concatenate([np.asarray(o, dtype=np.float32).ravel()
                                                  for o in obs]))
    misfit = np.zeros(1)
    gL = np.zeros((g.nz, g.nx))
    gM = np.zeros((g.nz, g.nx))
    gD = np.zeros((g.nz, g.nx))
    gS = np.zeros((G, g.nSteps))
    il = np.zeros((g.nz, g.nx)) if illumination else None
    if nthreads is None:
        nthreads = min(os.cpu_count() or 1, 16)
    cd, keep = None, []
    if cond is not None:
        cd = Cond()
        cd.if_win = int(bool(cond.get("if_win", False)))
        f = cond.get("filter")
        cd.if_filter = int(f is not None)
        if f is not None:
            for q in range(4):
                cd.filter[q] = float(f[q])
        cd.if_src_update = int(bool(cond.get("if_src_update", False)))
        if cd.if_win:
            for name in ("win_start", "win_end", "weights"):
                a = np.ascontiguousarray(np.concatenate([np.asarray(v, dtype=np.float64)
                                                         for v in cond[name]]))
                keep.append(a)
                setattr(cd, name, _p(a, C.c_double))
        cobs = np.zeros(tot, dtype=np.float32)
        csyn = np.zeros(tot, dtype=np.float32)
        sout = np.zeros((G, g.nSteps), dtype=np.float32)
        keep += [cobs, csyn, sout]
        cd.cond_obs, cd.cond_syn, cd.src_out = (_p(cobs, C.c_float), _p(csyn, C.c_float),
                                                _p(sout, C.c_float))
    rc = lib().oracle_cufd(precision, scheme, ckpt_every, calc_id, C.byref(g), G,
                           _p(zs, C.c_int), _p(xs, C.c_int), _p(nrec, C.c_int), _p(zr, C.c_int),
                           _p(xr, C.c_int), _p(L, C.c_double), _p(M, C.c_double),
                           _p(D, C.c_double), _p(S, C.c_double), _p(ob, C.c_float),
                           _p(syn, C.c_float), _p(res, C.c_float), _p(misfit, C.c_double),
                           _p(gL, C.c_double), _p(gM, C.c_double), _p(gD, C.c_double),
                           _p(gS, C.c_double), int(nthreads), _p(il, C.c_double),
                           C.byref(cd) if cd is not None else None)
    if rc != 0:
        raise RuntimeError(f"oracle_cufd failed with {rc}")
    offs = np.concatenate([[0], np.cumsum(nrec) * g.nSteps])
    split = lambda a: [a[offs[i]:offs[i + 1]].reshape(nrec[i], g.nSteps) for i in range(G)]
    out = dict(syn=split(syn), res=split(res), misfit=float(misfit[0]), grad_lambda=gL,
               grad_mu=gM, grad_den=gD, grad_stf=gS, illum=il)
    if cd is not None:
        out.update(cond_obs=split(cobs), cond_syn=split(csyn), src_updated=sout)
    return out


def read_json_line(path):
    """The reference reads only the first line of each JSON file (Parameter.cpp:28)."""
    with open(path) as f:
        return json.loads(f.readline())


def cufd(lam, mu, den, stf, calc_id, shot_ids, para_fname, precision=32, scheme=0,
         ckpt_every=0, nthreads=None, illumination=False):
    """File-driven restatement of ``cufd`` (libCUFD.cu:34-580).

    Returns (misfit, grad_lambda, grad_mu, grad_den, grad_stf, run_result).
    ``grad_stf`` has the shape of ``stf`` with rows 0..group-1 filled by group
    position (libCUFD.cu:454-456), the rest zero.
    """
    para = read_json_line(para_fname)
    survey = read_json_line(para["survey_fname"])
    nPml = para["nPoints_pml"]
    g = geom(para["nz"], para["nx"], nPml, para["nPad"], para["nSteps"], para["dz"],
             para["dx"], para["dt"], para["f0"])
    stf = np.asarray(stf, dtype=np.float64).reshape(-1, g.nSteps)
    src, recs, rows, obs = [], [], [], []
    cond = None
    if para.get("if_win") or "filter" in para or para.get("if_src_update"):
        cond = {"if_win": bool(para.get("if_win", False)), "filter": para.get("filter"),
                "if_src_update": bool(para.get("if_src_update", False)),
                "win_start": [], "win_end": [], "weights": []}
    for sid in shot_ids:
        sh = survey[f"shot{int(sid)}"]
        if cond is not None and cond["if_win"]:
            for name in ("win_start", "win_end", "weights"):
                cond[name].append(np.asarray(sh[name], dtype=np.float64))
        src.append((sh["z_src"] + nPml, sh["x_src"] + nPml))
        recs.append((np.asarray(sh["z_rec"]) + nPml, np.asarray(sh["x_rec"]) + nPml))
        rows.append(stf[int(sid)])
        if calc_id in (0, 1):
            fn = os.path.join(para["data_dir_name"], f"Shot{int(sid)}.bin")
            obs.append(np.fromfile(fn, dtype=np.float32).reshape(sh["nrec"], g.nSteps))
    r = run(g, calc_id, lam, mu, den, np.stack(rows), src, recs, obs or None,
            precision=precision, scheme=scheme, ckpt_every=ckpt_every, nthreads=nthreads,
            illumination=illumination, cond=cond)
    if calc_id == 2:
        for sid, d in zip(shot_ids, r["syn"]):
            d.astype(np.float32).tofile(os.path.join(para["data_dir_name"], f"Shot{int(sid)}.bin"))
    gS = np.zeros_like(stf)
    gS[:len(shot_ids)] = r["grad_stf"]
    return r["misfit"], r["grad_lambda"], r["grad_mu"], r["grad_den"], gS, r
