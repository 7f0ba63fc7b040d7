"""The operator: ``fwi_op`` / ``fwi_obs_op`` with the reference's signatures.

Reference: src/Core.jl:20-53 (Julia bindings), deps/CustomOps/FWI/FwiOp.cpp
(TF ops FwiOp / FwiOpGrad / FwiObsOp) and FwiOp.h:14-34 (forward / backward /
obscalc -> cufd(calc_id)).  Each call goes through the C ABI
(``fwiflow_cufd``) to the HIP kernels; there is no CPU path.

* ``fwi_op(lambda, mu, den, stf, gpu_id, shot_ids, para_fname) -> misfit``
  (calc_id 0).
* ``fwi_op_grad(...) -> (grad_lambda, grad_mu, grad_den, grad_stf)``
  (calc_id 1; FwiOpGrad, which ignores the upstream dy).
* ``fwi_obs_op(...) -> 0.0`` writes ``{data_dir}/Shot{id}.bin`` (calc_id 2).
* ``FwiOp`` : torch.autograd.Function whose backward is the gradient scaled by
  dy, the contract a ChainRulesCore ``rrule`` returns.
"""
from __future__ import annotations

import numpy as np

from . import _lib
from ._lib import check, f64, i32, ptr


def _args(lam, mu, den, stf, shot_ids):
    lam, mu, den, stf = f64(lam), f64(mu), f64(den), f64(np.atleast_2d(stf))
    if lam.ndim != 2 or mu.shape != lam.shape or den.shape != lam.shape:
        raise ValueError("lambda, mu and den must be 2-D arrays of one shape")
    return lam, mu, den, stf, i32(shot_ids)


def cufd(lam, mu, den, stf, calc_id, gpu_id, shot_ids, para_fname, noalloc=False):
    """Direct binding of ``fwiflow_cufd`` (libCUFD.cu:34-38); ``noalloc=True``
    calls ``fwiflow_cufd_noalloc``, which takes every device buffer from the
    library's fixed per-GPU scratch array instead of allocating."""
    L = _lib.load()
    lam, mu, den, stf, ids = _args(lam, mu, den, stf, shot_ids)
    misfit = np.zeros(1)
    gL = np.zeros_like(lam) if calc_id == 1 else None
    gM = np.zeros_like(lam) if calc_id == 1 else None
    gD = np.zeros_like(lam) if calc_id == 1 else None
    gS = np.zeros_like(stf) if calc_id == 1 else None
    D = _lib.C.c_double
    fn = L.fwiflow_cufd_noalloc if noalloc else L.fwiflow_cufd
    rc = fn(ptr(misfit, D), ptr(gL, D), ptr(gM, D), ptr(gD, D), ptr(gS, D),
            ptr(lam, D), ptr(mu, D), ptr(den, D), ptr(stf, D), int(calc_id),
            int(gpu_id), len(ids), ptr(ids, _lib.C.c_int), str(para_fname).encode())
    check(rc)
    return float(misfit[0]), gL, gM, gD, gS


def fwi_op(lam, mu, den, stf, gpu_id, shot_ids, para_fname):
    """Misfit 0.5*sum(res^2) over the shots (src/Core.jl:20-31, calc_id 0)."""
    return cufd(lam, mu, den, stf, 0, gpu_id, shot_ids, para_fname)[0]


def fwi_op_grad(lam, mu, den, stf, gpu_id, shot_ids, para_fname):
    """Gradient outputs of FwiOpGrad (FwiOp.cpp:140-257, calc_id 1)."""
    _, gL, gM, gD, gS = cufd(lam, mu, den, stf, 1, gpu_id, shot_ids, para_fname)
    return gL, gM, gD, gS


def fwi_obs_op(lam, mu, den, stf, gpu_id, shot_ids, para_fname):
    """Observed-data generation (src/Core.jl:42-53, calc_id 2); returns 0.0."""
    cufd(lam, mu, den, stf, 2, gpu_id, shot_ids, para_fname)
    return 0.0


def check_call(calc_id, shot_ids, para_fname, lam, mu, den):
    """Host-only validation (no GPU): raises FwiflowError like the compute path would."""
    L = _lib.load()
    lam, mu, den = f64(lam), f64(mu), f64(den)
    ids = i32(shot_ids)
    D = _lib.C.c_double
    check(L.fwiflow_check(int(calc_id), len(ids), ptr(ids, _lib.C.c_int),
                          str(para_fname).encode(), ptr(lam, D), ptr(mu, D), ptr(den, D)))


def host_tables(para_fname, nzA, nx, nSteps):
    """C-PML profiles and window tables the GPU path uses (no GPU needed)."""
    L = _lib.load()
    cz = np.zeros(6 * nzA, dtype=np.float32)
    cx = np.zeros(6 * nx, dtype=np.float32)
    wd = np.zeros(nSteps, dtype=np.float32)
    ws = np.zeros(nSteps, dtype=np.float32)
    F = _lib.C.c_float
    check(L.fwiflow_host_tables(str(para_fname).encode(), ptr(cz, F), ptr(cx, F), ptr(wd, F),
                                ptr(ws, F)))
    return cz.reshape(6, nzA), cx.reshape(6, nx), wd, ws


class Plan:
    """Persistent device-resident plan (fwiflow_plan_* in include/fwiflow.h)."""

    def __init__(self, para_fname, shot_ids, gpu_id=0):
        L = _lib.load()
        self._L = L
        self._h = _lib.C.c_void_p()
        self.shot_ids = i32(shot_ids)
        check(L.fwiflow_plan_create(_lib.C.byref(self._h), str(para_fname).encode(), int(gpu_id),
                                    len(self.shot_ids), ptr(self.shot_ids, _lib.C.c_int)))
        dims = np.zeros(6, dtype=np.int32)
        check(L.fwiflow_plan_dims(self._h, ptr(dims, _lib.C.c_int)))
        self.nz, self.nx, self.nSteps, self.ns, self.nrec_total, self.nPml = map(int, dims)

    def close(self):
        if self._h:
            self._L.fwiflow_plan_destroy(self._h)
            self._h = _lib.C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_model(self, lam, mu, den, zfast=False):
        lam, mu, den = f64(lam), f64(mu), f64(den)
        D = _lib.C.c_double
        fn = self._L.fwiflow_plan_set_model_zfast if zfast else self._L.fwiflow_plan_set_model
        check(fn(self._h, ptr(lam, D), ptr(mu, D), ptr(den, D)))

    def set_stf(self, stf):
        stf = f64(np.atleast_2d(stf))
        check(self._L.fwiflow_plan_set_stf(self._h, ptr(stf, _lib.C.c_double)))

    def set_obs(self, obs):
        """obs: concatenated f32 [nrec][nSteps] per shot (list of arrays or flat)."""
        if isinstance(obs, (list, tuple)):
            obs = np.concatenate([np.asarray(o, dtype=np.float32).ravel() for o in obs])
        obs = np.ascontiguousarray(obs, dtype=np.float32)
        if obs.size != self.nrec_total * self.nSteps:
            raise ValueError("obs size does not match the plan's receivers x nSteps")
        check(self._L.fwiflow_plan_set_obs(self._h, ptr(obs, _lib.C.c_float)))

    def load_obs(self):
        check(self._L.fwiflow_plan_load_obs(self._h))

    def run(self, calc_id):
        check(self._L.fwiflow_plan_run(self._h, int(calc_id)))

    def misfit(self):
        out = np.zeros(1)
        check(self._L.fwiflow_plan_misfit(self._h, ptr(out, _lib.C.c_double)))
        return float(out[0])

    def gradients(self):
        D = _lib.C.c_double
        gL = np.zeros((self.nz, self.nx))
        gM = np.zeros_like(gL)
        gD = np.zeros_like(gL)
        gS = np.zeros((self.ns, self.nSteps))
        check(self._L.fwiflow_plan_gradients(self._h, ptr(gL, D), ptr(gM, D), ptr(gD, D),
                                             ptr(gS, D)))
        return gL, gM, gD, gS

    def data(self):
        out = np.zeros(self.nrec_total * self.nSteps, dtype=np.float32)
        check(self._L.fwiflow_plan_data(self._h, ptr(out, _lib.C.c_float)))
        return out

    def write_data(self):
        check(self._L.fwiflow_plan_write_data(self._h))

    def residual(self):
        """Adjoint source of the last calc_id 0/1 run (after the conditioning
        stages), every shot [nrec][nSteps] concatenated."""
        out = np.zeros(self.nrec_total * self.nSteps, dtype=np.float32)
        check(self._L.fwiflow_plan_residual(self._h, ptr(out, _lib.C.c_float)))
        return out

    def source(self):
        """Source rows [group][nSteps] after the last run (updated by
        ``if_src_update``)."""
        out = np.zeros((self.ns, self.nSteps), dtype=np.float32)
        check(self._L.fwiflow_plan_source(self._h, ptr(out, _lib.C.c_float)))
        return out

    def set_checkpoint(self, k):
        check(self._L.fwiflow_plan_set_checkpoint(self._h, int(k)))

    def set_adjoint(self, mode):
        """'reference' (default) or 'exact': the transpose of the forward
        step, whose gradients are the derivatives of the discrete misfit."""
        modes = {"reference": 0, "exact": 1}
        if mode not in modes:
            raise ValueError(f"adjoint mode {mode!r}: expected one of {sorted(modes)}")
        check(self._L.fwiflow_plan_set_adjoint(self._h, modes[mode]))

    def set_timing(self, on=True):
        check(self._L.fwiflow_plan_set_timing(self._h, 1 if on else 0))

    def timing(self, name):
        ms = np.zeros(1)
        cnt = _lib.C.c_longlong(0)
        check(self._L.fwiflow_plan_timing(self._h, name.encode(), ptr(ms, _lib.C.c_double),
                                          _lib.C.byref(cnt)))
        return float(ms[0]), int(cnt.value)

    def device_bytes(self):
        b = _lib.C.c_longlong(0)
        check(self._L.fwiflow_plan_device_bytes(self._h, _lib.C.byref(b)))
        return int(b.value)

    def set_illumination(self, on=True):
        """Also sum the forward illumination (vz^2 + vx^2 over shots and steps)
        in later runs."""
        check(self._L.fwiflow_plan_set_illumination(self._h, 1 if on else 0))

    def illumination(self):
        """Illumination of the last run, [NZ][NX] float64."""
        out = np.zeros((self.nz, self.nx))
        check(self._L.fwiflow_plan_illumination(self._h, ptr(out, _lib.C.c_double)))
        return out

    def shots_per_launch(self):
        """(forward, backward) shots of the largest launch group of the last run."""
        f, b = _lib.C.c_int(0), _lib.C.c_int(0)
        check(self._L.fwiflow_plan_shots_per_launch(self._h, _lib.C.byref(f), _lib.C.byref(b)))
        return int(f.value), int(b.value)

    def launch_groups(self):
        """(forward, backward) shot groups, i.e. launches per time step, of the last run."""
        f, b = _lib.C.c_int(0), _lib.C.c_int(0)
        check(self._L.fwiflow_plan_launch_groups(self._h, _lib.C.byref(f), _lib.C.byref(b)))
        return int(f.value), int(b.value)

    def set_memory_budget(self, nbytes):
        """Device bytes a gradient run may hold for the boundary ring and checkpoints."""
        check(self._L.fwiflow_plan_set_memory_budget(self._h, float(nbytes)))

    def set_launch(self, tile_x, shots_fwd=-1, shots_bwd=-1):
        """Launch shape: tile width 1-4, shots per forward / backward launch
        (0: all, -1: automatic)."""
        check(self._L.fwiflow_plan_set_launch(self._h, int(tile_x), int(shots_fwd),
                                              int(shots_bwd)))

    def launch(self):
        """(tile width, shots per forward launch, shots per backward launch)."""
        t, f, b = _lib.C.c_int(0), _lib.C.c_int(0), _lib.C.c_int(0)
        check(self._L.fwiflow_plan_launch(self._h, _lib.C.byref(t), _lib.C.byref(f),
                                          _lib.C.byref(b)))
        return int(t.value), int(f.value), int(b.value)

    def memory_plan(self):
        """(checkpoint interval, shots per sweep, ring bytes, checkpoint bytes) of
        the last gradient run; interval 0 is the ring of every step."""
        k, s = _lib.C.c_int(0), _lib.C.c_int(0)
        rb, cb = _lib.C.c_longlong(0), _lib.C.c_longlong(0)
        check(self._L.fwiflow_plan_memory_plan(self._h, _lib.C.byref(k), _lib.C.byref(s),
                                               _lib.C.byref(rb), _lib.C.byref(cb)))
        return int(k.value), int(s.value), int(rb.value), int(cb.value)


class Job:
    """Several plans run together (fwiflow_job_* in include/fwiflow.h), e.g.
    baseline and monitor surveys on grids of different sizes: fused (every
    time-step launch covers a group of shots of every plan) where that keeps
    launches at least as large as the plans' own, else one plan after
    another.  The plans must share the GPU and nSteps and must outlive the
    job; results are read from each plan."""

    def __init__(self, plans=()):
        L = _lib.load()
        self._L = L
        self._h = _lib.C.c_void_p()
        self.plans = []
        check(L.fwiflow_job_create(_lib.C.byref(self._h)))
        for p in plans:
            self.add(p)

    def add(self, plan):
        check(self._L.fwiflow_job_add(self._h, plan._h))
        self.plans.append(plan)

    def run(self, calc_id):
        check(self._L.fwiflow_job_run(self._h, int(calc_id)))

    def fused(self):
        """Whether the last run fused every plan into each launch (else the
        plans ran one after another)."""
        f = _lib.C.c_int(0)
        check(self._L.fwiflow_job_fused(self._h, _lib.C.byref(f)))
        return bool(f.value)

    def close(self):
        if self._h:
            self._L.fwiflow_job_destroy(self._h)
            self._h = _lib.C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


try:  # torch is plumbing for autodiff callers; the op itself does not need it
    import torch

    class FwiOp(torch.autograd.Function):
        """misfit = FwiOp.apply(lam, mu, den, stf, gpu_id, shot_ids, para_fname).

        backward returns (dJ/dlam*dy, dJ/dmu*dy, dJ/dden*dy, dJ/dstf*dy, None...),
        the rrule contract.  Tensors are CPU float64 (the reference's op dtype).
        """

        @staticmethod
        def forward(ctx, lam, mu, den, stf, gpu_id, shot_ids, para_fname):
            args = [t.detach().cpu().numpy() for t in (lam, mu, den, stf)]
            ctx.saved = (args, gpu_id, np.asarray(shot_ids), para_fname)
            return torch.tensor(fwi_op(*args, gpu_id, shot_ids, para_fname),
                                dtype=torch.float64)

        @staticmethod
        def backward(ctx, dy):
            args, gpu_id, shot_ids, para_fname = ctx.saved
            gL, gM, gD, gS = fwi_op_grad(*args, gpu_id, shot_ids, para_fname)
            s = float(dy)
            t = lambda a: torch.from_numpy(a * s)
            return t(gL), t(gM), t(gD), t(gS), None, None, None
except ImportError:  # pragma: no cover
    FwiOp = None
