"""High-level FWI API of the reference (src/FWI.jl) over the HIP operator.

``FWI`` holds the geometry (nPad rule FWI.jl:12, mask FWI.jl:47-51) and writes
the parameter and survey files into a workspace (FWI.jl:54-58).
``compute_observation`` (FWI.jl:109-135) and ``compute_misfit``
(FWI.jl:156-189) keep the reference's semantics: shot_ids are 1-based here,
velocities are padded if needed, converted to moduli and passed to the op.
``compute_misfit_and_gradient`` adds the chain rule back to (cp, cs, rho),
which the reference gets from TensorFlow's autodiff of the same graph.
"""
from __future__ import annotations

import os
import tempfile
from dataclasses import dataclass, field

import numpy as np

from . import ops
from .utils import (moduli_to_velocity_grad, npad, padding, padding_adjoint, paraGen,
                    surveyGen, velocity_to_moduli)


@dataclass
class FWI:
    nz: int = 134
    nx: int = 384
    dz: float = 24.0
    dx: float = 24.0
    dt: float = 0.0025
    nSteps: int = 2000
    f0: float = 4.5
    nPml: int = 32
    nPad: int = -1
    para_fname: str = "para_file.json"
    survey_fname: str = "survey_file.json"
    data_dir_name: str = "Data"
    WORKSPACE: str = ""
    ind_src_x: np.ndarray = None
    ind_src_z: np.ndarray = None
    ind_rec_x: np.ndarray = None
    ind_rec_z: np.ndarray = None
    mask: np.ndarray = field(default=None, repr=False)
    mask_neg: np.ndarray = field(default=None, repr=False)

    def __post_init__(self):
        if self.nPad < 0:
            self.nPad = npad(self.nz, self.nPml)
        if not self.WORKSPACE:
            self.WORKSPACE = tempfile.mkdtemp(prefix="fwiflow_")

    @property
    def nz_pad(self):
        return self.nz + 2 * self.nPml + self.nPad

    @property
    def nx_pad(self):
        return self.nx + 2 * self.nPml

    @property
    def para_path(self):
        return os.path.join(self.WORKSPACE, self.para_fname)

    @property
    def data_dir(self):
        return os.path.join(self.WORKSPACE, self.data_dir_name)


def make_fwi(nz, nx, dz, dx, nSteps, dt, ind_src_z, ind_src_x, ind_rec_z, ind_rec_x, **kw):
    """FWI.jl:38-60: geometry, mask (interior = 1 except its top 10 rows), files."""
    fwi = FWI(nz=nz, nx=nx, dz=dz, dx=dx, nSteps=nSteps, dt=dt,
              ind_src_z=np.asarray(ind_src_z), ind_src_x=np.asarray(ind_src_x),
              ind_rec_z=np.asarray(ind_rec_z), ind_rec_x=np.asarray(ind_rec_x), **kw)
    Mask = np.zeros((fwi.nz_pad, fwi.nx_pad))
    Mask[fwi.nPml:fwi.nPml + nz, fwi.nPml:fwi.nPml + nx] = 1.0
    Mask[fwi.nPml:fwi.nPml + 10, :] = 0.0
    fwi.mask = Mask
    fwi.mask_neg = 1.0 - Mask
    assert len(ind_rec_x) == len(ind_rec_z)
    assert len(ind_src_x) == len(ind_src_z)
    paraGen(fwi.nz_pad, fwi.nx_pad, dz, dx, nSteps, dt, fwi.f0, fwi.nPml, fwi.nPad,
            fwi.para_path, os.path.join(fwi.WORKSPACE, fwi.survey_fname), fwi.data_dir)
    surveyGen(ind_src_z, ind_src_x, ind_rec_z, ind_rec_x,
              os.path.join(fwi.WORKSPACE, fwi.survey_fname))
    return fwi


def FWIExample():
    """FWI.jl:85-95."""
    ind_src_x = np.arange(4, 385, 8)
    ind_src_z = 2 * np.ones_like(ind_src_x)
    ind_rec_x = np.arange(3, 382)
    ind_rec_z = 2 * np.ones_like(ind_rec_x)
    return make_fwi(134, 384, 24.0, 24.0, 2000, 0.0025, ind_src_z, ind_src_x, ind_rec_z,
                    ind_rec_x)


def try_pad(fwi, a):
    """FWI.jl:216-223: an input of the padded size is returned as is, any
    other is padded.  The reference's padding (FWI.jl:193-204) computes a
    bilinear resize of an input that is not (nz, nx) but pads the input
    itself, so the result keeps the input's size plus the C-PML and pad
    cells; so does this one, and the operator calls below reject an array
    that is then not of the grid's size (``_on_grid``)."""
    a = np.asarray(a, dtype=np.float64)
    if a.shape == (fwi.nz_pad, fwi.nx_pad):
        return a
    return padding(a, fwi.nPml, fwi.nPad)


def _on_grid(fwi, *arrays):
    """The operator reads nz_pad x nx_pad values of each model array (the
    C ABI takes no sizes): refuse any other shape before the call."""
    for a in arrays:
        if a.shape != (fwi.nz_pad, fwi.nx_pad):
            raise ValueError(f"model array of shape {a.shape} after padding; the grid is "
                             f"{fwi.nz_pad}x{fwi.nx_pad} (pass (nz, nx) = ({fwi.nz}, {fwi.nx}) "
                             "or the padded size; inputs are not resized)")
    return arrays


def _stf_rows(fwi, stf_array):
    stf = np.asarray(stf_array, dtype=np.float64)
    if stf.ndim == 1 or stf.shape[0] == 1:  # FWI.jl:118-120
        stf = np.repeat(stf.reshape(1, -1), len(fwi.ind_src_z), axis=0)
    return stf


def compute_observation(fwi, cp, cs, rho, stf_array, shot_ids=None, gpu_id=0):
    """FWI.jl:109-135.  Returns data[g, nSteps, nrec] (float64 of the f32 files)."""
    cp_pad, cs_pad, rho_pad = _on_grid(fwi, *(try_pad(fwi, a) for a in (cp, cs, rho)))
    stf = _stf_rows(fwi, stf_array)
    lam, mu = velocity_to_moduli(cp_pad, cs_pad, rho_pad)
    if shot_ids is None:
        shot_ids = np.arange(1, len(fwi.ind_src_x) + 1)
    ids = np.asarray(shot_ids) - 1
    ops.fwi_obs_op(lam, mu, rho_pad, stf, gpu_id, ids, fwi.para_path)
    nrec = len(fwi.ind_rec_z)
    data = np.zeros((len(ids), fwi.nSteps, nrec))
    for i, sid in enumerate(ids):
        A = np.fromfile(os.path.join(fwi.data_dir, f"Shot{sid}.bin"), dtype=np.float32)
        data[i] = A.reshape(nrec, fwi.nSteps).T
    return data


def _masked(fwi, cp, cs, rho, is_masked, cp_ref, cs_ref, rho_ref):
    cp_pad, cs_pad, rho_pad = _on_grid(fwi, *(try_pad(fwi, a) for a in (cp, cs, rho)))
    if is_masked:
        return cp_pad, cs_pad, rho_pad
    if cp_ref is None or cs_ref is None or rho_ref is None:
        # FWI.jl:174-176 blends with the reference model, which must be given
        raise ValueError("is_masked=False needs cp_ref, cs_ref and rho_ref")
    refs = _on_grid(fwi, *(try_pad(fwi, a) for a in (cp_ref, cs_ref, rho_ref)))
    m, mn = fwi.mask, fwi.mask_neg
    return (cp_pad * m + refs[0] * mn, cs_pad * m + refs[1] * mn, rho_pad * m + refs[2] * mn)


def compute_misfit(fwi, cp, cs, rho, stf_array, shot_ids=None, gpu_id=0, is_masked=False,
                   cp_ref=None, cs_ref=None, rho_ref=None):
    """FWI.jl:156-189.  shot_ids 1-based; masked blend with the reference model."""
    cpm, csm, rhom = _masked(fwi, cp, cs, rho, is_masked, cp_ref, cs_ref, rho_ref)
    lam, mu = velocity_to_moduli(cpm, csm, rhom)
    stf = _stf_rows(fwi, stf_array)
    if shot_ids is None:
        shot_ids = np.arange(1, len(fwi.ind_src_x) + 1)
    ids = np.asarray(shot_ids) - 1
    return ops.fwi_op(lam, mu, rhom, stf, gpu_id, ids, fwi.para_path)


def compute_misfit_and_gradient(fwi, cp, cs, rho, stf_array, shot_ids=None, gpu_id=0,
                                is_masked=False, cp_ref=None, cs_ref=None, rho_ref=None):
    """Misfit and d(misfit)/d(cp, cs, rho) through mask, moduli and padding.

    Two operator calls (calc_id 0 then 1), as TF's evaluation of the loss and
    its gradient does with FwiOp / FwiOpGrad."""
    cpm, csm, rhom = _masked(fwi, cp, cs, rho, is_masked, cp_ref, cs_ref, rho_ref)
    lam, mu = velocity_to_moduli(cpm, csm, rhom)
    stf = _stf_rows(fwi, stf_array)
    if shot_ids is None:
        shot_ids = np.arange(1, len(fwi.ind_src_x) + 1)
    ids = np.asarray(shot_ids) - 1
    J = ops.fwi_op(lam, mu, rhom, stf, gpu_id, ids, fwi.para_path)
    gL, gM, gD, _ = ops.fwi_op_grad(lam, mu, rhom, stf, gpu_id, ids, fwi.para_path)
    g_cp, g_cs, g_rho = moduli_to_velocity_grad(cpm, csm, rhom, gL, gM, gD)
    if not is_masked:
        g_cp, g_cs, g_rho = (g * fwi.mask for g in (g_cp, g_cs, g_rho))
    out = []
    for g, a in zip((g_cp, g_cs, g_rho), (cp, cs, rho)):
        a = np.asarray(a)
        out.append(g if a.shape == g.shape else
                   padding_adjoint(g, fwi.nz, fwi.nx, fwi.nPml, fwi.nPad))
    return J, out[0], out[1], out[2]
