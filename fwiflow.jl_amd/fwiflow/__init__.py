"""fwiflow: MI355X (gfx950) elastic full-waveform-inversion operator.

Drop-in for FwiFlow.jl's ``fwi_op`` / ``fwi_obs_op`` path: the compute runs in
hand-written HIP kernels behind the C ABI of ``include/fwiflow.h``
(``lib/libfwiflow_hip.so``); this package is the host-side mirror of the
reference's Julia API (src/Core.jl, src/FWI.jl, src/Utils.jl).
"""
from .ops import Job, Plan, check_call, cufd, fwi_obs_op, fwi_op, fwi_op_grad, host_tables
from .utils import (moduli_to_velocity_grad, npad, padding, padding_adjoint, paraGen,
                    sourceGene, surveyGen, velocity_to_moduli)
from .fwi import (FWI, FWIExample, compute_misfit, compute_misfit_and_gradient,
                  compute_observation, make_fwi)
from .inversion import PlanObjective, lbfgs
from .precond import CNN, learned_descent, learned_lbfgs
from .tune import apply_tuned, autotune
from ._lib import FwiflowError, version

__all__ = [
    "apply_tuned", "autotune",
    "Job", "Plan", "check_call", "cufd", "fwi_obs_op", "fwi_op", "fwi_op_grad", "host_tables",
    "moduli_to_velocity_grad", "npad", "padding", "padding_adjoint", "paraGen", "sourceGene",
    "surveyGen", "velocity_to_moduli", "FWI", "FWIExample", "compute_misfit",
    "compute_misfit_and_gradient", "compute_observation", "make_fwi", "FwiflowError",
    "version", "PlanObjective", "lbfgs", "CNN", "learned_descent", "learned_lbfgs",
]
