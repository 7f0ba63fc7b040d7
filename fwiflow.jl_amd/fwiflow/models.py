"""Seeded synthetic elastic models and acquisition geometries.

Each ``case_*`` returns a ``Case`` with the interior model (cp, cs, rho of
shape (nz, nx)), the smoothed starting model, the acquisition (0-based indices
into the interior, as the survey file stores them), the source wavelet and the
C-PML settings.  Vs = Vp/sqrt(3); density from Gardner rho = 310 Vp^0.25.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
from scipy.ndimage import gaussian_filter

from .utils import npad, sourceGene


@dataclass
class Case:
    name: str
    nz: int
    nx: int
    dz: float
    dx: float
    nSteps: int
    dt: float
    f0: float
    nPml: int
    cp: np.ndarray
    cs: np.ndarray
    rho: np.ndarray
    cp0: np.ndarray
    cs0: np.ndarray
    rho0: np.ndarray
    src_z: np.ndarray
    src_x: np.ndarray
    rec_z: np.ndarray
    rec_x: np.ndarray
    stf: np.ndarray  # (1, nSteps), repeated per shot by the caller

    @property
    def nPad(self):
        return npad(self.nz, self.nPml)

    @property
    def NZ(self):
        return self.nz + 2 * self.nPml + self.nPad

    @property
    def NX(self):
        return self.nx + 2 * self.nPml

    @property
    def active_cells(self):
        return (self.NZ - self.nPad) * self.NX


def gardner(vp):
    return 310.0 * vp ** 0.25


def _elastic(vp, water_rows=0):
    vs = vp / np.sqrt(3.0)
    if water_rows:
        vs[:water_rows] = 0.0
    return vp, vs, gardner(vp)


def _smooth(vp, sigma, water_rows=0):
    v = gaussian_filter(vp, sigma, mode="nearest")
    return _elastic(v, water_rows)


def flat_layers(nz, nx, vmin, vmax, nlayers):
    vp = np.zeros((nz, nx))
    edges = np.linspace(0, nz, nlayers + 1).astype(int)
    for k, v in enumerate(np.linspace(vmin, vmax, nlayers)):
        vp[edges[k]:edges[k + 1]] = v
    return vp


def case_small(nz=40, nx=60, nSteps=300, nshots=2, seed=0, dh=20.0, dt=0.002, f0=10.0,
               nPml=12):
    """A few-second CPU case for parity tests: 3 layers + a Gaussian anomaly."""
    rng = np.random.default_rng(seed)
    vp = flat_layers(nz, nx, 2000.0, 3000.0, 3)
    zz, xx = np.mgrid[0:nz, 0:nx]
    cz, cx = rng.uniform(0.4, 0.7) * nz, rng.uniform(0.3, 0.7) * nx
    vp = vp * (1.0 - 0.1 * np.exp(-((zz - cz) ** 2 + (xx - cx) ** 2) / (2 * 5.0 ** 2)))
    cp, cs, rho = _elastic(vp)
    cp0, cs0, rho0 = _smooth(vp, 4.0)
    src_x = np.linspace(5, nx - 6, nshots).round().astype(int)
    rec_x = np.arange(0, nx)
    return Case("small", nz, nx, dh, dh, nSteps, dt, f0, nPml, cp, cs, rho, cp0, cs0, rho0,
                2 * np.ones(nshots, int), src_x, np.ones(nx, int), rec_x,
                sourceGene(f0, nSteps, dt))


def case1(seed=1):
    """100x200, 20 m, CPML 32, nt 1000, dt 2 ms, 5 flat layers 2000->4000 m/s,
    -10% Gaussian anomaly (sigma 10 cells) at a seeded position, 1 shot at
    (z, x) = (40 m, 2000 m), 200 receivers at z = 20 m, 10 Hz Ricker; starting
    model smoothed with sigma 8 cells."""
    nz, nx = 100, 200
    rng = np.random.default_rng(seed)
    vp = flat_layers(nz, nx, 2000.0, 4000.0, 5)
    zz, xx = np.mgrid[0:nz, 0:nx]
    cz, cx = rng.uniform(30, 70), rng.uniform(60, 140)
    vp = vp * (1.0 - 0.1 * np.exp(-((zz - cz) ** 2 + (xx - cx) ** 2) / (2 * 10.0 ** 2)))
    cp, cs, rho = _elastic(vp)
    cp0, cs0, rho0 = _smooth(vp, 8.0)
    return Case("c1", nz, nx, 20.0, 20.0, 1000, 0.002, 10.0, 32, cp, cs, rho, cp0, cs0, rho0,
                np.array([2]), np.array([100]), np.ones(nx, int), np.arange(nx),
                sourceGene(10.0, 1000, 0.002))


def case2(seed=2, water=True):
    """134x384, 24 m, CPML 32, nt 2000, dt 2 ms, 12 dipping layers 1500-4700 m/s
    with a seeded +/-5% smooth perturbation (correlation 15 cells), a 10-cell
    water layer (Vs = 0) on top, 30 shots along the surface, 384 receivers, 8 Hz.
    ``water=False`` keeps the top 10 rows elastic (Vs = Vp / sqrt(3)): the
    operator's adjoint grows without bound over long records when the model
    has a Vs = 0 layer (see the oracle tests), so gradients for inversion
    experiments use the elastic variant."""
    nz, nx = 134, 384
    rng = np.random.default_rng(seed)
    zz, xx = np.mgrid[0:nz, 0:nx]
    depth = zz + 0.08 * xx  # dipping interfaces
    edges = np.linspace(0, depth.max() + 1, 13)
    layer = np.clip(np.searchsorted(edges, depth, side="right") - 1, 0, 11)
    vp = np.linspace(1500.0, 4700.0, 12)[layer]
    pert = gaussian_filter(rng.standard_normal((nz, nx)), 15.0)
    pert = 0.05 * pert / np.abs(pert).max()
    vp = vp * (1.0 + pert)
    vp[:10] = 1500.0
    vp = np.minimum(vp, 4700.0)
    wr = 10 if water else 0
    cp, cs, rho = _elastic(vp, water_rows=wr)
    cp0, cs0, rho0 = _smooth(vp, 8.0, water_rows=wr)
    src_x = np.linspace(4, nx - 5, 30).round().astype(int)
    return Case("c2", nz, nx, 24.0, 24.0, 2000, 0.002, 8.0, 32, cp, cs, rho, cp0, cs0, rho0,
                2 * np.ones(30, int), src_x, 2 * np.ones(nx, int), np.arange(nx),
                sourceGene(8.0, 2000, 0.002))


def case3(seed=3, nSteps=4000, nshots=64):
    """500x1700, 10 m, CPML 40, nt 4000, dt 1 ms, 10 flat layers 1800->4500 m/s
    plus 3 seeded elliptical anomalies (+/-8% Vp), 64 shots at ind_z 2,
    ind_x = round(linspace(10, 1689, 64)), 1700 receivers at ind_z 2, 12 Hz."""
    nz, nx = 500, 1700
    rng = np.random.default_rng(seed)
    vp = flat_layers(nz, nx, 1800.0, 4500.0, 10)
    zz, xx = np.mgrid[0:nz, 0:nx]
    for sign in (1.0, -1.0, 1.0):
        cz, cx = rng.uniform(120, 420), rng.uniform(250, 1450)
        az, ax = rng.uniform(25, 60), rng.uniform(60, 180)
        inside = ((zz - cz) / az) ** 2 + ((xx - cx) / ax) ** 2 <= 1.0
        vp = np.where(inside, vp * (1.0 + sign * 0.08), vp)
    cp, cs, rho = _elastic(vp)
    cp0, cs0, rho0 = _smooth(vp, 8.0)
    src_x = np.round(np.linspace(10, 1689, nshots)).astype(int)
    return Case("c3", nz, nx, 10.0, 10.0, nSteps, 0.001, 12.0, 40, cp, cs, rho, cp0, cs0, rho0,
                2 * np.ones(nshots, int), src_x, 2 * np.ones(nx, int), np.arange(nx),
                sourceGene(12.0, nSteps, 0.001))


def case4(nz, nx, monitor=False, seed=4, nSteps=3000, nshots=16):
    """Time-lapse survey on one of the grids 200x600, 300x900, 450x1350 (12 m,
    CPML 32, nt 3000, dt 1.5 ms, 16 shots).  The monitor model differs from the
    baseline by a seeded -6% Vp / +2% density ellipse of 40x120 cells."""
    rng = np.random.default_rng(seed)
    zz, xx = np.mgrid[0:nz, 0:nx]
    depth = zz + 0.04 * xx
    edges = np.linspace(0, depth.max() + 1, 9)
    layer = np.clip(np.searchsorted(edges, depth, side="right") - 1, 0, 7)
    vp = np.linspace(1600.0, 4000.0, 8)[layer]
    cp, cs, rho = _elastic(vp)
    if monitor:
        cz, cx = rng.uniform(0.45, 0.65) * nz, rng.uniform(0.35, 0.65) * nx
        inside = ((zz - cz) / 20.0) ** 2 + ((xx - cx) / 60.0) ** 2 <= 1.0
        cp = np.where(inside, cp * 0.94, cp)
        cs = cp / np.sqrt(3.0)
        rho = np.where(inside, gardner(vp) * 1.02, gardner(vp))
    cp0, cs0, rho0 = _smooth(vp, 8.0)
    src_x = np.round(np.linspace(5, nx - 6, nshots)).astype(int)
    return Case(f"c4_{nz}x{nx}{'_mon' if monitor else ''}", nz, nx, 12.0, 12.0, nSteps, 0.0015,
                10.0, 32, cp, cs, rho, cp0, cs0, rho0, 2 * np.ones(nshots, int), src_x,
                2 * np.ones(nx, int), np.arange(nx), sourceGene(10.0, nSteps, 0.0015))


def write_case(case: Case, workspace: str, nshots=None):
    """Write the parameter and survey files of ``case`` into ``workspace``
    (through paraGen/surveyGen, as the FWI constructor does) and return
    (para_fname, stf[nShots, nSteps])."""
    import os

    from .utils import paraGen, surveyGen

    ns = len(case.src_x) if nshots is None else nshots
    os.makedirs(workspace, exist_ok=True)
    para = os.path.join(workspace, "para_file.json")
    survey = os.path.join(workspace, "survey_file.json")
    paraGen(case.NZ, case.NX, case.dz, case.dx, case.nSteps, case.dt, case.f0, case.nPml,
            case.nPad, para, survey, os.path.join(workspace, "Data"))
    surveyGen(case.src_z[:ns], case.src_x[:ns], case.rec_z, case.rec_x, survey)
    stf = np.repeat(case.stf.reshape(1, -1), ns, axis=0)
    return para, stf


def moduli(case: Case, which="true"):
    """Padded (lambda, mu, rho) of the true or starting model, row-major [NZ][NX]."""
    from .utils import padding, velocity_to_moduli

    cp, cs, rho = (case.cp, case.cs, case.rho) if which == "true" else (case.cp0, case.cs0,
                                                                        case.rho0)
    cp, cs, rho = (padding(a, case.nPml, case.nPad) for a in (cp, cs, rho))
    lam, mu = velocity_to_moduli(cp, cs, rho)
    return lam, mu, rho
