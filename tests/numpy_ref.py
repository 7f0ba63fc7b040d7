"""Second, independent restatement of the reference's elastic operator in
vectorised numpy (FP64), for tiny grids only.

Written from the reference source text (deps/CustomOps/FWI/Src/el_stress.cu,
el_velocity.cu, el_velocity_adj.cu, el_stress_adj.cu, utilities.cu,
libCUFD.cu) with whole-array slicing instead of the C++ oracle's per-cell
loops, so a transcription slip in either one (a dropped term, a wrong sign,
index or PML gate) shows up as a disagreement between the two.  Arrays are
indexed [z, x] here (the oracle stores x-major); the window, C-PML profiles
and model averages are taken from their own closed forms below.
"""
from __future__ import annotations

import numpy as np

C1, C2 = 9.0 / 8.0, 1.0 / 24.0


def cpml_1d(N, nPml, dh, f0, dt):
    """C-PML profiles K, a, b, K_half, a_half, b_half (utilities.cu:242-358)."""
    L = nPml * dh
    d0 = -(8 + 1) * 3000.0 * np.log(0.0008) / (2.0 * L)
    amax = np.pi * f0

    def prof(depth):
        xn = depth / L
        d = d0 * (0.25 * xn + 0.75 * xn ** 8)
        k = 1.0 + xn ** 8
        al = amax * (1.0 - xn)
        return d, k, al

    out = {k: np.zeros(N) for k in ("d", "K", "al", "dh", "Kh", "alh")}
    out["K"][:] = 1.0
    out["Kh"][:] = 1.0
    i = np.arange(N, dtype=float)
    for depth, (kd, kk, ka) in (((nPml - i) * dh, ("d", "K", "al")),
                                ((nPml - i - 0.5) * dh, ("dh", "Kh", "alh")),
                                ((nPml - N + i) * dh, ("d", "K", "al")),
                                ((nPml - N + i + 0.5) * dh, ("dh", "Kh", "alh"))):
        sel = depth >= 0
        d, k, al = prof(depth[sel])
        out[kd][sel], out[kk][sel], out[ka][sel] = d, k, al
    res = []
    for d, K, al in ((out["d"], out["K"], out["al"]), (out["dh"], out["Kh"], out["alh"])):
        al = np.maximum(al, 0.0)
        b = np.exp(-(d / K + al) * dt)
        with np.errstate(invalid="ignore", divide="ignore"):
            a = np.where(np.abs(d) > 1e-6, d * (b - 1.0) / (K * (d + K * al)), 0.0)
        res.append((K, a, b))
    (K, a, b), (Kh, ah, bh) = res
    return K, a, b, Kh, ah, bh


def window_sq(nt, dt, ratio):
    t = np.arange(nt) * dt
    T = nt * dt
    off = T * ratio
    if 2 * off >= T:
        return np.ones(nt)
    t1, t2 = off, T - off
    w = np.where(t < t1, np.sin(np.pi / 2 * t / t1),
                 np.where(t < t2, 1.0, np.where(t < T, np.cos(np.pi / 2 * (t - t2) / (T - t2)),
                                                0.0)))
    return w * w


class Op:
    """Forward + adjoint of the reference scheme on a [NZ, NX] grid."""

    def __init__(self, NZ, NX, nPml, nPad, dz, dx, dt, f0, lam, mu, den):
        self.NZ, self.NX, self.nPml, self.nPad = NZ, NX, nPml, nPad
        self.dz, self.dx, self.dt = dz, dx, dt
        self.lam = lam * 1e6
        self.mu = mu * 1e6
        self.den = den.copy()
        nzA = NZ - nPad
        self.nzA = nzA
        z = cpml_1d(nzA, nPml, dz, f0, dt)
        x = cpml_1d(NX, nPml, dx, f0, dt)
        pad = lambda v, fill: np.concatenate([v, np.full(nPad, fill)])
        self.Kz, self.az, self.bz, self.Kzh, self.azh, self.bzh = (
            pad(v, f)[:, None] for v, f in zip(z, (1, 0, 0, 1, 0, 0)))
        self.Kx, self.ax, self.bx, self.Kxh, self.axh, self.bxh = (v[None, :] for v in x)
        m = self.mu
        amu = np.zeros_like(m)
        ba = np.full_like(m, 1e-3)
        bb = np.full_like(m, 1e-3)
        s = (slice(2, NZ - 2), slice(2, NX - 2))
        A, B, C, D = m[2:NZ - 2, 2:NX - 2], m[3:NZ - 1, 2:NX - 2], m[2:NZ - 2, 3:NX - 1], m[3:NZ - 1, 3:NX - 1]
        with np.errstate(divide="ignore"):
            h = 4.0 / (1 / A + 1 / B + 1 / C + 1 / D)
        amu[s] = np.where((A == 0) | (B == 0) | (C == 0) | (D == 0), 0.0, h)
        ba[s] = 2.0 / (den[3:NZ - 1, 2:NX - 2] + den[2:NZ - 2, 2:NX - 2])
        bb[s] = 2.0 / (den[2:NZ - 2, 3:NX - 1] + den[2:NZ - 2, 2:NX - 2])
        self.amu, self.ba, self.bb = amu, ba, bb
        # masks
        zi = np.arange(NZ)[:, None]
        xi = np.arange(NX)[None, :]
        self.U = (zi >= 2) & (zi <= nzA - 3) & (xi >= 2) & (xi <= NX - 3)
        self.pz = (zi < nPml) | (zi > nzA - nPml - 1)
        self.px_s = (xi < nPml) | (xi > NX - nPml - 1)
        self.px_v = (xi < nPml) | (xi > NX - nPml)
        self.I = (zi >= nPml) & (zi <= nzA - 1 - nPml) & (xi >= nPml) & (xi <= NX - 1 - nPml)

    # shifted view helper: f[z+dz, x+dx] evaluated on the full grid (zeros outside)
    @staticmethod
    def sh(f, dz, dx):
        out = np.zeros_like(f)
        NZ, NX = f.shape
        zs = slice(max(0, -dz), NZ - max(0, dz))
        xs = slice(max(0, -dx), NX - max(0, dx))
        zt = slice(max(0, dz), NZ - max(0, -dz))
        xt = slice(max(0, dx), NX - max(0, -dx))
        out[zs, xs] = f[zt, xt]
        return out

    def Dz_f(self, f):  # forward-type z difference at (z, z-1) stencil: f(z)-f(z-1)
        S = self.sh
        return (C1 * (f - S(f, -1, 0)) - C2 * (S(f, 1, 0) - S(f, -2, 0))) / self.dz

    def Dz_b(self, f):  # f(z+1)-f(z)
        S = self.sh
        return (C1 * (S(f, 1, 0) - f) - C2 * (S(f, 2, 0) - S(f, -1, 0))) / self.dz

    def Dx_f(self, f):
        S = self.sh
        return (C1 * (f - S(f, 0, -1)) - C2 * (S(f, 0, 1) - S(f, 0, -2))) / self.dx

    def Dx_b(self, f):
        S = self.sh
        return (C1 * (S(f, 0, 1) - f) - C2 * (S(f, 0, 2) - S(f, 0, -1))) / self.dx

    def forward_step(self, st, psi):
        vz, vx, szz, sxx, sxz = st
        U, pz, pxs, pxv, dt = self.U, self.pz, self.px_s, self.px_v, self.dt
        # stress (el_stress.cu:50-88)
        dvz_dz, dvx_dx = self.Dz_f(vz), self.Dx_f(vx)
        dvx_dz, dvz_dx = self.Dz_b(vx), self.Dx_b(vz)
        m = U & pz
        psi["vzz"] = np.where(m, self.bz * psi["vzz"] + self.az * dvz_dz, psi["vzz"])
        dvz_dz = np.where(m, dvz_dz / self.Kz + psi["vzz"], dvz_dz)
        psi["vxz"] = np.where(m, self.bzh * psi["vxz"] + self.azh * dvx_dz, psi["vxz"])
        dvx_dz = np.where(m, dvx_dz / self.Kzh + psi["vxz"], dvx_dz)
        m = U & pxs
        psi["vxx"] = np.where(m, self.bx * psi["vxx"] + self.ax * dvx_dx, psi["vxx"])
        dvx_dx = np.where(m, dvx_dx / self.Kx + psi["vxx"], dvx_dx)
        psi["vzx"] = np.where(m, self.bxh * psi["vzx"] + self.axh * dvz_dx, psi["vzx"])
        dvz_dx = np.where(m, dvz_dx / self.Kxh + psi["vzx"], dvz_dx)
        l, mu = self.lam, self.mu
        szz = np.where(U, szz + ((l + 2 * mu) * dvz_dz + l * dvx_dx) * dt, szz)
        sxx = np.where(U, sxx + (l * dvz_dz + (l + 2 * mu) * dvx_dx) * dt, sxx)
        sxz = np.where(U, sxz + self.amu * (dvx_dz + dvz_dx) * dt, sxz)
        return [vz, vx, szz, sxx, sxz]

    def velocity_step(self, st, psi):
        vz, vx, szz, sxx, sxz = st
        U, pz, pxv, dt = self.U, self.pz, self.px_v, self.dt
        dszz_dz, dsxz_dx = self.Dz_b(szz), self.Dx_f(sxz)
        dsxz_dz, dsxx_dx = self.Dz_f(sxz), self.Dx_b(sxx)
        m = U & pz
        psi["szz"] = np.where(m, self.bzh * psi["szz"] + self.azh * dszz_dz, psi["szz"])
        dszz_dz = np.where(m, dszz_dz / self.Kzh + psi["szz"], dszz_dz)
        psi["sxzz"] = np.where(m, self.bz * psi["sxzz"] + self.az * dsxz_dz, psi["sxzz"])
        dsxz_dz = np.where(m, dsxz_dz / self.Kz + psi["sxzz"], dsxz_dz)
        m = U & pxv
        psi["sxzx"] = np.where(m, self.bx * psi["sxzx"] + self.ax * dsxz_dx, psi["sxzx"])
        dsxz_dx = np.where(m, dsxz_dx / self.Kx + psi["sxzx"], dsxz_dx)
        psi["sxx"] = np.where(m, self.bxh * psi["sxx"] + self.axh * dsxx_dx, psi["sxx"])
        dsxx_dx = np.where(m, dsxx_dx / self.Kxh + psi["sxx"], dsxx_dx)
        vz = np.where(U, vz + (dszz_dz + dsxz_dx) * self.ba * dt, vz)
        vx = np.where(U, vx + (dsxz_dz + dsxx_dx) * self.bb * dt, vx)
        return [vz, vx, szz, sxx, sxz]

    def run(self, src, recs, stf, obs=None):
        """src (z, x) and recs (zs, xs) are grid indices.  Returns syn [nrec, nt]
        and, when obs is given, (misfit, grads dict, stf grad)."""
        nt, dt = len(stf), self.dt
        NZ, NX = self.NZ, self.NX
        w_src = window_sq(nt, dt, 0.001)
        amp = stf * w_src
        zero = lambda: np.zeros((NZ, NX))
        st = [zero() for _ in range(5)]
        psi = {k: zero() for k in ("vzz", "vzx", "vxz", "vxx", "szz", "sxx", "sxzz", "sxzx")}
        zr, xr = recs
        syn = np.zeros((len(zr), nt))
        states = []  # (v^it, sigma^{it+1}) per step for the gradient
        self.illum = zero()  # sum over steps of vz^2 + vx^2 on the update region
        for it in range(nt - 1):
            vz_it, vx_it = st[0].copy(), st[1].copy()
            st = self.forward_step(st, psi)
            st[2][src] += 1500.0 ** 2 * amp[it] * dt
            st[3][src] += 3.0 * 1500.0 ** 2 * amp[it] * dt
            states.append((vz_it, vx_it, st[2].copy(), st[3].copy(), st[4].copy()))
            st = self.velocity_step(st, psi)
            self.illum += np.where(self.U, st[0] ** 2 + st[1] ** 2, 0.0)
            syn[:, it + 1] = st[2][zr, xr] + 3.0 * st[3][zr, xr]
        if obs is None:
            return syn
        w = window_sq(nt, dt, 0.005)
        res = obs * w - syn * w
        res[:, 0] = 0.0
        J = 0.5 * np.sum(res ** 2)
        res = res * w
        grads = self.adjoint(states, res, src, recs)
        return syn, J, grads

    def adjoint(self, states, res, src, recs):
        NZ, NX, dt = self.NZ, self.NX, self.dt
        nt = res.shape[1]
        zero = lambda: np.zeros((NZ, NX))
        vz, vx, szz, sxx, sxz = (zero() for _ in range(5))
        p = {k: zero() for k in ("vzz", "vzx", "vxz", "vxx", "szz", "sxx", "sxzz", "sxzx")}
        zr, xr = recs
        U, pz, pxs = self.U, self.pz, self.px_s
        l, mu, amu, ba, bb = self.lam, self.mu, self.amu, self.ba, self.bb

        def vel_adj(vz, vx):
            # el_velocity_adj.cu:57-100
            dpsixx_dx = -self.Dx_b(p["vxx"])  # (-c1(f(x+1)-f(x)) + c2(f(x+2)-f(x-1)))/dx
            dszz_dx, dsxx_dx = -self.Dx_b(szz), -self.Dx_b(sxx)
            dpsixz_dz, dsxz_dz = -self.Dz_f(p["vxz"]), -self.Dz_f(sxz)
            vx = np.where(U, vx + (self.ax * dpsixx_dx + l * dszz_dx / self.Kx * dt +
                                   (l + 2 * mu) * dsxx_dx / self.Kx * dt + self.azh * dpsixz_dz +
                                   amu / self.Kzh * dsxz_dz * dt), vx)
            p["sxx"] = np.where(U & pxs, self.bxh * p["sxx"] + bb * vx * dt, p["sxx"])
            p["sxzz"] = np.where(U & pz, self.bz * p["sxzz"] + bb * vx * dt, p["sxzz"])
            dpsizz_dz, dszz_dz, dsxx_dz = -self.Dz_b(p["vzz"]), -self.Dz_b(szz), -self.Dz_b(sxx)
            dpsizx_dx, dsxz_dx = -self.Dx_f(p["vzx"]), -self.Dx_f(sxz)
            vz = np.where(U, vz + (self.az * dpsizz_dz + (l + 2 * mu) * dszz_dz / self.Kz * dt +
                                   l * dsxx_dz / self.Kz * dt + self.axh * dpsizx_dx +
                                   amu / self.Kxh * dsxz_dx * dt), vz)
            p["sxzx"] = np.where(U & pxs, self.bx * p["sxzx"] + ba * vz * dt, p["sxzx"])
            p["szz"] = np.where(U & pz, self.bzh * p["szz"] + ba * vz * dt, p["szz"])
            return vz, vx

        def stress_adj(szz, sxx, sxz):
            # el_stress_adj.cu:160-203 (memory variables not gated)
            dphi_xz_x = -self.Dx_b(p["sxzx"])
            dvz_dx = -self.Dx_b(vz)
            dphi_xz_z = -self.Dz_b(p["sxzz"])
            dvx_dz = -self.Dz_b(vx)
            sxz = np.where(U, sxz + self.ax * dphi_xz_x + dvz_dx / self.Kx * ba * dt +
                           self.az * dphi_xz_z + dvx_dz / self.Kz * bb * dt, sxz)
            p["vzx"] = np.where(U, self.bxh * p["vzx"] + sxz * amu * dt, p["vzx"])
            p["vxz"] = np.where(U, self.bzh * p["vxz"] + sxz * amu * dt, p["vxz"])
            dphi_xx_x = -self.Dx_f(p["sxx"])
            dvx_dx = -self.Dx_f(vx)
            dphi_zz_z = -self.Dz_f(p["szz"])
            dvz_dz = -self.Dz_f(vz)
            sxx = np.where(U, sxx + self.axh * dphi_xx_x + bb * dvx_dx / self.Kxh * dt, sxx)
            szz = np.where(U, szz + self.azh * dphi_zz_z + ba * dvz_dz / self.Kzh * dt, szz)
            p["vxx"] = np.where(U, self.bx * p["vxx"] + l * szz * dt + (l + 2 * mu) * sxx * dt,
                                p["vxx"])
            p["vzz"] = np.where(U, self.bz * p["vzz"] + (l + 2 * mu) * szz * dt + l * sxx * dt,
                                p["vzz"])
            return szz, sxx, sxz

        def inject(szz, sxx, t):
            np.add.at(szz, (zr, xr), res[:, t])
            np.add.at(sxx, (zr, xr), 3.0 * res[:, t])

        gl, gm, gs, ga, gb = (zero() for _ in range(5))
        gstf = np.zeros(nt)
        vz, vx = vel_adj(vz, vx)
        inject(szz, sxx, nt - 1)
        szz, sxx, sxz = stress_adj(szz, sxx, sxz)
        I = self.I
        for it in range(nt - 2, -1, -1):
            gstf[it] = -(szz[src] + 3.0 * sxx[src]) * dt
            fvz, fvx, fszz, fsxx, fsxz = states[it]
            # density (el_velocity.cu:222-243): sigma^{it+1}
            a_ = np.where(I, -vz * (self.Dz_b(fszz) + self.Dx_f(fsxz)) * dt * (-ba ** 2 / 2), 0)
            b_ = np.where(I, -vx * (self.Dz_f(fsxz) + self.Dx_b(fsxx)) * dt * (-bb ** 2 / 2), 0)
            ga += a_
            gb += b_
            # lambda / mu (el_stress.cu:93-124): v^{it}
            dvz_dz, dvx_dx = self.Dz_f(fvz), self.Dx_f(fvx)
            dvx_dz, dvz_dx = self.Dz_b(fvx), self.Dx_b(fvz)
            gl += np.where(I, -(szz + sxx) * (dvz_dz + dvx_dx) * dt * 1e6, 0)
            gm += np.where(I, (-2 * szz * dvz_dz * dt - 2 * sxx * dvx_dx * dt) * 1e6, 0)
            gs += np.where(I, -sxz * (dvx_dz + dvz_dx) * dt, 0)
            vz, vx = vel_adj(vz, vx)
            inject(szz, sxx, it)
            szz, sxx, sxz = stress_adj(szz, sxx, sxz)
        # sprays onto the cells of the averages (el_stress.cu:113-124, el_velocity.cu:238-243)
        nzA, NXm = self.nzA, self.NX
        z1, x1 = nzA - 1 - self.nPml, NXm - 1 - self.nPml
        zi = np.arange(NZ)[:, None] + 0 * np.arange(NX)[None, :]
        xi = 0 * np.arange(NZ)[:, None] + np.arange(NX)[None, :]
        m = self.mu
        S = np.zeros_like(m)
        with np.errstate(divide="ignore", invalid="ignore"):
            S[:NZ - 1, :NX - 1] = (1 / m[:NZ - 1, :NX - 1] + 1 / m[1:, :NX - 1] +
                                   1 / m[:NZ - 1, 1:] + 1 / m[1:, 1:])
            T = np.where(I & (amu != 0), gs * amu / S * 1e6, 0.0)
            inv2 = np.where(m != 0, 1.0 / m ** 2, 0.0)
        gmu = gm.copy()
        gmu += T * inv2
        gmu[1:, :] += np.where(zi[1:, :] <= z1, T[:-1, :], 0) * inv2[1:, :]
        gmu[:, 1:] += T[:, :-1] * inv2[:, 1:]
        gmu[1:, 1:] += np.where((zi[1:, 1:] <= z1) & (xi[1:, 1:] <= x1), T[:-1, :-1], 0) * inv2[1:, 1:]
        gden = ga + gb
        gden[1:, :] += np.where(zi[1:, :] <= z1, ga[:-1, :], 0)
        gden[:, 1:] += gb[:, :-1]
        return dict(lam=gl, mu=gmu, den=gden, stf=gstf)
