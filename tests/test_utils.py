"""Host mirror of the reference's Julia helpers (src/FWI.jl, src/Utils.jl),
pinned by hand-built cases, adjoint identities and finite differences.  CPU
only."""
import numpy as np
import pytest

from fwiflow import fwi as F
from fwiflow.utils import (moduli_to_velocity_grad, npad, padding, padding_adjoint,
                           velocity_to_moduli)


def test_npad_rule():
    # FWI.jl:12: nPad = 32 - mod(nz + 2 nPml, 32); an aligned size still pads 32
    assert npad(134, 32) == 32 - (198 % 32)
    assert npad(64, 32) == 32
    assert (npad(500, 40) + 500 + 80) % 32 == 0


def test_padding_is_symmetric_with_edge_repeat():
    """tf.pad(..., "SYMMETRIC") repeats the edge value: [a b c] padded by 2 on
    each side is [b a a b c c b]."""
    a = np.array([[1.0, 2.0, 3.0]])
    p = padding(a, 2, 0)
    np.testing.assert_array_equal(p[2], [2, 1, 1, 2, 3, 3, 2])
    np.testing.assert_array_equal(p[:, 2], [1, 1, 1, 1, 1])
    # extra nPad rows go to the bottom only
    assert padding(np.zeros((4, 5)), 3, 2).shape == (4 + 3 + 3 + 2, 5 + 6)


def test_padding_adjoint_identity():
    rng = np.random.default_rng(0)
    nz, nx, P, pad = 7, 9, 4, 3
    x = rng.standard_normal((nz, nx))
    y = rng.standard_normal((nz + 2 * P + pad, nx + 2 * P))
    lhs = float(np.sum(padding(x, P, pad) * y))
    rhs = float(np.sum(x * padding_adjoint(y, nz, nx, P, pad)))
    assert lhs == pytest.approx(rhs, rel=1e-12)


def test_moduli_chain_rule_matches_finite_differences():
    """d/d(cp, cs, rho) of <gL, lambda> + <gM, mu> + <gD, rho>."""
    rng = np.random.default_rng(1)
    cp = rng.uniform(2000, 4000, (3, 4))
    cs = cp / np.sqrt(3.0)
    rho = rng.uniform(1800, 2600, (3, 4))
    gL, gM, gD = (rng.standard_normal((3, 4)) for _ in range(3))

    def J(cp_, cs_, rho_):
        lam, mu = velocity_to_moduli(cp_, cs_, rho_)
        return float(np.sum(gL * lam) + np.sum(gM * mu) + np.sum(gD * rho_))

    g_cp, g_cs, g_rho = moduli_to_velocity_grad(cp, cs, rho, gL, gM, gD)
    for k, (g, which) in enumerate(((g_cp, 0), (g_cs, 1), (g_rho, 2))):
        d = rng.standard_normal((3, 4))
        h = 1e-3
        args = [cp, cs, rho]
        up = [a + (h * d if i == which else 0) for i, a in enumerate(args)]
        dn = [a - (h * d if i == which else 0) for i, a in enumerate(args)]
        fd = (J(*up) - J(*dn)) / (2 * h)
        assert float(np.sum(g * d)) == pytest.approx(fd, rel=1e-8), k


def test_velocity_to_moduli_units():
    # Utils.jl:221-227: lambda, mu in MPa
    lam, mu = velocity_to_moduli(np.array([3000.0]), np.array([1500.0]), np.array([2000.0]))
    assert lam[0] == pytest.approx((9e6 - 2 * 2.25e6) * 2000 / 1e6)
    assert mu[0] == pytest.approx(2.25e6 * 2000 / 1e6)


def test_fwi_geometry_mask_and_padding(tmp_path):
    """FWI.jl:38-60: padded sizes, mask = 1 on the interior except its top 10
    rows, the parameter and survey files written into the workspace."""
    src_x = np.array([4, 20])
    rec_x = np.arange(3, 30)
    fwi = F.make_fwi(40, 32, 10.0, 10.0, 100, 1e-3, 2 * np.ones(2, int), src_x,
                     2 * np.ones(rec_x.size, int), rec_x, WORKSPACE=str(tmp_path), nPml=8)
    assert (fwi.nz_pad, fwi.nx_pad) == (40 + 16 + npad(40, 8), 32 + 16)
    m = fwi.mask
    assert m.shape == (fwi.nz_pad, fwi.nx_pad)
    assert m[8 + 10:8 + 40, 8:8 + 32].min() == 1.0
    assert m[8:8 + 10].max() == 0.0 and m[:, :8].max() == 0.0 and m[8 + 40:].max() == 0.0
    np.testing.assert_array_equal(fwi.mask_neg, 1.0 - m)
    assert (tmp_path / "para_file.json").exists() and (tmp_path / "survey_file.json").exists()
    # try_pad: padded input passes through, (nz, nx) is padded; another size
    # is padded unresized (FWI.jl:193-204 pads the input, not its resize),
    # and the operator calls refuse the result
    a = np.ones((40, 32))
    assert F.try_pad(fwi, a).shape == m.shape
    p = np.zeros(m.shape)
    assert F.try_pad(fwi, p) is p
    q = F.try_pad(fwi, np.arange(20 * 16, dtype=float).reshape(20, 16))
    assert q.shape == (20 + 16 + 8, 16 + 16)
    np.testing.assert_array_equal(q[8:28, 8:24], np.arange(20 * 16).reshape(20, 16))
    with pytest.raises(ValueError, match="after padding"):
        F._on_grid(fwi, q)
    # a 1-row source time function is repeated for every shot (FWI.jl:118-120)
    assert F._stf_rows(fwi, np.ones(100)).shape == (2, 100)
