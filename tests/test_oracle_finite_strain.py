"""Pins the oracle's finite-strain laws, which no reference test covers
(MaterialNeohookean tests are disabled in the reference:
test/test_model/test_solid_mechanics_model/test_materials/test_finite_def_materials.cc:48-82;
the neo-Hookean + damage composition has no counterpart at all).

Each check is a property the mathematics fixes and that a plausible slip in
the restatement (a dropped term, a transposed operand, a wrong sign, a wrong
Voigt index) breaks:
  * zero strain -> zero stress;
  * small strain -> the linear elastic law (lambda tr(eps) I + 2 mu eps);
  * S = 2 dW/dC (energy / stress consistency, by central differences);
  * objectivity: S(R F) = S(F) for a rotation R;
  * the finite-strain residual (B_L0 + B_L2)^T S equals the first
    Piola-Kirchhoff contraction F S grad N, for a non-symmetric gradient;
  * damage: d = 0 -> undamaged law, the Marigo branch, d monotone and <= 1.
"""
import numpy as np
import pytest

from oracle import oracle as orc

E, NU = 1e7, 0.45


def lame(E=E, nu=NU):
    return nu * E / ((1 + nu) * (1 - 2 * nu)), E / (2 * (1 + nu))


@pytest.mark.parametrize("d", [2, 3])
def test_zero_strain_zero_stress(d):
    S, _ = orc.point_stress(orc.NEOHOOKEAN, d, E, NU, np.zeros((d, d)))
    assert np.abs(S).max() == 0.0


@pytest.mark.parametrize("d", [2, 3])
def test_small_strain_limit_is_linear_elastic(d):
    rng = np.random.default_rng(1)
    G = rng.normal(size=(d, d))
    lam, mu = lame()
    for h in (1e-4, 1e-5, 1e-6):
        S, _ = orc.point_stress(orc.NEOHOOKEAN, d, E, NU, h * G)
        eps = 0.5 * (G + G.T)
        lin = lam * np.trace(eps) * np.eye(d) + 2 * mu * eps
        # the remainder is O(h^2): relative error O(h)
        assert np.abs(S / h - lin).max() / np.abs(lin).max() < 20 * h


@pytest.mark.parametrize("d", [2, 3])
def test_stress_is_energy_derivative(d):
    """S_ij = 2 dW/dC_ij, i.e. dW/dG_kl = (F S)_kl since C = F^T F, F = I + G."""
    rng = np.random.default_rng(2)
    G = 0.05 * rng.normal(size=(d, d))
    S, _ = orc.point_stress(orc.NEOHOOKEAN, d, E, NU, G)
    P = (np.eye(d) + G) @ S
    h = 1e-6
    num = np.zeros((d, d))
    for k in range(d):
        for l in range(d):
            Gp, Gm = G.copy(), G.copy()
            Gp[k, l] += h
            Gm[k, l] -= h
            num[k, l] = (orc.point_energy(orc.NEOHOOKEAN, d, E, NU, Gp) -
                         orc.point_energy(orc.NEOHOOKEAN, d, E, NU, Gm)) / (2 * h)
    assert np.abs(num - P).max() / np.abs(P).max() < 1e-6


def test_objectivity_large_rotation():
    """S depends on C = F^T F only: a rigid rotation of the deformed body leaves
    it unchanged (test_finite_deformation.cc:30-110 pattern)."""
    rng = np.random.default_rng(3)
    G = 0.2 * rng.normal(size=(3, 3))
    F = np.eye(3) + G
    S0, _ = orc.point_stress(orc.NEOHOOKEAN, 3, E, NU, G)
    for ang in (0.3, 1.2, 2.9):
        ax = rng.normal(size=3)
        ax /= np.linalg.norm(ax)
        K = np.array([[0, -ax[2], ax[1]], [ax[2], 0, -ax[0]], [-ax[1], ax[0], 0]])
        R = np.eye(3) + np.sin(ang) * K + (1 - np.cos(ang)) * K @ K
        S1, _ = orc.point_stress(orc.NEOHOOKEAN, 3, E, NU, R @ F - np.eye(3))
        assert np.abs(S1 - S0).max() / np.abs(S0).max() < 1e-12


@pytest.mark.parametrize("d,n", [(2, 4), (3, 8)])
def test_finite_residual_is_first_piola_contraction(d, n):
    """(B_L0 + B_L2(G))^T S_voigt == F S grad(N_a) for every node a, with a
    non-symmetric G (a transposed G in B_L2 or a wrong Voigt slot fails)."""
    rng = np.random.default_rng(4)
    B = rng.normal(size=(d, n))
    G = 0.3 * rng.normal(size=(d, d))
    A = rng.normal(size=(d, d))
    S = A + A.T
    r = orc.point_finite_residual(d, B, G, S)
    P = (np.eye(d) + G) @ S
    expect = (P @ B).T
    assert np.abs(r - expect).max() / np.abs(expect).max() < 1e-13


def test_damage_law_special_cases():
    rng = np.random.default_rng(5)
    G = 0.03 * rng.normal(size=(3, 3))
    W = orc.point_energy(orc.NEOHOOKEAN, 3, E, NU, G)
    S_nh, _ = orc.point_stress(orc.NEOHOOKEAN, 3, E, NU, G)
    # threshold above the energy: no damage, undamaged law
    S, dmg = orc.point_stress(orc.NEOHOOKEAN_DAMAGE, 3, E, NU, G, Sd=1e3, Yd=2 * W)
    assert dmg == 0.0 and np.array_equal(S, S_nh)
    # Marigo branch: d = (Y - Yd) / Sd, S = (1 - d) S_nh
    Sd, Yd = 10 * W, 0.5 * W
    S, dmg = orc.point_stress(orc.NEOHOOKEAN_DAMAGE, 3, E, NU, G, Sd=Sd, Yd=Yd)
    assert abs(dmg - (W - Yd) / Sd) < 1e-15
    assert np.abs(S - (1 - dmg) * S_nh).max() <= 1e-15 * np.abs(S_nh).max()
    # history: an existing larger damage is kept (Fd <= 0), and d <= 1
    S2, d2 = orc.point_stress(orc.NEOHOOKEAN_DAMAGE, 3, E, NU, G, Sd=Sd, Yd=Yd, damage=0.9)
    assert d2 == 0.9
    S3, d3 = orc.point_stress(orc.NEOHOOKEAN_DAMAGE, 3, E, NU, G, Sd=1e-3 * W, Yd=0.0)
    assert d3 == 1.0 and np.abs(S3).max() == 0.0
