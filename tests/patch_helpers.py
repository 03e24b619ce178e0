"""Shared set-up of the reference's linear patch test
(test/test_model/patch_tests/patch_test_linear_fixture.hh,
patch_test_linear_solid_mechanics_fixture.hh) for the oracle and GPU tests."""
import os

import numpy as np

from akantu_amd import mesh as akm

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

# patch_test_linear_fixture.hh:167-169
ALPHA = np.array([[0.01, 0.02, 0.03, 0.04],
                  [0.05, 0.06, 0.07, 0.08],
                  [0.09, 0.10, 0.11, 0.12]])

# data/material_check_stress_plane_strain.dat
STEEL = dict(E=2.1e11, nu=0.3, rho=7800.0, plane_stress=0)

PATCH_TYPES = ["triangle_3", "quadrangle_4", "tetrahedron_4", "hexahedron_8"]


def patch_mesh(name):
    return akm.read_msh(os.path.join(GOLDEN, "patch_%s.msh" % name))


def linear_field(X):
    d = X.shape[1]
    return ALPHA[:d, 0][None, :] + X @ ALPHA[:d, 1:d + 1].T


def prescribed_gradient(d):
    return ALPHA[:d, 1:d + 1]


def plane_strain_stress(grad, E=STEEL["E"], nu=STEEL["nu"]):
    d = grad.shape[0]
    strain = 0.5 * (grad + grad.T)
    lam = nu * E / ((1 + nu) * (1 - 2 * nu))
    mu = E / (2 * (1 + nu))
    return lam * np.trace(strain) * np.eye(d) + 2 * mu * strain


def boundary_traction_forces(mesh, sigma):
    """Nodal forces of the constant traction sigma.n on the boundary facets
    (the BC::Neumann::FromHigherDim(sigma) of the force check,
    patch_test_linear_solid_mechanics_fixture.hh:63-74), integrated exactly:
    segments / triangles: N_a integrates to |facet| / n; quadrangle facets with
    2 x 2 Gauss points."""
    facets, owner = akm.boundary_facets(mesh, with_owner=True)
    X = mesh.nodes
    d = mesh.dim
    f = np.zeros_like(X)
    cent = X[mesh.conn[owner].astype(np.int64)].mean(axis=1)
    for fac, c in zip(facets, cent):
        P = X[fac]
        mid = P.mean(axis=0)
        if d == 2:
            t = P[1] - P[0]
            nvec = np.array([t[1], -t[0]])  # |n| = length
            if np.dot(nvec, mid - c) < 0:
                nvec = -nvec
            for a in range(2):
                f[fac[a]] += sigma @ nvec / 2.0
        elif len(fac) == 3:
            nvec = 0.5 * np.cross(P[1] - P[0], P[2] - P[0])
            if np.dot(nvec, mid - c) < 0:
                nvec = -nvec
            for a in range(3):
                f[fac[a]] += sigma @ nvec / 3.0
        else:
            g = 1.0 / np.sqrt(3.0)
            for xi in (-g, g):
                for eta in (-g, g):
                    N = 0.25 * np.array([(1 - xi) * (1 - eta), (1 + xi) * (1 - eta),
                                         (1 + xi) * (1 + eta), (1 - xi) * (1 + eta)])
                    dxi = 0.25 * np.array([-(1 - eta), (1 - eta), (1 + eta), -(1 + eta)])
                    deta = 0.25 * np.array([-(1 - xi), -(1 + xi), (1 + xi), (1 - xi)])
                    t1, t2 = dxi @ P, deta @ P
                    nvec = np.cross(t1, t2)
                    if np.dot(nvec, mid - c) < 0:
                        nvec = -nvec
                    for a in range(4):
                        f[fac[a]] += N[a] * (sigma @ nvec)
    return f
