"""Writes tests/golden/homogeneous_states.npz: nodal forces of one tetrahedron under homogeneous
deformations whose Cauchy stress is known in closed form (the reference ships no golden vectors).

For a homogeneous deformation x = F X + t of a tetra, the nodal force on node a is
f_a = (1/3) sigma . (outward area vector of the face opposite a), sum_a f_a = 0 (the face-area
forces of InteractionFiniteStrainViscoplastic.cl:109-141; tetra.cl:159-168 is the same for
positively oriented tetra).  The stresses are textbook results, not the implementation's formulas:
- Neo-Hookean uniaxial stretch lambda (mu1 = mu, mu2 = 0):
  sigma_11 = 2/3 mu lambda^(-5/3) (lambda^2 - 1) + kappa (lambda - 1),
  sigma_22 = sigma_33 = -1/3 mu lambda^(-5/3) (lambda^2 - 1) + kappa (lambda - 1);
- Mooney-Rivlin hydrostatic stretch s: sigma = kappa (s^3 - 1) I;
- Mooney-Rivlin simple shear gamma (isochoric, B = [[1+g^2, g, 0], [g, 1, 0], [0, 0, 1]]):
  sigma = mu1 dev(B) + mu2 (dev of the second-invariant term), component by component below,
  checked against Rivlin's relations sigma_12 = (mu1 + mu2) gamma,
  sigma_11 - sigma_22 = (mu1 + mu2) gamma^2, sigma_22 - sigma_33 = -mu2 gamma^2.
Each state is rotated and translated rigidly (sigma -> Q sigma Q^T).
Run: ``python tests/golden/make_golden.py`` (numpy only).
"""

import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

REF = np.array([[0.1, -0.2, 0.05], [1.3, 0.1, -0.1], [0.2, 1.1, 0.3], [-0.1, 0.25, 1.2]])
E, NU = 0.01, 0.45


def moduli(ratio):
    mu = E / (2 * (1 + NU))
    return ratio * mu, (1 - ratio) * mu, E / (3 * (1 - 2 * NU))


def nodal_forces(x, sigma):
    """f_a = (1/3) sigma . A_a, A_a = outward area vector of the face opposite node a"""
    f = np.zeros((4, 3))
    c = x.mean(0)
    for a in range(4):
        i, j, k = [b for b in range(4) if b != a]
        A = 0.5 * np.cross(x[j] - x[i], x[k] - x[i])
        if np.dot(A, x[i] - c) < 0:
            A = -A
        f[a] = sigma @ A / 3.0
    return f


def states():
    out = []
    # uniaxial Neo-Hookean
    for lam in (0.9, 1.2):
        mu1, _, kappa = moduli(1.0)
        dev = mu1 * lam ** (-5 / 3) * (lam * lam - 1)
        sig = np.diag([2 / 3 * dev, -1 / 3 * dev, -1 / 3 * dev]) + kappa * (lam - 1) * np.eye(3)
        out.append(("nh_uniaxial", 1.0, np.diag([lam, 1, 1]), sig))
    # hydrostatic Mooney-Rivlin
    for s in (0.95, 1.08):
        _, _, kappa = moduli(0.75)
        out.append(("mr_hydrostatic", 0.75, s * np.eye(3), kappa * (s**3 - 1) * np.eye(3)))
    # simple shear Mooney-Rivlin: B = [[1+g^2, g, 0], [g, 1, 0], [0, 0, 1]], J = 1,
    # sigma = mu1 dev(B) + mu2 (trB B - B^2 - (trB^2 - trB2)/3 I)  ->  components in closed form
    for g in (0.15, 0.3):
        mu1, mu2, _ = moduli(0.6)
        s11 = mu1 * (2 * g * g / 3) + mu2 * (g * g / 3)
        s22 = mu1 * (-g * g / 3) + mu2 * (-2 * g * g / 3)
        s33 = mu1 * (-g * g / 3) + mu2 * (g * g / 3)
        sig = np.array([[s11, (mu1 + mu2) * g, 0], [(mu1 + mu2) * g, s22, 0], [0, 0, s33]])
        # Rivlin's universal relations for simple shear
        assert np.isclose(s11 - s22, (mu1 + mu2) * g * g) and np.isclose(s22 - s33, -mu2 * g * g)
        out.append(("mr_simple_shear", 0.6, np.array([[1, g, 0], [0, 1, 0], [0, 0, 1.0]]), sig))
    return out


def main():
    rng = np.random.default_rng(2024)
    names, ratios, Fs, xs, sigmas, forces = [], [], [], [], [], []
    for name, ratio, F, sig in states():
        # a rigid rotation and translation on top: sigma -> Q sigma Q^T (objectivity)
        ax = rng.normal(size=3)
        ax /= np.linalg.norm(ax)
        th = rng.uniform(0.2, 1.0)
        K = np.array([[0, -ax[2], ax[1]], [ax[2], 0, -ax[0]], [-ax[1], ax[0], 0]])
        Q = np.eye(3) + np.sin(th) * K + (1 - np.cos(th)) * K @ K
        x = REF @ (Q @ F).T + rng.normal(size=3)
        s = Q @ sig @ Q.T
        names.append(name)
        ratios.append(ratio)
        Fs.append(Q @ F)
        xs.append(x)
        sigmas.append(s)
        forces.append(nodal_forces(x, s))
    np.savez(os.path.join(HERE, "homogeneous_states.npz"), names=np.array(names), ratio=np.array(ratios),
             E=E, nu=NU, X=REF, F=np.array(Fs), x=np.array(xs), sigma=np.array(sigmas), forces=np.array(forces))


if __name__ == "__main__":
    main()
