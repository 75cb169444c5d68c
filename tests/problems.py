"""Synthetic test problems shared by the oracle and GPU tests."""
import numpy as np

from tissueorigami_amd import tissue, workloads
from tissueorigami_amd.workloads import c1_config, c5_config  # noqa: F401


def perturbed_problem(seed=0, visc=True, fixed=False, lx=4.0, ly=3.0, layers=2):
    """A small, irregular, pre-stressed tissue: random node perturbations,
    anisotropic growth and layer-wise materials, so that no entry of the
    Jacobian is accidentally near the 1e-13 threshold."""
    rng = np.random.default_rng(seed)
    m = tissue.rectangle(lx, ly, 1.0, 2.0, layers)
    nn = len(m["nodes"]); ne = len(m["conn"])
    X = m["nodes"] + rng.normal(0, 0.05, size=(nn, 3))
    refpos = X[m["conn"]].reshape(ne, 18)
    Fg = np.eye(3)[None] + 0.05 * rng.normal(size=(ne, 3, 3))
    invFg = np.linalg.inv(Fg).reshape(ne, 9)
    E = rng.uniform(100, 2000, ne)
    nu = rng.uniform(0.1, 0.45, ne)
    eta = rng.uniform(5, 20, ne) if visc else np.zeros(ne)
    ext = rng.uniform(5, 50, (nn, 3))
    un = (X + rng.normal(0, 0.02, size=X.shape)).reshape(-1)
    uk = un + rng.normal(0, 0.02, size=un.shape)
    area = rng.uniform(0.1, 1.0, nn)
    fx = None
    if fixed:
        fx = np.zeros((nn, 3), np.uint8)
        fx[0, :] = 1
        fx[5, 2] = 1
        fx[nn - 1, 1] = 1
    return dict(mesh=m, refpos=refpos, invFg=invFg, E=E, nu=nu, eta=eta, ext_visc=ext,
                un=un, uk=uk, area=area, fixed=fx, dt=7.0)


def growth_invFg(orc, Fg):
    return orc.inv3_many(Fg.reshape(-1, 9))


def growth_steps(orc, rate_per_hour, dt, n_steps):
    """growth_steps of the package with the oracle's 3x3 inverse."""
    return workloads.growth_steps(rate_per_hour, dt, n_steps, inv3=orc.inv3_many)
