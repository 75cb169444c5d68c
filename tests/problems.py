"""Synthetic test problems shared by the oracle and GPU tests."""
import numpy as np

from tissueorigami_amd import tissue


def c1_config(seed=0):
    """Plain-FP64 small rectangle: 10 x 5 um, side 1 um, 2 prism layers of 1 um
    (198 nodes, 200 prisms), E = 1000 Pa, nu = 0.3, no internal viscosity,
    in-plane growth 0.01/h on the basal layer, dt = 10 s.  External drag on
    the apical and basal surfaces keeps the system non-singular, so no DOF is
    fixed."""
    m = tissue.rectangle(10.0, 5.0, 1.0, 2.0, 2)
    ne = len(m["conn"])
    nn = len(m["nodes"])
    E = np.full(ne, 1000.0)
    nu = np.full(ne, 0.3)
    eta = np.zeros(ne)
    ext = np.zeros((nn, 3))
    ext[m["node_placement"] == 1] = 16000.0
    ext[m["node_placement"] == 0] = 10.0
    rate = np.where(m["layer"] == 0, 0.01, 0.0)
    place, spans = tissue.element_placement(m["conn"], m["node_placement"])
    etype = tissue.element_type(m["conn"], m["node_type"])
    return dict(mesh=m, E=E, nu=nu, eta=eta, ext_visc=ext, rate=rate, dt=10.0, n_steps=50,
                place=place, spans=spans, etype=etype)


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
