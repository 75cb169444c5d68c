"""Synthetic workloads of the benchmark and the profiling driver.

c1: rectangle 10 x 5 um, side 1 um, 2 layers (198 nodes, 200 prisms), E = 1000 Pa,
    nu = 0.3, no internal viscosity, growth 0.01/h on the basal layer, dt = 10 s.
c2: ellipse 100 x 70 um, side 2 um, 4 layers over 20 um; E = 1000 Pa apical
    layer, 100 Pa below; nu = 0.3; eta = 10 Pa s; growth from a seeded Gaussian
    random field (mean 0.02/h, sd 0.01/h, correlation length 15 um); dt = 60 s.
c3: ellipse 200 x 140 um, side 1 um, 8 layers over 30 um (1.62 M prisms,
    0.91 M nodes), same material and growth model as c2.
External drag 16000 on apical and 10 on basal surface nodes in all of them.
"""
import numpy as np

from . import tissue

CONFIGS = {
    # (kind, geometry, layers, thickness, dt, E_apical, E_other, nu, eta, growth)
    "c1": dict(kind="rectangle", lx=10.0, ly=5.0, side=1.0, layers=2, thickness=2.0, dt=10.0,
               E_ap=1000.0, E_other=1000.0, nu=0.3, eta=0.0, grf=False,
               desc="rectangle 10x5 um, side 1 um, 2 layers"),
    "c2": dict(kind="ellipse", a=100.0, b=70.0, side=2.0, layers=4, thickness=20.0, dt=60.0,
               E_ap=1000.0, E_other=100.0, nu=0.3, eta=10.0, grf=True,
               desc="ellipse 100x70 um, side 2 um, 4 layers over 20 um"),
    "c3": dict(kind="ellipse", a=200.0, b=140.0, side=1.0, layers=8, thickness=30.0, dt=60.0,
               E_ap=1000.0, E_other=100.0, nu=0.3, eta=10.0, grf=True,
               desc="ellipse 200x140 um, side 1 um, 8 layers over 30 um"),
}


def make_workload(name, seed=2024):
    c = CONFIGS[name]
    if c["kind"] == "rectangle":
        m = tissue.rectangle(c["lx"], c["ly"], c["side"], c["thickness"], c["layers"])
    else:
        m = tissue.ellipse(c["a"], c["b"], c["side"], c["thickness"], c["layers"])
    ne, nn = len(m["conn"]), len(m["nodes"])
    top = m["layer"] == m["n_layers"] - 1
    E = np.where(top, c["E_ap"], c["E_other"])
    nu = np.full(ne, c["nu"])
    eta = np.full(ne, c["eta"])
    ext = np.zeros((nn, 3))
    ext[m["node_placement"] == 1] = 16000.0
    ext[m["node_placement"] == 0] = 10.0
    if c["grf"]:
        cent = m["nodes"][m["conn"]].mean(1)[:, :2]
        rate = tissue.gaussian_random_field(cent, 0.02, 0.01, 15.0, seed)
    else:
        rate = np.where(m["layer"] == 0, 0.01, 0.0)
    place, spans = tissue.element_placement(m["conn"], m["node_placement"])
    etype = tissue.element_type(m["conn"], m["node_type"])
    inc = tissue.growth_increment(rate, c["dt"])
    rates = np.zeros((ne, 3))
    rates[:, 0] = rate / 3600.0
    rates[:, 1] = rate / 3600.0
    return dict(mesh=m, E=E, nu=nu, eta=eta, ext=ext, inc=inc, rates=rates, dt=c["dt"],
                place=place, spans=spans, etype=etype, cfg=c, seed=seed)


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


def c5_config(seed=5):
    """Adjoint configuration: the C1 mesh with an in-plane growth rate (1/h)
    given by 64 Gaussian bumps (sigma 12 um) on a seeded jittered 8 x 8 grid,
    coefficients ~ N(0.02, 0.005^2); 20 steps of 10 s; loss = sum over apical
    nodes of |x - target|^2 with a seeded target surface."""
    c = c1_config()
    m = c["mesh"]
    rng = np.random.default_rng(seed)
    gx, gy = np.meshgrid((np.arange(8) + 0.5) * 10.0 / 8, (np.arange(8) + 0.5) * 5.0 / 8)
    centres = np.stack([gx.ravel(), gy.ravel()], 1) + rng.uniform(-0.3, 0.3, (64, 2))
    coef = rng.normal(0.02, 0.005, 64)
    cent = m["nodes"][m["conn"]].mean(1)[:, :2]
    d2 = ((cent[:, None, :] - centres[None, :, :]) ** 2).sum(-1)
    phi = np.exp(-d2 / (2 * 12.0 ** 2))          # n_prisms x 64
    apical = np.where(m["node_placement"] == 1)[0]
    target = m["nodes"][apical] + rng.normal(0, 0.05, (len(apical), 3))
    c.update(phi=phi, coef=coef, apical=apical, target=target, n_steps=20, dt=10.0)
    return c


def growth_steps(rate_per_hour, dt, n_steps, inv3=None):
    """InvFg of every step for a constant in-plane rate, accumulated as
    ShapeBase::growShapeByFg does (F_g <- dF_g F_g, then its inverse), and its
    derivative with respect to the rate in 1/s: d/dr diag(g^-n, g^-n, 1) with
    g = exp(r dt).  inv3 inverts a stack of 3x3 matrices (numpy by default)."""
    inv3 = inv3 or (lambda F: np.linalg.inv(F).reshape(-1, 9))
    ne = len(rate_per_hour)
    inc = tissue.growth_increment(rate_per_hour, dt)
    Fg = np.tile(np.eye(3), (ne, 1, 1))
    inv = np.empty((n_steps, ne, 9))
    dinv = np.zeros((n_steps, ne, 9))
    r = np.asarray(rate_per_hour) / 3600.0
    for s in range(n_steps):
        Fg = inc @ Fg
        inv[s] = inv3(Fg)
        d = -(s + 1) * dt * np.exp(-(s + 1) * r * dt)
        dinv[s, :, 0] = d
        dinv[s, :, 4] = d
    return inv, dinv


def contact_workload(name="c2", gap=0.6, seed=2024):
    """Two copies of a workload's ellipse facing each other apical to apical
    (tissue.facing_slabs), `gap` um apart, with their basal surfaces clamped
    (every DOF of basal nodes fixed), the workload's materials and drag, and
    growth from the workload's random field over both copies.  Growth swells
    the apical layers into each other; node-node packing
    (Simulation::detectPacingNodes, Simulation.cpp:7997-8078) pushes them
    apart."""
    c = CONFIGS[name]
    m1 = tissue.ellipse(c["a"], c["b"], c["side"], c["thickness"], c["layers"])
    m = tissue.facing_slabs(m1, gap)
    ne, nn = len(m["conn"]), len(m["nodes"])
    top = m["layer"] == m1["n_layers"] - 1
    E = np.where(top, c["E_ap"], c["E_other"])
    ext = np.zeros((nn, 3))
    ext[m["node_placement"] == 1] = 16000.0
    ext[m["node_placement"] == 0] = 10.0
    fixed = np.zeros((nn, 3), np.uint8)
    fixed[m["node_placement"] == 0, :] = 1
    cent = m["nodes"][m["conn"]].mean(1)[:, :2]
    rate = tissue.gaussian_random_field(cent, 0.02, 0.01, 15.0, seed)
    place, spans = tissue.element_placement(m["conn"], m["node_placement"])
    etype = tissue.element_type(m["conn"], m["node_type"])
    rates = np.zeros((ne, 3))
    rates[:, 0] = rates[:, 1] = rate / 3600.0
    return dict(mesh=m, E=E, nu=np.full(ne, c["nu"]), eta=np.full(ne, c["eta"]), ext=ext,
                fixed=fixed, rates=rates, dt=c["dt"], place=place, spans=spans, etype=etype,
                cfg=c, seed=seed, gap=gap,
                desc=f"two {name} ellipses facing apical to apical, {gap} um apart, basal "
                     "surfaces clamped")
