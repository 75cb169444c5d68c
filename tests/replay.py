"""Per-step replay of the reference's Pardiso regression run07007.

Builds, from the committed fixture (tests/golden/run07007.npz), the inputs of
the NR step that produced frame k+1 from frame k:
  * u_n: frame k positions after the rigid pre-step alignment
    Simulation::alignTissueDVToXPositive (Simulation.cpp:8230-8272, with the
    ShapeBase rotation helpers at ShapeBase.cpp calculateRotationAngleSinCos /
    calculateRotationAxis / constructRotationMatrix / rotateVectorByRotationMatrix);
    run07007's tips are dorsal 0, ventral 9 (printed in expected_out/tmp);
  * F_g, effective E (getYoungModulus, ShapeBase.cpp:179-181), internal and
    external viscosities of frame k+1;
  * Poisson ratio 0.29, and 0 for ECM elements (ShapeBase::setECMMimicing,
    ShapeBase.cpp:983-989);
  * node viscosity surfaces from the aligned frame-k positions;
  * no fixed DOF (external viscosity is non-zero in x, y and z, so
    checkForZeroExternalViscosity, Simulation.cpp:4126-4200, fixes nothing) and
    no binding (Save_NodeBinding is empty at every frame).
"""
import os

import numpy as np

from tissueorigami_amd import tissue

HERE = os.path.dirname(os.path.abspath(__file__))
DORSAL, VENTRAL = 0, 9


def load():
    return np.load(os.path.join(HERE, "golden", "run07007.npz"))


def _normalise(v):
    mag2 = float(v @ v)
    if abs(mag2) > 1e-14 and abs(mag2 - 1.0) > 1e-14:
        return v / mag2 ** 0.5
    return v


def align_dv_to_x(P):
    P = P.copy()
    u = _normalise(P[VENTRAL] - P[DORSAL])
    v = np.array([1.0, 0.0, 0.0])
    c = float(u @ v)
    if c > 1.0:
        c, s = 1.0, 0.0
    elif c < -1.0:
        c, s = -1.0, 0.0
    else:
        s = np.sin(np.arccos(c))
    if c > -0.99998:
        ax = _normalise(np.cross(u, v))
    else:
        ax = np.array([1.0, 0.0, 0.0])
    R = np.array([
        [c + ax[0] * ax[0] * (1 - c), ax[0] * ax[1] * (1 - c) - ax[2] * s, ax[0] * ax[2] * (1 - c) + ax[1] * s],
        [ax[1] * ax[0] * (1 - c) + ax[2] * s, c + ax[1] * ax[1] * (1 - c), ax[1] * ax[2] * (1 - c) - ax[0] * s],
        [ax[2] * ax[0] * (1 - c) - ax[1] * s, ax[2] * ax[1] * (1 - c) + ax[0] * s, c + ax[2] * ax[2] * (1 - c)]])
    rel = P[1:] - P[0]
    P[1:] = rel @ R.T + P[0]
    return P


def step_inputs(d, k, orc):
    """Inputs of the step frame k -> frame k+1 (k = 0 .. n_frames-2)."""
    conn = d["conn"]
    ne = len(conn)
    nn = d["frames"].shape[1]
    place, spans = tissue.element_placement(conn, d["node_info"][:, 0])
    etype = tissue.element_type(conn, d["node_info"][:, 1])
    nu = np.full(ne, float(d["nu"]))
    nu[d["ecm"]] = 0.0
    un = align_dv_to_x(d["frames"][k]).reshape(-1)
    invFg = orc.inv3_many(d["Fg"][k + 1])
    lam, mu = orc.lame(d["E"][k + 1], nu)
    A = orc.visc_surfaces(nn, conn, un, place, etype, spans)
    return dict(conn=conn, refpos=d["refpos"], invFg=invFg, lam=lam, mu=mu,
                E=d["E"][k + 1], nu=nu, eta=d["eta"][k + 1], ext_visc=d["ext_visc"][k + 1],
                visc_surf=A, un=un, dt=float(d["dt"]), expected=d["frames"][k + 1].reshape(-1),
                forces=d["forces"][k + 1].reshape(-1),
                du_log=d["du_log"][k + 1][~np.isnan(d["du_log"][k + 1])])
