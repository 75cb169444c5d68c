"""The oracle's element strain (ShapeBase::Strain, Prism.cpp:826-837), pinned
against the reference's own Save_TensionCompression outputs of the Pardiso
regression runs run07007 and run07008 (tests/golden) and against the closed
forms of homogeneous deformations."""
import numpy as np
import pytest

import replay

# Save_TensionCompression holds exact doubles, the frames 10 significant
# digits: the strain of the saved positions differs from the saved strain by
# the position rounding (about 3e-9 in run07007, 4e-8 in run07008, whose
# prisms are 0.2 um thick ECM layers) plus the last NR increment (below the
# 1e-8 convergence bar).  Any other Gauss point, or F in place of Fe, is off
# by 1.5e-2 or more in every frame.
STRAIN_TOL = {"run07007": 5e-9, "run07008": 6e-8}


@pytest.fixture(scope="module")
def runs():
    return {r: replay.load(r) for r in ("run07007", "run07008")}


def _frame_strain(orc, d, k, gp=5, growth=True):
    pos = d["frames"][k][d["conn"]].reshape(-1, 18)
    ne = len(d["conn"])
    inv = orc.inv3_many(d["Fg"][k]) if growth else np.tile(np.eye(3).reshape(9), (ne, 1))
    return orc.element_strain(pos, d["refpos"], inv, gp)


@pytest.mark.parametrize("run", ["run07007", "run07008"])
def test_strain_matches_save_tension_compression(orc, runs, run):
    d = runs[run]
    for k in range(d["frames"].shape[0]):
        ref = d["strain"][k]
        err = np.abs(_frame_strain(orc, d, k) - ref).max()
        assert err < STRAIN_TOL[run], (k, err)
    # the saved strain is large enough to tell the Gauss points apart
    assert np.abs(d["strain"][-1]).max() > 4e-2


@pytest.mark.parametrize("run", ["run07007", "run07008"])
def test_saved_strain_is_the_last_gauss_point_of_fe(orc, runs, run):
    d = runs[run]
    for k in (1, d["frames"].shape[0] - 1):
        for gp in range(5):
            assert np.abs(_frame_strain(orc, d, k, gp) - d["strain"][k]).max() > 1e-2
        assert np.abs(_frame_strain(orc, d, k, growth=False) - d["strain"][k]).max() > 1e-2


def test_rest_mesh_has_zero_strain(orc, runs):
    d = runs["run07007"]
    assert np.abs(_frame_strain(orc, d, 0)).max() < 1e-15
    assert np.abs(d["strain"][0]).max() < 1e-15


def _prism(seed):
    rng = np.random.default_rng(seed)
    base = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [0, 1, 1]], float)
    return base + 0.1 * rng.standard_normal(base.shape)


def _rotation(seed):
    q, r = np.linalg.qr(np.random.default_rng(seed).standard_normal((3, 3)))
    q = q * np.sign(np.diag(r))
    return q if np.linalg.det(q) > 0 else -q


def _strain_all_points(orc, A, ref, invFg):
    pos = (ref @ A.T).reshape(1, 18)
    return np.stack([orc.element_strain(pos, ref.reshape(1, 18), invFg.reshape(1, 9), g)[0]
                     for g in range(6)])


def _voigt(E):
    return np.array([E[0, 0], E[1, 1], E[2, 2], 2 * E[0, 1], 2 * E[2, 1], 2 * E[0, 2]])


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_homogeneous_deformation_closed_forms(orc, seed):
    ref = _prism(seed)
    I = np.eye(3)
    # stretch: E = (s^2 - 1) / 2 on the diagonal at every Gauss point
    s = np.array([1.1, 0.9, 1.3])
    got = _strain_all_points(orc, np.diag(s), ref, I)
    assert np.allclose(got, np.r_[(s * s - 1) / 2, 0, 0, 0], atol=1e-13, rtol=0)
    # simple shear x += g y: engineering shear 2 E01 = g and E11 = g^2 / 2
    g = 0.3
    got = _strain_all_points(orc, np.array([[1, g, 0], [0, 1, 0], [0, 0, 1.0]]), ref, I)
    assert np.allclose(got, [0, g * g / 2, 0, g, 0, 0], atol=1e-13, rtol=0)
    # shear in z-y and z-x land in slots 4 and 5
    got = _strain_all_points(orc, np.array([[1, 0, 0], [0, 1, 0], [0, g, 1.0]]), ref, I)
    assert np.allclose(got, [0, g * g / 2, 0, 0, g, 0], atol=1e-13, rtol=0)
    got = _strain_all_points(orc, np.array([[1, 0, 0], [0, 1, 0], [g, 0, 1.0]]), ref, I)
    assert np.allclose(got, [g * g / 2, 0, 0, 0, 0, g], atol=1e-13, rtol=0)
    # a general map with growth: E = (Fe^T Fe - I) / 2, Fe = A Fg^-1
    rng = np.random.default_rng(seed + 10)
    A = I + 0.2 * rng.standard_normal((3, 3))
    Fg = I + 0.1 * rng.standard_normal((3, 3))
    Fe = A @ np.linalg.inv(Fg)
    want = _voigt((Fe.T @ Fe - I) / 2)
    got = _strain_all_points(orc, A, ref, np.linalg.inv(Fg))
    assert np.allclose(got, want, atol=1e-12, rtol=0)
    # growth that matches the deformation leaves no strain; a rigid rotation
    # of the current shape leaves the strain unchanged
    assert np.abs(_strain_all_points(orc, Fg, ref, np.linalg.inv(Fg))).max() < 1e-13
    R = _rotation(seed)
    got_r = _strain_all_points(orc, R @ A, ref, np.linalg.inv(Fg))
    assert np.allclose(got_r, want, atol=1e-12, rtol=0)
