"""Pins the oracle's whole NR step against the reference's own Pardiso
regression outputs (tests/sim_pardiso_reg/run07007/expected_out in the
reference repository, extracted into tests/golden/run07007.npz)."""
import numpy as np
import pytest

import replay

STEPS = list(range(10))


@pytest.fixture(scope="module")
def fx():
    return replay.load()


def _problem(orc, s):
    return orc.Problem(s["conn"], s["refpos"], s["invFg"], s["lam"], s["mu"], s["eta"],
                       s["ext_visc"], s["visc_surf"])


@pytest.mark.parametrize("k", STEPS)
def test_replay_step_positions_and_du_log(orc, fx, k):
    s = replay.step_inputs(fx, k, orc)
    out, du, conv = _problem(orc, s).nr_step_dense(s["un"], s["dt"], 1e-8, 20)
    assert conv
    # the reference's own acceptance tolerance is 1e-8 absolute
    # (tests/sim_pardiso_reg/test_SimulationPardiso.py:189); Save_Frame holds
    # 10 significant digits, so the replayed input itself carries ~5e-10.
    assert np.abs(out - s["expected"]).max() < 5e-9
    assert len(du) == len(s["du_log"])
    # printed with 6 significant digits; iterates below 1e-10 are dominated by
    # the direct solver's rounding and are compared only for their magnitude
    big = s["du_log"] > 1e-10
    assert np.allclose(du[big], s["du_log"][big], rtol=1e-5, atol=0)
    assert np.all(du[~big] < 1e-10)


@pytest.mark.parametrize("k", STEPS)
def test_save_force_fixture(orc, fx, k):
    """Save_Force holds the internal nodal forces (elastic + internal viscous,
    ShapeBase::writeInternalForcesTogeAndgv) at the last NR iterate of the
    step.  Evaluated at the saved positions (10 digits, last |du| < 1e-8) the
    forces agree to the input rounding times the stiffness (~1e-5)."""
    s = replay.step_inputs(fx, k, orc)
    K, g, gi = _problem(orc, s).assemble_dense(s["expected"], s["un"], s["dt"])
    F = s["forces"]
    assert np.abs(F).max() > 1.0
    assert np.abs(gi - F).max() < 2e-5


def test_first_step_is_at_rest(orc, fx):
    d = fx
    assert d["du_log"][0][0] < 1e-14
