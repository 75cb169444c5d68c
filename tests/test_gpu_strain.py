"""The device's per-prism strain output (tf_nr_element_strain: ShapeBase::Strain,
Prism.cpp:826-837, the rows of Save_TensionCompression) against the oracle and
against the reference's own saved strain of run07007 and run07008.
Every test here needs an MI355X."""
import numpy as np
import pytest

import problems
import replay

pytestmark = pytest.mark.gpu

# the oracle's own bounds against the same fixtures (positions saved with 10
# significant digits)
STRAIN_TOL = {"run07007": 5e-9, "run07008": 6e-8}


@pytest.fixture(scope="module")
def nr():
    from tissueorigami_amd import nr as _nr
    _nr.load()
    return _nr


def _solver(nr, p):
    m = p["mesh"]
    s = nr.NewtonRaphsonSolver(len(m["nodes"]), m["conn"], p["refpos"])
    s.set_material(p["E"], p["nu"], p["eta"], p["ext_visc"], p["fixed"])
    s.set_growth(p["invFg"])
    s.set_visc_surface(p["area"])
    return s


def _oracle_strain(orc, p, u):
    conn = p["mesh"]["conn"]
    return orc.element_strain(u.reshape(-1, 3)[conn].reshape(-1, 18), p["refpos"], p["invFg"])


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_assembly_strain_matches_oracle(orc, nr, precision):
    p = problems.perturbed_problem(seed=3, visc=True)
    s = _solver(nr, p)
    with pytest.raises(nr.TFError):
        s.element_strain()
    s.assemble(p["uk"], p["un"], p["dt"], precision)
    got = s.element_strain()
    want = _oracle_strain(orc, p, p["uk"])
    assert np.abs(want).max() > 1e-2
    # FP64 in both precisions (the mixed kernel's tangents are FP32, its
    # deformation gradient is not); fused multiply-adds move the last bits
    assert np.abs(got - want).max() < 1e-14
    s.close()


def test_step_strain_is_that_of_the_last_iterate(orc, nr):
    """The reference's strain comes from the last force evaluation of the
    step, before the final increment (Simulation.cpp:6640-6720): with one
    NR iteration that is the start of the step."""
    p = problems.perturbed_problem(seed=4, visc=True)
    s = _solver(nr, p)
    u1, st = s.step(p["un"], p["dt"], 1e-8, 0, "fp64", 1e-10, allow_not_converged=True)
    assert not st["converged"]
    got = s.element_strain()
    assert np.abs(got - _oracle_strain(orc, p, p["un"])).max() < 1e-14
    assert np.abs(got - _oracle_strain(orc, p, u1)).max() > 1e-6
    s.close()


def test_ablated_prisms_give_zero_strain(orc, nr):
    p = problems.perturbed_problem(seed=5, visc=True)
    s = _solver(nr, p)
    ne = len(p["mesh"]["conn"])
    abl = np.zeros(ne, np.uint8)
    abl[::7] = 1
    s.set_ablation(abl)
    s.assemble(p["uk"], p["un"], p["dt"], "mixed")
    got = s.element_strain()
    assert np.all(got[abl == 1] == 0.0)
    want = _oracle_strain(orc, p, p["uk"])
    assert np.abs(got[abl == 0] - want[abl == 0]).max() < 1e-14
    s.close()


@pytest.mark.parametrize("precision,rtol", [("fp64", 1e-13), ("mixed", 1e-3)])
@pytest.mark.parametrize("run,k", [("run07007", 0), ("run07007", 6), ("run07008", -1),
                                   ("run07008", 5)])
def test_replayed_step_strain_matches_save_tension_compression(orc, nr, run, k, precision,
                                                               rtol):
    d = replay.load(run)
    s_in = replay.step_inputs(d, k, orc)
    s = nr.NewtonRaphsonSolver(d["frames"].shape[1], s_in["conn"], s_in["refpos"])
    s.set_material(s_in["E"], s_in["nu"], s_in["eta"], s_in["ext_visc"], s_in["fixed"],
                   s_in["pairs"])
    s.set_growth(s_in["invFg"])
    s.set_visc_surface(s_in["visc_surf"])
    if precision == "mixed":
        s.set_krylov("mixed", "cg")
    u, st = s.step(s_in["un"], s_in["dt"], 1e-8, 20, precision, rtol)
    assert st["converged"]
    got = s.element_strain()
    ref = d["strain"][k + 1]
    assert np.abs(ref).max() > 4e-2
    assert np.abs(got - ref).max() < STRAIN_TOL[run]
    s.close()
