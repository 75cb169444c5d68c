"""The allocation-free entry point tf_nr_update_visc_surface_noalloc against
tf_nr_update_visc_surface, through the C ABI: two calls in a row on one
device, the second on a smaller mesh (the shared scratch array keeps the
first call's larger contents), and a batch handle.  The surfaces must be
equal bit for bit (the same kernel reads the same flags), and equal to the
oracle's restatement of Simulation::updateNodeViscositySurfaces.  Needs an
MI355X."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

_d = ctypes.POINTER(ctypes.c_double)
_i = ctypes.POINTER(ctypes.c_int32)
_u8 = ctypes.POINTER(ctypes.c_uint8)


@pytest.fixture(scope="module")
def nr():
    from tissueorigami_amd import nr as _nr
    _nr.load()
    return _nr


def _case(seed, a, b, side, layers):
    from tissueorigami_amd import tissue
    m = tissue.ellipse(a, b, side, 2.5 * layers, layers)
    rng = np.random.default_rng(seed)
    u = (m["nodes"] + rng.normal(0, 0.05 * side, m["nodes"].shape)).reshape(-1)
    place, spans = tissue.element_placement(m["conn"], m["node_placement"])
    etype = tissue.element_type(m["conn"], m["node_type"])
    return m, u, place.astype(np.int32), etype.astype(np.int32), spans.astype(np.uint8)


def _both(nr, s, u, place, etype, spans):
    L = nr.load()
    u = np.ascontiguousarray(u, dtype=np.float64)
    a_ref = np.full(s.n_nodes, np.nan)
    a_new = np.full(s.n_nodes, np.nan)
    args = (u.ctypes.data_as(_d), place.ctypes.data_as(_i), etype.ctypes.data_as(_i),
            spans.ctypes.data_as(_u8))
    assert L.tf_nr_update_visc_surface_noalloc(s._h, *args, a_new.ctypes.data_as(_d)) == 0, \
        L.tf_last_error()
    assert L.tf_nr_update_visc_surface(s._h, *args, a_ref.ctypes.data_as(_d)) == 0
    return a_new, a_ref


def test_noalloc_surfaces_equal_the_allocating_entry_point(orc, nr):
    big = _case(1, 40.0, 28.0, 1.0, 4)
    small = _case(2, 12.0, 9.0, 1.0, 2)
    assert len(small[0]["conn"]) < len(big[0]["conn"])
    for m, u, place, etype, spans in (big, small):
        s = nr.NewtonRaphsonSolver(len(m["nodes"]), m["conn"], m["refpos"])
        a_new, a_ref = _both(nr, s, u, place, etype, spans)
        assert np.array_equal(a_new, a_ref)
        a_o = orc.visc_surfaces(len(m["nodes"]), m["conn"], u, place, etype, spans)
        assert np.abs(a_new - a_o).max() <= 1e-12 * np.abs(a_o).max()
        assert np.count_nonzero(a_new) > 0
        # the surfaces now on the device are the ones the step uses
        s.close()


def test_noalloc_on_a_batch_handle(nr):
    cases = [_case(3, 15.0, 10.0, 1.0, 3), _case(4, 9.0, 6.0, 1.0, 2)]
    b = nr.TissueBatch([(len(c[0]["nodes"]), c[0]["conn"], c[0]["refpos"]) for c in cases])
    cat = lambda k: np.ascontiguousarray(np.concatenate([c[k] for c in cases]))
    a_new, a_ref = _both(nr, b, cat(1), cat(2), cat(3), cat(4))
    assert np.array_equal(a_new, a_ref)
    b.close()


def test_noalloc_step_uses_the_staged_surfaces(nr):
    """A step after the allocation-free update equals a step after the
    allocating one, bit for bit."""
    from tissueorigami_amd import tissue
    m, u, place, etype, spans = _case(5, 10.0, 7.0, 1.0, 2)
    ne = len(m["conn"])
    # the batched sweep's material: viscous, drag 16000 apical / 10 basal
    E = np.full(ne, 1000.0); nu = np.full(ne, 0.3); eta = np.full(ne, 10.0)
    ext = np.zeros((len(m["nodes"]), 3))
    ext[m["node_placement"] == 1] = 16000.0
    ext[m["node_placement"] == 0] = 10.0
    inv = np.linalg.inv(tissue.growth_increment(np.full(ne, 0.02), 60.0)).reshape(-1, 9)
    out = []
    for entry in ("tf_nr_update_visc_surface_noalloc", "tf_nr_update_visc_surface"):
        s = nr.NewtonRaphsonSolver(len(m["nodes"]), m["conn"], m["refpos"])
        s.set_material(E, nu, eta, ext)
        s.set_growth(inv)
        uu = np.ascontiguousarray(u)
        assert getattr(nr.load(), entry)(s._h, uu.ctypes.data_as(_d), place.ctypes.data_as(_i),
                                         etype.ctypes.data_as(_i), spans.ctypes.data_as(_u8),
                                         None) == 0
        x, st = s.step(uu, 60.0, 1e-8, 20, "fp64", 1e-10)
        assert st["converged"]
        out.append(x)
        s.close()
    assert np.array_equal(out[0], out[1])
