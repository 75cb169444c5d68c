"""ctypes binding of the CPU oracle (oracle/liboracle.so).

Test infrastructure only: imported by tests/, by __graft_entry__.smoke() as the
checker and by bench.py's cpu_baseline leg.  The product library never uses it.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")
_lib = None

_d = ctypes.POINTER(ctypes.c_double)
_i = ctypes.POINTER(ctypes.c_int32)
_u8 = ctypes.POINTER(ctypes.c_uint8)


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = ctypes.CDLL(_LIB)
        L.orc_lame.argtypes = [ctypes.c_int, _d, _d, _d, _d]
        L.orc_prism_geometry.argtypes = [_d, _d, _d]
        L.orc_inv3.argtypes = [_d, _d]
        L.orc_elements.argtypes = [ctypes.c_int, _d, _d, _d, _d, _d, _d, _d, ctypes.c_double,
                                   ctypes.c_int, _d, _d, _d, _d]
        L.orc_element_strain.argtypes = [ctypes.c_int, _d, _d, _d, ctypes.c_int, _d]
        L.orc_assemble_dense.argtypes = [ctypes.c_int, ctypes.c_int, _i, _d, _d, _d, _d, _d, _d,
                                         _d, _u8, _i, ctypes.c_int, _d, _d, ctypes.c_double,
                                         ctypes.c_int, _d, _d, _d]
        L.orc_csr_threshold.argtypes = [_d, ctypes.c_int, _i, _i, _d, ctypes.c_int]
        L.orc_csr_threshold.restype = ctypes.c_int
        L.orc_dense_solve_transposed.argtypes = [_d, ctypes.c_int, _d, _d]
        L.orc_dense_solve_transposed.restype = ctypes.c_int
        L.orc_dnrm2.argtypes = [_d, ctypes.c_int]
        L.orc_dnrm2.restype = ctypes.c_double
        L.orc_visc_surfaces.argtypes = [ctypes.c_int, ctypes.c_int, _i, _d, _i, _i, _u8, _d]
        step_args = [ctypes.c_int, ctypes.c_int, _i, _d, _d, _d, _d, _d, _d, _d, _u8, _i,
                     ctypes.c_int, _d, ctypes.c_double, ctypes.c_double, ctypes.c_int]
        L.orc_nr_step_dense.argtypes = step_args + [ctypes.c_int, _d, _d, _i]
        L.orc_nr_step_dense.restype = ctypes.c_int
        L.orc_nr_step_sparse.argtypes = step_args + [ctypes.c_double, ctypes.c_int, _d, _d, _i, _i]
        L.orc_nr_step_sparse.restype = ctypes.c_int
        L.orc_nr_iteration_sparse.argtypes = [ctypes.c_int, ctypes.c_int, _i, _d, _d, _d, _d, _d,
                                              _d, _d, _u8, _d, _d, ctypes.c_double,
                                              ctypes.c_double, ctypes.c_int, _d, _d, _d]
        L.orc_nr_iteration_sparse.restype = ctypes.c_int
        L.orc_assemble_bsr.argtypes = [ctypes.c_int, ctypes.c_int, _i, _d, _d, _d, _d, _d, _d,
                                       _d, _u8, _d, _d, ctypes.c_double, _i, _i, _d, _d]
        L.orc_assemble_bsr.restype = ctypes.c_long
        L.orc_num_threads.restype = ctypes.c_int
        L.orc_set_packing.argtypes = [_i, ctypes.c_int, _d, ctypes.c_int, ctypes.c_double]
        L.orc_set_ablation.argtypes = [_u8, ctypes.c_int, _i, ctypes.c_int]
        L.orc_packing_couple.argtypes = [_d, _d, ctypes.c_double, ctypes.c_double,
                                         ctypes.c_double, _d, _d]
        L.orc_reference_volumes.argtypes = [ctypes.c_int, _d, _d]
        L.orc_node_masses.argtypes = [ctypes.c_int, ctypes.c_int, _i, _d, _d, _d]
        L.orc_bounding_box.argtypes = [ctypes.c_int, _d, ctypes.c_int, _d]
        L.orc_stiffness_rates.argtypes = [ctypes.c_int, _d, ctypes.c_int, ctypes.c_int, _d,
                                          ctypes.c_int, _d]
        L.orc_stiffness_interval.argtypes = [ctypes.c_int, _d, ctypes.c_double]
        L.orc_stiffness_interval.restype = ctypes.c_int
        L.orc_stiffness_rate_at.argtypes = [_d, ctypes.c_int, ctypes.c_int, ctypes.c_int, _d]
        L.orc_stiffness_rate_at.restype = ctypes.c_double
        L.orc_stiffness_apply.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                          ctypes.c_int, _d, ctypes.c_int, ctypes.c_int, _d,
                                          ctypes.c_int, ctypes.c_int, _u8, _d, _d, _d, _d, _d, _d]
        L.orc_stiffness_apply.restype = ctypes.c_int
        L.orc_packing_thresholds.argtypes = [ctypes.c_int, ctypes.c_int, _i, _d, _i, _i, _u8,
                                             ctypes.c_int, _d, _d, _d]
        L.orc_detect_packing.argtypes = [ctypes.c_int, ctypes.c_int, _i, _d, _i, _u8, _d, _d, _d,
                                         _i, ctypes.c_int]
        L.orc_detect_packing.restype = ctypes.c_int
        L.orc_grid_compatible.argtypes = [ctypes.c_int, ctypes.c_int, _d, _d, _d, _d, _u8]
        L.orc_polar_rotation.argtypes = [_d, _d]
        L.orc_polar_rotation.restype = ctypes.c_int
        L.orc_growth_rotation.argtypes = [_d, _d, _d]
        L.orc_growth_rotation.restype = ctypes.c_int
        L.orc_relative_positions.argtypes = [ctypes.c_int, _d, _u8, ctypes.c_int, ctypes.c_int,
                                             _i, _i, _d]
        L.orc_growth_step.argtypes = [ctypes.c_int, _i, _d, _d, _i, _i, _u8, _d, _d,
                                      ctypes.c_int, ctypes.c_int, _i, _i, _d, _d, _d, _d, _i,
                                      _d, _d, _d, ctypes.c_double, ctypes.c_double,
                                      ctypes.c_int, _u8, ctypes.c_int, ctypes.c_int, _i, _d,
                                      _d, _d]
        _lib = L
    return _lib


def _p(a, t):
    if a is None:
        return None
    return a.ctypes.data_as(t)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def num_threads():
    return lib().orc_num_threads()


def lame(E, nu):
    E = _f64(E); nu = _f64(np.broadcast_to(nu, E.shape))
    lam = np.empty_like(E); mu = np.empty_like(E)
    lib().orc_lame(E.size, _p(E, _d), _p(nu, _d), _p(lam, _d), _p(mu, _d))
    return lam, mu


def inv3(A):
    A = _f64(A).reshape(9)
    X = np.empty(9)
    lib().orc_inv3(_p(A, _d), _p(X, _d))
    return X.reshape(3, 3)


def inv3_many(M):
    M = _f64(M).reshape(-1, 9)
    return np.stack([inv3(m).reshape(9) for m in M])


def prism_geometry(ref18):
    ref18 = _f64(ref18).reshape(18)
    inv = np.empty(54); det = np.empty(6)
    lib().orc_prism_geometry(_p(ref18, _d), _p(inv, _d), _p(det, _d))
    return inv.reshape(6, 3, 3), det


def element_strain(pos, refpos, invFg, gp=5):
    """Green-Lagrange strain of each prism's last Gauss point (n_el x 6, Voigt
    E00, E11, E22, 2E01, 2E21, 2E02): ShapeBase::Strain, Prism.cpp:826-837.
    pos and refpos: per-prism node positions (n_el x 18); invFg: n_el x 9;
    gp: the Gauss point evaluated (the reference keeps the last, 5)."""
    pos = _f64(pos).reshape(-1, 18); ne = pos.shape[0]
    refpos = _f64(refpos).reshape(ne, 18); invFg = _f64(invFg).reshape(ne, 9)
    out = np.empty((ne, 6))
    if not 0 <= int(gp) < 6:
        raise ValueError("gp must be 0..5")
    lib().orc_element_strain(ne, _p(pos, _d), _p(refpos, _d), _p(invFg, _d), int(gp), _p(out, _d))
    return out


def elements(pos, refpos, invFg, lam, mu, eta, vel=None, dt=1.0, tangent_mode=1):
    """Per-element forces (ge, gv: n_el x 18) and tangents (Ke, Kv: n_el x 18 x 18)."""
    pos = _f64(pos).reshape(-1, 18); ne = pos.shape[0]
    refpos = _f64(refpos).reshape(ne, 18); invFg = _f64(invFg).reshape(ne, 9)
    lam = _f64(lam); mu = _f64(mu); eta = _f64(eta)
    vel = None if vel is None else _f64(vel).reshape(ne, 18)
    ge = np.empty((ne, 18)); gv = np.empty((ne, 18))
    Ke = np.empty((ne, 18, 18)); Kv = np.empty((ne, 18, 18))
    lib().orc_elements(ne, _p(pos, _d), _p(refpos, _d), _p(invFg, _d), _p(lam, _d), _p(mu, _d),
                       _p(eta, _d), _p(vel, _d), float(dt), int(tangent_mode), _p(ge, _d),
                       _p(gv, _d), _p(Ke, _d), _p(Kv, _d))
    return ge, gv, Ke, Kv


class Problem:
    """Arrays of one tissue in the oracle's (and the C ABI's) layout."""

    def __init__(self, conn, refpos, invFg, lam, mu, eta, ext_visc, visc_surf, fixed=None,
                 pairs=None):
        self.conn = _i32(conn).reshape(-1, 6)
        self.n_el = self.conn.shape[0]
        self.refpos = _f64(refpos).reshape(self.n_el, 18)
        self.invFg = _f64(invFg).reshape(self.n_el, 9)
        self.lam = _f64(lam); self.mu = _f64(mu); self.eta = _f64(eta)
        self.ext_visc = _f64(ext_visc).reshape(-1, 3)
        self.n_nodes = self.ext_visc.shape[0]
        self.visc_surf = _f64(visc_surf)
        self.fixed = None if fixed is None else np.ascontiguousarray(fixed, dtype=np.uint8).reshape(-1)
        self.pairs = (np.zeros((0, 2), np.int32) if pairs is None else _i32(pairs).reshape(-1, 2))

    def _common(self):
        return (self.n_nodes, self.n_el, _p(self.conn, _i), _p(self.refpos, _d),
                _p(self.invFg, _d), _p(self.lam, _d), _p(self.mu, _d), _p(self.eta, _d),
                _p(self.ext_visc, _d), _p(self.visc_surf, _d), _p(self.fixed, _u8),
                _p(self.pairs, _i), self.pairs.shape[0])

    def assemble_dense(self, uk, un, dt, tangent_mode=1):
        n = 3 * self.n_nodes
        uk = _f64(uk).reshape(n); un = _f64(un).reshape(n)
        K = np.empty((n, n)); g = np.empty(n); gi = np.empty(n)
        lib().orc_assemble_dense(*self._common(), _p(uk, _d), _p(un, _d), float(dt),
                                 int(tangent_mode), _p(K, _d), _p(g, _d), _p(gi, _d))
        return K, g, gi

    def nr_step_dense(self, un, dt, tol=1e-8, maxit=20, tangent_mode=1):
        n = 3 * self.n_nodes
        un = _f64(un).reshape(n)
        out = np.empty(n); du = np.zeros(maxit + 2); conv = np.zeros(1, np.int32)
        it = lib().orc_nr_step_dense(*self._common(), _p(un, _d), float(dt), float(tol),
                                     int(maxit), int(tangent_mode), _p(out, _d), _p(du, _d),
                                     _p(conv, _i))
        if it < 0:
            raise RuntimeError("oracle dense solve failed (singular matrix)")
        return out, du[:it], bool(conv[0])

    def nr_step_sparse(self, un, dt, tol=1e-8, maxit=20, krylov_rtol=1e-12, krylov_maxit=20000):
        n = 3 * self.n_nodes
        un = _f64(un).reshape(n)
        out = np.empty(n); du = np.zeros(maxit + 2); kit = np.zeros(maxit + 2, np.int32)
        conv = np.zeros(1, np.int32)
        it = lib().orc_nr_step_sparse(*self._common(), _p(un, _d), float(dt), float(tol),
                                      int(maxit), float(krylov_rtol), int(krylov_maxit),
                                      _p(out, _d), _p(du, _d), _p(kit, _i), _p(conv, _i))
        if it < 0:
            raise RuntimeError("oracle BiCGStab breakdown")
        return out, du[:it], kit[:it], bool(conv[0])

    def assemble_bsr(self, uk, un, dt):
        """Sparse J = K^T (block rows, block columns, 3x3 blocks) and gSum."""
        n = 3 * self.n_nodes
        uk = _f64(uk).reshape(n); un = _f64(un).reshape(n)
        L = lib()
        args = (self.n_nodes, self.n_el, _p(self.conn, _i), _p(self.refpos, _d),
                _p(self.invFg, _d), _p(self.lam, _d), _p(self.mu, _d), _p(self.eta, _d),
                _p(self.ext_visc, _d), _p(self.visc_surf, _d), _p(self.fixed, _u8),
                _p(uk, _d), _p(un, _d), float(dt))
        nb = L.orc_assemble_bsr(*args, None, None, None, None)
        rowptr = np.empty(self.n_nodes + 1, np.int32); col = np.empty(nb, np.int32)
        val = np.empty((nb, 3, 3)); g = np.empty(n)
        L.orc_assemble_bsr(*args, _p(rowptr, _i), _p(col, _i), _p(val, _d), _p(g, _d))
        return rowptr, col, val, g

    def nr_iteration_sparse(self, uk, un, dt, krylov_rtol, krylov_maxit):
        n = 3 * self.n_nodes
        uk = _f64(uk).reshape(n); un = _f64(un).reshape(n)
        du = np.empty(n); ta = np.zeros(1); ts = np.zeros(1)
        k = lib().orc_nr_iteration_sparse(
            self.n_nodes, self.n_el, _p(self.conn, _i), _p(self.refpos, _d), _p(self.invFg, _d),
            _p(self.lam, _d), _p(self.mu, _d), _p(self.eta, _d), _p(self.ext_visc, _d),
            _p(self.visc_surf, _d), _p(self.fixed, _u8), _p(uk, _d), _p(un, _d), float(dt),
            float(krylov_rtol), int(krylov_maxit), _p(du, _d), _p(ta, _d), _p(ts, _d))
        return du, k, float(ta[0]), float(ts[0])


def csr_threshold(K):
    K = _f64(K); n = K.shape[0]
    nnz = lib().orc_csr_threshold(_p(K, _d), n, None, None, None, 0)
    ia = np.empty(n + 1, np.int32); ja = np.empty(nnz, np.int32); a = np.empty(nnz)
    lib().orc_csr_threshold(_p(K, _d), n, _p(ia, _i), _p(ja, _i), _p(a, _d), nnz)
    return ia, ja, a


def dense_solve_transposed(K, b):
    K = _f64(K); b = _f64(b); x = np.empty_like(b)
    if lib().orc_dense_solve_transposed(_p(K, _d), K.shape[0], _p(b, _d), _p(x, _d)) != 0:
        raise RuntimeError("singular")
    return x


def dnrm2(x):
    x = _f64(x)
    return lib().orc_dnrm2(_p(x, _d), x.size)


def visc_surfaces(n_nodes, conn, u, elem_placement, elem_type, spans):
    conn = _i32(conn); u = _f64(u).reshape(-1)
    ep = _i32(elem_placement); et = _i32(elem_type)
    sp = np.ascontiguousarray(spans, dtype=np.uint8)
    A = np.empty(n_nodes)
    lib().orc_visc_surfaces(n_nodes, conn.shape[0], _p(conn, _i), _p(u, _d), _p(ep, _i),
                            _p(et, _i), _p(sp, _u8), _p(A, _d))
    return A


# ---- node-node packing (Simulation.cpp:3197-3260, 7307-7572, 7997-8152) ----

class packing:
    """Context in which the oracle's dense assembly and NR step add the packing
    couples `pairs` (n x 2, detection order) with node masses and the packing
    threshold (Simulation::packingThreshold)."""

    def __init__(self, pairs, mass, threshold):
        self.pairs = _i32(pairs).reshape(-1, 2)
        self.mass = _f64(mass)
        self.threshold = float(threshold)

    def __enter__(self):
        lib().orc_set_packing(_p(self.pairs, _i), self.pairs.shape[0], _p(self.mass, _d),
                              self.mass.size, self.threshold)
        return self

    def __exit__(self, *exc):
        lib().orc_set_packing(None, 0, None, 0, 0.0)
        return False


class ablation:
    """Context in which the oracle's assemblies, NR steps and viscosity surfaces
    treat the prisms flagged in `prisms` (n_prisms, ShapeBase::IsAblated) as
    laser-ablated and apply checkJacobianForAblatedNodes to `nodes`
    (Simulation::AblatedNodes)."""

    def __init__(self, prisms=None, nodes=None):
        self.prisms = None if prisms is None else np.ascontiguousarray(prisms, dtype=np.uint8)
        self.nodes = None if nodes is None else _i32(nodes).reshape(-1)

    def __enter__(self):
        lib().orc_set_ablation(_p(self.prisms, _u8), 0 if self.prisms is None else self.prisms.size,
                               _p(self.nodes, _i), 0 if self.nodes is None else self.nodes.size)
        return self

    def __exit__(self, *exc):
        lib().orc_set_ablation(None, 0, None, 0)
        return False


def packing_couple(p0, p1, m0, m1, threshold):
    """Force on the first node of a couple and its 3x3 derivative dF0/dx0."""
    p0 = _f64(p0); p1 = _f64(p1); F = np.empty(3); D = np.empty(9)
    lib().orc_packing_couple(_p(p0, _d), _p(p1, _d), float(m0), float(m1), float(threshold),
                             _p(F, _d), _p(D, _d))
    return F, D.reshape(3, 3)


def reference_volumes(refpos):
    refpos = _f64(refpos).reshape(-1, 18); V = np.empty(refpos.shape[0])
    lib().orc_reference_volumes(refpos.shape[0], _p(refpos, _d), _p(V, _d))
    return V


def node_masses(n_nodes, conn, vol, Fg):
    conn = _i32(conn); vol = _f64(vol); Fg = _f64(Fg).reshape(-1, 9)
    m = np.empty(n_nodes)
    lib().orc_node_masses(n_nodes, vol.size, _p(conn, _i), _p(vol, _d), _p(Fg, _d), _p(m, _d))
    return m


def bounding_box(u, symmetric_y=False):
    u = _f64(u).reshape(-1); bb = np.empty(6)
    lib().orc_bounding_box(u.size // 3, _p(u, _d), int(symmetric_y), _p(bb, _d))
    return bb


def packing_thresholds(n_nodes, conn, u, elem_placement, elem_type, spans, symmetric_y=False):
    """(squared detection thresholds on the 10 x 5 grid, packing threshold, bounding box)."""
    conn = _i32(conn); u = _f64(u).reshape(-1)
    ep = _i32(elem_placement); et = _i32(elem_type)
    sp = np.ascontiguousarray(spans, dtype=np.uint8)
    grid = np.empty(50); thr = np.zeros(1); bb = np.empty(6)
    lib().orc_packing_thresholds(n_nodes, conn.size // 6, _p(conn, _i), _p(u, _d), _p(ep, _i),
                                 _p(et, _i), _p(sp, _u8), int(symmetric_y), _p(grid, _d),
                                 _p(thr, _d), _p(bb, _d))
    return grid.reshape(10, 5), float(thr[0]), bb


def detect_packing(n_nodes, conn, u, node_placement, mass, grid_sq, bb, node_excluded=None):
    """Couples (n x 2, int32) in the reference's detection order."""
    conn = _i32(conn); u = _f64(u).reshape(-1); npl = _i32(node_placement)
    ex = None if node_excluded is None else np.ascontiguousarray(node_excluded, dtype=np.uint8)
    mass = _f64(mass); grid = _f64(grid_sq).reshape(50); bb = _f64(bb)
    args = (n_nodes, conn.size // 6, _p(conn, _i), _p(u, _d), _p(npl, _i), _p(ex, _u8),
            _p(mass, _d), _p(grid, _d), _p(bb, _d))
    n = lib().orc_detect_packing(*args, None, 0)
    pairs = np.zeros((max(n, 1), 2), np.int32)
    lib().orc_detect_packing(*args, _p(pairs, _i), n)
    return pairs[:n]


# ---- Young's modulus time series (YoungsModulusModifier.cpp)
def stiffness_rates(times, grids, linear=True):
    """Change-rate grids, (n_times - 1, ny, nx), of grids (n_times, ny, nx)."""
    times = _f64(times); grids = _f64(grids)
    nt, ny, nx = grids.shape
    out = np.empty((nt - 1, ny, nx))
    lib().orc_stiffness_rates(nt, _p(times, _d), nx, ny, _p(grids, _d), int(bool(linear)),
                              _p(out, _d))
    return out


def stiffness_interval(times, t):
    times = _f64(times)
    return lib().orc_stiffness_interval(times.size, _p(times, _d), float(t))


def stiffness_rate_at(rate_t, method, rel):
    rate_t = _f64(rate_t); rel = _f64(rel).reshape(2)
    ny, nx = rate_t.shape
    return lib().orc_stiffness_rate_at(_p(rate_t, _d), nx, ny, int(method), _p(rel, _d))


class StiffnessSeries:
    """One Young's modulus series over a mesh's prisms, stepped on the CPU."""

    def __init__(self, times, grids, linear, method, applicable, rel_pos):
        self.times = _f64(times); self.grids = _f64(grids)
        self.linear = int(bool(linear)); self.method = int(method)
        self.rate = stiffness_rates(self.times, self.grids, linear)
        self.applicable = np.ascontiguousarray(applicable, np.uint8).reshape(-1)
        self.rel = _f64(rel_pos).reshape(-1, 2)

    def apply(self, t, dt, E, nu, stm, lam, mu):
        """Updates stm, lam, mu in place; returns the interval used (-1: none)."""
        E = _f64(E); nu = _f64(nu)
        nt, ny, nx = self.grids.shape
        return lib().orc_stiffness_apply(
            E.size, float(t), float(dt), nt, _p(self.times, _d), nx, ny, _p(self.rate, _d),
            self.linear, self.method, _p(self.applicable, _u8), _p(self.rel, _d), _p(E, _d),
            _p(nu, _d), _p(stm, _d), _p(lam, _d), _p(mu, _d))


# ---- growth producers (oracle_growth.cpp)

def grid_compatible(rates, angles_deg):
    """Compatible corner data of a growth grid: rates (nx, ny, 3), angles (nx, ny)
    in degrees -> (growths (cells, 4, 3), angles (cells, 4), eliminated (cells, 4)),
    cell ix * (ny - 1) + iy."""
    r = _f64(rates)
    a = _f64(angles_deg)
    nx, ny = a.shape
    cells = (nx - 1) * (ny - 1)
    cg = np.zeros((cells, 4, 3)); ca = np.zeros((cells, 4)); ce = np.zeros((cells, 4), np.uint8)
    lib().orc_grid_compatible(nx, ny, _p(r, _d), _p(a, _d), _p(cg, _d), _p(ca, _d), _p(ce, _u8))
    return cg, ca, ce.astype(bool)


def polar_rotation(F):
    F = _f64(F).reshape(3, 3)
    R = np.zeros((3, 3))
    return R if lib().orc_polar_rotation(_p(F, _d), _p(R, _d)) else None


def growth_rotation(ref18, pos18):
    """GrowthStrainsRotMat of a prism (3 x 3), or None when its growth is not rotated."""
    r = _f64(ref18).reshape(18)
    q = _f64(pos18).reshape(18)
    g = np.zeros((3, 3))
    return g if lib().orc_growth_rotation(_p(r, _d), _p(q, _d), _p(g, _d)) else None


def relative_positions(u, conn, excluded=None, symmetric_y=False, apical_of=None):
    u = _f64(u).reshape(-1)
    c = _i32(conn).reshape(-1, 6)
    ex = None if excluded is None else np.ascontiguousarray(excluded, np.uint8)
    ap = None if apical_of is None else _i32(apical_of)
    rel = np.zeros((len(c), 2))
    lib().orc_relative_positions(u.size // 3, _p(u, _d), _p(ex, _u8), int(bool(symmetric_y)),
                                 len(c), _p(c, _i), _p(ap, _i), _p(rel, _d))
    return rel


def growth_step(conn, refpos, u, prisms, fns, rel, t, dt, interp, Fg, ablated=None,
                columns=None):
    """One step of Simulation::calculateGrowth and growShapeByFg.  prisms: dict of
    type, place, flags, rel_z, weights (or None), explicit_ecm; fns: list of dicts
    with type (1 or 3), t (t0, t1), apply (bits), z (zmin, zmax) and rate (1/s),
    angle (radians) for type 1, rates (nx, ny, 3) in 1/s and angles_deg (nx, ny)
    for type 3; columns (n, layers) or None.  Returns (Fg, invFg), n x 9."""
    c = _i32(conn).reshape(-1, 6)
    ne = len(c)
    ref = _f64(refpos).reshape(ne, 18)
    u = _f64(u).reshape(-1)
    tt = _i32(prisms["type"]); pl = _i32(prisms["place"])
    fl = np.ascontiguousarray(prisms["flags"], np.uint8)
    w = prisms.get("weights")
    w = None if w is None else _f64(w).reshape(ne, 2)
    rz = _f64(prisms["rel_z"])
    n = len(fns)
    ftype = np.array([f["type"] for f in fns], np.int32)
    fapply = np.array([f["apply"] for f in fns], np.int32)
    ft = _f64([f["t"] for f in fns]).reshape(-1)
    fz = _f64([f.get("z", (0.0, 1.0)) for f in fns]).reshape(-1)
    frate = _f64([f.get("rate", (0.0, 0.0, 0.0)) for f in fns]).reshape(-1)
    fang = _f64([f.get("angle", 0.0) for f in fns])
    fgrid = np.zeros((max(n, 1), 3), np.int32)
    grates, gang, off = [], [], 0
    for k, f in enumerate(fns):
        if f["type"] == 3:
            a = _f64(f["angles_deg"])
            fgrid[k] = (a.shape[0], a.shape[1], off)
            grates.append(_f64(f["rates"]).reshape(-1))
            gang.append(a.reshape(-1))
            off += a.size
    grates = _f64(np.concatenate(grates) if grates else np.zeros(1))
    gang = _f64(np.concatenate(gang) if gang else np.zeros(1))
    rel = _f64(rel).reshape(ne, 2)
    ab = None if ablated is None else np.ascontiguousarray(ablated, np.uint8)
    cols = None if columns is None else _i32(columns).reshape(len(columns), -1)
    ncol, layers = (0, 0) if cols is None else cols.shape
    vref = reference_volumes(ref) if cols is not None else None
    Fo = _f64(Fg).reshape(ne, 9).copy()
    inv = np.zeros((ne, 9))
    lib().orc_growth_step(ne, _p(c, _i), _p(ref, _d), _p(u, _d), _p(tt, _i), _p(pl, _i),
                          _p(fl, _u8), _p(w, _d), _p(rz, _d), int(bool(prisms.get("explicit_ecm"))),
                          n, _p(ftype, _i), _p(fapply, _i), _p(ft, _d), _p(fz, _d),
                          _p(frate, _d), _p(fang, _d), _p(fgrid, _i), _p(grates, _d),
                          _p(gang, _d), _p(rel, _d), float(t), float(dt), int(interp),
                          _p(ab, _u8), ncol, layers, _p(cols, _i),
                          _p(None if vref is None else _f64(vref), _d), _p(Fo, _d), _p(inv, _d))
    return Fo, inv
