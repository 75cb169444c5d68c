"""Python binding of libtissue_nr.so (the C ABI of include/tissue_nr.h).

The class mirrors the calls Simulation::updateStepNR makes on the reference's
NewtonRaphsonSolver (TissueFolding/SourceCode/NewtonRaphsonSolver.h:51-90) at
the granularity the C ABI offers: set-up once per mesh, per-step inputs,
tf_nr_step for a whole implicit step, and assemble / solve / get_csr for
testing the pieces.  There is no CPU fallback: loading fails loudly when the
HIP library has not been built, and every call raises on a non-zero status.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtissue_nr.so")

TF_OK, TF_ERR_HIP, TF_ERR_NAN, TF_ERR_KRYLOV, TF_ERR_NOT_CONVERGED, TF_ERR_ARG, TF_ERR_NO_DEVICE = range(7)
PREC = {"fp64": 0, "mixed": 1}
MAX_NR = 32

EXPORTS = ["tf_nr_create", "tf_nr_set_material", "tf_nr_set_growth", "tf_nr_set_visc_surface",
           "tf_nr_update_visc_surface", "tf_nr_update_visc_surface_noalloc", "tf_nr_set_ablation", "tf_nr_step", "tf_nr_assemble", "tf_nr_solve",
           "tf_csr_pattern", "tf_nr_get_csr", "tf_nr_element_outputs", "tf_nr_element_strain",
           "tf_nr_profile",
           "tf_nr_kernel_times", "tf_nr_pattern_info", "tf_adj_forward", "tf_adj_backward",
           "tf_nr_set_growth_rates", "tf_nr_growth_tensor", "tf_nr_set_surface_flags", "tf_nr_advance",
           "tf_batch_create", "tf_nr_autotune", "tf_nr_set_tile_size", "tf_nr_set_graphs", "tf_nr_set_krylov", "tf_nr_assemble_ex",
           "tf_copy_bandwidth", "tf_nr_set_coarse", "tf_nr_set_coarse_bands", "tf_nr_set_coarse_lag", "tf_nr_coarse_info", "tf_nr_set_warm_start",
           "tf_nr_set_forcing_safeguard",
           "tf_nr_set_packing", "tf_nr_set_node_mass", "tf_nr_update_node_mass",
           "tf_nr_detect_packing", "tf_nr_get_packing",
           "tf_nr_add_stiffness_series", "tf_nr_apply_stiffness_series", "tf_nr_set_time",
           "tf_nr_get_time", "tf_nr_get_stiffness_multiplier",
           "tf_nr_set_growth_prisms", "tf_nr_add_uniform_growth", "tf_nr_add_grid_growth",
           "tf_nr_set_growth_options", "tf_nr_set_column_volume_conservation",
           "tf_nr_update_relative_positions", "tf_nr_relative_positions", "tf_nr_grow",
           "tf_nr_destroy", "tf_last_error"]


class StepStats(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int32), ("nr_iterations", ctypes.c_int32),
                ("converged", ctypes.c_int32), ("krylov_iterations_total", ctypes.c_int32),
                ("krylov_iterations", ctypes.c_int32 * MAX_NR),
                ("du_norms", ctypes.c_double * MAX_NR),
                ("ms_element", ctypes.c_double), ("ms_assembly", ctypes.c_double),
                ("ms_krylov", ctypes.c_double), ("ms_total", ctypes.c_double),
                ("fp64_fallbacks", ctypes.c_int32), ("cg_fallbacks", ctypes.c_int32),
                ("gmres_fallbacks", ctypes.c_int32)]

    def as_dict(self):
        n = self.nr_iterations
        return dict(status=self.status, nr_iterations=n, converged=bool(self.converged),
                    krylov_iterations_total=self.krylov_iterations_total,
                    krylov_iterations=list(self.krylov_iterations[:n]),
                    du_norms=list(self.du_norms[:n]), ms_element=self.ms_element,
                    ms_assembly=self.ms_assembly, ms_krylov=self.ms_krylov,
                    ms_total=self.ms_total, fp64_fallbacks=self.fp64_fallbacks,
                    cg_fallbacks=self.cg_fallbacks, gmres_fallbacks=self.gmres_fallbacks)


class KernelTimes(ctypes.Structure):
    _fields_ = [("spmv_ms", ctypes.c_double * 2), ("spmv_launches", ctypes.c_int64 * 2),
                ("spmv_t_ms", ctypes.c_double * 2), ("spmv_t_launches", ctypes.c_int64 * 2),
                ("element_ms", ctypes.c_double), ("assembly_ms", ctypes.c_double),
                ("assemblies", ctypes.c_int64)]


class TFError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"tissue_nr status {code}: {msg}")
        self.code = code


_lib = None
_d = ctypes.POINTER(ctypes.c_double)
_i = ctypes.POINTER(ctypes.c_int32)
_i64 = ctypes.POINTER(ctypes.c_int64)
_u8 = ctypes.POINTER(ctypes.c_uint8)
_h = ctypes.c_void_p


def load():
    """Loads the HIP library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with python -m tissueorigami_amd.build")
    L = ctypes.CDLL(LIB_PATH)
    L.tf_nr_create.argtypes = [ctypes.c_int32, ctypes.c_int32, _i, _d, ctypes.POINTER(_h)]
    L.tf_nr_set_material.argtypes = [_h, _d, _d, _d, _d, _u8, _i, ctypes.c_int32]
    L.tf_nr_set_growth.argtypes = [_h, _d]
    L.tf_nr_set_visc_surface.argtypes = [_h, _d]
    L.tf_nr_update_visc_surface.argtypes = [_h, _d, _i, _i, _u8, _d]
    L.tf_nr_update_visc_surface_noalloc.argtypes = [_h, _d, _i, _i, _u8, _d]
    L.tf_nr_set_ablation.argtypes = [_h, _u8, _i, ctypes.c_int32]
    L.tf_nr_step.argtypes = [_h, _d, ctypes.c_double, ctypes.c_double, ctypes.c_int32,
                             ctypes.c_int32, ctypes.c_double, _d, ctypes.POINTER(StepStats)]
    L.tf_nr_assemble.argtypes = [_h, _d, _d, ctypes.c_double, _d, _d]
    L.tf_nr_assemble_ex.argtypes = [_h, _d, _d, ctypes.c_double, ctypes.c_int32, _d, _d]
    L.tf_nr_solve.argtypes = [_h, _d, _d, ctypes.c_double, ctypes.c_int32, _i]
    L.tf_csr_pattern.argtypes = [_h, _i, _i, _i64]
    L.tf_nr_get_csr.argtypes = [_h, ctypes.c_double, _i, _i, _d, _i64]
    L.tf_nr_element_outputs.argtypes = [_h, _d, _d, _d, _d]
    L.tf_nr_element_strain.argtypes = [_h, _d]
    L.tf_nr_profile.argtypes = [_h, ctypes.c_int32]
    L.tf_nr_kernel_times.argtypes = [_h, ctypes.POINTER(KernelTimes)]
    L.tf_nr_pattern_info.argtypes = [_h, _i64, _i64]
    L.tf_adj_forward.argtypes = [_h, ctypes.c_int32, _d, _d, ctypes.c_double, ctypes.c_double,
                                 ctypes.c_int32, ctypes.c_double, _i, _i, _u8, _d,
                                 ctypes.POINTER(StepStats)]
    L.tf_adj_backward.argtypes = [_h, _d, _d, _d, ctypes.c_double, _d, ctypes.POINTER(StepStats)]
    L.tf_nr_set_growth_rates.argtypes = [_h, _d]
    L.tf_nr_growth_tensor.argtypes = [_h, _d, _d]
    L.tf_nr_set_surface_flags.argtypes = [_h, _i, _i, _u8]
    L.tf_nr_advance.argtypes = [_h, _d, ctypes.c_int32, ctypes.c_double, ctypes.c_double,
                                ctypes.c_int32, ctypes.c_int32, ctypes.c_double, _d,
                                ctypes.POINTER(StepStats)]
    L.tf_batch_create.argtypes = [ctypes.c_int32, _i, _i, _i, _d, ctypes.POINTER(_h)]
    L.tf_nr_autotune.argtypes = [_h, ctypes.c_int32, _i, _d]
    L.tf_nr_set_tile_size.argtypes = [_h, ctypes.c_int32]
    L.tf_nr_set_graphs.argtypes = [_h, ctypes.c_int32]
    L.tf_nr_set_krylov.argtypes = [_h, ctypes.c_int32, ctypes.c_int32]
    L.tf_copy_bandwidth.argtypes = [ctypes.c_int64, ctypes.c_int32, _d]
    L.tf_nr_set_coarse.argtypes = [_h, ctypes.c_double]
    L.tf_nr_set_coarse_bands.argtypes = [_h, ctypes.c_int32]
    L.tf_nr_set_coarse_lag.argtypes = [_h, ctypes.c_int32]
    L.tf_nr_set_warm_start.argtypes = [_h, ctypes.c_int32]
    L.tf_nr_set_forcing_safeguard.argtypes = [_h, ctypes.c_int32]
    L.tf_nr_coarse_info.argtypes = [_h, _i, _i]
    L.tf_nr_set_packing.argtypes = [_h, _i, _u8, ctypes.c_int32, ctypes.c_int32]
    L.tf_nr_set_node_mass.argtypes = [_h, _d]
    L.tf_nr_update_node_mass.argtypes = [_h, _d, _d]
    L.tf_nr_detect_packing.argtypes = [_h, _d, _i]
    L.tf_nr_get_packing.argtypes = [_h, _i, _i, _d, _d]
    L.tf_nr_add_stiffness_series.argtypes = [_h, ctypes.c_int32, _d, ctypes.c_int32,
                                             ctypes.c_int32, _d, ctypes.c_int32, ctypes.c_int32,
                                             _u8, _d]
    L.tf_nr_apply_stiffness_series.argtypes = [_h, ctypes.c_double, ctypes.c_double]
    L.tf_nr_set_time.argtypes = [_h, ctypes.c_double]
    L.tf_nr_get_time.argtypes = [_h, _d]
    L.tf_nr_get_stiffness_multiplier.argtypes = [_h, _d]
    L.tf_nr_set_growth_prisms.argtypes = [_h, _i, _i, _u8, _d, _d, _i, _u8, ctypes.c_int32,
                                          ctypes.c_int32]
    L.tf_nr_add_uniform_growth.argtypes = [_h, ctypes.c_double, ctypes.c_double, ctypes.c_int32,
                                           ctypes.c_double, ctypes.c_double, _d, ctypes.c_double]
    L.tf_nr_add_grid_growth.argtypes = [_h, ctypes.c_double, ctypes.c_double, ctypes.c_int32,
                                        ctypes.c_double, ctypes.c_double, ctypes.c_int32,
                                        ctypes.c_int32, _d, _d]
    L.tf_nr_set_growth_options.argtypes = [_h, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, _d]
    L.tf_nr_set_column_volume_conservation.argtypes = [_h, ctypes.c_int32, ctypes.c_int32, _i]
    L.tf_nr_update_relative_positions.argtypes = [_h, _d]
    L.tf_nr_relative_positions.argtypes = [_h, _d, _d]
    L.tf_nr_grow.argtypes = [_h, _d, ctypes.c_double, ctypes.c_double]
    L.tf_nr_destroy.argtypes = [_h]
    L.tf_last_error.restype = ctypes.c_char_p
    for name in EXPORTS[:-1]:
        getattr(L, name).restype = ctypes.c_int
    _lib = L
    return L


def _p(a, t):
    return None if a is None else a.ctypes.data_as(t)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _check(code):
    if code != TF_OK:
        raise TFError(code, load().tf_last_error().decode())


def copy_bandwidth(mib=1024, reps=10):
    """GB/s (read + write) of the library's plain copy kernel over mib MiB."""
    out = np.zeros(1)
    _check(load().tf_copy_bandwidth(int(mib) << 20, int(reps), _p(out, _d)))
    return float(out[0])


class NewtonRaphsonSolver:
    """One tissue on one device (a tf_handle)."""

    n_tissues = 1

    def __init__(self, n_nodes, conn, refpos):
        L = load()
        self.conn = np.ascontiguousarray(conn, dtype=np.int32).reshape(-1, 6)
        self.n_nodes = int(n_nodes)
        self.n_prisms = self.conn.shape[0]
        refpos = _f64(refpos).reshape(self.n_prisms, 18)
        h = _h()
        _check(L.tf_nr_create(self.n_nodes, self.n_prisms, _p(self.conn, _i), _p(refpos, _d),
                              ctypes.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            load().tf_nr_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ---- per-step inputs
    def set_material(self, E, nu, eta, ext_visc, fixed=None, pairs=None):
        E = _f64(np.broadcast_to(E, (self.n_prisms,)))
        nu = _f64(np.broadcast_to(nu, (self.n_prisms,)))
        eta = _f64(np.broadcast_to(eta, (self.n_prisms,)))
        ev = _f64(np.broadcast_to(ext_visc, (self.n_nodes, 3)))
        fx = None if fixed is None else np.ascontiguousarray(fixed, dtype=np.uint8).reshape(-1)
        pr = None if pairs is None or len(pairs) == 0 else \
            np.ascontiguousarray(pairs, dtype=np.int32).reshape(-1, 2)
        n_pairs = 0 if pr is None else pr.shape[0]
        _check(load().tf_nr_set_material(self._h, _p(E, _d), _p(nu, _d), _p(eta, _d),
                                         _p(ev, _d), _p(fx, _u8), _p(pr, _i), n_pairs))

    def set_growth(self, inv_Fg):
        a = _f64(inv_Fg).reshape(self.n_prisms, 9)
        _check(load().tf_nr_set_growth(self._h, _p(a, _d)))

    def set_visc_surface(self, area):
        a = _f64(area).reshape(self.n_nodes)
        _check(load().tf_nr_set_visc_surface(self._h, _p(a, _d)))

    def set_ablation(self, prisms=None, nodes=None):
        """Laser ablation: per-prism flags (ShapeBase::IsAblated) and the
        ablated node ids of checkJacobianForAblatedNodes; None clears."""
        pr = None if prisms is None else np.ascontiguousarray(prisms, dtype=np.uint8)
        nd = None if nodes is None else np.ascontiguousarray(nodes, dtype=np.int32).reshape(-1)
        _check(load().tf_nr_set_ablation(self._h, _p(pr, _u8), _p(nd, _i),
                                         0 if nd is None else int(nd.size)))

    def update_visc_surface(self, u, elem_placement, elem_type, spans):
        u = _f64(u).reshape(-1)
        pl = np.ascontiguousarray(elem_placement, dtype=np.int32)
        ty = np.ascontiguousarray(elem_type, dtype=np.int32)
        sp = np.ascontiguousarray(spans, dtype=np.uint8)
        out = np.empty(self.n_nodes)
        _check(load().tf_nr_update_visc_surface(self._h, _p(u, _d), _p(pl, _i), _p(ty, _i),
                                                _p(sp, _u8), _p(out, _d)))
        return out

    # ---- the step
    def step(self, u_n, dt, tol=1e-8, maxit=20, precision="fp64", krylov_rtol=1e-10,
             allow_not_converged=False):
        u_n = _f64(u_n).reshape(-1)
        out = np.empty_like(u_n)
        st = (StepStats * self.n_tissues)()
        code = load().tf_nr_step(self._h, _p(u_n, _d), float(dt), float(tol), int(maxit),
                                 PREC[precision], float(krylov_rtol), _p(out, _d), st)
        if not (code == TF_ERR_NOT_CONVERGED and allow_not_converged):
            _check(code)
        return out, self._stats(st, 0)

    def _stats(self, arr, step):
        """Stats of one step: a dict, or one dict per tissue for a batch."""
        nt = self.n_tissues
        if nt == 1:
            return arr[step].as_dict()
        return [arr[step * nt + t].as_dict() for t in range(nt)]

    # ---- pieces
    def assemble(self, uk, un, dt, precision="fp64"):
        uk = _f64(uk).reshape(-1); un = _f64(un).reshape(-1)
        g = np.empty_like(uk); gi = np.empty_like(uk)
        _check(load().tf_nr_assemble_ex(self._h, _p(uk, _d), _p(un, _d), float(dt),
                                        PREC[precision], _p(g, _d), _p(gi, _d)))
        return g, gi

    def solve(self, b, rtol=1e-10, maxit=20000):
        b = _f64(b).reshape(-1)
        x = np.empty_like(b)
        it = np.zeros(1, np.int32)
        _check(load().tf_nr_solve(self._h, _p(b, _d), _p(x, _d), float(rtol), int(maxit),
                                  _p(it, _i)))
        return x, int(it[0])

    def csr_pattern(self):
        n = 3 * self.n_nodes
        nnz = np.zeros(1, np.int64)
        _check(load().tf_csr_pattern(self._h, None, None, _p(nnz, _i64)))
        ia = np.empty(n + 1, np.int32); ja = np.empty(int(nnz[0]), np.int32)
        _check(load().tf_csr_pattern(self._h, _p(ia, _i), _p(ja, _i), _p(nnz, _i64)))
        return ia, ja

    def get_csr(self, threshold=1e-13):
        n = 3 * self.n_nodes
        ia = np.empty(n + 1, np.int32)
        nnz = np.zeros(1, np.int64)
        _check(load().tf_nr_get_csr(self._h, float(threshold), _p(ia, _i), None, None,
                                    _p(nnz, _i64)))
        ja = np.empty(int(nnz[0]), np.int32); a = np.empty(int(nnz[0]))
        _check(load().tf_nr_get_csr(self._h, float(threshold), _p(ia, _i), _p(ja, _i), _p(a, _d),
                                    _p(nnz, _i64)))
        return ia, ja, a

    # ---- device-resident time stepping
    def set_growth_rates(self, rates_xyz_per_s):
        r = _f64(rates_xyz_per_s).reshape(self.n_prisms, 3)
        _check(load().tf_nr_set_growth_rates(self._h, _p(r, _d)))

    def growth_tensor(self, Fg=None):
        if Fg is None:
            out = np.empty((self.n_prisms, 9))
            _check(load().tf_nr_growth_tensor(self._h, _p(out, _d), None))
            return out
        Fg = _f64(Fg).reshape(self.n_prisms, 9)
        _check(load().tf_nr_growth_tensor(self._h, None, _p(Fg, _d)))

    def set_surface_flags(self, elem_placement, elem_type, spans):
        pl = np.ascontiguousarray(elem_placement, dtype=np.int32)
        ty = np.ascontiguousarray(elem_type, dtype=np.int32)
        sp = np.ascontiguousarray(spans, dtype=np.uint8)
        _check(load().tf_nr_set_surface_flags(self._h, _p(pl, _i), _p(ty, _i), _p(sp, _u8)))

    # ---- Young's modulus time series (YoungsModulusModifier)
    def add_stiffness_series(self, times, grids, linear, method, applicable, rel_pos=None):
        """One series: grids (n_times, ny, nx) applied at times (s); method 0 nearest,
        1 bilinear, 2 lower corner; applicable (n_prisms) and rel_pos (n_prisms x 2)
        per prism (tissue.stiffness_applicable, tissue.relative_positions); rel_pos
        None once the growth prisms are set (the device keeps them current)."""
        times = _f64(times).reshape(-1)
        grids = _f64(grids)
        if grids.ndim != 3 or grids.shape[0] != times.size:
            raise ValueError("grids must be (n_times, ny, nx)")
        app = np.ascontiguousarray(applicable, np.uint8).reshape(self.n_prisms)
        rel = None if rel_pos is None else _f64(rel_pos).reshape(self.n_prisms, 2)
        _check(load().tf_nr_add_stiffness_series(
            self._h, times.size, _p(times, _d), grids.shape[2], grids.shape[1], _p(grids, _d),
            int(bool(linear)), int(method), _p(app, _u8), _p(rel, _d)))

    # ---- growth functions (Simulation::calculateGrowth) on the device
    def set_growth_prisms(self, tissue_type, placement, flags, rel_z, weights=None,
                          apical_of=None, node_excluded=None, symmetric_y=False,
                          explicit_ecm=False):
        """Per-prism data of the growth functions (tissue.growth_prisms builds them):
        tissue type, placement, flags (1 ECM, 2 ECM at circumference, 4 actin),
        rel_z, linker weights (n x 2), apical prism of the column (-1: own),
        nodes left out of the bounding box."""
        ne = self.n_prisms
        tt = np.ascontiguousarray(tissue_type, np.int32).reshape(ne)
        pl = np.ascontiguousarray(placement, np.int32).reshape(ne)
        fl = np.ascontiguousarray(flags, np.uint8).reshape(ne)
        rz = _f64(rel_z).reshape(ne)
        w = None if weights is None else _f64(weights).reshape(ne, 2)
        ap = None if apical_of is None else np.ascontiguousarray(apical_of, np.int32).reshape(ne)
        nx = None if node_excluded is None else \
            np.ascontiguousarray(node_excluded, np.uint8).reshape(self.n_nodes)
        _check(load().tf_nr_set_growth_prisms(self._h, _p(tt, _i), _p(pl, _i), _p(fl, _u8),
                                              _p(w, _d), _p(rz, _d), _p(ap, _i), _p(nx, _u8),
                                              int(bool(symmetric_y)), int(bool(explicit_ecm))))

    def add_uniform_growth(self, t0, t1, rate, angle, apply=1, z_range=(0.0, 1.0)):
        """UniformGrowthFunction: rate (x, y, z) in 1/s, angle in radians; apply
        1 columnar, 2 peripodial, 4 basal ECM, 8 lateral ECM."""
        r = _f64(rate).reshape(3)
        _check(load().tf_nr_add_uniform_growth(self._h, float(t0), float(t1), int(apply),
                                               float(z_range[0]), float(z_range[1]), _p(r, _d),
                                               float(angle)))

    def add_grid_growth(self, t0, t1, rates, angles_deg, apply=1, z_range=(0.0, 1.0)):
        """GridBasedGrowthFunction: rates (nx, ny, 3) in 1/s and angles (nx, ny) in
        degrees at the grid points, the first index along x."""
        r = _f64(rates)
        a = _f64(angles_deg)
        if r.ndim != 3 or r.shape[2] != 3 or a.shape != r.shape[:2]:
            raise ValueError("rates must be (nx, ny, 3) and angles (nx, ny)")
        _check(load().tf_nr_add_grid_growth(self._h, float(t0), float(t1), int(apply),
                                            float(z_range[0]), float(z_range[1]), r.shape[0],
                                            r.shape[1], _p(r, _d), _p(a, _d)))

    def set_growth_options(self, interpolation=1, pinned=False, pin_times=()):
        t = _f64(pin_times).reshape(-1)
        _check(load().tf_nr_set_growth_options(self._h, int(interpolation), int(bool(pinned)),
                                               t.size, _p(t, _d) if t.size else None))

    def set_column_volume_conservation(self, columns):
        """columns (n_columns, layers): the prisms whose column volume is conserved
        (-1 pads); an empty array turns it off."""
        c = np.ascontiguousarray(columns, np.int32)
        if c.size == 0:
            _check(load().tf_nr_set_column_volume_conservation(self._h, 0, 0, None))
            return
        c = c.reshape(c.shape[0], -1)
        _check(load().tf_nr_set_column_volume_conservation(self._h, c.shape[0], c.shape[1],
                                                           _p(c, _i)))

    def update_relative_positions(self, u=None):
        u = None if u is None else _f64(u).reshape(-1)
        _check(load().tf_nr_update_relative_positions(self._h, _p(u, _d)))

    def relative_positions(self):
        """(current, pinned) relative positions, each n_prisms x 2."""
        cur = np.empty((self.n_prisms, 2))
        pin = np.empty((self.n_prisms, 2))
        _check(load().tf_nr_relative_positions(self._h, _p(cur, _d), _p(pin, _d)))
        return cur, pin

    def grow(self, t, dt, u=None):
        """One step's growth at time t (s) from positions u (None: the device's)."""
        u = None if u is None else _f64(u).reshape(-1)
        _check(load().tf_nr_grow(self._h, _p(u, _d), float(t), float(dt)))

    def apply_stiffness_series(self, t, dt):
        """Simulation::ApplyAllYoungsModulusModifiers at time t (s)."""
        _check(load().tf_nr_apply_stiffness_series(self._h, float(t), float(dt)))

    def set_time(self, t):
        """Time (s) of the next step of advance(), which applies the series first."""
        _check(load().tf_nr_set_time(self._h, float(t)))

    def time(self):
        out = np.zeros(1)
        _check(load().tf_nr_get_time(self._h, _p(out, _d)))
        return float(out[0])

    def stiffness_multiplier(self):
        out = np.empty(self.n_prisms)
        _check(load().tf_nr_get_stiffness_multiplier(self._h, _p(out, _d)))
        return out

    def advance(self, n_steps, dt, u0=None, tol=1e-8, maxit=20, precision="fp64",
                krylov_rtol=1e-10, want_positions=True):
        """n_steps device-resident steps; returns (positions or None, per-step stats)."""
        u0 = None if u0 is None else _f64(u0).reshape(-1)
        out = np.empty(3 * self.n_nodes) if want_positions else None
        st = (StepStats * (max(n_steps, 1) * self.n_tissues))()
        code = load().tf_nr_advance(self._h, _p(u0, _d), int(n_steps), float(dt), float(tol),
                                    int(maxit), PREC[precision], float(krylov_rtol), _p(out, _d),
                                    st)
        if code not in (TF_OK, TF_ERR_NOT_CONVERGED):
            _check(code)
        return out, [self._stats(st, i) for i in range(n_steps)]

    # ---- node-node packing (Simulation::detectPacingNodes and the couple terms)
    def set_packing(self, node_placement, node_excluded=None, symmetric_y=False, capacity=2):
        """Turns packing on for the nodes' placements (0 basal, 1 apical), or
        off with node_placement None."""
        if node_placement is None:
            _check(load().tf_nr_set_packing(self._h, None, None, 0, 0))
            return
        pl = np.ascontiguousarray(node_placement, dtype=np.int32).reshape(self.n_nodes)
        ex = None if node_excluded is None else \
            np.ascontiguousarray(node_excluded, dtype=np.uint8).reshape(self.n_nodes)
        _check(load().tf_nr_set_packing(self._h, _p(pl, _i), _p(ex, _u8), int(bool(symmetric_y)),
                                        int(capacity)))

    def set_node_mass(self, mass):
        m = _f64(mass).reshape(self.n_nodes)
        _check(load().tf_nr_set_node_mass(self._h, _p(m, _d)))

    def update_node_mass(self, Fg=None):
        """Node masses from F_g (n_prisms x 9; None: the device-resident F_g)."""
        F = None if Fg is None else _f64(Fg).reshape(self.n_prisms, 9)
        out = np.empty(self.n_nodes)
        _check(load().tf_nr_update_node_mass(self._h, _p(F, _d), _p(out, _d)))
        return out

    def detect_packing(self, u=None):
        """Couples at positions u (None: the device-resident positions); returns their count."""
        u = None if u is None else _f64(u).reshape(3 * self.n_nodes)
        n = np.zeros(1, np.int32)
        _check(load().tf_nr_detect_packing(self._h, _p(u, _d), _p(n, _i)))
        return int(n[0])

    def packing_count(self):
        """Couples of the last detection."""
        n = np.zeros(1, np.int32)
        _check(load().tf_nr_get_packing(self._h, _p(n, _i), None, None, None))
        return int(n[0])

    def packing(self):
        """(couples n x 2, forces n x 6 of the last residual, packing threshold,
        squared detection thresholds 10 x 5, bounding box 6)."""
        L = load()
        n = np.zeros(1, np.int32)
        thr = np.empty(57)
        _check(L.tf_nr_get_packing(self._h, _p(n, _i), None, None, _p(thr, _d)))
        pairs = np.zeros((max(int(n[0]), 1), 2), np.int32)
        forces = np.zeros((max(int(n[0]), 1), 6))
        _check(L.tf_nr_get_packing(self._h, _p(n, _i), _p(pairs, _i), _p(forces, _d), None))
        k = int(n[0])
        return pairs[:k], forces[:k], float(thr[0]), thr[1:51].reshape(10, 5), thr[51:57]

    # ---- tiles of the fused Krylov solver
    def autotune(self, precision="mixed"):
        """Times the candidate tile sizes on the assembled operator, keeps the
        fastest; returns (slices per tile, ms per solver iteration)."""
        spt = np.zeros(1, np.int32); ms = np.zeros(1)
        _check(load().tf_nr_autotune(self._h, PREC[precision], _p(spt, _i), _p(ms, _d)))
        return int(spt[0]), float(ms[0])

    def set_tile_size(self, slices_per_tile):
        _check(load().tf_nr_set_tile_size(self._h, int(slices_per_tile)))

    def set_krylov(self, precision, method):
        """Krylov method ("bicgstab", "cg", or "gmres" for FP64) of one precision mode."""
        _check(load().tf_nr_set_krylov(self._h, PREC[precision],
                                       {"bicgstab": 0, "cg": 1, "gmres": 2}[method]))

    def set_coarse(self, cell_size=-1.0):
        """Two-level preconditioner of the mixed CG solve: aggregate cell size
        (negative: automatic, the default; 0: block Jacobi only)."""
        _check(load().tf_nr_set_coarse(self._h, float(cell_size)))

    def set_coarse_bands(self, z_bands=3):
        """Bands through the thickness per aggregate cell of the coarse space (default 3;
        1: whole columns)."""
        _check(load().tf_nr_set_coarse_bands(self._h, int(z_bands)))

    def set_coarse_lag(self, lag=True):
        """Form the coarse inverse once per step (True) or at every mixed assembly."""
        _check(load().tf_nr_set_coarse_lag(self._h, int(bool(lag))))

    def set_warm_start(self, enable=True):
        """Start each step's first solve of an advance trajectory from the
        extrapolated displacement of the last steps (True) or from zero."""
        _check(load().tf_nr_set_warm_start(self._h, int(bool(enable))))

    def set_forcing_safeguard(self, enable=True):
        """Do not oversolve a solve predicted to be a step's last (inexact Newton)."""
        _check(load().tf_nr_set_forcing_safeguard(self._h, int(bool(enable))))

    def coarse_info(self):
        """(aggregates of the last coarse build, whether the last mixed solve used it)."""
        a = np.zeros(1, np.int32); on = np.zeros(1, np.int32)
        _check(load().tf_nr_coarse_info(self._h, _p(a, _i), _p(on, _i)))
        return int(a[0]), bool(on[0])

    def set_graphs(self, enable=True):
        """Krylov iteration chunks replayed through captured hipGraphs (default on)."""
        _check(load().tf_nr_set_graphs(self._h, 1 if enable else 0))

    # ---- adjoint (implicit-function theorem per step)
    def adj_forward(self, u0, inv_Fg_steps, dt, elem_placement, elem_type, spans, tol=1e-8,
                    maxit=20, krylov_rtol=1e-10):
        u0 = _f64(u0).reshape(-1)
        inv = _f64(inv_Fg_steps).reshape(-1, self.n_prisms, 9)
        pl = np.ascontiguousarray(elem_placement, dtype=np.int32)
        ty = np.ascontiguousarray(elem_type, dtype=np.int32)
        sp = np.ascontiguousarray(spans, dtype=np.uint8)
        out = np.empty_like(u0)
        st = StepStats()
        _check(load().tf_adj_forward(self._h, inv.shape[0], _p(u0, _d), _p(inv, _d), float(dt),
                                     float(tol), int(maxit), float(krylov_rtol), _p(pl, _i),
                                     _p(ty, _i), _p(sp, _u8), _p(out, _d), ctypes.byref(st)))
        self._adj_inv = inv
        return out, st.as_dict()

    def adj_backward(self, dinv_Fg_steps, dL_du, krylov_rtol=1e-12):
        dinv = _f64(dinv_Fg_steps).reshape(self._adj_inv.shape)
        g = _f64(dL_du).reshape(-1)
        out = np.empty(self.n_prisms)
        st = StepStats()
        _check(load().tf_adj_backward(self._h, _p(self._adj_inv, _d), _p(dinv, _d), _p(g, _d),
                                      float(krylov_rtol), _p(out, _d), ctypes.byref(st)))
        return out, st.as_dict()

    def profile(self, enable=True):
        _check(load().tf_nr_profile(self._h, 1 if enable else 0))

    def kernel_times(self):
        t = KernelTimes()
        _check(load().tf_nr_kernel_times(self._h, ctypes.byref(t)))
        return dict(spmv_ms_fp64=t.spmv_ms[0], spmv_ms_fp32=t.spmv_ms[1],
                    spmv_launches_fp64=t.spmv_launches[0], spmv_launches_fp32=t.spmv_launches[1],
                    spmv_t_ms_fp64=t.spmv_t_ms[0], spmv_t_ms_fp32=t.spmv_t_ms[1],
                    spmv_t_launches_fp64=t.spmv_t_launches[0],
                    spmv_t_launches_fp32=t.spmv_t_launches[1],
                    element_ms=t.element_ms, assembly_ms=t.assembly_ms, assemblies=t.assemblies)

    def pattern_info(self):
        nb = np.zeros(1, np.int64); ns = np.zeros(1, np.int64)
        _check(load().tf_nr_pattern_info(self._h, _p(nb, _i64), _p(ns, _i64)))
        return int(nb[0]), int(ns[0])

    def element_outputs(self):
        ne = self.n_prisms
        ge = np.empty((ne, 18)); gv = np.empty((ne, 18))
        Ke = np.empty((ne, 18, 18)); Kv = np.empty((ne, 18, 18))
        _check(load().tf_nr_element_outputs(self._h, _p(ge, _d), _p(gv, _d), _p(Ke, _d),
                                            _p(Kv, _d)))
        return ge, gv, Ke, Kv

    def element_strain(self):
        """Green-Lagrange strain of each prism's last Gauss point at the
        positions of the last assembly, (n_prisms, 6) in the Voigt order
        E00, E11, E22, 2E01, 2E21, 2E02 of ShapeBase::Strain (Prism.cpp:826-837),
        the rows Simulation::writeTensionCompression saves (Simulation.cpp:8521)."""
        out = np.empty((self.n_prisms, 6))
        _check(load().tf_nr_element_strain(self._h, _p(out, _d)))
        return out


class TissueBatch(NewtonRaphsonSolver):
    """Many independent tissues in one handle (tf_batch_create).

    meshes: sequence of (n_nodes, conn, refpos) with node ids local to each
    tissue.  Per-node and per-prism arrays of every call are the tissues'
    arrays concatenated in order; step/advance return one stats dict per
    tissue.  No reference counterpart: the reference runs one tissue per
    process."""

    def __init__(self, meshes):
        L = load()
        nn = np.array([int(m[0]) for m in meshes], np.int32)
        conns = [np.ascontiguousarray(m[1], dtype=np.int32).reshape(-1, 6) for m in meshes]
        ne = np.array([c.shape[0] for c in conns], np.int32)
        self.n_tissues = len(meshes)
        self.node_offsets = np.concatenate([[0], np.cumsum(nn)]).astype(np.int64)
        self.prism_offsets = np.concatenate([[0], np.cumsum(ne)]).astype(np.int64)
        self.n_nodes = int(self.node_offsets[-1])
        self.n_prisms = int(self.prism_offsets[-1])
        self.conn = np.ascontiguousarray(np.concatenate(conns))
        refpos = np.ascontiguousarray(
            np.concatenate([_f64(m[2]).reshape(-1, 18) for m in meshes]))
        h = _h()
        _check(L.tf_batch_create(self.n_tissues, _p(nn, _i), _p(ne, _i), _p(self.conn, _i),
                                 _p(refpos, _d), ctypes.byref(h)))
        self._h = h

    def split_nodes(self, a):
        """Splits a concatenated per-node array (n_nodes x k) into tissues."""
        a = np.asarray(a).reshape(self.n_nodes, -1)
        return [a[self.node_offsets[t]:self.node_offsets[t + 1]] for t in range(self.n_tissues)]
