"""ctypes binding of the CPU FP64 oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: the checker for the HIP product.  Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg import this module; the
product (f90_dns_amd) never does.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")


class OrcParams(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
        ("xlen", ctypes.c_double), ("ylen", ctypes.c_double), ("re", ctypes.c_double), ("dt", ctypes.c_double),
        ("use_crank_nicolson", ctypes.c_int), ("use_dealias_23_rule", ctypes.c_int),
        ("flow_control_method", ctypes.c_int),
        ("target_pressure_gradient", ctypes.c_double), ("target_bulk_velocity", ctypes.c_double),
        ("controller_gain", ctypes.c_double), ("controller_update_freq", ctypes.c_int), ("istart", ctypes.c_int),
        ("u_centerline", ctypes.c_double),
        ("enable_perturbations", ctypes.c_int), ("perturbation_kind", ctypes.c_int),
        ("analytical_mode_x", ctypes.c_int), ("analytical_mode_y", ctypes.c_int),
        ("perturbation_amplitude", ctypes.c_double), ("seed", ctypes.c_ulonglong), ("noise_kmax", ctypes.c_int),
        ("refactor_every_step", ctypes.c_int), ("lgl_derivative", ctypes.c_int),
        ("use_weak_form_divergence", ctypes.c_int), ("fft_variant", ctypes.c_int),
        ("convection_scheme", ctypes.c_int), ("pressure_solver", ctypes.c_int), ("cgstol", ctypes.c_double),
    ]


class OrcPiState(ctypes.Structure):
    _fields_ = [("G", ctypes.c_double), ("Ub", ctypes.c_double), ("err", ctypes.c_double),
                ("err_int", ctypes.c_double), ("err_prev", ctypes.c_double), ("gsum", ctypes.c_double),
                ("gavg", ctypes.c_double), ("gcount", ctypes.c_int)]


class OrcDiag(ctypes.Structure):
    _fields_ = [
        ("umax", ctypes.c_double), ("vmax", ctypes.c_double), ("wmax", ctypes.c_double), ("energy", ctypes.c_double),
        ("pressure_gradient", ctypes.c_double), ("bulk_velocity", ctypes.c_double),
        ("avg_pressure_gradient", ctypes.c_double), ("bulk_error_integral", ctypes.c_double),
        ("bulk_error_prev", ctypes.c_double), ("time", ctypes.c_double), ("step", ctypes.c_int),
    ]


def build(force=False):
    """Compile oracle/liboracle.so with the Makefile beside this file."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-B" if force else "liboracle.so"], cwd=HERE, check=True)
    return LIB_PATH


_lib = None
_dp = ctypes.POINTER(ctypes.c_double)


def lib():
    global _lib
    if _lib is None:
        build()
        L = ctypes.CDLL(LIB_PATH)
        L.orc_create.restype = ctypes.c_void_p
        L.orc_create.argtypes = [ctypes.POINTER(OrcParams)]
        L.orc_destroy.argtypes = [ctypes.c_void_p]
        for name in ("orc_init_fields",):
            getattr(L, name).argtypes = [ctypes.c_void_p]
        L.orc_set_fields.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, _dp]
        L.orc_get_fields.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, _dp]
        L.orc_set_wall_gradients.argtypes = [ctypes.c_void_p, _dp, _dp]
        L.orc_get_wall_gradients.argtypes = [ctypes.c_void_p, _dp, _dp]
        L.orc_set_controller.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                         ctypes.c_double, ctypes.c_int]
        for name in ("orc_compute_source_terms", "orc_viscous_step", "orc_solve_pressure",
                     "orc_pressure_correction", "orc_apply_boundary_conditions"):
            getattr(L, name).argtypes = [ctypes.c_void_p]
        L.orc_step.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_diagnostics.argtypes = [ctypes.c_void_p, ctypes.POINTER(OrcDiag)]
        L.orc_max_divergence.argtypes = [ctypes.c_void_p]
        L.orc_max_divergence.restype = ctypes.c_double
        for name in ("orc_source_terms", "orc_star_fields"):
            getattr(L, name).argtypes = [ctypes.c_void_p, ctypes.c_int]
            getattr(L, name).restype = _dp
        L.orc_status.argtypes = [ctypes.c_void_p]
        L.orc_diff_matrix_exact.argtypes = [ctypes.c_int, _dp, _dp]
        L.orc_phi.argtypes = [ctypes.c_void_p]
        L.orc_phi.restype = _dp
        L.orc_lgl.argtypes = [ctypes.c_int, _dp, _dp]
        L.orc_diff_matrix.argtypes = [ctypes.c_int, _dp, _dp]
        L.orc_wavenumbers.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double, _dp, _dp]
        L.orc_dealias_mask.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_ubyte)]
        L.orc_fft_r2c.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, _dp]
        L.orc_fft_c2r.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, _dp]
        L.orc_make_initial_field.argtypes = [ctypes.POINTER(OrcParams), _dp, _dp, _dp]
        L.orc_pi_update.argtypes = [ctypes.POINTER(OrcParams), ctypes.POINTER(OrcPiState), ctypes.c_int,
                                    ctypes.c_double]
        L.orc_pi_update.restype = ctypes.c_int
        L.orc_bulk_velocity.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp]
        L.orc_bulk_velocity.restype = ctypes.c_double
        L.orc_jacgs.argtypes = [ctypes.c_int, _dp, _dp, _dp, ctypes.c_int, ctypes.c_double, ctypes.POINTER(ctypes.c_int)]
        L.orc_jacgs.restype = ctypes.c_int
        L.orc_jacgs_stats.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(ctypes.c_int)]
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(_dp)


# Config dictionaries use the reference's namelist names (DNS_pressure_BC_3D.f90:559-568).
DEFAULTS = dict(
    nx=32, ny=32, nz=33, xlen=2 * np.pi, ylen=np.pi, re=180.0, dt=0.002,
    use_crank_nicolson=1, use_dealias_23_rule=1, flow_control_method=1,
    target_pressure_gradient=2.0 / 180.0, target_bulk_velocity=2.0 / 3.0, controller_gain=0.15,
    controller_update_freq=7, istart=0, u_centerline=1.0,
    enable_perturbations=1, perturbation_kind=0, analytical_mode_x=1, analytical_mode_y=1,
    perturbation_amplitude=0.01, seed=0, noise_kmax=8, refactor_every_step=0,
    lgl_derivative=0, use_weak_form_divergence=0, fft_variant=0, convection_scheme=0, pressure_solver=0,
    cgstol=1e-6,
)


def make_params(**kw):
    d = dict(DEFAULTS)
    unknown = set(kw) - set(d)
    if unknown:
        raise KeyError(f"unknown oracle parameters {sorted(unknown)}")
    d.update(kw)
    p = OrcParams()
    for k, v in d.items():
        setattr(p, k, v)
    return p


class Oracle:
    """One oracle run: fields in Fortran order as numpy arrays of shape (nz, ny, nx)."""

    def __init__(self, **kw):
        self.params = make_params(**kw)
        self.nx, self.ny, self.nz = self.params.nx, self.params.ny, self.params.nz
        self._h = lib().orc_create(ctypes.byref(self.params))
        if not self._h:
            raise ValueError("oracle rejected the grid")

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().orc_destroy(h)
            self._h = None

    @property
    def shape(self):
        return (self.nz, self.ny, self.nx)

    def init_fields(self):
        lib().orc_init_fields(self._h)

    def set_fields(self, u=None, v=None, w=None, p=None):
        arrs = [None if a is None else np.ascontiguousarray(a, dtype=np.float64).reshape(self.shape)
                for a in (u, v, w, p)]
        lib().orc_set_fields(self._h, *[None if a is None else _ptr(a) for a in arrs])

    def get_fields(self):
        out = [np.empty(self.shape) for _ in range(4)]
        lib().orc_get_fields(self._h, *[_ptr(a) for a in out])
        return out

    def set_wall_gradients(self, gpx, gpy):
        gx = np.ascontiguousarray(gpx, dtype=np.float64).reshape(2, self.ny, self.nx)
        gy = np.ascontiguousarray(gpy, dtype=np.float64).reshape(2, self.ny, self.nx)
        lib().orc_set_wall_gradients(self._h, _ptr(gx), _ptr(gy))

    def get_wall_gradients(self):
        gx = np.empty((2, self.ny, self.nx)); gy = np.empty((2, self.ny, self.nx))
        lib().orc_get_wall_gradients(self._h, _ptr(gx), _ptr(gy))
        return gx, gy

    def set_controller(self, G, integral=0.0, prev=0.0, gsum=0.0, gcount=0):
        lib().orc_set_controller(self._h, G, integral, prev, gsum, gcount)

    def compute_source_terms(self):
        lib().orc_compute_source_terms(self._h)

    def viscous_step(self):
        lib().orc_viscous_step(self._h)

    def solve_pressure(self):
        lib().orc_solve_pressure(self._h)

    def pressure_correction(self):
        lib().orc_pressure_correction(self._h)

    def apply_boundary_conditions(self):
        lib().orc_apply_boundary_conditions(self._h)

    def step(self, n=1):
        lib().orc_step(self._h, n)

    def _view(self, p):
        return np.ctypeslib.as_array(p, shape=(self.nz * self.ny * self.nx,)).reshape(self.shape).copy()

    def source_terms(self):
        return [self._view(lib().orc_source_terms(self._h, c)) for c in range(3)]

    def star_fields(self):
        return [self._view(lib().orc_star_fields(self._h, c)) for c in range(3)]

    def phi(self):
        return self._view(lib().orc_phi(self._h))

    def diagnostics(self):
        d = OrcDiag()
        lib().orc_diagnostics(self._h, ctypes.byref(d))
        return {f: getattr(d, f) for f, _ in OrcDiag._fields_}

    def status(self):
        return lib().orc_status(self._h)

    def max_divergence(self):
        return lib().orc_max_divergence(self._h)

    def jacgs_stats(self):
        it = ctypes.c_longlong(); nf = ctypes.c_int()
        lib().orc_jacgs_stats(self._h, ctypes.byref(it), ctypes.byref(nf))
        return it.value, nf.value


def lgl(n):
    x = np.empty(n); w = np.empty(n)
    lib().orc_lgl(n, _ptr(x), _ptr(w))
    return x, w


def diff_matrix(n):
    x, _ = lgl(n)
    d = np.empty(n * n)
    lib().orc_diff_matrix(n, _ptr(x), _ptr(d))
    return d.reshape(n, n).T.copy()       # D[i, j] = D(i, j)


def diff_matrix_exact(n):
    x, _ = lgl(n)
    d = np.empty(n * n)
    lib().orc_diff_matrix_exact(n, _ptr(x), _ptr(d))
    return d.reshape(n, n).T.copy()


def wavenumbers(nx, ny, xlen, ylen):
    xw = np.empty(nx // 2 + 1); yw = np.empty(ny)
    lib().orc_wavenumbers(nx, ny, xlen, ylen, _ptr(xw), _ptr(yw))
    return xw, yw


def dealias_mask(nx, ny):
    m = np.empty((ny, nx // 2 + 1), dtype=np.uint8)
    lib().orc_dealias_mask(nx, ny, m.ctypes.data_as(ctypes.POINTER(ctypes.c_ubyte)))
    return m


def fft_r2c(f):
    """f: (nplanes, ny, nx) real -> (nplanes, ny, nx//2+1) complex, unnormalised."""
    f = np.ascontiguousarray(f, dtype=np.float64)
    npl, ny, nx = f.shape
    out = np.empty((npl, ny, nx // 2 + 1), dtype=np.complex128)
    if lib().orc_fft_r2c(nx, ny, npl, _ptr(f), out.ctypes.data_as(_dp)):
        raise ValueError("odd transform size")
    return out


def fft_c2r(F, nx):
    """F: (nplanes, ny, nx//2+1) complex -> (nplanes, ny, nx) real, divided by nx*ny."""
    F = np.ascontiguousarray(F, dtype=np.complex128)
    npl, ny, _ = F.shape
    out = np.empty((npl, ny, nx))
    if lib().orc_fft_c2r(nx, ny, npl, F.ctypes.data_as(_dp), _ptr(out)):
        raise ValueError("odd transform size")
    return out


def initial_field(**kw):
    p = make_params(**kw)
    shape = (p.nz, p.ny, p.nx)
    u, v, w = (np.empty(shape) for _ in range(3))
    lib().orc_make_initial_field(ctypes.byref(p), _ptr(u), _ptr(v), _ptr(w))
    return u, v, w


class PIController:
    """update_pressure_gradient_3d on a prescribed bulk-velocity sequence (no flow field):
    the state starts as read_input_3d leaves it (G = average = target_pressure_gradient)."""

    def __init__(self, **kw):
        self.params = make_params(**kw)
        self.state = OrcPiState(G=self.params.target_pressure_gradient, gavg=self.params.target_pressure_gradient)

    def update(self, istep, ub):
        ran = lib().orc_pi_update(ctypes.byref(self.params), ctypes.byref(self.state), istep, ub)
        return bool(ran)

    def __getattr__(self, name):
        st = self.__dict__.get("state")
        if st is not None and name in dict(OrcPiState._fields_):
            return getattr(st, name)
        raise AttributeError(name)


def bulk_velocity(u):
    """calculate_bulk_velocity_3d of u with shape (nz, ny, nx)."""
    u = np.ascontiguousarray(u, dtype=np.float64)
    nz, ny, nx = u.shape
    return lib().orc_bulk_velocity(nx, ny, nz, _ptr(u))


def jacgs(a, b, niter=500, tol=1e-6, x0=None):
    """orc_jacgs on a real (n, n) system: (x, info, iterations)."""
    a = np.asfortranarray(a, dtype=np.float64)
    n = a.shape[0]
    b = np.ascontiguousarray(b, dtype=np.float64)
    x = np.zeros(n) if x0 is None else np.array(x0, dtype=np.float64)
    it = ctypes.c_int()
    info = lib().orc_jacgs(n, a.ctypes.data_as(_dp), _ptr(x), _ptr(b), niter, tol, ctypes.byref(it))
    return x, info, it.value
