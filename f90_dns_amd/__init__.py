"""f90_dns_amd -- MI355X (gfx950) 3D channel-DNS time step behind a C ABI.

This module is the Python host binding of include/f90dns.h (ctypes, no torch types).  It
mirrors the reference's run configuration (the namelist variables of
DNS_pressure_BC_3D.f90:559-568) and its stage subroutines (:314-326).  Every entry point
runs on the GPU through libf90dns.so; if the library or a gfx950 device is missing, the
calls raise -- there is no CPU fallback.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libf90dns.so")

DNS_PREC_FP64 = 0
DNS_PREC_MIXED = 1
DNS_PREC_MIXED_F32D = 2   # mixed, and the vorticity's D1 contraction on FP32 MFMA (FP32 accumulation)
_ERRORS = {1: "EINVAL", 2: "ENOMEM", 3: "EDEVICE", 4: "ENONFINITE", 5: "ESOLVER", 6: "EUNSUPPORTED"}


class DnsError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{_ERRORS.get(code, code)}: {msg}")
        self.code = code


class DnsParams(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
        ("xlen", ctypes.c_double), ("ylen", ctypes.c_double), ("re", ctypes.c_double), ("dt", ctypes.c_double),
        ("use_crank_nicolson", ctypes.c_int), ("use_dealias_23_rule", ctypes.c_int),
        ("use_weak_form_divergence", ctypes.c_int), ("flow_control_method", ctypes.c_int),
        ("target_pressure_gradient", ctypes.c_double), ("target_bulk_velocity", ctypes.c_double),
        ("controller_gain", ctypes.c_double), ("controller_update_freq", ctypes.c_int), ("istart", ctypes.c_int),
        ("u_centerline", ctypes.c_double),
        ("enable_perturbations", ctypes.c_int), ("perturbation_kind", ctypes.c_int),
        ("analytical_mode_x", ctypes.c_int), ("analytical_mode_y", ctypes.c_int),
        ("perturbation_amplitude", ctypes.c_double), ("seed", ctypes.c_ulonglong), ("noise_kmax", ctypes.c_int),
        ("precision", ctypes.c_int), ("lgl_derivative", ctypes.c_int),
        ("convection_scheme", ctypes.c_int), ("pressure_solver", ctypes.c_int), ("cgstol", ctypes.c_double),
    ]


class DnsDiag(ctypes.Structure):
    _fields_ = [
        ("umax", ctypes.c_double), ("vmax", ctypes.c_double), ("wmax", ctypes.c_double), ("energy", ctypes.c_double),
        ("pressure_gradient", ctypes.c_double), ("bulk_velocity", ctypes.c_double),
        ("avg_pressure_gradient", ctypes.c_double), ("bulk_error_integral", ctypes.c_double),
        ("bulk_error_prev", ctypes.c_double), ("time", ctypes.c_double), ("step", ctypes.c_int),
    ]


class ServerStats(ctypes.Structure):
    _fields_ = [
        ("ticks", ctypes.c_longlong), ("job_steps", ctypes.c_longlong), ("grid_point_steps", ctypes.c_double),
        ("device_ms", ctypes.c_double), ("step_p50_ms", ctypes.c_double), ("step_p99_ms", ctypes.c_double),
        ("bytes", ctypes.c_size_t), ("peak_bytes", ctypes.c_size_t), ("max_running", ctypes.c_int),
        ("queued", ctypes.c_int), ("running", ctypes.c_int), ("done", ctypes.c_int), ("failed", ctypes.c_int),
        ("rejected", ctypes.c_int), ("changes", ctypes.c_longlong), ("captures", ctypes.c_longlong),
        ("change_ms", ctypes.c_double), ("capture_ms", ctypes.c_double),
    ]


JOB_STATES = {0: "queued", 1: "running", 2: "done", 3: "failed", 4: "rejected"}


# reference defaults (read_input_3d, DNS_pressure_BC_3D.f90:571-603) where the 3D step reads them
DEFAULTS = dict(
    nx=128, ny=32, nz=33, xlen=2 * np.pi, ylen=2 * np.pi, re=180.0, dt=0.01,
    use_crank_nicolson=1, use_dealias_23_rule=0, use_weak_form_divergence=0, flow_control_method=1,
    target_pressure_gradient=0.03, target_bulk_velocity=1.0, controller_gain=0.15, controller_update_freq=7,
    istart=0, u_centerline=1.5, enable_perturbations=0, perturbation_kind=0, analytical_mode_x=1,
    analytical_mode_y=1, perturbation_amplitude=0.01, seed=0, noise_kmax=8, precision=DNS_PREC_FP64,
    lgl_derivative=0, convection_scheme=0, pressure_solver=0, cgstol=1e-6,
)

_lib = None
_dp = ctypes.POINTER(ctypes.c_double)
_vp = ctypes.c_void_p


def library_path():
    return LIB_PATH


def lib():
    """Load libf90dns.so (built in-tree by f90_dns_amd/build.py); raise if it is missing."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise DnsError(3, f"{LIB_PATH} is not built; run python -m f90_dns_amd.build")
        L = ctypes.CDLL(LIB_PATH)
        L.dns_create.argtypes = [ctypes.POINTER(DnsParams), ctypes.POINTER(_vp)]
        L.dns_destroy.argtypes = [_vp]
        L.dns_last_error.argtypes = [_vp]
        L.dns_last_error.restype = ctypes.c_char_p
        L.dns_version.restype = ctypes.c_char_p
        for n in ("dns_init_fields", "dns_begin_step", "dns_compute_source_terms", "dns_viscous_step",
                  "dns_solve_pressure", "dns_pressure_correction", "dns_apply_boundary_conditions",
                  "dns_synchronize"):
            getattr(L, n).argtypes = [_vp]
        for n in ("dns_set_fields", "dns_get_fields", "dns_set_fields_device", "dns_get_fields_device"):
            getattr(L, n).argtypes = [_vp, _vp, _vp, _vp, _vp]
        L.dns_set_controller.argtypes = [_vp, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                         ctypes.c_int]
        L.dns_step.argtypes = [_vp, ctypes.c_int]
        L.dns_step_graph.argtypes = [_vp, ctypes.c_int]
        L.dns_diagnostics.argtypes = [_vp, ctypes.POINTER(DnsDiag)]
        L.dns_check_divergence.argtypes = [_vp, _dp]
        L.dns_step_timing.argtypes = [_vp, _dp]
        L.dns_stream.argtypes = [_vp]
        L.dns_stream.restype = _vp
        L.dns_fft_plan_create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(_vp)]
        L.dns_fft_plan_destroy.argtypes = [_vp]
        for n in ("dns_fft_forward", "dns_fft_backward"):
            getattr(L, n).argtypes = [_vp, _vp, _vp, ctypes.c_int]
        for n in ("dns_fft_forward_device", "dns_fft_backward_device"):
            getattr(L, n).argtypes = [_vp, _vp, _vp, ctypes.c_int, _vp]
        L.dns_fft_execute_device_noalloc.argtypes = [_vp, ctypes.c_int, _vp, _vp, ctypes.c_int, _vp]
        L.dns_debug_spectral.argtypes = [_vp, ctypes.c_int, _vp]
        L.dns_debug_zderiv.argtypes = [_vp, _vp, _vp]
        L.dns_debug_mfma_probe.argtypes = [_dp]
        L.dns_debug_mfma_probe_f32.argtypes = [_dp]
        L.dns_debug_operator_error.argtypes = [_vp, _dp]
        L.dns_debug_zmap.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_int), _vp, _vp, _vp]
        L.dns_debug_wavenumbers.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double, _vp, _vp]
        L.dns_debug_dealias_mask.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _vp]
        L.dns_debug_context_wavenumbers.argtypes = [_vp, _vp, _vp]
        L.dns_debug_jacgs_stats.argtypes = [_vp, ctypes.POINTER(ctypes.c_ulonglong), ctypes.POINTER(ctypes.c_int)]
        L.dns_profile_step.argtypes = [_vp, ctypes.c_int, _dp]
        L.dns_profile_kernel_name.argtypes = [ctypes.c_int]
        L.dns_profile_kernel_name.restype = ctypes.c_char_p
        L.dns_create_batch.argtypes = [ctypes.POINTER(DnsParams), ctypes.c_int, ctypes.POINTER(_vp)]
        L.dns_members.argtypes = [_vp]
        L.dns_device_bytes.argtypes = [_vp]
        L.dns_device_bytes.restype = ctypes.c_size_t
        for n in ("dns_set_member_fields", "dns_get_member_fields"):
            getattr(L, n).argtypes = [_vp, ctypes.c_int, _vp, _vp, _vp, _vp]
        L.dns_set_member_controller.argtypes = [_vp, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                                ctypes.c_double, ctypes.c_double, ctypes.c_int]
        L.dns_member_diagnostics.argtypes = [_vp, ctypes.c_int, ctypes.POINTER(DnsDiag)]
        L.dns_member_divergence.argtypes = [_vp, ctypes.c_int, _dp]
        L.dns_set_member_params.argtypes = [_vp, ctypes.c_int, ctypes.POINTER(DnsParams)]
        L.dns_init_member_fields.argtypes = [_vp, ctypes.c_int]
        L.dns_ensemble_create.argtypes = [ctypes.POINTER(DnsParams), ctypes.c_int, ctypes.POINTER(_vp)]
        L.dns_ensemble_destroy.argtypes = [_vp]
        L.dns_ensemble_last_error.argtypes = [_vp]
        L.dns_ensemble_last_error.restype = ctypes.c_char_p
        for n in ("dns_ensemble_members", "dns_ensemble_groups", "dns_ensemble_init_fields",
                  "dns_ensemble_synchronize"):
            getattr(L, n).argtypes = [_vp]
        L.dns_ensemble_device_bytes.argtypes = [_vp]
        L.dns_ensemble_device_bytes.restype = ctypes.c_size_t
        L.dns_ensemble_member_shape.argtypes = [_vp, ctypes.c_int, _vp, _vp, _vp]
        for n in ("dns_ensemble_set_member_fields", "dns_ensemble_get_member_fields"):
            getattr(L, n).argtypes = [_vp, ctypes.c_int, _vp, _vp, _vp, _vp]
        L.dns_ensemble_member_diagnostics.argtypes = [_vp, ctypes.c_int, ctypes.POINTER(DnsDiag)]
        L.dns_ensemble_step.argtypes = [_vp, ctypes.c_int, ctypes.c_int]
        L.dns_ensemble_step_timing.argtypes = [_vp, _dp]
        L.dns_ensemble_set_tuning.argtypes = [_vp, ctypes.c_int, ctypes.c_int, _vp, _vp]
        L.dns_ensemble_autotune.argtypes = [_vp, ctypes.c_int, ctypes.c_char_p, ctypes.c_size_t]
        L.dns_ensemble_set_tile_waves.argtypes = [_vp, _vp]
        L.dns_ensemble_set_promotion.argtypes = [_vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.c_int]
        L.dns_ensemble_launch_counts.argtypes = [_vp, ctypes.POINTER(ctypes.c_int)]
        L.dns_estimate_bytes.argtypes = [ctypes.POINTER(DnsParams), ctypes.c_int, ctypes.POINTER(ctypes.c_size_t)]
        L.dns_server_create.argtypes = [ctypes.c_size_t, ctypes.c_int, ctypes.c_int, ctypes.POINTER(_vp)]
        L.dns_server_destroy.argtypes = [_vp]
        L.dns_server_last_error.argtypes = [_vp]
        L.dns_server_last_error.restype = ctypes.c_char_p
        L.dns_server_submit.argtypes = [_vp, ctypes.POINTER(DnsParams), ctypes.c_int, ctypes.c_int,
                                        ctypes.POINTER(ctypes.c_int)]
        L.dns_server_tick.argtypes = [_vp, ctypes.POINTER(ctypes.c_int)]
        L.dns_server_drain.argtypes = [_vp]
        L.dns_server_job_status.argtypes = [_vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                            ctypes.POINTER(ctypes.c_int)]
        L.dns_server_job_message.argtypes = [_vp, ctypes.c_int]
        L.dns_server_job_message.restype = ctypes.c_char_p
        L.dns_server_job_result.argtypes = [_vp, ctypes.c_int, _vp, _vp, _vp, _vp, ctypes.POINTER(DnsDiag)]
        L.dns_server_job_times.argtypes = [_vp, ctypes.c_int, _dp]
        L.dns_server_get_stats.argtypes = [_vp, ctypes.POINTER(ServerStats)]
        L.dns_server_set_health_interval.argtypes = [_vp, ctypes.c_int]
        L.dns_server_release_idle.argtypes = [_vp]
        L.dns_server_pending.argtypes = [_vp]
        L.dns_debug_copy.argtypes = [ctypes.c_size_t, ctypes.c_int, _dp]
        L.dns_debug_initial_field.argtypes = [ctypes.POINTER(DnsParams), _vp, _vp, _vp]
        L.dns_enqueue_step.argtypes = [_vp, ctypes.c_int]
        L.dns_query.argtypes = [_vp]
        _lib = L
    return _lib


def make_params(**kw):
    d = dict(DEFAULTS)
    unknown = set(kw) - set(d)
    if unknown:
        raise KeyError(f"unknown parameters {sorted(unknown)}")
    d.update(kw)
    p = DnsParams()
    for k, v in d.items():
        setattr(p, k, v)
    return p


def _arr(a, shape):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if a.size != int(np.prod(shape)):
        raise ValueError(f"expected {shape} values, got {a.shape}")
    return a


class Simulation:
    """One resident context (a dns_ctx).  Physical fields are numpy (nz, ny, nx) arrays in the
    reference's Fortran order f(i,j,k).

    members: optional list of per-member parameter overrides (dicts) on top of **kw; the
    context then advances len(members) flows of one shape together (dns_create_batch).  The
    field, controller and diagnostics methods take the member index (default 0)."""

    def __init__(self, members=None, **kw):
        self.params = make_params(**kw)
        if members is None:
            plist = [self.params]
        else:
            plist = [make_params(**{**kw, **m}) for m in members]
            self.params = plist[0]
        self.member_params = plist
        self.members = len(plist)
        self.nx, self.ny, self.nz = self.params.nx, self.params.ny, self.params.nz
        h = _vp()
        L = lib()
        arr = (DnsParams * len(plist))(*plist)
        rc = L.dns_create_batch(arr, len(plist), ctypes.byref(h))
        if rc != 0:
            raise DnsError(rc, L.dns_last_error(None).decode())
        self._h = h

    @property
    def device_bytes(self):
        return lib().dns_device_bytes(self._h)

    def _check(self, rc):
        if rc != 0:
            raise DnsError(rc, lib().dns_last_error(self._h).decode())

    def close(self):
        if getattr(self, "_h", None):
            lib().dns_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def shape(self):
        return (self.nz, self.ny, self.nx)

    @property
    def spectral_shape(self):
        return (self.nz, self.ny, self.nx // 2 + 1)

    def init_fields(self):
        self._check(lib().dns_init_fields(self._h))

    def set_fields(self, u=None, v=None, w=None, p=None, member=0):
        arrs = [None if a is None else _arr(a, self.shape) for a in (u, v, w, p)]
        self._check(lib().dns_set_member_fields(self._h, member,
                                                *[None if a is None else a.ctypes.data for a in arrs]))

    def get_fields(self, member=0):
        out = [np.empty(self.shape) for _ in range(4)]
        self._check(lib().dns_get_member_fields(self._h, member, *[a.ctypes.data for a in out]))
        return out

    def set_controller(self, G, integral=0.0, prev=0.0, gsum=0.0, gcount=0, member=0):
        self._check(lib().dns_set_member_controller(self._h, member, G, integral, prev, gsum, gcount))

    # the stage subroutines, DNS_pressure_BC_3D.f90:314-326
    def begin_step(self):
        self._check(lib().dns_begin_step(self._h))

    def compute_source_terms(self):
        self._check(lib().dns_compute_source_terms(self._h))

    def viscous_step(self):
        self._check(lib().dns_viscous_step(self._h))

    def solve_pressure(self):
        self._check(lib().dns_solve_pressure(self._h))

    def pressure_correction(self):
        self._check(lib().dns_pressure_correction(self._h))

    def apply_boundary_conditions(self):
        self._check(lib().dns_apply_boundary_conditions(self._h))

    def step(self, n=1, graph=False):
        L = lib()
        self._check((L.dns_step_graph if graph else L.dns_step)(self._h, n))

    def synchronize(self):
        self._check(lib().dns_synchronize(self._h))

    def step_timing(self):
        out = (ctypes.c_double * 4)()
        self._check(lib().dns_step_timing(self._h, out))
        return dict(min_ms=out[0], p50_ms=out[1], p99_ms=out[2], mean_ms=out[3])

    def enqueue_step(self, graph=True):
        """One step enqueued on the context stream; returns without waiting."""
        self._check(lib().dns_enqueue_step(self._h, 1 if graph else 0))

    def busy(self):
        rc = lib().dns_query(self._h)
        if rc not in (0, 1):
            self._check(rc)
        return rc == 1

    def diagnostics(self, member=0):
        d = DnsDiag()
        self._check(lib().dns_member_diagnostics(self._h, member, ctypes.byref(d)))
        return {f: getattr(d, f) for f, _ in DnsDiag._fields_}

    def max_divergence(self, member=0):
        v = ctypes.c_double()
        self._check(lib().dns_member_divergence(self._h, member, ctypes.byref(v)))
        return v.value

    def spectral(self, which):
        """Copy a resident spectral buffer (see include/f90dns_debug.h) as complex (nz, ny, nx/2+1)."""
        out = np.empty(self.spectral_shape, dtype=np.complex128)
        self._check(lib().dns_debug_spectral(self._h, which, out.ctypes.data))
        return out

    def zderiv(self, spec):
        a = np.ascontiguousarray(spec, dtype=np.complex128).reshape(self.spectral_shape)
        out = np.empty_like(a)
        self._check(lib().dns_debug_zderiv(self._h, a.ctypes.data, out.ctypes.data))
        return out

    def profile_step(self, nsteps):
        """Mean per-kernel device time (ms) over nsteps eager steps, events between launches."""
        out = (ctypes.c_double * 7)()
        self._check(lib().dns_profile_step(self._h, nsteps, out))
        return {lib().dns_profile_kernel_name(i).decode(): out[i] for i in range(7)}

    def operator_error(self):
        out = (ctypes.c_double * 2)()
        self._check(lib().dns_debug_operator_error(self._h, out))
        return out[0], out[1]

    def jacgs_stats(self):
        """(total iterations, non-converged systems) of the last jacgs pressure solve."""
        it = ctypes.c_ulonglong(); nf = ctypes.c_int()
        self._check(lib().dns_debug_jacgs_stats(self._h, ctypes.byref(it), ctypes.byref(nf)))
        return it.value, nf.value

    def device_wavenumbers(self):
        """The context's device-resident xw, yw tables, copied back."""
        xw = np.empty(self.nx // 2 + 1); yw = np.empty(self.ny)
        self._check(lib().dns_debug_context_wavenumbers(self._h, xw.ctypes.data, yw.ctypes.data))
        return xw, yw


def host_initial_field(**kw):
    """initialize_fields_3d evaluated point by point on the host (include/f90dns_debug.h)."""
    p = make_params(**kw)
    out = [np.empty((p.nz, p.ny, p.nx)) for _ in range(3)]
    rc = lib().dns_debug_initial_field(ctypes.byref(p), *[a.ctypes.data for a in out])
    if rc != 0:
        raise DnsError(rc, lib().dns_last_error(None).decode())
    return out


def zmap(nx, ny):
    """The mode space of the per-mode solves (include/f90dns_debug.h dns_debug_zmap; host
    only): (so, s2o, s2s) arrays over its real columns."""
    H2 = 2 * (nx // 2 + 1) * ny
    so = np.empty(H2, dtype=np.int32)
    s2o = np.empty(H2, dtype=np.int32)
    s2s = np.empty(H2)
    n = ctypes.c_int()
    rc = lib().dns_debug_zmap(nx, ny, ctypes.byref(n), so.ctypes.data, s2o.ctypes.data, s2s.ctypes.data)
    if rc != 0:
        raise DnsError(rc, "dns_debug_zmap")
    return so[:n.value], s2o[:n.value], s2s[:n.value]


def wavenumbers(nx, ny, xlen, ylen):
    """xw(nx/2+1), yw(ny) as the library computes and uploads them (host only)."""
    xw = np.empty(nx // 2 + 1); yw = np.empty(ny)
    rc = lib().dns_debug_wavenumbers(nx, ny, xlen, ylen, xw.ctypes.data, yw.ctypes.data)
    if rc != 0:
        raise DnsError(rc, "dns_debug_wavenumbers")
    return xw, yw


def dealias_mask(nx, ny, which=1):
    """The 2/3 mask (ny, nx/2+1), 1 = kept: which 0 = the two-pass transforms' table, 1 = the
    plane kernels' bounds rule (host only)."""
    m = np.empty((ny, nx // 2 + 1), dtype=np.uint8)
    rc = lib().dns_debug_dealias_mask(nx, ny, which, m.ctypes.data)
    if rc != 0:
        raise DnsError(rc, "dns_debug_dealias_mask")
    return m


def estimate_bytes(nmembers=1, **kw):
    """HBM bytes a context of nmembers members of this shape holds (before creating it)."""
    p = make_params(**kw)
    out = ctypes.c_size_t()
    rc = lib().dns_estimate_bytes(ctypes.byref(p), nmembers, ctypes.byref(out))
    if rc != 0:
        raise DnsError(rc, lib().dns_ensemble_last_error(None).decode())
    return out.value


class Ensemble:
    """Members of different grids (include/f90dns_ensemble.h): one launch per stage and kernel
    class for all of them, the step replayed from one captured graph.  members: list of
    per-member parameter dicts on top of **common."""

    def __init__(self, members, **common):
        self.member_params = [make_params(**{**common, **m}) for m in members]
        arr = (DnsParams * len(members))(*self.member_params)
        h = _vp()
        L = lib()
        rc = L.dns_ensemble_create(arr, len(members), ctypes.byref(h))
        if rc != 0:
            raise DnsError(rc, L.dns_ensemble_last_error(None).decode())
        self._h = h
        self.members = len(members)

    def _check(self, rc):
        if rc != 0:
            raise DnsError(rc, lib().dns_ensemble_last_error(self._h).decode())

    def close(self):
        if getattr(self, "_h", None):
            lib().dns_ensemble_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def groups(self):
        return lib().dns_ensemble_groups(self._h)

    @property
    def device_bytes(self):
        return lib().dns_ensemble_device_bytes(self._h)

    def shape(self, m):
        nx, ny, nz = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        self._check(lib().dns_ensemble_member_shape(self._h, m, ctypes.byref(nx), ctypes.byref(ny), ctypes.byref(nz)))
        return (nz.value, ny.value, nx.value)

    def points(self, m):
        nz, ny, nx = self.shape(m)
        return nx * ny * nz

    def init_fields(self):
        self._check(lib().dns_ensemble_init_fields(self._h))

    def set_fields(self, m, u=None, v=None, w=None, p=None):
        sh = self.shape(m)
        arrs = [None if a is None else _arr(a, sh) for a in (u, v, w, p)]
        self._check(lib().dns_ensemble_set_member_fields(self._h, m,
                                                         *[None if a is None else a.ctypes.data for a in arrs]))

    def get_fields(self, m):
        out = [np.empty(self.shape(m)) for _ in range(4)]
        self._check(lib().dns_ensemble_get_member_fields(self._h, m, *[a.ctypes.data for a in out]))
        return out

    def diagnostics(self, m):
        d = DnsDiag()
        self._check(lib().dns_ensemble_member_diagnostics(self._h, m, ctypes.byref(d)))
        return {f: getattr(d, f) for f, _ in DnsDiag._fields_}

    def step(self, n=1, graph=True):
        self._check(lib().dns_ensemble_step(self._h, n, 1 if graph else 0))

    def synchronize(self):
        self._check(lib().dns_ensemble_synchronize(self._h))

    def step_timing(self):
        out = (ctypes.c_double * 4)()
        self._check(lib().dns_ensemble_step_timing(self._h, out))
        return dict(min_ms=out[0], p50_ms=out[1], p99_ms=out[2], mean_ms=out[3])

    def set_tuning(self, lanes=1, interleave=0, promote_fp64=(0, 0, 0, 0), promote_fp32=(0, 0, 0, 0)):
        a = (ctypes.c_int * 4)(*(list(promote_fp64) + [0] * 4)[:4])
        b = (ctypes.c_int * 4)(*(list(promote_fp32) + [0] * 4)[:4])
        self._check(lib().dns_ensemble_set_tuning(self._h, lanes, interleave, a, b))

    # plane-transform size classes in the order promotion walks (include/f90dns_ensemble.h)
    SIZE_CLASS_ORDER = (0, 1, 2, 5, 6, 3, 4)

    def set_promotion(self, fp32, levels):
        """Promotion levels of the FP64 (fp32=0) or FP32 (fp32=1) plane kernels per size class."""
        a = (ctypes.c_int * max(1, len(levels)))(*levels)
        self._check(lib().dns_ensemble_set_promotion(self._h, int(fp32), a, len(levels)))

    def launch_counts(self):
        """Launches of one step per stage: vorticity, c2r, r2c, Helmholtz, Poisson."""
        c = (ctypes.c_int * 5)()
        self._check(lib().dns_ensemble_launch_counts(self._h, c))
        return dict(zip(("zderiv", "c2r", "r2c", "helmholtz", "poisson"), list(c)))

    def set_tile_waves(self, zderiv=0, helmholtz=0, poisson=0):
        """Least waves per column block of the wall-normal launches (0 = natural, 2 or 4)."""
        w = (ctypes.c_int * 3)(zderiv, helmholtz, poisson)
        self._check(lib().dns_ensemble_set_tile_waves(self._h, w))

    def autotune(self, steps_per_trial=5):
        import json
        buf = ctypes.create_string_buffer(1 << 16)
        self._check(lib().dns_ensemble_autotune(self._h, steps_per_trial, buf, len(buf)))
        return json.loads(buf.value.decode())


class Server:
    """Serving loop (include/f90dns_server.h): jobs admitted by HBM budget into slots of a
    ragged ensemble, every running job advanced by each tick."""

    def __init__(self, hbm_budget_bytes, slots_per_group=4, graph=True):
        h = _vp()
        rc = lib().dns_server_create(int(hbm_budget_bytes), slots_per_group, 1 if graph else 0, ctypes.byref(h))
        if rc != 0:
            raise DnsError(rc, lib().dns_ensemble_last_error(None).decode())
        self._h = h
        self._shape = {}

    def _check(self, rc):
        if rc != 0:
            raise DnsError(rc, lib().dns_server_last_error(self._h).decode())

    def close(self):
        if getattr(self, "_h", None):
            lib().dns_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, nsteps, keep_fields=False, **kw):
        p = make_params(**kw)
        jid = ctypes.c_int()
        self._check(lib().dns_server_submit(self._h, ctypes.byref(p), nsteps, 1 if keep_fields else 0,
                                            ctypes.byref(jid)))
        self._shape[jid.value] = (p.nz, p.ny, p.nx)
        return jid.value

    def tick(self):
        n = ctypes.c_int()
        self._check(lib().dns_server_tick(self._h, ctypes.byref(n)))
        return n.value

    def drain(self):
        self._check(lib().dns_server_drain(self._h))

    def status(self, jid):
        st, done = ctypes.c_int(), ctypes.c_int()
        self._check(lib().dns_server_job_status(self._h, jid, ctypes.byref(st), ctypes.byref(done)))
        return JOB_STATES[st.value], done.value

    def message(self, jid):
        return lib().dns_server_job_message(self._h, jid).decode()

    def result(self, jid, fields=True):
        d = DnsDiag()
        if fields:
            out = [np.empty(self._shape[jid]) for _ in range(4)]
            self._check(lib().dns_server_job_result(self._h, jid, *[a.ctypes.data for a in out], ctypes.byref(d)))
        else:
            out = None
            self._check(lib().dns_server_job_result(self._h, jid, None, None, None, None, ctypes.byref(d)))
        return out, {f: getattr(d, f) for f, _ in DnsDiag._fields_}

    def times(self, jid):
        out = (ctypes.c_double * 5)()
        self._check(lib().dns_server_job_times(self._h, jid, out))
        return dict(queue_s=out[0], turnaround_s=out[1], step_p50_ms=out[2], step_p99_ms=out[3], steps=int(out[4]))

    def stats(self):
        st = ServerStats()
        self._check(lib().dns_server_get_stats(self._h, ctypes.byref(st)))
        return {f: getattr(st, f) for f, _ in ServerStats._fields_}

    def pending(self):
        return lib().dns_server_pending(self._h)

    def set_health_interval(self, ticks):
        self._check(lib().dns_server_set_health_interval(self._h, ticks))

    def release_idle(self):
        """Free the HBM of shape groups whose jobs have all finished."""
        self._check(lib().dns_server_release_idle(self._h))


class FFTPlan:
    """The FFT module of the reference (fftw3_dns_module.f90:19-152) on the GPU."""

    def __init__(self, nx, ny, precision=DNS_PREC_FP64):
        self.nx, self.ny = nx, ny
        h = _vp()
        rc = lib().dns_fft_plan_create(nx, ny, precision, ctypes.byref(h))
        if rc != 0:
            raise DnsError(rc, lib().dns_last_error(None).decode())
        self._h = h

    def __del__(self):
        if getattr(self, "_h", None):
            lib().dns_fft_plan_destroy(self._h)
            self._h = None

    def forward(self, f):
        """f: (nplanes, ny, nx) real -> (nplanes, ny, nx/2+1) complex, unnormalised."""
        f = np.ascontiguousarray(f, dtype=np.float64)
        npl = f.shape[0]
        out = np.empty((npl, self.ny, self.nx // 2 + 1), dtype=np.complex128)
        rc = lib().dns_fft_forward(self._h, f.ctypes.data, out.ctypes.data, npl)
        if rc:
            raise DnsError(rc, "dns_fft_forward")
        return out

    def backward(self, F):
        """F: (nplanes, ny, nx/2+1) complex -> (nplanes, ny, nx) real, divided by nx*ny."""
        F = np.ascontiguousarray(F, dtype=np.complex128)
        npl = F.shape[0]
        out = np.empty((npl, self.ny, self.nx))
        rc = lib().dns_fft_backward(self._h, F.ctypes.data, out.ctypes.data, npl)
        if rc:
            raise DnsError(rc, "dns_fft_backward")
        return out
