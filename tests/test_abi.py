"""CPU-side checks of the C ABI library: it builds, loads and exports every declared symbol."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for h in os.listdir(os.path.join(ROOT, "include")):
        if h.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", h)).read()
            txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
            names.update(re.findall(r"\b(dns_[a-z0-9_]+)\s*\(", txt))
    return names


def test_library_exports_every_declared_symbol():
    import f90_dns_amd
    path = f90_dns_amd.library_path()
    if not os.path.exists(path):
        pytest.skip("libf90dns.so not built")
    L = ctypes.CDLL(path)
    missing = [n for n in sorted(declared_symbols()) if not hasattr(L, n)]
    assert not missing, missing
    assert len(declared_symbols()) > 25


def test_version_string_without_gpu():
    import f90_dns_amd
    if not os.path.exists(f90_dns_amd.library_path()):
        pytest.skip("libf90dns.so not built")
    assert b"gfx950" in f90_dns_amd.lib().dns_version()


def test_params_struct_matches_header():
    """The ctypes mirror of dns_params lists the header's fields in order."""
    import f90_dns_amd
    txt = open(os.path.join(ROOT, "include", "f90dns.h")).read()
    body = txt[txt.index("typedef struct dns_params"):txt.index("} dns_params;")]
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    fields = []
    for decl in body.split(";"):
        decl = decl.strip()
        if not decl or decl.startswith("typedef"):
            decl = decl.split("{", 1)[-1].strip()
        if not decl:
            continue
        parts = decl.replace(",", " , ").split()
        names = [t for t in parts[1:] if t not in (",", "long", "unsigned", "int", "double")]
        fields += names
    mine = [f for f, _ in f90_dns_amd.DnsParams._fields_]
    assert fields == mine


def test_fortran_driver_builds_and_fails_cleanly_without_gpu(tmp_path):
    """The Fortran driver (ISO_C_BINDING over the C ABI) links and reports a missing device
    through dns_last_error instead of crashing."""
    import subprocess
    from f90_dns_amd import build as b
    if b.amdflang() is None:
        pytest.skip("no Fortran compiler")
    exes = b.build_fortran()
    assert any(e.endswith("dns_3d_gpu") for e in exes)
    r = subprocess.run([os.path.join(b.BIN, "dns_3d_gpu"), os.path.join(ROOT, "tests", "golden", "trace_input.dat")],
                       capture_output=True, text=True, cwd=tmp_path, timeout=60)
    if r.returncode == 0:
        pytest.skip("a GPU is present: covered by the GPU driver test")
    assert "dns_create" in r.stdout and "ERROR" in r.stdout


def test_batch_rejects_members_of_different_shape_before_touching_a_device():
    """dns_create_batch validates every member (and that they share grid, dt, domain, flags)
    before any HIP call, so these errors come back on a machine without a GPU too."""
    import ctypes
    import f90_dns_amd as dns
    if not os.path.exists(dns.library_path()):
        pytest.skip("libf90dns.so not built")
    L = dns.lib()
    a = dns.make_params(nx=32, ny=32, nz=33)
    b = dns.make_params(nx=32, ny=32, nz=65)
    arr = (dns.DnsParams * 2)(a, b)
    h = ctypes.c_void_p()
    assert L.dns_create_batch(arr, 2, ctypes.byref(h)) == 1
    assert b"member 1 differs" in L.dns_last_error(None)
    c = dns.make_params(nx=32, ny=32, nz=33, re=-1.0)
    arr = (dns.DnsParams * 2)(a, c)
    assert L.dns_create_batch(arr, 2, ctypes.byref(h)) == 1
    assert b"member 1" in L.dns_last_error(None)
    assert L.dns_create_batch(arr, 0, ctypes.byref(h)) == 1
    d = dns.make_params(nx=32, ny=32, nz=33, re=300.0, seed=5)
    arr = (dns.DnsParams * 2)(a, d)
    rc = L.dns_create_batch(arr, 2, ctypes.byref(h))
    assert rc in (0, 3)
    if rc == 0:
        L.dns_destroy(h)


def test_host_initial_field_matches_oracle_generator():
    """The library's host generator (the reference for the device assembly) against the
    oracle's own restatement of initialize_fields_3d, analytical and seeded-noise kinds."""
    import numpy as np
    import f90_dns_amd as dns
    from oracle import oracle as orc
    if not os.path.exists(dns.library_path()):
        pytest.skip("libf90dns.so not built")
    for kw in [dict(nx=32, ny=32, nz=33, xlen=2 * np.pi, ylen=np.pi, enable_perturbations=1, perturbation_kind=0,
                    u_centerline=1.0, perturbation_amplitude=0.01),
               dict(nx=48, ny=36, nz=17, xlen=4 * np.pi, ylen=np.pi, enable_perturbations=1, perturbation_kind=1,
                    seed=123, noise_kmax=5, perturbation_amplitude=0.05, u_centerline=1.5),
               dict(nx=128, ny=128, nz=129, xlen=4 * np.pi, ylen=2 * np.pi, enable_perturbations=1,
                    perturbation_kind=1, seed=2024, noise_kmax=8, perturbation_amplitude=0.05, u_centerline=1.0),
               dict(nx=96, ny=64, nz=65, xlen=2 * np.pi, ylen=np.pi, enable_perturbations=1, perturbation_kind=0,
                    analytical_mode_x=2, analytical_mode_y=3, perturbation_amplitude=0.03, u_centerline=1.0),
               dict(nx=16, ny=24, nz=9, enable_perturbations=0, u_centerline=1.0)]:
        u, v, w = dns.host_initial_field(**kw)
        o = orc.Oracle(**kw)
        o.init_fields()
        ou, ov, ow, _ = o.get_fields()
        # bit-identical: the same operations in the same order (initialize_fields_3d)
        np.testing.assert_array_equal(u, ou)
        np.testing.assert_array_equal(v, ov)
        np.testing.assert_array_equal(w, ow)


@pytest.mark.parametrize("nx,ny", [(4, 2), (8, 6), (32, 32), (48, 36), (64, 96), (128, 128), (192, 48), (32, 27),
                                   (16, 9)])
def test_solve_mode_space_covers_a_real_planes_transform(nx, ny):
    """The per-mode solves walk the independent coefficients of a real plane's transform: every
    storage column is written exactly once (by its slot, or as a slot's mirror or a zeroed
    imaginary part), and the slots alone rebuild numpy's rfft2 of a random real plane."""
    import numpy as np
    import f90_dns_amd as dns
    if not os.path.exists(dns.library_path()):
        pytest.skip("libf90dns.so not built")
    so, s2o, s2s = dns.zmap(nx, ny)
    nxh = nx // 2 + 1
    H2 = 2 * nxh * ny
    assert len(so) == (nx * ny if ny % 2 == 0 else H2)
    targets = np.concatenate([so, s2o[s2o >= 0]])
    assert np.array_equal(np.sort(targets), np.arange(H2))
    if ny % 2:
        assert np.array_equal(so, np.arange(H2)) and np.all(s2o < 0)
    X = np.fft.rfft2(np.random.default_rng(nx * ny).standard_normal((ny, nx)))   # [j, i]: mode i + nxh j
    R = np.empty(H2)
    R[0::2], R[1::2] = X.real.ravel(), X.imag.ravel()
    Y = np.full(H2, np.nan)
    Y[so] = R[so]
    m = s2o >= 0
    Y[s2o[m]] = s2s[m] * R[so[m]]
    np.testing.assert_allclose(Y, R, rtol=0, atol=1e-12 * np.abs(R).max())


@pytest.mark.parametrize("nx,ny,xlen,ylen", [(32, 24, 6.283185307179586, 3.141592653589793),
                                             (128, 96, 4 * 3.141592653589793, 2 * 3.141592653589793),
                                             (48, 36, 1.7, 0.9), (192, 64, 12.566370614359172, 3.0)])
def test_library_wavenumbers_bitwise_equal_oracle(nx, ny, xlen, ylen):
    """xw, yw of setup_spectral_3d (DNS_pressure_BC_3D.f90:780-796) as the library uploads them,
    bit for bit against the oracle's (ny != nx, so an x/y swap fails)."""
    import numpy as np
    import f90_dns_amd as dns
    from oracle import oracle as orc
    if not os.path.exists(dns.library_path()):
        pytest.skip("libf90dns.so not built")
    xw, yw = dns.wavenumbers(nx, ny, xlen, ylen)
    oxw, oyw = orc.wavenumbers(nx, ny, xlen, ylen)
    np.testing.assert_array_equal(xw, oxw)
    np.testing.assert_array_equal(yw, oyw)


@pytest.mark.parametrize("nx,ny", [(32, 24), (24, 32), (128, 96), (96, 64), (192, 128), (64, 192), (48, 36), (34, 22),
                                   (8, 6)])
def test_library_dealias_mask_bitwise_equal_oracle(nx, ny):
    """The 2/3 mask (apply_dealias_truncation, DNS_pressure_BC_3D.f90:1163-1193): the two-pass
    transforms' table and the plane kernels' bounds rule, both bit for bit against the oracle,
    with ny != nx (a kx/ky bound swap fails)."""
    import numpy as np
    import f90_dns_amd as dns
    from oracle import oracle as orc
    if not os.path.exists(dns.library_path()):
        pytest.skip("libf90dns.so not built")
    want = orc.dealias_mask(nx, ny)
    for which in (0, 1):
        np.testing.assert_array_equal(dns.dealias_mask(nx, ny, which), want)


def test_fortran_driver_command_line_without_gpu(tmp_path):
    """process_command_line_args (DNS_pressure_BC_3D.f90:445-528) before any device work: --help
    prints the usage and stops, an unknown argument stops with an error, no argument selects
    input_3d.dat, --restart [file.dat] is accepted (and then reset by read_input_3d, :626-627)."""
    import subprocess
    from f90_dns_amd import build as b
    if b.amdflang() is None:
        pytest.skip("no Fortran compiler")
    exe = os.path.join(b.BIN, "dns_3d_gpu")
    if not os.path.exists(exe):
        b.build_fortran()
    run = lambda *a: subprocess.run([exe, *a], capture_output=True, text=True, cwd=tmp_path, timeout=60)
    r = run("--help")
    assert r.returncode == 0 and "USAGE:" in r.stdout and "--restart [file]" in r.stdout
    assert "dns_create" not in r.stdout
    r = run("-h")
    assert "USAGE:" in r.stdout
    r = run("bogus")
    assert "ERROR: Unknown argument: bogus" in r.stdout and "dns_create" not in r.stdout
    r = run()
    assert "Using default input file: input_3d.dat" in r.stdout
    assert "WARNING: Could not open input file: input_3d.dat" in r.stdout
    r = run("--restart", "saved.dat", "case.dat")
    assert "Restart requested from file: saved.dat" in r.stdout
    assert "Using specified input file: case.dat" in r.stdout
    assert "Command line restart mode enabled" in r.stdout
    r = run("--restart")
    assert "Restart requested from default file: restart.dat" in r.stdout


def test_ensemble_rejects_the_numerics_variants_before_touching_a_device():
    import ctypes
    import f90_dns_amd as dns
    if not os.path.exists(dns.library_path()):
        pytest.skip("libf90dns.so not built")
    L = dns.lib()
    a = dns.make_params(nx=32, ny=32, nz=33, convection_scheme=1)
    arr = (dns.DnsParams * 1)(a)
    h = ctypes.c_void_p()
    assert L.dns_ensemble_create(arr, 1, ctypes.byref(h)) == 6
    assert b"single contexts and batches only" in L.dns_ensemble_last_error(None)
