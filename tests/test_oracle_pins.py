"""Pins of the CPU FP64 oracle against the reference's own trace, closed forms,
brute force and independent library routines (numpy FFT, numpy.linalg)."""
import json
import math
import os

import numpy as np
import pytest

from oracle import oracle as orc

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


# --- the reference's shipped 15-step traces (standard_output.txt:88-104 and
# weak_form_output.txt:88-104; the weak-form run departs from the standard one at step 1
# in |w|max and grows unstable by step 15, so it pins the weak-form pressure RHS alone) ---
@pytest.mark.parametrize("weak", [0, 1])
@pytest.mark.parametrize("dealias", [0, 1])
def test_trace_matches_reference_stdout(dealias, weak):
    name = "weak_form_output_trace.json" if weak else "standard_output_trace.json"
    fx = json.load(open(os.path.join(GOLDEN, name)))
    o = orc.Oracle(nx=32, ny=32, nz=17, xlen=2 * np.pi, ylen=np.pi, re=500.0, dt=0.002, u_centerline=1.5,
                   flow_control_method=1, target_pressure_gradient=0.006, use_dealias_23_rule=dealias,
                   enable_perturbations=1, perturbation_kind=0, perturbation_amplitude=0.01,
                   use_weak_form_divergence=weak)
    o.init_fields()
    for row in fx["rows"]:
        o.step(1)
        d = o.diagnostics()
        assert d["step"] == row["step"]
        got = {k: "%.4E" % d[f] for k, f in (("umax", "umax"), ("vmax", "vmax"), ("wmax", "wmax"),
                                              ("E", "energy"), ("dPdx", "pressure_gradient"))}
        want = {k: "%.4E" % float(row[k]) for k in got}
        assert got == want, f"step {row['step']}"


def test_initial_perturbation_scale_matches_reference():
    # standard_output.txt:45: "Scaled from max= 2.0000E+00"; :53: max |u'|, |v'| = 1e-2, 5e-3
    u, v, w = orc.initial_field(nx=32, ny=32, nz=17, xlen=2 * np.pi, ylen=np.pi, u_centerline=0.0,
                                enable_perturbations=1, perturbation_amplitude=0.01)
    assert "%.4E" % np.abs(u).max() == "1.0000E-02"
    assert "%.4E" % np.abs(v).max() == "5.0000E-03"
    assert np.all(w == 0.0)


# --- LGL nodes, weights, differentiation matrix (lgl_module.f90) -------------------------
def test_lgl_closed_forms():
    x, w = orc.lgl(3)
    np.testing.assert_allclose(x, [-1, 0, 1], atol=1e-16)
    np.testing.assert_allclose(w, [1 / 3, 4 / 3, 1 / 3], rtol=1e-15)
    x, w = orc.lgl(5)
    r = math.sqrt(3 / 7)
    np.testing.assert_allclose(x, [-1, -r, 0, r, 1], atol=1e-15)
    np.testing.assert_allclose(w, [1 / 10, 49 / 90, 32 / 45, 49 / 90, 1 / 10], rtol=1e-14)


@pytest.mark.parametrize("n", [17, 33, 65, 129])
def test_lgl_quadrature_exact(n):
    x, w = orc.lgl(n)
    assert abs(w.sum() - 2.0) < 1e-13
    for p in range(0, 2 * n - 2, 2):                 # exact for degree <= 2n-3
        assert abs(np.dot(w, x ** p) - 2.0 / (p + 1)) < 1e-12


@pytest.mark.parametrize("n", [17, 33])
def test_d1_exact_on_polynomials(n):
    x, _ = orc.lgl(n)
    D = orc.diff_matrix(n)
    for p in range(1, n):
        got = D @ (x ** p)
        want = p * x ** (p - 1)
        assert np.max(np.abs(got - want)) < 1e-10 * n * n * p, p
    # d2/dz2 of 1.5(1-z^2) = -3 (debug_flow_analysis.f90:59-67)
    np.testing.assert_allclose(D @ (D @ (1.5 * (1 - x ** 2))), -3.0, atol=1e-9)


@pytest.mark.parametrize("n", [17, 33, 49, 57, 65, 129])
def test_d1_denominator_threshold(n):
    # lgl_module.f90:254: D(i,j) = 0 when |prod_{k!=j}(x_j - x_k)| <= 1e-15.  Brute force the
    # products independently and compare which entries survive.
    x, _ = orc.lgl(n)
    D = orc.diff_matrix(n)
    prods = np.array([np.prod([x[j] - x[k] for k in range(n) if k != j]) for j in range(n)])
    for j in range(n):
        col = np.delete(D[:, j], j)
        if abs(prods[j]) > 1e-15:
            assert np.all(col != 0.0)
        else:
            assert np.all(col == 0.0)
    if n >= 65:                                      # the reference's D1 is diagonal here
        assert np.count_nonzero(D - np.diag(np.diag(D))) == 0


# --- wavenumbers and mask ----------------------------------------------------------------
def test_wavenumbers_closed_form():
    nx, ny, xlen, ylen = 32, 24, 4 * np.pi, 2 * np.pi
    xw, yw = orc.wavenumbers(nx, ny, xlen, ylen)
    np.testing.assert_array_equal(xw, np.arange(nx // 2 + 1) * (2 * np.pi / xlen))
    k = np.array([j if j <= ny // 2 else j - ny for j in range(ny)], dtype=float)
    np.testing.assert_allclose(yw, k * 2.0 * np.pi / ylen, rtol=1e-15)
    assert yw[ny // 2] > 0                            # Nyquist carries +ny/2


@pytest.mark.parametrize("nx,kxm", [(32, 10), (64, 21), (96, 32), (128, 42), (192, 64)])
@pytest.mark.parametrize("ny,kym", [(24, 8), (36, 12), (96, 32), (130, 43)])
def test_dealias_mask_bounds(nx, kxm, ny, kym):
    # the closed-form bounds (kx_max = (nx/2)*2/3 = 10, 21, 32, 42, 64) and the reference's loop
    # (1-based i, j; :1176-1190) written out element by element, with ny != nx
    m = orc.dealias_mask(nx, ny)
    want = np.ones((ny, nx // 2 + 1), dtype=bool)
    for j1 in range(1, ny + 1):
        for i1 in range(1, nx // 2 + 2):
            if i1 - 1 > kxm:
                want[j1 - 1, i1 - 1] = False
            if j1 <= ny // 2 + 1:
                if j1 - 1 > kym:
                    want[j1 - 1, i1 - 1] = False
            elif abs(j1 - 1 - ny) > kym:
                want[j1 - 1, i1 - 1] = False
    np.testing.assert_array_equal(m.astype(bool), want)


# --- 2D transforms (FFTW conventions) ----------------------------------------------------
def _dft_r2c_brute(f):
    ny, nx = f.shape
    X = np.zeros((ny, nx // 2 + 1), dtype=complex)
    y = np.arange(ny)[:, None]; x = np.arange(nx)[None, :]
    for ky in range(ny):
        for kx in range(nx // 2 + 1):
            X[ky, kx] = np.sum(f * np.exp(-2j * np.pi * (kx * x / nx + ky * y / ny)))
    return X


def _c2r_formula(X, nx):
    # full complex inverse along y, then half-complex inverse along x using Re of kx=0, nx/2
    ny = X.shape[0]
    g = np.fft.ifft(X, axis=0) * ny
    c = np.full(nx // 2 + 1, 2.0); c[0] = 1.0; c[nx // 2] = 1.0
    xs = np.arange(nx)
    out = np.zeros((ny, nx))
    for k in range(nx // 2 + 1):
        out += c[k] * np.real(g[:, k][:, None] * np.exp(2j * np.pi * k * xs[None, :] / nx))
    return out / (nx * ny)


@pytest.mark.parametrize("nx,ny", [(8, 6), (6, 10), (12, 4)])
def test_fft_r2c_brute_force(nx, ny):
    rng = np.random.default_rng(1)
    f = rng.standard_normal((2, ny, nx))
    X = orc.fft_r2c(f)
    for p in range(2):
        np.testing.assert_allclose(X[p], _dft_r2c_brute(f[p]), atol=1e-12)


@pytest.mark.parametrize("nx,ny", [(32, 32), (96, 64), (128, 128), (192, 96)])
def test_fft_vs_numpy(nx, ny):
    rng = np.random.default_rng(2)
    f = rng.standard_normal((3, ny, nx))
    X = orc.fft_r2c(f)
    np.testing.assert_allclose(X, np.fft.rfft2(f), atol=1e-11 * math.sqrt(nx * ny))
    back = orc.fft_c2r(X, nx)
    assert np.max(np.abs(back - f)) < 1e-13


@pytest.mark.parametrize("nx,ny", [(8, 6), (32, 32)])
def test_c2r_non_hermitian_semantics(nx, ny):
    rng = np.random.default_rng(3)
    X = rng.standard_normal((1, ny, nx // 2 + 1)) + 1j * rng.standard_normal((1, ny, nx // 2 + 1))
    np.testing.assert_allclose(orc.fft_c2r(X, nx)[0], _c2r_formula(X[0], nx), atol=1e-14)


def test_fft_roundtrip_known_answer():
    # test_fft_roundtrip (DNS_pressure_BC_3D.f90:2035-2058): sin(2 alpha x) cos(3 beta y), err < 1e-12
    nx, ny, xlen, ylen = 32, 32, 2 * np.pi, np.pi
    x = np.arange(nx) * xlen / nx; y = np.arange(ny) * ylen / ny
    f = np.sin(2 * (2 * np.pi / xlen) * x)[None, :] * np.cos(3 * (2 * np.pi / ylen) * y)[:, None]
    back = orc.fft_c2r(orc.fft_r2c(f[None]), nx)[0]
    assert np.max(np.abs(back - f)) < 1e-12


# --- whole-step invariants ---------------------------------------------------------------
def test_poiseuille_fixed_point():
    # U = uc(1-z^2), G = 2 uc/Re is a steady solution; E = 4 uc^2/15 = 0.6 for uc = 1.5
    uc, re = 1.5, 500.0
    o = orc.Oracle(nx=16, ny=16, nz=17, re=re, u_centerline=uc, target_pressure_gradient=2 * uc / re,
                   enable_perturbations=0, use_dealias_23_rule=1)
    o.init_fields()
    u0 = o.get_fields()[0]
    o.step(20)
    u, v, w, p = o.get_fields()
    assert np.max(np.abs(u - u0)) < 1e-11
    assert np.max(np.abs(v)) < 1e-12 and np.max(np.abs(w)) < 1e-12
    assert abs(o.diagnostics()["energy"] - 0.6) < 1e-12


def test_cached_factors_bitwise_equal_refactoring():
    kw = dict(nx=16, ny=16, nz=17, re=180.0, flow_control_method=2, use_dealias_23_rule=1,
              perturbation_kind=1, seed=7, noise_kmax=3, perturbation_amplitude=0.05)
    a = orc.Oracle(refactor_every_step=0, **kw); a.init_fields(); a.step(4)
    b = orc.Oracle(refactor_every_step=1, **kw); b.init_fields(); b.step(4)
    for fa, fb in zip(a.get_fields(), b.get_fields()):
        np.testing.assert_array_equal(fa, fb)


def test_helmholtz_double_lifting_against_numpy_solve():
    # Independent restatement of one viscous stage with numpy.fft + numpy.linalg.solve, with
    # nonzero wall values so the double lifting of :1494-1502 is exercised.
    nx, ny, nz, re, dt = 8, 8, 9, 100.0, 0.01
    o = orc.Oracle(nx=nx, ny=ny, nz=nz, re=re, dt=dt, flow_control_method=1, target_pressure_gradient=0.02,
                   use_dealias_23_rule=0, perturbation_kind=1, seed=3, noise_kmax=2, perturbation_amplitude=0.1)
    o.init_fields()
    rng = np.random.default_rng(5)
    gpx = rng.standard_normal((2, ny, nx)) * 0.1; gpy = rng.standard_normal((2, ny, nx)) * 0.1
    o.set_wall_gradients(gpx, gpy)
    u, v, w, _ = o.get_fields()
    o.compute_source_terms()
    su = o.source_terms()[0]
    o.viscous_step()
    ustar = o.star_fields()[0]
    z, wz = orc.lgl(nz)
    D = orc.diff_matrix(nz)
    f = 0.5 * dt / re
    kx = np.fft.rfftfreq(nx, 1.0 / nx) * (2 * np.pi / (2 * np.pi))
    ky = np.array([j if j <= ny // 2 else j - ny for j in range(ny)], float) * 2.0 * np.pi / np.pi
    uh = np.fft.rfft2(u); sh = np.fft.rfft2(su)
    lap = np.einsum("km,myx->kyx", D @ D, uh) - (kx[None, None, :] ** 2 + ky[None, :, None] ** 2) * uh
    rh = uh + dt * sh + f * lap
    b1 = np.fft.rfft2(-dt * gpx[0]); b2 = np.fft.rfft2(-dt * gpx[1])
    A = D.T @ np.diag(wz) @ D
    out = np.zeros_like(uh)
    for j in range(ny):
        for i in range(nx // 2 + 1):
            H = f * A + np.diag((1 + f * (kx[i] ** 2 + ky[j] ** 2)) * wz)
            I = slice(1, nz - 1)
            rhs = wz[I] * rh[I, j, i] - 2 * H[I, 0] * b1[j, i] - 2 * H[I, nz - 1] * b2[j, i]
            out[I, j, i] = np.linalg.solve(H[I, I], rhs)
            out[0, j, i] = b1[j, i]; out[nz - 1, j, i] = b2[j, i]
    want = np.fft.irfft2(out, s=(ny, nx))
    np.testing.assert_allclose(ustar, want, atol=1e-12 * np.abs(want).max())


def test_seeded_noise_is_solenoidal_and_scaled():
    kw = dict(nx=32, ny=32, nz=33, xlen=4 * np.pi, ylen=2 * np.pi, u_centerline=0.0, perturbation_kind=1,
              seed=12345, noise_kmax=8, perturbation_amplitude=0.05)
    u, v, w = orc.initial_field(**kw)
    assert abs(np.sqrt(u ** 2 + v ** 2).max() - 0.05) < 1e-15
    o = orc.Oracle(**kw); o.init_fields()
    assert o.max_divergence() < 1e-13
    u2, _, _ = orc.initial_field(**kw)
    np.testing.assert_array_equal(u, u2)


def test_reference_d1_nz65_singular_poisson_like_zgesv():
    # centre diagonal entry of the reference D1 is exactly 0 at nz = 65, so the mode-(0,0)
    # Poisson matrix is singular: the reference's zgesv returns info > 0 and the program stops
    D = orc.diff_matrix(65)
    assert np.count_nonzero(D - np.diag(np.diag(D))) == 0
    assert D[43, 43] == 0.0               # interior diagonal entries are rounding residue
    o = orc.Oracle(nx=8, ny=8, nz=65, enable_perturbations=0)
    o.init_fields()
    o.step(1)
    assert o.status() > 0


def test_reference_d1_nz129_blows_up_as_written():
    # nz = 129: interior diagonal entries of the reference D1 are rounding residue (<= 1.5e-10),
    # the mode-(0,0) Poisson solve amplifies rounding by ~1e10 and the plane-mean w doubles every
    # step; with the exact LGL D1 the same run stays bounded
    kw = dict(nx=8, ny=8, nz=129, xlen=4 * np.pi, ylen=2 * np.pi, flow_control_method=2, perturbation_kind=1,
              seed=2024, noise_kmax=3, perturbation_amplitude=0.05)
    a = orc.Oracle(lgl_derivative=0, **kw); a.init_fields(); a.step(40)
    b = orc.Oracle(lgl_derivative=1, **kw); b.init_fields(); b.step(40)
    assert a.diagnostics()["wmax"] > 1e3
    assert b.diagnostics()["wmax"] < 1e-3


@pytest.mark.parametrize("n", [33, 65, 129])
def test_exact_d1_polynomials(n):
    x, _ = orc.lgl(n)
    D = orc.diff_matrix_exact(n)
    for p in (1, 2, 5, n // 2, n - 1):
        assert np.max(np.abs(D @ x ** p - p * x ** (p - 1))) < 1e-11 * n * n * p
    if n == 33:
        np.testing.assert_allclose(D, orc.diff_matrix(n), atol=1e-9 * np.abs(D).max())


# --- PI controller and bulk velocity (DNS_pressure_BC_3D.f90:2434-2464, 2488-2549) ------
# Hand-worked recurrences on prescribed bulk velocities.  Re = 100 (base 3/Re = 0.03),
# controller_gain Kp = 0.2 (Ki = 0.05 Kp = 0.01), target U_b* = 0.5, dt = 0.01,
# target_pressure_gradient = 0.02 (the state read_input_3d leaves: G = average = 0.02).
PI = dict(re=100.0, controller_gain=0.2, target_bulk_velocity=0.5, dt=0.01, target_pressure_gradient=0.02)


def _close(a, b):
    assert abs(a - b) <= 1e-15 * max(1.0, abs(b)), (a, b)


def test_pi_euler_branch_and_clamps_by_hand():
    # istep = 0, freq 7: mod(0, 7) = 0 runs every call and 0 <= 7 takes the Euler branch
    c = orc.PIController(controller_update_freq=7, **PI)
    rows = [  # U_b, G, integral, prev, average of G
        (0.40, 0.05, 0.001, 0.1, 0.05),                       # e = .1: G = .03 + .2*.1 + .01*0
        (0.45, 0.04001, 0.0015, 0.05, 0.045005),              # e = .05: G = .03 + .01 + .01*.001
        (1.00, 0.0, 0.0015, -0.5, 0.09001 / 3),              # e = -.5: G_u = -.069985 -> 0, I frozen
        (-60.0, 10.0, 0.0015, 60.5, 10.09001 / 4),           # e = 60.5: G_u = 12.130015 -> 10, I frozen
        (0.50, 0.030015, 0.0015, 0.0, 10.120025 / 5),        # e = 0: G = .03 + .01*.0015
    ]
    for n, (ub, G, I, prev, avg) in enumerate(rows, 1):
        assert c.update(0, ub)
        _close(c.Ub, ub); _close(c.G, G); _close(c.err_int, I); _close(c.err_prev, prev); _close(c.gavg, avg)
        assert c.gcount == n


def test_pi_trapezoid_branch_by_hand():
    # istep = 4 > freq 2 with mod(4, 2) = 0: I += dt/2 (e_prev + e)
    c = orc.PIController(controller_update_freq=2, **PI)
    for ub, G, I in [(0.40, 0.05, 0.0005),                  # I = .005 * (0 + .1)
                     (0.45, 0.040005, 0.00125),             # G = .04 + .01*.0005; I += .005 * .15
                     (0.30, 0.0700125, 0.0025)]:            # G = .07 + .01*.00125; I += .005 * .25
        assert c.update(4, ub)
        _close(c.G, G); _close(c.err_int, I)
    # istep = freq is still the Euler branch (istep <= freq), 2*freq the trapezoid
    e = orc.PIController(controller_update_freq=7, **PI)
    e.update(7, 0.40); e.update(7, 0.45)
    _close(e.err_int, 0.0015)
    t = orc.PIController(controller_update_freq=7, **PI)
    t.update(14, 0.40); t.update(14, 0.45)
    _close(t.err_int, 0.00125)


def test_pi_skipped_when_istep_not_a_multiple_of_freq():
    c = orc.PIController(controller_update_freq=7, **PI)
    for _ in range(3):
        assert not c.update(3, 0.1)
    assert c.G == 0.02 and c.gavg == 0.02 and c.gcount == 0 and c.err_int == 0.0 and c.Ub == 0.0


def test_pi_skip_branch_keeps_forcing_in_a_run():
    # istart = 3, freq 7: the controller never runs, so G stays target_pressure_gradient
    o = orc.Oracle(nx=8, ny=8, nz=9, flow_control_method=2, istart=3, controller_update_freq=7,
                   target_pressure_gradient=0.0125, perturbation_kind=1, seed=3, noise_kmax=2)
    o.init_fields()
    o.step(3)
    d = o.diagnostics()
    assert d["pressure_gradient"] == 0.0125 and d["avg_pressure_gradient"] == 0.0125
    assert d["bulk_error_integral"] == 0.0


@pytest.mark.parametrize("nz", [17, 33, 129])
def test_laminar_bulk_velocity_is_two_thirds_of_centreline(nz):
    # U_b = (1/2) int_{-1}^{1} uc (1 - z^2) dz = 2 uc / 3; LGL quadrature is exact for degree 2.
    # A zero-mean x-periodic part leaves it unchanged.
    nx, ny, uc = 12, 10, 1.5
    z, _ = orc.lgl(nz)
    x = np.arange(nx) * 2 * np.pi / nx
    u = uc * (1 - z ** 2)[:, None, None] * np.ones((nz, ny, nx))
    assert abs(orc.bulk_velocity(u) - 2 * uc / 3) < 1e-12
    u2 = u + 0.3 * np.sin(x)[None, None, :] * (1 - z ** 2)[:, None, None]
    assert abs(orc.bulk_velocity(u2) - 2 * uc / 3) < 1e-12
    o = orc.Oracle(nx=nx, ny=ny, nz=nz, u_centerline=uc, enable_perturbations=0)
    o.init_fields()
    assert abs(o.diagnostics()["bulk_velocity"] - 2 * uc / 3) < 1e-12


# --- the rounding floor of w -------------------------------------------------------------
def test_w_rounding_floor_of_the_oracle_itself():
    """w is ~1e-6 of u in the 32x32x33 channel, so any reordering of the arithmetic moves it at the 1e-10
    level of its own norm.  The oracle against itself with every plane transform factored in
    radix 2 instead of 4, odd factors first (an equivalent transform, other rounding): after 100
    steps u, v, p agree to ~1e-13 while w's own relative difference is ~1e-10."""
    from tests.gpu_util import CHANNEL32, field_errors
    a = orc.Oracle(**CHANNEL32); a.init_fields()
    b = orc.Oracle(fft_variant=1, **CHANNEL32); b.init_fields()
    a.step(100); b.step(100)
    e = field_errors(b.get_fields(), a.get_fields())
    assert e["u"] < 1e-13 and e["v"] < 1e-12 and e["p"] < 1e-13, e
    assert 5e-11 < e["w_self"] < 5e-10, e


# --- optional numerics variants of the 2D solver on the 3D step (parity unpinned: no 3D
# trace of the reference exists for them) ------------------------------------------------
def test_jacgs_two_by_two_by_hand():
    # A = [[4, 1], [1, 3]], b = [1, 2]: Jacobi-preconditioned CG reaches the exact solution
    # x = A^-1 b = [1/11, 7/11] in n = 2 iterations (a wrong alfa, beta or preconditioner does not)
    x, info, it = orc.jacgs(np.array([[4.0, 1.0], [1.0, 3.0]]), np.array([1.0, 2.0]), tol=1e-14)
    assert info == 0 and it == 2
    np.testing.assert_allclose(x, [1 / 11, 7 / 11], rtol=1e-14)
    # first iterate by hand: r0 = b, q0 = p0 = D^-1 b = [1/4, 2/3], A p0 = [5/3, 9/4],
    # alfa = (q0.r0)/(p0.Ap0) = (19/12)/(23/12) = 19/23, x1 = alfa p0
    x1, _, it1 = orc.jacgs(np.array([[4.0, 1.0], [1.0, 3.0]]), np.array([1.0, 2.0]), niter=1, tol=1e-14)
    assert it1 == 1
    np.testing.assert_allclose(x1, [19 / 23 / 4, 19 / 23 * 2 / 3], rtol=1e-14)


def test_jacgs_limits_and_definite_systems():
    rng = np.random.default_rng(4)
    M = rng.standard_normal((30, 30))
    A = M @ M.T + 30 * np.eye(30)
    b = rng.standard_normal(30)
    x, info, it = orc.jacgs(A, b, tol=1e-10)
    assert info == 0 and it < 100
    assert np.linalg.norm(b - A @ x) <= 1e-10
    np.testing.assert_allclose(x, np.linalg.solve(A, b), rtol=1e-8)
    # negative definite (the Poisson operator's sign) converges the same way
    xn, info, _ = orc.jacgs(-A, b, tol=1e-10)
    assert info == 0
    np.testing.assert_allclose(xn, -x, rtol=1e-8)
    # ||b|| <= tol: no iteration, x stays 0; tol = 0: the iteration limit, info = -1
    x0, info, it = orc.jacgs(A, 1e-9 * b / np.linalg.norm(b), tol=1e-8)
    assert info == 0 and it == 0 and np.all(x0 == 0)
    _, info, it = orc.jacgs(A, b, niter=7, tol=0.0)
    assert info == -1 and it == 7


def test_jacgs_pressure_stage_matches_direct_solve_when_tight():
    kw = dict(nx=16, ny=12, nz=17, flow_control_method=2, perturbation_kind=1, seed=5, noise_kmax=3,
              perturbation_amplitude=0.05)
    a = orc.Oracle(**kw); a.init_fields(); a.step(3)
    b = orc.Oracle(pressure_solver=1, cgstol=1e-12, **kw); b.init_fields(); b.step(3)
    it, nf = b.jacgs_stats()
    assert nf == 0 and it > 0
    for name, x, y in zip("uvwp", b.get_fields(), a.get_fields()):
        assert np.linalg.norm(x - y) <= (1e-9 if name == "w" else 1e-12) * np.linalg.norm(y), name


def test_rk_convection_stages_compose_the_single_product():
    """convection_scheme 1 against the stage recurrence run by hand on the oracle's own single
    product (convection_scheme 0 on the stage fields): S4 = S(un), u3 = un + dt/4 S4, ...,
    su = S(u1) + G with u1 = un + dt/2 S2, walls zero after each stage, fields back to un."""
    kw = dict(nx=16, ny=12, nz=17, re=150.0, dt=0.004, flow_control_method=1, target_pressure_gradient=0.01,
              perturbation_kind=1, seed=9, noise_kmax=3, perturbation_amplitude=0.08)
    r = orc.Oracle(convection_scheme=1, **kw); r.init_fields()
    un = r.get_fields()
    r.compute_source_terms()
    got = r.source_terms()
    assert all(np.array_equal(x, y) for x, y in zip(r.get_fields(), un))
    s = orc.Oracle(**kw)
    f = [x.copy() for x in un[:3]]
    for irk in (4, 3, 2, 1):
        s.set_fields(*f, un[3])
        s.compute_source_terms()
        S = s.source_terms()
        S[0] = S[0] - 0.01                       # the single product's forcing, added once at the end
        f = [u0 + (1.0 / irk) * kw["dt"] * sc for u0, sc in zip(un[:3], S)]
        for x in f:
            x[0] = 0.0; x[-1] = 0.0
    S[0] = S[0] + 0.01
    for x, y in zip(got, S):
        np.testing.assert_allclose(x, y, rtol=0, atol=1e-15 * max(1.0, np.abs(y).max()))
    # and it is a different scheme from the single evaluation (O(dt) apart)
    e = orc.Oracle(**kw); e.init_fields(); e.compute_source_terms()
    d = np.abs(e.source_terms()[1] - got[1]).max()
    assert 1e-8 < d < 1e-2
