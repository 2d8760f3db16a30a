"""Parity of the HIP path (through the C ABI) with the CPU FP64 oracle."""
import sys

import numpy as np
import pytest

from oracle import oracle as orc
from tests.gpu_util import CHANNEL32, W_SELF_FP64, W_SELF_MIXED, check_parity, field_errors, max_err, pair, rel_l2

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def dns():
    import f90_dns_amd
    f90_dns_amd.lib()
    return f90_dns_amd


# 192x192, 192x96 and 256x128 exceed the LDS-resident FP64 plane and take the two-pass transforms
SMALL_AND_RADIX3 = [(8, 4), (12, 6), (18, 18), (24, 54), (54, 24), (162, 48)]


@pytest.mark.parametrize("nx,ny", [(32, 32), (64, 64), (96, 96), (128, 128), (128, 32), (96, 64), (48, 36),
                                   (192, 192), (192, 96), (256, 128)] + SMALL_AND_RADIX3)
def test_fft_forward_matches_numpy(dns, nx, ny):
    _say(f"fft forward {nx}x{ny}")
    rng = np.random.default_rng(nx + ny)
    f = rng.standard_normal((5, ny, nx))
    X = dns.FFTPlan(nx, ny).forward(f)
    ref = np.fft.rfft2(f)
    assert np.max(np.abs(X - ref)) < 1e-12 * np.sqrt(nx * ny) * np.abs(ref).max() / np.sqrt(nx * ny) + 1e-11


@pytest.mark.parametrize("nx,ny", [(32, 32), (64, 64), (96, 96), (128, 128), (128, 32), (48, 36), (192, 192),
                                   (192, 96), (256, 128)] + SMALL_AND_RADIX3)
def test_fft_backward_non_hermitian_matches_oracle(dns, nx, ny):
    _say(f"fft backward {nx}x{ny}")
    rng = np.random.default_rng(7 + nx)
    F = rng.standard_normal((3, ny, nx // 2 + 1)) + 1j * rng.standard_normal((3, ny, nx // 2 + 1))
    got = dns.FFTPlan(nx, ny).backward(F)
    want = orc.fft_c2r(F, nx)
    assert np.max(np.abs(got - want)) < 1e-14 * max(1.0, np.abs(F).max())


def test_mfma_f64_lane_layout(dns):
    """v_mfma_f64_16x16x4f64 operand/result lane layout assumed by the contractions:
    A(i,k) at lane i + 16k, B(k,j) at lane j + 16k, C(i,j) at lane j + 16(i%4), item i/4."""
    import ctypes
    out = np.empty(4096 * 256)
    assert dns.lib().dns_debug_mfma_probe(out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))) == 0
    P = out.reshape(64, 64, 64, 4)          # [la, lb, lane, item]
    want = np.zeros_like(P)
    for la in range(64):
        for lb in range(64):
            if la // 16 == lb // 16:
                i, j = la % 16, lb % 16
                want[la, lb, j + 16 * (i % 4), i // 4] = 1.0
    nz_idx = np.argwhere(P != 0)
    _say(f"probe nonzeros {len(nz_idx)}; first: {nz_idx[:8].tolist()}")
    np.testing.assert_array_equal(P, want)


def test_mfma_f32_lane_layout_probe(dns):
    """v_mfma_f32_16x16x4f32 lane layout assumed by the FP32 contraction: operands as in FP64
    (A(i,k) at lane i + 16k, B(k,j) at lane j + 16k), C(i,j) at lane j + 16(i/4), item i%4."""
    import ctypes
    out = np.empty(4096 * 256)
    assert dns.lib().dns_debug_mfma_probe_f32(out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))) == 0
    P = out.reshape(64, 64, 64, 4)
    want = np.zeros_like(P)
    for la in range(64):
        for lb in range(64):
            if la // 16 == lb // 16:
                i, j = la % 16, lb % 16
                want[la, lb, j + 16 * (i // 4), i % 4] = 1.0
    np.testing.assert_array_equal(P, want)


@pytest.mark.parametrize("nz,exact", [(17, 0), (33, 0), (65, 1), (129, 1), (129, 0), (9, 0), (48, 1)])
def test_zderiv_mfma_matches_numpy(dns, nz, exact):
    g = dns.Simulation(nx=32, ny=16, nz=nz, lgl_derivative=exact)
    rng = np.random.default_rng(nz)
    spec = rng.standard_normal(g.spectral_shape) + 1j * rng.standard_normal(g.spectral_shape)
    D = orc.diff_matrix_exact(nz) if exact else orc.diff_matrix(nz)
    want = np.einsum("km,myx->kyx", D, spec)
    got = g.zderiv(spec)
    err = np.max(np.abs(got - want)) / (np.abs(D).max() * np.abs(spec).max())
    _say(f"zderiv nz={nz} exact={exact}: scaled err {err:.3e}")
    assert err < 1e-13 * nz


def test_operator_bases_reconstruct(dns):
    for nz, exact in ((17, 0), (33, 0), (65, 1), (129, 1), (129, 0)):
        g = dns.Simulation(nx=16, ny=16, nz=nz, dt=0.002, re=180.0, lgl_derivative=exact)
        eh, ep = g.operator_error()
        _say(f"operator reconstruction nz={nz} exact={exact}: {eh:.2e} {ep:.2e}")
        assert eh < 1e-15 and ep < 1e-15, (nz, eh, ep)


def test_reference_d1_at_nz65_is_rejected_like_zgesv(dns):
    """With the reference D1, nz = 65 makes the mode-(0,0) Poisson matrix exactly singular,
    where the reference's zgesv stops (DNS_pressure_BC_3D.f90:1715-1718)."""
    with pytest.raises(dns.DnsError) as e:
        dns.Simulation(nx=16, ny=16, nz=65, lgl_derivative=0)
    assert e.value.code == 5


def test_stage_parity_channel32(dns):
    """Each stage of one step against the oracle, on identical fields."""
    _say("stage parity 32x32x33")
    o, g = pair(**CHANNEL32)
    nx, ny = CHANNEL32["nx"], CHANNEL32["ny"]
    # warm both up 3 steps so p, wall gradients and all modes are nonzero
    o.step(3)
    for _ in range(3):
        g.step(1)
    g.synchronize()
    e = field_errors(g.get_fields(), o.get_fields())
    _say(f"  after 3 steps {e}")
    check_parity(e, 1e-12, W_SELF_FP64)
    g.begin_step()
    o.compute_source_terms(); g.compute_source_terms()
    G = o.diagnostics()["pressure_gradient"]
    want = [orc.fft_r2c(s) for s in o.source_terms()]
    want[0][:, 0, 0] -= G * nx * ny
    got = [g.spectral(7 + c) for c in range(3)]
    scale = np.sqrt(sum(np.linalg.norm(w) ** 2 for w in want))
    for c in range(3):
        err = np.linalg.norm(got[c] - want[c]) / scale
        _say(f"  S_{c}: {err:.2e}")
        assert err < 1e-12, c
    o.viscous_step(); g.viscous_step()
    want = [orc.fft_r2c(s) for s in o.star_fields()]
    scale = np.sqrt(sum(np.linalg.norm(w) ** 2 for w in want))
    for c in range(3):
        err = np.linalg.norm(g.spectral(4 + c) - want[c]) / scale
        _say(f"  star_{c}: {err:.2e}")
        assert err < 1e-12, c
    o.solve_pressure(); g.solve_pressure()
    err = rel_l2(g.spectral(12), orc.fft_r2c(o.phi()))
    _say(f"  phi: {err:.2e}")
    assert err < 1e-10
    o.pressure_correction(); g.pressure_correction()
    o.apply_boundary_conditions(); g.apply_boundary_conditions()
    e = field_errors(g.get_fields(), o.get_fields())
    _say(f"  after stages {e}")
    check_parity(e, 1e-11, W_SELF_FP64)


def test_fused_step_equals_stages(dns):
    _say("fused vs stage API")
    _, g1 = pair(**CHANNEL32)
    _, g2 = pair(**CHANNEL32)
    for _ in range(4):
        g1.step(1)
        g2.begin_step(); g2.compute_source_terms(); g2.viscous_step(); g2.solve_pressure()
        g2.pressure_correction(); g2.apply_boundary_conditions()
    for a, b in zip(g1.get_fields(), g2.get_fields()):
        assert rel_l2(a, b) < 1e-14


def test_graph_step_equals_eager(dns):
    _say("graph vs eager")
    _, g1 = pair(**CHANNEL32)
    _, g2 = pair(**CHANNEL32)
    g1.step(5)
    g2.step(2, graph=True)
    g2.synchronize()
    g2.step(3, graph=True)
    for a, b in zip(g1.get_fields(), g2.get_fields()):
        np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("dealias", [1, 0])
def test_channel32_100_steps(dns, dealias):
    """The FP64 parity bar: u, v, w, p within relative L2 1e-10 of the FP64 oracle after 100 steps,
    with the same energy history and max divergence order."""
    _say(f"32x32x33 100 steps dealias={dealias}")
    kw = dict(CHANNEL32, use_dealias_23_rule=dealias)
    o, g = pair(**kw)
    # the oracle's own twin (plane transforms factored otherwise): w's rounding floor of this run
    t = orc.Oracle(fft_variant=1, **kw)
    t.init_fields()
    for block in range(10):
        o.step(10)
        t.step(10)
        g.step(10)
        do, dg = o.diagnostics(), g.diagnostics()
        assert abs(do["energy"] - dg["energy"]) < 1e-12 * do["energy"], block
        # |w|max ~ 1e-6 |u|max: compared on the velocity scale (its own last bits are rounding)
        assert abs(do["wmax"] - dg["wmax"]) < 1e-10 * do["umax"]
        assert abs(do["vmax"] - dg["vmax"]) < 1e-10 * do["umax"]
    fo = o.get_fields(); fg = g.get_fields()
    errs = field_errors(fg, fo)
    floor = field_errors(t.get_fields(), fo)
    _say(f"  parity after 100 steps {errs}")
    _say(f"  oracle vs its twin (rounding floor) {floor}")
    # u, v, p and w (velocity scale) at 1e-10; w's own error within 4x the oracle's own floor
    check_parity(errs, 1e-10, 4 * floor["w_self"])
    # same max-divergence order: both at the level of the collocation projection's residue
    do_, dg_ = o.max_divergence(), g.max_divergence()
    _say(f"  max divergence oracle {do_:.3e} gpu {dg_:.3e}")
    assert dg_ < 2 * do_ + 1e-13 and do_ < 2 * dg_ + 1e-13


@pytest.mark.parametrize("istart", [0, 14, 3])
def test_pi_controller_trace(dns, istart):
    """Method 2 (constant bulk velocity), the controller called with istart (freq 7): istart = 0
    runs it every step on its Euler branch, 14 every step on the trapezoid branch, 3 never
    (G stays the target)."""
    _say(f"PI controller istart={istart}")
    kw = dict(CHANNEL32, flow_control_method=2, perturbation_kind=1, seed=11, noise_kmax=4,
              perturbation_amplitude=0.05, istart=istart, controller_update_freq=7)
    o, g = pair(**kw)
    for _ in range(20):
        o.step(1); g.step(1)
        do, dg = o.diagnostics(), g.diagnostics()
        for key in ("pressure_gradient", "bulk_velocity", "avg_pressure_gradient", "bulk_error_integral"):
            assert abs(do[key] - dg[key]) < 1e-12 * max(1.0, abs(do[key])), key
    e = field_errors(g.get_fields(), o.get_fields())
    _say(f"  PI run {e}")
    check_parity(e, 1e-10, W_SELF_FP64)


def test_trace_config_matches_reference_stdout(dns):
    """The reference's own 15-step trace (standard_output.txt:88-104) from the GPU path."""
    import json
    import os
    fx = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "standard_output_trace.json")))
    g = dns.Simulation(nx=32, ny=32, nz=17, xlen=2 * np.pi, ylen=np.pi, re=500.0, dt=0.002, u_centerline=1.5,
                       flow_control_method=1, target_pressure_gradient=0.006, use_dealias_23_rule=1,
                       enable_perturbations=1, perturbation_kind=0, perturbation_amplitude=0.01)
    g.init_fields()
    for row in fx["rows"]:
        g.step(1)
        d = g.diagnostics()
        got = ["%.4E" % d[f] for f in ("umax", "vmax", "wmax", "energy", "pressure_gradient")]
        want = ["%.4E" % float(row[k]) for k in ("umax", "vmax", "wmax", "E", "dPdx")]
        assert got == want, row["step"]


def test_seeded_noise_production_shape_short(dns):
    """128x128x129, method 2, seeded noise, exact LGL D1: 3 steps against the oracle."""
    _say("128^2x129 3 steps")
    kw = dict(nx=128, ny=128, nz=129, xlen=4 * np.pi, ylen=2 * np.pi, re=180.0, dt=0.002, use_crank_nicolson=1,
              use_dealias_23_rule=1, flow_control_method=2, target_pressure_gradient=2.0 / 180.0,
              target_bulk_velocity=2.0 / 3.0, controller_gain=0.15, u_centerline=1.0, enable_perturbations=1,
              perturbation_kind=1, seed=2024, noise_kmax=8, perturbation_amplitude=0.05, lgl_derivative=1)
    o, g = pair(**kw)
    o.step(3); g.step(3)
    fo, fg = o.get_fields(), g.get_fields()
    errs = field_errors(fg, fo)
    _say(f"  128^2x129 after 3 steps {errs}")
    check_parity(errs, 1e-10, W_SELF_FP64)


@pytest.mark.parametrize("nx,ny", [(96, 96), (128, 96), (192, 48), (96, 144), (144, 96)])
def test_largest_class_step_shapes(dns, nx, ny):
    """FP64 steps on planes of the 576- and 768-thread transform classes (4608 and 6144 values:
    96 x 96, 192 x 48, 128 x 96) and of the largest class with other first-stage radices (6912
    values; the step's c2r runs one persistent block per CU there): 4 steps against the oracle."""
    _say(f"{nx}x{ny}x17 4 steps")
    kw = dict(CHANNEL32, nx=nx, ny=ny, nz=17, xlen=4 * np.pi, ylen=2 * np.pi, flow_control_method=2,
              perturbation_kind=1, seed=7, noise_kmax=6, perturbation_amplitude=0.05)
    o, g = pair(**kw)
    o.step(4); g.step(4)
    errs = field_errors(g.get_fields(), o.get_fields())
    _say(f"  {nx}x{ny}x17 after 4 steps {errs}")
    check_parity(errs, 1e-10, W_SELF_FP64)


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("nx,ny,nz", [(8, 4, 5), (8, 6, 9), (12, 6, 9), (18, 18, 17), (24, 54, 17), (54, 24, 17),
                                      (162, 48, 9), (8, 4, 129), (12, 6, 100), (18, 18, 144)])
def test_small_and_radix3_grids(dns, nx, ny, nz, precision):
    """The smallest grid accepted (8 x 4 x 5), planes far smaller than a wavefront's share,
    plane lengths with several radix-3 stages (nx/2 = 9, 27, 81; ny = 27 x 2), and tiny planes
    under the four-wave solves' packed mode space (nz = 100 to 144): 4 steps against the FP64
    oracle, FP64 at 1e-10 and mixed precision at 1e-6."""
    _say(f"{nx}x{ny}x{nz} precision {precision} 4 steps")
    kw = dict(CHANNEL32, nx=nx, ny=ny, nz=nz, xlen=4 * np.pi, ylen=2 * np.pi, flow_control_method=2,
              perturbation_kind=1, seed=21, noise_kmax=2, perturbation_amplitude=0.05, lgl_derivative=int(nz > 33))
    o = orc.Oracle(**kw)
    o.init_fields()
    g = dns.Simulation(precision=precision, **kw)
    g.set_fields(*o.get_fields())
    o.step(4); g.step(4)
    errs = field_errors(g.get_fields(), o.get_fields())
    _say(f"  {nx}x{ny}x{nz} after 4 steps {errs}")
    if precision == 0:
        check_parity(errs, 1e-10, W_SELF_FP64)
    else:
        check_parity(errs, 1e-6, W_SELF_MIXED)


@pytest.mark.parametrize("nz", [81, 97, 130, 132, 144])
def test_three_wave_and_two_pair_tiles(dns, nz):
    """Folded operators of three row-block pairs (nz = 81, 97: three-wave blocks, one wave per
    pair) and of five to eight (nz = 130 to 144: four waves with two pairs each; 144 is the
    largest nz accepted): 10 FP64 steps against the oracle."""
    _say(f"48x36x{nz} 10 steps")
    kw = dict(CHANNEL32, nx=48, ny=36, nz=nz, lgl_derivative=1, flow_control_method=2, perturbation_kind=1,
              seed=13, noise_kmax=5, perturbation_amplitude=0.05)
    o, g = pair(**kw)
    o.step(10); g.step(10)
    errs = field_errors(g.get_fields(), o.get_fields())
    _say(f"  48x36x{nz} after 10 steps {errs}")
    check_parity(errs, 1e-10, W_SELF_FP64)


@pytest.mark.parametrize("nz,exact", [(18, 0), (24, 0), (34, 1), (36, 1), (48, 1), (64, 1), (66, 1), (68, 1),
                                      (128, 1)])
def test_even_nz_steps(dns, nz, exact):
    """Even nz (no centre node: the folded operators have no centre row; one-, two-, three- and
    four-wave contraction kernels), including halves of 16 k + 1 rows (nz = 34, 66; the
    Helmholtz interior at nz = 36, 68), which once lost the odd part of their last row: 10 FP64
    steps against the oracle."""
    _say(f"48x36x{nz} exact={exact} 10 steps")
    kw = dict(CHANNEL32, nx=48, ny=36, nz=nz, lgl_derivative=exact, flow_control_method=2, perturbation_kind=1,
              seed=12, noise_kmax=5, perturbation_amplitude=0.05)
    o, g = pair(**kw)
    o.step(10); g.step(10)
    errs = field_errors(g.get_fields(), o.get_fields())
    _say(f"  48x36x{nz} after 10 steps {errs}")
    check_parity(errs, 1e-10, W_SELF_FP64)


@pytest.mark.parametrize("precision", [1, 2])
@pytest.mark.parametrize("nz", [18, 34, 64, 66, 128, 130])
def test_even_nz_mixed_precision(dns, nz, precision):
    """Mixed precision at even nz, with the vorticity's D1 in FP64 arithmetic (1) and on FP32
    MFMA (2, whose result layout differs): 10 steps within 1e-6 of the FP64 oracle."""
    _say(f"48x36x{nz} precision {precision} 10 steps")
    kw = dict(CHANNEL32, nx=48, ny=36, nz=nz, lgl_derivative=1, flow_control_method=2, perturbation_kind=1,
              seed=14, noise_kmax=5, perturbation_amplitude=0.05)
    o = orc.Oracle(**kw)
    o.init_fields()
    g = dns.Simulation(precision=precision, **kw)
    g.set_fields(*o.get_fields())
    o.step(10); g.step(10)
    errs = field_errors(g.get_fields(), o.get_fields())
    _say(f"  48x36x{nz} precision {precision} after 10 steps {errs}")
    check_parity(errs, 1e-6, W_SELF_MIXED)


@pytest.mark.parametrize("kw", [dict(CHANNEL32),
                                dict(CHANNEL32, nx=48, ny=36, nz=33, flow_control_method=2, perturbation_kind=1,
                                     seed=4, noise_kmax=5, perturbation_amplitude=0.05, use_dealias_23_rule=0),
                                dict(CHANNEL32, nx=64, ny=32, nz=65, lgl_derivative=1, use_crank_nicolson=0,
                                     perturbation_kind=1, seed=8, noise_kmax=6, perturbation_amplitude=0.05)])
def test_gpu_matches_cpu_emulation_of_same_algorithm(dns, kw):
    """Same algebra on both sides (the library's operator matrices): implementation check."""
    from tests.kernel_emulation import Emulator
    _say(f"gpu vs emulation {kw['nx']}x{kw['ny']}x{kw['nz']}")
    o = orc.Oracle(**kw)
    o.init_fields()
    e = Emulator(**kw)
    f0 = o.get_fields()
    e.set_fields(*f0)
    g = dns.Simulation(**kw)
    g.set_fields(*f0)
    for _ in range(5):
        e.step()
    g.step(5)
    errs = field_errors(g.get_fields(), e.get_fields())
    _say(f"  {errs}")
    check_parity(errs, 1e-12, W_SELF_FP64)


def test_packed_mode_space_stages_against_oracle(dns):
    """nz = 129 puts the per-mode solves on their four-wave kernels, which walk only the
    independent coefficients and store the mirrored modes as conjugates: every stage's spectral
    output (all modes, mirrored and self-conjugate ones included) against the oracle's r2c."""
    _say("packed mode space 48x36x129 stages")
    kw = dict(CHANNEL32, nx=48, ny=36, nz=129, xlen=4 * np.pi, ylen=2 * np.pi, lgl_derivative=1,
              flow_control_method=2, perturbation_kind=1, seed=21, noise_kmax=6, perturbation_amplitude=0.05)
    o, g = pair(**kw)
    o.step(2)
    g.step(2)
    g.begin_step()
    o.compute_source_terms(); g.compute_source_terms()
    G = o.diagnostics()["pressure_gradient"]
    want = [orc.fft_r2c(s) for s in o.source_terms()]
    want[0][:, 0, 0] -= G * kw["nx"] * kw["ny"]
    scale = np.sqrt(sum(np.linalg.norm(w) ** 2 for w in want))
    for c in range(3):
        err = np.linalg.norm(g.spectral(7 + c) - want[c]) / scale
        _say(f"  S_{c}: {err:.2e}")
        assert err < 1e-11, c
    o.viscous_step(); g.viscous_step()
    want = [orc.fft_r2c(s) for s in o.star_fields()]
    scale = np.sqrt(sum(np.linalg.norm(w) ** 2 for w in want))
    for c in range(3):
        got = g.spectral(4 + c)
        err = np.linalg.norm(got - want[c]) / scale
        _say(f"  star_{c}: {err:.2e}")
        assert err < 1e-11, c
        # mirrored modes are exact conjugates, self-conjugate modes exactly real
        ny = kw["ny"]
        for i in (0, kw["nx"] // 2):
            assert np.array_equal(got[:, ny - 1:ny // 2:-1, i], np.conj(got[:, 1:ny // 2, i]))
            assert np.all(got[:, 0, i].imag == 0) and np.all(got[:, ny // 2, i].imag == 0)
    o.solve_pressure(); g.solve_pressure()
    err = rel_l2(g.spectral(12), orc.fft_r2c(o.phi()))
    _say(f"  phi: {err:.2e}")
    assert err < 1e-10
    o.pressure_correction(); g.pressure_correction()
    o.apply_boundary_conditions(); g.apply_boundary_conditions()
    e = field_errors(g.get_fields(), o.get_fields())
    _say(f"  after stages {e}")
    check_parity(e, 1e-10, W_SELF_FP64)


@pytest.mark.parametrize("precision", [0, 1])
def test_packed_mode_space_batch_member_equals_alone(dns, precision):
    """Members of a batch at nz = 129 (packed mode space) are bitwise the runs alone, and graph
    replay is bitwise the eager launches."""
    base = dict(CHANNEL32, nx=32, ny=24, nz=129, lgl_derivative=1, precision=precision, perturbation_kind=1,
                noise_kmax=4, perturbation_amplitude=0.04)
    mem = [dict(re=180.0, seed=3), dict(re=260.0, seed=4, flow_control_method=2)]
    _say(f"packed batch precision {precision}")
    batch = dns.Simulation(members=mem, **base)
    # the packed mode space's tables are counted by the estimate the serving loop admits by
    assert batch.device_bytes == dns.estimate_bytes(nmembers=len(mem), **{**base, **mem[0]})
    batch.init_fields()
    batch.step(6, graph=True)
    eager = dns.Simulation(members=mem, **base)
    eager.init_fields()
    eager.step(6, graph=False)
    for b in range(len(mem)):
        for x, y in zip(batch.get_fields(member=b), eager.get_fields(member=b)):
            assert np.array_equal(x, y), ("graph vs eager", b)
    eager.close()
    for b, m in enumerate(mem):
        one = dns.Simulation(**{**base, **m})
        one.init_fields()
        one.step(6, graph=True)
        for f, (x, y) in enumerate(zip(batch.get_fields(member=b), one.get_fields())):
            assert np.array_equal(x, y), (b, f)
        one.close()
    batch.close()


def test_context_wavenumber_tables_bitwise_equal_oracle(dns):
    """The device-resident xw, yw of a context (ny != nx) bit for bit against the oracle."""
    g = dns.Simulation(nx=48, ny=36, nz=17, xlen=4 * np.pi, ylen=2 * np.pi)
    xw, yw = g.device_wavenumbers()
    oxw, oyw = orc.wavenumbers(48, 36, 4 * np.pi, 2 * np.pi)
    np.testing.assert_array_equal(xw, oxw)
    np.testing.assert_array_equal(yw, oyw)


@pytest.mark.parametrize("nz", [17, 33, 129])
def test_laminar_bulk_velocity_on_device(dns, nz):
    """calculate_bulk_velocity_3d of the Poiseuille profile: 2 uc / 3 (exact LGL quadrature)."""
    g = dns.Simulation(nx=32, ny=24, nz=nz, u_centerline=1.5, enable_perturbations=0, lgl_derivative=1)
    g.init_fields()
    assert abs(g.diagnostics()["bulk_velocity"] - 1.0) < 1e-12


def test_restart_with_pressure_keeps_zero_wall_gradient_for_first_step(dns):
    """Fields set with a nonzero p (a restart read): the reference's grad_p_prev stays zero until
    the first pressure correction (DNS_pressure_BC_3D.f90:957,1923), so the first step's
    Helmholtz wall values are zero.  The oracle from the same fields with zero wall gradients
    is the reference; one whose wall gradients follow p differs, so the test can tell."""
    _say("restart with p")
    kw = dict(CHANNEL32, flow_control_method=2, perturbation_kind=1, seed=19, noise_kmax=4,
              perturbation_amplitude=0.05)
    src = orc.Oracle(**kw)
    src.init_fields()
    src.step(4)
    f = src.get_fields()
    gx, gy = src.get_wall_gradients()
    assert np.abs(f[3]).max() > 0 and np.abs(gx).max() > 0
    o = orc.Oracle(**kw)
    o.set_fields(*f)                   # wall gradients stay zero, as after read_restart_file_3d
    alt = orc.Oracle(**kw)
    alt.set_fields(*f)
    alt.set_wall_gradients(gx, gy)     # what a step with gradients from p would do
    g = dns.Simulation(**kw)
    g.set_fields(*f)
    for _ in range(5):
        o.step(1); alt.step(1); g.step(1)
    e = field_errors(g.get_fields(), o.get_fields())
    e_alt = field_errors(alt.get_fields(), o.get_fields())
    _say(f"  restart parity {e}; gradient-from-p oracle differs by {e_alt}")
    assert max(e_alt["u"], e_alt["v"]) > 1e-8
    check_parity(e, 1e-10, W_SELF_FP64)
    # a second set_fields with p and graph replay: the flag applies to the first replay only
    g2 = dns.Simulation(**kw)
    g2.set_fields(*f)
    g2.step(2, graph=True)
    g2.step(3, graph=True)
    for a, b in zip(g2.get_fields(), g.get_fields()):
        np.testing.assert_array_equal(a, b)
