"""The Fortran driver and the drop-in FFT module on the GPU (ISO_C_BINDING -> C ABI -> HIP)."""
import json
import os
import re
import subprocess

import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "f90_dns_amd", "bin")


def _run(args, cwd):
    print("running", args, flush=True)
    return subprocess.run(args, capture_output=True, text=True, cwd=cwd, timeout=120)


def test_fortran_fft_module_known_answers(tmp_path):
    exe = os.path.join(BIN, "fft_module_check")
    if not os.path.exists(exe):
        pytest.skip("Fortran programs not built")
    r = _run([exe], tmp_path)
    print(r.stdout)
    assert r.returncode == 0 and "FFT MODULE PASSED" in r.stdout


@pytest.mark.parametrize("weak", [False, True])
@pytest.mark.parametrize("fused", [False, True])
def test_fortran_driver_reproduces_reference_trace(tmp_path, fused, weak):
    """Same input as the reference runs of standard_output.txt and weak_form_output.txt (the
    same namelist with use_weak_form_divergence = .true.); every printed diagnostic row must
    equal the reference's (:88-104 of each)."""
    exe = os.path.join(BIN, "dns_3d_gpu")
    if not os.path.exists(exe):
        pytest.skip("Fortran driver not built")
    src = open(os.path.join(ROOT, "tests", "golden", "trace_input.dat")).read()
    if weak:
        assert "use_weak_form_divergence = .false." in src
        src = src.replace("use_weak_form_divergence = .false.", "use_weak_form_divergence = .true.")
    if fused:
        src += "&gpu\n  fused_step = .true.\n/\n"
    inp = tmp_path / "input.dat"
    inp.write_text(src)
    r = _run([exe, str(inp)], tmp_path)
    print(r.stdout[-3000:])
    assert r.returncode == 0, r.stdout + r.stderr
    trace = "weak_form_output_trace.json" if weak else "standard_output_trace.json"
    fx = json.load(open(os.path.join(ROOT, "tests", "golden", trace)))
    lines = [l for l in r.stdout.splitlines() if re.match(r"\s*Step\s+\d+, t=", l)]
    assert len(lines) == len(fx["rows"])
    for line, row in zip(lines, fx["rows"]):
        want = (f" Step {row['step']:6d}, t={row['t']:12.6f}, |u|max={float(row['umax']):12.4E}, "
                f"|v|max={float(row['vmax']):12.4E}, |w|max={float(row['wmax']):12.4E}, E={float(row['E']):12.4E}, "
                f"dP/dx={float(row['dPdx']):12.4E}")
        norm = lambda s: re.sub(r"\s+", " ", s.strip().rstrip(","))
        assert norm(line) == norm(want), (line, want)
    import glob
    dirs = glob.glob(str(tmp_path / "input_????????_??????"))
    assert len(dirs) == 1
    restart = os.path.join(dirs[0], "restart.dat")
    n = 32 * 32 * 17
    assert os.path.getsize(restart) == 18 + 4 + 16 + 12 + 24 + 7 * 8 * n + 32
    # the reference's restart detection finds nothing in the new directory and starts new
    assert "WARNING: Restart file not found: " + dirs[0].split("/")[-1] + "/restart.dat" in r.stdout.replace(
        str(tmp_path) + "/", "")
    assert "STARTING NEW SIMULATION" in r.stdout


def test_fortran_driver_output_files(tmp_path):
    """output_time_series.dat and flow_statistics.dat in <input>_<YYYYMMDD_HHMMSS>/, with the
    reference's layouts (DNS_pressure_BC_3D.f90:2664-2694, 2702-2748, 2849-2948): the time
    series carries the printed diagnostics of every nwrt step, the statistics the final state,
    its wall shear from the restart file's u over the LGL nodes."""
    import glob

    import numpy as np

    from oracle import oracle as orc
    exe = os.path.join(BIN, "dns_3d_gpu")
    if not os.path.exists(exe):
        pytest.skip("Fortran driver not built")
    inp = tmp_path / "input.dat"
    inp.write_text(open(os.path.join(ROOT, "tests", "golden", "trace_input.dat")).read())
    r = _run([exe, str(inp)], tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    dirs = glob.glob(str(tmp_path / "input_????????_??????"))
    assert len(dirs) == 1, os.listdir(tmp_path)
    ts = open(os.path.join(dirs[0], "output_time_series.dat")).read().splitlines()
    assert ts[0] == "# DNS 3D Time Series Data"
    assert ts[1].startswith("# Columns: step, time, u_max")
    rows = [l.split() for l in ts[2:]]
    fx = json.load(open(os.path.join(ROOT, "tests", "golden", "standard_output_trace.json")))
    assert len(rows) == len(fx["rows"])
    for row, want in zip(rows, fx["rows"]):
        assert len(row) == 7 and int(row[0]) == int(want["step"])
        got = [float(x) for x in row[2:]]
        ref = [float(want[k]) for k in ("umax", "vmax", "wmax", "E", "dPdx")]
        for g, w_ in zip(got, ref):
            assert abs(g - w_) <= 5e-4 * abs(w_), (row, want)   # the trace holds 5 digits
    st = {}
    for l in open(os.path.join(dirs[0], "flow_statistics.dat")).read().splitlines():
        if " = " in l:
            k, v = l.split(" = ")
            st[k.strip()] = float(v)
    assert st["u_max"] == float(rows[-1][2]) and st["kinetic_energy"] == float(rows[-1][5])
    n = 32 * 32 * 17
    raw = open(os.path.join(dirs[0], "restart.dat"), "rb").read()
    u = np.frombuffer(raw, dtype="<f8", count=n, offset=18 + 4 + 16 + 12 + 24).reshape(17, 32, 32)
    z, _ = orc.lgl(17)
    nu = 1.0 / 500.0
    lower = nu * np.mean((u[1] - u[0]) / (z[1] - z[0]))
    upper = nu * np.mean((u[16] - u[15]) / (z[16] - z[15]))
    assert abs(st["tau_wall_lower"] - lower) <= 1e-6 * abs(lower), (st, lower)
    assert abs(st["tau_wall_upper"] - upper) <= 1e-6 * abs(upper), (st, upper)
    ut = np.sqrt((abs(lower) + abs(upper)) / 2)
    assert abs(st["re_tau"] - ut * 500.0) <= 1e-6 * ut * 500.0
    # every 10 nwrt steps (here step 10): a snapshot of every point and the plane-mean profile
    snap = np.loadtxt(os.path.join(dirs[0], "velocity_field_00010.dat"), comments="#")
    prof = np.loadtxt(os.path.join(dirs[0], "spanwise_profile_00010.dat"), comments="#")
    assert snap.shape == (n, 9) and prof.shape == (17, 5)
    i, j, k = snap[:, 0].astype(int), snap[:, 1].astype(int), snap[:, 2].astype(int)
    assert (i[:32] == np.arange(1, 33)).all() and k[0] == 1 and k[-1] == 17 and j[-1] == 32
    assert np.allclose(snap[:, 3], (i - 1) * 6.283185307179586 / 32, rtol=1e-6, atol=1e-12)
    assert np.allclose(snap[:, 5], z[k - 1], rtol=1e-6, atol=1e-12)
    assert np.allclose(prof[:, 1], z, rtol=1e-6, atol=1e-12)
    for c in range(3):   # plane means of the snapshot's own (rounded) values
        means = snap[:, 6 + c].reshape(17, 32 * 32).mean(axis=1)
        assert np.allclose(prof[:, 2 + c], means, rtol=1e-5, atol=1e-12), c


def _restart_fields(path, nx, ny, nz):
    import numpy as np
    raw = open(path, "rb").read()
    n = nx * ny * nz
    head = 18
    assert raw[:head] == b"DNS3D_RESTART_v1.0"
    step = int(np.frombuffer(raw, "<i4", 1, head)[0])
    t, dt = np.frombuffer(raw, "<f8", 2, head + 4)
    dims = np.frombuffer(raw, "<i4", 3, head + 20)
    off = head + 4 + 16 + 12 + 24
    f = [np.frombuffer(raw, "<f8", n, off + 8 * n * i).reshape(nz, ny, nx) for i in range(7)]
    ctl = np.frombuffer(raw, "<f8", 4, off + 56 * n)
    return step, t, tuple(int(x) for x in dims), f, ctl


def test_fortran_driver_restart_files_summary_and_resume(tmp_path):
    """restart.dat every restart_write_freq steps before the step body and after the last step
    (DNS_pressure_BC_3D.f90:307-309, 367), simulation_summary.txt (:2594-2659), the timing
    summary every 100 steps (:2207-2230), and a run resumed from a restart file (&gpu
    restart_file) against the oracle resumed from the same fields with zero wall gradients."""
    import glob

    import numpy as np

    from oracle import oracle as orc
    exe = os.path.join(BIN, "dns_3d_gpu")
    if not os.path.exists(exe):
        pytest.skip("Fortran driver not built")
    base = open(os.path.join(ROOT, "tests", "golden", "trace_input.dat")).read()
    assert "nsteps = 15" in base and "restart_write_freq" in base
    src = re.sub(r"restart_write_freq\s*=\s*\d+", "restart_write_freq = 4", base).replace("nsteps = 15", "nsteps = 10")
    a_dir = tmp_path / "a"
    a_dir.mkdir()
    (a_dir / "first.dat").write_text(src)
    r = _run([exe, "first.dat"], a_dir)
    assert r.returncode == 0, r.stdout + r.stderr
    written = re.findall(r"Restart file written at step (\d+), time =\s*([0-9.]+)", r.stdout)
    assert [int(s) for s, _ in written] == [4, 8, 10], written
    # periodic files carry the post-increment time of their step (:296, 307)
    assert abs(float(written[0][1]) - 4 * 0.002) < 1e-9 and abs(float(written[2][1]) - 10 * 0.002) < 1e-9
    dirs = glob.glob(str(a_dir / "first_????????_??????"))
    assert len(dirs) == 1
    summ = open(os.path.join(dirs[0], "simulation_summary.txt"), encoding="utf-8").read().splitlines()
    assert summ[1] == "   3D DNS SIMULATION SUMMARY" and " Input file: first.dat" in summ
    assert "   Grid size: 32 \u00d7 32 \u00d7 17" in summ and "   Total steps: 10" in summ
    assert "   Restart write frequency: 4" in summ and "   Restart mode: DISABLED (new simulation)" in summ
    step, t, dims, f, ctl = _restart_fields(os.path.join(dirs[0], "restart.dat"), 32, 32, 17)
    assert step == 10 and dims == (32, 32, 17) and abs(t - 0.02) < 1e-15
    # resume for 5 more steps from that file
    b_dir = tmp_path / "b"
    b_dir.mkdir()
    (b_dir / "second.dat").write_text(src.replace("nsteps = 10", "nsteps = 15") +
                                      "&gpu\n  restart_file = '%s'\n/\n" % os.path.join(dirs[0], "restart.dat"))
    r = _run([exe, "second.dat"], b_dir)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "RESTARTING FROM CHECKPOINT" in r.stdout and "Remaining steps: 5" in r.stdout
    steps = [int(m) for m in re.findall(r"^ Step\s+(\d+), t=", r.stdout, flags=re.M)]
    assert steps == [11, 12, 13, 14, 15]
    d2 = glob.glob(str(b_dir / "second_????????_??????"))[0]
    step2, t2, _, f2, _ = _restart_fields(os.path.join(d2, "restart.dat"), 32, 32, 17)
    assert step2 == 15 and abs(t2 - 15 * 0.002) < 1e-12
    kw = dict(nx=32, ny=32, nz=17, xlen=2 * np.pi, ylen=np.pi, re=500.0, dt=0.002, u_centerline=1.5,
              flow_control_method=1, target_pressure_gradient=0.006, use_dealias_23_rule=1,
              enable_perturbations=1, perturbation_kind=0, perturbation_amplitude=0.01)
    o = orc.Oracle(**kw)
    o.set_fields(f[0], f[1], f[2], f[6])          # wall gradients stay zero, as the reference's
    o.set_controller(ctl[0], ctl[2], ctl[3])
    o.step(5)
    ref = o.get_fields()
    for name, got, want in zip("uvwp", (f2[0], f2[1], f2[2], f2[6]), ref):
        err = np.linalg.norm(got - want) / np.linalg.norm(want)
        print(f"resumed {name}: {err:.2e}")
        assert err < (1e-8 if name == "w" else 1e-10), name
