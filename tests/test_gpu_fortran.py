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


@pytest.mark.parametrize("fused", [False, True])
def test_fortran_driver_reproduces_reference_trace(tmp_path, fused):
    """Same input as the reference run of standard_output.txt; every printed diagnostic row must
    equal the reference's (standard_output.txt:88-104)."""
    exe = os.path.join(BIN, "dns_3d_gpu")
    if not os.path.exists(exe):
        pytest.skip("Fortran driver not built")
    src = open(os.path.join(ROOT, "tests", "golden", "trace_input.dat")).read()
    if fused:
        src += "&gpu\n  fused_step = .true.\n/\n"
    inp = tmp_path / "input.dat"
    inp.write_text(src)
    r = _run([exe, str(inp)], tmp_path)
    print(r.stdout[-3000:])
    assert r.returncode == 0, r.stdout + r.stderr
    fx = json.load(open(os.path.join(ROOT, "tests", "golden", "standard_output_trace.json")))
    lines = [l for l in r.stdout.splitlines() if re.match(r"\s*Step\s+\d+, t=", l)]
    assert len(lines) == len(fx["rows"])
    for line, row in zip(lines, fx["rows"]):
        want = (f" Step {row['step']:6d}, t={row['t']:12.6f}, |u|max={float(row['umax']):12.4E}, "
                f"|v|max={float(row['vmax']):12.4E}, |w|max={float(row['wmax']):12.4E}, E={float(row['E']):12.4E}, "
                f"dP/dx={float(row['dPdx']):12.4E}")
        norm = lambda s: re.sub(r"\s+", " ", s.strip().rstrip(","))
        assert norm(line) == norm(want), (line, want)
    restart = tmp_path / "restart.dat"
    n = 32 * 32 * 17
    assert restart.stat().st_size == 18 + 4 + 16 + 12 + 24 + 7 * 8 * n + 32
