"""The rocprofv3 summaries behind profiles/ and bench.py's roofline entry, on small synthetic
result databases (the two rocpd tables the summary reads): calibrated HBM bytes from
FETCH_SIZE/WRITE_SIZE and the MFMA busy fraction."""
import contextlib
import io
import sqlite3
import sys

import pytest

from f90_dns_amd import rocprof_summary as rs
from f90_dns_amd.profile_step import COPY_DOUBLES

HELM = "void dns::k_helmholtz<1, false, 4>(dns::HelmArgs)"
COPY = "void dns::k_copy(double const*, double*, unsigned long)"


def _trace(path, rows):
    c = sqlite3.connect(path)
    c.execute("create table kernels (name text, duration integer)")
    c.executemany("insert into kernels values (?, ?)", rows)
    c.commit()
    c.close()


def _counters(path, rows):
    c = sqlite3.connect(path)
    c.execute("create table counters_collection (kernel_name text, counter_name text, value real)")
    c.executemany("insert into counters_collection values (?, ?, ?)", rows)
    c.commit()
    c.close()


def _run(args):
    out = io.StringIO()
    old = sys.argv
    sys.argv = ["rocprof_summary"] + args
    try:
        with contextlib.redirect_stdout(out):
            rs.main()
    finally:
        sys.argv = old
    return out.getvalue()


def test_traffic_is_calibrated_on_the_copy_kernel(tmp_path):
    copy = COPY_DOUBLES * 8   # bytes each way
    t, f, w = (str(tmp_path / n) for n in ("t.db", "f.db", "w.db"))
    _trace(t, [(HELM, 80000), (HELM, 82000), (COPY, 200000)])
    # the copy reads half its bytes on FETCH_SIZE and all of them on WRITE_SIZE (KiB)
    _counters(f, [(COPY, "FETCH_SIZE", copy / 2 / 1024), (HELM, "FETCH_SIZE", 50e3), (HELM, "FETCH_SIZE", 54e3)])
    _counters(w, [(COPY, "WRITE_SIZE", copy / 1024), (HELM, "WRITE_SIZE", 40e3)])
    text = _run(["--trace", t, "--fetch", f, "--write", w])
    assert "FETCH_SIZE x 2.000, WRITE_SIZE x 1.000" in text
    line = [l for l in text.splitlines() if l.startswith("traffic helmholtz")]
    assert len(line) == 1
    want = 52e3 * 1024 * 2.0 + 40e3 * 1024
    assert float(line[0].split()[2]) == pytest.approx(want, rel=1e-9)
    assert "| 81.0 | 80.0 | 82.0 |" in text   # mean, min, max in microseconds


def test_mfma_busy_fraction(tmp_path):
    t, m = str(tmp_path / "t.db"), str(tmp_path / "m.db")
    _trace(t, [(HELM, 80000)])   # 80 us
    busy = 0.41 * 1024 * 2.4e9 * 80e-6
    _counters(m, [(HELM, "SQ_VALU_MFMA_BUSY_CYCLES", busy), (HELM, "SQ_BUSY_CYCLES", 1.0)])
    text = _run(["--trace", t, "--mfma", m])
    row = [l for l in text.splitlines() if l.startswith(f"| `{rs.short(HELM)}`") and l.count("|") == 5]
    assert row and row[-1].rstrip().endswith("| 0.41 |"), row


def test_bench_reads_the_committed_helmholtz_traffic():
    """bench.py's roofline entry takes the dominant kernel's HBM bytes from the committed
    FP64 step profile; the figure must be there and lie between the kernel's algorithmic bytes
    (9 spectral fields of 17.2 MB at 128 x 128 x 129) and twice that."""
    import bench
    t = bench.profiled_traffic("step_128x128x129_fp64_kernels.md", "helmholtz")
    alg = 9 * 65 * 128 * 129 * 16
    assert t is not None and alg <= t < 2 * alg, (t, alg)
