"""Extract the reference's shipped 15-step diagnostic traces into JSON fixtures.

Reads /root/reference/standard_output.txt and weak_form_output.txt (the stdout of the
reference's 3D solver, 32x32x17, Re=500, dt=0.002, Lx=2pi, Ly=pi, u_centerline=1.5, constant
G=0.006, analytical perturbation mx=my=1, amplitude 0.01, with the standard and the weak-form
Poisson RHS; run settings printed at standard_output.txt:7-22 and weak_form_output.txt:5-22)
and writes the per-step rows (:88-104 of each) as data.  Run once in a container that has the
reference; the fixtures are committed.
"""
import json
import os
import re
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
row = re.compile(r"Step\s+(\d+), t=\s*([-\d.E+]+), \|u\|max=\s*([-\d.E+]+), \|v\|max=\s*([-\d.E+]+), "
                 r"\|w\|max=\s*([-\d.E+]+), E=\s*([-\d.E+]+), dP/dx=\s*([-\d.E+]+)")
for src, out, poisson in (("standard_output.txt", "standard_output_trace.json", "standard"),
                          ("weak_form_output.txt", "weak_form_output_trace.json", "weak form")):
    rows = []
    with open(os.path.join(REF, src), encoding="utf-8") as f:
        for line in f:
            m = row.search(line)
            if m:
                rows.append({"step": int(m.group(1)), "t": float(m.group(2)), "umax": m.group(3),
                             "vmax": m.group(4), "wmax": m.group(5), "E": m.group(6), "dPdx": m.group(7)})
    fixture = {
        "source": f"{src}:88-104 (reference stdout, values as printed, ES12.4)",
        "run": {"nx": 32, "ny": 32, "nz": 17, "re": 500.0, "dt": 0.002, "xlen": "2*pi", "ylen": "pi",
                "u_centerline": 1.5, "flow_control_method": 1, "target_pressure_gradient": 0.006,
                "perturbation": "analytical mx=1 my=1 amplitude 0.01", "poisson": poisson},
        "initial_perturbation_max_before_scaling": "2.0000E+00",
        "rows": rows,
    }
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), out)
    with open(path, "w") as f:
        json.dump(fixture, f, indent=1)
    print(f"wrote {len(rows)} rows to {path}")
