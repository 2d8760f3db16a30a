"""The HIP step's algebra (emulated on the CPU with the library's operator matrices) against
the oracle: pins the spectral-resident formulation and the eigenbasis solves."""
import os

import numpy as np
import pytest

from oracle import oracle as orc
from tests.gpu_util import CHANNEL32, field_errors, max_err

import f90_dns_amd

pytestmark = pytest.mark.skipif(not os.path.exists(f90_dns_amd.library_path()), reason="library not built")


def _run(steps, **kw):
    from tests.kernel_emulation import Emulator
    o = orc.Oracle(**kw)
    o.init_fields()
    e = Emulator(**kw)
    e.set_fields(*o.get_fields())
    for _ in range(steps):
        o.step(1)
        e.step()
    return o, e


def test_emulated_step_matches_oracle_channel32():
    o, e = _run(10, **CHANNEL32)
    errs = field_errors(e.get_fields(), o.get_fields())
    assert max_err(errs) < 1e-12, errs


def test_emulated_step_pi_control_and_walls():
    kw = dict(CHANNEL32, nx=16, ny=16, nz=17, flow_control_method=2, perturbation_kind=1, seed=3, noise_kmax=3,
              perturbation_amplitude=0.05, use_dealias_23_rule=0)
    o, e = _run(8, **kw)
    errs = field_errors(e.get_fields(), o.get_fields())
    assert max_err(errs) < 1e-12, errs
    assert abs(o.diagnostics()["pressure_gradient"] - e.G) < 1e-13


def test_emulated_step_exact_d1_nz65():
    kw = dict(CHANNEL32, nx=16, ny=16, nz=65, lgl_derivative=1, flow_control_method=2, perturbation_kind=1, seed=9,
              noise_kmax=3, perturbation_amplitude=0.05)
    o, e = _run(4, **kw)
    errs = field_errors(e.get_fields()[:3], o.get_fields()[:3])
    assert max_err(errs) < 1e-10, errs


@pytest.mark.parametrize("nz,exact", [(18, 0), (48, 1)])
def test_emulated_step_even_nz(nz, exact):
    """Even nz: the LGL nodes have no centre node, so the folds have no centre row and the
    Helmholtz panel pairs (v row c, u row c+1) cover every node pair."""
    kw = dict(CHANNEL32, nx=16, ny=12, nz=nz, lgl_derivative=exact, flow_control_method=2, perturbation_kind=1,
              seed=5, noise_kmax=3, perturbation_amplitude=0.05)
    o, e = _run(4, **kw)
    errs = field_errors(e.get_fields(), o.get_fields())
    assert max_err(errs) < 1e-12, errs
