"""Why the mixed path keeps FP64 arithmetic in the D1 contraction (FP32 output).

Emulated on the CPU with the oracle's D1 (numpy float32 matrix products accumulate in FP32, as
an FP32 MFMA chain does): on channel-like columns -- the laminar profile with smooth
perturbations -- an FP32 contraction errs by 0.4-1.3e-6 of the result's norm and a split
(hi + lo) FP32 x 3 contraction still by 0.2-0.5e-6, because the FP32 accumulation of D1's large
entries (up to N(N+1)/4 at the ends) dominates; FP64 arithmetic rounded once to FP32 errs by
2.5e-8.  The mixed path's product u x omega is formed in FP32 anyway (rounding ~6e-8 of its
terms), so an FP32 D1 would raise the vorticity error 20-50 fold above that level."""
import numpy as np
import pytest

from oracle import oracle as orc


def _columns(nz, seed=1, ncol=64):
    z, _ = orc.lgl(nz)
    rng = np.random.default_rng(seed)
    return np.stack([(1 - z ** 2) * (1 + 0.05 * np.cos(k * z + rng.uniform(0, 6))) * rng.standard_normal()
                     for k in range(ncol)], axis=1)


@pytest.mark.parametrize("nz,exact", [(33, 0), (65, 1), (129, 1)])
def test_fp32_d1_contraction_error_against_fp64_rounded_once(nz, exact):
    D = orc.diff_matrix_exact(nz) if exact else orc.diff_matrix(nz)
    F = _columns(nz)
    ref = D @ F
    rel = lambda g: np.linalg.norm(g - ref) / np.linalg.norm(ref)
    fp32 = rel((D.astype(np.float32) @ F.astype(np.float32)).astype(np.float64))
    Dh = D.astype(np.float32); Dl = (D - Dh).astype(np.float32)
    Fh = F.astype(np.float32); Fl = (F - Fh).astype(np.float32)
    split3 = rel((Dh @ Fh + Dh @ Fl + Dl @ Fh).astype(np.float64))
    once = rel((D @ F).astype(np.float32).astype(np.float64))
    assert once < 5e-8
    assert fp32 > 10 * once and split3 > 5 * once
    assert 1e-7 < fp32 < 5e-6
