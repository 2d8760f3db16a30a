"""Why the mixed path keeps FP64 arithmetic in the D1 contraction (FP32 output).

Emulated on the CPU with the oracle's D1 (numpy float32 matrix products accumulate in FP32, as
an FP32 MFMA chain does): on channel-like columns -- the laminar profile with smooth
perturbations -- an FP32 contraction errs by 0.4-1.3e-6 of the result's norm and a split
(hi + lo) FP32 x 3 contraction still by 0.2-0.5e-6, because the FP32 accumulation of D1's large
entries (up to N(N+1)/4 at the ends) dominates; FP64 arithmetic rounded once to FP32 errs by
2.5e-8.  The mixed path's product u x omega is formed in FP32 anyway (rounding ~6e-8 of its
terms), so an FP32 D1 would raise the vorticity error 20-50 fold above that level.

The split-BF16 x 3 contraction (D = Dh + Dl and x = xh + xl in bfloat16, Dh xh + Dh xl + Dl xh
summed in FP32, as bf16 MFMAs with an FP32 accumulator do) keeps about 16 bits of each operand:
it errs by ~1e-4 of the result's norm, 20-100 times the FP32 contraction, so it does not hold
the mixed path's error and is not built."""
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


def _bf16(x):
    """Round float32 values to bfloat16 (round to nearest even), returned as float32."""
    u = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    return ((u + ((u >> 16) & 1) + 0x7FFF) & 0xFFFF0000).astype(np.uint32).view(np.float32)


def test_bf16_rounding():
    x = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -9, -3.0e-3, 65504.0], dtype=np.float32)
    got = _bf16(x)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: to even (1); 1 + 3 * 2^-9 rounds up to 1 + 2^-7
    assert got[0] == 1.0 and got[1] == 1.0 and got[2] == np.float32(1.0 + 2.0 ** -7)
    assert abs(got[3] / x[3] - 1) < 2.0 ** -8 and abs(got[4] / x[4] - 1) < 2.0 ** -8
    assert np.all(_bf16(got) == got)


@pytest.mark.parametrize("nz,exact", [(33, 0), (65, 1), (129, 1)])
def test_split_bf16x3_d1_contraction_error(nz, exact):
    D = orc.diff_matrix_exact(nz) if exact else orc.diff_matrix(nz)
    F = _columns(nz)
    ref = D @ F
    rel = lambda g: np.linalg.norm(g - ref) / np.linalg.norm(ref)
    D32, F32 = D.astype(np.float32), F.astype(np.float32)
    fp32 = rel((D32 @ F32).astype(np.float64))
    Dh, Fh = _bf16(D32), _bf16(F32)
    Dl, Fl = _bf16(D32 - Dh), _bf16(F32 - Fh)
    # the bf16 products are exact in FP32; the three partial products accumulate in FP32
    bf3 = rel(((Dh @ Fh).astype(np.float32) + (Dh @ Fl).astype(np.float32) + (Dl @ Fh).astype(np.float32))
              .astype(np.float64))
    assert 1e-5 < bf3 < 1e-3
    assert bf3 > 20 * fp32
