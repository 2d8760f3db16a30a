"""Compact oracle fixtures for runs too long to repeat inside a GPU test.

A fixture holds what the FP64 oracle produced after a number of steps of one run, in a form
small enough to commit: for each field (u, v, w, p) its exact L2 norm, its L2 norm per z-plane,
one full z-plane, and a random sketch S f of the whole field.  S is a k x N Gaussian matrix
(entries N(0, 1) drawn in float32 from numpy's PCG64 with a fixed seed, in fixed column chunks),
so for any field difference d, |S d|^2 / k is an unbiased estimate of |d|^2 whose ratio to
the truth follows chi^2_k / k: with k = 128 the estimate of |d| lies within [0.78, 1.24] of
|d| with probability 0.999.  The sketch of a GPU field is taken with the same S, so
|S a - S b| / (sqrt(k) |b|) estimates the relative L2 error |a - b| / |b| of the whole field.

Fixtures also keep the diagnostics trace and, where the generator ran it, the same estimate
between the oracle and its twin with reordered plane transforms (oracle fft_variant = 1):
the rounding floor of each field for that run.
"""
import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
SKETCH_ROWS = 128
SKETCH_SEED = 20261204
SKETCH_CHUNK = 1 << 16
FIELDS = ("u", "v", "w", "p")


def sketch(fields, k=SKETCH_ROWS, seed=SKETCH_SEED, chunk=SKETCH_CHUNK):
    """S f for each field of the list (all of one size N): array (len(fields), k)."""
    F = np.stack([np.ascontiguousarray(f, dtype=np.float64).ravel() for f in fields], axis=1)
    N = F.shape[0]
    rng = np.random.Generator(np.random.PCG64(seed))
    out = np.zeros((k, F.shape[1]))
    for c0 in range(0, N, chunk):
        c1 = min(N, c0 + chunk)
        S = rng.standard_normal((k, c1 - c0), dtype=np.float32).astype(np.float64)
        out += S @ F[c0:c1]
    return out.T.copy()


def summarize(fields, plane):
    """The fixture record of (u, v, w, p), each of shape (nz, ny, nx)."""
    rec = {}
    sk = sketch(fields)
    for name, f, s in zip(FIELDS, fields, sk):
        rec[f"{name}_norm"] = np.array(np.linalg.norm(f))
        rec[f"{name}_planes"] = np.linalg.norm(f.reshape(f.shape[0], -1), axis=1)
        rec[f"{name}_sketch"] = s
        rec[f"{name}_plane"] = np.ascontiguousarray(f[plane])
    return rec


def path(name):
    return os.path.join(GOLDEN, f"oracle_{name}.npz")


def save(name, meta, snapshots):
    """snapshots: {step: record}; meta: JSON-able dict (params, trace, floors)."""
    arrays = {"meta": np.array(json.dumps(meta))}
    for step, rec in snapshots.items():
        for key, val in rec.items():
            arrays[f"s{step}_{key}"] = val
    np.savez_compressed(path(name), **arrays)


class Fixture:
    def __init__(self, name):
        z = np.load(path(name), allow_pickle=False)
        self.meta = json.loads(str(z["meta"]))
        self.params = self.meta["params"]
        self._z = z

    def steps(self):
        return sorted({int(k.split("_", 1)[0][1:]) for k in self._z.files if k.startswith("s")})

    def record(self, step):
        pre = f"s{step}_"
        return {k[len(pre):]: self._z[k] for k in self._z.files if k.startswith(pre)}

    def floor(self, step):
        """Estimated relative error between the oracle and its reordered twin (or None)."""
        f = self.meta.get("floor", {}).get(str(step))
        return f

    def trace(self):
        return self.meta.get("trace", [])

    def compare(self, fields, step):
        """Estimated errors of GPU fields (u, v, w, p) against the oracle after `step` steps:
        relative L2 for u, v, p and w ('w_self'), w's error on the velocity scale ('w'), the
        exact relative L2 over the stored plane ('<f>_plane'), and the worst per-plane norm
        mismatch ('<f>_planes')."""
        rec = self.record(step)
        k = rec["u_sketch"].shape[0]
        sk = sketch(fields, k=k)
        vel = np.sqrt(sum(float(rec[f"{c}_norm"]) ** 2 for c in "uvw"))
        plane = self.meta["plane"]
        out = {}
        for i, name in enumerate(FIELDS):
            d = np.linalg.norm(sk[i] - rec[f"{name}_sketch"]) / np.sqrt(k)
            nb = float(rec[f"{name}_norm"])
            rel = d / nb if nb > 0 else d
            if name == "w":
                out["w_self"] = rel
                out["w"] = d / vel
            else:
                out[name] = rel
            ref_plane = rec[f"{name}_plane"]
            got_plane = np.asarray(fields[i])[plane]
            npl = np.linalg.norm(ref_plane)
            out[f"{name}_plane"] = np.linalg.norm(got_plane - ref_plane) / (npl if npl > 0 else 1.0)
            pn = np.linalg.norm(np.asarray(fields[i]).reshape(len(rec[f"{name}_planes"]), -1), axis=1)
            scale = max(float(np.max(rec[f"{name}_planes"])), 1e-300)
            out[f"{name}_planes"] = float(np.max(np.abs(pn - rec[f"{name}_planes"]))) / scale
        return out


def oracle_params(p):
    """The oracle's keyword arguments for a fixture's parameter dict (drops GPU-only keys)."""
    return {k: v for k, v in p.items() if k not in ("precision",)}
