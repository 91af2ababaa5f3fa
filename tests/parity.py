"""Shared comparison helpers for the parity tests.

Tolerances (FP64 path):
  - state vectors: bit-identical (the kernels repeat the reference's
    per-site arithmetic in the same order, without contraction);
  - moments: |mu - mu_ref| <= 1e-12 * sum_n |L_n||psi_n| bound, taken as
    1e-12 * |mu_0| scaled, plus 1e-10 * |mu_ref| -- the sums differ only by
    their reduction order;
  - rows (DOS, VAC, MSD, ...): normwise ||a - b||_inf <= 1e-10 ||b||_inf and,
    element-wise, |a - b| <= 1e-10 |b| wherever |b| > 1e-6 ||b||_inf.
"""
import numpy as np

ROW_RTOL = 1e-10


def assert_bits(a, b, what=""):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    assert a.shape == b.shape, what
    same = (a.view(np.uint64) == b.view(np.uint64))
    if not same.all():
        i = int(np.argmin(same))
        raise AssertionError(f"{what}: {int((~same).sum())} of {a.size} differ; first at {i}: "
                             f"{a[i]!r} vs {b[i]!r}")


def assert_moments(mu, ref, what="", rtol=1e-10, scale_tol=1e-12):
    mu = np.asarray(mu)
    ref = np.asarray(ref)
    scale = max(abs(ref[0]), np.abs(ref).max())
    err = np.abs(mu - ref)
    bound = rtol * np.abs(ref) + scale_tol * scale
    if not (err <= bound).all():
        i = int(np.argmax(err - bound))
        raise AssertionError(f"{what}: moment {i}: {mu[i]!r} vs {ref[i]!r} (err {err[i]:.3e}, "
                             f"bound {bound[i]:.3e})")


def assert_rows(a, b, what="", rtol=ROW_RTOL, floor=1e-6):
    a = np.atleast_2d(np.asarray(a, dtype=np.float64))
    b = np.atleast_2d(np.asarray(b, dtype=np.float64))
    assert a.shape == b.shape, f"{what}: shape {a.shape} vs {b.shape}"
    for r in range(a.shape[0]):
        bmax = np.abs(b[r]).max()
        if bmax == 0:
            assert np.abs(a[r]).max() == 0, what
            continue
        normwise = np.abs(a[r] - b[r]).max() / bmax
        assert normwise <= rtol, f"{what} row {r}: normwise error {normwise:.3e}"
        mask = np.abs(b[r]) > floor * bmax
        elem = (np.abs(a[r] - b[r])[mask] / np.abs(b[r])[mask]).max() if mask.any() else 0.0
        assert elem <= rtol, f"{what} row {r}: element-wise error {elem:.3e}"


def read_out(path):
    """Parses a .out file (rows of space-separated values)."""
    rows = []
    with open(path) as f:
        for line in f:
            if line.strip():
                rows.append([float(x) for x in line.split()])
    return np.array(rows)
