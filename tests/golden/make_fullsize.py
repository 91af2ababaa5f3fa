"""Writes the full-size fixtures of the benchmarked workloads (test data).

Runs the oracle -- the C restatement of the reference CPU build, pinned bit
for bit against the reference on the small golden cases -- on the inputs of
the benchmarked workloads at their real sizes, with all host threads, and
stores what the full-size GPU tests compare against
(tests/test_gpu_fullsize.py):

  config2.npz  8.0M-site honeycomb (2000 x 1000 cells, Anderson U[-1, 1]),
               E_max 10, M = 3000: the DOS moment set and row of the first
               random vector; the state after evolve(+1, 20) and after
               evolvex(+1, 20) of that vector (sha256 of its FP64 bytes plus
               sampled values); find_msd over two steps of dt = 2 (the
               moment sets and rows).
  cube.npz     160^3 Peierls cube, E_max 15, M = 2048: the DOS moment sets
               and rows of vectors 0 and 1 of the seeded phases.
  batch8.npz   realisation 8 of the uneven batch (106,128 sites, charged
               impurities), M = 1500, 16 steps of dt = 2: find_vac and
               find_msd rows of random vectors 0 and 1.

With --reference (where the reference program was built into oracle/_ref
by `make -C oracle ref`), the reference program also runs the DOS of the
8.0M-site honeycomb at M = 3000, and its dos.out must equal the oracle's
row printed the same way (6 significant digits): a full-size pin of the
oracle itself.  That run takes about six minutes on one core.

Usage: python tests/golden/make_fullsize.py [--reference] [--only NAME]
"""
from __future__ import annotations

import argparse
import hashlib
import os
import sys
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from gpuqt_amd import inputs as gi  # noqa: E402
from gpuqt_amd import workloads as wl  # noqa: E402

OUT = os.path.join(HERE, "fullsize")
SAMPLE_COUNT = 64


def state_digest(re, im):
    """sha256 of the FP64 bytes of (re, im), and sampled entries."""
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(re, dtype="<f8").tobytes())
    h.update(np.ascontiguousarray(im, dtype="<f8").tobytes())
    n = len(re)
    idx = np.unique(np.concatenate([[0, 1, n // 2, n - 1],
                                    np.random.default_rng(99).integers(0, n, SAMPLE_COUNT)]))
    return h.hexdigest(), idx.astype(np.int64), re[idx].copy(), im[idx].copy()


def log(msg):
    print(f"[fullsize] {msg}", flush=True)


def config2(tmp, with_reference):
    d = os.path.join(tmp, "config2")
    os.makedirs(d)
    wl.write_config2(d, time_steps=2)
    t0 = time.perf_counter()
    m = oracle.OracleModel(d)
    log(f"config2: {m.n} sites built in {time.perf_counter() - t0:.1f} s")
    r = m.random_state()
    t0 = time.perf_counter()
    dos_rows, dos_mu = m.find_dos(r)
    log(f"config2: DOS M={m.number_of_moments} in {time.perf_counter() - t0:.1f} s")
    t = float(m.time_step[0] * m.energy_max)
    ev = state_digest(*m.evolve(1, t, *r))
    evx = state_digest(*m.evolvex(1, t, *r))
    t0 = time.perf_counter()
    msd_rows, msd_mu = m.find_msd(r)
    log(f"config2: MSD over {len(m.time_step)} steps in {time.perf_counter() - t0:.1f} s")
    out = dict(n=m.n, moments=m.number_of_moments, t_scaled=t, energies=m.energy,
               dos_mu=dos_mu[0], dos_row=dos_rows[0], msd_mu=msd_mu, msd_rows=msd_rows,
               evolve_sha=np.array(ev[0]), evolve_idx=ev[1], evolve_re=ev[2], evolve_im=ev[3],
               evolvex_sha=np.array(evx[0]), evolvex_idx=evx[1], evolvex_re=evx[2],
               evolvex_im=evx[3])
    if with_reference:
        check_reference_dos(tmp, dos_rows[0])
    return out


def check_reference_dos(tmp, row):
    """The reference program's dos.out of the same DOS job against the
    oracle's row at the reference's output precision."""
    ref = oracle.reference_binary()
    if ref is None:
        raise SystemExit("--reference: the reference program is not built (make -C oracle ref)")
    d = os.path.join(tmp, "config2_ref")
    os.makedirs(d)
    wl.write_config2(d, time_steps=2, msd=False)
    drv = os.path.join(tmp, "ref_input.txt")
    with open(drv, "w") as f:
        f.write(d + "\n")
    t0 = time.perf_counter()
    r = oracle.run_reference(drv)
    if r.returncode != 0:
        raise SystemExit("reference run failed:\n" + r.stdout[-2000:] + r.stderr[-2000:])
    log(f"config2: reference program DOS in {time.perf_counter() - t0:.1f} s")
    with open(os.path.join(d, "dos.out")) as f:
        got = f.read().split()
    want = [f"{v:g}" for v in row]
    if got != want:
        bad = [i for i, (a, b) in enumerate(zip(got, want)) if a != b]
        raise SystemExit(f"reference dos.out differs from the oracle's row at {len(bad)} of "
                         f"{len(want)} energies (first {bad[:5]})")
    log("config2: reference dos.out equals the oracle's DOS row at 6 digits")


def cube(tmp):
    arrays, n = gi.cubic_peierls_arrays(wl.CUBE_L)
    oh = oracle.OracleHam(arrays, n, 6, wl.CUBE_EMAX)
    re, im = wl.cube_phases(2, n)
    M = 2048
    mus, rows = [], []
    vol = float(wl.CUBE_L ** 3)
    for r in range(2):
        t0 = time.perf_counter()
        mr, mi = oh.moments((re[r], im[r]), (re[r], im[r]), M)
        log(f"cube: vector {r}, M={M} in {time.perf_counter() - t0:.1f} s")
        mu = mr + 1j * mi
        mus.append(mu)
        dr, di = oracle.damping(mr, mi)
        rows.append(oracle.summation(dr, wl.CUBE_ENERGIES, wl.CUBE_EMAX, vol))
    return dict(n=n, moments=M, energies=wl.CUBE_ENERGIES, mu=np.array(mus),
                rows=np.array(rows))


def batch8(tmp):
    k = 8
    nx, ny = wl.batch_cells(64)[k]
    d = os.path.join(tmp, f"batch{k}")
    wl.write_batch_realisation(d, nx, ny)
    m = oracle.OracleModel(d)
    log(f"batch realisation {k}: {m.n} sites")
    vecs = [m.random_state() for _ in range(2)]
    vac, msd = [], []
    for r, v in enumerate(vecs):
        t0 = time.perf_counter()
        vac.append(m.find_vac(v)[0])
        msd.append(m.find_msd(v)[0])
        log(f"batch: vector {r} VAC + MSD in {time.perf_counter() - t0:.1f} s")
    return dict(realisation=k, cells=np.array([nx, ny]), n=m.n, moments=m.number_of_moments,
                energies=m.energy, time_step=m.time_step, vac_rows=np.array(vac),
                msd_rows=np.array(msd))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--reference", action="store_true")
    ap.add_argument("--only", choices=("config2", "cube", "batch8"))
    args = ap.parse_args()
    os.makedirs(OUT, exist_ok=True)
    log(f"oracle threads: {oracle.num_threads()}")
    with tempfile.TemporaryDirectory() as tmp:
        for name, fn in (("config2", lambda: config2(tmp, args.reference)),
                         ("cube", lambda: cube(tmp)), ("batch8", lambda: batch8(tmp))):
            if args.only and args.only != name:
                continue
            data = fn()
            np.savez_compressed(os.path.join(OUT, f"{name}.npz"), **data)
            log(f"wrote {name}.npz")


if __name__ == "__main__":
    main()
