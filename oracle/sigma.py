"""Conductivity post-processing restated in numpy (test infrastructure only,
like the rest of oracle/).

Follows the reference's example scripts line by line:
examples/fan2018cpc/lattice/diffusive/plot_results.m:16-44 (averages over the
random vectors, sigma from VAC by cumtrapz, sigma from MSD by differences)
and examples/review/square/diffusive/plot_results.m:33-44 (the same, and the
"true" VAC / MSD as the averages divided by the DOS).  The integral uses
scipy's cumulative trapezoid (MATLAB's cumtrapz), an implementation
independent of the product's loop.
"""
from __future__ import annotations

import numpy as np
from scipy.integrate import cumulative_trapezoid

PI = 3.141592653589793


def conductivity(time_step, dos=None, vac=None, msd=None):
    """time_step [Nt]; dos [Ns, Ne]; vac / msd [Ns*Nt, Ne].  Returns a dict
    with t_vac, t_msd and the [Nt, Ne] arrays sigma_vac, sigma_msd, vac, msd
    that the inputs allow."""
    ts = np.asarray(time_step, dtype=float)
    Nt = len(ts)
    out = {"t_vac": np.concatenate([[0.0], np.cumsum(ts[:-1])]),   # plot_results.m:30
           "t_msd": np.cumsum(ts) - ts / 2}                          # :34 (uniform steps)
    dos_ave = None if dos is None else np.atleast_2d(dos).mean(axis=0)   # :17
    if vac is not None:
        v = np.asarray(vac, dtype=float)
        ave = v.reshape(-1, Nt, v.shape[1]).mean(axis=0)              # :18-27
        out["sigma_vac"] = 2 * PI * cumulative_trapezoid(ave, out["t_vac"], axis=0,
                                                          initial=0)  # :31
        if dos_ave is not None:
            out["vac"] = ave / dos_ave                                 # :42-45
    if msd is not None:
        m = np.asarray(msd, dtype=float)
        ave = m.reshape(-1, Nt, m.shape[1]).mean(axis=0)
        prev = np.vstack([np.zeros((1, ave.shape[1])), ave[:-1]])
        out["sigma_msd"] = PI * (ave - prev) / ts[:, None]             # :35-38
        if dos_ave is not None:
            out["msd"] = ave / dos_ave
    return out
