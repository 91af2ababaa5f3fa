"""The drop-in's Model with its lattice kept for the GPU (CPU): every lattice
golden case and the fan2018cpc examples -- vacancies, Anderson and charged
disorder included -- give the same site count, potential and random vectors
as the host-expanded Model, and expanding the kept description on the host
gives its neighbour arrays bit for bit.  General-model cases keep no
description."""
import os

import numpy as np
import pytest

from conftest import CASES_DIR, GOLDEN, golden_cases
from gpuqt_amd import hostlib


def _quiet(fn):
    devnull = os.open(os.devnull, os.O_WRONLY)
    saved = os.dup(1)
    os.dup2(devnull, 1)
    try:
        return fn()
    finally:
        os.dup2(saved, 1)
        os.close(devnull)
        os.close(saved)


CASE_DIRS = [os.path.join(CASES_DIR, c) for c in golden_cases()] + \
    [os.path.join(GOLDEN, "fan2018cpc", e) for e in ("diffusive", "localized")]


@pytest.mark.parametrize("d", CASE_DIRS, ids=os.path.basename)
def test_deferred_lattice_equals_host_model(d):
    lattice_model = os.path.exists(os.path.join(d, "lattice.in"))
    a = _quiet(lambda: hostlib.HostModel(d))
    b = _quiet(lambda: hostlib.HostModel(d, lattice_on_device=True))
    assert b.n == a.n and b.max_neighbor == a.max_neighbor
    got = b.lattice()
    if not lattice_model:
        assert got is None
        return
    lat, site_index = got
    n_pristine = int(np.prod(lat["cells"])) * lat["orbitals"]
    if site_index is None:
        assert n_pristine == a.n
    else:
        assert site_index.shape == (n_pristine,)
        kept = site_index[site_index >= 0]
        np.testing.assert_array_equal(kept, np.arange(a.n))
    ref = a.arrays()
    np.testing.assert_array_equal(b.potential(), ref["potential"])
    for _ in range(2):
        ra, rb = a.random_state(), b.random_state()
        np.testing.assert_array_equal(ra[0], rb[0])
        np.testing.assert_array_equal(ra[1], rb[1])
    with pytest.raises(hostlib.InputError):
        b.arrays()
    b.expand_lattice()
    assert b.lattice() is None
    got_arrays = b.arrays()
    for k, v in ref.items():
        np.testing.assert_array_equal(got_arrays[k].view(np.uint8), v.view(np.uint8), err_msg=k)
