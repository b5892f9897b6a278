"""The allocation-free entry point cemc_batch_get_cf_noalloc against
cemc_batch_get_cf, through the C ABI: the same correlation functions and
energy, bit for bit (both run the same kernels in the same order), on two
calls in a row with the second on a smaller lattice (so the scratch array
still holds the first call's larger contents), and a clear error for a
request beyond the scratch array's size."""
import ctypes

import numpy as np
import pytest

from cemc_amd import _capi
from cemc_amd.batch import ChainBatch
from cemc_amd.lattice import FCCSettings, random_occupation, seeded_ecis
from cemc_amd.model import CEModel

pytestmark = pytest.mark.gpu


def _batch(L, species, shells, kind, seed):
    s = FCCSettings(L, species, shells)
    model = CEModel(s, seeded_ecis(s, seed, 0.02))
    occ = random_occupation(s.n_sites, len(species), seed)
    chain = dict(kind=kind, temperature=900.0, occupation=occ, seed=seed, uid=0)
    return ChainBatch([model], [chain])


def _both(batch, chain=0):
    lib = _capi.load()
    n = batch.models[chain].n_cf
    want, got = np.zeros(n), np.full(n, np.nan)
    ew, eg = ctypes.c_double(0.0), ctypes.c_double(np.nan)
    _capi.check(lib.cemc_batch_get_cf(batch.handle, chain, want.ctypes.data_as(_capi._dp), ctypes.byref(ew)))
    _capi.check(lib.cemc_batch_get_cf_noalloc(batch.handle, chain, got.ctypes.data_as(_capi._dp),
                                              ctypes.byref(eg)))
    return want, ew.value, got, eg.value


def test_noalloc_cf_equals_allocating_entry_point(gpu):
    big = _batch(12, ["A", "B", "C"], {2: 4, 3: 3, 4: 2}, 0, 11)
    small = _batch(6, ["A", "B"], {2: 4, 3: 2}, 1, 12)
    for b in (big, small):
        b.run(3000)
    print("no-alloc: ternary FCC 12^3 (%d CFs), then binary FCC 6^3 (%d CFs)"
          % (big.models[0].n_cf, small.models[0].n_cf), flush=True)
    for b in (big, small):
        want, ew, got, eg = _both(b)
        assert np.array_equal(want, got)
        assert ew == eg
        # the energy the batch tracks move by move agrees to rounding
        assert abs(eg - b.energies()[0]) <= 1e-10 * b.models[0].n_sites
    # energy only (cf NULL)
    lib = _capi.load()
    e = ctypes.c_double(np.nan)
    _capi.check(lib.cemc_batch_get_cf_noalloc(small.handle, 0, None, ctypes.byref(e)))
    assert e.value == _both(small)[1]
    big.close()
    small.close()


def test_noalloc_rejects_requests_beyond_its_scratch(gpu):
    # quaternary 36^3: 855 CFs x 183 blocks of partial sums exceed 1 MiB
    b = _batch(36, ["A", "B", "C", "D"], {2: 3, 3: 3, 4: 3}, 1, 13)
    lib = _capi.load()
    cf = np.zeros(b.models[0].n_cf)
    e = ctypes.c_double(0.0)
    rc = lib.cemc_batch_get_cf_noalloc(b.handle, 0, cf.ctypes.data_as(_capi._dp), ctypes.byref(e))
    assert rc == _capi.CEMC_EINVAL
    assert "scratch" in lib.cemc_last_error().decode()
    # the allocating entry point still serves it
    _capi.check(lib.cemc_batch_get_cf(b.handle, 0, cf.ctypes.data_as(_capi._dp), ctypes.byref(e)))
    assert abs(e.value - b.energies()[0]) <= 1e-10 * b.models[0].n_sites
    b.close()
