"""GPU counterpart of the reference's updater binding ``PyCEUpdater``
(cemc/cpp_ext/pyce_updater.pyx:5-60 over class CEUpdater,
cpp/include/ce_updater.hpp:47-213).

Same constructor and methods, same exceptions (ValueError for
std::invalid_argument, RuntimeError otherwise).  The correlation-function
arithmetic, history slots and occupation live on the GPU; this class converts
symbols to species ids at the boundary and keeps the caller's atoms object in
step with the lattice, as the reference does on every flip and undo
(cpp/src/ce_updater.cpp:338-347, 435-446).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _capi
from .model import CEModel


_MAX_UNDO = 999


class PyCEUpdater:
    def __init__(self, atoms, bc, corr_func, eci, model=None):
        lib = _capi.load()
        if bc is None:
            raise ValueError("CEBulk object is nullptr!")
        symbols = [a.symbol for a in atoms]
        unknown = sorted(set(symbols) - set(bc.unique_elements))
        if unknown:
            raise ValueError("symbols %s are not among the settings' unique_elements" % unknown)
        self.model = model if model is not None else CEModel(bc, eci)
        if sorted(eci.keys()) != self.model.names:
            raise ValueError("All ECIs does not correspond to a correlation function!")
        self.atoms = atoms
        self.bc = bc
        self.corr_func = corr_func
        self.eci = dict(eci)
        self._num_threads = 1
        self._vib = None
        occ = self.model.species_ids(symbols)
        eci_arr = np.array([float(eci[n]) for n in self.model.names])
        o, op = _capi.i8(occ)
        e, ep = _capi.f64(eci_arr)
        cp = None
        if corr_func is not None:
            # CFHistoryTracker::insert keeps only names that have an ECI and
            # leaves the others at zero (cpp/src/cf_history_tracker.cpp:82-95)
            cf0 = np.array([float(corr_func.get(n, 0.0)) for n in self.model.names])
            c, cp = _capi.f64(cf0)
        # corr_func None: the correlation functions are computed on the GPU
        h = ctypes.c_void_p()
        _capi.check(lib.cemc_updater_create(self.model.handle, op, cp, ep, ctypes.byref(h)))
        self.handle = h
        self._lib = lib
        self._occ = occ.copy()
        # calculate's argument buffers, per number of changes (the library
        # copies them during the call), and its energy
        self._bufs = {}
        self._e = ctypes.c_double(0.0)
        self._e_ref = ctypes.byref(self._e)
        # changes in the device's history since the last clear, as
        # (site, old symbol, new symbol): an undo restores exactly these on
        # the atoms object; None when unknown (after an error), and then the
        # atoms object is compared with the lattice instead
        self._pending = []

    # --------------------------------------------------------------- helpers
    def _ids(self, changes):
        sid = self.model.sid
        try:
            site = np.array([int(c[0]) for c in changes], dtype=np.int32)
            old = np.array([sid[c[1]] for c in changes], dtype=np.int32)
            new = np.array([sid[c[2]] for c in changes], dtype=np.int32)
        except KeyError as exc:
            raise ValueError("unknown symbol %s" % exc)
        return site, old, new

    def _sync_atoms(self):
        """Write species that changed on the device back into the atoms object."""
        occ = np.zeros(self.model.n_sites, dtype=np.int8)
        _capi.check(self._lib.cemc_updater_get_occupation(self.handle, occ.ctypes.data_as(_capi._i8p)))
        diff = np.nonzero(occ != self._occ)[0]
        for i in diff:
            self.atoms[int(i)].symbol = self.model.species[int(occ[i])]
        self._occ = occ

    # ------------------------------------------------------------- interface
    def clear_history(self):
        rc = self._lib.cemc_updater_clear_history(self.handle)
        if rc:
            _capi.check(rc)
        self._pending = []

    def undo_changes(self):
        rc = self._lib.cemc_updater_undo_changes(self.handle)
        if rc:
            _capi.check(rc)
        pending = self._pending
        # the history holds at most 999 changes (ce_updater's ring of 1000 slots)
        if pending is None or len(pending) > _MAX_UNDO:
            self._sync_atoms()
        else:
            atoms, occ, sid = self.atoms, self._occ, self.model.sid
            for i, old, _ in reversed(pending):
                atoms[i].symbol = old
                occ[i] = sid[old]
        self._pending = []

    def _applied(self, changes):
        """Mirror changes the device made (no-op changes, old == new, are
        not recorded, ce_updater.cpp:315-318) on the atoms object and in the
        pending list."""
        atoms, occ, sid, pending = self.atoms, self._occ, self.model.sid, self._pending
        for c in changes:
            if c[1] == c[2]:
                continue
            i = int(c[0])
            atoms[i].symbol = c[2]
            occ[i] = sid[c[2]]
            if pending is not None:
                pending.append((i, c[1], c[2]))

    def update_cf(self, system_changes):
        site, old, new = self._ids([system_changes])
        rc = self._lib.cemc_updater_update_cf(self.handle, int(site[0]), int(old[0]), int(new[0]))
        if rc == _capi.CEMC_OK:
            self._applied([system_changes])
        else:
            self._pending = None
            self._sync_atoms()
        _capi.check(rc)

    def calculate(self, system_changes):
        changes = list(system_changes)
        n = len(changes)
        sid = self.model.sid
        bufs = self._bufs.get(n)
        if bufs is None:
            arr = ctypes.c_int32 * n
            bufs = self._bufs[n] = (arr(), arr(), arr())
        site, old, new = bufs
        try:
            for k, c in enumerate(changes):
                site[k] = int(c[0])
                old[k] = sid[c[1]]
                new[k] = sid[c[2]]
        except KeyError as exc:
            raise ValueError("unknown symbol %s" % exc)
        e = self._e
        e.value = 0.0
        rc = self._lib.cemc_updater_calculate(self.handle, n, site, old, new, self._e_ref)
        if rc == _capi.CEMC_OK:
            # the changes were made as given: set the same symbols in atoms
            # (ce_updater.cpp:338-347) without reading the lattice back
            # (no-op changes, old == new, are not recorded, :315-318)
            atoms, occ, pending = self.atoms, self._occ, self._pending
            for c in changes:
                if c[1] != c[2]:
                    i = int(c[0])
                    atoms[i].symbol = c[2]
                    occ[i] = sid[c[2]]
                    if pending is not None:
                        pending.append((i, c[1], c[2]))
            return e.value
        self._pending = None
        self._sync_atoms()
        _capi.check(rc)
        return e.value

    def get_energy(self):
        e = ctypes.c_double(0.0)
        _capi.check(self._lib.cemc_updater_get_energy(self.handle, ctypes.byref(e)))
        return e.value

    def get_cf_array(self):
        out = np.zeros(self.model.n_cf)
        _capi.check(self._lib.cemc_updater_get_cf(self.handle, out.ctypes.data_as(_capi._dp)))
        return out

    def get_cf(self):
        return dict(zip(self.model.names, self.get_cf_array().tolist()))

    def set_ecis(self, ecis):
        """ECI update by name (cpp/src/ce_updater.cpp:615-629)."""
        for name in ecis:
            if name not in self.eci:
                raise ValueError("No name corresponding to " + name)
        self.eci.update({k: float(v) for k, v in ecis.items()})
        a, p = _capi.f64(np.array([self.eci[n] for n in self.model.names]))
        _capi.check(self._lib.cemc_updater_set_ecis(self.handle, p))

    def get_singlets(self):
        n = ctypes.c_int32(0)
        _capi.check(self._lib.cemc_updater_get_singlets(self.handle, None, 0, ctypes.byref(n)))
        out = np.zeros(n.value)
        _capi.check(self._lib.cemc_updater_get_singlets(self.handle, out.ctypes.data_as(_capi._dp), n.value, None))
        return out

    def get_symbols(self):
        occ = np.zeros(self.model.n_sites, dtype=np.int8)
        _capi.check(self._lib.cemc_updater_get_occupation(self.handle, occ.ctypes.data_as(_capi._i8p)))
        return self.model.symbols(occ)

    def get_occupation(self):
        occ = np.zeros(self.model.n_sites, dtype=np.int8)
        _capi.check(self._lib.cemc_updater_get_occupation(self.handle, occ.ctypes.data_as(_capi._i8p)))
        return occ

    def set_occupation(self, occupation):
        """Replace the whole occupation (species ids); CFs are recomputed on
        the GPU and the history is cleared."""
        occ = np.ascontiguousarray(occupation, dtype=np.int8)
        if occ.shape != (self.model.n_sites,):
            raise ValueError("Length of the symbols array has to match the length of the atoms object.!")
        rc = self._lib.cemc_updater_set_occupation(self.handle, occ.ctypes.data_as(_capi._i8p))
        _capi.check(rc)
        self._pending = []  # the history is cleared
        self._sync_atoms()

    def set_symbols(self, symbols):
        """CE.set_symbols (cemc/ce_calculator.py:436-448) in one device call."""
        if len(symbols) != self.model.n_sites:
            raise ValueError("Length of the symbols array has to match the length of the atoms object.!")
        try:
            occ = self.model.species_ids(symbols)
        except KeyError as exc:
            raise ValueError("unknown symbol %s" % exc)
        self.set_occupation(occ)

    def history_size(self):
        v = ctypes.c_int32(0)
        _capi.check(self._lib.cemc_updater_history_size(self.handle, ctypes.byref(v)))
        return v.value

    def set_resident(self, on=True):
        """Calls as commands to a resident GPU kernel (the default) or as
        one kernel launch each; the results are identical."""
        _capi.check(self._lib.cemc_updater_set_resident(self.handle, 1 if on else 0))

    def set_num_threads(self, num_threads):
        """Stored for interface compatibility; the work is spread over GPU lanes."""
        if int(num_threads) <= 0:
            raise ValueError("num_threads must be positive")
        _capi.check(self._lib.cemc_updater_set_num_threads(self.handle, int(num_threads)))
        self._num_threads = int(num_threads)

    def add_linear_vib_correction(self, eci_per_kbT):
        """LinearVibCorrection (cpp/src/linear_vib_correction.cpp:6-16)."""
        pos = {n: k for k, n in enumerate(self.model.names)}
        unknown = sorted(set(eci_per_kbT) - set(pos))
        if unknown:
            raise ValueError("no correlation function named %s" % unknown[0])
        idx, ip = _capi.i32([pos[n] for n in eci_per_kbT])
        val, vp = _capi.f64([float(v) for v in eci_per_kbT.values()])
        _capi.check(self._lib.cemc_updater_add_linear_vib_correction(self.handle, len(idx), ip, vp))
        self._vib = dict(eci_per_kbT)

    def vib_energy(self, T):
        e = ctypes.c_double(0.0)
        _capi.check(self._lib.cemc_updater_vib_energy(self.handle, float(T), ctypes.byref(e)))
        return e.value

    def close(self):
        if getattr(self, "handle", None):
            self._lib.cemc_updater_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
