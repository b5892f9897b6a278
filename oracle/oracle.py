"""Test infrastructure: Python side of the CPU oracle (libce_oracle.so).

``OracleUpdater`` parses a settings object the way ``CEUpdater::init`` does
(cpp/src/ce_updater.cpp:32-234: sorted species ids, translation groups,
background sites, cluster info, basis functions, used translation-matrix
columns, ECIs in sorted-name order) and hands flat tables to the C
restatement in ce_oracle.c.  Its methods mirror the reference binding
(cemc/cpp_ext/pyce_updater.pyx:26-60).  Only tests, the smoke check and the
CPU-baseline leg of the benchmark use this module.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from .brute_force import equivalent_decorations

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libce_oracle.so")
_lib = None

CF_C0, CF_SINGLET, CF_CLUSTER = 0, 1, 2
KIND_CANONICAL, KIND_SGC = 0, 1


def load_library():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        raise ImportError("oracle library not built: run `make -C oracle`")
    lib = ctypes.CDLL(_LIB_PATH)
    P = ctypes.c_void_p
    ip = ctypes.POINTER(ctypes.c_int)
    dp = ctypes.POINTER(ctypes.c_double)
    lib.ceo_model_create.restype = P
    lib.ceo_model_create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, dp, ctypes.c_int, ip, ip,
                                     ctypes.c_int, ip, ip, ip, ip, ip, ctypes.c_int, ip, ip, ip, ip, ip,
                                     ip, ip, ctypes.c_int, ip]
    lib.ceo_model_destroy.argtypes = [P]
    lib.ceo_chain_create.restype = P
    lib.ceo_chain_create.argtypes = [P, ctypes.POINTER(ctypes.c_int8), dp, dp]
    lib.ceo_chain_destroy.argtypes = [P]
    lib.ceo_chain_set_threads.argtypes = [P, ctypes.c_int]
    lib.ceo_update_cf.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    lib.ceo_calculate.argtypes = [P, ctypes.c_int, ip, ip, ip, dp]
    lib.ceo_undo_changes.argtypes = [P]
    lib.ceo_clear_history.argtypes = [P]
    lib.ceo_get_energy.restype = ctypes.c_double
    lib.ceo_get_energy.argtypes = [P]
    lib.ceo_get_cf.argtypes = [P, dp]
    lib.ceo_set_ecis.argtypes = [P, dp]
    lib.ceo_get_occupation.argtypes = [P, ctypes.POINTER(ctypes.c_int8)]
    lib.ceo_history_size.restype = ctypes.c_int
    lib.ceo_history_size.argtypes = [P]
    lib.ceo_last_error.restype = ctypes.c_char_p
    lib.ceo_philox4x64.argtypes = [ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64),
                                   ctypes.POINTER(ctypes.c_uint64)]
    lib.ceo_run.argtypes = [P, ctypes.c_int, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64,
                            ctypes.c_int64, ctypes.c_double, ctypes.POINTER(ctypes.c_uint8), dp,
                            ctypes.c_int64, ctypes.POINTER(ctypes.c_int64)]
    lib.ceo_run_many.argtypes = [ctypes.POINTER(P), ctypes.c_int, ctypes.c_int, ctypes.c_uint64,
                                 ctypes.POINTER(ctypes.c_uint64), ctypes.c_uint64, ctypes.c_int64,
                                 dp, ctypes.c_int, ctypes.POINTER(ctypes.c_int64)]
    lib.ceo_run_hooked.argtypes = [P, ctypes.c_int, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64,
                                   ctypes.c_int64, ctypes.c_double, ctypes.POINTER(ctypes.c_uint8), dp,
                                   ctypes.c_int64, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(Hooks)]
    lib.ceo_wl_run.argtypes = [P, ctypes.POINTER(WLState), ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64,
                               ctypes.c_int64, ctypes.c_int64, ctypes.POINTER(ctypes.c_uint8), ctypes.c_int64,
                               ctypes.POINTER(ctypes.c_int64)]
    _lib = lib
    return lib


class Hooks(ctypes.Structure):
    """ceo_hooks of ce_oracle.h."""
    _fields_ = [("fixed_species", ctypes.c_uint32), ("max_attempts", ctypes.c_int32),
                ("allowed", ctypes.POINTER(ctypes.c_uint8)), ("pair_e1", ctypes.c_int32),
                ("pair_e2", ctypes.c_int32), ("pair_nm", ctypes.c_int32), ("pair_nb", ctypes.POINTER(ctypes.c_int32)),
                ("bias_species", ctypes.c_int32), ("bias_table", ctypes.POINTER(ctypes.c_double)),
                ("failed_move", ctypes.c_int64), ("recycle", ctypes.c_int32), ("pad", ctypes.c_int32),
                ("rec_sums", ctypes.POINTER(ctypes.c_double))]


class WLState(ctypes.Structure):
    """ceo_wl of ce_oracle.h."""
    _fields_ = [("n_bins", ctypes.c_int32), ("cur_bin", ctypes.c_int32), ("is_first", ctypes.c_int32),
                ("converged", ctypes.c_int32), ("e_min", ctypes.c_double), ("e_max", ctypes.c_double),
                ("f", ctypes.c_double), ("f_min", ctypes.c_double), ("flatness", ctypes.c_double),
                ("logdos", ctypes.POINTER(ctypes.c_double)), ("hist", ctypes.POINTER(ctypes.c_int32)),
                ("known", ctypes.POINTER(ctypes.c_uint8)), ("n_moves", ctypes.c_int64),
                ("n_accept", ctypes.c_int64), ("n_outside", ctypes.c_int64), ("n_self", ctypes.c_int64),
                ("n_flat", ctypes.c_int64),
                ("inverse_time", ctypes.c_int32), ("adaptive", ctypes.c_int32), ("min_width", ctypes.c_int32),
                ("inv_active", ctypes.c_int32), ("upper", ctypes.c_int32), ("n_overlap", ctypes.c_int32),
                ("n_edges", ctypes.c_int32), ("has_first", ctypes.c_int32), ("inv_factor", ctypes.c_double),
                ("bin_change", ctypes.c_double), ("cur_energy", ctypes.c_double), ("steps", ctypes.c_int64),
                ("since_flat", ctypes.c_int64), ("prev_conv", ctypes.c_int64), ("n_walk", ctypes.c_int64),
                ("edges", ctypes.POINTER(ctypes.c_int32)), ("edge_n", ctypes.POINTER(ctypes.c_int32)),
                ("edge_vals", ctypes.POINTER(ctypes.c_double)), ("saved", ctypes.POINTER(ctypes.c_uint8)),
                ("snap_occ", ctypes.POINTER(ctypes.c_int8)), ("snap_cf", ctypes.POINTER(ctypes.c_double)),
                ("snap_trk", ctypes.POINTER(ctypes.c_int32)), ("snap_bin", ctypes.POINTER(ctypes.c_int32)),
                ("sub_hist", ctypes.POINTER(ctypes.c_uint32)), ("kind", ctypes.c_int32), ("pad", ctypes.c_int32)]


class OracleWL:
    """A Wang-Landau walker of the oracle: histogram arrays owned here.

    inverse_time, adaptive and min_width choose the schemes of ce_oracle.h;
    n_sites and n_cf size the state records of adaptive windows; kind 1
    makes semi-grand-canonical flips (the updater's ECIs carry the chemical
    potentials)."""

    def __init__(self, n_bins, e_min, e_max, f, f_min=1e-6, flatness=0.8, cur_bin=0, is_first=True,
                 logdos=None, hist=None, known=None, inverse_time=False, adaptive=False, min_width=10,
                 n_sites=0, n_cf=0, kind=0):
        nb = int(n_bins)
        self.logdos = np.zeros(nb) if logdos is None else np.array(logdos, dtype=np.float64)
        self.hist = np.zeros(nb, dtype=np.int32) if hist is None else np.array(hist, dtype=np.int32)
        self.known = np.zeros(nb, dtype=np.uint8) if known is None else np.array(known, dtype=np.uint8)
        rows = nb + 1 if adaptive else 1
        self.edges = np.zeros(nb, dtype=np.int32)
        self.edge_n = np.zeros(nb, dtype=np.int32)
        self.edge_vals = np.zeros(nb * nb if adaptive else 1)
        self.saved = np.zeros(nb, dtype=np.uint8)
        self.snap_occ = np.zeros(rows * max(int(n_sites), 1), dtype=np.int8)
        self.snap_cf = np.zeros(rows * max(int(n_cf), 1))
        self.snap_trk = np.zeros(rows * max(int(n_sites), 1), dtype=np.int32)
        self.snap_bin = np.zeros(nb + 1, dtype=np.int32)
        self.sub_hist = np.zeros(nb * 10, dtype=np.uint32)
        if adaptive and (n_sites <= 0 or n_cf <= 0):
            raise ValueError("adaptive windows need n_sites and n_cf for the state records")
        i32 = ctypes.POINTER(ctypes.c_int32)
        self.state = WLState(nb, int(cur_bin), int(bool(is_first)), 0, float(e_min), float(e_max),
                             float(f), float(f_min), float(flatness), _dp(self.logdos),
                             self.hist.ctypes.data_as(i32),
                             self.known.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), 0, 0, 0, 0, 0,
                             int(bool(inverse_time)), int(bool(adaptive)), int(min_width), 0, nb, 7, 0, 0,
                             0.0, 0.0, 0.0, 0, 0, -1, 0,
                             self.edges.ctypes.data_as(i32), self.edge_n.ctypes.data_as(i32),
                             _dp(self.edge_vals), self.saved.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
                             self.snap_occ.ctypes.data_as(ctypes.POINTER(ctypes.c_int8)), _dp(self.snap_cf),
                             self.snap_trk.ctypes.data_as(i32), self.snap_bin.ctypes.data_as(i32),
                             self.sub_hist.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), int(kind), 0)

    def sub_bins(self):
        """Sub-bin visit counts, (n_bins, 10)."""
        return self.sub_hist.reshape(-1, 10)


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def philox4x64(ctr, key):
    """One Philox4x64-10 block from the C oracle."""
    lib = load_library()
    c = (ctypes.c_uint64 * 4)(*[int(v) & 0xFFFFFFFFFFFFFFFF for v in ctr])
    k = (ctypes.c_uint64 * 2)(*[int(v) & 0xFFFFFFFFFFFFFFFF for v in key])
    out = (ctypes.c_uint64 * 4)()
    lib.ceo_philox4x64(c, k, out)
    return [int(v) for v in out]


def _raise(lib, rc):
    msg = lib.ceo_last_error().decode()
    if rc == 1:
        raise ValueError(msg)
    raise RuntimeError(msg)


class OracleModel:
    """Flat tables parsed from a settings object for a given ECI name set."""

    def __init__(self, bc, names):
        lib = load_library()
        self.names = sorted(names)
        uniq = sorted(set(bc.unique_elements))
        self.species = uniq
        self.sid = {s: i for i, s in enumerate(uniq)}
        n_s = len(uniq)
        n_bfs = int(bc.num_unique_elements) - 1
        basis = np.zeros((n_bfs, n_s))
        for d, bf in enumerate(bc.basis_functions[:n_bfs]):
            for s, v in bf.items():
                basis[d, self.sid[s]] = v
        n_sites = len(bc.trans_matrix)
        self.n_sites = n_sites
        bkg = set(int(i) for i in bc.background_indices)
        site_group = np.full(n_sites, -1, dtype=np.int32)
        for g, sites in enumerate(bc.index_by_trans_symm):
            for i in sites:
                site_group[int(i)] = g
        background = np.zeros(n_sites, dtype=np.int32)
        background[sorted(bkg)] = 1
        self.site_group = site_group
        self.background = background
        n_groups = len(bc.cluster_info)
        cols = sorted({int(c) for info in bc.cluster_info for cl in info.values()
                       for sub in cl["indices"] for c in sub})
        slot = {c: j for j, c in enumerate(cols)}
        K = len(cols)
        tm = bc.trans_matrix
        trans = np.zeros((n_sites, max(K, 1)), dtype=np.int32)
        if isinstance(tm, np.ndarray):
            trans[:, :K] = tm[:, cols]
        else:
            for i in range(n_sites):
                if background[i]:
                    continue
                row = tm[i]
                for j, c in enumerate(cols):
                    trans[i, j] = row[c]
        # clusters, one entry per (group, prefix)
        cl_size, cl_M, cl_norm, mem_off, ord_off, members, order = [], [], [], [0], [0], [], []
        cl_index = {}
        norm = {}
        for info in bc.cluster_info:
            for name, cl in info.items():
                norm[name] = norm.get(name, 0) + len(cl["indices"])
        equiv = {}
        for g, info in enumerate(bc.cluster_info):
            for name, cl in info.items():
                n = int(cl["size"])
                if n < 2:
                    continue
                cl_index[(g, name)] = len(cl_size)
                subs = cl["indices"]
                cl_size.append(n)
                cl_M.append(len(subs))
                cl_norm.append(norm[name])
                for sub in subs:
                    members.extend(slot[int(c)] for c in sub)
                for o in cl["order"]:
                    order.extend(int(v) for v in o)
                mem_off.append(len(members))
                ord_off.append(len(order))
                equiv[name] = [list(e) for e in cl["equiv_sites"]]
        n_cf = len(self.names)
        kind = np.zeros(n_cf, dtype=np.int32)
        dec = np.zeros(n_cf, dtype=np.int32)
        cf_cluster = np.full((n_groups, n_cf), -1, dtype=np.int32)
        eq_off = [0]
        eq_decos = []
        for k, name in enumerate(self.names):
            if name.startswith("c0"):
                kind[k] = CF_C0
            elif name.startswith("c1"):
                kind[k] = CF_SINGLET
                dec[k] = int(name.rsplit("_", 1)[1][0])
            else:
                kind[k] = CF_CLUSTER
                prefix, d = name.rsplit("_", 1)
                digits = [int(ch) for ch in d]
                for g in range(n_groups):
                    if (g, prefix) in cl_index:
                        cf_cluster[g, k] = cl_index[(g, prefix)]
                eqs = equivalent_decorations(digits, equiv[prefix]) if equiv.get(prefix) else [digits]
                for e in eqs:
                    eq_decos.extend(list(e) + [0] * (4 - len(e)))
            eq_off.append(len(eq_decos) // 4)
        arrs = dict(
            basis=np.ascontiguousarray(basis.ravel()), site_group=site_group, background=background,
            kind=kind, dec=dec, cf_cluster=np.ascontiguousarray(cf_cluster.ravel()),
            eq_off=np.array(eq_off, dtype=np.int32), eq_decos=np.array(eq_decos + [0] * 4, dtype=np.int32),
            cl_size=np.array(cl_size + [0], dtype=np.int32), cl_M=np.array(cl_M + [0], dtype=np.int32),
            cl_norm=np.array(cl_norm + [0], dtype=np.int32), mem_off=np.array(mem_off, dtype=np.int32),
            members=np.array(members + [0], dtype=np.int32), ord_off=np.array(ord_off, dtype=np.int32),
            order=np.array(order + [0], dtype=np.int32), trans=np.ascontiguousarray(trans))
        self._arrs = arrs
        self.n_cf = n_cf
        self.n_species = n_s
        self.handle = lib.ceo_model_create(
            n_sites, n_s, n_bfs, _dp(arrs["basis"]), n_groups, _ip(site_group), _ip(background), n_cf, _ip(kind),
            _ip(dec), _ip(arrs["cf_cluster"]), _ip(arrs["eq_off"]), _ip(arrs["eq_decos"]), len(cl_size),
            _ip(arrs["cl_size"]), _ip(arrs["cl_M"]), _ip(arrs["cl_norm"]), _ip(arrs["mem_off"]),
            _ip(arrs["members"]), _ip(arrs["ord_off"]), _ip(arrs["order"]), K, _ip(arrs["trans"]))
        if not self.handle:
            raise MemoryError("oracle model allocation failed")

    def __del__(self):
        if getattr(self, "handle", None) and _lib is not None:
            _lib.ceo_model_destroy(self.handle)
            self.handle = None


class OracleUpdater:
    """CPU restatement with the method set of the reference binding."""

    def __init__(self, model, occupation, cf, eci):
        lib = load_library()
        self.model = model
        occ = np.ascontiguousarray(np.asarray(occupation, dtype=np.int8))
        cfv = np.array([cf[n] for n in model.names], dtype=np.float64)
        ev = np.array([eci[n] for n in model.names], dtype=np.float64)
        self._eci = ev
        self.handle = lib.ceo_chain_create(model.handle, occ.ctypes.data_as(ctypes.POINTER(ctypes.c_int8)),
                                           _dp(cfv), _dp(ev))
        if not self.handle:
            raise MemoryError("oracle chain allocation failed")

    def __del__(self):
        if getattr(self, "handle", None) and _lib is not None:
            _lib.ceo_chain_destroy(self.handle)
            self.handle = None

    def _ids(self, changes):
        sid = self.model.sid
        site = np.array([int(c[0]) for c in changes], dtype=np.int32)
        old = np.array([sid[c[1]] if isinstance(c[1], str) else int(c[1]) for c in changes], dtype=np.int32)
        new = np.array([sid[c[2]] if isinstance(c[2], str) else int(c[2]) for c in changes], dtype=np.int32)
        return site, old, new

    def set_num_threads(self, n):
        _lib.ceo_chain_set_threads(self.handle, int(n))

    def update_cf(self, change):
        s, o, n = self._ids([change])
        rc = _lib.ceo_update_cf(self.handle, int(s[0]), int(o[0]), int(n[0]))
        if rc:
            _raise(_lib, rc)

    def calculate(self, changes):
        s, o, n = self._ids(changes)
        e = ctypes.c_double(0.0)
        rc = _lib.ceo_calculate(self.handle, len(changes), _ip(s), _ip(o), _ip(n), ctypes.byref(e))
        if rc:
            _raise(_lib, rc)
        return e.value

    def undo_changes(self):
        _lib.ceo_undo_changes(self.handle)

    def clear_history(self):
        _lib.ceo_clear_history(self.handle)

    def get_energy(self):
        return _lib.ceo_get_energy(self.handle)

    def get_cf(self):
        out = np.zeros(self.model.n_cf)
        _lib.ceo_get_cf(self.handle, _dp(out))
        return dict(zip(self.model.names, out.tolist()))

    def get_cf_array(self):
        out = np.zeros(self.model.n_cf)
        _lib.ceo_get_cf(self.handle, _dp(out))
        return out

    def set_ecis(self, eci):
        """Update ECIs by name (cpp/src/ce_updater.cpp:615-629)."""
        pos = {n: k for k, n in enumerate(self.model.names)}
        for name, val in eci.items():
            if name not in pos:
                raise ValueError("No name corresponding to " + name)
            self._eci[pos[name]] = float(val)
        _lib.ceo_set_ecis(self.handle, _dp(self._eci))

    def get_singlets(self):
        cf = self.get_cf()
        return np.array([cf[n] for n in self.model.names if n.startswith("c1")])

    def get_occupation(self):
        out = np.zeros(self.model.n_sites, dtype=np.int8)
        _lib.ceo_get_occupation(self.handle, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int8)))
        return out

    def get_symbols(self):
        return [self.model.species[s] for s in self.get_occupation()]

    def history_size(self):
        return _lib.ceo_history_size(self.handle)

    def run(self, kind, seed, chain_uid, n_moves, temperature, move_begin=0, trace_every=0):
        """Fed-stream Metropolis; returns (accept bits packed LSB-first, energy trace, n_accept)."""
        bits = np.zeros((n_moves + 7) // 8, dtype=np.uint8)
        n_tr = n_moves // trace_every if trace_every else 0
        trace = np.zeros(max(n_tr, 1))
        acc = ctypes.c_int64(0)
        rc = _lib.ceo_run(self.handle, int(kind), int(seed), int(chain_uid), int(move_begin), int(n_moves),
                          float(temperature), bits.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
                          _dp(trace), int(trace_every), ctypes.byref(acc))
        if rc:
            _raise(_lib, rc)
        return bits, trace[:n_tr], acc.value


def wl_run(updater, wl, seed, chain_uid, n_checks, check_every, move_begin=0, counters=None):
    """Wang-Landau walker `wl` (OracleWL) on the canonical chain of `updater`;
    returns the accept bits of the steps made (packed LSB-first, bit t for
    Philox counter move_begin + t).  counters: a list that receives the
    number of Philox counters used (steps plus moves back into a window)."""
    lib = load_library()
    n = int(n_checks) * int(check_every)
    bits = np.zeros(max((n + 7) // 8, 1), dtype=np.uint8)
    used = ctypes.c_int64(0)
    rc = lib.ceo_wl_run(updater.handle, ctypes.byref(wl.state), int(seed), int(chain_uid), int(move_begin),
                        int(n_checks), int(check_every), bits.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
                        int(8 * bits.size), ctypes.byref(used))
    if counters is not None:
        counters.append(int(used.value))
    if rc:
        _raise(lib, rc)
    return bits


def run_hooked(updater, kind, seed, chain_uid, n_moves, temperature, fixed_species=0, max_attempts=10000,
               allowed=None, pair=None, pair_neighbours=None, bias_species=-1, bias_table=None, move_begin=0,
               trace_every=0, recycle_sums=None):
    """Fed-stream run with constraints and bias; returns (accept bits, energy
    trace, n_accept, failed_move or None).  pair: (e1, e2) species ids with
    pair_neighbours an (n_sites, n_members) array.  recycle_sums: a float64
    array of 3 + 3 n_bfs accumulators switches waste recycling on."""
    lib = load_library()
    keep = []
    h = Hooks()
    h.fixed_species = int(fixed_species)
    h.max_attempts = int(max_attempts)
    if allowed is not None:
        a = np.ascontiguousarray(allowed, dtype=np.uint8)
        keep.append(a)
        h.allowed = a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))
    h.pair_e1, h.pair_e2 = (-1, -1) if pair is None else (int(pair[0]), int(pair[1]))
    if pair is not None:
        nb = np.ascontiguousarray(pair_neighbours, dtype=np.int32)
        keep.append(nb)
        h.pair_nm = nb.shape[1]
        h.pair_nb = nb.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
    if recycle_sums is not None:
        if recycle_sums.dtype != np.float64 or not recycle_sums.flags["C_CONTIGUOUS"]:
            raise ValueError("recycle_sums must be a contiguous float64 array")
        h.recycle = 1
        h.rec_sums = _dp(recycle_sums)
    h.bias_species = int(bias_species)
    if bias_species >= 0:
        bt = np.ascontiguousarray(bias_table, dtype=np.float64)
        keep.append(bt)
        h.bias_table = _dp(bt)
    bits = np.zeros((n_moves + 7) // 8 + 1, dtype=np.uint8)
    n_tr = n_moves // trace_every if trace_every else 0
    trace = np.zeros(max(n_tr, 1))
    acc = ctypes.c_int64(0)
    rc = lib.ceo_run_hooked(updater.handle, int(kind), int(seed), int(chain_uid), int(move_begin), int(n_moves),
                            float(temperature), bits.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), _dp(trace),
                            int(trace_every), ctypes.byref(acc), ctypes.byref(h))
    failed = None if h.failed_move < 0 else int(h.failed_move)
    if rc and failed is None:
        _raise(lib, rc)
    return bits, trace[:n_tr], acc.value, failed


def run_many(updaters, kind, seed, chain_uids, n_moves, temperatures, n_threads, move_begin=0):
    """Advance independent oracle chains in parallel; returns accept counts."""
    lib = load_library()
    n = len(updaters)
    handles = (ctypes.c_void_p * n)(*[u.handle for u in updaters])
    uids = (ctypes.c_uint64 * n)(*[int(v) for v in chain_uids])
    temps = np.ascontiguousarray(np.asarray(temperatures, dtype=np.float64))
    acc = (ctypes.c_int64 * n)()
    rc = lib.ceo_run_many(handles, n, int(kind), int(seed), uids, int(move_begin), int(n_moves), _dp(temps),
                          int(n_threads), acc)
    if rc:
        raise RuntimeError("oracle run failed")
    return [int(v) for v in acc]
