"""Device model: a settings object + ECIs flattened into the C ABI's model
description and uploaded to the GPU.

The parsing mirrors what ``CEUpdater::init`` reads (cpp/src/ce_updater.cpp:32-234):
species ids in sorted-name order (cpp/src/symbols_with_numbers.cpp:11-21),
translation groups and background sites (ce_updater.cpp:733-776, 1015-1027),
one cluster entry per (group, name) of size >= 2 with its sub-cluster
``indices``/``order`` (cpp/src/cluster.cpp:185-232), the cluster normalisation
``C_p`` summed over groups (ce_updater.cpp:137-144), equivalent decorations
(cpp/src/cluster.cpp:41-128), the used translation-matrix columns
(ce_updater.cpp:899-979) and ECIs in sorted-name order
(cpp/src/named_array.cpp:18-23).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _capi
from .lattice import equivalent_decorations


def _parse_name(name):
    prefix, deco = name.rsplit("_", 1)
    return prefix, [int(ch) for ch in deco]


class CEModel:
    """Lattice, clusters, basis and ECIs resident on the GPU."""

    def __init__(self, settings, eci):
        lib = _capi.load()
        self.settings = settings
        names = sorted(eci.keys())
        self.names = names
        self.species = sorted(set(settings.unique_elements))
        self.sid = {s: i for i, s in enumerate(self.species)}
        n_s = len(self.species)
        self.n_species = n_s
        n_bfs = int(settings.num_unique_elements) - 1
        if n_bfs != n_s - 1:
            raise ValueError("num_unique_elements does not match unique_elements")
        basis = np.zeros((n_bfs, n_s))
        for d, bf in enumerate(settings.basis_functions[:n_bfs]):
            for sym, val in bf.items():
                basis[d, self.sid[sym]] = float(val)
        self.basis = basis
        n_sites = len(settings.trans_matrix) if not hasattr(settings, "n_sites") else int(settings.n_sites)
        self.n_sites = n_sites
        site_group = np.full(n_sites, -1, dtype=np.int32)
        for g, sites in enumerate(settings.index_by_trans_symm):
            for i in sites:
                if site_group[int(i)] != -1:
                    raise RuntimeError("One site appears to be present in more than one translation symmetry group!")
                site_group[int(i)] = g
        background = np.zeros(n_sites, dtype=bool)
        background[[int(i) for i in settings.background_indices]] = True
        if np.any((site_group < 0) & ~background):
            raise RuntimeError("a site has not been assigned to any translational symmetry group")
        # background sites keep their group: trans_symm_group_count includes
        # them (cpp/src/ce_updater.cpp:768-775)
        self.site_group = site_group
        self.background = background
        n_groups = len(settings.cluster_info)

        # cluster entries
        norm = {}
        for info in settings.cluster_info:
            for cname, cl in info.items():
                norm[cname] = norm.get(cname, 0) + len(cl["indices"])
        cols = sorted({int(c) for info in settings.cluster_info for cl in info.values()
                       for sub in cl["indices"] for c in sub})
        if not cols:
            raise RuntimeError("It looks like no clusters are present. Max lookup index was 0 for ref_indx 0")
        slot = {c: j for j, c in enumerate(cols)}
        ent_group, ent_size, ent_M, ent_norm, mem_off, ord_off, members, order = [], [], [], [], [0], [0], [], []
        entry_of = {}
        equiv = {}
        for g, info in enumerate(settings.cluster_info):
            for cname, cl in info.items():
                n = int(cl["size"])
                if n < 2:
                    continue
                subs = cl["indices"]
                for sub in subs:
                    if int(cl["ref_indx"]) in [int(v) for v in sub] or len(set(sub)) != len(sub):
                        raise ValueError("self interaction: a site appears twice in one cluster")
                entry_of[(g, cname)] = len(ent_group)
                ent_group.append(g)
                ent_size.append(n)
                ent_M.append(len(subs))
                ent_norm.append(norm[cname])
                for sub in subs:
                    members.extend(slot[int(c)] for c in sub)
                for o in cl["order"]:
                    order.extend(int(v) for v in o)
                mem_off.append(len(members))
                ord_off.append(len(order))
                equiv[cname] = [list(e) for e in cl["equiv_sites"]]

        # correlation functions in sorted-name order
        n_cf = len(names)
        kind = np.zeros(n_cf, dtype=np.int32)
        dec = np.zeros(n_cf, dtype=np.int32)
        cf_entry = np.full((n_groups, n_cf), -1, dtype=np.int32)
        eq_off = [0]
        eq = []
        for k, name in enumerate(names):
            if name.startswith("c0"):
                kind[k] = _capi.CF_C0
            elif name.startswith("c1"):
                kind[k] = _capi.CF_SINGLET
                dec[k] = _parse_name(name)[1][0]
            else:
                kind[k] = _capi.CF_CLUSTER
                prefix, digits = _parse_name(name)
                for g in range(n_groups):
                    if (g, prefix) in entry_of:
                        cf_entry[g, k] = entry_of[(g, prefix)]
                groups = equiv.get(prefix, [])
                decos = equivalent_decorations(digits, groups) if groups else [digits]
                for e in decos:
                    eq.extend(list(e) + [0] * (4 - len(e)))
            eq_off.append(len(eq) // 4)
        self.cf_kind = kind
        self.cf_dec = dec
        self.singlet_names = [n for n in names if n.startswith("c1")]

        # translation table restricted to used columns
        K = len(cols)
        if hasattr(settings, "trans_table") and hasattr(settings, "trans_cols") and \
                list(np.asarray(settings.trans_cols)) == cols:
            trans = np.ascontiguousarray(settings.trans_table, dtype=np.int32)
        else:
            tm = settings.trans_matrix
            trans = np.zeros((n_sites, K), dtype=np.int32)
            if isinstance(tm, np.ndarray):
                if tm.shape[1] < cols[-1] + 1:
                    raise ValueError("Something is wrong with the translation matrix passed.")
                trans[:] = tm[:, cols]
            else:
                for i in range(n_sites):
                    row = tm[i]
                    for j, c in enumerate(cols):
                        if c in row:
                            trans[i, j] = row[c]
                        elif not background[i]:
                            # background rows may be incomplete (ce_updater.cpp:924-927)
                            raise ValueError("Requested value %d is not a key in the dictionary!" % c)
        dims = cells = None
        if hasattr(settings, "dims") and hasattr(settings, "cells"):
            dims = np.array(settings.dims, dtype=np.int32)
            cells = np.ascontiguousarray(np.asarray(settings.cells)[cols], dtype=np.int32)

        self.eci = np.array([float(eci[n]) for n in names])
        self._keep = []

        def arr(a, kind_):
            conv = {"i32": _capi.i32, "f64": _capi.f64}[kind_]
            a2, ptr = conv(a)
            self._keep.append(a2)
            return ptr

        desc = _capi.ModelDesc()
        desc.n_sites = n_sites
        desc.n_species = n_s
        desc.n_groups = n_groups
        desc.n_cf = n_cf
        desc.n_entries = len(ent_group)
        desc.n_trans_cols = K
        desc.basis = arr(basis.ravel(), "f64")
        desc.site_group = arr(site_group, "i32")
        desc.cf_kind = arr(kind, "i32")
        desc.cf_dec = arr(dec, "i32")
        desc.cf_entry = arr(cf_entry.ravel(), "i32")
        desc.cf_eq_off = arr(np.array(eq_off), "i32")
        desc.eq_decos = arr(np.array(eq + [0] * 4), "i32")
        desc.entry_group = arr(np.array(ent_group + [0]), "i32")
        desc.entry_size = arr(np.array(ent_size + [0]), "i32")
        desc.entry_M = arr(np.array(ent_M + [0]), "i32")
        desc.entry_norm = arr(np.array(ent_norm + [0]), "i32")
        desc.entry_mem_off = arr(np.array(mem_off), "i32")
        desc.members = arr(np.array(members + [0]), "i32")
        desc.entry_ord_off = arr(np.array(ord_off), "i32")
        desc.order = arr(np.array(order + [0]), "i32")
        desc.trans = arr(trans.ravel(), "i32")
        desc.periodic_dims = arr(dims, "i32") if dims is not None else None
        desc.col_cells = arr(cells.ravel(), "i32") if cells is not None else None
        desc.eci = arr(self.eci, "f64")
        desc.background = arr(background.astype(np.int32), "i32")
        handle = ctypes.c_void_p()
        _capi.check(lib.cemc_model_create(ctypes.byref(desc), ctypes.byref(handle)))
        self._keep = []
        self.handle = handle
        self.n_cf = n_cf
        self.K = K
        self.n_groups = n_groups
        self.n_subclusters = sum(ent_M)

    @property
    def periodic(self):
        flag = ctypes.c_int(0)
        _capi.check(_capi.load().cemc_model_is_periodic(self.handle, ctypes.byref(flag)))
        return bool(flag.value)

    def set_ecis(self, eci):
        """Replace ECIs by name (the batched path's energy tables are rebuilt)."""
        pos = {n: k for k, n in enumerate(self.names)}
        new = self.eci.copy()
        for name, val in eci.items():
            if name not in pos:
                raise ValueError("All ECIs has to correspond to a correlation function!")
            new[pos[name]] = float(val)
        a, p = _capi.f64(new)
        _capi.check(_capi.load().cemc_model_set_ecis(self.handle, p))
        self.eci = a

    def species_ids(self, symbols):
        return np.array([self.sid[s] for s in symbols], dtype=np.int8)

    def symbols(self, occupation):
        return [self.species[int(s)] for s in occupation]

    def gathers_per_flip(self):
        """Non-changed cluster members read per flip, summed over sub-clusters."""
        tot = 0
        for info in self.settings.cluster_info[:1]:
            for cl in info.values():
                if int(cl["size"]) >= 2:
                    tot += len(cl["indices"]) * (int(cl["size"]) - 1)
        return tot

    def close(self):
        if getattr(self, "handle", None):
            _capi.load().cemc_model_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
