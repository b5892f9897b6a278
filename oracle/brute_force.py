"""Test infrastructure: correlation functions computed from scratch.

Independent of the incremental update: every correlation function is
evaluated as the normalised sum, over all sites of each translation group and
every sub-cluster listed for that site, of the decorated basis-function product
(the definition the reference tests compare the updater against,
tests/test_CE_updater.py:40-115, where CLEASE's CorrFunction plays this role).

    cf[c0]      = 1
    cf[c1_d]    = (1/N) * sum_{i not background} phi_d(s_i)
    cf[p_delta] = sum_g 1/(|E| C_p T_g) * sum_{i in g} sum_{m in sub_g(p)}
                  sum_{e in E(delta)} prod_k phi_{e_k}(s_{order_m(i, trans(i, members_m))_k})

with E the equivalent decorations of delta, C_p the number of sub-clusters of
orbit p summed over groups and T_g the number of sites in group g
(normalisation used by cpp/src/ce_updater.cpp:137-144, 386-404).

Pure numpy; runs on the interpreter that drives the compiled reference too.
"""
from __future__ import annotations

import itertools

import numpy as np


def equivalent_decorations(deco, equiv_sites):
    """Permutations of decoration digits inside each group of equivalent sites,
    Cartesian product over groups, duplicates kept."""
    out = [list(deco)]
    for group in equiv_sites:
        nxt = []
        for d in out:
            for perm in itertools.permutations([d[p] for p in group]):
                e = list(d)
                for p, v in zip(group, perm):
                    e[p] = v
                nxt.append(e)
        out = nxt
    return out


def _species_ids(bc):
    uniq = sorted(set(bc.unique_elements))
    return {s: i for i, s in enumerate(uniq)}, uniq


def _basis_table(bc, uniq):
    return np.array([[bf[s] for s in uniq] for bf in bc.basis_functions], dtype=np.float64)


def _trans_array(bc, cols):
    tm = bc.trans_matrix
    if isinstance(tm, np.ndarray):
        return np.ascontiguousarray(tm[:, cols].astype(np.int64))
    n = len(tm)
    out = np.full((n, len(cols)), -1, dtype=np.int64)
    for i in range(n):
        row = tm[i]
        for j, c in enumerate(cols):
            out[i, j] = row.get(c, -1) if isinstance(row, dict) else row[c]
    return out


def brute_force_cf(bc, occupation, names, chunk=4096):
    """Correlation functions ``names`` for species ids ``occupation``."""
    sid, uniq = _species_ids(bc)
    phi = _basis_table(bc, uniq)
    occ = np.asarray(occupation, dtype=np.int64)
    n_sites = len(occ)
    bkg = np.zeros(n_sites, dtype=bool)
    bkg[list(bc.background_indices)] = True
    groups = [np.array(sorted(g), dtype=np.int64) for g in bc.index_by_trans_symm]
    cols = sorted({c for info in bc.cluster_info for cl in info.values() for sub in cl["indices"] for c in sub})
    col_slot = {c: j for j, c in enumerate(cols)}
    trans = _trans_array(bc, cols) if cols else np.zeros((n_sites, 0), dtype=np.int64)

    norm = {}
    for info in bc.cluster_info:
        for name, cl in info.items():
            norm[name] = norm.get(name, 0) + len(cl["indices"])

    out = {}
    # cluster names grouped by prefix: the member species of a prefix's
    # sub-clusters are gathered once and shared by all its decorations
    by_prefix = {}
    for name in names:
        if name.startswith("c0"):
            out[name] = 1.0
            continue
        prefix, dec = name.rsplit("_", 1)
        digits = [int(ch) for ch in dec]
        if name.startswith("c1"):
            vals = phi[digits[0], occ]
            out[name] = float(vals[~bkg].sum() / n_sites)
            continue
        by_prefix.setdefault(prefix, []).append((name, digits))
    for prefix, items in by_prefix.items():
        total = {name: 0.0 for name, _ in items}
        for g, info in enumerate(bc.cluster_info):
            if prefix not in info:
                continue
            cl = info[prefix]
            n = int(cl["size"])
            members = np.array(cl["indices"], dtype=np.int64).reshape(-1, n - 1)
            order = np.array(cl["order"], dtype=np.int64).reshape(-1, n)
            eqs = [equivalent_decorations(d, cl["equiv_sites"]) if cl["equiv_sites"] else [d] for _, d in items]
            slots = np.vectorize(col_slot.get)(members) if members.size else members
            sites_g = groups[g]
            acc = [0.0] * len(items)
            for lo in range(0, len(sites_g), chunk):
                ref = sites_g[lo:lo + chunk]
                idx = np.empty((len(ref), len(members), n), dtype=np.int64)
                idx[:, :, 0] = ref[:, None]
                for j in range(n - 1):
                    idx[:, :, j + 1] = trans[ref][:, slots[:, j]]
                ordered = np.take_along_axis(idx, np.broadcast_to(order[None], idx.shape), axis=2)
                spec = occ[ordered]
                for t, eq in enumerate(eqs):
                    for e in eq:
                        prod = np.ones(spec.shape[:2])
                        for k in range(n):
                            prod = prod * phi[e[k], spec[:, :, k]]
                        acc[t] += prod.sum()
            for t, (name, _) in enumerate(items):
                total[name] += acc[t] / (len(eqs[t]) * norm[prefix] * len(sites_g))
        for name, _ in items:
            out[name] = float(total[name])
    return {name: out[name] for name in names}
