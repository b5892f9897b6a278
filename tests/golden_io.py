"""Loading of the golden fixtures written by tests/golden/make_fixtures.py."""
from __future__ import annotations

import json
import os

import numpy as np

from cemc_amd.lattice import CrystalSettings, FCCSettings, settings_from_kwargs

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

CHAIN_FIXTURES = [
    "binary_canonical_plain.npz",
    "binary_sgc_chain0.npz",
    "binary_sgc_chain1.npz",
    "binary_sgc_chain2.npz",
    "binary_sgc_chain3.npz",
    "ternary_canonical.npz",
    "ternary_sgc.npz",
    "quaternary_canonical.npz",
]

# two-sublattice crystals: two translation groups, or one group and a
# background sublattice
CRYSTAL_FIXTURES = [
    "rocksalt_canonical.npz",
    "rocksalt_background_sgc.npz",
    "cscl_sgc.npz",
    # hexagonal close packing (two translation groups related by the 6_3
    # screw) and a space-group-217 crystal (29 translation groups)
    "hcp_ternary_canonical.npz",
    "hcp_binary_sgc.npz",
    "sg217_binary_sgc.npz",
    "sg217_binary_canonical.npz",
]


class ChainFixture:
    def __init__(self, fname):
        z = np.load(os.path.join(GOLDEN, fname), allow_pickle=False)
        self.fname = fname
        self.meta = json.loads(str(z["meta"]))
        m = self.meta
        shells = {int(k): v for k, v in m["shells"].items()}
        if m.get("crystal") and "classtype" in m["crystal"]:
            self.settings = settings_from_kwargs(dict(m["crystal"], symmetrize_decorations=m["symmetrize"]))
        elif m.get("crystal"):
            cr = m["crystal"]
            self.settings = CrystalSettings(m["lattice"], cr["basis_elements"], shells, cr["structure"],
                                            symmetrize_decorations=m["symmetrize"])
        else:
            self.settings = FCCSettings(m["lattice"], m["species"], shells, symmetrize_decorations=m["symmetrize"])
        self.names = m["names"]
        assert self.names == self.settings.cf_names()
        self.eci = dict(zip(self.names, z["eci"].tolist()))
        self.eci_run = dict(self.eci)
        for key in sorted(m["mu"]):
            self.eci_run[key] = self.eci_run[key] - m["mu"][key]
        self.occ0 = z["occ0"].astype(np.int8)
        self.cf0 = dict(zip(self.names, z["cf0"].tolist()))
        self.accept_bits = z["accept_bits"]
        self.e_trace = z["e_trace"]
        self.cf_checkpoints = z["cf_checkpoints"]
        self.occ_final = z["occ_final"].astype(np.int8)
        self.kind = m["kind"]
        self.seed = m["seed"]
        self.chain_uid = m["chain_uid"]
        self.n_moves = m["n_moves"]
        self.temperature = m["temperature"]
        self.trace_every = m["trace_every"]
        self.cf_every = m["cf_every"]
        self.n_accept = m["n_accept"]

    @property
    def n_sites(self):
        return self.settings.n_sites


def load_api_fixture():
    with open(os.path.join(GOLDEN, "updater_api_ternary.json")) as fh:
        meta = json.load(fh)
    bc = FCCSettings(meta["lattice"], meta["species"], {int(k): v for k, v in meta["shells"].items()},
                     symmetrize_decorations=meta["symmetrize"])
    bc.background_indices = list(meta["background"])
    return meta, bc
