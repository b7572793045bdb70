"""Save / load a World (pdpd_amd.world) to a self-contained .npz fixture.

Plain arrays plus one JSON string for the polymer types; loaded with
numpy.load(allow_pickle=False)."""
import json

import numpy as np

from pdpd_amd.world import PolymerType, World

ARRAYS = ("box", "conservative", "dissipative", "polymer_type_of", "polymer_bead_start", "polymer_bead_ids",
          "x", "v", "f", "bead_type", "is_monomer", "polymer_id", "polymer_offset")


def world_to_arrays(w, prefix="w_"):
    out = {prefix + k: np.asarray(getattr(w, k)) for k in ARRAYS}
    if w.stationary is not None:
        out[prefix + "stationary"] = np.asarray(w.stationary, np.uint8)
    meta = dict(dt=w.dt, lambda_=w.lambda_, seed=int(w.seed), t=int(w.t), bead_type_names=w.bead_type_names,
                polymer_types=[dict(name=p.name, bead_types=list(map(int, p.bead_types)),
                                    bonds=[list(b) for b in p.bonds], bond_pairs=[list(b) for b in p.bond_pairs])
                               for p in w.polymer_types])
    out[prefix + "meta"] = np.array(json.dumps(meta))
    return out


def world_from_arrays(z, prefix="w_"):
    meta = json.loads(str(z[prefix + "meta"]))
    kw = {k: np.array(z[prefix + k]) for k in ARRAYS}
    pts = [PolymerType(p["name"], p["bead_types"], [tuple(b) for b in p["bonds"]],
                       [tuple(b) for b in p["bond_pairs"]]) for p in meta["polymer_types"]]
    stat = np.array(z[prefix + "stationary"]) if prefix + "stationary" in z else None
    return World(dt=meta["dt"], lambda_=meta["lambda_"], seed=meta["seed"], t=meta["t"],
                 bead_type_names=meta["bead_type_names"], polymer_types=pts, stationary=stat, **kw)
