"""Extract input fixtures from the reference's shipped data files.

The reference ships HDF5 inputs but no tests; this script turns one
geometry (createdgeometries/G1/G1_nonP_0.h5, 84^3 voxels at 12 nm), two
species files (parameters/nonP_blanket/parameters_sr1100_ryr55.h5,
parameters/phosphorylation/parameters_sr900_ryr55_pryr25.h5) and G1's
phosphorylation masks (createdgeometries/G1/G1_*.npy, 168x168 pixels of
0 / 1 RyR / 2 pRyR, stored as the pixel coordinates of each value) into
plain data (npz + json) so tests can use them where the reference tree is
absent.  Run in the build container:  python tests/golden/make_fixtures.py
"""
import dataclasses
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "hpc-ca-phospattern_amd"))
import csparks as cs  # noqa: E402

REF = "/root/reference"
GEOMS = {"g1_nonP_0": "createdgeometries/G1/G1_nonP_0.h5"}
MASKS = ["G1_nonP_0", "G1_inner_10", "G1_inner_25", "G1_outer_10", "G1_outer_25", "G1_uniform_10",
         "G1_uniform_25"]
MODELS = {"params_sr1100_ryr55": "parameters/nonP_blanket/parameters_sr1100_ryr55.h5",
          "params_sr900_ryr55_pryr25": "parameters/phosphorylation/parameters_sr900_ryr55_pryr25.h5"}


def main():
    for name, rel in GEOMS.items():
        g = cs.read_geometry(os.path.join(REF, rel))
        arrays = {"voxels": g.voxels, "domain_connections": g.domain_connections}
        meta = {"h": g.h, "global_size": list(g.global_size), "domain_names": g.domain_names,
                "boundaries": [{"name": b.name, "type": b.type} for b in g.boundaries], "source": rel}
        for b in g.boundaries:
            arrays["boundary_" + b.name] = b.pairs
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **arrays)
        with open(os.path.join(HERE, name + ".json"), "w") as f:
            json.dump(meta, f, indent=1)
    for name, rel in MODELS.items():
        m = cs.read_species(os.path.join(REF, rel))
        d = dataclasses.asdict(m)
        d["init"] = np.asarray(m.init).tolist()
        d["source"] = rel
        with open(os.path.join(HERE, name + ".json"), "w") as f:
            json.dump(d, f, indent=1)
    masks = {}
    for name in MASKS:
        # numpy files read without pickles (allow_pickle stays False)
        m = np.load(os.path.join(REF, "createdgeometries/G1", name + ".npy"))
        masks[name] = {"shape": list(m.shape),
                       "ryr": np.argwhere(m == 1).tolist(), "p_ryr": np.argwhere(m == 2).tolist(),
                       "source": f"createdgeometries/G1/{name}.npy"}
    with open(os.path.join(HERE, "g1_phospho_masks.json"), "w") as f:
        json.dump(masks, f)


if __name__ == "__main__":
    main()
