"""Regenerate the committed workload asset descriptions from the reference's asset files.

Reads genesis/assets/xml/franka_emika_panda/panda.xml and genesis/assets/urdf/go2/urdf/go2.urdf from a reference
checkout (default /root/reference) with the loaders of this package and writes panda.json / go2.json next to this
file. The JSON files are data (link tree, inertials, joint limits and gains, collision capsules); the step never
reads the original asset files at run time.
"""
import json
import os
import sys

from genesis_amd.loaders import entity_to_json, parse_mjcf, parse_urdf

HERE = os.path.dirname(os.path.abspath(__file__))


def main(ref_root="/root/reference"):
    assets = os.path.join(ref_root, "genesis", "assets")
    panda = parse_mjcf(os.path.join(assets, "xml", "franka_emika_panda", "panda.xml"))
    go2 = parse_urdf(os.path.join(assets, "urdf", "go2", "urdf", "go2.urdf"))
    for name, ent in (("panda", panda), ("go2", go2)):
        with open(os.path.join(HERE, f"{name}.json"), "w") as f:
            json.dump(entity_to_json(ent), f, indent=1)
        print(name, len(ent.links), "links", sum(len(l.geoms) for l in ent.links), "geoms")


if __name__ == "__main__":
    main(*sys.argv[1:])
