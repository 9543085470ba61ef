"""Extract the reference's published result data into small JSON fixtures.

Reads the JSON payloads embedded in the reference's result viewers (data files, not
code): results/3d/stone_wall_3d.html:150 and results/3d/stone_wall_scene3d.html:142.
Writes tests/golden/reference_results.json with, per case (solid / p20 / p40 at H=40,
seed 7, grid 800x160x321): the front silhouette bits (packbits of
build_wall(...).any(axis=0)[:, :50], transposed and flipped, gzip + base64), porosity,
top-line heights, the force statistics, the smoothed Cd series and the recirculation
area.  Run in a container that holds the reference:  python tests/golden/make_reference_fixtures.py
"""
import json
import os
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_results.json")


def payload(path, line):
    s = open(path, encoding="utf-8").read().split("\n")[line - 1]
    return json.loads(s[s.index("{"):s.rindex("}") + 1])


def main():
    a = payload(os.path.join(REF, "results/3d/stone_wall_3d.html"), 150)
    b = payload(os.path.join(REF, "results/3d/stone_wall_scene3d.html"), 142)
    scene = {c["key"]: c for c in b["cases"]}
    out = {"source": {"viewer": "results/3d/stone_wall_3d.html:150", "scene": "results/3d/stone_wall_scene3d.html:142"},
           "info": a["info"], "cases": []}
    for c in a["cases"]:
        sc = scene[c["key"]]
        out["cases"].append(dict(
            key=c["key"], porosity=c["porosity"], porosity_target=sc.get("porosityTarget"),
            topMean=c["topMean"], topMax=c["topMax"], stats=c["stats"], scene_stats=sc["stats"],
            cd=c["cd"], cdUnits=c["cdUnits"], front=c["front"],
            reverse_area=sc["reverse"]["area"] if sc.get("reverse") else None,
            vortex_dims=sc["vortex"].get("dims") if sc.get("vortex") else None))
    with open(OUT, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", OUT, os.path.getsize(OUT), "bytes")


if __name__ == "__main__":
    main()
