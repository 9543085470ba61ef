"""Re-run the reference's published H = 40 study through this package's CLI and compare.

The reference publishes, for its three H = 40 cases (solid, 20 % and 40 % porosity; grid
800 x 160 x 321, Re 15000, seed 7, 20 H/U spin-up + 30 H/U averaging), the force
statistics of postprocess.py:65-82 (results/3d/stone_wall_3d.html:150, `cases[].stats`)
and the recirculation area of export3d.py:169-184 (results/3d/stone_wall_scene3d.html:142).
This runs the same cases with `stonewall3d.lbm3d.main` (the reference's flags, no
checkpoints), recomputes the same statistics from the written .npz and prints, per case,
one JSON line with ours, the published value and the relative difference.

The init noise differs from the reference's (Warp's RNG cannot be reproduced), so the
runs are different realisations of the same flow: agreement is statistical.

    python -m stonewall3d.validate --published ../tests/golden/reference_results.json \\
        --out /tmp/refcases [--dtype f16] [--cases solid,p20,p40]
"""
import argparse
import json
import os
import time

import numpy as np

from . import lbm3d, stats

KEYS = ("Cd", "CdMax", "CdStd", "Cm", "CmMax", "Cl")


def published_cases(path):
    """{case: dict(porosity_target, Cd, ..., area)} from the published-results fixture."""
    with open(path) as f:
        d = json.load(f)
    out = {}
    for c in d["cases"]:
        ref = {k: float(c["stats"][k]) for k in KEYS}
        ref["area"] = float(c["reverse_area"])
        ref["porosity_target"] = float(c["porosity_target"])
        out[c["key"]] = ref
    return out, d["info"]


def run_case(case, ref, H, out, dtype, keep=False):
    t0 = time.time()
    lbm3d.main(["--case", case, "--porosity", str(ref["porosity_target"]), "--H", str(H), "--seed", "7",
                "--out", out, "--ckpt-every", "0", "--no-resume", "--dtype", dtype])
    path = os.path.join(out, f"{case}.npz")
    d = np.load(path)
    m = json.loads(str(d["meta"]))
    st = stats.force_stats(d["Cd"], d["Cm"], d["Cl"], m["n_spin"])
    st["area"] = stats.recirculation_area_ds2(d["mean3d_ds2"][0], d["wall_ds2"], H, m["X0"], m["UH"], m["ny"])
    if not keep:
        os.remove(path)
    rel = {k: (st[k] - ref[k]) / abs(ref[k]) for k in st if ref.get(k)}
    return dict(case=case, dtype=dtype, seconds=round(time.time() - t0, 1), porosity=m.get("porosity"),
                ours={k: round(v, 4) for k, v in st.items()}, published={k: ref[k] for k in st},
                rel_diff={k: round(v, 4) for k, v in rel.items()})


def main(argv=None):
    ap = argparse.ArgumentParser(description="re-run the reference's published H=40 cases and compare statistics")
    ap.add_argument("--published", required=True, help="JSON of the reference's published results")
    ap.add_argument("--out", required=True)
    ap.add_argument("--dtype", default="f32", choices=["f32", "f16", "bf16", "f64"])
    ap.add_argument("--cases", default="solid,p20,p40")
    ap.add_argument("--keep", action="store_true", help="keep the .npz files")
    a = ap.parse_args(argv)
    refs, info = published_cases(a.published)
    os.makedirs(a.out, exist_ok=True)
    for case in a.cases.split(","):
        print(json.dumps(run_case(case, refs[case], int(info["H"]), a.out, a.dtype, a.keep)), flush=True)


if __name__ == "__main__":
    main()
