"""Checkpoint / resume in the reference's directory format (stonewall3d/lbm3d.py:320-355).

<case>.ckpt/ holds
  f.npy         populations, reference layout (19, nx, ny, nz); float32 for the FP32 form
                (the reference's format), float64 for FP64, raw uint16 bits for FP16 / BF16
  force.npy     (n_total, nz, 2) float32 wall force per step and height, as the reference
                writes it (its rows are FP32 atomic sums)
  force64.npy   the same rows in float64 (this package's FP64 reduction), so that a resume
                continues the exact record; the reference ignores the file, and a directory
                without it (one the reference wrote) resumes from force.npy
  uacc.npy      (4, nx, ny, nz) float32 velocity accumulator
  snaps_xz.npy, snaps_xy.npy, snaps_q.npy   snapshot stacks
  state.json    {"t": next step, "n_acc": samples in uacc, "sig": run signature}
A save writes <case>.ckpt.tmp/, then swaps it in with os.replace, so an interrupted save
leaves the previous checkpoint intact (lbm3d.py:352-354).  Resume refuses a checkpoint
whose signature differs from the run's (lbm3d.py:326-330).  For the reference's settings
(FP32 storage, periodic sides, steady inlet) the signature has exactly the reference's keys,
so each side resumes the other's checkpoints; other settings add keys (lbm3d.run_signature).
"""
import json
import os
import shutil

import numpy as np


class SignatureMismatch(SystemExit):
    pass


def save(sim, path, t_next, n_acc, sig, snaps=None):
    """Write sim's state at step t_next into the checkpoint directory `path`."""
    tmp = path + ".tmp"
    if os.path.isdir(tmp):
        shutil.rmtree(tmp)
    os.makedirs(tmp)
    np.save(os.path.join(tmp, "f.npy"), sim.populations())
    force = np.zeros((sim.n_steps, sim.nz, 2), np.float64)
    rows = sim.force_rows()
    force[:rows.shape[0]] = rows
    np.save(os.path.join(tmp, "force.npy"), force.astype(np.float32))
    np.save(os.path.join(tmp, "force64.npy"), force)
    np.save(os.path.join(tmp, "uacc.npy"), sim.accumulator())
    snaps = snaps or {}
    for k in ("snaps_xz", "snaps_xy", "snaps_q"):
        np.save(os.path.join(tmp, f"{k}.npy"), np.asarray(snaps.get(k, np.zeros((0,), np.float16))))
    with open(os.path.join(tmp, "state.json"), "w") as f:
        json.dump(dict(t=int(t_next), n_acc=int(n_acc), sig=sig), f)
    if os.path.isdir(path):
        shutil.rmtree(path)
    os.replace(tmp, path)


def read_state(path):
    with open(os.path.join(path, "state.json")) as f:
        return json.load(f)


def load(sim, path, sig):
    """Restore sim from `path`; returns (t, n_acc, snaps).  Raises SignatureMismatch if the
    checkpoint was made with different settings."""
    info = read_state(path)
    if info["sig"] != sig:
        raise SignatureMismatch(f"checkpoint {path} was made with different settings; use the same options, "
                                f"or --no-resume to start over")
    sim.load_populations(np.load(os.path.join(path, "f.npy")))
    t = int(info["t"])
    f64 = os.path.join(path, "force64.npy")
    force = np.load(f64 if os.path.exists(f64) else os.path.join(path, "force.npy"))
    sim.set_force_rows(force[:t].astype(np.float64))
    sim.set_accumulator(np.load(os.path.join(path, "uacc.npy")), int(info["n_acc"]))
    sim.t = t
    snaps = {k: np.load(os.path.join(path, f"{k}.npy")) for k in ("snaps_xz", "snaps_xy", "snaps_q")}
    return t, int(info["n_acc"]), snaps
