"""Checkpoints and resume (io.jl:1-27,81-101; simulation.jl:96-148).

The reference writes JLD2 (HDF5) files `step_XXXXXXXX.jld2`.  HDF5 is not available to this
package, so a checkpoint is a numpy `.npz` archive `step_XXXXXXXX.npz` with the same fields:
psi (the array in the reference's memory order: numpy shape (D, nz, ny, nx), C order, is
byte-identical to Julia's column-major (nx, ny, nz, D)), t, step, grid_n_points,
grid_box_size, atom_name, c0, c1, c_lhy, c_extra, zeeman_p, zeeman_q, c_dd, dt,
imaginary_time.  Files are read with allow_pickle=False.
"""
from __future__ import annotations

import dataclasses
import os
import re

import numpy as np

from .api import Workspace, run_simulation

_STEP_FILE = re.compile(r"^step_(\d{8})\.npz$")


def save_state(filename: str, ws: Workspace, inst: int = 0):
    """save_state (io.jl:1-27) of one instance."""
    inter = ws.interactions[inst]
    z = ws.zeeman[inst]
    p = ws.params[inst]
    np.savez(filename,
             psi=ws.get_psi(inst),
             t=np.float64(ws.t[inst]),
             step=np.int64(ws.step[inst]),
             grid_n_points=np.asarray(ws.grid.config.n_points, dtype=np.int64),
             grid_box_size=np.asarray(ws.grid.config.box_size, dtype=np.float64),
             atom_name=np.str_(getattr(getattr(ws, "atom", None), "name", "")),
             c0=np.float64(inter.c0), c1=np.float64(inter.c1), c_lhy=np.float64(inter.c_lhy),
             c_extra=np.asarray(inter.c_extra, dtype=np.float64),
             zeeman_p=np.float64(z.p), zeeman_q=np.float64(z.q),
             c_dd=np.float64(p.c_dd if ws.enable_ddi else 0.0),
             dt=np.float64(p.dt),
             imaginary_time=np.bool_(ws.sim_params.imaginary_time))


def load_state(filename: str) -> dict:
    """load_state (io.jl:81-101); absent optional fields take the reference's defaults."""
    with np.load(filename, allow_pickle=False) as d:
        def get(k, default):
            return d[k] if k in d.files else default
        return dict(psi=np.array(d["psi"]), t=float(d["t"]), step=int(d["step"]),
                    grid_n_points=tuple(int(v) for v in d["grid_n_points"]),
                    grid_box_size=tuple(float(v) for v in d["grid_box_size"]),
                    atom_name=str(d["atom_name"]),
                    c0=float(get("c0", np.nan)), c1=float(get("c1", np.nan)),
                    c_lhy=float(get("c_lhy", 0.0)),
                    c_extra=tuple(float(v) for v in get("c_extra", np.zeros(0))),
                    zeeman_p=float(get("zeeman_p", np.nan)),
                    zeeman_q=float(get("zeeman_q", np.nan)),
                    c_dd=float(get("c_dd", 0.0)), dt=float(get("dt", np.nan)),
                    imaginary_time=bool(get("imaginary_time", False)))


def checkpoint_name(checkpoint_dir: str, step: int) -> str:
    return os.path.join(checkpoint_dir, f"step_{step:08d}.npz")


def latest_checkpoint(checkpoint_dir: str):
    """The checkpoint of the highest step in the directory (simulation.jl:105-110), or None."""
    if not os.path.isdir(checkpoint_dir):
        return None
    names = sorted(f for f in os.listdir(checkpoint_dir) if _STEP_FILE.match(f))
    return os.path.join(checkpoint_dir, names[-1]) if names else None


def run_simulation_checkpointed(ws: Workspace, *, checkpoint_dir="checkpoints",
                                checkpoint_every=1000, callback=None, resume=False,
                                keep_snapshots=True):
    """run_simulation_checkpointed! (simulation.jl:96-148): run_simulation! over the steps
    that remain, writing a checkpoint at every save point whose global step is a multiple of
    checkpoint_every; resume=True first restores psi, t and step from the latest checkpoint."""
    os.makedirs(checkpoint_dir, exist_ok=True)
    if resume:
        latest = latest_checkpoint(checkpoint_dir)
        if latest is not None:
            data = load_state(latest)
            ws.set_psi(data["psi"], 0)
            ws.t[0] = data["t"]
            ws.step[0] = data["step"]
    start = int(ws.step[0])
    sp = ws.sim_params
    remaining = sp.n_steps - start
    if remaining <= 0:
        return run_simulation(ws, callback=callback, keep_snapshots=keep_snapshots)

    def cb(w, step):
        g = start + step
        if g % checkpoint_every == 0:
            save_state(checkpoint_name(checkpoint_dir, g), w)
        if callback is not None:
            callback(w, step)

    ws.sim_params = dataclasses.replace(sp, n_steps=remaining)
    try:
        res = run_simulation(ws, callback=cb, keep_snapshots=keep_snapshots)
    finally:
        ws.sim_params = sp
    ws.step[0] = start + remaining
    return res
