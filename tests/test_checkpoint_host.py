"""Checkpoint files (becgpu.io): naming, discovery of the latest step, loading with the
reference's defaults for absent fields (io.jl:81-101).  No device calls."""
import os

import numpy as np

from becgpu.io import checkpoint_name, latest_checkpoint, load_state


def test_latest_checkpoint_and_defaults(tmp_path):
    d = str(tmp_path)
    assert latest_checkpoint(d) is None
    psi = (np.arange(3 * 16) + 1j).reshape(3, 16)
    for step in (50, 1000, 150):
        np.savez(checkpoint_name(d, step), psi=psi * step, t=np.float64(step * 0.01),
                 step=np.int64(step), grid_n_points=np.array([16]), grid_box_size=np.array([10.0]),
                 atom_name=np.str_("Rb87"))
    open(os.path.join(d, "notes.txt"), "w").close()
    open(os.path.join(d, "step_12.npz"), "w").close()            # not an 8-digit step file
    latest = latest_checkpoint(d)
    assert os.path.basename(latest) == "step_00001000.npz"
    s = load_state(latest)
    assert s["step"] == 1000 and s["t"] == 10.0 and s["atom_name"] == "Rb87"
    assert np.array_equal(s["psi"], psi * 1000)
    assert s["grid_n_points"] == (16,) and s["grid_box_size"] == (10.0,)
    assert np.isnan(s["c0"]) and np.isnan(s["zeeman_p"]) and np.isnan(s["dt"])
    assert s["c_lhy"] == 0.0 and s["c_dd"] == 0.0 and s["c_extra"] == () and s["imaginary_time"] is False
