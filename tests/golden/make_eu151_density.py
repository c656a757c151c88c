"""Extract the data of the reference's golden ground state
examples/eu151/density_64.json (3-D F=6 64^3 find_ground_state output,
examples/eu151/density_64.jl:11-59, eu151_setup.jl:4-24) into a small fixture.
Run in the build container only (reads /root/reference)."""
import json
import os

SRC = "/root/reference/examples/eu151/density_64.json"
DST = os.path.join(os.path.dirname(os.path.abspath(__file__)), "eu151_density_64.json")

d = json.load(open(SRC))
keep = {k: d[k] for k in ("N", "box", "n_peak", "comp_pops", "n_x", "n_y", "n_z", "x",
                          "slice_xy", "slice_xy_shape")}
json.dump(keep, open(DST, "w"))
print("wrote", DST)
