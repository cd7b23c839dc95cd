"""Convert the reference's two HDF5 golden fixtures into JSON text fixtures.

Source files (data only, read through the `h5dump` tool; h5py is absent):
  tests/data/noise_golden.h5  -- written by tests/data/gen_golden.py
  tests/data/lbfgs_golden.h5  -- written by tests/data/gen_lbfgs_golden.py

Every value is printed by h5dump with "%.17g", which round-trips IEEE doubles
exactly, so the JSON files hold the fixtures bit for bit.

Run from the repository root (needs the reference checkout and h5dump):
    python tests/golden/convert_h5_fixtures.py
"""
import json
import os
import re
import shutil
import subprocess
import sys

REF = "/root/reference/tests/data"
OUT = os.path.dirname(os.path.abspath(__file__))
H5DUMP = shutil.which("h5dump") or "/opt/conda/bin/h5dump"


def _values(text):
    body = text.split("DATA {", 1)[1].split("}", 1)[0]
    body = re.sub(r"\(\d+(,\d+)*\):", " ", body)
    return [float(tok) for tok in body.replace(",", " ").split()]


def dump(path, obj, attr=False):
    flag = "-a" if attr else "-d"
    out = subprocess.run([H5DUMP, flag, obj, "-m", "%.17g", "-w", "0", path],
                         check=True, capture_output=True, text=True).stdout
    dims = re.search(r"DATASPACE\s+(SCALAR|SIMPLE \{ \(([^)]*)\))", out)
    vals = _values(out)
    if dims.group(1) == "SCALAR":
        return vals[0]
    shape = [int(s) for s in dims.group(2).split(",")]
    return {"shape": shape, "data": vals}


def main():
    noise = os.path.join(REF, "noise_golden.h5")
    lbfgs = os.path.join(REF, "lbfgs_golden.h5")
    fx = {k: dump(noise, "/" + k, attr=True)
          for k in ("nx", "ny", "Lx", "Ly", "amplitude", "xi_x", "xi_y")}
    for k in ("field", "x", "y", "v", "dvdx", "dvdy"):
        fx[k] = dump(noise, "/" + k)
    fx["source"] = "reference tests/data/noise_golden.h5 (tests/data/gen_golden.py)"
    with open(os.path.join(OUT, "noise_golden.json"), "w") as f:
        json.dump(fx, f, indent=0)
    gx = {k: dump(lbfgs, "/" + k, attr=True)
          for k in ("nx", "ny", "Lx", "Ly", "line_tension", "e_min")}
    for k in ("field", "h0", "h_min"):
        gx[k] = dump(lbfgs, "/" + k)
    gx["source"] = "reference tests/data/lbfgs_golden.h5 (tests/data/gen_lbfgs_golden.py)"
    with open(os.path.join(OUT, "lbfgs_golden.json"), "w") as f:
        json.dump(gx, f, indent=0)
    print("wrote noise_golden.json, lbfgs_golden.json", file=sys.stderr)


if __name__ == "__main__":
    main()
