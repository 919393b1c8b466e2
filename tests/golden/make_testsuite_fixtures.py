"""Extract the reference test suite's data into small committed fixtures.

Run once in a container that has the reference checkout and an h5py
(conda's python3.9 here):

    /opt/conda/bin/python3.9 tests/golden/make_testsuite_fixtures.py /root/reference

Writes, next to this script:
  * testsuite_br_<name>.npz   -- each deck's Br input map (dataset 'Data'
                                 with its 'dim1' (theta) / 'dim2' (phi) scales),
                                 as float64 arrays in (theta, phi) order;
  * testsuite_reference.json  -- per deck: the namelist values the driver
                                 needs, the energies (ES22.16), the CG
                                 iteration count, the residual history, the
                                 printed r/t/p mesh tables, and readbr's
                                 printed flux balance (positive and negative
                                 flux, BR00 when the flux was balanced).
Only data is copied: inputs and the outputs the reference recorded.
"""
import json
import os
import re
import sys

import h5py
import numpy as np

DECKS = ["potential_field_source_surface", "potential_field_current_sheet", "open_field"]


def parse_namelist(text):
    vals = {}
    for line in text.splitlines():
        line = line.split("!")[0].strip()
        m = re.match(r"^(\w+)\s*=\s*(.+)$", line)
        if not m:
            continue
        key, v = m.group(1).lower(), m.group(2).strip().rstrip(",")
        if v.startswith(("'", '"')):
            vals[key] = v.strip("'\"")
        elif v.lower() in (".true.", ".false."):
            vals[key] = v.lower() == ".true."
        else:
            parts = [p.strip() for p in v.split(",") if p.strip()]
            nums = [float(p.lower().replace("d", "e")) for p in parts]
            vals[key] = nums if len(nums) > 1 else nums[0]
    return vals


def parse_out(text):
    res = {}
    e = {}
    for key, pat in (("br2", r"Energy in Br\*\*2 =\s*(\S+)"), ("bt2", r"Energy in Bt\*\*2 =\s*(\S+)"),
                     ("bp2", r"Energy in Bp\*\*2 =\s*(\S+)"), ("total", r"Magnetic energy =\s*(\S+)")):
        e[key] = float(re.search(pat, text).group(1))
    res["energy"] = e
    res["iterations"] = int(re.search(r"Number of iterations =\s*(\d+)", text).group(1))
    hist = [(int(a), float(b)) for a, b in re.findall(r"Iteration:\s+(\d+)\s+Residual:\s+(\S+)", text)]
    res["residual_history"] = hist
    meshes = {}
    for blk in re.split(r"### Mesh information for coordinate ", text)[1:]:
        coord = blk[0]
        body = blk.split("Mesh-point locations:")[1]
        rows = []
        for line in body.splitlines()[2:]:
            parts = line.split()
            if not parts or not parts[0].isdigit():
                if rows:
                    break
                continue
            rows.append([int(parts[0]), float(parts[1])])
        meshes[coord] = rows
    res["mesh"] = meshes
    return res


def parse_log(text):
    """readbr's flux prints (pot3d.log): Positive/Negative flux and, when
    the flux was balanced, BR00."""
    res = {"fluxp": float(re.search(r"Positive flux =\s*(\S+)", text).group(1)),
           "fluxm": float(re.search(r"Negative flux =\s*(\S+)", text).group(1))}
    m = re.search(r"BR00 \(monopole Br field magnitude\) =\s*(\S+)", text)
    res["br00err"] = float(m.group(1)) if m else None
    return res


def main(ref):
    here = os.path.dirname(os.path.abspath(__file__))
    out = {}
    for deck in DECKS:
        base = os.path.join(ref, "testsuite", deck)
        nl = parse_namelist(open(os.path.join(base, "input", "pot3d.dat")).read())
        with h5py.File(os.path.join(base, "input", nl["br0file"]), "r") as h:
            data = np.asarray(h["Data"][()], dtype=np.float64).T  # Fortran (theta,phi)
            tn = np.asarray(h["dim1"][()], dtype=np.float64)
            pn = np.asarray(h["dim2"][()], dtype=np.float64)
        np.savez_compressed(os.path.join(here, f"testsuite_br_{deck}.npz"), data=data, tn=tn, pn=pn)
        rec = parse_out(open(os.path.join(base, "reference", "pot3d.out")).read())
        rec["namelist"] = nl
        rec["readbr"] = parse_log(open(os.path.join(base, "reference", "pot3d.log")).read())
        out[deck] = rec
    with open(os.path.join(here, "testsuite_reference.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
