"""Writes the golden fixtures of tests/golden/ (run from the repository root:
``python tests/golden/make_golden.py``).

  atoms5k_box37_avg.npz, atoms5k_box37_given.npz
      The 5,000-atom configuration (37 A box, H/C/O/N/S 50/30/15/4/1 %,
      seed 1, order 4, 64^3 cell, sigma 2 -> 128^3 FFT, 0.01 A^-1 bins
      requested, q <= 0.5): positions, element ids, the FP64 oracle's
      histogram H_b, counts C_b and output rows, once with the border-mean
      pad (padding::avg) and once with the explicit-water pad of SPC/E
      (padding::given, 0.334 e/A^3).
  reference_solvents_db.json
      The layout of the reference's solvent-profile database
      (pycusaxs/data/reference_solvents.db, read with sqlite3 in read-only
      mode) and, per stored profile, its key parameters, its profile_hash, its
      point count and first points: pins the profile store's format.
  ref_helpers.json
      Outputs of the reference's own host helpers compiled from its sources
      by oracle/build_ref.sh (oracle/_ref/ref_driver): B-spline weights and
      derivatives (BSpline.cpp), modulus tables (BSpmod.cpp), the box
      inverse (Cell.cpp; orthorhombic, and triclinic through the transposed
      box) and explicit-solvent counts (AtomCounter.cpp).  Kept
      as data so the pins hold where the reference tree is absent.

The oracle is the checker; these files pin it (and through it the engine) on
inputs that later changes cannot silently move."""
from __future__ import annotations

import json
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import saxs_oracle as O  # noqa: E402

ATOMS5K = dict(n_atoms=5000, L=37.0, seed=1, n=64, sigma=2.0, dq=0.01, qcut=0.5, order=4)
REF_DRIVER = os.path.join(ROOT, "oracle", "_ref", "ref_driver")

BSPLINE_W = (0.0, 0.125, 0.3, 0.5, 0.77, 0.999)
BSPMOD = ((4, 12, 30, 128), (6, 16, 21, 64))
CELLS = ((37.0, 37.0, 37.0), (37.0, 41.5, 100.0))
TRICLINIC = (((30.0, 0.0, 0.0), (7.5, 28.0, 0.0), (-5.0, 6.25, 33.0)),
             ((61.3, 0.0, 0.0), (-20.4, 57.7, 0.0), (-20.4, -28.9, 50.0)))
SOLVENT = (((37.0, 37.0, 37.0), (64, 64, 64), "SPCE", {}),
           ((100.0, 100.0, 100.0), (128, 128, 128), "TIP3P", {"Na": 150, "Cl": 150}),
           ((80.5, 61.25, 90.0), (96, 72, 108), "TIP4P", {"K": 7, "Cl": 9}))


def atoms5k_fixture(pad_given):
    c = ATOMS5K
    pos, syms = O.synthetic_system(c["n_atoms"], c["L"], seed=c["seed"])
    symbols, ids = O.element_ids(syms)
    box = np.diag([c["L"]] * 3)
    plan = O.make_plan((c["n"],) * 3, box, c["order"], c["sigma"], None, c["dq"], c["qcut"],
                       pad_given=pad_given)
    H, C = O.frame_histogram(plan, pos, ids, symbols, box)
    rows = np.array(O.saxs_rows(plan, H, C), dtype=np.float64)
    pads = np.array([(pad_given or {}).get(s, 0.0) for s in symbols], dtype=np.float64)
    return dict(positions=pos.astype(np.float32), element_ids=ids.astype(np.int8),
                symbols=np.array(symbols), H=H, C=C, rows=rows, pads=pads,
                box_len=np.float64(c["L"]), n=np.int32(c["n"]), sigma=np.float64(c["sigma"]),
                dq=np.float64(c["dq"]), qcut=np.float64(c["qcut"]), order=np.int32(c["order"]),
                given=np.bool_(pad_given is not None))


def ref(*args):
    out = subprocess.run([REF_DRIVER, *map(str, args)], capture_output=True, text=True, check=True)
    return out.stdout.strip().splitlines()


def ref_helpers():
    d = {"bspline": {}, "bspmod": [], "cell": [], "triclinic": [], "solvent": []}
    for p in (4, 6):
        rows = []
        for line in ref("bspline", p, *BSPLINE_W):
            th, dth = line.split("|")
            rows.append({"theta": [float(v) for v in th.split()],
                         "dtheta": [float(v) for v in dth.split()]})
        d["bspline"][str(p)] = {"w": list(BSPLINE_W), "values": rows}
    for p, nx, ny, nz in BSPMOD:
        mods = [[float(v) for v in line.split()] for line in ref("bspmod", p, nx, ny, nz)]
        d["bspmod"].append({"order": p, "sizes": [nx, ny, nz], "mod": mods})
    for lengths in CELLS:
        line = ref("cell", lengths[0], 0, 0, 0, lengths[1], 0, 0, 0, lengths[2])[0]
        d["cell"].append({"lengths": list(lengths), "oc": [float(v) for v in line.split()]})
    for B in TRICLINIC:
        # Cell.cpp's upper-triangular inverse applied to B^T (rows a, b, c -> columns)
        bt = np.asarray(B, dtype=np.float64).T.ravel()
        line = ref("cell", *bt)[0]
        d["triclinic"].append({"box": [list(r) for r in B], "inverse": [float(v) for v in line.split()]})
    for lengths, n, model, ions in SOLVENT:
        args = [*lengths, model, *n]
        for k, v in ions.items():
            args += [k, v]
        vals = {}
        for line in ref("atomcount", *args):
            sym, val = line.split()
            vals[sym] = float(val)
        d["solvent"].append({"lengths": list(lengths), "n": list(n), "model": model,
                             "ions": ions, "pads": vals})
    return d


REF_DB = "/root/reference/pycusaxs/data/reference_solvents.db"


def reference_db_fixture():
    """The reference's solvent-profile database, opened read-only: its table
    layout (PRAGMA table_info), and per row the key parameters, the stored
    profile_hash, the point count and the first ten (q, I) points."""
    import sqlite3
    con = sqlite3.connect(f"file:{REF_DB}?mode=ro", uri=True)
    cols = [list(r) for r in con.execute("PRAGMA table_info(saxs_profiles)")]
    rows = []
    for r in con.execute("SELECT profile_hash, water_model, n_water_molecules, ion_counts, "
                         "other_molecules, box_x, box_y, box_z, box_volume, supercell_scale, "
                         "supercell_volume, simulation_time_ps, n_frames_analyzed, frame_stride, "
                         "grid_size, bin_size, qcut, spline_order, density_g_cm3, n_atoms, "
                         "profile_data FROM saxs_profiles ORDER BY id"):
        data = json.loads(r[20])
        rows.append({"profile_hash": r[0], "water_model": r[1], "n_water_molecules": r[2],
                     "ion_counts": json.loads(r[3]), "other_molecules": json.loads(r[4] or "{}"),
                     "box": [r[5], r[6], r[7]], "box_volume": r[8], "supercell_scale": r[9],
                     "supercell_volume": r[10], "simulation_time_ps": r[11],
                     "n_frames_analyzed": r[12], "frame_stride": r[13],
                     "grid_size": json.loads(r[14]), "bin_size": r[15], "qcut": r[16],
                     "spline_order": r[17], "density_g_cm3": r[18], "n_atoms": r[19],
                     "n_points": len(data), "first_points": data[:10]})
    con.close()
    return {"columns": cols, "rows": rows}


def main():
    np.savez_compressed(os.path.join(HERE, "atoms5k_box37_avg.npz"), **atoms5k_fixture(None))
    given = O.solvent_pads((ATOMS5K["L"],) * 3, (ATOMS5K["n"],) * 3, "SPCE", {})
    np.savez_compressed(os.path.join(HERE, "atoms5k_box37_given.npz"), **atoms5k_fixture(given))
    if os.path.exists(REF_DB):
        with open(os.path.join(HERE, "reference_solvents_db.json"), "w") as fh:
            json.dump(reference_db_fixture(), fh, indent=1)
    if os.path.exists(REF_DRIVER):
        with open(os.path.join(HERE, "ref_helpers.json"), "w") as fh:
            json.dump(ref_helpers(), fh, indent=1)
    else:
        print("oracle/_ref/ref_driver not built (oracle/build_ref.sh); ref_helpers.json kept")
    print("golden fixtures written to", HERE)


if __name__ == "__main__":
    main()
