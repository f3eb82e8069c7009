"""Profile store and solvent subtraction (pycusaxs_amd.profiles), CPU only.

Pins: the table layout and every stored profile_hash of the reference's own
solvent-profile database (tests/golden/reference_solvents_db.json, written
by tests/golden/make_golden.py); a database written here and read back;
the subtraction and resampling arithmetic against closed forms (linear data
is reproduced exactly by both interpolations, cubic data by the cubic
spline); the command-line tool end to end."""
import json
import os
import sqlite3

import numpy as np
import pytest

from pycusaxs_amd import profiles as P

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def refdb():
    with open(os.path.join(GOLDEN, "reference_solvents_db.json")) as fh:
        return json.load(fh)


def _save(db, row, points, **over):
    kw = dict(water_model=row["water_model"], n_water_molecules=row["n_water_molecules"],
              ion_counts=row["ion_counts"], box_x=row["box"][0], box_y=row["box"][1],
              box_z=row["box"][2], box_volume=row["box_volume"],
              supercell_scale=row["supercell_scale"], supercell_volume=row["supercell_volume"],
              simulation_time_ps=row["simulation_time_ps"],
              n_frames_analyzed=row["n_frames_analyzed"], grid_size=row["grid_size"],
              frame_stride=row["frame_stride"], bin_size=row["bin_size"], qcut=row["qcut"],
              order=row["spline_order"], density_g_cm3=row["density_g_cm3"],
              n_atoms=row["n_atoms"], other_molecules=row["other_molecules"],
              notes="test")
    kw.update(over)
    return db.save_profile(points, **kw)


def test_profile_hash_reproduces_reference_database(refdb):
    for row in refdb["rows"]:
        h = P.profile_hash(row["water_model"], row["ion_counts"], row["box"],
                           row["supercell_scale"], row["grid_size"], row["simulation_time_ps"])
        assert h == row["profile_hash"], row


def test_table_layout_matches_reference(refdb, tmp_path):
    db = P.SaxsDatabase(tmp_path / "a.db")
    cols = [list(r) for r in db.conn.execute("PRAGMA table_info(saxs_profiles)")]
    db.close()
    assert [c[1:] for c in cols] == [c[1:] for c in refdb["columns"]]
    assert tuple(c[1] for c in cols) == P.COLUMNS


def test_save_find_get_list_roundtrip(refdb, tmp_path):
    path = tmp_path / "b.db"
    with P.SaxsDatabase(path) as db:
        ids = [_save(db, row, row["first_points"]) for row in refdb["rows"]]
        assert ids == [1, 2, 3, 4, 5]
        # the same key keeps its row
        assert _save(db, refdb["rows"][2], [[0.1, 1.0]]) == 3
        r = refdb["rows"][2]
        got = db.find_profile(r["water_model"], r["ion_counts"], *r["box"], r["supercell_scale"],
                              r["grid_size"], r["simulation_time_ps"])
        assert got["id"] == 3 and got["profile_hash"] == r["profile_hash"]
        assert got["profile_data"] == r["first_points"]
        assert got["ion_counts"] == {"Na": 42, "Cl": 42} and got["grid_size"] == [128, 128, 128]
        assert db.get_profile(99) is None
        assert len(db.list_profiles()) == 5 and len(db.list_profiles("SPCE")) == 0
    # stored as the reference stores it: JSON text columns
    con = sqlite3.connect(str(path))
    ions, grid, data = con.execute("SELECT ion_counts, grid_size, profile_data FROM "
                                   "saxs_profiles WHERE id = 3").fetchone()
    con.close()
    assert json.loads(ions) == {"Na": 42, "Cl": 42}
    assert grid == "[128, 128, 128]"
    assert json.loads(data) == r["first_points"]


def test_export_csv(refdb, tmp_path):
    with P.SaxsDatabase(tmp_path / "c.db") as db:
        pid = _save(db, refdb["rows"][0], [[0.01, 2.5], [0.02, 1.25]])
        db.export_profile_csv(pid, tmp_path / "p.csv")
        with pytest.raises(ValueError):
            db.export_profile_csv(42, tmp_path / "q.csv")
    lines = (tmp_path / "p.csv").read_text().splitlines()
    assert lines[0] == "# SAXS Solvent Profile"
    assert lines[4] == "# Box Size: 90.000 x 90.000 x 90.000 Å"
    assert lines[-2:] == ["0.010000,2.500000e+00", "0.020000,1.250000e+00"]


def _prof(q, i, pid=1):
    return {"id": pid, "profile_data": np.stack([q, i], axis=1).tolist(), "water_model": "TIP3P",
            "ion_counts": {}, "other_molecules": {}, "grid_size": [64, 64, 64],
            "supercell_scale": 2.0, "box_x": 30.0, "box_y": 30.0, "box_z": 30.0,
            "box_volume": 27000.0, "density_g_cm3": None, "simulation_time_ps": 10.0}


def test_subtract_same_grid_is_exact():
    q = np.linspace(0.0, 1.0, 11)
    u, r = _prof(q, 3.0 + q), _prof(q, 1.0 + 2.0 * q)
    qq, out = P.subtract_profiles(u, r, 0.5)
    assert np.array_equal(qq, q)
    assert np.allclose(out, 3.0 + q - 0.5 * (1.0 + 2.0 * q), rtol=0, atol=1e-15)


@pytest.mark.parametrize("method", ["linear", "cubic"])
def test_subtract_interpolates_on_overlap_only(method):
    qu = np.linspace(0.05, 1.5, 30)
    qr = np.linspace(0.0, 1.0, 17)
    ref = 2.0 - 0.7 * qr if method == "linear" else 1.0 + qr - 0.3 * qr ** 2 + 0.1 * qr ** 3
    qq, out = P.subtract_profiles(_prof(qu, np.full_like(qu, 5.0)), _prof(qr, ref), 2.0, method)
    inside = qu <= 1.0
    exact = 2.0 - 0.7 * qu if method == "linear" else 1.0 + qu - 0.3 * qu ** 2 + 0.1 * qu ** 3
    assert np.allclose(out[inside], 5.0 - 2.0 * exact[inside], rtol=0, atol=1e-12)
    assert np.array_equal(out[~inside], np.full((~inside).sum(), 5.0))


def test_subtract_without_overlap_raises():
    with pytest.raises(ValueError, match="No overlap"):
        P.subtract_profiles(_prof(np.linspace(1.0, 2.0, 5), np.ones(5)),
                            _prof(np.linspace(0.0, 0.5, 5), np.ones(5)), 1.0)


def test_resample_uniform_grid():
    q = np.array([0.0, 0.1, 0.25, 0.4, 0.7])
    qq, ii = P.resample_profile(q, 3.0 * q + 1.0, 0.1, "linear")
    assert len(qq) == 8 and qq[0] == 0.0 and qq[-1] == 0.7      # ceil(0.7 / 0.1) + 1
    assert np.allclose(ii, 3.0 * qq + 1.0, rtol=0, atol=1e-14)


def test_command_line_tool(refdb, tmp_path):
    row = refdb["rows"][0]
    q = np.linspace(0.02, 1.0, 50)
    with P.SaxsDatabase(tmp_path / "user.db") as db:
        uid = _save(db, row, np.stack([q, 10.0 - q], axis=1).tolist())
    with P.SaxsDatabase(tmp_path / "ref.db") as db:
        rid = _save(db, row, np.stack([q, 4.0 + 0 * q], axis=1).tolist(), notes="ref")
    out = tmp_path / "sub.dat"
    rc = P.main(["--db", str(tmp_path / "user.db"), "--id", str(uid), "--ref-db",
                 str(tmp_path / "ref.db"), "--ref-id", str(rid), "--scale", "0.5", "-o", str(out)])
    assert rc == 0
    lines = out.read_text().splitlines()
    assert lines[0] == "# SAXS Profile after Solvent Subtraction"
    assert "#   Scaling Factor: 0.500000" in lines
    vals = np.array([[float(v) for v in l.split()] for l in lines if not l.startswith("#")])
    assert np.allclose(vals[:, 0], np.round(q, 6)) and np.allclose(vals[:, 1], 8.0 - q, rtol=1e-6)
    assert P.main(["--db", str(tmp_path / "user.db"), "--id", "7", "--ref-db",
                   str(tmp_path / "ref.db"), "--ref-id", str(rid), "--scale", "1"]) == 1
