"""SAXS profile store and solvent subtraction: the on-disk formats that the
tools downstream of the scattering path consume.

``SaxsDatabase`` keeps the reference's SQLite layout (one ``saxs_profiles``
table, saxs_database.py:42-101), its profile key (SHA-256 of a canonical
string of water model, sorted ion counts, box, supercell scale, grid and
analysed time: saxs_database.py:103-127), its JSON-encoded fields and its CSV
export (:312-350), so databases written here and by the reference can be
read by either.  ``subtract_profiles`` and ``resample_profile`` restate the
arithmetic of the subtraction tool (saxs_subtract.py:126-248): the scaled
reference is subtracted on the overlap of the two q ranges only, after it is
interpolated onto the user's q grid when the grids differ (a cubic or linear
spline through the reference points, extrapolated at the ends); a profile can
be resampled onto a uniform q grid.  ``main`` is the tool's command line
without its interactive prompts (saxs_subtract.py:375-538).
"""
from __future__ import annotations

import argparse
import hashlib
import json
import sqlite3
import sys
from pathlib import Path
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

# Column order of the reference table (saxs_database.py:47-98); the CREATE
# statement below declares the same columns, types and constraints.
COLUMNS = ("id", "profile_hash", "water_model", "n_water_molecules", "ion_counts",
           "other_molecules", "box_x", "box_y", "box_z", "box_volume", "supercell_scale",
           "supercell_volume", "simulation_time_ps", "n_frames_analyzed", "frame_stride",
           "grid_size", "bin_size", "qcut", "spline_order", "density_g_cm3", "n_atoms",
           "created_timestamp", "notes", "profile_data")

_SCHEMA = """
CREATE TABLE IF NOT EXISTS saxs_profiles (
    id INTEGER PRIMARY KEY AUTOINCREMENT,
    profile_hash TEXT UNIQUE NOT NULL,
    water_model TEXT NOT NULL,
    n_water_molecules INTEGER NOT NULL,
    ion_counts TEXT NOT NULL,
    other_molecules TEXT,
    box_x REAL NOT NULL,
    box_y REAL NOT NULL,
    box_z REAL NOT NULL,
    box_volume REAL NOT NULL,
    supercell_scale REAL NOT NULL,
    supercell_volume REAL NOT NULL,
    simulation_time_ps REAL NOT NULL,
    n_frames_analyzed INTEGER NOT NULL,
    frame_stride INTEGER,
    grid_size TEXT NOT NULL,
    bin_size REAL,
    qcut REAL,
    spline_order INTEGER,
    density_g_cm3 REAL,
    n_atoms INTEGER,
    created_timestamp TEXT DEFAULT CURRENT_TIMESTAMP,
    notes TEXT,
    profile_data TEXT NOT NULL
)
"""
_INDEXES = ("CREATE INDEX IF NOT EXISTS idx_profile_hash ON saxs_profiles(profile_hash)",
            "CREATE INDEX IF NOT EXISTS idx_water_model ON saxs_profiles(water_model)")


def profile_key(water_model: str, ion_counts: Dict[str, int], box: Sequence[float],
                supercell_scale: float, grid_size: Sequence[int],
                simulation_time_ps: float) -> str:
    """The identifying string of a solvent profile; its SHA-256 is the
    table's ``profile_hash`` (saxs_database.py:118-127)."""
    ions = sorted(ion_counts.items())
    return (f"{water_model}|{ions}|{box[0]:.3f},{box[1]:.3f},{box[2]:.3f}|"
            f"{supercell_scale:.4f}|{grid_size[0]}x{grid_size[1]}x{grid_size[2]}|"
            f"{simulation_time_ps:.2f}")


def profile_hash(*args, **kwargs) -> str:
    return hashlib.sha256(profile_key(*args, **kwargs).encode()).hexdigest()


class SaxsDatabase:
    """SQLite store of SAXS profiles (by default in the file
    ``saxs_profiles.db``), one row per (solvent, box, grid, analysed time)."""

    def __init__(self, db_path="saxs_profiles.db", read_only: bool = False):
        self.db_path = Path(db_path)
        if read_only:
            self.conn = sqlite3.connect(f"file:{self.db_path}?mode=ro", uri=True)
        else:
            self.conn = sqlite3.connect(str(self.db_path))
        self.conn.row_factory = sqlite3.Row
        if not read_only:
            with self.conn:
                self.conn.execute(_SCHEMA)
                for sql in _INDEXES:
                    self.conn.execute(sql)

    # -- writing --------------------------------------------------------------
    def save_profile(self, profile_data, water_model: str, n_water_molecules: int,
                     ion_counts: Dict[str, int], box_x: float, box_y: float, box_z: float,
                     box_volume: float, supercell_scale: float, supercell_volume: float,
                     simulation_time_ps: float, n_frames_analyzed: int,
                     grid_size: Sequence[int], frame_stride: int = 1,
                     bin_size: Optional[float] = None, qcut: Optional[float] = None,
                     order: Optional[int] = None, density_g_cm3: Optional[float] = None,
                     n_atoms: Optional[int] = None,
                     other_molecules: Optional[Dict[str, int]] = None,
                     notes: Optional[str] = None) -> int:
        """Stores a profile (a sequence of (q, I(q)) pairs) and returns its
        row id; a profile with the same key keeps its row and that id is
        returned (saxs_database.py:179-221)."""
        h = profile_hash(water_model, ion_counts, (box_x, box_y, box_z), supercell_scale,
                         grid_size, simulation_time_ps)
        rows = [[float(q), float(i)] for q, i in profile_data]
        values = (h, water_model, int(n_water_molecules), json.dumps(ion_counts),
                  json.dumps(other_molecules or {}), box_x, box_y, box_z, box_volume,
                  supercell_scale, supercell_volume, simulation_time_ps, int(n_frames_analyzed),
                  frame_stride, json.dumps(list(grid_size)), bin_size, qcut, order,
                  density_g_cm3, n_atoms, notes, json.dumps(rows))
        try:
            with self.conn:
                cur = self.conn.execute(
                    f"INSERT INTO saxs_profiles ({', '.join(COLUMNS[1:21])}, notes, profile_data) "
                    f"VALUES ({', '.join('?' * 22)})", values)
            return int(cur.lastrowid)
        except sqlite3.IntegrityError:
            print(f"Warning: Profile with hash {h[:16]}... already exists in database")
            return int(self.conn.execute("SELECT id FROM saxs_profiles WHERE profile_hash = ?",
                                         (h,)).fetchone()[0])

    # -- reading --------------------------------------------------------------
    @staticmethod
    def _decode(row: sqlite3.Row) -> Dict:
        d = dict(row)
        d["ion_counts"] = json.loads(d["ion_counts"])
        d["other_molecules"] = json.loads(d["other_molecules"]) if d.get("other_molecules") else {}
        d["grid_size"] = json.loads(d["grid_size"])
        d["profile_data"] = json.loads(d["profile_data"])
        return d

    def find_profile(self, water_model: str, ion_counts: Dict[str, int], box_x: float,
                     box_y: float, box_z: float, supercell_scale: float,
                     grid_size: Sequence[int], simulation_time_ps: float,
                     tolerance: float = 0.01) -> Optional[Dict]:
        """The profile with this key, or None (the key fixes the box to 1e-3 A,
        so ``tolerance`` is accepted for interface parity and not used, as in
        saxs_database.py:223-257)."""
        h = profile_hash(water_model, ion_counts, (box_x, box_y, box_z), supercell_scale,
                         grid_size, simulation_time_ps)
        row = self.conn.execute("SELECT * FROM saxs_profiles WHERE profile_hash = ?",
                                (h,)).fetchone()
        return self._decode(row) if row else None

    def get_profile(self, profile_id: int) -> Optional[Dict]:
        row = self.conn.execute("SELECT * FROM saxs_profiles WHERE id = ?",
                                (int(profile_id),)).fetchone()
        return self._decode(row) if row else None

    def list_profiles(self, water_model: Optional[str] = None) -> List[Dict]:
        """All profiles, newest first, optionally of one water model."""
        if water_model:
            cur = self.conn.execute("SELECT * FROM saxs_profiles WHERE water_model = ? "
                                    "ORDER BY created_timestamp DESC", (water_model,))
        else:
            cur = self.conn.execute("SELECT * FROM saxs_profiles ORDER BY created_timestamp DESC")
        return [self._decode(r) for r in cur.fetchall()]

    def export_profile_csv(self, profile_id: int, output_path) -> None:
        """CSV with a commented metadata header (saxs_database.py:312-350)."""
        p = self.get_profile(profile_id)
        if p is None:
            raise ValueError(f"Profile with ID {profile_id} not found")
        dens = p.get("density_g_cm3")
        lines = ["# SAXS Solvent Profile",
                 f"# Water Model: {p['water_model']}",
                 f"# Water Molecules: {p['n_water_molecules']}",
                 f"# Ion Counts: {p['ion_counts']}",
                 f"# Box Size: {p['box_x']:.3f} x {p['box_y']:.3f} x {p['box_z']:.3f} Å",
                 f"# Supercell Scale: {p['supercell_scale']:.4f}",
                 f"# Simulation Time: {p['simulation_time_ps']:.2f} ps",
                 f"# Frames Analyzed: {p['n_frames_analyzed']}",
                 f"# Density: {f'{dens:.4f}' if dens is not None else 'N/A'} g/cm³",
                 "#",
                 "# q (1/Å), I(q) (1/Å³)"]
        lines += [f"{q:.6f},{i:.6e}" for q, i in p["profile_data"]]
        Path(output_path).write_text("\n".join(lines) + "\n")

    def close(self):
        self.conn.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


# ---------------------------------------------------------------- subtraction --

def interpolate_profile(q_ref, iq_ref, q_target, method: str = "cubic") -> np.ndarray:
    """I_ref on q_target: the cubic (not-a-knot) or linear spline through
    the reference points, extrapolated beyond them (saxs_subtract.py:64-86)."""
    from scipy.interpolate import interp1d
    kind = "cubic" if method == "cubic" else "linear"
    f = interp1d(np.asarray(q_ref, dtype=np.float64), np.asarray(iq_ref, dtype=np.float64),
                 kind=kind, bounds_error=False, fill_value="extrapolate")
    return f(np.asarray(q_target, dtype=np.float64))


def subtract_profiles(user_profile: Dict, ref_profile: Dict, scale: float,
                      interp_method: str = "cubic") -> Tuple[np.ndarray, np.ndarray]:
    """(q_user, I_user - scale I_ref) with the subtraction applied only where
    q_user lies in [max(q_user[0], q_ref[0]), min(q_user[-1], q_ref[-1])];
    the reference is interpolated onto q_user unless both grids agree to
    1e-6 relative (saxs_subtract.py:126-198).  Raises ValueError when the
    ranges do not overlap."""
    u = np.asarray(user_profile["profile_data"], dtype=np.float64)
    r = np.asarray(ref_profile["profile_data"], dtype=np.float64)
    qu, iu, qr, ir = u[:, 0], u[:, 1], r[:, 0], r[:, 1]
    lo, hi = max(qu[0], qr[0]), min(qu[-1], qr[-1])
    if hi <= lo:
        raise ValueError("No overlap between user and reference q-ranges!")
    mask = (qu >= lo) & (qu <= hi)
    same = len(qu) == len(qr) and np.allclose(qu, qr, rtol=1e-6)
    ir_on_u = ir if same else interpolate_profile(qr, ir, qu, interp_method)
    out = iu.copy()
    out[mask] = iu[mask] - scale * ir_on_u[mask]
    return qu, out


def resample_profile(q, iq, dq: float, method: str = "cubic") -> Tuple[np.ndarray, np.ndarray]:
    """The profile on ceil((q[-1] - q[0]) / dq) + 1 evenly spaced points from
    q[0] to q[-1] (saxs_subtract.py:201-232)."""
    q = np.asarray(q, dtype=np.float64)
    n = int(np.ceil((q[-1] - q[0]) / dq)) + 1
    qu = np.linspace(q[0], q[-1], n)
    return qu, interpolate_profile(q, iq, qu, method)


def _ions_text(d: Dict) -> str:
    s = ", ".join(f"{k}: {v}" for k, v in d.items() if v > 0) if d else ""
    return s or "none"


def _profile_header(title: str, p: Dict) -> List[str]:
    dens = p.get("density_g_cm3")
    other = p.get("other_molecules") or {}
    return [f"# {title}:",
            f"#   ID: {p['id']}",
            f"#   Water Model: {p['water_model']}",
            f"#   Ions: {_ions_text(p.get('ion_counts') or {})}",
            f"#   Other Molecules: {', '.join(f'{k}: {v}' for k, v in other.items()) or 'none'}",
            f"#   Grid: {p['grid_size']}",
            f"#   Supercell Scale: {p['supercell_scale']:.4f}",
            f"#   Box: {p['box_x']:.3f} x {p['box_y']:.3f} x {p['box_z']:.3f} Å",
            f"#   Volume: {p['box_volume']:.2f} Ų",
            f"#   Density: {f'{dens:.4f}' if dens is not None else 'N/A'} g/cm³",
            f"#   Simulation Time: {p['simulation_time_ps']:.2f} ps",
            "#"]


def save_subtracted_profile(output_path, q, iq, user_profile: Dict, ref_profile: Dict,
                            scale: float, resampled: bool = False,
                            original_points: Optional[int] = None) -> None:
    """The subtracted profile with the metadata header of
    saxs_subtract.py:235-333 ("q  I(q)" rows as %.6f  %.6e)."""
    lines = ["# SAXS Profile after Solvent Subtraction", "#"]
    lines += _profile_header("User Profile", user_profile)
    lines += _profile_header("Reference Profile (subtracted)", ref_profile)
    lines += ["# Subtraction Parameters:",
              f"#   Scaling Factor: {scale:.6f}",
              "#   Note: Subtraction applied only in overlap region between profiles",
              "#"]
    if resampled and original_points is not None:
        lines += ["# Output Processing:",
                  f"#   Original data points: {original_points}",
                  f"#   Resampled to: {len(q)} points",
                  f"#   Q spacing (dq): {(q[1] - q[0]):.6f} Å⁻¹",
                  "#"]
    lines.append("# q (Å⁻¹)    I(q) [subtracted]")
    lines += [f"{a:.6f}  {b:.6e}" for a, b in zip(q, iq)]
    Path(output_path).write_text("\n".join(lines) + "\n")


def main(argv: Optional[Sequence[str]] = None) -> int:
    """saxs-subtract without prompts: --id and --scale are required (the
    reference asks for them interactively when absent)."""
    ap = argparse.ArgumentParser(description="Subtract a reference solvent SAXS profile "
                                             "from a simulated profile")
    ap.add_argument("--db", required=True, help="database holding the simulated profile")
    ap.add_argument("--id", type=int, required=True, help="profile id in --db")
    ap.add_argument("--ref-db", required=True, help="database holding the solvent profile")
    ap.add_argument("--ref-id", type=int, required=True, help="profile id in --ref-db")
    ap.add_argument("--scale", type=float, required=True, help="scaling factor (> 0)")
    ap.add_argument("-o", "--output", help="output file (default subtracted_<id>.dat)")
    ap.add_argument("--dq", type=float, help="resample onto a uniform grid of this spacing")
    ap.add_argument("--interp", choices=("linear", "cubic"), default="cubic")
    a = ap.parse_args(argv)
    if not a.scale > 0:
        print("Error: scaling factor must be positive.", file=sys.stderr)
        return 1
    for path in (a.db, a.ref_db):
        if not Path(path).exists():
            print(f"Error: database not found: {path}", file=sys.stderr)
            return 1
    with SaxsDatabase(a.db, read_only=True) as udb, SaxsDatabase(a.ref_db, read_only=True) as rdb:
        user, ref = udb.get_profile(a.id), rdb.get_profile(a.ref_id)
    if user is None or ref is None:
        print(f"Error: profile {a.id if user is None else a.ref_id} not found.", file=sys.stderr)
        return 1
    q, iq = subtract_profiles(user, ref, a.scale, a.interp)
    n0 = len(q)
    if a.dq:
        q, iq = resample_profile(q, iq, a.dq, a.interp)
    out = Path(a.output) if a.output else Path(f"subtracted_{a.id}.dat")
    save_subtracted_profile(out, q, iq, user, ref, a.scale, resampled=bool(a.dq),
                            original_points=n0 if a.dq else None)
    print(f"Subtracted profile written to {out}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
