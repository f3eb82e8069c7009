"""The pycusaxs_cuda.run boundary: argument validation and error mapping
(cuda_bindings.cpp:32-167, cudaSAXS.cu:50-65), output formatting
(RunSaxs.cu:224-227) and element grouping (RunSaxs.cu:108-115).  These run
on the CPU: every check here fires before the engine touches the GPU."""
import numpy as np
import pytest

import pycusaxs_cuda
from pycusaxs_amd import backend
from tests.fake_topology import FakeTopology


def call(**kw):
    args = dict(obj_topology=None, topology="sys.tpr", trajectory="traj.xtc", grid=[32])
    args.update(kw)
    return pycusaxs_cuda.run(**args)


@pytest.mark.parametrize("kw,msg", [
    (dict(grid=[]), "at least one dimension"),
    (dict(grid=[0]), "must be positive"),
    (dict(grid=[32, 32]), "either 1 or 3"),
    (dict(grid=[32, -1, 32]), "must be positive"),
    (dict(scaled_grid=[0]), "Scaled grid dimension must be positive"),
    (dict(scaled_grid=[64, 64]), "either 1 or 3"),
    (dict(stride=0), "Stride must be positive"),
    (dict(order=0), "between 1 and 8"),
    (dict(order=9), "between 1 and 8"),
    (dict(bin_size=-0.1), "Bin size"),
    (dict(qcut=-1.0), "Q cutoff"),
    (dict(scale_factor=-1.0), "Scale factor must be non-negative"),
    (dict(scale_factor=0.0), "Scale factor must be positive"),
    (dict(begin=-1), "non-negative"),
    (dict(begin=3, end=2), "Begin frame"),
    (dict(topology=""), "Topology path"),
    (dict(trajectory=""), "Trajectory path"),
    (dict(order=5), "not supported"),
    (dict(padding="none"), "padding must be"),
    (dict(padding="given"), "needs a water model"),
    (dict(padding="given", water_model="TIP5P"), "needs a water model"),
])
def test_invalid_arguments_raise_value_error(kw, msg):
    with pytest.raises(ValueError, match=msg):
        call(**kw)


def test_keyword_defaults_match_reference():
    import inspect
    sig = inspect.signature(pycusaxs_cuda.run)
    d = {k: v.default for k, v in sig.parameters.items()}
    assert list(sig.parameters)[:17] == [
        "obj_topology", "topology", "trajectory", "grid", "scaled_grid", "begin", "end",
        "stride", "output", "order", "scale_factor", "bin_size", "qcut", "water_model",
        "sodium", "chlorine", "simulation"]
    assert (d["begin"], d["end"], d["stride"], d["output"], d["order"], d["scale_factor"],
            d["bin_size"], d["qcut"], d["water_model"], d["sodium"], d["chlorine"],
            d["simulation"]) == (0, 0, 1, "", 4, 1.0, 0.0, 0.0, "", 0, 0, "")
    assert list(d["scaled_grid"]) == []


def test_row_format():
    txt = backend.format_rows([(0.08491, 822274.6), (0.16982, 1.5e-3)])
    assert txt == "   0.08491 8.22275e+05\n   0.16982 1.50000e-03\n"


def test_element_groups_sorted_like_std_map():
    syms, idx, ids = backend.element_groups({"O": [2, 5], "H": [0, 1, 3], "Cl": [4], "C": [6]})
    assert syms == ["C", "Cl", "H", "O"]
    assert idx.tolist() == [6, 4, 0, 1, 3, 2, 5]
    assert ids.tolist() == [0, 1, 2, 2, 2, 3, 3]


def test_unknown_element_is_an_error():
    with pytest.raises(ValueError, match="form factor"):
        backend.element_groups({"H": [0], "Qq": [1]})


def test_malformed_frame_record_is_an_error():
    with pytest.raises(ValueError):
        backend._frame_fields({"frame": 0, "time": 0.0, "positions": np.zeros((3, 2))})
    with pytest.raises(ValueError):
        backend._frame_fields({"frame": 0})
    f, t, p, b = backend._frame_fields({"frame": 3, "time": 1.5, "positions": np.zeros((4, 3)),
                                        "box": np.eye(3)})
    assert (f, t, p.shape, b.shape) == (3, 1.5, (4, 3), (3, 3))


def test_summary_mentions_configuration():
    s = backend.summary_text("a.tpr", "b.xtc", [64] * 3, [], 0, 9, 2, 4, 2.0, 0.01, 0.5, "",
                             {"Na": 3, "Cl": 0})
    assert "Total Frames  : 5" in s and "Grid Size     : (64, 64, 64)" in s
    assert "Na⁺ atoms     : 3" in s and "Cl" not in s.split("Solvent Model")[1]
