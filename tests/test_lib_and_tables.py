"""CPU checks of the engine's C ABI library and host-side tables.

The library must load and export every symbol include/saxs_engine.h
declares (no compute call is made without a GPU).  The host tables the
engine hands to the GPU (FFT sizes, exact shell bins, moduli, form factors)
must match the oracle and, where it is built, the reference's own helpers
compiled in oracle/_ref."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

from oracle import saxs_oracle as O
from pycusaxs_amd import _lib, formfactors
from pycusaxs_amd import geometry as G

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_DRIVER = os.path.join(ROOT, "oracle", "_ref", "ref_driver")
HEADER = os.path.join(ROOT, "include", "saxs_engine.h")


def declared_symbols():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|int32_t|int64_t|const char\*)\s+(saxs_\w+)\(",
                                 txt, flags=re.M)))


def test_header_and_binding_agree():
    assert set(declared_symbols()) == set(_lib.EXPORTED)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(_lib.LIB_PATH))
    for name in declared_symbols():
        assert hasattr(lib, name), name
    L = _lib.load()
    assert L.saxs_abi_version() == _lib.ABI_VERSION
    assert L.saxs_last_error() == b""


def test_plan_create_rejects_bad_desc_without_gpu():
    L = _lib.load()
    d = _lib.SaxsDesc()
    d.n[:] = [32, 32, 32]
    d.nn[:] = [64, 64, 64]
    d.order = 5
    plan = ctypes.c_void_p()
    rc = L.saxs_plan_create(ctypes.byref(d), ctypes.byref(plan))
    assert rc == _lib.SAXS_EUNSUP and b"order" in L.saxs_last_error()
    with pytest.raises(ValueError):
        _lib.check(rc, "saxs_plan_create")


@pytest.mark.parametrize("n,sigma", [(64, 2.0), (128, 2.0), (27, 2.0), (21, 2.5), (37, 2.0),
                                     (100, 2.5), (50, 1.7), (192, 2.0)])
def test_fft_size_matches_oracle(n, sigma):
    assert G.fft_size(n, sigma) == O.find_closest_product(n, sigma)


@pytest.mark.parametrize("L,n,dq,qcut,policy", [(37.0, 64, 0.01, 0.5, "reference"),
                                                (100.0, 128, 0.005, 0.6, "reference"),
                                                (80.0, 128, 0.005, 0.5, "exact"),
                                                (25.0, 21, 0.0, 100.0, "reference"),
                                                (46.4, 75, 0.02, 0.9, "reference")])
def test_shell_bins_match_oracle(L, n, dq, qcut, policy):
    oc = G.inverse_box_diag(np.diag([L] * 3))
    nn = (G.fft_size(n, 2.0),) * 3
    h = G.hist_params((n,) * 3, nn, oc, 2.0, dq, qcut, policy)
    ho = O.hist_params((n,) * 3, nn, oc, 2.0, dq, qcut, policy)
    assert (np.float32(h.dq), np.float32(h.qcut), h.num_bins) == (ho.dq, ho.qcut, ho.num_bins)
    b = G.shell_bins(h.spacing[0], h.dq, h.qcut, h.num_bins)
    bo = [O.bin_of_n(v, ho.spacing[0], ho.dq, ho.qcut, ho.num_bins) for v in range(len(b) + 5)]
    assert list(b) == bo[:len(b)]
    assert all(v == -1 for v in bo[len(b):])


@pytest.mark.parametrize("N,p", [(128, 4), (256, 4), (135, 4), (64, 6), (140, 6)])
def test_moduli_match_oracle(N, p):
    assert np.allclose(G.bspline_moduli(N, p), O.bsp_moduli(N, p), rtol=1e-13)


def test_form_factor_table_matches_oracle():
    syms = list(O.FORM_FACTORS)
    ff = G.form_factor_table(syms, 0.05, 400)
    q = 0.05 * np.sqrt(np.arange(400))
    for e, s in enumerate(syms):
        assert np.allclose(ff[e], O.form_factor(s, q), rtol=1e-14, atol=1e-14)


def test_unknown_element_raises():
    with pytest.raises(ValueError):
        formfactors.coefficients("Xx")


def test_non_lower_triangular_box_rejected():
    with pytest.raises(ValueError):
        G.inverse_box_matrix([[30, 5, 0], [0, 30, 0], [0, 0, 30]])


def test_orthorhombic_helper_rejects_triclinic():
    with pytest.raises(ValueError):
        G.inverse_box_diag([[30, 0, 0], [5, 30, 0], [0, 0, 30]])


# ---- against the reference's own helpers (compiled from its sources) ----

needs_ref = pytest.mark.skipif(not os.path.exists(REF_DRIVER), reason="oracle/_ref not built")


def ref(*args):
    out = subprocess.run([REF_DRIVER, *map(str, args)], capture_output=True, text=True, check=True)
    return out.stdout.strip().splitlines()


@needs_ref
def test_form_factor_data_matches_reference_table():
    syms = list(formfactors.table())
    for chunk in range(0, len(syms), 40):
        part = syms[chunk:chunk + 40]
        lines = ref("scatter", *part)
        for s, line in zip(part, lines):
            vals = np.array([float(v) for v in line.split()])
            assert np.array_equal(vals.astype(np.float32),
                                  np.array(formfactors.coefficients(s), dtype=np.float32)), s
            assert np.array_equal(vals.astype(np.float32),
                                  np.array(O.FORM_FACTORS.get(s, formfactors.coefficients(s)),
                                           dtype=np.float32)), s


@needs_ref
@pytest.mark.parametrize("p", [4, 6])
def test_bspline_weights_match_reference(p):
    ws = [0.0, 0.125, 0.3, 0.5, 0.77, 0.999]
    for w, line in zip(ws, ref("bspline", p, *ws)):
        x, dx = line.split("|")
        x = np.array([float(v) for v in x.split()])
        dx = np.array([float(v) for v in dx.split()])
        w32 = float(np.float32(w))
        assert np.allclose(x, O.bspline_theta(w32, p), atol=3e-7)
        assert np.allclose(dx, O.bspline_dtheta(w32, p), atol=3e-7)


@needs_ref
@pytest.mark.parametrize("N", [64, 128, 256, 135])
def test_moduli_match_reference(N):
    lines = ref("bspmod", 4, N, N, N)
    mod = np.array([float(v) for v in lines[0].split()])
    assert np.allclose(mod, O.bsp_moduli(N, 4), rtol=2e-5)
    assert np.allclose(mod, G.bspline_moduli(N, 4), rtol=2e-5)


@needs_ref
def test_cell_inverse_matches_reference():
    line = ref("cell", 37, 0, 0, 0, 41.5, 0, 0, 0, 100)[0]
    oc = np.array([float(v) for v in line.split()]).reshape(3, 3)
    assert np.array_equal(np.diag(oc).astype(np.float32), G.inverse_box_diag(np.diag([37, 41.5, 100])))
    assert np.array_equal(np.diag(oc).astype(np.float32), O.inverse_box(np.diag([37, 41.5, 100])))


@needs_ref
@pytest.mark.parametrize("lengths,n,model,ions", [
    ((37.0, 37.0, 37.0), (64, 64, 64), "SPCE", {}),
    ((100.0, 100.0, 100.0), (128, 128, 128), "TIP3P", {"Na": 150, "Cl": 150}),
    ((80.5, 61.25, 90.0), (96, 72, 108), "TIP4P", {"K": 7, "Cl": 9, "Ca": 1}),
    ((215.0, 215.0, 215.0), (256, 256, 256), "SPC", {"Mg": 3}),
])
def test_solvent_pads_match_reference(lengths, n, model, ions):
    """padding::given values, bit for bit in float32, from the engine's host
    code and from the oracle against AtomCounter compiled from its source."""
    args = [*lengths, model, *n]
    for k, v in ions.items():
        args += [k, v]
    got = {}
    for line in ref("atomcount", *args):
        sym, val = line.split()
        got[sym] = np.float32(float(val))
    mine = G.explicit_solvent_pads(lengths, n, model, ions)
    orc = O.solvent_pads(lengths, n, model, ions)
    assert set(got) == set(mine) == set(orc)
    for s in got:
        assert np.float32(mine[s]) == got[s], s
        assert np.float32(orc[s]) == got[s], s


@needs_ref
def test_unknown_water_model_fails_like_reference():
    assert ref("atomcount", 37, 37, 37, "XYZ", 64, 64, 64) == ["Invalid water model 'XYZ'. Using SPCE as default.", "error"]
    with pytest.raises(ValueError):
        G.explicit_solvent_pads((37.0,) * 3, (64,) * 3, "XYZ", {})


@pytest.mark.parametrize("Ls,n,dq,qcut,policy", [((40.0, 50.0, 60.0), 32, 0.0, 0.9, "reference"),
                                                 ((33.0, 40.0, 37.0), 30, 0.0, 1.2, "reference"),
                                                 ((60.0, 70.0, 80.0), 64, 0.01, 0.6, "exact"),
                                                 ((40.0, 40.0, 20.0), 32, 0.0, 0.9, "reference")])
def test_point_bins_match_oracle(Ls, n, dq, qcut, policy):
    """Point-keyed tables of orthorhombic cells (engine host code, float64
    with an exact integer decision near edges) against the oracle's exact
    rational rule, point by point.  The 40:40:20 cell puts many points exactly
    on bin edges (dq is twice the two equal spacings)."""
    oc = G.inverse_box_diag(np.diag(Ls))
    nn = (2 * n,) * 3
    h = G.hist_params((n,) * 3, nn, oc, 2.0, dq, qcut, policy)
    km = G.point_kmax(h.spacing, h.qcut)
    ho = O.hist_params((n,) * 3, nn, O.inverse_box(np.diag(Ls)), 2.0, dq, qcut, policy)
    assert tuple(float(v) for v in ho.spacing) == tuple(h.spacing)
    for a in range(3):      # kmax is the largest |m_a| inside q <= qcut on its own
        s = float(h.spacing[a])
        assert (km[a] * s) ** 2 <= float(h.qcut) ** 2 * (1 + 1e-12)
        assert ((km[a] + 1) * s) ** 2 > float(h.qcut) ** 2
    bins = G.point_bins(h.spacing, h.dq, h.qcut, h.num_bins, km)
    ff = G.form_factor_points(["C", "O"], h.spacing, km)
    i = 0
    for ax in range(km[0] + 1):
        for ay in range(km[1] + 1):
            for az in range(km[2] + 1):
                assert bins[i] == O.bin_of_m((ax, ay, az), ho.spacing, ho.dq, ho.qcut, ho.num_bins)
                q = np.sqrt(sum((m * float(s)) ** 2 for m, s in zip((ax, ay, az), h.spacing)))
                assert ff[1, i] == pytest.approx(O.form_factor("O", q), rel=1e-12)
                i += 1
    assert i == len(bins) and (bins >= 0).sum() > 50


@needs_ref
@pytest.mark.parametrize("B", [
    [[37.0, 0, 0], [0, 41.5, 0], [0, 0, 100.0]],
    [[30.0, 0, 0], [7.5, 28.0, 0], [-5.0, 6.25, 33.0]],
    [[61.3, 0, 0], [-20.4, 57.7, 0], [-20.4, -28.9, 50.0]],      # truncated-octahedron-like
])
def test_triclinic_inverse_matches_reference(B):
    """M = Cell.cpp's upper-triangular inverse applied to the transposed box,
    bit for bit in float32 (the reference helper is given B^T)."""
    B = np.asarray(B, dtype=np.float32)
    line = ref("cell", *[float(v) for v in B.T.ravel()])[0]
    M_ref = np.array([float(v) for v in line.split()], dtype=np.float64).reshape(3, 3)
    assert np.array_equal(M_ref.astype(np.float32), G.inverse_box_matrix(B))
    assert np.array_equal(M_ref.astype(np.float32), O.inverse_box_matrix(B))


def test_triclinic_bins_match_oracle():
    """Signed point tables (engine host code, integers) against the oracle's
    exact rational rule, and kmax covers every point inside qcut."""
    B = np.array([[30.0, 0, 0], [7.5, 28.0, 0], [-5.0, 6.25, 33.0]])
    n, nn = (24, 24, 24), (48, 48, 48)
    M = G.inverse_box_matrix(B)
    assert G.is_triclinic(M)
    h = G.hist_params(n, nn, np.diag(M), 2.0, 0.0, 0.7, "reference")
    basis = G.reciprocal_basis(M, n, nn)
    assert np.array_equal(basis, O.reciprocal_basis(O.inverse_box_matrix(B), n, nn))
    km = G.triclinic_kmax(basis, h.qcut, nn)
    bins = G.triclinic_bins(basis, h.dq, h.qcut, h.num_bins, km)
    ff = G.form_factor_vectors(["S"], basis, km)
    i = 0
    for mx in range(-km[0], km[0] + 1):
        for my in range(-km[1], km[1] + 1):
            for mz in range(-km[2], km[2] + 1):
                assert bins[i] == O.bin_of_vec((mx, my, mz), basis, h.dq, h.qcut, h.num_bins)
                q = np.linalg.norm(mx * basis[0].astype(float) + my * basis[1].astype(float)
                                   + mz * basis[2].astype(float))
                assert ff[0, i] == pytest.approx(O.form_factor("S", q), rel=1e-12)
                i += 1
    assert (bins >= 0).sum() > 100
    # nothing inside qcut lies outside the box
    for a in range(3):
        for extra in (km[a] + 1, -km[a] - 1):
            m = [0, 0, 0]
            m[a] = extra
            for u in range(-km[(a + 1) % 3], km[(a + 1) % 3] + 1):
                m[(a + 1) % 3] = u
                for v in range(-km[(a + 2) % 3], km[(a + 2) % 3] + 1):
                    m[(a + 2) % 3] = v
                    assert O.bin_of_vec(tuple(m), basis, h.dq, h.qcut, h.num_bins) == -1


def test_capacity_error_maps_to_value_error_subclass():
    """SAXS_ECAPACITY (a box whose q tables need more workspace) is a
    ValueError to callers of the ABI, and a distinct class the engine
    catches to create a plan for that box."""
    with pytest.raises(_lib.CapacityError):
        _lib.check(_lib.SAXS_ECAPACITY, "saxs_plan_set_tables")
    assert issubclass(_lib.CapacityError, ValueError)
    txt = open(HEADER).read()
    assert "#define SAXS_ECAPACITY (-5)" in txt


@pytest.mark.parametrize("nz", [16, 128, 512])
def test_given_pad_bound(nz):
    """A given pad times 2^32 fixed-point units, summed over a column of n_z
    voxels, stays below 2^63: the bound is 2^31 / n_z atoms per voxel."""
    lim = G.max_given_pad(nz)
    assert lim * 2.0 ** 32 * nz <= 2.0 ** 63
    assert np.nextafter(lim, np.inf) * 2.0 ** 32 * nz > 2.0 ** 63
