"""CPU-side checks: the C-ABI library loads and exports every declared symbol, and the
host logic of the SOAP mirror (parameters, basis tables, feature layout, errors)
agrees with the oracle.  No GPU calls."""

import math
import os
import re

import numpy as np
import pytest

from cusoap_amd import SOAP, _lib, systems
from oracle import soap_oracle as so

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    header = open(os.path.join(ROOT, "include", "cusoap_amd.h")).read()
    declared = set(re.findall(r"\b(cusoap_[a-z0-9_]+)\s*\(", header))
    lib = _lib.load()
    for name in declared:
        assert hasattr(lib, name), name
    assert declared == set(_lib.exported_symbols())
    assert lib.cusoap_max_l() == 9 and lib.cusoap_max_n() == 16


def test_constructor_errors_match_reference():
    with pytest.raises(ValueError):
        SOAP(r_cut=5.0, n_max=None, l_max=2, species=[1])
    with pytest.raises(ValueError):
        SOAP(r_cut=5.0, n_max=2, l_max=2, species=[])
    with pytest.raises(ValueError):
        SOAP(r_cut=5.0, n_max=2, l_max=2, species=[1], average="bogus")
    with pytest.raises(ValueError):
        SOAP(r_cut=5.0, n_max=2, l_max=2, species=[1], rbf="bessel")
    with pytest.raises(ValueError):
        SOAP(n_max=2, l_max=2, species=[1])
    with pytest.raises(ValueError):
        SOAP(r_cut=5.0, n_max=2, l_max=2, species=[1], dtype="float16")


def test_no_cpu_path():
    s = SOAP(r_cut=5.0, n_max=2, l_max=2, species=[1, 8], device="cpu")
    with pytest.raises(RuntimeError):
        s.create(systems.water((1, 1, 1)))


@pytest.mark.parametrize("cross", [True, False])
def test_feature_layout_matches_oracle(cross):
    kw = dict(r_cut=5.0, n_max=3, l_max=2, sigma=0.5, species=["H", "C", "O"], crossover=cross)
    g = SOAP(device="cpu", **kw)
    o = so.OracleSOAP(**dict(kw, species=[1, 6, 8]))
    assert g.get_number_of_features() == o.get_number_of_features()
    np.testing.assert_array_equal(g._feat_rows, o.feature_table())
    for a in (1, 6, 8):
        for b in (1, 6, 8):
            if cross or a == b:
                assert g.get_location((a, b)) == o.get_location((a, b))
    assert SOAP(device="cpu", average="cc", **kw).get_number_of_features() == 3 * 3 * 9


def test_host_basis_tables_match_oracle():
    for rbf in ("gto", "polynomial"):
        kw = dict(r_cut=6.0, n_max=5, l_max=4, sigma=0.5, species=[1, 8], rbf=rbf,
                  weighting=dict(function="pow", r0=1.0, c=1.0, d=1.0, m=2.0, w0=0.5))
        g = SOAP(device="cpu", **kw)
        o = so.OracleSOAP(**kw)
        assert g._cutoff == o.cutoff
        np.testing.assert_allclose(g._self_l0, o.self_l0, rtol=1e-12)
        if rbf == "gto":
            np.testing.assert_array_equal(g._alphas, o.alphas)
            gamma, pref = o.gto_radial_tables()
            np.testing.assert_array_equal(g._gto_gamma, gamma)
            np.testing.assert_array_equal(g._gto_pref, pref)
        else:
            np.testing.assert_array_equal(g._gss, o.gss)
            B0, Bh0 = o.poly_B_Bhat_r0()
            np.testing.assert_array_equal(g._poly_B0, B0)
            np.testing.assert_array_equal(g._poly_Bh0, Bh0)


def test_weighting_rcut_inference():
    g = SOAP(device="cpu", n_max=2, l_max=1, species=[1],
             weighting=dict(function="pow", r0=1.0, c=1.0, d=1.0, m=2.0, threshold=1e-2))
    np.testing.assert_allclose(g._r_cut, math.sqrt(99.0))


def test_pbc_shifts_match_oracle():
    cell = np.array([[5.0, 0, 0], [2.5, 4.0, 0], [0.3, 0.2, 6.0]])
    from cusoap_amd.soap import pbc_image_shifts
    pbc = np.array([True, True, True])
    offs = so.pbc_offsets(cell, pbc, 7.0)
    np.testing.assert_array_equal(pbc_image_shifts(cell, pbc, 7.0), so.image_shifts(cell, offs))


def test_xyz_reader_on_reference_format(tmp_path):
    p = tmp_path / "w.xyz"
    p.write_text("3\n\n  O  0.0 0.0 0.0\n  H  0.9572 0.0 0.0\n  H  -0.24 0.92 0.0\n")
    fr = systems.read_xyz(str(p))
    assert len(fr) == 1 and fr[0]["atomic_numbers"].tolist() == [8, 1, 1]


def test_workload_generators():
    s = systems.water_cluster_300()
    assert s["positions"].shape == (300, 3)
    d = np.linalg.norm(s["positions"][1] - s["positions"][0])
    np.testing.assert_allclose(d, 0.9572, rtol=1e-12)
    v1 = s["positions"][1] - s["positions"][0]
    v2 = s["positions"][2] - s["positions"][0]
    ang = math.degrees(math.acos(v1 @ v2 / (np.linalg.norm(v1) * np.linalg.norm(v2))))
    np.testing.assert_allclose(ang, 104.52, rtol=1e-10)
    np.testing.assert_allclose(systems.water_box_length(32 ** 3), 99.338, atol=1e-3)
    b = systems.random_box(n_atoms=500, density=0.1, seed=1)
    L = b["cell"][0, 0]
    dd = b["positions"][:, None] - b["positions"][None]
    dd -= L * np.round(dd / L)
    r = np.sqrt((dd ** 2).sum(-1)) + np.eye(500) * 10
    assert r.min() >= 0.9
