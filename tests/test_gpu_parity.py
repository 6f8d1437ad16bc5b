"""HIP path against the FP64 oracle on the same seeded inputs.

Tolerances: the float64 path is held to 1e-10 relative error per feature column
(max_c |gpu - ref| / max_c |ref|); neighbour lists must be identical.
"""

import math
import sys

import numpy as np
import pytest
import torch

from cusoap_amd import SOAP, systems
from oracle import soap_oracle as so

pytestmark = pytest.mark.gpu

RTOL64 = 1e-10


def colerr(a, b):
    a = a.detach().cpu().double().numpy() if isinstance(a, torch.Tensor) else np.asarray(a)
    b = np.asarray(b, dtype=np.float64)
    a = a.reshape(-1, a.shape[-1])
    b = b.reshape(-1, b.shape[-1])
    scale = np.abs(b).max(axis=0)
    scale = np.where(scale > 0, scale, 1.0)
    return float((np.abs(a - b).max(axis=0) / scale).max())


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def pair(params, **kw):
    o = so.OracleSOAP(**params)
    extra = {"_betas": o.betas} if params.get("rbf", "gto") == "gto" else {}
    g = SOAP(device="cuda:0", dtype="float64", **params, **extra, **kw)
    return o, g


W300 = dict(r_cut=6.0, n_max=4, l_max=3, sigma=0.5, species=[1, 8])


# ---- neighbour list -------------------------------------------------------------

@pytest.mark.parametrize("drop_self", [True, False])
def test_nl_bit_exact_cluster(drop_self):
    s = systems.water_cluster_300(seed=0)
    o, g = pair(W300)
    say("nl cluster300 drop_self", drop_self)
    off, nbr = g.neighbor_list(s, drop_self=drop_self)
    ci, ni, _, _ = o.neighbour_list(s["positions"], s["atomic_numbers"], s["positions"], drop_self=drop_self)
    counts = np.diff(off.cpu().numpy())
    np.testing.assert_array_equal(counts, np.bincount(ci, minlength=300))
    np.testing.assert_array_equal(nbr.cpu().numpy(), ni)


def test_nl_bit_exact_periodic_small_cell():
    s = systems.water((3, 3, 3), spacing=3.1, seed=2, cell_length=9.3)
    p = dict(W300, periodic=True)
    o, g = pair(p)
    say("nl periodic 81 atoms, L=9.3")
    off, nbr = g.neighbor_list(s, drop_self=True)
    ci, ni, _, _ = o.neighbour_list(s["positions"], s["atomic_numbers"], s["positions"], s["cell"], s["pbc"],
                                    drop_self=True)
    np.testing.assert_array_equal(np.diff(off.cpu().numpy()), np.bincount(ci, minlength=81))
    np.testing.assert_array_equal(nbr.cpu().numpy(), ni)


def test_nl_bit_exact_3000_and_coordinate_centers():
    s = systems.water_cluster_3000(seed=1)
    o, g = pair(dict(W300, r_cut=10.0))
    say("nl cluster3000 r_cut=10")
    off, nbr = g.neighbor_list(s, drop_self=True)
    ci, ni, _, _ = o.neighbour_list(s["positions"], s["atomic_numbers"], s["positions"], drop_self=True)
    np.testing.assert_array_equal(nbr.cpu().numpy(), ni)
    cen = [[1.0, 2.0, 3.0], 17, [10.5, 11.25, 0.0]]
    off, nbr = g.neighbor_list(s, centers=cen, drop_self=True)
    cx, _ = o.prepare_centers(s["positions"], cen)
    ci, ni, _, _ = o.neighbour_list(s["positions"], s["atomic_numbers"], cx, drop_self=True)
    np.testing.assert_array_equal(nbr.cpu().numpy(), ni)


def test_nl_bit_exact_over_1024_neighbours():
    """Centers with more than 1024 edges: the workgroup-per-center fill with 4096 slots
    (the 3,000-atom test at r_cut 10 takes the 1024-slot one; the 300-atom tests take
    the rows that the count pass sorted)."""
    s = systems.water_cluster_3000(seed=3)
    o, g = pair(dict(W300, r_cut=20.0))
    say("nl cluster3000 r_cut=20")
    off, nbr = g.neighbor_list(s, drop_self=True)
    counts = np.diff(off.cpu().numpy())
    assert counts.max() > 1024
    ci, ni, _, _ = o.neighbour_list(s["positions"], s["atomic_numbers"], s["positions"], drop_self=True)
    np.testing.assert_array_equal(counts, np.bincount(ci, minlength=3000))
    np.testing.assert_array_equal(nbr.cpu().numpy(), ni)


# ---- descriptors ------------------------------------------------------------------

@pytest.mark.parametrize("average", ["off", "inner", "outer", "cc"])
def test_create_gto_f64(average):
    s = systems.water_cluster_300(seed=0)
    o, g = pair(dict(W300, average=average))
    say("create gto 300", average)
    ref = o.create(s)
    out = g.create(s)
    assert out.shape == ref.shape and out.dtype == torch.float64
    assert colerr(out, ref) < RTOL64


def test_create_gto_f32_output_and_centers():
    s = systems.water_cluster_300(seed=3)
    o = so.OracleSOAP(**W300)
    g = SOAP(device="cuda:0", dtype="float32", _betas=o.betas, **W300)
    cen = [5, 0, [1.0, 1.0, 1.0], 299]
    say("create f32 output, explicit centers")
    out = g.create(s, centers=cen)
    assert out.dtype == torch.float32
    ref = o.create(s, centers=cen)
    assert colerr(out, ref) < 1e-6


def test_create_poly_f64():
    s = systems.water_cluster_300(seed=4)
    p = dict(W300, rbf="polynomial")
    o, g = pair(p)
    say("create poly 300")
    assert colerr(g.create(s), o.create(s)) < RTOL64


@pytest.mark.parametrize("wd", [dict(function="pow", r0=2.0, c=1.0, d=1.0, m=2.0, w0=0.5),
                                dict(function="poly", r0=5.5, c=1.0, m=2.0),
                                dict(function="exp", r0=1.0, c=1.5, d=0.5)])
def test_create_weighting_f64(wd):
    s = systems.water_cluster_300(seed=5)
    o, g = pair(dict(W300, weighting=wd))
    say("create weighting", wd["function"])
    assert colerr(g.create(s), o.create(s)) < RTOL64


def test_create_periodic_f64_crossover_off():
    s = systems.water((3, 3, 3), spacing=3.1, seed=6, cell_length=9.3)
    o, g = pair(dict(W300, periodic=True, crossover=False))
    say("create periodic crossover=False")
    assert colerr(g.create(s), o.create(s)) < RTOL64


def test_create_multi_species_l6():
    s = systems.random_box(n_atoms=400, density=0.1, seed=3)
    p = dict(r_cut=5.0, n_max=8, l_max=6, sigma=0.4, species=[1, 6, 7, 8], periodic=True)
    o, g = pair(p)
    say("create 4-species periodic n8 l6")
    assert colerr(g.create(s), o.create(s)) < RTOL64


def test_create_list_of_systems_and_sparse():
    s1, s2 = systems.water_cluster_300(seed=0), systems.water_cluster_300(seed=1)
    o, g = pair(W300)
    out = g.create([s1, s2])
    assert out.shape == (2, 300, o.get_number_of_features())
    g.sparse = True
    sp = g.create(s1)
    assert sp.is_sparse
    assert colerr(sp.to_dense(), o.create(s1)) < RTOL64


# ---- precision forms ----------------------------------------------------------------

def test_mixed_and_f32_precision_errors():
    s = systems.water_cluster_3000(seed=0)
    p = dict(r_cut=10.0, n_max=7, l_max=3, sigma=0.5, species=[1, 8])
    o = so.OracleSOAP(**p)
    ref = o.create(s)
    errs = {}
    for prec in ("float64", "mixed", "float32"):
        g = SOAP(device="cuda:0", dtype="float64", precision=prec, _betas=o.betas, **p)
        errs[prec] = colerr(g.create(s), ref)
    say("precision errors (n7 l3, 3000 atoms):", errs)
    assert errs["float64"] < RTOL64
    assert errs["mixed"] < 1e-4


# ---- dense Jacobian ---------------------------------------------------------------------

@pytest.mark.parametrize("average", ["off", "cc"])
def test_jacobian_closed_form_f64(average):
    s = systems.water_cluster_300(seed=0)
    o, g = pair(dict(W300, average=average))
    cen = list(range(0, 300, 10))
    say("jacobian closed form", average)
    J, d = g.derivatives(s, centers=cen, method="analytical")
    Jr, dr = o.derivatives(s, centers=cen)
    assert J.shape == Jr.shape
    assert colerr(J, Jr) < RTOL64
    assert colerr(d, dr) < RTOL64


def test_jacobian_include_exclude_f32():
    s = systems.water_cluster_300(seed=2)
    o = so.OracleSOAP(**W300)
    g = SOAP(device="cuda:0", dtype="float32", _betas=o.betas, **W300)
    inc = [250, 3, 7, 0, 100, 101, 102]
    say("jacobian include/exclude f32")
    J = g.derivatives(s, centers=[0, 100, 250], include=inc, exclude=[7], return_descriptor=False)
    Jr = o.derivatives(s, centers=[0, 100, 250], include=inc, exclude=[7], return_descriptor=False)
    assert J.dtype == torch.float32
    assert colerr(J, Jr) < 1e-6


def test_jacobian_poly_f64():
    s = systems.water_cluster_300(seed=1)
    o, g = pair(dict(W300, rbf="polynomial"))
    say("jacobian poly")
    cen = [0, 1, 2, 150]
    assert colerr(g.derivatives(s, centers=cen, return_descriptor=False),
                  o.derivatives(s, centers=cen, return_descriptor=False)) < RTOL64


def test_jacobian_general_periodic_weighting_attach():
    s = systems.water((2, 2, 2), spacing=3.0, seed=3, cell_length=6.0)
    p = dict(r_cut=4.0, n_max=3, l_max=3, sigma=0.5, species=[1, 8], periodic=True,
             weighting=dict(function="pow", r0=1.5, c=1.0, d=1.0, m=2.0, w0=0.7))
    o, g = pair(p)
    for attach in (False, True):
        say("jacobian periodic+weighting attach", attach)
        J, d = g.derivatives(s, centers=[0, 4, 9], attach=attach)
        Jr, dr = o.derivatives(s, centers=[0, 4, 9], attach=attach)
        assert colerr(J, Jr) < RTOL64
        assert colerr(d, dr) < RTOL64


def test_jacobian_mixed_f32_rows():
    """Mixed precision with float32 output: float32 edge records and rows, against float64."""
    s = systems.water_cluster_300(seed=4)
    cen = list(range(0, 300, 7))
    for average in ("off", "cc"):
        p = dict(W300, average=average)
        o = so.OracleSOAP(**p)
        g = SOAP(device="cuda:0", dtype="float32", precision="mixed", _betas=o.betas, **p)
        say("jacobian mixed f32 cluster", average)
        J, d = g.derivatives(s, centers=cen)
        Jr, dr = o.derivatives(s, centers=cen)
        assert J.dtype == torch.float32
        say("  err", colerr(J, Jr))
        assert colerr(J, Jr) < 1e-5
        assert colerr(d, dr) < 1e-5
    # periodic images sharing one row, the weighting slope and the attach row
    s = systems.water((2, 2, 2), spacing=3.0, seed=3, cell_length=6.0)
    p = dict(r_cut=4.0, n_max=3, l_max=3, sigma=0.5, species=[1, 8], periodic=True,
             weighting=dict(function="pow", r0=1.5, c=1.0, d=1.0, m=2.0, w0=0.7))
    o = so.OracleSOAP(**p)
    g = SOAP(device="cuda:0", dtype="float32", precision="mixed", _betas=o.betas, **p)
    for attach in (False, True):
        say("jacobian mixed f32 periodic attach", attach)
        J = g.derivatives(s, centers=[0, 4, 9], attach=attach, return_descriptor=False)
        Jr = o.derivatives(s, centers=[0, 4, 9], attach=attach, return_descriptor=False)
        say("  err", colerr(J, Jr))
        assert colerr(J, Jr) < 1e-5


# ---- adjoint --------------------------------------------------------------------------------

@pytest.mark.parametrize("average", ["off", "cc"])
def test_vjp_matches_oracle(average):
    s = systems.water_cluster_300(seed=7)
    o, g = pair(dict(W300, average=average))
    w = np.random.default_rng(0).normal(size=o.get_number_of_features())
    say("vjp", average)
    F = g.vjp(s, torch.tensor(w))
    Fr = o.vjp(s, w[None, :].repeat(300, 0))
    np.testing.assert_allclose(F.cpu().numpy(), Fr, atol=1e-10 * np.abs(Fr).max())


def test_vjp_periodic_weighting_poly():
    s = systems.water((3, 3, 3), spacing=3.1, seed=8, cell_length=9.3)
    p = dict(r_cut=4.0, n_max=3, l_max=2, sigma=0.5, species=[1, 8], periodic=True, rbf="polynomial",
             weighting=dict(function="exp", r0=1.0, c=1.5, d=0.5))
    o, g = pair(p)
    G = np.random.default_rng(1).normal(size=(81, o.get_number_of_features()))
    say("vjp periodic weighting poly")
    F = g.vjp(s, torch.tensor(G))
    Fr = o.vjp(s, G)
    np.testing.assert_allclose(F.cpu().numpy(), Fr, atol=1e-10 * np.abs(Fr).max())


def test_autograd_function():
    s = systems.water_cluster_300(seed=9)
    o, g = pair(W300)
    pos = torch.tensor(s["positions"], device="cuda:0", requires_grad=True)
    desc = g.create_autograd(pos, s["atomic_numbers"])
    w = torch.randn(desc.shape[1], dtype=desc.dtype, device=desc.device, generator=torch.Generator("cuda").manual_seed(0))
    (desc * w).sum().backward()
    Fr = o.vjp(s, w.cpu().numpy()[None, :].repeat(300, 0))
    np.testing.assert_allclose(pos.grad.cpu().numpy(), Fr, atol=1e-10 * np.abs(Fr).max())


# ---- basis ---------------------------------------------------------------------------------------

def test_lowdin_kernel_against_lapack():
    for (rc, n, l) in [(6.0, 4, 3), (10.0, 7, 3), (6.0, 8, 6)]:
        g = SOAP(device="cuda:0", r_cut=rc, n_max=n, l_max=l, sigma=0.5, species=[1])
        alphas, betas = so.gto_basis(rc, n, l)
        for ll in range(l + 1):
            S = so.gto_overlap(alphas[ll], ll)
            kappa = np.linalg.cond(S)
            b = g._betas[ll]
            err = np.abs(b @ S @ b - np.eye(n)).max()
            say(f"lowdin rc={rc} n={n} l={ll} cond={kappa:.1e} |bSb-I|={err:.1e}")
            assert err < 1e-14 * kappa + 1e-12
            np.testing.assert_allclose(b, betas[ll], atol=1e-15 * kappa * np.abs(betas[ll]).max())


def test_default_basis_create_matches_oracle_cfg1():
    s = systems.water_cluster_300(seed=0)
    o = so.OracleSOAP(**W300)
    g = SOAP(device="cuda:0", dtype="float64", **W300)   # beta from the device Jacobi kernel
    assert colerr(g.create(s), o.create(s)) < RTOL64


@pytest.mark.parametrize("prec,tol", [("mixed", 1e-5), ("float32", 5e-3)])
def test_vjp_precision_forms(prec, tol):
    s = systems.water_cluster_300(seed=11)
    p = dict(r_cut=6.0, n_max=6, l_max=4, sigma=0.5, species=[1, 8])
    o = so.OracleSOAP(**p)
    g = SOAP(device="cuda:0", dtype="float64", precision=prec, _betas=o.betas, **p)
    w = np.random.default_rng(2).normal(size=o.get_number_of_features())
    say("vjp precision", prec)
    d, F = g.descriptor_and_vjp(s, torch.tensor(w))
    Fr = o.vjp(s, w[None, :].repeat(300, 0))
    err = np.abs(F.cpu().numpy() - Fr).max() / np.abs(Fr).max()
    say("  force rel err", err, "desc err", colerr(d, o.create(s)))
    assert err < tol
    assert colerr(d, o.create(s)) < tol


def test_descriptor_and_vjp_periodic_weighted_f64():
    s = systems.water((3, 3, 3), spacing=3.1, seed=12, cell_length=9.3)
    p = dict(r_cut=4.0, n_max=4, l_max=3, sigma=0.5, species=[1, 8], periodic=True,
             weighting=dict(function="pow", r0=1.0, c=1.0, d=1.0, m=2.0, w0=0.5))
    o, g = pair(p)
    G = np.random.default_rng(3).normal(size=(81, o.get_number_of_features()))
    d, F = g.descriptor_and_vjp(s, torch.tensor(G))
    np.testing.assert_allclose(F.cpu().numpy(), o.vjp(s, G), atol=1e-10 * np.abs(o.vjp(s, G)).max())
    assert colerr(d, o.create(s)) < RTOL64


def _fused_d2(c, n):
    """d^2 of the reference formula with the dot product fused (fma(cz, nz, fma(cy, ny, cx nx)))."""
    from fractions import Fraction as Fr
    p = float(Fr(c[0]) * Fr(n[0]))
    p = float(Fr(c[1]) * Fr(n[1]) + Fr(p))
    dot = float(Fr(c[2]) * Fr(n[2]) + Fr(p))
    return ((dot * -2.0) + float(so.sqnorm3(c[None])[0])) + float(so.sqnorm3(n[None])[0])


def test_nl_distance_rounding_at_cutoff():
    """A pair whose separately rounded d^2 and fused-dot d^2 fall on opposite sides of cutoff^2
    must be an edge exactly when the oracle's (separately rounded) formula says so."""
    o = so.OracleSOAP(**W300)
    cut2 = o.cutoff * o.cutoff
    rng = np.random.default_rng(42)
    found = None
    for _ in range(200000):
        c = rng.uniform(20.0, 60.0, 3)
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        n = c + d * o.cutoff * (1.0 + rng.uniform(-1e-13, 1e-13))
        ci, ni = so.radius_edges(c[None], n[None], o.cutoff)
        sep_in = len(ci) == 1
        if sep_in != (_fused_d2(c, n) <= cut2):
            found = (c, n, sep_in)
            break
    assert found is not None, "no straddling pair found"
    c, n, sep_in = found
    g = SOAP(device="cuda:0", dtype="float64", _betas=o.betas, **W300)
    s = {"positions": np.stack([c, n]), "atomic_numbers": np.array([8, 1])}
    say("straddling pair: oracle edge =", sep_in)
    off, nbr = g.neighbor_list(s, drop_self=True)
    counts = np.diff(off.cpu().numpy())
    assert counts[0] == (1 if sep_in else 0)


def test_context_cache_follows_system_changes():
    """Reused species/cell/shift tensors (same composition, cell and pbc as the last call)
    must give what a fresh object gives, and any change in species, cell or pbc must be
    picked up."""
    p = dict(species=[1, 8], r_cut=4.0, n_max=4, l_max=3, sigma=0.5, periodic=True)
    s1 = systems.water_box(4, seed=11)
    g = SOAP(device="cuda:0", dtype="float64", **p)
    g.create(s1)
    d1 = g.create(dict(s1, positions=np.asarray(s1["positions"]) + 0.01))       # cached path
    ref = SOAP(device="cuda:0", dtype="float64", **p).create(dict(s1, positions=np.asarray(s1["positions"]) + 0.01))
    assert torch.equal(d1, ref)
    Z2 = np.asarray(s1["atomic_numbers"]).copy()
    Z2[[0, 1]] = Z2[[1, 0]]                                                       # same shape, other species
    s2 = dict(s1, atomic_numbers=Z2)
    assert torch.equal(g.create(s2), SOAP(device="cuda:0", dtype="float64", **p).create(s2))
    cell3 = np.asarray(s1["cell"], dtype=np.float64) * 1.05                        # other cell
    s3 = dict(s2, cell=cell3)
    assert torch.equal(g.create(s3), SOAP(device="cuda:0", dtype="float64", **p).create(s3))
    s4 = dict(s3, pbc=[True, True, False])                                          # other pbc
    assert torch.equal(g.create(s4), SOAP(device="cuda:0", dtype="float64", **p).create(s4))
