"""The benchmark configurations on the GPU against the float64 oracle.

- (H2O)_1000 sweep, n_max 1..7 x l_max 0..3 (r_cut 10, sigma 0.5): create in float64
  (1e-10), mixed and plain float32 against the oracle, and the dense float32 Jacobian of
  the plain float32 form on 16 centers x all atoms at three sweep points.  Stated bounds
  for the plain float32 form: 1e-4 on descriptors and on Jacobians (its Loewdin
  contraction is done in float32; cond(S_l) stays below ~1e6 at r_cut 10, n_max <= 7).
- The adjoint at the 1,000,002-atom cluster's basis (n_max 7, l_max 3, r_cut 6,
  non-periodic) on its first 1,000 molecules (3,000 atoms): forces of a Gaussian readout
  in float64 (1e-10), mixed + table and plain float32 (1e-4) against the oracle's adjoint,
  against central differences (h = 1e-4 A) of the float64 descriptor on 64 seeded atoms,
  and against the GPU float64 dense Jacobian contracted with the readout.
- The adjoint with index centers that move with their atoms (attach=True), against the
  oracle Jacobian contracted with a per-center gradient.
- cc features and cc Jacobians at l_max 6; the device Loewdin matrices at n_max 8,
  l_max 6; the neighbour fill of centers with more than 4096 edges.

Per-feature errors are max over rows of |gpu - ref| divided by max |ref| of the column;
force errors max |F - F_ref| / max |F_ref|.
"""

import math
import sys

import numpy as np
import pytest
import torch

from cusoap_amd import SOAP, systems
from oracle import soap_oracle as so

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def colerr(a, b):
    a = a.detach().cpu().double().numpy() if isinstance(a, torch.Tensor) else np.asarray(a)
    b = np.asarray(b, dtype=np.float64)
    a, b = a.reshape(-1, a.shape[-1]), b.reshape(-1, b.shape[-1])
    scale = np.abs(b).max(axis=0)
    return float((np.abs(a - b).max(axis=0) / np.where(scale > 0, scale, 1.0)).max())


def ferr(F, Fr):
    F = F.detach().cpu().numpy() if isinstance(F, torch.Tensor) else np.asarray(F)
    return float(np.abs(F - Fr).max() / np.abs(Fr).max())


# ---- (H2O)_1000 sweep ----------------------------------------------------------------------

CL = systems.water_cluster_3000(seed=0)
SWEEP = dict(r_cut=10.0, sigma=0.5, species=[1, 8])
_SWEEP_CACHE = {}


def _sweep_oracle(n):
    """Oracle coefficients of the whole cluster at (n, l_max 3); lower l_max are slices."""
    if n not in _SWEEP_CACHE:
        o = so.OracleSOAP(n_max=n, l_max=3, **SWEEP)
        _, c = o.create(CL, return_coeffs=True)
        _SWEEP_CACHE[n] = (o, c)
    return _SWEEP_CACHE[n]


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6, 7])
def test_sweep_create_against_oracle(n):
    o3, c3 = _sweep_oracle(n)
    for l in range(4):
        ol = so.OracleSOAP(n_max=n, l_max=l, **SWEEP)
        ref = ol.descriptor(np.ascontiguousarray(c3[..., :(l + 1) ** 2]))
        errs = {}
        for prec in ("float64", "mixed", "float32"):
            g = SOAP(device=DEV, dtype="float64" if prec == "float64" else "float32", precision=prec,
                     n_max=n, l_max=l, _betas=o3.betas[:l + 1], **SWEEP)
            errs[prec] = colerr(g.create(CL), ref)
        say(f"sweep n{n} l{l}:", {k: f"{v:.1e}" for k, v in errs.items()})
        assert errs["float64"] < 1e-10
        assert errs["mixed"] < 1e-5
        assert errs["float32"] < 1e-4


@pytest.mark.parametrize("n,l", [(1, 0), (4, 2), (7, 3)])
def test_sweep_dense_jacobian_float32(n, l):
    o = so.OracleSOAP(n_max=n, l_max=l, **SWEEP)
    cen = list(range(0, 3000, 188))[:16]
    J_ref, d_ref = o.derivatives(CL, centers=cen)
    for prec, tol in (("float32", 1e-4), ("mixed", 1e-5)):
        g = SOAP(device=DEV, dtype="float32", precision=prec, n_max=n, l_max=l, _betas=o.betas, **SWEEP)
        J, d = g.derivatives(CL, centers=cen, method="analytical")
        ej, ed = colerr(J, J_ref), colerr(d, d_ref)
        say(f"sweep jacobian n{n} l{l} {prec}: jacobian {ej:.1e} descriptor {ed:.1e}")
        assert J.shape == J_ref.shape and J.dtype == torch.float32
        assert ej < tol and ed < tol


# ---- adjoint at the 1M-atom cluster's basis ----------------------------------------------------

P5 = dict(r_cut=6.0, n_max=7, l_max=3, sigma=0.5, species=[1, 8])
SUB = systems.water((70, 70, 69), seed=0, n_mol=1000)        # first 1,000 molecules of the 1M cluster


def _oracle_vjp(o, s, w, chunk=500):
    n = len(s["positions"])
    return sum(o.vjp(s, w, centers=list(range(c0, min(c0 + chunk, n)))) for c0 in range(0, n, chunk))


@pytest.fixture(scope="module")
def cfg5():
    o = so.OracleSOAP(**P5)
    w = np.random.default_rng(5).normal(size=o.get_number_of_features())
    say("oracle adjoint, 3,000-atom sub-cluster, n7 l3")
    return o, w, _oracle_vjp(o, SUB, w)


@pytest.mark.parametrize("prec,radial,tol", [("float64", "exact", 1e-10), ("mixed", "table", 1e-4),
                                             ("float32", "exact", 1e-4)])
def test_cfg5_adjoint_against_oracle(cfg5, prec, radial, tol):
    o, w, F_ref = cfg5
    g = SOAP(device=DEV, dtype="float32", precision=prec, radial=radial, _betas=o.betas, **P5)
    F = g.vjp(SUB, torch.tensor(w))
    say(f"cfg5 adjoint {prec}/{radial}: force err {ferr(F, F_ref):.2e}")
    assert ferr(F, F_ref) < tol


def test_cfg5_adjoint_against_finite_differences(cfg5):
    """Central differences of L = sum_c w.p_c in float64 with the centers fixed (the
    convention of the adjoint), h = 1e-4 A, on 64 seeded atoms; only centers within the
    cutoff (plus h) of the moved atom change, so only they are recomputed."""
    o, w, _ = cfg5
    g = SOAP(device=DEV, dtype="float64", precision="float64", _betas=o.betas, **P5)
    wt = torch.tensor(w, device=DEV)
    pos = np.asarray(SUB["positions"])
    atoms = np.random.default_rng(64).choice(len(pos), 64, replace=False)
    h = 1e-4
    F_fd = np.zeros((64, 3))
    for i, a in enumerate(atoms):
        near = np.flatnonzero(((pos - pos[a]) ** 2).sum(1) <= (g._cutoff + 2 * h) ** 2)
        cen = [pos[c].copy() for c in near]
        for d in range(3):
            Ls = []
            for sgn in (1.0, -1.0):
                p = pos.copy()
                p[a, d] += sgn * h
                Ls.append(float((g.create({"positions": p, "atomic_numbers": SUB["atomic_numbers"]},
                                          centers=cen) * wt).sum()))
            F_fd[i, d] = (Ls[0] - Ls[1]) / (2 * h)
    for prec, radial, tol in (("float64", "exact", 1e-7), ("mixed", "table", 1e-4), ("float32", "exact", 1e-4)):
        ga = SOAP(device=DEV, dtype="float32", precision=prec, radial=radial, _betas=o.betas, **P5)
        F = ga.vjp(SUB, wt).cpu().numpy()[atoms]
        e = float(np.abs(F - F_fd).max() / np.abs(F_fd).max())
        say(f"cfg5 adjoint vs central differences ({prec}/{radial}): {e:.2e}")
        assert e < tol


def test_cfg5_adjoint_against_dense_jacobian(cfg5):
    """The adjoint against the GPU float64 dense Jacobian (itself pinned to the oracle at
    1e-10) contracted with the readout, center block by center block."""
    o, w, _ = cfg5
    g = SOAP(device=DEV, dtype="float64", precision="float64", _betas=o.betas, **P5)
    wt = torch.tensor(w, device=DEV)
    n = len(SUB["positions"])
    F_dense = torch.zeros((n, 3), dtype=torch.float64, device=DEV)
    for c0 in range(0, n, 250):
        J = g.derivatives(SUB, centers=list(range(c0, min(c0 + 250, n))), method="analytical",
                          return_descriptor=False)
        F_dense += torch.einsum("cadf,f->ad", J, wt)
        del J
    F_dense = F_dense.cpu().numpy()
    for prec, radial, tol in (("float64", "exact", 1e-10), ("mixed", "table", 1e-4), ("float32", "exact", 1e-4)):
        ga = SOAP(device=DEV, dtype="float32", precision=prec, radial=radial, _betas=o.betas, **P5)
        e = ferr(ga.vjp(SUB, wt), F_dense)
        say(f"cfg5 adjoint vs dense Jacobian ({prec}/{radial}): {e:.2e}")
        assert e < tol


# ---- adjoint with attached centers -----------------------------------------------------------

W300 = dict(r_cut=6.0, n_max=4, l_max=3, sigma=0.5, species=[1, 8])
_ATTACH_REF = {}


@pytest.mark.parametrize("form", [("float64", "exact", "gto", 1e-10), ("mixed", "exact", "gto", 1e-5),
                                  ("mixed", "table", "gto", 2e-5), ("mixed", "chebyshev", "gto", 2e-5),
                                  ("float64", "exact", "polynomial", 1e-10)])
@pytest.mark.parametrize("centers", [None, [5, [1.0, 2.0, 3.0], 17, 250, [0.5, 0.5, 0.5], 0]])
def test_vjp_attach_against_oracle_jacobian(form, centers):
    prec, radial, rbf, tol = form
    s = systems.water_cluster_300(seed=13)
    p = dict(W300, rbf=rbf)
    o = so.OracleSOAP(**p)
    extra = {"_betas": o.betas} if rbf == "gto" else {}
    g = SOAP(device=DEV, dtype="float64", precision=prec, radial=radial, **p, **extra)
    n_c = 300 if centers is None else len(centers)
    G = np.random.default_rng(14).normal(size=(n_c, o.get_number_of_features()))
    key = (rbf, centers is None)
    if key not in _ATTACH_REF:
        J = o.derivatives(s, centers=centers, attach=True, return_descriptor=False)
        _ATTACH_REF[key] = np.einsum("cadf,cf->ad", J, G)
        del J
    F_ref = _ATTACH_REF[key]
    F = g.vjp(s, torch.tensor(G), centers=centers, attach=True)
    say(f"vjp attach {prec}/{radial}/{rbf} centers={'all' if centers is None else len(centers)}: "
        f"{ferr(F, F_ref):.2e}")
    assert ferr(F, F_ref) < tol


# ---- cc at l_max 6, device Loewdin at n8 l6, large neighbour counts ---------------------------------

def test_cc_create_and_jacobian_l6():
    s = systems.water_cluster_300(seed=15)
    p = dict(r_cut=5.0, n_max=3, l_max=6, sigma=0.5, species=[1, 8], average="cc")
    o = so.OracleSOAP(**p)
    g = SOAP(device=DEV, dtype="float64", _betas=o.betas, **p)
    assert colerr(g.create(s), o.create(s)) < 1e-10
    cen = [0, 7, 151, 299]
    J, d = g.derivatives(s, centers=cen)
    J_ref, d_ref = o.derivatives(s, centers=cen)
    say(f"cc l6: jacobian {colerr(J, J_ref):.1e}")
    assert colerr(J, J_ref) < 1e-10 and colerr(d, d_ref) < 1e-10


def test_default_loewdin_n8_l6_create_and_vjp():
    """The shipped default: Loewdin matrices from the device Jacobi kernel (no oracle beta
    injected) at n_max 8, l_max 6, where cond(S_l) reaches ~1e9.  The two S^-1/2 agree to
    about cond(S) eps, so the bound is 1e-14 cond(S) (2e-5 here), not 1e-10."""
    s = systems.random_box(n_atoms=400, density=0.1, seed=3)
    p = dict(r_cut=5.0, n_max=8, l_max=6, sigma=0.4, species=[1, 6, 7, 8], periodic=True)
    o = so.OracleSOAP(**p)
    kappa = max(np.linalg.cond(so.gto_overlap(o.alphas[l], l)) for l in range(7))
    g = SOAP(device=DEV, dtype="float64", **p)
    ed = colerr(g.create(s), o.create(s))
    w = np.random.default_rng(16).normal(size=o.get_number_of_features())
    ef = ferr(g.vjp(s, torch.tensor(w)), o.vjp(s, w))
    say(f"default beta n8 l6: cond {kappa:.1e}, descriptor {ed:.1e}, forces {ef:.1e}")
    assert ed < 1e-14 * kappa and ef < 1e-14 * kappa


def test_nl_bit_exact_over_4096_neighbours():
    """Centers with more than 4096 edges (the 16384-slot fill)."""
    s = systems.water_box(20, seed=17)                       # 24,000 atoms, L = 62.1 A
    p = dict(r_cut=22.0, n_max=2, l_max=0, sigma=0.5, species=[1, 8], periodic=True)
    o = so.OracleSOAP(**p)
    g = SOAP(device=DEV, dtype="float64", _betas=o.betas, **p)
    cen = [0, 1, 2, 5000, 12345, 23999]
    off, nbr = g.neighbor_list(s, centers=cen, drop_self=True)
    counts = np.diff(off.cpu().numpy())
    say("nl counts", counts)
    assert counts.min() > 4096
    cx, _ = o.prepare_centers(s["positions"], cen)
    ci, ni, _, _ = o.neighbour_list(s["positions"], s["atomic_numbers"], cx, s["cell"], s["pbc"], drop_self=True)
    np.testing.assert_array_equal(counts, np.bincount(ci, minlength=len(cen)))
    np.testing.assert_array_equal(nbr.cpu().numpy(), ni)
