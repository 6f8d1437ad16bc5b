"""Chebyshev radial form (radial="chebyshev"): the matrix-core coefficient kernel
(fp16 hi/lo products, float32 accumulation) against the float64 oracle.

Tolerances: descriptors within 1e-5 relative per feature column (max over columns of
max|d - ref| / max|ref|; 2e-5 for H/C/N/O) and forces within 2e-5 of max|F|.  The
series itself is exact to ~1e-11 of each radial function; what remains is the float32
operands (t, T_j, Y_lm) and the fp16 hi/lo products (~22 bits per operand), an absolute
error of ~1e-7 of each radial function's maximum per edge.  A column whose values are
far below that scale on every center sees it as a larger relative error, so centre
subsets are measured against the column scale of the whole system.
"""

import sys

import numpy as np
import pytest
import torch

from cusoap_amd import SOAP, systems
from oracle import soap_oracle as so

pytestmark = pytest.mark.gpu

DEV = "cuda:0"
WB = dict(r_cut=6.0, n_max=8, l_max=6, sigma=0.5, species=[1, 8])
HCNO = dict(r_cut=5.0, n_max=8, l_max=6, sigma=0.4, species=[1, 6, 7, 8], periodic=True)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def colerr(a, b):
    a = a.detach().cpu().double().numpy() if isinstance(a, torch.Tensor) else np.asarray(a)
    b = np.asarray(b, dtype=np.float64)
    a, b = a.reshape(-1, a.shape[-1]), b.reshape(-1, b.shape[-1])
    scale = np.abs(b).max(axis=0)
    return float((np.abs(a - b).max(axis=0) / np.where(scale > 0, scale, 1.0)).max())


@pytest.mark.parametrize("n_max,l_max", [(8, 6), (7, 3), (6, 5), (5, 4), (4, 2), (3, 0), (8, 1)])
def test_chebyshev_create_and_forces_water_box(n_max, l_max):
    s = systems.water_box(5, seed=4)
    p = dict(WB, periodic=True, n_max=n_max, l_max=l_max)
    o = so.OracleSOAP(**p)
    say(f"chebyshev create n{n_max} l{l_max}, water box {len(s['positions'])} atoms")
    g = SOAP(device=DEV, dtype="float64", precision="mixed", radial="chebyshev", _betas=o.betas, **p)
    w = np.random.default_rng(5).normal(size=o.get_number_of_features())
    ref = o.create(s)
    Fr = o.vjp(s, w[None, :].repeat(len(s["positions"]), 0))
    d = g.create(s)
    d2, F = g.descriptor_and_vjp(s, torch.tensor(w))
    ed, ef = colerr(d, ref), np.abs(F.cpu().numpy() - Fr).max() / np.abs(Fr).max()
    say(f"  K {g._cheb_terms}, fit {g._cheb_fit_err}, desc err {ed:.2e}, force err {ef:.2e}")
    assert ed < 1e-5 and ef < 2e-5
    assert colerr(d2, ref) < 1e-5


def subset_err(d, ref, ref_all):
    d = d.detach().cpu().double().numpy()
    sc = np.maximum(np.abs(ref).max(0), np.abs(ref_all).max(0))
    return float((np.abs(d - ref).max(0) / np.where(sc > 0, sc, 1.0)).max())


def test_chebyshev_cluster_and_coordinate_centers():
    """Non-periodic cluster, centers mixing atom indices and coordinates (the coordinate-center
    slot), and the create path's self term."""
    s = systems.water_cluster_300(seed=2)
    p = dict(r_cut=6.0, n_max=7, l_max=3, sigma=0.5, species=[1, 8])
    o = so.OracleSOAP(**p)
    g = SOAP(device=DEV, dtype="float64", precision="mixed", radial="chebyshev", _betas=o.betas, **p)
    cen = [0, 17, [1.0, 2.0, 3.0], 299, [4.25, 0.5, 7.75], 5]
    ref_all = o.create(s)
    e_all = colerr(g.create(s), ref_all)
    e_cen = subset_err(g.create(s, centers=cen), o.create(s, centers=cen), ref_all)
    say(f"chebyshev cluster: desc {e_all:.2e}, mixed centers {e_cen:.2e}")
    assert e_all < 1e-5 and e_cen < 1e-5


def test_chebyshev_four_species():
    """H/C/N/O basis: four species, eight (species, interval) groups per tile, masked blocks."""
    s = systems.random_box(n_atoms=400, density=0.1, seed=3)
    o = so.OracleSOAP(**HCNO)
    g = SOAP(device=DEV, dtype="float64", precision="mixed", radial="chebyshev", _betas=o.betas, **HCNO)
    cen = [0, 17, [1.0, 2.0, 3.0], 399, [10.25, 0.5, 7.75], 5]
    w = np.random.default_rng(8).normal(size=o.get_number_of_features())
    Fr = o.vjp(s, w[None, :].repeat(len(s["positions"]), 0))
    ref_all = o.create(s)
    d, F = g.descriptor_and_vjp(s, torch.tensor(w))
    ed, ef = colerr(d, ref_all), np.abs(F.cpu().numpy() - Fr).max() / np.abs(Fr).max()
    ec = subset_err(g.create(s, centers=cen), o.create(s, centers=cen), ref_all)
    say(f"chebyshev H/C/N/O: desc {ed:.2e}, forces {ef:.2e}, mixed centers {ec:.2e}")
    assert ed < 2e-5 and ef < 2e-5 and ec < 2e-5


def test_chebyshev_centers_with_several_rounds():
    """r_cut 9 (cutoff 10.9 A): ~540 edges per center, so the kernel runs several 256-edge rounds per center
    (groups accumulated across rounds, the contraction scale from the whole center's counts);
    create and forces against the float64 oracle."""
    s = systems.water_box(6, seed=6)
    p = dict(r_cut=9.0, n_max=4, l_max=3, sigma=0.5, species=[1, 8], periodic=True)
    o = so.OracleSOAP(**p)
    g = SOAP(device=DEV, dtype="float64", precision="mixed", radial="chebyshev", _betas=o.betas, **p)
    nl = g._neighbors(g._context(s, None), drop_self=False)
    per_center = nl.n_edges / len(s["positions"])
    w = np.random.default_rng(7).normal(size=o.get_number_of_features())
    ref = o.create(s)
    Fr = o.vjp(s, w[None, :].repeat(len(s["positions"]), 0))
    d, F = g.descriptor_and_vjp(s, torch.tensor(w))
    ed, ef = colerr(d, ref), np.abs(F.cpu().numpy() - Fr).max() / np.abs(Fr).max()
    say(f"chebyshev r_cut 9: {per_center:.0f} edges per center (max {nl.max_count}), desc {ed:.2e}, forces {ef:.2e}")
    assert nl.max_count > 2 * 256 and ed < 1e-5 and ef < 2e-5
