"""Power spectrum on the matrix cores against the float64 contraction of the same
coefficients, and the full descriptor against the float64 oracle.

Per-feature error = max_c |p - ref| / max_c |ref| for each feature column; the
report prints the worst column and the median column.  Bounds: fp16 2e-3,
bf16 2e-2 (8-bit mantissa), hi/lo split fp16x3 1e-5 and bf16x3 1e-4.
"""

import sys

import numpy as np
import pytest
import torch

from cusoap_amd import SOAP, systems
from oracle import soap_oracle as so

pytestmark = pytest.mark.gpu

BOUND = {"fp16": 2e-3, "bf16": 2e-2, "fp16x3": 1e-5, "bf16x3": 1e-4}


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def colerrs(a, b):
    a = a.detach().cpu().double().numpy()
    b = b.detach().cpu().double().numpy() if isinstance(b, torch.Tensor) else np.asarray(b)
    scale = np.abs(b).max(axis=0)
    e = np.abs(a - b).max(axis=0) / np.where(scale > 0, scale, 1.0)
    return float(e.max()), float(np.median(e))


CASES = [
    ("water n8 l6", dict(r_cut=6.0, n_max=8, l_max=6, sigma=0.5, species=[1, 8]), "water"),
    ("water n7 l3", dict(r_cut=10.0, n_max=7, l_max=3, sigma=0.5, species=[1, 8]), "water"),
    ("hcno n8 l6", dict(r_cut=5.0, n_max=8, l_max=6, sigma=0.4, species=[1, 6, 7, 8], periodic=True), "hcno"),
    ("water n4 l8 no crossover", dict(r_cut=6.0, n_max=4, l_max=8, sigma=0.5, species=[1, 8], crossover=False),
     "water"),
]


@pytest.mark.parametrize("name,params,kind", CASES, ids=[c[0] for c in CASES])
def test_mfma_spectrum_against_float64_contraction(name, params, kind):
    s = systems.water_cluster_300(seed=2) if kind == "water" else systems.random_box(n_atoms=400, density=0.10,
                                                                                      seed=3)
    ref_soap = SOAP(device="cuda:0", dtype="float64", **params)
    coeffs = ref_soap.coefficients(s)
    ref = ref_soap._descriptor(coeffs)
    report = {}
    for mode, bound in BOUND.items():
        g = SOAP(device="cuda:0", dtype="float64", spectrum=mode, _betas=ref_soap._betas, **params)
        worst, med = colerrs(g._descriptor(coeffs), ref)
        report[mode] = (f"{worst:.2e}", f"{med:.2e}")
        assert worst < bound, (mode, worst)
    say(name, "per-feature error (worst, median):", report)


def test_mfma_descriptor_against_oracle():
    s = systems.water_cluster_300(seed=4)
    p = dict(r_cut=6.0, n_max=4, l_max=3, sigma=0.5, species=[1, 8])
    o = so.OracleSOAP(**p)
    ref = o.create(s)
    for mode in ("fp16x3", "bf16x3"):
        g = SOAP(device="cuda:0", dtype="float64", spectrum=mode, _betas=o.betas, **p)
        worst, _ = colerrs(g.create(s), ref)
        say("create", mode, "vs oracle", worst)
        assert worst < BOUND[mode]
