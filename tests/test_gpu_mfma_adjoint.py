"""Matrix-core power-spectrum adjoint (float32 coefficients, shared descriptor gradient).

The mixed and radial-table forms form dL/dc = W_l c on the matrix cores
(v_mfma_f32_16x16x4_f32); the vector kernel (CUSOAP_PS_ADJOINT=vector) computes the same
float32 products in another order.  Both are held against each other (float32 sums of
up to S n_max terms: 1e-6 relative) and the forces against the float64 oracle at the
mixed form's 1e-5 bar.  Species x n_max of 16 (one row tile, four K steps), 20 (two row
tiles, padded K) and 6 (one padded tile) are covered.
"""

import os
import sys

import numpy as np
import pytest
import torch

from cusoap_amd import SOAP, systems
from oracle import soap_oracle as so

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def forces(g, s, w, kernel):
    old = os.environ.get("CUSOAP_PS_ADJOINT")
    try:
        if kernel == "vector":
            os.environ["CUSOAP_PS_ADJOINT"] = "vector"
        else:
            os.environ.pop("CUSOAP_PS_ADJOINT", None)
        d, F = g.descriptor_and_vjp(s, w)
        torch.cuda.synchronize()
        return d, F
    finally:
        if old is None:
            os.environ.pop("CUSOAP_PS_ADJOINT", None)
        else:
            os.environ["CUSOAP_PS_ADJOINT"] = old


CASES = [
    ("water300_n8l6_table", lambda: systems.water_cluster_300(seed=4),
     dict(r_cut=5.0, n_max=8, l_max=6, sigma=0.5, species=[1, 8]), dict(radial="table")),
    ("hcno_n5l3_mixed", lambda: systems.random_box(400, seed=2),
     dict(r_cut=4.0, n_max=5, l_max=3, sigma=0.4, species=[1, 6, 7, 8], periodic=True), {}),
    ("water300_n3l2_mixed", lambda: systems.water_cluster_300(seed=5),
     dict(r_cut=4.5, n_max=3, l_max=2, sigma=0.5, species=[1, 8]), {}),
]


@pytest.mark.parametrize("name,make,p,kw", CASES, ids=[c[0] for c in CASES])
def test_mfma_ps_adjoint(name, make, p, kw):
    s = make()
    o = so.OracleSOAP(**p)
    g = SOAP(device="cuda:0", dtype="float32", precision="mixed", _betas=o.betas, **p, **kw)
    w = torch.tensor(np.random.default_rng(3).normal(size=o.get_number_of_features()), device="cuda:0")
    say(name, "atoms", s["positions"].shape[0])
    d1, F1 = forces(g, s, w, "mfma")
    d2, F2 = forces(g, s, w, "vector")
    assert torch.equal(d1, d2)
    dv = (F1 - F2).abs().max().item() / F2.abs().max().item()
    Fr = o.vjp(s, np.repeat(w.cpu().numpy()[None, :], s["positions"].shape[0], 0))
    err = np.abs(F1.cpu().numpy() - Fr).max() / np.abs(Fr).max()
    say("  mfma vs vector", dv, "mfma vs oracle", err)
    assert dv < 1e-6
    assert err < 1e-5
