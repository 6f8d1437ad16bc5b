"""cusoap_vjp_noalloc (static scratch, no allocation) against cusoap_vjp, through the C ABI.

Two calls in a row on one GPU: first a 24,000-atom periodic water box whose adjoint rows
(n_max 8, l_max 6, two species) need three chunks of the 64 MiB scratch, then a
375-atom box (one chunk), each compared with the workspace entry point on the same
inputs.  The float64 form sums every contribution in float64 (tolerance 1e-12 of
max|F|); the table form sums float contributions (1e-5 of max|F|, against the 1e-4
bound of the mixed forms).
"""

import ctypes as C
import sys

import numpy as np
import pytest
import torch

from cusoap_amd import SOAP, _lib, systems

pytestmark = pytest.mark.gpu
DEV = "cuda:0"
P = dict(r_cut=6.0, n_max=8, l_max=6, sigma=0.5, species=[1, 8], periodic=True)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _both(g, s, grad_desc, centers=None, attach=False):
    """Forces from cusoap_vjp (via the SOAP object) and from cusoap_vjp_noalloc on the same inputs."""
    lib = _lib.load()
    x = g._context(s, centers)
    nl = g._neighbors(x, drop_self=False)
    coeffs = g._vjp_coefficients(x, nl)
    gd = grad_desc.to(DEV).contiguous()
    F_ws = g._vjp_ctx(x, gd, centers, attach, coeffs=coeffs, nl_keep=nl)
    att = g._attach_array(x, centers, attach)
    F_na = torch.full((x.n_atoms, 3), float("nan"), dtype=torch.float64, device=DEV)
    b = g._basis()
    rc = lib.cusoap_vjp_noalloc(
        C.byref(b), _lib.MODE_PS, {"float64": _lib.PREC_F64, "mixed": _lib.PREC_MIXED}[g.precision],
        x.pos.data_ptr(), x.n_atoms, x.sp.data_ptr(), x.shifts.data_ptr(), x.n_img, x.centers.data_ptr(),
        x.n_centers, att.data_ptr() if att is not None else None, x.self_sp.data_ptr(), nl.offsets.data_ptr(),
        nl.nbr.data_ptr(), coeffs.data_ptr(), _lib.DT_F64 if coeffs.dtype == torch.float64 else _lib.DT_F32,
        g._keep["pf"].data_ptr(), gd.data_ptr(), _lib.DT_F64 if gd.dtype == torch.float64 else _lib.DT_F32,
        1 if gd.dim() == 1 else 0, g.get_number_of_features(), F_na.data_ptr(), g._stream())
    _lib.check(rc, "vjp_noalloc")
    torch.cuda.synchronize()
    return F_ws.cpu().numpy(), F_na.cpu().numpy()


def _rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


@pytest.mark.parametrize("precision,radial,tol", [("float64", "exact", 1e-12), ("mixed", "table", 1e-5)])
def test_noalloc_vjp_matches_workspace_entry(precision, radial, tol):
    g = SOAP(device=DEV, dtype="float64", precision=precision, radial=radial, **P)
    nf = g.get_number_of_features()
    big = systems.water_box(20, seed=21)                       # 24,000 atoms: three scratch chunks
    rows = 8 * 2 * 8 * 49
    assert len(big["positions"]) * rows > 2 * (64 << 20)
    w = torch.tensor(np.random.default_rng(1).normal(size=nf))
    say(f"noalloc vjp, {precision}/{radial}: 24,000-atom water box, shared gradient")
    F_ws, F_na = _both(g, big, w)
    say("  rel diff", _rel(F_na, F_ws))
    assert np.isfinite(F_na).all() and _rel(F_na, F_ws) < tol
    small = systems.water_box(5, seed=22)                      # second call, smaller input: one chunk
    G = torch.tensor(np.random.default_rng(2).normal(size=(len(small["positions"]), nf)))
    say(f"noalloc vjp, {precision}/{radial}: 375-atom water box, per-center gradient")
    F_ws, F_na = _both(g, small, G)
    say("  rel diff", _rel(F_na, F_ws))
    assert np.isfinite(F_na).all() and _rel(F_na, F_ws) < tol


def test_noalloc_vjp_explicit_centers_attach():
    """Index centers that move with their atoms (center_attach) on the chunked path's offsets."""
    g = SOAP(device=DEV, dtype="float64", precision="float64", **P)
    s = systems.water_box(5, seed=23)
    cen = [3, 100, 7, 250, 374, 0]
    G = torch.tensor(np.random.default_rng(3).normal(size=(len(cen), g.get_number_of_features())))
    say("noalloc vjp: explicit centers with attach")
    F_ws, F_na = _both(g, s, G, centers=cen, attach=True)
    say("  rel diff", _rel(F_na, F_ws))
    assert _rel(F_na, F_ws) < 1e-12
