"""Learned radial model on the GPU: kernels against the CPU restatement
(oracle/mlp_oracle.py), the trained model against the exact radial functions
(its gate), and the learned descriptor path against the float64 oracle.

Tolerances: float32 network kernels vs the float64 restatement at 1e-5 of the
output scale (2e-4 for gradients, which sum float32 products over a tile).
The trained networks must pass the fit stage of their gate: held-out error of g_nl
<= 2e-5 and of its slope <= 1e-4, relative to each function's maximum.  Learned
descriptors and forces are held to 1e-4 relative against the float64 oracle (max
over feature columns of max|d - ref| / max|ref|; forces max|F - ref| / max|ref|):
the end-to-end stage of the gate returns pairs to the exact function until a
calibration system meets half of that.  The plain float32 method misses 1e-4 on
descriptors by 8x, and the test requires the learned form to beat it there.
"""

import ctypes as C
import sys

import numpy as np
import pytest
import torch

from cusoap_amd import SOAP, _lib, systems
from cusoap_amd.surrogate import HIDDEN, exact_radial
from oracle import mlp_oracle as mo
from oracle import soap_oracle as so

pytestmark = pytest.mark.gpu

WB = dict(r_cut=6.0, n_max=8, l_max=6, sigma=0.5, species=[1, 8])
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


def _mix32(h):
    h = np.uint64(h) & np.uint64(0xFFFFFFFF)
    h ^= h >> np.uint64(16)
    h = (h * np.uint64(0x85EBCA6B)) & np.uint64(0xFFFFFFFF)
    h ^= h >> np.uint64(13)
    h = (h * np.uint64(0xC2B2AE35)) & np.uint64(0xFFFFFFFF)
    h ^= h >> np.uint64(16)
    return h


def _sample_index(seed, step, b, n):
    """The minibatch rule of the training kernel (murmur3 finaliser of seed, step, position)."""
    lo = (np.uint64(step) * np.uint64(0x9E3779B9) + np.uint64(step >> 32)) & np.uint64(0xFFFFFFFF)
    h = _mix32(np.uint64(seed) ^ _mix32(lo))
    h = _mix32(h ^ ((np.uint64(b) * np.uint64(0x27D4EB2F) + np.uint64(0x165667B1)) & np.uint64(0xFFFFFFFF)))
    return int(h % np.uint64(n))


def _t(a, dtype=torch.float32):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype, device=DEV)


def test_radial_exact_matches_oracle():
    o = so.OracleSOAP(**WB)
    g = SOAP(device=DEV, _betas=o.betas, **WB)
    rng = np.random.default_rng(0)
    r = np.concatenate([[0.0, 1e-7], np.linspace(0.0, o.cutoff, 501), rng.uniform(0.0, o.cutoff, 500)])
    B, dB = exact_radial(g, r)
    Bo, Bho = o.radial_B_Bhat(r)                                    # (P, n, l)
    Bo = Bo.transpose(0, 2, 1).reshape(len(r), -1)
    dBo = (Bho * r[:, None, None]).transpose(0, 2, 1).reshape(len(r), -1)
    assert colerr(B, Bo) < 1e-10
    assert colerr(dB, dBo) < 1e-10


def test_mlp_eval_and_loss_grad_match_oracle():
    H, O, n_pairs, n_samples = HIDDEN, 56, 3, 1000
    rng = np.random.default_rng(1)
    params = np.stack([mo.init_params(H, O, rng) for _ in range(n_pairs)]).astype(np.float32)
    x = rng.uniform(-1, 1, n_samples).astype(np.float32)
    t = rng.normal(size=(n_pairs, n_samples, O)).astype(np.float32)
    dt = rng.normal(size=(n_pairs, n_samples, O)).astype(np.float32)
    wv = rng.uniform(0.5, 1.5, O).astype(np.float32)
    wd = rng.uniform(0.01, 0.1, O).astype(np.float32)
    lib = _lib.load()
    st = torch.cuda.current_stream().cuda_stream
    pd, xd, td, dtd, wvd, wdd = _t(params), _t(x), _t(t), _t(dt), _t(wv), _t(wd)
    # inference
    y = torch.empty((n_pairs, n_samples, O), dtype=torch.float32, device=DEV)
    dy = torch.empty_like(y)
    _lib.check(lib.cusoap_mlp_eval(pd.data_ptr(), n_pairs, H, O, xd.data_ptr(), n_samples, y.data_ptr(),
                                   dy.data_ptr(), st), "mlp_eval")
    for q in range(n_pairs):
        yo, dyo, _ = mo.forward(params[q].astype(np.float64), x.astype(np.float64), H, O)
        assert np.abs(y[q].cpu().numpy() - yo).max() < 1e-5 * np.abs(yo).max()
        assert np.abs(dy[q].cpu().numpy() - dyo).max() < 1e-5 * np.abs(dyo).max()
    # loss and gradient on a hashed minibatch, per-pair targets
    batch, seed, step = 96, 7, 5
    nbytes = C.c_size_t(0)
    _lib.check(lib.cusoap_mlp_train_workspace_size(n_pairs, H, O, batch, C.byref(nbytes)), "ws")
    ws = torch.empty(int(nbytes.value), dtype=torch.uint8, device=DEV)
    grad = torch.empty_like(pd)
    loss = torch.empty(n_pairs, dtype=torch.float32, device=DEV)
    _lib.check(lib.cusoap_mlp_loss_grad(pd.data_ptr(), n_pairs, H, O, xd.data_ptr(), td.data_ptr(), dtd.data_ptr(),
                                        n_samples, n_samples * O, wvd.data_ptr(), wdd.data_ptr(), batch, seed, step,
                                        ws.data_ptr(), ws.numel(), grad.data_ptr(), loss.data_ptr(), st),
               "mlp_loss_grad")
    idx = np.array([_sample_index(seed, step, b, n_samples) for b in range(batch)])
    for q in range(n_pairs):
        L, gref = mo.loss_and_grad(params[q].astype(np.float64), x[idx].astype(np.float64),
                                   t[q][idx].astype(np.float64), dt[q][idx].astype(np.float64),
                                   wv.astype(np.float64), wd.astype(np.float64), H, O)
        gq = grad[q].cpu().numpy()
        say(f"pair {q}: loss {loss[q].item():.6g} vs {L:.6g}; grad err {np.abs(gq - gref).max() / np.abs(gref).max():.2e}")
        assert abs(loss[q].item() - L) < 1e-5 * L
        assert np.abs(gq - gref).max() < 2e-4 * np.abs(gref).max()


def test_adam_matches_oracle():
    rng = np.random.default_rng(2)
    n = 5000
    p = rng.normal(size=n)
    m, v = np.zeros(n), np.zeros(n)
    pd, md, vd = _t(p), _t(m), _t(v)
    lib = _lib.load()
    st = torch.cuda.current_stream().cuda_stream
    for k in range(1, 4):
        g = rng.normal(size=n)
        gd = _t(g)
        _lib.check(lib.cusoap_adam_step(pd.data_ptr(), gd.data_ptr(), md.data_ptr(), vd.data_ptr(), n, 1e-3, 0.9,
                                        0.999, 1e-8, k, st), "adam")
        p, m, v = mo.adam_step(p.astype(np.float32).astype(np.float64), g.astype(np.float32).astype(np.float64),
                               m, v, k, 1e-3)
    assert np.abs(pd.cpu().numpy() - p).max() < 1e-6


def test_refit_fits_a_representable_output_layer():
    """Targets made by the same hidden layers are reproduced by the float64 least-squares refit."""
    H, O, n_pairs, n_samples = HIDDEN, 20, 2, 4096
    rng = np.random.default_rng(3)
    params = np.stack([mo.init_params(H, O, rng) for _ in range(n_pairs)])
    x = rng.uniform(-1, 1, n_samples)
    t = np.empty((n_pairs, n_samples, O))
    dt = np.empty_like(t)
    for q in range(n_pairs):
        tp = params[q].copy()
        W3, b3 = mo.unpack(tp, H, O)[4:]
        W3[:] = rng.normal(size=W3.shape)
        b3[:] = rng.normal(size=b3.shape)
        t[q], dt[q], _ = mo.forward(tp, x, H, O)
    wd = np.full(O, 0.05)
    lib = _lib.load()
    st = torch.cuda.current_stream().cuda_stream
    pd, xd, td, dtd, wdd = _t(params), _t(x), _t(t), _t(dt), _t(wd)
    nbytes = C.c_size_t(0)
    _lib.check(lib.cusoap_mlp_refit_workspace_size(n_pairs, H, O, n_samples, C.byref(nbytes)), "ws")
    ws = torch.empty(int(nbytes.value), dtype=torch.uint8, device=DEV)
    _lib.check(lib.cusoap_mlp_refit_output(pd.data_ptr(), n_pairs, H, O, xd.data_ptr(), td.data_ptr(),
                                           dtd.data_ptr(), n_samples, n_samples * O, wdd.data_ptr(), 0.0,
                                           ws.data_ptr(), ws.numel(), st), "refit")
    fitted = pd.cpu().numpy().astype(np.float64)
    for q in range(n_pairs):
        y, dy, _ = mo.forward(fitted[q], x, H, O)
        say(f"refit pair {q}: value err {np.abs(y - t[q]).max() / np.abs(t[q]).max():.2e}")
        assert np.abs(y - t[q]).max() < 1e-4 * np.abs(t[q]).max()
        assert np.abs(dy - dt[q]).max() < 1e-3 * np.abs(dt[q]).max()


@pytest.fixture(scope="module")
def learned():
    o = so.OracleSOAP(**WB)
    say("training the radial networks (n8 l6, 4 pairs)")
    g = SOAP(device=DEV, dtype="float64", precision="mixed", radial="learned", _betas=o.betas, **WB)
    say("  trained in", round(g.surrogate.train_seconds, 2), "s; errors", g.surrogate.errors)
    return o, g


def test_training_meets_the_gate(learned):
    """Every pair passes the fit stage; the end-to-end stage leaves the calibration errors
    within its bound (0.35e-4) with channels still on the networks."""
    o, g = learned
    e = g.surrogate.errors
    rep = g.surrogate.gate_report()
    say("gate report:", rep)
    assert e["value"].max() <= 2e-5 and e["slope"].max() <= 1e-4
    assert all(u == 1 for u in rep["radial_fit_gate"])
    assert rep["radial_calibration_err"]["desc"] <= 0.35e-4 and rep["radial_calibration_err"]["force"] <= 0.35e-4
    assert g.surrogate.use_pair.any()


@pytest.mark.parametrize("n_max,l_max", [(8, 6), (7, 3), (6, 5), (5, 4), (4, 2), (3, 4)])
def test_exact_table_interpolation_matches_oracle(n_max, l_max):
    """The table path with exact node values isolates the interpolation from the network
    (edge pass: radial vectors in registers at n_max 4..8, in LDS at n_max 3)."""
    s = systems.water_box(5, seed=4)
    p = dict(WB, periodic=True, n_max=n_max, l_max=l_max)
    o = so.OracleSOAP(**p)
    w = np.random.default_rng(5).normal(size=o.get_number_of_features())
    Fr = o.vjp(s, w[None, :].repeat(len(s["positions"]), 0))
    ref = o.create(s)
    # forces: the cubic-Hermite interpolation of the table itself (1024 intervals in
    # float64, 512 in mixed) reaches ~1e-5 of max|F| on the smallest basis (n 4, l 2)
    for prec, tol in (("float64", (1e-6, 2e-5)), ("mixed", (1e-5, 2e-5))):
        g = SOAP(device=DEV, dtype="float64", precision=prec, radial="table", _betas=o.betas, **p)
        d, F = g.descriptor_and_vjp(s, torch.tensor(w))
        ed, ef = colerr(d, ref), np.abs(F.cpu().numpy() - Fr).max() / np.abs(Fr).max()
        say(f"exact table n{n_max} l{l_max} ({prec}): desc err {ed:.2e}, force err {ef:.2e}")
        assert ed < tol[0] and ef < tol[1]


def test_learned_descriptor_and_forces_within_bound(learned):
    """Learned radial + float32 accumulation against the float64 oracle, and against the plain
    float32 method (same kernels, exact radial, float32 Loewdin contraction)."""
    o, g = learned
    s = systems.water_box(5, seed=6)
    p = dict(WB, periodic=True)
    op = so.OracleSOAP(**p)
    gl = SOAP(device=DEV, dtype="float64", precision="mixed", radial="learned", surrogate=g.surrogate,
              _betas=o.betas, **p)
    g32 = SOAP(device=DEV, dtype="float64", precision="float32", _betas=o.betas, **p)
    w = np.random.default_rng(7).normal(size=op.get_number_of_features())
    ref = op.create(s)
    Fr = op.vjp(s, w[None, :].repeat(len(s["positions"]), 0))
    errs = {}
    for name, soap in (("learned", gl), ("float32", g32)):
        d, F = soap.descriptor_and_vjp(s, torch.tensor(w))
        errs[name] = (colerr(d, ref), np.abs(F.cpu().numpy() - Fr).max() / np.abs(Fr).max())
    say("desc/force errors vs float64 oracle:", errs, "pairs learned:", gl._learned_use_pair)
    assert errs["learned"][0] <= 1e-4 and errs["learned"][1] <= 1e-4
    assert errs["learned"][0] < errs["float32"][0]
    assert gl._learned_use_pair.any()


def test_learned_full_box_within_bound(learned):
    """Every atom of a 12,288-atom periodic water box, against the float64 GPU form (itself
    within 1e-7 / 1e-10 of the float64 oracle): descriptors and forces of the learned form
    (default 256-interval table) within 1e-4.  A check on a few hundred atoms can miss the
    tail: with a 1024-interval table the learned forces were within 8e-5 on the first 4,096
    centers of the 98,304-atom box but 1.9e-4 on all atoms (1% of them above 1e-4), from
    float32 node noise that finer tables amplify into slopes."""
    o, g = learned
    s = systems.water_box(16, seed=12)
    p = dict(WB, periodic=True)
    gl = SOAP(device=DEV, dtype="float64", precision="mixed", radial="learned", surrogate=g.surrogate,
              _betas=o.betas, **p)
    g64 = SOAP(device=DEV, dtype="float64", precision="float64", _betas=o.betas, **p)
    w = torch.tensor(np.random.default_rng(13).normal(size=gl.get_number_of_features()))
    d64, F64 = g64.descriptor_and_vjp(s, w)
    d, F = gl.descriptor_and_vjp(s, w)
    ed = ((d - d64).abs().amax(0) / d64.abs().amax(0).clamp_min(1e-300)).max().item()
    ef = ((F - F64).abs().max() / F64.abs().max()).item()
    say(f"learned, all {len(s['positions'])} atoms ({gl._rad_intervals} intervals): desc {ed:.2e}, forces {ef:.2e}")
    assert gl._rad_intervals == 256
    assert ed <= 1e-4 and ef <= 1e-4


def test_checkpoint_roundtrip(learned, tmp_path):
    o, g = learned
    path = str(tmp_path / "radial.npz")
    g.surrogate.save(path)
    g2 = SOAP(device=DEV, dtype="float64", precision="mixed", radial="learned", surrogate=path,
              _betas=o.betas, **WB)
    assert torch.equal(g2._rad_table, g._rad_table)
    s = systems.water_cluster_300(seed=3)
    assert torch.equal(g2.create(s), g.create(s))


HCNO = dict(r_cut=5.0, n_max=8, l_max=6, sigma=0.4, species=[1, 6, 7, 8], periodic=True)


def test_exact_table_multi_species_and_coordinate_centers():
    """Four species (the 4-species accumulation and the edge pass's 5-species groups) and
    centers that mix atom indices with coordinates (the table's coordinate-center slot)."""
    s = systems.random_box(n_atoms=400, density=0.1, seed=3)
    o = so.OracleSOAP(**HCNO)
    w = np.random.default_rng(8).normal(size=o.get_number_of_features())
    Fr = o.vjp(s, w[None, :].repeat(len(s["positions"]), 0))
    ref = o.create(s)
    cen = [0, 17, [1.0, 2.0, 3.0], 399, [10.25, 0.5, 7.75], 5]
    ref_c = o.create(s, centers=cen)
    for prec, tol in (("float64", (1e-6, 2e-5)), ("mixed", (1e-5, 2e-5))):
        g = SOAP(device=DEV, dtype="float64", precision=prec, radial="table", _betas=o.betas, **HCNO)
        d, F = g.descriptor_and_vjp(s, torch.tensor(w))
        ed, ef = colerr(d, ref), np.abs(F.cpu().numpy() - Fr).max() / np.abs(Fr).max()
        ec = colerr(g.create(s, centers=cen), ref_c)
        say(f"exact table H/C/N/O ({prec}): desc {ed:.2e}, forces {ef:.2e}, mixed centers {ec:.2e}")
        assert ed < tol[0] and ef < tol[1] and ec < tol[0]


def test_learned_hcno_sixteen_pairs():
    """The H/C/N/O basis (r_cut 5, sigma 0.4, n 8, l 6): 16 pair networks trained on the
    device, gated, against the float64 oracle on a 400-atom periodic box."""
    o = so.OracleSOAP(**HCNO)
    say("training 16 radial networks (H/C/N/O, n8 l6)")
    g = SOAP(device=DEV, dtype="float64", precision="mixed", radial="learned", _betas=o.betas, **HCNO)
    rep = g.surrogate.gate_report()
    say("  trained in", round(g.surrogate.train_seconds, 1), "s; gate", rep)
    s = systems.random_box(n_atoms=400, density=0.1, seed=9)
    w = np.random.default_rng(10).normal(size=o.get_number_of_features())
    ref = o.create(s)
    Fr = o.vjp(s, w[None, :].repeat(len(s["positions"]), 0))
    d, F = g.descriptor_and_vjp(s, torch.tensor(w))
    ed, ef = colerr(d, ref), np.abs(F.cpu().numpy() - Fr).max() / np.abs(Fr).max()
    g32 = SOAP(device=DEV, dtype="float64", precision="float32", _betas=o.betas, **HCNO)
    ed32 = colerr(g32.create(s), ref)
    say(f"learned H/C/N/O: desc {ed:.2e} (plain float32 {ed32:.2e}), forces {ef:.2e}, "
        f"channels learned {rep['radial_channels_learned']} of {rep['radial_channels_total']}")
    assert ed <= 1e-4 and ef <= 1e-4 and ed < ed32
    assert rep["radial_channels_learned"] > 0
