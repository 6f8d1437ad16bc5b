"""Fixed-order force sums of the adjoint (every atom its own center; SOAP.ordered_forces).

With the centers being the atoms, cusoap_vjp_ordered stores each edge's contribution at
its own index and sums, for each atom a, the contributions of the edges c -> a (found
through the reverse-edge index of cusoap_edge_reverse) in a fixed order, so the forces
must be the same bits on every call.  Checked per form (float64, plain
float32, mixed, mixed + table, mixed + Chebyshev, polynomial basis):
- three calls give bitwise equal forces;
- they agree with the float64 oracle's adjoint (the tolerance of each form);
- they agree with the atomic sums, reached two ways: the same centers listed in
  reverse order (centers[c] != positions[c], so no edge has a reverse edge and every
  contribution is added atomically inside cusoap_vjp_ordered) and cusoap_vjp;
- with attach=True (each center moves with its atom) against the oracle Jacobian
  contracted with per-center gradients, and bitwise repeatable.
"""

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


def ferr(F, Fr):
    F = F.detach().cpu().numpy() if isinstance(F, torch.Tensor) else np.asarray(F)
    return float(np.abs(F - Fr).max() / np.abs(Fr).max())


BOX = systems.water_box(4, seed=11)           # 192 atoms, periodic
P = dict(r_cut=5.0, n_max=4, l_max=3, sigma=0.5, species=[1, 8], periodic=True)
_ORACLE = {}


def oracle(rbf="gto"):
    if rbf not in _ORACLE:
        o = so.OracleSOAP(rbf=rbf, **P)
        w = np.random.default_rng(3).normal(size=(len(BOX["positions"]), o.get_number_of_features()))
        _ORACLE[rbf] = (o, w, o.vjp(BOX, w))
    return _ORACLE[rbf]


# (dtype, precision, radial, rbf, tolerance against the float64 oracle)
FORMS = [
    ("float64", "float64", "exact", "gto", 1e-10),
    ("float32", "float32", "exact", "gto", 1e-4),
    ("float64", "mixed", "exact", "gto", 2e-5),
    ("float64", "mixed", "table", "gto", 2e-5),
    ("float64", "mixed", "chebyshev", "gto", 2e-5),
    ("float64", "float64", "exact", "polynomial", 1e-10),
]


@pytest.mark.parametrize("dtype,prec,radial,rbf,tol", FORMS, ids=[f"{f[1]}-{f[2]}-{f[3]}" for f in FORMS])
def test_forces_are_bitwise_repeatable(dtype, prec, radial, rbf, tol):
    o, w, Fr = oracle(rbf)
    kw = dict(P, rbf=rbf)
    if rbf == "gto":
        kw["_betas"] = o.betas
    g = SOAP(device=DEV, dtype=dtype, precision=prec, radial=radial, **kw)
    g.ordered_forces = True
    G = torch.tensor(w, dtype=torch.float64, device=DEV)
    runs = [g.vjp(BOX, G) for _ in range(3)]
    same = all(torch.equal(runs[0], r) for r in runs[1:])
    n = len(BOX["positions"])
    rev = list(range(n - 1, -1, -1))
    F_atomic = g.vjp(BOX, G[rev], centers=rev)
    x = g._context(BOX, None)
    F_small = g._vjp_ctx(x, G, None, False, ordered=False)
    e_or, e_at = ferr(runs[0], Fr), ferr(runs[0], F_atomic.cpu().numpy())
    e_sm = ferr(runs[0], F_small.cpu().numpy())
    say(f"{prec}/{radial}/{rbf}: repeatable {same}, oracle {e_or:.1e}, vs reversed centers {e_at:.1e}, "
        f"vs cusoap_vjp {e_sm:.1e}")
    assert same
    assert e_or < tol
    # the same sums in another order: rounding of the edge pass's type at most
    agree = 1e-12 if prec == "float64" else 1e-5
    assert e_at < agree and e_sm < agree


@pytest.mark.parametrize("prec,radial", [("float64", "exact"), ("mixed", "table")])
def test_attached_forces_are_bitwise_repeatable(prec, radial):
    o, _, _ = oracle("gto")
    g = SOAP(device=DEV, dtype="float64", precision=prec, radial=radial, _betas=o.betas, **P)
    g.ordered_forces = True
    n = len(BOX["positions"])
    G = np.random.default_rng(5).normal(size=(n, o.get_number_of_features()))
    J = o.derivatives(BOX, attach=True, return_descriptor=False)          # (C, A, 3, F)
    Fr = np.einsum("caxf,cf->ax", J, G)
    Gt = torch.tensor(G, device=DEV)
    runs = [g.vjp(BOX, Gt, attach=True) for _ in range(3)]
    same = all(torch.equal(runs[0], r) for r in runs[1:])
    e = ferr(runs[0], Fr)
    say(f"attach {prec}/{radial}: repeatable {same}, oracle {e:.1e}")
    assert same and e < (1e-10 if prec == "float64" else 2e-5)


@pytest.mark.parametrize("system", ["box", "small_cell", "cluster"])
def test_edge_reverse_is_the_reverse_edge(system):
    """rev[e] against a host search of the same lists: for e = (c -> a, image s) the position of
    (a -> c, image -s) in a's range, or -1; an involution on the matched edges.  The small
    cell (9.3 A, r_cut 4) holds several images of one atom in a list."""
    if system == "box":
        s, p = BOX, P
    elif system == "small_cell":
        s = systems.water((3, 3, 3), spacing=3.1, seed=12, cell_length=9.3)
        p = dict(r_cut=4.0, n_max=2, l_max=1, sigma=0.5, species=[1, 8], periodic=True)
    else:
        s, p = systems.water_cluster_300(seed=4), dict(r_cut=5.0, n_max=2, l_max=1, sigma=0.5, species=[1, 8])
    g = SOAP(device=DEV, dtype="float64", **p)
    g.ordered_forces = True
    x = g._context(s, None)
    nl = g._neighbors(x, drop_self=False)
    rev, ev = g._edge_reverse(x, nl)
    torch.cuda.current_stream().wait_event(ev)
    r = rev.cpu().numpy()
    off = nl.offsets.cpu().numpy()
    nbr = nl.nbr[:nl.n_edges].cpu().numpy().astype(np.int64)
    n = x.n_atoms
    sh = x.shifts.cpu().numpy()
    neg = np.array([int(np.nonzero((sh == -sh[i]).all(1))[0][0]) for i in range(len(sh))])
    img, a = nbr // n, nbr % n
    c = np.repeat(np.arange(n), np.diff(off))
    pos_of = {}
    for e in range(len(nbr)):
        pos_of[(int(c[e]), int(nbr[e]))] = e
    want = np.array([pos_of.get((int(a[e]), int(neg[img[e]] * n + c[e])), -1) for e in range(len(nbr))])
    matched = r >= 0
    say(f"edge_reverse {system}: {len(nbr)} edges, {len(sh)} images, {int((~matched).sum())} without reverse")
    assert np.array_equal(r, want)
    assert np.array_equal(r[r[matched]], np.nonzero(matched)[0])


@pytest.mark.parametrize("prec,radial,tol", [("float64", "exact", 1e-12), ("mixed", "table", 1e-6)])
def test_forces_do_not_depend_on_atom_order(prec, radial, tol):
    """The edge pass groups centers along a Morton curve of their coordinates: the same box
    with its atoms shuffled gives the same forces (permuted back) to rounding of the sums."""
    s = systems.water_box(6, seed=9)                     # 648 atoms
    p = dict(P, r_cut=5.0)
    g = SOAP(device=DEV, dtype="float64", precision=prec, radial=radial, **p)
    n = len(s["positions"])
    w = torch.tensor(np.random.default_rng(2).normal(size=g.get_number_of_features()), device=DEV)
    F = g.vjp(s, w).cpu().numpy()
    perm = np.random.default_rng(8).permutation(n)
    sp = dict(s, positions=np.asarray(s["positions"])[perm], atomic_numbers=np.asarray(s["atomic_numbers"])[perm])
    Fp = g.vjp(sp, w).cpu().numpy()
    inv = np.empty(n, dtype=np.int64)
    inv[perm] = np.arange(n)
    e = ferr(Fp[inv], F)
    say(f"shuffled atoms {prec}/{radial}: {e:.1e}")
    assert e < tol
