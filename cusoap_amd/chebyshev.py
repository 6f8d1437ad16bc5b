"""Chebyshev form of the orthonormal GTO radial functions, for the matrix-core kernels.

No reference counterpart: the reference evaluates the radial integral of every edge
(soap_generator.py:1358-1391 for create(), 2389-2405 for the derivative path).  Here
each radial function that multiplies the unit-vector harmonics,

    f_nl(r) = r^l B_nl(r),   B_nl(r) = sum_j beta_l[n][j] pref_lj exp(-gamma_lj r^2),

is expanded once per basis in Chebyshev polynomials on two intervals, [0, r_s] and
[r_s, cutoff] with r_s = 0.38 cutoff, 32 terms each:

    f_nl(r) = sum_{j<32} a_lnpj T_j(t_p(r)),   t_p = 2 (r - lo_p) / (hi_p - lo_p) - 1.

The coefficient kernel then forms, per center, neighbour species and interval, the
moments Q[j][lm] = sum_e T_j(t_e) Y_lm(r_e / |r_e|) as one dense product over the edges
(no l-dependence left in it, so it runs on the matrix cores), and c[n][lm] =
sum_pj a_lnpj Q_pj[lm] afterwards.  The derivative df/dr is the exact derivative of the
truncated series (its own coefficients, chebder), so the adjoint differentiates the
function the forward pass evaluates.

Why two intervals.  A series over all of [0, cutoff] needs ~48 terms, and every edge then
adds to every moment: a radial function that lives below ~2.5 A (the narrow Gaussians of
high l) gets its coefficient from the cancellation of ~200 O(1) moments, and float32
accumulation leaves ~1e-4 of it (measured on the GPU and reproduced in numpy).  With the
split at 0.38 cutoff (3.0 A for the water box) the inner interval holds ~6% of the edges,
the error falls to ~4e-6, each edge touches 32 rows instead of 48, and the fits are
better: 6e-12 (values) and 1.4e-9 (slopes) of each function's maximum for the water box
(r_cut 6, sigma 0.5, n 8, l 6).
"""
from __future__ import annotations

import math
from typing import Optional, Tuple

import numpy as np
import torch

INTERVALS = 2             # [0, r_s] and [r_s, cutoff]
SPLIT = 0.38              # r_s = SPLIT * cutoff
TERMS = 32                # terms per interval (two 16-row tiles of the matrix-core product)
N_PAD = 8                 # n padded to 8 in the coefficient rows
TOL_VALUE = 1e-7          # fit bound on f, relative to each function's maximum on [0, cutoff]
TOL_SLOPE = 1e-5          # fit bound on df/dr, relative to its maximum


def gto_projections(soap, r: np.ndarray) -> Tuple[np.ndarray, np.ndarray]:
    """f = r^l B_nl(r) and df/dr in float64, shape (P, l_max+1, n_max)."""
    r = np.asarray(r, dtype=np.float64)
    L1, n = soap._l_max + 1, soap._n_max
    g, pref, beta = soap._gto_gamma, soap._gto_pref, soap._betas          # (L1,n), (L1,n), (L1,n,n)
    e = np.exp(-g[None] * (r * r)[:, None, None]) * pref[None]               # (P, L1, j)
    B = np.einsum("lnj,plj->pln", beta, e)
    dB = np.einsum("lnj,plj->pln", beta, e * (-2.0 * g[None]) * r[:, None, None])
    lv = np.arange(L1, dtype=np.float64)[None, :, None]
    rl = r[:, None, None] ** lv
    rlm1 = np.where(lv > 0, r[:, None, None] ** np.maximum(lv - 1, 0), 0.0)
    return rl * B, rl * dB + lv * rlm1 * B


def nodes(K: int, cutoff: float) -> Tuple[np.ndarray, np.ndarray]:
    """Chebyshev points of the first kind x_j and their radii r_j = cutoff (x_j + 1) / 2."""
    x = np.cos(np.pi * (np.arange(K) + 0.5) / K)
    return x, 0.5 * cutoff * (x + 1.0)


def interpolate(values: np.ndarray, K: int) -> np.ndarray:
    """Coefficients (..., K) of the interpolant through values (K, ...) at the first-kind nodes."""
    j = np.arange(K) + 0.5
    Tm = np.cos(np.pi * np.outer(np.arange(K), j) / K)                      # (k, j)
    a = (2.0 / K) * np.tensordot(Tm, values, axes=(1, 0))                     # (K, ...)
    a[0] *= 0.5
    return np.moveaxis(a, 0, -1)


def derivative(a: np.ndarray, cutoff: float) -> np.ndarray:
    """Coefficients of d/dr of the series a (..., K) in t = 2 r / cutoff - 1 (same length K)."""
    K = a.shape[-1]
    d = np.zeros_like(a)
    for k in range(K - 1, 0, -1):
        # c'_{k-1} = c'_{k+1} + 2 k c_k (and halved for k - 1 = 0)
        d[..., k - 1] = (d[..., k + 1] if k + 1 < K else 0.0) + 2.0 * k * a[..., k]
    d[..., 0] *= 0.5
    return d * (2.0 / cutoff)


def evaluate(a: np.ndarray, r: np.ndarray, cutoff: float) -> np.ndarray:
    """Series a (..., K) at radii r -> (P, ...)."""
    t = 2.0 * np.asarray(r, dtype=np.float64) / cutoff - 1.0
    K = a.shape[-1]
    T = np.empty((K, t.size))
    T[0] = 1.0
    if K > 1:
        T[1] = t
    for k in range(2, K):
        T[k] = 2.0 * t * T[k - 1] - T[k - 2]
    return np.tensordot(T.T, np.moveaxis(a, -1, 0), axes=(1, 0))


def split_radius(cutoff: float) -> float:
    return SPLIT * cutoff


def interval_edges(cutoff: float) -> np.ndarray:
    return np.array([0.0, split_radius(cutoff), cutoff])


def fit_function(fun, cutoff: float, D: int = TERMS):
    """Per-interval coefficients a, d (..., INTERVALS, D) of fun(r) -> (..,) values and of its
    derivative, by interpolation at each interval's first-kind Chebyshev points."""
    e = interval_edges(cutoff)
    a, d = [], []
    for q in range(INTERVALS):
        lo, hi = e[q], e[q + 1]
        x = np.cos(np.pi * (np.arange(D) + 0.5) / D)
        aq = interpolate(fun(lo + 0.5 * (hi - lo) * (x + 1.0)), D)
        a.append(aq)
        d.append(derivative(aq, hi - lo))
    return np.stack(a, -2), np.stack(d, -2)


def evaluate_pieces(a: np.ndarray, r: np.ndarray, cutoff: float) -> np.ndarray:
    """Piecewise series a (..., INTERVALS, D) at radii r -> (P, ...)."""
    e = interval_edges(cutoff)
    r = np.asarray(r, dtype=np.float64)
    q = (r >= e[1]).astype(int)
    out = np.empty((r.size,) + a.shape[:-2])
    for k in range(INTERVALS):
        m = q == k
        if m.any():
            out[m] = evaluate(a[..., k, :], r[m] - e[k], e[k + 1] - e[k])
    return out


def fit_errors(soap, a, d, grid: int = 4001):
    """Max errors of the series a (values) and d (slopes) against the exact functions, each
    relative to that function's maximum on [0, cutoff]."""
    cutoff = soap._cutoff
    r = np.linspace(0.0, cutoff, grid)
    f_ref, d_ref = gto_projections(soap, r)
    fmax = np.abs(f_ref).max(0)
    dmax = np.abs(d_ref).max(0)
    ef = np.abs(evaluate_pieces(a, r, cutoff) - f_ref).max(0) / np.where(fmax > 0, fmax, 1.0)
    ed = np.abs(evaluate_pieces(d, r, cutoff) - d_ref).max(0) / np.where(dmax > 0, dmax, 1.0)
    return float(ef.max()), float(ed.max())


def fit_exact(soap):
    """Coefficients (L1, n, INTERVALS, TERMS) of f and df/dr of the exact radial functions and
    the measured fit errors (values, slopes)."""
    a, d = fit_function(lambda r: gto_projections(soap, r)[0], soap._cutoff)
    ef, ed = fit_errors(soap, a, d)
    return a, d, ef, ed


def fit_learned(soap, model, channels: np.ndarray):
    """Per-slot coefficients (slots, L1, n, INTERVALS, TERMS) of the gated learned model: the
    network's f = r_fit^l g_net(r) on [0, r_fit] for the channels (pair, l, n) in ``channels``,
    the exact function elsewhere (other channels, radii beyond r_fit, slots of centers that
    are not atoms).  The model jumps by its own error at r_fit; the series spreads that jump
    over the outer interval at a fraction of its size."""
    S, L1, n = soap.n_species, soap._l_max + 1, soap._n_max
    slots = (S + 1) * S
    pairs = channels.shape[0]
    lv = np.arange(L1, dtype=np.float64)

    def fun(r):
        r = np.asarray(r, dtype=np.float64)
        f_exact, _ = gto_projections(soap, r)                                   # (P, L1, n)
        out = np.repeat(f_exact[:, None], slots, axis=1)                        # (P, slots, L1, n)
        g, _ = model.evaluate(r)                                                # (pairs, P, O)
        g = g.cpu().numpy().reshape(pairs, r.size, L1, n)
        f_net = (model.r_fit ** lv)[None, None, :, None] * g
        inside = r <= model.r_fit * (1.0 + 1e-12)
        for pr in range(pairs):
            m = np.asarray(channels[pr], dtype=bool)
            sel = out[inside, pr]
            sel[:, m] = f_net[pr, inside][:, m]
            out[inside, pr] = sel
        return out

    return fit_function(fun, soap._cutoff)


def device_table(soap, a: np.ndarray, d: np.ndarray) -> torch.Tensor:
    """float32 ((S+1) S, 2, L1, INTERVALS, TERMS, 8) device table: the same exact coefficients
    in every slot, or per-slot coefficients when a, d are (slots, L1, n, INTERVALS, TERMS)."""
    S, L1, n = soap.n_species, soap._l_max + 1, soap._n_max
    slots = (S + 1) * S
    if a.ndim == 4:
        a = np.broadcast_to(a, (slots,) + a.shape)
        d = np.broadcast_to(d, (slots,) + d.shape)
    tab = np.zeros((slots, 2, L1, INTERVALS, TERMS, N_PAD), dtype=np.float64)
    tab[:, 0, :, :, :, :n] = np.moveaxis(a, 2, -1)
    tab[:, 1, :, :, :, :n] = np.moveaxis(d, 2, -1)
    return torch.from_numpy(tab.astype(np.float32)).to(soap.device)


def contraction_fragments(soap, a: np.ndarray):
    """fp16 operands of the contraction c[n][lm] = sum_pj a_lnpj Q_p[j][lm] on the matrix cores.

    a: (slots, L1, n, INTERVALS, TERMS) or (L1, n, INTERVALS, TERMS).  Returns a float16 device
    tensor (slots, L1, INTERVALS, 64 lanes, 16) holding, for lane = row + 16 g, the hi halves
    then the lo halves of a_l[row][p][k] * 2^e_lrow at k = 4 g + (j % 4) + 16 (j / 4), j = 0..7
    (the A operand of v_mfma_f32_16x16x32_f16, rows = n, contraction index matching the
    accumulator lanes of Q), and the float32 inverse scales 2^-e_ln (slots, L1, 8).  e_ln puts
    each row's largest coefficient in [2^13, 2^14)."""
    S, L1, n = soap.n_species, soap._l_max + 1, soap._n_max
    slots = (S + 1) * S
    if a.ndim == 4:
        a = np.broadcast_to(a, (slots,) + a.shape)
    amax = np.abs(a).max(axis=(3, 4))                                       # (slots, L1, n)
    e = np.where(amax > 0, 13 - np.floor(np.log2(np.where(amax > 0, amax, 1.0))), 0.0)
    scaled = a * np.exp2(e)[..., None, None]
    lane = np.arange(64)
    row, g = lane % 16, lane // 16
    j = np.arange(8)
    k = 4 * g[:, None] + (j % 4)[None, :] + 16 * (j // 4)[None, :]           # (64, 8)
    frag = np.zeros((slots, L1, INTERVALS, 64, 16), dtype=np.float16)
    valid = row < n
    rr = np.where(valid, row, 0)
    for q in range(INTERVALS):
        v = scaled[:, :, rr[:, None], q, k] * valid[:, None]                 # (slots, L1, 64, 8)
        hi = v.astype(np.float16)
        lo = (v - hi.astype(np.float64)).astype(np.float16)
        frag[:, :, q, :, :8] = hi
        frag[:, :, q, :, 8:] = lo
    inv = np.zeros((slots, L1, N_PAD), dtype=np.float32)
    inv[:, :, :n] = np.exp2(-e)
    dev = soap.device
    return torch.from_numpy(frag).to(dev), torch.from_numpy(inv).to(dev)


def supported(soap) -> Optional[str]:
    """None when the matrix-core Chebyshev kernels cover this basis, else the reason."""
    if soap._rbf != "gto":
        return "the Chebyshev radial form covers the GTO basis"
    if soap._weighting is not None:
        return "the Chebyshev radial form covers unweighted bases"
    if soap._n_max > N_PAD:
        return f"the Chebyshev radial form covers n_max <= {N_PAD}"
    if soap._l_max > 6:
        return "the Chebyshev radial form covers l_max <= 6"
    if soap.n_species > 4:
        return "the Chebyshev radial form covers at most 4 species"
    return None
