"""Host side of the Chebyshev radial form (cusoap_amd/chebyshev.py), on the CPU.

The projections f_nl = r^l B_nl and their slopes are pinned against the oracle's radial
functions (B and Bhat = (1/r) dB/dr, soap_generator.py:2389-2405); the interpolation and
the derivative recurrence against numpy's Chebyshev routines; the fits of the bases the
benchmarks use against the stated bounds (1e-7 values, 1e-5 slopes).
"""

import numpy as np
import pytest

from cusoap_amd import SOAP
from cusoap_amd import chebyshev as ch
from oracle import soap_oracle as so

BASES = [
    dict(r_cut=6.0, n_max=8, l_max=6, sigma=0.5, species=[1, 8]),          # water box
    dict(r_cut=5.0, n_max=8, l_max=6, sigma=0.4, species=[1, 6, 7, 8]),    # H/C/N/O box
    dict(r_cut=6.0, n_max=7, l_max=3, sigma=0.5, species=[1, 8]),          # 1M cluster
]


def _soap(p):
    o = so.OracleSOAP(**p)
    return o, SOAP(device="cpu", dtype="float64", _betas=o.betas, **p)


@pytest.mark.parametrize("p", BASES)
def test_projections_match_oracle_radial_functions(p):
    o, g = _soap(p)
    r = np.linspace(1e-3, g._cutoff, 777)
    f, df = ch.gto_projections(g, r)                      # (P, L1, n)
    B, Bh = o.radial_B_Bhat(r)                            # (P, n, L1)
    lv = np.arange(p["l_max"] + 1)[None, :, None]
    rl = r[:, None, None] ** lv
    f_ref = rl * np.swapaxes(B, 1, 2)
    df_ref = rl * r[:, None, None] * np.swapaxes(Bh, 1, 2) + lv * r[:, None, None] ** np.maximum(lv - 1, 0) * \
        np.swapaxes(B, 1, 2)
    assert np.abs(f - f_ref).max() <= 1e-12 * np.abs(f_ref).max()
    assert np.abs(df - df_ref).max() <= 1e-11 * np.abs(df_ref).max()


def test_interpolation_and_derivative_match_numpy():
    rng = np.random.default_rng(0)
    a = rng.normal(size=(3, 48))
    d = ch.derivative(a, 7.0)
    ref = np.polynomial.chebyshev.chebder(a, axis=-1) * (2.0 / 7.0)
    assert np.abs(d[..., :47] - ref).max() < 1e-11 and not d[..., 47].any()
    x, r = ch.nodes(24, 5.0)
    v = np.stack([np.exp(x), np.cos(3 * x)], axis=1)
    c = ch.interpolate(v, 24)
    for j in range(2):
        assert np.abs(c[j] - np.polynomial.chebyshev.chebfit(x, v[:, j], 23)).max() < 1e-12
    # evaluate() is the series in t = 2 r / cutoff - 1
    rr = np.linspace(0, 5.0, 11)
    ev = ch.evaluate(c, rr, 5.0)
    assert np.abs(ev - np.stack([np.polynomial.chebyshev.chebval(2 * rr / 5 - 1, c[j]) for j in range(2)], 1)).max() \
        < 1e-12


@pytest.mark.parametrize("p", BASES)
def test_fit_meets_bounds(p):
    _, g = _soap(p)
    a, d, ef, ed = ch.fit_exact(g)
    assert ef <= ch.TOL_VALUE and ed <= ch.TOL_SLOPE
    assert a.shape == (p["l_max"] + 1, p["n_max"], ch.INTERVALS, ch.TERMS)
    # the series is continuous across the split to the fit accuracy
    rs = ch.split_radius(g._cutoff)
    lo = ch.evaluate_pieces(a, np.array([rs * (1 - 1e-12)]), g._cutoff)
    hi = ch.evaluate_pieces(a, np.array([rs]), g._cutoff)
    fmax = np.abs(ch.gto_projections(g, np.linspace(0, g._cutoff, 2001))[0]).max(0)
    assert (np.abs(lo - hi)[0] / fmax).max() < 1e-8
    # a planted error in one coefficient is seen by the bound check
    a2 = a.copy()
    a2[3 % (p["l_max"] + 1), 0, 1, 20] += 1e-5 * np.abs(a[3 % (p["l_max"] + 1), 0]).max()
    assert ch.fit_errors(g, a2, d)[0] > ch.TOL_VALUE


def test_supported_reasons():
    _, g = _soap(dict(r_cut=5.0, n_max=9, l_max=3, sigma=0.5, species=[1]))
    assert "n_max" in ch.supported(g)
    _, g = _soap(dict(r_cut=5.0, n_max=4, l_max=7, sigma=0.5, species=[1]))
    assert "l_max" in ch.supported(g)
    _, g = _soap(dict(r_cut=5.0, n_max=4, l_max=3, sigma=0.5, species=[1, 8]))
    assert ch.supported(g) is None


def test_device_layouts_decode_to_the_coefficients():
    """device_table and contraction_fragments hold the fitted coefficients in the layouts the
    kernels read: the table as (slot, value/slope, l, interval, term, n padded to 8), the
    fragments as hi + lo halves of a_l[n][p][k] 2^e_ln at lane n + 16 g, element j with
    k = 4 g + (j % 4) + 16 (j / 4), and rowinv = 2^-e_ln putting each row maximum in
    [2^13, 2^14)."""
    p = BASES[0]
    _, g = _soap(p)
    a, d, _, _ = ch.fit_exact(g)
    S, L1, n = g.n_species, p["l_max"] + 1, p["n_max"]
    slots = (S + 1) * S
    tab = ch.device_table(g, a, d).numpy()
    assert tab.shape == (slots, 2, L1, ch.INTERVALS, ch.TERMS, ch.N_PAD)
    for sl in (0, slots - 1):
        assert np.abs(tab[sl, 0, :, :, :, :n] - np.moveaxis(a, 1, -1)).max() <= 1e-7 * np.abs(a).max()
        assert np.abs(tab[sl, 1, :, :, :, :n] - np.moveaxis(d, 1, -1)).max() <= 1e-7 * np.abs(d).max()
    assert not tab[..., n:].any()
    frag, rowinv = ch.contraction_fragments(g, a)
    frag, rowinv = frag.numpy().astype(np.float64), rowinv.numpy().astype(np.float64)
    assert frag.shape == (slots, L1, ch.INTERVALS, 64, 16) and rowinv.shape == (slots, L1, ch.N_PAD)
    lane = np.arange(64)
    row, grp = lane % 16, lane // 16
    j = np.arange(8)
    k = 4 * grp[:, None] + (j % 4)[None, :] + 16 * (j // 4)[None, :]           # (64, 8)
    for sl in (0, slots - 1):
        for l in range(L1):
            for q in range(ch.INTERVALS):
                val = (frag[sl, l, q, :, :8] + frag[sl, l, q, :, 8:]) * 1.0     # hi + lo, scaled
                for ln in range(64):
                    if row[ln] >= n:
                        assert not frag[sl, l, q, ln].any()
                        continue
                    ref = a[l, row[ln], q, k[ln]]
                    got = val[ln] * rowinv[sl, l, row[ln]]
                    assert np.abs(got - ref).max() <= 2.0 ** -21 * np.abs(a[l, row[ln]]).max() + 1e-30
            rmax = np.abs(a[l, :, :, :]).max(axis=(1, 2)) / rowinv[sl, l, :n]    # scaled row maxima
            assert np.all((rmax >= 2.0 ** 13) & (rmax < 2.0 ** 14))
