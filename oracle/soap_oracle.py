"""FP64 CPU oracle for the SOAP hot path -- TEST INFRASTRUCTURE ONLY.

This module restates, in plain numpy float64, the algorithm of the reference
``SOAP`` class (``src/cusoap/soap_generator.py``, abbreviated SG below).  It is
the checker the HIP path is compared against.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it; the product package (``cusoap_amd``) never does, and has no CPU
fallback through it.

What each function follows (file:line in the reference) and what pins it is
stated per function.  The one deliberate substitution: the reference takes its
real spherical / solid harmonics from the external ``sphericart`` library
(SG:675-686); here they come from our own Cartesian recurrence
(:func:`solid_harmonics`), pinned by the closed forms of the real-harmonic
table for l<=3, by scipy's complex harmonics converted to the real basis
without the Condon-Shortley phase (the convention sphericart documents) for
every degree up to 9 (sign- and m-order-sensitive), by orthonormality by
quadrature, the addition theorem and finite differences.  sphericart itself is
not installed, so the identity with its implementation rests on that
documented convention.

Neighbour-list distances use the reference's formula
d^2 = (|c|^2 + |n|^2) - 2 c.n (SG:1091-1113), with the sums written out in a
fixed order ((x*x + y*y) + z*z) and no fused multiply-add, which is the
definition of "bit-exact" the HIP kernels follow.
"""

from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Any, List, Optional, Sequence, Tuple

import numpy as np

PI = math.pi
Y00 = 1.0 / math.sqrt(4.0 * math.pi)


# ---------------------------------------------------------------------------
# Basis construction
# ---------------------------------------------------------------------------

def lowdin_invsqrt(S: np.ndarray, eps: float = 1e-12) -> np.ndarray:
    """S^{-1/2} via symmetric eigendecomposition, eigenvalues clamped at eps.

    Follows SG:339-343 (``_lowdin_invsqrt``).  Pinned in tests by
    beta S beta = I and beta = beta^T > 0.
    """
    w, V = np.linalg.eigh(S)
    w = np.maximum(w, eps)
    return (V * (1.0 / np.sqrt(w))[None, :]) @ V.T


def torch_linspace(start: float, end: float, n: int) -> np.ndarray:
    """torch.linspace(start, end, n) in float64, as used at SG:694.

    step = (end-start)/(n-1); entries below n//2 are fma(step, i, start), the rest
    fma(-step, n-1-i, end), each rounded once (computed exactly with fractions).
    Pinned against torch.linspace itself in the tests.
    """
    from fractions import Fraction
    if n == 1:
        return np.array([float(start)])
    step = (end - start) / (n - 1)
    return np.array([float(Fraction(start) + Fraction(step) * i) if i < n // 2
                     else float(Fraction(end) - Fraction(step) * (n - 1 - i)) for i in range(n)])


def gto_basis(r_cut: float, n_max: int, l_max: int) -> Tuple[np.ndarray, np.ndarray]:
    """GTO exponents alpha (l_max+1, n_max) and Loewdin betas (l_max+1, n, n).

    Follows SG:690-708: a = torch.linspace(1, r_cut, n_max) (restated by
    :func:`torch_linspace`),
    alpha_l = -ln(1e-3 / a^l) / a^2, S_l,ij = Gamma(l+3/2)/2 (a_i+a_j)^-(l+3/2).
    """
    a = torch_linspace(1.0, r_cut, n_max)
    alphas, betas = [], []
    for l in range(l_max + 1):
        alpha_l = -np.log(1e-3 / (a ** l)) / (a ** 2)
        m = alpha_l[:, None] + alpha_l[None, :]
        gamma_val = math.exp(math.lgamma(l + 1.5))
        S = 0.5 * gamma_val * (m ** (-(l + 1.5)))
        alphas.append(alpha_l)
        betas.append(lowdin_invsqrt(S))
    return np.stack(alphas), np.stack(betas)


def gto_overlap(alpha_l: np.ndarray, l: int) -> np.ndarray:
    """Overlap matrix of the un-normalised GTO primitives r^l exp(-alpha r^2)."""
    m = alpha_l[:, None] + alpha_l[None, :]
    return 0.5 * math.exp(math.lgamma(l + 1.5)) * (m ** (-(l + 1.5)))


def gauss_legendre(n: int) -> Tuple[np.ndarray, np.ndarray]:
    """Gauss-Legendre nodes/weights on [-1, 1] by Golub-Welsch (SG:350-362)."""
    k = np.arange(1, n, dtype=np.float64)
    beta = k / np.sqrt(4 * k * k - 1)
    J = np.diag(beta, 1) + np.diag(beta, -1)
    x, V = np.linalg.eigh(J)
    w = 2 * (V[0, :] ** 2)
    return x, w


def poly_basis(r_cut: float, n_max: int, quad_n: int = 100):
    """Polynomial basis on the Gauss-Legendre grid (SG:710-761).

    Returns (rx, wr, gss, betas) with gss = betas @ fs, fs[k,q] = (rc-rx_q)^(k+3).
    betas = scipy.linalg.sqrtm(inv(S)) exactly as SG:735-736.
    """
    from scipy.linalg import sqrtm

    n = n_max
    rc = r_cut
    S = np.zeros((n, n))
    for i in range(1, n + 1):
        for j in range(1, n + 1):
            S[i - 1, j - 1] = (2 * rc ** (7 + i + j)) / ((5 + i + j) * (6 + i + j) * (7 + i + j))
    betas = sqrtm(np.linalg.inv(S))
    if np.iscomplexobj(betas):
        raise ValueError("Could not calculate normalization factors for the radial basis "
                         "in the domain of real numbers.")
    betas = np.ascontiguousarray(betas)
    x, w = gauss_legendre(quad_n)
    rx = rc * 0.5 * (x + 1.0)
    wr = rc * 0.5 * w
    rclip = np.clip(rx, 0.0, rc)
    fs = np.zeros((n, quad_n))
    for k in range(1, n + 1):
        fs[k - 1] = (rc - rclip) ** (k + 2)
    gss = betas @ fs
    return rx, wr, gss, betas


# ---------------------------------------------------------------------------
# Weighting (SG:403-470)
# ---------------------------------------------------------------------------

@dataclass
class Weighting:
    function: Optional[str] = None
    r0: float = 1.0
    c: float = 1.0
    d: float = 1.0
    m: float = 1.0
    w0: float = 1.0

    def infer_rcut(self, threshold: float = 1e-2) -> float:
        """SG:423-440."""
        if self.function is None:
            raise ValueError("Cannot infer r_cut when weighting.function is None; provide r_cut explicitly.")
        fn = self.function
        if fn == "poly":
            return float(self.r0)
        if fn in ("pow", "exp"):
            val = self.c / threshold - self.d
            if val <= 0:
                raise ValueError(f"Invalid weighting: c/threshold - d must be >0 for {fn}.")
            if fn == "pow":
                return float(self.r0 * (val ** (1.0 / self.m)))
            return float(-self.r0 * math.log(val))
        raise ValueError(f"Unknown weighting function: {fn}")

    def __call__(self, r: np.ndarray) -> np.ndarray:
        """w(r) with w0 at r < 1e-12 (SG:442-470)."""
        r = np.asarray(r, dtype=np.float64)
        if self.function is None:
            w = np.ones_like(r)
        elif self.function == "poly":
            rr0 = r / self.r0
            base = 1.0 + 2.0 * (rr0 * rr0 * rr0) - 3.0 * (rr0 * rr0)
            w = np.where(r <= self.r0, self.c * np.maximum(base, 0.0) ** self.m, 0.0)
        elif self.function == "pow":
            w = self.c / (self.d + (r / self.r0) ** self.m)
        elif self.function == "exp":
            w = self.c / (self.d + np.exp(-r / self.r0))
        else:
            raise ValueError(f"Unknown weighting function: {self.function}")
        return np.where(r < 1e-12, self.w0, w)

    def deriv(self, r: np.ndarray) -> np.ndarray:
        """dw/dr (0 where w is the constant w0, i.e. r < 1e-12).

        The reference obtains this by autograd through ``__call__``; this is the
        hand derivative of the same expressions (pinned by FD in tests).
        """
        r = np.asarray(r, dtype=np.float64)
        if self.function is None:
            dw = np.zeros_like(r)
        elif self.function == "poly":
            x = r / self.r0
            base = 1.0 + 2.0 * x ** 3 - 3.0 * x ** 2
            dbase = (6.0 * x ** 2 - 6.0 * x) / self.r0
            with np.errstate(divide="ignore", invalid="ignore"):
                dw = np.where((r <= self.r0) & (base > 0.0),
                              self.c * self.m * np.maximum(base, 0.0) ** (self.m - 1.0) * dbase, 0.0)
        elif self.function == "pow":
            x = r / self.r0
            with np.errstate(divide="ignore", invalid="ignore"):
                dx = np.where(r > 0, self.m * x ** (self.m - 1.0) / self.r0, 0.0)
            dw = -self.c * dx / (self.d + x ** self.m) ** 2
        elif self.function == "exp":
            e = np.exp(-r / self.r0)
            dw = self.c * e / self.r0 / (self.d + e) ** 2
        else:
            raise ValueError(f"Unknown weighting function: {self.function}")
        return np.where(r < 1e-12, 0.0, dw)


# ---------------------------------------------------------------------------
# Real solid / spherical harmonics (replaces sphericart, SG:675-686)
# ---------------------------------------------------------------------------

def _klm(l: int, m: int) -> float:
    return math.sqrt((2 * l + 1) / (4 * math.pi) * math.factorial(l - m) / math.factorial(l + m))


def solid_harmonics(l_max: int, xyz: np.ndarray, grad: bool = False):
    """Real solid harmonics r^l Y_lm(r^) (E, (l_max+1)^2), optional gradients (E,3,L).

    Ordering l-major, m=-l..l (index l*l + l + m).  Convention: positive real
    harmonics without Condon-Shortley phase (Y_1,-1..1 proportional to y,z,x),
    the convention sphericart documents.  Built from the associated-Legendre
    recurrence in (z, r^2) and (x+iy)^m; gradients by forward differentiation
    of the same recurrence.
    """
    xyz = np.asarray(xyz, dtype=np.float64)
    x, y, z = xyz[:, 0], xyz[:, 1], xyz[:, 2]
    E = xyz.shape[0]
    r2 = x * x + y * y + z * z
    L = (l_max + 1) ** 2
    Y = np.zeros((E, L))
    G = np.zeros((E, 3, L)) if grad else None
    # c_m + i s_m = (x + i y)^m
    c = [np.ones(E)]
    s = [np.zeros(E)]
    for m in range(1, l_max + 1):
        c.append(x * c[m - 1] - y * s[m - 1])
        s.append(x * s[m - 1] + y * c[m - 1])
    for m in range(0, l_max + 1):
        # Q_l^m(z, r^2) with derivatives dQ/dz (at fixed r2) and dQ/dr2
        dfact = 1.0
        for k in range(1, 2 * m, 2):
            dfact *= k
        Q = {m: np.full(E, dfact)}
        Qz = {m: np.zeros(E)}
        Qr = {m: np.zeros(E)}
        if m + 1 <= l_max:
            Q[m + 1] = (2 * m + 1) * z * Q[m]
            Qz[m + 1] = (2 * m + 1) * Q[m]
            Qr[m + 1] = np.zeros(E)
        for l in range(m + 2, l_max + 1):
            a = 2 * l - 1
            b = l + m - 1
            Q[l] = (a * z * Q[l - 1] - b * r2 * Q[l - 2]) / (l - m)
            Qz[l] = (a * (Q[l - 1] + z * Qz[l - 1]) - b * r2 * Qz[l - 2]) / (l - m)
            Qr[l] = (a * z * Qr[l - 1] - b * (Q[l - 2] + r2 * Qr[l - 2])) / (l - m)
        for l in range(m, l_max + 1):
            if m == 0:
                K = _klm(l, 0)
                Y[:, l * l + l] = K * Q[l]
                if grad:
                    G[:, 0, l * l + l] = K * 2 * x * Qr[l]
                    G[:, 1, l * l + l] = K * 2 * y * Qr[l]
                    G[:, 2, l * l + l] = K * (Qz[l] + 2 * z * Qr[l])
            else:
                K = math.sqrt(2.0) * _klm(l, m)
                ip = l * l + l + m
                im = l * l + l - m
                Y[:, ip] = K * Q[l] * c[m]
                Y[:, im] = K * Q[l] * s[m]
                if grad:
                    dQx = 2 * x * Qr[l]
                    dQy = 2 * y * Qr[l]
                    dQz = Qz[l] + 2 * z * Qr[l]
                    G[:, 0, ip] = K * (dQx * c[m] + Q[l] * m * c[m - 1])
                    G[:, 1, ip] = K * (dQy * c[m] - Q[l] * m * s[m - 1])
                    G[:, 2, ip] = K * dQz * c[m]
                    G[:, 0, im] = K * (dQx * s[m] + Q[l] * m * s[m - 1])
                    G[:, 1, im] = K * (dQy * s[m] + Q[l] * m * c[m - 1])
                    G[:, 2, im] = K * dQz * s[m]
    return (Y, G) if grad else Y


def spherical_harmonics(l_max: int, unit: np.ndarray) -> np.ndarray:
    """Real spherical harmonics Y_lm(u) for unit vectors u (= solid harmonics at |u|=1)."""
    return solid_harmonics(l_max, unit)


# ---------------------------------------------------------------------------
# Periodic images and neighbour search
# ---------------------------------------------------------------------------

def perpendicular_lengths(cell: np.ndarray) -> np.ndarray:
    """Cell heights |V| / |b x c| ... (SG:369-379)."""
    a, b, c = cell[0], cell[1], cell[2]
    v = float(np.dot(a, np.cross(b, c)))
    eps = 1e-12
    return np.array([abs(v) / (np.linalg.norm(np.cross(b, c)) + eps),
                     abs(v) / (np.linalg.norm(np.cross(c, a)) + eps),
                     abs(v) / (np.linalg.norm(np.cross(a, b)) + eps)])


def pbc_offsets(cell: np.ndarray, pbc: np.ndarray, cutoff: float) -> np.ndarray:
    """Integer image offsets, meshgrid 'ij' order over -n..n per periodic axis (SG:382-396)."""
    perps = perpendicular_lengths(cell)
    n_rep = [int(math.ceil(cutoff / (perps[i] + 1e-12))) if bool(pbc[i]) else 0 for i in range(3)]
    ranges = [np.arange(-n, n + 1) for n in n_rep]
    ox, oy, oz = np.meshgrid(*ranges, indexing="ij")
    return np.stack([ox.reshape(-1), oy.reshape(-1), oz.reshape(-1)], axis=-1)


def image_shifts(cell: np.ndarray, offsets: np.ndarray) -> np.ndarray:
    """shift = offsets @ cell, summed in the fixed order (o0 c0 + o1 c1) + o2 c2 (SG:1017)."""
    o = offsets.astype(np.float64)
    return (o[:, 0:1] * cell[0][None, :] + o[:, 1:2] * cell[1][None, :]) + o[:, 2:3] * cell[2][None, :]


def extend_periodic(positions, Z, cell, pbc, cutoff):
    """All atoms replicated over the image offsets: index img*N + i (SG:1015-1021)."""
    offsets = pbc_offsets(cell, pbc, cutoff)
    shift = image_shifts(cell, offsets)
    pos_ext = (positions[None, :, :] + shift[:, None, :]).reshape(-1, 3)
    Z_ext = np.tile(Z, offsets.shape[0])
    return pos_ext, Z_ext, shift


def sqnorm3(p: np.ndarray) -> np.ndarray:
    return (p[:, 0] * p[:, 0] + p[:, 1] * p[:, 1]) + p[:, 2] * p[:, 2]


def radius_edges(centers: np.ndarray, neigh_pos: np.ndarray, cutoff: float,
                 chunk: int = 1024) -> Tuple[np.ndarray, np.ndarray]:
    """Edges (center, neighbour) with d^2 <= cutoff^2 in (center, neighbour) order.

    Follows the GEMM branch of SG:1077-1125: d^2 = ((-2 c.n) + |c|^2) + |n|^2
    in float64, ``nonzero`` order.  The dot product is summed as
    (cx nx + cy ny) + cz nz with separate roundings.
    """
    C, M = centers.shape[0], neigh_pos.shape[0]
    if C == 0 or M == 0:
        return np.zeros(0, np.int64), np.zeros(0, np.int64)
    cutoff2 = float(cutoff) * float(cutoff)
    # Candidates outside the centers' bounding box grown by cutoff + 1e-6 cannot pass
    # the test: their d^2 exceeds cutoff^2 by >= 2 cutoff 1e-6, far above the ~1e-11
    # rounding of the formula at these coordinates.  Dropping them first leaves the
    # same edges in the same order (np.nonzero keeps ascending candidate indices).
    lo = centers.min(0) - (cutoff + 1e-6)
    hi = centers.max(0) + (cutoff + 1e-6)
    keep = np.nonzero(np.all((neigh_pos >= lo) & (neigh_pos <= hi), axis=1))[0]
    if keep.size < M:
        ci, ni = radius_edges(centers, neigh_pos[keep], cutoff, chunk)
        return ci, keep[ni]
    nn = sqnorm3(neigh_pos)
    ci, ni = [], []
    for s in range(0, C, chunk):
        cp = centers[s:s + chunk]
        cn = sqnorm3(cp)
        dot = (cp[:, 0:1] * neigh_pos[None, :, 0] + cp[:, 1:2] * neigh_pos[None, :, 1]) \
            + cp[:, 2:3] * neigh_pos[None, :, 2]
        d2 = ((dot * -2.0) + cn[:, None]) + nn[None, :]
        a, b = np.nonzero(d2 <= cutoff2)
        ci.append(a + s)
        ni.append(b)
    return np.concatenate(ci).astype(np.int64), np.concatenate(ni).astype(np.int64)


def cutoff_band_pairs(centers, neigh_pos, cutoff, rel=1e-9):
    """Number of pairs whose d^2 lies within a relative band of cutoff^2 (ordering-sensitive zone)."""
    cutoff2 = cutoff * cutoff
    nn = sqnorm3(neigh_pos)
    cnt = 0
    for s in range(0, centers.shape[0], 1024):
        cp = centers[s:s + 1024]
        d2 = ((cp[:, None, :] - neigh_pos[None, :, :]) ** 2).sum(-1)
        cnt += int(np.count_nonzero(np.abs(d2 - cutoff2) <= rel * cutoff2))
    return cnt


# ---------------------------------------------------------------------------
# The SOAP oracle
# ---------------------------------------------------------------------------

class OracleSOAP:
    """FP64 restatement of the reference ``SOAP`` (SG:515-2699), CPU, numpy.

    Parameters mirror SG:524-541 (device/dtype/sparse are irrelevant here:
    every output is float64).
    """

    def __init__(self, r_cut=None, n_max=None, l_max=None, sigma=1.0, rbf="gto",
                 weighting=None, crossover=True, average="off", species=None,
                 periodic=False, quad_n=100, betas=None):
        if n_max is None or l_max is None:
            raise ValueError("n_max and l_max must be provided.")
        if species is None or len(species) == 0:
            raise ValueError("species must be provided (list of atomic numbers or symbols).")
        self.species_Z = sorted(set(int(s) for s in species))
        self.n_species = len(self.species_Z)
        self.Z_to_sp = {z: i for i, z in enumerate(self.species_Z)}
        self.n_max, self.l_max = int(n_max), int(l_max)
        self.sigma = float(sigma)
        self.eta = 1.0 / (2.0 * self.sigma * self.sigma)
        self.rbf = str(rbf).lower()
        self.crossover = bool(crossover)
        self.average = str(average)
        self.periodic = bool(periodic)
        self.quad_n = int(quad_n)
        if self.average not in ("off", "inner", "outer", "cc"):
            raise ValueError("average must be one of: 'off', 'inner', 'outer', 'cc'.")
        thr = 1e-2
        if weighting is not None:
            wd = dict(weighting)
            thr = float(wd.pop("threshold", 1e-2))
            self.weighting = Weighting(**wd)
        else:
            self.weighting = None
        if r_cut is None:
            if self.weighting is None:
                raise ValueError("Either r_cut or weighting must be provided.")
            self.r_cut = float(self.weighting.infer_rcut(threshold=thr))
        else:
            self.r_cut = float(r_cut)
        self.cutoff = self.r_cut + float(self.sigma * math.sqrt(-2.0 * math.log(1e-3)))
        if self.rbf == "gto":
            self.alphas, self.betas = gto_basis(self.r_cut, self.n_max, self.l_max)
            if betas is not None:
                self.betas = np.asarray(betas, dtype=np.float64)
        elif self.rbf == "polynomial":
            self.rx, self.wr, self.gss, self.poly_betas = poly_basis(self.r_cut, self.n_max, self.quad_n)
        else:
            raise ValueError("rbf must be 'gto' or 'polynomial'.")
        self._init_self()
        self.feat_slices = self._feature_slices()

    # ---- constants -------------------------------------------------------
    @property
    def w0(self) -> float:
        return float(self.weighting.w0) if self.weighting is not None else 1.0

    def _init_self(self):
        """SG:765-819."""
        eta = self.eta
        if self.rbf == "gto":
            p = self.alphas[0] + eta
            prim = self.w0 * (PI ** 1.5) * p ** -1.5
            self.self_l0 = (self.betas[0] @ prim) * Y00
            if self.l_max >= 1:
                p1 = self.alphas[1] + eta
                self.self_l1 = self.w0 * (self.betas[1] @ ((PI ** 1.5) * (eta / p1) * p1 ** -1.5))
        else:
            common0 = (4.0 * PI) * (self.rx ** 2) * np.exp(-eta * self.rx ** 2) * self.wr
            self.self_l0 = self.w0 * (common0 @ self.gss.T) * Y00
            if self.l_max >= 1:
                common1 = common0 * (2.0 * eta * self.rx / 3.0)
                self.self_l1 = self.w0 * (common1 @ self.gss.T)

    def gto_radial_tables(self):
        """gamma_kl, pref_kl (l+1, n): prim_kl(r) = pref_kl r^l exp(-gamma_kl r^2) (SG:1358-1368)."""
        p = self.alphas + self.eta
        gamma = self.alphas * self.eta / p
        lv = np.arange(self.l_max + 1, dtype=np.float64)[:, None]
        pref = (PI ** 1.5) * (self.eta / p) ** lv * p ** -1.5
        return gamma, pref

    # ---- features ----------------------------------------------------------
    def n_power_features(self) -> int:
        """SG:913-919."""
        if self.crossover:
            ch = self.n_species * self.n_max
            return (ch * (ch + 1) // 2) * (self.l_max + 1)
        return self.n_species * (self.l_max + 1) * (self.n_max * (self.n_max + 1) // 2)

    def get_number_of_features(self) -> int:
        """SG:908-911."""
        if self.average == "cc":
            return self.n_species * self.n_max * (self.l_max + 1) ** 2
        return self.n_power_features()

    def _feature_slices(self):
        """(j, jd, l, is_diag, start, end) in DScribe order (SG:1215-1245)."""
        out, off = [], 0
        n = self.n_max
        for j in range(self.n_species):
            jd_lim = self.n_species if self.crossover else j + 1
            for jd in range(j, jd_lim):
                for l in range(self.l_max + 1):
                    size = n * (n + 1) // 2 if j == jd else n * n
                    out.append((j, jd, l, j == jd, off, off + size))
                    off += size
        assert off == self.n_power_features()
        return out

    def feature_table(self) -> np.ndarray:
        """Per feature (j, jd, n, nd, l) in the order of SG:1247-1291."""
        rows = []
        for (j, jd, l, diag, s0, s1) in self.feat_slices:
            pairs = [(a, b) for a in range(self.n_max) for b in range(a if diag else 0, self.n_max)]
            rows.extend((j, jd, a, b, l) for a, b in pairs)
        return np.array(rows, dtype=np.int64).reshape(-1, 5)

    def get_location(self, pair):
        """SG:921-962 (atomic numbers only)."""
        if self.average == "cc":
            raise ValueError("get_location is not defined for average='cc' (projection coefficients).")
        Z1, Z2 = int(pair[0]), int(pair[1])
        i, j = self.Z_to_sp[Z1], self.Z_to_sp[Z2]
        if i > j:
            i, j = j, i
        symm = (self.n_max * (self.n_max + 1) // 2) * (self.l_max + 1)
        unsymm = self.n_max * self.n_max * (self.l_max + 1)
        if self.crossover:
            m_symm = i + int(j > i)
            m_unsymm = j + i * self.n_species - i * (i + 1) / 2 - m_symm
            start = int(m_symm * symm + m_unsymm * unsymm)
            return slice(start, start + (symm if i == j else unsymm))
        if i != j:
            raise ValueError("crossover=False: no cross-species blocks exist.")
        return slice(i * symm, i * symm + symm)

    # ---- system handling ---------------------------------------------------
    def species_of(self, Z: np.ndarray) -> np.ndarray:
        lut = {z: i for i, z in enumerate(self.species_Z)}
        return np.array([lut.get(int(z), -1) for z in Z], dtype=np.int64)

    @staticmethod
    def prepare_centers(positions: np.ndarray, centers) -> Tuple[np.ndarray, np.ndarray]:
        """Centers as (C,3) and atom index or -1 (SG:966-1003; snap within 1e-12)."""
        if centers is None:
            return positions.copy(), np.arange(positions.shape[0], dtype=np.int64)
        xyz, idx = [], []
        for c in centers:
            if isinstance(c, int):
                xyz.append(positions[c])
                idx.append(int(c))
            else:
                cc = np.asarray(c, dtype=np.float64).reshape(3)
                xyz.append(cc)
                j = -1
                if positions.shape[0] > 0:
                    d2 = ((positions - cc[None, :]) ** 2).sum(1)
                    k = int(np.argmin(d2))
                    if d2[k] <= 1e-12 * 1e-12:
                        j = k
                idx.append(j)
        return np.stack(xyz), np.array(idx, dtype=np.int64)

    def neighbour_positions(self, positions, Z, cell, pbc):
        if self.periodic:
            if cell is None or pbc is None:
                raise ValueError("periodic=True requires cell and pbc.")
            pos_ext, Z_ext, shift = extend_periodic(positions, Z, np.asarray(cell, np.float64),
                                                    np.asarray(pbc, bool), self.cutoff)
            return pos_ext, Z_ext, shift.shape[0]
        return positions, Z, 1

    def neighbour_list(self, positions, Z, centers_xyz, cell=None, pbc=None, drop_self=True):
        """Edges (center, ext index) ordered as the reference (SG:1128-1181).

        drop_self: drop edges with r <= 1e-12 (the create() path); the
        closed-form derivative path keeps them (SG:2372-2373).
        """
        pos_ext, Z_ext, n_img = self.neighbour_positions(positions, Z, cell, pbc)
        ci, ni = radius_edges(centers_xyz, pos_ext, self.cutoff)
        rvec = pos_ext[ni] - centers_xyz[ci]
        if drop_self:
            r = np.sqrt((rvec ** 2).sum(1))
            keep = r > 1e-12
            ci, ni, rvec = ci[keep], ni[keep], rvec[keep]
        return ci, ni, rvec, self.species_of(Z_ext[ni]) if ni.size else np.zeros(0, np.int64)

    # ---- radial functions ----------------------------------------------------
    def poly_I_J(self, r: np.ndarray, want_J: bool):
        """Quadrature radial integrals with the clamped Bessel recurrence (SG:1437-1507)."""
        eta, rx, wr = self.eta, self.rx, self.wr
        nmax, Lp1 = self.n_max, self.l_max + 1
        gssT = self.gss.T
        W = np.concatenate([gssT, gssT * rx[:, None]], 1) if want_J else gssT
        E = r.shape[0]
        I = np.empty((E, nmax, Lp1))
        J = np.empty((E, nmax, Lp1)) if want_J else None
        L_top = Lp1 if want_J else Lp1 - 1
        rxw = rx * rx * wr
        chunk = max(1, (1 << 20) // rx.size)
        for e0 in range(0, E, chunk):
            re_ = r[e0:e0 + chunk]
            tt = np.maximum((2.0 * eta) * rx[None, :] * re_[:, None], 1e-12)
            e2 = np.exp(-2.0 * tt)
            oOt = 1.0 / tt
            common = (4.0 * PI) * rxw[None, :] * np.exp(-eta * (rx[None, :] - re_[:, None]) ** 2)
            Ppp = Pp = None
            for l in range(L_top + 1):
                if l == 0:
                    P = common * ((1.0 - e2) * (0.5 * oOt))
                elif l == 1:
                    ie1 = np.maximum((1.0 + e2) * (0.5 * oOt) - (1.0 - e2) * (0.5 * oOt * oOt), 0.0)
                    P = common * ie1
                else:
                    P = np.minimum(np.maximum(Ppp - (2 * l - 1) * oOt * Pp, 0.0), Pp)
                out = P @ W
                if l < Lp1:
                    I[e0:e0 + chunk, :, l] = out[:, :nmax]
                if want_J and l >= 1:
                    J[e0:e0 + chunk, :, l - 1] = out[:, nmax:]
                Ppp, Pp = Pp, P
        return I, J

    def poly_B_Bhat_r0(self):
        """Exact r->0 limits of B and Bhat for the polynomial basis (SG:1509-1546)."""
        eta, rx, wr = self.eta, self.rx, self.wr
        gssT = self.gss.T
        common0 = (4.0 * PI) * rx * rx * wr * np.exp(-eta * rx * rx)
        B0 = np.empty((self.n_max, self.l_max + 1))
        Bh0 = np.empty_like(B0)
        fl = np.ones_like(rx)
        df = 1.0
        for l in range(self.l_max + 1):
            dfn = df * (2 * l + 3)
            B0[:, l] = (common0 * fl / df) @ gssT
            Bh0[:, l] = (2.0 * eta) * ((common0 * fl * (2.0 * eta) * rx * rx / dfn) @ gssT - B0[:, l])
            fl = fl * (2.0 * eta) * rx
            df = dfn
        return B0, Bh0

    def radial_B_Bhat(self, r: np.ndarray):
        """B[e,n,l] and Bhat = (1/r) dB/dr of c_nlm(edge) = B * Ysolid (SG:2389-2405, 1548-1595)."""
        if self.rbf == "gto":
            gamma, pref = self.gto_radial_tables()
            Q = pref[None] * np.exp(-gamma[None] * (r * r)[:, None, None])     # (E,l,k)
            B = np.einsum("elk,lnk->enl", Q, self.betas)
            Bh = np.einsum("elk,lnk->enl", (-2.0 * gamma[None]) * Q, self.betas)
            return B, Bh
        tiny = r < 1e-4
        rs = np.where(tiny, 1.0, r)
        I, J = self.poly_I_J(rs, want_J=True)
        Lp1 = self.l_max + 1
        rl = rs[:, None] ** np.arange(Lp1)[None, :]
        B = I / rl[:, None, :]
        Bh = (2.0 * self.eta) * (J - rs[:, None, None] * I) / (rl * rs[:, None])[:, None, :]
        if tiny.any():
            B0, Bh0 = self.poly_B_Bhat_r0()
            B[tiny] = B0
            Bh[tiny] = Bh0
        return B, Bh

    # ---- coefficients --------------------------------------------------------
    def coefficients_create(self, ci, sp, rvec, n_centers):
        """c[C,S,n,L] of the create() path (SG:1294-1435): unit-vector Y times radial."""
        E = rvec.shape[0]
        S, nmax, Lp1 = self.n_species, self.n_max, self.l_max + 1
        L = Lp1 * Lp1
        r = np.sqrt((rvec ** 2).sum(1))
        w = self.weighting(r) if self.weighting is not None else np.ones(E)
        unit = rvec / np.maximum(r, 1e-12)[:, None]
        Y = spherical_harmonics(self.l_max, unit)
        row = ci * S + sp
        out = np.zeros((n_centers * S, nmax, L))
        if self.rbf == "gto":
            gamma, pref = self.gto_radial_tables()
            for l in range(Lp1):
                prim = w[:, None] * (pref[l][None, :] * (r ** l)[:, None] * np.exp(-gamma[l][None, :] * (r * r)[:, None]))
                contrib = prim[:, :, None] * Y[:, None, l * l:(l + 1) ** 2]
                acc = np.zeros((n_centers * S, nmax, 2 * l + 1))
                np.add.at(acc, row, contrib)
                out[:, :, l * l:(l + 1) ** 2] = np.einsum("ab,rbm->ram", self.betas[l], acc)
        else:
            I, _ = self.poly_I_J(r, want_J=False)
            for l in range(Lp1):
                contrib = (w[:, None] * I[:, :, l])[:, :, None] * Y[:, None, l * l:(l + 1) ** 2]
                acc = np.zeros((n_centers * S, nmax, 2 * l + 1))
                np.add.at(acc, row, contrib)
                out[:, :, l * l:(l + 1) ** 2] = acc
        return out.reshape(n_centers, S, nmax, L)

    def add_self_terms(self, c, center_idx, Z):
        """SG:821-850: l=0 self term at (center, species of its atom)."""
        for k, a in enumerate(center_idx):
            if a >= 0:
                s = self.Z_to_sp.get(int(Z[a]), -1)
                if s >= 0:
                    c[k, s, :, 0] += self.self_l0

    def coefficients_solid(self, ci, sp, rvec, n_centers):
        """c[C,S,n,L] of the derivative path (SG:2414-2423): B * Ysolid, self edges kept."""
        S, nmax, Lp1 = self.n_species, self.n_max, self.l_max + 1
        L = Lp1 * Lp1
        r = np.sqrt((rvec ** 2).sum(1))
        B, _ = self.radial_B_Bhat(r)
        w = self.weighting(r) if self.weighting is not None else np.ones(r.shape[0])
        Ys = solid_harmonics(self.l_max, rvec)
        lofm = np.array([l for l in range(Lp1) for _ in range(2 * l + 1)])
        cE = w[:, None, None] * B[:, :, lofm] * Ys[:, None, :]
        acc = np.zeros((n_centers * S, nmax * L))
        np.add.at(acc, ci * S + sp, cE.reshape(cE.shape[0], -1))
        return acc.reshape(n_centers, S, nmax, L)

    # ---- output assembly -----------------------------------------------------
    @staticmethod
    def prefactor(l: int) -> float:
        """sqrt(8 pi^2 / (2l+1)) (SG:1599-1601)."""
        return math.sqrt(8.0 * PI * PI / (2 * l + 1))

    def cc_perm_scale(self):
        """DScribe cc convention: l=1 stored (z,x,y); pi^-3/2 for l>=2 (SG:1603-1622)."""
        L = (self.l_max + 1) ** 2
        perm = np.arange(L)
        if self.l_max >= 1:
            perm[1:4] = [2, 3, 1]
        scale = np.ones(L)
        scale[4:] = PI ** -1.5
        return perm, scale

    def projection_cc(self, c):
        """SG:1624-1632."""
        perm, scale = self.cc_perm_scale()
        return (c[..., perm] * scale).reshape(c.shape[0], -1)

    def power_spectrum(self, c):
        """Per-center p_nn'l in DScribe order (SG:1634-1689)."""
        C = c.shape[0]
        out = np.empty((C, self.n_power_features()))
        iu, ju = np.triu_indices(self.n_max)
        for l in range(self.l_max + 1):
            cl = c[..., l * l:(l + 1) ** 2]
            P = self.prefactor(l) * np.einsum("cjnm,ckdm->cjknd", cl, cl)
            for (j, jd, l2, diag, s0, s1) in self.feat_slices:
                if l2 != l:
                    continue
                out[:, s0:s1] = P[:, j, j][:, iu, ju] if diag else P[:, j, jd].reshape(C, -1)
        return out

    def descriptor(self, c):
        return self.projection_cc(c) if self.average == "cc" else self.power_spectrum(c)

    # ---- public API ------------------------------------------------------------
    def _unpack(self, system):
        if isinstance(system, dict):
            pos = np.asarray(system["positions"], np.float64)
            Z = np.asarray(system["atomic_numbers"], np.int64)
            cell = system.get("cell")
            pbc = system.get("pbc")
        else:
            pos = np.asarray(system[0], np.float64)
            Z = np.asarray(system[1], np.int64)
            cell = system[2] if len(system) > 2 else None
            pbc = system[3] if len(system) > 3 else None
        bad = set(Z.tolist()) - set(self.species_Z)
        if bad:
            raise ValueError(f"System contains atomic numbers not in species list: {sorted(bad)}")
        return pos, Z, cell, pbc

    def create(self, system, centers=None, return_coeffs=False):
        """The create()/_evaluate forward (SG:1759-1810), float64."""
        pos, Z, cell, pbc = self._unpack(system)
        cxyz, cidx = self.prepare_centers(pos, centers)
        ci, ni, rvec, sp = self.neighbour_list(pos, Z, cxyz, cell, pbc, drop_self=True)
        c = self.coefficients_create(ci, sp, rvec, cxyz.shape[0])
        self.add_self_terms(c, cidx, Z)
        d = self.descriptor(c)
        return (d, c) if return_coeffs else d

    def _included(self, n_atoms, include, exclude):
        idx = list(range(n_atoms)) if include is None else [int(i) for i in include]
        if exclude is not None:
            ex = set(int(i) for i in exclude)
            idx = [i for i in idx if i not in ex]
        return idx

    def edge_gradients(self, rvec, r):
        """g[e,d,n,lm] = d/d rvec_d [w B_nl Ysolid_lm] (SG:2563-2567, plus the weighting term)."""
        Lp1 = self.l_max + 1
        B, Bh = self.radial_B_Bhat(r)
        lofm = np.array([l for l in range(Lp1) for _ in range(2 * l + 1)])
        Ys, Yg = solid_harmonics(self.l_max, rvec, grad=True)
        if self.weighting is not None:
            w = self.weighting(r)
            dw = self.weighting.deriv(r)
            with np.errstate(divide="ignore", invalid="ignore"):
                dwr = np.where(r > 0, dw / np.where(r > 0, r, 1.0), 0.0)
            Bh = w[:, None, None] * Bh + dwr[:, None, None] * B
            B = w[:, None, None] * B
        Bl, Bhl = B[:, :, lofm], Bh[:, :, lofm]
        g = rvec[:, :, None, None] * (Bhl[:, None] * Ys[:, None, None, :]) + Bl[:, None] * Yg[:, :, None, :]
        return g   # (E,3,n,L)

    def derivatives(self, system, centers=None, include=None, exclude=None, attach=False,
                    return_descriptor=True):
        """Dense analytical Jacobian (C, n_inc, 3, n_feat), float64.

        Non-periodic, unweighted, attach=False: the closed form of
        SG:2270-2626 (self edges kept, fixed centers).  Otherwise the
        same closed form generalised to what the reference obtains by
        autograd (SG:2007-2113): every periodic image edge of atom a adds to
        row a, the weighting derivative enters the edge gradient, and with
        attach=True an index center moves with its atom (minus the sum of its
        edge gradients).  The descriptor returned is the one each reference
        path returns (closed form: from the kept-self-edge coefficients;
        general path: the create() forward).
        """
        pos, Z, cell, pbc = self._unpack(system)
        n_atoms = pos.shape[0]
        idx = self._included(n_atoms, include, exclude)
        n_inc = len(idx)
        inv = -np.ones(n_atoms, np.int64)
        inv[idx] = np.arange(n_inc)
        cxyz, cidx = self.prepare_centers(pos, centers)
        if attach and centers is not None:
            cidx_attach = np.array([int(c) if isinstance(c, int) else -1 for c in centers], np.int64)
        elif attach:
            cidx_attach = np.arange(n_atoms)
        else:
            cidx_attach = None
        C = cxyz.shape[0]
        S, nmax, Lp1 = self.n_species, self.n_max, self.l_max + 1
        L = Lp1 * Lp1
        closed = (not self.periodic) and self.weighting is None and not attach
        ci, ni, rvec, sp = self.neighbour_list(pos, Z, cxyz, cell, pbc, drop_self=False)
        atom = ni % n_atoms
        r = np.sqrt((rvec ** 2).sum(1))
        if closed:
            c = self.coefficients_solid(ci, sp, rvec, C)
        else:
            keep = r > 1e-12
            c = self.coefficients_create(ci[keep], sp[keep], rvec[keep], C)
            self.add_self_terms(c, cidx, Z)
        g = self.edge_gradients(rvec, r)                       # (E,3,n,L)
        D = np.zeros((C * n_inc * S, 3, nmax, L))
        op = inv[atom]
        v = op >= 0
        np.add.at(D, ((ci * n_inc + op) * S + sp)[v], g[v])
        if cidx_attach is not None:
            # center-motion term: d rvec / d x_center = -I for every edge of that center
            ca = cidx_attach[ci]
            opc = np.where(ca >= 0, inv[np.maximum(ca, 0)], -1)
            vc = opc >= 0
            np.add.at(D, ((ci * n_inc + opc) * S + sp)[vc], -g[vc])
        D = D.reshape(C, n_inc, S, 3, nmax, L)
        n_feat = self.get_number_of_features()
        if self.average == "cc":
            perm, scale = self.cc_perm_scale()
            deriv = (D[..., perm] * scale).transpose(0, 1, 3, 2, 4, 5).reshape(C, n_inc, 3, n_feat)
        else:
            deriv = np.zeros((C, n_inc, 3, n_feat))
            iu, ju = np.triu_indices(nmax)
            for l in range(Lp1):
                Cl = c[..., l * l:(l + 1) ** 2]
                Dl = D[..., l * l:(l + 1) ** 2]
                X = np.einsum("cajdnm,ckqm->cadjnkq", Dl, Cl)
                PD = self.prefactor(l) * (X + X.transpose(0, 1, 2, 5, 6, 3, 4))
                for (j, jd, l2, diag, s0, s1) in self.feat_slices:
                    if l2 != l:
                        continue
                    blk = PD[:, :, :, j, :, jd, :]
                    deriv[..., s0:s1] = blk[..., iu, ju] if diag else blk.reshape(C, n_inc, 3, -1)
        if not return_descriptor:
            return deriv
        return deriv, self.descriptor(c)

    def derivatives_numerical(self, system, centers=None, include=None, exclude=None,
                              attach=False, d=1e-4):
        """Central finite differences with fixed centers (SG:1871-2003)."""
        pos, Z, cell, pbc = self._unpack(system)
        n_atoms = pos.shape[0]
        idx = self._included(n_atoms, include, exclude)
        cxyz, _ = self.prepare_centers(pos, centers)
        base = [cxyz[k].copy() for k in range(cxyz.shape[0])]
        if attach:
            attached = list(range(cxyz.shape[0])) if centers is None else \
                [int(c) if isinstance(c, int) else -1 for c in centers]
        else:
            attached = [-1] * cxyz.shape[0]
        out = np.zeros((cxyz.shape[0], len(idx), 3, self.get_number_of_features()))
        for ai, a in enumerate(idx):
            for comp in range(3):
                fs = []
                for sgn in (1.0, -1.0):
                    p = pos.copy()
                    p[a, comp] += sgn * d
                    sysd = {"positions": p, "atomic_numbers": Z, "cell": cell, "pbc": pbc}
                    cen = [p[t] if t == a else b for t, b in zip(attached, base)]
                    fs.append(self.create(sysd, cen))
                out[:, ai, comp, :] = (fs[0] - fs[1]) / (2.0 * d)
        return out

    def vjp(self, system, grad_desc, centers=None):
        """Adjoint of the descriptor w.r.t. positions: sum_{c,f} grad[c,f] dp_cf/dx (N,3).

        No reference counterpart; pinned by contracting :meth:`derivatives`
        with grad_desc.  Fixed centers, self edges kept (closed-form path).
        """
        pos, Z, cell, pbc = self._unpack(system)
        n_atoms = pos.shape[0]
        cxyz, cidx = self.prepare_centers(pos, centers)
        C = cxyz.shape[0]
        S, nmax, Lp1 = self.n_species, self.n_max, self.l_max + 1
        L = Lp1 * Lp1
        ci, ni, rvec, sp = self.neighbour_list(pos, Z, cxyz, cell, pbc, drop_self=False)
        r = np.sqrt((rvec ** 2).sum(1))
        if (not self.periodic) and self.weighting is None:
            c = self.coefficients_solid(ci, sp, rvec, C)
        else:
            keep = r > 1e-12
            c = self.coefficients_create(ci[keep], sp[keep], rvec[keep], C)
            self.add_self_terms(c, cidx, Z)
        G = self.coeff_adjoint(c, grad_desc)                  # (C,S,n,L)
        g = self.edge_gradients(rvec, r)                       # (E,3,n,L)
        contrib = np.einsum("ednl,enl->ed", g, G[ci, sp])
        F = np.zeros((n_atoms, 3))
        np.add.at(F, ni % n_atoms, contrib)
        return F

    def coeff_adjoint(self, c, grad_desc):
        """dL/dc for L = sum grad_desc * descriptor(c)."""
        C = c.shape[0]
        S, nmax, Lp1 = self.n_species, self.n_max, self.l_max + 1
        grad_desc = np.broadcast_to(np.asarray(grad_desc, np.float64), (C, self.get_number_of_features()))
        if self.average == "cc":
            perm, scale = self.cc_perm_scale()
            G = np.zeros_like(c)
            gd = grad_desc.reshape(C, S, nmax, -1)
            G[..., perm] = gd * scale
            return G
        G = np.zeros_like(c)
        ft = self.feature_table()
        for f, (j, jd, a, b, l) in enumerate(ft):
            pf = self.prefactor(l)
            sl = slice(l * l, (l + 1) ** 2)
            wv = grad_desc[:, f][:, None] * pf
            G[:, j, a, sl] += wv * c[:, jd, b, sl]
            G[:, jd, b, sl] += wv * c[:, j, a, sl]
        return G
