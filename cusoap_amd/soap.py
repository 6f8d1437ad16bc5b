"""MI355X-native drop-in for the reference ``SOAP`` class (src/cusoap/soap_generator.py).

Host-side parameter resolution, basis tables and input adaptation mirror the
reference (soap_generator.py:265-1013); every per-atom stage -- neighbour
search, density coefficients, power spectrum / cc, dense Jacobian and the
adjoint -- runs in the hand-written HIP kernels of ``libcusoap_amd.so``
through its C ABI.  There is no CPU fallback: ``create`` and ``derivatives``
raise when no HIP device or no built library is available.
"""

from __future__ import annotations

import math
import time
from dataclasses import dataclass
from typing import Any, Dict, List, Optional, Sequence, Tuple, Union

import ctypes as C
import numpy as np
import torch

from . import _lib
from .systems import SYMBOLS

PI = math.pi
Y00 = 1.0 / math.sqrt(4.0 * math.pi)


# ----------------------------------------------------------------------------
# host helpers (restate soap_generator.py:265-470)
# ----------------------------------------------------------------------------

def _torch_dtype_from_str(dtype: str) -> torch.dtype:
    """soap_generator.py:265-270."""
    if dtype == "float32":
        return torch.float32
    if dtype == "float64":
        return torch.float64
    raise ValueError("dtype must be 'float32' or 'float64'")


def _host(x, dtype=None) -> np.ndarray:
    if isinstance(x, torch.Tensor):
        x = x.detach().cpu().numpy()
    return np.asarray(x, dtype=dtype)


def system_to_arrays(system: Any, keep_device: bool = False):
    """ASE-like object, dict or (positions, Z[, cell[, pbc]]) -> arrays (soap_generator.py:291-332).

    Positions come back as float64 host arrays, or -- with keep_device and a
    GPU tensor as input -- as the float64 device tensor itself (no host copy).
    """
    if isinstance(system, (tuple, list)):
        if len(system) < 2:
            raise ValueError("Tuple/list system must be (positions, atomic_numbers, [cell], [pbc]).")
        pos, Z = system[0], system[1]
        cell = system[2] if len(system) >= 3 else None
        pbc = system[3] if len(system) >= 4 else None
    elif isinstance(system, dict):
        pos, Z = system["positions"], system["atomic_numbers"]
        cell, pbc = system.get("cell"), system.get("pbc")
    elif hasattr(system, "get_positions") and hasattr(system, "get_atomic_numbers"):
        pos, Z = system.get_positions(), system.get_atomic_numbers()
        cell = pbc = None
        if hasattr(system, "get_cell"):
            cl = system.get_cell()
            cell = getattr(cl, "array", cl)
        if hasattr(system, "get_pbc"):
            pbc = system.get_pbc()
    else:
        raise TypeError("Unsupported system type. Provide ASE.Atoms-like, dict, or (positions, Z, cell, pbc).")
    if keep_device and isinstance(pos, torch.Tensor) and pos.is_cuda:
        pos = pos.detach().to(torch.float64).reshape(-1, 3).contiguous()
    else:
        pos = np.ascontiguousarray(_host(pos, np.float64).reshape(-1, 3))
    Z = _host(Z, np.int64).reshape(-1)
    cell = None if cell is None else _host(cell, np.float64).reshape(3, 3)
    pbc = None if pbc is None else _host(pbc, bool).reshape(3)
    return pos, Z, cell, pbc


def linspace_ref(start: float, end: float, n: int) -> np.ndarray:
    """torch.linspace in float64 (the reference builds its GTO widths with it, soap_generator.py:694).

    step = (end - start) / (n - 1); below the midpoint fma(step, i, start), above it
    fma(-step, n - 1 - i, end), each correctly rounded once (evaluated exactly here).
    """
    from fractions import Fraction
    if n == 1:
        return np.array([float(start)])
    step = (end - start) / (n - 1)
    half = n // 2
    return np.array([float(Fraction(start) + Fraction(step) * i) if i < half
                     else float(Fraction(end) - Fraction(step) * (n - 1 - i)) for i in range(n)])


def gauss_legendre(n: int) -> Tuple[np.ndarray, np.ndarray]:
    """Golub-Welsch nodes and weights on [-1, 1] (soap_generator.py:350-362)."""
    k = np.arange(1, n, dtype=np.float64)
    beta = k / np.sqrt(4 * k * k - 1)
    J = np.diag(beta, 1) + np.diag(beta, -1)
    x, V = np.linalg.eigh(J)
    return x, 2 * (V[0, :] ** 2)


def perpendicular_lengths(cell: np.ndarray) -> np.ndarray:
    """soap_generator.py:369-379."""
    a, b, c = cell
    v = float(np.dot(a, np.cross(b, c)))
    return np.array([abs(v) / (np.linalg.norm(np.cross(b, c)) + 1e-12),
                     abs(v) / (np.linalg.norm(np.cross(c, a)) + 1e-12),
                     abs(v) / (np.linalg.norm(np.cross(a, b)) + 1e-12)])


def pbc_image_shifts(cell: np.ndarray, pbc: np.ndarray, cutoff: float) -> np.ndarray:
    """Cartesian image shifts in meshgrid-'ij' order over -n..n (soap_generator.py:382-396, 1016-1017)."""
    h = perpendicular_lengths(cell)
    n_rep = [int(math.ceil(cutoff / (h[i] + 1e-12))) if bool(pbc[i]) else 0 for i in range(3)]
    ox, oy, oz = np.meshgrid(*[np.arange(-n, n + 1) for n in n_rep], indexing="ij")
    o = np.stack([ox.ravel(), oy.ravel(), oz.ravel()], 1).astype(np.float64)
    return np.ascontiguousarray((o[:, 0:1] * cell[0] + o[:, 1:2] * cell[1]) + o[:, 2:3] * cell[2])


@dataclass
class Weighting:
    """DScribe-style weighting (soap_generator.py:403-470); evaluated on the device."""
    function: Optional[str] = None
    r0: float = 1.0
    c: float = 1.0
    d: float = 1.0
    m: float = 1.0
    w0: float = 1.0

    def infer_rcut(self, threshold: float = 1e-2) -> float:
        if self.function is None:
            raise ValueError("Cannot infer r_cut when weighting.function is None; provide r_cut explicitly.")
        if self.function == "poly":
            return float(self.r0)
        if self.function in ("pow", "exp"):
            val = self.c / threshold - self.d
            if val <= 0:
                raise ValueError(f"Invalid weighting: c/threshold - d must be >0 for {self.function}.")
            if self.function == "pow":
                return float(self.r0 * (val ** (1.0 / self.m)))
            return float(-self.r0 * math.log(val))
        raise ValueError(f"Unknown weighting function: {self.function}")

    @property
    def code(self) -> int:
        return {None: _lib.W_NONE, "poly": _lib.W_POLY, "pow": _lib.W_POW, "exp": _lib.W_EXP}[self.function]


# ----------------------------------------------------------------------------
# the SOAP class
# ----------------------------------------------------------------------------

_PRECISIONS = {"float64": _lib.PREC_F64, "mixed": _lib.PREC_MIXED, "float32": _lib.PREC_F32}
_SPECTRA = {"exact": None, "fp16": (_lib.MFMA_F16, 0), "fp16x3": (_lib.MFMA_F16, 1),
            "bf16": (_lib.MFMA_BF16, 0), "bf16x3": (_lib.MFMA_BF16, 1)}


class _NL:
    """CSR neighbour list on the device: offsets (C+1) int64, nbr (E) int32 ext indices."""

    def __init__(self, offsets, nbr, n_edges, max_count):
        self.offsets, self.nbr, self.n_edges, self.max_count = offsets, nbr, n_edges, max_count


class _Ctx:
    """Device-resident inputs of one system."""


class SOAP:
    """DScribe-compatible SOAP on MI355X (signature of soap_generator.py:524-541).

    Extra keyword ``precision`` selects the arithmetic of the coefficient stage:
    "float64" (default, what the reference computes internally), "mixed"
    (float32 harmonics and accumulation, float64 radial functions) or
    "float32".  Output tensors are ``dtype`` as in the reference.

    Extra keyword ``radial``: "exact" (default), "table", "chebyshev" or "learned".
    "table" tabulates the exact float64 GTO radial functions and slopes once
    (``table_intervals`` cubic-Hermite intervals on [0, cutoff]; by default 256 for the
    learned model and 512 for the exact functions in the float precisions, whose float32
    slopes get noisier on finer tables, and 1024 in float64:
    cusoap_amd.surrogate.default_table_intervals).  The coefficient
    and adjoint kernels then interpolate per edge, with no exponentials and no
    Loewdin contraction per edge.  "chebyshev" (float precisions, unweighted GTO,
    n_max <= 8, l_max <= 6, at most 4 species) expands the exact functions in
    piecewise Chebyshev series so that the coefficient stage runs on the matrix
    cores (cusoap_amd.chebyshev); the adjoint uses the table.  "learned" trains (or
    loads, ``surrogate=<checkpoint path>``) per-species-pair networks for the radial
    functions on the device and feeds the gated model to both stages the same ways
    (cusoap_amd.surrogate).  ``surrogate`` may also be a RadialSurrogate or a dict
    of its constructor and train() options.  The dense Jacobian always uses the
    exact radial functions.

    Attribute ``ordered_forces`` (default False): when True and every atom is a
    center, ``vjp`` / ``descriptor_and_vjp`` sum each atom's force in a fixed order
    (cusoap_edge_reverse + cusoap_vjp_ordered), so the forces are the same bits on
    every run; otherwise they are summed with float64 atomics.

    Extra keyword ``spectrum``: arithmetic of the power-spectrum contraction.
    "exact" (default) uses the vector units in the coefficient precision.
    "fp16" / "bf16" use the matrix cores with fp16 / bf16 operands and float32
    accumulation.  "fp16x3" / "bf16x3" split each operand into hi + lo and sum
    three products (cusoap_power_spectrum_mfma).
    """

    def __init__(self, r_cut=None, n_max=None, l_max=None, sigma=1.0, rbf="gto", weighting=None,
                 crossover=True, average="off", species=None, periodic=False, sparse=False,
                 dtype="float32", device=None, quad_n=100, max_num_neighbors=None,
                 precision="float64", radial="exact", surrogate=None, table_intervals=None, spectrum="exact",
                 _betas=None):
        if n_max is None or l_max is None:
            raise ValueError("n_max and l_max must be provided.")
        if species is None or len(species) == 0:
            raise ValueError("species must be provided (list of atomic numbers or symbols).")
        # constructor arguments, for objects of the same basis in another form (_sibling)
        self._init_kwargs = dict(r_cut=r_cut, n_max=n_max, l_max=l_max, sigma=sigma, rbf=rbf, weighting=weighting,
                                 crossover=crossover, average=average, species=list(species), periodic=periodic,
                                 sparse=sparse, dtype=dtype, device=device, quad_n=quad_n,
                                 max_num_neighbors=max_num_neighbors, precision=precision,
                                 table_intervals=table_intervals, spectrum=spectrum, _betas=_betas)
        self._dtype_str = dtype
        self.dtype = _torch_dtype_from_str(dtype)
        self.device = torch.device(device) if device is not None else \
            torch.device("cuda" if torch.cuda.is_available() else "cpu")
        zs = []
        for s in species:
            if isinstance(s, str):
                if s not in SYMBOLS:
                    raise ValueError(f"Cannot resolve species symbol '{s}'. Use atomic numbers.")
                zs.append(SYMBOLS[s])
            else:
                zs.append(int(s))
        self._species_Z = sorted(set(zs))
        self._atomic_number_set = set(self._species_Z)
        self._Z_to_species_index = {z: i for i, z in enumerate(self._species_Z)}
        self.n_species = len(self._species_Z)
        self._n_max, self._l_max = int(n_max), int(l_max)
        self._sigma = float(sigma)
        self._eta = 1.0 / (2.0 * self._sigma * self._sigma)
        self._rbf = str(rbf).lower()
        self.crossover = bool(crossover)
        self.average = str(average)
        self.periodic = bool(periodic)
        self.sparse = bool(sparse)
        self.quad_n = int(quad_n)
        self.max_num_neighbors = max_num_neighbors   # exhaustive search: accepted, no truncation
        if precision not in _PRECISIONS:
            raise ValueError("precision must be 'float64', 'mixed' or 'float32'")
        self.precision = precision
        if self.average not in ("off", "inner", "outer", "cc"):
            raise ValueError("average must be one of: 'off', 'inner', 'outer', 'cc' (cc = projection coeffs).")
        self._weighting_dict = dict(weighting) if weighting is not None else None
        thr = 1e-2
        if self._weighting_dict is not None:
            thr = float(self._weighting_dict.get("threshold", 1e-2))
            self._weighting = Weighting(**{k: v for k, v in self._weighting_dict.items() if k != "threshold"})
        else:
            self._weighting = None
        if r_cut is None:
            if self._weighting is None:
                raise ValueError("Either r_cut or weighting must be provided.")
            self._r_cut = float(self._weighting.infer_rcut(threshold=thr))
        else:
            self._r_cut = float(r_cut)
        self._cutoff_padding = float(self._sigma * math.sqrt(-2.0 * math.log(1e-3)))
        self._cutoff = self._r_cut + self._cutoff_padding
        if self._n_max > 16 or self._l_max > 9:
            raise NotImplementedError("cusoap_amd kernels are compiled for n_max <= 16 and l_max <= 9")
        if self._rbf not in ("gto", "polynomial"):
            raise ValueError("rbf must be 'gto' or 'polynomial'.")
        self._on_gpu = self.device.type == "cuda" and torch.cuda.is_available()
        if self._rbf == "gto":
            self._init_gto_basis(_betas)
        else:
            self._init_poly_basis()
        self._init_self_contributions()
        self.n_features = int(self.get_number_of_features())
        self._build_feature_tables()
        self._basis_struct = None
        self._basis_learned = None
        self._species_cache = None   # device tensors derived from (Z, cell, pbc) of the last system
        self._last_nl_backend = "hip-cell-list"
        # neighbour-search buffer sizes (workspace bytes, edges) of the last search per
        # (centers, atoms, images, cutoff, drop_self): lets one call build the next list
        self._nl_sizes = {}
        if radial not in ("exact", "table", "learned", "chebyshev"):
            raise ValueError("radial must be 'exact', 'table', 'learned' or 'chebyshev'")
        if spectrum not in _SPECTRA:
            raise ValueError(f"spectrum must be one of {sorted(_SPECTRA)}")
        self.spectrum = spectrum
        self.radial = radial
        # True: fixed-order force sums in the adjoint when every atom is a center (forces
        # the same bits on every run; cusoap_edge_reverse + cusoap_vjp_ordered, ~0.35 ms more
        # on a 98,304-atom water box step); False (default): float64 atomics
        self.ordered_forces = False
        self._side_stream = None
        self.surrogate = None
        self._rad_table = None
        self._rad_cheb = None
        if table_intervals is None:
            from .surrogate import default_table_intervals
            table_intervals = default_table_intervals(radial, precision)
        if radial == "learned":
            self._init_learned_radial(surrogate, int(table_intervals))
        elif radial == "table":
            from .surrogate import radial_table
            self._require_gpu()
            if self._rbf != "gto":
                raise NotImplementedError("radial='table' covers the GTO basis")
            self._rad_intervals = int(table_intervals)
            self._rad_table = radial_table(self, self._rad_intervals)
        elif radial == "chebyshev":
            from .surrogate import radial_table
            self._init_chebyshev_radial()
            # the adjoint edge pass reads the exact functions from the cubic-Hermite table
            self._rad_intervals = int(table_intervals)
            self._rad_table = radial_table(self, self._rad_intervals)

    def _sibling(self, **overrides) -> "SOAP":
        """A SOAP of the same basis (same Loewdin matrices) with some arguments replaced."""
        kw = dict(self._init_kwargs, _betas=self._betas)
        kw.update(overrides)
        return SOAP(**kw)

    def _init_learned_radial(self, surrogate, n_intervals: int):
        """Train or load the radial networks, gate them, tabulate the radial model."""
        from .surrogate import RadialSurrogate, masked_table, radial_table
        self._require_gpu()
        if self._rbf != "gto":
            raise NotImplementedError("radial='learned' covers the GTO basis")
        if isinstance(surrogate, RadialSurrogate):
            model = surrogate
        else:
            opts = dict(surrogate) if isinstance(surrogate, dict) else {}
            ckpt = surrogate if isinstance(surrogate, str) else opts.pop("checkpoint", None)
            init_keys = ("r_fit", "n_train", "n_heldout", "seed", "deriv_weight", "tol", "tol_slope")
            gate_keys = ("calibration", "bar", "margin")
            gate_opts = {k: opts.pop(k) for k in gate_keys if k in opts}
            model = RadialSurrogate(self, **{k: opts.pop(k) for k in init_keys if k in opts})
            if ckpt:
                model.load(ckpt)
            else:
                model.train(**opts)
            model.gate_options = gate_opts
        self._rad_intervals = n_intervals
        self.surrogate = model
        channels = model.gate(self, n_intervals)
        self._learned_channels = channels
        self._learned_use_pair = channels.reshape(channels.shape[0], -1).any(1).astype(np.int32)
        exact = radial_table(self, n_intervals)
        net = radial_table(self, n_intervals, model=model, use_pair=self._learned_use_pair)
        self._rad_table = masked_table(exact, net, channels)
        # the float coefficient stage reads the same gated model as Chebyshev series on the
        # matrix cores (the adjoint edge pass keeps the table) for one or two species; with
        # more, the eight or more (species, interval) groups of a center each pad to whole
        # operand blocks and the table kernel was measured faster (H/C/N/O box: 0.43 against
        # 0.54 ms; water box 2.11 against 1.53 ms and the 1M cluster 17.2 against 13.1 ms
        # the other way)
        from . import chebyshev as cheb
        if self.precision != "float64" and cheb.supported(self) is None and self.n_species <= 2:
            self._init_chebyshev_radial(coefficients=cheb.fit_learned(self, model, channels))

    def _init_chebyshev_radial(self, coefficients=None):
        """Chebyshev series of the radial functions for the matrix-core kernels (exact
        functions, or the given per-slot coefficients (a, d, K))."""
        from . import chebyshev as cheb
        self._require_gpu()
        why = cheb.supported(self)
        if why is not None:
            raise NotImplementedError(why)
        if self.precision == "float64":
            raise ValueError("the Chebyshev radial form computes in float32 (fp16 hi/lo matrix-core products): "
                             "use precision='mixed'")
        if coefficients is None:
            a, d, ef, ed = cheb.fit_exact(self)
            if ef > cheb.TOL_VALUE or ed > cheb.TOL_SLOPE:
                raise NotImplementedError(f"Chebyshev radial form: {cheb.TERMS} terms per interval fit this basis to "
                                          f"{ef:.1e} (values) and {ed:.1e} (slopes) only; use radial='table'")
            self._cheb_fit_err = (ef, ed)
        else:
            a, d = coefficients
        self._cheb_terms = cheb.TERMS
        self._cheb_split = cheb.split_radius(self._cutoff)
        self._rad_cheb = cheb.device_table(self, a, d)
        self._cheb_frag, self._cheb_rowinv = cheb.contraction_fragments(self, a)
        self._basis_learned = None

    # ---- basis (soap_generator.py:690-819) ---------------------------------
    def _init_gto_basis(self, betas=None):
        a = linspace_ref(1.0, self._r_cut, self._n_max)
        alphas, S = [], []
        for l in range(self._l_max + 1):
            alpha_l = -np.log(1e-3 / (a ** l)) / (a ** 2)
            m = alpha_l[:, None] + alpha_l[None, :]
            S.append(0.5 * math.exp(math.lgamma(l + 1.5)) * (m ** (-(l + 1.5))))
            alphas.append(alpha_l)
        self._alphas = np.stack(alphas)
        S = np.ascontiguousarray(np.stack(S))
        if betas is not None:
            self._betas = np.asarray(betas, dtype=np.float64)
        elif self._on_gpu:
            self._betas = self._lowdin_gpu(S)
        else:
            w, V = np.linalg.eigh(S)
            w = np.maximum(w, 1e-12)
            self._betas = np.einsum("lik,lk,ljk->lij", V, 1.0 / np.sqrt(w), V)
        p = self._alphas + self._eta
        self._gto_gamma = self._alphas * self._eta / p
        lv = np.arange(self._l_max + 1, dtype=np.float64)[:, None]
        self._gto_pref = (PI ** 1.5) * (self._eta / p) ** lv * p ** -1.5

    def _lowdin_gpu(self, S: np.ndarray) -> np.ndarray:
        """Loewdin S^-1/2 in float64 on the device (replaces soap_generator.py:339-343)."""
        lib = _lib.load()
        Sd = torch.from_numpy(S).to(self.device)
        out = torch.empty_like(Sd)
        _lib.check(lib.cusoap_lowdin_invsqrt_f64(Sd.data_ptr(), S.shape[0], S.shape[1], 1e-12, out.data_ptr(),
                                                 self._stream()), "lowdin_invsqrt_f64")
        return out.cpu().numpy()

    def _init_poly_basis(self):
        """soap_generator.py:710-761: beta = scipy sqrtm(inv(S)) on the host, GL(quad_n) grid."""
        from scipy.linalg import sqrtm
        n, rc = self._n_max, self._r_cut
        S = np.zeros((n, n))
        for i in range(1, n + 1):
            for j in range(1, n + 1):
                S[i - 1, j - 1] = (2 * rc ** (7 + i + j)) / ((5 + i + j) * (6 + i + j) * (7 + i + j))
        betas = sqrtm(np.linalg.inv(S))
        if np.iscomplexobj(betas):
            raise ValueError("Could not calculate normalization factors for the radial basis in the domain of "
                             "real numbers. Lowering the number of radial basis functions (n_max) or increasing "
                             "the radial cutoff (r_cut) is advised.")
        x, w = gauss_legendre(self.quad_n)
        self._rx = rc * 0.5 * (x + 1.0)
        self._wr = rc * 0.5 * w
        rclip = np.clip(self._rx, 0.0, rc)
        fs = np.stack([(rc - rclip) ** (k + 2) for k in range(1, n + 1)])
        self._gss = np.ascontiguousarray(betas) @ fs
        # r -> 0 limits of B and Bhat (soap_generator.py:1509-1546)
        eta, rx, wr = self._eta, self._rx, self._wr
        common0 = (4.0 * PI) * rx * rx * wr * np.exp(-eta * rx * rx)
        L1 = self._l_max + 1
        B0, Bh0 = np.empty((n, L1)), np.empty((n, L1))
        fl, df = np.ones_like(rx), 1.0
        for l in range(L1):
            dfn = df * (2 * l + 3)
            B0[:, l] = (common0 * fl / df) @ self._gss.T
            Bh0[:, l] = (2.0 * eta) * ((common0 * fl * (2.0 * eta) * rx * rx / dfn) @ self._gss.T - B0[:, l])
            fl = fl * (2.0 * eta) * rx
            df = dfn
        self._poly_B0, self._poly_Bh0 = B0, Bh0

    def _init_self_contributions(self):
        """l=0 self coefficient of a Gaussian at the center (soap_generator.py:765-795)."""
        w0 = float(self._weighting.w0) if self._weighting is not None else 1.0
        if self._rbf == "gto":
            p = self._alphas[0] + self._eta
            self._self_l0 = (self._betas[0] @ (w0 * (PI ** 1.5) * p ** -1.5)) * Y00
        else:
            common0 = (4.0 * PI) * (self._rx * self._rx) * np.exp(-self._eta * self._rx ** 2) * self._wr
            self._self_l0 = w0 * (common0 @ self._gss.T) * Y00

    # ---- features (soap_generator.py:908-962, 1215-1291) -----------------------
    def get_number_of_features(self) -> int:
        if self.average == "cc":
            return self.n_species * self._n_max * (self._l_max + 1) ** 2
        return self._n_power_features()

    def _n_power_features(self) -> int:
        if self.crossover:
            ch = self.n_species * self._n_max
            return int((ch * (ch + 1) // 2) * (self._l_max + 1))
        return int(self.n_species * (self._l_max + 1) * (self._n_max * (self._n_max + 1) // 2))

    def _build_feature_tables(self):
        n, S, L1 = self._n_max, self.n_species, self._l_max + 1
        rows = []
        for j in range(S):
            for jd in range(j, S if self.crossover else j + 1):
                for l in range(L1):
                    for a in range(n):
                        for b in range(a if j == jd else 0, n):
                            rows.append((j, jd, a, b, l))
        ft = np.array(rows, dtype=np.int64).reshape(-1, 5)
        if ft.shape[0] != self._n_power_features():
            raise RuntimeError("Feature table size mismatch.")
        self._feat_rows = ft
        self._feat_packed = (ft[:, 0] | (ft[:, 1] << 8) | (ft[:, 2] << 16) | (ft[:, 3] << 22)
                             | (ft[:, 4] << 28)).astype(np.uint32)
        SN = S * n
        pf = -np.ones((L1, SN, SN), dtype=np.int32)
        for f, (j, jd, a, b, l) in enumerate(ft):
            pf[l, j * n + a, jd * n + b] = f
            pf[l, jd * n + b, j * n + a] = f
        self._pair_feature = pf

    def get_location(self, species_pair) -> slice:
        """DScribe-style feature slice of a species pair (soap_generator.py:921-962)."""
        if self.average == "cc":
            raise ValueError("get_location is not defined for average='cc' (projection coefficients).")
        if len(species_pair) != 2:
            raise ValueError("species_pair must be length-2.")

        def toZ(x):
            if isinstance(x, str):
                if x not in SYMBOLS:
                    raise ValueError(f"Cannot resolve species symbol '{x}'.")
                return SYMBOLS[x]
            return int(x)

        Z1, Z2 = toZ(species_pair[0]), toZ(species_pair[1])
        if Z1 not in self._atomic_number_set or Z2 not in self._atomic_number_set:
            raise ValueError("Species not present in descriptor species list.")
        i, j = self._Z_to_species_index[Z1], self._Z_to_species_index[Z2]
        if i > j:
            i, j = j, i
        symm = (self._n_max * (self._n_max + 1) // 2) * (self._l_max + 1)
        unsymm = (self._n_max * self._n_max) * (self._l_max + 1)
        if self.crossover:
            m_symm = i + int(j > i)
            m_unsymm = j + i * self.n_species - i * (i + 1) / 2 - m_symm
            start = int(m_symm * symm + m_unsymm * unsymm)
            return slice(start, start + (symm if i == j else unsymm))
        if i != j:
            raise ValueError("crossover=False: no cross-species blocks exist.")
        return slice(i * symm, i * symm + symm)

    # ---- device plumbing ------------------------------------------------------
    def _require_gpu(self):
        if not self._on_gpu:
            raise RuntimeError("cusoap_amd runs on a HIP GPU only (device=%s, cuda available=%s); there is no CPU path"
                               % (self.device, torch.cuda.is_available()))
        _lib.load()

    def _stream(self) -> int:
        return torch.cuda.current_stream(self.device).cuda_stream

    def _dev(self, a: np.ndarray, dtype=None) -> torch.Tensor:
        t = torch.from_numpy(np.ascontiguousarray(a))
        if dtype is not None:
            t = t.to(dtype)
        return t.to(self.device, non_blocking=False)

    def _basis(self) -> "_lib.Basis":
        """Basis tables for the kernels, with the radial table when radial='learned'."""
        b = self._basis_exact()
        if self._rad_table is None and self._rad_cheb is None:
            return b
        if self._basis_learned is None:
            bl = _lib.Basis.from_buffer_copy(b)
            if self._rad_cheb is not None:
                bl.rad_cheb = self._rad_cheb.data_ptr()
                bl.cheb_terms = self._cheb_terms
                bl.cheb_split = self._cheb_split
                bl.cheb_frag = self._cheb_frag.data_ptr()
                bl.cheb_rowinv = self._cheb_rowinv.data_ptr()
            if self._rad_table is not None:
                bl.rad_table = self._rad_table.data_ptr()
                bl.rad_intervals = self._rad_intervals
                bl.rad_rmax = self._cutoff
            self._basis_learned = bl
        return self._basis_learned

    def _basis_exact(self) -> "_lib.Basis":
        if self._basis_struct is not None:
            return self._basis_struct
        self._require_gpu()
        b = _lib.Basis()
        b.n_species, b.n_max, b.l_max = self.n_species, self._n_max, self._l_max
        b.rbf = _lib.RBF_GTO if self._rbf == "gto" else _lib.RBF_POLY
        b.eta, b.cutoff = self._eta, self._cutoff
        w = self._weighting
        b.weighting = _lib.W_NONE if w is None else w.code
        if w is not None:
            b.w_r0, b.w_c, b.w_d, b.w_m, b.w0 = w.r0, w.c, w.d, w.m, w.w0
        else:
            b.w_r0, b.w_c, b.w_d, b.w_m, b.w0 = 1.0, 1.0, 1.0, 1.0, 1.0
        keep = {}
        if self._rbf == "gto":
            keep["gamma"] = self._dev(self._gto_gamma)
            keep["pref"] = self._dev(self._gto_pref)
            keep["beta"] = self._dev(self._betas)
            b.gto_gamma, b.gto_pref, b.gto_beta = (keep[k].data_ptr() for k in ("gamma", "pref", "beta"))
            b.quad_n = 0
        else:
            keep["rx"] = self._dev(self._rx)
            keep["rxw"] = self._dev((4.0 * PI) * (self._rx * self._rx * self._wr))
            keep["gss"] = self._dev(self._gss)
            keep["B0"] = self._dev(self._poly_B0)
            keep["Bh0"] = self._dev(self._poly_Bh0)
            b.poly_rx, b.poly_rxw, b.poly_gss = keep["rx"].data_ptr(), keep["rxw"].data_ptr(), keep["gss"].data_ptr()
            b.poly_B0, b.poly_Bh0 = keep["B0"].data_ptr(), keep["Bh0"].data_ptr()
            b.quad_n = self.quad_n
        keep["self_l0"] = self._dev(self._self_l0)
        b.self_l0 = keep["self_l0"].data_ptr()
        keep["ft"] = self._dev(self._feat_packed.view(np.int32))
        keep["pf"] = self._dev(self._pair_feature)
        self._keep = keep
        self._basis_struct = b
        return b

    # ---- centers (soap_generator.py:966-1013) ----------------------------------
    def prepare_centers(self, positions: np.ndarray, centers) -> Tuple[np.ndarray, np.ndarray]:
        """Centers as (C,3) float64 and the atom index of each (or -1); coordinates snap within 1e-12."""
        if centers is None:
            return positions, np.arange(positions.shape[0], dtype=np.int64)
        centers = list(centers)
        if len(centers) == 0:
            raise RuntimeError("stack expects a non-empty TensorList")
        xyz = np.empty((len(centers), 3))
        idx = np.empty(len(centers), dtype=np.int64)
        for k, c in enumerate(centers):
            if isinstance(c, int):
                xyz[k] = positions[c]
                idx[k] = int(c)
            else:
                cc = _host(c, np.float64)
                if cc.size != 3:
                    raise RuntimeError(f"shape '[3]' is invalid for input of size {cc.size}")
                xyz[k] = cc.reshape(3)
                idx[k] = -1
                if positions.shape[0] > 0:
                    d2 = ((positions - xyz[k][None, :]) ** 2).sum(1)
                    j = int(np.argmin(d2))
                    if d2[j] <= 1e-12 * 1e-12:
                        idx[k] = j
        return xyz, idx

    def _context(self, system, centers, mark=None) -> _Ctx:
        """Device-resident inputs; ``mark(stage)`` (optional) is called after the input
        conversion ("to_tensors") and after the centers ("centers")."""
        self._require_gpu()
        pos, Z, cell, pbc = system_to_arrays(system, keep_device=True)
        if centers is None and self._species_cache is not None:
            x = self._context_cached(pos, Z, cell, pbc)
            if x is not None:
                if mark:
                    mark("to_tensors")
                    mark("centers")
                return x
        bad = set(np.unique(Z).tolist()) - self._atomic_number_set
        if bad:
            raise ValueError(f"System contains atomic numbers not in species list: {sorted(bad)}")
        x = _Ctx()
        if isinstance(pos, torch.Tensor):
            x.pos = pos.to(self.device)
            x.n_atoms = pos.shape[0]
        else:
            x.pos = self._dev(pos)
            x.n_atoms = pos.shape[0]
        if mark:
            mark("to_tensors")
        if centers is None:
            x.centers = x.pos
            cidx = np.arange(x.n_atoms, dtype=np.int64)
        else:
            pos_h = pos.cpu().numpy() if isinstance(pos, torch.Tensor) else pos
            cxyz, cidx = self.prepare_centers(pos_h, centers)
            x.centers = self._dev(np.ascontiguousarray(cxyz))
        if mark:
            mark("centers")
        if self.periodic:
            if cell is None or pbc is None:
                raise ValueError("periodic=True requires cell and pbc.")
            shifts = pbc_image_shifts(cell, pbc, self._cutoff)
        else:
            shifts = np.zeros((1, 3))
        lut = np.full(max(self._species_Z + [int(Z.max()) if Z.size else 0]) + 1, -1, dtype=np.int32)
        for z, i in self._Z_to_species_index.items():
            lut[z] = i
        sp = lut[Z] if Z.size else np.zeros(0, np.int32)
        x.n_centers, x.n_img = int(cidx.shape[0]), shifts.shape[0]
        x.Z_h, x.cidx_h, x.sp_h = Z, cidx, sp
        x.sp = self._dev(sp.astype(np.int32))
        x.shifts = self._dev(shifts)
        self_sp = np.where(cidx >= 0, sp[np.maximum(cidx, 0)] if sp.size else -1, -1).astype(np.int32)
        x.self_sp = self._dev(self_sp)
        if centers is None:
            # the same composition, cell and pbc on the next call (an MD or benchmark loop)
            # reuse these without the host work and the synchronous host-to-device copies
            self._species_cache = (np.array(Z, copy=True),
                                   None if cell is None else np.array(cell, dtype=np.float64, copy=True),
                                   None if pbc is None else np.array(pbc, dtype=bool, copy=True), x)
        return x

    def _context_cached(self, pos, Z, cell, pbc):
        """Context of a system with the cached composition, cell and pbc, else None."""
        Zc, cellc, pbcc, xc = self._species_cache
        if Z.shape != Zc.shape or not np.array_equal(Z, Zc):
            return None
        if self.periodic:
            if cell is None or pbc is None or cellc is None or pbcc is None:
                return None
            if not (np.array_equal(np.asarray(cell, dtype=np.float64), cellc) and
                    np.array_equal(np.asarray(pbc, dtype=bool), pbcc)):
                return None
        x = _Ctx()
        x.pos = pos.to(self.device) if isinstance(pos, torch.Tensor) else self._dev(pos)
        x.n_atoms = pos.shape[0]
        x.centers = x.pos
        for k in ("n_centers", "n_img", "Z_h", "cidx_h", "sp_h", "sp", "shifts", "self_sp"):
            setattr(x, k, getattr(xc, k))
        return x

    # ---- pipeline stages ----------------------------------------------------------
    def _neighbors(self, x: _Ctx, drop_self: bool, cutoff: Optional[float] = None) -> _NL:
        lib = _lib.load()
        st = self._stream()
        cutoff = self._cutoff if cutoff is None else float(cutoff)
        plan = _lib.NLPlan()
        tmp = torch.empty(16384, dtype=torch.uint8, device=self.device)
        offsets = torch.empty(x.n_centers + 1, dtype=torch.int64, device=self.device)
        total, mx = C.c_int64(0), C.c_int32(0)
        key = (x.n_centers, x.n_atoms, x.n_img, cutoff, bool(drop_self))
        sizes = self._nl_sizes.get(key)
        ws = nbr = None
        if sizes is not None:
            # plan, count and fill in one call (no host work between them), with the
            # buffers sized from the last search on this shape
            ws = torch.empty(sizes[0], dtype=torch.uint8, device=self.device)
            nbr = torch.empty(sizes[1], dtype=torch.int32, device=self.device)
            rc = lib.cusoap_nl_build(x.pos.data_ptr(), x.n_atoms, x.shifts.data_ptr(), x.n_img, x.centers.data_ptr(),
                                     x.n_centers, cutoff, int(drop_self), tmp.data_ptr(), ws.data_ptr(), ws.numel(),
                                     offsets.data_ptr(), nbr.data_ptr(), nbr.numel(), C.byref(plan), C.byref(total),
                                     C.byref(mx), st)
            if rc == _lib.OK and int(total.value) <= nbr.numel():
                self._nl_sizes[key] = (sizes[0], max(sizes[1], self._nbr_room(int(total.value))))
                return _NL(offsets, nbr, int(total.value), int(mx.value))
            if rc not in (_lib.OK, _lib.ERR_WORKSPACE):
                _lib.check(rc, "nl_build")
            if rc == _lib.ERR_WORKSPACE:
                ws = None                      # the plan is filled; count with a workspace of its size
        else:
            _lib.check(lib.cusoap_nl_plan(x.centers.data_ptr(), x.n_centers, x.n_atoms, x.n_img, cutoff,
                                          tmp.data_ptr(), C.byref(plan), st), "nl_plan")
        if ws is None:
            ws = torch.empty(int(plan.ws_bytes), dtype=torch.uint8, device=self.device)
            _lib.check(lib.cusoap_nl_count(C.byref(plan), x.pos.data_ptr(), x.n_atoms, x.shifts.data_ptr(),
                                           x.n_img, x.centers.data_ptr(), x.n_centers, cutoff, int(drop_self),
                                           ws.data_ptr(), ws.numel(), offsets.data_ptr(), C.byref(total),
                                           C.byref(mx), st), "nl_count")
        nbr = torch.empty(max(int(total.value), 1), dtype=torch.int32, device=self.device)
        _lib.check(lib.cusoap_nl_fill(C.byref(plan), x.pos.data_ptr(), x.n_atoms, x.shifts.data_ptr(), x.n_img,
                                      x.centers.data_ptr(), x.n_centers, cutoff, int(drop_self),
                                      ws.data_ptr(), ws.numel(), offsets.data_ptr(), mx.value, nbr.data_ptr(), st),
                   "nl_fill")
        self._nl_sizes[key] = (int(plan.ws_bytes) + int(plan.ws_bytes) // 16, self._nbr_room(int(total.value)))
        return _NL(offsets, nbr, int(total.value), int(mx.value))

    @staticmethod
    def _nbr_room(n_edges: int) -> int:
        """Edge capacity for the next search on the same shape: 1/8 more, at least 1024."""
        return n_edges + n_edges // 8 + 1024

    def _coeff_dtype(self, precision: str) -> torch.dtype:
        return torch.float64 if precision == "float64" else torch.float32

    def _coefficients(self, x: _Ctx, nl: _NL, create_path: bool, precision: Optional[str] = None,
                      exact: bool = False) -> torch.Tensor:
        precision = precision or self.precision
        lib = _lib.load()
        b = self._basis_exact() if exact else self._basis()
        L2 = (self._l_max + 1) ** 2
        out = torch.empty((x.n_centers, self.n_species, self._n_max, L2), dtype=self._coeff_dtype(precision),
                          device=self.device)
        _lib.check(lib.cusoap_coefficients(C.byref(b), _PRECISIONS[precision], x.pos.data_ptr(), x.n_atoms,
                                           x.sp.data_ptr(), x.shifts.data_ptr(), x.n_img, x.centers.data_ptr(),
                                           x.n_centers, x.self_sp.data_ptr(), nl.offsets.data_ptr(),
                                           nl.nbr.data_ptr(), int(create_path), out.data_ptr(), self._stream()),
                   "coefficients")
        return out

    def _descriptor(self, coeffs: torch.Tensor, average: Optional[str] = None, dtype: Optional[torch.dtype] = None,
                    exact_spectrum: bool = False) -> torch.Tensor:
        avg = self.average if average is None else average
        dtype = self.dtype if dtype is None else dtype
        lib = _lib.load()
        b = self._basis()
        cdt = _lib.DT_F64 if coeffs.dtype == torch.float64 else _lib.DT_F32
        odt = _lib.DT_F64 if dtype == torch.float64 else _lib.DT_F32
        C_ = coeffs.shape[0]
        if avg == "cc":
            out = torch.empty((C_, self.n_species * self._n_max * (self._l_max + 1) ** 2), dtype=dtype,
                              device=self.device)
            _lib.check(lib.cusoap_cc_projection(C.byref(b), coeffs.data_ptr(), cdt, C_, out.data_ptr(), odt,
                                                self._stream()), "cc_projection")
            return out
        nf = self._n_power_features()
        out = torch.empty((C_, nf), dtype=dtype, device=self.device)
        if self.spectrum != "exact" and not exact_spectrum:
            kind, split = _SPECTRA[self.spectrum]
            _lib.check(lib.cusoap_power_spectrum_mfma(C.byref(b), coeffs.data_ptr(), cdt, C_,
                                                      self._keep["pf"].data_ptr(), nf, kind, split, out.data_ptr(),
                                                      odt, self._stream()), "power_spectrum_mfma")
            return out
        _lib.check(lib.cusoap_power_spectrum(C.byref(b), coeffs.data_ptr(), cdt, C_, self._keep["ft"].data_ptr(),
                                             nf, out.data_ptr(), odt, self._stream()), "power_spectrum")
        return out

    # ---- public forward API (soap_generator.py:1691-1810) -----------------------
    def generate(self, system, centers=None, n_jobs: int = 1, only_physical_cores: bool = False,
                 verbose: bool = False):
        if isinstance(system, list):
            systems = system
            if centers is None:
                centers_list = [None] * len(systems)
            else:
                if not isinstance(centers, list):
                    raise ValueError("For multiple systems, centers must be a list with one entry per system.")
                centers_list = centers
                if len(centers_list) != len(systems):
                    raise ValueError("centers list must match number of systems.")
            outs = [self.generate(s, c) for s, c in zip(systems, centers_list)]
            try:
                return torch.stack(outs, dim=0)
            except Exception:
                return outs
        out = self._evaluate(system, centers)
        return out.to_sparse_coo() if self.sparse else out

    def create(self, system, centers=None, n_jobs: int = 1, only_physical_cores: bool = False,
               verbose: bool = False):
        return self.generate(system, centers, n_jobs=n_jobs, only_physical_cores=only_physical_cores,
                             verbose=verbose)

    def _evaluate(self, system, centers=None, average: Optional[str] = None) -> torch.Tensor:
        x = self._context(system, centers)
        nl = self._neighbors(x, drop_self=True)
        coeffs = self._coefficients(x, nl, create_path=True)
        return self._descriptor(coeffs, average)

    def coefficients(self, system, centers=None, create_path: bool = True) -> torch.Tensor:
        """c[center, species, n, lm] (float64 unless precision='float32')."""
        x = self._context(system, centers)
        nl = self._neighbors(x, drop_self=create_path)
        return self._coefficients(x, nl, create_path=create_path)

    def neighbor_list(self, system, centers=None, drop_self: bool = True):
        """(offsets, ext_indices) of the device neighbour list, as int64 / int32 tensors."""
        x = self._context(system, centers)
        nl = self._neighbors(x, drop_self=drop_self)
        return nl.offsets, nl.nbr[:nl.n_edges]

    # stages of the reference's coefficient step (soap_generator.py:1328-1391, 821-850) that
    # one fused kernel performs here
    _FUSED_STAGES = ("ylm", "weights", "radial", "scatter", "beta", "self_terms")

    @torch.no_grad()
    def create_with_profile(self, system, centers=None):
        """Descriptor, per-stage seconds and run statistics (soap_generator.py:1815-1868).

        The keys are the reference's: to_tensors, centers, neighbor, ylm, weights, radial,
        scatter, beta, self_terms, power_spectrum, total.  to_tensors, centers and neighbor
        (which synchronises on its edge count) are wall times with the stream synchronised
        at each end, as the reference's Profiler measures them; the device stages are the
        HIP-event times of their kernels on the launch stream.  The reference's ylm ...
        self_terms run inside one kernel here: its time is reported under "coefficients",
        and those keys hold 0.0 (``extra["fused_into_coefficients"]`` lists them).
        """
        times: Dict[str, float] = {}
        clock = [0.0]

        def sync():
            torch.cuda.synchronize(self.device)

        def mark(name):
            sync()
            now = time.perf_counter()
            times[name] = times.get(name, 0.0) + (now - clock[0])
            clock[0] = now

        kernels = ("nl_count", "nl_fill", "coefficients", "power_spectrum", "cc_projection")
        _lib.profile_enable(True)
        try:
            for k in kernels:
                _lib.profile_get(k)
            sync()
            t_all = clock[0] = time.perf_counter()
            x = self._context(system, centers, mark=mark)
            nl = self._neighbors(x, drop_self=True)
            mark("neighbor")
            coeffs = self._coefficients(x, nl, create_path=True)
            out = self._descriptor(coeffs)
            sync()
            times["total"] = time.perf_counter() - t_all
            kms = {k: _lib.profile_get(k)[0] for k in kernels}
        finally:
            _lib.profile_enable(False)
        for k in self._FUSED_STAGES:
            times[k] = 0.0
        times["coefficients"] = kms["coefficients"] * 1e-3
        times["power_spectrum"] = (kms["power_spectrum"] + kms["cc_projection"]) * 1e-3
        extra = {"n_atoms": x.n_atoms, "n_centers": x.n_centers, "E_edges": nl.n_edges,
                 "neighbors_per_center": nl.n_edges / x.n_centers if x.n_centers else 0.0,
                 "nl_backend": self._last_nl_backend, "kernel_ms": kms,
                 "fused_into_coefficients": list(self._FUSED_STAGES)}
        return (out.to_sparse_coo() if self.sparse else out), times, extra

    # ---- derivatives (soap_generator.py:1871-2699) ------------------------------------
    @staticmethod
    def _included(n_atoms, include, exclude) -> List[int]:
        idx = list(range(n_atoms)) if include is None else [int(i) for i in include]
        if exclude is not None:
            ex = set(int(i) for i in exclude)
            idx = [i for i in idx if i not in ex]
        return idx

    def _attach_array(self, x: _Ctx, centers, attach: bool) -> Optional[torch.Tensor]:
        if not attach:
            return None
        if centers is None:
            a = np.arange(x.n_centers, dtype=np.int32)
        else:
            a = np.array([int(c) if isinstance(c, int) else -1 for c in centers], dtype=np.int32)
        return self._dev(a)

    def _dense_jacobian(self, system, centers, include, exclude, attach, return_descriptor):
        x = self._context(system, centers)
        lib = _lib.load()
        b = self._basis_exact()
        idx = self._included(x.n_atoms, include, exclude)
        n_inc = len(idx)
        inv = -np.ones(x.n_atoms, dtype=np.int32)
        if n_inc:
            inv[np.asarray(idx)] = np.arange(n_inc, dtype=np.int32)
        closed = (not self.periodic) and self._weighting is None and not attach
        nl_keep = self._neighbors(x, drop_self=False)
        # the dense Jacobian kernels take float64 coefficients
        prec = "mixed" if self.precision == "mixed" else "float64"
        coeffs = self._coefficients(x, nl_keep, create_path=not closed, precision=prec, exact=True)
        if coeffs.dtype != torch.float64:
            coeffs = coeffs.double()
        mode = _lib.MODE_CC if self.average == "cc" else _lib.MODE_PS
        n_feat = self.get_number_of_features()
        deriv = torch.empty((x.n_centers, n_inc, 3, n_feat), dtype=self.dtype, device=self.device)
        nbytes = C.c_size_t(0)
        _lib.check(lib.cusoap_jacobian_workspace_size(C.byref(b), mode, nl_keep.n_edges, x.n_centers,
                                                      n_inc, C.byref(nbytes)),
                   "jacobian_workspace_size")
        ws = torch.empty(int(nbytes.value), dtype=torch.uint8, device=self.device)
        inv_d = self._dev(inv)
        att = self._attach_array(x, centers, attach)
        odt = _lib.DT_F64 if self.dtype == torch.float64 else _lib.DT_F32
        cdt = _lib.DT_F64 if (self.precision == "float64" or odt == _lib.DT_F64) else _lib.DT_F32
        if deriv.numel() > 0:
            _lib.check(lib.cusoap_jacobian_dense(
                C.byref(b), mode, x.pos.data_ptr(), x.n_atoms, x.sp.data_ptr(), x.shifts.data_ptr(), x.n_img,
                x.centers.data_ptr(), x.n_centers, att.data_ptr() if att is not None else None,
                nl_keep.offsets.data_ptr(), nl_keep.nbr.data_ptr(), nl_keep.n_edges, coeffs.data_ptr(),
                inv_d.data_ptr(), n_inc, self._keep["ft"].data_ptr(), n_feat, ws.data_ptr(), ws.numel(),
                deriv.data_ptr(), odt, cdt, self._stream()), "jacobian_dense")
        if return_descriptor:
            return deriv, self._descriptor(coeffs)
        return deriv

    def derivatives_analytical_ps(self, system, centers=None, include=None, exclude=None,
                                  return_descriptor=True):
        """Closed-form Jacobian, fixed centers (soap_generator.py:2270-2626)."""
        if self._rbf not in ("gto", "polynomial"):
            raise NotImplementedError("closed-form ps derivative needs rbf='gto' or 'polynomial'; use derivatives_analytical.")
        if self.periodic:
            raise NotImplementedError("closed-form ps derivative is non-periodic; use derivatives_analytical.")
        if self._weighting is not None:
            raise NotImplementedError("closed-form ps derivative assumes weighting=None; use derivatives_analytical.")
        return self._dense_jacobian(system, centers, include, exclude, False, return_descriptor)

    def derivatives_analytical(self, system, centers=None, include=None, exclude=None,
                               return_descriptor=True, attach=False):
        """Exact Jacobian for every configuration (the reference's autograd path,
        soap_generator.py:2007-2113), here in closed form on the device."""
        return self._dense_jacobian(system, centers, include, exclude, attach, return_descriptor)

    def derivatives_numerical(self, system, centers=None, include=None, exclude=None, method="auto",
                              return_descriptor=True, attach=False, n_jobs=1, only_physical_cores=False,
                              verbose=False, d=1e-4):
        """Central differences with fixed centers, float64 evaluations (soap_generator.py:1871-2003).

        The reference evaluates the whole system twice per (atom, component).  Here every
        displacement of a batch of atoms runs in one launch sequence: each displaced copy
        of an atom is appended to the atom array as an atom of its own, and each center
        that the displacement can change (one within cutoff + 2d of the atom, or attached
        to it) is evaluated once per displacement, with the other copies and the atom's
        undisplaced position filtered out of its neighbour list.  A center the
        displacement cannot reach has equal descriptors at +d and -d, so its row is zero
        as in the reference.  Coefficients, power spectrum and differences are float64
        with the exact radial functions; the result is cast to ``dtype``.
        """
        if method not in ("auto", "numerical"):
            raise ValueError("method must be 'auto' or 'numerical'.")
        desc0 = self.create(system, centers) if return_descriptor else None
        deriv = self._fd_batched(system, centers, include, exclude, attach, float(d))
        deriv = deriv.to(self.dtype)
        return (deriv, desc0) if return_descriptor else deriv

    # displaced evaluations per launch sequence of the batched finite differences
    _FD_MAX_ROWS = 1 << 15

    @torch.no_grad()
    def _fd_batched(self, system, centers, include, exclude, attach, d) -> torch.Tensor:
        pos, Z, cell, pbc = system_to_arrays(system)
        n_atoms = pos.shape[0]
        idx = self._included(n_atoms, include, exclude)
        cxyz, _ = self.prepare_centers(pos, centers)
        n_c = cxyz.shape[0]
        if attach:
            attached = np.arange(n_c, dtype=np.int64) if centers is None else \
                np.array([int(c) if isinstance(c, int) else -1 for c in centers], dtype=np.int64)
        else:
            attached = -np.ones(n_c, dtype=np.int64)
        n_feat = self.get_number_of_features()
        deriv = torch.zeros((n_c, len(idx), 3, n_feat), dtype=torch.float64, device=self.device)
        if not idx or n_atoms == 0:
            return deriv
        # the undisplaced system with the centers as fixed points: which (center, atom)
        # pairs a displacement of d can change
        base = self._context({"positions": pos, "atomic_numbers": Z, "cell": cell, "pbc": pbc},
                             [cxyz[k] for k in range(n_c)])
        nl = self._neighbors(base, drop_self=False, cutoff=self._cutoff + 2.0 * d)
        counts = (nl.offsets[1:] - nl.offsets[:-1]).cpu().numpy()
        nbr_atom = (nl.nbr[:nl.n_edges].to(torch.int64) % n_atoms).cpu().numpy()
        pair_c = np.repeat(np.arange(n_c, dtype=np.int64), counts)
        key = np.unique(pair_c * n_atoms + nbr_atom)
        att = np.nonzero(attached >= 0)[0]
        key = np.union1d(key, att * n_atoms + attached[att])
        inc_pos = -np.ones(n_atoms, dtype=np.int64)
        inc_pos[np.asarray(idx)] = np.arange(len(idx))
        pc, pa = key // n_atoms, key % n_atoms
        keep = inc_pos[pa] >= 0
        pc, pa = pc[keep], pa[keep]
        order = np.argsort(inc_pos[pa], kind="stable")        # pairs grouped by included atom
        pc, pa = pc[order], pa[order]
        # batches of whole atoms, at most _FD_MAX_ROWS displaced centers each
        start = 0
        n_pairs = pc.shape[0]
        while start < n_pairs:
            stop = start
            while stop < n_pairs and (stop == start or 6 * (stop - start) < self._FD_MAX_ROWS):
                a = pa[stop]
                while stop < n_pairs and pa[stop] == a:
                    stop += 1
            self._fd_batch(pos, Z, cell, pbc, cxyz, attached, inc_pos, pc[start:stop], pa[start:stop], d, deriv)
            start = stop
        return deriv

    def _fd_batch(self, pos, Z, cell, pbc, cxyz, attached, inc_pos, pc, pa, d, deriv):
        """(f(+d) - f(-d)) / 2d of the (center pc[k], atom pa[k]) pairs, all 3 components."""
        n_atoms = pos.shape[0]
        atoms = np.unique(pa)
        slot = -np.ones(n_atoms, dtype=np.int64)
        slot[atoms] = np.arange(atoms.shape[0])
        # displaced copy (atom, comp, sign) at index n_atoms + 6 slot + 2 comp + sign
        disp = np.repeat(pos[atoms], 6, axis=0).reshape(-1, 6, 3)
        for comp in range(3):
            disp[:, 2 * comp, comp] = pos[atoms, comp] + d
            disp[:, 2 * comp + 1, comp] = pos[atoms, comp] - d
        disp = disp.reshape(-1, 3)
        n_ext = n_atoms + disp.shape[0]
        ext_pos = np.concatenate([pos, disp], 0)
        ext_Z = np.concatenate([Z, np.repeat(Z[atoms], 6)])
        # one displaced evaluation of center pc[k] per (comp, sign): row 6k + 2 comp + sign
        P = pc.shape[0]
        rows_c = np.repeat(pc, 6)
        rows_a = np.repeat(pa, 6)
        rows_v = n_atoms + 6 * np.repeat(slot[pa], 6) + np.tile(np.arange(6), P)
        moved = attached[rows_c] == rows_a
        rows_xyz = np.where(moved[:, None], ext_pos[rows_v], cxyz[rows_c])
        x = self._fd_context(ext_pos, ext_Z, cell, pbc, rows_xyz)
        # a displaced center snaps to an atom of its own displaced system (within 1e-12),
        # as prepare_centers does, for its self term
        x.self_sp = self._fd_self_species(x, n_atoms, rows_xyz, rows_a, rows_v)
        nl = self._neighbors(x, drop_self=True)
        nl = self._fd_filter(nl, n_ext, n_atoms, rows_a, rows_v)
        coeffs = self._coefficients(x, nl, create_path=True, precision="float64", exact=True)
        f = self._descriptor(coeffs, dtype=torch.float64, exact_spectrum=True).view(P, 3, 2, -1)
        df = (f[:, :, 0, :] - f[:, :, 1, :]) / (2.0 * d)
        ci = torch.from_numpy(pc).to(self.device)
        ai = torch.from_numpy(inc_pos[pa]).to(self.device)
        deriv[ci, ai] = df

    def _fd_context(self, ext_pos, ext_Z, cell, pbc, rows_xyz) -> _Ctx:
        x = _Ctx()
        x.pos = self._dev(ext_pos)
        x.n_atoms = ext_pos.shape[0]
        x.centers = self._dev(np.ascontiguousarray(rows_xyz))
        x.n_centers = rows_xyz.shape[0]
        if self.periodic:
            if cell is None or pbc is None:
                raise ValueError("periodic=True requires cell and pbc.")
            shifts = pbc_image_shifts(cell, pbc, self._cutoff)
        else:
            shifts = np.zeros((1, 3))
        x.shifts = self._dev(shifts)
        x.n_img = shifts.shape[0]
        lut = np.full(max(self._species_Z + [int(ext_Z.max())]) + 1, -1, dtype=np.int32)
        for z, i in self._Z_to_species_index.items():
            lut[z] = i
        x.sp_h = lut[ext_Z]
        x.sp = self._dev(x.sp_h.astype(np.int32))
        return x

    def _fd_self_species(self, x, n_atoms, rows_xyz, rows_a, rows_v) -> torch.Tensor:
        """Species of the atom each displaced center coincides with (argmin, d^2 <= 1e-24), or -1.

        The displaced system of row v is the undisplaced atoms with atom rows_a[v] at the
        position of its copy rows_v[v]."""
        pos = x.pos[:n_atoms]
        q = x.centers
        a = torch.from_numpy(rows_a).to(self.device)
        v = torch.from_numpy(rows_v).to(self.device)
        sp = x.sp[:n_atoms].to(torch.int64)
        out = torch.empty(q.shape[0], dtype=torch.int32, device=self.device)
        step = max(1, (1 << 23) // max(n_atoms, 1))
        for s0 in range(0, q.shape[0], step):
            s1 = min(q.shape[0], s0 + step)
            d2 = ((pos[None, :, :] - q[s0:s1, None, :]) ** 2).sum(-1)
            da = ((x.pos[v[s0:s1]] - q[s0:s1]) ** 2).sum(-1)
            d2.scatter_(1, a[s0:s1, None], da[:, None])
            m, j = d2.min(1)
            out[s0:s1] = torch.where(m <= 1e-24, sp[j], torch.full_like(j, -1)).to(torch.int32)
        return out

    def _fd_filter(self, nl: _NL, n_ext, n_atoms, rows_a, rows_v) -> _NL:
        """Drop from row v's edges every image of atom rows_a[v] and of the copies other than rows_v[v]."""
        n_rows = nl.offsets.shape[0] - 1
        e = nl.nbr[:nl.n_edges].to(torch.int64)
        cnt = nl.offsets[1:] - nl.offsets[:-1]
        row = torch.repeat_interleave(torch.arange(n_rows, device=self.device), cnt)
        atom = e % n_ext
        a = torch.from_numpy(rows_a).to(self.device)[row]
        v = torch.from_numpy(rows_v).to(self.device)[row]
        keep = (atom != a) & ((atom < n_atoms) | (atom == v))
        nbr = nl.nbr[:nl.n_edges][keep].contiguous()
        cnt = torch.bincount(row[keep], minlength=n_rows)
        offsets = torch.zeros(n_rows + 1, dtype=torch.int64, device=self.device)
        offsets[1:] = torch.cumsum(cnt, 0)
        n_edges = int(nbr.shape[0])
        if n_edges == 0:
            nbr = torch.zeros(1, dtype=torch.int32, device=self.device)
        return _NL(offsets, nbr, n_edges, int(cnt.max().item()) if n_rows else 0)

    def derivatives(self, system, centers=None, include=None, exclude=None, method="auto",
                    return_descriptor=True, attach=False, n_jobs=1, only_physical_cores=False, verbose=False):
        """Dispatch as the reference (soap_generator.py:2628-2699)."""
        if method == "numerical":
            return self.derivatives_numerical(system=system, centers=centers, include=include, exclude=exclude,
                                              method="numerical", return_descriptor=return_descriptor,
                                              attach=attach)
        if method in ("auto", "analytical"):
            if (not attach and not self.periodic and self._weighting is None
                    and self.average in ("off", "outer", "cc", "inner")):
                return self.derivatives_analytical_ps(system=system, centers=centers, include=include,
                                                      exclude=exclude, return_descriptor=return_descriptor)
            return self.derivatives_analytical(system=system, centers=centers, include=include, exclude=exclude,
                                               return_descriptor=return_descriptor, attach=attach)
        raise ValueError("method must be one of: 'auto', 'analytical', 'numerical'.")

    # ---- adjoint (no reference counterpart) ------------------------------------------------
    def vjp(self, system, grad_desc, centers=None, attach=False) -> torch.Tensor:
        """sum_{c,f} grad_desc[c,f] d desc[c,f] / d positions, (n_atoms, 3) float64, without dp/dR.

        grad_desc: (n_centers, n_features) or a single (n_features) row shared by
        every center (a linear readout).  Fixed centers unless attach=True.
        """
        x = self._context(system, centers)
        return self._vjp_ctx(x, grad_desc, centers, attach)

    def _vjp_coefficients(self, x: _Ctx, nl_keep: _NL) -> torch.Tensor:
        """Coefficients for the descriptor a derivative differentiates, on the kept-self-edge list.

        Closed-form convention (non-periodic, unweighted): the self edge counts as
        an edge.  Otherwise the create() convention (self edge skipped, self term
        added), as the reference's autograd path differentiates create().
        """
        closed = (not self.periodic) and self._weighting is None
        return self._coefficients(x, nl_keep, create_path=not closed)

    def _edge_reverse(self, x: _Ctx, nl: _NL):
        """(rev, event) for the fixed-order force sums, or None when they do not apply.

        rev[e] is the position of the reverse edge of e (cusoap_edge_reverse).  It depends on
        the neighbour lists only, so it runs on a second stream (default priority) while the
        coefficients are formed; the adjoint waits on the event.
        """
        if not (self.ordered_forces and x.n_centers == x.n_atoms and x.n_centers > 0 and nl.n_edges > 0):
            return None
        lib = _lib.load()
        main = torch.cuda.current_stream(self.device)
        if self._side_stream is None:
            self._side_stream = torch.cuda.Stream(self.device)
        side = self._side_stream
        side.wait_stream(main)
        rev = torch.empty(nl.n_edges, dtype=torch.int32, device=self.device)
        nbytes = C.c_size_t(0)
        _lib.check(lib.cusoap_edge_reverse_workspace_size(x.n_centers, C.byref(nbytes)), "edge_reverse_workspace_size")
        ws = torch.empty(int(nbytes.value), dtype=torch.uint8, device=self.device)
        _lib.check(lib.cusoap_edge_reverse(x.pos.data_ptr(), x.n_atoms, x.shifts.data_ptr(), x.n_img,
                                           x.centers.data_ptr(), x.n_centers, nl.offsets.data_ptr(),
                                           nl.nbr.data_ptr(), ws.data_ptr(), ws.numel(), rev.data_ptr(),
                                           side.cuda_stream), "edge_reverse")
        for t in (rev, ws, x.pos, x.shifts, x.centers, nl.offsets, nl.nbr):
            t.record_stream(side)
        ev = torch.cuda.Event()
        ev.record(side)
        return rev, ev

    def _vjp_ctx(self, x: _Ctx, grad_desc, centers, attach, coeffs=None, nl_keep=None, ordered=True, rev=None):
        lib = _lib.load()
        b = self._basis()
        n_feat = self.get_number_of_features()
        g = grad_desc if isinstance(grad_desc, torch.Tensor) else torch.as_tensor(np.asarray(grad_desc))
        g = g.to(self.device)
        if g.dtype not in (torch.float32, torch.float64):
            g = g.to(torch.float64)
        g = g.contiguous()
        if g.dim() == 1:
            if g.numel() != n_feat:
                raise ValueError("grad_desc row must have n_features entries")
            bcast = 1
        else:
            if tuple(g.shape) != (x.n_centers, n_feat):
                raise ValueError(f"grad_desc must be ({x.n_centers}, {n_feat})")
            bcast = 0
        if nl_keep is None:
            nl_keep = self._neighbors(x, drop_self=False)
        if rev is None and ordered:
            rev = self._edge_reverse(x, nl_keep)
        if coeffs is None:
            coeffs = self._vjp_coefficients(x, nl_keep)
        grad = torch.empty((x.n_atoms, 3), dtype=torch.float64, device=self.device)
        att = self._attach_array(x, centers, attach)
        mode = _lib.MODE_CC if self.average == "cc" else _lib.MODE_PS
        prec = _PRECISIONS[self.precision]
        nbytes = C.c_size_t(0)
        head = (C.byref(b), mode, prec, x.pos.data_ptr(), x.n_atoms, x.sp.data_ptr(),
                x.shifts.data_ptr(), x.n_img, x.centers.data_ptr(), x.n_centers,
                att.data_ptr() if att is not None else None, x.self_sp.data_ptr(), nl_keep.offsets.data_ptr(),
                nl_keep.nbr.data_ptr())
        tail = (coeffs.data_ptr(), _lib.DT_F64 if coeffs.dtype == torch.float64 else _lib.DT_F32,
                self._keep["pf"].data_ptr(), g.data_ptr(), _lib.DT_F64 if g.dtype == torch.float64 else _lib.DT_F32,
                bcast, n_feat)
        if ordered and rev is not None:
            rev_t, ev = rev
            torch.cuda.current_stream(self.device).wait_event(ev)
            _lib.check(lib.cusoap_vjp_ordered_workspace_size(C.byref(b), prec, x.n_centers, nl_keep.n_edges,
                                                             C.byref(nbytes)), "vjp_ordered_workspace_size")
            ws = torch.empty(int(nbytes.value), dtype=torch.uint8, device=self.device)
            _lib.check(lib.cusoap_vjp_ordered(*head, rev_t.data_ptr(), nl_keep.n_edges, *tail, ws.data_ptr(),
                                              ws.numel(), grad.data_ptr(), self._stream()), "vjp_ordered")
        else:
            _lib.check(lib.cusoap_vjp_workspace_size(C.byref(b), x.n_centers, C.byref(nbytes)), "vjp_workspace_size")
            ws = torch.empty(int(nbytes.value), dtype=torch.uint8, device=self.device)
            _lib.check(lib.cusoap_vjp(*head, *tail, ws.data_ptr(), ws.numel(), grad.data_ptr(), self._stream()),
                       "vjp")
        return grad

    def descriptor_and_vjp(self, system, grad_desc, centers=None, attach=False):
        """Descriptor and its adjoint from one neighbour search and one coefficient pass."""
        x = self._context(system, centers)
        nl_keep = self._neighbors(x, drop_self=False)
        rev = self._edge_reverse(x, nl_keep)
        coeffs = self._vjp_coefficients(x, nl_keep)
        desc = self._descriptor(coeffs)
        return desc, self._vjp_ctx(x, grad_desc, centers, attach, coeffs=coeffs, nl_keep=nl_keep, rev=rev)

    def create_autograd(self, positions: torch.Tensor, atomic_numbers, cell=None, pbc=None, centers=None,
                        attach=False) -> torch.Tensor:
        """Descriptor as a differentiable function of ``positions`` (backward = the adjoint kernel)."""
        return _SOAPFunction.apply(positions, self, atomic_numbers, cell, pbc, centers, attach)


class _SOAPFunction(torch.autograd.Function):
    """desc = SOAP(positions); backward through cusoap_vjp (no dp/dR is formed)."""

    @staticmethod
    def forward(ctx, positions, soap, Z, cell, pbc, centers, attach):
        system = {"positions": positions.detach(), "atomic_numbers": Z, "cell": cell, "pbc": pbc}
        desc = soap.create(system, centers)
        ctx.soap, ctx.system, ctx.centers, ctx.attach = soap, system, centers, attach
        ctx.pos_dtype, ctx.pos_device = positions.dtype, positions.device
        return desc

    @staticmethod
    def backward(ctx, grad_out):
        g = ctx.soap.vjp(ctx.system, grad_out.contiguous(), ctx.centers, ctx.attach)
        return g.to(device=ctx.pos_device, dtype=ctx.pos_dtype), None, None, None, None, None, None
