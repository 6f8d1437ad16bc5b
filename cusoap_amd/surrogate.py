"""Learned radial model: per-species-pair networks for the orthonormal GTO radial
functions, trained on the device, gated by their held-out error, tabulated.

No reference counterpart.  The reference evaluates the exact radial integral
for every edge (soap_generator.py:1358-1391 for create(), 2389-2405 for the
derivative).  Here one network per (center species, neighbour species) pair,
x = 2 r / r_fit - 1 -> 64 tanh -> 64 tanh -> n_max (l_max+1), is fitted to the
exact float64 radial projections g_nl(r) = (r / r_fit)^l B_nl(r) and their slopes
on [0, r_fit] by the library's own kernels.  g_nl is what the reference multiplies
with the unit-vector harmonics (up to the constant r_fit^l); B_nl multiplies the
solid harmonics r^l Y_lm in these kernels.  An error measured on B_nl would
overweight small r, where r^l removes it.  The kernels used are:
  - exact samples (cusoap_radial_exact);
  - minibatch Adam on values plus slopes (cusoap_mlp_loss_grad, cusoap_adam_step);
  - a final exact least-squares refit of the output layer (cusoap_mlp_refit_output).
A pair whose held-out maximum error, relative to each function's maximum,
exceeds the bound keeps the exact function.  So do radii beyond r_fit and
centers that are not atoms.

Why a table.  The radial functions depend on r only.  The kernels read them
from a cubic-Hermite table of the model's values and slopes (TABLE_INTERVALS =
256 intervals on [0, cutoff]; cusoap_radial_table).  This costs one network
evaluation per node instead of per edge.  The per-edge work is then float32
interpolation of already-orthonormal functions.  There is no Loewdin cancellation
and no float64 exponential per edge.

Why 256 intervals.  The float32 node values carry rounding noise (the network's
float32 evaluation, amplified by 1/u^l in B = g / u^l), and the interpolant's slope
between two nodes takes that noise divided by the interval width h.  Finer tables
therefore give noisier forces.  On the 98,304-atom water box (n 8, l 6, r_cut 6)
the learned forces were 1.9e-4 of max|F| from the float64 path at 1024 intervals
(1,016 atoms above 1e-4), 4.9e-5 at 256 and 5.3e-5 at 128. The exact table gave
5.2e-6 at 1024, 2.3e-6 at 512 and 5.9e-6 at 256.  The smaller table (4x fewer rows
for the L2 caches) also made the edge pass ~4% faster.  In float64 (precision=
"float64") there is no such noise, and SOAP keeps 1024 intervals there: at 256 the
interpolation error alone reaches 1.6e-6 of the descriptors' scale.
"""
from __future__ import annotations

import ctypes as C
import math
import time
from typing import Optional

import numpy as np
import torch

from . import _lib

TABLE_INTERVALS = 256        # default intervals of the learned model's table (the module docstring says why)


def default_table_intervals(radial: str, precision: str) -> int:
    """Intervals of the radial table when the caller gives none: 256 for the learned model
    (its node noise dominates, see the module docstring), 512 for the exact functions in
    the float precisions (their float32 node rounding against the interpolation error:
    forces 2.3e-6 of max|F| on the water box at 512, 5.9e-6 at 256 and 5.2e-6 at 1024;
    the H/C/N/O basis at 256 reached 1.5e-5 on descriptors), 1024 in float64."""
    if precision == "float64":
        return 1024
    return TABLE_INTERVALS if radial == "learned" else 512


HIDDEN = 64


def param_count(n_out: int) -> int:
    return 3 * HIDDEN + HIDDEN * HIDDEN + n_out * HIDDEN + n_out


def init_params(n_pairs: int, n_out: int, seed: int) -> np.ndarray:
    """First layer spread over the input interval; Xavier-scaled hidden/output layers."""
    rng = np.random.default_rng(seed)
    out = []
    H = HIDDEN
    for _ in range(n_pairs):
        out.append(np.concatenate([
            rng.uniform(-4.0, 4.0, H), rng.uniform(-4.0, 4.0, H),
            rng.normal(0.0, 1.0 / math.sqrt(H), H * H), np.zeros(H),
            rng.normal(0.0, 1.0 / math.sqrt(H), n_out * H), np.zeros(n_out)]))
    return np.stack(out).astype(np.float32)


class RadialSurrogate:
    """Networks, training state and gate of one SOAP basis.

    soap: a cusoap_amd.SOAP with rbf="gto" on a HIP device.
    r_fit: upper end of the training interval (default r_cut: samples r in [0, r_cut]).
    tol, tol_slope: gate bounds on the held-out max error of g_nl and of its slope,
         each relative to that function's maximum over [0, r_fit].
    """

    def __init__(self, soap, r_fit: Optional[float] = None, n_train: int = 65536, n_heldout: int = 16384,
                 seed: int = 0, deriv_weight: float = 1.0, tol: float = 2e-5, tol_slope: float = 1e-4):
        if soap._rbf != "gto":
            raise NotImplementedError("the learned radial model covers the GTO basis")
        self.soap = soap
        self.device = soap.device
        self.S = soap.n_species
        self.n_pairs = self.S * self.S
        self.n, self.L1 = soap._n_max, soap._l_max + 1
        self.O = self.n * self.L1
        if self.O > 128:
            raise NotImplementedError("learned radial model: n_max (l_max+1) must be <= 128")
        self.r_fit = float(r_fit if r_fit is not None else soap._r_cut)
        self.n_train, self.n_heldout, self.seed = int(n_train), int(n_heldout), int(seed)
        self.deriv_weight, self.tol, self.tol_slope = float(deriv_weight), float(tol), float(tol_slope)
        self.P = param_count(self.O)
        lib = _lib.load()
        if lib.cusoap_mlp_param_count(HIDDEN, self.O) != self.P:
            raise RuntimeError("network layout mismatch between library and host")
        self.params = torch.from_numpy(init_params(self.n_pairs, self.O, self.seed)).to(self.device)
        self.m = torch.zeros_like(self.params)
        self.v = torch.zeros_like(self.params)
        self.step = 0
        self.history = []
        self.train_seconds = 0.0
        self._make_data()
        self.errors = None
        self.use_pair = np.zeros(self.n_pairs, dtype=np.int32)

    # ---- data -------------------------------------------------------------------
    def exact(self, r: np.ndarray):
        return exact_radial(self.soap, r)

    def projections(self, r: np.ndarray):
        """Exact g = u^l B and dg/dr = u^l dB/dr + l u^(l-1) B / r_fit (P, O), u = r / r_fit."""
        B, dB = self.exact(r)
        u = torch.as_tensor(np.asarray(r, dtype=np.float64) / self.r_fit, device=self.device)[:, None]
        lv = torch.arange(self.L1, device=self.device, dtype=torch.float64).repeat_interleave(self.n)[None]
        ul = u ** lv
        ulm1 = torch.where(lv > 0, u ** (lv - 1).clamp_min(0), torch.zeros_like(ul))
        return ul * B, ul * dB + lv * ulm1 * B / self.r_fit

    def _make_data(self):
        rng = np.random.default_rng(self.seed + 1)
        r = rng.uniform(0.0, self.r_fit, self.n_train)
        B, dB = self.projections(r)
        self.scale = B.abs().amax(0).clamp_min(1e-300)                    # (O,) float64
        half = self.r_fit / 2.0
        self.x = torch.as_tensor(2.0 * r / self.r_fit - 1.0, device=self.device).float().contiguous()
        self.t = (B / self.scale).float().contiguous()
        dt = dB * half / self.scale
        self.dscale = dt.abs().amax(0).clamp_min(1e-300)
        self.dt = dt.float().contiguous()
        self.w_val = torch.ones(self.O, dtype=torch.float32, device=self.device)
        self.w_der = (self.deriv_weight / (self.dscale * self.dscale)).float().contiguous()
        # held-out: fresh uniform samples plus a fine uniform grid (both ends included)
        rng2 = np.random.default_rng(self.seed + 2)
        r_ho = np.concatenate([rng2.uniform(0.0, self.r_fit, self.n_heldout), np.linspace(0.0, self.r_fit, 8193)])
        self.r_ho = r_ho
        self.B_ho, self.dB_ho = self.projections(r_ho)
        self.x_ho = torch.as_tensor(2.0 * r_ho / self.r_fit - 1.0, device=self.device).float().contiguous()

    # ---- training ---------------------------------------------------------------
    def train(self, steps: int = 60000, batch: int = 4096, lr: float = 3e-3, lr_min: float = 1e-5,
              refit: bool = True, refit_every: int = 0, log_every: int = 0, seed: int = 12345,
              eps: float = 1e-12):
        """Adam with a cosine learning-rate schedule, then the output-layer refit
        (also every ``refit_every`` steps when set: Adam then only has to move the
        hidden layers towards features the exact output layer fits best)."""
        lib = _lib.load()
        st = self.soap._stream()
        nbytes = C.c_size_t(0)
        _lib.check(lib.cusoap_mlp_train_workspace_size(self.n_pairs, HIDDEN, self.O, batch, C.byref(nbytes)),
                   "mlp_train_workspace_size")
        ws = torch.empty(int(nbytes.value), dtype=torch.uint8, device=self.device)
        grad = torch.empty_like(self.params)
        loss = torch.empty(self.n_pairs, dtype=torch.float32, device=self.device)
        torch.cuda.synchronize(self.device)
        t0 = time.perf_counter()
        for k in range(steps):
            self.step += 1
            lr_k = lr_min + 0.5 * (lr - lr_min) * (1.0 + math.cos(math.pi * k / max(steps, 1)))
            _lib.check(lib.cusoap_mlp_loss_grad(
                self.params.data_ptr(), self.n_pairs, HIDDEN, self.O, self.x.data_ptr(), self.t.data_ptr(),
                self.dt.data_ptr(), self.n_train, 0, self.w_val.data_ptr(), self.w_der.data_ptr(), batch,
                seed, self.step, ws.data_ptr(), ws.numel(), grad.data_ptr(), loss.data_ptr(), st), "mlp_loss_grad")
            _lib.check(lib.cusoap_adam_step(self.params.data_ptr(), grad.data_ptr(), self.m.data_ptr(),
                                            self.v.data_ptr(), self.params.numel(), lr_k, 0.9, 0.999, eps,
                                            self.step, st), "adam_step")
            if refit_every and (k + 1) % refit_every == 0 and k + 1 < steps:
                self.refit()
            if log_every and (k % log_every == 0 or k == steps - 1):
                self.history.append((self.step, loss.cpu().numpy().tolist()))
        if refit:
            self.refit()
        torch.cuda.synchronize(self.device)
        self.train_seconds += time.perf_counter() - t0
        return self

    def refit(self, ridge: float = 1e-12):
        lib = _lib.load()
        nbytes = C.c_size_t(0)
        _lib.check(lib.cusoap_mlp_refit_workspace_size(self.n_pairs, HIDDEN, self.O, self.n_train,
                                                       C.byref(nbytes)), "mlp_refit_workspace_size")
        ws = torch.empty(int(nbytes.value), dtype=torch.uint8, device=self.device)
        _lib.check(lib.cusoap_mlp_refit_output(
            self.params.data_ptr(), self.n_pairs, HIDDEN, self.O, self.x.data_ptr(), self.t.data_ptr(),
            self.dt.data_ptr(), self.n_train, 0, self.w_der.data_ptr(), ridge, ws.data_ptr(), ws.numel(),
            self.soap._stream()), "mlp_refit_output")

    def loss(self, step: Optional[int] = None, batch: int = 4096, seed: int = 12345) -> np.ndarray:
        """Minibatch loss per pair at the current parameters (no update)."""
        lib = _lib.load()
        nbytes = C.c_size_t(0)
        _lib.check(lib.cusoap_mlp_train_workspace_size(self.n_pairs, HIDDEN, self.O, batch, C.byref(nbytes)),
                   "mlp_train_workspace_size")
        ws = torch.empty(int(nbytes.value), dtype=torch.uint8, device=self.device)
        grad = torch.empty_like(self.params)
        loss = torch.empty(self.n_pairs, dtype=torch.float32, device=self.device)
        _lib.check(lib.cusoap_mlp_loss_grad(
            self.params.data_ptr(), self.n_pairs, HIDDEN, self.O, self.x.data_ptr(), self.t.data_ptr(),
            self.dt.data_ptr(), self.n_train, 0, self.w_val.data_ptr(), self.w_der.data_ptr(), batch, seed,
            self.step if step is None else step, ws.data_ptr(), ws.numel(), grad.data_ptr(), loss.data_ptr(),
            self.soap._stream()), "mlp_loss_grad")
        return loss.cpu().numpy()

    # ---- inference and gate -----------------------------------------------------
    def evaluate(self, r: np.ndarray):
        """Network g (pairs, P, O) and dg/dr (pairs, P, O), float32 -> float64, on the device."""
        lib = _lib.load()
        x = torch.as_tensor(2.0 * np.asarray(r, dtype=np.float64) / self.r_fit - 1.0,
                            device=self.device).float().contiguous()
        y = torch.empty((self.n_pairs, x.numel(), self.O), dtype=torch.float32, device=self.device)
        dy = torch.empty_like(y)
        _lib.check(lib.cusoap_mlp_eval(self.params.data_ptr(), self.n_pairs, HIDDEN, self.O, x.data_ptr(),
                                       x.numel(), y.data_ptr(), dy.data_ptr(), self.soap._stream()), "mlp_eval")
        B = y.double() * self.scale
        dB = dy.double() * self.scale * (2.0 / self.r_fit)
        return B, dB

    def heldout_errors(self):
        """Per pair: max over functions of max|model - exact| / max|exact| of g_nl on the
        held-out radii, for values and for slopes."""
        B, dB = self.evaluate(self.r_ho)
        vmax = self.B_ho.abs().amax(0)
        dmax = self.dB_ho.abs().amax(0)
        ev = ((B - self.B_ho[None]).abs().amax(1) / vmax).amax(1)
        ed = ((dB - self.dB_ho[None]).abs().amax(1) / dmax).amax(1)
        self.errors = {"value": ev.cpu().numpy(), "slope": ed.cpu().numpy()}
        return self.errors

    def gate(self, soap=None, n_intervals: int = TABLE_INTERVALS) -> np.ndarray:
        """Channels (species pair, l, n) whose network the radial table of ``soap`` may use.

        Returns a boolean mask (S*S, l_max+1, n_max).  Two stages, the exact function
        judging both:
          1. fit: the channels of a pair pass when the pair's held-out errors are within
             tol (values) and tol_slope (slopes), relative to each function's maximum;
          2. end to end: on a small calibration system of the basis' species (a periodic
             water box for H/O, else a random box; ``gate_options["calibration"]``
             replaces it) the descriptors and the forces of a seeded Gaussian readout
             from the learned table are compared with those from the exact table (same
             kernels, float64 arithmetic).  Each channel's own contribution to those errors
             is measured once (table with that channel alone on its network); the errors
             of a set of channels are, to first order, the sums of these contributions.
             While the predicted larger error exceeds margin * bar (default 0.35 * 1e-4:
             the largest errors of other systems of the same kind were measured at up to
             ~2.4x those of the calibration system),
             the channel whose return to the exact function lowers it most goes back;
             the chosen set is then evaluated for real and trimmed further if needed.
        """
        soap = self.soap if soap is None else soap
        opts = getattr(self, "gate_options", {}) or {}
        bar, margin = float(opts.get("bar", 1e-4)), float(opts.get("margin", 0.35))
        bound = margin * bar
        e = self.heldout_errors()
        fit = (e["value"] <= self.tol) & (e["slope"] <= self.tol_slope)
        mask = np.repeat(fit[:, None, None], self.L1 * self.n).reshape(self.n_pairs, self.L1, self.n)
        rep = {"radial_heldout_value_err": float(e["value"].max()),
               "radial_heldout_slope_err": float(e["slope"].max()),
               "radial_fit_gate": [int(u) for u in fit], "radial_gate_bound": bound}
        if mask.any():
            calib = opts.get("calibration") or calibration_system(soap)
            # float64 arithmetic: the deviations are the network's alone, and repeat bit for
            # bit closely enough that the selection does not depend on float32 rounding
            ref = soap._sibling(radial="table", table_intervals=n_intervals, dtype="float64",
                                precision="float64", average=soap.average)
            exact_tab = ref._rad_table
            net_tab = radial_table(soap, n_intervals, model=self, use_pair=fit.astype(np.int32))
            w = torch.as_tensor(np.random.default_rng(777).normal(size=ref.get_number_of_features()),
                                device=soap.device)
            d_ref, F_ref = ref.descriptor_and_vjp(calib, w)
            d_ref = d_ref.double()
            d_sc = d_ref.abs().amax(0).clamp_min(1e-300)
            F_sc = F_ref.abs().max().clamp_min(1e-300)

            def deviation(m):
                """Normalised descriptor and force deviations of the table with channels m."""
                ref._rad_table = masked_table(exact_tab, net_tab, m)
                ref._basis_learned = None
                d, F = ref.descriptor_and_vjp(calib, w)
                return (d.double() - d_ref) / d_sc, ((F - F_ref) / F_sc).reshape(-1)

            def norms(dv, fv):
                return float(dv.abs().max()), float(fv.abs().max())

            err = norms(*deviation(mask))
            rep["radial_calibration_err_all_fit_channels"] = {"desc": err[0], "force": err[1]}
            chans = [tuple(c) for c in np.argwhere(mask)]
            if max(err) > bound:
                # one deviation per channel; the set's deviation is their sum (first order);
                # descriptor rows of at most 256 centers enter the model, all rows the checks
                rows = min(256, int(d_ref.shape[0]))
                Ed, Ef = [], []
                for c in chans:
                    one = np.zeros_like(mask)
                    one[c] = True
                    dv, fv = deviation(one)
                    Ed.append(dv[:rows].reshape(-1).float())     # the first rows stand for all
                    Ef.append(fv.float())
                Ed, Ef = torch.stack(Ed), torch.stack(Ef)
                keep = torch.ones(len(chans), dtype=torch.bool, device=soap.device)
                Sd, Sf = Ed.sum(0), Ef.sum(0)
                while bool(keep.any()):
                    cur = max(float(Sd.abs().max()), float(Sf.abs().max()))
                    if cur <= bound:
                        m = np.zeros_like(mask)
                        for k in torch.nonzero(keep).flatten().tolist():
                            m[chans[k]] = True
                        err = norms(*deviation(m))
                        if max(err) <= bound:
                            break
                    # the kept channel whose removal leaves the smallest larger error
                    cand = torch.nonzero(keep).flatten()
                    after = torch.maximum((Sd[None] - Ed[cand]).abs().amax(1), (Sf[None] - Ef[cand]).abs().amax(1))
                    k = int(cand[int(torch.argmin(after))])
                    keep[k] = False
                    Sd, Sf = Sd - Ed[k], Sf - Ef[k]
                mask = np.zeros_like(mask)
                for k in torch.nonzero(keep).flatten().tolist():
                    mask[chans[k]] = True
                if not mask.any():
                    err = (0.0, 0.0)
                del Ed, Ef
            rep["radial_calibration_err"] = {"desc": err[0], "force": err[1]}
            rep["radial_calibration_atoms"] = int(len(calib["positions"]))
            ref._rad_table, ref._basis_learned = exact_tab, None
        rep["radial_channels_learned"] = int(mask.sum())
        rep["radial_channels_total"] = int(mask.size)
        self.channels = mask
        self.use_pair = mask.reshape(self.n_pairs, -1).any(1).astype(np.int32)
        self.report = rep
        return mask.copy()

    def gate_report(self) -> dict:
        """What the last gate measured: held-out errors, the fit gate, calibration errors."""
        return dict(getattr(self, "report", {}))

    # ---- tabulation -------------------------------------------------------------
    def table(self, n_intervals: int = TABLE_INTERVALS, r_max: Optional[float] = None, exact: bool = False) -> torch.Tensor:
        """Radial table for basis.rad_table: network nodes where gated in, exact elsewhere."""
        return radial_table(self.soap, n_intervals, r_max, None if exact else self)

    # ---- checkpoint ---------------------------------------------------------------
    def _meta(self):
        s = self.soap
        return np.array([s._r_cut, s._sigma, s._n_max, s._l_max, self.S, self.r_fit, HIDDEN, self.seed,
                         self.n_train], dtype=np.float64)

    def save(self, path: str):
        np.savez(path, params=self.params.cpu().numpy(), meta=self._meta(), species=np.array(self.soap._species_Z),
                 scale=self.scale.cpu().numpy(), dscale=self.dscale.cpu().numpy(), step=np.array(self.step),
                 use_pair=self.use_pair, betas=np.asarray(self.soap._betas))

    def load(self, path: str):
        d = np.load(path, allow_pickle=False)
        if not np.array_equal(d["meta"], self._meta()) or list(d["species"]) != list(self.soap._species_Z):
            raise ValueError("checkpoint was trained for a different basis")
        if np.abs(d["betas"] - np.asarray(self.soap._betas)).max() > 1e-9 * np.abs(d["betas"]).max():
            raise ValueError("checkpoint was trained for a different orthonormalisation")
        # the output scales are recomputed from the same seeded samples in __init__; a
        # checkpoint whose saved scales differ would rescale the network outputs
        for key, mine in (("scale", self.scale), ("dscale", self.dscale)):
            saved = torch.as_tensor(d[key], device=self.device, dtype=torch.float64)
            if saved.shape != mine.shape or (saved - mine).abs().max() > 1e-12 * mine.abs().max():
                raise ValueError(f"checkpoint {key} does not match this basis and training set")
            mine.copy_(saved)
        self.params = torch.from_numpy(d["params"].astype(np.float32)).to(self.device).contiguous()
        self.step = int(d["step"])
        return self


def masked_table(exact: torch.Tensor, net: torch.Tensor, channels: np.ndarray) -> torch.Tensor:
    """Radial table with the network's nodes on the given (pair, l, n) channels and the
    exact nodes everywhere else (slots of coordinate centers stay exact)."""
    n_pairs = channels.shape[0]
    full = torch.zeros(exact.shape[0], *channels.shape[1:], dtype=torch.bool, device=exact.device)
    full[:n_pairs] = torch.as_tensor(channels, device=exact.device)
    return torch.where(full[:, None, :, :, None], net, exact).contiguous()


def calibration_system(soap, seed: int = 0) -> dict:
    """Small system of the basis' species for the end-to-end gate: a periodic 8^3 water box
    (1,536 atoms, 1.0 g/cm^3) for H/O, else a random periodic box of 1,536 atoms at
    0.10 A^-3 (H/C/N/O in the fractions 0.50/0.25/0.10/0.15, other species in equal parts).
    Without periodic=True the same atoms are taken as a cluster."""
    from . import systems
    Z = list(soap._species_Z)
    if Z == [1, 8]:
        s = systems.water_box(8, seed=seed)
    else:
        fr = (0.50, 0.25, 0.10, 0.15) if Z == [1, 6, 7, 8] else tuple([1.0 / len(Z)] * len(Z))
        s = systems.random_box(n_atoms=1536, fractions=fr, Z=Z, density=0.10, seed=seed)
    if not soap.periodic:
        s = {"positions": s["positions"], "atomic_numbers": s["atomic_numbers"]}
    return s


def exact_radial(soap, r: np.ndarray):
    """Exact float64 B (P, O) and dB/dr (P, O) of a GTO SOAP at radii r, outputs l-major, on the device."""
    lib = _lib.load()
    O = soap._n_max * (soap._l_max + 1)
    rd = torch.as_tensor(np.ascontiguousarray(r, dtype=np.float64), device=soap.device)
    B = torch.empty((rd.numel(), O), dtype=torch.float64, device=soap.device)
    dB = torch.empty_like(B)
    _lib.check(lib.cusoap_radial_exact(C.byref(soap._basis_exact()), rd.data_ptr(), rd.numel(), B.data_ptr(),
                                       dB.data_ptr(), soap._stream()), "radial_exact")
    return B, dB


def radial_table(soap, n_intervals: int = TABLE_INTERVALS, r_max: Optional[float] = None,
                 model: Optional[RadialSurrogate] = None, use_pair: Optional[np.ndarray] = None) -> torch.Tensor:
    """(S+1) S x (n_intervals+1) x (l_max+1) x n_max x 2 float32 table on the device.

    The last axis holds (B_nl(r_i), dB_nl/dr(r_i)) at node r_i.  With a model, the nodes
    of the pairs in ``use_pair`` (default: the model's last gate) come from the networks."""
    lib = _lib.load()
    b = soap._basis_exact()
    r_max = float(soap._cutoff if r_max is None else r_max)
    S, L1, n = soap.n_species, soap._l_max + 1, soap._n_max
    table = torch.empty(((S + 1) * S, n_intervals + 1, L1, n, 2), dtype=torch.float32, device=soap.device)
    if model is None:
        _lib.check(lib.cusoap_radial_table(C.byref(b), None, HIDDEN, None, None, 1.0, n_intervals, r_max, None, 0,
                                           table.data_ptr(), soap._stream()), "radial_table")
        return table
    nbytes = C.c_size_t(0)
    _lib.check(lib.cusoap_radial_table_workspace_size(C.byref(b), n_intervals, C.byref(nbytes)),
               "radial_table_workspace_size")
    ws = torch.empty(int(nbytes.value), dtype=torch.uint8, device=soap.device)
    use = torch.as_tensor(np.asarray(model.use_pair if use_pair is None else use_pair, dtype=np.int32),
                          device=soap.device)
    scale = model.scale.float().expand(model.n_pairs, model.O).contiguous()
    _lib.check(lib.cusoap_radial_table(C.byref(b), model.params.data_ptr(), HIDDEN, use.data_ptr(),
                                       scale.data_ptr(), model.r_fit, n_intervals, r_max, ws.data_ptr(), ws.numel(),
                                       table.data_ptr(), soap._stream()), "radial_table")
    return table
