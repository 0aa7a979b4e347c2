"""Natural cubic splines for the background / thermal-history tables.

Host-side table producer support (not the GPU hot path).  Semantics follow
``src/discoeb/spline_interpolation.py`` of DISCO-EB:

* construction (``spline_interpolation.__init__``, lines 8-113): NaN fill-forward
  of x (factor 1.01) and y (factor 1.0), second derivatives ``S`` from a Thomas
  solve with natural end conditions ``S[0] = S[-1] = 0``;
* evaluation (``evaluate``, lines 130-153): interval index
  ``clip(searchsorted(x, x_new, 'left') - 1, 0, n - 2)`` and the standard
  ``A y_i + B y_{i+1} + ((A^3-A) S_i + (B^3-B) S_{i+1}) h^2 / 6`` form, which
  extrapolates with the end pieces outside the knot range;
* ``derivative`` (lines 191-212) and ``integral`` (lines 155-189).

Everything is vectorised over a leading batch axis and works on complex
arrays, so the whole table producer can be differentiated by the complex-step
method.  Interval search always uses the real part (a branch decision, like a
``searchsorted`` under JAX forward-mode differentiation).
"""
from __future__ import annotations

import numpy as np

__all__ = ["Spline", "interval_index"]


def _fill_forward(v: np.ndarray, fac: float) -> np.ndarray:
    """Replace NaNs by ``fac`` times the last observed value (scan from the start)."""
    if not np.any(np.isnan(v.real)):
        return v
    out = v.copy()
    last = out[..., 0].copy()
    for i in range(out.shape[-1]):
        nan = np.isnan(out[..., i].real)
        out[..., i] = np.where(nan, last * fac, out[..., i])
        last = out[..., i]
    return out


def interval_index(x: np.ndarray, xn: np.ndarray, shared: bool = False) -> np.ndarray:
    """``clip(searchsorted(x, xn, side='left') - 1, 0, n - 2)`` row-wise.

    ``x`` has shape (B, n); ``xn`` has shape (B,) or (B, m).  ``shared`` means
    every row of ``x`` holds the same knots.
    """
    xr = np.asarray(x).real
    xnr = np.asarray(xn).real
    n = xr.shape[-1]
    if shared:
        cnt = np.searchsorted(xr[0], xnr, side="left")
    else:
        cnt = np.empty(xnr.shape, dtype=np.int64)
        for b in range(xr.shape[0]):
            cnt[b] = np.searchsorted(xr[b], xnr[b], side="left")
    return np.clip(cnt - 1, 0, n - 2)


def _take(a: np.ndarray, idx: np.ndarray) -> np.ndarray:
    if idx.ndim == 1:
        return np.take_along_axis(a, idx[:, None], axis=-1)[:, 0]
    return np.take_along_axis(a, idx, axis=-1)


class Spline:
    """Batched natural cubic spline ``y(x)`` with knots along the last axis."""

    def __init__(self, x, y, integrate_from_start: bool = True):
        x = np.asarray(x)
        y = np.asarray(y)
        self.shared_x = x.ndim == 1 or x.shape[0] == 1
        if x.ndim == 1:
            x = x[None, :]
        if y.ndim == 1:
            y = y[None, :]
        B = max(x.shape[0], y.shape[0])
        x = np.broadcast_to(x, (B, x.shape[-1]))
        y = np.broadcast_to(y, (B, y.shape[-1]))
        n = y.shape[-1]
        if n < 2:
            raise ValueError("There must be at least two data points.")
        y = _fill_forward(np.array(y), 1.0)
        x = _fill_forward(np.array(x), 1.01)
        dtype = np.result_type(x.dtype, y.dtype, np.float64)
        x = x.astype(dtype)
        y = y.astype(dtype)
        h = x[:, 1:] - x[:, :-1]
        m = n - 2
        S = np.zeros((B, n), dtype=dtype)
        if m > 0:
            a = h[:, :-1]
            bd = 2.0 * (h[:, :-1] + h[:, 1:])
            c = h[:, 1:]
            d = 6.0 * ((y[:, 2:] - y[:, 1:-1]) / h[:, 1:] - (y[:, 1:-1] - y[:, :-2]) / h[:, :-1])
            cp = np.zeros((B, m), dtype=dtype)
            dp = np.zeros((B, m), dtype=dtype)
            cp[:, 0] = c[:, 0] / bd[:, 0]
            dp[:, 0] = d[:, 0] / bd[:, 0]
            for i in range(1, m):
                denom = bd[:, i] - a[:, i] * cp[:, i - 1]
                cp[:, i] = c[:, i] / denom if i < m - 1 else 0.0
                dp[:, i] = (d[:, i] - a[:, i] * dp[:, i - 1]) / denom
            Si = np.zeros((B, m), dtype=dtype)
            Si[:, m - 1] = dp[:, m - 1]
            for idx in range(m - 2, -1, -1):
                Si[:, idx] = dp[:, idx] - cp[:, idx] * Si[:, idx + 1]
            S[:, 1:-1] = Si
        ai = y[:, :-1]
        bi = (y[:, 1:] - y[:, :-1]) / h - h * (S[:, 1:] + 2.0 * S[:, :-1]) / 6.0
        ci = S[:, :-1] / 2.0
        di = (S[:, 1:] - S[:, :-1]) / (6.0 * h)
        I_full = ai * h + bi * h**2 / 2 + ci * h**3 / 3 + di * h**4 / 4
        self.I_cum = np.concatenate([np.zeros((B, 1), dtype=dtype), np.cumsum(I_full, axis=-1)], axis=-1)
        self.I_total = self.I_cum[:, -1]
        self.x, self.y, self.S = x, y, S
        self.integrate_from_start = integrate_from_start

    @property
    def n(self) -> int:
        return self.x.shape[-1]

    def _local(self, xn):
        xn = np.asarray(xn)
        scalar = xn.ndim == 0
        if scalar:
            xn = np.broadcast_to(xn, (self.x.shape[0],))
        idx = interval_index(self.x, xn, self.shared_x)
        x0 = _take(self.x, idx)
        x1 = _take(self.x, idx + 1)
        return xn, idx, x0, x1

    def evaluate(self, xn):
        """Spline value(s); ``xn`` of shape (), (B,) or (B, m)."""
        xn, idx, x0, x1 = self._local(xn)
        h = x1 - x0
        t = (xn - x0) / h
        A = 1 - t
        Bt = t
        y0 = _take(self.y, idx)
        y1 = _take(self.y, idx + 1)
        s0 = _take(self.S, idx)
        s1 = _take(self.S, idx + 1)
        return A * y0 + Bt * y1 + ((A**3 - A) * s0 + (Bt**3 - Bt) * s1) * (h**2) / 6.0

    def derivative(self, xn):
        xn, idx, x0, x1 = self._local(xn)
        h = x1 - x0
        dv = xn - x0
        y0 = _take(self.y, idx)
        y1 = _take(self.y, idx + 1)
        s0 = _take(self.S, idx)
        s1 = _take(self.S, idx + 1)
        b = (y1 - y0) / h - h * (s1 + 2 * s0) / 6.0
        c = s0 / 2.0
        d = (s1 - s0) / (6.0 * h)
        return b + 2 * c * dv + 3 * d * dv**2

    def integral(self, xn):
        xn, idx, x0, x1 = self._local(xn)
        h = x1 - x0
        dv = xn - x0
        y0 = _take(self.y, idx)
        y1 = _take(self.y, idx + 1)
        s0 = _take(self.S, idx)
        s1 = _take(self.S, idx + 1)
        b = (y1 - y0) / h - h * (s1 + 2 * s0) / 6.0
        c = s0 / 2.0
        d = (s1 - s0) / (6.0 * h)
        part = y0 * dv + b * dv**2 / 2 + c * dv**3 / 3 + d * dv**4 / 4
        fwd = _take(self.I_cum, idx) + part
        if self.integrate_from_start:
            return fwd
        tot = self.I_total if xn.ndim == 1 else self.I_total[:, None]
        return tot - fwd

    def pack(self) -> np.ndarray:
        """(B, 3, n) array of (x, y, S) knots: the device table format."""
        return np.stack([self.x, self.y, self.S], axis=1)
