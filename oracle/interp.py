"""Uniform-grid SK table interpolation (oracle, test infrastructure only).

Restates PolyInterpU.__call__ (slakonet/interpolation.py:107-238), poly_interp
(:255-311) and poly5_zero (:241-252):

* table row j holds the value at r = (j + 1) * dr;
* ind = floor(r / dr); for ind <= n_grid the 8-point stencil ends at
  ind_last = clamp(ind + 5, 9, n_grid) and uses rows ind_last-9 .. ind_last-2
  (grid points (ind_last-8 .. ind_last-1) * dr), evaluated by Neville's
  tableau along the top path (the closest-point search of :277-284 never moves,
  so the path index stays 0);
* for n_grid <= ind <= n_grid - 1 + int(tail / dr) the DFTB+ fifth-order tail
  to zero at r_max = (n_grid - 1) dr + tail overwrites the polynomial value;
* beyond that the value is 0.
"""
import numpy as np


def neville_top(xa, yb, r):
    """Neville tableau along index 0 (slakonet/interpolation.py:286-311).

    xa [m, 8], yb [m, 8, k], r [m] -> [m, k]
    """
    n = xa.shape[1]
    cc = yb.copy()
    dd = yb.copy()
    y = yb[:, 0].copy()
    for mm in range(n - 1):
        for ii in range(n - mm - 1):
            r0 = xa[:, ii] - xa[:, ii + mm + 1]
            r1 = (cc[:, ii + 1] - dd[:, ii]) / r0[:, None]
            cc[:, ii] = (xa[:, ii] - r)[:, None] * r1
            dd[:, ii] = (xa[:, ii + mm + 1] - r)[:, None] * r1
        y = y + cc[:, 0]
    return y


def poly5_zero(y0, y0p, y0pp, xx, dx):
    """DFTB+ poly5ToZero (slakonet/interpolation.py:241-252)."""
    dx1 = y0p * dx
    dx2 = y0pp * dx * dx
    dd = 10.0 * y0 - 4.0 * dx1 + 0.5 * dx2
    ee = -15.0 * y0 + 7.0 * dx1 - 1.0 * dx2
    ff = 6.0 * y0 - 3.0 * dx1 + 0.5 * dx2
    xr = xx / dx
    return ((ff * xr + ee) * xr + dd) * xr * xr * xr


def poly_interp_u(yy, grid_step, r, tail=1.0, delta_r=1e-5, n_interp=8, n_interp_r=4):
    """Interpolate a table yy [n_grid, k] at distances r [m] -> [m, k]."""
    yy = np.asarray(yy, dtype=np.float64)
    if yy.ndim == 1:
        yy = yy[:, None]
    r = np.asarray(r, dtype=np.float64).reshape(-1)
    n_grid = yy.shape[0]
    r_max = (n_grid - 1) * grid_step + tail
    ind = np.floor(r / grid_step).astype(np.int64)
    out = np.zeros((len(r), yy.shape[1]))
    mask = ind <= n_grid
    if mask.any():
        il = ind[mask] + n_interp_r + 1
        il[il > n_grid] = n_grid
        il[il < n_interp + 1] = n_interp + 1
        xa = (il[:, None] - n_interp + np.arange(n_interp)[None, :]) * grid_step
        rows = il[:, None] - n_interp - 1 + np.arange(n_interp)[None, :]
        yb = yy[rows]
        out[mask] = neville_top(xa, yb, r[mask])
    max_ind = n_grid - 1 + int(tail / grid_step)
    is_tail = (ind >= n_grid) & (ind <= max_ind)
    if is_tail.any():
        dr = r[is_tail] - r_max
        ilast = n_grid
        xa = (ilast - n_interp + np.arange(n_interp)) * grid_step
        yb = yy[ilast - n_interp - 1: ilast - 1]
        m = 1
        xa2 = np.repeat(xa[None], m, 0)
        yb2 = np.repeat(yb[None], m, 0)
        y0 = neville_top(xa2, yb2, xa2[:, n_interp - 1] - delta_r)[0]
        y2 = neville_top(xa2, yb2, xa2[:, n_interp - 1] + delta_r)[0]
        y1 = yy[ilast - 2]
        y1p = (y2 - y0) / (2.0 * delta_r)
        y1pp = (y2 + y0 - 2.0 * y1) / (delta_r * delta_r)
        out[is_tail] = poly5_zero(y1[None], y1p[None], y1pp[None], dr[:, None], -1.0 * tail)
    return out
