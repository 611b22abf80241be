"""Slater-Koster sub-blocks by frame rotation (oracle, test infrastructure only).

Restates sub_block_rot / sub_block_ref (slakonet/slaterkoster.py:1013-1122)
and the s/p/d rotation matrices _rot_{s,p,d}_{yz,xy}
(slakonet/slaterkoster.py:1130-1372):

* orbital order p = (y, z, x), d = (xy, yz, z^2, xz, x^2-y^2);
* integrals are ordered sigma, pi, delta; the reference-frame block for
  l1 <= l2 carries [.., delta, pi, sigma, pi, delta ..] on the diagonal with
  offset l2 - l1;
* lab block = R_l1(u) . I . R_l2(u)^T (the einsum "ji,ik,ck->jc");
* the yz-type rotation is used when |u_y| > |u_z|, else the xy-type.

The matrices below are written column by column (each listed vector is one
column), which is how the reference stacks them.
"""
import numpy as np

SQR3 = np.sqrt(3.0)


def _cols(*cols):
    # each col: [m] arrays list of length k -> R [m, k, ncols]
    return np.stack([np.stack(c, -1) for c in cols], -1)


def rot_s(u):
    return np.linalg.norm(u, axis=-1)[:, None, None]


def rot_p_yz(u):
    x, y, z = u[:, 0], u[:, 1], u[:, 2]
    zero = np.zeros_like(x)
    a = np.sqrt(1.0 - z * z)
    return _cols([x / a, zero, -y / a], [y, z, x], [y * z / a, -a, x * z / a])


def rot_p_xy(u):
    x, y, z = u[:, 0], u[:, 1], u[:, 2]
    zero = np.zeros_like(x)
    a = np.sqrt(1.0 - y * y)
    return _cols([a, -y * z / a, -x * y / a], [y, z, x], [zero, -x / a, z / a])


def rot_d_yz(u):
    x, y, z = u[:, 0], u[:, 1], u[:, 2]
    zero = np.zeros_like(x)
    a = 1.0 - z * z
    b = np.sqrt(a)
    xz, xy, yz = x * z, y * x, z * y
    xyz = x * yz
    x2 = x * x
    return _cols(
        [-z + 2.0 * x2 * z / a, -x, zero, y, -2.0 * xyz / a],
        [-b + 2.0 * x2 / b, xz / b, zero, -yz / b, -2.0 * xy / b],
        [xy * SQR3, yz * SQR3, 1.0 - 1.5 * a, xz * SQR3, SQR3 * (-0.5 * a + x2)],
        [2.0 * xyz / b, -2.0 * y * b + y / b, -SQR3 * z * b, -2.0 * x * b + x / b,
         -z * b + 2.0 * x2 * z / b],
        [-xy + 2.0 * xy / a, -yz, 0.5 * SQR3 * a, -xz, 0.5 * a - 1.0 + x2 * (-1.0 + 2.0 / a)],
    )


def rot_d_xy(u):
    x, y, z = u[:, 0], u[:, 1], u[:, 2]
    a = 1.0 - y * y
    b = np.sqrt(a)
    xz, xy, yz = x * z, y * x, z * y
    xyz = x * yz
    z2 = z * z
    return _cols(
        [z, -x, xyz * SQR3 / a, y * (1 - 2 * z2 / a), -xyz / a],
        [x * (2 * b - 1 / b), z * (2 * b - 1.0 / b), -y * z2 * SQR3 / b, -2 * xyz / b,
         y * (-2 * b + z2 / b)],
        [xy * SQR3, yz * SQR3, 1.5 * z2 - 0.5, xz * SQR3, 0.5 * SQR3 * (2 * a - z2 - 1)],
        [yz / b, -xy / b, -xz * SQR3 / b, -b + 2 * z2 / b, xz / b],
        [xy, yz, SQR3 * (0.5 * (z2 + 1) - z2 / a), xz - 2 * xz / a, a - 0.5 * z2 - 0.5 + z2 / a],
    )


_YZ = {0: rot_s, 1: rot_p_yz, 2: rot_d_yz}
_XY = {0: rot_s, 1: rot_p_xy, 2: rot_d_xy}


def sub_block_ref(l1, l2, integrals):
    """Reference-frame block [m, 2l1+1, 2l2+1] for l1 <= l2 (slaterkoster.py:1075-1122)."""
    m = integrals.shape[0]
    blk = np.zeros((m, 2 * l1 + 1, 2 * l2 + 1))
    size = l1 + 1
    diag = np.zeros((m, 2 * l1 + 1))
    diag[:, -size:] = integrals
    if size > 1:
        diag[:, : size - 1] = integrals[:, 1:][:, ::-1]
    for t in range(2 * l1 + 1):
        blk[:, t, t + l2 - l1] = diag[:, t]
    return blk


def sub_block_rot(l1, l2, u, integrals):
    """Lab-frame SK block for l1 <= l2, unit vectors u [m,3] row->column atom."""
    assert l1 <= l2
    u = np.asarray(u, dtype=np.float64)
    integrals = np.asarray(integrals, dtype=np.float64).reshape(len(u), l1 + 1)
    if l1 == 0 and l2 == 0:
        return integrals[:, None, :]
    imat = sub_block_ref(l1, l2, integrals)
    out = np.zeros((len(u), 2 * l1 + 1, 2 * l2 + 1))
    yz = np.abs(u[:, 1]) > np.abs(u[:, 2])
    for rots, mask in ((_YZ, yz), (_XY, ~yz)):
        if mask.any():
            ua = u[mask]
            ra = rots[l1](ua)
            rb = rots[l2](ua)
            out[mask] = np.einsum("mji,mik,mck->mjc", ra, imat[mask], rb)
    return out
