"""do_BCs for 2D arrays (the reference's jsolver/boundary_handler_2d.py:3-29).

Returns a new float64 GPU tensor whose first and last ``rim + 1`` columns
(direction 0), rows (direction 1) or both (-1) are zero or the given values:
vals_xb / vals_xe have shape ``arr.shape[0]`` (one value per row, set into
the leading / trailing columns), vals_yb / vals_ye have shape ``arr.shape[1]``.
Columns are written before rows, as in the reference.  Runs the HIP kernel
jsv_do_bcs_2d; the input is not modified.
"""
from . import _lib
from ._device import device, ptr, stream_ptr, to_device


def do_BCs(arr, rim=1, direction=-1, vals_xb=None, vals_xe=None, vals_yb=None, vals_ye=None):
    dev = device()
    out = to_device(arr, dev).clone()
    if out.ndim != 2:
        raise ValueError(f"do_BCs (2D) needs a two-dimensional array, got shape {tuple(out.shape)}")
    ny, nx = out.shape
    v = [None if a is None else to_device(a, dev) for a in (vals_xb, vals_xe, vals_yb, vals_ye)]
    for a, n in zip(v, (ny, ny, nx, nx)):
        if a is not None and a.numel() != n:
            raise ValueError(f"boundary values of size {a.numel()} do not match {n}")
    _lib.check(_lib.load().jsv_do_bcs_2d(ptr(out), ny, nx, int(rim), int(direction),
                                         *(ptr(a) for a in v), stream_ptr()))
    return out
