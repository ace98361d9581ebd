"""do_BCs for 3D arrays (the reference's jsolver/boundary_handler_3d.py:3-35).

Returns a new float64 GPU tensor with the first and last ``rim + 1`` layers
along x (direction 0), y (1), z (2) or all (-1) set to zero or to the given
values: vals_xb / vals_xe shaped ``(nz, ny)``, vals_yb / vals_ye shaped
``(nz, nx)``.  Faithful to the reference, the z faces also take vals_yb /
vals_ye (read as ``(ny, nx)``) and vals_zb / vals_ze are accepted but unused
(boundary_handler_3d.py:31-33).  The solver never passes values, so this only
concerns direct callers.  Runs the HIP kernel jsv_do_bcs_3d.
"""
from . import _lib
from ._device import device, ptr, stream_ptr, to_device


def do_BCs(arr, rim=1, direction=-1, vals_xb=None, vals_xe=None, vals_yb=None, vals_ye=None,
           vals_zb=None, vals_ze=None):
    dev = device()
    out = to_device(arr, dev).clone()
    if out.ndim != 3:
        raise ValueError(f"do_BCs (3D) needs a three-dimensional array, got shape {tuple(out.shape)}")
    nz, ny, nx = out.shape
    v = [None if a is None else to_device(a, dev) for a in (vals_xb, vals_xe, vals_yb, vals_ye)]
    if (vals_yb is not None or vals_ye is not None) and direction in (-1, 2) and nz != ny:
        raise ValueError("vals_yb/vals_ye are also used for the z faces and need nz == ny")
    _lib.check(_lib.load().jsv_do_bcs_3d(ptr(out), nz, ny, nx, int(rim), int(direction),
                                         *(ptr(a) for a in v), stream_ptr()))
    return out
