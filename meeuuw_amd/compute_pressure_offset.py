"""Drop-in compute_pressure_offset (compute_pressure_offset.py:11-66), FP64 on the GPU."""
import numpy as np

from . import _dev as dev
from . import _lib

MODES = {"none": 0, "surface": 1, "volume": 2}


def compute_pressure_offset(geometry, pressure_normalisation, axisymmetric, top_element, nel, nelx,
                            nq_per_element, N_P, JxWq, p, icon_P, theta_P, xq, volume):
    if pressure_normalisation not in MODES:
        raise ValueError("compute_pressure_offset: unknown pressure_normalisation value")
    mode = MODES[pressure_normalisation]
    numpy_out = not dev.is_dev(p)
    ip = dev.to_dev(icon_P, np.int32)
    pd = dev.to_dev(p, np.float64)
    top = dev.to_dev(top_element, np.uint8)
    Np = dev.to_dev(N_P, np.float64)
    jw = dev.to_dev(JxWq, np.float64)
    out, work = dev.empty(1), dev.empty(1024)
    if axisymmetric:  # compute_pressure_offset.py:33-39, 49-55
        if (mode == 1 and theta_P is None) or (mode == 2 and xq is None):
            raise ValueError("compute_pressure_offset: axisymmetric needs theta_P (surface) or xq "
                             "(volume)")
        th = dev.to_dev(theta_P, np.float64) if mode == 1 else None
        xqd = dev.to_dev(xq, np.float64) if mode == 2 else None
        _lib.call("mw_pressure_offset_axisymmetric", mode, dev.ptr(top), int(nel), dev.ptr(Np),
                  dev.ptr(jw), dev.ptr(pd), dev.ptr(ip), dev.ptr(th), dev.ptr(xqd), dev.ptr(out),
                  dev.ptr(work), dev.stream())
    else:
        _lib.call("mw_pressure_offset", mode, dev.ptr(top), int(nel), int(nelx), dev.ptr(Np),
                  dev.ptr(jw), dev.ptr(pd), dev.ptr(ip), float(volume), dev.ptr(out),
                  dev.ptr(work), dev.stream())
    return float(dev.host(out)[0]) if numpy_out else out[0]
