"""Differentiable TD objective on top of the libtorch autograd Function ``_qbh_torch``.

``td_objective`` returns J(v_ext, A0) = sum_n w_n e.j(psi_n, A0 * shape_n) of the GPU
propagator (``TDPropagator``), differentiable in ``vext`` (device float64 grid) and ``A0``
(scalar float64 tensor); the gradients come from the hand-written adjoint in libqbh.so.
The extension module must have been built (``make -C qball_amd/csrc torch``): there is no
fallback.
"""
import numpy as np
import torch

from .workload import dipole_weights, field_shape  # noqa: F401  (part of this module's API)

try:
    from . import _qbh_torch
except ImportError as exc:  # the product path has no substitute
    raise ImportError("qball_amd/_qbh_torch*.so is not built: run make -C qball_amd/csrc torch") from exc


def td_objective(prop, psi0_ptr, ld, vext, A0, nsteps, dt, shape, w, e=(1.0, 0.0, 0.0), order=4,
                 checkpoint=20, flags=0):
    """J as a 0-d float64 tensor with autograd through ``vext`` and ``A0``."""
    shape = torch.as_tensor(np.ascontiguousarray(shape, dtype=np.float64)).reshape(nsteps + 1, 3)
    w = torch.as_tensor(np.ascontiguousarray(w, dtype=np.float64)).reshape(nsteps + 1)
    e = torch.as_tensor(np.ascontiguousarray(e, dtype=np.float64)).reshape(3)
    return _qbh_torch.td_objective(vext, A0, int(prop.h.value), int(psi0_ptr), int(ld), int(nsteps),
                                   float(dt), shape, w, e, int(order), int(checkpoint), int(flags))
