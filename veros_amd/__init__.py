"""veros_amd — MI355X (gfx950) implementation of veros's per-timestep mixing hot path.

Public API (same names and argument meaning as the reference):

* ``utilities.create_water_masks``, ``utilities.solve_implicit``
  (veros/core/utilities.py:38-59), plus ``utilities.solve_implicit_kbot``;
* ``isoneutral.isoneutral_diffusion_pre`` (veros/core/isoneutral/isoneutral.py:18-229);
* ``isoneutral.isoneutral_diffusion`` / ``isoneutral_diffusion_kernel``
  (veros/core/isoneutral/diffusion.py:210-295), ``isoneutral_skew_diffusion``
  (:298-307) and the fused temperature + salinity pass;
* ``batched.TdmaBatch``: many grids' TDMA in one persistent launch;
* ``autodiff``: torch.autograd wrappers over the adjoint kernels.

All compute runs in libvxmix.so (hand-written HIP kernels behind the C ABI
in include/vxmix.h); importing a compute function without the built library
raises, there is no CPU fallback.
"""
from . import workloads  # noqa: F401  (host-side input generators, no GPU needed)

__all__ = ["utilities", "isoneutral", "state", "workloads", "batched", "autodiff"]


def __getattr__(name):
    if name in ("utilities", "isoneutral", "state", "batched", "autodiff"):
        import importlib
        return importlib.import_module(f".{name}", __name__)
    raise AttributeError(name)
