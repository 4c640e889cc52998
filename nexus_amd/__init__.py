"""MI355X-native (gfx950) sparse-grid MLS-MPM substep with the nexus_mpm pipeline API.

Layout:
  csrc/            HIP kernels (mpm_kernels.hip) and host runtime + C ABI (mpm_api.hip)
  mpm.py           Python mirror of MpmState / MpmPipeline over the C ABI
  scenes.py        seeded scene generators for the workloads
  build.py         in-tree hipcc build of libnexus_mpm.so
"""
from . import scenes  # noqa: F401

__all__ = ["scenes"]
