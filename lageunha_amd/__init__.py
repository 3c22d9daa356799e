"""lageunha_amd — MI355X-native TreePM gravity step (EUNHA/GOTPM) behind a C ABI.

The compute path is liblagpm.so (HIP for gfx950 + rocFFT); `lagpm` is the ctypes mirror of
the reference's entry points and `ics` builds the synthetic workloads.
"""
from . import ics, lagpm  # noqa: F401

__all__ = ["ics", "lagpm"]
