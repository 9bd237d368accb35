"""MI355X (gfx950) implementation of ParaTC's per-time-step hot path.

The compute path is the HIP library ``libparatc_gpu.so`` (sources in
``csrc/``, C-ABI in ``include/paratc_gpu.h``).  This package holds the host
side: case set-up (``case``), seeded inputs (``rng``), the ctypes binding
(``capi``) and the mirror of the reference's procedures (``solver``).
"""
from .case import Case, MeshSection, case_channel, case_gravity_current, case_poisson, clc_ycoord  # noqa: F401

__all__ = ["Case", "MeshSection", "case_channel", "case_gravity_current", "case_poisson", "clc_ycoord"]
