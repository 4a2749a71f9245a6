"""MI355X-native space-charge step of IMPACT-Z (host side).

``impactz.hip`` binds the HIP C ABI (include/impactz_hip.h) and mirrors the
reference's BeamBunch / CompDom / FieldQuant procedures; ``impactz.beam``
defines beams, lattices and the seeded synthetic particles; ``impactz.sim``
runs a lattice in the reference's segment order.
"""
from . import beam  # noqa: F401

__all__ = ["beam", "hip", "sim"]
