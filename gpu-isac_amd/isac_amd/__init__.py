"""MI355X-native polar-ring 2D alignment (GPU-ISAC hot path).

``binding`` holds the ctypes layer over gpu_aln_pack.so (C ABI of the
reference library); ``synth`` the seeded synthetic-particle generator.
"""
from . import binding, synth  # noqa: F401

__all__ = ["binding", "synth"]
