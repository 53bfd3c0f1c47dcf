"""CloverLeaf3D hydro timestep kernels, MI355X-native (gfx950 HIP).

Modules:
  deck    the reference's input-deck reader (read_input.f90 / parse.f90)
  chunk   the device-resident hydro loop over the C ABI
  _lib    the ctypes binding of include/clover_hip.h
"""
from .deck import Deck, parse_deck, cuboid_deck, DEFAULT_DECK  # noqa: F401

__all__ = ["Deck", "parse_deck", "cuboid_deck", "DEFAULT_DECK", "Chunk"]


def __getattr__(name):
    if name == "Chunk":
        from .chunk import Chunk
        return Chunk
    raise AttributeError(name)
