"""spc5_amd -- SPC5 beta(r,VS) sparse matrix-vector products on MI355X (gfx950).

Layers:
  include/spc5_hip.h    C ABI of libspc5_hip.so (reference ABI + device API)
  include/spc5_hip.hpp  C++ drop-in for the reference's spc5.hpp
  spc5_amd.spc5         Python mirror of the spc5.hpp API (host arrays)
  spc5_amd.device       device-resident matrices on torch tensors
"""
from ._lib import Spc5Error, lib  # noqa: F401

__all__ = ["Spc5Error", "lib"]
