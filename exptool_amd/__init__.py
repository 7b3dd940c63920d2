"""MI355X-native basis-function-expansion (BFE) hot path of exptool.

Sub-packages mirror the reference layout for the path:
  exptool_amd.basis.spheresl / eof / potential / compatibility / tablegen
  exptool_amd.utils.halo_methods / integrate / samplers
  exptool_amd.io.particle
The device work goes through libexptool_bfe.so (C ABI in include/exptool_bfe.h)
via exptool_amd.bfe.BFESystem.
"""
__version__ = "0.1.0"
