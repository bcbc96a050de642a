"""hexdyn: MPAS-A v8.4.2 split-explicit nonhydrostatic dycore large step for
AMD MI355X (gfx950), behind a C-ABI library (libhexdyn.so) with a Python
driver over PyTorch-ROCm tensors.

Modules:
  mesh     icosahedral SCVT meshes with MPAS connectivity and TRiSK weights
  initial  vertical grid, terrain, reference state, synthetic initial states
  setup    RK schedule, off-centring, w damping, advection coefficients
  case     named synthetic workloads
  config   DycoreConfig (MPAS namelist names, refusals)
  capi     ctypes mirrors of include/hexdyn.h
  lib      libhexdyn.so loader (no CPU fallback)
  driver   HexDycoreDriver (from_host / step_device / step)
  bytes_model  algorithmic HBM bytes per large step
"""

__version__ = "0.1.0"
