"""svfsi_amd: MI355X (gfx950) implementation of svFSI's per-Newton-step
linear-system path -- element tangent/residual construction, DOASSEM scatter
into svFSILS block-CSR, and the svFSILS Krylov solve -- behind a C ABI
(include/svfsi_amd.h) that keeps the svFSILS call signatures.

Modules:
  _lib      ctypes binding of libsvfsi_amd.so (no CPU fallback)
  fsils     FSILS_LHS_CREATE / BC_CREATE / LS_CREATE / SOLVE mirror
  assembly  device-resident element and face assembly
  mesh      synthetic tet4 / tet10 meshes
  problems  inputs of the benchmark configurations
  step      one Newton step (assemble + solve) on the GPU
  build     hipcc build of the library
"""
__version__ = "0.1.0"
