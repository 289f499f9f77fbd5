"""alt3 host side: the THERMACOND three-phonon path on MI355X.

Modules
  capi     ctypes binding of libalt3.so (include/alt3.h) and the Context wrapper
  refio    THERMACOND input/output file formats
  models   model crystals (Stillinger-Weber Si, Born-Mayer binaries) and their FCs
  build    hipcc build of libalt3.so (gfx950)
  kappa    job-1 driver mirroring THERMACOND/kappa.f90:389-871 on the C ABI
  serving  continuous-batching kappa-solve server
"""
