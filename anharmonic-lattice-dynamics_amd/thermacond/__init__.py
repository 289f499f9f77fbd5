"""alt3 host side: the THERMACOND three-phonon path on MI355X.

Modules
  capi        ctypes binding of libalt3.so (include/alt3.h) and the Context wrapper
  refio       THERMACOND input/output file formats
  splitio     the split-job (job 4/5) |V3|^2 and collision-matrix files
  models      model crystals (Stillinger-Weber Si, Born-Mayer binaries) and their FCs
  selfenergy  job 2: three- and four-phonon self-energies at given q
  serving     continuous-batching kappa-solve server with shared |V3|^2 launches
  workloads   the benchmark workloads (rock-salt 12^3, the SW-Si sweep, the request stream)
  profiling   summaries of rocprofv3 output for the library's kernels
  build       hipcc build of libalt3.so (gfx950) and of the Fortran job-1 driver

The job-1 driver over the C ABI is Fortran: fortran/alt3_kappa.f90.
"""
