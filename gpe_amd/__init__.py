"""MI355X (gfx950, HIP) time-stepping path of the Gross-Pitaevskii solver.

Modules:
  sim        Simulation: one field resident on the GPU (solver steps, norm,
             energy, line length, vortex cores, condensate and spectrum,
             filtered field, imprinting, file dumps)
  batch      Batch: many uneven simulations advanced together
  serve      Server: continuous batching of a request stream
  config     RunConfig: the reference's run parameters
  ic         host-side initial states
  state_io   the reference's end_state, dens, zeros, mass, spectrum and
             filtered file formats
  tuning     launch-configuration table (tile variant, z-chunk per stage)
  accuracy   compensated FP32 against FP64 (psi, energy, line length, cores)
  workloads  the batch and serving workloads
  build      hipcc build of libgpehip.so
  _lib       ctypes binding of include/gpe_hip.h

The library must be built (python -m gpe_amd.build); nothing falls back to
a host implementation.
"""
