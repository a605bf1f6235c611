"""dohp_amd -- MI355X-native matrix-free hp element operator of dohp (jedbrown/dohp).

Submodules:
  capi      ctypes binding of the HIP library's C ABI (include/dohp_amd.h); raises on
            import when dohp_amd/libdohp_amd.so has not been built.
  mesh      synthetic hexahedral block meshes (vertex coordinates, seeded perturbation).
  fields    seeded random coefficient fields (log-normal k).
"""
