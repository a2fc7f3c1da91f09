"""emerald_amd: the reference sph2d_box 2D SPH step on MI355X (gfx950).

The compute path is the HIP library libemerald_sph.so (C ABI in
include/emerald_sph.h); this package is its Python host mirror:
  sph2d_box   Parameters, Simulation_config, State, solids/dam-break inputs,
              <method>_simulation_step, EZ_EXAMPLE_SIM
  search      stand-alone z-index sort + neighbour search
  scenes      seeded synthetic workloads
  build       in-tree hipcc build of the library
"""

from ._capi import EmsphError, lib  # noqa: F401

__all__ = ["EmsphError", "lib"]
