"""MI355X (gfx950) W-projection gridder: host-side Python interface.

- ``wproj``     numpy mirror of oskar_grid_wproj_f/_d (host pointers, drop-in)
- ``torch_ops`` device-resident gridding / degridding on torch tensors and the
                ``GridWProj`` autograd function (import needs torch)
- ``synth``     seeded synthetic workloads of the reference data sets' shapes
- ``_lib``      ctypes bindings of include/mi355_gridder.h
"""
from ._lib import GridderError, Context, load  # noqa: F401
from . import wproj, synth  # noqa: F401
