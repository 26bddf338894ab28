"""em2d: the reference's Python front-end (openmp/python/em2d) on the MI355X
engine.  Same package layout and class names as the reference's Cython
module: em2d.Simulation, em2d.Species, em2d.EMF, em2d.Current, em2d.Vec3Grid,
em2d.zpic, and the em2d.density / em2d.udist / em2d.filter / em2d.laser
classes.  The core (em2d.em2d) is a Cython extension over the C ABI of
libzpic_amd.so (include/zpic_amd.h); build it with
`python zpic-parallel_amd/build.py`."""
from .em2d import *  # noqa: F401,F403
from . import density, udist, filter, laser  # noqa: F401
