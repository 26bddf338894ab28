"""Density profiles (the reference's em2d.density: None, Uniform, Step, Slab;
density.h:30-132), plus Sphere (in the reference's C++ API) and Ramp (no
reference counterpart: linear from 0 at `begin` to n0 at `end`).  Plain
parameter holders; Species.set_density hands them to the engine."""

_DIRS = {"x": 0, "y": 1}


class _Profile:
    kind = 1
    dir = 0
    n0 = 1.0
    pos = begin = end = radius = 0.0
    center = (0.0, 0.0)


class None_(_Profile):
    """None(): no particles."""
    kind = 0

    def __init__(self):
        self.n0 = 0.0


class Uniform(_Profile):
    """Uniform( n0 )"""
    kind = 1

    def __init__(self, n0):
        self.n0 = float(n0)


class Step(_Profile):
    """Step( dir, n0, pos ): density n0 beyond `pos` along dir ('x' or 'y')."""
    kind = 2

    def __init__(self, dir, n0, pos):
        self.dir = _DIRS[dir]
        self.n0 = float(n0)
        self.pos = float(pos)


class Slab(_Profile):
    """Slab( dir, n0, begin, end ): density n0 in [begin, end) along dir."""
    kind = 3

    def __init__(self, dir, n0, begin, end):
        self.dir = _DIRS[dir]
        self.n0 = float(n0)
        self.begin = float(begin)
        self.end = float(end)


class Sphere(_Profile):
    """Sphere( n0, center, radius ) (Density::Sphere, density.h:117-130)."""
    kind = 4

    def __init__(self, n0, center, radius):
        self.n0 = float(n0)
        self.center = (float(center[0]), float(center[1]))
        self.radius = float(radius)


class Ramp(_Profile):
    """Ramp( dir, n0, begin, end ): no reference counterpart; density rising
    linearly from 0 at `begin` to n0 at `end` along dir, n0 beyond."""
    kind = 5

    def __init__(self, dir, n0, begin, end):
        if not end > begin:
            raise ValueError("Invalid ramp, end must be > begin")
        self.dir = _DIRS[dir]
        self.n0 = float(n0)
        self.begin = float(begin)
        self.end = float(end)


globals()["None"] = None_
