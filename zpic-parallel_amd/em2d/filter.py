"""Digital current filters (the reference's em2d.filter: None, Binomial,
Compensated; filter.h:9-78)."""

_DIRS = {"x": 0, "y": 1}


class None_:
    """None(): no filtering."""
    kind = 0
    dir = 0
    order = 1


class Binomial:
    """Binomial( dir, level = 1 ): `level` passes of [1/4, 1/2, 1/4] along dir."""
    kind = 1

    def __init__(self, dir, level=1):
        self.dir = _DIRS[dir]
        self.order = max(1, int(level))


class Compensated(Binomial):
    """Compensated( dir, level = 1 ): binomial passes then the compensator."""
    kind = 2


globals()["None"] = None_
