"""Momentum distributions (the reference's em2d.udist: None, Cold, Thermal,
ThermalCorr; udist.h:13-58)."""


class _UDist:
    kind = 0
    uth = (0.0, 0.0, 0.0)
    ufl = (0.0, 0.0, 0.0)
    npmin = 2


class None_(_UDist):
    """None(): zero momentum."""
    kind = 0


class Cold(_UDist):
    """Cold( ufl ): every particle at the fluid momentum ufl."""
    kind = 1

    def __init__(self, ufl):
        self.ufl = tuple(float(v) for v in ufl)


class Thermal(_UDist):
    """Thermal( uth, ufl ): Gaussian momenta of width uth around ufl."""
    kind = 2

    def __init__(self, uth, ufl):
        self.uth = tuple(float(v) for v in uth)
        self.ufl = tuple(float(v) for v in ufl)


class ThermalCorr(_UDist):
    """ThermalCorr( uth, ufl, npmin = 2 ): thermal, with the mean of every
    cell holding more than npmin particles set to ufl."""
    kind = 3

    def __init__(self, uth, ufl, npmin=2):
        if npmin <= 1:
            raise ValueError("invalid npmin parameter, must be > 1")
        self.uth = tuple(float(v) for v in uth)
        self.ufl = tuple(float(v) for v in ufl)
        self.npmin = int(npmin)


globals()["None"] = None_
