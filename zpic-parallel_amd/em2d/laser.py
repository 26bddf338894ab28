"""Laser pulses (the reference's em2d.laser: PlaneWave, Gaussian; laser.h).
add( emf ) adds the pulse to the E and B fields of an EMF object."""


class PlaneWave:
    """PlaneWave( start = 0, fwhm = 0, rise = 0, flat = 0, fall = 0, a0 = 0,
    omega0 = 0, polarization = 0, cos_pol = 0, sin_pol = 0 )"""
    kind = 0

    def __init__(self, *, start=0.0, fwhm=0.0, rise=0.0, flat=0.0, fall=0.0, a0=0.0, omega0=0.0,
                 polarization=0.0, cos_pol=0.0, sin_pol=0.0):
        self.start, self.fwhm, self.rise, self.flat, self.fall = start, fwhm, rise, flat, fall
        self.a0, self.omega0 = a0, omega0
        self.polarization, self.cos_pol, self.sin_pol = polarization, cos_pol, sin_pol
        self.filter = 1
        self.W0 = self.focus = self.axis = 0.0

    def add(self, emf):
        """add( emf ): adds the pulse to the EM fields."""
        emf._add_laser(self)


class Gaussian(PlaneWave):
    """Gaussian( ..., W0 = 0, focus = 0, axis = 0 ): Gaussian beam of waist W0
    focused at x = focus on the axis y = axis."""
    kind = 1

    def __init__(self, *, W0=0.0, focus=0.0, axis=0.0, **kw):
        super().__init__(**kw)
        self.W0, self.focus, self.axis = W0, focus, axis
