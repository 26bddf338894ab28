"""The em2d Python front-end (Cython over the C ABI) loads and mirrors the
reference's Python API (openmp/python/em2d): module layout, class names,
constructor keywords and methods.  No GPU calls."""
import os

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def em2d():
    import em2d
    return em2d


def test_module_layout(em2d):
    for name in ("Simulation", "Species", "EMF", "Current", "Vec3Grid", "Grid", "zpic"):
        assert hasattr(em2d, name), name
    assert em2d.em2d.__file__.endswith(".so")        # the compiled extension, not a Python shim
    for mod, names in (("density", ("Uniform", "Step", "Slab", "Sphere", "Ramp")),
                       ("udist", ("Cold", "Thermal", "ThermalCorr")),
                       ("filter", ("Binomial", "Compensated")),
                       ("laser", ("PlaneWave", "Gaussian"))):
        m = getattr(em2d, mod)
        for n in names:
            assert hasattr(m, n), (mod, n)
        assert getattr(m, "None", None) is not None or mod == "laser"


def test_method_surface(em2d):
    """The methods and properties the reference's classes define
    (em2d.pyx:280-1558), minus the plotting helpers."""
    surface = {
        "Vec3Grid": ["buffer_size", "zero", "set", "add", "gather", "copy_to_gc", "add_from_gc", "x_shift_left",
                     "kernel_x", "kernel_y", "save"],
        "Current": ["J", "n", "set_filter", "save"],
        "EMF": ["E", "B", "n", "box", "save", "advance"],
        "Species": ["name", "set_udist", "set_density", "save", "save_charge", "gather", "get_charge",
                    "get_phasespace", "save_phasespace"],
        "Simulation": ["t", "n", "ntiles", "box", "dt", "emf", "current", "add", "advance", "energy_info"],
    }
    for cls, names in surface.items():
        for n in names:
            assert hasattr(getattr(em2d, cls), n), (cls, n)


def test_courant_matches_float32_formula(em2d):
    """zpic::courant (zpic.h:64-91): sqrt(1 / (1/dx^2 + 1/dy^2)) in float."""
    f = np.float32
    for dx, dy in ((0.1, 0.1), (0.05, 0.2), (0.02, 0.2)):
        want = np.sqrt(f(1) / (f(1) / (f(dx) * f(dx)) + f(1) / (f(dy) * f(dy))))
        assert em2d.zpic.courant(dx=[dx, dy]) == float(want)
    # the LWFA driver: 1024 x 128 cells in a 20.48 x 25.6 box (main.cpp:548-553)
    a = em2d.zpic.courant(dims=[1024, 128], box=[20.48, 25.6])
    b = em2d.zpic.courant(ntiles=[32, 4], nx=[32, 32], box=[20.48, 25.6])
    assert a == b and abs(a - 1.0 / np.sqrt(1 / 0.02 ** 2 + 1 / 0.2 ** 2)) < 1e-6


def test_parameter_classes(em2d):
    d = em2d.density.Step("x", 1.0, 20.48)
    assert (d.kind, d.dir, d.n0, d.pos) == (2, 0, 1.0, 20.48)
    s = em2d.density.Slab("y", 0.5, 1.0, 2.0)
    assert (s.kind, s.dir, s.begin, s.end) == (3, 1, 1.0, 2.0)
    with pytest.raises(ValueError):
        em2d.density.Ramp("x", 1.0, 2.0, 1.0)
    u = em2d.udist.ThermalCorr([0.1, 0.1, 0.1], [0, 0, 0.6])
    assert (u.kind, u.npmin, u.ufl[2]) == (3, 2, 0.6)
    with pytest.raises(ValueError):
        em2d.udist.ThermalCorr([0.1] * 3, [0] * 3, npmin=1)
    f = em2d.filter.Compensated("x", 4)
    assert (f.kind, f.dir, f.order) == (2, 0, 4)
    g = em2d.laser.Gaussian(start=17.0, fwhm=2.0, a0=1.0, omega0=10.0, W0=4.0, focus=20.0, axis=12.8, sin_pol=1)
    assert (g.kind, g.W0, g.sin_pol, g.filter) == (1, 4.0, 1, 1)
