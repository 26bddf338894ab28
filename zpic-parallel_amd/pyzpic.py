"""Python front-end of the MI355X em2d library (libzpic_amd.so) over ctypes.

Mirrors the reference's Python layer (openmp/python/em2d/em2d.pyx:
Simulation, Species, EMF, Current) on top of the object entry points of
include/zpic_amd.h.  There is no CPU fallback: loading fails loudly when the
library is missing, and creating a Simulation fails when no GPU is present.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libzpic_amd.so")

UDIST = {"none": 0, "cold": 1, "thermal": 2, "thermal_corr": 3}
FILTER = {"none": 0, "binomial": 1, "compensated": 2}
STAGES = {"push": 0, "move": 1, "sort": 2, "species": 3, "zero_j": 4, "current": 5, "yee": 6,
          "add_gc": 7, "copy_gc": 8, "filter": 9, "emf_bc": 10, "yee_vacuum": 11}


class ZpicError(RuntimeError):
    pass


class Geom(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in
                ("ntiles_x", "ntiles_y", "nx", "ny", "gc_lower", "gc_upper", "ext_x", "ext_y", "tile_vol", "offset")]


_lib = None


def lib():
    """Loads libzpic_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ZpicError("libzpic_amd.so not built: run `python zpic-parallel_amd/build.py` "
                        "(or __graft_entry__.build())")
    L = ctypes.CDLL(LIB_PATH)
    vp, i, u, f, d = ctypes.c_void_p, ctypes.c_int, ctypes.c_uint, ctypes.c_float, ctypes.c_double
    u64 = ctypes.c_uint64
    sig = {
        "zpic_last_error": (ctypes.c_char_p, []),
        "zpic_geom_init": (i, [vp, i, i, i, i, i, i]),
        "zpic_simulation_new": (vp, [u, u, u, u, f, f, d]),
        "zpic_simulation_free": (None, [vp]),
        "zpic_species_new": (vp, [ctypes.c_char_p, f, u, u]),
        "zpic_species_free": (None, [vp]),
        "zpic_species_set_udist": (i, [vp, i, vp, vp, i]),
        "zpic_species_set_density_uniform": (i, [vp, f]),
        "zpic_species_set_seed": (i, [vp, u]),
        "zpic_species_set_bc": (i, [vp, vp]),
        "zpic_emf_set_bc": (i, [vp, vp]),
        "zpic_current_set_bc": (i, [vp, vp]),
        "zpic_simulation_add_species": (i, [vp, vp]),
        "zpic_simulation_advance": (i, [vp, i]),
        "zpic_simulation_stage": (i, [vp, i, i]),
        "zpic_current_set_filter": (i, [vp, i, i, i]),
        "zpic_simulation_set_deterministic": (i, [vp, i]),
        "zpic_simulation_get_iter": (u, [vp]),
        "zpic_simulation_get_t": (d, [vp]),
        "zpic_simulation_get_nmove": (u64, [vp]),
        "zpic_species_get_energy": (d, [vp]),
        "zpic_emf_get_energy": (i, [vp, vp, vp]),
        "zpic_species_get_nmove": (u64, [vp]),
        "zpic_species_np_total": (u64, [vp]),
        "zpic_species_get_particles": (i, [vp, vp, vp, vp, vp, vp]),
        "zpic_species_set_particles": (i, [vp, vp, vp, vp, vp]),
        "zpic_field_get": (i, [vp, i, vp]),
        "zpic_field_set": (i, [vp, i, vp]),
        "zpic_field_geom": (i, [vp, vp]),
        "zpic_species_charge": (ctypes.c_longlong, [vp, vp]),
        "zpic_species_gather": (i, [vp, i, vp]),
        "zpic_simulation_set_timing": (i, [vp, i]),
        "zpic_simulation_get_timing": (i, [vp, vp]),
        "zpic_simulation_sync": (i, [vp]),
        "zpic_batch_new": (vp, [i]),
        "zpic_batch_free": (None, [vp]),
        "zpic_batch_add": (i, [vp, vp]),
        "zpic_batch_set_graph": (i, [vp, i]),
        "zpic_batch_advance": (i, [vp, i]),
        "zpic_batch_set_persistent": (i, [vp, i]),
        "zpic_batch_add_threads": (i, [vp, vp, i]),
        "zpic_batch_status": (i, [vp]),
        "zpic_set_threads_per_tile": (i, [i]),
        "zpic_get_threads_per_tile": (i, []),
        "zpic_set_sort_order": (i, [i]),
        "zpic_get_sort_order": (i, []),
        "zpic_emf_save": (i, [vp, i, i]),
        "zpic_current_save": (i, [vp, i]),
        "zpic_field_save_file": (i, [vp, i, i, ctypes.c_char_p]),
        "zpic_species_save": (i, [vp]),
        "zpic_species_save_charge": (i, [vp]),
        "zpic_species_save_phasespace": (i, [vp, i, f, f, i, i, f, f, i]),
        "zpic_zdf_save_grid": (i, [vp, i, vp, vp, ctypes.c_char_p]),
        "zpic_fitted_threads_per_tile": (i, [i, i]),
        "zpic_denoiser_new": (vp, [i, i, i]),
        "zpic_denoiser_free": (None, [vp]),
        "zpic_denoiser_set_params": (i, [vp, vp]),
        "zpic_denoiser_get_params": (i, [vp, vp]),
        "zpic_denoiser_set_scale": (i, [vp, f]),
        "zpic_denoiser_get_scale": (f, [vp]),
        "zpic_denoiser_size": (i, [vp, vp, vp]),
        "zpic_denoiser_forward_host": (i, [vp, vp, vp]),
        "zpic_denoiser_gradient_host": (i, [vp, vp, vp, vp, vp]),
        "zpic_denoiser_num_pairs": (i, [vp]),
        "zpic_denoiser_clear_pairs": (i, [vp]),
        "zpic_denoiser_get_pair": (i, [vp, i, vp, vp]),
        "zpic_denoiser_target_rms": (i, [vp, vp]),
        "zpic_denoiser_train": (i, [vp, i, f, u, i, vp]),
        "zpic_denoiser_eval_pair": (i, [vp, i, vp, vp]),
        "zpic_simulation_add_pair": (i, [vp, vp, vp]),
        "zpic_current_set_learned_filter": (i, [vp, vp]),
        "zpic_current_set_fallback": (i, [vp, i]),
        "zpic_current_get_fallback": (i, [vp]),
        "zpic_simulation_copy_fields": (i, [vp, vp]),
        "zpic_species_subsample": (i, [vp, vp, i, i]),
        "zpic_autotune": (i, [u, u, u, u, i, ctypes.c_longlong, ctypes.c_char_p, vp]),
        "zpic_species_set_density": (i, [vp, vp]),
        "zpic_simulation_set_moving_window": (i, [vp]),
        "zpic_simulation_advance_mov_window": (i, [vp, i]),
        "zpic_simulation_window_moves": (i, [vp, vp]),
        "zpic_laser_add": (i, [vp, vp]),
        "zpic_emf_set_absorbing": (i, [vp, i, f]),
        "zpic_species_reinject": (i, [vp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _chk(rc):
    if rc != 0:
        raise ZpicError(lib().zpic_last_error().decode())
    return rc


PHASESPACE = {"x": 0, "y": 1, "ux": 2, "uy": 3, "uz": 4}


class ZdfAxis(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char_p), ("type", ctypes.c_int), ("min", ctypes.c_double),
                ("max", ctypes.c_double), ("label", ctypes.c_char_p), ("units", ctypes.c_char_p)]


class ZdfGridInfo(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char_p), ("ndims", ctypes.c_uint32), ("count", ctypes.c_uint64 * 3),
                ("label", ctypes.c_char_p), ("units", ctypes.c_char_p), ("axis", ctypes.POINTER(ZdfAxis))]


class ZdfIteration(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char_p), ("n", ctypes.c_int32), ("t", ctypes.c_double),
                ("time_units", ctypes.c_char_p)]


def _b(s):
    return None if s is None else s.encode()


def zdf_save_grid(data, name, path, label=None, units=None, axes=None, iteration=(None, 0, 0.0, None)):
    """zdf_save_grid on a host array (float32 / float64, C order [ny, nx]):
    writes <path>/<name>-<n:06>.zdf.  axes: list of (name, min, max, label,
    units) from the fastest axis; iteration: (name, n, t, time_units)."""
    data = np.ascontiguousarray(data)
    dtype = {np.dtype(np.float32): 9, np.dtype(np.float64): 10, np.dtype(np.int32): 5}[data.dtype]
    info = ZdfGridInfo()
    info.name = _b(name)
    info.ndims = data.ndim
    for k, n in enumerate(reversed(data.shape)):
        info.count[k] = n
    info.label, info.units = _b(label), _b(units)
    keep = []
    if axes:
        arr = (ZdfAxis * len(axes))()
        for k, (an, amin, amax, alab, aun) in enumerate(axes):
            arr[k] = ZdfAxis(_b(an), 0, amin, amax, _b(alab), _b(aun))
        info.axis = ctypes.cast(arr, ctypes.POINTER(ZdfAxis))
        keep.append(arr)
    it = ZdfIteration(_b(iteration[0]), iteration[1], iteration[2], _b(iteration[3]))
    _chk(lib().zpic_zdf_save_grid(_p(data), dtype, ctypes.byref(info), ctypes.byref(it), path.encode()))


class _Density(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("dir", ctypes.c_int), ("n0", ctypes.c_float), ("pos", ctypes.c_float),
                ("begin", ctypes.c_float), ("end", ctypes.c_float), ("center", ctypes.c_float * 2),
                ("radius", ctypes.c_float)]


_AXIS = {"x": 0, "y": 1, 0: 0, 1: 1}


class Density:
    """Density profiles (density.h:30-132; the reference's Cython classes
    em2d.density.None / Uniform / Step / Slab, plus Sphere)."""

    @staticmethod
    def none():
        return _Density(kind=0)

    @staticmethod
    def uniform(n0=1.0):
        return _Density(kind=1, n0=n0)

    @staticmethod
    def step(dir, n0, pos):
        return _Density(kind=2, dir=_AXIS[dir], n0=n0, pos=pos)

    @staticmethod
    def slab(dir, n0, begin, end):
        return _Density(kind=3, dir=_AXIS[dir], n0=n0, begin=begin, end=end)

    @staticmethod
    def sphere(n0, center, radius):
        return _Density(kind=4, n0=n0, center=(ctypes.c_float * 2)(*center), radius=radius)

    @staticmethod
    def ramp(dir, n0, begin, end):
        """No reference counterpart: 0 at begin rising linearly to n0 at end."""
        return _Density(kind=5, dir=_AXIS[dir], n0=n0, begin=begin, end=end)


class Species:
    """Species(name, m_q, ppc) with the reference's udist / density setters.
    density: a Density profile (default uniform n0 = 1, or n0 if given)."""

    def __init__(self, name, m_q, ppc, udist="none", uth=(0, 0, 0), ufl=(0, 0, 0), npmin=2, n0=None, seed=0,
                 density=None):
        L = lib()
        h = L.zpic_species_new(name.encode(), m_q, ppc[0], ppc[1])
        if not h:
            raise ZpicError(L.zpic_last_error().decode())
        self.h = h
        self.name = name
        self.sim = None
        if n0 is not None:
            _chk(L.zpic_species_set_density_uniform(h, n0))
        if density is not None:
            self.set_density(density)
        if seed:
            _chk(L.zpic_species_set_seed(h, seed))
        self.set_udist(udist, uth, ufl, npmin)

    def set_density(self, d):
        """Species::set_density (before add_species)."""
        _chk(lib().zpic_species_set_density(self.h, ctypes.byref(d)))

    def set_udist(self, kind, uth=(0, 0, 0), ufl=(0, 0, 0), npmin=2):
        uth = np.asarray(uth, np.float32)
        ufl = np.asarray(ufl, np.float32)
        _chk(lib().zpic_species_set_udist(self.h, UDIST[kind], _p(uth), _p(ufl), npmin))

    def __del__(self):
        if getattr(self, "h", None) and getattr(self, "sim", None) is None:
            lib().zpic_species_free(self.h)
            self.h = None

    def get_energy(self):
        return float(lib().zpic_species_get_energy(self.h))

    def get_nmove(self):
        return int(lib().zpic_species_get_nmove(self.h))

    def np_total(self):
        return int(lib().zpic_species_np_total(self.h))

    def particles(self):
        """Host copy in tile order: np, offset per tile, ix [n,2], x [n,2], u [n,3]."""
        n = self.np_total()
        nt = self.sim.ntiles
        out = {"np": np.zeros(nt, np.int32), "offset": np.zeros(nt, np.int32),
               "ix": np.zeros((n, 2), np.int32), "x": np.zeros((n, 2), np.float32),
               "u": np.zeros((n, 3), np.float32)}
        _chk(lib().zpic_species_get_particles(self.h, _p(out["np"]), _p(out["offset"]), _p(out["ix"]),
                                              _p(out["x"]), _p(out["u"])))
        return out

    def set_particles(self, np_, ix, x, u):
        np_ = np.ascontiguousarray(np_, np.int32)
        ix = np.ascontiguousarray(ix, np.int32)
        x = np.ascontiguousarray(x, np.float32)
        u = np.ascontiguousarray(u, np.float32)
        _chk(lib().zpic_species_set_particles(self.h, _p(np_), _p(ix), _p(x), _p(u)))

    def charge(self):
        """Charge density on the (gc 0,1) tile-major grid after the guard add."""
        n = lib().zpic_species_charge(self.h, None)
        if n < 0:
            raise ZpicError(lib().zpic_last_error().decode())
        out = np.zeros(n, np.float32)
        lib().zpic_species_charge(self.h, _p(out))
        return out

    def gather(self, quant):
        q = {"x": 0, "y": 1, "ux": 2, "uy": 3, "uz": 4}[quant]
        out = np.zeros(self.np_total(), np.float32)
        _chk(lib().zpic_species_gather(self.h, q, _p(out)))
        return out

    # ZDF diagnostics, written under the current working directory
    def set_bc(self, x=("periodic", "periodic"), y=("periodic", "periodic")):
        """Species::set_bc: each side "open", "periodic" or "reflecting"."""
        names = {"open": 0, "periodic": 1, "reflecting": 2}
        b = np.array([names[x[0]], names[x[1]], names[y[0]], names[y[1]]], np.int32)
        _chk(lib().zpic_species_set_bc(self.h, _p(b)))

    def reinject(self):
        """Particles replaced by a device injection of the density profile
        over the whole box (the moving window's kernels), momenta 0."""
        _chk(lib().zpic_species_reinject(self.h))

    def subsample_from(self, src, stride, phase=0):
        """Every stride-th particle of each of src's tiles (same tiling)."""
        _chk(lib().zpic_species_subsample(self.h, src.h, stride, phase))

    def save(self):
        _chk(lib().zpic_species_save(self.h))

    def save_charge(self):
        _chk(lib().zpic_species_save_charge(self.h))

    def save_phasespace(self, quant0, range0, size0, quant1=None, range1=(0.0, 1.0), size1=0):
        q1 = PHASESPACE[quant1] if quant1 else 0
        _chk(lib().zpic_species_save_phasespace(self.h, PHASESPACE[quant0], range0[0], range0[1], size0, q1,
                                                range1[0], range1[1], size1 if quant1 else 0))


class Simulation:
    """Simulation(ntiles, nx, box, dt) (simulation.h:39-43) on the GPU.

    As the reference's Cython Simulation (em2d.pyx:1449-1550), species may
    be given here, and moving_window=True turns the moving window on (after
    the species are added) and makes advance() call advance_mov_window()."""

    def __init__(self, ntiles, nx, box, dt, filter=None, deterministic=False, species=None, moving_window=False):
        L = lib()
        h = L.zpic_simulation_new(ntiles[0], ntiles[1], nx[0], nx[1], float(box[0]), float(box[1]), float(dt))
        if not h:
            raise ZpicError(L.zpic_last_error().decode())
        self.h = h
        self.species = []
        g = Geom()
        _chk(L.zpic_field_geom(h, ctypes.byref(g)))
        self.geom = g
        self.ntiles = g.ntiles_x * g.ntiles_y
        self.mov_window = False
        if filter:
            self.set_filter(*filter)
        if deterministic:
            self.set_deterministic(True)
        for sp in ( [species] if isinstance(species, Species) else (species or []) ):
            self.add_species(sp)
        if moving_window:
            self.set_moving_window()

    def set_moving_window(self):
        """Simulation::set_moving_window (before the first step)."""
        _chk(lib().zpic_simulation_set_moving_window(self.h))
        self.mov_window = True

    def advance_mov_window(self, n=1):
        _chk(lib().zpic_simulation_advance_mov_window(self.h, n))

    def window_moves(self):
        n = ctypes.c_uint(0)
        _chk(lib().zpic_simulation_window_moves(self.h, ctypes.byref(n)))
        return int(n.value)

    def add_laser(self, laser):
        """Simulation.add_laser (em2d.pyx:1521-1536): adds a PlaneWave or
        Gaussian pulse to E and B."""
        laser.add(self)

    def close(self):
        if getattr(self, "h", None):
            lib().zpic_simulation_free(self.h)
            self.h = None
            for s in self.species:
                lib().zpic_species_free(s.h)
                s.h = None

    def __del__(self):
        self.close()

    def add_species(self, sp):
        _chk(lib().zpic_simulation_add_species(self.h, sp.h))
        sp.sim = self
        self.species.append(sp)

    def set_filter(self, kind, direction=0, order=1):
        _chk(lib().zpic_current_set_filter(self.h, FILTER[kind] if isinstance(kind, str) else kind,
                                           direction, order))

    def set_deterministic(self, on=True):
        _chk(lib().zpic_simulation_set_deterministic(self.h, 1 if on else 0))

    def advance(self, n=1):
        if self.mov_window:
            _chk(lib().zpic_simulation_advance_mov_window(self.h, n))
        else:
            _chk(lib().zpic_simulation_advance(self.h, n))

    def stage(self, what, k=0):
        _chk(lib().zpic_simulation_stage(self.h, STAGES[what], k))

    def sync(self):
        _chk(lib().zpic_simulation_sync(self.h))

    def get_iter(self):
        return int(lib().zpic_simulation_get_iter(self.h))

    def get_t(self):
        return float(lib().zpic_simulation_get_t(self.h))

    def get_nmove(self):
        return int(lib().zpic_simulation_get_nmove(self.h))

    def field(self, which):
        g = self.geom
        out = np.zeros((self.ntiles * g.tile_vol, 3), np.float32)
        _chk(lib().zpic_field_get(self.h, "EBJ".index(which), _p(out)))
        return out

    def set_field(self, which, data):
        data = np.ascontiguousarray(data, np.float32)
        assert data.size == self.ntiles * self.geom.tile_vol * 3
        _chk(lib().zpic_field_set(self.h, "EBJ".index(which), _p(data)))

    def emf_energy(self):
        e = np.zeros(3)
        b = np.zeros(3)
        _chk(lib().zpic_emf_get_energy(self.h, _p(e), _p(b)))
        return e, b

    def energies(self):
        """Kinetic energy per species, then E.x E.y E.z B.x B.y B.z."""
        e, b = self.emf_energy()
        return np.array([s.get_energy() for s in self.species] + list(e) + list(b))

    # ZDF diagnostics, written under the current working directory
    def save_emf(self, field, fc):
        _chk(lib().zpic_emf_save(self.h, "EB".index(field), "xyz".index(fc)))

    def save_current(self, fc):
        _chk(lib().zpic_current_save(self.h, "xyz".index(fc)))

    def save_field_file(self, which, fc, filename):
        _chk(lib().zpic_field_save_file(self.h, "EBJ".index(which), "xyz".index(fc), filename.encode()))

    def set_emf_bc(self, x=("periodic", "periodic"), y=("periodic", "periodic")):
        """EMF::set_bc: each side "none", "periodic", "pec" or "pmc", or
        "absorbing" (damping layer, no reference counterpart)."""
        names = {"none": 0, "periodic": 1, "pec": 2, "pmc": 3, "absorbing": 4}
        b = np.array([names[x[0]], names[x[1]], names[y[0]], names[y[1]]], np.int32)
        _chk(lib().zpic_emf_set_bc(self.h, _p(b)))

    def set_absorbing(self, width=20, alpha=0.015):
        """Width (cells) and strength of the absorbing field layers."""
        _chk(lib().zpic_emf_set_absorbing(self.h, int(width), float(alpha)))

    def set_current_bc(self, x=("periodic", "periodic"), y=("periodic", "periodic")):
        """Current::set_bc: each side "none", "periodic" or "reflecting"."""
        names = {"none": 0, "periodic": 1, "reflecting": 2}
        b = np.array([names[x[0]], names[x[1]], names[y[0]], names[y[1]]], np.int32)
        _chk(lib().zpic_current_set_bc(self.h, _p(b)))

    # learned current filter (see Denoiser)
    def set_learned_filter(self, dn):
        _chk(lib().zpic_current_set_learned_filter(self.h, dn.h))
        self._dn = dn  # keep it alive while the filter uses it

    def set_fallback(self, on=True):
        _chk(lib().zpic_current_set_fallback(self.h, 1 if on else 0))

    def get_fallback(self):
        return int(lib().zpic_current_get_fallback(self.h))

    def copy_fields_from(self, other):
        _chk(lib().zpic_simulation_copy_fields(self.h, other.h))

    def add_pair_to(self, dn, target_sim):
        """Stores (this simulation's J, target_sim's J) in dn's pair store."""
        _chk(lib().zpic_simulation_add_pair(dn.h, self.h, target_sim.h))

    def set_timing(self, on=True):
        _chk(lib().zpic_simulation_set_timing(self.h, 1 if on else 0))

    def get_timing(self):
        out = np.zeros(5)
        _chk(lib().zpic_simulation_get_timing(self.h, _p(out)))
        return dict(zip(("push_move", "sort", "current", "yee", "step"), out.tolist()))


class _Laser(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("start", ctypes.c_float), ("fwhm", ctypes.c_float),
                ("rise", ctypes.c_float), ("flat", ctypes.c_float), ("fall", ctypes.c_float),
                ("a0", ctypes.c_float), ("omega0", ctypes.c_float), ("polarization", ctypes.c_float),
                ("cos_pol", ctypes.c_float), ("sin_pol", ctypes.c_float), ("filter", ctypes.c_uint),
                ("W0", ctypes.c_float), ("focus", ctypes.c_float), ("axis", ctypes.c_float)]


class PlaneWave:
    """Laser::PlaneWave (the reference's Cython em2d.laser.PlaneWave)."""
    kind = 0

    def __init__(self, *, start=0.0, fwhm=0.0, rise=0.0, flat=0.0, fall=0.0, a0=0.0, omega0=0.0,
                 polarization=0.0, cos_pol=0.0, sin_pol=0.0, filter=1):
        self.p = dict(start=start, fwhm=fwhm, rise=rise, flat=flat, fall=fall, a0=a0, omega0=omega0,
                      polarization=polarization, cos_pol=cos_pol, sin_pol=sin_pol, filter=filter)

    def _struct(self):
        return _Laser(kind=self.kind, **self.p)

    def add(self, sim):
        """Pulse::add: raises ZpicError for invalid pulse parameters."""
        _chk(lib().zpic_laser_add(sim.h, ctypes.byref(self._struct())))


class Gaussian(PlaneWave):
    """Laser::Gaussian (the reference's Cython em2d.laser.Gaussian)."""
    kind = 1

    def __init__(self, *, W0=0.0, focus=0.0, axis=0.0, **kw):
        super().__init__(**kw)
        self.p.update(W0=W0, focus=focus, axis=axis)


class Batch:
    """Independent simulations advanced together: one launch per stage over
    every instance, and one hipGraph launch per step when graph=True.

    The simulations are held by reference and must stay open while the batch
    lives; each must hold the same number of species."""

    def __init__(self, sims=(), threads_per_tile=256, graph=False, persistent=False):
        L = lib()
        h = L.zpic_batch_new(threads_per_tile)
        if not h:
            raise ZpicError(L.zpic_last_error().decode())
        self.h = h
        self.sims = []
        for s in sims:
            self.add(s)
        self.set_graph(graph)
        self.set_persistent(persistent)

    def set_persistent(self, on=True):
        """One persistent launch per step over every instance (see
        zpic_batch_step) instead of one launch per stage."""
        _chk(lib().zpic_batch_set_persistent(self.h, 1 if on else 0))

    def status(self):
        """Synchronises; raises if a persistent step reported an error."""
        _chk(lib().zpic_batch_status(self.h))

    def add(self, sim, threads_per_tile=0):
        """Adds an instance; threads_per_tile: its particle pass's threads
        (0 = the batch's)."""
        _chk(lib().zpic_batch_add_threads(self.h, sim.h, threads_per_tile))
        self.sims.append(sim)

    def set_graph(self, on=True):
        _chk(lib().zpic_batch_set_graph(self.h, 1 if on else 0))

    def advance(self, n=1):
        _chk(lib().zpic_batch_advance(self.h, n))

    def close(self):
        if getattr(self, "h", None):
            lib().zpic_batch_free(self.h)
            self.h = None

    def __del__(self):
        self.close()


class TuneResult(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_uint), ("ny", ctypes.c_uint), ("threads", ctypes.c_int),
                ("particles_per_ns", ctypes.c_double), ("from_cache", ctypes.c_int)]


def autotune(grid, ppc, steps=3, max_cells=0, cache=None):
    """Fastest tile shape and threads per tile for a grid and ppc (timed on
    the GPU; see zpic_autotune).  Returns a dict; cache: optional file path."""
    r = TuneResult()
    _chk(lib().zpic_autotune(grid[0], grid[1], ppc[0], ppc[1], steps, max_cells,
                             cache.encode() if cache else None, ctypes.byref(r)))
    return {"nx": (r.nx, r.ny), "threads": r.threads, "particles_per_ns": r.particles_per_ns,
            "from_cache": bool(r.from_cache)}


def fitted_threads_per_tile(threads, tile_vol):
    return int(lib().zpic_fitted_threads_per_tile(threads, tile_vol))


class Denoiser:
    """Learned current denoiser for an H x W grid (see zpic_amd.h): 8,835
    FP32 parameters, BF16 matrix-core kernels, training on stored pairs."""

    NPARAM = 8835

    def __init__(self, H, W, max_pairs=0, params=None, scale=None):
        L = lib()
        h = L.zpic_denoiser_new(H, W, max_pairs)
        if not h:
            raise ZpicError(L.zpic_last_error().decode())
        self.h = h
        self.H, self.W = H, W
        if params is not None:
            self.set_params(params)
        if scale is not None:
            self.set_scale(scale)

    def close(self):
        if getattr(self, "h", None):
            lib().zpic_denoiser_free(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.NPARAM
        _chk(lib().zpic_denoiser_set_params(self.h, _p(p)))

    def params(self):
        out = np.zeros(self.NPARAM, np.float32)
        _chk(lib().zpic_denoiser_get_params(self.h, _p(out)))
        return out

    def set_scale(self, s):
        _chk(lib().zpic_denoiser_set_scale(self.h, float(s)))

    @property
    def scale(self):
        return float(lib().zpic_denoiser_get_scale(self.h))

    def forward(self, x):
        """x: [H, W, 3] float32 (physical units) -> network output, same shape."""
        x = np.ascontiguousarray(x, np.float32)
        assert x.shape == (self.H, self.W, 3)
        out = np.zeros_like(x)
        _chk(lib().zpic_denoiser_forward_host(self.h, _p(x), _p(out)))
        return out

    def gradient(self, x, target):
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(target, np.float32)
        g = np.zeros(self.NPARAM, np.float32)
        loss = ctypes.c_double()
        _chk(lib().zpic_denoiser_gradient_host(self.h, _p(x), _p(t), _p(g), ctypes.byref(loss)))
        return loss.value, g

    def num_pairs(self):
        return int(lib().zpic_denoiser_num_pairs(self.h))

    def clear_pairs(self):
        _chk(lib().zpic_denoiser_clear_pairs(self.h))

    def pair(self, k):
        a = np.zeros((self.H, self.W, 3), np.float32)
        b = np.zeros_like(a)
        _chk(lib().zpic_denoiser_get_pair(self.h, k, _p(a), _p(b)))
        return a, b

    def target_rms(self):
        r = ctypes.c_double()
        _chk(lib().zpic_denoiser_target_rms(self.h, ctypes.byref(r)))
        return r.value

    def train(self, steps, lr=1e-3, seed=0, use=0):
        """Adam steps on the first `use` stored pairs (all if 0); returns the
        per-step losses."""
        hist = np.zeros(max(steps, 1), np.float64)
        _chk(lib().zpic_denoiser_train(self.h, steps, float(lr), seed, use, _p(hist)))
        return hist[:steps]

    def eval_pair(self, k):
        a, b = ctypes.c_double(), ctypes.c_double()
        _chk(lib().zpic_denoiser_eval_pair(self.h, k, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def save(self, path):
        np.savez(path, params=self.params(), scale=np.float32(self.scale))

    @classmethod
    def load(cls, path, H, W, max_pairs=0):
        d = np.load(path, allow_pickle=False)
        return cls(H, W, max_pairs, params=d["params"], scale=float(d["scale"]))


def set_threads_per_tile(n):
    """Threads per tile of the fast particle pass (multiple of 64 in [64, 512])."""
    _chk(lib().zpic_set_threads_per_tile(n))


def get_threads_per_tile():
    return int(lib().zpic_get_threads_per_tile())


def set_sort_order(reference):
    """Fast-mode tile sort: True the reference's within-tile order (the
    deterministic mode always keeps it), False (default) order-free."""
    _chk(lib().zpic_set_sort_order(1 if reference else 0))


def get_sort_order():
    return bool(lib().zpic_get_sort_order())


def weibel(ntiles=(8, 8), nx=(16, 16), ppc=(8, 8), box=None, dt=0.07, uth=(0.1, 0.1, 0.1),
           ufl=(0.0, 0.0, 0.6), filter=None, deterministic=False, seed=0):
    """The reference Weibel setup (openmp/em2d/main.cpp:179-259) on the GPU.
    seed != 0 offsets the momentum RNG seeds (an ensemble member)."""
    if box is None:
        box = (np.float32(ntiles[0] * nx[0] / 10.0), np.float32(ntiles[1] * nx[1] / 10.0))
    sim = Simulation(ntiles, nx, box, float(np.float32(dt)), filter=filter, deterministic=deterministic)
    sim.add_species(Species("electrons", -1.0, ppc, "thermal_corr", uth, ufl, seed=16 * seed))
    sim.add_species(Species("positrons", +1.0, ppc, "thermal_corr", uth, tuple(-v for v in ufl), seed=16 * seed))
    return sim


def batch64_instances(n=64, seed=2024):
    """The batched Weibel parameter scan: n instances with the grid drawn from
    {64, 128, 256, 512}^2, ppc from {2, 4, 8}^2, drift uz from U[0.2, 0.9] and
    thermal spread uth from U[0.02, 0.2] (numpy RandomState(seed), drawn in
    that order per instance); dx = dy = 0.1, dt = 0.07, periodic."""
    rng = np.random.RandomState(seed)
    out = []
    for _ in range(n):
        g = int(rng.choice([64, 128, 256, 512]))
        p = int(rng.choice([2, 4, 8]))
        uz = float(rng.uniform(0.2, 0.9))
        uth = float(rng.uniform(0.02, 0.2))
        out.append(dict(grid=(g, g), ppc=(p, p), uth=(uth, uth, uth), ufl=(0.0, 0.0, uz)))
    return out


def tune_instances(instances, steps=2, max_cells=1024 * 1024, cache=None):
    """Autotuned tile shape and threads per tile (zpic_autotune) for a
    parameter scan; adds "nx", "ntiles" and "threads" to each instance (in
    place) and returns them.  The autotuner's samples do not depend on the
    grid beyond which tile shapes divide it, so instances are tuned once per
    ppc and set of dividing shapes."""
    tuned = {}
    for c in instances:
        fits = tuple(tuple(s for s in (8, 16, 32) if n % s == 0) for n in c["grid"])
        key = (c["ppc"], fits)
        if key not in tuned:
            tuned[key] = autotune(c["grid"], c["ppc"], steps=steps, max_cells=max_cells, cache=cache)
        r = tuned[key]
        c["nx"] = tuple(r["nx"])
        c["ntiles"] = (c["grid"][0] // r["nx"][0], c["grid"][1] // r["nx"][1])
        c["threads"] = r["threads"]
    return instances


def scan_simulation(c):
    """One instance of a parameter scan (see batch64_instances) with its tile
    shape c["nx"] (default 16x16)."""
    nx = tuple(c.get("nx", (16, 16)))
    ntiles = tuple(c.get("ntiles", (c["grid"][0] // nx[0], c["grid"][1] // nx[1])))
    return weibel(ntiles=ntiles, nx=nx, ppc=c["ppc"], uth=c["uth"], ufl=c["ufl"])
