# cython: language_level=3
"""Core classes of the em2d front-end (the reference's openmp/python/em2d/
em2d.pyx:1-1558) over the C ABI of libzpic_amd.so: Simulation, Species, EMF,
Current, Vec3Grid (views of a simulation's grids), Grid (charge densities)
and zpic.courant.  Same constructors, properties and methods as the
reference's classes; the plotting helpers (which need matplotlib / visxd)
are not part of this build.  Errors raise RuntimeError with the library's
message."""
import numpy as np

from libc.stdint cimport uint64_t


cdef extern from "zpic_amd.h":
    ctypedef struct zpic_simulation:
        pass
    ctypedef struct zpic_species:
        pass
    ctypedef struct zpic_geom:
        int ntiles_x
        int ntiles_y
        int nx
        int ny
        int gc_lower
        int gc_upper
        int ext_x
        int ext_y
        int tile_vol
        int offset
    ctypedef struct zpic_density:
        int kind
        int dir
        float n0
        float pos
        float begin
        float end
        float center[2]
        float radius
    ctypedef struct zpic_laser:
        int kind
        float start, fwhm, rise, flat, fall
        float a0, omega0
        float polarization, cos_pol, sin_pol
        unsigned filter
        float W0, focus, axis

    const char * zpic_last_error()
    zpic_simulation * zpic_simulation_new(unsigned ntx, unsigned nty, unsigned nx, unsigned ny, float box_x,
                                          float box_y, double dt)
    void zpic_simulation_free(zpic_simulation * sim)
    zpic_species * zpic_species_new(const char * name, float m_q, unsigned ppc_x, unsigned ppc_y)
    void zpic_species_free(zpic_species * sp)
    int zpic_species_set_udist(zpic_species * sp, int kind, const float * uth, const float * ufl, int npmin)
    int zpic_species_set_density(zpic_species * sp, const zpic_density * d)
    int zpic_simulation_add_species(zpic_simulation * sim, zpic_species * sp)
    int zpic_simulation_advance(zpic_simulation * sim, int n)
    int zpic_simulation_advance_mov_window(zpic_simulation * sim, int n)
    int zpic_simulation_set_moving_window(zpic_simulation * sim)
    int zpic_simulation_set_deterministic(zpic_simulation * sim, int on)
    int zpic_simulation_stage(zpic_simulation * sim, int stage, int species)
    unsigned zpic_simulation_get_iter(const zpic_simulation * sim)
    double zpic_simulation_get_t(const zpic_simulation * sim)
    uint64_t zpic_simulation_get_nmove(const zpic_simulation * sim)
    int zpic_simulation_info(const zpic_simulation * sim, unsigned * ntiles, unsigned * nx, float * box, double * dt)
    int zpic_simulation_energy_info(zpic_simulation * sim)
    int zpic_emf_get_iter(const zpic_simulation * sim)
    int zpic_current_get_iter(const zpic_simulation * sim)
    int zpic_current_set_filter(zpic_simulation * sim, int kind, int dir, int order)
    int zpic_emf_get_energy(zpic_simulation * sim, double * ene_E, double * ene_B)
    int zpic_emf_save(zpic_simulation * sim, int field, int fc)
    int zpic_current_save(zpic_simulation * sim, int fc)
    int zpic_field_save_file(zpic_simulation * sim, int which, int fc, const char * filename)
    int zpic_field_get(zpic_simulation * sim, int which, float * out)
    int zpic_field_set(zpic_simulation * sim, int which, const float * data)
    int zpic_field_device(zpic_simulation * sim, int which, float ** buf, zpic_geom * g, int * periodic)
    int zpic_laser_add(zpic_simulation * sim, const zpic_laser * laser)
    double zpic_species_get_energy(zpic_species * sp)
    uint64_t zpic_species_get_nmove(zpic_species * sp)
    uint64_t zpic_species_np_total(zpic_species * sp)
    int zpic_species_gather(zpic_species * sp, int quant, float * out)
    long long zpic_species_charge(zpic_species * sp, float * out)
    int zpic_species_get_phasespace(zpic_species * sp, int q0, float min0, float max0, int size0, int q1,
                                    float min1, float max1, int size1, float * out)
    int zpic_species_save(zpic_species * sp)
    int zpic_species_save_charge(zpic_species * sp)
    int zpic_species_save_phasespace(zpic_species * sp, int q0, float min0, float max0, int size0, int q1,
                                     float min1, float max1, int size1)
    int zpic_fill(float * buf, size_t n, float value, void * stream)
    int zpic_grid_add(float * dst, const float * src, size_t n, void * stream)
    int zpic_copy_to_gc(float * buf, const zpic_geom * g, int ncomp, int px, int py, void * stream)
    int zpic_add_from_gc(float * buf, const zpic_geom * g, int ncomp, int px, int py, void * stream)
    int zpic_kernel3(float * buf, const zpic_geom * g, int dir, float a, float b, float c, int px, int py,
                     void * stream)
    int zpic_x_shift_left(float * buf, const zpic_geom * g, int ncomp, int shift, int px, int py, void * stream)


cdef int _chk(int rc) except -1:
    if rc != 0:
        raise RuntimeError(zpic_last_error().decode())
    return 0


_FC = {'x': 0, 'y': 1, 'z': 2}
_QUANT = {'x': 0, 'y': 1, 'ux': 2, 'uy': 3, 'uz': 4}


class zpic:
    """ZPIC utilities (zpic.h:64-91)."""

    @staticmethod
    def sys_info():
        """sys_info(): prints the build this front-end runs on."""
        print("em2d on libzpic_amd.so: HIP kernels for AMD Instinct MI355X (gfx950), FP32")

    @staticmethod
    def courant(list dx=None, list dims=None, list ntiles=None, list nx=None, list box=None):
        """courant( dx, dims, ntiles, nx, box ): CFL time-step limit
        sqrt(1 / (1/dx^2 + 1/dy^2)) in single precision, from the cell size,
        or the grid size and box, or tiles, tile size and box."""
        f = np.float32
        if dx is None:
            if ntiles is None:
                d = (f(box[0]) / f(dims[0]), f(box[1]) / f(dims[1]))
            else:
                d = (f(box[0]) / f(nx[0] * ntiles[0]), f(box[1]) / f(nx[1] * ntiles[1]))
        else:
            d = (f(dx[0]), f(dx[1]))
        one = f(1.0)
        return float(np.sqrt(one / (one / (d[0] * d[0]) + one / (d[1] * d[1]))))


cdef class _Handle:
    """Owner of a C simulation handle (kept alive by every view of it)."""
    cdef zpic_simulation * sim
    cdef object species

    def __cinit__(self):
        self.sim = NULL
        self.species = []

    def __dealloc__(self):
        if self.sim != NULL:
            zpic_simulation_free(self.sim)
            self.sim = NULL


cdef _Handle _new_handle(list ntiles, list nx, list box, double dt):
    cdef _Handle h = _Handle()
    h.sim = zpic_simulation_new(ntiles[0], ntiles[1], nx[0], nx[1], box[0], box[1], dt)
    if h.sim == NULL:
        raise RuntimeError(zpic_last_error().decode())
    return h


cdef class Vec3Grid:
    """Vec3Grid: view of a float3 tiled grid of a simulation (E, B or J).
    Same methods as the reference's Vec3Grid (em2d.pyx:280-500); the grid
    itself stays in GPU memory."""
    cdef _Handle h
    cdef int which

    def __cinit__(self):
        self.h = None
        self.which = 0

    cdef void _bind(self, _Handle h, int which):
        self.h = h
        self.which = which

    cdef float * _dev(self, zpic_geom * g, int * per) except NULL:
        cdef float * buf = NULL
        _chk(zpic_field_device(self.h.sim, self.which, &buf, g, per))
        return buf

    def buffer_size(self):
        """Grid buffer size (elements, all tiles with guard cells)."""
        cdef zpic_geom g
        cdef int per[2]
        self._dev(&g, per)
        return g.tile_vol * g.ntiles_x * g.ntiles_y

    def zero(self):
        cdef zpic_geom g
        cdef int per[2]
        cdef float * b = self._dev(&g, per)
        _chk(zpic_fill(b, 3 * <size_t> (g.tile_vol * g.ntiles_x * g.ntiles_y), 0.0, NULL))

    def set(self, val):
        """set( val ): every element set to the float3 value val."""
        n = self.buffer_size()
        a = np.empty((n, 3), np.float32)
        a[:] = np.asarray(val, np.float32).reshape(1, 3)
        cdef float[:, ::1] m = a
        _chk(zpic_field_set(self.h.sim, self.which, &m[0, 0]))

    def add(self, Vec3Grid rhs):
        """add( rhs ): element-wise sum with another grid of the same shape."""
        cdef zpic_geom g, g2
        cdef int per[2]
        cdef float * b = self._dev(&g, per)
        cdef float * r = rhs._dev(&g2, per)
        if g.tile_vol * g.ntiles_x * g.ntiles_y != g2.tile_vol * g2.ntiles_x * g2.ntiles_y:
            raise RuntimeError("grids differ in size")
        _chk(zpic_grid_add(b, r, 3 * <size_t> (g.tile_vol * g.ntiles_x * g.ntiles_y), NULL))

    def _tiles(self):
        cdef zpic_geom g
        cdef int per[2]
        self._dev(&g, per)
        a = np.empty((g.ntiles_x * g.ntiles_y * g.tile_vol, 3), np.float32)
        cdef float[:, ::1] m = a
        _chk(zpic_field_get(self.h.sim, self.which, &m[0, 0]))
        return a, (g.ntiles_x, g.ntiles_y, g.nx, g.ny, g.gc_lower, g.ext_x, g.tile_vol)

    def gather(self, str fc):
        """gather( fc ): component 'x', 'y' or 'z' of the interior cells as a
        contiguous [ny, nx] array."""
        a, (ntx, nty, nx, ny, gl, ext, vol) = self._tiles()
        return _gather(a[:, _FC[fc]], ntx, nty, nx, ny, gl, ext, vol)

    def copy_to_gc(self):
        cdef zpic_geom g
        cdef int per[2]
        cdef float * b = self._dev(&g, per)
        _chk(zpic_copy_to_gc(b, &g, 3, per[0], per[1], NULL))

    def add_from_gc(self):
        cdef zpic_geom g
        cdef int per[2]
        cdef float * b = self._dev(&g, per)
        _chk(zpic_add_from_gc(b, &g, 3, per[0], per[1], NULL))

    def x_shift_left(self, unsigned int shift):
        cdef zpic_geom g
        cdef int per[2]
        cdef float * b = self._dev(&g, per)
        _chk(zpic_x_shift_left(b, &g, 3, shift, per[0], per[1], NULL))

    def kernel_x(self, float a, float b, float c):
        """kernel_x( a, b, c ): one [a, b, c] convolution pass along x."""
        cdef zpic_geom g
        cdef int per[2]
        cdef float * d = self._dev(&g, per)
        _chk(zpic_kernel3(d, &g, 0, a, b, c, per[0], per[1], NULL))

    def kernel_y(self, float a, float b, float c):
        cdef zpic_geom g
        cdef int per[2]
        cdef float * d = self._dev(&g, per)
        _chk(zpic_kernel3(d, &g, 1, a, b, c, per[0], per[1], NULL))

    def save(self, str fc, str filename):
        """save( fc, filename ): component written to a ZDF file."""
        _chk(zpic_field_save_file(self.h.sim, self.which, _FC[fc], filename.encode()))


def _gather(comp, ntx, nty, nx, ny, gl, ext, vol):
    """Interior cells of a tile-major component array as a contiguous
    [nty * ny, ntx * nx] array (Vec3Grid::gather, vec3grid.h:76-158)."""
    t = np.asarray(comp).reshape(nty, ntx, vol)
    rows = np.stack([t[:, :, (gl + j) * ext + gl:(gl + j) * ext + gl + nx] for j in range(ny)], axis=2)
    return np.ascontiguousarray(rows.transpose(0, 2, 1, 3).reshape(nty * ny, ntx * nx), dtype=np.float32)


class Grid:
    """Grid: a scalar tiled grid held on the host (charge densities from
    Species.get_charge); gather() gives the contiguous interior."""

    def __init__(self, data, ntiles, nx, gc):
        self.data = data
        self.ntiles = tuple(ntiles)
        self.nx = tuple(nx)
        self.gc = gc
        ext = gc[0][0] + nx[0] + gc[0][1]
        self._ext = ext
        self._vol = (ext * (gc[1][0] + nx[1] + gc[1][1]) + 3) & ~3

    def buffer_size(self):
        return self.data.size

    def gather(self):
        return _gather(self.data, self.ntiles[0], self.ntiles[1], self.nx[0], self.nx[1], self.gc[0][0],
                       self._ext, self._vol)


cdef class Current:
    """Current( ntiles = None, nx = None, box = None, dt = 0 ): current
    density; a view of a simulation's current when made by Simulation."""
    cdef _Handle h
    cdef Vec3Grid _J

    def __cinit__(self, list ntiles=None, list nx=None, list box=None, double dt=0):
        self._J = Vec3Grid()
        self.h = None
        if ntiles is not None:
            self._bind(_new_handle(ntiles, nx, box, dt))

    cdef void _bind(self, _Handle h):
        self.h = h
        self._J._bind(h, 2)

    @property
    def J(self):
        """Vec3Grid with the current density"""
        return self._J

    @property
    def n(self):
        """Iteration number"""
        return zpic_current_get_iter(self.h.sim)

    def set_filter(self, filter):
        """set_filter( filter ): em2d.filter.None, Binomial or Compensated."""
        _chk(zpic_current_set_filter(self.h.sim, filter.kind, filter.dir, filter.order))

    def save(self, fc):
        """save( fc ): component 'x', 'y' or 'z' to CURRENT/J<c>-<n>.zdf"""
        _chk(zpic_current_save(self.h.sim, _FC[fc]))

    def advance(self):
        """advance(): guard exchange, boundaries and filter (Current::advance)."""
        _chk(zpic_simulation_stage(self.h.sim, 5, 0))

    def zero(self):
        _chk(zpic_simulation_stage(self.h.sim, 4, 0))


cdef class EMF:
    """EMF( ntiles = None, nx = None, box = None, dt = 0 ): electromagnetic
    fields; a view of a simulation's fields when made by Simulation."""
    cdef _Handle h
    cdef Vec3Grid _E
    cdef Vec3Grid _B

    def __cinit__(self, list ntiles=None, list nx=None, list box=None, double dt=0):
        self._E = Vec3Grid()
        self._B = Vec3Grid()
        self.h = None
        if ntiles is not None:
            self._bind(_new_handle(ntiles, nx, box, dt))

    cdef void _bind(self, _Handle h):
        self.h = h
        self._E._bind(h, 0)
        self._B._bind(h, 1)

    @property
    def E(self):
        return self._E

    @property
    def B(self):
        return self._B

    @property
    def n(self):
        return zpic_emf_get_iter(self.h.sim)

    @property
    def box(self):
        cdef float b[2]
        _chk(zpic_simulation_info(self.h.sim, NULL, NULL, b, NULL))
        return [b[0], b[1]]

    def save(self, fld, fc):
        """save( fld, fc ): 'E' or 'B', component 'x', 'y' or 'z'."""
        _chk(zpic_emf_save(self.h.sim, {'E': 0, 'B': 1}[fld], _FC[fc]))

    def advance(self, Current current=None):
        """advance( current = None ): one Yee step, with the current density
        of `current` (which must be this simulation's) or without current."""
        if current is None:
            _chk(zpic_simulation_stage(self.h.sim, 11, 0))
        else:
            if current.h is not self.h:
                raise RuntimeError("EMF.advance(current): the current must belong to the same simulation")
            _chk(zpic_simulation_stage(self.h.sim, 6, 0))

    def get_energy(self):
        """Field energy per component: ([Ex, Ey, Ez], [Bx, By, Bz])."""
        cdef double e[3]
        cdef double b[3]
        _chk(zpic_emf_get_energy(self.h.sim, e, b))
        return [e[0], e[1], e[2]], [b[0], b[1], b[2]]

    def _add_laser(self, pulse):
        cdef zpic_laser l
        l.kind = pulse.kind
        l.start = pulse.start
        l.fwhm = pulse.fwhm
        l.rise = pulse.rise
        l.flat = pulse.flat
        l.fall = pulse.fall
        l.a0 = pulse.a0
        l.omega0 = pulse.omega0
        l.polarization = pulse.polarization
        l.cos_pol = pulse.cos_pol
        l.sin_pol = pulse.sin_pol
        l.filter = pulse.filter
        l.W0 = pulse.W0
        l.focus = pulse.focus
        l.axis = pulse.axis
        _chk(zpic_laser_add(self.h.sim, &l))


cdef class Species:
    """Species( name, m_q, ppc, *, udist = None, density = None )"""
    cdef zpic_species * sp
    cdef object _name
    cdef object _grid   # (ntiles, nx) once added to a simulation

    def __cinit__(self, str name=None, float m_q=0, list ppc=None, *, udist=None, density=None):
        self.sp = NULL
        self._grid = None
        self._name = name
        if ppc is not None:
            self.sp = zpic_species_new(name.encode('utf-8'), m_q, ppc[0], ppc[1])
            if self.sp == NULL:
                raise RuntimeError(zpic_last_error().decode())
            if udist is not None:
                self.set_udist(udist)
            if density is not None:
                self.set_density(density)

    def __dealloc__(self):
        # a species added to a simulation is referenced by the simulation's
        # handle, so it is released only after the simulation itself
        if self.sp != NULL:
            zpic_species_free(self.sp)
        self.sp = NULL

    @property
    def name(self):
        return self._name

    def set_udist(self, udist):
        """set_udist( udist ): em2d.udist.None, Cold, Thermal or ThermalCorr."""
        cdef float th[3]
        cdef float fl[3]
        for i in range(3):
            th[i] = udist.uth[i]
            fl[i] = udist.ufl[i]
        _chk(zpic_species_set_udist(self.sp, udist.kind, th, fl, udist.npmin))

    def set_density(self, density):
        """set_density( density ): em2d.density.None, Uniform, Step, Slab,
        Sphere or Ramp."""
        cdef zpic_density d
        d.kind = density.kind
        d.dir = density.dir
        d.n0 = density.n0
        d.pos = density.pos
        d.begin = density.begin
        d.end = density.end
        d.center[0] = density.center[0]
        d.center[1] = density.center[1]
        d.radius = density.radius
        _chk(zpic_species_set_density(self.sp, &d))

    def save(self):
        _chk(zpic_species_save(self.sp))

    def save_charge(self):
        _chk(zpic_species_save_charge(self.sp))

    def np_total(self):
        return zpic_species_np_total(self.sp)

    def get_energy(self):
        return zpic_species_get_energy(self.sp)

    def gather(self, q):
        """gather( q ): 'x', 'y', 'ux', 'uy' or 'uz' of every particle."""
        a = np.zeros(zpic_species_np_total(self.sp), np.float32)
        if a.size == 0:
            return a
        cdef float[::1] m = a
        _chk(zpic_species_gather(self.sp, _QUANT[q], &m[0]))
        return a

    def get_charge(self):
        """get_charge(): charge density (Grid with guard cells [[0,1],[0,1]])."""
        n = zpic_species_charge(self.sp, NULL)
        if n < 0:
            raise RuntimeError(zpic_last_error().decode())
        a = np.zeros(n, np.float32)
        cdef float[::1] m = a
        zpic_species_charge(self.sp, &m[0])
        nt, nx = self._grid
        return Grid(a, list(nt), list(nx), [[0, 1], [0, 1]])

    def get_phasespace(self, quant0, range0, size0, quant1=None, range1=None, size1=None):
        """get_phasespace( quant0, range0, size0, quant1 = None, range1 = None,
        size1 = None ): 1-D [size0] or 2-D [size1, size0] phase-space density."""
        if quant1 is None:
            pha = np.empty(size0, np.float32)
            q1, r1, s1 = -1, (0.0, 0.0), 0
        else:
            if quant1 == quant0:
                raise RuntimeError("Cannot generate a 2d phasespace with the same quantity on both axis")
            pha = np.empty((size1, size0), np.float32)
            q1, r1, s1 = _QUANT[quant1], range1, size1
        cdef float[::1] m = pha.reshape(-1)
        _chk(zpic_species_get_phasespace(self.sp, _QUANT[quant0], range0[0], range0[1], size0, q1, r1[0], r1[1],
                                         s1, &m[0]))
        return pha

    def save_phasespace(self, quant0, range0, size0, quant1=None, range1=None, size1=None):
        if quant1 is None:
            _chk(zpic_species_save_phasespace(self.sp, _QUANT[quant0], range0[0], range0[1], size0, 0, 0.0, 0.0, 0))
        else:
            _chk(zpic_species_save_phasespace(self.sp, _QUANT[quant0], range0[0], range0[1], size0, _QUANT[quant1],
                                              range1[0], range1[1], size1))


cdef class Simulation:
    """Simulation( ntiles, nx, box, dt, *, species = None, moving_window = False )"""
    cdef _Handle h
    cdef EMF _emf
    cdef Current _current
    cdef object mov_window

    def __cinit__(self, list ntiles, list nx, list box, double dt, *, species=None, moving_window=False):
        self.h = _new_handle(ntiles, nx, box, dt)
        self._emf = EMF()
        self._emf._bind(self.h)
        self._current = Current()
        self._current._bind(self.h)
        if isinstance(species, Species):
            self._add_species(species)
        elif isinstance(species, (list, tuple)):
            for s in species:
                self._add_species(s)
        self.mov_window = bool(moving_window)
        if self.mov_window:
            _chk(zpic_simulation_set_moving_window(self.h.sim))

    cdef _add_species(self, Species s):
        _chk(zpic_simulation_add_species(self.h.sim, s.sp))
        cdef unsigned nt[2]
        cdef unsigned nx[2]
        _chk(zpic_simulation_info(self.h.sim, nt, nx, NULL, NULL))
        s._grid = ((nt[0], nt[1]), (nx[0], nx[1]))
        self.h.species.append(s)

    @property
    def t(self):
        return zpic_simulation_get_t(self.h.sim)

    @property
    def n(self):
        return zpic_simulation_get_iter(self.h.sim)

    @property
    def ntiles(self):
        cdef unsigned nt[2]
        _chk(zpic_simulation_info(self.h.sim, nt, NULL, NULL, NULL))
        return (nt[0], nt[1])

    @property
    def box(self):
        cdef float b[2]
        _chk(zpic_simulation_info(self.h.sim, NULL, NULL, b, NULL))
        return (b[0], b[1])

    @property
    def dt(self):
        cdef double d
        _chk(zpic_simulation_info(self.h.sim, NULL, NULL, NULL, &d))
        return d

    @property
    def emf(self):
        return self._emf

    @property
    def current(self):
        return self._current

    @property
    def species(self):
        return list(self.h.species)

    def add(self, src):
        """add( src ): adds a laser pulse (em2d.laser.PlaneWave / Gaussian)."""
        if hasattr(src, "omega0"):
            src.add(self._emf)
        else:
            raise RuntimeError("Invalid src object type")

    def advance(self):
        """advance(): one step (advance_mov_window with a moving window)."""
        if self.mov_window:
            _chk(zpic_simulation_advance_mov_window(self.h.sim, 1))
        else:
            _chk(zpic_simulation_advance(self.h.sim, 1))

    def energy_info(self):
        _chk(zpic_simulation_energy_info(self.h.sim))

    def get_nmove(self):
        return zpic_simulation_get_nmove(self.h.sim)

    def set_deterministic(self, on=True):
        """Reference single-thread order for the deposit and the sums (no
        reference counterpart: bit-identical results, slower)."""
        _chk(zpic_simulation_set_deterministic(self.h.sim, 1 if on else 0))
