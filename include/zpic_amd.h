/*
 * zpic_amd.h -- C ABI of the MI355X em2d particle-in-cell timestep.
 *
 * Two layers, both plain C (pointers, sizes, PODs; no HIP or torch types):
 *
 *  1. Stage entry points on device buffers.  Each one is one stage of the
 *     reference's Simulation::advance() and takes an opaque HIP stream
 *     (`void *stream`, NULL = the library's default stream).  They return 0 on
 *     success and a negative ZPIC_E* code otherwise; zpic_last_error() gives
 *     the message.  Work is enqueued only; nothing synchronises the host.
 *
 *  2. Object entry points mirroring the reference C++ classes
 *     (Simulation / Species / EMF / Current), i.e. exactly what the reference's
 *     own FFI (the Cython declarations openmp/python/em2d/simulation.pxd,
 *     species.pxd, emf.pxd, current.pxd) binds.  They own device memory and one
 *     HIP stream per simulation.
 *
 * Layouts (the reference's, openmp/em2d/particles.h:82-105, grid.h:77-92):
 *   particles  tile-major SoA: ix int2[max_part] (cell inside the tile),
 *              x float2[max_part] (offset in [-0.5,0.5) from the cell centre),
 *              u float3[max_part] (proper velocity u = gamma v / c);
 *              tile t owns [offset[t], offset[t] + np[t]).
 *   fields     tile-major float3 with guard cells: tile t occupies elements
 *              [t*tile_vol, (t+1)*tile_vol); cell (i,j) of tile t is element
 *              t*tile_vol + offset + j*ext_x + i, i in [-gc_lower, nx+gc_upper).
 */
#ifndef ZPIC_AMD_H_
#define ZPIC_AMD_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ codes */
#define ZPIC_OK 0
#define ZPIC_EINVAL (-1)   /* invalid argument / configuration             */
#define ZPIC_EHIP (-2)     /* HIP runtime error                             */
#define ZPIC_ERANGE (-3)   /* size beyond a buffer or a stated maximum      */
#define ZPIC_ESTATE (-4)   /* call not valid in the object's current state  */

/* Last error message of the calling thread ("" if none). */
const char *zpic_last_error(void);

/* ------------------------------------------------------------- geometry */
/* Tiled grid geometry (grid<T> fields, grid.h:47-92). */
typedef struct {
    int ntiles_x, ntiles_y; /* number of tiles                               */
    int nx, ny;             /* cells per tile                                */
    int gc_lower, gc_upper; /* guard cells below / above (same in x and y)   */
    int ext_x, ext_y;       /* tile size including guard cells               */
    int tile_vol;           /* elements per tile: ext_x*ext_y rounded to 4   */
    int offset;             /* element of cell (0,0): gc_lower*ext_x+gc_lower */
} zpic_geom;

/* Fills a geometry from tile count, tile size and guard cells. */
int zpic_geom_init(zpic_geom *g, int ntiles_x, int ntiles_y, int nx, int ny, int gc_lower, int gc_upper);

/* Device particle buffer (Particles / ParticleData, particles.h:82-206). */
typedef struct {
    int *np;        /* [ntiles] particles per tile                */
    int *offset;    /* [ntiles] first particle of each tile       */
    int *ix;        /* int2[max_part]                             */
    float *x;       /* float2[max_part]                           */
    float *u;       /* float3[max_part]                           */
    uint32_t max_part;
} zpic_particles;

/* Device scratch of the tile sort (ParticleSort, particles.h:20-65): leaver
 * lists, per-tile direction counts, new counts, per-wave segment counts and
 * the staged-particle counts, with the sizes given per member. */
typedef struct {
    int *idx;       /* [max_part]   leaver list (tile-local index << 4 | direction) */
    int *npt;       /* [9*ntiles]   particles per direction (4 = staying)           */
    int *new_np;    /* [ntiles]     particles per tile after the sort               */
    int *seg;       /* [16*ntiles]  leavers found by each wave segment              */
    int *tmp_np;    /* [ntiles]     staged particles per tile                       */
    int *total;     /* [1]          particles after the sort                        */
    unsigned long long *nmove; /* [1] running sum of the particles each sort's
                                  pass moved (Species::get_nmove); may be NULL */
    /* Order-free sort of the fast mode (may be NULL): buffers of the
     * particle buffer's size (int2, float2, float3 per particle; the
     * staging buffer `tmp` serves).  When set, a fast-mode particle pass
     * compacts each wave segment's stayers to the segment's front and writes
     * the segment's leaver records (cell in the source tile's frame) to these
     * buffers at the same segment positions, and writes no leaver list; the
     * next zpic_tile_sort / zpic_tile_sort_copy mode 0 places particles
     * without keeping the reference's within-tile order; counts, offsets and
     * the per-tile multisets are the reference's. */
    int *stage_ix;
    float *stage_x;
    float *stage_u;
} zpic_sort;

/* ----------------------------------------------- stage entry points (1) */

/* Species::push (species.cpp:2302-2335, push_kernel :1626-1682): linear E/B
 * interpolation + relativistic Boris push of every particle; writes each
 * tile's sum of u^2/(gamma+1) (double) into tile_energy[ntiles]. */
int zpic_push(const zpic_particles *p, const zpic_geom *fld, const float *E, const float *B,
              float alpha, double *tile_energy, int deterministic, void *stream);

/* Species::move(J) (species.cpp:2138-2164, move_deposit_kernel :1400-1493):
 * position advance, split-segment charge-conserving current deposit into the
 * tile's J region (accumulated on top of J), cell update, and classification
 * of tile leavers into `s` for the following zpic_tile_sort. */
int zpic_move_deposit(const zpic_particles *p, const zpic_geom *fld, float *J,
                      float dt_dx, float dt_dy, float q, float qnx, float qny,
                      zpic_sort *s, int deterministic, void *stream);

/* push + move of Species::advance(emf, current) (species.cpp:2051-2068) fused
 * into one pass over the particles (same results as zpic_push then
 * zpic_move_deposit). */
int zpic_push_move_deposit(const zpic_particles *p, const zpic_geom *fld, const float *E,
                           const float *B, float *J, float alpha, float dt_dx, float dt_dy,
                           float q, float qnx, float qny, double *tile_energy, zpic_sort *s,
                           int deterministic, void *stream);

/* As zpic_push_move_deposit, then Species::process_bc (species.cpp:1854-1989)
 * on the moved particles: reflect is a mask of reflecting walls (1 x lower,
 * 2 x upper, 4 y lower, 8 y upper); particles crossing an open wall are
 * dropped by the next zpic_tile_sort with that axis not periodic. */
int zpic_push_move_deposit_bc(const zpic_particles *p, const zpic_geom *fld, const float *E, const float *B,
                              float *J, float alpha, float dt_dx, float dt_dy, float q, float qnx, float qny,
                              double *tile_energy, zpic_sort *s, int reflect, int deterministic, void *stream);

/* As zpic_push_move_deposit_bc with the cell shift of
 * move_deposit_shift_kernel (species.cpp:1508-1597): (shift_x, shift_y) is
 * added to every particle's new cell before the boundary processing.  The
 * moving window of Species::advance_mov_window (species.cpp:2083-2128)
 * shifts by (-1, 0) on the steps where the window moves. */
int zpic_push_move_deposit_shift(const zpic_particles *p, const zpic_geom *fld, const float *E, const float *B,
                                 float *J, float alpha, float dt_dx, float dt_dy, float q, float qnx, float qny,
                                 double *tile_energy, zpic_sort *s, int reflect, int shift_x, int shift_y,
                                 int deterministic, void *stream);
/* No reference counterpart: zpic_push_move_deposit_shift in the fast mode
 * with each tile's region of J (guard cells included) written instead of
 * added to, for the first species of a step in place of zeroing J first
 * (Simulation::advance: current.zero(), then the species' advance); the
 * result is the same, and a J fill and a J read are saved.  The tile
 * padding of J is left as it is. */
int zpic_push_move_deposit_overwrite(const zpic_particles *p, const zpic_geom *fld, const float *E, const float *B,
                                     float *J, float alpha, float dt_dx, float dt_dy, float q, float qnx, float qny,
                                     double *tile_energy, zpic_sort *s, int reflect, int shift_x, int shift_y,
                                     void *stream);

/* Particles::tile_sort (particles.cpp:855-904: bnd_check, update_tile_info,
 * copy_out, copy_in) as per-tile counting + prefix scan.  Uses the
 * classification left in `s` by zpic_move_deposit / zpic_push_move_deposit.
 * `tmp` supplies the staging buffers and receives the new offsets.  The
 * resulting within-tile order is the reference's single-thread order. */
int zpic_tile_sort(zpic_particles *p, zpic_particles *tmp, zpic_sort *s, int ntiles_x, int ntiles_y,
                   int nx, int ny, int periodic_x, int periodic_y, void *stream);

/* Particles::tile_sort(tmp, sort, extra) (particles.cpp:855-904), the sort
 * that leaves room for particles injected next, in two calls so that the
 * caller can pick the path from the total between them:
 *  zpic_tile_sort_count: new particle counts, and new offsets in tmp->offset
 *    that reserve extra[t] free slots after tile t's particles
 *    (update_tile_info with extra, :336-357; extra may be NULL); the total,
 *    reserved slots included, goes to s->total[0].
 *  zpic_tile_sort_copy, mode 0: the particles move inside p (copy_out +
 *    copy_in, :379-614; zpic_tile_sort is count + copy mode 0).  Mode 1: every
 *    particle is copied into tmp at the new offsets and tmp->np is set
 *    (copy_sorted, :628-763), the path the reference takes when the total
 *    exceeds p->max_part; the caller then swaps the two buffers
 *    (swap_buffers, particles.h:195-206).  In mode 1 tmp must hold at least
 *    s->total[0] particles, and s->stage_ix must be NULL (mode 1 reads the
 *    particles in place; a particle pass that staged leavers is answered
 *    with ZPIC_EINVAL). */
int zpic_tile_sort_count(zpic_particles *p, zpic_particles *tmp, zpic_sort *s, int ntiles_x, int ntiles_y,
                         int nx, int ny, int periodic_x, int periodic_y, const int *extra, void *stream);
int zpic_tile_sort_copy(zpic_particles *p, zpic_particles *tmp, zpic_sort *s, int ntiles_x, int ntiles_y,
                        int nx, int ny, int periodic_x, int periodic_y, int mode, void *stream);

/* Density profiles of particle injection (Density::Uniform / Step / Slab /
 * Sphere, density.h:30-132) as plain data, lengths relative to the window
 * origin: a step keeps the lattice points past `start` (in cells) along
 * `dir`, a slab those in [start, end), a sphere those within `radius` of the
 * centre (simulation units, cell size dx, dy).  A ramp (no reference
 * counterpart) rises linearly from 0 at `start` to full density at `end`
 * (cells along dir) by keeping a fraction of each cell's lattice points. */
#define ZPIC_DENSITY_NONE 0
#define ZPIC_DENSITY_UNIFORM 1
#define ZPIC_DENSITY_STEP 2
#define ZPIC_DENSITY_SLAB 3
#define ZPIC_DENSITY_SPHERE 4
#define ZPIC_DENSITY_RAMP 5   /* no reference counterpart: linear ramp      */
typedef struct {
    int kind;                          /* ZPIC_DENSITY_*                      */
    int dir;                           /* step / slab axis: 0 = x, 1 = y      */
    float start, end;                  /* step / slab bounds, in cells        */
    float center_x, center_y, radius;  /* sphere                              */
    float dx, dy;                      /* cell size                           */
} zpic_inject_profile;

/* Density::*::np_inject (density.cpp:231-298, 444-573, 720-854, 969-1067):
 * number of particles the profile places in each tile's part of the global
 * cell range [range[0], range[1]] x [range[2], range[3]] (ppc_x * ppc_y
 * lattice points per cell), into np[ntiles]. */
int zpic_np_inject(const zpic_inject_profile *d, int ntiles_x, int ntiles_y, int nx, int ny, int ppc_x,
                   int ppc_y, const int range[4], int *np, void *stream);

/* Density::*::inject (density.cpp:103-218, 315-442, 575-718, 856-967): those
 * particles (u = 0) appended after each tile's p->np[t] particles, in the
 * reference's order, and p->np[t] increased.  The slots must be free (the
 * extra of zpic_tile_sort_count). */
int zpic_inject(const zpic_inject_profile *d, const zpic_particles *p, int ntiles_x, int ntiles_y, int nx,
                int ny, int ppc_x, int ppc_y, const int range[4], void *stream);

/* grid<T>::add_from_gc / copy_to_gc (grid.h:253-445) on a tile-major grid
 * of `ncomp` floats per element (3 for vec3grid<float3>, 1 for grid<float>).
 * copy_to_gc (the x pass then the y pass) runs as one launch. */
int zpic_add_from_gc(float *buf, const zpic_geom *g, int ncomp, int periodic_x, int periodic_y, void *stream);
int zpic_copy_to_gc(float *buf, const zpic_geom *g, int ncomp, int periodic_x, int periodic_y, void *stream);

/* copy_to_gc of two grids of the same geometry (E and B after the Yee
 * update, emf.cpp:569-570) in one launch. */
int zpic_copy_to_gc2(float *buf0, float *buf1, const zpic_geom *g, int ncomp, int periodic_x, int periodic_y,
                     void *stream);

/* add_from_gc followed by copy_to_gc (Current::advance, current.cpp:230-231):
 * the two add passes, then the copy in one launch. */
int zpic_add_copy_gc(float *buf, const zpic_geom *g, int ncomp, int periodic_x, int periodic_y, void *stream);

/* grid<T>::kernel3_x / kernel3_y (grid.h:495-593) on vec3 grids: one pass of
 * the 3-point convolution [a,b,c] along dir (0 = x, 1 = y), then the guard
 * copy along that direction. */
int zpic_kernel3(float *buf, const zpic_geom *g, int dir, float a, float b, float c,
                 int periodic_x, int periodic_y, void *stream);

/* `passes` kernel3_x passes (coef[3p..3p+2] = a, b, c of pass p), each
 * followed by the x guard copy: Filter::Binomial / Compensated along x
 * (filter.h:22-78).  Two to 16 passes on tiles at least gc + passes wide run
 * as one tile-parallel launch into a library-held scratch grid (one per
 * device and stream, allocated on first use) and a device copy back, with
 * results identical to the pass-by-pass sequence (every element computed as
 * A[i-1]*a + A[i]*b + A[i+1]*c); otherwise pass by pass. */
int zpic_filter_x(float *buf, const zpic_geom *g, int passes, const float *coef, int periodic_x, int periodic_y,
                  void *stream);

/* Current::advance (current.cpp:227-246) for periodic boundaries: guard add,
 * guard copy and the digital filter (filter.h:9-78).  filter_kind 0 none,
 * 1 binomial, 2 compensated; filter_dir 0 x / 1 y. */
int zpic_current_advance(float *J, const zpic_geom *g, int filter_kind, int filter_dir,
                         int filter_order, int periodic_x, int periodic_y, void *stream);

/* Current::advance (current.cpp:226-246) on a grid periodic along x and y
 * with no filter or a filter along x (filter.h:22-78): add_from_gc,
 * copy_to_gc and `passes` kernel3_x passes (coef[3p..3p+2] = a, b, c), each
 * followed by copy_to_gc_x, as one tile-parallel launch that reads the
 * deposit `dep` and writes the result to `J` (a different buffer of the same
 * geometry; `dep` is left as it is).  Every element of J, guards and tile
 * padding included, is identical to what that sequence leaves in place.
 * Needs nx >= gc_lower + gc_upper + passes and ny >= gc_lower + gc_upper. */
int zpic_current_update(const float *dep, float *J, const zpic_geom *g, int passes, const float *coef,
                        void *stream);

/* EMF::advance(Current&) (emf.cpp:513-580): B half step, E step with J,
 * B half step per tile, then guard copy of E and B. */
int zpic_yee(float *E, float *B, const float *J, const zpic_geom *g, float dt_dx, float dt_dy,
             float dt, int periodic_x, int periodic_y, void *stream);

/* EMF::get_energy (emf.cpp:662-713): sums of E_c^2, B_c^2 over interior
 * cells in the reference's order.  out6 is a device pointer to 6 doubles;
 * tile_scratch holds 6*ntiles doubles. */
int zpic_field_energy(const float *E, const float *B, const zpic_geom *g, double *tile_scratch,
                      double *out6, void *stream);

/* No counterpart in the reference.  zpic_field_energy for callers that
 * cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc): the per-tile sums live in
 * a library-held device array of 6 * 65536 doubles (3 MiB per device, kept
 * until the process ends), so no tile_scratch argument and no allocation.
 * Grids of more than 65536 tiles give ZPIC_ERANGE.  Results equal
 * zpic_field_energy's bit for bit.  Every call on one device shares that
 * array: calls that overlap on one device (other streams, other host
 * threads) race and must be serialised by the caller.  `stream` must belong
 * to the current device. */
int zpic_field_energy_noalloc(const float *E, const float *B, const zpic_geom *g, double *out6, void *stream);

/* Sum of per-tile doubles (the OpenMP reduction of Species::push,
 * species.cpp:2311): in tile order when deterministic != 0 (the reference's
 * single-thread result), as a parallel tree otherwise.  out is a device
 * pointer. */
int zpic_sum_tiles(const double *tile_values, int ntiles, double *out, int deterministic, void *stream);

/* Species::deposit_charge (species.cpp:2348-2410) into a scalar grid with
 * guard cells (accumulates), e.g. the gc {0,1} grid of save_charge. */
int zpic_deposit_charge(const zpic_particles *p, const zpic_geom *g, float *rho, float q,
                        int deterministic, void *stream);

/* Species::dep_phasespace (species.cpp:2579-2710 1-D, :2786-3020 2-D): phase-
 * space density of the particles into out[size0] (1-D, size1 = 0) or
 * out[size1][size0] (2-D, quant1 > quant0).  Quantities 0..4 = x, y, ux, uy,
 * uz; the ranges of x and y are in cells (the reference divides them by dx
 * first); norm is the caller's normalisation.  out is cleared first.  The
 * deterministic mode adds in the reference's particle order (one wavefront). */
int zpic_phasespace(const zpic_particles *p, int ntiles_x, int ntiles_y, int nx, int ny, int quant0,
                    float min0, float max0, int size0, int quant1, float min1, float max1, int size1,
                    float norm, float *out, int deterministic, void *stream);

/* Physical boundaries of the fields, after their guard cells were filled
 * with the axis' periodic flag off: EMF::process_bc (emf.cpp:207-471; bc
 * values 2 pec, 3 pmc) on E and B, Current::process_bc (current.cpp:53-218;
 * 2 reflecting) on J.  bc[4] = {x lower, x upper, y lower, y upper} (0 none,
 * 1 periodic: nothing to do). */
int zpic_emf_bc(float *E, float *B, const zpic_geom *g, const int bc[4], void *stream);
int zpic_current_bc(float *J, const zpic_geom *g, const int bc[4], void *stream);

/* Absorbing field boundary (no reference counterpart; the EMF boundary
 * value ZPIC_EMF_ABSORBING next to the reference's none 0, periodic 1, pec
 * 2, pmc 3): within `width` cells of each absorbing wall E and B are scaled
 * by exp(-(alpha (width - i))^2) every step, i the distance in cells from the
 * wall (the mask of Cerjan et al., Geophysics 50 (1985) 705); guard cells
 * included.  bc[4] = {x lower, x upper, y lower, y upper}. */
#define ZPIC_EMF_ABSORBING 4
int zpic_emf_absorb(float *E, float *B, const zpic_geom *g, const int bc[4], int width, float alpha,
                    void *stream);

/* vec3grid::gather (vec3grid.h:76-158): component fc (0,1,2) of the interior
 * cells into a contiguous [ny*ntiles_y][nx*ntiles_x] device array. */
int zpic_gather_component(const float *buf, const zpic_geom *g, int ncomp, int fc, float *out, void *stream);

/* Fills n floats with a value (grid::zero / set). */
int zpic_fill(float *buf, size_t n, float value, void *stream);

/* grid<T>::add (grid.h:208-212): dst += src over n floats (whole buffers,
 * guard cells included). */
int zpic_grid_add(float *dst, const float *src, size_t n, void *stream);

/* Measurement only, no reference counterpart: reads and writes back every
 * record of n particles (int2 ix, float2 x, float3 u: 28 B each, modified),
 * the particle pass's HBM traffic without its arithmetic, to time the
 * ceiling that traffic allows on the device. */
int zpic_particle_stream_probe(int *ix, float *x, float *u, size_t n, void *stream);

/* Measurement only, no reference counterpart: kind 0 copies n 16-byte words
 * from a to b; kind 1 reads and writes back n particle records (a = int2 ix,
 * b = float2 x, c = float3 u).  loads_per_lane (1, 2, 4 or 8) independent
 * records or words are in flight per lane; `blocks` workgroups of 256 lanes
 * stride over the buffers.  Kinds 2 and 3: the records as `blocks` tiles of
 * n / blocks records, one workgroup per tile; kind 2 with each of the four
 * waves streaming a contiguous quarter of the tile (the particle pass's
 * order), kind 3 with the waves walking the tile together (loads_per_lane
 * 1, 2 or 4).  Times the HBM ceiling of an access pattern. */
int zpic_bandwidth_probe(int kind, void *a, void *b, void *c, size_t n, int loads_per_lane, int blocks,
                         void *stream);

/* grid<T>::x_shift_left (grid.h:455-485): every row of every tile, guard
 * cells included, shifted left by `shift` cells with zeros entering at the
 * right, then the x guard copy.  Requires 0 < shift < gc_upper. */
int zpic_x_shift_left(float *buf, const zpic_geom *g, int ncomp, int shift, int periodic_x, int periodic_y,
                      void *stream);

/* zpic_x_shift_left of one or two float3 grids of the same geometry (buf1
 * may be NULL), both grids per launch (the shift, then the x guard copy),
 * with the same results: EMF::move_window (emf.cpp:120-129) shifts E and B
 * together. */
int zpic_x_shift_left2(float *buf0, float *buf1, const zpic_geom *g, int shift, int periodic_x, void *stream);

/* div_corr_x (laser.cpp:20-70): Ex and Bx of the interior cells set so that
 * the discrete divergence of E and B vanishes, accumulated in double from
 * the right edge of each row of cells leftwards (zero field assumed beyond
 * it), dx/dy the cell size.  Guard cells are left as they are. */
int zpic_div_corr_x(float *E, float *B, const zpic_geom *g, float dx, float dy, void *stream);

/* Threads per tile of the fast particle pass: a multiple of 64 in [64, 512]
 * (default 256).  Library-wide; the deterministic mode always uses 64. */
int zpic_set_threads_per_tile(int threads);
int zpic_get_threads_per_tile(void);

/* Tile sort of the fast mode, library-wide: 0 (default) = order-free (the
 * species objects stage leavers in the particle pass, see
 * zpic_sort.stage_ix, and one placement pass fills each tile; counts,
 * offsets and per-tile multisets are the reference's, the order within a
 * tile is not); 1 = the reference's single-thread within-tile order
 * (particles.cpp:379-614), as in the deterministic mode, which always keeps
 * it. */
int zpic_set_sort_order(int reference_order);
int zpic_get_sort_order(void);
/* Threads per tile the fast particle pass actually uses for tiles of
 * tile_vol elements when `threads` is requested (fewer wavefronts when the
 * LDS staging of a large tile would not fit); 0 if nothing fits. */
int zpic_fitted_threads_per_tile(int threads, int tile_vol);

/* Autotuner (no reference counterpart): times short Weibel-like runs at ppc
 * for every tile shape that divides the grid_x x grid_y grid ({8,16,32} x
 * {8,16,32} cells) and every threads-per-tile in {64, 128, 256, 512} that
 * fits, each on a sample of about four tiles per compute unit (the shapes
 * are compared at full occupancy, whatever the grid's own size), at most
 * max_cells cells (0 = 1024*1024), and returns the fastest (particles per ns
 * over push, deposit, sort and field stages).  Results are cached in the text file cache_path (NULL = none)
 * keyed by device name, grid and ppc; from_cache reports a hit.  Leaves the
 * library's threads per tile unchanged. */
typedef struct {
    unsigned nx, ny;
    int threads;
    double particles_per_ns;
    int from_cache;
} zpic_tune_result;
int zpic_autotune(unsigned grid_x, unsigned grid_y, unsigned ppc_x, unsigned ppc_y, int steps, long long max_cells,
                  const char *cache_path, zpic_tune_result *out);

/* --------------------------------------------------- batched stages
 * Many independent simulations advanced with one launch per stage over all
 * their tiles (a parameter scan).  No reference counterpart: the reference
 * runs one simulation per process; each instance of a batch computes what its
 * own single-instance run computes (same kernels, same per-tile work). */

/* One species of one instance. */
typedef struct {
    zpic_particles p, tmp;  /* particle and staging buffers              */
    zpic_sort s;            /* sort scratch                              */
    zpic_geom fld;          /* field geometry of the instance (gc {1,2}) */
    const float *E, *B;     /* fields of the instance                    */
    float *J;               /* current of the instance                   */
    double *tile_energy;    /* [ntiles] scratch                          */
    double *energy_out;     /* [1] device: sum of the push energies      */
    float alpha, dt_dx, dt_dy, q, qnx, qny;
    int periodic_x, periodic_y;
    int threads;            /* threads per tile of this item's particle pass
                               (fitted to its tiles); 0: the plan's          */
} zpic_batch_species;

typedef struct zpic_batch_particles zpic_batch_particles;

/* Builds the device tables for advancing `n` species (at most one per
 * instance: species of one instance share J and must be advanced one after
 * the other, so use one plan per species slot). */
int zpic_batch_particles_create(zpic_batch_particles **plan, const zpic_batch_species *items, int n,
                                int threads_per_tile);
/* Species::advance(emf, current) of every item: fused push + move + deposit,
 * tile sort, energy sums (fast mode). */
int zpic_batch_particles_advance(zpic_batch_particles *plan, void *stream);
void zpic_batch_particles_free(zpic_batch_particles *plan);

/* Fields of one instance. */
typedef struct {
    float *E, *B, *J;
    zpic_geom g;                  /* gc {1,2}                                */
    float dt_dx, dt_dy, dt;       /* as EMF::advance narrows them            */
    int filter_kind, filter_dir, filter_order;   /* as zpic_current_advance  */
    int periodic_x, periodic_y;
} zpic_batch_fields_item;

typedef struct zpic_batch_fields zpic_batch_fields;

int zpic_batch_fields_create(zpic_batch_fields **plan, const zpic_batch_fields_item *items, int n);
/* Current::zero of every instance. */
int zpic_batch_zero_current(zpic_batch_fields *plan, void *stream);
/* Current::advance of every instance (guard add, guard copy, filter). */
int zpic_batch_current_advance(zpic_batch_fields *plan, void *stream);
/* EMF::advance(current) of every instance (Yee + guard copy of E, B). */
int zpic_batch_yee(zpic_batch_fields *plan, void *stream);
void zpic_batch_fields_free(zpic_batch_fields *plan);

/* The whole step of every instance in one persistent launch (no reference
 * counterpart; Simulation::advance, simulation.h:90-107, of each instance).
 * species[k * n + i] is species slot k of instance i (nslots slots),
 * fields[i] the fields of instance i.  Every instance runs the particle
 * pass with the threads per tile that fit its tiles (as
 * zpic_fitted_threads_per_tile(threads_per_tile, tile_vol)), so each computes
 * what its own Simulation::advance computes with that setting.  Fast mode
 * and periodic boundaries only; the sort is the order-free one when every
 * species item's stage buffers are set, the reference-order one when none
 * is.  Workgroups walk a list of (stage, instance,
 * tiles) work items in dependency order and wait on per-(stage, instance)
 * completion counters, so instances overlap; advance enqueues a counter
 * reset and the persistent kernel (two launches, graph-capturable).
 * status synchronises the device and reports ZPIC_EHIP if a dependency wait
 * ever exceeded its spin limit (the results of that step are then invalid);
 * info reports the launch shape (workgroups, threads) and the work items. */
typedef struct zpic_batch_step zpic_batch_step;
int zpic_batch_step_create(zpic_batch_step **plan, const zpic_batch_species *species, int nslots,
                           const zpic_batch_fields_item *fields, int n, int threads_per_tile);
int zpic_batch_step_advance(zpic_batch_step *plan, void *stream);
int zpic_batch_step_status(zpic_batch_step *plan);
int zpic_batch_step_info(const zpic_batch_step *plan, int *workgroups, int *threads, int *items);
void zpic_batch_step_free(zpic_batch_step *plan);

/* --------------------------------------------- object entry points (2) */
typedef struct zpic_simulation zpic_simulation;
typedef struct zpic_species zpic_species;

/* Simulation(ntiles, nx, box, dt) (simulation.h:39-43).  NULL on error
 * (e.g. the Courant condition of EMF::EMF, emf.cpp:22-28). */
zpic_simulation *zpic_simulation_new(unsigned ntiles_x, unsigned ntiles_y, unsigned nx, unsigned ny,
                                     float box_x, float box_y, double dt);
void zpic_simulation_free(zpic_simulation *sim);

/* Species(name, m_q, ppc) (species.h:147) with its density / udist setters
 * (species.h:172-194, udist.h:19-58, density.h:32-60).
 * udist kind: 0 None, 1 Cold(ufl), 2 Thermal(uth, ufl), 3 ThermalCorr(uth, ufl, npmin). */
zpic_species *zpic_species_new(const char *name, float m_q, unsigned ppc_x, unsigned ppc_y);
void zpic_species_free(zpic_species *sp);
int zpic_species_set_udist(zpic_species *sp, int kind, const float uth[3], const float ufl[3], int npmin);
/* Offset added to the species id in the momentum RNG seed (0 = the
 * reference's seeding, udist.cpp:73,94); before add_species.  For seeded
 * ensembles of otherwise identical runs.  No reference counterpart. */
int zpic_species_set_seed(zpic_species *sp, unsigned offset);
int zpic_species_set_density_uniform(zpic_species *sp, float n0);

/* Species::set_density (species.h:172-175) with any profile of
 * density.h:30-132 (the Cython classes openmp/python/em2d/density/defs.pxd
 * bind None, Uniform, Step and Slab): kind ZPIC_DENSITY_*, density n0; a
 * step starts at `pos` along dir (0 x, 1 y), a slab spans [begin, end), a
 * sphere has centre and radius; a ramp (ZPIC_DENSITY_RAMP, no reference
 * counterpart) rises from 0 at `begin` to n0 at `end` along dir.  Before
 * add_species. */
typedef struct {
    int kind;
    int dir;
    float n0;
    float pos;
    float begin, end;
    float center[2];
    float radius;
} zpic_density;
int zpic_species_set_density(zpic_species *sp, const zpic_density *d);

/* Simulation::add_species (simulation.h:68-71): the simulation keeps a
 * non-owning reference; the species must outlive it. */
int zpic_simulation_add_species(zpic_simulation *sim, zpic_species *sp);

/* Simulation::advance (simulation.h:90-107), n times. */
int zpic_simulation_advance(zpic_simulation *sim, int n);

/* Replaces the species' particles by a device injection of its density
 * profile over the whole box (Species::np_inject + inject,
 * species.cpp:1808-1842, through zpic_np_inject / zpic_inject), momenta 0:
 * the particles a cold initial load places, made by the kernels of the
 * moving window.  For checking those kernels against the initial load. */
int zpic_species_reinject(zpic_species *sp);

/* Simulation::set_moving_window (simulation.h:56-62; Cython:
 * Simulation(..., moving_window=True), em2d.pyx:1478-1480): fields and
 * current lose their x periodicity, species get open x walls, and the window
 * moves one cell whenever the time passes the next cell.  ZPIC_ESTATE after
 * the first step. */
int zpic_simulation_set_moving_window(zpic_simulation *sim);

/* Simulation::advance_mov_window (simulation.h:113-127), n times: on the
 * steps where the window moves, particles shift one cell left, the species'
 * density profile is injected into the rightmost column of cells, and E and
 * B shift one cell left after the field update. */
int zpic_simulation_advance_mov_window(zpic_simulation *sim, int n);

/* Cells the window has moved (EMF's MovingWindow::n_move, moving_window.h). */
int zpic_simulation_window_moves(const zpic_simulation *sim, unsigned *n_move);

/* Device buffer, geometry and periodic flags (x, y) of a field of the
 * simulation (which: 0 E, 1 B, 2 J), for the stage entry points (guard
 * exchange, filters, shifts) on the simulation's own grids; the buffer stays
 * owned by the simulation.  The Cython Vec3Grid view (em2d.pyx:280-500)
 * works through it. */
int zpic_field_device(zpic_simulation *sim, int which, float **buf, zpic_geom *g, int periodic[2]);

/* Simulation geometry: tiles, cells per tile, box, time step
 * (Simulation::ntiles / nx / box / dt, simulation.h:22-25). */
int zpic_simulation_info(const zpic_simulation *sim, unsigned ntiles[2], unsigned nx[2], float box[2],
                         double *dt);

/* EMF::get_iter / Current::get_iter (emf.h:91, current.h:63). */
int zpic_emf_get_iter(const zpic_simulation *sim);
int zpic_current_get_iter(const zpic_simulation *sim);

/* Simulation::energy_info (simulation.h:150-169): the energy report on stdout. */
int zpic_simulation_energy_info(zpic_simulation *sim);

/* Species::dep_phasespace (species.cpp:2579-2620, 2786-2850) into a host
 * array: [size0] (size1 <= 0) or [size1][size0]; q 0 x, 1 y, 2 ux, 3 uy,
 * 4 uz (Cython Species.get_phasespace, em2d.pyx:1217-1288). */
int zpic_species_get_phasespace(zpic_species *sp, int q0, float min0, float max0, int size0, int q1, float min1,
                                float max1, int size1, float *out);

/* Laser pulses (Laser::PlaneWave / Laser::Gaussian, laser.h; Cython
 * openmp/python/em2d/laser/defs.pxd): kind 0 plane wave, 1 Gaussian beam;
 * fwhm > 0 overrides rise / flat / fall; cos_pol = sin_pol = 0 takes the
 * polarisation angle; filter = compensated x filter passes (reference
 * default 1); W0 / focus / axis for the Gaussian beam. */
typedef struct {
    int kind;
    float start, fwhm, rise, flat, fall;
    float a0, omega0;
    float polarization, cos_pol, sin_pol;
    unsigned filter;
    float W0, focus, axis;
} zpic_laser;

/* Pulse::add (laser.h:67-81, launch laser.cpp:224-437): adds the pulse to the
 * simulation's E and B.  ZPIC_EINVAL for invalid pulse parameters (the
 * reference's -1). */
int zpic_laser_add(zpic_simulation *sim, const zpic_laser *laser);

/* Stages of one Simulation::advance, for stage-level parity checks:
 * 0 push(k), 1 move(k), 2 tile_sort(k), 3 Species::advance(k),
 * 4 Current::zero, 5 Current::advance, 6 EMF::advance(current),
 * 7 J add_from_gc, 8 J copy_to_gc, 9 J filter, 10 EMF::process_bc,
 * 11 EMF::advance() (no current). */
int zpic_simulation_stage(zpic_simulation *sim, int stage, int species);

/* Current::set_filter (current.h:137) with Filter::None/Binomial/Compensated. */
int zpic_current_set_filter(zpic_simulation *sim, int kind, int dir, int order);

/* 1 = deterministic deposit and energy order (bit-for-bit the reference's
 * single-thread results), 0 = fast atomic order (default). */
int zpic_simulation_set_deterministic(zpic_simulation *sim, int on);

unsigned zpic_simulation_get_iter(const zpic_simulation *sim);
double zpic_simulation_get_t(const zpic_simulation *sim);
uint64_t zpic_simulation_get_nmove(zpic_simulation *sim);

/* Species::get_energy (species.h:365-368) and EMF::get_energy (emf.cpp:662). */
double zpic_species_get_energy(zpic_species *sp);
int zpic_emf_get_energy(zpic_simulation *sim, double ene_E[3], double ene_B[3]);

uint64_t zpic_species_get_nmove(const zpic_species *sp);
uint64_t zpic_species_np_total(zpic_species *sp);

/* Host copies of a species' particles in tile order (compact): np/offset
 * [ntiles], ix [2n], x [2n], u [3n]; any pointer may be NULL. */
int zpic_species_get_particles(zpic_species *sp, int *np, int *offset, int *ix, float *x, float *u);
/* Replaces a species' particles (same compact layout). */
int zpic_species_set_particles(zpic_species *sp, const int *np, const int *ix, const float *x, const float *u);

/* Whole tile-major field buffers (guards included): which 0 = E, 1 = B, 2 = J. */
int zpic_field_get(zpic_simulation *sim, int which, float *host_out);
int zpic_field_set(zpic_simulation *sim, int which, const float *host_in);
int zpic_field_geom(const zpic_simulation *sim, zpic_geom *g);

/* Species::save_charge's density (species.cpp:2508-2522): tile-major
 * gc {0,1} scalar grid after the guard add, host copy; returns the element
 * count or a negative code. */
long long zpic_species_charge(zpic_species *sp, float *host_out);

/* Species::gather(part::quant, float*) (species.h:515): x, y (global
 * position in simulation units), ux, uy, uz for quant 0..4, host copy. */
int zpic_species_gather(zpic_species *sp, int quant, float *host_out);

/* Average kernel time (ms) of the last step's stages measured with HIP
 * events when timing is enabled: out[0] push+move+deposit, out[1] sort,
 * out[2] current, out[3] yee, out[4] whole step. */
int zpic_simulation_set_timing(zpic_simulation *sim, int on);
int zpic_simulation_get_timing(zpic_simulation *sim, double out[5]);

/* ZDF diagnostics, written under the current working directory as the
 * reference does.  EMF::save(field 0 E / 1 B, fc) (emf.cpp:589-655),
 * Current::save(fc) (current.cpp:263-313), vec3grid::save(fc, filename)
 * with minimal metadata (which 0 E, 1 B, 2 J; vec3grid.h:422-437),
 * Species::save (species.cpp:2417-2501), save_charge (:2508-2560) and
 * save_phasespace 1-D (:2717-2764) / 2-D (:3026-3100; pass size1 = 0 for
 * 1-D). */
int zpic_emf_save(zpic_simulation *sim, int field, int fc);
int zpic_current_save(zpic_simulation *sim, int fc);
int zpic_field_save_file(zpic_simulation *sim, int which, int fc, const char *filename);
int zpic_species_save(zpic_species *sp);
int zpic_species_save_charge(zpic_species *sp);
int zpic_species_save_phasespace(zpic_species *sp, int quant0, float min0, float max0, int size0, int quant1,
                                 float min1, float max1, int size1);

/* zdf_save_grid (zdf.c:1515-1585) on host data: writes
 * "<path>/<info.name>-<it.n:06>.zdf" holding TYPE "grid", the grid info, the
 * iteration and one dataset of `data_type` (9 = float32, 10 = float64, ...;
 * the zdf_data_type numbering).  Pure host code: needs no GPU. */
typedef struct {
    const char *name;
    int type;              /* 0 linear, 1 log10, 2 log2 */
    double min, max;
    const char *label;
    const char *units;
} zpic_zdf_axis;
typedef struct {
    const char *name;
    uint32_t ndims;
    uint64_t count[3];
    const char *label;
    const char *units;
    const zpic_zdf_axis *axis;  /* ndims entries, or NULL */
} zpic_zdf_grid_info;
typedef struct {
    const char *name;      /* NULL writes "Iteration" */
    int32_t n;
    double t;
    const char *time_units;
} zpic_zdf_iteration;
int zpic_zdf_save_grid(const void *data, int data_type, const zpic_zdf_grid_info *info,
                       const zpic_zdf_iteration *it, const char *path);

/* Learned current denoiser (no reference counterpart): 3 convolution layers
 * of 5x5 taps, 3 -> 16 -> 16 -> 3 channels with ReLU between them, periodic
 * padding, BF16 operands on the matrix cores with FP32 accumulation.  The
 * 8,835 parameters are W1[16][3][5][5], b1[16], W2[16][16][5][5], b2[16],
 * W3[3][16][5][5], b3[3] (FP32).  The network sees J / scale and its output
 * is multiplied by scale.  An object is sized for one H x W grid (W a
 * multiple of 32) and stores up to max_pairs (input, target) J pairs for
 * training on the GPU (Adam).  Host arrays are [y][x][3] FP32. */
typedef struct zpic_denoiser zpic_denoiser;
zpic_denoiser *zpic_denoiser_new(int H, int W, int max_pairs);
void zpic_denoiser_free(zpic_denoiser *d);
int zpic_denoiser_set_params(zpic_denoiser *d, const float *host);
int zpic_denoiser_get_params(zpic_denoiser *d, float *host);
int zpic_denoiser_set_scale(zpic_denoiser *d, float scale);
float zpic_denoiser_get_scale(const zpic_denoiser *d);
int zpic_denoiser_size(const zpic_denoiser *d, int *H, int *W);
int zpic_denoiser_forward_host(zpic_denoiser *d, const float *in, float *out);
int zpic_denoiser_gradient_host(zpic_denoiser *d, const float *in, const float *target, float *grad,
                                double *loss);
int zpic_denoiser_add_pair(zpic_denoiser *d, const float *J_in, const float *J_target, const zpic_geom *g,
                           void *stream);
int zpic_denoiser_num_pairs(const zpic_denoiser *d);
int zpic_denoiser_clear_pairs(zpic_denoiser *d);
int zpic_denoiser_get_pair(zpic_denoiser *d, int k, float *in, float *target);
int zpic_denoiser_target_rms(zpic_denoiser *d, double *rms);
int zpic_denoiser_train(zpic_denoiser *d, int steps, float lr, unsigned seed, int use, double *loss_hist);
int zpic_denoiser_eval_pair(zpic_denoiser *d, int k, double *mse_net, double *mse_input);
int zpic_denoiser_apply(zpic_denoiser *d, float *J, const zpic_geom *g, void *stream);
int zpic_denoiser_reldiff(zpic_denoiser *d, const float *a, const float *b, size_t n, void *stream, double *out);
/* Stores (in's current J, target's current J) as a training pair; both
 * simulations must use the same grid and tiling. */
int zpic_simulation_add_pair(zpic_denoiser *d, zpic_simulation *in, zpic_simulation *target);

/* Boundary conditions (Species::set_bc species.h:210-237, EMF::set_bc
 * emf.h:112-143, Current::set_bc current.h:84-114): bc[4] = {x lower,
 * x upper, y lower, y upper}; species 0 open, 1 periodic, 2 reflecting; EMF
 * 0 none, 1 periodic, 2 pec, 3 pmc (4 absorbing: no reference
 * counterpart, see zpic_emf_absorb); current 0 none, 1 periodic, 2
 * reflecting.  Periodic must be set on both sides of an axis. */
int zpic_species_set_bc(zpic_species *sp, const int bc[4]);
int zpic_emf_set_bc(zpic_simulation *sim, const int bc[4]);
/* Width (cells, default 20) and strength (default 0.015) of the absorbing
 * field layers (no reference counterpart). */
int zpic_emf_set_absorbing(zpic_simulation *sim, int width, float alpha);
int zpic_current_set_bc(zpic_simulation *sim, const int bc[4]);

/* Learned-filter hooks: the simulation's current filter becomes the
 * denoiser (not owned); set_fallback switches it to the binomial x order-1
 * filter (on = 1) or back (on = 0); get_fallback reports the state (-1 if the
 * filter is not the learned one). */
int zpic_current_set_learned_filter(zpic_simulation *sim, zpic_denoiser *d);
int zpic_current_set_fallback(zpic_simulation *sim, int on);
int zpic_current_get_fallback(zpic_simulation *sim);

/* Copies E and B of src into dst (same grid). */
int zpic_simulation_copy_fields(zpic_simulation *dst, const zpic_simulation *src);

/* Every stride-th particle (from index phase) of each tile of src, in tile
 * order, into dst; dst's counts and offsets are rewritten.  src_geom and
 * dst_geom describe the two species' tilings (tile counts and cells per
 * tile), which must be equal (ZPIC_EINVAL otherwise): the tile tables of both
 * hold ntiles_x * ntiles_y entries.  Species form: the species' particle
 * buffers. */
int zpic_subsample(const zpic_particles *src, const zpic_geom *src_geom, zpic_particles *dst,
                   const zpic_geom *dst_geom, int stride, int phase, void *stream);
int zpic_species_subsample(zpic_species *dst, zpic_species *src, int stride, int phase);

/* Device synchronisation of the simulation's stream. */
int zpic_simulation_sync(zpic_simulation *sim);

/* Batch of independent simulations advanced together (one launch per stage
 * over all instances, optionally one hipGraph launch per step).  No reference
 * counterpart.  Simulations are added by reference and must outlive the
 * batch; each must hold the same number of species. */
typedef struct zpic_batch zpic_batch;
zpic_batch *zpic_batch_new(int threads_per_tile);
void zpic_batch_free(zpic_batch *b);
int zpic_batch_add(zpic_batch *b, zpic_simulation *sim);
/* As zpic_batch_add, with this instance's threads per tile (0 = the
 * batch's): its particle pass runs as its own Simulation::advance does
 * after zpic_set_threads_per_tile(threads_per_tile). */
int zpic_batch_add_threads(zpic_batch *b, zpic_simulation *sim, int threads_per_tile);
int zpic_batch_set_graph(zpic_batch *b, int on);
/* on = 1: every step is one persistent launch over all instances
 * (zpic_batch_step; periodic boundaries, the library's fast-mode sort
 * order); on = 0 (default): one launch per stage. */
int zpic_batch_set_persistent(zpic_batch *b, int on);
/* Synchronises; ZPIC_EHIP if a persistent step reported an error. */
int zpic_batch_status(zpic_batch *b);
int zpic_batch_advance(zpic_batch *b, int n);

#ifdef __cplusplus
}
#endif

#endif /* ZPIC_AMD_H_ */
