// zpic C++ API for the MI355X em2d step: the reference's Simulation /
// Species / EMF / Current / vec3grid / Filter / UDistribution / Density
// classes (openmp/em2d) with the same names, signatures and argument
// meanings, backed by device-resident buffers and the stage entry points of
// zpic_amd.h.  Compile user code that includes this header with hipcc
// (it uses HIP's int2/uint2/float2/float3/double3 vector types).
//
// Error behaviour: where the reference prints "(*error*) ..." and exits,
// these classes throw std::runtime_error carrying the same message, so that
// a library never terminates its host process.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "zdf.hpp"
#include "zpic_amd.h"

namespace coord {
enum cart { x = 0, y };
}

namespace fcomp {
enum cart { x = 0, y, z };
}

namespace part {
enum quant { x = 0, y, ux, uy, uz };
}

namespace phasespace {
enum quant { x = 0, y, ux, uy, uz };
// Name, label and units of a phase-space axis (species.h:20-38)
inline void qinfo( quant q, std::string & name, std::string & label, std::string & units ) {
    static const char * n[] = { "x", "y", "ux", "uy", "uz" };
    static const char * l[] = { "x", "y", "u_x", "u_y", "u_z" };
    name = n[q];
    label = l[q];
    units = ( q <= y ) ? "c/\\omega_n" : "c";
}
}  // namespace phasespace

// MovingWindow (moving_window.h): number of cells the window has moved
class MovingWindow {
    bool active_;
    unsigned int n_move_;
    double dx_;

  public:
    MovingWindow() : active_( false ), n_move_( 0 ), dx_( 0 ) {}
    void init( float const dx ) {
        active_ = true;
        n_move_ = 0;
        dx_ = dx;
    }
    bool active() const { return active_; }
    unsigned int n_move() const { return n_move_; }
    int advance() {
        if ( active_ ) n_move_++;
        return n_move_;
    }
    double motion() const { return n_move_ * dx_; }
    bool needs_move( double const t ) const { return active_ && ( t > dx_ * ( n_move_ + 1 ) ); }
};

// bnd<T> (bnd.h:21-78): per-axis lower / upper values
template <typename T>
class bnd {
  public:
    struct { T lower; T upper; } x;
    struct { T lower; T upper; } y;
    bnd() : x{ T( 0 ), T( 0 ) }, y{ T( 0 ), T( 0 ) } {}
    bnd( T v ) : x{ v, v }, y{ v, v } {}
};

namespace zpic {

// Library stream shared by all objects (default priority).
hipStream_t stream();
// Second library stream (default priority) for work that may run beside the
// first's: the last species' sort while the current and fields advance.
hipStream_t side_stream();
// While alive, stream() returns `s` on the constructing thread.
class StreamScope {
    hipStream_t prev;

  public:
    explicit StreamScope( hipStream_t s );
    ~StreamScope();
    StreamScope( const StreamScope & ) = delete;
    StreamScope & operator=( const StreamScope & ) = delete;
};

// Throws std::runtime_error for a non-zero zpic_* return code.
void check( int rc );
void hip_check( hipError_t e, const char * what );

template <class T>
T * device_alloc( size_t n ) {
    void * p = nullptr;
    if ( n ) hip_check( hipMalloc( &p, n * sizeof( T ) ), "hipMalloc" );
    return static_cast<T *>( p );
}
inline void device_free( void * p ) { if ( p ) (void) hipFree( p ); }

}  // namespace zpic

// ---------------------------------------------------------------- grid<T>
// Tile-major grid with guard cells (grid.h:14-638); T is float or float3.
template <class T>
class grid {
  protected:
    uint2 ntiles;
    int2 periodic;
    const uint2 dims;

  public:
    T * d_buffer;
    const uint2 nx;
    const bnd<unsigned int> gc;
    const uint2 ext_nx;
    const unsigned int offset;
    const unsigned int tile_vol;
    std::string name;

    grid( uint2 const ntiles_, uint2 const nx_, bnd<unsigned int> const gc_ )
        : ntiles( ntiles_ ), periodic( make_int2( 1, 1 ) ), dims( make_uint2( ntiles_.x * nx_.x, ntiles_.y * nx_.y ) ),
          d_buffer( nullptr ), nx( nx_ ), gc( gc_ ),
          ext_nx( make_uint2( gc_.x.lower + nx_.x + gc_.x.upper, gc_.y.lower + nx_.y + gc_.y.upper ) ),
          offset( gc_.y.lower * ext_nx.x + gc_.x.lower ), tile_vol( ( ext_nx.x * ext_nx.y + 3 ) & ~3u ), name( "grid" ) {
        if ( ntiles.x == 0 || ntiles.y == 0 ) throw std::runtime_error( "Invalid number of tiles" );
        if ( nx.x == 0 || nx.y == 0 ) throw std::runtime_error( "Invalid tile size" );
        if ( gc.x.lower != gc.y.lower || gc.x.upper != gc.y.upper )
            throw std::runtime_error( "guard cells must be the same along x and y" );
        d_buffer = zpic::device_alloc<T>( buffer_size() );
    }
    grid( uint2 const ntiles_, uint2 const nx_ ) : grid( ntiles_, nx_, bnd<unsigned int>( 0 ) ) {}
    ~grid() { zpic::device_free( d_buffer ); }
    grid( const grid & ) = delete;
    grid & operator=( const grid & ) = delete;

    static constexpr int ncomp = sizeof( T ) / sizeof( float );

    uint2 get_ntiles() const { return ntiles; }
    uint2 get_dims() const { return dims; }
    int2 get_periodic() const { return periodic; }
    void set_periodic( int2 p ) { periodic = p; }
    std::size_t buffer_size() const { return static_cast<std::size_t>( tile_vol ) * ( ntiles.x * ntiles.y ); }

    zpic_geom geom() const {
        zpic_geom g;
        zpic_geom_init( &g, ntiles.x, ntiles.y, nx.x, nx.y, gc.x.lower, gc.x.upper );
        return g;
    }
    float * fbuf() { return reinterpret_cast<float *>( d_buffer ); }
    const float * fbuf() const { return reinterpret_cast<const float *>( d_buffer ); }

    int zero() {
        zpic::check( zpic_fill( fbuf(), buffer_size() * ncomp, 0.0f, zpic::stream() ) );
        return 0;
    }
    void copy_to_gc() {
        zpic_geom g = geom();
        zpic::check( zpic_copy_to_gc( fbuf(), &g, ncomp, periodic.x, periodic.y, zpic::stream() ) );
    }
    void add_from_gc() {
        zpic_geom g = geom();
        zpic::check( zpic_add_from_gc( fbuf(), &g, ncomp, periodic.x, periodic.y, zpic::stream() ) );
    }
    // copy_to_gc of this grid and of `other` (same geometry), one launch
    void copy_to_gc( grid & other ) {
        zpic_geom g = geom();
        zpic::check( zpic_copy_to_gc2( fbuf(), other.fbuf(), &g, ncomp, periodic.x, periodic.y, zpic::stream() ) );
    }
    // add_from_gc followed by copy_to_gc; the guard sums the copy replaces
    // are not kept
    void add_copy_gc() {
        zpic_geom g = geom();
        zpic::check( zpic_add_copy_gc( fbuf(), &g, ncomp, periodic.x, periodic.y, zpic::stream() ) );
    }
    void kernel3_x( float a, float b, float c ) {
        zpic_geom g = geom();
        zpic::check( zpic_kernel3( fbuf(), &g, 0, a, b, c, periodic.x, periodic.y, zpic::stream() ) );
    }
    void kernel3_y( float a, float b, float c ) {
        zpic_geom g = geom();
        zpic::check( zpic_kernel3( fbuf(), &g, 1, a, b, c, periodic.x, periodic.y, zpic::stream() ) );
    }
    // kernel3_x passes with coefficients coef[3p .. 3p+2], one launch
    void filter_x( const std::vector<float> & coef ) {
        zpic_geom g = geom();
        zpic::check( zpic_filter_x( fbuf(), &g, (int) coef.size() / 3, coef.data(), periodic.x, periodic.y,
                                    zpic::stream() ) );
    }
    // Whole tile-major buffer to / from the host (synchronous).
    void download( T * host ) const {
        zpic::hip_check( hipMemcpyAsync( host, d_buffer, buffer_size() * sizeof( T ), hipMemcpyDeviceToHost, zpic::stream() ), "download" );
        zpic::hip_check( hipStreamSynchronize( zpic::stream() ), "sync" );
    }
    void upload( const T * host ) {
        zpic::hip_check( hipMemcpyAsync( d_buffer, host, buffer_size() * sizeof( T ), hipMemcpyHostToDevice, zpic::stream() ), "upload" );
        zpic::hip_check( hipStreamSynchronize( zpic::stream() ), "sync" );
    }
    // grid::gather (grid.h:222-247): interior cells, contiguous [dims.y][dims.x], host output
    unsigned int gather( float * host_out, int fc = 0 ) const {
        zpic_geom g = geom();
        float * d = zpic::device_alloc<float>( (size_t) dims.x * dims.y );
        zpic::check( zpic_gather_component( fbuf(), &g, ncomp, fc, d, zpic::stream() ) );
        zpic::hip_check( hipMemcpyAsync( host_out, d, sizeof( float ) * dims.x * dims.y, hipMemcpyDeviceToHost, zpic::stream() ), "gather" );
        zpic::hip_check( hipStreamSynchronize( zpic::stream() ), "sync" );
        zpic::device_free( d );
        return dims.x * dims.y;
    }
    // grid::save (grid.h:601-633): interior of component fc as a 2-D ZDF grid
    void save( zdf::grid_info & info, zdf::iteration & iter, std::string path, int fc = 0 ) const {
        info.ndims = 2;
        info.count[0] = dims.x;
        info.count[1] = dims.y;
        std::vector<float> h( (size_t) dims.x * dims.y );
        gather( h.data(), fc );
        zdf::save_grid( h.data(), info, iter, path );
    }
    void save( std::string filename, int fc = 0 ) const {
        std::vector<float> h( (size_t) dims.x * dims.y );
        gather( h.data(), fc );
        const uint64_t gdims[2] = { dims.x, dims.y };
        zdf::save_grid( h.data(), zdf::zdf_float32, 2, gdims, name, filename );
    }
};

// vec3grid<float3> (vec3grid.h:38-442)
template <class V>
class vec3grid : public grid<V> {
  public:
    using grid<V>::grid;
    // component fc of the interior cells, contiguous (vec3grid.h:76-158)
    unsigned int gather( const fcomp::cart fc, float * host_out ) const { return grid<V>::gather( host_out, fc ); }
    // vec3grid::save (vec3grid.h:398-437)
    void save( const fcomp::cart fc, zdf::grid_info & info, zdf::iteration & iter, std::string path ) const {
        grid<V>::save( info, iter, path, fc );
    }
    void save( const fcomp::cart fc, std::string filename ) const {
        static const char * comp[] = { "x", "y", "z" };
        std::vector<float> h( (size_t) this->dims.x * this->dims.y );
        grid<V>::gather( h.data(), fc );
        const uint64_t gdims[2] = { this->dims.x, this->dims.y };
        zdf::save_grid( h.data(), zdf::zdf_float32, 2, gdims, this->name + "-" + comp[fc], filename );
    }
};

// ----------------------------------------------------------- Filter:: (filter.h)
namespace Filter {
class Digital {
  public:
    virtual Digital * clone() const = 0;
    virtual void apply( vec3grid<float3> & fld ) = 0;
    virtual int kind() const = 0;
    virtual ~Digital() = default;
};
class None : public Digital {
  public:
    None * clone() const override { return new None(); }
    void apply( vec3grid<float3> & ) override {}
    int kind() const override { return 0; }
};
class Binomial : public Digital {
  protected:
    unsigned int order;
    coord::cart dir;

  public:
    Binomial( coord::cart dir_, unsigned int order_ = 0 ) : order( order_ > 0 ? order_ : 1 ), dir( dir_ ) {}
    Binomial * clone() const override { return new Binomial( dir, order ); }
    // along x all passes run as one fused launch (zpic_filter_x)
    void apply( vec3grid<float3> & fld ) override {
        if ( dir == coord::x ) {
            fld.filter_x( passes() );
            return;
        }
        for ( unsigned i = 0; i < order; i++ ) fld.kernel3_y( 0.25f, 0.5f, 0.25f );
    }
    int kind() const override { return 1; }
    // coefficients (a, b, c) of the passes of this filter
    virtual std::vector<float> passes() const {
        std::vector<float> c;
        for ( unsigned i = 0; i < order; i++ ) c.insert( c.end(), { 0.25f, 0.5f, 0.25f } );
        return c;
    }
    unsigned get_order() const { return order; }
    coord::cart get_dir() const { return dir; }
};
class Compensated : public Binomial {
  public:
    Compensated( coord::cart dir_, unsigned int order_ = 0 ) : Binomial( dir_, order_ ) {}
    Compensated * clone() const override { return new Compensated( dir, order ); }
    void apply( vec3grid<float3> & fld ) override {
        if ( dir == coord::x ) {
            fld.filter_x( passes() );
            return;
        }
        Binomial::apply( fld );
        const std::vector<float> c = passes();
        fld.kernel3_y( c[c.size() - 3], c[c.size() - 2], c[c.size() - 1] );
    }
    int kind() const override { return 2; }
    std::vector<float> passes() const override {
        std::vector<float> c = Binomial::passes();
        const float a = -1.0f;
        const float b = ( 4.0 + 2.0 * order ) / order;
        const float norm = 2 * a + b;
        c.insert( c.end(), { a / norm, b / norm, a / norm } );
        return c;
    }
};
// Learned denoiser in place of the binomial filter (no reference
// counterpart): applies a zpic_denoiser (not owned; it must outlive the
// filter and match the grid) to J's interior, then refreshes the guards.
// While `fallback` is set (by a failed shadow check) the binomial filter
// along x of order 1 runs instead.
class Learned : public Digital {
    zpic_denoiser * dn;
    bool fallback;

  public:
    explicit Learned( zpic_denoiser * d ) : dn( d ), fallback( false ) {}
    Learned * clone() const override {
        auto * c = new Learned( dn );
        c->fallback = fallback;
        return c;
    }
    void apply( vec3grid<float3> & fld ) override {
        if ( fallback || !dn ) {
            Binomial( coord::x, 1 ).apply( fld );
            return;
        }
        zpic_geom g = fld.geom();
        zpic::check( zpic_denoiser_apply( dn, fld.fbuf(), &g, zpic::stream() ) );
        fld.copy_to_gc();
    }
    int kind() const override { return 3; }
    void set_fallback( bool on ) { fallback = on; }
    bool get_fallback() const { return fallback; }
    zpic_denoiser * denoiser() const { return dn; }
};
}  // namespace Filter

// -------------------------------------------------------------- Particles
class Particles {
  public:
    const uint2 ntiles;
    const uint2 nx;
    int2 periodic;
    const uint2 dims;
    int * np;
    int * offset;
    int2 * ix;
    float2 * x;
    float3 * u;
    uint32_t max_part;

    Particles( uint2 const ntiles, uint2 const nx, uint32_t const max_part );
    ~Particles();
    Particles( const Particles & ) = delete;
    Particles & operator=( const Particles & ) = delete;

    zpic_particles view() const;
    uint64_t np_total() const;
    // Particles::grow_buffer (particles.h:174-187): reallocates the particle
    // arrays (contents discarded) to new_max rounded up to 64 Ki particles,
    // if new_max exceeds the current size
    void grow_buffer( uint32_t new_max );
    // swap_buffers (particles.h:195-206): exchanges the particle arrays, the
    // counts and offsets, and the sizes
    friend void swap_buffers( Particles & a, Particles & b );
    std::vector<int> get_np() const;
    std::vector<int> get_offset() const;
};

class ParticleSort {
  public:
    zpic_sort buf;
    const int ntiles;
    ParticleSort( uint2 const ntiles, uint32_t const max_part );
    ~ParticleSort();
    // leaver list sized for max_part particles (grown with the particle buffer)
    void grow( uint32_t max_part );
    uint32_t idx_size;
    ParticleSort( const ParticleSort & ) = delete;
    ParticleSort & operator=( const ParticleSort & ) = delete;
};

// --------------------------------------------------------- UDistribution
namespace UDistribution {
class Type {
  public:
    virtual Type * clone() const = 0;
    // Fills the momenta of particles sorted by tile (host arrays), seed = species id.
    virtual void set( int ntiles, const int * np, const int * offset, const int2 * ix, uint2 nx, float3 * u,
                      unsigned int seed ) const = 0;
    virtual ~Type() = default;
};
class None : public Type {
  public:
    None * clone() const override { return new None(); }
    void set( int, const int *, const int *, const int2 *, uint2, float3 *, unsigned int ) const override;
};
class Cold : public Type {
  public:
    const float3 ufl;
    Cold( float3 const ufl_ ) : ufl( ufl_ ) {}
    Cold * clone() const override { return new Cold( ufl ); }
    void set( int, const int *, const int *, const int2 *, uint2, float3 *, unsigned int ) const override;
};
class Thermal : public Type {
  public:
    const float3 uth, ufl;
    Thermal( float3 const uth_, float3 const ufl_ ) : uth( uth_ ), ufl( ufl_ ) {}
    Thermal * clone() const override { return new Thermal( uth, ufl ); }
    void set( int, const int *, const int *, const int2 *, uint2, float3 *, unsigned int ) const override;
};
class ThermalCorr : public Type {
  public:
    const float3 uth, ufl;
    const int npmin;
    ThermalCorr( float3 const uth_, float3 const ufl_, int const npmin_ = 2 ) : uth( uth_ ), ufl( ufl_ ), npmin( npmin_ ) {
        if ( npmin <= 1 ) throw std::runtime_error( "(*error*) invalid npmin parameter, must be > 1" );
    }
    ThermalCorr * clone() const override { return new ThermalCorr( uth, ufl, npmin ); }
    void set( int, const int *, const int *, const int2 *, uint2, float3 *, unsigned int ) const override;
};
}  // namespace UDistribution

// --------------------------------------------------------------- Density
// Density profiles (density.h:11-132): the density n0 sets the charge per
// particle; the profile selects which points of the regular injection
// lattice are kept.  desc() gives the selection as plain data with lengths
// relative to the window origin `ref`, as the reference's inject /
// np_inject compute them (density.cpp:415-442, 674-718, 950-967): a step or
// a slab along either axis is placed with ref.x and dx.x.
namespace Density {
class Profile {
  public:
    const float n0;
    Profile( float const n0_ ) : n0( n0_ < 0 ? -n0_ : n0_ ) {}
    virtual Profile * clone() const = 0;
    virtual zpic_inject_profile desc( float2 const dx, float2 const ref ) const = 0;
    virtual ~Profile() = default;
};
class None : public Profile {
  public:
    None() : Profile( 0.f ) {}
    None * clone() const override { return new None(); }
    zpic_inject_profile desc( float2 const dx, float2 const ref ) const override;
};
class Uniform : public Profile {
  public:
    Uniform( float const n0_ ) : Profile( n0_ ) {}
    Uniform * clone() const override { return new Uniform( n0 ); }
    zpic_inject_profile desc( float2 const dx, float2 const ref ) const override;
};
class Step : public Profile {
  public:
    const float pos;
    const coord::cart dir;
    Step( coord::cart dir_, float const n0_, float const pos_ ) : Profile( n0_ ), pos( pos_ ), dir( dir_ ) {}
    Step * clone() const override { return new Step( dir, n0, pos ); }
    zpic_inject_profile desc( float2 const dx, float2 const ref ) const override;
};
class Slab : public Profile {
  public:
    const float begin, end;
    const coord::cart dir;
    Slab( coord::cart dir_, float const n0_, float begin_, float end_ )
        : Profile( n0_ ), begin( begin_ ), end( end_ ), dir( dir_ ) {}
    Slab * clone() const override { return new Slab( dir, n0, begin, end ); }
    zpic_inject_profile desc( float2 const dx, float2 const ref ) const override;
};
// No reference counterpart: density rising linearly from 0 at `begin` to n0
// at `end` along dir, n0 beyond (equal-charge particles, thinned lattice)
class Ramp : public Profile {
  public:
    const float begin, end;
    const coord::cart dir;
    Ramp( coord::cart dir_, float const n0_, float begin_, float end_ )
        : Profile( n0_ ), begin( begin_ ), end( end_ ), dir( dir_ ) {
        if ( !( end > begin ) ) throw std::runtime_error( "(*error*) Invalid ramp, end must be > begin" );
    }
    Ramp * clone() const override { return new Ramp( dir, n0, begin, end ); }
    zpic_inject_profile desc( float2 const dx, float2 const ref ) const override;
};
class Sphere : public Profile {
  public:
    const float2 center;
    const float radius;
    Sphere( float const n0_, float2 center_, float radius_ ) : Profile( n0_ ), center( center_ ), radius( radius_ ) {}
    Sphere * clone() const override { return new Sphere( n0, center, radius ); }
    zpic_inject_profile desc( float2 const dx, float2 const ref ) const override;
};

// Host injection of a profile over the global cell range `range`
// ({x lower, x upper, y lower, y upper}, inclusive): particles per tile, and
// their cells and positions at [offset[t], offset[t] + np[t]).
void np_inject_host( zpic_inject_profile const & d, uint2 ntiles, uint2 nx, uint2 ppc, const int range[4], int * np );
void inject_host( zpic_inject_profile const & d, uint2 ntiles, uint2 nx, uint2 ppc, const int range[4],
                  const int * offset, int2 * ix, float2 * x );
}  // namespace Density

namespace species {
enum pusher { boris, euler };
namespace bc {
enum type { open = 0, periodic, reflecting };
}
typedef bnd<bc::type> bc_type;
}  // namespace species

namespace emf {
enum field { e = 0, b };
namespace bc {
// absorbing: no reference counterpart (damping layer, zpic_emf_absorb)
enum type { none = 0, periodic, pec, pmc, absorbing };
}
typedef bnd<bc::type> bc_type;
}  // namespace emf

namespace current {
namespace bc {
enum type { none = 0, periodic, reflecting };
}
typedef bnd<bc::type> bc_type;
}  // namespace current

// ------------------------------------------------------------------ Current
class Current {
    float2 dx;
    double dt;
    current::bc_type bc;
    int iter;
    // Second buffer of the one-pass update (zpic_current_update): on periodic
    // grids with no filter or a filter along x, advance() forms the new
    // current from J's deposit into this grid's buffer in one launch, then
    // J and this grid exchange buffers (allocated on first use).
    vec3grid<float3> * Jalt;
    // whether advance() may run as one pass, and the filter's x passes
    bool one_pass_ok( std::vector<float> & coef ) const;

  public:
    vec3grid<float3> * J;
    const float2 box;
    Filter::Digital * filter;
    MovingWindow moving_window;

    Current( uint2 const ntiles, uint2 const nx, float2 const box, double const dt );
    ~Current();
    int get_iter() const { return iter; }
    double get_dt() const { return dt; }
    current::bc_type get_bc() const { return bc; }
    void set_bc( current::bc_type new_bc );
    void set_filter( Filter::Digital const & new_filter ) {
        delete filter;
        filter = new_filter.clone();
    }
    void advance();
    void zero();
    // Current::set_moving_window (current.h:123-136): 0, or -1 after the first step
    int set_moving_window();
    // Current::save (current.cpp:263-313): CURRENT/J<c>-<iter>.zdf
    void save( fcomp::cart const jc );
    void note_step() { iter++; }
    float2 get_dx() const { return dx; }
};

// ---------------------------------------------------------------------- EMF
class EMF {
    emf::bc_type bc;
    const float2 dx;
    const double dt;
    int iter;
    double * d_energy;   // [6*ntiles + 6] device scratch

  public:
    vec3grid<float3> * E;
    vec3grid<float3> * B;
    const float2 box;
    MovingWindow moving_window;

    EMF( uint2 const ntiles, uint2 const nx, float2 const box, double const dt );
    ~EMF();
    int get_iter() const { return iter; }
    double get_dt() const { return dt; }
    emf::bc_type get_bc() const { return bc; }
    void set_bc( emf::bc_type new_bc );
    void advance();
    void advance( Current & current );
    void get_energy( double3 & ene_E, double3 & ene_B );
    // EMF::process_bc (emf.cpp:431-471), then the absorbing layers
    void process_bc();
    // Width (cells) and strength of absorbing layers (no reference counterpart)
    void set_absorbing( int width, float alpha );
    int abs_width = 20;
    float abs_alpha = 0.015f;
    // EMF::set_moving_window (emf.h:152-165): 0, or -1 after the first step
    int set_moving_window();
    // EMF::move_window (emf.cpp:120-129): E and B shifted one cell left
    // when the window is due to move
    void move_window();
    // EMF::save (emf.cpp:589-655): EMF/<E|B><c>-<iter>.zdf
    void save( emf::field const field, fcomp::cart const fc );
    void note_step() { iter++; }
    float2 get_dx() const { return dx; }
};

// ------------------------------------------------------------------ Species
class Species {
  protected:
    int id;
    uint2 ppc;
    float q;
    float2 dx;
    float2 box;
    float dt;
    int iter;
    Particles * particles;
    Particles * tmp;
    ParticleSort * sort;
    Density::Profile * density;

  private:
    species::bc_type bc;
    UDistribution::Type * udist;
    double * d_tile_energy;  // [ntiles] per-tile push energy sums
    double * d_energy_dev;   // [1]
    uint64_t d_nmove;
    // Upper bound of the particles held: set by every load, raised by the
    // largest possible count of every injection, lowered to the exact total
    // whenever that is read back (particles leave only through open walls)
    uint64_t np_bound;
    int * d_np_inj;          // [ntiles] particles to inject (moving window)
    bool deterministic;
    unsigned seed_offset;

  public:
    const std::string name;
    const float m_q;
    species::pusher push_type;
    MovingWindow moving_window;

    Species( std::string const name, float const m_q, uint2 const ppc );
    virtual ~Species();
    virtual void initialize( float2 const box, uint2 const ntiles, uint2 const nx, double const dt, int const id );

    virtual void set_density( Density::Profile const & d ) { delete density; density = d.clone(); }
    Density::Profile & get_density() { return *density; }
    virtual void set_udist( UDistribution::Type const & u ) { delete udist; udist = u.clone(); }
    UDistribution::Type & get_udist() { return *udist; }
    void set_bc( species::bc_type new_bc );
    species::bc_type get_bc() const { return bc; }
    // Offset added to the species id in the momentum RNG seed (0: the
    // reference's seeding, udist.cpp:73,94); must be set before initialize
    void set_seed( unsigned offset ) { seed_offset = offset; }

    void push( vec3grid<float3> * const E, vec3grid<float3> * const B );
    void move( vec3grid<float3> * const J );
    void tile_sort();
    // Particles::tile_sort(tmp, sort, extra) (particles.cpp:855-904): room
    // for extra[t] (device array) more particles after each tile; max_extra
    // bounds their sum.  Takes the copy_sorted path, growing the buffers,
    // when the total exceeds the particle buffer, as the reference does.
    void tile_sort( const int * extra, uint64_t max_extra );
    virtual void advance( EMF const & emf, Current & current );
    // advance() (window: advance_mov_window()) recording `after_push` (may be
    // null) between the particle pass and the sort
    // set_j (fast mode): the particle pass writes J's tile regions instead
    // of adding to them, for the first species of a step whose J was not
    // zeroed
    void advance_timed( EMF const & emf, Current & current, hipEvent_t after_push, bool window = false,
                        bool set_j = false );
    // Species::advance_mov_window (species.cpp:2083-2128): on the steps where
    // the window moves, the particles shift one cell left and the density
    // profile is injected into the rightmost column of cells
    virtual void advance_mov_window( EMF const & emf, Current & current );
    // Species::set_moving_window (species.h:254-265): 0, or -1 after the first step
    virtual int set_moving_window();
    // Species::np_inject / inject over a global cell range
    // ({x lower, x upper, y lower, y upper}) on the device (species.cpp:1819-1842)
    void np_inject( const int range[4], int * d_np );
    void inject( const int range[4] );
    // The particles replaced by a device injection of the density profile
    // over the whole box (np_inject, exclusive scan, inject; momenta 0): the
    // same particles a cold initial load places, by the window's kernels
    void reinject();

    // The two halves of advance(): the particle pass (push, move, deposit,
    // leaver classification; with the window's cell shift when `shift`),
    // and the sort, the injection when `shift`, the iteration count and the
    // energy sum.
    void particle_pass( EMF const & emf, Current & current, bool shift, bool set_j = false );
    void set_staging( zpic_sort & s, bool window = false ) const;
    void finish_step( bool shift );
    // Whether this step shifts the particles (the window moves), for a step
    // of advance_mov_window (window) or advance (not window)
    bool shifts( bool window ) const;

    // Deterministic (reference single-thread) deposit / energy order
    void set_deterministic( bool on ) { deterministic = on; }
    bool is_deterministic() const { return deterministic; }

    void deposit_charge( grid<float> & charge ) const;
    // Diagnostics (species.cpp:2417-3100): PARTICLES/<name>/particles-<name>-<iter>.zdf,
    // CHARGE/<name>/<name>-charge-<iter>.zdf and PHASESPACE/<name>/<name>-<q>-<iter>.zdf
    void save() const;
    void save_charge() const;
    // Phase-space density into the device buffer d_data (size, or size0*size1, floats)
    void dep_phasespace( float * const d_data, phasespace::quant q, float2 range, unsigned const size ) const;
    void dep_phasespace( float * const d_data, phasespace::quant quant0, float2 range0, unsigned const size0,
                         phasespace::quant quant1, float2 range1, unsigned const size1 ) const;
    void save_phasespace( phasespace::quant quant, float2 const range, int const size ) const;
    void save_phasespace( phasespace::quant quant0, float2 const range0, int const size0, phasespace::quant quant1,
                          float2 const range1, int const size1 ) const;
    double get_energy() const;
    // particles moved so far (device counter; synchronises)
    uint64_t get_nmove() const;
    // reflecting walls of set_bc as the particle pass's mask
    int reflect_mask() const;
    int get_iter() const { return iter; }
    float get_dt() const { return dt; }
    uint2 get_ntiles() const { return particles->ntiles; }
    uint2 get_nx() const { return particles->nx; }
    float2 get_box() const { return box; }
    uint64_t np_total() const { return particles->np_total(); }
    float get_q() const { return q; }
    float2 get_dx() const { return dx; }
    uint2 get_ppc() const { return ppc; }

    // Host copies in tile order (compact)
    void get_particles( int * np, int * offset, int * ix, float * x, float * u ) const;
    void set_particles( const int * np, const int * ix, const float * x, const float * u );
    // Every stride-th particle of each of src's tiles (same tiling), from index phase
    void subsample_from( const Species & src, int stride, int phase );
    // Species::gather (species.h:515): x, y in simulation units, or ux, uy, uz
    void gather( part::quant quant, float * host_out ) const;

    // Device access for callers that drive the stages themselves
    Particles * get_particles_buffer() { return particles; }

    // Arguments of this species for a batched advance with the given fields
    zpic_batch_species batch_item( EMF const & emf, Current & current ) const;
    // device particle buffers (they change when the buffers are swapped)
    void buffers( std::vector<const void *> & out ) const {
        out.push_back( particles->ix );
        out.push_back( tmp->ix );
    }
    // Host-side bookkeeping of one step advanced by a batch (iteration, nmove)
    void note_step() { iter++; }

  private:
    void step( EMF const & emf, Current & current, hipEvent_t after_push, bool window, bool set_j = false );
};

// --------------------------------------------------------------- Simulation
class Simulation {
    unsigned int iter;
    bool timing;
    // stage timing without a host wait per step: the events of the last few
    // steps, read when their slot comes round again or in get_timing()
    static constexpr int TSLOTS = 3;
    std::vector<hipEvent_t> ev[TSLOTS];
    bool ev_pending[TSLOTS] = { false, false, false };
    bool ev_overlap[TSLOTS] = { false, false, false };
    size_t ev_nsp[TSLOTS] = { 0, 0, 0 };  // species count of the recorded step
    // ordering events between the two streams (no timing)
    hipEvent_t dep[2] = { nullptr, nullptr };
    long long t_ring;
    double t_acc[5];
    long long t_steps;
    void accumulate( int slot );

  public:
    const uint2 ntiles;
    const uint2 nx;
    const float2 box;
    const double dt;
    EMF emf;
    Current current;
    std::vector<Species *> species;

    Simulation( uint2 const ntiles, uint2 const nx, float2 const box, double dt );
    ~Simulation();

    void add_species( Species & s ) {
        species.push_back( &s );
        s.initialize( box, ntiles, nx, dt, species.size() );
    }
    Species * get_species( std::string name ) {
        for ( auto * s : species ) if ( s->name == name ) return s;
        return nullptr;
    }
    void advance();
    // Simulation::set_moving_window / advance_mov_window (simulation.h:56-62, 113-127)
    void set_moving_window();
    void advance_mov_window();
    unsigned int get_iter() const { return iter; }
    double get_t() const { return iter * double( dt ); }
    void energy_info();
    uint64_t get_nmove() const {
        uint64_t n = 0;
        for ( auto * s : species ) n += s->get_nmove();
        return n;
    }
    void set_deterministic( bool on ) { for ( auto * s : species ) s->set_deterministic( on ); }
    // HIP-event timing of the stages (push+move, sort, current, yee, step)
    void set_timing( bool on );
    void get_timing( double out[5] );
    // Host-side bookkeeping of one step advanced by a batch
    void note_step();
    // Arguments of this simulation's fields for a batched advance
    zpic_batch_fields_item batch_fields() const;

  private:
    void step( bool window );
};

// ------------------------------------------------------------------- Batch
// Many independent simulations advanced together: every stage of
// Simulation::advance is one launch over the tiles of all instances, and the
// whole step can be captured once in a hipGraph and replayed.  Instances keep
// their own buffers; each computes exactly what its own Simulation::advance
// computes with the same threads per tile (fast mode).  No reference
// counterpart (the reference advances one simulation per process).
class Batch {
    std::vector<Simulation *> sims;
    std::vector<int> sim_threads;               // per instance, 0 = `threads`
    std::vector<zpic_batch_particles *> slots;  // one plan per species slot
    zpic_batch_fields * fields;
    zpic_batch_step * persistent_step;          // the whole step in one launch
    int threads;
    bool use_graph;
    bool persistent;
    hipGraph_t graph;
    hipGraphExec_t exec;
    std::vector<const void *> bound;  // device buffers the tables were built on
    void release();
    void step_kernels();
    std::vector<const void *> buffers() const;

  public:
    explicit Batch( int threads_per_tile = 256 );
    ~Batch();
    Batch( const Batch & ) = delete;
    Batch & operator=( const Batch & ) = delete;

    // Non-owning; all instances must hold the same number of species.
    // threads_per_tile: this instance's particle pass (0 = the batch's).
    void add( Simulation & sim, int threads_per_tile = 0 );
    // Builds the launch tables (called by the first advance if needed).
    void build();
    void advance();
    // Capture the step in a hipGraph on the first advance and replay it.
    void set_graph( bool on );
    // One persistent launch per step (zpic_batch_step) instead of one launch
    // per stage (periodic boundaries).
    void set_persistent( bool on );
    // Synchronises and throws if a persistent step reported an error.
    void check_status();
    size_t size() const { return sims.size(); }
};

// -------------------------------------------------------------------- Laser
// Laser pulses (laser.h, laser.cpp:172-437).  launch() fills E and B with
// the pulse on the host, in the reference's float / double arithmetic and
// with the host math library (as the reference, which builds it on the CPU),
// then the guard copy, the compensated x filter (Pulse::filter passes, if
// > 0), and for the Gaussian beam the divergence correction (div_corr_x) run
// on the device; add() launches into temporary grids and adds them to the
// simulation fields.
namespace Laser {
class Pulse {
  public:
    float start;   // front edge of the pulse
    float fwhm;    // if > 0: rise = fall = fwhm, flat = 0
    float rise, flat, fall;
    float a0;      // normalised peak vector potential
    float omega0;  // frequency
    float polarization;
    float cos_pol, sin_pol;  // both 0: taken from `polarization`
    unsigned int filter;

    Pulse()
        : start( 0 ), fwhm( 0 ), rise( 0 ), flat( 0 ), fall( 0 ), a0( 0 ), omega0( 0 ), polarization( 0 ), cos_pol( 0 ),
          sin_pol( 0 ), filter( 1 ) {}
    virtual ~Pulse() = default;
    // longitudinal envelope (laser.h:38-63)
    float lon_env( const float z ) const;
    // 0, or -1 for invalid parameters (laser.cpp:178-210)
    virtual int validate();
    virtual int launch( vec3grid<float3> & E, vec3grid<float3> & B, float2 box ) = 0;
    // Pulse::add (laser.h:67-81): 0, or -1 for invalid parameters
    virtual int add( EMF & emf );
};

class PlaneWave : public Pulse {
  public:
    int launch( vec3grid<float3> & E, vec3grid<float3> & B, float2 box ) override;
};

class Gaussian : public Pulse {
  public:
    float W0;     // waist
    float focus;  // focal plane position (x)
    float axis;   // beam axis position (y)
    Gaussian() : Pulse(), W0( 0 ), focus( 0 ), axis( 0 ) {}
    int validate() override;
    int launch( vec3grid<float3> & E, vec3grid<float3> & B, float2 box ) override;
};
}  // namespace Laser
