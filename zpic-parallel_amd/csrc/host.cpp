// Host side of the zpic C++ API (zpic.h): object lifetimes, device buffers,
// initial particle load, and the call sequence of the reference timestep,
// each stage dispatched to a HIP kernel through the stage entry points of
// zpic_amd.h.  The initial load (density injection and momentum
// distribution) runs on the host exactly as the reference does, one worker
// per tile range, and is uploaded once; every per-step stage runs on the GPU.
#include "common.h"
#include "zpic.h"
#include "density.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <fstream>
#include <iostream>
#include <limits>
#include <mutex>
#include <thread>
#include <cstdlib>

namespace zpic {

namespace {
thread_local hipStream_t t_stream_override = nullptr;
}

StreamScope::StreamScope( hipStream_t s ) : prev( t_stream_override ) { t_stream_override = s; }
StreamScope::~StreamScope() { t_stream_override = prev; }

hipStream_t side_stream() {
    static hipStream_t s = nullptr;
    static std::once_flag once;
    std::call_once( once, [] {
        if ( hipStreamCreateWithFlags( &s, hipStreamNonBlocking ) != hipSuccess ) s = nullptr;
    } );
    if ( !s ) throw std::runtime_error( "(*error*) no HIP device available: the zpic library needs an MI355X GPU" );
    return s;
}

hipStream_t stream() {
    if ( t_stream_override ) return t_stream_override;
    static hipStream_t s = nullptr;
    static std::once_flag once;
    std::call_once( once, [] {
        if ( hipStreamCreateWithFlags( &s, hipStreamNonBlocking ) != hipSuccess ) s = nullptr;
    } );
    if ( !s ) throw std::runtime_error( "(*error*) no HIP device available: the zpic library needs an MI355X GPU" );
    return s;
}

hipStream_t default_stream_or_null() {
    try {
        return stream();
    } catch ( ... ) {
        return nullptr;
    }
}

void check( int rc ) {
    if ( rc != ZPIC_OK ) throw std::runtime_error( zpic_last_error() );
}

void hip_check( hipError_t e, const char * what ) {
    if ( e != hipSuccess ) throw std::runtime_error( std::string( what ) + ": " + hipGetErrorString( e ) );
}

namespace {

template <class T>
void to_device( T * dst, const T * src, size_t n ) {
    if ( n ) hip_check( hipMemcpyAsync( dst, src, n * sizeof( T ), hipMemcpyHostToDevice, stream() ), "upload" );
}
template <class T>
void to_host( T * dst, const T * src, size_t n ) {
    if ( n ) hip_check( hipMemcpyAsync( dst, src, n * sizeof( T ), hipMemcpyDeviceToHost, stream() ), "download" );
}
void sync() { hip_check( hipStreamSynchronize( stream() ), "hipStreamSynchronize" ); }

// Runs fn(t) for t in [0, n) on the host worker threads.
// Worker count: ZPIC_HOST_THREADS if set, else the hardware threads capped at 16.
template <class F>
void parallel_tiles( int n, F && fn ) {
    int nw = (int) std::thread::hardware_concurrency();
    if ( nw > 16 ) nw = 16;
    if ( const char * env = std::getenv( "ZPIC_HOST_THREADS" ) ) nw = std::atoi( env );
    if ( nw < 1 ) nw = 1;
    if ( nw > n ) nw = n;
    if ( nw <= 1 ) { for ( int t = 0; t < n; t++ ) fn( t ); return; }
    std::vector<std::thread> th;
    for ( int w = 0; w < nw; w++ )
        th.emplace_back( [&, w] { for ( int t = w; t < n; t += nw ) fn( t ); } );
    for ( auto & x : th ) x.join();
}

// Marsaglia multiply-with-carry generator and polar Box-Muller normal
// deviates (zrandom, random.h:23-155); one independent stream per tile.
class TileRng {
    uint32_t a, b;
    double spare;
    static constexpr double none = std::numeric_limits<double>::max();

  public:
    TileRng( unsigned tile, uint32_t sa, uint32_t sb )
        : a( sa + tile ), b( sb + ( tile ^ 0xFFFFFFFFu ) ), spare( none ) {}
    uint32_t bits() {
        a = 36969u * ( a & 0xFFFFu ) + ( a >> 16 );
        b = 18000u * ( b & 0xFFFFu ) + ( b >> 16 );
        return ( a << 16 ) + b;
    }
    double normal() {
        if ( spare != none ) {
            const double r = spare;
            spare = none;
            return r;
        }
        double p, q, s;
        do {
            p = ( bits() + 1.0 ) * 0x1.fffffffep-33 * 2.0 - 1.0;
            q = ( bits() + 1.0 ) * 0x1.fffffffep-33 * 2.0 - 1.0;
            s = p * p + q * q;
        } while ( s == 0.0 || s >= 1.0 );
        const double f = std::sqrt( -2.0 * std::log( s ) / s );
        spare = p * f;
        return q * f;
    }
};

}  // namespace
}  // namespace zpic

using namespace zpic;

// ------------------------------------------------------------------ Particles
Particles::Particles( uint2 const ntiles_, uint2 const nx_, uint32_t const max_part_ )
    : ntiles( ntiles_ ), nx( nx_ ), periodic( make_int2( 1, 1 ) ),
      dims( make_uint2( ntiles_.x * nx_.x, ntiles_.y * nx_.y ) ), max_part( max_part_ ) {
    const size_t nt = (size_t) ntiles.x * ntiles.y;
    np = device_alloc<int>( nt );
    offset = device_alloc<int>( nt );
    ix = device_alloc<int2>( max_part );
    x = device_alloc<float2>( max_part );
    u = device_alloc<float3>( max_part );
    hip_check( hipMemsetAsync( np, 0, nt * sizeof( int ), stream() ), "memset" );
    hip_check( hipMemsetAsync( offset, 0, nt * sizeof( int ), stream() ), "memset" );
}

Particles::~Particles() {
    device_free( u ); device_free( x ); device_free( ix );
    device_free( offset ); device_free( np );
}

zpic_particles Particles::view() const {
    zpic_particles v;
    v.np = np; v.offset = offset;
    v.ix = reinterpret_cast<int *>( ix );
    v.x = reinterpret_cast<float *>( x );
    v.u = reinterpret_cast<float *>( u );
    v.max_part = max_part;
    return v;
}

std::vector<int> Particles::get_np() const {
    std::vector<int> h( ntiles.x * ntiles.y );
    to_host( h.data(), np, h.size() );
    sync();
    return h;
}
std::vector<int> Particles::get_offset() const {
    std::vector<int> h( ntiles.x * ntiles.y );
    to_host( h.data(), offset, h.size() );
    sync();
    return h;
}
uint64_t Particles::np_total() const {
    uint64_t s = 0;
    for ( int v : get_np() ) s += v;
    return s;
}

void Particles::grow_buffer( uint32_t new_max ) {
    if ( new_max <= max_part ) return;
    sync();  // no kernel may still use the old arrays
    device_free( u ); device_free( x ); device_free( ix );
    u = nullptr; x = nullptr; ix = nullptr;
    max_part = ( new_max + 65535u ) & ~65535u;
    ix = device_alloc<int2>( max_part );
    x = device_alloc<float2>( max_part );
    u = device_alloc<float3>( max_part );
}

void swap_buffers( Particles & a, Particles & b ) {
    std::swap( a.ix, b.ix );
    std::swap( a.x, b.x );
    std::swap( a.u, b.u );
    std::swap( a.max_part, b.max_part );
    std::swap( a.np, b.np );
    std::swap( a.offset, b.offset );
}

ParticleSort::ParticleSort( uint2 const nt2, uint32_t const max_part )
    : buf{}, ntiles( nt2.x * nt2.y ), idx_size( max_part ) {
    buf.idx = device_alloc<int>( max_part );
    buf.npt = device_alloc<int>( 9 * (size_t) ntiles );
    buf.new_np = device_alloc<int>( ntiles );
    buf.seg = device_alloc<int>( 16 * (size_t) ntiles );
    buf.tmp_np = device_alloc<int>( ntiles );
    buf.total = device_alloc<int>( 1 );
    buf.nmove = device_alloc<unsigned long long>( 1 );
    hip_check( hipMemsetAsync( buf.nmove, 0, sizeof( unsigned long long ), stream() ), "hipMemsetAsync" );
}
void ParticleSort::grow( uint32_t max_part ) {
    if ( max_part <= idx_size ) return;
    sync();
    device_free( buf.idx );
    buf.idx = nullptr;
    buf.idx = device_alloc<int>( max_part );
    idx_size = max_part;
}

ParticleSort::~ParticleSort() {
    device_free( buf.nmove );
    device_free( buf.total ); device_free( buf.tmp_np ); device_free( buf.seg );
    device_free( buf.new_np ); device_free( buf.npt ); device_free( buf.idx );
}

// ----------------------------------------------------------------- Density
zpic_inject_profile Density::None::desc( float2 const dx, float2 const ) const {
    zpic_inject_profile d{};
    d.kind = ZPIC_DENSITY_NONE;
    d.dx = dx.x; d.dy = dx.y;
    return d;
}

zpic_inject_profile Density::Uniform::desc( float2 const dx, float2 const ) const {
    zpic_inject_profile d{};
    d.kind = ZPIC_DENSITY_UNIFORM;
    d.dx = dx.x; d.dy = dx.y;
    return d;
}

// Density::Step::inject / np_inject (density.cpp:415-442, 528-573):
// step_pos = (pos - ref.x) / dx.x
zpic_inject_profile Density::Step::desc( float2 const dx, float2 const ref ) const {
    zpic_inject_profile d{};
    d.kind = ZPIC_DENSITY_STEP;
    d.dir = ( dir == coord::x ) ? 0 : 1;
    d.start = ( pos - ref.x ) / dx.x;
    d.dx = dx.x; d.dy = dx.y;
    return d;
}

// Density::Slab::inject / np_inject (density.cpp:674-718, 807-854)
zpic_inject_profile Density::Slab::desc( float2 const dx, float2 const ref ) const {
    zpic_inject_profile d{};
    d.kind = ZPIC_DENSITY_SLAB;
    d.dir = ( dir == coord::x ) ? 0 : 1;
    d.start = ( begin - ref.x ) / dx.x;
    d.end = ( end - ref.x ) / dx.x;
    d.dx = dx.x; d.dy = dx.y;
    return d;
}

// Density::Sphere::inject / np_inject (density.cpp:950-967, 1050-1067)
zpic_inject_profile Density::Sphere::desc( float2 const dx, float2 const ref ) const {
    zpic_inject_profile d{};
    d.kind = ZPIC_DENSITY_SPHERE;
    d.center_x = center.x - ref.x;
    d.center_y = center.y - ref.y;
    d.radius = radius;
    d.dx = dx.x; d.dy = dx.y;
    return d;
}

// Density::Ramp (no reference counterpart): along x placed like a step
// (window origin subtracted), along y in y cells
zpic_inject_profile Density::Ramp::desc( float2 const dx, float2 const ref ) const {
    zpic_inject_profile d{};
    d.kind = ZPIC_DENSITY_RAMP;
    d.dir = ( dir == coord::x ) ? 0 : 1;
    if ( dir == coord::x ) {
        d.start = ( begin - ref.x ) / dx.x;
        d.end = ( end - ref.x ) / dx.x;
    } else {
        d.start = ( begin - ref.y ) / dx.y;
        d.end = ( end - ref.y ) / dx.y;
    }
    d.dx = dx.x; d.dy = dx.y;
    return d;
}

// The lattice points of the range kept by the profile, per tile, in the
// reference's order (lattice point i1, i0, then cell in the clipped range).
void Density::np_inject_host( zpic_inject_profile const & d, uint2 ntiles, uint2 nx, uint2 ppc, const int range[4],
                              int * np ) {
    parallel_tiles( ntiles.x * ntiles.y, [&]( int t ) {
        const int tx = t % ntiles.x, ty = t / ntiles.x;
        TileRange r;
        int n = 0;
        const int npc = ppc.x * ppc.y, stride = ramp_stride( npc );
        if ( d.kind != ZPIC_DENSITY_NONE && tile_range( range, tx, ty, nx.x, nx.y, r ) ) {
            for ( unsigned i1 = 0; i1 < ppc.y; i1++ )
                for ( unsigned i0 = 0; i0 < ppc.x; i0++ ) {
                    const float px = lattice_pos( i0, ppc.x ), py = lattice_pos( i1, ppc.y );
                    const int k = i1 * ppc.x + i0;
                    for ( int c = 0; c < r.vol; c++ )
                        n += profile_keeps( d, tx * nx.x + c % r.row + r.i0, ty * nx.y + c / r.row + r.j0, px, py, k,
                                            npc, stride );
                }
        }
        np[t] = n;
    } );
}

void Density::inject_host( zpic_inject_profile const & d, uint2 ntiles, uint2 nx, uint2 ppc, const int range[4],
                           const int * offset, int2 * ix, float2 * x ) {
    parallel_tiles( ntiles.x * ntiles.y, [&]( int t ) {
        const int tx = t % ntiles.x, ty = t / ntiles.x;
        TileRange r;
        if ( d.kind == ZPIC_DENSITY_NONE || !tile_range( range, tx, ty, nx.x, nx.y, r ) ) return;
        int n = offset[t];
        const int npc = ppc.x * ppc.y, stride = ramp_stride( npc );
        for ( unsigned i1 = 0; i1 < ppc.y; i1++ )
            for ( unsigned i0 = 0; i0 < ppc.x; i0++ ) {
                const float2 pos = make_float2( lattice_pos( i0, ppc.x ), lattice_pos( i1, ppc.y ) );
                const int k = i1 * ppc.x + i0;
                for ( int c = 0; c < r.vol; c++ ) {
                    const int2 cell = make_int2( c % r.row + r.i0, c / r.row + r.j0 );
                    if ( profile_keeps( d, tx * nx.x + cell.x, ty * nx.y + cell.y, pos.x, pos.y, k, npc, stride ) ) {
                        ix[n] = cell;
                        x[n] = pos;
                        n++;
                    }
                }
            }
    } );
}

// --------------------------------------------------------- UDistribution
// (udist.cpp:17-136).  Normal deviates of one particle are drawn in the order
// z, y, x, the evaluation order of the reference build's make_float3 call.
void UDistribution::None::set( int nt, const int * np, const int * off, const int2 *, uint2, float3 * u,
                               unsigned ) const {
    for ( int t = 0; t < nt; t++ )
        for ( int i = 0; i < np[t]; i++ ) u[off[t] + i] = make_float3( 0, 0, 0 );
}

void UDistribution::Cold::set( int nt, const int * np, const int * off, const int2 *, uint2, float3 * u,
                               unsigned ) const {
    for ( int t = 0; t < nt; t++ )
        for ( int i = 0; i < np[t]; i++ ) u[off[t] + i] = ufl;
}

void UDistribution::Thermal::set( int nt, const int * np, const int * off, const int2 *, uint2, float3 * u,
                                  unsigned seed ) const {
    parallel_tiles( nt, [&]( int t ) {
        TileRng rng( t, 12345u + seed, 67890u );
        for ( int i = 0; i < np[t]; i++ ) {
            const double nz = rng.normal(), ny = rng.normal(), nxv = rng.normal();
            u[off[t] + i] = make_float3( (float) ( ufl.x + uth.x * nxv ), (float) ( ufl.y + uth.y * ny ),
                                         (float) ( ufl.z + uth.z * nz ) );
        }
    } );
}

void UDistribution::ThermalCorr::set( int nt, const int * np, const int * off, const int2 * ix, uint2 nx, float3 * u,
                                      unsigned seed ) const {
    const int cells = nx.x * nx.y;
    parallel_tiles( nt, [&]( int t ) {
        TileRng rng( t, 12345u + seed, 67890u );
        std::vector<float> sx( cells, 0.f ), sy( cells, 0.f ), sz( cells, 0.f );
        std::vector<int> cnt( cells, 0 );
        float3 * ut = u + off[t];
        const int2 * it = ix + off[t];
        for ( int i = 0; i < np[t]; i++ ) {
            const double nz = rng.normal(), ny = rng.normal(), nxv = rng.normal();
            const float3 v = make_float3( (float) ( uth.x * nxv ), (float) ( uth.y * ny ), (float) ( uth.z * nz ) );
            ut[i] = v;
            const int c = it[i].x + (int) nx.x * it[i].y;
            cnt[c] += 1;
            sx[c] += v.x; sy[c] += v.y; sz[c] += v.z;
        }
        for ( int c = 0; c < cells; c++ ) {
            if ( cnt[c] > npmin ) {
                sx[c] /= cnt[c]; sy[c] /= cnt[c]; sz[c] /= cnt[c];
            } else {
                sx[c] = sy[c] = sz[c] = 0.f;
            }
        }
        for ( int i = 0; i < np[t]; i++ ) {
            const int c = it[i].x + (int) nx.x * it[i].y;
            ut[i].x += ufl.x - sx[c];
            ut[i].y += ufl.y - sy[c];
            ut[i].z += ufl.z - sz[c];
        }
    } );
}

// ------------------------------------------------------------------ Current
Current::Current( uint2 const ntiles, uint2 const nx, float2 const box_, double const dt_ )
    : dx( make_float2( box_.x / ( nx.x * ntiles.x ), box_.y / ( nx.y * ntiles.y ) ) ), dt( dt_ ), iter( 0 ),
      Jalt( nullptr ), box( box_ ) {
    bnd<unsigned int> gc;
    gc.x = { 1, 2 };
    gc.y = { 1, 2 };
    J = new vec3grid<float3>( ntiles, nx, gc );
    J->name = "Current density";
    J->zero();
    bc = current::bc_type( current::bc::periodic );
    filter = new Filter::None();
}

Current::~Current() {
    delete J;
    delete Jalt;
    delete filter;
}

// Current::set_bc (current.h:84-114)
void Current::set_bc( current::bc_type new_bc ) {
    if ( ( new_bc.x.lower == current::bc::periodic || new_bc.x.upper == current::bc::periodic ) &&
         new_bc.x.lower != new_bc.x.upper )
        throw std::runtime_error( "(*error*) Current boundary type mismatch along x." );
    if ( ( new_bc.y.lower == current::bc::periodic || new_bc.y.upper == current::bc::periodic ) &&
         new_bc.y.lower != new_bc.y.upper )
        throw std::runtime_error( "(*error*) Current boundary type mismatch along y." );
    bc = new_bc;
    J->set_periodic( make_int2( bc.x.lower == current::bc::periodic, bc.y.lower == current::bc::periodic ) );
}

bool Current::one_pass_ok( std::vector<float> & coef ) const {
    coef.clear();
    const int2 p = J->get_periodic();
    if ( !p.x || !p.y ) return false;
    if ( filter->kind() == 1 || filter->kind() == 2 ) {
        const auto * f = dynamic_cast<const Filter::Binomial *>( filter );
        if ( !f || f->get_dir() != coord::x ) return false;
        coef = f->passes();
    } else if ( filter->kind() != 0 ) {
        return false;
    }
    const int passes = (int) coef.size() / 3;
    const unsigned gw = J->gc.x.lower + J->gc.x.upper;
    const size_t lds = sizeof( float ) * 6 * ( J->ext_nx.x + 2 * passes ) * J->ext_nx.y;
    return passes <= 16 && J->nx.x >= gw + passes && J->nx.y >= gw && lds <= 65536;
}

// Current::advance (current.cpp:227-246)
void Current::advance() {
    std::vector<float> coef;
    if ( one_pass_ok( coef ) ) {
        // guard add, guard copy and the x filter in one launch into the
        // second buffer, which then becomes J's
        if ( !Jalt ) {
            Jalt = new vec3grid<float3>( J->get_ntiles(), J->nx, J->gc );
            Jalt->name = J->name;
        }
        zpic_geom g = J->geom();
        check( zpic_current_update( J->fbuf(), Jalt->fbuf(), &g, (int) coef.size() / 3, coef.data(), stream() ) );
        std::swap( J->d_buffer, Jalt->d_buffer );
        iter++;
        if ( moving_window.needs_move( iter * dt ) ) moving_window.advance();
        return;
    }
    J->add_copy_gc();
    if ( bc.x.lower > current::bc::periodic || bc.x.upper > current::bc::periodic ||
         bc.y.lower > current::bc::periodic || bc.y.upper > current::bc::periodic ) {
        const int b[4] = { bc.x.lower, bc.x.upper, bc.y.lower, bc.y.upper };
        zpic_geom g = J->geom();
        check( zpic_current_bc( J->fbuf(), &g, b, stream() ) );
    }
    filter->apply( *J );
    iter++;
    // only the axis range of the current's diagnostics follows this window
    // (current.cpp:240-245)
    if ( moving_window.needs_move( iter * dt ) ) moving_window.advance();
}

int Current::set_moving_window() {
    if ( iter == 0 ) {
        moving_window.init( dx.x );
        bc.x.lower = bc.x.upper = current::bc::none;
        J->set_periodic( make_int2( 0, J->get_periodic().y ) );
        return 0;
    }
    std::cerr << "(*error*) set_moving_window() called with iter != 0\n";
    return -1;
}

void Current::zero() { J->zero(); }

// ---------------------------------------------------------------------- EMF
EMF::EMF( uint2 const ntiles, uint2 const nx, float2 const box_, double const dt_ )
    : dx( make_float2( box_.x / ( nx.x * ntiles.x ), box_.y / ( nx.y * ntiles.y ) ) ), dt( dt_ ), iter( 0 ),
      box( box_ ) {
    const float cour = std::sqrt( 1.0f / ( 1.0f / ( dx.x * dx.x ) + 1.0f / ( dx.y * dx.y ) ) );
    if ( dt >= cour )
        throw std::runtime_error( "(*error*) Invalid timestep, courant condition violation. For the current "
                                  "resolution the maximum timestep is dt = " + std::to_string( cour ) );
    bnd<unsigned int> gc;
    gc.x = { 1, 2 };
    gc.y = { 1, 2 };
    E = new vec3grid<float3>( ntiles, nx, gc );
    E->name = "Electric field";
    B = new vec3grid<float3>( ntiles, nx, gc );
    B->name = "Magnetic field";
    E->zero();
    B->zero();
    bc = emf::bc_type( emf::bc::periodic );
    d_energy = device_alloc<double>( 6 * (size_t) ntiles.x * ntiles.y + 6 );
}

EMF::~EMF() {
    device_free( d_energy );
    delete E;
    delete B;
}

// EMF::set_bc (emf.h:112-143)
void EMF::set_bc( emf::bc_type new_bc ) {
    if ( ( new_bc.x.lower == emf::bc::periodic || new_bc.x.upper == emf::bc::periodic ) &&
         new_bc.x.lower != new_bc.x.upper )
        throw std::runtime_error( "(*error*) EMF boundary type mismatch along x." );
    if ( ( new_bc.y.lower == emf::bc::periodic || new_bc.y.upper == emf::bc::periodic ) &&
         new_bc.y.lower != new_bc.y.upper )
        throw std::runtime_error( "(*error*) EMF boundary type mismatch along y." );
    bc = new_bc;
    const int2 p = make_int2( bc.x.lower == emf::bc::periodic, bc.y.lower == emf::bc::periodic );
    E->set_periodic( p );
    B->set_periodic( p );
}

// EMF::process_bc (emf.cpp:431-471)
void EMF::process_bc() {
    if ( bc.x.lower > emf::bc::periodic || bc.x.upper > emf::bc::periodic || bc.y.lower > emf::bc::periodic ||
         bc.y.upper > emf::bc::periodic ) {
        const int b[4] = { bc.x.lower, bc.x.upper, bc.y.lower, bc.y.upper };
        zpic_geom g = E->geom();
        check( zpic_emf_bc( E->fbuf(), B->fbuf(), &g, b, stream() ) );
        if ( b[0] == emf::bc::absorbing || b[1] == emf::bc::absorbing || b[2] == emf::bc::absorbing ||
             b[3] == emf::bc::absorbing )
            check( zpic_emf_absorb( E->fbuf(), B->fbuf(), &g, b, abs_width, abs_alpha, stream() ) );
    }
}

void EMF::set_absorbing( int width, float alpha ) {
    if ( width < 1 || !( alpha >= 0 ) )
        throw std::runtime_error( "(*error*) absorbing layer needs width >= 1 and alpha >= 0" );
    abs_width = width;
    abs_alpha = alpha;
}

// EMF::advance(Current&) (emf.cpp:513-580): dt/dx is evaluated in double and
// narrowed, dt is narrowed to float, as in the reference.
void EMF::advance( Current & current ) {
    const float dtdx = dt / dx.x, dtdy = dt / dx.y;
    zpic_geom g = E->geom();
    const int2 p = E->get_periodic();
    check( zpic_yee( E->fbuf(), B->fbuf(), current.J->fbuf(), &g, dtdx, dtdy, (float) dt, p.x, p.y, stream() ) );
    process_bc();
    iter += 1;
    if ( moving_window.active() ) move_window();
}

// EMF::advance() (emf.cpp:135-193): no current
void EMF::advance() {
    const float dtdx = dt / dx.x, dtdy = dt / dx.y;
    zpic_geom g = E->geom();
    const int2 p = E->get_periodic();
    check( zpic_yee( E->fbuf(), B->fbuf(), nullptr, &g, dtdx, dtdy, (float) dt, p.x, p.y, stream() ) );
    process_bc();
    iter += 1;
    if ( moving_window.active() ) move_window();
}

int EMF::set_moving_window() {
    if ( iter == 0 ) {
        moving_window.init( dx.x );
        bc.x.lower = bc.x.upper = emf::bc::none;
        const int2 p = make_int2( 0, E->get_periodic().y );
        E->set_periodic( p );
        B->set_periodic( p );
        return 0;
    }
    std::cerr << "(*error*) set_moving_window() called with iter != 0\n";
    return -1;
}

void EMF::move_window() {
    if ( moving_window.needs_move( iter * dt ) ) {
        zpic_geom g = E->geom();
        const int2 p = E->get_periodic();
        check( zpic_x_shift_left2( E->fbuf(), B->fbuf(), &g, 1, p.x, stream() ) );
        moving_window.advance();
    }
}

void EMF::get_energy( double3 & ene_E, double3 & ene_B ) {
    zpic_geom g = E->geom();
    const size_t nt = (size_t) g.ntiles_x * g.ntiles_y;
    check( zpic_field_energy( E->fbuf(), B->fbuf(), &g, d_energy, d_energy + 6 * nt, stream() ) );
    double h[6];
    to_host( h, d_energy + 6 * nt, 6 );
    sync();
    ene_E = make_double3( h[0], h[1], h[2] );
    ene_B = make_double3( h[3], h[4], h[5] );
}

// ------------------------------------------------------------------ Species
Species::Species( std::string const name_, float const m_q_, uint2 const ppc_ )
    : ppc( ppc_ ), particles( nullptr ), tmp( nullptr ), sort( nullptr ), d_tile_energy( nullptr ),
      d_energy_dev( nullptr ), d_nmove( 0 ), np_bound( 0 ), d_np_inj( nullptr ), deterministic( false ), seed_offset( 0 ), name( name_ ),
      m_q( m_q_ ),
      push_type( species::boris ) {
    if ( m_q == 0 ) throw std::runtime_error( "(*error*) Invalid m_q value, must be not 0, aborting..." );
    if ( ppc.x < 1 || ppc.y < 1 )
        throw std::runtime_error( "(*error*) Invalid ppc value, must be >= 1 in all directions" );
    density = new Density::Uniform( 1.0 );
    udist = new UDistribution::None();
    bc = species::bc_type( species::bc::periodic );
    id = 0; q = 0; dt = 0; iter = 0;
    dx = make_float2( 0, 0 );
    box = make_float2( 0, 0 );
}

Species::~Species() {
    device_free( d_np_inj );
    device_free( d_energy_dev );
    device_free( d_tile_energy );
    delete sort;
    delete tmp;
    delete particles;
    delete density;
    delete udist;
}

// Species::set_bc (species.h:210-237)
void Species::set_bc( species::bc_type new_bc ) {
    if ( ( new_bc.x.lower == species::bc::periodic || new_bc.x.upper == species::bc::periodic ) &&
         new_bc.x.lower != new_bc.x.upper )
        throw std::runtime_error( "(*error*) Species boundary type mismatch along x." );
    if ( ( new_bc.y.lower == species::bc::periodic || new_bc.y.upper == species::bc::periodic ) &&
         new_bc.y.lower != new_bc.y.upper )
        throw std::runtime_error( "(*error*) Species boundary type mismatch along y." );
    bc = new_bc;
    if ( particles ) {
        particles->periodic.x = ( bc.x.lower == species::bc::periodic );
        particles->periodic.y = ( bc.y.lower == species::bc::periodic );
    }
}

int Species::reflect_mask() const {
    return ( bc.x.lower == species::bc::reflecting ? 1 : 0 ) | ( bc.x.upper == species::bc::reflecting ? 2 : 0 ) |
           ( bc.y.lower == species::bc::reflecting ? 4 : 0 ) | ( bc.y.upper == species::bc::reflecting ? 8 : 0 );
}

uint64_t Species::get_nmove() const {
    if ( !sort ) return 0;
    unsigned long long n = 0;
    to_host( &n, sort->buf.nmove, 1 );
    sync();
    return n;
}

// Species::initialize (species.cpp:1730-1788)
void Species::initialize( float2 const box_, uint2 const ntiles, uint2 const nx, double const dt_, int const id_ ) {
    if ( push_type != species::boris )
        throw std::runtime_error( "(*error*) only the Boris pusher is available in this build" );
    box = box_;
    dt = dt_;
    id = id_;
    q = std::copysign( density->n0, m_q ) / ( ppc.x * ppc.y );
    const float2 dims = make_float2( nx.x * ntiles.x, nx.y * ntiles.y );
    dx.x = box.x / dims.x;
    dx.y = box.y / dims.y;
    const uint32_t max_part = 1.2 * dims.x * dims.y * ppc.x * ppc.y;

    particles = new Particles( ntiles, nx, max_part );
    particles->periodic = make_int2( bc.x.lower == species::bc::periodic, bc.y.lower == species::bc::periodic );
    tmp = new Particles( ntiles, nx, max_part );
    sort = new ParticleSort( ntiles, max_part );
    const int nt = ntiles.x * ntiles.y;
    d_tile_energy = device_alloc<double>( nt );
    d_energy_dev = device_alloc<double>( 1 );
    d_np_inj = device_alloc<int>( nt );
    hip_check( hipMemsetAsync( d_tile_energy, 0, nt * sizeof( double ), stream() ), "memset" );
    hip_check( hipMemsetAsync( d_energy_dev, 0, sizeof( double ), stream() ), "memset" );
    d_nmove = 0;
    iter = 0;

    // Species::inject over the whole box (species.cpp:1808-1813), window origin
    // at its current motion
    const int range[4] = { 0, (int) ( nx.x * ntiles.x ) - 1, 0, (int) ( nx.y * ntiles.y ) - 1 };
    const zpic_inject_profile prof = density->desc( dx, make_float2( moving_window.motion(), 0 ) );
    std::vector<int> np( nt ), off( nt );
    Density::np_inject_host( prof, ntiles, nx, ppc, range, np.data() );
    size_t total = 0;
    for ( int t = 0; t < nt; t++ ) { off[t] = (int) total; total += np[t]; }
    if ( total > max_part ) throw std::runtime_error( "(*error*) particle buffer too small for the injection" );
    std::vector<int2> hix( total );
    std::vector<float2> hx( total );
    std::vector<float3> hu( total );
    Density::inject_host( prof, ntiles, nx, ppc, range, off.data(), hix.data(), hx.data() );
    udist->set( nt, np.data(), off.data(), hix.data(), nx, hu.data(), id + seed_offset );

    to_device( particles->np, np.data(), nt );
    to_device( particles->offset, off.data(), nt );
    to_device( particles->ix, hix.data(), total );
    to_device( particles->x, hx.data(), total );
    to_device( particles->u, hu.data(), total );
    sync();
    np_bound = total;
}

// Species::push (species.cpp:2302-2335)
void Species::push( vec3grid<float3> * const E, vec3grid<float3> * const B ) {
    const float alpha = 0.5 * dt / m_q;
    zpic_geom g = E->geom();
    zpic_particles p = particles->view();
    check( zpic_push( &p, &g, E->fbuf(), B->fbuf(), alpha, d_tile_energy, deterministic, stream() ) );
    check( zpic_sum_tiles( d_tile_energy, g.ntiles_x * g.ntiles_y, d_energy_dev, deterministic, stream() ) );
}

// Species::move(J) (species.cpp:2138-2164)
void Species::move( vec3grid<float3> * const J ) {
    const float2 dt_dx = make_float2( dt / dx.x, dt / dx.y );
    const float2 qnx = make_float2( q * dx.x / dt, q * dx.y / dt );
    zpic_geom g = J->geom();
    zpic_particles p = particles->view();
    set_staging( sort->buf );  // the following tile_sort places what this pass stages
    check( zpic_move_deposit( &p, &g, J->fbuf(), dt_dx.x, dt_dx.y, q, qnx.x, qnx.y, &sort->buf, deterministic,
                              stream() ) );
}

void Species::tile_sort() { tile_sort( nullptr, 0 ); }

void Species::tile_sort( const int * extra, uint64_t max_extra ) {
    zpic_particles p = particles->view(), t = tmp->view();
    const uint2 nt = particles->ntiles, nx = particles->nx;
    const int2 per = particles->periodic;
    check( zpic_tile_sort_count( &p, &t, &sort->buf, nt.x, nt.y, nx.x, nx.y, per.x, per.y, extra, stream() ) );
    int mode = 0;
    int total = 0;
    if ( extra ) {
        // the reference compares the total (reserved slots included) with
        // the buffer size; it is read back only when the bound could exceed it
        np_bound += max_extra;
        if ( np_bound > particles->max_part ) {
            to_host( &total, sort->buf.total, 1 );
            sync();
            np_bound = (uint32_t) total;
            if ( (uint32_t) total > particles->max_part ) mode = 1;
        }
    }
    if ( mode == 0 ) {
        check( zpic_tile_sort_copy( &p, &t, &sort->buf, nt.x, nt.y, nx.x, nx.y, per.x, per.y, 0, stream() ) );
        return;
    }
    // copy_sorted path (particles.cpp:868-881): grow the staging buffer, copy
    // every particle into it in order, swap.  The staging buffer and the
    // leaver list are then grown to the new size, so that later sorts (whose
    // staging writes reach the new total) stay inside them.
    tmp->grow_buffer( (uint32_t) total );
    t = tmp->view();
    check( zpic_tile_sort_copy( &p, &t, &sort->buf, nt.x, nt.y, nx.x, nx.y, per.x, per.y, 1, stream() ) );
    swap_buffers( *particles, *tmp );
    tmp->grow_buffer( particles->max_part );
    sort->grow( particles->max_part );
}

void Species::np_inject( const int range[4], int * d_np ) {
    const zpic_inject_profile prof = density->desc( dx, make_float2( moving_window.motion(), 0 ) );
    const uint2 nt = particles->ntiles, nx = particles->nx;
    check( zpic_np_inject( &prof, nt.x, nt.y, nx.x, nx.y, ppc.x, ppc.y, range, d_np, stream() ) );
}

void Species::inject( const int range[4] ) {
    const zpic_inject_profile prof = density->desc( dx, make_float2( moving_window.motion(), 0 ) );
    const uint2 nt = particles->ntiles, nx = particles->nx;
    zpic_particles p = particles->view();
    check( zpic_inject( &prof, &p, nt.x, nt.y, nx.x, nx.y, ppc.x, ppc.y, range, stream() ) );
}

void Species::reinject() {
    const uint2 nt = particles->ntiles, gd = particles->dims;
    const int n = nt.x * nt.y;
    const int range[4] = { 0, (int) gd.x - 1, 0, (int) gd.y - 1 };
    np_inject( range, d_np_inj );
    std::vector<int> np( n ), off( n );
    to_host( np.data(), d_np_inj, n );
    sync();
    size_t total = 0;
    for ( int t = 0; t < n; t++ ) { off[t] = (int) total; total += np[t]; }
    if ( total > particles->max_part ) throw std::runtime_error( "(*error*) particle buffer too small for the injection" );
    std::vector<int> zero( n, 0 );
    to_device( particles->offset, off.data(), n );
    to_device( particles->np, zero.data(), n );
    inject( range );
    sync();
    np_bound = total;
}

int Species::set_moving_window() {
    if ( iter == 0 ) {
        moving_window.init( dx.x );
        bc.x.lower = bc.x.upper = species::bc::open;
        particles->periodic.x = false;
        return 0;
    }
    std::cerr << "(*error*) set_moving_window() called with iter != 0\n";
    return -1;
}

// Species::advance(emf, current) (species.cpp:2051-2068): the push and the
// move run as one pass over the particles.
void Species::advance( EMF const & emf, Current & current ) { step( emf, current, nullptr, false ); }

void Species::advance_timed( EMF const & emf, Current & current, hipEvent_t after_push, bool window, bool set_j ) {
    step( emf, current, after_push, window, set_j );
}

void Species::advance_mov_window( EMF const & emf, Current & current ) { step( emf, current, nullptr, true ); }

// Fast mode with the order-free sort: the particle pass compacts stayers and
// stages leaver records in the sort buffer (`tmp`), and the sort places them
// from there.  Not on moving-window steps: their sort may have to take the
// copy_sorted path, which reads the particles in place, and where it need not
// (the injection cannot outgrow the buffer) the reference order measured
// faster on the LWFA set-up (0.411 against 0.432 ms per step), whose tiles
// hold few particles and all of whose particles move one cell on such steps.
void Species::set_staging( zpic_sort & s, bool window ) const {
    const bool staged = !deterministic && !window && zpic_get_sort_order() == 0;
    s.stage_ix = staged ? reinterpret_cast<int *>( tmp->ix ) : nullptr;
    s.stage_x = staged ? reinterpret_cast<float *>( tmp->x ) : nullptr;
    s.stage_u = staged ? reinterpret_cast<float *>( tmp->u ) : nullptr;
}

void Species::particle_pass( EMF const & emf, Current & current, bool shift, bool set_j ) {
    const float alpha = 0.5 * dt / m_q;
    const float2 dt_dx = make_float2( dt / dx.x, dt / dx.y );
    const float2 qnx = make_float2( q * dx.x / dt, q * dx.y / dt );
    zpic_geom g = emf.E->geom();
    zpic_particles p = particles->view();
    set_staging( sort->buf, shift );
    if ( set_j && !deterministic )
        check( zpic_push_move_deposit_overwrite( &p, &g, emf.E->fbuf(), emf.B->fbuf(), current.J->fbuf(), alpha,
                                                 dt_dx.x, dt_dx.y, q, qnx.x, qnx.y, d_tile_energy, &sort->buf,
                                                 reflect_mask(), shift ? -1 : 0, 0, stream() ) );
    else
        check( zpic_push_move_deposit_shift( &p, &g, emf.E->fbuf(), emf.B->fbuf(), current.J->fbuf(), alpha,
                                             dt_dx.x, dt_dx.y, q, qnx.x, qnx.y, d_tile_energy, &sort->buf,
                                             reflect_mask(), shift ? -1 : 0, 0, deterministic, stream() ) );
}

void Species::finish_step( bool shift ) {
    if ( shift ) {
        // inject into the rightmost column of cells (species.cpp:2094-2112)
        const uint2 gd = particles->dims;
        const int range[4] = { (int) gd.x - 1, (int) gd.x - 1, 0, (int) gd.y - 1 };
        np_inject( range, d_np_inj );
        tile_sort( d_np_inj, (uint64_t) gd.y * ppc.x * ppc.y );
        inject( range );
        moving_window.advance();
    } else {
        tile_sort();
    }
    iter++;
    const uint2 nt = particles->ntiles;
    check( zpic_sum_tiles( d_tile_energy, (int) ( nt.x * nt.y ), d_energy_dev, deterministic, stream() ) );
}

// species.cpp:2088: the window time is (iter + 1) * dt in float
bool Species::shifts( bool window ) const { return window && moving_window.needs_move( ( iter + 1 ) * dt ); }

void Species::step( EMF const & emf, Current & current, hipEvent_t after_push, bool window, bool set_j ) {
    const bool shift = shifts( window );
    particle_pass( emf, current, shift, set_j );
    if ( after_push ) hip_check( hipEventRecord( after_push, stream() ), "hipEventRecord" );
    finish_step( shift );
}

double Species::get_energy() const {
    double e = 0;
    to_host( &e, d_energy_dev, 1 );
    sync();
    return e * q * m_q * dx.x * dx.y;
}

void Species::deposit_charge( grid<float> & charge ) const {
    zpic_geom g = charge.geom();
    zpic_particles p = particles->view();
    check( zpic_deposit_charge( &p, &g, charge.fbuf(), q, deterministic, stream() ) );
}

void Species::get_particles( int * np, int * offset, int * ix, float * x, float * u ) const {
    const int nt = particles->ntiles.x * particles->ntiles.y;
    std::vector<int> hnp = particles->get_np(), hoff = particles->get_offset();
    size_t n = 0;
    for ( int t = 0; t < nt; t++ ) {
        if ( np ) np[t] = hnp[t];
        if ( offset ) offset[t] = (int) n;
        if ( ix ) to_host( reinterpret_cast<int2 *>( ix ) + n, particles->ix + hoff[t], hnp[t] );
        if ( x ) to_host( reinterpret_cast<float2 *>( x ) + n, particles->x + hoff[t], hnp[t] );
        if ( u ) to_host( reinterpret_cast<float3 *>( u ) + n, particles->u + hoff[t], hnp[t] );
        n += hnp[t];
    }
    sync();
}

void Species::set_particles( const int * np, const int * ix, const float * x, const float * u ) {
    const int nt = particles->ntiles.x * particles->ntiles.y;
    std::vector<int> off( nt );
    size_t n = 0;
    for ( int t = 0; t < nt; t++ ) {
        if ( np[t] < 0 ) throw std::invalid_argument( "(*error*) negative particle count for a tile" );
        off[t] = (int) n;
        n += np[t];
    }
    if ( n > particles->max_part ) throw std::runtime_error( "(*error*) too many particles for the buffer" );
    // The particle pass stages a tile's fields and current in LDS around its
    // own cells and the sort moves particles at most one tile per step: every
    // cell index must lie inside its tile.
    const int nxt = (int) particles->nx.x, nyt = (int) particles->nx.y;
    for ( size_t k = 0; k < n; k++ ) {
        const int cx = ix[2 * k], cy = ix[2 * k + 1];
        if ( cx < 0 || cx >= nxt || cy < 0 || cy >= nyt )
            throw std::invalid_argument( "(*error*) particle cell index outside its tile" );
    }
    to_device( particles->np, np, nt );
    to_device( particles->offset, off.data(), nt );
    to_device( particles->ix, reinterpret_cast<const int2 *>( ix ), n );
    to_device( particles->x, reinterpret_cast<const float2 *>( x ), n );
    to_device( particles->u, reinterpret_cast<const float3 *>( u ), n );
    sync();
    np_bound = n;
}

void Species::subsample_from( const Species & src, int stride, int phase ) {
    if ( src.particles->ntiles.x != particles->ntiles.x || src.particles->ntiles.y != particles->ntiles.y ||
         src.particles->nx.x != particles->nx.x || src.particles->nx.y != particles->nx.y )
        throw std::runtime_error( "(*error*) subsample needs species with the same tiling" );
    zpic_particles s = src.particles->view(), d = particles->view();
    zpic_geom gs{}, gd{};
    check( zpic_geom_init( &gs, (int) src.particles->ntiles.x, (int) src.particles->ntiles.y,
                           (int) src.particles->nx.x, (int) src.particles->nx.y, 0, 0 ) );
    check( zpic_geom_init( &gd, (int) particles->ntiles.x, (int) particles->ntiles.y, (int) particles->nx.x,
                           (int) particles->nx.y, 0, 0 ) );
    check( zpic_subsample( &s, &gs, &d, &gd, stride, phase, stream() ) );
    np_bound = particles->np_total();
}

// Particles::gather (particles.cpp:88-150) with Species::gather's scales
void Species::gather( part::quant quant, float * out ) const {
    const int nt = particles->ntiles.x * particles->ntiles.y;
    const uint64_t n = particles->np_total();
    std::vector<int> np( nt ), off( nt ), ix( 2 * n );
    std::vector<float> x( 2 * n ), u( 3 * n );
    get_particles( np.data(), off.data(), ix.data(), x.data(), u.data() );
    const uint2 tnx = particles->nx;
    for ( int t = 0; t < nt; t++ ) {
        const unsigned tx = t % particles->ntiles.x, ty = t / particles->ntiles.x;
        for ( int i = off[t]; i < off[t] + np[t]; i++ ) {
            float v = 0;
            switch ( quant ) {
            case part::x: v = std::fma( dx.x, (float) ( tx * tnx.x + ix[2 * i] ) + ( 0.5f + x[2 * i] ), 0.0f ); break;
            case part::y: v = std::fma( dx.y, (float) ( ty * tnx.y + ix[2 * i + 1] ) + ( 0.5f + x[2 * i + 1] ), 0.0f ); break;
            case part::ux: v = u[3 * i]; break;
            case part::uy: v = u[3 * i + 1]; break;
            case part::uz: v = u[3 * i + 2]; break;
            }
            out[i] = v;
        }
    }
}

// --------------------------------------------------------------- Simulation
Simulation::Simulation( uint2 const ntiles_, uint2 const nx_, float2 const box_, double dt_ )
    : iter( 0 ), timing( false ), t_ring( 0 ), t_steps( 0 ), ntiles( ntiles_ ), nx( nx_ ), box( box_ ), dt( dt_ ),
      emf( ntiles_, nx_, box_, dt_ ), current( ntiles_, nx_, box_, dt_ ) {
    for ( double & v : t_acc ) v = 0;
}

Simulation::~Simulation() {
    for ( auto & v : ev )
        for ( auto e : v ) (void) hipEventDestroy( e );
    for ( hipEvent_t e : dep )
        if ( e ) (void) hipEventDestroy( e );
}

void Simulation::set_timing( bool on ) {
    timing = on;
    for ( double & v : t_acc ) v = 0;
    t_steps = 0;
    for ( bool & p : ev_pending ) p = false;
}

// Simulation::advance (simulation.h:90-107) and advance_mov_window
// (simulation.h:113-127)
void Simulation::advance() { step( false ); }
void Simulation::advance_mov_window() { step( true ); }

void Simulation::set_moving_window() {
    emf.set_moving_window();
    current.set_moving_window();
    for ( auto * s : species ) s->set_moving_window();
}

// One step.  With the moving window, the last species' second half (sort,
// injection when the window moves, energy sum) runs on the side stream
// beside the current and field updates: it reads and writes only that
// species' particles, sort scratch and energy, and those updates only J, E
// and B.  (Measured: 5.7 % off the 4096x512 LWFA step, whose field kernels
// are small and latency-bound; 0.5 % slower on the periodic 2048^2 Weibel
// step, whose field kernels then compete with the sort for HBM, so periodic
// steps stay serial.)  The main stream
// waits for the side stream at the end of the step, so that the step's
// results are complete wherever the main stream is synchronised (host code
// of the sort that reads a count back does so on the side stream).  (A
// species' sort beside the next species' particle pass gained nothing: that
// pass is bound by HBM traffic and slowed by as much as the sort hid.)
// Stage timing records events in one of TSLOTS slots and reads a slot's
// elapsed times only when the slot comes round again (or in get_timing()),
// so that timing adds no host wait per step.
void Simulation::step( bool window ) {
    const size_t nsp = species.size();
    const bool overlap = window && nsp > 0;
    const size_t nev = 2 * nsp + 4;
    const int slot = (int) ( t_ring % TSLOTS );
    if ( timing ) {
        if ( ev_pending[slot] ) accumulate( slot );
        while ( ev[slot].size() < nev ) {
            hipEvent_t e;
            hip_check( hipEventCreate( &e ), "hipEventCreate" );
            ev[slot].push_back( e );
        }
    }
    if ( overlap && !dep[0] )
        for ( hipEvent_t & e : dep ) hip_check( hipEventCreateWithFlags( &e, hipEventDisableTiming ), "hipEventCreate" );
    std::vector<hipEvent_t> & E = ev[slot];
    auto rec = [&]( size_t i, hipStream_t st ) {
        if ( timing ) hip_check( hipEventRecord( E[i], st ), "hipEventRecord" );
    };
    // fast mode: the first species' particle pass writes J instead of a
    // fill zeroing it first
    const bool set_j = nsp > 0 && !species[0]->is_deterministic();
    if ( !set_j ) current.zero();
    rec( 0, stream() );
    for ( size_t k = 0; k + ( overlap ? 1 : 0 ) < nsp; k++ ) {
        species[k]->advance_timed( emf, current, timing ? E[1 + 2 * k] : nullptr, window, set_j && k == 0 );
        rec( 2 + 2 * k, stream() );
    }
    if ( overlap ) {
        Species & sp = *species[nsp - 1];
        const bool shift = sp.shifts( window );
        sp.particle_pass( emf, current, shift, set_j && nsp == 1 );
        rec( 2 * nsp - 1, stream() );
        hipStream_t side = side_stream();
        hip_check( hipEventRecord( dep[0], stream() ), "hipEventRecord" );
        hip_check( hipStreamWaitEvent( side, dep[0], 0 ), "hipStreamWaitEvent" );
        {
            StreamScope scope( side );
            sp.finish_step( shift );
            rec( 2 * nsp, side );
            hip_check( hipEventRecord( dep[1], side ), "hipEventRecord" );
        }
    }
    current.advance();
    rec( 2 * nsp + 1, stream() );
    emf.advance( current );
    rec( 2 * nsp + 2, stream() );
    if ( overlap ) hip_check( hipStreamWaitEvent( stream(), dep[1], 0 ), "hipStreamWaitEvent" );
    rec( 2 * nsp + 3, stream() );
    if ( timing ) {
        ev_pending[slot] = true;
        ev_overlap[slot] = overlap;
        ev_nsp[slot] = nsp;
        t_ring++;
    }
    iter++;
}

// Stage times of the step recorded in `slot` (waits for its last event):
// particle passes, sorts, current update, field update, whole step.
void Simulation::accumulate( int slot ) {
    const size_t nsp = ev_nsp[slot];
    const std::vector<hipEvent_t> & E = ev[slot];
    hip_check( hipEventSynchronize( E[2 * nsp + 3] ), "hipEventSynchronize" );
    auto el = [&]( size_t a, size_t b ) {
        float ms = 0;
        hip_check( hipEventElapsedTime( &ms, E[a], E[b] ), "elapsed" );
        return (double) ms;
    };
    double pm = 0, so = 0;
    for ( size_t k = 0; k < nsp; k++ ) {
        const size_t start = k == 0 ? 0 : 2 * k;  // after the previous species' sort
        pm += el( start, 1 + 2 * k );
        so += el( 1 + 2 * k, 2 + 2 * k );
    }
    t_acc[0] += pm;
    t_acc[1] += so;
    // the current update starts after the last particle pass when the last
    // sort runs beside it
    t_acc[2] += el( ev_overlap[slot] ? 2 * nsp - 1 : 2 * nsp, 2 * nsp + 1 );
    t_acc[3] += el( 2 * nsp + 1, 2 * nsp + 2 );
    t_acc[4] += el( 0, 2 * nsp + 3 );
    t_steps++;
    ev_pending[slot] = false;
}

void Simulation::get_timing( double out[5] ) {
    for ( int s = 0; s < TSLOTS; s++ )
        if ( ev_pending[s] ) accumulate( s );
    for ( int i = 0; i < 5; i++ ) out[i] = t_steps ? t_acc[i] / t_steps : 0.0;
}

void Simulation::energy_info() {
    std::cout << "(*info*) Energy at n = " << iter << ", t = " << iter * double( dt ) << '\n';
    double part_ene = 0;
    for ( auto * s : species ) {
        const double kin = s->get_energy();
        std::cout << "(*info*) " << s->name << " = " << kin << '\n';
        part_ene += kin;
    }
    if ( species.size() > 1 ) std::cout << "(*info*) Total particle energy = " << part_ene << '\n';
    double3 eE, eB;
    emf.get_energy( eE, eB );
    std::cout << "(*info*) Electric field = " << eE.x + eE.y + eE.z << '\n';
    std::cout << "(*info*) Magnetic field = " << eB.x + eB.y + eB.z << '\n';
    std::cout << "(*info*) total = " << part_ene + eE.x + eE.y + eE.z + eB.x + eB.y + eB.z << '\n';
}

// ------------------------------------------------------------------- Batch
zpic_batch_species Species::batch_item( EMF const & emf, Current & current ) const {
    zpic_batch_species it{};
    it.p = particles->view();
    it.tmp = tmp->view();
    it.s = sort->buf;
    set_staging( it.s );
    it.fld = emf.E->geom();
    it.E = emf.E->fbuf();
    it.B = emf.B->fbuf();
    it.J = current.J->fbuf();
    it.tile_energy = d_tile_energy;
    it.energy_out = d_energy_dev;
    it.alpha = 0.5 * dt / m_q;
    it.dt_dx = dt / dx.x;
    it.dt_dy = dt / dx.y;
    it.q = q;
    it.qnx = q * dx.x / dt;
    it.qny = q * dx.y / dt;
    it.periodic_x = particles->periodic.x;
    it.periodic_y = particles->periodic.y;
    return it;
}

void Simulation::note_step() {
    for ( auto * s : species ) s->note_step();
    current.note_step();
    emf.note_step();
    iter++;
}

zpic_batch_fields_item Simulation::batch_fields() const {
    zpic_batch_fields_item it;
    it.E = emf.E->fbuf();
    it.B = emf.B->fbuf();
    it.J = current.J->fbuf();
    it.g = emf.E->geom();
    const float2 d = emf.get_dx();
    it.dt_dx = dt / d.x;
    it.dt_dy = dt / d.y;
    it.dt = (float) dt;
    it.filter_kind = current.filter->kind();
    it.filter_dir = 0;
    it.filter_order = 1;
    if ( auto * b = dynamic_cast<const Filter::Binomial *>( current.filter ) ) {
        it.filter_dir = b->get_dir() == coord::x ? 0 : 1;
        it.filter_order = b->get_order();
    }
    const int2 p = emf.E->get_periodic();
    it.periodic_x = p.x;
    it.periodic_y = p.y;
    return it;
}

Batch::Batch( int threads_per_tile )
    : fields( nullptr ), persistent_step( nullptr ), threads( threads_per_tile ), use_graph( false ),
      persistent( false ), graph( nullptr ), exec( nullptr ) {}

Batch::~Batch() { release(); }

void Batch::release() {
    if ( exec ) (void) hipGraphExecDestroy( exec );
    if ( graph ) (void) hipGraphDestroy( graph );
    exec = nullptr;
    graph = nullptr;
    for ( auto * p : slots ) zpic_batch_particles_free( p );
    slots.clear();
    zpic_batch_fields_free( fields );
    fields = nullptr;
    zpic_batch_step_free( persistent_step );
    persistent_step = nullptr;
}

void Batch::add( Simulation & sim, int threads_per_tile ) {
    if ( threads_per_tile != 0 && ( threads_per_tile < 64 || threads_per_tile > 512 || threads_per_tile % 64 ) )
        throw std::invalid_argument( "(*error*) threads per tile must be 0 or a multiple of 64 in [64, 512]" );
    if ( !sims.empty() && sim.species.size() != sims[0]->species.size() )
        throw std::runtime_error( "(*error*) all instances of a batch must hold the same number of species" );
    auto periodic = []( int lo, int hi ) { return lo == 1 && hi == 1; };
    const auto eb = sim.emf.get_bc();
    const auto cb = sim.current.get_bc();
    bool ok = periodic( eb.x.lower, eb.x.upper ) && periodic( eb.y.lower, eb.y.upper ) &&
              periodic( cb.x.lower, cb.x.upper ) && periodic( cb.y.lower, cb.y.upper );
    for ( auto * s : sim.species ) {
        const auto sb = s->get_bc();
        ok = ok && periodic( sb.x.lower, sb.x.upper ) && periodic( sb.y.lower, sb.y.upper );
    }
    if ( !ok ) throw std::runtime_error( "(*error*) batches run periodic boundaries only" );
    release();
    sims.push_back( &sim );
    sim_threads.push_back( threads_per_tile );
}

void Batch::set_graph( bool on ) {
    if ( exec ) (void) hipGraphExecDestroy( exec );
    if ( graph ) (void) hipGraphDestroy( graph );
    exec = nullptr;
    graph = nullptr;
    use_graph = on;
}

void Batch::set_persistent( bool on ) {
    release();
    persistent = on;
}

void Batch::check_status() {
    if ( persistent_step ) check( zpic_batch_step_status( persistent_step ) );
}

std::vector<const void *> Batch::buffers() const {
    std::vector<const void *> out;
    for ( auto * s : sims ) {
        out.push_back( s->current.J->fbuf() );
        for ( auto * sp : s->species ) sp->buffers( out );
    }
    return out;
}

void Batch::build() {
    if ( sims.empty() ) throw std::runtime_error( "(*error*) empty batch" );
    release();
    const size_t nsp = sims[0]->species.size();
    if ( persistent ) {
        std::vector<zpic_batch_species> items;
        for ( size_t k = 0; k < nsp; k++ )
            for ( size_t i = 0; i < sims.size(); i++ ) {
                Simulation * s = sims[i];
                items.push_back( s->species[k]->batch_item( s->emf, s->current ) );
                items.back().threads = sim_threads[i];
            }
        std::vector<zpic_batch_fields_item> fitems;
        for ( auto * s : sims ) fitems.push_back( s->batch_fields() );
        check( zpic_batch_step_create( &persistent_step, items.data(), (int) nsp, fitems.data(), (int) fitems.size(),
                                       threads ) );
        bound = buffers();
        return;
    }
    for ( size_t k = 0; k < nsp; k++ ) {
        std::vector<zpic_batch_species> items;
        for ( size_t i = 0; i < sims.size(); i++ ) {
            Simulation * s = sims[i];
            items.push_back( s->species[k]->batch_item( s->emf, s->current ) );
            items.back().threads = sim_threads[i];
        }
        zpic_batch_particles * plan = nullptr;
        check( zpic_batch_particles_create( &plan, items.data(), (int) items.size(), threads ) );
        slots.push_back( plan );
    }
    std::vector<zpic_batch_fields_item> fitems;
    for ( auto * s : sims ) fitems.push_back( s->batch_fields() );
    check( zpic_batch_fields_create( &fields, fitems.data(), (int) fitems.size() ) );
    bound = buffers();
}

// Simulation::advance (simulation.h:90-107) for every instance: each stage is
// one launch over all instances.
void Batch::step_kernels() {
    if ( persistent_step ) {
        check( zpic_batch_step_advance( persistent_step, stream() ) );
        return;
    }
    check( zpic_batch_zero_current( fields, stream() ) );
    for ( auto * plan : slots ) check( zpic_batch_particles_advance( plan, stream() ) );
    check( zpic_batch_current_advance( fields, stream() ) );
    check( zpic_batch_yee( fields, stream() ) );
}

void Batch::advance() {
    // rebuilt when an instance's buffers moved since (single-instance steps
    // exchange the current's buffers and may grow or swap particle buffers)
    if ( ( fields || persistent_step ) && buffers() != bound ) build();
    if ( !fields && !persistent_step ) build();
    if ( use_graph ) {
        if ( !exec ) {
            hip_check( hipStreamBeginCapture( stream(), hipStreamCaptureModeThreadLocal ), "hipStreamBeginCapture" );
            try {
                step_kernels();
            } catch ( ... ) {
                hipGraph_t g = nullptr;
                (void) hipStreamEndCapture( stream(), &g );
                if ( g ) (void) hipGraphDestroy( g );
                throw;
            }
            hip_check( hipStreamEndCapture( stream(), &graph ), "hipStreamEndCapture" );
            hip_check( hipGraphInstantiate( &exec, graph, nullptr, nullptr, 0 ), "hipGraphInstantiate" );
        }
        hip_check( hipGraphLaunch( exec, stream() ), "hipGraphLaunch" );
    } else {
        step_kernels();
    }
    for ( auto * s : sims ) s->note_step();
}

// ------------------------------------------------------------- Diagnostics
void EMF::save( emf::field const field, fcomp::cart const fc ) {
    if ( fc < 0 || fc > 2 ) {
        std::cerr << "(*error*) Invalid field component (fc) selected, returning" << std::endl;
        return;
    }
    static const char comp[] = { 'x', 'y', 'z' };
    char vfname[16], vflabel[16];
    vec3grid<float3> * f;
    switch ( field ) {
    case emf::e: f = E; break;
    case emf::b: f = B; break;
    default:
        std::cerr << "(*error*) Invalid field type selected, returning..." << std::endl;
        return;
    }
    const char F = ( field == emf::e ) ? 'E' : 'B';
    std::snprintf( vfname, 16, "%c%c", F, comp[fc] );
    std::snprintf( vflabel, 16, "%c_%c", F, comp[fc] );

    zdf::grid_axis axis[2];
    axis[0] = zdf::grid_axis{ (char *) "x", zdf::zdf_linear, 0.0 + moving_window.motion(), box.x + moving_window.motion(),
                              (char *) "x", (char *) "c/\\omega_n" };
    axis[1] = zdf::grid_axis{ (char *) "y", zdf::zdf_linear, 0.0, box.y, (char *) "y", (char *) "c/\\omega_n" };
    zdf::grid_info info{};
    info.name = vfname;
    info.ndims = 2;
    info.label = vflabel;
    info.units = (char *) "m_e c \\omega_n e^{-1}";
    info.axis = axis;
    zdf::iteration iteration{ nullptr, iter, iter * dt, (char *) "1/\\omega_n" };
    f->save( fc, info, iteration, "EMF" );
}

void Current::save( fcomp::cart const jc ) {
    if ( jc < 0 || jc > 2 ) {
        std::cerr << "(*error*) Invalid current component (jc) selected, returning" << std::endl;
        return;
    }
    static const char comp[] = { 'x', 'y', 'z' };
    char vfname[16], vflabel[16];
    std::snprintf( vfname, 16, "J%c", comp[jc] );
    std::snprintf( vflabel, 16, "J_%c", comp[jc] );
    // The reference shifts the lower y bound (not the upper x bound) with the
    // window (current.cpp:282-296); kept as is.
    zdf::grid_axis axis[2];
    axis[0] = zdf::grid_axis{ (char *) "x", zdf::zdf_linear, 0.0 + moving_window.motion(), box.x, (char *) "x",
                              (char *) "c/\\omega_n" };
    axis[1] = zdf::grid_axis{ (char *) "y", zdf::zdf_linear, 0.0 + moving_window.motion(), box.y, (char *) "y",
                              (char *) "c/\\omega_n" };
    zdf::grid_info info{};
    info.name = vfname;
    info.ndims = 2;
    info.label = vflabel;
    info.units = (char *) "e \\omega_n^2 / c";
    info.axis = axis;
    zdf::iteration iteration{ nullptr, iter, iter * dt, (char *) "1/\\omega_n" };
    J->save( jc, info, iteration, "CURRENT" );
}

void Species::save() const {
    const std::string path = "PARTICLES/" + name;
    static const char * quants[] = { "x", "y", "ux", "uy", "uz" };
    static const char * qlabels[] = { "x", "y", "u_x", "u_y", "u_z" };
    static const char * qunits[] = { "c/\\omega_n", "c/\\omega_n", "c", "c", "c" };
    zdf::iteration iter_info{ nullptr, iter, (double) ( iter * dt ), (char *) "1/\\omega_n" };
    zdf::part_info info{};
    info.name = (char *) name.c_str();
    info.label = (char *) name.c_str();
    info.nquants = 5;
    info.quants = (char **) quants;
    info.qlabels = (char **) qlabels;
    info.qunits = (char **) qunits;
    const uint32_t np = np_total();
    info.np = np;
    zdf::file f;
    if ( zdf::open_part_file( f, info, iter_info, path ) != 1 )
        throw std::runtime_error( "(*error*) Unable to open particle file under " + path );
    std::vector<float> data( np );
    for ( int q = 0; q < 5; q++ ) {
        if ( np > 0 ) gather( static_cast<part::quant>( q ), data.data() );
        zdf::add_quant_part_file( f, quants[q], data.data(), np );
    }
    f.close();
}

void Species::save_charge() const {
    bnd<unsigned int> gc;
    gc.x = { 0, 1 };
    gc.y = { 0, 1 };
    grid<float> charge( particles->ntiles, particles->nx, gc );
    charge.zero();
    deposit_charge( charge );
    charge.add_from_gc();

    zdf::grid_axis axis[2];
    axis[0] = zdf::grid_axis{ (char *) "x", zdf::zdf_linear, 0. + moving_window.motion(), box.x + moving_window.motion(),
                              (char *) "x", (char *) "c/\\omega_n" };
    axis[1] = zdf::grid_axis{ (char *) "y", zdf::zdf_linear, 0., box.y, (char *) "y", (char *) "c/\\omega_n" };
    const std::string grid_name = name + "-charge";
    const std::string grid_label = name + " \\rho";
    zdf::grid_info info{};
    info.name = (char *) grid_name.c_str();
    info.label = (char *) grid_label.c_str();
    info.units = (char *) "n_e";
    info.axis = axis;
    zdf::iteration iter_info{ (char *) "ITERATION", iter, (double) ( iter * dt ), (char *) "1/\\omega_n" };
    charge.save( info, iter_info, "CHARGE/" + name );
}

// The normalisation is evaluated as the reference writes it: fabs(q) is the
// double-precision overload there, so the products run in double and the
// result is rounded to float once (species.cpp:2636-2637, 2880-2882).
void Species::dep_phasespace( float * const d_data, phasespace::quant quant, float2 range, unsigned const size ) const {
    const float norm = std::fabs( (double) q ) * ( dx.x * dx.y ) * size / ( range.y - range.x );
    if ( quant == phasespace::x ) { range.y /= dx.x; range.x /= dx.x; }
    if ( quant == phasespace::y ) { range.y /= dx.y; range.x /= dx.y; }
    zpic_particles p = particles->view();
    check( zpic_phasespace( &p, particles->ntiles.x, particles->ntiles.y, particles->nx.x, particles->nx.y, quant, range.x,
                            range.y, size, -1, 0.f, 0.f, 0, norm, d_data, deterministic, stream() ) );
}

void Species::dep_phasespace( float * const d_data, phasespace::quant quant0, float2 range0, unsigned const size0,
                              phasespace::quant quant1, float2 range1, unsigned const size1 ) const {
    const float norm = std::fabs( (double) q ) * ( dx.x * dx.y ) * ( size0 / ( range0.y - range0.x ) ) *
                       ( size1 / ( range1.y - range1.x ) );
    auto scale = [&]( phasespace::quant qn, float2 & r ) {
        if ( qn == phasespace::x ) { r.y /= dx.x; r.x /= dx.x; }
        if ( qn == phasespace::y ) { r.y /= dx.y; r.x /= dx.y; }
    };
    scale( quant0, range0 );
    scale( quant1, range1 );
    zpic_particles p = particles->view();
    check( zpic_phasespace( &p, particles->ntiles.x, particles->ntiles.y, particles->nx.x, particles->nx.y, quant0,
                            range0.x, range0.y, size0, quant1, range1.x, range1.y, size1, norm, d_data, deterministic,
                            stream() ) );
}

void Species::save_phasespace( phasespace::quant quant, float2 const range, int const size ) const {
    std::string qname, qlabel, qunits;
    phasespace::qinfo( quant, qname, qlabel, qunits );
    zdf::grid_axis axis{ (char *) qname.c_str(), zdf::zdf_linear, range.x, range.y, (char *) qlabel.c_str(),
                         (char *) qunits.c_str() };
    if ( quant == phasespace::x ) {
        axis.min += moving_window.motion();
        axis.max += moving_window.motion();
    }
    const std::string pha_name = name + "-" + qname;
    const std::string pha_label = name + "\\,(" + qlabel + ")";
    zdf::grid_info info{};
    info.name = (char *) pha_name.c_str();
    info.ndims = 1;
    info.label = (char *) pha_label.c_str();
    info.units = (char *) "n_e";
    info.axis = &axis;
    info.count[0] = size;
    zdf::iteration iter_info{ (char *) "ITERATION", iter, (double) ( iter * dt ), (char *) "1/\\omega_n" };

    float * d = zpic::device_alloc<float>( size );
    std::vector<float> h( size );
    dep_phasespace( d, quant, range, size );
    to_host( h.data(), d, size );
    sync();
    zpic::device_free( d );
    zdf::save_grid( h.data(), info, iter_info, "PHASESPACE/" + name );
}

void Species::save_phasespace( phasespace::quant quant0, float2 const range0, int const size0, phasespace::quant quant1,
                               float2 const range1, int const size1 ) const {
    if ( quant0 >= quant1 ) {
        std::cerr << "(*error*) for 2D phasespaces, the 2nd quantity must be indexed higher than the first one\n";
        return;
    }
    std::string qname0, qlabel0, qunits0, qname1, qlabel1, qunits1;
    phasespace::qinfo( quant0, qname0, qlabel0, qunits0 );
    phasespace::qinfo( quant1, qname1, qlabel1, qunits1 );
    zdf::grid_axis axis[2] = {
        zdf::grid_axis{ (char *) qname0.c_str(), zdf::zdf_linear, range0.x, range0.y, (char *) qlabel0.c_str(),
                        (char *) qunits0.c_str() },
        zdf::grid_axis{ (char *) qname1.c_str(), zdf::zdf_linear, range1.x, range1.y, (char *) qlabel1.c_str(),
                        (char *) qunits1.c_str() } };
    if ( quant0 == phasespace::x ) {
        axis[0].min += moving_window.motion();
        axis[0].max += moving_window.motion();
    }
    const std::string pha_name = name + "-" + qname0 + qname1;
    const std::string pha_label = name + " \\,(" + qlabel0 + "\\rm{-}" + qlabel1 + ")";
    zdf::grid_info info{};
    info.name = (char *) pha_name.c_str();
    info.ndims = 2;
    info.count[0] = static_cast<unsigned>( size0 );
    info.count[1] = static_cast<unsigned>( size1 );
    info.label = (char *) pha_label.c_str();
    info.units = (char *) "n_e";
    info.axis = axis;
    zdf::iteration iter_info{ (char *) "ITERATION", iter, (double) ( iter * dt ), (char *) "1/\\omega_n" };

    const size_t n = (size_t) size0 * size1;
    float * d = zpic::device_alloc<float>( n );
    std::vector<float> h( n );
    dep_phasespace( d, quant0, range0, size0, quant1, range1, size1 );
    to_host( h.data(), d, n );
    sync();
    zpic::device_free( d );
    zdf::save_grid( h.data(), info, iter_info, "PHASESPACE/" + name );
}

// --------------------------------------------------------------- Autotuner
namespace {

std::string tune_key( uint2 grid, uint2 ppc ) {
    hipDeviceProp_t prop;
    int dev = 0;
    (void) hipGetDevice( &dev );
    std::string arch = "unknown";
    if ( hipGetDeviceProperties( &prop, dev ) == hipSuccess ) arch = prop.gcnArchName;
    for ( auto & ch : arch )
        if ( ch == ' ' ) ch = '_';
    return arch + ":" + std::to_string( grid.x ) + "x" + std::to_string( grid.y ) + ":" + std::to_string( ppc.x ) + "x" +
           std::to_string( ppc.y );
}

bool tune_lookup( const std::string & path, const std::string & key, zpic_tune_result & out ) {
    std::ifstream in( path );
    std::string k;
    zpic_tune_result r{};
    while ( in >> k >> r.nx >> r.ny >> r.threads >> r.particles_per_ns ) {
        if ( k == key ) {
            out = r;
            out.from_cache = 1;
            return true;
        }
    }
    return false;
}

void tune_store( const std::string & path, const std::string & key, const zpic_tune_result & r ) {
    std::ofstream out( path, std::ios::app );
    out << key << ' ' << r.nx << ' ' << r.ny << ' ' << r.threads << ' ' << r.particles_per_ns << '\n';
}

// particles / ns of `steps` steps of a two-species thermal plasma with the
// given tiling, after two warm-up steps
double tune_trial( uint2 ntiles, uint2 nx, uint2 ppc, int threads, int steps ) {
    check( zpic_set_threads_per_tile( threads ) );
    const float2 box = make_float2( ntiles.x * nx.x * 0.1f, ntiles.y * nx.y * 0.1f );
    Simulation sim( ntiles, nx, box, 0.07 );
    Species e( "electrons", -1.0f, ppc ), p( "positrons", 1.0f, ppc );
    e.set_udist( UDistribution::Thermal( make_float3( 0.1f, 0.1f, 0.1f ), make_float3( 0, 0, 0.6f ) ) );
    p.set_udist( UDistribution::Thermal( make_float3( 0.1f, 0.1f, 0.1f ), make_float3( 0, 0, -0.6f ) ) );
    sim.add_species( e );
    sim.add_species( p );
    for ( int i = 0; i < 2; i++ ) sim.advance();
    hipEvent_t t0, t1;
    hip_check( hipEventCreate( &t0 ), "hipEventCreate" );
    hip_check( hipEventCreate( &t1 ), "hipEventCreate" );
    const uint64_t n0 = sim.get_nmove();
    hip_check( hipEventRecord( t0, stream() ), "hipEventRecord" );
    for ( int i = 0; i < steps; i++ ) sim.advance();
    hip_check( hipEventRecord( t1, stream() ), "hipEventRecord" );
    hip_check( hipEventSynchronize( t1 ), "hipEventSynchronize" );
    float ms = 0;
    hip_check( hipEventElapsedTime( &ms, t0, t1 ), "hipEventElapsedTime" );
    (void) hipEventDestroy( t0 );
    (void) hipEventDestroy( t1 );
    return ( sim.get_nmove() - n0 ) / ( ms * 1e6 );
}

}  // namespace

extern "C" int zpic_autotune( unsigned grid_x, unsigned grid_y, unsigned ppc_x, unsigned ppc_y, int steps,
                              long long max_cells, const char * cache_path, zpic_tune_result * out ) {
    if ( !out || grid_x < 8 || grid_y < 8 || ppc_x < 1 || ppc_y < 1 || steps < 1 )
        return zpic::set_error( ZPIC_EINVAL, "invalid autotune request" );
    if ( max_cells <= 0 ) max_cells = 1024 * 1024;
    const int saved = zpic_get_threads_per_tile();
    try {
        const uint2 grid = make_uint2( grid_x, grid_y ), ppc = make_uint2( ppc_x, ppc_y );
        const std::string key = tune_key( grid, ppc );
        if ( cache_path && tune_lookup( cache_path, key, *out ) ) return ZPIC_OK;
        int dev = 0, cus = 0;
        hip_check( hipGetDevice( &dev ), "hipGetDevice" );
        hip_check( hipDeviceGetAttribute( &cus, hipDeviceAttributeMultiprocessorCount, dev ), "hipDeviceGetAttribute" );
        // A tile shape is judged at full occupancy, as it runs in a batch of
        // many instances or in a large grid: the sample has about four tiles
        // per compute unit whatever the instance's own grid (which only has
        // to be divisible by the shape), within max_cells cells.  (A sample
        // cut from a small grid leaves most CUs idle and favours small tiles,
        // which are the slower ones when the GPU is full.)
        const long long want = 4LL * std::max( cus, 1 );
        zpic_tune_result best{};
        const unsigned sizes[] = { 8, 16, 32 };
        for ( unsigned sy : sizes )
            for ( unsigned sx : sizes ) {
                if ( grid_x % sx || grid_y % sy ) continue;
                uint2 nt = make_uint2( 2, 2 );
                while ( (long long) nt.x * nt.y < want ) {
                    const uint2 next = nt.x <= nt.y ? make_uint2( 2 * nt.x, nt.y ) : make_uint2( nt.x, 2 * nt.y );
                    if ( (long long) next.x * sx * next.y * sy > max_cells ) break;
                    nt = next;
                }
                const int vol = ( ( sx + 3 ) * ( sy + 3 ) + 3 ) & ~3;
                for ( int threads : { 64, 128, 256, 512 } ) {
                    if ( zpic_fitted_threads_per_tile( threads, vol ) != threads ) continue;
                    const double pns = tune_trial( nt, make_uint2( sx, sy ), ppc, threads, steps );
                    if ( pns > best.particles_per_ns ) best = zpic_tune_result{ sx, sy, threads, pns, 0 };
                }
            }
        (void) zpic_set_threads_per_tile( saved );
        if ( best.threads == 0 ) return zpic::set_error( ZPIC_ERANGE, "no tile shape fits the grid" );
        *out = best;
        if ( cache_path ) tune_store( cache_path, key, best );
        return ZPIC_OK;
    } catch ( std::exception & e ) {
        (void) zpic_set_threads_per_tile( saved );
        return zpic::set_error( ZPIC_EHIP, e.what() );
    }
}

// -------------------------------------------------------------------- Laser
// Field values are computed on the host with the reference's types: the
// envelope's sine in double (M_PI_2 * csi / rise), the Gaussian beam's Gouy
// phase with the double atan2 that laser.cpp's unqualified call resolves
// to, every other function in float.
float Laser::Pulse::lon_env( const float z ) const {
    if ( z > start ) return 0.0;
    if ( z > start - rise ) {
        const float csi = z - start;
        const float e = std::sin( M_PI_2 * csi / rise );
        return e * e;
    }
    if ( z > start - ( rise + flat ) ) return 1.0;
    if ( z > start - ( rise + flat + fall ) ) {
        const float csi = z - ( start - rise - flat - fall );
        const float e = std::sin( M_PI_2 * csi / fall );
        return e * e;
    }
    return 0.0;
}

int Laser::Pulse::validate() {
    if ( a0 <= 0 ) {
        std::cerr << "(*error*) Invalid laser a0, must be > 0\n";
        return -1;
    }
    if ( omega0 <= 0 ) {
        std::cerr << "(*error*) Invalid laser OMEGA0, must be > 0\n";
        return -1;
    }
    if ( fwhm > 0 ) {
        rise = fwhm;
        fall = fwhm;
        flat = 0.;
    } else {
        if ( rise <= 0 ) {
            std::cerr << "(*error*) Invalid laser RISE, must be > 0\n";
            return -1;
        }
        if ( flat < 0 ) {
            std::cerr << "(*error*) Invalid laser FLAT, must be >= 0\n";
            return -1;
        }
        if ( fall <= 0 ) {
            std::cerr << "(*error*) Invalid laser FALL, must be > 0\n";
            return -1;
        }
    }
    return 0;
}

int Laser::Pulse::add( EMF & emf ) {
    vec3grid<float3> tmp_E( emf.E->get_ntiles(), emf.E->nx, emf.E->gc );
    vec3grid<float3> tmp_B( emf.E->get_ntiles(), emf.B->nx, emf.B->gc );
    const int ierr = launch( tmp_E, tmp_B, emf.box );
    if ( !ierr ) {
        check( zpic_grid_add( emf.E->fbuf(), tmp_E.fbuf(), emf.E->buffer_size() * 3, stream() ) );
        check( zpic_grid_add( emf.B->fbuf(), tmp_B.fbuf(), emf.B->buffer_size() * 3, stream() ) );
        sync();
    }
    return ierr;
}

namespace {

// Guard copy, compensated x filter and upload-side steps shared by the
// pulses (laser.cpp:276-286, 420-432).
void finish_launch( vec3grid<float3> & E, vec3grid<float3> & B, std::vector<float3> & hE, std::vector<float3> & hB,
                    unsigned filter ) {
    E.upload( hE.data() );
    B.upload( hB.data() );
    E.copy_to_gc( B );
    if ( filter > 0 ) {
        Filter::Compensated fcomp( coord::x, filter );
        fcomp.apply( E );
        fcomp.apply( B );
    }
}

// gauss_phase (laser.cpp:303-316)
float gauss_phase( const float omega0, const float W0, const float z, const float r ) {
    const float z0 = omega0 * ( W0 * W0 ) / 2;
    const float rho2 = r * r;
    const float curv = rho2 * z / ( z0 * z0 + z * z );
    const float rWl2 = ( z0 * z0 ) / ( z0 * z0 + z * z );
    const float gouy_shift = (float) std::atan2( (double) z, (double) z0 );
    return std::sqrt( std::sqrt( rWl2 ) ) * std::exp( -rho2 * rWl2 / ( W0 * W0 ) ) *
           std::cos( omega0 * ( z + curv ) - gouy_shift );
}

}  // namespace

// Laser::PlaneWave::launch (laser.cpp:224-289)
int Laser::PlaneWave::launch( vec3grid<float3> & E, vec3grid<float3> & B, float2 box ) {
    if ( validate() < 0 ) return -1;
    if ( ( cos_pol == 0 ) && ( sin_pol == 0 ) ) {
        cos_pol = std::cos( polarization );
        sin_pol = std::sin( polarization );
    }
    const uint2 g_nx = E.get_dims();
    const float2 dx = make_float2( box.x / g_nx.x, box.y / g_nx.y );
    const uint2 ntiles = E.get_ntiles();
    const uint2 nx = E.nx;
    const int ystride = E.ext_nx.x;
    const float k = omega0;
    const float amp = omega0 * a0;
    std::vector<float3> hE( E.buffer_size(), make_float3( 0, 0, 0 ) ), hB( B.buffer_size(), make_float3( 0, 0, 0 ) );
    for ( unsigned ty = 0; ty < ntiles.y; ty++ )
        for ( unsigned tx = 0; tx < ntiles.x; tx++ ) {
            const unsigned tid = ty * ntiles.x + tx;
            float3 * tile_E = &hE[(size_t) tid * E.tile_vol + E.offset];
            float3 * tile_B = &hB[(size_t) tid * B.tile_vol + B.offset];
            const int ix0 = tx * nx.x;
            for ( unsigned iy = 0; iy < nx.y; iy++ )
                for ( unsigned ix = 0; ix < nx.x; ix++ ) {
                    const float z = ( ix0 + ix ) * dx.x;
                    const float z_2 = ( ix0 + ix + 0.5 ) * dx.x;
                    const float lenv = amp * lon_env( z );
                    const float lenv_2 = amp * lon_env( z_2 );
                    tile_E[ix + iy * ystride] = make_float3( 0, +lenv * std::cos( k * z ) * cos_pol,
                                                             +lenv * std::cos( k * z ) * sin_pol );
                    tile_B[ix + iy * ystride] = make_float3( 0, -lenv_2 * std::cos( k * z_2 ) * sin_pol,
                                                             +lenv_2 * std::cos( k * z_2 ) * cos_pol );
                }
        }
    finish_launch( E, B, hE, hB, filter );
    return 0;
}

int Laser::Gaussian::validate() {
    if ( Laser::Pulse::validate() < 0 ) return -1;
    if ( W0 <= 0 ) {
        std::cerr << "(*error*) Invalid laser W0, must be > 0\n";
        return -1;
    }
    return 0;
}

// Laser::Gaussian::launch (laser.cpp:361-437)
int Laser::Gaussian::launch( vec3grid<float3> & E, vec3grid<float3> & B, float2 box ) {
    if ( validate() < 0 ) return -1;
    if ( ( cos_pol == 0 ) && ( sin_pol == 0 ) ) {
        cos_pol = std::cos( polarization );
        sin_pol = std::sin( polarization );
    }
    const uint2 g_nx = E.get_dims();
    const float2 dx = make_float2( box.x / g_nx.x, box.y / g_nx.y );
    const uint2 ntiles = E.get_ntiles();
    const uint2 nx = E.nx;
    const int ystride = E.ext_nx.x;
    const float amp = omega0 * a0;
    std::vector<float3> hE( E.buffer_size(), make_float3( 0, 0, 0 ) ), hB( B.buffer_size(), make_float3( 0, 0, 0 ) );
    for ( unsigned ty = 0; ty < ntiles.y; ty++ )
        for ( unsigned tx = 0; tx < ntiles.x; tx++ ) {
            const unsigned tid = ty * ntiles.x + tx;
            float3 * tile_E = &hE[(size_t) tid * E.tile_vol + E.offset];
            float3 * tile_B = &hB[(size_t) tid * B.tile_vol + B.offset];
            const int ix0 = tx * nx.x;
            const int iy0 = ty * nx.y;
            for ( unsigned iy = 0; iy < nx.y; iy++ )
                for ( unsigned ix = 0; ix < nx.x; ix++ ) {
                    const float z = ( ix0 + ix ) * dx.x;
                    const float z_2 = ( ix0 + ix + 0.5 ) * dx.x;
                    const float r = ( iy0 + iy ) * dx.y - axis;
                    const float r_2 = ( iy0 + iy + 0.5 ) * dx.y - axis;
                    const float lenv = amp * lon_env( z );
                    const float lenv_2 = amp * lon_env( z_2 );
                    tile_E[ix + iy * ystride] =
                        make_float3( 0, +lenv * gauss_phase( omega0, W0, z - focus, r_2 ) * cos_pol,
                                     +lenv * gauss_phase( omega0, W0, z - focus, r ) * sin_pol );
                    tile_B[ix + iy * ystride] =
                        make_float3( 0, -lenv_2 * gauss_phase( omega0, W0, z_2 - focus, r ) * sin_pol,
                                     +lenv_2 * gauss_phase( omega0, W0, z_2 - focus, r_2 ) * cos_pol );
                }
        }
    finish_launch( E, B, hE, hB, filter );
    zpic_geom g = E.geom();
    check( zpic_div_corr_x( E.fbuf(), B.fbuf(), &g, dx.x, dx.y, stream() ) );
    return 0;
}
