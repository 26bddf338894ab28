// oracle_capi.cpp -- TEST INFRASTRUCTURE, NOT PRODUCT CODE.
//
// C entry points over the CPU restatement (zpic_oracle.hpp) so that the
// Python tests and the cpu_baseline leg of bench.py can drive it through
// ctypes.  `prec` selects the real type: 4 = float (bit-exact reference
// arithmetic), 8 = double (FP64 evaluation of the same algorithm).  Floating
// point arrays are exchanged in the handle's precision.

#include "zpic_oracle.hpp"

#include <cstdio>
#include <memory>
#include <string>

using namespace oracle;

namespace {

thread_local std::string g_err;

struct Handle {
    int prec;
    std::unique_ptr<Sim<float>> f;
    std::unique_ptr<Sim<double>> d;
};

template <class F>
int guard( F && fn ) {
    try { return fn(); } catch ( std::exception & e ) { g_err = e.what(); return -1; }
}

template <class R>
int get_particles( Sim<R> & s, int k, int * np, int * off, int * ix, R * x, R * u ) {
    if ( k < 0 || k >= (int) s.sp.size() ) { g_err = "bad species index"; return -1; }
    auto & P = s.sp[k].p;
    size_t n = 0;
    for ( int t = 0; t < P.ntiles; t++ ) {
        if ( np ) np[t] = P.np[t];
        if ( off ) off[t] = (int) n;
        for ( int i = P.off[t]; i < P.off[t] + P.np[t]; i++, n++ ) {
            if ( ix ) { ix[2 * n] = P.ix[i].x; ix[2 * n + 1] = P.ix[i].y; }
            if ( x ) { x[2 * n] = P.x[i].x; x[2 * n + 1] = P.x[i].y; }
            if ( u ) { u[3 * n] = P.u[i].x; u[3 * n + 1] = P.u[i].y; u[3 * n + 2] = P.u[i].z; }
        }
    }
    return 0;
}

template <class R>
int set_particles( Sim<R> & s, int k, const int * np, const int * ix, const R * x, const R * u ) {
    if ( k < 0 || k >= (int) s.sp.size() ) { g_err = "bad species index"; return -1; }
    auto & sp = s.sp[k];
    auto & P = sp.p;
    size_t n = 0;
    for ( int t = 0; t < P.ntiles; t++ ) { P.off[t] = (int) n; P.np[t] = np[t]; n += np[t]; }
    if ( n > sp.max_part ) { g_err = "too many particles"; return -1; }
    for ( size_t i = 0; i < n; i++ ) {
        P.ix[i] = { ix[2 * i], ix[2 * i + 1] };
        P.x[i] = { x[2 * i], x[2 * i + 1] };
        P.u[i] = { u[3 * i], u[3 * i + 1], u[3 * i + 2] };
    }
    return 0;
}

template <class R>
Grid<V3<R>> * field( Sim<R> & s, int which ) {
    switch ( which ) {
    case 0: return &s.E;
    case 1: return &s.B;
    case 2: return &s.J;
    }
    return nullptr;
}

template <class R>
int stage( Sim<R> & s, int what, int k ) {
    if ( what <= 3 && ( k < 0 || k >= (int) s.sp.size() ) ) { g_err = "bad species index"; return -1; }
    switch ( what ) {
    case 0: s.push( s.sp[k] ); break;
    case 1: s.move_deposit( s.sp[k] ); break;
    case 2: s.sp[k].iter++; s.tile_sort( s.sp[k] ); break;
    case 3: s.species_advance( s.sp[k] ); break;
    case 4: std::fill( s.J.d.begin(), s.J.d.end(), V3<R>{ 0, 0, 0 } ); break;
    case 5: s.J.add_from_gc(); s.J.copy_to_gc(); apply_filter<R>( s.J, s.filter ); break;
    case 6: s.yee_step(); break;
    case 7: s.J.add_from_gc(); break;
    case 8: s.J.copy_to_gc(); break;
    case 9: apply_filter<R>( s.J, s.filter ); break;
    default: g_err = "bad stage"; return -1;
    }
    return 0;
}

} // namespace

extern "C" {

const char * orc_last_error() { return g_err.c_str(); }

void * orc_create( int prec, int ntx, int nty, int nx, int ny, float boxx, float boxy, double dt ) {
    try {
        auto h = new Handle();
        h->prec = prec;
        if ( prec == 8 ) h->d.reset( new Sim<double>( ntx, nty, nx, ny, boxx, boxy, dt ) );
        else h->f.reset( new Sim<float>( ntx, nty, nx, ny, boxx, boxy, dt ) );
        return h;
    } catch ( std::exception & e ) {
        g_err = e.what();
        return nullptr;
    }
}

void orc_destroy( void * p ) { delete (Handle *) p; }

int orc_set_filter( void * p, int kind, int dir, int order ) {
    Handle * h = (Handle *) p;
    FilterSpec fs{ kind, dir, order };
    if ( h->f ) h->f->filter = fs; else h->d->filter = fs;
    return 0;
}

int orc_add_species( void * p, float m_q, int ppcx, int ppcy, int udist, const float * uth, const float * ufl,
                     int npmin, float n0 ) {
    Handle * h = (Handle *) p;
    return guard( [&] {
        auto fill = [&]( auto & s ) {
            s.m_q = m_q; s.ppcx = ppcx; s.ppcy = ppcy; s.n0 = std::fabs( n0 );
            s.ud.kind = udist; s.ud.npmin = npmin;
            for ( int i = 0; i < 3; i++ ) { s.ud.uth[i] = uth ? uth[i] : 0; s.ud.ufl[i] = ufl ? ufl[i] : 0; }
        };
        if ( m_q == 0 ) throw std::runtime_error( "Invalid m_q value, must be not 0" );
        if ( ppcx < 1 || ppcy < 1 ) throw std::runtime_error( "Invalid ppc value, must be >= 1 in all directions" );
        if ( h->f ) { Species<float> s; fill( s ); h->f->add_species( s ); }
        else { Species<double> s; fill( s ); h->d->add_species( s ); }
        return 0;
    } );
}

// orc_add_species with a density profile: kind 0 none, 1 uniform, 2 step,
// 3 slab, 4 sphere; prm = {pos or begin, end, centre x, centre y, radius}.
int orc_add_species_profile( void * p, float m_q, int ppcx, int ppcy, int udist, const float * uth, const float * ufl,
                             int npmin, float n0, int kind, int dir, const float * prm ) {
    Handle * h = (Handle *) p;
    return guard( [&] {
        auto fill = [&]( auto & s ) {
            s.m_q = m_q; s.ppcx = ppcx; s.ppcy = ppcy; s.n0 = std::fabs( n0 );
            s.ud.kind = udist; s.ud.npmin = npmin;
            for ( int i = 0; i < 3; i++ ) { s.ud.uth[i] = uth ? uth[i] : 0; s.ud.ufl[i] = ufl ? ufl[i] : 0; }
            s.prof.kind = kind; s.prof.dir = dir;
            s.prof.pos = s.prof.begin = prm[0]; s.prof.end = prm[1];
            s.prof.cx = prm[2]; s.prof.cy = prm[3]; s.prof.radius = prm[4];
        };
        if ( m_q == 0 ) throw std::runtime_error( "Invalid m_q value, must be not 0" );
        if ( ppcx < 1 || ppcy < 1 ) throw std::runtime_error( "Invalid ppc value, must be >= 1 in all directions" );
        if ( kind < 0 || kind > 4 || !prm ) throw std::runtime_error( "invalid density profile" );
        if ( h->f ) { Species<float> s; fill( s ); h->f->add_species( s ); }
        else { Species<double> s; fill( s ); h->d->add_species( s ); }
        return 0;
    } );
}

int orc_advance( void * p, int nsteps ) {
    Handle * h = (Handle *) p;
    return guard( [&] {
        for ( int i = 0; i < nsteps; i++ ) { if ( h->f ) h->f->advance(); else h->d->advance(); }
        return 0;
    } );
}

int orc_stage( void * p, int what, int k ) {
    Handle * h = (Handle *) p;
    return guard( [&] { return h->f ? stage( *h->f, what, k ) : stage( *h->d, what, k ); } );
}

// geometry: ntx, nty, nx, ny, ext_x, ext_y, vol, off
int orc_geom( void * p, int * out ) {
    Handle * h = (Handle *) p;
    const Geom & g = h->f ? h->f->g : h->d->g;
    int v[8] = { g.ntx, g.nty, g.nx, g.ny, g.ext_x, g.ext_y, g.vol, g.off };
    for ( int i = 0; i < 8; i++ ) out[i] = v[i];
    return 0;
}

long long orc_np_total( void * p, int k ) {
    Handle * h = (Handle *) p;
    if ( h->f ) return k < (int) h->f->sp.size() ? (long long) h->f->sp[k].p.total() : -1;
    return k < (int) h->d->sp.size() ? (long long) h->d->sp[k].p.total() : -1;
}

unsigned long long orc_nmove( void * p, int k ) {
    Handle * h = (Handle *) p;
    return h->f ? h->f->sp[k].nmove : h->d->sp[k].nmove;
}

int orc_get_particles( void * p, int k, int * np, int * off, int * ix, void * x, void * u ) {
    Handle * h = (Handle *) p;
    if ( h->f ) return get_particles( *h->f, k, np, off, ix, (float *) x, (float *) u );
    return get_particles( *h->d, k, np, off, ix, (double *) x, (double *) u );
}

int orc_set_particles( void * p, int k, const int * np, const int * ix, const void * x, const void * u ) {
    Handle * h = (Handle *) p;
    if ( h->f ) return set_particles( *h->f, k, np, ix, (const float *) x, (const float *) u );
    return set_particles( *h->d, k, np, ix, (const double *) x, (const double *) u );
}

// which: 0 = E, 1 = B, 2 = J; buffers are tile-major, 3 reals per element
int orc_get_field( void * p, int which, void * out ) {
    Handle * h = (Handle *) p;
    if ( h->f ) {
        auto * g = field( *h->f, which ); if ( !g ) return -1;
        memcpy( out, g->d.data(), g->d.size() * sizeof( V3<float> ) );
    } else {
        auto * g = field( *h->d, which ); if ( !g ) return -1;
        memcpy( out, g->d.data(), g->d.size() * sizeof( V3<double> ) );
    }
    return 0;
}

int orc_set_field( void * p, int which, const void * in ) {
    Handle * h = (Handle *) p;
    if ( h->f ) {
        auto * g = field( *h->f, which ); if ( !g ) return -1;
        memcpy( g->d.data(), in, g->d.size() * sizeof( V3<float> ) );
    } else {
        auto * g = field( *h->d, which ); if ( !g ) return -1;
        memcpy( g->d.data(), in, g->d.size() * sizeof( V3<double> ) );
    }
    return 0;
}

// out: kinetic energy per species, then E.x E.y E.z B.x B.y B.z
int orc_energies( void * p, double * out ) {
    Handle * h = (Handle *) p;
    if ( h->f ) {
        size_t n = h->f->sp.size();
        for ( size_t k = 0; k < n; k++ ) out[k] = h->f->kinetic_energy( h->f->sp[k] );
        h->f->field_energy( out + n );
    } else {
        size_t n = h->d->sp.size();
        for ( size_t k = 0; k < n; k++ ) out[k] = h->d->kinetic_energy( h->d->sp[k] );
        h->d->field_energy( out + n );
    }
    return 0;
}

// Raw per-species push energy accumulator (sum over particles of u^2/(gamma+1))
double orc_raw_energy( void * p, int k ) {
    Handle * h = (Handle *) p;
    return h->f ? h->f->sp[k].energy : h->d->sp[k].energy;
}

// Charge density of species k on a (gc 0 lower, 1 upper) tile-major grid
// after the guard-cell add; returns the number of reals written.
long long orc_charge( void * p, int k, void * out ) {
    Handle * h = (Handle *) p;
    if ( h->f ) {
        auto g = h->f->charge( h->f->sp[k] );
        memcpy( out, g.d.data(), g.d.size() * sizeof( float ) );
        return (long long) g.d.size();
    }
    auto g = h->d->charge( h->d->sp[k] );
    memcpy( out, g.d.data(), g.d.size() * sizeof( double ) );
    return (long long) g.d.size();
}

int orc_iter( void * p ) {
    Handle * h = (Handle *) p;
    return h->f ? h->f->iter : h->d->iter;
}

} // extern "C"
