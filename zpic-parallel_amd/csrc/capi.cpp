// Object entry points of zpic_amd.h: C handles over the C++ classes of
// zpic.h, the surface the reference's Cython layer binds
// (openmp/python/em2d/simulation.pxd, species.pxd, emf.pxd, current.pxd).
// Exceptions never cross the boundary: they become ZPIC_E* codes with the
// message kept for zpic_last_error().
#include "common.h"
#include "zpic.h"

#include <cstring>
#include <memory>

struct zpic_simulation {
    std::unique_ptr<Simulation> sim;
    bool deterministic = false;
};

struct zpic_species {
    std::unique_ptr<Species> sp;
    std::unique_ptr<UDistribution::Type> udist;
    bool added = false;
};

namespace {

template <class F>
int guarded( F && fn ) {
    try {
        return fn();
    } catch ( std::exception & e ) {
        return zpic::set_error( ZPIC_EINVAL, e.what() );
    }
}

}  // namespace

extern "C" {

zpic_simulation * zpic_simulation_new( unsigned ntx, unsigned nty, unsigned nx, unsigned ny, float box_x, float box_y,
                                       double dt ) {
    try {
        auto * h = new zpic_simulation();
        h->sim.reset( new Simulation( make_uint2( ntx, nty ), make_uint2( nx, ny ), make_float2( box_x, box_y ), dt ) );
        return h;
    } catch ( std::exception & e ) {
        zpic::set_error( ZPIC_EINVAL, e.what() );
        return nullptr;
    }
}

void zpic_simulation_free( zpic_simulation * h ) { delete h; }

zpic_species * zpic_species_new( const char * name, float m_q, unsigned ppc_x, unsigned ppc_y ) {
    try {
        auto * h = new zpic_species();
        h->sp.reset( new Species( name ? name : "species", m_q, make_uint2( ppc_x, ppc_y ) ) );
        return h;
    } catch ( std::exception & e ) {
        zpic::set_error( ZPIC_EINVAL, e.what() );
        return nullptr;
    }
}

void zpic_species_free( zpic_species * h ) { delete h; }

int zpic_species_set_udist( zpic_species * h, int kind, const float uth[3], const float ufl[3], int npmin ) {
    return guarded( [&] {
        if ( !h ) return zpic::set_error( ZPIC_EINVAL, "null species" );
        if ( h->added ) return zpic::set_error( ZPIC_ESTATE, "udist must be set before add_species" );
        const float3 th = uth ? make_float3( uth[0], uth[1], uth[2] ) : make_float3( 0, 0, 0 );
        const float3 fl = ufl ? make_float3( ufl[0], ufl[1], ufl[2] ) : make_float3( 0, 0, 0 );
        switch ( kind ) {
        case 0: h->sp->set_udist( UDistribution::None() ); break;
        case 1: h->sp->set_udist( UDistribution::Cold( fl ) ); break;
        case 2: h->sp->set_udist( UDistribution::Thermal( th, fl ) ); break;
        case 3: h->sp->set_udist( UDistribution::ThermalCorr( th, fl, npmin ) ); break;
        default: return zpic::set_error( ZPIC_EINVAL, "unknown udist kind" );
        }
        return ZPIC_OK;
    } );
}

int zpic_species_set_seed( zpic_species * h, unsigned offset ) {
    if ( !h ) return zpic::set_error( ZPIC_EINVAL, "null species" );
    if ( h->added ) return zpic::set_error( ZPIC_ESTATE, "seed must be set before add_species" );
    h->sp->set_seed( offset );
    return ZPIC_OK;
}

int zpic_species_set_density_uniform( zpic_species * h, float n0 ) {
    return guarded( [&] {
        if ( !h ) return zpic::set_error( ZPIC_EINVAL, "null species" );
        if ( h->added ) return zpic::set_error( ZPIC_ESTATE, "density must be set before add_species" );
        h->sp->set_density( Density::Uniform( n0 ) );
        return ZPIC_OK;
    } );
}

int zpic_species_set_density( zpic_species * h, const zpic_density * d ) {
    return guarded( [&] {
        if ( !h || !d ) return zpic::set_error( ZPIC_EINVAL, "null argument" );
        if ( h->added ) return zpic::set_error( ZPIC_ESTATE, "density must be set before add_species" );
        if ( ( d->kind == ZPIC_DENSITY_STEP || d->kind == ZPIC_DENSITY_SLAB || d->kind == ZPIC_DENSITY_RAMP ) &&
             d->dir != 0 && d->dir != 1 )
            return zpic::set_error( ZPIC_EINVAL, "profile direction must be 0 (x) or 1 (y)" );
        const coord::cart dir = d->dir == 1 ? coord::y : coord::x;
        switch ( d->kind ) {
        case ZPIC_DENSITY_NONE: h->sp->set_density( Density::None() ); break;
        case ZPIC_DENSITY_UNIFORM: h->sp->set_density( Density::Uniform( d->n0 ) ); break;
        case ZPIC_DENSITY_STEP: h->sp->set_density( Density::Step( dir, d->n0, d->pos ) ); break;
        case ZPIC_DENSITY_SLAB: h->sp->set_density( Density::Slab( dir, d->n0, d->begin, d->end ) ); break;
        case ZPIC_DENSITY_SPHERE:
            h->sp->set_density( Density::Sphere( d->n0, make_float2( d->center[0], d->center[1] ), d->radius ) );
            break;
        case ZPIC_DENSITY_RAMP: h->sp->set_density( Density::Ramp( dir, d->n0, d->begin, d->end ) ); break;
        default: return zpic::set_error( ZPIC_EINVAL, "unknown density profile" );
        }
        return ZPIC_OK;
    } );
}

int zpic_simulation_add_species( zpic_simulation * s, zpic_species * h ) {
    return guarded( [&] {
        if ( !s || !h ) return zpic::set_error( ZPIC_EINVAL, "null handle" );
        if ( h->added ) return zpic::set_error( ZPIC_ESTATE, "species already added to a simulation" );
        s->sim->add_species( *h->sp );
        h->sp->set_deterministic( s->deterministic );
        h->added = true;
        return ZPIC_OK;
    } );
}

int zpic_simulation_advance( zpic_simulation * s, int n ) {
    return guarded( [&] {
        if ( !s ) return zpic::set_error( ZPIC_EINVAL, "null simulation" );
        for ( int i = 0; i < n; i++ ) s->sim->advance();
        return ZPIC_OK;
    } );
}

int zpic_species_reinject( zpic_species * h ) {
    return guarded( [&] {
        if ( !h ) return zpic::set_error( ZPIC_EINVAL, "null species" );
        if ( !h->added ) return zpic::set_error( ZPIC_ESTATE, "species not added to a simulation" );
        h->sp->reinject();
        return ZPIC_OK;
    } );
}

int zpic_simulation_set_moving_window( zpic_simulation * s ) {
    return guarded( [&] {
        if ( !s ) return zpic::set_error( ZPIC_EINVAL, "null simulation" );
        if ( s->sim->get_iter() != 0 )
            return zpic::set_error( ZPIC_ESTATE, "(*error*) set_moving_window() called with iter != 0" );
        s->sim->set_moving_window();
        return ZPIC_OK;
    } );
}

int zpic_simulation_advance_mov_window( zpic_simulation * s, int n ) {
    return guarded( [&] {
        if ( !s ) return zpic::set_error( ZPIC_EINVAL, "null simulation" );
        for ( int i = 0; i < n; i++ ) s->sim->advance_mov_window();
        return ZPIC_OK;
    } );
}

int zpic_simulation_window_moves( const zpic_simulation * s, unsigned * n_move ) {
    if ( !s || !n_move ) return zpic::set_error( ZPIC_EINVAL, "null argument" );
    *n_move = s->sim->emf.moving_window.n_move();
    return ZPIC_OK;
}

int zpic_laser_add( zpic_simulation * s, const zpic_laser * l ) {
    return guarded( [&] {
        if ( !s || !l ) return zpic::set_error( ZPIC_EINVAL, "null argument" );
        std::unique_ptr<Laser::Pulse> p;
        if ( l->kind == 0 ) {
            p.reset( new Laser::PlaneWave() );
        } else if ( l->kind == 1 ) {
            auto * g = new Laser::Gaussian();
            g->W0 = l->W0;
            g->focus = l->focus;
            g->axis = l->axis;
            p.reset( g );
        } else {
            return zpic::set_error( ZPIC_EINVAL, "unknown laser kind (0 plane wave, 1 Gaussian)" );
        }
        p->start = l->start;
        p->fwhm = l->fwhm;
        p->rise = l->rise;
        p->flat = l->flat;
        p->fall = l->fall;
        p->a0 = l->a0;
        p->omega0 = l->omega0;
        p->polarization = l->polarization;
        p->cos_pol = l->cos_pol;
        p->sin_pol = l->sin_pol;
        p->filter = l->filter;
        if ( p->add( s->sim->emf ) != 0 ) return zpic::set_error( ZPIC_EINVAL, "(*error*) invalid laser parameters" );
        return ZPIC_OK;
    } );
}

int zpic_simulation_stage( zpic_simulation * s, int stage, int k ) {
    return guarded( [&] {
        if ( !s ) return zpic::set_error( ZPIC_EINVAL, "null simulation" );
        Simulation & sim = *s->sim;
        if ( stage <= 3 && ( k < 0 || k >= (int) sim.species.size() ) )
            return zpic::set_error( ZPIC_EINVAL, "bad species index" );
        switch ( stage ) {
        case 0: sim.species[k]->push( sim.emf.E, sim.emf.B ); break;
        case 1: sim.species[k]->move( sim.current.J ); break;
        case 2: sim.species[k]->tile_sort(); break;
        case 3: sim.species[k]->advance( sim.emf, sim.current ); break;
        case 4: sim.current.zero(); break;
        case 5: sim.current.advance(); break;
        case 6: sim.emf.advance( sim.current ); break;
        case 7: sim.current.J->add_from_gc(); break;
        case 8: sim.current.J->copy_to_gc(); break;
        case 9: sim.current.filter->apply( *sim.current.J ); break;
        case 10: sim.emf.process_bc(); break;
        case 11: sim.emf.advance(); break;
        default: return zpic::set_error( ZPIC_EINVAL, "bad stage" );
        }
        return ZPIC_OK;
    } );
}

int zpic_current_set_filter( zpic_simulation * s, int kind, int dir, int order ) {
    return guarded( [&] {
        if ( !s ) return zpic::set_error( ZPIC_EINVAL, "null simulation" );
        const coord::cart d = dir ? coord::y : coord::x;
        switch ( kind ) {
        case 0: s->sim->current.set_filter( Filter::None() ); break;
        case 1: s->sim->current.set_filter( Filter::Binomial( d, order ) ); break;
        case 2: s->sim->current.set_filter( Filter::Compensated( d, order ) ); break;
        default: return zpic::set_error( ZPIC_EINVAL, "unknown filter kind" );
        }
        return ZPIC_OK;
    } );
}

int zpic_simulation_set_deterministic( zpic_simulation * s, int on ) {
    if ( !s ) return zpic::set_error( ZPIC_EINVAL, "null simulation" );
    s->deterministic = on != 0;
    s->sim->set_deterministic( on != 0 );
    return ZPIC_OK;
}

unsigned zpic_simulation_get_iter( const zpic_simulation * s ) { return s ? s->sim->get_iter() : 0; }
double zpic_simulation_get_t( const zpic_simulation * s ) { return s ? s->sim->get_t() : 0.0; }
uint64_t zpic_simulation_get_nmove( zpic_simulation * s ) { return s ? s->sim->get_nmove() : 0; }

double zpic_species_get_energy( zpic_species * h ) {
    try {
        return h ? h->sp->get_energy() : 0.0;
    } catch ( std::exception & e ) {
        zpic::set_error( ZPIC_EHIP, e.what() );
        return 0.0;
    }
}

int zpic_emf_get_energy( zpic_simulation * s, double ene_E[3], double ene_B[3] ) {
    return guarded( [&] {
        if ( !s ) return zpic::set_error( ZPIC_EINVAL, "null simulation" );
        double3 e, b;
        s->sim->emf.get_energy( e, b );
        ene_E[0] = e.x; ene_E[1] = e.y; ene_E[2] = e.z;
        ene_B[0] = b.x; ene_B[1] = b.y; ene_B[2] = b.z;
        return ZPIC_OK;
    } );
}

uint64_t zpic_species_get_nmove( const zpic_species * h ) { return h ? h->sp->get_nmove() : 0; }

uint64_t zpic_species_np_total( zpic_species * h ) {
    try {
        return h && h->added ? h->sp->np_total() : 0;
    } catch ( std::exception & e ) {
        zpic::set_error( ZPIC_EHIP, e.what() );
        return 0;
    }
}

int zpic_species_get_particles( zpic_species * h, int * np, int * offset, int * ix, float * x, float * u ) {
    return guarded( [&] {
        if ( !h || !h->added ) return zpic::set_error( ZPIC_ESTATE, "species not initialised" );
        h->sp->get_particles( np, offset, ix, x, u );
        return ZPIC_OK;
    } );
}

int zpic_species_set_particles( zpic_species * h, const int * np, const int * ix, const float * x, const float * u ) {
    return guarded( [&] {
        if ( !h || !h->added ) return zpic::set_error( ZPIC_ESTATE, "species not initialised" );
        if ( !np || !ix || !x || !u ) return zpic::set_error( ZPIC_EINVAL, "null buffer" );
        h->sp->set_particles( np, ix, x, u );
        return ZPIC_OK;
    } );
}

static vec3grid<float3> * field_of( zpic_simulation * s, int which ) {
    switch ( which ) {
    case 0: return s->sim->emf.E;
    case 1: return s->sim->emf.B;
    case 2: return s->sim->current.J;
    }
    return nullptr;
}

int zpic_field_device( zpic_simulation * s, int which, float ** buf, zpic_geom * g, int periodic[2] ) {
    return guarded( [&] {
        vec3grid<float3> * f = s ? field_of( s, which ) : nullptr;
        if ( !f || !buf || !g || !periodic ) return zpic::set_error( ZPIC_EINVAL, "bad field request" );
        *buf = f->fbuf();
        *g = f->geom();
        const int2 p = f->get_periodic();
        periodic[0] = p.x;
        periodic[1] = p.y;
        return ZPIC_OK;
    } );
}

int zpic_simulation_info( const zpic_simulation * s, unsigned ntiles[2], unsigned nx[2], float box[2], double * dt ) {
    if ( !s ) return zpic::set_error( ZPIC_EINVAL, "null simulation" );
    const Simulation & sim = *s->sim;
    if ( ntiles ) { ntiles[0] = sim.ntiles.x; ntiles[1] = sim.ntiles.y; }
    if ( nx ) { nx[0] = sim.nx.x; nx[1] = sim.nx.y; }
    if ( box ) { box[0] = sim.box.x; box[1] = sim.box.y; }
    if ( dt ) *dt = sim.dt;
    return ZPIC_OK;
}

int zpic_emf_get_iter( const zpic_simulation * s ) { return s ? s->sim->emf.get_iter() : -1; }
int zpic_current_get_iter( const zpic_simulation * s ) { return s ? s->sim->current.get_iter() : -1; }

int zpic_simulation_energy_info( zpic_simulation * s ) {
    return guarded( [&] {
        if ( !s ) return zpic::set_error( ZPIC_EINVAL, "null simulation" );
        s->sim->energy_info();
        return ZPIC_OK;
    } );
}

int zpic_species_get_phasespace( zpic_species * h, int q0, float min0, float max0, int size0, int q1, float min1,
                                 float max1, int size1, float * out ) {
    return guarded( [&] {
        if ( !h || !h->added ) return zpic::set_error( ZPIC_ESTATE, "species not initialised" );
        if ( !out ) return zpic::set_error( ZPIC_EINVAL, "null output" );
        if ( q0 < 0 || q0 > 4 || size0 < 1 || !( max0 > min0 ) ) return zpic::set_error( ZPIC_EINVAL, "bad axis 0" );
        const size_t n = (size_t) size0 * ( size1 > 0 ? size1 : 1 );
        float * d = zpic::device_alloc<float>( n );
        try {
            if ( size1 <= 0 ) {
                h->sp->dep_phasespace( d, static_cast<phasespace::quant>( q0 ), make_float2( min0, max0 ), size0 );
            } else {
                if ( q1 < 0 || q1 > 4 || q1 == q0 || !( max1 > min1 ) )
                    throw std::runtime_error( "Cannot generate a 2d phasespace with these axes" );
                h->sp->dep_phasespace( d, static_cast<phasespace::quant>( q0 ), make_float2( min0, max0 ), size0,
                                       static_cast<phasespace::quant>( q1 ), make_float2( min1, max1 ), size1 );
            }
            zpic::hip_check( hipMemcpyAsync( out, d, n * sizeof( float ), hipMemcpyDeviceToHost, zpic::stream() ),
                             "phasespace copy" );
            zpic::hip_check( hipStreamSynchronize( zpic::stream() ), "sync" );
        } catch ( ... ) {
            zpic::device_free( d );
            throw;
        }
        zpic::device_free( d );
        return ZPIC_OK;
    } );
}

int zpic_field_get( zpic_simulation * s, int which, float * out ) {
    return guarded( [&] {
        vec3grid<float3> * f = s ? field_of( s, which ) : nullptr;
        if ( !f || !out ) return zpic::set_error( ZPIC_EINVAL, "bad field request" );
        f->download( reinterpret_cast<float3 *>( out ) );
        return ZPIC_OK;
    } );
}

int zpic_field_set( zpic_simulation * s, int which, const float * in ) {
    return guarded( [&] {
        vec3grid<float3> * f = s ? field_of( s, which ) : nullptr;
        if ( !f || !in ) return zpic::set_error( ZPIC_EINVAL, "bad field request" );
        f->upload( reinterpret_cast<const float3 *>( in ) );
        return ZPIC_OK;
    } );
}

int zpic_field_geom( const zpic_simulation * s, zpic_geom * g ) {
    if ( !s || !g ) return zpic::set_error( ZPIC_EINVAL, "null handle" );
    *g = s->sim->emf.E->geom();
    return ZPIC_OK;
}

long long zpic_species_charge( zpic_species * h, float * out ) {
    try {
        if ( !h || !h->added ) return zpic::set_error( ZPIC_ESTATE, "species not initialised" );
        bnd<unsigned int> gc;
        gc.x = { 0, 1 };
        gc.y = { 0, 1 };
        grid<float> rho( h->sp->get_ntiles(), h->sp->get_nx(), gc );
        rho.zero();
        h->sp->deposit_charge( rho );
        rho.add_from_gc();
        if ( out ) rho.download( out );
        return (long long) rho.buffer_size();
    } catch ( std::exception & e ) {
        return zpic::set_error( ZPIC_EINVAL, e.what() );
    }
}

int zpic_species_gather( zpic_species * h, int quant, float * out ) {
    return guarded( [&] {
        if ( !h || !h->added ) return zpic::set_error( ZPIC_ESTATE, "species not initialised" );
        if ( quant < 0 || quant > 4 || !out ) return zpic::set_error( ZPIC_EINVAL, "bad gather request" );
        h->sp->gather( static_cast<part::quant>( quant ), out );
        return ZPIC_OK;
    } );
}

int zpic_simulation_set_timing( zpic_simulation * s, int on ) {
    if ( !s ) return zpic::set_error( ZPIC_EINVAL, "null simulation" );
    s->sim->set_timing( on != 0 );
    return ZPIC_OK;
}

int zpic_simulation_get_timing( zpic_simulation * s, double out[5] ) {
    if ( !s ) return zpic::set_error( ZPIC_EINVAL, "null simulation" );
    s->sim->get_timing( out );
    return ZPIC_OK;
}

int zpic_simulation_sync( zpic_simulation * s ) {
    (void) s;
    return guarded( [&] {
        zpic::hip_check( hipStreamSynchronize( zpic::stream() ), "hipStreamSynchronize" );
        return ZPIC_OK;
    } );
}

}  // extern "C"

struct zpic_batch {
    std::unique_ptr<Batch> b;
};

extern "C" {

zpic_batch * zpic_batch_new( int threads_per_tile ) {
    if ( threads_per_tile < 64 || threads_per_tile > 512 || threads_per_tile % 64 ) {
        zpic::set_error( ZPIC_EINVAL, "threads per tile must be a multiple of 64 in [64, 512]" );
        return nullptr;
    }
    auto * h = new zpic_batch();
    h->b.reset( new Batch( threads_per_tile ) );
    return h;
}

void zpic_batch_free( zpic_batch * b ) { delete b; }

int zpic_batch_add( zpic_batch * b, zpic_simulation * s ) { return zpic_batch_add_threads( b, s, 0 ); }

int zpic_batch_add_threads( zpic_batch * b, zpic_simulation * s, int threads_per_tile ) {
    return guarded( [&] {
        if ( !b || !s ) return zpic::set_error( ZPIC_EINVAL, "null handle" );
        if ( s->deterministic ) return zpic::set_error( ZPIC_EINVAL, "batches run the fast mode only" );
        b->b->add( *s->sim, threads_per_tile );
        return ZPIC_OK;
    } );
}

int zpic_batch_set_graph( zpic_batch * b, int on ) {
    if ( !b ) return zpic::set_error( ZPIC_EINVAL, "null batch" );
    b->b->set_graph( on != 0 );
    return ZPIC_OK;
}

int zpic_batch_set_persistent( zpic_batch * b, int on ) {
    return guarded( [&] {
        if ( !b ) return zpic::set_error( ZPIC_EINVAL, "null batch" );
        b->b->set_persistent( on != 0 );
        return ZPIC_OK;
    } );
}

int zpic_batch_status( zpic_batch * b ) {
    return guarded( [&] {
        if ( !b ) return zpic::set_error( ZPIC_EINVAL, "null batch" );
        b->b->check_status();
        return ZPIC_OK;
    } );
}

int zpic_batch_advance( zpic_batch * b, int n ) {
    return guarded( [&] {
        if ( !b ) return zpic::set_error( ZPIC_EINVAL, "null batch" );
        for ( int i = 0; i < n; i++ ) b->b->advance();
        return ZPIC_OK;
    } );
}

}  // extern "C"

extern "C" {

int zpic_emf_save( zpic_simulation * s, int field, int fc ) {
    return guarded( [&] {
        if ( !s ) return zpic::set_error( ZPIC_EINVAL, "null simulation" );
        if ( field < 0 || field > 1 || fc < 0 || fc > 2 ) return zpic::set_error( ZPIC_EINVAL, "bad field or component" );
        s->sim->emf.save( field ? emf::b : emf::e, static_cast<fcomp::cart>( fc ) );
        return ZPIC_OK;
    } );
}

int zpic_current_save( zpic_simulation * s, int fc ) {
    return guarded( [&] {
        if ( !s ) return zpic::set_error( ZPIC_EINVAL, "null simulation" );
        if ( fc < 0 || fc > 2 ) return zpic::set_error( ZPIC_EINVAL, "bad component" );
        s->sim->current.save( static_cast<fcomp::cart>( fc ) );
        return ZPIC_OK;
    } );
}

int zpic_field_save_file( zpic_simulation * s, int which, int fc, const char * filename ) {
    return guarded( [&] {
        if ( !s || !filename ) return zpic::set_error( ZPIC_EINVAL, "null argument" );
        if ( which < 0 || which > 2 || fc < 0 || fc > 2 ) return zpic::set_error( ZPIC_EINVAL, "bad field or component" );
        vec3grid<float3> * f = which == 0 ? s->sim->emf.E : which == 1 ? s->sim->emf.B : s->sim->current.J;
        f->save( static_cast<fcomp::cart>( fc ), std::string( filename ) );
        return ZPIC_OK;
    } );
}

int zpic_species_save( zpic_species * h ) {
    return guarded( [&] {
        if ( !h || !h->added ) return zpic::set_error( ZPIC_ESTATE, "species not initialised" );
        h->sp->save();
        return ZPIC_OK;
    } );
}

int zpic_species_save_charge( zpic_species * h ) {
    return guarded( [&] {
        if ( !h || !h->added ) return zpic::set_error( ZPIC_ESTATE, "species not initialised" );
        h->sp->save_charge();
        return ZPIC_OK;
    } );
}

int zpic_species_save_phasespace( zpic_species * h, int q0, float min0, float max0, int size0, int q1, float min1,
                                  float max1, int size1 ) {
    return guarded( [&] {
        if ( !h || !h->added ) return zpic::set_error( ZPIC_ESTATE, "species not initialised" );
        if ( q0 < 0 || q0 > 4 || size0 < 1 || !( max0 > min0 ) ) return zpic::set_error( ZPIC_EINVAL, "bad axis 0" );
        const auto p0 = static_cast<phasespace::quant>( q0 );
        if ( size1 <= 0 ) {
            h->sp->save_phasespace( p0, make_float2( min0, max0 ), size0 );
            return ZPIC_OK;
        }
        if ( q1 <= q0 || q1 > 4 || !( max1 > min1 ) )
            return zpic::set_error( ZPIC_EINVAL, "bad axis 1 (its quantity must come after axis 0's)" );
        h->sp->save_phasespace( p0, make_float2( min0, max0 ), size0, static_cast<phasespace::quant>( q1 ),
                                make_float2( min1, max1 ), size1 );
        return ZPIC_OK;
    } );
}

}  // extern "C"

extern "C" int zpic_zdf_save_grid( const void * data, int data_type, const zpic_zdf_grid_info * info,
                                   const zpic_zdf_iteration * it, const char * path ) {
    if ( !data || !info || !it || !path ) return zpic::set_error( ZPIC_EINVAL, "null argument" );
    if ( info->ndims < 1 || info->ndims > zdf::max_dims || zdf::size_of( static_cast<zdf::data_type>( data_type ) ) == 0 )
        return zpic::set_error( ZPIC_EINVAL, "bad grid dimensions or data type" );
    zdf::grid_axis axis[zdf::max_dims];
    zdf::grid_info g{};
    g.name = const_cast<char *>( info->name );
    g.ndims = info->ndims;
    for ( unsigned i = 0; i < info->ndims; i++ ) g.count[i] = info->count[i];
    g.label = const_cast<char *>( info->label );
    g.units = const_cast<char *>( info->units );
    if ( info->axis ) {
        for ( unsigned i = 0; i < info->ndims; i++ ) {
            const zpic_zdf_axis & a = info->axis[i];
            axis[i] = zdf::grid_axis{ const_cast<char *>( a.name ), static_cast<zdf::axis_type>( a.type ), a.min, a.max,
                                      const_cast<char *>( a.label ), const_cast<char *>( a.units ) };
        }
        g.axis = axis;
    }
    zdf::iteration iter{ const_cast<char *>( it->name ), it->n, it->t, const_cast<char *>( it->time_units ) };
    if ( zdf::save_grid( data, static_cast<zdf::data_type>( data_type ), g, iter, path ) != 1 )
        return zpic::set_error( ZPIC_EINVAL, std::string( "unable to write ZDF grid file under " ) + path );
    return ZPIC_OK;
}

extern "C" {

int zpic_current_set_learned_filter( zpic_simulation * s, zpic_denoiser * d ) {
    return guarded( [&] {
        if ( !s || !d ) return zpic::set_error( ZPIC_EINVAL, "null handle" );
        int H = 0, W = 0;
        zpic_denoiser_size( d, &H, &W );
        const uint2 dims = s->sim->current.J->get_dims();
        if ( (int) dims.x != W || (int) dims.y != H ) return zpic::set_error( ZPIC_EINVAL, "denoiser size does not match the grid" );
        s->sim->current.set_filter( Filter::Learned( d ) );
        return ZPIC_OK;
    } );
}

int zpic_current_set_fallback( zpic_simulation * s, int on ) {
    if ( !s ) return zpic::set_error( ZPIC_EINVAL, "null simulation" );
    auto * f = dynamic_cast<Filter::Learned *>( s->sim->current.filter );
    if ( !f ) return zpic::set_error( ZPIC_ESTATE, "the current filter is not the learned one" );
    f->set_fallback( on != 0 );
    return ZPIC_OK;
}

int zpic_current_get_fallback( zpic_simulation * s ) {
    if ( !s ) return -1;
    auto * f = dynamic_cast<Filter::Learned *>( s->sim->current.filter );
    return f ? ( f->get_fallback() ? 1 : 0 ) : -1;
}

int zpic_simulation_copy_fields( zpic_simulation * dst, const zpic_simulation * src ) {
    return guarded( [&] {
        if ( !dst || !src ) return zpic::set_error( ZPIC_EINVAL, "null handle" );
        const auto & a = *dst->sim->emf.E;
        const auto & b = *src->sim->emf.E;
        if ( a.buffer_size() != b.buffer_size() || a.get_dims().x != b.get_dims().x || a.get_dims().y != b.get_dims().y ||
             a.nx.x != b.nx.x || a.nx.y != b.nx.y )
            return zpic::set_error( ZPIC_EINVAL, "grids differ" );
        const size_t bytes = a.buffer_size() * sizeof( float3 );
        zpic::hip_check( hipMemcpyAsync( dst->sim->emf.E->d_buffer, src->sim->emf.E->d_buffer, bytes,
                                         hipMemcpyDeviceToDevice, zpic::stream() ), "copy E" );
        zpic::hip_check( hipMemcpyAsync( dst->sim->emf.B->d_buffer, src->sim->emf.B->d_buffer, bytes,
                                         hipMemcpyDeviceToDevice, zpic::stream() ), "copy B" );
        return ZPIC_OK;
    } );
}

int zpic_species_subsample( zpic_species * dst, zpic_species * src, int stride, int phase ) {
    return guarded( [&] {
        if ( !dst || !src || !dst->added || !src->added ) return zpic::set_error( ZPIC_ESTATE, "species not initialised" );
        dst->sp->subsample_from( *src->sp, stride, phase );
        return ZPIC_OK;
    } );
}

}  // extern "C"

extern "C" int zpic_simulation_add_pair( zpic_denoiser * d, zpic_simulation * in, zpic_simulation * target ) {
    return guarded( [&] {
        if ( !d || !in || !target ) return zpic::set_error( ZPIC_EINVAL, "null handle" );
        zpic_geom g = in->sim->current.J->geom(), gt = target->sim->current.J->geom();
        if ( g.nx != gt.nx || g.ny != gt.ny || g.ntiles_x != gt.ntiles_x || g.ntiles_y != gt.ntiles_y )
            return zpic::set_error( ZPIC_EINVAL, "pair grids must use the same grid and tiling" );
        return zpic_denoiser_add_pair( d, in->sim->current.J->fbuf(), target->sim->current.J->fbuf(), &g,
                                       zpic::stream() );
    } );
}

extern "C" {

int zpic_species_set_bc( zpic_species * h, const int bc[4] ) {
    return guarded( [&] {
        if ( !h || !bc ) return zpic::set_error( ZPIC_EINVAL, "null argument" );
        for ( int k = 0; k < 4; k++ )
            if ( bc[k] < 0 || bc[k] > 2 ) return zpic::set_error( ZPIC_EINVAL, "species bc must be 0 open, 1 periodic, 2 reflecting" );
        species::bc_type b;
        b.x.lower = (species::bc::type) bc[0];
        b.x.upper = (species::bc::type) bc[1];
        b.y.lower = (species::bc::type) bc[2];
        b.y.upper = (species::bc::type) bc[3];
        h->sp->set_bc( b );
        return ZPIC_OK;
    } );
}

int zpic_emf_set_absorbing( zpic_simulation * s, int width, float alpha ) {
    return guarded( [&] {
        if ( !s ) return zpic::set_error( ZPIC_EINVAL, "null simulation" );
        s->sim->emf.set_absorbing( width, alpha );
        return ZPIC_OK;
    } );
}

int zpic_emf_set_bc( zpic_simulation * s, const int bc[4] ) {
    return guarded( [&] {
        if ( !s || !bc ) return zpic::set_error( ZPIC_EINVAL, "null argument" );
        for ( int k = 0; k < 4; k++ )
            if ( bc[k] < 0 || bc[k] > ZPIC_EMF_ABSORBING )
                return zpic::set_error( ZPIC_EINVAL, "emf bc must be 0 none, 1 periodic, 2 pec, 3 pmc, 4 absorbing" );
        emf::bc_type b;
        b.x.lower = (emf::bc::type) bc[0];
        b.x.upper = (emf::bc::type) bc[1];
        b.y.lower = (emf::bc::type) bc[2];
        b.y.upper = (emf::bc::type) bc[3];
        s->sim->emf.set_bc( b );
        return ZPIC_OK;
    } );
}

int zpic_current_set_bc( zpic_simulation * s, const int bc[4] ) {
    return guarded( [&] {
        if ( !s || !bc ) return zpic::set_error( ZPIC_EINVAL, "null argument" );
        for ( int k = 0; k < 4; k++ )
            if ( bc[k] < 0 || bc[k] > 2 ) return zpic::set_error( ZPIC_EINVAL, "current bc must be 0 none, 1 periodic, 2 reflecting" );
        current::bc_type b;
        b.x.lower = (current::bc::type) bc[0];
        b.x.upper = (current::bc::type) bc[1];
        b.y.lower = (current::bc::type) bc[2];
        b.y.upper = (current::bc::type) bc[3];
        s->sim->current.set_bc( b );
        return ZPIC_OK;
    } );
}

}  // extern "C"
