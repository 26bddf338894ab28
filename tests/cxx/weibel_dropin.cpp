// The reference's Weibel driver (openmp/em2d/main.cpp:179-259) compiled
// against this library's C++ API (zpic-parallel_amd/csrc/zpic.h) instead of
// the reference's simulation.h: same classes, constructors and calls.  The
// set-up is the 32x32 case of the golden vectors (2x2 tiles of 16x16 cells,
// box 3.2, ppc 2x2), with the deterministic mode switched on (the one line
// the reference does not have) so that the run reproduces the reference's
// single-thread results bit for bit.
//
//     weibel_dropin <outdir>
//
// writes, after steps 0, 1 and 5, every array of the golden vectors as raw
// little-endian files <outdir>/s<step>__<name>.bin (E, B, J with guards;
// per species np, offset, ix, x, u, charge, energy, nmove; field energies).
#include "zpic.h"

#include <cstdio>
#include <string>
#include <vector>

namespace {

std::string g_out;

template <class T>
void put( const std::string & name, const T * data, size_t n ) {
    const std::string path = g_out + "/" + name + ".bin";
    FILE * f = std::fopen( path.c_str(), "wb" );
    if ( !f || std::fwrite( data, sizeof( T ), n, f ) != n ) {
        std::fprintf( stderr, "cannot write %s\n", path.c_str() );
        std::exit( 1 );
    }
    std::fclose( f );
}

void dump( Simulation & sim, std::vector<Species *> & sp, int step ) {
    const std::string s = "s" + std::to_string( step ) + "__";
    const char * names[3] = { "E", "B", "J" };
    vec3grid<float3> * grids[3] = { sim.emf.E, sim.emf.B, sim.current.J };
    for ( int g = 0; g < 3; g++ ) {
        std::vector<float3> h( grids[g]->buffer_size() );
        grids[g]->download( h.data() );
        put( s + names[g], reinterpret_cast<const float *>( h.data() ), 3 * h.size() );
    }
    double3 eE, eB;
    sim.emf.get_energy( eE, eB );
    const double fe[6] = { eE.x, eE.y, eE.z, eB.x, eB.y, eB.z };
    put( s + "field_energy", fe, 6 );
    for ( size_t k = 0; k < sp.size(); k++ ) {
        const std::string p = s + "sp" + std::to_string( k ) + "__";
        const size_t n = sp[k]->np_total();
        const uint2 nt = sim.ntiles;
        std::vector<int> np( nt.x * nt.y ), off( nt.x * nt.y ), ix( 2 * n );
        std::vector<float> x( 2 * n ), u( 3 * n );
        sp[k]->get_particles( np.data(), off.data(), ix.data(), x.data(), u.data() );
        put( p + "np", np.data(), np.size() );
        put( p + "offset", off.data(), off.size() );
        put( p + "ix", ix.data(), ix.size() );
        put( p + "x", x.data(), x.size() );
        put( p + "u", u.data(), u.size() );
        // charge density on a (gc 0,1) grid after the guard add (Species::save_charge)
        bnd<unsigned int> gc;
        gc.x.lower = gc.y.lower = 0;
        gc.x.upper = gc.y.upper = 1;
        grid<float> charge( sim.ntiles, sim.nx, gc );
        charge.zero();
        sp[k]->deposit_charge( charge );
        charge.add_from_gc();
        std::vector<float> rho( charge.buffer_size() );
        charge.download( rho.data() );
        put( p + "charge", rho.data(), rho.size() );
        const double e = sp[k]->get_energy();
        put( p + "energy", &e, 1 );
        const uint64_t nm = sp[k]->get_nmove();
        put( p + "nmove", &nm, 1 );
    }
}

}  // namespace

int main( int argc, char ** argv ) {
    if ( argc < 2 ) {
        std::fprintf( stderr, "usage: %s <outdir>\n", argv[0] );
        return 2;
    }
    g_out = argv[1];
    try {
        // main.cpp:179-259 with the 32x32 golden set-up
        uint2 ntiles = make_uint2( 2, 2 );
        uint2 nx = make_uint2( 16, 16 );
        float2 box = make_float2( 3.2, 3.2 );
        float dt = 0.07;
        Simulation sim( ntiles, nx, box, dt );
        uint2 ppc = make_uint2( 2, 2 );
        Species electrons( "electrons", -1.0f, ppc );
        electrons.set_udist( UDistribution::ThermalCorr( make_float3( 0.1, 0.1, 0.1 ), make_float3( 0, 0, 0.6 ) ) );
        sim.add_species( electrons );
        Species positrons( "positrons", +1.0f, ppc );
        positrons.set_udist( UDistribution::ThermalCorr( make_float3( 0.1, 0.1, 0.1 ), make_float3( 0, 0, -0.6 ) ) );
        sim.add_species( positrons );
        sim.set_deterministic( true );
        std::vector<Species *> sp = { &electrons, &positrons };
        dump( sim, sp, 0 );
        for ( int step = 1; step <= 5; step++ ) {
            sim.advance();
            if ( step == 1 || step == 5 ) dump( sim, sp, step );
        }
        std::printf( "weibel_dropin: %u steps, t = %g, nmove %llu\n", sim.get_iter(), sim.get_t(),
                     (unsigned long long) sim.get_nmove() );
    } catch ( std::exception & e ) {
        std::fprintf( stderr, "%s\n", e.what() );
        return 1;
    }
    return 0;
}
