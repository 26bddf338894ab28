// Test-infrastructure driver for the REFERENCE em2d OpenMP code.
//
// This file is compiled against the unmodified reference sources
// (openmp/em2d/*.cpp, read in place from the reference checkout by
// oracle/Makefile) into oracle/_ref/ref_driver.  It is a checker only: it
// produces golden vectors (dump mode) that pin the oracle restatement, and it
// times the reference CPU path (bench mode) as the CPU baseline.  Nothing in
// the product links or runs it.
//
// It reproduces the call sequence of the reference Weibel driver
// (openmp/em2d/main.cpp:179-259): Simulation(ntiles, nx, box, dt) with
// `float dt` widened to double, two ThermalCorr species added in the order
// electrons (id 1) then positrons (id 2), then Simulation::advance() in a loop.
//
// Usage:
//   ref_driver dump  <out.bin> key=value ...
//   ref_driver bench key=value ...
//   ref_driver zdf   <outdir> key=value ...   (advance `steps` steps, then
//                    write the diagnostics set below into outdir)
// keys: ntiles=X,Y nx=X,Y ppc=X,Y box=X,Y dt=F uth=X,Y,Z ufl=X,Y,Z
//       filter=none|bx<N>|by<N>|cx<N>|cy<N>  snap=s0,s1,...  steps=N warmup=N
//       energies=1 (record energies after every step up to the last snapshot)
//       sbc=xl,xu,yl,yu  ebc=...  cbc=...  boundary conditions of the species
//       (0 open, 1 periodic, 2 reflecting), the EMF (0 none, 1 periodic, 2 pec,
//       3 pmc) and the current (0 none, 1 periodic, 2 reflecting)
//       plasma=weibel|electrons  two ThermalCorr species (default), or one
//       electron species (q/m = -1) with udist cold (ufl) or thermal (uth > 0)
//       dens=uniform | step:D,n0,pos | slab:D,n0,begin,end | sphere:n0,cx,cy,r
//       (D = x or y) density profile of every species
//       laser=none | plane:start,fwhm,a0,omega0,cos_pol,sin_pol
//       | gauss:start,fwhm,a0,omega0,W0,focus,axis,cos_pol,sin_pol
//       window=1  Simulation::set_moving_window, advance_mov_window steps;
//       wsbc=xl,xu,yl,yu  species boundaries set after set_moving_window
// Dumped grids have their alignment padding (elements ext_x*ext_y ..
// tile_vol of every tile, never read) zeroed: the laser's temporary grids
// leave uninitialised values there.

#include "simulation.h"
#include "species.h"
#include "udist.h"
#include "filter.h"
#include "density.h"
#include "laser.h"

#include <chrono>
#include <unistd.h>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>
#include <map>

#ifdef _OPENMP
#include <omp.h>
#endif

// Exposes the protected particle buffer of the reference Species class.
class ProbeSpecies : public Species {
  public:
    using Species::Species;
    Particles * parts() { return particles; }
    float2 cell() const { return dx; }
};

// The diagnostics set written by `zdf` mode with the reference's own writers:
// EMF::save (all components), Current::save, the minimal-metadata
// vec3grid::save(fc, filename), and per species save_charge, 1-D and 2-D
// save_phasespace.  Species::save casts std::string arrays to char**
// (species.cpp:2445-2447) and faults in zdf_add_part_info; its particle file
// is written here with the metadata that code intends (names "x" "y" "ux"
// "uy" "uz", labels "x" "y" "u_x" "u_y" "u_z", units "c/\\omega_n" twice then
// "c" three times), through the same zdf::open_part_file /
// add_quant_part_file calls and Particles::gather scales.
static void write_zdf( Simulation & sim, std::vector<ProbeSpecies*> & sp ) {
    for ( int f = 0; f < 2; f++ )
        for ( int c = 0; c < 3; c++ ) sim.emf.save( f ? emf::b : emf::e, (fcomp::cart) c );
    for ( int c = 0; c < 3; c++ ) sim.current.save( (fcomp::cart) c );
    sim.emf.E->save( fcomp::y, "E-y-min.zdf" );
    sim.current.J->save( fcomp::z, "J-z-min.zdf" );
    for ( auto * s : sp ) {
        s->save_charge();
        const float2 box = s->get_box();
        s->save_phasespace( phasespace::x, make_float2( 0.0f, box.x ), 64 );
        s->save_phasespace( phasespace::y, make_float2( 0.5f, box.y - 0.3f ), 50 );
        s->save_phasespace( phasespace::ux, make_float2( -0.3f, 0.3f ), 40 );
        s->save_phasespace( phasespace::uz, make_float2( -1.0f, 1.0f ), 32 );
        s->save_phasespace( phasespace::x, make_float2( 0.0f, box.x ), 32, phasespace::uz, make_float2( -1.0f, 1.0f ), 24 );
        s->save_phasespace( phasespace::ux, make_float2( -0.25f, 0.25f ), 20, phasespace::uy, make_float2( -0.25f, 0.25f ), 20 );
        s->save_phasespace( phasespace::y, make_float2( 0.0f, box.y ), 16, phasespace::ux, make_float2( -0.4f, 0.4f ), 12 );

        char qn[5][3] = { "x", "y", "ux", "uy", "uz" };
        char ql[5][4] = { "x", "y", "u_x", "u_y", "u_z" };
        char qu[5][12] = { "c/\\omega_n", "c/\\omega_n", "c", "c", "c" };
        char * quants[5], * qlabels[5], * qunits[5];
        for ( int i = 0; i < 5; i++ ) { quants[i] = qn[i]; qlabels[i] = ql[i]; qunits[i] = qu[i]; }
        zdf::part_info info = {
            .name = (char *) s->name.c_str(),
            .label = (char *) s->name.c_str(),
            .nquants = 5,
            .quants = quants,
            .qlabels = qlabels,
            .qunits = qunits,
        };
        Particles * p = s->parts();
        const uint32_t np = p->np_total();
        info.np = np;
        zdf::iteration it = { .n = s->get_iter(), .t = s->get_iter() * s->get_dt(), .time_units = (char *) "1/\\omega_n" };
        zdf::file f;
        zdf::open_part_file( f, info, it, "PARTICLES/" + s->name );
        std::vector<float> data( np );
        const float2 d = s->cell();
        p->gather( part::quant::x, data.data(), make_float2( d.x, 0 ) );
        zdf::add_quant_part_file( f, "x", data.data(), np );
        p->gather( part::quant::y, data.data(), make_float2( d.y, 0 ) );
        zdf::add_quant_part_file( f, "y", data.data(), np );
        p->gather( part::quant::ux, data.data() );
        zdf::add_quant_part_file( f, "ux", data.data(), np );
        p->gather( part::quant::uy, data.data() );
        zdf::add_quant_part_file( f, "uy", data.data(), np );
        p->gather( part::quant::uz, data.data() );
        zdf::add_quant_part_file( f, "uz", data.data(), np );
        zdf::close_file( f );
    }
}

struct Args {
    uint2 ntiles = make_uint2( 8, 8 );
    uint2 nx = make_uint2( 16, 16 );
    uint2 ppc = make_uint2( 8, 8 );
    float2 box = make_float2( 12.8f, 12.8f );
    float dt = 0.07f;
    float3 uth = make_float3( 0.1f, 0.1f, 0.1f );
    float3 ufl = make_float3( 0.0f, 0.0f, 0.6f );
    std::string filter = "none";
    std::vector<int> snap;
    int steps = 10;
    int warmup = 1;
    int energies = 0;
    int sbc[4] = { 1, 1, 1, 1 };
    int ebc[4] = { 1, 1, 1, 1 };
    int cbc[4] = { 1, 1, 1, 1 };
    std::string plasma = "weibel";
    std::string dens = "uniform";
    std::string laser = "none";
    int window = 0;
    int wsbc[4] = { -1, -1, -1, -1 };
};

static std::vector<std::string> split( const std::string & s, char c ) {
    std::vector<std::string> out; std::string cur;
    for ( char ch : s ) { if ( ch == c ) { out.push_back( cur ); cur.clear(); } else cur += ch; }
    out.push_back( cur );
    return out;
}

static Args parse( int argc, char ** argv, int first ) {
    Args a;
    for ( int i = first; i < argc; i++ ) {
        std::string kv = argv[i];
        auto p = kv.find( '=' );
        if ( p == std::string::npos ) { fprintf( stderr, "bad arg %s\n", argv[i] ); exit( 2 ); }
        std::string k = kv.substr( 0, p ), v = kv.substr( p + 1 );
        auto f = split( v, ',' );
        if ( k == "ntiles" ) a.ntiles = make_uint2( atoi( f[0].c_str() ), atoi( f[1].c_str() ) );
        else if ( k == "nx" ) a.nx = make_uint2( atoi( f[0].c_str() ), atoi( f[1].c_str() ) );
        else if ( k == "ppc" ) a.ppc = make_uint2( atoi( f[0].c_str() ), atoi( f[1].c_str() ) );
        else if ( k == "box" ) a.box = make_float2( strtof( f[0].c_str(), 0 ), strtof( f[1].c_str(), 0 ) );
        else if ( k == "dt" ) a.dt = strtof( v.c_str(), 0 );
        else if ( k == "uth" ) a.uth = make_float3( strtof( f[0].c_str(), 0 ), strtof( f[1].c_str(), 0 ), strtof( f[2].c_str(), 0 ) );
        else if ( k == "ufl" ) a.ufl = make_float3( strtof( f[0].c_str(), 0 ), strtof( f[1].c_str(), 0 ), strtof( f[2].c_str(), 0 ) );
        else if ( k == "filter" ) a.filter = v;
        else if ( k == "snap" ) { for ( auto & s : f ) a.snap.push_back( atoi( s.c_str() ) ); }
        else if ( k == "steps" ) a.steps = atoi( v.c_str() );
        else if ( k == "warmup" ) a.warmup = atoi( v.c_str() );
        else if ( k == "energies" ) a.energies = atoi( v.c_str() );
        else if ( k == "sbc" || k == "ebc" || k == "cbc" || k == "wsbc" ) {
            int * dst = ( k == "sbc" ) ? a.sbc : ( k == "ebc" ) ? a.ebc : ( k == "cbc" ) ? a.cbc : a.wsbc;
            for ( int i = 0; i < 4; i++ ) dst[i] = atoi( f[i].c_str() );
        }
        else if ( k == "plasma" ) a.plasma = v;
        else if ( k == "dens" ) a.dens = v;
        else if ( k == "laser" ) a.laser = v;
        else if ( k == "window" ) a.window = atoi( v.c_str() );
        else { fprintf( stderr, "unknown key %s\n", k.c_str() ); exit( 2 ); }
    }
    return a;
}

// Tagged record: u32 name length, name bytes, 1 dtype char, u64 count, payload.
static void put( FILE * fp, const std::string & name, char dtype, uint64_t count, const void * data ) {
    uint32_t n = name.size();
    size_t esize = ( dtype == 'd' || dtype == 'q' ) ? 8 : 4;
    fwrite( &n, 4, 1, fp ); fwrite( name.data(), 1, n, fp );
    fwrite( &dtype, 1, 1, fp ); fwrite( &count, 8, 1, fp );
    if ( count ) fwrite( data, esize, count, fp );
}

static void set_filter( Simulation & sim, const std::string & f ) {
    if ( f == "none" ) return;
    unsigned order = atoi( f.c_str() + 2 );
    coord::cart dir = ( f[1] == 'x' ) ? coord::x : coord::y;
    if ( f[0] == 'b' ) sim.current.set_filter( Filter::Binomial( dir, order ) );
    else if ( f[0] == 'c' ) sim.current.set_filter( Filter::Compensated( dir, order ) );
    else { fprintf( stderr, "bad filter %s\n", f.c_str() ); exit( 2 ); }
}

// Copy of a tile-major grid with the alignment padding of every tile zeroed.
template <class T>
static std::vector<T> unpadded( grid<T> & g ) {
    std::vector<T> out( g.d_buffer, g.d_buffer + g.buffer_size() );
    const size_t used = (size_t) g.ext_nx.x * g.ext_nx.y;
    const size_t nt = g.buffer_size() / g.tile_vol;
    T zero;
    std::memset( (void *) &zero, 0, sizeof( T ) );
    for ( size_t t = 0; t < nt; t++ )
        for ( size_t i = used; i < g.tile_vol; i++ ) out[t * g.tile_vol + i] = zero;
    return out;
}

static Density::Profile * make_density( const std::string & d ) {
    if ( d == "uniform" ) return new Density::Uniform( 1.0f );
    auto c = d.find( ':' );
    if ( c == std::string::npos ) { fprintf( stderr, "bad dens %s\n", d.c_str() ); exit( 2 ); }
    std::string kind = d.substr( 0, c );
    auto f = split( d.substr( c + 1 ), ',' );
    auto fl = [&]( size_t i ) { return strtof( f.at( i ).c_str(), 0 ); };
    if ( kind == "step" ) return new Density::Step( f.at( 0 ) == "y" ? coord::y : coord::x, fl( 1 ), fl( 2 ) );
    if ( kind == "slab" ) return new Density::Slab( f.at( 0 ) == "y" ? coord::y : coord::x, fl( 1 ), fl( 2 ), fl( 3 ) );
    if ( kind == "sphere" ) return new Density::Sphere( fl( 0 ), make_float2( fl( 1 ), fl( 2 ) ), fl( 3 ) );
    fprintf( stderr, "bad dens %s\n", d.c_str() );
    exit( 2 );
}

static void add_laser( Simulation & sim, const std::string & l ) {
    if ( l == "none" ) return;
    auto c = l.find( ':' );
    std::string kind = l.substr( 0, c );
    auto f = split( l.substr( c + 1 ), ',' );
    auto fl = [&]( size_t i ) { return strtof( f.at( i ).c_str(), 0 ); };
    int rc;
    if ( kind == "plane" ) {
        Laser::PlaneWave w;
        w.start = fl( 0 ); w.fwhm = fl( 1 ); w.a0 = fl( 2 ); w.omega0 = fl( 3 );
        w.cos_pol = fl( 4 ); w.sin_pol = fl( 5 );
        rc = w.add( sim.emf );
    } else if ( kind == "gauss" ) {
        Laser::Gaussian w;
        w.start = fl( 0 ); w.fwhm = fl( 1 ); w.a0 = fl( 2 ); w.omega0 = fl( 3 );
        w.W0 = fl( 4 ); w.focus = fl( 5 ); w.axis = fl( 6 );
        w.cos_pol = fl( 7 ); w.sin_pol = fl( 8 );
        rc = w.add( sim.emf );
    } else {
        fprintf( stderr, "bad laser %s\n", l.c_str() );
        exit( 2 );
    }
    if ( rc ) { fprintf( stderr, "laser add failed\n" ); exit( 1 ); }
}

static void snapshot( FILE * fp, int step, Simulation & sim, std::vector<ProbeSpecies*> & sp ) {
    std::string pre = "s" + std::to_string( step ) + "/";
    for ( size_t k = 0; k < sp.size(); k++ ) {
        Particles * p = sp[k]->parts();
        unsigned nt = p->ntiles.x * p->ntiles.y;
        std::string sn = pre + "sp" + std::to_string( k ) + "/";
        put( fp, sn + "np", 'i', nt, p->np );
        put( fp, sn + "offset", 'i', nt, p->offset );
        std::vector<int> ix; std::vector<float> x, u;
        for ( unsigned t = 0; t < nt; t++ ) {
            for ( int i = p->offset[t]; i < p->offset[t] + p->np[t]; i++ ) {
                ix.push_back( p->ix[i].x ); ix.push_back( p->ix[i].y );
                x.push_back( p->x[i].x ); x.push_back( p->x[i].y );
                u.push_back( p->u[i].x ); u.push_back( p->u[i].y ); u.push_back( p->u[i].z );
            }
        }
        put( fp, sn + "ix", 'i', ix.size(), ix.data() );
        put( fp, sn + "x", 'f', x.size(), x.data() );
        put( fp, sn + "u", 'f', u.size(), u.data() );
        double e = sp[k]->get_energy();
        put( fp, sn + "energy", 'd', 1, &e );
        uint64_t nm = sp[k]->get_nmove();
        put( fp, sn + "nmove", 'q', 1, &nm );
        // charge density as Species::save_charge builds it (species.cpp:2508-2522)
        bnd<unsigned int> cgc;
        cgc.x = { 0, 1 };
        cgc.y = { 0, 1 };
        grid<float> charge( p->ntiles, p->nx, cgc );
        charge.zero();
        sp[k]->deposit_charge( charge );
        charge.add_from_gc();
        auto ch = unpadded( charge );
        put( fp, sn + "charge", 'f', ch.size(), ch.data() );
    }
    size_t fsz = sim.emf.E->buffer_size() * 3;
    auto vE = unpadded( *sim.emf.E ), vB = unpadded( *sim.emf.B ), vJ = unpadded( *sim.current.J );
    put( fp, pre + "E", 'f', fsz, vE.data() );
    put( fp, pre + "B", 'f', fsz, vB.data() );
    put( fp, pre + "J", 'f', fsz, vJ.data() );
    double3 eE, eB;
    sim.emf.get_energy( eE, eB );
    double fe[6] = { eE.x, eE.y, eE.z, eB.x, eB.y, eB.z };
    put( fp, pre + "field_energy", 'd', 6, fe );
}

int main( int argc, char ** argv ) {
    if ( argc < 2 ) { fprintf( stderr, "usage: ref_driver dump|bench ...\n" ); return 2; }
    std::string mode = argv[1];
    bool dump = ( mode == "dump" );
    bool zdfmode = ( mode == "zdf" );
    Args a = parse( argc, argv, ( dump || zdfmode ) ? 3 : 2 );

    Simulation sim( a.ntiles, a.nx, a.box, (double) a.dt );
    set_filter( sim, a.filter );

    const bool weibel = ( a.plasma == "weibel" );
    if ( !weibel && a.plasma != "electrons" ) { fprintf( stderr, "bad plasma %s\n", a.plasma.c_str() ); return 2; }
    Density::Profile * dens = make_density( a.dens );

    ProbeSpecies electrons( "electrons", -1.0f, a.ppc );
    electrons.set_density( *dens );
    if ( weibel ) electrons.set_udist( UDistribution::ThermalCorr( a.uth, a.ufl ) );
    else if ( a.uth.x > 0 || a.uth.y > 0 || a.uth.z > 0 ) electrons.set_udist( UDistribution::Thermal( a.uth, a.ufl ) );
    else electrons.set_udist( UDistribution::Cold( a.ufl ) );
    sim.add_species( electrons );

    ProbeSpecies positrons( "positrons", +1.0f, a.ppc );
    std::vector<ProbeSpecies*> sp = { &electrons };
    if ( weibel ) {
        positrons.set_density( *dens );
        positrons.set_udist( UDistribution::ThermalCorr( a.uth, make_float3( -a.ufl.x, -a.ufl.y, -a.ufl.z ) ) );
        sim.add_species( positrons );
        sp.push_back( &positrons );
    }
    delete dens;

    {
        species::bc_type sb;
        sb.x = { (species::bc::type) a.sbc[0], (species::bc::type) a.sbc[1] };
        sb.y = { (species::bc::type) a.sbc[2], (species::bc::type) a.sbc[3] };
        for ( auto * s : sp ) s->set_bc( sb );
        emf::bc_type eb;
        eb.x = { (emf::bc::type) a.ebc[0], (emf::bc::type) a.ebc[1] };
        eb.y = { (emf::bc::type) a.ebc[2], (emf::bc::type) a.ebc[3] };
        sim.emf.set_bc( eb );
        current::bc_type cb;
        cb.x = { (current::bc::type) a.cbc[0], (current::bc::type) a.cbc[1] };
        cb.y = { (current::bc::type) a.cbc[2], (current::bc::type) a.cbc[3] };
        sim.current.set_bc( cb );
    }

    add_laser( sim, a.laser );
    if ( a.window ) {
        sim.set_moving_window();
        if ( a.wsbc[0] >= 0 ) {
            species::bc_type sb;
            sb.x = { (species::bc::type) a.wsbc[0], (species::bc::type) a.wsbc[1] };
            sb.y = { (species::bc::type) a.wsbc[2], (species::bc::type) a.wsbc[3] };
            for ( auto * s : sp ) s->set_bc( sb );
        }
    }
    auto advance = [&]() {
        if ( a.window ) sim.advance_mov_window();
        else sim.advance();
    };

    if ( zdfmode ) {
        for ( int i = 0; i < a.steps; i++ ) advance();
        if ( chdir( argv[2] ) ) { perror( "chdir" ); return 1; }
        write_zdf( sim, sp );
        return 0;
    }

    if ( dump ) {
        FILE * fp = fopen( argv[2], "wb" );
        if ( !fp ) { perror( "fopen" ); return 1; }
        int last = 0;
        for ( int s : a.snap ) last = std::max( last, s );
        std::vector<double> ene;
        for ( int step = 0; step <= last; step++ ) {
            if ( step > 0 ) advance();
            for ( int s : a.snap ) if ( s == step ) snapshot( fp, step, sim, sp );
            if ( a.energies && step > 0 ) {
                double3 eE, eB;
                sim.emf.get_energy( eE, eB );
                for ( auto * s : sp ) ene.push_back( s->get_energy() );
                ene.push_back( eE.x ); ene.push_back( eE.y ); ene.push_back( eE.z );
                ene.push_back( eB.x ); ene.push_back( eB.y ); ene.push_back( eB.z );
            }
        }
        if ( a.energies ) put( fp, "energies", 'd', ene.size(), ene.data() );
        fclose( fp );
        return 0;
    }

    // bench: warmup steps, then `steps` timed steps of Simulation::advance()
    for ( int i = 0; i < a.warmup; i++ ) advance();
    uint64_t n0 = sim.get_nmove();
    auto t0 = std::chrono::steady_clock::now();
    for ( int i = 0; i < a.steps; i++ ) advance();
    auto t1 = std::chrono::steady_clock::now();
    uint64_t moved = sim.get_nmove() - n0;
    double sec = std::chrono::duration<double>( t1 - t0 ).count();
    int threads = 1;
#ifdef _OPENMP
    threads = omp_get_max_threads();
#endif
    printf( "{\"particles_per_ns\": %.6g, \"ms_per_step\": %.6g, \"steps\": %d, \"threads\": %d, "
            "\"particles\": %llu}\n",
            moved / sec / 1e9, 1e3 * sec / a.steps, a.steps, threads,
            (unsigned long long) ( electrons.np_total() + ( weibel ? positrons.np_total() : 0 ) ) );
    return 0;
}
