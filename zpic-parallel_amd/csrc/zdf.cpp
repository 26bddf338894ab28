// ZDF writer (format of openmp/em2d/zdf.c; see zdf.hpp for the layout).
#include "zdf.hpp"

#include <cerrno>
#include <cstring>
#include <sys/stat.h>
#include <sys/types.h>

namespace zdf {

namespace {

// Record ids | versions (zdf.c:37-52)
constexpr uint32_t ID_INT32 = 0x00010000;
constexpr uint32_t ID_DOUBLE = 0x00020000;
constexpr uint32_t ID_STRING = 0x00030000;
constexpr uint32_t ID_DATASET = 0x00100002;
constexpr uint32_t ID_ITERATION = 0x00200001;
constexpr uint32_t ID_GRID_INFO = 0x00210001;
constexpr uint32_t ID_PART_INFO = 0x00220002;

constexpr size_t UNIT = 4;  // every string is padded to a multiple of 4 bytes
const char MAGIC[4] = { 'Z', 'D', 'F', '1' };

size_t pad4( size_t n ) { return ( n + UNIT - 1 ) / UNIT * UNIT; }

// Bytes a string takes in a record payload.
size_t str_size( const char * s ) {
    const size_t n = s ? std::strlen( s ) : 0;
    return 4 + ( n ? pad4( n ) : 0 );
}

size_t grid_info_size( const grid_info & g ) {
    size_t n = 4 + 8 * (size_t) g.ndims + str_size( g.label ) + str_size( g.units ) + 4;
    if ( g.axis )
        for ( unsigned i = 0; i < g.ndims; i++ )
            n += str_size( g.axis[i].name ) + 4 + 16 + str_size( g.axis[i].label ) + str_size( g.axis[i].units );
    return n;
}

size_t part_info_size( const part_info & p ) {
    size_t n = str_size( p.label ) + 8 + 4;
    for ( unsigned i = 0; i < p.nquants; i++ )
        n += str_size( p.quants[i] ) + str_size( p.qlabels[i] ) + str_size( p.qunits[i] );
    return n;
}

}  // namespace

size_t size_of( data_type t ) {
    switch ( t ) {
    case zdf_int8: case zdf_uint8: return 1;
    case zdf_int16: case zdf_uint16: return 2;
    case zdf_int32: case zdf_uint32: case zdf_float32: return 4;
    case zdf_int64: case zdf_uint64: case zdf_float64: case zdf_complex64: return 8;
    case zdf_complex128: return 16;
    default: return 0;
    }
}

int create_path( const std::string & path ) {
    if ( path.empty() ) return 0;
    if ( mkdir( path.c_str(), 0755 ) == 0 || errno == EEXIST ) return 0;
    if ( errno != ENOENT ) return errno;
    const size_t cut = path.find_last_of( '/' );
    if ( cut == std::string::npos || cut == 0 ) return ENOENT;
    if ( int rc = create_path( path.substr( 0, cut ) ) ) return rc;
    return ( mkdir( path.c_str(), 0755 ) == 0 || errno == EEXIST ) ? 0 : errno;
}

bool file::raw( const void * p, size_t n ) { return n == 0 || std::fwrite( p, 1, n, fp ) == n; }

bool file::str( const char * s ) {
    const uint32_t n = s ? (uint32_t) std::strlen( s ) : 0;
    if ( !u32( n ) ) return false;
    if ( n == 0 ) return true;
    static const char zeros[UNIT] = { 0, 0, 0, 0 };
    return raw( s, n ) && raw( zeros, pad4( n ) - n );
}

bool file::header( uint32_t id, const char * name, uint64_t length ) { return u32( id ) && str( name ) && u64( length ); }

bool file::create( const std::string & filename ) {
    close();
    fp = std::fopen( filename.c_str(), "w+b" );
    if ( !fp ) return false;
    ndatasets = 0;
    if ( !raw( MAGIC, sizeof MAGIC ) ) {
        close();
        return false;
    }
    return true;
}

bool file::close() {
    if ( !fp ) return true;
    const bool ok = std::fclose( fp ) == 0;
    fp = nullptr;
    return ok;
}

bool file::add_string( const char * name, const char * value ) {
    return header( ID_STRING, name, str_size( value ) ) && str( value );
}

bool file::add_int32( const char * name, int32_t value ) { return header( ID_INT32, name, 4 ) && i32( value ); }

bool file::add_double( const char * name, double value ) { return header( ID_DOUBLE, name, 8 ) && f64( value ); }

bool file::add_iteration( const iteration & it ) {
    const char * name = it.name ? it.name : "Iteration";
    return header( ID_ITERATION, name, 4 + 8 + str_size( it.time_units ) ) && i32( it.n ) && f64( it.t ) &&
           str( it.time_units );
}

bool file::add_grid_info( const grid_info & g ) {
    if ( !header( ID_GRID_INFO, g.name, grid_info_size( g ) ) || !u32( g.ndims ) ) return false;
    for ( unsigned i = 0; i < g.ndims; i++ )
        if ( !u64( g.count[i] ) ) return false;
    if ( !str( g.label ) || !str( g.units ) || !i32( g.axis != nullptr ) ) return false;
    if ( g.axis )
        for ( unsigned i = 0; i < g.ndims; i++ ) {
            const grid_axis & a = g.axis[i];
            if ( !str( a.name ) || !i32( a.type ) || !f64( a.min ) || !f64( a.max ) || !str( a.label ) ||
                 !str( a.units ) )
                return false;
        }
    return true;
}

bool file::add_part_info( const part_info & p ) {
    if ( !header( ID_PART_INFO, p.name, part_info_size( p ) ) || !str( p.label ) || !u64( p.np ) ||
         !u32( p.nquants ) )
        return false;
    for ( unsigned i = 0; i < p.nquants; i++ )
        if ( !str( p.quants[i] ) ) return false;
    for ( unsigned i = 0; i < p.nquants; i++ )
        if ( !str( p.qlabels[i] ) ) return false;
    for ( unsigned i = 0; i < p.nquants; i++ )
        if ( !str( p.qunits[i] ) ) return false;
    return true;
}

bool file::add_dataset( const char * name, data_type type, uint32_t ndims, const uint64_t * count, const void * data ) {
    uint64_t n = 1;
    for ( unsigned i = 0; i < ndims; i++ ) n *= count[i];
    const size_t bytes = size_of( type ) * n;
    const size_t hdr = 4 + 4 + 4 + 8 * (size_t) ndims;
    if ( !header( ID_DATASET, name, bytes + hdr ) ) return false;
    if ( !u32( ++ndatasets ) || !i32( type ) || !u32( ndims ) ) return false;
    for ( unsigned i = 0; i < ndims; i++ )
        if ( !u64( count[i] ) ) return false;
    return raw( data, bytes );
}

namespace {

std::string numbered( const std::string & dir, const std::string & stem, int n ) {
    char num[16];
    std::snprintf( num, sizeof num, "%06u", (unsigned) n );
    return dir + "/" + stem + "-" + num + ".zdf";
}

}  // namespace

int save_grid( const void * data, data_type type, const grid_info & info, const iteration & it,
               const std::string & path ) {
    create_path( path );
    file f;
    if ( !f.create( numbered( path, info.name ? info.name : "", it.n ) ) ) return 0;
    if ( !f.add_string( "TYPE", "grid" ) || !f.add_grid_info( info ) || !f.add_iteration( it ) ) return 0;
    if ( !f.add_dataset( info.name, type, info.ndims, info.count, data ) ) return 0;
    return f.close() ? 1 : 0;
}

int save_grid( const void * data, data_type type, unsigned ndims, const uint64_t * dims, const std::string & name,
               const std::string & filename ) {
    file f;
    if ( !f.create( filename ) ) return -1;
    grid_info info{};
    info.name = const_cast<char *>( name.c_str() );
    info.ndims = ndims;
    for ( unsigned i = 0; i < ndims && i < max_dims; i++ ) info.count[i] = dims[i];
    if ( !f.add_string( "TYPE", "grid" ) || !f.add_grid_info( info ) ) return 0;
    if ( !f.add_dataset( info.name, type, ndims, info.count, data ) ) return 0;
    return f.close() ? 1 : 0;
}

int open_part_file( file & f, const part_info & info, const iteration & it, const std::string & path ) {
    create_path( path );
    if ( !f.create( numbered( path, std::string( "particles-" ) + ( info.name ? info.name : "" ), it.n ) ) )
        return -1;
    if ( !f.add_string( "TYPE", "particles" ) || !f.add_part_info( info ) || !f.add_iteration( it ) ) return 0;
    return 1;
}

int add_quant_part_file( file & f, const std::string & name, const float * data, uint64_t np ) {
    return f.add_dataset( name.c_str(), zdf_float32, 1, &np, data ) ? 1 : 0;
}

}  // namespace zdf
