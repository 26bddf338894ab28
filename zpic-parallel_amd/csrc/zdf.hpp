// ZDF file writer: the reference's diagnostic file format (openmp/em2d/zdf.c,
// zdf.h, zdf-cpp.h), written byte for byte the same way.
//
// A file is the 4-byte magic "ZDF1" followed by records.  Every record opens
// with a header {u32 id|version, string name, u64 payload length}; strings are
// a u32 length followed by the bytes padded with zeros to a multiple of 4;
// scalars and arrays are little-endian.  Record kinds written here: string,
// int32, double, iteration, grid info, particle info and dataset (zdf.c:37-52
// hold the ids).
//
// The metadata structs keep the reference's field names and order so that
// code written against zdf::grid_info / zdf::iteration / zdf::part_info
// (designated initialisers included) compiles unchanged.
#pragma once

#include <cstdint>
#include <cstdio>
#include <string>

namespace zdf {

enum data_type {
    zdf_null = 0, zdf_int8, zdf_uint8, zdf_int16, zdf_uint16, zdf_int32, zdf_uint32,
    zdf_int64, zdf_uint64, zdf_float32, zdf_float64, zdf_complex64, zdf_complex128
};

enum axis_type { zdf_linear = 0, zdf_log10, zdf_log2 };

constexpr unsigned max_dims = 3;

struct grid_axis {
    char * name;
    axis_type type;
    double min;
    double max;
    char * label;
    char * units;
};

struct grid_info {
    char * name;
    uint32_t ndims;
    uint64_t count[max_dims];
    char * label;
    char * units;
    grid_axis * axis;
};

struct iteration {
    char * name;
    int32_t n;
    double t;
    char * time_units;
};

struct part_info {
    char * name;
    char * label;
    uint64_t np;
    uint32_t nquants;
    char ** quants;
    char ** qlabels;
    char ** qunits;
};

size_t size_of( data_type t );

template <typename T> constexpr data_type type_of() { return zdf_null; }
template <> constexpr data_type type_of<int8_t>() { return zdf_int8; }
template <> constexpr data_type type_of<uint8_t>() { return zdf_uint8; }
template <> constexpr data_type type_of<int16_t>() { return zdf_int16; }
template <> constexpr data_type type_of<uint16_t>() { return zdf_uint16; }
template <> constexpr data_type type_of<int32_t>() { return zdf_int32; }
template <> constexpr data_type type_of<uint32_t>() { return zdf_uint32; }
template <> constexpr data_type type_of<float>() { return zdf_float32; }
template <> constexpr data_type type_of<double>() { return zdf_float64; }

// Creates every missing directory of `path` (zdf.c:101-129); 0 on success.
int create_path( const std::string & path );

// One open ZDF file.  Every add_* returns false on an I/O error.
class file {
    FILE * fp = nullptr;
    uint32_t ndatasets = 0;

    bool raw( const void * p, size_t n );
    bool u32( uint32_t v ) { return raw( &v, 4 ); }
    bool i32( int32_t v ) { return raw( &v, 4 ); }
    bool u64( uint64_t v ) { return raw( &v, 8 ); }
    bool f64( double v ) { return raw( &v, 8 ); }
    bool str( const char * s );
    bool header( uint32_t id, const char * name, uint64_t length );

  public:
    file() = default;
    file( const file & ) = delete;
    file & operator=( const file & ) = delete;
    ~file() { close(); }

    bool create( const std::string & filename );
    bool close();
    bool is_open() const { return fp != nullptr; }

    bool add_string( const char * name, const char * value );
    bool add_int32( const char * name, int32_t value );
    bool add_double( const char * name, double value );
    bool add_iteration( const iteration & it );
    bool add_grid_info( const grid_info & info );
    bool add_part_info( const part_info & info );
    // Dataset record: header {id, type, ndims, count[]} then the raw data.
    bool add_dataset( const char * name, data_type type, uint32_t ndims, const uint64_t * count, const void * data );
};

// zdf_save_grid (zdf.c:1515-1585): "<path>/<name>-<n:06>.zdf" holding TYPE =
// "grid", the grid info, the iteration and one dataset.  1 on success.
int save_grid( const void * data, data_type type, const grid_info & info, const iteration & it,
               const std::string & path );

template <typename T>
int save_grid( const T * data, const grid_info & info, const iteration & it, const std::string & path ) {
    static_assert( type_of<T>() != zdf_null, "unsupported ZDF data type" );
    return save_grid( data, type_of<T>(), info, it, path );
}

// Minimal-metadata grid file (zdf-cpp.h save_grid(buffer, ndims, dims, name,
// filename)): TYPE, grid info without label/units/axes, one dataset.
int save_grid( const void * data, data_type type, unsigned ndims, const uint64_t * dims, const std::string & name,
               const std::string & filename );

// zdf_open_part_file (zdf.c:1587-1622): "<path>/particles-<name>-<n:06>.zdf"
// holding TYPE = "particles", the particle info and the iteration.
int open_part_file( file & f, const part_info & info, const iteration & it, const std::string & path );

// One particle quantity as a 1-D float32 dataset (zdf.c:1624-1640).
int add_quant_part_file( file & f, const std::string & name, const float * data, uint64_t np );

}  // namespace zdf
