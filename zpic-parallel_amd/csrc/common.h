// Shared helpers of the zpic HIP library: error reporting and launch checks.
#pragma once

#include <hip/hip_runtime.h>

#include <string>

#include "zpic_amd.h"

namespace zpic {

// Records the message returned by zpic_last_error() and passes the code through.
int set_error( int code, const std::string & msg );

inline int hip_error( hipError_t e, const char * where ) {
    return set_error( ZPIC_EHIP, std::string( where ) + ": " + hipGetErrorString( e ) );
}

// Global-memory (address space 1) view of a buffer.  Pointers that may come
// from an argument table in memory are generic to the compiler, and generic
// accesses become flat instructions, which also count against the LDS wait
// counter: every LDS wait would then wait for the outstanding loads too.
// The host pass of hipcc also parses device code, for a target without that
// address space: there the view is a plain pointer.
#if defined( __HIP_DEVICE_COMPILE__ )
template <class T> using gptr = __attribute__( ( address_space( 1 ) ) ) T *;
#else
template <class T> using gptr = T *;
#endif
template <class T> __device__ inline gptr<T> gbl( T * p ) { return (gptr<T>) p; }

// The library's stream (default priority, non-blocking), or null if no
// device is present (launches then fail and report the error).
hipStream_t default_stream_or_null();

// The stream of a stage entry point: the caller's, or the library's for NULL.
inline hipStream_t as_stream( void * s ) { return s ? (hipStream_t) s : default_stream_or_null(); }

}  // namespace zpic

#define ZPIC_HIP_TRY( call )                                               \
    do {                                                                   \
        hipError_t e_ = ( call );                                          \
        if ( e_ != hipSuccess ) return zpic::hip_error( e_, #call );       \
    } while ( 0 )

// Checks the launch that was just enqueued.
#define ZPIC_LAUNCH_CHECK( name )                                          \
    do {                                                                   \
        hipError_t e_ = hipGetLastError();                                 \
        if ( e_ != hipSuccess ) return zpic::hip_error( e_, name );        \
    } while ( 0 )
