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
