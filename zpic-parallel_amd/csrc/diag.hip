// Diagnostic deposits of a species on gfx950: 1-D and 2-D phase-space
// densities (dep_pha1_kernel / dep_pha2_kernel, species.cpp:2579-2620 and
// :2786-2850) with the reference's arithmetic: bin coordinate
// n = (d - min) * (size / (max - min)) - 0.5, k = int(n + 1) - 1, linear
// weights (1 - w, w), the last bin of each axis never written.
//
// Fast mode: one workgroup per tile; the tile's contributions go to a
// workgroup-private histogram in LDS when it fits (float LDS atomics), and
// are added to the output with global float atomics.  Deterministic mode: one
// wavefront walks the tiles and their particles in the reference's order and
// adds them in that order, so the result is bit-identical to the reference's
// (single-thread) deposit.
#include "common.h"

#include <vector>

namespace zpic {
namespace {

constexpr int WAVE = 64;
constexpr int PHA_BLOCK = 256;
constexpr int PHA_LDS_BINS = 12288;  // 48 KiB of float bins

struct PhaAxis {
    int quant;   // 0 x, 1 y, 2 ux, 3 uy, 4 uz
    float min;   // range in cells for x and y, in c for u
    float rdx;   // size / (max - min)
    int size;
};

struct PhaArgs {
    const int * np;
    const int * off;
    const int2 * ix;
    const float2 * x;
    const float3 * u;
    int ntx, nx, ny;
    PhaAxis a0, a1;  // a1.size == 0 for a 1-D phase space
    float norm;
};

__device__ inline float pha_value( int quant, int shiftx, int shifty, int2 c, float2 x, float3 u ) {
    switch ( quant ) {
    case 0: return (float) ( shiftx + c.x ) + ( x.x + 0.5f );
    case 1: return (float) ( shifty + c.y ) + ( x.y + 0.5f );
    case 2: return u.x;
    case 3: return u.y;
    default: return u.z;
    }
}

// Bin index k and weight w of a value along one axis.
__device__ inline void pha_bin( const PhaAxis & a, float d, int & k, float & w ) {
    const float n = ( d - a.min ) * a.rdx - 0.5f;
    k = int( n + 1 ) - 1;
    w = n - k;
}

__device__ inline bool in_axis( int k, int size ) { return k >= 0 && k < size - 1; }

// The (up to four) contributions of one particle, in the reference's order.
template <class ADD>
__device__ inline void pha_deposit( const PhaArgs & a, int shiftx, int shifty, int2 c, float2 x, float3 u, ADD add ) {
    int k0;
    float w0;
    pha_bin( a.a0, pha_value( a.a0.quant, shiftx, shifty, c, x, u ), k0, w0 );
    if ( a.a1.size == 0 ) {
        if ( in_axis( k0, a.a0.size ) ) add( k0, ( 1 - w0 ) * a.norm );
        if ( in_axis( k0 + 1, a.a0.size ) ) add( k0 + 1, w0 * a.norm );
        return;
    }
    int k1;
    float w1;
    pha_bin( a.a1, pha_value( a.a1.quant, shiftx, shifty, c, x, u ), k1, w1 );
    const int s0 = a.a0.size;
    if ( in_axis( k0, s0 ) && in_axis( k1, a.a1.size ) ) add( k1 * s0 + k0, ( 1 - w0 ) * ( 1 - w1 ) * a.norm );
    if ( in_axis( k0 + 1, s0 ) && in_axis( k1, a.a1.size ) ) add( k1 * s0 + k0 + 1, w0 * ( 1 - w1 ) * a.norm );
    if ( in_axis( k0, s0 ) && in_axis( k1 + 1, a.a1.size ) ) add( ( k1 + 1 ) * s0 + k0, ( 1 - w0 ) * w1 * a.norm );
    if ( in_axis( k0 + 1, s0 ) && in_axis( k1 + 1, a.a1.size ) )
        add( ( k1 + 1 ) * s0 + k0 + 1, w0 * w1 * a.norm );
}

__global__ void __launch_bounds__( PHA_BLOCK ) k_phasespace( PhaArgs a, float * __restrict__ out, int bins ) {
    extern __shared__ float hist[];
    const int t = blockIdx.x;
    const bool local = bins <= PHA_LDS_BINS;
    if ( local ) {
        for ( int k = threadIdx.x; k < bins; k += blockDim.x ) hist[k] = 0.0f;
        __syncthreads();
    }
    const int shiftx = ( t % a.ntx ) * a.nx, shifty = ( t / a.ntx ) * a.ny;
    const int off = a.off[t], np = a.np[t];
    for ( int i = threadIdx.x; i < np; i += blockDim.x ) {
        const int2 c = a.ix[off + i];
        const float2 x = a.x[off + i];
        const float3 u = a.u[off + i];
        if ( local )
            pha_deposit( a, shiftx, shifty, c, x, u, [&]( int k, float v ) { atomicAdd( &hist[k], v ); } );
        else
            pha_deposit( a, shiftx, shifty, c, x, u, [&]( int k, float v ) { atomicAdd( &out[k], v ); } );
    }
    if ( local ) {
        __syncthreads();
        for ( int k = threadIdx.x; k < bins; k += blockDim.x )
            if ( hist[k] != 0.0f ) atomicAdd( &out[k], hist[k] );
    }
}

// One wavefront: particles of each tile in order; lane l's contributions are
// added after those of lanes < l.
__global__ void __launch_bounds__( WAVE ) k_phasespace_det( PhaArgs a, float * __restrict__ out, int ntiles ) {
    const int lane = threadIdx.x;
    for ( int t = 0; t < ntiles; t++ ) {
        const int shiftx = ( t % a.ntx ) * a.nx, shifty = ( t / a.ntx ) * a.ny;
        const int off = a.off[t], np = a.np[t];
        for ( int i0 = 0; i0 < np; i0 += WAVE ) {
            const int i = i0 + lane;
            int2 c = make_int2( 0, 0 );
            float2 x = make_float2( 0.f, 0.f );
            float3 u = make_float3( 0.f, 0.f, 0.f );
            if ( i < np ) {
                c = a.ix[off + i];
                x = a.x[off + i];
                u = a.u[off + i];
            }
            for ( int l = 0; l < WAVE && i0 + l < np; l++ ) {
                if ( lane == l )
                    pha_deposit( a, shiftx, shifty, c, x, u, [&]( int k, float v ) {
                        volatile float * o = out + k;
                        *o = *o + v;
                    } );
                __builtin_amdgcn_wave_barrier();
            }
        }
    }
}

}  // namespace
}  // namespace zpic

using namespace zpic;

extern "C" int zpic_phasespace( const zpic_particles * p, int ntiles_x, int ntiles_y, int nx, int ny, int quant0,
                                float min0, float max0, int size0, int quant1, float min1, float max1, int size1,
                                float norm, float * out, int deterministic, void * stream ) {
    if ( !p || !out ) return set_error( ZPIC_EINVAL, "null buffer" );
    if ( ntiles_x < 1 || ntiles_y < 1 || nx < 1 || ny < 1 ) return set_error( ZPIC_EINVAL, "invalid tile geometry" );
    if ( quant0 < 0 || quant0 > 4 || size0 < 1 || !( max0 > min0 ) )
        return set_error( ZPIC_EINVAL, "invalid phase-space axis 0" );
    const bool two = size1 > 0;
    if ( two && ( quant1 <= quant0 || quant1 > 4 || !( max1 > min1 ) ) )
        return set_error( ZPIC_EINVAL, "invalid phase-space axis 1 (its quantity must come after axis 0's)" );
    PhaArgs a;
    a.np = p->np; a.off = p->offset; a.ix = (const int2 *) p->ix; a.x = (const float2 *) p->x;
    a.u = (const float3 *) p->u;
    a.ntx = ntiles_x; a.nx = nx; a.ny = ny;
    a.a0 = PhaAxis{ quant0, min0, size0 / ( max0 - min0 ), size0 };
    a.a1 = two ? PhaAxis{ quant1, min1, size1 / ( max1 - min1 ), size1 } : PhaAxis{ 0, 0.f, 0.f, 0 };
    a.norm = norm;
    const long long bins = (long long) size0 * ( two ? size1 : 1 );
    if ( bins > ( 1ll << 30 ) ) return set_error( ZPIC_ERANGE, "phase-space grid too large" );
    hipStream_t st = as_stream( stream );
    ZPIC_HIP_TRY( hipMemsetAsync( out, 0, sizeof( float ) * bins, st ) );
    const int nt = ntiles_x * ntiles_y;
    if ( deterministic ) {
        hipLaunchKernelGGL( k_phasespace_det, dim3( 1 ), dim3( WAVE ), 0, st, a, out, nt );
    } else {
        const size_t lds = bins <= PHA_LDS_BINS ? sizeof( float ) * bins : 0;
        hipLaunchKernelGGL( k_phasespace, dim3( nt ), dim3( PHA_BLOCK ), lds, st, a, out, (int) bins );
    }
    ZPIC_LAUNCH_CHECK( "k_phasespace" );
    return ZPIC_OK;
}

// --------------------------------------------------------------------------
// Every stride-th particle (from index `phase`) of each tile of src into dst,
// in tile order; dst->np / dst->offset receive the counts and their
// exclusive scan (computed on the host from src's counts).
namespace zpic {
namespace {

__global__ void __launch_bounds__( 256 ) k_subsample( const int * __restrict__ snp, const int * __restrict__ soff,
                                                      const int2 * __restrict__ six, const float2 * __restrict__ sx,
                                                      const float3 * __restrict__ su, const int * __restrict__ dnp,
                                                      const int * __restrict__ doff, int2 * __restrict__ dix,
                                                      float2 * __restrict__ dx, float3 * __restrict__ du, int stride,
                                                      int phase ) {
    const int t = blockIdx.x;
    const int so = soff[t], n = dnp[t], d0 = doff[t];
    (void) snp;
    for ( int i = threadIdx.x; i < n; i += blockDim.x ) {
        const int k = so + phase + i * stride;
        dix[d0 + i] = six[k];
        dx[d0 + i] = sx[k];
        du[d0 + i] = su[k];
    }
}

}  // namespace
}  // namespace zpic

extern "C" int zpic_subsample( const zpic_particles * src, const zpic_geom * src_geom, zpic_particles * dst,
                               const zpic_geom * dst_geom, int stride, int phase, void * stream ) {
    if ( !src || !dst || !src_geom || !dst_geom || stride < 1 || phase < 0 || phase >= stride )
        return set_error( ZPIC_EINVAL, "invalid subsample request" );
    if ( src_geom->ntiles_x != dst_geom->ntiles_x || src_geom->ntiles_y != dst_geom->ntiles_y ||
         src_geom->nx != dst_geom->nx || src_geom->ny != dst_geom->ny )
        return set_error( ZPIC_EINVAL, "subsample needs species with the same tiling" );
    const long long nt = (long long) src_geom->ntiles_x * src_geom->ntiles_y;
    if ( src_geom->ntiles_x < 1 || src_geom->ntiles_y < 1 || nt > ( 1 << 30 ) )
        return set_error( ZPIC_EINVAL, "invalid tile count" );
    const int ntiles = (int) nt;
    hipStream_t st = as_stream( stream );
    std::vector<int> np( ntiles ), off( ntiles ), dnp( ntiles ), doff( ntiles );
    ZPIC_HIP_TRY( hipMemcpyAsync( np.data(), src->np, sizeof( int ) * ntiles, hipMemcpyDeviceToHost, st ) );
    ZPIC_HIP_TRY( hipMemcpyAsync( off.data(), src->offset, sizeof( int ) * ntiles, hipMemcpyDeviceToHost, st ) );
    ZPIC_HIP_TRY( hipStreamSynchronize( st ) );
    long long total = 0;
    for ( int t = 0; t < ntiles; t++ ) {
        dnp[t] = np[t] > phase ? ( np[t] - phase + stride - 1 ) / stride : 0;
        doff[t] = (int) total;
        total += dnp[t];
    }
    if ( total > (long long) dst->max_part ) return set_error( ZPIC_ERANGE, "subsample exceeds the destination's capacity" );
    ZPIC_HIP_TRY( hipMemcpyAsync( dst->np, dnp.data(), sizeof( int ) * ntiles, hipMemcpyHostToDevice, st ) );
    ZPIC_HIP_TRY( hipMemcpyAsync( dst->offset, doff.data(), sizeof( int ) * ntiles, hipMemcpyHostToDevice, st ) );
    hipLaunchKernelGGL( k_subsample, dim3( ntiles ), dim3( 256 ), 0, st, src->np, src->offset, (const int2 *) src->ix,
                        (const float2 *) src->x, (const float3 *) src->u, dst->np, dst->offset, (int2 *) dst->ix,
                        (float2 *) dst->x, (float3 *) dst->u, stride, phase );
    ZPIC_LAUNCH_CHECK( "k_subsample" );
    // the host count vectors must outlive the copies
    ZPIC_HIP_TRY( hipStreamSynchronize( st ) );
    return ZPIC_OK;
}
