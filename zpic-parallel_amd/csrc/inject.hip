// Particle injection on gfx950 (Density::*::np_inject / inject,
// density.cpp:103-1067): the lattice points of a cell range that a density
// profile keeps, counted or appended per tile.  One workgroup per tile walks
// the candidates in the reference's order (lattice point, then cell) 256 at
// a time; a workgroup prefix count of the kept ones gives every kept point
// its slot, so the appended order is the reference's serial order.
#include "common.h"
#include "density.hpp"

namespace zpic {
namespace {

constexpr int WAVE = 64;
constexpr int INJ_BLOCK = 256;

struct InjArgs {
    zpic_inject_profile d;
    int range[4];
    int ntx, nx, ny, ppcx, ppcy;
    int stride;  // ramp_stride(ppcx * ppcy)
};

__device__ inline unsigned long long lanes_below() {
    const unsigned lane = threadIdx.x % WAVE;
    return lane ? ( ~0ull >> ( WAVE - lane ) ) : 0ull;
}

// Candidate k of a tile's range r: lattice point k / vol (i0 fastest), cell
// k % vol (row-major in the range); true if the profile keeps it.
__device__ inline bool candidate( const InjArgs & a, const TileRange & r, int tx, int ty, int k, int2 & cell,
                                  float2 & pos ) {
    const int p = k / r.vol, idx = k - p * r.vol;
    const int i1 = p / a.ppcx, i0 = p - i1 * a.ppcx;
    cell = make_int2( idx % r.row + r.i0, idx / r.row + r.j0 );
    pos = make_float2( lattice_pos( i0, a.ppcx ), lattice_pos( i1, a.ppcy ) );
    return profile_keeps( a.d, tx * a.nx + cell.x, ty * a.ny + cell.y, pos.x, pos.y, p, a.ppcx * a.ppcy, a.stride );
}

// Kept candidates of this workgroup's chunk: returns the number kept by the
// whole chunk and sets `before` to the number kept by lower threads.
__device__ inline int block_count( bool keep, int & before, int * s_wave ) {
    const int w = threadIdx.x / WAVE;
    const unsigned long long m = __ballot( keep );
    if ( threadIdx.x % WAVE == 0 ) s_wave[w] = __popcll( m );
    __syncthreads();
    int tot = 0, pre = 0;
    for ( int v = 0; v < INJ_BLOCK / WAVE; v++ ) {
        if ( v < w ) pre += s_wave[v];
        tot += s_wave[v];
    }
    before = pre + __popcll( m & lanes_below() );
    __syncthreads();
    return tot;
}

__global__ void __launch_bounds__( INJ_BLOCK ) k_np_inject( InjArgs a, int * np ) {
    __shared__ int s_wave[INJ_BLOCK / WAVE];
    const int t = blockIdx.x;
    const int tx = t % a.ntx, ty = t / a.ntx;
    TileRange r;
    int total = 0;
    if ( tile_range( a.range, tx, ty, a.nx, a.ny, r ) ) {
        const int n = a.ppcx * a.ppcy * r.vol;
        for ( int k0 = 0; k0 < n; k0 += INJ_BLOCK ) {
            const int k = k0 + threadIdx.x;
            int2 cell;
            float2 pos;
            const bool keep = k < n && candidate( a, r, tx, ty, k, cell, pos );
            int before;
            total += block_count( keep, before, s_wave );
        }
    }
    if ( threadIdx.x == 0 ) gbl( np )[t] = total;
}

// Slots at or beyond `cap` (the buffer size) are never written.
__global__ void __launch_bounds__( INJ_BLOCK ) k_inject( InjArgs a, int * np, const int * offset, int2 * ix,
                                                         float2 * x, float3 * u, unsigned cap ) {
    __shared__ int s_wave[INJ_BLOCK / WAVE];
    const int t = blockIdx.x;
    const int tx = t % a.ntx, ty = t / a.ntx;
    TileRange r;
    if ( !tile_range( a.range, tx, ty, a.nx, a.ny, r ) ) return;
    const int np0 = gbl( np )[t];
    const int base = gbl( offset )[t] + np0;
    const gptr<int2> gix = gbl( ix );
    const gptr<float2> gx = gbl( x );
    const gptr<float3> gu = gbl( u );
    const int n = a.ppcx * a.ppcy * r.vol;
    int done = 0;
    for ( int k0 = 0; k0 < n; k0 += INJ_BLOCK ) {
        const int k = k0 + threadIdx.x;
        int2 cell = make_int2( 0, 0 );
        float2 pos = make_float2( 0.f, 0.f );
        const bool keep = k < n && candidate( a, r, tx, ty, k, cell, pos );
        int before;
        const int tot = block_count( keep, before, s_wave );
        const int dst = base + done + before;
        if ( keep && dst >= 0 && (unsigned) dst < cap ) {
            gix[dst] = cell;
            gx[dst] = pos;
            gu[dst] = make_float3( 0.f, 0.f, 0.f );
        }
        done += tot;
    }
    if ( threadIdx.x == 0 ) gbl( np )[t] = np0 + done;
}

int check_inject( const zpic_inject_profile * d, int ntx, int nty, int nx, int ny, int ppcx, int ppcy,
                  const int range[4] ) {
    if ( !d || !range ) return set_error( ZPIC_EINVAL, "null argument" );
    if ( ntx < 1 || nty < 1 || nx < 1 || ny < 1 ) return set_error( ZPIC_EINVAL, "invalid tiling" );
    if ( ppcx < 1 || ppcy < 1 ) return set_error( ZPIC_EINVAL, "Invalid ppc value, must be >= 1 in all directions" );
    if ( d->kind < ZPIC_DENSITY_NONE || d->kind > ZPIC_DENSITY_RAMP )
        return set_error( ZPIC_EINVAL, "unknown density profile" );
    if ( d->kind == ZPIC_DENSITY_RAMP && !( d->end > d->start ) )
        return set_error( ZPIC_EINVAL, "ramp end must lie beyond its start" );
    if ( d->dir != 0 && d->dir != 1 ) return set_error( ZPIC_EINVAL, "profile direction must be 0 (x) or 1 (y)" );
    if ( range[0] > range[1] || range[2] > range[3] ) return set_error( ZPIC_EINVAL, "empty injection range" );
    if ( (long long) ppcx * ppcy * nx * ny > ( 1ll << 30 ) ) return set_error( ZPIC_ERANGE, "too many lattice points per tile" );
    return ZPIC_OK;
}

InjArgs inj_args( const zpic_inject_profile * d, int ntx, int nx, int ny, int ppcx, int ppcy, const int range[4] ) {
    InjArgs a;
    a.d = *d;
    for ( int i = 0; i < 4; i++ ) a.range[i] = range[i];
    a.ntx = ntx; a.nx = nx; a.ny = ny; a.ppcx = ppcx; a.ppcy = ppcy;
    a.stride = ramp_stride( ppcx * ppcy );
    return a;
}

}  // namespace
}  // namespace zpic

using namespace zpic;

extern "C" {

int zpic_np_inject( const zpic_inject_profile * d, int ntx, int nty, int nx, int ny, int ppcx, int ppcy,
                    const int range[4], int * np, void * stream ) {
    if ( int rc = check_inject( d, ntx, nty, nx, ny, ppcx, ppcy, range ) ) return rc;
    if ( !np ) return set_error( ZPIC_EINVAL, "null np buffer" );
    const InjArgs a = inj_args( d, ntx, nx, ny, ppcx, ppcy, range );
    hipLaunchKernelGGL( k_np_inject, dim3( ntx * nty ), dim3( INJ_BLOCK ), 0, as_stream( stream ), a, np );
    ZPIC_LAUNCH_CHECK( "k_np_inject" );
    return ZPIC_OK;
}

int zpic_inject( const zpic_inject_profile * d, const zpic_particles * p, int ntx, int nty, int nx, int ny, int ppcx,
                 int ppcy, const int range[4], void * stream ) {
    if ( int rc = check_inject( d, ntx, nty, nx, ny, ppcx, ppcy, range ) ) return rc;
    if ( !p || !p->np || !p->offset || !p->ix || !p->x || !p->u ) return set_error( ZPIC_EINVAL, "null particle buffer" );
    if ( d->kind == ZPIC_DENSITY_NONE ) return ZPIC_OK;
    const InjArgs a = inj_args( d, ntx, nx, ny, ppcx, ppcy, range );
    hipLaunchKernelGGL( k_inject, dim3( ntx * nty ), dim3( INJ_BLOCK ), 0, as_stream( stream ), a, p->np, p->offset,
                        (int2 *) p->ix, (float2 *) p->x, (float3 *) p->u, (unsigned) p->max_part );
    ZPIC_LAUNCH_CHECK( "k_inject" );
    return ZPIC_OK;
}

}  // extern "C"
