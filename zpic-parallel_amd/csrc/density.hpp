// Particle injection by density profile, shared by the host (initial
// injection) and the device (injection at the leading edge of a moving
// window).
//
// Density::Uniform / Step / Slab / Sphere (density.cpp:103-1067) all inject
// the same regular lattice of ppc.x * ppc.y positions per cell, restricted to
// a cell range clipped to each tile; they differ only in which lattice points
// are kept.  The order is the reference's: for each lattice point (i1 < ppc.y,
// i0 < ppc.x), for each cell of the clipped range in row-major order, the
// accepted points are appended to the tile.
#pragma once

#include <hip/hip_runtime.h>

#include "zpic_amd.h"

namespace zpic {

// Clipped injection range of tile (tx, ty) in tile-local cells; false if the
// range misses the tile (inject_uniform_kernel, density.cpp:103-145).
struct TileRange {
    int i0, j0, row, vol;
};

__host__ __device__ inline bool tile_range( const int range[4], int tx, int ty, int nx, int ny, TileRange & r ) {
    int ri0 = range[0] - tx * nx, ri1 = range[1] - tx * nx;
    int rj0 = range[2] - ty * ny, rj1 = range[3] - ty * ny;
    if ( !( ri0 < nx && ri1 >= 0 && rj0 < ny && rj1 >= 0 ) ) return false;
    if ( ri0 < 0 ) ri0 = 0;
    if ( rj0 < 0 ) rj0 = 0;
    if ( ri1 >= nx ) ri1 = nx - 1;
    if ( rj1 >= ny ) rj1 = ny - 1;
    r.i0 = ri0;
    r.j0 = rj0;
    r.row = ri1 - ri0 + 1;
    r.vol = ( rj1 - rj0 + 1 ) * r.row;
    return true;
}

// In-cell lattice position of point i of n along one axis: the reference
// evaluates 1/n * (i + 0.5) - 0.5 in double and rounds once to float.
__host__ __device__ inline float lattice_pos( unsigned i, unsigned n ) {
    const double d = 1.0 / n;
    return (float) ( d * ( i + 0.5 ) - 0.5 );
}

// Rank of lattice point k among the npc points of a cell, for the ramp's
// thinning: k * stride mod npc with stride coprime to npc near 0.618 npc, so
// that points of consecutive rank lie apart in the cell.
__host__ __device__ inline int ramp_stride( int npc ) {
    int s = (int) ( 0.618034f * npc );
    if ( s < 1 ) s = 1;
    for ( ; s > 1; s-- ) {
        int a = s, b = npc;
        while ( b ) {
            const int r = a % b;
            a = b;
            b = r;
        }
        if ( a == 1 ) break;
    }
    return s;
}

// Whether the profile keeps the lattice point at in-cell position (px, py)
// of global cell (gx, gy) (cell indices counted from the window origin); k
// is the point's index i1 * ppc.x + i0 among the npc points of its cell.
//  step:   t = cell + (pos + 0.5) along dir, kept if t > start
//          (inject_step_kernel, density.cpp:315-413)
//  slab:   kept if start <= t < end (inject_slab_kernel, :575-672)
//  sphere: kept if |((cell + pos + 0.5) * dx) - center|^2 < radius^2,
//          in float (inject_sphere_kernel, :856-948)
//  ramp:   no reference counterpart: density rising linearly from 0 at
//          `start` to n0 at `end` (cells along dir), n0 beyond.  Particles
//          keep the equal charge of the other profiles; a point is kept when
//          its rank (ramp_stride) satisfies rank + 1/2 < f npc, f the ramp
//          fraction at the point, so a cell holds about f npc particles.
// t and the sphere coordinates are formed in double from the integer cell
// and the float position and rounded to float, as in the reference.
__host__ __device__ inline bool profile_keeps( const zpic_inject_profile & d, int gx, int gy, float px, float py,
                                               int k, int npc, int stride ) {
    switch ( d.kind ) {
    case ZPIC_DENSITY_UNIFORM:
        return true;
    case ZPIC_DENSITY_STEP:
    case ZPIC_DENSITY_SLAB: {
        const float t = ( d.dir == 0 ) ? (float) ( gx + ( px + 0.5 ) ) : (float) ( gy + ( py + 0.5 ) );
        if ( d.kind == ZPIC_DENSITY_STEP ) return t > d.start;
        return ( t >= d.start ) && ( t < d.end );
    }
    case ZPIC_DENSITY_RAMP: {
        const float t = ( d.dir == 0 ) ? (float) ( gx + ( px + 0.5 ) ) : (float) ( gy + ( py + 0.5 ) );
        const float f = ( t - d.start ) / ( d.end - d.start );
        const int rank = ( k * stride ) % npc;
        return ( rank + 0.5f ) < f * npc;
    }
    case ZPIC_DENSITY_SPHERE: {
        const float x = (float) ( ( gx + ( px + 0.5 ) ) * d.dx );
        const float y = (float) ( ( gy + ( py + 0.5 ) ) * d.dy );
        const float r2 = d.radius * d.radius;
        const float ax = x - d.center_x, ay = y - d.center_y;
        return ax * ax + ay * ay < r2;
    }
    default:
        return false;
    }
}

}  // namespace zpic
