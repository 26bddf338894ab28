// Grid stages of the em2d step on gfx950: guard-cell exchange, digital
// filter, Yee FDTD update, field energy, gather.  Tile-major float3 grids
// with guard cells, one workgroup per tile; stencils read the tile (guards
// included) from LDS.  All arithmetic keeps the reference's operation order
// (built with -ffp-contract=off), so every result is bit-identical to the
// reference's.
#include "common.h"
#include "grid_ops.h"

#include <algorithm>
#include <map>
#include <mutex>
#include <memory>
#include <vector>

namespace zpic {

namespace {
thread_local std::string g_error;
}

int set_error( int code, const std::string & msg ) {
    g_error = msg;
    return code;
}

namespace {

using namespace grid_ops;

template <class A>
__device__ inline int select_instance( A & a, const A * tab, const int2 * map ) {
    if ( map ) {
        const int2 r = gbl( map )[blockIdx.x];
        a = gbl( tab )[r.x];
        return r.y;
    }
    return blockIdx.x;
}

template <int NC>
__global__ void __launch_bounds__( GC_BLOCK ) k_copy_gc_x( GcArgs a, const GcArgs * tab, const int2 * map ) {
    const int t = select_instance( a, tab, map );
    copy_gc_x_tile<NC>( a, t );
}

// copy_to_gc_x of two grids of one geometry (blockIdx.y selects the grid).
template <int NC>
__global__ void __launch_bounds__( GC_BLOCK ) k_copy_gc_x2( GcArgs a0, GcArgs a1 ) {
    copy_gc_x_tile<NC>( blockIdx.y ? a1 : a0, blockIdx.x );
}

template <int NC>
__global__ void __launch_bounds__( GC_BLOCK ) k_copy_gc_xy( GcArgs a0, GcArgs a1 ) {
    copy_gc_xy_tile<NC>( blockIdx.y ? a1 : a0, blockIdx.x );
}

template <int NC>
__global__ void __launch_bounds__( GC_BLOCK ) k_add_gc_x( GcArgs a, const GcArgs * tab, const int2 * map ) {
    const int t = select_instance( a, tab, map );
    add_gc_x_tile<NC>( a, t );
}

template <int NC>
__global__ void __launch_bounds__( GC_BLOCK ) k_add_gc_y( GcArgs a, const GcArgs * tab, const int2 * map ) {
    const int t = select_instance( a, tab, map );
    add_gc_y_tile<NC>( a, t );
}

template <int NC>
__global__ void __launch_bounds__( GC_BLOCK ) k_copy_gc_y( GcArgs a, const GcArgs * tab, const int2 * map ) {
    const int t = select_instance( a, tab, map );
    copy_gc_y_tile<NC>( a, t );
}

__global__ void __launch_bounds__( TILE_BLOCK ) k_kernel3( K3Args args, const K3Args * tab, const int2 * map ) {
    extern __shared__ float A[];
    const int t = select_instance( args, tab, map );
    kernel3_tile( args, t, A );
}

// 128 threads and CUR_TILES tiles per workgroup: 104-108 us on the 2048^2
// grid, against 113 us with one tile per workgroup and 115 / 139 us with
// 256 / 512 threads (the fixed costs of a workgroup amortised; the same
// for the Yee update and the guard copy gained nothing or lost)
constexpr int CUR_BLOCK = 128;
constexpr int CUR_TILES = 4;
__global__ void __launch_bounds__( CUR_BLOCK ) k_current( CurArgs a ) {
    extern __shared__ float sm[];
    for ( int t = blockIdx.x; t < a.g.ntx * a.g.nty; t += gridDim.x ) {
        if ( t != (int) blockIdx.x ) __syncthreads();  // LDS of the previous tile read out
        current_tile( a, t, sm );
    }
}

__global__ void __launch_bounds__( TILE_BLOCK ) k_yee( YeeArgs a, const YeeArgs * tab, const int2 * map ) {
    extern __shared__ float sm[];
    const int t = select_instance( a, tab, map );
    yee_tile( a, t, sm );
}

// EMF::get_energy per tile (emf.cpp:680-705): one thread per (tile,
// component), cells in the reference's order; the inner bound is nx.y as in
// the reference loop.
__global__ void k_energy_tiles( const float * __restrict__ E, const float * __restrict__ B, G g,
                                double * __restrict__ out ) {
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    const int nt = g.ntx * g.nty;
    if ( gid >= 6 * nt ) return;
    const int t = gid / 6, comp = gid % 6;
    const float * F = ( comp < 3 ) ? E : B;
    const int c = comp % 3;
    const float * base = F + ( (size_t) t * g.vol + g.off ) * 3;
    double s = 0;
    for ( int iy = 0; iy < g.ny; iy++ )
        for ( int ix = 0; ix < g.ny; ix++ ) {
            const float v = base[( iy * g.ext_x + ix ) * 3 + c];
            s += v * v;
        }
    out[6 * t + comp] = s;
}

// Sums tile values in tile order, out[c] = 0 + ((v0 + v1) + ...): the block
// stages 1024 values at a time in LDS and thread c adds them in order.
constexpr int SUM_BLOCK = 1024;
__global__ void __launch_bounds__( SUM_BLOCK ) k_sum_tiles( const double * __restrict__ v, int n, int stride,
                                                            double * __restrict__ out ) {
    __shared__ double buf[SUM_BLOCK];
    for ( int c = 0; c < stride; c++ ) {
        double acc = 0;
        for ( int base = 0; base < n; base += SUM_BLOCK ) {
            const int t = base + threadIdx.x;
            buf[threadIdx.x] = t < n ? v[(size_t) t * stride + c] : 0.0;
            __syncthreads();
            if ( threadIdx.x == 0 ) {
                const int m = min( SUM_BLOCK, n - base );
                for ( int k = 0; k < m; k++ ) acc += buf[k];
            }
            __syncthreads();
        }
        if ( threadIdx.x == 0 ) out[c] = 0.0 + acc;
    }
}

// Unordered parallel sum for the fast mode: per-thread runs (four partial
// sums, so that four loads are in flight), a butterfly within each wave,
// then the wave sums in wave order; one workgroup barrier.
__global__ void __launch_bounds__( SUM_BLOCK ) k_sum_tiles_fast( const double * __restrict__ v, int n,
                                                                 double * __restrict__ out ) {
    constexpr int W64 = 64;
    __shared__ double wsum[SUM_BLOCK / W64];
    double a0 = 0, a1 = 0, a2 = 0, a3 = 0;
    int t = threadIdx.x;
    for ( ; t + 3 * SUM_BLOCK < n; t += 4 * SUM_BLOCK ) {
        a0 += v[t];
        a1 += v[t + SUM_BLOCK];
        a2 += v[t + 2 * SUM_BLOCK];
        a3 += v[t + 3 * SUM_BLOCK];
    }
    for ( ; t < n; t += SUM_BLOCK ) a0 += v[t];
    double acc = ( a0 + a1 ) + ( a2 + a3 );
    for ( int o = W64 / 2; o > 0; o >>= 1 ) acc += __shfl_xor( acc, o );
    if ( threadIdx.x % W64 == 0 ) wsum[threadIdx.x / W64] = acc;
    __syncthreads();
    if ( threadIdx.x == 0 ) {
        double s = 0;
        for ( int w = 0; w < SUM_BLOCK / W64; w++ ) s += wsum[w];
        out[0] = s;
    }
}

__global__ void k_gather( const float * __restrict__ buf, G g, int ncomp, int fc, float * __restrict__ out ) {
    const int dimx = g.nx * g.ntx;
    const size_t n = (size_t) dimx * g.ny * g.nty;
    for ( size_t k = blockIdx.x * (size_t) blockDim.x + threadIdx.x; k < n; k += (size_t) gridDim.x * blockDim.x ) {
        const int gx = k % dimx, gy = k / dimx;
        const int tx = gx / g.nx, ty = gy / g.ny, ix = gx % g.nx, iy = gy % g.ny;
        const size_t t = (size_t) ty * g.ntx + tx;
        out[k] = buf[( t * g.vol + g.off + iy * g.ext_x + ix ) * ncomp + fc];
    }
}

__global__ void k_fill( float * __restrict__ buf, size_t n, float v ) {
    for ( size_t k = blockIdx.x * (size_t) blockDim.x + threadIdx.x; k < n; k += (size_t) gridDim.x * blockDim.x )
        buf[k] = v;
}

__global__ void k_grid_add( float * __restrict__ dst, const float * __restrict__ src, size_t n ) {
    for ( size_t k = blockIdx.x * (size_t) blockDim.x + threadIdx.x; k < n; k += (size_t) gridDim.x * blockDim.x )
        dst[k] += src[k];
}

// x_shift_left (grid.h:455-485) of one tile: rows of ext_x elements,
// guard rows and columns included, staged in LDS and written back shifted.
// blockIdx.y selects the grid (buf0 or buf1).
template <int NC>
__global__ void __launch_bounds__( TILE_BLOCK ) k_x_shift_left( float * buf0, float * buf1, G g, int shift ) {
    extern __shared__ float A[];
    const gptr<float> tb = gbl( blockIdx.y ? buf1 : buf0 ) + (size_t) blockIdx.x * g.vol * NC;
    const int n = g.ext_x * g.ext_y * NC;
    for ( int k = threadIdx.x; k < n; k += blockDim.x ) A[k] = tb[k];
    __syncthreads();
    const int keep = ( g.ext_x - shift ) * NC;
    for ( int k = threadIdx.x; k < n; k += blockDim.x ) {
        const int r = k % ( g.ext_x * NC );
        tb[k] = ( r < keep ) ? A[k + shift * NC] : 0.0f;
    }
}

// Several kernel3_x passes, each followed by copy_to_gc_x (grid.h:495-539
// and 253-295), in one launch, tile-parallel and out of place: workgroup t
// stages its tile's rows (all components, guard rows included) widened by
// `passes` columns of the x neighbours' interiors on each side, runs the
// passes in LDS and writes the whole tile (interior and x guards) to `dst`.
// Every loaded column is one of three tiles' data (lower neighbour L, the
// tile T, upper neighbour R), and the passes reproduce the pass-by-pass
// results exactly: the first pass reads each tile's own x guards as they
// are (the element next to a tile edge uses that tile's guard, as kernel3_x
// does); later passes read the neighbour's new edge value, which is what
// the guard copy between passes puts there; at a non-periodic edge the
// guards keep their values through all passes.  After the last pass the x
// guards hold the neighbours' final edge values, as the last copy_to_gc_x
// leaves them.  Needs gc_lower + passes <= nx and gc_upper + passes <= nx
// (the widened rows reach no further than the neighbours).
constexpr int MAX_PASSES = 16;
constexpr int TF_BLOCK = 180;  // threads per workgroup at most (3 x window width x rows at once; 96-1024 measured)
struct TileFilterArgs {
    const float * src;
    float * dst;
    G g;
    int px;
    int passes;
    float a[MAX_PASSES], b[MAX_PASSES], c[MAX_PASSES];
};

__global__ void __launch_bounds__( TF_BLOCK ) k_filter_tiles_x( TileFilterArgs a ) {
    // thread (q, r): float q = 3 w + comp of every blockDim.y-th row
    extern __shared__ float lds[];
    const G g = a.g;
    const int t = blockIdx.x;
    const int tx = t % g.ntx, ty = t / g.ntx;
    const int n = a.passes;
    const int WL = g.gl + n;                   // window columns left of the tile's first cell
    const int W = WL + g.nx + g.gu + n;        // window width
    const int Q = 3 * W;                       // floats per window row
    const int NE = g.ext_y * Q;                // floats per window copy
    float * A = lds;
    float * B = lds + NE;
    float * bg = lds + 2 * NE;                 // per row and component: L's upper, T's lower, T's upper, R's lower guard
    const int lo = nbr( tx, ty, -1, 0, g, a.px, 0 ), hi = nbr( tx, ty, +1, 0, g, a.px, 0 );
    const gptr<const float> src = gbl( a.src );
    const int q = threadIdx.x, r0 = threadIdx.y, dr = blockDim.y;
    const int w = q / 3, comp = q - 3 * w;
    const int ix = w - WL;
    // the column's source: (tile, column) or none
    int stile = -1, scol = 0;
    if ( ix >= 0 && ix < g.nx ) { stile = t; scol = g.gl + ix; }
    else if ( ix < 0 ) {
        if ( lo >= 0 ) { stile = lo; scol = g.gl + g.nx + ix; }
        else if ( ix >= -g.gl ) { stile = t; scol = g.gl + ix; }
    } else {
        if ( hi >= 0 ) { stile = hi; scol = g.gl + ix - g.nx; }
        else if ( ix < g.nx + g.gu ) { stile = t; scol = g.gl + ix; }
    }
    const size_t sbase = ( (size_t) ( stile < 0 ? 0 : stile ) * g.vol + scol ) * 3 + comp;
    for ( int j = r0; j < g.ext_y; j += dr ) A[j * Q + q] = stile >= 0 ? src[sbase + (size_t) j * g.ext_x * 3] : 0.0f;
    if ( q < 12 ) {
        const int side = q / 3;  // 0 L upper, 1 T lower, 2 T upper, 3 R lower
        const int gt = side == 0 ? lo : side == 3 ? hi : t;
        const int gcol = ( side == 0 || side == 2 ) ? g.gl + g.nx : g.gl - 1;
        for ( int j = r0; j < g.ext_y; j += dr )
            bg[j * 12 + q] = gt >= 0 ? src[( (size_t) gt * g.vol + gcol + j * g.ext_x ) * 3 + comp] : 0.0f;
    }
    __syncthreads();
    // held columns: the window's ends and the guards at a non-periodic edge
    const bool held = ( lo < 0 && ix < 0 ) || ( hi < 0 && ix >= g.nx ) || w == 0 || w == W - 1;
    const int side = ix == -1 ? 0 : ix == 0 ? 1 : ix == g.nx - 1 ? 2 : ix == g.nx ? 3 : -1;
    for ( int p = 0; p < n; p++ ) {
        const float ca = a.a[p], cb = a.b[p], cc = a.c[p];
        for ( int j = r0; j < g.ext_y; j += dr ) {
            const int k = j * Q + q;
            if ( held ) {
                B[k] = A[k];
                continue;
            }
            float l = A[k - 3], r = A[k + 3];
            if ( p == 0 && side >= 0 ) {
                // each tile's own guard next to its edge
                const float gv = bg[j * 12 + side * 3 + comp];
                if ( side == 0 || side == 2 ) r = gv;
                else l = gv;
            }
            B[k] = l * ca + A[k] * cb + r * cc;
        }
        __syncthreads();
        float * tmp = A;
        A = B;
        B = tmp;
    }
    // the tile: interior and x guards, every row
    const gptr<float> dst = gbl( a.dst ) + (size_t) t * g.vol * 3;
    const int E = g.ext_x;
    if ( q < 3 * E ) {
        const int col = q / 3, wo = col - g.gl + WL;
        for ( int j = r0; j < g.ext_y; j += dr ) dst[( col + j * E ) * 3 + comp] = A[j * Q + wo * 3 + comp];
    }
    // tile padding as it was
    const int tid = threadIdx.y * blockDim.x + threadIdx.x, nth = blockDim.x * blockDim.y;
    for ( int k = g.ext_y * E * 3 + tid; k < g.vol * 3; k += nth ) dst[k] = src[(size_t) t * g.vol * 3 + k];
}

// Absorbing layers (no reference counterpart): E and B scaled by
// exp(-(alpha (L - i))^2) at i cells from an absorbing wall, i < L (the mask
// of Cerjan, Kosloff, Kosloff & Reshef, Geophysics 50 (1985) 705), every
// element of the tile including guard cells, with factors of both axes
// multiplied.  abs[4] = {x lower, x upper, y lower, y upper}: 1 if absorbing.
struct AbsArgs {
    float * E;
    float * B;
    G g;
    int abs[4];
    int width;
    float alpha;
};

__device__ inline float layer_factor( int i, int width, float alpha ) {
    if ( i >= width ) return 1.0f;
    const float a = alpha * (float) ( width - i );
    return expf( -a * a );
}

__global__ void k_emf_absorb( AbsArgs a ) {
    const G g = a.g;
    const int t = blockIdx.x;
    const int tx = t % g.ntx, ty = t / g.ntx;
    const int X = g.ntx * g.nx, Y = g.nty * g.ny;
    // skip tiles (with their guards) entirely outside every layer
    const int x0 = tx * g.nx - g.gl, x1 = tx * g.nx + g.nx + g.gu - 1;
    const int y0 = ty * g.ny - g.gl, y1 = ty * g.ny + g.ny + g.gu - 1;
    const bool hit = ( a.abs[0] && x0 < a.width ) || ( a.abs[1] && X - 1 - x1 < a.width ) ||
                     ( a.abs[2] && y0 < a.width ) || ( a.abs[3] && Y - 1 - y1 < a.width );
    if ( !hit ) return;
    const gptr<float> Et = gbl( a.E ) + (size_t) t * g.vol * 3;
    const gptr<float> Bt = gbl( a.B ) + (size_t) t * g.vol * 3;
    for ( int e = threadIdx.x; e < g.ext_x * g.ext_y; e += blockDim.x ) {
        const int gx = tx * g.nx + e % g.ext_x - g.gl, gy = ty * g.ny + e / g.ext_x - g.gl;
        float f = 1.0f;
        if ( a.abs[0] ) f *= layer_factor( gx, a.width, a.alpha );
        if ( a.abs[1] ) f *= layer_factor( X - 1 - gx, a.width, a.alpha );
        if ( a.abs[2] ) f *= layer_factor( gy, a.width, a.alpha );
        if ( a.abs[3] ) f *= layer_factor( Y - 1 - gy, a.width, a.alpha );
        if ( f == 1.0f ) continue;
        for ( int c = 0; c < 3; c++ ) {
            Et[3 * e + c] *= f;
            Bt[3 * e + c] *= f;
        }
    }
}

// div_corr_x (laser.cpp:20-70): one thread per row of cells, tiles right to
// left and cells right to left, the divergence accumulated in double.
__global__ void k_div_corr_x( float * E, float * B, G g, double dx_dy ) {
    const int row = blockIdx.x * blockDim.x + threadIdx.x;
    if ( row >= g.nty * g.ny ) return;
    const int ty = row / g.ny, iy = row % g.ny;
    const gptr<float> Eg = gbl( E ), Bg = gbl( B );
    const int ys = g.ext_x;
    double dE = 0.0, dB = 0.0;
    for ( int tx = g.ntx - 1; tx >= 0; tx-- ) {
        const size_t base = ( (size_t) ( ty * g.ntx + tx ) * g.vol + g.off ) * 3;
        const gptr<float> te = Eg + base, tbb = Bg + base;
        for ( int ix = g.nx - 1; ix >= 0; ix-- ) {
            const float de = te[3 * ( ix + 1 + iy * ys ) + 1] - te[3 * ( ix + 1 + ( iy - 1 ) * ys ) + 1];
            dE += dx_dy * de;
            te[3 * ( ix + iy * ys )] = (float) dE;
            const float db = tbb[3 * ( ix + ( iy + 1 ) * ys ) + 1] - tbb[3 * ( ix + iy * ys ) + 1];
            dB += dx_dy * db;
            tbb[3 * ( ix + iy * ys )] = (float) dB;
        }
    }
}

// ------------------------------------------------------ physical boundaries
// Field values of a tile through its (x = gc lower) or (y = gc lower) origin,
// as the reference's emf_bcx / emf_bcy / current_bcx / current_bcy address
// them (emf.cpp:207-471, current.cpp:53-176).  bc[4] = {x lower, x upper,
// y lower, y upper} with the reference's enum values (emf: 2 pec, 3 pmc;
// current: 2 reflecting).  One workgroup per boundary tile; the x pass runs
// before the y pass (the y pass reads the guard columns the x pass wrote).
// A single tile along an axis takes only the lower branch, and takes it
// twice, as the reference's `for tx : {0, ntiles.x - 1}` with
// `tile_idx.x == 0` does (twice matters for the current's reflecting wall,
// whose update is not idempotent); every row (x pass) or column (y pass)
// only touches its own cells, so the thread of a row repeats it in order.
struct BcArgs {
    float * E;  // E, or J for the current
    float * B;  // null for the current
    G g;
    int lo, hi;  // bc of the lower / upper side of this axis
    int reps;    // 2 for a single tile along the axis, else 1
};

__device__ inline float3 ld3( gptr<float> p, int k ) { return make_float3( p[3 * k], p[3 * k + 1], p[3 * k + 2] ); }
__device__ inline void st3( gptr<float> p, int k, float3 v ) {
    p[3 * k] = v.x;
    p[3 * k + 1] = v.y;
    p[3 * k + 2] = v.z;
}

// x pass: blockIdx.y = tile row, blockIdx.x = 0 lower tile, 1 upper tile
__global__ void k_emf_bcx( BcArgs a ) {
    const G g = a.g;
    const int tx = blockIdx.x == 0 ? 0 : g.ntx - 1;
    const int t = blockIdx.y * g.ntx + tx;
    const int ys = g.ext_x, nx = g.nx;
    const gptr<float> E = gbl( a.E ) + ( (size_t) t * g.vol + g.gl ) * 3;
    const gptr<float> B = gbl( a.B ) + ( (size_t) t * g.vol + g.gl ) * 3;
    for ( int idx = threadIdx.x; idx < g.ext_y; idx += blockDim.x )
        for ( int r = 0; r < a.reps; r++ ) {
        const int iy = idx * ys;
        if ( tx == 0 ) {
            if ( a.lo == 3 ) {  // pmc
                float3 e = ld3( E, -1 + iy ), b = ld3( B, -1 + iy );
                e.x = -E[3 * ( 0 + iy )];
                e.y = E[3 * ( 1 + iy ) + 1];
                e.z = E[3 * ( 1 + iy ) + 2];
                b.x = B[3 * ( 1 + iy )];
                b.y = -B[3 * ( 0 + iy ) + 1];
                b.z = -B[3 * ( 0 + iy ) + 2];
                st3( E, -1 + iy, e );
                st3( B, -1 + iy, b );
            } else if ( a.lo == 2 ) {  // pec
                float3 e = ld3( E, -1 + iy ), b = ld3( B, -1 + iy );
                e.x = E[3 * ( 0 + iy )];
                e.y = -E[3 * ( 1 + iy ) + 1];
                e.z = -E[3 * ( 1 + iy ) + 2];
                b.x = -B[3 * ( 1 + iy )];
                b.y = B[3 * ( 0 + iy ) + 1];
                b.z = B[3 * ( 0 + iy ) + 2];
                st3( E, -1 + iy, e );
                st3( B, -1 + iy, b );
                E[3 * ( 0 + iy ) + 1] = 0.0f;
                E[3 * ( 0 + iy ) + 2] = 0.0f;
                B[3 * ( 0 + iy )] = 0.0f;
            }
        } else {
            if ( a.hi == 3 ) {  // pmc
                E[3 * ( nx + iy )] = -E[3 * ( nx - 1 + iy )];
                E[3 * ( nx + 1 + iy )] = -E[3 * ( nx - 2 + iy )];
                E[3 * ( nx + 1 + iy ) + 1] = E[3 * ( nx - 1 + iy ) + 1];
                E[3 * ( nx + 1 + iy ) + 2] = E[3 * ( nx - 1 + iy ) + 2];
                B[3 * ( nx + iy ) + 1] = -B[3 * ( nx - 1 + iy ) + 1];
                B[3 * ( nx + iy ) + 2] = -B[3 * ( nx - 1 + iy ) + 2];
                B[3 * ( nx + 1 + iy )] = B[3 * ( nx - 1 + iy )];
                B[3 * ( nx + 1 + iy ) + 1] = -B[3 * ( nx - 2 + iy ) + 1];
                B[3 * ( nx + 1 + iy ) + 2] = -B[3 * ( nx - 2 + iy ) + 2];
            } else if ( a.hi == 2 ) {  // pec
                E[3 * ( nx + iy )] = E[3 * ( nx - 1 + iy )];
                E[3 * ( nx + iy ) + 1] = 0.0f;
                E[3 * ( nx + iy ) + 2] = 0.0f;
                E[3 * ( nx + 1 + iy )] = E[3 * ( nx - 2 + iy )];
                E[3 * ( nx + 1 + iy ) + 1] = -E[3 * ( nx - 1 + iy ) + 1];
                E[3 * ( nx + 1 + iy ) + 2] = -E[3 * ( nx - 1 + iy ) + 2];
                B[3 * ( nx + iy )] = 0.0f;
                B[3 * ( nx + iy ) + 1] = B[3 * ( nx - 1 + iy ) + 1];
                B[3 * ( nx + iy ) + 2] = B[3 * ( nx - 1 + iy ) + 2];
                B[3 * ( nx + 1 + iy )] = -B[3 * ( nx - 1 + iy )];
                B[3 * ( nx + 1 + iy ) + 1] = B[3 * ( nx - 2 + iy ) + 1];
                B[3 * ( nx + 1 + iy ) + 2] = B[3 * ( nx - 2 + iy ) + 2];
            }
        }
    }
}

// y pass: blockIdx.x = tile column, blockIdx.y = 0 lower tile, 1 upper tile.
// The upper formulas keep the reference's rows (some are indexed with nx.x
// where nx.y is meant, emf.cpp:400-402 and 420-422; equal for square tiles).
__global__ void k_emf_bcy( BcArgs a ) {
    const G g = a.g;
    const int ty = blockIdx.y == 0 ? 0 : g.nty - 1;
    const int t = ty * g.ntx + blockIdx.x;
    const int ys = g.ext_x, ny = g.ny, nxx = g.nx;
    const gptr<float> E = gbl( a.E ) + ( (size_t) t * g.vol + (size_t) g.gl * ys ) * 3;
    const gptr<float> B = gbl( a.B ) + ( (size_t) t * g.vol + (size_t) g.gl * ys ) * 3;
    for ( int ix = threadIdx.x; ix < g.ext_x; ix += blockDim.x )
        for ( int r = 0; r < a.reps; r++ ) {
        if ( ty == 0 ) {
            if ( a.lo == 3 ) {  // pmc
                E[3 * ( ix - ys )] = E[3 * ( ix + ys )];
                E[3 * ( ix - ys ) + 1] = -E[3 * ix + 1];
                E[3 * ( ix - ys ) + 2] = E[3 * ( ix + ys ) + 2];
                B[3 * ( ix - ys )] = -B[3 * ix];
                B[3 * ( ix - ys ) + 1] = B[3 * ( ix + ys ) + 1];
                B[3 * ( ix - ys ) + 2] = -B[3 * ix + 2];
            } else if ( a.lo == 2 ) {  // pec
                E[3 * ( ix - ys )] = -E[3 * ( ix + ys )];
                E[3 * ( ix - ys ) + 1] = E[3 * ix + 1];
                E[3 * ( ix - ys ) + 2] = -E[3 * ( ix + ys ) + 2];
                E[3 * ix] = 0.0f;
                E[3 * ix + 2] = 0.0f;
                B[3 * ( ix - ys )] = B[3 * ix];
                B[3 * ( ix - ys ) + 1] = -B[3 * ( ix + ys ) + 1];
                B[3 * ( ix - ys ) + 2] = B[3 * ix + 2];
                B[3 * ix + 1] = 0.0f;
            }
        } else {
            if ( a.hi == 3 ) {  // pmc
                E[3 * ( ix + ny * ys ) + 1] = -E[3 * ( ix + ( ny - 1 ) * ys ) + 1];
                E[3 * ( ix + ( ny + 1 ) * ys )] = E[3 * ( ix + ( ny - 1 ) * ys )];
                E[3 * ( ix + ( ny + 1 ) * ys ) + 1] = -E[3 * ( ix + ( ny - 2 ) * ys ) + 1];
                E[3 * ( ix + ( ny + 1 ) * ys ) + 2] = E[3 * ( ix + ( ny - 1 ) * ys ) + 2];
                B[3 * ( ix + ny * ys )] = -B[3 * ( ix + ( ny - 1 ) * ys )];
                B[3 * ( ix + ny * ys ) + 2] = -B[3 * ( ix + ( ny - 1 ) * ys ) + 2];
                B[3 * ( ix + ( ny + 1 ) * ys )] = -B[3 * ( ix + ( nxx - 2 ) * ys )];
                B[3 * ( ix + ( ny + 1 ) * ys ) + 1] = B[3 * ( ix + ( nxx - 1 ) * ys ) + 1];
                B[3 * ( ix + ( ny + 1 ) * ys ) + 2] = -B[3 * ( ix + ( nxx - 2 ) * ys ) + 2];
            } else if ( a.hi == 2 ) {  // pec
                E[3 * ( ix + ny * ys )] = 0.0f;
                E[3 * ( ix + ny * ys ) + 1] = E[3 * ( ix + ( ny - 1 ) * ys ) + 1];
                E[3 * ( ix + ny * ys ) + 2] = 0.0f;
                E[3 * ( ix + ( ny + 1 ) * ys )] = -E[3 * ( ix + ( nxx - 1 ) * ys )];
                E[3 * ( ix + ( ny + 1 ) * ys ) + 1] = E[3 * ( ix + ( nxx - 2 ) * ys ) + 1];
                E[3 * ( ix + ( ny + 1 ) * ys ) + 2] = -E[3 * ( ix + ( nxx - 1 ) * ys ) + 2];
                B[3 * ( ix + ny * ys )] = B[3 * ( ix + ( ny - 1 ) * ys )];
                B[3 * ( ix + ny * ys ) + 1] = 0.0f;
                B[3 * ( ix + ny * ys ) + 2] = B[3 * ( ix + ( ny - 1 ) * ys ) + 2];
                B[3 * ( ix + ( ny + 1 ) * ys )] = B[3 * ( ix + ( ny - 2 ) * ys )];
                B[3 * ( ix + ( ny + 1 ) * ys ) + 1] = -B[3 * ( ix + ( ny - 1 ) * ys ) + 1];
                B[3 * ( ix + ( ny + 1 ) * ys ) + 2] = B[3 * ( ix + ( ny - 2 ) * ys ) + 2];
            }
        }
    }
}

// current reflecting boundaries (current.cpp:53-124 x, 127-176 y)
__global__ void k_current_bcx( BcArgs a ) {
    const G g = a.g;
    const int tx = blockIdx.x == 0 ? 0 : g.ntx - 1;
    const int t = blockIdx.y * g.ntx + tx;
    const int ys = g.ext_x, nx = g.nx;
    const gptr<float> J = gbl( a.E ) + ( (size_t) t * g.vol + g.gl ) * 3;
    for ( int idx = threadIdx.x; idx < g.ext_y; idx += blockDim.x )
        for ( int r = 0; r < a.reps; r++ ) {
        const int iy = idx * ys;
        if ( tx == 0 ) {
            if ( a.lo == 2 ) {
                const float jx0 = -J[3 * ( -1 + iy )] + J[3 * ( 0 + iy )];
                const float jy1 = J[3 * ( -1 + iy ) + 1] + J[3 * ( 1 + iy ) + 1];
                const float jz1 = J[3 * ( -1 + iy ) + 2] + J[3 * ( 1 + iy ) + 2];
                J[3 * ( -1 + iy )] = J[3 * ( 0 + iy )] = jx0;
                J[3 * ( -1 + iy ) + 1] = J[3 * ( 1 + iy ) + 1] = jy1;
                J[3 * ( -1 + iy ) + 2] = J[3 * ( 1 + iy ) + 2] = jz1;
            }
        } else if ( a.hi == 2 ) {
            const float jx0 = J[3 * ( nx - 1 + iy )] - J[3 * ( nx + iy )];
            const float jy1 = J[3 * ( nx - 1 + iy ) + 1] + J[3 * ( nx + 1 + iy ) + 1];
            const float jz1 = J[3 * ( nx - 1 + iy ) + 2] + J[3 * ( nx + 1 + iy ) + 2];
            J[3 * ( nx - 1 + iy )] = J[3 * ( nx + iy )] = jx0;
            J[3 * ( nx - 1 + iy ) + 1] = J[3 * ( nx + 1 + iy ) + 1] = jy1;
            J[3 * ( nx - 1 + iy ) + 2] = J[3 * ( nx + 1 + iy ) + 2] = jz1;
        }
    }
}

__global__ void k_current_bcy( BcArgs a ) {
    const G g = a.g;
    const int ty = blockIdx.y == 0 ? 0 : g.nty - 1;
    const int t = ty * g.ntx + blockIdx.x;
    const int ys = g.ext_x, ny = g.ny;
    const gptr<float> J = gbl( a.E ) + ( (size_t) t * g.vol + (size_t) g.gl * ys ) * 3;
    for ( int ix = threadIdx.x; ix < g.ext_x; ix += blockDim.x )
        for ( int r = 0; r < a.reps; r++ ) {
        if ( ty == 0 ) {
            if ( a.lo == 2 ) {
                const float jx1 = J[3 * ( ix - ys )] + J[3 * ( ix + ys )];
                const float jy0 = -J[3 * ( ix - ys ) + 1] + J[3 * ix + 1];
                const float jz1 = J[3 * ( ix - ys ) + 2] + J[3 * ( ix + ys ) + 2];
                J[3 * ( ix - ys )] = J[3 * ( ix + ys )] = jx1;
                J[3 * ( ix - ys ) + 1] = J[3 * ix + 1] = jy0;
                J[3 * ( ix - ys ) + 2] = J[3 * ( ix + ys ) + 2] = jz1;
            }
        } else if ( a.hi == 2 ) {
            const float jx1 = J[3 * ( ix + ( ny - 1 ) * ys )] + J[3 * ( ix + ( ny + 1 ) * ys )];
            const float jy0 = J[3 * ( ix + ( ny - 1 ) * ys ) + 1] - J[3 * ( ix + ny * ys ) + 1];
            const float jz1 = J[3 * ( ix + ( ny - 1 ) * ys ) + 2] + J[3 * ( ix + ( ny + 1 ) * ys ) + 2];
            J[3 * ( ix + ( ny - 1 ) * ys )] = J[3 * ( ix + ( ny + 1 ) * ys )] = jx1;
            J[3 * ( ix + ( ny - 1 ) * ys ) + 1] = J[3 * ( ix + ny * ys ) + 1] = jy0;
            J[3 * ( ix + ( ny - 1 ) * ys ) + 2] = J[3 * ( ix + ( ny + 1 ) * ys ) + 2] = jz1;
        }
    }
}

int check( const zpic_geom * g ) {
    if ( !g || g->ntiles_x < 1 || g->ntiles_y < 1 || g->nx < 1 || g->ny < 1 || g->gc_lower < 0 || g->gc_upper < 0 )
        return set_error( ZPIC_EINVAL, "invalid grid geometry" );
    if ( g->gc_upper > g->nx || g->gc_lower > g->nx || g->gc_upper > g->ny || g->gc_lower > g->ny )
        return set_error( ZPIC_EINVAL, "guard cells wider than the tile" );
    return ZPIC_OK;
}

template <int NC>
int gc_pass( int which, float * buf, const zpic_geom * gg, int px, int py, hipStream_t st ) {
    const G g = make( gg );
    const int nt = g.ntx * g.nty;
    const GcArgs a{ buf, g, px, py };
    const GcArgs * tab = nullptr;
    const int2 * map = nullptr;
    switch ( which ) {
    case 0: hipLaunchKernelGGL( k_add_gc_x<NC>, dim3( nt ), dim3( GC_BLOCK ), 0, st, a, tab, map ); break;
    case 1: hipLaunchKernelGGL( k_add_gc_y<NC>, dim3( nt ), dim3( GC_BLOCK ), 0, st, a, tab, map ); break;
    case 2: hipLaunchKernelGGL( k_copy_gc_x<NC>, dim3( nt ), dim3( GC_BLOCK ), 0, st, a, tab, map ); break;
    case 3: hipLaunchKernelGGL( k_copy_gc_y<NC>, dim3( nt ), dim3( GC_BLOCK ), 0, st, a, tab, map ); break;
    }
    ZPIC_LAUNCH_CHECK( "guard cell kernel" );
    return ZPIC_OK;
}

int gc( int which, float * buf, const zpic_geom * g, int ncomp, int px, int py, void * stream ) {
    if ( int rc = check( g ) ) return rc;
    if ( !buf ) return set_error( ZPIC_EINVAL, "null buffer" );
    hipStream_t st = as_stream( stream );
    if ( ncomp == 3 ) return gc_pass<3>( which, buf, g, px ? 1 : 0, py ? 1 : 0, st );
    if ( ncomp == 1 ) return gc_pass<1>( which, buf, g, px ? 1 : 0, py ? 1 : 0, st );
    return set_error( ZPIC_EINVAL, "ncomp must be 1 or 3" );
}

// copy_to_gc of one grid (buf1 null) or of two grids of one geometry, one launch.
int copy_xy( float * buf0, float * buf1, const zpic_geom * gg, int ncomp, int px, int py, void * stream ) {
    if ( int rc = check( gg ) ) return rc;
    if ( !buf0 ) return set_error( ZPIC_EINVAL, "null buffer" );
    const G g = make( gg );
    const GcArgs a0{ buf0, g, px ? 1 : 0, py ? 1 : 0 };
    const GcArgs a1{ buf1 ? buf1 : buf0, g, px ? 1 : 0, py ? 1 : 0 };
    const dim3 grid( g.ntx * g.nty, buf1 ? 2 : 1 );
    hipStream_t st = as_stream( stream );
    if ( ncomp == 3 ) hipLaunchKernelGGL( k_copy_gc_xy<3>, grid, dim3( GC_BLOCK ), 0, st, a0, a1 );
    else if ( ncomp == 1 ) hipLaunchKernelGGL( k_copy_gc_xy<1>, grid, dim3( GC_BLOCK ), 0, st, a0, a1 );
    else return set_error( ZPIC_EINVAL, "ncomp must be 1 or 3" );
    ZPIC_LAUNCH_CHECK( "k_copy_gc_xy" );
    return ZPIC_OK;
}

// ------------------------------------------------------------------ batches
// Fill of whole buffers, one row of blocks (blockIdx.y) per entry.
struct FillArgs {
    float * buf;
    size_t n;
};

__global__ void k_fill_batch( const FillArgs * tab ) {
    const FillArgs a = gbl( tab )[blockIdx.y];
    const gptr<float> buf = gbl( a.buf );
    for ( size_t k = blockIdx.x * (size_t) blockDim.x + threadIdx.x; k < a.n; k += (size_t) gridDim.x * blockDim.x )
        buf[k] = 0.0f;
}

// One batched launch: a device table of per-entry arguments and the
// block -> (entry, tile) map.
struct Group {
    int kind;  // 0 add_gc_x, 1 add_gc_y, 2 copy_gc_x, 3 copy_gc_y, 4 kernel3, 5 yee
    void * tab;
    int2 * map;
    int nblocks;
    size_t lds;
};

template <class A>
int make_group( Group & grp, int kind, const std::vector<A> & entries, const std::vector<int> & tiles, size_t lds ) {
    std::vector<int2> map;
    for ( size_t e = 0; e < entries.size(); e++ )
        for ( int t = 0; t < tiles[e]; t++ ) map.push_back( make_int2( (int) e, t ) );
    grp.kind = kind;
    grp.nblocks = (int) map.size();
    grp.lds = lds;
    grp.tab = nullptr;
    grp.map = nullptr;
    if ( map.empty() ) return ZPIC_OK;
    ZPIC_HIP_TRY( hipMalloc( &grp.tab, sizeof( A ) * entries.size() ) );
    ZPIC_HIP_TRY( hipMalloc( (void **) &grp.map, sizeof( int2 ) * map.size() ) );
    ZPIC_HIP_TRY( hipMemcpy( grp.tab, entries.data(), sizeof( A ) * entries.size(), hipMemcpyHostToDevice ) );
    ZPIC_HIP_TRY( hipMemcpy( grp.map, map.data(), sizeof( int2 ) * map.size(), hipMemcpyHostToDevice ) );
    return ZPIC_OK;
}

int run_group( const Group & grp, hipStream_t st ) {
    if ( grp.nblocks == 0 ) return ZPIC_OK;
    const dim3 nb( grp.nblocks );
    switch ( grp.kind ) {
    case 0: hipLaunchKernelGGL( k_add_gc_x<3>, nb, dim3( GC_BLOCK ), 0, st, GcArgs{}, (const GcArgs *) grp.tab, grp.map ); break;
    case 1: hipLaunchKernelGGL( k_add_gc_y<3>, nb, dim3( GC_BLOCK ), 0, st, GcArgs{}, (const GcArgs *) grp.tab, grp.map ); break;
    case 2: hipLaunchKernelGGL( k_copy_gc_x<3>, nb, dim3( GC_BLOCK ), 0, st, GcArgs{}, (const GcArgs *) grp.tab, grp.map ); break;
    case 3: hipLaunchKernelGGL( k_copy_gc_y<3>, nb, dim3( GC_BLOCK ), 0, st, GcArgs{}, (const GcArgs *) grp.tab, grp.map ); break;
    case 4: hipLaunchKernelGGL( k_kernel3, nb, dim3( TILE_BLOCK ), grp.lds, st, K3Args{}, (const K3Args *) grp.tab, grp.map ); break;
    case 5: hipLaunchKernelGGL( k_yee, nb, dim3( TILE_BLOCK ), grp.lds, st, YeeArgs{}, (const YeeArgs *) grp.tab, grp.map ); break;
    }
    ZPIC_LAUNCH_CHECK( "batched grid kernel" );
    return ZPIC_OK;
}

}  // namespace

struct BatchFieldPlan {
    int n = 0;
    FillArgs * d_fill = nullptr;
    size_t max_fill = 0;
    std::vector<Group> current;  // add, copy, filter passes
    std::vector<Group> yee;      // Yee, then E and B guard copies
    ~BatchFieldPlan() {
        (void) hipFree( d_fill );
        for ( auto * v : { &current, &yee } )
            for ( auto & g : *v ) { (void) hipFree( g.tab ); (void) hipFree( g.map ); }
    }
};

int batch_fields_create( BatchFieldPlan ** out, const zpic_batch_fields_item * items, int n ) {
    if ( !out || !items || n < 1 ) return set_error( ZPIC_EINVAL, "empty batch" );
    auto plan = std::make_unique<BatchFieldPlan>();
    plan->n = n;
    std::vector<FillArgs> fills( n );
    std::vector<GcArgs> jgc( n ), ebgc;
    std::vector<YeeArgs> yees( n );
    std::vector<int> tiles( n ), tiles2;
    int max_vol = 0;
    // filter passes of each instance: (dir, a, b, c)
    struct Pass { int dir; float a, b, c; };
    std::vector<std::vector<Pass>> passes( n );
    size_t max_passes = 0;
    for ( int i = 0; i < n; i++ ) {
        const zpic_batch_fields_item & it = items[i];
        if ( int rc = check( &it.g ) ) return rc;
        if ( it.g.gc_lower < 1 || it.g.gc_upper < 2 ) return set_error( ZPIC_EINVAL, "batch fields need guard cells {1,2}" );
        if ( it.filter_kind < 0 || it.filter_kind > 2 )
            return set_error( ZPIC_EINVAL, "batches run the none, binomial and compensated current filters only" );
        const G g = make( &it.g );
        const int px = it.periodic_x ? 1 : 0, py = it.periodic_y ? 1 : 0;
        tiles[i] = g.ntx * g.nty;
        max_vol = std::max( max_vol, g.vol );
        fills[i] = FillArgs{ it.J, (size_t) g.vol * tiles[i] * 3 };
        plan->max_fill = std::max( plan->max_fill, fills[i].n );
        jgc[i] = GcArgs{ it.J, g, px, py };
        ebgc.push_back( GcArgs{ it.E, g, px, py } );
        ebgc.push_back( GcArgs{ it.B, g, px, py } );
        tiles2.push_back( tiles[i] );
        tiles2.push_back( tiles[i] );
        yees[i] = YeeArgs{ it.E, it.B, it.J, g, it.dt_dx, it.dt_dy, it.dt };
        if ( it.filter_kind ) {
            const unsigned order = it.filter_order > 0 ? it.filter_order : 1;
            for ( unsigned k = 0; k < order; k++ ) passes[i].push_back( Pass{ it.filter_dir, 0.25f, 0.5f, 0.25f } );
            if ( it.filter_kind == 2 ) {
                // Filter::Compensated (filter.h:58-76)
                const float a = -1.0f;
                const float b = ( 4.0 + 2.0 * order ) / order;
                const float norm = 2 * a + b;
                passes[i].push_back( Pass{ it.filter_dir, a / norm, b / norm, a / norm } );
            }
        }
        max_passes = std::max( max_passes, passes[i].size() );
    }
    if ( sizeof( float ) * 6 * max_vol > 65536 ) return set_error( ZPIC_ERANGE, "tile too large for LDS staging" );
    ZPIC_HIP_TRY( hipMalloc( (void **) &plan->d_fill, sizeof( FillArgs ) * n ) );
    ZPIC_HIP_TRY( hipMemcpy( plan->d_fill, fills.data(), sizeof( FillArgs ) * n, hipMemcpyHostToDevice ) );

    for ( int kind = 0; kind < 4; kind++ ) {
        Group grp;
        if ( int rc = make_group( grp, kind, jgc, tiles, 0 ) ) return rc;
        plan->current.push_back( grp );
    }
    for ( size_t p = 0; p < max_passes; p++ ) {
        for ( int dir = 0; dir < 2; dir++ ) {
            std::vector<K3Args> k3;
            std::vector<GcArgs> cg;
            std::vector<int> nt;
            for ( int i = 0; i < n; i++ ) {
                if ( p < passes[i].size() && passes[i][p].dir == dir ) {
                    const Pass & q = passes[i][p];
                    k3.push_back( K3Args{ jgc[i].buf, jgc[i].g, dir, q.a, q.b, q.c } );
                    cg.push_back( jgc[i] );
                    nt.push_back( tiles[i] );
                }
            }
            if ( k3.empty() ) continue;
            Group g1, g2;
            if ( int rc = make_group( g1, 4, k3, nt, sizeof( float ) * 3 * max_vol ) ) return rc;
            if ( int rc = make_group( g2, dir == 0 ? 2 : 3, cg, nt, 0 ) ) return rc;
            plan->current.push_back( g1 );
            plan->current.push_back( g2 );
        }
    }
    {
        Group gy, gx, gyy;
        if ( int rc = make_group( gy, 5, yees, tiles, sizeof( float ) * 6 * max_vol ) ) return rc;
        if ( int rc = make_group( gx, 2, ebgc, tiles2, 0 ) ) return rc;
        if ( int rc = make_group( gyy, 3, ebgc, tiles2, 0 ) ) return rc;
        plan->yee.push_back( gy );
        plan->yee.push_back( gx );
        plan->yee.push_back( gyy );
    }
    *out = plan.release();
    return ZPIC_OK;
}

}  // namespace zpic

using namespace zpic;

struct zpic_batch_fields {
    BatchFieldPlan * plan;
};

extern "C" {

const char * zpic_last_error( void ) { return g_error.c_str(); }

int zpic_geom_init( zpic_geom * g, int ntx, int nty, int nx, int ny, int gl, int gu ) {
    if ( !g || ntx < 1 || nty < 1 || nx < 1 || ny < 1 || gl < 0 || gu < 0 )
        return set_error( ZPIC_EINVAL, "invalid grid geometry" );
    g->ntiles_x = ntx; g->ntiles_y = nty; g->nx = nx; g->ny = ny;
    g->gc_lower = gl; g->gc_upper = gu;
    g->ext_x = gl + nx + gu;
    g->ext_y = gl + ny + gu;
    g->tile_vol = ( g->ext_x * g->ext_y + 3 ) & ~3;
    g->offset = gl * g->ext_x + gl;
    return ZPIC_OK;
}

int zpic_add_from_gc( float * buf, const zpic_geom * g, int ncomp, int px, int py, void * stream ) {
    if ( int rc = gc( 0, buf, g, ncomp, px, py, stream ) ) return rc;
    return gc( 1, buf, g, ncomp, px, py, stream );
}

int zpic_copy_to_gc( float * buf, const zpic_geom * g, int ncomp, int px, int py, void * stream ) {
    return copy_xy( buf, nullptr, g, ncomp, px, py, stream );
}

int zpic_copy_to_gc2( float * buf0, float * buf1, const zpic_geom * g, int ncomp, int px, int py, void * stream ) {
    if ( !buf1 ) return set_error( ZPIC_EINVAL, "null buffer" );
    return copy_xy( buf0, buf1, g, ncomp, px, py, stream );
}

int zpic_add_copy_gc( float * buf, const zpic_geom * gg, int ncomp, int px, int py, void * stream ) {
    if ( int rc = zpic_add_from_gc( buf, gg, ncomp, px, py, stream ) ) return rc;
    return copy_xy( buf, nullptr, gg, ncomp, px, py, stream );
}

int zpic_kernel3( float * buf, const zpic_geom * gg, int dir, float a, float b, float c, int px, int py,
                  void * stream ) {
    if ( int rc = check( gg ) ) return rc;
    if ( gg->gc_lower < 1 || gg->gc_upper < 1 )
        return set_error( ZPIC_EINVAL, "kernel3 requires at least 1 guard cell at both boundaries" );
    const G g = make( gg );
    const size_t lds = sizeof( float ) * 3 * g.vol;
    if ( lds > 65536 ) return set_error( ZPIC_ERANGE, "tile too large for LDS staging" );
    hipStream_t st = as_stream( stream );
    const K3Args args{ buf, g, dir, a, b, c };
    hipLaunchKernelGGL( k_kernel3, dim3( g.ntx * g.nty ), dim3( TILE_BLOCK ), lds, st, args, (const K3Args *) nullptr,
                        (const int2 *) nullptr );
    ZPIC_LAUNCH_CHECK( "k_kernel3" );
    return gc( dir == 0 ? 2 : 3, buf, gg, 3, px, py, stream );
}

}  // extern "C"

namespace zpic {
namespace {
// Device scratch of at least `bytes`, one per (device, stream) so that work
// on different streams never shares it, grown on demand and kept for the
// process (staging of the tile-parallel x filter).
float * filter_scratch( size_t bytes, hipStream_t st ) {
    static std::mutex m;
    static std::map<std::pair<int, hipStream_t>, std::pair<void *, size_t>> cache;
    int dev = 0;
    if ( hipGetDevice( &dev ) != hipSuccess ) return nullptr;
    std::lock_guard<std::mutex> lk( m );
    auto & e = cache[std::make_pair( dev, st )];
    if ( e.second < bytes ) {
        if ( e.first ) (void) hipFree( e.first );
        e.first = nullptr;
        e.second = 0;
        if ( hipMalloc( &e.first, bytes ) != hipSuccess ) return nullptr;
        e.second = bytes;
    }
    return static_cast<float *>( e.first );
}
}  // namespace
}  // namespace zpic

extern "C" {

int zpic_filter_x( float * buf, const zpic_geom * gg, int passes, const float * coef, int px, int py,
                   void * stream ) {
    if ( int rc = check( gg ) ) return rc;
    if ( !buf || !coef ) return set_error( ZPIC_EINVAL, "null argument" );
    if ( passes < 1 ) return set_error( ZPIC_EINVAL, "filter passes must be >= 1" );
    if ( gg->gc_lower < 1 || gg->gc_upper < 1 )
        return set_error( ZPIC_EINVAL, "kernel3 requires at least 1 guard cell at both boundaries" );
    const G g = make( gg );
    // one pass, or tiles too narrow for the widened rows: pass by pass
    const int W = g.gl + g.nx + g.gu + 2 * passes;
    const size_t lds = sizeof( float ) * ( 2 * (size_t) g.ext_y * W * 3 + (size_t) g.ext_y * 12 );
    const size_t bytes = sizeof( float ) * 3 * (size_t) g.vol * g.ntx * g.nty;
    hipStream_t st = as_stream( stream );
    float * tmp = nullptr;
    if ( passes > 1 && passes <= MAX_PASSES && g.gl + passes <= g.nx && g.gu + passes <= g.nx && lds <= 65536 &&
         3 * W <= TF_BLOCK )
        tmp = filter_scratch( bytes, st );
    if ( !tmp ) {
        for ( int p = 0; p < passes; p++ )
            if ( int rc = zpic_kernel3( buf, gg, 0, coef[3 * p], coef[3 * p + 1], coef[3 * p + 2], px, py, stream ) )
                return rc;
        return ZPIC_OK;
    }
    TileFilterArgs a;
    a.src = buf;
    a.dst = tmp;
    a.g = g;
    a.px = px ? 1 : 0;
    a.passes = passes;
    for ( int p = 0; p < passes; p++ ) {
        a.a[p] = coef[3 * p];
        a.b[p] = coef[3 * p + 1];
        a.c[p] = coef[3 * p + 2];
    }
    const int Q = 3 * W;
    const dim3 block( Q, std::max( 1, std::min( g.ext_y, TF_BLOCK / Q ) ) );
    hipLaunchKernelGGL( k_filter_tiles_x, dim3( g.ntx * g.nty ), block, lds, st, a );
    ZPIC_LAUNCH_CHECK( "k_filter_tiles_x" );
    ZPIC_HIP_TRY( hipMemcpyAsync( buf, tmp, bytes, hipMemcpyDeviceToDevice, st ) );
    (void) py;
    return ZPIC_OK;
}

int zpic_current_advance( float * J, const zpic_geom * g, int kind, int dir, int order, int px, int py,
                          void * stream ) {
    if ( int rc = zpic_add_copy_gc( J, g, 3, px, py, stream ) ) return rc;
    if ( kind == 0 ) return ZPIC_OK;
    const unsigned n = order > 0 ? order : 1;
    std::vector<float> coef;
    for ( unsigned i = 0; i < n; i++ ) coef.insert( coef.end(), { 0.25f, 0.5f, 0.25f } );
    if ( kind == 2 ) {
        // Filter::Compensated (filter.h:58-76)
        const float a = -1.0f;
        const float b = ( 4.0 + 2.0 * n ) / n;
        const float norm = 2 * a + b;
        coef.insert( coef.end(), { a / norm, b / norm, a / norm } );
    }
    const int passes = (int) coef.size() / 3;
    if ( dir == 0 ) return zpic_filter_x( J, g, passes, coef.data(), px, py, stream );
    for ( int p = 0; p < passes; p++ )
        if ( int rc = zpic_kernel3( J, g, dir, coef[3 * p], coef[3 * p + 1], coef[3 * p + 2], px, py, stream ) )
            return rc;
    return ZPIC_OK;
}

int zpic_current_update( const float * dep, float * J, const zpic_geom * gg, int passes, const float * coef,
                         void * stream ) {
    if ( int rc = check( gg ) ) return rc;
    if ( !dep || !J || ( passes > 0 && !coef ) ) return set_error( ZPIC_EINVAL, "null argument" );
    if ( dep == J ) return set_error( ZPIC_EINVAL, "current update: deposit and current buffers must differ" );
    if ( passes < 0 || passes > CUR_MAX_PASSES ) return set_error( ZPIC_EINVAL, "current update: 0 to 16 filter passes" );
    if ( passes > 0 && ( gg->gc_lower < 1 || gg->gc_upper < 1 ) )
        return set_error( ZPIC_EINVAL, "kernel3 requires at least 1 guard cell at both boundaries" );
    const G g = make( gg );
    if ( g.nx < g.gl + g.gu + passes || g.ny < g.gl + g.gu )
        return set_error( ZPIC_EINVAL, "current update: tiles narrower than both guard bands plus the filter passes" );
    if ( 3 * ( g.vol - g.ext_x * g.ext_y ) > CUR_BLOCK ) return set_error( ZPIC_EINVAL, "current update: tile padding" );
    const size_t lds = current_lds_bytes( g, passes );
    if ( lds > 65536 ) return set_error( ZPIC_ERANGE, "tile too large for LDS staging" );
    CurArgs a{};
    a.dep = dep;
    a.J = J;
    a.g = g;
    a.passes = passes;
    for ( int p = 0; p < passes; p++ ) {
        a.a[p] = coef[3 * p];
        a.b[p] = coef[3 * p + 1];
        a.c[p] = coef[3 * p + 2];
    }
    hipStream_t st = as_stream( stream );
    hipLaunchKernelGGL( k_current, dim3( ( g.ntx * g.nty + CUR_TILES - 1 ) / CUR_TILES ), dim3( CUR_BLOCK ), lds, st, a );
    ZPIC_LAUNCH_CHECK( "k_current" );
    return ZPIC_OK;
}

int zpic_yee( float * E, float * B, const float * J, const zpic_geom * gg, float dtdx, float dtdy, float dt, int px,
              int py, void * stream ) {
    if ( int rc = check( gg ) ) return rc;
    if ( gg->gc_lower < 1 || gg->gc_upper < 2 ) return set_error( ZPIC_EINVAL, "Yee update needs guard cells {1,2}" );
    if ( !E || !B ) return set_error( ZPIC_EINVAL, "null buffer" );
    const G g = make( gg );
    const size_t lds = sizeof( float ) * 6 * g.vol;
    if ( lds > 65536 ) return set_error( ZPIC_ERANGE, "tile too large for LDS staging" );
    hipStream_t st = as_stream( stream );
    const YeeArgs args{ E, B, J, g, dtdx, dtdy, dt };
    hipLaunchKernelGGL( k_yee, dim3( g.ntx * g.nty ), dim3( TILE_BLOCK ), lds, st, args, (const YeeArgs *) nullptr,
                        (const int2 *) nullptr );
    ZPIC_LAUNCH_CHECK( "k_yee" );
    return zpic_copy_to_gc2( E, B, gg, 3, px, py, stream );
}

// EMF::process_bc / Current::process_bc (emf.cpp:431-471, current.cpp:178-218).
static int physical_bc( float * F, float * Bf, const zpic_geom * gg, const int bc[4], bool emf, void * stream ) {
    if ( int rc = check( gg ) ) return rc;
    if ( !F || ( emf && !Bf ) || !bc ) return set_error( ZPIC_EINVAL, "null buffer" );
    if ( gg->gc_lower < 1 || gg->gc_upper < 2 ) return set_error( ZPIC_EINVAL, "boundaries need guard cells {1,2}" );
    if ( gg->nx < 2 || gg->ny < 2 ) return set_error( ZPIC_EINVAL, "boundaries need tiles of at least 2x2 cells" );
    // the reference's upper-y EMF formulas read rows nx.x - 1, nx.x - 2, which
    // must stay inside the tile (rows up to ny + 2)
    if ( emf && ( bc[3] == 2 || bc[3] == 3 ) && gg->nx > gg->ny + 2 )
        return set_error( ZPIC_EINVAL, "upper y pec/pmc needs tile nx <= ny + 2 (reference row indexing)" );
    const G g = make( gg );
    hipStream_t st = as_stream( stream );
    // bc values: 0 none, 1 periodic, >1 physical
    if ( bc[0] > 1 || bc[1] > 1 ) {
        const BcArgs a{ F, Bf, g, bc[0], bc[1], g.ntx > 1 ? 1 : 2 };
        const dim3 grid( g.ntx > 1 ? 2 : 1, g.nty );
        if ( emf ) hipLaunchKernelGGL( k_emf_bcx, grid, dim3( 64 ), 0, st, a );
        else hipLaunchKernelGGL( k_current_bcx, grid, dim3( 64 ), 0, st, a );
        ZPIC_LAUNCH_CHECK( "boundary x" );
    }
    if ( bc[2] > 1 || bc[3] > 1 ) {
        const BcArgs a{ F, Bf, g, bc[2], bc[3], g.nty > 1 ? 1 : 2 };
        const dim3 grid( g.ntx, g.nty > 1 ? 2 : 1 );
        if ( emf ) hipLaunchKernelGGL( k_emf_bcy, grid, dim3( 64 ), 0, st, a );
        else hipLaunchKernelGGL( k_current_bcy, grid, dim3( 64 ), 0, st, a );
        ZPIC_LAUNCH_CHECK( "boundary y" );
    }
    return ZPIC_OK;
}

int zpic_emf_bc( float * E, float * B, const zpic_geom * g, const int bc[4], void * stream ) {
    return physical_bc( E, B, g, bc, true, stream );
}

int zpic_current_bc( float * J, const zpic_geom * g, const int bc[4], void * stream ) {
    return physical_bc( J, nullptr, g, bc, false, stream );
}

}  // extern "C"

namespace zpic {
namespace {

int field_energy_args( const float * E, const float * B, const zpic_geom * gg, const double * out6 ) {
    if ( int rc = check( gg ) ) return rc;
    if ( !E || !B || !out6 ) return set_error( ZPIC_EINVAL, "null buffer" );
    // the reference's inner bound nx.y must stay inside the tile
    if ( gg->offset + ( gg->ny - 1 ) * gg->ext_x + gg->ny - 1 >= gg->tile_vol )
        return set_error( ZPIC_ERANGE, "field energy loop would leave the tile (ny too large for nx)" );
    return ZPIC_OK;
}

// Per-tile sums into tile_sums[6 * ntiles] (every entry written), then their
// sums in tile order into out6.
int field_energy_launch( const float * E, const float * B, const zpic_geom * gg, double * tile_sums, double * out6,
                         hipStream_t st ) {
    const G g = make( gg );
    const int nt = g.ntx * g.nty;
    hipLaunchKernelGGL( k_energy_tiles, dim3( ( 6 * nt + 255 ) / 256 ), dim3( 256 ), 0, st, E, B, g, tile_sums );
    ZPIC_LAUNCH_CHECK( "k_energy_tiles" );
    hipLaunchKernelGGL( k_sum_tiles, dim3( 1 ), dim3( SUM_BLOCK ), 0, st, tile_sums, nt, 6, out6 );
    ZPIC_LAUNCH_CHECK( "k_sum_tiles" );
    return ZPIC_OK;
}

// Tile sums of the allocation-free field energy (zpic_field_energy_noalloc):
// 6 doubles for each of at most NOALLOC_MAX_TILES tiles, 3 MiB, held by every
// device the library runs on until the process ends.  Host code never uses
// the symbol's name as an address: scratch_address() looks up each device's
// own copy.
constexpr int NOALLOC_MAX_TILES = 65536;
static __device__ __align__( 256 ) double scratch[6 * NOALLOC_MAX_TILES];

// Device address of `scratch` on the current device, looked up with
// hipGetSymbolAddress on the first call made with that device current and
// kept per device; null (and rc set) on failure.
double * scratch_address( int & rc ) {
    static std::mutex m;
    static std::map<int, double *> addr;
    int dev = 0;
    hipError_t e = hipGetDevice( &dev );
    if ( e != hipSuccess ) {
        rc = hip_error( e, "hipGetDevice" );
        return nullptr;
    }
    std::lock_guard<std::mutex> lk( m );
    auto it = addr.find( dev );
    if ( it != addr.end() ) return it->second;
    double * p = nullptr;
    e = hipGetSymbolAddress( (void **) &p, HIP_SYMBOL( scratch ) );
    if ( e != hipSuccess || !p ) {
        rc = e != hipSuccess ? hip_error( e, "hipGetSymbolAddress(scratch)" )
                             : set_error( ZPIC_EHIP, "hipGetSymbolAddress(scratch) returned null" );
        return nullptr;
    }
    addr[dev] = p;
    return p;
}

}  // namespace
}  // namespace zpic

extern "C" {

int zpic_field_energy( const float * E, const float * B, const zpic_geom * gg, double * scratch_, double * out6,
                       void * stream ) {
    if ( int rc = field_energy_args( E, B, gg, out6 ) ) return rc;
    if ( !scratch_ ) return set_error( ZPIC_EINVAL, "null buffer" );
    return field_energy_launch( E, B, gg, scratch_, out6, as_stream( stream ) );
}

int zpic_field_energy_noalloc( const float * E, const float * B, const zpic_geom * gg, double * out6, void * stream ) {
    if ( int rc = field_energy_args( E, B, gg, out6 ) ) return rc;
    if ( (long long) gg->ntiles_x * gg->ntiles_y > NOALLOC_MAX_TILES )
        return set_error( ZPIC_ERANGE, "zpic_field_energy_noalloc: more than 65536 tiles" );
    int rc = ZPIC_OK;
    double * tile_sums = scratch_address( rc );
    if ( !tile_sums ) return rc;
    return field_energy_launch( E, B, gg, tile_sums, out6, as_stream( stream ) );
}

int zpic_sum_tiles( const double * v, int n, double * out, int deterministic, void * stream ) {
    if ( !v || !out || n < 0 ) return set_error( ZPIC_EINVAL, "invalid arguments" );
    if ( deterministic )
        hipLaunchKernelGGL( k_sum_tiles, dim3( 1 ), dim3( SUM_BLOCK ), 0, as_stream( stream ), v, n, 1, out );
    else
        hipLaunchKernelGGL( k_sum_tiles_fast, dim3( 1 ), dim3( SUM_BLOCK ), 0, as_stream( stream ), v, n, out );
    ZPIC_LAUNCH_CHECK( "k_sum_tiles" );
    return ZPIC_OK;
}

int zpic_gather_component( const float * buf, const zpic_geom * gg, int ncomp, int fc, float * out, void * stream ) {
    if ( int rc = check( gg ) ) return rc;
    if ( !buf || !out || fc < 0 || fc >= ncomp ) return set_error( ZPIC_EINVAL, "invalid gather arguments" );
    const G g = make( gg );
    hipLaunchKernelGGL( k_gather, dim3( 1024 ), dim3( 256 ), 0, as_stream( stream ), buf, g, ncomp, fc, out );
    ZPIC_LAUNCH_CHECK( "k_gather" );
    return ZPIC_OK;
}

int zpic_fill( float * buf, size_t n, float v, void * stream ) {
    if ( !buf && n ) return set_error( ZPIC_EINVAL, "null buffer" );
    if ( n == 0 ) return ZPIC_OK;
    size_t blocks = ( n + 255 ) / 256;
    if ( blocks > 8192 ) blocks = 8192;
    hipLaunchKernelGGL( k_fill, dim3( blocks ), dim3( 256 ), 0, as_stream( stream ), buf, n, v );
    ZPIC_LAUNCH_CHECK( "k_fill" );
    return ZPIC_OK;
}

int zpic_grid_add( float * dst, const float * src, size_t n, void * stream ) {
    if ( ( !dst || !src ) && n ) return set_error( ZPIC_EINVAL, "null buffer" );
    if ( n == 0 ) return ZPIC_OK;
    size_t blocks = ( n + 255 ) / 256;
    if ( blocks > 8192 ) blocks = 8192;
    hipLaunchKernelGGL( k_grid_add, dim3( blocks ), dim3( 256 ), 0, as_stream( stream ), dst, src, n );
    ZPIC_LAUNCH_CHECK( "k_grid_add" );
    return ZPIC_OK;
}

int zpic_x_shift_left2( float * buf0, float * buf1, const zpic_geom * gg, int shift, int px, void * stream ) {
    if ( int rc = check( gg ) ) return rc;
    if ( !buf0 ) return set_error( ZPIC_EINVAL, "null buffer" );
    if ( !( shift > 0 && shift < gg->gc_upper ) )
        return set_error( ZPIC_EINVAL, "x_shift_left(), invalid shift value, must be 0 < shift < gc.x.upper" );
    const G g = make( gg );
    hipStream_t st = as_stream( stream );
    // both grids shifted in one launch (tile by tile), then both x guard copies in one
    const size_t lds = sizeof( float ) * 3 * g.ext_x * g.ext_y;
    if ( lds > 65536 ) return set_error( ZPIC_ERANGE, "tile too large for LDS staging" );
    const dim3 grid( g.ntx * g.nty, buf1 ? 2 : 1 );
    hipLaunchKernelGGL( k_x_shift_left<3>, grid, dim3( TILE_BLOCK ), lds, st, buf0, buf1 ? buf1 : buf0, g, shift );
    ZPIC_LAUNCH_CHECK( "k_x_shift_left" );
    const GcArgs a0{ buf0, g, px ? 1 : 0, 0 }, a1{ buf1 ? buf1 : buf0, g, px ? 1 : 0, 0 };
    hipLaunchKernelGGL( k_copy_gc_x2<3>, grid, dim3( GC_BLOCK ), 0, st, a0, a1 );
    ZPIC_LAUNCH_CHECK( "k_copy_gc_x2" );
    return ZPIC_OK;
}

int zpic_x_shift_left( float * buf, const zpic_geom * gg, int ncomp, int shift, int px, int py, void * stream ) {
    if ( int rc = check( gg ) ) return rc;
    if ( !buf ) return set_error( ZPIC_EINVAL, "null buffer" );
    if ( ncomp != 1 && ncomp != 3 ) return set_error( ZPIC_EINVAL, "ncomp must be 1 or 3" );
    if ( !( shift > 0 && shift < gg->gc_upper ) )
        return set_error( ZPIC_EINVAL, "x_shift_left(), invalid shift value, must be 0 < shift < gc.x.upper" );
    const G g = make( gg );
    const size_t lds = sizeof( float ) * ncomp * g.ext_x * g.ext_y;
    if ( lds > 65536 ) return set_error( ZPIC_ERANGE, "tile too large for LDS staging" );
    hipStream_t st = as_stream( stream );
    if ( ncomp == 3 )
        hipLaunchKernelGGL( k_x_shift_left<3>, dim3( g.ntx * g.nty ), dim3( TILE_BLOCK ), lds, st, buf, buf, g, shift );
    else
        hipLaunchKernelGGL( k_x_shift_left<1>, dim3( g.ntx * g.nty ), dim3( TILE_BLOCK ), lds, st, buf, buf, g, shift );
    ZPIC_LAUNCH_CHECK( "k_x_shift_left" );
    return gc( 2, buf, gg, ncomp, px, py, stream );
}

int zpic_emf_absorb( float * E, float * B, const zpic_geom * gg, const int bc[4], int width, float alpha,
                     void * stream ) {
    if ( int rc = check( gg ) ) return rc;
    if ( !E || !B || !bc ) return set_error( ZPIC_EINVAL, "null argument" );
    if ( width < 1 || !( alpha >= 0 ) ) return set_error( ZPIC_EINVAL, "absorbing layer needs width >= 1, alpha >= 0" );
    AbsArgs a{ E, B, make( gg ), { 0, 0, 0, 0 }, width, alpha };
    bool any = false;
    for ( int i = 0; i < 4; i++ ) {
        a.abs[i] = ( bc[i] == ZPIC_EMF_ABSORBING );
        any = any || a.abs[i];
    }
    if ( !any ) return ZPIC_OK;
    hipLaunchKernelGGL( k_emf_absorb, dim3( a.g.ntx * a.g.nty ), dim3( TILE_BLOCK ), 0, as_stream( stream ), a );
    ZPIC_LAUNCH_CHECK( "k_emf_absorb" );
    return ZPIC_OK;
}

int zpic_div_corr_x( float * E, float * B, const zpic_geom * gg, float dx, float dy, void * stream ) {
    if ( int rc = check( gg ) ) return rc;
    if ( !E || !B ) return set_error( ZPIC_EINVAL, "null buffer" );
    if ( gg->gc_lower < 1 || gg->gc_upper < 1 ) return set_error( ZPIC_EINVAL, "div_corr_x needs guard cells" );
    if ( !( dx > 0 ) || !( dy > 0 ) ) return set_error( ZPIC_EINVAL, "cell size must be positive" );
    const G g = make( gg );
    const int rows = g.nty * g.ny;
    const double dx_dy = ( (double) dx ) / ( (double) dy );
    hipLaunchKernelGGL( k_div_corr_x, dim3( ( rows + 63 ) / 64 ), dim3( 64 ), 0, as_stream( stream ), E, B, g,
                        dx_dy );
    ZPIC_LAUNCH_CHECK( "k_div_corr_x" );
    return ZPIC_OK;
}

}  // extern "C"

extern "C" {

int zpic_batch_fields_create( zpic_batch_fields ** out, const zpic_batch_fields_item * items, int n ) {
    if ( !out ) return set_error( ZPIC_EINVAL, "null output" );
    BatchFieldPlan * plan = nullptr;
    if ( int rc = batch_fields_create( &plan, items, n ) ) return rc;
    *out = new zpic_batch_fields{ plan };
    return ZPIC_OK;
}

int zpic_batch_zero_current( zpic_batch_fields * b, void * stream ) {
    if ( !b || !b->plan ) return set_error( ZPIC_EINVAL, "null batch" );
    size_t blocks = ( b->plan->max_fill + 255 ) / 256;
    if ( blocks > 1024 ) blocks = 1024;
    hipLaunchKernelGGL( k_fill_batch, dim3( blocks, b->plan->n ), dim3( 256 ), 0, as_stream( stream ),
                        (const FillArgs *) b->plan->d_fill );
    ZPIC_LAUNCH_CHECK( "k_fill_batch" );
    return ZPIC_OK;
}

int zpic_batch_current_advance( zpic_batch_fields * b, void * stream ) {
    if ( !b || !b->plan ) return set_error( ZPIC_EINVAL, "null batch" );
    for ( const Group & g : b->plan->current )
        if ( int rc = run_group( g, as_stream( stream ) ) ) return rc;
    return ZPIC_OK;
}

int zpic_batch_yee( zpic_batch_fields * b, void * stream ) {
    if ( !b || !b->plan ) return set_error( ZPIC_EINVAL, "null batch" );
    for ( const Group & g : b->plan->yee )
        if ( int rc = run_group( g, as_stream( stream ) ) ) return rc;
    return ZPIC_OK;
}

void zpic_batch_fields_free( zpic_batch_fields * b ) {
    if ( !b ) return;
    delete b->plan;
    delete b;
}

}  // extern "C"
