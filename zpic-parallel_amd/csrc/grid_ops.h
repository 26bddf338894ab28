// Per-tile bodies of the grid stages (guard-cell exchange, digital filter
// pass, Yee update) on tile-major float grids with guard cells.  Each body
// is run by one workgroup for tile t, with any workgroup size; the
// single-stage kernels (fields.hip) and the persistent batched step
// (particles.hip) both call them.  Arithmetic keeps the reference's
// operation order (built with -ffp-contract=off).
#pragma once

#include "common.h"

namespace zpic {
namespace grid_ops {

struct G {
    int ntx, nty, nx, ny, gl, gu, ext_x, ext_y, vol, off;
};

inline G make( const zpic_geom * g ) {
    return G{ g->ntiles_x, g->ntiles_y, g->nx, g->ny, g->gc_lower, g->gc_upper,
              g->ext_x, g->ext_y, g->tile_vol, g->offset };
}

__device__ inline int nbr( int tx, int ty, int dx, int dy, const G & g, int px, int py ) {
    int x = tx + dx, y = ty + dy;
    if ( px ) { if ( x < 0 ) x += g.ntx; if ( x >= g.ntx ) x -= g.ntx; }
    if ( py ) { if ( y < 0 ) y += g.nty; if ( y >= g.nty ) y -= g.nty; }
    if ( x < 0 || x >= g.ntx || y < 0 || y >= g.nty ) return -1;
    return y * g.ntx + x;
}

// Arguments of one grid operand; batched launches pass a table of them and a
// map from block to (entry, tile), single launches pass nullptr for both.
struct GcArgs {
    float * buf;
    G g;
    int px, py;
};

struct K3Args {
    float * buf;
    G g;
    int dir;
    float a, b, c;
};

struct YeeArgs {
    float * E;
    float * B;
    const float * J;
    G g;
    float dtdx, dtdy, dt;
};

constexpr int GC_BLOCK = 128;

// add_from_gc, x pass (grid.h:367-401): lower neighbour's upper guards into
// the first interior columns, then upper neighbour's lower guards into the
// last interior columns.
template <int NC>
__device__ inline void add_gc_x_tile( const GcArgs & a, const int t ) {
    const gptr<float> buf = gbl( a.buf );
    const G g = a.g;
    const int px = a.px, py = a.py;
    const int tx = t % g.ntx, ty = t / g.ntx;
    const gptr<float> loc = buf + (size_t) t * g.vol * NC;
    const int lo = nbr( tx, ty, -1, 0, g, px, py ), hi = nbr( tx, ty, +1, 0, g, px, py );
    const int E = g.ext_x;
    if ( lo >= 0 ) {
        const gptr<const float> src = buf + (size_t) lo * g.vol * NC;
        for ( int k = threadIdx.x; k < g.ext_y * g.gu * NC; k += blockDim.x ) {
            const int c = k % NC, e = k / NC, i = e % g.gu, j = e / g.gu;
            loc[( g.gl + i + j * E ) * NC + c] += src[( g.gl + g.nx + i + j * E ) * NC + c];
        }
    }
    __syncthreads();
    if ( hi >= 0 ) {
        const gptr<const float> src = buf + (size_t) hi * g.vol * NC;
        for ( int k = threadIdx.x; k < g.ext_y * g.gl * NC; k += blockDim.x ) {
            const int c = k % NC, e = k / NC, i = e % g.gl, j = e / g.gl;
            loc[( g.nx + i + j * E ) * NC + c] += src[( i + j * E ) * NC + c];
        }
    }
}

// add_from_gc, y pass (grid.h:406-444): full rows of the y neighbours' guards.
template <int NC>
__device__ inline void add_gc_y_tile( const GcArgs & a, const int t ) {
    const gptr<float> buf = gbl( a.buf );
    const G g = a.g;
    const int px = a.px, py = a.py;
    const int tx = t % g.ntx, ty = t / g.ntx;
    const gptr<float> loc = buf + (size_t) t * g.vol * NC;
    const int lo = nbr( tx, ty, 0, -1, g, px, py ), hi = nbr( tx, ty, 0, +1, g, px, py );
    const int E = g.ext_x;
    if ( lo >= 0 ) {
        const gptr<const float> src = buf + (size_t) lo * g.vol * NC;
        for ( int k = threadIdx.x; k < g.gu * E * NC; k += blockDim.x ) {
            const int c = k % NC, e = k / NC, i = e % E, j = e / E;
            loc[( i + ( g.gl + j ) * E ) * NC + c] += src[( i + ( g.gl + g.ny + j ) * E ) * NC + c];
        }
    }
    __syncthreads();
    if ( hi >= 0 ) {
        const gptr<const float> src = buf + (size_t) hi * g.vol * NC;
        for ( int k = threadIdx.x; k < g.gl * E * NC; k += blockDim.x ) {
            const int c = k % NC, e = k / NC, i = e % E, j = e / E;
            loc[( i + ( g.ny + j ) * E ) * NC + c] += src[( i + j * E ) * NC + c];
        }
    }
}

// copy_to_gc_x (grid.h:253-297) of tile t: x guard columns from the
// neighbours' edges.
template <int NC>
__device__ inline void copy_gc_x_tile( const GcArgs & a, int t ) {
    const gptr<float> buf = gbl( a.buf );
    const G g = a.g;
    const int px = a.px, py = a.py;
    const int tx = t % g.ntx, ty = t / g.ntx;
    const gptr<float> loc = buf + (size_t) t * g.vol * NC;
    const int lo = nbr( tx, ty, -1, 0, g, px, py ), hi = nbr( tx, ty, +1, 0, g, px, py );
    const int E = g.ext_x;
    if ( lo >= 0 ) {
        const gptr<const float> src = buf + (size_t) lo * g.vol * NC;
        for ( int k = threadIdx.x; k < g.ext_y * g.gl * NC; k += blockDim.x ) {
            const int c = k % NC, e = k / NC, i = e % g.gl, j = e / g.gl;
            loc[( i + j * E ) * NC + c] = src[( g.nx + i + j * E ) * NC + c];
        }
    }
    if ( hi >= 0 ) {
        const gptr<const float> src = buf + (size_t) hi * g.vol * NC;
        for ( int k = threadIdx.x; k < g.ext_y * g.gu * NC; k += blockDim.x ) {
            const int c = k % NC, e = k / NC, i = e % g.gu, j = e / g.gu;
            loc[( g.gl + g.nx + i + j * E ) * NC + c] = src[( g.gl + i + j * E ) * NC + c];
        }
    }
}

// copy_to_gc_y (grid.h:303-340): y guard rows (full width) from the neighbours.
template <int NC>
__device__ inline void copy_gc_y_tile( const GcArgs & a, const int t ) {
    const gptr<float> buf = gbl( a.buf );
    const G g = a.g;
    const int px = a.px, py = a.py;
    const int tx = t % g.ntx, ty = t / g.ntx;
    const gptr<float> loc = buf + (size_t) t * g.vol * NC;
    const int lo = nbr( tx, ty, 0, -1, g, px, py ), hi = nbr( tx, ty, 0, +1, g, px, py );
    const int E = g.ext_x;
    if ( lo >= 0 ) {
        const gptr<const float> src = buf + (size_t) lo * g.vol * NC;
        for ( int k = threadIdx.x; k < g.gl * E * NC; k += blockDim.x ) {
            const int c = k % NC, e = k / NC, i = e % E, j = e / E;
            loc[( i + j * E ) * NC + c] = src[( i + ( g.ny + j ) * E ) * NC + c];
        }
    }
    if ( hi >= 0 ) {
        const gptr<const float> src = buf + (size_t) hi * g.vol * NC;
        for ( int k = threadIdx.x; k < g.gu * E * NC; k += blockDim.x ) {
            const int c = k % NC, e = k / NC, i = e % E, j = e / E;
            loc[( i + ( g.gl + g.ny + j ) * E ) * NC + c] = src[( i + ( g.gl + j ) * E ) * NC + c];
        }
    }
}

// Where element (i, j) of tile t takes its value from in copy_to_gc_x
// followed by copy_to_gc_y (grid.h:253-340): the y pass copies the y
// neighbour's row (as its own x pass left it), the x pass the x neighbour's
// column, so a guard element resolves first along y (tile row, row index),
// then along x within that tile row.  Returns false when the element keeps
// its own value (interior, or a guard at a non-periodic edge).  Every source
// it names is an interior element or a guard element that keeps its value,
// so one launch may copy all guards of all tiles at once.
__device__ inline bool copy_source( const G & g, int px, int py, int t, int i, int j, int & st, int & si, int & sj ) {
    int tx = t % g.ntx, ty = t / g.ntx;
    si = i;
    sj = j;
    if ( j < g.gl ) {
        const int n = nbr( tx, ty, 0, -1, g, px, py );
        if ( n >= 0 ) { ty = n / g.ntx; sj = g.ny + j; }
    } else if ( j >= g.gl + g.ny ) {
        const int n = nbr( tx, ty, 0, +1, g, px, py );
        if ( n >= 0 ) { ty = n / g.ntx; sj = j - g.ny; }
    }
    if ( i < g.gl ) {
        const int n = nbr( tx, ty, -1, 0, g, px, py );
        if ( n >= 0 ) { tx = n % g.ntx; si = g.nx + i; }
    } else if ( i >= g.gl + g.nx ) {
        const int n = nbr( tx, ty, +1, 0, g, px, py );
        if ( n >= 0 ) { tx = n % g.ntx; si = i - g.nx; }
    }
    st = ty * g.ntx + tx;
    return st != t || si != i || sj != j;
}

// copy_to_gc (x then y passes, grid.h:253-340) of tile t in one pass: every
// guard element of the tile copied from where copy_source resolves it (all
// tiles of a grid may run at once).
template <int NC>
__device__ inline void copy_gc_xy_tile( const GcArgs & a, const int t ) {
    const G g = a.g;
    const gptr<float> buf = gbl( a.buf );
    const int E = g.ext_x, gw = g.gl + g.gu;
    const int nrow = gw * E, ncell = nrow + g.ny * gw;
    for ( int k = threadIdx.x; k < ncell * NC; k += blockDim.x ) {
        const int c = k % NC, e = k / NC;
        int i, j;
        if ( e < g.gl * E ) {
            j = e / E;
            i = e % E;
        } else if ( e < nrow ) {
            const int f = e - g.gl * E;
            j = g.gl + g.ny + f / E;
            i = f % E;
        } else {
            const int f = e - nrow, q = f % gw;
            j = g.gl + f / gw;
            i = q < g.gl ? q : g.nx + q;
        }
        int st, si, sj;
        if ( copy_source( g, a.px, a.py, t, i, j, st, si, sj ) )
            buf[( (size_t) t * g.vol + i + j * E ) * NC + c] = buf[( (size_t) st * g.vol + si + sj * E ) * NC + c];
    }
}

constexpr int TILE_BLOCK = 256;

// kernel3_x / kernel3_y (grid.h:495-593): B = A[-1]*a + A[0]*b + A[+1]*c per
// component over all rows (x) or all columns (y) of the tile, A in LDS.
__device__ inline void kernel3_tile( const K3Args & args, const int t, float * A ) {
    const gptr<float> buf = gbl( args.buf );
    const G g = args.g;
    const int dir = args.dir;
    const float a = args.a, b = args.b, c = args.c;
    const gptr<float> tb = buf + (size_t) t * g.vol * 3;
    for ( int k = threadIdx.x; k < 3 * g.vol; k += blockDim.x ) A[k] = tb[k];
    __syncthreads();
    const int E = g.ext_x;
    const int step = ( dir == 0 ) ? 3 : 3 * E;
    const int x0 = ( dir == 0 ) ? g.gl : 0, nxr = ( dir == 0 ) ? g.nx : E;
    const int y0 = ( dir == 0 ) ? 0 : g.gl, nyr = ( dir == 0 ) ? g.ext_y : g.ny;
    for ( int k = threadIdx.x; k < nxr * nyr * 3; k += blockDim.x ) {
        const int comp = k % 3, e = k / 3;
        const int ix = x0 + e % nxr, iy = y0 + e / nxr;
        const int q = ( iy * E + ix ) * 3 + comp;
        tb[q] = A[q - step] * a + A[q] * b + A[q + step] * c;
    }
}

// EMF::advance(Current&) per tile (emf.cpp:513-566, yee_b :67-81,
// yeeJ_e :486-506): E and B of the tile in LDS planes, J from global.
__device__ inline void yee_tile( const YeeArgs & a, const int t, float * sm ) {
    const gptr<float> Eg = gbl( a.E );
    const gptr<float> Bg = gbl( a.B );
    const gptr<const float> Jg = gbl( a.J );
    const G g = a.g;
    const float dtdx = a.dtdx, dtdy = a.dtdy, dt = a.dt;
    const int vol = g.vol, ys = g.ext_x;
    float * Ex = sm; float * Ey = Ex + vol; float * Ez = Ey + vol;
    float * Bx = Ez + vol; float * By = Bx + vol; float * Bz = By + vol;
    const gptr<float> Et = Eg + (size_t) t * vol * 3;
    const gptr<float> Bt = Bg + (size_t) t * vol * 3;
    const gptr<const float> Jt = a.J ? Jg + (size_t) t * vol * 3 : nullptr;
    for ( int k = threadIdx.x; k < 3 * vol; k += blockDim.x ) {
        const int el = k / 3, c = k - 3 * el;
        Ex[c * vol + el] = Et[k];
        Bx[c * vol + el] = Bt[k];
    }
    __syncthreads();
    const float hx = dtdx / 2, hy = dtdy / 2;
    const int nb = ( g.nx + 2 ) * ( g.ny + 2 );
    auto bstep = [&]() {
        for ( int e = threadIdx.x; e < nb; e += blockDim.x ) {
            const int ix = e % ( g.nx + 2 ) - 1, iy = e / ( g.nx + 2 ) - 1;
            const int k = g.off + ix + iy * ys;
            Bx[k] += ( -hy * ( Ez[k + ys] - Ez[k] ) );
            By[k] += ( hx * ( Ez[k + 1] - Ez[k] ) );
            Bz[k] += ( -hx * ( Ey[k + 1] - Ey[k] ) + hy * ( Ex[k + ys] - Ex[k] ) );
        }
    };
    bstep();
    __syncthreads();
    for ( int e = threadIdx.x; e < nb; e += blockDim.x ) {
        const int ix = e % ( g.nx + 2 ), iy = e / ( g.nx + 2 );
        const int k = g.off + ix + iy * ys;
        const float jx = Jt ? Jt[3 * k] : 0.0f, jy = Jt ? Jt[3 * k + 1] : 0.0f, jz = Jt ? Jt[3 * k + 2] : 0.0f;
        Ex[k] += ( dtdy * ( Bz[k] - Bz[k - ys] ) ) - dt * jx;
        Ey[k] += ( -dtdx * ( Bz[k] - Bz[k - 1] ) ) - dt * jy;
        Ez[k] += ( dtdx * ( By[k] - By[k - 1] ) - dtdy * ( Bx[k] - Bx[k - ys] ) ) - dt * jz;
    }
    __syncthreads();
    bstep();
    __syncthreads();
    for ( int k = threadIdx.x; k < 3 * vol; k += blockDim.x ) {
        const int el = k / 3, c = k - 3 * el;
        Et[k] = Ex[c * vol + el];
        Bt[k] = Bx[c * vol + el];
    }
}

// Current::advance on a grid periodic along x and y (current.cpp:226-246:
// add_from_gc, copy_to_gc, then the filter's kernel3_x passes, each followed
// by copy_to_gc_x, grid.h:367-444, 253-340, 495-545) for tile t in one pass,
// from the deposit buffer `dep` into `J` (distinct buffers: neighbours read
// `dep` while J is written).  After that sequence on a periodic grid every
// element of a tile's region, guards included, holds the value of the global
// cell it stands for, so tile t computes its whole region itself:
//  * the summed current of every cell in columns [-gl - P, nx + gu + P) and
//    rows [-gl, ny + gu) (tile-local), with the additions of the x pass
//    (own value, + lower neighbour's upper guard, + upper neighbour's lower
//    guard) then the y pass (+ the y neighbours' guard rows as their own x
//    pass left them), in that order;
//  * P filter passes in LDS, A[i-1]*a + A[i]*b + A[i+1]*c, the valid columns
//    shrinking by one on each side per pass;
//  * the region [-gl, nx + gu) x [-gl, ny + gu) written, and the tile's
//    padding elements copied from `dep` (the sequence never touches them).
// With nx >= gl + gu + P and ny >= gl + gu every value a cell of the region
// needs lies in tile t's 3x3 neighbourhood.  LDS: current_lds_bytes.
constexpr int CUR_MAX_PASSES = 16;
struct CurArgs {
    const float * dep;
    float * J;
    G g;
    int passes;
    float a[CUR_MAX_PASSES], b[CUR_MAX_PASSES], c[CUR_MAX_PASSES];
};

inline size_t current_lds_bytes( const G & g, int passes ) {
    return sizeof( float ) * 2 * 3 * (size_t) ( g.nx + g.gl + g.gu + 2 * passes ) * g.ext_y;
}

// Loop over the n = rows x len elements of a 2-D range, row-major, thread
// tid taking elements tid, tid + blockDim.x, ...: (row, col) carried without
// a division per element.
template <class F>
__device__ inline void for_2d( int rows, int len, F && f ) {
    const int step = blockDim.x, dcol = step % len, drow = step / len;
    int row = threadIdx.x / len, col = threadIdx.x % len;
    for ( ; row < rows; ) {
        f( row, col );
        col += dcol;
        row += drow;
        if ( col >= len ) {
            col -= len;
            row++;
        }
    }
}

__device__ inline void current_tile( const CurArgs & a, const int t, float * sm ) {
    const G g = a.g;
    const int P = a.passes;
    const int E = g.ext_x;
    const int XW = g.nx + g.gl + g.gu + 2 * P, plane = XW * g.ext_y;
    float * A = sm;
    float * Bf = sm + 3 * plane;
    const gptr<const float> dep = gbl( a.dep );
    // first float of each deposit region of the 3x3 neighbourhood, index
    // (dy + 1) * 3 + dx + 1
    __shared__ long long s_nb[9];
    if ( threadIdx.x < 9 ) {
        const int k = threadIdx.x;
        s_nb[k] = (long long) nbr( t % g.ntx, t / g.ntx, k % 3 - 1, k / 3 - 1, g, 1, 1 ) * g.vol * 3;
    }
    __syncthreads();
    // deposit of neighbour k at tile-local (x, y), component c
    auto raw = [&]( int k, int x, int y, int c ) { return dep[s_nb[k] + ( ( x + g.gl ) + ( y + g.gl ) * E ) * 3 + c]; };
    // neighbour k (at dx, dy) at interior column i, any row y, after the x
    // pass; its x neighbours are k - 1 and k + 1 (inside the 3x3 block
    // whenever the terms apply)
    auto xsum = [&]( int k, int i, int y, int c ) {
        float v = raw( k, i, y, c );
        if ( i < g.gu ) v += raw( k - 1, g.nx + i, y, c );
        if ( i >= g.nx - g.gl ) v += raw( k + 1, i - g.nx, y, c );
        return v;
    };
    // summed current on the region, SoA planes A[c][row][col]
    for_2d( g.ext_y, 3 * XW, [&]( int row, int q ) {
        const int col = q / 3, c = q - 3 * col;
        const int xx = col - g.gl - P, yy = row - g.gl;
        const int dx = xx < 0 ? -1 : ( xx >= g.nx ? 1 : 0 );
        const int dy = yy < 0 ? -1 : ( yy >= g.ny ? 1 : 0 );
        const int i = xx - dx * g.nx, j = yy - dy * g.ny;
        const int k = ( dy + 1 ) * 3 + dx + 1;
        float s = xsum( k, i, j, c );
        if ( j < g.gu ) s += xsum( k - 3, i, g.ny + j, c );
        if ( j >= g.ny - g.gl ) s += xsum( k + 3, i, j - g.ny, c );
        A[c * plane + row * XW + col] = s;
    } );
    for ( int p = 0; p < P; p++ ) {
        __syncthreads();
        const float ca = a.a[p], cb = a.b[p], cc = a.c[p];
        float * const In = A;
        float * const Out = Bf;
        for_2d( 3 * g.ext_y, XW, [&]( int r, int x ) {
            if ( x > p && x < XW - 1 - p ) {
                const int k = r * XW + x;
                Out[k] = In[k - 1] * ca + In[k] * cb + In[k + 1] * cc;
            }
        } );
        A = Out;
        Bf = In;
    }
    __syncthreads();
    const gptr<float> Jt = gbl( a.J ) + (size_t) t * g.vol * 3;
    float * const R = A;
    for_2d( g.ext_y, 3 * E, [&]( int row, int q ) {
        const int col = q / 3, c = q - 3 * col;
        Jt[row * 3 * E + q] = R[c * plane + row * XW + col + P];
    } );
    // padding
    const int npad = 3 * ( g.vol - E * g.ext_y );
    if ( (int) threadIdx.x < npad ) {
        const size_t k = (size_t) 3 * E * g.ext_y + threadIdx.x;
        Jt[k] = dep[s_nb[4] + k];
    }
}

// Current::zero (grid.h:169-172) of tile t: every element of its region,
// guards and padding included.
__device__ inline void zero_tile( float * buf, const int vol3, const int t ) {
    const gptr<float> b = gbl( buf ) + (size_t) t * vol3;
    for ( int k = threadIdx.x; k < vol3; k += blockDim.x ) b[k] = 0.0f;
}

}  // namespace grid_ops
}  // namespace zpic
