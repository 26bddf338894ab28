// zpic_oracle.hpp -- TEST INFRASTRUCTURE, NOT PRODUCT CODE.
//
// A clean-room CPU restatement of the reference em2d timestep
// (openmp/em2d, scalar path) used only as the parity checker by tests/,
// __graft_entry__.smoke() and the cpu_baseline leg of bench.py.  The product
// (zpic-parallel_amd/) never includes, links or calls this file.
//
// Templated on the real type R:
//   R = float  : bit-for-bit the reference's FP32 arithmetic (same operation
//                order, std::fma exactly where the reference calls ops::fma,
//                correctly rounded sqrt / division, no contraction: build with
//                -ffp-contract=off).  Pinned against golden vectors dumped from
//                the compiled reference (tests/golden/).
//   R = double : the same algorithm evaluated in FP64, used to measure the
//                discrete charge-continuity residual of the FP32 paths.
//
// Where the reference runs with OpenMP, the order of per-tile results it
// produces with ONE thread is reproduced here for any thread count: every
// per-tile stage writes only its own tile, the tile sort is executed in tile
// order, and reductions over tiles are summed in tile order.
//
// Functions follow (reference file:line):
//   interpolate_fld ........ species.cpp:43-92
//   dudt_boris ............. species.cpp:117-178
//   rgamma ................. species.cpp:22-26
//   dep_current_seg ........ species.cpp:267-310
//   split2d ................ species.cpp:336-430
//   move_deposit_kernel .... species.cpp:1400-1493, Species::move :2138-2164
//   push_kernel ............ species.cpp:1626-1682, Species::push :2302-2335
//   tile_sort .............. particles.cpp:212-614, 855-904
//   add_from_gc/copy_to_gc . grid.h:253-445
//   kernel3_x / kernel3_y .. grid.h:495-593, Filter:: filter.h:9-78
//   yee_b / yeeJ_e / advance emf.cpp:67-81, 486-580
//   get_energy ............. emf.cpp:662-713, species.h:365-368
//   initialize / inject .... species.cpp:1730-1788, density.cpp:103-1067
//                            (Uniform, Step, Slab, Sphere; whole box, window
//                            origin at 0)
//   ThermalCorr / Thermal .. udist.cpp:42-136, zrandom random.h:23-155
//   dep_charge_kernel ...... species.cpp:2348-2410
#pragma once

#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <stdexcept>
#include <string>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace oracle {

template <class R> struct V2 { R x, y; };
template <class R> struct V3 { R x, y, z; };
struct I2 { int x, y; };

// ---------------------------------------------------------------------------
// Tiled grid geometry: tile-major storage, each tile (nx.x+gl+gu) x (nx.y+gl+gu)
// padded to a multiple of 4 elements (grid.h:77-92, utils.h:24-27).
struct Geom {
    int ntx, nty, nx, ny;
    int glx, gux, gly, guy;
    int ext_x, ext_y, vol, off;
    Geom() = default;
    Geom( int ntx_, int nty_, int nx_, int ny_, int gl, int gu )
        : ntx( ntx_ ), nty( nty_ ), nx( nx_ ), ny( ny_ ), glx( gl ), gux( gu ), gly( gl ), guy( gu ) {
        ext_x = glx + nx + gux;
        ext_y = gly + ny + guy;
        vol = ( ext_x * ext_y + 3 ) & ~3;
        off = gly * ext_x + glx;
    }
    int ntiles() const { return ntx * nty; }
    size_t size() const { return (size_t) vol * ntiles(); }
};

template <class T>
struct Grid {
    Geom g;
    bool per_x = true, per_y = true;
    std::vector<T> d;
    Grid() = default;
    explicit Grid( const Geom & g_ ) : g( g_ ), d( g_.size(), T{} ) {}
    T * tile( int t ) { return d.data() + (size_t) t * g.vol; }

    // Neighbour tile index along x / y with optional periodic wrap, -1 if none.
    int nbr_x( int tx, int ty, int dx ) const {
        int n = tx + dx;
        if ( per_x ) { if ( n < 0 ) n += g.ntx; if ( n >= g.ntx ) n -= g.ntx; }
        if ( n < 0 || n >= g.ntx ) return -1;
        return ty * g.ntx + n;
    }
    int nbr_y( int tx, int ty, int dy ) const {
        int n = ty + dy;
        if ( per_y ) { if ( n < 0 ) n += g.nty; if ( n >= g.nty ) n -= g.nty; }
        if ( n < 0 || n >= g.nty ) return -1;
        return n * g.ntx + tx;
    }

    // grid.h:253-354 -- guard cells take the neighbours' edge values, x then y
    void copy_to_gc_x() {
        const int E = g.ext_x;
        for ( int t = 0; t < g.ntiles(); t++ ) {
            int tx = t % g.ntx, ty = t / g.ntx;
            T * loc = tile( t );
            int lo = nbr_x( tx, ty, -1 ), hi = nbr_x( tx, ty, +1 );
            if ( lo >= 0 ) {
                const T * src = tile( lo );
                for ( int j = 0; j < g.ext_y; j++ )
                    for ( int i = 0; i < g.glx; i++ ) loc[i + j * E] = src[g.nx + i + j * E];
            }
            if ( hi >= 0 ) {
                const T * src = tile( hi );
                for ( int j = 0; j < g.ext_y; j++ )
                    for ( int i = 0; i < g.gux; i++ ) loc[g.glx + g.nx + i + j * E] = src[g.glx + i + j * E];
            }
        }
    }
    void copy_to_gc_y() {
        const int E = g.ext_x;
        for ( int t = 0; t < g.ntiles(); t++ ) {
            int tx = t % g.ntx, ty = t / g.ntx;
            T * loc = tile( t );
            int lo = nbr_y( tx, ty, -1 ), hi = nbr_y( tx, ty, +1 );
            if ( lo >= 0 ) {
                const T * src = tile( lo );
                for ( int j = 0; j < g.gly; j++ )
                    for ( int i = 0; i < E; i++ ) loc[i + j * E] = src[i + ( g.ny + j ) * E];
            }
            if ( hi >= 0 ) {
                const T * src = tile( hi );
                for ( int j = 0; j < g.guy; j++ )
                    for ( int i = 0; i < E; i++ ) loc[i + ( g.gly + g.ny + j ) * E] = src[i + ( g.gly + j ) * E];
            }
        }
    }
    void copy_to_gc() { copy_to_gc_x(); copy_to_gc_y(); }

    // grid.h:360-445 -- neighbours' guard deposits are added into the interior,
    // a complete x pass over all tiles before the y pass.
    void add_from_gc() {
        const int E = g.ext_x;
        for ( int t = 0; t < g.ntiles(); t++ ) {
            int tx = t % g.ntx, ty = t / g.ntx;
            T * loc = tile( t );
            int lo = nbr_x( tx, ty, -1 ), hi = nbr_x( tx, ty, +1 );
            if ( lo >= 0 ) {
                const T * src = tile( lo );
                for ( int j = 0; j < g.ext_y; j++ )
                    for ( int i = 0; i < g.gux; i++ ) add( loc[g.glx + i + j * E], src[g.glx + g.nx + i + j * E] );
            }
            if ( hi >= 0 ) {
                const T * src = tile( hi );
                for ( int j = 0; j < g.ext_y; j++ )
                    for ( int i = 0; i < g.glx; i++ ) add( loc[g.nx + i + j * E], src[i + j * E] );
            }
        }
        for ( int t = 0; t < g.ntiles(); t++ ) {
            int tx = t % g.ntx, ty = t / g.ntx;
            T * loc = tile( t );
            int lo = nbr_y( tx, ty, -1 ), hi = nbr_y( tx, ty, +1 );
            if ( lo >= 0 ) {
                const T * src = tile( lo );
                for ( int j = 0; j < g.guy; j++ )
                    for ( int i = 0; i < E; i++ ) add( loc[i + ( g.gly + j ) * E], src[i + ( g.gly + g.ny + j ) * E] );
            }
            if ( hi >= 0 ) {
                const T * src = tile( hi );
                for ( int j = 0; j < g.gly; j++ )
                    for ( int i = 0; i < E; i++ ) add( loc[i + ( g.ny + j ) * E], src[i + j * E] );
            }
        }
    }

    template <class S> static void add( V3<S> & a, const V3<S> & b ) { a.x += b.x; a.y += b.y; a.z += b.z; }
    template <class S> static void add( S & a, const S & b ) { a += b; }
};

// 3-point convolution [a,b,c] along x over all rows of every tile, then x guards
// (grid.h:495-541); along y over all columns, then y guards (grid.h:549-593).
// Per element: (A[-1]*a + A[0]*b) + A[+1]*c, component-wise.
template <class R>
void kernel3( Grid<V3<R>> & f, int dir, R a, R b, R c ) {
    const Geom & g = f.g;
    const int E = g.ext_x;
    const int step = ( dir == 0 ) ? 1 : E;
    std::vector<V3<R>> A( g.vol );
    for ( int t = 0; t < g.ntiles(); t++ ) {
        V3<R> * buf = f.tile( t );
        std::copy( buf, buf + g.vol, A.begin() );
        int y0 = ( dir == 0 ) ? 0 : g.gly, y1 = ( dir == 0 ) ? g.ext_y : g.gly + g.ny;
        int x0 = ( dir == 0 ) ? g.glx : 0, x1 = ( dir == 0 ) ? g.glx + g.nx : E;
        for ( int iy = y0; iy < y1; iy++ )
            for ( int ix = x0; ix < x1; ix++ ) {
                int k = iy * E + ix;
                const V3<R> & m = A[k - step], & o = A[k], & p = A[k + step];
                buf[k].x = m.x * a + o.x * b + p.x * c;
                buf[k].y = m.y * a + o.y * b + p.y * c;
                buf[k].z = m.z * a + o.z * b + p.z * c;
            }
    }
    if ( dir == 0 ) f.copy_to_gc_x(); else f.copy_to_gc_y();
}

// Filter description: kind 0 none, 1 binomial, 2 compensated; dir 0 x / 1 y.
struct FilterSpec { int kind = 0, dir = 0, order = 1; };

template <class R>
void apply_filter( Grid<V3<R>> & J, const FilterSpec & fs ) {
    if ( fs.kind == 0 ) return;
    unsigned order = fs.order > 0 ? fs.order : 1;
    for ( unsigned i = 0; i < order; i++ ) kernel3<R>( J, fs.dir, R( 0.25f ), R( 0.5f ), R( 0.25f ) );
    if ( fs.kind == 2 ) {
        // filter.h:63-66: a = -1, b = (4 + 2 n) / n (double, narrowed), norm = 2a + b
        R a = R( -1.0f );
        R b = R( ( 4.0 + 2.0 * order ) / order );
        R norm = R( 2 ) * a + b;
        kernel3<R>( J, fs.dir, a / norm, b / norm, a / norm );
    }
}

// ---------------------------------------------------------------------------
// Marsaglia MWC + polar Box-Muller (random.h:23-155)
struct Rng {
    uint32_t sx, sy;
    double cached;
    Rng( int id, uint32_t seedx, uint32_t seedy ) {
        sx = seedx + (uint32_t) id;
        sy = seedy + ( ( (uint32_t) id ) ^ 0xFFFFFFFFu );
        cached = std::numeric_limits<double>::max();
    }
    uint32_t next() {
        sx = 36969u * ( sx & 65535u ) + ( sx >> 16 );
        sy = 18000u * ( sy & 65535u ) + ( sy >> 16 );
        return ( sx << 16 ) + sy;
    }
    double open01() { return ( next() + 1.0 ) * 0x1.fffffffep-33; }
    double normal() {
        if ( cached != std::numeric_limits<double>::max() ) {
            double r = cached;
            cached = std::numeric_limits<double>::max();
            return r;
        }
        double v1, v2, r2;
        do {
            v1 = open01() * 2.0 - 1.0;
            v2 = open01() * 2.0 - 1.0;
            r2 = v1 * v1 + v2 * v2;
        } while ( r2 == 0.0 || r2 >= 1.0 );
        double f = std::sqrt( -2.0 * std::log( r2 ) / r2 );
        cached = v1 * f;
        return v2 * f;
    }
};

// Momentum distributions (udist.h/.cpp). kind: 0 none, 1 cold, 2 thermal, 3 thermal_corr
struct UDist {
    int kind = 0;
    float uth[3] = { 0, 0, 0 };
    float ufl[3] = { 0, 0, 0 };
    int npmin = 2;
};

// Density profile (density.h:30-132): 0 none, 1 uniform, 2 step, 3 slab,
// 4 sphere; dir 0 = x, 1 = y; lengths in simulation units.
struct Profile {
    int kind = 1;
    int dir = 0;
    float pos = 0, begin = 0, end = 0;
    float cx = 0, cy = 0, radius = 0;
};

// ---------------------------------------------------------------------------
template <class R>
struct Particles {
    int ntiles = 0;
    std::vector<int> np, off;
    std::vector<I2> ix;
    std::vector<V2<R>> x;
    std::vector<V3<R>> u;
    void alloc( int nt, size_t max_part ) {
        ntiles = nt;
        np.assign( nt, 0 );
        off.assign( nt, 0 );
        ix.resize( max_part );
        x.resize( max_part );
        u.resize( max_part );
    }
    size_t total() const { size_t s = 0; for ( int v : np ) s += v; return s; }
};

template <class R>
struct Species {
    std::string name;
    float m_q;
    int ppcx, ppcy;
    int id = 0;
    float n0 = 1.0f;
    UDist ud;
    Profile prof;
    int per_x = 1, per_y = 1;

    // derived at initialize()
    R q, dt;
    V2<R> dx;
    V2<R> box;
    int ntx, nty, nx, ny;
    size_t max_part;
    Particles<R> p, tmp;
    // sort scratch
    std::vector<int> leav, nleav, npt, new_np;
    double energy = 0;
    uint64_t nmove = 0;
    int iter = 0;
};

// ---------------------------------------------------------------------------
// Per-particle arithmetic.  Every expression keeps the reference's operation
// order; std::fma appears exactly where the reference uses ops::fma.

template <class R>
inline void interp( const V3<R> * E, const V3<R> * B, int ys, I2 ix, V2<R> x, V3<R> & e, V3<R> & b ) {
    const int i = ix.x, j = ix.y;
    const R s0x = R( 0.5f ) - x.x, s1x = R( 0.5f ) + x.x;
    const R s0y = R( 0.5f ) - x.y, s1y = R( 0.5f ) + x.y;
    const int hx = x.x < 0, hy = x.y < 0;
    const int ih = i - hx, jh = j - hy;
    const R s0xh = R( 1 - hx ) - x.x, s1xh = R( hx ) + x.x;
    const R s0yh = R( 1 - hy ) - x.y, s1yh = R( hy ) + x.y;
    auto at = [ys]( const V3<R> * F, int a, int c ) -> const V3<R> & { return F[a + c * ys]; };

    e.x = ( at( E, ih, j ).x * s0xh + at( E, ih + 1, j ).x * s1xh ) * s0y +
          ( at( E, ih, j + 1 ).x * s0xh + at( E, ih + 1, j + 1 ).x * s1xh ) * s1y;
    e.y = ( at( E, i, jh ).y * s0x + at( E, i + 1, jh ).y * s1x ) * s0yh +
          ( at( E, i, jh + 1 ).y * s0x + at( E, i + 1, jh + 1 ).y * s1x ) * s1yh;
    e.z = ( at( E, i, j ).z * s0x + at( E, i + 1, j ).z * s1x ) * s0y +
          ( at( E, i, j + 1 ).z * s0x + at( E, i + 1, j + 1 ).z * s1x ) * s1y;

    b.x = ( at( B, i, jh ).x * s0x + at( B, i + 1, jh ).x * s1x ) * s0yh +
          ( at( B, i, jh + 1 ).x * s0x + at( B, i + 1, jh + 1 ).x * s1x ) * s1yh;
    b.y = ( at( B, ih, j ).y * s0xh + at( B, ih + 1, j ).y * s1xh ) * s0y +
          ( at( B, ih, j + 1 ).y * s0xh + at( B, ih + 1, j + 1 ).y * s1xh ) * s1y;
    b.z = ( at( B, ih, jh ).z * s0xh + at( B, ih + 1, jh ).z * s1xh ) * s0yh +
          ( at( B, ih, jh + 1 ).z * s0xh + at( B, ih + 1, jh + 1 ).z * s1xh ) * s1yh;
}

template <class R>
inline V3<R> boris( R alpha, V3<R> e, V3<R> b, V3<R> u, double & energy ) {
    e.x *= alpha; e.y *= alpha; e.z *= alpha;
    V3<R> ut{ u.x + e.x, u.y + e.y, u.z + e.z };
    {
        const R usq = std::fma( ut.z, ut.z, std::fma( ut.y, ut.y, ut.x * ut.x ) );
        const R g = std::sqrt( R( 1 ) + usq );
        energy += usq / ( g + R( 1 ) );
        const R ag = alpha / g;
        b.x *= ag; b.y *= ag; b.z *= ag;
    }
    V3<R> v;
    v.x = std::fma( b.z, ut.y, ut.x );
    v.y = std::fma( b.x, ut.z, ut.y );
    v.z = std::fma( b.y, ut.x, ut.z );
    v.x = std::fma( -b.y, ut.z, v.x );
    v.y = std::fma( -b.z, ut.x, v.y );
    v.z = std::fma( -b.x, ut.y, v.z );
    {
        const R s = R( 2 ) / std::fma( b.z, b.z, std::fma( b.y, b.y, std::fma( b.x, b.x, R( 1 ) ) ) );
        b.x *= s; b.y *= s; b.z *= s;
    }
    ut.x = std::fma( b.z, v.y, ut.x );
    ut.y = std::fma( b.x, v.z, ut.y );
    ut.z = std::fma( b.y, v.x, ut.z );
    ut.x = std::fma( -b.y, v.z, ut.x );
    ut.y = std::fma( -b.z, v.x, ut.y );
    ut.z = std::fma( -b.x, v.y, ut.z );
    ut.x += e.x; ut.y += e.y; ut.z += e.z;
    return ut;
}

template <class R>
inline R inv_gamma( V3<R> u ) {
    return R( 1 ) / std::sqrt( std::fma( u.z, u.z, std::fma( u.y, u.y, std::fma( u.x, u.x, R( 1 ) ) ) ) );
}

// One straight segment inside one cell: 8 current contributions
// (Jx at 2 nodes, Jy at 2 nodes, Jz at 4 nodes).  J points at cell (0,0).
template <class R>
inline void deposit_segment( I2 c, V2<R> a, V2<R> b, V2<R> qnx, R qvz, V3<R> * J, int ys ) {
    const R h = R( 0.5f );
    const R A0x = h - a.x, A1x = h + a.x, B0x = h - b.x, B1x = h + b.x;
    const R A0y = h - a.y, A1y = h + a.y, B0y = h - b.y, B1y = h + b.y;
    const R jx = qnx.x * ( b.x - a.x );
    const R jy = qnx.y * ( b.y - a.y );
    const R wy0 = h * ( A0y + B0y ), wy1 = h * ( A1y + B1y );
    const R wx0 = h * ( A0x + B0x ), wx1 = h * ( A1x + B1x );
    V3<R> * n00 = J + c.x + ys * c.y;
    V3<R> * n10 = n00 + 1;
    V3<R> * n01 = n00 + ys;
    V3<R> * n11 = n01 + 1;
    const R two = R( 2.0f );
    n00->x += jx * wy0;
    n00->y += jy * wx0;
    n00->z += qvz * ( A0x * A0y + B0x * B0y + ( A0x * B0y - B0x * A0y ) / two );
    n10->y += jy * wx1;
    n10->z += qvz * ( A1x * A0y + B1x * B0y + ( A1x * B0y - B1x * A0y ) / two );
    n01->x += jx * wy1;
    n01->z += qvz * ( A0x * A1y + B0x * B1y + ( A0x * B1y - B0x * A1y ) / two );
    n11->z += qvz * ( A1x * A1y + B1x * B1y + ( A1x * B1y - B1x * A1y ) / two );
}

template <class R>
struct Seg { I2 c; V2<R> a, b; R qvz; };

// Splits the motion x0 -> x1 (= x0 + d) at cell edges into up to 3 segments.
// Returns the number of segments; di receives the cell displacement.
template <class R>
inline int split_path( I2 c0, V2<R> x0, V2<R> x1, V2<R> d, R qvz, I2 & di, Seg<R> s[3] ) {
    const R h = R( 0.5f );
    di.x = ( x1.x >= h ) - ( x1.x < -h );
    di.y = ( x1.y >= h ) - ( x1.y < -h );
    const bool cx = di.x != 0, cy = di.y != 0;
    s[0] = { c0, x0, x1, qvz };
    if ( !cx && !cy ) return 1;
    if ( !cx ) {
        // single y crossing
        const R yb = h * di.y;
        const R f = ( yb - x0.y ) / d.y;
        const R xb = x0.x + d.x * f;
        s[0].b = { xb, yb };
        s[0].qvz = qvz * f;
        s[1] = { { c0.x, c0.y + di.y }, { xb, -yb }, { x1.x, x1.y - di.y }, qvz * ( R( 1 ) - f ) };
        return 2;
    }
    // x crossing
    const R xb = h * di.x;
    R f = ( xb - x0.x ) / d.x;
    const R yb = x0.y + d.y * f;
    s[0].b = { xb, yb };
    s[0].qvz = qvz * f;
    s[1] = { { c0.x + di.x, c0.y }, { -xb, yb }, { x1.x - di.x, x1.y }, qvz * ( R( 1 ) - f ) };
    if ( !cy ) return 2;
    const R yb2 = h * di.y;
    if ( yb >= -h && yb < h ) {
        // the y crossing happens on the second segment
        f = ( yb2 - yb ) / ( x1.y - yb );
        const R xb2 = -xb + ( x1.x - xb ) * f;
        s[2] = { { s[1].c.x, s[1].c.y + di.y }, { xb2, -yb2 }, { s[1].b.x, s[1].b.y - di.y }, s[1].qvz * ( R( 1 ) - f ) };
        s[1].b = { xb2, yb2 };
        s[1].qvz *= f;
    } else {
        // the y crossing happens on the first segment
        f = ( yb2 - x0.y ) / ( yb - x0.y );
        const R xb2 = x0.x + ( xb - x0.x ) * f;
        s[2] = { { s[0].c.x, s[0].c.y + di.y }, { xb2, -yb2 }, { s[0].b.x, s[0].b.y - di.y }, s[0].qvz * ( R( 1 ) - f ) };
        s[0].b = { xb2, yb2 };
        s[0].qvz *= f;
        s[1].c.y += di.y;
        s[1].a.y -= di.y;
        s[1].b.y -= di.y;
    }
    return 3;
}

// ---------------------------------------------------------------------------
template <class R>
struct Sim {
    Geom g;
    V2<R> box;
    double dt;
    V2<R> dx;
    Grid<V3<R>> E, B, J;
    FilterSpec filter;
    std::vector<Species<R>> sp;
    int iter = 0;

    Sim( int ntx, int nty, int nx, int ny, float boxx, float boxy, double dt_ )
        : g( ntx, nty, nx, ny, 1, 2 ), dt( dt_ ), E( g ), B( g ), J( g ) {
        box = { R( boxx ), R( boxy ) };
        dx = { box.x / R( (unsigned) ( nx * ntx ) ), box.y / R( (unsigned) ( ny * nty ) ) };
        // emf.cpp:22-28 Courant condition (computed in float as the reference)
        float fdx = (float) dx.x, fdy = (float) dx.y;
        float cour = std::sqrt( 1.0f / ( 1.0f / ( fdx * fdx ) + 1.0f / ( fdy * fdy ) ) );
        if ( dt >= cour ) throw std::runtime_error( "Invalid timestep, courant condition violation" );
    }

    // Simulation::add_species -> Species::initialize (species.cpp:1730-1788)
    void add_species( Species<R> s ) {
        s.id = (int) sp.size() + 1;
        s.ntx = g.ntx; s.nty = g.nty; s.nx = g.nx; s.ny = g.ny;
        s.box = box;
        s.dt = R( (float) dt );
        if ( sizeof( R ) == 8 ) s.dt = R( dt );
        s.q = std::copysign( R( s.n0 ), R( s.m_q ) ) / R( (unsigned) ( s.ppcx * s.ppcy ) );
        V2<R> dims{ R( (unsigned) ( g.nx * g.ntx ) ), R( (unsigned) ( g.ny * g.nty ) ) };
        s.dx = { box.x / dims.x, box.y / dims.y };
        s.max_part = (size_t) ( 1.2 * (double) dims.x * (double) dims.y * s.ppcx * s.ppcy );
        const int nt = g.ntiles();
        s.p.alloc( nt, s.max_part );
        s.tmp.alloc( nt, s.max_part );
        s.leav.resize( s.max_part );
        s.nleav.assign( nt, 0 );
        s.npt.assign( 9 * nt, 0 );
        s.new_np.assign( nt, 0 );
        inject_profile( s );
        set_udist( s );
        sp.push_back( s );
    }

    // Density::*::np_inject + exclusive scan + Density::*::inject over the
    // whole box (density.cpp:103-1067).  Per tile, lattice point (j, i) outer,
    // cell inner (the "mk2" order); a point is kept when the profile says so:
    //   step  (:316-573):  t = float(cell + (pos + 0.5)) > (pos0 - 0) / dx.x
    //   slab  (:575-854):  (begin / dx.x) <= t < (end / dx.x)
    //   sphere (:856-1067): |float((cell + pos + 0.5) * dx) - centre|^2 < r^2
    // the step and the slab are placed with dx.x along either axis, as the
    // reference does; t and the sphere point are formed in double and
    // rounded to float.
    static bool keeps( const Profile & pr, float dxx, float dxy, int gx, int gy, float px, float py ) {
        switch ( pr.kind ) {
        case 1: return true;
        case 2:
        case 3: {
            const float t = ( pr.dir == 0 ) ? (float) ( (double) gx + ( (double) px + 0.5 ) )
                                            : (float) ( (double) gy + ( (double) py + 0.5 ) );
            if ( pr.kind == 2 ) return t > ( pr.pos - 0.0f ) / dxx;
            return t >= ( pr.begin - 0.0f ) / dxx && t < ( pr.end - 0.0f ) / dxx;
        }
        case 4: {
            const float X = (float) ( ( (double) gx + ( (double) px + 0.5 ) ) * (double) dxx );
            const float Y = (float) ( ( (double) gy + ( (double) py + 0.5 ) ) * (double) dxy );
            const float r2 = pr.radius * pr.radius;
            const float ax = X - pr.cx, ay = Y - pr.cy;
            return ax * ax + ay * ay < r2;
        }
        default: return false;
        }
    }

    void inject_profile( Species<R> & s ) {
        const int nt = g.ntiles();
        const int vol = g.nx * g.ny;
        const double fx = 1.0 / s.ppcx, fy = 1.0 / s.ppcy;
        const float dxx = (float) s.dx.x, dxy = (float) s.dx.y;
        // lattice positions as float (the reference's make_float2), as R for storage
        std::vector<float> lx( s.ppcx ), ly( s.ppcy );
        for ( int i = 0; i < s.ppcx; i++ ) lx[i] = (float) ( fx * ( i + 0.5 ) - 0.5 );
        for ( int j = 0; j < s.ppcy; j++ ) ly[j] = (float) ( fy * ( j + 0.5 ) - 0.5 );
        int acc = 0;
        for ( int t = 0; t < nt; t++ ) {
            const int tx = t % g.ntx, ty = t / g.ntx;
            int n = 0;
            for ( int j = 0; j < s.ppcy; j++ )
                for ( int i = 0; i < s.ppcx; i++ )
                    for ( int c = 0; c < vol; c++ )
                        n += keeps( s.prof, dxx, dxy, tx * g.nx + c % g.nx, ty * g.ny + c / g.nx, lx[i], ly[j] );
            s.p.off[t] = acc;
            s.p.np[t] = n;
            acc += n;
        }
        if ( (size_t) acc > s.max_part ) throw std::runtime_error( "particle buffer too small for the injection" );
        for ( int t = 0; t < nt; t++ ) {
            const int tx = t % g.ntx, ty = t / g.ntx;
            int n = s.p.off[t];
            for ( int j = 0; j < s.ppcy; j++ )
                for ( int i = 0; i < s.ppcx; i++ ) {
                    V2<R> pos{ R( lx[i] ), R( ly[j] ) };
                    if ( sizeof( R ) == 8 ) pos = { R( fx * ( i + 0.5 ) - 0.5 ), R( fy * ( j + 0.5 ) - 0.5 ) };
                    for ( int c = 0; c < vol; c++ ) {
                        if ( !keeps( s.prof, dxx, dxy, tx * g.nx + c % g.nx, ty * g.ny + c / g.nx, lx[i], ly[j] ) )
                            continue;
                        s.p.ix[n] = { c % g.nx, c / g.nx };
                        s.p.x[n] = pos;
                        s.p.u[n] = { 0, 0, 0 };
                        n++;
                    }
                }
        }
    }

    // udist.cpp:42-136.  The three normal deviates of one particle are drawn
    // z first, then y, then x: that is the argument evaluation order of
    // make_float3(...) in the reference build (g++, x86-64).
    void set_udist( Species<R> & s ) {
        const UDist & ud = s.ud;
        const int nt = g.ntiles();
        const uint32_t seedx = 12345u + (uint32_t) s.id, seedy = 67890u;
        const R fl[3] = { R( ud.ufl[0] ), R( ud.ufl[1] ), R( ud.ufl[2] ) };
        for ( int t = 0; t < nt; t++ ) {
            const int o = s.p.off[t], n = s.p.np[t];
            if ( ud.kind == 0 ) { for ( int i = 0; i < n; i++ ) s.p.u[o + i] = { 0, 0, 0 }; continue; }
            if ( ud.kind == 1 ) { for ( int i = 0; i < n; i++ ) s.p.u[o + i] = { fl[0], fl[1], fl[2] }; continue; }
            Rng rng( t, seedx, seedy );
            auto draw = [&]( V3<R> & v ) {
                double rz = rng.normal();
                double ry = rng.normal();
                double rx = rng.normal();
                v.x = R( (float) ( ud.uth[0] * rx ) );
                v.y = R( (float) ( ud.uth[1] * ry ) );
                v.z = R( (float) ( ud.uth[2] * rz ) );
                if ( sizeof( R ) == 8 ) v = { R( ud.uth[0] * rx ), R( ud.uth[1] * ry ), R( ud.uth[2] * rz ) };
            };
            if ( ud.kind == 2 ) {
                // Thermal: ufl + uth * N(0,1) evaluated in double, then narrowed
                for ( int i = 0; i < n; i++ ) {
                    double rz = rng.normal();
                    double ry = rng.normal();
                    double rx = rng.normal();
                    double vx = ud.ufl[0] + ud.uth[0] * rx, vy = ud.ufl[1] + ud.uth[1] * ry, vz = ud.ufl[2] + ud.uth[2] * rz;
                    if ( sizeof( R ) == 8 ) s.p.u[o + i] = { R( vx ), R( vy ), R( vz ) };
                    else s.p.u[o + i] = { R( (float) vx ), R( (float) vy ), R( (float) vz ) };
                }
                continue;
            }
            // ThermalCorr: remove the per-cell mean when a cell holds > npmin particles
            const int cells = g.nx * g.ny;
            std::vector<V3<R>> mean( cells, V3<R>{ 0, 0, 0 } );
            std::vector<int> cnt( cells, 0 );
            for ( int i = 0; i < n; i++ ) {
                V3<R> th; draw( th );
                s.p.u[o + i] = th;
                const int c = s.p.ix[o + i].x + g.nx * s.p.ix[o + i].y;
                cnt[c] += 1;
                mean[c].x += th.x; mean[c].y += th.y; mean[c].z += th.z;
            }
            for ( int c = 0; c < cells; c++ ) {
                if ( cnt[c] > ud.npmin ) {
                    mean[c].x /= R( cnt[c] ); mean[c].y /= R( cnt[c] ); mean[c].z /= R( cnt[c] );
                } else mean[c] = { 0, 0, 0 };
            }
            for ( int i = 0; i < n; i++ ) {
                const int c = s.p.ix[o + i].x + g.nx * s.p.ix[o + i].y;
                V3<R> & v = s.p.u[o + i];
                v.x += fl[0] - mean[c].x;
                v.y += fl[1] - mean[c].y;
                v.z += fl[2] - mean[c].z;
            }
        }
    }

    // ---- Species::push (species.cpp:2302-2335, 1626-1682)
    void push( Species<R> & s ) {
        const R alpha = ( sizeof( R ) == 8 ) ? R( 0.5 * s.dt / s.m_q )
                                              : R( (float) ( 0.5 * (double) s.dt / (double) s.m_q ) );
        const int nt = g.ntiles();
        std::vector<double> et( nt, 0.0 );
        #pragma omp parallel for schedule(dynamic)
        for ( int t = 0; t < nt; t++ ) {
            std::vector<V3<R>> El( E.tile( t ), E.tile( t ) + g.vol ), Bl( B.tile( t ), B.tile( t ) + g.vol );
            const V3<R> * e0 = El.data() + g.off;
            const V3<R> * b0 = Bl.data() + g.off;
            double en = 0;
            const int o = s.p.off[t], n = s.p.np[t];
            for ( int i = 0; i < n; i++ ) {
                V3<R> e, b;
                interp<R>( e0, b0, g.ext_x, s.p.ix[o + i], s.p.x[o + i], e, b );
                s.p.u[o + i] = boris<R>( alpha, e, b, s.p.u[o + i], en );
            }
            et[t] = en;
        }
        double acc = 0;
        for ( int t = 0; t < nt; t++ ) acc += et[t];
        s.energy = 0 + acc;
    }

    // ---- Species::move(J) (species.cpp:2138-2164, 1400-1493)
    void move_deposit( Species<R> & s ) {
        const V2<R> dt_dx{ s.dt / s.dx.x, s.dt / s.dx.y };
        const V2<R> qnx{ s.q * s.dx.x / s.dt, s.q * s.dx.y / s.dt };
        const int nt = g.ntiles();
        #pragma omp parallel for schedule(dynamic)
        for ( int t = 0; t < nt; t++ ) {
            std::vector<V3<R>> Jl( g.vol, V3<R>{ 0, 0, 0 } );
            V3<R> * j0 = Jl.data() + g.off;
            const int o = s.p.off[t], n = s.p.np[t];
            for ( int i = 0; i < n; i++ ) {
                const V3<R> pu = s.p.u[o + i];
                const V2<R> x0 = s.p.x[o + i];
                const I2 c0 = s.p.ix[o + i];
                const R rg = inv_gamma<R>( pu );
                const V2<R> d{ dt_dx.x * rg * pu.x, dt_dx.y * rg * pu.y };
                const R qvz = s.q * pu.z * rg * R( 0.5f );
                V2<R> x1{ x0.x + d.x, x0.y + d.y };
                I2 di;
                Seg<R> sg[3];
                const int ns = split_path<R>( c0, x0, x1, d, qvz, di, sg );
                for ( int k = 0; k < ns; k++ ) deposit_segment<R>( sg[k].c, sg[k].a, sg[k].b, qnx, sg[k].qvz, j0, g.ext_x );
                x1.x -= di.x; x1.y -= di.y;
                s.p.x[o + i] = x1;
                s.p.ix[o + i] = { c0.x + di.x, c0.y + di.y };
            }
            V3<R> * dst = J.tile( t );
            for ( int k = 0; k < g.ext_x * g.ext_y; k++ ) {
                dst[k].x += Jl[k].x; dst[k].y += Jl[k].y; dst[k].z += Jl[k].z;
            }
        }
        for ( int t = 0; t < nt; t++ ) s.nmove += s.p.np[t];
    }

    // ---- Particles::tile_sort (particles.cpp:212-614, 855-904), no injection room
    //
    // Particles that left their tile are classified into 8 directions.  Each
    // tile's new slot range is the exclusive scan of the new counts.  Moving
    // particles are staged in `tmp`, where each target tile receives, in
    // source-tile order, the source's shifted stayers (if it is itself) and
    // then one block per direction 0..8.  Holes left below n0 are filled with
    // the surviving particles above n0 in increasing order; the staged
    // particles are then appended (or prepended into the gap when the tile's
    // range moved down).
    int dir_target( const Species<R> & s, int t, int d ) const {
        int tx = t % g.ntx + ( d % 3 - 1 ), ty = t / g.ntx + ( d / 3 - 1 );
        if ( s.per_x ) { if ( tx < 0 ) tx += g.ntx; if ( tx >= g.ntx ) tx -= g.ntx; }
        if ( s.per_y ) { if ( ty < 0 ) ty += g.nty; if ( ty >= g.nty ) ty -= g.nty; }
        if ( tx < 0 || tx >= g.ntx || ty < 0 || ty >= g.nty ) return -1;
        return ty * g.ntx + tx;
    }
    int cross_dir( I2 c ) const {
        int xc = ( c.x >= g.nx ) - ( c.x < 0 );
        int yc = ( c.y >= g.ny ) - ( c.y < 0 );
        return ( yc + 1 ) * 3 + ( xc + 1 );
    }

    void tile_sort( Species<R> & s ) {
        Particles<R> & P = s.p;
        Particles<R> & T = s.tmp;
        const int nt = g.ntiles();
        std::fill( s.new_np.begin(), s.new_np.end(), 0 );
        // classification
        for ( int t = 0; t < nt; t++ ) {
            int * cnt = &s.npt[9 * t];
            for ( int d = 0; d < 9; d++ ) cnt[d] = 0;
            int nl = 0;
            const int o = P.off[t];
            for ( int i = 0; i < P.np[t]; i++ ) {
                int d = cross_dir( P.ix[o + i] );
                if ( d != 4 ) { cnt[d]++; s.leav[o + nl++] = i; }
            }
            s.nleav[t] = nl;
            cnt[4] = P.np[t] - nl;
        }
        for ( int t = 0; t < nt; t++ )
            for ( int d = 0; d < 9; d++ ) {
                int tg = dir_target( s, t, d );
                if ( tg >= 0 ) s.new_np[tg] += s.npt[9 * t + d];
            }
        int acc = 0;
        for ( int t = 0; t < nt; t++ ) { T.off[t] = acc; T.np[t] = 0; acc += s.new_np[t]; }
        if ( (size_t) acc > s.max_part ) throw std::runtime_error( "particle buffer overflow in tile_sort" );

        // staging (tile order fixes the staging order)
        for ( int t = 0; t < nt; t++ ) {
            const int o = P.off[t], no = T.off[t];
            const int n0 = s.npt[9 * t + 4];
            const int nnew = s.new_np[t];
            int nshift;
            if ( no >= o ) nshift = no - o;
            else { nshift = ( o + n0 ) - ( no + nnew ); if ( nshift < 0 ) nshift = 0; }
            if ( nshift > n0 ) nshift = n0;
            int slot[9];
            slot[4] = T.off[t] + T.np[t]; T.np[t] += nshift;
            for ( int d = 0; d < 9; d++ ) {
                if ( d == 4 ) continue;
                int tg = dir_target( s, t, d );
                if ( tg >= 0 ) { slot[d] = T.off[tg] + T.np[tg]; T.np[tg] += s.npt[9 * t + d]; }
                else slot[d] = -1;
            }
            // leavers out
            for ( int l = 0; l < s.nleav[t]; l++ ) {
                const int k = o + s.leav[o + l];
                I2 c = P.ix[k];
                int xc = ( c.x >= g.nx ) - ( c.x < 0 );
                int yc = ( c.y >= g.ny ) - ( c.y < 0 );
                int d = ( yc + 1 ) * 3 + ( xc + 1 );
                if ( slot[d] >= 0 ) {
                    const int dst = slot[d]++;
                    T.ix[dst] = { c.x - xc * g.nx, c.y - yc * g.ny };
                    T.x[dst] = P.x[k];
                    T.u[dst] = P.u[k];
                }
            }
            // holes below n0 <- surviving particles above n0, in order
            int src = n0;
            for ( int l = 0; l < s.nleav[t]; l++ ) {
                const int k = s.leav[o + l];
                if ( k >= n0 ) break;
                while ( cross_dir( P.ix[o + src] ) != 4 ) src++;
                P.ix[o + k] = P.ix[o + src]; P.x[o + k] = P.x[o + src]; P.u[o + k] = P.u[o + src];
                src++;
            }
            // shifted stayers
            const int first = ( no >= o ) ? 0 : n0 - nshift;
            for ( int i = 0; i < nshift; i++ ) {
                T.ix[slot[4] + i] = P.ix[o + first + i];
                T.x[slot[4] + i] = P.x[o + first + i];
                T.u[slot[4] + i] = P.u[o + first + i];
            }
            P.np[t] = n0 - nshift;
        }
        // staged particles in
        for ( int t = 0; t < nt; t++ ) {
            const int o = P.off[t], n = P.np[t], no = T.off[t], nin = T.np[t];
            auto put = [&]( int dst, int src ) { P.ix[dst] = T.ix[src]; P.x[dst] = T.x[src]; P.u[dst] = T.u[src]; };
            int front = 0;
            if ( no < o ) { front = o - no; if ( front > nin ) front = nin; }
            for ( int i = 0; i < front; i++ ) put( no + i, no + i );
            for ( int i = front; i < nin; i++ ) put( no + n + i, no + i );
            P.np[t] = n + nin;
            P.off[t] = no;
        }
    }

    // ---- Species::advance(emf, current) (species.cpp:2051-2068)
    void species_advance( Species<R> & s ) {
        push( s );
        move_deposit( s );
        s.iter++;
        tile_sort( s );
    }

    // ---- EMF::advance(Current&) (emf.cpp:513-580)
    void yee_step() {
        const V2<R> dtdx = ( sizeof( R ) == 8 ) ? V2<R>{ R( dt / dx.x ), R( dt / dx.y ) }
                                                : V2<R>{ R( (float) ( dt / (double) dx.x ) ), R( (float) ( dt / (double) dx.y ) ) };
        const V2<R> h{ dtdx.x / R( 2 ), dtdx.y / R( 2 ) };
        const R dtf = ( sizeof( R ) == 8 ) ? R( dt ) : R( (float) dt );
        const int nt = g.ntiles(), ys = g.ext_x;
        #pragma omp parallel for
        for ( int t = 0; t < nt; t++ ) {
            std::vector<V3<R>> El( E.tile( t ), E.tile( t ) + g.vol ), Bl( B.tile( t ), B.tile( t ) + g.vol );
            V3<R> * e = El.data() + g.off;
            V3<R> * b = Bl.data() + g.off;
            const V3<R> * j = J.tile( t ) + g.off;
            auto bstep = [&]( V2<R> c ) {
                for ( int iy = -1; iy < g.ny + 1; iy++ )
                    for ( int ix = -1; ix < g.nx + 1; ix++ ) {
                        const int k = ix + iy * ys;
                        b[k].x += ( -c.y * ( e[k + ys].z - e[k].z ) );
                        b[k].y += ( c.x * ( e[k + 1].z - e[k].z ) );
                        b[k].z += ( -c.x * ( e[k + 1].y - e[k].y ) + c.y * ( e[k + ys].x - e[k].x ) );
                    }
            };
            bstep( h );
            for ( int iy = 0; iy < g.ny + 2; iy++ )
                for ( int ix = 0; ix < g.nx + 2; ix++ ) {
                    const int k = ix + iy * ys;
                    e[k].x += ( dtdx.y * ( b[k].z - b[k - ys].z ) ) - dtf * j[k].x;
                    e[k].y += ( -dtdx.x * ( b[k].z - b[k - 1].z ) ) - dtf * j[k].y;
                    e[k].z += ( dtdx.x * ( b[k].y - b[k - 1].y ) - dtdx.y * ( b[k].x - b[k - ys].x ) ) - dtf * j[k].z;
                }
            bstep( h );
            std::copy( El.begin(), El.end(), E.tile( t ) );
            std::copy( Bl.begin(), Bl.end(), B.tile( t ) );
        }
        E.copy_to_gc();
        B.copy_to_gc();
    }

    // ---- Simulation::advance (simulation.h:90-107)
    void advance() {
        std::fill( J.d.begin(), J.d.end(), V3<R>{ 0, 0, 0 } );
        for ( auto & s : sp ) species_advance( s );
        J.add_from_gc();
        J.copy_to_gc();
        apply_filter<R>( J, filter );
        yee_step();
        iter++;
    }

    // ---- EMF::get_energy (emf.cpp:662-713): note the inner loop bound nx.y
    void field_energy( double out[6] ) const {
        for ( int k = 0; k < 6; k++ ) out[k] = 0;
        for ( int t = 0; t < g.ntiles(); t++ ) {
            const V3<R> * e = E.d.data() + (size_t) t * g.vol + g.off;
            const V3<R> * b = B.d.data() + (size_t) t * g.vol + g.off;
            double te[6] = { 0, 0, 0, 0, 0, 0 };
            for ( int iy = 0; iy < g.ny; iy++ )
                for ( int ix = 0; ix < g.ny; ix++ ) {
                    const V3<R> ev = e[iy * g.ext_x + ix], bv = b[iy * g.ext_x + ix];
                    te[0] += ev.x * ev.x; te[1] += ev.y * ev.y; te[2] += ev.z * ev.z;
                    te[3] += bv.x * bv.x; te[4] += bv.y * bv.y; te[5] += bv.z * bv.z;
                }
            for ( int k = 0; k < 6; k++ ) out[k] += te[k];
        }
    }

    // species.h:365-368
    double kinetic_energy( const Species<R> & s ) const {
        return s.energy * s.q * s.m_q * s.dx.x * s.dx.y;
    }

    // ---- charge deposit on a grid with gc {0,1} (species.cpp:2348-2410, 2508-2522)
    // Returns the tile-major buffer after add_from_gc.
    Grid<R> charge( const Species<R> & s ) const {
        Geom cg( g.ntx, g.nty, g.nx, g.ny, 0, 1 );
        Grid<R> rho( cg );
        const int ys = cg.ext_x;
        for ( int t = 0; t < g.ntiles(); t++ ) {
            std::vector<R> loc( cg.vol, R( 0 ) );
            R * c0 = loc.data() + cg.off;
            const int o = s.p.off[t], n = s.p.np[t];
            for ( int i = 0; i < n; i++ ) {
                const I2 c = s.p.ix[o + i];
                const V2<R> x = s.p.x[o + i];
                const int k = c.y * ys + c.x;
                const R s0x = R( 0.5f ) - x.x, s1x = R( 0.5f ) + x.x;
                const R s0y = R( 0.5f ) - x.y, s1y = R( 0.5f ) + x.y;
                c0[k] += s0y * s0x * s.q;
                c0[k + 1] += s0y * s1x * s.q;
                c0[k + ys] += s1y * s0x * s.q;
                c0[k + 1 + ys] += s1y * s1x * s.q;
            }
            R * dst = rho.tile( t );
            for ( int k = 0; k < cg.ext_x * cg.ext_y; k++ ) dst[k] += loc[k];
        }
        rho.add_from_gc();
        return rho;
    }
};

} // namespace oracle
