// Particle stages of the em2d step on gfx950: field interpolation + Boris
// push, split-segment charge-conserving current deposit, and the tile sort.
//
// One workgroup per tile.  The tile's E and B (with guard cells) are staged in
// LDS as six float planes.  Each wavefront accumulates its particles' current
// in its own three J planes in LDS with plain read-add-writes, made
// conflict-free by a per-round owner election (LDS float atomics run at about
// one lane per clock on this part); the waves' planes are summed and flushed
// once, with plain adds, into the tile's own region of the global J grid (each
// tile owns its guard cells, so the flush needs no global atomics).
//
// Arithmetic is the reference's, operation for operation: no contraction
// (built with -ffp-contract=off), fmaf exactly where the reference calls
// ops::fma (species.cpp:22-26, 117-178), correctly rounded division and
// square root.  Per-particle results are therefore bit-identical to the
// reference's; only the order of the J additions (and of the energy sum)
// differs in the fast mode.  The deterministic mode runs one wavefront per
// tile and serialises the deposit and the energy sum over lanes in particle
// order, which reproduces the reference's single-thread results exactly.
#include "common.h"
#include "grid_ops.h"

#include <algorithm>
#include <memory>
#include <vector>

namespace zpic {
namespace {

constexpr int WAVE = 64;
constexpr int FAST_BLOCK = 512;  // 8 wavefronts per tile in the fast mode
constexpr int MAX_SEG = 15;      // wave segments recorded per tile (slot 15 = chunk size)

struct PartView {
    int * np;
    int * off;
    int2 * ix;
    float2 * x;
    float3 * u;
};

struct SortView {
    int * idx;
    int * npt;
    int * new_np;
    int * seg;
    int * tmp_np;
    int * total;
    unsigned long long * nmove;
    int2 * stage_ix;  // order-free sort staging (fast mode), or null
    float2 * stage_x;
    float3 * stage_u;
};

struct Geo {
    int nx, ny, ext_x, ext_y, vol, off, ntx, nty;
};

// Number of set bits of m below this lane (v_mbcnt_lo / v_mbcnt_hi).
__device__ inline int count_below( unsigned long long m ) {
    return (int) __builtin_amdgcn_mbcnt_hi( (unsigned) ( m >> 32 ), __builtin_amdgcn_mbcnt_lo( (unsigned) m, 0u ) );
}

// ---- interpolate_fld (species.cpp:43-92) on LDS float planes
__device__ inline void interpolate( const float * __restrict__ Ex, const float * __restrict__ Ey,
                                    const float * __restrict__ Ez, const float * __restrict__ Bx,
                                    const float * __restrict__ By, const float * __restrict__ Bz, int ys,
                                    int2 c, float2 x, float3 & e, float3 & b ) {
    const int i = c.x, j = c.y;
    const float s0x = 0.5f - x.x, s1x = 0.5f + x.x;
    const float s0y = 0.5f - x.y, s1y = 0.5f + x.y;
    const int hx = x.x < 0, hy = x.y < 0;
    const int ih = i - hx, jh = j - hy;
    const float s0xh = ( 1 - hx ) - x.x, s1xh = hx + x.x;
    const float s0yh = ( 1 - hy ) - x.y, s1yh = hy + x.y;

    // row offsets by 24-bit multiplies (cell rows and tile widths are small)
    const int jr = __mul24( j, ys ), jhr = __mul24( jh, ys );
    const int k00 = ih + jr;   // (ih, j)
    const int kjh = i + jhr;   // (i, jh)
    const int kij = i + jr;    // (i, j)
    const int khh = ih + jhr;  // (ih, jh)

    e.x = ( Ex[k00] * s0xh + Ex[k00 + 1] * s1xh ) * s0y + ( Ex[k00 + ys] * s0xh + Ex[k00 + ys + 1] * s1xh ) * s1y;
    e.y = ( Ey[kjh] * s0x + Ey[kjh + 1] * s1x ) * s0yh + ( Ey[kjh + ys] * s0x + Ey[kjh + ys + 1] * s1x ) * s1yh;
    e.z = ( Ez[kij] * s0x + Ez[kij + 1] * s1x ) * s0y + ( Ez[kij + ys] * s0x + Ez[kij + ys + 1] * s1x ) * s1y;

    b.x = ( Bx[kjh] * s0x + Bx[kjh + 1] * s1x ) * s0yh + ( Bx[kjh + ys] * s0x + Bx[kjh + ys + 1] * s1x ) * s1yh;
    b.y = ( By[k00] * s0xh + By[k00 + 1] * s1xh ) * s0y + ( By[k00 + ys] * s0xh + By[k00 + ys + 1] * s1xh ) * s1y;
    b.z = ( Bz[khh] * s0xh + Bz[khh + 1] * s1xh ) * s0yh + ( Bz[khh + ys] * s0xh + Bz[khh + ys + 1] * s1xh ) * s1yh;
}

// Correctly rounded square root of x >= 1: the hardware square root moved by
// one ulp down or up from the signs of the two residuals x - s(s -+ ulp),
// which is the compiler's correctly rounded sequence without its rescaling
// of inputs below 2^-96 and its +-0 / inf / NaN passthrough, neither of which
// applies to x >= 1.
__device__ inline float sqrt_ge1( float x ) {
    const float s = __builtin_amdgcn_sqrtf( x );
    const float dn = __int_as_float( __float_as_int( s ) - 1 );
    const float up = __int_as_float( __float_as_int( s ) + 1 );
    const float lo = ( fmaf( -dn, s, x ) <= 0.0f ) ? dn : s;
    return ( fmaf( -up, s, x ) > 0.0f ) ? up : lo;
}

// Correctly rounded a / b for |b| >= 1 and normal a / b: the hardware
// reciprocal refined once, then the quotient corrected twice with fused
// residuals, which is the compiler's division sequence when its operand
// scaling and special-case fixup leave the values unchanged (they only act
// on denormal, huge or special operands and results).
__device__ inline float div_ge1( float a, float b ) {
    float r = __builtin_amdgcn_rcpf( b );
    r = fmaf( fmaf( -b, r, 1.0f ), r, r );
    float q = a * r;
    q = fmaf( fmaf( -b, q, a ), r, q );
    return fmaf( fmaf( -b, q, a ), r, q );
}

// ---- dudt_boris (species.cpp:117-178); returns u and the energy term.
// EXACT_ENE: the energy term |u|^2 / (gamma + 1) by the general division, as
// the reference (the deterministic mode); otherwise by div_ge1, correctly
// rounded except for a subnormal quotient (|u| below 1e-19, where the general
// division's rescaling matters), in 6 instructions instead of 13.
template <bool EXACT_ENE>
__device__ inline float3 boris( float alpha, float3 e, float3 b, float3 u, float & ene ) {
    e.x *= alpha; e.y *= alpha; e.z *= alpha;
    float3 ut = make_float3( u.x + e.x, u.y + e.y, u.z + e.z );
    {
        const float usq = fmaf( ut.z, ut.z, fmaf( ut.y, ut.y, ut.x * ut.x ) );
        const float g = sqrt_ge1( 1.0f + usq );
        if constexpr ( EXACT_ENE ) ene = usq / ( g + 1.0f );  // usq may be 0 or tiny: general division
        else ene = div_ge1( usq, g + 1.0f );
        const float ag = div_ge1( alpha, g );
        b.x *= ag; b.y *= ag; b.z *= ag;
    }
    float3 v;
    v.x = fmaf( b.z, ut.y, ut.x );
    v.y = fmaf( b.x, ut.z, ut.y );
    v.z = fmaf( b.y, ut.x, ut.z );
    v.x = fmaf( -b.y, ut.z, v.x );
    v.y = fmaf( -b.z, ut.x, v.y );
    v.z = fmaf( -b.x, ut.y, v.z );
    {
        const float s = div_ge1( 2.0f, fmaf( b.z, b.z, fmaf( b.y, b.y, fmaf( b.x, b.x, 1.0f ) ) ) );
        b.x *= s; b.y *= s; b.z *= s;
    }
    ut.x = fmaf( b.z, v.y, ut.x );
    ut.y = fmaf( b.x, v.z, ut.y );
    ut.z = fmaf( b.y, v.x, ut.z );
    ut.x = fmaf( -b.y, v.z, ut.x );
    ut.y = fmaf( -b.z, v.x, ut.y );
    ut.z = fmaf( -b.x, v.y, ut.z );
    ut.x += e.x; ut.y += e.y; ut.z += e.z;
    return ut;
}

// ---- rgamma (species.cpp:22-26): true division by a rounded sqrt, not rsqrt
__device__ inline float inv_gamma( float3 u ) {
    return div_ge1( 1.0f, sqrt_ge1( fmaf( u.z, u.z, fmaf( u.y, u.y, fmaf( u.x, u.x, 1.0f ) ) ) ) );
}

struct Seg {
    int2 c;
    float2 a, b;
    float qvz;
};

// ---- split2d (species.cpp:336-430): up to three straight segments, one per cell
__device__ inline int split_path( int2 c0, float2 x0, float2 x1, float2 d, float qvz, int2 & di, Seg s[3] ) {
    di.x = ( x1.x >= 0.5f ) - ( x1.x < -0.5f );
    di.y = ( x1.y >= 0.5f ) - ( x1.y < -0.5f );
    const bool cx = di.x != 0, cy = di.y != 0;
    s[0].c = c0; s[0].a = x0; s[0].b = x1; s[0].qvz = qvz;
    if ( !cx && !cy ) return 1;
    if ( !cx ) {
        const float yb = 0.5f * di.y;
        const float f = ( yb - x0.y ) / d.y;
        const float xb = x0.x + d.x * f;
        s[0].b = make_float2( xb, yb );
        s[0].qvz = qvz * f;
        s[1].c = make_int2( c0.x, c0.y + di.y );
        s[1].a = make_float2( xb, -yb );
        s[1].b = make_float2( x1.x, x1.y - di.y );
        s[1].qvz = qvz * ( 1 - f );
        return 2;
    }
    const float xb = 0.5f * di.x;
    float f = ( xb - x0.x ) / d.x;
    const float yb = x0.y + d.y * f;
    s[0].b = make_float2( xb, yb );
    s[0].qvz = qvz * f;
    s[1].c = make_int2( c0.x + di.x, c0.y );
    s[1].a = make_float2( -xb, yb );
    s[1].b = make_float2( x1.x - di.x, x1.y );
    s[1].qvz = qvz * ( 1 - f );
    if ( !cy ) return 2;
    const float yb2 = 0.5f * di.y;
    if ( yb >= -0.5f && yb < 0.5f ) {
        // y crossing on the second segment
        f = ( yb2 - yb ) / ( x1.y - yb );
        const float xb2 = -xb + ( x1.x - xb ) * f;
        s[2].c = make_int2( s[1].c.x, s[1].c.y + di.y );
        s[2].a = make_float2( xb2, -yb2 );
        s[2].b = make_float2( s[1].b.x, s[1].b.y - di.y );
        s[2].qvz = s[1].qvz * ( 1 - f );
        s[1].b = make_float2( xb2, yb2 );
        s[1].qvz *= f;
    } else {
        // y crossing on the first segment
        f = ( yb2 - x0.y ) / ( yb - x0.y );
        const float xb2 = x0.x + ( xb - x0.x ) * f;
        s[2].c = make_int2( s[0].c.x, s[0].c.y + di.y );
        s[2].a = make_float2( xb2, -yb2 );
        s[2].b = make_float2( s[0].b.x, s[0].b.y - di.y );
        s[2].qvz = s[0].qvz * ( 1 - f );
        s[0].b = make_float2( xb2, yb2 );
        s[0].qvz *= f;
        s[1].c.y += di.y;
        s[1].a.y -= di.y;
        s[1].b.y -= di.y;
    }
    return 3;
}

// ---- dep_current_seg (species.cpp:267-310): the 8 current contributions of
// one segment, in the reference's order: Jx(k), Jy(k), Jz(k), Jy(k+1),
// Jz(k+1), Jx(k+ys), Jz(k+ys), Jz(k+ys+1) with k the segment's cell.
__device__ inline void segment_current( const Seg & s, float qnx, float qny, float v[8] ) {
    const float A0x = 0.5f - s.a.x, A1x = 0.5f + s.a.x, B0x = 0.5f - s.b.x, B1x = 0.5f + s.b.x;
    const float A0y = 0.5f - s.a.y, A1y = 0.5f + s.a.y, B0y = 0.5f - s.b.y, B1y = 0.5f + s.b.y;
    const float jx = qnx * ( s.b.x - s.a.x );
    const float jy = qny * ( s.b.y - s.a.y );
    const float wy0 = 0.5f * ( A0y + B0y ), wy1 = 0.5f * ( A1y + B1y );
    const float wx0 = 0.5f * ( A0x + B0x ), wx1 = 0.5f * ( A1x + B1x );
    const float q = s.qvz;
    v[0] = jx * wy0;
    v[1] = jy * wx0;
    v[2] = q * ( A0x * A0y + B0x * B0y + ( A0x * B0y - B0x * A0y ) / 2.0f );
    v[3] = jy * wx1;
    v[4] = q * ( A1x * A0y + B1x * B0y + ( A1x * B0y - B1x * A0y ) / 2.0f );
    v[5] = jx * wy1;
    v[6] = q * ( A0x * A1y + B0x * B1y + ( A0x * B1y - B0x * A1y ) / 2.0f );
    v[7] = q * ( A1x * A1y + B1x * B1y + ( A1x * B1y - B1x * A1y ) / 2.0f );
}

// The same eight contributions with fused multiply-adds, for the fast mode
// (whose J is compared to a tolerance: the deposit order differs anyway):
// wy0 = (A0y + B0y) / 2 = 1/2 - (a.y + b.y) / 2, wy1 = 1/2 + (a.y + b.y) / 2,
// and Jz = q (Ax (Ay + By / 2) + Bx (By - Ay / 2)), the reference's
// Ax Ay + Bx By + (Ax By - Bx Ay) / 2 regrouped: 40 instead of 58 operations.
__device__ inline void segment_current_fma( const Seg & s, float qnx, float qny, float v[8] ) {
    const float A0x = 0.5f - s.a.x, A1x = 0.5f + s.a.x, B0x = 0.5f - s.b.x, B1x = 0.5f + s.b.x;
    const float A0y = 0.5f - s.a.y, A1y = 0.5f + s.a.y, B0y = 0.5f - s.b.y, B1y = 0.5f + s.b.y;
    const float jx = qnx * ( s.b.x - s.a.x );
    const float jy = qny * ( s.b.y - s.a.y );
    const float mx = 0.5f * ( s.a.x + s.b.x ), my = 0.5f * ( s.a.y + s.b.y );
    v[0] = jx * ( 0.5f - my );
    v[5] = jx * ( 0.5f + my );
    v[1] = jy * ( 0.5f - mx );
    v[3] = jy * ( 0.5f + mx );
    const float q = s.qvz;
    const float p0 = fmaf( 0.5f, B0y, A0y ), r0 = fmaf( -0.5f, A0y, B0y );
    const float p1 = fmaf( 0.5f, B1y, A1y ), r1 = fmaf( -0.5f, A1y, B1y );
    v[2] = q * fmaf( A0x, p0, B0x * r0 );
    v[4] = q * fmaf( A1x, p0, B1x * r0 );
    v[6] = q * fmaf( A0x, p1, B0x * r1 );
    v[7] = q * fmaf( A1x, p1, B1x * r1 );
}

// LDS (address space 3) element types: volatile accesses through generic
// pointers would be lowered to flat instructions.
typedef __attribute__( ( address_space( 3 ) ) ) float lds_f32;
typedef __attribute__( ( address_space( 3 ) ) ) unsigned char lds_u8;

// Read-add-write of the 8 contributions.  The accesses are volatile so that
// they are issued in program order: another lane's contribution to the same
// element through a different slot (Jx(k+ys) of one cell is Jx(k) of the cell
// above) then always sees the previous write (LDS executes one wave's
// instructions in order).
__device__ inline void add_current( volatile lds_f32 * Jx, volatile lds_f32 * Jy, volatile lds_f32 * Jz, int k,
                                    int ys, const float v[8] ) {
    Jx[k] = Jx[k] + v[0];
    Jy[k] = Jy[k] + v[1];
    Jz[k] = Jz[k] + v[2];
    Jy[k + 1] = Jy[k + 1] + v[3];
    Jz[k + 1] = Jz[k + 1] + v[4];
    Jx[k + ys] = Jx[k + ys] + v[5];
    Jz[k + ys] = Jz[k + ys] + v[6];
    Jz[k + ys + 1] = Jz[k + ys + 1] + v[7];
}

// Read-add-write of the 8 contributions of lanes that own distinct cells, in
// four batches by node offset (0, +1, +ys, +ys+1): inside a batch the lanes'
// addresses are distinct, so each batch issues its reads together; a later
// batch may touch what an earlier one wrote (Jx(k+ys) of one cell is Jx(k)
// of the cell above) and, the accesses being volatile and LDS in order per
// wave, sees those writes.
__device__ inline void add_current_batched( volatile lds_f32 * Jx, volatile lds_f32 * Jy, volatile lds_f32 * Jz,
                                            int k, int ys, const float v[8] ) {
    {
        const float a = Jx[k], b = Jy[k], c = Jz[k];
        Jx[k] = a + v[0];
        Jy[k] = b + v[1];
        Jz[k] = c + v[2];
    }
    {
        const float a = Jy[k + 1], b = Jz[k + 1];
        Jy[k + 1] = a + v[3];
        Jz[k + 1] = b + v[4];
    }
    {
        const float a = Jx[k + ys], b = Jz[k + ys];
        Jx[k + ys] = a + v[5];
        Jz[k + ys] = b + v[6];
    }
    Jz[k + ys + 1] = Jz[k + ys + 1] + v[7];
}

// Fast-mode deposit of one segment per pending lane into the wave's private J
// planes, without atomics: each round, every pending lane writes its lane id
// into the tag byte of its cell and reads it back; the lane whose id survived
// owns the cell for this round, so the owners of a round hold distinct cells
// and their read-add-writes cannot collide.  Rounds repeat while lanes are
// pending (as many rounds as the most repeated cell in the wave).
__device__ inline void deposit_elected( bool pend, const Seg & s, float qnx, float qny, volatile lds_f32 * Jx,
                                        volatile lds_f32 * Jy, volatile lds_f32 * Jz, volatile lds_u8 * tag, int ys,
                                        unsigned lane ) {
    // computed on every lane (SIMD lanes cost the same masked or not); used
    // only by pending ones
    float v[8];
    segment_current_fma( s, qnx, qny, v );
    const int k = s.c.x + __mul24( ys, s.c.y );
    while ( __ballot( pend ) ) {
        if ( pend ) tag[k] = (unsigned char) lane;
        if ( pend && tag[k] == (unsigned char) lane ) {
            add_current_batched( Jx, Jy, Jz, k, ys, v );
            pend = false;
        }
    }
}

struct MoveArgs {
    float dtdx, dtdy, q, qnx, qny;
    int refl;      // reflecting walls: 1 x lower, 2 x upper, 4 y lower, 8 y upper
    int shx, shy;  // cell shift added to every moved particle (moving window)
};

// Species::process_bc (species.cpp:1854-1989) for one moved particle of
// tile (tx, ty): a reflecting wall mirrors the cell index, the in-cell
// position and the momentum component.  As in the reference's loops over
// tiles {0, ntiles - 1}, a single tile along an axis only sees its lower wall.
__device__ inline bool reflect_walls( int refl, int tx, int ty, int ntx, int nty, int nx, int ny, int2 & c,
                                      float2 & x, float3 & u ) {
    bool r = false;
    if ( tx == 0 ) {
        if ( ( refl & 1 ) && c.x < 0 ) { c.x += 1; x.x = -x.x; u.x = -u.x; r = true; }
    } else if ( tx == ntx - 1 ) {
        if ( ( refl & 2 ) && c.x >= nx ) { c.x -= 1; x.x = -x.x; u.x = -u.x; r = true; }
    }
    if ( ty == 0 ) {
        if ( ( refl & 4 ) && c.y < 0 ) { c.y += 1; x.y = -x.y; u.y = -u.y; r = true; }
    } else if ( ty == nty - 1 ) {
        if ( ( refl & 8 ) && c.y >= ny ) { c.y -= 1; x.y = -x.y; u.y = -u.y; r = true; }
    }
    return r;
}

// ---------------------------------------------------------------------------
// Particle pass.  PUSH: interpolation + Boris.  MOVE: position advance,
// deposit, cell update and leaver classification.  DET: one wavefront, lane
// serialised deposit and energy sum (reference order).
//
// Wave w of W handles the contiguous chunk [w*chunk, (w+1)*chunk) of the
// tile's particles.  With the reference order the leaver indices it appends
// to the tile's list (idx[offset + w*chunk ...]) are in increasing order;
// with the order-free sort it compacts its stayers to the chunk's front and
// stages its leavers at the chunk's first positions.  seg[16 t + w] keeps
// the count of each wave's leavers, seg[16 t + 15] the chunk.
//
// LDS: E and B planes (6*vol floats, PUSH), then W private J plane triples
// (3*vol floats each, MOVE), then W tag byte arrays (vol bytes each, MOVE),
// then W crossing queues (QCAP entries of QWORDS words, MOVE), then W
// leaver buffers (LBCAP records of LBWORDS words, order-free sort, lbuf).
//
// Fast mode: a particle that stays in its cell deposits its single segment
// in the iteration that moves it; one that crosses a cell edge (about one in
// ten) is queued with its path, and the wave splits and deposits 64 queued
// paths at a time with every lane busy, so the rarely needed split and the
// second and third segments do not occupy the whole wave every iteration.
constexpr int QCAP = WAVE;  // crossing-queue entries per wave
constexpr int QWORDS = 6;   // cell, x0.x, x0.y, d.x, d.y, qvz
constexpr int PREFETCH = 2;  // iterations of 64 particles loaded ahead
// Order-free sort: where the LDS allows it without costing a workgroup per
// CU, a wave buffers the leavers it finds and writes them to the staging
// arrays LBCAP at a time, LBWORDS words each (cell x, y, position x, y,
// momentum x, y, z): whole cache lines instead of a few records per
// iteration, whose lines leave the L2 partly written.
constexpr int LBCAP = 32;
constexpr int LBWORDS = 7;

// LDS of the particle pass: E and B planes (PUSH), W private J triples, W
// tag arrays and W crossing queues (MOVE), W leaver buffers (lbuf).
__host__ __device__ inline size_t particles_lds_bytes( int vol, int W, bool push, bool move, bool lbuf = false ) {
    return sizeof( float ) * (size_t) vol * ( ( push ? 6 : 0 ) + ( move ? 3 * W : 0 ) ) +
           ( move ? ( ( (size_t) W * vol + 3 ) & ~(size_t) 3 ) + sizeof( float ) * (size_t) W * QCAP * QWORDS : 0 ) +
           ( move && lbuf ? sizeof( float ) * (size_t) W * LBCAP * LBWORDS : 0 );
}

// Arguments of one instance of the particle pass.
struct PartArgs {
    PartView P;
    Geo g;
    const float3 * E;
    const float3 * B;
    float3 * J;
    float alpha;
    MoveArgs m;
    double * tile_energy;
    SortView S;
    int waves;  // waves of the workgroup that take particles (0: all)
    int lbuf;   // order-free sort: leavers buffered in LDS (LDS sized for it)
    int jset;   // fast mode: the tile's J region is written, not added to
};

struct PartG {
    gptr<int> np;
    gptr<int> off;
    gptr<int2> ix;
    gptr<float2> x;
    gptr<float3> u;
};
__device__ inline PartG gview( const PartView & v ) { return PartG{ gbl( v.np ), gbl( v.off ), gbl( v.ix ), gbl( v.x ), gbl( v.u ) }; }

struct SortG {
    gptr<int> idx;
    gptr<int> npt;
    gptr<int> new_np;
    gptr<int> seg;
    gptr<int> tmp_np;
    gptr<int> total;
    gptr<unsigned long long> nmove;
};
__device__ inline SortG gview( const SortView & v ) {
    return SortG{ gbl( v.idx ), gbl( v.npt ), gbl( v.new_np ), gbl( v.seg ), gbl( v.tmp_np ), gbl( v.total ),
                  gbl( v.nmove ) };
}

// float3 records element by element: an aggregate copy of a float3 between
// global arrays is otherwise staged in a private temporary, which the
// compiler places in LDS (12 bytes per thread) or scratch.
__device__ inline float3 ld3( gptr<const float3> p ) {
    const gptr<const float> f = (gptr<const float>) p;
    return make_float3( f[0], f[1], f[2] );
}
__device__ inline float3 ld3( gptr<float3> p ) { return ld3( (gptr<const float3>) p ); }
__device__ inline void st3( gptr<float3> p, const float3 v ) {
    const gptr<float> f = (gptr<float>) p;
    f[0] = v.x;
    f[1] = v.y;
    f[2] = v.z;
}

// Element i of one tile's particle array through the tile's (uniform) base
// address and a 32-bit byte offset: the global load / store form with a
// scalar base and a vector offset, no 64-bit address arithmetic per lane.
template <class T>
struct TileArray {
    gptr<char> base;
    __device__ T get( unsigned i ) const { return load( i * (unsigned) sizeof( T ) ); }
    __device__ void set( unsigned i, const T & v ) const { store( i * (unsigned) sizeof( T ), v ); }
    // by byte offset
    __device__ T load( unsigned b ) const { return *(gptr<T>) ( base + b ); }
    __device__ void store( unsigned b, const T & v ) const { *(gptr<T>) ( base + b ) = v; }
};

// Single-instance launches pass map == nullptr and use `a` with tile =
// blockIdx.x; batched launches pass a table of per-instance arguments and a
// map from block to (instance, tile).
template <class A>
__device__ inline int select_instance( A & a, const A * tab, const int2 * map ) {
    if ( map ) {
        const int2 r = gbl( map )[blockIdx.x];
        a = gbl( tab )[r.x];
        return r.y;
    }
    return blockIdx.x;
}

// The same for kernels that take one item per wave: item g of the launch.
template <class A>
__device__ inline int select_item( A & a, const A * tab, const int2 * map, int g ) {
    if ( map ) {
        const int2 r = gbl( map )[g];
        a = gbl( tab )[r.x];
        return r.y;
    }
    return g;
}

// The particle pass of tile t (one workgroup).  `smem` is the workgroup's
// LDS of particles_lds_bytes( vol, W ) bytes, W = the waves that take
// particles: all of the workgroup's, or a.waves if that is smaller (the
// other waves then only join the workgroup barriers).
template <bool PUSH, bool MOVE, bool DET>
__device__ inline void particles_tile( const PartArgs & a, const int t, float * smem ) {
    const PartView P = a.P;
    const Geo g = a.g;
    const float3 * __restrict__ E = a.E;
    const float3 * __restrict__ B = a.B;
    float3 * __restrict__ J = a.J;
    const float alpha = a.alpha;
    const MoveArgs m = a.m;
    double * __restrict__ tile_energy = a.tile_energy;
    const SortView S = a.S;
    const int tid = threadIdx.x;
    const int nthr = blockDim.x;
    const int vol = g.vol;
    const int ys = g.ext_x;
    const int W = ( a.waves > 0 && a.waves < nthr / WAVE ) ? a.waves : nthr / WAVE;
    const int w = tid / WAVE;
    const int lane = tid % WAVE;

    float * Ex = smem;
    float * Ey = Ex + vol;
    float * Ez = Ey + vol;
    float * Bx = Ez + vol;
    float * By = Bx + vol;
    float * Bz = By + vol;
    float * Jall = PUSH ? Bz + vol : smem;        // W triples of planes
    float * Jx = Jall + 3 * vol * w;              // this wave's planes
    float * Jy = Jx + vol;
    float * Jz = Jy + vol;
    unsigned char * tags = reinterpret_cast<unsigned char *>( Jall + 3 * vol * W ) + (size_t) vol * w;
    float * const qbase = reinterpret_cast<float *>( reinterpret_cast<unsigned char *>( Jall + 3 * vol * W ) +
                                                     ( ( (size_t) W * vol + 3 ) & ~(size_t) 3 ) );
    float * queue = qbase + (size_t) w * QCAP * QWORDS;
    float * lbuf = qbase + (size_t) W * QCAP * QWORDS + (size_t) w * LBCAP * LBWORDS;  // if a.lbuf
    __shared__ int s_cnt[9];
    __shared__ double s_en[FAST_BLOCK / WAVE];
    __shared__ int s_nl[FAST_BLOCK / WAVE];

    if constexpr ( PUSH ) {
        const gptr<const float> Et = gbl( reinterpret_cast<const float *>( E + (size_t) t * vol ) );
        const gptr<const float> Bt = gbl( reinterpret_cast<const float *>( B + (size_t) t * vol ) );
        for ( int k = tid; k < 3 * vol; k += nthr ) {
            const int el = k / 3, c = k - 3 * el;
            Ex[c * vol + el] = Et[k];
            Bx[c * vol + el] = Bt[k];
        }
    }
    if constexpr ( MOVE ) {
        for ( int k = tid; k < 3 * vol * W; k += nthr ) Jall[k] = 0.0f;
        if ( tid < 9 ) s_cnt[tid] = 0;
    }
    __syncthreads();

    const int off = gbl( P.off )[t];
    const int np = gbl( P.np )[t];
    const int chunk = ( ( np + W - 1 ) / W + WAVE - 1 ) & ~( WAVE - 1 );
    const int beg = w * chunk;
    const int end = min( np, beg + chunk );

    // tile base pointers (uniform) + 32-bit particle index
    const TileArray<int2> pix{ (gptr<char>) gbl( P.ix + off ) };
    const TileArray<float2> px{ (gptr<char>) gbl( P.x + off ) };
    const TileArray<float3> pu{ (gptr<char>) gbl( P.u + off ) };
    const gptr<int> sidx = gbl( S.idx + off );
    // Order-free sort (fast mode): each wave compacts its stayers to the front
    // of its chunk and writes its leavers, cell still in this tile's frame, to
    // the staging arrays at the same chunk positions; no leaver list.
    const bool stage = MOVE && !DET && S.stage_ix != nullptr;
    const TileArray<int2> six{ (gptr<char>) gbl( S.stage_ix + off ) };
    const TileArray<float2> sx{ (gptr<char>) gbl( S.stage_x + off ) };
    const TileArray<float3> su{ (gptr<char>) gbl( S.stage_u + off ) };

    const float * e0x = Ex + g.off; const float * e0y = Ey + g.off; const float * e0z = Ez + g.off;
    const float * b0x = Bx + g.off; const float * b0y = By + g.off; const float * b0z = Bz + g.off;
    volatile lds_f32 * j0x = (volatile lds_f32 *) ( Jx + g.off );
    volatile lds_f32 * j0y = (volatile lds_f32 *) ( Jy + g.off );
    volatile lds_f32 * j0z = (volatile lds_f32 *) ( Jz + g.off );
    volatile lds_u8 * tag0 = (volatile lds_u8 *) ( tags + g.off );

    double en = 0.0;
    int nl = 0;
    int qn = 0;  // queued crossing paths of this wave
    volatile lds_f32 * q = (volatile lds_f32 *) queue;
    int nst = 0;  // order-free sort: stayers compacted so far
    const bool lbon = stage && a.lbuf;
    int lbn = 0;  // leavers in the LDS buffer
    int lbw = 0;  // buffered leavers written to the staging arrays
    volatile lds_f32 * lb = (volatile lds_f32 *) lbuf;

    // the first `count` buffered leavers to the staging arrays at chunk
    // positions beg + lbw, ...
    auto flush_leavers = [&]( int count ) {
        if ( lane < count ) {
            const unsigned k = (unsigned) ( beg + lbw + lane );
            six.set( k, make_int2( __float_as_int( lb[lane] ), __float_as_int( lb[LBCAP + lane] ) ) );
            sx.set( k, make_float2( lb[2 * LBCAP + lane], lb[3 * LBCAP + lane] ) );
            su.set( k, make_float3( lb[4 * LBCAP + lane], lb[5 * LBCAP + lane], lb[6 * LBCAP + lane] ) );
        }
        lbw += count;
    };

    // split + deposit of the first `count` queued paths, one per lane
    auto flush_queue = [&]( int count ) {
        Seg sg[3] = {};
        int ns = 0;
        if ( lane < count ) {
            const int cp = __float_as_int( q[lane] );
            const int2 c0 = make_int2( cp & 0xffff, cp >> 16 );
            const float2 x0 = make_float2( q[QCAP + lane], q[2 * QCAP + lane] );
            const float2 d = make_float2( q[3 * QCAP + lane], q[4 * QCAP + lane] );
            const float qvz = q[5 * QCAP + lane];
            const float2 x1 = make_float2( x0.x + d.x, x0.y + d.y );
            int2 di;
            ns = split_path( c0, x0, x1, d, qvz, di, sg );
        }
        deposit_elected( ns > 0, sg[0], m.qnx, m.qny, j0x, j0y, j0z, tag0, ys, lane );
        if ( __ballot( ns > 1 ) ) deposit_elected( ns > 1, sg[1], m.qnx, m.qny, j0x, j0y, j0z, tag0, ys, lane );
        if ( __ballot( ns > 2 ) ) deposit_elected( ns > 2, sg[2], m.qnx, m.qny, j0x, j0y, j0z, tag0, ys, lane );
    };

    // The particles of the next PREFETCH iterations are loaded while the
    // current ones are processed (more bytes in flight per wave: the pass is
    // bound by memory latency at the occupancy the LDS allows).  Byte offsets
    // of the lane's particle in the 8-byte (ix, x) and 12-byte (u) arrays are
    // carried across iterations; prefetch offsets are clamped to the chunk's
    // last particle, so that every lane loads (lanes past the end load data
    // that they never store or deposit).
    int2 cq[PREFETCH] = {};
    float2 xq[PREFETCH] = {};
    float3 uq[PREFETCH] = {};
    unsigned o8 = (unsigned) ( beg + lane ) * 8u, o12 = (unsigned) ( beg + lane ) * 12u;
    const unsigned last8 = (unsigned) ( end - 1 ) * 8u, last12 = (unsigned) ( end - 1 ) * 12u;
    if ( beg < end ) {
#pragma unroll
        for ( int d = 0; d < PREFETCH; d++ ) {
            const unsigned n8 = min( o8 + 8u * WAVE * d, last8 ), n12 = min( o12 + 12u * WAVE * d, last12 );
            cq[d] = pix.load( n8 );
            xq[d] = px.load( n8 );
            uq[d] = pu.load( n12 );
        }
    }

    for ( int i0 = beg; i0 < end; i0 += WAVE, o8 += 8u * WAVE, o12 += 12u * WAVE ) {
        const int i = i0 + lane;
        const bool act = i < end;
        int2 c = cq[0];
        float2 x = xq[0];
        float3 u = uq[0];
#pragma unroll
        for ( int d = 0; d + 1 < PREFETCH; d++ ) {
            cq[d] = cq[d + 1];
            xq[d] = xq[d + 1];
            uq[d] = uq[d + 1];
        }
        {
            const unsigned n8 = min( o8 + 8u * WAVE * PREFETCH, last8 );
            const unsigned n12 = min( o12 + 12u * WAVE * PREFETCH, last12 );
            cq[PREFETCH - 1] = pix.load( n8 );
            xq[PREFETCH - 1] = px.load( n8 );
            uq[PREFETCH - 1] = pu.load( n12 );
        }
        if constexpr ( PUSH ) {
            float ene = 0.0f;
            if ( act ) {
                float3 e, b;
                interpolate( e0x, e0y, e0z, b0x, b0y, b0z, ys, c, x, e, b );
                u = boris<DET>( alpha, e, b, u, ene );
                if ( !stage ) pu.store( o12, u );  // order-free: written where the particle goes
            }
            if constexpr ( DET ) {
                // energy += term, in particle order
                for ( int l = 0; l < WAVE; l++ ) {
                    const float v = __shfl( ene, l );
                    if ( i0 + l < end ) en += v;
                }
            } else {
                en += ene;
            }
        }
        if constexpr ( MOVE && !DET ) {
            // computed on every lane (lanes past the end hold some particle's
            // data); stored and deposited by active lanes only
            const float rg = inv_gamma( u );
            const float2 d = make_float2( m.dtdx * rg * u.x, m.dtdy * rg * u.y );
            Seg s0;
            s0.c = c;
            s0.a = x;
            s0.b = make_float2( x.x + d.x, x.y + d.y );
            s0.qvz = m.q * u.z * rg * 0.5f;
            // deltai (species.cpp:342-343)
            const int2 di = make_int2( ( s0.b.x >= 0.5f ) - ( s0.b.x < -0.5f ), ( s0.b.y >= 0.5f ) - ( s0.b.y < -0.5f ) );
            const bool cross = act && ( di.x != 0 || di.y != 0 );
            int2 c1 = make_int2( c.x + di.x + m.shx, c.y + di.y + m.shy );
            float2 x1 = make_float2( s0.b.x - di.x, s0.b.y - di.y );
            if ( act ) {
                const bool refl =
                    m.refl && reflect_walls( m.refl, t % g.ntx, t / g.ntx, g.ntx, g.nty, g.nx, g.ny, c1, x1, u );
                if ( !stage ) {
                    if ( refl ) pu.store( o12, u );
                    px.store( o8, x1 );
                    pix.store( o8, c1 );
                }
            }
            const int xc = ( c1.x >= g.nx ) - ( c1.x < 0 );
            const int yc = ( c1.y >= g.ny ) - ( c1.y < 0 );
            const int dir = act ? ( yc + 1 ) * 3 + ( xc + 1 ) : 4;
            // the stores go out before the deposit: the iteration ends by
            // waiting for the wave's outstanding memory operations (one
            // counter tracks loads and stores), and the deposit hides them
            const bool leave = act && dir != 4;
            const unsigned long long mask = __ballot( leave );
            if ( leave ) atomicAdd( &s_cnt[dir], 1 );
            if ( !stage ) {
                if ( leave ) sidx[beg + nl + count_below( mask )] = ( i << 4 ) | dir;
            } else {
                // leavers: to the staging arrays at chunk positions beg + nl, ...
                if ( !lbon ) {
                    if ( leave ) {
                        const unsigned k = (unsigned) ( beg + nl + count_below( mask ) );
                        six.set( k, c1 );
                        sx.set( k, x1 );
                        su.set( k, u );
                    }
                } else if ( mask ) {
                    // through the LDS buffer, flushed whenever full
                    const int nlm = __popcll( mask );
                    const int pos = lbn + count_below( mask );
                    int base = 0;
                    for ( ;; ) {
                        if ( leave && pos >= base && pos < base + LBCAP ) {
                            const int k = pos - base;
                            lb[k] = __int_as_float( c1.x );
                            lb[LBCAP + k] = __int_as_float( c1.y );
                            lb[2 * LBCAP + k] = x1.x;
                            lb[3 * LBCAP + k] = x1.y;
                            lb[4 * LBCAP + k] = u.x;
                            lb[5 * LBCAP + k] = u.y;
                            lb[6 * LBCAP + k] = u.z;
                        }
                        if ( lbn + nlm - base < LBCAP ) break;
                        flush_leavers( LBCAP );
                        base += LBCAP;
                    }
                    lbn += nlm - base;
                }
                // stayers: compacted to the front of the chunk (positions at
                // or below the ones already read)
                const bool stay = act && dir == 4;
                const unsigned long long sm = __ballot( stay );
                if ( stay ) {
                    const unsigned r = (unsigned) ( beg + nst + count_below( sm ) );
                    pix.set( r, c1 );
                    px.set( r, x1 );
                    pu.set( r, u );
                }
                nst += __popcll( sm );
            }
            nl += __popcll( mask );
            // paths inside one cell: their single segment now
            deposit_elected( act && !cross, s0, m.qnx, m.qny, j0x, j0y, j0z, tag0, ys, lane );
            // paths that cross: queued, split and deposited 64 at a time
            const unsigned long long cm = __ballot( cross );
            if ( cm ) {
                const int nc = __popcll( cm );
                const int pos = qn + count_below( cm );
                auto put = [&]( int k ) {
                    q[k] = __int_as_float( ( s0.c.y << 16 ) | s0.c.x );
                    q[QCAP + k] = s0.a.x;
                    q[2 * QCAP + k] = s0.a.y;
                    q[3 * QCAP + k] = d.x;
                    q[4 * QCAP + k] = d.y;
                    q[5 * QCAP + k] = s0.qvz;
                };
                // positions [base, base + QCAP) of the unbounded queue are
                // its slots; a full queue is flushed and the window moves on
                int base = 0;
                for ( ;; ) {
                    if ( cross && pos >= base && pos < base + QCAP ) put( pos - base );
                    if ( qn + nc - base < QCAP ) break;
                    flush_queue( QCAP );
                    base += QCAP;
                }
                qn += nc - base;
            }
        }
        if constexpr ( MOVE && DET ) {
            int dir = 4;
            int ns = 0;
            Seg sg[3];
            if ( act ) {
                const float rg = inv_gamma( u );
                const float2 d = make_float2( m.dtdx * rg * u.x, m.dtdy * rg * u.y );
                const float qvz = m.q * u.z * rg * 0.5f;
                float2 x1 = make_float2( x.x + d.x, x.y + d.y );
                int2 di;
                ns = split_path( c, x, x1, d, qvz, di, sg );
                x1.x -= di.x;
                x1.y -= di.y;
                int2 c1 = make_int2( c.x + di.x + m.shx, c.y + di.y + m.shy );
                if ( m.refl && reflect_walls( m.refl, t % g.ntx, t / g.ntx, g.ntx, g.nty, g.nx, g.ny, c1, x1, u ) )
                    pu.store( o12, u );
                px.store( o8, x1 );
                pix.store( o8, c1 );
                const int xc = ( c1.x >= g.nx ) - ( c1.x < 0 );
                const int yc = ( c1.y >= g.ny ) - ( c1.y < 0 );
                dir = ( yc + 1 ) * 3 + ( xc + 1 );
            }
            {
                // particle order: lane l deposits all its segments before lane l+1
                for ( int l = 0; l < WAVE; l++ ) {
                    if ( lane == l && ns > 0 ) {
                        float v[8];
                        segment_current( sg[0], m.qnx, m.qny, v );
                        add_current( j0x, j0y, j0z, sg[0].c.x + ys * sg[0].c.y, ys, v );
                        if ( ns > 1 ) {
                            segment_current( sg[1], m.qnx, m.qny, v );
                            add_current( j0x, j0y, j0z, sg[1].c.x + ys * sg[1].c.y, ys, v );
                        }
                        if ( ns > 2 ) {
                            segment_current( sg[2], m.qnx, m.qny, v );
                            add_current( j0x, j0y, j0z, sg[2].c.x + ys * sg[2].c.y, ys, v );
                        }
                    }
                }
            }
            const bool leave = act && dir != 4;
            const unsigned long long mask = __ballot( leave );
            if ( leave ) {
                const int pos = nl + count_below( mask );
                sidx[beg + pos] = ( i << 4 ) | dir;
                atomicAdd( &s_cnt[dir], 1 );
            }
            nl += __popcll( mask );
        }
    }
    if constexpr ( MOVE && !DET ) {
        if ( qn > 0 ) flush_queue( qn );
        if ( lbon && lbn > 0 ) flush_leavers( lbn );
    }

    if constexpr ( PUSH ) {
        if constexpr ( !DET ) {
            for ( int o = WAVE / 2; o > 0; o >>= 1 ) en += __shfl_xor( en, o );
        }
        if ( lane == 0 ) s_en[w] = en;
    }
    if constexpr ( MOVE ) {
        if ( lane == 0 ) s_nl[w] = nl;
    }
    __syncthreads();

    if constexpr ( MOVE ) {
        // flush: J(tile) += sum of the waves' J (LDS), guard cells included
        // (species.cpp:1488-1492)
        const gptr<float> Jt = gbl( reinterpret_cast<float *>( J + (size_t) t * vol ) );
        for ( int k = tid; k < 3 * g.ext_x * g.ext_y; k += nthr ) {
            const int el = k / 3, c = k - 3 * el;
            float s = Jall[c * vol + el];
            for ( int v = 1; v < W; v++ ) s += Jall[3 * vol * v + c * vol + el];
            // jset (the first species of a step, fast mode): the region is
            // written instead of zeroed and added to; the same value, since
            // s is never -0 (every plane element starts at +0)
            if ( !DET && a.jset ) Jt[k] = s;
            else Jt[k] += s;
        }
        if ( tid == 0 ) {
            const gptr<int> seg = gbl( S.seg + 16 * t );
            const gptr<int> npt = gbl( S.npt + 9 * t );
            int tot = 0;
            for ( int v = 0; v < W; v++ ) {
                seg[v] = s_nl[v];
                tot += s_nl[v];
            }
            for ( int v = W; v < MAX_SEG; v++ ) seg[v] = 0;
            seg[MAX_SEG] = chunk;
            for ( int d = 0; d < 9; d++ ) npt[d] = s_cnt[d];
            npt[4] = np - tot;
        }
    }
    if constexpr ( PUSH ) {
        if ( tid == 0 ) {
            double s = 0.0;
            for ( int v = 0; v < W; v++ ) s += s_en[v];
            gbl( tile_energy )[t] = s;
        }
    }
}

template <bool PUSH, bool MOVE, bool DET>
__global__ void __launch_bounds__( FAST_BLOCK ) k_particles( PartArgs a, const PartArgs * tab, const int2 * map ) {
    extern __shared__ float smem[];
    const int t = select_instance( a, tab, map );
    particles_tile<PUSH, MOVE, DET>( a, t, smem );
}

// ---------------------------------------------------------------------------
// Tile sort.  Directions d = 0..8 encode the motion (d%3-1, d/3-1) of a
// particle from its source tile to its target tile; d = 4 is "stays".

__device__ inline int wrap_tile( int tx, int ty, int ntx, int nty, int px, int py ) {
    if ( px ) { if ( tx < 0 ) tx += ntx; if ( tx >= ntx ) tx -= ntx; }
    if ( py ) { if ( ty < 0 ) ty += nty; if ( ty >= nty ) ty -= nty; }
    if ( tx < 0 || tx >= ntx || ty < 0 || ty >= nty ) return -1;
    return ty * ntx + tx;
}

struct SortGeo {
    int ntx, nty, nx, ny, px, py;
};

// Tile receiving the particles that tile t sends in direction d (-1 if none).
__device__ inline int target_of( int t, int d, const SortGeo & q ) {
    return wrap_tile( t % q.ntx + ( d % 3 - 1 ), t / q.ntx + ( d / 3 - 1 ), q.ntx, q.nty, q.px, q.py );
}
// Tile that sends particles in direction d into tile t (-1 if none).
__device__ inline int source_of( int t, int d, const SortGeo & q ) {
    return wrap_tile( t % q.ntx - ( d % 3 - 1 ), t / q.ntx - ( d / 3 - 1 ), q.ntx, q.nty, q.px, q.py );
}

// Arguments of one instance of the tile sort.
struct SortArgs {
    PartView P;  // particles (offsets are the old ones until k_sort_in)
    PartView T;  // staging buffer; T.off receives the new offsets
    SortView S;
    SortGeo q;
    int ntiles;
    const int * extra;  // free slots reserved after each tile (may be null)
};

// New particle count per tile (bnd_check's new_np, particles.cpp:212-303):
// the tile gathers the direction counts its 9 source tiles send to it.
// Batched launches index the instance with blockIdx.y.
__device__ inline void sort_newnp_tile( const SortArgs & a, const int t ) {
    const SortG S = gview( a.S );
    int n = 0;
    for ( int d = 0; d < 9; d++ ) {
        const int s = source_of( t, d, a.q );
        if ( s >= 0 ) n += S.npt[9 * s + d];
    }
    S.new_np[t] = n;
}

__global__ void k_sort_newnp( SortArgs a, const SortArgs * tab ) {
    if ( tab ) a = gbl( tab )[blockIdx.y];
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if ( t >= a.ntiles ) return;
    sort_newnp_tile( a, t );
}

// Exclusive scan of the new counts into the new tile offsets
// (update_tile_info, particles.cpp:314-357).  One workgroup per instance;
// each thread takes SCAN_K consecutive tiles (16-byte loads and stores when
// the arrays allow), sums them, and one workgroup scan of the thread sums
// (wave shuffles, then the wave totals by the first wave) gives every
// thread its start; rounds of BLOCK * SCAN_K tiles follow with a running
// carry.  All loads of a round are issued before its scan.
// BLOCK: the workgroup size (a multiple of 64, at most 64 waves).
constexpr int SCAN_BLOCK = 1024;
constexpr int SCAN_K = 16;
template <int BLOCK>
__device__ inline void sort_count_body( const SortArgs & a ) {
    const SortG S = gview( a.S );
    const gptr<int> new_off = gbl( a.T.off );
    const gptr<const int> extra = gbl( a.extra );
    const gptr<const int> np_old = gbl( (const int *) a.P.np );
    const bool count_moved = a.S.nmove != nullptr;
    const int ntiles = a.ntiles;
    const int tid = threadIdx.x, lane = tid % WAVE, w = tid / WAVE;
    __shared__ int s_wave[BLOCK / WAVE];
    __shared__ unsigned long long s_moved;
    if ( tid == 0 ) s_moved = 0;
    // 16-byte accesses need every array 16-byte aligned (tile starts are
    // multiples of SCAN_K)
    const bool vec = ( ( reinterpret_cast<uintptr_t>( a.S.new_np ) | reinterpret_cast<uintptr_t>( a.T.off ) |
                         reinterpret_cast<uintptr_t>( a.extra ) | reinterpret_cast<uintptr_t>( a.P.np ) ) & 15 ) == 0;
    int carry = 0;
    unsigned long long moved = 0;
    for ( int base = 0; base < ntiles; base += BLOCK * SCAN_K ) {
        const int t0 = base + tid * SCAN_K;
        int v[SCAN_K];
        if ( vec && t0 + SCAN_K <= ntiles ) {
#pragma unroll
            for ( int q = 0; q < SCAN_K / 4; q++ ) {
                const int4 n4 = *(const gptr<int4>) ( S.new_np + t0 + 4 * q );
                int4 e4 = make_int4( 0, 0, 0, 0 );
                if ( a.extra ) e4 = *(const gptr<int4>) ( extra + t0 + 4 * q );
                v[4 * q] = n4.x + e4.x;
                v[4 * q + 1] = n4.y + e4.y;
                v[4 * q + 2] = n4.z + e4.z;
                v[4 * q + 3] = n4.w + e4.w;
                if ( count_moved ) {
                    const int4 m4 = *(const gptr<int4>) ( np_old + t0 + 4 * q );
                    moved += (unsigned) m4.x + (unsigned long long) (unsigned) m4.y + (unsigned) m4.z + (unsigned) m4.w;
                }
            }
        } else {
#pragma unroll
            for ( int k = 0; k < SCAN_K; k++ ) {
                const int t = t0 + k;
                v[k] = 0;
                if ( t < ntiles ) {
                    v[k] = S.new_np[t] + ( a.extra ? extra[t] : 0 );
                    if ( count_moved ) moved += (unsigned) np_old[t];
                }
            }
        }
        int sum = 0;
#pragma unroll
        for ( int k = 0; k < SCAN_K; k++ ) sum += v[k];
        // inclusive scan of the thread sums within the wave
        int x = sum;
        for ( int o = 1; o < WAVE; o <<= 1 ) {
            const int y = __shfl_up( x, o );
            if ( lane >= o ) x += y;
        }
        if ( lane == WAVE - 1 ) s_wave[w] = x;
        __syncthreads();
        if ( w == 0 ) {
            int z = lane < BLOCK / WAVE ? s_wave[lane] : 0;
            for ( int o = 1; o < BLOCK / WAVE; o <<= 1 ) {
                const int y = __shfl_up( z, o );
                if ( lane >= o ) z += y;
            }
            if ( lane < BLOCK / WAVE ) s_wave[lane] = z;  // inclusive wave totals
        }
        __syncthreads();
        int off = carry + ( w > 0 ? s_wave[w - 1] : 0 ) + x - sum;
        int o[SCAN_K];
#pragma unroll
        for ( int k = 0; k < SCAN_K; k++ ) {
            o[k] = off;
            off += v[k];
        }
        if ( vec && t0 + SCAN_K <= ntiles ) {
#pragma unroll
            for ( int q = 0; q < SCAN_K / 4; q++ )
                *(gptr<int4>) ( new_off + t0 + 4 * q ) = make_int4( o[4 * q], o[4 * q + 1], o[4 * q + 2], o[4 * q + 3] );
        } else {
#pragma unroll
            for ( int k = 0; k < SCAN_K; k++ )
                if ( t0 + k < ntiles ) new_off[t0 + k] = o[k];
        }
        carry += s_wave[BLOCK / WAVE - 1];
        __syncthreads();  // s_wave is rewritten by the next round
    }
    if ( tid == 0 ) S.total[0] = carry;
    // Species::move's d_nmove += np (species.cpp:2162): the counts the
    // particle pass just moved, i.e. before this sort
    if ( count_moved ) {
        __syncthreads();  // s_moved cleared
        for ( int o = WAVE / 2; o > 0; o >>= 1 ) moved += __shfl_xor( moved, o );
        if ( lane == 0 && moved ) atomicAdd( &s_moved, moved );
        __syncthreads();
        if ( tid == 0 ) S.nmove[0] += s_moved;
    }
}

__global__ void __launch_bounds__( SCAN_BLOCK ) k_sort_count( SortArgs a, const SortArgs * tab ) {
    if ( tab ) a = gbl( tab )[blockIdx.x];
    sort_count_body<SCAN_BLOCK>( a );
}

// Stayers that move to the staging buffer because the tile's range moved
// (copy_out, particles.cpp:409-423).
__device__ inline int shifted_stayers( int t, gptr<int> old_off, gptr<int> new_off, const SortG & S ) {
    const int o = old_off[t], no = new_off[t], n0 = S.npt[9 * t + 4], nn = S.new_np[t];
    int ns;
    if ( no >= o ) ns = no - o;
    else { ns = ( o + n0 ) - ( no + nn ); if ( ns < 0 ) ns = 0; }
    return ns > n0 ? n0 : ns;
}

// Position j of the tile's ordered leaver list; pre[] holds the segments'
// exclusive prefix counts (non-decreasing), in registers.
__device__ inline int leaver_at( int j, const int ( &pre )[MAX_SEG + 1], int chunk, gptr<int> idx, int off ) {
    int w = 0;
#pragma unroll
    for ( int v = 1; v < MAX_SEG; v++ ) w += pre[v] <= j;
    int base = 0;
#pragma unroll
    for ( int v = 0; v < MAX_SEG; v++ ) base = ( v == w ) ? pre[v] : base;
    return idx[off + w * chunk + ( j - base )];
}

constexpr int SORT_BLOCK = 256;
constexpr int SORT_CACHE = 4;  // iterations per wave whose list entries and cells stay in registers

// copy_out (particles.cpp:379-550): leavers to their staging slots, holes
// below n0 refilled from the stayers above n0, shifted stayers staged.  The
// staging slots of each target tile are filled in the reference's
// single-thread reservation order: by source tile, the source's own shifted
// stayers first, then directions 0..8.
// SORTED: copy_sorted (particles.cpp:628-763) instead: every particle goes to
// T at the new offsets, the n0 stayers (after the hole fill) reserved as
// direction 4 in the same by-source, by-direction order, and T.np is set.
// Tile s's part (one workgroup of at least 128 threads, any number of waves).
template <bool SORTED>
__device__ inline void sort_out_tile( const SortArgs & a, const int s ) {
    __shared__ int s_start[9];
    __shared__ int s_pre[MAX_SEG + 1];
    __shared__ int s_hdr[4];
    const PartG P = gview( a.P ), T = gview( a.T );
    const SortG S = gview( a.S );
    const SortGeo q = a.q;
    const int tid = threadIdx.x;

    // Set-up with every global load issued at once: thread 9 d + e loads the
    // count source e sends to this tile's target in direction d, thread
    // 81 + d that target's new offset, threads 96.. the wave segment counts
    // and the tile header; then nine threads form the staging starts.
    __shared__ int s_cnt9[9][9];
    __shared__ int s_src9[9][9];
    __shared__ int s_toff[9];
    __shared__ int s_seg[MAX_SEG];
    if ( tid < 81 ) {
        const int d = tid / 9, e = tid - 9 * ( tid / 9 );
        const int tg = ( d == 4 ) ? s : target_of( s, d, q );
        int se = -1, cnt = 0;
        if ( tg >= 0 ) {
            se = ( e == 4 ) ? tg : source_of( tg, e, q );
            if ( se >= 0 ) {
                if ( e != 4 ) cnt = S.npt[9 * se + e];
                else if ( SORTED ) cnt = S.npt[9 * tg + 4];
                else cnt = shifted_stayers( tg, P.off, T.off, S );
            }
        }
        s_cnt9[d][e] = cnt;
        s_src9[d][e] = se;
    } else if ( tid < 90 ) {
        const int d = tid - 81;
        const int tg = ( d == 4 ) ? s : target_of( s, d, q );
        s_toff[d] = tg >= 0 ? T.off[tg] : 0;
    } else if ( tid >= 96 && tid < 96 + MAX_SEG ) {
        s_seg[tid - 96] = S.seg[16 * s + ( tid - 96 )];
    } else if ( tid >= 112 && tid < 116 ) {
        const int h = tid - 112;
        s_hdr[h] = h == 0 ? P.off[s] : h == 1 ? P.np[s] : h == 2 ? S.npt[9 * s + 4] : S.seg[16 * s + MAX_SEG];
    }
    __syncthreads();
    if ( tid < 9 ) {
        const int d = tid;
        const int tg = ( d == 4 ) ? s : target_of( s, d, q );
        int start = -1;
        if ( tg >= 0 ) {
            const int mine = ( d == 4 && !SORTED ) ? -1 : d;
            int before = 0, total = 0;
            for ( int e = 0; e < 9; e++ ) {
                const int se = s_src9[d][e];
                if ( se < 0 ) continue;
                const int cnt = s_cnt9[d][e];
                const int rank = ( e == 4 && !SORTED ) ? -1 : e;
                if ( se < s || ( se == s && rank < mine ) ) before += cnt;
                total += cnt;
            }
            start = s_toff[d] + before;
            if ( d == 4 && !SORTED ) S.tmp_np[s] = total;
        }
        s_start[d] = start;
    }
    if ( tid == 0 ) {
        int acc = 0;
        for ( int w = 0; w < MAX_SEG; w++ ) { s_pre[w] = acc; acc += s_seg[w]; }
        s_pre[MAX_SEG] = acc;
    }
    __syncthreads();
    int pre[MAX_SEG + 1];
#pragma unroll
    for ( int v = 0; v <= MAX_SEG; v++ ) pre[v] = s_pre[v];

    const int o = s_hdr[0], np = s_hdr[1], n0 = s_hdr[2], chunk = s_hdr[3];
    const int nl = np - n0;

    // The leaver list and the tail [n0, np) (both nl long) are split into W
    // contiguous quarters, one per wave.  Pass 1 counts, per quarter, the
    // leavers of each direction and the stayers; the exclusive sums over the
    // preceding quarters give each wave its starting ranks.
    const int W = blockDim.x / WAVE, w = tid / WAVE, lane = tid % WAVE;
    const int qlen = ( ( nl + W - 1 ) / W + WAVE - 1 ) & ~( WAVE - 1 );
    const int j0 = min( nl, w * qlen ), j1 = min( nl, j0 + qlen );
    __shared__ int s_dir[FAST_BLOCK / WAVE][9];
    __shared__ int s_stay[FAST_BLOCK / WAVE];
    // The first SORT_CACHE iterations of a wave keep their leaver entries and
    // tail cells in registers from pass 1 for passes 2 and 3 (a wave's
    // quarter is rarely longer), so that those passes start from the records'
    // loads instead of reloading the list first.
    int ent_c[SORT_CACHE];
    int2 tix_c[SORT_CACHE];
    auto in_tile = [&]( int2 c ) { return c.x >= 0 && c.x < q.nx && c.y >= 0 && c.y < q.ny; };
    {
        int cnt[9] = { 0, 0, 0, 0, 0, 0, 0, 0, 0 };
        int st = 0;
        auto pass1 = [&]( int jb, int & ent, int2 & c ) {
            const int j = jb + lane;
            const bool valid = j < j1;
            ent = valid ? leaver_at( j, pre, chunk, S.idx, o ) : 4;
            c = valid ? P.ix[o + n0 + j] : make_int2( -1, -1 );
            const int d = ent & 15;
            for ( int dd = 0; dd < 9; dd++ ) cnt[dd] += __popcll( __ballot( d == dd && valid ) );
            st += __popcll( __ballot( valid && in_tile( c ) ) );
        };
#pragma unroll
        for ( int it = 0; it < SORT_CACHE; it++ ) {
            ent_c[it] = 4;
            tix_c[it] = make_int2( -1, -1 );
            if ( j0 + it * WAVE < j1 ) pass1( j0 + it * WAVE, ent_c[it], tix_c[it] );
        }
        for ( int jb = j0 + SORT_CACHE * WAVE; jb < j1; jb += WAVE ) {
            int ent;
            int2 c;
            pass1( jb, ent, c );
        }
        if ( lane == 0 ) {
            for ( int dd = 0; dd < 9; dd++ ) s_dir[w][dd] = cnt[dd];
            s_stay[w] = st;
        }
    }
    __syncthreads();

    // Pass 2: leavers to staging, rank within each direction = order in the list
    {
        int run[9];
        for ( int dd = 0; dd < 9; dd++ ) {
            run[dd] = 0;
            for ( int v = 0; v < w; v++ ) run[dd] += s_dir[v][dd];
        }
        auto pass2 = [&]( int jb, int ent ) {
            const int j = jb + lane;
            const bool valid = j < j1;
            const int d = ent & 15;
            int rank = 0;
            for ( int dd = 0; dd < 9; dd++ ) {
                if ( dd == 4 ) continue;
                const unsigned long long mk = __ballot( valid && d == dd );
                if ( d == dd ) rank = run[dd] + count_below( mk );
                run[dd] += __popcll( mk );
            }
            if ( valid && s_start[d] >= 0 ) {
                const int k = o + ( ent >> 4 );
                const int2 c = P.ix[k];
                const float2 x = P.x[k];
                const float3 u = ld3( P.u + k );
                const int xc = ( c.x >= q.nx ) - ( c.x < 0 );
                const int yc = ( c.y >= q.ny ) - ( c.y < 0 );
                const int dst = s_start[d] + rank;
                T.ix[dst] = make_int2( c.x - xc * q.nx, c.y - yc * q.ny );
                T.x[dst] = x;
                st3( T.u + dst, u );
            }
        };
#pragma unroll
        for ( int it = 0; it < SORT_CACHE; it++ )
            if ( j0 + it * WAVE < j1 ) pass2( j0 + it * WAVE, ent_c[it] );
        for ( int jb = j0 + SORT_CACHE * WAVE; jb < j1; jb += WAVE ) {
            const int j = jb + lane;
            pass2( jb, j < j1 ? leaver_at( j, pre, chunk, S.idx, o ) : 4 );
        }
    }
    // every leaver is read before any hole (a leaver slot below n0) is overwritten
    __syncthreads();

    // Pass 3: holes below n0 <- stayers above n0, both in increasing order
    {
        int r0 = 0;
        for ( int v = 0; v < w; v++ ) r0 += s_stay[v];
        auto pass3 = [&]( int jb, int2 c ) {
            const int j = jb + lane;
            const bool stay = j < j1 && in_tile( c );
            const unsigned long long mk = __ballot( stay );
            if ( stay ) {
                const float2 x = P.x[o + n0 + j];
                const float3 u = ld3( P.u + ( o + n0 + j ) );
                const int r = r0 + count_below( mk );
                const int hole = o + ( leaver_at( r, pre, chunk, S.idx, o ) >> 4 );
                P.ix[hole] = c;
                P.x[hole] = x;
                st3( P.u + hole, u );
            }
            r0 += __popcll( mk );
        };
#pragma unroll
        for ( int it = 0; it < SORT_CACHE; it++ )
            if ( j0 + it * WAVE < j1 ) pass3( j0 + it * WAVE, tix_c[it] );
        for ( int jb = j0 + SORT_CACHE * WAVE; jb < j1; jb += WAVE ) {
            const int j = jb + lane;
            pass3( jb, j < j1 ? P.ix[o + n0 + j] : make_int2( -1, -1 ) );
        }
    }
    __syncthreads();

    if constexpr ( SORTED ) {
        // all stayers (after the hole fill)
        const int dst0 = s_start[4];
        for ( int i = tid; i < n0; i += blockDim.x ) {
            T.ix[dst0 + i] = P.ix[o + i];
            T.x[dst0 + i] = P.x[o + i];
            st3( T.u + ( dst0 + i ), ld3( P.u + ( o + i ) ) );
        }
        if ( tid == 0 ) T.np[s] = S.new_np[s];
    } else {
        // shifted stayers (after the hole fill)
        const int ns = shifted_stayers( s, P.off, T.off, S );
        const int first = ( T.off[s] >= o ) ? 0 : n0 - ns;
        const int dst0 = s_start[4];
        for ( int i = tid; i < ns; i += blockDim.x ) {
            T.ix[dst0 + i] = P.ix[o + first + i];
            T.x[dst0 + i] = P.x[o + first + i];
            st3( T.u + ( dst0 + i ), ld3( P.u + ( o + first + i ) ) );
        }
        __syncthreads();
        if ( tid == 0 ) P.np[s] = n0 - ns;
    }
}

template <bool SORTED>
__global__ void __launch_bounds__( SORT_BLOCK ) k_sort_out( SortArgs a, const SortArgs * tab, const int2 * map ) {
    const int s = select_instance( a, tab, map );
    sort_out_tile<SORTED>( a, s );
}

// copy_in (particles.cpp:559-614): staged particles appended after the
// in-place stayers, or first into the gap when the range moved down.
__device__ inline void sort_in_tile( const SortArgs & a, const int t ) {
    __shared__ int s_hdr[4];
    const PartG P = gview( a.P ), T = gview( a.T );
    const SortG S = gview( a.S );
    if ( threadIdx.x == 0 ) {
        s_hdr[0] = P.off[t];
        s_hdr[1] = P.np[t];
        s_hdr[2] = T.off[t];
        s_hdr[3] = S.tmp_np[t];
    }
    __syncthreads();
    const int o = s_hdr[0], n = s_hdr[1], no = s_hdr[2], nin = s_hdr[3];
    int front = 0;
    if ( no < o ) { front = o - no; if ( front > nin ) front = nin; }
    for ( int i = threadIdx.x; i < nin; i += blockDim.x ) {
        const int dst = ( i < front ) ? no + i : no + n + i;
        P.ix[dst] = T.ix[no + i];
        P.x[dst] = T.x[no + i];
        st3( P.u + dst, ld3( T.u + ( no + i ) ) );
    }
    __syncthreads();
    if ( threadIdx.x == 0 ) {
        P.np[t] = n + nin;
        P.off[t] = no;
    }
}

__global__ void __launch_bounds__( SORT_BLOCK ) k_sort_in( SortArgs a, const SortArgs * tab, const int2 * map ) {
    const int t = select_instance( a, tab, map );
    sort_in_tile( a, t );
}

// ---------------------------------------------------------------------------
// Order-free tile sort of the fast mode.  The particle pass has compacted
// each wave chunk's stayers to the chunk's front (particle arrays) and
// written its leavers, cell still in the source tile's frame, to the
// staging arrays at the chunk's first positions; seg[v] counts chunk v's
// leavers and npt[9 t + d] the tile's leavers of direction d.  After the new
// counts and offsets (k_sort_newnp, k_sort_count):
//   k_sort_shift: each tile stages the stayers that lie outside its new range
//     into the free slots of its staging region (behind each chunk's
//     leavers), keeps its old offset and count in seg[SEG_OLD_OFF],
//     seg[SEG_OLD_NP], and buckets its leavers by direction: idx[o + start(d)
//     + r] = the chunk-relative staging position of the r-th leaver of
//     direction d in staging order, start(d) = the tile's leavers of the
//     directions before d;
//   k_sort_in_fast: each tile fills the slots of its new range that do not
//     hold one of its stayers with its staged stayers, then with the leavers
//     of its eight sources' buckets towards it, in a fixed enumeration order.
// Every record is read once and written once after the particle pass; the
// within-tile order is not the reference's (counts, offsets and per-tile
// multisets are).  One wave sorts one tile: the work per tile is a few
// hundred records behind a chain of dependent loads, so the kernels are
// bound by latency and gain from as many tiles in flight as possible.
constexpr int SEG_OLD_OFF = 13, SEG_OLD_NP = 14;
constexpr int SORT_WAVES = 4;       // tiles per workgroup of k_sort_shift / k_sort_in_fast
constexpr int SORT_WAVE_LDS = 384;  // ints of LDS per wave
constexpr int SORT_BATCH = 4;       // records per lane loaded before any is stored

// Exclusive prefix sum of x over the wave's lanes; `total` gets the sum.
__device__ inline int wave_excl_scan( int x, int & total ) {
    const int lane = threadIdx.x % WAVE;
    int s = x;
#pragma unroll
    for ( int o = 1; o < WAVE; o <<= 1 ) {
        const int y = __shfl_up( s, o );
        if ( lane >= o ) s += y;
    }
    total = __shfl( s, WAVE - 1 );
    return s - x;
}

// Orders the wave's LDS writes before its following LDS reads (the wave's
// LDS accesses execute in order; this keeps the compiler from moving them).
__device__ inline void wave_sync() {
    __builtin_amdgcn_fence( __ATOMIC_RELEASE, "wavefront" );
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence( __ATOMIC_ACQUIRE, "wavefront" );
}

// Segment v of a prefix table pre[0..MAX_SEG] (non-decreasing) that holds
// position k (the last v with pre[v] <= k).
__device__ inline int segment_of( int k, const int * pre ) {
    int w = 0;
#pragma unroll
    for ( int v = 1; v < MAX_SEG; v++ ) w += pre[v] <= k;
    return w;
}

// Direction of a cell relative to the tile (4: inside).
__device__ inline int cell_dir( int2 c, const SortGeo & q ) {
    return ( ( c.y >= q.ny ) - ( c.y < 0 ) + 1 ) * 3 + ( ( c.x >= q.nx ) - ( c.x < 0 ) + 1 );
}

// k_sort_shift for tile t by one wave; L = the wave's SORT_WAVE_LDS ints.
__device__ inline void sort_shift_wave( const SortArgs & a, const int t, int * L ) {
    int * const seg = L;        // [16] leavers per chunk, [MAX_SEG] = chunk
    int * const gpre = L + 16;  // [16] free-slot prefix of the staging region
    int * const pre = L + 32;   // [16] prefix of the leaver counts
    const PartG P = gview( a.P );
    const SortG S = gview( a.S );
    const gptr<int2> Gix = gbl( a.S.stage_ix );
    const gptr<float2> Gx = gbl( a.S.stage_x );
    const gptr<float3> Gu = gbl( a.S.stage_u );
    const gptr<const int> new_off = gbl( (const int *) a.T.off );
    const int lane = threadIdx.x % WAVE;
    wave_sync();  // the wave's previous tile has read L

    // lanes 0-15: seg; 16-19: old offset, old count, new offset, new count;
    // 20-28: direction counts
    int val = 0;
    if ( lane <= MAX_SEG ) val = S.seg[16 * t + lane];
    else if ( lane == 16 ) val = P.off[t];
    else if ( lane == 17 ) val = P.np[t];
    else if ( lane == 18 ) val = new_off[t];
    else if ( lane == 19 ) val = S.new_np[t];
    else if ( lane < 29 ) val = S.npt[9 * t + lane - 20];
    const int o = __builtin_amdgcn_readlane( val, 16 ), np = __builtin_amdgcn_readlane( val, 17 );
    const int no = __builtin_amdgcn_readlane( val, 18 ), nn = __builtin_amdgcn_readlane( val, 19 );
    const int chunk = __builtin_amdgcn_readlane( val, MAX_SEG );
    {
        // chunk v = [v chunk, min(np, (v + 1) chunk)), its first seg[v]
        // staging slots hold leavers, the rest are free
        const int v = lane;
        const int s = v < MAX_SEG ? val : 0;
        const int len = v < MAX_SEG ? min( max( np - v * chunk, 0 ), chunk ) : 0;
        int nl, nfree;
        const int p = wave_excl_scan( s, nl );
        const int f = wave_excl_scan( v < MAX_SEG ? max( len - s, 0 ) : 0, nfree );
        if ( v <= MAX_SEG ) {
            seg[v] = val;
            pre[v] = v < MAX_SEG ? p : nl;
            gpre[v] = v < MAX_SEG ? f : nfree;
        }
    }
    if ( lane == 0 ) {
        S.seg[16 * t + SEG_OLD_OFF] = o;
        S.seg[16 * t + SEG_OLD_NP] = np;
    }
    // bucket starts: the leavers of the directions before (lane 20 + d)
    int dstart, dtot;
    {
        const int d = lane - 20;
        dstart = wave_excl_scan( ( d >= 0 && d < 9 && d != 4 ) ? val : 0, dtot );
    }
    wave_sync();

    // the leavers' cells, SORT_BATCH per lane at a time; the first batch is
    // loaded before the stayer copies below (which write other staging
    // slots), so that the two round trips overlap
    const int nl = pre[MAX_SEG];
    int pos[SORT_BATCH];
    int2 cl[SORT_BATCH];
    auto load_batch = [&]( int j0 ) {
#pragma unroll
        for ( int i = 0; i < SORT_BATCH; i++ ) {
            const int j = j0 + i * WAVE + lane;
            pos[i] = -1;
            cl[i] = make_int2( 0, 0 );
            if ( j < nl ) {
                const int v = segment_of( j, pre );
                pos[i] = v * chunk + ( j - pre[v] );
                cl[i] = Gix[o + pos[i]];
            }
        }
    };
    load_batch( 0 );

    // stayers at old-range positions below and above the new range: to the
    // free staging slots in order
    const int a0 = o, la = max( min( o + np, no ) - o, 0 );
    const int b0 = max( o, no + nn ), lb = max( o + np - b0, 0 );
    const int nout = la + lb;
    int ns = 0;
    for ( int j0 = 0; j0 < nout; j0 += WAVE ) {
        const int j = j0 + lane;
        bool stay = false;
        int p = 0;
        if ( j < nout ) {
            p = j < la ? a0 + j : b0 + ( j - la );
            const int r = p - o, v = r / chunk, len = min( np - v * chunk, chunk );
            stay = r - v * chunk < len - seg[v];
        }
        const unsigned long long m = __ballot( stay );
        if ( stay ) {
            const int k = ns + count_below( m );
            const int w = segment_of( k, gpre );
            const int dst = o + w * chunk + seg[w] + ( k - gpre[w] );
            const int2 c = P.ix[p];
            const float2 x = P.x[p];
            const float3 u = ld3( P.u + p );
            Gix[dst] = c;
            Gx[dst] = x;
            st3( Gu + dst, u );
        }
        ns += __popcll( m );
    }
    if ( lane == 0 ) S.tmp_np[t] = ns;

    // direction buckets of the leavers, staging order kept in each
    int run[9];
#pragma unroll
    for ( int d = 0; d < 9; d++ ) run[d] = __builtin_amdgcn_readlane( dstart, 20 + d );
    for ( int j0 = 0; j0 < nl; j0 += SORT_BATCH * WAVE ) {
        if ( j0 > 0 ) load_batch( j0 );
        int dir[SORT_BATCH];
#pragma unroll
        for ( int i = 0; i < SORT_BATCH; i++ ) dir[i] = pos[i] >= 0 ? cell_dir( cl[i], a.q ) : 4;
#pragma unroll
        for ( int i = 0; i < SORT_BATCH; i++ ) {
            int at = -1;
#pragma unroll
            for ( int d = 0; d < 9; d++ ) {
                if ( d == 4 ) continue;
                const unsigned long long m = __ballot( dir[i] == d );
                if ( dir[i] == d ) at = run[d] + count_below( m );
                run[d] += __popcll( m );
            }
            if ( at >= 0 ) S.idx[o + at] = pos[i];
        }
    }
}

__global__ void __launch_bounds__( SORT_WAVES * WAVE ) k_sort_shift( SortArgs a, const SortArgs * tab, const int2 * map, int n ) {
    __shared__ int s_lds[SORT_WAVES][SORT_WAVE_LDS];
    const int w = __builtin_amdgcn_readfirstlane( (int) threadIdx.x / WAVE );
    const int g = blockIdx.x * SORT_WAVES + w;
    if ( g >= n ) return;
    const int t = select_item( a, tab, map, g );
    sort_shift_wave( a, t, s_lds[w] );
}

// k_sort_in_fast for tile t by one wave; L = the wave's SORT_WAVE_LDS ints.
// The slots of the new range [no, no + nn) that none of the tile's stayers
// holds are runs: the part before the old range, the tail of every chunk
// (behind its compacted stayers) and the part after the old range.  They
// are filled in order with the tile's staged stayers, then with its eight
// sources' leavers towards it, source by source (each in its staging order,
// from the source's direction bucket).
__device__ inline void sort_in_fast_wave( const SortArgs & a, const int t, int * L ) {
    int * const seg = L;          // [9][16] leavers per chunk of the 9 sources (e = 4: the tile)
    int * const so_s = L + 144;   // [9] source tile
    int * const so_o = L + 153;   // [9] its old offset
    int * const so_n = L + 162;   // [9] its old count
    int * const npt = L + 171;    // [9][9] direction counts of the 9 sources
    int * const hdr = L + 252;    // [4] new offset, new count, staged stayers
    int * const gpre = L + 256;   // [16] free-slot prefix of the tile's staging region
    int * const rb = L + 272;     // [17] hole run starts, in position order
    int * const rpre = L + 289;   // [18] prefix of their lengths
    int * const bst = L + 307;    // [9] start of source e's bucket towards the tile
    int * const ebase = L + 316;  // [10] first rank of source e's particles; [9] = all
    const PartG P = gview( a.P );
    const SortG S = gview( a.S );
    const gptr<const int2> Gix = gbl( (const int2 *) a.S.stage_ix );
    const gptr<const float2> Gx = gbl( (const float2 *) a.S.stage_x );
    const gptr<const float3> Gu = gbl( (const float3 *) a.S.stage_u );
    const gptr<const int> new_off = gbl( (const int *) a.T.off );
    const SortGeo q = a.q;
    const int lane = threadIdx.x % WAVE;
    wave_sync();  // the wave's previous tile has read L

    // set-up: every load at once (228 values, four per lane)
#pragma unroll
    for ( int i = 0; i < 4; i++ ) {
        const int k = i * WAVE + lane;
        if ( k < 144 ) {
            const int e = k / 16, v = k % 16;
            const int s = ( e == 4 ) ? t : source_of( t, e, q );
            const int x = s >= 0 ? S.seg[16 * s + v] : 0;
            seg[k] = ( v == SEG_OLD_OFF || v == SEG_OLD_NP ) ? 0 : x;
            if ( v == SEG_OLD_OFF ) so_o[e] = x;
            if ( v == SEG_OLD_NP ) so_n[e] = x;
            if ( v == 0 ) so_s[e] = s;
        } else if ( k < 144 + 81 ) {
            const int e = ( k - 144 ) / 9, d = ( k - 144 ) % 9;
            const int s = source_of( t, e, q );
            npt[k - 144] = ( e != 4 && s >= 0 ) ? S.npt[9 * s + d] : 0;
        } else if ( k < 144 + 81 + 3 ) {
            const int h = k - 225;
            hdr[h] = h == 0 ? new_off[t] : h == 1 ? S.new_np[t] : S.tmp_np[t];
        }
    }
    wave_sync();

    const int o = so_o[4], np = so_n[4], chunk = seg[4 * 16 + MAX_SEG];
    const int no = hdr[0], nn = hdr[1], ns = hdr[2];
    {
        // lanes 0-8: source e's bucket towards the tile and its first rank
        const int e = lane;
        int st = 0, cin = 0;
        if ( e < 9 && e != 4 ) {
            for ( int d = 0; d < e; d++ ) st += d == 4 ? 0 : npt[9 * e + d];
            cin = npt[9 * e + e];
        }
        int all;
        const int eb = ns + wave_excl_scan( e < 9 ? cin : 0, all );
        if ( e < 9 ) {
            bst[e] = st;
            ebase[e] = eb;
        }
        if ( e == 9 ) ebase[9] = ns + all;
    }
    {
        // lanes 0-14: chunk v's free staging slots (behind its leavers)
        const int v = lane;
        const int len = v < MAX_SEG && chunk > 0 ? min( max( np - v * chunk, 0 ), chunk ) : 0;
        const int sv = v < MAX_SEG ? seg[4 * 16 + v] : 0;
        int nfree;
        const int f = wave_excl_scan( v < MAX_SEG ? max( len - sv, 0 ) : 0, nfree );
        if ( v <= MAX_SEG ) gpre[v] = v < MAX_SEG ? f : nfree;
    }
    {
        // lanes 0-16: hole runs of the new range, clipped to it: before the
        // old range, behind chunk r - 1's stayers, after the old range
        const int r = lane;
        int b = 0, en = 0;
        if ( r == 0 ) {
            b = no;
            en = o;
        } else if ( r <= MAX_SEG ) {
            const int v = r - 1;
            const int len = chunk > 0 ? min( max( np - v * chunk, 0 ), chunk ) : 0;
            b = o + v * chunk + ( len - seg[4 * 16 + v] );
            en = o + v * chunk + len;
        } else if ( r == MAX_SEG + 1 ) {
            b = o + np;
            en = no + nn;
        }
        b = max( b, no );
        en = min( en, no + nn );
        int nholes;
        const int rp = wave_excl_scan( r <= MAX_SEG + 1 ? max( en - b, 0 ) : 0, nholes );
        if ( r <= MAX_SEG + 1 ) {
            rb[r] = b;
            rpre[r] = rp;
        }
        if ( r == MAX_SEG + 2 ) rpre[r] = nholes;
    }
    wave_sync();

    const int nholes = rpre[MAX_SEG + 2], ntot = ebase[9];
    // position of fill rank k (-1 past the holes: the counts always balance)
    auto slot = [&]( int k ) {
        if ( k >= nholes ) return -1;
        int r = 0;
#pragma unroll
        for ( int v = 1; v < MAX_SEG + 2; v++ ) r += rpre[v] <= k;
        return rb[r] + ( k - rpre[r] );
    };
    // ranks [0, ns): the tile's staged stayers; [ebase[e], ebase[e + 1]):
    // source e's bucket towards the tile.  SORT_BATCH ranks per lane are
    // resolved and loaded before any is written.
    for ( int k0 = 0; k0 < ntot; k0 += SORT_BATCH * WAVE ) {
        int src[SORT_BATCH], dst[SORT_BATCH], sh[SORT_BATCH];
#pragma unroll
        for ( int i = 0; i < SORT_BATCH; i++ ) {
            const int k = k0 + i * WAVE + lane;
            src[i] = -1;
            dst[i] = -1;
            sh[i] = 4;
            if ( k < ntot ) {
                dst[i] = slot( k );
                if ( k < ns ) {
                    const int v = segment_of( k, gpre );
                    src[i] = o + v * chunk + seg[4 * 16 + v] + ( k - gpre[v] );
                } else {
                    int e = 0;
#pragma unroll
                    for ( int f = 1; f < 9; f++ ) e = ebase[f] <= k ? f : e;
                    const int oe = so_o[e];
                    const int pos = S.idx[oe + bst[e] + ( k - ebase[e] )];
                    sh[i] = e;
                    src[i] = ( pos >= 0 && pos < so_n[e] ) ? oe + pos : -1;
                }
            }
        }
        int2 c[SORT_BATCH];
        float2 x[SORT_BATCH];
        float3 u[SORT_BATCH];
#pragma unroll
        for ( int i = 0; i < SORT_BATCH; i++ ) {
            c[i] = make_int2( 0, 0 );
            x[i] = make_float2( 0.f, 0.f );
            u[i] = make_float3( 0.f, 0.f, 0.f );
            if ( src[i] >= 0 && dst[i] >= 0 ) {
                c[i] = Gix[src[i]];
                x[i] = Gx[src[i]];
                u[i] = ld3( Gu + src[i] );
            }
        }
#pragma unroll
        for ( int i = 0; i < SORT_BATCH; i++ ) {
            if ( src[i] >= 0 && dst[i] >= 0 ) {
                // a source's leavers move into this tile's frame (e = 4: none)
                const int dx = sh[i] % 3 - 1, dy = sh[i] / 3 - 1;
                P.ix[dst[i]] = make_int2( c[i].x - dx * q.nx, c[i].y - dy * q.ny );
                P.x[dst[i]] = x[i];
                st3( P.u + dst[i], u[i] );
            }
        }
    }
    if ( lane == 0 ) {
        P.np[t] = nn;
        P.off[t] = no;
    }
}

__global__ void __launch_bounds__( SORT_WAVES * WAVE ) k_sort_in_fast( SortArgs a, const SortArgs * tab, const int2 * map, int n ) {
    __shared__ int s_lds[SORT_WAVES][SORT_WAVE_LDS];
    const int w = __builtin_amdgcn_readfirstlane( (int) threadIdx.x / WAVE );
    const int g = blockIdx.x * SORT_WAVES + w;
    if ( g >= n ) return;
    const int t = select_item( a, tab, map, g );
    sort_in_fast_wave( a, t, s_lds[w] );
}

// ---------------------------------------------------------------------------
// dep_charge_kernel (species.cpp:2348-2395) into a scalar grid with guards.
template <bool DET>
__global__ void __launch_bounds__( FAST_BLOCK ) k_charge( PartView P, Geo g, float * __restrict__ rho, float q ) {
    extern __shared__ float smem[];
    const int t = blockIdx.x;
    const int vol = g.vol, ys = g.ext_x;
    for ( int k = threadIdx.x; k < vol; k += blockDim.x ) smem[k] = 0.0f;
    __syncthreads();
    float * c0 = smem + g.off;
    const int off = P.off[t], np = P.np[t];
    const int lane = threadIdx.x % WAVE;
    const int nthr = blockDim.x;
    for ( int i0 = ( threadIdx.x / WAVE ) * WAVE; i0 < np; i0 += nthr ) {
        const int i = i0 + lane;
        const bool act = i < np;
        int k = 0;
        float w00 = 0, w10 = 0, w01 = 0, w11 = 0;
        if ( act ) {
            const int2 c = P.ix[off + i];
            const float2 x = P.x[off + i];
            k = c.y * ys + c.x;
            const float s0x = 0.5f - x.x, s1x = 0.5f + x.x, s0y = 0.5f - x.y, s1y = 0.5f + x.y;
            w00 = s0y * s0x * q; w10 = s0y * s1x * q; w01 = s1y * s0x * q; w11 = s1y * s1x * q;
        }
        if constexpr ( DET ) {
            for ( int l = 0; l < WAVE; l++ ) {
                if ( lane == l && act ) {
                    c0[k] += w00; c0[k + 1] += w10; c0[k + ys] += w01; c0[k + ys + 1] += w11;
                }
            }
        } else if ( act ) {
            // diagnostic path: LDS float atomics are adequate here
            __hip_atomic_fetch_add( &c0[k], w00, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP );
            __hip_atomic_fetch_add( &c0[k + 1], w10, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP );
            __hip_atomic_fetch_add( &c0[k + ys], w01, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP );
            __hip_atomic_fetch_add( &c0[k + ys + 1], w11, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP );
        }
    }
    __syncthreads();
    float * dst = rho + (size_t) t * vol;
    for ( int k = threadIdx.x; k < g.ext_x * g.ext_y; k += blockDim.x ) dst[k] += smem[k];
}

Geo make_geo( const zpic_geom * f ) {
    Geo g;
    g.nx = f->nx; g.ny = f->ny; g.ext_x = f->ext_x; g.ext_y = f->ext_y; g.vol = f->tile_vol; g.off = f->offset;
    g.ntx = f->ntiles_x; g.nty = f->ntiles_y;
    return g;
}

PartView view( const zpic_particles * p ) {
    PartView v;
    v.np = p->np; v.off = p->offset;
    v.ix = reinterpret_cast<int2 *>( p->ix );
    v.x = reinterpret_cast<float2 *>( p->x );
    v.u = reinterpret_cast<float3 *>( p->u );
    return v;
}

SortView view( const zpic_sort * s ) {
    SortView v;
    if ( s ) {
        v.idx = s->idx; v.npt = s->npt; v.new_np = s->new_np; v.seg = s->seg; v.tmp_np = s->tmp_np; v.total = s->total;
        v.nmove = s->nmove;
        const bool st = s->stage_ix && s->stage_x && s->stage_u;
        v.stage_ix = st ? reinterpret_cast<int2 *>( s->stage_ix ) : nullptr;
        v.stage_x = st ? reinterpret_cast<float2 *>( s->stage_x ) : nullptr;
        v.stage_u = st ? reinterpret_cast<float3 *>( s->stage_u ) : nullptr;
    } else {
        v.idx = v.npt = v.new_np = v.seg = v.tmp_np = v.total = nullptr;
        v.nmove = nullptr;
        v.stage_ix = nullptr; v.stage_x = nullptr; v.stage_u = nullptr;
    }
    return v;
}


int check_geom( const zpic_geom * g ) {
    if ( !g || g->ntiles_x < 1 || g->ntiles_y < 1 || g->nx < 1 || g->ny < 1 )
        return set_error( ZPIC_EINVAL, "invalid grid geometry" );
    if ( g->gc_lower < 1 || g->gc_upper < 2 )
        return set_error( ZPIC_EINVAL, "particle stages need guard cells {1,2}" );
    return ZPIC_OK;
}

// Threads per tile of the fast particle pass (a multiple of 64, at most
// FAST_BLOCK); the deterministic mode always uses one wavefront.  Default 256:
// on the 2048^2 Weibel step the particle pass took 7.08 ms with 4 waves per
// tile against 7.36 ms with 8 (fewer private J planes to clear and reduce).
int g_threads = 256;  // library-wide setting
// Fast-mode tile sort: 0 (default) = order-free (leavers staged by the
// particle pass, k_sort_shift + k_sort_in_fast), 1 = the reference's
// within-tile order (k_sort_out + k_sort_in).  The deterministic mode always
// keeps the reference's.  On the 2048^2 Weibel step the order-free sort
// takes about 0.11 ms per species against 0.30 ms, at the same particle-pass
// time.
int g_reference_order = 0;

constexpr size_t LDS_LIMIT = 65536;  // dynamic LDS per workgroup

// Threads per tile the particle pass runs with for tiles of `vol` elements
// when asked for `threads` (reduced until the LDS staging fits).
int fitted_threads( int threads, int vol ) {
    while ( threads > WAVE && particles_lds_bytes( vol, threads / WAVE, true, true ) > LDS_LIMIT ) threads -= WAVE;
    return particles_lds_bytes( vol, threads / WAVE, true, true ) > LDS_LIMIT ? 0 : threads;
}

int check_threads( int threads ) {
    if ( threads < WAVE || threads > FAST_BLOCK || threads % WAVE )
        return set_error( ZPIC_EINVAL, "threads per tile must be a multiple of 64 in [64, 512]" );
    return ZPIC_OK;
}

template <bool PUSH, bool MOVE>
int launch_particles( const zpic_particles * p, const zpic_geom * fld, const float * E, const float * B, float * J,
                      float alpha, MoveArgs m, double * tile_energy, zpic_sort * s, int det, void * stream,
                      int jset = 0 ) {
    if ( int rc = check_geom( fld ) ) return rc;
    if ( !p || ( PUSH && ( !E || !B || !tile_energy ) ) || ( MOVE && ( !J || !s ) ) )
        return set_error( ZPIC_EINVAL, "null buffer" );
    const int nt = fld->ntiles_x * fld->ntiles_y;
    // Large tiles: fewer wavefronts (each holds a private J copy) until LDS fits.
    int threads = det ? WAVE : g_threads;
    while ( threads > WAVE && particles_lds_bytes( fld->tile_vol, threads / WAVE, PUSH, MOVE ) > LDS_LIMIT )
        threads -= WAVE;
    size_t lds = particles_lds_bytes( fld->tile_vol, threads / WAVE, PUSH, MOVE );
    if ( lds > LDS_LIMIT ) return set_error( ZPIC_ERANGE, "tile too large for LDS staging" );
    PartArgs a{};
    a.P = view( p ); a.g = make_geo( fld );
    a.E = (const float3 *) E; a.B = (const float3 *) B; a.J = (float3 *) J;
    a.alpha = alpha; a.m = m; a.tile_energy = tile_energy; a.S = view( s );
    a.jset = MOVE && !det && jset;
    if ( MOVE && !det && s->stage_ix ) {
        // the leaver buffers, where they cost no workgroup per CU
        const size_t with = particles_lds_bytes( fld->tile_vol, threads / WAVE, PUSH, MOVE, true );
        int dev = 0, cu_lds = 0;
        ZPIC_HIP_TRY( hipGetDevice( &dev ) );
        ZPIC_HIP_TRY( hipDeviceGetAttribute( &cu_lds, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev ) );
        const size_t stat = 256;  // the kernel's static LDS, rounded up
        if ( with <= LDS_LIMIT && cu_lds / ( with + stat ) >= cu_lds / ( lds + stat ) ) {
            a.lbuf = 1;
            lds = with;
        }
    }
    hipStream_t st = as_stream( stream );
    if ( det )
        hipLaunchKernelGGL( ( k_particles<PUSH, MOVE, true> ), dim3( nt ), dim3( threads ), lds, st, a,
                            (const PartArgs *) nullptr, (const int2 *) nullptr );
    else
        hipLaunchKernelGGL( ( k_particles<PUSH, MOVE, false> ), dim3( nt ), dim3( threads ), lds, st, a,
                            (const PartArgs *) nullptr, (const int2 *) nullptr );
    ZPIC_LAUNCH_CHECK( "k_particles" );
    return ZPIC_OK;
}

SortArgs sort_args( zpic_particles * p, zpic_particles * tmp, zpic_sort * s, int ntx, int nty, int nx, int ny, int px,
                    int py ) {
    SortArgs a;
    a.P = view( p ); a.T = view( tmp ); a.S = view( s );
    a.q = SortGeo{ ntx, nty, nx, ny, px ? 1 : 0, py ? 1 : 0 };
    a.ntiles = ntx * nty;
    a.extra = nullptr;
    return a;
}

}  // namespace

// ------------------------------------------------------------------ batches
// Per entry: sum of its per-tile push energies (parallel tree, fast mode),
// written to the entry's energy output.  One workgroup per entry.
struct EnergyArgs {
    const double * tile;
    int n;
    double * out;
};

// Device tables for one launch per stage over many independent instances:
// entry e of a table holds the arguments of one (instance, species) pair,
// map[b] = (entry, tile) for block b.
struct BatchParticlePlan {
    int nentries = 0;
    int nblocks = 0;
    int max_tiles = 0;
    int max_vol = 0;
    int threads = FAST_BLOCK;  // workgroup size: the largest of the entries' fitted threads
    size_t lds = 0;            // the largest of the entries' LDS needs
    bool staged = false;       // every entry uses the order-free sort
    PartArgs * d_part = nullptr;
    SortArgs * d_sort = nullptr;
    int2 * d_map = nullptr;
    EnergyArgs * d_energy = nullptr;  // [nentries]
    std::vector<PartArgs> h_part;
};

int batch_particles_create( BatchParticlePlan ** out, const zpic_batch_species * items, int n, int threads ) {
    if ( !out || !items || n < 1 ) return set_error( ZPIC_EINVAL, "empty batch" );
    if ( int rc = check_threads( threads ) ) return rc;
    auto * plan = new BatchParticlePlan();
    plan->threads = WAVE;
    plan->nentries = n;
    std::vector<SortArgs> sorts( n );
    std::vector<EnergyArgs> energies( n );
    std::vector<int2> map;
    for ( int e = 0; e < n; e++ ) {
        const zpic_batch_species & it = items[e];
        if ( int rc = check_geom( &it.fld ) ) { delete plan; return rc; }
        PartArgs a{};
        a.P = view( &it.p ); a.g = make_geo( &it.fld );
        a.E = (const float3 *) it.E; a.B = (const float3 *) it.B; a.J = (float3 *) it.J;
        a.alpha = it.alpha; a.m = MoveArgs{ it.dt_dx, it.dt_dy, it.q, it.qnx, it.qny };
        a.tile_energy = it.tile_energy;
        zpic_sort s = it.s;
        a.S = view( &s );
        // each entry runs with the threads that fit its own tiles, as its own
        // single-instance pass does
        if ( it.threads && check_threads( it.threads ) ) { delete plan; return ZPIC_EINVAL; }
        const int th = fitted_threads( it.threads ? it.threads : threads, it.fld.tile_vol );
        if ( th == 0 ) { delete plan; return set_error( ZPIC_ERANGE, "tile too large for LDS staging" ); }
        a.waves = th / WAVE;
        plan->threads = std::max( plan->threads, th );
        plan->lds = std::max( plan->lds, particles_lds_bytes( it.fld.tile_vol, a.waves, true, true ) );
        plan->h_part.push_back( a );
        zpic_particles p = it.p, t = it.tmp;
        sorts[e] = sort_args( &p, &t, &s, it.fld.ntiles_x, it.fld.ntiles_y, it.fld.nx, it.fld.ny, it.periodic_x,
                              it.periodic_y );
        const int nt = it.fld.ntiles_x * it.fld.ntiles_y;
        energies[e] = EnergyArgs{ it.tile_energy, nt, it.energy_out };
        for ( int t2 = 0; t2 < nt; t2++ ) map.push_back( make_int2( e, t2 ) );
        plan->max_tiles = std::max( plan->max_tiles, nt );
        plan->max_vol = std::max( plan->max_vol, it.fld.tile_vol );
    }
    plan->nblocks = (int) map.size();
    {
        int nst = 0;
        for ( int e = 0; e < n; e++ ) nst += items[e].s.stage_ix != nullptr;
        if ( nst != 0 && nst != n ) {
            delete plan;
            return set_error( ZPIC_EINVAL, "batch entries must all use the same sort (staged or not)" );
        }
        plan->staged = nst == n;
    }
    hipError_t e1 = hipMalloc( (void **) &plan->d_part, sizeof( PartArgs ) * n );
    hipError_t e2 = hipMalloc( (void **) &plan->d_sort, sizeof( SortArgs ) * n );
    hipError_t e3 = hipMalloc( (void **) &plan->d_map, sizeof( int2 ) * map.size() );
    hipError_t e4 = hipMalloc( (void **) &plan->d_energy, sizeof( EnergyArgs ) * n );
    if ( e1 || e2 || e3 || e4 ) {
        (void) hipFree( plan->d_part ); (void) hipFree( plan->d_sort ); (void) hipFree( plan->d_map );
        (void) hipFree( plan->d_energy );
        delete plan;
        return set_error( ZPIC_EHIP, "hipMalloc failed for batch tables" );
    }
    ZPIC_HIP_TRY( hipMemcpy( plan->d_part, plan->h_part.data(), sizeof( PartArgs ) * n, hipMemcpyHostToDevice ) );
    ZPIC_HIP_TRY( hipMemcpy( plan->d_sort, sorts.data(), sizeof( SortArgs ) * n, hipMemcpyHostToDevice ) );
    ZPIC_HIP_TRY( hipMemcpy( plan->d_map, map.data(), sizeof( int2 ) * map.size(), hipMemcpyHostToDevice ) );
    ZPIC_HIP_TRY( hipMemcpy( plan->d_energy, energies.data(), sizeof( EnergyArgs ) * n, hipMemcpyHostToDevice ) );
    *out = plan;
    return ZPIC_OK;
}

void batch_particles_destroy( BatchParticlePlan * plan ) {
    if ( !plan ) return;
    (void) hipFree( plan->d_part ); (void) hipFree( plan->d_sort ); (void) hipFree( plan->d_map );
    (void) hipFree( plan->d_energy );
    delete plan;
}

// Sum of an entry's per-tile push energies: per-thread runs, a butterfly
// within each wave, then the wave sums in wave order (any workgroup size).
__device__ inline void batch_energy_body( const EnergyArgs & a ) {
    __shared__ double s_wsum[FAST_BLOCK / WAVE];
    const gptr<const double> tile = gbl( a.tile );
    double acc = 0;
    for ( int t = threadIdx.x; t < a.n; t += blockDim.x ) acc += tile[t];
    for ( int o = WAVE / 2; o > 0; o >>= 1 ) acc += __shfl_xor( acc, o );
    if ( threadIdx.x % WAVE == 0 ) s_wsum[threadIdx.x / WAVE] = acc;
    __syncthreads();
    if ( threadIdx.x == 0 ) {
        double s = 0;
        for ( int v = 0; v < (int) blockDim.x / WAVE; v++ ) s += s_wsum[v];
        gbl( a.out )[0] = s;
    }
    __syncthreads();  // s_wsum may be reused by the next body in the workgroup
}

__global__ void __launch_bounds__( 256 ) k_batch_energy( const EnergyArgs * tab ) {
    batch_energy_body( gbl( tab )[blockIdx.x] );
}

int batch_particles_step( BatchParticlePlan * plan, void * stream ) {
    hipStream_t st = as_stream( stream );
    PartArgs dummy{};
    hipLaunchKernelGGL( ( k_particles<true, true, false> ), dim3( plan->nblocks ), dim3( plan->threads ), plan->lds, st,
                        dummy, (const PartArgs *) plan->d_part, (const int2 *) plan->d_map );
    ZPIC_LAUNCH_CHECK( "k_particles (batch)" );
    SortArgs sd{};
    hipLaunchKernelGGL( k_sort_newnp, dim3( ( plan->max_tiles + 255 ) / 256, plan->nentries ), dim3( 256 ), 0, st, sd,
                        (const SortArgs *) plan->d_sort );
    ZPIC_LAUNCH_CHECK( "k_sort_newnp (batch)" );
    hipLaunchKernelGGL( k_sort_count, dim3( plan->nentries ), dim3( SCAN_BLOCK ), 0, st, sd,
                        (const SortArgs *) plan->d_sort );
    ZPIC_LAUNCH_CHECK( "k_sort_count (batch)" );
    if ( plan->staged ) {
        const dim3 grid( ( plan->nblocks + SORT_WAVES - 1 ) / SORT_WAVES );
        hipLaunchKernelGGL( k_sort_shift, grid, dim3( SORT_WAVES * WAVE ), 0, st, sd, (const SortArgs *) plan->d_sort,
                            (const int2 *) plan->d_map, plan->nblocks );
        ZPIC_LAUNCH_CHECK( "k_sort_shift (batch)" );
        hipLaunchKernelGGL( k_sort_in_fast, grid, dim3( SORT_WAVES * WAVE ), 0, st, sd, (const SortArgs *) plan->d_sort,
                            (const int2 *) plan->d_map, plan->nblocks );
        ZPIC_LAUNCH_CHECK( "k_sort_in_fast (batch)" );
    } else {
        hipLaunchKernelGGL( k_sort_out<false>, dim3( plan->nblocks ), dim3( SORT_BLOCK ), 0, st, sd,
                            (const SortArgs *) plan->d_sort, (const int2 *) plan->d_map );
        ZPIC_LAUNCH_CHECK( "k_sort_out (batch)" );
        hipLaunchKernelGGL( k_sort_in, dim3( plan->nblocks ), dim3( SORT_BLOCK ), 0, st, sd,
                            (const SortArgs *) plan->d_sort, (const int2 *) plan->d_map );
        ZPIC_LAUNCH_CHECK( "k_sort_in (batch)" );
    }
    hipLaunchKernelGGL( k_batch_energy, dim3( plan->nentries ), dim3( 256 ), 0, st, (const EnergyArgs *) plan->d_energy );
    ZPIC_LAUNCH_CHECK( "k_batch_energy" );
    return ZPIC_OK;
}

namespace {

// ---------------------------------------------------------------------------
// Persistent batched step: Simulation::advance (simulation.h:90-107) of every
// instance of a batch in one launch.  The step is a list of work items
// (stage, instance, tiles) ordered so that every item's predecessors come
// before it.  Workgroups claim items in that order from STEP_QUEUES queues
// (their home queue first, then the others), wait until the stages their item
// depends on are complete for that instance (a counter of finished items per
// (stage, instance)), run it and count it as finished.  A workgroup only ever
// waits for items claimed before its own by workgroups that are running, so
// the step needs no co-resident grid and always drains; a wait longer than
// the spin limit (which would be a bug) raises an error flag and goes on
// instead of hanging.  Different instances overlap freely: a small instance
// runs its sort or its field update while a large one is still in its
// particle pass.
enum StepKind : int {
    SK_ZERO = 0,  // Current::zero of tiles [t0, t1)
    SK_PART,      // particle pass of tiles [t0, t1), one after the other
    SK_NEWNP,     // new per-tile counts of tiles [t0, t1)
    SK_COUNT,     // offsets scan, nmove and push energy of the species
    SK_SHIFT,     // order-free sort: stayers outside the new range staged (tiles [t0, t1))
    SK_PLACE,     // order-free sort: placement
    SK_OUT,       // reference-order sort: copy_out
    SK_IN,        // reference-order sort: copy_in
    SK_ADD_X,     // add_from_gc x pass of J (tiles [t0, t1))
    SK_ADD_Y,     // add_from_gc y pass of J
    SK_COPY_XY,   // copy_to_gc of one grid (tiles [t0, t1))
    SK_COPY_X,    // copy_to_gc_x of J after an x filter pass
    SK_COPY_Y,    // copy_to_gc_y of J after a y filter pass
    SK_K3,        // one filter pass (tiles [t0, t1))
    SK_YEE,       // Yee update (tiles [t0, t1))
};

struct StepItem {
    int kind, arg, t0, t1;
    int done;                      // counter of this item's (stage, instance)
    int dep0, need0, dep1, need1;  // counters to wait for (-1: none), final values
    int pad[3];
};

constexpr int STEP_QUEUES = 8;
constexpr int STEP_SPIN_LIMIT = 1 << 22;  // polls of one wait (seconds), then the error flag

struct StepTables {
    const StepItem * items;
    int qoff[STEP_QUEUES + 1];  // queue q holds items [qoff[q], qoff[q + 1])
    int * next;                 // [STEP_QUEUES] claim counters, then the stage counters
    int * counters;
    int * error;
    const PartArgs * part;
    const SortArgs * sort;
    const EnergyArgs * energy;
    const grid_ops::GcArgs * gc;
    const grid_ops::K3Args * k3;
    const grid_ops::YeeArgs * yee;
};

// A table entry read into scalar registers: the entry is the same for the
// whole workgroup, and values the compiler cannot prove uniform (loads after
// stores in the kernel) would otherwise live in vector registers.
template <class X>
__device__ inline X uniform_load( const X * p ) {
    static_assert( sizeof( X ) % 4 == 0, "4-byte words" );
    X out;
    const gptr<const int> src = gbl( reinterpret_cast<const int *>( p ) );
    int * dst = reinterpret_cast<int *>( &out );
#pragma unroll
    for ( int i = 0; i < (int) ( sizeof( X ) / 4 ); i++ ) dst[i] = __builtin_amdgcn_readfirstlane( src[i] );
    return out;
}

// Polls with relaxed loads: an acquire load invalidates the L2 of the
// polling workgroup's XCD on every poll, which evicted every other
// workgroup's data and made the step an order of magnitude slower; the
// caller issues one acquire fence after the waits instead.
__device__ inline void step_wait( const StepTables & T, int c, int need ) {
    if ( c < 0 ) return;
    int spins = 0;
    while ( __hip_atomic_load( T.counters + c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT ) < need ) {
        if ( ++spins > STEP_SPIN_LIMIT ) {
            __hip_atomic_fetch_or( T.error, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT );
            return;
        }
        __builtin_amdgcn_s_sleep( 4 );
    }
}

// A workgroup runs up to STEP_CHAIN items of one (stage, instance) in a row
// behind one acquire and publishes them with one release: such items wait for
// the same counters (so the acquire of the first covers the others) and do
// not read each other's output.  Finished items are published before the
// workgroup waits for anything else, so that it never waits for itself.
constexpr int STEP_CHAIN = 8;

template <int BLOCK>
__global__ void __launch_bounds__( BLOCK, 4 ) k_batch_step( StepTables T ) {
    extern __shared__ float smem[];
    __shared__ int s_item;
    const int home = blockIdx.x % STEP_QUEUES;
    int q = 0;  // queues exhausted so far, counted from the home queue
    int pending = 0;      // finished items not yet published
    int pend_done = -1;   // their counter
    int pend_dep[4] = { -1, 0, -1, 0 };  // their dependencies
    // every wave's stores complete at the L2, then one device-scope release
    // (an L2 write-back) and the counter increment
    auto publish = [&]() {
        if ( pending == 0 ) return;
        asm volatile( "s_waitcnt vmcnt(0)" ::: "memory" );
        __syncthreads();
        if ( threadIdx.x == 0 ) {
            __builtin_amdgcn_fence( __ATOMIC_RELEASE, "agent" );
            __hip_atomic_fetch_add( T.counters + pend_done, pending, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT );
        }
        pending = 0;
    };
    for ( ;; ) {
        if ( threadIdx.x == 0 ) {
            int k = -1;
            while ( q < STEP_QUEUES ) {
                const int qq = ( home + q ) % STEP_QUEUES;
                const int n = T.qoff[qq + 1] - T.qoff[qq];
                if ( n > 0 ) {
                    const int c = __hip_atomic_fetch_add( T.next + qq, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT );
                    if ( c < n ) {
                        k = T.qoff[qq] + c;
                        break;
                    }
                }
                q++;
            }
            s_item = k;
        }
        __syncthreads();
        // uniform across the workgroup: item and argument loads become scalar
        const int k = __builtin_amdgcn_readfirstlane( s_item );
        __syncthreads();
        if ( k < 0 ) break;
        const StepItem it = uniform_load( T.items + k );
        const bool chained = pending > 0 && pending < STEP_CHAIN && it.done == pend_done && it.dep0 == pend_dep[0] &&
                             it.need0 == pend_dep[1] && it.dep1 == pend_dep[2] && it.need1 == pend_dep[3];
        if ( !chained ) {
            publish();
            if ( threadIdx.x == 0 ) {
                step_wait( T, it.dep0, it.need0 );
                step_wait( T, it.dep1, it.need1 );
                // the caches this workgroup reads through (its CU's L1, its
                // XCD's L2) are dropped once, before any wave loads the
                // item's inputs
                __builtin_amdgcn_fence( __ATOMIC_ACQUIRE, "agent" );
            }
            pend_done = it.done;
            pend_dep[0] = it.dep0;
            pend_dep[1] = it.need0;
            pend_dep[2] = it.dep1;
            pend_dep[3] = it.need1;
        }
        __syncthreads();
        switch ( it.kind ) {
        case SK_ZERO: {
            const grid_ops::GcArgs g = uniform_load( T.gc + it.arg );
            for ( int t = it.t0; t < it.t1; t++ ) grid_ops::zero_tile( g.buf, 3 * g.g.vol, t );
            break;
        }
        case SK_PART: {
            const PartArgs a = uniform_load( T.part + it.arg );
            for ( int t = it.t0; t < it.t1; t++ ) {
                if ( t > it.t0 ) __syncthreads();  // LDS of the previous tile read out
                particles_tile<true, true, false>( a, t, smem );
            }
            break;
        }
        case SK_NEWNP: {
            const SortArgs a = uniform_load( T.sort + it.arg );
            for ( int t = it.t0 + (int) threadIdx.x; t < it.t1; t += BLOCK ) sort_newnp_tile( a, t );
            break;
        }
        case SK_COUNT:
            sort_count_body<BLOCK>( uniform_load( T.sort + it.arg ) );
            __syncthreads();
            batch_energy_body( uniform_load( T.energy + it.arg ) );
            break;
        case SK_SHIFT:
        case SK_PLACE: {
            // one tile per wave, in the wave's part of the LDS
            const SortArgs a = uniform_load( T.sort + it.arg );
            const int w = __builtin_amdgcn_readfirstlane( (int) threadIdx.x / WAVE );
            int * L = reinterpret_cast<int *>( smem ) + w * SORT_WAVE_LDS;
            for ( int t = it.t0 + w; t < it.t1; t += BLOCK / WAVE ) {
                if ( it.kind == SK_SHIFT ) sort_shift_wave( a, t, L );
                else sort_in_fast_wave( a, t, L );
            }
            break;
        }
        case SK_OUT:
        case SK_IN: {
            const SortArgs a = uniform_load( T.sort + it.arg );
            for ( int t = it.t0; t < it.t1; t++ ) {
                if ( t > it.t0 ) __syncthreads();
                if ( it.kind == SK_OUT ) sort_out_tile<false>( a, t );
                else sort_in_tile( a, t );
            }
            break;
        }
        case SK_ADD_X:
        case SK_ADD_Y:
        case SK_COPY_XY:
        case SK_COPY_X:
        case SK_COPY_Y: {
            const grid_ops::GcArgs g = uniform_load( T.gc + it.arg );
            for ( int t = it.t0; t < it.t1; t++ ) {
                if ( it.kind == SK_ADD_X ) grid_ops::add_gc_x_tile<3>( g, t );
                else if ( it.kind == SK_ADD_Y ) grid_ops::add_gc_y_tile<3>( g, t );
                else if ( it.kind == SK_COPY_XY ) grid_ops::copy_gc_xy_tile<3>( g, t );
                else if ( it.kind == SK_COPY_X ) grid_ops::copy_gc_x_tile<3>( g, t );
                else grid_ops::copy_gc_y_tile<3>( g, t );
            }
            break;
        }
        case SK_K3: {
            const grid_ops::K3Args a = uniform_load( T.k3 + it.arg );
            for ( int t = it.t0; t < it.t1; t++ ) {
                if ( t > it.t0 ) __syncthreads();
                grid_ops::kernel3_tile( a, t, smem );
            }
            break;
        }
        case SK_YEE: {
            const grid_ops::YeeArgs a = uniform_load( T.yee + it.arg );
            for ( int t = it.t0; t < it.t1; t++ ) {
                if ( t > it.t0 ) __syncthreads();
                grid_ops::yee_tile( a, t, smem );
            }
            break;
        }
        default: break;
        }
        // published with the chain (one L2 write-back per chain, not per item
        // or per wave: each is a whole-cache operation that serialises in the
        // L2)
        pending++;
        __syncthreads();  // LDS of this item read out before the next one
    }
    publish();
}

// Clears the claim and stage counters before each step (a kernel, so that the
// step can be captured in a graph without memset nodes).
__global__ void k_step_reset( int * v, int n ) {
    for ( int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x ) gbl( v )[k] = 0;
}

struct BatchStepPlan {
    int block = 256;
    size_t lds = 0;
    int grid = 0;
    int nsync = 0;  // claim counters + stage counters
    StepTables T{};
    std::vector<void *> allocs;
    ~BatchStepPlan() {
        for ( void * p : allocs ) (void) hipFree( p );
    }
    template <class X>
    int upload( X ** dst, const std::vector<X> & v ) {
        *dst = nullptr;
        if ( v.empty() ) return ZPIC_OK;
        void * p = nullptr;
        ZPIC_HIP_TRY( hipMalloc( &p, sizeof( X ) * v.size() ) );
        allocs.push_back( p );
        ZPIC_HIP_TRY( hipMemcpy( p, v.data(), sizeof( X ) * v.size(), hipMemcpyHostToDevice ) );
        *dst = static_cast<X *>( p );
        return ZPIC_OK;
    }
};

// Filter passes of Current::advance (filter.h:22-78) as (dir, a, b, c).
struct FilterPass {
    int dir;
    float a, b, c;
};
std::vector<FilterPass> filter_passes( const zpic_batch_fields_item & it ) {
    std::vector<FilterPass> out;
    if ( it.filter_kind == 0 ) return out;
    const unsigned order = it.filter_order > 0 ? it.filter_order : 1;
    for ( unsigned k = 0; k < order; k++ ) out.push_back( FilterPass{ it.filter_dir, 0.25f, 0.5f, 0.25f } );
    if ( it.filter_kind == 2 ) {
        const float a = -1.0f;
        const float b = ( 4.0 + 2.0 * order ) / order;
        const float norm = 2 * a + b;
        out.push_back( FilterPass{ it.filter_dir, a / norm, b / norm, a / norm } );
    }
    return out;
}

int batch_step_create( BatchStepPlan ** out, const zpic_batch_species * sp, int nslots, const zpic_batch_fields_item * fl,
                       int n, int threads ) {
    if ( !out || !sp || !fl || n < 1 || nslots < 1 ) return set_error( ZPIC_EINVAL, "empty batch" );
    if ( int rc = check_threads( threads ) ) return rc;
    auto plan = std::make_unique<BatchStepPlan>();
    // instances by decreasing particle count: the big ones claim first
    std::vector<int> order( n );
    std::vector<long long> weight( n, 0 );
    std::vector<int> waves( n );
    int block = 256;
    size_t lds = 0;
    // order-free sort (every item's staging buffers set) or the reference order
    const bool staged = sp[0].s.stage_ix && sp[0].s.stage_x && sp[0].s.stage_u;
    for ( int i = 0; i < n; i++ ) {
        order[i] = i;
        const zpic_batch_fields_item & f = fl[i];
        if ( int rc = check_geom( &f.g ) ) return rc;
        if ( !f.E || !f.B || !f.J ) return set_error( ZPIC_EINVAL, "null field buffer" );
        if ( f.filter_kind < 0 || f.filter_kind > 2 )
            return set_error( ZPIC_EINVAL, "batches run the none, binomial and compensated current filters only" );
        if ( !f.periodic_x || !f.periodic_y ) return set_error( ZPIC_EINVAL, "the persistent step runs periodic boundaries only" );
        const int vol = f.g.tile_vol;
        const int req = sp[i].threads ? sp[i].threads : threads;  // slot 0 of the instance
        if ( int rc = check_threads( req ) ) return rc;
        const int th = fitted_threads( req, vol );
        if ( th == 0 ) return set_error( ZPIC_ERANGE, "tile too large for LDS staging" );
        waves[i] = th / WAVE;
        block = std::max( block, th );
        lds = std::max( { lds, particles_lds_bytes( vol, waves[i], true, true ), sizeof( float ) * 6 * (size_t) vol } );
        if ( staged ) lds = std::max( lds, sizeof( int ) * SORT_WAVE_LDS * ( 512 / WAVE ) );  // order-free sort, a tile per wave
        for ( int k = 0; k < nslots; k++ ) {
            const zpic_batch_species & s = sp[k * n + i];
            const bool st = s.s.stage_ix && s.s.stage_x && s.s.stage_u;
            if ( st != staged ) return set_error( ZPIC_EINVAL, "batch entries must all use the same sort (staged or not)" );
            if ( s.fld.ntiles_x != f.g.ntiles_x || s.fld.ntiles_y != f.g.ntiles_y || s.fld.nx != f.g.nx ||
                 s.fld.ny != f.g.ny || s.E != f.E || s.B != f.B || s.J != f.J )
                return set_error( ZPIC_EINVAL, "species and fields of an instance do not match" );
            if ( !s.periodic_x || !s.periodic_y ) return set_error( ZPIC_EINVAL, "the persistent step runs periodic boundaries only" );
        }
        // work of the instance: its particles (buffer sizes) and its cells
        weight[i] = 4LL * f.g.ntiles_x * f.g.ntiles_y * f.g.tile_vol;
        for ( int k = 0; k < nslots; k++ ) weight[i] += sp[k * n + i].p.max_part;
    }
    if ( lds > LDS_LIMIT ) return set_error( ZPIC_ERANGE, "tile too large for LDS staging" );
    std::stable_sort( order.begin(), order.end(), [&]( int a, int b ) { return weight[a] > weight[b]; } );
    plan->block = block;
    plan->lds = lds;

    // argument tables
    std::vector<PartArgs> part( (size_t) nslots * n );
    std::vector<SortArgs> sort( (size_t) nslots * n );
    std::vector<EnergyArgs> ener( (size_t) nslots * n );
    std::vector<grid_ops::GcArgs> gc( 3 * (size_t) n );
    std::vector<grid_ops::YeeArgs> yee( n );
    std::vector<grid_ops::K3Args> k3;
    std::vector<std::vector<int>> k3_of( n );  // k3 table index of each pass of each instance
    std::vector<std::vector<FilterPass>> passes( n );
    size_t max_passes = 0;
    for ( int i = 0; i < n; i++ ) {
        const zpic_batch_fields_item & f = fl[i];
        const grid_ops::G g = grid_ops::make( &f.g );
        gc[3 * i + 0] = grid_ops::GcArgs{ f.J, g, 1, 1 };
        gc[3 * i + 1] = grid_ops::GcArgs{ f.E, g, 1, 1 };
        gc[3 * i + 2] = grid_ops::GcArgs{ f.B, g, 1, 1 };
        yee[i] = grid_ops::YeeArgs{ f.E, f.B, f.J, g, f.dt_dx, f.dt_dy, f.dt };
        passes[i] = filter_passes( f );
        max_passes = std::max( max_passes, passes[i].size() );
        for ( const FilterPass & p : passes[i] ) {
            k3_of[i].push_back( (int) k3.size() );
            k3.push_back( grid_ops::K3Args{ f.J, g, p.dir, p.a, p.b, p.c } );
        }
        for ( int k = 0; k < nslots; k++ ) {
            const zpic_batch_species & it = sp[k * n + i];
            PartArgs a{};
            a.P = view( &it.p );
            a.g = make_geo( &it.fld );
            a.E = (const float3 *) it.E; a.B = (const float3 *) it.B; a.J = (float3 *) it.J;
            a.alpha = it.alpha;
            a.m = MoveArgs{ it.dt_dx, it.dt_dy, it.q, it.qnx, it.qny };
            a.tile_energy = it.tile_energy;
            zpic_sort s = it.s;
            a.S = view( &s );
            a.waves = waves[i];
            part[k * n + i] = a;
            zpic_particles p = it.p, t = it.tmp;
            sort[k * n + i] = sort_args( &p, &t, &s, it.fld.ntiles_x, it.fld.ntiles_y, it.fld.nx, it.fld.ny, 1, 1 );
            ener[k * n + i] = EnergyArgs{ it.tile_energy, it.fld.ntiles_x * it.fld.ntiles_y, it.energy_out };
        }
    }

    // stages in claim order; counters[stage * n + instance]
    struct Stage {
        int kind;
        int slot;  // species slot, filter pass or grid (1 E+B) as the kind needs
    };
    std::vector<Stage> stages;
    stages.push_back( { SK_ZERO, 0 } );
    for ( int k = 0; k < nslots; k++ ) stages.push_back( { SK_PART, k } );
    for ( int k = 0; k < nslots; k++ ) {
        stages.push_back( { SK_NEWNP, k } );
        stages.push_back( { SK_COUNT, k } );
        stages.push_back( { staged ? SK_SHIFT : SK_OUT, k } );
        stages.push_back( { staged ? SK_PLACE : SK_IN, k } );
    }
    stages.push_back( { SK_ADD_X, 0 } );
    stages.push_back( { SK_ADD_Y, 0 } );
    stages.push_back( { SK_COPY_XY, 0 } );
    for ( size_t p = 0; p < max_passes; p++ ) {
        stages.push_back( { SK_K3, (int) p } );
        stages.push_back( { SK_COPY_X, (int) p } );  // x or y guard copy by the pass's direction
    }
    stages.push_back( { SK_YEE, 0 } );
    stages.push_back( { SK_COPY_XY, 1 } );
    const int ns = (int) stages.size();

    // tiles per item of the light stages
    constexpr int ZERO_TILES = 32, NEWNP_TILES = 1024, GC_TILES = 16;
    constexpr long long PART_GRAIN = 65536;
    constexpr int CELL_GRAIN = 4096;
    // tiles per item of the per-tile stages: small tiles are grouped so that
    // an item moves at least ~65536 particles (particle pass, sort) or covers
    // at least 4096 cells (Yee, filter pass), amortising the claim, the waits
    // and above all the fences (an L2 write-back each) over more work (on the
    // 64-instance scan: 17.8 ms per step at 16384 / 1024, 14.1 ms at 65536 /
    // 4096, 18.7 ms at 262144 / 16384, where too few items remain to fill the
    // device)
    std::vector<int> part_group( n ), cell_group( n );
    for ( int i = 0; i < n; i++ ) {
        const int cells = fl[i].g.nx * fl[i].g.ny;
        long long ppt = 0;  // particles per tile of the largest species slot
        for ( int k = 0; k < nslots; k++ ) {
            const zpic_batch_species & it = sp[k * n + i];
            const long long nt = (long long) it.fld.ntiles_x * it.fld.ntiles_y;
            ppt = std::max( ppt, (long long) it.p.max_part / std::max( nt, 1LL ) );
        }
        part_group[i] = (int) std::min( 1024LL, std::max( 1LL, PART_GRAIN / std::max( ppt, 1LL ) ) );
        cell_group[i] = std::min( 256, std::max( 1, CELL_GRAIN / std::max( cells, 1 ) ) );
    }
    std::vector<int> need( (size_t) ns * n, 0 );  // items per (stage, instance)
    std::vector<std::vector<StepItem>> per_stage( ns );
    std::vector<int> last( n, -1 );  // last stage with items, per instance (current chain)
    std::vector<int> part_stage( n, -1 );
    for ( int si = 0; si < ns; si++ ) {
        const Stage st = stages[si];
        for ( int oi = 0; oi < n; oi++ ) {
            const int i = order[oi];
            const zpic_batch_fields_item & f = fl[i];
            const int nt = f.g.ntiles_x * f.g.ntiles_y;
            const int ctr = si * n + i;
            auto push = [&]( int kind, int arg, int t0, int t1, int dep0 ) {
                StepItem it{};
                it.kind = kind; it.arg = arg; it.t0 = t0; it.t1 = t1; it.done = ctr;
                it.dep0 = dep0; it.dep1 = -1; it.need1 = 0;
                per_stage[si].push_back( it );
                need[ctr]++;
            };
            // predecessor counter of this stage for instance i (-1: none)
            int dep = -1;
            switch ( st.kind ) {
            case SK_ZERO:
                for ( int t = 0; t < nt; t += ZERO_TILES ) push( SK_ZERO, 3 * i, t, std::min( nt, t + ZERO_TILES ), -1 );
                last[i] = si;
                break;
            case SK_PART:
                dep = last[i] * n + i;
                for ( int t = 0; t < nt; t += part_group[i] )
                    push( SK_PART, st.slot * n + i, t, std::min( nt, t + part_group[i] ), dep );
                last[i] = si;
                part_stage[i] = si;
                break;
            case SK_NEWNP:
                // after the species' own particle pass
                dep = ( 1 + st.slot ) * n + i;
                for ( int t = 0; t < nt; t += NEWNP_TILES ) push( SK_NEWNP, st.slot * n + i, t, std::min( nt, t + NEWNP_TILES ), dep );
                break;
            case SK_COUNT:
            case SK_SHIFT:
            case SK_PLACE:
            case SK_OUT:
            case SK_IN:
                dep = ( si - 1 ) * n + i;
                if ( st.kind == SK_COUNT ) push( SK_COUNT, st.slot * n + i, 0, nt, dep );
                else {
                    // the order-free sort takes a tile per wave: at least two per wave and item
                    const int grp = ( st.kind == SK_SHIFT || st.kind == SK_PLACE ) ? std::max( part_group[i], 16 ) : part_group[i];
                    for ( int t = 0; t < nt; t += grp ) push( st.kind, st.slot * n + i, t, std::min( nt, t + grp ), dep );
                }
                break;
            case SK_ADD_X:
            case SK_ADD_Y:
            case SK_COPY_XY:
            case SK_YEE: {
                dep = last[i] * n + i;
                const int arg = st.kind == SK_YEE ? i : 3 * i + st.slot;
                if ( st.kind == SK_COPY_XY && st.slot == 1 ) {
                    // E and B after the Yee update: two items per tile range
                    for ( int t = 0; t < nt; t += GC_TILES ) {
                        push( SK_COPY_XY, 3 * i + 1, t, std::min( nt, t + GC_TILES ), dep );
                        push( SK_COPY_XY, 3 * i + 2, t, std::min( nt, t + GC_TILES ), dep );
                    }
                } else if ( st.kind == SK_YEE ) {
                    for ( int t = 0; t < nt; t += cell_group[i] ) push( SK_YEE, arg, t, std::min( nt, t + cell_group[i] ), dep );
                } else {
                    for ( int t = 0; t < nt; t += GC_TILES ) push( st.kind, arg, t, std::min( nt, t + GC_TILES ), dep );
                }
                last[i] = si;
                break;
            }
            case SK_K3:
            case SK_COPY_X: {
                if ( (size_t) st.slot >= passes[i].size() ) break;  // no such pass: the chain skips it
                dep = last[i] * n + i;
                if ( st.kind == SK_K3 ) {
                    for ( int t = 0; t < nt; t += cell_group[i] )
                        push( SK_K3, k3_of[i][st.slot], t, std::min( nt, t + cell_group[i] ), dep );
                } else {
                    const int kind = passes[i][st.slot].dir == 0 ? SK_COPY_X : SK_COPY_Y;
                    for ( int t = 0; t < nt; t += GC_TILES ) push( kind, 3 * i, t, std::min( nt, t + GC_TILES ), dep );
                }
                last[i] = si;
                break;
            }
            }
        }
    }
    // final counts of the predecessors; queues: item j of a stage to queue j % STEP_QUEUES
    std::vector<std::vector<StepItem>> queues( STEP_QUEUES );
    for ( int si = 0; si < ns; si++ ) {
        int j = 0;
        for ( StepItem it : per_stage[si] ) {
            it.need0 = it.dep0 >= 0 ? need[it.dep0] : 0;
            queues[j++ % STEP_QUEUES].push_back( it );
        }
    }
    std::vector<StepItem> items;
    for ( int qq = 0; qq < STEP_QUEUES; qq++ ) {
        plan->T.qoff[qq] = (int) items.size();
        items.insert( items.end(), queues[qq].begin(), queues[qq].end() );
    }
    plan->T.qoff[STEP_QUEUES] = (int) items.size();
    plan->nsync = STEP_QUEUES + ns * n;

    StepItem * d_items = nullptr;
    if ( int rc = plan->upload( &d_items, items ) ) return rc;
    plan->T.items = d_items;
    std::vector<int> sync( plan->nsync + 1, 0 );  // claim counters, stage counters, error flag
    int * d_sync = nullptr;
    if ( int rc = plan->upload( &d_sync, sync ) ) return rc;
    plan->T.next = d_sync;
    plan->T.counters = d_sync + STEP_QUEUES;
    plan->T.error = d_sync + plan->nsync;
    PartArgs * d_part = nullptr;
    SortArgs * d_sort = nullptr;
    EnergyArgs * d_ener = nullptr;
    grid_ops::GcArgs * d_gc = nullptr;
    grid_ops::K3Args * d_k3 = nullptr;
    grid_ops::YeeArgs * d_yee = nullptr;
    if ( int rc = plan->upload( &d_part, part ) ) return rc;
    if ( int rc = plan->upload( &d_sort, sort ) ) return rc;
    if ( int rc = plan->upload( &d_ener, ener ) ) return rc;
    if ( int rc = plan->upload( &d_gc, gc ) ) return rc;
    if ( int rc = plan->upload( &d_k3, k3 ) ) return rc;
    if ( int rc = plan->upload( &d_yee, yee ) ) return rc;
    plan->T.part = d_part; plan->T.sort = d_sort; plan->T.energy = d_ener;
    plan->T.gc = d_gc; plan->T.k3 = d_k3; plan->T.yee = d_yee;

    // the dynamic LDS beside the kernel's static LDS must fit one workgroup
    int dev = 0, cus = 0, per_cu = 0, max_lds = 0;
    ZPIC_HIP_TRY( hipGetDevice( &dev ) );
    ZPIC_HIP_TRY( hipDeviceGetAttribute( &max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev ) );
    hipFuncAttributes fa{};
    if ( block == 256 ) ZPIC_HIP_TRY( hipFuncGetAttributes( &fa, reinterpret_cast<const void *>( &k_batch_step<256> ) ) );
    else ZPIC_HIP_TRY( hipFuncGetAttributes( &fa, reinterpret_cast<const void *>( &k_batch_step<512> ) ) );
    if ( lds + fa.sharedSizeBytes > (size_t) max_lds )
        return set_error( ZPIC_ERANGE, "persistent step: a tile's LDS staging does not fit one workgroup" );
    // as many workgroups as fit on the device at once
    ZPIC_HIP_TRY( hipDeviceGetAttribute( &cus, hipDeviceAttributeMultiprocessorCount, dev ) );
    if ( block == 256 )
        ZPIC_HIP_TRY( hipOccupancyMaxActiveBlocksPerMultiprocessor( &per_cu, k_batch_step<256>, 256, lds ) );
    else
        ZPIC_HIP_TRY( hipOccupancyMaxActiveBlocksPerMultiprocessor( &per_cu, k_batch_step<512>, 512, lds ) );
    plan->grid = std::max( 1, std::min( cus * std::max( per_cu, 1 ), (int) items.size() ) );
    *out = plan.release();
    return ZPIC_OK;
}

int batch_step_run( BatchStepPlan * plan, void * stream ) {
    hipStream_t st = as_stream( stream );
    hipLaunchKernelGGL( k_step_reset, dim3( ( plan->nsync + 255 ) / 256 ), dim3( 256 ), 0, st, plan->T.next, plan->nsync );
    ZPIC_LAUNCH_CHECK( "k_step_reset" );
    if ( plan->block == 256 )
        hipLaunchKernelGGL( k_batch_step<256>, dim3( plan->grid ), dim3( 256 ), plan->lds, st, plan->T );
    else
        hipLaunchKernelGGL( k_batch_step<512>, dim3( plan->grid ), dim3( 512 ), plan->lds, st, plan->T );
    ZPIC_LAUNCH_CHECK( "k_batch_step" );
    return ZPIC_OK;
}

int batch_step_status( BatchStepPlan * plan ) {
    int err = 0;
    ZPIC_HIP_TRY( hipDeviceSynchronize() );
    ZPIC_HIP_TRY( hipMemcpy( &err, plan->T.error, sizeof( int ), hipMemcpyDeviceToHost ) );
    if ( err ) {
        const int zero = 0;
        ZPIC_HIP_TRY( hipMemcpy( plan->T.error, &zero, sizeof( int ), hipMemcpyHostToDevice ) );
        return set_error( ZPIC_EHIP, "persistent step: a dependency wait exceeded its spin limit" );
    }
    return ZPIC_OK;
}

}  // namespace

// The particle pass's HBM traffic without its arithmetic: every particle
// record (int2 cell, float2 position, float3 momentum: 28 B) read and written
// back changed, grid-stride.  A measurement of the ceiling that traffic
// allows on the device, not part of the timestep.
__global__ void __launch_bounds__( 256 ) k_stream_probe( int2 * __restrict__ ix, float2 * __restrict__ x,
                                                         float3 * __restrict__ u, size_t n ) {
    const gptr<int2> gix = gbl( ix );
    const gptr<float2> gx = gbl( x );
    const gptr<float3> gu = gbl( u );
    for ( size_t i = blockIdx.x * (size_t) blockDim.x + threadIdx.x; i < n; i += (size_t) gridDim.x * blockDim.x ) {
        int2 c = gix[i];
        float2 p = gx[i];
        float3 v = gu[i];
        c.x ^= 1;
        p.x = p.x + 1.0f;
        v.z = v.z + 1.0f;
        gix[i] = c;
        gx[i] = p;
        gu[i] = v;
    }
}

// Device copy of n 16-byte words with U independent loads in flight per
// lane: a workgroup of 256 lanes moves 256 U words per iteration (lane l
// takes words l, l + 256, ..., contiguous 4 KiB runs per load instruction),
// grid-stride over the buffer.  Measurement only.
template <int U>
__global__ void __launch_bounds__( 256 ) k_copy_probe( const float4 * __restrict__ src, float4 * __restrict__ dst,
                                                       size_t n ) {
    const gptr<const float4> s = gbl( src );
    const gptr<float4> d = gbl( dst );
    const size_t step = (size_t) gridDim.x * 256 * U;
    for ( size_t base = (size_t) blockIdx.x * 256 * U + threadIdx.x; base < n; base += step ) {
        float4 v[U];
#pragma unroll
        for ( int k = 0; k < U; k++ ) {
            const size_t i = base + (size_t) k * 256;
            v[k] = i < n ? (float4) s[i] : make_float4( 0.f, 0.f, 0.f, 0.f );
        }
#pragma unroll
        for ( int k = 0; k < U; k++ ) {
            const size_t i = base + (size_t) k * 256;
            if ( i < n ) d[i] = v[k];
        }
    }
}

// The particle records with U records in flight per lane (same access order
// as k_copy_probe, read and written back modified).  Measurement only.
template <int U>
__global__ void __launch_bounds__( 256 ) k_stream_probe_u( int2 * __restrict__ ix, float2 * __restrict__ x,
                                                           float3 * __restrict__ u, size_t n ) {
    const gptr<int2> gix = gbl( ix );
    const gptr<float2> gx = gbl( x );
    const gptr<float3> gu = gbl( u );
    const size_t step = (size_t) gridDim.x * 256 * U;
    for ( size_t base = (size_t) blockIdx.x * 256 * U + threadIdx.x; base < n; base += step ) {
        int2 c[U];
        float2 p[U];
        float3 v[U];
#pragma unroll
        for ( int k = 0; k < U; k++ ) {
            const size_t i = min( base + (size_t) k * 256, n - 1 );
            c[k] = gix[i];
            p[k] = gx[i];
            v[k] = gu[i];
        }
#pragma unroll
        for ( int k = 0; k < U; k++ ) {
            const size_t i = base + (size_t) k * 256;
            if ( i < n ) {
                c[k].x ^= 1;
                p[k].x = p[k].x + 1.0f;
                v[k].z = v[k].z + 1.0f;
                gix[i] = c[k];
                gx[i] = p[k];
                gu[i] = v[k];
            }
        }
    }
}

// The particle records of `gridDim.x` tiles of tnp records each, read and
// written back changed, one 256-lane workgroup per tile, U iterations of 64
// records in flight per wave.  ROUNDS = false: wave w streams the contiguous
// quarter [w c, (w + 1) c) of its tile (the particle pass's order); ROUNDS =
// true: the four waves walk the tile together, 256 records per round.
// Measurement only.
template <int U, bool ROUNDS>
__global__ void __launch_bounds__( 256 ) k_tile_probe( int2 * __restrict__ ix, float2 * __restrict__ x,
                                                       float3 * __restrict__ u, int tnp ) {
    const size_t base = (size_t) blockIdx.x * tnp;
    const gptr<int2> gix = gbl( ix + base );
    const gptr<float2> gx = gbl( x + base );
    const gptr<float3> gu = gbl( u + base );
    const int w = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    const int chunk = ROUNDS ? tnp : ( ( ( tnp + 3 ) / 4 + WAVE - 1 ) & ~( WAVE - 1 ) );
    const int beg = ROUNDS ? w * WAVE : w * chunk;
    const int end = ROUNDS ? tnp : min( tnp, beg + chunk );
    const int step = ROUNDS ? 4 * WAVE : WAVE;
    for ( int i0 = beg; i0 < end; i0 += step * U ) {
        int2 c[U];
        float2 p[U];
        float3 v[U];
#pragma unroll
        for ( int k = 0; k < U; k++ ) {
            const int i = min( i0 + k * step + lane, end - 1 );
            c[k] = gix[i];
            p[k] = gx[i];
            v[k] = gu[i];
        }
#pragma unroll
        for ( int k = 0; k < U; k++ ) {
            const int i = i0 + k * step + lane;
            if ( i < end ) {
                c[k].x ^= 1;
                p[k].x = p[k].x + 1.0f;
                v[k].z = v[k].z + 1.0f;
                gix[i] = c[k];
                gx[i] = p[k];
                gu[i] = v[k];
            }
        }
    }
}

}  // namespace zpic

using namespace zpic;

extern "C" {

int zpic_push( const zpic_particles * p, const zpic_geom * fld, const float * E, const float * B, float alpha,
               double * tile_energy, int deterministic, void * stream ) {
    MoveArgs m{};
    return launch_particles<true, false>( p, fld, E, B, nullptr, alpha, m, tile_energy, nullptr, deterministic, stream );
}

int zpic_move_deposit( const zpic_particles * p, const zpic_geom * fld, float * J, float dt_dx, float dt_dy, float q,
                       float qnx, float qny, zpic_sort * s, int deterministic, void * stream ) {
    MoveArgs m{ dt_dx, dt_dy, q, qnx, qny };
    return launch_particles<false, true>( p, fld, nullptr, nullptr, J, 0.f, m, nullptr, s, deterministic, stream );
}

int zpic_push_move_deposit( const zpic_particles * p, const zpic_geom * fld, const float * E, const float * B,
                            float * J, float alpha, float dt_dx, float dt_dy, float q, float qnx, float qny,
                            double * tile_energy, zpic_sort * s, int deterministic, void * stream ) {
    MoveArgs m{ dt_dx, dt_dy, q, qnx, qny };
    return launch_particles<true, true>( p, fld, E, B, J, alpha, m, tile_energy, s, deterministic, stream );
}

int zpic_push_move_deposit_bc( const zpic_particles * p, const zpic_geom * fld, const float * E, const float * B,
                               float * J, float alpha, float dt_dx, float dt_dy, float q, float qnx, float qny,
                               double * tile_energy, zpic_sort * s, int reflect, int deterministic, void * stream ) {
    if ( reflect < 0 || reflect > 15 ) return set_error( ZPIC_EINVAL, "reflect mask must be in [0, 15]" );
    MoveArgs m{ dt_dx, dt_dy, q, qnx, qny, reflect };
    return launch_particles<true, true>( p, fld, E, B, J, alpha, m, tile_energy, s, deterministic, stream );
}

int zpic_push_move_deposit_shift( const zpic_particles * p, const zpic_geom * fld, const float * E, const float * B,
                                  float * J, float alpha, float dt_dx, float dt_dy, float q, float qnx, float qny,
                                  double * tile_energy, zpic_sort * s, int reflect, int shift_x, int shift_y,
                                  int deterministic, void * stream ) {
    if ( reflect < 0 || reflect > 15 ) return set_error( ZPIC_EINVAL, "reflect mask must be in [0, 15]" );
    if ( shift_x < -1 || shift_x > 1 || shift_y < -1 || shift_y > 1 )
        return set_error( ZPIC_EINVAL, "cell shift must be -1, 0 or 1 per axis" );
    MoveArgs m{ dt_dx, dt_dy, q, qnx, qny, reflect, shift_x, shift_y };
    return launch_particles<true, true>( p, fld, E, B, J, alpha, m, tile_energy, s, deterministic, stream );
}

int zpic_push_move_deposit_overwrite( const zpic_particles * p, const zpic_geom * fld, const float * E,
                                      const float * B, float * J, float alpha, float dt_dx, float dt_dy, float q,
                                      float qnx, float qny, double * tile_energy, zpic_sort * s, int reflect,
                                      int shift_x, int shift_y, void * stream ) {
    if ( reflect < 0 || reflect > 15 ) return set_error( ZPIC_EINVAL, "reflect mask must be in [0, 15]" );
    if ( shift_x < -1 || shift_x > 1 || shift_y < -1 || shift_y > 1 )
        return set_error( ZPIC_EINVAL, "cell shift must be -1, 0 or 1 per axis" );
    MoveArgs m{ dt_dx, dt_dy, q, qnx, qny, reflect, shift_x, shift_y };
    return launch_particles<true, true>( p, fld, E, B, J, alpha, m, tile_energy, s, 0, stream, 1 );
}

int zpic_tile_sort_count( zpic_particles * p, zpic_particles * tmp, zpic_sort * s, int ntiles_x, int ntiles_y, int nx,
                          int ny, int periodic_x, int periodic_y, const int * extra, void * stream ) {
    if ( !p || !tmp || !s || ntiles_x < 1 || ntiles_y < 1 || nx < 1 || ny < 1 )
        return set_error( ZPIC_EINVAL, "invalid tile_sort arguments" );
    const int nt = ntiles_x * ntiles_y;
    SortArgs a = sort_args( p, tmp, s, ntiles_x, ntiles_y, nx, ny, periodic_x, periodic_y );
    a.extra = extra;
    const SortArgs * tab = nullptr;
    hipStream_t st = as_stream( stream );
    hipLaunchKernelGGL( k_sort_newnp, dim3( ( nt + 255 ) / 256 ), dim3( 256 ), 0, st, a, tab );
    ZPIC_LAUNCH_CHECK( "k_sort_newnp" );
    hipLaunchKernelGGL( k_sort_count, dim3( 1 ), dim3( SCAN_BLOCK ), 0, st, a, tab );
    ZPIC_LAUNCH_CHECK( "k_sort_count" );
    return ZPIC_OK;
}

int zpic_tile_sort_copy( zpic_particles * p, zpic_particles * tmp, zpic_sort * s, int ntiles_x, int ntiles_y, int nx,
                         int ny, int periodic_x, int periodic_y, int mode, void * stream ) {
    if ( !p || !tmp || !s || ntiles_x < 1 || ntiles_y < 1 || nx < 1 || ny < 1 )
        return set_error( ZPIC_EINVAL, "invalid tile_sort arguments" );
    if ( mode != 0 && mode != 1 ) return set_error( ZPIC_EINVAL, "tile_sort copy mode must be 0 or 1" );
    const int nt = ntiles_x * ntiles_y;
    const SortArgs a = sort_args( p, tmp, s, ntiles_x, ntiles_y, nx, ny, periodic_x, periodic_y );
    const SortArgs * tab = nullptr;
    const int2 * map = nullptr;
    hipStream_t st = as_stream( stream );
    if ( mode == 1 ) {
        // copy_sorted reads every particle in place: not after a particle
        // pass that staged leavers
        if ( a.S.stage_ix ) return set_error( ZPIC_EINVAL, "copy_sorted after a staging particle pass" );
        hipLaunchKernelGGL( k_sort_out<true>, dim3( nt ), dim3( SORT_BLOCK ), 0, st, a, tab, map );
        ZPIC_LAUNCH_CHECK( "k_sort_out (copy_sorted)" );
        return ZPIC_OK;
    }
    if ( a.S.stage_ix ) {
        // order-free placement from the particle pass's staging
        const dim3 grid( ( nt + SORT_WAVES - 1 ) / SORT_WAVES );
        hipLaunchKernelGGL( k_sort_shift, grid, dim3( SORT_WAVES * WAVE ), 0, st, a, tab, map, nt );
        ZPIC_LAUNCH_CHECK( "k_sort_shift" );
        hipLaunchKernelGGL( k_sort_in_fast, grid, dim3( SORT_WAVES * WAVE ), 0, st, a, tab, map, nt );
        ZPIC_LAUNCH_CHECK( "k_sort_in_fast" );
        return ZPIC_OK;
    }
    hipLaunchKernelGGL( k_sort_out<false>, dim3( nt ), dim3( SORT_BLOCK ), 0, st, a, tab, map );
    ZPIC_LAUNCH_CHECK( "k_sort_out" );
    hipLaunchKernelGGL( k_sort_in, dim3( nt ), dim3( SORT_BLOCK ), 0, st, a, tab, map );
    ZPIC_LAUNCH_CHECK( "k_sort_in" );
    return ZPIC_OK;
}

int zpic_tile_sort( zpic_particles * p, zpic_particles * tmp, zpic_sort * s, int ntiles_x, int ntiles_y, int nx,
                    int ny, int periodic_x, int periodic_y, void * stream ) {
    if ( int rc = zpic_tile_sort_count( p, tmp, s, ntiles_x, ntiles_y, nx, ny, periodic_x, periodic_y, nullptr,
                                        stream ) )
        return rc;
    return zpic_tile_sort_copy( p, tmp, s, ntiles_x, ntiles_y, nx, ny, periodic_x, periodic_y, 0, stream );
}

int zpic_set_threads_per_tile( int threads ) {
    if ( int rc = check_threads( threads ) ) return rc;
    g_threads = threads;
    return ZPIC_OK;
}

int zpic_get_threads_per_tile( void ) { return g_threads; }

int zpic_set_sort_order( int reference ) {
    if ( reference != 0 && reference != 1 ) return set_error( ZPIC_EINVAL, "sort order must be 0 or 1" );
    g_reference_order = reference;
    return ZPIC_OK;
}

int zpic_get_sort_order( void ) { return g_reference_order; }

int zpic_fitted_threads_per_tile( int threads, int tile_vol ) {
    if ( threads < WAVE || threads > FAST_BLOCK || threads % WAVE || tile_vol < 1 ) return 0;
    return fitted_threads( threads, tile_vol );
}

struct zpic_batch_particles {
    BatchParticlePlan * plan;
};

int zpic_batch_particles_create( zpic_batch_particles ** out, const zpic_batch_species * items, int n,
                                 int threads_per_tile ) {
    if ( !out ) return set_error( ZPIC_EINVAL, "null output" );
    BatchParticlePlan * plan = nullptr;
    if ( int rc = batch_particles_create( &plan, items, n, threads_per_tile ) ) return rc;
    *out = new zpic_batch_particles{ plan };
    return ZPIC_OK;
}

int zpic_batch_particles_advance( zpic_batch_particles * b, void * stream ) {
    if ( !b || !b->plan ) return set_error( ZPIC_EINVAL, "null batch" );
    return batch_particles_step( b->plan, stream );
}

void zpic_batch_particles_free( zpic_batch_particles * b ) {
    if ( !b ) return;
    batch_particles_destroy( b->plan );
    delete b;
}

struct zpic_batch_step {
    BatchStepPlan * plan;
};

int zpic_batch_step_create( zpic_batch_step ** out, const zpic_batch_species * species, int nslots,
                            const zpic_batch_fields_item * fields, int n, int threads_per_tile ) {
    if ( !out ) return set_error( ZPIC_EINVAL, "null output" );
    BatchStepPlan * plan = nullptr;
    if ( int rc = batch_step_create( &plan, species, nslots, fields, n, threads_per_tile ) ) return rc;
    *out = new zpic_batch_step{ plan };
    return ZPIC_OK;
}

int zpic_batch_step_advance( zpic_batch_step * b, void * stream ) {
    if ( !b || !b->plan ) return set_error( ZPIC_EINVAL, "null batch step" );
    return batch_step_run( b->plan, stream );
}

int zpic_batch_step_status( zpic_batch_step * b ) {
    if ( !b || !b->plan ) return set_error( ZPIC_EINVAL, "null batch step" );
    return batch_step_status( b->plan );
}

int zpic_batch_step_info( const zpic_batch_step * b, int * workgroups, int * threads, int * items ) {
    if ( !b || !b->plan ) return set_error( ZPIC_EINVAL, "null batch step" );
    if ( workgroups ) *workgroups = b->plan->grid;
    if ( threads ) *threads = b->plan->block;
    if ( items ) *items = b->plan->T.qoff[STEP_QUEUES];
    return ZPIC_OK;
}

void zpic_batch_step_free( zpic_batch_step * b ) {
    if ( !b ) return;
    delete b->plan;
    delete b;
}

int zpic_deposit_charge( const zpic_particles * p, const zpic_geom * g, float * rho, float q, int deterministic,
                         void * stream ) {
    if ( !p || !g || !rho ) return set_error( ZPIC_EINVAL, "null buffer" );
    const int nt = g->ntiles_x * g->ntiles_y;
    const size_t lds = sizeof( float ) * g->tile_vol;
    if ( lds > 65536 ) return set_error( ZPIC_ERANGE, "tile too large for LDS staging" );
    hipStream_t st = as_stream( stream );
    if ( deterministic )
        hipLaunchKernelGGL( k_charge<true>, dim3( nt ), dim3( WAVE ), lds, st, view( p ), make_geo( g ), rho, q );
    else
        hipLaunchKernelGGL( k_charge<false>, dim3( nt ), dim3( FAST_BLOCK ), lds, st, view( p ), make_geo( g ), rho, q );
    ZPIC_LAUNCH_CHECK( "k_charge" );
    return ZPIC_OK;
}

int zpic_particle_stream_probe( int * ix, float * x, float * u, size_t n, void * stream ) {
    if ( ( !ix || !x || !u ) && n ) return set_error( ZPIC_EINVAL, "null buffer" );
    if ( n == 0 ) return ZPIC_OK;
    size_t blocks = ( n + 255 ) / 256;
    if ( blocks > 65536 ) blocks = 65536;
    hipLaunchKernelGGL( k_stream_probe, dim3( blocks ), dim3( 256 ), 0, as_stream( stream ), (int2 *) ix, (float2 *) x,
                        (float3 *) u, n );
    ZPIC_LAUNCH_CHECK( "k_stream_probe" );
    return ZPIC_OK;
}

int zpic_bandwidth_probe( int kind, void * a, void * b, void * c, size_t n, int loads_per_lane, int blocks,
                          void * stream ) {
    if ( kind < 0 || kind > 3 )
        return set_error( ZPIC_EINVAL, "probe kind must be 0 (copy), 1 (particle records), 2 or 3 (tiles)" );
    if ( ( !a || !b || ( kind >= 1 && !c ) ) && n ) return set_error( ZPIC_EINVAL, "null buffer" );
    if ( blocks < 1 ) return set_error( ZPIC_EINVAL, "blocks must be positive" );
    if ( n == 0 ) return ZPIC_OK;
    hipStream_t st = as_stream( stream );
    if ( kind >= 2 ) {
        // blocks = tiles of n / blocks records each
        if ( n % (size_t) blocks || n / (size_t) blocks > ( 1u << 30 ) )
            return set_error( ZPIC_EINVAL, "tile probe: n must be a multiple of blocks" );
        const int tnp = (int) ( n / (size_t) blocks );
#define ZPIC_TPROBE( U )                                                                                               \
    case U:                                                                                                            \
        if ( kind == 2 )                                                                                               \
            hipLaunchKernelGGL( ( k_tile_probe<U, false> ), dim3( blocks ), dim3( 256 ), 0, st, (int2 *) a,            \
                                (float2 *) b, (float3 *) c, tnp );                                                     \
        else                                                                                                           \
            hipLaunchKernelGGL( ( k_tile_probe<U, true> ), dim3( blocks ), dim3( 256 ), 0, st, (int2 *) a,             \
                                (float2 *) b, (float3 *) c, tnp );                                                     \
        break;
        switch ( loads_per_lane ) {
            ZPIC_TPROBE( 1 )
            ZPIC_TPROBE( 2 )
            ZPIC_TPROBE( 4 )
        default: return set_error( ZPIC_EINVAL, "loads_per_lane must be 1, 2 or 4 for tile probes" );
        }
#undef ZPIC_TPROBE
        ZPIC_LAUNCH_CHECK( "tile probe" );
        return ZPIC_OK;
    }
#define ZPIC_PROBE( U )                                                                                                \
    case U:                                                                                                            \
        if ( kind == 0 )                                                                                               \
            hipLaunchKernelGGL( k_copy_probe<U>, dim3( blocks ), dim3( 256 ), 0, st, (const float4 *) a, (float4 *) b, \
                                n );                                                                                   \
        else                                                                                                           \
            hipLaunchKernelGGL( k_stream_probe_u<U>, dim3( blocks ), dim3( 256 ), 0, st, (int2 *) a, (float2 *) b,     \
                                (float3 *) c, n );                                                                     \
        break;
    switch ( loads_per_lane ) {
        ZPIC_PROBE( 1 )
        ZPIC_PROBE( 2 )
        ZPIC_PROBE( 4 )
        ZPIC_PROBE( 8 )
    default: return set_error( ZPIC_EINVAL, "loads_per_lane must be 1, 2, 4 or 8" );
    }
#undef ZPIC_PROBE
    ZPIC_LAUNCH_CHECK( "bandwidth probe" );
    return ZPIC_OK;
}

}  // extern "C"
