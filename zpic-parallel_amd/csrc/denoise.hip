// Learned current denoiser (no reference counterpart): a 3-layer 5x5
// convolutional network, 3 -> 16 -> 16 -> 3 channels with ReLU between
// layers (8,835 parameters), applied to the periodic global current J
// (interior cells, [y][x][3]) in place of the binomial filter.
//
// Every convolution is an implicit GEMM on the gfx950 matrix cores
// (v_mfma_f32_16x16x32_bf16: BF16 operands, FP32 accumulation):
//
//   forward / input gradient   D[co][pixel] = W[co][k] x im2col[k][pixel]
//       16 output channels x 16 pixels per instruction, k = tap * C + ci
//       (C = 8 for the 3-channel input padded to 8, 16 otherwise), the
//       weight operand held in registers for the whole launch, the image
//       operand loaded as one 16-byte vector per lane (activations are
//       pixel-major [y][x][C] BF16, periodic wrap in the index);
//   weight gradient            D[co][kcol] = dZ[co][pixel] x im2col[pixel][kcol]
//       16 output channels x 16 weight columns, 32 pixels per instruction,
//       operands staged in LDS for one image row at a time; one partial
//       sum per row, reduced in row order (deterministic).
//
// Operand lane layouts of v_mfma_f32_16x16x32_bf16 (wave64): lane l holds
// A[row l%16][k (l/16)*8 .. +7], B[k (l/16)*8 .. +7][col l%16], and the
// result D[row (l/16)*4 + i][col l%16], i = 0..3.
//
// Parameters (FP32 master copy, the layout of the C ABI): W1[16][3][5][5],
// b1[16], W2[16][16][5][5], b2[16], W3[3][16][5][5], b3[3].  Training is
// Adam on the FP32 master copy; the BF16 operand packs are rebuilt after
// every update.
#include "common.h"

#include <algorithm>
#include <cmath>
#include <random>
#include <vector>

namespace zpic {
namespace {

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__( ( ext_vector_type( 8 ) ) );
typedef float f32x4 __attribute__( ( ext_vector_type( 4 ) ) );

constexpr int WAVE = 64;
constexpr int NPARAM = 8835;
constexpr int OFF_W1 = 0, OFF_B1 = 1200, OFF_W2 = 1216, OFF_B2 = 7616, OFF_W3 = 7632, OFF_B3 = 8832;
constexpr int KS1 = 7;   // layer 1: 25 taps x 8 channels = 200 -> 224 = 7 x 32
constexpr int KS2 = 13;  // 16-channel layers: 25 taps x 16 = 400 -> 416 = 13 x 32
constexpr int PACK_FWD1 = 0, PACK_FWD2 = 1, PACK_FWD3 = 2, PACK_BWD3 = 3, PACK_BWD2 = 4, NPACK = 5;

__host__ __device__ inline int pack_ksteps( int which ) { return which == PACK_FWD1 ? KS1 : KS2; }

// Weight-operand value of pack `which` for row (output channel) r and k.
__device__ inline float pack_value( const float * w, int which, int r, int k ) {
    if ( which == PACK_FWD1 ) {
        const int tap = k / 8, ci = k % 8;
        return ( tap < 25 && ci < 3 ) ? w[OFF_W1 + ( r * 3 + ci ) * 25 + tap] : 0.0f;
    }
    const int tap = k / 16, c = k % 16;
    if ( tap >= 25 ) return 0.0f;
    switch ( which ) {
    case PACK_FWD2: return w[OFF_W2 + ( r * 16 + c ) * 25 + tap];
    case PACK_FWD3: return r < 3 ? w[OFF_W3 + ( r * 16 + c ) * 25 + tap] : 0.0f;
    // input gradients: rows are the layer's input channels, k runs over
    // (flipped tap, output channel)
    case PACK_BWD3: return c < 3 ? w[OFF_W3 + ( c * 16 + r ) * 25 + ( 24 - tap )] : 0.0f;
    default: return w[OFF_W2 + ( c * 16 + r ) * 25 + ( 24 - tap )];
    }
}

// pack[which][ks][lane][8]: lane l holds row l%16, k = ks*32 + (l/16)*8 + j
__global__ void k_dn_pack( const float * __restrict__ w, bf16 * __restrict__ pack ) {
    const int which = blockIdx.y;
    const int ks = blockIdx.x;
    if ( ks >= pack_ksteps( which ) ) return;
    const int l = threadIdx.x;
    bf16 * dst = pack + ( (size_t) which * KS2 + ks ) * WAVE * 8 + l * 8;
    for ( int j = 0; j < 8; j++ ) dst[j] = (bf16) pack_value( w, which, l % 16, ks * 32 + ( l / 16 ) * 8 + j );
}

// Convolution D[co][p] = sum_k W[co][k] im2col[k][p] (+ bias, ReLU, mask):
// CIN input channels per pixel (8 or 16), KS k-steps.  Each wave owns blocks
// of 16 consecutive pixels (grid-stride).  Output: BF16 [P][16], or FP32
// [P][cout] times out_scale.
template <int CIN, int KS, bool RELU, bool MASK, bool F32OUT>
__global__ void __launch_bounds__( 256 ) k_dn_conv( const bf16 * __restrict__ in, const bf16 * __restrict__ pack,
                                                     const float * __restrict__ bias, int cout,
                                                     const bf16 * __restrict__ mask, void * __restrict__ out, int H,
                                                     int W, float out_scale ) {
    const int l = threadIdx.x % WAVE;
    const int g = l / 16;
    bf16x8 wa[KS];
    for ( int ks = 0; ks < KS; ks++ ) wa[ks] = *reinterpret_cast<const bf16x8 *>( pack + ( ks * WAVE + l ) * 8 );
    float bv[4];
    for ( int i = 0; i < 4; i++ ) {
        const int co = g * 4 + i;
        bv[i] = ( bias && co < cout ) ? bias[co] : 0.0f;
    }
    const int P = H * W;
    const int nblk = P / 16;
    const int waves = gridDim.x * ( blockDim.x / WAVE );
    for ( int blk = blockIdx.x * ( blockDim.x / WAVE ) + threadIdx.x / WAVE; blk < nblk; blk += waves ) {
        const int p = blk * 16 + l % 16;
        const int y = p / W, x = p - y * W;
        f32x4 acc = { 0.f, 0.f, 0.f, 0.f };
        for ( int ks = 0; ks < KS; ks++ ) {
            int tap, c0;
            if constexpr ( CIN == 16 ) {
                tap = ks * 2 + g / 2;
                c0 = ( g % 2 ) * 8;
            } else {
                tap = ks * 4 + g;
                c0 = 0;
            }
            bf16x8 b;
            if ( tap < 25 ) {
                int yy = y + tap / 5 - 2, xx = x + tap % 5 - 2;
                yy += ( yy < 0 ) ? H : ( yy >= H ? -H : 0 );
                xx += ( xx < 0 ) ? W : ( xx >= W ? -W : 0 );
                b = *reinterpret_cast<const bf16x8 *>( in + ( (size_t) yy * W + xx ) * CIN + c0 );
            } else {
                for ( int j = 0; j < 8; j++ ) b[j] = (bf16) 0.0f;
            }
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16( wa[ks], b, acc, 0, 0, 0 );
        }
        for ( int i = 0; i < 4; i++ ) {
            const int co = g * 4 + i;
            float v = acc[i] + bv[i];
            if constexpr ( RELU ) v = v > 0.0f ? v : 0.0f;
            if constexpr ( MASK ) v = ( (float) mask[(size_t) p * 16 + co] > 0.0f ) ? v : 0.0f;
            if constexpr ( F32OUT ) {
                if ( co < cout ) reinterpret_cast<float *>( out )[(size_t) p * cout + co] = v * out_scale;
            } else {
                reinterpret_cast<bf16 *>( out )[(size_t) p * 16 + co] = (bf16) v;
            }
        }
    }
}

// Weight and bias gradients of one layer for one image row per workgroup
// (one wave): part[row][co][kcol] and part[row][co][NK + 0] (bias).  CIN
// input channels, NKB blocks of 16 weight columns (k = tap*CIN + ci).
template <int CIN, int NKB>
__global__ void __launch_bounds__( WAVE ) k_dn_wgrad( const bf16 * __restrict__ in, const bf16 * __restrict__ dz,
                                                      float * __restrict__ part, int H, int W ) {
    constexpr int SPAN = 32 + 4;  // pixels of a strip plus the 2-cell halo on each side
    __shared__ bf16 s_in[5][SPAN][CIN];
    __shared__ bf16 s_dz[32][16];
    const int y = blockIdx.x;
    const int l = threadIdx.x;
    const int g = l / 16, r = l % 16;
    f32x4 acc[NKB];
    for ( int kb = 0; kb < NKB; kb++ ) acc[kb] = f32x4{ 0.f, 0.f, 0.f, 0.f };
    float dbias = 0.0f;
    for ( int x0 = 0; x0 < W; x0 += 32 ) {
        __syncthreads();
        for ( int e = l; e < 5 * SPAN * CIN; e += WAVE ) {
            const int c = e % CIN, t = e / CIN, sx = t % SPAN, sy = t / SPAN;
            int yy = y + sy - 2, xx = x0 + sx - 2;
            yy += ( yy < 0 ) ? H : ( yy >= H ? -H : 0 );
            xx += ( xx < 0 ) ? W : ( xx >= W ? -W : 0 );
            s_in[sy][sx][c] = in[( (size_t) yy * W + xx ) * CIN + c];
        }
        for ( int e = l; e < 32 * 16; e += WAVE ) s_dz[e / 16][e % 16] = dz[( (size_t) y * W + x0 ) * 16 + e];
        __syncthreads();
        // A = dz^T: row co = r, pixels g*8 .. g*8+7
        bf16x8 a;
        for ( int j = 0; j < 8; j++ ) a[j] = s_dz[g * 8 + j][r];
        for ( int j = 0; j < 8; j++ ) dbias += (float) a[j];
        for ( int kb = 0; kb < NKB; kb++ ) {
            const int k = kb * 16 + r;  // this lane's weight column
            const int tap = k / CIN, ci = k % CIN;
            bf16x8 b;
            if ( tap < 25 ) {
                const int dy = tap / 5, dx = tap % 5;
                for ( int j = 0; j < 8; j++ ) b[j] = s_in[dy][g * 8 + j + dx][ci];
            } else {
                for ( int j = 0; j < 8; j++ ) b[j] = (bf16) 0.0f;
            }
            acc[kb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16( a, b, acc[kb], 0, 0, 0 );
        }
    }
    // D[co = g*4+i][kcol = kb*16 + r]
    constexpr int NK = NKB * 16;
    float * dst = part + (size_t) y * 16 * ( NK + 1 );
    for ( int kb = 0; kb < NKB; kb++ )
        for ( int i = 0; i < 4; i++ ) dst[( g * 4 + i ) * ( NK + 1 ) + kb * 16 + r] = acc[kb][i];
    // bias: lanes r (same co) over the four pixel groups
    dbias += __shfl_xor( dbias, 16 );
    dbias += __shfl_xor( dbias, 32 );
    if ( g == 0 ) dst[r * ( NK + 1 ) + NK] = dbias;
}

// Sums the per-row partials in row order into the FP32 gradient (layout of
// the parameters): layer 1 (CIN 8, 13 blocks), 2 and 3 (CIN 16, 25 blocks).
__global__ void k_dn_wgrad_reduce( const float * __restrict__ part1, const float * __restrict__ part2,
                                   const float * __restrict__ part3, int H, float * __restrict__ grad ) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if ( i >= NPARAM ) return;
    const float * part;
    int nk, co, col;
    if ( i < OFF_B1 ) {  // W1[co][ci][tap] <- column tap*8 + ci
        part = part1; nk = 13 * 16;
        co = i / 75; const int ci = ( i / 25 ) % 3, tap = i % 25;
        col = tap * 8 + ci;
    } else if ( i < OFF_W2 ) {
        part = part1; nk = 13 * 16; co = i - OFF_B1; col = nk;
    } else if ( i < OFF_B2 ) {
        const int j = i - OFF_W2;
        part = part2; nk = 25 * 16;
        co = j / 400; const int ci = ( j / 25 ) % 16, tap = j % 25;
        col = tap * 16 + ci;
    } else if ( i < OFF_W3 ) {
        part = part2; nk = 25 * 16; co = i - OFF_B2; col = nk;
    } else if ( i < OFF_B3 ) {
        const int j = i - OFF_W3;
        part = part3; nk = 25 * 16;
        co = j / 400; const int ci = ( j / 25 ) % 16, tap = j % 25;
        col = tap * 16 + ci;
    } else {
        part = part3; nk = 25 * 16; co = i - OFF_B3; col = nk;
    }
    double s = 0.0;
    for ( int y = 0; y < H; y++ ) s += part[( (size_t) y * 16 + co ) * ( nk + 1 ) + col];
    grad[i] = (float) s;
}

// Gathers the interior of a tile-major float3 grid into [y][x][8] BF16
// (channels 3..7 zero) scaled by inv_scale, and optionally a FP32 [y][x][3]
// copy (unscaled).
__global__ void k_dn_gather( const float * __restrict__ J, zpic_geom g, float inv_scale, bf16 * __restrict__ x0,
                             float * __restrict__ f32 ) {
    const int W = g.nx * g.ntiles_x, H = g.ny * g.ntiles_y;
    const size_t P = (size_t) W * H;
    for ( size_t p = blockIdx.x * (size_t) blockDim.x + threadIdx.x; p < P; p += (size_t) gridDim.x * blockDim.x ) {
        const int y = p / W, x = p % W;
        const int t = ( y / g.ny ) * g.ntiles_x + x / g.nx;
        const float * c = J + ( (size_t) t * g.tile_vol + g.offset + ( y % g.ny ) * g.ext_x + x % g.nx ) * 3;
        const float v0 = c[0], v1 = c[1], v2 = c[2];
        if ( x0 ) {
            bf16 * d = x0 + p * 8;
            d[0] = (bf16) ( v0 * inv_scale );
            d[1] = (bf16) ( v1 * inv_scale );
            d[2] = (bf16) ( v2 * inv_scale );
            for ( int k = 3; k < 8; k++ ) d[k] = (bf16) 0.0f;
        }
        if ( f32 ) {
            f32[p * 3] = v0;
            f32[p * 3 + 1] = v1;
            f32[p * 3 + 2] = v2;
        }
    }
}

// [y][x][3] FP32 into the interior of a tile-major float3 grid.
__global__ void k_dn_scatter( const float * __restrict__ v, zpic_geom g, float * __restrict__ J ) {
    const int W = g.nx * g.ntiles_x, H = g.ny * g.ntiles_y;
    const size_t P = (size_t) W * H;
    for ( size_t p = blockIdx.x * (size_t) blockDim.x + threadIdx.x; p < P; p += (size_t) gridDim.x * blockDim.x ) {
        const int y = p / W, x = p % W;
        const int t = ( y / g.ny ) * g.ntiles_x + x / g.nx;
        float * c = J + ( (size_t) t * g.tile_vol + g.offset + ( y % g.ny ) * g.ext_x + x % g.nx ) * 3;
        c[0] = v[p * 3];
        c[1] = v[p * 3 + 1];
        c[2] = v[p * 3 + 2];
    }
}

// [y][x][3] FP32 -> [y][x][8] BF16 scaled
__global__ void k_dn_input( const float * __restrict__ v, size_t P, float inv_scale, bf16 * __restrict__ x0 ) {
    for ( size_t p = blockIdx.x * (size_t) blockDim.x + threadIdx.x; p < P; p += (size_t) gridDim.x * blockDim.x ) {
        bf16 * d = x0 + p * 8;
        for ( int k = 0; k < 3; k++ ) d[k] = (bf16) ( v[p * 3 + k] * inv_scale );
        for ( int k = 3; k < 8; k++ ) d[k] = (bf16) 0.0f;
    }
}

// Mean-square loss against the target (both [P][3] FP32, the target in
// physical units): dz3 = d loss / d out3 (network units) as BF16 [P][16];
// per-block loss partial sums.
__global__ void __launch_bounds__( 256 ) k_dn_loss( const float * __restrict__ out, const float * __restrict__ tgt,
                                                    size_t P, float scale, bf16 * __restrict__ dz3,
                                                    double * __restrict__ partial ) {
    __shared__ double s[256];
    double acc = 0.0;
    const double n = 3.0 * P;
    for ( size_t p = blockIdx.x * (size_t) blockDim.x + threadIdx.x; p < P; p += (size_t) gridDim.x * blockDim.x ) {
        for ( int k = 0; k < 3; k++ ) {
            const float d = out[p * 3 + k] - tgt[p * 3 + k];  // physical units
            acc += (double) d * d;
            // out = scale * net  =>  d loss / d net = 2 d scale / n
            dz3[p * 16 + k] = (bf16) ( 2.0 * d * scale / n );
        }
        for ( int k = 3; k < 16; k++ ) dz3[p * 16 + k] = (bf16) 0.0f;
    }
    s[threadIdx.x] = acc;
    __syncthreads();
    for ( int o = 128; o > 0; o >>= 1 ) {
        if ( (int) threadIdx.x < o ) s[threadIdx.x] += s[threadIdx.x + o];
        __syncthreads();
    }
    if ( threadIdx.x == 0 ) partial[blockIdx.x] = s[0];
}

__global__ void k_dn_adam( float * __restrict__ w, float * __restrict__ m, float * __restrict__ v,
                           const float * __restrict__ grad, float lr, float b1, float b2, float eps, float c1, float c2 ) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if ( i >= NPARAM ) return;
    const float gi = grad[i];
    const float mi = b1 * m[i] + ( 1.0f - b1 ) * gi;
    const float vi = b2 * v[i] + ( 1.0f - b2 ) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * ( mi / c1 ) / ( sqrtf( vi / c2 ) + eps );
}

// Low-pass comparison for the shadow check: sum over cells of |a - b|^2 and
// |b|^2 of [P][3] arrays.
__global__ void __launch_bounds__( 256 ) k_dn_diff( const float * __restrict__ a, const float * __restrict__ b, size_t n,
                                                    double * __restrict__ partial ) {
    __shared__ double s0[256], s1[256];
    double d2 = 0.0, b2 = 0.0;
    for ( size_t i = blockIdx.x * (size_t) blockDim.x + threadIdx.x; i < n; i += (size_t) gridDim.x * blockDim.x ) {
        const double d = (double) a[i] - b[i];
        d2 += d * d;
        b2 += (double) b[i] * b[i];
    }
    s0[threadIdx.x] = d2;
    s1[threadIdx.x] = b2;
    __syncthreads();
    for ( int o = 128; o > 0; o >>= 1 ) {
        if ( (int) threadIdx.x < o ) {
            s0[threadIdx.x] += s0[threadIdx.x + o];
            s1[threadIdx.x] += s1[threadIdx.x + o];
        }
        __syncthreads();
    }
    if ( threadIdx.x == 0 ) {
        partial[2 * blockIdx.x] = s0[0];
        partial[2 * blockIdx.x + 1] = s1[0];
    }
}

constexpr int RED_BLOCKS = 256;

}  // namespace
}  // namespace zpic

using namespace zpic;

struct zpic_denoiser {
    int H = 0, W = 0, max_pairs = 0, npairs = 0;
    float scale = 1.0f;
    int adam_t = 0;
    hipStream_t st = nullptr;
    float * w = nullptr;       // [NPARAM] master
    float * m = nullptr;       // Adam moments
    float * v = nullptr;
    float * grad = nullptr;    // [NPARAM]
    bf16 * pack = nullptr;     // [NPACK][KS2][64][8]
    bf16 * x0 = nullptr;       // [P][8]
    bf16 * a1 = nullptr;       // [P][16]
    bf16 * a2 = nullptr;
    bf16 * dz3 = nullptr;
    bf16 * dz2 = nullptr;
    bf16 * dz1 = nullptr;
    float * out = nullptr;     // [P][3]
    float * part1 = nullptr;   // [H][16][13*16+1]
    float * part2 = nullptr;   // [H][16][25*16+1]
    float * part3 = nullptr;
    double * red = nullptr;    // [2*RED_BLOCKS]
    float * pairs_in = nullptr;   // [max_pairs][P][3]
    float * pairs_tgt = nullptr;  // [max_pairs][P][3]
    std::vector<void *> owned;
};

namespace {

template <class T>
int dn_alloc( zpic_denoiser * d, T ** p, size_t n ) {
    void * q = nullptr;
    if ( n == 0 ) { *p = nullptr; return ZPIC_OK; }
    ZPIC_HIP_TRY( hipMalloc( &q, n * sizeof( T ) ) );
    d->owned.push_back( q );
    *p = static_cast<T *>( q );
    return ZPIC_OK;
}

size_t dn_P( const zpic_denoiser * d ) { return (size_t) d->H * d->W; }

int dn_repack( zpic_denoiser * d ) {
    hipLaunchKernelGGL( k_dn_pack, dim3( KS2, NPACK ), dim3( WAVE ), 0, d->st, d->w, d->pack );
    ZPIC_LAUNCH_CHECK( "k_dn_pack" );
    return ZPIC_OK;
}

int conv_grid( const zpic_denoiser * d ) {
    const int nblk = (int) ( dn_P( d ) / 16 );
    return std::min( 4096, ( nblk + 3 ) / 4 );
}

const bf16 * packp( const zpic_denoiser * d, int which ) { return d->pack + (size_t) which * KS2 * WAVE * 8; }

// Forward pass from d->x0; output d->out (physical units) = scale * net.
int dn_forward( zpic_denoiser * d ) {
    const int H = d->H, W = d->W, gr = conv_grid( d );
    hipLaunchKernelGGL( ( k_dn_conv<8, KS1, true, false, false> ), dim3( gr ), dim3( 256 ), 0, d->st, d->x0,
                        packp( d, PACK_FWD1 ), d->w + OFF_B1, 16, (const bf16 *) nullptr, (void *) d->a1, H, W, 1.0f );
    hipLaunchKernelGGL( ( k_dn_conv<16, KS2, true, false, false> ), dim3( gr ), dim3( 256 ), 0, d->st, d->a1,
                        packp( d, PACK_FWD2 ), d->w + OFF_B2, 16, (const bf16 *) nullptr, (void *) d->a2, H, W, 1.0f );
    hipLaunchKernelGGL( ( k_dn_conv<16, KS2, false, false, true> ), dim3( gr ), dim3( 256 ), 0, d->st, d->a2,
                        packp( d, PACK_FWD3 ), d->w + OFF_B3, 3, (const bf16 *) nullptr, (void *) d->out, H, W,
                        d->scale );
    ZPIC_LAUNCH_CHECK( "k_dn_conv (forward)" );
    return ZPIC_OK;
}

// Loss against tgt ([P][3] physical) and the full gradient into d->grad.
int dn_backward( zpic_denoiser * d, const float * tgt, double * loss ) {
    const int H = d->H, W = d->W, gr = conv_grid( d );
    const size_t P = dn_P( d );
    hipLaunchKernelGGL( k_dn_loss, dim3( RED_BLOCKS ), dim3( 256 ), 0, d->st, d->out, tgt, P, d->scale, d->dz3, d->red );
    // input gradients through layers 3 and 2, ReLU masks from the activations
    hipLaunchKernelGGL( ( k_dn_conv<16, KS2, false, true, false> ), dim3( gr ), dim3( 256 ), 0, d->st, d->dz3,
                        packp( d, PACK_BWD3 ), (const float *) nullptr, 16, d->a2, (void *) d->dz2, H, W, 1.0f );
    hipLaunchKernelGGL( ( k_dn_conv<16, KS2, false, true, false> ), dim3( gr ), dim3( 256 ), 0, d->st, d->dz2,
                        packp( d, PACK_BWD2 ), (const float *) nullptr, 16, d->a1, (void *) d->dz1, H, W, 1.0f );
    ZPIC_LAUNCH_CHECK( "k_dn_conv (input gradients)" );
    hipLaunchKernelGGL( ( k_dn_wgrad<8, 13> ), dim3( H ), dim3( WAVE ), 0, d->st, d->x0, d->dz1, d->part1, H, W );
    hipLaunchKernelGGL( ( k_dn_wgrad<16, 25> ), dim3( H ), dim3( WAVE ), 0, d->st, d->a1, d->dz2, d->part2, H, W );
    hipLaunchKernelGGL( ( k_dn_wgrad<16, 25> ), dim3( H ), dim3( WAVE ), 0, d->st, d->a2, d->dz3, d->part3, H, W );
    hipLaunchKernelGGL( k_dn_wgrad_reduce, dim3( ( NPARAM + 255 ) / 256 ), dim3( 256 ), 0, d->st, d->part1, d->part2,
                        d->part3, H, d->grad );
    ZPIC_LAUNCH_CHECK( "k_dn_wgrad" );
    if ( loss ) {
        std::vector<double> h( RED_BLOCKS );
        ZPIC_HIP_TRY( hipMemcpyAsync( h.data(), d->red, sizeof( double ) * RED_BLOCKS, hipMemcpyDeviceToHost, d->st ) );
        ZPIC_HIP_TRY( hipStreamSynchronize( d->st ) );
        double s = 0;
        for ( double x : h ) s += x;
        *loss = s / ( 3.0 * P );
    }
    return ZPIC_OK;
}

int dn_input_from_f32( zpic_denoiser * d, const float * v ) {
    hipLaunchKernelGGL( k_dn_input, dim3( 1024 ), dim3( 256 ), 0, d->st, v, dn_P( d ), 1.0f / d->scale, d->x0 );
    ZPIC_LAUNCH_CHECK( "k_dn_input" );
    return ZPIC_OK;
}

int dn_check_geom( const zpic_denoiser * d, const zpic_geom * g ) {
    if ( !g || g->nx * g->ntiles_x != d->W || g->ny * g->ntiles_y != d->H )
        return set_error( ZPIC_EINVAL, "grid does not match the denoiser's size" );
    return ZPIC_OK;
}

}  // namespace

extern "C" {

zpic_denoiser * zpic_denoiser_new( int H, int W, int max_pairs ) {
    if ( H < 5 || W < 32 || W % 32 || max_pairs < 0 ) {
        set_error( ZPIC_EINVAL, "denoiser grid must be at least 5 rows and a multiple of 32 columns" );
        return nullptr;
    }
    auto * d = new zpic_denoiser();
    d->H = H;
    d->W = W;
    d->max_pairs = max_pairs;
    const size_t P = (size_t) H * W;
    int rc = ZPIC_OK;
    if ( hipStreamCreateWithFlags( &d->st, hipStreamNonBlocking ) != hipSuccess ) rc = set_error( ZPIC_EHIP, "stream" );
    if ( !rc ) rc = dn_alloc( d, &d->w, NPARAM );
    if ( !rc ) rc = dn_alloc( d, &d->m, NPARAM );
    if ( !rc ) rc = dn_alloc( d, &d->v, NPARAM );
    if ( !rc ) rc = dn_alloc( d, &d->grad, NPARAM );
    if ( !rc ) rc = dn_alloc( d, &d->pack, (size_t) NPACK * KS2 * WAVE * 8 );
    if ( !rc ) rc = dn_alloc( d, &d->x0, P * 8 );
    if ( !rc ) rc = dn_alloc( d, &d->a1, P * 16 );
    if ( !rc ) rc = dn_alloc( d, &d->a2, P * 16 );
    if ( !rc ) rc = dn_alloc( d, &d->dz3, P * 16 );
    if ( !rc ) rc = dn_alloc( d, &d->dz2, P * 16 );
    if ( !rc ) rc = dn_alloc( d, &d->dz1, P * 16 );
    if ( !rc ) rc = dn_alloc( d, &d->out, P * 3 );
    if ( !rc ) rc = dn_alloc( d, &d->part1, (size_t) H * 16 * ( 13 * 16 + 1 ) );
    if ( !rc ) rc = dn_alloc( d, &d->part2, (size_t) H * 16 * ( 25 * 16 + 1 ) );
    if ( !rc ) rc = dn_alloc( d, &d->part3, (size_t) H * 16 * ( 25 * 16 + 1 ) );
    if ( !rc ) rc = dn_alloc( d, &d->red, 2 * RED_BLOCKS );
    if ( !rc ) rc = dn_alloc( d, &d->pairs_in, (size_t) max_pairs * P * 3 );
    if ( !rc ) rc = dn_alloc( d, &d->pairs_tgt, (size_t) max_pairs * P * 3 );
    if ( !rc && hipMemsetAsync( d->pack, 0, sizeof( bf16 ) * NPACK * KS2 * WAVE * 8, d->st ) != hipSuccess )
        rc = set_error( ZPIC_EHIP, "hipMemsetAsync" );
    if ( !rc ) {
        // He-style initialisation, fixed seed; biases zero
        std::vector<float> h( NPARAM, 0.0f );
        std::mt19937 gen( 12345 );
        auto fill = [&]( int off, int n, int fan_in ) {
            std::normal_distribution<float> nd( 0.0f, std::sqrt( 2.0f / fan_in ) );
            for ( int i = 0; i < n; i++ ) h[off + i] = nd( gen );
        };
        fill( OFF_W1, 1200, 75 );
        fill( OFF_W2, 6400, 400 );
        fill( OFF_W3, 1200, 400 );
        rc = zpic_denoiser_set_params( d, h.data() );
    }
    if ( rc ) {
        zpic_denoiser_free( d );
        return nullptr;
    }
    return d;
}

void zpic_denoiser_free( zpic_denoiser * d ) {
    if ( !d ) return;
    if ( d->st ) (void) hipStreamSynchronize( d->st );
    for ( void * p : d->owned ) (void) hipFree( p );
    if ( d->st ) (void) hipStreamDestroy( d->st );
    delete d;
}

int zpic_denoiser_set_params( zpic_denoiser * d, const float * host ) {
    if ( !d || !host ) return set_error( ZPIC_EINVAL, "null argument" );
    ZPIC_HIP_TRY( hipMemcpyAsync( d->w, host, sizeof( float ) * NPARAM, hipMemcpyHostToDevice, d->st ) );
    ZPIC_HIP_TRY( hipMemsetAsync( d->m, 0, sizeof( float ) * NPARAM, d->st ) );
    ZPIC_HIP_TRY( hipMemsetAsync( d->v, 0, sizeof( float ) * NPARAM, d->st ) );
    d->adam_t = 0;
    if ( int rc = dn_repack( d ) ) return rc;
    ZPIC_HIP_TRY( hipStreamSynchronize( d->st ) );
    return ZPIC_OK;
}

int zpic_denoiser_get_params( zpic_denoiser * d, float * host ) {
    if ( !d || !host ) return set_error( ZPIC_EINVAL, "null argument" );
    ZPIC_HIP_TRY( hipMemcpyAsync( host, d->w, sizeof( float ) * NPARAM, hipMemcpyDeviceToHost, d->st ) );
    ZPIC_HIP_TRY( hipStreamSynchronize( d->st ) );
    return ZPIC_OK;
}

int zpic_denoiser_set_scale( zpic_denoiser * d, float scale ) {
    if ( !d || !( scale > 0.0f ) ) return set_error( ZPIC_EINVAL, "scale must be positive" );
    d->scale = scale;
    return ZPIC_OK;
}

float zpic_denoiser_get_scale( const zpic_denoiser * d ) { return d ? d->scale : 0.0f; }

int zpic_denoiser_forward_host( zpic_denoiser * d, const float * in, float * out ) {
    if ( !d || !in || !out ) return set_error( ZPIC_EINVAL, "null argument" );
    const size_t P = dn_P( d );
    float * tmp = nullptr;
    ZPIC_HIP_TRY( hipMalloc( (void **) &tmp, sizeof( float ) * P * 3 ) );
    int rc = ZPIC_OK;
    if ( hipMemcpyAsync( tmp, in, sizeof( float ) * P * 3, hipMemcpyHostToDevice, d->st ) != hipSuccess )
        rc = set_error( ZPIC_EHIP, "copy in" );
    if ( !rc ) rc = dn_input_from_f32( d, tmp );
    if ( !rc ) rc = dn_forward( d );
    if ( !rc && hipMemcpyAsync( out, d->out, sizeof( float ) * P * 3, hipMemcpyDeviceToHost, d->st ) != hipSuccess )
        rc = set_error( ZPIC_EHIP, "copy out" );
    if ( !rc && hipStreamSynchronize( d->st ) != hipSuccess ) rc = set_error( ZPIC_EHIP, "sync" );
    (void) hipFree( tmp );
    return rc;
}

int zpic_denoiser_gradient_host( zpic_denoiser * d, const float * in, const float * tgt, float * grad, double * loss ) {
    if ( !d || !in || !tgt || !grad ) return set_error( ZPIC_EINVAL, "null argument" );
    const size_t P = dn_P( d );
    float * tmp = nullptr;
    ZPIC_HIP_TRY( hipMalloc( (void **) &tmp, sizeof( float ) * P * 6 ) );
    int rc = ZPIC_OK;
    if ( hipMemcpyAsync( tmp, in, sizeof( float ) * P * 3, hipMemcpyHostToDevice, d->st ) != hipSuccess ||
         hipMemcpyAsync( tmp + P * 3, tgt, sizeof( float ) * P * 3, hipMemcpyHostToDevice, d->st ) != hipSuccess )
        rc = set_error( ZPIC_EHIP, "copy in" );
    if ( !rc ) rc = dn_input_from_f32( d, tmp );
    if ( !rc ) rc = dn_forward( d );
    if ( !rc ) rc = dn_backward( d, tmp + P * 3, loss );
    if ( !rc && hipMemcpyAsync( grad, d->grad, sizeof( float ) * NPARAM, hipMemcpyDeviceToHost, d->st ) != hipSuccess )
        rc = set_error( ZPIC_EHIP, "copy out" );
    if ( !rc && hipStreamSynchronize( d->st ) != hipSuccess ) rc = set_error( ZPIC_EHIP, "sync" );
    (void) hipFree( tmp );
    return rc;
}

// Stores a training pair from two tile-major J grids (device): the noisy
// input and the target.
int zpic_denoiser_add_pair( zpic_denoiser * d, const float * J_in, const float * J_target, const zpic_geom * g,
                            void * stream ) {
    if ( !d || !J_in || !J_target ) return set_error( ZPIC_EINVAL, "null argument" );
    if ( int rc = dn_check_geom( d, g ) ) return rc;
    if ( d->npairs >= d->max_pairs ) return set_error( ZPIC_ERANGE, "pair store full" );
    const size_t P = dn_P( d );
    // the grids were written on the caller's stream
    ZPIC_HIP_TRY( hipStreamSynchronize( as_stream( stream ) ) );
    float * pi = d->pairs_in + (size_t) d->npairs * P * 3;
    float * pt = d->pairs_tgt + (size_t) d->npairs * P * 3;
    hipLaunchKernelGGL( k_dn_gather, dim3( 1024 ), dim3( 256 ), 0, d->st, J_in, *g, 1.0f, (bf16 *) nullptr, pi );
    hipLaunchKernelGGL( k_dn_gather, dim3( 1024 ), dim3( 256 ), 0, d->st, J_target, *g, 1.0f, (bf16 *) nullptr, pt );
    ZPIC_LAUNCH_CHECK( "k_dn_gather" );
    ZPIC_HIP_TRY( hipStreamSynchronize( d->st ) );
    d->npairs++;
    return ZPIC_OK;
}

int zpic_denoiser_num_pairs( const zpic_denoiser * d ) { return d ? d->npairs : 0; }

int zpic_denoiser_clear_pairs( zpic_denoiser * d ) {
    if ( !d ) return set_error( ZPIC_EINVAL, "null denoiser" );
    d->npairs = 0;
    return ZPIC_OK;
}

// Root-mean-square of the stored targets (a natural scale for the network).
int zpic_denoiser_target_rms( zpic_denoiser * d, double * rms ) {
    if ( !d || !rms || d->npairs == 0 ) return set_error( ZPIC_EINVAL, "no pairs stored" );
    const size_t n = (size_t) d->npairs * dn_P( d ) * 3;
    hipLaunchKernelGGL( k_dn_diff, dim3( RED_BLOCKS ), dim3( 256 ), 0, d->st, d->pairs_tgt, d->pairs_tgt, n, d->red );
    ZPIC_LAUNCH_CHECK( "k_dn_diff" );
    std::vector<double> h( 2 * RED_BLOCKS );
    ZPIC_HIP_TRY( hipMemcpyAsync( h.data(), d->red, sizeof( double ) * 2 * RED_BLOCKS, hipMemcpyDeviceToHost, d->st ) );
    ZPIC_HIP_TRY( hipStreamSynchronize( d->st ) );
    double s = 0;
    for ( int i = 0; i < RED_BLOCKS; i++ ) s += h[2 * i + 1];
    *rms = std::sqrt( s / n );
    return ZPIC_OK;
}

// `steps` Adam steps over the first `use` stored pairs (all if use <= 0),
// pair order drawn with `seed`; loss_hist[steps] receives each step's loss
// (may be null).
int zpic_denoiser_train( zpic_denoiser * d, int steps, float lr, unsigned seed, int use, double * loss_hist ) {
    if ( !d || steps < 0 || !( lr > 0.0f ) ) return set_error( ZPIC_EINVAL, "invalid training request" );
    if ( d->npairs == 0 ) return set_error( ZPIC_ESTATE, "no training pairs stored" );
    if ( use <= 0 || use > d->npairs ) use = d->npairs;
    const size_t P = dn_P( d );
    std::mt19937 gen( seed );
    std::uniform_int_distribution<int> pick( 0, use - 1 );
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    for ( int s = 0; s < steps; s++ ) {
        const int k = pick( gen );
        if ( int rc = dn_input_from_f32( d, d->pairs_in + (size_t) k * P * 3 ) ) return rc;
        if ( int rc = dn_forward( d ) ) return rc;
        double loss = 0;
        if ( int rc = dn_backward( d, d->pairs_tgt + (size_t) k * P * 3, loss_hist ? &loss : nullptr ) ) return rc;
        if ( loss_hist ) loss_hist[s] = loss;
        d->adam_t++;
        const float c1 = 1.0f - std::pow( b1, (float) d->adam_t ), c2 = 1.0f - std::pow( b2, (float) d->adam_t );
        hipLaunchKernelGGL( k_dn_adam, dim3( ( NPARAM + 255 ) / 256 ), dim3( 256 ), 0, d->st, d->w, d->m, d->v, d->grad,
                            lr, b1, b2, eps, c1, c2 );
        ZPIC_LAUNCH_CHECK( "k_dn_adam" );
        if ( int rc = dn_repack( d ) ) return rc;
    }
    ZPIC_HIP_TRY( hipStreamSynchronize( d->st ) );
    return ZPIC_OK;
}

// Mean-square error of the network output (and of the unfiltered input)
// against the target of stored pair k.
int zpic_denoiser_eval_pair( zpic_denoiser * d, int k, double * mse_net, double * mse_input ) {
    if ( !d || k < 0 || k >= d->npairs ) return set_error( ZPIC_EINVAL, "bad pair index" );
    const size_t P = dn_P( d );
    const float * in = d->pairs_in + (size_t) k * P * 3;
    const float * tg = d->pairs_tgt + (size_t) k * P * 3;
    if ( int rc = dn_input_from_f32( d, in ) ) return rc;
    if ( int rc = dn_forward( d ) ) return rc;
    std::vector<double> h( 2 * RED_BLOCKS );
    auto mse = [&]( const float * a, double * out ) -> int {
        hipLaunchKernelGGL( k_dn_diff, dim3( RED_BLOCKS ), dim3( 256 ), 0, d->st, a, tg, P * 3, d->red );
        ZPIC_LAUNCH_CHECK( "k_dn_diff" );
        ZPIC_HIP_TRY( hipMemcpyAsync( h.data(), d->red, sizeof( double ) * 2 * RED_BLOCKS, hipMemcpyDeviceToHost, d->st ) );
        ZPIC_HIP_TRY( hipStreamSynchronize( d->st ) );
        double s = 0;
        for ( int i = 0; i < RED_BLOCKS; i++ ) s += h[2 * i];
        *out = s / ( 3.0 * P );
        return ZPIC_OK;
    };
    if ( mse_net )
        if ( int rc = mse( d->out, mse_net ) ) return rc;
    if ( mse_input )
        if ( int rc = mse( in, mse_input ) ) return rc;
    return ZPIC_OK;
}

// Host copies of stored pair k ([P][3] each).
int zpic_denoiser_get_pair( zpic_denoiser * d, int k, float * in, float * tgt ) {
    if ( !d || k < 0 || k >= d->npairs ) return set_error( ZPIC_EINVAL, "bad pair index" );
    const size_t P = dn_P( d );
    if ( in )
        ZPIC_HIP_TRY( hipMemcpyAsync( in, d->pairs_in + (size_t) k * P * 3, sizeof( float ) * P * 3,
                                      hipMemcpyDeviceToHost, d->st ) );
    if ( tgt )
        ZPIC_HIP_TRY( hipMemcpyAsync( tgt, d->pairs_tgt + (size_t) k * P * 3, sizeof( float ) * P * 3,
                                      hipMemcpyDeviceToHost, d->st ) );
    ZPIC_HIP_TRY( hipStreamSynchronize( d->st ) );
    return ZPIC_OK;
}

// In-place denoising of the interior of a tile-major J grid (device) issued
// on `stream`; guard cells are left to the caller (copy_to_gc).
int zpic_denoiser_apply( zpic_denoiser * d, float * J, const zpic_geom * g, void * stream ) {
    if ( !d || !J ) return set_error( ZPIC_EINVAL, "null argument" );
    if ( int rc = dn_check_geom( d, g ) ) return rc;
    hipStream_t caller = as_stream( stream );
    hipStream_t own = d->st;
    d->st = caller;  // run on the caller's stream: ordered with the step
    hipLaunchKernelGGL( k_dn_gather, dim3( 1024 ), dim3( 256 ), 0, d->st, J, *g, 1.0f / d->scale, d->x0,
                        (float *) nullptr );
    int rc = dn_forward( d );
    if ( !rc ) {
        hipLaunchKernelGGL( k_dn_scatter, dim3( 1024 ), dim3( 256 ), 0, d->st, d->out, *g, J );
        hipError_t e = hipGetLastError();
        if ( e != hipSuccess ) rc = zpic::hip_error( e, "k_dn_scatter" );
    }
    d->st = own;
    return rc;
}

// Relative L2 difference ||a - b|| / ||b|| of two [n] float arrays (device).
int zpic_denoiser_reldiff( zpic_denoiser * d, const float * a, const float * b, size_t n, void * stream,
                           double * out ) {
    if ( !d || !a || !b || !out ) return set_error( ZPIC_EINVAL, "null argument" );
    hipStream_t st = as_stream( stream );
    hipLaunchKernelGGL( k_dn_diff, dim3( RED_BLOCKS ), dim3( 256 ), 0, st, a, b, n, d->red );
    ZPIC_LAUNCH_CHECK( "k_dn_diff" );
    std::vector<double> h( 2 * RED_BLOCKS );
    ZPIC_HIP_TRY( hipMemcpyAsync( h.data(), d->red, sizeof( double ) * 2 * RED_BLOCKS, hipMemcpyDeviceToHost, st ) );
    ZPIC_HIP_TRY( hipStreamSynchronize( st ) );
    double s0 = 0, s1 = 0;
    for ( int i = 0; i < RED_BLOCKS; i++ ) {
        s0 += h[2 * i];
        s1 += h[2 * i + 1];
    }
    *out = s1 > 0 ? std::sqrt( s0 / s1 ) : std::sqrt( s0 );
    return ZPIC_OK;
}

int zpic_denoiser_size( const zpic_denoiser * d, int * H, int * W ) {
    if ( !d ) return set_error( ZPIC_EINVAL, "null denoiser" );
    if ( H ) *H = d->H;
    if ( W ) *W = d->W;
    return ZPIC_OK;
}

}  // extern "C"
