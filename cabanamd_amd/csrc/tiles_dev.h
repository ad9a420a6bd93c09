// tiles_dev.h -- device helpers shared by the tiled list build (tiles.hip)
// and the tiled force kernel (lj.hip).  See tiles.hip for the layout.
#pragma once

#include "internal.h"

namespace cmd
{
static constexpr int TI_WORDS = 44;
enum
{
    TI_LB = 0,     // [9] first local index of each stencil column run
    TI_LLEN = 9,   // [9] local run lengths
    TI_GB = 18,    // [9] first sorted-ghost slot of each column run
    TI_GLEN = 27,  // [9] ghost run lengths
    TI_TL = 36,    // first local atom of the tile
    TI_NAT = 37,   // atoms in the tile
    TI_TOTAL = 38, // staged positions
    TI_EA = 39,    // first staged cell (e0 - 1, clamped)
    TI_NST = 40    // staged cells per column
};

constexpr int TILE_STRIP = 4;

__device__ __forceinline__ void tile_coords( const TileGrid &T, int t, int &a,
                                             int &b, int &e0, int &e1 )
{
    // columns in strips of TILE_STRIP rows of a, b outer and a inner within
    // a strip: the 3 x 3 columns around a column run close together in
    // time, so a staged column is mostly re-read from L2
    const int col = t / T.tpc, z = t % T.tpc;
    const int strip = col / ( TILE_STRIP * T.nc1 ), r = col % ( TILE_STRIP * T.nc1 );
    const int h = min( TILE_STRIP, T.nc0 - strip * TILE_STRIP );
    a = 1 + strip * TILE_STRIP + r % h;
    b = 1 + r / h;
    e0 = 1 + z * T.E;
    e1 = min( e0 + T.E, T.g.n[2] - 1 );
}

__device__ __forceinline__ int col_base( const TileGrid &T, int a, int b )
{
    return ( a * T.g.n[1] + b ) * T.g.n[2];
}

// Round-robin dispatch puts block b on XCD b % 8; give XCD k the k-th
// contiguous slice of tiles.
__device__ __forceinline__ int xcd_tile( int b, int nb )
{
    const int q = nb >> 3, r = nb & 7;
    const int xcd = b & 7, k = b >> 3;
    return ( xcd < r ) ? xcd * ( q + 1 ) + k : r * ( q + 1 ) + ( xcd - r ) * q + k;
}

// Shared tile header: info words and the 19 segment offsets
// (segment 2q = locals of column q, 2q+1 = its ghosts).
struct TileHead
{
    int info[TI_WORDS];
    int segoff[19];
};

__device__ inline void load_head( const int *__restrict__ info, int t, TileHead &H )
{
    const int *I = info + (size_t)t * TI_WORDS;
    for ( int k = threadIdx.x; k < TI_WORDS; k += blockDim.x )
        H.info[k] = I[k];
    __syncthreads();
    if ( threadIdx.x == 0 )
    {
        int off = 0;
        for ( int q = 0; q < 9; ++q )
        {
            H.segoff[2 * q] = off;
            off += H.info[TI_LLEN + q];
            H.segoff[2 * q + 1] = off;
            off += H.info[TI_GLEN + q];
        }
        H.segoff[18] = off;
    }
    __syncthreads();
}

// Position of a tile's sentinel slot (slot TOTAL, just past the staged
// atoms): list rows are padded with it, and its distance to any atom is far
// beyond every cutoff, so a padded pair contributes exact zeros.
static constexpr double TILE_FAR = 1.0e30;

// Bytes of one staged position in the force kernel's LDS records (x, y, z
// doubles).  Tiled list entries are byte offsets of records (slot * 24), so
// a gather is one address add and two LDS reads with immediate offsets; the
// largest tile (MAX_SLOTS staged atoms plus the sentinel) must keep every
// offset inside 16 bits.
static constexpr int TILE_REC = 24;

// Lanes per atom of slab wave w of a tile with nat atoms, for tiled lists
// built with G lanes per atom (64 / G atoms per full wave): G, doubled up to
// 8 while the wave's atoms still fill the 64 lanes.  A tile's last wave is
// usually part full; spreading its rows over more lanes shortens its walk.
__device__ __forceinline__ int wave_lanes( int G, int nat, int w )
{
    const int na = min( 64 / G, nat - w * ( 64 / G ) );
    int g = G;
    while ( g < 8 && na * g * 2 <= 64 )
        g *= 2;
    return g;
}

// Copy the 18 runs of positions (locals from x, ghosts from xg) into LDS as
// 24-byte records, followed by the sentinel record; with WITH_TYPE the types
// follow as ints, `cap` records on.
template <bool WITH_TYPE>
__device__ inline void stage_records( const TileHead &H, const double4 *__restrict__ x,
                                      const double4 *__restrict__ xg,
                                      double *__restrict__ s_rec, int cap )
{
    int *s_t = reinterpret_cast<int *>( s_rec + 3 * cap );
    const int total = H.segoff[18];
    constexpr int B = 2; // slots per thread whose loads are in flight together
    int seg = 0;
    for ( int k0 = threadIdx.x; k0 < total; k0 += B * blockDim.x )
    {
        double4 p[B];
#pragma unroll
        for ( int u = 0; u < B; ++u )
        {
            // past the end: reload the last slot (in bounds)
            const int k = min( k0 + u * (int)blockDim.x, total - 1 );
            while ( k >= H.segoff[seg + 1] )
                ++seg;
            const int q = seg >> 1, off = k - H.segoff[seg];
            const double4 *__restrict__ base = ( seg & 1 ) ? xg : x;
            const int idx = ( ( seg & 1 ) ? H.info[TI_GB + q] : H.info[TI_LB + q] ) + off;
            p[u] = base[idx];
        }
        // stores past the end go to the sentinel record, rewritten below
#pragma unroll
        for ( int u = 0; u < B; ++u )
        {
            const int k = min( k0 + u * (int)blockDim.x, total );
            s_rec[3 * k] = p[u].x;
            s_rec[3 * k + 1] = p[u].y;
            s_rec[3 * k + 2] = p[u].z;
            if ( WITH_TYPE )
                s_t[k] = (int)p[u].w;
        }
    }
    __syncthreads();
    if ( threadIdx.x == 0 )
    {
        s_rec[3 * total] = TILE_FAR;
        s_rec[3 * total + 1] = TILE_FAR;
        s_rec[3 * total + 2] = TILE_FAR;
        if ( WITH_TYPE )
            s_t[total] = 0;
    }
    __syncthreads();
}

// The same staging as separate coordinate planes (x, y, z doubles, and the
// type as an int when more than one type is present), `cap` slots each: a
// gather then reads three 8-byte words, which spread over the LDS banks
// better than one 32-byte record.
template <bool WITH_TYPE>
__device__ inline void stage_planes( const TileHead &H, const double4 *__restrict__ x,
                                     const double4 *__restrict__ xg,
                                     double *__restrict__ s_x, int cap )
{
    double *s_y = s_x + cap, *s_z = s_x + 2 * cap;
    int *s_t = reinterpret_cast<int *>( s_x + 3 * cap );
    const int total = H.segoff[18];
    constexpr int B = 4;
    int seg = 0;
    for ( int k0 = threadIdx.x; k0 < total; k0 += B * blockDim.x )
    {
        double4 p[B];
#pragma unroll
        for ( int u = 0; u < B; ++u )
        {
            const int k = min( k0 + u * (int)blockDim.x, total - 1 );
            while ( k >= H.segoff[seg + 1] )
                ++seg;
            const int q = seg >> 1, off = k - H.segoff[seg];
            const double4 *__restrict__ base = ( seg & 1 ) ? xg : x;
            const int idx = ( ( seg & 1 ) ? H.info[TI_GB + q] : H.info[TI_LB + q] ) + off;
            p[u] = base[idx];
        }
#pragma unroll
        for ( int u = 0; u < B; ++u )
        {
            const int k = min( k0 + u * (int)blockDim.x, total );
            s_x[k] = p[u].x;
            s_y[k] = p[u].y;
            s_z[k] = p[u].z;
            if ( WITH_TYPE )
                s_t[k] = (int)p[u].w;
        }
    }
    __syncthreads();
    if ( threadIdx.x == 0 )
    {
        s_x[total] = TILE_FAR;
        s_y[total] = TILE_FAR;
        s_z[total] = TILE_FAR;
        if ( WITH_TYPE )
            s_t[total] = 0;
    }
    __syncthreads();
}

} // namespace cmd
