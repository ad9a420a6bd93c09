// tiles_dev.h -- device helpers shared by the tiled list build (tiles.hip)
// and the tiled force kernels (lj.hip).  See tiles.hip for the layout.
#pragma once

#include "internal.h"

namespace cmd
{
// A tile is W x W interior columns of the binning grid (W = 1, 2 or 3) over E
// cells [e0, e1).  Its atoms are the local atoms of those cells: one
// contiguous run per atom column (the cell-sorted order runs along a
// column), taken in atom-column order c = ca * W + cb.  Every neighbour
// candidate of those atoms lies in the (W + 2) x (W + 2) staged columns
// around them, cells [e0 - 1, e1]: per staged column one contiguous run of
// locals and one of cell-sorted ghosts (segments 2q and 2q + 1 of staged
// column q = da * (W + 2) + db).  Staged columns past the padded grid (the
// right edge of a tile whose second atom column does not exist) are empty.
constexpr int TILE_WMAX = 3;
constexpr int TILE_QMAX = ( TILE_WMAX + 2 ) * ( TILE_WMAX + 2 ); // staged columns
constexpr int TILE_CMAX = TILE_WMAX * TILE_WMAX;                 // atom columns
static constexpr int TI_WORDS = 4 * TILE_QMAX + 2 * TILE_CMAX + 4;
enum
{
    TI_LB = 0,                         // [Q] first local index of each staged column run
    TI_LLEN = TILE_QMAX,               // [Q] local run lengths
    TI_GB = 2 * TILE_QMAX,             // [Q] first sorted-ghost slot of each run
    TI_GLEN = 3 * TILE_QMAX,           // [Q] ghost run lengths
    TI_AL = 4 * TILE_QMAX,             // [C] first local atom of each atom column
    TI_AN = 4 * TILE_QMAX + TILE_CMAX, // [C] atoms of each atom column
    TI_NAT = 4 * TILE_QMAX + 2 * TILE_CMAX, // atoms in the tile
    TI_TOTAL,                               // staged positions
    TI_EA,                                  // first staged cell (e0 - 1, clamped)
    TI_NST                                  // staged cells per column
};

constexpr int TILE_STRIP = 4;

__host__ __device__ constexpr int tile_q( int W ) { return ( W + 2 ) * ( W + 2 ); }

__device__ __forceinline__ void tile_coords( const TileGrid &T, int t, int &a0,
                                             int &b0, int &e0, int &e1 )
{
    // tile columns in strips of TILE_STRIP rows of a, b outer and a inner
    // within a strip: the columns around a tile are staged by its neighbours
    // close together in time, so a staged column is mostly re-read from L2
    const int col = t / T.tpc, z = t % T.tpc;
    const int strip = col / ( TILE_STRIP * T.nt1 ), r = col % ( TILE_STRIP * T.nt1 );
    const int h = min( TILE_STRIP, T.nt0 - strip * TILE_STRIP );
    a0 = 1 + T.W * ( strip * TILE_STRIP + r % h );
    b0 = 1 + T.W * ( r / h );
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

// Shared tile header: info words, the segment offsets in LDS slots
// (segment 2q = locals of staged column q, 2q + 1 = its ghosts) and the
// tile-atom offsets of the atom columns.
struct TileHead
{
    int info[TI_WORDS];
    int segoff[2 * TILE_QMAX + 1];
    int aoff[TILE_CMAX + 1];
    int nseg; // 2 (W + 2)^2
};

__device__ inline void load_head( const int *__restrict__ info, int t, int W, TileHead &H )
{
    const int *I = info + (size_t)t * TI_WORDS;
    for ( int k = threadIdx.x; k < TI_WORDS; k += blockDim.x )
        H.info[k] = I[k];
    __syncthreads();
    if ( threadIdx.x == 0 )
    {
        const int nq = tile_q( W );
        int off = 0;
        for ( int q = 0; q < nq; ++q )
        {
            H.segoff[2 * q] = off;
            off += H.info[TI_LLEN + q];
            H.segoff[2 * q + 1] = off;
            off += H.info[TI_GLEN + q];
        }
        H.segoff[2 * nq] = off;
        H.nseg = 2 * nq;
        int a = 0;
        for ( int c = 0; c < W * W; ++c )
        {
            H.aoff[c] = a;
            a += H.info[TI_AN + c];
        }
        H.aoff[W * W] = a;
    }
    __syncthreads();
}

// System index of tile atom ta (ta < the tile's atom count).
__device__ __forceinline__ int tile_atom( const TileHead &H, int W, int ta )
{
    int c = 0;
    while ( c + 1 < W * W && ta >= H.aoff[c + 1] )
        ++c;
    return H.info[TI_AL + c] + ( ta - H.aoff[c] );
}

// Staged column of atom column c.
__device__ __forceinline__ int atom_col_q( int W, int c )
{
    return ( c / W + 1 ) * ( W + 2 ) + ( c % W + 1 );
}

// Position of a tile's sentinel slot (slot TOTAL, just past the staged
// atoms): list rows are padded with it, and its distance to any atom is far
// beyond every cutoff, so a padded pair contributes exact zeros.
static constexpr double TILE_FAR = 1.0e30;

// Record sizes (TILE_REC, TILE_REC32) are defined with the neighbour list
// (internal.h): tiled list entries are byte offsets of records.

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

// The segment s (0 <= s < nseg) with off[s] <= k < off[s + 1], for
// off[0] <= k < off[nseg]: a binary search, so that the searches of a
// thread's several slots are independent chains of a few LDS reads (a walk
// from segment to segment would chain one read per segment passed).
__device__ __forceinline__ int find_seg( const int *off, int nseg, int k )
{
    int lo = 0, hi = nseg;
    while ( hi - lo > 1 )
    {
        const int mid = ( lo + hi ) >> 1;
        if ( off[mid] <= k )
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

// Copy the staged runs of positions (locals from x, ghosts from xg) into LDS
// as 24-byte records, followed by the sentinel record; with WITH_TYPE the
// types follow as ints, `cap` records on.  Each thread has up to B loads in
// flight, so a tile's neighbourhood (about a thousand records) arrives in
// one round of memory latency for blocks of 256 threads and more.  The
// `nzero` doubles at `zero` (an LDS area of the caller's) are cleared while
// the first round's loads are in flight.  Ends with a barrier.
template <bool WITH_TYPE>
__device__ inline void stage_records( const TileHead &H, const double4 *__restrict__ x,
                                      const double4 *__restrict__ xg,
                                      double *__restrict__ s_rec, int cap,
                                      double *__restrict__ zero = nullptr, int nzero = 0 )
{
    int *s_t = reinterpret_cast<int *>( s_rec + 3 * cap );
    const int total = H.segoff[H.nseg];
    constexpr int B = 8; // slots per thread whose loads are in flight together
    auto clear = [&]()
    {
        for ( int k = threadIdx.x; k < nzero; k += blockDim.x )
            zero[k] = 0.0;
    };
    for ( int k0 = threadIdx.x; k0 < total; k0 += B * blockDim.x )
    {
        double4 p[B];
#pragma unroll
        for ( int u = 0; u < B; ++u )
        {
            // past the end: reload the last slot (in bounds)
            const int k = min( k0 + u * (int)blockDim.x, total - 1 );
            const int seg = find_seg( H.segoff, H.nseg, k );
            const int q = seg >> 1, off = k - H.segoff[seg];
            const double4 *__restrict__ base = ( seg & 1 ) ? xg : x;
            const int idx = ( ( seg & 1 ) ? H.info[TI_GB + q] : H.info[TI_LB + q] ) + off;
            p[u] = base[idx];
        }
        if ( k0 == (int)threadIdx.x )
            clear();
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
    // threads past the records clear their share too
    if ( (int)threadIdx.x >= total )
        clear();
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

// The same neighbourhood staged as 16-byte FP32 records for the
// mixed-precision full-list kernel: x, y, z about the tile's centre (ox, oy,
// oz; the staged region spans a few cells, so the coordinates keep about
// 1e-7 of a cell in absolute precision) and the type's bits in w, followed
// by the sentinel record.  Ends with a barrier.
__device__ inline void stage_records32( const TileHead &H, const double4 *__restrict__ x,
                                        const double4 *__restrict__ xg,
                                        float4 *__restrict__ s_rec, double ox, double oy,
                                        double oz )
{
    const int total = H.segoff[H.nseg];
    constexpr int B = 8; // slots per thread whose loads are in flight together
    for ( int k0 = threadIdx.x; k0 < total; k0 += B * blockDim.x )
    {
        double4 p[B];
#pragma unroll
        for ( int u = 0; u < B; ++u )
        {
            // past the end: reload the last slot (in bounds)
            const int k = min( k0 + u * (int)blockDim.x, total - 1 );
            const int seg = find_seg( H.segoff, H.nseg, k );
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
            s_rec[k] = make_float4( (float)( p[u].x - ox ), (float)( p[u].y - oy ),
                                    (float)( p[u].z - oz ), __int_as_float( (int)p[u].w ) );
        }
    }
    __syncthreads();
    if ( threadIdx.x == 0 )
        s_rec[total] = make_float4( (float)TILE_FAR, (float)TILE_FAR, (float)TILE_FAR,
                                    __int_as_float( 0 ) );
    __syncthreads();
}

// Centre of tile t's staged region, the origin of its FP32 records.
__device__ __forceinline__ void tile_centre( const TileGrid &T, int t, double &ox, double &oy,
                                             double &oz )
{
    int a0, b0, e0, e1;
    tile_coords( T, t, a0, b0, e0, e1 );
    // columns a0 - 1 .. a0 + W, cells e0 - 1 .. e1
    ox = T.g.min[0] + ( a0 - 1 + 0.5 * ( T.W + 2 ) ) / T.g.rd[0];
    oy = T.g.min[1] + ( b0 - 1 + 0.5 * ( T.W + 2 ) ) / T.g.rd[1];
    oz = T.g.min[2] + 0.5 * ( e0 + e1 ) / T.g.rd[2];
}

} // namespace cmd
