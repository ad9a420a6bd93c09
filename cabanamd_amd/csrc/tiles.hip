// tiles.hip -- LDS-tiled Verlet lists and LJ force for binned systems.
//
// Same list as NeighborVerlet::create (neighbor_verlet.h:43-62) and the same
// forces as ForceLJ::compute_force_full (force_lj_cabana_neigh_impl.h:
// 151-203); only the work decomposition and the list encoding differ.
//
// A tile is W x W interior columns (W = 1, 2 or 3) of the binning grid (cells
// >= the neighbour radius) over E cells [e0, e1).  Its atoms are the local
// atoms of those cells, one contiguous run per atom column (locals are cell
// sorted by create_binning), and every neighbour candidate lies in the
// (W + 2) x (W + 2) staged columns around them, cells [e0 - 1, e1]: per
// staged column one contiguous range of local atoms and one of ghosts
// (ghosts are cell sorted here).  The force block stages those ranges of
// positions in LDS and the list stores, for each neighbour, the 16-bit byte
// offset of its LDS record (slot * TILE_REC).  Rows are walked in the order
// the oracle uses (stencil columns, then cells, then atoms by index), so
// lists and FP64 sums are unchanged.  Wider tiles stage fewer records per
// atom (2 x 2 x 3 cells: 6.9 per atom against 12 for 1 x 1 x 6) and give a
// block more atoms for the same LDS.
//
// Slab of a tile: the tile's atoms in atom-column order, in waves of
// APW = 64/G atoms; with gw lanes per atom in wave w (G, or up to 8 in a
// part-full wave: wave_lanes), neighbour k of atom a of the wave goes to row
// r = k / gw, lane (k % gw) * (64 / gw) + a, and four consecutive rows of a
// lane are packed in 8 bytes:
//   tslab[tile_base + ((w * R4 + r / 4) * 64 + lane) * 4 + r % 4],
// R4 = ceil(ceil(cap / G) / 4), so one 8-byte load per lane feeds four pairs.
// Every 8-byte word belongs to one atom, so the rows of different atom
// columns of a wave are written by different blocks without overlap.
#include "internal.h"
#include "tiles_dev.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>

namespace cmd
{
static constexpr int TPB = 256;
static constexpr int BTPB = 512;       // list build block
static constexpr int EMAX = 16;
static constexpr int MAX_SLOTS = 2720; // largest staged neighbourhood (65 KB of records)
static_assert( ( MAX_SLOTS + 1 ) * TILE_REC + 16 < 65536,
               "record offsets must fit the 16-bit list entries" );
// The same bound for 16-byte FP32 records: the sentinel's byte offset
// (slot MAX_SLOTS32) still fits the 16-bit entries.
static constexpr int MAX_SLOTS32 = ( 65536 - 16 ) / TILE_REC32 - 1;
static_assert( MAX_SLOTS32 * TILE_REC32 < 65536, "record offsets must fit the 16-bit list entries" );

// Largest staged neighbourhood of a force tile for list entries addressing
// records of `rec` bytes.
static int max_force_slots( int rec )
{
    return rec == TILE_REC32 ? MAX_SLOTS32 : MAX_SLOTS;
}

// Per-tile runs, atom columns and staged count; slab waves per tile.
// slots[0] gets the largest staged count of a tile, slots[1] the number of
// local atoms the tiles cover (all of them only when every local atom sits
// in an interior cell of the grid), slots[2] the most slab waves of one
// tile, slots[3] the largest staged count of one atom column's 3 x 3
// columns (the list build's neighbourhood), slots[4] the most atoms of one
// atom column (the list build's rows).
__global__ void k_tile_info( TileGrid T, const int *__restrict__ ls,
                             const int *__restrict__ gs, int apw,
                             int *__restrict__ info, int *__restrict__ waves,
                             int *__restrict__ slots )
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    int total = 0, nat = 0, nwv = 0, bmax = 0, cmax = 0;
    if ( t < T.ntiles )
    {
        int a0, b0, e0, e1;
        tile_coords( T, t, a0, b0, e0, e1 );
        const int ea = max( e0 - 1, 0 ), eb = min( e1 + 1, T.g.n[2] );
        const int S = T.W + 2;
        int *I = info + (size_t)t * TI_WORDS;
        int len[TILE_QMAX];
        for ( int q = 0; q < TILE_QMAX; ++q )
        {
            const int a = a0 - 1 + q / S, b = b0 - 1 + q % S;
            int lb = 0, ll = 0, gb = 0, gl = 0;
            if ( q < S * S && a < T.g.n[0] && b < T.g.n[1] )
            {
                const int c = col_base( T, a, b );
                lb = ls[c + ea];
                ll = ls[c + eb] - lb;
                gb = gs[c + ea];
                gl = gs[c + eb] - gb;
            }
            I[TI_LB + q] = lb;
            I[TI_LLEN + q] = ll;
            I[TI_GB + q] = gb;
            I[TI_GLEN + q] = gl;
            len[q] = ll + gl;
            total += ll + gl;
        }
        for ( int c = 0; c < TILE_CMAX; ++c )
        {
            const int ca = c / T.W, cb = c % T.W;
            int al = 0, an = 0;
            if ( c < T.W * T.W && a0 + ca <= T.g.n[0] - 2 && b0 + cb <= T.g.n[1] - 2 )
            {
                const int cc = col_base( T, a0 + ca, b0 + cb );
                al = ls[cc + e0];
                an = ls[cc + e1] - al;
                int s9 = 0;
                for ( int da = 0; da < 3; ++da )
                    for ( int db = 0; db < 3; ++db )
                        s9 += len[( ca + da ) * S + cb + db];
                bmax = max( bmax, s9 );
            }
            I[TI_AL + c] = al;
            I[TI_AN + c] = an;
            nat += an;
            cmax = max( cmax, an );
        }
        I[TI_NAT] = nat;
        I[TI_TOTAL] = total;
        I[TI_EA] = ea;
        I[TI_NST] = eb - ea;
        nwv = ( nat + apw - 1 ) / apw;
        waves[t] = nwv;
    }
    int m = total;
#pragma unroll
    for ( int o = 32; o > 0; o >>= 1 )
    {
        m = max( m, __shfl_xor( m, o, 64 ) );
        nat += __shfl_xor( nat, o, 64 );
        nwv = max( nwv, __shfl_xor( nwv, o, 64 ) );
        bmax = max( bmax, __shfl_xor( bmax, o, 64 ) );
        cmax = max( cmax, __shfl_xor( cmax, o, 64 ) );
    }
    if ( ( threadIdx.x & 63 ) == 0 )
    {
        if ( m > 0 )
            atomicMax( &slots[0], m );
        if ( nat > 0 )
            atomicAdd( &slots[1], nat );
        if ( nwv > 0 )
            atomicMax( &slots[2], nwv );
        if ( bmax > 0 )
            atomicMax( &slots[3], bmax );
        if ( cmax > 0 )
            atomicMax( &slots[4], cmax );
    }
}

// List build, one block per atom column of a tile and one wave per cell of
// that column.  All atoms of a cell share the same 27-cell stencil, so the
// wave loads the stencil's candidates into registers, BCH chunks of 64 at a
// time (lane = candidate, positions from LDS planes), and tests every atom
// of the cell against them in turn, the atom's position and row length being
// wave-uniform (read out of the lanes that hold them); ballot + mbcnt place each
// accepted candidate in the atom's row, so rows come out in walk order.  The
// block stages the 3 x 3 columns around its atom column (a subset of the
// tile's staged columns, the same cell range) as coordinate planes; each
// candidate carries the byte offset of its record in the force tile's
// staging, which is what the row stores.  The stencil is a list of
// contiguous LDS ranges (per column: the locals of its three cells, or
// locals and ghosts cell by cell where ghosts are present), concatenated so
// that chunks run across range ends; the range table sits in registers (lane
// r = range r), and each pass of chunks applies the ranges that start inside
// it in turn, so a lane ends with the range holding its candidate.  Rows are
// assembled in LDS one after another (pre-filled with the sentinel), so an
// entry's address is the row start plus its index, and are transposed into
// the slab layout on the way out.
static constexpr int MAXR = 9 * 3 * 2; // stencil ranges of one cell
static constexpr int BCH = 10;         // candidate chunks of 64 held in registers

// Bytes of the coordinate planes of `slots` staged atoms (elements of type
// PT), rounded up so that what follows is 16-byte aligned.
template <class PT>
__host__ __device__ constexpr size_t tile_planes_bytes( int slots )
{
    return ( (size_t)slots * 3 * sizeof( PT ) + 15 ) & ~(size_t)15;
}

// Entries of one LDS row of the list build: the slab's entries per atom
// (R4 groups of 4 G) plus 2, an odd number of 4-byte words.
__host__ __device__ constexpr int tile_row_stride( int R4, int G )
{
    return R4 * 4 * G + 2;
}

// The list build holds coordinates in FP32 about the block's origin (half
// the LDS and registers of FP64).  A candidate whose FP32 r^2 lies at or
// below lo (rc^2 less the FP32 error bound) is within rc in FP64 too; one
// above hi (rc^2 plus the bound) is not; one in between is tested again in
// FP64 on the positions in memory, as the reference evaluates it.  The
// half list's direction test (the candidate above the atom in x, then y,
// then z, comparing FP64 coordinates) follows the sign of the FP32 x
// difference where it exceeds eps (its error bound), and is settled in
// FP64 with the distance otherwise.
// What the re-test needs: the positions, the block's 9-column segment
// offsets (LDS) and the tile head's words, to map a build slot to its atom.
struct ExactCtx
{
    const double4 *x, *xg;
    const int *bseg;
    const int *info;
    int ca, cb, S;
    double rc2;
};

// Of the lanes in q (candidates at build slot sj, the atom at local index
// ia), those within rc in FP64 (and, HALF, above the atom).
template <bool HALF>
__device__ __forceinline__ unsigned long long exact_lanes( const ExactCtx &X,
                                                           unsigned long long q, int sj, int ia )
{
    bool ok = false;
    if ( __builtin_amdgcn_inverse_ballot_w64( q ) )
    {
        const int seg = find_seg( X.bseg, 18, sj );
        const int qb = seg >> 1, off = sj - X.bseg[seg];
        const int qf = ( X.ca + qb / 3 ) * X.S + X.cb + qb % 3;
        const double4 c = ( seg & 1 ) ? X.xg[X.info[TI_GB + qf] + off]
                                      : X.x[X.info[TI_LB + qf] + off];
        const double4 a = X.x[ia];
        const double dx = a.x - c.x;
        const double dy = a.y - c.y;
        const double dz = a.z - c.z;
        ok = dx * dx + dy * dy + dz * dz <= X.rc2;
        if ( HALF )
            ok = ok && ( ( c.x > a.x ) |
                         ( ( c.x == a.x ) & ( ( c.y > a.y ) | ( ( c.y == a.y ) & ( c.z > a.z ) ) ) ) );
    }
    return __builtin_amdgcn_ballot_w64( ok ) & q;
}

// The FP32 tests of two atoms (candidates vs and vs + 1 of the pass, local
// indices ia and ia + 1) against NC register chunks, their operations
// interleaved so that each dependent operation has the other atom's
// independent one beside it; lanes in the band are settled by exact_lanes.
template <int NC, bool HALF>
__device__ __forceinline__ void test_pair32( const float ( &px )[BCH], const float ( &py )[BCH],
                                             const float ( &pz )[BCH], const float4 pa,
                                             const float4 pb, int vs, float lo, float hi,
                                             float eps, const ExactCtx &X,
                                             const int ( &dj )[BCH], int v0, int ia,
                                             unsigned long long ( &ma )[NC],
                                             unsigned long long ( &mb )[NC] )
{
    const int csa = vs >> 6, csb = ( vs + 1 ) >> 6;
    const unsigned long long nsa = ~( 1ull << ( vs & 63 ) );
    const unsigned long long nsb = ~( 1ull << ( ( vs + 1 ) & 63 ) );
    const int lane = threadIdx.x & 63;
#pragma unroll
    for ( int c = 0; c < NC; ++c )
    {
        const float dxa = pa.x - px[c], dxb = pb.x - px[c];
        const float dya = pa.y - py[c], dyb = pb.y - py[c];
        const float dza = pa.z - pz[c], dzb = pb.z - pz[c];
        const float ra = fmaf( dza, dza, fmaf( dya, dya, dxa * dxa ) );
        const float rb = fmaf( dzb, dzb, fmaf( dyb, dyb, dxb * dxb ) );
        unsigned long long sa, sb, qa, qb;
        if ( HALF )
        {
            // settled: inside lo with a certain x direction
            const bool ka = ( ra <= lo ) & ( fabsf( dxa ) > eps );
            const bool kb = ( rb <= lo ) & ( fabsf( dxb ) > eps );
            sa = __builtin_amdgcn_ballot_w64( ka & ( dxa < 0.f ) );
            sb = __builtin_amdgcn_ballot_w64( kb & ( dxb < 0.f ) );
            qa = __builtin_amdgcn_ballot_w64( ra <= hi ) & ~__builtin_amdgcn_ballot_w64( ka );
            qb = __builtin_amdgcn_ballot_w64( rb <= hi ) & ~__builtin_amdgcn_ballot_w64( kb );
        }
        else
        {
            sa = __builtin_amdgcn_ballot_w64( ra <= lo );
            sb = __builtin_amdgcn_ballot_w64( rb <= lo );
            qa = __builtin_amdgcn_ballot_w64( ra <= hi ) & ~sa;
            qb = __builtin_amdgcn_ballot_w64( rb <= hi ) & ~sb;
        }
        if ( qa | qb )
        {
            const int sj = v0 + c * 64 + lane + dj[c];
            if ( qa )
                sa |= exact_lanes<HALF>( X, qa, sj, ia );
            if ( qb )
                sb |= exact_lanes<HALF>( X, qb, sj, ia + 1 );
        }
        ma[c] = sa & ( c == csa ? nsa : ~0ull );
        mb[c] = sb & ( c == csb ? nsb : ~0ull );
    }
}

template <int NC, bool HALF>
__device__ __forceinline__ void test_chunks32( const float ( &px )[BCH], const float ( &py )[BCH],
                                               const float ( &pz )[BCH], const float4 pa, int vs,
                                               float lo, float hi, float eps, const ExactCtx &X,
                                               const int ( &dj )[BCH], int v0, int ia,
                                               unsigned long long ( &m )[NC] )
{
    const int cself = vs >> 6;
    const unsigned long long notself = ~( 1ull << ( vs & 63 ) );
    const int lane = threadIdx.x & 63;
#pragma unroll
    for ( int c = 0; c < NC; ++c )
    {
        const float dx = pa.x - px[c];
        const float dy = pa.y - py[c];
        const float dz = pa.z - pz[c];
        const float r = fmaf( dz, dz, fmaf( dy, dy, dx * dx ) );
        unsigned long long sa, q;
        if ( HALF )
        {
            const bool k = ( r <= lo ) & ( fabsf( dx ) > eps );
            sa = __builtin_amdgcn_ballot_w64( k & ( dx < 0.f ) );
            q = __builtin_amdgcn_ballot_w64( r <= hi ) & ~__builtin_amdgcn_ballot_w64( k );
        }
        else
        {
            sa = __builtin_amdgcn_ballot_w64( r <= lo );
            q = __builtin_amdgcn_ballot_w64( r <= hi ) & ~sa;
        }
        if ( q )
            sa |= exact_lanes<HALF>( X, q, v0 + c * 64 + lane + dj[c], ia );
        m[c] = sa & ( c == cself ? notself : ~0ull );
    }
}

// The accepted candidates of one atom's masks appended to its row in chunk
// order.  A row holds cap entries and a spare slot at index cap: lanes that
// do not accept, and accepted entries that would fall past the row's
// capacity (the row length is still counted, and the list is rebuilt with
// larger rows), store into the spare slot, so the stores need no branch.
// Returns the new row length.
template <int NC>
__device__ __forceinline__ unsigned place_row( const int ( &po )[BCH],
                                               const unsigned long long ( &m )[NC],
                                               unsigned short *row, unsigned k0, unsigned cap )
{
#pragma unroll
    for ( int c = 0; c < NC; ++c )
    {
        unsigned e = cap;
        if ( __builtin_amdgcn_inverse_ballot_w64( m[c] ) )
            e = min( __builtin_amdgcn_mbcnt_hi( (unsigned)( m[c] >> 32 ),
                                                __builtin_amdgcn_mbcnt_lo( (unsigned)m[c], k0 ) ),
                     cap );
        row[e] = (unsigned short)po[c];
        k0 += (unsigned)__popcll( m[c] );
    }
    return k0;
}

// Lane l's value of v, as a wave-uniform (scalar) value.
__device__ __forceinline__ double readlane_f64( double v, int l )
{
    const unsigned long long b = (unsigned long long)__double_as_longlong( v );
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane( (int)(unsigned)b, l );
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane( (int)(unsigned)( b >> 32 ), l );
    return __longlong_as_double( (long long)( ( (unsigned long long)hi << 32 ) | lo ) );
}

__device__ __forceinline__ double4 lane_pos( const double4 &a, int l )
{
    return make_double4( readlane_f64( a.x, l ), readlane_f64( a.y, l ),
                         readlane_f64( a.z, l ), 0.0 );
}

__device__ __forceinline__ float readlane_f32( float v, int l )
{
    return __int_as_float( __builtin_amdgcn_readlane( __float_as_int( v ), l ) );
}

__device__ __forceinline__ float4 lane_pos( const float4 &a, int l )
{
    return make_float4( readlane_f32( a.x, l ), readlane_f32( a.y, l ),
                        readlane_f32( a.z, l ), 0.0f );
}

// The atoms 0 .. na - 1 of a group of a cell's atoms (lane l of `at` holds
// atom l's position; candidates vs0 + l of this pass, LDS rows row0 + l * RS)
// against NC register chunks of candidates, two atoms at a time.  An atom's
// position and row length are read out of their lanes as wave-uniform
// values, so the loop waits on no memory.  Lane l of cnt holds atom l's row
// length.  Coordinates are FP32 about the block's origin, with the band of
// B (and, HALF, the x-direction bound eps; pass start v0, local index i0 of
// the group's first atom) settled in FP64 (ExactCtx, dj: the chunks'
// candidate-to-slot offsets).
struct BandArgs
{
    float lo, hi, eps;
    int v0, i0;
};

template <int NC, int G, bool HALF, class PT, class P4>
__device__ __forceinline__ void build_atoms( const PT ( &px )[BCH], const PT ( &py )[BCH],
                                             const PT ( &pz )[BCH], const int ( &po )[BCH],
                                             const P4 &at, int na, int vs0,
                                             unsigned short *row0, int RS, double rc2,
                                             unsigned cap, int &cnt, const ExactCtx &X,
                                             const int ( &dj )[BCH], const BandArgs &B )
{
    int l = 0;
    for ( ; l + 1 < na; l += 2 )
    {
        const P4 pa = lane_pos( at, l ), pb = lane_pos( at, l + 1 );
        const unsigned ka0 = (unsigned)__builtin_amdgcn_readlane( cnt, l );
        const unsigned kb0 = (unsigned)__builtin_amdgcn_readlane( cnt, l + 1 );
        unsigned long long ma[NC], mb[NC];
        test_pair32<NC, HALF>( px, py, pz, pa, pb, vs0 + l, B.lo, B.hi, B.eps, X, dj, B.v0,
                               B.i0 + l, ma, mb );
        const unsigned ka = place_row<NC>( po, ma, row0 + l * RS, ka0, cap );
        const unsigned kb = place_row<NC>( po, mb, row0 + ( l + 1 ) * RS, kb0, cap );
        const int lane = threadIdx.x & 63;
        cnt = lane == l ? (int)ka : ( lane == l + 1 ? (int)kb : cnt );
    }
    if ( l < na )
    {
        const P4 pa = lane_pos( at, l );
        unsigned long long ma[NC];
        test_chunks32<NC, HALF>( px, py, pz, pa, vs0 + l, B.lo, B.hi, B.eps, X, dj, B.v0,
                                 B.i0 + l, ma );
        const unsigned ka = place_row<NC>( po, ma, row0 + l * RS,
                                           (unsigned)__builtin_amdgcn_readlane( cnt, l ), cap );
        if ( ( threadIdx.x & 63 ) == l )
            cnt = (int)ka;
    }
}

// LDS keeps the build at about three waves per SIMD, so the compiler may
// spend registers on overlapping the candidate chunks' independent tests
// instead of on occupancy it cannot reach.
template <int G, bool HALF>
__global__ __launch_bounds__( BTPB ) __attribute__( ( amdgpu_waves_per_eu( 2, 4 ) ) ) void k_tile_build(
    TileGrid T, const int *__restrict__ ls, const int *__restrict__ gs,
    const int *__restrict__ info, const int *__restrict__ woff,
    const double4 *__restrict__ x, const double4 *__restrict__ xg, double rc2, int rec,
    int cap, int slot_cap, int rows_cap, int waves_cap, unsigned short *__restrict__ slab,
    int *__restrict__ counts, int *__restrict__ max_row, int *__restrict__ over,
    int *__restrict__ waves_total, float band_lo, float band_hi, float band_eps )
{
    // coordinate planes in FP32 about the block's origin (tests above)
    using PT = float;
    using P4 = float4;
    extern __shared__ double s_dyn[];
    PT *s_x = reinterpret_cast<PT *>( s_dyn );
    __shared__ TileHead H;
    __shared__ int s_bseg[19];                // the block's 9-column segment offsets
    __shared__ int s_dl[9], s_dg[9];          // force slot - build slot, per column
    __shared__ int s_lcs[9][EMAX + 4], s_gcs[9][EMAX + 4];
    __shared__ int s_rs[BTPB / 64][MAXR];     // range starts (LDS slots)
    __shared__ int s_rf[BTPB / 64][MAXR];     // range starts in the force tile's slots
    __shared__ int s_rp[BTPB / 64][MAXR + 1]; // prefix of range lengths
    constexpr int APW = 64 / G;
    const int W = T.W, S = W + 2, NA = W * W;
    const int sub = xcd_tile( blockIdx.x, gridDim.x );
    const int t = sub / NA, c = sub % NA, ca = c / W, cb = c % W;
    if ( blockIdx.x == 0 && threadIdx.x == 0 )
        *waves_total = woff[T.ntiles]; // read back with the build's results
    load_head( info, t, W, H );
    const int nat_c = H.info[TI_AN + c];
    if ( nat_c == 0 )
        return; // block-uniform: an atom column outside the grid, or empty
    int a0, b0, e0, e1;
    tile_coords( T, t, a0, b0, e0, e1 );
    const int a = a0 + ca, b = b0 + cb;
    const int ea = H.info[TI_EA], nst = H.info[TI_NST];
    // the FP32 planes' origin: the centre of the atom column's staged cells
    const double ox = T.g.min[0] + ( a + 0.5 ) / T.g.rd[0];
    const double oy = T.g.min[1] + ( b + 0.5 ) / T.g.rd[1];
    const double oz = T.g.min[2] + ( ea + 0.5 * nst ) / T.g.rd[2];
    // the block's 3 x 3 columns: build column qb = da * 3 + db is staged
    // column (ca + da) * S + cb + db of the tile
    if ( threadIdx.x == 0 )
    {
        int off = 0;
        for ( int qb = 0; qb < 9; ++qb )
        {
            const int qf = ( ca + qb / 3 ) * S + cb + qb % 3;
            s_bseg[2 * qb] = off;
            s_dl[qb] = H.segoff[2 * qf] - off;
            off += H.info[TI_LLEN + qf];
            s_bseg[2 * qb + 1] = off;
            s_dg[qb] = H.segoff[2 * qf + 1] - off;
            off += H.info[TI_GLEN + qf];
        }
        s_bseg[18] = off;
    }
    __syncthreads();
    // the launch's sizes (planes, rows, slab) may come from an earlier
    // build: a block whose needs exceed them writes nothing and flags the
    // build, which is then redone with this build's sizes
    if ( s_bseg[18] + 1 > slot_cap || nat_c > rows_cap || woff[t + 1] > waves_cap )
    {
        if ( threadIdx.x == 0 )
            atomicOr( over, 1 );
        return;
    }
    // cell starts of the 3 x 3 columns in build slots (one entry per
    // thread: 9 (nst + 1) <= 9 (E + 3) <= 64 min(E, 8) threads); the
    // loads are issued here and their results stored after the staging, so
    // that both sets of loads are in flight together
    const int kq = threadIdx.x;
    const bool has_cs = kq < 9 * ( nst + 1 );
    int cs_l = 0, cs_g = 0;
    if ( has_cs )
    {
        const int qb = kq / ( nst + 1 ), ce = kq % ( nst + 1 );
        const int cl = col_base( T, a + qb / 3 - 1, b + qb % 3 - 1 ) + ea + ce;
        cs_l = ls[cl];
        cs_g = gs[cl];
    }
    // coordinate planes of the 3 x 3 columns' runs, then the sentinel
    PT *s_y = s_x + slot_cap, *s_z = s_x + 2 * slot_cap;
    const int btotal = s_bseg[18];
    {
        constexpr int B = 8;
        for ( int k0 = threadIdx.x; k0 < btotal; k0 += B * blockDim.x )
        {
            double4 p[B];
#pragma unroll
            for ( int u = 0; u < B; ++u )
            {
                const int k = min( k0 + u * (int)blockDim.x, btotal - 1 );
                const int seg = find_seg( s_bseg, 18, k );
                const int qb = seg >> 1, off = k - s_bseg[seg];
                const int qf = ( ca + qb / 3 ) * S + cb + qb % 3;
                p[u] = ( seg & 1 ) ? xg[H.info[TI_GB + qf] + off] : x[H.info[TI_LB + qf] + off];
            }
#pragma unroll
            for ( int u = 0; u < B; ++u )
            {
                const int k = min( k0 + u * (int)blockDim.x, btotal );
                s_x[k] = (PT)( p[u].x - ox );
                s_y[k] = (PT)( p[u].y - oy );
                s_z[k] = (PT)( p[u].z - oz );
            }
        }
        if ( has_cs )
        {
            const int qb = kq / ( nst + 1 ), ce = kq % ( nst + 1 );
            const int qf = ( ca + qb / 3 ) * S + cb + qb % 3;
            s_lcs[qb][ce] = s_bseg[2 * qb] + ( cs_l - H.info[TI_LB + qf] );
            s_gcs[qb][ce] = s_bseg[2 * qb + 1] + ( cs_g - H.info[TI_GB + qf] );
        }
        __syncthreads();
        if ( threadIdx.x == 0 )
        {
            s_x[btotal] = (PT)TILE_FAR;
            s_y[btotal] = (PT)TILE_FAR;
            s_z[btotal] = (PT)TILE_FAR;
        }
    }
    // after the planes (rounded to 16 bytes): the rows of this atom column's
    // atoms
    unsigned short *chunk = reinterpret_cast<unsigned short *>(
        reinterpret_cast<char *>( s_x ) + tile_planes_bytes<PT>( slot_cap ) );
    const int R4 = ( ( cap + G - 1 ) / G + 3 ) / 4;
    const int wsz = R4 * 256; // slab entries of one slab wave
    // the stride is an odd number of words, which spreads the copy-out's
    // strided reads over the banks
    const int RS = tile_row_stride( R4, G );
    const int al = H.info[TI_AL + c];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int nwv = blockDim.x >> 6;
    const int fsent = H.segoff[H.nseg]; // the force tile's sentinel slot
    // rows start as the force tile's sentinel record
    {
        const unsigned sent = (unsigned)( fsent * rec );
        const unsigned sent2 = sent | ( sent << 16 );
        const uint4 fill = make_uint4( sent2, sent2, sent2, sent2 );
        const int n16 = ( nat_c * RS + 7 ) / 8;
        for ( int k = threadIdx.x; k < n16; k += blockDim.x )
            reinterpret_cast<uint4 *>( chunk )[k] = fill;
    }
    __syncthreads();
    int mrow = 0;
    int *rs = s_rs[wv], *rf = s_rf[wv], *rp = s_rp[wv];
    const int bsent = btotal;
    const ExactCtx X{ x, xg, s_bseg, H.info, ca, cb, S, rc2 };
    for ( int e = e0 + wv; e < e1; e += nwv )
    {
        const int ce = e - ea;
        // stencil cells of this cell inside the staged range
        const int elo = max( ce - 1, 0 ), ehi = min( ce + 1, nst - 1 );
        // the stencil's ranges, lane r = range r: its first candidate index
        // (t_p) and the offsets that turn a candidate index into its build
        // slot (t_dj) and into its force record's byte offset (t_df).  Where
        // none of the 9 columns holds ghosts in the stencil's cells, column q
        // is the single range q (its locals of the three cells); otherwise
        // the ranges are laid out through LDS (per column the locals of its
        // three cells, or locals and ghosts cell by cell).
        bool gh = false;
        if ( lane < 9 )
            gh = s_gcs[lane][elo] != s_gcs[lane][ehi + 1];
        int nr, total, vbase;
        int t_p = 0, t_dj = 0, t_df = 0;
        if ( __builtin_amdgcn_ballot_w64( gh ) == 0 )
        {
            int len = 0, s0 = 0;
            if ( lane < 9 )
            {
                s0 = s_lcs[lane][elo];
                len = s_lcs[lane][ehi + 1] - s0;
            }
            int inc = len;
#pragma unroll
            for ( int o = 1; o < 16; o <<= 1 )
            {
                const int u = __shfl_up( inc, o, 64 );
                if ( lane >= o )
                    inc += u;
            }
            nr = 9;
            total = __builtin_amdgcn_readlane( inc, 8 );
            if ( lane < 9 )
            {
                t_p = inc - len;
                t_dj = s0 - t_p;
                t_df = ( s0 + s_dl[lane] - t_p ) * rec;
            }
            // candidate index of the cell's first atom: in the centre
            // column's (range 4's) locals, cell ce
            int vb = 0;
            if ( lane == 4 )
                vb = t_p + ( s_lcs[4][ce] - s0 );
            vbase = __builtin_amdgcn_readlane( vb, 4 );
        }
        else
        {
            // column q's ranges (lane q < 9), then their offsets in the list
            const int nr_q = lane < 9 ? ( gh ? 2 * ( ehi - elo + 1 ) : 1 ) : 0;
            int off = nr_q;
#pragma unroll
            for ( int o = 1; o < 16; o <<= 1 )
            {
                const int u = __shfl_up( off, o, 64 );
                if ( lane >= o )
                    off += u;
            }
            nr = __builtin_amdgcn_readlane( off, 8 );
            off -= nr_q;
            const int off0 = off;
            if ( lane < 9 )
            {
                if ( !gh )
                {
                    rs[off] = s_lcs[lane][elo];
                    rf[off] = s_lcs[lane][elo] + s_dl[lane];
                    rp[off + 1] = s_lcs[lane][ehi + 1] - s_lcs[lane][elo];
                }
                else
                {
                    for ( int c2 = elo; c2 <= ehi; ++c2 )
                    {
                        rs[off] = s_lcs[lane][c2];
                        rf[off] = s_lcs[lane][c2] + s_dl[lane];
                        rp[off + 1] = s_lcs[lane][c2 + 1] - s_lcs[lane][c2];
                        rs[off + 1] = s_gcs[lane][c2];
                        rf[off + 1] = s_gcs[lane][c2] + s_dg[lane];
                        rp[off + 2] = s_gcs[lane][c2 + 1] - s_gcs[lane][c2];
                        off += 2;
                    }
                }
            }
            __builtin_amdgcn_fence( __ATOMIC_SEQ_CST, "wavefront" );
            // prefix of the range lengths (nr <= MAXR < 64: one range per lane)
            const int len = lane < nr ? rp[lane + 1] : 0;
            int inc = len;
#pragma unroll
            for ( int o = 1; o < 64; o <<= 1 )
            {
                const int u = __shfl_up( inc, o, 64 );
                if ( lane >= o )
                    inc += u;
            }
            total = __builtin_amdgcn_readfirstlane( __shfl( inc, nr - 1, 64 ) );
            if ( lane < nr )
            {
                t_p = inc - len;
                t_dj = rs[lane] - t_p;
                t_df = ( rf[lane] - t_p ) * rec;
            }
            // candidate index of the cell's first atom, in the centre
            // column's (lane 4's) ranges: the locals of cell ce, its range
            // 2 (ce - elo) where the column holds ghosts, else at offset
            // ce - elo cells into its single range
            int vb = 0, vo = 0;
            if ( lane == 4 )
            {
                vb = gh ? off0 + 2 * ( ce - elo ) : off0;
                vo = gh ? 0 : s_lcs[4][ce] - s_lcs[4][elo];
            }
            const int rc = __builtin_amdgcn_readlane( vb, 4 );
            vbase = __builtin_amdgcn_readlane( t_p, rc ) + __builtin_amdgcn_readlane( vo, 4 );
            __builtin_amdgcn_fence( __ATOMIC_SEQ_CST, "wavefront" );
        }
        // the cell's atoms, up to 64 at a time (lane l holds atom a0 + l),
        // from the centre column's planes: build slot = atom index - dl4
        const int dl4 = H.info[TI_LB + ( ca + 1 ) * S + cb + 1] - s_bseg[8];
        // wave-uniform (LDS reads land in vector registers): kept scalar so
        // that the per-chunk self masks and row offsets stay scalar work
        const int c0 = __builtin_amdgcn_readfirstlane( s_lcs[4][ce] + dl4 );
        const int c1 = __builtin_amdgcn_readfirstlane( s_lcs[4][ce + 1] + dl4 );
        for ( int i0 = c0; i0 < c1; i0 += 64 )
        {
            const int na = min( 64, c1 - i0 );
            const int sa = i0 + min( lane, na - 1 ) - dl4;
            P4 at; // lane l: atom i0 + l
            at.x = s_x[sa];
            at.y = s_y[sa];
            at.z = s_z[sa];
            at.w = 0;
            int cnt = 0; // row length of atom i0 + lane
            int cur = 0; // range holding the pass's first candidate
            for ( int v0 = 0; v0 < total; v0 += BCH * 64 )
            {
                // BCH chunks of candidates in registers (lane = candidate);
                // past the end: the sentinel position, which no test accepts.
                // A candidate belongs to the last range that starts at or
                // before it (ranges in order; an empty range is passed over
                // by the next one, which starts at the same index): the
                // ranges starting inside the pass are applied in turn.
                while ( cur + 1 < nr && __builtin_amdgcn_readlane( t_p, cur + 1 ) <= v0 )
                    ++cur;
                int dj[BCH], df[BCH];
                {
                    const int d = __builtin_amdgcn_readlane( t_dj, cur );
                    const int f = __builtin_amdgcn_readlane( t_df, cur );
#pragma unroll
                    for ( int ch = 0; ch < BCH; ++ch )
                    {
                        dj[ch] = d;
                        df[ch] = f;
                    }
                }
                for ( int r = cur + 1; r < nr; ++r )
                {
                    const int pr = __builtin_amdgcn_readlane( t_p, r );
                    if ( pr >= v0 + BCH * 64 )
                        break;
                    const int d = __builtin_amdgcn_readlane( t_dj, r );
                    const int f = __builtin_amdgcn_readlane( t_df, r );
#pragma unroll
                    for ( int ch = 0; ch < BCH; ++ch )
                    {
                        const bool in = v0 + ch * 64 + lane >= pr;
                        dj[ch] = in ? d : dj[ch];
                        df[ch] = in ? f : df[ch];
                    }
                }
                PT px[BCH], py[BCH], pz[BCH];
                int po[BCH]; // the force tile's record byte offset, the value a row stores
#pragma unroll
                for ( int ch = 0; ch < BCH; ++ch )
                {
                    const int v = v0 + ch * 64 + lane;
                    const bool ok = v < total;
                    const int sj = ok ? v + dj[ch] : bsent;
                    po[ch] = ok ? v * rec + df[ch] : fsent * rec;
                    px[ch] = s_x[sj];
                    py[ch] = s_y[sj];
                    pz[ch] = s_z[sj];
                }
                const int nch = min( BCH, ( total - v0 + 63 ) >> 6 );
                const int vs0 = __builtin_amdgcn_readfirstlane( vbase + ( i0 - c0 ) - v0 );
                unsigned short *row0 = chunk + ( i0 - al ) * RS;
                const unsigned ucap = (unsigned)cap;
                const BandArgs B{ band_lo, band_hi, band_eps, v0, i0 };
                if ( nch <= 4 )
                    build_atoms<4, G, HALF>( px, py, pz, po, at, na, vs0, row0, RS, rc2,
                                             ucap, cnt, X, dj, B );
                else if ( nch <= 6 )
                    build_atoms<6, G, HALF>( px, py, pz, po, at, na, vs0, row0, RS, rc2,
                                             ucap, cnt, X, dj, B );
                else if ( nch <= 8 )
                    build_atoms<8, G, HALF>( px, py, pz, po, at, na, vs0, row0, RS, rc2,
                                             ucap, cnt, X, dj, B );
                else if ( nch == 9 )
                    build_atoms<9, G, HALF>( px, py, pz, po, at, na, vs0, row0, RS, rc2,
                                             ucap, cnt, X, dj, B );
                else
                    build_atoms<BCH, G, HALF>( px, py, pz, po, at, na, vs0, row0, RS,
                                               rc2, ucap, cnt, X, dj, B );
            }
            if ( lane < na )
            {
                counts[i0 + lane] = cnt;
                mrow = max( mrow, cnt );
            }
        }
        // the range tables are rewritten for the next cell
        __builtin_amdgcn_fence( __ATOMIC_SEQ_CST, "wavefront" );
    }
    __syncthreads();
    {
        // slab layout: per slab wave (gw = wave_lanes lanes per atom, 64 / gw
        // atoms), R4 groups of 64 eight-byte words; word (group g4, lane L)
        // holds entries (4 g4 + j) gw + L / (64 / gw), j = 0..3, of the
        // wave's atom L % (64 / gw).  Only the words of this atom column's
        // atoms (tile atoms ta0 .. ta0 + nat_c - 1) are written; entries at
        // or past cap are the sentinel.
        const int nat = H.info[TI_NAT], ta0 = H.aoff[c];
        const int w0 = ta0 / APW, w1 = ( ta0 + nat_c - 1 ) / APW;
        const unsigned short sent = (unsigned short)( fsent * rec );
        for ( int sw = w0; sw <= w1; ++sw )
        {
            // lanes per atom of slab wave sw (a power of two): lane L of a
            // word group holds the atom L % apw, entry column L / apw
            const int lg = __builtin_ctz( (unsigned)wave_lanes( G, nat, sw ) );
            const int sh = 6 - lg; // apw = 1 << sh
            const int aw = lane & ( ( 1 << sh ) - 1 ), sub = lane >> sh;
            const int ta = sw * APW + aw;
            if ( aw >= APW || ta < ta0 || ta >= ta0 + nat_c )
                continue;
            const unsigned short *row = chunk + ( ta - ta0 ) * RS;
            uint2 *dst = reinterpret_cast<uint2 *>( slab + (size_t)( woff[t] + sw ) * wsz );
            for ( int g4 = wv; g4 < R4; g4 += nwv )
            {
                unsigned e4[4];
#pragma unroll
                for ( int j = 0; j < 4; ++j )
                {
                    const int en = ( ( 4 * g4 + j ) << lg ) + sub;
                    e4[j] = en < cap ? row[en] : sent;
                }
                dst[g4 * 64 + lane] = make_uint2( e4[0] | ( e4[1] << 16 ), e4[2] | ( e4[3] << 16 ) );
            }
        }
    }
#pragma unroll
    for ( int o = 32; o > 0; o >>= 1 )
        mrow = max( mrow, __shfl_xor( mrow, o, 64 ) );
    if ( lane == 0 && mrow > 0 )
        atomicMax( max_row, mrow );
}

// CSR view: slots mapped back to system indices (locals by run start,
// ghosts through the ghost sort permutation).
template <int G>
__global__ __launch_bounds__( TPB ) void k_tile_to_csr(
    TileGrid T, const int *__restrict__ info, const int *__restrict__ woff,
    int cap, int n_local, int rec, const unsigned short *__restrict__ slab,
    const int *__restrict__ counts, const int *__restrict__ gperm,
    const int64_t *__restrict__ offsets, int *__restrict__ idx )
{
    __shared__ TileHead H;
    constexpr int APW = 64 / G;
    const int t = blockIdx.x;
    load_head( info, t, T.W, H );
    const int R4 = ( ( cap + G - 1 ) / G + 3 ) / 4;
    const unsigned short *st = slab + (size_t)woff[t] * R4 * 256;
    const int nat = H.info[TI_NAT];
    for ( int ta = threadIdx.x; ta < nat; ta += blockDim.x )
    {
        const int i = tile_atom( H, T.W, ta );
        const int w = ta / APW, aw = ta % APW;
        const int gw = wave_lanes( G, nat, w ), apw = 64 / gw;
        const int n = counts[i];
        const int64_t o = offsets[i];
        for ( int k = 0; k < n; ++k )
        {
            const int r = k / gw, lane = ( k % gw ) * apw + aw;
            const int sj =
                st[( ( w * R4 + ( r >> 2 ) ) * 64 + lane ) * 4 + ( r & 3 )] / rec;
            const int seg = find_seg( H.segoff, H.nseg, sj );
            const int q = seg >> 1, off = sj - H.segoff[seg];
            idx[o + k] = ( seg & 1 ) ? n_local + gperm[H.info[TI_GB + q] + off]
                                     : H.info[TI_LB + q] + off;
        }
    }
}

// Ghost refresh for sorted ghosts: the value of update_halo
// (comm_mpi_impl.h:369-408) written both to the ghost's system slot and to
// its sorted slot.
__global__ void k_ghost_refresh_sorted( int n_local, int n_ghost,
                                        double4 *__restrict__ x,
                                        double4 *__restrict__ xg,
                                        const int *__restrict__ gperm,
                                        const int *__restrict__ root,
                                        const int *__restrict__ code, double Lx,
                                        double Ly, double Lz )
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if ( k >= n_ghost )
        return;
    const int g = gperm[k];
    double4 p = x[root[g]];
    const int c = code[g];
    if ( c & 1 )
        p.x = p.x - Lx;
    if ( c & 2 )
        p.x = p.x + Lx;
    if ( c & 4 )
        p.y = p.y - Ly;
    if ( c & 8 )
        p.y = p.y + Ly;
    if ( c & 16 )
        p.z = p.z - Lz;
    if ( c & 32 )
        p.z = p.z + Lz;
    x[n_local + g] = p;
    xg[k] = p;
}

__global__ void k_gather_ghosts( int n_local, int n_ghost,
                                 const int *__restrict__ gperm,
                                 const double4 *__restrict__ x,
                                 double4 *__restrict__ xg )
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if ( k < n_ghost )
        xg[k] = x[n_local + gperm[k]];
}

// Root of the ghost at each sorted slot: the key that groups the slots by
// local atom (the images the fused step writes).
__global__ void k_image_roots( int n_ghost, const int *__restrict__ gperm,
                               const int *__restrict__ root, int *__restrict__ key )
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if ( k < n_ghost )
        key[k] = root[gperm[k]];
}

// Image entries in root order: sorted slot and shift code in one word.
__global__ void k_image_pack( int n_ghost, const int *__restrict__ slot,
                              const int *__restrict__ gperm,
                              const int *__restrict__ code, int *__restrict__ pk )
{
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if ( m < n_ghost )
    {
        const int k = slot[m];
        pk[m] = (int)( (unsigned)k | ( (unsigned)code[gperm[k]] << IMG_CODE_SHIFT ) );
    }
}

cmd_status tiles_prepare_ghosts( cmd_system *S )
{
    S->xg_valid = false;
    if ( !S->bin_grid_valid )
        return CMD_OK;
    const int ng = S->n_ghost;
    CMD_TRY( cell_sort( S->x.p + S->n_local, ng, S->bin_grid, S->gsort, S->sc,
                        S->stream ) );
    CMD_TRY( S->xg.reserve( std::max( ng, 1 ) ) );
    CMD_TRY( S->xg_alt.reserve( std::max( ng, 1 ) ) );
    if ( ng > 0 )
    {
        k_gather_ghosts<<<( ng + TPB - 1 ) / TPB, TPB, 0, S->stream>>>(
            S->n_local, ng, S->gsort.perm.p, S->x.p, S->xg.p );
        CMD_LAUNCH_CHECK();
    }
    // sorted ghost slots grouped by their local atom
    CMD_TRY( S->gimg.cell_of.reserve( std::max( ng, 1 ) ) );
    if ( ng > 0 )
    {
        k_image_roots<<<( ng + TPB - 1 ) / TPB, TPB, 0, S->stream>>>(
            ng, S->gsort.perm.p, S->ghost_root.p, S->gimg.cell_of.p );
        CMD_LAUNCH_CHECK();
    }
    CMD_TRY( cell_sort_ids( ng, std::max( S->n_local, 1 ), S->gimg, S->sc, S->stream ) );
    S->img_valid = ng <= IMG_SLOT_MASK + 1;
    if ( S->img_valid )
    {
        CMD_TRY( S->img_pk.reserve( std::max( ng, 1 ) ) );
        if ( ng > 0 )
        {
            k_image_pack<<<( ng + TPB - 1 ) / TPB, TPB, 0, S->stream>>>(
                ng, S->gimg.perm.p, S->gsort.perm.p, S->ghost_code.p, S->img_pk.p );
            CMD_LAUNCH_CHECK();
        }
    }
    S->xg_valid = true;
    S->ghosts_x_stale = false;
    return CMD_OK;
}

cmd_status tiles_ghost_refresh( cmd_system *S )
{
    S->ghosts_x_stale = false;
    if ( S->n_ghost == 0 )
        return CMD_OK;
    k_ghost_refresh_sorted<<<( S->n_ghost + TPB - 1 ) / TPB, TPB, 0, S->stream>>>(
        S->n_local, S->n_ghost, S->x.p, S->xg.p, S->gsort.perm.p, S->ghost_root.p,
        S->ghost_code.p, S->dom.L[0], S->dom.L[1], S->dom.L[2] );
    CMD_LAUNCH_CHECK();
    return CMD_OK;
}

// Tile info for width W and height E, launched without waiting: the tile
// table, slab waves per tile and their scan (total at start[ntiles]), and
// the maxima in the result words: [0] largest staged count of a tile, [1]
// local atoms the tiles cover, [2] most slab waves of a tile, [3] largest
// staged count of one atom column's 3 x 3 columns, [4] most atoms of one
// atom column.
static cmd_status tile_info_launch( cmd_neigh *N, cmd_system *S, int W, int E )
{
    TileGrid &T = N->tg;
    T.g = S->bin_grid;
    T.nc0 = T.g.n[0] - 2;
    T.nc1 = T.g.n[1] - 2;
    T.W = W;
    T.nt0 = ( T.nc0 + W - 1 ) / W;
    T.nt1 = ( T.nc1 + W - 1 ) / W;
    T.E = E;
    T.tpc = ( T.g.n[2] - 2 + E - 1 ) / E;
    T.ntiles = T.nt0 * T.nt1 * T.tpc;
    CMD_REQUIRE( T.nc0 > 0 && T.nc1 > 0 && T.tpc > 0, CMD_ERR_ARG,
                 "binning grid has no interior cells" );
    CMD_TRY( N->tile_info.reserve( (size_t)T.ntiles * TI_WORDS ) );
    CMD_TRY( S->sc.ires.reserve( TILE_RES_WORDS ) );
    DevBuf<int> &waves = N->grid_sort.cursor; // scratch: waves per tile
    CMD_TRY( waves.reserve( T.ntiles + 1 ) );
    CMD_TRY( N->grid_sort.start.reserve( T.ntiles + 1 ) ); // scan of waves
    CMD_TRY( fill_i32( S->sc.ires.p + TILE_RES_INFO, 0, 5, S->stream ) );
    const int apw = 64 / N->lanes;
    // one wave per block: a thread does a whole tile (about a hundred
    // dependent loads), so the tiles are spread over as many CUs as possible
    k_tile_info<<<( T.ntiles + 63 ) / 64, 64, 0, S->stream>>>(
        T, S->bins.start.p, S->gsort.start.p, apw, N->tile_info.p, waves.p,
        S->sc.ires.p + TILE_RES_INFO );
    CMD_LAUNCH_CHECK();
    CMD_TRY( scan_exclusive( waves.p, N->grid_sort.start.p, T.ntiles, S->sc,
                             S->stream, nullptr ) );
    return CMD_OK;
}

// Reads the result words (res, TILE_RES_WORDS of them) and the slab waves
// of all tiles (h5) in one wait, through the system's pinned words: after a
// build (built), which stores the slab waves among the result words, in one
// copy.
static cmd_status tile_results( cmd_neigh *N, cmd_system *S, int *res, int *h5, bool built )
{
    CMD_TRY( S->hw.init( HW_WORDS ) );
    int *hr = S->hw.p + HW_TILE;
    CMD_HIP( hipMemcpyAsync( hr, S->sc.ires.p, TILE_RES_WORDS * sizeof( int ),
                             hipMemcpyDeviceToHost, S->stream ) );
    if ( !built )
        CMD_HIP( hipMemcpyAsync( hr + TILE_RES_WAVES, N->grid_sort.start.p + N->tg.ntiles,
                                 sizeof( int ), hipMemcpyDeviceToHost, S->stream ) );
    CMD_HIP( hipStreamSynchronize( S->stream ) );
    for ( int k = 0; k < TILE_RES_WORDS; ++k )
        res[k] = hr[k];
    *h5 = res[TILE_RES_WAVES];
    return CMD_OK;
}

// Waves per block of the force kernels: as tuned, else one per slab wave of
// the largest tile, at most 8 (a tile with more loops).
static int force_waves( const cmd_neigh *N, int most_waves )
{
    return N->tile_bw > 0 ? N->tile_bw : std::min( std::max( most_waves, 1 ), 8 );
}

static size_t tile_slab_entries( const cmd_neigh *N, int nw )
{
    const int G = N->lanes;
    const size_t R4 = (size_t)( ( ( N->cap + G - 1 ) / G + 3 ) / 4 );
    return std::max<size_t>( (size_t)nw * R4 * 256, 1 );
}

// LDS of one block (hard limit, less the static tables) and of one CU.
// The device attributes are read once per device.
static void lds_limits( size_t *block, size_t *cu )
{
    constexpr int NDEV = 64;
    static std::atomic<int> cached_lim[NDEV], cached_lcu[NDEV];
    int dev = 0, lim = 0, lcu = 0;
    if ( hipGetDevice( &dev ) != hipSuccess || dev < 0 || dev >= NDEV )
        dev = -1;
    if ( dev >= 0 )
    {
        lim = cached_lim[dev].load();
        lcu = cached_lcu[dev].load();
    }
    if ( lim <= 0 || lcu <= 0 )
    {
        if ( dev < 0 ||
             hipDeviceGetAttribute( &lim, hipDeviceAttributeMaxSharedMemoryPerBlock,
                                    dev ) != hipSuccess ||
             lim <= 0 )
            lim = 65536;
        if ( dev < 0 ||
             hipDeviceGetAttribute( &lcu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor,
                                    dev ) != hipSuccess ||
             lcu < lim )
            lcu = lim;
        if ( dev >= 0 )
        {
            cached_lim[dev].store( lim );
            cached_lcu[dev].store( lcu );
        }
    }
    const size_t statics = 8192; // tile header, cell and range tables
    *block = (size_t)lim > statics ? (size_t)lim - statics : 0;
    *cu = (size_t)lcu;
}

// LDS of the force kernels on tiles with `slots` staged atoms: a position
// record and, with several types, a type per slot (lj.hip; FP32 records
// hold the type); the half list adds a reaction record and an energy weight
// per slot.
static size_t force_lds( int slots, bool half, int rec )
{
    const size_t per = rec == TILE_REC32 ? (size_t)TILE_REC32 : TILE_REC + sizeof( int );
    return (size_t)( slots + 1 ) * per * ( half ? 2 : 1 ) + 16;
}

template <int G>
static void launch_build( cmd_neigh *N, cmd_system *S, int slot_cap, int rows_cap,
                          int waves_cap, size_t lds )
{
    const double rc2 = N->neigh_cut * N->neigh_cut;
    // the list build tests in FP32 about each block's origin: every
    // staged coordinate lies within amax of it (1.5 cells across, half the
    // staged cells along the column), so the FP32 r^2 is within
    // 128 amax^2 2^-24 of the FP64 one (roundings of the coordinates, their
    // differences, the squares and the sums); candidates within that band of
    // rc^2 are tested again in FP64
    const CellGrid &g = N->tg.g;
    const double amax = 1.001 * std::max( { 1.5 / g.rd[0], 1.5 / g.rd[1],
                                            0.5 * ( N->tg.E + 2 ) / g.rd[2] } );
    double band = 128.0 * amax * amax * std::ldexp( 1.0, -24 ) + 1e-6 * rc2;
    // CMD_BUILD_BAND widens the band (never narrows it), so that tests can
    // send most candidates through the FP64 re-test
    if ( const char *e = std::getenv( "CMD_BUILD_BAND" ) )
        band = std::max( band, std::atof( e ) );
    const float lo = (float)( rc2 - band ), hi = (float)( rc2 + band );
    // the half list's x direction: an FP32 difference of two coordinates
    // within amax of the origin is within 4 amax 2^-24 of the FP64 one
    const float eps = (float)( 8.0 * amax * std::ldexp( 1.0, -24 ) );
    int *res = S->sc.ires.p;
    // one block per atom column of a tile, one wave per cell of the column
    const int threads = 64 * std::min( std::max( N->tg.E, 1 ), BTPB / 64 );
    const int nb = N->tg.ntiles * N->tg.W * N->tg.W;
    if ( N->half )
        k_tile_build<G, true><<<nb, threads, lds, S->stream>>>(
            N->tg, S->bins.start.p, S->gsort.start.p, N->tile_info.p,
            N->grid_sort.start.p, S->x.p, S->xg.p, rc2, N->rec, N->cap, slot_cap, rows_cap,
            waves_cap, N->tslab.p, N->counts.p, res + TILE_RES_MAXROW, res + TILE_RES_OVER,
            res + TILE_RES_WAVES, lo, hi, eps );
    else
        k_tile_build<G, false><<<nb, threads, lds, S->stream>>>(
            N->tg, S->bins.start.p, S->gsort.start.p, N->tile_info.p,
            N->grid_sort.start.p, S->x.p, S->xg.p, rc2, N->rec, N->cap, slot_cap, rows_cap,
            waves_cap, N->tslab.p, N->counts.p, res + TILE_RES_MAXROW, res + TILE_RES_OVER,
            res + TILE_RES_WAVES, lo, hi, eps );
}

static cmd_status launch_build_g( cmd_neigh *N, cmd_system *S, int slot_cap, int rows_cap,
                                  int waves_cap, size_t lds )
{
    // the build's result words start at zero (the device test of the
    // bounding box, when used, is left as it was written)
    CMD_TRY( fill_i32( S->sc.ires.p + TILE_RES_MAXROW, 0, 2, S->stream ) );
    switch ( N->lanes )
    {
    case 2:
        launch_build<2>( N, S, slot_cap, rows_cap, waves_cap, lds );
        break;
    case 4:
        launch_build<4>( N, S, slot_cap, rows_cap, waves_cap, lds );
        break;
    default:
        launch_build<8>( N, S, slot_cap, rows_cap, waves_cap, lds );
        break;
    }
    CMD_LAUNCH_CHECK();
    return CMD_OK;
}

// LDS of the build block: coordinate planes of its 3 x 3 columns and the
// sentinel slot, then the rows of its atom column (each with a spare slot
// past its cap entries).
static size_t build_lds_cap( int G, int cap, int slot_cap, int rows )
{
    const int R4 = ( ( cap + G - 1 ) / G + 3 ) / 4;
    return tile_planes_bytes<float>( slot_cap ) +
           ( (size_t)std::max( rows, 1 ) * tile_row_stride( R4, G ) + 8 ) * sizeof( unsigned short );
}

static size_t build_lds( const cmd_neigh *N, int slot_cap, int rows )
{
    return build_lds_cap( N->lanes, N->cap, slot_cap, rows );
}

// Tile height asked for: the set one, else 4 cells for single columns and 3
// for wider tiles.
static int requested_height( const cmd_neigh *N, int W )
{
    return std::min( N->tile_e > 0 ? N->tile_e : ( W == 1 ? 4 : 3 ), EMAX - 2 );
}

// Whether a layout with tile maxima h and build LDS `lds` fits the device
// (fits) and keeps the force kernel's blocks per CU (blocks) at the tuned
// level; returns both tests together.
static bool layout_ok( const cmd_neigh *N, const int h[6], size_t lds, size_t hard,
                       size_t lds_cu, int E, bool *fits, int *blocks )
{
    const size_t flds = force_lds( h[0], N->half != 0, N->rec );
    *fits = h[0] <= max_force_slots( N->rec ) && lds <= hard && flds <= hard;
    // blocks per CU of the force kernel (its static tables included)
    *blocks = (int)std::min( lds_cu / ( flds + 1024 ), (size_t)( 32 / N->tg.bw ) );
    return *fits && ( N->occ_blocks == 0 || *blocks >= N->occ_blocks || E == 1 );
}

// Sizes for the next build's launch: this build's, with headroom for the
// atoms that move between columns meanwhile.  The row capacity grows ahead
// of the rows: a start from a lattice melts, and its longest row keeps
// growing for many rebuilds; a row past the capacity would send the next
// build through the exact path (a second build).  mr is this build's
// longest row.
static void record_spec( cmd_neigh *N, int W, int E, const int h[6], int mr, size_t hard )
{
    cmd_neigh::Spec &P = N->spec;
    P.valid = false;
    if ( E != requested_height( N, W ) )
        return; // a lowered height is chosen afresh at each build
    P.G = N->lanes;
    P.half = N->half;
    P.W = W;
    P.E = E;
    P.bw_req = N->tile_bw;
    P.cap = N->cap;
    if ( mr + std::max( 4, mr / 16 ) > P.cap )
        P.cap = mr + std::max( 8, mr / 8 );
    P.slot_cap = std::min( h[3] + 1 + std::max( 16, h[3] / 32 ), MAX_SLOTS + 1 );
    P.rows = h[4] + std::max( 2, h[4] / 20 );
    P.slab_waves = h[5] + std::max( 64, h[5] / 50 );
    P.lds = build_lds_cap( N->lanes, P.cap, P.slot_cap, P.rows );
    P.valid = P.lds <= hard;
}

// With CMD_VERBOSE set: the layout a tiled build chose and the sizes that
// set the force kernel's LDS (largest staged neighbourhood) and blocks per CU.
static void report_layout( const cmd_neigh *N, const char *how, int W, int E, const int h[6],
                           size_t build_lds )
{
    if ( !std::getenv( "CMD_VERBOSE" ) )
        return;
    std::fprintf( stderr,
                  "# %s tiled build %dx%dx%d: tile slots max %d, 3x3 column slots max %d, "
                  "column atoms max %d, slab waves %d, force blocks/CU %d, waves/block %d, "
                  "row capacity %d, build LDS %zu B\n",
                  how, W, W, E, h[0], h[3], h[4], h[5], N->tile_blocks, N->tg.bw, N->cap,
                  build_lds );
}

cmd_status tiles_build( cmd_neigh *N, cmd_system *S, bool *built )
{
    *built = false;
    N->spec.valid = false;
    if ( N->half )
        N->rec = TILE_REC; // the half-list kernel stages FP64 records
    if ( !S->bin_grid_valid || !S->xg_valid || N->lanes < 2 )
        return CMD_OK;
    if ( N->cap < 1 )
        N->cap = std::max( 1, N->max_neigh_guess );
    size_t hard = 0, lds_cu = 0;
    lds_limits( &hard, &lds_cu );
    // tile width and height: the requested ones (2 x 2 columns, 3 cells by
    // default), lowered while the neighbourhood and rows do not fit in LDS,
    // or while the force kernel would fit fewer blocks per CU than on the
    // layout the autotuner chose (occ_blocks): a denser neighbourhood than
    // the tuned one (the perfect lattice at step 0) would otherwise lose a
    // block of occupancy
    const int W = N->tile_w > 0 ? N->tile_w : 2;
    int E = requested_height( N, W );
    int h[6] = { 0, 0, 0, 0, 0, 0 };
    int res[TILE_RES_WORDS];
    for ( int attempt = 0; attempt < 4; ++attempt )
    {
        size_t lds = 0;
        for ( ;; )
        {
            CMD_TRY( tile_info_launch( N, S, W, E ) );
            CMD_TRY( tile_results( N, S, res, &h[5], false ) );
            for ( int k = 0; k < 5; ++k )
                h[k] = res[TILE_RES_INFO + k];
            N->tg.bw = force_waves( N, h[2] );
            // local atoms in the grid's padding cells (not wrapped into the
            // box) belong to no tile: use the global layout
            if ( h[1] != S->n_local )
                return CMD_OK;
            lds = build_lds( N, h[3] + 1, h[4] );
            bool fits = false;
            int blocks = 0;
            if ( layout_ok( N, h, lds, hard, lds_cu, E, &fits, &blocks ) )
            {
                N->tile_blocks = blocks;
                break;
            }
            if ( E == 1 )
                return CMD_OK; // too dense for LDS tiles: global path
            E = fits ? E - 1 : std::max( 1, E * 2 / 3 );
        }
        N->max_tile_slots = h[0];
        // the slab sized for the next build's launch too (record_spec)
        CMD_TRY( N->tslab.reserve( tile_slab_entries( N, h[5] + std::max( 64, h[5] / 50 ) ) ) );
        CMD_TRY( launch_build_g( N, S, h[3] + 1, h[4], h[5], lds ) );
        CMD_TRY( tile_results( N, S, res, &h[5], true ) );
        const int mr = res[TILE_RES_MAXROW];
        N->max_row = mr;
        if ( mr > N->max_neigh_guess )
            N->max_neigh_guess = (int)( mr * 1.1 );
        if ( mr <= N->cap )
        {
            *built = true;
            N->tiled = true;
            N->built_rec = N->rec;
            N->gperm_view = S->gsort.perm.p;
            record_spec( N, W, E, h, mr, hard );
            report_layout( N, "exact", W, E, h, lds );
            return CMD_OK;
        }
        // rows overflowed: larger rows, and the LDS fit rechecked
        N->cap = std::max( N->max_neigh_guess, mr );
    }
    set_error( "neighbour rows kept growing during the tiled build" );
    return CMD_ERR_CAPACITY;
}

// The tiled build launched with the sizes of the last one (record_spec) and
// read back once, after the build: tile maxima, longest row, the capacity
// flag of the build blocks and the device's bounding-box test (written by
// the caller into TILE_RES_INSIDE).  Built only when every test passes and
// the layout is the one tiles_build would choose from these maxima;
// otherwise *built stays false and the caller builds afresh.
cmd_status tiles_build_spec( cmd_neigh *N, cmd_system *S, bool *built )
{
    *built = false;
    const cmd_neigh::Spec P = N->spec;
    N->spec.valid = false;
    if ( !P.valid || !S->bin_grid_valid || !S->xg_valid || P.G != N->lanes ||
         P.half != N->half || P.cap < N->cap || P.bw_req != N->tile_bw )
        return CMD_OK;
    const int W = N->tile_w > 0 ? N->tile_w : 2;
    if ( P.W != W || P.E != requested_height( N, W ) )
        return CMD_OK;
    N->cap = P.cap; // grown ahead of the rows (record_spec)
    CMD_TRY( tile_info_launch( N, S, P.W, P.E ) );
    CMD_TRY( N->tslab.reserve( tile_slab_entries( N, P.slab_waves ) ) );
    CMD_TRY( launch_build_g( N, S, P.slot_cap, P.rows, P.slab_waves, P.lds ) );
    int res[TILE_RES_WORDS], h[6] = { 0, 0, 0, 0, 0, 0 };
    CMD_TRY( tile_results( N, S, res, &h[5], true ) );
    for ( int k = 0; k < 5; ++k )
        h[k] = res[TILE_RES_INFO + k];
    if ( res[TILE_RES_INSIDE] != 1 || res[TILE_RES_OVER] != 0 || h[1] != S->n_local )
        return CMD_OK;
    size_t hard = 0, lds_cu = 0;
    lds_limits( &hard, &lds_cu );
    N->tg.bw = force_waves( N, h[2] );
    bool fits = false;
    int blocks = 0;
    if ( !layout_ok( N, h, build_lds( N, h[3] + 1, h[4] ), hard, lds_cu, P.E, &fits, &blocks ) )
        return CMD_OK;
    const int mr = res[TILE_RES_MAXROW];
    N->max_row = mr;
    if ( mr > N->max_neigh_guess )
        N->max_neigh_guess = (int)( mr * 1.1 );
    if ( mr > N->cap )
    {
        // rows overflowed: the caller's build starts at the grown rows
        N->cap = std::max( N->max_neigh_guess, mr );
        return CMD_OK;
    }
    N->tile_blocks = blocks;
    N->max_tile_slots = h[0];
    N->tiled = true;
    N->built_rec = N->rec;
    N->gperm_view = S->gsort.perm.p;
    *built = true;
    record_spec( N, P.W, P.E, h, mr, hard );
    report_layout( N, "speculative", P.W, P.E, h, P.lds );
    return CMD_OK;
}

cmd_status tiles_csr( cmd_neigh *N, int64_t *nnz_out )
{
    // offsets already scanned by the caller into N->offsets
    const int nt = N->tg.ntiles;
    switch ( N->lanes )
    {
    case 2:
        k_tile_to_csr<2><<<nt, TPB, 0, N->stream>>>(
            N->tg, N->tile_info.p, N->grid_sort.start.p, N->cap, N->n_rows, N->built_rec,
            N->tslab.p, N->counts.p, N->gperm_view, N->offsets.p, N->idx.p );
        break;
    case 4:
        k_tile_to_csr<4><<<nt, TPB, 0, N->stream>>>(
            N->tg, N->tile_info.p, N->grid_sort.start.p, N->cap, N->n_rows, N->built_rec,
            N->tslab.p, N->counts.p, N->gperm_view, N->offsets.p, N->idx.p );
        break;
    default:
        k_tile_to_csr<8><<<nt, TPB, 0, N->stream>>>(
            N->tg, N->tile_info.p, N->grid_sort.start.p, N->cap, N->n_rows, N->built_rec,
            N->tslab.p, N->counts.p, N->gperm_view, N->offsets.p, N->idx.p );
        break;
    }
    CMD_LAUNCH_CHECK();
    (void)nnz_out;
    return CMD_OK;
}

} // namespace cmd
