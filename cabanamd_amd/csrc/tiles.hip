// tiles.hip -- LDS-tiled Verlet lists and LJ force for binned systems.
//
// Same list as NeighborVerlet::create (neighbor_verlet.h:43-62) and the same
// forces as ForceLJ::compute_force_full (force_lj_cabana_neigh_impl.h:
// 151-203); only the work decomposition and the list encoding differ.
//
// A tile is a run of E cells [e0, e1) of one interior column (a, b) of the
// binning grid (cells >= the neighbour radius).  Its atoms are the contiguous
// local range of those cells, and every neighbour candidate lies in the 27-cell
// stencils of those cells: 9 columns (a+-1, b+-1) x cells [e0-1, e1].  Per
// column that is one contiguous range of local atoms (locals are cell sorted
// by create_binning) and one contiguous range of ghosts (ghosts are cell
// sorted here).  A block stages those 18 ranges of positions in LDS and the
// list stores, for each neighbour, the 16-bit byte offset of its LDS record
// (slot * TILE_REC).  Rows are walked in
// the order the oracle uses (stencil columns, then cells, then atoms by
// index), so lists and FP64 sums are unchanged.
//
// Slab of a tile: waves of APW = 64/G atoms; with gw lanes per atom in wave
// w (G, or up to 8 in a part-full wave: wave_lanes), neighbour k of atom a
// of the wave goes to row r = k / gw, lane (k % gw) * (64 / gw) + a, and
// four consecutive rows of a lane are packed in 8 bytes:
//   tslab[tile_base + ((w * R4 + r / 4) * 64 + lane) * 4 + r % 4],
// R4 = ceil(ceil(cap / G) / 4), so one 8-byte load per lane feeds four pairs.
#include "internal.h"
#include "tiles_dev.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>

namespace cmd
{
static constexpr int TPB = 256;
static constexpr int BTPB = 512;       // list build block
static constexpr int EMAX = 16;
static constexpr int MAX_SLOTS = 1792; // 42 KB of 24-byte records per block
static_assert( ( MAX_SLOTS + 1 ) * TILE_REC + 16 < 65536,
               "record offsets must fit the 16-bit list entries" );
// Per-tile runs, atom range and staged count; waves per tile for the slab.
// slots[0] gets the largest staged count, slots[1] the number of local atoms
// the tiles cover (all of them only when every local atom sits in an
// interior cell of the grid), slots[2] the most slab waves of one tile.
__global__ void k_tile_info( TileGrid T, const int *__restrict__ ls,
                             const int *__restrict__ gs, int apw,
                             int *__restrict__ info, int *__restrict__ waves,
                             int *__restrict__ slots )
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    int total = 0, nat = 0, nwv = 0;
    if ( t < T.ntiles )
    {
        int a, b, e0, e1;
        tile_coords( T, t, a, b, e0, e1 );
        const int ea = max( e0 - 1, 0 ), eb = min( e1 + 1, T.g.n[2] );
        int *I = info + (size_t)t * TI_WORDS;
        int q = 0;
        for ( int da = -1; da <= 1; ++da )
            for ( int db = -1; db <= 1; ++db )
            {
                const int c = col_base( T, a + da, b + db );
                I[TI_LB + q] = ls[c + ea];
                I[TI_LLEN + q] = ls[c + eb] - ls[c + ea];
                I[TI_GB + q] = gs[c + ea];
                I[TI_GLEN + q] = gs[c + eb] - gs[c + ea];
                total += I[TI_LLEN + q] + I[TI_GLEN + q];
                ++q;
            }
        const int cc = col_base( T, a, b );
        I[TI_TL] = ls[cc + e0];
        I[TI_NAT] = ls[cc + e1] - ls[cc + e0];
        I[TI_TOTAL] = total;
        I[TI_EA] = ea;
        I[TI_NST] = eb - ea;
        nat = I[TI_NAT];
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
    }
    if ( ( threadIdx.x & 63 ) == 0 )
    {
        if ( m > 0 )
            atomicMax( &slots[0], m );
        if ( nat > 0 )
            atomicAdd( &slots[1], nat );
        if ( nwv > 0 )
            atomicMax( &slots[2], nwv );
    }
}

template <bool HALF>
__device__ __forceinline__ bool tile_accept( const double4 &pi, const double4 &pj,
                                             double rc2 )
{
    const double dx = pi.x - pj.x;
    const double dy = pi.y - pj.y;
    const double dz = pi.z - pj.z;
    const double rsq = dx * dx + dy * dy + dz * dz;
    bool ok = rsq <= rc2;
    if ( HALF )
        ok = ok && ( ( pj.x > pi.x ) |
                     ( ( pj.x == pi.x ) & ( ( pj.y > pi.y ) |
                                            ( ( pj.y == pi.y ) & ( pj.z > pi.z ) ) ) ) );
    return ok;
}

// List build, one wave per cell of the tile.  All atoms of a cell share the
// same 27-cell stencil, so the wave loads the stencil's candidates into
// registers, BCH chunks of 64 at a time (lane = candidate, positions from
// LDS planes), and tests every atom of the cell against them in turn, the
// atom's position and row length being wave-uniform (scalar loads and
// registers); ballot + mbcnt place each accepted candidate in the atom's
// row, so rows come out in walk order.  The
// stencil is a list of contiguous LDS ranges (per column: the locals of its
// three cells, or locals and ghosts cell by cell where ghosts are present),
// concatenated so that chunks run across range ends.  Rows are assembled in
// LDS one after another (pre-filled with the sentinel slot), so an entry's
// address is the row start plus its index, and are transposed into the slab
// layout on the way out.
static constexpr int MAXR = 9 * 3 * 2; // stencil ranges of one cell
static constexpr int BCH = 10;         // candidate chunks of 64 held in registers

// Bytes of the coordinate planes of `slots` staged atoms, rounded up so that
// what follows is 16-byte aligned.
__host__ __device__ constexpr size_t tile_planes_bytes( int slots )
{
    return ( (size_t)slots * 3 * sizeof( double ) + 15 ) & ~(size_t)15;
}

// Entries of one LDS row of the list build: the slab's entries per atom
// (R4 groups of 4 G) plus 2, an odd number of 4-byte words.
__host__ __device__ constexpr int tile_row_stride( int R4, int G )
{
    return R4 * 4 * G + 2;
}

// One atom against NC register chunks of candidates: every chunk's test
// first (independent chains, one mask each), then the row entries in chunk
// order.  A lane that does not accept writes its own slot of a trash area
// instead, so the stores need no branch.  Returns the atom's new row length.
template <int NC, int G, bool HALF>
__device__ __forceinline__ unsigned build_row( const double ( &px )[BCH], const double ( &py )[BCH],
                                               const double ( &pz )[BCH], const int ( &po )[BCH],
                                               const double4 pa, int vs, unsigned short *row,
                                               unsigned k0, double rc2, int cap,
                                               unsigned short *trash )
{
    const int lane = threadIdx.x & 63;
    // the atom itself: candidate vs of this pass (chunk vs >> 6, lane
    // vs & 63; outside the pass when the chunk is not in 0 .. NC - 1)
    const int cself = vs >> 6;
    const unsigned long long notself = ~( 1ull << ( vs & 63 ) );
    unsigned long long m[NC];
#pragma unroll
    for ( int c = 0; c < NC; ++c )
    {
        const double dx = pa.x - px[c];
        const double dy = pa.y - py[c];
        const double dz = pa.z - pz[c];
        const double r = dx * dx + dy * dy + dz * dz;
        // ballots of the plain compares are their lane masks as they stand
        // (a ballot of a combined predicate is rematerialised first)
        m[c] = __builtin_amdgcn_ballot_w64( r <= rc2 ) & ( c == cself ? notself : ~0ull );
        if ( HALF )
        {
            const bool up = ( px[c] > pa.x ) |
                            ( ( px[c] == pa.x ) &
                              ( ( py[c] > pa.y ) | ( ( py[c] == pa.y ) & ( pz[c] > pa.z ) ) ) );
            m[c] &= __builtin_amdgcn_ballot_w64( up );
        }
    }
#pragma unroll
    for ( int c = 0; c < NC; ++c )
    {
        const unsigned k = __builtin_amdgcn_mbcnt_hi(
            (unsigned)( m[c] >> 32 ), __builtin_amdgcn_mbcnt_lo( (unsigned)m[c], k0 ) );
        const bool put =
            __builtin_amdgcn_inverse_ballot_w64( m[c] & __builtin_amdgcn_ballot_w64( k < (unsigned)cap ) );
        *( put ? row + k : trash + lane ) = (unsigned short)po[c];
        k0 += (unsigned)__popcll( m[c] );
    }
    return k0;
}

// The atoms a0 .. a0 + na - 1 of a cell (candidates vs0 + l of this pass,
// LDS rows row0 + l * RS) against NC register chunks of candidates, one
// atom at a time: its
// position and row length are wave-uniform (scalar loads and registers).
// Lane l of cnt holds atom a0 + l's row length.
template <int NC, int G, bool HALF>
__device__ __forceinline__ void build_atoms( const double ( &px )[BCH], const double ( &py )[BCH],
                                             const double ( &pz )[BCH], const int ( &po )[BCH],
                                             const double4 *__restrict__ x, int a0, int na,
                                             int vs0, unsigned short *row0, int RS, double rc2,
                                             int cap, unsigned short *trash, int &cnt )
{
    const int lane = threadIdx.x & 63;
    double4 pn = x[a0];
    for ( int l = 0; l < na; ++l )
    {
        const double4 pa = pn;
        pn = x[a0 + min( l + 1, na - 1 )];
        unsigned k0 = (unsigned)__builtin_amdgcn_readlane( cnt, l );
        k0 = build_row<NC, G, HALF>( px, py, pz, po, pa, vs0 + l, row0 + l * RS, k0,
                                     rc2, cap, trash );
        if ( lane == l )
            cnt = (int)k0;
    }
}

template <int G, bool HALF>
__global__ __launch_bounds__( BTPB ) void k_tile_build(
    TileGrid T, const int *__restrict__ ls, const int *__restrict__ gs,
    const int *__restrict__ info, const int *__restrict__ woff,
    const double4 *__restrict__ x, const double4 *__restrict__ xg, double rc2,
    int cap, int slot_cap, unsigned short *__restrict__ slab,
    int *__restrict__ counts, int *__restrict__ max_row )
{
    extern __shared__ double s_x[];
    __shared__ TileHead H;
    __shared__ int s_lcs[9][EMAX + 4], s_gcs[9][EMAX + 4];
    __shared__ int s_rs[BTPB / 64][MAXR];     // range starts (LDS slots)
    __shared__ int s_rp[BTPB / 64][MAXR + 1]; // prefix of range lengths
    constexpr int APW = 64 / G;
    const int t = xcd_tile( blockIdx.x, gridDim.x );
    load_head( info, t, H );
    int a, b, e0, e1;
    tile_coords( T, t, a, b, e0, e1 );
    const int ea = H.info[TI_EA], nst = H.info[TI_NST];
    for ( int k = threadIdx.x; k < 9 * ( nst + 1 ); k += blockDim.x )
    {
        const int q = k / ( nst + 1 ), ce = k % ( nst + 1 );
        const int c = col_base( T, a + q / 3 - 1, b + q % 3 - 1 ) + ea + ce;
        s_lcs[q][ce] = H.segoff[2 * q] + ( ls[c] - H.info[TI_LB + q] );
        s_gcs[q][ce] = H.segoff[2 * q + 1] + ( gs[c] - H.info[TI_GB + q] );
    }
    stage_planes<false>( H, x, xg, s_x, slot_cap );
    const double *s_y = s_x + slot_cap, *s_z = s_x + 2 * slot_cap;
    // after the planes (rounded to 16 bytes): a 64-entry trash area, then
    // the rows
    unsigned short *trash = reinterpret_cast<unsigned short *>(
        reinterpret_cast<char *>( s_x ) + tile_planes_bytes( slot_cap ) );
    unsigned short *chunk = trash + 64;
    const int R4 = ( ( cap + G - 1 ) / G + 3 ) / 4;
    const int wsz = R4 * 256; // slab entries of one slab wave
    // LDS rows: one per atom of the tile's slab waves, entries in order;
    // the stride is an odd number of words, which spreads the copy-out's
    // strided reads over the banks
    const int RS = tile_row_stride( R4, G );
    const int nw = woff[t + 1] - woff[t];
    const int tl = H.info[TI_TL];
    const int si0 = H.segoff[8] - H.info[TI_LB + 4]; // LDS slot of local i = si0 + i
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int nwv = blockDim.x >> 6;
    const int sent_slot = H.segoff[18];
    // rows start as the sentinel record (the tile's slot TOTAL)
    {
        const unsigned sent = (unsigned)( sent_slot * TILE_REC );
        const unsigned sent2 = sent | ( sent << 16 );
        const uint4 fill = make_uint4( sent2, sent2, sent2, sent2 );
        const int n16 = nw * APW * RS / 8;
        for ( int k = threadIdx.x; k < n16; k += blockDim.x )
            reinterpret_cast<uint4 *>( chunk )[k] = fill;
    }
    __syncthreads();
    int mrow = 0;
    int *rs = s_rs[wv], *rp = s_rp[wv];
    for ( int e = e0 + wv; e < e1; e += nwv )
    {
        const int ce = e - ea;
        // stencil cells of this cell inside the staged range
        const int elo = max( ce - 1, 0 ), ehi = min( ce + 1, nst - 1 );
        // column q's ranges (lane q < 9), then their offsets in the list
        int nr_q = 0;
        bool gh = false;
        if ( lane < 9 )
        {
            gh = s_gcs[lane][elo] != s_gcs[lane][ehi + 1];
            nr_q = gh ? 2 * ( ehi - elo + 1 ) : 1;
        }
        int off = nr_q;
#pragma unroll
        for ( int o = 1; o < 16; o <<= 1 )
        {
            const int u = __shfl_up( off, o, 64 );
            if ( lane >= o )
                off += u;
        }
        const int nr = __builtin_amdgcn_readfirstlane( __shfl( off, 8, 64 ) );
        off -= nr_q;
        const int off0 = off;
        if ( lane < 9 )
        {
            if ( !gh )
            {
                rs[off] = s_lcs[lane][elo];
                rp[off + 1] = s_lcs[lane][ehi + 1] - s_lcs[lane][elo];
            }
            else
            {
                for ( int c2 = elo; c2 <= ehi; ++c2 )
                {
                    rs[off] = s_lcs[lane][c2];
                    rp[off + 1] = s_lcs[lane][c2 + 1] - s_lcs[lane][c2];
                    rs[off + 1] = s_gcs[lane][c2];
                    rp[off + 2] = s_gcs[lane][c2 + 1] - s_gcs[lane][c2];
                    off += 2;
                }
            }
        }
        __builtin_amdgcn_fence( __ATOMIC_SEQ_CST, "wavefront" );
        // prefix of the range lengths (nr <= MAXR < 64: one range per lane)
        int len = lane < nr ? rp[lane + 1] : 0;
        int inc = len;
#pragma unroll
        for ( int o = 1; o < 64; o <<= 1 )
        {
            const int u = __shfl_up( inc, o, 64 );
            if ( lane >= o )
                inc += u;
        }
        if ( lane < nr )
            rp[lane + 1] = inc;
        if ( lane == 0 )
            rp[0] = 0;
        const int total = __builtin_amdgcn_readfirstlane( __shfl( inc, nr - 1, 64 ) );
        __builtin_amdgcn_fence( __ATOMIC_SEQ_CST, "wavefront" );
        // candidate index of the cell's first atom: its slot lies in the
        // centre column's (lane 4's) local range of cell ce
        int vb = 0;
        if ( lane == 4 )
            vb = gh ? rp[off0 + 2 * ( ce - elo )]
                    : rp[off0] + ( s_lcs[4][ce] - s_lcs[4][elo] );
        const int vbase = __builtin_amdgcn_readlane( vb, 4 );
        // the cell's atoms, up to 64 at a time (lane l holds atom a0 + l)
        const int cc = col_base( T, a, b ) + e;
        const int c0 = __builtin_amdgcn_readfirstlane( ls[cc] );
        const int c1 = __builtin_amdgcn_readfirstlane( ls[cc + 1] );
        for ( int a0 = c0; a0 < c1; a0 += 64 )
        {
            const int na = min( 64, c1 - a0 );
            int cnt = 0; // row length of atom a0 + lane
            for ( int v0 = 0; v0 < total; v0 += BCH * 64 )
            {
                // BCH chunks of candidates in registers (lane = candidate);
                // past the end: the sentinel position, which no test accepts
                double px[BCH], py[BCH], pz[BCH];
                int po[BCH]; // record byte offset, the value a row stores
#pragma unroll
                for ( int c = 0; c < BCH; ++c )
                {
                    const int v = v0 + c * 64 + lane;
                    int sj = sent_slot;
                    if ( v0 + c * 64 < total )
                    {
                        // range holding candidate v: binary search over the
                        // prefix of range lengths
                        int lo = 0, hi = nr;
                        while ( hi - lo > 1 )
                        {
                            const int mid = ( lo + hi ) >> 1;
                            if ( rp[mid] <= v )
                                lo = mid;
                            else
                                hi = mid;
                        }
                        if ( v < total )
                            sj = rs[lo] + ( v - rp[lo] );
                    }
                    px[c] = s_x[sj];
                    py[c] = s_y[sj];
                    pz[c] = s_z[sj];
                    po[c] = sj * TILE_REC;
                }
                const int nch = min( BCH, ( total - v0 + 63 ) >> 6 );
                const int vs0 = vbase + ( a0 - c0 ) - v0;
                unsigned short *row0 = chunk + ( a0 - tl ) * RS;
                if ( nch <= 4 )
                    build_atoms<4, G, HALF>( px, py, pz, po, x, a0, na, vs0, row0, RS, rc2,
                                             cap, trash, cnt );
                else if ( nch <= 6 )
                    build_atoms<6, G, HALF>( px, py, pz, po, x, a0, na, vs0, row0, RS, rc2,
                                             cap, trash, cnt );
                else if ( nch <= 8 )
                    build_atoms<8, G, HALF>( px, py, pz, po, x, a0, na, vs0, row0, RS, rc2,
                                             cap, trash, cnt );
                else if ( nch == 9 )
                    build_atoms<9, G, HALF>( px, py, pz, po, x, a0, na, vs0, row0, RS, rc2,
                                             cap, trash, cnt );
                else
                    build_atoms<BCH, G, HALF>( px, py, pz, po, x, a0, na, vs0, row0, RS,
                                               rc2, cap, trash, cnt );
            }
            if ( lane < na )
            {
                counts[a0 + lane] = cnt;
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
        // wave's atom L % (64 / gw).  Words of lanes past the tile's atoms
        // are never read and not written; entries at or past cap are the
        // sentinel.
        const int nwords = nw * R4 * 64, nat = H.info[TI_NAT];
        const unsigned short sent = (unsigned short)( sent_slot * TILE_REC );
        uint2 *dst = reinterpret_cast<uint2 *>( slab + (size_t)woff[t] * wsz );
        for ( int w = threadIdx.x; w < nwords; w += blockDim.x )
        {
            const int sw = w / ( R4 * 64 ), rem = w - sw * ( R4 * 64 );
            const int g4 = rem >> 6, L = rem & 63;
            const int gw = wave_lanes( G, nat, sw ), apw = 64 / gw;
            const int ta = sw * APW + L % apw;
            if ( L % apw >= min( APW, nat - sw * APW ) )
                continue;
            const unsigned short *row = chunk + ta * RS;
            unsigned e4[4];
#pragma unroll
            for ( int j = 0; j < 4; ++j )
            {
                const int e = ( 4 * g4 + j ) * gw + L / apw;
                e4[j] = e < cap ? row[e] : sent;
            }
            dst[w] = make_uint2( e4[0] | ( e4[1] << 16 ), e4[2] | ( e4[3] << 16 ) );
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
    int cap, int n_local, const unsigned short *__restrict__ slab,
    const int *__restrict__ counts, const int *__restrict__ gperm,
    const int64_t *__restrict__ offsets, int *__restrict__ idx )
{
    __shared__ TileHead H;
    constexpr int APW = 64 / G;
    const int t = blockIdx.x;
    load_head( info, t, H );
    const int R4 = ( ( cap + G - 1 ) / G + 3 ) / 4;
    const unsigned short *st = slab + (size_t)woff[t] * R4 * 256;
    const int nat = H.info[TI_NAT], tl = H.info[TI_TL];
    for ( int ta = threadIdx.x; ta < nat; ta += blockDim.x )
    {
        const int i = tl + ta;
        const int w = ta / APW, aw = ta % APW;
        const int gw = wave_lanes( G, nat, w ), apw = 64 / gw;
        const int n = counts[i];
        const int64_t o = offsets[i];
        for ( int k = 0; k < n; ++k )
        {
            const int r = k / gw, lane = ( k % gw ) * apw + aw;
            const int sj =
                st[( ( w * R4 + ( r >> 2 ) ) * 64 + lane ) * 4 + ( r & 3 )] / TILE_REC;
            int seg = 0;
            while ( sj >= H.segoff[seg + 1] )
                ++seg;
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

static cmd_status tile_info_for( cmd_neigh *N, cmd_system *S, int E, int h[4] )
{
    TileGrid &T = N->tg;
    T.g = S->bin_grid;
    T.nc0 = T.g.n[0] - 2;
    T.nc1 = T.g.n[1] - 2;
    T.E = E;
    T.tpc = ( T.g.n[2] - 2 + E - 1 ) / E;
    T.ntiles = T.nc0 * T.nc1 * T.tpc;
    CMD_REQUIRE( T.nc0 > 0 && T.nc1 > 0 && T.tpc > 0, CMD_ERR_ARG,
                 "binning grid has no interior cells" );
    CMD_TRY( N->tile_info.reserve( (size_t)T.ntiles * TI_WORDS ) );
    CMD_TRY( S->sc.ires.reserve( 4 ) );
    DevBuf<int> &waves = N->grid_sort.cursor; // scratch: waves per tile
    CMD_TRY( waves.reserve( T.ntiles + 1 ) );
    CMD_TRY( N->grid_sort.start.reserve( T.ntiles + 1 ) ); // scan of waves
    CMD_TRY( fill_i32( S->sc.ires.p, 0, 3, S->stream ) );
    const int apw = 64 / N->lanes;
    k_tile_info<<<( T.ntiles + TPB - 1 ) / TPB, TPB, 0, S->stream>>>(
        T, S->bins.start.p, S->gsort.start.p, apw, N->tile_info.p, waves.p,
        S->sc.ires.p );
    CMD_LAUNCH_CHECK();
    CMD_TRY( scan_exclusive( waves.p, N->grid_sort.start.p, T.ntiles, S->sc,
                             S->stream, nullptr ) );
    // h[0..2] from the info kernel, h[3] = slab waves of all tiles (the
    // scan's total), in one wait
    CMD_HIP( hipMemcpyAsync( h, S->sc.ires.p, 3 * sizeof( int ),
                             hipMemcpyDeviceToHost, S->stream ) );
    CMD_HIP( hipMemcpyAsync( h + 3, N->grid_sort.start.p + T.ntiles, sizeof( int ),
                             hipMemcpyDeviceToHost, S->stream ) );
    CMD_HIP( hipStreamSynchronize( S->stream ) );
    return CMD_OK;
}

static size_t tile_slab_entries( const cmd_neigh *N, int nw )
{
    const int G = N->lanes;
    const size_t R4 = (size_t)( ( ( N->cap + G - 1 ) / G + 3 ) / 4 );
    return std::max<size_t>( (size_t)nw * R4 * 256, 1 );
}

// LDS limits of one list-build block: the hard per-block limit, and the
// size that leaves room for four blocks (32 waves) per CU, which the row
// chunk is sized to when the staged positions allow it.
static void build_lds_limits( size_t *hard, size_t *target )
{
    int dev = 0, lim = 0;
    if ( hipGetDevice( &dev ) != hipSuccess ||
         hipDeviceGetAttribute( &lim, hipDeviceAttributeMaxSharedMemoryPerBlock,
                                dev ) != hipSuccess ||
         lim <= 0 )
        lim = 65536;
    const size_t statics = 4096; // tile header and cell tables
    *hard = (size_t)lim > statics ? (size_t)lim - statics : 0;
    *target = (size_t)lim >= 2 * 65536 ? (size_t)lim / 4 - statics : *hard / 2;
}

template <int G>
static void launch_build( cmd_neigh *N, cmd_system *S, int slot_cap, size_t lds,
                          int *max_row )
{
    const double rc2 = N->neigh_cut * N->neigh_cut;
    // one wave per cell of the tile
    const int threads = 64 * std::min( std::max( N->tg.E, 1 ), BTPB / 64 );
    if ( N->half )
        k_tile_build<G, true><<<N->tg.ntiles, threads, lds, S->stream>>>(
            N->tg, S->bins.start.p, S->gsort.start.p, N->tile_info.p,
            N->grid_sort.start.p, S->x.p, S->xg.p, rc2, N->cap, slot_cap,
            N->tslab.p, N->counts.p, max_row );
    else
        k_tile_build<G, false><<<N->tg.ntiles, threads, lds, S->stream>>>(
            N->tg, S->bins.start.p, S->gsort.start.p, N->tile_info.p,
            N->grid_sort.start.p, S->x.p, S->xg.p, rc2, N->cap, slot_cap,
            N->tslab.p, N->counts.p, max_row );
}

cmd_status tiles_build( cmd_neigh *N, cmd_system *S, bool *built )
{
    *built = false;
    if ( !S->bin_grid_valid || !S->xg_valid || N->lanes < 2 )
        return CMD_OK;
    const int G = N->lanes;
    if ( N->cap < 1 )
        N->cap = std::max( 1, N->max_neigh_guess );
    size_t hard = 0, target = 0;
    build_lds_limits( &hard, &target );
    (void)target;
    // tile height: the requested one, or the tallest whose neighbourhood and
    // rows fit in LDS
    int E = N->tile_e > 0 ? N->tile_e : 4;
    int h[4] = { 0, 0, 0, 0 };
    for ( int attempt = 0; attempt < 4; ++attempt )
    {
        size_t lds = 0;
        for ( ;; )
        {
            CMD_TRY( tile_info_for( N, S, std::min( E, EMAX - 2 ), h ) );
            // local atoms in the grid's padding cells (not wrapped into the
            // box) belong to no tile: use the global layout
            if ( h[1] != S->n_local )
                return CMD_OK;
            const int R4 = ( ( N->cap + G - 1 ) / G + 3 ) / 4;
            const size_t wbytes =
                (size_t)( 64 / G ) * tile_row_stride( R4, G ) * sizeof( unsigned short );
            // coordinate planes of the staged atoms and the sentinel slot,
            // the trash area of rejected entries, then the rows of every
            // slab wave of the tile
            lds = tile_planes_bytes( h[0] + 1 ) + 64 * sizeof( unsigned short ) +
                  (size_t)std::max( h[2], 1 ) * wbytes;
            // the force kernels' records on the same tiles (lj.hip): a
            // position record and, with several types, a type per slot;
            // the half list adds a reaction record and an energy weight
            const size_t slot_bytes = TILE_REC + sizeof( int );
            const size_t force_lds =
                (size_t)( h[0] + 1 ) * slot_bytes * ( N->half ? 2 : 1 ) + 8;
            if ( h[0] <= MAX_SLOTS && lds <= hard && force_lds <= hard )
                break;
            if ( E == 1 )
                return CMD_OK; // too dense for LDS tiles: global path
            E = std::max( 1, E * 2 / 3 );
        }
        N->max_tile_slots = h[0];
        CMD_TRY( N->tslab.reserve( tile_slab_entries( N, h[3] ) ) );
        CMD_TRY( fill_i32( S->sc.ires.p, 0, 1, S->stream ) );
        const int slot_cap = h[0] + 1;
        switch ( G )
        {
        case 2:
            launch_build<2>( N, S, slot_cap, lds, S->sc.ires.p );
            break;
        case 4:
            launch_build<4>( N, S, slot_cap, lds, S->sc.ires.p );
            break;
        default:
            launch_build<8>( N, S, slot_cap, lds, S->sc.ires.p );
            break;
        }
        CMD_LAUNCH_CHECK();
        int mr = 0;
        CMD_HIP( hipMemcpyAsync( &mr, S->sc.ires.p, sizeof( int ),
                                 hipMemcpyDeviceToHost, S->stream ) );
        CMD_HIP( hipStreamSynchronize( S->stream ) );
        N->max_row = mr;
        if ( mr > N->max_neigh_guess )
            N->max_neigh_guess = (int)( mr * 1.1 );
        if ( mr <= N->cap )
        {
            *built = true;
            N->tiled = true;
            N->gperm_view = S->gsort.perm.p;
            return CMD_OK;
        }
        // rows overflowed: larger rows, and the LDS fit rechecked
        N->cap = std::max( N->max_neigh_guess, mr );
    }
    set_error( "neighbour rows kept growing during the tiled build" );
    return CMD_ERR_CAPACITY;
}

cmd_status tiles_csr( cmd_neigh *N, int64_t *nnz_out )
{
    // offsets already scanned by the caller into N->offsets
    const int nt = N->tg.ntiles;
    switch ( N->lanes )
    {
    case 2:
        k_tile_to_csr<2><<<nt, TPB, 0, N->stream>>>(
            N->tg, N->tile_info.p, N->grid_sort.start.p, N->cap, N->n_rows,
            N->tslab.p, N->counts.p, N->gperm_view, N->offsets.p, N->idx.p );
        break;
    case 4:
        k_tile_to_csr<4><<<nt, TPB, 0, N->stream>>>(
            N->tg, N->tile_info.p, N->grid_sort.start.p, N->cap, N->n_rows,
            N->tslab.p, N->counts.p, N->gperm_view, N->offsets.p, N->idx.p );
        break;
    default:
        k_tile_to_csr<8><<<nt, TPB, 0, N->stream>>>(
            N->tg, N->tile_info.p, N->grid_sort.start.p, N->cap, N->n_rows,
            N->tslab.p, N->counts.p, N->gperm_view, N->offsets.p, N->idx.p );
        break;
    }
    CMD_LAUNCH_CHECK();
    (void)nnz_out;
    return CMD_OK;
}

} // namespace cmd
