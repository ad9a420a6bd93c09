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
// Slab of a tile: waves of APW = 64/G atoms; neighbour k of atom a of wave w
// goes to row r = k / G, lane (k % G) * APW + a, and four consecutive rows of
// a lane are packed in 8 bytes:
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
// same 27-cell stencil, so the wave walks the stencil's candidates once, 64
// at a time (lane = candidate, positions from LDS planes), and tests each
// chunk against every atom of the cell in turn; ballot + mbcnt place each
// accepted candidate in the atom's row, so rows come out in walk order.  The
// stencil is a list of contiguous LDS ranges (per column: the locals of its
// three cells, or locals and ghosts cell by cell where ghosts are present),
// concatenated so that chunks run across range ends.  Rows are assembled in
// LDS in the slab layout, pre-filled with the sentinel slot, and written out
// with 16-byte stores.
static constexpr int MAXR = 9 * 3 * 2; // stencil ranges of one cell

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
    unsigned short *chunk = reinterpret_cast<unsigned short *>( s_x + 3 * slot_cap );
    const int R4 = ( ( cap + G - 1 ) / G + 3 ) / 4;
    const int wsz = R4 * 256; // slab entries of one slab wave
    const int nw = woff[t + 1] - woff[t];
    const int tl = H.info[TI_TL];
    const int si0 = H.segoff[8] - H.info[TI_LB + 4]; // LDS slot of local i = si0 + i
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int nwv = blockDim.x >> 6;
    // rows start as the sentinel record (the tile's slot TOTAL)
    {
        const unsigned sent = (unsigned)( H.segoff[18] * TILE_REC );
        const unsigned sent2 = sent | ( sent << 16 );
        const uint4 fill = make_uint4( sent2, sent2, sent2, sent2 );
        const int n16 = nw * wsz / 8;
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
        // the cell's atoms, up to 64 at a time (lane l holds atom a0 + l)
        const int cc = col_base( T, a, b ) + e;
        const int c0 = __builtin_amdgcn_readfirstlane( ls[cc] );
        const int c1 = __builtin_amdgcn_readfirstlane( ls[cc + 1] );
        for ( int a0 = c0; a0 < c1; a0 += 64 )
        {
            const int na = min( 64, c1 - a0 );
            int cnt = 0; // row length of atom a0 + lane
            for ( int v0 = 0; v0 < total; v0 += 64 )
            {
                const int v = v0 + lane;
                const bool valid = v < total;
                // range holding candidate v: binary search over the prefix
                int lo = 0, hi = nr;
                while ( hi - lo > 1 )
                {
                    const int mid = ( lo + hi ) >> 1;
                    if ( rp[mid] <= v )
                        lo = mid;
                    else
                        hi = mid;
                }
                const int sj = valid ? rs[lo] + ( v - rp[lo] ) : si0 + a0;
                const double pjx = s_x[sj], pjy = s_y[sj], pjz = s_z[sj];
                // two atoms per pass: independent distance chains in flight
                for ( int l = 0; l < na; l += 2 )
                {
                    const bool two = l + 1 < na;
                    const int l2 = two ? l + 1 : l;
                    const int s1 = si0 + a0 + l, s2 = si0 + a0 + l2;
                    const double ax = s_x[s1], ay = s_y[s1], az = s_z[s1];
                    const double bx = s_x[s2], by = s_y[s2], bz = s_z[s2];
                    const double dx1 = ax - pjx, dy1 = ay - pjy, dz1 = az - pjz;
                    const double dx2 = bx - pjx, dy2 = by - pjy, dz2 = bz - pjz;
                    const double r1 = dx1 * dx1 + dy1 * dy1 + dz1 * dz1;
                    const double r2 = dx2 * dx2 + dy2 * dy2 + dz2 * dz2;
                    bool acc1 = valid & ( sj != s1 ) & ( r1 <= rc2 );
                    bool acc2 = two & valid & ( sj != s2 ) & ( r2 <= rc2 );
                    if ( HALF )
                    {
                        acc1 = acc1 & ( ( pjx > ax ) |
                                        ( ( pjx == ax ) &
                                          ( ( pjy > ay ) | ( ( pjy == ay ) & ( pjz > az ) ) ) ) );
                        acc2 = acc2 & ( ( pjx > bx ) |
                                        ( ( pjx == bx ) &
                                          ( ( pjy > by ) | ( ( pjy == by ) & ( pjz > bz ) ) ) ) );
                    }
                    const unsigned long long m1 = __ballot( acc1 );
                    const unsigned long long m2 = __ballot( acc2 );
                    const int c1 = __builtin_amdgcn_readlane( cnt, l );
                    const int c2 = __builtin_amdgcn_readlane( cnt, l2 );
                    const unsigned k1 =
                        (unsigned)c1 + __builtin_amdgcn_mbcnt_hi(
                                           (unsigned)( m1 >> 32 ),
                                           __builtin_amdgcn_mbcnt_lo( (unsigned)m1, 0u ) );
                    const unsigned k2 =
                        (unsigned)c2 + __builtin_amdgcn_mbcnt_hi(
                                           (unsigned)( m2 >> 32 ),
                                           __builtin_amdgcn_mbcnt_lo( (unsigned)m2, 0u ) );
                    auto put = [&]( int ll, unsigned k )
                    {
                        const unsigned ta = (unsigned)( a0 + ll - tl );
                        const unsigned r = k / G;
                        const unsigned L = ( k % G ) * APW + ta % APW;
                        chunk[( ( ( ta / APW ) * R4 + ( r >> 2 ) ) * 64 + L ) * 4 + ( r & 3 )] =
                            (unsigned short)( sj * TILE_REC );
                    };
                    if ( acc1 && k1 < (unsigned)cap )
                        put( l, k1 );
                    if ( acc2 && k2 < (unsigned)cap )
                        put( l2, k2 );
                    if ( lane == l )
                        cnt = c1 + (int)__popcll( m1 );
                    if ( two && lane == l2 )
                        cnt = c2 + (int)__popcll( m2 );
                }
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
        const int n16 = nw * wsz / 8;
        const uint4 *src = reinterpret_cast<const uint4 *>( chunk );
        uint4 *dst = reinterpret_cast<uint4 *>( slab + (size_t)woff[t] * wsz );
        for ( int k = threadIdx.x; k < n16; k += blockDim.x )
            dst[k] = src[k];
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
        const int n = counts[i];
        const int64_t o = offsets[i];
        for ( int k = 0; k < n; ++k )
        {
            const int r = k / G, lane = ( k % G ) * APW + aw;
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

cmd_status tiles_prepare_ghosts( cmd_system *S )
{
    S->xg_valid = false;
    if ( !S->bin_grid_valid )
        return CMD_OK;
    const int ng = S->n_ghost;
    CMD_TRY( cell_sort( S->x.p + S->n_local, ng, S->bin_grid, S->gsort, S->sc,
                        S->stream ) );
    CMD_TRY( S->xg.reserve( std::max( ng, 1 ) ) );
    if ( ng > 0 )
    {
        k_gather_ghosts<<<( ng + TPB - 1 ) / TPB, TPB, 0, S->stream>>>(
            S->n_local, ng, S->gsort.perm.p, S->x.p, S->xg.p );
        CMD_LAUNCH_CHECK();
    }
    S->xg_valid = true;
    return CMD_OK;
}

cmd_status tiles_ghost_refresh( cmd_system *S )
{
    if ( S->n_ghost == 0 )
        return CMD_OK;
    k_ghost_refresh_sorted<<<( S->n_ghost + TPB - 1 ) / TPB, TPB, 0, S->stream>>>(
        S->n_local, S->n_ghost, S->x.p, S->xg.p, S->gsort.perm.p, S->ghost_root.p,
        S->ghost_code.p, S->dom.L[0], S->dom.L[1], S->dom.L[2] );
    CMD_LAUNCH_CHECK();
    return CMD_OK;
}

static cmd_status tile_info_for( cmd_neigh *N, cmd_system *S, int E, int h[3] )
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
    CMD_HIP( hipMemcpyAsync( h, S->sc.ires.p, 3 * sizeof( int ),
                             hipMemcpyDeviceToHost, S->stream ) );
    CMD_HIP( hipStreamSynchronize( S->stream ) );
    return CMD_OK;
}

static size_t tile_slab_entries( cmd_neigh *N, hipStream_t st )
{
    int nw = 0;
    (void)hipMemcpyAsync( &nw, N->grid_sort.start.p + N->tg.ntiles, sizeof( int ),
                          hipMemcpyDeviceToHost, st );
    (void)hipStreamSynchronize( st );
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
    int h[3] = { 0, 0, 0 };
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
            const size_t wbytes =
                (size_t)( ( ( N->cap + G - 1 ) / G + 3 ) / 4 ) * 256 * sizeof( unsigned short );
            // coordinate planes of the staged atoms and the sentinel slot,
            // then the rows of every slab wave of the tile
            lds = (size_t)( h[0] + 1 ) * 3 * sizeof( double ) +
                  (size_t)std::max( h[2], 1 ) * wbytes;
            if ( h[0] <= MAX_SLOTS && lds <= hard )
                break;
            if ( E == 1 )
                return CMD_OK; // too dense for LDS tiles: global path
            E = std::max( 1, E * 2 / 3 );
        }
        N->max_tile_slots = h[0];
        CMD_TRY( N->tslab.reserve( tile_slab_entries( N, S->stream ) ) );
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
