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
// list stores, for each neighbour, its 16-bit LDS slot.  Rows are walked in
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
static constexpr int EMAX = 16;
static constexpr int MAX_SLOTS = 1792; // 56 KB of double4 per block
// Per-tile runs, atom range and staged count; waves per tile for the slab.
// slots[0] gets the largest staged count, slots[1] the number of local atoms
// the tiles cover (all of them only when every local atom sits in an
// interior cell of the grid).
__global__ void k_tile_info( TileGrid T, const int *__restrict__ ls,
                             const int *__restrict__ gs, int apw,
                             int *__restrict__ info, int *__restrict__ waves,
                             int *__restrict__ slots )
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    int total = 0, nat = 0;
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
        waves[t] = ( nat + apw - 1 ) / apw;
    }
    int m = total;
#pragma unroll
    for ( int o = 32; o > 0; o >>= 1 )
    {
        m = max( m, __shfl_xor( m, o, 64 ) );
        nat += __shfl_xor( nat, o, 64 );
    }
    if ( ( threadIdx.x & 63 ) == 0 )
    {
        if ( m > 0 )
            atomicMax( &slots[0], m );
        if ( nat > 0 )
            atomicAdd( &slots[1], nat );
    }
}

__device__ __forceinline__ bool tile_accept( const double4 &pi, const double4 &pj,
                                             double rc2, int half )
{
    if ( half &&
         !( ( pj.x > pi.x ) ||
            ( ( pj.x == pi.x ) &&
              ( ( pj.y > pi.y ) || ( ( pj.y == pi.y ) && ( pj.z > pi.z ) ) ) ) ) )
        return false;
    const double dx = pi.x - pj.x;
    const double dy = pi.y - pj.y;
    const double dz = pi.z - pj.z;
    const double rsq = dx * dx + dy * dy + dz * dz;
    return rsq <= rc2;
}

// List build: one lane per atom walks its stencil in LDS.
template <int G>
__global__ __launch_bounds__( TPB ) void k_tile_build(
    TileGrid T, const int *__restrict__ ls, const int *__restrict__ gs,
    const int *__restrict__ info, const int *__restrict__ woff,
    const double4 *__restrict__ x, const double4 *__restrict__ xg, double rc2,
    int half, int cap, unsigned short *__restrict__ slab,
    int *__restrict__ counts, int *__restrict__ max_row )
{
    extern __shared__ double4 s_pos[];
    __shared__ TileHead H;
    __shared__ int s_lcs[9][EMAX + 4], s_gcs[9][EMAX + 4];
    constexpr int APW = 64 / G;
    const int t = blockIdx.x;
    load_head( info, t, H );
    int a, b, e0, e1;
    tile_coords( T, t, a, b, e0, e1 );
    const int ea = H.info[TI_EA], nst = H.info[TI_NST];
    // LDS slot of the first atom of each staged cell, locals and ghosts
    for ( int k = threadIdx.x; k < 9 * ( nst + 1 ); k += blockDim.x )
    {
        const int q = k / ( nst + 1 ), ce = k % ( nst + 1 );
        const int c = col_base( T, a + q / 3 - 1, b + q % 3 - 1 ) + ea + ce;
        s_lcs[q][ce] = H.segoff[2 * q] + ( ls[c] - H.info[TI_LB + q] );
        s_gcs[q][ce] = H.segoff[2 * q + 1] + ( gs[c] - H.info[TI_GB + q] );
    }
    stage_positions( H, x, xg, s_pos );
    const int R4 = ( ( cap + G - 1 ) / G + 3 ) / 4;
    unsigned short *st = slab + (size_t)woff[t] * R4 * 256;
    const int nat = H.info[TI_NAT], tl = H.info[TI_TL];
    int mrow = 0;
    for ( int ta = threadIdx.x; ta < nat; ta += blockDim.x )
    {
        const int i = tl + ta;
        const int si = H.segoff[8] + ( i - H.info[TI_LB + 4] );
        const double4 pi = s_pos[si];
        const int ei = cell_coord( pi.z, T.g.min[2], T.g.rd[2], T.g.n[2],
                                   T.g.floor_mode );
        // stencil cells of atom i inside the staged range (always true when
        // the atoms have not moved since create_binning; clamped regardless)
        const int elo = max( max( ei - 1, 0 ) - ea, 0 );
        const int ehi = min( min( ei + 1, T.g.n[2] - 1 ) - ea, nst - 1 );
        const int w = ta / APW, aw = ta % APW;
        int cnt = 0;
        auto take = [&]( int sj )
        {
            if ( sj == si )
                return;
            if ( tile_accept( pi, s_pos[sj], rc2, half ) )
            {
                if ( cnt < cap )
                {
                    const int r = cnt / G, lane = ( cnt % G ) * APW + aw;
                    st[( ( w * R4 + ( r >> 2 ) ) * 64 + lane ) * 4 + ( r & 3 )] =
                        (unsigned short)sj;
                }
                ++cnt;
            }
        };
        for ( int q = 0; q < 9; ++q )
            for ( int ce = elo; ce <= ehi; ++ce )
            {
                for ( int sj = s_lcs[q][ce]; sj < s_lcs[q][ce + 1]; ++sj )
                    take( sj );
                for ( int sj = s_gcs[q][ce]; sj < s_gcs[q][ce + 1]; ++sj )
                    take( sj );
            }
        counts[i] = cnt;
        mrow = max( mrow, cnt );
    }
#pragma unroll
    for ( int o = 32; o > 0; o >>= 1 )
        mrow = max( mrow, __shfl_xor( mrow, o, 64 ) );
    if ( ( threadIdx.x & 63 ) == 0 && mrow > 0 )
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
            const int sj = st[( ( w * R4 + ( r >> 2 ) ) * 64 + lane ) * 4 + ( r & 3 )];
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

static cmd_status tile_info_for( cmd_neigh *N, cmd_system *S, int E,
                                 int *max_total, int *covered )
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
    CMD_TRY( fill_i32( S->sc.ires.p, 0, 2, S->stream ) );
    const int apw = 64 / N->lanes;
    k_tile_info<<<( T.ntiles + TPB - 1 ) / TPB, TPB, 0, S->stream>>>(
        T, S->bins.start.p, S->gsort.start.p, apw, N->tile_info.p, waves.p,
        S->sc.ires.p );
    CMD_LAUNCH_CHECK();
    CMD_TRY( scan_exclusive( waves.p, N->grid_sort.start.p, T.ntiles, S->sc,
                             S->stream, nullptr ) );
    int h[2] = { 0, 0 };
    CMD_HIP( hipMemcpyAsync( h, S->sc.ires.p, sizeof( h ),
                             hipMemcpyDeviceToHost, S->stream ) );
    CMD_HIP( hipStreamSynchronize( S->stream ) );
    *max_total = h[0];
    *covered = h[1];
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

template <int G>
static void launch_build( cmd_neigh *N, cmd_system *S, size_t lds, int *max_row )
{
    const double rc2 = N->neigh_cut * N->neigh_cut;
    k_tile_build<G><<<N->tg.ntiles, TPB, lds, S->stream>>>(
        N->tg, S->bins.start.p, S->gsort.start.p, N->tile_info.p,
        N->grid_sort.start.p, S->x.p, S->xg.p, rc2, N->half, N->cap, N->tslab.p,
        N->counts.p, max_row );
}

cmd_status tiles_build( cmd_neigh *N, cmd_system *S, bool *built )
{
    *built = false;
    if ( !S->bin_grid_valid || !S->xg_valid || N->lanes < 2 )
        return CMD_OK;
    // tile height: the requested one, or the tallest whose neighbourhood
    // fits the LDS budget
    int E = N->tile_e > 0 ? N->tile_e : 6;
    int max_total = 0, covered = 0;
    for ( ;; )
    {
        CMD_TRY( tile_info_for( N, S, std::min( E, EMAX - 2 ), &max_total,
                                &covered ) );
        // local atoms in the grid's padding cells (not wrapped into the box)
        // belong to no tile: use the global layout
        if ( covered != S->n_local )
            return CMD_OK;
        if ( max_total <= MAX_SLOTS )
            break;
        if ( E == 1 )
            return CMD_OK; // too dense for LDS tiles: global path
        E = std::max( 1, E * 2 / 3 );
    }
    N->max_tile_slots = max_total;
    const size_t lds = (size_t)std::max( max_total, 1 ) * sizeof( double4 );
    if ( N->cap < 1 )
        N->cap = std::max( 1, N->max_neigh_guess );
    for ( int attempt = 0; attempt < 3; ++attempt )
    {
        CMD_TRY( N->tslab.reserve( tile_slab_entries( N, S->stream ) ) );
        CMD_TRY( fill_i32( S->sc.ires.p, 0, 1, S->stream ) );
        switch ( N->lanes )
        {
        case 2:
            launch_build<2>( N, S, lds, S->sc.ires.p );
            break;
        case 4:
            launch_build<4>( N, S, lds, S->sc.ires.p );
            break;
        default:
            launch_build<8>( N, S, lds, S->sc.ires.p );
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
