// neighbor.hip -- Verlet list build.
//
// Stands in for NeighborVerlet::create (neighbor_verlet.h:43-62), which
// builds a Cabana::VerletList over rows [0,N_local) and candidates
// [0,N_local+N_ghost) with radius neigh_cut.  Accept rule (Cabana
// FullNeighborTag / HalfNeighborTag discriminators): j != i, rsq <= rc^2
// with rsq = dx*dx + dy*dy + dz*dz and d = x_i - x_j evaluated without
// contraction; half lists keep j only if x_j is lexicographically greater
// than x_i (x, then y, then z).
//
// Layout ("slab"): atoms are grouped per wavefront, APW = 64/G atoms per
// wave with G lanes per atom.  Neighbour k of atom a of wave w is stored at
//   slab[((w * R + k / G) * 64) + (k % G) * APW + a],   R = ceil(cap / G),
// so that at every step of the force loop the 64 lanes of a wave read 64
// consecutive ints.  cap is the per-atom capacity, i.e. the reference's
// max_neigh_guess of the 2D Verlet layout: a build that finds a longer row
// grows cap to 1.1x the longest row and runs again (neighbor_verlet.h:58-61).
// A CSR view (int64 offsets, int32 indices) is derived on demand.
//
// Build: bounding box of all atoms, stable counting sort of all atoms into a
// grid with cells >= rc, then one lane per local atom walks its 27-cell
// stencil (x-cell, y-cell, z-cell order; index order inside a cell) and
// appends accepted neighbours.  The row order is therefore deterministic.
#include "internal.h"

#include <algorithm>
#include <cfloat>
#include <cmath>

namespace cmd
{
static constexpr int TPB = 256;

__global__ void k_bbox( int n, const double4 *__restrict__ x,
                        double *__restrict__ part )
{
    double lo[3] = { DBL_MAX, DBL_MAX, DBL_MAX };
    double hi[3] = { -DBL_MAX, -DBL_MAX, -DBL_MAX };
    for ( int i = blockIdx.x * blockDim.x + threadIdx.x; i < n;
          i += gridDim.x * blockDim.x )
    {
        const double4 p = x[i];
        lo[0] = fmin( lo[0], p.x );
        lo[1] = fmin( lo[1], p.y );
        lo[2] = fmin( lo[2], p.z );
        hi[0] = fmax( hi[0], p.x );
        hi[1] = fmax( hi[1], p.y );
        hi[2] = fmax( hi[2], p.z );
    }
    __shared__ double red[6][TPB / 64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for ( int k = 0; k < 3; ++k )
    {
        double a = lo[k], b = hi[k];
#pragma unroll
        for ( int o = 32; o > 0; o >>= 1 )
        {
            a = fmin( a, __shfl_down( a, o, 64 ) );
            b = fmax( b, __shfl_down( b, o, 64 ) );
        }
        if ( lane == 0 )
        {
            red[k][wid] = a;
            red[3 + k][wid] = b;
        }
    }
    __syncthreads();
    if ( threadIdx.x < 6 )
    {
        const int k = threadIdx.x;
        double a = red[k][0];
        for ( int w = 1; w < TPB / 64; ++w )
            a = ( k < 3 ) ? fmin( a, red[k][w] ) : fmax( a, red[k][w] );
        part[blockIdx.x * 6 + k] = a;
    }
}

// One 64-lane wave: lane t folds partials t, t+64, ... then a butterfly.
// With `inside`, it also records whether the box lies in [glo, ghi) on
// every axis (1) or not (0), the test the tiled build's grid needs.
__global__ void k_bbox_final( int nb, const double *__restrict__ part,
                              double *__restrict__ out, double3 glo, double3 ghi,
                              int *__restrict__ inside )
{
    const int t = threadIdx.x;
    double v[6];
    for ( int k = 0; k < 6; ++k )
        v[k] = ( k < 3 ) ? DBL_MAX : -DBL_MAX;
    for ( int b = t; b < nb; b += 64 )
        for ( int k = 0; k < 6; ++k )
            v[k] = ( k < 3 ) ? fmin( v[k], part[b * 6 + k] )
                             : fmax( v[k], part[b * 6 + k] );
    for ( int k = 0; k < 6; ++k )
    {
#pragma unroll
        for ( int o = 32; o > 0; o >>= 1 )
        {
            const double u = __shfl_xor( v[k], o, 64 );
            v[k] = ( k < 3 ) ? fmin( v[k], u ) : fmax( v[k], u );
        }
    }
    if ( t == 0 )
    {
        for ( int k = 0; k < 6; ++k )
            out[k] = v[k];
        if ( inside )
            *inside = ( v[0] >= glo.x && v[1] >= glo.y && v[2] >= glo.z && v[3] < ghi.x &&
                        v[4] < ghi.y && v[5] < ghi.z )
                          ? 1
                          : 0;
    }
}

__global__ void k_gather_sorted( int n, const int *__restrict__ perm,
                                 const double4 *__restrict__ x,
                                 double4 *__restrict__ xs )
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if ( k >= n )
        return;
    const int i = perm[k];
    double4 p = x[i];
    p.w = (double)i;
    xs[k] = p;
}

__device__ __forceinline__ bool accept_pair( const double4 &pi, const double4 &pj,
                                             int i, int j, double rc2, int half )
{
    if ( i == j )
        return false;
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


// One lane per local atom i; lanes of a wave are consecutive atoms, which
// after binning sit in neighbouring cells and walk overlapping stencils.
template <int G>
__global__ __launch_bounds__( TPB ) void k_neigh_slab(
    int n_local, const double4 *__restrict__ x, const double4 *__restrict__ xs,
    const int *__restrict__ start, CellGrid g, double rc2, int half, int cap,
    int *__restrict__ counts, int *__restrict__ slab, int *__restrict__ max_row )
{
    constexpr int APW = 64 / G;
    const int i = blockIdx.x * TPB + (int)threadIdx.x;
    int cnt = 0;
    if ( i < n_local )
    {
        const double4 pi = x[i];
        const int a = cell_coord( pi.x, g.min[0], g.rd[0], g.n[0], g.floor_mode );
        const int b = cell_coord( pi.y, g.min[1], g.rd[1], g.n[1], g.floor_mode );
        const int e = cell_coord( pi.z, g.min[2], g.rd[2], g.n[2], g.floor_mode );
        const int e0 = e > 0 ? e - 1 : 0;
        const int e1 = e < g.n[2] - 1 ? e + 1 : g.n[2] - 1;
        const int R = ( cap + G - 1 ) / G;
        const int w = i / APW, ai = i % APW;
        int *out = slab + (size_t)w * R * 64 + ai;
        for ( int aa = a - 1; aa <= a + 1; ++aa )
        {
            if ( aa < 0 || aa >= g.n[0] )
                continue;
            for ( int bb = b - 1; bb <= b + 1; ++bb )
            {
                if ( bb < 0 || bb >= g.n[1] )
                    continue;
                const int base = ( aa * g.n[1] + bb ) * g.n[2];
                const int s1 = start[base + e1 + 1];
                for ( int s = start[base + e0]; s < s1; ++s )
                {
                    const double4 pj = xs[s];
                    const int j = (int)pj.w;
                    if ( accept_pair( pi, pj, i, j, rc2, half ) )
                    {
                        if ( cnt < cap )
                            out[( cnt / G ) * 64 + ( cnt % G ) * APW] = j;
                        ++cnt;
                    }
                }
            }
        }
        counts[i] = cnt;
    }
    int m = cnt;
#pragma unroll
    for ( int o = 32; o > 0; o >>= 1 )
        m = max( m, __shfl_xor( m, o, 64 ) );
    if ( ( threadIdx.x & 63 ) == 0 && m > 0 )
        atomicMax( max_row, m );
}

// CSR view: row i of the slab copied to idx[offsets[i] ...].
template <int G>
__global__ void k_slab_to_csr( int n_local, int cap, const int *__restrict__ counts,
                               const int *__restrict__ slab,
                               const int64_t *__restrict__ offsets,
                               int *__restrict__ idx )
{
    constexpr int APW = 64 / G;
    const int i = blockIdx.x * TPB + (int)threadIdx.x;
    if ( i >= n_local )
        return;
    const int R = ( cap + G - 1 ) / G;
    const int w = i / APW, ai = i % APW;
    const int *row = slab + (size_t)w * R * 64 + ai;
    const int n = counts[i];
    const int64_t o = offsets[i];
    for ( int k = 0; k < n; ++k )
        idx[o + k] = row[( k / G ) * 64 + ( k % G ) * APW];
}

// Grid for the candidate search (mirrored by the oracle): cells of at least
// rc per dimension over the bounding box, at most 1024 per dimension, and at
// most max(64, 2 n) cells in total (halving the largest dimension).
static void neigh_grid( const double lo[3], const double hi[3], double rc,
                        int n_total, int ng[3], double rcell[3] )
{
    double ext[3];
    for ( int k = 0; k < 3; ++k )
    {
        ext[k] = hi[k] - lo[k];
        ng[k] = std::max( 1, (int)std::floor( ext[k] / rc ) );
        ng[k] = std::min( ng[k], 1024 );
    }
    const int64_t limit = std::max<int64_t>( 64, 2 * (int64_t)n_total );
    while ( (int64_t)ng[0] * ng[1] * ng[2] > limit )
    {
        int m = 0;
        if ( ng[1] > ng[m] )
            m = 1;
        if ( ng[2] > ng[m] )
            m = 2;
        ng[m] = std::max( 1, ng[m] / 2 );
    }
    for ( int k = 0; k < 3; ++k )
        rcell[k] = ng[k] / std::max( ext[k], 1e-300 );
}

static cmd_status launch_slab( cmd_neigh *N, cmd_system *S, const CellGrid &g,
                               int n_local, int cap, int *max_row_dev )
{
    const double rc2 = N->neigh_cut * N->neigh_cut;
    const int nb = std::max( 1, ( n_local + TPB - 1 ) / TPB );
#define CMD_SLAB( GG )                                                        \
    k_neigh_slab<GG><<<nb, TPB, 0, S->stream>>>(                              \
        n_local, S->x.p, N->xs.p, N->grid_sort.start.p, g, rc2, N->half, cap, \
        N->counts.p, N->slab.p, max_row_dev )
    switch ( N->lanes )
    {
    case 1:
        CMD_SLAB( 1 );
        break;
    case 2:
        CMD_SLAB( 2 );
        break;
    case 4:
        CMD_SLAB( 4 );
        break;
    case 8:
        CMD_SLAB( 8 );
        break;
    default:
        set_error( "unsupported neighbour layout lanes %d", N->lanes );
        return CMD_ERR_ARG;
    }
#undef CMD_SLAB
    CMD_LAUNCH_CHECK();
    return CMD_OK;
}

size_t slab_entries( int n_local, int cap, int lanes )
{
    const int apw = 64 / lanes;
    const size_t waves = (size_t)( n_local + apw - 1 ) / apw;
    const size_t R = (size_t)( cap + lanes - 1 ) / lanes;
    return std::max<size_t>( waves * R * 64, 1 );
}

cmd_status neigh_build( cmd_neigh *N, cmd_system *S )
{
    hipStream_t st = S->stream;
    N->stream = st;
    const int n_local = S->n_local;
    const int nt = S->n_local + S->n_ghost;
    N->n_rows = n_local;
    N->csr_valid = false;
    N->tiled = false;
    CMD_TRY( N->counts.reserve( std::max( n_local, 1 ) ) );
    if ( nt == 0 || n_local == 0 )
    {
        N->nnz = 0;
        N->max_row = 0;
        return CMD_OK;
    }
    const int nbb = std::min( 1024, std::max( 1, ( nt + TPB - 1 ) / TPB ) );
    CMD_TRY( S->sc.dpart.reserve( 6 * (size_t)nbb ) );
    CMD_TRY( S->sc.dres.reserve( 8 ) );
    // (0) the tiled layout of the last build, rebuilt without host waits
    // before the build's own (tiles_build_spec): the bounding-box test runs
    // on the device and is read back with the build's results; anything
    // that does not hold sends the build down the path below
    if ( N->allow_tiles && N->lanes >= 2 && S->bin_grid_valid && N->spec.valid )
    {
        const CellGrid &bg = S->bin_grid;
        bool cells_ok = true;
        double3 glo, ghi;
        double lo[3], hi[3];
        for ( int k = 0; k < 3; ++k )
        {
            lo[k] = bg.min[k];
            hi[k] = bg.min[k] + bg.n[k] / bg.rd[k];
            if ( !( 1.0 / bg.rd[k] >= N->neigh_cut ) )
                cells_ok = false;
        }
        if ( cells_ok )
        {
            glo = make_double3( lo[0], lo[1], lo[2] );
            ghi = make_double3( hi[0], hi[1], hi[2] );
            CMD_TRY( S->sc.ires.reserve( TILE_RES_WORDS ) );
            k_bbox<<<nbb, TPB, 0, st>>>( nt, S->x.p, S->sc.dpart.p );
            CMD_LAUNCH_CHECK();
            k_bbox_final<<<1, 64, 0, st>>>( nbb, S->sc.dpart.p, S->sc.dres.p, glo, ghi,
                                            S->sc.ires.p + TILE_RES_INSIDE );
            CMD_LAUNCH_CHECK();
            if ( !S->xg_valid )
                CMD_TRY( tiles_prepare_ghosts( S ) );
            bool built = false;
            CMD_TRY( tiles_build_spec( N, S, &built ) );
            if ( built )
            {
                N->used_bin_grid = true;
                N->nnz = -1;
                return CMD_OK;
            }
        }
    }
    // (1) bounding box
    k_bbox<<<nbb, TPB, 0, st>>>( nt, S->x.p, S->sc.dpart.p );
    CMD_LAUNCH_CHECK();
    k_bbox_final<<<1, 64, 0, st>>>( nbb, S->sc.dpart.p, S->sc.dres.p, double3{}, double3{},
                                    nullptr );
    CMD_LAUNCH_CHECK();
    double bb[6];
    CMD_HIP( hipMemcpyAsync( bb, S->sc.dres.p, sizeof( bb ),
                             hipMemcpyDeviceToHost, st ) );
    CMD_HIP( hipStreamSynchronize( st ) );
    // (2) grid and counting sort of all atoms.  The grid of the last sorting
    // create_binning is used when its cells are at least rc wide and it holds
    // every atom (local and ghost): rows then walk cells in the order the
    // local atoms are stored.  Otherwise a grid over the bounding box.
    CellGrid g;
    bool use_bin = S->bin_grid_valid;
    if ( use_bin )
    {
        const CellGrid &bg = S->bin_grid;
        for ( int k = 0; k < 3; ++k )
        {
            const double lo = bg.min[k], hi = bg.min[k] + bg.n[k] / bg.rd[k];
            if ( !( 1.0 / bg.rd[k] >= N->neigh_cut ) || !( bb[k] >= lo ) ||
                 !( bb[3 + k] < hi ) )
                use_bin = false;
        }
    }
    if ( use_bin && N->allow_tiles && N->lanes >= 2 )
    {
        if ( !S->xg_valid )
            CMD_TRY( tiles_prepare_ghosts( S ) );
        bool built = false;
        CMD_TRY( tiles_build( N, S, &built ) );
        if ( built )
        {
            N->used_bin_grid = true;
            N->nnz = -1;
            return CMD_OK;
        }
    }
    N->tiled = false;
    if ( use_bin )
        g = S->bin_grid;
    else
    {
        int ng[3];
        double rcell[3];
        neigh_grid( bb, bb + 3, N->neigh_cut, nt, ng, rcell );
        g.floor_mode = 0;
        for ( int k = 0; k < 3; ++k )
        {
            g.min[k] = bb[k];
            g.rd[k] = rcell[k];
            g.n[k] = ng[k];
        }
    }
    for ( int k = 0; k < 3; ++k )
        N->ng[k] = g.n[k];
    N->used_bin_grid = use_bin;
    CMD_TRY( cell_sort( S->x.p, nt, g, N->grid_sort, S->sc, st ) );
    CMD_TRY( N->xs.reserve( nt ) );
    k_gather_sorted<<<( nt + TPB - 1 ) / TPB, TPB, 0, st>>>(
        nt, N->grid_sort.perm.p, S->x.p, N->xs.p );
    CMD_LAUNCH_CHECK();
    // (3) single pass into the slab at the current capacity; regrow and
    // repeat if a row was longer
    CMD_TRY( S->sc.ires.reserve( 4 ) );
    if ( N->cap < 1 )
        N->cap = std::max( 1, N->max_neigh_guess );
    for ( int attempt = 0; attempt < 3; ++attempt )
    {
        CMD_TRY( N->slab.reserve( slab_entries( n_local, N->cap, N->lanes ) ) );
        CMD_TRY( fill_i32( S->sc.ires.p, 0, 1, st ) );
        CMD_TRY( launch_slab( N, S, g, n_local, N->cap, S->sc.ires.p ) );
        int mr = 0;
        CMD_HIP( hipMemcpyAsync( &mr, S->sc.ires.p, sizeof( int ),
                                 hipMemcpyDeviceToHost, st ) );
        CMD_HIP( hipStreamSynchronize( st ) );
        N->max_row = mr;
        // NeighborVerlet::create: guess = 1.1 * max when max > guess
        if ( mr > N->max_neigh_guess )
            N->max_neigh_guess = (int)( mr * 1.1 );
        if ( mr <= N->cap )
            break;
        N->cap = std::max( N->max_neigh_guess, mr );
    }
    CMD_REQUIRE( N->max_row <= N->cap, CMD_ERR_CAPACITY,
                 "neighbour rows kept growing during the build" );
    N->nnz = -1; // known once the CSR view is made
    return CMD_OK;
}

cmd_status neigh_csr( cmd_neigh *N )
{
    if ( N->csr_valid )
        return CMD_OK;
    const int n = N->n_rows;
    hipStream_t st = N->stream;
    CMD_TRY( N->offsets.reserve( (size_t)n + 1 ) );
    if ( n == 0 )
    {
        CMD_TRY( fill_i32( (int *)N->offsets.p, 0, 2, st ) );
        N->nnz = 0;
        N->csr_valid = true;
        return CMD_OK;
    }
    Scratch sc;
    int64_t nnz = 0;
    CMD_TRY( scan_exclusive( N->counts.p, N->offsets.p, n, sc, st, &nnz ) );
    CMD_TRY( N->idx.reserve( (size_t)std::max<int64_t>( nnz, 1 ) ) );
    const int nb = ( n + TPB - 1 ) / TPB;
    if ( N->tiled )
    {
        CMD_TRY( tiles_csr( N, nullptr ) );
        CMD_HIP( hipStreamSynchronize( st ) );
        N->nnz = nnz;
        N->csr_valid = true;
        return CMD_OK;
    }
    switch ( N->lanes )
    {
    case 1:
        k_slab_to_csr<1><<<nb, TPB, 0, st>>>( n, N->cap, N->counts.p, N->slab.p,
                                              N->offsets.p, N->idx.p );
        break;
    case 2:
        k_slab_to_csr<2><<<nb, TPB, 0, st>>>( n, N->cap, N->counts.p, N->slab.p,
                                              N->offsets.p, N->idx.p );
        break;
    case 8:
        k_slab_to_csr<8><<<nb, TPB, 0, st>>>( n, N->cap, N->counts.p, N->slab.p,
                                              N->offsets.p, N->idx.p );
        break;
    default:
        k_slab_to_csr<4><<<nb, TPB, 0, st>>>( n, N->cap, N->counts.p, N->slab.p,
                                              N->offsets.p, N->idx.p );
        break;
    }
    CMD_LAUNCH_CHECK();
    CMD_HIP( hipStreamSynchronize( st ) );
    N->nnz = nnz;
    N->csr_valid = true;
    return CMD_OK;
}

} // namespace cmd

using namespace cmd;

extern "C" {

cmd_status cmd_neigh_create( double neigh_cut, int half, int max_neigh_guess,
                             cmd_neigh **out )
{
    CMD_REQUIRE( out, CMD_ERR_ARG, "out is NULL" );
    CMD_REQUIRE( neigh_cut > 0, CMD_ERR_ARG, "neigh_cut must be positive" );
    cmd_neigh *n = new cmd_neigh;
    n->neigh_cut = neigh_cut;
    n->half = half ? 1 : 0;
    n->max_neigh_guess = std::max( 1, max_neigh_guess );
    *out = n;
    return CMD_OK;
}

cmd_status cmd_neigh_destroy( cmd_neigh *n )
{
    delete n;
    return CMD_OK;
}

cmd_status cmd_neigh_set_tile( cmd_neigh *n, int height )
{
    CMD_REQUIRE( n, CMD_ERR_ARG, "neighbor list is NULL" );
    CMD_REQUIRE( height >= 0 && height <= 14, CMD_ERR_ARG,
                 "tile height must be 0 (choose) or 1..14 cells" );
    if ( height != n->tile_e )
    {
        n->tile_e = height;
        n->n_rows = 0; // layout changed: must rebuild
        n->csr_valid = false;
    }
    return CMD_OK;
}

cmd_status cmd_neigh_set_lanes( cmd_neigh *n, int lanes )
{
    CMD_REQUIRE( n, CMD_ERR_ARG, "neighbor list is NULL" );
    CMD_REQUIRE( lanes == 1 || lanes == 2 || lanes == 4 || lanes == 8,
                 CMD_ERR_ARG, "lanes per atom must be 1, 2, 4 or 8" );
    if ( lanes != n->lanes )
    {
        n->lanes = lanes;
        n->n_rows = 0; // layout changed: must rebuild
        n->csr_valid = false;
    }
    return CMD_OK;
}

cmd_status cmd_neigh_build( cmd_neigh *n, cmd_system *s )
{
    CMD_REQUIRE( n && s, CMD_ERR_ARG, "NULL argument" );
    return neigh_build( n, s );
}

cmd_status cmd_neigh_info( const cmd_neigh *n, int *n_rows, int64_t *nnz,
                           int *max_row, int *max_neigh_guess )
{
    CMD_REQUIRE( n, CMD_ERR_ARG, "neighbor list is NULL" );
    if ( nnz )
    {
        CMD_TRY( neigh_csr( const_cast<cmd_neigh *>( n ) ) );
        *nnz = n->nnz;
    }
    if ( n_rows )
        *n_rows = n->n_rows;
    if ( max_row )
        *max_row = n->max_row;
    if ( max_neigh_guess )
        *max_neigh_guess = n->max_neigh_guess;
    return CMD_OK;
}

cmd_status cmd_neigh_set_tile_width( cmd_neigh *n, int width )
{
    CMD_REQUIRE( n, CMD_ERR_ARG, "neighbor list is NULL" );
    CMD_REQUIRE( width >= 0 && width <= 3, CMD_ERR_ARG,
                 "tile width must be 0 (choose), 1, 2 or 3 columns" );
    if ( width != n->tile_w )
    {
        n->tile_w = width;
        n->n_rows = 0; // layout changed: must rebuild
        n->csr_valid = false;
    }
    return CMD_OK;
}

cmd_status cmd_neigh_tile_shape( const cmd_neigh *n, int *width, int *height )
{
    CMD_REQUIRE( n, CMD_ERR_ARG, "neighbor list is NULL" );
    if ( width )
        *width = n->tiled ? n->tg.W : 0;
    if ( height )
        *height = n->tiled ? n->tg.E : 0;
    return CMD_OK;
}

cmd_status cmd_neigh_layout( const cmd_neigh *n, int *tiled, int *tile_height,
                             int *lanes )
{
    CMD_REQUIRE( n, CMD_ERR_ARG, "neighbor list is NULL" );
    if ( tiled )
        *tiled = n->tiled ? 1 : 0;
    if ( tile_height )
        *tile_height = n->tiled ? n->tg.E : 0;
    if ( lanes )
        *lanes = n->lanes;
    return CMD_OK;
}

cmd_status cmd_neigh_download( cmd_neigh *n, int64_t *offsets, int *idx )
{
    CMD_REQUIRE( n, CMD_ERR_ARG, "neighbor list is NULL" );
    CMD_REQUIRE( n->stream || n->n_rows == 0, CMD_ERR_STATE,
                 "neighbor list was never built" );
    CMD_TRY( neigh_csr( n ) );
    if ( offsets )
        CMD_HIP( hipMemcpyAsync( offsets, n->offsets.p,
                                 sizeof( int64_t ) * ( (size_t)n->n_rows + 1 ),
                                 hipMemcpyDeviceToHost, n->stream ) );
    if ( idx && n->nnz > 0 )
        CMD_HIP( hipMemcpyAsync( idx, n->idx.p, sizeof( int ) * (size_t)n->nnz,
                                 hipMemcpyDeviceToHost, n->stream ) );
    CMD_HIP( hipStreamSynchronize( n->stream ) );
    return CMD_OK;
}

} // extern "C"
