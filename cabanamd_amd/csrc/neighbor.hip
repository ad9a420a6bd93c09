// neighbor.hip -- CSR Verlet list build.
//
// Stands in for NeighborVerlet::create (neighbor_verlet.h:43-62), which
// builds a Cabana::VerletList over rows [0,N_local) and candidates
// [0,N_local+N_ghost) with radius neigh_cut.  Accept rule (Cabana
// FullNeighborTag / HalfNeighborTag discriminators): j != i, rsq <= rc^2
// with rsq = dx*dx + dy*dy + dz*dz and d = x_i - x_j evaluated without
// contraction; half lists keep j only if x_j is lexicographically greater
// than x_i (x, then y, then z).
//
// Build: (1) bounding box of all atoms, (2) stable counting sort of all
// atoms into a grid with cells >= rc, (3) one 64-lane wave per cell stages
// the 27-cell candidate set in LDS in tiles and, for each local atom of the
// cell, tests 64 candidates per step, counting by ballot/popcount (count
// pass) or writing by ballot prefix (fill pass); (4) scan of the per-row
// counts into int64 CSR offsets between the passes.  Rows come out in
// stencil-cell order (x-cell, y-cell, z-cell) and index order inside a
// cell, so a build is deterministic.
#include "internal.h"

#include <algorithm>
#include <cfloat>
#include <cmath>

namespace cmd
{
static constexpr int TPB = 256;
static constexpr int NB_TILE = 512; // candidates staged per LDS tile
static constexpr int NB_ATOMS = 64; // row counters kept per chunk of a cell

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

__global__ void k_bbox_final( int nb, const double *__restrict__ part,
                              double *__restrict__ out )
{
    if ( threadIdx.x < 6 )
    {
        const int k = threadIdx.x;
        double a = part[k];
        for ( int b = 1; b < nb; ++b )
            a = ( k < 3 ) ? fmin( a, part[b * 6 + k] ) : fmax( a, part[b * 6 + k] );
        out[k] = a;
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

template <bool FILL>
__global__ __launch_bounds__( 64 ) void k_neigh_rows(
    const double4 *__restrict__ xs, const int *__restrict__ start, int ng0,
    int ng1, int ng2, int n_local, double rc2, int half,
    int *__restrict__ counts, const int64_t *__restrict__ offsets,
    int *__restrict__ idx, int *__restrict__ max_row )
{
    const int c = blockIdx.x;
    const int cs = start[c], ce = start[c + 1];
    if ( cs == ce )
        return;
    const int lane = threadIdx.x;
    __shared__ int s_rb[9], s_roff[10];
    __shared__ double4 s_c[NB_TILE];
    __shared__ int s_cnt[NB_ATOMS];
    if ( lane == 0 )
    {
        const int a = c / ( ng1 * ng2 ), b = ( c / ng2 ) % ng1, e = c % ng2;
        const int e0 = e > 0 ? e - 1 : 0;
        const int e1 = e < ng2 - 1 ? e + 1 : ng2 - 1;
        int q = 0, tot = 0;
        for ( int da = -1; da <= 1; ++da )
            for ( int db = -1; db <= 1; ++db )
            {
                const int aa = a + da, bb = b + db;
                int rb = 0, re = 0;
                if ( aa >= 0 && aa < ng0 && bb >= 0 && bb < ng1 )
                {
                    const int base = ( aa * ng1 + bb ) * ng2;
                    rb = start[base + e0];
                    re = start[base + e1 + 1];
                }
                s_rb[q] = rb;
                s_roff[q] = tot;
                tot += re - rb;
                ++q;
            }
        s_roff[9] = tot;
    }
    __syncthreads();
    const int ncand = s_roff[9];
    for ( int a0 = cs; a0 < ce; a0 += NB_ATOMS )
    {
        const int na = min( NB_ATOMS, ce - a0 );
        if ( lane < na )
            s_cnt[lane] = 0;
        __syncthreads();
        for ( int t0 = 0; t0 < ncand; t0 += NB_TILE )
        {
            const int nt = min( NB_TILE, ncand - t0 );
            for ( int q = lane; q < nt; q += 64 )
            {
                const int qq = t0 + q;
                int r = 0;
                while ( qq >= s_roff[r + 1] )
                    ++r;
                s_c[q] = xs[s_rb[r] + ( qq - s_roff[r] )];
            }
            __syncthreads();
            for ( int t = 0; t < na; ++t )
            {
                const double4 pi = xs[a0 + t];
                const int i = (int)pi.w;
                if ( i >= n_local )
                    continue;
                int cnt = s_cnt[t];
                const int64_t rbase = FILL ? offsets[i] : 0;
                for ( int q0 = 0; q0 < nt; q0 += 64 )
                {
                    const int q = q0 + lane;
                    bool acc = false;
                    int j = 0;
                    if ( q < nt )
                    {
                        const double4 pj = s_c[q];
                        j = (int)pj.w;
                        acc = accept_pair( pi, pj, i, j, rc2, half );
                    }
                    const uint64_t m = __ballot( acc );
                    if ( FILL && acc )
                    {
                        const int pre = __builtin_amdgcn_mbcnt_hi(
                            (unsigned)( m >> 32 ),
                            __builtin_amdgcn_mbcnt_lo( (unsigned)m, 0u ) );
                        idx[rbase + cnt + pre] = j;
                    }
                    cnt += __popcll( m );
                }
                if ( lane == 0 )
                    s_cnt[t] = cnt;
            }
            __syncthreads();
        }
        if ( !FILL && lane < na )
        {
            const int i = (int)xs[a0 + lane].w;
            if ( i < n_local )
            {
                counts[i] = s_cnt[lane];
                atomicMax( max_row, s_cnt[lane] );
            }
        }
        __syncthreads();
    }
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

static cmd_status neigh_build( cmd_neigh *N, cmd_system *S )
{
    hipStream_t st = S->stream;
    N->stream = st;
    const int n_local = S->n_local;
    const int nt = S->n_local + S->n_ghost;
    N->n_rows = n_local;
    CMD_TRY( N->offsets.reserve( (size_t)n_local + 1 ) );
    CMD_TRY( N->counts.reserve( std::max( n_local, 1 ) ) );
    if ( nt == 0 || n_local == 0 )
    {
        CMD_TRY( fill_i32( (int *)N->offsets.p, 0, 2 * ( (size_t)n_local + 1 ),
                           st ) );
        N->nnz = 0;
        N->max_row = 0;
        return CMD_OK;
    }
    // (1) bounding box
    const int nbb = std::min( 1024, std::max( 1, ( nt + TPB - 1 ) / TPB ) );
    CMD_TRY( S->sc.dpart.reserve( 6 * (size_t)nbb ) );
    CMD_TRY( S->sc.dres.reserve( 8 ) );
    k_bbox<<<nbb, TPB, 0, st>>>( nt, S->x.p, S->sc.dpart.p );
    CMD_LAUNCH_CHECK();
    k_bbox_final<<<1, 64, 0, st>>>( nbb, S->sc.dpart.p, S->sc.dres.p );
    CMD_LAUNCH_CHECK();
    double bb[6];
    CMD_HIP( hipMemcpyAsync( bb, S->sc.dres.p, sizeof( bb ),
                             hipMemcpyDeviceToHost, st ) );
    CMD_HIP( hipStreamSynchronize( st ) );
    // (2) grid and counting sort of all atoms
    int ng[3];
    double rcell[3];
    neigh_grid( bb, bb + 3, N->neigh_cut, nt, ng, rcell );
    CellGrid g;
    g.floor_mode = 0;
    for ( int k = 0; k < 3; ++k )
    {
        g.min[k] = bb[k];
        g.rd[k] = rcell[k];
        g.n[k] = ng[k];
        N->ng[k] = ng[k];
    }
    CMD_TRY( cell_sort( S->x.p, nt, g, N->grid_sort, S->sc, st ) );
    CMD_TRY( N->xs.reserve( nt ) );
    k_gather_sorted<<<( nt + TPB - 1 ) / TPB, TPB, 0, st>>>(
        nt, N->grid_sort.perm.p, S->x.p, N->xs.p );
    CMD_LAUNCH_CHECK();
    // (3) count pass
    const int ncell = N->grid_sort.ncell;
    const double rc2 = N->neigh_cut * N->neigh_cut;
    CMD_TRY( S->sc.ires.reserve( 4 ) );
    CMD_TRY( fill_i32( S->sc.ires.p, 0, 1, st ) );
    k_neigh_rows<false><<<ncell, 64, 0, st>>>(
        N->xs.p, N->grid_sort.start.p, ng[0], ng[1], ng[2], n_local, rc2,
        N->half, N->counts.p, nullptr, nullptr, S->sc.ires.p );
    CMD_LAUNCH_CHECK();
    // (4) offsets
    int64_t nnz = 0;
    CMD_TRY( scan_exclusive( N->counts.p, N->offsets.p, n_local, S->sc, st,
                             &nnz ) );
    int mr = 0;
    CMD_HIP( hipMemcpyAsync( &mr, S->sc.ires.p, sizeof( int ),
                             hipMemcpyDeviceToHost, st ) );
    CMD_HIP( hipStreamSynchronize( st ) );
    N->nnz = nnz;
    N->max_row = mr;
    CMD_TRY( N->idx.reserve( (size_t)std::max<int64_t>( nnz, 1 ) +
                             (size_t)( nnz / 10 ) ) );
    // (5) fill pass
    k_neigh_rows<true><<<ncell, 64, 0, st>>>(
        N->xs.p, N->grid_sort.start.p, ng[0], ng[1], ng[2], n_local, rc2,
        N->half, nullptr, N->offsets.p, N->idx.p, nullptr );
    CMD_LAUNCH_CHECK();
    // NeighborVerlet::create regrows its capacity guess (neighbor_verlet.h:58-61)
    if ( N->max_row > N->max_neigh_guess )
        N->max_neigh_guess = (int)( N->max_row * 1.1 );
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
    n->max_neigh_guess = max_neigh_guess;
    *out = n;
    return CMD_OK;
}

cmd_status cmd_neigh_destroy( cmd_neigh *n )
{
    delete n;
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
    if ( n_rows )
        *n_rows = n->n_rows;
    if ( nnz )
        *nnz = n->nnz;
    if ( max_row )
        *max_row = n->max_row;
    if ( max_neigh_guess )
        *max_neigh_guess = n->max_neigh_guess;
    return CMD_OK;
}

cmd_status cmd_neigh_download( cmd_neigh *n, int64_t *offsets, int *idx )
{
    CMD_REQUIRE( n, CMD_ERR_ARG, "neighbor list is NULL" );
    CMD_REQUIRE( n->offsets.p, CMD_ERR_STATE, "neighbor list was never built" );
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
