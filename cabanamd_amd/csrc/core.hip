// core.hip -- device management, System state, scans, linked-cell sort,
// velocity-Verlet, periodic wrap, 6-phase ghosts and thermo reductions.
//
// Reference lines restated here (CabanaMD source tree):
//   integrator_nve.h:91-110, integrator_nve_impl.h:49-83   (NVE)
//   comm_mpi.h:141-168, comm_mpi_impl.h:191-441            (wrap, ghosts)
//   binning_cabana_impl.h:57-113                           (cell sort)
//   property_temperature*.h, property_kine*.h              (thermo)
#include "internal.h"

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstring>
#include <mutex>

namespace cmd
{
static thread_local std::string g_err;

void set_error( const char *fmt, ... )
{
    char buf[1024];
    va_list ap;
    va_start( ap, fmt );
    vsnprintf( buf, sizeof( buf ), fmt, ap );
    va_end( ap );
    g_err = buf;
}

static constexpr int TPB = 256;
static inline int nblocks_for( int64_t n, int tpb = TPB )
{
    return (int)std::max<int64_t>( 1, ( n + tpb - 1 ) / tpb );
}

// ---------------------------------------------------------------- fills
template <class T>
__global__ void k_fill( T *__restrict__ p, T v, size_t n )
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for ( ; i < n; i += stride )
        p[i] = v;
}

cmd_status fill_i32( int *p, int v, size_t n, hipStream_t st )
{
    if ( n == 0 )
        return CMD_OK;
    int nb = (int)std::min<size_t>( ( n + TPB - 1 ) / TPB, 65536 );
    k_fill<int><<<nb, TPB, 0, st>>>( p, v, n );
    CMD_LAUNCH_CHECK();
    return CMD_OK;
}

__global__ void k_zero2( int *__restrict__ p, int *__restrict__ q, size_t n )
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for ( ; i < n; i += stride )
    {
        p[i] = 0;
        q[i] = 0;
    }
}

cmd_status zero2_i32( int *p, int *q, size_t n, hipStream_t st )
{
    if ( n == 0 )
        return CMD_OK;
    int nb = (int)std::min<size_t>( ( n + TPB - 1 ) / TPB, 65536 );
    k_zero2<<<nb, TPB, 0, st>>>( p, q, n );
    CMD_LAUNCH_CHECK();
    return CMD_OK;
}

cmd_status fill_f64( double *p, double v, size_t n, hipStream_t st )
{
    if ( n == 0 )
        return CMD_OK;
    int nb = (int)std::min<size_t>( ( n + TPB - 1 ) / TPB, 65536 );
    k_fill<double><<<nb, TPB, 0, st>>>( p, v, n );
    CMD_LAUNCH_CHECK();
    return CMD_OK;
}

// ---------------------------------------------------------------- scans
// Reduce-then-scan over tiles of SCAN_TILE elements: per-tile sums, one
// block scans the tile sums, then each tile rescans with its offset.
static constexpr int SCAN_ITEMS = 8;
static constexpr int SCAN_TILE = TPB * SCAN_ITEMS;

// Exclusive scan of one value per thread over a 256-thread block.
__device__ int64_t block_exclusive_scan( int64_t v, int64_t *total )
{
    __shared__ int64_t wsum[TPB / 64];
    const int lane = threadIdx.x & 63;
    const int wid = threadIdx.x >> 6;
    int64_t s = v;
#pragma unroll
    for ( int o = 1; o < 64; o <<= 1 )
    {
        int64_t t = __shfl_up( s, o, 64 );
        if ( lane >= o )
            s += t;
    }
    if ( lane == 63 )
        wsum[wid] = s;
    __syncthreads();
    int64_t woff = 0, tot = 0;
#pragma unroll
    for ( int w = 0; w < TPB / 64; ++w )
    {
        if ( w < wid )
            woff += wsum[w];
        tot += wsum[w];
    }
    __syncthreads();
    *total = tot;
    return woff + s - v;
}

template <class InT>
__global__ void k_scan_tile_sums( const InT *__restrict__ in, int n,
                                  int64_t *__restrict__ blk )
{
    const int64_t base = (int64_t)blockIdx.x * SCAN_TILE;
    int64_t s = 0;
    for ( int k = 0; k < SCAN_ITEMS; ++k )
    {
        int64_t i = base + (int64_t)k * TPB + threadIdx.x;
        if ( i < n )
            s += in[i];
    }
    int64_t tot;
    block_exclusive_scan( s, &tot );
    if ( threadIdx.x == 0 )
        blk[blockIdx.x] = tot;
}

template <class OutT>
__global__ void k_scan_tile_offsets( int64_t *__restrict__ blk, int nblk,
                                     OutT *__restrict__ out, int n,
                                     int64_t *__restrict__ total_slot )
{
    int64_t carry = 0;
    for ( int base = 0; base < nblk; base += TPB )
    {
        int i = base + threadIdx.x;
        int64_t v = ( i < nblk ) ? blk[i] : 0;
        int64_t tot;
        int64_t e = block_exclusive_scan( v, &tot );
        if ( i < nblk )
            blk[i] = carry + e;
        carry += tot;
    }
    if ( threadIdx.x == 0 )
    {
        out[n] = (OutT)carry;
        if ( total_slot )
            *total_slot = carry;
    }
}

template <class InT, class OutT>
__global__ void k_scan_tile_apply( const InT *__restrict__ in, int n,
                                   const int64_t *__restrict__ blk,
                                   OutT *__restrict__ out )
{
    const int64_t base = (int64_t)blockIdx.x * SCAN_TILE +
                         (int64_t)threadIdx.x * SCAN_ITEMS;
    InT vals[SCAN_ITEMS];
    int64_t s = 0;
#pragma unroll
    for ( int k = 0; k < SCAN_ITEMS; ++k )
    {
        int64_t i = base + k;
        vals[k] = ( i < n ) ? in[i] : (InT)0;
        s += vals[k];
    }
    int64_t tot;
    int64_t e = block_exclusive_scan( s, &tot ) + blk[blockIdx.x];
#pragma unroll
    for ( int k = 0; k < SCAN_ITEMS; ++k )
    {
        int64_t i = base + k;
        if ( i < n )
            out[i] = (OutT)e;
        e += vals[k];
    }
}

template <class OutT>
static cmd_status scan_impl( const int *in, OutT *out, int n, Scratch &sc,
                             hipStream_t st, int64_t *total_host )
{
    const int nblk = std::max( 1, ( n + SCAN_TILE - 1 ) / SCAN_TILE );
    CMD_TRY( sc.blk.reserve( nblk + 2 ) );
    int64_t *slot = sc.blk.p + nblk + 1;
    if ( n > 0 )
    {
        k_scan_tile_sums<int><<<nblk, TPB, 0, st>>>( in, n, sc.blk.p );
        CMD_LAUNCH_CHECK();
    }
    else
    {
        CMD_TRY( fill_i32( (int *)sc.blk.p, 0, 2, st ) );
    }
    k_scan_tile_offsets<OutT><<<1, TPB, 0, st>>>( sc.blk.p, n > 0 ? nblk : 0,
                                                  out, n, slot );
    CMD_LAUNCH_CHECK();
    if ( n > 0 )
    {
        k_scan_tile_apply<int, OutT><<<nblk, TPB, 0, st>>>( in, n, sc.blk.p,
                                                            out );
        CMD_LAUNCH_CHECK();
    }
    if ( total_host )
    {
        CMD_HIP( hipMemcpyAsync( total_host, slot, sizeof( int64_t ),
                                 hipMemcpyDeviceToHost, st ) );
        CMD_HIP( hipStreamSynchronize( st ) );
    }
    return CMD_OK;
}

cmd_status scan_exclusive( const int *in, int *out, int n, Scratch &sc,
                           hipStream_t st, int64_t *total_host )
{
    return scan_impl<int>( in, out, n, sc, st, total_host );
}

cmd_status scan_exclusive( const int *in, int64_t *out, int n, Scratch &sc,
                           hipStream_t st, int64_t *total_host )
{
    return scan_impl<int64_t>( in, out, n, sc, st, total_host );
}

// ---------------------------------------------------------------- cell sort
__device__ __forceinline__ int cell_of_pos( const CellGrid &g, double4 p )
{
    int a = cell_coord( p.x, g.min[0], g.rd[0], g.n[0], g.floor_mode );
    int b = cell_coord( p.y, g.min[1], g.rd[1], g.n[1], g.floor_mode );
    int c = cell_coord( p.z, g.min[2], g.rd[2], g.n[2], g.floor_mode );
    return ( a * g.n[1] + b ) * g.n[2] + c;
}

// Atoms arrive mostly in cell order (they were sorted at the last rebuild
// and have moved less than a cell since), so a wave's lanes form runs of
// equal cell ids.  One lane per run does the atomic for the whole run.
// Valid lanes are a prefix of the wave.
struct CellRun
{
    bool head;    // first lane of its run
    int head_lane;
    int len;      // run length (meaningful on the head lane)
};

__device__ __forceinline__ CellRun cell_run( int c, bool valid )
{
    const int lane = threadIdx.x & 63;
    const int prev = __shfl_up( c, 1, 64 );
    CellRun R;
    R.head = valid && ( lane == 0 || prev != c );
    const unsigned long long hm = __builtin_amdgcn_ballot_w64( R.head );
    const int nvalid = (int)__popcll( __builtin_amdgcn_ballot_w64( valid ) );
    const unsigned long long upto = ( 2ull << lane ) - 1ull; // lanes 0..lane
    const unsigned long long below = hm & upto, above = hm & ~upto;
    R.head_lane = below ? 63 - __clzll( (long long)below ) : lane;
    R.len = ( above ? __ffsll( (long long)above ) - 1 : nvalid ) - lane;
    return R;
}

__global__ void k_cell_count( const double4 *__restrict__ x, int n,
                              CellGrid g, int *__restrict__ cell_of,
                              int *__restrict__ count )
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = i < n;
    int c = -1;
    if ( valid )
    {
        c = cell_of_pos( g, x[i] );
        cell_of[i] = c;
    }
    const CellRun R = cell_run( c, valid );
    if ( R.head )
        atomicAdd( &count[c], R.len );
}

__global__ void k_cell_count_ids( int n, const int *__restrict__ cell_of,
                                  int *__restrict__ count )
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = i < n;
    const int c = valid ? cell_of[i] : -1;
    const CellRun R = cell_run( c, valid );
    if ( R.head )
        atomicAdd( &count[c], R.len );
}

// Slots inside each cell in arrival order: a run keeps its order, runs of
// one cell land in the order their atomics happen.
__global__ void k_cell_fill( int n, const int *__restrict__ cell_of,
                             const int *__restrict__ start,
                             int *__restrict__ cursor, int *__restrict__ out )
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = i < n;
    const int c = valid ? cell_of[i] : -1;
    const CellRun R = cell_run( c, valid );
    int b = 0;
    if ( R.head )
        b = atomicAdd( &cursor[c], R.len );
    b = __shfl( b, R.head_lane, 64 );
    if ( valid )
        out[start[c] + b + ( ( (int)threadIdx.x & 63 ) - R.head_lane )] = i;
}

// Index order inside each cell, which makes the sort stable and
// deterministic: an atom's place is the number of its cell's atoms with a
// smaller index (cells hold tens of atoms at the densities binned here).
__global__ void k_cell_rank( int n, const int *__restrict__ cell_of,
                             const int *__restrict__ start,
                             const int *__restrict__ in, int *__restrict__ perm )
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if ( k >= n )
        return;
    const int v = in[k];
    const int c = cell_of[v];
    const int b = start[c], e = start[c + 1];
    int r = 0;
    for ( int m = b; m < e; ++m )
        r += in[m] < v;
    perm[b + r] = v;
}

// Shared tail of both sorts: scan of the counts, fill, rank.
static cmd_status cell_sort_tail( int n, int ncell, CellSort &cs, Scratch &sc,
                                  hipStream_t st )
{
    CMD_TRY( scan_exclusive( cs.count.p, cs.start.p, ncell, sc, st, nullptr ) );
    if ( n > 0 )
    {
        k_cell_fill<<<nblocks_for( n ), TPB, 0, st>>>( n, cs.cell_of.p,
                                                      cs.start.p, cs.cursor.p,
                                                      cs.tmp.p );
        CMD_LAUNCH_CHECK();
        k_cell_rank<<<nblocks_for( n ), TPB, 0, st>>>( n, cs.cell_of.p,
                                                      cs.start.p, cs.tmp.p,
                                                      cs.perm.p );
        CMD_LAUNCH_CHECK();
    }
    return CMD_OK;
}

static cmd_status cell_sort_reserve( int n, int ncell, CellSort &cs, hipStream_t st )
{
    cs.ncell = ncell;
    CMD_TRY( cs.count.reserve( ncell ) );
    CMD_TRY( cs.cursor.reserve( ncell ) );
    CMD_TRY( cs.start.reserve( ncell + 1 ) );
    CMD_TRY( cs.perm.reserve( std::max( n, 1 ) ) );
    CMD_TRY( cs.tmp.reserve( std::max( n, 1 ) ) );
    CMD_TRY( zero2_i32( cs.count.p, cs.cursor.p, ncell, st ) );
    return CMD_OK;
}

cmd_status cell_sort( const double4 *x, int n, const CellGrid &g, CellSort &cs,
                      Scratch &sc, hipStream_t st )
{
    const int ncell = g.n[0] * g.n[1] * g.n[2];
    CMD_REQUIRE( ncell > 0, CMD_ERR_ARG, "cell grid is empty" );
    CMD_TRY( cs.cell_of.reserve( std::max( n, 1 ) ) );
    CMD_TRY( cell_sort_reserve( n, ncell, cs, st ) );
    if ( n > 0 )
    {
        k_cell_count<<<nblocks_for( n ), TPB, 0, st>>>( x, n, g, cs.cell_of.p,
                                                       cs.count.p );
        CMD_LAUNCH_CHECK();
    }
    return cell_sort_tail( n, ncell, cs, sc, st );
}

cmd_status cell_sort_ids( int n, int ncell, CellSort &cs, Scratch &sc,
                          hipStream_t st )
{
    CMD_REQUIRE( ncell > 0, CMD_ERR_ARG, "cell grid is empty" );
    CMD_REQUIRE( (int64_t)cs.cell_of.n >= n, CMD_ERR_STATE,
                 "cell ids were not computed" );
    CMD_TRY( cell_sort_reserve( n, ncell, cs, st ) );
    if ( n > 0 )
    {
        k_cell_count_ids<<<nblocks_for( n ), TPB, 0, st>>>( n, cs.cell_of.p,
                                                           cs.count.p );
        CMD_LAUNCH_CHECK();
    }
    return cell_sort_tail( n, ncell, cs, sc, st );
}

// ---------------------------------------------------------------- NVE
// integrator_nve.h:91-101 (TagInitial) and :103-110 (TagFinal).
__global__ void k_nve_initial( int n, int cap, double4 *__restrict__ x,
                               double *__restrict__ v,
                               const double *__restrict__ f, MassTable mt,
                               double dtf, double dtv )
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if ( i >= n )
        return;
    double4 p = x[i];
    const double dtfm = dtf / mt.m[(int)p.w];
    double vx = v[i] + dtfm * f[i];
    double vy = v[cap + i] + dtfm * f[cap + i];
    double vz = v[2 * cap + i] + dtfm * f[2 * cap + i];
    v[i] = vx;
    v[cap + i] = vy;
    v[2 * cap + i] = vz;
    p.x = p.x + dtv * vx;
    p.y = p.y + dtv * vy;
    p.z = p.z + dtv * vz;
    x[i] = p;
}

__global__ void k_nve_final( int n, int cap, const double4 *__restrict__ x,
                             double *__restrict__ v,
                             const double *__restrict__ f, MassTable mt,
                             double dtf )
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if ( i >= n )
        return;
    const double dtfm = dtf / mt.m[(int)x[i].w];
    v[i] = v[i] + dtfm * f[i];
    v[cap + i] = v[cap + i] + dtfm * f[cap + i];
    v[2 * cap + i] = v[2 * cap + i] + dtfm * f[2 * cap + i];
}

// Partial sums of one block (256 threads) into partial[blockIdx*4 + k].
__device__ void block_partials4( double a, double b, double c, double d,
                                 double *__restrict__ partial )
{
    __shared__ double red[4][TPB / 64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for ( int o = 32; o > 0; o >>= 1 )
    {
        a += __shfl_down( a, o, 64 );
        b += __shfl_down( b, o, 64 );
        c += __shfl_down( c, o, 64 );
        d += __shfl_down( d, o, 64 );
    }
    if ( lane == 0 )
    {
        red[0][wid] = a;
        red[1][wid] = b;
        red[2][wid] = c;
        red[3][wid] = d;
    }
    __syncthreads();
    if ( threadIdx.x < 4 )
    {
        double s = 0.0;
        for ( int w = 0; w < TPB / 64; ++w )
            s += red[threadIdx.x][w];
        partial[blockIdx.x * 4 + threadIdx.x] = s;
    }
}

// Fused final kick of this step, thermo partials, and (optionally) the next
// step's half kick + drift (+ wrap) into x_next.  Used after half-list
// forces, whose per-atom totals are only known after the atomics finish.
__global__ void k_final_initial( int n, int cap, const double4 *__restrict__ x,
                                 double4 *__restrict__ x_next,
                                 double *__restrict__ v,
                                 double *__restrict__ f, MassTable mt,
                                 StepCtl ctl, double *__restrict__ partial )
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    double mv2 = 0.0;
    if ( i < n )
    {
        double4 p = x[i];
        const double m = mt.m[(int)p.w];
        const double dtfm = ctl.dtf / m;
        const double fx = f[i], fy = f[cap + i], fz = f[2 * cap + i];
        if ( ctl.zero_f )
        {
            f[i] = 0.0;
            f[cap + i] = 0.0;
            f[2 * cap + i] = 0.0;
        }
        double vx = v[i] + dtfm * fx;
        double vy = v[cap + i] + dtfm * fy;
        double vz = v[2 * cap + i] + dtfm * fz;
        mv2 = ( vx * vx + vy * vy + vz * vz ) * m;
        if ( ctl.do_next_initial )
        {
            vx = vx + dtfm * fx;
            vy = vy + dtfm * fy;
            vz = vz + dtfm * fz;
            p.x = p.x + ctl.dtv * vx;
            p.y = p.y + ctl.dtv * vy;
            p.z = p.z + ctl.dtv * vz;
            if ( ctl.wrap_next )
            {
                const double cx = p.x, cy = p.y, cz = p.z;
                if ( cx > ctl.L[0] )
                    p.x = cx - ctl.L[0];
                if ( cx < 0 )
                    p.x = cx + ctl.L[0];
                if ( cy > ctl.L[1] )
                    p.y = cy - ctl.L[1];
                if ( cy < 0 )
                    p.y = cy + ctl.L[1];
                if ( cz > ctl.L[2] )
                    p.z = cz - ctl.L[2];
                if ( cz < 0 )
                    p.z = cz + ctl.L[2];
            }
            x_next[i] = p;
            // the next step's ghost refresh (update_halo,
            // comm_mpi_impl.h:369-408): the drifted position, shifted per
            // image, at the image's sorted ghost slot
            if ( ctl.xg_next )
                for ( int m = ctl.img_start[i]; m < ctl.img_start[i + 1]; ++m )
                {
                    const unsigned e = (unsigned)ctl.img_pk[m];
                    const unsigned code = e >> IMG_CODE_SHIFT;
                    double4 g = p;
                    if ( code & 1 )
                        g.x = g.x - ctl.L[0];
                    if ( code & 2 )
                        g.x = g.x + ctl.L[0];
                    if ( code & 4 )
                        g.y = g.y - ctl.L[1];
                    if ( code & 8 )
                        g.y = g.y + ctl.L[1];
                    if ( code & 16 )
                        g.z = g.z - ctl.L[2];
                    if ( code & 32 )
                        g.z = g.z + ctl.L[2];
                    ctl.xg_next[e & IMG_SLOT_MASK] = g;
                }
        }
        v[i] = vx;
        v[cap + i] = vy;
        v[2 * cap + i] = vz;
    }
    if ( ctl.thermo )
        block_partials4( mv2, 0.0, 0.0, 0.0, partial );
}

cmd_status nve_final_initial( cmd_system *S, const StepCtl &ctl,
                              double4 *x_next, double *partial,
                              int *nblocks_out )
{
    const int n = S->n_local;
    const int nb = nblocks_for( n );
    if ( nblocks_out )
        *nblocks_out = nb;
    k_final_initial<<<nb, TPB, 0, S->stream>>>( n, S->cap, S->x.p, x_next,
                                                S->v.p, S->f.p,
                                                S->mass_table(), ctl, partial );
    CMD_LAUNCH_CHECK();
    return CMD_OK;
}

// Deterministic reduction of nblocks rows of 4 partials (fixed order).
__global__ void k_reduce_partials( const double *__restrict__ partial,
                                   int nblocks, double *__restrict__ out )
{
    double s[4] = { 0.0, 0.0, 0.0, 0.0 };
    for ( int b = threadIdx.x; b < nblocks; b += TPB )
        for ( int k = 0; k < 4; ++k )
            s[k] += partial[b * 4 + k];
    double tmp[4];
    for ( int k = 0; k < 4; ++k )
        tmp[k] = s[k];
    // reuse the block helper with blockIdx.x == 0 semantics via out
    __shared__ double red[4][TPB / 64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for ( int k = 0; k < 4; ++k )
    {
        double a = tmp[k];
#pragma unroll
        for ( int o = 32; o > 0; o >>= 1 )
            a += __shfl_down( a, o, 64 );
        if ( lane == 0 )
            red[k][wid] = a;
    }
    __syncthreads();
    if ( threadIdx.x < 4 )
    {
        double a = 0.0;
        for ( int w = 0; w < TPB / 64; ++w )
            a += red[threadIdx.x][w];
        out[threadIdx.x] = a;
    }
}

cmd_status reduce_partials( const double *partial, int nblocks, double *out,
                            hipStream_t st )
{
    k_reduce_partials<<<1, TPB, 0, st>>>( partial, nblocks, out );
    CMD_LAUNCH_CHECK();
    return CMD_OK;
}

// ---------------------------------------------------------------- wrap
// comm_mpi.h:141-168: per dimension, x > L -> x - L, x < 0 -> x + L, both
// tested on the value before the update.
__global__ void k_wrap( int n, double4 *__restrict__ x, double Lx, double Ly,
                        double Lz )
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if ( i >= n )
        return;
    double4 p = x[i];
    const double cx = p.x, cy = p.y, cz = p.z;
    if ( cx > Lx )
        p.x = cx - Lx;
    if ( cx < 0 )
        p.x = cx + Lx;
    if ( cy > Ly )
        p.y = cy - Ly;
    if ( cy < 0 )
        p.y = cy + Ly;
    if ( cz > Lz )
        p.z = cz - Lz;
    if ( cz < 0 )
        p.z = cz + Lz;
    x[i] = p;
}

cmd_status wrap_local( cmd_system *S )
{
    if ( S->n_local == 0 )
        return CMD_OK;
    k_wrap<<<nblocks_for( S->n_local ), TPB, 0, S->stream>>>(
        S->n_local, S->x.p, S->dom.L[0], S->dom.L[1], S->dom.L[2] );
    CMD_LAUNCH_CHECK();
    return CMD_OK;
}

// ---------------------------------------------------------------- ghosts
// The six-phase halo build (comm_mpi_impl.h:280-367, TagHaloPack
// comm_mpi.h:240-319, TagHaloPBC :323-353) in one pass.  Phase p works on
// dimension p / 2: an even phase sends the atoms of the current set with
// x >= hi - depth, shifted by -L; an odd phase those with x <= lo + depth,
// shifted by +L, skipping the ghosts the even phase of its dimension has
// just made.  Ghosts of earlier phases are sent on, and their coordinate in
// the phase's dimension is their root's.  So a local atom has one image per
// set of phases (at most one per dimension) whose conditions its own
// coordinates meet, and since each phase appends in the order of its
// sources, the ghosts come out grouped by that set, the sets in the order
// below, each group in ascending root index.  Bit p of a set is phase p.
static constexpr int HALO_SETS = 26;
__host__ __device__ constexpr int halo_set( int j )
{
    constexpr int s[HALO_SETS] = { 1,  2,  4,  5,  6,  8,  9,  10, 16, 17, 18, 20, 21,
                                   22, 24, 25, 26, 32, 33, 34, 36, 37, 38, 40, 41, 42 };
    return s[j];
}
// the set an image was sent on from (its set without its last phase), as
// an index into the table above; -1 for the local atom itself
__host__ __device__ constexpr int halo_source( int j )
{
    constexpr int s[HALO_SETS] = { -1, -1, -1, 0,  1, -1, 0, 1, -1, 0, 1, 2, 3,
                                   4,  5,  6,  7, -1, 0, 1, 2, 3, 4, 5, 6, 7 };
    return s[j];
}
__host__ __device__ constexpr int halo_phase( int j )
{
    return halo_set( j ) >= 32 ? 5 : halo_set( j ) >= 16 ? 4 : halo_set( j ) >= 8 ? 3
         : halo_set( j ) >= 4 ? 2 : halo_set( j ) >= 2 ? 1 : 0;
}

struct HaloFaces
{
    double hi[3]; // even phases send x >= hi[d]
    double lo[3]; // odd phases send x <= lo[d]
    double L[3];
};

__device__ __forceinline__ int halo_mask( const HaloFaces &H, double4 p )
{
    const double c[3] = { p.x, p.y, p.z };
    int m = 0;
#pragma unroll
    for ( int d = 0; d < 3; ++d )
        m |= ( c[d] >= H.hi[d] ? 1 : 0 ) << ( 2 * d ) | ( c[d] <= H.lo[d] ? 1 : 0 ) << ( 2 * d + 1 );
    return m;
}

static HaloFaces halo_faces( const cmd_system *s, double comm_depth )
{
    HaloFaces H;
    for ( int d = 0; d < 3; ++d )
    {
        H.hi[d] = s->dom.local_hi[d] - comm_depth;
        H.lo[d] = s->dom.local_lo[d] + comm_depth;
        H.L[d] = s->dom.L[d];
    }
    return H;
}

// The totals per set (tot, unless null) and, in the pass over sorted atoms
// (MASKS), each atom's phase mask and the images per set and block
// (cnt[j * nblk + block]).  That pass takes one atom per thread; the
// totals-only pass strides over the atoms with as many blocks as launched
// (fewer blocks, fewer global atomics on the 26 totals).
template <bool MASKS>
__global__ __launch_bounds__( TPB ) void k_halo_count( int n, const double4 *__restrict__ x,
                                                       HaloFaces H, int *__restrict__ mask,
                                                       int nblk, int *__restrict__ cnt,
                                                       int *__restrict__ tot )
{
    __shared__ int s_c[HALO_SETS];
    const int lane = threadIdx.x & 63;
    if ( threadIdx.x < HALO_SETS )
        s_c[threadIdx.x] = 0;
    __syncthreads();
    const int stride = MASKS ? n : gridDim.x * blockDim.x;
    for ( int b0 = blockIdx.x * blockDim.x; b0 < n; b0 += stride )
    {
        const int i = b0 + threadIdx.x;
        int m = 0;
        if ( i < n )
        {
            m = halo_mask( H, x[i] );
            if ( MASKS )
                mask[i] = m;
        }
        if ( __builtin_amdgcn_ballot_w64( m != 0 ) )
        {
#pragma unroll
            for ( int j = 0; j < HALO_SETS; ++j )
            {
                const unsigned long long b =
                    __builtin_amdgcn_ballot_w64( ( m & halo_set( j ) ) == halo_set( j ) );
                if ( lane == 0 && b )
                    atomicAdd( &s_c[j], (int)__popcll( b ) );
            }
        }
    }
    __syncthreads();
    if ( threadIdx.x < HALO_SETS )
    {
        const int c = s_c[threadIdx.x];
        if ( MASKS )
            cnt[threadIdx.x * nblk + blockIdx.x] = c;
        if ( c && tot )
            atomicAdd( &tot[threadIdx.x], c );
    }
}

// Writes every image: ghost slot n_local + k of set j's k-th image, where
// off[j * nblk + block] (the exclusive scan of the counts) is where the
// block's images of set j start.  send_idx records each ghost's source as
// the phase saw it (the local atom, or the ghost it was sent on from).
//
// With totals counted before the sort (pre non-null, ng their sum): a
// block writes nothing unless the scan's total is ng, so no image lands
// past the ghost rows the host sized, and block 0 sets *flag when a set's
// images differ from its total.  The images' velocity and force rows are
// zeroed (cap: the planes' stride).
__global__ __launch_bounds__( TPB ) void k_halo_fill( int n, int cap, double4 *__restrict__ x,
                                                      double *__restrict__ v,
                                                      double *__restrict__ f,
                                                      int *__restrict__ id,
                                                      const int *__restrict__ mask,
                                                      int nblk, const int *__restrict__ off,
                                                      HaloFaces H, int *__restrict__ send_idx,
                                                      int *__restrict__ root,
                                                      int *__restrict__ code,
                                                      const int *__restrict__ pre, int ng,
                                                      int *__restrict__ flag )
{
    __shared__ int s_w[HALO_SETS][TPB / 64];
    if ( pre )
    {
        if ( blockIdx.x == 0 && threadIdx.x < HALO_SETS )
        {
            const int j = threadIdx.x;
            if ( off[( j + 1 ) * nblk] - off[j * nblk] != pre[j] )
                *flag = 1;
        }
        if ( off[HALO_SETS * nblk] != ng )
        {
            if ( threadIdx.x == 0 )
                *flag = 1;
            return; // block-uniform
        }
    }
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int m = i < n ? mask[i] : 0;
    const bool any = __builtin_amdgcn_ballot_w64( m != 0 ) != 0;
#pragma unroll
    for ( int j = 0; j < HALO_SETS; ++j )
    {
        const unsigned long long b =
            any ? __builtin_amdgcn_ballot_w64( ( m & halo_set( j ) ) == halo_set( j ) ) : 0ull;
        if ( lane == 0 )
            s_w[j][wv] = (int)__popcll( b );
    }
    __syncthreads();
    if ( !any )
        return;
    double4 p0 = make_double4( 0.0, 0.0, 0.0, 0.0 );
    int idv = 0;
    if ( m )
    {
        p0 = x[i];
        idv = id[i];
    }
    int pos[HALO_SETS];
#pragma unroll
    for ( int j = 0; j < HALO_SETS; ++j )
    {
        const int c = halo_set( j );
        const bool has = ( m & c ) == c;
        const unsigned long long b = __builtin_amdgcn_ballot_w64( has );
        int k = off[j * nblk + blockIdx.x];
        for ( int w = 0; w < wv; ++w )
            k += s_w[j][w];
        k += (int)__builtin_amdgcn_mbcnt_hi( (unsigned)( b >> 32 ),
                                             __builtin_amdgcn_mbcnt_lo( (unsigned)b, 0u ) );
        pos[j] = k;
        if ( has )
        {
            double4 p = p0;
            if ( c & 1 )
                p.x -= H.L[0];
            if ( c & 2 )
                p.x += H.L[0];
            if ( c & 4 )
                p.y -= H.L[1];
            if ( c & 8 )
                p.y += H.L[1];
            if ( c & 16 )
                p.z -= H.L[2];
            if ( c & 32 )
                p.z += H.L[2];
            x[n + k] = p;
            id[n + k] = idv;
#pragma unroll
            for ( int d = 0; d < 3; ++d )
            {
                v[(size_t)d * cap + n + k] = 0.0;
                f[(size_t)d * cap + n + k] = 0.0;
            }
            root[k] = i;
            code[k] = c;
            send_idx[k] = halo_source( j ) < 0 ? i : n + pos[halo_source( j ) < 0 ? 0 : halo_source( j )];
        }
    }
}

cmd_status halo_precount( cmd_system *s, double comm_depth )
{
    s->halo_pre = false;
    const int n = s->n_local;
    if ( n == 0 || !s->has_domain )
        return CMD_OK;
    CMD_TRY( s->hw.init( HW_WORDS ) );
    CMD_TRY( s->halo_pre_tot.reserve( HALO_SETS ) );
    CMD_TRY( fill_i32( s->halo_pre_tot.p, 0, HALO_SETS, s->stream ) );
    k_halo_count<false><<<std::min( nblocks_for( n ), 1024 ), TPB, 0, s->stream>>>(
        n, s->x.p, halo_faces( s, comm_depth ), nullptr, 0, nullptr, s->halo_pre_tot.p );
    CMD_LAUNCH_CHECK();
    CMD_HIP( hipMemcpyAsync( s->hw.p + HW_HALO, s->halo_pre_tot.p, HALO_SETS * sizeof( int ),
                             hipMemcpyDeviceToHost, s->stream ) );
    CMD_HIP( hipEventRecord( s->hw.ev, s->stream ) );
    s->halo_pre = true;
    s->halo_pre_n = n;
    s->halo_pre_depth = comm_depth;
    return CMD_OK;
}

cmd_status halo_precount_check( cmd_system *s )
{
    if ( !s->halo_pre_check )
        return CMD_OK;
    s->halo_pre_check = false;
    CMD_HIP( hipEventSynchronize( s->hw.ev ) );
    CMD_REQUIRE( s->hw.p[HW_HALO_FLAG] == 0, CMD_ERR_STATE,
                 "the halo pass over sorted atoms found other ghost totals than the "
                 "count before the sort" );
    return CMD_OK;
}

// update_halo (comm_mpi_impl.h:369-408), resolved: every ghost coordinate is
// its root's coordinate with at most one +-L per dimension, exactly the
// value the phase-ordered gather + shift produces.
__global__ void k_ghost_refresh( int n_local, int n_ghost,
                                 double4 *__restrict__ x,
                                 const int *__restrict__ root,
                                 const int *__restrict__ code, double Lx,
                                 double Ly, double Lz )
{
    int g = blockIdx.x * blockDim.x + threadIdx.x;
    if ( g >= n_ghost )
        return;
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
}

// update_force (comm_mpi_impl.h:410-441): add ghost forces into their roots.
__global__ void k_ghost_fold( int n_local, int n_ghost, int cap,
                              double *__restrict__ f,
                              const int *__restrict__ root )
{
    int g = blockIdx.x * blockDim.x + threadIdx.x;
    if ( g >= n_ghost )
        return;
    const int r = root[g];
    const int s = n_local + g;
    atomicAdd( &f[r], f[s] );
    atomicAdd( &f[cap + r], f[cap + s] );
    atomicAdd( &f[2 * cap + r], f[2 * cap + s] );
}

cmd_status ghost_refresh( cmd_system *S )
{
    if ( S->n_ghost == 0 )
        return CMD_OK;
    k_ghost_refresh<<<nblocks_for( S->n_ghost ), TPB, 0, S->stream>>>(
        S->n_local, S->n_ghost, S->x.p, S->ghost_root.p, S->ghost_code.p,
        S->dom.L[0], S->dom.L[1], S->dom.L[2] );
    CMD_LAUNCH_CHECK();
    return CMD_OK;
}

// ---------------------------------------------------------------- permute
__global__ void k_permute( int n, int cap, const int *__restrict__ perm,
                           const double4 *__restrict__ x,
                           double4 *__restrict__ xo,
                           const double *__restrict__ v,
                           double *__restrict__ vo,
                           const double *__restrict__ f,
                           double *__restrict__ fo, const int *__restrict__ id,
                           int *__restrict__ ido )
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if ( k >= n )
        return;
    const int s = perm[k];
    xo[k] = x[s];
    vo[k] = v[s];
    vo[cap + k] = v[cap + s];
    vo[2 * cap + k] = v[2 * cap + s];
    if ( f ) // forces that a caller still reads
    {
        fo[k] = f[s];
        fo[cap + k] = f[cap + s];
        fo[2 * cap + k] = f[2 * cap + s];
    }
    ido[k] = id[s];
}

cmd_status permute_local( cmd_system *s, const int *perm )
{
    const int n = s->n_local;
    if ( n == 0 )
        return CMD_OK;
    CMD_TRY( s->x_alt.reserve( s->cap ) );
    CMD_TRY( s->id_alt.reserve( s->cap ) );
    CMD_TRY( s->v_alt.reserve( 3 * (size_t)s->cap ) );
    CMD_TRY( s->f_alt.reserve( 3 * (size_t)s->cap ) );
    k_permute<<<nblocks_for( n ), TPB, 0, s->stream>>>(
        n, s->cap, perm, s->x.p, s->x_alt.p, s->v.p, s->v_alt.p, s->f.p,
        s->f_alt.p, s->id.p, s->id_alt.p );
    CMD_LAUNCH_CHECK();
    s->x.swap( s->x_alt );
    s->v.swap( s->v_alt );
    s->f.swap( s->f_alt );
    s->id.swap( s->id_alt );
    return CMD_OK;
}

// ---------------------------------------------------------------- thermo
__global__ void k_mv2_partials( int n, int cap, const double4 *__restrict__ x,
                                const double *__restrict__ v, MassTable mt,
                                double *__restrict__ partial )
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    double s = 0.0;
    if ( i < n )
    {
        const double vx = v[i], vy = v[cap + i], vz = v[2 * cap + i];
        s = ( vx * vx + vy * vy + vz * vz ) * mt.m[(int)x[i].w];
    }
    block_partials4( s, 0.0, 0.0, 0.0, partial );
}

static cmd_status sum_mv2( cmd_system *S, double *out )
{
    const int n = S->n_local;
    const int nb = nblocks_for( n );
    CMD_TRY( S->sc.dpart.reserve( 4 * (size_t)nb ) );
    CMD_TRY( S->sc.dres.reserve( 8 ) );
    k_mv2_partials<<<nb, TPB, 0, S->stream>>>( n, S->cap, S->x.p, S->v.p,
                                              S->mass_table(), S->sc.dpart.p );
    CMD_LAUNCH_CHECK();
    CMD_TRY( reduce_partials( S->sc.dpart.p, nb, S->sc.dres.p, S->stream ) );
    double h[4];
    CMD_HIP( hipMemcpyAsync( h, S->sc.dres.p, sizeof( h ),
                             hipMemcpyDeviceToHost, S->stream ) );
    CMD_HIP( hipStreamSynchronize( S->stream ) );
    *out = h[0];
    return CMD_OK;
}

// ------------------------------------------------ no-allocation thermo
// Scratch of cmd_thermo_kinetic_noalloc: four partial sums per 256-atom
// block, then the four reduced sums.  Each device holds its own copy for the
// life of the process (1 MiB + 32 B); host code reaches it only through the
// address hipGetSymbolAddress returns for the current device.
static constexpr int NOALLOC_MAX_BLOCKS = 32768; // 8,388,608 local atoms
static __device__ __align__( 256 ) double scratch[4 * NOALLOC_MAX_BLOCKS + 4];

static constexpr int MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static double *g_scratch_ptr[MAX_DEVICES] = {};

// The current device's copy of `scratch`, looked up once per device.
static cmd_status noalloc_scratch( double **out )
{
    int dev = 0;
    CMD_HIP( hipGetDevice( &dev ) );
    CMD_REQUIRE( dev >= 0 && dev < MAX_DEVICES, CMD_ERR_ARG, "device %d beyond %d", dev,
                 MAX_DEVICES );
    std::lock_guard<std::mutex> lock( g_scratch_mu );
    if ( !g_scratch_ptr[dev] )
    {
        void *ptr = nullptr;
        CMD_HIP( hipGetSymbolAddress( &ptr, HIP_SYMBOL( scratch ) ) );
        g_scratch_ptr[dev] = static_cast<double *>( ptr );
    }
    *out = g_scratch_ptr[dev];
    return CMD_OK;
}

// Streaming copy used to calibrate byte counters and achievable bandwidth.
__global__ void k_copy_probe( const double4 *__restrict__ a,
                              double4 *__restrict__ b, size_t n )
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for ( ; i < n; i += stride )
        b[i] = a[i];
}

} // namespace cmd

using namespace cmd;

// =================================================================== ABI
extern "C" {

const char *cmd_last_error( void ) { return g_err.c_str(); }

int cmd_version( void ) { return 1; }

cmd_status cmd_device_count( int *count )
{
    CMD_REQUIRE( count, CMD_ERR_ARG, "count is NULL" );
    int c = 0;
    hipError_t e = hipGetDeviceCount( &c );
    if ( e != hipSuccess )
        c = 0;
    *count = c;
    return CMD_OK;
}

cmd_status cmd_device_init( int device )
{
    int c = 0;
    if ( hipGetDeviceCount( &c ) != hipSuccess || c == 0 )
    {
        set_error( "no HIP device visible" );
        return CMD_ERR_NO_DEVICE;
    }
    CMD_REQUIRE( device >= 0 && device < c, CMD_ERR_ARG,
                 "device %d out of range (%d visible)", device, c );
    CMD_HIP( hipSetDevice( device ) );
    return CMD_OK;
}

cmd_status cmd_bandwidth_probe( int64_t bytes, int reps, double *gbs,
                                double *ms_per_copy )
{
    CMD_REQUIRE( bytes >= 32 && reps >= 1, CMD_ERR_ARG, "bad probe size" );
    const size_t n = (size_t)bytes / sizeof( double4 );
    DevBuf<double4> a, b;
    CMD_TRY( a.reserve( n ) );
    CMD_TRY( b.reserve( n ) );
    hipStream_t st;
    CMD_HIP( hipStreamCreateWithFlags( &st, hipStreamNonBlocking ) );
    const int grid = 256 * 8;
    k_fill<double><<<grid, TPB, 0, st>>>( (double *)a.p, 1.0, 4 * n );
    k_copy_probe<<<grid, TPB, 0, st>>>( a.p, b.p, n );
    hipEvent_t e0, e1;
    CMD_HIP( hipEventCreate( &e0 ) );
    CMD_HIP( hipEventCreate( &e1 ) );
    CMD_HIP( hipEventRecord( e0, st ) );
    for ( int r = 0; r < reps; ++r )
        k_copy_probe<<<grid, TPB, 0, st>>>( a.p, b.p, n );
    CMD_HIP( hipEventRecord( e1, st ) );
    CMD_HIP( hipEventSynchronize( e1 ) );
    float ms = 0.0f;
    CMD_HIP( hipEventElapsedTime( &ms, e0, e1 ) );
    (void)hipEventDestroy( e0 );
    (void)hipEventDestroy( e1 );
    (void)hipStreamDestroy( st );
    const double per = ms / reps;
    if ( ms_per_copy )
        *ms_per_copy = per;
    if ( gbs )
        *gbs = 2.0 * n * sizeof( double4 ) / ( per * 1e-3 ) / 1e9;
    return CMD_OK;
}

cmd_status cmd_device_synchronize( void )
{
    CMD_HIP( hipDeviceSynchronize() );
    return CMD_OK;
}

cmd_status cmd_domain_create( const double low[3], const double high[3],
                              double ghost_cutoff, cmd_domain *d )
{
    CMD_REQUIRE( low && high && d, CMD_ERR_ARG, "NULL argument" );
    for ( int k = 0; k < 3; ++k )
    {
        CMD_REQUIRE( high[k] > low[k], CMD_ERR_ARG,
                     "domain high <= low in dim %d", k );
        d->low[k] = low[k];
        d->high[k] = high[k];
        d->L[k] = high[k] - low[k];
        d->cell[k] = ( high[k] - low[k] ) / 100.0;
    }
    const double min_cell = std::min( std::min( d->cell[0], d->cell[1] ),
                                      d->cell[2] );
    d->halo_width = (int)std::ceil( ghost_cutoff / min_cell );
    for ( int k = 0; k < 3; ++k )
    {
        d->local_lo[k] = low[k] + d->cell[k] * 0.0;
        d->local_hi[k] = low[k] + d->cell[k] * 100.0;
        d->local_mesh[k] = d->local_hi[k] - d->local_lo[k];
        d->ghost_lo[k] = low[k] + d->cell[k] * ( 0.0 - d->halo_width );
        d->ghost_hi[k] = low[k] + d->cell[k] * ( 100.0 + d->halo_width );
    }
    return CMD_OK;
}

cmd_status cmd_system_create( int ntypes, cmd_system **out )
{
    CMD_REQUIRE( out, CMD_ERR_ARG, "out is NULL" );
    CMD_REQUIRE( ntypes >= 1 && ntypes <= CMD_MAX_TYPES, CMD_ERR_ARG,
                 "ntypes %d outside [1,%d]", ntypes, CMD_MAX_TYPES );
    int c = 0;
    if ( hipGetDeviceCount( &c ) != hipSuccess || c == 0 )
    {
        set_error( "no HIP device visible" );
        return CMD_ERR_NO_DEVICE;
    }
    cmd_system *s = new cmd_system;
    s->ntypes = ntypes;
    for ( int k = 0; k < CMD_MAX_TYPES; ++k )
        s->mass[k] = 1.0;
    hipError_t e = hipStreamCreateWithFlags( &s->stream, hipStreamNonBlocking );
    if ( e != hipSuccess )
    {
        delete s;
        set_error( "hipStreamCreate failed: %s", hipGetErrorString( e ) );
        return CMD_ERR_HIP;
    }
    s->own_stream = true;
    *out = s;
    return CMD_OK;
}

cmd_status cmd_system_destroy( cmd_system *s )
{
    if ( !s )
        return CMD_OK;
    if ( s->stream )
        (void)hipStreamSynchronize( s->stream );
    if ( s->own_stream && s->stream )
        (void)hipStreamDestroy( s->stream );
    delete s;
    return CMD_OK;
}

cmd_status cmd_system_set_stream( cmd_system *s, void *hip_stream )
{
    CMD_REQUIRE( s, CMD_ERR_ARG, "system is NULL" );
    CMD_HIP( hipStreamSynchronize( s->stream ) );
    if ( s->own_stream && s->stream )
        (void)hipStreamDestroy( s->stream );
    s->stream = (hipStream_t)hip_stream;
    s->own_stream = false;
    return CMD_OK;
}

cmd_status cmd_system_set_units( cmd_system *s, double dt, double mvv2e,
                                 double boltz )
{
    CMD_REQUIRE( s, CMD_ERR_ARG, "system is NULL" );
    CMD_REQUIRE( mvv2e > 0 && boltz > 0, CMD_ERR_ARG,
                 "mvv2e and boltz must be positive" );
    s->dt = dt;
    s->mvv2e = mvv2e;
    s->boltz = boltz;
    return CMD_OK;
}

cmd_status cmd_system_set_mass( cmd_system *s, int ntypes, const double *mass )
{
    CMD_REQUIRE( s && mass, CMD_ERR_ARG, "NULL argument" );
    CMD_REQUIRE( ntypes >= 1 && ntypes <= CMD_MAX_TYPES, CMD_ERR_ARG,
                 "ntypes %d outside [1,%d]", ntypes, CMD_MAX_TYPES );
    for ( int k = 0; k < ntypes; ++k )
    {
        CMD_REQUIRE( mass[k] > 0, CMD_ERR_ARG, "mass of type %d is not positive",
                     k + 1 );
        s->mass[k] = mass[k];
    }
    return CMD_OK;
}

cmd_status cmd_system_set_domain( cmd_system *s, const cmd_domain *d )
{
    CMD_REQUIRE( s && d, CMD_ERR_ARG, "NULL argument" );
    s->dom = *d;
    s->has_domain = true;
    return CMD_OK;
}

cmd_status cmd_system::reserve( int n )
{
    if ( n <= cap )
        return CMD_OK;
    const int newcap = std::max( n, cap + cap / 4 );
    const int used = n_local + n_ghost;
    // x, id, ghost maps: linear layout, copy the used prefix
    CMD_TRY( x.reserve( newcap, stream, true, used ) );
    CMD_TRY( x_alt.reserve( newcap ) );
    CMD_TRY( id.reserve( newcap, stream, true, used ) );
    CMD_TRY( id_alt.reserve( newcap ) );
    CMD_TRY( ghost_root.reserve( newcap, stream, true, n_ghost ) );
    CMD_TRY( ghost_code.reserve( newcap, stream, true, n_ghost ) );
    CMD_TRY( send_idx.reserve( newcap, stream, true, n_ghost ) );
    // planar v and f: copy plane by plane
    auto regrow_planes = [&]( DevBuf<double> &b ) -> cmd_status
    {
        DevBuf<double> nb;
        CMD_TRY( nb.reserve( 3 * (size_t)newcap ) );
        if ( b.p && used > 0 )
        {
            for ( int k = 0; k < 3; ++k )
                CMD_HIP( hipMemcpyAsync( nb.p + (size_t)k * newcap,
                                         b.p + (size_t)k * cap,
                                         sizeof( double ) * used,
                                         hipMemcpyDeviceToDevice, stream ) );
            CMD_HIP( hipStreamSynchronize( stream ) );
        }
        b.swap( nb );
        return CMD_OK;
    };
    CMD_TRY( regrow_planes( v ) );
    CMD_TRY( regrow_planes( f ) );
    v_alt.release();
    f_alt.release();
    CMD_TRY( v_alt.reserve( 3 * (size_t)newcap ) );
    CMD_TRY( f_alt.reserve( 3 * (size_t)newcap ) );
    cap = newcap;
    return CMD_OK;
}

cmd_status cmd_system_reserve( cmd_system *s, int n )
{
    CMD_REQUIRE( s && n >= 0, CMD_ERR_ARG, "bad argument" );
    return s->reserve( n );
}

cmd_status cmd_system_upload( cmd_system *s, int n_local, int n_ghost,
                              const double *x, const double *v,
                              const double *f, const int *type, const int *id )
{
    CMD_REQUIRE( s && x && type, CMD_ERR_ARG, "system, x and type are required" );
    CMD_REQUIRE( n_local >= 0 && n_ghost >= 0, CMD_ERR_ARG, "negative count" );
    const int n = n_local + n_ghost;
    for ( int i = 0; i < n; ++i )
        CMD_REQUIRE( type[i] >= 0 && type[i] < s->ntypes, CMD_ERR_ARG,
                     "atom %d has type %d outside [0,%d)", i, type[i],
                     s->ntypes );
    s->n_local = 0;
    s->n_ghost = 0;
    CMD_TRY( s->reserve( std::max( n, 1 ) ) );
    const int cap = s->cap;
    std::vector<double4> hx( std::max( n, 1 ) );
    std::vector<double> hv( 3 * (size_t)cap, 0.0 ), hf( 3 * (size_t)cap, 0.0 );
    std::vector<int> hid( std::max( n, 1 ) );
    for ( int i = 0; i < n; ++i )
    {
        hx[i] = make_double4( x[3 * i], x[3 * i + 1], x[3 * i + 2],
                              (double)type[i] );
        for ( int k = 0; k < 3; ++k )
        {
            hv[(size_t)k * cap + i] = v ? v[3 * i + k] : 0.0;
            hf[(size_t)k * cap + i] = f ? f[3 * i + k] : 0.0;
        }
        hid[i] = id ? id[i] : i + 1;
    }
    if ( n > 0 )
    {
        CMD_HIP( hipMemcpyAsync( s->x.p, hx.data(), sizeof( double4 ) * n,
                                 hipMemcpyHostToDevice, s->stream ) );
        CMD_HIP( hipMemcpyAsync( s->id.p, hid.data(), sizeof( int ) * n,
                                 hipMemcpyHostToDevice, s->stream ) );
    }
    CMD_HIP( hipMemcpyAsync( s->v.p, hv.data(), sizeof( double ) * 3 * cap,
                             hipMemcpyHostToDevice, s->stream ) );
    CMD_HIP( hipMemcpyAsync( s->f.p, hf.data(), sizeof( double ) * 3 * cap,
                             hipMemcpyHostToDevice, s->stream ) );
    CMD_HIP( hipStreamSynchronize( s->stream ) );
    s->n_local = n_local;
    s->n_ghost = n_ghost;
    s->N = n_local;
    s->bin_grid_valid = false;
    s->xg_valid = false;
    for ( int k = 0; k < 6; ++k )
        s->counts[k] = 0;
    return CMD_OK;
}

cmd_status cmd_system_download( cmd_system *s, double *x, double *v, double *f,
                                int *type, int *id )
{
    CMD_REQUIRE( s, CMD_ERR_ARG, "system is NULL" );
    const int n = s->n_local + s->n_ghost;
    const int cap = s->cap;
    if ( n == 0 )
        return CMD_OK;
    if ( x || type )
    {
        std::vector<double4> hx( n );
        CMD_HIP( hipMemcpyAsync( hx.data(), s->x.p, sizeof( double4 ) * n,
                                 hipMemcpyDeviceToHost, s->stream ) );
        CMD_HIP( hipStreamSynchronize( s->stream ) );
        for ( int i = 0; i < n; ++i )
        {
            if ( x )
            {
                x[3 * i] = hx[i].x;
                x[3 * i + 1] = hx[i].y;
                x[3 * i + 2] = hx[i].z;
            }
            if ( type )
                type[i] = (int)hx[i].w;
        }
    }
    auto planes = [&]( const double *d, double *h ) -> cmd_status
    {
        std::vector<double> t( 3 * (size_t)cap );
        CMD_HIP( hipMemcpyAsync( t.data(), d, sizeof( double ) * 3 * cap,
                                 hipMemcpyDeviceToHost, s->stream ) );
        CMD_HIP( hipStreamSynchronize( s->stream ) );
        for ( int i = 0; i < n; ++i )
            for ( int k = 0; k < 3; ++k )
                h[3 * i + k] = t[(size_t)k * cap + i];
        return CMD_OK;
    };
    if ( v )
        CMD_TRY( planes( s->v.p, v ) );
    if ( f )
        CMD_TRY( planes( s->f.p, f ) );
    if ( id )
    {
        CMD_HIP( hipMemcpyAsync( id, s->id.p, sizeof( int ) * n,
                                 hipMemcpyDeviceToHost, s->stream ) );
        CMD_HIP( hipStreamSynchronize( s->stream ) );
    }
    return CMD_OK;
}

cmd_status cmd_system_counts( const cmd_system *s, int *n_global, int *n_local,
                              int *n_ghost )
{
    CMD_REQUIRE( s, CMD_ERR_ARG, "system is NULL" );
    if ( n_global )
        *n_global = s->N;
    if ( n_local )
        *n_local = s->n_local;
    if ( n_ghost )
        *n_ghost = s->n_ghost;
    return CMD_OK;
}

cmd_status cmd_system_footprint( const cmd_system *s, int64_t *bytes )
{
    CMD_REQUIRE( s && bytes, CMD_ERR_ARG, "NULL argument" );
    // x, x_alt (32 B each), v, v_alt, f, f_alt (24 B each), id, id_alt,
    // ghost maps, flags, positions (4 B each)
    *bytes = (int64_t)s->cap * ( 2 * 32 + 4 * 24 + 7 * 4 );
    return CMD_OK;
}

cmd_status cmd_nve_initial( cmd_system *s )
{
    CMD_REQUIRE( s, CMD_ERR_ARG, "system is NULL" );
    if ( s->n_local == 0 )
        return CMD_OK;
    const double dtf = 0.5 * s->dt / s->mvv2e;
    s->bin_grid_valid = false; // atoms leave the cells they were sorted into
    k_nve_initial<<<nblocks_for( s->n_local ), TPB, 0, s->stream>>>(
        s->n_local, s->cap, s->x.p, s->v.p, s->f.p, s->mass_table(), dtf,
        s->dt );
    CMD_LAUNCH_CHECK();
    return CMD_OK;
}

cmd_status cmd_nve_final( cmd_system *s )
{
    CMD_REQUIRE( s, CMD_ERR_ARG, "system is NULL" );
    if ( s->n_local == 0 )
        return CMD_OK;
    const double dtf = 0.5 * s->dt / s->mvv2e;
    k_nve_final<<<nblocks_for( s->n_local ), TPB, 0, s->stream>>>(
        s->n_local, s->cap, s->x.p, s->v.p, s->f.p, s->mass_table(), dtf );
    CMD_LAUNCH_CHECK();
    return CMD_OK;
}

cmd_status cmd_comm_exchange( cmd_system *s )
{
    CMD_REQUIRE( s, CMD_ERR_ARG, "system is NULL" );
    CMD_REQUIRE( s->has_domain, CMD_ERR_STATE, "system has no domain" );
    s->n_ghost = 0; // Comm::exchange resizes to N_local (comm_mpi_impl.h:270-271)
    s->bin_grid_valid = false;
    s->xg_valid = false;
    return wrap_local( s );
}

cmd_status cmd_comm_exchange_halo( cmd_system *s, double comm_depth )
{
    CMD_REQUIRE( s, CMD_ERR_ARG, "system is NULL" );
    CMD_REQUIRE( s->has_domain, CMD_ERR_STATE, "system has no domain" );
    s->xg_valid = false;
    const int n = s->n_local;
    // totals counted before the sort of these positions (halo_precount)
    const bool pre = s->halo_pre && s->halo_pre_n == n && s->halo_pre_depth == comm_depth;
    s->halo_pre = false;
    s->n_ghost = 0;
    for ( int k = 0; k < 6; ++k )
        s->counts[k] = 0;
    if ( n == 0 )
        return CMD_OK;
    const HaloFaces H = halo_faces( s, comm_depth );
    const int nblk = nblocks_for( n );
    const size_t nc = (size_t)HALO_SETS * nblk;
    // counts, their scan (nc + 1 entries: the scan appends the total), the
    // totals and the atoms' phase masks (kept apart from the system's
    // arrays, which may be regrown below)
    CMD_TRY( s->halo_blk.reserve( 2 * nc + 1 + HALO_SETS + n ) );
    int *cnt = s->halo_blk.p, *off = cnt + nc, *tot = off + nc + 1, *mask = tot + HALO_SETS;
    if ( !pre )
        CMD_TRY( fill_i32( tot, 0, HALO_SETS, s->stream ) );
    k_halo_count<true><<<nblk, TPB, 0, s->stream>>>( n, s->x.p, H, mask, nblk, cnt,
                                                     pre ? nullptr : tot );
    CMD_LAUNCH_CHECK();
    CMD_TRY( scan_exclusive( cnt, off, (int)nc, s->sc, s->stream, nullptr ) );
    int th[HALO_SETS];
    if ( pre )
    {
        // the count before the sort is long done by now; the sorted pass
        // above stays queued behind the sort
        CMD_HIP( hipEventSynchronize( s->hw.ev ) );
        for ( int j = 0; j < HALO_SETS; ++j )
            th[j] = s->hw.p[HW_HALO + j];
        s->hw.p[HW_HALO_FLAG] = 0;
    }
    else
    {
        CMD_HIP( hipMemcpyAsync( th, tot, sizeof( th ), hipMemcpyDeviceToHost, s->stream ) );
        CMD_HIP( hipStreamSynchronize( s->stream ) );
    }
    int ng = 0;
    for ( int j = 0; j < HALO_SETS; ++j )
    {
        s->counts[halo_phase( j )] += th[j];
        ng += th[j];
    }
    if ( n + ng > s->cap )
        CMD_TRY( s->reserve( (int)( ( n + ng ) * 1.1 ) + 64 ) );
    // the halo gathers positions and types only (comm_mpi_impl.h:346-347);
    // the ghosts' velocity and force rows are set to zero so that downloads
    // are defined and full-list kernels, which never write them, leave them 0
    if ( ng > 0 || pre )
    {
        k_halo_fill<<<nblk, TPB, 0, s->stream>>>(
            n, s->cap, s->x.p, s->v.p, s->f.p, s->id.p, mask, nblk, off, H, s->send_idx.p,
            s->ghost_root.p, s->ghost_code.p, pre ? s->halo_pre_tot.p : nullptr, ng,
            pre ? s->hw.d + HW_HALO_FLAG : nullptr );
        CMD_LAUNCH_CHECK();
    }
    if ( pre )
    {
        CMD_HIP( hipEventRecord( s->hw.ev, s->stream ) );
        s->halo_pre_check = true;
    }
    s->n_ghost = ng;
    return CMD_OK;
}

cmd_status cmd_comm_update_halo( cmd_system *s )
{
    CMD_REQUIRE( s, CMD_ERR_ARG, "system is NULL" );
    return s->xg_valid ? tiles_ghost_refresh( s ) : ghost_refresh( s );
}

cmd_status cmd_comm_update_force( cmd_system *s )
{
    CMD_REQUIRE( s, CMD_ERR_ARG, "system is NULL" );
    if ( s->n_ghost == 0 )
        return CMD_OK;
    k_ghost_fold<<<nblocks_for( s->n_ghost ), TPB, 0, s->stream>>>(
        s->n_local, s->n_ghost, s->cap, s->f.p, s->ghost_root.p );
    CMD_LAUNCH_CHECK();
    return CMD_OK;
}

cmd_status cmd_comm_ghost_lists( cmd_system *s, int *send_idx, int counts[6] )
{
    CMD_REQUIRE( s && counts, CMD_ERR_ARG, "NULL argument" );
    int tot = 0;
    for ( int k = 0; k < 6; ++k )
    {
        counts[k] = s->counts[k];
        tot += s->counts[k];
    }
    if ( send_idx && tot > 0 )
    {
        CMD_HIP( hipMemcpyAsync( send_idx, s->send_idx.p, sizeof( int ) * tot,
                                 hipMemcpyDeviceToHost, s->stream ) );
        CMD_HIP( hipStreamSynchronize( s->stream ) );
    }
    return CMD_OK;
}

// Binning::create_binning (binning_cabana_impl.h:57-113) with the Cabana
// LinkedCellList grid (floor, top-edge clamp, cardinal (i*ny + j)*nz + k).
cmd_status cmd_binning_create( cmd_system *s, double dx_in, double dy_in,
                               double dz_in, int halo_depth, int do_local,
                               int do_ghost, int sort, int *perm_out )
{
    return binning_create( s, dx_in, dy_in, dz_in, halo_depth, do_local, do_ghost, sort,
                           perm_out, true );
}

} // extern "C"

// The same, with the forces left out of the permutation when keep_f is
// false: the time loop recomputes them after every list rebuild before
// anything reads them, so their rows are left in the old order.
cmd_status cmd::binning_create( cmd_system *s, double dx_in, double dy_in, double dz_in,
                                int halo_depth, int do_local, int do_ghost, int sort,
                                int *perm_out, bool keep_f )
{
    CMD_REQUIRE( s, CMD_ERR_ARG, "system is NULL" );
    CMD_REQUIRE( s->has_domain, CMD_ERR_STATE, "system has no domain" );
    CMD_REQUIRE( dx_in > 0 && dy_in > 0 && dz_in > 0, CMD_ERR_ARG,
                 "bin sizes must be positive" );
    if ( !( do_local || do_ghost ) )
        return CMD_OK;
    CMD_REQUIRE( do_local && !do_ghost, CMD_ERR_ARG,
                 "only local-atom binning (do_local=1, do_ghost=0) is supported, "
                 "as used by CbnMD" );
    const double din[3] = { dx_in, dy_in, dz_in };
    int nbin[3];
    double delta[3];
    for ( int k = 0; k < 3; ++k )
    {
        nbin[k] = (int)( s->dom.local_mesh[k] / din[k] );
        if ( nbin[k] == 0 )
            nbin[k] = 1;
        delta[k] = s->dom.local_mesh[k] / nbin[k];
    }
    const double eps = delta[0] / 1000;
    CellGrid g;
    g.floor_mode = 1;
    for ( int k = 0; k < 3; ++k )
    {
        const double mn = -delta[k] * halo_depth - eps + s->dom.local_lo[k];
        const double mx = delta[k] * halo_depth + eps + s->dom.local_hi[k];
        g.min[k] = mn;
        g.n[k] = (int)std::floor( ( mx - mn ) * ( 1.0 / delta[k] ) );
        CMD_REQUIRE( g.n[k] >= 1, CMD_ERR_ARG, "empty bin grid" );
        const double c = ( mx - mn ) / g.n[k];
        g.rd[k] = 1.0 / c;
    }
    const int n = s->n_local;
    CMD_TRY( cell_sort( s->x.p, n, g, s->bins, s->sc, s->stream ) );
    if ( sort && n > 0 )
    {
        const int cap = s->cap;
        k_permute<<<nblocks_for( n ), TPB, 0, s->stream>>>(
            n, cap, s->bins.perm.p, s->x.p, s->x_alt.p, s->v.p, s->v_alt.p,
            keep_f ? s->f.p : nullptr, s->f_alt.p, s->id.p, s->id_alt.p );
        CMD_LAUNCH_CHECK();
        s->x.swap( s->x_alt );
        s->v.swap( s->v_alt );
        if ( keep_f )
            s->f.swap( s->f_alt );
        s->id.swap( s->id_alt );
        s->n_ghost = 0; // ghosts are not part of the permuted range
        s->bin_grid = g;
        s->bin_grid_valid = true;
        s->xg_valid = false;
    }
    if ( perm_out && n > 0 )
    {
        CMD_HIP( hipMemcpyAsync( perm_out, s->bins.perm.p, sizeof( int ) * n,
                                 hipMemcpyDeviceToHost, s->stream ) );
        CMD_HIP( hipStreamSynchronize( s->stream ) );
    }
    return CMD_OK;
}

extern "C" {

cmd_status cmd_thermo_temperature( cmd_system *s, double *T )
{
    CMD_REQUIRE( s && T, CMD_ERR_ARG, "NULL argument" );
    double mv2 = 0.0;
    CMD_TRY( sum_mv2( s, &mv2 ) );
    const int dof = 3 * s->N - 3;
    const double factor = s->mvv2e / ( 1.0 * dof * s->boltz );
    *T = mv2 * factor;
    return CMD_OK;
}

cmd_status cmd_thermo_kinetic( cmd_system *s, double *ke )
{
    CMD_REQUIRE( s && ke, CMD_ERR_ARG, "NULL argument" );
    double mv2 = 0.0;
    CMD_TRY( sum_mv2( s, &mv2 ) );
    *ke = mv2 * ( 0.5 * s->mvv2e );
    return CMD_OK;
}

// cmd_thermo_kinetic with its partial sums in the file-scope `scratch`
// array instead of the system's lazily grown buffers: the same kernels in
// the same order (identical bits), no allocation.  Every slot the reduction
// reads is written first by the same call, so nothing needs clearing; calls
// on one device share the array, so overlapping calls race.
cmd_status cmd_thermo_kinetic_noalloc( cmd_system *s, double *ke )
{
    CMD_REQUIRE( s && ke, CMD_ERR_ARG, "NULL argument" );
    const int n = s->n_local;
    const int nb = nblocks_for( n );
    CMD_REQUIRE( nb <= NOALLOC_MAX_BLOCKS, CMD_ERR_CAPACITY,
                 "cmd_thermo_kinetic_noalloc takes at most %d local atoms (%d given)",
                 NOALLOC_MAX_BLOCKS * TPB, n );
    double *scr = nullptr;
    CMD_TRY( noalloc_scratch( &scr ) );
    double *res = scr + 4 * (size_t)NOALLOC_MAX_BLOCKS;
    k_mv2_partials<<<nb, TPB, 0, s->stream>>>( n, s->cap, s->x.p, s->v.p, s->mass_table(),
                                              scr );
    CMD_LAUNCH_CHECK();
    CMD_TRY( reduce_partials( scr, nb, res, s->stream ) );
    double h[4];
    CMD_HIP( hipMemcpyAsync( h, res, sizeof( h ), hipMemcpyDeviceToHost, s->stream ) );
    CMD_HIP( hipStreamSynchronize( s->stream ) );
    *ke = h[0] * ( 0.5 * s->mvv2e );
    return CMD_OK;
}

} // extern "C"
