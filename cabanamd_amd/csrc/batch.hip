// batch.hip -- many independent LJ systems stepped together.
//
// CabanaMD runs one system per process (CbnMD::run, cabanamd_impl.h:285-399).
// Here a batch holds any number of independent periodic boxes of uneven size
// and steps them with the same launches: their owned atoms are concatenated
// into one flat range [0, N) (system k owns [a0_k, a0_k + n_k)), their ghost
// images follow at [N, N + G), and every per-atom kernel runs over the flat
// index.  Only the rebuild and the thermo reductions need to know the
// system of an atom; they read it from a per-atom system index and a table
// of per-system boxes and cell grids.  Each system follows the reference's
// step order exactly:
//   initial_integrate; rebuild (wrap, binning sort, 6-face ghost images,
//   Verlet list) every `every` steps, else ghost refresh; force;
//   final_integrate; thermo.
// The force, both half-kicks and the drift run in the fused LJ kernel of the
// single-system engine (lj.hip) on the flat range, so one launch advances
// every system by a step.
//
// Independence: a system's atoms, ghosts and list rows depend only on that
// system (cells belong to one system, sorts are stable by flat index, ghost
// images are emitted per atom in a fixed direction order, reductions run one
// block per system in a fixed order), so its trajectory is bit-identical
// whether it runs alone or beside others.  Systems join and leave only at a
// rebuild boundary.
#include "internal.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <vector>

namespace cmd
{
static constexpr int TPB = 256;

static inline int nblk( int64_t n ) { return (int)std::max<int64_t>( 1, ( n + TPB - 1 ) / TPB ); }

// One system of the batch as the kernels see it.
struct BSys
{
    double lo[3], hi[3], L[3];
    double gmin[3], grd[3]; // cell grid over the box plus the ghost shell
    int gn[3];
    int cell_base;          // first cell of this system in the batch's cells
    int a0, na;             // owned atoms in the flat range
    int pad;
};

__device__ __forceinline__ int bcell( const BSys &s, double4 p )
{
    const int a = cell_coord( p.x, s.gmin[0], s.grd[0], s.gn[0], 1 );
    const int b = cell_coord( p.y, s.gmin[1], s.grd[1], s.gn[1], 1 );
    const int c = cell_coord( p.z, s.gmin[2], s.grd[2], s.gn[2], 1 );
    return s.cell_base + ( a * s.gn[1] + b ) * s.gn[2] + c;
}

// comm_mpi.h:141-168 per system: x > hi -> x - L, x < lo -> x + L, both
// tested on the value before the update (lo = 0 gives the reference's form).
__global__ void k_b_wrap( int n, double4 *__restrict__ x, const int *__restrict__ sys,
                          const BSys *__restrict__ T )
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if ( i >= n )
        return;
    const BSys &s = T[sys[i]];
    double4 p = x[i];
    double c[3] = { p.x, p.y, p.z };
    for ( int d = 0; d < 3; ++d )
    {
        const double q = c[d];
        if ( q > s.hi[d] )
            c[d] = q - s.L[d];
        if ( q < s.lo[d] )
            c[d] = q + s.L[d];
    }
    p.x = c[0];
    p.y = c[1];
    p.z = c[2];
    x[i] = p;
}

__global__ void k_b_cell_local( int n, const double4 *__restrict__ x,
                                const int *__restrict__ sys, const BSys *__restrict__ T,
                                int *__restrict__ cell_of )
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if ( i >= n )
        return;
    cell_of[i] = bcell( T[sys[i]], x[i] );
}

__global__ void k_b_cell_all( int n_local, int n_total, const double4 *__restrict__ x,
                              const int *__restrict__ sys, const int *__restrict__ root,
                              const BSys *__restrict__ T, int *__restrict__ cell_of )
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if ( i >= n_total )
        return;
    const int k = i < n_local ? sys[i] : sys[root[i - n_local]];
    cell_of[i] = bcell( T[k], x[i] );
}

// Shift choices per dimension of an owned atom: 0 always, -L when within
// depth of the high face (phase 2d of TagHaloPack, comm_mpi.h:240-319),
// +L when within depth of the low face (phase 2d+1).  The ghosts of an atom
// are every combination except no shift at all -- the set the six
// sequential phases produce.
__device__ __forceinline__ void shift_sets( const BSys &s, double4 p, double depth,
                                            int nopt[3], int opt[3][3] )
{
    const double c[3] = { p.x, p.y, p.z };
    for ( int d = 0; d < 3; ++d )
    {
        int k = 0;
        opt[d][k++] = 0;
        if ( c[d] >= s.hi[d] - depth )
            opt[d][k++] = 1 << ( 2 * d );
        if ( c[d] <= s.lo[d] + depth )
            opt[d][k++] = 1 << ( 2 * d + 1 );
        nopt[d] = k;
    }
}

__global__ void k_b_ghost_count( int n, const double4 *__restrict__ x,
                                 const int *__restrict__ sys, const BSys *__restrict__ T,
                                 double depth, int *__restrict__ cnt )
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if ( i >= n )
        return;
    int nopt[3], opt[3][3];
    shift_sets( T[sys[i]], x[i], depth, nopt, opt );
    cnt[i] = nopt[0] * nopt[1] * nopt[2] - 1;
}

__device__ __forceinline__ double4 apply_code( double4 p, int c, const double L[3] )
{
    if ( c & 1 )
        p.x = p.x - L[0];
    if ( c & 2 )
        p.x = p.x + L[0];
    if ( c & 4 )
        p.y = p.y - L[1];
    if ( c & 8 )
        p.y = p.y + L[1];
    if ( c & 16 )
        p.z = p.z - L[2];
    if ( c & 32 )
        p.z = p.z + L[2];
    return p;
}

// Ghosts of atom i at ofs[i] .. in (z, y, x) shift-option order.
__global__ void k_b_ghost_fill( int n, double4 *__restrict__ x, int *__restrict__ id,
                                const int *__restrict__ sys, const BSys *__restrict__ T,
                                double depth, const int *__restrict__ ofs,
                                int *__restrict__ root, int *__restrict__ code )
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if ( i >= n )
        return;
    const BSys &s = T[sys[i]];
    const double4 p = x[i];
    int nopt[3], opt[3][3];
    shift_sets( s, p, depth, nopt, opt );
    int g = ofs[i];
    for ( int c = 0; c < nopt[2]; ++c )
        for ( int b = 0; b < nopt[1]; ++b )
            for ( int a = 0; a < nopt[0]; ++a )
            {
                const int cd = opt[0][a] | opt[1][b] | opt[2][c];
                if ( cd == 0 )
                    continue;
                x[n + g] = apply_code( p, cd, s.L );
                id[n + g] = id[i];
                root[g] = i;
                code[g] = cd;
                ++g;
            }
}

// update_halo (comm_mpi_impl.h:369-408): ghost = root shifted by its code.
__global__ void k_b_ghost_refresh( int n_local, int n_ghost, double4 *__restrict__ x,
                                   const int *__restrict__ sys, const int *__restrict__ root,
                                   const int *__restrict__ code, const BSys *__restrict__ T )
{
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if ( g >= n_ghost )
        return;
    const int r = root[g];
    x[n_local + g] = apply_code( x[r], code[g], T[sys[r]].L );
}

__global__ void k_b_gather( int n, const int *__restrict__ perm, const double4 *__restrict__ x,
                            double4 *__restrict__ xs )
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if ( k >= n )
        return;
    const int i = perm[k];
    double4 p = x[i];
    p.w = (double)i;
    xs[k] = p;
}

// Full Verlet list of the owned atoms over their system's 27-cell stencil
// (neighbor_verlet.h:43-62; accept j != i and rsq <= rc^2), written in the
// lane-interleaved slab layout of the single-system engine (neighbor.hip):
// entry k of atom a of wave w at ((w R + k / G) 64) + (k % G) APW + a.
// One wave per cell (cells of every system in one space): the cell's owned
// atoms share its stencil, which is at most 9 contiguous runs of the
// cell-sorted candidates (one per column, cells e-1 .. e+1, columns in
// (x, y) order).  The wave loads the candidates once into registers, CHB
// chunks of 64 at a time (lane = candidate), and tests the cell's atoms one
// by one (position and row length wave-uniform); ballot + mbcnt place the
// accepted candidates in each row in walk order.
static constexpr int CHB = 8;

template <int G>
__global__ __launch_bounds__( 256 ) void k_b_neigh_cell(
    int ncell, const double4 *__restrict__ x, const double4 *__restrict__ xs,
    const int *__restrict__ lstart, const int *__restrict__ astart,
    const int *__restrict__ sys, const BSys *__restrict__ T, double rc2, int cap,
    int *__restrict__ counts, int *__restrict__ slab, int *__restrict__ max_row )
{
    constexpr int APW = 64 / G;
    const int lane = threadIdx.x & 63;
    const int c = blockIdx.x * ( blockDim.x >> 6 ) + ( threadIdx.x >> 6 );
    int mrow = 0;
    const int c0 = c < ncell ? __builtin_amdgcn_readfirstlane( lstart[c] ) : 0;
    const int c1 = c < ncell ? __builtin_amdgcn_readfirstlane( lstart[c + 1] ) : 0;
    if ( c1 > c0 )
    {
        const BSys &s = T[__builtin_amdgcn_readfirstlane( sys[c0] )];
        const int lc = c - s.cell_base;
        const int a = lc / ( s.gn[1] * s.gn[2] ), b = ( lc / s.gn[2] ) % s.gn[1],
                  e = lc % s.gn[2];
        const int e0 = e > 0 ? e - 1 : 0, e1 = e < s.gn[2] - 1 ? e + 1 : s.gn[2] - 1;
        // column q = lane (< 9): its run [rs, rs + len) of candidates
        int rs = 0, len = 0;
        if ( lane < 9 )
        {
            const int aa = a + lane / 3 - 1, bb = b + lane % 3 - 1;
            if ( aa >= 0 && aa < s.gn[0] && bb >= 0 && bb < s.gn[1] )
            {
                const int base = s.cell_base + ( aa * s.gn[1] + bb ) * s.gn[2];
                rs = astart[base + e0];
                len = astart[base + e1 + 1] - rs;
            }
        }
        int inc = len;
#pragma unroll
        for ( int o = 1; o < 16; o <<= 1 )
        {
            const int u = __shfl_up( inc, o, 64 );
            if ( lane >= o )
                inc += u;
        }
        // run q starts at candidate P[q] of the walk; candidate v of run q is
        // xs[v + D[q]] with D[q] = rs_q - P[q]
        int P[9], D[9];
#pragma unroll
        for ( int q = 0; q < 9; ++q )
        {
            const int iq = __builtin_amdgcn_readlane( inc, q );
            const int lq = __builtin_amdgcn_readlane( len, q );
            P[q] = iq - lq;
            D[q] = __builtin_amdgcn_readlane( rs, q ) - P[q];
        }
        const int total = __builtin_amdgcn_readlane( inc, 8 );
        const int R = ( cap + G - 1 ) / G;
        for ( int a0 = c0; a0 < c1; a0 += 64 )
        {
            const int na = min( 64, c1 - a0 );
            int cnt = 0; // row length of atom a0 + lane
            for ( int v0 = 0; v0 < total; v0 += CHB * 64 )
            {
                double px[CHB], py[CHB], pz[CHB];
                int pj[CHB];
#pragma unroll
                for ( int k = 0; k < CHB; ++k )
                {
                    const int v = v0 + k * 64 + lane;
                    px[k] = 1.0e30;
                    py[k] = 1.0e30;
                    pz[k] = 1.0e30;
                    pj[k] = -1;
                    if ( v < total )
                    {
                        int d = D[0];
#pragma unroll
                        for ( int q = 1; q < 9; ++q )
                            d = v >= P[q] ? D[q] : d;
                        const double4 p = xs[v + d];
                        px[k] = p.x;
                        py[k] = p.y;
                        pz[k] = p.z;
                        pj[k] = (int)p.w;
                    }
                }
                const int nch = min( CHB, ( total - v0 + 63 ) >> 6 );
                double4 pn = x[a0];
                for ( int l = 0; l < na; ++l )
                {
                    const double4 pa = pn;
                    pn = x[a0 + min( l + 1, na - 1 )];
                    const int i = a0 + l;
                    int *out = slab + (size_t)( i / APW ) * R * 64 + i % APW;
                    unsigned k0 = (unsigned)__builtin_amdgcn_readlane( cnt, l );
#pragma unroll
                    for ( int k = 0; k < CHB; ++k )
                    {
                        if ( k < nch )
                        {
                            const double dx = pa.x - px[k];
                            const double dy = pa.y - py[k];
                            const double dz = pa.z - pz[k];
                            const double rsq = dx * dx + dy * dy + dz * dz;
                            const unsigned long long m =
                                __builtin_amdgcn_ballot_w64( rsq <= rc2 ) &
                                __builtin_amdgcn_ballot_w64( pj[k] != i );
                            const unsigned kk = __builtin_amdgcn_mbcnt_hi(
                                (unsigned)( m >> 32 ),
                                __builtin_amdgcn_mbcnt_lo( (unsigned)m, k0 ) );
                            if ( __builtin_amdgcn_inverse_ballot_w64( m ) && kk < (unsigned)cap )
                                out[( kk / G ) * 64 + ( kk % G ) * APW] = pj[k];
                            k0 += (unsigned)__popcll( m );
                        }
                    }
                    if ( lane == l )
                        cnt = (int)k0;
                }
            }
            if ( lane < na )
            {
                counts[a0 + lane] = cnt;
                mrow = max( mrow, cnt );
            }
        }
    }
#pragma unroll
    for ( int o = 32; o > 0; o >>= 1 )
        mrow = max( mrow, __shfl_xor( mrow, o, 64 ) );
    if ( lane == 0 && mrow > 0 )
        atomicMax( max_row, mrow );
}

// Per-atom thermo terms in FP64: pe_i = 1/2 sum_j shifted LJ over the full
// list (ForceLJ::compute_energy_full, force_lj_cabana_neigh_impl.h:261-315)
// and m v^2 (property_kine_impl.h:55-76).
template <int G>
__global__ void k_b_atom_thermo( int n_local, int scap, int cap, const double4 *__restrict__ x,
                                 const double *__restrict__ v, const int *__restrict__ counts,
                                 const int *__restrict__ slab, const LJParams *__restrict__ P,
                                 MassTable mt, double *__restrict__ pe, double *__restrict__ mv2 )
{
    constexpr int APW = 64 / G;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if ( i >= n_local )
        return;
    const double4 pi = x[i];
    const int ti = (int)pi.w;
    const int R = ( cap + G - 1 ) / G;
    const int *row = slab + (size_t)( i / APW ) * R * 64 + i % APW;
    const int n = counts[i];
    double e = 0.0;
    for ( int k = 0; k < n; ++k )
    {
        const double4 pj = x[row[( k / G ) * 64 + ( k % G ) * APW]];
        const double dx = pi.x - pj.x;
        const double dy = pi.y - pj.y;
        const double dz = pi.z - pj.z;
        const double rsq = dx * dx + dy * dy + dz * dz;
        const int t = ti * P->ntypes + (int)pj.w;
        if ( rsq < P->cutsq[t] )
        {
            const double r2inv = 1.0 / rsq;
            const double r6inv = r2inv * r2inv * r2inv;
            e += 0.5 * ( r6inv * ( 0.5 * P->lj1[t] * r6inv - P->lj2[t] ) / 6.0 - P->eshift[t] );
        }
    }
    pe[i] = e;
    const double m = mt.m[ti];
    const double vx = v[i], vy = v[scap + i], vz = v[2 * scap + i];
    mv2[i] = ( vx * vx + vy * vy + vz * vz ) * m;
}

__global__ void k_b_mv2( int n_local, int scap, const double4 *__restrict__ x,
                         const double *__restrict__ v, MassTable mt, double *__restrict__ mv2 )
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if ( i >= n_local )
        return;
    const double m = mt.m[(int)x[i].w];
    const double vx = v[i], vy = v[scap + i], vz = v[2 * scap + i];
    mv2[i] = ( vx * vx + vy * vy + vz * vz ) * m;
}

// One block per system: sums of pe and m v^2 over its atoms, in an order
// fixed by the system alone.
__global__ void k_b_sys_sum( const BSys *__restrict__ T, const double *__restrict__ pe,
                             const double *__restrict__ mv2, double *__restrict__ out )
{
    const BSys &s = T[blockIdx.x];
    double a = 0.0, b = 0.0;
    for ( int k = threadIdx.x; k < s.na; k += TPB )
    {
        a += pe[s.a0 + k];
        b += mv2[s.a0 + k];
    }
    __shared__ double ra[TPB], rb[TPB];
    ra[threadIdx.x] = a;
    rb[threadIdx.x] = b;
    __syncthreads();
    for ( int o = TPB / 2; o > 0; o >>= 1 )
    {
        if ( (int)threadIdx.x < o )
        {
            ra[threadIdx.x] += ra[threadIdx.x + o];
            rb[threadIdx.x] += rb[threadIdx.x + o];
        }
        __syncthreads();
    }
    if ( threadIdx.x == 0 )
    {
        out[2 * blockIdx.x] = ra[0];
        out[2 * blockIdx.x + 1] = rb[0];
    }
}

__global__ void k_b_fill_sys( int a0, int n, int k, int *__restrict__ sys )
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if ( i < n )
        sys[a0 + i] = k;
}

// Host copies of one new system, waiting to join.
struct BPending
{
    int handle;
    std::vector<double4> x;
    std::vector<double> v; // 3 planes of n
    std::vector<int> id;
    double lo[3], hi[3];
};

struct BMember
{
    int handle;
    int n;
    double lo[3], hi[3];
    int join_step;
    int a0;
    std::vector<double> thermo; // rows: local step, T, PE/N, (PE+KE)/N
};
} // namespace cmd

using namespace cmd;

struct cmd_batch
{
    cmd_params p{};
    cmd_system *S = nullptr;
    cmd_neigh *N = nullptr;
    cmd_force *F = nullptr;
    hipStream_t stream = nullptr;
    bool nnp = false; // NNP systems (force_type 2): unfused step order
    int prec = CMD_PREC_FP64;
    int lanes = 0; // 0: autotune at the first build
    int step = 0;
    int next_handle = 1;
    std::vector<BMember> members;
    std::vector<BPending> pending;
    std::vector<int> leaving;
    std::vector<BSys> htab;
    DevBuf<BSys> tab;
    DevBuf<int> sys;
    CellSort lsort, asort;
    DevBuf<int> gcnt, gofs, max_row;
    DevBuf<double4> xs;
    DevBuf<double> pe, mv2, sums;
    DevBuf<LJParams> Pd;
    DevBuf<double> partial;
    std::vector<double> hsums;
    int ncell = 0;
    double timers[4] = { 0, 0, 0, 0 }; // rebuild, steps, thermo, join/leave
    // Two plain LJ steps (ghost refresh + fused step, into x_alt and back)
    // captured as one HIP graph, one per position buffer the pair starts
    // from (an odd number of eager steps between replays swaps them); a
    // graph is valid while its epoch (bumped at every rebuild, where
    // buffers, counts and lists change) is the batch's.
    struct PairGraph
    {
        hipGraphExec_t exec = nullptr;
        const void *x = nullptr; // S->x.p at capture
        int epoch = -1;
    };
    bool graphs = true;
    PairGraph pg[2];
    int epoch = 0;
    long long graph_steps = 0; // steps run by graph replays
};

namespace
{
double now_s()
{
    return std::chrono::duration<double>( std::chrono::steady_clock::now().time_since_epoch() )
        .count();
}

// Cell grid of one system: cells of at least neigh_cut over [lo - depth,
// hi + depth] (the system's ghost shell), at most 1024 per dimension.
void grid_of( BSys &s, double depth, double cut )
{
    for ( int d = 0; d < 3; ++d )
    {
        const double mn = s.lo[d] - depth, ext = s.L[d] + 2.0 * depth;
        s.gn[d] = std::min( 1024, std::max( 1, (int)std::floor( ext / cut ) ) );
        s.gmin[d] = mn;
        s.grd[d] = s.gn[d] / ext;
    }
}

cmd_status upload_table( cmd_batch *B )
{
    const double depth = B->N->neigh_cut;
    int cb = 0, a0 = 0;
    B->htab.resize( B->members.size() );
    for ( size_t k = 0; k < B->members.size(); ++k )
    {
        BMember &m = B->members[k];
        BSys &s = B->htab[k];
        std::memset( &s, 0, sizeof( s ) );
        for ( int d = 0; d < 3; ++d )
        {
            s.lo[d] = m.lo[d];
            s.hi[d] = m.hi[d];
            s.L[d] = m.hi[d] - m.lo[d];
        }
        grid_of( s, depth, B->N->neigh_cut );
        s.cell_base = cb;
        cb += s.gn[0] * s.gn[1] * s.gn[2];
        s.a0 = a0;
        s.na = m.n;
        m.a0 = a0;
        a0 += m.n;
    }
    B->ncell = cb;
    CMD_TRY( B->tab.reserve( std::max<size_t>( 1, B->htab.size() ) ) );
    if ( !B->htab.empty() )
        CMD_HIP( hipMemcpyAsync( B->tab.p, B->htab.data(), sizeof( BSys ) * B->htab.size(),
                                 hipMemcpyHostToDevice, B->stream ) );
    return CMD_OK;
}

// Stable counting sort of n cell ids already in cs.cell_of.
cmd_status sort_cells( cmd_batch *B, CellSort &cs, int n )
{
    return cell_sort_ids( n, B->ncell, cs, B->S->sc, B->stream );
}

cmd_status build_list( cmd_batch *B, int G )
{
    cmd_system *S = B->S;
    cmd_neigh *N = B->N;
    const int n = S->n_local;
    const double rc2 = N->neigh_cut * N->neigh_cut;
    N->lanes = G;
    N->half = 0;
    N->tiled = false;
    N->csr_valid = false;
    N->nnz = -1;
    N->n_rows = n;
    if ( N->cap <= 0 )
        N->cap = std::max( 1, N->max_neigh_guess );
    CMD_TRY( N->counts.reserve( std::max( n, 1 ) ) );
    CMD_TRY( B->max_row.reserve( 1 ) );
    for ( int attempt = 0; attempt < 4; ++attempt )
    {
        CMD_TRY( N->slab.reserve( slab_entries( n, N->cap, G ) ) );
        CMD_TRY( fill_i32( B->max_row.p, 0, 1, B->stream ) );
        if ( n > 0 )
        {
#define CMD_BN( GG )                                                                   \
    k_b_neigh_cell<GG><<<( B->ncell + 3 ) / 4, 256, 0, B->stream>>>(                    \
        B->ncell, S->x.p, B->xs.p, B->lsort.start.p, B->asort.start.p, B->sys.p, B->tab.p, \
        rc2, N->cap, N->counts.p, N->slab.p, B->max_row.p )
            switch ( G )
            {
            case 1: CMD_BN( 1 ); break;
            case 2: CMD_BN( 2 ); break;
            case 4: CMD_BN( 4 ); break;
            default: CMD_BN( 8 ); break;
            }
#undef CMD_BN
            CMD_LAUNCH_CHECK();
        }
        int mr = 0;
        CMD_HIP( hipMemcpyAsync( &mr, B->max_row.p, sizeof( int ), hipMemcpyDeviceToHost,
                                 B->stream ) );
        CMD_HIP( hipStreamSynchronize( B->stream ) );
        N->max_row = mr;
        if ( mr <= N->cap )
        {
            // neighbor_verlet.h:58-61: the guess follows the longest row
            if ( mr > N->max_neigh_guess )
                N->max_neigh_guess = (int)( 1.1 * mr );
            return CMD_OK;
        }
        N->cap = (int)( 1.1 * mr ) + 1;
    }
    set_error( "batch neighbour list did not converge" );
    return CMD_ERR_STATE;
}

// Rebuild of every system from wrapped positions: binning sort of the owned
// atoms, ghost images, cell sort of all atoms, Verlet list.
cmd_status rebuild( cmd_batch *B, int G )
{
    ++B->epoch; // captured steps hold this build's buffers and sizes
    const double t0 = now_s();
    cmd_system *S = B->S;
    const hipStream_t st = B->stream;
    const int n = S->n_local;
    const double depth = B->N->neigh_cut;
    S->n_ghost = 0;
    if ( n == 0 )
        return CMD_OK;
    // Binning::create_binning + System::permute (binning_cabana_impl.h:57-113)
    CMD_TRY( B->lsort.cell_of.reserve( n ) );
    k_b_cell_local<<<nblk( n ), TPB, 0, st>>>( n, S->x.p, B->sys.p, B->tab.p,
                                               B->lsort.cell_of.p );
    CMD_LAUNCH_CHECK();
    CMD_TRY( sort_cells( B, B->lsort, n ) );
    CMD_TRY( permute_local( S, B->lsort.perm.p ) );
    // ghost images (exchange_halo, comm_mpi_impl.h:280-367)
    CMD_TRY( B->gcnt.reserve( n ) );
    CMD_TRY( B->gofs.reserve( n + 1 ) );
    k_b_ghost_count<<<nblk( n ), TPB, 0, st>>>( n, S->x.p, B->sys.p, B->tab.p, depth,
                                                B->gcnt.p );
    CMD_LAUNCH_CHECK();
    int64_t ng = 0;
    CMD_TRY( scan_exclusive( B->gcnt.p, B->gofs.p, n, S->sc, st, &ng ) );
    CMD_REQUIRE( (int64_t)n + ng < ( int64_t( 1 ) << 31 ), CMD_ERR_OOM,
                 "batch of %d atoms has %lld ghosts: over the int range", n, (long long)ng );
    CMD_TRY( S->reserve( n + (int)ng ) );
    CMD_TRY( S->ghost_root.reserve( std::max<int64_t>( 1, ng ) ) );
    CMD_TRY( S->ghost_code.reserve( std::max<int64_t>( 1, ng ) ) );
    k_b_ghost_fill<<<nblk( n ), TPB, 0, st>>>( n, S->x.p, S->id.p, B->sys.p, B->tab.p, depth,
                                               B->gofs.p, S->ghost_root.p, S->ghost_code.p );
    CMD_LAUNCH_CHECK();
    S->n_ghost = (int)ng;
    // candidates: all atoms in their system's cells
    const int nt = n + (int)ng;
    CMD_TRY( B->asort.cell_of.reserve( nt ) );
    k_b_cell_all<<<nblk( nt ), TPB, 0, st>>>( n, nt, S->x.p, B->sys.p, S->ghost_root.p,
                                              B->tab.p, B->asort.cell_of.p );
    CMD_LAUNCH_CHECK();
    CMD_TRY( sort_cells( B, B->asort, nt ) );
    CMD_TRY( B->xs.reserve( nt ) );
    k_b_gather<<<nblk( nt ), TPB, 0, st>>>( nt, B->asort.perm.p, S->x.p, B->xs.p );
    CMD_LAUNCH_CHECK();
    CMD_TRY( build_list( B, G ) );
    B->timers[0] += now_s() - t0;
    return CMD_OK;
}

cmd_status refresh( cmd_batch *B )
{
    cmd_system *S = B->S;
    if ( S->n_ghost == 0 )
        return CMD_OK;
    k_b_ghost_refresh<<<nblk( S->n_ghost ), TPB, 0, B->stream>>>(
        S->n_local, S->n_ghost, S->x.p, B->sys.p, S->ghost_root.p, S->ghost_code.p, B->tab.p );
    CMD_LAUNCH_CHECK();
    return CMD_OK;
}

cmd_status wrap( cmd_batch *B, double4 *x )
{
    const int n = B->S->n_local;
    if ( n == 0 )
        return CMD_OK;
    k_b_wrap<<<nblk( n ), TPB, 0, B->stream>>>( n, x, B->sys.p, B->tab.p );
    CMD_LAUNCH_CHECK();
    return CMD_OK;
}

// Per-system thermo of the current state into B->hsums (pe, m v^2 pairs);
// rows are appended to the members selected by `which` (nullptr = all).
cmd_status thermo( cmd_batch *B, const std::vector<char> *which )
{
    const double t0 = now_s();
    cmd_system *S = B->S;
    cmd_neigh *N = B->N;
    const int n = S->n_local;
    const int ns = (int)B->members.size();
    if ( ns == 0 )
        return CMD_OK;
    CMD_TRY( B->pe.reserve( std::max( n, 1 ) ) );
    CMD_TRY( B->mv2.reserve( std::max( n, 1 ) ) );
    CMD_TRY( B->sums.reserve( 2 * (size_t)ns ) );
    if ( n > 0 && B->nnp )
    {
        // ForceNNP::compute_energy (force_nnp_cabana_neigh_impl.h:101-130)
        // per system: (sum_i E_i + n * offset) * 27.211384021355236
        CMD_TRY( nnp_atom_energy( B->F->nnp, n, B->pe.p, B->stream ) );
        k_b_mv2<<<nblk( n ), TPB, 0, B->stream>>>( n, S->cap, S->x.p, S->v.p, S->mass_table(),
                                                   B->mv2.p );
        CMD_LAUNCH_CHECK();
    }
    else if ( n > 0 )
    {
#define CMD_BT( GG )                                                                        \
    k_b_atom_thermo<GG><<<nblk( n ), TPB, 0, B->stream>>>( n, S->cap, N->cap, S->x.p, S->v.p, \
                                                           N->counts.p, N->slab.p, B->Pd.p,   \
                                                           S->mass_table(), B->pe.p, B->mv2.p )
        switch ( N->lanes )
        {
        case 1: CMD_BT( 1 ); break;
        case 2: CMD_BT( 2 ); break;
        case 4: CMD_BT( 4 ); break;
        default: CMD_BT( 8 ); break;
        }
#undef CMD_BT
        CMD_LAUNCH_CHECK();
    }
    k_b_sys_sum<<<ns, TPB, 0, B->stream>>>( B->tab.p, B->pe.p, B->mv2.p, B->sums.p );
    CMD_LAUNCH_CHECK();
    B->hsums.resize( 2 * (size_t)ns );
    CMD_HIP( hipMemcpyAsync( B->hsums.data(), B->sums.p, sizeof( double ) * 2 * ns,
                             hipMemcpyDeviceToHost, B->stream ) );
    CMD_HIP( hipStreamSynchronize( B->stream ) );
    for ( int k = 0; k < ns; ++k )
    {
        if ( which && !( *which )[k] )
            continue;
        BMember &m = B->members[k];
        double pe = B->hsums[2 * k];
        const double mv2 = B->hsums[2 * k + 1];
        if ( B->nnp )
            pe = ( pe + m.n * nnp_energy_offset( B->F->nnp ) ) * 27.211384021355236;
        const double ke = 0.5 * S->mvv2e * mv2;
        // property_temperature_impl.h:55-77 (dof 3N - 3), cabanamd_impl.h:238-242
        const double T = mv2 * S->mvv2e / ( ( 3.0 * m.n - 3.0 ) * S->boltz );
        m.thermo.push_back( (double)( B->step - m.join_step ) );
        m.thermo.push_back( T );
        m.thermo.push_back( pe / m.n );
        m.thermo.push_back( ( pe + ke ) / m.n );
    }
    B->timers[2] += now_s() - t0;
    return CMD_OK;
}


// Forces of the current positions with the current list (CbnMD::init's
// first compute; for NNP also the ghost-force fold, cabanamd_impl.h:218-223).
cmd_status forces( cmd_batch *B )
{
    if ( B->S->n_local == 0 )
        return CMD_OK;
    if ( B->nnp )
    {
        CMD_TRY( cmd_force_compute( B->F, B->S, B->N ) );
        return cmd_comm_update_force( B->S );
    }
    StepCtl ctl{};
    return lj_full( B->F, B->S, B->N, ctl, nullptr, nullptr, nullptr );
}

// Times the force kernel on each list layout and keeps the fastest.
cmd_status autotune( cmd_batch *B )
{
    const int cands[] = { 1, 2, 4, 8 };
    hipEvent_t e0, e1;
    CMD_HIP( hipEventCreate( &e0 ) );
    CMD_HIP( hipEventCreate( &e1 ) );
    float best = 1e30f;
    int bestG = 1;
    StepCtl ctl{};
    cmd_status st = CMD_OK;
    for ( int G : cands )
    {
        if ( ( st = build_list( B, G ) ) != CMD_OK )
            break;
        if ( ( st = lj_full( B->F, B->S, B->N, ctl, nullptr, nullptr, nullptr ) ) != CMD_OK )
            break;
        (void)hipEventRecord( e0, B->stream );
        for ( int r = 0; r < 3; ++r )
            (void)lj_full( B->F, B->S, B->N, ctl, nullptr, nullptr, nullptr );
        (void)hipEventRecord( e1, B->stream );
        (void)hipEventSynchronize( e1 );
        float ms = 0.f;
        (void)hipEventElapsedTime( &ms, e0, e1 );
        if ( ms < best )
        {
            best = ms;
            bestG = G;
        }
    }
    (void)hipEventDestroy( e0 );
    (void)hipEventDestroy( e1 );
    CMD_TRY( st );
    B->lanes = bestG;
    return build_list( B, bestG );
}

// Membership changes at a rebuild boundary: leaving systems are dropped,
// waiting ones appended; then the whole batch is rebuilt and its forces
// recomputed.  For systems that stay, positions are those of the rebuild at
// this same step, so the sort is the identity and lists and forces come out
// bit-identical to what they had.
cmd_status apply_changes( cmd_batch *B )
{
    ++B->epoch; // membership changes move every buffer
    if ( B->pending.empty() && B->leaving.empty() )
        return CMD_OK;
    const double t0 = now_s();
    cmd_system *S = B->S;
    const hipStream_t st = B->stream;
    std::vector<BMember> keep;
    for ( BMember &m : B->members )
        if ( std::find( B->leaving.begin(), B->leaving.end(), m.handle ) == B->leaving.end() )
            keep.push_back( m );
    int total = 0;
    for ( const BMember &m : keep )
        total += m.n;
    for ( const BPending &q : B->pending )
        total += (int)q.id.size();
    // new flat arrays in the alternate buffers
    CMD_TRY( S->reserve( std::max( total, 1 ) ) );
    CMD_TRY( S->x_alt.reserve( S->cap ) );
    CMD_TRY( S->id_alt.reserve( S->cap ) );
    CMD_TRY( S->v_alt.reserve( 3 * (size_t)S->cap ) );
    CMD_TRY( S->f_alt.reserve( 3 * (size_t)S->cap ) );
    CMD_TRY( B->sys.reserve( std::max( total, 1 ) ) );
    const int cap = S->cap;
    int a = 0;
    for ( BMember &m : keep )
    {
        if ( m.n > 0 && a != m.a0 )
        {
            CMD_HIP( hipMemcpyAsync( S->x_alt.p + a, S->x.p + m.a0, sizeof( double4 ) * m.n,
                                     hipMemcpyDeviceToDevice, st ) );
            CMD_HIP( hipMemcpyAsync( S->id_alt.p + a, S->id.p + m.a0, sizeof( int ) * m.n,
                                     hipMemcpyDeviceToDevice, st ) );
            for ( int c = 0; c < 3; ++c )
                CMD_HIP( hipMemcpyAsync( S->v_alt.p + (size_t)c * cap + a,
                                         S->v.p + (size_t)c * cap + m.a0,
                                         sizeof( double ) * m.n, hipMemcpyDeviceToDevice, st ) );
        }
        else if ( m.n > 0 )
        {
            CMD_HIP( hipMemcpyAsync( S->x_alt.p + a, S->x.p + a, sizeof( double4 ) * m.n,
                                     hipMemcpyDeviceToDevice, st ) );
            CMD_HIP( hipMemcpyAsync( S->id_alt.p + a, S->id.p + a, sizeof( int ) * m.n,
                                     hipMemcpyDeviceToDevice, st ) );
            for ( int c = 0; c < 3; ++c )
                CMD_HIP( hipMemcpyAsync( S->v_alt.p + (size_t)c * cap + a,
                                         S->v.p + (size_t)c * cap + a, sizeof( double ) * m.n,
                                         hipMemcpyDeviceToDevice, st ) );
        }
        m.a0 = a;
        a += m.n;
    }
    for ( BPending &q : B->pending )
    {
        const int n = (int)q.id.size();
        BMember m;
        m.handle = q.handle;
        m.n = n;
        for ( int d = 0; d < 3; ++d )
        {
            m.lo[d] = q.lo[d];
            m.hi[d] = q.hi[d];
        }
        m.join_step = B->step;
        m.a0 = a;
        if ( n > 0 )
        {
            CMD_HIP( hipMemcpyAsync( S->x_alt.p + a, q.x.data(), sizeof( double4 ) * n,
                                     hipMemcpyHostToDevice, st ) );
            CMD_HIP( hipMemcpyAsync( S->id_alt.p + a, q.id.data(), sizeof( int ) * n,
                                     hipMemcpyHostToDevice, st ) );
            for ( int c = 0; c < 3; ++c )
                CMD_HIP( hipMemcpyAsync( S->v_alt.p + (size_t)c * cap + a,
                                         q.v.data() + (size_t)c * n, sizeof( double ) * n,
                                         hipMemcpyHostToDevice, st ) );
        }
        keep.push_back( m );
        a += n;
    }
    // the host staging above must outlive the copies
    CMD_HIP( hipStreamSynchronize( st ) );
    S->x.swap( S->x_alt );
    S->id.swap( S->id_alt );
    S->v.swap( S->v_alt );
    const int nnew = (int)B->pending.size();
    B->pending.clear();
    B->leaving.clear();
    B->members.swap( keep );
    S->n_local = total;
    S->N = total;
    S->n_ghost = 0;
    CMD_TRY( upload_table( B ) );
    for ( size_t k = 0; k < B->members.size(); ++k )
    {
        const BMember &m = B->members[k];
        if ( m.n > 0 )
        {
            k_b_fill_sys<<<nblk( m.n ), TPB, 0, st>>>( m.a0, m.n, (int)k, B->sys.p );
            CMD_LAUNCH_CHECK();
        }
    }
    // the new systems start from wrapped positions (CbnMD::init's exchange)
    CMD_TRY( wrap( B, S->x.p ) );
    if ( B->nnp && B->lanes == 0 )
        B->lanes = 1; // the NNP kernels read the list as CSR
    const bool tune = B->lanes == 0 && S->n_local > 0;
    CMD_TRY( rebuild( B, tune ? 1 : B->lanes ) );
    if ( tune )
        CMD_TRY( autotune( B ) );
    // step forces of every system (CbnMD::init's first compute for new ones)
    CMD_TRY( forces( B ) );
    std::vector<char> fresh( B->members.size(), 0 );
    for ( size_t k = B->members.size() - nnew; k < B->members.size(); ++k )
        fresh[k] = 1;
    CMD_TRY( thermo( B, &fresh ) );
    B->timers[3] += now_s() - t0;
    return CMD_OK;
}
} // namespace

extern "C" {

cmd_status cmd_batch_create( const cmd_params *p, cmd_batch **out )
{
    CMD_REQUIRE( p && out, CMD_ERR_ARG, "params and out are required" );
    CMD_REQUIRE( p->force_type == 0 || p->force_type == 2, CMD_ERR_ARG,
                 "a batch runs lj/cut (force_type 0) or nnp (2) systems" );
    CMD_REQUIRE( !p->half, CMD_ERR_ARG, "a batch uses full neighbour lists" );
    CMD_REQUIRE( p->every > 0, CMD_ERR_ARG, "neigh_modify every must be positive" );
    cmd_batch *B = new cmd_batch();
    B->p = *p;
    cmd_status s = cmd_system_create( p->ntypes, &B->S );
    if ( s == CMD_OK )
        s = cmd_system_set_units( B->S, p->dt, p->mvv2e, p->boltz );
    if ( s == CMD_OK )
        s = cmd_system_set_mass( B->S, p->ntypes, p->mass );
    if ( s == CMD_OK )
        s = cmd_neigh_create( p->force_cutoff + p->skin, 0, p->max_neigh_guess, &B->N );
    B->nnp = p->force_type == 2;
    if ( s == CMD_OK && !B->nnp )
        s = cmd_force_lj_create( p->ntypes, &B->F );
    for ( int c = 0; s == CMD_OK && !B->nnp && c < p->ncoeff; ++c )
        s = cmd_force_lj_init_coeff( B->F, p->coeff_i[c], p->coeff_j[c], p->coeff_eps[c],
                                     p->coeff_sigma[c], p->coeff_cut[c] );
    if ( s == CMD_OK )
    {
        B->stream = B->S->stream;
        B->N->allow_tiles = false;
    }
    if ( s == CMD_OK && !B->nnp )
    {
        const LJParams P = lj_params( B->F );
        s = B->Pd.reserve( 1 );
        if ( s == CMD_OK && hipMemcpy( B->Pd.p, &P, sizeof( P ), hipMemcpyHostToDevice ) !=
                                hipSuccess )
        {
            set_error( "upload of LJ parameters failed" );
            s = CMD_ERR_HIP;
        }
    }
    if ( s != CMD_OK )
    {
        cmd_batch_destroy( B );
        return s;
    }
    *out = B;
    return CMD_OK;
}

cmd_status cmd_batch_destroy( cmd_batch *b )
{
    if ( !b )
        return CMD_OK;
    if ( b->S && b->S->stream )
        (void)hipStreamSynchronize( b->S->stream );
    for ( auto &g : b->pg )
        if ( g.exec )
            (void)hipGraphExecDestroy( g.exec );
    if ( b->F )
        cmd_force_destroy( b->F );
    if ( b->N )
        cmd_neigh_destroy( b->N );
    b->Pd.release();
    b->tab.release();
    if ( b->S )
        cmd_system_destroy( b->S );
    delete b;
    return CMD_OK;
}

cmd_status cmd_batch_set_precision( cmd_batch *b, int prec )
{
    CMD_REQUIRE( b, CMD_ERR_ARG, "batch is NULL" );
    CMD_REQUIRE( prec == CMD_PREC_FP64 || prec == CMD_PREC_MIXED, CMD_ERR_ARG,
                 "unknown precision %d", prec );
    if ( prec != b->prec )
        ++b->epoch; // captured steps launch the other precision's kernel
    b->prec = prec;
    return b->F ? cmd_force_set_precision( b->F, prec ) : CMD_OK;
}

cmd_status cmd_batch_set_nnp_model( cmd_batch *b, const cmd_nnp_model *m )
{
    CMD_REQUIRE( b && m, CMD_ERR_ARG, "batch and model are required" );
    CMD_REQUIRE( b->nnp, CMD_ERR_STATE, "the batch runs lj/cut systems" );
    CMD_REQUIRE( b->members.empty() && b->pending.empty(), CMD_ERR_STATE,
                 "the model is set before systems are added" );
    if ( b->F )
    {
        cmd_force_destroy( b->F );
        b->F = nullptr;
    }
    CMD_TRY( cmd_force_nnp_create( m, &b->F ) );
    return cmd_force_set_precision( b->F, b->prec );
}

cmd_status cmd_batch_set_graphs( cmd_batch *b, int on )
{
    CMD_REQUIRE( b, CMD_ERR_ARG, "batch is NULL" );
    b->graphs = on != 0;
    return CMD_OK;
}

cmd_status cmd_batch_graph_steps( const cmd_batch *b, long long *steps )
{
    CMD_REQUIRE( b && steps, CMD_ERR_ARG, "NULL argument" );
    *steps = b->graph_steps;
    return CMD_OK;
}

cmd_status cmd_batch_set_lanes( cmd_batch *b, int lanes )
{
    CMD_REQUIRE( b, CMD_ERR_ARG, "batch is NULL" );
    CMD_REQUIRE( lanes == 0 || lanes == 1 || lanes == 2 || lanes == 4 || lanes == 8,
                 CMD_ERR_ARG, "lanes must be 0 (tune), 1, 2, 4 or 8" );
    CMD_REQUIRE( b->members.empty(), CMD_ERR_STATE,
                 "the list layout is fixed once systems are running" );
    b->lanes = lanes;
    return CMD_OK;
}

cmd_status cmd_batch_add( cmd_batch *b, int n, const double *x, const double *v,
                          const int *type, const int *id, const double lo[3],
                          const double hi[3], int *handle )
{
    CMD_REQUIRE( b && x && type && lo && hi && n > 0, CMD_ERR_ARG,
                 "batch, n > 0, x, type, lo and hi are required" );
    CMD_REQUIRE( b->F, CMD_ERR_STATE, "an NNP batch needs its model first" );
    const double depth = b->N->neigh_cut;
    for ( int d = 0; d < 3; ++d )
        CMD_REQUIRE( hi[d] - lo[d] > 2.0 * depth, CMD_ERR_ARG,
                     "box length %g in dimension %d must exceed twice the ghost depth %g",
                     hi[d] - lo[d], d, depth );
    BPending q;
    q.handle = b->next_handle++;
    q.x.resize( n );
    q.v.assign( 3 * (size_t)n, 0.0 );
    q.id.resize( n );
    for ( int i = 0; i < n; ++i )
    {
        CMD_REQUIRE( type[i] >= 0 && type[i] < b->p.ntypes, CMD_ERR_ARG,
                     "atom %d has type %d outside [0,%d)", i, type[i], b->p.ntypes );
        q.x[i] = make_double4( x[3 * i], x[3 * i + 1], x[3 * i + 2], (double)type[i] );
        q.id[i] = id ? id[i] : i + 1;
        if ( v )
            for ( int c = 0; c < 3; ++c )
                q.v[(size_t)c * n + i] = v[3 * i + c];
    }
    for ( int d = 0; d < 3; ++d )
    {
        q.lo[d] = lo[d];
        q.hi[d] = hi[d];
    }
    if ( handle )
        *handle = q.handle;
    b->pending.push_back( std::move( q ) );
    return CMD_OK;
}

cmd_status cmd_batch_remove( cmd_batch *b, int handle )
{
    CMD_REQUIRE( b, CMD_ERR_ARG, "batch is NULL" );
    for ( size_t k = 0; k < b->pending.size(); ++k )
        if ( b->pending[k].handle == handle )
        {
            b->pending.erase( b->pending.begin() + k );
            return CMD_OK;
        }
    for ( const BMember &m : b->members )
        if ( m.handle == handle )
        {
            b->leaving.push_back( handle );
            return CMD_OK;
        }
    set_error( "no system with handle %d", handle );
    return CMD_ERR_ARG;
}

// Systems added or removed since the last call join or leave now when the
// batch sits on a rebuild boundary (else at the first call that starts on
// one).  Thermo rows are taken every thermo_rate steps and at the last step.
// Capture two plain steps (refresh ghosts, fused step into x_alt; the
// same from x_alt back into x) as one graph.  The capture runs nothing; the
// host-side buffer swaps it makes are undone by the second.  Only kernels
// are recorded (no copies, fills or host waits).
static cmd_status capture_plain_pair( cmd_batch *b, const StepCtl &ctl,
                                      cmd_batch::PairGraph &pg )
{
    cmd_system *S = b->S;
    if ( pg.exec )
    {
        CMD_HIP( hipGraphExecDestroy( pg.exec ) );
        pg.exec = nullptr;
    }
    pg.x = S->x.p;
    CMD_HIP( hipStreamBeginCapture( b->stream, hipStreamCaptureModeThreadLocal ) );
    cmd_status rc = CMD_OK;
    for ( int k = 0; k < 2; ++k )
    {
        if ( rc == CMD_OK )
            rc = refresh( b );
        if ( rc == CMD_OK )
            rc = lj_full( b->F, S, b->N, ctl, S->x_alt.p, nullptr, nullptr );
        S->x.swap( S->x_alt );
    }
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture( b->stream, &g );
    if ( rc != CMD_OK )
    {
        if ( e == hipSuccess && g )
            (void)hipGraphDestroy( g );
        return rc;
    }
    CMD_HIP( e );
    const hipError_t ei = hipGraphInstantiate( &pg.exec, g, nullptr, nullptr, 0 );
    (void)hipGraphDestroy( g );
    CMD_HIP( ei );
    pg.epoch = b->epoch;
    return CMD_OK;
}

cmd_status cmd_batch_run( cmd_batch *b, int nsteps )
{
    CMD_REQUIRE( b && nsteps >= 0, CMD_ERR_ARG, "bad argument" );
    const cmd_params &p = b->p;
    if ( b->step % p.every == 0 )
        CMD_TRY( apply_changes( b ) );
    cmd_system *S = b->S;
    cmd_neigh *N = b->N;
    if ( nsteps == 0 || S->n_local == 0 )
    {
        b->step += nsteps;
        return CMD_OK;
    }
    const double t0 = now_s();
    const hipStream_t st = b->stream;
    const int G = N->lanes;
    const int s0 = b->step + 1, s1 = b->step + nsteps;
    if ( b->nnp )
    {
        // CbnMD::run's order (cabanamd_impl.h:285-399) step by step
        for ( int s = s0; s <= s1; ++s )
        {
            b->step = s;
            CMD_TRY( cmd_nve_initial( S ) );
            if ( s % p.every == 0 )
            {
                CMD_TRY( wrap( b, S->x.p ) );
                CMD_TRY( rebuild( b, G ) );
            }
            else
                CMD_TRY( refresh( b ) );
            CMD_TRY( forces( b ) );
            CMD_TRY( cmd_nve_final( S ) );
            if ( s == s1 || ( p.thermo_rate > 0 && s % p.thermo_rate == 0 ) )
                CMD_TRY( thermo( b, nullptr ) );
        }
        CMD_HIP( hipStreamSynchronize( st ) );
        b->timers[1] += now_s() - t0;
        return CMD_OK;
    }
    StepCtl ctl{};
    ctl.dtf = 0.5 * S->dt / S->mvv2e;
    ctl.dtv = S->dt;
    CMD_TRY( S->x_alt.reserve( S->cap ) );
    CMD_TRY( cmd_nve_initial( S ) );
    if ( s0 % p.every == 0 )
        CMD_TRY( wrap( b, S->x.p ) );
    // a plain step: no rebuild, no thermo row, no wrap for the next rebuild,
    // not the last of the call
    auto plain = [&]( int t )
    {
        return t < s1 && t % p.every != 0 && ( t + 1 ) % p.every != 0 &&
               !( p.thermo_rate > 0 && t % p.thermo_rate == 0 );
    };
    // several types: the coefficient tables go to the device now, so the
    // captured steps launch with no host copy (a batch's coefficients are
    // set at its creation and do not change)
    CMD_TRY( lj_params_upload( b->F, st ) );
    const bool graphs = b->graphs;
    for ( int s = s0; s <= s1; ++s )
    {
        b->step = s;
        if ( graphs && plain( s ) && plain( s + 1 ) )
        {
            StepCtl gc = ctl;
            gc.do_final = 1;
            gc.do_next_initial = 1;
            gc.write_f = 0;
            gc.thermo = 0;
            gc.wrap_next = 0;
            // the graph of the buffer the pair starts from; else capture
            // over a slot of another epoch, or the other buffer's
            cmd_batch::PairGraph *use = nullptr;
            for ( auto &g : b->pg )
                if ( g.exec && g.epoch == b->epoch && g.x == S->x.p )
                    use = &g;
            if ( !use )
            {
                use = &b->pg[0];
                for ( auto &g : b->pg )
                    if ( !g.exec || g.epoch != b->epoch )
                        use = &g;
                CMD_TRY( capture_plain_pair( b, gc, *use ) );
            }
            CMD_HIP( hipGraphLaunch( use->exec, st ) );
            b->graph_steps += 2;
            b->step = ++s; // the pair ends with x back in place
            continue;
        }
        if ( s % p.every == 0 )
            CMD_TRY( rebuild( b, G ) );
        else
            CMD_TRY( refresh( b ) );
        const bool last = s == s1;
        const bool th = last || ( p.thermo_rate > 0 && s % p.thermo_rate == 0 );
        ctl.do_final = 1;
        ctl.thermo = 0;
        ctl.wrap_next = 0;
        if ( th )
        {
            // final half-kick only, so the thermo sees this step's velocities
            ctl.do_next_initial = 0;
            ctl.write_f = 1;
            CMD_TRY( S->x_alt.reserve( S->cap ) );
            CMD_TRY( lj_full( b->F, S, N, ctl, S->x_alt.p, nullptr, nullptr ) );
            CMD_TRY( thermo( b, nullptr ) );
            if ( !last )
            {
                CMD_TRY( cmd_nve_initial( S ) );
                if ( ( s + 1 ) % p.every == 0 )
                    CMD_TRY( wrap( b, S->x.p ) );
            }
        }
        else
        {
            ctl.do_next_initial = 1;
            ctl.write_f = 0;
            CMD_TRY( S->x_alt.reserve( S->cap ) );
            CMD_TRY( lj_full( b->F, S, N, ctl, S->x_alt.p, nullptr, nullptr ) );
            if ( ( s + 1 ) % p.every == 0 )
                CMD_TRY( wrap( b, S->x_alt.p ) );
            // ghosts are refreshed (or rebuilt) into the new buffer next step
            S->x.swap( S->x_alt );
        }
    }
    CMD_HIP( hipStreamSynchronize( st ) );
    b->timers[1] += now_s() - t0;
    return CMD_OK;
}

cmd_status cmd_batch_counts( const cmd_batch *b, int *nsys, int64_t *natoms, int *step,
                             int *lanes )
{
    CMD_REQUIRE( b, CMD_ERR_ARG, "batch is NULL" );
    if ( nsys )
        *nsys = (int)b->members.size();
    if ( natoms )
        *natoms = b->S->n_local;
    if ( step )
        *step = b->step;
    if ( lanes )
        *lanes = b->N->lanes;
    return CMD_OK;
}

static const BMember *find_member( const cmd_batch *b, int handle )
{
    for ( const BMember &m : b->members )
        if ( m.handle == handle )
            return &m;
    return nullptr;
}

cmd_status cmd_batch_system_info( const cmd_batch *b, int handle, int *n, int *local_step,
                                  int *state )
{
    CMD_REQUIRE( b, CMD_ERR_ARG, "batch is NULL" );
    int st = 0, nn = 0, ls = 0;
    for ( const BPending &q : b->pending )
        if ( q.handle == handle )
        {
            st = 1;
            nn = (int)q.id.size();
        }
    if ( const BMember *m = find_member( b, handle ) )
    {
        st = std::find( b->leaving.begin(), b->leaving.end(), handle ) != b->leaving.end() ? 3
                                                                                           : 2;
        nn = m->n;
        ls = b->step - m->join_step;
    }
    if ( n )
        *n = nn;
    if ( local_step )
        *local_step = ls;
    if ( state )
        *state = st;
    return CMD_OK;
}

cmd_status cmd_batch_download( cmd_batch *b, int handle, double *x, double *v, double *f,
                               int *type, int *id )
{
    CMD_REQUIRE( b, CMD_ERR_ARG, "batch is NULL" );
    const BMember *m = find_member( b, handle );
    CMD_REQUIRE( m, CMD_ERR_ARG, "system %d is not running", handle );
    cmd_system *S = b->S;
    const int n = m->n;
    std::vector<double4> hx( n );
    std::vector<double> hv( 3 * (size_t)n ), hf( 3 * (size_t)n );
    std::vector<int> hid( n );
    CMD_HIP( hipStreamSynchronize( b->stream ) );
    CMD_HIP( hipMemcpy( hx.data(), S->x.p + m->a0, sizeof( double4 ) * n,
                        hipMemcpyDeviceToHost ) );
    CMD_HIP( hipMemcpy( hid.data(), S->id.p + m->a0, sizeof( int ) * n,
                        hipMemcpyDeviceToHost ) );
    for ( int c = 0; c < 3; ++c )
    {
        CMD_HIP( hipMemcpy( hv.data() + (size_t)c * n, S->v.p + (size_t)c * S->cap + m->a0,
                            sizeof( double ) * n, hipMemcpyDeviceToHost ) );
        CMD_HIP( hipMemcpy( hf.data() + (size_t)c * n, S->f.p + (size_t)c * S->cap + m->a0,
                            sizeof( double ) * n, hipMemcpyDeviceToHost ) );
    }
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
        if ( id )
            id[i] = hid[i];
        for ( int c = 0; c < 3; ++c )
        {
            if ( v )
                v[3 * i + c] = hv[(size_t)c * n + i];
            if ( f )
                f[3 * i + c] = hf[(size_t)c * n + i];
        }
    }
    return CMD_OK;
}

cmd_status cmd_batch_thermo( const cmd_batch *b, int handle, int max_rows, double *rows,
                             int *rows_out )
{
    CMD_REQUIRE( b, CMD_ERR_ARG, "batch is NULL" );
    const BMember *m = find_member( b, handle );
    CMD_REQUIRE( m, CMD_ERR_ARG, "system %d is not running", handle );
    const int nr = (int)( m->thermo.size() / 4 );
    if ( rows_out )
        *rows_out = nr;
    if ( rows )
        std::memcpy( rows, m->thermo.data(),
                     sizeof( double ) * 4 * (size_t)std::min( nr, std::max( max_rows, 0 ) ) );
    return CMD_OK;
}

cmd_status cmd_batch_timers( const cmd_batch *b, double out[4] )
{
    CMD_REQUIRE( b && out, CMD_ERR_ARG, "bad argument" );
    for ( int k = 0; k < 4; ++k )
        out[k] = b->timers[k];
    return CMD_OK;
}

} // extern "C"
