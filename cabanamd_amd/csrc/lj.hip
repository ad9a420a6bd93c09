// lj.hip -- 12-6 Lennard-Jones on lane-interleaved Verlet lists.
//
// Restates force_types/force_lj_cabana_neigh_impl.h (CabanaMD):
//   init_coeff :62-89   lj1 = 48 eps sigma^12, lj2 = 24 eps sigma^6, cutsq
//   full force :151-203 (pair force applied for rsq < cutsq, unshifted)
//   half force :205-259 (the negated pair force also lands on j)
//   energy     :261-377 (shifted; full weight 1/2, half 1 or 1/2 for ghost j)
//
// Kernel mapping: G lanes per atom, APW = 64/G atoms per wave, following the
// list's slab layout (neighbor.hip), so each neighbour step of a wave is one
// 256-byte coalesced index load plus one 32-byte position gather per lane.
// Pair maths is branch-free (cutoff applied as a select) and neighbours are
// processed four at a time so that four index loads and four gathers are in
// flight per lane.  Block indices are remapped so each of the 8 XCDs walks a
// contiguous range of cell-sorted atoms (its x_j gathers stay in its L2).
#include "internal.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>

namespace cmd
{
static constexpr int TPB = 256;

// Round-robin dispatch puts block b on XCD b % 8; give XCD k the k-th
// contiguous slice of logical blocks instead.
__device__ __forceinline__ int xcd_remap( int b, int nb )
{
    const int q = nb >> 3, r = nb & 7;
    const int xcd = b & 7, k = b >> 3;
    return ( xcd < r ) ? xcd * ( q + 1 ) + k : r * ( q + 1 ) + ( xcd - r ) * q + k;
}

LJParams lj_params( const cmd_force *F )
{
    LJParams P;
    const int nt = F->ntypes;
    P.ntypes = nt;
    for ( int a = 0; a < CMD_MAX_TYPES * CMD_MAX_TYPES; ++a )
    {
        P.lj1[a] = P.lj2[a] = P.cutsq[a] = P.eshift[a] = 0.0;
        P.lj1f[a] = P.lj2f[a] = 0.0f;
    }
    for ( int a = 0; a < nt * nt; ++a )
    {
        P.lj1[a] = F->lj1[a];
        P.lj2[a] = F->lj2[a];
        P.cutsq[a] = F->cutsq[a];
        P.lj1f[a] = (float)F->lj1[a];
        P.lj2f[a] = (float)F->lj2[a];
        if ( F->cutsq[a] > 0 )
        {
            const double r2invc = 1.0 / F->cutsq[a];
            const double r6invc = r2invc * r2invc * r2invc;
            P.eshift[a] = r6invc * ( 0.5 * F->lj1[a] * r6invc - F->lj2[a] ) / 6.0;
        }
    }
    return P;
}

struct PairCoef
{
    double lj1, lj2, cs, esh;
    float lj1f, lj2f;
};

template <bool ONE_TYPE>
__device__ __forceinline__ PairCoef coef( const LJParams *__restrict__ Pd,
                                          const PairCoef &one, int ti, int tj )
{
    if ( ONE_TYPE )
        return one;
    const int t = ti * Pd->ntypes + tj;
    PairCoef c;
    c.lj1 = Pd->lj1[t];
    c.lj2 = Pd->lj2[t];
    c.cs = Pd->cutsq[t];
    c.esh = Pd->eshift[t];
    c.lj1f = Pd->lj1f[t];
    c.lj2f = Pd->lj2f[t];
    return c;
}

// Force factor of one pair (F_i = d * fpair), zero outside the cutoff; adds
// the shifted pair energy and r.F to e and w when ENERGY.
template <int PREC, bool ENERGY>
__device__ __forceinline__ double lj_pair( double rsq, const PairCoef &c,
                                           double &e, double &w )
{
    const bool in = rsq < c.cs;
    if ( PREC == CMD_PREC_FP64 )
    {
        const double r2inv = 1.0 / rsq;
        const double r6inv = r2inv * r2inv * r2inv;
        const double fp = ( r6inv * ( c.lj1 * r6inv - c.lj2 ) ) * r2inv;
        const double fpair = in ? fp : 0.0;
        if ( ENERGY )
        {
            const double ep = r6inv * ( 0.5 * c.lj1 * r6inv - c.lj2 ) / 6.0 - c.esh;
            e += in ? ep : 0.0;
            w += rsq * fpair;
        }
        return fpair;
    }
    else
    {
        const float r2inv = __builtin_amdgcn_rcpf( (float)rsq );
        const float r6inv = r2inv * r2inv * r2inv;
        const float fp = ( r6inv * __builtin_fmaf( c.lj1f, r6inv, -c.lj2f ) ) * r2inv;
        const double fpair = in ? (double)fp : 0.0;
        if ( ENERGY )
        {
            const float ep =
                r6inv * __builtin_fmaf( 0.5f * c.lj1f, r6inv, -c.lj2f ) * ( 1.0f / 6.0f );
            e += in ? (double)ep - c.esh : 0.0;
            w += rsq * fpair;
        }
        return fpair;
    }
}

template <int PREC>
__device__ __forceinline__ void accumulate( double &fx, double &fy, double &fz,
                                            double dx, double dy, double dz,
                                            double fpair )
{
    if ( PREC == CMD_PREC_FP64 )
    {
        // the reference's f_i += d * fpair, multiply then add
        fx += dx * fpair;
        fy += dy * fpair;
        fz += dz * fpair;
    }
    else
    {
        fx = __builtin_fma( dx, fpair, fx );
        fy = __builtin_fma( dy, fpair, fy );
        fz = __builtin_fma( dz, fpair, fz );
    }
}

__device__ void block_sum4( double a, double b, double c, double d,
                            double *__restrict__ out )
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
        out[threadIdx.x] = s;
    }
}

// Position of a lane in the slab: wave index, atom slot, sub-lane, atom.
struct SlabLane
{
    int wave, a, s, i;
};

template <int G>
__device__ __forceinline__ SlabLane slab_lane( int block )
{
    constexpr int APW = 64 / G;
    SlabLane L;
    const int t = block * TPB + (int)threadIdx.x;
    L.wave = t >> 6;
    const int lane = t & 63;
    L.a = lane % APW;
    L.s = lane / APW;
    L.i = L.wave * APW + L.a;
    return L;
}

// Row walk of one lane: neighbours k = r*G + s for r < nr at row[r*64]; the
// body gets (j, x_j).  Four index loads then four gathers are issued before
// any arithmetic.
template <int G, class Body>
__device__ __forceinline__ void walk_row( const int *__restrict__ row, int nr,
                                          const double4 *__restrict__ x,
                                          Body body )
{
    int r = 0;
    for ( ; r + 4 <= nr; r += 4 )
    {
        const int j0 = row[( r + 0 ) * 64];
        const int j1 = row[( r + 1 ) * 64];
        const int j2 = row[( r + 2 ) * 64];
        const int j3 = row[( r + 3 ) * 64];
        const double4 p0 = x[j0];
        const double4 p1 = x[j1];
        const double4 p2 = x[j2];
        const double4 p3 = x[j3];
        body( j0, p0 );
        body( j1, p1 );
        body( j2, p2 );
        body( j3, p3 );
    }
    for ( ; r < nr; ++r )
    {
        const int j = row[r * 64];
        body( j, x[j] );
    }
}

// Full list.  Without FUSED: f_i = sum_j d_ij fpair (ghost rows untouched).
// With FUSED the lane then finishes velocity-Verlet for atom i (final half
// kick, thermo partials) and, when asked, starts the next step (half kick,
// drift, wrap) writing positions to x_next.
template <int G, int PREC, bool ONE_TYPE, bool FUSED>
__global__ __launch_bounds__( TPB ) void k_lj_full(
    int n_local, int scap, int lcap, const double4 *__restrict__ x,
    double4 *__restrict__ x_next, double *__restrict__ v,
    double *__restrict__ f, const int *__restrict__ counts,
    const int *__restrict__ slab, const LJParams *__restrict__ Pd,
    PairCoef one, MassTable mt, StepCtl ctl, double *__restrict__ partial )
{
    constexpr int APW = 64 / G;
    const SlabLane L = slab_lane<G>( xcd_remap( blockIdx.x, gridDim.x ) );
    const int i = L.i;
    const bool active = i < n_local;
    const bool energy = FUSED && ctl.thermo;
    double fx = 0.0, fy = 0.0, fz = 0.0, pe = 0.0, vir = 0.0;
    double4 xi = make_double4( 0.0, 0.0, 0.0, 0.0 );
    if ( active )
    {
        xi = x[i];
        const int ti = (int)xi.w;
        const int n = counts[i];
        const int nr = n > L.s ? ( n - L.s + G - 1 ) / G : 0;
        const int R = ( lcap + G - 1 ) / G;
        const int *row = slab + (size_t)L.wave * R * 64 + L.s * APW + L.a;
        auto body = [&]( int, const double4 &xj )
        {
            const double dx = xi.x - xj.x;
            const double dy = xi.y - xj.y;
            const double dz = xi.z - xj.z;
            const double rsq = dx * dx + dy * dy + dz * dz;
            const PairCoef c = coef<ONE_TYPE>( Pd, one, ti, (int)xj.w );
            const double fpair = energy ? lj_pair<PREC, true>( rsq, c, pe, vir )
                                        : lj_pair<PREC, false>( rsq, c, pe, vir );
            accumulate<PREC>( fx, fy, fz, dx, dy, dz, fpair );
        };
        walk_row<G>( row, nr, x, body );
    }
#pragma unroll
    for ( int o = APW; o < 64; o <<= 1 )
    {
        fx += __shfl_xor( fx, o, 64 );
        fy += __shfl_xor( fy, o, 64 );
        fz += __shfl_xor( fz, o, 64 );
    }
    double mv2 = 0.0;
    if ( active && L.s == 0 )
    {
        if ( !FUSED || ctl.write_f )
        {
            f[i] = fx;
            f[scap + i] = fy;
            f[2 * scap + i] = fz;
        }
        if ( FUSED && ctl.do_final )
        {
            const double m = mt.m[(int)xi.w];
            const double dtfm = ctl.dtf / m;
            double vx = v[i] + dtfm * fx;
            double vy = v[scap + i] + dtfm * fy;
            double vz = v[2 * scap + i] + dtfm * fz;
            mv2 = ( vx * vx + vy * vy + vz * vz ) * m;
            if ( ctl.do_next_initial )
            {
                vx = vx + dtfm * fx;
                vy = vy + dtfm * fy;
                vz = vz + dtfm * fz;
                double4 p = xi;
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
            }
            v[i] = vx;
            v[scap + i] = vy;
            v[2 * scap + i] = vz;
        }
    }
    if ( energy )
        block_sum4( mv2, 0.5 * pe, 0.5 * vir, 0.0, partial + 4 * blockIdx.x );
}

// Half list: each listed pair once; the reaction goes to j by atomics
// (j may be a ghost, folded later by update_force).  Energy weights follow
// compute_energy_half: 1 for a local j, 1/2 for a ghost j.
template <int G, int PREC, bool ONE_TYPE>
__global__ __launch_bounds__( TPB ) void k_lj_half(
    int n_local, int scap, int lcap, const double4 *__restrict__ x,
    double *__restrict__ f, const int *__restrict__ counts,
    const int *__restrict__ slab, const LJParams *__restrict__ Pd,
    PairCoef one, int thermo, double *__restrict__ partial )
{
    constexpr int APW = 64 / G;
    const SlabLane L = slab_lane<G>( xcd_remap( blockIdx.x, gridDim.x ) );
    const int i = L.i;
    double pe = 0.0, vir = 0.0;
    if ( i < n_local )
    {
        const double4 xi = x[i];
        const int ti = (int)xi.w;
        const int n = counts[i];
        const int nr = n > L.s ? ( n - L.s + G - 1 ) / G : 0;
        const int R = ( lcap + G - 1 ) / G;
        const int *row = slab + (size_t)L.wave * R * 64 + L.s * APW + L.a;
        double fx = 0.0, fy = 0.0, fz = 0.0;
        auto body = [&]( int j, const double4 &xj )
        {
            const double dx = xi.x - xj.x;
            const double dy = xi.y - xj.y;
            const double dz = xi.z - xj.z;
            const double rsq = dx * dx + dy * dy + dz * dz;
            const PairCoef c = coef<ONE_TYPE>( Pd, one, ti, (int)xj.w );
            double ep = 0.0, wp = 0.0;
            const double fpair = thermo ? lj_pair<PREC, true>( rsq, c, ep, wp )
                                        : lj_pair<PREC, false>( rsq, c, ep, wp );
            if ( rsq < c.cs )
            {
                const double ax = dx * fpair, ay = dy * fpair, az = dz * fpair;
                fx += ax;
                fy += ay;
                fz += az;
                unsafeAtomicAdd( &f[j], -ax );
                unsafeAtomicAdd( &f[scap + j], -ay );
                unsafeAtomicAdd( &f[2 * scap + j], -az );
            }
            pe += ( ( j < n_local ) ? 1.0 : 0.5 ) * ep;
            vir += wp;
        };
        walk_row<G>( row, nr, x, body );
        unsafeAtomicAdd( &f[i], fx );
        unsafeAtomicAdd( &f[scap + i], fy );
        unsafeAtomicAdd( &f[2 * scap + i], fz );
    }
    if ( thermo )
        block_sum4( 0.0, pe, vir, 0.0, partial + 4 * blockIdx.x );
}

// compute_energy (reference list weights) and pair virial, FP64.
template <int G, bool ONE_TYPE>
__global__ __launch_bounds__( TPB ) void k_lj_energy(
    int n_local, int lcap, const double4 *__restrict__ x,
    const int *__restrict__ counts, const int *__restrict__ slab,
    const LJParams *__restrict__ Pd, PairCoef one, int half,
    double *__restrict__ partial )
{
    constexpr int APW = 64 / G;
    const SlabLane L = slab_lane<G>( blockIdx.x );
    const int i = L.i;
    double pe = 0.0, vir = 0.0;
    if ( i < n_local )
    {
        const double4 xi = x[i];
        const int ti = (int)xi.w;
        const int n = counts[i];
        const int nr = n > L.s ? ( n - L.s + G - 1 ) / G : 0;
        const int R = ( lcap + G - 1 ) / G;
        const int *row = slab + (size_t)L.wave * R * 64 + L.s * APW + L.a;
        for ( int r = 0; r < nr; ++r )
        {
            const int j = row[r * 64];
            const double4 xj = x[j];
            const double dx = xi.x - xj.x;
            const double dy = xi.y - xj.y;
            const double dz = xi.z - xj.z;
            const double rsq = dx * dx + dy * dy + dz * dz;
            const PairCoef c = coef<ONE_TYPE>( Pd, one, ti, (int)xj.w );
            if ( rsq < c.cs )
            {
                const double fac = half ? ( ( j < n_local ) ? 1.0 : 0.5 ) : 0.5;
                const double r2inv = 1.0 / rsq;
                const double r6inv = r2inv * r2inv * r2inv;
                pe += fac * r6inv * ( 0.5 * c.lj1 * r6inv - c.lj2 ) / 6.0;
                pe -= fac * c.esh;
                const double fpair = ( r6inv * ( c.lj1 * r6inv - c.lj2 ) ) * r2inv;
                vir += ( half ? 1.0 : 0.5 ) * rsq * fpair;
            }
        }
    }
    block_sum4( 0.0, pe, vir, 0.0, partial + 4 * blockIdx.x );
}

// --------------------------------------------------------------- dispatch
static cmd_status upload_params( const cmd_force *F, hipStream_t st,
                                 LJParams **dptr )
{
    cmd_force *mf = const_cast<cmd_force *>( F );
    CMD_TRY( mf->dparams.reserve( 1 ) );
    static thread_local LJParams P;
    P = lj_params( F );
    CMD_HIP( hipMemcpyAsync( mf->dparams.p, &P, sizeof( P ),
                             hipMemcpyHostToDevice, st ) );
    CMD_HIP( hipStreamSynchronize( st ) );
    *dptr = mf->dparams.p;
    return CMD_OK;
}

static PairCoef one_coef( const LJParams &P )
{
    PairCoef c;
    c.lj1 = P.lj1[0];
    c.lj2 = P.lj2[0];
    c.cs = P.cutsq[0];
    c.esh = P.eshift[0];
    c.lj1f = P.lj1f[0];
    c.lj2f = P.lj2f[0];
    return c;
}

static int blocks_for( int n, int lanes )
{
    const int atoms_per_block = TPB / lanes;
    return std::max( 1, ( n + atoms_per_block - 1 ) / atoms_per_block );
}

template <int G, int PREC, bool ONE, bool FUSED>
static void launch_full_t( int nb, hipStream_t st, int n, int scap, int lcap,
                           const double4 *x, double4 *xn, double *v, double *f,
                           const int *cnt, const int *slab, const LJParams *Pd,
                           const PairCoef &c, const MassTable &mt,
                           const StepCtl &ctl, double *partial )
{
    k_lj_full<G, PREC, ONE, FUSED><<<nb, TPB, 0, st>>>(
        n, scap, lcap, x, xn, v, f, cnt, slab, Pd, c, mt, ctl, partial );
}

template <int PREC, bool ONE, bool FUSED>
static cmd_status launch_full_g( int G, int nb, hipStream_t st, int n, int scap,
                                 int lcap, const double4 *x, double4 *xn,
                                 double *v, double *f, const int *cnt,
                                 const int *slab, const LJParams *Pd,
                                 const PairCoef &c, const MassTable &mt,
                                 const StepCtl &ctl, double *partial )
{
    switch ( G )
    {
    case 1:
        launch_full_t<1, PREC, ONE, FUSED>( nb, st, n, scap, lcap, x, xn, v, f,
                                            cnt, slab, Pd, c, mt, ctl, partial );
        break;
    case 2:
        launch_full_t<2, PREC, ONE, FUSED>( nb, st, n, scap, lcap, x, xn, v, f,
                                            cnt, slab, Pd, c, mt, ctl, partial );
        break;
    case 4:
        launch_full_t<4, PREC, ONE, FUSED>( nb, st, n, scap, lcap, x, xn, v, f,
                                            cnt, slab, Pd, c, mt, ctl, partial );
        break;
    case 8:
        launch_full_t<8, PREC, ONE, FUSED>( nb, st, n, scap, lcap, x, xn, v, f,
                                            cnt, slab, Pd, c, mt, ctl, partial );
        break;
    default:
        set_error( "unsupported lanes per atom %d", G );
        return CMD_ERR_ARG;
    }
    CMD_LAUNCH_CHECK();
    return CMD_OK;
}

cmd_status lj_full( const cmd_force *F, cmd_system *S, const cmd_neigh *N,
                    const StepCtl &ctl, double4 *x_next, double *partial,
                    int *nblocks_out )
{
    const int n = S->n_local;
    const int G = N->lanes;
    const int nb = blocks_for( n, G );
    if ( nblocks_out )
        *nblocks_out = nb;
    if ( n == 0 )
        return CMD_OK;
    CMD_REQUIRE( N->n_rows == n, CMD_ERR_STATE,
                 "neighbor list has %d rows but the system has %d local atoms",
                 N->n_rows, n );
    const LJParams P = lj_params( F );
    const PairCoef c = one_coef( P );
    LJParams *Pd = nullptr;
    const bool one = F->ntypes == 1;
    if ( !one )
        CMD_TRY( upload_params( F, S->stream, &Pd ) );
    const bool fused = ( ctl.do_final || ctl.thermo );
    const MassTable mt = S->mass_table();
#define CMD_FULL( PREC, ONE, FUS )                                             \
    return launch_full_g<PREC, ONE, FUS>( G, nb, S->stream, n, S->cap, N->cap, \
                                          S->x.p, x_next, S->v.p, S->f.p,      \
                                          N->counts.p, N->slab.p, Pd, c, mt,   \
                                          ctl, partial )
    if ( F->prec == CMD_PREC_FP64 )
    {
        if ( one )
        {
            if ( fused )
                CMD_FULL( CMD_PREC_FP64, true, true );
            CMD_FULL( CMD_PREC_FP64, true, false );
        }
        if ( fused )
            CMD_FULL( CMD_PREC_FP64, false, true );
        CMD_FULL( CMD_PREC_FP64, false, false );
    }
    if ( one )
    {
        if ( fused )
            CMD_FULL( CMD_PREC_MIXED, true, true );
        CMD_FULL( CMD_PREC_MIXED, true, false );
    }
    if ( fused )
        CMD_FULL( CMD_PREC_MIXED, false, true );
    CMD_FULL( CMD_PREC_MIXED, false, false );
#undef CMD_FULL
}

template <int G>
static void launch_half_t( const cmd_force *F, cmd_system *S, const cmd_neigh *N,
                           int nb, const LJParams *Pd, const PairCoef &c,
                           int thermo, double *partial )
{
    const int n = S->n_local;
    const bool one = F->ntypes == 1;
    if ( F->prec == CMD_PREC_FP64 )
    {
        if ( one )
            k_lj_half<G, CMD_PREC_FP64, true><<<nb, TPB, 0, S->stream>>>(
                n, S->cap, N->cap, S->x.p, S->f.p, N->counts.p, N->slab.p, Pd,
                c, thermo, partial );
        else
            k_lj_half<G, CMD_PREC_FP64, false><<<nb, TPB, 0, S->stream>>>(
                n, S->cap, N->cap, S->x.p, S->f.p, N->counts.p, N->slab.p, Pd,
                c, thermo, partial );
    }
    else
    {
        if ( one )
            k_lj_half<G, CMD_PREC_MIXED, true><<<nb, TPB, 0, S->stream>>>(
                n, S->cap, N->cap, S->x.p, S->f.p, N->counts.p, N->slab.p, Pd,
                c, thermo, partial );
        else
            k_lj_half<G, CMD_PREC_MIXED, false><<<nb, TPB, 0, S->stream>>>(
                n, S->cap, N->cap, S->x.p, S->f.p, N->counts.p, N->slab.p, Pd,
                c, thermo, partial );
    }
}

cmd_status lj_half( const cmd_force *F, cmd_system *S, const cmd_neigh *N,
                    int thermo, double *partial, int *nblocks_out )
{
    const int n = S->n_local;
    const int G = N->lanes;
    const int nb = blocks_for( n, G );
    if ( nblocks_out )
        *nblocks_out = nb;
    if ( n == 0 )
        return CMD_OK;
    CMD_REQUIRE( N->n_rows == n, CMD_ERR_STATE, "neighbor list is stale" );
    const LJParams P = lj_params( F );
    const PairCoef c = one_coef( P );
    LJParams *Pd = nullptr;
    if ( F->ntypes != 1 )
        CMD_TRY( upload_params( F, S->stream, &Pd ) );
    switch ( G )
    {
    case 1:
        launch_half_t<1>( F, S, N, nb, Pd, c, thermo, partial );
        break;
    case 2:
        launch_half_t<2>( F, S, N, nb, Pd, c, thermo, partial );
        break;
    case 4:
        launch_half_t<4>( F, S, N, nb, Pd, c, thermo, partial );
        break;
    case 8:
        launch_half_t<8>( F, S, N, nb, Pd, c, thermo, partial );
        break;
    default:
        set_error( "unsupported lanes per atom %d", G );
        return CMD_ERR_ARG;
    }
    CMD_LAUNCH_CHECK();
    return CMD_OK;
}

cmd_status lj_energy_virial( const cmd_force *F, cmd_system *S,
                             const cmd_neigh *N, double *pe, double *w )
{
    const int n = S->n_local;
    const int G = N->lanes;
    const int nb = blocks_for( n, G );
    CMD_REQUIRE( N->n_rows == n, CMD_ERR_STATE, "neighbor list is stale" );
    CMD_TRY( S->sc.dpart.reserve( 4 * (size_t)nb ) );
    CMD_TRY( S->sc.dres.reserve( 8 ) );
    const LJParams P = lj_params( F );
    const PairCoef c = one_coef( P );
    LJParams *Pd = nullptr;
    const bool one = F->ntypes == 1;
    if ( !one )
        CMD_TRY( upload_params( F, S->stream, &Pd ) );
#define CMD_EN( GG )                                                          \
    if ( one )                                                                \
        k_lj_energy<GG, true><<<nb, TPB, 0, S->stream>>>(                     \
            n, N->cap, S->x.p, N->counts.p, N->slab.p, Pd, c, N->half,        \
            S->sc.dpart.p );                                                  \
    else                                                                      \
        k_lj_energy<GG, false><<<nb, TPB, 0, S->stream>>>(                    \
            n, N->cap, S->x.p, N->counts.p, N->slab.p, Pd, c, N->half,        \
            S->sc.dpart.p )
    if ( G == 1 )
    {
        CMD_EN( 1 );
    }
    else if ( G == 2 )
    {
        CMD_EN( 2 );
    }
    else if ( G == 4 )
    {
        CMD_EN( 4 );
    }
    else
    {
        CMD_EN( 8 );
    }
#undef CMD_EN
    CMD_LAUNCH_CHECK();
    CMD_TRY( reduce_partials( S->sc.dpart.p, nb, S->sc.dres.p, S->stream ) );
    double h[4];
    CMD_HIP( hipMemcpyAsync( h, S->sc.dres.p, sizeof( h ),
                             hipMemcpyDeviceToHost, S->stream ) );
    CMD_HIP( hipStreamSynchronize( S->stream ) );
    if ( pe )
        *pe = h[1];
    if ( w )
        *w = h[2];
    return CMD_OK;
}

// Build the list in each layout (1, 2, 4, 8 lanes per atom), time the force
// kernel on it and keep the fastest; the list is left built in that layout.
cmd_status autotune_lanes( cmd_force *F, cmd_system *S, cmd_neigh *N )
{
    const int cands[4] = { 1, 2, 4, 8 };
    StepCtl ctl{};
    hipEvent_t e0, e1;
    CMD_HIP( hipEventCreate( &e0 ) );
    CMD_HIP( hipEventCreate( &e1 ) );
    float best = 1e30f;
    int best_g = 1;
    cmd_status st = CMD_OK;
    for ( int c = 0; c < 4 && st == CMD_OK; ++c )
    {
        const int G = cands[c];
        N->lanes = G;
        st = neigh_build( N, S );
        if ( st != CMD_OK )
            break;
        st = lj_full( F, S, N, ctl, nullptr, nullptr, nullptr );
        if ( st != CMD_OK )
            break;
        (void)hipEventRecord( e0, S->stream );
        for ( int r = 0; r < 3; ++r )
            (void)lj_full( F, S, N, ctl, nullptr, nullptr, nullptr );
        (void)hipEventRecord( e1, S->stream );
        (void)hipEventSynchronize( e1 );
        float ms = 0.0f;
        (void)hipEventElapsedTime( &ms, e0, e1 );
        if ( std::getenv( "CMD_VERBOSE" ) )
            std::fprintf( stderr, "# autotune lanes=%d: %.4f ms/launch\n", G,
                          ms / 3.0f );
        if ( ms < best )
        {
            best = ms;
            best_g = G;
        }
    }
    (void)hipEventDestroy( e0 );
    (void)hipEventDestroy( e1 );
    CMD_TRY( st );
    F->lanes = best_g;
    F->tuned = true;
    if ( N->lanes != best_g )
    {
        N->lanes = best_g;
        CMD_TRY( neigh_build( N, S ) );
    }
    return CMD_OK;
}

cmd_status force_compute_plain( cmd_force *F, cmd_system *S, cmd_neigh *N )
{
    const int nt = S->n_local + S->n_ghost;
    // f = 0 over local and ghost atoms (cabanamd_impl.h:336-338)
    for ( int k = 0; k < 3; ++k )
        CMD_TRY( fill_f64( S->f.p + (size_t)k * S->cap, 0.0, nt, S->stream ) );
    if ( N->half )
        return lj_half( F, S, N, 0, nullptr, nullptr );
    StepCtl ctl{};
    return lj_full( F, S, N, ctl, nullptr, nullptr, nullptr );
}

} // namespace cmd

using namespace cmd;

extern "C" {

cmd_status cmd_force_lj_create( int ntypes, cmd_force **out )
{
    CMD_REQUIRE( out, CMD_ERR_ARG, "out is NULL" );
    CMD_REQUIRE( ntypes >= 1 && ntypes <= CMD_MAX_TYPES, CMD_ERR_ARG,
                 "ntypes %d outside [1,%d]", ntypes, CMD_MAX_TYPES );
    cmd_force *f = new cmd_force;
    f->kind = 0;
    f->ntypes = ntypes;
    for ( int a = 0; a < CMD_MAX_TYPES * CMD_MAX_TYPES; ++a )
        f->lj1[a] = f->lj2[a] = f->cutsq[a] = 0.0;
    f->lanes = 0;
    f->tuned = false;
    *out = f;
    return CMD_OK;
}

cmd_status cmd_force_destroy( cmd_force *f )
{
    delete f;
    return CMD_OK;
}

cmd_status cmd_force_lj_init_coeff( cmd_force *f, int i, int j, double eps,
                                    double sigma, double cut )
{
    CMD_REQUIRE( f, CMD_ERR_ARG, "force is NULL" );
    CMD_REQUIRE( i >= 1 && i <= f->ntypes && j >= 1 && j <= f->ntypes,
                 CMD_ERR_ARG, "pair_coeff types %d %d outside [1,%d]", i, j,
                 f->ntypes );
    const int nt = f->ntypes;
    const int a = ( i - 1 ) * nt + ( j - 1 ), b = ( j - 1 ) * nt + ( i - 1 );
    f->lj1[a] = 48.0 * eps * std::pow( sigma, 12.0 );
    f->lj2[a] = 24.0 * eps * std::pow( sigma, 6.0 );
    f->cutsq[a] = cut * cut;
    f->lj1[b] = f->lj1[a];
    f->lj2[b] = f->lj2[a];
    f->cutsq[b] = f->cutsq[a];
    return CMD_OK;
}

cmd_status cmd_force_set_precision( cmd_force *f, int prec )
{
    CMD_REQUIRE( f, CMD_ERR_ARG, "force is NULL" );
    CMD_REQUIRE( prec == CMD_PREC_FP64 || prec == CMD_PREC_MIXED, CMD_ERR_ARG,
                 "unknown precision %d", prec );
    f->prec = prec;
    return CMD_OK;
}

cmd_status cmd_force_set_lanes( cmd_force *f, int lanes )
{
    CMD_REQUIRE( f, CMD_ERR_ARG, "force is NULL" );
    CMD_REQUIRE( lanes == 0 || lanes == 1 || lanes == 2 || lanes == 4 ||
                     lanes == 8,
                 CMD_ERR_ARG, "lanes must be 0 (autotune), 1, 2, 4 or 8" );
    f->lanes = lanes;
    f->tuned = lanes != 0;
    return CMD_OK;
}

cmd_status cmd_force_get_lanes( const cmd_force *f, int *lanes )
{
    CMD_REQUIRE( f && lanes, CMD_ERR_ARG, "NULL argument" );
    *lanes = f->lanes;
    return CMD_OK;
}

cmd_status cmd_force_compute( cmd_force *f, cmd_system *s, cmd_neigh *n )
{
    CMD_REQUIRE( f && s && n, CMD_ERR_ARG, "NULL argument" );
    CMD_REQUIRE( n->n_rows == s->n_local, CMD_ERR_STATE,
                 "neighbor list has %d rows but the system has %d local atoms",
                 n->n_rows, s->n_local );
    return force_compute_plain( f, s, n );
}

cmd_status cmd_force_energy( cmd_force *f, cmd_system *s, cmd_neigh *n,
                             double *pe )
{
    CMD_REQUIRE( f && s && n && pe, CMD_ERR_ARG, "NULL argument" );
    return lj_energy_virial( f, s, n, pe, nullptr );
}

cmd_status cmd_force_virial( cmd_force *f, cmd_system *s, cmd_neigh *n,
                             double *w )
{
    CMD_REQUIRE( f && s && n && w, CMD_ERR_ARG, "NULL argument" );
    return lj_energy_virial( f, s, n, nullptr, w );
}

} // extern "C"
