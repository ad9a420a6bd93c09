// lj.hip -- 12-6 Lennard-Jones on CSR Verlet lists.
//
// Restates force_types/force_lj_cabana_neigh_impl.h (CabanaMD):
//   init_coeff :62-89   lj1 = 48 eps sigma^12, lj2 = 24 eps sigma^6, cutsq
//   full force :151-203 (pair force applied for rsq < cutsq, unshifted)
//   half force :205-259 (the negated pair force also lands on j)
//   energy     :261-377 (shifted; full weight 1/2, half 1 or 1/2 for ghost j)
//
// Kernel mapping: G lanes cooperate on one atom's CSR row (G = 1..16, the
// "lanes per atom" the autotuner picks); partial sums are combined with a
// butterfly inside the G-lane group.  Block indices are remapped so each of
// the 8 XCDs walks a contiguous range of cell-sorted atoms, which keeps the
// x_j gathers of neighbouring blocks in the same L2.
#include "internal.h"

#include <algorithm>
#include <cmath>

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

// One pair.  Returns the force factor (F_i = d * fpair) and adds the pair
// energy (unweighted, shifted) to e when ENERGY.
template <int PREC, bool ENERGY>
__device__ __forceinline__ double lj_pair( double rsq, double lj1, double lj2,
                                           float lj1f, float lj2f,
                                           double eshift, double &e )
{
    if ( PREC == CMD_PREC_FP64 )
    {
        const double r2inv = 1.0 / rsq;
        const double r6inv = r2inv * r2inv * r2inv;
        const double fpair = ( r6inv * ( lj1 * r6inv - lj2 ) ) * r2inv;
        if ( ENERGY )
            e += r6inv * ( 0.5 * lj1 * r6inv - lj2 ) / 6.0 - eshift;
        return fpair;
    }
    else
    {
        const float r2f = (float)rsq;
        const float r2inv = __builtin_amdgcn_rcpf( r2f );
        const float r6inv = r2inv * r2inv * r2inv;
        const float fpair = ( r6inv * __builtin_fmaf( lj1f, r6inv, -lj2f ) ) * r2inv;
        if ( ENERGY )
            e += (double)( r6inv * __builtin_fmaf( 0.5f * lj1f, r6inv, -lj2f ) *
                           ( 1.0f / 6.0f ) ) -
                 eshift;
        return (double)fpair;
    }
}

template <int G>
__device__ __forceinline__ double group_sum( double v )
{
#pragma unroll
    for ( int o = G / 2; o > 0; o >>= 1 )
        v += __shfl_xor( v, o, G );
    return v;
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

// Full list.  Without FUSED: f_i = sum_j d_ij fpair (ghost rows untouched).
// With FUSED the same thread then finishes velocity-Verlet for atom i
// (final half kick, thermo partials) and, when asked, starts the next step
// (half kick, drift, wrap) writing positions to x_next.
template <int G, int PREC, bool ONE_TYPE, bool FUSED>
__global__ __launch_bounds__( TPB ) void k_lj_full(
    int n_local, int cap, const double4 *__restrict__ x,
    double4 *__restrict__ x_next, double *__restrict__ v,
    double *__restrict__ f, const int64_t *__restrict__ off,
    const int *__restrict__ idx, const LJParams *__restrict__ Pd, double lj1s,
    double lj2s, double cuts, double eshs, MassTable mt, StepCtl ctl,
    double *__restrict__ partial )
{
    const int b = xcd_remap( blockIdx.x, gridDim.x );
    const int sub = threadIdx.x % G;
    const int i = ( b * TPB + (int)threadIdx.x ) / G;
    const bool active = i < n_local;
    const bool energy = FUSED && ctl.thermo;
    double fx = 0.0, fy = 0.0, fz = 0.0, pe = 0.0, vir = 0.0;
    double4 xi = make_double4( 0.0, 0.0, 0.0, 0.0 );
    if ( active )
    {
        xi = x[i];
        const int ti = (int)xi.w;
        const int64_t e = off[i + 1];
        for ( int64_t k = off[i] + sub; k < e; k += G )
        {
            const int j = idx[k];
            const double4 xj = x[j];
            const double dx = xi.x - xj.x;
            const double dy = xi.y - xj.y;
            const double dz = xi.z - xj.z;
            const double rsq = dx * dx + dy * dy + dz * dz;
            double lj1 = lj1s, lj2 = lj2s, cs = cuts, esh = eshs;
            float lj1f = (float)lj1s, lj2f = (float)lj2s;
            if ( !ONE_TYPE )
            {
                const int t = ti * Pd->ntypes + (int)xj.w;
                lj1 = Pd->lj1[t];
                lj2 = Pd->lj2[t];
                cs = Pd->cutsq[t];
                esh = Pd->eshift[t];
                lj1f = Pd->lj1f[t];
                lj2f = Pd->lj2f[t];
            }
            if ( rsq < cs )
            {
                double fpair;
                if ( energy )
                {
                    fpair = lj_pair<PREC, true>( rsq, lj1, lj2, lj1f, lj2f,
                                                 esh, pe );
                    vir += rsq * fpair;
                }
                else
                    fpair = lj_pair<PREC, false>( rsq, lj1, lj2, lj1f, lj2f,
                                                  esh, pe );
                if ( PREC == CMD_PREC_FP64 )
                {
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
        }
    }
    if ( G > 1 )
    {
        fx = group_sum<G>( fx );
        fy = group_sum<G>( fy );
        fz = group_sum<G>( fz );
    }
    double mv2 = 0.0;
    if ( active && sub == 0 )
    {
        if ( !FUSED || ctl.write_f )
        {
            f[i] = fx;
            f[cap + i] = fy;
            f[2 * cap + i] = fz;
        }
        if ( FUSED && ctl.do_final )
        {
            const double m = mt.m[(int)xi.w];
            const double dtfm = ctl.dtf / m;
            double vx = v[i] + dtfm * fx;
            double vy = v[cap + i] + dtfm * fy;
            double vz = v[2 * cap + i] + dtfm * fz;
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
            v[cap + i] = vy;
            v[2 * cap + i] = vz;
        }
    }
    if ( energy )
        block_sum4( mv2, 0.5 * pe, 0.5 * vir, 0.0, partial + 4 * blockIdx.x );
}

// Half list: each listed pair once; the reaction goes to j by atomics
// (j may be a ghost, folded later by update_force).
template <int PREC, bool ONE_TYPE>
__global__ __launch_bounds__( TPB ) void k_lj_half(
    int n_local, int cap, const double4 *__restrict__ x,
    double *__restrict__ f, const int64_t *__restrict__ off,
    const int *__restrict__ idx, const LJParams *__restrict__ Pd, double lj1s,
    double lj2s, double cuts, double eshs, int thermo,
    double *__restrict__ partial )
{
    const int b = xcd_remap( blockIdx.x, gridDim.x );
    const int i = b * TPB + (int)threadIdx.x;
    double pe = 0.0, vir = 0.0;
    if ( i < n_local )
    {
        const double4 xi = x[i];
        const int ti = (int)xi.w;
        double fx = 0.0, fy = 0.0, fz = 0.0;
        const int64_t e = off[i + 1];
        for ( int64_t k = off[i]; k < e; ++k )
        {
            const int j = idx[k];
            const double4 xj = x[j];
            const double dx = xi.x - xj.x;
            const double dy = xi.y - xj.y;
            const double dz = xi.z - xj.z;
            const double rsq = dx * dx + dy * dy + dz * dz;
            double lj1 = lj1s, lj2 = lj2s, cs = cuts, esh = eshs;
            float lj1f = (float)lj1s, lj2f = (float)lj2s;
            if ( !ONE_TYPE )
            {
                const int t = ti * Pd->ntypes + (int)xj.w;
                lj1 = Pd->lj1[t];
                lj2 = Pd->lj2[t];
                cs = Pd->cutsq[t];
                esh = Pd->eshift[t];
                lj1f = Pd->lj1f[t];
                lj2f = Pd->lj2f[t];
            }
            if ( rsq < cs )
            {
                double ep = 0.0;
                const double fpair =
                    thermo ? lj_pair<PREC, true>( rsq, lj1, lj2, lj1f, lj2f,
                                                  esh, ep )
                           : lj_pair<PREC, false>( rsq, lj1, lj2, lj1f, lj2f,
                                                   esh, ep );
                const double wgt = ( j < n_local ) ? 1.0 : 0.5;
                pe += wgt * ep;
                vir += rsq * fpair;
                const double ax = dx * fpair, ay = dy * fpair, az = dz * fpair;
                fx += ax;
                fy += ay;
                fz += az;
                unsafeAtomicAdd( &f[j], -ax );
                unsafeAtomicAdd( &f[cap + j], -ay );
                unsafeAtomicAdd( &f[2 * cap + j], -az );
            }
        }
        unsafeAtomicAdd( &f[i], fx );
        unsafeAtomicAdd( &f[cap + i], fy );
        unsafeAtomicAdd( &f[2 * cap + i], fz );
    }
    if ( thermo )
        block_sum4( 0.0, pe, vir, 0.0, partial + 4 * blockIdx.x );
}

// compute_energy (reference weights) and pair virial, FP64.
template <bool ONE_TYPE>
__global__ __launch_bounds__( TPB ) void k_lj_energy(
    int n_local, const double4 *__restrict__ x,
    const int64_t *__restrict__ off, const int *__restrict__ idx,
    const LJParams *__restrict__ Pd, double lj1s, double lj2s, double cuts,
    double eshs, int half, double *__restrict__ partial )
{
    const int i = blockIdx.x * TPB + (int)threadIdx.x;
    double pe = 0.0, vir = 0.0;
    if ( i < n_local )
    {
        const double4 xi = x[i];
        const int ti = (int)xi.w;
        for ( int64_t k = off[i]; k < off[i + 1]; ++k )
        {
            const int j = idx[k];
            const double4 xj = x[j];
            const double dx = xi.x - xj.x;
            const double dy = xi.y - xj.y;
            const double dz = xi.z - xj.z;
            const double rsq = dx * dx + dy * dy + dz * dz;
            double lj1 = lj1s, lj2 = lj2s, cs = cuts, esh = eshs;
            if ( !ONE_TYPE )
            {
                const int t = ti * Pd->ntypes + (int)xj.w;
                lj1 = Pd->lj1[t];
                lj2 = Pd->lj2[t];
                cs = Pd->cutsq[t];
                esh = Pd->eshift[t];
            }
            if ( rsq < cs )
            {
                const double fac = half ? ( ( j < n_local ) ? 1.0 : 0.5 ) : 0.5;
                const double r2inv = 1.0 / rsq;
                const double r6inv = r2inv * r2inv * r2inv;
                pe += fac * r6inv * ( 0.5 * lj1 * r6inv - lj2 ) / 6.0;
                pe -= fac * esh;
                const double fpair = ( r6inv * ( lj1 * r6inv - lj2 ) ) * r2inv;
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

template <int G, int PREC, bool ONE, bool FUSED>
static void launch_full_g( int nb, hipStream_t st, int n, int cap,
                           const double4 *x, double4 *xn, double *v, double *f,
                           const int64_t *off, const int *idx,
                           const LJParams *Pd, const LJParams &P,
                           const MassTable &mt, const StepCtl &ctl,
                           double *partial )
{
    k_lj_full<G, PREC, ONE, FUSED><<<nb, TPB, 0, st>>>(
        n, cap, x, xn, v, f, off, idx, Pd, P.lj1[0], P.lj2[0], P.cutsq[0],
        P.eshift[0], mt, ctl, partial );
}

template <int PREC, bool ONE, bool FUSED>
static cmd_status launch_full_p( int G, int nb, hipStream_t st, int n, int cap,
                                 const double4 *x, double4 *xn, double *v,
                                 double *f, const int64_t *off, const int *idx,
                                 const LJParams *Pd, const LJParams &P,
                                 const MassTable &mt, const StepCtl &ctl,
                                 double *partial )
{
    switch ( G )
    {
    case 1:
        launch_full_g<1, PREC, ONE, FUSED>( nb, st, n, cap, x, xn, v, f, off,
                                            idx, Pd, P, mt, ctl, partial );
        break;
    case 2:
        launch_full_g<2, PREC, ONE, FUSED>( nb, st, n, cap, x, xn, v, f, off,
                                            idx, Pd, P, mt, ctl, partial );
        break;
    case 4:
        launch_full_g<4, PREC, ONE, FUSED>( nb, st, n, cap, x, xn, v, f, off,
                                            idx, Pd, P, mt, ctl, partial );
        break;
    case 8:
        launch_full_g<8, PREC, ONE, FUSED>( nb, st, n, cap, x, xn, v, f, off,
                                            idx, Pd, P, mt, ctl, partial );
        break;
    case 16:
        launch_full_g<16, PREC, ONE, FUSED>( nb, st, n, cap, x, xn, v, f, off,
                                             idx, Pd, P, mt, ctl, partial );
        break;
    default:
        set_error( "unsupported lanes per atom %d", G );
        return CMD_ERR_ARG;
    }
    CMD_LAUNCH_CHECK();
    return CMD_OK;
}

cmd_status lj_full( const cmd_force *F, cmd_system *S, const cmd_neigh *N,
                    int lanes, const StepCtl &ctl, double4 *x_next,
                    double *partial, int *nblocks_out )
{
    const int n = S->n_local;
    const int G = lanes;
    const int atoms_per_block = TPB / G;
    const int nb = std::max( 1, ( n + atoms_per_block - 1 ) / atoms_per_block );
    if ( nblocks_out )
        *nblocks_out = nb;
    if ( n == 0 )
        return CMD_OK;
    LJParams P = lj_params( F );
    LJParams *Pd = nullptr;
    const bool one = F->ntypes == 1;
    if ( !one )
        CMD_TRY( upload_params( F, S->stream, &Pd ) );
    const bool fused = ( ctl.do_final || ctl.thermo );
    const MassTable mt = S->mass_table();
#define CMD_FULL( PREC, ONE, FUS )                                            \
    return launch_full_p<PREC, ONE, FUS>( G, nb, S->stream, n, S->cap, S->x.p, \
                                          x_next, S->v.p, S->f.p,             \
                                          N->offsets.p, N->idx.p, Pd, P, mt,  \
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

cmd_status lj_half( const cmd_force *F, cmd_system *S, const cmd_neigh *N,
                    int thermo, double *partial, int *nblocks_out )
{
    const int n = S->n_local;
    const int nb = std::max( 1, ( n + TPB - 1 ) / TPB );
    if ( nblocks_out )
        *nblocks_out = nb;
    if ( n == 0 )
        return CMD_OK;
    LJParams P = lj_params( F );
    LJParams *Pd = nullptr;
    const bool one = F->ntypes == 1;
    if ( !one )
        CMD_TRY( upload_params( F, S->stream, &Pd ) );
#define CMD_HALF( PREC, ONE )                                                  \
    k_lj_half<PREC, ONE><<<nb, TPB, 0, S->stream>>>(                           \
        n, S->cap, S->x.p, S->f.p, N->offsets.p, N->idx.p, Pd, P.lj1[0],       \
        P.lj2[0], P.cutsq[0], P.eshift[0], thermo, partial )
    if ( F->prec == CMD_PREC_FP64 )
    {
        if ( one )
            CMD_HALF( CMD_PREC_FP64, true );
        else
            CMD_HALF( CMD_PREC_FP64, false );
    }
    else
    {
        if ( one )
            CMD_HALF( CMD_PREC_MIXED, true );
        else
            CMD_HALF( CMD_PREC_MIXED, false );
    }
#undef CMD_HALF
    CMD_LAUNCH_CHECK();
    return CMD_OK;
}

cmd_status lj_energy_virial( const cmd_force *F, cmd_system *S,
                             const cmd_neigh *N, double *pe, double *w )
{
    const int n = S->n_local;
    const int nb = std::max( 1, ( n + TPB - 1 ) / TPB );
    CMD_TRY( S->sc.dpart.reserve( 4 * (size_t)nb ) );
    CMD_TRY( S->sc.dres.reserve( 8 ) );
    LJParams P = lj_params( F );
    LJParams *Pd = nullptr;
    const bool one = F->ntypes == 1;
    if ( !one )
        CMD_TRY( upload_params( F, S->stream, &Pd ) );
    if ( one )
        k_lj_energy<true><<<nb, TPB, 0, S->stream>>>(
            n, S->x.p, N->offsets.p, N->idx.p, Pd, P.lj1[0], P.lj2[0],
            P.cutsq[0], P.eshift[0], N->half, S->sc.dpart.p );
    else
        k_lj_energy<false><<<nb, TPB, 0, S->stream>>>(
            n, S->x.p, N->offsets.p, N->idx.p, Pd, P.lj1[0], P.lj2[0],
            P.cutsq[0], P.eshift[0], N->half, S->sc.dpart.p );
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

// Time each lanes-per-atom variant on the current list and keep the fastest.
static cmd_status autotune_lanes( cmd_force *F, cmd_system *S,
                                  const cmd_neigh *N )
{
    const int cands[5] = { 1, 2, 4, 8, 16 };
    StepCtl ctl{};
    hipEvent_t e0, e1;
    CMD_HIP( hipEventCreate( &e0 ) );
    CMD_HIP( hipEventCreate( &e1 ) );
    float best = 1e30f;
    int best_g = 4;
    for ( int c = 0; c < 5; ++c )
    {
        const int G = cands[c];
        cmd_status st = lj_full( F, S, N, G, ctl, nullptr, nullptr, nullptr );
        if ( st != CMD_OK )
        {
            (void)hipEventDestroy( e0 );
            (void)hipEventDestroy( e1 );
            return st;
        }
        (void)hipEventRecord( e0, S->stream );
        for ( int r = 0; r < 3; ++r )
            (void)lj_full( F, S, N, G, ctl, nullptr, nullptr, nullptr );
        (void)hipEventRecord( e1, S->stream );
        (void)hipEventSynchronize( e1 );
        float ms = 0.0f;
        (void)hipEventElapsedTime( &ms, e0, e1 );
        if ( ms < best )
        {
            best = ms;
            best_g = G;
        }
    }
    (void)hipEventDestroy( e0 );
    (void)hipEventDestroy( e1 );
    F->lanes = best_g;
    F->tuned = true;
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
    if ( !F->tuned && F->lanes == 0 )
        CMD_TRY( autotune_lanes( F, S, N ) );
    StepCtl ctl{};
    return lj_full( F, S, N, F->lanes, ctl, nullptr, nullptr, nullptr );
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
    f->tuned = false;
    return CMD_OK;
}

cmd_status cmd_force_set_lanes( cmd_force *f, int lanes )
{
    CMD_REQUIRE( f, CMD_ERR_ARG, "force is NULL" );
    CMD_REQUIRE( lanes == 0 || lanes == 1 || lanes == 2 || lanes == 4 ||
                     lanes == 8 || lanes == 16,
                 CMD_ERR_ARG, "lanes must be 0 (autotune), 1, 2, 4, 8 or 16" );
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
    CMD_REQUIRE( n->n_rows == s->n_local, CMD_ERR_STATE,
                 "neighbor list is stale" );
    return lj_energy_virial( f, s, n, pe, nullptr );
}

cmd_status cmd_force_virial( cmd_force *f, cmd_system *s, cmd_neigh *n,
                             double *w )
{
    CMD_REQUIRE( f && s && n && w, CMD_ERR_ARG, "NULL argument" );
    CMD_REQUIRE( n->n_rows == s->n_local, CMD_ERR_STATE,
                 "neighbor list is stale" );
    return lj_energy_virial( f, s, n, nullptr, w );
}

} // extern "C"
