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
#include "tiles_dev.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

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
        // outside the cutoff the reciprocal of +inf is 0, so fp = -0 and
        // the accumulation adds nothing: one select on the way in
        const float r2inv = __builtin_amdgcn_rcpf( in ? (float)rsq : __builtin_inff() );
        const float r6inv = r2inv * r2inv * r2inv;
        const float fp = ( r6inv * __builtin_fmaf( c.lj1f, r6inv, -c.lj2f ) ) * r2inv;
        const double fpair = (double)fp;
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

// Sum of four values over the block into out[0..3]; red holds 4 x 8
// doubles of LDS (up to 8 waves per block).
__device__ void block_sum4_in( double a, double b, double c, double d,
                               double *__restrict__ out, double ( *red )[8] );

__device__ void block_sum4( double a, double b, double c, double d,
                            double *__restrict__ out )
{
    __shared__ double red[4][8];
    block_sum4_in( a, b, c, d, out, red );
}

__device__ void block_sum4_in( double a, double b, double c, double d,
                               double *__restrict__ out, double ( *red )[8] )
{
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
        for ( int w = 0; w < (int)( blockDim.x >> 6 ); ++w )
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
// body gets (j, x_j).  Software-pipelined: the index loads of the next four
// rows are issued before the gathers and arithmetic of the current four, so
// the index stream's latency overlaps the gathers.  Index loads are
// non-temporal: the stream is read once and must not evict the reused x_j
// lines from L2.
// Position gather: x, y, z only when the type slot is not needed (one atom
// type), which saves a quarter of the address/data traffic per pair.
template <bool XYZ_ONLY>
__device__ __forceinline__ double4 gather_pos( const double4 *__restrict__ x,
                                               int j )
{
    if ( XYZ_ONLY )
    {
        const double2 a = *reinterpret_cast<const double2 *>( x + j );
        const double c = reinterpret_cast<const double *>( x + j )[2];
        return make_double4( a.x, a.y, c, 0.0 );
    }
    return x[j];
}

template <int G, bool XYZ_ONLY, class Body>
__device__ __forceinline__ void walk_row( const int *__restrict__ row, int nr,
                                          const double4 *__restrict__ x,
                                          Body body )
{
    constexpr int U = 4;
    int jn[U];
#pragma unroll
    for ( int u = 0; u < U; ++u )
        jn[u] = ( u < nr ) ? __builtin_nontemporal_load( &row[u * 64] ) : -1;
    for ( int r = 0; r < nr; r += U )
    {
        int jc[U];
#pragma unroll
        for ( int u = 0; u < U; ++u )
        {
            jc[u] = jn[u];
            const int rn = r + U + u;
            jn[u] = ( rn < nr ) ? __builtin_nontemporal_load( &row[rn * 64] ) : -1;
        }
        double4 p[U];
#pragma unroll
        for ( int u = 0; u < U; ++u )
            p[u] = gather_pos<XYZ_ONLY>( x, jc[u] >= 0 ? jc[u] : 0 );
#pragma unroll
        for ( int u = 0; u < U; ++u )
            if ( jc[u] >= 0 )
                body( jc[u], p[u] );
    }
}

// After the force of atom i is known: store f, finish this step's
// velocity (final half kick) and start the next one (half kick, drift, wrap)
// -- integrator_nve.h:91-110 and comm_mpi.h:141-168 -- adding m v^2 to mv2.
// With G >= 4 lanes per atom lane s handles component s; otherwise lane 0
// does all three.  Both orders run the same operations on the same values.
// vin, when given, holds the velocity components the lane would load
// (component s, or all three), read early so their latency is hidden.
template <int G, bool FUSED>
__device__ __forceinline__ void finish_atom( bool active, int s, int i,
                                             const double4 &xi, double fx,
                                             double fy, double fz, int scap,
                                             double4 *__restrict__ x_next,
                                             double *__restrict__ v,
                                             double *__restrict__ f,
                                             const MassTable &mt,
                                             const StepCtl &ctl, double &mv2,
                                             const double *vin = nullptr,
                                             int img_b = 0, int img_e = 0 )
{
    if ( G >= 4 )
    {
        // every lane of the atom holds the totals: lane s finishes component
        // s (same operations as the one-lane epilogue, so identical bits)
        const int c = s;
        if ( active && c < 4 )
        {
            const double fc = c == 0 ? fx : ( c == 1 ? fy : fz );
            if ( c < 3 && ( !FUSED || ctl.write_f ) )
                __builtin_nontemporal_store( fc, &f[c * scap + i] );
            if ( FUSED && ctl.do_final )
            {
                double *xn = reinterpret_cast<double *>( x_next + i );
                if ( c < 3 )
                {
                    const double m = mt.m[(int)xi.w];
                    const double dtfm = ctl.dtf / m;
                    double vc = ( vin ? vin[0] : v[c * scap + i] ) + dtfm * fc;
                    mv2 += vc * vc * m;
                    if ( ctl.do_next_initial )
                    {
                        vc = vc + dtfm * fc;
                        const double xc = c == 0 ? xi.x : ( c == 1 ? xi.y : xi.z );
                        double pc = xc + ctl.dtv * vc;
                        if ( ctl.wrap_next )
                        {
                            const double q = pc;
                            if ( q > ctl.L[c] )
                                pc = q - ctl.L[c];
                            if ( q < 0 )
                                pc = q + ctl.L[c];
                        }
                        xn[c] = pc;
                        for ( int m = img_b; m < img_e; ++m )
                        {
                            // the ghost refresh's shift of this component
                            const unsigned e = (unsigned)ctl.img_pk[m];
                            const unsigned code = e >> ( IMG_CODE_SHIFT + 2 * c );
                            double g = pc;
                            if ( code & 1 )
                                g = g - ctl.L[c];
                            if ( code & 2 )
                                g = g + ctl.L[c];
                            reinterpret_cast<double *>( ctl.xg_next + ( e & IMG_SLOT_MASK ) )[c] = g;
                        }
                    }
                    v[c * scap + i] = vc;
                }
                else if ( ctl.do_next_initial )
                {
                    xn[3] = xi.w;
                    for ( int m = img_b; m < img_e; ++m )
                        reinterpret_cast<double *>(
                            ctl.xg_next + ( (unsigned)ctl.img_pk[m] & IMG_SLOT_MASK ) )[3] = xi.w;
                }
            }
        }
    }
    else if ( active && s == 0 )
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
            double vx = ( vin ? vin[0] : v[i] ) + dtfm * fx;
            double vy = ( vin ? vin[1] : v[scap + i] ) + dtfm * fy;
            double vz = ( vin ? vin[2] : v[2 * scap + i] ) + dtfm * fz;
            mv2 += ( vx * vx + vy * vy + vz * vz ) * m;
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
                for ( int m = img_b; m < img_e; ++m )
                {
                    // the ghost refresh's shifts (one per dimension)
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
            v[scap + i] = vy;
            v[2 * scap + i] = vz;
        }
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
        walk_row<G, ONE_TYPE>( row, nr, x, body );
    }
#pragma unroll
    for ( int o = APW; o < 64; o <<= 1 )
    {
        fx += __shfl_xor( fx, o, 64 );
        fy += __shfl_xor( fy, o, 64 );
        fz += __shfl_xor( fz, o, 64 );
    }
    double mv2 = 0.0;
    finish_atom<G, FUSED>( active, L.s, i, xi, fx, fy, fz, scap, x_next, v, f, mt,
                           ctl, mv2 );
    if ( energy )
        block_sum4( mv2, 0.5 * pe, 0.5 * vir, 0.0, partial + 4 * blockIdx.x );
}

// Tiled full list (tiles.hip): the block stages its tile's stencil
// neighbourhood in LDS and walks 16-bit slots, four rows per 8-byte load.
// The list pads every row of a slab wave up to the wave's longest row with
// the tile's sentinel slot (a position far outside any cutoff), so all lanes
// run the same number of iterations with no per-pair bounds test; inactive
// lanes sit at another far point.  A padded pair adds exact zeros.
// LDS keeps the tiled kernels at three or four waves per SIMD, so the
// compiler may use registers for more loads and pair chains in flight.
template <int G, int PREC, bool ONE_TYPE, bool FUSED, bool ENERGY, bool F32REC>
__global__ __launch_bounds__( 512 ) __attribute__( ( amdgpu_waves_per_eu( 2, 4 ) ) ) void k_tile_force(
    TileGrid T, const int *__restrict__ info, const int *__restrict__ woff,
    int lcap, int scap, int slot_cap, const double4 *__restrict__ x,
    const double4 *__restrict__ xg, double4 *__restrict__ x_next,
    double *__restrict__ v, double *__restrict__ f,
    const int *__restrict__ counts, const unsigned short *__restrict__ slab,
    const LJParams *__restrict__ Pd, PairCoef one, MassTable mt, StepCtl ctl,
    double *__restrict__ partial )
{
    extern __shared__ double s_rec[];
    __shared__ TileHead H;
    constexpr int APW = 64 / G;
    const int t = xcd_tile( blockIdx.x, gridDim.x );
    // staging waves issue first: a block's tile head and records are on
    // the critical path of all its waves (1 % faster than equal priority)
    __builtin_amdgcn_s_setprio( 2 );
    load_head( info, t, T.W, H );
    const char *s_byte = reinterpret_cast<const char *>( s_rec );
    const int *s_t = reinterpret_cast<const int *>( s_rec + 3 * slot_cap );
    const int R4 = ( ( lcap + G - 1 ) / G + 3 ) / 4;
    const int wbase = woff[t];
    const unsigned short *st = slab + (size_t)wbase * R4 * 256;
    const int nw = woff[t + 1] - wbase;
    const int nat = H.info[TI_NAT];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int nwb = blockDim.x >> 6;
    // four sentinel entries: the word a lane past the tile's atoms walks
    constexpr int REC = F32REC ? TILE_REC32 : TILE_REC;
    const unsigned long long sent4 =
        0x0001000100010001ull * (unsigned)( H.segoff[H.nseg] * REC );
    // FP32 records: coordinates about the centre of the staged region
    double ox = 0.0, oy = 0.0, oz = 0.0;
    if constexpr ( F32REC )
        tile_centre( T, t, ox, oy, oz );
    double pe = 0.0, vir = 0.0, mv2 = 0.0;
    // What a slab wave reads from global memory before its pair loop: the
    // atom, its row length and the first three list words.  The first
    // wave's reads are issued before the block stages its records, so both
    // sets of loads are in flight together.
    struct WaveIn
    {
        double4 xi;
        int n;
        unsigned long long q0, q1, q2;
    };
    auto lane_map = [&]( int w, int &gw, int &apw, int &sub, int &i, bool &active )
    {
        // gw lanes per atom in this slab wave (G, or more in a part-full wave)
        gw = wave_lanes( G, nat, w );
        apw = 64 / gw;
        const int a = lane % apw;
        sub = lane / apw;
        const int ta = w * APW + a;
        active = a < APW && ta < nat;
        // lanes past the atoms take the tile's first atom (nat >= 1 here)
        i = tile_atom( H, T.W, active ? ta : 0 );
    };
    auto row_of = [&]( int w ) -> const unsigned long long *
    {
        return reinterpret_cast<const unsigned long long *>(
            st + ( (size_t)w * R4 * 64 + lane ) * 4 );
    };
    auto prep = [&]( int w, WaveIn &P )
    {
        int gw, apw, sub, i;
        bool active;
        lane_map( w, gw, apw, sub, i, active );
        // loaded unconditionally (lanes past the atoms read the tile's first
        // atom) and replaced after staging, so no wait is needed here
        P.xi = x[i];
        P.n = counts[i];
        // the first three words of every lane's column (R4 >= 1 words each;
        // words past the row's end are loaded but not used)
        const unsigned long long *row8 = row_of( w );
        P.q0 = __builtin_nontemporal_load( row8 );
        __builtin_amdgcn_sched_barrier( 0 );
        P.q1 = __builtin_nontemporal_load( row8 + (size_t)min( 1, R4 - 1 ) * 64 );
        __builtin_amdgcn_sched_barrier( 0 );
        P.q2 = __builtin_nontemporal_load( row8 + (size_t)min( 2, R4 - 1 ) * 64 );
        __builtin_amdgcn_sched_barrier( 0 );
    };
    WaveIn P;
    if ( wid < nw )
        prep( wid, P );
    if constexpr ( F32REC )
        stage_records32( H, x, xg, reinterpret_cast<float4 *>( s_rec ), ox, oy, oz );
    else
        stage_records<!ONE_TYPE>( H, x, xg, s_rec, slot_cap );
    __builtin_amdgcn_s_setprio( 0 );
    for ( int w = wid; w < nw; w += nwb )
    {
        if ( w != wid )
            prep( w, P );
        int gw, apw, sub, i;
        bool active;
        lane_map( w, gw, apw, sub, i, active );
        const double4 xi =
            active ? P.xi : make_double4( -TILE_FAR, -TILE_FAR, -TILE_FAR, 0.0 );
        const int ti = (int)xi.w;
        const int n = active ? P.n : 0;
        double fx = 0.0, fy = 0.0, fz = 0.0;
        int nr = n > sub ? ( n - sub + gw - 1 ) / gw : 0;
#pragma unroll
        for ( int o = 32; o > 0; o >>= 1 )
            nr = max( nr, __shfl_xor( nr, o, 64 ) );
        nr = __builtin_amdgcn_readfirstlane( nr );
        const unsigned long long *row8 = row_of( w );
        // the velocities and ghost image range the epilogue needs, loaded
        // before the pair loop
        double vpre[3] = { 0.0, 0.0, 0.0 };
        int img_b = 0, img_e = 0;
        if ( FUSED && active && ctl.do_final )
        {
            if ( G >= 4 )
            {
                if ( sub < 3 )
                    vpre[0] = v[sub * scap + i];
            }
            else if ( sub == 0 )
            {
                vpre[0] = v[i];
                vpre[1] = v[scap + i];
                vpre[2] = v[2 * scap + i];
            }
            if ( ctl.xg_next && ( G >= 4 ? sub < 4 : sub == 0 ) )
            {
                img_b = ctl.img_start[i];
                img_e = ctl.img_start[i + 1];
            }
        }
        // the atom about the tile's centre for FP32 records (inactive lanes
        // at the far point, whose distance to every record overflows to
        // +inf and is outside every cutoff)
        const float xfx = active ? (float)( xi.x - ox ) : -(float)TILE_FAR;
        const float xfy = active ? (float)( xi.y - oy ) : -(float)TILE_FAR;
        const float xfz = active ? (float)( xi.z - oz ) : -(float)TILE_FAR;
        const float csf = (float)one.cs;
        // FP32 records: the pair in FP32 from the coordinate differences
        // on, the force summed in FP64
        auto body32 = [&]( const float4 &xj )
        {
            const float dx = xfx - xj.x;
            const float dy = xfy - xj.y;
            const float dz = xfz - xj.z;
            const float rsq = __builtin_fmaf( dx, dx, __builtin_fmaf( dy, dy, dz * dz ) );
            const PairCoef c = coef<ONE_TYPE>( Pd, one, ti, __float_as_int( xj.w ) );
            const bool in = rsq < ( ONE_TYPE ? csf : (float)c.cs );
            const float r2inv = __builtin_amdgcn_rcpf( in ? rsq : __builtin_inff() );
            const float r6inv = r2inv * r2inv * r2inv;
            const float fp = ( r6inv * __builtin_fmaf( c.lj1f, r6inv, -c.lj2f ) ) * r2inv;
            if ( ENERGY )
            {
                const float ep =
                    r6inv * __builtin_fmaf( 0.5f * c.lj1f, r6inv, -c.lj2f ) * ( 1.0f / 6.0f );
                pe += in ? (double)ep - c.esh : 0.0;
                // a far record's distance overflows to +inf in FP32
                vir += in ? (double)rsq * (double)fp : 0.0;
            }
            const double fpd = (double)fp;
            fx = __builtin_fma( (double)dx, fpd, fx );
            fy = __builtin_fma( (double)dy, fpd, fy );
            fz = __builtin_fma( (double)dz, fpd, fz );
        };
        auto body = [&]( const double4 &xj )
        {
            const double dx = xi.x - xj.x;
            const double dy = xi.y - xj.y;
            const double dz = xi.z - xj.z;
            const double rsq = ( PREC == CMD_PREC_FP64 )
                                   ? dx * dx + dy * dy + dz * dz
                                   : __builtin_fma( dx, dx, __builtin_fma( dy, dy, dz * dz ) );
            const PairCoef c = coef<ONE_TYPE>( Pd, one, ti, (int)xj.w );
            const double fpair = lj_pair<PREC, ENERGY>( rsq, c, pe, vir );
            accumulate<PREC>( fx, fy, fz, dx, dy, dz, fpair );
        };
        // four records per 8-byte word of byte offsets; all four LDS reads
        // are issued before the arithmetic that consumes them
        auto fetch = [&]( unsigned long long wd, int q ) -> double4
        {
            const unsigned off = (unsigned)( ( wd >> ( 16 * q ) ) & 0xffffull );
            const double *p = reinterpret_cast<const double *>( s_byte + off );
            return make_double4( p[0], p[1], p[2],
                                 ONE_TYPE ? 0.0 : (double)s_t[off / TILE_REC] );
        };
        // list words three iterations ahead: under load the memory latency
        // spans several iterations of all waves of a SIMD.  The loop is
        // unrolled by three so each word stays in its register until used
        // (a rotation through registers would wait for the newest load).
        // Every lane loads from its own column of the wave's slab, which
        // holds at least max(nq, 1) words for all 64 lanes; lanes past the
        // tile's atoms use the sentinel word in place of what their column
        // holds.  Loads are issued unconditionally (indexes past the last
        // word re-read it), so the count of loads in flight is the same on
        // every path and the wait before a use covers only that word.
        const int nq = ( nr + 3 ) / 4;
        const int klast = max( nq - 1, 0 );
        auto word = [&]( int k ) -> unsigned long long
        { return __builtin_nontemporal_load( row8 + (size_t)min( k, klast ) * 64 ); };
        auto fetch32 = [&]( unsigned long long wd, int q ) -> float4
        {
            const unsigned off = (unsigned)( ( wd >> ( 16 * q ) ) & 0xffffull );
            return *reinterpret_cast<const float4 *>( s_byte + off );
        };
        auto quad = [&]( unsigned long long wd )
        {
            const unsigned long long cur = active ? wd : sent4;
            if constexpr ( F32REC )
            {
                const float4 p0 = fetch32( cur, 0 );
                const float4 p1 = fetch32( cur, 1 );
                const float4 p2 = fetch32( cur, 2 );
                const float4 p3 = fetch32( cur, 3 );
                body32( p0 );
                body32( p1 );
                body32( p2 );
                body32( p3 );
            }
            else
            {
                const double4 p0 = fetch( cur, 0 );
                const double4 p1 = fetch( cur, 1 );
                const double4 p2 = fetch( cur, 2 );
                const double4 p3 = fetch( cur, 3 );
                body( p0 );
                body( p1 );
                body( p2 );
                body( p3 );
            }
        };
        unsigned long long q0 = P.q0, q1 = P.q1, q2 = P.q2;
        for ( int k = 0; k < nq; k += 3 )
        {
            quad( q0 );
            q0 = word( k + 3 );
            if ( k + 1 < nq )
                quad( q1 );
            q1 = word( k + 4 );
            if ( k + 2 < nq )
                quad( q2 );
            q2 = word( k + 5 );
        }
        for ( int o = apw; o < 64; o <<= 1 )
        {
            fx += __shfl_xor( fx, o, 64 );
            fy += __shfl_xor( fy, o, 64 );
            fz += __shfl_xor( fz, o, 64 );
        }
        finish_atom<G, FUSED>( active, sub, i, xi, fx, fy, fz, scap, x_next, v,
                               f, mt, ctl, mv2, vpre, img_b, img_e );
    }
    if ( FUSED && ENERGY )
    {
        // the staged records are dead once every wave is past its rows:
        // their first 256 bytes hold the block sums (no static LDS)
        __syncthreads();
        block_sum4_in( mv2, 0.5 * pe, 0.5 * vir, 0.0, partial + 4 * blockIdx.x,
                       reinterpret_cast<double( * )[8]>( s_rec ) );
    }
}

// LDS bytes before the reaction records of the tiled half-list kernel: the
// position records and, with several types, an int type per slot.
__host__ __device__ constexpr size_t half_acc_offset( int slot_cap, bool one_type )
{
    return ( (size_t)slot_cap * TILE_REC + ( one_type ? 0 : (size_t)slot_cap * sizeof( int ) ) +
             7 ) & ~(size_t)7;
}

// Tiled half list (compute_force_half, force_lj_cabana_neigh_impl.h:205-259):
// each listed pair once, both sides accumulated.  The block stages its
// tile's neighbourhood as records, with an FP64 reaction record per staged
// slot in LDS at the same byte offset in a second area.  A pair adds d fpair
// to atom i (registers) and subtracts it from slot j's reaction record (LDS
// atomics); the atom's own total joins its slot's record.  When every row
// of the tile is done, each slot with a nonzero record adds it to f of its
// atom with global atomics (an atom is staged by several tiles); ghost
// slots go to the local atoms they image, so update_force's fold has
// nothing left to add.  Energy weights follow compute_energy_half: 1 for a
// local j, 1/2 for a ghost j.
template <int G, int PREC, bool ONE_TYPE, bool ENERGY>
__global__ __launch_bounds__( 512 ) void k_tile_half(
    TileGrid T, const int *__restrict__ info, const int *__restrict__ woff,
    int lcap, int scap, int slot_cap, const double4 *__restrict__ x,
    const double4 *__restrict__ xg, const int *__restrict__ gperm,
    const int *__restrict__ ghost_root, double *__restrict__ f,
    const int *__restrict__ counts, const unsigned short *__restrict__ slab,
    const LJParams *__restrict__ Pd, PairCoef one, double *__restrict__ partial )
{
    extern __shared__ double s_rec[];
    __shared__ TileHead H;
    constexpr int APW = 64 / G;
    const int t = xcd_tile( blockIdx.x, gridDim.x );
    load_head( info, t, T.W, H );
    char *s_byte = reinterpret_cast<char *>( s_rec );
    const int *s_t = reinterpret_cast<const int *>( s_rec + 3 * slot_cap );
    // reaction records, then (energy only) a ghost flag per slot
    char *s_acc = s_byte + half_acc_offset( slot_cap, ONE_TYPE );
    int *s_gh = reinterpret_cast<int *>( s_acc + (size_t)slot_cap * TILE_REC );
    const int total = H.segoff[H.nseg];
    if ( ENERGY )
    {
        for ( int k = threadIdx.x; k < slot_cap; k += blockDim.x )
            s_gh[k] = k < total && ( find_seg( H.segoff, H.nseg, k ) & 1 );
    }
    const int R4 = ( ( lcap + G - 1 ) / G + 3 ) / 4;
    const unsigned short *st = slab + (size_t)woff[t] * R4 * 256;
    const int nw = woff[t + 1] - woff[t];
    const int nat = H.info[TI_NAT];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int nwb = blockDim.x >> 6;
    const unsigned long long sent4 =
        0x0001000100010001ull * (unsigned)( H.segoff[H.nseg] * TILE_REC );
    // the lane's atom in slab wave w: gw lanes per atom (G, or more in a
    // part-full wave), the atom, and the LDS slot it is staged at (its
    // staged column's local run); lanes past the atoms take the tile's
    // first atom
    auto lane_map = [&]( int w, int &gw, int &apw, int &sub, int &i, int &islot, bool &active )
    {
        gw = wave_lanes( G, nat, w );
        apw = 64 / gw;
        const int a = lane % apw;
        sub = lane / apw;
        const int ta = w * APW + a;
        active = a < APW && ta < nat;
        const int tq = active ? ta : 0;
        int ac = 0;
        while ( ac + 1 < T.W * T.W && tq >= H.aoff[ac + 1] )
            ++ac;
        i = H.info[TI_AL + ac] + ( tq - H.aoff[ac] );
        const int qc = atom_col_q( T.W, ac );
        islot = H.segoff[2 * qc] + ( i - H.info[TI_LB + qc] );
    };
    auto row_of = [&]( int w ) -> const unsigned long long *
    {
        return reinterpret_cast<const unsigned long long *>(
            st + ( (size_t)w * R4 * 64 + lane ) * 4 );
    };
    // the atom's position, row length and first three list words, loaded
    // before the staging so that their memory latency overlaps it (words
    // past a row's end are loaded but not used)
    struct WaveIn
    {
        double4 xi;
        int n;
        unsigned long long q0, q1, q2;
    };
    auto prep = [&]( int w, WaveIn &P )
    {
        int gw, apw, sub, i, islot;
        bool active;
        lane_map( w, gw, apw, sub, i, islot, active );
        P.xi = x[i];
        P.n = counts[i];
        const unsigned long long *row8 = row_of( w );
        P.q0 = __builtin_nontemporal_load( row8 );
        __builtin_amdgcn_sched_barrier( 0 );
        P.q1 = __builtin_nontemporal_load( row8 + (size_t)min( 1, R4 - 1 ) * 64 );
        __builtin_amdgcn_sched_barrier( 0 );
        P.q2 = __builtin_nontemporal_load( row8 + (size_t)min( 2, R4 - 1 ) * 64 );
        __builtin_amdgcn_sched_barrier( 0 );
    };
    WaveIn P;
    if ( wid < nw )
        prep( wid, P );
    // the reaction records are cleared while the staging loads are in flight
    stage_records<!ONE_TYPE>( H, x, xg, s_rec, slot_cap, reinterpret_cast<double *>( s_acc ),
                              3 * slot_cap ); // ends with a barrier
    double pe = 0.0, vir = 0.0;
    for ( int w = wid; w < nw; w += nwb )
    {
        if ( w != wid )
            prep( w, P );
        int gw, apw, sub, i, islot;
        bool active;
        lane_map( w, gw, apw, sub, i, islot, active );
        double fx = 0.0, fy = 0.0, fz = 0.0;
        const double4 xi =
            active ? P.xi : make_double4( -TILE_FAR, -TILE_FAR, -TILE_FAR, 0.0 );
        const int ti = (int)xi.w;
        const int n = active ? P.n : 0;
        int nr = n > sub ? ( n - sub + gw - 1 ) / gw : 0;
#pragma unroll
        for ( int o = 32; o > 0; o >>= 1 )
            nr = max( nr, __shfl_xor( nr, o, 64 ) );
        nr = __builtin_amdgcn_readfirstlane( nr );
        const unsigned long long *row8 = row_of( w );
        auto body = [&]( unsigned off )
        {
            const double *p = reinterpret_cast<const double *>( s_byte + off );
            const double dx = xi.x - p[0];
            const double dy = xi.y - p[1];
            const double dz = xi.z - p[2];
            const double rsq = dx * dx + dy * dy + dz * dz;
            const PairCoef c =
                coef<ONE_TYPE>( Pd, one, ti, ONE_TYPE ? 0 : s_t[off / TILE_REC] );
            double ep = 0.0, wp = 0.0;
            const double fpair = lj_pair<PREC, ENERGY>( rsq, c, ep, wp );
            if ( ENERGY )
            {
                pe += ( s_gh[off / TILE_REC] ? 0.5 : 1.0 ) * ep;
                vir += wp;
            }
            if ( rsq < c.cs )
            {
                const double ax = dx * fpair, ay = dy * fpair, az = dz * fpair;
                fx += ax;
                fy += ay;
                fz += az;
                double *r = reinterpret_cast<double *>( s_acc + off );
                unsafeAtomicAdd( r, -ax );
                unsafeAtomicAdd( r + 1, -ay );
                unsafeAtomicAdd( r + 2, -az );
            }
        };
        // list words three loads ahead, as in the full-list kernel: every
        // lane loads from its own column of the wave's slab (indexes past
        // the row's last word re-read it) and lanes past the tile's atoms
        // walk the sentinel in place of what their column holds
        const int nq = ( nr + 3 ) / 4;
        const int klast = max( nq - 1, 0 );
        auto word = [&]( int k ) -> unsigned long long
        { return __builtin_nontemporal_load( row8 + (size_t)min( k, klast ) * 64 ); };
        auto quad = [&]( unsigned long long w8 )
        {
            const unsigned long long wd = active ? w8 : sent4;
            body( (unsigned)( wd & 0xffffull ) );
            body( (unsigned)( ( wd >> 16 ) & 0xffffull ) );
            body( (unsigned)( ( wd >> 32 ) & 0xffffull ) );
            body( (unsigned)( wd >> 48 ) );
        };
        unsigned long long q0 = P.q0, q1 = P.q1, q2 = P.q2;
        for ( int k = 0; k < nq; k += 3 )
        {
            quad( q0 );
            q0 = word( k + 3 );
            if ( k + 1 < nq )
                quad( q1 );
            q1 = word( k + 4 );
            if ( k + 2 < nq )
                quad( q2 );
            q2 = word( k + 5 );
        }
        for ( int o = apw; o < 64; o <<= 1 )
        {
            fx += __shfl_xor( fx, o, 64 );
            fy += __shfl_xor( fy, o, 64 );
            fz += __shfl_xor( fz, o, 64 );
        }
        if ( active && sub == 0 )
        {
            double *r = reinterpret_cast<double *>( s_acc + islot * TILE_REC );
            unsafeAtomicAdd( r, fx );
            unsafeAtomicAdd( r + 1, fy );
            unsafeAtomicAdd( r + 2, fz );
        }
    }
    __syncthreads();
    // every staged slot with a reaction: locals by run start, ghosts through
    // the ghost sort permutation to the local atoms they image (a ghost's
    // force is its source atom's, which is what update_force's fold of the
    // ghost rows gives: the ghost rows are left untouched, and a fold adds
    // nothing)
    {
        for ( int k = threadIdx.x; k < total; k += blockDim.x )
        {
            const double *r = reinterpret_cast<const double *>( s_acc + k * TILE_REC );
            const double ax = r[0], ay = r[1], az = r[2];
            if ( ax == 0.0 && ay == 0.0 && az == 0.0 )
                continue;
            const int seg = find_seg( H.segoff, H.nseg, k );
            const int q = seg >> 1, o = k - H.segoff[seg];
            const int idx = ( seg & 1 ) ? ghost_root[gperm[H.info[TI_GB + q] + o]]
                                        : H.info[TI_LB + q] + o;
            unsafeAtomicAdd( &f[idx], ax );
            unsafeAtomicAdd( &f[scap + idx], ay );
            unsafeAtomicAdd( &f[2 * scap + idx], az );
        }
    }
    if ( ENERGY )
        block_sum4( 0.0, pe, vir, 0.0, partial + 4 * blockIdx.x );
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
        walk_row<G, ONE_TYPE>( row, nr, x, body );
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

// compute_energy over a CSR view (used for tiled lists).
template <bool ONE_TYPE>
__global__ __launch_bounds__( TPB ) void k_lj_energy_csr(
    int n_local, const double4 *__restrict__ x, const int64_t *__restrict__ off,
    const int *__restrict__ idx, const LJParams *__restrict__ Pd, PairCoef one,
    int half, double *__restrict__ partial )
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
// The device copy of the multi-type tables, uploaded once after the
// coefficients change (the copy waits for the stream, so the host table it
// reads can go); launches between changes reuse it without a host wait.
static cmd_status upload_params( const cmd_force *F, hipStream_t st,
                                 LJParams **dptr )
{
    cmd_force *mf = const_cast<cmd_force *>( F );
    if ( mf->dparams_dirty || !mf->dparams.p )
    {
        CMD_TRY( mf->dparams.reserve( 1 ) );
        const LJParams P = lj_params( F );
        CMD_HIP( hipMemcpyAsync( mf->dparams.p, &P, sizeof( P ),
                                 hipMemcpyHostToDevice, st ) );
        CMD_HIP( hipStreamSynchronize( st ) );
        mf->dparams_dirty = false;
    }
    *dptr = mf->dparams.p;
    return CMD_OK;
}

cmd_status lj_params_upload( const cmd_force *F, hipStream_t st )
{
    LJParams *d = nullptr;
    return F->ntypes > 1 ? upload_params( F, st, &d ) : CMD_OK;
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

template <int G, int PREC, bool ONE, bool FUSED, bool ENERGY>
static void launch_tile_k( cmd_neigh *N, cmd_system *S, double4 *xn,
                           const LJParams *Pd, const PairCoef &c,
                           const MassTable &mt, const StepCtl &ctl,
                           double *partial )
{
    // staged records plus the sentinel (and an int type per slot)
    const int slot_cap = std::max( N->max_tile_slots, 1 ) + 1;
    if constexpr ( PREC == CMD_PREC_MIXED )
    {
        if ( N->built_rec == TILE_REC32 )
        {
            // FP32 records (the type inside); at least the 4 x 8 doubles of
            // the thermo block sums, which reuse them
            const size_t lds =
                std::max<size_t>( (size_t)slot_cap * TILE_REC32, 32 * sizeof( double ) );
            k_tile_force<G, PREC, ONE, FUSED, ENERGY, true>
                <<<N->tg.ntiles, 64 * N->tg.bw, lds, S->stream>>>(
                    N->tg, N->tile_info.p, N->grid_sort.start.p, N->cap, S->cap, slot_cap,
                    S->x.p, S->xg.p, xn, S->v.p, S->f.p, N->counts.p, N->tslab.p, Pd, c, mt,
                    ctl, partial );
            return;
        }
    }
    // at least the 4 x 8 doubles of the thermo block sums, which reuse it
    const size_t lds = std::max<size_t>(
        (size_t)slot_cap * ( TILE_REC + ( ONE ? 0 : sizeof( int ) ) ), 32 * sizeof( double ) );
    k_tile_force<G, PREC, ONE, FUSED, ENERGY, false><<<N->tg.ntiles, 64 * N->tg.bw, lds, S->stream>>>(
        N->tg, N->tile_info.p, N->grid_sort.start.p, N->cap, S->cap, slot_cap,
        S->x.p, S->xg.p, xn, S->v.p, S->f.p, N->counts.p, N->tslab.p, Pd, c, mt,
        ctl, partial );
}

template <int G, int PREC, bool ONE, bool FUSED>
static void launch_tile_t( cmd_neigh *N, cmd_system *S, double4 *xn,
                           const LJParams *Pd, const PairCoef &c,
                           const MassTable &mt, const StepCtl &ctl,
                           double *partial )
{
    if constexpr ( FUSED )
    {
        if ( ctl.thermo )
        {
            launch_tile_k<G, PREC, ONE, FUSED, true>( N, S, xn, Pd, c, mt, ctl, partial );
            return;
        }
    }
    launch_tile_k<G, PREC, ONE, FUSED, false>( N, S, xn, Pd, c, mt, ctl, partial );
}

template <int PREC, bool ONE, bool FUSED>
static cmd_status launch_tile_g( cmd_neigh *N, cmd_system *S, double4 *xn,
                                 const LJParams *Pd, const PairCoef &c,
                                 const MassTable &mt, const StepCtl &ctl,
                                 double *partial )
{
    switch ( N->lanes )
    {
    case 2:
        launch_tile_t<2, PREC, ONE, FUSED>( N, S, xn, Pd, c, mt, ctl, partial );
        break;
    case 4:
        launch_tile_t<4, PREC, ONE, FUSED>( N, S, xn, Pd, c, mt, ctl, partial );
        break;
    case 8:
        launch_tile_t<8, PREC, ONE, FUSED>( N, S, xn, Pd, c, mt, ctl, partial );
        break;
    default:
        set_error( "tiled lists need 2, 4 or 8 lanes per atom" );
        return CMD_ERR_ARG;
    }
    CMD_LAUNCH_CHECK();
    return CMD_OK;
}

static cmd_status launch_tile( int prec, bool one, bool fused,
                               const cmd_neigh *Nc, cmd_system *S, double4 *xn,
                               const LJParams *Pd, const PairCoef &c,
                               const MassTable &mt, const StepCtl &ctl,
                               double *partial )
{
    cmd_neigh *N = const_cast<cmd_neigh *>( Nc );
    CMD_REQUIRE( N->built_rec == TILE_REC || prec == CMD_PREC_MIXED, CMD_ERR_STATE,
                 "the tiled list addresses FP32 records (mixed precision); rebuild it for "
                 "FP64 forces" );
    if ( prec == CMD_PREC_FP64 )
    {
        if ( one )
            return fused ? launch_tile_g<CMD_PREC_FP64, true, true>( N, S, xn, Pd, c, mt, ctl, partial )
                         : launch_tile_g<CMD_PREC_FP64, true, false>( N, S, xn, Pd, c, mt, ctl, partial );
        return fused ? launch_tile_g<CMD_PREC_FP64, false, true>( N, S, xn, Pd, c, mt, ctl, partial )
                     : launch_tile_g<CMD_PREC_FP64, false, false>( N, S, xn, Pd, c, mt, ctl, partial );
    }
    if ( one )
        return fused ? launch_tile_g<CMD_PREC_MIXED, true, true>( N, S, xn, Pd, c, mt, ctl, partial )
                     : launch_tile_g<CMD_PREC_MIXED, true, false>( N, S, xn, Pd, c, mt, ctl, partial );
    return fused ? launch_tile_g<CMD_PREC_MIXED, false, true>( N, S, xn, Pd, c, mt, ctl, partial )
                 : launch_tile_g<CMD_PREC_MIXED, false, false>( N, S, xn, Pd, c, mt, ctl, partial );
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
    if ( N->tiled )
    {
        CMD_REQUIRE( S->xg_valid, CMD_ERR_STATE,
                     "tiled list needs the sorted ghosts of its build" );
        if ( nblocks_out )
            *nblocks_out = N->tg.ntiles;
        return launch_tile( F->prec, one, fused, N, S, x_next, Pd, c, mt, ctl,
                            partial );
    }
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

template <int G, int PREC, bool ONE, bool ENERGY>
static void launch_tile_half_k( const cmd_neigh *N, cmd_system *S, const LJParams *Pd,
                                const PairCoef &c, double *partial )
{
    const int slot_cap = std::max( N->max_tile_slots, 1 ) + 1;
    const size_t lds = half_acc_offset( slot_cap, ONE ) + (size_t)slot_cap * TILE_REC +
                       ( ENERGY ? (size_t)slot_cap * sizeof( int ) : 0 );
    k_tile_half<G, PREC, ONE, ENERGY><<<N->tg.ntiles, 64 * N->tg.bw, lds, S->stream>>>(
        N->tg, N->tile_info.p, N->grid_sort.start.p, N->cap, S->cap, slot_cap, S->x.p,
        S->xg.p, S->gsort.perm.p, S->ghost_root.p, S->f.p, N->counts.p, N->tslab.p, Pd, c,
        partial );
}

template <int G, int PREC>
static void launch_tile_half_t( const cmd_neigh *N, cmd_system *S, bool one, bool thermo,
                                const LJParams *Pd, const PairCoef &c, double *partial )
{
    if ( one )
    {
        if ( thermo )
            launch_tile_half_k<G, PREC, true, true>( N, S, Pd, c, partial );
        else
            launch_tile_half_k<G, PREC, true, false>( N, S, Pd, c, partial );
    }
    else
    {
        if ( thermo )
            launch_tile_half_k<G, PREC, false, true>( N, S, Pd, c, partial );
        else
            launch_tile_half_k<G, PREC, false, false>( N, S, Pd, c, partial );
    }
}

template <int PREC>
static cmd_status launch_tile_half( const cmd_neigh *N, cmd_system *S, bool one,
                                    bool thermo, const LJParams *Pd, const PairCoef &c,
                                    double *partial )
{
    switch ( N->lanes )
    {
    case 2:
        launch_tile_half_t<2, PREC>( N, S, one, thermo, Pd, c, partial );
        break;
    case 4:
        launch_tile_half_t<4, PREC>( N, S, one, thermo, Pd, c, partial );
        break;
    case 8:
        launch_tile_half_t<8, PREC>( N, S, one, thermo, Pd, c, partial );
        break;
    default:
        set_error( "tiled lists need 2, 4 or 8 lanes per atom" );
        return CMD_ERR_ARG;
    }
    CMD_LAUNCH_CHECK();
    return CMD_OK;
}

cmd_status lj_half( const cmd_force *F, cmd_system *S, const cmd_neigh *N,
                    int thermo, double *partial, int *nblocks_out )
{
    const int n = S->n_local;
    const int G = N->lanes;
    const int nb = N->tiled ? N->tg.ntiles : blocks_for( n, G );
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
    if ( N->tiled )
    {
        CMD_REQUIRE( S->xg_valid, CMD_ERR_STATE,
                     "tiled list needs the sorted ghosts of its build" );
        const bool one = F->ntypes == 1;
        if ( F->prec == CMD_PREC_FP64 )
            return launch_tile_half<CMD_PREC_FP64>( N, S, one, thermo != 0, Pd, c, partial );
        return launch_tile_half<CMD_PREC_MIXED>( N, S, one, thermo != 0, Pd, c, partial );
    }
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
    int nb = blocks_for( n, G );
    CMD_REQUIRE( N->n_rows == n, CMD_ERR_STATE, "neighbor list is stale" );
    CMD_TRY( S->sc.dpart.reserve( 4 * (size_t)std::max( nb, blocks_for( n, 1 ) ) ) );
    CMD_TRY( S->sc.dres.reserve( 8 ) );
    const LJParams P = lj_params( F );
    const PairCoef c = one_coef( P );
    LJParams *Pd = nullptr;
    const bool one = F->ntypes == 1;
    if ( !one )
        CMD_TRY( upload_params( F, S->stream, &Pd ) );
    if ( N->tiled )
    {
        CMD_TRY( neigh_csr( const_cast<cmd_neigh *>( N ) ) );
        nb = blocks_for( n, 1 );
        if ( one )
            k_lj_energy_csr<true><<<nb, TPB, 0, S->stream>>>(
                n, S->x.p, N->offsets.p, N->idx.p, Pd, c, N->half, S->sc.dpart.p );
        else
            k_lj_energy_csr<false><<<nb, TPB, 0, S->stream>>>(
                n, S->x.p, N->offsets.p, N->idx.p, Pd, c, N->half, S->sc.dpart.p );
    }
    else
    {
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
    }
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

// Build the list in each layout -- lanes per atom 1, 2, 4, 8 and, for the
// LDS-tiled layout, tiles of 2 x 2 columns 2 to 4 cells tall and of one
// column 4 to 8 cells tall -- time the force kernel on it and keep the
// fastest; the list is left built in that layout, and later builds keep at
// least its force-kernel blocks per CU.
cmd_status autotune_lanes( cmd_force *F, cmd_system *S, cmd_neigh *N, int every )
{
    struct Cand
    {
        int G, W, E, bw;
    };
    // lanes per atom x tile width x tile height x waves per block (0: one
    // per slab wave of the largest tile)
    static const Cand all[] = { { 1, 0, 0, 0 }, { 2, 2, 2, 4 }, { 2, 2, 2, 2 }, { 2, 2, 3, 4 },
                                { 2, 2, 3, 2 }, { 2, 2, 3, 8 }, { 2, 2, 4, 4 }, { 2, 2, 4, 2 },
                                { 2, 2, 4, 8 }, { 2, 2, 5, 4 }, { 2, 2, 6, 4 }, { 2, 2, 6, 8 },
                                { 2, 1, 4, 4 }, { 2, 1, 6, 4 }, { 2, 1, 6, 2 }, { 4, 2, 3, 4 },
                                { 4, 2, 4, 4 }, { 4, 1, 6, 4 }, { 8, 2, 3, 4 } };
    // CMD_TUNE=G,W,E,bw limits the search to one layout (for measurements)
    std::vector<Cand> cv( std::begin( all ), std::end( all ) );
    if ( const char *e = std::getenv( "CMD_TUNE" ) )
    {
        Cand c{};
        if ( std::sscanf( e, "%d,%d,%d,%d", &c.G, &c.W, &c.E, &c.bw ) == 4 )
            cv.assign( 1, c );
    }
    const Cand *cands = cv.data();
    const size_t ncand = cv.size();
    N->occ_blocks = 0;
    // time the fused step the run loop launches (force, both half-kicks and
    // the drift into x_alt); the velocities it kicks are restored after
    StepCtl ctl{};
    ctl.dtf = 0.5 * S->dt / S->mvv2e;
    ctl.dtv = S->dt;
    for ( int k = 0; k < 3; ++k )
        ctl.L[k] = S->dom.L[k];
    ctl.do_final = 1;
    ctl.do_next_initial = 1;
    CMD_TRY( S->x_alt.reserve( S->cap ) );
    CMD_TRY( S->v_alt.reserve( 3 * (size_t)S->cap ) );
    const size_t vbytes = sizeof( double ) * 3 * (size_t)S->cap;
    CMD_HIP( hipMemcpyAsync( S->v_alt.p, S->v.p, vbytes, hipMemcpyDeviceToDevice, S->stream ) );
    hipEvent_t e0, e1, b0, b1;
    CMD_HIP( hipEventCreate( &e0 ) );
    CMD_HIP( hipEventCreate( &e1 ) );
    CMD_HIP( hipEventCreate( &b0 ) );
    CMD_HIP( hipEventCreate( &b1 ) );
    every = std::max( every, 1 );
    float best = 1e30f;
    Cand bestc = { 1, 0, 0, 0 };
    std::vector<float> tms( ncand, 1e30f );
    cmd_status st = CMD_OK;
    for ( size_t ci = 0; ci < ncand; ++ci )
    {
        const Cand &cd = cands[ci];
        tms[ci] = 1e30f;
        N->lanes = cd.G;
        N->tile_w = cd.W;
        N->tile_e = cd.E;
        N->tile_bw = cd.bw;
        (void)hipEventRecord( b0, S->stream );
        st = neigh_build( N, S );
        if ( st != CMD_OK )
            break;
        (void)hipEventRecord( b1, S->stream );
        if ( cd.E > 0 && !N->tiled )
            continue; // the tiled layout does not apply here
        constexpr int reps = 6;
        for ( int r = 0; r < 2; ++r )
            st = lj_full( F, S, N, ctl, S->x_alt.p, nullptr, nullptr );
        if ( st != CMD_OK )
            break;
        (void)hipEventRecord( e0, S->stream );
        for ( int r = 0; r < reps; ++r )
            (void)lj_full( F, S, N, ctl, S->x_alt.p, nullptr, nullptr );
        (void)hipEventRecord( e1, S->stream );
        (void)hipEventSynchronize( e1 );
        (void)hipMemcpyAsync( S->v.p, S->v_alt.p, vbytes, hipMemcpyDeviceToDevice, S->stream );
        float ms = 0.0f, bms = 0.0f;
        (void)hipEventElapsedTime( &ms, e0, e1 );
        (void)hipEventElapsedTime( &bms, b0, b1 );
        if ( std::getenv( "CMD_VERBOSE" ) )
            std::fprintf( stderr,
                          "# autotune lanes=%d tile=%dx%dx%d waves/block=%d: %.4f ms/step, "
                          "list build %.3f ms\n",
                          cd.G, N->tiled ? N->tg.W : 0, N->tiled ? N->tg.W : 0,
                          N->tiled ? N->tg.E : 0, N->tiled ? N->tg.bw : 0, ms / reps, bms );
        // the step time plus the build's share of the steps it serves
        ms += bms * reps / every;
        tms[ci] = ms;
        if ( ms < best )
        {
            best = ms;
            bestc = cd;
        }
    }
    // Near the lattice sites (init, first rebuild) more lanes per atom
    // measure within a few percent of fewer lanes but fall behind as the
    // structure disorders (melted LJ, 2,048,000 atoms: 2 lanes 0.253 ms,
    // 4 lanes 0.270 ms per fused step); so among candidates within 3 % of
    // the fastest the fewest lanes win, the faster of those first.
    for ( size_t ci = 0; ci < ncand; ++ci )
    {
        const Cand &cd = cands[ci];
        if ( tms[ci] > 1.03f * best )
            continue;
        if ( cd.G < bestc.G )
        {
            bestc = cd;
            // the fastest candidate with these lanes
            for ( size_t cj = 0; cj < ncand; ++cj )
                if ( cands[cj].G == cd.G && tms[cj] < tms[ci] )
                    bestc = cands[cj];
        }
    }
    (void)hipEventDestroy( e0 );
    (void)hipEventDestroy( e1 );
    (void)hipEventDestroy( b0 );
    (void)hipEventDestroy( b1 );
    CMD_HIP( hipMemcpyAsync( S->v.p, S->v_alt.p, vbytes, hipMemcpyDeviceToDevice, S->stream ) );
    CMD_TRY( st );
    F->lanes = bestc.G;
    F->tuned = true;
    N->lanes = bestc.G;
    N->tile_w = bestc.W;
    N->tile_e = bestc.E;
    N->tile_bw = bestc.bw;
    CMD_TRY( neigh_build( N, S ) );
    N->occ_blocks = N->tiled ? N->tile_blocks : 0;
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
    if ( f && f->nnp )
        nnp_destroy( f->nnp );
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
    f->dparams_dirty = true;
    return CMD_OK;
}

cmd_status cmd_force_set_precision( cmd_force *f, int prec )
{
    CMD_REQUIRE( f, CMD_ERR_ARG, "force is NULL" );
    CMD_REQUIRE( prec == CMD_PREC_FP64 || prec == CMD_PREC_MIXED, CMD_ERR_ARG,
                 "unknown precision %d", prec );
    f->prec = prec;
    if ( f->kind == 2 )
        return nnp_set_precision( f->nnp, prec );
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
    if ( f->kind == 2 )
        return nnp_compute( f->nnp, s, n );
    return force_compute_plain( f, s, n );
}

cmd_status cmd_force_energy( cmd_force *f, cmd_system *s, cmd_neigh *n,
                             double *pe )
{
    CMD_REQUIRE( f && s && n && pe, CMD_ERR_ARG, "NULL argument" );
    if ( f->kind == 2 )
        return nnp_reference_energy( f->nnp, s, pe );
    return lj_energy_virial( f, s, n, pe, nullptr );
}

cmd_status cmd_force_virial( cmd_force *f, cmd_system *s, cmd_neigh *n,
                             double *w )
{
    CMD_REQUIRE( f && s && n && w, CMD_ERR_ARG, "NULL argument" );
    CMD_REQUIRE( f->kind == 0, CMD_ERR_ARG, "the virial is computed for LJ only" );
    return lj_energy_virial( f, s, n, nullptr, w );
}

} // extern "C"
