/*
 * oracle.cpp -- FP64 CPU restatement of the CabanaMD timestep path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP engine and the CPU
 * baseline for bench.py.  Nothing in the product (cabanamd_amd/) links or
 * loads this file.  Built with g++ -O3 -fopenmp -ffp-contract=off.
 *
 * Citations are to the CabanaMD reference tree (src/...).  Where the
 * reference delegates to the external Cabana 0.6.1 library (LinkedCellList,
 * VerletList, Halo gather/scatter, Cabana::Grid local mesh), the published
 * algorithm of that library is restated as noted at each function.
 */
#include "oracle.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <cstdint>
#include <limits>
#include <string>
#include <unordered_map>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

static const int RED_CHUNK = 4096; /* atoms per deterministic partial sum */

/* Neumaier compensated sum: the energy sums run over ~10^6 pair terms with
 * a systematic shift term, where a plain running sum drifts by ~1e-12
 * relative; compensation keeps the checker within a few ulp of the exact
 * sum, whatever the summation order of the code it checks. */
struct CompSum
{
    double s = 0.0, c = 0.0;
    void add( double v )
    {
        const double t = s + v;
        if ( std::fabs( s ) >= std::fabs( v ) )
            c += ( s - t ) + v;
        else
            c += ( v - t ) + s;
        s = t;
    }
    double value() const { return s + c; }
};

namespace
{
using clk = std::chrono::steady_clock;
double seconds_since( clk::time_point t0 )
{
    return std::chrono::duration<double>( clk::now() - t0 ).count();
}
} // namespace

/* ------------------------------------------------------------------------ */
/* Domain: system.h:140-205.  Cabana::Grid uniform global mesh with 100 cells
 * per dimension per rank; one rank owns everything.  Own corners are
 * low + cell*offset and low + cell*(offset+ncell); ghost corners extend by
 * halo_width = ceil(ghost_cutoff / min cell size) cells. */
void orc_create_domain( const double low[3], const double high[3],
                        double ghost_cutoff, orc_domain *d )
{
    for ( int k = 0; k < 3; ++k )
    {
        d->low[k] = low[k];
        d->high[k] = high[k];
        d->L[k] = high[k] - low[k];
        d->cell[k] = ( high[k] - low[k] ) / 100.0;
    }
    double min_cell = std::min( std::min( d->cell[0], d->cell[1] ), d->cell[2] );
    d->halo_width = (int)std::ceil( ghost_cutoff / min_cell );
    for ( int k = 0; k < 3; ++k )
    {
        d->local_lo[k] = low[k] + d->cell[k] * 0.0;
        d->local_hi[k] = low[k] + d->cell[k] * 100.0;
        d->local_mesh[k] = d->local_hi[k] - d->local_lo[k];
        d->ghost_lo[k] = low[k] + d->cell[k] * ( 0.0 - d->halo_width );
        d->ghost_hi[k] = low[k] + d->cell[k] * ( 100.0 + d->halo_width );
    }
}

/* ------------------------------------------------------------------------ */
/* LAMMPS_RandomVelocityGeom, inputFile.h:66-148.  Park-Miller minimal
 * standard generator with Schrage factorisation, Marsaglia polar gaussian,
 * and a Jenkins one-at-a-time hash of the seed bytes then the 24 coordinate
 * bytes (bytes read as plain char, i.e. signed on x86). */
static const int RNG_IA = 16807;
static const int RNG_IM = 2147483647;
static const int RNG_IQ = 127773;
static const int RNG_IR = 2836;

double orc_rng_uniform( int *seed )
{
    int k = *seed / RNG_IQ;
    *seed = RNG_IA * ( *seed - k * RNG_IQ ) - RNG_IR * k;
    if ( *seed < 0 )
        *seed += RNG_IM;
    return ( 1.0 / RNG_IM ) * *seed;
}

double orc_rng_gaussian( int *seed )
{
    double v1, v2, rsq;
    do
    {
        v1 = 2.0 * orc_rng_uniform( seed ) - 1.0;
        v2 = 2.0 * orc_rng_uniform( seed ) - 1.0;
        rsq = v1 * v1 + v2 * v2;
    } while ( ( rsq >= 1.0 ) || ( rsq == 0.0 ) );
    const double fac = std::sqrt( -2.0 * std::log( rsq ) / rsq );
    return v2 * fac;
}

int orc_rng_reset( int ibase, const double coord[3] )
{
    unsigned int hash = 0;
    const char *str = (const char *)&ibase;
    for ( int i = 0; i < (int)sizeof( int ); ++i )
    {
        hash += str[i];
        hash += ( hash << 10 );
        hash ^= ( hash >> 6 );
    }
    str = (const char *)coord;
    for ( int i = 0; i < 3 * (int)sizeof( double ); ++i )
    {
        hash += str[i];
        hash += ( hash << 10 );
        hash ^= ( hash >> 6 );
    }
    hash += ( hash << 3 );
    hash ^= ( hash >> 11 );
    hash += ( hash << 15 );
    int seed = hash & 0x7ffffff; /* 27-bit mask, as in the reference */
    if ( !seed )
        seed = 1;
    for ( int i = 0; i < 5; ++i )
        orc_rng_uniform( &seed );
    return seed;
}

/* ------------------------------------------------------------------------ */
/* Lattice: inputFile_impl.h:562-792.  A single region block equal to the
 * box with create_atoms type 1, so every accepted site is type 0 and the
 * reference's rand() % 1 pick is always the one region. */
namespace
{
struct LatticeLoop
{
    int ix0, ix1, iy0, iy1, iz0, iz1;
    double maxv[3];
};

LatticeLoop lattice_loop( const orc_params *p, const orc_domain *d )
{
    const double a = p->lattice_constant;
    LatticeLoop l;
    for ( int k = 0; k < 3; ++k )
        l.maxv[k] = a * p->region_hi[k];
    int st[3], en[3];
    for ( int k = 0; k < 3; ++k )
    {
        st[k] = (int)( d->local_lo[k] / a - 0.5 );
        en[k] = (int)std::max( std::min( l.maxv[k] / a, d->local_hi[k] / a + 0.5 ),
                               static_cast<double>( st[k] ) );
        if ( st[k] == en[k] )
            en[k] -= 1;
    }
    l.ix0 = st[0];
    l.ix1 = en[0];
    l.iy0 = st[1];
    l.iy1 = en[1];
    l.iz0 = st[2];
    l.iz1 = en[2];
    return l;
}

bool lattice_accept( const orc_domain *d, const LatticeLoop &l, double xt,
                     double yt, double zt )
{
    return ( xt >= d->local_lo[0] ) && ( yt >= d->local_lo[1] ) &&
           ( zt >= d->local_lo[2] ) && ( xt < d->local_hi[0] ) &&
           ( yt < d->local_hi[1] ) && ( zt < d->local_hi[2] ) &&
           ( xt < l.maxv[0] ) && ( yt < l.maxv[1] ) && ( zt < l.maxv[2] );
}

template <class Visit>
int lattice_walk( const orc_params *p, const orc_domain *d, Visit visit )
{
    const double a = p->lattice_constant;
    LatticeLoop l = lattice_loop( p, d );
    int n = 0;
    if ( p->lattice_style == 0 )
    {
        for ( int iz = l.iz0; iz <= l.iz1; ++iz )
        {
            double zt = a * ( iz + p->lattice_offset[2] );
            for ( int iy = l.iy0; iy <= l.iy1; ++iy )
            {
                double yt = a * ( iy + p->lattice_offset[1] );
                for ( int ix = l.ix0; ix <= l.ix1; ++ix )
                {
                    double xt = a * ( ix + p->lattice_offset[0] );
                    if ( lattice_accept( d, l, xt, yt, zt ) )
                        visit( n++, xt, yt, zt );
                }
            }
        }
        return n;
    }
    double basis[4][3] = { { 0.0, 0.0, 0.0 },
                           { 0.5, 0.5, 0.0 },
                           { 0.5, 0.0, 0.5 },
                           { 0.0, 0.5, 0.5 } };
    for ( int b = 0; b < 4; ++b )
        for ( int k = 0; k < 3; ++k )
            basis[b][k] += p->lattice_offset[k];
    for ( int iz = l.iz0; iz <= l.iz1; ++iz )
        for ( int iy = l.iy0; iy <= l.iy1; ++iy )
            for ( int ix = l.ix0; ix <= l.ix1; ++ix )
                for ( int b = 0; b < 4; ++b )
                {
                    double xt = a * ( 1.0 * ix + basis[b][0] );
                    double yt = a * ( 1.0 * iy + basis[b][1] );
                    double zt = a * ( 1.0 * iz + basis[b][2] );
                    if ( lattice_accept( d, l, xt, yt, zt ) )
                        visit( n++, xt, yt, zt );
                }
    return n;
}
} // namespace

/* glibc rand() from seed 1: r[0] = 1, r[i] = 16807 r[i-1] mod (2^31-1)
 * (i < 31), r[i] = r[i-31] (31 <= i < 34), r[i] = r[i-31] + r[i-3] mod 2^32
 * afterwards; call k returns r[344 + k] >> 1.  Kept as one growing array. */
struct GlibcRand
{
    std::vector<uint32_t> r;
    size_t k = 0;
    GlibcRand()
    {
        r.resize( 344 );
        r[0] = 1;
        for ( int i = 1; i < 31; ++i )
            r[i] = (uint32_t)( ( 16807ull * r[i - 1] ) % 2147483647ull );
        for ( int i = 31; i < 34; ++i )
            r[i] = r[i - 31];
        for ( int i = 34; i < 344; ++i )
            r[i] = r[i - 31] + r[i - 3];
    }
    int operator()()
    {
        const size_t i = 344 + k++;
        r.push_back( r[i - 31] + r[i - 3] );
        return (int)( r[i] >> 1 );
    }
};

void orc_crand( int n, int *out )
{
    GlibcRand g;
    for ( int i = 0; i < n; ++i )
        out[i] = g();
}

/* Types of the lattice atoms (inputFile_impl.h:651, 755-777): sc sites draw
 * rand() % ntypes; an fcc site is created when a region with a
 * create_atoms type holds it, the type of one such region picked by
 * rand() % count among them in the regions map's iteration order
 * (std::unordered_map<std::string, Block>, inputFile.h:172). */
struct LatticeTyping
{
    const orc_params *p;
    std::vector<int> order; /* region slots in map iteration order */
    GlibcRand g;
    explicit LatticeTyping( const orc_params *pp ) : p( pp )
    {
        std::unordered_map<std::string, int> m;
        for ( int k = 0; k < p->nregions; ++k )
            m[p->region_name[k]] = k;
        for ( auto it = m.begin(); it != m.end(); ++it )
            order.push_back( it->second );
    }
    bool in_block( int k, double xt, double yt, double zt ) const
    {
        const double a = p->lattice_constant;
        const double *b = p->region_block[k];
        return xt >= a * b[0] && yt >= a * b[2] && zt >= a * b[4] && xt < a * b[1] &&
               yt < a * b[3] && zt < a * b[5];
    }
    /* -1: no atom at this site; else the 0-based type (draws when asked) */
    int type_at( double xt, double yt, double zt, bool draw )
    {
        if ( p->lattice_style == 0 )
            return draw ? g() % p->ntypes : 0;
        if ( p->nregions == 0 )
        {
            if ( draw )
                g();
            return 0;
        }
        std::vector<int> ids;
        for ( int k : order )
            if ( p->region_type[k] != 0 && in_block( k, xt, yt, zt ) )
                ids.push_back( k );
        if ( ids.empty() )
            return -1;
        if ( !draw )
            return 0;
        return p->region_type[ids[g() % (int)ids.size()]] - 1;
    }
};

int orc_lattice_count( const orc_params *p, const orc_domain *d )
{
    LatticeTyping ty( p );
    int n = 0;
    lattice_walk( p, d,
                  [&]( int, double xt, double yt, double zt )
                  {
                      if ( ty.type_at( xt, yt, zt, false ) >= 0 )
                          ++n;
                  } );
    return n;
}

int orc_lattice_fill( const orc_params *p, const orc_domain *d, double *x,
                      int *type, int *id )
{
    LatticeTyping ty( p );
    int n = 0;
    lattice_walk( p, d,
                  [&]( int, double xt, double yt, double zt )
                  {
                      const int t = ty.type_at( xt, yt, zt, true );
                      if ( t < 0 )
                          return;
                      x[3 * n + 0] = xt;
                      x[3 * n + 1] = yt;
                      x[3 * n + 2] = zt;
                      type[n] = t;
                      id[n] = n + 1;
                      ++n;
                  } );
    return n;
}

/* Velocities: inputFile_impl.h:795-865.  type_to_temperature holds only the
 * 'velocity all' entry (keyed as type 1, inputFile_impl.h:404-409); other
 * types get the map's default {temp 0, seed 0}. */
void orc_velocity_create( const orc_params *p, int n, const double *x,
                          const int *type, double *v )
{
    double total_mass = 0.0, px = 0.0, py = 0.0, pz = 0.0;
    for ( int i = 0; i < n; ++i )
    {
        int key = type[i] + 1;
        int seed_base = ( key == 1 ) ? p->temp_seed : p->type_seed[key - 1];
        int seed = orc_rng_reset( seed_base, &x[3 * i] );
        double mass_i = p->mass[type[i]];
        double vx, vy, vz;
        if ( p->vel_gaussian )
        {
            vx = orc_rng_gaussian( &seed );
            vy = orc_rng_gaussian( &seed );
            vz = orc_rng_gaussian( &seed );
        }
        else
        {
            vx = orc_rng_uniform( &seed ) - 0.5;
            vy = orc_rng_uniform( &seed ) - 0.5;
            vz = orc_rng_uniform( &seed ) - 0.5;
        }
        v[3 * i + 0] = vx / std::sqrt( mass_i );
        v[3 * i + 1] = vy / std::sqrt( mass_i );
        v[3 * i + 2] = vz / std::sqrt( mass_i );
        total_mass += mass_i;
        px += mass_i * v[3 * i + 0];
        py += mass_i * v[3 * i + 1];
        pz += mass_i * v[3 * i + 2];
    }
    double sx = px / total_mass, sy = py / total_mass, sz = pz / total_mass;
    for ( int i = 0; i < n; ++i )
    {
        v[3 * i + 0] -= sx;
        v[3 * i + 1] -= sy;
        v[3 * i + 2] -= sz;
    }
    double T = orc_temperature( n, n, v, type, p->mass, p->mvv2e, p->boltz );
    for ( int i = 0; i < n; ++i )
    {
        double target = ( type[i] + 1 == 1 ) ? p->temp_target : p->type_temp[type[i]];
        double s = std::sqrt( target / T );
        v[3 * i + 0] *= s;
        v[3 * i + 1] *= s;
        v[3 * i + 2] *= s;
    }
}

/* ------------------------------------------------------------------------ */
/* NVE: integrator_nve.h:91-110; dtf = 0.5*dt/mvv2e, dtv = dt
 * (integrator_nve_impl.h:49-54). */
void orc_nve_initial( int n, double *x, double *v, const double *f,
                      const int *type, const double *mass, double dt,
                      double mvv2e )
{
    const double dtf = 0.5 * dt / mvv2e;
    const double dtv = dt;
#pragma omp parallel for schedule( static )
    for ( int i = 0; i < n; ++i )
    {
        const double dtfm = dtf / mass[type[i]];
        v[3 * i + 0] += dtfm * f[3 * i + 0];
        v[3 * i + 1] += dtfm * f[3 * i + 1];
        v[3 * i + 2] += dtfm * f[3 * i + 2];
        x[3 * i + 0] += dtv * v[3 * i + 0];
        x[3 * i + 1] += dtv * v[3 * i + 1];
        x[3 * i + 2] += dtv * v[3 * i + 2];
    }
}

void orc_nve_final( int n, double *v, const double *f, const int *type,
                    const double *mass, double dt, double mvv2e )
{
    const double dtf = 0.5 * dt / mvv2e;
#pragma omp parallel for schedule( static )
    for ( int i = 0; i < n; ++i )
    {
        const double dtfm = dtf / mass[type[i]];
        v[3 * i + 0] += dtfm * f[3 * i + 0];
        v[3 * i + 1] += dtfm * f[3 * i + 1];
        v[3 * i + 2] += dtfm * f[3 * i + 2];
    }
}

/* Wrap: comm_mpi.h:141-168 (TagExchangeSelf, proc_grid == 1 in every dim). */
void orc_wrap( int n, double *x, const double L[3] )
{
#pragma omp parallel for schedule( static )
    for ( int i = 0; i < n; ++i )
        for ( int k = 0; k < 3; ++k )
        {
            const double c = x[3 * i + k];
            if ( c > L[k] )
                x[3 * i + k] -= L[k];
            if ( c < 0 )
                x[3 * i + k] += L[k];
        }
}

/* ------------------------------------------------------------------------ */
/* Binning: binning_cabana_impl.h:57-113, called with (neigh_cut x3, 1, true,
 * false, true).  The Cabana LinkedCellList grid: n = floor((max-min)/delta),
 * cell = (max-min)/n, index = floor((x-min)/cell) clamped to n-1, cardinal id
 * (i*ny + j)*nz + k; the permutation is the counting sort by cell id, in
 * index order within a cell (the Serial-backend order). */
void orc_bin_grid( const orc_domain *d, const double delta_in[3],
                   int halo_depth, double gmin[3], double rd[3], int ng[3] )
{
    int nbin[3];
    double delta[3];
    for ( int k = 0; k < 3; ++k )
    {
        nbin[k] = (int)( d->local_mesh[k] / delta_in[k] );
        if ( nbin[k] == 0 )
            nbin[k] = 1;
        delta[k] = d->local_mesh[k] / nbin[k];
    }
    const double eps = delta[0] / 1000;
    for ( int k = 0; k < 3; ++k )
    {
        gmin[k] = -delta[k] * halo_depth - eps + d->local_lo[k];
        const double gmax = delta[k] * halo_depth + eps + d->local_hi[k];
        ng[k] = (int)std::floor( ( gmax - gmin[k] ) * ( 1.0 / delta[k] ) );
        double c = ( gmax - gmin[k] ) / ng[k];
        rd[k] = 1.0 / c;
    }
}

static inline int grid_cell_coord( double c, double mn, double rd, int n,
                                   int floor_mode )
{
    const double t = ( c - mn ) * rd;
    int ic = floor_mode ? (int)std::floor( t ) : (int)t;
    if ( ic == n )
        ic = n - 1;
    /* out-of-grid coordinates (undefined in the reference) clamp */
    return std::min( std::max( ic, 0 ), n - 1 );
}

int orc_bin_sort( int n, const double *x, const orc_domain *d,
                  const double delta_in[3], int halo_depth, int *perm,
                  int *cell_of )
{
    double gmin[3], rd[3];
    int ng[3];
    orc_bin_grid( d, delta_in, halo_depth, gmin, rd, ng );
    const int ncell = ng[0] * ng[1] * ng[2];
    std::vector<int> cid( n );
    std::vector<int> count( ncell + 1, 0 );
    for ( int i = 0; i < n; ++i )
    {
        int c3[3];
        for ( int k = 0; k < 3; ++k )
            c3[k] = grid_cell_coord( x[3 * i + k], gmin[k], rd[k], ng[k], 1 );
        cid[i] = ( c3[0] * ng[1] + c3[1] ) * ng[2] + c3[2];
        count[cid[i]]++;
    }
    std::vector<int> off( ncell + 1, 0 );
    for ( int c = 0; c < ncell; ++c )
        off[c + 1] = off[c] + count[c];
    for ( int i = 0; i < n; ++i )
    {
        int slot = off[cid[i]]++;
        perm[slot] = i;
        if ( cell_of )
            cell_of[slot] = cid[i];
    }
    return ncell;
}

void orc_permute( int n, const int *perm, double *x, double *v, int *type,
                  int *id )
{
    std::vector<double> tx( 3 * (size_t)n ), tv( 3 * (size_t)n );
    std::vector<int> tt( n ), ti( n );
    for ( int i = 0; i < n; ++i )
    {
        int s = perm[i];
        for ( int k = 0; k < 3; ++k )
        {
            if ( x )
                tx[3 * i + k] = x[3 * s + k];
            if ( v )
                tv[3 * i + k] = v[3 * s + k];
        }
        if ( type )
            tt[i] = type[s];
        if ( id )
            ti[i] = id[s];
    }
    if ( x )
        std::memcpy( x, tx.data(), sizeof( double ) * 3 * n );
    if ( v )
        std::memcpy( v, tv.data(), sizeof( double ) * 3 * n );
    if ( type )
        std::memcpy( type, tt.data(), sizeof( int ) * n );
    if ( id )
        std::memcpy( id, ti.data(), sizeof( int ) * n );
}

/* ------------------------------------------------------------------------ */
/* Ghost build: comm_mpi_impl.h:280-367 with TagHaloPack comm_mpi.h:240-319
 * and TagHaloPBC comm_mpi.h:323-353, one rank.  Phase p packs atoms of the
 * current set (odd phases exclude the ghosts received in phase p-1) whose
 * coordinate d = p/2 satisfies x_d >= hi_d - depth (even p) or
 * x_d <= lo_d + depth (odd p); the appended ghost is x_src shifted by -L_d
 * (even) or +L_d (odd).  Pack order is index order (Serial backend). */
int orc_ghost_build( int n_local, int cap, double *x, int *type, int *id,
                     const orc_domain *d, double depth, int *send_idx,
                     int counts[6] )
{
    int n_ghost = 0;
    int nsend_total = 0;
    for ( int phase = 0; phase < 6; ++phase )
    {
        const int dim = phase / 2;
        const int nparticles =
            n_local + n_ghost - ( ( phase % 2 == 1 ) ? counts[phase - 1] : 0 );
        int cnt = 0;
        const int base = n_local + n_ghost;
        for ( int i = 0; i < nparticles; ++i )
        {
            bool take = ( phase % 2 == 0 )
                            ? ( x[3 * i + dim] >= d->local_hi[dim] - depth )
                            : ( x[3 * i + dim] <= d->local_lo[dim] + depth );
            if ( !take )
                continue;
            if ( base + cnt >= cap || nsend_total + cnt >= cap )
                return -1;
            send_idx[nsend_total + cnt] = i;
            cnt++;
        }
        for ( int k = 0; k < cnt; ++k )
        {
            const int s = send_idx[nsend_total + k];
            const int g = base + k;
            for ( int c = 0; c < 3; ++c )
                x[3 * g + c] = x[3 * s + c];
            type[g] = type[s];
            if ( id )
                id[g] = id[s];
            if ( phase % 2 == 0 )
                x[3 * g + dim] -= d->L[dim];
            else
                x[3 * g + dim] += d->L[dim];
        }
        counts[phase] = cnt;
        nsend_total += cnt;
        n_ghost += cnt;
    }
    return n_ghost;
}

/* Ghost refresh: comm_mpi_impl.h:369-408 (gather x, then PBC shift, phase by
 * phase so chained ghosts see refreshed sources). */
void orc_ghost_refresh( int n_local, double *x, const orc_domain *d,
                        const int *send_idx, const int counts[6] )
{
    int base = n_local, off = 0;
    for ( int phase = 0; phase < 6; ++phase )
    {
        const int dim = phase / 2;
        for ( int k = 0; k < counts[phase]; ++k )
        {
            const int s = send_idx[off + k];
            const int g = base + k;
            for ( int c = 0; c < 3; ++c )
                x[3 * g + c] = x[3 * s + c];
            if ( phase % 2 == 0 )
                x[3 * g + dim] -= d->L[dim];
            else
                x[3 * g + dim] += d->L[dim];
        }
        base += counts[phase];
        off += counts[phase];
    }
}

/* Ghost force fold: comm_mpi_impl.h:410-441 (Cabana::scatter, phases 5..0,
 * adds each ghost's force into its source). */
void orc_ghost_fold( int n_local, double *f, const int *send_idx,
                     const int counts[6] )
{
    int starts[6], offs[6];
    int base = n_local, off = 0;
    for ( int p = 0; p < 6; ++p )
    {
        starts[p] = base;
        offs[p] = off;
        base += counts[p];
        off += counts[p];
    }
    for ( int phase = 5; phase >= 0; --phase )
        for ( int k = 0; k < counts[phase]; ++k )
        {
            const int s = send_idx[offs[phase] + k];
            const int g = starts[phase] + k;
            for ( int c = 0; c < 3; ++c )
                f[3 * s + c] += f[3 * g + c];
        }
}

/* ------------------------------------------------------------------------ */
/* Verlet list accept rule (Cabana VerletList, FullNeighborTag /
 * HalfNeighborTag discriminators; list inclusion uses rsq <= rc^2 as the
 * reference's own brute force does, tstNeighbor.hpp:97,130). */
static inline bool neigh_accept( const double *x, int i, int j, double rc2,
                                 int half )
{
    if ( i == j )
        return false;
    const double xi = x[3 * i], yi = x[3 * i + 1], zi = x[3 * i + 2];
    const double xj = x[3 * j], yj = x[3 * j + 1], zj = x[3 * j + 2];
    if ( half &&
         !( ( xj > xi ) ||
            ( ( xj == xi ) && ( ( yj > yi ) || ( ( yj == yi ) && ( zj > zi ) ) ) ) ) )
        return false;
    const double dx = xi - xj;
    const double dy = yi - yj;
    const double dz = zi - zj;
    const double rsq = dx * dx + dy * dy + dz * dz;
    return rsq <= rc2;
}

int64_t orc_neigh_brute( int n_local, int n_total, const double *x, double rc,
                         int half, int64_t *offsets, int *idx )
{
    const double rc2 = rc * rc;
    int64_t nnz = 0;
    for ( int i = 0; i < n_local; ++i )
    {
        if ( offsets )
            offsets[i] = nnz;
        for ( int j = 0; j < n_total; ++j )
            if ( neigh_accept( x, i, j, rc2, half ) )
            {
                if ( idx )
                    idx[nnz] = j;
                nnz++;
            }
    }
    if ( offsets )
        offsets[n_local] = nnz;
    return nnz;
}

/* Cell-list build on a given grid: rows walk the 27-cell stencil in
 * (x-cell, y-cell, z-cell) order, atoms of a cell in index order. */
int64_t orc_neigh_build_grid( int n_local, int n_total, const double *x,
                              double rc, int half, const double gmin[3],
                              const double rd[3], const int ng[3],
                              int floor_mode, int64_t *offsets, int *idx )
{
    const double rc2 = rc * rc;
    const int ncell = ng[0] * ng[1] * ng[2];
    auto cell3 = [&]( int i, int c[3] )
    {
        for ( int k = 0; k < 3; ++k )
            c[k] = grid_cell_coord( x[3 * i + k], gmin[k], rd[k], ng[k],
                                    floor_mode );
    };
    std::vector<int> cnt( ncell + 1, 0 ), start( ncell + 1, 0 ),
        sorted( std::max( n_total, 1 ) );
    std::vector<int> cid( std::max( n_total, 1 ) );
    for ( int i = 0; i < n_total; ++i )
    {
        int c[3];
        cell3( i, c );
        cid[i] = ( c[0] * ng[1] + c[1] ) * ng[2] + c[2];
        cnt[cid[i]]++;
    }
    for ( int c = 0; c < ncell; ++c )
        start[c + 1] = start[c] + cnt[c];
    {
        std::vector<int> cur( start.begin(), start.end() - 1 );
        for ( int i = 0; i < n_total; ++i )
            sorted[cur[cid[i]]++] = i;
    }
    auto visit_row = [&]( int i, int *out ) -> int
    {
        int c[3];
        cell3( i, c );
        int m = 0;
        for ( int a = std::max( c[0] - 1, 0 ); a <= std::min( c[0] + 1, ng[0] - 1 ); ++a )
            for ( int b = std::max( c[1] - 1, 0 ); b <= std::min( c[1] + 1, ng[1] - 1 ); ++b )
                for ( int e = std::max( c[2] - 1, 0 ); e <= std::min( c[2] + 1, ng[2] - 1 ); ++e )
                {
                    int cc = ( a * ng[1] + b ) * ng[2] + e;
                    for ( int s = start[cc]; s < start[cc + 1]; ++s )
                    {
                        int j = sorted[s];
                        if ( neigh_accept( x, i, j, rc2, half ) )
                        {
                            if ( out )
                                out[m] = j;
                            m++;
                        }
                    }
                }
        return m;
    };
    std::vector<int64_t> rowcnt( n_local + 1, 0 );
#pragma omp parallel for schedule( dynamic, 256 )
    for ( int i = 0; i < n_local; ++i )
        rowcnt[i] = visit_row( i, nullptr );
    int64_t nnz = 0;
    for ( int i = 0; i < n_local; ++i )
    {
        if ( offsets )
            offsets[i] = nnz;
        nnz += rowcnt[i];
    }
    if ( offsets )
        offsets[n_local] = nnz;
    if ( idx && offsets )
    {
#pragma omp parallel for schedule( dynamic, 256 )
        for ( int i = 0; i < n_local; ++i )
            visit_row( i, idx + offsets[i] );
    }
    return nnz;
}

static void bbox_of( int n, const double *x, double lo[3], double hi[3] )
{
    for ( int k = 0; k < 3; ++k )
    {
        lo[k] = std::numeric_limits<double>::max();
        hi[k] = -std::numeric_limits<double>::max();
    }
    for ( int i = 0; i < n; ++i )
        for ( int k = 0; k < 3; ++k )
        {
            lo[k] = std::min( lo[k], x[3 * i + k] );
            hi[k] = std::max( hi[k], x[3 * i + k] );
        }
}

/* Bounding-box grid: cells >= rc per dimension, at most 1024 per dimension
 * and at most max(64, 2 n_total) in total (halving the largest dimension). */
static void bbox_grid( const double lo[3], const double hi[3], double rc,
                       int n_total, double gmin[3], double rd[3], int ng[3] )
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
    {
        gmin[k] = lo[k];
        rd[k] = ng[k] / std::max( ext[k], 1e-300 );
    }
}

int64_t orc_neigh_build( int n_local, int n_total, const double *x, double rc,
                         int half, int64_t *offsets, int *idx )
{
    if ( n_total == 0 )
    {
        if ( offsets )
            offsets[0] = 0;
        return 0;
    }
    double lo[3], hi[3], gmin[3], rd[3];
    int ng[3];
    bbox_of( n_total, x, lo, hi );
    bbox_grid( lo, hi, rc, n_total, gmin, rd, ng );
    return orc_neigh_build_grid( n_local, n_total, x, rc, half, gmin, rd, ng, 0,
                                 offsets, idx );
}

/* ------------------------------------------------------------------------ */
/* LJ coefficients: force_lj_cabana_neigh_impl.h:62-89. */
void orc_lj_coeff( int ntypes, int ncoeff, const int *ci, const int *cj,
                   const double *eps, const double *sigma, const double *cut,
                   double *lj1, double *lj2, double *cutsq )
{
    for ( int a = 0; a < ntypes * ntypes; ++a )
        lj1[a] = lj2[a] = cutsq[a] = 0.0;
    for ( int a = 0; a < ncoeff; ++a )
    {
        int i = ci[a] - 1, j = cj[a] - 1;
        lj1[i * ntypes + j] = 48.0 * eps[a] * std::pow( sigma[a], 12.0 );
        lj2[i * ntypes + j] = 24.0 * eps[a] * std::pow( sigma[a], 6.0 );
        cutsq[i * ntypes + j] = cut[a] * cut[a];
        lj1[j * ntypes + i] = lj1[i * ntypes + j];
        lj2[j * ntypes + i] = lj2[i * ntypes + j];
        cutsq[j * ntypes + i] = cutsq[i * ntypes + j];
    }
}

/* Full: force_lj_cabana_neigh_impl.h:151-203 (force on i only; the pair
 * force is unshifted and applied for rsq < cutsq).  Half: :205-259 (the
 * negated pair force also lands on j, possibly a ghost).  f is accumulated
 * into, as the reference does after zeroing it (cabanamd_impl.h:336-341). */
void orc_lj_force( int n_local, int n_total, const double *x, const int *type,
                   int ntypes, const double *lj1, const double *lj2,
                   const double *cutsq, const int64_t *offsets,
                   const int *idx, int half, double *f )
{
    (void)n_total;
    auto pair = [&]( int i, int j, double &dx, double &dy, double &dz,
                     double &fpair ) -> bool
    {
        dx = x[3 * i] - x[3 * j];
        dy = x[3 * i + 1] - x[3 * j + 1];
        dz = x[3 * i + 2] - x[3 * j + 2];
        const int t = type[i] * ntypes + type[j];
        const double rsq = dx * dx + dy * dy + dz * dz;
        if ( !( rsq < cutsq[t] ) )
            return false;
        double r2inv = 1.0 / rsq;
        double r6inv = r2inv * r2inv * r2inv;
        fpair = ( r6inv * ( lj1[t] * r6inv - lj2[t] ) ) * r2inv;
        return true;
    };
    if ( !half )
    {
#pragma omp parallel for schedule( static )
        for ( int i = 0; i < n_local; ++i )
        {
            for ( int64_t k = offsets[i]; k < offsets[i + 1]; ++k )
            {
                double dx, dy, dz, fpair;
                double fx = 0.0, fy = 0.0, fz = 0.0;
                if ( pair( i, idx[k], dx, dy, dz, fpair ) )
                {
                    fx += dx * fpair;
                    fy += dy * fpair;
                    fz += dz * fpair;
                }
                f[3 * i] += fx;
                f[3 * i + 1] += fy;
                f[3 * i + 2] += fz;
            }
        }
        return;
    }
    for ( int i = 0; i < n_local; ++i )
        for ( int64_t k = offsets[i]; k < offsets[i + 1]; ++k )
        {
            const int j = idx[k];
            double dx, dy, dz, fpair;
            double fx = 0.0, fy = 0.0, fz = 0.0;
            if ( pair( i, j, dx, dy, dz, fpair ) )
            {
                fx += dx * fpair;
                fy += dy * fpair;
                fz += dz * fpair;
                f[3 * j] -= dx * fpair;
                f[3 * j + 1] -= dy * fpair;
                f[3 * j + 2] -= dz * fpair;
            }
            f[3 * i] += fx;
            f[3 * i + 1] += fy;
            f[3 * i + 2] += fz;
        }
}

/* Shifted LJ energy: full :261-315 (weight 1/2 per listed pair), half
 * :317-377 (weight 1 if j is local, 1/2 if j is a ghost). */
double orc_lj_energy( int n_local, const double *x, const int *type,
                      int ntypes, const double *lj1, const double *lj2,
                      const double *cutsq, const int64_t *offsets,
                      const int *idx, int half )
{
    // fixed chunks summed in order: the same result for any thread count
    const int nchunk = ( n_local + RED_CHUNK - 1 ) / RED_CHUNK;
    std::vector<double> part( nchunk, 0.0 );
#pragma omp parallel for schedule( static )
    for ( int c = 0; c < nchunk; ++c )
    {
        CompSum pe;
        const int i1 = std::min( n_local, ( c + 1 ) * RED_CHUNK );
        for ( int i = c * RED_CHUNK; i < i1; ++i )
            for ( int64_t k = offsets[i]; k < offsets[i + 1]; ++k )
            {
                const int j = idx[k];
                const double dx = x[3 * i] - x[3 * j];
                const double dy = x[3 * i + 1] - x[3 * j + 1];
                const double dz = x[3 * i + 2] - x[3 * j + 2];
                const int t = type[i] * ntypes + type[j];
                const double rsq = dx * dx + dy * dy + dz * dz;
                if ( rsq < cutsq[t] )
                {
                    const double fac =
                        half ? ( ( j < n_local ) ? 1.0 : 0.5 ) : 0.5;
                    double r2inv = 1.0 / rsq;
                    double r6inv = r2inv * r2inv * r2inv;
                    pe.add( fac * r6inv * ( 0.5 * lj1[t] * r6inv - lj2[t] ) / 6.0 );
                    double r2invc = 1.0 / cutsq[t];
                    double r6invc = r2invc * r2invc * r2invc;
                    pe.add( -( fac * r6invc * ( 0.5 * lj1[t] * r6invc - lj2[t] ) / 6.0 ) );
                }
            }
        part[c] = pe.value();
    }
    CompSum pe;
    for ( int c = 0; c < nchunk; ++c )
        pe.add( part[c] );
    return pe.value();
}

double orc_lj_virial( int n_local, const double *x, const int *type,
                      int ntypes, const double *lj1, const double *lj2,
                      const double *cutsq, const int64_t *offsets,
                      const int *idx, int half )
{
    const double fac = half ? 1.0 : 0.5;
    const int nchunk = ( n_local + RED_CHUNK - 1 ) / RED_CHUNK;
    std::vector<double> part( nchunk, 0.0 );
#pragma omp parallel for schedule( static )
    for ( int c = 0; c < nchunk; ++c )
    {
        CompSum w;
        const int i1 = std::min( n_local, ( c + 1 ) * RED_CHUNK );
        for ( int i = c * RED_CHUNK; i < i1; ++i )
            for ( int64_t k = offsets[i]; k < offsets[i + 1]; ++k )
            {
                const int j = idx[k];
                const double dx = x[3 * i] - x[3 * j];
                const double dy = x[3 * i + 1] - x[3 * j + 1];
                const double dz = x[3 * i + 2] - x[3 * j + 2];
                const int t = type[i] * ntypes + type[j];
                const double rsq = dx * dx + dy * dy + dz * dz;
                if ( rsq < cutsq[t] )
                {
                    double r2inv = 1.0 / rsq;
                    double r6inv = r2inv * r2inv * r2inv;
                    double fpair = ( r6inv * ( lj1[t] * r6inv - lj2[t] ) ) * r2inv;
                    w.add( fac * rsq * fpair );
                }
            }
        part[c] = w.value();
    }
    CompSum w;
    for ( int c = 0; c < nchunk; ++c )
        w.add( part[c] );
    return w.value();
}

/* ------------------------------------------------------------------------ */
/* Thermo: property_temperature.h:71-77 + impl :55-77; property_kine_impl.h:55-76
 * (compensated sequential sums here; the reference's reduction order is
 * unspecified). */
double orc_temperature( int n_local, int n_global, const double *v,
                        const int *type, const double *mass, double mvv2e,
                        double boltz )
{
    CompSum T;
    for ( int i = 0; i < n_local; ++i )
        T.add( ( v[3 * i] * v[3 * i] + v[3 * i + 1] * v[3 * i + 1] +
                 v[3 * i + 2] * v[3 * i + 2] ) *
               mass[type[i]] );
    const int dof = 3 * n_global - 3;
    const double factor = mvv2e / ( 1.0 * dof * boltz );
    return T.value() * factor;
}

double orc_kinetic_energy( int n_local, const double *v, const int *type,
                           const double *mass, double mvv2e )
{
    CompSum ke;
    for ( int i = 0; i < n_local; ++i )
        ke.add( ( v[3 * i] * v[3 * i] + v[3 * i + 1] * v[3 * i + 1] +
                  v[3 * i + 2] * v[3 * i + 2] ) *
                mass[type[i]] );
    return ke.value() * ( 0.5 * mvv2e );
}

/* ------------------------------------------------------------------------ */
/* Timestep loop: CbnMD::init cabanamd_impl.h:70-255 and CbnMD::run :257-432. */
struct orc_md
{
    orc_params p;
    orc_domain d;
    int N = 0, n_local = 0, n_ghost = 0, cap = 0;
    double neigh_cut = 0.0;
    std::vector<double> x, v, f;
    std::vector<int> type, id, send_idx;
    int counts[6] = { 0, 0, 0, 0, 0, 0 };
    std::vector<double> lj1, lj2, cutsq;
    std::vector<int64_t> off;
    std::vector<int> idx;
    std::vector<double> thermo;
    double t_force = 0, t_neigh = 0, t_comm = 0, t_int = 0, t_other = 0,
           t_total = 0;

    void grow( int need )
    {
        if ( need <= cap )
            return;
        cap = need;
        x.resize( 3 * (size_t)cap );
        v.resize( 3 * (size_t)cap );
        f.resize( 3 * (size_t)cap );
        type.resize( cap );
        id.resize( cap );
        send_idx.resize( cap );
    }
    void ghosts()
    {
        for ( ;; )
        {
            int ng = orc_ghost_build( n_local, cap, x.data(), type.data(),
                                      id.data(), &d, neigh_cut,
                                      send_idx.data(), counts );
            if ( ng >= 0 )
            {
                n_ghost = ng;
                return;
            }
            grow( (int)( cap * 1.5 ) + 64 );
        }
    }
    /* The engine's grid rule: the binning grid when its cells are >= rc and
     * it holds every atom, otherwise a grid over the bounding box. */
    void neighbors()
    {
        int nt = n_local + n_ghost;
        off.resize( n_local + 1 );
        double lo[3], hi[3], gmin[3], rd[3];
        int ng[3], fm = 1;
        bbox_of( nt, x.data(), lo, hi );
        double delta[3] = { neigh_cut, neigh_cut, neigh_cut };
        orc_bin_grid( &d, delta, 1, gmin, rd, ng );
        bool use_bin = true;
        for ( int k = 0; k < 3; ++k )
            if ( !( 1.0 / rd[k] >= neigh_cut ) || !( lo[k] >= gmin[k] ) ||
                 !( hi[k] < gmin[k] + ng[k] / rd[k] ) )
                use_bin = false;
        if ( !use_bin )
        {
            bbox_grid( lo, hi, neigh_cut, nt, gmin, rd, ng );
            fm = 0;
        }
        int64_t nnz = orc_neigh_build_grid( n_local, nt, x.data(), neigh_cut,
                                            p.half, gmin, rd, ng, fm, off.data(),
                                            nullptr );
        idx.resize( std::max<int64_t>( nnz, 1 ) );
        orc_neigh_build_grid( n_local, nt, x.data(), neigh_cut, p.half, gmin, rd,
                              ng, fm, off.data(), idx.data() );
    }
    void sort()
    {
        double delta[3] = { neigh_cut, neigh_cut, neigh_cut };
        std::vector<int> perm( n_local );
        orc_bin_sort( n_local, x.data(), &d, delta, 1, perm.data(), nullptr );
        orc_permute( n_local, perm.data(), x.data(), v.data(), type.data(),
                     id.data() );
    }
    void force()
    {
        std::fill( f.begin(), f.begin() + 3 * (size_t)( n_local + n_ghost ),
                   0.0 );
        orc_lj_force( n_local, n_local + n_ghost, x.data(), type.data(),
                      p.ntypes, lj1.data(), lj2.data(), cutsq.data(),
                      off.data(), idx.data(), p.half, f.data() );
    }
    void record( int step )
    {
        double T = orc_temperature( n_local, N, v.data(), type.data(), p.mass,
                                    p.mvv2e, p.boltz );
        double PE = orc_lj_energy( n_local, x.data(), type.data(), p.ntypes,
                                   lj1.data(), lj2.data(), cutsq.data(),
                                   off.data(), idx.data(), p.half ) /
                    N;
        double KE = orc_kinetic_energy( n_local, v.data(), type.data(), p.mass,
                                        p.mvv2e ) /
                    N;
        thermo.push_back( step );
        thermo.push_back( T );
        thermo.push_back( PE );
        thermo.push_back( PE + KE );
    }
};

orc_md *orc_md_create( const orc_params *p )
{
    orc_md *md = new orc_md;
    md->p = *p;
    return md;
}

void orc_md_destroy( orc_md *md ) { delete md; }

int orc_md_init( orc_md *md )
{
    const orc_params &p = md->p;
    md->neigh_cut = p.force_cutoff + p.skin;
    const double a = p.lattice_constant;
    double low[3], high[3];
    for ( int k = 0; k < 3; ++k )
    {
        low[k] = a * p.region_lo[k];
        high[k] = a * p.region_hi[k];
    }
    orc_create_domain( low, high, p.comm_ghost_cutoff, &md->d );
    int n = orc_lattice_count( &p, &md->d );
    md->N = md->n_local = n;
    md->grow( n + n / 2 + 1024 );
    orc_lattice_fill( &p, &md->d, md->x.data(), md->type.data(),
                      md->id.data() );
    orc_velocity_create( &p, n, md->x.data(), md->type.data(), md->v.data() );
    md->lj1.assign( p.ntypes * p.ntypes, 0.0 );
    md->lj2.assign( p.ntypes * p.ntypes, 0.0 );
    md->cutsq.assign( p.ntypes * p.ntypes, 0.0 );
    orc_lj_coeff( p.ntypes, p.ncoeff, p.coeff_i, p.coeff_j, p.coeff_eps,
                  p.coeff_sigma, p.coeff_cut, md->lj1.data(), md->lj2.data(),
                  md->cutsq.data() );
    orc_wrap( md->n_local, md->x.data(), md->d.L );
    md->n_ghost = 0;
    md->sort();
    md->ghosts();
    md->neighbors();
    md->force();
    if ( p.half )
        orc_ghost_fold( md->n_local, md->f.data(), md->send_idx.data(),
                        md->counts );
    if ( p.thermo_rate > 0 )
        md->record( 0 );
    return 0;
}

int orc_md_run( orc_md *md, int nsteps )
{
    const orc_params &p = md->p;
    auto t_all = clk::now();
    for ( int step = 1; step <= nsteps; ++step )
    {
        auto t0 = clk::now();
        orc_nve_initial( md->n_local, md->x.data(), md->v.data(), md->f.data(),
                         md->type.data(), p.mass, p.dt, p.mvv2e );
        md->t_int += seconds_since( t0 );
        if ( step % p.every == 0 )
        {
            t0 = clk::now();
            orc_wrap( md->n_local, md->x.data(), md->d.L );
            md->n_ghost = 0;
            md->t_comm += seconds_since( t0 );
            t0 = clk::now();
            md->sort();
            md->t_other += seconds_since( t0 );
            t0 = clk::now();
            md->ghosts();
            md->t_comm += seconds_since( t0 );
            t0 = clk::now();
            md->neighbors();
            md->t_neigh += seconds_since( t0 );
        }
        else
        {
            t0 = clk::now();
            orc_ghost_refresh( md->n_local, md->x.data(), &md->d,
                               md->send_idx.data(), md->counts );
            md->t_comm += seconds_since( t0 );
        }
        t0 = clk::now();
        md->force();
        md->t_force += seconds_since( t0 );
        if ( p.half )
        {
            t0 = clk::now();
            orc_ghost_fold( md->n_local, md->f.data(), md->send_idx.data(),
                            md->counts );
            md->t_comm += seconds_since( t0 );
        }
        t0 = clk::now();
        orc_nve_final( md->n_local, md->v.data(), md->f.data(),
                       md->type.data(), p.mass, p.dt, p.mvv2e );
        md->t_int += seconds_since( t0 );
        t0 = clk::now();
        if ( p.thermo_rate > 0 && step % p.thermo_rate == 0 )
            md->record( step );
        md->t_other += seconds_since( t0 );
    }
    md->t_total += seconds_since( t_all );
    return 0;
}

int orc_md_n_local( const orc_md *md ) { return md->n_local; }
int orc_md_n_ghost( const orc_md *md ) { return md->n_ghost; }
int64_t orc_md_nnz( const orc_md *md ) { return md->off.empty() ? 0 : md->off.back(); }

void orc_md_get( const orc_md *md, double *x, double *v, double *f, int *type,
                 int *id )
{
    const size_t n = md->n_local + md->n_ghost;
    if ( x )
        std::memcpy( x, md->x.data(), sizeof( double ) * 3 * n );
    if ( v )
        std::memcpy( v, md->v.data(), sizeof( double ) * 3 * n );
    if ( f )
        std::memcpy( f, md->f.data(), sizeof( double ) * 3 * n );
    if ( type )
        std::memcpy( type, md->type.data(), sizeof( int ) * n );
    if ( id )
        std::memcpy( id, md->id.data(), sizeof( int ) * n );
}

int orc_md_thermo( const orc_md *md, int max_rows, double *rows )
{
    int n = (int)( md->thermo.size() / 4 );
    int m = std::min( n, max_rows );
    if ( rows )
        std::memcpy( rows, md->thermo.data(), sizeof( double ) * 4 * m );
    return n;
}

void orc_md_timers( const orc_md *md, double out[6] )
{
    out[0] = md->t_force;
    out[1] = md->t_neigh;
    out[2] = md->t_comm;
    out[3] = md->t_int;
    out[4] = md->t_other;
    out[5] = md->t_total;
}
