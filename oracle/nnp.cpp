/*
 * nnp.cpp -- FP64 restatement of the Behler-Parrinello neural-network
 * potential used by ForceNNP (force_types/force_nnp_cabana_neigh_impl.h).
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * The reference delegates all NNP arithmetic to n2p2 v2.0.1 (nnp::ModeCabana,
 * not present here); its call sequence is force_nnp_cabana_neigh_impl.h:
 * 68-99 (symmetry functions -> atomic networks -> forces) and :101-130
 * (energy assembly).  The definitions restated here are n2p2's published
 * ones for the input.nn keywords the reference fixture uses
 * (input/nnp/input.nn):
 *   cutoff_type 1         fc(r) = (cos(pi r / rc) + 1) / 2 for r < rc
 *   symfunction type 2    G = sum_j exp(-eta (r_ij - rs)^2) fc(r_ij)
 *                          over neighbours j of element e1 (e1 = -1: any)
 *   symfunction type 3    G = 2^(1-zeta) sum_{j<k} (1 + lambda cos t_ijk)^zeta
 *                          exp(-eta sum_{ij,ik,jk} (r - rs)^2)
 *                          fc(r_ij) fc(r_ik) fc(r_jk)
 *                          over pairs whose elements are {e1, e2}
 *                          (e1 = e2 = -1: any pair)
 *   scale_symmetry_functions  Gs = smin + (smax - smin)(G - Gmin)/(Gmax - Gmin)
 *   global_activation t / l   tanh / identity per layer
 * Forces are the analytic gradient F_k = -sum_i sum_s dE_i/dG_is dG_is/dx_k.
 * Parity with n2p2 itself is unpinned (the library cannot be imported);
 * these functions are pinned by finite differences, invariances and an
 * independent brute-force evaluation in the tests.
 */
#include "oracle.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <vector>

namespace
{
const double PI = 3.14159265358979323846;

struct SF
{
    int type, e1, e2;
    double eta, rs, lambda, zeta, rc;
};

struct Element
{
    std::vector<SF> sf;
    std::vector<double> gmin, gmax;
    std::vector<std::vector<double>> W, b; // per layer: W [out][in], b [out]
    double atom_energy = 0.0;
};
} // namespace

struct orc_nnp
{
    int nelem = 0;
    int nlayers = 0;          // weight layers
    std::vector<int> nodes;   // nodes[0] = inputs (per element: nsf) ...
    std::vector<char> act;    // per weight layer: 't' or 'l'
    double smin = 0.0, smax = 1.0;
    int scale = 1;
    std::vector<Element> el;
};

namespace
{
inline double fcut( double r, double rc )
{
    return r < rc ? 0.5 * ( std::cos( PI * r / rc ) + 1.0 ) : 0.0;
}
inline double dfcut( double r, double rc )
{
    return r < rc ? -0.5 * PI / rc * std::sin( PI * r / rc ) : 0.0;
}
inline bool pair_match( const SF &s, int ej, int ek )
{
    if ( s.e1 < 0 )
        return true;
    return ( ej == s.e1 && ek == s.e2 ) || ( ej == s.e2 && ek == s.e1 );
}

struct Nb
{
    int j;
    double d[3]; // x_j - x_i
    double r;
};

// Neighbours of atom i inside the largest symmetry-function cutoff.
void neighbours( const orc_nnp *m, int i, const double *x, const int64_t *off,
                 const int *idx, double rcmax, std::vector<Nb> &out )
{
    out.clear();
    (void)m;
    for ( int64_t k = off[i]; k < off[i + 1]; ++k )
    {
        const int j = idx[k];
        Nb n;
        n.j = j;
        for ( int c = 0; c < 3; ++c )
            n.d[c] = x[3 * j + c] - x[3 * i + c];
        n.r = std::sqrt( n.d[0] * n.d[0] + n.d[1] * n.d[1] + n.d[2] * n.d[2] );
        if ( n.r < rcmax )
            out.push_back( n );
    }
}

double rc_max( const orc_nnp *m )
{
    double r = 0.0;
    for ( const auto &e : m->el )
        for ( const auto &s : e.sf )
            r = std::max( r, s.rc );
    return r;
}

// G_s of atom i and, when dG is given, dG_s/dx for i (slot 0) and each
// neighbour (slot 1 + n): dG[s][(1 + nn) * 3].
void symmetry_functions( const orc_nnp *m, int ei, const std::vector<Nb> &nb,
                         const int *type, std::vector<double> &G,
                         std::vector<std::vector<double>> *dG )
{
    const Element &E = m->el[ei];
    const int ns = (int)E.sf.size();
    const int nn = (int)nb.size();
    G.assign( ns, 0.0 );
    if ( dG )
    {
        dG->assign( ns, std::vector<double>( 3 * ( nn + 1 ), 0.0 ) );
    }
    for ( int s = 0; s < ns; ++s )
    {
        const SF &f = E.sf[s];
        if ( f.type == 2 )
        {
            for ( int a = 0; a < nn; ++a )
            {
                if ( ( f.e1 >= 0 && type[nb[a].j] != f.e1 ) || nb[a].r >= f.rc )
                    continue;
                const double r = nb[a].r;
                const double ex = std::exp( -f.eta * ( r - f.rs ) * ( r - f.rs ) );
                G[s] += ex * fcut( r, f.rc );
                if ( dG )
                {
                    const double g1 = ex * ( -2.0 * f.eta * ( r - f.rs ) * fcut( r, f.rc ) +
                                             dfcut( r, f.rc ) );
                    for ( int c = 0; c < 3; ++c )
                    {
                        const double u = nb[a].d[c] / r; // d r / d x_j
                        ( *dG )[s][3 * ( 1 + a ) + c] += g1 * u;
                        ( *dG )[s][c] -= g1 * u;
                    }
                }
            }
        }
        else
        {
            const double pref = std::pow( 2.0, 1.0 - f.zeta );
            for ( int a = 0; a < nn; ++a )
                for ( int q = a + 1; q < nn; ++q )
                {
                    if ( !pair_match( f, type[nb[a].j], type[nb[q].j] ) )
                        continue;
                    const double *A = nb[a].d, *B = nb[q].d;
                    const double r1 = nb[a].r, r2 = nb[q].r;
                    double Cv[3];
                    for ( int c = 0; c < 3; ++c )
                        Cv[c] = B[c] - A[c];
                    const double r3 =
                        std::sqrt( Cv[0] * Cv[0] + Cv[1] * Cv[1] + Cv[2] * Cv[2] );
                    if ( r1 >= f.rc || r2 >= f.rc || r3 >= f.rc )
                        continue;
                    const double ab = A[0] * B[0] + A[1] * B[1] + A[2] * B[2];
                    const double cs = ab / ( r1 * r2 );
                    const double base = 1.0 + f.lambda * cs;
                    const double P = std::pow( base, f.zeta );
                    const double ex = std::exp(
                        -f.eta * ( ( r1 - f.rs ) * ( r1 - f.rs ) + ( r2 - f.rs ) * ( r2 - f.rs ) +
                                   ( r3 - f.rs ) * ( r3 - f.rs ) ) );
                    const double f1 = fcut( r1, f.rc ), f2 = fcut( r2, f.rc ),
                                 f3 = fcut( r3, f.rc );
                    const double F = f1 * f2 * f3;
                    G[s] += pref * P * ex * F;
                    if ( !dG )
                        continue;
                    // d/dA and d/dB of P * ex * F
                    const double dP = f.zeta * f.lambda *
                                      ( f.zeta == 1.0 ? 1.0 : std::pow( base, f.zeta - 1.0 ) );
                    const double d1 = dfcut( r1, f.rc ), d2 = dfcut( r2, f.rc ),
                                 d3 = dfcut( r3, f.rc );
                    for ( int c = 0; c < 3; ++c )
                    {
                        const double dcos_a = B[c] / ( r1 * r2 ) - cs * A[c] / ( r1 * r1 );
                        const double dcos_b = A[c] / ( r1 * r2 ) - cs * B[c] / ( r2 * r2 );
                        const double u1 = A[c] / r1, u2 = B[c] / r2, u3 = Cv[c] / r3;
                        // r3 = |B - A|: dr3/dA = -u3, dr3/dB = +u3
                        const double dex_a =
                            ex * ( -2.0 * f.eta ) * ( ( r1 - f.rs ) * u1 - ( r3 - f.rs ) * u3 );
                        const double dex_b =
                            ex * ( -2.0 * f.eta ) * ( ( r2 - f.rs ) * u2 + ( r3 - f.rs ) * u3 );
                        const double dF_a = d1 * u1 * f2 * f3 - f1 * f2 * d3 * u3;
                        const double dF_b = f1 * d2 * u2 * f3 + f1 * f2 * d3 * u3;
                        const double ga =
                            pref * ( dP * dcos_a * ex * F + P * dex_a * F + P * ex * dF_a );
                        const double gb =
                            pref * ( dP * dcos_b * ex * F + P * dex_b * F + P * ex * dF_b );
                        ( *dG )[s][3 * ( 1 + a ) + c] += ga;
                        ( *dG )[s][3 * ( 1 + q ) + c] += gb;
                        ( *dG )[s][c] -= ga + gb;
                    }
                }
        }
    }
}

// Network energy of one atom from its raw symmetry functions; dEdG gets
// dE/dG (raw, i.e. including the scaling derivative).
double network( const orc_nnp *m, int ei, const std::vector<double> &G,
                std::vector<double> *dEdG )
{
    const Element &E = m->el[ei];
    const int L = m->nlayers;
    std::vector<std::vector<double>> h( L + 1 ), z( L );
    const int ns = (int)E.sf.size();
    h[0].resize( ns );
    for ( int s = 0; s < ns; ++s )
    {
        if ( m->scale )
        {
            const double w = E.gmax[s] - E.gmin[s];
            h[0][s] = w > 0.0 ? m->smin + ( m->smax - m->smin ) * ( G[s] - E.gmin[s] ) / w
                              : m->smin;
        }
        else
            h[0][s] = G[s];
    }
    for ( int l = 0; l < L; ++l )
    {
        const auto &W = E.W[l];
        const auto &b = E.b[l];
        const int no = (int)b.size();
        z[l].assign( no, 0.0 );
        h[l + 1].assign( no, 0.0 );
        for ( int o = 0; o < no; ++o )
        {
            double acc = b[o];
            for ( size_t k = 0; k < h[l].size(); ++k )
                acc += W[(size_t)o * h[l].size() + k] * h[l][k];
            z[l][o] = acc;
            h[l + 1][o] = m->act[l] == 't' ? std::tanh( acc ) : acc;
        }
    }
    const double e = h[L][0];
    if ( dEdG )
    {
        std::vector<double> d( 1, 1.0 ); // dE/dh[L]
        for ( int l = L - 1; l >= 0; --l )
        {
            const auto &W = E.W[l];
            const size_t ni = h[l].size();
            std::vector<double> dz( d.size() );
            for ( size_t o = 0; o < d.size(); ++o )
                dz[o] = m->act[l] == 't' ? d[o] * ( 1.0 - h[l + 1][o] * h[l + 1][o] ) : d[o];
            std::vector<double> dn( ni, 0.0 );
            for ( size_t o = 0; o < dz.size(); ++o )
                for ( size_t k = 0; k < ni; ++k )
                    dn[k] += W[o * ni + k] * dz[o];
            d.swap( dn );
        }
        dEdG->assign( ns, 0.0 );
        for ( int s = 0; s < ns; ++s )
        {
            double sc = 1.0;
            if ( m->scale )
            {
                const double w = E.gmax[s] - E.gmin[s];
                sc = w > 0.0 ? ( m->smax - m->smin ) / w : 0.0;
            }
            ( *dEdG )[s] = d[s] * sc;
        }
    }
    return e;
}
} // namespace

extern "C" {

orc_nnp *orc_nnp_create( int nelem, int nlayers, const int *nodes_hidden,
                         const char *activations )
{
    orc_nnp *m = new orc_nnp;
    m->nelem = nelem;
    m->nlayers = nlayers;
    m->nodes.assign( nlayers + 1, 0 );
    for ( int l = 1; l < nlayers; ++l )
        m->nodes[l] = nodes_hidden[l - 1];
    m->nodes[nlayers] = 1;
    m->act.assign( activations, activations + nlayers );
    m->el.resize( nelem );
    return m;
}

void orc_nnp_destroy( orc_nnp *m ) { delete m; }

void orc_nnp_set_scaling_range( orc_nnp *m, int scale, double smin, double smax )
{
    m->scale = scale;
    m->smin = smin;
    m->smax = smax;
}

int orc_nnp_add_sf( orc_nnp *m, int elem, int type, int e1, int e2, double eta,
                    double rs, double lambda, double zeta, double rc,
                    double gmin, double gmax )
{
    Element &E = m->el[elem];
    E.sf.push_back( SF{ type, e1, e2, eta, rs, lambda, zeta, rc } );
    E.gmin.push_back( gmin );
    E.gmax.push_back( gmax );
    return (int)E.sf.size() - 1;
}

/* W is [out][in] row-major, in = nsf of the element for layer 0. */
void orc_nnp_set_layer( orc_nnp *m, int elem, int layer, int n_out, int n_in,
                        const double *W, const double *b )
{
    Element &E = m->el[elem];
    if ( (int)E.W.size() < m->nlayers )
    {
        E.W.resize( m->nlayers );
        E.b.resize( m->nlayers );
    }
    E.W[layer].assign( W, W + (size_t)n_out * n_in );
    E.b[layer].assign( b, b + n_out );
}

void orc_nnp_set_atom_energy( orc_nnp *m, int elem, double e )
{
    m->el[elem].atom_energy = e;
}

int orc_nnp_nsf( const orc_nnp *m, int elem ) { return (int)m->el[elem].sf.size(); }

/* Symmetry functions of the local atoms: G[i * ld + s] (ld >= max nsf). */
void orc_nnp_symmetry_functions( const orc_nnp *m, int n_local, const double *x,
                                 const int *type, const int64_t *offsets,
                                 const int *idx, int ld, double *G )
{
    const double rcm = rc_max( m );
#pragma omp parallel
    {
        std::vector<Nb> nb;
        std::vector<double> g;
#pragma omp for schedule( dynamic, 16 )
        for ( int i = 0; i < n_local; ++i )
        {
            neighbours( m, i, x, offsets, idx, rcm, nb );
            symmetry_functions( m, type[i], nb, type, g, nullptr );
            for ( size_t s = 0; s < g.size(); ++s )
                G[(size_t)i * ld + s] = g[s];
        }
    }
}

/* Atomic energies E_i (network outputs), dE/dG (raw) and forces on every
 * atom in [0, n_total) (ghost rows receive the reactions of their images;
 * fold them with orc_ghost_fold).  f is accumulated (+=).  Returns sum E_i. */
double orc_nnp_compute( const orc_nnp *m, int n_local, int n_total,
                        const double *x, const int *type, const int64_t *offsets,
                        const int *idx, int ld, double *G, double *dEdG,
                        double *Ei, double *f )
{
    (void)n_total;
    const double rcm = rc_max( m );
    std::vector<double> e( n_local, 0.0 );
    // forces: per-atom contributions are gathered serially after a parallel
    // pass so that the sum order is fixed
    std::vector<std::vector<std::pair<int, std::array<double, 3>>>> contrib( n_local );
#pragma omp parallel
    {
        std::vector<Nb> nb;
        std::vector<double> g, dedg;
        std::vector<std::vector<double>> dg;
#pragma omp for schedule( dynamic, 16 )
        for ( int i = 0; i < n_local; ++i )
        {
            neighbours( m, i, x, offsets, idx, rcm, nb );
            symmetry_functions( m, type[i], nb, type, g, &dg );
            e[i] = network( m, type[i], g, &dedg );
            const int ns = (int)g.size();
            if ( G )
                for ( int s = 0; s < ns; ++s )
                    G[(size_t)i * ld + s] = g[s];
            if ( dEdG )
                for ( int s = 0; s < ns; ++s )
                    dEdG[(size_t)i * ld + s] = dedg[s];
            auto &cv = contrib[i];
            cv.resize( nb.size() + 1 );
            for ( size_t a = 0; a <= nb.size(); ++a )
            {
                cv[a].first = a == 0 ? i : nb[a - 1].j;
                for ( int c = 0; c < 3; ++c )
                {
                    double acc = 0.0;
                    for ( int s = 0; s < ns; ++s )
                        acc += dedg[s] * dg[s][3 * a + c];
                    cv[a].second[c] = -acc;
                }
            }
        }
    }
    double etot = 0.0;
    for ( int i = 0; i < n_local; ++i )
    {
        if ( Ei )
            Ei[i] = e[i];
        etot += e[i];
        for ( const auto &p : contrib[i] )
            for ( int c = 0; c < 3; ++c )
                f[3 * p.first + c] += p.second[c];
    }
    return etot;
}

/* ForceNNP::compute_energy (force_nnp_cabana_neigh_impl.h:101-130) for a
 * model without normalisation (convEnergy 1, meanEnergy 0): the sum of
 * atomic energies plus N_local times the offset of element 0, converted
 * with the reference's hard-coded Hartree-to-eV factor. */
double orc_nnp_reference_energy( const orc_nnp *m, int n_local, double sum_ei )
{
    return ( sum_ei + n_local * m->el[0].atom_energy ) * 27.211384021355236;
}

} /* extern "C" */
