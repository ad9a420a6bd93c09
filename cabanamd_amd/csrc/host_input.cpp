// host_input.cpp -- input deck, run parameters and the initial state.
//
// Restates InputFile<System> (inputFile.h:150-275, inputFile_impl.h:56-865)
// on the host: the LAMMPS-subset parser, the sc/fcc lattice generator and
// the LAMMPS "loop geom" velocity generator, so a deck produces the same
// atoms, ids, types and velocities as the reference.
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <limits>
#include <sstream>
#include <string>
#include <vector>
#include <map>
#include <unordered_map>
#include <cstdint>

#include "../../include/cabanamd_amd.h"

namespace cmd
{
void set_error( const char *fmt, ... );
}

namespace
{
// split on spaces, tabs, CR and LF, dropping empty words
// (inputFile_impl.h:56-69)
std::vector<std::string> split_words( const std::string &line )
{
    std::vector<std::string> w;
    std::string cur;
    for ( char c : line )
    {
        if ( c == ' ' || c == '\t' || c == '\r' || c == '\n' )
        {
            if ( !cur.empty() )
                w.push_back( cur );
            cur.clear();
        }
        else
            cur.push_back( c );
    }
    if ( !cur.empty() )
        w.push_back( cur );
    return w;
}

// ---- LAMMPS_RandomVelocityGeom (inputFile.h:66-148) ------------------------
// Park-Miller minimal standard generator (Schrage's method), Marsaglia polar
// gaussian, seeded by a Jenkins one-at-a-time hash of the base seed's bytes
// and the atom's 24 coordinate bytes (plain char: signed on x86), keeping 27
// bits, followed by 5 warm-up draws.
class VelocityRng
{
    int seed_ = 0;

  public:
    double uniform()
    {
        const int IA = 16807, IM = 2147483647, IQ = 127773, IR = 2836;
        int k = seed_ / IQ;
        seed_ = IA * ( seed_ - k * IQ ) - IR * k;
        if ( seed_ < 0 )
            seed_ += IM;
        return ( 1.0 / IM ) * seed_;
    }
    double gaussian()
    {
        double v1, v2, rsq;
        do
        {
            v1 = 2.0 * uniform() - 1.0;
            v2 = 2.0 * uniform() - 1.0;
            rsq = v1 * v1 + v2 * v2;
        } while ( ( rsq >= 1.0 ) || ( rsq == 0.0 ) );
        return v2 * std::sqrt( -2.0 * std::log( rsq ) / rsq );
    }
    void reset( int ibase, const double *coord )
    {
        unsigned int h = 0;
        auto mix = [&h]( const char *bytes, int nb )
        {
            for ( int i = 0; i < nb; ++i )
            {
                h += bytes[i];
                h += ( h << 10 );
                h ^= ( h >> 6 );
            }
        };
        mix( reinterpret_cast<const char *>( &ibase ), (int)sizeof( int ) );
        mix( reinterpret_cast<const char *>( coord ), 3 * (int)sizeof( double ) );
        h += ( h << 3 );
        h ^= ( h >> 11 );
        h += ( h << 15 );
        seed_ = (int)( h & 0x7ffffff );
        if ( !seed_ )
            seed_ = 1;
        for ( int i = 0; i < 5; ++i )
            uniform();
    }
};

// rand() of the C library as a fresh reference process sees it (never
// seeded, so seed 1).  glibc's generator is additive feedback of degree 31
// and separation 3 (TYPE_3 of random_r.c), restated from its published
// description: r_0 = seed; r_i = 16807 r_{i-1} mod (2^31 - 1) for
// 0 < i < 31; r_i = r_{i-31} for 31 <= i < 34; then r_i = r_{i-31} + r_{i-3}
// (mod 2^32) and the k-th call returns r_{k+344} >> 1.  The engine keeps its
// own copy, restarted for every lattice, so its types never depend on other
// users of rand() in the process.
class CRand
{
    uint32_t r_[34];
    int i_ = 0; // index of the next term, modulo 34

    uint32_t step()
    {
        const uint32_t v = r_[( i_ + 34 - 31 ) % 34] + r_[( i_ + 34 - 3 ) % 34];
        r_[i_] = v;
        i_ = ( i_ + 1 ) % 34;
        return v;
    }

  public:
    explicit CRand( uint32_t seed = 1 )
    {
        int32_t r = (int32_t)( seed == 0 ? 1 : seed );
        r_[0] = (uint32_t)r;
        for ( int i = 1; i < 31; ++i )
        {
            r = (int32_t)( ( 16807LL * r ) % 2147483647LL );
            if ( r < 0 )
                r += 2147483647;
            r_[i] = (uint32_t)r;
        }
        for ( int i = 31; i < 34; ++i )
            r_[i] = r_[i - 31];
        i_ = 0; // r_34 overwrites the slot of r_0
        for ( int i = 34; i < 344; ++i )
            step();
    }
    int next() { return (int)( step() >> 1 ); }
};

// The deck's named region blocks in the iteration order of the reference's
// std::unordered_map<std::string, Block> (inputFile.h:172-173), which
// decides both which region 'create_atoms T box' names (the map's first
// entry, inputFile_impl.h:287-292) and the order get_regions lists the
// regions of a site in (inputFile.h:203-216).  The same container type,
// filled with the same names in the same order, iterates in the same order.
std::vector<int> region_order( const cmd_params *p )
{
    std::unordered_map<std::string, int> m;
    for ( int k = 0; k < p->nregions; ++k )
        m[std::string( p->region_name[k] )] = k;
    std::vector<int> order;
    for ( const auto &kv : m )
        order.push_back( kv.second );
    return order;
}

// Lattice loop bounds and the half-open site test (inputFile_impl.h:562-588,
// 598-606, 692-699).
struct SiteWalk
{
    int lo[3], hi[3];
    double maxv[3];
};

SiteWalk site_walk( const cmd_params *p, const cmd_domain *d )
{
    SiteWalk w;
    const double a = p->lattice_constant;
    for ( int k = 0; k < 3; ++k )
    {
        w.maxv[k] = a * p->region_hi[k];
        w.lo[k] = (int)( d->local_lo[k] / a - 0.5 );
        w.hi[k] = (int)std::max( std::min( w.maxv[k] / a, d->local_hi[k] / a + 0.5 ),
                                 static_cast<double>( w.lo[k] ) );
        if ( w.lo[k] == w.hi[k] )
            w.hi[k] -= 1;
    }
    return w;
}

inline bool site_inside( const cmd_domain *d, const SiteWalk &w, double x,
                         double y, double z )
{
    return ( x >= d->local_lo[0] ) && ( y >= d->local_lo[1] ) &&
           ( z >= d->local_lo[2] ) && ( x < d->local_hi[0] ) &&
           ( y < d->local_hi[1] ) && ( z < d->local_hi[2] ) && ( x < w.maxv[0] ) &&
           ( y < w.maxv[1] ) && ( z < w.maxv[2] );
}

// Every lattice site inside the box, in the reference's loop order, passed
// to visit(x, y, z, sc) (sc: the simple cubic lattice, whose sites take a
// random type instead of a region's).
template <class Visit>
void walk_sites( const cmd_params *p, const cmd_domain *d, Visit visit )
{
    const double a = p->lattice_constant;
    const SiteWalk w = site_walk( p, d );
    if ( p->lattice_style == 0 )
    {
        // simple cubic (inputFile_impl.h:590-655)
        for ( int iz = w.lo[2]; iz <= w.hi[2]; ++iz )
        {
            const double z = a * ( iz + p->lattice_offset[2] );
            for ( int iy = w.lo[1]; iy <= w.hi[1]; ++iy )
            {
                const double y = a * ( iy + p->lattice_offset[1] );
                for ( int ix = w.lo[0]; ix <= w.hi[0]; ++ix )
                {
                    const double x = a * ( ix + p->lattice_offset[0] );
                    if ( site_inside( d, w, x, y, z ) )
                        visit( x, y, z, true );
                }
            }
        }
        return;
    }
    // face-centred cubic, 4-site basis (inputFile_impl.h:658-792)
    const double base[4][3] = { { 0.0, 0.0, 0.0 },
                                { 0.5, 0.5, 0.0 },
                                { 0.5, 0.0, 0.5 },
                                { 0.0, 0.5, 0.5 } };
    double basis[4][3];
    for ( int b = 0; b < 4; ++b )
        for ( int k = 0; k < 3; ++k )
            basis[b][k] = base[b][k] + p->lattice_offset[k];
    for ( int iz = w.lo[2]; iz <= w.hi[2]; ++iz )
        for ( int iy = w.lo[1]; iy <= w.hi[1]; ++iy )
            for ( int ix = w.lo[0]; ix <= w.hi[0]; ++ix )
                for ( int b = 0; b < 4; ++b )
                {
                    const double x = a * ( 1.0 * ix + basis[b][0] );
                    const double y = a * ( 1.0 * iy + basis[b][1] );
                    const double z = a * ( 1.0 * iz + basis[b][2] );
                    if ( site_inside( d, w, x, y, z ) )
                        visit( x, y, z, false );
                }
}

// Atom creation over the sites (inputFile_impl.h:658-792): an fcc site
// becomes an atom when it lies in a region that create_atoms gave a type;
// of several such regions, rand() picks one in the map's iteration order
// (one draw per atom).  sc sites are all created with type rand() % ntypes.
// emit(k, x, y, z, type0) gets the atom's index and 0-based type.
template <class Emit>
int create_sites( const cmd_params *p, const cmd_domain *d, bool draw, Emit emit )
{
    const double a = p->lattice_constant;
    const std::vector<int> order = region_order( p );
    CRand crand;
    std::vector<int> hits;
    int n = 0;
    walk_sites( p, d,
                [&]( double x, double y, double z, bool sc )
                {
                    if ( sc )
                    {
                        const int t = draw ? crand.next() % p->ntypes : 0;
                        emit( n++, x, y, z, t );
                        return;
                    }
                    int type1 = 1; // the single block of type 1
                    if ( p->nregions > 0 )
                    {
                        hits.clear();
                        for ( int k : order )
                        {
                            const double *b = p->region_block[k];
                            if ( p->region_type[k] > 0 && x >= a * b[0] && y >= a * b[2] &&
                                 z >= a * b[4] && x < a * b[1] && y < a * b[3] && z < a * b[5] )
                                hits.push_back( k );
                        }
                        if ( hits.empty() )
                            return;
                        if ( draw )
                            type1 = p->region_type[hits[crand.next() % (int)hits.size()]];
                    }
                    else if ( draw )
                        crand.next();
                    emit( n++, x, y, z, type1 - 1 );
                } );
    return n;
}

cmd_status deck_error( const std::string &msg )
{
    cmd::set_error( "%s", msg.c_str() );
    return CMD_ERR_IO;
}

} // namespace

extern "C" {

// Defaults of InputFile (inputFile.h:150-275, inputFile_impl.h:70-84) with
// units lj applied.
cmd_status cmd_params_default( cmd_params *p )
{
    if ( !p )
    {
        cmd::set_error( "params is NULL" );
        return CMD_ERR_ARG;
    }
    std::memset( p, 0, sizeof( *p ) );
    p->units = 1;
    p->lattice_style = 1;
    p->lattice_constant = 0.8442;
    p->ntypes = 1;
    for ( int k = 0; k < CMD_MAX_TYPES; ++k )
        p->mass[k] = 1.0;
    p->force_cutoff = 2.5;
    p->skin = 0.0;
    p->every = 20;
    p->max_neigh_guess = 50;
    p->comm_ghost_cutoff = std::pow( ( 4.0 / 0.8442 ), ( 1.0 / 3.0 ) ) * 20.0;
    p->half = 0;
    p->nsteps = 100;
    p->thermo_rate = 10;
    p->dt = 0.005;
    p->mvv2e = 1.0;
    p->boltz = 1.0;
    p->force_type = 0;
    p->vacuum_rate = 1.0;
    return CMD_OK;
}

// check_lammps_command (inputFile_impl.h:125-534) over every line of a deck.
// force_iteration: -1 keeps the deck's newton setting, 0 forces a full list,
// 1 a half list (the --force-iteration override, inputFile_impl.h:489-515).
cmd_status cmd_params_parse_deck( const char *path, int force_iteration,
                                  cmd_params *p )
{
    if ( !path || !p )
    {
        cmd::set_error( "NULL argument" );
        return CMD_ERR_ARG;
    }
    std::ifstream in( path );
    if ( !in )
        return deck_error( std::string( "cannot open input deck " ) + path );
    cmd_params_default( p );
    bool timestepflag = false;
    bool have_region = false;
    // the reference's regions map (inputFile.h:172): id -> slot of p
    std::unordered_map<std::string, int> regions;
    std::string line;
    int lineno = 0;
    while ( std::getline( in, line ) )
    {
        ++lineno;
        auto w = split_words( line );
        if ( w.empty() || w[0][0] == '#' )
            continue;
        const std::string &key = w[0];
        auto need = [&]( size_t n ) -> bool { return w.size() >= n; };
        auto bad = [&]( const std::string &msg ) -> cmd_status
        {
            std::ostringstream os;
            os << "LAMMPS-Command: " << msg << " (line " << lineno << ": " << line
               << ")";
            return deck_error( os.str() );
        };
        try
        {
            if ( key == "variable" )
                return bad( "'variable' keyword is not supported in CabanaMD" );
            else if ( key == "units" )
            {
                if ( !need( 2 ) )
                    return bad( "'units' needs a style" );
                if ( w[1] == "metal" )
                {
                    p->units = 2;
                    p->boltz = 8.617343e-5;
                    p->mvv2e = 1.0364269e-4;
                    p->dt = 0.001;
                }
                else if ( w[1] == "real" )
                {
                    p->units = 0;
                    p->boltz = 0.0019872067;
                    p->mvv2e = 48.88821291 * 48.88821291;
                    if ( !timestepflag )
                        p->dt = 1.0;
                }
                else if ( w[1] == "lj" )
                {
                    p->units = 1;
                    p->boltz = 1.0;
                    p->mvv2e = 1.0;
                    if ( !timestepflag )
                        p->dt = 0.005;
                }
                else
                    return bad( "'units' command only supports 'metal', 'real', "
                                "and 'lj' in CabanaMD" );
            }
            else if ( key == "atom_style" )
            {
                if ( !need( 2 ) || ( w[1] != "atomic" && w[1] != "charge" ) )
                    return bad( "'atom_style' command only supports 'atomic' and "
                                "'charge' in CabanaMD" );
                p->atom_style_charge = w[1] == "charge";
            }
            else if ( key == "lattice" )
            {
                if ( !need( 3 ) )
                    return bad( "'lattice' needs a style and a value" );
                if ( w[1] == "sc" )
                {
                    p->lattice_style = 0;
                    p->lattice_constant = std::stod( w[2] );
                }
                else if ( w[1] == "fcc" )
                {
                    p->lattice_style = 1;
                    p->lattice_constant =
                        ( p->units == 1 )
                            ? std::pow( ( 4.0 / std::stod( w[2] ) ), ( 1.0 / 3.0 ) )
                            : std::stod( w[2] );
                }
                else
                    return bad( "'lattice' command only supports 'sc' and 'fcc' "
                                "in CabanaMD" );
                if ( w.size() > 3 )
                {
                    if ( w[3] == "origin" && need( 7 ) )
                    {
                        p->lattice_offset[0] = std::stod( w[4] );
                        p->lattice_offset[1] = std::stod( w[5] );
                        p->lattice_offset[2] = std::stod( w[6] );
                    }
                    else
                        return bad( "'lattice' command only supports 'origin' "
                                    "additional option in CabanaMD" );
                }
            }
            else if ( key == "region" )
            {
                // inputFile_impl.h:241-266: regions[id] = block (a second
                // definition of an id replaces its block), and the bounds
                // of all blocks
                if ( !need( 9 ) || w[2] != "block" )
                    return bad( "'region' command only supports 'block' option "
                                "in CabanaMD" );
                if ( w[1].size() >= sizeof( p->region_name[0] ) )
                    return bad( "region id too long" );
                auto it = regions.find( w[1] );
                int k = 0;
                if ( it == regions.end() )
                {
                    if ( p->nregions >= CMD_MAX_REGIONS )
                        return bad( "too many regions" );
                    k = p->nregions++;
                    regions[w[1]] = k;
                    std::snprintf( p->region_name[k], sizeof( p->region_name[k] ), "%s",
                                   w[1].c_str() );
                }
                else
                    k = it->second;
                for ( int c = 0; c < 6; ++c )
                    p->region_block[k][c] = std::stod( w[3 + c] );
                for ( int c = 0; c < 3; ++c )
                {
                    const double lo = p->region_block[k][2 * c], hi = p->region_block[k][2 * c + 1];
                    p->region_lo[c] = have_region ? std::min( p->region_lo[c], lo ) : lo;
                    p->region_hi[c] = have_region ? std::max( p->region_hi[c], hi ) : hi;
                }
                have_region = true;
            }
            else if ( key == "create_box" )
            {
                if ( !need( 2 ) )
                    return bad( "'create_box' needs a type count" );
                p->ntypes = std::stoi( w[1] );
                if ( p->ntypes < 1 || p->ntypes > CMD_MAX_TYPES )
                    return bad( "create_box type count outside [1,8]" );
            }
            else if ( key == "create_atoms" )
            {
                // inputFile_impl.h:271-297: regions_to_type[id] = type; 'box'
                // names the first region of the map
                if ( !need( 3 ) || ( w[2] != "box" && w[2] != "region" ) )
                    return bad( "'create_atoms' command only supports 'region' "
                                "option in CabanaMD" );
                const int t = std::stoi( w[1] );
                if ( t < 1 || t > CMD_MAX_TYPES )
                    return bad( "create_atoms type outside [1,8]" );
                int k = -1;
                if ( w[2] == "region" )
                {
                    if ( !need( 4 ) || regions.find( w[3] ) == regions.end() )
                        return bad( "region '" + ( need( 4 ) ? w[3] : std::string() ) +
                                    "' is not defined" );
                    k = regions[w[3]];
                }
                else
                {
                    if ( regions.empty() )
                        return bad( "create_atoms box before any region" );
                    k = regions.begin()->second;
                }
                p->region_type[k] = t;
            }
            else if ( key == "mass" )
            {
                if ( !need( 3 ) )
                    return bad( "'mass' needs a type and a value" );
                const int t = std::stoi( w[1] ) - 1;
                if ( t < 0 || t >= CMD_MAX_TYPES )
                    return bad( "mass type outside [1,8]" );
                p->mass[t] = std::stod( w[2] );
            }
            else if ( key == "read_data" || key == "write_data" )
            {
                // inputFile_impl.h:315-326
                if ( !need( 2 ) )
                    return bad( "'" + key + "' needs a file name" );
                char *dst = key == "read_data" ? p->data_in : p->data_out;
                if ( w[1].size() >= sizeof( p->data_in ) )
                    return bad( "data file name too long" );
                std::snprintf( dst, sizeof( p->data_in ), "%s", w[1].c_str() );
            }
            else if ( key == "dump" )
            {
                // VTK output is outside this engine; the command is accepted
                // with the reference's argument checks and otherwise ignored.
                if ( !need( 6 ) || w[3] != "vtk" )
                    return bad( "'dump' command only supports 'vtk' in CabanaMD" );
                if ( w[2] != "all" )
                    return bad( "'dump' command only supports dumping 'all' types "
                                "in CabanaMD" );
            }
            else if ( key == "pair_style" )
            {
                if ( !need( 2 ) )
                    return bad( "'pair_style' needs a style" );
                if ( w[1] == "lj/cut" && need( 3 ) )
                {
                    p->force_type = 0;
                    p->force_cutoff = std::stod( w[2] );
                }
                else if ( w[1] == "nnp" )
                {
                    p->force_type = 2;
                    if ( need( 4 ) )
                        std::snprintf( p->nnp_dir, sizeof( p->nnp_dir ), "%s",
                                       w[3].c_str() );
                }
                else
                    return bad( "'pair_style' command only supports 'lj/cut' and "
                                "'nnp' style in CabanaMD" );
            }
            else if ( key == "pair_coeff" )
            {
                if ( p->force_type == 2 )
                {
                    if ( !need( 4 ) )
                        return bad( "'pair_coeff' for nnp needs a cutoff" );
                    p->force_cutoff = std::stod( w[3] );
                }
                else
                {
                    if ( !need( 6 ) )
                        return bad( "'pair_coeff' needs i j eps sigma cut" );
                    if ( p->ncoeff >= 16 )
                        return bad( "too many pair_coeff lines" );
                    const int c = p->ncoeff++;
                    p->coeff_i[c] = std::stoi( w[1] );
                    p->coeff_j[c] = std::stoi( w[2] );
                    p->coeff_eps[c] = std::stod( w[3] );
                    p->coeff_sigma[c] = std::stod( w[4] );
                    p->coeff_cut[c] = std::stod( w[5] );
                }
            }
            else if ( key == "velocity" )
            {
                if ( !need( 5 ) || w[2] != "create" )
                    return bad( "'velocity' command can only be used with option "
                                "'create' in CabanaMD" );
                // type_to_temperature[type] (inputFile_impl.h:396-410);
                // 'all' is keyed as type 1
                const int t = ( w[1] == "all" ) ? 1 : std::stoi( w[1] );
                if ( t < 1 || t > CMD_MAX_TYPES )
                    return bad( "velocity type outside [1,8]" );
                if ( t == 1 )
                {
                    p->temp_target = std::stod( w[3] );
                    p->temp_seed = std::stoi( w[4] );
                }
                else
                {
                    p->type_temp[t - 1] = std::stod( w[3] );
                    p->type_seed[t - 1] = std::stoi( w[4] );
                }
                for ( size_t k = 5; k + 1 < w.size(); ++k )
                    if ( w[k] == "dist" && w[k + 1] == "gaussian" )
                        p->vel_gaussian = 1;
            }
            else if ( key == "neighbor" )
            {
                if ( !need( 2 ) )
                    return bad( "'neighbor' needs a skin" );
                p->skin = std::stod( w[1] );
            }
            else if ( key == "neigh_modify" )
            {
                size_t i = 1;
                while ( i < w.size() )
                {
                    if ( w[i] == "every" && i + 1 < w.size() )
                    {
                        p->every = std::stoi( w[i + 1] );
                        i += 2;
                    }
                    else if ( w[i] == "one" && i + 1 < w.size() )
                    {
                        p->max_neigh_guess = std::stoi( w[i + 1] );
                        i += 2;
                    }
                    else
                        return bad( "'neigh_modify' only supports 'every' and "
                                    "'one' in CabanaMD" );
                }
            }
            else if ( key == "comm_modify" )
            {
                if ( !need( 4 ) || w[1] != "cutoff" || w[2] != "*" )
                    return bad( "'comm_modify' command only supports single "
                                "cutoff 'cutoff' for all atom types '*' in "
                                "CabanaMD" );
                p->comm_ghost_cutoff = std::stod( w[3] );
            }
            else if ( key == "fix" )
            {
                if ( !need( 4 ) || w[3] != "nve" )
                    return bad( "'fix' command only supports 'nve' style in "
                                "CabanaMD" );
            }
            else if ( key == "run" )
            {
                if ( !need( 2 ) )
                    return bad( "'run' needs a step count" );
                p->nsteps = std::stoi( w[1] );
            }
            else if ( key == "thermo" )
            {
                if ( !need( 2 ) )
                    return bad( "'thermo' needs a rate" );
                p->thermo_rate = std::stoi( w[1] );
            }
            else if ( key == "timestep" )
            {
                if ( !need( 2 ) )
                    return bad( "'timestep' needs a value" );
                p->dt = std::stod( w[1] );
                timestepflag = true;
            }
            else if ( key == "newton" )
            {
                if ( !need( 2 ) || ( w[1] != "on" && w[1] != "off" ) )
                    return bad( "'newton' must be followed by 'on' or 'off'" );
                if ( force_iteration < 0 )
                    p->half = ( w[1] == "on" ) ? 1 : 0;
            }
            else if ( key == "group" )
            {
                if ( !need( 4 ) || w[2] != "region" )
                    return bad( "'group' command can only be used with 'region' "
                                "in CabanaMD" );
            }
            else
                return bad( "Unknown input file keyword" );
        }
        catch ( const std::exception & )
        {
            return bad( "malformed number" );
        }
    }
    if ( force_iteration >= 0 )
        p->half = force_iteration ? 1 : 0;
    // atoms come from a region lattice or from read_data
    if ( !have_region && !p->data_in[0] )
        return deck_error( "the deck defines no region block" );
    return CMD_OK;
}

// create_lattice (inputFile_impl.h:536-865): sites, ids and types (by
// region), then velocities (loop-geom uniform or gaussian, seeded per
// type), zero net momentum and a rescale of each atom to its type's target
// temperature.
cmd_status cmd_lattice_create( const cmd_params *p, const cmd_domain *d, int *n,
                               double *x, double *v, int *type, int *id )
{
    if ( !p || !d || !n )
    {
        cmd::set_error( "NULL argument" );
        return CMD_ERR_ARG;
    }
    if ( p->lattice_constant <= 0 )
    {
        cmd::set_error( "lattice constant must be positive" );
        return CMD_ERR_ARG;
    }
    const int count =
        create_sites( p, d, false, []( int, double, double, double, int ) {} );
    *n = count;
    if ( !x )
        return CMD_OK;
    std::vector<int> types( count, 0 );
    create_sites( p, d, true,
                  [&]( int k, double xs, double ys, double zs, int t )
                  {
                      x[3 * k + 0] = xs;
                      x[3 * k + 1] = ys;
                      x[3 * k + 2] = zs;
                      types[k] = t;
                      if ( type )
                          type[k] = t;
                      if ( id )
                          id[k] = k + 1;
                  } );
    for ( int t : types )
        if ( t < 0 || t >= p->ntypes )
        {
            cmd::set_error( "create_atoms type %d outside create_box's %d types", t + 1,
                            p->ntypes );
            return CMD_ERR_ARG;
        }
    if ( !v )
        return CMD_OK;
    // type_to_temperature: type 1 from temp_target/temp_seed, others per type
    auto vel_seed = [&]( int t ) { return t == 0 ? p->temp_seed : p->type_seed[t]; };
    auto vel_temp = [&]( int t ) { return t == 0 ? p->temp_target : p->type_temp[t]; };
    double total_mass = 0.0, px = 0.0, py = 0.0, pz = 0.0;
    VelocityRng rng;
    for ( int i = 0; i < count; ++i )
    {
        const int t = types[i];
        rng.reset( vel_seed( t ), &x[3 * i] );
        const double m = p->mass[t];
        double vx, vy, vz;
        if ( p->vel_gaussian )
        {
            vx = rng.gaussian();
            vy = rng.gaussian();
            vz = rng.gaussian();
        }
        else
        {
            vx = rng.uniform() - 0.5;
            vy = rng.uniform() - 0.5;
            vz = rng.uniform() - 0.5;
        }
        v[3 * i + 0] = vx / std::sqrt( m );
        v[3 * i + 1] = vy / std::sqrt( m );
        v[3 * i + 2] = vz / std::sqrt( m );
        total_mass += m;
        px += m * v[3 * i + 0];
        py += m * v[3 * i + 1];
        pz += m * v[3 * i + 2];
    }
    const double sx = px / total_mass, sy = py / total_mass,
                 sz = pz / total_mass;
    double mv2 = 0.0;
    for ( int i = 0; i < count; ++i )
    {
        v[3 * i + 0] -= sx;
        v[3 * i + 1] -= sy;
        v[3 * i + 2] -= sz;
    }
    // sum m v^2 with Neumaier compensation (the reference's device reduction
    // has no fixed order; the compensated sum is within an ulp of exact)
    double comp = 0.0;
    for ( int i = 0; i < count; ++i )
    {
        const int t = types[i];
        const double term = ( v[3 * i] * v[3 * i] + v[3 * i + 1] * v[3 * i + 1] +
                              v[3 * i + 2] * v[3 * i + 2] ) *
                            p->mass[t];
        const double sum = mv2 + term;
        if ( std::fabs( mv2 ) >= std::fabs( term ) )
            comp += ( mv2 - sum ) + term;
        else
            comp += ( term - sum ) + mv2;
        mv2 = sum;
    }
    mv2 = mv2 + comp;
    const int dof = 3 * count - 3;
    const double T = mv2 * ( p->mvv2e / ( 1.0 * dof * p->boltz ) );
    for ( int i = 0; i < count; ++i )
    {
        const double s = std::sqrt( vel_temp( types[i] ) / T );
        v[3 * i + 0] *= s;
        v[3 * i + 1] *= s;
        v[3 * i + 2] *= s;
    }
    return CMD_OK;
}

} // extern "C"

// ---- LAMMPS data files (read_data.h:100-422, restated)
namespace
{
// Next line that is not blank after removing a # comment; false at EOF.
bool data_line( std::ifstream &in, std::string &line )
{
    while ( std::getline( in, line ) )
    {
        line = line.substr( 0, line.find( '#' ) );
        if ( line.find_first_not_of( " \r\t\n" ) != std::string::npos )
            return true;
    }
    return false;
}

std::string trimmed( const std::string &s )
{
    const size_t a = s.find_first_not_of( " \r\t\n" );
    if ( a == std::string::npos )
        return "";
    const size_t b = s.find_last_not_of( " \r\t\n" );
    return s.substr( a, b - a + 1 );
}

struct DataFile
{
    int natoms = 0, ntypes = 0;
    double lo[3] = { 0, 0, 0 }, hi[3] = { 0, 0, 0 };
    std::vector<double> x, v, q, mass;
    std::vector<int> type, id;
    bool has_v = false;
};

cmd_status parse_data( const char *path, int charge, DataFile &D )
{
    std::ifstream in( path );
    if ( !in.good() )
        return deck_error( std::string( "could not read data file " ) + path );
    std::string line;
    if ( !std::getline( in, line ) ) // the first line is a title
        return deck_error( "empty data file" );
    // header: until "zlo zhi" (read_lammps_header)
    bool box[3] = { false, false, false };
    while ( !box[2] )
    {
        if ( !data_line( in, line ) )
            return deck_error( "data file header ends before 'zlo zhi'" );
        const char *c = line.c_str();
        if ( line.find( "atom types" ) != std::string::npos )
            std::sscanf( c, "%d", &D.ntypes );
        else if ( line.find( "atoms" ) != std::string::npos )
            std::sscanf( c, "%d", &D.natoms );
        else if ( line.find( "xlo xhi" ) != std::string::npos )
            box[0] = std::sscanf( c, "%lg %lg", &D.lo[0], &D.hi[0] ) == 2;
        else if ( line.find( "ylo yhi" ) != std::string::npos )
            box[1] = std::sscanf( c, "%lg %lg", &D.lo[1], &D.hi[1] ) == 2;
        else if ( line.find( "zlo zhi" ) != std::string::npos )
            box[2] = std::sscanf( c, "%lg %lg", &D.lo[2], &D.hi[2] ) == 2;
    }
    if ( !box[0] || !box[1] || D.natoms < 0 || D.ntypes < 1 || D.ntypes > CMD_MAX_TYPES )
        return deck_error( "data file header needs atoms, atom types (1..8) and the box" );
    D.mass.assign( D.ntypes, std::numeric_limits<double>::quiet_NaN() );
    bool have_atoms = false;
    while ( data_line( in, line ) )
    {
        const std::string key = trimmed( line );
        if ( key == "Atoms" )
        {
            for ( int k = 0; k < D.natoms; ++k )
            {
                if ( !data_line( in, line ) )
                    return deck_error( "data file ends inside Atoms" );
                int idv = 0, t = 0;
                double a = 0, b = 0, c = 0, qq = 0;
                const int got = charge ? std::sscanf( line.c_str(), "%d %d %lg %lg %lg %lg",
                                                      &idv, &t, &qq, &a, &b, &c )
                                       : std::sscanf( line.c_str(), "%d %d %lg %lg %lg", &idv,
                                                      &t, &a, &b, &c );
                if ( got != ( charge ? 6 : 5 ) )
                    return deck_error( "malformed Atoms line: " + line );
                if ( t < 1 || t > D.ntypes )
                    return deck_error( "atom type out of range: " + line );
                // kept when inside [lo, hi) (one rank owns the whole box)
                if ( a >= D.lo[0] && b >= D.lo[1] && c >= D.lo[2] && a < D.hi[0] &&
                     b < D.hi[1] && c < D.hi[2] )
                {
                    D.id.push_back( idv );
                    D.type.push_back( t - 1 );
                    D.x.push_back( a );
                    D.x.push_back( b );
                    D.x.push_back( c );
                    D.q.push_back( qq );
                }
            }
            D.v.assign( D.x.size(), 0.0 );
            have_atoms = true;
        }
        else if ( key == "Velocities" )
        {
            if ( !have_atoms )
                return deck_error( "Must read Atoms before Velocities" );
            std::map<int, size_t> where;
            for ( size_t k = 0; k < D.id.size(); ++k )
                where[D.id[k]] = k;
            for ( int k = 0; k < D.natoms; ++k )
            {
                if ( !data_line( in, line ) )
                    return deck_error( "data file ends inside Velocities" );
                int idv = 0;
                double a = 0, b = 0, c = 0;
                if ( std::sscanf( line.c_str(), "%d %lg %lg %lg", &idv, &a, &b, &c ) != 4 )
                    return deck_error( "malformed Velocities line: " + line );
                const auto it = where.find( idv );
                if ( it != where.end() )
                {
                    D.v[3 * it->second] = a;
                    D.v[3 * it->second + 1] = b;
                    D.v[3 * it->second + 2] = c;
                }
            }
            D.has_v = true;
        }
        else if ( key == "Masses" )
        {
            for ( int k = 0; k < D.ntypes; ++k )
            {
                if ( !data_line( in, line ) )
                    return deck_error( "data file ends inside Masses" );
                int t = 0;
                double m = 0;
                if ( std::sscanf( line.c_str(), "%d %lg", &t, &m ) != 2 || t < 1 ||
                     t > D.ntypes )
                    return deck_error( "malformed Masses line: " + line );
                D.mass[t - 1] = m;
            }
        }
        else if ( key == "Pair Coeffs" )
        {
            // ignored: pair_coeff comes from the input deck
            for ( int k = 0; k < D.ntypes; ++k )
                if ( !data_line( in, line ) )
                    return deck_error( "data file ends inside Pair Coeffs" );
        }
        else
            return deck_error( "Unknown identifier in data file: " + key );
    }
    if ( !have_atoms )
        return deck_error( "data file has no Atoms section" );
    if ( (int)D.id.size() != D.natoms )
        return deck_error( "Created incorrect # of atoms." );
    return CMD_OK;
}
} // namespace

extern "C" {

cmd_status cmd_data_read( const char *path, int charge_style, int *n, int *ntypes,
                          double lo[3], double hi[3], double *x, double *v, double *q,
                          int *type, int *id, double *mass )
{
    if ( !path || !n )
    {
        cmd::set_error( "path and n are required" );
        return CMD_ERR_ARG;
    }
    DataFile D;
    const cmd_status st = parse_data( path, charge_style, D );
    if ( st != CMD_OK )
        return st;
    *n = (int)D.id.size();
    if ( ntypes )
        *ntypes = D.ntypes;
    for ( int k = 0; k < 3; ++k )
    {
        if ( lo )
            lo[k] = D.lo[k];
        if ( hi )
            hi[k] = D.hi[k];
    }
    if ( !x )
        return CMD_OK;
    std::copy( D.x.begin(), D.x.end(), x );
    if ( v )
        std::copy( D.v.begin(), D.v.end(), v );
    if ( q )
        std::copy( D.q.begin(), D.q.end(), q );
    if ( type )
        std::copy( D.type.begin(), D.type.end(), type );
    if ( id )
        std::copy( D.id.begin(), D.id.end(), id );
    if ( mass )
        for ( int t = 0; t < D.ntypes; ++t )
            if ( D.mass[t] == D.mass[t] )
                mass[t] = D.mass[t];
    return CMD_OK;
}

cmd_status cmd_data_write( const char *path, int n, int ntypes, const double lo[3],
                           const double hi[3], const double *x, const double *v,
                           const int *type, const int *id )
{
    if ( !path || n < 0 || !lo || !hi || ( n > 0 && ( !x || !type || !id ) ) )
    {
        cmd::set_error( "path, box, and x, type, id for n > 0 are required" );
        return CMD_ERR_ARG;
    }
    std::FILE *f = std::fopen( path, "w" );
    if ( !f )
    {
        cmd::set_error( "cannot write %s", path );
        return CMD_ERR_IO;
    }
    std::fprintf( f, "LAMMPS data file from CabanaMD\n\n%d atoms\n%d atom types\n\n", n,
                  ntypes );
    const char *ax[3] = { "x", "y", "z" };
    for ( int k = 0; k < 3; ++k )
        std::fprintf( f, "%.17g %.17g %slo %shi\n", lo[k], hi[k], ax[k], ax[k] );
    std::fprintf( f, "\nAtoms # atomic\n\n" );
    // positions wrapped into [lo, hi) so that read_data keeps every atom
    // (between rebuilds atoms may sit just outside the box)
    auto wrap = [&]( double c, int k )
    {
        const double L = hi[k] - lo[k];
        double w = c - L * std::floor( ( c - lo[k] ) / L );
        if ( !( w < hi[k] ) || w < lo[k] )
            w = lo[k];
        return w;
    };
    for ( int i = 0; i < n; ++i )
        std::fprintf( f, "%d %d %.17g %.17g %.17g\n", id[i], type[i] + 1,
                      wrap( x[3 * i], 0 ), wrap( x[3 * i + 1], 1 ), wrap( x[3 * i + 2], 2 ) );
    std::fprintf( f, "\nVelocities\n\n" );
    for ( int i = 0; i < n; ++i )
        std::fprintf( f, "%d %.17g %.17g %.17g\n", id[i], v ? v[3 * i] : 0.0,
                      v ? v[3 * i + 1] : 0.0, v ? v[3 * i + 2] : 0.0 );
    const bool ok = std::fclose( f ) == 0;
    if ( !ok )
    {
        cmd::set_error( "write to %s failed", path );
        return CMD_ERR_IO;
    }
    return CMD_OK;
}

} // extern "C"
