/*
 * cabanamd_amd.hpp -- C++ host classes over the C ABI (cabanamd_amd.h), with
 * the names, arguments and error behaviour of the CabanaMD modules they
 * replace, so that code written against the reference's classes reads the
 * same.  Header only; link with -lcabanamd_amd.
 *
 *   System          system.h:65-279 (device SoA instead of AoSoA slices)
 *   Comm            comm_mpi.h:60-358 (one rank: wrap, ghosts, fold)
 *   Binning         binning_cabana.h:54-68
 *   NeighborVerlet  neighbor_verlet.h:19-70 (full or half)
 *   ForceLJ         force_lj_cabana_neigh.h / _impl.h:49-377
 *   Integrator      integrator_nve.h:56-110
 *   Temperature, KinE, PotE   property_*.h
 *   InputFile       inputFile.h:150-275 (deck subset)
 *   CbnMD           cabanamd.h:65-104 (init / run / thermo rows)
 *
 * Errors: the reference's log_err prints and throws std::runtime_error
 * (output.h:37-54); every method here throws cabanamd_amd::Error (a
 * std::runtime_error) carrying the status and cmd_last_error()'s message.
 * Device state lives behind the handles; there are no slices to re-fetch
 * after resize or permute.
 */
#ifndef CABANAMD_AMD_HPP
#define CABANAMD_AMD_HPP

#include "cabanamd_amd.h"

#include <array>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace cabanamd_amd
{

struct Error : std::runtime_error
{
    cmd_status status;
    Error( cmd_status s, const std::string &msg )
        : std::runtime_error( msg )
        , status( s )
    {
    }
};

inline void check( cmd_status s )
{
    if ( s != CMD_OK )
    {
        const char *m = cmd_last_error();
        throw Error( s, m ? m : "cabanamd_amd error" );
    }
}

// Movable owner of one C handle.
template <class H, cmd_status ( *Destroy )( H * )>
class Handle
{
  public:
    Handle() = default;
    explicit Handle( H *h )
        : h_( h )
    {
    }
    Handle( const Handle & ) = delete;
    Handle &operator=( const Handle & ) = delete;
    Handle( Handle &&o ) noexcept
        : h_( std::exchange( o.h_, nullptr ) )
    {
    }
    Handle &operator=( Handle &&o ) noexcept
    {
        if ( this != &o )
        {
            reset();
            h_ = std::exchange( o.h_, nullptr );
        }
        return *this;
    }
    ~Handle() { reset(); }
    void reset()
    {
        if ( h_ )
            (void)Destroy( h_ );
        h_ = nullptr;
    }
    H *get() const { return h_; }

  private:
    H *h_ = nullptr;
};

class System
{
  public:
    explicit System( int ntypes = 1 )
    {
        cmd_system *s = nullptr;
        check( cmd_system_create( ntypes, &s ) );
        h_ = Handle<cmd_system, cmd_system_destroy>( s );
        this->ntypes = ntypes;
    }

    int ntypes = 1;
    double dt = 0.005, mvv2e = 1.0, boltz = 1.0;

    // System::create_domain (system.h:149-205)
    cmd_domain create_domain( std::array<double, 3> low,
                              std::array<double, 3> high, double ghost_cutoff )
    {
        cmd_domain d;
        check( cmd_domain_create( low.data(), high.data(), ghost_cutoff, &d ) );
        check( cmd_system_set_domain( h_.get(), &d ) );
        return d;
    }
    void set_units( double dt_, double mvv2e_, double boltz_ )
    {
        dt = dt_;
        mvv2e = mvv2e_;
        boltz = boltz_;
        check( cmd_system_set_units( h_.get(), dt, mvv2e, boltz ) );
    }
    void set_mass( const std::vector<double> &mass )
    {
        check( cmd_system_set_mass( h_.get(), (int)mass.size(), mass.data() ) );
    }
    // System::resize (system_1aosoa.h:63-70)
    void resize( int n ) { check( cmd_system_reserve( h_.get(), n ) ); }
    void upload( int n_local, int n_ghost, const double *x, const double *v,
                 const double *f, const int *type, const int *id )
    {
        check( cmd_system_upload( h_.get(), n_local, n_ghost, x, v, f, type, id ) );
    }
    void download( double *x, double *v, double *f, int *type, int *id )
    {
        check( cmd_system_download( h_.get(), x, v, f, type, id ) );
    }
    int N() const { return counts()[0]; }
    int N_local() const { return counts()[1]; }
    int N_ghost() const { return counts()[2]; }
    std::array<int, 3> counts() const
    {
        std::array<int, 3> c{};
        check( cmd_system_counts( h_.get(), &c[0], &c[1], &c[2] ) );
        return c;
    }
    void set_stream( void *hip_stream )
    {
        check( cmd_system_set_stream( h_.get(), hip_stream ) );
    }
    const char *name() const { return "System:HIP-SoA"; }
    cmd_system *handle() const { return h_.get(); }

  private:
    Handle<cmd_system, cmd_system_destroy> h_;
};

class Comm
{
  public:
    Comm( System *s, double comm_depth )
        : s_( s )
        , depth_( comm_depth )
    {
    }
    void exchange() { check( cmd_comm_exchange( s_->handle() ) ); }
    void exchange_halo() { check( cmd_comm_exchange_halo( s_->handle(), depth_ ) ); }
    void update_halo() { check( cmd_comm_update_halo( s_->handle() ) ); }
    void update_force() { check( cmd_comm_update_force( s_->handle() ) ); }
    int num_processes() const { return 1; }

  private:
    System *s_;
    double depth_;
};

class Binning
{
  public:
    explicit Binning( System *s )
        : s_( s )
    {
    }
    void create_binning( double dx, double dy, double dz, int halo_depth,
                         bool do_local, bool do_ghost, bool sort )
    {
        check( cmd_binning_create( s_->handle(), dx, dy, dz, halo_depth,
                                   do_local, do_ghost, sort, nullptr ) );
    }
    const char *name() const { return "Binning:HIP-CountingSort"; }

  private:
    System *s_;
};

class NeighborVerlet
{
  public:
    NeighborVerlet( double neigh_cut_, bool half_neigh_, int max_neigh_guess_ )
        : neigh_cut( neigh_cut_ )
        , half_neigh( half_neigh_ )
        , max_neigh_guess( max_neigh_guess_ )
    {
        cmd_neigh *n = nullptr;
        check( cmd_neigh_create( neigh_cut, half_neigh, max_neigh_guess, &n ) );
        h_ = Handle<cmd_neigh, cmd_neigh_destroy>( n );
    }
    double neigh_cut;
    bool half_neigh;
    int max_neigh_guess;

    // NeighborVerlet::create (neighbor_verlet.h:43-62)
    void create( System *s )
    {
        check( cmd_neigh_build( h_.get(), s->handle() ) );
        int rows, maxr, guess;
        int64_t nnz;
        check( cmd_neigh_info( h_.get(), &rows, &nnz, &maxr, &guess ) );
        max_neigh_guess = guess;
    }
    void set_lanes( int lanes ) { check( cmd_neigh_set_lanes( h_.get(), lanes ) ); }
    // CSR copy of the list: offsets[N_local + 1], idx[nnz]
    void get( std::vector<int64_t> &offsets, std::vector<int> &idx )
    {
        int rows, maxr, guess;
        int64_t nnz;
        check( cmd_neigh_info( h_.get(), &rows, &nnz, &maxr, &guess ) );
        offsets.resize( rows + 1 );
        idx.resize( nnz );
        check( cmd_neigh_download( h_.get(), offsets.data(), idx.data() ) );
    }
    const char *name() const
    {
        return half_neigh ? "Neighbor:HIPVerletHalf" : "Neighbor:HIPVerletFull";
    }
    cmd_neigh *handle() const { return h_.get(); }

  private:
    Handle<cmd_neigh, cmd_neigh_destroy> h_;
};

class ForceLJ
{
  public:
    explicit ForceLJ( System *s )
    {
        cmd_force *f = nullptr;
        check( cmd_force_lj_create( s->ntypes, &f ) );
        h_ = Handle<cmd_force, cmd_force_destroy>( f );
    }
    // init_coeff (force_lj_cabana_neigh_impl.h:62-89): each entry is a
    // pair_coeff line, {"pair_coeff", i, j, eps, sigma, cut}
    void init_coeff( std::vector<std::vector<std::string>> args )
    {
        for ( const auto &pair : args )
            check( cmd_force_lj_init_coeff(
                h_.get(), std::stoi( pair.at( 1 ) ), std::stoi( pair.at( 2 ) ),
                std::stod( pair.at( 3 ) ), std::stod( pair.at( 4 ) ),
                std::stod( pair.at( 5 ) ) ) );
    }
    void set_precision( int prec ) { check( cmd_force_set_precision( h_.get(), prec ) ); }
    void compute( System *s, NeighborVerlet *n )
    {
        check( cmd_force_compute( h_.get(), s->handle(), n->handle() ) );
    }
    double compute_energy( System *s, NeighborVerlet *n )
    {
        double pe = 0.0;
        check( cmd_force_energy( h_.get(), s->handle(), n->handle(), &pe ) );
        return pe;
    }
    const char *name() const { return "Force:LJCabanaVerletHIP"; }
    const char *system_name() const { return "ForceSystemNone"; }
    cmd_force *handle() const { return h_.get(); }

  private:
    Handle<cmd_force, cmd_force_destroy> h_;
};

class Integrator
{
  public:
    explicit Integrator( System * ) {}
    void initial_integrate( System *s ) { check( cmd_nve_initial( s->handle() ) ); }
    void final_integrate( System *s ) { check( cmd_nve_final( s->handle() ) ); }
};

class Temperature
{
  public:
    explicit Temperature( Comm * = nullptr ) {}
    double compute( System *s )
    {
        double T = 0.0;
        check( cmd_thermo_temperature( s->handle(), &T ) );
        return T;
    }
};

class KinE
{
  public:
    explicit KinE( Comm * = nullptr ) {}
    double compute( System *s )
    {
        double ke = 0.0;
        check( cmd_thermo_kinetic( s->handle(), &ke ) );
        return ke;
    }
};

class PotE
{
  public:
    explicit PotE( Comm * = nullptr ) {}
    double compute( System *s, ForceLJ *f, NeighborVerlet *n )
    {
        return f->compute_energy( s, n );
    }
};

// InputFile::read_file (inputFile_impl.h:125-534) for the deck subset the
// timestep path needs; force_iteration 1 = NEIGH_HALF (newton on).
class InputFile
{
  public:
    cmd_params params;
    InputFile() { check( cmd_params_default( &params ) ); }
    void read_file( const char *path, int force_iteration = -1 )
    {
        check( cmd_params_parse_deck( path, force_iteration, &params ) );
    }
};

class CbnMD
{
  public:
    explicit CbnMD( const cmd_params &p )
    {
        cmd_md *m = nullptr;
        check( cmd_md_create( &p, &m ) );
        h_ = Handle<cmd_md, cmd_md_destroy>( m );
    }
    void set_precision( int prec ) { check( cmd_md_set_precision( h_.get(), prec ) ); }
    void set_lanes( int lanes ) { check( cmd_md_set_lanes( h_.get(), lanes ) ); }
    void set_tile( int height ) { check( cmd_md_set_tile( h_.get(), height ) ); }
    void set_ghost_fusion( bool on ) { check( cmd_md_set_ghost_fusion( h_.get(), on ? 1 : 0 ) ); }
    void init() { check( cmd_md_init( h_.get() ) ); }
    void run( int nsteps ) { check( cmd_md_run( h_.get(), nsteps ) ); }
    // rows of (step, T, PE/N, (PE+KE)/N, W)
    std::vector<std::array<double, 5>> thermo() const
    {
        int n = 0;
        check( cmd_md_thermo( h_.get(), 0, nullptr, &n ) );
        std::vector<std::array<double, 5>> rows( n );
        if ( n > 0 )
            check( cmd_md_thermo( h_.get(), n, rows[0].data(), &n ) );
        return rows;
    }
    // force, neigh, comm, integrate, other, total (seconds)
    std::array<double, 6> timers() const
    {
        std::array<double, 6> t{};
        check( cmd_md_timers( h_.get(), t.data() ) );
        return t;
    }
    cmd_md *handle() const { return h_.get(); }

  private:
    Handle<cmd_md, cmd_md_destroy> h_;
};

// The model pair_style nnp loads through n2p2 (force_nnp_cabana_neigh_impl.h
// :41-66): symmetry functions, scaling and per-element networks.
class NNPModel
{
  public:
    // n2p2 directory (input.nn, scaling.data, weights.NNN.data) with the
    // element names and atomic numbers of the system's types
    static NNPModel load_n2p2( const char *dir, const std::vector<const char *> &names,
                               const std::vector<int> &znum )
    {
        cmd_nnp_model *m = nullptr;
        check( cmd_nnp_model_load_n2p2( dir, (int)names.size(), names.data(), znum.data(),
                                        &m ) );
        return NNPModel( m );
    }
    cmd_nnp_model *handle() const { return h_.get(); }

  private:
    explicit NNPModel( cmd_nnp_model *m )
        : h_( m )
    {
    }
    Handle<cmd_nnp_model, cmd_nnp_model_destroy> h_;
};

// ForceNNP (force_nnp_cabana_neigh_impl.h:41-130).  compute() writes local
// and ghost forces; fold them with Comm::update_force as CbnMD does.
class ForceNNP
{
  public:
    explicit ForceNNP( const NNPModel &m )
    {
        cmd_force *f = nullptr;
        check( cmd_force_nnp_create( m.handle(), &f ) );
        h_ = Handle<cmd_force, cmd_force_destroy>( f );
    }
    void set_precision( int prec ) { check( cmd_force_set_precision( h_.get(), prec ) ); }
    void compute( System *s, NeighborVerlet *n )
    {
        check( cmd_force_compute( h_.get(), s->handle(), n->handle() ) );
    }
    double compute_energy( System *s, NeighborVerlet *n )
    {
        double pe = 0.0;
        check( cmd_force_energy( h_.get(), s->handle(), n->handle(), &pe ) );
        return pe;
    }
    const char *name() const { return "Force:NNPCabanaVerletHIP"; }
    cmd_force *handle() const { return h_.get(); }

  private:
    Handle<cmd_force, cmd_force_destroy> h_;
};

// Many independent systems stepped together (no reference counterpart).
class Batch
{
  public:
    explicit Batch( const cmd_params &p )
    {
        cmd_batch *b = nullptr;
        check( cmd_batch_create( &p, &b ) );
        h_ = Handle<cmd_batch, cmd_batch_destroy>( b );
    }
    void set_precision( int prec ) { check( cmd_batch_set_precision( h_.get(), prec ) ); }
    void set_lanes( int lanes ) { check( cmd_batch_set_lanes( h_.get(), lanes ) ); }
    void set_nnp_model( const NNPModel &m )
    {
        check( cmd_batch_set_nnp_model( h_.get(), m.handle() ) );
    }
    int add( int n, const double *x, const double *v, const int *type, const int *id,
             const double lo[3], const double hi[3] )
    {
        int h = 0;
        check( cmd_batch_add( h_.get(), n, x, v, type, id, lo, hi, &h ) );
        return h;
    }
    void remove( int handle ) { check( cmd_batch_remove( h_.get(), handle ) ); }
    void run( int nsteps ) { check( cmd_batch_run( h_.get(), nsteps ) ); }
    cmd_batch *handle() const { return h_.get(); }

  private:
    Handle<cmd_batch, cmd_batch_destroy> h_;
};

// MD request server with continuous batching (no reference counterpart).
class Server
{
  public:
    Server( const cmd_params &lj, const cmd_params *nnp, const NNPModel *model,
            int64_t hbm_budget, int max_jobs )
    {
        cmd_server *s = nullptr;
        check( cmd_server_create( &lj, nnp, model ? model->handle() : nullptr, hbm_budget,
                                  max_jobs, &s ) );
        h_ = Handle<cmd_server, cmd_server_destroy>( s );
    }
    int submit( int kind, int n, const double *x, const double *v, const int *type,
                const int *id, const double lo[3], const double hi[3], int nsteps,
                double arrival )
    {
        int j = 0;
        check( cmd_server_submit( h_.get(), kind, n, x, v, type, id, lo, hi, nsteps, arrival,
                                  &j ) );
        return j;
    }
    void drain( double timeout_s ) { check( cmd_server_drain( h_.get(), timeout_s ) ); }
    cmd_server *handle() const { return h_.get(); }

  private:
    Handle<cmd_server, cmd_server_destroy> h_;
};

} // namespace cabanamd_amd

#endif
