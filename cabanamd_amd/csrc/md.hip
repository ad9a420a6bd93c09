// md.hip -- the timestep loop: CbnMD::init and CbnMD::run
// (cabanamd_impl.h:70-255 and :257-432) on one GPU.
//
// Step order kept from the reference (cabanamd_impl.h:285-399):
//   initial_integrate; rebuild (wrap, sort, ghosts, list) every `every`
//   steps, else ghost refresh; f = 0 + force; ghost force fold for half
//   lists; final_integrate; thermo every thermo_rate steps.
// Fusion: for full lists one kernel computes the force of step s, finishes
// step s (final kick, thermo partials) and starts step s+1 (half kick,
// drift, and the periodic wrap when s+1 rebuilds), writing positions to the
// ping-pong buffer.  Every value is produced by the same operations in the
// same order as the unfused kernels, so both paths give identical bits.
#include "internal.h"

#include <algorithm>
#include <string>
#include <map>
#include <chrono>
#include <cmath>
#include <cstring>

namespace cmd
{
enum
{
    PH_FORCE = 0,
    PH_NEIGH = 1,
    PH_COMM = 2,
    PH_INT = 3,
    PH_OTHER = 4
};
} // namespace cmd

struct cmd_md
{
    cmd_params p{};
    cmd_domain d{};
    cmd_system *S = nullptr;
    cmd_neigh *N = nullptr;
    cmd_force *F = nullptr;
    double neigh_cut = 0.0;
    int step = 0;
    bool inited = false;
    int prec = CMD_PREC_FP64;
    int lanes = 0;
    int tile = 0;   // tile height of the tiled list (0 = default or tuned)
    int tile_w = 0; // tile width of the tiled list (0 = default or tuned)
    bool fuse_ghosts = true; // update_halo inside the fused step (tiled lists)
    // set for one rebuild: the list is built by timing every layout
    bool retune = false;
    bool relist = false; // rebuild the list at the next step (its records changed)
    double tune_s = 0.0; // host seconds spent choosing the layout at init
    std::vector<double> thermo; // rows of 5: step, T, PE/N, (PE+KE)/N, W
    cmd_nnp_model *nnp_model = nullptr; // injected NNP model (owned)
    bool custom_state = false;          // initial state given by the caller
    std::vector<double> init_x, init_v;
    std::vector<int> init_type, init_id;
    double init_lo[3] = { 0, 0, 0 }, init_hi[3] = { 0, 0, 0 };
    cmd::DevBuf<double> partA, partB, trows;
    double timers[6] = { 0, 0, 0, 0, 0, 0 };
    bool profile = false;
    // above 1: events time one force phase in prof_every and no other phase,
    // so that their markers stay out of most steps
    int prof_every = 1;
    int prof_seen = 0;
    int prof_launches = 0;
    double prof_ms = 0.0;
    struct Ev
    {
        hipEvent_t a, b;
        int phase;
    };
    std::vector<Ev> events;
    std::vector<hipEvent_t> pool;

    ~cmd_md()
    {
        for ( auto &e : events )
        {
            pool.push_back( e.a );
            pool.push_back( e.b );
        }
        for ( auto e : pool )
            (void)hipEventDestroy( e );
        cmd_force_destroy( F );
        cmd_neigh_destroy( N );
        cmd_system_destroy( S );
        cmd_nnp_model_destroy( nnp_model );
    }
    hipEvent_t get_event()
    {
        if ( !pool.empty() )
        {
            hipEvent_t e = pool.back();
            pool.pop_back();
            return e;
        }
        return new_event();
    }
    // timing only: no system-scope fence at the marker, which would write
    // back and invalidate the caches between the timed launches
    static hipEvent_t new_event()
    {
        hipEvent_t e = nullptr;
        if ( hipEventCreateWithFlags( &e, hipEventDisableSystemFence ) != hipSuccess )
            return nullptr;
        return e;
    }
    // Phase timing by HIP events on the system's stream (profile mode).
    int begin( int phase )
    {
        if ( !profile )
            return -1;
        if ( prof_every > 1 && ( phase != cmd::PH_FORCE || prof_seen++ % prof_every != 0 ) )
            return -1;
        Ev e{ get_event(), get_event(), phase };
        (void)hipEventRecord( e.a, S->stream );
        events.push_back( e );
        return (int)events.size() - 1;
    }
    void end( int k )
    {
        if ( k >= 0 )
            (void)hipEventRecord( events[k].b, S->stream );
    }
    void collect()
    {
        for ( auto &e : events )
        {
            float ms = 0.0f;
            (void)hipEventSynchronize( e.b );
            (void)hipEventElapsedTime( &ms, e.a, e.b );
            timers[e.phase] += ms * 1e-3;
            if ( e.phase == cmd::PH_FORCE )
            {
                prof_launches++;
                prof_ms += ms;
            }
            pool.push_back( e.a );
            pool.push_back( e.b );
        }
        events.clear();
    }
};

namespace cmd
{

static cmd_status rebuild( cmd_md *md )
{
    cmd_system *S = md->S;
    int k = md->begin( PH_OTHER );
    // the ghost totals of these positions, counted ahead of the sort (which
    // only permutes them): the halo pass then has its ghost count without
    // waiting for the sort, so the rebuild's launches queue up behind it
    CMD_TRY( halo_precount( S, md->neigh_cut ) );
    // the time loop computes the forces afresh after a rebuild before it
    // reads them: they are not carried through the sort
    const cmd_status bs = binning_create( S, md->neigh_cut, md->neigh_cut, md->neigh_cut, 1,
                                          1, 0, 1, nullptr, false );
    if ( bs != CMD_OK )
    {
        S->halo_pre = false; // no halo pass consumes these totals
        return bs;
    }
    md->end( k );
    k = md->begin( PH_COMM );
    // ghost velocity and force rows come out zero (full-list kernels never
    // write the force rows)
    CMD_TRY( cmd_comm_exchange_halo( S, md->neigh_cut ) );
    md->end( k );
    k = md->begin( PH_NEIGH );
    if ( md->retune )
    {
        md->retune = false;
        CMD_TRY( autotune_lanes( md->F, S, md->N, md->p.every ) ); // leaves the list built
        md->F->lanes = md->N->lanes;
    }
    else
        CMD_TRY( cmd_neigh_build( md->N, S ) );
    md->end( k );
    // the list build has waited on the device: the halo's check is in
    return halo_precount_check( S );
}

static void thermo_row( cmd_md *md, int step, double mv2, double pe, double w )
{
    cmd_system *S = md->S;
    const int dof = 3 * S->N - 3;
    const double T = mv2 * ( S->mvv2e / ( 1.0 * dof * S->boltz ) );
    const double KE = mv2 * ( 0.5 * S->mvv2e );
    md->thermo.push_back( step );
    md->thermo.push_back( T );
    md->thermo.push_back( pe / S->N );
    md->thermo.push_back( ( pe + KE ) / S->N );
    md->thermo.push_back( w );
}

static cmd_status thermo_now( cmd_md *md, int step )
{
    cmd_system *S = md->S;
    double T = 0.0, pe = 0.0, w = 0.0, ke = 0.0;
    CMD_TRY( cmd_thermo_temperature( S, &T ) );
    if ( md->F->kind == 2 )
        CMD_TRY( cmd_force_energy( md->F, S, md->N, &pe ) );
    else
        CMD_TRY( lj_energy_virial( md->F, S, md->N, &pe, &w ) );
    CMD_TRY( cmd_thermo_kinetic( S, &ke ) );
    thermo_row( md, step, ke / ( 0.5 * S->mvv2e ), pe, w );
    return CMD_OK;
}

// The reference's step order without fusion (cabanamd_impl.h:285-399), used
// for the NNP force, whose cost dwarfs the integrator's.
static cmd_status run_steps_plain( cmd_md *md, int nsteps )
{
    cmd_system *S = md->S;
    const cmd_params &p = md->p;
    const int s0 = md->step + 1, s1 = md->step + nsteps;
    auto t0 = std::chrono::steady_clock::now();
    for ( int s = s0; s <= s1; ++s )
    {
        int k = md->begin( PH_INT );
        CMD_TRY( cmd_nve_initial( S ) );
        md->end( k );
        if ( s % p.every == 0 )
        {
            k = md->begin( PH_COMM );
            CMD_TRY( cmd_comm_exchange( S ) );
            md->end( k );
            CMD_TRY( rebuild( md ) );
        }
        else
        {
            k = md->begin( PH_COMM );
            CMD_TRY( cmd_comm_update_halo( S ) );
            md->end( k );
        }
        k = md->begin( PH_FORCE );
        CMD_TRY( cmd_force_compute( md->F, S, md->N ) );
        md->end( k );
        k = md->begin( PH_COMM );
        CMD_TRY( cmd_comm_update_force( S ) );
        md->end( k );
        k = md->begin( PH_INT );
        CMD_TRY( cmd_nve_final( S ) );
        md->end( k );
        if ( p.thermo_rate > 0 && s % p.thermo_rate == 0 )
            CMD_TRY( thermo_now( md, s ) );
    }
    CMD_HIP( hipStreamSynchronize( S->stream ) );
    md->timers[5] +=
        std::chrono::duration<double>( std::chrono::steady_clock::now() - t0 )
            .count();
    md->collect();
    md->step = s1;
    return CMD_OK;
}

static cmd_status run_steps( cmd_md *md, int nsteps )
{
    if ( md->F->kind == 2 )
        return run_steps_plain( md, nsteps );
    cmd_system *S = md->S;
    cmd_neigh *N = md->N;
    cmd_force *F = md->F;
    const cmd_params &p = md->p;
    const hipStream_t st = S->stream;
    const bool half = N->half != 0;
    const int s0 = md->step + 1, s1 = md->step + nsteps;
    const int nthermo = p.thermo_rate > 0 ? nsteps / p.thermo_rate + 2 : 0;
    CMD_TRY( md->trows.reserve( 8 * (size_t)std::max( nthermo, 1 ) ) );
    std::vector<int> thermo_steps;
    StepCtl ctl{};
    ctl.dtf = 0.5 * S->dt / S->mvv2e;
    ctl.dtv = S->dt;
    for ( int k = 0; k < 3; ++k )
        ctl.L[k] = S->dom.L[k];

    auto t0 = std::chrono::steady_clock::now();
    // a list that addresses the other precision's records is rebuilt at
    // this call's first step
    const bool relist = md->relist;
    md->relist = false;
    // initial_integrate of the first step of this call (in place)
    {
        int k = md->begin( PH_INT );
        CMD_TRY( cmd_nve_initial( S ) );
        if ( s0 % p.every == 0 || relist )
            CMD_TRY( wrap_local( S ) );
        md->end( k );
    }
    // the last fused step wrote this step's sorted ghosts (update_halo fused
    // into the step's epilogue, tiled full lists)
    bool ghosts_next = false;
    for ( int s = s0; s <= s1; ++s )
    {
        const bool rb = ( s % p.every == 0 ) || ( relist && s == s0 );
        if ( rb )
        {
            S->n_ghost = 0;
            CMD_TRY( rebuild( md ) );
        }
        else if ( ghosts_next )
        {
            S->xg.swap( S->xg_alt );
            S->ghosts_x_stale = true; // the system's ghost rows lag until a refresh
        }
        else
        {
            int k = md->begin( PH_COMM );
            CMD_TRY( cmd_comm_update_halo( S ) );
            md->end( k );
        }
        const bool th = p.thermo_rate > 0 && ( s % p.thermo_rate == 0 );
        const bool last = ( s == s1 );
        ctl.do_final = 1;
        ctl.do_next_initial = last ? 0 : 1;
        ctl.wrap_next = ( !last && ( ( s + 1 ) % p.every == 0 ) ) ? 1 : 0;
        ctl.thermo = th ? 1 : 0;
        // the fused kernel keeps this step's forces in registers; they are
        // stored only at the end of the call, for the next call's first
        // half-kick and for downloads
        ctl.write_f = half || last ? 1 : 0;
        const int n = S->n_local;
        const int G = N->lanes;
        const int nbA = std::max( N->tiled ? N->tg.ntiles : 1,
                                  ( n + 255 ) / 256 * G );
        CMD_TRY( md->partA.reserve( 4 * (size_t)nbA + 4 ) );
        CMD_TRY( md->partB.reserve( 4 * (size_t)( ( n + 255 ) / 256 + 1 ) ) );
        CMD_TRY( S->x_alt.reserve( S->cap ) );
        int nb_force = 0, nb_int = 0;
        ctl.xg_next = nullptr;
        if ( md->fuse_ghosts && N->tiled && S->xg_valid && S->img_valid &&
             S->n_ghost > 0 && ctl.do_next_initial && !ctl.wrap_next )
        {
            CMD_TRY( S->xg_alt.reserve( S->n_ghost ) );
            ctl.xg_next = S->xg_alt.p;
            ctl.img_start = S->gimg.start.p;
            ctl.img_pk = S->img_pk.p;
        }
        ghosts_next = ctl.xg_next != nullptr;
        if ( !half )
        {
            int k = md->begin( PH_FORCE );
            CMD_TRY( lj_full( F, S, N, ctl, S->x_alt.p, md->partA.p,
                              &nb_force ) );
            md->end( k );
        }
        else
        {
            // tiled half lists: the force kernel adds each ghost's reaction
            // to the atom it images (nothing left to fold), and the
            // integrator clears the forces it has read for the next step's
            // sums, so f is cleared here only at the call's first step (it
            // held the last call's forces for the first half-kick) and after
            // a list rebuild (new ghost rows)
            const bool tiled = N->tiled;
            int k = md->begin( PH_FORCE );
            if ( !tiled || rb || s == s0 )
            {
                const int nt = S->n_local + S->n_ghost;
                for ( int c = 0; c < 3; ++c )
                    CMD_TRY( fill_f64( S->f.p + (size_t)c * S->cap, 0.0, nt, st ) );
            }
            CMD_TRY( lj_half( F, S, N, ctl.thermo, md->partA.p, &nb_force ) );
            md->end( k );
            if ( !tiled )
            {
                k = md->begin( PH_COMM );
                CMD_TRY( cmd_comm_update_force( S ) );
                md->end( k );
            }
            k = md->begin( PH_INT );
            ctl.zero_f = tiled && !last ? 1 : 0;
            CMD_TRY( nve_final_initial( S, ctl, S->x_alt.p, md->partB.p,
                                        &nb_int ) );
            ctl.zero_f = 0;
            md->end( k );
        }
        if ( th )
        {
            const int r = (int)thermo_steps.size();
            CMD_TRY( reduce_partials( md->partA.p, nb_force,
                                      md->trows.p + 8 * r, st ) );
            if ( half )
                CMD_TRY( reduce_partials( md->partB.p, nb_int,
                                          md->trows.p + 8 * r + 4, st ) );
            else
                CMD_TRY( fill_f64( md->trows.p + 8 * r + 4, 0.0, 4, st ) );
            thermo_steps.push_back( s );
        }
        if ( !last )
            S->x.swap( S->x_alt );
    }
    // ghost rows of the system at the final positions (downloads, the next
    // call's first step)
    if ( S->ghosts_x_stale )
        CMD_TRY( tiles_ghost_refresh( S ) );
    CMD_HIP( hipStreamSynchronize( st ) );
    md->timers[5] +=
        std::chrono::duration<double>( std::chrono::steady_clock::now() - t0 )
            .count();
    md->collect();
    md->step = s1;
    if ( !thermo_steps.empty() )
    {
        std::vector<double> h( 8 * thermo_steps.size() );
        CMD_HIP( hipMemcpy( h.data(), md->trows.p, sizeof( double ) * h.size(),
                            hipMemcpyDeviceToHost ) );
        for ( size_t r = 0; r < thermo_steps.size(); ++r )
        {
            const double *row = &h[8 * r];
            // full: [mv2, pe, w, 0 | 0...]; half: [0, pe, w, 0 | mv2, 0, 0, 0]
            thermo_row( md, thermo_steps[r], row[0] + row[4], row[1], row[2] );
        }
    }
    return CMD_OK;
}

// Choose lanes per atom and tile height (autotune_lanes) on a short scratch
// run from the initial state, then put the initial state back.  Timed on the
// perfect lattice, layouts with more lanes per atom measure within a few
// percent of fewer lanes but fall behind once the atoms leave their sites
// (melted LJ, 2,048,000 atoms: 2 lanes 0.253 ms, 4 lanes 0.270 ms per fused
// step), so the layouts are timed at the scratch run's list rebuild.  The
// run proper then starts from the uploaded state exactly as an untuned run
// would, and no tuning falls inside CbnMD::run's timed steps.
static cmd_status tune_layout( cmd_md *md, int n, const std::vector<double> &x,
                               const std::vector<double> &v, const std::vector<int> &type,
                               const std::vector<int> &id )
{
    cmd_system *S = md->S;
    const auto t0 = std::chrono::steady_clock::now();
    CMD_TRY( cmd_comm_exchange( S ) );
    CMD_TRY( cmd_binning_create( S, md->neigh_cut, md->neigh_cut, md->neigh_cut, 1, 1, 0, 1,
                                 nullptr ) );
    CMD_TRY( cmd_comm_exchange_halo( S, md->neigh_cut ) );
    md->N->lanes = 2; // the tiled layout for the scratch steps
    md->N->tile_w = 0;
    md->N->tile_e = 0;
    md->N->tile_bw = 0;
    CMD_TRY( cmd_neigh_build( md->N, S ) );
    md->F->lanes = md->N->lanes;
    CMD_TRY( cmd_force_compute( md->F, S, md->N ) );
    // the scratch steps end on a rebuild step, where the layouts are timed
    const int k = std::min( md->p.every, 20 );
    const int every = md->p.every, thermo = md->p.thermo_rate;
    md->p.every = k;
    md->p.thermo_rate = 0;
    md->retune = true;
    cmd_status st = run_steps( md, k );
    md->p.every = every;
    md->p.thermo_rate = thermo;
    md->retune = false;
    CMD_TRY( st );
    md->F->lanes = md->N->lanes;
    // back to the initial state; the caller rebuilds in the chosen layout
    md->step = 0;
    md->thermo.clear();
    for ( double &t : md->timers )
        t = 0.0;
    md->prof_launches = 0;
    md->prof_ms = 0.0;
    CMD_TRY( cmd_system_upload( S, n, 0, x.data(), v.data(), nullptr, type.data(), id.data() ) );
    CMD_HIP( hipStreamSynchronize( S->stream ) );
    md->tune_s = std::chrono::duration<double>( std::chrono::steady_clock::now() - t0 ).count();
    return CMD_OK;
}

} // namespace cmd

using namespace cmd;

extern "C" {

cmd_status cmd_md_create( const cmd_params *p, cmd_md **out )
{
    CMD_REQUIRE( p && out, CMD_ERR_ARG, "NULL argument" );
    CMD_REQUIRE( p->force_type == 0 || p->force_type == 2, CMD_ERR_ARG,
                 "pair_style must be lj/cut or nnp" );
    CMD_REQUIRE( p->every > 0, CMD_ERR_ARG, "neigh_modify every must be > 0" );
    CMD_REQUIRE( p->ntypes >= 1 && p->ntypes <= CMD_MAX_TYPES, CMD_ERR_ARG,
                 "ntypes %d outside [1,%d]", p->ntypes, CMD_MAX_TYPES );
    cmd_md *md = new cmd_md;
    md->p = *p;
    *out = md;
    return CMD_OK;
}

cmd_status cmd_md_destroy( cmd_md *md )
{
    delete md;
    return CMD_OK;
}

cmd_status cmd_md_set_precision( cmd_md *md, int prec )
{
    CMD_REQUIRE( md, CMD_ERR_ARG, "md is NULL" );
    CMD_REQUIRE( prec == CMD_PREC_FP64 || prec == CMD_PREC_MIXED, CMD_ERR_ARG,
                 "unknown precision %d", prec );
    md->prec = prec;
    if ( md->F )
        CMD_TRY( cmd_force_set_precision( md->F, prec ) );
    // full lists address FP32 records in mixed precision, FP64 ones
    // otherwise: a built list of the other kind is rebuilt before its next use
    if ( md->N && !md->N->half && md->F && md->F->kind != 2 )
    {
        const int rec = prec == CMD_PREC_MIXED ? TILE_REC32 : TILE_REC;
        if ( md->N->rec != rec )
        {
            md->N->rec = rec;
            md->N->spec.valid = false;
            if ( md->N->tiled )
                md->relist = true;
        }
    }
    return CMD_OK;
}

cmd_status cmd_md_set_lanes( cmd_md *md, int lanes )
{
    CMD_REQUIRE( md, CMD_ERR_ARG, "md is NULL" );
    md->lanes = lanes;
    if ( md->F )
        CMD_TRY( cmd_force_set_lanes( md->F, lanes ) );
    return CMD_OK;
}

cmd_status cmd_md_set_tile( cmd_md *md, int height )
{
    CMD_REQUIRE( md, CMD_ERR_ARG, "md is NULL" );
    CMD_REQUIRE( !md->inited, CMD_ERR_STATE, "the tile height is set before init" );
    CMD_REQUIRE( height >= 0 && height <= 14, CMD_ERR_ARG,
                 "tile height must be 0 (choose) or 1..14 cells" );
    md->tile = height;
    return CMD_OK;
}

cmd_status cmd_md_set_tile_width( cmd_md *md, int width )
{
    CMD_REQUIRE( md, CMD_ERR_ARG, "md is NULL" );
    CMD_REQUIRE( !md->inited, CMD_ERR_STATE, "the tile width is set before init" );
    CMD_REQUIRE( width >= 0 && width <= 3, CMD_ERR_ARG,
                 "tile width must be 0 (choose), 1, 2 or 3 columns" );
    md->tile_w = width;
    return CMD_OK;
}

cmd_status cmd_md_set_ghost_fusion( cmd_md *md, int on )
{
    CMD_REQUIRE( md, CMD_ERR_ARG, "md is NULL" );
    md->fuse_ghosts = on != 0;
    return CMD_OK;
}

cmd_status cmd_md_set_nnp_model( cmd_md *md, const cmd_nnp_model *m )
{
    CMD_REQUIRE( md && m, CMD_ERR_ARG, "NULL argument" );
    CMD_REQUIRE( !md->inited, CMD_ERR_STATE, "the model is set before init" );
    cmd_nnp_model_destroy( md->nnp_model );
    md->nnp_model = nullptr;
    return nnp_model_copy( m, &md->nnp_model );
}

cmd_status cmd_md_set_initial_state( cmd_md *md, int n, const double *x,
                                     const double *v, const int *type,
                                     const int *id, const double lo[3],
                                     const double hi[3] )
{
    CMD_REQUIRE( md && x && type && lo && hi && n > 0, CMD_ERR_ARG,
                 "bad initial state" );
    CMD_REQUIRE( !md->inited, CMD_ERR_STATE, "the state is set before init" );
    md->custom_state = true;
    md->init_x.assign( x, x + 3 * (size_t)n );
    md->init_v.assign( 3 * (size_t)n, 0.0 );
    if ( v )
        md->init_v.assign( v, v + 3 * (size_t)n );
    md->init_type.assign( type, type + n );
    md->init_id.resize( n );
    for ( int i = 0; i < n; ++i )
        md->init_id[i] = id ? id[i] : i + 1;
    for ( int k = 0; k < 3; ++k )
    {
        md->init_lo[k] = lo[k];
        md->init_hi[k] = hi[k];
    }
    return CMD_OK;
}

cmd_status cmd_md_init( cmd_md *md )
{
    CMD_REQUIRE( md, CMD_ERR_ARG, "md is NULL" );
    CMD_REQUIRE( !md->inited, CMD_ERR_STATE, "init called twice" );
    const cmd_params &p = md->p;
    md->neigh_cut = p.force_cutoff + p.skin;
    double low[3], high[3];
    for ( int k = 0; k < 3; ++k )
    {
        low[k] = p.lattice_constant * p.region_lo[k];
        high[k] = p.lattice_constant * p.region_hi[k];
        // --vacuum (inputFile_impl.h:553-559): the box grows, the lattice
        // keeps the regions' bounds
        if ( p.vacuum_rate > 1.0 )
            high[k] *= p.vacuum_rate;
    }
    if ( !md->custom_state && p.data_in[0] )
    {
        // read_lammps_data_file (read_data.h:310-385) when the deck has
        // read_data: atoms, velocities, masses and box from the file
        int n = 0, nt = 0;
        double lo[3], hi[3];
        CMD_TRY( cmd_data_read( p.data_in, p.atom_style_charge, &n, &nt, lo, hi, nullptr,
                                nullptr, nullptr, nullptr, nullptr, nullptr ) );
        CMD_REQUIRE( nt <= p.ntypes || p.ntypes == 1, CMD_ERR_ARG,
                     "data file has %d atom types", nt );
        md->p.ntypes = std::max( p.ntypes, nt );
        md->init_x.resize( 3 * (size_t)n );
        md->init_v.resize( 3 * (size_t)n );
        md->init_type.resize( n );
        md->init_id.resize( n );
        CMD_TRY( cmd_data_read( p.data_in, p.atom_style_charge, &n, &nt, lo, hi,
                                md->init_x.data(), md->init_v.data(), nullptr,
                                md->init_type.data(), md->init_id.data(), md->p.mass ) );
        for ( int k = 0; k < 3; ++k )
        {
            md->init_lo[k] = lo[k];
            md->init_hi[k] = hi[k];
        }
        md->custom_state = true;
    }
    if ( md->custom_state )
        for ( int k = 0; k < 3; ++k )
        {
            low[k] = md->init_lo[k];
            high[k] = md->init_hi[k];
        }
    CMD_TRY( cmd_domain_create( low, high, p.comm_ghost_cutoff, &md->d ) );
    int n = 0;
    std::vector<double> x, v;
    std::vector<int> type, id;
    if ( md->custom_state )
    {
        n = (int)md->init_type.size();
        x = md->init_x;
        v = md->init_v;
        type = md->init_type;
        id = md->init_id;
    }
    else
    {
        CMD_TRY( cmd_lattice_create( &p, &md->d, &n, nullptr, nullptr, nullptr,
                                     nullptr ) );
        CMD_REQUIRE( n > 0, CMD_ERR_ARG, "the lattice creates no atoms" );
        x.resize( 3 * (size_t)n );
        v.resize( 3 * (size_t)n );
        type.resize( n );
        id.resize( n );
        CMD_TRY( cmd_lattice_create( &p, &md->d, &n, x.data(), v.data(),
                                     type.data(), id.data() ) );
    }
    CMD_TRY( cmd_system_create( p.ntypes, &md->S ) );
    cmd_system *S = md->S;
    CMD_TRY( cmd_system_set_units( S, p.dt, p.mvv2e, p.boltz ) );
    CMD_TRY( cmd_system_set_mass( S, p.ntypes, p.mass ) );
    CMD_TRY( cmd_system_set_domain( S, &md->d ) );
    CMD_TRY( cmd_system_reserve( S, n + n / 2 + 1024 ) );
    CMD_TRY( cmd_system_upload( S, n, 0, x.data(), v.data(), nullptr,
                                type.data(), id.data() ) );
    const bool nnp = p.force_type == 2;
    if ( nnp )
    {
        // ForceNNP::init_coeff (force_nnp_cabana_neigh_impl.h:41-66): the
        // model directory of pair_style nnp, unless a model was given
        if ( !md->nnp_model )
            CMD_TRY( nnp_load_deck_model( p.nnp_dir, &md->nnp_model ) );
        CMD_TRY( cmd_force_nnp_create( md->nnp_model, &md->F ) );
    }
    else
    {
        CMD_TRY( cmd_force_lj_create( p.ntypes, &md->F ) );
        for ( int c = 0; c < p.ncoeff; ++c )
            CMD_TRY( cmd_force_lj_init_coeff( md->F, p.coeff_i[c], p.coeff_j[c],
                                              p.coeff_eps[c], p.coeff_sigma[c],
                                              p.coeff_cut[c] ) );
        CMD_TRY( cmd_force_set_lanes( md->F, md->lanes ) );
    }
    CMD_TRY( cmd_force_set_precision( md->F, md->prec ) );
    CMD_TRY( cmd_neigh_create( md->neigh_cut, p.half, p.max_neigh_guess,
                               &md->N ) );
    if ( nnp )
        md->N->allow_tiles = false; // the NNP kernels read the CSR view
    else if ( md->lanes > 0 )
        CMD_TRY( cmd_neigh_set_lanes( md->N, md->lanes ) );
    else if ( p.half )
        CMD_TRY( cmd_neigh_set_lanes( md->N, 2 ) ); // tiled half list
    if ( !nnp && md->tile > 0 )
        CMD_TRY( cmd_neigh_set_tile( md->N, md->tile ) );
    if ( !nnp && md->tile_w > 0 )
        CMD_TRY( cmd_neigh_set_tile_width( md->N, md->tile_w ) );
    // the mixed-precision full-list kernel stages FP32 records
    if ( !nnp && !p.half && md->prec == CMD_PREC_MIXED )
        md->N->rec = TILE_REC32;
    if ( md->lanes == 0 && !p.half && !nnp )
        CMD_TRY( tune_layout( md, n, x, v, type, id ) );
    // cabanamd_impl.h:197-223: exchange, sort, halo, list, f = 0, force
    CMD_TRY( cmd_comm_exchange( S ) );
    CMD_TRY( cmd_binning_create( S, md->neigh_cut, md->neigh_cut,
                                 md->neigh_cut, 1, 1, 0, 1, nullptr ) );
    CMD_TRY( cmd_comm_exchange_halo( S, md->neigh_cut ) );
    CMD_TRY( cmd_neigh_build( md->N, S ) );
    md->F->lanes = md->N->lanes;
    CMD_TRY( cmd_force_compute( md->F, S, md->N ) );
    if ( p.half || nnp )
        CMD_TRY( cmd_comm_update_force( S ) );
    if ( p.thermo_rate > 0 )
        CMD_TRY( thermo_now( md, 0 ) );
    CMD_HIP( hipStreamSynchronize( S->stream ) );
    md->inited = true;
    return CMD_OK;
}

// write_data at the end of CbnMD::run (cabanamd_impl.h:430-431): the owned
// atoms in atom-id order.
static cmd_status write_data_file( cmd_md *md )
{
    cmd_system *S = md->S;
    const int n = S->n_local + S->n_ghost;
    std::vector<double> x( 3 * (size_t)n ), v( 3 * (size_t)n );
    std::vector<int> type( n ), id( n );
    CMD_TRY( cmd_system_download( S, x.data(), v.data(), nullptr, type.data(), id.data() ) );
    const int nl = S->n_local;
    std::vector<int> o( nl );
    for ( int k = 0; k < nl; ++k )
        o[k] = k;
    std::sort( o.begin(), o.end(), [&]( int a, int b ) { return id[a] < id[b]; } );
    std::vector<double> xs( 3 * (size_t)nl ), vs( 3 * (size_t)nl );
    std::vector<int> ts( nl ), is( nl );
    for ( int k = 0; k < nl; ++k )
    {
        for ( int c = 0; c < 3; ++c )
        {
            xs[3 * k + c] = x[3 * (size_t)o[k] + c];
            vs[3 * k + c] = v[3 * (size_t)o[k] + c];
        }
        ts[k] = type[o[k]];
        is[k] = id[o[k]];
    }
    return cmd_data_write( md->p.data_out, nl, md->p.ntypes, md->d.low, md->d.high,
                           xs.data(), vs.data(), ts.data(), is.data() );
}

// Owned atoms with x, v, f in atom-id order (host copies).
struct Snapshot
{
    std::vector<int> id, type;
    std::vector<double> x, v, f;
};

static cmd_status snapshot( cmd_md *md, Snapshot &out )
{
    cmd_system *S = md->S;
    const int n = S->n_local + S->n_ghost;
    std::vector<double> x( 3 * (size_t)n ), v( 3 * (size_t)n ), f( 3 * (size_t)n );
    std::vector<int> type( n ), id( n );
    CMD_TRY( cmd_system_download( S, x.data(), v.data(), f.data(), type.data(), id.data() ) );
    const int nl = S->n_local;
    std::vector<int> o( nl );
    for ( int k = 0; k < nl; ++k )
        o[k] = k;
    std::sort( o.begin(), o.end(), [&]( int a, int b ) { return id[a] < id[b]; } );
    out.id.resize( nl );
    out.type.resize( nl );
    out.x.resize( 3 * (size_t)nl );
    out.v.resize( 3 * (size_t)nl );
    out.f.resize( 3 * (size_t)nl );
    for ( int k = 0; k < nl; ++k )
    {
        out.id[k] = id[o[k]];
        out.type[k] = type[o[k]];
        for ( int c = 0; c < 3; ++c )
        {
            out.x[3 * k + c] = x[3 * (size_t)o[k] + c];
            out.v[3 * k + c] = v[3 * (size_t)o[k] + c];
            out.f[3 * k + c] = f[3 * (size_t)o[k] + c];
        }
    }
    return CMD_OK;
}

static std::string dump_name( const char *dir, int step )
{
    char b[64];
    std::snprintf( b, sizeof( b ), "/output.%010d.%03d", step, 0 );
    return std::string( dir ) + b;
}

extern "C" cmd_status cmd_md_dump( cmd_md *md, const char *dir, int step )
{
    CMD_REQUIRE( md && dir, CMD_ERR_ARG, "md and dir are required" );
    CMD_REQUIRE( md->inited, CMD_ERR_STATE, "dump before init" );
    Snapshot s;
    CMD_TRY( snapshot( md, s ) );
    const std::string name = dump_name( dir, step );
    std::FILE *fp = std::fopen( name.c_str(), "wb" );
    CMD_REQUIRE( fp, CMD_ERR_IO, "Cannot open dump file: %s", name.c_str() );
    const int n = (int)s.id.size();
    bool ok = std::fwrite( "CMDDUMP1", 1, 8, fp ) == 8;
    ok = ok && std::fwrite( &n, sizeof( int ), 1, fp ) == 1;
    ok = ok && std::fwrite( &step, sizeof( int ), 1, fp ) == 1;
    for ( int i = 0; ok && i < n; ++i )
    {
        const double q = 0.0;
        ok = std::fwrite( &s.id[i], sizeof( int ), 1, fp ) == 1 &&
             std::fwrite( &s.type[i], sizeof( int ), 1, fp ) == 1 &&
             std::fwrite( &q, sizeof( double ), 1, fp ) == 1 &&
             std::fwrite( &s.x[3 * i], sizeof( double ), 3, fp ) == 3 &&
             std::fwrite( &s.v[3 * i], sizeof( double ), 3, fp ) == 3 &&
             std::fwrite( &s.f[3 * i], sizeof( double ), 3, fp ) == 3;
    }
    ok = ( std::fclose( fp ) == 0 ) && ok;
    CMD_REQUIRE( ok, CMD_ERR_IO, "write to %s failed", name.c_str() );
    return CMD_OK;
}

extern "C" cmd_status cmd_md_check_correctness( cmd_md *md, const char *refdir, int step,
                                                const char *file, double out[6] )
{
    CMD_REQUIRE( md && refdir, CMD_ERR_ARG, "md and refdir are required" );
    CMD_REQUIRE( md->inited, CMD_ERR_STATE, "check before init" );
    Snapshot s;
    CMD_TRY( snapshot( md, s ) );
    const std::string name = dump_name( refdir, step );
    std::FILE *fp = std::fopen( name.c_str(), "rb" );
    CMD_REQUIRE( fp, CMD_ERR_IO, "Cannot open input file: %s", name.c_str() );
    char magic[8] = {};
    int n = 0, st = 0;
    bool ok = std::fread( magic, 1, 8, fp ) == 8 && std::memcmp( magic, "CMDDUMP1", 8 ) == 0 &&
              std::fread( &n, sizeof( int ), 1, fp ) == 1 &&
              std::fread( &st, sizeof( int ), 1, fp ) == 1;
    if ( !ok )
    {
        std::fclose( fp );
        set_error( "%s is not a snapshot file", name.c_str() );
        return CMD_ERR_IO;
    }
    if ( n != (int)s.id.size() )
    {
        std::fclose( fp );
        set_error( "Mismatch in current and reference atom counts" );
        return CMD_ERR_STATE;
    }
    std::map<int, int> where;
    for ( int k = 0; k < n; ++k )
        where[s.id[k]] = k;
    double sr = 0, sv = 0, sf = 0, mr = 0, mv = 0, mf = 0;
    for ( int i = 0; i < n; ++i )
    {
        int rid = 0, rt = 0;
        double q = 0, rx[3], rv[3], rf[3];
        ok = std::fread( &rid, sizeof( int ), 1, fp ) == 1 &&
             std::fread( &rt, sizeof( int ), 1, fp ) == 1 &&
             std::fread( &q, sizeof( double ), 1, fp ) == 1 &&
             std::fread( rx, sizeof( double ), 3, fp ) == 3 &&
             std::fread( rv, sizeof( double ), 3, fp ) == 3 &&
             std::fread( rf, sizeof( double ), 3, fp ) == 3;
        if ( !ok )
            break;
        const auto it = where.find( rid );
        if ( it == where.end() )
        {
            std::fclose( fp );
            set_error( "Unable to find current id matching reference id: %d", rid );
            return CMD_ERR_STATE;
        }
        const int k = it->second;
        double dr = 0, dv = 0, df = 0;
        for ( int c = 0; c < 3; ++c )
        {
            const double a = s.x[3 * k + c] - rx[c], b = s.v[3 * k + c] - rv[c],
                         d = s.f[3 * k + c] - rf[c];
            dr += a * a;
            dv += b * b;
            df += d * d;
            mr = std::max( mr, std::fabs( a ) );
            mv = std::max( mv, std::fabs( b ) );
            mf = std::max( mf, std::fabs( d ) );
        }
        sr += dr;
        sv += dv;
        sf += df;
    }
    std::fclose( fp );
    CMD_REQUIRE( ok, CMD_ERR_IO, "Error reading reference data." );
    const double res[6] = { std::sqrt( sr ), mr, std::sqrt( sv ), mv, std::sqrt( sf ), mf };
    if ( out )
        for ( int k = 0; k < 6; ++k )
            out[k] = res[k];
    if ( file )
    {
        std::FILE *fo = std::fopen( file, step == 0 ? "w" : "a" );
        CMD_REQUIRE( fo, CMD_ERR_IO, "cannot write %s", file );
        if ( step == 0 )
            std::fprintf( fo, "# timestep deltarnorm maxdelr deltavnorm maxdelv "
                              "deltafnorm maxdelf\n" );
        std::fprintf( fo, "%d %g %g %g %g %g %g\n", step, res[0], res[1], res[2], res[3],
                      res[4], res[5] );
        std::fclose( fo );
    }
    return CMD_OK;
}

cmd_status cmd_md_run( cmd_md *md, int nsteps )
{
    CMD_REQUIRE( md, CMD_ERR_ARG, "md is NULL" );
    CMD_REQUIRE( md->inited, CMD_ERR_STATE, "run before init" );
    if ( nsteps <= 0 )
        return CMD_OK;
    CMD_TRY( run_steps( md, nsteps ) );
    if ( md->p.data_out[0] )
        CMD_TRY( write_data_file( md ) );
    return CMD_OK;
}

cmd_status cmd_md_system( cmd_md *md, cmd_system **s )
{
    CMD_REQUIRE( md && s, CMD_ERR_ARG, "NULL argument" );
    *s = md->S;
    return CMD_OK;
}

cmd_status cmd_md_neigh( cmd_md *md, cmd_neigh **n )
{
    CMD_REQUIRE( md && n, CMD_ERR_ARG, "NULL argument" );
    *n = md->N;
    return CMD_OK;
}

cmd_status cmd_md_force( cmd_md *md, cmd_force **f )
{
    CMD_REQUIRE( md && f, CMD_ERR_ARG, "NULL argument" );
    *f = md->F;
    return CMD_OK;
}

cmd_status cmd_md_thermo( const cmd_md *md, int max_rows, double *rows,
                          int *rows_out )
{
    CMD_REQUIRE( md, CMD_ERR_ARG, "md is NULL" );
    const int n = (int)( md->thermo.size() / 5 );
    if ( rows_out )
        *rows_out = n;
    if ( rows )
    {
        const int m = std::min( n, max_rows );
        std::memcpy( rows, md->thermo.data(), sizeof( double ) * 5 * m );
    }
    return CMD_OK;
}

cmd_status cmd_md_timers( const cmd_md *md, double out[6] )
{
    CMD_REQUIRE( md && out, CMD_ERR_ARG, "NULL argument" );
    for ( int k = 0; k < 6; ++k )
        out[k] = md->timers[k];
    return CMD_OK;
}

cmd_status cmd_md_profile( cmd_md *md, int enable )
{
    CMD_REQUIRE( md, CMD_ERR_ARG, "md is NULL" );
    md->profile = enable != 0;
    md->prof_every = std::max( enable, 1 );
    md->prof_seen = 0;
    // events are created here rather than inside the steps they time (a
    // creation costs tens of microseconds of host time); more are created
    // on demand
    if ( md->profile )
        while ( md->pool.size() < 512 )
        {
            hipEvent_t e = md->new_event();
            if ( !e )
                break;
            md->pool.push_back( e );
        }
    md->prof_launches = 0;
    md->prof_ms = 0.0;
    for ( int k = 0; k < 6; ++k )
        md->timers[k] = 0.0;
    return CMD_OK;
}

cmd_status cmd_md_tune_info( const cmd_md *md, int *lanes, int *tile_width, int *tile,
                             double *seconds )
{
    CMD_REQUIRE( md, CMD_ERR_ARG, "md is NULL" );
    CMD_REQUIRE( md->inited, CMD_ERR_STATE, "tune info before init" );
    if ( lanes )
        *lanes = md->N->lanes;
    if ( tile_width )
        *tile_width = md->N->tiled ? md->N->tg.W : 0;
    if ( tile )
        *tile = md->N->tiled ? md->N->tg.E : 0;
    if ( seconds )
        *seconds = md->tune_s;
    return CMD_OK;
}

cmd_status cmd_md_kernel_stats( const cmd_md *md, int *launches,
                                double *mean_ms )
{
    CMD_REQUIRE( md, CMD_ERR_ARG, "md is NULL" );
    if ( launches )
        *launches = md->prof_launches;
    if ( mean_ms )
        *mean_ms = md->prof_launches ? md->prof_ms / md->prof_launches : 0.0;
    return CMD_OK;
}

} // extern "C"
