// serve.cpp -- MD request server with continuous batching.
//
// No reference counterpart (CabanaMD runs one deck per process).  Jobs are
// independent MD runs (LJ or NNP systems with a step count) that arrive over
// time.  The server keeps two batches resident on the device, one per force
// (batch.hip), and loops:
//   admit   -- jobs whose arrival time has passed join their batch in arrival
//              order while their estimated device footprint fits the budget
//              and a slot is free (first come, first served);
//   advance -- both batches step together, each on its own stream from its
//              own host thread, so LJ and NNP work overlaps on the GPU.  A
//              chunk ends at the next rebuild boundary or at the earliest job
//              end, whichever comes first; new jobs join at a boundary;
//   retire  -- a job that has run its steps is copied out (positions and
//              velocities by atom id, thermo rows), timestamped and removed.
// A job's trajectory does not depend on the other jobs (batch.hip), so its
// result is the same alone or in any mix.
#include "../../include/cabanamd_amd.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdarg>
#include <cstring>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

namespace cmd
{
void set_error( const char *fmt, ... );
}

namespace
{
// Device bytes per atom a job is charged at admission: positions, velocities
// and forces with their alternates, ids, ghost images, the list (about 100
// entries), sort buffers; NNP adds G and dE/dG rows (96 floats each) and the
// CSR copy of the list.
constexpr int64_t kLjBytesPerAtom = 800;
constexpr int64_t kNnpBytesPerAtom = 2400;

enum JobState
{
    QUEUED = 1,
    RUNNING = 2,
    DONE = 3
};

struct Job
{
    int kind = 0; // 0 lj, 2 nnp
    int n = 0, nsteps = 0;
    double t_arrival = 0, t_admit = -1, t_done = -1;
    int64_t bytes = 0;
    int state = QUEUED;
    int handle = 0;
    std::vector<double> x, v;
    std::vector<int> type, id;
    double lo[3], hi[3];
    // results
    std::vector<double> rx, rv, thermo;
    std::vector<int> rid;
};

double clock_s()
{
    return std::chrono::duration<double>( std::chrono::steady_clock::now().time_since_epoch() )
        .count();
}
} // namespace

struct cmd_server
{
    cmd_batch *lj = nullptr, *nnp = nullptr;
    int every = 20;
    int64_t budget = 0, used = 0;
    int max_jobs = 0, running = 0;
    double t0 = 0;
    std::vector<Job> jobs;
    size_t next_admit[2] = { 0, 0 }; // per kind (lj, nnp), in submission order
    double atom_steps = 0, busy_s = 0;
    int64_t ticks = 0;
};

namespace
{
cmd_batch *batch_of( cmd_server *s, const Job &j ) { return j.kind == 2 ? s->nnp : s->lj; }

// Admits arrived jobs of one kind in arrival order while memory and slots
// allow (the head of a kind's queue blocks the rest of that kind only).
cmd_status admit( cmd_server *s, int kind, double now )
{
    size_t &next = s->next_admit[kind == 2];
    for ( ; next < s->jobs.size(); ++next )
    {
        Job &j = s->jobs[next];
        if ( j.kind != kind )
            continue;
        if ( j.t_arrival > now || s->running >= s->max_jobs || s->used + j.bytes > s->budget )
            break;
        cmd_batch *b = batch_of( s, j );
        int h = 0;
        cmd_status st = cmd_batch_add( b, j.n, j.x.data(), j.v.data(), j.type.data(),
                                       j.id.data(), j.lo, j.hi, &h );
        if ( st != CMD_OK )
            return st;
        j.handle = h;
        j.state = RUNNING;
        j.t_admit = now;
        s->used += j.bytes;
        ++s->running;
        // the batch holds its own copy now
        std::vector<double>().swap( j.x );
        std::vector<double>().swap( j.v );
        std::vector<int>().swap( j.type );
        std::vector<int>().swap( j.id );
    }
    return CMD_OK;
}

cmd_status retire( cmd_server *s, Job &j )
{
    cmd_batch *b = batch_of( s, j );
    const int n = j.n;
    std::vector<double> x( 3 * (size_t)n ), v( 3 * (size_t)n );
    std::vector<int> id( n );
    cmd_status st = cmd_batch_download( b, j.handle, x.data(), v.data(), nullptr, nullptr,
                                        id.data() );
    if ( st != CMD_OK )
        return st;
    std::vector<int> o( n );
    std::iota( o.begin(), o.end(), 0 );
    std::stable_sort( o.begin(), o.end(), [&]( int a, int c ) { return id[a] < id[c]; } );
    j.rx.resize( 3 * (size_t)n );
    j.rv.resize( 3 * (size_t)n );
    j.rid.resize( n );
    for ( int k = 0; k < n; ++k )
    {
        for ( int c = 0; c < 3; ++c )
        {
            j.rx[3 * k + c] = x[3 * (size_t)o[k] + c];
            j.rv[3 * k + c] = v[3 * (size_t)o[k] + c];
        }
        j.rid[k] = id[o[k]];
    }
    int rows = 0;
    if ( ( st = cmd_batch_thermo( b, j.handle, 0, nullptr, &rows ) ) != CMD_OK )
        return st;
    j.thermo.resize( 4 * (size_t)rows );
    if ( ( st = cmd_batch_thermo( b, j.handle, rows, j.thermo.data(), &rows ) ) != CMD_OK )
        return st;
    if ( ( st = cmd_batch_remove( b, j.handle ) ) != CMD_OK )
        return st;
    j.state = DONE;
    j.t_done = clock_s() - s->t0;
    s->atom_steps += (double)j.n * j.nsteps;
    s->used -= j.bytes;
    --s->running;
    return CMD_OK;
}

// Length of the next chunk of one kind's batch: up to its next rebuild
// boundary, no further than the earliest end of its running jobs; 0 when
// it has nothing to step.  Jobs waiting to join only move the batch to its
// boundary.
cmd_status chunk( cmd_server *s, int kind, int *k_out )
{
    *k_out = 0;
    cmd_batch *b = kind == 2 ? s->nnp : s->lj;
    if ( !b )
        return CMD_OK;
    int step = 0;
    cmd_status st = cmd_batch_counts( b, nullptr, nullptr, &step, nullptr );
    if ( st != CMD_OK )
        return st;
    const bool boundary = step % s->every == 0;
    int k = s->every - step % s->every;
    bool work = false;
    for ( Job &j : s->jobs )
    {
        if ( j.state != RUNNING || j.kind != kind )
            continue;
        int n = 0, ls = 0, state = 0;
        if ( ( st = cmd_batch_system_info( b, j.handle, &n, &ls, &state ) ) != CMD_OK )
            return st;
        work = true;
        if ( state == 1 && !boundary )
            continue;
        k = std::min( k, j.nsteps - ls );
    }
    *k_out = work ? std::max( k, 1 ) : 0;
    return CMD_OK;
}

cmd_status retire_finished( cmd_server *s, int kind )
{
    cmd_batch *b = kind == 2 ? s->nnp : s->lj;
    for ( Job &j : s->jobs )
    {
        if ( j.state != RUNNING || j.kind != kind )
            continue;
        int n = 0, ls = 0, state = 0;
        cmd_status st = cmd_batch_system_info( b, j.handle, &n, &ls, &state );
        if ( st != CMD_OK )
            return st;
        if ( state == 2 && ls >= j.nsteps )
            if ( ( st = retire( s, j ) ) != CMD_OK )
                return st;
    }
    return CMD_OK;
}

// One round.  The NNP batch steps one chunk on a worker thread while this
// thread keeps admitting, stepping and retiring LJ jobs chunk by chunk, so
// short LJ jobs do not wait behind long NNP chunks; NNP jobs are admitted
// and retired between rounds (the worker owns the NNP batch meanwhile).
cmd_status advance( cmd_server *s, bool *progress )
{
    *progress = false;
    const double t0 = clock_s();
    cmd_status st = CMD_OK, st_nnp = CMD_OK;
    int kn = 0;
    if ( s->nnp )
    {
        if ( ( st = admit( s, 2, t0 - s->t0 ) ) != CMD_OK )
            return st;
        if ( ( st = chunk( s, 2, &kn ) ) != CMD_OK )
            return st;
    }
    std::atomic<bool> busy( kn > 0 );
    std::thread th;
    if ( kn > 0 )
        th = std::thread( [&] {
            st_nnp = cmd_batch_run( s->nnp, kn );
            busy = false;
        } );
    do
    {
        if ( ( st = admit( s, 0, clock_s() - s->t0 ) ) != CMD_OK )
            break;
        int kl = 0;
        if ( ( st = chunk( s, 0, &kl ) ) != CMD_OK )
            break;
        if ( kl > 0 )
        {
            if ( ( st = cmd_batch_run( s->lj, kl ) ) != CMD_OK )
                break;
            if ( ( st = retire_finished( s, 0 ) ) != CMD_OK )
                break;
            *progress = true;
        }
        else if ( busy )
            std::this_thread::sleep_for( std::chrono::microseconds( 200 ) );
    } while ( busy );
    if ( th.joinable() )
        th.join();
    s->busy_s += clock_s() - t0;
    ++s->ticks;
    if ( st != CMD_OK )
        return st;
    if ( st_nnp != CMD_OK )
        return st_nnp;
    if ( kn > 0 )
    {
        *progress = true;
        if ( ( st = retire_finished( s, 2 ) ) != CMD_OK )
            return st;
    }
    return CMD_OK;
}
} // namespace

extern "C" {

cmd_status cmd_server_create( const cmd_params *lj, const cmd_params *nnp,
                              const cmd_nnp_model *model, int64_t hbm_budget, int max_jobs,
                              cmd_server **out )
{
    if ( !lj || !out || hbm_budget <= 0 || max_jobs <= 0 )
    {
        cmd::set_error( "LJ parameters, a positive budget, max_jobs and out are required" );
        return CMD_ERR_ARG;
    }
    if ( nnp && ( !model || nnp->every != lj->every ) )
    {
        cmd::set_error( "NNP jobs need a model and the LJ jobs' rebuild interval" );
        return CMD_ERR_ARG;
    }
    cmd_server *s = new cmd_server();
    s->every = lj->every;
    s->budget = hbm_budget;
    s->max_jobs = max_jobs;
    s->t0 = clock_s();
    cmd_status st = cmd_batch_create( lj, &s->lj );
    if ( st == CMD_OK && nnp )
        st = cmd_batch_create( nnp, &s->nnp );
    if ( st == CMD_OK && nnp )
        st = cmd_batch_set_nnp_model( s->nnp, model );
    if ( st != CMD_OK )
    {
        cmd_server_destroy( s );
        return st;
    }
    *out = s;
    return CMD_OK;
}

cmd_status cmd_server_destroy( cmd_server *s )
{
    if ( !s )
        return CMD_OK;
    cmd_batch_destroy( s->lj );
    cmd_batch_destroy( s->nnp );
    delete s;
    return CMD_OK;
}

cmd_status cmd_server_configure( cmd_server *s, int prec, int lanes )
{
    if ( !s )
        return CMD_ERR_ARG;
    cmd_status st = cmd_batch_set_precision( s->lj, prec );
    if ( st == CMD_OK )
        st = cmd_batch_set_lanes( s->lj, lanes );
    if ( st == CMD_OK && s->nnp )
        st = cmd_batch_set_precision( s->nnp, prec );
    return st;
}

cmd_status cmd_server_clock( const cmd_server *s, double *t )
{
    if ( !s || !t )
        return CMD_ERR_ARG;
    *t = clock_s() - s->t0;
    return CMD_OK;
}

cmd_status cmd_server_submit( cmd_server *s, int kind, int n, const double *x, const double *v,
                              const int *type, const int *id, const double lo[3],
                              const double hi[3], int nsteps, double arrival, int *job )
{
    if ( !s || !x || !type || !lo || !hi || n <= 0 || nsteps <= 0 )
    {
        cmd::set_error( "server, n > 0, x, type, lo, hi and nsteps > 0 are required" );
        return CMD_ERR_ARG;
    }
    if ( kind != 0 && !( kind == 2 && s->nnp ) )
    {
        cmd::set_error( "job kind %d is not served (0 lj%s)", kind, s->nnp ? ", 2 nnp" : "" );
        return CMD_ERR_ARG;
    }
    if ( !s->jobs.empty() && arrival < s->jobs.back().t_arrival )
    {
        cmd::set_error( "jobs are submitted in arrival order" );
        return CMD_ERR_ARG;
    }
    Job j;
    j.kind = kind;
    j.n = n;
    j.nsteps = nsteps;
    j.t_arrival = arrival;
    j.bytes = (int64_t)n * ( kind == 2 ? kNnpBytesPerAtom : kLjBytesPerAtom );
    if ( j.bytes > s->budget )
    {
        cmd::set_error( "a job of %d atoms needs about %lld bytes, over the budget %lld", n,
                        (long long)j.bytes, (long long)s->budget );
        return CMD_ERR_CAPACITY;
    }
    j.x.assign( x, x + 3 * (size_t)n );
    j.v.assign( 3 * (size_t)n, 0.0 );
    if ( v )
        j.v.assign( v, v + 3 * (size_t)n );
    j.type.assign( type, type + n );
    j.id.resize( n );
    for ( int i = 0; i < n; ++i )
        j.id[i] = id ? id[i] : i + 1;
    for ( int d = 0; d < 3; ++d )
    {
        j.lo[d] = lo[d];
        j.hi[d] = hi[d];
    }
    s->jobs.push_back( std::move( j ) );
    if ( job )
        *job = (int)s->jobs.size() - 1;
    return CMD_OK;
}

// One scheduling round: admit, advance one chunk, retire.
cmd_status cmd_server_poll( cmd_server *s, int *queued, int *running, int *done )
{
    if ( !s )
        return CMD_ERR_ARG;
    bool progress = false;
    cmd_status st = advance( s, &progress );
    int q = 0, r = 0, d = 0;
    for ( const Job &j : s->jobs )
        ( j.state == QUEUED ? q : ( j.state == RUNNING ? r : d ) )++;
    if ( queued )
        *queued = q;
    if ( running )
        *running = r;
    if ( done )
        *done = d;
    return st;
}

// Polls until every submitted job is done, sleeping until the next arrival
// when nothing runs; CMD_ERR_STATE after timeout_s seconds.
cmd_status cmd_server_drain( cmd_server *s, double timeout_s )
{
    if ( !s )
        return CMD_ERR_ARG;
    const double start = clock_s();
    for ( ;; )
    {
        int q = 0, r = 0, d = 0;
        cmd_status st = cmd_server_poll( s, &q, &r, &d );
        if ( st != CMD_OK )
            return st;
        if ( q == 0 && r == 0 )
            return CMD_OK;
        if ( clock_s() - start > timeout_s )
        {
            cmd::set_error( "server drain timed out with %d queued and %d running jobs", q, r );
            return CMD_ERR_STATE;
        }
        if ( r == 0 )
        {
            // idle: sleep until the next arrival
            double next = 1e300;
            for ( const Job &j : s->jobs )
                if ( j.state == QUEUED )
                    next = std::min( next, j.t_arrival );
            const double wait = next - ( clock_s() - s->t0 );
            if ( wait > 0 && next < 1e299 )
                std::this_thread::sleep_for( std::chrono::duration<double>( wait ) );
        }
    }
}

cmd_status cmd_server_job( const cmd_server *s, int job, int *state, double times[3],
                           int *n )
{
    if ( !s || job < 0 || job >= (int)s->jobs.size() )
    {
        cmd::set_error( "no job %d", job );
        return CMD_ERR_ARG;
    }
    const Job &j = s->jobs[job];
    if ( state )
        *state = j.state;
    if ( times )
    {
        times[0] = j.t_arrival;
        times[1] = j.t_admit;
        times[2] = j.t_done;
    }
    if ( n )
        *n = j.n;
    return CMD_OK;
}

cmd_status cmd_server_result( const cmd_server *s, int job, double *x, double *v, int *id,
                              int max_rows, double *thermo, int *rows )
{
    if ( !s || job < 0 || job >= (int)s->jobs.size() )
    {
        cmd::set_error( "no job %d", job );
        return CMD_ERR_ARG;
    }
    const Job &j = s->jobs[job];
    if ( j.state != DONE )
    {
        cmd::set_error( "job %d has not finished", job );
        return CMD_ERR_STATE;
    }
    if ( x )
        std::memcpy( x, j.rx.data(), sizeof( double ) * j.rx.size() );
    if ( v )
        std::memcpy( v, j.rv.data(), sizeof( double ) * j.rv.size() );
    if ( id )
        std::memcpy( id, j.rid.data(), sizeof( int ) * j.rid.size() );
    const int nr = (int)( j.thermo.size() / 4 );
    if ( rows )
        *rows = nr;
    if ( thermo )
        std::memcpy( thermo, j.thermo.data(),
                     sizeof( double ) * 4 * (size_t)std::min( nr, std::max( max_rows, 0 ) ) );
    return CMD_OK;
}

// out: atom-steps of finished jobs, seconds spent stepping, scheduling
// rounds, device bytes charged now, running jobs, finished jobs.
cmd_status cmd_server_stats( const cmd_server *s, double out[6] )
{
    if ( !s || !out )
        return CMD_ERR_ARG;
    int d = 0;
    for ( const Job &j : s->jobs )
        d += j.state == DONE;
    out[0] = s->atom_steps;
    out[1] = s->busy_s;
    out[2] = (double)s->ticks;
    out[3] = (double)s->used;
    out[4] = s->running;
    out[5] = d;
    return CMD_OK;
}

} // extern "C"
