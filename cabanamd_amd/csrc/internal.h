// internal.h -- shared declarations of the HIP engine behind cabanamd_amd.h.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <utility>
#include <vector>

#include "../../include/cabanamd_amd.h"

namespace cmd
{
void set_error( const char *fmt, ... );
} // namespace cmd

#define CMD_HIP( call )                                                      \
    do                                                                       \
    {                                                                        \
        hipError_t e_ = ( call );                                            \
        if ( e_ != hipSuccess )                                              \
        {                                                                    \
            cmd::set_error( "%s failed: %s (%s:%d)", #call,                  \
                            hipGetErrorString( e_ ), __FILE__, __LINE__ );   \
            return CMD_ERR_HIP;                                              \
        }                                                                    \
    } while ( 0 )

#define CMD_TRY( call )                                                      \
    do                                                                       \
    {                                                                        \
        cmd_status s_ = ( call );                                            \
        if ( s_ != CMD_OK )                                                  \
            return s_;                                                       \
    } while ( 0 )

#define CMD_REQUIRE( cond, code, ... )                                       \
    do                                                                       \
    {                                                                        \
        if ( !( cond ) )                                                     \
        {                                                                    \
            cmd::set_error( __VA_ARGS__ );                                   \
            return code;                                                     \
        }                                                                    \
    } while ( 0 )

#define CMD_LAUNCH_CHECK()                                                   \
    do                                                                       \
    {                                                                        \
        hipError_t e_ = hipGetLastError();                                   \
        if ( e_ != hipSuccess )                                              \
        {                                                                    \
            cmd::set_error( "kernel launch failed: %s (%s:%d)",              \
                            hipGetErrorString( e_ ), __FILE__, __LINE__ );   \
            return CMD_ERR_HIP;                                              \
        }                                                                    \
    } while ( 0 )

namespace cmd
{

// Owning device buffer.  reserve() never shrinks; keep=true copies the first
// `used` elements into the new allocation.
template <class T>
struct DevBuf
{
    T *p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf( const DevBuf & ) = delete;
    DevBuf &operator=( const DevBuf & ) = delete;
    ~DevBuf() { release(); }
    void release()
    {
        if ( p )
            (void)hipFree( p );
        p = nullptr;
        n = 0;
    }
    cmd_status reserve( size_t m, hipStream_t st = nullptr, bool keep = false,
                        size_t used = 0 )
    {
        if ( m <= n )
            return CMD_OK;
        T *q = nullptr;
        hipError_t e = hipMalloc( (void **)&q, ( m ? m : 1 ) * sizeof( T ) );
        if ( e != hipSuccess )
        {
            set_error( "hipMalloc of %zu bytes failed: %s", m * sizeof( T ),
                       hipGetErrorString( e ) );
            return CMD_ERR_OOM;
        }
        if ( keep && p && used )
        {
            e = hipMemcpyAsync( q, p, used * sizeof( T ),
                                hipMemcpyDeviceToDevice, st );
            if ( e == hipSuccess )
                e = hipStreamSynchronize( st );
            if ( e != hipSuccess )
            {
                (void)hipFree( q );
                set_error( "buffer regrow copy failed: %s",
                           hipGetErrorString( e ) );
                return CMD_ERR_HIP;
            }
        }
        release();
        p = q;
        n = m;
        return CMD_OK;
    }
    void swap( DevBuf &o )
    {
        std::swap( p, o.p );
        std::swap( n, o.n );
    }
};

// Reusable temporary storage for device-wide scans and reductions.
// Pinned host words that the device writes (by copies, or by kernels through
// the device address d), and an event recorded after the last such write.
struct HostWords
{
    int *p = nullptr; // host address
    int *d = nullptr; // device address of the same words
    hipEvent_t ev = nullptr;
    HostWords() = default;
    HostWords( const HostWords & ) = delete;
    HostWords &operator=( const HostWords & ) = delete;
    ~HostWords()
    {
        if ( ev )
            (void)hipEventDestroy( ev );
        if ( p )
            (void)hipHostFree( p );
    }
    // allocates the words (zeroed) and the event on first use
    cmd_status init( size_t words )
    {
        if ( p )
            return CMD_OK;
        void *h = nullptr;
        hipError_t e = hipHostMalloc( &h, words * sizeof( int ), hipHostMallocDefault );
        if ( e != hipSuccess )
        {
            set_error( "hipHostMalloc of %zu bytes failed: %s", words * sizeof( int ),
                       hipGetErrorString( e ) );
            return CMD_ERR_OOM;
        }
        p = static_cast<int *>( h );
        for ( size_t k = 0; k < words; ++k )
            p[k] = 0;
        void *dp = nullptr;
        e = hipHostGetDevicePointer( &dp, h, 0 );
        if ( e == hipSuccess )
            e = hipEventCreateWithFlags( &ev, hipEventDisableTiming );
        if ( e != hipSuccess )
        {
            set_error( "pinned result words: %s", hipGetErrorString( e ) );
            return CMD_ERR_HIP;
        }
        d = static_cast<int *>( dp );
        return CMD_OK;
    }
};

// Layout of a system's pinned result words (cmd_system::hw).
enum
{
    HW_HALO = 0,       // 26 ghost totals per phase set, counted before the sort
    HW_HALO_FLAG = 26, // 1: the sorted halo pass found other totals
    HW_TILE = 32,      // the tiled build's result words (TILE_RES_*)
    HW_WORDS = 64
};

struct Scratch
{
    DevBuf<int64_t> blk;      // per-block partial sums of scans
    DevBuf<double> dpart;     // per-block partials of FP64 reductions
    DevBuf<double> dres;      // small result slots
    DevBuf<int> ires;         // small int result slots
    std::vector<double> hres; // host staging
};

// Device-wide exclusive scans (scan.hip).  out has n+1 entries, out[n] is
// the total.  If total_host is non-null the total is copied back (sync).
cmd_status scan_exclusive( const int *in, int *out, int n, Scratch &sc,
                           hipStream_t st, int64_t *total_host );
cmd_status scan_exclusive( const int *in, int64_t *out, int n, Scratch &sc,
                           hipStream_t st, int64_t *total_host );

// Fill kernels (graph-safe replacement for hipMemsetAsync).
cmd_status fill_i32( int *p, int v, size_t n, hipStream_t st );
// p[0 .. n) = q[0 .. n) = 0 in one launch.
cmd_status zero2_i32( int *p, int *q, size_t n, hipStream_t st );
cmd_status fill_f64( double *p, double v, size_t n, hipStream_t st );

struct MassTable
{
    double m[CMD_MAX_TYPES];
};

// Linked-cell grid: cell = floor((x - min) * rd) per dimension, with the
// reference's top-edge clamp and a guard against out-of-range coordinates.
struct CellGrid
{
    double min[3];
    double rd[3];
    int n[3];
    int floor_mode; // 1: floor() (Cabana LinkedCellList); 0: truncation
};

__device__ __forceinline__ int cell_coord( double c, double mn, double rd,
                                           int n, int floor_mode )
{
    const double t = ( c - mn ) * rd;
    int ic = floor_mode ? (int)floor( t ) : (int)t;
    if ( ic == n )
        ic = n - 1;
    ic = ic < 0 ? 0 : ( ic > n - 1 ? n - 1 : ic );
    return ic;
}

// Bytes of one staged position in the tiled force kernels' LDS records:
// TILE_REC for FP64 records (x, y, z doubles), TILE_REC32 for the FP32
// records of the mixed-precision full-list kernel (x, y, z floats about the
// tile's centre and the type).  Tiled list entries are byte offsets of
// records (slot * record size), so a gather is one address add and LDS reads
// with immediate offsets; the largest tile (MAX_SLOTS staged atoms plus the
// sentinel) must keep every offset inside 16 bits.
constexpr int TILE_REC = 24;
constexpr int TILE_REC32 = 16;

// Tiles of the binning grid: a tile is the run of cells [e0, e1) of one
// interior column (a, b); ntiles = nc0 * nc1 * tpc.
struct TileGrid
{
    CellGrid g;
    int nc0, nc1; // interior columns: a in [1, n0-2], b in [1, n1-2]
    int W;        // tile width: W x W interior columns per tile (1, 2 or 3)
    int nt0, nt1; // tile columns: ceil(nc0 / W) x ceil(nc1 / W)
    int E, tpc;   // tile height in cells, tiles per column
    int ntiles;
    int bw;       // waves per block of the tiled force kernels
};

// Counting sort of n positions into grid cells, stable by index.  Produces
// perm (sorted slot -> index) and start (ncell+1 cell offsets).
struct CellSort
{
    DevBuf<int> cell_of, count, cursor, start, perm;
    DevBuf<int> tmp; // per-cell fill in arrival order, before the rank pass
    int ncell = 0;
};
cmd_status cell_sort( const double4 *x, int n, const CellGrid &g, CellSort &cs,
                      Scratch &sc, hipStream_t st );
// The same sort for cell ids the caller has already written to cs.cell_of.
cmd_status cell_sort_ids( int n, int ncell, CellSort &cs, Scratch &sc,
                          hipStream_t st );

} // namespace cmd

// The System: device SoA state of one simulated box.
struct cmd_system
{
    int ntypes = 1;
    double mass[CMD_MAX_TYPES];
    double dt = 0.005, mvv2e = 1.0, boltz = 1.0;
    cmd_domain dom{};
    bool has_domain = false;
    int N = 0, n_local = 0, n_ghost = 0;
    int cap = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;

    // positions (x, y, z, type as double); x_alt is the ping-pong partner
    cmd::DevBuf<double4> x, x_alt;
    // velocities and forces: three planes of cap doubles each
    cmd::DevBuf<double> v, v_alt, f, f_alt;
    cmd::DevBuf<int> id, id_alt;

    // ghosts: per ghost its local root and the bit set of applied shifts
    // (bit p = phase p: -Lx, +Lx, -Ly, +Ly, -Lz, +Lz); send_idx holds the
    // reference's per-phase source lists back to back.
    cmd::DevBuf<int> ghost_root, ghost_code, send_idx;
    int counts[6] = { 0, 0, 0, 0, 0, 0 };
    // halo build scratch: per-block image counts of the 26 phase sets, their
    // scan, the totals and the atoms' phase masks
    cmd::DevBuf<int> halo_blk;
    // ghost totals per phase set counted before a sort (halo_precount), for
    // positions that only the sort changes until the halo pass consumes
    // them; hw holds their host copy
    cmd::DevBuf<int> halo_pre_tot;
    bool halo_pre = false;       // the totals await the next halo pass
    bool halo_pre_check = false; // that pass's mismatch flag is still unread
    int halo_pre_n = 0;
    double halo_pre_depth = 0.0;
    cmd::HostWords hw;

    cmd::CellSort bins;
    // grid of the last sorting create_binning (valid until atoms change)
    cmd::CellGrid bin_grid{};
    bool bin_grid_valid = false;
    // ghosts sorted into the binning grid (for tiled lists): gsort.perm maps
    // a sorted slot to a ghost index; xg holds ghost positions in that order
    cmd::CellSort gsort;
    cmd::DevBuf<double4> xg;
    bool xg_valid = false;
    // the fused step writes the next step's sorted ghosts into xg_alt: local
    // atom i has images img_pk[gimg.start[i] .. gimg.start[i + 1]), each the
    // sorted slot | shift code << IMG_CODE_SHIFT; the system's own ghost rows
    // are then stale until the next refresh
    cmd::DevBuf<double4> xg_alt;
    cmd::CellSort gimg;
    cmd::DevBuf<int> img_pk;
    bool img_valid = false; // img_pk matches the current ghosts
    bool ghosts_x_stale = false;
    cmd::Scratch sc;

    cmd::MassTable mass_table() const
    {
        cmd::MassTable t;
        for ( int k = 0; k < CMD_MAX_TYPES; ++k )
            t.m[k] = mass[k];
        return t;
    }
    cmd_status reserve( int n ); // keeps the first n_local+n_ghost atoms
};

struct cmd_neigh
{
    double neigh_cut = 0.0;
    int half = 0;
    int max_neigh_guess = 50;
    int lanes = 1;       // G: lanes per atom of the slab layout (1, 2, 4)
    int cap = 0;         // per-atom slab capacity (>= longest row)
    int n_rows = 0;
    int64_t nnz = -1;    // set when the CSR view is made
    int max_row = 0;
    cmd::DevBuf<int> slab;   // lane-interleaved rows, see neighbor.hip
    cmd::DevBuf<int> counts; // per-row lengths
    bool csr_valid = false;
    cmd::DevBuf<int64_t> offsets; // CSR view (on demand)
    cmd::DevBuf<int> idx;
    cmd::DevBuf<double4> xs; // positions in grid order, w = atom index
    cmd::CellSort grid_sort;
    int ng[3] = { 0, 0, 0 };
    bool used_bin_grid = false;
    hipStream_t stream = nullptr; // stream of the last build

    // Tiled layout (full lists on the binning grid, tiles.hip): rows hold
    // 16-bit slots into the LDS copy of a tile's stencil neighbourhood.
    bool allow_tiles = true;
    bool tiled = false;          // the last build produced the tiled layout
    int tile_e = 0;              // tile height in cells (0 = choose)
    int tile_w = 0;              // tile width in columns (0 = choose)
    int tile_bw = 0;             // waves per force-kernel block (0 = one per slab wave, <= 8)
    int tile_blocks = 0;         // force-kernel blocks per CU of the last tiled build
    int occ_blocks = 0;          // keep at least this many (0: any), set by the autotuner
    cmd::TileGrid tg{};
    cmd::DevBuf<int> tile_info;  // TI_WORDS ints per tile
    cmd::DevBuf<int64_t> tile_off; // slab offset of each tile (16-bit units)
    cmd::DevBuf<unsigned short> tslab;
    int max_tile_slots = 0;
    // Bytes per staged record that the tiled list's entries count in (an
    // entry is slot * rec): 24 for the FP64 records (x, y, z doubles), 16
    // for the FP32 records of the mixed-precision full-list kernel (x, y, z
    // floats about the tile's centre and the type)
    int rec = cmd::TILE_REC;       // for the next build
    int built_rec = cmd::TILE_REC; // of the tiled list as built
    const int *gperm_view = nullptr; // the system's ghost sort at build time
    // Sizes of the last tiled build (with some headroom), with which the
    // next build is launched before its own tile sizes are read back
    // (tiles_build_spec); valid only for the layout they were taken on.
    struct Spec
    {
        bool valid = false;
        int G = 0, half = 0, W = 0, E = 0, bw_req = 0, cap = 0;
        int slot_cap = 0;   // build planes: staged slots + sentinel
        int rows = 0;       // build rows: atoms of one atom column
        int slab_waves = 0; // slab waves of all tiles
        size_t lds = 0;     // build block's dynamic LDS
    } spec;
};

// LJ parameters passed by value to kernels.
struct LJParams
{
    double lj1[CMD_MAX_TYPES * CMD_MAX_TYPES];
    double lj2[CMD_MAX_TYPES * CMD_MAX_TYPES];
    double cutsq[CMD_MAX_TYPES * CMD_MAX_TYPES];
    float lj1f[CMD_MAX_TYPES * CMD_MAX_TYPES];
    float lj2f[CMD_MAX_TYPES * CMD_MAX_TYPES];
    double eshift[CMD_MAX_TYPES * CMD_MAX_TYPES]; // energy at the cutoff
    int ntypes;
};

struct cmd_nnp;

struct cmd_force
{
    int kind = 0; // 0 lj, 2 nnp
    cmd_nnp *nnp = nullptr; // NNP model, plan and per-atom buffers (nnp.hip)
    int ntypes = 1;
    double lj1[CMD_MAX_TYPES * CMD_MAX_TYPES];
    double lj2[CMD_MAX_TYPES * CMD_MAX_TYPES];
    double cutsq[CMD_MAX_TYPES * CMD_MAX_TYPES];
    int prec = CMD_PREC_FP64;
    int lanes = 4;
    bool tuned = false;
    cmd::DevBuf<LJParams> dparams; // device copy for multi-type kernels
    bool dparams_dirty = true;     // coefficients changed since the copy
};

namespace cmd
{
LJParams lj_params( const cmd_force *f );
// The device copy of a multi-type force's coefficient tables, uploaded now
// if they changed since the last upload (a copy and a wait on st); the
// force kernels then use it without a host copy, so they may be captured.
cmd_status lj_params_upload( const cmd_force *f, hipStream_t st );

// A ghost image entry packs its sorted slot (below 2^26) and its 6-bit
// shift code.
constexpr int IMG_CODE_SHIFT = 26;
constexpr int IMG_SLOT_MASK = ( 1 << IMG_CODE_SHIFT ) - 1;

// Integration constants for the fused force epilogue.
struct StepCtl
{
    double dtf, dtv;      // 0.5*dt/mvv2e and dt
    double L[3];          // box lengths for the wrap
    int do_final;         // finish this step's velocities
    int do_next_initial;  // start the next step (half kick + drift)
    int wrap_next;        // wrap the drifted positions (next step rebuilds)
    int thermo;           // accumulate KE, PE and virial partials
    int write_f;          // store f
    int zero_f;           // (integrator) clear f after reading it: the next
                          // step's forces accumulate into it
    // when set (with do_next_initial and no wrap), the drifted position of
    // every local atom is also written, shifted per its images' codes, to
    // xg_next at its sorted ghost slots: the next step's ghost refresh
    double4 *xg_next = nullptr;
    const int *img_start = nullptr; // per local atom, into img_pk
    const int *img_pk = nullptr;    // sorted slot | code << IMG_CODE_SHIFT
};

// LJ kernels (lj.hip)
cmd_status lj_full( const cmd_force *F, cmd_system *S, const cmd_neigh *N,
                    const StepCtl &ctl, double4 *x_next, double *partial,
                    int *nblocks_out );
cmd_status lj_half( const cmd_force *F, cmd_system *S, const cmd_neigh *N,
                    int thermo, double *partial, int *nblocks_out );
cmd_status lj_energy_virial( const cmd_force *F, cmd_system *S,
                             const cmd_neigh *N, double *pe, double *w );

// NNP force (nnp.hip)
cmd_status nnp_compute( cmd_nnp *M, cmd_system *S, cmd_neigh *N );
cmd_status nnp_reference_energy( cmd_nnp *M, cmd_system *S, double *pe );
cmd_status nnp_energy_sum( cmd_nnp *M, cmd_system *S, double *sum );
// Atomic energies E_i of the last compute as FP64 into out[0..n) (device).
cmd_status nnp_atom_energy( cmd_nnp *M, int n, double *out, hipStream_t st );
// The per-atom offset ForceNNP::compute_energy adds (element 0's atom energy).
double nnp_energy_offset( const cmd_nnp *M );
cmd_status nnp_destroy( cmd_nnp *M );
cmd_status nnp_set_precision( cmd_nnp *M, int prec );
cmd_status nnp_model_copy( const cmd_nnp_model *m, cmd_nnp_model **out );
// The model of a pair_style nnp deck: the n2p2 directory, elements from its
// input.nn 'elements' line, weights files by atomic number.
cmd_status nnp_load_deck_model( const char *dir, cmd_nnp_model **out );

// Integrator kernels (core.hip)
cmd_status nve_final_initial( cmd_system *S, const StepCtl &ctl,
                              double4 *x_next, double *partial,
                              int *nblocks_out );
// Reduce nblocks rows of 4 partials into out[4] (device) deterministically.
cmd_status reduce_partials( const double *partial, int nblocks, double *out,
                            hipStream_t st );
cmd_status ghost_refresh( cmd_system *S );
// System::permute of the owned atoms: slot k takes atom perm[k] (x, v, f, id).
cmd_status permute_local( cmd_system *S, const int *perm );
cmd_status wrap_local( cmd_system *S );
// The halo's ghost totals of the current positions, counted before a sort
// that only permutes them: the next cmd_comm_exchange_halo at the same depth
// takes its ghost count from the pinned copy of these totals (waiting for
// this count's event, not for the sorted pass) and checks on the device that
// the sorted pass finds the same totals; halo_precount_check reads that
// result once the stream has passed it.
cmd_status halo_precount( cmd_system *S, double comm_depth );
cmd_status halo_precount_check( cmd_system *S );
cmd_status force_compute_plain( cmd_force *F, cmd_system *S, cmd_neigh *N );
cmd_status neigh_build( cmd_neigh *N, cmd_system *S );
cmd_status neigh_csr( cmd_neigh *N );
size_t slab_entries( int n_local, int cap, int lanes );
// Layout search for the fused full-list step; a layout scores its step
// time plus its list build time over the `every` steps the list serves.
cmd_status autotune_lanes( cmd_force *F, cmd_system *S, cmd_neigh *N, int every );
// tiles.hip
cmd_status tiles_prepare_ghosts( cmd_system *S );
// cmd_binning_create with the forces optionally left out of the
// permutation (keep_f false: the caller recomputes them before reading).
cmd_status binning_create( cmd_system *s, double dx_in, double dy_in, double dz_in,
                           int halo_depth, int do_local, int do_ghost, int sort, int *perm_out,
                           bool keep_f );
cmd_status tiles_build( cmd_neigh *N, cmd_system *S, bool *built );
cmd_status tiles_build_spec( cmd_neigh *N, cmd_system *S, bool *built );
// Result words of the tiled build in the system's int scratch (sc.ires):
// the tile-info maxima, then the build's longest row, its capacity flag and
// the bounding-box test.
enum
{
    TILE_RES_INFO = 0,     // 5 words, k_tile_info
    TILE_RES_MAXROW = 5,   // longest row of the build
    TILE_RES_OVER = 6,     // 1: a block's needs exceeded the launch sizes
    TILE_RES_INSIDE = 7,   // 1: every position lies in the binning grid
    TILE_RES_WAVES = 8,    // slab waves of all tiles (written by the build)
    TILE_RES_WORDS = 9
};
static_assert( HW_TILE + TILE_RES_WORDS <= HW_WORDS, "tile results fit the pinned words" );
cmd_status tiles_csr( cmd_neigh *N, int64_t *nnz );
cmd_status tiles_ghost_refresh( cmd_system *S );
} // namespace cmd
