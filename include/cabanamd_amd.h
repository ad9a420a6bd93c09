/*
 * cabanamd_amd.h -- C ABI of the MI355X-native CabanaMD timestep engine.
 *
 * One shared library (libcabanamd_amd.so) exports these entry points.  They
 * replace the reference's C++ module classes on the timestep path; each block
 * below names the reference interface it stands in for (paths relative to the
 * CabanaMD source tree).  The C++ classes in cabanamd_amd.hpp wrap this ABI
 * with the reference's own class and method names.
 *
 * Conventions
 *  - Opaque handles own device-resident state (HBM).  Host arrays crossing
 *    the ABI are plain pointers: positions, velocities and forces are
 *    row-major (n,3) doubles, types are 0-based ints, ids are ints.
 *  - Every call returns cmd_status; CMD_OK is 0.  On error
 *    cmd_last_error() returns a message (thread-local).  Nothing throws
 *    across the ABI.  The reference reports errors by throwing
 *    std::runtime_error from log_err (output.h:37-54).
 *  - Calls are ordered on the handle's HIP stream (cmd_system_set_stream);
 *    calls that return host data synchronise that stream.
 *  - One host thread per handle.
 */
#ifndef CABANAMD_AMD_H
#define CABANAMD_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum cmd_status
{
    CMD_OK = 0,
    CMD_ERR_ARG = 1,      /* bad argument */
    CMD_ERR_HIP = 2,      /* HIP runtime error */
    CMD_ERR_OOM = 3,      /* device allocation failed */
    CMD_ERR_STATE = 4,    /* call out of order */
    CMD_ERR_CAPACITY = 5, /* a fixed capacity was exceeded */
    CMD_ERR_IO = 6,       /* file or deck error */
    CMD_ERR_NO_DEVICE = 7 /* no GPU visible */
} cmd_status;

enum
{
    CMD_MAX_TYPES = 8,
    CMD_MAX_REGIONS = 16
};

/* Force-kernel precision.  FP64: pair maths in double (the reference's).
 * MIXED: separations and cutoff tests in double, pair maths in float,
 * per-atom accumulation in double. */
enum
{
    CMD_PREC_FP64 = 0,
    CMD_PREC_MIXED = 1
};

const char *cmd_last_error( void );
int cmd_version( void );

/* ---- device --------------------------------------------------------------
 * Replaces Kokkos::initialize / ScopeGuard (bin/main.cpp:64). */
cmd_status cmd_device_init( int device );
cmd_status cmd_device_count( int *count );
cmd_status cmd_device_synchronize( void );
/* Streaming-copy probe (no reference counterpart): copies `bytes` between two
 * device buffers `reps` times; reports read+write GB/s and ms per copy.  Used
 * to calibrate hardware byte counters and the achievable HBM bandwidth. */
cmd_status cmd_bandwidth_probe( int64_t bytes, int reps, double *gbs,
                                double *ms_per_copy );

/* ---- domain --------------------------------------------------------------
 * System::create_domain (system.h:149-205): a periodic box of 100 mesh
 * cells per dimension, one rank, halo = ceil(ghost_cutoff / cell) cells. */
typedef struct cmd_domain
{
    double low[3], high[3];
    double L[3]; /* global_mesh_{x,y,z} */
    double cell[3];
    double local_lo[3], local_hi[3], local_mesh[3];
    double ghost_lo[3], ghost_hi[3];
    int halo_width;
} cmd_domain;

cmd_status cmd_domain_create( const double low[3], const double high[3],
                              double ghost_cutoff, cmd_domain *out );

/* ---- System --------------------------------------------------------------
 * System<Device,layout> (system.h:65-279, system_types/system_1aosoa.h):
 * N, N_local, N_ghost, ntypes, mass, dt, mvv2e, boltz, mesh bounds and the
 * x/v/f/type/id slices, held here as device SoA. */
typedef struct cmd_system cmd_system;

cmd_status cmd_system_create( int ntypes, cmd_system **out );
cmd_status cmd_system_destroy( cmd_system *s );
cmd_status cmd_system_set_stream( cmd_system *s, void *hip_stream );
cmd_status cmd_system_set_units( cmd_system *s, double dt, double mvv2e,
                                 double boltz );
cmd_status cmd_system_set_mass( cmd_system *s, int ntypes,
                                const double *mass );
cmd_status cmd_system_set_domain( cmd_system *s, const cmd_domain *d );
/* Grow device capacity to at least n atoms (System::resize, system_1aosoa.h:63-70). */
cmd_status cmd_system_reserve( cmd_system *s, int n );
/* Upload n_local owned atoms followed by n_ghost ghosts.  v, f, id may be
 * NULL (zero-filled; id defaults to index+1).  N (global) = n_local. */
cmd_status cmd_system_upload( cmd_system *s, int n_local, int n_ghost,
                              const double *x, const double *v,
                              const double *f, const int *type,
                              const int *id );
/* Download the first n = n_local + n_ghost atoms (any pointer may be NULL).
 * Ghost rows made by cmd_comm_exchange_halo carry positions, types and ids,
 * with zero velocities (the halo gathers x and type only,
 * comm_mpi_impl.h:346-347). */
cmd_status cmd_system_download( cmd_system *s, double *x, double *v,
                                double *f, int *type, int *id );
cmd_status cmd_system_counts( const cmd_system *s, int *n_global,
                              int *n_local, int *n_ghost );
/* Per-atom device storage in bytes (for admission by HBM footprint). */
cmd_status cmd_system_footprint( const cmd_system *s, int64_t *bytes );

/* ---- Integrator -----------------------------------------------------------
 * Integrator<System> (integrator_nve.h:56-110; integrator_nve_impl.h:49-83):
 * velocity-Verlet over [0, N_local). */
cmd_status cmd_nve_initial( cmd_system *s );
cmd_status cmd_nve_final( cmd_system *s );

/* ---- Comm (one rank) -----------------------------------------------------
 * Comm<System> (comm_mpi_impl.h:191-441): exchange = periodic wrap (and
 * drop ghosts), exchange_halo = 6-phase ghost build at depth comm_depth,
 * update_halo = ghost position refresh, update_force = ghost force fold. */
cmd_status cmd_comm_exchange( cmd_system *s );
cmd_status cmd_comm_exchange_halo( cmd_system *s, double comm_depth );
cmd_status cmd_comm_update_halo( cmd_system *s );
cmd_status cmd_comm_update_force( cmd_system *s );
/* Ghost bookkeeping in the reference's terms: the concatenated per-phase
 * source lists (pack_indicies) and per-phase counts. send_idx may be NULL. */
cmd_status cmd_comm_ghost_lists( cmd_system *s, int *send_idx, int counts[6] );

/* ---- Binning --------------------------------------------------------------
 * Binning<System>::create_binning (binning_cabana_impl.h:57-113): linked-cell
 * counting sort of the local atoms and permutation of every per-atom array
 * into cell order.  perm_out (optional, host, N_local ints) receives the
 * permutation: new slot k holds old atom perm_out[k]. */
cmd_status cmd_binning_create( cmd_system *s, double dx, double dy,
                               double dz, int halo_depth, int do_local,
                               int do_ghost, int sort, int *perm_out );

/* ---- Neighbor --------------------------------------------------------------
 * NeighborVerlet<System,Full|Half,...> (neighbor_verlet.h:19-70): CSR Verlet
 * list over rows [0,N_local), candidates [0,N_local+N_ghost), radius
 * neigh_cut.  max_neigh_guess keeps the reference's regrow-by-1.1 semantics
 * for the reported capacity. */
typedef struct cmd_neigh cmd_neigh;

cmd_status cmd_neigh_create( double neigh_cut, int half, int max_neigh_guess,
                             cmd_neigh **out );
cmd_status cmd_neigh_destroy( cmd_neigh *n );
/* Lanes per atom of the list's device layout (1, 2, 4 or 8; default 1).  The
 * force kernels follow the layout of the list they are given. */
cmd_status cmd_neigh_set_lanes( cmd_neigh *n, int lanes );
/* Tiled layout control (no reference counterpart; for tuning and tests):
 * tile height in binning cells (0 = choose) and tile width, 1, 2 or 3 binning
 * columns per side (0 = choose: 2); the shape of the last tiled build
 * (0, 0 when the list is not tiled). */
cmd_status cmd_neigh_set_tile( cmd_neigh *n, int height );
cmd_status cmd_neigh_set_tile_width( cmd_neigh *n, int width );
cmd_status cmd_neigh_tile_shape( const cmd_neigh *n, int *width, int *height );
cmd_status cmd_neigh_build( cmd_neigh *n, cmd_system *s );
cmd_status cmd_neigh_info( const cmd_neigh *n, int *n_rows, int64_t *nnz,
                           int *max_row, int *max_neigh_guess );
/* Layout of the last build (no reference counterpart; for tuning and
 * tests): tiled = 1 for the LDS-tiled layout, with its tile height in
 * cells; lanes per atom.  Any pointer may be NULL. */
cmd_status cmd_neigh_layout( const cmd_neigh *n, int *tiled, int *tile_height,
                             int *lanes );
/* offsets: n_rows+1 int64; idx: nnz ints (either may be NULL). */
cmd_status cmd_neigh_download( cmd_neigh *n, int64_t *offsets, int *idx );

/* ---- Force -----------------------------------------------------------------
 * Force<System,Neighbor> (force.h:57-74) with ForceLJ
 * (force_types/force_lj_cabana_neigh_impl.h:49-377). */
typedef struct cmd_force cmd_force;

cmd_status cmd_force_lj_create( int ntypes, cmd_force **out );
cmd_status cmd_force_destroy( cmd_force *f );
/* pair_coeff i j eps sigma cut, 1-based types (impl.h:62-89). */
cmd_status cmd_force_lj_init_coeff( cmd_force *f, int i, int j, double eps,
                                    double sigma, double cut );
cmd_status cmd_force_set_precision( cmd_force *f, int prec );
/* Lanes per atom the timestep loop lays its list out for (1, 2, 4);
 * 0 = autotune at init by timing each layout. */
cmd_status cmd_force_set_lanes( cmd_force *f, int lanes );
cmd_status cmd_force_get_lanes( const cmd_force *f, int *lanes );
/* f = 0 over local+ghost, then accumulate (cabanamd_impl.h:336-341). */
cmd_status cmd_force_compute( cmd_force *f, cmd_system *s, cmd_neigh *n );
/* Shifted pair energy with the reference's list weights (impl.h:261-377). */
cmd_status cmd_force_energy( cmd_force *f, cmd_system *s, cmd_neigh *n,
                             double *pe );
/* Pair virial sum r.F (not in the reference; used for pressure). */
cmd_status cmd_force_virial( cmd_force *f, cmd_system *s, cmd_neigh *n,
                             double *w );

/* ---- NNP force --------------------------------------------------------------
 * ForceNNP (force_types/force_nnp_cabana_neigh_impl.h:41-130).  The model is
 * what init_coeff loads through n2p2 (:41-66): symmetry functions (types 2
 * and 3, cosine cutoff), min-max scaling and per-element networks.  Forces
 * are written to local and ghost rows; fold them with cmd_comm_update_force,
 * as the reference does after an NNP force (cabanamd_impl.h:346-353).
 * cmd_force_set_precision: CMD_PREC_MIXED runs the network layers as BF16
 * MFMA GEMMs with FP32 accumulation (default), CMD_PREC_FP64 runs them in
 * FP32 on the vector units; symmetry functions and forces are FP32 either
 * way, forces are accumulated in FP64. */
typedef struct cmd_nnp_model cmd_nnp_model;

cmd_status cmd_nnp_model_create( int nelem, int n_hidden, const int *hidden,
                                 const char *activations, cmd_nnp_model **out );
cmd_status cmd_nnp_model_destroy( cmd_nnp_model *m );
cmd_status cmd_nnp_model_set_scaling( cmd_nnp_model *m, int scale, double smin,
                                      double smax );
/* type 2: e1 = neighbour element (-1 any); type 3: e1 = e2 = -1 (any pair). */
cmd_status cmd_nnp_model_add_sf( cmd_nnp_model *m, int elem, int type, int e1,
                                 int e2, double eta, double rs, double lambda,
                                 double zeta, double rc, double gmin,
                                 double gmax );
/* W is [n_out][n_in] row-major. */
cmd_status cmd_nnp_model_set_layer( cmd_nnp_model *m, int elem, int layer,
                                    int n_out, int n_in, const double *W,
                                    const double *b );
cmd_status cmd_nnp_model_set_atom_energy( cmd_nnp_model *m, int elem,
                                          double e );
/* n2p2 model directory: input.nn, scaling.data, weights.%03d.data (by
 * atomic number), as ForceNNP::init_coeff loads it (:48-65). */
cmd_status cmd_nnp_model_load_n2p2( const char *dir, int nelem,
                                    const char *const *names,
                                    const int *atomic_numbers,
                                    cmd_nnp_model **out );
/* Seeded two-element benchmark model (no reference counterpart): 30 radial
 * + 40 angular functions per element, network 70-25-25-1. */
cmd_status cmd_nnp_model_synthetic( uint64_t seed, double rc,
                                    cmd_nnp_model **out );
cmd_status cmd_nnp_model_info( const cmd_nnp_model *m, int elem, int *nsf,
                               int *n_hidden, int *hidden,
                               double *atom_energy );
/* params: type, e1, e2, eta, rs, lambda, zeta, rc, gmin, gmax */
cmd_status cmd_nnp_model_get_sf( const cmd_nnp_model *m, int elem, int k,
                                 double params[10] );
cmd_status cmd_nnp_model_get_layer( const cmd_nnp_model *m, int elem,
                                    int layer, double *W, double *b );
cmd_status cmd_force_nnp_create( const cmd_nnp_model *m, cmd_force **out );
/* Set every scaling range to the min/max over the current frame's atoms. */
cmd_status cmd_force_nnp_fit_scaling( cmd_force *f, cmd_system *s,
                                      cmd_neigh *n );
cmd_status cmd_force_nnp_model( const cmd_force *f, const cmd_nnp_model **m );
/* Work of the last compute (no reference counterpart: the measurement of
 * the NNP force against the FP32 roofline).  out[0] neighbours inside rc,
 * [1] neighbour pairs with their third side inside rc, [2] FP32 flops of
 * the symmetry functions, [3] of the force chain, [4] of the network
 * layers (forward and backward), [5] atoms; flops under the per-term model
 * of the engine's work model (add, multiply 1, fused multiply-add 2,
 * transcendental 1). */
cmd_status cmd_force_nnp_work( cmd_force *f, double out[6] );
/* Per-atom G, dE/dG (ld floats per atom) and atomic energies of the last
 * compute (System_NNP's G, dEdG and E, system_nnp_1aosoa.h:19-57). */
cmd_status cmd_force_nnp_atomic( cmd_force *f, cmd_system *s, int ld,
                                 float *G, float *dEdG, float *Ei );
/* Jittered simple-cubic lattice for the NNP benchmark (no reference
 * counterpart): nx*ny*nz sites at spacing s, each coordinate displaced by
 * U(-jitter, jitter)*s, element 1 on round(n * frac1) sites chosen by a
 * seeded shuffle, element 0 elsewhere.  Two-call protocol (x == NULL). */
cmd_status cmd_nnp_lattice_create( int nx, int ny, int nz, double s,
                                   double jitter, double frac1, uint64_t seed,
                                   int *n, double *x, int *type, int *id );

/* ---- thermo ----------------------------------------------------------------
 * Temperature / KinE (property_temperature_impl.h:55-77,
 * property_kine_impl.h:55-76). */
cmd_status cmd_thermo_temperature( cmd_system *s, double *T );
cmd_status cmd_thermo_kinetic( cmd_system *s, double *ke );
/* No reference counterpart: cmd_thermo_kinetic without any device
 * allocation, for callers that cannot allocate (HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  Its partial sums live in
 * a fixed per-device array (1 MiB); at most 8,388,608 local atoms, more is
 * CMD_ERR_CAPACITY.  Same kernels and result bits as cmd_thermo_kinetic;
 * calls that overlap on one device race on the array. */
cmd_status cmd_thermo_kinetic_noalloc( cmd_system *s, double *ke );

/* ---- run parameters and input deck ----------------------------------------
 * InputFile<System> (inputFile.h:150-275; inputFile_impl.h:125-534). */
typedef struct cmd_params
{
    int units;               /* 0 real, 1 lj, 2 metal (types.h:52-57) */
    int lattice_style;       /* 0 sc, 1 fcc */
    double lattice_constant; /* converted: fcc + lj gives (4/rho)^(1/3) */
    double lattice_offset[3];
    double region_lo[3], region_hi[3]; /* lattice units */
    int ntypes;
    double mass[CMD_MAX_TYPES];
    double temp_target;
    int temp_seed;
    int vel_gaussian; /* 0 loop-geom uniform (reference), 1 gaussian */
    double force_cutoff, skin;
    int every;           /* neigh_modify every */
    int max_neigh_guess; /* neigh_modify one */
    double comm_ghost_cutoff;
    int half;            /* newton on */
    int nsteps, thermo_rate;
    double dt, mvv2e, boltz;
    int ncoeff;
    int coeff_i[16], coeff_j[16];
    double coeff_eps[16], coeff_sigma[16], coeff_cut[16];
    int force_type; /* 0 lj/cut, 2 nnp (types.h:80-85) */
    char nnp_dir[256];
    int atom_style_charge; /* atom_style charge (else atomic) */
    char data_in[256];     /* read_data file ("" = lattice) */
    char data_out[256];    /* write_data file at the end of each run */
    /* Named region blocks in deck order (inputFile_impl.h:241-266) and the
     * create_atoms type of each (1-based, 0 = no atoms; :271-297).  With no
     * named region, region_lo/hi is one block of type 1.  region_lo/hi
     * always hold the bounding box of the blocks. */
    int nregions;
    char region_name[CMD_MAX_REGIONS][32];
    double region_block[CMD_MAX_REGIONS][6]; /* xlo xhi ylo yhi zlo zhi */
    int region_type[CMD_MAX_REGIONS];
    /* velocity create per type (type_to_temperature, :396-410): type 1 (or
     * 'all') is temp_target/temp_seed; type k > 1 is type_temp[k-1] and
     * type_seed[k-1]; a type never named keeps temperature 0, seed 0. */
    double type_temp[CMD_MAX_TYPES];
    int type_seed[CMD_MAX_TYPES];
    double vacuum_rate; /* --vacuum: box high corner scaled (inputCL.cpp:243-250); 1 = off */
} cmd_params;

/* Parse a LAMMPS-subset deck (inputFile_impl.h:141-533).  Unknown or
 * unsupported keywords return CMD_ERR_IO with the reference's message. */
cmd_status cmd_params_default( cmd_params *p );
cmd_status cmd_params_parse_deck( const char *path, int force_iteration,
                                  cmd_params *p );

/* Initial state on the host (inputFile_impl.h:536-865): lattice sites,
 * ids, types and velocities.  Two-call protocol: x == NULL returns the count. */
cmd_status cmd_lattice_create( const cmd_params *p, const cmd_domain *d,
                               int *n, double *x, double *v, int *type,
                               int *id );

/* LAMMPS data files (read_data.h:100-385, restated).  Header keywords
 * "atoms", "atom types", "xlo xhi", "ylo yhi", "zlo zhi"; sections Atoms
 * (atomic: id type x y z; charge: id type q x y z), Velocities (id vx vy vz),
 * Masses (type mass) and Pair Coeffs (skipped, as the reference does).
 * Types in the file are 1-based and returned 0-based.  Atoms outside
 * [lo, hi) are dropped like the reference's single-rank read, which then
 * reports a wrong atom count as an error.  Two-call protocol: with x == NULL
 * only n, ntypes, lo and hi are returned.  mass (ntypes entries, may be
 * NULL) gets the Masses section or is left unchanged; q may be NULL. */
cmd_status cmd_data_read( const char *path, int charge_style, int *n, int *ntypes,
                          double lo[3], double hi[3], double *x, double *v,
                          double *q, int *type, int *id, double *mass );
/* write_data (read_data.h:387-422): header, "Atoms # atomic" and
 * "Velocities", with every value at full double precision (the reference
 * prints 6 significant digits) and positions wrapped into [lo, hi), so the
 * file reads back whole at any step (the reference writes them unwrapped). */
cmd_status cmd_data_write( const char *path, int n, int ntypes, const double lo[3],
                           const double hi[3], const double *x, const double *v,
                           const int *type, const int *id );

/* ---- timestep loop ---------------------------------------------------------
 * CbnMD<System,Neighbor>::init / run (cabanamd_impl.h:70-432).  run() fuses
 * the steps on the device: force + final integrate + next initial integrate
 * (with wrap before a rebuild) in one kernel, then ghost refresh or the
 * rebuild sequence. */
typedef struct cmd_md cmd_md;

cmd_status cmd_md_create( const cmd_params *p, cmd_md **out );
cmd_status cmd_md_destroy( cmd_md *md );
cmd_status cmd_md_set_precision( cmd_md *md, int prec );
cmd_status cmd_md_set_lanes( cmd_md *md, int lanes );
/* Tile height (cells) of the tiled list when lanes are fixed (no reference
 * counterpart; for tuning): 0 keeps the default. */
cmd_status cmd_md_set_tile( cmd_md *md, int height );
cmd_status cmd_md_set_tile_width( cmd_md *md, int width );
/* Ghost refresh fused into the step (no reference counterpart; on by
 * default): on tiled full lists the fused step also writes the next step's
 * sorted ghost positions, in place of a separate update_halo pass.  Off
 * gives bit-identical results through the separate pass. */
cmd_status cmd_md_set_ghost_fusion( cmd_md *md, int on );
/* Before init (no reference counterpart): an NNP model to use instead of
 * loading pair_style nnp's directory, and an initial state (atoms and box)
 * to use instead of the deck's lattice and velocities. */
cmd_status cmd_md_set_nnp_model( cmd_md *md, const cmd_nnp_model *m );
cmd_status cmd_md_set_initial_state( cmd_md *md, int n, const double *x,
                                     const double *v, const int *type,
                                     const int *id, const double lo[3],
                                     const double hi[3] );
cmd_status cmd_md_init( cmd_md *md );
cmd_status cmd_md_run( cmd_md *md, int nsteps );
cmd_status cmd_md_system( cmd_md *md, cmd_system **s );
cmd_status cmd_md_neigh( cmd_md *md, cmd_neigh **n );
cmd_status cmd_md_force( cmd_md *md, cmd_force **f );
/* Thermo rows (step, T, PE/N, (PE+KE)/N, virial W); returns row count in
 * *rows_out and copies min(max_rows, count) rows (5 doubles each). */
cmd_status cmd_md_thermo( const cmd_md *md, int max_rows, double *rows,
                          int *rows_out );
/* Binary snapshot of the owned atoms (CbnMD::dump_binary,
 * cabanamd_impl.h:434-481) in a fixed format: file
 * <dir>/output.<step as %010d>.000 holding the 8 bytes "CMDDUMP1", int32 n,
 * int32 step, then per atom in atom-id order: int32 id, int32 type, and
 * doubles q (0), x[3], v[3], f[3].  f is the last computed force. */
cmd_status cmd_md_dump( cmd_md *md, const char *dir, int step );
/* CbnMD::check_correctness (cabanamd_impl.h:485-642): compare the current
 * owned atoms with the snapshot of `step` in refdir, matched by atom id.
 * out[6] = sqrt(sum |dx|^2), max |dx_k|, the same for v and f.  When file is
 * not NULL the reference's line "step dr maxdr dv maxdv df maxdf" is
 * appended (the file is started with its header at step 0). */
cmd_status cmd_md_check_correctness( cmd_md *md, const char *refdir, int step,
                                     const char *file, double out[6] );
/* Phase times in seconds (CbnMD::run timers, cabanamd_impl.h:273-282):
 * force, neigh, comm, integrate, other, total. */
cmd_status cmd_md_timers( const cmd_md *md, double out[6] );
/* Per-launch HIP-event timing of the fused force kernel on the system's
 * stream: enable before run(); returns launches and mean ms.  enable = 1
 * times every phase of every step (the phase timers); enable = k > 1 times
 * one force phase in k and no other phase, which keeps the events' markers
 * (a few microseconds of idle device each) out of most steps. */
cmd_status cmd_md_profile( cmd_md *md, int enable );
cmd_status cmd_md_kernel_stats( const cmd_md *md, int *launches,
                                double *mean_ms );
/* The list layout init chose (lanes per atom, tile width in columns and
 * height in cells, 0 for the global slab) and the host seconds init spent
 * choosing it; no
 * reference counterpart (the reference's layout is fixed at compile time,
 * neighbor_verlet.h:30-32). */
cmd_status cmd_md_tune_info( const cmd_md *md, int *lanes, int *tile_width, int *tile,
                             double *seconds );

/* ---- batched ensemble -------------------------------------------------------
 * No reference counterpart: CabanaMD steps one system per process.  A batch
 * steps many independent periodic LJ systems of uneven size with the same
 * launches (owned atoms of all systems in one flat range, ghosts after them),
 * each following CbnMD::run's step order (cabanamd_impl.h:285-399) with the
 * batch's shared parameters (cmd_params: pair_coeff, cutoff, skin, every,
 * dt, units, masses, thermo_rate).  Full lists only.  A system's trajectory
 * does not depend on the others in the batch.  Systems added or removed
 * join or leave at the start of the next cmd_batch_run that begins on a
 * rebuild step (step % every == 0). */
typedef struct cmd_batch cmd_batch;

cmd_status cmd_batch_create( const cmd_params *p, cmd_batch **out );
cmd_status cmd_batch_destroy( cmd_batch *b );
cmd_status cmd_batch_set_precision( cmd_batch *b, int prec );
/* A batch created with force_type 2 runs NNP systems (ForceNNP, unfused
 * step order) with this model; set before adding systems.  Thermo PE is
 * ForceNNP::compute_energy's per system (force_nnp_cabana_neigh_impl.h:
 * 101-130). */
cmd_status cmd_batch_set_nnp_model( cmd_batch *b, const cmd_nnp_model *m );
/* Lanes per atom of the list layout (1, 2, 4, 8); 0 = time each layout on
 * the first systems and keep the fastest.  Fixed once systems run. */
cmd_status cmd_batch_set_lanes( cmd_batch *b, int lanes );
/* HIP graphs for the plain LJ steps between rebuilds (on by default): two
 * consecutive steps (ghost refresh and the fused step, each way of the
 * position ping-pong) are captured once per rebuild and replayed.  No
 * reference counterpart; off runs the same kernels launched one by one.
 * cmd_batch_graph_steps counts the steps that graph replays ran. */
cmd_status cmd_batch_set_graphs( cmd_batch *b, int on );
cmd_status cmd_batch_graph_steps( const cmd_batch *b, long long *steps );
/* Queue a system: n atoms (x, v as n x 3; v may be NULL), types 0-based,
 * ids (NULL = 1..n) and its box [lo, hi), each side longer than twice the
 * ghost depth (cutoff + skin).  *handle names it in later calls. */
cmd_status cmd_batch_add( cmd_batch *b, int n, const double *x, const double *v,
                          const int *type, const int *id, const double lo[3],
                          const double hi[3], int *handle );
cmd_status cmd_batch_remove( cmd_batch *b, int handle );
/* Advance every running system by nsteps; thermo rows every thermo_rate
 * steps of the batch and at the last step. */
cmd_status cmd_batch_run( cmd_batch *b, int nsteps );
cmd_status cmd_batch_counts( const cmd_batch *b, int *nsys, int64_t *natoms,
                             int *step, int *lanes );
/* state: 0 unknown, 1 waiting to join, 2 running, 3 leaving; local_step is
 * the number of steps the system has run. */
cmd_status cmd_batch_system_info( const cmd_batch *b, int handle, int *n,
                                  int *local_step, int *state );
/* The system's atoms in their current (cell-sorted) order; any pointer may
 * be NULL.  f holds the forces of the last step. */
cmd_status cmd_batch_download( cmd_batch *b, int handle, double *x, double *v,
                               double *f, int *type, int *id );
/* Rows (local step, T, PE/N, (PE+KE)/N); a row at join (step 0), then as
 * cmd_batch_run takes them. */
cmd_status cmd_batch_thermo( const cmd_batch *b, int handle, int max_rows,
                             double *rows, int *rows_out );
/* Host seconds in rebuilds, step loops, thermo and joins/leaves. */
cmd_status cmd_batch_timers( const cmd_batch *b, double out[4] );

/* ---- MD request server ------------------------------------------------------
 * No reference counterpart.  Serves independent MD jobs (LJ, and NNP when
 * nnp parameters and a model are given) with continuous batching: two
 * device-resident batches (one per force) step together on their own
 * streams from their own host threads; arriving jobs are admitted in
 * arrival order while their estimated device footprint fits hbm_budget and
 * fewer than max_jobs run, join at the next rebuild boundary, and leave
 * with their results copied out the step they finish.  A job's result does
 * not depend on the other jobs.  Both parameter sets must share `every`. */
typedef struct cmd_server cmd_server;

cmd_status cmd_server_create( const cmd_params *lj, const cmd_params *nnp,
                              const cmd_nnp_model *model, int64_t hbm_budget,
                              int max_jobs, cmd_server **out );
cmd_status cmd_server_destroy( cmd_server *s );
/* Precision of both batches and lanes of the LJ list (0 = autotune). */
cmd_status cmd_server_configure( cmd_server *s, int prec, int lanes );
/* Seconds since the server was created (the clock of arrival times). */
cmd_status cmd_server_clock( const cmd_server *s, double *t );
/* A job: kind 0 (LJ) or 2 (NNP), n atoms (x, v as n x 3; v may be NULL),
 * types, ids (NULL = 1..n), box [lo, hi), nsteps, and the arrival time on
 * the server clock (non-decreasing over submissions).  *job indexes it. */
cmd_status cmd_server_submit( cmd_server *s, int kind, int n, const double *x,
                              const double *v, const int *type, const int *id,
                              const double lo[3], const double hi[3],
                              int nsteps, double arrival, int *job );
/* One round: admit arrived jobs, advance one chunk, retire finished jobs. */
cmd_status cmd_server_poll( cmd_server *s, int *queued, int *running,
                            int *done );
/* Poll until every submitted job is done (waiting for arrivals). */
cmd_status cmd_server_drain( cmd_server *s, double timeout_s );
/* state 1 queued, 2 running, 3 done; times = arrival, admission, finish. */
cmd_status cmd_server_job( const cmd_server *s, int job, int *state,
                           double times[3], int *n );
/* Final positions and velocities ordered by atom id, the ids, and thermo
 * rows (local step, T, PE/N, (PE+KE)/N). */
cmd_status cmd_server_result( const cmd_server *s, int job, double *x,
                              double *v, int *id, int max_rows, double *thermo,
                              int *rows );
/* out: finished atom-steps, seconds stepping, rounds, bytes charged, running
 * jobs, finished jobs. */
cmd_status cmd_server_stats( const cmd_server *s, double out[6] );

#ifdef __cplusplus
}
#endif
#endif
