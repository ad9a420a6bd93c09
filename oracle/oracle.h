/*
 * oracle.h -- FP64 CPU restatement of the CabanaMD timestep path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP engine
 * (cabanamd_amd/) is compared against, and the CPU baseline that bench.py
 * times.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load it.  The product never links, loads or calls it.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the CabanaMD source tree).  Functions whose reference behaviour lives in the
 * external Cabana library (LinkedCellList, VerletList, Halo) restate its
 * published algorithm; see the parity notes next to each declaration.
 *
 * Conventions: positions/velocities/forces are row-major (n,3) doubles, types
 * are 0-based ints, neighbour lists are CSR (offsets[n_local+1], idx[nnz]).
 * All FP64 arithmetic is compiled with -ffp-contract=off so that every
 * expression is evaluated as written (separate multiply and add).
 */
#ifndef CABANAMD_ORACLE_H
#define CABANAMD_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- domain (system.h:140-205 + Cabana::Grid uniform mesh) ---------------- */
typedef struct orc_domain
{
    double low[3], high[3];      /* global mesh corners */
    double L[3];                 /* global_mesh_{x,y,z} = high - low */
    double cell[3];              /* global mesh cell size, 100 cells per dim */
    double local_lo[3], local_hi[3], local_mesh[3];
    double ghost_lo[3], ghost_hi[3];
    int halo_width;              /* in mesh cells */
} orc_domain;

void orc_create_domain( const double low[3], const double high[3],
                        double ghost_cutoff, orc_domain *d );

/* ---- LAMMPS loop-geom velocity RNG (inputFile.h:66-148) ------------------- */
int orc_rng_reset( int ibase, const double coord[3] ); /* returns the seed */
double orc_rng_uniform( int *seed );
double orc_rng_gaussian( int *seed );

/* ---- initial state (inputFile_impl.h:536-865) ----------------------------- */
typedef struct orc_params
{
    int units;                 /* 0 real, 1 lj, 2 metal (types.h:52-57) */
    int lattice_style;         /* 0 sc, 1 fcc */
    double lattice_constant;   /* already converted (fcc+lj: (4/rho)^(1/3)) */
    double lattice_offset[3];
    double region_lo[3], region_hi[3]; /* region block, lattice units */
    int ntypes;
    double mass[8];
    double temp_target;
    int temp_seed;
    int vel_gaussian;          /* 0: uniform loop geom (reference); 1: gaussian */
    double force_cutoff, skin;
    int every;                 /* neigh_modify every (comm_exchange_rate) */
    int max_neigh_guess;
    double comm_ghost_cutoff;
    int half;                  /* newton on -> half list */
    int nsteps, thermo_rate;
    double dt, mvv2e, boltz;
    int ncoeff;
    int coeff_i[16], coeff_j[16]; /* 1-based type ids as in pair_coeff */
    double coeff_eps[16], coeff_sigma[16], coeff_cut[16];
    /* named region blocks in deck order and their create_atoms types
     * (1-based, 0 none); nregions == 0: region_lo/hi is one block of type 1
     * (inputFile_impl.h:241-297) */
    int nregions;
    char region_name[16][32];
    double region_block[16][6]; /* xlo xhi ylo yhi zlo zhi, lattice units */
    int region_type[16];
    /* velocity create of types 2..8 (type 1: temp_target/temp_seed),
     * inputFile_impl.h:396-410 */
    double type_temp[8];
    int type_seed[8];
} orc_params;

/* The first n values of the C library's rand() after srand(1) (glibc's
 * TYPE_3 additive feedback generator), which the reference's lattice draws
 * types with (inputFile_impl.h:651,766). */
void orc_crand( int n, int *out );

/* Count atoms the lattice would create (first pass), then fill. */
int orc_lattice_count( const orc_params *p, const orc_domain *d );
int orc_lattice_fill( const orc_params *p, const orc_domain *d, double *x,
                      int *type, int *id );
/* Velocities: RNG per atom, zero momentum, rescale to temp_target. */
void orc_velocity_create( const orc_params *p, int n, const double *x,
                          const int *type, double *v );

/* ---- integrator (integrator_nve.h:91-110, integrator_nve_impl.h:49-83) ---- */
void orc_nve_initial( int n, double *x, double *v, const double *f,
                      const int *type, const double *mass, double dt,
                      double mvv2e );
void orc_nve_final( int n, double *v, const double *f, const int *type,
                    const double *mass, double dt, double mvv2e );

/* ---- periodic wrap (comm_mpi.h:141-168) ----------------------------------- */
void orc_wrap( int n, double *x, const double L[3] );

/* ---- binning (binning_cabana_impl.h:57-113 + Cabana LinkedCellList) ------- */
/* Returns number of grid cells; perm[k] = old index of the atom at new slot k
 * (stable by old index within a cell).  cell_of (optional) gets each sorted
 * atom's cardinal cell id. */
int orc_bin_sort( int n, const double *x, const orc_domain *d,
                  const double delta_in[3], int halo_depth, int *perm,
                  int *cell_of );
/* The LinkedCellList grid create_binning sorts with: lower corner, reciprocal
 * cell size and cell counts (floor-based cell index, top-edge clamp). */
void orc_bin_grid( const orc_domain *d, const double delta_in[3],
                   int halo_depth, double gmin[3], double rd[3], int ng[3] );
/* Apply perm to x (n,3), v (n,3), type, id in place. */
void orc_permute( int n, const int *perm, double *x, double *v, int *type,
                  int *id );

/* ---- ghosts (comm_mpi_impl.h:280-441, comm_mpi.h:240-353) ----------------- */
/* Appends ghosts after n_local in x/type/id (capacity cap). send_idx gets the
 * concatenated per-phase source lists (capacity cap), counts[6] their sizes.
 * Returns n_ghost, or -1 if cap is too small. */
int orc_ghost_build( int n_local, int cap, double *x, int *type, int *id,
                     const orc_domain *d, double depth, int *send_idx,
                     int counts[6] );
void orc_ghost_refresh( int n_local, double *x, const orc_domain *d,
                        const int *send_idx, const int counts[6] );
void orc_ghost_fold( int n_local, double *f, const int *send_idx,
                     const int counts[6] );

/* ---- Verlet lists (neighbor_verlet.h:43-62 + Cabana VerletList) ----------- */
/* Rows for [0,n_local), candidates [0,n_total).  Accept j != i and
 * rsq <= rc^2 (rsq = dx*dx + dy*dy + dz*dz, dx = x_i - x_j), plus for half
 * lists the coordinate-lexicographic discriminator.  Two-call protocol:
 * call with idx == NULL to get nnz and offsets, then with idx to fill. */
int64_t orc_neigh_build( int n_local, int n_total, const double *x, double rc,
                         int half, int64_t *offsets, int *idx );
/* Same accept rule and row order, candidate search on a given grid
 * (floor_mode 1: cell = floor((x - min) * rd) with the top-edge clamp, as the
 * binning grid; 0: truncation).  The grid cells must be >= rc wide and hold
 * every atom. */
int64_t orc_neigh_build_grid( int n_local, int n_total, const double *x,
                              double rc, int half, const double gmin[3],
                              const double rd[3], const int ng[3],
                              int floor_mode, int64_t *offsets, int *idx );
/* N^2 brute force with the same accept rule (tstNeighbor.hpp:76-141). */
int64_t orc_neigh_brute( int n_local, int n_total, const double *x, double rc,
                         int half, int64_t *offsets, int *idx );

/* ---- LJ (force_lj_cabana_neigh_impl.h:62-377) ----------------------------- */
void orc_lj_coeff( int ntypes, int ncoeff, const int *ci, const int *cj,
                   const double *eps, const double *sigma, const double *cut,
                   double *lj1, double *lj2, double *cutsq );
void orc_lj_force( int n_local, int n_total, const double *x, const int *type,
                   int ntypes, const double *lj1, const double *lj2,
                   const double *cutsq, const int64_t *offsets,
                   const int *idx, int half, double *f );
double orc_lj_energy( int n_local, const double *x, const int *type,
                      int ntypes, const double *lj1, const double *lj2,
                      const double *cutsq, const int64_t *offsets,
                      const int *idx, int half );
/* Pair virial sum_{pairs} r.F (the reference has none); full lists count each
 * pair from both sides with weight 1/2, half lists with weight 1. */
double orc_lj_virial( int n_local, const double *x, const int *type,
                      int ntypes, const double *lj1, const double *lj2,
                      const double *cutsq, const int64_t *offsets,
                      const int *idx, int half );

/* ---- thermo (property_temperature_impl.h:55-77, property_kine_impl.h:55-76)*/
double orc_temperature( int n_local, int n_global, const double *v,
                        const int *type, const double *mass, double mvv2e,
                        double boltz );
double orc_kinetic_energy( int n_local, const double *v, const int *type,
                           const double *mass, double mvv2e );

/* ---- whole timestep loop (cabanamd_impl.h:70-432) ------------------------- */
typedef struct orc_md orc_md;
orc_md *orc_md_create( const orc_params *p );
void orc_md_destroy( orc_md *md );
/* Runs init() (lattice, wrap, sort, ghosts, list, step-0 force, thermo). */
int orc_md_init( orc_md *md );
/* Runs nsteps of run(); thermo rows are appended every thermo_rate steps. */
int orc_md_run( orc_md *md, int nsteps );
int orc_md_n_local( const orc_md *md );
int orc_md_n_ghost( const orc_md *md );
/* Copies out the current arrays (any pointer may be NULL). */
void orc_md_get( const orc_md *md, double *x, double *v, double *f, int *type,
                 int *id );
/* Thermo rows: step, T, PE/N, (PE+KE)/N.  Returns number of rows. */
int orc_md_thermo( const orc_md *md, int max_rows, double *rows );
/* Phase timers (s): force, neigh, comm, integrate, other, total. */
void orc_md_timers( const orc_md *md, double out[6] );
int64_t orc_md_nnz( const orc_md *md );

/* ---- NNP (force_nnp_cabana_neigh_impl.h:41-130 + n2p2 definitions) -------
 * See nnp.cpp for the symmetry-function, scaling and network definitions.
 * Elements are 0-based; e1 = e2 = -1 on a type-3 function accepts any pair. */
typedef struct orc_nnp orc_nnp;
orc_nnp *orc_nnp_create( int nelem, int nlayers, const int *nodes_hidden,
                         const char *activations );
void orc_nnp_destroy( orc_nnp *m );
void orc_nnp_set_scaling_range( orc_nnp *m, int scale, double smin, double smax );
int orc_nnp_add_sf( orc_nnp *m, int elem, int type, int e1, int e2, double eta,
                    double rs, double lambda, double zeta, double rc,
                    double gmin, double gmax );
void orc_nnp_set_layer( orc_nnp *m, int elem, int layer, int n_out, int n_in,
                        const double *W, const double *b );
void orc_nnp_set_atom_energy( orc_nnp *m, int elem, double e );
int orc_nnp_nsf( const orc_nnp *m, int elem );
void orc_nnp_symmetry_functions( const orc_nnp *m, int n_local, const double *x,
                                 const int *type, const int64_t *offsets,
                                 const int *idx, int ld, double *G );
double orc_nnp_compute( const orc_nnp *m, int n_local, int n_total,
                        const double *x, const int *type,
                        const int64_t *offsets, const int *idx, int ld,
                        double *G, double *dEdG, double *Ei, double *f );
double orc_nnp_reference_energy( const orc_nnp *m, int n_local, double sum_ei );

#ifdef __cplusplus
}
#endif
#endif
