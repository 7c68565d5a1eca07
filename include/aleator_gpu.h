/*
 * aleator_gpu.h -- C ABI of the MI355X (gfx950) adsorption hot path.
 *
 * The three reference operations this library replaces, each behind plain
 * pointers and sizes so that any host (C++, Python ctypes, ...) can bind it:
 *
 *   framework energy grid   FrameworkEnergyGrid::FrameworkEnergyGrid
 *                           (src/engines/monte_carlo/framework_energy_grid.cc:193-249),
 *                           extended with a Coulomb (Ewald real + reciprocal) table;
 *   Widom insertion         MonteCarloEngine::widomInsertionHenryCoefficient
 *                           (src/engines/monte_carlo/monte_carlo_engine.cc:850-917);
 *   S(k) incremental state  forcefield::EwaldIncrementalState
 *                           (src/forcefield/electrostatics/ewald_incremental_state.hpp:70-171).
 *
 * Conventions follow the reference exactly: lattice rows are the lattice
 * vectors in Angstrom (core/geometry/lattice.hpp:25-26), energies are in K,
 * charges in e, grid node (ix,iy,iz) sits at fractional (ix/nx, iy/ny, iz/nz)
 * and tables are flattened (iz*ny + iy)*nx + ix
 * (framework_energy_grid.hpp:185-194).
 *
 * Errors: every entry point returns an aleator_status; the message of the last
 * failure on the calling thread is available from aleator_gpu_last_error().
 * The status codes map one-to-one onto the reference's exception types
 * (std::invalid_argument, std::out_of_range, std::runtime_error,
 * aleator::NotImplemented; see src/core/exceptions.hpp:13-17).
 *
 * Threading: a context, and every object created from it, is used by one host
 * thread at a time (the reference is single-threaded by contract,
 * src/io/config.cc:131-135).
 */
#ifndef ALEATOR_GPU_H
#define ALEATOR_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    ALEATOR_OK = 0,
    ALEATOR_INVALID_ARGUMENT = 1, /* std::invalid_argument */
    ALEATOR_OUT_OF_RANGE = 2,     /* std::out_of_range */
    ALEATOR_RUNTIME_ERROR = 3,    /* std::runtime_error */
    ALEATOR_NOT_IMPLEMENTED = 4,  /* aleator::NotImplemented */
    ALEATOR_DEVICE_ERROR = 5,     /* HIP runtime failure (no reference counterpart) */
    ALEATOR_PARSE_ERROR = 6       /* io::CifParseError (io/cif_error.hpp:8-16) */
} aleator_status;

/* LennardJonesTruncation (pairwise/lennard_jones.hpp:21-38) */
typedef enum { ALEATOR_LJ_TRUNCATED = 0, ALEATOR_LJ_SHIFTED = 1, ALEATOR_LJ_LINEAR_FORCE_SHIFTED = 2 } aleator_lj_truncation;
/* MixingRule (pairwise/lennard_jones.hpp:42-50) */
typedef enum { ALEATOR_MIX_LORENTZ_BERTHELOT = 0, ALEATOR_MIX_GEOMETRIC = 1 } aleator_mixing_rule;
/* Arithmetic of the pair terms.  FP64: every term in double.  MIXED: geometry and
 * pair terms in FP32, accumulation in FP64, reciprocal space always FP64. */
typedef enum { ALEATOR_PRECISION_FP64 = 0, ALEATOR_PRECISION_MIXED = 1 } aleator_precision;

typedef struct aleator_ctx aleator_ctx;
typedef struct aleator_framework aleator_framework;
typedef struct aleator_ewald_state aleator_ewald_state;

/* core::Lattice matrix: m[i] = lattice vector a_i (lattice.hpp:21-30). */
typedef struct {
    double m[3][3];
} aleator_lattice;

/* core::ParticleData subset the hot path reads (particle_data.hpp:20-99): host pointers. */
typedef struct {
    const double* x;
    const double* y;
    const double* z;
    const double* charge;    /* may be NULL: all zero */
    const uint32_t* species; /* index into aleator_lj_params */
    uint64_t n;
} aleator_particles;

/* LennardJones constructor arguments (lennard_jones.hpp:64-66). */
typedef struct {
    const double* epsilon; /* K, per species */
    const double* sigma;   /* Angstrom, per species */
    uint32_t n_species;
    double cutoff;
    int32_t truncation; /* aleator_lj_truncation */
    int32_t mixing;     /* aleator_mixing_rule */
} aleator_lj_params;

/* Ewald constructor arguments (ewald.hpp:76-77); k-set is the reference's
 * spherical rule n0^2+n1^2+n2^2 < kmax^2+2, n != 0 (ewald.cc:137-161). */
typedef struct {
    double alpha;
    double real_space_cutoff;
    int32_t kmax;
} aleator_ewald_params;

/* forcefield::ChargedSite (ewald_incremental_state.hpp:17-22). */
typedef struct {
    double x, y, z;
    double charge;
} aleator_charged_site;

/* engines::MoleculeSite (molecule_species.hpp:14-21), mass omitted (unused on the path). */
typedef struct {
    double x, y, z;
    uint32_t lj_species;
    double charge;
} aleator_molecule_site;

/* ---- context --------------------------------------------------------- */
const char* aleator_gpu_last_error(void);
/* No reference counterpart: binds the calling thread to `device` and holds its stream. */
aleator_status aleator_gpu_init(int32_t device, aleator_ctx** out);
aleator_status aleator_gpu_destroy(aleator_ctx* ctx);
/* The HIP stream (hipStream_t) every launch of this context goes to. */
void* aleator_gpu_stream(aleator_ctx* ctx);
/* No reference counterpart (profiling): synchronises the stream, returns the summed
 * HIP-event time (ms) and count of the dominant-kernel launches recorded since the
 * previous call (the grid tile kernel of each grid build, the pair kernels of each
 * Widom batch), clears them, and enables (non-zero) or disables further recording. */
aleator_status aleator_gpu_timing(aleator_ctx* ctx, int32_t enable, double* total_ms, uint64_t* launches);

/* ---- device-resident framework --------------------------------------- */
/* Uploads the rigid framework (atoms periodic in `lattice`), precomputes the per
 * (guest, framework type) mixed LJ table (lennard_jones.cc:49-68) and, when `ewald`
 * is non-NULL, the reference k-set and the framework structure factor S(k)
 * (EwaldIncrementalState ctor, ewald_incremental_state.cc:55-83).  Validates what
 * the reference validates: positive LJ cutoff (lennard_jones.cc:39-47), Ewald
 * parameters (ewald.cc:199-212), charge neutrality when charged (ewald.cc:181-195),
 * cutoffs <= L_perp/2 (lattice.cc:286-300). */
aleator_status aleator_gpu_framework_create(aleator_ctx* ctx, const aleator_lattice* lattice,
                                            const aleator_particles* atoms,
                                            const aleator_lj_params* lj,
                                            const aleator_ewald_params* ewald,
                                            aleator_framework** out);
aleator_status aleator_gpu_framework_destroy(aleator_framework* fw);
/* Number of reciprocal terms of the framework's k-set (0 when uncharged). */
aleator_status aleator_gpu_framework_num_k(const aleator_framework* fw, uint64_t* nk);

/* ---- framework energy grid ------------------------------------------- */
/* Grid point counts n_axis = max(1, lround(L_perp(axis)/spacing))
 * (framework_energy_grid.cc:207-211). */
aleator_status aleator_gpu_grid_point_counts(const aleator_lattice* grid_lattice, double spacing,
                                             int32_t n_points[3]);
/* The analytic LJ tail correction of one insertion, off by default (the reference adds it to
 * neither its grid nor its Widom loop).  When enabled, the grid builds add to every LJ table
 * the tail of one site of that table's species, before the cap, and Widom adds to every
 * trial's dU the tail of the whole molecule: tailEnergyCorrection (lennard_jones.cc:218-257)
 * of the framework plus the inserted sites minus that of the framework alone (the same
 * quantity as aleator_lj_tail_insertion_delta). */
aleator_status aleator_gpu_framework_set_lj_tail(aleator_framework* fw, int32_t enable);
/* No reference counterpart (a NaN / Inf sentinel; the reference writes such values silently):
 * the number of non-finite table values (NaN, +-Inf) that the grid builds of this framework
 * have written since it was created, over every entry point that builds its tables (per call,
 * allocation-free and batched plans).  The values themselves are written as computed (an atom
 * exactly on a node gives an infinite Coulomb value, as in the reference); callers compare the
 * count before and after a build.  Waits for the context stream. */
aleator_status aleator_gpu_framework_nonfinite_count(aleator_framework* fw, uint64_t* count);
/* Builds one table per entry of `guest_species` (FrameworkEnergyGrid ctor body,
 * framework_energy_grid.cc:221-248): E(node) = min(sum_f pairEnergy(r, mix(guest,f)), cap).
 * When `coulomb` is non-zero and the framework is charged, one more table follows: the
 * Ewald potential per unit charge (K/e), real space kC*sum q erfc(alpha r)/r (ewald.cc:223-249)
 * plus reciprocal space, the term of EwaldIncrementalState::computeDelta linear in the
 * inserted charge (ewald_incremental_state.cc:117-121); it is not capped.
 * `grid_lattice` NULL means the framework lattice (the reference's behaviour); otherwise its
 * rows must be the framework lattice rows divided by positive integers (a primitive cell of
 * a replicated framework), and the nodes are that cell's nodes.
 * `out_host`: (n_guest + coulomb) * nz*ny*nx doubles, table-major. */
aleator_status aleator_gpu_grid_build(aleator_framework* fw, const aleator_lattice* grid_lattice,
                                      const uint32_t* guest_species, int32_t n_guest,
                                      int32_t coulomb, const int32_t n_points[3],
                                      int32_t precision, double cap, double* out_host);
/* Same, writing to device memory on the context stream; returns without synchronising. */
aleator_status aleator_gpu_grid_build_device(aleator_framework* fw,
                                             const aleator_lattice* grid_lattice,
                                             const uint32_t* guest_species, int32_t n_guest,
                                             int32_t coulomb, const int32_t n_points[3],
                                             int32_t precision, double cap, double* out_device);
/* No counterpart in the reference: for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops).  Same as aleator_gpu_grid_build_device, with the
 * build's scratch (constants, reciprocal contraction tables) in a fixed 32 MiB device array of
 * the library instead of the context's growable buffer: no hipMalloc or hipFree.  A build whose
 * scratch would exceed 32 MiB fails with ALEATOR_INVALID_ARGUMENT (the c1 three-table grid needs
 * about 5.5 MB).  All calls on one GPU share the array: calls that overlap on one GPU race. */
aleator_status aleator_gpu_grid_build_device_noalloc(aleator_framework* fw,
                                                     const aleator_lattice* grid_lattice,
                                                     const uint32_t* guest_species, int32_t n_guest,
                                                     int32_t coulomb, const int32_t n_points[3],
                                                     int32_t precision, double cap, double* out_device);

/* A reusable build of the grids of one or many frameworks (no reference counterpart as one
 * call; each framework's tables equal aleator_gpu_grid_build_device's).  Creation uploads every
 * constant once; execute only launches kernels -- one tile launch over all frameworks -- and,
 * with use_graph, replays them from a captured hipGraph.  tile: a tile-shape index in
 * [0, 4), or -1 to time every shape that fits, with every allowed number of atoms per list pass,
 * on the context stream and keep the fastest (the choice is also used by later per-call builds
 * on the context).  Per framework f:
 * grid_lattices[f] (or NULL for all: the framework lattices), n_points[3f..3f+2], and
 * out_device[f] with (n_guest + coulomb) * nodes doubles. */
typedef struct aleator_grid_plan aleator_grid_plan;
aleator_status aleator_gpu_grid_plan_create(aleator_ctx* ctx, int32_t n_frameworks,
                                            aleator_framework* const* frameworks,
                                            const aleator_lattice* grid_lattices,
                                            const uint32_t* guest_species, int32_t n_guest,
                                            int32_t coulomb, const int32_t* n_points, int32_t precision,
                                            double cap, double* const* out_device, int32_t tile,
                                            int32_t use_graph, aleator_grid_plan** out);
aleator_status aleator_gpu_grid_plan_execute(aleator_grid_plan* plan);
/* The tile shape in use and, when the plan autotuned, the median time (ms) of each shape at
 * its best atoms-per-pass (-1 where a shape was not tried); tuning_ms may be NULL, else holds
 * 4 floats. */
aleator_status aleator_gpu_grid_plan_tile(const aleator_grid_plan* plan, int32_t* tile, float* tuning_ms);
/* Framework atoms staged per pass of the tile kernel's LDS list build, tuned with the shape
 * (128; 256, 512 or 1024 for the frameworks of which at most one image of an atom can be
 * listed, the others keeping 128). */
aleator_status aleator_gpu_grid_plan_atoms_per_pass(const aleator_grid_plan* plan, int32_t* atoms);
/* No reference counterpart (profiling): degrees of the erf(alpha r)/r polynomials of the plan's
 * Coulomb tables (0 without one): `degree` for the list entries inside both cutoffs for every
 * node of a tile, `shell_degree` for the others (equal to `degree` when the list is not split). */
aleator_status aleator_gpu_grid_plan_erf_degree(const aleator_grid_plan* plan, int32_t* degree, int32_t* shell_degree);
aleator_status aleator_gpu_grid_plan_destroy(aleator_grid_plan* plan);

/* ---- energy-grid cache file, format v1 -------------------------------------------------
 * Byte-compatible with the reference's cache (framework_energy_grid.cc:31-191 key and
 * helpers, :286-313 serialize, :315-322 cacheFilePath, :357-408 loadFromCache): grids built
 * here load in the reference's engine and reference-built files load here.  The key is
 * makeFrameworkEnergyGridCacheKey's (framework_energy_grid.cc:171-191): the framework atoms
 * as given (Cartesian x, y, z and species; charges are not part of it), the lattice, every
 * species' (epsilon, sigma), the cutoff, the truncation, the guest species (sorted and
 * deduplicated here, as there) and the spacing; the mixing rule is not part of it, as in
 * the reference. */
typedef struct {
    const aleator_lattice* lattice;
    const aleator_particles* atoms;
    const aleator_lj_params* lj;
    const uint32_t* guest_species;
    int32_t n_guest;
    double spacing;
} aleator_grid_cache_key;
/* FrameworkEnergyGridCacheKey::hash (framework_energy_grid.cc:158-163): FNV-1a 64 of the key. */
aleator_status aleator_grid_cache_key_hash(const aleator_grid_cache_key* key, uint64_t* hash);
/* FrameworkEnergyGrid::cacheFilePath: directory/framework_energy_grid_v1_<hash hex>.bin. */
aleator_status aleator_grid_cache_file_path(const char* directory, const aleator_grid_cache_key* key,
                                            char* out, uint64_t capacity);
/* FrameworkEnergyGrid::serialize: tables [n_tables][nz*ny*nx] (host), one per species,
 * written in ascending species order; parent directories are created. */
aleator_status aleator_grid_cache_serialize(const char* path, const aleator_grid_cache_key* key,
                                            const int32_t n_points[3], const uint32_t* table_species,
                                            int32_t n_tables, const double* tables);
/* FrameworkEnergyGrid::loadFromCache for the file of `key` in `directory`: *hit = 0 when no
 * such file exists (a miss, status OK); ALEATOR_RUNTIME_ERROR when it exists but its version
 * or stored key differs, or it is truncated.  On a hit, n_points, n_tables and, when `tables`
 * is not NULL (capacity in doubles), table_species and the tables in ascending species order. */
aleator_status aleator_grid_cache_load(const char* directory, const aleator_grid_cache_key* key,
                                       int32_t n_points[3], uint32_t* table_species, int32_t* n_tables,
                                       double* tables, uint64_t capacity, int32_t* hit);
/* FrameworkEnergyGrid::interpolate (framework_energy_grid.cc:251-284) on the device: trilinear
 * interpolation of one table [nz][ny][nx] over the nodes (i/n) of `lattice` at n Cartesian
 * positions (device, [n][3], any periodic image); writes n doubles to out_device on the
 * context stream (asynchronous).  The reference's arithmetic, in its order. */
aleator_status aleator_gpu_grid_interpolate(aleator_ctx* ctx, const aleator_lattice* lattice,
                                            const int32_t n_points[3], const double* table_device,
                                            const double* positions_device, uint64_t n,
                                            double* out_device);
/* Same with host buffers (uploads the table and positions, synchronises). */
aleator_status aleator_gpu_grid_interpolate_host(aleator_ctx* ctx, const aleator_lattice* lattice,
                                                 const int32_t n_points[3], const double* table,
                                                 const double* positions, uint64_t n, double* out);

/* ---- Widom insertion -------------------------------------------------- */
/* Batched Widom test-particle insertion of a rigid molecule, as the reference loop
 * (monte_carlo_engine.cc:880-916): trial t takes uniform draws from
 * Philox4x32Rng(seed, stream) at call indices first_call_index + t*d + j, with d = 3
 * for one site and d = 6 otherwise (3 for the position, 3 for the Shoemake quaternion),
 * in that order.  dU = LJ + [real-space + reciprocal + self + exclusion when charged].
 * Writes per-trial dU (nullable) and sum_t exp(-dU/T); K_H = sum/(n_trials*T). */
aleator_status aleator_gpu_widom(aleator_framework* fw, const aleator_molecule_site* sites,
                                 int32_t n_sites, uint64_t seed, uint64_t stream,
                                 uint64_t first_call_index, uint64_t n_trials, double temperature,
                                 int32_t precision, double* du_host, double* sum_boltzmann);
/* Device variant: du_device (nullable) and sum_device (1 double) in device memory. */
aleator_status aleator_gpu_widom_device(aleator_framework* fw, const aleator_molecule_site* sites,
                                        int32_t n_sites, uint64_t seed, uint64_t stream,
                                        uint64_t first_call_index, uint64_t n_trials,
                                        double temperature, int32_t precision, double* du_device,
                                        double* sum_device);

/* ---- EwaldIncrementalState -------------------------------------------- */
/* Initialised from the framework's atoms (ctor, ewald_incremental_state.cc:55-83). */
aleator_status aleator_gpu_ewald_state_create(aleator_framework* fw, aleator_ewald_state** out);
aleator_status aleator_gpu_ewald_state_destroy(aleator_ewald_state* st);
/* proposeMove / proposeInsertion / proposeDeletion (ewald_incremental_state.cc:147-159):
 * n_removed == 0 is an insertion, n_added == 0 a deletion.  Never mutates the state. */
aleator_status aleator_gpu_ewald_state_propose(aleator_ewald_state* st,
                                               const aleator_charged_site* removed,
                                               int32_t n_removed,
                                               const aleator_charged_site* added, int32_t n_added,
                                               double* delta_energy);
/* commitMove / commitInsertion / commitDeletion (ewald_incremental_state.cc:161-172). */
aleator_status aleator_gpu_ewald_state_commit(aleator_ewald_state* st,
                                              const aleator_charged_site* removed,
                                              int32_t n_removed, const aleator_charged_site* added,
                                              int32_t n_added);
/* Many independent proposals against the same committed state in one launch:
 * proposal p removes removed[p*n_removed .. ] and adds added[p*n_added .. ].
 * No reference counterpart as a single call; each element equals propose(). */
aleator_status aleator_gpu_ewald_state_propose_batch(aleator_ewald_state* st, uint64_t n_proposals,
                                                     const aleator_charged_site* removed,
                                                     int32_t n_removed,
                                                     const aleator_charged_site* added,
                                                     int32_t n_added, double* delta_energy);
/* committedEnergy (ewald_incremental_state.cc:85-93). */
aleator_status aleator_gpu_ewald_state_committed_energy(aleator_ewald_state* st, double* energy);
/* resync / recomputeFromScratch (ewald_incremental_state.cc:174-247); molecules as CSR
 * (mol_offsets[n_molecules+1] into mol_sites, indices into `particles`). */
aleator_status aleator_gpu_ewald_state_resync(aleator_ewald_state* st,
                                              const aleator_particles* particles,
                                              const uint64_t* mol_offsets,
                                              const uint64_t* mol_sites, uint64_t n_molecules);
aleator_status aleator_gpu_ewald_state_recompute(aleator_ewald_state* st,
                                                 const aleator_particles* particles,
                                                 const uint64_t* mol_offsets,
                                                 const uint64_t* mol_sites, uint64_t n_molecules,
                                                 double* energy);
/* rawStateForTesting (ewald_incremental_state.cc:249-257): s_re[Nk], s_im[Nk], self,
 * exclusion, in the reference's k order; `capacity` doubles available. */
aleator_status aleator_gpu_ewald_state_raw(aleator_ewald_state* st, double* out, uint64_t capacity,
                                           uint64_t* written);

/* ---- batched grand-canonical Monte Carlo chains ------------------------ */
/* Many independent single-species GCMC chains against one device framework, each with its
 * own Philox4x32Rng(seed, stream), molecule list and S(k) state, advanced together in one
 * persistent launch.  Each chain follows MonteCarloEngine::run (monte_carlo_engine.cc:405-439)
 * with one species: translation / rotation / insertion / deletion at 25 % each, the
 * reference's draw order per move (:441-671), acceptance by gcmc_acceptance.hpp:28-101,
 * Ewald resync every 100 accepted moves (:380-403).  Energies are the exact direct sums (no
 * energy grid): LJ against framework and guests, real-space Ewald, S(k) reciprocal, self and
 * exclusion terms. */
typedef struct aleator_gcmc aleator_gcmc;
typedef struct {
    double fugacity_pa; /* reservoir fugacity (e.g. aleator_peng_robinson_fugacity) */
    double temperature;
    uint64_t seed;
    uint64_t stream;
} aleator_gcmc_chain;
typedef struct {
    int32_t n_molecules;
    int32_t error; /* 1: the chain needed more than max_molecules molecules and stopped */
    uint64_t steps;
    double sum_molecules; /* sum over steps of the molecule count after the step */
    uint64_t attempted[4]; /* translation, rotation, insertion, deletion */
    uint64_t accepted[4];
} aleator_gcmc_stats;
aleator_status aleator_gpu_gcmc_create(aleator_framework* fw, const aleator_molecule_site* sites,
                                       int32_t n_sites, const aleator_gcmc_chain* chains,
                                       int32_t n_chains, double max_translation, double max_rotation,
                                       int32_t max_molecules, aleator_gcmc** out);
/* Advances every chain by n_steps in one launch on the context stream (asynchronous).
 * trace_device (nullable): int32 [n_chains][n_steps] molecule count after each step. */
aleator_status aleator_gpu_gcmc_run(aleator_gcmc* gcmc, uint64_t n_steps, int32_t* trace_device);
/* Whether run without a trace replays a captured hipGraph of its launch (recaptured whenever the
 * launch's arguments change: steps, proposal, grid, cells) or launches directly (default 0).  The
 * chains' state lives in device memory, so replays continue them exactly as launches do. */
aleator_status aleator_gpu_gcmc_use_graph(aleator_gcmc* gcmc, int32_t enable);
aleator_status aleator_gpu_gcmc_stats(aleator_gcmc* gcmc, aleator_gcmc_stats* out /* [n_chains] */);
aleator_status aleator_gpu_gcmc_reset_stats(aleator_gcmc* gcmc);
/* Site positions of one chain in molecule order, [n_molecules][n_sites][3]; xyz may be NULL
 * to query the count. */
aleator_status aleator_gpu_gcmc_positions(aleator_gcmc* gcmc, int32_t chain, double* xyz,
                                          uint64_t capacity, int32_t* n_molecules);
/* committedEnergy of the chain's Ewald state (ewald_incremental_state.cc:85-93). */
aleator_status aleator_gpu_gcmc_electrostatic_energy(aleator_gcmc* gcmc, int32_t chain, double* energy);
/* Consistency check at every Ewald resync, the reference's drift gate
 * (test_ewald_incremental_state.cc:137-222, tolerance 1e-10) run inside the chains: before each
 * resync replaces the committed S(k), self and exclusion terms by the from-scratch ones
 * (ewald_incremental_state.cc:174-247), the kernel records |fresh - committed| /
 * max(1, |fresh|) of the Ewald energy.  Per chain [n_chains]: the largest such drift (NaN once
 * a non-finite energy was seen) and the number of resyncs it covers, since creation or
 * aleator_gpu_gcmc_reset_stats.  Uncharged chains never resync and report 0. */
aleator_status aleator_gpu_gcmc_resync_check(aleator_gcmc* gcmc, double* max_drift, uint64_t* resyncs);
/* No reference counterpart (the reference keeps no GCMC checkpoint; its RNG state is resumable
 * by construction, counter_based_rng.hpp:59-66): one chain's whole state as a flat record, to
 * resume a run or move a chain between chain sets of the same framework and species.
 * Little-endian layout, byte offsets:
 *   0 u32 magic 0x4b434741 ("AGCK")   4 u32 version 1
 *   8 i32 n_sites                    12 i32 n_molecules
 *  16 i32 nk_half (0 uncharged)      20 i32 accepted moves since the last resync
 *  24 u64 Philox call index of the next draw
 *  32 f64 self energy   40 f64 exclusion energy   48 f64 sum of molecule counts   56 u64 steps
 *  64 u64 attempted[4]  96 u64 accepted[4]  (translation, rotation, insertion, deletion)
 * 128 f64 max resync drift          136 u64 resyncs
 * 144 f64 S_re[nk_half], f64 S_im[nk_half] (the first half of the reference-order k-set),
 *     then f64 xyz[n_molecules][n_sites][3] (Cartesian, molecule order).
 * checkpoint_size gives the record's size for the chain's current molecule count.  load_chain
 * checks magic, version, site count, k-set size and max_molecules, and clears the chain's
 * capacity error. */
aleator_status aleator_gpu_gcmc_checkpoint_size(aleator_gcmc* gcmc, int32_t chain, uint64_t* bytes);
aleator_status aleator_gpu_gcmc_save_chain(aleator_gcmc* gcmc, int32_t chain, void* buffer, uint64_t capacity,
                                           uint64_t* written);
aleator_status aleator_gpu_gcmc_load_chain(aleator_gcmc* gcmc, int32_t chain, const void* buffer, uint64_t size);
aleator_status aleator_gpu_gcmc_destroy(aleator_gcmc* gcmc);
/* No reference counterpart (learned proposal): makes every later insertion draw its anchor
 * point from a piecewise-constant density instead of uniformly, and corrects both insertion
 * and deletion for it (R_ins x 1/(V p(r)), R_del x V p(r_del), on top of
 * gcmc_acceptance.hpp:28-51), so detailed balance holds with the exact energies.  The density
 * lives on nb[0] x nb[1] x nb[2] equal bins of the cell divided by reps[i] along axis i and
 * repeats over the reps[0]*reps[1]*reps[2] images (a table over the primitive cell of a
 * supercell): p(bin) = (1 - uniform_mix) * w[bin] / sum(w) + uniform_mix / n_bins, with
 * w >= 0 (row-major, axis 2 fastest).  Every bin must keep p > 0.  The anchor of a molecule is
 * the centroid of its sites, so the species' site offsets must be centred (mean 0).
 * Insertion draw order: bin (1 uniform, inverse CDF), image (1), position in the bin (3),
 * orientation as before, acceptance.  w = NULL restores uniform insertion. */
aleator_status aleator_gpu_gcmc_set_proposal(aleator_gcmc* gcmc, const double* w, const int32_t nb[3],
                                             const int32_t reps[3], double uniform_mix);
/* The frameworkEnergyGrid argument of MonteCarloEngine's constructor
 * (monte_carlo_engine.hpp:129-151, used at monte_carlo_engine.cc:341-374): from the next run on,
 * a site's framework LJ energy is its species' table interpolated at the site
 * (FrameworkEnergyGrid::interpolate, framework_energy_grid.cc:251-284) instead of the sum over
 * framework atoms; real-space Coulomb with the framework and every guest-guest term stay exact.
 * tables_device: FP64 [n_tables][n2][n1][n0] (aleator_gpu_grid_build_device's layout, built
 * on the framework's own cell) for LJ species `species[t]`; it is copied, so the caller may free
 * it after the call.  ALEATOR_OUT_OF_RANGE when a site's species has no table (the reference
 * throws std::out_of_range at the first interpolation).  tables_device = NULL restores the
 * direct scan. */
aleator_status aleator_gpu_gcmc_set_energy_grid(aleator_gcmc* gcmc, const int32_t n_points[3],
                                                const uint32_t* species, int32_t n_tables,
                                                const double* tables_device);
/* Same with the tables in host memory (a FrameworkEnergyGrid's, or a cache file's as
 * aleator_grid_cache_load returns them). */
aleator_status aleator_gpu_gcmc_set_energy_grid_host(aleator_gcmc* gcmc, const int32_t n_points[3],
                                                     const uint32_t* species, int32_t n_tables,
                                                     const double* tables);
/* Guest-guest candidates through guest cells, the device form of the engine's guest CellList
 * (MonteCarloEngine::moleculeDispersionEnergy with guestCellList_.forEachIndexNear and the
 * ascending sort, monte_carlo_engine.cc:341-374; CellList, cell_list.hpp:26-123).  Cells at
 * least the guest cutoff wide along each lattice axis; a guest site is a candidate of a trial
 * when its cell neighbours a trial site's cell, compacted in ascending site order, so the
 * pairs within the cutoff are those of the full scan.  mode 0 (the default): full scan, the
 * faster one at the occupancies measured (a filter pass costs about what the out-of-range pairs
 * it drops cost); 1: cells when every axis has at least 4, in the steps that start with at
 * least 1024 guest sites; 2: cells in every step whenever max_molecules x sites <= 4096
 * (ALEATOR_OUT_OF_RANGE above).  Takes effect at the next run. */
aleator_status aleator_gpu_gcmc_set_guest_cells(aleator_gcmc* gcmc, int32_t mode);
/* cells per lattice axis in use (0, 0, 0: full scan) */
aleator_status aleator_gpu_gcmc_guest_cells(const aleator_gcmc* gcmc, int32_t cells[3]);

/* ---- learned insertion density: Fourier-feature MLP (no reference counterpart) ----
 * 64 features (sin, cos of 2 pi k.f for 32 integer vectors k, aleator_gpu_mlp_frequencies)
 * of the fractional position f -> 4 x 256 SiLU -> 1, trained and evaluated on the device with
 * BF16 MFMA GEMMs accumulating in FP32; FP32 master weights and Adam.  Parameters are one
 * packed float vector: W1[64][256] b1[256] W2[256][256] b2 W3 b3 W4 b4 w5[256] b5[1]. */
typedef struct aleator_mlp aleator_mlp;
int64_t aleator_gpu_mlp_num_params(void);
/* frequency vectors, int8 [32][3] */
aleator_status aleator_gpu_mlp_frequencies(int8_t* out);
/* max_rows: rows per device pass (activation memory ~ 3.1 KB per row) */
aleator_status aleator_gpu_mlp_create(aleator_ctx* ctx, int32_t max_rows, aleator_mlp** out);
aleator_status aleator_gpu_mlp_destroy(aleator_mlp* mlp);
/* He-uniform weights from Philox4x32Rng(seed, 0), zero biases; resets Adam */
aleator_status aleator_gpu_mlp_init(aleator_mlp* mlp, uint64_t seed);
aleator_status aleator_gpu_mlp_set_params(aleator_mlp* mlp, const float* params /* host */);
aleator_status aleator_gpu_mlp_get_params(aleator_mlp* mlp, float* params /* host */);
/* gradient of the last loss_grad / train iteration, host */
aleator_status aleator_gpu_mlp_get_grads(aleator_mlp* mlp, float* grads);
/* out_device[i] = net(frac_device[i]) for n rows of fractional xyz (double [n][3]) */
aleator_status aleator_gpu_mlp_predict(aleator_mlp* mlp, const double* frac_device, int64_t n, float* out_device);
/* mean squared error over n <= max_rows rows and its gradient (kept on the device) */
aleator_status aleator_gpu_mlp_loss_grad(aleator_mlp* mlp, const double* frac_device, const double* target_device,
                                         int32_t n, float* loss);
/* n_iters Adam steps (lr) on minibatches of `batch` rows drawn with replacement from the
 * n_samples rows (row = floor(u n), u = Philox4x32Rng(seed, stream) at call iter*batch + r);
 * loss_trace (host, nullable) receives each iteration's minibatch loss. */
aleator_status aleator_gpu_mlp_train(aleator_mlp* mlp, const double* frac_device, const double* target_device,
                                     int64_t n_samples, int32_t batch, int32_t n_iters, float lr, uint64_t seed,
                                     uint64_t stream, float* loss_trace);
/* Whether train replays one captured hipGraph of an iteration's kernels (default 1) or launches
 * them one by one (0); both give the same results bit for bit. */
aleator_status aleator_gpu_mlp_use_graph(aleator_mlp* mlp, int32_t enable);

/* ---- host utilities of the path (no device work) ------------------------ */
/* Philox4x32Rng(seed, stream).nextUniform() at call indices first .. first+n-1
 * (core/math/counter_based_rng.cc:56-79), on the host; the device kernels draw the same values. */
aleator_status aleator_philox_uniforms(uint64_t seed, uint64_t stream, uint64_t first_call_index,
                                      uint64_t n, double* out);
/* minimumSupercellReplication (core/geometry/supercell.cc:7-20): smallest n_i >= 1 with
 * n_i * L_perp_i > 2 * cutoff. */
aleator_status aleator_supercell_replication(const aleator_lattice* lattice, double cutoff, int32_t n[3]);
/* replicateSupercell (core/geometry/supercell.cc:22-60): images in i, j, k order, each a copy
 * of `in` translated by i*a0 + j*a1 + k*a2; output arrays hold in->n*nx*ny*nz entries. */
aleator_status aleator_replicate_supercell(const aleator_particles* in, const aleator_lattice* lattice,
                                           int32_t nx, int32_t ny, int32_t nz, double* x, double* y,
                                           double* z, double* charge, uint32_t* species,
                                           aleator_lattice* out_lattice);
/* PengRobinson::fugacityPascal (engines/monte_carlo/peng_robinson.cc:10-110). */
aleator_status aleator_peng_robinson_fugacity(double critical_temperature, double critical_pressure_pa,
                                              double acentric_factor, double temperature,
                                              double pressure_pa, double* fugacity_pa);
/* LennardJones::tailEnergyCorrection (pairwise/lennard_jones.cc:218-257) of the framework plus
 * one inserted molecule minus that of the framework alone: the constant analytic tail term of
 * one insertion.  The reference's Widom loop does not add it; a caller that wants it multiplies
 * K_H by exp(-delta/T). */
aleator_status aleator_lj_tail_insertion_delta(const aleator_lj_params* lj, const aleator_lattice* lattice,
                                               const uint32_t* framework_species, uint64_t n_framework,
                                               const aleator_molecule_site* sites, int32_t n_sites,
                                               double* delta);

/* The erf kernel of the FP64 pair loops (no reference counterpart: it is how the device evaluates
 * Ewald::realSpaceParticleEnergy's erfc(alpha r)/r, ewald.cc:223-249, as 1/r - erf(alpha r)/r):
 * erf(alpha r)/r = sum_k coef[k] x^k, x = r^2 xs - 1, over r in [0, rmax], with the smallest
 * degree of 18, 24 and 28 (at least min_degree) whose dropped Chebyshev terms sum below tol.
 * coef holds 29 doubles; *degree receives the degree. */
aleator_status aleator_erf_kernel_poly(double alpha, double rmax, double tol, int32_t min_degree, double* coef,
                                       double* xs, int32_t* degree);

/* ---- structure files: CIF reader (host only) ---------------------------- */
/* io::readCif (io/structure_io.hpp:39-50, structure_io.cc:173-250): cell from _cell_length_*
 * and _cell_angle_* (a along x, b in the xy plane), symmetry operations from the first loop
 * with _symmetry_equiv_pos_as_xyz, else _space_group_symop_operation_xyz, else identity, and
 * the _atom_site loop expanded through every operation: images wrapped into [0,1), an image
 * within 1e-3 A of one already kept for that site dropped.  Atoms are emitted site by site in
 * file order, each site's images in operation order; species are numbered by first
 * appearance of each element; positions are Cartesian, masses from the element table, charge
 * 0 and occupancy 1 when absent.  Malformed input, or a file that cannot be opened, returns
 * ALEATOR_PARSE_ERROR. */
typedef struct aleator_structure aleator_structure;
aleator_status aleator_read_cif(const char* path, aleator_structure** out);
/* The same from CIF text in memory. */
aleator_status aleator_parse_cif(const char* text, uint64_t length, aleator_structure** out);
aleator_status aleator_structure_destroy(aleator_structure* structure);
aleator_status aleator_structure_size(const aleator_structure* structure, uint64_t* n_atoms,
                                      uint32_t* n_species);
aleator_status aleator_structure_lattice(const aleator_structure* structure, aleator_lattice* lattice);
/* Copies the per-atom arrays (n_atoms entries each); any pointer may be NULL to skip it
 * (StructureData::particles, ::occupancy; structure_io.hpp:18-37). */
aleator_status aleator_structure_atoms(const aleator_structure* structure, double* x, double* y,
                                       double* z, double* mass, double* charge, uint32_t* species,
                                       double* occupancy);
/* StructureData::speciesSymbols[species] and ::sourceLabel[atom]; the strings live as long
 * as the structure. */
aleator_status aleator_structure_species_symbol(const aleator_structure* structure, uint32_t species,
                                                const char** symbol);
aleator_status aleator_structure_source_label(const aleator_structure* structure, uint64_t atom,
                                              const char** label);

/* io::parseCifDocument (io/cif_document.hpp:26-50, cif_document.cc:185-350): the first data
 * block's single tags (lowercased) and loop_ tables, values as tokenized (quotes removed). */
typedef struct aleator_cif_document aleator_cif_document;
aleator_status aleator_cif_document_parse(const char* text, uint64_t length, aleator_cif_document** out);
aleator_status aleator_cif_document_destroy(aleator_cif_document* doc);
aleator_status aleator_cif_document_block_name(const aleator_cif_document* doc, const char** name);
/* CifDocument::findTag: *value is NULL when the tag is absent (the lookup is case-blind). */
aleator_status aleator_cif_document_find_tag(const aleator_cif_document* doc, const char* tag,
                                             const char** value);
/* The single tags in ascending (byte) order of their lowercased names. */
aleator_status aleator_cif_document_num_tags(const aleator_cif_document* doc, uint64_t* n_tags);
aleator_status aleator_cif_document_tag(const aleator_cif_document* doc, uint64_t index, const char** tag,
                                        const char** value);
aleator_status aleator_cif_document_num_loops(const aleator_cif_document* doc, uint64_t* n_loops);
/* CifDocument::findLoopWithColumns: index of the first loop holding every tag (lowercase), or -1. */
aleator_status aleator_cif_document_find_loop(const aleator_cif_document* doc, const char* const* tags,
                                              uint32_t n_tags, int64_t* loop);
aleator_status aleator_cif_document_loop_shape(const aleator_cif_document* doc, uint64_t loop,
                                               uint64_t* n_columns, uint64_t* n_rows);
aleator_status aleator_cif_document_loop_column(const aleator_cif_document* doc, uint64_t loop,
                                                uint64_t column, const char** tag);
aleator_status aleator_cif_document_loop_value(const aleator_cif_document* doc, uint64_t loop,
                                               uint64_t row, uint64_t column, const char** value);
/* tryParseCifNumber / parseCifNumber (cif_document.hpp:52-61): "?" and "." are absent, a
 * "(esd)" suffix is dropped.  The second returns ALEATOR_PARSE_ERROR where the first reports
 * no value. */
aleator_status aleator_cif_try_parse_number(const char* token, double* value, int32_t* has_value);
aleator_status aleator_cif_parse_number(const char* token, double* value);
/* parseSymmetryOperation (io/symmetry_operation.hpp:20-28): x' = M x + t, M row-major. */
aleator_status aleator_parse_symmetry_operation(const char* text, double matrix[9], double translation[3]);
/* elementFromLabel / atomicMass (io/element_table.hpp:9-21); symbol is NUL-terminated. */
aleator_status aleator_element_from_label(const char* label, char symbol[4], int32_t* found);
aleator_status aleator_atomic_mass(const char* symbol, double* mass, int32_t* found);

/* ---- synthetic frameworks (no reference counterpart) ------------------ */
/* Uniform fractional placement from Philox4x32Rng(seed, stream), rejecting a candidate
 * closer than min_dist (minimum image) to an earlier atom; species assigned in index order
 * with count round(p_t*N) for all but the last type; charges N(0, charge_sigma) by
 * Box-Muller on Philox4x32Rng(seed, charge_stream) minus their mean (none if sigma <= 0). */
aleator_status aleator_synth_framework(const aleator_lattice* lattice, uint64_t n_atoms,
                                       uint64_t seed, uint64_t stream, double min_dist,
                                       const double* proportions, uint32_t n_types,
                                       double charge_sigma, uint64_t charge_stream, double* x,
                                       double* y, double* z, double* charge, uint32_t* species);

#ifdef __cplusplus
}
#endif
#endif
