/*
 * oracle.h -- FP64 CPU restatement of the adsorption hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP
 * implementation in aleator_amd/ is compared against.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it; the
 * product path never calls it and fails loudly when its HIP library is
 * missing.
 *
 * Every function restates the reference algorithm in the same arithmetic and
 * loop order (compiled with -ffp-contract=off), citing the reference file and
 * line it follows.  Pinning of each function (golden vectors, the compiled
 * reference in oracle/_ref, closed forms) is listed in the header comment of
 * oracle.c.
 */
#ifndef ALEATOR_ORACLE_H
#define ALEATOR_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- Philox4x32-10 (counter_based_rng.cc:13-79) ----------------------- */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

typedef struct {
    uint32_t key[2];
    uint32_t ctr[4];
    uint64_t call_index;
} orc_rng;

void orc_rng_seed(orc_rng* rng, uint64_t seed, uint64_t stream);
uint64_t orc_rng_next_u64(orc_rng* rng);
double orc_rng_next_uniform(orc_rng* rng);
/* n uniforms of stream (seed, stream) starting at call index `first`. */
void orc_uniforms(uint64_t seed, uint64_t stream, uint64_t first, uint64_t n, double* out);

/* ---- Lattice (lattice.cc) -------------------------------------------- */
typedef struct {
    double m[3][3];       /* row i = lattice vector a_i */
    int orthorhombic;     /* computeIsOrthorhombic, lattice.cc:115-132 */
    double reduced[3][3]; /* greedyReduce, lattice.cc:70-101 */
} orc_lattice;

void orc_lattice_init(orc_lattice* lat, const double m[3][3]);
double orc_lattice_volume(const orc_lattice* lat);
void orc_frac_to_cart(const orc_lattice* lat, const double f[3], double c[3]);
int orc_cart_to_frac(const orc_lattice* lat, const double c[3], double f[3]);
/* d = nearest image of (b - a); returns 0 or -1 for a degenerate cell. */
int orc_minimum_image(const orc_lattice* lat, const double a[3], const double b[3], double d[3]);
double orc_perpendicular_width(const orc_lattice* lat, int axis);
/* 0 if cutoff <= L_perp/2 on every axis, else the 1-based offending axis. */
int orc_validate_cutoff(const orc_lattice* lat, double cutoff);

/* ---- Particles and supercell (supercell.cc:7-60) --------------------- */
typedef struct {
    size_t n;
    double *x, *y, *z, *q;
    uint32_t* species;
} orc_particles;

void orc_min_supercell_replication(const orc_lattice* lat, double cutoff, int n[3]);
/* Allocates out->x..species (free with orc_particles_free). */
int orc_replicate_supercell(const orc_particles* in, const orc_lattice* lat, int nx, int ny, int nz,
                            orc_particles* out, double out_m[3][3]);
void orc_particles_free(orc_particles* p);

/* ---- Lennard-Jones (lennard_jones.cc) -------------------------------- */
enum { ORC_TRUNCATED = 0, ORC_SHIFTED = 1, ORC_LFS = 2 };
enum { ORC_MIX_LB = 0, ORC_MIX_GEOMETRIC = 1 };

typedef struct {
    const double* epsilon;
    const double* sigma;
    uint32_t n_species;
    double cutoff;
    int truncation;
    int mixing;
} orc_lj;

void orc_lj_mixed(const orc_lj* lj, uint32_t si, uint32_t sj, double* eps, double* sigma);
double orc_lj_pair_energy(const orc_lj* lj, double r, double eps, double sigma);
double orc_lj_particle_energy(const orc_lj* lj, size_t index, const orc_particles* p,
                              const orc_lattice* lat, const size_t* excluded, size_t n_excluded);
double orc_lj_tail_energy(const orc_lj* lj, const orc_particles* p, const orc_lattice* lat);
/* Whole-system pair energy over i<j (index order); tests only. */
double orc_lj_total_energy(const orc_lj* lj, const orc_particles* p, const orc_lattice* lat);
double orc_lj_total_virial(const orc_lj* lj, const orc_particles* p, const orc_lattice* lat);

/* ---- Ewald (ewald.cc) ------------------------------------------------ */
#define ORC_KCOULOMB 167100.9468982874

typedef struct {
    size_t n;
    double* k;      /* [n][3] */
    double* weight; /* [n] */
    int32_t* nvec;  /* [n][3] integer index vectors */
} orc_kset;

int orc_reciprocal_terms(const orc_lattice* lat, double alpha, int kmax, orc_kset* out);
void orc_kset_free(orc_kset* ks);
double orc_ewald_real_particle_energy(double alpha, double rc, size_t index, const orc_particles* p,
                                      const orc_lattice* lat, const size_t* excluded,
                                      size_t n_excluded);
typedef struct {
    double real, reciprocal, self, exclusion, surface;
} orc_ewald_breakdown;
/* exclusions: n_excl pairs (i,j).  Real space over i<j pairs in index order. */
int orc_ewald_breakdown_energy(double alpha, double rc, int kmax, const orc_particles* p,
                               const orc_lattice* lat, const size_t* excl_pairs, size_t n_excl,
                               orc_ewald_breakdown* out);

/* ---- EwaldIncrementalState (ewald_incremental_state.cc) -------------- */
typedef struct {
    double x, y, z, charge;
} orc_site;

typedef struct {
    orc_lattice lat;
    double alpha;
    orc_kset terms;
    double volume;
    double* s_re;
    double* s_im;
    double self_energy;
    double exclusion_energy;
} orc_ewald_state;

int orc_ewald_state_init(orc_ewald_state* st, double alpha, int kmax, const orc_particles* p,
                         const orc_lattice* lat);
void orc_ewald_state_free(orc_ewald_state* st);
double orc_ewald_state_committed(const orc_ewald_state* st);
/* computeDelta total; if d_re/d_im non-null they receive the per-k delta. */
double orc_ewald_state_delta(const orc_ewald_state* st, const orc_site* removed, size_t nr,
                             const orc_site* added, size_t na, double* d_re, double* d_im,
                             double* recip, double* self, double* excl);
void orc_ewald_state_commit(orc_ewald_state* st, const orc_site* removed, size_t nr,
                            const orc_site* added, size_t na);
/* molecules given as CSR: mol_offsets[n_mol+1] into mol_sites. */
void orc_ewald_state_resync(orc_ewald_state* st, const orc_particles* p, const size_t* mol_offsets,
                            const size_t* mol_sites, size_t n_mol);
double orc_ewald_state_recompute(const orc_ewald_state* st, const orc_particles* p,
                                 const size_t* mol_offsets, const size_t* mol_sites, size_t n_mol);
/* 2*Nk + 2 doubles: s_re, s_im, self, exclusion. */
void orc_ewald_state_raw(const orc_ewald_state* st, double* out);

/* ---- FrameworkEnergyGrid (framework_energy_grid.cc:193-288) ---------- */
void orc_grid_point_counts(const orc_lattice* lat, double spacing, int32_t n[3]);
/* out[g][nz][ny][nx] for each guest in `guests` (caller passes the sorted unique list). */
int orc_grid_build(const orc_lj* lj, const orc_particles* fw, const orc_lattice* lat,
                   const uint32_t* guests, int n_guest, const int32_t n[3], double cap, double* out);
/* Generalised node set: nodes at (i/n)*grid_m, atoms periodic in `lat`.  Nodes listed
 * explicitly by index (flat (iz*ny+iy)*nx+ix) so tests can sample a subset. */
int orc_grid_nodes_lj(const orc_lj* lj, const orc_particles* fw, const orc_lattice* lat,
                      const double grid_m[3][3], const int32_t n[3], uint32_t guest,
                      const int64_t* nodes, size_t n_nodes, double cap, double* out);
/* Coulomb potential per unit charge (K/e) at explicit nodes: real space (erfc) + reciprocal
 * over the full k-set of st.  Composed from ewald.cc:223-249 and the linear term of
 * ewald_incremental_state.cc:117-121. */
int orc_grid_nodes_coulomb(const orc_ewald_state* st, double rc_real, const orc_particles* fw,
                           const double grid_m[3][3], const int32_t n[3], const int64_t* nodes,
                           size_t n_nodes, double* out_real, double* out_recip);
double orc_grid_interpolate(const orc_lattice* lat, const int32_t n[3], const double* table,
                            const double pos[3]);

/* ---- Widom (monte_carlo_engine.cc:850-917) --------------------------- */
typedef struct {
    double x, y, z;
    uint32_t lj_species;
    double charge;
} orc_mol_site;

/* Returns K_H; du_out[t] per trial if non-null.  ewald_state may be NULL (uncharged). */
double orc_widom(const orc_lj* lj, const orc_particles* fw, const orc_lattice* lat,
                 const orc_mol_site* sites, size_t n_sites, const orc_ewald_state* ewald_state,
                 double rc_real, uint64_t seed, uint64_t stream, uint64_t first_call,
                 uint64_t n_trials, double temperature, double* du_out);

/* ---- GCMC move loop (monte_carlo_engine.cc:380-439, 441-671, 814-848) --- */
typedef struct {
    orc_lj lj;
    double *eps, *sig;
    orc_lattice lat;
    size_t n, cap, n_fw;
    double *x, *y, *z, *q;
    uint32_t* sp;
    size_t n_mol, mol_cap, ns;
    size_t* mol; /* [n_mol][ns] particle indices, molecule order as the reference's vector */
    orc_mol_site* sites;
    double fugacity_internal, temperature, max_translation, max_rotation;
    int charged;
    double alpha, rc_real;
    orc_ewald_state ew;
    size_t moves_since_resync;
    orc_rng rng;
    uint64_t attempted[4], accepted[4]; /* translate, rotate, insert, delete */
    /* insertion density (no reference counterpart); prop_cdf NULL = uniform insertion */
    double *prop_cdf, *prop_fac;
    int prop_nb[3], prop_reps[3];
    size_t prop_nbins;
    double prop_inv[3][3];
    /* framework energy grid (monte_carlo_engine.cc:341-374 with frameworkEnergyGrid_): NULL =
     * direct scan.  grid_tables[k] is the [n2][n1][n0] table of species grid_species[k]. */
    double* grid_tables;
    uint32_t* grid_species;
    int32_t grid_n_tables;
    int32_t grid_n[3];
} orc_gcmc;

int orc_gcmc_init(orc_gcmc* g, const orc_lj* lj, const orc_particles* fw, const orc_lattice* lat,
                  const orc_mol_site* sites, size_t n_sites, double fugacity_pa, double temperature,
                  double max_translation, double max_rotation, double alpha, double rc_real, int kmax,
                  int charged, uint64_t seed, uint64_t stream);
void orc_gcmc_run(orc_gcmc* g, uint64_t steps, double* n_trace);
/* Biased insertion from a piecewise-constant density over nb bins of the cell divided by reps,
 * repeated over the reps images: p(bin) = (1 - mix) w / sum(w) + mix / n_bins.  Insertion
 * acceptance is divided by V p(anchor), deletion acceptance multiplied by V p(centroid).
 * w = NULL restores uniform insertion.  Returns -1 on invalid input. */
int orc_gcmc_set_proposal(orc_gcmc* g, const double* w, const int nb[3], const int reps[3], double mix);
double orc_gcmc_electrostatic_energy(const orc_gcmc* g);
void orc_gcmc_free(orc_gcmc* g);
/* Framework LJ from interpolated tables (copied), as the reference engine given a
 * FrameworkEnergyGrid; tables == NULL restores the direct scan.  -1 if a site species has no
 * table (FrameworkEnergyGrid::interpolate's std::out_of_range). */
int orc_gcmc_set_grid(orc_gcmc* g, const int32_t n[3], const uint32_t* species, int32_t n_tables,
                      const double* tables);

/* ---- Synthetic frameworks (no reference counterpart) ----------------- */
/* Uniform fractional placement from Philox(seed, stream), rejecting candidates
 * closer than min_dist to an earlier atom; types in index order by proportion;
 * charges N(0, charge_sigma) shifted to zero mean when charge_sigma > 0. */
int orc_synth_framework(const double m[3][3], size_t n_atoms, uint64_t seed, uint64_t stream,
                        double min_dist, const double* proportions, uint32_t n_types,
                        double charge_sigma, uint64_t charge_stream, double* x, double* y,
                        double* z, double* q, uint32_t* species);

/* ---- Learned insertion density network (no reference counterpart) ---- */
#define ORC_MLP_IN 64
#define ORC_MLP_FREQS 32
#define ORC_MLP_HIDDEN 256
#define ORC_MLP_PARAMS (ORC_MLP_IN * ORC_MLP_HIDDEN + ORC_MLP_HIDDEN + 3 * (ORC_MLP_HIDDEN * ORC_MLP_HIDDEN + ORC_MLP_HIDDEN) + ORC_MLP_HIDDEN + 1)
void orc_mlp_frequencies(int8_t out[3 * ORC_MLP_FREQS]);
/* y[r] = network(frac[r]) for n rows of fractional xyz; bf16 != 0 rounds GEMM operands to
 * bfloat16 where the device does */
void orc_mlp_forward(const double* params, const double* frac, size_t n, int bf16, double* y);
/* mean squared error against target and its gradient grad[ORC_MLP_PARAMS] */
double orc_mlp_loss_grad(const double* params, const double* frac, const double* target, size_t n, int bf16,
                         double* grad);

/* ---- Peng-Robinson (peng_robinson.cc:10-110) ------------------------- */
double orc_pr_fugacity(double tc, double pc, double omega, double temperature, double pressure);

#ifdef __cplusplus
}
#endif
#endif
