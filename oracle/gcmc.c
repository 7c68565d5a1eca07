/*
 * gcmc.c -- FP64 restatement of the reference's single-species grand-canonical Monte Carlo
 * move loop (TEST INFRASTRUCTURE ONLY; see oracle.h), the caller of the S(k) update:
 *   run                      monte_carlo_engine.cc:405-439 (one species: 25 % each move)
 *   attemptTranslation       :441-484
 *   attemptRotation          :486-554
 *   attemptInsertion         :556-619
 *   attemptDeletion          :621-671
 *   maybeResyncEwald         :380-403 (every 100 accepted moves)
 *   removeMolecule / removeParticleIndices  :814-848 (swap-with-last particle removal)
 *   acceptance ratios        gcmc_acceptance.hpp:28-101
 * Guest-guest dispersion is a full scan in particle-index order, which is what the reference's
 * guest cell list reproduces bit for bit (monte_carlo_engine.cc:316-374).  With an energy grid
 * (orc_gcmc_set_grid) the framework part is the grid's interpolation instead.
 * Nested-argument draws are taken left to right (see oracle.c, Widom).
 *
 * Pinned bitwise against the reference engine compiled in oracle/_ref
 * (tests/test_oracle_vs_reference.py).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define ORC_PI 3.14159265358979323846

static int grow(orc_gcmc* g, size_t need) {
    if (need <= g->cap) return 0;
    size_t cap = g->cap ? 2 * g->cap : 1024;
    while (cap < need) cap *= 2;
    g->x = (double*)realloc(g->x, cap * sizeof(double));
    g->y = (double*)realloc(g->y, cap * sizeof(double));
    g->z = (double*)realloc(g->z, cap * sizeof(double));
    g->q = (double*)realloc(g->q, cap * sizeof(double));
    g->sp = (uint32_t*)realloc(g->sp, cap * sizeof(uint32_t));
    g->cap = cap;
    return (g->x && g->y && g->z && g->q && g->sp) ? 0 : -1;
}

static size_t push_particle(orc_gcmc* g, double x, double y, double z, double q, uint32_t sp) {
    grow(g, g->n + 1);
    g->x[g->n] = x;
    g->y[g->n] = y;
    g->z[g->n] = z;
    g->q[g->n] = q;
    g->sp[g->n] = sp;
    return g->n++;
}

static orc_particles view(const orc_gcmc* g) {
    orc_particles p = {g->n, g->x, g->y, g->z, g->q, g->sp};
    return p;
}

int orc_gcmc_init(orc_gcmc* g, const orc_lj* lj, const orc_particles* fw, const orc_lattice* lat,
                  const orc_mol_site* sites, size_t n_sites, double fugacity_pa, double temperature,
                  double max_translation, double max_rotation, double alpha, double rc_real, int kmax,
                  int charged, uint64_t seed, uint64_t stream) {
    memset(g, 0, sizeof(*g));
    g->eps = (double*)malloc(lj->n_species * sizeof(double));
    g->sig = (double*)malloc(lj->n_species * sizeof(double));
    memcpy(g->eps, lj->epsilon, lj->n_species * sizeof(double));
    memcpy(g->sig, lj->sigma, lj->n_species * sizeof(double));
    g->lj = *lj;
    g->lj.epsilon = g->eps;
    g->lj.sigma = g->sig;
    g->lat = *lat;
    g->ns = n_sites;
    g->sites = (orc_mol_site*)malloc(n_sites * sizeof(orc_mol_site));
    memcpy(g->sites, sites, n_sites * sizeof(orc_mol_site));
    g->fugacity_internal = fugacity_pa * (1.0 / (1.380649e-23 * 1.0e30));
    g->temperature = temperature;
    g->max_translation = max_translation;
    g->max_rotation = max_rotation;
    if (grow(g, fw->n + 64)) return -1;
    for (size_t i = 0; i < fw->n; ++i) push_particle(g, fw->x[i], fw->y[i], fw->z[i], fw->q ? fw->q[i] : 0.0, fw->species[i]);
    g->n_fw = fw->n;
    g->charged = charged;
    g->alpha = alpha;
    g->rc_real = rc_real;
    if (charged) {
        orc_particles p = view(g);
        if (orc_ewald_state_init(&g->ew, alpha, kmax, &p, lat)) return -1;
    }
    orc_rng_seed(&g->rng, seed, stream);
    return 0;
}

void orc_gcmc_free(orc_gcmc* g) {
    free(g->x);
    free(g->y);
    free(g->z);
    free(g->q);
    free(g->sp);
    free(g->mol);
    free(g->sites);
    free(g->eps);
    free(g->sig);
    free(g->prop_cdf);
    free(g->prop_fac);
    free(g->grid_tables);
    free(g->grid_species);
    if (g->charged) orc_ewald_state_free(&g->ew);
    memset(g, 0, sizeof(*g));
}

static double uni(orc_gcmc* g) { return orc_rng_next_uniform(&g->rng); }

/* computeParticleEnergyOverCandidates over the guests, ascending index (lennard_jones.cc:184-216) */
static double guest_lj(orc_gcmc* g, size_t index, const size_t* excl, size_t ne) {
    double energy = 0.0;
    const double pi[3] = {g->x[index], g->y[index], g->z[index]};
    for (size_t j = g->n_fw; j < g->n; ++j) {
        if (j == index) continue;
        int skip = 0;
        for (size_t k = 0; k < ne && !skip; ++k) skip = excl[k] == j;
        if (skip) continue;
        double eps, sig, d[3];
        orc_lj_mixed(&g->lj, g->sp[index], g->sp[j], &eps, &sig);
        const double pj[3] = {g->x[j], g->y[j], g->z[j]};
        orc_minimum_image(&g->lat, pi, pj, d);
        energy += orc_lj_pair_energy(&g->lj, sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]), eps, sig);
    }
    return energy;
}

/* moleculeDispersionEnergy (monte_carlo_engine.cc:341-374): per site, the framework part
 * (direct scan, or the grid's interpolation) and then the guests */
static double dispersion(orc_gcmc* g, const size_t* sites, size_t ns, const size_t* also, size_t na) {
    size_t excl[64];
    size_t ne = 0;
    for (size_t i = 0; i < ns; ++i) excl[ne++] = sites[i];
    for (size_t i = 0; i < na; ++i) excl[ne++] = also[i];
    double e = 0.0;
    if (g->grid_tables) {
        const size_t cells = (size_t)g->grid_n[0] * g->grid_n[1] * g->grid_n[2];
        for (size_t s = 0; s < ns; ++s) {
            const size_t i = sites[s];
            int32_t k = 0;
            while (g->grid_species[k] != g->sp[i]) ++k; /* present: checked by orc_gcmc_set_grid */
            const double pos[3] = {g->x[i], g->y[i], g->z[i]};
            e += orc_grid_interpolate(&g->lat, g->grid_n, g->grid_tables + (size_t)k * cells, pos);
            e += guest_lj(g, i, excl, ne);
        }
        return e;
    }
    orc_particles p = view(g);
    for (size_t s = 0; s < ns; ++s) e += orc_lj_particle_energy(&g->lj, sites[s], &p, &g->lat, excl, ne);
    return e;
}

int orc_gcmc_set_grid(orc_gcmc* g, const int32_t n[3], const uint32_t* species, int32_t n_tables,
                      const double* tables) {
    free(g->grid_tables);
    free(g->grid_species);
    g->grid_tables = NULL;
    g->grid_species = NULL;
    g->grid_n_tables = 0;
    if (!tables) return 0;
    for (size_t s = 0; s < g->ns; ++s) {
        int found = 0;
        for (int32_t k = 0; k < n_tables; ++k) found |= species[k] == g->sites[s].lj_species;
        if (!found) return -1;
    }
    const size_t cells = (size_t)n[0] * n[1] * n[2];
    g->grid_tables = (double*)malloc((size_t)n_tables * cells * sizeof(double));
    g->grid_species = (uint32_t*)malloc((size_t)n_tables * sizeof(uint32_t));
    memcpy(g->grid_tables, tables, (size_t)n_tables * cells * sizeof(double));
    memcpy(g->grid_species, species, (size_t)n_tables * sizeof(uint32_t));
    g->grid_n_tables = n_tables;
    for (int a = 0; a < 3; ++a) g->grid_n[a] = n[a];
    return 0;
}

/* moleculeElectrostaticRealSpaceEnergy (monte_carlo_engine.cc:299-314) */
static double electro_real(orc_gcmc* g, const size_t* sites, size_t ns) {
    if (!g->charged) return 0.0;
    orc_particles p = view(g);
    double e = 0.0;
    for (size_t s = 0; s < ns; ++s)
        e += orc_ewald_real_particle_energy(g->alpha, g->rc_real, sites[s], &p, &g->lat, sites, ns);
    return e;
}

static void charged_sites_of(const orc_gcmc* g, const size_t* sites, size_t ns, orc_site* out) {
    for (size_t s = 0; s < ns; ++s) {
        out[s].x = g->x[sites[s]];
        out[s].y = g->y[sites[s]];
        out[s].z = g->z[sites[s]];
        out[s].charge = g->q[sites[s]];
    }
}

static void resync(orc_gcmc* g) {
    size_t* off = (size_t*)malloc((g->n_mol + 1) * sizeof(size_t));
    for (size_t m = 0; m <= g->n_mol; ++m) off[m] = m * g->ns;
    orc_particles p = view(g);
    orc_ewald_state_resync(&g->ew, &p, off, g->mol, g->n_mol);
    free(off);
}

/* maybeResyncEwald (monte_carlo_engine.cc:380-403) */
static void maybe_resync(orc_gcmc* g) {
    if (!g->charged) return;
    ++g->moves_since_resync;
    if (g->moves_since_resync >= 100) {
        resync(g);
        g->moves_since_resync = 0;
    }
}

static double clamp_prob(double r) { return r < 1.0 ? r : 1.0; }

/* attemptTranslation (monte_carlo_engine.cc:441-484) */
static void attempt_translation(orc_gcmc* g) {
    if (g->n_mol == 0) return;
    const size_t mi = (size_t)(uni(g) * (double)g->n_mol);
    size_t* sites = g->mol + mi * g->ns;
    const size_t ns = g->ns;
    const double u_before = dispersion(g, sites, ns, NULL, 0);
    const double e_before = electro_real(g, sites, ns);
    orc_site old_cs[16], new_cs[16];
    if (g->charged) charged_sites_of(g, sites, ns, old_cs);
    const double d0 = (2.0 * uni(g) - 1.0) * g->max_translation;
    const double d1 = (2.0 * uni(g) - 1.0) * g->max_translation;
    const double d2 = (2.0 * uni(g) - 1.0) * g->max_translation;
    double orig[16][3];
    for (size_t k = 0; k < ns; ++k) {
        const size_t s = sites[k];
        orig[k][0] = g->x[s];
        orig[k][1] = g->y[s];
        orig[k][2] = g->z[s];
        g->x[s] = orig[k][0] + d0;
        g->y[s] = orig[k][1] + d1;
        g->z[s] = orig[k][2] + d2;
    }
    const double u_after = dispersion(g, sites, ns, NULL, 0);
    const double e_after = electro_real(g, sites, ns);
    double recip = 0.0;
    if (g->charged) {
        charged_sites_of(g, sites, ns, new_cs);
        recip = orc_ewald_state_delta(&g->ew, old_cs, ns, new_cs, ns, NULL, NULL, NULL, NULL, NULL);
    }
    const double du = (u_after - u_before) + (e_after - e_before) + recip;
    const double ratio = exp(-du / g->temperature);
    g->attempted[0]++;
    if (uni(g) < clamp_prob(ratio)) {
        if (g->charged) {
            orc_ewald_state_commit(&g->ew, old_cs, ns, new_cs, ns);
            maybe_resync(g);
        }
        g->accepted[0]++;
        return;
    }
    for (size_t k = 0; k < ns; ++k) {
        g->x[sites[k]] = orig[k][0];
        g->y[sites[k]] = orig[k][1];
        g->z[sites[k]] = orig[k][2];
    }
}

typedef struct {
    double v[3];
} v3t;
static v3t mk(double a, double b, double c) {
    v3t r = {{a, b, c}};
    return r;
}
static v3t add3(v3t a, v3t b) { return mk(a.v[0] + b.v[0], a.v[1] + b.v[1], a.v[2] + b.v[2]); }
static v3t sub3(v3t a, v3t b) { return mk(a.v[0] - b.v[0], a.v[1] - b.v[1], a.v[2] - b.v[2]); }
static v3t scale3(v3t a, double s) { return mk(a.v[0] * s, a.v[1] * s, a.v[2] * s); }
static double dot3(v3t a, v3t b) { return a.v[0] * b.v[0] + a.v[1] * b.v[1] + a.v[2] * b.v[2]; }
static v3t cross3(v3t a, v3t b) {
    return mk(a.v[1] * b.v[2] - a.v[2] * b.v[1], a.v[2] * b.v[0] - a.v[0] * b.v[2],
              a.v[0] * b.v[1] - a.v[1] * b.v[0]);
}

/* rotateAboutAxis (monte_carlo_engine.cc:40-44) */
static v3t rotate_about_axis(v3t v, v3t axis, double theta) {
    const double c = cos(theta);
    const double s = sin(theta);
    return add3(add3(scale3(v, c), scale3(cross3(axis, v), s)), scale3(axis, dot3(axis, v) * (1.0 - c)));
}

/* attemptRotation (monte_carlo_engine.cc:486-554), single species */
static void attempt_rotation(orc_gcmc* g) {
    if (g->n_mol == 0) return;
    if (g->ns < 2) return;
    const size_t mi = (size_t)(uni(g) * (double)g->n_mol);
    size_t* sites = g->mol + mi * g->ns;
    const size_t ns = g->ns;
    const double u_before = dispersion(g, sites, ns, NULL, 0);
    const double e_before = electro_real(g, sites, ns);
    orc_site old_cs[16], new_cs[16];
    if (g->charged) charged_sites_of(g, sites, ns, old_cs);
    v3t centroid = mk(0.0, 0.0, 0.0);
    for (size_t k = 0; k < ns; ++k) centroid = add3(centroid, mk(g->x[sites[k]], g->y[sites[k]], g->z[sites[k]]));
    centroid = scale3(centroid, 1.0 / (double)ns);
    const double cos_t = 2.0 * uni(g) - 1.0;
    const double one_minus = 1.0 - cos_t * cos_t;
    const double sin_t = sqrt((0.0 < one_minus) ? one_minus : 0.0); /* std::max(0.0, .) */
    const double phi = 2.0 * ORC_PI * uni(g);
    const v3t axis = mk(sin_t * cos(phi), sin_t * sin(phi), cos_t);
    const double angle = (2.0 * uni(g) - 1.0) * g->max_rotation;
    double orig[16][3];
    for (size_t k = 0; k < ns; ++k) {
        const size_t s = sites[k];
        const v3t p = mk(g->x[s], g->y[s], g->z[s]);
        orig[k][0] = p.v[0];
        orig[k][1] = p.v[1];
        orig[k][2] = p.v[2];
        const v3t np = add3(centroid, rotate_about_axis(sub3(p, centroid), axis, angle));
        g->x[s] = np.v[0];
        g->y[s] = np.v[1];
        g->z[s] = np.v[2];
    }
    const double u_after = dispersion(g, sites, ns, NULL, 0);
    const double e_after = electro_real(g, sites, ns);
    double recip = 0.0;
    if (g->charged) {
        charged_sites_of(g, sites, ns, new_cs);
        recip = orc_ewald_state_delta(&g->ew, old_cs, ns, new_cs, ns, NULL, NULL, NULL, NULL, NULL);
    }
    const double du = (u_after - u_before) + (e_after - e_before) + recip;
    const double ratio = exp(-du / g->temperature);
    g->attempted[1]++;
    if (uni(g) < clamp_prob(ratio)) {
        if (g->charged) {
            orc_ewald_state_commit(&g->ew, old_cs, ns, new_cs, ns);
            maybe_resync(g);
        }
        g->accepted[1]++;
        return;
    }
    for (size_t k = 0; k < ns; ++k) {
        g->x[sites[k]] = orig[k][0];
        g->y[sites[k]] = orig[k][1];
        g->z[sites[k]] = orig[k][2];
    }
}

/* randomUnitQuaternion / rotateByQuaternion (monte_carlo_engine.cc:56-72) */
static void quat_rotate(double u0, double u1, double u2, const double v[3], double out[3]) {
    const double r1 = sqrt(1.0 - u0);
    const double r2 = sqrt(u0);
    const double th1 = 2.0 * ORC_PI * u1;
    const double th2 = 2.0 * ORC_PI * u2;
    const double w = r2 * cos(th2), x = r1 * sin(th1), y = r1 * cos(th1), z = r2 * sin(th2);
    const v3t qv = mk(x, y, z);
    const v3t vv = mk(v[0], v[1], v[2]);
    const v3t t = scale3(cross3(qv, vv), 2.0);
    const v3t r = add3(add3(vv, scale3(t, w)), cross3(qv, t));
    out[0] = r.v[0];
    out[1] = r.v[1];
    out[2] = r.v[2];
}

/* ---- insertion density (no reference counterpart) ---------------------------------------- */

/* smallest bin with u < cdf[bin] (cdf[last] = 1) */
static size_t proposal_draw_bin(const orc_gcmc* g, double u) {
    size_t lo = 0, hi = g->prop_nbins - 1;
    while (lo < hi) {
        const size_t mid = (lo + hi) / 2;
        if (u < g->prop_cdf[mid])
            hi = mid;
        else
            lo = mid + 1;
    }
    return lo;
}

/* bin of a molecule's anchor: centroid of its sites -> fractional (c . inv, inv = cofactor
 * rows / volume) -> folded into one primitive image -> binned */
static size_t proposal_anchor_bin(const orc_gcmc* g, const size_t* sites) {
    double cen[3] = {0.0, 0.0, 0.0};
    for (size_t s = 0; s < g->ns; ++s) {
        cen[0] = cen[0] + g->x[sites[s]];
        cen[1] = cen[1] + g->y[sites[s]];
        cen[2] = cen[2] + g->z[sites[s]];
    }
    const double inv_n = 1.0 / (double)g->ns;
    size_t bin = 0;
    for (int j = 0; j < 3; ++j) {
        const double f = (cen[0] * inv_n) * g->prop_inv[0][j] + (cen[1] * inv_n) * g->prop_inv[1][j] +
                         (cen[2] * inv_n) * g->prop_inv[2][j];
        double t = f * (double)g->prop_reps[j];
        t = t - floor(t);
        const double kf = t * (double)g->prop_nb[j];
        int k = kf >= 0.0 ? (int)kf : 0;
        if (k > g->prop_nb[j] - 1) k = g->prop_nb[j] - 1;
        bin = bin * (size_t)g->prop_nb[j] + (size_t)k;
    }
    return bin;
}

int orc_gcmc_set_proposal(orc_gcmc* g, const double* w, const int nb[3], const int reps[3], double mix) {
    free(g->prop_cdf);
    free(g->prop_fac);
    g->prop_cdf = g->prop_fac = NULL;
    g->prop_nbins = 0;
    if (!w) return 0;
    if (!(mix >= 0.0 && mix <= 1.0)) return -1;
    size_t n = 1;
    for (int i = 0; i < 3; ++i) {
        if (nb[i] < 1 || reps[i] < 1) return -1;
        n *= (size_t)nb[i];
    }
    double total = 0.0;
    for (size_t b = 0; b < n; ++b) {
        if (!(w[b] >= 0.0) || !isfinite(w[b])) return -1;
        total += w[b];
    }
    if (!(total > 0.0)) return -1;
    double* cdf = (double*)malloc(n * sizeof(double));
    double* fac = (double*)malloc(n * sizeof(double));
    double run = 0.0;
    for (size_t b = 0; b < n; ++b) {
        run += (1.0 - mix) * (w[b] / total) + mix / (double)n;
        cdf[b] = run;
    }
    cdf[n - 1] = 1.0;
    double prev = 0.0;
    for (size_t b = 0; b < n; ++b) {
        if (!(cdf[b] > prev)) {
            free(cdf);
            free(fac);
            return -1;
        }
        /* the probability the inverse-CDF draw gives bin b */
        fac[b] = (double)n * (cdf[b] - prev);
        prev = cdf[b];
    }
    g->prop_cdf = cdf;
    g->prop_fac = fac;
    g->prop_nbins = n;
    for (int i = 0; i < 3; ++i) {
        g->prop_nb[i] = nb[i];
        g->prop_reps[i] = reps[i];
    }
    /* inverse of the row-vector cell: column j = (a_{j+1} x a_{j+2}) / V */
    const double(*m)[3] = g->lat.m;
    const double cx = m[1][1] * m[2][2] - m[1][2] * m[2][1];
    const double cy = m[1][2] * m[2][0] - m[1][0] * m[2][2];
    const double cz = m[1][0] * m[2][1] - m[1][1] * m[2][0];
    const double vol = m[0][0] * cx + m[0][1] * cy + m[0][2] * cz;
    for (int j = 0; j < 3; ++j) {
        const double* a = m[(j + 1) % 3];
        const double* b = m[(j + 2) % 3];
        const double c[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
        for (int r = 0; r < 3; ++r) g->prop_inv[r][j] = c[r] / vol;
    }
    return 0;
}

/* attemptInsertion (monte_carlo_engine.cc:556-619), single species */
static void attempt_insertion(orc_gcmc* g) {
    const double(*m)[3] = g->lat.m;
    double ref[3];
    double bias = 1.0;
    if (g->prop_cdf) {
        const size_t bin = proposal_draw_bin(g, uni(g));
        bias = g->prop_fac[bin];
        const int nb1 = g->prop_nb[1], nb2 = g->prop_nb[2];
        const int bn[3] = {(int)(bin / (size_t)(nb1 * nb2)), (int)((bin / (size_t)nb2) % (size_t)nb1),
                           (int)(bin % (size_t)nb2)};
        const int r1 = g->prop_reps[1], r2 = g->prop_reps[2];
        const int n_img = g->prop_reps[0] * r1 * r2;
        int img = (int)(uni(g) * (double)n_img);
        if (img > n_img - 1) img = n_img - 1;
        const int im[3] = {img / (r1 * r2), (img / r2) % r1, img % r2};
        double f[3];
        for (int j = 0; j < 3; ++j) {
            const double u = uni(g);
            f[j] = ((double)im[j] + ((double)bn[j] + u) / (double)g->prop_nb[j]) / (double)g->prop_reps[j];
        }
        for (int j = 0; j < 3; ++j) ref[j] = m[0][j] * f[0] + m[1][j] * f[1] + m[2][j] * f[2];
    } else {
        const double u0 = uni(g), u1 = uni(g), u2 = uni(g);
        for (int j = 0; j < 3; ++j) ref[j] = m[0][j] * u0 + m[1][j] * u1 + m[2][j] * u2;
    }
    double q0 = 0, q1 = 0, q2 = 0;
    if (g->ns > 1) {
        q0 = uni(g);
        q1 = uni(g);
        q2 = uni(g);
    }
    size_t new_sites[16];
    for (size_t k = 0; k < g->ns; ++k) {
        const double local[3] = {g->sites[k].x, g->sites[k].y, g->sites[k].z};
        double rot[3] = {local[0], local[1], local[2]};
        if (g->ns > 1) quat_rotate(q0, q1, q2, local, rot);
        new_sites[k] = push_particle(g, ref[0] + rot[0], ref[1] + rot[1], ref[2] + rot[2], g->sites[k].charge,
                                     g->sites[k].lj_species);
    }
    const double disp = dispersion(g, new_sites, g->ns, NULL, 0);
    const double real = electro_real(g, new_sites, g->ns);
    orc_site cs[16];
    double recip = 0.0;
    if (g->charged) {
        charged_sites_of(g, new_sites, g->ns, cs);
        recip = orc_ewald_state_delta(&g->ew, NULL, 0, cs, g->ns, NULL, NULL, NULL, NULL, NULL);
    }
    const double du = disp + real + recip;
    const double volume = fabs(orc_lattice_volume(&g->lat));
    double ratio = (g->fugacity_internal * volume / g->temperature) / (double)(g->n_mol + 1) *
                   exp(-du / g->temperature);
    if (g->prop_cdf) ratio = ratio / bias;
    g->attempted[2]++;
    if (uni(g) < clamp_prob(ratio)) {
        if (g->charged) orc_ewald_state_commit(&g->ew, NULL, 0, cs, g->ns);
        if (g->n_mol == g->mol_cap) {
            g->mol_cap = g->mol_cap ? 2 * g->mol_cap : 64;
            g->mol = (size_t*)realloc(g->mol, g->mol_cap * g->ns * sizeof(size_t));
        }
        memcpy(g->mol + g->n_mol * g->ns, new_sites, g->ns * sizeof(size_t));
        g->n_mol++;
        maybe_resync(g);
        g->accepted[2]++;
        return;
    }
    g->n -= g->ns;
}

/* removeParticleIndices (monte_carlo_engine.cc:825-848) */
static void remove_particle_indices(orc_gcmc* g, size_t* idx, size_t cnt) {
    for (size_t k = 0; k < cnt; ++k) {
        const size_t site = idx[k];
        const size_t last = g->n - 1;
        if (site != last) {
            g->x[site] = g->x[last];
            g->y[site] = g->y[last];
            g->z[site] = g->z[last];
            g->q[site] = g->q[last];
            g->sp[site] = g->sp[last];
            for (size_t i = 0; i < g->n_mol * g->ns; ++i)
                if (g->mol[i] == last) g->mol[i] = site;
            for (size_t k2 = k + 1; k2 < cnt; ++k2)
                if (idx[k2] == last) idx[k2] = site;
        }
        g->n--;
    }
}

/* attemptDeletion (monte_carlo_engine.cc:621-671), single species */
static void attempt_deletion(orc_gcmc* g) {
    if (g->n_mol == 0) return;
    const size_t mi = (size_t)(uni(g) * (double)g->n_mol);
    const size_t count_before = g->n_mol;
    size_t* sites = g->mol + mi * g->ns;
    const double u_int = dispersion(g, sites, g->ns, NULL, 0);
    const double real = electro_real(g, sites, g->ns);
    orc_site cs[16];
    double recip = 0.0;
    if (g->charged) {
        charged_sites_of(g, sites, g->ns, cs);
        recip = orc_ewald_state_delta(&g->ew, cs, g->ns, NULL, 0, NULL, NULL, NULL, NULL, NULL);
    }
    const double du = -(u_int + real) + recip;
    const double volume = fabs(orc_lattice_volume(&g->lat));
    double ratio = (double)count_before / (g->fugacity_internal * volume / g->temperature) *
                   exp(-du / g->temperature);
    if (g->prop_cdf) ratio = ratio * g->prop_fac[proposal_anchor_bin(g, sites)];
    g->attempted[3]++;
    if (uni(g) < clamp_prob(ratio)) {
        if (g->charged) orc_ewald_state_commit(&g->ew, cs, g->ns, NULL, 0);
        size_t idx[16];
        memcpy(idx, sites, g->ns * sizeof(size_t));
        memmove(g->mol + mi * g->ns, g->mol + (mi + 1) * g->ns, (g->n_mol - mi - 1) * g->ns * sizeof(size_t));
        g->n_mol--;
        remove_particle_indices(g, idx, g->ns);
        maybe_resync(g);
        g->accepted[3]++;
    }
}

/* run (monte_carlo_engine.cc:405-439), single species */
void orc_gcmc_run(orc_gcmc* g, uint64_t steps, double* n_trace) {
    for (uint64_t step = 0; step < steps; ++step) {
        const double pick = uni(g);
        if (pick < 0.25)
            attempt_translation(g);
        else if (pick < 0.5)
            attempt_rotation(g);
        else if (pick < 0.75)
            attempt_insertion(g);
        else
            attempt_deletion(g);
        if (n_trace) n_trace[step] = (double)g->n_mol;
    }
}

double orc_gcmc_electrostatic_energy(const orc_gcmc* g) {
    return g->charged ? orc_ewald_state_committed(&g->ew) : 0.0;
}
