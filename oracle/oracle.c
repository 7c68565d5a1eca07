/*
 * oracle.c -- FP64 CPU restatement of the adsorption hot path (TEST
 * INFRASTRUCTURE ONLY; see oracle.h).  Built with -ffp-contract=off so that
 * every expression rounds exactly as the reference's non-FMA x86-64 build.
 *
 * How each part is pinned (tests/test_oracle_*.py):
 *   philox            Random123 known-answer vectors (test_philox_random123_vectors.cc:22-44),
 *                     bitwise against the compiled reference (oracle/_ref).
 *   lattice           brute force over 125 images; frac<->cart round trip; bitwise vs _ref.
 *   supercell         closed-form counts, bitwise vs _ref.
 *   lennard-jones     NIST SRSW U_pair/W_pair/U_LRC (tests/golden/nist_lj), mixing closed
 *                     forms, tail vs quadrature, bitwise vs _ref.
 *   ewald             Madelung constant (rock salt), NIST SPC/E per-term energies, alpha
 *                     independence, k-set counts 948/2204, bitwise vs _ref.
 *   ewald state       delta == breakdown(after) - breakdown(before); insert/delete
 *                     antisymmetry; rollback; bitwise vs _ref.
 *   grid              bitwise vs _ref grid build (nodes read back from the reference's own
 *                     cache file); interpolation exact on nodes / linear fields.
 *   coulomb grid      phi_real == reference realSpaceParticleEnergy of a unit charge;
 *                     phi_recip == (proposeInsertion(+1) - proposeInsertion(-1)) / 2.
 *   widom             K_H bitwise vs the reference engine (_ref) for single- and
 *                     multi-site species, charged and uncharged.
 *   synth framework   no reference counterpart: minimum-distance and composition
 *                     invariants, and equality with the product generator.
 *   peng-robinson     ideal-gas limit and bitwise vs _ref.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define ORC_PI 3.14159265358979323846

/* ======================================================================
 * Philox4x32-10 -- counter_based_rng.cc:13-54 (constants :13-17, round :29-38)
 * ====================================================================== */
void orc_philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4]) {
    uint32_t c[4] = {ctr_in[0], ctr_in[1], ctr_in[2], ctr_in[3]};
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int round = 0; round < 10; ++round) {
        if (round > 0) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        const uint64_t p0 = (uint64_t)0xD2511F53u * (uint64_t)c[0];
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * (uint64_t)c[2];
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        const uint32_t n0 = hi1 ^ c[1] ^ k0;
        const uint32_t n2 = hi0 ^ c[3] ^ k1;
        c[0] = n0;
        c[1] = lo1;
        c[2] = n2;
        c[3] = lo0;
    }
    memcpy(out, c, sizeof(c));
}

/* Philox4x32Rng::seed/nextU64/nextUniform -- counter_based_rng.cc:56-79 */
void orc_rng_seed(orc_rng* rng, uint64_t seed, uint64_t stream) {
    rng->key[0] = (uint32_t)seed;
    rng->key[1] = (uint32_t)(seed >> 32);
    rng->ctr[0] = 0;
    rng->ctr[1] = 0;
    rng->ctr[2] = (uint32_t)stream;
    rng->ctr[3] = (uint32_t)(stream >> 32);
    rng->call_index = 0;
}

uint64_t orc_rng_next_u64(orc_rng* rng) {
    rng->ctr[0] = (uint32_t)rng->call_index;
    rng->ctr[1] = (uint32_t)(rng->call_index >> 32);
    ++rng->call_index;
    uint32_t out[4];
    orc_philox4x32_10(rng->ctr, rng->key, out);
    return ((uint64_t)out[0] << 32) | (uint64_t)out[1];
}

double orc_rng_next_uniform(orc_rng* rng) {
    return (double)(orc_rng_next_u64(rng) >> 11) * (1.0 / 9007199254740992.0);
}

void orc_uniforms(uint64_t seed, uint64_t stream, uint64_t first, uint64_t n, double* out) {
    orc_rng rng;
    orc_rng_seed(&rng, seed, stream);
    rng.call_index = first;
    for (uint64_t i = 0; i < n; ++i) out[i] = orc_rng_next_uniform(&rng);
}

/* ======================================================================
 * Lattice -- lattice.cc
 * ====================================================================== */
typedef struct {
    double v[3];
} vec3;

static vec3 v3(double a, double b, double c) {
    vec3 r = {{a, b, c}};
    return r;
}
static vec3 vadd(vec3 a, vec3 b) { return v3(a.v[0] + b.v[0], a.v[1] + b.v[1], a.v[2] + b.v[2]); }
static vec3 vsub(vec3 a, vec3 b) { return v3(a.v[0] - b.v[0], a.v[1] - b.v[1], a.v[2] - b.v[2]); }
static vec3 vscale(vec3 a, double s) { return v3(a.v[0] * s, a.v[1] * s, a.v[2] * s); }
static double vdot(vec3 a, vec3 b) { return a.v[0] * b.v[0] + a.v[1] * b.v[1] + a.v[2] * b.v[2]; }
static vec3 vcross(vec3 a, vec3 b) {
    return v3(a.v[1] * b.v[2] - a.v[2] * b.v[1], a.v[2] * b.v[0] - a.v[0] * b.v[2],
              a.v[0] * b.v[1] - a.v[1] * b.v[0]);
}
static double vnorm2(vec3 a) { return vdot(a, a); }
static vec3 row(const double m[3][3], int i) { return v3(m[i][0], m[i][1], m[i][2]); }

#define ORC_DEGENERATE 1e-10

/* fractionalOf -- lattice.cc:54-60 */
static vec3 fractional_of(vec3 c, vec3 b0, vec3 b1, vec3 b2, double volume) {
    return v3(vdot(c, vcross(b1, b2)) / volume, vdot(c, vcross(b2, b0)) / volume,
              vdot(c, vcross(b0, b1)) / volume);
}

/* greedyReduce -- lattice.cc:70-101.  std::sort on three elements is libstdc++'s
 * insertion sort (stable), restated here with the same strict comparator. */
static void greedy_reduce(vec3 v[3]) {
    for (int iter = 0; iter < 50; ++iter) {
        int changed = 0;
        for (int i = 0; i < 3; ++i) {
            for (int j = 0; j < 3; ++j) {
                if (i == j) continue;
                const double nj2 = vnorm2(v[j]);
                if (nj2 < ORC_DEGENERATE) continue;
                const double t = round(vdot(v[i], v[j]) / nj2);
                if (t != 0.0) {
                    v[i] = vsub(v[i], vscale(v[j], t));
                    changed = 1;
                }
            }
        }
        for (int i = 1; i < 3; ++i) {
            vec3 val = v[i];
            int k = i;
            while (k > 0 && vnorm2(val) < vnorm2(v[k - 1])) {
                v[k] = v[k - 1];
                --k;
            }
            v[k] = val;
        }
        if (!changed) break;
    }
}

/* computeIsOrthorhombic -- lattice.cc:115-132 */
static int compute_is_orthorhombic(const double m[3][3]) {
    for (int i = 0; i < 3; ++i) {
        for (int j = 0; j < 3; ++j) {
            if (i == j) continue;
            double scale = fabs(m[i][i]);
            if (fabs(m[j][j]) > scale) scale = fabs(m[j][j]);
            if (1.0 > scale) scale = 1.0;
            if (fabs(m[i][j]) > 1e-9 * scale) return 0;
        }
    }
    return 1;
}

void orc_lattice_init(orc_lattice* lat, const double m[3][3]) {
    memcpy(lat->m, m, sizeof(lat->m));
    lat->orthorhombic = compute_is_orthorhombic(m);
    vec3 v[3] = {row(m, 0), row(m, 1), row(m, 2)};
    greedy_reduce(v);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) lat->reduced[i][j] = v[i].v[j];
}

/* Lattice::volume -- lattice.hpp:35-41 */
double orc_lattice_volume(const orc_lattice* lat) {
    const double(*m)[3] = lat->m;
    const double cx = m[1][1] * m[2][2] - m[1][2] * m[2][1];
    const double cy = m[1][2] * m[2][0] - m[1][0] * m[2][2];
    const double cz = m[1][0] * m[2][1] - m[1][1] * m[2][0];
    return m[0][0] * cx + m[0][1] * cy + m[0][2] * cz;
}

/* fractionalToCartesian -- lattice.cc:157-164 */
void orc_frac_to_cart(const orc_lattice* lat, const double f[3], double c[3]) {
    const double(*m)[3] = lat->m;
    c[0] = f[0] * m[0][0] + f[1] * m[1][0] + f[2] * m[2][0];
    c[1] = f[0] * m[0][1] + f[1] * m[1][1] + f[2] * m[2][1];
    c[2] = f[0] * m[0][2] + f[1] * m[1][2] + f[2] * m[2][2];
}

/* cartesianToFractional -- lattice.cc:166-175 */
int orc_cart_to_frac(const orc_lattice* lat, const double c[3], double f[3]) {
    const double vol = orc_lattice_volume(lat);
    if (fabs(vol) < ORC_DEGENERATE) return -1;
    const vec3 r = fractional_of(v3(c[0], c[1], c[2]), row(lat->m, 0), row(lat->m, 1),
                                 row(lat->m, 2), vol);
    f[0] = r.v[0];
    f[1] = r.v[1];
    f[2] = r.v[2];
    return 0;
}

/* minimumImageDisplacement -- lattice.cc:185-255 */
int orc_minimum_image(const orc_lattice* lat, const double a[3], const double b[3], double out[3]) {
    const double vol = orc_lattice_volume(lat);
    if (fabs(vol) < ORC_DEGENERATE) return -1;
    const vec3 d = vsub(v3(b[0], b[1], b[2]), v3(a[0], a[1], a[2]));
    if (lat->orthorhombic) {
        vec3 w = d;
        for (int axis = 0; axis < 3; ++axis) {
            const double length = lat->m[axis][axis];
            w.v[axis] -= length * round(w.v[axis] / length);
        }
        out[0] = w.v[0];
        out[1] = w.v[1];
        out[2] = w.v[2];
        return 0;
    }
    const vec3 r0 = row(lat->reduced, 0), r1 = row(lat->reduced, 1), r2 = row(lat->reduced, 2);
    const double reduced_volume = vdot(r0, vcross(r1, r2));
    const vec3 nearest = fractional_of(d, r0, r1, r2, reduced_volume);
    const long n0 = lround(nearest.v[0]);
    const long n1 = lround(nearest.v[1]);
    const long n2 = lround(nearest.v[2]);
    vec3 best = d;
    double best_n2 = vnorm2(d);
    for (long dn0 = -2; dn0 <= 2; ++dn0) {
        for (long dn1 = -2; dn1 <= 2; ++dn1) {
            for (long dn2 = -2; dn2 <= 2; ++dn2) {
                const double c0 = (double)(n0 + dn0);
                const double c1 = (double)(n1 + dn1);
                const double c2 = (double)(n2 + dn2);
                const vec3 t = vadd(vadd(vscale(r0, c0), vscale(r1, c1)), vscale(r2, c2));
                const vec3 cand = vsub(d, t);
                const double cn2 = vnorm2(cand);
                if (cn2 < best_n2) {
                    best_n2 = cn2;
                    best = cand;
                }
            }
        }
    }
    out[0] = best.v[0];
    out[1] = best.v[1];
    out[2] = best.v[2];
    return 0;
}

/* perpendicularWidth -- lattice.cc:257-284 */
double orc_perpendicular_width(const orc_lattice* lat, int axis) {
    const vec3 a0 = row(lat->m, 0), a1 = row(lat->m, 1), a2 = row(lat->m, 2);
    vec3 o0, o1;
    if (axis == 0) {
        o0 = a1;
        o1 = a2;
    } else if (axis == 1) {
        o0 = a2;
        o1 = a0;
    } else {
        o0 = a0;
        o1 = a1;
    }
    const double face = sqrt(vnorm2(vcross(o0, o1)));
    if (face < ORC_DEGENERATE) return -1.0;
    return fabs(orc_lattice_volume(lat)) / face;
}

/* validateCutoff -- lattice.cc:286-300 */
int orc_validate_cutoff(const orc_lattice* lat, double cutoff) {
    for (int axis = 0; axis < 3; ++axis) {
        const double h = orc_perpendicular_width(lat, axis);
        if (cutoff > 0.5 * h) return axis + 1;
    }
    return 0;
}

/* ======================================================================
 * Supercell -- supercell.cc:7-60
 * ====================================================================== */
void orc_min_supercell_replication(const orc_lattice* lat, double cutoff, int n[3]) {
    for (int axis = 0; axis < 3; ++axis) {
        n[axis] = 1;
        const double width = orc_perpendicular_width(lat, axis);
        while ((double)n[axis] * width <= 2.0 * cutoff) ++n[axis];
    }
}

static int particles_alloc(orc_particles* p, size_t n) {
    p->n = n;
    p->x = (double*)calloc(n ? n : 1, sizeof(double));
    p->y = (double*)calloc(n ? n : 1, sizeof(double));
    p->z = (double*)calloc(n ? n : 1, sizeof(double));
    p->q = (double*)calloc(n ? n : 1, sizeof(double));
    p->species = (uint32_t*)calloc(n ? n : 1, sizeof(uint32_t));
    return (p->x && p->y && p->z && p->q && p->species) ? 0 : -1;
}

void orc_particles_free(orc_particles* p) {
    free(p->x);
    free(p->y);
    free(p->z);
    free(p->q);
    free(p->species);
    memset(p, 0, sizeof(*p));
}

int orc_replicate_supercell(const orc_particles* in, const orc_lattice* lat, int nx, int ny, int nz,
                            orc_particles* out, double out_m[3][3]) {
    if (nx < 1 || ny < 1 || nz < 1) return -1;
    const double(*m)[3] = lat->m;
    if (particles_alloc(out, in->n * (size_t)nx * (size_t)ny * (size_t)nz)) return -1;
    size_t w = 0;
    for (int i = 0; i < nx; ++i) {
        for (int j = 0; j < ny; ++j) {
            for (int k = 0; k < nz; ++k) {
                const double ox = i * m[0][0] + j * m[1][0] + k * m[2][0];
                const double oy = i * m[0][1] + j * m[1][1] + k * m[2][1];
                const double oz = i * m[0][2] + j * m[1][2] + k * m[2][2];
                for (size_t p = 0; p < in->n; ++p, ++w) {
                    out->x[w] = in->x[p] + ox;
                    out->y[w] = in->y[p] + oy;
                    out->z[w] = in->z[p] + oz;
                    out->q[w] = in->q[p];
                    out->species[w] = in->species[p];
                }
            }
        }
    }
    for (int d = 0; d < 3; ++d) {
        out_m[0][d] = nx * m[0][d];
        out_m[1][d] = ny * m[1][d];
        out_m[2][d] = nz * m[2][d];
    }
    return 0;
}

/* ======================================================================
 * Lennard-Jones -- lennard_jones.cc
 * ====================================================================== */
/* mixedParameters -- lennard_jones.cc:49-68 */
void orc_lj_mixed(const orc_lj* lj, uint32_t si, uint32_t sj, double* eps, double* sigma) {
    if (si == sj) {
        *eps = lj->epsilon[si];
        *sigma = lj->sigma[si];
        return;
    }
    *eps = sqrt(lj->epsilon[si] * lj->epsilon[sj]);
    if (lj->mixing == ORC_MIX_LB)
        *sigma = 0.5 * (lj->sigma[si] + lj->sigma[sj]);
    else
        *sigma = sqrt(lj->sigma[si] * lj->sigma[sj]);
}

/* rawPotential / rawDerivative -- lennard_jones.cc:15-30 */
static double raw_potential(double r, double eps, double sigma) {
    const double sr = sigma / r;
    const double sr6 = sr * sr * sr * sr * sr * sr;
    const double sr12 = sr6 * sr6;
    return 4.0 * eps * (sr12 - sr6);
}

static double raw_derivative(double r, double eps, double sigma) {
    const double sr = sigma / r;
    const double sr6 = sr * sr * sr * sr * sr * sr;
    const double sr12 = sr6 * sr6;
    return (24.0 * eps / r) * (sr6 - 2.0 * sr12);
}

/* pairEnergy -- lennard_jones.cc:70-86 */
double orc_lj_pair_energy(const orc_lj* lj, double r, double eps, double sigma) {
    if (r > lj->cutoff) return 0.0;
    switch (lj->truncation) {
        case ORC_TRUNCATED:
            return raw_potential(r, eps, sigma);
        case ORC_SHIFTED:
            return raw_potential(r, eps, sigma) - raw_potential(lj->cutoff, eps, sigma);
        case ORC_LFS: {
            const double v_rc = raw_potential(lj->cutoff, eps, sigma);
            const double dv_rc = raw_derivative(lj->cutoff, eps, sigma);
            return raw_potential(r, eps, sigma) - v_rc - (r - lj->cutoff) * dv_rc;
        }
    }
    return 0.0;
}

/* pairForceOverR -- lennard_jones.cc:88-104 */
static double lj_pair_force_over_r(const orc_lj* lj, double r, double eps, double sigma) {
    if (r > lj->cutoff) return 0.0;
    double dvdr = raw_derivative(r, eps, sigma);
    if (lj->truncation == ORC_LFS) dvdr = raw_derivative(r, eps, sigma) - raw_derivative(lj->cutoff, eps, sigma);
    return -dvdr / r;
}

static double norm3(const double d[3]) { return sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]); }

static int is_excluded(size_t j, const size_t* ex, size_t nex) {
    for (size_t e = 0; e < nex; ++e)
        if (ex[e] == j) return 1;
    return 0;
}

/* computeParticleEnergy -- lennard_jones.cc:158-182 */
double orc_lj_particle_energy(const orc_lj* lj, size_t index, const orc_particles* p,
                              const orc_lattice* lat, const size_t* excluded, size_t n_excluded) {
    double energy = 0.0;
    const double pi[3] = {p->x[index], p->y[index], p->z[index]};
    for (size_t j = 0; j < p->n; ++j) {
        if (j == index) continue;
        if (is_excluded(j, excluded, n_excluded)) continue;
        double eps, sig;
        orc_lj_mixed(lj, p->species[index], p->species[j], &eps, &sig);
        const double pj[3] = {p->x[j], p->y[j], p->z[j]};
        double d[3];
        orc_minimum_image(lat, pi, pj, d);
        energy += orc_lj_pair_energy(lj, norm3(d), eps, sig);
    }
    return energy;
}

/* tailEnergyCorrection -- lennard_jones.cc:218-257 */
double orc_lj_tail_energy(const orc_lj* lj, const orc_particles* p, const orc_lattice* lat) {
    const double volume = fabs(orc_lattice_volume(lat));
    const double rc = lj->cutoff;
    size_t* counts = (size_t*)calloc(lj->n_species, sizeof(size_t));
    for (size_t i = 0; i < p->n; ++i) counts[p->species[i]] += 1;
    double total = 0.0;
    for (uint32_t a = 0; a < lj->n_species; ++a) {
        if (counts[a] == 0) continue;
        for (uint32_t b = 0; b < lj->n_species; ++b) {
            if (counts[b] == 0) continue;
            double eps, sig;
            orc_lj_mixed(lj, a, b, &eps, &sig);
            const double sr3 = sig * sig * sig;
            const double sr9 = sr3 * sr3 * sr3 / (rc * rc * rc * rc * rc * rc * rc * rc * rc);
            const double sr3_over_rc3 = sr3 / (rc * rc * rc);
            const double integral =
                eps * sr3 * ((4.0 / 9.0) * sr9 - (4.0 / 3.0) * sr3_over_rc3);
            total += (double)counts[a] * (double)counts[b] * integral;
        }
    }
    free(counts);
    return (2.0 * ORC_PI / volume) * total;
}

double orc_lj_total_energy(const orc_lj* lj, const orc_particles* p, const orc_lattice* lat) {
    double e = 0.0;
    for (size_t i = 0; i < p->n; ++i) {
        const double pi[3] = {p->x[i], p->y[i], p->z[i]};
        for (size_t j = i + 1; j < p->n; ++j) {
            double eps, sig, d[3];
            orc_lj_mixed(lj, p->species[i], p->species[j], &eps, &sig);
            const double pj[3] = {p->x[j], p->y[j], p->z[j]};
            orc_minimum_image(lat, pi, pj, d);
            e += orc_lj_pair_energy(lj, norm3(d), eps, sig);
        }
    }
    return e;
}

/* computeVirial -- lennard_jones.cc:113-127 */
double orc_lj_total_virial(const orc_lj* lj, const orc_particles* p, const orc_lattice* lat) {
    double w = 0.0;
    for (size_t i = 0; i < p->n; ++i) {
        const double pi[3] = {p->x[i], p->y[i], p->z[i]};
        for (size_t j = i + 1; j < p->n; ++j) {
            double eps, sig, d[3];
            orc_lj_mixed(lj, p->species[i], p->species[j], &eps, &sig);
            const double pj[3] = {p->x[j], p->y[j], p->z[j]};
            orc_minimum_image(lat, pi, pj, d);
            const double r = norm3(d);
            w += lj_pair_force_over_r(lj, r, eps, sig) * r * r;
        }
    }
    return w;
}

/* ======================================================================
 * Ewald -- ewald.cc
 * ====================================================================== */
/* reciprocalVectors + buildReciprocalTerms -- ewald.cc:33-45, 137-161 */
int orc_reciprocal_terms(const orc_lattice* lat, double alpha, int kmax, orc_kset* out) {
    const vec3 a0 = row(lat->m, 0), a1 = row(lat->m, 1), a2 = row(lat->m, 2);
    const double volume = orc_lattice_volume(lat);
    const double two_pi_over_v = 2.0 * ORC_PI / volume;
    const vec3 b0 = vscale(vcross(a1, a2), two_pi_over_v);
    const vec3 b1 = vscale(vcross(a2, a0), two_pi_over_v);
    const vec3 b2 = vscale(vcross(a0, a1), two_pi_over_v);
    const double four_alpha_sq = 4.0 * alpha * alpha;
    const int bound = kmax * kmax + 2;
    const size_t cap = (size_t)(2 * kmax + 1) * (2 * kmax + 1) * (2 * kmax + 1);
    out->k = (double*)malloc(cap * 3 * sizeof(double));
    out->weight = (double*)malloc(cap * sizeof(double));
    out->nvec = (int32_t*)malloc(cap * 3 * sizeof(int32_t));
    if (!out->k || !out->weight || !out->nvec) return -1;
    size_t n = 0;
    for (int n0 = -kmax; n0 <= kmax; ++n0) {
        for (int n1 = -kmax; n1 <= kmax; ++n1) {
            for (int n2 = -kmax; n2 <= kmax; ++n2) {
                if (n0 == 0 && n1 == 0 && n2 == 0) continue;
                if (n0 * n0 + n1 * n1 + n2 * n2 >= bound) continue;
                const vec3 k = vadd(vadd(vscale(b0, n0), vscale(b1, n1)), vscale(b2, n2));
                const double k2 = vdot(k, k);
                out->k[3 * n + 0] = k.v[0];
                out->k[3 * n + 1] = k.v[1];
                out->k[3 * n + 2] = k.v[2];
                out->weight[n] = exp(-k2 / four_alpha_sq) / k2;
                out->nvec[3 * n + 0] = n0;
                out->nvec[3 * n + 1] = n1;
                out->nvec[3 * n + 2] = n2;
                ++n;
            }
        }
    }
    out->n = n;
    return 0;
}

void orc_kset_free(orc_kset* ks) {
    free(ks->k);
    free(ks->weight);
    free(ks->nvec);
    memset(ks, 0, sizeof(*ks));
}

/* erfOverR -- ewald.cc:69-79 / ewald_incremental_state.cc:24-30 */
static double erf_over_r(double r, double alpha) {
    if (r < 1e-8) return 2.0 * alpha / sqrt(ORC_PI);
    return erf(alpha * r) / r;
}

/* realSpaceParticleEnergy -- ewald.cc:223-249 */
double orc_ewald_real_particle_energy(double alpha, double rc, size_t index, const orc_particles* p,
                                      const orc_lattice* lat, const size_t* excluded,
                                      size_t n_excluded) {
    double energy = 0.0;
    const double pi[3] = {p->x[index], p->y[index], p->z[index]};
    for (size_t j = 0; j < p->n; ++j) {
        if (j == index) continue;
        if (is_excluded(j, excluded, n_excluded)) continue;
        const double pj[3] = {p->x[j], p->y[j], p->z[j]};
        double d[3];
        orc_minimum_image(lat, pi, pj, d);
        const double r = norm3(d);
        if (r > rc) continue;
        energy += p->q[index] * p->q[j] * (erfc(alpha * r) / r);
    }
    return energy * ORC_KCOULOMB;
}

/* computeEnergyBreakdown -- ewald.cc:251-313 (real space over i<j in index order) */
int orc_ewald_breakdown_energy(double alpha, double rc, int kmax, const orc_particles* p,
                               const orc_lattice* lat, const size_t* excl_pairs, size_t n_excl,
                               orc_ewald_breakdown* out) {
    double total_q = 0.0;
    for (size_t i = 0; i < p->n; ++i) total_q += p->q[i];
    if (fabs(total_q) > 1e-8) return -1;
    double u_real = 0.0;
    for (size_t i = 0; i < p->n; ++i) {
        const double pi[3] = {p->x[i], p->y[i], p->z[i]};
        for (size_t j = i + 1; j < p->n; ++j) {
            int skip = 0;
            for (size_t e = 0; e < n_excl; ++e) {
                const size_t a = excl_pairs[2 * e], b = excl_pairs[2 * e + 1];
                if ((a == i && b == j) || (a == j && b == i)) skip = 1;
            }
            if (skip) continue;
            const double pj[3] = {p->x[j], p->y[j], p->z[j]};
            double d[3];
            orc_minimum_image(lat, pi, pj, d);
            const double r = norm3(d);
            if (r > rc) continue;
            u_real += p->q[i] * p->q[j] * (erfc(alpha * r) / r);
        }
    }
    out->real = u_real * ORC_KCOULOMB;
    const double volume = fabs(orc_lattice_volume(lat));
    orc_kset ks;
    if (orc_reciprocal_terms(lat, alpha, kmax, &ks)) return -1;
    double u_recip = 0.0;
    for (size_t t = 0; t < ks.n; ++t) {
        const double* k = ks.k + 3 * t;
        double re = 0.0, im = 0.0;
        for (size_t i = 0; i < p->n; ++i) {
            const double kr = k[0] * p->x[i] + k[1] * p->y[i] + k[2] * p->z[i];
            re += p->q[i] * cos(kr);
            im += p->q[i] * sin(kr);
        }
        u_recip += ks.weight[t] * (re * re + im * im);
    }
    orc_kset_free(&ks);
    out->reciprocal = u_recip * ORC_KCOULOMB * (2.0 * ORC_PI / volume);
    double q2 = 0.0;
    for (size_t i = 0; i < p->n; ++i) q2 += p->q[i] * p->q[i];
    out->self = -ORC_KCOULOMB * (alpha / sqrt(ORC_PI)) * q2;
    double u_excl = 0.0;
    for (size_t e = 0; e < n_excl; ++e) {
        const size_t i = excl_pairs[2 * e], j = excl_pairs[2 * e + 1];
        const double pi[3] = {p->x[i], p->y[i], p->z[i]};
        const double pj[3] = {p->x[j], p->y[j], p->z[j]};
        double d[3];
        orc_minimum_image(lat, pi, pj, d);
        u_excl += p->q[i] * p->q[j] * erf_over_r(norm3(d), alpha);
    }
    out->exclusion = -u_excl * ORC_KCOULOMB;
    out->surface = 0.0;
    return 0;
}

/* ======================================================================
 * EwaldIncrementalState -- ewald_incremental_state.cc:55-257
 * ====================================================================== */
static double site_q2(const orc_site* s, size_t n) {
    double sum = 0.0;
    for (size_t i = 0; i < n; ++i) sum += s[i].charge * s[i].charge;
    return sum;
}

/* pairExclusionSum -- ewald_incremental_state.cc:32-43 */
static double pair_exclusion_sum(const orc_site* s, size_t n, const orc_lattice* lat, double alpha) {
    double sum = 0.0;
    for (size_t i = 0; i < n; ++i) {
        for (size_t j = i + 1; j < n; ++j) {
            const double a[3] = {s[i].x, s[i].y, s[i].z};
            const double b[3] = {s[j].x, s[j].y, s[j].z};
            double d[3];
            orc_minimum_image(lat, a, b, d);
            sum += s[i].charge * s[j].charge * erf_over_r(norm3(d), alpha);
        }
    }
    return sum;
}

static void structure_factors(const orc_kset* ks, const orc_particles* p, double* re_out,
                              double* im_out) {
    for (size_t t = 0; t < ks->n; ++t) {
        const double* k = ks->k + 3 * t;
        double re = 0.0, im = 0.0;
        for (size_t i = 0; i < p->n; ++i) {
            const double kr = k[0] * p->x[i] + k[1] * p->y[i] + k[2] * p->z[i];
            re += p->q[i] * cos(kr);
            im += p->q[i] * sin(kr);
        }
        re_out[t] = re;
        im_out[t] = im;
    }
}

/* constructor -- ewald_incremental_state.cc:55-83 */
int orc_ewald_state_init(orc_ewald_state* st, double alpha, int kmax, const orc_particles* p,
                         const orc_lattice* lat) {
    memset(st, 0, sizeof(*st));
    st->lat = *lat;
    st->alpha = alpha;
    if (orc_reciprocal_terms(lat, alpha, kmax, &st->terms)) return -1;
    st->volume = fabs(orc_lattice_volume(lat));
    st->s_re = (double*)calloc(st->terms.n, sizeof(double));
    st->s_im = (double*)calloc(st->terms.n, sizeof(double));
    if (!st->s_re || !st->s_im) return -1;
    structure_factors(&st->terms, p, st->s_re, st->s_im);
    double q2 = 0.0;
    for (size_t i = 0; i < p->n; ++i) q2 += p->q[i] * p->q[i];
    st->self_energy = -ORC_KCOULOMB * (alpha / sqrt(ORC_PI)) * q2;
    st->exclusion_energy = 0.0;
    return 0;
}

void orc_ewald_state_free(orc_ewald_state* st) {
    orc_kset_free(&st->terms);
    free(st->s_re);
    free(st->s_im);
    st->s_re = st->s_im = NULL;
}

/* committedEnergy -- ewald_incremental_state.cc:85-93 */
double orc_ewald_state_committed(const orc_ewald_state* st) {
    double sum = 0.0;
    for (size_t t = 0; t < st->terms.n; ++t)
        sum += st->terms.weight[t] * (st->s_re[t] * st->s_re[t] + st->s_im[t] * st->s_im[t]);
    const double reciprocal = sum * ORC_KCOULOMB * (2.0 * ORC_PI / st->volume);
    return reciprocal + st->self_energy + st->exclusion_energy;
}

/* computeDelta -- ewald_incremental_state.cc:95-136 */
double orc_ewald_state_delta(const orc_ewald_state* st, const orc_site* removed, size_t nr,
                             const orc_site* added, size_t na, double* d_re, double* d_im,
                             double* recip_out, double* self_out, double* excl_out) {
    const double prefactor = ORC_KCOULOMB * (2.0 * ORC_PI / st->volume);
    double wsum = 0.0;
    for (size_t t = 0; t < st->terms.n; ++t) {
        const double* k = st->terms.k + 3 * t;
        double dre = 0.0, dim = 0.0;
        for (size_t s = 0; s < na; ++s) {
            const double kr = k[0] * added[s].x + k[1] * added[s].y + k[2] * added[s].z;
            dre += added[s].charge * cos(kr);
            dim += added[s].charge * sin(kr);
        }
        for (size_t s = 0; s < nr; ++s) {
            const double kr = k[0] * removed[s].x + k[1] * removed[s].y + k[2] * removed[s].z;
            dre -= removed[s].charge * cos(kr);
            dim -= removed[s].charge * sin(kr);
        }
        const double new_re = st->s_re[t] + dre;
        const double new_im = st->s_im[t] + dim;
        const double old_mag = st->s_re[t] * st->s_re[t] + st->s_im[t] * st->s_im[t];
        const double new_mag = new_re * new_re + new_im * new_im;
        wsum += st->terms.weight[t] * (new_mag - old_mag);
        if (d_re) d_re[t] = dre;
        if (d_im) d_im[t] = dim;
    }
    const double recip = prefactor * wsum;
    const double self_pref = -ORC_KCOULOMB * (st->alpha / sqrt(ORC_PI));
    const double self = self_pref * (site_q2(added, na) - site_q2(removed, nr));
    const double add_ex = pair_exclusion_sum(added, na, &st->lat, st->alpha);
    const double rem_ex = pair_exclusion_sum(removed, nr, &st->lat, st->alpha);
    const double excl = -ORC_KCOULOMB * (add_ex - rem_ex);
    if (recip_out) *recip_out = recip;
    if (self_out) *self_out = self;
    if (excl_out) *excl_out = excl;
    return recip + self + excl;
}

/* apply + commit* -- ewald_incremental_state.cc:138-172 */
void orc_ewald_state_commit(orc_ewald_state* st, const orc_site* removed, size_t nr,
                            const orc_site* added, size_t na) {
    double* dre = (double*)malloc(st->terms.n * sizeof(double));
    double* dim = (double*)malloc(st->terms.n * sizeof(double));
    double recip, self, excl;
    orc_ewald_state_delta(st, removed, nr, added, na, dre, dim, &recip, &self, &excl);
    for (size_t t = 0; t < st->terms.n; ++t) {
        st->s_re[t] += dre[t];
        st->s_im[t] += dim[t];
    }
    st->self_energy += self;
    st->exclusion_energy += excl;
    free(dre);
    free(dim);
}

static size_t gather_molecule(const orc_particles* p, const size_t* off, const size_t* sites,
                              size_t m, orc_site* buf) {
    size_t n = 0;
    for (size_t s = off[m]; s < off[m + 1]; ++s, ++n) {
        const size_t i = sites[s];
        buf[n].x = p->x[i];
        buf[n].y = p->y[i];
        buf[n].z = p->z[i];
        buf[n].charge = p->q[i];
    }
    return n;
}

static size_t max_molecule(const size_t* off, size_t n_mol) {
    size_t mx = 1;
    for (size_t m = 0; m < n_mol; ++m)
        if (off[m + 1] - off[m] > mx) mx = off[m + 1] - off[m];
    return mx;
}

/* resync -- ewald_incremental_state.cc:214-247 */
void orc_ewald_state_resync(orc_ewald_state* st, const orc_particles* p, const size_t* off,
                            const size_t* sites, size_t n_mol) {
    structure_factors(&st->terms, p, st->s_re, st->s_im);
    double q2 = 0.0;
    for (size_t i = 0; i < p->n; ++i) q2 += p->q[i] * p->q[i];
    st->self_energy = -ORC_KCOULOMB * (st->alpha / sqrt(ORC_PI)) * q2;
    orc_site* buf = (orc_site*)malloc(max_molecule(off, n_mol) * sizeof(orc_site));
    double exclusion = 0.0;
    for (size_t m = 0; m < n_mol; ++m) {
        const size_t n = gather_molecule(p, off, sites, m, buf);
        exclusion += pair_exclusion_sum(buf, n, &st->lat, st->alpha);
    }
    free(buf);
    st->exclusion_energy = -ORC_KCOULOMB * exclusion;
}

/* recomputeFromScratch -- ewald_incremental_state.cc:174-212 */
double orc_ewald_state_recompute(const orc_ewald_state* st, const orc_particles* p,
                                 const size_t* off, const size_t* sites, size_t n_mol) {
    double rsum = 0.0;
    for (size_t t = 0; t < st->terms.n; ++t) {
        const double* k = st->terms.k + 3 * t;
        double re = 0.0, im = 0.0;
        for (size_t i = 0; i < p->n; ++i) {
            const double kr = k[0] * p->x[i] + k[1] * p->y[i] + k[2] * p->z[i];
            re += p->q[i] * cos(kr);
            im += p->q[i] * sin(kr);
        }
        rsum += st->terms.weight[t] * (re * re + im * im);
    }
    const double reciprocal = rsum * ORC_KCOULOMB * (2.0 * ORC_PI / st->volume);
    double q2 = 0.0;
    for (size_t i = 0; i < p->n; ++i) q2 += p->q[i] * p->q[i];
    const double self = -ORC_KCOULOMB * (st->alpha / sqrt(ORC_PI)) * q2;
    orc_site* buf = (orc_site*)malloc(max_molecule(off, n_mol) * sizeof(orc_site));
    double exclusion = 0.0;
    for (size_t m = 0; m < n_mol; ++m) {
        const size_t n = gather_molecule(p, off, sites, m, buf);
        exclusion += -ORC_KCOULOMB * pair_exclusion_sum(buf, n, &st->lat, st->alpha);
    }
    free(buf);
    return reciprocal + self + exclusion;
}

/* rawStateForTesting -- ewald_incremental_state.cc:249-257 */
void orc_ewald_state_raw(const orc_ewald_state* st, double* out) {
    const size_t n = st->terms.n;
    memcpy(out, st->s_re, n * sizeof(double));
    memcpy(out + n, st->s_im, n * sizeof(double));
    out[2 * n] = st->self_energy;
    out[2 * n + 1] = st->exclusion_energy;
}

/* ======================================================================
 * FrameworkEnergyGrid -- framework_energy_grid.cc:193-288
 * ====================================================================== */
/* point counts -- framework_energy_grid.cc:207-211 */
void orc_grid_point_counts(const orc_lattice* lat, double spacing, int32_t n[3]) {
    for (int axis = 0; axis < 3; ++axis) {
        const double width = orc_perpendicular_width(lat, axis);
        const int v = (int)lround(width / spacing);
        n[axis] = v < 1 ? 1 : v;
    }
}

static double capped(double e, double cap) { return (cap < e) ? cap : e; }

/* constructor body -- framework_energy_grid.cc:221-248 */
int orc_grid_build(const orc_lj* lj, const orc_particles* fw, const orc_lattice* lat,
                   const uint32_t* guests, int n_guest, const int32_t n[3], double cap, double* out) {
    const size_t total = (size_t)n[0] * (size_t)n[1] * (size_t)n[2];
    for (int g = 0; g < n_guest; ++g) {
        double* table = out + (size_t)g * total;
        for (int iz = 0; iz < n[2]; ++iz) {
            for (int iy = 0; iy < n[1]; ++iy) {
                for (int ix = 0; ix < n[0]; ++ix) {
                    const double f[3] = {(double)ix / (double)n[0], (double)iy / (double)n[1],
                                         (double)iz / (double)n[2]};
                    double cart[3];
                    orc_frac_to_cart(lat, f, cart);
                    double energy = 0.0;
                    for (size_t a = 0; a < fw->n; ++a) {
                        double eps, sig, d[3];
                        orc_lj_mixed(lj, guests[g], fw->species[a], &eps, &sig);
                        const double pa[3] = {fw->x[a], fw->y[a], fw->z[a]};
                        orc_minimum_image(lat, cart, pa, d);
                        const double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
                        energy += orc_lj_pair_energy(lj, r, eps, sig);
                    }
                    table[((size_t)iz * (size_t)n[1] + (size_t)iy) * (size_t)n[0] + (size_t)ix] =
                        capped(energy, cap);
                }
            }
        }
    }
    return 0;
}

static void node_cart(const double gm[3][3], const int32_t n[3], int64_t node, double cart[3]) {
    const int64_t ix = node % n[0];
    const int64_t iy = (node / n[0]) % n[1];
    const int64_t iz = node / ((int64_t)n[0] * n[1]);
    const double f[3] = {(double)ix / (double)n[0], (double)iy / (double)n[1],
                         (double)iz / (double)n[2]};
    cart[0] = f[0] * gm[0][0] + f[1] * gm[1][0] + f[2] * gm[2][0];
    cart[1] = f[0] * gm[0][1] + f[1] * gm[1][1] + f[2] * gm[2][1];
    cart[2] = f[0] * gm[0][2] + f[1] * gm[1][2] + f[2] * gm[2][2];
}

int orc_grid_nodes_lj(const orc_lj* lj, const orc_particles* fw, const orc_lattice* lat,
                      const double grid_m[3][3], const int32_t n[3], uint32_t guest,
                      const int64_t* nodes, size_t n_nodes, double cap, double* out) {
    for (size_t i = 0; i < n_nodes; ++i) {
        double cart[3];
        node_cart(grid_m, n, nodes[i], cart);
        double energy = 0.0;
        for (size_t a = 0; a < fw->n; ++a) {
            double eps, sig, d[3];
            orc_lj_mixed(lj, guest, fw->species[a], &eps, &sig);
            const double pa[3] = {fw->x[a], fw->y[a], fw->z[a]};
            orc_minimum_image(lat, cart, pa, d);
            const double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
            energy += orc_lj_pair_energy(lj, r, eps, sig);
        }
        out[i] = capped(energy, cap);
    }
    return 0;
}

int orc_grid_nodes_coulomb(const orc_ewald_state* st, double rc_real, const orc_particles* fw,
                           const double grid_m[3][3], const int32_t n[3], const int64_t* nodes,
                           size_t n_nodes, double* out_real, double* out_recip) {
    const double pref = ORC_KCOULOMB * (4.0 * ORC_PI / st->volume);
    for (size_t i = 0; i < n_nodes; ++i) {
        double cart[3];
        node_cart(grid_m, n, nodes[i], cart);
        double real = 0.0;
        for (size_t a = 0; a < fw->n; ++a) {
            double d[3];
            const double pa[3] = {fw->x[a], fw->y[a], fw->z[a]};
            orc_minimum_image(&st->lat, cart, pa, d);
            const double r = norm3(d);
            if (r > rc_real) continue;
            real += 1.0 * fw->q[a] * (erfc(st->alpha * r) / r);
        }
        out_real[i] = real * ORC_KCOULOMB;
        double rec = 0.0;
        for (size_t t = 0; t < st->terms.n; ++t) {
            const double* k = st->terms.k + 3 * t;
            const double kr = k[0] * cart[0] + k[1] * cart[1] + k[2] * cart[2];
            rec += st->terms.weight[t] * (st->s_re[t] * cos(kr) + st->s_im[t] * sin(kr));
        }
        out_recip[i] = pref * rec;
    }
    return 0;
}

/* interpolate -- framework_energy_grid.cc:251-288 */
double orc_grid_interpolate(const orc_lattice* lat, const int32_t n[3], const double* table,
                            const double pos[3]) {
    double f[3];
    orc_cart_to_frac(lat, pos, f);
    for (int a = 0; a < 3; ++a) f[a] -= floor(f[a]);
    int i0[3], i1[3];
    double t[3];
    for (int a = 0; a < 3; ++a) {
        const double g = f[a] * (double)n[a];
        const int base = (int)floor(g);
        t[a] = g - (double)base;
        const int wrapped = ((base % n[a]) + n[a]) % n[a];
        i0[a] = wrapped;
        i1[a] = (wrapped + 1) % n[a];
    }
#define V(ix, iy, iz) table[((size_t)(iz) * (size_t)n[1] + (size_t)(iy)) * (size_t)n[0] + (size_t)(ix)]
    const double c00 = V(i0[0], i0[1], i0[2]) * (1.0 - t[0]) + V(i1[0], i0[1], i0[2]) * t[0];
    const double c10 = V(i0[0], i1[1], i0[2]) * (1.0 - t[0]) + V(i1[0], i1[1], i0[2]) * t[0];
    const double c01 = V(i0[0], i0[1], i1[2]) * (1.0 - t[0]) + V(i1[0], i0[1], i1[2]) * t[0];
    const double c11 = V(i0[0], i1[1], i1[2]) * (1.0 - t[0]) + V(i1[0], i1[1], i1[2]) * t[0];
#undef V
    const double c0 = c00 * (1.0 - t[1]) + c10 * t[1];
    const double c1 = c01 * (1.0 - t[1]) + c11 * t[1];
    return c0 * (1.0 - t[2]) + c1 * t[2];
}

/* ======================================================================
 * Widom -- monte_carlo_engine.cc:850-917 (helpers :56-72)
 * Draw order: the reference writes the draws as nested call arguments; this
 * restatement takes them left to right (the order clang evaluates them in).
 * ====================================================================== */
typedef struct {
    double w, x, y, z;
} quat;

/* randomUnitQuaternion -- monte_carlo_engine.cc:56-63 */
static quat random_unit_quaternion(double u0, double u1, double u2) {
    const double r1 = sqrt(1.0 - u0);
    const double r2 = sqrt(u0);
    const double th1 = 2.0 * ORC_PI * u1;
    const double th2 = 2.0 * ORC_PI * u2;
    quat q = {r2 * cos(th2), r1 * sin(th1), r1 * cos(th1), r2 * sin(th2)};
    return q;
}

/* rotateByQuaternion -- monte_carlo_engine.cc:68-72 */
static vec3 rotate_by_quaternion(vec3 v, quat q) {
    const vec3 qv = v3(q.x, q.y, q.z);
    const vec3 t = vscale(vcross(qv, v), 2.0);
    return vadd(vadd(v, vscale(t, q.w)), vcross(qv, t));
}

double orc_widom(const orc_lj* lj, const orc_particles* fw, const orc_lattice* lat,
                 const orc_mol_site* sites, size_t n_sites, const orc_ewald_state* ewald_state,
                 double rc_real, uint64_t seed, uint64_t stream, uint64_t first_call,
                 uint64_t n_trials, double temperature, double* du_out) {
    orc_particles work;
    particles_alloc(&work, fw->n + n_sites);
    memcpy(work.x, fw->x, fw->n * sizeof(double));
    memcpy(work.y, fw->y, fw->n * sizeof(double));
    memcpy(work.z, fw->z, fw->n * sizeof(double));
    memcpy(work.q, fw->q, fw->n * sizeof(double));
    memcpy(work.species, fw->species, fw->n * sizeof(uint32_t));
    size_t* trial_idx = (size_t*)malloc(n_sites * sizeof(size_t));
    orc_site* charged = (orc_site*)malloc(n_sites * sizeof(orc_site));
    const vec3 a0 = row(lat->m, 0), a1 = row(lat->m, 1), a2 = row(lat->m, 2);
    orc_rng rng;
    orc_rng_seed(&rng, seed, stream);
    rng.call_index = first_call;
    double sum_boltzmann = 0.0;
    for (uint64_t trial = 0; trial < n_trials; ++trial) {
        const double u0 = orc_rng_next_uniform(&rng);
        const double u1 = orc_rng_next_uniform(&rng);
        const double u2 = orc_rng_next_uniform(&rng);
        const vec3 ref = vadd(vadd(vscale(a0, u0), vscale(a1, u1)), vscale(a2, u2));
        quat q = {1.0, 0.0, 0.0, 0.0};
        if (n_sites > 1) {
            const double q0 = orc_rng_next_uniform(&rng);
            const double q1 = orc_rng_next_uniform(&rng);
            const double q2 = orc_rng_next_uniform(&rng);
            q = random_unit_quaternion(q0, q1, q2);
        }
        work.n = fw->n + n_sites;
        for (size_t s = 0; s < n_sites; ++s) {
            const vec3 local = v3(sites[s].x, sites[s].y, sites[s].z);
            const vec3 rot = n_sites > 1 ? rotate_by_quaternion(local, q) : local;
            const vec3 wp = vadd(ref, rot);
            const size_t idx = fw->n + s;
            work.x[idx] = wp.v[0];
            work.y[idx] = wp.v[1];
            work.z[idx] = wp.v[2];
            work.q[idx] = sites[s].charge;
            work.species[idx] = sites[s].lj_species;
            trial_idx[s] = idx;
        }
        double du = 0.0;
        for (size_t s = 0; s < n_sites; ++s)
            du += orc_lj_particle_energy(lj, trial_idx[s], &work, lat, trial_idx, n_sites);
        if (ewald_state) {
            double real = 0.0;
            for (size_t s = 0; s < n_sites; ++s)
                real += orc_ewald_real_particle_energy(ewald_state->alpha, rc_real, trial_idx[s],
                                                       &work, lat, trial_idx, n_sites);
            du += real;
            for (size_t s = 0; s < n_sites; ++s) {
                charged[s].x = work.x[trial_idx[s]];
                charged[s].y = work.y[trial_idx[s]];
                charged[s].z = work.z[trial_idx[s]];
                charged[s].charge = work.q[trial_idx[s]];
            }
            du += orc_ewald_state_delta(ewald_state, NULL, 0, charged, n_sites, NULL, NULL, NULL,
                                        NULL, NULL);
        }
        if (du_out) du_out[trial] = du;
        sum_boltzmann += exp(-du / temperature);
    }
    work.n = fw->n + n_sites;
    orc_particles_free(&work);
    free(trial_idx);
    free(charged);
    return (sum_boltzmann / (double)n_trials) / temperature;
}

/* ======================================================================
 * Synthetic frameworks (no reference counterpart).
 * The rejection test uses the fractional-rounding minimum image, which is exact
 * for distances below half the smallest perpendicular width; the product
 * generator makes the identical decisions with the identical arithmetic.
 * ====================================================================== */
static void inverse3(const double m[3][3], double inv[3][3]) {
    const double det = m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]) -
                       m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]) +
                       m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]);
    inv[0][0] = (m[1][1] * m[2][2] - m[1][2] * m[2][1]) / det;
    inv[0][1] = (m[0][2] * m[2][1] - m[0][1] * m[2][2]) / det;
    inv[0][2] = (m[0][1] * m[1][2] - m[0][2] * m[1][1]) / det;
    inv[1][0] = (m[1][2] * m[2][0] - m[1][0] * m[2][2]) / det;
    inv[1][1] = (m[0][0] * m[2][2] - m[0][2] * m[2][0]) / det;
    inv[1][2] = (m[0][2] * m[1][0] - m[0][0] * m[1][2]) / det;
    inv[2][0] = (m[1][0] * m[2][1] - m[1][1] * m[2][0]) / det;
    inv[2][1] = (m[0][1] * m[2][0] - m[0][0] * m[2][1]) / det;
    inv[2][2] = (m[0][0] * m[1][1] - m[0][1] * m[1][0]) / det;
}

int orc_synth_framework(const double m[3][3], size_t n_atoms, uint64_t seed, uint64_t stream,
                        double min_dist, const double* proportions, uint32_t n_types,
                        double charge_sigma, uint64_t charge_stream, double* x, double* y,
                        double* z, double* q, uint32_t* species) {
    double inv[3][3];
    inverse3(m, inv);
    double* fx = (double*)malloc((n_atoms ? n_atoms : 1) * 3 * sizeof(double));
    orc_rng rng;
    orc_rng_seed(&rng, seed, stream);
    const double md2 = min_dist * min_dist;
    size_t placed = 0;
    uint64_t attempts = 0;
    while (placed < n_atoms) {
        if (++attempts > 1000ull * (uint64_t)n_atoms + 1000ull) {
            free(fx);
            return -1;
        }
        const double f0 = orc_rng_next_uniform(&rng);
        const double f1 = orc_rng_next_uniform(&rng);
        const double f2 = orc_rng_next_uniform(&rng);
        int ok = 1;
        for (size_t j = 0; j < placed && ok; ++j) {
            double d0 = fx[3 * j + 0] - f0, d1 = fx[3 * j + 1] - f1, d2 = fx[3 * j + 2] - f2;
            d0 -= rint(d0);
            d1 -= rint(d1);
            d2 -= rint(d2);
            const double cx = d0 * m[0][0] + d1 * m[1][0] + d2 * m[2][0];
            const double cy = d0 * m[0][1] + d1 * m[1][1] + d2 * m[2][1];
            const double cz = d0 * m[0][2] + d1 * m[1][2] + d2 * m[2][2];
            if (cx * cx + cy * cy + cz * cz < md2) ok = 0;
        }
        if (!ok) continue;
        fx[3 * placed + 0] = f0;
        fx[3 * placed + 1] = f1;
        fx[3 * placed + 2] = f2;
        x[placed] = f0 * m[0][0] + f1 * m[1][0] + f2 * m[2][0];
        y[placed] = f0 * m[0][1] + f1 * m[1][1] + f2 * m[2][1];
        z[placed] = f0 * m[0][2] + f1 * m[1][2] + f2 * m[2][2];
        ++placed;
    }
    (void)inv;
    free(fx);
    /* types in index order: count_t = round(p_t * N) for all but the last type */
    size_t w = 0;
    for (uint32_t t = 0; t < n_types; ++t) {
        size_t cnt;
        if (t + 1 < n_types) {
            cnt = (size_t)llround(proportions[t] * (double)n_atoms);
            if (w + cnt > n_atoms) cnt = n_atoms - w;
        } else {
            cnt = n_atoms - w;
        }
        for (size_t i = 0; i < cnt; ++i) species[w++] = t;
    }
    /* charges: Box-Muller on Philox(seed, charge_stream), then shift to zero mean */
    if (charge_sigma > 0.0) {
        orc_rng crng;
        orc_rng_seed(&crng, seed, charge_stream);
        double mean = 0.0;
        for (size_t i = 0; i < n_atoms; ++i) {
            const double u1 = orc_rng_next_uniform(&crng);
            const double u2 = orc_rng_next_uniform(&crng);
            const double g = sqrt(-2.0 * log(1.0 - u1)) * cos(2.0 * ORC_PI * u2);
            q[i] = charge_sigma * g;
            mean += q[i];
        }
        mean /= (double)n_atoms;
        for (size_t i = 0; i < n_atoms; ++i) q[i] -= mean;
    } else {
        for (size_t i = 0; i < n_atoms; ++i) q[i] = 0.0;
    }
    return 0;
}

/* ======================================================================
 * Peng-Robinson -- peng_robinson.cc:10-110
 * ====================================================================== */
double orc_pr_fugacity(double tc, double pc, double omega, double T, double P) {
    const double r = 8.31446261815324;
    const double a = 0.45724 * r * r * tc * tc / pc;
    const double b = 0.07780 * r * tc / pc;
    const double kappa = 0.37464 + 1.54226 * omega - 0.26992 * omega * omega;
    const double term = 1.0 + kappa * (1.0 - sqrt(T / tc));
    const double alpha_t = term * term;
    const double a_alpha = a * alpha_t;
    const double A = a_alpha * P / (r * r * T * T);
    const double B = b * P / (r * T);
    const double c2 = -(1.0 - B);
    const double c1 = A - 3.0 * B * B - 2.0 * B;
    const double c0 = -(A * B - B * B - B * B * B);
    double zz = 1.0;
    int converged = 0;
    for (int iter = 0; iter < 100; ++iter) {
        const double fz = zz * zz * zz + c2 * zz * zz + c1 * zz + c0;
        if (fabs(fz) < 1e-14) {
            converged = 1;
            break;
        }
        const double fpz = 3.0 * zz * zz + 2.0 * c2 * zz + c1;
        if (fpz == 0.0) break;
        zz -= fz / fpz;
        if (!(zz > B)) zz = B + 0.5 * (1.0 - B > 0 ? 1.0 - B : 1.0);
    }
    if (!converged && fabs(zz * zz * zz + c2 * zz * zz + c1 * zz + c0) > 1e-8) return NAN;
    const double sqrt2 = 1.41421356237309504880;
    const double ln_phi = (zz - 1.0) - log(zz - B) -
                          (A / (2.0 * sqrt2 * B)) *
                              log((zz + (1.0 + sqrt2) * B) / (zz + (1.0 - sqrt2) * B));
    return exp(ln_phi) * P;
}
