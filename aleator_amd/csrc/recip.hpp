// recip.hpp -- reciprocal-space pieces shared by the S(k) state and Widom kernels.
//
// Reference k-set: n = (n0,n1,n2), n0^2+n1^2+n2^2 < kmax^2+2, n != 0, in
// lexicographic order (ewald.cc:137-161).  The first half of that order is the set of
// lexicographically negative n; the second half is its mirror -n in reverse order.
// Because the charges are real, S(-k) = conj S(k) and a single-site phase factor
// obeys the same symmetry, so every sum over the full set equals twice the sum over
// the first half: the device stores and updates only that half.
//
// exp(i k.r) with k = n.B and r = f.A is exp(2 pi i n.f): a product of three per-axis
// factors exp(2 pi i n_a f_a).  A wave stages those factors for each site in LDS
// (sincospi on the fractional coordinate, (2K+1) per axis) and then evaluates every
// k with two complex multiplies per site.
#pragma once

#include "device.hpp"

namespace aleator_gpu {

// Per-wave phase staging: E[(axis*nkk + m)*ns + s] = exp(2 pi i (m-K) f_{s,axis}).
// `site4` holds ns sites as (f0, f1, f2, q).  Called by the 64 lanes of the wave;
// the caller places a workgroup barrier before E is read.
__device__ inline void stage_phases(int lane, int ns, int K, const double* site4, cplx* E) {
    const int nkk = 2 * K + 1;
    const int total = 3 * nkk * ns;
    for (int i = lane; i < total; i += 64) {
        const int s = i % ns;
        const int m = (i / ns) % nkk;
        const int axis = i / (ns * nkk);
        double sv, cv;
        sincospi(2.0 * (double)(m - K) * site4[4 * s + axis], &sv, &cv);
        E[i] = {cv, sv};
    }
}

// sum over the half k-set of w * (|S + dS|^2 - |S|^2), dS = sum_s q_s exp(i k.r_s),
// by one wave; sites [0, n_rem) enter with -q (removed), the rest with +q (added).
// With commit, S += dS is written back.  Returns the wave total in every lane.
__device__ inline double wave_recip_delta(int lane, int ns, int n_rem, int K, const cplx* E,
                                          const double* site4, const int4* kn, const double* kw,
                                          double* sre, double* sim, int nk_half, bool commit) {
    const int nkk = 2 * K + 1;
    double acc = 0.0;
    for (int t = lane; t < nk_half; t += 64) {
        const int4 n = kn[t];
        double dre = 0.0, dim = 0.0;
        for (int s = 0; s < ns; ++s) {
            const cplx e0 = E[(0 * nkk + n.x + K) * ns + s];
            const cplx e1 = E[(1 * nkk + n.y + K) * ns + s];
            const cplx e2 = E[(2 * nkk + n.z + K) * ns + s];
            const cplx e = cmul(cmul(e0, e1), e2);
            const double qs = s < n_rem ? -site4[4 * s + 3] : site4[4 * s + 3];
            dre += qs * e.re;
            dim += qs * e.im;
        }
        const double sr = sre[t], si = sim[t];
        acc += kw[t] * (2.0 * (sr * dre + si * dim) + (dre * dre + dim * dim));
        if (commit) {
            sre[t] = sr + dre;
            sim[t] = si + dim;
        }
    }
    return wave_sum(acc);
}

// exp(2 pi i m f) for m = -K..K into E[0..2K], from one sincospi and its powers (negative m
// by conjugation); the powers carry at most K roundings.
__device__ inline void phase_row(double f, int K, cplx* E) {
    double sv, cv;
    sincospi(2.0 * f, &sv, &cv);
    const cplx w = {cv, sv};
    cplx p = {1.0, 0.0};
    E[K] = p;
    for (int m = 1; m <= K; ++m) {
        p = cmul(p, w);
        E[K + m] = p;
        E[K - m] = {p.re, -p.im};
    }
}

// One molecule's share of sum_k w(k) (2 Re(conj(S) dS) + |dS|^2) over the half k-set, taken
// as runs of consecutive n2 (see Framework::d_kseg): the caller's thread g of rg takes runs
// g, g + rg, ... in order.  E: the molecule's per-axis rows, [NS][3][2K + 1]; q: signed charges
// (removed sites negative).  Per run and site, exp(i k.r) = E0[n0] E1[n1] E2[n2] is formed once
// and stepped along n2 by one complex multiply per term.
template <int NS>
__device__ inline double runs_recip_sum(const cplx* E, const double (&q)[NS], int K, const int2* __restrict__ kseg,
                                        int n_seg, int g, int rg, const double* __restrict__ kw,
                                        const double* __restrict__ sre, const double* __restrict__ sim) {
    const int nkk = 2 * K + 1;
    double acc = 0.0;
    for (int sg = g; sg < n_seg; sg += rg) {
        const int2 d = kseg[sg];
        const int i0 = d.y & 255, i1 = (d.y >> 8) & 255, i2 = (d.y >> 16) & 255, len = d.y >> 24;
        cplx e[NS], w[NS];
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            const cplx* Es = E + (size_t)s * 3 * nkk;
            e[s] = cmul(cmul(Es[i0], Es[nkk + i1]), Es[2 * nkk + i2]);
            w[s] = Es[2 * nkk + K + 1];
        }
        for (int i = 0; i < len; ++i) {
            double dre = 0.0, dim = 0.0;
#pragma unroll
            for (int s = 0; s < NS; ++s) {
                dre += q[s] * e[s].re;
                dim += q[s] * e[s].im;
                e[s] = cmul(e[s], w[s]);
            }
            const int k = d.x + i;
            acc += kw[k] * (2.0 * (sre[k] * dre + sim[k] * dim) + (dre * dre + dim * dim));
        }
    }
    return acc;
}

// Molecules per workgroup of 256 threads for the run sum: 16, 8 or 4, so that each molecule's
// 3 * ns phase rows have a thread each and the rows fit in 64 KB of LDS.
inline int runs_molecules_per_block(int ns, int K) {
    const size_t per = (size_t)ns * 3 * (2 * K + 1) * sizeof(cplx);
    for (int rt : {16, 8, 4})
        if (256 / rt >= 3 * ns && per * rt <= 65536) return rt;
    return 0;  // the run kernels do not apply
}

}  // namespace aleator_gpu
