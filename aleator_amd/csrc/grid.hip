// grid.hip -- framework energy grids on gfx950.
//
// Reference: FrameworkEnergyGrid constructor body
// (src/engines/monte_carlo/framework_energy_grid.cc:221-248): for every node,
// E = min(sum over framework atoms of pairEnergy(|minimum image|, mix), cap).
//
// MI355X design.  One workgroup (256 threads) owns a tile of TX x TY x (TZ*NPT) nodes;
// each thread owns NPT nodes stacked along z, so every LDS entry it reads is reused NPT
// times from registers.  The workgroup walks the framework atoms in chunks of 128, and for
// each atom tests the 8 periodic images whose fractional offset from the tile centre lies in
// [-1,1)^3; every image within (cutoff + tile radius) of the tile centre is appended to an
// LDS list (structure of arrays: dx, dy, dz relative to the tile centre, charge, type)
// through a deterministic block scan, in atom order.  Because the cutoff is at most half the
// smallest perpendicular width, at most one image of an atom lies within the cutoff of any
// node, so summing every listed image inside the cutoff is the reference's minimum-image
// sum.  The list is in atom order, so each node adds its terms in atom order and a build is
// bitwise reproducible run to run.  Every lane reads the same list entry (an LDS broadcast).  The work is VALU-bound: the list
// is reused by all nodes of the tile and the output is 8 B per node per table.
//
// Many frameworks (or one) are built by one launch: workgroup b finds its framework in a
// prefix table of tile counts and reads that framework's descriptor from device memory.
//
// Coulomb table: real space as Ewald::realSpaceParticleEnergy of a unit charge
// (ewald.cc:223-249), reciprocal space as the linear term of computeDelta
// (ewald_incremental_state.cc:117-121).  Every node sits at fractional coordinates
// i/(n*rep) of the k-set's lattice, so exp(i k.r) factorises per axis and the sum over the
// (2K+1)^3 index cube is three small separable contractions in FP64: two pre-kernels, the
// last one fused into the tile kernel.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <type_traits>
#include <vector>

#include "device.hpp"
#include "grid.hpp"
#include "internal.hpp"

namespace aleator_gpu {

namespace {

constexpr int kThreads = 256;
constexpr int kCap = 1024;   // LDS list capacity (entries)
// atoms staged per pass of the list build: 128 (each may add 8 images, 1024 <= kCap), or, when
// only the nearest image can be within the search radius, 256, 512 or 1024 (1, 2 or 4 atoms per
// thread); the plan's autotuner picks it with the tile shape
constexpr int kChunkSmall = 128;
constexpr int kErfcIntervals = 256;  // FP32 erfc(alpha r) table: cubic Hermite pieces over [0, alpha rc]

// LDS image list: position, charge and compact LJ type of each atom image.  FP64 keeps
// structure-of-arrays; FP32 packs (x, y, z, q) in one float4 so an entry is two LDS reads.
template <typename R>
struct Entries {
    R x[kCap], y[kCap], z[kCap], q[kCap];
    int t[kCap];
    __device__ void put(int w, double x_, double y_, double z_, double q_, int t_, bool) {
        x[w] = (R)x_;
        y[w] = (R)y_;
        z[w] = (R)z_;
        q[w] = (R)q_;
        t[w] = t_;
    }
};
template <>
struct Entries<float> {
    float4 p[kCap];
    int t[kCap];
    // with_charge false: the fourth word carries the type bits instead of the charge
    __device__ void put(int w, double x_, double y_, double z_, double q_, int t_, bool with_charge) {
        p[w] = make_float4((float)x_, (float)y_, (float)z_, with_charge ? (float)q_ : __int_as_float(t_));
        t[w] = t_;
    }
};

template <typename R>
__device__ inline R rcp(R v);
template <>
__device__ inline double rcp<double>(double v) {
    return 1.0 / v;
}
template <>
__device__ inline float rcp<float>(float v) {
    return __builtin_amdgcn_rcpf(v);
}

// LFS: some framework of the launch uses linear-force-shifted LJ (FP32 path: the slope term is
// compiled in; the FP64 path reads the truncation per framework at run time).
template <typename R, int NG, bool COUL, bool LFS, int TX, int TY, int TZ, int NPT>
__global__ __launch_bounds__(kThreads) void grid_tile_kernel(const GridArgs* __restrict__ descs,
                                                             const int* __restrict__ prefix, int n_desc) {
    static_assert(TX * TY * TZ == kThreads, "one thread per (x, y, z-group)");
    constexpr bool kF32 = std::is_same_v<R, float>;
    constexpr int NZ = TZ * NPT;  // nodes per tile along z
    __shared__ Entries<R> ent;
    __shared__ double tab[4][kMaxGuests][kMaxTypes];
    // FP32 pair constants per (guest, type): A = 4 eps sigma^12, B = 4 eps sigma^6, the shift
    // V(rc) - rc V'(rc) and the slope V'(rc) (shift 0 when truncated, slope 0 unless
    // linear-force-shifted): V = A/r^12 - (B/r^6 + shift + r slope), the slope term only with LFS
    __shared__ float4 tabf[kMaxGuests][kMaxTypes];
    __shared__ float4 erfc_tab[(std::is_same_v<R, float> && COUL) ? kErfcIntervals : 1];
    __shared__ int scan_tmp[kThreads / 64 + 1];

    // framework of this workgroup
    const int b = blockIdx.x;
    int lo = 0, hi = n_desc - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (prefix[mid] <= b) lo = mid;
        else hi = mid - 1;
    }
    const GridArgs& a = descs[lo];
    const int tile = b - prefix[lo];

    const int tid = threadIdx.x;
    if constexpr (kF32) {
        constexpr int kPairs = kMaxGuests * kMaxTypes;
        for (int i = tid; i < kPairs; i += kThreads) {
            const double s2 = a.ljtab[i], e4 = a.ljtab[kPairs + i];
            const double s6 = s2 * s2 * s2;
            const double vrc = a.trunc != ALEATOR_LJ_TRUNCATED ? a.ljtab[2 * kPairs + i] : 0.0;
            const double slope = a.trunc == ALEATOR_LJ_LINEAR_FORCE_SHIFTED ? a.ljtab[3 * kPairs + i] : 0.0;
            (&tabf[0][0])[i] =
                make_float4((float)(e4 * s6 * s6), (float)(e4 * s6), (float)(vrc - a.rc_lj * slope), (float)slope);
        }
    } else {
        for (int i = tid; i < 4 * kMaxGuests * kMaxTypes; i += kThreads) (&tab[0][0][0])[i] = a.ljtab[i];
    }

    // erfc(alpha r) for the FP32 pair loop: piece k covers z = alpha r in [k h, (k+1) h],
    // h = alpha rc / (kErfcIntervals - 1), as the cubic Hermite interpolant of the FP64 erfc and
    // its derivative -2/sqrt(pi) exp(-z^2) at both ends, in powers of the fraction of the piece.
    // Interpolation error below 1.3e-7 relative on [0, 3.2] (emulated against the FP64 erfc).
    float erfc_scale = 0.0f;  // r -> piece coordinate u = z / h
    if constexpr (kF32 && COUL) {
        static_assert(kErfcIntervals == kThreads, "one piece per thread");
        const double h = a.alpha * sqrt(a.rc2_real) / (kErfcIntervals - 1);
        const double z0 = tid * h, z1 = z0 + h;
        const double f0 = erfc(z0), f1 = erfc(z1);
        const double d0 = -1.1283791670955126 * exp(-z0 * z0) * h, d1 = -1.1283791670955126 * exp(-z1 * z1) * h;
        erfc_tab[tid] = make_float4((float)f0, (float)d0, (float)(3.0 * (f1 - f0) - (2.0 * d0 + d1)),
                                    (float)(2.0 * (f0 - f1) + (d0 + d1)));
        erfc_scale = (float)(a.alpha / h);
    }

    const int n0 = a.n[0], n1 = a.n[1], n2 = a.n[2];
    const int t0 = (n0 + TX - 1) / TX, t1 = (n1 + TY - 1) / TY;
    const int bx = tile % t0;
    const int by = (tile / t0) % t1;
    const int bz = tile / (t0 * t1);
    const int tx = tid % TX, ty = (tid / TX) % TY, tz = tid / (TX * TY);
    const int ix = bx * TX + tx, iy = by * TY + ty, iz0 = bz * NZ + tz * NPT;

    // tile centre in grid and image fractional coordinates
    const double cg0 = (bx * TX + 0.5 * (TX - 1)) / n0;
    const double cg1 = (by * TY + 0.5 * (TY - 1)) / n1;
    const double cg2 = (bz * NZ + 0.5 * (NZ - 1)) / n2;
    const double ci0 = cg0 / a.rep[0], ci1 = cg1 / a.rep[1], ci2 = cg2 / a.rep[2];
    // node offsets from the tile centre (Cartesian)
    R ox[NPT], oy[NPT], oz[NPT];
    bool valid[NPT];
    const double d0 = (tx - 0.5 * (TX - 1)) / n0;
    const double d1 = (ty - 0.5 * (TY - 1)) / n1;
#pragma unroll
    for (int p = 0; p < NPT; ++p) {
        const double d2 = (tz * NPT + p - 0.5 * (NZ - 1)) / n2;
        ox[p] = (R)(d0 * a.mg[0] + d1 * a.mg[3] + d2 * a.mg[6]);
        oy[p] = (R)(d0 * a.mg[1] + d1 * a.mg[4] + d2 * a.mg[7]);
        oz[p] = (R)(d0 * a.mg[2] + d1 * a.mg[5] + d2 * a.mg[8]);
        valid[p] = ix < n0 && iy < n1 && iz0 + p < n2;
    }
    const bool any_valid = valid[0];
    // node p of a thread sits p steps s (one grid step along the third lattice row) past node 0,
    // so |e - o_p|^2 = |e - o_0|^2 - 2p (e - o_0).s + p^2 |s|^2: two instructions per extra node
    // instead of six, and no cancellation because |s| is a grid spacing
    float sx = 0.0f, sy = 0.0f, sz = 0.0f, s2p[NPT], m2p[NPT];
    if constexpr (kF32) {
        const double gx = a.mg[6] / n2, gy = a.mg[7] / n2, gz = a.mg[8] / n2;
        sx = (float)gx;
        sy = (float)gy;
        sz = (float)gz;
#pragma unroll
        for (int p = 0; p < NPT; ++p) {
            s2p[p] = (float)((double)(p * p) * (gx * gx + gy * gy + gz * gz));
            m2p[p] = (float)(-2 * p);
        }
    }

    double acc[NPT][NG > 0 ? NG : 1];
    double accq[NPT];
#pragma unroll
    for (int p = 0; p < NPT; ++p) {
        accq[p] = 0.0;
#pragma unroll
        for (int g = 0; g < (NG > 0 ? NG : 1); ++g) acc[p][g] = 0.0;
    }

    const R rc2 = (R)a.rc2_lj;
    const R rcr2 = (R)a.rc2_real;
    const R alpha = (R)a.alpha;
    const R rc = (R)a.rc_lj;
    const int trunc = a.trunc;

    auto compute = [&](int count) {
        __syncthreads();
        if constexpr (kF32) {
            // 1/r from one reciprocal square root: r = r2 / sqrt(r2), 1/r^2 = (1/r)^2.  When the LJ
            // and real-space cutoffs are equal (or only one term is built) one branch covers both
            // terms; otherwise each term has its own branch.
            auto pairs = [&](auto same_cutoffs) {
                constexpr bool kOne = decltype(same_cutoffs)::value;
#pragma unroll 2
                for (int e = 0; e < count; ++e) {
                    const float4 P = ent.p[e];
                    // without a Coulomb table the entry's fourth word holds its type (one LDS read)
                    const int t = COUL ? ent.t[e] : __float_as_int(P.w);
                    float4 cg[NG > 0 ? NG : 1];
#pragma unroll
                    for (int g = 0; g < NG; ++g) cg[g] = tabf[g][t];
                    const float dx = P.x - ox[0];
                    const float dy = P.y - oy[0];
                    const float dz = P.z - oz[0];
                    const float r20 = dx * dx + dy * dy + dz * dz;
                    const float tds = NPT > 1 ? dx * sx + dy * sy + dz * sz : 0.0f;
#pragma unroll
                    for (int p = 0; p < NPT; ++p) {
                        const float r2 = p == 0 ? r20 : fmaf(tds, m2p[p], r20 + s2p[p]);
                        // LJ from 1/r^2 (and r, linear-force-shifted only)
                        auto lj = [&](float ir2, float r) {
                            const float ir6 = ir2 * ir2 * ir2;
                            const float ir12 = ir6 * ir6;
#pragma unroll
                            for (int g = 0; g < NG; ++g) {
                                float w = fmaf(cg[g].y, ir6, cg[g].z);
                                if constexpr (LFS) w = fmaf(r, cg[g].w, w);
                                acc[p][g] += (double)fmaf(cg[g].x, ir12, -w);
                            }
                        };
                        auto coul = [&](float ir) {
                            // r <= rc(1 + 2^-22) here, so 0 <= u < kErfcIntervals and the piece
                            // index (int)u = floor(u) stays inside the table
                            const float u = (r2 * ir) * erfc_scale;
                            const float f = __builtin_amdgcn_fractf(u);
                            const float4 c = erfc_tab[(int)u];
                            const float erfcz = fmaf(fmaf(fmaf(c.w, f, c.z), f, c.y), f, c.x);
                            accq[p] += (double)(P.w * (erfcz * ir));
                        };
                        if constexpr (kOne) {
                            if (r2 <= (NG > 0 ? rc2 : rcr2)) {
                                if constexpr (!COUL && !LFS) {
                                    lj(__builtin_amdgcn_rcpf(r2), 0.0f);  // LJ alone needs no 1/r
                                } else {
                                    const float ir = __builtin_amdgcn_rsqf(r2);
                                    if constexpr (NG > 0) lj(ir * ir, r2 * ir);
                                    if constexpr (COUL) coul(ir);
                                }
                            }
                        } else {
                            const float ir = __builtin_amdgcn_rsqf(r2);
                            if (r2 <= rc2) lj(ir * ir, r2 * ir);
                            if (r2 <= rcr2) coul(ir);
                        }
                    }
                }
            };
            if (any_valid) {
                if (NG == 0 || !COUL || rc2 == rcr2)
                    pairs(std::true_type{});
                else
                    pairs(std::false_type{});
            }
        } else if (any_valid) {
#pragma unroll 2
            for (int e = 0; e < count; ++e) {
                const R ex = ent.x[e], ey = ent.y[e], ez = ent.z[e];
                const int t = ent.t[e];
                R s2c[NG > 0 ? NG : 1], e4c[NG > 0 ? NG : 1];
#pragma unroll
                for (int g = 0; g < NG; ++g) {
                    s2c[g] = (R)tab[0][g][t];
                    e4c[g] = (R)tab[1][g][t];
                }
#pragma unroll
                for (int p = 0; p < NPT; ++p) {
                    const R dx = ex - ox[p];
                    const R dy = ey - oy[p];
                    const R dz = ez - oz[p];
                    const R r2 = dx * dx + dy * dy + dz * dz;
                    if (NG > 0 && r2 <= rc2) {
                        const R ir2 = rcp<R>(r2);
                        R r = 0;
                        if (trunc == ALEATOR_LJ_LINEAR_FORCE_SHIFTED) r = sqrt(r2);
#pragma unroll
                        for (int g = 0; g < NG; ++g) {
                            const R s2 = s2c[g] * ir2;
                            const R s6 = s2 * s2 * s2;
                            R v = e4c[g] * (s6 * s6 - s6);
                            if (trunc == ALEATOR_LJ_SHIFTED) {
                                v -= (R)tab[2][g][t];
                            } else if (trunc == ALEATOR_LJ_LINEAR_FORCE_SHIFTED) {
                                v = v - (R)tab[2][g][t] - (r - rc) * (R)tab[3][g][t];
                            }
                            acc[p][g] += (double)v;
                        }
                    }
                    if (COUL && r2 <= rcr2) {
                        const R r = sqrt(r2);
                        accq[p] += (double)(ent.q[e] * erfc_over_r<R>(alpha, r));
                    }
                }
            }
        }
        __syncthreads();
    };

    int pending = 0;
    const int natoms = a.natoms;
    const int chunk = a.chunk;
    const int apt = chunk > kThreads ? chunk / kThreads : 1;  // atoms per thread and pass (contiguous,
                                                              // so the list stays in atom order)
    for (int base = 0; base < natoms; base += chunk) {
        int cnt = 0;
        unsigned mask[4] = {0u, 0u, 0u, 0u};  // images of each atom within the search radius
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int at = base + tid * apt + j;
            if (j >= apt || tid * apt >= chunk || at >= natoms) continue;
            double h0 = a.fx[at] - ci0, h1 = a.fy[at] - ci1, h2 = a.fz[at] - ci2;
            h0 -= floor(h0);
            h1 -= floor(h1);
            h2 -= floor(h2);
            if (a.single_image) {
                // the search radius is below L_perp/2: only the nearest image (fractional
                // offsets in [-1/2, 1/2)) can be within it
                const int s = (h0 >= 0.5 ? 1 : 0) | (h1 >= 0.5 ? 2 : 0) | (h2 >= 0.5 ? 4 : 0);
                const double g0 = h0 - (s & 1), g1 = h1 - ((s >> 1) & 1), g2 = h2 - (s >> 2);
                const double cx = g0 * a.mi[0] + g1 * a.mi[3] + g2 * a.mi[6];
                const double cy = g0 * a.mi[1] + g1 * a.mi[4] + g2 * a.mi[7];
                const double cz = g0 * a.mi[2] + g1 * a.mi[5] + g2 * a.mi[8];
                if (cx * cx + cy * cy + cz * cz <= a.search2) {
                    mask[j] = 1u << s;
                    ++cnt;
                }
            } else {
#pragma unroll
                for (int s = 0; s < 8; ++s) {
                    const double g0 = h0 - (s & 1), g1 = h1 - ((s >> 1) & 1), g2 = h2 - (s >> 2);
                    const double cx = g0 * a.mi[0] + g1 * a.mi[3] + g2 * a.mi[6];
                    const double cy = g0 * a.mi[1] + g1 * a.mi[4] + g2 * a.mi[7];
                    const double cz = g0 * a.mi[2] + g1 * a.mi[5] + g2 * a.mi[8];
                    if (cx * cx + cy * cy + cz * cz <= a.search2) {
                        mask[j] |= 1u << s;
                        ++cnt;
                    }
                }
            }
        }
        int total;
        const int off = block_exclusive_scan(cnt, scan_tmp, &total);
        if (pending + total > kCap) {
            compute(pending);
            pending = 0;
        }
        int w = pending + off;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            if (!mask[j]) continue;
            const int at = base + tid * apt + j;
            const double qa = a.q[at];
            const int ta = a.type[at];
            // the wrapped offsets again (not kept across the scan, to spare registers)
            double h0 = a.fx[at] - ci0, h1 = a.fy[at] - ci1, h2 = a.fz[at] - ci2;
            h0 -= floor(h0);
            h1 -= floor(h1);
            h2 -= floor(h2);
            for (int s = 0; s < 8; ++s) {
                if (!(mask[j] & (1u << s))) continue;
                const double g0 = h0 - (s & 1), g1 = h1 - ((s >> 1) & 1), g2 = h2 - (s >> 2);
                ent.put(w, g0 * a.mi[0] + g1 * a.mi[3] + g2 * a.mi[6], g0 * a.mi[1] + g1 * a.mi[4] + g2 * a.mi[7],
                        g0 * a.mi[2] + g1 * a.mi[5] + g2 * a.mi[8], qa, ta, COUL);
                ++w;
            }
        }
        pending += total;
    }
    compute(pending);

    const int nk = 2 * a.kmax + 1;
#pragma unroll
    for (int p = 0; p < NPT; ++p) {
        if (!valid[p]) continue;
        const int iz = iz0 + p;
        const long long flat = ((long long)iz * n1 + iy) * n0 + ix;
#pragma unroll
        for (int g = 0; g < NG; ++g) {
            const double e = acc[p][g];
            a.out[g * a.stride + flat] = (a.cap < e) ? a.cap : e;  // std::min(e, cap), NaN kept
        }
        if (COUL) {
            double rec = 0.0;
            for (int m = 0; m < nk; ++m) {
                const double2 tt = a.t2[((long long)m * n2 + iz) * n1 + iy];
                const double2 ee = a.e0[(long long)m * n0 + ix];
                rec += tt.x * ee.x - tt.y * ee.y;
            }
            a.out[NG * a.stride + flat] = a.kc * accq[p] + a.recip_pref * rec;
        }
    }
}

// T1[n0][n1][iz] = sum_n2 C[n0][n1][n2] * E2[n2][iz]
__global__ void recip_stage_a(const double2* c, const double2* e2, int nk, int nz, double2* t1) {
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long long)nk * nk * nz) return;
    const int iz = idx % nz;
    const long long n01 = idx / nz;
    double re = 0.0, im = 0.0;
    for (int n2 = 0; n2 < nk; ++n2) {
        const double2 a = c[n01 * nk + n2];
        const double2 b = e2[(long long)n2 * nz + iz];
        re += a.x * b.x - a.y * b.y;
        im += a.x * b.y + a.y * b.x;
    }
    t1[idx] = make_double2(re, im);
}

// T2[n0][iz][iy] = sum_n1 T1[n0][n1][iz] * E1[n1][iy]
__global__ void recip_stage_b(const double2* t1, const double2* e1, int nk, int ny, int nz, double2* t2) {
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long long)nk * nz * ny) return;
    const int iy = idx % ny;
    const int iz = (idx / ny) % nz;
    const int n0 = idx / ((long long)ny * nz);
    double re = 0.0, im = 0.0;
    for (int n1 = 0; n1 < nk; ++n1) {
        const double2 a = t1[((long long)n0 * nk + n1) * nz + iz];
        const double2 b = e1[(long long)n1 * ny + iy];
        re += a.x * b.x - a.y * b.y;
        im += a.x * b.y + a.y * b.x;
    }
    t2[idx] = make_double2(re, im);
}

// ---------------------------------------------------------------- tile shapes
struct TileShape {
    int tx, ty, tz, npt;
};
constexpr TileShape kShapes[kNumTileShapes] = {{8, 8, 4, 1}, {8, 8, 4, 2}, {8, 8, 4, 4}, {4, 8, 8, 2}};

template <typename R, int NG, bool COUL, bool LFS>
void launch_shape(int shape, const GridArgs* d, const int* prefix, int n_desc, int blocks, hipStream_t s) {
    switch (shape) {
#define SHAPE(I, X, Y, Z, P)                                                                                   \
    case I:                                                                                                    \
        hipLaunchKernelGGL((grid_tile_kernel<R, NG, COUL, LFS, X, Y, Z, P>), dim3(blocks), dim3(kThreads), 0, s, \
                           d, prefix, n_desc);                                                                 \
        return;
        SHAPE(0, 8, 8, 4, 1)
        SHAPE(1, 8, 8, 4, 2)
        SHAPE(2, 8, 8, 4, 4)
        SHAPE(3, 4, 8, 8, 2)
#undef SHAPE
    }
    fail(ALEATOR_INVALID_ARGUMENT, "grid: unknown tile shape");
}

template <typename R, bool LFS>
void dispatch(int shape, int ng, bool coul, const GridArgs* d, const int* prefix, int n_desc, int blocks,
              hipStream_t s) {
    if (coul) {
        switch (ng) {
            case 0: return launch_shape<R, 0, true, LFS>(shape, d, prefix, n_desc, blocks, s);
            case 1: return launch_shape<R, 1, true, LFS>(shape, d, prefix, n_desc, blocks, s);
            case 2: return launch_shape<R, 2, true, LFS>(shape, d, prefix, n_desc, blocks, s);
            case 3: return launch_shape<R, 3, true, LFS>(shape, d, prefix, n_desc, blocks, s);
            case 4: return launch_shape<R, 4, true, LFS>(shape, d, prefix, n_desc, blocks, s);
        }
    } else {
        switch (ng) {
            case 1: return launch_shape<R, 1, false, LFS>(shape, d, prefix, n_desc, blocks, s);
            case 2: return launch_shape<R, 2, false, LFS>(shape, d, prefix, n_desc, blocks, s);
            case 3: return launch_shape<R, 3, false, LFS>(shape, d, prefix, n_desc, blocks, s);
            case 4: return launch_shape<R, 4, false, LFS>(shape, d, prefix, n_desc, blocks, s);
        }
    }
    fail(ALEATOR_INVALID_ARGUMENT, "grid: unsupported guest count");
}

double norm3(const double* v) { return std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); }

size_t align256(size_t v) { return (v + 255) & ~size_t(255); }

}  // namespace

int tile_blocks(const int32_t n[3], int shape) {
    const TileShape& t = kShapes[shape];
    const long long b = (long long)((n[0] + t.tx - 1) / t.tx) * ((n[1] + t.ty - 1) / t.ty) *
                        ((n[2] + t.tz * t.npt - 1) / (t.tz * t.npt));
    if (b > 0x3fffffff) fail(ALEATOR_INVALID_ARGUMENT, "grid: too many nodes");
    return (int)b;
}

// ---------------------------------------------------------------- job: one framework's build
void GridJob::prepare(Framework& fw, const GridRequest& req) {
    ng = (int)req.guests.size();
    coulomb = req.coulomb;
    precision = req.precision;
    if (ng > kMaxGuests) fail(ALEATOR_INVALID_ARGUMENT, "grid: at most 4 guest species per build");
    if (ng == 0 && !req.coulomb) fail(ALEATOR_INVALID_ARGUMENT, "grid: nothing to build");
    if (req.coulomb && !fw.charged)
        fail(ALEATOR_INVALID_ARGUMENT, "grid: a Coulomb table needs a framework created with Ewald parameters");
    if (req.precision != ALEATOR_PRECISION_FP64 && req.precision != ALEATOR_PRECISION_MIXED)
        fail(ALEATOR_INVALID_ARGUMENT, "grid: unknown precision");
    for (int i = 0; i < 3; ++i)
        if (req.n[i] < 1) fail(ALEATOR_INVALID_ARGUMENT, "grid: point counts must be >= 1");
    GridArgs& a = args;
    a = GridArgs{};
    for (int i = 0; i < 3; ++i) {
        a.n[i] = req.n[i];
        for (int j = 0; j < 3; ++j) a.mi[3 * i + j] = fw.lat.m[i][j];
    }
    // grid lattice: the framework lattice or a primitive cell whose rows are the framework rows
    // divided by integers
    double gm[9];
    if (req.grid_m) {
        std::memcpy(gm, req.grid_m, sizeof(gm));
        for (int i = 0; i < 3; ++i) {
            const double rep = std::round(norm3(&a.mi[3 * i]) / norm3(&gm[3 * i]));
            if (rep < 1.0) fail(ALEATOR_INVALID_ARGUMENT, "grid: grid lattice larger than the framework lattice");
            for (int j = 0; j < 3; ++j)
                if (std::abs(gm[3 * i + j] * rep - a.mi[3 * i + j]) > 1e-9 * norm3(&a.mi[3 * i]))
                    fail(ALEATOR_INVALID_ARGUMENT,
                         "grid: grid lattice rows must be the framework lattice rows divided by integers");
            a.rep[i] = rep;
        }
    } else {
        std::memcpy(gm, a.mi, sizeof(gm));
        a.rep[0] = a.rep[1] = a.rep[2] = 1.0;
    }
    std::memcpy(a.mg, gm, sizeof(gm));
    // the tile radius is checked against L_perp for the largest tile of every shape
    rmax = std::max(ng > 0 ? fw.rc : 0.0, req.coulomb ? fw.rc_real : 0.0);
    for (int i = 0; i < 3; ++i) lperp[i] = fw.lat.lperp[i];
    a.fx = fw.d_fx;
    a.fy = fw.d_fy;
    a.fz = fw.d_fz;
    a.q = fw.d_q;
    a.type = fw.d_type;
    a.natoms = (int)fw.n;
    a.rc_lj = fw.rc;
    a.rc2_lj = fw.rc * fw.rc;
    a.rc2_real = fw.rc_real * fw.rc_real;
    a.alpha = fw.alpha;
    a.kc = kCoulomb;
    a.cap = req.cap;
    a.trunc = fw.trunc;
    a.kmax = req.coulomb ? fw.kmax : 0;
    a.out = req.out_device;
    a.stride = (long long)req.n[0] * req.n[1] * req.n[2];

    // LJ constants per (guest slot, framework type)
    ljtab.assign(4 * kMaxGuests * kMaxTypes, 0.0);
    for (int g = 0; g < ng; ++g) {
        if (req.guests[g] >= fw.eps.size()) fail(ALEATOR_OUT_OF_RANGE, "guest species out of range");
        for (size_t t = 0; t < fw.type_species.size(); ++t) {
            double s2, e4, v, dv;
            lj_pair_constants(fw, req.guests[g], fw.type_species[t], s2, e4, v, dv);
            ljtab[(0 * kMaxGuests + g) * kMaxTypes + t] = s2;
            ljtab[(1 * kMaxGuests + g) * kMaxTypes + t] = e4;
            ljtab[(2 * kMaxGuests + g) * kMaxTypes + t] = v;
            ljtab[(3 * kMaxGuests + g) * kMaxTypes + t] = dv;
        }
    }
    bytes = align256(ljtab.size() * sizeof(double));
    nk = 2 * fw.kmax + 1;
    if (req.coulomb) {
        cube.assign((size_t)nk * nk * nk, make_double2(0.0, 0.0));
        // C(n) = w(n) conj(S(n)) over the full k-set (S(-n) = conj S(n))
        const size_t nkt = fw.kset.size();
        for (size_t t = 0; t < nkt; ++t) {
            const bool low = (int)t < fw.nk_half;
            const size_t h = low ? t : nkt - 1 - t;
            const double re = fw.h_sre[h], im = low ? fw.h_sim[h] : -fw.h_sim[h];
            const int* nv = &fw.kset.n[3 * t];
            const size_t ci = ((size_t)(nv[0] + fw.kmax) * nk + (nv[1] + fw.kmax)) * nk + (nv[2] + fw.kmax);
            cube[ci] = make_double2(fw.kset.weight[t] * re, -fw.kset.weight[t] * im);
        }
        auto phase_table = [&](int axis, std::vector<double2>& e) {
            const long long period = (long long)req.n[axis] * (long long)a.rep[axis];
            e.resize((size_t)nk * req.n[axis]);
            for (int m = 0; m < nk; ++m)
                for (int i = 0; i < req.n[axis]; ++i) {
                    long long p = ((long long)(m - fw.kmax) * i) % period;
                    if (p < 0) p += period;
                    const double ang = 2.0 * kPi * (double)p / (double)period;
                    e[(size_t)m * req.n[axis] + i] = make_double2(std::cos(ang), std::sin(ang));
                }
        };
        phase_table(0, e0);
        phase_table(1, e1);
        phase_table(2, e2);
        off_c = bytes;
        bytes += align256(cube.size() * sizeof(double2));
        off_e0 = bytes;
        bytes += align256(e0.size() * sizeof(double2));
        off_e1 = bytes;
        bytes += align256(e1.size() * sizeof(double2));
        off_e2 = bytes;
        bytes += align256(e2.size() * sizeof(double2));
        off_t1 = bytes;
        bytes += align256((size_t)nk * nk * req.n[2] * sizeof(double2));
        off_t2 = bytes;
        bytes += align256((size_t)nk * req.n[2] * req.n[1] * sizeof(double2));
        a.recip_pref = kCoulomb * (4.0 * kPi / std::abs(fw.lat.volume));
    }
}

double GridJob::tile_radius(int shape) const {
    const TileShape& t = kShapes[shape];
    const int ext[3] = {t.tx, t.ty, t.tz * t.npt};
    double radius = 0.0;
    for (int c = 0; c < 8; ++c) {
        double d[3];
        for (int i = 0; i < 3; ++i) d[i] = (((c >> i) & 1) ? 0.5 : -0.5) * (ext[i] - 1) / args.n[i];
        double v[3];
        for (int j = 0; j < 3; ++j) v[j] = d[0] * args.mg[j] + d[1] * args.mg[3 + j] + d[2] * args.mg[6 + j];
        radius = std::max(radius, norm3(v));
    }
    return radius;
}

bool GridJob::shape_fits(int shape) const {
    const double sr = rmax + tile_radius(shape) + 1e-6 * (1.0 + rmax);
    for (int i = 0; i < 3; ++i)
        if (sr >= lperp[i]) return false;
    return true;
}

void GridJob::set_shape(int s, int chunk) {
    // the 8-image scheme needs every image within the search radius to have fractional offsets
    // inside (-1, 1): guaranteed when the search radius is below L_perp
    if (s < 0 || s >= kNumTileShapes) fail(ALEATOR_INVALID_ARGUMENT, "grid: unknown tile shape");
    if (!shape_fits(s))
        fail(ALEATOR_INVALID_ARGUMENT, "grid: cutoff plus tile radius exceeds the perpendicular width");
    const double sr = rmax + tile_radius(s) + 1e-6 * (1.0 + rmax);
    args.search2 = sr * sr;
    // below half the smallest perpendicular width at most one image of an atom, the nearest,
    // is within the search radius (the fractional rounding image is then the nearest one)
    args.single_image = (2.0 * sr < std::min(lperp[0], std::min(lperp[1], lperp[2]))) ? 1 : 0;
    // more than 128 atoms per pass only with one image per atom (a pass then adds <= 1024 entries)
    if (!(chunk == kChunkSmall || (args.single_image && (chunk == 256 || chunk == 512 || chunk == 1024))))
        fail(ALEATOR_INVALID_ARGUMENT,
             "grid: atoms per list pass must be 128, or 256, 512 or 1024 with one image per atom");
    args.chunk = chunk;
    blocks = tile_blocks(args.n, s);
    shape = s;
}

bool GridJob::chunk_fits(int s, int chunk) const {
    if (chunk == kChunkSmall) return true;
    if (chunk != 256 && chunk != 512 && chunk != 1024) return false;
    const double sr = rmax + tile_radius(s) + 1e-6 * (1.0 + rmax);
    return 2.0 * sr < std::min(lperp[0], std::min(lperp[1], lperp[2]));
}

void GridJob::bind(char* base) {
    dev_base = base;
    args.ljtab = reinterpret_cast<const double*>(base);
    if (coulomb) {
        args.t2 = reinterpret_cast<const double2*>(base + off_t2);
        args.e0 = reinterpret_cast<const double2*>(base + off_e0);
    }
}

void GridJob::upload(hipStream_t s) const {
    ALEATOR_HIP_CHECK(hipMemcpyAsync(dev_base, ljtab.data(), ljtab.size() * sizeof(double), hipMemcpyHostToDevice, s));
    if (coulomb) {
        ALEATOR_HIP_CHECK(hipMemcpyAsync(dev_base + off_c, cube.data(), cube.size() * sizeof(double2),
                                         hipMemcpyHostToDevice, s));
        ALEATOR_HIP_CHECK(hipMemcpyAsync(dev_base + off_e0, e0.data(), e0.size() * sizeof(double2),
                                         hipMemcpyHostToDevice, s));
        ALEATOR_HIP_CHECK(hipMemcpyAsync(dev_base + off_e1, e1.data(), e1.size() * sizeof(double2),
                                         hipMemcpyHostToDevice, s));
        ALEATOR_HIP_CHECK(hipMemcpyAsync(dev_base + off_e2, e2.data(), e2.size() * sizeof(double2),
                                         hipMemcpyHostToDevice, s));
    }
}

void GridJob::launch_pre(hipStream_t s) const {
    if (!coulomb) return;
    auto* d_c = reinterpret_cast<const double2*>(dev_base + off_c);
    auto* d_e1 = reinterpret_cast<const double2*>(dev_base + off_e1);
    auto* d_e2 = reinterpret_cast<const double2*>(dev_base + off_e2);
    auto* d_t1 = reinterpret_cast<double2*>(dev_base + off_t1);
    auto* d_t2 = reinterpret_cast<double2*>(dev_base + off_t2);
    const long long na = (long long)nk * nk * args.n[2];
    hipLaunchKernelGGL(recip_stage_a, dim3((unsigned)((na + 255) / 256)), dim3(256), 0, s, d_c, d_e2, nk, args.n[2],
                       d_t1);
    const long long nb = (long long)nk * args.n[2] * args.n[1];
    hipLaunchKernelGGL(recip_stage_b, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0, s, d_t1, d_e1, nk, args.n[1],
                       args.n[2], d_t2);
}

void launch_tiles(int precision, int shape, int ng, bool coulomb, bool lfs, const GridArgs* d_descs,
                  const int* d_prefix, int n_desc, int blocks, hipStream_t s) {
    if (precision == ALEATOR_PRECISION_FP64)
        dispatch<double, false>(shape, ng, coulomb, d_descs, d_prefix, n_desc, blocks, s);
    else if (lfs)
        dispatch<float, true>(shape, ng, coulomb, d_descs, d_prefix, n_desc, blocks, s);
    else
        dispatch<float, false>(shape, ng, coulomb, d_descs, d_prefix, n_desc, blocks, s);
}

// single build through the context scratch buffer (the per-call entry points)
void launch_grid(Framework& fw, const GridRequest& req) {
    Context& ctx = *fw.ctx;
    GridJob job;
    job.prepare(fw, req);
    const int shape = job.shape_fits(ctx.tile_shape) ? ctx.tile_shape : 0;
    job.set_shape(shape, job.chunk_fits(shape, ctx.tile_chunk) ? ctx.tile_chunk : 128);
    const size_t off_desc = job.bytes, off_prefix = align256(off_desc + sizeof(GridArgs));
    char* base = static_cast<char*>(ctx.ensure_scratch(off_prefix + 256));
    job.bind(base);
    job.upload(ctx.stream);
    const int prefix[2] = {0, job.blocks};
    ALEATOR_HIP_CHECK(hipMemcpyAsync(base + off_desc, &job.args, sizeof(GridArgs), hipMemcpyHostToDevice, ctx.stream));
    ALEATOR_HIP_CHECK(hipMemcpyAsync(base + off_prefix, prefix, sizeof(prefix), hipMemcpyHostToDevice, ctx.stream));
    job.launch_pre(ctx.stream);
    ctx.time_begin();
    launch_tiles(job.precision, job.shape_used(), job.ng, job.coulomb, job.args.trunc == ALEATOR_LJ_LINEAR_FORCE_SHIFTED,
                 reinterpret_cast<const GridArgs*>(base + off_desc),
                 reinterpret_cast<const int*>(base + off_prefix), 1, job.blocks, ctx.stream);
    ctx.time_end();
    ALEATOR_HIP_CHECK(hipGetLastError());
}

}  // namespace aleator_gpu
