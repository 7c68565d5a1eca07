// grid_tile.hpp -- the framework energy-grid tile kernel (instantiated by grid_tile_fp64.hip and
// grid_tile_mixed.hip; host side in grid.hip).
//
// Reference: FrameworkEnergyGrid constructor body
// (src/engines/monte_carlo/framework_energy_grid.cc:221-248): for every node,
// E = min(sum over framework atoms of pairEnergy(|minimum image|, mix), cap).
//
// MI355X design.  One workgroup (256 threads) owns a tile of TX x TY x (TZ*NPT) nodes; each
// thread owns NPT nodes stacked along the third lattice row, so every LDS entry it reads is
// reused NPT times from registers.  The workgroup walks the framework atoms in passes, and for
// each atom tests the periodic images whose fractional offset from the tile centre lies in
// [-1,1)^3 (only the nearest one when no other can be in range); every image within
// (cutoff + tile radius) of the tile centre is appended to an LDS list (position relative to the
// tile centre, charge, type) through a deterministic block scan, in atom order.  Because the
// cutoff is at most half the smallest perpendicular width, at most one image of an atom lies
// within the cutoff of any node, so summing every listed image inside the cutoff is the
// reference's minimum-image sum.  The list has two parts: images within (cutoff - tile radius)
// of the centre are inside the cutoff for every node of the tile and are summed without a test;
// the others (the shell) are tested per node.  Every lane reads the same entry (an LDS
// broadcast).  The work is VALU-bound: the list is reused by all nodes of the tile and the
// output is 8 B per node per table.
//
// Pair arithmetic is FP64 (pair.hpp): 1/r or 1/r^2 from the hardware approximation refined by NR
// Newton steps (1: MIXED, 2: FP64), LJ as r^-6 (A r^-6 - B) [- shift [- slope r]] with
// A = 4 eps sigma^12, B = 4 eps sigma^6 per (guest, type) (two FMAs per table when truncated),
// real-space Coulomb as 1/r - erf(alpha r)/r with the erf kernel a degree-DEG polynomial in r^2
// (DEG 0: no Coulomb table).
//
// Many frameworks (or one) are built by one launch: workgroup b finds its framework in a prefix
// table of tile counts and reads that framework's descriptor from device memory.  Workgroups are
// dispatched to the 8 XCDs round robin (b mod 8), so the tile of workgroup b is renumbered to
// give each XCD a contiguous range of tiles: neighbouring tiles then share their output cache
// lines and their rows of the reciprocal table inside one L2.
//
// Coulomb table: real space as Ewald::realSpaceParticleEnergy of a unit charge
// (ewald.cc:223-249), reciprocal space as the linear term of computeDelta
// (ewald_incremental_state.cc:117-121) from the separable contraction of grid.hip (the last of
// its three stages is fused here: 2K + 1 complex products per node).
#pragma once

#include <type_traits>

#include "device.hpp"
#include "grid.hpp"
#include "internal.hpp"
#include "pair.hpp"

namespace aleator_gpu {

namespace tile_detail {

constexpr int kThreads = 256;
constexpr int kCap = 1024;  // LDS list capacity (entries)

// contiguous tile ranges per XCD for workgroups dispatched round robin over 8 XCDs (a bijection
// of [0, n))
__device__ inline int xcd_contiguous(int b, int n) {
    constexpr int kXcd = 8;
    const int per = n / kXcd, rem = n % kXcd, x = b % kXcd;
    return x * per + min(x, rem) + b / kXcd;
}

}  // namespace tile_detail

// EXACT: FP64 precision (two Newton steps), else MIXED (one).  LJK: the most general LJ truncation
// of the launch's frameworks (0 truncated, 1 shifted, 2 linear-force-shifted): the shift term is
// compiled in from 1 on, the slope term at 2.
template <bool EXACT, int NG, int DEG, int LJK, int TX, int TY, int TZ, int NPT>
__global__ __launch_bounds__(tile_detail::kThreads) void grid_tile_kernel(const GridArgs* __restrict__ descs,
                                                                          const int* __restrict__ prefix,
                                                                          int n_desc) {
    using namespace tile_detail;
    static_assert(TX * TY * TZ == kThreads, "one thread per (x, y, z-group)");
    constexpr bool COUL = DEG > 0;
    constexpr int DEGF = DEG > kFarDrop ? DEG - kFarDrop : DEG;  // the shell entries' erf degree
    constexpr bool LFS = LJK == 2;
    constexpr int NR = EXACT ? 2 : 1;
    constexpr int NZ = TZ * NPT;  // nodes per tile along z
    constexpr int NGA = NG > 0 ? NG : 1;
    __shared__ double2 e_xy[kCap];
    __shared__ double2 e_zq[kCap];
    __shared__ int e_t[kCap];
    // per (guest, type): A, B, shift = V(rc) - rc V'(rc) (0 when truncated; V(rc) when shifted),
    // slope = V'(rc) (linear-force-shifted only)
    __shared__ double4 tab[NGA][kMaxTypes];
    __shared__ int scan_tmp[kThreads / 64 + 1];
    __shared__ double s_poly[DEG + 1];
    __shared__ double s_polyf[DEGF + 1];

    // framework and tile of this workgroup
    const int b = xcd_contiguous(blockIdx.x, gridDim.x);
    int lo = 0, hi = n_desc - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (prefix[mid] <= b) lo = mid;
        else hi = mid - 1;
    }
    const GridArgs& a = descs[lo];
    const int tile = b - prefix[lo];

    const int tid = threadIdx.x;
    for (int i = tid; i < NG * kMaxTypes; i += kThreads) {
        constexpr int kPairs = kMaxGuests * kMaxTypes;
        const double s2 = a.ljtab[i], e4 = a.ljtab[kPairs + i];
        const double s6 = s2 * s2 * s2;
        const double vrc = a.trunc != ALEATOR_LJ_TRUNCATED ? a.ljtab[2 * kPairs + i] : 0.0;
        const double slope = a.trunc == ALEATOR_LJ_LINEAR_FORCE_SHIFTED ? a.ljtab[3 * kPairs + i] : 0.0;
        (&tab[0][0])[i] = make_double4(e4 * s6 * s6, e4 * s6, vrc - a.rc_lj * slope, slope);
    }
    // the erf polynomials live in LDS and the pair loops read them there (the compiler keeps the
    // current loop's coefficients in vector registers; held as two full register copies they
    // took the kernel to 162 VGPRs and 3 waves per SIMD, now 126 and 4)
    if (tid <= DEG) s_poly[tid] = COUL ? a.poly[tid] : 0.0;
    if (tid <= DEGF) s_polyf[tid] = COUL ? a.poly_far[tid] : 0.0;
    const bool far_ok = COUL && a.far_ok;

    const int n0 = a.n[0], n1 = a.n[1], n2 = a.n[2];
    const int t0 = (n0 + TX - 1) / TX, t1 = (n1 + TY - 1) / TY;
    const int bx = tile % t0;
    const int by = (tile / t0) % t1;
    const int bz = tile / (t0 * t1);
    // threads cover the tile's valid nodes first: in an edge tile (the grid size need not be a
    // multiple of the tile) the valid x, y columns and z groups are packed into the lowest thread
    // ids, so its waves are full or skip the pair loop instead of running with idle lanes
    const int vx = min(TX, n0 - bx * TX), vy = min(TY, n1 - by * TY);
    const int vz = min(TZ, (n2 - bz * NZ + NPT - 1) / NPT);
    const bool in_tile = tid < vx * vy * vz;
    const int tx = in_tile ? tid % vx : 0, ty = in_tile ? (tid / vx) % vy : 0, tz = in_tile ? tid / (vx * vy) : 0;
    const int ix = bx * TX + tx, iy = by * TY + ty, iz0 = bz * NZ + tz * NPT;

    // tile centre in grid and image fractional coordinates
    const double cg0 = (bx * TX + 0.5 * (TX - 1)) / n0;
    const double cg1 = (by * TY + 0.5 * (TY - 1)) / n1;
    const double cg2 = (bz * NZ + 0.5 * (NZ - 1)) / n2;
    const double ci0 = cg0 / a.rep[0], ci1 = cg1 / a.rep[1], ci2 = cg2 / a.rep[2];
    // node 0's offset from the tile centre (Cartesian)
    double ox, oy, oz;
    {
        const double d0 = (tx - 0.5 * (TX - 1)) / n0;
        const double d1 = (ty - 0.5 * (TY - 1)) / n1;
        const double d2 = (tz * NPT - 0.5 * (NZ - 1)) / n2;
        ox = d0 * a.mg[0] + d1 * a.mg[3] + d2 * a.mg[6];
        oy = d0 * a.mg[1] + d1 * a.mg[4] + d2 * a.mg[7];
        oz = d0 * a.mg[2] + d1 * a.mg[5] + d2 * a.mg[8];
    }
    bool valid[NPT];
#pragma unroll
    for (int p = 0; p < NPT; ++p) valid[p] = in_tile && iz0 + p < n2;
    const bool any_valid = valid[0];
    // node p of a thread sits p steps s (one grid step along the third lattice row) past node 0,
    // so |e - o_p|^2 = |e - o_0|^2 - 2p (e - o_0).s + p^2 |s|^2: two instructions per extra node
    const double sx = a.mg[6] / n2, sy = a.mg[7] / n2, sz = a.mg[8] / n2;
    double s2p[NPT], m2p[NPT];
#pragma unroll
    for (int p = 0; p < NPT; ++p) {
        s2p[p] = (double)(p * p) * (sx * sx + sy * sy + sz * sz);
        m2p[p] = (double)(-2 * p);
    }

    double acc[NPT][NGA];
    double accq[NPT];
#pragma unroll
    for (int p = 0; p < NPT; ++p) {
        accq[p] = 0.0;
#pragma unroll
        for (int g = 0; g < NGA; ++g) acc[p][g] = 0.0;
    }

    const double rc2 = a.rc2_lj, rcr2 = a.rc2_real;
    // one cutoff test covers both terms when the cutoffs agree or only one term is built
    const bool one_cut = NG == 0 || !COUL || rc2 == rcr2;
    const double cut2 = NG > 0 ? rc2 : rcr2;

    // entries [lo_e, hi_e); INNER: every node of the tile is within both cutoffs; FAR: the shell's
    // lower-degree erf polynomial
    auto run = [&](auto inner_tag, auto far_tag, int lo_e, int hi_e) {
        constexpr bool INNER = decltype(inner_tag)::value;
        constexpr bool FAR = decltype(far_tag)::value;
        constexpr int D = FAR ? DEGF : DEG;
        const double* pc = FAR ? s_polyf : s_poly;
        const double pxs = FAR ? a.poly_far_xs : a.poly_xs;
        const double pxo = FAR ? a.poly_far_xo : -1.0;
#pragma unroll 2
        for (int e = lo_e; e < hi_e; ++e) {
            const double2 xy = e_xy[e], zq = e_zq[e];
            const int t = e_t[e];
            double4 cg[NGA];
#pragma unroll
            for (int g = 0; g < NG; ++g) cg[g] = tab[g][t];
            const double dx = xy.x - ox, dy = xy.y - oy, dz = zq.x - oz;
            const double r20 = fma(dz, dz, fma(dy, dy, dx * dx));
            const double tds = NPT > 1 ? fma(dz, sz, fma(dy, sy, dx * sx)) : 0.0;
#pragma unroll
            for (int p = 0; p < NPT; ++p) {
                const double r2 = p == 0 ? r20 : fma(tds, m2p[p], r20 + s2p[p]);
                auto lj = [&](double ir2, double r) {
                    const double ir6 = ir2 * ir2 * ir2;
#pragma unroll
                    for (int g = 0; g < NG; ++g) {
                        double v = fma(ir6, fma(cg[g].x, ir6, -cg[g].y), acc[p][g]);
                        if constexpr (LJK == 1) v -= cg[g].z;
                        if constexpr (LFS) v -= fma(r, cg[g].w, cg[g].z);
                        acc[p][g] = v;
                    }
                };
                auto coul = [&](double ir) {
                    if constexpr (COUL) {
                        const double x = fma(r2, pxs, pxo);
                        double pp = pc[D];
#pragma unroll
                        for (int k = D - 1; k >= 0; --k) pp = fma(pp, x, pc[k]);
                        accq[p] = fma(zq.y, ir - pp, accq[p]);
                    }
                };
                if (INNER || one_cut) {
                    if (INNER || r2 <= cut2) {
                        if constexpr (!COUL && !LFS) {
                            lj(rcp_refined<NR>(r2), 0.0);  // LJ alone needs no 1/r
                        } else {
                            const double ir = rsq_refined<NR>(r2);
                            if constexpr (NG > 0) lj(ir * ir, r2 * ir);
                            coul(ir);
                        }
                    }
                } else {
                    const double ir = rsq_refined<NR>(r2);
                    if (r2 <= rc2) lj(ir * ir, r2 * ir);
                    if (r2 <= rcr2) coul(ir);
                }
            }
        }
    };
    auto compute = [&](int n_in, int n_out) {
        __syncthreads();
        if (any_valid) {
            run(std::true_type{}, std::false_type{}, 0, n_in);
            if (far_ok) run(std::false_type{}, std::true_type{}, kCap - n_out, kCap);
            else run(std::false_type{}, std::false_type{}, kCap - n_out, kCap);
        }
        __syncthreads();
    };

    int pend_in = 0, pend_out = 0;
    const int natoms = a.natoms;
    const int chunk = a.chunk;
    const int apt = chunk > kThreads ? chunk / kThreads : 1;  // atoms per thread and pass (contiguous,
                                                              // so the list stays in atom order)
    const double inner2 = a.inner2;
    for (int base = 0; base < natoms; base += chunk) {
        int cnt = 0;                          // inner | shell << 16
        unsigned mask[4] = {0u, 0u, 0u, 0u};  // images of each atom within the search radius
        unsigned in_mask[4] = {0u, 0u, 0u, 0u};
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int at = base + tid * apt + j;
            if (j >= apt || tid * apt >= chunk || at >= natoms) continue;
            double h0 = a.fx[at] - ci0, h1 = a.fy[at] - ci1, h2 = a.fz[at] - ci2;
            h0 -= floor(h0);
            h1 -= floor(h1);
            h2 -= floor(h2);
            auto test = [&](int s) {
                const double g0 = h0 - (s & 1), g1 = h1 - ((s >> 1) & 1), g2 = h2 - (s >> 2);
                const double cx = g0 * a.mi[0] + g1 * a.mi[3] + g2 * a.mi[6];
                const double cy = g0 * a.mi[1] + g1 * a.mi[4] + g2 * a.mi[7];
                const double cz = g0 * a.mi[2] + g1 * a.mi[5] + g2 * a.mi[8];
                const double d2 = cx * cx + cy * cy + cz * cz;
                if (d2 <= a.search2) {
                    mask[j] |= 1u << s;
                    if (d2 <= inner2) {
                        in_mask[j] |= 1u << s;
                        cnt += 1;
                    } else {
                        cnt += 1 << 16;
                    }
                }
            };
            if (a.single_image) {
                // the search radius is below L_perp/2: only the nearest image (fractional
                // offsets in [-1/2, 1/2)) can be within it
                test((h0 >= 0.5 ? 1 : 0) | (h1 >= 0.5 ? 2 : 0) | (h2 >= 0.5 ? 4 : 0));
            } else {
#pragma unroll
                for (int s = 0; s < 8; ++s) test(s);
            }
        }
        int total;
        const int off = block_exclusive_scan(cnt, scan_tmp, &total);
        const int tot_in = total & 0xffff, tot_out = total >> 16;
        if (pend_in + pend_out + tot_in + tot_out > kCap) {
            compute(pend_in, pend_out);
            pend_in = pend_out = 0;
        }
        int w_in = pend_in + (off & 0xffff), w_out = kCap - 1 - (pend_out + (off >> 16));
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
                const int w = (in_mask[j] & (1u << s)) ? w_in++ : w_out--;
                e_xy[w] = make_double2(g0 * a.mi[0] + g1 * a.mi[3] + g2 * a.mi[6],
                                       g0 * a.mi[1] + g1 * a.mi[4] + g2 * a.mi[7]);
                e_zq[w] = make_double2(g0 * a.mi[2] + g1 * a.mi[5] + g2 * a.mi[8], qa);
                e_t[w] = ta;
            }
        }
        pend_in += tot_in;
        pend_out += tot_out;
    }
    compute(pend_in, pend_out);

    const int nk = 2 * a.kmax + 1;
    unsigned nonfinite = 0;  // NaN / Inf sentinel: counted, the values are written as computed
#pragma unroll
    for (int p = 0; p < NPT; ++p) {
        if (!valid[p]) continue;
        const int iz = iz0 + p;
        const long long flat = ((long long)iz * n1 + iy) * n0 + ix;
#pragma unroll
        for (int g = 0; g < NG; ++g) {
            const double e = acc[p][g] + a.tail[g];
            const double v = (a.cap < e) ? a.cap : e;  // std::min(e, cap), NaN kept
            nonfinite += !isfinite(v);
            a.out[g * a.stride + flat] = v;
        }
        if constexpr (COUL) {
            double rec = 0.0;
            for (int m = 0; m < nk; ++m) {
                const double2 tt = a.t2[((long long)m * n2 + iz) * n1 + iy];
                const double2 ee = a.e0[(long long)m * n0 + ix];
                rec += tt.x * ee.x - tt.y * ee.y;
            }
            const double v = a.kc * accq[p] + a.recip_pref * rec;
            nonfinite += !isfinite(v);
            a.out[NG * a.stride + flat] = v;
        }
    }
    if (nonfinite) atomicAdd(a.nonfinite, (unsigned long long)nonfinite);
}

// ---------------------------------------------------------------- instantiation and dispatch
struct TileShape {
    int tx, ty, tz, npt;
};
constexpr TileShape kShapes[kNumTileShapes] = {{8, 8, 4, 1}, {8, 8, 4, 2}, {8, 8, 4, 4}, {4, 8, 8, 2}};

template <bool EXACT, int NG, int DEG, int LJK>
void launch_shape(int shape, const GridArgs* d, const int* prefix, int n_desc, int blocks, hipStream_t s) {
    switch (shape) {
#define ALEATOR_TILE_SHAPE(I, X, Y, Z, P)                                                                       \
    case I:                                                                                                     \
        hipLaunchKernelGGL((grid_tile_kernel<EXACT, NG, DEG, LJK, X, Y, Z, P>), dim3(blocks),                   \
                           dim3(tile_detail::kThreads), 0, s, d, prefix, n_desc);                               \
        return;
        ALEATOR_TILE_SHAPE(0, 8, 8, 4, 1)
        ALEATOR_TILE_SHAPE(1, 8, 8, 4, 2)
        ALEATOR_TILE_SHAPE(2, 8, 8, 4, 4)
        ALEATOR_TILE_SHAPE(3, 4, 8, 8, 2)
#undef ALEATOR_TILE_SHAPE
    }
    fail(ALEATOR_INVALID_ARGUMENT, "grid: unknown tile shape");
}

template <bool EXACT, int DEG, int LJK>
void dispatch_ng(int shape, int ng, const GridArgs* d, const int* prefix, int n_desc, int blocks, hipStream_t s) {
    switch (ng) {
        case 0:
            if constexpr (DEG > 0) return launch_shape<EXACT, 0, DEG, LJK>(shape, d, prefix, n_desc, blocks, s);
            break;
        case 1: return launch_shape<EXACT, 1, DEG, LJK>(shape, d, prefix, n_desc, blocks, s);
        case 2: return launch_shape<EXACT, 2, DEG, LJK>(shape, d, prefix, n_desc, blocks, s);
        case 3: return launch_shape<EXACT, 3, DEG, LJK>(shape, d, prefix, n_desc, blocks, s);
        case 4: return launch_shape<EXACT, 4, DEG, LJK>(shape, d, prefix, n_desc, blocks, s);
    }
    fail(ALEATOR_INVALID_ARGUMENT, "grid: unsupported guest count");
}

template <bool EXACT, int LJK>
void dispatch_deg(int shape, int ng, int deg, const GridArgs* d, const int* prefix, int n_desc, int blocks,
                  hipStream_t s) {
    switch (deg) {
        case 0: return dispatch_ng<EXACT, 0, LJK>(shape, ng, d, prefix, n_desc, blocks, s);
        case 18: return dispatch_ng<EXACT, 18, LJK>(shape, ng, d, prefix, n_desc, blocks, s);
        case 24: return dispatch_ng<EXACT, 24, LJK>(shape, ng, d, prefix, n_desc, blocks, s);
        case kErfPolyMax: return dispatch_ng<EXACT, kErfPolyMax, LJK>(shape, ng, d, prefix, n_desc, blocks, s);
    }
    fail(ALEATOR_INVALID_ARGUMENT, "grid: unsupported erf polynomial degree");
}

}  // namespace aleator_gpu
