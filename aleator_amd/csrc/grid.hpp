// grid.hpp -- descriptors and jobs of the grid build (see grid.hip).
#pragma once

#include <vector>

#include "internal.hpp"

namespace aleator_gpu {

constexpr int kNumTileShapes = 4;
constexpr int kChunkSmall = 128;  // atoms per list pass when an atom may add 8 images
constexpr int kFarDrop = 7;  // the shell entries' erf polynomial is this many degrees lower

// Everything one framework's tile kernel reads; stored in device memory, one per framework.
struct GridArgs {
    int n[3];
    double mg[9];   // grid lattice rows
    double mi[9];   // image (framework) lattice rows
    double rep[3];  // image fractional = grid fractional / rep
    const double* fx;
    const double* fy;
    const double* fz;
    const double* q;
    const int32_t* type;
    int natoms;
    double search2;
    int single_image;  // search radius < L_perp/2: test only the nearest image of each atom
    int chunk;         // atoms staged per list-build pass: 128, or 256 / 512 / 1024 with single_image
    double rc_lj, rc2_lj;
    double rc2_real, alpha, kc;
    double cap;
    int trunc;
    const double* ljtab;  // [4][kMaxGuests][kMaxTypes]: sigma^2, 4 eps, V(rc), V'(rc)
    const double2* t2;    // [2K+1][nz][ny]
    const double2* e0;    // [2K+1][nx]
    int kmax;
    double recip_pref;
    double* out;
    long long stride;
    double inner2;                    // images within sqrt(inner2) of the tile centre are inside
                                      // every node's cutoffs (-1: none)
    double poly[kErfPolyMax + 1];     // erf(alpha r)/r in x = r^2 poly_xs - 1 (pair.hpp)
    double poly_xs;
    // the same kernel fitted over the distances the shell entries can have (r >= sqrt(inner2) -
    // tile radius), kFarDrop degrees lower, in x = r^2 poly_far_xs + poly_far_xo; used when far_ok
    double poly_far[kErfPolyMax + 1];
    double poly_far_xs, poly_far_xo;
    int far_ok;
    double tail[kMaxGuests];          // analytic LJ tail of one guest site, added before the cap
    unsigned long long* nonfinite;    // running count of non-finite values written (the framework's)
};

// One framework's build: host-side constants, their device placement, pre-kernels.
struct GridJob {
    GridArgs args{};
    int ng = 0;
    bool coulomb = false;
    int precision = 0;
    int deg = 0;  // erf polynomial degree (0 without a Coulomb table)
    int blocks = 0;
    int shape = 0;
    int nk = 0;
    double rmax = 0.0, lperp[3] = {0, 0, 0};
    std::vector<double> ljtab;
    std::vector<double2> cube, e0, e1, e2;
    size_t bytes = 0, off_c = 0, off_e0 = 0, off_e1 = 0, off_e2 = 0, off_t1 = 0, off_t2 = 0;
    char* dev_base = nullptr;

    void prepare(Framework& fw, const GridRequest& req);
    void fit_poly(int min_deg);                 // erf polynomial of degree >= min_deg (Coulomb only)
    // FP64: the dropped part of the erf kernel below the last bits of a term; MIXED: below 1e-12
    // of it (a term of 3e4 K then carries under 3e-8 K)
    double poly_tolerance() const { return precision == ALEATOR_PRECISION_FP64 ? 1e-17 : 1e-12; }
    double tile_radius(int shape) const;
    bool shape_fits(int shape) const;
    void set_shape(int s, int chunk);           // chunk: atoms per list-build pass
    bool chunk_fits(int s, int chunk) const;    // allowed with tile shape s
    int shape_used() const { return shape; }
    void bind(char* base);                 // device addresses of this job's `bytes`
    void upload(hipStream_t s) const;      // host constants -> device
    void launch_pre(hipStream_t s) const;  // reciprocal contractions (Coulomb only)
};

int tile_blocks(const int32_t n[3], int shape);
// deg: erf polynomial degree of every framework of the launch (0: no Coulomb table); ljk: the most
// general LJ truncation of the launch's frameworks (0 truncated, 1 shifted, 2 linear-force-shifted)
void launch_tiles(int precision, int shape, int ng, int deg, int ljk, const GridArgs* d_descs, const int* d_prefix,
                  int n_desc, int blocks, hipStream_t s);
// truncation kind of one framework (ALEATOR_LJ_*) in the order used by launch_tiles
inline int lj_kind(int trunc) {
    return trunc == ALEATOR_LJ_LINEAR_FORCE_SHIFTED ? 2 : trunc == ALEATOR_LJ_SHIFTED ? 1 : 0;
}
// the tile kernels of each precision and LJ truncation kind (grid_tile_*.hip)
void launch_tiles_fp64_trunc(int shape, int ng, int deg, const GridArgs* d, const int* prefix, int n_desc, int blocks,
                             hipStream_t s);
void launch_tiles_fp64_shift(int shape, int ng, int deg, const GridArgs* d, const int* prefix, int n_desc, int blocks,
                             hipStream_t s);
void launch_tiles_fp64_lfs(int shape, int ng, int deg, const GridArgs* d, const int* prefix, int n_desc, int blocks,
                           hipStream_t s);
void launch_tiles_mixed_trunc(int shape, int ng, int deg, const GridArgs* d, const int* prefix, int n_desc, int blocks,
                              hipStream_t s);
void launch_tiles_mixed_shift(int shape, int ng, int deg, const GridArgs* d, const int* prefix, int n_desc, int blocks,
                              hipStream_t s);
void launch_tiles_mixed_lfs(int shape, int ng, int deg, const GridArgs* d, const int* prefix, int n_desc, int blocks,
                            hipStream_t s);

}  // namespace aleator_gpu
