// plx_amg.hpp -- smoothed-aggregation algebraic multigrid for the assembled elasticity operator
// (host set-up in plx_amg.cpp, device V-cycle in plx_gf.hip).
//
// The role of PETSc's GAMG preconditioner that PyLith selects for elasticity
// (share/settings/solver_elasticity.cfg, materials/Elasticity.cc:307-318): aggregates of
// strongly connected nodes, the six rigid-body modes as near-null space (translations and
// rotations from the node coordinates), a per-aggregate QR for the tentative prolongator, one
// damped-Jacobi smoothing step of it, Galerkin coarse operators, Chebyshev smoothing on every level
// and a dense inverse on the coarsest.
#pragma once

#include <cstdint>
#include <memory>
#include <string>
#include <utility>
#include <vector>

namespace plx {

// Allocator that default-initialises (leaves doubles and ints unset) on resize: the set-up's
// large arrays are written in full right after they are sized, so a zero fill would only cost
// a serial pass over gigabytes.
template <typename T>
struct DefaultInitAlloc : std::allocator<T> {
    template <typename U>
    struct rebind {
        using other = DefaultInitAlloc<U>;
    };
    DefaultInitAlloc() = default;
    template <typename U>
    DefaultInitAlloc(const DefaultInitAlloc<U> &) noexcept {}
    template <typename U>
    void construct(U *p) noexcept {
        ::new (static_cast<void *>(p)) U;
    }
    template <typename U, typename... Args>
    void construct(U *p, Args &&...args) {
        ::new (static_cast<void *>(p)) U(std::forward<Args>(args)...);
    }
};
using AmgIdx = std::vector<int32_t, DefaultInitAlloc<int32_t>>;
using AmgVal = std::vector<double, DefaultInitAlloc<double>>;
using AmgValF = std::vector<float, DefaultInitAlloc<float>>;

struct AmgLevel {
    int n = 0;       // block rows
    int b = 0;       // block size (3 on the finest level, 6 below)
    int nc = 0;      // block rows of the next level
    int bc = 6;      // block size of the next level (3 below a P2 -> P1 level)
    // operator of this level (block CSR, b x b blocks row-major); level 0 uses the caller's copy
    AmgIdx Aptr, Acol;
    AmgVal Aval;
    // prolongator (n rows, b x bc blocks) and its transpose (nc rows, bc x b blocks)
    AmgIdx Pptr, Pcol, Rptr, Rcol;
    AmgVal Pval, Rval;
    AmgVal Dinv;  // [n][b][b] inverse diagonal blocks
    double lmax = 0.0;         // upper bound of the spectrum of D^-1 A
};

struct AmgHierarchy {
    std::vector<AmgLevel> levels;  // levels[l] maps to levels[l+1] (or to the coarsest)
    int coarse_n = 0;              // dofs of the coarsest operator
    std::vector<double> coarse_inv;  // dense inverse of the coarsest operator [coarse_n][coarse_n]
};

// Build the hierarchy of the 3x3-block operator (ptr, col, val) on n nodes.  fixed[3 n] marks
// identity rows/columns (constrained dofs, positive split nodes); coords [n][3].  Aggregates
// grow over the strong couplings of each level, ||A_ij||_F >= th sqrt(||A_ii||_F ||A_jj||_F)
// with th = theta on the finest level and theta / 2 below, and the prolongator is smoothed with
// the operator filtered to those couplings (theta = 0: every coupling of the node graph).
int amg_setup(int n, const int32_t *ptr, const int32_t *col, const double *val, const double *coords,
              const uint8_t *fixed, int coarse_max, int max_levels, AmgHierarchy *h, std::string *err,
              double theta = 0.0);

// P2 operators: a first level coarsens the P2 nodes onto the mesh vertices by the P1
// interpolation (a vertex node takes its vertex's value, an edge node the mean of its two
// vertices': parents par[par_ptr[i] ..] with weights parw), the Galerkin product gives the P1
// operator, and amg_setup builds the rest of the hierarchy on it (vertex coordinates vxyz
// [nv][3], threshold theta).  A vertex dof is fixed when its own node's dof is.  lmax0 > 0: the
// P2 level's lambda_max estimate of D^-1 A, computed by the caller (else by power iteration
// here, as level_lmax does for every level).
template <typename V>  // operator values: double or float (instantiated for both)
int amg_setup_p2(int n, const int32_t *ptr, const int32_t *col, const V *val, const uint8_t *fixed, int nv,
                 const int32_t *par_ptr, const int32_t *par, const double *parw, const double *vxyz, int coarse_max,
                 int max_levels, AmgHierarchy *h, std::string *err, double theta = 0.0, double lmax0 = 0.0);

// Start vector of the lambda_max power iterations (n * b entries; the same on every host).
std::vector<double> lmax_start_vector(size_t n);

}  // namespace plx
