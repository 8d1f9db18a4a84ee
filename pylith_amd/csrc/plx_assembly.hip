// plx_assembly.hip -- gfx950 kernels of the residual / Jacobian-action apply.
//
// One persistent workgroup walks a sequence of tiles of cells (one or two cells per thread):
//   1. the tile's nodal field and vertex coordinates are staged in LDS through the tile's
//      node / vertex lists (coalesced where node numbers run contiguously); the next tile's
//      gathers are issued into registers before the current tile's arithmetic;
//   2. each thread evaluates its element vectors (Q1 hex: sum-factorized 2x2x2 Gauss or the
//      modal form on affine cells; P2 tet: 14-point rule or its P1-moment form) from the
//      device twin of the pointwise kernel;
//   3. element vectors are accumulated into an LDS node buffer either colour by colour (fixed
//      summation order, bitwise reproducible) or with LDS atomic adds into an FP64 accumulator
//      (the default; the order of the additions into one node, and so the last bits, may vary);
//   4. nodes owned by the tile are written straight to the output vector; shared nodes leave
//      as partial sums in the halo buffer, reduced in slot order by a node-ordered second kernel.
// No global atomics are used.
#include <algorithm>
#include <map>
#include <mutex>

#include "plx_element.hpp"
#include "plx_internal.hpp"

namespace plx {

// Reference tables of the P2 element in constant memory (each device holds its own copy).
__constant__ P2Tables c_p2;
constexpr int kMaxDevices = 64;
static bool g_tables_ready[kMaxDevices] = {};
static std::mutex g_device_mu;

struct DevTab {
    __device__ static double w(int q) { return c_p2.w[q]; }
    __device__ static double dphi(int q, int b, int k) { return c_p2.dphi[q][b][k]; }
    __device__ static double lam(int q, int v) { return c_p2.lam[q][v]; }
};

// Called with the DM's device current (plx_dm_create).
static int ensure_tables(plx_dm dm) {
    std::lock_guard<std::mutex> lock(g_device_mu);
    if (g_tables_ready[dm->device]) return 0;
    P2Tables t;
    p2_fill_tables(&t);
    PLX_HIP(dm, hipMemcpyToSymbol(HIP_SYMBOL(c_p2), &t, sizeof(t)));
    g_tables_ready[dm->device] = true;
    return 0;
}

// --------------------------------------------------------------------------------------------
// Scratch of the allocation-free residual entry point (plx_compute_residual_noalloc): the
// shared-node partial sums of one apply live in this file-scope device array instead of the
// key's own halo buffer.  Each device holds its own copy until the process ends; host code
// reaches it only through the address hipGetSymbolAddress returns on that device (the array's
// name is a host-side stand-in).  Every call on one device shares the array, so calls that
// overlap on one device race.  The apply writes every halo slot before the reduction reads it,
// so nothing needs clearing between calls.
// --------------------------------------------------------------------------------------------
constexpr size_t kScratchBytes = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];
static void *g_scratch_ptr[kMaxDevices] = {};

int device_setup(plx_dm dm) {
    int dev = 0;
    PLX_HIP(dm, hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return set_error(dm, 1, "device index beyond the supported range");
    dm->device = dev;
    {
        std::lock_guard<std::mutex> lock(g_device_mu);
        if (!g_scratch_ptr[dev]) {
            void *p = nullptr;
            PLX_HIP(dm, hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
            if (!p) return set_error(dm, 76, "scratch array has no device address");
            g_scratch_ptr[dev] = p;
        }
    }
    return ensure_tables(dm);
}

// --------------------------------------------------------------------------------------------
// Tile kernel.
// --------------------------------------------------------------------------------------------
struct ApplyArgs {
    TilePlanDev p;
    const uint8_t *color;   // [stride] colour of each key cell (tile-local greedy colouring)
    const int32_t *tile_ncolor;
    const void *mat;        // [2][stride] lambda, mu
    const void *u;
    void *r;
    void *halo;
    int mode;
    double scale;           // multiplies the element vectors (e.g. -1 for the SPD operator -J)
    const int32_t *tile_list;  // tiles this launch covers (nullptr: all, in order)
    int num_list;
    const void *records;       // packed 16-bit local closure (+ cone) per key cell
    const double *prestress;   // Q1: per-point stress added to sigma, [6][8][stride] (nullptr: none)
};

// Cell solution values read from the tile's LDS buffer (component c, local basis b).
template <typename Real, int NB>
struct LdsAcc {
    const Real *sU;
    int lc[NB];
    __device__ __forceinline__ Real operator()(int c, int b) const { return sU[3 * lc[b] + c]; }
};
template <typename Real>
struct LdsAccQ1 {
    const Real *sU;
    int lc[8];
    __device__ __forceinline__ Real operator()(int c, int l) const { return sU[3 * lc[q1_lex(l)] + c]; }
};

constexpr int kAccColored = 0, kAccAtomic = 1;

// Map slot i to a tile so that each of the 8 XCDs (round-robin workgroup dispatch) walks a
// contiguous range of tiles: neighbouring tiles then share that XCD's L2.
__device__ __forceinline__ int xcd_tile(int i, int n) {
    const int per = n >> 3, rem = n & 7;
    const int x = i & 7, j = i >> 3;
    return x * per + min(x, rem) + j;
}

// Workgroup shape per cell type and vector type: NT threads, CPT cells per thread, PFN node
// slots and PFV vertex slots per thread for the register prefetch of the next tile.
template <int CT, typename Real>
struct Shape;
template <typename Real>
struct Shape<PLX_HEX_Q1, Real> {
    // RECW: uint4 words per cell record (8 x uint16 closure)
    static constexpr int NT = 256, CPT = 2, NB = 8, NV = 8, PFN = 3, PFV = 0, RECW = 1;
};
// P2 FP64: 4-wave workgroups over the 384-cell tile (cells 0-255 on the first pass, 256-383 on
// the second), two workgroups per CU at the kernel's register count (2 waves per SIMD: every
// wave slot of the CU busy; measured 6 % faster than 3-wave workgroups with two full passes).
// FP32 takes the same shape (three workgroups per CU at its 142 VGPRs): two cells a thread give
// the two-phase fetch its split point (3 % faster than 6-wave workgroups of one cell a thread).
// record: 10 x uint16 closure, 4 x uint16 cone, 2 padding = 2 uint4 words
template <>
struct Shape<PLX_TET_P2, double> {
    static constexpr int NT = 256, CPT = 2, NB = 10, NV = 4, PFN = 3, PFV = 1, RECW = 2;
};
template <>
struct Shape<PLX_TET_P2, float> {
    static constexpr int NT = 256, CPT = 2, NB = 10, NV = 4, PFN = 3, PFV = 1, RECW = 2;
};

// Entry k of a packed 16-bit cell record.
template <int W>
__device__ __forceinline__ int rec_get(const uint4 (&r)[W], int k) {
    const uint4 &w = r[k >> 3];
    const int h = (k >> 1) & 3;
    const unsigned v = h == 0 ? w.x : h == 1 ? w.y : h == 2 ? w.z : w.w;
    return (k & 1) ? (int)(v >> 16) : (int)(v & 0xffffu);
}

// One tile's data on its way from HBM: gathered node values and coordinates, the cells'
// local closures, moduli and colours (registers; the next tile's loads overlap the current
// tile's arithmetic).  The gathers go out in two phases: the node / vertex ids (with the cells'
// records and moduli) before the current tile's first cell, the values at those ids after it,
// so the ids' latency hides behind that cell's arithmetic.
template <typename Real, int CT>
struct Stage {
    using S = Shape<CT, Real>;
    int tile, c0, nc, n0, nn, v0, nvt, nown, hoff;
    int gn[S::PFN];                          // node ids between the two phases (-1: past the tile)
    int gv[CT == PLX_HEX_Q1 ? 1 : S::PFV];   // P2 vertex ids
    Real pu[S::PFN][3];
    double px[CT == PLX_HEX_Q1 ? S::PFN : S::PFV][3];
    uint4 rec[S::CPT][S::RECW];  // packed 16-bit local closure (and cone) of each cell
    Real lam[S::CPT], mu[S::CPT];
    int color[S::CPT];
};

// First phase of the two-phase fetch: extents, node (and P2 vertex) ids, the cells' records and
// moduli.
template <typename Real, int CT, int ACC>
__device__ __forceinline__ void fetch_ids(Stage<Real, CT> &T, const ApplyArgs &A, int tile, int tid) {
    using S = Shape<CT, Real>;
    const TilePlanDev &P = A.p;
    T.tile = tile;
    T.c0 = P.tile_cell_ptr[tile];
    T.nc = P.tile_cell_ptr[tile + 1] - T.c0;
    T.n0 = P.tile_node_ptr[tile];
    T.nn = P.tile_node_ptr[tile + 1] - T.n0;
    T.v0 = P.tile_vert_ptr[tile];
    T.nvt = P.tile_vert_ptr[tile + 1] - T.v0;
    T.nown = P.tile_nown[tile];
    T.hoff = P.tile_halo_off[tile];
#pragma unroll
    for (int j = 0; j < S::PFN; ++j) {
        const int k = tid + j * S::NT;
        T.gn[j] = k < T.nn ? P.tile_nodes[T.n0 + k] : -1;
    }
    if constexpr (CT != PLX_HEX_Q1) {
#pragma unroll
        for (int j = 0; j < S::PFV; ++j) {
            const int k = tid + j * S::NT;
            T.gv[j] = k < T.nvt ? P.tile_verts[T.v0 + k] : -1;
        }
    }
    const Real *mat = static_cast<const Real *>(A.mat);
#pragma unroll
    for (int i = 0; i < S::CPT; ++i) {
        const int cell = tid + i * S::NT;
        if (cell < T.nc) {
            const int pos = T.c0 + cell;
            const uint4 *recs = reinterpret_cast<const uint4 *>(A.records);
#pragma unroll
            for (int w = 0; w < S::RECW; ++w) T.rec[i][w] = recs[(size_t)pos * S::RECW + w];
            // unscaled here (scaled where used): the fetch then has no use of these loads to wait for
            T.lam[i] = mat[pos];
            T.mu[i] = mat[P.stride + pos];
            if (ACC == kAccColored) T.color[i] = A.color[pos];
        }
    }
}

// Second phase: the values and coordinates at the ids of fetch_ids.
template <typename Real, int CT>
__device__ __forceinline__ void fetch_values(Stage<Real, CT> &T, const ApplyArgs &A) {
    using S = Shape<CT, Real>;
    const TilePlanDev &P = A.p;
    const Real *u = static_cast<const Real *>(A.u);
#pragma unroll
    for (int j = 0; j < S::PFN; ++j) {
        if (T.gn[j] >= 0) {
            const size_t g = (size_t)T.gn[j];
#pragma unroll
            for (int c = 0; c < 3; ++c) T.pu[j][c] = u[3 * g + c];
            if constexpr (CT == PLX_HEX_Q1) {
#pragma unroll
                for (int d = 0; d < 3; ++d) T.px[j][d] = P.coords[3 * g + d];
            }
        }
    }
    if constexpr (CT != PLX_HEX_Q1) {
#pragma unroll
        for (int j = 0; j < S::PFV; ++j) {
            if (T.gv[j] >= 0) {
                const size_t g = (size_t)T.gv[j];
#pragma unroll
                for (int d = 0; d < 3; ++d) T.px[j][d] = P.coords[3 * g + d];
            }
        }
    }
}

// One phase (the first tile of every kernel; the kernels that keep one phase throughout).
template <typename Real, int CT, int ACC>
__device__ __forceinline__ void fetch(Stage<Real, CT> &T, const ApplyArgs &A, int tile, int tid) {
    using S = Shape<CT, Real>;
    const TilePlanDev &P = A.p;
    const Real *u = static_cast<const Real *>(A.u);
    T.tile = tile;
    T.c0 = P.tile_cell_ptr[tile];
    T.nc = P.tile_cell_ptr[tile + 1] - T.c0;
    T.n0 = P.tile_node_ptr[tile];
    T.nn = P.tile_node_ptr[tile + 1] - T.n0;
    T.v0 = P.tile_vert_ptr[tile];
    T.nvt = P.tile_vert_ptr[tile + 1] - T.v0;
    T.nown = P.tile_nown[tile];
    T.hoff = P.tile_halo_off[tile];
    if constexpr (CT == PLX_HEX_Q1) {
        // every node id first, then the values: one wait between the two, not one per slot
        int gn[S::PFN];
#pragma unroll
        for (int j = 0; j < S::PFN; ++j) {
            const int k = tid + j * S::NT;
            gn[j] = k < T.nn ? P.tile_nodes[T.n0 + k] : -1;
        }
#pragma unroll
        for (int j = 0; j < S::PFN; ++j) {
            if (gn[j] >= 0) {
                const size_t g = (size_t)gn[j];
#pragma unroll
                for (int c = 0; c < 3; ++c) T.pu[j][c] = u[3 * g + c];
#pragma unroll
                for (int d = 0; d < 3; ++d) T.px[j][d] = P.coords[3 * g + d];
            }
        }
    } else {
        // P2: id and values slot by slot (measured faster here than all ids first)
#pragma unroll
        for (int j = 0; j < S::PFN; ++j) {
            const int k = tid + j * S::NT;
            if (k < T.nn) {
                const size_t g = (size_t)P.tile_nodes[T.n0 + k];
#pragma unroll
                for (int c = 0; c < 3; ++c) T.pu[j][c] = u[3 * g + c];
            }
        }
#pragma unroll
        for (int j = 0; j < S::PFV; ++j) {
            const int k = tid + j * S::NT;
            if (k < T.nvt) {
                const size_t g = (size_t)P.tile_verts[T.v0 + k];
#pragma unroll
                for (int d = 0; d < 3; ++d) T.px[j][d] = P.coords[3 * g + d];
            }
        }
    }
    const Real *mat = static_cast<const Real *>(A.mat);
#pragma unroll
    for (int i = 0; i < S::CPT; ++i) {
        const int cell = tid + i * S::NT;
        if (cell < T.nc) {
            const int pos = T.c0 + cell;
            const uint4 *recs = reinterpret_cast<const uint4 *>(A.records);
#pragma unroll
            for (int w = 0; w < S::RECW; ++w) T.rec[i][w] = recs[(size_t)pos * S::RECW + w];
            // unscaled here (scaled where used): the fetch then has no use of these loads to wait for
            T.lam[i] = mat[pos];
            T.mu[i] = mat[P.stride + pos];
            if (ACC == kAccColored) T.color[i] = A.color[pos];
        }
    }
}

// Element paths of the tile kernel: the factorized form (kernels linear in the gradient with
// per-cell coefficients, affine cells), the quadrature-point form on affine cells, and the
// quadrature-point form on any Q1 cell (P2 tets are affine: their quadrature-point form is
// kPathQpAffine).
enum { kPathFactored = 0, kPathQpAffine = 1, kPathQpGeneral = 2 };

// Persistent, software-pipelined tile kernel: each workgroup walks its slots; while it works
// on one tile, the gathers of its next tile are in flight.
template <typename Real, int CT, int PATH, int ACC, int MINW>
__global__ void __launch_bounds__((Shape<CT, Real>::NT), MINW) apply_tiles(ApplyArgs A) {
    using S = Shape<CT, Real>;
    constexpr int NT = S::NT, NB = S::NB, NV = S::NV;
    static_assert(S::PFN * NT >= 768 && S::CPT * NT >= (CT == PLX_HEX_Q1 ? 512 : 384), "tile limits");
    static_assert(CT == PLX_HEX_Q1 || S::PFV * NT >= 192, "vertex limit");
    extern __shared__ __align__(16) unsigned char smem[];
    const int tid = threadIdx.x;
    const int G = gridDim.x;
    int slot = blockIdx.x;
    if (slot >= A.num_list) return;
    auto tile_of = [&](int s) {
        const int t = xcd_tile(s, A.num_list);
        return A.tile_list ? A.tile_list[t] : t;
    };
    Stage<Real, CT> cur;
    fetch<Real, CT, ACC>(cur, A, tile_of(slot), tid);
    // Q1: the first tile's loads complete here, as every later tile's complete before the
    // loop's back edge, so the loop is entered with nothing in flight.  Otherwise the wait
    // analysis, merging this path into the loop, puts a wait for them inside the cell
    // arithmetic, where in every later tile it also waits for the next tile's gathers
    // (Q1 192^3 FP64: 0.300 -> 0.275 ms; the P2 kernels measured no gain)
    if constexpr (CT == PLX_HEX_Q1) __builtin_amdgcn_s_waitcnt(0);
    Real *r = static_cast<Real *>(A.r);
    Real *halo = static_cast<Real *>(A.halo);
    const bool insert = A.mode == PLX_INSERT_VALUES;
    // LDS atomics accumulate in FP64 whatever the vector type (native ds_add_f64; FP32 calls
    // round once at the write-out)
    using Acc = typename std::conditional<ACC == kAccAtomic, double, Real>::type;
    while (true) {
        // LDS layout of this tile: coordinates, the accumulator, then nodal values
        const int nxv = CT == PLX_HEX_Q1 ? cur.nn : cur.nvt;
        double *sX = reinterpret_cast<double *>(smem);
        Acc *sR = reinterpret_cast<Acc *>(sX + 3 * nxv);
        Real *sU = reinterpret_cast<Real *>(sR + 3 * cur.nn);
#pragma unroll
        for (int j = 0; j < S::PFN; ++j) {
            const int k = tid + j * NT;
            if (k < cur.nn) {
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    sU[3 * k + c] = cur.pu[j][c];
                    sR[3 * k + c] = Acc(0);
                }
                if constexpr (CT == PLX_HEX_Q1) {
#pragma unroll
                    for (int d = 0; d < 3; ++d) sX[3 * k + d] = cur.px[j][d];
                }
            }
        }
        if constexpr (CT != PLX_HEX_Q1) {
#pragma unroll
            for (int j = 0; j < S::PFV; ++j) {
                const int k = tid + j * NT;
                if (k < cur.nvt) {
#pragma unroll
                    for (int d = 0; d < 3; ++d) sX[3 * k + d] = cur.px[j][d];
                }
            }
        }
        __syncthreads();

        // The next tile's loads go out in two phases (see Stage) in the factored Q1 kernels and
        // the P2 kernels (measured -3 % each for the factored kernels with LDS atomics or colour
        // phases, FP64 and FP32, -1.7 % for the P2 quadrature-point kernel); the Q1
        // quadrature-point kernels overflow their registers with it and keep one phase.  The P2 kernel makes room by
        // loading its write-out ids after the cells (below) instead of early.
        constexpr bool P2SPLIT = CT != PLX_HEX_Q1 && (PATH == kPathFactored || PATH == kPathQpAffine);
        constexpr bool SPLIT = (CT == PLX_HEX_Q1 && PATH == kPathFactored && S::CPT > 1) || P2SPLIT;
        // node ids of this tile's owned write-out, loaded before the next tile's gathers so that
        // waiting for them does not wait for those (not for the FP64 Q1 kernel: its registers
        // are full, and the extra ids spill; not for the two-phase P2 kernel)
        constexpr bool EARLY = (CT != PLX_HEX_Q1 || sizeof(Real) == 4) && !P2SPLIT;
        constexpr int NW = EARLY ? S::PFN : 1;
        int wn[NW];
        if constexpr (EARLY) {
#pragma unroll
            for (int j = 0; j < NW; ++j) {
                const int k = tid + j * NT;
                wn[j] = k < cur.nown ? A.p.tile_nodes[cur.n0 + k] : 0;
            }
        }
        const int next_slot = slot + G;
        const bool has_next = next_slot < A.num_list;
        Stage<Real, CT> nxt;
        // the next tile's loads are issued at a raised wave priority, so the SIMD issues them
        // ahead of the arithmetic of its other waves (P2 FP64 kernel -1.3 %; the write-out below
        // at priority 1: Q1 -1.4 %)
        __builtin_amdgcn_s_setprio(2);
        if (has_next) {
            if constexpr (SPLIT) fetch_ids<Real, CT, ACC>(nxt, A, tile_of(next_slot), tid);
            else fetch<Real, CT, ACC>(nxt, A, tile_of(next_slot), tid);
        }
        __builtin_amdgcn_s_setprio(0);

#pragma unroll
        for (int i = 0; i < S::CPT; ++i) {
            const bool active = tid + i * NT < cur.nc;
            int lc[NB];
#pragma unroll
            for (int b = 0; b < NB; ++b) lc[b] = rec_get(cur.rec[i], b);
            Real re[3][NB];
            if (active) {
                const Real lam = (Real)A.scale * cur.lam[i], mu = (Real)A.scale * cur.mu[i];
                if constexpr (CT == PLX_HEX_Q1) {
                    Real rl[3][8];
                    const LdsAccQ1<double> xacc{sX, {lc[0], lc[1], lc[2], lc[3], lc[4], lc[5], lc[6], lc[7]}};
                    const LdsAccQ1<Real> acc{sU, {lc[0], lc[1], lc[2], lc[3], lc[4], lc[5], lc[6], lc[7]}};
                    if constexpr (PATH == kPathFactored) {
                        // affine tiles only (the plan routes other tiles to the quadrature-point kernel)
                        double J[9];
                        q1_affine_jacobian(xacc, J);
                        q1_element_affine<Real>(J, acc, lam, mu, rl);
                    } else {
                        const double *pre = A.prestress ? A.prestress + cur.c0 + tid + i * NT : nullptr;
                        q1_element_qp<Real, PATH == kPathQpAffine>(xacc, acc, lam, mu, rl, pre, (size_t)A.p.stride,
                                                                   (Real)A.scale);
                    }
#pragma unroll
                    for (int l = 0; l < 8; ++l)
#pragma unroll
                        for (int c = 0; c < 3; ++c) re[c][q1_lex(l)] = rl[c][l];
                } else {
                    double X[4][3];
#pragma unroll
                    for (int v = 0; v < NV; ++v) {
                        const int lv = rec_get(cur.rec[i], NB + v);
#pragma unroll
                        for (int d = 0; d < 3; ++d) X[v][d] = sX[3 * lv + d];
                    }
                    if constexpr (PATH == kPathFactored) {
                        LdsAcc<Real, 10> acc{sU, {lc[0], lc[1], lc[2], lc[3], lc[4], lc[5], lc[6], lc[7], lc[8], lc[9]}};
                        p2_element_linear<Real>(X, acc, lam, mu, c_p2.wdiag, c_p2.woff, re);
                    } else {
                        Real U[3][10];
#pragma unroll
                        for (int b = 0; b < NB; ++b)
#pragma unroll
                            for (int c = 0; c < 3; ++c) U[c][b] = sU[3 * lc[b] + c];
                        p2_element<Real, DevTab>(X, U, lam, mu, re);
                    }
                }
            }
            if constexpr (ACC == kAccAtomic) {
                // LDS atomics: no colour phases; the order of the additions into a node varies
                if (active) {
#pragma unroll
                    for (int b = 0; b < NB; ++b)
#pragma unroll
                        for (int c = 0; c < 3; ++c) atomicAdd(&sR[3 * lc[b] + c], (Acc)re[c][b]);
                }
            } else {
                // cells of one colour share no node: fixed summation order (cells i = 0 of the
                // workgroup colour by colour, then cells i = 1), reproducible bits
                const int ncolor = A.tile_ncolor[cur.tile];
                for (int col = 0; col < ncolor; ++col) {
                    __syncthreads();
                    if (active && cur.color[i] == col) {
#pragma unroll
                        for (int b = 0; b < NB; ++b)
#pragma unroll
                            for (int c = 0; c < 3; ++c) sR[3 * lc[b] + c] += re[c][b];
                    }
                }
            }
            if constexpr (SPLIT) {
                if (i == 0 && has_next) {
                    __builtin_amdgcn_s_setprio(2);
                    fetch_values<Real, CT>(nxt, A);
                    __builtin_amdgcn_s_setprio(0);
                }
            }
        }
        __syncthreads();
        __builtin_amdgcn_s_setprio(1);

        if constexpr (EARLY) {
#pragma unroll
            for (int j = 0; j < NW; ++j) {
                const int k = tid + j * NT;
                if (k < cur.nown) {
                    const size_t g = (size_t)wn[j];
#pragma unroll
                    for (int c = 0; c < 3; ++c) {
                        const Real v = (Real)sR[3 * k + c];
                        r[3 * g + c] = insert ? v : r[3 * g + c] + v;
                    }
                }
            }
        } else if constexpr (P2SPLIT) {
            // the write-out ids, all slots at once (clamped positions, no branch)
            int wl[S::PFN];
#pragma unroll
            for (int j = 0; j < S::PFN; ++j) wl[j] = A.p.tile_nodes[cur.n0 + min(tid + j * NT, max(cur.nown - 1, 0))];
#pragma unroll
            for (int j = 0; j < S::PFN; ++j) {
                const int k = tid + j * NT;
                if (k < cur.nown) {
                    const size_t g = (size_t)wl[j];
#pragma unroll
                    for (int c = 0; c < 3; ++c) {
                        const Real v = (Real)sR[3 * k + c];
                        r[3 * g + c] = insert ? v : r[3 * g + c] + v;
                    }
                }
            }
        } else {
            for (int k = tid; k < cur.nown; k += NT) {
                const size_t g = (size_t)A.p.tile_nodes[cur.n0 + k];
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    const Real v = (Real)sR[3 * k + c];
                    r[3 * g + c] = insert ? v : r[3 * g + c] + v;
                }
            }
        }
        const size_t hoff = (size_t)cur.hoff;
        for (int k = cur.nown + tid; k < cur.nn; k += NT) {
            const size_t s = hoff + (size_t)(k - cur.nown);
#pragma unroll
            for (int c = 0; c < 3; ++c) halo[3 * s + c] = (Real)sR[3 * k + c];
        }
        __builtin_amdgcn_s_setprio(0);
        if (!has_next) break;
        __syncthreads();  // every read of this tile's LDS precedes the next tile's staging
        cur = nxt;
        slot = next_slot;
    }
}

// Shared nodes: sum the tiles' partials in tile order.
template <typename Real>
__global__ void __launch_bounds__(256) reduce_halo(int nshared, const int32_t *__restrict__ node,
                                                   const int32_t *__restrict__ ptr,
                                                   const int32_t *__restrict__ slot,
                                                   const Real *__restrict__ halo, Real *r, int mode) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nshared) return;
    const int j0 = ptr[s], j1 = ptr[s + 1];
    const size_t g = 3 * (size_t)node[s];
    Real a0 = Real(0), a1 = Real(0), a2 = Real(0);
    // slots four at a time: their ids, then their values, then the sums in slot order (the
    // ids are read at clamped positions inside this node's range; extra lanes are not summed)
    for (int jb = j0; jb < j1; jb += 4) {
        int sl[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) sl[q] = slot[min(jb + q, j1 - 1)];
        Real v[4][3];
#pragma unroll
        for (int q = 0; q < 4; ++q)
#pragma unroll
            for (int c = 0; c < 3; ++c) v[q][c] = halo[3 * (size_t)sl[q] + c];
#pragma unroll
        for (int q = 0; q < 4; ++q)
            if (jb + q < j1) {
                a0 += v[q][0];
                a1 += v[q][1];
                a2 += v[q][2];
            }
    }
    if (mode == PLX_INSERT_VALUES) {
        r[g + 0] = a0;
        r[g + 1] = a1;
        r[g + 2] = a2;
    } else {
        r[g + 0] += a0;
        r[g + 1] += a1;
        r[g + 2] += a2;
    }
}

static TilePlanDev plan_dev(plx_dm dm, const KeyPlan *kp) {
    TilePlanDev p;
    p.num_tiles = kp->num_tiles;
    p.stride = (int)kp->cells.size();
    p.tile_cell_ptr = kp->d_tile_cell_ptr;
    p.tile_node_ptr = kp->d_tile_node_ptr;
    p.tile_nown = kp->d_tile_nown;
    p.tile_nodes = kp->d_tile_nodes;
    p.tile_vert_ptr = kp->d_tile_vert_ptr;
    p.tile_verts = kp->d_tile_verts;
    p.tile_halo_off = kp->d_tile_halo_off;
    p.lclosure = kp->d_lclosure;
    p.lcone = kp->d_lcone;
    p.coords = dm->d_coords;
    return p;
}

// Resident workgroups per CU of a persistent tile kernel at this LDS size, from the occupancy
// calculator (registers, LDS and waves), cached per kernel and size; at most 4.
static int resident_per_cu(const void *kernel, int nt, size_t lds, int fallback) {
    static std::mutex mu;
    static std::map<std::pair<const void *, size_t>, int> cache;
    std::lock_guard<std::mutex> lock(mu);
    const auto key = std::make_pair(kernel, lds);
    const auto it = cache.find(key);
    if (it != cache.end()) return it->second;
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kernel, nt, lds) != hipSuccess || n < 1) n = fallback;
    n = std::min(n, 4);
    cache[key] = n;
    return n;
}

template <typename Real>
static int apply_typed(plx_dm dm, KeyPlan *kp, const void *u, void *r, int mode, double scale,
                       const double *prestress, void *halo, bool log) {
    ApplyArgs A;
    A.scale = scale;
    A.prestress = prestress;
    A.p = plan_dev(dm, kp);
    A.color = kp->d_color;
    A.tile_ncolor = kp->d_tile_ncolor;
    A.mat = std::is_same<Real, double>::value ? (const void *)kp->d_mat64 : (const void *)kp->d_mat32;
    A.u = u;
    A.r = r;
    A.halo = halo ? halo : kp->d_halo;
    A.mode = mode;
    A.records = kp->d_records;
    // a per-point prestress needs the quadrature-point path
    const bool linear = dm->algorithm != PLX_ALGO_QUADRATURE_POINTS && !prestress;
    // AUTO: LDS atomics (FP64 accumulator, native ds_add_f64) for both vector types
    const bool atomic = dm->accumulate != PLX_ACCUMULATE_COLORED;
    const size_t lds = (size_t)kp->max_verts * 3 * sizeof(double) +
                       (size_t)kp->max_nodes * 3 * (sizeof(Real) + (atomic ? sizeof(double) : sizeof(Real)));
    // persistent grid: as many workgroups per CU as the kernel's registers and LDS let reside
    // (a multiple of 8 in all keeps each workgroup on the XCD whose tile range it walks), never
    // more than there are tiles
    if (dm->num_cu <= 0) PLX_HIP(dm, hipDeviceGetAttribute(&dm->num_cu, hipDeviceAttributeMultiprocessorCount, 0));
    const int ncu = std::max(dm->num_cu, 8);
    {
    ScopedEvent ev(dm, log ? "plx:apply_tiles" : nullptr);
#define PLX_LAUNCH_W(CT, PATH, ACC, W, N)                                                             \
    do {                                                                                             \
        const void *kfn = reinterpret_cast<const void *>(&apply_tiles<Real, CT, PATH, ACC, W>);       \
        const int per_cu = resident_per_cu(kfn, Shape<CT, Real>::NT, lds, dm->wg_per_cu);            \
        const int resident = ((ncu * per_cu + 7) / 8) * 8;                                           \
        hipLaunchKernelGGL((apply_tiles<Real, CT, PATH, ACC, W>), dim3(std::min(N, resident)),         \
                           dim3(Shape<CT, Real>::NT), lds, dm->stream, A);                           \
    } while (0)
#define PLX_LAUNCH(CT, PATH, ACC, LIST, N)                                                           \
    do {                                                                                             \
        A.tile_list = LIST;                                                                          \
        A.num_list = N;                                                                              \
        if (N > 0) {                                                                                 \
            if (minw <= 2) PLX_LAUNCH_W(CT, PATH, ACC, 2, N);                                          \
            else if (minw == 3) PLX_LAUNCH_W(CT, PATH, ACC, 3, N);                                     \
            else PLX_LAUNCH_W(CT, PATH, ACC, 4, N);                                                    \
        }                                                                                            \
    } while (0)
    const int minw = dm->min_waves;
    if (dm->cell_type == PLX_HEX_Q1) {
        // affine tiles: the factorized kernel (or the quadrature-point one on affine cells); the
        // rest: the quadrature-point kernel with the Jacobian at each point
        if (linear) {
            if (atomic) {
                PLX_LAUNCH(PLX_HEX_Q1, kPathFactored, kAccAtomic, kp->d_tiles_affine, kp->num_tiles_affine);
                PLX_LAUNCH(PLX_HEX_Q1, kPathQpGeneral, kAccAtomic, kp->d_tiles_general, kp->num_tiles_general);
            } else {
                PLX_LAUNCH(PLX_HEX_Q1, kPathFactored, kAccColored, kp->d_tiles_affine, kp->num_tiles_affine);
                PLX_LAUNCH(PLX_HEX_Q1, kPathQpGeneral, kAccColored, kp->d_tiles_general, kp->num_tiles_general);
            }
        } else {
            if (atomic) {
                PLX_LAUNCH(PLX_HEX_Q1, kPathQpAffine, kAccAtomic, kp->d_tiles_affine, kp->num_tiles_affine);
                PLX_LAUNCH(PLX_HEX_Q1, kPathQpGeneral, kAccAtomic, kp->d_tiles_general, kp->num_tiles_general);
            } else {
                PLX_LAUNCH(PLX_HEX_Q1, kPathQpAffine, kAccColored, kp->d_tiles_affine, kp->num_tiles_affine);
                PLX_LAUNCH(PLX_HEX_Q1, kPathQpGeneral, kAccColored, kp->d_tiles_general, kp->num_tiles_general);
            }
        }
    } else {
        if (linear) {
            if (atomic) PLX_LAUNCH(PLX_TET_P2, kPathFactored, kAccAtomic, nullptr, kp->num_tiles);
            else PLX_LAUNCH(PLX_TET_P2, kPathFactored, kAccColored, nullptr, kp->num_tiles);
        } else {
            if (atomic) PLX_LAUNCH(PLX_TET_P2, kPathQpAffine, kAccAtomic, nullptr, kp->num_tiles);
            else PLX_LAUNCH(PLX_TET_P2, kPathQpAffine, kAccColored, nullptr, kp->num_tiles);
        }
    }
#undef PLX_LAUNCH_W
#undef PLX_LAUNCH
    }
    PLX_HIP(dm, hipGetLastError());
    if (kp->num_shared > 0) {
        ScopedEvent ev(dm, log ? "plx:reduce_halo" : nullptr);
        hipLaunchKernelGGL(reduce_halo<Real>, dim3((kp->num_shared + 255) / 256), dim3(256), 0,
                           dm->stream, kp->num_shared, kp->d_shared_node, kp->d_shared_ptr,
                           kp->d_shared_slot, static_cast<const Real *>(A.halo),
                           static_cast<Real *>(r), mode);
        PLX_HIP(dm, hipGetLastError());
    }
    return 0;
}

// Residual and Jacobian action share the element operator: for isotropic linear elasticity
// J v = F(v); for generalized Maxwell F(u) = J u + (the state term), with J the elastic operator
// at the effective shear modulus (the residual adds the per-point state stress).
int launch_apply(plx_dm dm, KeyPlan *kp, int scalar_type, bool jacobian, const void *u, void *r,
                 int mode, double scale, void *halo, bool log) {
    if (mode != PLX_ADD_VALUES && mode != PLX_INSERT_VALUES) return set_error(dm, 1, "bad insert mode");
    const double *pre = (!jacobian && kp->genmax) ? kp->d_prestress : nullptr;
    if (pre && scalar_type != PLX_FP64) return set_error(dm, 1, "the viscoelastic residual is FP64");
    if (scalar_type == PLX_FP64) return apply_typed<double>(dm, kp, u, r, mode, scale, pre, halo, log);
    if (scalar_type == PLX_FP32) return apply_typed<float>(dm, kp, u, r, mode, scale, pre, halo, log);
    return set_error(dm, 1, "bad scalar type");
}

// Body force / gravity load (fixed per node): r[node] += load.
template <typename Real>
__global__ void __launch_bounds__(256) add_load(int n, const int32_t *__restrict__ node,
                                                const Real *__restrict__ load, Real *r) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const size_t g = 3 * (size_t)node[i];
    r[g + 0] += load[3 * (size_t)i + 0];
    r[g + 1] += load[3 * (size_t)i + 1];
    r[g + 2] += load[3 * (size_t)i + 2];
}

int launch_add_load(plx_dm dm, KeyPlan *kp, int scalar_type, void *r) {
    const int n = kp->num_load;
    if (scalar_type == PLX_FP64)
        hipLaunchKernelGGL(add_load<double>, dim3((n + 255) / 256), dim3(256), 0, dm->stream, n,
                           kp->d_load_node, kp->d_load64, static_cast<double *>(r));
    else
        hipLaunchKernelGGL(add_load<float>, dim3((n + 255) / 256), dim3(256), 0, dm->stream, n,
                           kp->d_load_node, kp->d_load32, static_cast<float *>(r));
    PLX_HIP(dm, hipGetLastError());
    return 0;
}

// --------------------------------------------------------------------------------------------
// Point-block diagonal of J (FP64): D_b[f][g] = -sum_q w|J| [(lambda+mu) dphi_f dphi_g
//   + mu delta_fg |grad phi|^2] for isotropic C, i.e. the diagonal blocks of the element
// matrix of Jf3 = -C.  One thread per cell, global FP64 atomics (set-up path, not timed).
// --------------------------------------------------------------------------------------------
template <int CT>
__global__ void __launch_bounds__(256) pbdiag_cells(TilePlanDev P, const int32_t *cells,
                                                    const int32_t *closure, const int32_t *cone,
                                                    const double *mat, int ncells, double *blocks) {
    constexpr int NB = CT == PLX_HEX_Q1 ? 8 : 10;
    constexpr int NV = CT == PLX_HEX_Q1 ? 8 : 4;
    const int pos = blockIdx.x * blockDim.x + threadIdx.x;
    if (pos >= ncells) return;
    const int cell = cells[pos];
    const double lam = mat[pos], mu = mat[P.stride + pos];
    double X[NV][3];
#pragma unroll
    for (int v = 0; v < NV; ++v)
#pragma unroll
        for (int d = 0; d < 3; ++d) X[v][d] = P.coords[3 * (size_t)cone[(size_t)cell * NV + v] + d];
    double D[NB][6];
#pragma unroll
    for (int b = 0; b < NB; ++b)
#pragma unroll
        for (int i = 0; i < 6; ++i) D[b][i] = 0.0;
    const int nq = CT == PLX_HEX_Q1 ? 8 : kMaxQuad;
    const double g = 0.57735026918962576451;
    const double ref[8][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
                              {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1}};
    for (int q = 0; q < nq; ++q) {
        double dphi[NB][3], J[9] = {0}, Ji[9], w;
        if constexpr (CT == PLX_HEX_Q1) {
            const double xi[3] = {(q & 1) ? g : -g, (q & 2) ? g : -g, (q & 4) ? g : -g};
            for (int b = 0; b < 8; ++b) {
                const double a0 = 1 + ref[b][0] * xi[0], a1 = 1 + ref[b][1] * xi[1], a2 = 1 + ref[b][2] * xi[2];
                dphi[b][0] = 0.125 * ref[b][0] * a1 * a2;
                dphi[b][1] = 0.125 * a0 * ref[b][1] * a2;
                dphi[b][2] = 0.125 * a0 * a1 * ref[b][2];
            }
            for (int v = 0; v < 8; ++v)
                for (int d = 0; d < 3; ++d)
                    for (int k = 0; k < 3; ++k) J[3 * d + k] += X[v][d] * dphi[v][k];
            w = 1.0;
        } else {
            for (int b = 0; b < 10; ++b)
                for (int k = 0; k < 3; ++k) dphi[b][k] = c_p2.dphi[q][b][k];
            for (int d = 0; d < 3; ++d)
                for (int k = 0; k < 3; ++k) J[3 * d + k] = X[k + 1][d] - X[0][d];
            w = c_p2.w[q];
        }
        const double wd = w * inv3(J, Ji);
        for (int b = 0; b < NB; ++b) {
            double gp[3];
            for (int d = 0; d < 3; ++d) gp[d] = dphi[b][0] * Ji[d] + dphi[b][1] * Ji[3 + d] + dphi[b][2] * Ji[6 + d];
            const double n2 = gp[0] * gp[0] + gp[1] * gp[1] + gp[2] * gp[2];
            const double lm = lam + mu;
            D[b][0] -= wd * (lm * gp[0] * gp[0] + mu * n2);
            D[b][1] -= wd * (lm * gp[1] * gp[1] + mu * n2);
            D[b][2] -= wd * (lm * gp[2] * gp[2] + mu * n2);
            D[b][3] -= wd * lm * gp[0] * gp[1];
            D[b][4] -= wd * lm * gp[1] * gp[2];
            D[b][5] -= wd * lm * gp[0] * gp[2];
        }
    }
    for (int b = 0; b < NB; ++b) {
        double *B = blocks + 9 * (size_t)closure[(size_t)cell * NB + b];
        atomicAdd(&B[0], D[b][0]);
        atomicAdd(&B[4], D[b][1]);
        atomicAdd(&B[8], D[b][2]);
        atomicAdd(&B[1], D[b][3]);
        atomicAdd(&B[3], D[b][3]);
        atomicAdd(&B[5], D[b][4]);
        atomicAdd(&B[7], D[b][4]);
        atomicAdd(&B[2], D[b][5]);
        atomicAdd(&B[6], D[b][5]);
    }
}

// Device copies of the closure and cone index (set-up kernels only), each uploaded on first use.
int ensure_mesh_indices(plx_dm dm) {
    if (!dm->d_closure) {
        PLX_HIP(dm, hipMalloc(&dm->d_closure, dm->closure.size() * sizeof(int32_t)));
        PLX_HIP(dm, hipMemcpy(dm->d_closure, dm->closure.data(), dm->closure.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    }
    if (!dm->d_cone) {
        PLX_HIP(dm, hipMalloc(&dm->d_cone, dm->cone.size() * sizeof(int32_t)));
        PLX_HIP(dm, hipMemcpy(dm->d_cone, dm->cone.data(), dm->cone.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    }
    return 0;
}

int launch_pbdiag(plx_dm dm, KeyPlan *kp, double *blocks) {
    const int err = ensure_mesh_indices(dm);
    if (err) return err;
    const int n = (int)kp->cells.size();
    const TilePlanDev P = plan_dev(dm, kp);
    if (dm->cell_type == PLX_HEX_Q1)
        hipLaunchKernelGGL(pbdiag_cells<PLX_HEX_Q1>, dim3((n + 255) / 256), dim3(256), 0, dm->stream, P,
                           kp->d_cells, dm->d_closure, dm->d_cone, kp->d_mat64, n, blocks);
    else
        hipLaunchKernelGGL(pbdiag_cells<PLX_TET_P2>, dim3((n + 255) / 256), dim3(256), 0, dm->stream, P,
                           kp->d_cells, dm->d_closure, dm->d_cone, kp->d_mat64, n, blocks);
    PLX_HIP(dm, hipGetLastError());
    return 0;
}

// --------------------------------------------------------------------------------------------
// Geometry validation: count cells with a non-positive Jacobian determinant at a quadrature
// point (DMPlexComputeCellGeometryFEM rejects them).
// --------------------------------------------------------------------------------------------
template <int CT>
__global__ void __launch_bounds__(256) validate_cells(int ncells, const int32_t *cone,
                                                      const double *coords, int *bad) {
    constexpr int NV = CT == PLX_HEX_Q1 ? 8 : 4;
    const int cell = blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= ncells) return;
    double X[NV][3];
    for (int v = 0; v < NV; ++v)
        for (int d = 0; d < 3; ++d) X[v][d] = coords[3 * (size_t)cone[(size_t)cell * NV + v] + d];
    double minDet = 1e300;
    if constexpr (CT == PLX_HEX_Q1) {
        const double g = 0.57735026918962576451;
        const double ref[8][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
                                  {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1}};
        for (int q = 0; q < 8; ++q) {
            const double xi[3] = {(q & 1) ? g : -g, (q & 2) ? g : -g, (q & 4) ? g : -g};
            double J[9] = {0}, Ji[9];
            for (int b = 0; b < 8; ++b) {
                const double a0 = 1 + ref[b][0] * xi[0], a1 = 1 + ref[b][1] * xi[1], a2 = 1 + ref[b][2] * xi[2];
                const double dp[3] = {0.125 * ref[b][0] * a1 * a2, 0.125 * a0 * ref[b][1] * a2,
                                      0.125 * a0 * a1 * ref[b][2]};
                for (int d = 0; d < 3; ++d)
                    for (int k = 0; k < 3; ++k) J[3 * d + k] += X[b][d] * dp[k];
            }
            minDet = fmin(minDet, inv3(J, Ji));
        }
    } else {
        double J[9], Ji[9];
        for (int d = 0; d < 3; ++d)
            for (int k = 0; k < 3; ++k) J[3 * d + k] = X[k + 1][d] - X[0][d];
        minDet = inv3(J, Ji);
    }
    if (!(minDet > 0.0)) atomicAdd(bad, 1);
}

int validate_geometry(plx_dm dm, int *num_bad) {
    int32_t *d_cone = nullptr;
    int *d_bad = nullptr;
    PLX_HIP(dm, hipMalloc(&d_cone, dm->cone.size() * sizeof(int32_t)));
    PLX_HIP(dm, hipMemcpy(d_cone, dm->cone.data(), dm->cone.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    PLX_HIP(dm, hipMalloc(&d_bad, sizeof(int)));
    PLX_HIP(dm, hipMemset(d_bad, 0, sizeof(int)));
    const int n = dm->num_cells;
    if (dm->cell_type == PLX_HEX_Q1)
        hipLaunchKernelGGL(validate_cells<PLX_HEX_Q1>, dim3((n + 255) / 256), dim3(256), 0, 0, n, d_cone, dm->d_coords, d_bad);
    else
        hipLaunchKernelGGL(validate_cells<PLX_TET_P2>, dim3((n + 255) / 256), dim3(256), 0, 0, n, d_cone, dm->d_coords, d_bad);
    PLX_HIP(dm, hipGetLastError());
    PLX_HIP(dm, hipMemcpy(num_bad, d_bad, sizeof(int), hipMemcpyDeviceToHost));
    (void)hipFree(d_cone);
    (void)hipFree(d_bad);
    return 0;
}

}  // namespace plx

using namespace plx;

extern "C" {

int plx_compute_residual_noalloc(plx_dm dm, int label_value, int part, int scalar_type, double t,
                                 const void *locX, const void *locX_t, void *locF, int mode) {
    (void)t;
    (void)locX_t;
    KeyPlan *kp = nullptr;
    if (const int e = residual_key(dm, label_value, part, locX, locF, &kp)) return e;
    if (kp->pf_residual || !kp->f1)
        return set_error(dm, 1, "PetscPointFn-form keys are not served allocation-free: use plx_compute_residual");
    if (scalar_type != PLX_FP64 && scalar_type != PLX_FP32) return set_error(dm, 1, "bad scalar type");
    const size_t need = (size_t)kp->halo_slots * 3 * (scalar_type == PLX_FP64 ? sizeof(double) : sizeof(float));
    if (need > kScratchBytes)
        return set_error(dm, 1, "the key's shared-node partial sums (" + std::to_string(need) +
                                    " bytes) exceed the allocation-free scratch (" +
                                    std::to_string(kScratchBytes) + " bytes): use plx_compute_residual");
    void *scr = g_scratch_ptr[dm->device];
    if (!scr) return set_error(dm, 76, "scratch array not set up on the DM's device");
    const int err = launch_apply(dm, kp, scalar_type, false, locX, locF, mode, 1.0, scr, false);
    if (err || kp->num_load == 0) return err;
    return launch_add_load(dm, kp, scalar_type, locF);
}

}  // extern "C"
