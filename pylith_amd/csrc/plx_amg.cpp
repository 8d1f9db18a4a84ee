// plx_amg.cpp -- host set-up of the smoothed-aggregation hierarchy (see plx_amg.hpp).
#include "plx_amg.hpp"

#include <algorithm>
#include <cmath>
#include <random>
#include <thread>

namespace plx {
namespace {

int num_threads() {
    const unsigned hc = std::thread::hardware_concurrency();
    return (int)std::max(1u, std::min(16u, hc));
}

// f(lo, hi, tid) over T contiguous chunks of [0, n) on T threads
template <typename F>
void parallel_chunks(int n, int T, F f) {
    if (T <= 1 || n < 1024) {
        f(0, n, 0);
        return;
    }
    std::vector<std::thread> th;
    for (int t = 0; t < T; ++t) {
        const int lo = (int)((int64_t)n * t / T), hi = (int)((int64_t)n * (t + 1) / T);
        th.emplace_back([=, &f] { f(lo, hi, t); });
    }
    for (auto &x : th) x.join();
}

// Inverse of a b x b block (b <= 6) by Gauss-Jordan with partial pivoting; a singular block
// (zero pivot) yields the identity (its rows are decoupled dummy dofs).
void block_inverse(const double *A, int b, double *Ai) {
    // rows equilibrated to unit max norm first (solve (S A) X = S): a partly constrained node's
    // block holds an identity row of 1 next to stiffness rows many orders larger
    alignas(64) double M[6][12];
    bool singular = false;
    for (int i = 0; i < b; ++i) {
        double rs = 0.0;
        for (int j = 0; j < b; ++j) rs = std::max(rs, std::fabs(A[i * b + j]));
        singular = singular || !(rs > 0.0);
        const double s = rs > 0.0 ? 1.0 / rs : 0.0;
        for (int j = 0; j < b; ++j) {
            M[i][j] = A[i * b + j] * s;
            M[i][b + j] = i == j ? s : 0.0;
        }
    }
    for (int c = 0; c < b; ++c) {
        int p = c;
        for (int r = c + 1; r < b; ++r)
            if (std::fabs(M[r][c]) > std::fabs(M[p][c])) p = r;
        if (singular || !(std::fabs(M[p][c]) > 1e-14)) {
            for (int i = 0; i < b * b; ++i) Ai[i] = (i % (b + 1) == 0) ? 1.0 : 0.0;
            return;
        }
        if (p != c)
            for (int j = 0; j < 2 * b; ++j) std::swap(M[p][j], M[c][j]);
        const double inv = 1.0 / M[c][c];
        for (int j = 0; j < 2 * b; ++j) M[c][j] *= inv;
        for (int r = 0; r < b; ++r)
            if (r != c) {
                const double f = M[r][c];
                if (f != 0.0)
                    for (int j = 0; j < 2 * b; ++j) M[r][j] -= f * M[c][j];
            }
    }
    for (int i = 0; i < b; ++i)
        for (int j = 0; j < b; ++j) Ai[i * b + j] = M[i][b + j];
}

// y (b x 6) += alpha * A (b x b) * X (b x 6)
inline void gemm_b6(int b, double alpha, const double *A, const double *X, double *y) {
    for (int i = 0; i < b; ++i)
        for (int k = 0; k < b; ++k) {
            const double a = alpha * A[i * b + k];
            if (a == 0.0) continue;
            for (int j = 0; j < 6; ++j) y[i * 6 + j] += a * X[k * 6 + j];
        }
}

// y (B x 6) += A (B x B) X (B x 6), fixed sizes (unrolled and vectorised)
template <int B>
inline void gemm_bx6(const double *__restrict__ A, const double *__restrict__ X, double *__restrict__ y) {
    for (int i = 0; i < B; ++i)
        for (int k = 0; k < B; ++k) {
            const double a = A[i * B + k];
            for (int j = 0; j < 6; ++j) y[i * 6 + j] += a * X[k * 6 + j];
        }
}

// o (6 x 6) += P^T X for P, X (B x 6)
template <int B>
inline void gemm_t6(const double *__restrict__ P, const double *__restrict__ X, double *__restrict__ o) {
    for (int k = 0; k < B; ++k)
        for (int r = 0; r < 6; ++r) {
            const double p = P[k * 6 + r];
            for (int q = 0; q < 6; ++q) o[r * 6 + q] += p * X[k * 6 + q];
        }
}

// Inverse diagonal blocks of a b x b block operator; false when a row has no diagonal block.
template <typename V>
bool level_dinv(int n, int b, const int32_t *Ap, const int32_t *Ac, const V *Av, AmgVal *Dinv) {
    const int T = num_threads(), bb = b * b;
    Dinv->assign((size_t)n * bb, 0.0);
    std::vector<uint8_t> no_diag(T, 0);
    parallel_chunks(n, T, [&](int lo, int hi, int t) {
        for (int i = lo; i < hi; ++i) {
            int d = -1;
            for (int e = Ap[i]; e < Ap[i + 1]; ++e)
                if (Ac[e] == i) d = e;
            if (d < 0) {
                no_diag[t] = 1;
                return;
            }
            double blk[36];
            for (int k = 0; k < bb; ++k) blk[k] = (double)Av[(size_t)d * bb + k];
            block_inverse(blk, b, Dinv->data() + (size_t)i * bb);
        }
    });
    return std::find(no_diag.begin(), no_diag.end(), 1) == no_diag.end();
}

// 1.1 x the largest eigenvalue of D^-1 A, 12 power iterations from a fixed random vector.
template <typename V>
double level_lmax(int n, int b, const int32_t *Ap, const int32_t *Ac, const V *Av, const AmgVal &Dinv) {
    const int T = num_threads(), bb = b * b;
    const size_t nd_ = (size_t)n * b;
    std::vector<double> v = lmax_start_vector(nd_), w(nd_);
    double lam = 0.0;
    std::vector<double> part(T);
    double vn = 0.0;
    for (double x : v) vn += x * x;
    vn = std::sqrt(vn);
    for (int it = 0; it < 12; ++it) {
        std::fill(part.begin(), part.end(), 0.0);
        parallel_chunks(n, T, [&](int lo, int hi, int t) {
            double acc = 0.0;
            for (int i = lo; i < hi; ++i) {
                double y[6] = {0, 0, 0, 0, 0, 0};
                for (int e = Ap[i]; e < Ap[i + 1]; ++e) {
                    const V *A = &Av[(size_t)e * bb];
                    const double *x = &v[(size_t)Ac[e] * b];
                    for (int r = 0; r < b; ++r) {
                        double sacc = 0.0;
                        for (int c = 0; c < b; ++c) sacc += A[r * b + c] * x[c];
                        y[r] += sacc;
                    }
                }
                const double *Di = &Dinv[(size_t)i * bb];
                for (int r = 0; r < b; ++r) {
                    double sacc = 0.0;
                    for (int c = 0; c < b; ++c) sacc += Di[r * b + c] * y[c];
                    w[(size_t)i * b + r] = sacc / vn;
                    acc += sacc * sacc;
                }
            }
            part[t] = acc;
        });
        double wn = 0.0;
        for (double x : part) wn += x;
        wn = std::sqrt(wn) / vn;  // || D^-1 A v || / || v || with v scaled to unit norm
        lam = wn;
        v.swap(w);
        vn = wn;
    }
    return 1.1 * lam;
}

struct Csr6 {  // block rows of b x 6 blocks, sorted columns
    AmgIdx ptr, col;
    AmgVal val;
};

// Row-parallel construction of a block-CSR matrix: row(i, col, val, tid) appends row i's
// blocks (bs values each); chunks are built per thread and concatenated in row order.
template <typename RowFn>
void build_rows(int n, int bs, Csr6 *out, RowFn row) {
    const int T = num_threads();
    std::vector<Csr6> part(T);
    parallel_chunks(n, T, [&](int lo, int hi, int t) {
        Csr6 &P = part[t];
        P.ptr.assign(1, 0);
        for (int i = lo; i < hi; ++i) {
            row(i, P.col, P.val, t);
            P.ptr.push_back((int32_t)P.col.size());
        }
    });
    // each part's place in the result, then the parts copied in parallel
    std::vector<size_t> coff(T + 1, 0);
    std::vector<int> roff(T + 1, 0);
    for (int t = 0; t < T; ++t) {
        coff[t + 1] = coff[t] + part[t].col.size();
        roff[t + 1] = roff[t] + (part[t].ptr.empty() ? 0 : (int)part[t].ptr.size() - 1);
    }
    out->ptr.assign(n + 1, 0);
    out->col.resize(coff[T]);
    out->val.resize(coff[T] * bs);
    std::vector<std::thread> th;
    for (int t = 0; t < T; ++t)
        th.emplace_back([&, t] {
            Csr6 &P = part[t];
            if (P.ptr.empty()) return;
            for (size_t r = 1; r < P.ptr.size(); ++r) out->ptr[roff[t] + r] = (int32_t)(coff[t] + P.ptr[r]);
            std::copy(P.col.begin(), P.col.end(), out->col.begin() + coff[t]);
            std::copy(P.val.begin(), P.val.end(), out->val.begin() + coff[t] * bs);
            AmgIdx().swap(P.ptr);
            AmgIdx().swap(P.col);
            AmgVal().swap(P.val);
        });
    for (auto &x : th) x.join();
}

// Sparse product rows: out_i = sum_j A_ij * M_j for M with b x 6 blocks.
void spgemm_A_M(int n, int b, const int32_t *Aptr, const int32_t *Acol, const double *Aval, const Csr6 &M,
                int ncols, Csr6 *out) {
    const int T = num_threads();
    std::vector<std::vector<int>> marks(T, std::vector<int>(ncols, -1)), lists(T), orders(T);
    std::vector<std::vector<double>> accs(T);
    const int bb = b * 6;
    build_rows(n, bb, out, [&](int i, AmgIdx &ocol, AmgVal &oval, int t) {
        std::vector<int> &mark = marks[t], &list = lists[t], &order = orders[t];
        std::vector<double> &acc = accs[t];
        list.clear();
        acc.clear();
        for (int e = Aptr[i]; e < Aptr[i + 1]; ++e) {
            const int j = Acol[e];
            const double *A = Aval + (size_t)e * b * b;
            for (int f = M.ptr[j]; f < M.ptr[j + 1]; ++f) {
                const int c = M.col[f];
                if (mark[c] < 0) {
                    mark[c] = (int)list.size();
                    list.push_back(c);
                    acc.resize(acc.size() + bb, 0.0);
                }
                if (b == 3) gemm_bx6<3>(A, M.val.data() + (size_t)f * bb, acc.data() + (size_t)mark[c] * bb);
                else gemm_bx6<6>(A, M.val.data() + (size_t)f * bb, acc.data() + (size_t)mark[c] * bb);
            }
        }
        order.resize(list.size());
        for (size_t u = 0; u < order.size(); ++u) order[u] = (int)u;
        std::sort(order.begin(), order.end(), [&](int x, int y) { return list[x] < list[y]; });
        for (int u : order) {
            ocol.push_back(list[u]);
            oval.insert(oval.end(), acc.begin() + (size_t)u * bb, acc.begin() + (size_t)(u + 1) * bb);
        }
        for (int c : list) mark[c] = -1;
    });
}

}  // namespace

std::vector<double> lmax_start_vector(size_t n) {
    std::mt19937_64 rng(7);
    std::normal_distribution<double> nd;
    std::vector<double> v(n);
    for (auto &x : v) x = nd(rng);
    return v;
}

int amg_setup(int n0, const int32_t *ptr0, const int32_t *col0, const double *val0, const double *coords,
              const uint8_t *fixed0, int coarse_max, int max_levels, AmgHierarchy *h, std::string *err,
              double theta) {
    h->levels.clear();
    h->coarse_inv.clear();
    // current level operator
    int n = n0, b = 3;
    // the current level's operator: the caller's arrays on level 0, the Galerkin products below
    AmgIdx Aptr_v, Acol_v;
    AmgVal Aval_v;
    const int32_t *Ap = ptr0, *Ac = col0;
    const double *Av = val0;
    size_t nnz = (size_t)ptr0[n0];
    std::vector<uint8_t> fixed(fixed0, fixed0 + 3 * (size_t)n0);
    std::vector<double> B;  // near-null space [n*b][6], built per aggregate on level 0
    const int T = num_threads();
    for (int lev = 0;; ++lev) {
        AmgLevel L;
        L.n = n;
        L.b = b;
        const int bb = b * b;
        // inverse diagonal blocks
        if (!level_dinv(n, b, Ap, Ac, Av, &L.Dinv)) {
            *err = "amg: missing diagonal block";
            return 1;
        }
        // aggregation on the node graph; nodes with every dof fixed stay out
        std::vector<uint8_t> excl(n, 0);
        for (int i = 0; i < n; ++i) {
            bool all = true;
            for (int c = 0; c < b; ++c) all = all && fixed[(size_t)i * b + c];
            excl[i] = all;
        }
        // strong couplings: ||A_ij||_F >= th sqrt(||A_ii||_F ||A_jj||_F), th = theta on the finest
        // level and theta / 2 below (th = 0: the whole node graph)
        const double th = lev == 0 ? theta : 0.5 * theta;
        std::vector<uint8_t> strong(nnz, 1);
        if (th > 0.0) {
            std::vector<double> dn(n, 0.0);
            auto fro = [&](size_t e) {
                double t = 0.0;
                for (int k = 0; k < bb; ++k) t += Av[e * bb + k] * Av[e * bb + k];
                return std::sqrt(t);
            };
            for (int i = 0; i < n; ++i)
                for (int e = Ap[i]; e < Ap[i + 1]; ++e)
                    if (Ac[e] == i) dn[i] = fro((size_t)e);
            for (int i = 0; i < n; ++i)
                for (int e = Ap[i]; e < Ap[i + 1]; ++e)
                    strong[e] = Ac[e] == i || fro((size_t)e) >= th * std::sqrt(dn[i] * dn[Ac[e]]);
        }
        std::vector<int> agg(n, -1);
        int na = 0;
        for (int i = 0; i < n; ++i) {
            if (excl[i] || agg[i] >= 0) continue;
            bool free_nbhd = true;
            for (int e = Ap[i]; e < Ap[i + 1] && free_nbhd; ++e) {
                const int j = Ac[e];
                if (strong[e] && !excl[j] && agg[j] >= 0) free_nbhd = false;
            }
            if (!free_nbhd) continue;
            for (int e = Ap[i]; e < Ap[i + 1]; ++e)
                if (strong[e] && !excl[Ac[e]]) agg[Ac[e]] = na;
            agg[i] = na++;
        }
        for (int i = 0; i < n; ++i) {
            if (excl[i] || agg[i] >= 0) continue;
            for (int e = Ap[i]; e < Ap[i + 1]; ++e) {
                const int j = Ac[e];
                if (strong[e] && !excl[j] && agg[j] >= 0) {
                    agg[i] = agg[j];
                    break;
                }
            }
            if (agg[i] < 0) agg[i] = na++;
        }
        if (na == 0) {
            *err = "amg: no free nodes";
            return 1;
        }
        std::vector<std::vector<int>> members(na);
        for (int i = 0; i < n; ++i)
            if (agg[i] >= 0) members[agg[i]].push_back(i);
        // near-null space: rigid-body modes about each aggregate's centroid (level 0), the QR
        // factors of the finer level below
        if (lev == 0) {
            B.assign((size_t)n * 3 * 6, 0.0);
            for (int a = 0; a < na; ++a) {
                double cen[3] = {0, 0, 0};
                for (int i : members[a])
                    for (int d = 0; d < 3; ++d) cen[d] += coords[3 * (size_t)i + d] / members[a].size();
                for (int i : members[a]) {
                    const double x = coords[3 * (size_t)i] - cen[0], y = coords[3 * (size_t)i + 1] - cen[1],
                                 z = coords[3 * (size_t)i + 2] - cen[2];
                    double *r0 = &B[(3 * (size_t)i) * 6], *r1 = r0 + 6, *r2 = r1 + 6;
                    r0[0] = 1.0;
                    r1[1] = 1.0;
                    r2[2] = 1.0;
                    r0[3] = -y;  // rotation about z
                    r1[3] = x;
                    r1[4] = -z;  // about x
                    r2[4] = y;
                    r0[5] = z;  // about y
                    r2[5] = -x;
                }
            }
            for (size_t d = 0; d < (size_t)n * 3; ++d)
                if (fixed[d])
                    for (int k = 0; k < 6; ++k) B[d * 6 + k] = 0.0;
        }
        // tentative prolongator: per-aggregate modified Gram-Schmidt (twice) of the null space rows
        Csr6 P0;
        P0.ptr.assign(n + 1, 0);
        std::vector<double> Q((size_t)n * b * 6, 0.0);  // P0 block of each fine row
        std::vector<double> Bc((size_t)na * 6 * 6, 0.0);
        parallel_chunks(na, T, [&](int alo, int ahi, int) {
          for (int a = alo; a < ahi; ++a) {
            const auto &mem = members[a];
            const int rows = (int)mem.size() * b;
            std::vector<double> M((size_t)rows * 6);
            for (size_t t = 0; t < mem.size(); ++t)
                for (int c = 0; c < b; ++c)
                    for (int k = 0; k < 6; ++k) M[(t * b + c) * 6 + k] = B[((size_t)mem[t] * b + c) * 6 + k];
            double colmax = 0.0;
            for (int k = 0; k < 6; ++k) {
                double s = 0.0;
                for (int r = 0; r < rows; ++r) s += M[(size_t)r * 6 + k] * M[(size_t)r * 6 + k];
                colmax = std::max(colmax, std::sqrt(s));
            }
            double R[6][6] = {};
            for (int k = 0; k < 6; ++k) {
                for (int pass = 0; pass < 2; ++pass)
                    for (int j = 0; j < k; ++j) {
                        double s = 0.0;
                        for (int r = 0; r < rows; ++r) s += M[(size_t)r * 6 + j] * M[(size_t)r * 6 + k];
                        R[j][k] += s;
                        for (int r = 0; r < rows; ++r) M[(size_t)r * 6 + k] -= s * M[(size_t)r * 6 + j];
                    }
                double nk = 0.0;
                for (int r = 0; r < rows; ++r) nk += M[(size_t)r * 6 + k] * M[(size_t)r * 6 + k];
                nk = std::sqrt(nk);
                if (nk > 1e-10 * colmax) {
                    R[k][k] = nk;
                    for (int r = 0; r < rows; ++r) M[(size_t)r * 6 + k] /= nk;
                } else {  // rank-deficient aggregate: a dummy coarse dof
                    for (int r = 0; r < rows; ++r) M[(size_t)r * 6 + k] = 0.0;
                    for (int j = 0; j < 6; ++j) R[k][j] = 0.0;
                }
            }
            for (size_t t = 0; t < mem.size(); ++t)
                for (int c = 0; c < b; ++c)
                    for (int k = 0; k < 6; ++k) Q[((size_t)mem[t] * b + c) * 6 + k] = M[(t * b + c) * 6 + k];
            for (int k = 0; k < 6; ++k)
                for (int j = 0; j < 6; ++j) Bc[((size_t)a * 6 + k) * 6 + j] = R[k][j];
          }
        });
        for (int i = 0; i < n; ++i) P0.ptr[i + 1] = P0.ptr[i] + (agg[i] >= 0 ? 1 : 0);
        P0.col.resize(P0.ptr[n]);
        P0.val.resize((size_t)P0.ptr[n] * b * 6);
        for (int i = 0; i < n; ++i)
            if (agg[i] >= 0) {
                P0.col[P0.ptr[i]] = agg[i];
                std::copy(&Q[(size_t)i * b * 6], &Q[(size_t)(i + 1) * b * 6], &P0.val[(size_t)P0.ptr[i] * b * 6]);
            }
        // spectrum bound of D^-1 A (power iteration, 1.1 x the estimate)
        L.lmax = level_lmax(n, b, Ap, Ac, Av, L.Dinv);
        // smoothed prolongator P = P0 - omega D^-1 A_F P0.  With a threshold, A_F is the filtered
        // operator: the strong blocks, the weak blocks of a row lumped onto its diagonal block
        // (row sums kept, so translations stay in the range of P); it keeps P's stencil, and
        // with it the coarse operator, from growing with the weak couplings
        Csr6 AP0, P;
        if (th > 0.0) {
            std::vector<int32_t> Fptr(n + 1, 0), Fcol;
            std::vector<double> Fval;
            Fcol.reserve(nnz);
            Fval.reserve(nnz * bb);
            for (int i = 0; i < n; ++i) {
                double lump[36] = {};
                size_t dpos = 0;
                for (int e = Ap[i]; e < Ap[i + 1]; ++e) {
                    const double *blk = &Av[(size_t)e * bb];
                    if (Ac[e] == i) dpos = Fcol.size();
                    if (strong[e]) {
                        Fcol.push_back(Ac[e]);
                        Fval.insert(Fval.end(), blk, blk + bb);
                    } else {
                        for (int k = 0; k < bb; ++k) lump[k] += blk[k];
                    }
                }
                for (int k = 0; k < bb; ++k) Fval[dpos * bb + k] += lump[k];
                Fptr[i + 1] = (int32_t)Fcol.size();
            }
            spgemm_A_M(n, b, Fptr.data(), Fcol.data(), Fval.data(), P0, na, &AP0);
        } else {
            spgemm_A_M(n, b, Ap, Ac, Av, P0, na, &AP0);
        }
        const double omega = 4.0 / (3.0 * L.lmax);
        // P0's own aggregate is always among AP0's columns when row i is free (A_ii P0_i); a
        // fixed row keeps P = 0 there (its A row is the identity and P0 row is zero)
        build_rows(n, b * 6, &P, [&](int i, AmgIdx &ocol, AmgVal &oval, int) {
            for (int f = AP0.ptr[i]; f < AP0.ptr[i + 1]; ++f) {
                const int c = AP0.col[f];
                double blk[36] = {};
                if (agg[i] == c) std::copy(&P0.val[(size_t)P0.ptr[i] * b * 6], &P0.val[(size_t)(P0.ptr[i] + 1) * b * 6], blk);
                gemm_b6(b, -omega, &L.Dinv[(size_t)i * bb], &AP0.val[(size_t)f * b * 6], blk);
                ocol.push_back(c);
                oval.insert(oval.end(), blk, blk + b * 6);
            }
        });
        // Galerkin coarse operator Ac = P^T (A P)
        Csr6 AP;
        spgemm_A_M(n, b, Ap, Ac, Av, P, na, &AP);
        // transpose of P: for coarse row a, the fine rows i with their b x 6 blocks
        AmgIdx Rptr(na + 1, 0), Rrow(P.col.size()), Rsrc(P.col.size());
        for (int32_t c : P.col) Rptr[c + 1]++;
        for (int a = 0; a < na; ++a) Rptr[a + 1] += Rptr[a];
        {
            std::vector<int32_t> fill(Rptr.begin(), Rptr.end() - 1);
            for (int i = 0; i < n; ++i)
                for (int f = P.ptr[i]; f < P.ptr[i + 1]; ++f) {
                    Rrow[fill[P.col[f]]] = i;
                    Rsrc[fill[P.col[f]]++] = f;
                }
        }
        Csr6 C;
        {
            std::vector<std::vector<int>> marks(T, std::vector<int>(na, -1)), lists(T), orders(T);
            std::vector<std::vector<double>> accs(T);
            build_rows(na, 36, &C, [&](int a, AmgIdx &ocol, AmgVal &oval, int t) {
                std::vector<int> &mark = marks[t], &list = lists[t], &order = orders[t];
                std::vector<double> &acc = accs[t];
                list.clear();
                acc.clear();
                for (int u = Rptr[a]; u < Rptr[a + 1]; ++u) {
                    const int i = Rrow[u];
                    const double *Pia = &P.val[(size_t)Rsrc[u] * b * 6];  // b x 6
                    for (int f = AP.ptr[i]; f < AP.ptr[i + 1]; ++f) {
                        const int c = AP.col[f];
                        if (mark[c] < 0) {
                            mark[c] = (int)list.size();
                            list.push_back(c);
                            acc.resize(acc.size() + 36, 0.0);
                        }
                        const double *APic = &AP.val[(size_t)f * b * 6];  // b x 6
                        double *o = &acc[(size_t)mark[c] * 36];
                        if (b == 3) gemm_t6<3>(Pia, APic, o);
                        else gemm_t6<6>(Pia, APic, o);
                    }
                }
                order.resize(list.size());
                for (size_t u = 0; u < order.size(); ++u) order[u] = (int)u;
                std::sort(order.begin(), order.end(), [&](int x, int y) { return list[x] < list[y]; });
                bool has_diag = false;
                for (int u : order) {
                    double *o = &acc[(size_t)u * 36];
                    if (list[u] == a) {
                        has_diag = true;
                        for (int r = 0; r < 6; ++r)  // dummy (rank-deficient) coarse dofs: identity
                            if (o[r * 6 + r] == 0.0) o[r * 6 + r] = 1.0;
                    }
                    ocol.push_back(list[u]);
                    oval.insert(oval.end(), o, o + 36);
                }
                if (!has_diag) {  // an aggregate disconnected from everything: identity block
                    // keep the columns sorted: insert the diagonal at its place
                    size_t pos = ocol.size() - order.size();
                    while (pos < ocol.size() && ocol[pos] < a) ++pos;
                    ocol.insert(ocol.begin() + pos, a);
                    double I6[36];
                    for (int r = 0; r < 36; ++r) I6[r] = r % 7 == 0 ? 1.0 : 0.0;
                    oval.insert(oval.begin() + pos * 36, I6, I6 + 36);
                }
                for (int c : list) mark[c] = -1;
            });
        }
        AmgIdx &Cptr = C.ptr, &Ccol = C.col;
        AmgVal &Cval = C.val;
        // store this level's transfer operators
        L.nc = na;
        L.Rval.resize((size_t)P.col.size() * 6 * b);
        parallel_chunks((int)Rrow.size(), T, [&](int lo, int hi, int) {
            for (int t = lo; t < hi; ++t) {
                const double *Pia = &P.val[(size_t)Rsrc[t] * b * 6];
                for (int r = 0; r < 6; ++r)
                    for (int k = 0; k < b; ++k) L.Rval[(size_t)t * 6 * b + r * b + k] = Pia[k * 6 + r];
            }
        });
        L.Pptr = std::move(P.ptr);
        L.Pcol = std::move(P.col);
        L.Pval = std::move(P.val);
        L.Rptr = std::move(Rptr);
        L.Rcol = std::move(Rrow);
        if (lev > 0) {  // the next level's operator replaces these below
            L.Aptr = std::move(Aptr_v);
            L.Acol = std::move(Acol_v);
            L.Aval = std::move(Aval_v);
        }
        h->levels.push_back(std::move(L));
        // next level
        n = na;
        b = 6;
        Aptr_v.swap(Cptr);
        Acol_v.swap(Ccol);
        Aval_v.swap(Cval);
        Ap = Aptr_v.data();
        Ac = Acol_v.data();
        Av = Aval_v.data();
        nnz = Acol_v.size();
        B.swap(Bc);
        fixed.assign((size_t)n * 6, 0);
        if (6 * n <= coarse_max || (int)h->levels.size() >= max_levels) break;
    }
    // coarsest: dense inverse (Gauss-Jordan with partial pivoting)
    const int m = 6 * n;
    if (m > 4096) {
        *err = "amg: coarsest level too large";
        return 1;
    }
    std::vector<double> D((size_t)m * m, 0.0), I((size_t)m * m, 0.0);
    for (int i = 0; i < n; ++i)
        for (int e = Ap[i]; e < Ap[i + 1]; ++e)
            for (int r = 0; r < 6; ++r)
                for (int c = 0; c < 6; ++c) D[(size_t)(6 * i + r) * m + 6 * Ac[e] + c] = Av[(size_t)e * 36 + r * 6 + c];
    for (int i = 0; i < m; ++i) I[(size_t)i * m + i] = 1.0;
    for (int c = 0; c < m; ++c) {
        int p = c;
        for (int r = c + 1; r < m; ++r)
            if (std::fabs(D[(size_t)r * m + c]) > std::fabs(D[(size_t)p * m + c])) p = r;
        if (D[(size_t)p * m + c] == 0.0) {
            *err = "amg: singular coarsest operator";
            return 1;
        }
        if (p != c)
            for (int j = 0; j < m; ++j) {
                std::swap(D[(size_t)p * m + j], D[(size_t)c * m + j]);
                std::swap(I[(size_t)p * m + j], I[(size_t)c * m + j]);
            }
        const double inv = 1.0 / D[(size_t)c * m + c];
        for (int j = 0; j < m; ++j) {
            D[(size_t)c * m + j] *= inv;
            I[(size_t)c * m + j] *= inv;
        }
        for (int r = 0; r < m; ++r)
            if (r != c) {
                const double f = D[(size_t)r * m + c];
                if (f == 0.0) continue;
                for (int j = 0; j < m; ++j) {
                    D[(size_t)r * m + j] -= f * D[(size_t)c * m + j];
                    I[(size_t)r * m + j] -= f * I[(size_t)c * m + j];
                }
            }
    }
    h->coarse_n = m;
    h->coarse_inv.swap(I);
    return 0;
}

template <typename V>
int amg_setup_p2(int n, const int32_t *ptr, const int32_t *col, const V *val, const uint8_t *fixed, int nv,
                 const int32_t *par_ptr, const int32_t *par, const double *parw, const double *vxyz, int coarse_max,
                 int max_levels, AmgHierarchy *h, std::string *err, double theta, double lmax0) {
    h->levels.clear();
    h->coarse_inv.clear();
    if (max_levels < 2) {
        *err = "amg: a P2 hierarchy needs at least two levels";
        return 1;
    }
    AmgLevel L0;
    L0.n = n;
    L0.b = 3;
    L0.nc = nv;
    L0.bc = 3;
    if (!level_dinv(n, 3, ptr, col, val, &L0.Dinv)) {
        *err = "amg: missing diagonal block";
        return 1;
    }
    L0.lmax = lmax0 > 0.0 ? lmax0 : level_lmax(n, 3, ptr, col, val, L0.Dinv);
    // fixed vertex dofs: those of the vertex's own node (the node with that single parent)
    std::vector<uint8_t> fixed1(3 * (size_t)nv, 0);
    for (int i = 0; i < n; ++i) {
        if (par_ptr[i + 1] - par_ptr[i] != 1) continue;
        const int I = par[par_ptr[i]];
        if (I < 0 || I >= nv) {
            *err = "amg: P2 parent out of range";
            return 1;
        }
        for (int c = 0; c < 3; ++c)
            if (fixed[3 * (size_t)i + c]) fixed1[3 * (size_t)I + c] = 1;
    }
    // prolongator: diagonal 3 x 3 blocks, zero where the fine or the coarse dof is fixed
    const int np = par_ptr[n];
    L0.Pptr.assign(par_ptr, par_ptr + n + 1);
    L0.Pcol.resize(np);
    L0.Pval.assign((size_t)np * 9, 0.0);
    for (int i = 0; i < n; ++i)
        for (int k = par_ptr[i]; k < par_ptr[i + 1]; ++k) {
            const int I = par[k];
            if (I < 0 || I >= nv) {
                *err = "amg: P2 parent out of range";
                return 1;
            }
            L0.Pcol[k] = I;
            for (int c = 0; c < 3; ++c)
                L0.Pval[(size_t)k * 9 + 4 * c] = (fixed[3 * (size_t)i + c] || fixed1[3 * (size_t)I + c]) ? 0.0 : parw[k];
        }
    // restriction P^T: vertex I's children (fine rows) with the same diagonal blocks
    L0.Rptr.assign(nv + 1, 0);
    for (int k = 0; k < np; ++k) L0.Rptr[L0.Pcol[k] + 1]++;
    for (int I = 0; I < nv; ++I) L0.Rptr[I + 1] += L0.Rptr[I];
    L0.Rcol.resize(np);
    L0.Rval.resize((size_t)np * 9);
    AmgIdx rsrc(np);
    {
        std::vector<int32_t> fill(L0.Rptr.begin(), L0.Rptr.end() - 1);
        for (int i = 0; i < n; ++i)
            for (int k = par_ptr[i]; k < par_ptr[i + 1]; ++k) {
                const int at = fill[L0.Pcol[k]]++;
                L0.Rcol[at] = i;
                rsrc[at] = k;
                std::copy(&L0.Pval[(size_t)k * 9], &L0.Pval[(size_t)(k + 1) * 9], &L0.Rval[(size_t)at * 9]);
            }
    }
    // Galerkin product A1 = P^T A P: coarse row I over its children i, their rows and the parents
    // of each column
    Csr6 A1;
    {
        const int T = num_threads();
        std::vector<std::vector<int>> marks(T, std::vector<int>(nv, -1)), lists(T), orders(T);
        std::vector<std::vector<double>> accs(T);
        build_rows(nv, 9, &A1, [&](int I, AmgIdx &ocol, AmgVal &oval, int t) {
            std::vector<int> &mark = marks[t], &list = lists[t], &order = orders[t];
            std::vector<double> &acc = accs[t];
            list.clear();
            acc.clear();
            for (int u = L0.Rptr[I]; u < L0.Rptr[I + 1]; ++u) {
                const int i = L0.Rcol[u];
                const double *pi = &L0.Pval[(size_t)rsrc[u] * 9];
                for (int e = ptr[i]; e < ptr[i + 1]; ++e) {
                    const int j = col[e];
                    const V *A = val + (size_t)e * 9;
                    for (int k = par_ptr[j]; k < par_ptr[j + 1]; ++k) {
                        const int J = L0.Pcol[k];
                        const double *pj = &L0.Pval[(size_t)k * 9];
                        if (mark[J] < 0) {
                            mark[J] = (int)list.size();
                            list.push_back(J);
                            acc.resize(acc.size() + 9, 0.0);
                        }
                        double *o = &acc[(size_t)mark[J] * 9];
                        for (int c = 0; c < 3; ++c)
                            for (int d = 0; d < 3; ++d) o[3 * c + d] += pi[4 * c] * A[3 * c + d] * pj[4 * d];
                    }
                }
            }
            order.resize(list.size());
            for (size_t u = 0; u < order.size(); ++u) order[u] = (int)u;
            std::sort(order.begin(), order.end(), [&](int x, int y) { return list[x] < list[y]; });
            for (int u : order) {
                const int J = list[u];
                double *o = &acc[(size_t)u * 9];
                for (int c = 0; c < 3; ++c)
                    for (int d = 0; d < 3; ++d)
                        if (fixed1[3 * (size_t)I + c] || fixed1[3 * (size_t)J + d])
                            o[3 * c + d] = (I == J && c == d && fixed1[3 * (size_t)I + c]) ? 1.0 : 0.0;
                ocol.push_back(J);
                oval.insert(oval.end(), o, o + 9);
            }
            for (int J : list) mark[J] = -1;
        });
    }
    AmgHierarchy h1;
    if (amg_setup(nv, A1.ptr.data(), A1.col.data(), A1.val.data(), vxyz, fixed1.data(), coarse_max, max_levels - 1,
                  &h1, err, theta))
        return 1;
    // amg_setup leaves its finest operator with the caller: here it is level 1's
    h1.levels[0].Aptr = std::move(A1.ptr);
    h1.levels[0].Acol = std::move(A1.col);
    h1.levels[0].Aval = std::move(A1.val);
    h->levels.push_back(std::move(L0));
    for (auto &L : h1.levels) h->levels.push_back(std::move(L));
    h->coarse_n = h1.coarse_n;
    h->coarse_inv.swap(h1.coarse_inv);
    return 0;
}

template int amg_setup_p2<double>(int, const int32_t *, const int32_t *, const double *, const uint8_t *, int,
                                  const int32_t *, const int32_t *, const double *, const double *, int, int,
                                  AmgHierarchy *, std::string *, double, double);
template int amg_setup_p2<float>(int, const int32_t *, const int32_t *, const float *, const uint8_t *, int,
                                 const int32_t *, const int32_t *, const double *, const double *, int, int,
                                 AmgHierarchy *, std::string *, double, double);

}  // namespace plx
