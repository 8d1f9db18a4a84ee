// Host-only test shim over the multigrid set-up (pylith_amd/csrc/plx_amg.cpp), built with g++ into
// tests/cpp/libamg_probe.so and driven by tests/test_amg_host.py: builds a hierarchy, exports its
// arrays, and applies the V-cycle the device runs (damped block Jacobi s = 2 / (lmax (1 + 1/30))
// before and after the coarse correction, dense solve on the coarsest level) in FP64.
#include <cstring>
#include <string>
#include <vector>

#include "plx_amg.hpp"

using plx::AmgHierarchy;
using plx::AmgLevel;

namespace {
AmgHierarchy H;
std::vector<int32_t> A0p, A0c;
std::vector<double> A0v;
std::string last_error;
double theta = 0.0;  // strength-of-connection threshold of the next build

void spmm(int nr, int br, int bc, const int32_t *p, const int32_t *c, const double *v, const double *x,
          double *y, bool acc) {
    for (int i = 0; i < nr; ++i) {
        double s[6] = {0, 0, 0, 0, 0, 0};
        for (int e = p[i]; e < p[i + 1]; ++e)
            for (int a = 0; a < br; ++a)
                for (int k = 0; k < bc; ++k) s[a] += v[(size_t)e * br * bc + a * bc + k] * x[(size_t)c[e] * bc + k];
        for (int a = 0; a < br; ++a) y[(size_t)i * br + a] = (acc ? y[(size_t)i * br + a] : 0.0) + s[a];
    }
}

void level_A(int l, const double *x, double *y) {
    const AmgLevel &L = H.levels[l];
    if (l == 0) spmm(L.n, 3, 3, A0p.data(), A0c.data(), A0v.data(), x, y, false);
    else spmm(L.n, L.b, L.b, L.Aptr.data(), L.Acol.data(), L.Aval.data(), x, y, false);
}

void dinv(int l, double s, const double *r, double *z) {
    const AmgLevel &L = H.levels[l];
    const int b = L.b;
    for (int i = 0; i < L.n; ++i)
        for (int a = 0; a < b; ++a) {
            double t = 0.0;
            for (int c = 0; c < b; ++c) t += L.Dinv[(size_t)i * b * b + a * b + c] * r[(size_t)i * b + c];
            z[(size_t)i * b + a] = s * t;
        }
}

void vcycle(int l, const double *bv, double *x) {
    const int nl = (int)H.levels.size();
    if (l == nl) {
        const int m = H.coarse_n;
        for (int i = 0; i < m; ++i) {
            double t = 0.0;
            for (int j = 0; j < m; ++j) t += H.coarse_inv[(size_t)i * m + j] * bv[j];
            x[i] = t;
        }
        return;
    }
    const AmgLevel &L = H.levels[l];
    const size_t nd = (size_t)L.n * L.b;
    const double s = 2.0 / (L.lmax * (1.0 + 1.0 / 30.0));
    std::vector<double> xl(nd), r(nd), t(nd), z(nd), cb((size_t)L.nc * L.bc), cx((size_t)L.nc * L.bc);
    dinv(l, s, bv, xl.data());
    level_A(l, xl.data(), t.data());
    for (size_t i = 0; i < nd; ++i) r[i] = bv[i] - t[i];
    spmm(L.nc, L.bc, L.b, L.Rptr.data(), L.Rcol.data(), L.Rval.data(), r.data(), cb.data(), false);
    vcycle(l + 1, cb.data(), cx.data());
    spmm(L.n, L.b, L.bc, L.Pptr.data(), L.Pcol.data(), L.Pval.data(), cx.data(), xl.data(), true);
    level_A(l, xl.data(), t.data());
    for (size_t i = 0; i < nd; ++i) r[i] = bv[i] - t[i];
    dinv(l, s, r.data(), z.data());
    for (size_t i = 0; i < nd; ++i) x[i] = xl[i] + z[i];
}

template <typename V>
int copy_out(const V &v, void *dst) {
    if (dst) std::memcpy(dst, v.data(), v.size() * sizeof(v[0]));
    return (int)v.size();
}
}  // namespace

extern "C" {

int amg_probe_build(int n, const int32_t *ptr, const int32_t *col, const double *val, const double *coords,
                    const uint8_t *fixed, int coarse_max, int max_levels) {
    A0p.assign(ptr, ptr + n + 1);
    A0c.assign(col, col + ptr[n]);
    A0v.assign(val, val + (size_t)ptr[n] * 9);
    if (plx::amg_setup(n, ptr, col, val, coords, fixed, coarse_max, max_levels, &H, &last_error, theta)) return -1;
    return (int)H.levels.size();
}

void amg_probe_set_theta(double t) { theta = t; }

// The P2 form (a P2 -> P1 first level, then the aggregation hierarchy of the P1 operator).
int amg_probe_build_p2(int n, const int32_t *ptr, const int32_t *col, const double *val, const uint8_t *fixed,
                       int nv, const int32_t *par_ptr, const int32_t *par, const double *parw, const double *vxyz,
                       int coarse_max, int max_levels) {
    A0p.assign(ptr, ptr + n + 1);
    A0c.assign(col, col + ptr[n]);
    A0v.assign(val, val + (size_t)ptr[n] * 9);
    if (plx::amg_setup_p2(n, ptr, col, val, fixed, nv, par_ptr, par, parw, vxyz, coarse_max, max_levels, &H,
                          &last_error, theta))
        return -1;
    return (int)H.levels.size();
}

// The same from FP32 operator values (the set-up's device path reads the operator's FP32 copy).
int amg_probe_build_p2f(int n, const int32_t *ptr, const int32_t *col, const float *val, const uint8_t *fixed,
                        int nv, const int32_t *par_ptr, const int32_t *par, const double *parw, const double *vxyz,
                        int coarse_max, int max_levels) {
    A0p.assign(ptr, ptr + n + 1);
    A0c.assign(col, col + ptr[n]);
    A0v.assign(val, val + (size_t)ptr[n] * 9);
    if (plx::amg_setup_p2(n, ptr, col, val, fixed, nv, par_ptr, par, parw, vxyz, coarse_max, max_levels, &H,
                          &last_error, theta))
        return -1;
    return (int)H.levels.size();
}

const char *amg_probe_error() { return last_error.c_str(); }

// info: n, b, nc; lmax
int amg_probe_level(int l, int *info, double *lmax) {
    const AmgLevel &L = H.levels[l];
    info[0] = L.n;
    info[1] = L.b;
    info[2] = L.nc;
    *lmax = L.lmax;
    return 0;
}

// which: 0 Aptr 1 Acol 2 Aval 3 Pptr 4 Pcol 5 Pval 6 Rptr 7 Rcol 8 Rval 9 Dinv; returns the count
int amg_probe_array(int l, int which, void *dst) {
    const AmgLevel &L = H.levels[l];
    switch (which) {
    case 0: return copy_out(L.Aptr, dst);
    case 1: return copy_out(L.Acol, dst);
    case 2: return copy_out(L.Aval, dst);
    case 3: return copy_out(L.Pptr, dst);
    case 4: return copy_out(L.Pcol, dst);
    case 5: return copy_out(L.Pval, dst);
    case 6: return copy_out(L.Rptr, dst);
    case 7: return copy_out(L.Rcol, dst);
    case 8: return copy_out(L.Rval, dst);
    case 9: return copy_out(L.Dinv, dst);
    default: return -1;
    }
}

int amg_probe_coarse(double *dst) { return copy_out(H.coarse_inv, dst); }
int amg_probe_coarse_n() { return H.coarse_n; }
int amg_probe_level_bc(int l) { return H.levels[l].bc; }

void amg_probe_vcycle(const double *r, double *z) { vcycle(0, r, z); }

}  // extern "C"
