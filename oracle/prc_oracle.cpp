// prc_oracle.cpp -- TEST INFRASTRUCTURE ONLY (the oracle).
//
// CPU restatement of the reference positron-range-correction algorithm (prc.cu of
// jlherraiz/PRC_MLEM), used as the checker of the HIP path. Only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it; the product
// (libprc_amd.so and the prc executable) never links, loads or calls it.
//
// Two precisions, one algorithm:
//   * FP32 "reference semantics": the arithmetic of prc.cu, with nvcc's default
//     contraction of a*b+-c into fmaf written out explicitly where prc.cu has such an
//     expression on the ray/DDA path (ray() 1-z*z, trace_ray's depth update, the
//     adjoint's accumulation). The survival table is built from host libm exactly as
//     prc.cu:625-633 builds it on the host.
//   * FP64 accuracy oracle: the same algorithm in double with the survival function
//     evaluated analytically (no table) -- the reference for NRMSE / count error.
//
// Pinning (what checks this file; tests/test_oracle.py):
//   isotope model + table: 1-S(cutoff) = 0.9995, table error <= 2e-5 vs FP64 for all
//     8 isotopes (tests/test_rdp.py:18-24 of the reference);
//   ray set: unit directions, mirrored pairs, first moments of the water kernel;
//   DDA: brute-force voxel walk by exact plane intersection on tiny grids, and the
//     deposit sum = 1 - S(exit) telescoping identity;
//   projector: A^T 1 = 1, sum(Ax) = sum(x), <Ax,y> = <x,A^T y>, Ax = 0 off support,
//     water kernel vs an independent shell x direction quadrature (< 1% of the blur);
//   RDP update: finite differences, scalar minimisation, the 0.5 / 0 known answers
//     (tests/test_rdp.py:27-90 of the reference);
//   solver: count conservation, zero-iteration identity, warm-up equality, unit scaling.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include <omp.h>

namespace {

// prc.cu:38-57 -- Paneque-Yunta et al., Phys. Med. Biol. 70 (2025) 165013, Tables 4-5.
struct Br { float w, a, b, c, f; };
struct Iso { const char* name; int n; Br br[3]; };
const Iso kIso[8] = {
    {"11C", 1, {{1.000f, 0.777f, 0.468f, 0.217f, 0.010f}}},
    {"13N", 1, {{1.000f, 0.537f, 0.347f, 0.193f, 0.050f}}},
    {"15O", 1, {{1.000f, 0.295f, 0.221f, 0.133f, 0.118f}}},
    {"18F", 1, {{1.000f, 1.612f, 0.703f, 0.526f, 0.011f}}},
    {"64Cu", 1, {{1.000f, 1.494f, 0.000f, 0.631f, 0.251f}}},
    {"68Ga", 2, {{0.014f, 1.048f, 0.000f, 0.475f, 0.271f}, {0.986f, 0.289f, 0.137f, 0.159f, 0.322f}}},
    {"82Rb", 2, {{0.136f, 0.191f, 0.084f, 0.114f, 0.408f}, {0.857f, 0.135f, 0.063f, 0.087f, 0.451f}}},
    {"124I", 3, {{0.013f, 1.007f, 0.000f, 0.421f, 0.443f}, {0.515f, 0.412f, 0.000f, 0.210f, 0.458f},
                 {0.471f, 0.241f, 0.000f, 0.150f, 0.369f}}},
};
constexpr int kBins = 4096;                 // prc.cu:61
constexpr float kMinRho = 1.0e-3f;          // prc.cu:102
constexpr int kNR = 2, kNS = 5, kNQ = 125;  // prc.cu:103
constexpr int kNearRays = 2048;             // prc.cu:104

// prc.cu:577-584: the smallest weight absorbs 1 - sum(w) in FP32 (first one on ties).
Iso renormalised(int idx) {
    Iso m = kIso[idx];
    float wsum = 0.0f;
    int smallest = 0;
    for (int b = 0; b < m.n; ++b) {
        wsum += m.br[b].w;
        if (m.br[b].w < m.br[smallest].w) smallest = b;
    }
    m.br[smallest].w += 1.0f - wsum;
    return m;
}

// prc.cu:65-72, host flavour (no contraction), float.
float survival_f(const Iso& m, float r) {
    float s = 0.0f;
    for (int i = 0; i < m.n; ++i) {
        const float x = m.br[i].b * r, y = m.br[i].c * r;
        const float p = powf(m.br[i].a * r, 1.0f + m.br[i].f);
        const float q = x * x * sqrtf(x);
        const float u = y * y * y * sqrtf(y);
        const float e = expf(-((p + q) + u));
        const float t = m.br[i].w * e;
        s = s + t;
    }
    return s;
}

// The same survival in double with the FP32 coefficients promoted (tests/reference.py:20-26).
double survival_d(const Iso& m, double r) {
    double s = 0.0;
    for (int i = 0; i < m.n; ++i) {
        const double w = m.br[i].w, a = m.br[i].a, b = m.br[i].b, c = m.br[i].c, f = m.br[i].f;
        const double x = b * r, y = c * r;
        s += w * std::exp(-(std::pow(a * r, 1.0 + f) + x * x * std::sqrt(x) + y * y * y * std::sqrt(y)));
    }
    return s;
}

// prc.cu:83-91 (FP32, fraction 0.9995 at prc.cu:587).
float range_cutoff_f(const Iso& m, float fraction) {
    float lo = 0.0f, hi = 1.0f;
    while (1.0f - survival_f(m, hi) < fraction && hi < 256.0f) hi *= 2.0f;
    for (int it = 0; it < 64; ++it) {
        const float mid = 0.5f * (lo + hi);
        if (1.0f - survival_f(m, mid) < fraction) lo = mid;
        else hi = mid;
    }
    return hi;
}

template <class T>
struct Grid {
    int nx, ny, nz;
    T sx, sy, sz, cutoff, support_rho, air_cap;
};

// Survival lookup: FP32 table lerp (prc.cu:74-80) or analytic S.
struct Range {
    const Iso* model;
    const float* table;  // kBins + 1 entries, or nullptr = analytic
    float scale;
    float at(float w) const {
        if (!table) return survival_f(*model, w);
        const float u = w * scale;
        const int i = std::min((int)u, kBins - 1);
        const float lo = table[i], hi = table[i + 1];
        return fmaf(fminf(u - (float)i, 1.0f), hi - lo, lo);
    }
    double at(double w) const { return survival_d(*model, w); }
};

template <class T> struct V3 { T x, y, z; };

// prc.cu:118-129. sincospif is evaluated as the rounded double sin/cos of pi*arg.
void ray_f(int k, int n, V3<float>* start, V3<float>* dir) {
    const unsigned m = (unsigned)k >> 1;
    const float sign = (k & 1) ? -1.0f : 1.0f;
    const float z = 1.0f - ((float)m + 0.5f) / (float)((n + 1) >> 1);
    const float r = sqrtf(fmaxf(0.0f, fmaf(-z, z, 1.0f)));
    const float to_unit = 2.3283064e-10f;
    const float arg = 2.0f * to_unit * (float)(m * 2654435769u);
    const double pa = 3.14159265358979323846 * (double)arg;
    const float sn = (float)std::sin(pa), cs = (float)std::cos(pa);
    *dir = {sign * r * cs, sign * r * sn, sign * z};
    *start = {sign * fmaf(to_unit, (float)(0x80000000u + m * 3518319155u), -0.5f),
              sign * fmaf(to_unit, (float)(0x80000000u + m * 2882110345u), -0.5f),
              sign * fmaf(to_unit, (float)(0x80000000u + m * 2360945575u), -0.5f)};
}

void ray_d(int k, int n, V3<double>* start, V3<double>* dir) {
    const unsigned m = (unsigned)k >> 1;
    const double sign = (k & 1) ? -1.0 : 1.0;
    const double z = 1.0 - ((double)m + 0.5) / (double)((n + 1) >> 1);
    const double r = std::sqrt(std::max(0.0, 1.0 - z * z));
    const double to_unit = 1.0 / 4294967296.0;
    const double pa = 3.14159265358979323846 * 2.0 * to_unit * (double)(m * 2654435769u);
    *dir = {sign * r * std::cos(pa), sign * r * std::sin(pa), sign * z};
    *start = {sign * (to_unit * (double)(0x80000000u + m * 3518319155u) - 0.5),
              sign * (to_unit * (double)(0x80000000u + m * 2882110345u) - 0.5),
              sign * (to_unit * (double)(0x80000000u + m * 2360945575u) - 0.5)};
}

inline float fma_t(float a, float b, float c) { return fmaf(a, b, c); }
inline double fma_t(double a, double b, double c) { return a * b + c; }

inline bool inside(int nx, int ny, int nz, int i, int j, int k) {
    return i >= 0 && i < nx && j >= 0 && j < ny && k >= 0 && k < nz;
}

// prc.cu:134-165. Returns the number of voxel visits (VRS); calls visit(i,j,k,p) per deposit.
template <class T, class Visit>
long trace(const Grid<T>& g, const T* rho, const Range& rg, int vx, int vy, int vz, V3<T> x, V3<T> d,
           bool near, Visit&& visit) {
    const T big = (T)1.0e30;
    const int di = d.x > 0 ? 1 : -1, dj = d.y > 0 ? 1 : -1, dk = d.z > 0 ? 1 : -1;
    T tx = d.x != 0 ? fma_t((T)0.5, (T)di, -x.x) * g.sx / d.x : big, dtx = d.x != 0 ? g.sx / std::fabs(d.x) : big;
    T ty = d.y != 0 ? fma_t((T)0.5, (T)dj, -x.y) * g.sy / d.y : big, dty = d.y != 0 ? g.sy / std::fabs(d.y) : big;
    T tz = d.z != 0 ? fma_t((T)0.5, (T)dk, -x.z) * g.sz / d.z : big, dtz = d.z != 0 ? g.sz / std::fabs(d.z) : big;
    int i = vx, j = vy, k = vz;
    T t = 0, w = 0, surv = 1, air = 0;
    long steps = 0;
    for (;;) {
        const bool in_block = std::abs(i - vx) <= kNR && std::abs(j - vy) <= kNR && std::abs(k - vz) <= kNR;
        if (near && !in_block) return steps;
        ++steps;
        const T t_next = std::min(tx, std::min(ty, tz));
        const T r = rho[((size_t)k * g.ny + j) * g.nx + i];
        const T w_next = std::min(fma_t(r, t_next - t, w), g.cutoff);
        const T surv_next = rg.at(w_next);
        if (r >= g.support_rho) {
            air = 0;
            if (in_block == near && surv > surv_next) visit(i, j, k, surv - surv_next);
        } else if ((air += t_next - t) > g.air_cap) {
            return steps;
        }
        if (w_next >= g.cutoff) return steps;
        t = t_next, w = w_next, surv = surv_next;
        if (t_next == tx) i += di, tx += dtx;
        else if (t_next == ty) j += dj, ty += dty;
        else k += dk, tz += dtz;
        if (!inside(g.nx, g.ny, g.nz, i, j, k)) return steps;
    }
}

template <class T>
void ray_t(int k, int n, V3<T>* s, V3<T>* d);
template <> void ray_t<float>(int k, int n, V3<float>* s, V3<float>* d) { ray_f(k, n, s, d); }
template <> void ray_t<double>(int k, int n, V3<double>* s, V3<double>* d) { ray_d(k, n, s, d); }

// prc.cu:178-267 with the near cache on: W and the 125 near weights are built once.
template <class T>
struct Projector {
    Grid<T> g;
    std::vector<T> rho;
    std::vector<int> active;
    int n_rays = 0;
    Range rg{};
    std::vector<V3<T>> fs, fd;  // far ray set
    std::vector<T> W, nearw;    // n_a, n_a * 125
    std::atomic<long> vrs_near{0}, vrs_far{0};

    V3<int> coord(int s) const { return {s % g.nx, (s / g.nx) % g.ny, s / (g.nx * g.ny)}; }

    void build() {
        fs.resize(n_rays), fd.resize(n_rays);
        for (int r = 0; r < n_rays; ++r) ray_t<T>(r, n_rays, &fs[r], &fd[r]);
        std::vector<V3<T>> ns(kNearRays), nd(kNearRays);
        for (int r = 0; r < kNearRays; ++r) ray_t<T>(r, kNearRays, &ns[r], &nd[r]);
        const int na = (int)active.size();
        W.assign(na, 0), nearw.assign((size_t)na * kNQ, 0);
        long vn = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : vn)
        for (int a = 0; a < na; ++a) {
            const V3<int> v = coord(active[a]);
            T sum[kNQ];
            for (int q = 0; q < kNQ; ++q) sum[q] = 0;
            for (int r = 0; r < kNearRays; ++r)
                vn += trace(g, rho.data(), rg, v.x, v.y, v.z, ns[r], nd[r], true, [&](int i, int j, int k, T p) {
                    sum[((k - v.z + kNR) * kNS + j - v.y + kNR) * kNS + i - v.x + kNR] += p;
                });
            const T near_scale = (T)n_rays / (T)kNearRays;
            T acc = 0;
            for (int q = 0; q < kNQ; ++q) {
                const int i = v.x + q % kNS - kNR, j = v.y + q / kNS % kNS - kNR, k = v.z + q / (kNS * kNS) - kNR;
                if (!inside(g.nx, g.ny, g.nz, i, j, k)) continue;
                const T w = sum[q] * near_scale;
                nearw[(size_t)a * kNQ + q] = w;
                if (w > 0) acc += w;
            }
            for (int r = 0; r < n_rays; ++r)
                trace(g, rho.data(), rg, v.x, v.y, v.z, fs[r], fd[r], false, [&](int, int, int, T p) { acc += p; });
            W[a] = acc;
        }
        vrs_near = vn;
    }

    // Visit every far deposit of source a (ray order) and the near weights first.
    template <class F>
    long for_each(int a, F&& f) const {
        const V3<int> v = coord(active[a]);
        for (int q = 0; q < kNQ; ++q) {
            const int i = v.x + q % kNS - kNR, j = v.y + q / kNS % kNS - kNR, k = v.z + q / (kNS * kNS) - kNR;
            if (!inside(g.nx, g.ny, g.nz, i, j, k)) continue;
            const T w = nearw[(size_t)a * kNQ + q];
            if (w > 0) f(((size_t)k * g.ny + j) * g.nx + i, w);
        }
        long steps = 0;
        for (int r = 0; r < n_rays; ++r)
            steps += trace(g, rho.data(), rg, v.x, v.y, v.z, fs[r], fd[r], false,
                           [&](int i, int j, int k, T p) { f(((size_t)k * g.ny + j) * g.nx + i, p); });
        return steps;
    }

    // prc.cu:239-254, 430-434: out = A f (overwrites out).
    void forward(const T* f, T* out) {
        const size_t n = (size_t)g.nx * g.ny * g.nz;
        const int nt = omp_get_max_threads();
        std::vector<T> part((size_t)nt * n, 0);
        long vf = 0;
#pragma omp parallel reduction(+ : vf)
        {
            T* mine = part.data() + (size_t)omp_get_thread_num() * n;
#pragma omp for schedule(dynamic, 64)
            for (int a = 0; a < (int)active.size(); ++a) {
                const T fs_ = f[active[a]], ws = W[a];
                const T scale = ws > 0 ? fs_ / ws : 0;
                if (!(scale > 0)) continue;
                vf += for_each(a, [&](size_t t, T w) { mine[t] += scale * w; });
            }
        }
        vrs_far += vf;
#pragma omp parallel for schedule(static)
        for (long long t = 0; t < (long long)n; ++t) {
            T s = 0;
            for (int p = 0; p < nt; ++p) s += part[(size_t)p * n + t];
            out[t] = s;
        }
    }

    // prc.cu:256-267, 436: out_s = sum_t w_st in_t / W_s on active s (others untouched).
    void adjoint(const T* in, T* out) {
        long vf = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : vf)
        for (int a = 0; a < (int)active.size(); ++a) {
            const int s = active[a];
            if (!(W[a] > 0)) { out[s] = 0; continue; }
            T acc = 0;
            vf += for_each(a, [&](size_t t, T w) { acc = fma_t(w, in[t], acc); });
            out[s] = acc / W[a];
        }
        vrs_far += vf;
    }
};

// prc.cu:279-309 -- RDP MAP-EM surrogate update for one voxel.
template <class T>
T rdp_update(const Grid<T>& g, const T* f, const T* back, const unsigned char* sup, size_t idx, T beta, T gamma,
             T eps) {
    if (!sup[idx]) return 0;
    const T fj = f[idx], f_em = fj * back[idx];
    if (beta == 0) return std::max(f_em, (T)0);
    const int i = (int)(idx % g.nx), j = (int)((idx / g.nx) % g.ny), k = (int)(idx / ((size_t)g.nx * g.ny));
    T grad = 0, curv = 0;
    const T smin = std::min(g.sx, std::min(g.sy, g.sz));
    for (int dz = -1; dz <= 1; ++dz)
        for (int dy = -1; dy <= 1; ++dy)
            for (int dx = -1; dx <= 1; ++dx) {
                if (!(dx | dy | dz) || !inside(g.nx, g.ny, g.nz, i + dx, j + dy, k + dz)) continue;
                const size_t nb = idx + ((long)dz * g.ny + dy) * g.nx + dx;
                if (!sup[nb]) continue;
                const T mx = dx * g.sx, my = dy * g.sy, mz = dz * g.sz;
                const T w = smin / std::sqrt(mx * mx + my * my + mz * mz);
                const T fk = f[nb], d = fj - fk, ad = std::fabs(d);
                const T D = fj + fk + gamma * ad + eps, invD = (T)1 / D;
                grad += w * d * (fj + (T)3 * fk + gamma * ad + (T)2 * eps) * invD * invD;
                const T half_quad = ((T)2 * (fj + fk + eps) + gamma * ad) * invD * invD;
                const T hess = (T)2 * ((T)2 * fk + eps) * ((T)2 * fk + eps) * invD * invD * invD;
                curv += (T)2 * w * std::max(half_quad, hess);
            }
    const T a = beta * curv, B = (T)1 + beta * (grad - curv * fj);
    const T root = std::sqrt(B * B + (T)4 * a * f_em);
    return B >= 0 ? (f_em > 0 ? (T)2 * f_em / (B + root) : (T)0) : (root - B) / ((T)2 * a);
}

template <class T>
double dsum(const T* a, size_t n) {  // prc.cu:392-395 (double accumulator, index order)
    double s = 0;
    for (size_t i = 0; i < n; ++i) s += (double)a[i];
    return s;
}
template <class T>
double ddot(const T* a, const T* b, size_t n) {  // prc.cu:397-400
    double s = 0;
    for (size_t i = 0; i < n; ++i) s += (double)a[i] * (double)b[i];
    return s;
}

// prc.cu:659-710 -- the MAP-EM / Biggs-Andrews loop on preprocessed data g.
// trace_cb (optional) receives the estimate after every iteration.
template <class T>
void solve(Projector<T>& A, const T* g_in, T* out, int iterations, T beta, T gamma, bool accel,
           const unsigned char* sup, double support_sum, void (*cb)(int, const T*, void*), void* user) {
    const size_t n = (size_t)A.g.nx * A.g.ny * A.g.nz;
    std::vector<T> f(g_in, g_in + n), prev(n, 0), y(accel ? n : 0, 0), hprev(accel ? n : 0, 0), est(n, 0);
    const T eps = (T)(1.0e-6 * support_sum / (double)A.active.size());
    if (iterations > 0) {
        for (size_t i = 0; i < n; ++i)
            if (sup[i]) f[i] = std::max(f[i], (T)1000 * eps);
        const T sc = (T)(support_sum / dsum(f.data(), n));
        for (size_t i = 0; i < n; ++i) f[i] = f[i] * sc;
    }
    T lambda = 0;
    T *pf = f.data(), *pprev = prev.data(), *py = accel ? y.data() : nullptr, *ph = accel ? hprev.data() : nullptr,
      *pest = est.data();
    for (int it = 1; it <= iterations; ++it) {
        const T* start = pf;
        if (accel) {
            // prc.cu:318-321; nvcc contracts f + lambda*(f - prev) into one fma.
            for (size_t i = 0; i < n; ++i) py[i] = std::max(fma_t(lambda, pf[i] - pprev[i], pf[i]), (T)0.5 * pf[i]);
            start = py;
        }
        A.forward(start, pest);
        for (size_t i = 0; i < n; ++i) pest[i] = sup[i] ? g_in[i] / std::max(pest[i], eps) : (T)0;
        A.adjoint(pest, pprev);
        std::vector<T> next(n);
#pragma omp parallel for schedule(static)
        for (long long i = 0; i < (long long)n; ++i)
            next[i] = rdp_update(A.g, start, pprev, sup, (size_t)i, beta, gamma, eps);
        std::copy(next.begin(), next.end(), pest);
        T scale = 1;
        if (beta > 0) {
            const double s = dsum(pest, n);
            if (s > 0) scale = (T)(support_sum / s);
        }
        if (beta > 0 || accel)
            for (size_t i = 0; i < n; ++i) {
                const T v = pest[i] * scale;
                pest[i] = v;
                if (py) py[i] = v - py[i];
            }
        if (accel) {
            const double hh = ddot(ph, ph, n);
            lambda = hh > 0 ? (T)std::min(1.0, std::max(0.0, ddot(py, ph, n) / hh)) : (T)0;
            std::swap(py, ph);
            std::swap(pprev, pf);
        }
        std::swap(pf, pest);
        if (cb) cb(it, pf, user);
    }
    std::copy(pf, pf + n, out);
}

struct Handle {
    int prec = 32;
    Iso model{};
    float table[kBins + 1];
    float scale = 0;
    float cutoff = 0;
    bool exact = false;
    Projector<float> pf;
    Projector<double> pd;
};

}  // namespace

extern "C" {

int orc_isotope_index(const char* name) {
    for (int i = 0; i < 8; ++i)
        if (!strcmp(name, kIso[i].name)) return i;
    return -1;
}

// Renormalised branches: out[5*b + {w,a,b,c,f}], returns the branch count.
int orc_model(int iso, float* out) {
    const Iso m = renormalised(iso);
    for (int b = 0; b < m.n; ++b) {
        out[5 * b] = m.br[b].w, out[5 * b + 1] = m.br[b].a, out[5 * b + 2] = m.br[b].b;
        out[5 * b + 3] = m.br[b].c, out[5 * b + 4] = m.br[b].f;
    }
    return m.n;
}

float orc_survival_f32(int iso, float r) { return survival_f(renormalised(iso), r); }
double orc_survival_f64(int iso, double r) { return survival_d(renormalised(iso), r); }
float orc_range_cutoff(int iso, float fraction) { return range_cutoff_f(renormalised(iso), fraction); }

// prc.cu:625-633: scale = 4096/cutoff, table[j] = S(j/scale), j = 0..4096.
float orc_range_table(int iso, float cutoff, float* table) {
    const Iso m = renormalised(iso);
    const float scale = (float)kBins / cutoff;
    for (int j = 0; j <= kBins; ++j) table[j] = survival_f(m, (float)j / scale);
    return scale;
}

float orc_survival_at(const float* table, float scale, float w) {
    Range rg{nullptr, table, scale};
    return rg.at(w);
}

void orc_ray_f32(int k, int n, float* start, float* dir) {
    V3<float> s, d;
    ray_f(k, n, &s, &d);
    start[0] = s.x, start[1] = s.y, start[2] = s.z, dir[0] = d.x, dir[1] = d.y, dir[2] = d.z;
}

void orc_ray_f64(int k, int n, double* start, double* dir) {
    V3<double> s, d;
    ray_d(k, n, &s, &d);
    start[0] = s.x, start[1] = s.y, start[2] = s.z, dir[0] = d.x, dir[1] = d.y, dir[2] = d.z;
}

// One FP32 trace (prc.cu:134-165) with an explicit table (nullptr = analytic S of isotope iso).
// Writes up to cap deposits (linear voxel index, weight); returns the voxel-visit count and
// stores the deposit count in *n_dep.
long orc_trace_f32(int nx, int ny, int nz, float sx, float sy, float sz, float cutoff, float support_rho,
                   float air_cap, const float* rho, int iso, const float* table, float scale, int vx, int vy,
                   int vz, const float* start, const float* dir, int near, int* idx, float* wts, int cap,
                   int* n_dep) {
    const Iso m = renormalised(iso);
    Range rg{&m, table, scale};
    Grid<float> g{nx, ny, nz, sx, sy, sz, cutoff, support_rho, air_cap};
    int c = 0;
    const long steps = trace(g, rho, rg, vx, vy, vz, V3<float>{start[0], start[1], start[2]},
                             V3<float>{dir[0], dir[1], dir[2]}, near != 0, [&](int i, int j, int k, float p) {
                                 if (c < cap) idx[c] = (k * ny + j) * nx + i, wts[c] = p;
                                 ++c;
                             });
    *n_dep = c;
    return steps;
}

// Projector construction (prc.cu:405-422). rho must be preprocessed (clamped to >= 1e-3).
// prec 32: reference semantics (table unless exact != 0; table_override replaces the host
// table when non-null, e.g. with a device-built one). prec 64: analytic FP64.
void* orc_create(int prec, int iso, int nx, int ny, int nz, float sx, float sy, float sz, float support_rho,
                 float air_cap, int n_rays, int exact, const float* rho, const float* table_override) {
    Handle* h = new Handle;
    h->prec = prec;
    h->model = renormalised(iso);
    h->cutoff = range_cutoff_f(h->model, 0.9995f);
    h->exact = exact != 0;
    h->scale = (float)kBins / h->cutoff;
    for (int j = 0; j <= kBins; ++j) h->table[j] = survival_f(h->model, (float)j / h->scale);
    if (table_override) std::memcpy(h->table, table_override, sizeof(h->table));
    const size_t n = (size_t)nx * ny * nz;
    std::vector<int> active;
    for (size_t i = 0; i < n; ++i)
        if (rho[i] >= support_rho) active.push_back((int)i);
    if (prec == 64) {
        auto& p = h->pd;
        p.g = {nx, ny, nz, sx, sy, sz, h->cutoff, support_rho, air_cap};
        p.rho.assign(rho, rho + n);
        p.active = active;
        p.n_rays = n_rays;
        p.rg = Range{&h->model, nullptr, h->scale};
        p.build();
    } else {
        auto& p = h->pf;
        p.g = {nx, ny, nz, sx, sy, sz, h->cutoff, support_rho, air_cap};
        p.rho.assign(rho, rho + n);
        p.active = active;
        p.n_rays = n_rays;
        p.rg = Range{&h->model, h->exact ? nullptr : h->table, h->scale};
        p.build();
    }
    return h;
}

void orc_destroy(void* hv) { delete (Handle*)hv; }

float orc_cutoff(void* hv) { return ((Handle*)hv)->cutoff; }
int orc_n_active(void* hv) {
    Handle* h = (Handle*)hv;
    return (int)(h->prec == 64 ? h->pd.active.size() : h->pf.active.size());
}
void orc_table(void* hv, float* out) { std::memcpy(out, ((Handle*)hv)->table, sizeof(((Handle*)hv)->table)); }

// W per active voxel (ascending linear index), as double.
void orc_norm(void* hv, double* W) {
    Handle* h = (Handle*)hv;
    if (h->prec == 64) std::copy(h->pd.W.begin(), h->pd.W.end(), W);
    else std::copy(h->pf.W.begin(), h->pf.W.end(), W);
}

void orc_near_weights(void* hv, double* out) {
    Handle* h = (Handle*)hv;
    if (h->prec == 64) std::copy(h->pd.nearw.begin(), h->pd.nearw.end(), out);
    else std::copy(h->pf.nearw.begin(), h->pf.nearw.end(), out);
}

// Voxel visits: [0] near (normalisation), [1] far (accumulated over all passes so far).
void orc_vrs(void* hv, long long* out) {
    Handle* h = (Handle*)hv;
    out[0] = h->prec == 64 ? h->pd.vrs_near.load() : h->pf.vrs_near.load();
    out[1] = h->prec == 64 ? h->pd.vrs_far.load() : h->pf.vrs_far.load();
}

void orc_forward_f32(void* hv, const float* in, float* out) { ((Handle*)hv)->pf.forward(in, out); }
void orc_adjoint_f32(void* hv, const float* in, float* out) { ((Handle*)hv)->pf.adjoint(in, out); }
void orc_forward_f64(void* hv, const double* in, double* out) { ((Handle*)hv)->pd.forward(in, out); }
void orc_adjoint_f64(void* hv, const double* in, double* out) { ((Handle*)hv)->pd.adjoint(in, out); }

// Far-field deposit list of active source a: (target, weight) in ray order, near taps first.
long orc_source_column(void* hv, int a, int* idx, double* wts, int cap) {
    Handle* h = (Handle*)hv;
    long c = 0;
    auto put = [&](size_t t, double w) {
        if (c < cap) idx[c] = (int)t, wts[c] = w;
        ++c;
    };
    if (h->prec == 64) h->pd.for_each(a, [&](size_t t, double w) { put(t, w); });
    else h->pf.for_each(a, [&](size_t t, float w) { put(t, w); });
    return c;
}

typedef void (*orc_cb_f32)(int, const float*, void*);
typedef void (*orc_cb_f64)(int, const double*, void*);

// prc.cu:659-710 on preprocessed data g (zero off support). support_sum as prc.cu:597-613.
void orc_solve_f32(void* hv, const float* g, float* out, int iterations, float beta, float gamma, int accel,
                   orc_cb_f32 cb, void* user) {
    Handle* h = (Handle*)hv;
    auto& A = h->pf;
    const size_t n = (size_t)A.g.nx * A.g.ny * A.g.nz;
    std::vector<unsigned char> sup(n);
    double support_sum = 0;
    for (size_t i = 0; i < n; ++i) {
        sup[i] = A.rho[i] >= A.g.support_rho;
        if (sup[i]) support_sum += g[i];
    }
    solve<float>(A, g, out, iterations, beta, gamma, accel != 0, sup.data(), support_sum, cb, user);
}

void orc_solve_f64(void* hv, const double* g, double* out, int iterations, double beta, double gamma, int accel,
                   orc_cb_f64 cb, void* user) {
    Handle* h = (Handle*)hv;
    auto& A = h->pd;
    const size_t n = (size_t)A.g.nx * A.g.ny * A.g.nz;
    std::vector<unsigned char> sup(n);
    double support_sum = 0;
    for (size_t i = 0; i < n; ++i) {
        sup[i] = A.rho[i] >= A.g.support_rho;
        if (sup[i]) support_sum += g[i];
    }
    solve<double>(A, g, out, iterations, beta, gamma, accel != 0, sup.data(), support_sum, cb, user);
}

// One RDP surrogate update of voxel idx (prc.cu:279-309), FP64, for the mathematics tests.
double orc_rdp_update_f64(int nx, int ny, int nz, double sx, double sy, double sz, const double* f,
                          const double* back, const unsigned char* sup, long idx, double beta, double gamma,
                          double eps) {
    Grid<double> g{nx, ny, nz, sx, sy, sz, 0, 0, 0};
    return rdp_update(g, f, back, sup, (size_t)idx, beta, gamma, eps);
}

float orc_rdp_update_f32(int nx, int ny, int nz, float sx, float sy, float sz, const float* f, const float* back,
                         const unsigned char* sup, long idx, float beta, float gamma, float eps) {
    Grid<float> g{nx, ny, nz, sx, sy, sz, 0, 0, 0};
    return rdp_update(g, f, back, sup, (size_t)idx, beta, gamma, eps);
}

int orc_threads(void) { return omp_get_max_threads(); }

void orc_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }

}  // extern "C"
