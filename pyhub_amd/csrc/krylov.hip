// krylov.hip -- Lanczos ground state, continued-fraction chains and the adjoint
// MINRES solve, driven from the host over device-resident vectors.
//
// Reference: LANCZOS (pyhub/libs/src/lanczos.f90:1-163) -- a plain three-term
// recurrence, Ritz values every NADD = 5 steps, the NDEG degeneracy logic
// (:74-89), ground vector psi = V y / |V y| (compute_solver.f90:282-284).
//
// Modes
//   PLAIN64: the reference algorithm on FP64 vectors (3 resident), psi rebuilt by
//            a second pass of the same recurrence (the reference stores every
//            vector in HDF5 instead, :49).
//   FULL64 : FP64 vectors, full reorthogonalisation against the stored basis.
//   MIXED  : FP32 vectors, FP64 alpha/beta/reorth coefficients and tridiagonal,
//            FP64 reconstruction of psi, then FP64 Rayleigh-Ritz refinement steps.
// Every Krylov vector x_k is kept unnormalised with a scale sig_k (x_k = sig_k * xt_k),
// so the normalisation is folded into the next fused H*v instead of a pass.
#include "phub_internal.h"

#include <algorithm>
#include <chrono>
#include <type_traits>
#include <cstdio>
#include <cstdlib>
#include <cmath>
#include <cstring>
#include <vector>

// ---------------------------------------------------------------------------
// host tridiagonal eigensolver (implicit QL, Wilkinson shift)
// ---------------------------------------------------------------------------
// d[m] diagonal, e[m-1] off-diagonal.  Returns eigenvalues ascending in d; if z
// is non-null it receives the leading zr components (zr <= 0: all m) of the
// eigenvectors, vector j = z[j*zr .. j*zr+zr-1] (zr = 1: their first components, in
// O(m^2)).
static int tridiag_ql(int m, double *d, const double *e_in, double *z, int zr = 0) {
    if (zr <= 0 || zr > m) zr = m;
    std::vector<double> e(m + 1, 0.0);
    for (int i = 0; i + 1 < m; ++i) e[i] = e_in[i];
    if (z) {
        std::fill(z, z + (size_t)m * zr, 0.0);
        for (int i = 0; i < zr; ++i) z[(size_t)i * zr + i] = 1.0;
    }
    for (int l = 0; l < m; ++l) {
        int it = 0;
        while (true) {
            int mm = l;
            for (; mm < m - 1; ++mm) {
                double dd = std::fabs(d[mm]) + std::fabs(d[mm + 1]);
                if (std::fabs(e[mm]) <= 2.2e-16 * dd + 1e-300) break;
            }
            if (mm == l) break;
            if (++it > 300) return 1;
            double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
            double r = std::hypot(g, 1.0);
            g = d[mm] - d[l] + e[l] / (g + std::copysign(r, g));
            double s = 1.0, c = 1.0, p = 0.0;
            int i = mm - 1;
            bool early = false;
            for (; i >= l; --i) {
                double f = s * e[i], b = c * e[i];
                r = std::hypot(f, g);
                e[i + 1] = r;
                if (r == 0.0) {
                    d[i + 1] -= p;
                    e[mm] = 0.0;
                    early = true;
                    break;
                }
                s = f / r;
                c = g / r;
                g = d[i + 1] - p;
                r = (d[i] - g) * s + 2.0 * c * b;
                p = s * r;
                d[i + 1] = g + p;
                g = c * r - b;
                if (z) {
                    double *zi = z + (size_t)i * zr, *zi1 = z + (size_t)(i + 1) * zr;
                    for (int k = 0; k < zr; ++k) {
                        double t = zi1[k];
                        zi1[k] = s * zi[k] + c * t;
                        zi[k] = c * zi[k] - s * t;
                    }
                }
            }
            if (early) continue;
            d[l] -= p;
            e[l] = g;
            e[mm] = 0.0;
        }
    }
    // selection sort ascending
    for (int i = 0; i + 1 < m; ++i) {
        int k = i;
        for (int j = i + 1; j < m; ++j)
            if (d[j] < d[k]) k = j;
        if (k != i) {
            std::swap(d[i], d[k]);
            if (z)
                for (int q = 0; q < zr; ++q) std::swap(z[(size_t)i * zr + q], z[(size_t)k * zr + q]);
        }
    }
    return 0;
}

// lowest Ritz pair of the m x m Lanczos tridiagonal (alpha[0..m-1], beta[1..m-1])
static int lowest_ritz(int m, const double *alpha, const double *beta, double *e0, double *e1,
                       std::vector<double> *y) {
    std::vector<double> d(alpha, alpha + m), e(std::max(m - 1, 1), 0.0);
    for (int i = 0; i + 1 < m; ++i) e[i] = beta[i + 1];
    std::vector<double> z;
    if (y) z.resize((size_t)m * m);
    if (tridiag_ql(m, d.data(), e.data(), y ? z.data() : nullptr)) return 1;
    *e0 = d[0];
    if (e1) *e1 = (m > 1) ? d[1] : d[0];
    if (y) y->assign(z.begin(), z.begin() + m);
    return 0;
}

// ---------------------------------------------------------------------------
// vector kernels
// ---------------------------------------------------------------------------
__device__ inline uint64_t splitmix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

template <typename T>
__global__ void k_randvec(T *__restrict__ x, int64_t n, uint64_t seed) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        uint64_t r = splitmix64(seed * 0x632BE59BD9B4E019ull + (uint64_t)i);
        x[i] = (T)((double)(r >> 11) * (1.0 / 9007199254740992.0) - 0.5);
    }
}

template <typename T>
__global__ void k_convert(const double *__restrict__ a, T *__restrict__ b, int64_t n, double s) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = (T)(s * a[i]);
}

__device__ __forceinline__ double wave_sum(double a) {
    for (int off = 32; off > 0; off >>= 1) a += __shfl_xor(a, off);
    return a;
}

// 16-byte packs (four FP32 elements per lane access) for the FP32 y = a x + b y pass,
// where they measured faster (0.47 -> 0.40 ms on 1.66e8 elements).  The passes with
// more streams or FP64 elements measured slower with them (the MINRES passes 2.61 ->
// 2.96 ms per iteration, the FP64 y = a x + b y 0.72 -> 0.80 ms) and loop per element.
// The pack loop runs when every pointer is 16-byte aligned (a wave-uniform test); the
// last n mod 4 elements, or all of them otherwise, go element by element.
template <typename T, int W>
using vpk = T __attribute__((ext_vector_type(W)));
template <typename V>
constexpr int vpk_w() { return 16 / (int)sizeof(V); }
__device__ __forceinline__ bool al16(const void *p) { return ((uintptr_t)p & 15) == 0; }

// per-block partial sums of a.b over a grid-stride range
template <typename TA, typename TB>
__global__ __launch_bounds__(256) void k_dot(const TA *__restrict__ a, const TB *__restrict__ b,
                                             int64_t n, double *__restrict__ part) {
    __shared__ double sh[4];
    double s = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        s += (double)a[i] * (double)b[i];
    s = wave_sum(s);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0] + sh[1] + sh[2] + sh[3];
}

// y = a*x + b*y ; partial |y|^2
template <typename TX, typename TY>
__global__ __launch_bounds__(256) void k_axpby(double a, const TX *__restrict__ x, double b,
                                               TY *__restrict__ y, int64_t n,
                                               double *__restrict__ part,
                                               const double *__restrict__ dsc = nullptr) {
    __shared__ double sh[4];
    // dsc: a graph-captured Lanczos step takes a = -alpha_k * scale from device memory
    if (dsc) a = -dsc[2] * dsc[0];
    double s = 0.0;
    const int64_t t0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t st = (int64_t)gridDim.x * blockDim.x;
    int64_t i0 = 0;
    if constexpr (sizeof(TX) == 4 && sizeof(TY) == 4) {
        constexpr int W = vpk_w<TY>();
        using PX = vpk<TX, W>;
        using PY = vpk<TY, W>;
        if (al16(x) && al16(y)) {
            const int64_t nq = n / W;
            for (int64_t q = t0; q < nq; q += st) {
                PY yv = reinterpret_cast<const PY *>(y)[q];
                PX xv{};
                if (x) xv = reinterpret_cast<const PX *>(x)[q];
#pragma unroll
                for (int e = 0; e < W; ++e) {
                    const TY t = (TY)((x ? a * (double)xv[e] : 0.0) + b * (double)yv[e]);
                    yv[e] = t;
                    s += (double)t * (double)t;
                }
                reinterpret_cast<PY *>(y)[q] = yv;
            }
            i0 = nq * W;
        }
    }
    for (int64_t i = i0 + t0; i < n; i += st) {
        double v = (x ? a * (double)x[i] : 0.0) + b * (double)y[i];
        TY t = (TY)v;
        y[i] = t;
        s += (double)t * (double)t;
    }
    if (part) {
        s = wave_sum(s);
        if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
        __syncthreads();
        if (threadIdx.x == 0) part[blockIdx.x] = sh[0] + sh[1] + sh[2] + sh[3];
    }
}

// Kernels of a graph-captured three-term Lanczos step, whose scalars live in device
// memory: dsc[0] the scale of x_k, dsc[1] beta_k times the scale of x_{k-1},
// dsc[2] alpha_k = <scale x_k, out>, dsc[3] |out|^2 after the alpha update (the
// alpha update itself is k_axpby with its coefficient from dsc).
// one thread: alpha_k and beta_{k+1} into their arrays, the scales of the next step
// (beta_{k+1} = |out|, scale 1 / beta_{k+1}), k advanced
__global__ void k_lcz_step_end(double *__restrict__ dsc, double *__restrict__ alpha,
                               double *__restrict__ beta, int *__restrict__ steps) {
    const int k = steps[0];
    const double a = dsc[2], b = sqrt(fmax(dsc[3], 0.0));
    alpha[k] = a;
    beta[k + 1] = b;
    dsc[1] = b * dsc[0];
    dsc[0] = (b > 0.0) ? 1.0 / b : 0.0;
    steps[0] = k + 1;
}

// the new Lanczos vector into its slot of the store: slots[k] after the step ended
template <typename T>
__global__ __launch_bounds__(256) void k_store_vec(const T *__restrict__ src,
                                                   T *const *__restrict__ slots,
                                                   const int *__restrict__ steps, int64_t n) {
    T *dst = slots[steps[0]];
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

// r = a * psi + b * y (psi FP64, y and r of type V)
template <typename V>
__global__ __launch_bounds__(256) void k_axpby_to(double a, const V *__restrict__ psi,
                                                  double b, const V *__restrict__ y,
                                                  V *__restrict__ r, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        r[i] = (V)(a * (double)psi[i] + b * (double)y[i]);
}

// Adjoint MINRES, the vector work of one iteration in three passes.  Vectors of
// type V (FP64 for the adjoint solve, FP32 for the loose correction solves of
// jd_refine, which pass an FP32 copy of psi as well), every sum in FP64.
// k_minres_scale: v = s * y.
template <typename V>
__global__ __launch_bounds__(256) void k_minres_scale(V *__restrict__ v, const V *__restrict__ y,
                                                      double s, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        v[i] = (V)(s * (double)y[i]);
}

// preconditioned form: v = s * (minv * y - gam * psi), the preconditioned residual
// P M^-1 y (y orthogonal to psi, gam = psi . (minv * y)) scaled by s
template <typename V>
__global__ __launch_bounds__(256) void k_minres_pscale(V *__restrict__ v, const V *__restrict__ y,
                                                       const V *__restrict__ minv,
                                                       const V *__restrict__ psi, double gam,
                                                       double s, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        v[i] = (V)(s * ((double)minv[i] * (double)y[i] - gam * (double)psi[i]));
}

// after t = H v: t <- t - e0 v - c1 r1 (no r1 term in the first iteration); partials
// [block][3] of (psi.(t - e0 v), v.t, v.psi).  psi is projected out of (H - e0) v
// before the r1 term: taking it from the whole of t would also strip r1's own
// small psi component, which breaks the three-term relation the MINRES update
// relies on (slower convergence, 1e-11 instead of 1e-13 in a 60-state test).  R1: r1
// is read (every iteration after the first); a template flag, so that no load in the
// loop is conditional (a conditional load measured 30 % slower in the next pass).
template <typename V, bool R1>
__global__ __launch_bounds__(256) void k_minres_a(V *__restrict__ t, const V *__restrict__ v,
                                                  const V *__restrict__ r1,
                                                  const V *__restrict__ psi, double e0,
                                                  double c1, int64_t n,
                                                  double *__restrict__ part) {
    __shared__ double sh[3][4];
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    auto one = [&](double ti, double vi, double pi, double ri) -> V {
        ti -= e0 * vi;
        s0 += pi * ti;
        ti -= c1 * ri;
        const V tv = (V)ti;
        s1 += vi * (double)tv;
        s2 += vi * pi;
        return tv;
    };
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        t[i] = one((double)t[i], (double)v[i], (double)psi[i], R1 ? (double)r1[i] : 0.0);
    s0 = wave_sum(s0);
    s1 = wave_sum(s1);
    s2 = wave_sum(s2);
    if ((threadIdx.x & 63) == 0) {
        sh[0][threadIdx.x >> 6] = s0;
        sh[1][threadIdx.x >> 6] = s1;
        sh[2][threadIdx.x >> 6] = s2;
    }
    __syncthreads();
    if (threadIdx.x < 3)
        part[(size_t)blockIdx.x * 3 + threadIdx.x] =
            sh[threadIdx.x][0] + sh[threadIdx.x][1] + sh[threadIdx.x][2] + sh[threadIdx.x][3];
}

// r = t - p psi - c2 r2 (r may not alias t, psi or r2); partials [block][2]: without
// the preconditioner (PC false, minv unused) |r|^2 and 0, with it r . (minv r) and
// psi . (minv r)
template <typename V, bool PC>
__global__ __launch_bounds__(256) void k_minres_b(const V *__restrict__ t,
                                                  const V *__restrict__ psi,
                                                  const V *__restrict__ r2, double p, double c2,
                                                  const V *__restrict__ minv,
                                                  V *__restrict__ r, int64_t n,
                                                  double *__restrict__ part) {
    __shared__ double sh[2][4];
    double s0 = 0.0, s1 = 0.0;
    auto one = [&](double ti, double pi, double r2i, double mi) -> V {
        const V ri = (V)(ti - p * pi - c2 * r2i);
        const double rd = (double)ri;
        if constexpr (PC) {
            const double mr = mi * rd;
            s0 += rd * mr;
            s1 += pi * mr;
        } else {
            s0 += rd * rd;
        }
        return ri;
    };
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        r[i] = one((double)t[i], (double)psi[i], (double)r2[i], PC ? (double)minv[i] : 0.0);
    s0 = wave_sum(s0);
    s1 = wave_sum(s1);
    if ((threadIdx.x & 63) == 0) {
        sh[0][threadIdx.x >> 6] = s0;
        sh[1][threadIdx.x >> 6] = s1;
    }
    __syncthreads();
    if (threadIdx.x < 2)
        part[(size_t)blockIdx.x * 2 + threadIdx.x] =
            sh[threadIdx.x][0] + sh[threadIdx.x][1] + sh[threadIdx.x][2] + sh[threadIdx.x][3];
}

// w = (v - oe wp - de wl) * ig, x += phi w
template <typename V>
__global__ __launch_bounds__(256) void k_minres_c(const V *__restrict__ v, const V *__restrict__ wp,
                                                  const V *__restrict__ wl, double oe, double de,
                                                  double ig, double phi, V *__restrict__ w,
                                                  V *__restrict__ x, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const double wi = (((double)v[i] - oe * (double)wp[i]) - de * (double)wl[i]) * ig;
        w[i] = (V)wi;
        x[i] = (V)((double)x[i] + phi * wi);
    }
}

// second pass of the plain recurrence: out = (out - a*x) * binv ; psi += y*out
template <typename T>
__global__ __launch_bounds__(256) void k_second_pass(T *__restrict__ out, const T *__restrict__ x,
                                                     double a, double binv,
                                                     double *__restrict__ psi, double y,
                                                     int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        double v = ((double)out[i] - a * (double)x[i]) * binv;
        T t = (T)v;
        out[i] = t;
        psi[i] += y * (double)t;
    }
}

// multi-dot against a stored basis: part[block*k + j] = sum_slice xt_j . w
// The block stages its slice of w in LDS; each wave takes basis vectors
// j = wave, wave+4, ... and reduces its dot with DPP shuffles.
constexpr int MD_SLICE = 2048;
// The staged vector is w - a_eff * xk (formed in FP64): projecting the dominant
// local component out before the global coefficients are taken keeps the
// reorthogonalisation stable with one pass (otherwise the overlap with older
// vectors is multiplied by alpha/beta every step).
template <typename T, typename TW>
__global__ __launch_bounds__(256) void k_multidot(const T *__restrict__ basis, int64_t ld, int k,
                                                  const TW *__restrict__ w,
                                                  const T *__restrict__ xk, double a_eff,
                                                  int64_t n, double *__restrict__ part) {
    __shared__ double ws[MD_SLICE];
    extern __shared__ double acc[];   // k doubles
    for (int j = threadIdx.x; j < k; j += blockDim.x) acc[j] = 0.0;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int64_t base = (int64_t)blockIdx.x * MD_SLICE; base < n;
         base += (int64_t)gridDim.x * MD_SLICE) {
        const int len = (int)std::min<int64_t>(MD_SLICE, n - base);
        __syncthreads();
        for (int i = threadIdx.x; i < MD_SLICE; i += blockDim.x)
            ws[i] = (i < len) ? (double)w[base + i] - (xk ? a_eff * (double)xk[base + i] : 0.0) : 0.0;
        __syncthreads();
        for (int j = wave; j < k; j += 4) {
            const T *vj = basis + (size_t)j * ld + base;
            double s = 0.0;
            for (int i = lane; i < len; i += 64) s += (double)vj[i] * ws[i];
            s = wave_sum(s);
            if (lane == 0) acc[j] += s;
        }
    }
    __syncthreads();
    for (int j = threadIdx.x; j < k; j += blockDim.x) part[(size_t)blockIdx.x * k + j] = acc[j];
}

// Vectorised forms of the two kernels above for 16-byte aligned vectors (the
// Krylov basis is allocated with a leading dimension padded to 64 elements).
// k_multidot_v: each thread keeps its 16 / sizeof(T) * NV elements of the
// staged w - a_eff * xk in FP64 registers and streams every basis vector
// through 16-byte loads (two vectors per iteration, so eight loads are in
// flight); per-wave partial dots go to LDS, one slot per wave and vector.
template <typename T, typename TW>
__global__ __launch_bounds__(256) void k_multidot_v(const T *__restrict__ basis, int64_t ld, int k,
                                                    const TW *__restrict__ w,
                                                    const T *__restrict__ xk, double a_eff,
                                                    int64_t n, double *__restrict__ part,
                                                    const T *const *__restrict__ ptrs) {
    // vector j: ptrs[j] when a pointer table is given, else basis + j * ld
    auto vec = [&](int j) { return ptrs ? ptrs[j] : basis + (size_t)j * ld; };
    constexpr int VEC = 16 / sizeof(T);
    constexpr int NV = 4;
    constexpr int E = VEC * NV;
    constexpr int64_t S = 256 * E;
    extern __shared__ double acc[];   // 4 * k
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int j = threadIdx.x; j < 4 * k; j += 256) acc[j] = 0.0;
    __syncthreads();
    for (int64_t base = (int64_t)blockIdx.x * S; base < n; base += (int64_t)gridDim.x * S) {
        const bool full = base + S <= n;
        double ws[E];
#pragma unroll
        for (int q = 0; q < NV; ++q)
#pragma unroll
            for (int e = 0; e < VEC; ++e) {
                const int64_t i = base + ((int64_t)q * 256 + threadIdx.x) * VEC + e;
                ws[q * VEC + e] = (full || i < n)
                                      ? (double)w[i] - (xk ? a_eff * (double)xk[i] : 0.0)
                                      : 0.0;
            }
        auto dot = [&](const T *vj) -> double {
            double sacc = 0.0;
            if (full) {
                using V4 = typename std::conditional<sizeof(T) == 4, float4, double2>::type;
                V4 v[NV];
#pragma unroll
                for (int q = 0; q < NV; ++q)
                    v[q] = *reinterpret_cast<const V4 *>(vj + base + ((int64_t)q * 256 + threadIdx.x) * VEC);
#pragma unroll
                for (int q = 0; q < NV; ++q) {
                    const T *pv = reinterpret_cast<const T *>(&v[q]);
#pragma unroll
                    for (int e = 0; e < VEC; ++e) sacc += ws[q * VEC + e] * (double)pv[e];
                }
            } else {
#pragma unroll
                for (int q = 0; q < NV; ++q)
#pragma unroll
                    for (int e = 0; e < VEC; ++e) {
                        const int64_t i = base + ((int64_t)q * 256 + threadIdx.x) * VEC + e;
                        if (i < n) sacc += ws[q * VEC + e] * (double)vj[i];
                    }
            }
            return sacc;
        };
        int j = 0;
        for (; j + 1 < k; j += 2) {
            double s0 = dot(vec(j));
            double s1 = dot(vec(j + 1));
            s0 = wave_sum(s0);
            s1 = wave_sum(s1);
            if (lane == 0) {
                acc[wave * k + j] += s0;
                acc[wave * k + j + 1] += s1;
            }
        }
        if (j < k) {
            const double s0 = wave_sum(dot(vec(j)));
            if (lane == 0) acc[wave * k + j] += s0;
        }
    }
    __syncthreads();
    for (int j = threadIdx.x; j < k; j += 256)
        part[(size_t)blockIdx.x * k + j] = acc[j] + acc[k + j] + acc[2 * k + j] + acc[3 * k + j];
}

// k_combine_v: w -= sum_j coef_j v_j over 16 / sizeof(T) * 4 elements per thread
// with 16-byte loads (two basis vectors per iteration, eight loads in flight);
// psi_mode as k_combine.
template <typename T, typename TW>
__global__ __launch_bounds__(256) void k_combine_v(const T *__restrict__ basis, int64_t ld, int k,
                                                   const double *__restrict__ coef,
                                                   TW *__restrict__ w, int64_t n, int psi_mode,
                                                   double *__restrict__ part,
                                                   const T *const *__restrict__ ptrs) {
    constexpr int VEC = 16 / sizeof(T);
    constexpr int NV = 4;
    constexpr int E = VEC * NV;
    constexpr int64_t S = 256 * E;
    using V4 = typename std::conditional<sizeof(T) == 4, float4, double2>::type;
    extern __shared__ double cs[];
    __shared__ double sh[4];
    for (int j = threadIdx.x; j < k; j += blockDim.x) cs[j] = coef[j];
    __syncthreads();
    double s2 = 0.0;
    for (int64_t base = (int64_t)blockIdx.x * S; base < n; base += (int64_t)gridDim.x * S) {
        const bool full = base + S <= n;
        double sum[E];
#pragma unroll
        for (int e = 0; e < E; ++e) sum[e] = 0.0;
        auto add = [&](int j) {
            const T *vj = (ptrs ? ptrs[j] : basis + (size_t)j * ld) + base;
            const double cj = cs[j];
            if (full) {
                V4 v[NV];
#pragma unroll
                for (int q = 0; q < NV; ++q)
                    v[q] = *reinterpret_cast<const V4 *>(vj + ((int64_t)q * 256 + threadIdx.x) * VEC);
#pragma unroll
                for (int q = 0; q < NV; ++q) {
                    const T *pv = reinterpret_cast<const T *>(&v[q]);
#pragma unroll
                    for (int e = 0; e < VEC; ++e) sum[q * VEC + e] += cj * (double)pv[e];
                }
            } else {
#pragma unroll
                for (int q = 0; q < NV; ++q)
#pragma unroll
                    for (int e = 0; e < VEC; ++e) {
                        const int64_t o = ((int64_t)q * 256 + threadIdx.x) * VEC + e;
                        if (base + o < n) sum[q * VEC + e] += cj * (double)vj[o];
                    }
            }
        };
        int j = 0;
        for (; j + 1 < k; j += 2) {
            add(j);
            add(j + 1);
        }
        if (j < k) add(j);
#pragma unroll
        for (int q = 0; q < NV; ++q)
#pragma unroll
            for (int e = 0; e < VEC; ++e) {
                const int64_t i = base + ((int64_t)q * 256 + threadIdx.x) * VEC + e;
                if (full || i < n) {
                    const double v = psi_mode ? sum[q * VEC + e] : (double)w[i] - sum[q * VEC + e];
                    const TW t = (TW)v;
                    w[i] = t;
                    s2 += (double)t * (double)t;
                }
            }
    }
    if (part) {
        s2 = wave_sum(s2);
        if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s2;
        __syncthreads();
        if (threadIdx.x == 0) part[blockIdx.x] = sh[0] + sh[1] + sh[2] + sh[3];
    }
}

// w -= sum_j coef_j * xt_j  (coef in FP64); partial |w|^2.  When psi_mode,
// w = sum_j coef_j xt_j (w need not be initialised).
template <typename T, typename TW>
__global__ __launch_bounds__(256) void k_combine(const T *__restrict__ basis, int64_t ld, int k,
                                                 const double *__restrict__ coef,
                                                 TW *__restrict__ w, int64_t n, int psi_mode,
                                                 double *__restrict__ part) {
    extern __shared__ double cs[];
    __shared__ double sh[4];
    for (int j = threadIdx.x; j < k; j += blockDim.x) cs[j] = coef[j];
    __syncthreads();
    double s2 = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        double v = psi_mode ? 0.0 : (double)w[i];
        double sum = 0.0;
        for (int j = 0; j < k; ++j) sum += cs[j] * (double)basis[(size_t)j * ld + i];
        v = psi_mode ? sum : v - sum;
        TW t = (TW)v;
        w[i] = t;
        s2 += (double)t * (double)t;
    }
    if (part) {
        s2 = wave_sum(s2);
        if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s2;
        __syncthreads();
        if (threadIdx.x == 0) part[blockIdx.x] = sh[0] + sh[1] + sh[2] + sh[3];
    }
}

// ---------------------------------------------------------------------------
// small host helpers
// ---------------------------------------------------------------------------
static const int VEC_GRID = 4096;

static inline unsigned grid_for(int64_t n) {
    int64_t b = (n + 255) / 256;
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>(b, VEC_GRID));
}

template <typename TA, typename TB>
static int dev_dot(phub_ctx *c, const TA *a, const TB *b, int64_t n, double *out) {
    unsigned g = grid_for(n);
    int rc = phub_ensure_red(c, g + 64);
    if (rc) return rc;
    hipLaunchKernelGGL((k_dot<TA, TB>), dim3(g), dim3(256), 0, c->stream, a, b, n, c->d_red);
    PHUB_HIP(c, hipGetLastError());
    return phub_reduce_partials(c, (int)g, 1, out);
}

template <typename TX, typename TY>
static int dev_axpby(phub_ctx *c, double a, const TX *x, double b, TY *y, int64_t n,
                     double *norm2) {
    unsigned g = grid_for(n);
    int rc = phub_ensure_red(c, g + 64);
    if (rc) return rc;
    hipLaunchKernelGGL((k_axpby<TX, TY>), dim3(g), dim3(256), 0, c->stream, a, x, b, y, n,
                       norm2 ? c->d_red : nullptr);
    PHUB_HIP(c, hipGetLastError());
    if (norm2) return phub_reduce_partials(c, (int)g, 1, norm2);
    return PHUB_OK;
}

// 16-byte aligned pointer and a leading dimension that keeps every vector aligned
static inline bool vec_ok(const void *p, int64_t ld) {
    return ((uintptr_t)p % 16 == 0) && (ld % 4 == 0);
}

// leading dimension of a stored Krylov basis: padded to 64 elements (256 B in FP32)
static inline int64_t basis_ld(int64_t n) { return (n + 63) / 64 * 64; }

template <typename T, typename TW>
static int dev_multidot(phub_ctx *c, const T *basis, int64_t ld, int k, const TW *w, int64_t n,
                        double *out, const T *xk = nullptr, double a_eff = 0.0) {
    if (k > 4096) return phub_fail(c, PHUB_ERR_UNSUPPORTED, "too many basis vectors");
    unsigned g;
    if (vec_ok(basis, ld) && vec_ok(w, 0) && (!xk || vec_ok(xk, 0)) && k <= 1536) {
        const int64_t slices = (n + 256 * 64 / (int64_t)sizeof(T) - 1) / (256 * 64 / (int64_t)sizeof(T));
        g = (unsigned)std::max<int64_t>(1, std::min<int64_t>(slices, 2048));
        int rc = phub_ensure_red(c, (size_t)g * k + k + 64);
        if (rc) return rc;
        hipLaunchKernelGGL((k_multidot_v<T, TW>), dim3(g), dim3(256), 4 * k * sizeof(double),
                           c->stream, basis, ld, k, w, xk, a_eff, n, c->d_red,
                           (const T *const *)nullptr);
    } else {
        const int64_t slices = (n + MD_SLICE - 1) / MD_SLICE;
        g = (unsigned)std::max<int64_t>(1, std::min<int64_t>(slices, 1024));
        int rc = phub_ensure_red(c, (size_t)g * k + k + 64);
        if (rc) return rc;
        hipLaunchKernelGGL((k_multidot<T, TW>), dim3(g), dim3(256), k * sizeof(double), c->stream,
                           basis, ld, k, w, xk, a_eff, n, c->d_red);
    }
    PHUB_HIP(c, hipGetLastError());
    // reduce k components; host buffer may need to grow
    if ((size_t)k > c->h_red_cap) {
        if (c->h_red) hipHostFree(c->h_red);
        PHUB_HIP(c, phub_host_malloc(&c->h_red, (size_t)k * sizeof(double), hipHostMallocCoherent));
        c->h_red_cap = k;
    }
    return phub_reduce_partials(c, (int)g, k, out);
}

template <typename T, typename TW>
static int dev_combine(phub_ctx *c, const T *basis, int64_t ld, int k, const double *coef_host,
                       TW *w, int64_t n, bool psi_mode, double *norm2, double *d_coef) {
    PHUB_HIP(c, hipMemcpyAsync(d_coef, coef_host, k * sizeof(double), hipMemcpyHostToDevice,
                               c->stream));
    unsigned g = grid_for(n);
    int rc = phub_ensure_red(c, g + 64);
    if (rc) return rc;
    if (vec_ok(basis, ld) && vec_ok(w, 0) && sizeof(TW) == sizeof(T))
        hipLaunchKernelGGL((k_combine_v<T, TW>), dim3(g), dim3(256), k * sizeof(double), c->stream,
                           basis, ld, k, d_coef, w, n, psi_mode ? 1 : 0,
                           norm2 ? c->d_red : nullptr, (const T *const *)nullptr);
    else
        hipLaunchKernelGGL((k_combine<T, TW>), dim3(g), dim3(256), k * sizeof(double), c->stream,
                           basis, ld, k, d_coef, w, n, psi_mode ? 1 : 0,
                           norm2 ? c->d_red : nullptr);
    PHUB_HIP(c, hipGetLastError());
    if (norm2) return phub_reduce_partials(c, (int)g, 1, norm2);
    // keep d_coef alive until the kernel has read it
    PHUB_HIP(c, hipStreamSynchronize(c->stream));
    return PHUB_OK;
}

// psi-mode combine over vectors at arbitrary (16-byte aligned) addresses:
// w = sum_j coef_j vs[j]; d_ptrs holds at least k pointers
template <typename T, typename TW>
static int dev_combine_ptrs(phub_ctx *c, T *const *vs, int k, const double *coef_host, TW *w,
                            int64_t n, double *d_coef, const T **d_ptrs) {
    PHUB_HIP(c, hipMemcpyAsync(d_ptrs, vs, k * sizeof(T *), hipMemcpyHostToDevice, c->stream));
    PHUB_HIP(c, hipMemcpyAsync(d_coef, coef_host, k * sizeof(double), hipMemcpyHostToDevice,
                               c->stream));
    hipLaunchKernelGGL((k_combine_v<T, TW>), dim3(grid_for(n)), dim3(256), k * sizeof(double),
                       c->stream, (const T *)nullptr, (int64_t)0, k, d_coef, w, n, 1,
                       (double *)nullptr, (const T *const *)d_ptrs);
    PHUB_HIP(c, hipGetLastError());
    // the host arrays must stay valid until the uploads have been consumed
    PHUB_HIP(c, hipStreamSynchronize(c->stream));
    return PHUB_OK;
}

struct EvTimer {
    phub_ctx *c;
    hipEvent_t a, b;
    double total_ms = 0.0;
    int n = 0;
    explicit EvTimer(phub_ctx *cc) : c(cc) {
        hipEventCreate(&a);
        hipEventCreate(&b);
    }
    ~EvTimer() {
        hipEventDestroy(a);
        hipEventDestroy(b);
    }
    void start() { hipEventRecord(a, c->stream); }
    void stop() {
        hipEventRecord(b, c->stream);
        hipEventSynchronize(b);
        float ms = 0;
        hipEventElapsedTime(&ms, a, b);
        total_ms += ms;
        ++n;
    }
};

// fused step: out = sig_k H xt_k - beta_eff xt_{k-1}; returns <x_k,out>, |out|^2
template <typename T>
static int fused_hv(phub_ctx *c, const phub_op *op, const T *xk, double sig, const T *xp,
                    double beta_eff, T *out, double *dots, EvTimer *tm) {
    int nb = 0;
    if (tm) tm->start();
    int rc = phub_hv_launch<T>(c, op, xk, out, sig, xp, beta_eff, true, &nb);
    if (rc) return rc;
    if (tm) tm->stop();
    return phub_reduce_partials(c, nb, 2, dots);
}

// ---------------------------------------------------------------------------
// Lanczos
// ---------------------------------------------------------------------------
struct LczState {
    std::vector<double> alpha, beta;   // beta[0] = start norm, beta[k] couples k-1,k
    int steps = 0;
    int ndeg = 0;
    bool converged = false;
    bool breakdown = false;
    bool norm_lost = false;
    double ev = 0.0, e0 = 0.0;
    std::vector<double> prev;          // lowest Ritz values at the previous check
    // three-term recurrence: Ritz values within ghost_tol (relative) of a lower one are
    // copies of it (0: full reorthogonalisation, where a repeat is a degenerate level)
    double ghost_tol = 0.0;
};

// convergence test in the reference's form (lanczos.f90:74-89) on the Ritz values
static void ref_convergence(LczState &s, double tol, int max_steps, const double *ritz) {
    if (std::fabs(ritz[s.ndeg] - s.ev) < tol) {
        if (std::fabs(ritz[s.ndeg] - ritz[0]) > tol && (s.ndeg + 1) > 1)
            s.converged = true;
        else
            s.ndeg += 1;
    } else {
        if (s.steps <= max_steps - 10) s.ev = ritz[s.ndeg];
    }
}

// The positions of the nst lowest genuine Ritz values (ascending d, u1 the first
// components of their tridiagonal eigenvectors).  In exact arithmetic the Ritz values
// of a single-vector Lanczos run are simple; without reorthogonalisation, rounding
// adds copies of the values that have converged (Paige's ghosts) and, while a copy
// forms, a value whose eigenvector has no weight on the start vector (Cullum and
// Willoughby's spurious values: also eigenvalues of T without its first row and
// column).  Either would pass for an excited state.  So: values within tol (relative)
// of a cluster's first member are one cluster (the copies agree to about 1e-8 with FP64
// vectors, 1e-5 with FP32 ones), represented by its member of largest |u1|; a value
// alone with |u1| < 1e-10 is spurious (measured: below 1e-18 for the spurious values,
// above 1e-6 for the weakest genuine cluster of a 300-state test).  The lowest value is
// always the ground state, with its own vector (the energy reported is that value).
// When fewer genuine
// values exist than asked for, the rest fill up in ascending order.  tol = 0 (full
// reorthogonalisation, where a repeated value is a degenerate level): the nst lowest.
static std::vector<int> genuine_ritz(const double *d, const double *u1, int m, int nst, double tol) {
    std::vector<int> keep;
    if (!(tol > 0.0) || !u1) {
        for (int i = 0; i < std::min(m, nst); ++i) keep.push_back(i);
        return keep;
    }
    std::vector<char> used(std::max(m, 1), 0);
    for (int i = 0; i < m && (int)keep.size() < nst;) {
        int j = i + 1, best = i;
        while (j < m && d[j] - d[i] <= tol * std::max(1.0, std::fabs(d[j]))) {
            if (std::fabs(u1[j]) > std::fabs(u1[best])) best = j;
            ++j;
        }
        if (i == 0) best = 0;   // the ground state: the lowest value and its vector, as ever
        if (i == 0 || j - i > 1 || std::fabs(u1[i]) >= 1e-10) {
            keep.push_back(best);
            used[best] = 1;
        }
        i = j;
    }
    for (int i = 0; i < m && (int)keep.size() < nst; ++i)
        if (!used[i]) keep.push_back(i);
    std::sort(keep.begin(), keep.end());
    return keep;
}

// Ritz values of the m x m Lanczos tridiagonal and the first components of their
// eigenvectors
static int ritz_first(int m, const double *alpha, const double *beta, std::vector<double> &vals,
                      std::vector<double> &u1) {
    vals.assign(alpha, alpha + m);
    std::vector<double> e(std::max(m - 1, 1), 0.0);
    for (int i = 0; i + 1 < m; ++i) e[i] = beta[i + 1];
    u1.assign(std::max(m, 1), 0.0);
    return tridiag_ql(m, vals.data(), e.data(), u1.data(), 1);
}

// With several requested states the run also continues until each of the nst lowest
// distinct Ritz values has moved by less than tol since the previous check (the
// reference's test alone follows the two lowest distinct levels).  nst = 1 is the
// reference's test.
static void convergence(LczState &s, double tol, int max_steps, const std::vector<double> &ritz,
                        int nst) {
    ref_convergence(s, tol, max_steps, ritz.data());
    if (nst <= 1) return;
    std::vector<double> vals = ritz, u1;
    if (s.ghost_tol > 0.0 && ritz_first(s.steps, s.alpha.data(), s.beta.data(), vals, u1)) u1.clear();
    const std::vector<int> k_of = genuine_ritz(vals.data(), u1.empty() ? nullptr : u1.data(),
                                               (int)vals.size(), nst, s.ghost_tol);
    std::vector<double> cur;
    for (int k : k_of) cur.push_back(vals[k]);
    bool stable = (int)cur.size() >= nst && (int)s.prev.size() >= nst;
    for (int k = 0; stable && k < nst; ++k) stable = std::fabs(cur[k] - s.prev[k]) < tol;
    s.prev = cur;
    if (!stable) s.converged = false;
}

// With a psi_tol, convergence also needs the Lanczos residual estimate of the lowest
// Ritz pair, beta_m |y_m| = ||H x - theta x|| in exact arithmetic, below psi_tol.
// FP32 vectors cannot represent psi below about 1e-6 relative, so their target is
// raised to that floor (the FP64 corrections take psi the rest of the way).
static void residual_rule(LczState &s, double psi_tol, bool fp32) {
    if (!s.converged || !(psi_tol > 0.0) || s.breakdown) return;
    double e0 = 0, e1 = 0;
    std::vector<double> y;
    if (lowest_ritz(s.steps, s.alpha.data(), s.beta.data(), &e0, &e1, &y)) return;
    const double target = fp32 ? std::max(psi_tol, 1e-6 * std::max(1.0, std::fabs(e0))) : psi_tol;
    if (std::fabs(s.beta[s.steps] * y[s.steps - 1]) > target) s.converged = false;
}

static int all_ritz(int m, const double *alpha, const double *beta, std::vector<double> &out) {
    out.assign(alpha, alpha + m);
    std::vector<double> e(std::max(m - 1, 1), 0.0);
    for (int i = 0; i + 1 < m; ++i) e[i] = beta[i + 1];
    return tridiag_ql(m, out.data(), e.data(), nullptr);
}

// the nst lowest Ritz values and their tridiagonal eigenvectors (ys[k][j], j < m)
static int ritz_states(const LczState &s, int nst, std::vector<double> &vals,
                       std::vector<std::vector<double>> &ys) {
    const int m = s.steps;
    std::vector<double> d(s.alpha.begin(), s.alpha.begin() + m), e(std::max(m - 1, 1), 0.0);
    for (int i = 0; i + 1 < m; ++i) e[i] = s.beta[i + 1];
    std::vector<double> z((size_t)m * m);
    if (tridiag_ql(m, d.data(), e.data(), z.data())) return 1;
    nst = std::max(1, std::min(nst, m));
    std::vector<double> u1(m);
    for (int k = 0; k < m; ++k) u1[k] = z[(size_t)k * m];
    const std::vector<int> k_of = genuine_ritz(d.data(), u1.data(), m, nst, s.ghost_tol);
    vals.resize(nst);
    ys.assign(nst, std::vector<double>(m));
    for (int k = 0; k < nst; ++k) {
        vals[k] = d[k_of[k]];
        for (int j = 0; j < m; ++j) ys[k][j] = z[(size_t)k_of[k] * m + j];
    }
    return 0;
}

// Device memory for Krylov vectors: carved from a caller workspace while it lasts
// (256-byte aligned pieces), then allocated here and freed with the arena.
struct Arena {
    phub_ctx *c;
    char *base = nullptr;
    size_t size = 0, used = 0;
    std::vector<void *> owned;
    Arena(phub_ctx *ctx, void *ws, size_t bytes) : c(ctx) {
        if (ws) {
            const uintptr_t a = ((uintptr_t)ws + 255) & ~(uintptr_t)255;
            const size_t skip = a - (uintptr_t)ws;
            if (bytes > skip) {
                base = (char *)a;
                size = bytes - skip;
            }
        }
    }
    ~Arena() { release(); }
    Arena(const Arena &) = delete;
    Arena &operator=(const Arena &) = delete;
    size_t room() const {
        const size_t off = (used + 255) & ~(size_t)255;
        return base && off < size ? size - off : 0;
    }
    void *take(size_t n) {
        if (base && n <= room()) {
            const size_t off = (used + 255) & ~(size_t)255;
            used = off + n;
            return base + off;
        }
        void *p = nullptr;
        if (phub_malloc(&p, n) != hipSuccess) {
            (void)hipGetLastError();   // a failed allocation is not a sticky error here
            return nullptr;
        }
        owned.push_back(p);
        return p;
    }
    // every piece handed out so far becomes invalid
    void release() {
        for (void *p : owned) phub_dfree(c, p);
        owned.clear();
        used = 0;
    }
};

// Vectors of the plain recurrence, kept while device memory allows: the reference
// keeps every Lanczos vector (lanczos.f90:49) and rebuilds psi from them
// (compute_solver.f90:271-285).  slot[k] holds vector k while `growing`; chunks of
// up to 8 vectors are added on demand.  When a chunk cannot be allocated the store
// stops growing and the recurrence rotates three slots other than slot 0 (the
// start vector), so psi can still be rebuilt by a second pass.  Vectors come from
// an Arena: with a caller workspace the store never grows beyond it.
template <typename T>
struct KrylovStore {
    Arena &ar;
    std::vector<T *> slot;
    int64_t ld = 0;
    bool growing = false;
    bool keep_first = false;   // slot 0 (the start vector) is never overwritten
    int cap = 0;               // at most this many vectors (0: no limit)
    KrylovStore(Arena &a, int64_t n) : ar(a), ld(basis_ld(n)) {}
    bool add(int cnt) {
        T *p = (T *)ar.take((size_t)cnt * ld * sizeof(T));
        if (!p) return false;
        for (int i = 0; i < cnt; ++i) slot.push_back(p + (size_t)i * ld);
        return true;
    }
    bool grow() {
        if (cap > 0 && (int)slot.size() >= cap) return false;
        int cnt = cap > 0 ? std::min(8, cap - (int)slot.size()) : 8;
        if (ar.base) {
            // a caller workspace bounds the store: nothing is allocated for it
            const size_t fit = ar.room() / ((size_t)ld * sizeof(T));
            return fit >= 1 && add((int)std::min<size_t>(cnt, fit));
        }
        return add(cnt) || add(1);
    }
};

// plain three-term recurrence (no reorthogonalisation): three rotating buffers
// (v0t, b1, b2), or the vectors of a KrylovStore (st: v0t = st->slot[0])
template <typename T>
static int lanczos_plain(phub_ctx *c, const phub_op *op, const phub_lanczos_opts *o,
                         T *v0t, T *b1, T *b2, LczState &s, EvTimer *tm, bool fixed,
                         KrylovStore<T> *st = nullptr) {
    const int64_t n = op->dim;
    const int maxs = o->max_steps;
    const int every = o->check_every > 0 ? o->check_every : 5;
    s.alpha.assign(maxs + 2, 0.0);
    s.beta.assign(maxs + 2, 0.0);
    double n2 = 0.0;
    int rc = dev_dot<T, T>(c, v0t, v0t, n, &n2);
    if (rc) return rc;
    if (!(n2 > 0.0)) return phub_fail(c, PHUB_ERR_ARG, "start vector is zero");
    s.beta[0] = std::sqrt(n2);
    T *xp = b1, *xk = v0t, *out = b2;   // xp content irrelevant while beta_eff = 0
    int ip = -1, ik = 0;                 // slot indices of xp, xk (with a store)
    double sig = 1.0 / s.beta[0], sigp = 0.0;
    std::vector<double> ritz;
    // PHUB_PROGRESS=k: one line on stderr every k steps (long solves)
    const char *pe = std::getenv("PHUB_PROGRESS");
    const int progress = pe ? std::atoi(pe) : 0;
    const auto tp0 = std::chrono::steady_clock::now();
    while (!s.converged && !s.breakdown && s.steps < maxs) {
        int io = -1;
        if (st) {
            if (st->growing && s.steps + 1 >= (int)st->slot.size() && !st->grow())
                st->growing = false;
            if (st->growing) {
                io = s.steps + 1;
            } else {
                for (int q = st->keep_first ? 1 : 0; q < (int)st->slot.size() && io < 0; ++q)
                    if (q != ip && q != ik) io = q;
                if (io < 0) return phub_fail(c, PHUB_ERR_ARG, "Krylov store has too few slots");
            }
            out = st->slot[io];
        }
        double dots[2];
        rc = fused_hv<T>(c, op, xk, sig, s.steps ? xp : nullptr, s.beta[s.steps] * sigp, out,
                         dots, tm);
        if (rc) return rc;
        const double a = dots[0];
        s.alpha[s.steps] = a;
        double nn = 0.0;
        rc = dev_axpby<T, T>(c, -a * sig, xk, 1.0, out, n, &nn);
        if (rc) return rc;
        const double b = std::sqrt(std::max(nn, 0.0));
        s.steps++;
        s.beta[s.steps] = b;
        std::swap(xp, xk);
        std::swap(xk, out);
        ip = ik;
        ik = io;
        sigp = sig;
        sig = (b > 0.0) ? 1.0 / b : 0.0;
        if (!(b > 1e-13 * std::max(1.0, std::fabs(a)))) s.breakdown = true;
        if (progress > 0 && s.steps % progress == 0) {
            double e = 0.0;
            if (all_ritz(s.steps, s.alpha.data(), s.beta.data(), ritz) == 0) e = ritz[0];
            std::fprintf(stderr, "phub lanczos: step %d  lowest Ritz %.12f  %.1f s\n", s.steps, e,
                         std::chrono::duration<double>(std::chrono::steady_clock::now() - tp0).count());
            std::fflush(stderr);
        }
        if (!fixed && (s.steps % every == 0 || s.breakdown || s.steps == maxs)) {
            if (all_ritz(s.steps, s.alpha.data(), s.beta.data(), ritz))
                return phub_fail(c, PHUB_ERR_HIP, "tridiagonal QL failed");
            s.e0 = ritz[0];
            if (s.breakdown) { s.converged = true; break; }
            convergence(s, o->tol, maxs, ritz, o->n_states);
            residual_rule(s, o->psi_tol, sizeof(T) == 4);
        }
    }
    if (fixed || s.steps == 0 || !s.converged) {
        if (all_ritz(s.steps, s.alpha.data(), s.beta.data(), ritz) == 0 && s.steps) s.e0 = ritz[0];
    }
    return PHUB_OK;
}

// The three-term recurrence of lanczos_plain with each step captured once into a
// hipGraph and replayed: the step's scalars stay in device memory (GraphBufs) and
// alpha, beta go to the context's coherent pinned array, so a step is one graph
// launch and the host waits only at the convergence checks (every check_every
// steps).  Three graphs, one per rotation
// of the three work vectors r[0..2]; with a store, each step also copies its new
// vector into slots[k] (a device table of the store's slot pointers, every slot
// allocated before the run).  The arithmetic is that of lanczos_plain, so both give
// the same bits.  Steps past a breakdown inside a block of check_every steps are
// discarded at the check.  Used for the Hubbard operators (their fused epilogue reads
// device scalars); step 0 runs outside the graphs (it tunes the operator's tiles and
// sizes the reduction buffer).
struct GraphBufs {
    double *dsc;       // 8 scalars
    int *steps;        // 1
    void **slots;      // max_steps + 2 (store mode)
};

static bool graph_capable(const phub_op *op) {
    const char *e = std::getenv("PHUB_LANCZOS_GRAPH");   // 0: never, 1: any size
    if (e && e[0] == '0') return false;
    if (op->kind != 0 || op->d_ul) return false;
    return (e && e[0] == '1') || op->dim <= ((int64_t)1 << 24);
}

template <typename T>
static int lanczos_plain_graph(phub_ctx *c, const phub_op *op, const phub_lanczos_opts *o,
                               const T *start, T *const r[3], LczState &s, EvTimer *tm,
                               bool fixed, KrylovStore<T> *st, const GraphBufs &gb) {
    const int64_t n = op->dim;
    const int maxs = o->max_steps;
    const int every = o->check_every > 0 ? o->check_every : 5;
    const unsigned g = grid_for(n);
    s.alpha.assign(maxs + 2, 0.0);
    s.beta.assign(maxs + 2, 0.0);
    double n2 = 0.0;
    int rc = dev_dot<T, T>(c, start, start, n, &n2);
    if (rc) return rc;
    if (!(n2 > 0.0)) return phub_fail(c, PHUB_ERR_ARG, "start vector is zero");
    s.beta[0] = std::sqrt(n2);
    // step 0 as lanczos_plain takes it (outside the graphs: it also tunes the operator's
    // tiles and sizes the reduction buffer), into r[2] with r[0] = x_0
    PHUB_HIP(c, hipMemcpyAsync(r[0], start, n * sizeof(T), hipMemcpyDeviceToDevice, c->stream));
    double sig = 1.0 / s.beta[0];
    double dots[2];
    int nb = 0;
    {
        if (tm) tm->start();
        rc = phub_hv_launch<T>(c, op, r[0], r[2], sig, (const T *)nullptr, 0.0, true, &nb);
        if (rc) return rc;
        if (tm) tm->stop();
        rc = phub_reduce_partials(c, nb, 2, dots);
        if (rc) return rc;
    }
    const double a0 = dots[0];
    s.alpha[0] = a0;
    double nn = 0.0;
    rc = dev_axpby<T, T>(c, -a0 * sig, r[0], 1.0, r[2], n, &nn);
    if (rc) return rc;
    const double b1 = std::sqrt(std::max(nn, 0.0));
    s.steps = 1;
    s.beta[1] = b1;
    if (st) {
        PHUB_HIP(c, hipMemcpyAsync(st->slot[1], r[2], n * sizeof(T), hipMemcpyDeviceToDevice,
                                   c->stream));
    }
    if (!(b1 > 1e-13 * std::max(1.0, std::fabs(a0)))) s.breakdown = true;
    rc = phub_ensure_red(c, (size_t)std::max<int64_t>(2 * (int64_t)nb, g) + 64);
    if (rc) return rc;
    // alpha and beta go straight to coherent pinned host memory: a check only waits
    if (c->h_lcz_cap < 2 * (size_t)(maxs + 2)) {
        if (c->h_lcz) hipHostFree(c->h_lcz);
        c->h_lcz = nullptr;
        c->h_lcz_cap = 0;
        PHUB_HIP(c, phub_host_malloc(&c->h_lcz, 2 * (size_t)(maxs + 2) * sizeof(double),
                                     hipHostMallocCoherent));
        c->h_lcz_cap = 2 * (size_t)(maxs + 2);
    }
    double *h_alpha = c->h_lcz, *h_beta = c->h_lcz + maxs + 2;
    // device scalars after step 0: scale 1/beta_1, beta_1 * scale of x_0
    const double sig1 = (b1 > 0.0) ? 1.0 / b1 : 0.0;
    const double init[4] = {sig1, b1 * sig, 0.0, 0.0};
    PHUB_HIP(c, hipMemcpyAsync(gb.dsc, init, sizeof(init), hipMemcpyHostToDevice, c->stream));
    h_alpha[0] = s.alpha[0];
    h_beta[0] = s.beta[0];
    h_beta[1] = s.beta[1];
    const int one = 1;
    PHUB_HIP(c, hipMemcpyAsync(gb.steps, &one, sizeof(int), hipMemcpyHostToDevice, c->stream));
    if (st) {
        PHUB_HIP(c, hipMemcpyAsync(gb.slots, st->slot.data(), st->slot.size() * sizeof(void *),
                                   hipMemcpyHostToDevice, c->stream));
    }
    PHUB_HIP(c, hipStreamSynchronize(c->stream));
    if (!c->capture_stream)
        PHUB_HIP(c, hipStreamCreateWithFlags(&c->capture_stream, hipStreamNonBlocking));
    // (x_k, x_{k-1}, out) of step k: rotation phase k % 3
    const int rot[3][3] = {{0, 1, 2}, {2, 0, 1}, {1, 2, 0}};
    hipGraphExec_t exec[3] = {nullptr, nullptr, nullptr};
    struct Cleanup {
        hipGraphExec_t *e;
        ~Cleanup() {
            for (int i = 0; i < 3; ++i)
                if (e[i]) hipGraphExecDestroy(e[i]);
        }
    } cleanup{exec};
    hipStream_t user = c->stream;
    for (int ph = 0; ph < 3; ++ph) {
        T *xk = r[rot[ph][0]], *xp = r[rot[ph][1]], *out = r[rot[ph][2]];
        c->stream = c->capture_stream;
        c->dev_scalars = gb.dsc;
        hipError_t e = hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal);
        int nb2 = 0;
        if (e == hipSuccess) {
            rc = phub_hv_launch<T>(c, op, xk, out, 1.0, xp, 0.0, true, &nb2);
            if (!rc) rc = phub_reduce_partials_dev(c, nb2, 2, gb.dsc + 2);
            if (!rc) {
                hipLaunchKernelGGL((k_axpby<T, T>), dim3(g), dim3(256), 0, c->stream, 0.0,
                                   (const T *)xk, 1.0, out, n, c->d_red, (const double *)gb.dsc);
                rc = phub_reduce_partials_dev(c, (int)g, 1, gb.dsc + 3);
            }
            if (!rc) {
                hipLaunchKernelGGL(k_lcz_step_end, dim3(1), dim3(1), 0, c->stream, gb.dsc, h_alpha,
                                   h_beta, gb.steps);
                if (st)
                    hipLaunchKernelGGL(k_store_vec<T>, dim3(g), dim3(256), 0, c->stream, out,
                                       (T *const *)gb.slots, gb.steps, n);
            }
        }
        hipGraph_t graph = nullptr;
        const hipError_t ee = hipStreamEndCapture(c->stream, &graph);
        c->stream = user;
        c->dev_scalars = nullptr;
        if (e != hipSuccess) return phub_hip_fail(c, e, "hipStreamBeginCapture");
        if (rc) {
            if (graph) hipGraphDestroy(graph);
            return rc;
        }
        if (ee != hipSuccess) return phub_hip_fail(c, ee, "hipStreamEndCapture");
        if (nb2 != nb) {
            hipGraphDestroy(graph);
            return phub_fail(c, PHUB_ERR_HIP, "captured H*v changed its grid");
        }
        e = hipGraphInstantiate(&exec[ph], graph, nullptr, nullptr, 0);
        hipGraphDestroy(graph);
        if (e != hipSuccess) return phub_hip_fail(c, e, "hipGraphInstantiate");
    }
    std::vector<double> ritz;
    int launched = 1;
    if (!fixed && s.breakdown) {
        if (all_ritz(s.steps, s.alpha.data(), s.beta.data(), ritz))
            return phub_fail(c, PHUB_ERR_HIP, "tridiagonal QL failed");
        s.e0 = ritz[0];
        s.converged = true;
    }
    while (!s.converged && !s.breakdown && launched < maxs) {
        const int blk = std::min(every - launched % every, maxs - launched);
        for (int q = 0; q < blk; ++q, ++launched)
            PHUB_HIP(c, hipGraphLaunch(exec[launched % 3], c->stream));
        PHUB_HIP(c, hipStreamSynchronize(c->stream));
        std::memcpy(&s.alpha[0], h_alpha, launched * sizeof(double));
        std::memcpy(&s.beta[1], h_beta + 1, launched * sizeof(double));
        // the step count of the host-driven loop: up to the first breakdown
        int k = s.steps;
        while (k < launched) {
            ++k;
            if (!(s.beta[k] > 1e-13 * std::max(1.0, std::fabs(s.alpha[k - 1])))) {
                s.breakdown = true;
                break;
            }
        }
        s.steps = k;
        if (!fixed && (s.steps % every == 0 || s.breakdown || s.steps == maxs)) {
            if (all_ritz(s.steps, s.alpha.data(), s.beta.data(), ritz))
                return phub_fail(c, PHUB_ERR_HIP, "tridiagonal QL failed");
            s.e0 = ritz[0];
            if (s.breakdown) { s.converged = true; break; }
            convergence(s, o->tol, maxs, ritz, o->n_states);
            residual_rule(s, o->psi_tol, sizeof(T) == 4);
        }
    }
    for (int k = s.steps + 1; k < maxs + 2; ++k) s.beta[k] = 0.0;
    for (int k = s.steps; k < maxs + 2; ++k) s.alpha[k] = 0.0;
    if (fixed || s.steps == 0 || !s.converged) {
        if (all_ritz(s.steps, s.alpha.data(), s.beta.data(), ritz) == 0 && s.steps) s.e0 = ritz[0];
    }
    return PHUB_OK;
}

// second pass of the plain recurrence to rebuild psi = sum_k y_k x_k
template <typename T>
static int plain_second_pass(phub_ctx *c, const phub_op *op, const T *v0t, T *b1, T *b2, T *b3,
                             const LczState &s, const std::vector<std::vector<double>> &ys,
                             double *psi) {
    // psi_k = sum_j ys[k][j] x_j for every requested state k (state k at psi + k*n)
    const int64_t n = op->dim;
    const unsigned g = grid_for(n);
    const int nst = (int)ys.size();
    PHUB_HIP(c, hipMemsetAsync(psi, 0, (size_t)nst * n * sizeof(double), c->stream));
    // x_0 = v0 / beta0 in b1
    hipLaunchKernelGGL((k_axpby<T, T>), dim3(g), dim3(256), 0, c->stream, 1.0 / s.beta[0], v0t,
                       0.0, b1, n, (double *)nullptr);
    for (int k = 0; k < nst; ++k)
        hipLaunchKernelGGL((k_axpby<T, double>), dim3(g), dim3(256), 0, c->stream, ys[k][0], b1,
                           1.0, psi + (size_t)k * n, n, (double *)nullptr);
    T *xp = b2, *xk = b1, *out = b3;
    const int m = (int)ys[0].size();
    for (int k = 0; k + 1 < m; ++k) {
        int rc = phub_hv_launch<T>(c, op, xk, out, 1.0, k ? xp : nullptr, s.beta[k], false, nullptr);
        if (rc) return rc;
        hipLaunchKernelGGL((k_second_pass<T>), dim3(g), dim3(256), 0, c->stream, out, xk,
                           s.alpha[k], 1.0 / s.beta[k + 1], psi, ys[0][k + 1], n);
        for (int q = 1; q < nst; ++q)
            hipLaunchKernelGGL((k_axpby<T, double>), dim3(g), dim3(256), 0, c->stream,
                               ys[q][k + 1], out, 1.0, psi + (size_t)q * n, n, (double *)nullptr);
        PHUB_HIP(c, hipGetLastError());
        std::swap(xp, xk);
        std::swap(xk, out);
    }
    return PHUB_OK;
}

// normalise the reconstructed states; the ground state's drift is the reference's
// NORM LOST check (lanczos.f90:74-76)
static int normalise_states(phub_ctx *c, double *psi, int nst, int64_t n, LczState &s) {
    for (int k = 0; k < nst; ++k) {
        double *pk = psi + (size_t)k * n;
        double nn = 0;
        int rc = dev_dot<double, double>(c, pk, pk, n, &nn);
        if (rc) return rc;
        if (k == 0 && std::fabs(1.0 - std::sqrt(nn)) > 1e-5) s.norm_lost = true;
        rc = dev_axpby<double, double>(c, 0.0, (const double *)nullptr, 1.0 / std::sqrt(nn), pk,
                                       n, nullptr);
        if (rc) return rc;
    }
    return PHUB_OK;
}

// Lanczos with a stored basis and full reorthogonalisation (classical Gram-Schmidt,
// repeated once when the norm drops by more than 1/sqrt(2)).
template <typename T>
static int lanczos_full(phub_ctx *c, const phub_op *op, const phub_lanczos_opts *o, T *basis,
                        int64_t ld, LczState &s, std::vector<double> &sig, double *d_coef,
                        EvTimer *tm, bool fixed) {
    const int64_t n = op->dim;
    const int maxs = o->max_steps;
    const int every = o->check_every > 0 ? o->check_every : 5;
    s.alpha.assign(maxs + 2, 0.0);
    s.beta.assign(maxs + 2, 0.0);
    sig.assign(maxs + 2, 0.0);
    double n2 = 0.0;
    int rc = dev_dot<T, T>(c, basis, basis, n, &n2);
    if (rc) return rc;
    if (!(n2 > 0.0)) return phub_fail(c, PHUB_ERR_ARG, "start vector is zero");
    s.beta[0] = std::sqrt(n2);
    sig[0] = 1.0 / s.beta[0];
    std::vector<double> cf(maxs + 2), ritz;
    while (!s.converged && !s.breakdown && s.steps < maxs) {
        const int k = s.steps;
        T *xk = basis + (size_t)k * ld;
        T *out = basis + (size_t)(k + 1) * ld;
        const T *xp = k ? basis + (size_t)(k - 1) * ld : nullptr;
        double dots[2];
        rc = fused_hv<T>(c, op, xk, sig[k], xp, k ? s.beta[k] * sig[k - 1] : 0.0, out, dots, tm);
        if (rc) return rc;
        // alpha from the fused pass; global coefficients on (out - alpha x_k)
        const double a0 = dots[0];
        double nbefore = dots[1], nafter = 0.0;
        rc = dev_multidot<T, T>(c, basis, ld, k + 1, out, n, cf.data(), xk, a0 * sig[k]);
        if (rc) return rc;
        for (int j = 0; j <= k; ++j) cf[j] *= sig[j];   // <x_j, out - a0 x_k>
        s.alpha[k] = a0 + cf[k];
        // norm left after the local three-term part: the DGKS test compares against it
        const double nlocal = std::max(nbefore - a0 * a0, 0.0);
        for (int j = 0; j <= k; ++j) cf[j] *= sig[j];   // coefficients on xt_j
        cf[k] += a0 * sig[k];
        rc = dev_combine<T, T>(c, basis, ld, k + 1, cf.data(), out, n, false, &nafter, d_coef);
        if (rc) return rc;
        if (nafter < 0.25 * nlocal) {
            double nb2 = nafter;
            rc = dev_multidot<T, T>(c, basis, ld, k + 1, out, n, cf.data());
            if (rc) return rc;
            for (int j = 0; j <= k; ++j) cf[j] *= sig[j];
            s.alpha[k] += cf[k];
            for (int j = 0; j <= k; ++j) cf[j] *= sig[j];
            rc = dev_combine<T, T>(c, basis, ld, k + 1, cf.data(), out, n, false, &nafter, d_coef);
            if (rc) return rc;
            (void)nb2;
        }
        const double b = std::sqrt(std::max(nafter, 0.0));
        s.steps++;
        s.beta[s.steps] = b;
        sig[s.steps] = (b > 0.0) ? 1.0 / b : 0.0;
        if (!(b > 1e-13 * std::max(1.0, std::fabs(s.alpha[k])))) s.breakdown = true;
        if (!fixed && (s.steps % every == 0 || s.breakdown || s.steps == maxs)) {
            if (all_ritz(s.steps, s.alpha.data(), s.beta.data(), ritz))
                return phub_fail(c, PHUB_ERR_HIP, "tridiagonal QL failed");
            s.e0 = ritz[0];
            if (s.breakdown) { s.converged = true; break; }
            convergence(s, o->tol, maxs, ritz, o->n_states);
            residual_rule(s, o->psi_tol, sizeof(T) == 4);
        }
    }
    if (s.steps && all_ritz(s.steps, s.alpha.data(), s.beta.data(), ritz) == 0) s.e0 = ritz[0];
    return PHUB_OK;
}

// E = <psi|H|psi>, residual |H psi - E psi| (psi normalised), FP64; tmp: n doubles
static int rq_residual(phub_ctx *c, const phub_op *op, const double *psi, double *tmp,
                       double *e, double *res) {
    const int64_t n = op->dim;
    int rc = phub_hv_launch<double>(c, op, psi, tmp, 1.0, (const double *)nullptr, 0.0, false,
                                    nullptr);
    if (rc) return rc;
    double pp = 0, ph = 0;
    rc = dev_dot<double, double>(c, psi, psi, n, &pp);
    if (!rc) rc = dev_dot<double, double>(c, psi, tmp, n, &ph);
    if (rc) return rc;
    const double E = ph / pp;
    double r2 = 0;
    rc = dev_axpby<double, double>(c, -E, psi, 1.0, tmp, n, &r2);
    if (rc) return rc;
    *e = E;
    *res = std::sqrt(r2 / pp);
    return PHUB_OK;
}

// Jacobi-Davidson corrections of a normalised FP64 psi (Sleijpen and van der Vorst,
// SIAM J. Matrix Anal. Appl. 17, 401 (1996)): with E = <psi|H|psi> and
// r = H psi - E psi (orthogonal to psi), both in FP64, solve the correction equation
// (H - E) d = -r on the complement of psi loosely, by MINRES with FP32 vectors and
// FP32 H*v, and take psi <- (psi + d) / |psi + d|.  The residual falls by about the
// inner tolerance per step, so that tolerance aims at psi_tol in one step
// (between 1e-5 and 1e-1 relative; at most 120 inner iterations).  At most rs
// steps, fewer once |r| <= psi_tol.  On return e0 is the Rayleigh quotient of psi
// and res its residual.
template <typename V>
static int minres_core(phub_ctx *c, const phub_op *op, const V *psi, double e0, const V *b,
                       V *x, int max_iter, double tol, int *iters, double *rel_res, Arena &ar,
                       const V *minv);

__global__ __launch_bounds__(256) void k_jd_rhs(const double *__restrict__ r, float *__restrict__ b,
                                                int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        b[i] = (float)(-r[i]);
}

// The MINRES solves on (H - e) (the corrections of psi, the adjoint solve) take the
// Jacobi preconditioner 1 / (diag(H) - e) on one-body Hubbard operators: their diagonal,
// U times the double occupancy, spreads the spectrum of H - e far more than the hops
// do.  PHUB_JD_PRECOND=0 solves unpreconditioned.
bool phub_minres_precond(const phub_op *op) {
    const char *pe = std::getenv("PHUB_JD_PRECOND");
    return op->kind == 0 && !op->d_ul && !(pe && pe[0] == '0');
}

static int jd_refine(phub_ctx *c, const phub_op *op, double *psi, Arena &ar, int rs,
                     double psi_tol, double *e0, double *res, phub_lanczos_result *out) {
    const int64_t n = op->dim;
    const unsigned g = grid_for(n);
    double *tmp = (double *)ar.take(n * sizeof(double));
    float *b32 = (float *)ar.take(n * sizeof(float));
    float *x32 = (float *)ar.take(n * sizeof(float));
    if (!tmp || !b32 || !x32) return phub_fail(c, PHUB_ERR_OOM, "FP64 correction vectors do not fit");
    // Jacobi preconditioner of the inner solves on one-body Hubbard operators
    float *psi32 = (float *)ar.take(n * sizeof(float));
    if (!psi32) return phub_fail(c, PHUB_ERR_OOM, "FP64 correction vectors do not fit");
    float *minv = nullptr;
    if (phub_minres_precond(op)) {
        minv = (float *)ar.take(n * sizeof(float));
        if (!minv) return phub_fail(c, PHUB_ERR_OOM, "FP64 correction vectors do not fit");
    }
    int rc = rq_residual(c, op, psi, tmp, e0, res);
    if (rc) return rc;
    for (int it = 0; it < rs && !(psi_tol > 0.0 && *res <= psi_tol); ++it) {
        // tmp holds r = H psi - E psi; the right-hand side is -r
        hipLaunchKernelGGL(k_jd_rhs, dim3(g), dim3(256), 0, c->stream, tmp, b32, n);
        PHUB_HIP(c, hipGetLastError());
        const double tol_in = psi_tol > 0.0 ? std::min(1e-1, std::max(1e-5, 0.3 * psi_tol / *res))
                                            : 1e-2;
        int iters = 0;
        double rr = 0.0;
        if (minv) {
            rc = phub_hub_minv(c, op, *e0, minv);
            if (rc) return rc;
        }
        // the inner solve projects with an FP32 copy of psi: a psi component it leaves
        // in the correction only rescales psi + d, which is normalised below
        hipLaunchKernelGGL(k_convert<float>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream,
                           psi, psi32, n, 1.0);   // one element per thread: a full grid
        PHUB_HIP(c, hipGetLastError());
        const size_t mark = ar.used;
        rc = minres_core<float>(c, op, (const float *)psi32, *e0, b32, x32, 120, tol_in, &iters,
                                &rr, ar, (const float *)minv);
        ar.used = mark;   // the seven MINRES vectors are free again
        if (rc) return rc;
        out->corrections += 1;
        out->correction_iters += iters;
        double nn = 0.0;
        rc = dev_axpby<float, double>(c, 1.0, x32, 1.0, psi, n, &nn);
        if (rc) return rc;
        rc = dev_axpby<double, double>(c, 0.0, (const double *)nullptr, 1.0 / std::sqrt(nn), psi, n,
                                       nullptr);
        if (rc) return rc;
        rc = rq_residual(c, op, psi, tmp, e0, res);
        if (rc) return rc;
    }
    return PHUB_OK;
}

template <typename T>
struct DevBuf {
    phub_ctx *c;
    T *p = nullptr;
    explicit DevBuf(phub_ctx *ctx) : c(ctx) {}
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { phub_dfree(c, p); }
};

// The graph buffers inside the Lanczos arena block (3 max_steps + 40 doubles):
// coefficients [0, max_steps + 8), psi pointers [max_steps + 8, 2 max_steps + 9),
// then the scalars, the step counter and the slot table.
static GraphBufs graph_bufs(double *coef, int max_steps) {
    double *p = coef + 2 * max_steps + 16;
    GraphBufs g;
    g.dsc = p;
    g.steps = (int *)(p + 8);
    g.slots = (void **)(p + 10);
    return g;
}

// plain recurrence with 3 resident vectors of type T (FP64 scalars); psi by a
// second pass from a kept copy of the start vector
template <typename T>
static int plain_mode(phub_ctx *c, const phub_op *op, const phub_lanczos_opts *o,
                      const phub_lanczos_opts *oo, const double *d_v0, double *d_psi,
                      LczState &s, EvTimer *tm, bool fixed, phub_lanczos_result *res,
                      std::chrono::steady_clock::time_point t0, double *t_lcz, double *coef,
                      Arena &ar) {
    const int64_t n = op->dim;
    const unsigned g = (unsigned)((n + 255) / 256);
    // with psi wanted the vectors are kept (slot 0 = start vector survives for a
    // second pass if the store stops growing); otherwise three rotating slots
    KrylovStore<T> st(ar, n);
    if (!st.add(o->want_psi ? 4 : 3))
        return phub_fail(c, PHUB_ERR_OOM, "Lanczos vectors do not fit in device memory");
    st.growing = st.keep_first = o->want_psi != 0;
    // PHUB_LANCZOS_KEEP=k caps the vectors kept (device memory); psi then comes from a
    // second pass of the recurrence
    if (const char *ek = std::getenv("PHUB_LANCZOS_KEEP")) st.cap = std::max(4, std::atoi(ek));
    if (d_v0)
        hipLaunchKernelGGL(k_convert<T>, dim3(g), dim3(256), 0, c->stream, d_v0, st.slot[0], n, 1.0);
    else
        hipLaunchKernelGGL(k_randvec<T>, dim3(g), dim3(256), 0, c->stream, st.slot[0], n, o->seed);
    PHUB_HIP(c, hipGetLastError());
    int rc = PHUB_OK;
    // graph-captured steps where the operator allows them and every vector the run
    // may keep fits (the store cannot grow between graph launches)
    bool graphed = false;
    if (graph_capable(op) && !st.cap) {
        T *r[3];
        bool ok = true;
        for (int i = 0; i < 3 && ok; ++i) ok = (r[i] = (T *)ar.take(n * sizeof(T))) != nullptr;
        if (ok && st.growing)
            while ((int)st.slot.size() < oo->max_steps + 1 && st.grow()) {}
        if (ok && (!st.growing || (int)st.slot.size() >= oo->max_steps + 1)) {
            rc = lanczos_plain_graph<T>(c, op, oo, st.slot[0], r, s, tm, fixed,
                                        st.growing ? &st : nullptr, graph_bufs(coef, oo->max_steps));
            if (rc) return rc;
            graphed = true;
        }
    }
    if (!graphed) {
        rc = lanczos_plain<T>(c, op, oo, st.slot[0], st.slot[1], st.slot[2], s, tm, fixed, &st);
        if (rc) return rc;
    }
    *t_lcz = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    res->e0_ritz = res->e0 = s.e0;
    if (o->want_psi) {
        std::vector<double> vals;
        std::vector<std::vector<double>> ys;
        if (ritz_states(s, std::max(1, o->n_states), vals, ys))
            return phub_fail(c, PHUB_ERR_HIP, "tridiagonal QL failed");
        if (st.growing) {
            // every vector is stored: psi_q = sum_j y_qj x_j / beta_j in one pass each
            // the pointer table sits in the coefficient block, after its max_steps + 8
            struct { const T **p; } dptr{(const T **)(coef + oo->max_steps + 8)};
            std::vector<double> cf(s.steps);
            for (size_t q = 0; q < ys.size(); ++q) {
                for (int j = 0; j < s.steps; ++j)
                    cf[j] = s.beta[j] > 0.0 ? ys[q][j] / s.beta[j] : 0.0;
                rc = dev_combine_ptrs<T, double>(c, st.slot.data(), s.steps, cf.data(),
                                                 d_psi + q * n, n, coef, dptr.p);
                if (rc) return rc;
            }
        } else {
            rc = plain_second_pass<T>(c, op, st.slot[0], st.slot[1], st.slot[2], st.slot[3], s,
                                      ys, d_psi);
            if (rc) return rc;
        }
        if (sizeof(T) == 4) {
            LczState tmp = s;   // FP32 vectors drift more: no norm check here
            rc = normalise_states(c, d_psi, (int)ys.size(), n, tmp);
        } else {
            rc = normalise_states(c, d_psi, (int)ys.size(), n, s);
        }
        if (rc) return rc;
        const int rs = sizeof(T) == 4 ? std::max(0, std::min(o->refine_steps, 8)) : 0;
        // the Krylov vectors are no longer needed: their memory serves what follows
        PHUB_HIP(c, hipStreamSynchronize(c->stream));
        st.slot.clear();
        ar.release();
        if (rs > 0) {
            // FP32 recurrence, then FP64 corrections of the ground state (as MIXED)
            rc = jd_refine(c, op, d_psi, ar, rs, o->psi_tol, &res->e0, &res->residual, res);
            if (rc) return rc;
            PHUB_HIP(c, hipStreamSynchronize(c->stream));
            res->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            return PHUB_OK;
        }
        res->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        double *tmp = (double *)ar.take(n * sizeof(double));
        if (!tmp) return phub_fail(c, PHUB_ERR_OOM, "residual vector does not fit");
        double erq = 0;   // e0 stays the Ritz value, as the reference reports it
        rc = rq_residual(c, op, d_psi, tmp, &erq, &res->residual);
        if (rc) return rc;
    }
    return PHUB_OK;
}

extern "C" int phub_lanczos_workspace_bytes(const phub_op *op, const phub_lanczos_opts *o,
                                            size_t *bytes) {
    if (!op || !o || !bytes) return PHUB_ERR_ARG;
    const int64_t n = op->dim;
    const size_t ld = (size_t)basis_ld(n);
    int steps = o->fixed_steps > 0 ? o->fixed_steps : o->max_steps;
    steps = (int)std::min<int64_t>(std::max(steps, 1), std::max<int64_t>(n, 1));
    const size_t m1 = (size_t)steps + 1;
    // jd_refine: one FP64 vector, eleven FP32 ones (the preconditioner and the FP32
    // copy of psi included)
    const size_t refine = ((ld * sizeof(double) + 255) & ~(size_t)255) +
                          11 * ((ld * sizeof(float) + 255) & ~(size_t)255) + 512;
    size_t b = 0;
    switch (o->mode) {
    case PHUB_LCZ_FULL64: b = m1 * ld * sizeof(double); break;
    case PHUB_LCZ_MIXED:
        b = std::max(m1 * ld * sizeof(float), (o->want_psi || o->refine_steps > 0) ? refine : 0);
        break;
    // the plain modes: the kept vectors, plus three work vectors when the steps run as
    // graphs (graph_capable)
    case PHUB_LCZ_PLAIN64:
        b = (o->want_psi ? std::max(m1, (size_t)4) : 3) * ld * sizeof(double) +
            (graph_capable(op) ? 3 * ld * sizeof(double) + 768 : 0);
        break;
    case PHUB_LCZ_PLAIN32:
        b = o->want_psi ? std::max(std::max(m1, (size_t)4) * ld * sizeof(float),
                                   o->refine_steps > 0 ? refine : ld * sizeof(double))
                        : 3 * ld * sizeof(float);
        b += graph_capable(op) ? 3 * ld * sizeof(float) + 768 : 0;
        break;
    default: return PHUB_ERR_ARG;
    }
    *bytes = b + 256;   // alignment slack
    return PHUB_OK;
}

extern "C" int phub_lanczos(phub_ctx *c, const phub_op *op, const phub_lanczos_opts *o,
                            const double *d_v0, double *alpha, double *beta,
                            phub_lanczos_result *res, double *d_psi) {
    if (!c || !op || !o || !res) return phub_fail(c, PHUB_ERR_ARG, "null argument");
    if (o->max_steps < 1) return phub_fail(c, PHUB_ERR_ARG, "max_steps must be >= 1");
    if (o->mode < 0 || o->mode > 3) return phub_fail(c, PHUB_ERR_ARG, "unknown Lanczos mode");
    if (o->want_psi && !d_psi) return phub_fail(c, PHUB_ERR_ARG, "want_psi needs d_psi");
    if (o->n_states < 0 || o->n_states > PHUB_MAX_STATES)
        return phub_fail(c, PHUB_ERR_ARG, "n_states must be in [0, PHUB_MAX_STATES]");
    std::memset(res, 0, sizeof(*res));
    PHUB_HIP(c, hipSetDevice(c->device));
    const int64_t n = op->dim;
    if (n == 0) return phub_fail(c, PHUB_ERR_ARG, "empty sector");
    const bool fixed = o->fixed_steps > 0;
    phub_lanczos_opts oo = *o;
    if (fixed) oo.max_steps = o->fixed_steps;
    oo.max_steps = (int)std::min<int64_t>(oo.max_steps, std::max<int64_t>(n, 1));
    const unsigned g = (unsigned)((n + 255) / 256);
    auto t0 = std::chrono::steady_clock::now();
    EvTimer tm(c);
    LczState s;
    s.ghost_tol = o->mode == PHUB_LCZ_PLAIN64 ? 1e-8 : o->mode == PHUB_LCZ_PLAIN32 ? 1e-5 : 0.0;
    // combine coefficients (max_steps + 8 doubles), the table of basis-vector pointers
    // of the one-pass psi (max_steps + 1) and the buffers of the graph-captured plain
    // recurrence (graph_bufs): the context arena
    struct { double *p; } coef{nullptr};
    int rc = phub_arena(c, (size_t)(3 * oo.max_steps + 40) * sizeof(double), (void **)&coef.p);
    if (rc) return rc;
    double t_lcz = 0.0;
    Arena ar(c, o->workspace, o->workspace_bytes);

    if (o->mode == PHUB_LCZ_PLAIN64) {
        rc = plain_mode<double>(c, op, o, &oo, d_v0, d_psi, s, &tm, fixed, res, t0, &t_lcz, coef.p,
                                ar);
        if (rc) return rc;
    } else if (o->mode == PHUB_LCZ_FULL64) {
        const int64_t ld = basis_ld(n);
        size_t bytes = (size_t)(oo.max_steps + 1) * ld * sizeof(double);
        double *basis = (double *)ar.take(bytes);
        if (!basis)
            return phub_fail(c, PHUB_ERR_OOM, "FP64 Krylov basis does not fit: " +
                                                  std::to_string(bytes / 1e9) + " GB requested");
        if (d_v0)
            PHUB_HIP(c, hipMemcpyAsync(basis, d_v0, n * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
        else
            hipLaunchKernelGGL(k_randvec<double>, dim3(g), dim3(256), 0, c->stream, basis, n, o->seed);
        std::vector<double> sig;
        rc = lanczos_full<double>(c, op, &oo, basis, ld, s, sig, coef.p, &tm, fixed);
        if (rc) return rc;
        t_lcz = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        res->e0_ritz = res->e0 = s.e0;
        if (o->want_psi) {
            std::vector<double> vals;
            std::vector<std::vector<double>> ys;
            if (ritz_states(s, std::max(1, o->n_states), vals, ys))
                return phub_fail(c, PHUB_ERR_HIP, "tridiagonal QL failed");
            for (size_t k = 0; k < ys.size(); ++k) {
                std::vector<double> cf(s.steps);
                for (int j = 0; j < s.steps; ++j) cf[j] = ys[k][j] * sig[j];
                double nn = 0;
                rc = dev_combine<double, double>(c, basis, ld, s.steps, cf.data(), d_psi + k * n,
                                                 n, true, &nn, coef.p);
                if (rc) return rc;
            }
            rc = normalise_states(c, d_psi, (int)ys.size(), n, s);
            if (rc) return rc;
            PHUB_HIP(c, hipStreamSynchronize(c->stream));
            res->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            double E = 0;
            rc = rq_residual(c, op, d_psi, basis, &E, &res->residual);
            if (rc) return rc;
        }
    } else if (o->mode == PHUB_LCZ_PLAIN32) {
        rc = plain_mode<float>(c, op, o, &oo, d_v0, d_psi, s, &tm, fixed, res, t0, &t_lcz, coef.p,
                               ar);
        if (rc) return rc;
    } else {   // MIXED
        const int64_t ld = basis_ld(n);
        size_t bytes = (size_t)(oo.max_steps + 1) * ld * sizeof(float);
        float *basis = (float *)ar.take(bytes);
        if (!basis)
            return phub_fail(c, PHUB_ERR_OOM, "FP32 Krylov basis does not fit: " +
                                                  std::to_string(bytes / 1e9) + " GB requested");
        if (d_v0)
            hipLaunchKernelGGL(k_convert<float>, dim3(g), dim3(256), 0, c->stream, d_v0, basis, n, 1.0);
        else
            hipLaunchKernelGGL(k_randvec<float>, dim3(g), dim3(256), 0, c->stream, basis, n, o->seed);
        std::vector<double> sig;
        rc = lanczos_full<float>(c, op, &oo, basis, ld, s, sig, coef.p, &tm, fixed);
        if (rc) return rc;
        t_lcz = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        res->e0_ritz = res->e0 = s.e0;
        if (o->want_psi || o->refine_steps > 0) {
            if (!d_psi) return phub_fail(c, PHUB_ERR_ARG, "refinement needs d_psi");
            std::vector<double> vals;
            std::vector<std::vector<double>> ys;
            if (ritz_states(s, std::max(1, o->n_states), vals, ys))
                return phub_fail(c, PHUB_ERR_HIP, "tridiagonal QL failed");
            for (size_t k = 0; k < ys.size(); ++k) {
                std::vector<double> cf(s.steps);
                for (int j = 0; j < s.steps; ++j) cf[j] = ys[k][j] * sig[j];
                double nn = 0;
                rc = dev_combine<float, double>(c, basis, ld, s.steps, cf.data(), d_psi + k * n,
                                                n, true, &nn, coef.p);
                if (rc) return rc;
            }
            {
                LczState tmp = s;   // the FP32 basis drifts more: no norm check here
                rc = normalise_states(c, d_psi, (int)ys.size(), n, tmp);
                if (rc) return rc;
            }
            // the FP32 basis is no longer needed: its memory serves the FP64 work
            PHUB_HIP(c, hipStreamSynchronize(c->stream));
            ar.release();
            const int rs = std::max(0, std::min(o->refine_steps, 8));
            rc = jd_refine(c, op, d_psi, ar, rs, o->psi_tol, &res->e0, &res->residual, res);
            if (rc) return rc;
            PHUB_HIP(c, hipStreamSynchronize(c->stream));
            res->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        }
    }
    PHUB_HIP(c, hipStreamSynchronize(c->stream));
    if (!o->want_psi && !(o->mode == PHUB_LCZ_MIXED && o->refine_steps > 0))
        res->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    res->seconds_lanczos = t_lcz;
    {
        std::vector<double> rv;
        std::vector<double> u1;
        if (s.steps > 0 && ritz_first(s.steps, s.alpha.data(), s.beta.data(), rv, u1) == 0) {
            const std::vector<int> k_of = genuine_ritz(rv.data(), u1.data(), (int)rv.size(),
                                                       std::min(PHUB_MAX_STATES, s.steps), s.ghost_tol);
            for (size_t k = 0; k < k_of.size(); ++k) res->e_states[k] = rv[k_of[k]];
        }
        if (o->mode == PHUB_LCZ_MIXED || (o->mode == PHUB_LCZ_PLAIN32 && o->want_psi &&
                                          o->refine_steps > 0))
            res->e_states[0] = res->e0;
    }
    res->steps = s.steps;
    res->converged = (s.converged || fixed) ? 1 : 0;
    res->ndeg = s.ndeg;
    res->hv_ms_avg = tm.n ? tm.total_ms / tm.n : 0.0;
    if (alpha)
        for (int i = 0; i < s.steps; ++i) alpha[i] = s.alpha[i];
    if (beta)
        for (int i = 0; i <= s.steps; ++i) beta[i] = s.beta[i];
    res->norm_lost = s.norm_lost ? 1 : 0;
    if (!res->converged)
        return phub_fail(c, PHUB_ERR_NOCONV, "Lanczos reached max_steps without convergence");
    return PHUB_OK;
}

// ---------------------------------------------------------------------------
// continued-fraction chain (single-vector GF_BANDLANCZOS, spgf_bandlanczos.f90:309-448)
// ---------------------------------------------------------------------------
template <typename T>
static int cf_chain_t(phub_ctx *c, const phub_op *op, const double *start, int nsteps,
                      double *alpha, double *beta, double *norm2, void *ws, size_t ws_bytes) {
    const int64_t n = op->dim;
    const unsigned g = (unsigned)((n + 255) / 256);
    // three vectors: from the caller's workspace, else allocated for the call
    Arena ar(c, ws, ws_bytes);
    struct { T *p; } v0{(T *)ar.take(n * sizeof(T))}, b1{(T *)ar.take(n * sizeof(T))},
        b2{(T *)ar.take(n * sizeof(T))};
    if (!v0.p || !b1.p || !b2.p) return phub_fail(c, PHUB_ERR_OOM, "phub_cf_chain: vectors");
    double n2 = 0;
    int rc = dev_dot<double, double>(c, start, start, n, &n2);
    if (rc) return rc;
    *norm2 = n2;
    for (int i = 0; i < nsteps; ++i) alpha[i] = beta[i] = 0.0;
    if (!(n2 > 0.0)) return PHUB_OK;
    hipLaunchKernelGGL(k_convert<T>, dim3(g), dim3(256), 0, c->stream, start, v0.p, n, 1.0);
    phub_lanczos_opts o{};
    o.max_steps = nsteps;
    LczState s;
    rc = lanczos_plain<T>(c, op, &o, v0.p, b1.p, b2.p, s, nullptr, true);
    if (rc) return rc;
    for (int i = 0; i < s.steps && i < nsteps; ++i) {
        alpha[i] = s.alpha[i];
        beta[i] = s.beta[i + 1];
    }
    return PHUB_OK;
}

// ---------------------------------------------------------------------------
// dense symmetric eigensolver for the band matrix (host): Householder reduction
// to tridiagonal form, then the implicit QL above.  A is row-major m x m and
// destroyed; w ascending; Z (optional) column j = eigenvector j (Z[j*m + i]).
// ---------------------------------------------------------------------------
static int sym_eig(int m, std::vector<double> &A, std::vector<double> &w, std::vector<double> *Z) {
    w.assign(m, 0.0);
    if (m == 0) return 0;
    std::vector<double> Qm;
    if (Z) {
        Qm.assign((size_t)m * m, 0.0);
        for (int i = 0; i < m; ++i) Qm[(size_t)i * m + i] = 1.0;
    }
    std::vector<double> v(m), pv(m), q(m);
    for (int k = 0; k + 2 < m; ++k) {
        double sn = 0.0;
        for (int i = k + 1; i < m; ++i) sn += A[(size_t)i * m + k] * A[(size_t)i * m + k];
        if (sn == 0.0) continue;
        double al = std::sqrt(sn);
        if (A[(size_t)(k + 1) * m + k] > 0.0) al = -al;
        std::fill(v.begin(), v.end(), 0.0);
        v[k + 1] = A[(size_t)(k + 1) * m + k] - al;
        for (int i = k + 2; i < m; ++i) v[i] = A[(size_t)i * m + k];
        double vn2 = 0.0;
        for (int i = k + 1; i < m; ++i) vn2 += v[i] * v[i];
        if (vn2 == 0.0) continue;
        const double bt = 2.0 / vn2;
        // A <- H A H with H = I - bt v v^T:  p = bt A v, q = p - (bt v.p / 2) v,
        // A -= v q^T + q v^T
        double vp = 0.0;
        for (int i = k; i < m; ++i) {
            double t = 0.0;
            for (int j = k + 1; j < m; ++j) t += A[(size_t)i * m + j] * v[j];
            pv[i] = bt * t;
            vp += v[i] * pv[i];
        }
        const double K = 0.5 * bt * vp;
        for (int i = k; i < m; ++i) q[i] = pv[i] - K * v[i];
        for (int i = k; i < m; ++i)
            for (int j = k; j < m; ++j) A[(size_t)i * m + j] -= v[i] * q[j] + q[i] * v[j];
        if (Z)
            for (int r = 0; r < m; ++r) {
                double t = 0.0;
                for (int j = k + 1; j < m; ++j) t += Qm[(size_t)r * m + j] * v[j];
                t *= bt;
                for (int j = k + 1; j < m; ++j) Qm[(size_t)r * m + j] -= t * v[j];
            }
    }
    std::vector<double> d(m), e(std::max(m - 1, 1), 0.0), z;
    for (int i = 0; i < m; ++i) d[i] = A[(size_t)i * m + i];
    for (int i = 0; i + 1 < m; ++i) e[i] = A[(size_t)(i + 1) * m + i];
    if (Z) z.resize((size_t)m * m);
    if (tridiag_ql(m, d.data(), e.data(), Z ? z.data() : nullptr)) return 1;
    w = d;
    if (Z) {
        Z->assign((size_t)m * m, 0.0);
        for (int jv = 0; jv < m; ++jv)
            for (int i = 0; i < m; ++i) {
                double t = 0.0;
                for (int r = 0; r < m; ++r) t += Qm[(size_t)i * m + r] * z[(size_t)jv * m + r];
                (*Z)[(size_t)jv * m + i] = t;
            }
    }
    return 0;
}

// lowest eigenvalue of a symmetric band matrix (bandwidth bw, row-major m x m) by
// bisection on the inertia of A - sigma I (number of negative pivots of its
// LDL^T, Sylvester's law), O(m bw^2) per shift
static double band_lowest(int m, int bw, const std::vector<double> &A) {
    double lo = 1e300, hi = -1e300;
    for (int i = 0; i < m; ++i) {
        double r = 0.0;
        for (int j = std::max(0, i - bw); j <= std::min(m - 1, i + bw); ++j)
            if (j != i) r += std::fabs(A[(size_t)i * m + j]);
        lo = std::min(lo, A[(size_t)i * m + i] - r);
        hi = std::max(hi, A[(size_t)i * m + i] + r);
    }
    const int w = bw + 1;
    std::vector<double> L((size_t)m * w, 0.0), D(m);
    auto below = [&](double sg) {
        int neg = 0;
        for (int i = 0; i < m; ++i) {
            const int j0 = std::max(0, i - bw);
            for (int j = j0; j < i; ++j) {
                double t = A[(size_t)i * m + j];
                for (int k = std::max(j0, j - bw); k < j; ++k)
                    t -= L[(size_t)i * w + (i - k)] * L[(size_t)j * w + (j - k)] * D[k];
                L[(size_t)i * w + (i - j)] = t / D[j];
            }
            double d = A[(size_t)i * m + i] - sg;
            for (int k = j0; k < i; ++k) d -= L[(size_t)i * w + (i - k)] * L[(size_t)i * w + (i - k)] * D[k];
            if (d == 0.0) d = -1e-300;
            D[i] = d;
            if (d < 0.0) ++neg;
        }
        return neg;
    };
    for (int it = 0; it < 200 && hi - lo > 1e-15 * std::max(1.0, std::fabs(lo)); ++it) {
        const double mid = 0.5 * (lo + hi);
        if (below(mid) >= 1) hi = mid;
        else lo = mid;
    }
    return 0.5 * (lo + hi);
}

// band matrix of the recurrence with the deflated indices dropped
// (spgf_bandlanczos.f90:396-431): ht[(m x m) row-major], returns m
static int band_matrix(int nb, int p, const std::vector<double> &alpha,
                       const std::vector<double> &beta2, const std::vector<int> &def,
                       std::vector<double> &ht) {
    std::vector<int> keep;
    for (int i = 1; i <= nb; ++i) {     // 1-based index as DEF
        bool drop = false;
        for (int k = 0; k < p; ++k)
            if (def[k] != 0 && i >= def[k] && (i - def[k]) % p == 0) drop = true;
        if (!drop) keep.push_back(i - 1);
    }
    const int m = (int)keep.size();
    std::vector<int> pos(nb, -1);
    for (int a = 0; a < m; ++a) pos[keep[a]] = a;
    ht.assign((size_t)m * m, 0.0);
    for (int i = 0; i < nb; ++i) {
        if (pos[i] >= 0) ht[(size_t)pos[i] * m + pos[i]] = alpha[i];
        for (int j = 1; j <= p && i + j < nb; ++j) {
            if (pos[i] < 0 || pos[i + j] < 0) continue;
            const double b = beta2[(size_t)i * p + (j - 1)];
            ht[(size_t)pos[i + j] * m + pos[i]] = b;
            ht[(size_t)pos[i] * m + pos[i + j]] = b;
        }
    }
    return m;
}

extern "C" int phub_band_lanczos_workspace_bytes(const phub_op *op, int p, size_t *bytes) {
    if (!op || !bytes || p < 1) return PHUB_ERR_ARG;
    *bytes = (size_t)(2 * p + 1) * basis_ld(op->dim) * sizeof(double) + 256;
    return PHUB_OK;
}

extern "C" int phub_band_lanczos(phub_ctx *c, const phub_op *op, const double *d_V, int p,
                                 const phub_band_opts *o, double *positions, double *Q,
                                 int *npoles) {
    if (!c || !op || !d_V || !o || !positions || !Q || !npoles)
        return phub_fail(c, PHUB_ERR_ARG, "null argument");
    if (p < 1 || p > 256) return phub_fail(c, PHUB_ERR_ARG, "p must be in [1, 256]");
    if (o->max_steps < p + 11) return phub_fail(c, PHUB_ERR_ARG, "max_steps too small for p");
    PHUB_HIP(c, hipSetDevice(c->device));
    *npoles = 0;
    const int64_t n = op->dim, ld = basis_ld(n);
    const double DF = 1e-10;   // deflation threshold (spgf_bandlanczos.f90:312)
    // 2p + 1 vector slots: the p previous vectors, the current one and its p - 1
    // successors, and the work vector; rotated by pointer, never copied
    // the slots from the caller's workspace (else allocated for the call); the
    // coefficients and the pointer table from the context arena
    Arena ar(c, o->workspace, o->workspace_bytes);
    struct { double *p; } slots{(double *)ar.take((size_t)(2 * p + 1) * ld * sizeof(double))};
    if (!slots.p) return phub_fail(c, PHUB_ERR_OOM, "phub_band_lanczos: vectors");
    struct { double *p; } coefd{nullptr};
    int arc = phub_arena(c, (size_t)(4 * p + 9) * sizeof(double), (void **)&coefd.p);
    if (arc) return arc;
    struct { const double **p; } dptr{(const double **)(coefd.p + 2 * p + 8)};
    PHUB_HIP(c, hipMemsetAsync(slots.p, 0, (size_t)(2 * p + 1) * ld * sizeof(double), c->stream));
    std::vector<double *> ring(2 * p);       // [0,p): previous block V0, [p,2p): V
    for (int i = 0; i < 2 * p; ++i) ring[i] = slots.p + (size_t)i * ld;
    double *wv = slots.p + (size_t)(2 * p) * ld;
    for (int i = 0; i < p; ++i)
        PHUB_HIP(c, hipMemcpyAsync(ring[p + i], d_V + (size_t)i * n, n * sizeof(double),
                                   hipMemcpyDeviceToDevice, c->stream));
    // multidot / combine over a list of vectors (pointer table on the device)
    std::vector<const double *> hp(2 * p + 1);
    auto mdot = [&](const std::vector<double *> &vs, int first, int cnt, const double *w,
                    double *out) -> int {
        for (int j = 0; j < cnt; ++j) hp[j] = vs[first + j];
        PHUB_HIP(c, hipMemcpyAsync(dptr.p, hp.data(), cnt * sizeof(double *), hipMemcpyHostToDevice, c->stream));
        const int64_t slices = (n + 2048 - 1) / 2048;
        const unsigned g = (unsigned)std::max<int64_t>(1, std::min<int64_t>(slices, 2048));
        int rc = phub_ensure_red(c, (size_t)g * cnt + cnt + 64);
        if (rc) return rc;
        hipLaunchKernelGGL((k_multidot_v<double, double>), dim3(g), dim3(256), 4 * cnt * sizeof(double),
                           c->stream, (const double *)nullptr, (int64_t)0, cnt, w,
                           (const double *)nullptr, 0.0, n, c->d_red, (const double *const *)dptr.p);
        PHUB_HIP(c, hipGetLastError());
        return phub_reduce_partials(c, (int)g, cnt, out);
    };
    auto msub = [&](const std::vector<double *> &vs, int first, int cnt, const double *coef,
                    double *w, double *norm2) -> int {
        for (int j = 0; j < cnt; ++j) hp[j] = vs[first + j];
        PHUB_HIP(c, hipMemcpyAsync(dptr.p, hp.data(), cnt * sizeof(double *), hipMemcpyHostToDevice, c->stream));
        PHUB_HIP(c, hipMemcpyAsync(coefd.p, coef, cnt * sizeof(double), hipMemcpyHostToDevice, c->stream));
        const unsigned g = grid_for(n);
        int rc = phub_ensure_red(c, g + 64);
        if (rc) return rc;
        hipLaunchKernelGGL((k_combine_v<double, double>), dim3(g), dim3(256), cnt * sizeof(double),
                           c->stream, (const double *)nullptr, (int64_t)0, cnt, coefd.p, w, n, 0,
                           norm2 ? c->d_red : nullptr, (const double *const *)dptr.p);
        PHUB_HIP(c, hipGetLastError());
        if (norm2) return phub_reduce_partials(c, (int)g, 1, norm2);
        PHUB_HIP(c, hipStreamSynchronize(c->stream));   // coef / pointer uploads stay valid
        return PHUB_OK;
    };
    // Gram-Schmidt of the start block (spgf_bandlanczos.f90:331-346): V_i = sum_j R(i,j) q_j
    std::vector<double> R((size_t)p * p, 0.0), cf(2 * p + 8);
    for (int i = 0; i < p; ++i) {
        double *vi = ring[p + i];
        double nn = 0.0;
        if (i > 0) {
            int rc = mdot(ring, p, i, vi, cf.data());
            if (rc) return rc;
            for (int j = 0; j < i; ++j) R[(size_t)i * p + j] = cf[j];
            rc = msub(ring, p, i, cf.data(), vi, &nn);
            if (rc) return rc;
        } else {
            int rc = dev_dot<double, double>(c, vi, vi, n, &nn);
            if (rc) return rc;
        }
        const double nr = std::sqrt(std::max(nn, 0.0));
        R[(size_t)i * p + i] = nr;
        if (!(nr > 0.0)) return phub_fail(c, PHUB_ERR_ARG, "start vectors are linearly dependent");
        int rc = dev_axpby<double, double>(c, 0.0, (const double *)nullptr, 1.0 / nr, vi, n, nullptr);
        if (rc) return rc;
    }
    // band recurrence (spgf_bandlanczos.f90:352-390)
    const int cap = o->max_steps + p;
    std::vector<double> alpha(cap, 0.0), beta2((size_t)cap * p, 0.0), ht, val, vec;
    std::vector<int> def(p, 0);
    int nb = 0, m = 0;
    bool cv = false;
    double ev = 0.0;
    while (!cv && nb < o->max_steps - 10) {
        for (int i = 0; i < p; ++i) {
            double *cur = ring[p];
            // w = H v, alpha = <v, w>
            int nblk = 0;
            int rc = phub_hv_launch<double>(c, op, cur, wv, 1.0, (const double *)nullptr, 0.0, true, &nblk);
            if (rc) return rc;
            double dots[2];
            rc = phub_reduce_partials(c, nblk, 2, dots);
            if (rc) return rc;
            const double a = dots[0];
            alpha[nb] = a;
            // against the previous block (coefficients not kept, :56-59)
            rc = mdot(ring, 0, p, wv, cf.data());
            if (rc) return rc;
            rc = msub(ring, 0, p, cf.data(), wv, nullptr);
            if (rc) return rc;
            // against the successors: band entries (:60-63)
            if (p > 1) {
                rc = mdot(ring, p + 1, p - 1, wv, cf.data());
                if (rc) return rc;
                for (int j = 0; j < p - 1; ++j) beta2[(size_t)nb * p + j] = cf[j];
                rc = msub(ring, p + 1, p - 1, cf.data(), wv, nullptr);
                if (rc) return rc;
            }
            double nn = 0.0;
            rc = dev_axpby<double, double>(c, -a, cur, 1.0, wv, n, &nn);
            if (rc) return rc;
            const double b = std::sqrt(std::max(nn, 0.0));
            beta2[(size_t)nb * p + (p - 1)] = b;
            ++nb;
            if (b <= DF && def[i] == 0) def[i] = nb + p;   // deflation (:66-72), 1-based
            // shift: the oldest vector's slot receives the new last vector w / b
            double *freed = ring[0];
            for (int j = 0; j + 1 < 2 * p; ++j) ring[j] = ring[j + 1];
            ring[2 * p - 1] = wv;
            rc = dev_axpby<double, double>(c, 0.0, (const double *)nullptr, b > 0.0 ? 1.0 / b : 0.0,
                                           wv, n, nullptr);
            if (rc) return rc;
            wv = freed;
        }
        m = band_matrix(nb, p, alpha, beta2, def, ht);
        const double e1 = band_lowest(m, p, ht);
        if (std::fabs(e1 - ev) < o->tol) cv = true;
        else ev = e1;
    }
    // final poles and weights
    m = band_matrix(nb, p, alpha, beta2, def, ht);
    if (sym_eig(m, ht, val, &vec)) return phub_fail(c, PHUB_ERR_HIP, "band eigensolver failed");
    for (int k = 0; k < m; ++k) {
        positions[k] = val[k];
        for (int i = 0; i < p; ++i) {
            double t = 0.0;
            for (int j = 0; j <= i; ++j) t += R[(size_t)i * p + j] * vec[(size_t)k * m + j];
            Q[(size_t)k * p + i] = t;
        }
    }
    *npoles = m;
    if (!cv) return phub_fail(c, PHUB_ERR_NOCONV, "FORTRAN BANDLANCZOS ERROR: MAX LCZ STEP REACH equivalent");
    return PHUB_OK;
}

extern "C" int phub_cf_chain_workspace_bytes(const phub_op *op, int precision, size_t *bytes) {
    if (!op || !bytes || (precision != 32 && precision != 64)) return PHUB_ERR_ARG;
    const size_t v = ((size_t)op->dim * (precision / 8) + 255) & ~(size_t)255;
    *bytes = 3 * v + 256;
    return PHUB_OK;
}

extern "C" int phub_cf_chain(phub_ctx *c, const phub_op *op, const double *d_start, int nsteps,
                             int precision, double *alpha, double *beta, double *norm2,
                             void *workspace, size_t workspace_bytes) {
    if (!c || !op || !d_start || !alpha || !beta || !norm2 || nsteps < 1)
        return phub_fail(c, PHUB_ERR_ARG, "bad argument");
    if (op->dim == 0) {
        *norm2 = 0;
        for (int i = 0; i < nsteps; ++i) alpha[i] = beta[i] = 0.0;
        return PHUB_OK;
    }
    if (precision == 64)
        return cf_chain_t<double>(c, op, d_start, nsteps, alpha, beta, norm2, workspace,
                                  workspace_bytes);
    if (precision == 32)
        return cf_chain_t<float>(c, op, d_start, nsteps, alpha, beta, norm2, workspace,
                                 workspace_bytes);
    return phub_fail(c, PHUB_ERR_ARG, "precision must be 32 or 64");
}

// ---------------------------------------------------------------------------
// adjoint MINRES: (H - e0) x = b on the complement of psi (Paige-Saunders MINRES)
// ---------------------------------------------------------------------------
extern "C" int phub_adjoint_minres_workspace_bytes(const phub_op *op, size_t *bytes) {
    if (!op || !bytes) return PHUB_ERR_ARG;
    // seven work vectors and the preconditioner's diagonal
    *bytes = 8 * (((size_t)op->dim * sizeof(double) + 255) & ~(size_t)255) + 256;
    return PHUB_OK;
}

// MINRES on (H - e0) restricted to the complement of psi (Paige-Saunders), vectors
// of type V: x (zeroed here) solves it for the right-hand side b projected off psi,
// to the relative residual tol or max_iter iterations.  Seven work vectors from ar.
// With minv (a positive diagonal M^-1) the iteration is preconditioned by P M^-1 P,
// P = 1 - psi psi^T: symmetric and positive definite on the complement, where every
// iterate lies.  The residual vectors r stay orthogonal to psi, so the preconditioned
// residual is minv * r - (psi . minv r) psi and its norm^2 is r . (minv r); both dots
// come from the pass that forms r, and the projection happens where v is formed.  tol
// then bounds the residual in the M^-1 norm.
template <typename V>
static int minres_core(phub_ctx *c, const phub_op *op, const V *psi, double e0, const V *b,
                       V *x, int max_iter, double tol, int *iters, double *rel_res, Arena &ar,
                       const V *minv) {
    const int64_t n = op->dim;
    // r2 is also the unnormalised Lanczos vector y of the iteration; wl / wp are the
    // last two search directions and wf the slot the next one is written to
    struct Vec { V *p; };
    Vec r1{}, r2{}, v{}, t{}, wl{}, wp{}, wf{};
    for (Vec *bf : {&r1, &r2, &v, &t, &wl, &wp, &wf}) {
        bf->p = (V *)ar.take(n * sizeof(V));
        if (!bf->p) return phub_fail(c, PHUB_ERR_OOM, "MINRES work vectors");
    }
    const unsigned g = grid_for(n);
    int rc = phub_ensure_red(c, (size_t)g * 3 + 3 + 64);
    if (rc) return rc;
    PHUB_HIP(c, hipMemsetAsync(x, 0, n * sizeof(V), c->stream));
    PHUB_HIP(c, hipMemsetAsync(wl.p, 0, n * sizeof(V), c->stream));
    PHUB_HIP(c, hipMemsetAsync(wp.p, 0, n * sizeof(V), c->stream));
    // r1 = b - (psi . b) psi
    double pb = 0;
    rc = dev_dot<V, V>(c, psi, b, n, &pb);
    if (rc) return rc;
    hipLaunchKernelGGL(k_axpby_to<V>, dim3(g), dim3(256), 0, c->stream, -pb, psi, 1.0, b, r1.p, n);
    PHUB_HIP(c, hipGetLastError());
    // r2 = r1, with r1 . (M^-1 r1) and psi . (M^-1 r1) (or |r1|^2)
    if (minv)
        hipLaunchKernelGGL((k_minres_b<V, true>), dim3(g), dim3(256), 0, c->stream, (const V *)r1.p,
                           psi, (const V *)r1.p, 0.0, 0.0, minv, r2.p, n, c->d_red);
    else
        hipLaunchKernelGGL((k_minres_b<V, false>), dim3(g), dim3(256), 0, c->stream, (const V *)r1.p,
                           psi, (const V *)r1.p, 0.0, 0.0, minv, r2.p, n, c->d_red);
    PHUB_HIP(c, hipGetLastError());
    double d2[2];
    rc = phub_reduce_partials(c, (int)g, 2, d2);
    if (rc) return rc;
    const double beta1 = std::sqrt(std::max(d2[0], 0.0));
    double gam = d2[1];
    if (iters) *iters = 0;
    if (rel_res) *rel_res = 0.0;
    if (!(beta1 > 0.0)) return PHUB_OK;
    double oldb = 0, beta = beta1, dbar = 0, epsln = 0, phibar = beta1, cs = -1, sn = 0;
    int it = 0;
    double rnorm = beta1;
    for (it = 1; it <= max_iter; ++it) {
        // v = y / beta (y = r2, or its preconditioned form); t = (H - e0) v - (beta / oldb) r1;
        // p = psi . (H - e0) v
        if (minv)
            hipLaunchKernelGGL(k_minres_pscale<V>, dim3(g), dim3(256), 0, c->stream, v.p,
                               (const V *)r2.p, minv, psi, gam, 1.0 / beta, n);
        else
            hipLaunchKernelGGL(k_minres_scale<V>, dim3(g), dim3(256), 0, c->stream, v.p, r2.p,
                               1.0 / beta, n);
        rc = phub_hv_launch<V>(c, op, v.p, t.p, 1.0, (const V *)nullptr, 0.0, false, nullptr);
        if (rc) return rc;
        if (it >= 2)
            hipLaunchKernelGGL((k_minres_a<V, true>), dim3(g), dim3(256), 0, c->stream, t.p, v.p,
                               (const V *)r1.p, psi, e0, beta / oldb, n, c->d_red);
        else
            hipLaunchKernelGGL((k_minres_a<V, false>), dim3(g), dim3(256), 0, c->stream, t.p, v.p,
                               (const V *)nullptr, psi, e0, 0.0, n, c->d_red);
        PHUB_HIP(c, hipGetLastError());
        double d3[3];
        rc = phub_reduce_partials(c, (int)g, 3, d3);
        if (rc) return rc;
        const double p = d3[0];                    // psi . (H - e0) v
        const double alfa = d3[1] - p * d3[2];     // v . (t - p psi)
        // y_new = t - p psi - (alfa / beta) r2, into r1's slot; then r1 <- r2, r2 <- y_new
        if (minv)
            hipLaunchKernelGGL((k_minres_b<V, true>), dim3(g), dim3(256), 0, c->stream, (const V *)t.p,
                               psi, (const V *)r2.p, p, alfa / beta, minv, r1.p, n, c->d_red);
        else
            hipLaunchKernelGGL((k_minres_b<V, false>), dim3(g), dim3(256), 0, c->stream, (const V *)t.p,
                               psi, (const V *)r2.p, p, alfa / beta, minv, r1.p, n, c->d_red);
        PHUB_HIP(c, hipGetLastError());
        double yy[2];
        rc = phub_reduce_partials(c, (int)g, 2, yy);
        if (rc) return rc;
        std::swap(r1.p, r2.p);
        oldb = beta;
        beta = std::sqrt(std::max(yy[0], 0.0));
        gam = yy[1];
        const double oldeps = epsln;
        const double delta = cs * dbar + sn * alfa;
        const double gbar = sn * dbar - cs * alfa;
        epsln = sn * beta;
        dbar = -cs * beta;
        double gamma = std::max(std::hypot(gbar, beta), 1e-300);
        cs = gbar / gamma;
        sn = beta / gamma;
        const double phi = cs * phibar;
        phibar = sn * phibar;
        // w = (v - oldeps wp - delta wl) / gamma, x += phi w; wp <- wl, wl <- w
        hipLaunchKernelGGL(k_minres_c<V>, dim3(g), dim3(256), 0, c->stream, v.p, wp.p, wl.p,
                           oldeps, delta, 1.0 / gamma, phi, wf.p, x, n);
        PHUB_HIP(c, hipGetLastError());
        // (wf, wp, wl) <- (old wp, old wl, the one just written)
        std::swap(wf.p, wp.p);
        std::swap(wp.p, wl.p);
        rnorm = phibar;
        if (rnorm / beta1 < tol || !(beta > 0.0)) break;
    }
    PHUB_HIP(c, hipStreamSynchronize(c->stream));
    if (iters) *iters = std::min(it, max_iter);
    if (rel_res) *rel_res = rnorm / beta1;
    return PHUB_OK;
}

extern "C" int phub_adjoint_minres(phub_ctx *c, const phub_op *op, const double *psi, double e0,
                                   const double *b, double *x, int max_iter, double tol,
                                   int *iters, double *rel_res, void *workspace,
                                   size_t workspace_bytes) {
    if (!c || !op || !psi || !b || !x) return phub_fail(c, PHUB_ERR_ARG, "null argument");
    // the work vectors come from the caller's workspace, else are allocated for the call
    Arena ar(c, workspace, workspace_bytes);
    double *minv = nullptr;
    if (phub_minres_precond(op)) {
        minv = (double *)ar.take(op->dim * sizeof(double));
        if (!minv) return phub_fail(c, PHUB_ERR_OOM, "MINRES preconditioner vector");
        int rc = phub_hub_minv(c, op, e0, minv);
        if (rc) return rc;
    }
    return minres_core<double>(c, op, psi, e0, b, x, max_iter, tol, iters, rel_res, ar,
                               (const double *)minv);
}
