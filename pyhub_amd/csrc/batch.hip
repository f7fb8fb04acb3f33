// batch.hip -- many small Hubbard clusters in one call.
//
// Work: for every cluster a ground state (Lanczos in FP64 vectors, stopped on the
// Ritz residual, psi0 rebuilt by a second pass; the chains' accuracy is set by
// psi0, so it is not computed in FP32), then for every site i and spin s the
// particle chain c^dag_is psi0 and the hole chain c_is psi0 (continued-fraction
// Lanczos, the single-vector case of GF_BANDLANCZOS, spgf_bandlanczos.f90:309-448;
// excitation signs as GF_E_* / GF_H_*, spgf_ed.f90:304-429) in FP32 vectors
// with FP64 recurrence scalars.
//
// Two launches on one stream, each a persistent grid of one workgroup per CU
// pulling tasks from an atomic counter: the first solves every ground state
// (largest cluster first), the second runs every chain (largest sector first).
// Stream order makes the ground states visible to the chains, so no workgroup
// ever waits on another.  A ground-state sector of BIG_GS states or more would
// hold one workgroup far longer than the rest of the launch, so those ground
// states are solved instead by the full-device Lanczos (FP64 vectors, the same
// three-term recurrence) from host threads on streams of their own while the first
// launch runs; the chain launch is queued after they finish.
#include "phub_internal.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

namespace {

constexpr int BT = 1024;          // threads per workgroup
constexpr int MAXN = 16;          // sites per cluster in the batched path
constexpr int MAXSTEPS = 256;     // Lanczos steps per task
// ground states of at least this many states: full device (a sweep of the 256-cluster
// workload: 0.905 s at 100,000, 0.880 s at 40,000 and 60,000, 0.91 s at 25,000)
constexpr int64_t BIG_GS = 50000;
// dynamic LDS of the persistent workgroup (one per CU) for the H*v's row and column
// staging: 19,968 FP64 or 39,936 FP32 elements, at least one row or column of a
// 16-site sector (12,870)
constexpr int BATCH_STG_BYTES = 156 * 1024;

struct SectorDev {
    int64_t nu, nd;
    int nup, ndn;
    const uint32_t *mu, *md;
    const double *eu, *ed;
    const int32_t *cu, *iu;
    const float *cfu;
    const double *cfu64;
    const int32_t *cd, *id;
    const float *cfd;
    const double *cfd64;
    int maxh_dn;
};

template <typename V>
__device__ __forceinline__ V cf_up(const SectorDev &s, size_t i) {
    if constexpr (sizeof(V) == 8) return s.cfu64[i];
    else return s.cfu[i];
}
template <typename V>
__device__ __forceinline__ V cf_dn(const SectorDev &s, size_t i) {
    if constexpr (sizeof(V) == 8) return s.cfd64[i];
    else return s.cfd[i];
}

struct ClusterDev {
    int n;
    int u_uniform;
    double u0;
    double U[MAXN];
    SectorDev sec[5];      // 0 ground, 1 up-1, 2 up+1, 3 dn-1, 4 dn+1
    double *psi0;          // ground-state vector (sec[0] dimension), FP64
    int64_t psi_off;
};

struct Task {
    int cluster;
    int kind;              // 0 ground state, 1 chain
    int site, spin, create;
    int slot;              // output slot of the chain
};

struct BatchDev {
    const ClusterDev *cl;
    const Task *tasks;
    int ntasks;
    int gs_steps, gf_steps;
    double gs_tol;
    uint64_t seed;
    unsigned int *counter;
    double *scratch;       // per workgroup: 3 vectors of dmax doubles (FP32 chains use the front)
    int64_t dmax;
    double *e0;            // per cluster
    double *ga, *gb, *gn;  // chain outputs [slot * gf_steps], norm2 [slot]
    const uint64_t *binom;
    double *tri;           // per workgroup FP64 scratch: alpha/beta of up to MAXSTEPS
    int first;             // task range of this launch: [first, last)
    int last;
};

__device__ inline uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// workgroup-wide sum of two doubles (BT threads, 16 waves)
__device__ inline void wg_sum2(double &a, double &b, double *sh) {
    for (int off = 32; off > 0; off >>= 1) {
        a += __shfl_xor(a, off);
        b += __shfl_xor(b, off);
    }
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    __syncthreads();
    if (l == 0) { sh[2 * w] = a; sh[2 * w + 1] = b; }
    __syncthreads();
    double s0 = 0.0, s1 = 0.0;
    for (int i = 0; i < BT / 64; ++i) { s0 += sh[2 * i]; s1 += sh[2 * i + 1]; }
    a = s0;
    b = s1;
}

// out = scale * (H x) - bprev * xp ; returns (<scale x, out>, |out|^2) workgroup-wide.
// Two passes through the workgroup's dynamic LDS (stg, cap elements), so that every
// gather reads LDS:
//  1. blocks of whole rows of x: the diagonal and the up hops (a source is in the
//     state's own row), one thread per column and group of rows so that each
//     up-table entry serves the group; out holds their sum;
//  2. blocks of whole columns (every row's slice of ncol columns, stg[d][c]): the down
//     hops (a source is in the state's own column), then the epilogue and the dots.
// Each pass reads x once more than a single pass would (and out once), in place of the
// down hops' row gathers from global memory.  Indices within a block fit 32 bits.
// Between the passes, out written by one thread is read by another of the workgroup
// after a barrier (workgroup-scope visibility, as between the steps of a chain).
template <typename V>
__device__ void wg_hv(const ClusterDev &c, const SectorDev &s, const V *__restrict__ x,
                      double scale, const V *__restrict__ xp, double bprev,
                      V *__restrict__ out, double &d_xo, double &d_oo, double *sh,
                      V *__restrict__ stg, int cap) {
    const uint32_t nu = (uint32_t)s.nu, nd = (uint32_t)s.nd;
    const uint32_t rb = (uint32_t)cap / nu;    // rows per block (>= 1: cap >= nu)
    for (uint32_t d0 = 0; d0 < nd; d0 += rb) {
        const uint32_t nrow = min(rb, nd - d0);
        const uint32_t nel = nrow * nu;
        const V *__restrict__ xb = x + (int64_t)d0 * nu;
        __syncthreads();                               // the previous block is read
        for (uint32_t l = threadIdx.x; l < nel; l += BT) stg[l] = xb[l];
        __syncthreads();
        // one thread per column and group of RG rows: each up-table entry (loaded once)
        // serves the RG rows (8 in FP32; 4 in FP64, whose path already spills registers)
        constexpr uint32_t RG = sizeof(V) == 4 ? 8 : 4;
        const uint32_t ngrp = (nrow + RG - 1) / RG, nitem = ngrp * nu;
        for (uint32_t l = threadIdx.x; l < nitem; l += BT) {
            const uint32_t g4 = l / nu, u = l - g4 * nu;
            const uint32_t r0 = RG * g4;
            const uint32_t xu = s.mu[u];
            const double eu = s.eu[u];
            V acc[RG];
#pragma unroll
            for (int r = 0; r < (int)RG; ++r) {
                const uint32_t rr = min(r0 + r, nrow - 1);   // rows past the block repeat the last
                const uint32_t d = d0 + rr;
                double dg = eu + s.ed[d];
                const uint32_t both = xu & s.md[d];
                if (c.u_uniform) {
                    dg += c.u0 * (double)__popc(both);
                } else {
                    uint32_t bb = both;
                    while (bb) {
                        const int bt = __ffs(bb) - 1;
                        dg += c.U[c.n - 1 - bt];
                        bb &= bb - 1;
                    }
                }
                acc[r] = (V)dg * stg[rr * nu + u];
            }
            const V *sr[RG];
#pragma unroll
            for (int r = 0; r < (int)RG; ++r) sr[r] = stg + min(r0 + r, nrow - 1) * nu;
            const int cu = s.cu[u];
            for (int k = 0; k < cu; ++k) {
                const int32_t j = s.iu[(size_t)k * nu + u];
                const V cf = cf_up<V>(s, (size_t)k * nu + u);
#pragma unroll
                for (int r = 0; r < (int)RG; ++r) acc[r] += cf * sr[r][j];
            }
#pragma unroll
            for (int r = 0; r < (int)RG; ++r)
                if (r0 + r < nrow) out[(int64_t)(d0 + r0 + r) * nu + u] = acc[r];
        }
    }
    const uint32_t cb = (uint32_t)cap / nd;    // columns per block (>= 1: cap >= nd)
    double a = 0.0, b = 0.0;
    for (uint32_t u0 = 0; u0 < nu; u0 += cb) {
        const uint32_t ncol = min(cb, nu - u0);
        const uint32_t nel = ncol * nd;
        __syncthreads();                               // the previous block (and pass 1) done
        for (uint32_t l = threadIdx.x; l < nel; l += BT) {
            const uint32_t d = l / ncol, cc = l - d * ncol;
            stg[l] = x[(int64_t)d * nu + u0 + cc];
        }
        __syncthreads();
        for (uint32_t l = threadIdx.x; l < nel; l += BT) {
            const uint32_t d = l / ncol, cc = l - d * ncol;
            const int64_t e = (int64_t)d * nu + u0 + cc;
            V acc = out[e];
            const int cd = s.cd[d];
            const int32_t *idr = s.id + (size_t)d * s.maxh_dn;
            int k = 0;
            for (; k + 4 <= cd; k += 4) {
                int32_t j[4];
                V cf[4];
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    j[q] = idr[k + q];
                    cf[q] = cf_dn<V>(s, (size_t)d * s.maxh_dn + k + q);
                }
#pragma unroll
                for (int q = 0; q < 4; ++q) acc += cf[q] * stg[(uint32_t)j[q] * ncol + cc];
            }
            for (; k < cd; ++k)
                acc += cf_dn<V>(s, (size_t)d * s.maxh_dn + k) * stg[(uint32_t)idr[k] * ncol + cc];
            double v = scale * (double)acc;
            if (xp) v -= bprev * (double)xp[e];
            const V vt = (V)v;
            out[e] = vt;
            a += scale * (double)stg[l] * (double)vt;
            b += (double)vt * (double)vt;
        }
    }
    wg_sum2(a, b, sh);
    d_xo = a;
    d_oo = b;
}

// y = y - a * x (x scaled by sx) ; returns |y|^2 workgroup-wide
template <typename V>
__device__ double wg_update(V *__restrict__ y, const V *__restrict__ x, double a, int64_t D,
                            double *sh) {
    double s = 0.0, dummy = 0.0;
    for (int64_t e = threadIdx.x; e < D; e += BT) {
        const V t = (V)((double)y[e] - a * (double)x[e]);
        y[e] = t;
        s += (double)t * (double)t;
    }
    wg_sum2(s, dummy, sh);
    return s;
}

// Lanczos chain of `steps` steps from v (overwritten).  alpha[k], beta[k] (beta[k]
// couples k and k+1).  Three buffers rotate; vectors are kept unnormalised with
// scale sig.  If psi is non-null, the pass accumulates psi += y[k] * x_k.
__device__ double tri_lowest(int m, const double *al, const double *be, double *y, double *wk);

// With tol > 0 (ground states) the recurrence stops once the Ritz residual
// |beta_m * y_m| of the lowest pair drops below tol * max(1, |E|), tested every
// 10 steps from step 20 (y and wk: FP64 scratch of MAXSTEPS each).
template <typename V>
__device__ int wg_lanczos(const ClusterDev &c, const SectorDev &s, V *v0, V *b1, V *b2,
                          int steps, double *alpha, double *beta, double norm0, const double *y,
                          V *psi, double *sh, V *stg, int cap, double tol = 0.0,
                          double *ty = nullptr, double *twk = nullptr) {
    __shared__ int stop_sh;
    const int64_t D = s.nu * s.nd;
    V *xp = b1, *xk = v0, *out = b2;
    double sig = 1.0 / norm0, sigp = 0.0, bk = 0.0;
    int m = 0;
    for (int k = 0; k < steps; ++k) {
        if (psi) {
            for (int64_t e = threadIdx.x; e < D; e += BT) psi[e] += (V)(y[k] * sig * (double)xk[e]);
        }
        if (k == steps - 1 && psi) { m = steps; break; }
        double dxo, doo;
        wg_hv(c, s, xk, sig, k ? xp : nullptr, bk * sigp, out, dxo, doo, sh, stg, cap);
        const double a = dxo;
        const double nn = wg_update(out, xk, a * sig, D, sh);
        const double b = sqrt(fmax(nn, 0.0));
        if (!psi && threadIdx.x == 0) { alpha[k] = a; beta[k] = b; }
        m = k + 1;
        // rotate
        V *t = xp;
        xp = xk;
        xk = out;
        out = t;
        sigp = sig;
        bk = psi ? beta[k] : b;
        sig = (bk > 0.0) ? 1.0 / bk : 0.0;
        __syncthreads();
        if (!(b > 1e-10 * fmax(1.0, fabs(a)))) {        // invariant subspace
            // the chain ends here: no coupling to a further level
            if (!psi && threadIdx.x == 0) beta[k] = 0.0;
            break;
        }
        if (tol > 0.0 && !psi && m >= 20 && m % 10 == 0 && m < steps) {
            if (threadIdx.x == 0) {
                const double e = tri_lowest(m, alpha, beta, ty, twk);
                stop_sh = fabs(b * ty[m - 1]) < tol * fmax(1.0, fabs(e));
            }
            __syncthreads();
            const int stop = stop_sh;
            __syncthreads();
            if (stop) break;
        }
    }
    return m;
}

// lowest eigenvector of the tridiagonal (alpha[0..m-1], beta[0..m-2]) by
// bisection + inverse iteration, single thread (m <= MAXSTEPS)
__device__ double tri_lowest(int m, const double *al, const double *be, double *y, double *wk) {
    double lo = al[0], hi = al[0];
    for (int i = 0; i < m; ++i) {
        const double r = (i > 0 ? fabs(be[i - 1]) : 0.0) + (i < m - 1 ? fabs(be[i]) : 0.0);
        lo = fmin(lo, al[i] - r);
        hi = fmax(hi, al[i] + r);
    }
    for (int it = 0; it < 200; ++it) {          // Sturm count: eigenvalues below mid
        const double mid = 0.5 * (lo + hi);
        int cnt = 0;
        double q = al[0] - mid;
        if (q < 0) ++cnt;
        for (int i = 1; i < m; ++i) {
            q = al[i] - mid - be[i - 1] * be[i - 1] / (q != 0.0 ? q : 1e-300);
            if (q < 0) ++cnt;
        }
        if (cnt >= 1) hi = mid; else lo = mid;
        if (hi - lo <= 1e-15 * fmax(1.0, fabs(mid))) break;
    }
    const double lam = 0.5 * (lo + hi) - 1e-13 * fmax(1.0, fabs(lo));
    // inverse iteration (T - lam) y = y, Thomas algorithm, 3 sweeps
    for (int i = 0; i < m; ++i) y[i] = 1.0;
    double *cp = wk, *dp = wk + m;
    for (int sweep = 0; sweep < 3; ++sweep) {
        cp[0] = (m > 1 ? be[0] : 0.0) / (al[0] - lam);
        dp[0] = y[0] / (al[0] - lam);
        for (int i = 1; i < m; ++i) {
            const double den = (al[i] - lam) - be[i - 1] * cp[i - 1];
            cp[i] = (i < m - 1 ? be[i] : 0.0) / den;
            dp[i] = (y[i] - be[i - 1] * dp[i - 1]) / den;
        }
        y[m - 1] = dp[m - 1];
        for (int i = m - 2; i >= 0; --i) y[i] = dp[i] - cp[i] * y[i + 1];
        double nrm = 0.0;
        for (int i = 0; i < m; ++i) nrm += y[i] * y[i];
        nrm = sqrt(nrm);
        for (int i = 0; i < m; ++i) y[i] /= nrm;
    }
    return 0.5 * (lo + hi);
}

// excitation c^dag / c of psi0 into sector t (pull form), returns |out|^2
__device__ double wg_excite(const BatchDev &B, const ClusterDev &c, const SectorDev &g,
                            const SectorDev &t, int site, int spin, int create, float *out,
                            double *sh) {
    const int n = c.n, bit = n - site;
    const uint32_t bm = 1u << bit;
    const uint32_t above = (bit + 1 < 32) ? ~((bm << 1) - 1u) : 0u;
    const double parity = (spin == 0 && (g.ndn & 1)) ? -1.0 : 1.0;
    const int64_t D = t.nu * t.nd;
    double s = 0.0, dummy = 0.0;
    for (int64_t e = threadIdx.x; e < D; e += BT) {
        const int64_t d = e / t.nu, u = e - d * t.nu;
        float v = 0.f;
        if (spin == 0) {
            const uint32_t xt = t.mu[u];
            const bool occ = (xt & bm) != 0;
            if (create ? occ : !occ) {
                const uint32_t xs = xt ^ bm;
                const int64_t us = rank_ref_dev(B.binom, xs, n, g.nup);
                const double ac = (__popc(xs & above) & 1) ? -1.0 : 1.0;
                v = (float)(parity * ac * (double)c.psi0[d * g.nu + us]);
            }
        } else {
            const uint32_t xt = t.md[d];
            const bool occ = (xt & bm) != 0;
            if (create ? occ : !occ) {
                const uint32_t xs = xt ^ bm;
                const int64_t ds = rank_ref_dev(B.binom, xs, n, g.ndn);
                const double ac = (__popc(xs & above) & 1) ? -1.0 : 1.0;
                v = (float)(ac * (double)c.psi0[ds * g.nu + u]);
            }
        }
        out[e] = v;
        s += (double)v * (double)v;
    }
    wg_sum2(s, dummy, sh);
    return s;
}

// KIND 0: ground-state tasks only (FP64); KIND 1: chain tasks only (FP32).  Two
// instantiations, so that each gets the registers of its own path.
template <int KIND>
__global__ __launch_bounds__(BT) void k_batch(BatchDev B) {
    __shared__ double sh[2 * BT / 64];
    __shared__ int task_sh;
    // row staging of the H*v: the dynamic LDS (BATCH_STG_BYTES), FP64 for the ground
    // states, FP32 for the chains
    extern __shared__ __attribute__((aligned(16))) unsigned char bstg[];
    double *stg64 = reinterpret_cast<double *>(bstg);
    float *stg32 = reinterpret_cast<float *>(bstg);
    constexpr int cap64 = BATCH_STG_BYTES / 8, cap32 = BATCH_STG_BYTES / 4;
    // FP64 view for the ground state, FP32 view (front half) for the chains
    double *d0 = B.scratch + (size_t)blockIdx.x * 3 * B.dmax;
    double *d1 = d0 + B.dmax, *d2 = d1 + B.dmax;
    float *w0 = reinterpret_cast<float *>(d0);
    float *w1 = w0 + B.dmax, *w2 = w1 + B.dmax;
    double *tri = B.tri + (size_t)blockIdx.x * 6 * MAXSTEPS;
    while (true) {
        if (threadIdx.x == 0) task_sh = B.first + (int)atomicAdd(B.counter, 1u);
        __syncthreads();
        const int ti = task_sh;
        __syncthreads();
        if (ti >= B.last) break;
        const Task tk = B.tasks[ti];
        const ClusterDev &c = B.cl[tk.cluster];
        const SectorDev &g = c.sec[0];
        const int64_t Dg = g.nu * g.nd;
        if constexpr (KIND == 0) {
            // ground state in FP64 vectors: random start, plain recurrence, second pass for psi0
            double n2 = 0.0, dummy = 0.0;
            for (int64_t e = threadIdx.x; e < Dg; e += BT) {
                const uint64_t r = mix64(B.seed * 0x9E3779B97F4A7C15ull + (uint64_t)tk.cluster * 0x100000001B3ull + (uint64_t)e);
                const double v = (double)(r >> 11) * (1.0 / 9007199254740992.0) - 0.5;
                d0[e] = v;
                c.psi0[e] = v;   // keep the start vector for the second pass
                n2 += v * v;
            }
            wg_sum2(n2, dummy, sh);
            const double nrm = sqrt(n2);
            double *al = tri, *be = tri + MAXSTEPS, *y = tri + 2 * MAXSTEPS, *wk = tri + 3 * MAXSTEPS;
            const int steps = (int)min<int64_t>(B.gs_steps, Dg);
            const int m = wg_lanczos<double>(c, g, d0, d1, d2, steps, al, be, nrm, nullptr, nullptr, sh,
                                             stg64, cap64, B.gs_tol, y, wk);
            __syncthreads();
            if (threadIdx.x == 0) B.e0[tk.cluster] = tri_lowest(m, al, be, y, wk);
            __syncthreads();
            // second pass: d0 <- start, psi0 accumulated in place (psi0 holds the start)
            for (int64_t e = threadIdx.x; e < Dg; e += BT) {
                d0[e] = c.psi0[e];
                c.psi0[e] = 0.0;
            }
            __syncthreads();
            wg_lanczos<double>(c, g, d0, d1, d2, m, al, be, nrm, y, c.psi0, sh, stg64, cap64);
            __syncthreads();
            // normalise psi0
            double p2 = 0.0, d2s = 0.0;
            for (int64_t e = threadIdx.x; e < Dg; e += BT) p2 += c.psi0[e] * c.psi0[e];
            wg_sum2(p2, d2s, sh);
            const double inv = 1.0 / sqrt(p2);
            for (int64_t e = threadIdx.x; e < Dg; e += BT) c.psi0[e] *= inv;
            __syncthreads();
        } else {
            const int sidx = (tk.spin == 0) ? (tk.create ? 2 : 1) : (tk.create ? 4 : 3);
            const SectorDev &t = c.sec[sidx];
            const int64_t Dt = t.nu * t.nd;
            double *ga = B.ga + (size_t)tk.slot * B.gf_steps;
            double *gb = B.gb + (size_t)tk.slot * B.gf_steps;
            if (Dt == 0) {
                if (threadIdx.x == 0) B.gn[tk.slot] = 0.0;
                for (int k = threadIdx.x; k < B.gf_steps; k += BT) { ga[k] = 0.0; gb[k] = 0.0; }
                continue;
            }
            const double n2 = wg_excite(B, c, g, t, tk.site, tk.spin, tk.create, w0, sh);
            if (threadIdx.x == 0) B.gn[tk.slot] = n2;
            for (int k = threadIdx.x; k < B.gf_steps; k += BT) { ga[k] = 0.0; gb[k] = 0.0; }
            __syncthreads();
            if (n2 > 0.0) {
                // every chain runs gf_steps steps, also on sectors with fewer states: in
                // FP32 the basis loses orthogonality before the sector is exhausted, so
                // stopping at the sector dimension would keep an O(1) last beta (and
                // drop poles not yet converged); the extra steps only add ghost copies
                // of converged poles, whose weights sum to the pole's own
                wg_lanczos<float>(c, t, w0, w1, w2, B.gf_steps, ga, gb, sqrt(n2), nullptr, nullptr, sh,
                                  stg32, cap32);
            }
            __syncthreads();
        }
    }
}

template <typename T>
struct DB {
    T *p = nullptr;
    ~DB() {
        if (p) hipFree(p);
    }
};

}  // namespace

// Ground states of the clusters in `big` by phub_lanczos (PLAIN64: the three-term
// recurrence of the persistent path, with the vectors kept and psi combined from
// them; psi into the cluster's psi0 slot, e0 into d_e0), beside the persistent
// launch of the small ground states on c->stream.  Each solve is bound by its host
// round trips, so up to BIG_THREADS host threads run them at once, each on a context
// with a stream of its own; every thread synchronises its stream before it ends, so
// the chain launch queued after the join sees psi0 and e0.  The stopping rule is the
// Ritz value's change, 1e-3 of gs_tol, which leaves psi0 at least as converged as the
// persistent path's residual rule.
//
// hipMemcpy and hipFree wait for kernels on every stream, the persistent launch
// included, so the workers' streams, contexts (whose creation copies with hipMemcpy)
// and Krylov workspaces are made before that launch (BigWorkers::prepare), their
// contexts keep freed memory until they are destroyed (defer_free), and all of it is
// released after the chains (the destructor).
constexpr int BIG_THREADS = 8;

struct BigWorkers {
    struct W {
        hipStream_t s = nullptr;
        phub_ctx *c = nullptr;
        void *ws = nullptr;
        size_t ws_bytes = 0;
    };
    std::vector<W> w;
    int device = 0;
    ~BigWorkers() {
        hipSetDevice(device);
        for (W &x : w) {
            if (x.s) hipStreamSynchronize(x.s);
            if (x.c) phub_ctx_destroy(x.c);
            if (x.ws) hipFree(x.ws);
            if (x.s) hipStreamDestroy(x.s);
        }
    }

    static phub_lanczos_opts opts(const phub_batch_opts *o, int i) {
        phub_lanczos_opts lo{};
        lo.mode = PHUB_LCZ_PLAIN64;
        lo.max_steps = o->gs_steps;
        lo.tol = 1e-3 * o->gs_tol;
        lo.check_every = 5;
        lo.seed = o->seed + 0x9E3779B97F4A7C15ull * (uint64_t)(i + 1);
        lo.want_psi = 1;
        lo.n_states = 1;
        return lo;
    }

    // worker k solves clusters big[k], big[k + nth], ...
    int prepare(phub_ctx *c, const std::vector<int> &big, const std::vector<phub_op *> &ops_gs,
                const phub_batch_opts *o) {
        device = c->device;
        const int nth = std::min<int>(BIG_THREADS, (int)big.size());
        w.resize(nth);
        for (int k = 0; k < nth; ++k) {
            W &x = w[k];
            PHUB_HIP(c, hipStreamCreateWithFlags(&x.s, hipStreamNonBlocking));
            if (phub_ctx_create(c->device, (void *)x.s, &x.c) != PHUB_OK)
                return phub_fail(c, PHUB_ERR_HIP, "context for the large ground states: " +
                                                      std::string(x.c ? x.c->err : ""));
            x.c->defer_free = true;
            for (size_t j = k; j < big.size(); j += nth) {
                const phub_lanczos_opts lo = opts(o, big[j]);
                size_t b = 0;
                if (phub_lanczos_workspace_bytes(ops_gs[big[j]], &lo, &b) != PHUB_OK)
                    return phub_fail(c, PHUB_ERR_ARG, "large ground state workspace");
                x.ws_bytes = std::max(x.ws_bytes, b);
            }
            PHUB_HIP(c, phub_malloc(&x.ws, x.ws_bytes));
        }
        return PHUB_OK;
    }

    int run(phub_ctx *c, const std::vector<int> &big, const std::vector<phub_op *> &ops_gs,
            const std::vector<ClusterDev> &hc, const phub_batch_opts *o, double *d_e0) {
        const int nth = (int)w.size();
        std::vector<int> rcs(nth, PHUB_OK);
        std::vector<std::string> msgs(nth);
        auto worker = [&](int k) {
            W &x = w[k];
            int rc = hipSetDevice(device) == hipSuccess ? PHUB_OK : PHUB_ERR_HIP;
            std::vector<double> al(MAXSTEPS + 2), be(MAXSTEPS + 3);
            for (size_t j = k; j < big.size() && rc == PHUB_OK; j += nth) {
                const int i = big[j];
                phub_lanczos_opts lo = opts(o, i);
                lo.workspace = x.ws;
                lo.workspace_bytes = x.ws_bytes;
                phub_lanczos_result res{};
                rc = phub_lanczos(x.c, ops_gs[i], &lo, nullptr, al.data(), be.data(), &res, hc[i].psi0);
                if (rc == PHUB_ERR_NOCONV) rc = PHUB_OK;   // the step cap, as in the persistent path
                if (rc == PHUB_OK) {
                    const double e = res.e0;
                    if (hipMemcpyAsync(d_e0 + i, &e, sizeof(double), hipMemcpyHostToDevice, x.s) != hipSuccess ||
                        hipStreamSynchronize(x.s) != hipSuccess)
                        rc = PHUB_ERR_HIP;
                }
            }
            if (hipStreamSynchronize(x.s) != hipSuccess && rc == PHUB_OK) rc = PHUB_ERR_HIP;
            rcs[k] = rc;
            if (rc) msgs[k] = x.c->err;
        };
        std::vector<std::thread> th;
        std::vector<int> here{0};   // workers run on this thread (all of them if threads fail)
        for (int k = 1; k < nth; ++k) {
            try {
                th.emplace_back(worker, k);
            } catch (...) {
                here.push_back(k);
            }
        }
        for (int k : here) worker(k);
        for (auto &t : th) t.join();
        for (int k = 0; k < nth; ++k)
            if (rcs[k]) return phub_fail(c, rcs[k], msgs[k].empty() ? "large ground state" : msgs[k]);
        return PHUB_OK;
    }
};

extern "C" int phub_batch_hubbard(phub_ctx *c, const phub_cluster *cls, int ncl,
                                  const phub_batch_opts *o, double *e0, double *gf_alpha,
                                  double *gf_beta, double *gf_norm2) {
    if (!c || !cls || ncl < 1 || !o || !e0) return phub_fail(c, PHUB_ERR_ARG, "bad argument");
    if (o->gs_steps < 2 || o->gs_steps > MAXSTEPS || o->gf_steps < 1 || o->gf_steps > MAXSTEPS)
        return phub_fail(c, PHUB_ERR_ARG, "steps must be in [2, 256]");
    PHUB_HIP(c, hipSetDevice(c->device));
    std::vector<phub_op *> ops;
    HopBatch tabs;
    auto cleanup = [&]() {
        phub_free_hop_batch(&tabs);
        for (auto *p : ops) phub_op_destroy(p);
    };
    std::vector<ClusterDev> hc(ncl);
    std::vector<phub_op *> ops_gs(ncl, nullptr);   // ground-sector operator of the large clusters
    // the per-species tables of every cluster: up tables of N_up - 1, N_up, N_up + 1
    // (k-major) and down tables of N_dn - 1, N_dn, N_dn + 1 (row-major), shared by the
    // five sectors, all built in one batch
    std::vector<HopSpec> specs;
    std::vector<int> tab_of(ncl * 6, -1);   // [cluster][0..2 up, 3..5 down] -> table
    for (int i = 0; i < ncl; ++i) {
        const phub_cluster &q = cls[i];
        if (q.n_sites < 1 || q.n_sites > MAXN) { cleanup(); return phub_fail(c, PHUB_ERR_UNSUPPORTED, "batched clusters need n <= 16"); }
        if (q.n_up < 0 || q.n_up > q.n_sites || q.n_dn < 0 || q.n_dn > q.n_sites) {
            cleanup();
            return phub_fail(c, PHUB_ERR_ARG, "particle numbers out of range");
        }
        for (int j = 0; j < q.n_sites * q.n_sites; ++j)
            if (!std::isfinite(q.t[j])) { cleanup(); return phub_fail(c, PHUB_ERR_ARG, "t_matrix has non-finite entries"); }
        for (int j = 0; j < q.n_sites; ++j)
            if (!std::isfinite(q.u[j])) { cleanup(); return phub_fail(c, PHUB_ERR_ARG, "U has non-finite entries"); }
        for (int J = 0; J < q.n_sites; ++J)
            for (int K = 0; K < J; ++K)
                if (q.t[J * q.n_sites + K] != q.t[K * q.n_sites + J]) {
                    cleanup();
                    return phub_fail(c, PHUB_ERR_ARG, "t_matrix must be symmetric");
                }
        for (int sp = 0; sp < 2; ++sp)
            for (int dk = -1; dk <= 1; ++dk) {
                const int k = (sp == 0 ? q.n_up : q.n_dn) + dk;
                if (k < 0 || k > q.n_sites) continue;
                tab_of[i * 6 + sp * 3 + dk + 1] = (int)specs.size();
                specs.push_back(HopSpec{q.n_sites, k, q.t, sp == 0});
            }
    }
    {
        int rc = phub_build_hop_tables(c, specs, &tabs);
        if (rc) { cleanup(); return rc; }
    }
    int64_t dmax = 1, psi_total = 0;
    int64_t big_gs = BIG_GS;   // PHUB_BATCH_BIG_GS overrides the threshold
    if (const char *e = std::getenv("PHUB_BATCH_BIG_GS"))
        if (std::atoll(e) > 0) big_gs = std::atoll(e);
    for (int i = 0; i < ncl; ++i) {
        const phub_cluster &q = cls[i];
        ClusterDev &h = hc[i];
        std::memset(&h, 0, sizeof(h));
        h.n = q.n_sites;
        h.u_uniform = 1;
        h.u0 = q.u[0];
        for (int j = 0; j < h.n; ++j) {
            h.U[j] = q.u[j];
            if (q.u[j] != q.u[0]) h.u_uniform = 0;
        }
        // sector s: (up offset, down offset) from (N_up, N_dn)
        const int secs[5][2] = {{0, 0}, {-1, 0}, {1, 0}, {0, -1}, {0, 1}};
        for (int s = 0; s < 5; ++s) {
            SectorDev &sd = h.sec[s];
            sd.nup = q.n_up + secs[s][0];
            sd.ndn = q.n_dn + secs[s][1];
            const int tu = tab_of[i * 6 + secs[s][0] + 1], tdn = tab_of[i * 6 + 3 + secs[s][1] + 1];
            if (tu < 0 || tdn < 0) continue;   // empty sector
            const HopTable &up = tabs.tabs[tu], &dn = tabs.tabs[tdn];
            sd.nu = up.dim;
            sd.nd = dn.dim;
            sd.mu = up.masks;
            sd.md = dn.masks;
            sd.eu = up.eloc;
            sd.ed = dn.eloc;
            sd.cu = up.cnt;
            sd.iu = up.idx;
            sd.cfu = up.c32;
            sd.cfu64 = up.c64;
            sd.cd = dn.cnt;
            sd.id = dn.idx;
            sd.cfd = dn.c32;
            sd.cfd64 = dn.c64;
            sd.maxh_dn = std::max(dn.maxh, 1);
            dmax = std::max<int64_t>(dmax, sd.nu * sd.nd);
        }
        h.psi_off = psi_total;
        psi_total += std::max<int64_t>(h.sec[0].nu * h.sec[0].nd, 1);
        if (h.sec[0].nu * h.sec[0].nd >= big_gs) {
            phub_hubbard_desc desc{q.n_sites, q.n_up, q.n_dn, q.t, q.u, nullptr};
            phub_op *op = nullptr;
            int rc = phub_op_hubbard(c, &desc, &op);
            if (rc) { cleanup(); return rc; }
            ops.push_back(op);
            ops_gs[i] = op;
        }
    }
    // tasks: ground states (largest first), then chains (largest first)
    std::vector<Task> tasks;
    std::vector<int> order(ncl);
    std::iota(order.begin(), order.end(), 0);
    std::sort(order.begin(), order.end(), [&](int a, int b) {
        return hc[a].sec[0].nu * hc[a].sec[0].nd > hc[b].sec[0].nu * hc[b].sec[0].nd;
    });
    std::vector<int> big;   // clusters whose ground state runs on the full device
    for (int i : order) {
        if (hc[i].sec[0].nu * hc[i].sec[0].nd >= big_gs) big.push_back(i);
        else tasks.push_back(Task{i, 0, 0, 0, 0, 0});
    }
    const int ngs = (int)tasks.size();
    std::vector<Task> chains;
    for (int i = 0; i < ncl; ++i)
        for (int site = 1; site <= hc[i].n; ++site)
            for (int spin = 0; spin < 2; ++spin)
                for (int kind = 0; kind < 2; ++kind) {
                    const int slot = ((i * 32 + (site - 1)) * 2 + spin) * 2 + kind;
                    chains.push_back(Task{i, 1, site, spin, kind == 0 ? 1 : 0, slot});
                }
    auto tdim = [&](const Task &t) {
        const int s = (t.spin == 0) ? (t.create ? 2 : 1) : (t.create ? 4 : 3);
        return hc[t.cluster].sec[s].nu * hc[t.cluster].sec[s].nd;
    };
    std::stable_sort(chains.begin(), chains.end(),
                     [&](const Task &a, const Task &b) { return tdim(a) > tdim(b); });
    tasks.insert(tasks.end(), chains.begin(), chains.end());
    const int nslots = ncl * 32 * 2 * 2;
    // device buffers
    int dev_cu = 0;
    hipDeviceGetAttribute(&dev_cu, hipDeviceAttributeMultiprocessorCount, c->device);
    const int grid = std::max(1, std::min<int>(dev_cu > 0 ? dev_cu : 256, (int)tasks.size()));
    DB<double> psi, scratch;
    DB<ClusterDev> dcl;
    DB<Task> dtasks;
    DB<unsigned int> counter;
    DB<double> de0, ga, gb, gn, tri;
    PHUB_HIP(c, phub_malloc(&psi.p, psi_total * sizeof(double)));
    for (int i = 0; i < ncl; ++i) hc[i].psi0 = psi.p + hc[i].psi_off;
    PHUB_HIP(c, phub_malloc(&scratch.p, (size_t)grid * 3 * dmax * sizeof(double)));
    PHUB_HIP(c, phub_malloc(&dcl.p, ncl * sizeof(ClusterDev)));
    PHUB_HIP(c, phub_malloc(&dtasks.p, tasks.size() * sizeof(Task)));
    PHUB_HIP(c, phub_malloc(&counter.p, 2 * sizeof(unsigned int)));
    PHUB_HIP(c, phub_malloc(&de0.p, ncl * sizeof(double)));
    PHUB_HIP(c, phub_malloc(&ga.p, (size_t)nslots * o->gf_steps * sizeof(double)));
    PHUB_HIP(c, phub_malloc(&gb.p, (size_t)nslots * o->gf_steps * sizeof(double)));
    PHUB_HIP(c, phub_malloc(&gn.p, nslots * sizeof(double)));
    PHUB_HIP(c, phub_malloc(&tri.p, (size_t)grid * 6 * MAXSTEPS * sizeof(double)));
    PHUB_HIP(c, phub_copy(c, dcl.p, hc.data(), ncl * sizeof(ClusterDev), hipMemcpyHostToDevice));
    PHUB_HIP(c, phub_copy(c, dtasks.p, tasks.data(), tasks.size() * sizeof(Task), hipMemcpyHostToDevice));
    PHUB_HIP(c, phub_clear(c, counter.p, 2 * sizeof(unsigned int)));
    PHUB_HIP(c, phub_clear(c, gn.p, nslots * sizeof(double)));
    PHUB_HIP(c, phub_clear(c, ga.p, (size_t)nslots * o->gf_steps * sizeof(double)));
    PHUB_HIP(c, phub_clear(c, gb.p, (size_t)nslots * o->gf_steps * sizeof(double)));
    BatchDev B{};
    B.cl = dcl.p;
    B.tasks = dtasks.p;
    B.ntasks = (int)tasks.size();
    B.gs_steps = o->gs_steps;
    B.gs_tol = o->gs_tol;
    B.gf_steps = o->gf_steps;
    B.seed = o->seed;
    B.scratch = scratch.p;
    B.dmax = dmax;
    B.e0 = de0.p;
    B.ga = ga.p;
    B.gb = gb.p;
    B.gn = gn.p;
    B.binom = c->d_binom;
    B.tri = tri.p;
    // launch 1: the small ground states; meanwhile the large ones on a second stream;
    // launch 2: chains, after both (stream order and an event publish psi0 and e0)
    B.counter = counter.p;
    B.first = 0;
    B.last = ngs;
    BigWorkers bw;   // declared after the buffers: destroyed (synchronised) before them
    if (!big.empty()) {
        int rc = bw.prepare(c, big, ops_gs, o);
        if (rc) { cleanup(); return rc; }
    }
    {
        static bool attr_set[PHUB_MAX_DEVICES];   // once per device and process
        if (c->device < 0 || c->device >= PHUB_MAX_DEVICES) {
            cleanup();
            return phub_fail(c, PHUB_ERR_UNSUPPORTED, "device index beyond PHUB_MAX_DEVICES");
        }
        if (!attr_set[c->device]) {
            PHUB_HIP(c, hipFuncSetAttribute((const void *)k_batch<0>,
                                            hipFuncAttributeMaxDynamicSharedMemorySize, BATCH_STG_BYTES));
            PHUB_HIP(c, hipFuncSetAttribute((const void *)k_batch<1>,
                                            hipFuncAttributeMaxDynamicSharedMemorySize, BATCH_STG_BYTES));
            attr_set[c->device] = true;
        }
    }
    if (ngs > 0)
        hipLaunchKernelGGL(k_batch<0>, dim3(std::max(1, std::min(grid, ngs))), dim3(BT), BATCH_STG_BYTES,
                           c->stream, B);
    PHUB_HIP(c, hipGetLastError());
    if (!big.empty()) {
        int rc = bw.run(c, big, ops_gs, hc, o, de0.p);
        if (rc) {
            // the first launch still reads the buffers freed on return
            hipStreamSynchronize(c->stream);
            cleanup();
            return rc;
        }
    }
    B.counter = counter.p + 1;
    B.first = ngs;
    B.last = (int)tasks.size();
    if (B.last > B.first)
        hipLaunchKernelGGL(k_batch<1>, dim3(std::max(1, std::min(grid, B.last - B.first))), dim3(BT),
                           BATCH_STG_BYTES, c->stream, B);
    PHUB_HIP(c, hipGetLastError());
    PHUB_HIP(c, hipStreamSynchronize(c->stream));
    PHUB_HIP(c, phub_copy(c, e0, de0.p, ncl * sizeof(double), hipMemcpyDeviceToHost));
    if (gf_alpha) PHUB_HIP(c, phub_copy(c, gf_alpha, ga.p, (size_t)nslots * o->gf_steps * sizeof(double), hipMemcpyDeviceToHost));
    if (gf_beta) PHUB_HIP(c, phub_copy(c, gf_beta, gb.p, (size_t)nslots * o->gf_steps * sizeof(double), hipMemcpyDeviceToHost));
    if (gf_norm2) PHUB_HIP(c, phub_copy(c, gf_norm2, gn.p, nslots * sizeof(double), hipMemcpyDeviceToHost));
    cleanup();
    return PHUB_OK;
}
