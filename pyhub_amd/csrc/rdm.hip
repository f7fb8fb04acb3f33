// rdm.hip -- reduced quantities <bra| O |ket> and Hellmann-Feynman gradients.
//
// Reference: ONE_RDM, TWO_RDM (double occupancy NI) and J_SPIN
// (pyhub/libs/src/static_rq.f90:153-395) and the Heisenberg J_SPIN
// (static_rq_heis.f90:58-116).  The reference loops over site pairs with a linear
// search per hop; here every quantity is one or two passes over the state:
//   * hop expectations <bra| c^dag_K c_J |ket> for all J != K:
//       up spin   -- lanes own a column u, read their all-pairs entries once and
//                    accumulate over a chunk of rows in registers;
//       down spin -- entries are uniform over a row, block-reduced per entry;
//   * diagonal quantities (densities, double occupancy, <n n> correlations) from
//     row sums p(d), column sums q(u) and R[d][J] = sum_u w[d][u] n_J(u);
//   * spin-flip part of <S_J.S_K>: per element, up hop J->K and down hop K->J.
#include "phub_internal.h"

#include <algorithm>
#include <cstdlib>
#include <cmath>
#include <cstring>
#include <vector>

__device__ __forceinline__ double wsum(double a) {
    for (int off = 32; off > 0; off >>= 1) a += __shfl_xor(a, off);
    return a;
}

// block-wide sum of one value (blockDim = 256) -> returned to every thread
__device__ __forceinline__ double block_sum(double a, double *sh) {
    a = wsum(a);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = a;
    __syncthreads();
    return sh[0] + sh[1] + sh[2] + sh[3];
}

constexpr int RDM_ROWS = 64;   // rows per workgroup in the up-hop pass
constexpr int RDM_CHUNK = 8;   // entries accumulated in registers at once

// up-spin hop expectations: out[p] += sign * sum_d bra[d][u] * ket[d][u'] over entries
__global__ __launch_bounds__(256) void k_rdm_up(const int64_t nu, const int64_t nd,
                                                const int32_t *__restrict__ cnt,
                                                const int32_t *__restrict__ idx,
                                                const double *__restrict__ sgn,
                                                const int16_t *__restrict__ pair,
                                                const double *__restrict__ bra,
                                                const double *__restrict__ ket, int np,
                                                double *__restrict__ out, int ncb) {
    extern __shared__ double acc[];   // np doubles
    for (int i = threadIdx.x; i < np; i += blockDim.x) acc[i] = 0.0;
    __syncthreads();
    const int cb = blockIdx.x % ncb;
    const int64_t d0 = (int64_t)(blockIdx.x / ncb) * RDM_ROWS;
    const int64_t d1 = min(d0 + RDM_ROWS, nd);
    const int64_t u = (int64_t)cb * 256 + threadIdx.x;
    if (u < nu) {
        const int c = cnt[u];
        for (int k0 = 0; k0 < c; k0 += RDM_CHUNK) {
            double a[RDM_CHUNK];
            int32_t t[RDM_CHUNK];
#pragma unroll
            for (int e = 0; e < RDM_CHUNK; ++e) {
                a[e] = 0.0;
                t[e] = (k0 + e < c) ? idx[(size_t)(k0 + e) * nu + u] : (int32_t)u;
            }
            for (int64_t d = d0; d < d1; ++d) {
                const double b = bra[d * nu + u];
#pragma unroll
                for (int e = 0; e < RDM_CHUNK; ++e) a[e] += b * ket[d * nu + t[e]];
            }
#pragma unroll
            for (int e = 0; e < RDM_CHUNK; ++e)
                if (k0 + e < c) {
                    const size_t o = (size_t)(k0 + e) * nu + u;
                    atomicAdd(&acc[pair[o]], sgn[o] * a[e]);
                }
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < np; i += blockDim.x)
        if (acc[i] != 0.0) atomicAdd(&out[i], acc[i]);
}

// down-spin hop expectations: one workgroup per row d, entries uniform; RDM_CHUNK
// entries per pass over the row (the bra row read once for them), their sums
// reduced together
__global__ __launch_bounds__(256) void k_rdm_dn(const int64_t nu, const int64_t nd,
                                                const int32_t *__restrict__ cnt,
                                                const int32_t *__restrict__ idx,
                                                const double *__restrict__ sgn,
                                                const int16_t *__restrict__ pair, int maxh,
                                                const double *__restrict__ bra,
                                                const double *__restrict__ ket,
                                                double *__restrict__ out) {
    __shared__ double sh[4 * RDM_CHUNK];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int64_t d = blockIdx.x; d < nd; d += gridDim.x) {
        const int c = cnt[d];
        for (int k0 = 0; k0 < c; k0 += RDM_CHUNK) {
            int64_t dp[RDM_CHUNK];
            double a[RDM_CHUNK];
#pragma unroll
            for (int e = 0; e < RDM_CHUNK; ++e) {
                dp[e] = (k0 + e < c) ? (int64_t)idx[(size_t)d * maxh + k0 + e] : d;
                a[e] = 0.0;
            }
            for (int64_t u = threadIdx.x; u < nu; u += blockDim.x) {
                const double b = bra[d * nu + u];
#pragma unroll
                for (int e = 0; e < RDM_CHUNK; ++e) a[e] += b * ket[dp[e] * nu + u];
            }
#pragma unroll
            for (int e = 0; e < RDM_CHUNK; ++e) a[e] = wsum(a[e]);
            __syncthreads();   // sh is free (the previous chunk's readers are done)
            if (lane == 0)
#pragma unroll
                for (int e = 0; e < RDM_CHUNK; ++e) sh[wave * RDM_CHUNK + e] = a[e];
            __syncthreads();
            if (threadIdx.x < RDM_CHUNK && k0 + (int)threadIdx.x < c) {
                const int e = threadIdx.x;
                const double s = sh[e] + sh[RDM_CHUNK + e] + sh[2 * RDM_CHUNK + e] + sh[3 * RDM_CHUNK + e];
                const size_t o = (size_t)d * maxh + k0 + e;
                atomicAdd(&out[pair[o]], sgn[o] * s);
            }
        }
    }
}

// diagonal pass, rows: one workgroup per row d gives p[d] = sum_u w[d][u] and
// R[d][b] = sum_u w[d][u] n_b(u) (bit position b), w = bra * ket, all n + 1 sums
// reduced together
__global__ __launch_bounds__(256) void k_rdm_diag_rows(const int64_t nu, const int64_t nd, int n,
                                                       const uint32_t *__restrict__ mu,
                                                       const double *__restrict__ bra,
                                                       const double *__restrict__ ket,
                                                       double *__restrict__ R, double *__restrict__ p) {
    __shared__ double sh[4][33];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int64_t d = blockIdx.x; d < nd; d += gridDim.x) {
        double ps = 0.0, r[32];
#pragma unroll
        for (int b = 0; b < 32; ++b) r[b] = 0.0;
        for (int64_t u = threadIdx.x; u < nu; u += blockDim.x) {
            const double w = bra[d * nu + u] * ket[d * nu + u];
            const uint32_t x = mu[u];
            ps += w;
#pragma unroll
            for (int b = 0; b < 32; ++b) r[b] += ((x >> b) & 1u) ? w : 0.0;
        }
        ps = wsum(ps);
#pragma unroll
        for (int b = 0; b < 32; ++b)
            if (b < n) r[b] = wsum(r[b]);
        __syncthreads();   // sh is free (the previous row's readers are done)
        if (lane == 0) {
            sh[wave][32] = ps;
#pragma unroll
            for (int b = 0; b < 32; ++b) sh[wave][b] = r[b];
        }
        __syncthreads();
        if ((int)threadIdx.x <= n) {
            const int b = threadIdx.x == (unsigned)n ? 32 : (int)threadIdx.x;
            const double v = sh[0][b] + sh[1][b] + sh[2][b] + sh[3][b];
            if (b == 32) p[d] = v;
            else R[d * n + b] = v;
        }
    }
}

// diagonal pass, columns: q[u] = sum_d w[d][u], lanes over u, a chunk of rows per
// workgroup
__global__ __launch_bounds__(256) void k_rdm_diag_cols(const int64_t nu, const int64_t nd,
                                                       const double *__restrict__ bra,
                                                       const double *__restrict__ ket,
                                                       double *__restrict__ q, int ncb) {
    const int cb = blockIdx.x % ncb;
    const int64_t d0 = (int64_t)(blockIdx.x / ncb) * RDM_ROWS;
    const int64_t d1 = min(d0 + RDM_ROWS, nd);
    const int64_t u = (int64_t)cb * 256 + threadIdx.x;
    if (u >= nu) return;
    double qs = 0.0;
    for (int64_t d = d0; d < d1; ++d) qs += bra[d * nu + u] * ket[d * nu + u];
    atomicAdd(&q[u], qs);
}

// Diagonal quantities from the row sums p, column sums q and R[d][b] (site J is bit
// n - J), added to out:
//   [0, n) densities up, [n, 2n) densities down, [2n, 3n) NI, then n x n each of
//   <n_J,up n_K,up>, <n_K,up n_J,dn> at [(K-1) n + J-1], <n_J,dn n_K,dn>
// One workgroup per chunk of ASM_CHUNK states (the up states, then the down states,
// in one index range), staged in LDS; one thread per output sum walks the chunk and
// adds its register total to out once.  Launch asm_grid(nu + nd) workgroups of
// ASM_THREADS; n <= 32.
constexpr int ASM_CHUNK = 256;
constexpr int ASM_THREADS = 1024;
__global__ __launch_bounds__(ASM_THREADS) void k_rdm_assemble(int n, int64_t nu, int64_t nd,
                                                              const uint32_t *__restrict__ mu,
                                                              const uint32_t *__restrict__ md,
                                                              const double *__restrict__ q,
                                                              const double *__restrict__ p,
                                                              const double *__restrict__ R,
                                                              double *__restrict__ out) {
    __shared__ uint32_t sm[ASM_CHUNK];
    __shared__ double sw[ASM_CHUNK];
    __shared__ double sr[ASM_CHUNK * 16];   // R rows of the chunk's down states, n <= 16
    const int nn = n * n, tot = 3 * n + 3 * nn;
    const int64_t g0 = (int64_t)blockIdx.x * ASM_CHUNK;
    const int cnt = (int)min<int64_t>(ASM_CHUNK, nu + nd - g0);
    const bool rows_in_lds = n <= 16;
    for (int i = threadIdx.x; i < cnt; i += blockDim.x) {
        const int64_t gi = g0 + i;
        if (gi < nu) {
            sm[i] = mu[gi];
            sw[i] = q[gi];
        } else {
            sm[i] = md[gi - nu];
            sw[i] = p[gi - nu];
        }
    }
    if (rows_in_lds)
        for (int i = threadIdx.x; i < cnt * n; i += blockDim.x) {
            const int64_t gi = g0 + i / n;
            sr[i] = gi >= nu ? R[(gi - nu) * n + i % n] : 0.0;
        }
    __syncthreads();
    // states [0, nup_here) of the chunk are up states, the rest down states
    const int nup_here = (int)max<int64_t>(0, min<int64_t>(cnt, nu - g0));
    for (int t = threadIdx.x; t < tot; t += blockDim.x) {
        double acc = 0.0;
        auto rrow = [&](int i, int b) -> double {   // R[d][b] of chunk state i (a down state)
            return rows_in_lds ? sr[i * n + b] : R[(g0 + i - nu) * n + b];
        };
        if (t < n) {                       // density up, site J = t + 1
            const uint32_t bj = 1u << (n - 1 - t);
            for (int i = 0; i < nup_here; ++i)
                if (sm[i] & bj) acc += sw[i];
        } else if (t < 2 * n) {            // density down
            const uint32_t bj = 1u << (n - 1 - (t - n));
            for (int i = nup_here; i < cnt; ++i)
                if (sm[i] & bj) acc += sw[i];
        } else if (t < 3 * n) {            // NI: R[d][n - J] over down states with J occupied
            const int J = t - 2 * n + 1;
            const uint32_t bj = 1u << (n - J);
            for (int i = nup_here; i < cnt; ++i)
                if (sm[i] & bj) acc += rrow(i, n - J);
        } else {
            const int u = t - 3 * n, blk = u / nn, jk = u % nn;
            const int a = jk / n + 1, b = jk % n + 1;
            if (blk == 0) {                // <n_J,up n_K,up>: J = a, K = b
                const uint32_t m2 = (1u << (n - a)) | (1u << (n - b));
                for (int i = 0; i < nup_here; ++i)
                    if ((sm[i] & m2) == m2) acc += sw[i];
            } else if (blk == 1) {         // [(K-1) n + J-1] = R[d][n - K] over J occupied
                const int K = a, J = b;
                const uint32_t bj = 1u << (n - J);
                for (int i = nup_here; i < cnt; ++i)
                    if (sm[i] & bj) acc += rrow(i, n - K);
            } else {                       // <n_J,dn n_K,dn>
                const uint32_t m2 = (1u << (n - a)) | (1u << (n - b));
                for (int i = nup_here; i < cnt; ++i)
                    if ((sm[i] & m2) == m2) acc += sw[i];
            }
        }
        if (acc != 0.0) atomicAdd(&out[t], acc);
    }
}

static inline unsigned asm_grid(int64_t states) {
    return (unsigned)std::max<int64_t>(1, (states + ASM_CHUNK - 1) / ASM_CHUNK);
}

// the down-spin pair map: c^dag_J c_K applied to every down state, for every
// ordered pair p = (J-1)*n + (K-1) (J == K: -1); layout [d][p], so that the lanes
// of a workgroup (different pairs, the same rows) read the same few rows
__global__ __launch_bounds__(256) void k_pairmap(const uint64_t *__restrict__ binom, int n, int k,
                                                 int64_t nd, const uint32_t *__restrict__ md,
                                                 int32_t *__restrict__ pm) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)n * n * nd) return;
    const int64_t d = i / (n * n);
    const int p = (int)(i - d * n * n);
    const int J = p / n + 1, K = p % n + 1;
    const int bJ = n - J, bK = n - K;
    const uint32_t xd = md[d];
    int32_t e = -1;
    if (J != K && ((xd >> bK) & 1u) && !((xd >> bJ) & 1u)) {
        const int lo = min(bJ, bK), hi = max(bJ, bK);
        const uint32_t between = (hi - lo > 1) ? (((1u << (hi - lo - 1)) - 1u) << (lo + 1)) : 0u;
        const uint32_t xn = xd ^ (1u << bK) ^ (1u << bJ);
        const int64_t dp = rank_ref_dev(binom, xn, n, k);
        e = (int32_t)((dp << 1) | (__popc(xd & between) & 1));
    }
    pm[i] = e;
}

// spin-flip part of J_SPIN (static_rq.f90:345-383): for J != K,
//   F[J][K] = sum ac*ac2 * bra[d][u] * ket[d'][u'], u' = c^dag_K c_J u (up),
//   d' = c^dag_J c_K d (down).  Accumulated in LDS per ordered pair.
__global__ __launch_bounds__(256) void k_rdm_flip(const uint64_t *__restrict__ binom,
                                                  const int64_t nu, const int64_t nd, int n,
                                                  int ndn, const uint32_t *__restrict__ mu,
                                                  const uint32_t *__restrict__ md,
                                                  const int32_t *__restrict__ cnt,
                                                  const int32_t *__restrict__ idx,
                                                  const double *__restrict__ sgn,
                                                  const int16_t *__restrict__ pair,
                                                  const double *__restrict__ bra,
                                                  const double *__restrict__ ket,
                                                  double *__restrict__ out, int ncb,
                                                  const int32_t *__restrict__ pm) {
    extern __shared__ double acc[];   // n*n
    for (int i = threadIdx.x; i < n * n; i += blockDim.x) acc[i] = 0.0;
    __syncthreads();
    const int cb = blockIdx.x % ncb;
    const int64_t d0 = (int64_t)(blockIdx.x / ncb) * RDM_ROWS;
    const int64_t d1 = min(d0 + RDM_ROWS, nd);
    const int64_t u = (int64_t)cb * 256 + threadIdx.x;
    if (u < nu) {
        const int c = cnt[u];
        for (int k = 0; k < c; ++k) {
            const size_t o = (size_t)k * nu + u;
            const int p = pair[o];
            const int J = p / n + 1, K = p % n + 1;
            const int bJ = n - J, bK = n - K;
            const int64_t up2 = idx[o];
            const double ac = sgn[o];
            const int lo = min(bJ, bK), hi = max(bJ, bK);
            const uint32_t between = (hi - lo > 1) ? (((1u << (hi - lo - 1)) - 1u) << (lo + 1)) : 0u;
            double s = 0.0;
            if (pm) {   // the down hop from the pair map: one load per row
                for (int64_t d = d0; d < d1; ++d) {
                    const int32_t e = pm[(size_t)d * n * n + p];
                    if (e >= 0)
                        s += ((e & 1) ? -1.0 : 1.0) * bra[d * nu + u] * ket[(int64_t)(e >> 1) * nu + up2];
                }
            } else
            for (int64_t d = d0; d < d1; ++d) {
                const uint32_t xd = md[d];
                if (((xd >> bK) & 1u) && !((xd >> bJ) & 1u)) {
                    const uint32_t xn = xd ^ (1u << bK) ^ (1u << bJ);
                    const double ac2 = (__popc(xd & between) & 1) ? -1.0 : 1.0;
                    const int64_t dp = rank_ref_dev(binom, xn, n, ndn);
                    s += ac2 * bra[d * nu + u] * ket[dp * nu + up2];
                }
            }
            if (s != 0.0) atomicAdd(&acc[p], ac * s);
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < n * n; i += blockDim.x)
        if (acc[i] != 0.0) atomicAdd(&out[i], acc[i]);
}

// ---------------------------------------------------------------------------
// host assembly
// ---------------------------------------------------------------------------
int phub_ensure_pair_tables(phub_ctx *c, phub_op *op) {
    if (!op->rup.idx) {
        std::vector<double> ones(op->n * op->n, 1.0);
        int rc = phub_build_hop_table(c, op->n, op->nup, ones.data(), true, true, &op->rup);
        if (rc) return rc;
    }
    if (!op->rdn.idx) {
        std::vector<double> ones(op->n * op->n, 1.0);
        int rc = phub_build_hop_table(c, op->n, op->ndn, ones.data(), false, true, &op->rdn);
        if (rc) return rc;
    }
    return PHUB_OK;
}

// Raw expectation values for the Hubbard operator:
//  hop_up[J][K] = <bra| c^dag_K c_J (up) |ket>, J != K (raw ordered pair, not symmetrised)
//  hop_dn[J][K] likewise for down; nu_[J], nd_[J] densities; ni[J] double occupancy;
//  nn_uu, nn_ud, nn_dd [J][K] = <n_Ja n_Kb>; flip[J][K] as in J_SPIN.
struct HubRaw {
    std::vector<double> hop_up, hop_dn, dens_up, dens_dn, ni, nn_uu, nn_ud, nn_dd, flip;
};

static int hubbard_raw(phub_ctx *c, phub_op *op, const double *bra, const double *ket,
                       HubRaw &r, bool want_flip) {
    const int n = op->n;
    const int64_t nu = op->dim_up, nd = op->dim_dn;
    int rc = phub_ensure_pair_tables(c, op);
    if (rc) return rc;
    const int nn = n * n;
    const int ndiag = 3 * n + 3 * nn;
    // one arena block, carved: hop / flip sums, per-row R, p, q, the diagonal sums
    const size_t tot = (size_t)3 * nn + (size_t)nd * n + (size_t)nd + (size_t)nu + (size_t)ndiag;
    void *ab = nullptr;
    rc = phub_arena(c, tot * sizeof(double), &ab);
    if (rc) return rc;
    struct { double *p; } dout{(double *)ab}, dR{dout.p + 3 * nn}, dp{dR.p + nd * n},
        dq{dp.p + nd}, ddiag{dq.p + nu};
    PHUB_HIP(c, hipMemsetAsync(ab, 0, tot * sizeof(double), c->stream));
    const int ncb = (int)((nu + 255) / 256);
    const int64_t nrg = (nd + RDM_ROWS - 1) / RDM_ROWS;
    const unsigned grid = (unsigned)(ncb * nrg);
    if (op->rup.maxh > 0)
        hipLaunchKernelGGL(k_rdm_up, dim3(grid), dim3(256), nn * sizeof(double), c->stream, nu, nd,
                           op->rup.cnt, op->rup.idx, op->rup.c64, op->rup.pair, bra, ket, nn,
                           dout.p, ncb);
    if (op->rdn.maxh > 0)
        hipLaunchKernelGGL(k_rdm_dn, dim3((unsigned)std::min<int64_t>(nd, 8192)), dim3(256), 0,
                           c->stream, nu, nd, op->rdn.cnt, op->rdn.idx, op->rdn.c64,
                           op->rdn.pair, op->rdn.maxh, bra, ket, dout.p + nn);
    hipLaunchKernelGGL(k_rdm_diag_rows, dim3((unsigned)std::min<int64_t>(nd, 16384)), dim3(256), 0,
                       c->stream, nu, nd, n, op->up.masks, bra, ket, dR.p, dp.p);
    hipLaunchKernelGGL(k_rdm_diag_cols, dim3(grid), dim3(256), 0, c->stream, nu, nd, bra, ket, dq.p,
                       ncb);
    if (want_flip && op->rup.maxh > 0) {
        // the pair map when it fits 1 GiB and its entries 31 bits; else ranks per row
        const size_t pm_bytes = (size_t)nn * (size_t)nd * sizeof(int32_t);
        const char *env = std::getenv("PHUB_RDM_PAIRMAP");   // "0": ranks per row
        const bool use_pm = !(env && env[0] == '0');
        if (use_pm && !op->dn_pairmap && pm_bytes <= (size_t(1) << 30) && nd < (1ll << 30)) {
            PHUB_HIP(c, phub_malloc(&op->dn_pairmap, pm_bytes));
            hipLaunchKernelGGL(k_pairmap, dim3((unsigned)(((size_t)nn * nd + 255) / 256)), dim3(256), 0,
                               c->stream, c->d_binom, n, op->ndn, nd, op->dn.masks, op->dn_pairmap);
        }
        hipLaunchKernelGGL(k_rdm_flip, dim3(grid), dim3(256), nn * sizeof(double), c->stream,
                           c->d_binom, nu, nd, n, op->ndn, op->up.masks, op->dn.masks,
                           op->rup.cnt, op->rup.idx, op->rup.c64, op->rup.pair, bra, ket,
                           dout.p + 2 * nn, ncb, use_pm ? (const int32_t *)op->dn_pairmap : nullptr);
    }
    // the diagonal quantities assembled on the device (site J <-> bit n-J)
    hipLaunchKernelGGL(k_rdm_assemble, dim3(asm_grid(nu + nd)), dim3(ASM_THREADS), 0, c->stream, n,
                       nu, nd, op->up.masks, op->dn.masks, dq.p, dp.p, dR.p, ddiag.p);
    PHUB_HIP(c, hipGetLastError());
    std::vector<double> hout(3 * nn), hdiag(ndiag);
    PHUB_HIP(c, hipMemcpyAsync(hout.data(), dout.p, 3 * nn * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    PHUB_HIP(c, hipMemcpyAsync(hdiag.data(), ddiag.p, ndiag * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    PHUB_HIP(c, hipStreamSynchronize(c->stream));
    r.hop_up.assign(hout.begin(), hout.begin() + nn);
    r.hop_dn.assign(hout.begin() + nn, hout.begin() + 2 * nn);
    r.flip.assign(hout.begin() + 2 * nn, hout.end());
    r.dens_up.assign(hdiag.begin(), hdiag.begin() + n);
    r.dens_dn.assign(hdiag.begin() + n, hdiag.begin() + 2 * n);
    r.ni.assign(hdiag.begin() + 2 * n, hdiag.begin() + 3 * n);
    r.nn_uu.assign(hdiag.begin() + 3 * n, hdiag.begin() + 3 * n + nn);
    r.nn_ud.assign(hdiag.begin() + 3 * n + nn, hdiag.begin() + 3 * n + 2 * nn);
    r.nn_dd.assign(hdiag.begin() + 3 * n + 2 * nn, hdiag.end());
    return PHUB_OK;
}

extern "C" int phub_rdm_hubbard(phub_ctx *c, const phub_op *op_, const double *bra,
                                const double *ket, double *rdm_up, double *rdm_dn, double *ni,
                                double *sisj) {
    if (!c || !op_ || !bra || !ket) return phub_fail(c, PHUB_ERR_ARG, "null argument");
    if (op_->kind != 0) return phub_fail(c, PHUB_ERR_ARG, "not a Hubbard operator");
    phub_op *op = const_cast<phub_op *>(op_);
    HubRaw r;
    int rc = hubbard_raw(c, op, bra, ket, r, sisj != nullptr);
    if (rc) return rc;
    const int n = op->n;
    // ONE_RDM (static_rq.f90:161-243): element (J,K), K <= J, mirrored to (K,J);
    // when N_up = N_dn the reference copies the up matrix into the down one (:212-241)
    const bool copy_up = (op->nup == op->ndn);
    for (int J = 1; J <= n; ++J)
        for (int K = 1; K <= J; ++K) {
            double vu = (J == K) ? r.dens_up[J - 1] : r.hop_up[(J - 1) * n + (K - 1)];
            double vd = (J == K) ? r.dens_dn[J - 1] : r.hop_dn[(J - 1) * n + (K - 1)];
            if (copy_up) vd = vu;
            if (rdm_up) rdm_up[(J - 1) * n + (K - 1)] = rdm_up[(K - 1) * n + (J - 1)] = vu;
            if (rdm_dn) rdm_dn[(J - 1) * n + (K - 1)] = rdm_dn[(K - 1) * n + (J - 1)] = vd;
        }
    if (ni)
        for (int J = 0; J < n; ++J) ni[J] = r.ni[J];
    // J_SPIN (static_rq.f90:323-395): K <= J, mirrored
    if (sisj)
        for (int J = 1; J <= n; ++J)
            for (int K = 1; K <= J; ++K) {
                const int jk = (J - 1) * n + (K - 1), kj = (K - 1) * n + (J - 1);
                double szsz = 0.25 * (r.nn_uu[jk] - r.nn_ud[jk] - r.nn_ud[kj] + r.nn_dd[jk]);
                double v;
                if (J == K) {
                    const double single = r.dens_up[J - 1] + r.dens_dn[J - 1] - 2.0 * r.ni[J - 1];
                    v = 0.5 * single + szsz;
                } else {
                    v = -r.flip[jk] + szsz;
                }
                sisj[jk] = sisj[kj] = v;
            }
    return PHUB_OK;
}

extern "C" int phub_one_rdm_spin(phub_ctx *c, const phub_op *op_, const double *bra,
                                 const double *ket, double *g_up, double *g_dn) {
    if (!c || !op_ || !bra || !ket) return phub_fail(c, PHUB_ERR_ARG, "null argument");
    if (op_->kind != 0) return phub_fail(c, PHUB_ERR_ARG, "not a Hubbard operator");
    phub_op *op = const_cast<phub_op *>(op_);
    HubRaw r;
    // hop(b = ket, k = bra)[K][J] = <bra| c^dag_J c_K |ket>
    int rc = hubbard_raw(c, op, ket, bra, r, false);
    if (rc) return rc;
    const int n = op->n;
    for (int J = 1; J <= n; ++J)
        for (int K = 1; K <= n; ++K) {
            const int jk = (J - 1) * n + (K - 1), kj = (K - 1) * n + (J - 1);
            if (g_up) g_up[jk] = (J == K) ? r.dens_up[J - 1] : r.hop_up[kj];
            if (g_dn) g_dn[jk] = (J == K) ? r.dens_dn[J - 1] : r.hop_dn[kj];
        }
    return PHUB_OK;
}

// <bra| dH/dT(J,K) |ket> = sum_s <bra| c^dag_K c_J |ket>, <bra| dH/dT(J,J) |ket> = sum_s <n_J>,
// <bra| dH/dU_J |ket> = <n_Ju n_Jd>.  hubbard_raw(b, k) contracts <k| O |b>.
extern "C" int phub_grad_hubbard(phub_ctx *c, const phub_op *op_, const double *bra,
                                 const double *ket, double *grad_t, double *grad_u) {
    if (!c || !op_ || !bra || !ket) return phub_fail(c, PHUB_ERR_ARG, "null argument");
    if (op_->kind != 0) return phub_fail(c, PHUB_ERR_ARG, "not a Hubbard operator");
    phub_op *op = const_cast<phub_op *>(op_);
    HubRaw r;
    int rc = hubbard_raw(c, op, ket, bra, r, false);
    if (rc) return rc;
    const int n = op->n;
    if (grad_t)
        for (int J = 1; J <= n; ++J)
            for (int K = 1; K <= n; ++K) {
                const int jk = (J - 1) * n + (K - 1);
                grad_t[jk] = (J == K) ? r.dens_up[J - 1] + r.dens_dn[J - 1] : r.hop_up[jk] + r.hop_dn[jk];
            }
    if (grad_u)
        for (int J = 0; J < n; ++J) grad_u[J] = r.ni[J];
    return PHUB_OK;
}

// Hellmann-Feynman: dE/dT(J,K) = sum_s <c^dag_K c_J>, dE/dT(J,J) = <n_J>, dE/dU_J = NI(J)
extern "C" int phub_hf_grad_hubbard(phub_ctx *c, const phub_op *op, const double *psi,
                                    double *grad_t, double *grad_u) {
    return phub_grad_hubbard(c, op, psi, psi, grad_t, grad_u);
}

// ---------------------------------------------------------------------------
// DBL_OCC of TWO_RDM (static_rq.f90:247-306), push form: the source (l1, i) with
// amplitude bra, every up move c^dag_K c_J i (K <= J, J = K the occupation) and down
// move c^dag_N c_M l1 (N <= M) adds s_up s_dn bra ket[target] to (J,K,M,N).
__global__ __launch_bounds__(256) void k_two_rdm(int n, int64_t nu, int64_t nd,
                                                 const uint32_t *__restrict__ mu,
                                                 const uint32_t *__restrict__ md,
                                                 const int32_t *__restrict__ cu,
                                                 const int32_t *__restrict__ iu,
                                                 const double *__restrict__ su,
                                                 const int16_t *__restrict__ pu,
                                                 const int32_t *__restrict__ cd,
                                                 const int32_t *__restrict__ id,
                                                 const double *__restrict__ sd,
                                                 const int16_t *__restrict__ pd, int maxh_d,
                                                 const double *__restrict__ bra,
                                                 const double *__restrict__ ket,
                                                 double *__restrict__ out) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nu * nd) return;
    const double b = bra[e];
    if (b == 0.0) return;
    const int64_t l1 = e / nu, i = e - l1 * nu;
    const uint32_t xu = mu[i], xd = md[l1];
    const int cntu = cu[i], cntd = cd[l1];
    for (int ia = 0; ia < cntu + n; ++ia) {
        int J, K;
        double s1;
        int64_t pos;
        if (ia < cntu) {
            const size_t o = (size_t)ia * nu + i;
            const int pr = pu[o];     // (J-1)*n + (K-1) of c^dag_K c_J
            J = pr / n;
            K = pr % n;
            if (K > J) continue;
            s1 = su[o];
            pos = iu[o];
        } else {
            const int site = ia - cntu;
            if (!((xu >> (n - 1 - site)) & 1u)) continue;
            J = K = site;
            s1 = 1.0;
            pos = i;
        }
        for (int ib = 0; ib < cntd + n; ++ib) {
            int M, N;
            double s2;
            int64_t pos2;
            if (ib < cntd) {
                const size_t o = (size_t)l1 * maxh_d + ib;
                const int q = pd[o];
                M = q / n;
                N = q % n;
                if (N > M) continue;
                s2 = sd[o];
                pos2 = id[o];
            } else {
                const int site = ib - cntd;
                if (!((xd >> (n - 1 - site)) & 1u)) continue;
                M = N = site;
                s2 = 1.0;
                pos2 = l1;
            }
            atomicAdd(&out[(((size_t)J * n + K) * n + M) * n + N], s1 * s2 * b * ket[pos2 * nu + pos]);
        }
    }
}

extern "C" int phub_two_rdm_hubbard(phub_ctx *c, const phub_op *op, const double *bra,
                                    const double *ket, double *out) {
    if (!c || !op || !bra || !ket || !out) return phub_fail(c, PHUB_ERR_ARG, "null argument");
    if (op->kind != 0) return phub_fail(c, PHUB_ERR_ARG, "not a Hubbard operator");
    phub_op *mop = const_cast<phub_op *>(op);
    int rc = phub_ensure_pair_tables(c, mop);
    if (rc) return rc;
    const int n = op->n;
    const size_t n4 = (size_t)n * n * n * n;
    struct { double *p; } d{nullptr};
    rc = phub_arena(c, n4 * sizeof(double), (void **)&d.p);
    if (rc) return rc;
    PHUB_HIP(c, hipMemsetAsync(d.p, 0, n4 * sizeof(double), c->stream));
    const int64_t nb = (op->dim + 255) / 256;
    if (nb > INT32_MAX) return phub_fail(c, PHUB_ERR_UNSUPPORTED, "grid too large");
    hipLaunchKernelGGL(k_two_rdm, dim3((unsigned)nb), dim3(256), 0, c->stream, n, op->dim_up,
                       op->dim_dn, op->up.masks, op->dn.masks, op->rup.cnt, op->rup.idx,
                       op->rup.c64, op->rup.pair, op->rdn.cnt, op->rdn.idx, op->rdn.c64,
                       op->rdn.pair, std::max(op->rdn.maxh, 1), bra, ket, d.p);
    PHUB_HIP(c, hipGetLastError());
    std::vector<double> h(n4);
    PHUB_HIP(c, hipMemcpyAsync(h.data(), d.p, n4 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    PHUB_HIP(c, hipStreamSynchronize(c->stream));
    // computed for K <= J, N <= M; the reference copies to the other orderings
    auto ix = [n](int a, int b, int cc, int dd) { return (((size_t)a * n + b) * n + cc) * n + dd; };
    for (int J = 0; J < n; ++J)
        for (int K = 0; K <= J; ++K)
            for (int M = 0; M < n; ++M)
                for (int N = 0; N <= M; ++N) {
                    const double v = h[ix(J, K, M, N)];
                    out[ix(J, K, M, N)] = v;
                    out[ix(J, K, N, M)] = v;
                    out[ix(K, J, M, N)] = v;
                    out[ix(K, J, N, M)] = v;
                }
    return PHUB_OK;
}

// Heisenberg J_SPIN (static_rq_heis.f90:58-116) and dE/dJ
// ---------------------------------------------------------------------------
// flip[b] = sum over states i with site J up, K down of bra[pos] * ket[i]
// (pos = flipped state), for bond b = {J<K}; and the reverse (K up, J down).
// zz[b] = sum_i bra[i]*ket[i] * (same spin ? +1 : -1); norm = sum bra*ket.
__global__ __launch_bounds__(256) void k_heis_bonds(int64_t dim, int L,
                                                    const uint64_t *__restrict__ masks,
                                                    const int64_t *__restrict__ first,
                                                    const uint32_t *__restrict__ pos, int nb,
                                                    const int *__restrict__ bj,
                                                    const int *__restrict__ bk,
                                                    const double *__restrict__ bra,
                                                    const double *__restrict__ ket,
                                                    double *__restrict__ out) {
    extern __shared__ double acc[];   // 3*nb + 1
    __shared__ double sh[4];
    for (int i = threadIdx.x; i < 3 * nb + 1; i += blockDim.x) acc[i] = 0.0;
    __syncthreads();
    const uint64_t lmask = (1ull << L) - 1ull;
    double nrm = 0.0;
    for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < dim;
         s += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t x = masks[s];
        const double k = ket[s];
        const double w = bra[s] * k;
        nrm += w;
        for (int b = 0; b < nb; ++b) {
            const int j = bj[b], kk = bk[b];
            const bool diff = ((x >> j) ^ (x >> kk)) & 1ull;
            atomicAdd(&acc[3 * b + 2], diff ? -w : w);
            if (diff) {
                // the index of x2: its low-half block's first index plus the position
                // of its high half within the block
                const uint64_t x2 = x ^ ((1ull << j) | (1ull << kk));
                const int64_t i2 = first[x2 & lmask] + (int64_t)pos[x2 >> L];
                // source i = s has J up (bit j set): contributes to ordered (J,K)
                atomicAdd(&acc[3 * b + (((x >> j) & 1ull) ? 0 : 1)], bra[i2] * k);
            }
        }
    }
    nrm = wsum(nrm);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = nrm;
    __syncthreads();
    if (threadIdx.x == 0) acc[3 * nb] += sh[0] + sh[1] + sh[2] + sh[3];
    __syncthreads();
    for (int i = threadIdx.x; i < 3 * nb + 1; i += blockDim.x) atomicAdd(&out[i], acc[i]);
}

static int heis_bonds(phub_ctx *c, const phub_op *op, const double *bra, const double *ket,
                      std::vector<double> &h) {
    const HeisTables &t = op->h;
    if (!t.masks)
        return phub_fail(c, PHUB_ERR_UNSUPPORTED, "spin correlations need dim <= 2^28 (mask table)");
    const int nb = t.nb;
    struct { double *p; } d{nullptr};
    int rc = phub_arena(c, (3 * nb + 1) * sizeof(double), (void **)&d.p);
    if (rc) return rc;
    PHUB_HIP(c, hipMemsetAsync(d.p, 0, (3 * nb + 1) * sizeof(double), c->stream));
    const unsigned grid = (unsigned)std::max<int64_t>(1, std::min<int64_t>((t.dim + 255) / 256, 2048));
    hipLaunchKernelGGL(k_heis_bonds, dim3(grid), dim3(256), (3 * nb + 1) * sizeof(double),
                       c->stream, t.dim, t.stg.L, t.masks, t.stg.first, t.stg.pos, nb,
                       t.bj, t.bk, bra, ket, d.p);
    PHUB_HIP(c, hipGetLastError());
    h.resize(3 * nb + 1);
    PHUB_HIP(c, hipMemcpyAsync(h.data(), d.p, (3 * nb + 1) * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    PHUB_HIP(c, hipStreamSynchronize(c->stream));
    return PHUB_OK;
}

// The all-pairs view of a Heisenberg operator (every J < K a bond, same sign
// convention): built on the first spin-correlation call and kept with the
// operator, so later calls allocate nothing.
static int heis_all_pairs(phub_ctx *c, const phub_op *op, const phub_op **out) {
    phub_op *mop = const_cast<phub_op *>(op);
    if (!mop->heis_all) {
        const int n = op->n;
        std::vector<double> ones(n * n, 1.0);
        for (int i = 0; i < n; ++i) ones[i * n + i] = 0.0;
        phub_heis_desc hd{n, ones.data(), op->sign_convention};
        int rc = phub_op_heisenberg(c, &hd, &mop->heis_all);
        if (rc) return rc;
    }
    *out = mop->heis_all;
    return PHUB_OK;
}

// host copy of the bond list in the same order phub_build_heis uses
static void bond_list(const phub_op *op, std::vector<int> &J1, std::vector<int> &K1) {
    const int n = op->n;
    for (int J = 1; J <= n; ++J)
        for (int K = J + 1; K <= n; ++K) {
            const double jk = op->J[(J - 1) * n + (K - 1)], kj = op->J[(K - 1) * n + (J - 1)];
            if (jk == 0.0 && kj == 0.0) continue;
            J1.push_back(J);
            K1.push_back(K);
        }
}

extern "C" int phub_rdm_heisenberg(phub_ctx *c, const phub_op *op, const double *bra,
                                   const double *ket, double *sisj) {
    if (!c || !op || !bra || !ket || !sisj) return phub_fail(c, PHUB_ERR_ARG, "null argument");
    if (op->kind != 1) return phub_fail(c, PHUB_ERR_ARG, "not a Heisenberg operator");
    // J_SPIN runs over every ordered pair; pairs without a bond still need their
    // <S_J.S_K>: computed with the operator's all-pairs view
    const int n = op->n;
    const phub_op *all = nullptr;
    int rc = heis_all_pairs(c, op, &all);
    if (rc) return rc;
    std::vector<double> h;
    rc = heis_bonds(c, all, bra, ket, h);
    if (rc) return rc;
    std::vector<int> J1, K1;
    for (int J = 1; J <= n; ++J)
        for (int K = J + 1; K <= n; ++K) { J1.push_back(J); K1.push_back(K); }
    const double nrm = h[3 * (int)J1.size()];
    for (int J = 0; J < n; ++J) sisj[J * n + J] = 0.75 * nrm;   // 0.5*P(single) + 0.25
    for (size_t b = 0; b < J1.size(); ++b) {
        const int J = J1[b], K = K1[b];
        // flip term: -AC * bra(pos) * ket(i), AC = -(-1)^{J+K} (four ANTICOMs)
        const double sgn = ((J + K) & 1) ? -1.0 : 1.0;   // (-1)^{J+K}
        const double zz = 0.25 * h[3 * b + 2];
        sisj[(J - 1) * n + (K - 1)] = sgn * h[3 * b + 0] + zz;   // J up in the ket state
        sisj[(K - 1) * n + (J - 1)] = sgn * h[3 * b + 1] + zz;
    }
    return PHUB_OK;
}

// dE/dJ(J,K) for the operator as built (both sign conventions), ordered pairs
extern "C" int phub_hf_grad_heisenberg(phub_ctx *c, const phub_op *op, const double *psi,
                                       double *grad_j) {
    if (!c || !op || !psi || !grad_j) return phub_fail(c, PHUB_ERR_ARG, "null argument");
    if (op->kind != 1) return phub_fail(c, PHUB_ERR_ARG, "not a Heisenberg operator");
    const int n = op->n;
    const phub_op *all = nullptr;
    int rc = heis_all_pairs(c, op, &all);
    if (rc) return rc;
    std::vector<double> h;
    rc = heis_bonds(c, all, psi, psi, h);
    if (rc) return rc;
    const double nrm = h[h.size() - 1];
    int b = 0;
    for (int J = 1; J <= n; ++J) grad_j[(J - 1) * n + (J - 1)] = 0.75 * nrm;
    for (int J = 1; J <= n; ++J)
        for (int K = J + 1; K <= n; ++K, ++b) {
            const double pr = ((J + K) & 1) ? -1.0 : 1.0;
            const double fs = (op->sign_convention == 0) ? -pr : pr;   // flip element sign
            // H_JK = J(J,K) [sz sz + flip(source J up)] ; derivative = <.>
            // in pull form element(source J up -> target) = fs*J(J,K): <psi|.|psi> =
            // fs * sum_{i: J up} psi[pos] psi[i]
            const double zz = 0.25 * h[3 * b + 2];
            grad_j[(J - 1) * n + (K - 1)] = fs * h[3 * b + 0] + zz;
            grad_j[(K - 1) * n + (J - 1)] = fs * h[3 * b + 1] + zz;
        }
    return PHUB_OK;
}
