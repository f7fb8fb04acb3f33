// hv.hip -- on-the-fly H*v for the Hubbard and Heisenberg operators.
//
// Hubbard (reference HPROD, pyhub/libs/src/hprod_hubbard.f90:17-118):
//   the state is a dim_dn x dim_up matrix V[d][u] (u fastest, index d*dim_up+u).
//   w[d][u] = (sum_J U_J n_Ju n_Jd + eps_u + eps_d) v[d][u]            (diagonal, :27-35)
//           + sum_{u'} A_up[u][u'] v[d][u']                             (up hops,  :38-61)
//           + sum_{d'} A_dn[d][d'] v[d'][u]                             (down hops, :93-114)
//   written in pull form (each output element is owned by one lane: no atomics).
//   Up-hop targets vary per lane (k-major table, coalesced table reads, gathers
//   inside the row d).  Down-hop targets are uniform over a workgroup's row, so
//   they are scalar table loads and fully coalesced vector loads of other rows.
//   A workgroup owns R rows x 256 columns: each lane reads its up-hop list once
//   and applies it to R rows.
// Heisenberg (reference HPROD, pyhub/libs/src/hprod_heis.f90:1-61): one lane per
//   state, bonds uniform, the flipped state's index from two-half colex tables.
//
// Fused Lanczos epilogue (optional): out = scale*(H x) - beta*prev and per-block
// partial sums <scale*x, out>, <out, out> (FP64) written to the context's
// reduction buffer.
#include "phub_internal.h"

#include <algorithm>

template <typename T>
struct HubArgs {
    int64_t dim_up, dim_dn;
    const uint32_t *mu, *md;
    const double *eu, *ed;
    const int32_t *cu, *iu;
    const T *cfu;
    const int32_t *cd, *id;
    const T *cfd;
    int maxh_dn;
    const int32_t *rp_dn;     // CSR row pointers of the down table (compressed entries)
    const uint32_t *csr_dn;
    const T *cls_dn;
    int ncls_dn;
    int n;
    int u_uniform;
    double u0;
    double U[32];
};

template <typename T>
__device__ __forceinline__ void block_sum2(double a, double b, double *part) {
    __shared__ double sh[2][8];
    for (int off = 32; off > 0; off >>= 1) {
        a += __shfl_xor(a, off);
        b += __shfl_xor(b, off);
    }
    int wv = threadIdx.x >> 6, ln = threadIdx.x & 63;
    if (ln == 0) { sh[0][wv] = a; sh[1][wv] = b; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double s0 = 0.0, s1 = 0.0;
        for (int i = 0; i < (int)(blockDim.x >> 6); ++i) { s0 += sh[0][i]; s1 += sh[1][i]; }
        part[2 * (size_t)blockIdx.x] = s0;
        part[2 * (size_t)blockIdx.x + 1] = s1;
    }
}

template <typename T, int R, bool EPI>
__global__ __launch_bounds__(256) void k_hv_hub(HubArgs<T> a, const T *__restrict__ x,
                                                T *__restrict__ out, double scale,
                                                const T *__restrict__ prev, double beta,
                                                double *__restrict__ part, int ncb) {
    const int cb = blockIdx.x % ncb;
    const int64_t d0 = (int64_t)(blockIdx.x / ncb) * R;
    const int64_t u = (int64_t)cb * 256 + threadIdx.x;
    const int64_t nu = a.dim_up;
    double s_xo = 0.0, s_oo = 0.0;
    if (u < nu) {
        const uint32_t xu = a.mu[u];
        const double eu = a.eu[u];
        T acc[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int64_t d = d0 + r;
            acc[r] = (T)0;
            if (d < a.dim_dn) {
                const uint32_t both = xu & a.md[d];
                double dg = eu + a.ed[d];
                if (a.u_uniform) {
                    dg += a.u0 * (double)__popc(both);
                } else {
                    uint32_t bb = both;
                    while (bb) {
                        int b = __ffs(bb) - 1;
                        dg += a.U[a.n - 1 - b];
                        bb &= bb - 1;
                    }
                }
                acc[r] = (T)dg * x[d * nu + u];
            }
        }
        // up hops: lane-varying targets inside each row
        const int cnt = a.cu[u];
        for (int k = 0; k < cnt; ++k) {
            const int32_t j = a.iu[(size_t)k * nu + u];
            const T c = a.cfu[(size_t)k * nu + u];
#pragma unroll
            for (int r = 0; r < R; ++r)
                if (d0 + r < a.dim_dn) acc[r] += c * x[(d0 + r) * nu + j];
        }
        // down hops: row-uniform targets, coalesced rows
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int64_t d = d0 + r;
            if (d < a.dim_dn) {
                const int cd = a.cd[d];
                const int32_t *idr = a.id + (size_t)d * a.maxh_dn;
                const T *cdr = a.cfd + (size_t)d * a.maxh_dn;
                T s = acc[r];
                for (int k = 0; k < cd; ++k) s += cdr[k] * x[(int64_t)idr[k] * nu + u];
                const size_t o = (size_t)d * nu + u;
                if (EPI) {
                    double val = scale * (double)s;
                    if (prev) val -= beta * (double)prev[o];
                    const T vt = (T)val;
                    out[o] = vt;
                    const double xs = scale * (double)x[o];
                    s_xo += xs * (double)vt;
                    s_oo += (double)vt * (double)vt;
                } else {
                    out[o] = s;
                }
            }
        }
    }
    if (EPI) block_sum2<T>(s_xo, s_oo, part);
}

// ---------------------------------------------------------------------------
// Two-kernel Hubbard H*v (used when R rows of the state fit in LDS):
//  k_hv_up : up hops + diagonal.  A workgroup stages R rows d0..d0+R-1 in LDS;
//            each lane walks its columns u, reads each compressed hop entry
//            (idx << 8 | class) once and gathers from the R staged rows.
//            out[d][u] = diag * x[d][u] + sum_up.
//  k_hv_dn : down hops + fused epilogue.  Each XCD owns 64-column slabs of the
//            state (workgroup b runs on XCD b % 8; slabs are dealt so that the
//            XCD's concurrent workgroups share one slab, which then stays in
//            its L2).  A wave takes one row d: its hop list is wave-uniform
//            (scalar loads) and every gather is a coalesced 64-lane row segment.
//            out = scale * (out + sum_dn) - beta * prev, dot partials.
// ---------------------------------------------------------------------------
constexpr int UP_THREADS = 1024;
constexpr int DN_ROWS = 16;     // rows per workgroup in k_hv_dn: many workgroups per slab
constexpr int UP_UNROLL = 8;    // independent hop-entry loads in flight per lane
constexpr int DN_UNROLL = 8;    // independent row gathers in flight per lane

template <typename T, int R>
__global__ __launch_bounds__(UP_THREADS) void k_hv_up(HubArgs<T> a,
                                                      const uint32_t *__restrict__ packed,
                                                      const T *__restrict__ cls, int ncls,
                                                      const T *__restrict__ x,
                                                      T *__restrict__ out) {
    extern __shared__ double smem_d[];
    T *rows = reinterpret_cast<T *>(smem_d);
    const int64_t nu = a.dim_up;
    T *cl = rows + R * nu;
    const int64_t d0 = (int64_t)blockIdx.x * R;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const bool ok = d0 + r < a.dim_dn;
        const T *src = x + (d0 + r) * nu;
        for (int64_t u = threadIdx.x; u < nu; u += UP_THREADS)
            rows[r * nu + u] = ok ? __builtin_nontemporal_load(src + u) : (T)0;
    }
    for (int i = threadIdx.x; i < ncls; i += UP_THREADS) cl[i] = cls[i];
    __syncthreads();
    uint32_t xd[R];
    double ed[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const bool ok = d0 + r < a.dim_dn;
        xd[r] = ok ? a.md[d0 + r] : 0u;
        ed[r] = ok ? a.ed[d0 + r] : 0.0;
    }
    for (int64_t u = threadIdx.x; u < nu; u += UP_THREADS) {
        const uint32_t xu = a.mu[u];
        const double eu = a.eu[u];
        T acc[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const uint32_t both = xu & xd[r];
            double dg = eu + ed[r];
            if (a.u_uniform) {
                dg += a.u0 * (double)__popc(both);
            } else {
                uint32_t bb = both;
                while (bb) {
                    const int b = __ffs(bb) - 1;
                    dg += a.U[a.n - 1 - b];
                    bb &= bb - 1;
                }
            }
            acc[r] = (T)dg * rows[r * nu + u];
        }
        const int cnt = a.cu[u];
        const uint32_t *pe = packed + u;
        int k = 0;
        for (; k + UP_UNROLL <= cnt; k += UP_UNROLL) {
            uint32_t e[UP_UNROLL];
#pragma unroll
            for (int q = 0; q < UP_UNROLL; ++q) e[q] = pe[(size_t)(k + q) * nu];
#pragma unroll
            for (int q = 0; q < UP_UNROLL; ++q) {
                const uint32_t j = e[q] >> 8;
                const T c = cl[e[q] & 255u];
#pragma unroll
                for (int r = 0; r < R; ++r) acc[r] += c * rows[r * nu + j];
            }
        }
        for (; k < cnt; ++k) {
            const uint32_t e = pe[(size_t)k * nu];
            const uint32_t j = e >> 8;
            const T c = cl[e & 255u];
#pragma unroll
            for (int r = 0; r < R; ++r) acc[r] += c * rows[r * nu + j];
        }
#pragma unroll
        for (int r = 0; r < R; ++r)
            if (d0 + r < a.dim_dn) __builtin_nontemporal_store(acc[r], out + (d0 + r) * nu + u);
    }
}

// COLS columns per row segment (64 FP32 = 256 B, 32 FP64 = 256 B): a wave covers
// 64/COLS rows at once.
template <typename T, bool EPI, int COLS>
__global__ __launch_bounds__(256) void k_hv_dn(HubArgs<T> a, const T *__restrict__ x,
                                               T *__restrict__ out, double scale,
                                               const T *__restrict__ prev, double beta,
                                               double *__restrict__ part, int nchunks, int nrb) {
    constexpr int RPW = 64 / COLS;
    const int b = blockIdx.x;
    const int xcd = b & 7, i = b >> 3;
    const int chunk = (i / nrb) * 8 + xcd;
    const int rb = i % nrb;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int sub = lane / COLS, cidx = lane % COLS;
    const int64_t nu = a.dim_up;
    const int64_t col = (int64_t)chunk * COLS + cidx;
    const bool colok = (chunk < nchunks) && (col < nu);
    __shared__ T cls[256];
    for (int q = threadIdx.x; q < a.ncls_dn; q += 256) cls[q] = a.cls_dn[q];
    __syncthreads();
    double s_xo = 0.0, s_oo = 0.0;
    if (colok) {
        for (int rr = wave * RPW + sub; rr < DN_ROWS; rr += 4 * RPW) {
            const int64_t d = (int64_t)rb * DN_ROWS + rr;
            if (d >= a.dim_dn) break;
            const int32_t r0 = a.rp_dn[d], r1 = a.rp_dn[d + 1];
            const T *xc = x + col;
            T s = (T)0;
            int32_t k = r0;
            for (; k + DN_UNROLL <= r1; k += DN_UNROLL) {
                uint32_t e[DN_UNROLL];
                T v[DN_UNROLL];
#pragma unroll
                for (int q = 0; q < DN_UNROLL; ++q) e[q] = a.csr_dn[k + q];
#pragma unroll
                for (int q = 0; q < DN_UNROLL; ++q) v[q] = xc[(int64_t)(e[q] >> 8) * nu];
#pragma unroll
                for (int q = 0; q < DN_UNROLL; ++q) s += cls[e[q] & 255u] * v[q];
            }
            for (; k < r1; ++k) {
                const uint32_t e = a.csr_dn[k];
                s += cls[e & 255u] * xc[(int64_t)(e >> 8) * nu];
            }
            const size_t o = (size_t)d * nu + col;
            if (EPI) {
                // streamed once: non-temporal, so the slab keeps the L2
                double val = scale * ((double)__builtin_nontemporal_load(out + o) + (double)s);
                if (prev) val -= beta * (double)__builtin_nontemporal_load(prev + o);
                const T vt = (T)val;
                __builtin_nontemporal_store(vt, out + o);
                s_xo += scale * (double)x[o] * (double)vt;
                s_oo += (double)vt * (double)vt;
            } else {
                __builtin_nontemporal_store(__builtin_nontemporal_load(out + o) + s, out + o);
            }
        }
    }
    if (EPI) block_sum2<T>(s_xo, s_oo, part);
}

struct HeisArgs {
    int64_t dim;
    int n, L, nb;
    uint64_t total_m1;   // C(n, n/2) - 1
    const uint64_t *masks, *tabA, *tabB;
    const int *bj, *bk;
    const double *bJ, *bJr, *bsign, *ising;
    double dconst;
};

template <typename T, bool EPI>
__global__ __launch_bounds__(256) void k_hv_heis(HeisArgs a, const T *__restrict__ x,
                                                 T *__restrict__ out, double scale,
                                                 const T *__restrict__ prev, double beta,
                                                 double *__restrict__ part) {
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double s_xo = 0.0, s_oo = 0.0;
    if (s < a.dim) {
        const uint64_t xs = a.masks[s];
        const uint64_t lmask = (1ull << a.L) - 1ull;
        double dg = a.dconst;
        T acc = (T)0;
        for (int b = 0; b < a.nb; ++b) {
            const int bj = a.bj[b], bk = a.bk[b];
            if (((xs >> bj) ^ (xs >> bk)) & 1ull) {
                dg -= 0.5 * a.ising[b];
                const uint64_t x2 = xs ^ ((1ull << bj) | (1ull << bk));
                const uint64_t xh = x2 >> a.L;
                const uint64_t colex = a.tabA[xh] + a.tabB[((uint64_t)__popcll(xh) << a.L) + (x2 & lmask)];
                const int64_t r = (int64_t)(a.total_m1 - colex);
                const double cf = a.bsign[b] * (((xs >> bj) & 1ull) ? a.bJr[b] : a.bJ[b]);
                acc += (T)cf * x[r];
            }
        }
        acc += (T)dg * x[s];
        if (EPI) {
            double val = scale * (double)acc;
            if (prev) val -= beta * (double)prev[s];
            const T vt = (T)val;
            out[s] = vt;
            s_xo += scale * (double)x[s] * (double)vt;
            s_oo += (double)vt * (double)vt;
        } else {
            out[s] = acc;
        }
    }
    if (EPI) block_sum2<T>(s_xo, s_oo, part);
}

template <typename T>
static HubArgs<T> hub_args(const phub_op *op) {
    HubArgs<T> a{};
    a.dim_up = op->dim_up;
    a.dim_dn = op->dim_dn;
    a.mu = op->up.masks;
    a.md = op->dn.masks;
    a.eu = op->up.eloc;
    a.ed = op->dn.eloc;
    a.cu = op->up.cnt;
    a.iu = op->up.idx;
    a.cd = op->dn.cnt;
    a.id = op->dn.idx;
    if constexpr (sizeof(T) == 8) {
        a.cfu = (const T *)op->up.c64;
        a.cfd = (const T *)op->dn.c64;
    } else {
        a.cfu = (const T *)op->up.c32;
        a.cfd = (const T *)op->dn.c32;
    }
    a.maxh_dn = std::max(op->dn.maxh, 1);
    a.rp_dn = op->dn.rowptr;
    a.csr_dn = op->dn.csr;
    a.cls_dn = (sizeof(T) == 8) ? (const T *)op->dn.cls64 : (const T *)op->dn.cls32;
    a.ncls_dn = op->dn.ncls;
    a.n = op->n;
    a.u_uniform = op->u_uniform;
    a.u0 = op->u0;
    for (int i = 0; i < 32; ++i) a.U[i] = (i < op->n) ? op->U[i] : 0.0;
    return a;
}

static HeisArgs heis_args(const phub_op *op) {
    const HeisTables &h = op->h;
    HeisArgs a{};
    a.dim = h.dim;
    a.n = h.n;
    a.L = h.L;
    a.nb = h.nb;
    a.total_m1 = (uint64_t)h.dim - 1;
    a.masks = h.masks;
    a.tabA = h.tabA;
    a.tabB = h.tabB;
    a.bj = h.bj;
    a.bk = h.bk;
    a.bJ = h.bJ;
    a.bJr = h.bJr;
    a.bsign = h.bsign;
    a.ising = h.bond_ising;
    a.dconst = h.diag_const;
    return a;
}

constexpr int HUB_R = 4;

template <typename T>
int phub_hv_launch(phub_ctx *c, const phub_op *op, const T *x, T *out, double scale,
                   const T *prev, double beta, bool want_dots, int *nblocks_out) {
    int nblocks = 0;
    // the epilogue kernel applies scale / prev and writes dot partials
    const bool epi = want_dots || prev != nullptr || scale != 1.0;
    if (op->kind == 0 && op->up.packed && op->dn.csr) {
        // two-kernel path: R rows of the state in LDS for the up hops
        constexpr int R = sizeof(T) == 4 ? 2 : 1;
        const size_t lds = ((size_t)R * op->dim_up + op->up.ncls) * sizeof(T);
        if (lds <= 150 * 1024) {
            HubArgs<T> a = hub_args<T>(op);
            static bool attr_set[2] = {false, false};
            if (!attr_set[sizeof(T) == 8]) {
                PHUB_HIP(c, hipFuncSetAttribute((const void *)k_hv_up<T, R>,
                                                hipFuncAttributeMaxDynamicSharedMemorySize,
                                                160 * 1024));
                attr_set[sizeof(T) == 8] = true;
            }
            const T *cls = (sizeof(T) == 8) ? (const T *)op->up.cls64 : (const T *)op->up.cls32;
            const unsigned g1 = (unsigned)((op->dim_dn + R - 1) / R);
            hipLaunchKernelGGL((k_hv_up<T, R>), dim3(g1), dim3(UP_THREADS), lds, c->stream, a,
                               op->up.packed, cls, op->up.ncls, x, out);
            PHUB_HIP(c, hipGetLastError());
            constexpr int COLS = sizeof(T) == 4 ? 32 : 16;
            const int nchunks = (int)((op->dim_up + COLS - 1) / COLS);
            const int nrb = (int)((op->dim_dn + DN_ROWS - 1) / DN_ROWS);
            const int64_t g2 = (int64_t)8 * ((nchunks + 7) / 8) * nrb;
            if (g2 > INT32_MAX) return phub_fail(c, PHUB_ERR_UNSUPPORTED, "grid too large");
            nblocks = (int)g2;
            if (epi) {
                int rc = phub_ensure_red(c, (size_t)nblocks * 2 + 64);
                if (rc) return rc;
                hipLaunchKernelGGL((k_hv_dn<T, true, COLS>), dim3(nblocks), dim3(256), 0, c->stream, a,
                                   x, out, scale, prev, beta, c->d_red, nchunks, nrb);
            } else {
                hipLaunchKernelGGL((k_hv_dn<T, false, COLS>), dim3(nblocks), dim3(256), 0, c->stream,
                                   a, x, out, scale, prev, beta, c->d_red, nchunks, nrb);
            }
            PHUB_HIP(c, hipGetLastError());
            if (nblocks_out) *nblocks_out = nblocks;
            return PHUB_OK;
        }
    }
    if (op->kind == 0) {
        if (op->n > 32) return phub_fail(c, PHUB_ERR_UNSUPPORTED, "Hubbard H*v needs n <= 32");
        const int ncb = (int)((op->dim_up + 255) / 256);
        const int64_t nrg = (op->dim_dn + HUB_R - 1) / HUB_R;
        const int64_t nb = (int64_t)ncb * nrg;
        if (nb > INT32_MAX) return phub_fail(c, PHUB_ERR_UNSUPPORTED, "grid too large");
        nblocks = (int)nb;
        if (epi) {
            int rc = phub_ensure_red(c, (size_t)nblocks * 2 + 64);
            if (rc) return rc;
        }
        HubArgs<T> a = hub_args<T>(op);
        if (epi)
            hipLaunchKernelGGL((k_hv_hub<T, HUB_R, true>), dim3(nblocks), dim3(256), 0, c->stream,
                               a, x, out, scale, prev, beta, c->d_red, ncb);
        else
            hipLaunchKernelGGL((k_hv_hub<T, HUB_R, false>), dim3(nblocks), dim3(256), 0,
                               c->stream, a, x, out, scale, prev, beta, c->d_red, ncb);
    } else {
        const int64_t nb = (op->dim + 255) / 256;
        if (nb > INT32_MAX) return phub_fail(c, PHUB_ERR_UNSUPPORTED, "grid too large");
        nblocks = (int)nb;
        if (epi) {
            int rc = phub_ensure_red(c, (size_t)nblocks * 2 + 64);
            if (rc) return rc;
        }
        HeisArgs a = heis_args(op);
        if (epi)
            hipLaunchKernelGGL((k_hv_heis<T, true>), dim3(nblocks), dim3(256), 0, c->stream, a,
                               x, out, scale, prev, beta, c->d_red);
        else
            hipLaunchKernelGGL((k_hv_heis<T, false>), dim3(nblocks), dim3(256), 0, c->stream, a,
                               x, out, scale, prev, beta, c->d_red);
    }
    PHUB_HIP(c, hipGetLastError());
    if (nblocks_out) *nblocks_out = nblocks;
    return PHUB_OK;
}

template int phub_hv_launch<double>(phub_ctx *, const phub_op *, const double *, double *, double,
                                    const double *, double, bool, int *);
template int phub_hv_launch<float>(phub_ctx *, const phub_op *, const float *, float *, double,
                                   const float *, double, bool, int *);

template <typename T>
static int hv_entry(phub_ctx *c, const phub_op *op, const T *v, T *w) {
    if (!c || !op || !v || !w) return phub_fail(c, PHUB_ERR_ARG, "null argument");
    if ((const void *)v == (const void *)w) return phub_fail(c, PHUB_ERR_ARG, "w must not alias v");
    if (op->dim == 0) return PHUB_OK;
    PHUB_HIP(c, hipEventRecord(c->ev0, c->stream));
    int rc = phub_hv_launch<T>(c, op, v, w, 1.0, (const T *)nullptr, 0.0, false, nullptr);
    if (rc) return rc;
    PHUB_HIP(c, hipEventRecord(c->ev1, c->stream));
    PHUB_HIP(c, hipEventSynchronize(c->ev1));
    float ms = 0.f;
    PHUB_HIP(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->last_ms = ms;
    return PHUB_OK;
}

extern "C" int phub_hv_f64(phub_ctx *c, const phub_op *op, const double *v, double *w) {
    return hv_entry<double>(c, op, v, w);
}

extern "C" int phub_hv_f32(phub_ctx *c, const phub_op *op, const float *v, float *w) {
    return hv_entry<float>(c, op, v, w);
}
