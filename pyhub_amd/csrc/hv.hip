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
// Heisenberg (reference HPROD, pyhub/libs/src/hprod_heis.f90:1-61): one workgroup
//   per block of states sharing their low-half pattern, the block and one
//   neighbouring block at a time staged in LDS, no mask table (k_hv_heis_stg below).
//
// Fused Lanczos epilogue (optional): out = scale*(H x) - beta*prev and per-block
// partial sums <scale*x, out>, <out, out> (FP64) written to the context's
// reduction buffer.
#include "phub_internal.h"

#include <algorithm>
#include <cstdio>
#include <cstdlib>

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
    // when set, the fused epilogue takes its scale and beta from device memory
    // (dsc[0], dsc[1]) instead of the kernel arguments: graph-captured Lanczos steps
    const double *dsc;
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
                    const double sc = a.dsc ? a.dsc[0] : scale, bt = a.dsc ? a.dsc[1] : beta;
                    double val = sc * (double)s;
                    if (prev) val -= bt * (double)prev[o];
                    const T vt = (T)val;
                    out[o] = vt;
                    const double xs = sc * (double)x[o];
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
constexpr int UP_CLS_BYTES = 256 * 8;   // coefficient slots ahead of the rows in LDS
constexpr int UP_LDS_MAX = 159 * 1024;  // coefficient slots + staged rows (of 160 KB per CU)
constexpr int UP_UNROLL = 8;    // independent hop-entry loads in flight per lane

template <typename T, int R>
__global__ __launch_bounds__(UP_THREADS) void k_hv_up(HubArgs<T> a,
                                                      const uint32_t *__restrict__ packed,
                                                      const T *__restrict__ cls, int ncls,
                                                      const T *__restrict__ x,
                                                      T *__restrict__ out) {
    // LDS: 256 coefficient slots of 8 bytes (UP_CLS_BYTES), then the R rows
    // interleaved: element u of row r at rows[u * R + r], so one gather of column j
    // fetches all R rows (a single 64-bit LDS read for two FP32 rows).  Three FP32
    // rows: rows 0 and 1 interleaved as for two, row 2 in an array of its own after
    // them (a 64-bit and a 32-bit read per column); every hop entry then serves three
    // rows, so the entry table streams through L2 a third less often
    // a static array (not dynamic shared memory): its LDS address is a compile-time
    // constant, so both decoded offsets go straight into the ds_read address
    __shared__ __attribute__((aligned(16))) double smem_up[UP_LDS_MAX / 8];
    char *lds = reinterpret_cast<char *>(smem_up);
    T *rows = reinterpret_cast<T *>(lds + UP_CLS_BYTES);
    const int nu = (int)a.dim_up;
    const int64_t d0 = (int64_t)blockIdx.x * R;
    for (int u = threadIdx.x; u < nu; u += UP_THREADS) {
        T v[R];
#pragma unroll
        for (int r = 0; r < R; ++r)
            v[r] = (d0 + r < a.dim_dn) ? __builtin_nontemporal_load(x + (d0 + r) * (int64_t)nu + u) : (T)0;
        if constexpr (R == 3) {
            reinterpret_cast<float2 *>(rows)[u] = make_float2(v[0], v[1]);
            rows[2 * nu + u] = v[2];
        } else if constexpr (R == 2 && sizeof(T) == 4) {
            reinterpret_cast<float2 *>(rows)[u] = make_float2(v[0], v[1]);
        } else {
#pragma unroll
            for (int r = 0; r < R; ++r) rows[u * R + r] = v[r];
        }
    }
    for (int i = threadIdx.x; i < ncls; i += UP_THREADS) *reinterpret_cast<T *>(lds + 8 * i) = cls[i];
    __syncthreads();
    // one column of the R staged rows is 8 bytes (two FP32 or one FP64) or 16 bytes
    // (two FP64); a packed entry carries both LDS byte offsets for 8-byte columns:
    // e = (j * 8) << 11 | class * 8, so the decode is one shift and one mask (one
    // more shift for 16-byte columns)
    static_assert(R * sizeof(T) == 8 || R * sizeof(T) == 16 || (R == 3 && sizeof(T) == 4),
                  "8- or 16-byte LDS columns, or three FP32 rows");
    constexpr int CSH = R * sizeof(T) == 16 ? 1 : 0;   // column offset = (j * 8) << CSH
    const char *rowb = lds + UP_CLS_BYTES;
    const char *rowb2 = rowb + 8 * (size_t)nu;          // row 2 of three FP32 rows: 4 j

    auto gather = [&](uint32_t off, T *v) {
        if constexpr (R == 3) {
            const float2 p = *reinterpret_cast<const float2 *>(rowb + off);
            v[0] = p.x;
            v[1] = p.y;
            v[2] = *reinterpret_cast<const float *>(rowb2 + (off >> 1));
        } else if constexpr (R == 2 && sizeof(T) == 4) {
            const float2 p = *reinterpret_cast<const float2 *>(rowb + off);
            v[0] = p.x;
            v[1] = p.y;
        } else if constexpr (R == 2) {
            const double2 p = *reinterpret_cast<const double2 *>(rowb + off);
            v[0] = p.x;
            v[1] = p.y;
        } else {
            v[0] = *reinterpret_cast<const T *>(rowb + off);
        }
    };
    uint32_t xd[R];
    double ed[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const bool ok = d0 + r < a.dim_dn;
        xd[r] = ok ? a.md[d0 + r] : 0u;
        ed[r] = ok ? a.ed[d0 + r] : 0.0;
    }
    for (int u = threadIdx.x; u < nu; u += UP_THREADS) {
        const uint32_t xu = a.mu[u];
        const double eu = a.eu[u];
        T acc[R], self[R];
        gather(((uint32_t)u * 8u) << CSH, self);
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
            acc[r] = (T)dg * self[r];
        }
        const int cnt = a.cu[u];
        // row k of the k-major entry table is a wave-uniform (SGPR) base and the
        // lane's column a 32-bit byte offset: saddr loads, no per-entry address VALU
        const uint32_t uoff = (uint32_t)u * 4u;
        int k = 0;
        for (; k + UP_UNROLL <= cnt; k += UP_UNROLL) {
            uint32_t e[UP_UNROLL];
#pragma unroll
            for (int q = 0; q < UP_UNROLL; ++q)
                e[q] = *(const uint32_t *)((const char *)(packed + (size_t)(k + q) * nu) + uoff);
#pragma unroll
            for (int q = 0; q < UP_UNROLL; ++q) {
                const T c = *reinterpret_cast<const T *>(lds + (e[q] & 0x7f8u));
                T v[R];
                gather((e[q] >> 11) << CSH, v);
#pragma unroll
                for (int r = 0; r < R; ++r) acc[r] += c * v[r];
            }
        }
        for (; k < cnt; ++k) {
            const uint32_t e = *(const uint32_t *)((const char *)(packed + (size_t)k * nu) + uoff);
            const T c = *reinterpret_cast<const T *>(lds + (e & 0x7f8u));
            T v[R];
            gather((e >> 11) << CSH, v);
#pragma unroll
            for (int r = 0; r < R; ++r) acc[r] += c * v[r];
        }
#pragma unroll
        for (int r = 0; r < R; ++r)
            if (d0 + r < a.dim_dn) __builtin_nontemporal_store(acc[r], out + (d0 + r) * (int64_t)nu + u);
    }
}

// COLS columns per row segment, ROWS rows per workgroup, VW columns per lane
// (spaced COLS/VW apart, so every load instruction is one coalesced segment of
// COLS/VW elements); a wave covers 64*VW/COLS rows at once.  The tile is picked
// per operator by the tuner in phub_hv_launch.
template <typename T, bool EPI, int COLS, int ROWS, int VW, int UN>
__global__ __launch_bounds__(256) void k_hv_dn(HubArgs<T> a, const T *__restrict__ x,
                                               T *__restrict__ out, double scale,
                                               const T *__restrict__ prev, double beta,
                                               double *__restrict__ part, int nchunks, int nrb) {
    constexpr int LPR = COLS / VW;          // lanes per row
    constexpr int RPW = 64 / LPR;           // rows per wave
    static_assert(LPR * VW == COLS && RPW * LPR == 64, "tile");
    const int b = blockIdx.x;
    const int xcd = b & 7, i = b >> 3;
    const int chunk = (i / nrb) * 8 + xcd;
    const int rb = i % nrb;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int sub = lane / LPR, cidx = lane % LPR;
    const int64_t nu = a.dim_up;
    const int64_t col0 = (int64_t)chunk * COLS + cidx;
    bool ok[VW];
    bool any = false;
#pragma unroll
    for (int j = 0; j < VW; ++j) {
        ok[j] = (chunk < nchunks) && (col0 + j * LPR < nu);
        any |= ok[j];
    }
    __shared__ T cls[256];
    for (int q = threadIdx.x; q < a.ncls_dn; q += 256) cls[q] = a.cls_dn[q];
    __syncthreads();
    const double sc = (EPI && a.dsc) ? a.dsc[0] : scale, bt = (EPI && a.dsc) ? a.dsc[1] : beta;
    double s_xo = 0.0, s_oo = 0.0;
    if (any) {
        // columns past the end read column col0 (in range) and are not stored
        int64_t cj[VW];
#pragma unroll
        for (int j = 0; j < VW; ++j) cj[j] = ok[j] ? col0 + j * LPR : col0;
        if constexpr (RPW == 1) {
            // one row per wave: its hop list is wave-uniform, so the entries, the
            // coefficients and each source row's base address live in SGPRs (scalar
            // loads through the constant address space; the tables are read-only
            // here) and a gather is a saddr load with the lane's 32-bit byte offset
            using ci32 = const __attribute__((address_space(4))) int32_t;
            using cu32 = const __attribute__((address_space(4))) uint32_t;
            using cT = const __attribute__((address_space(4))) T;
            ci32 *rp = (ci32 *)a.rp_dn;
            cu32 *csr = (cu32 *)a.csr_dn;
            cT *cl = (cT *)a.cls_dn;
            const int wv = __builtin_amdgcn_readfirstlane(wave);
            uint32_t cb[VW];
#pragma unroll
            for (int j = 0; j < VW; ++j) cb[j] = (uint32_t)cj[j] * (uint32_t)sizeof(T);
            for (int rr = wv; rr < ROWS; rr += 4) {
                const int64_t d = (int64_t)rb * ROWS + rr;
                if (d >= a.dim_dn) break;
                const int32_t r0 = rp[d], r1 = rp[d + 1];
                T sacc[VW];
#pragma unroll
                for (int j = 0; j < VW; ++j) sacc[j] = (T)0;
                int32_t k = r0;
                for (; k + UN <= r1; k += UN) {
                    uint32_t e[UN];
#pragma unroll
                    for (int q = 0; q < UN; ++q) e[q] = csr[k + q];
                    T v[UN][VW];
#pragma unroll
                    for (int q = 0; q < UN; ++q) {
                        const char *row = (const char *)(x + (int64_t)(e[q] >> 8) * nu);
#pragma unroll
                        for (int j = 0; j < VW; ++j) v[q][j] = *(const T *)(row + cb[j]);
                    }
#pragma unroll
                    for (int q = 0; q < UN; ++q) {
                        const T c = cl[e[q] & 255u];
#pragma unroll
                        for (int j = 0; j < VW; ++j) sacc[j] += c * v[q][j];
                    }
                }
                for (; k < r1; ++k) {
                    const uint32_t e = csr[k];
                    const char *row = (const char *)(x + (int64_t)(e >> 8) * nu);
                    const T c = cl[e & 255u];
#pragma unroll
                    for (int j = 0; j < VW; ++j) sacc[j] += c * *(const T *)(row + cb[j]);
                }
#pragma unroll
                for (int j = 0; j < VW; ++j) {
                    if (!ok[j]) continue;
                    const size_t o = (size_t)d * nu + cj[j];
                    if (EPI) {
                        double val = sc * ((double)__builtin_nontemporal_load(out + o) + (double)sacc[j]);
                        if (prev) val -= bt * (double)__builtin_nontemporal_load(prev + o);
                        const T vt = (T)val;
                        __builtin_nontemporal_store(vt, out + o);
                        s_xo += sc * (double)x[o] * (double)vt;
                        s_oo += (double)vt * (double)vt;
                    } else {
                        __builtin_nontemporal_store(__builtin_nontemporal_load(out + o) + sacc[j], out + o);
                    }
                }
            }
        } else
        for (int rr = wave * RPW + sub; rr < ROWS; rr += 4 * RPW) {
            const int64_t d = (int64_t)rb * ROWS + rr;
            if (d >= a.dim_dn) break;
            const int32_t r0 = a.rp_dn[d], r1 = a.rp_dn[d + 1];
            T s[VW];
#pragma unroll
            for (int j = 0; j < VW; ++j) s[j] = (T)0;
            int32_t k = r0;
            for (; k + UN <= r1; k += UN) {
                uint32_t e[UN];
                T v[UN][VW];
#pragma unroll
                for (int q = 0; q < UN; ++q) e[q] = a.csr_dn[k + q];
#pragma unroll
                for (int q = 0; q < UN; ++q) {
                    const T *row = x + (int64_t)(e[q] >> 8) * nu;
#pragma unroll
                    for (int j = 0; j < VW; ++j) v[q][j] = row[cj[j]];
                }
#pragma unroll
                for (int q = 0; q < UN; ++q) {
                    const T c = cls[e[q] & 255u];
#pragma unroll
                    for (int j = 0; j < VW; ++j) s[j] += c * v[q][j];
                }
            }
            for (; k < r1; ++k) {
                const uint32_t e = a.csr_dn[k];
                const T *row = x + (int64_t)(e >> 8) * nu;
                const T c = cls[e & 255u];
#pragma unroll
                for (int j = 0; j < VW; ++j) s[j] += c * row[cj[j]];
            }
#pragma unroll
            for (int j = 0; j < VW; ++j) {
                if (!ok[j]) continue;
                const size_t o = (size_t)d * nu + cj[j];
                if (EPI) {
                    // streamed once: non-temporal, so the slab keeps the L2
                    double val = sc * ((double)__builtin_nontemporal_load(out + o) + (double)s[j]);
                    if (prev) val -= bt * (double)__builtin_nontemporal_load(prev + o);
                    const T vt = (T)val;
                    __builtin_nontemporal_store(vt, out + o);
                    s_xo += sc * (double)x[o] * (double)vt;
                    s_oo += (double)vt * (double)vt;
                } else {
                    __builtin_nontemporal_store(__builtin_nontemporal_load(out + o) + s[j], out + o);
                }
            }
        }
    }
    if (EPI) block_sum2<T>(s_xo, s_oo, part);
}

// Down hops with buffer loads (k_hv_dn2): one row per wave, VW columns per lane spaced
// 64 apart, the same XCD-aware slabs as k_hv_dn.  A hop entry carries the source row's
// byte offset in the state and its coefficient (DnEnt), so it is loaded by the scalar
// unit straight into the SGPRs that the gathers use as their buffer soffset and the FMAs
// as their operand: no per-entry decode.  Entries are taken UN at a time with two groups
// in flight (the gathers of group g + 1 are issued before group g is summed).
template <typename T>
__device__ __forceinline__ T dn_bload(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
    if constexpr (sizeof(T) == 4)
        return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, (int)soff, 0));
    else
        return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, (int)voff, (int)soff, 0));
}

template <typename T, bool EPI, int ROWS, int VW, int UN>
__global__ __launch_bounds__(256) void k_hv_dn2(HubArgs<T> a, const DnEnt<T> *__restrict__ ent,
                                                uint32_t nrec, const T *__restrict__ x,
                                                T *__restrict__ out, double scale,
                                                const T *__restrict__ prev, double beta,
                                                double *__restrict__ part, int nchunks, int nrb) {
    constexpr int COLS = 64 * VW;
    const int b = blockIdx.x;
    const int xcd = b & 7, i = b >> 3;
    const int chunk = (i / nrb) * 8 + xcd;
    const int rb = i % nrb;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t nu = a.dim_up;
    const int64_t col0 = (int64_t)chunk * COLS + lane;
    bool ok[VW];
    bool any = false;
#pragma unroll
    for (int j = 0; j < VW; ++j) {
        ok[j] = (chunk < nchunks) && (col0 + 64 * j < nu);
        any |= ok[j];
    }
    const double sc = (EPI && a.dsc) ? a.dsc[0] : scale, bt = (EPI && a.dsc) ? a.dsc[1] : beta;
    double s_xo = 0.0, s_oo = 0.0;
    if (any) {
        // columns past the end read column col0 (in range) and are not stored
        int64_t cj[VW];
        uint32_t cb[VW];
#pragma unroll
        for (int j = 0; j < VW; ++j) {
            cj[j] = ok[j] ? col0 + 64 * j : col0;
            cb[j] = (uint32_t)cj[j] * (uint32_t)sizeof(T);
        }
        const __amdgpu_buffer_rsrc_t rs =
            __builtin_amdgcn_make_buffer_rsrc((void *)x, (short)0, (int)nrec, 0x00020000);
        using ci32 = const __attribute__((address_space(4))) int32_t;
        using cE = const __attribute__((address_space(4))) DnEnt<T>;
        ci32 *rp = (ci32 *)a.rp_dn;
        cE *E = (cE *)ent;
        const int wv = __builtin_amdgcn_readfirstlane(wave);
        for (int rr = wv; rr < ROWS; rr += 4) {
            const int64_t d = (int64_t)rb * ROWS + rr;
            if (d >= a.dim_dn) break;
            const int32_t r0 = rp[d], r1 = rp[d + 1];
            T acc[VW];
#pragma unroll
            for (int j = 0; j < VW; ++j) acc[j] = (T)0;
            auto issue = [&](int32_t k, T (&v)[UN][VW], T (&c)[UN]) {
#pragma unroll
                for (int q = 0; q < UN; ++q) {
                    const uint32_t off = E[k + q].off;
                    c[q] = E[k + q].c;
#pragma unroll
                    for (int j = 0; j < VW; ++j) v[q][j] = dn_bload<T>(rs, cb[j], off);
                }
            };
            auto consume = [&](const T (&v)[UN][VW], const T (&c)[UN]) {
                // keep the gathers just issued ahead of these sums
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int q = 0; q < UN; ++q)
#pragma unroll
                    for (int j = 0; j < VW; ++j) acc[j] += c[q] * v[q][j];
            };
            const int nfull = (r1 - r0) / UN;
            if (nfull > 0) {
                // the steady-state body issues unconditionally, so the wait before each
                // sum leaves the other group's gathers in flight
                T va[UN][VW], vb[UN][VW], ca[UN], cbq[UN];
                int g = 0;
                issue(r0, va, ca);
                while (g + 2 < nfull) {
                    issue(r0 + (g + 1) * UN, vb, cbq);
                    consume(va, ca);
                    issue(r0 + (g + 2) * UN, va, ca);
                    consume(vb, cbq);
                    g += 2;
                }
                if (g + 1 < nfull) {
                    issue(r0 + (g + 1) * UN, vb, cbq);
                    consume(va, ca);
                    consume(vb, cbq);
                } else {
                    consume(va, ca);
                }
            }
            for (int32_t k = r0 + nfull * UN; k < r1; ++k) {
                const uint32_t off = E[k].off;
                const T c = E[k].c;
#pragma unroll
                for (int j = 0; j < VW; ++j) acc[j] += c * dn_bload<T>(rs, cb[j], off);
            }
#pragma unroll
            for (int j = 0; j < VW; ++j) {
                if (!ok[j]) continue;
                const size_t o = (size_t)d * nu + cj[j];
                if (EPI) {
                    double val = sc * ((double)__builtin_nontemporal_load(out + o) + (double)acc[j]);
                    if (prev) val -= bt * (double)__builtin_nontemporal_load(prev + o);
                    const T vt = (T)val;
                    __builtin_nontemporal_store(vt, out + o);
                    s_xo += sc * (double)x[o] * (double)vt;
                    s_oo += (double)vt * (double)vt;
                } else {
                    __builtin_nontemporal_store(__builtin_nontemporal_load(out + o) + acc[j], out + o);
                }
            }
        }
    }
    if (EPI) block_sum2<T>(s_xo, s_oo, part);
}

// Staged Heisenberg H*v (replaces HPROD, hprod_heis.f90:1-61).  A state is (xh, xl),
// xh the nh top bits and xl the L low bits.  The states of one xl form a contiguous
// block of C(nh, m) states (m = n/2 - popc(xl)), state xh at position j(xh) from the
// block's first index.  One workgroup owns one block and keeps it in LDS (S0):
//   high-high flips  read S0 through a per-pattern table of targets, two 16-bit
//                    targets per word (a bond that does not flip points at a zero
//                    slot after the block);
//   low-low flips    read the same position in another block of the same popcount:
//                    a coalesced global read whose existence is uniform;
//   cross flips      one phase per low site of the cross bonds: the neighbouring
//                    block (popcount m +- 1) is staged in S1 and read through a
//                    per-pattern table of targets for the target's low bit (the zero
//                    slot where the flip does not exist).
// Every global read is coalesced, every scattered read is an LDS read, and every
// address is a wave-uniform base plus a 32-bit byte offset.  Targets are 4 * position:
// LDS byte offsets in FP32, doubled in FP64.  The diagonal is dconst minus half the
// Ising weight of each antiparallel bond: uniform for the low-low bonds, a per-pattern
// table for the high-high bonds, and two LDS tables indexed by the cross-site bits of
// xh for the cross bonds (built per block from the low bits of xl).  Matrix elements
// as HPROD's: the pull coefficient of a flip is bsign * J(J,K) (J symmetric), bsign
// the bond's sign convention factor.
struct HeisStgArgs {
    int nh, k, cntmax, ncx, nlo, nll, npair, nphase, nx;
    int order;          // workgroup -> block map: 0 each XCD takes one contiguous eighth of
                        // the list, 1 round robin, 2 / 3 runs of 16 / 64 consecutive blocks
                        // per XCD, the XCDs taking runs in turn
    int64_t nblk, GS;
    const int32_t *cnt;
    const int64_t *goff, *first;
    const uint32_t *blk, *hht, *cx;
    const uint16_t *xt;
    const float *ehh32;
    const double *ehh64;
    const HeisStgLL *ll;
    const HeisStgPair *hp;
    const HeisStgX *xb;
    const HeisStgPhase *ph;
    double dconst;
};

template <typename T, typename B>
__device__ __forceinline__ T stg_c(const B &b) {
    if constexpr (sizeof(T) == 4) return b.c32;
    else return b.c64;
}
template <typename T, typename B>
__device__ __forceinline__ T stg_h(const B &b) {
    if constexpr (sizeof(T) == 4) return b.h32;
    else return b.h64;
}
// element at 4-byte-unit offset o4 (4 * position) of an array of T
template <typename T>
__device__ __forceinline__ const T &stg_at(const T *base, uint32_t o4) {
    if constexpr (sizeof(T) == 4) return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) + o4);
    else return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) + (o4 << 1));
}
template <typename T>
__device__ __forceinline__ T &stg_at(T *base, uint32_t o4) {
    if constexpr (sizeof(T) == 4) return *reinterpret_cast<T *>(reinterpret_cast<char *>(base) + o4);
    else return *reinterpret_cast<T *>(reinterpret_cast<char *>(base) + (o4 << 1));
}
// a value every lane holds alike (read from one LDS address), moved to scalar registers
__device__ __forceinline__ int64_t stg_uniform(int64_t v) {
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)v);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)((uint64_t)v >> 32));
    return (int64_t)(((uint64_t)hi << 32) | lo);
}
__device__ __forceinline__ uint32_t stg_word(const uint32_t *base, uint32_t o4) {
    return *reinterpret_cast<const uint32_t *>(reinterpret_cast<const char *>(base) + o4);
}
__device__ __forceinline__ uint32_t stg_half(const uint16_t *base, uint32_t o4) {
    return *reinterpret_cast<const uint16_t *>(reinterpret_cast<const char *>(base) + (o4 >> 1));
}

// dynamic LDS of k_hv_heis_stg: two block buffers of C(15, 7) + 1 elements, the cross
// Ising tables and the per-block bond lists (at most 160 KB less the static LDS)
constexpr int HEIS_STG_SLOT = HEIS_STG_CNTMAX + 1;
constexpr size_t HEIS_STG_LDS_MAX = 160 * 1024 - 512;

// Copies the n <= C(15, 7) elements of a block into LDS: every load is issued before
// the first store (a loop with a run-time bound would wait on each load in turn).
template <typename T, int NT>
__device__ __forceinline__ void stg_stage(T *dst, const T *__restrict__ src, uint32_t n) {
    constexpr int K = (HEIS_STG_CNTMAX + NT - 1) / NT;
    const uint32_t tid = threadIdx.x;
    T v[K];
#pragma unroll
    for (int q = 0; q < K; ++q) {
        const uint32_t j = (uint32_t)q * NT + tid;
        if ((uint32_t)q * NT < n) v[q] = src[j < n ? j : n - 1u];
    }
#pragma unroll
    for (int q = 0; q < K; ++q) {
        const uint32_t j = (uint32_t)q * NT + tid;
        if ((uint32_t)q * NT < n && j < n) dst[j] = v[q];
    }
}

// The per-state work of one block, SPL states per lane (lanes past the block work on
// position cnt-1 and store nothing): no branch depends on the state index, so the
// accumulators stay in registers.
template <typename T, int NT, int SPL, bool EPI>
__device__ __forceinline__ void heis_stg_body(const HeisStgArgs &a, const T *__restrict__ x,
                                              T *__restrict__ out, double scale,
                                              const T *__restrict__ prev, double beta,
                                              uint32_t xl, int m, uint32_t cnt, int64_t go,
                                              int64_t first, T dgu, T *S0, T *S1,
                                              const T *Tlo, const T *Thi, const int64_t *llf,
                                              const T *llc, int nact, const int64_t *phf,
                                              double &s_xo, double &s_oo) {
    const uint32_t tid = threadIdx.x;
    uint32_t o4[SPL];
#pragma unroll
    for (int q = 0; q < SPL; ++q) {
        const uint32_t j = (uint32_t)q * NT + tid;
        o4[q] = 4u * (j < cnt ? j : cnt - 1u);
    }
    const uint32_t mlo = (1u << a.nlo) - 1u;
    T acc[SPL];
    {
        const uint32_t *cxb = a.cx + go;
        const float *e32 = a.ehh32 + go;
        const double *e64 = a.ehh64 + go;
#pragma unroll
        for (int q = 0; q < SPL; ++q) {
            const uint32_t cv = stg_word(cxb, o4[q]);
            T eh;
            if constexpr (sizeof(T) == 4) eh = stg_at<float>(e32, o4[q]);
            else eh = stg_at<double>(e64, o4[q]);
            acc[q] = (dgu - eh - Tlo[cv & mlo] - Thi[cv >> a.nlo]) * stg_at<T>(S0, o4[q]);
        }
    }
    // high-high bonds, two words (four bonds) per pass so that their loads overlap
    for (int p = 0; p < a.npair; p += 2) {
        const int p1 = p + 1 < a.npair ? p + 1 : p;   // an odd last pass repeats p with zeros
        const HeisStgPair h0 = a.hp[p], h1 = a.hp[p1];
        T c0, c1, c2, c3;
        if constexpr (sizeof(T) == 4) { c0 = h0.c32[0]; c1 = h0.c32[1]; c2 = h1.c32[0]; c3 = h1.c32[1]; }
        else { c0 = h0.c64[0]; c1 = h0.c64[1]; c2 = h1.c64[0]; c3 = h1.c64[1]; }
        if (p1 == p) { c2 = (T)0; c3 = (T)0; }
        const uint32_t *tb0 = a.hht + (int64_t)p * a.GS + go;
        const uint32_t *tb1 = a.hht + (int64_t)p1 * a.GS + go;
        uint32_t w0[SPL], w1[SPL];
#pragma unroll
        for (int q = 0; q < SPL; ++q) {
            w0[q] = stg_word(tb0, o4[q]);
            w1[q] = stg_word(tb1, o4[q]);
        }
#pragma unroll
        for (int q = 0; q < SPL; ++q)
            acc[q] += c0 * stg_at<T>(S0, w0[q] & 0xffffu) + c1 * stg_at<T>(S0, w0[q] >> 16) +
                      c2 * stg_at<T>(S0, w1[q] & 0xffffu) + c3 * stg_at<T>(S0, w1[q] >> 16);
    }
    // the antiparallel low-low bonds (compacted at the block start), two at a time
    {
        int b = 0;
        for (; b + 3 < nact; b += 4) {
            const T c0 = llc[b], c1 = llc[b + 1], c2 = llc[b + 2], c3 = llc[b + 3];
            const T *__restrict__ xs0 = x + stg_uniform(llf[b]);
            const T *__restrict__ xs1 = x + stg_uniform(llf[b + 1]);
            const T *__restrict__ xs2 = x + stg_uniform(llf[b + 2]);
            const T *__restrict__ xs3 = x + stg_uniform(llf[b + 3]);
            T v0[SPL], v1[SPL], v2[SPL], v3[SPL];
#pragma unroll
            for (int q = 0; q < SPL; ++q) {
                v0[q] = stg_at<T>(xs0, o4[q]);
                v1[q] = stg_at<T>(xs1, o4[q]);
                v2[q] = stg_at<T>(xs2, o4[q]);
                v3[q] = stg_at<T>(xs3, o4[q]);
            }
#pragma unroll
            for (int q = 0; q < SPL; ++q) acc[q] += (c0 * v0[q] + c1 * v1[q]) + (c2 * v2[q] + c3 * v3[q]);
        }
        for (; b < nact; ++b) {
            const T c = llc[b];
            const T *__restrict__ xs = x + stg_uniform(llf[b]);
#pragma unroll
            for (int q = 0; q < SPL; ++q) acc[q] += c * stg_at<T>(xs, o4[q]);
        }
    }
    // cross phases: one neighbouring block (high popcount m +- 1) staged in S1 each
    const uint32_t zero = 4u * (uint32_t)a.cntmax;
    for (int ph = 0; ph < a.nphase; ++ph) {
        const int64_t f2 = stg_uniform(phf[ph]);
        if (f2 < 0) continue;                // no source block: no lane flips
        const HeisStgPhase P = a.ph[ph];
        const uint32_t tb = (xl >> P.kk) & 1u;
        const uint32_t flip = tb << 15;
        __syncthreads();                     // the previous phase's reads of S1 are done
        stg_stage<T, NT>(S1, x + f2, (uint32_t)a.cnt[tb ? m + 1 : m - 1]);
        __syncthreads();
        for (int b = P.b0; b < P.b0 + P.nb; ++b) {
            const HeisStgX xb = a.xb[b];
            const T c = stg_c<T>(xb);
            const uint16_t *tt = a.xt + (int64_t)xb.jx * a.GS + go;
#pragma unroll
            for (int q = 0; q < SPL; ++q)
                acc[q] += c * stg_at<T>(S1, min(stg_half(tt, o4[q]) ^ flip, zero));
        }
    }
    T *__restrict__ ob = out + first;
    const T *__restrict__ pb = prev ? prev + first : nullptr;
#pragma unroll
    for (int q = 0; q < SPL; ++q) {
        const uint32_t j = (uint32_t)q * NT + tid;
        if (j < cnt) {
            if (EPI) {
                double val = scale * (double)acc[q];
                if (pb) val -= beta * (double)stg_at<T>(pb, o4[q]);
                const T vt = (T)val;
                stg_at<T>(ob, o4[q]) = vt;
                s_xo += scale * (double)stg_at<T>(S0, o4[q]) * (double)vt;
                s_oo += (double)vt * (double)vt;
            } else {
                stg_at<T>(ob, o4[q]) = acc[q];
            }
        }
    }
}

// states per lane of the five bodies: ceil(C(15, m) / NT) for m = 3..7, so that the
// blocks of a 15-site high half run without idle slots
template <int NT, int I>
constexpr int heis_stg_spl() {
    constexpr int c[5] = {455, 1365, 3003, 5005, HEIS_STG_CNTMAX};
    return (c[I] + NT - 1) / NT;
}

// FP32 runs two 512-thread workgroups per CU (four waves per SIMD: at most 128 VGPRs);
// FP64 one (its LDS holds one)
template <typename T, int NT, bool EPI>
__global__ __launch_bounds__(NT, sizeof(T) == 4 ? 4 : 2) void k_hv_heis_stg(HeisStgArgs a, const T *__restrict__ x,
                                                    T *__restrict__ out, double scale,
                                                    const T *__restrict__ prev, double beta,
                                                    double *__restrict__ part) {
    extern __shared__ __attribute__((aligned(16))) unsigned char stg_smem[];
    T *S0 = reinterpret_cast<T *>(stg_smem);
    T *S1 = S0 + HEIS_STG_SLOT;
    T *Wt = S1 + HEIS_STG_SLOT;           // Wt[hb][jx]: cross Ising weight for high bit hb
    T *Tlo = Wt + 2 * HEIS_STG_NXMAX;
    T *Thi = Tlo + (1 << a.nlo);
    // after the T arrays (rounded to 8 bytes): the compacted low-low sources and the
    // source block of every cross phase, then the low-low coefficients
    const size_t tbytes = ((2 * (size_t)HEIS_STG_SLOT + 2 * HEIS_STG_NXMAX + (1u << a.nlo) +
                            (1u << (a.ncx - a.nlo))) * sizeof(T) + 7) & ~(size_t)7;
    int64_t *llf = reinterpret_cast<int64_t *>(stg_smem + tbytes);
    int64_t *phf = llf + a.nll;
    T *llc = reinterpret_cast<T *>(phf + a.nphase);
    __shared__ int nact_sh;
    __shared__ T dgu_sh;
    const uint32_t tid = threadIdx.x;
    const int64_t bid = blockIdx.x, xcd = bid & 7, k = bid >> 3;
    int64_t it;
    if (a.order == 0) {
        const int64_t per = (a.nblk + 7) / 8;
        it = k < per ? xcd * per + k : a.nblk;
    }
    else if (a.order == 1) it = bid;
    else {
        const int64_t g = a.order == 2 ? 16 : 64;
        it = ((k / g) * 8 + xcd) * g + k % g;
    }
    double s_xo = 0.0, s_oo = 0.0;
    if (it < a.nblk) {
        const uint32_t xl = a.blk[it];
        const int m = a.k - __popc(xl);
        const uint32_t cnt = (uint32_t)a.cnt[m];
        const int64_t go = a.goff[m];
        const int64_t first = a.first[xl];
        const int nhi = a.ncx - a.nlo;
        stg_stage<T, NT>(S0, x + first, cnt);
        if (tid == 0) {
            S0[a.cntmax] = (T)0;
            S1[a.cntmax] = (T)0;
        }
        if ((int)tid < a.ncx) {
            T w0 = (T)0, w1 = (T)0;
            for (int b = 0; b < a.nx; ++b) {
                const HeisStgX xb = a.xb[b];
                if (xb.jx != (int)tid) continue;
                if ((xl >> xb.kk) & 1u) w0 += stg_h<T>(xb);   // high bit 0 against low bit 1
                else w1 += stg_h<T>(xb);
            }
            Wt[tid] = w0;
            Wt[HEIS_STG_NXMAX + tid] = w1;
        }
        if (tid < 64) {
            // wave 0 compacts the antiparallel low-low bonds (their source blocks are
            // looked up here in parallel) and sums their Ising weight
            T dsum = (T)0;
            int base = 0;
            for (int b0 = 0; b0 < a.nll; b0 += 64) {
                const int b = b0 + (int)tid;
                bool f = false;
                HeisStgLL l{};
                if (b < a.nll) {
                    l = a.ll[b];
                    f = ((xl >> l.p) ^ (xl >> l.q)) & 1u;
                }
                const uint64_t bal = __ballot(f);
                if (f) {
                    const int slot = base + __popcll(bal & ((1ull << tid) - 1ull));
                    llf[slot] = a.first[xl ^ ((1u << l.p) | (1u << l.q))];
                    llc[slot] = stg_c<T>(l);
                    dsum += stg_h<T>(l);
                }
                base += __popcll(bal);
            }
            for (int off = 32; off > 0; off >>= 1) dsum += __shfl_xor(dsum, off);
            if (tid == 0) {
                nact_sh = base;
                dgu_sh = (T)a.dconst - dsum;
            }
        }
        for (int p = (int)tid; p < a.nphase; p += NT) {
            const HeisStgPhase P = a.ph[p];
            const int m2 = ((xl >> P.kk) & 1u) ? m + 1 : m - 1;   // the source's high popcount
            phf[p] = (m2 < 0 || m2 > a.nh) ? (int64_t)-1 : a.first[xl ^ (1u << P.kk)];
        }
        __syncthreads();
        const T dgu = dgu_sh;
        const int nact = nact_sh;
        for (uint32_t sx = tid; sx < (1u << a.nlo); sx += NT) {
            T v = (T)0;
            for (int j = 0; j < a.nlo; ++j) v += Wt[((sx >> j) & 1u) * HEIS_STG_NXMAX + j];
            Tlo[sx] = v;
        }
        for (uint32_t sx = tid; sx < (1u << nhi); sx += NT) {
            T v = (T)0;
            for (int j = 0; j < nhi; ++j) v += Wt[((sx >> j) & 1u) * HEIS_STG_NXMAX + a.nlo + j];
            Thi[sx] = v;
        }
        __syncthreads();
        const uint32_t qn = (cnt + NT - 1) / NT;
#define PHUB_STG_BODY(I)                                                                      \
    heis_stg_body<T, NT, heis_stg_spl<NT, I>(), EPI>(a, x, out, scale, prev, beta, xl, m, cnt, \
                                                     go, first, dgu, S0, S1, Tlo, Thi, llf,   \
                                                     llc, nact, phf, s_xo, s_oo)
        if (qn <= (uint32_t)heis_stg_spl<NT, 0>()) PHUB_STG_BODY(0);
        else if (qn <= (uint32_t)heis_stg_spl<NT, 1>()) PHUB_STG_BODY(1);
        else if (qn <= (uint32_t)heis_stg_spl<NT, 2>()) PHUB_STG_BODY(2);
        else if (qn <= (uint32_t)heis_stg_spl<NT, 3>()) PHUB_STG_BODY(3);
        else PHUB_STG_BODY(4);
#undef PHUB_STG_BODY
    }
    if (EPI) block_sum2<T>(s_xo, s_oo, part);
}

// 4-index interaction (is_Uloc = 3, hprod_hubbard.f90:62-89) in pull form: the
// element (d', u') collects UL(J,K,M,N) s_up s_dn x[d][u] over the sources with
// u' = c^dag_{K up} c_{J up} u and d' = c^dag_{N dn} c_{M dn} d.  The sources are
// the all-pairs table entries of u' and d' read backwards (u = c^dag_J c_K u', same
// sign) plus the occupied sites (J = K, the number operator).  Added to out;
// with EPI the Lanczos epilogue is applied to the sum.
struct UlArgs {
    int n;
    int64_t nu, nd;
    const uint32_t *mu, *md;
    const int32_t *cu, *iu;   // up all-pairs table, k-major
    const double *su;
    const int16_t *pu;
    const int32_t *cd, *id;   // down all-pairs table, row-major
    const double *sd;
    const int16_t *pd;
    int maxh_d;
    const double *ul;         // n^4, thresholds applied
};

template <typename T, bool EPI>
__global__ __launch_bounds__(256) void k_hv_ul(UlArgs a, const T *__restrict__ x,
                                               T *__restrict__ out, double scale,
                                               const T *__restrict__ prev, double beta,
                                               double *__restrict__ part) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double s_xo = 0.0, s_oo = 0.0;
    if (e < a.nu * a.nd) {
        const int n = a.n, n2 = n * n;
        const int64_t dp = e / a.nu, upr = e - dp * a.nu;
        const uint32_t xu = a.mu[upr], xd = a.md[dp];
        const int cntu = a.cu[upr], cntd = a.cd[dp];
        double acc = 0.0;
        for (int ia = 0; ia < cntu + n; ++ia) {
            int J, K;
            double sg;
            int64_t us;
            if (ia < cntu) {
                const size_t o = (size_t)ia * a.nu + upr;
                const int pr = a.pu[o];
                J = pr % n;
                K = pr / n;
                sg = a.su[o];
                us = a.iu[o];
            } else {
                const int site = ia - cntu;
                if (!((xu >> (n - 1 - site)) & 1u)) continue;
                J = K = site;
                sg = 1.0;
                us = upr;
            }
            const double *ulJK = a.ul + ((size_t)J * n + K) * n2;
            for (int ib = 0; ib < cntd + n; ++ib) {
                int M, N;
                double sg2;
                int64_t ds;
                if (ib < cntd) {
                    const size_t o = (size_t)dp * a.maxh_d + ib;
                    const int q = a.pd[o];
                    M = q % n;
                    N = q / n;
                    sg2 = a.sd[o];
                    ds = a.id[o];
                } else {
                    const int site = ib - cntd;
                    if (!((xd >> (n - 1 - site)) & 1u)) continue;
                    M = N = site;
                    sg2 = 1.0;
                    ds = dp;
                }
                const double w = ulJK[M * n + N];
                if (w != 0.0) acc += w * sg * sg2 * (double)x[ds * a.nu + us];
            }
        }
        if (EPI) {
            double val = scale * ((double)out[e] + acc);
            if (prev) val -= beta * (double)prev[e];
            const T vt = (T)val;
            out[e] = vt;
            s_xo += scale * (double)x[e] * (double)vt;
            s_oo += (double)vt * (double)vt;
        } else {
            out[e] = (T)((double)out[e] + acc);
        }
    }
    if (EPI) block_sum2<T>(s_xo, s_oo, part);
}

// Row-staged form of the same sum.  A workgroup takes one target row d' and 256
// up targets u'.  The down pairs of d' (c^dag_N c_M d = d', occupied sites as M = N)
// are uniform over the workgroup: RB of them at a time, their source rows x[d][:]
// and the UL slices UL(:, :, M, N) times the down sign are staged in LDS, then each
// lane walks its own up pairs (J, K, u, sign) once per batch and reads both
// operands from LDS.  Used when RB >= 2 rows fit the LDS budget.
constexpr int UL_LDS_BUDGET = 48 * 1024;
constexpr int UL_MAXPAIRS = 288;   // pairs of one species state, n <= 32: N (n - N + 1) <= 272

template <typename T, bool EPI, int RB>
__global__ __launch_bounds__(256) void k_hv_ul_rows(UlArgs a, int nchunk, const T *__restrict__ x,
                                                    T *__restrict__ out, double scale,
                                                    const T *__restrict__ prev, double beta,
                                                    double *__restrict__ part) {
    // LDS, batch-interleaved so a lane reads the RB values of one (J,K) or one u
    // contiguously: s_ul[jk * RB + b] = UL(J,K,M_b,N_b) * sign_b, s_x[u * RB + b] = x[d_b][u]
    extern __shared__ double smem_ul[];
    __shared__ int s_mn[UL_MAXPAIRS];
    __shared__ int64_t s_ds[UL_MAXPAIRS];
    __shared__ double s_sg[UL_MAXPAIRS];
    __shared__ int s_cnt;
    const int n = a.n, n2 = n * n;
    const int64_t nu = a.nu;
    const int64_t dp = blockIdx.x / nchunk;
    const int64_t upr = (int64_t)(blockIdx.x % nchunk) * 256 + threadIdx.x;
    const bool live = upr < nu;
    double *s_ul = smem_ul;
    T *s_x = reinterpret_cast<T *>(smem_ul + (size_t)RB * n2);
    if (threadIdx.x == 0) {
        // the down pairs of d', in the order of the pull kernel (table, then sites)
        const uint32_t xd = a.md[dp];
        const int cntd = a.cd[dp];
        int c = 0;
        for (int ib = 0; ib < cntd; ++ib) {
            const size_t o = (size_t)dp * a.maxh_d + ib;
            const int q = a.pd[o];
            s_mn[c] = (q % n) * n + q / n;
            s_ds[c] = a.id[o];
            s_sg[c] = a.sd[o];
            ++c;
        }
        for (int site = 0; site < n; ++site)
            if ((xd >> (n - 1 - site)) & 1u) {
                s_mn[c] = site * n + site;
                s_ds[c] = dp;
                s_sg[c] = 1.0;
                ++c;
            }
        s_cnt = c;
    }
    __syncthreads();
    const int nbp = s_cnt;
    uint32_t xu = 0;
    int cntu = 0;
    if (live) {
        xu = a.mu[upr];
        cntu = a.cu[upr];
    }
    double acc = 0.0;
    for (int b0 = 0; b0 < nbp; b0 += RB) {
        const int nb = min(RB, nbp - b0);
        // batch slots past nb hold zeros (coefficient and value), so they add nothing
        for (int t = threadIdx.x; t < RB * n2; t += 256) {
            const int jk = t / RB, b = t - jk * RB;
            s_ul[t] = b < nb ? a.ul[(size_t)jk * n2 + s_mn[b0 + b]] * s_sg[b0 + b] : 0.0;
        }
        for (int b = 0; b < RB; ++b) {
            if (b < nb) {
                const T *row = x + s_ds[b0 + b] * nu;
                for (int64_t u = threadIdx.x; u < nu; u += 256) s_x[(size_t)u * RB + b] = row[u];
            } else {
                for (int64_t u = threadIdx.x; u < nu; u += 256) s_x[(size_t)u * RB + b] = (T)0;
            }
        }
        __syncthreads();
        if (live) {
            for (int ia = 0; ia < cntu + n; ++ia) {
                int jk;
                double sg;
                int64_t us;
                if (ia < cntu) {
                    const size_t o = (size_t)ia * nu + upr;
                    const int pr = a.pu[o];
                    jk = (pr % n) * n + pr / n;
                    sg = a.su[o];
                    us = a.iu[o];
                } else {
                    const int site = ia - cntu;
                    if (!((xu >> (n - 1 - site)) & 1u)) continue;
                    jk = site * n + site;
                    sg = 1.0;
                    us = upr;
                }
                const double *cu = s_ul + (size_t)jk * RB;
                const T *xv = s_x + (size_t)us * RB;
                double t = 0.0;
#pragma unroll
                for (int b = 0; b < RB; ++b) t += cu[b] * (double)xv[b];
                acc += sg * t;
            }
        }
        __syncthreads();
    }
    double s_xo = 0.0, s_oo = 0.0;
    if (live) {
        const int64_t e = dp * nu + upr;
        if (EPI) {
            double val = scale * ((double)out[e] + acc);
            if (prev) val -= beta * (double)prev[e];
            const T vt = (T)val;
            out[e] = vt;
            s_xo += scale * (double)x[e] * (double)vt;
            s_oo += (double)vt * (double)vt;
        } else {
            out[e] = (T)((double)out[e] + acc);
        }
    }
    if (EPI) block_sum2<T>(s_xo, s_oo, part);
}

template <typename T, bool EPI>
static void launch_ul_rows(phub_ctx *c, int rb, unsigned grid, size_t lds, const UlArgs &a,
                           int nchunk, const T *x, T *out, double scale, const T *prev,
                           double beta) {
    switch (rb) {
    case 16:
        hipLaunchKernelGGL((k_hv_ul_rows<T, EPI, 16>), dim3(grid), dim3(256), lds, c->stream, a,
                           nchunk, x, out, scale, prev, beta, c->d_red);
        break;
    case 8:
        hipLaunchKernelGGL((k_hv_ul_rows<T, EPI, 8>), dim3(grid), dim3(256), lds, c->stream, a,
                           nchunk, x, out, scale, prev, beta, c->d_red);
        break;
    case 4:
        hipLaunchKernelGGL((k_hv_ul_rows<T, EPI, 4>), dim3(grid), dim3(256), lds, c->stream, a,
                           nchunk, x, out, scale, prev, beta, c->d_red);
        break;
    default:
        hipLaunchKernelGGL((k_hv_ul_rows<T, EPI, 2>), dim3(grid), dim3(256), lds, c->stream, a,
                           nchunk, x, out, scale, prev, beta, c->d_red);
        break;
    }
}

template <typename T>
int phub_hv_ul(phub_ctx *c, const phub_op *op, const T *x, T *out, double scale, const T *prev,
               double beta, bool epi, int *nblocks_out) {
    UlArgs a{};
    a.n = op->n;
    a.nu = op->dim_up;
    a.nd = op->dim_dn;
    a.mu = op->up.masks;
    a.md = op->dn.masks;
    a.cu = op->rup.cnt;
    a.iu = op->rup.idx;
    a.su = op->rup.c64;
    a.pu = op->rup.pair;
    a.cd = op->rdn.cnt;
    a.id = op->rdn.idx;
    a.sd = op->rdn.c64;
    a.pd = op->rdn.pair;
    a.maxh_d = std::max(op->rdn.maxh, 1);
    a.ul = op->d_ul;
    const int n2 = op->n * op->n;
    // rows per batch: the largest power of two up to 16 that fits the LDS budget
    const int64_t fit = UL_LDS_BUDGET / ((int64_t)op->dim_up * sizeof(T) + (int64_t)n2 * 8);
    int rb = 16;
    while (rb > 1 && rb > fit) rb /= 2;
    const char *force_old = std::getenv("PHUB_UL_PULL");   // 1: the one-lane-per-target kernel
    if (rb >= 2 && op->rdn.maxh + op->n <= UL_MAXPAIRS && !(force_old && force_old[0] == '1')) {
        const int64_t nchunk = (op->dim_up + 255) / 256;
        const int64_t nbr = nchunk * op->dim_dn;
        if (nbr > INT32_MAX) return phub_fail(c, PHUB_ERR_UNSUPPORTED, "grid too large");
        const size_t lds = (size_t)rb * n2 * 8 + (size_t)rb * op->dim_up * sizeof(T);
        if (epi) {
            int rc = phub_ensure_red(c, (size_t)nbr * 2 + 64);
            if (rc) return rc;
            launch_ul_rows<T, true>(c, rb, (unsigned)nbr, lds, a, (int)nchunk, x, out, scale,
                                    prev, beta);
        } else {
            launch_ul_rows<T, false>(c, rb, (unsigned)nbr, lds, a, (int)nchunk, x, out, scale,
                                     prev, beta);
        }
        PHUB_HIP(c, hipGetLastError());
        if (nblocks_out) *nblocks_out = (int)nbr;
        return PHUB_OK;
    }
    const int64_t nb = (op->dim + 255) / 256;
    if (nb > INT32_MAX) return phub_fail(c, PHUB_ERR_UNSUPPORTED, "grid too large");
    if (epi) {
        int rc = phub_ensure_red(c, (size_t)nb * 2 + 64);
        if (rc) return rc;
        hipLaunchKernelGGL((k_hv_ul<T, true>), dim3((unsigned)nb), dim3(256), 0, c->stream, a, x,
                           out, scale, prev, beta, c->d_red);
    } else {
        hipLaunchKernelGGL((k_hv_ul<T, false>), dim3((unsigned)nb), dim3(256), 0, c->stream, a, x,
                           out, scale, prev, beta, c->d_red);
    }
    PHUB_HIP(c, hipGetLastError());
    if (nblocks_out) *nblocks_out = (int)nb;
    return PHUB_OK;
}

template int phub_hv_ul<double>(phub_ctx *, const phub_op *, const double *, double *, double,
                                const double *, double, bool, int *);
template int phub_hv_ul<float>(phub_ctx *, const phub_op *, const float *, float *, double,
                               const float *, double, bool, int *);

template <typename T>
static HubArgs<T> hub_args(const phub_op *op, const phub_ctx *c = nullptr) {
    HubArgs<T> a{};
    a.dsc = c ? c->dev_scalars : nullptr;
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

// minv[d * dim_up + u] = 1 / max(H_ii - e, floor): the inverse shifted diagonal of the
// one-body Hubbard operator (the diagonal the H*v kernels add: eloc of both species
// plus U over the doubly occupied sites), a Jacobi preconditioner for H - e near the
// ground state (every H_ii >= e0, so the entries are positive)
template <typename V>
__global__ __launch_bounds__(256) void k_hub_minv(HubArgs<float> a, double e, double floor_,
                                                 V *__restrict__ minv) {
    const int64_t n = a.dim_up * a.dim_dn;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t d = i / a.dim_up, u = i - d * a.dim_up;
        const uint32_t both = a.mu[u] & a.md[d];
        double dg = a.eu[u] + a.ed[d];
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
        minv[i] = (V)(1.0 / fmax(dg - e, floor_));
    }
}

template <typename V>
int phub_hub_minv(phub_ctx *c, const phub_op *op, double e, V *minv) {
    if (op->kind != 0 || op->d_ul)
        return phub_fail(c, PHUB_ERR_UNSUPPORTED, "diagonal preconditioner: one-body Hubbard operators only");
    const int64_t n = op->dim;
    const unsigned g = (unsigned)std::min<int64_t>((n + 255) / 256, 65536);
    hipLaunchKernelGGL(k_hub_minv<V>, dim3(g), dim3(256), 0, c->stream, hub_args<float>(op), e, 1e-3,
                       minv);
    PHUB_HIP(c, hipGetLastError());
    return PHUB_OK;
}
template int phub_hub_minv<float>(phub_ctx *, const phub_op *, double, float *);
template int phub_hub_minv<double>(phub_ctx *, const phub_op *, double, double *);

constexpr int HUB_R = 4;

// Candidate tiles of the down-hop kernels: (columns per row segment, rows per
// workgroup, columns per lane, row gathers in flight per lane, kernel).  Kernel 0 is
// k_hv_dn (compressed CSR entries, any tile); kernel 1 is k_hv_dn2 (byte-offset
// entries, buffer loads, one row per wave: columns = 64 * columns per lane).  FP32 and
// FP64 list the same byte widths, from 64 B to 1 KB per row segment.
struct DnTile {
    int cols, rows, vw, un, v2;
};
// The three with 4 or 8 rows per workgroup keep fewer rows in flight per XCD: their
// down-hop sources then span a narrower window of the state, more of it still in that
// XCD's L2 (4x4 FP32: 512 B x 4 rows, 8 in flight, 1.63 ms against 1.69 ms at 16 rows).
constexpr int DN_NCFG = 25;
constexpr DnTile DN_F32[DN_NCFG] = {
    {16, 16, 1, 8, 0},  {32, 16, 1, 8, 0},  {32, 32, 1, 8, 0},  {64, 8, 1, 8, 0},
    {64, 16, 1, 8, 0},  {64, 32, 1, 8, 0},  {128, 16, 2, 4, 0}, {128, 32, 2, 4, 0},
    {256, 16, 4, 2, 0}, {256, 32, 4, 2, 0}, {128, 16, 2, 8, 0}, {128, 32, 2, 8, 0},
    {256, 16, 4, 4, 0}, {256, 32, 4, 4, 0}, {128, 4, 2, 8, 0},  {128, 8, 2, 8, 0},
    {256, 4, 4, 4, 0},
    {64, 4, 1, 8, 1},   {64, 8, 1, 8, 1},   {64, 16, 1, 8, 1},  {128, 4, 2, 8, 1},
    {128, 8, 2, 8, 1},  {128, 16, 2, 4, 1}, {256, 4, 4, 4, 1},  {256, 8, 4, 4, 1}};
constexpr DnTile DN_F64[DN_NCFG] = {
    {8, 16, 1, 8, 0},   {16, 16, 1, 8, 0},  {16, 32, 1, 8, 0},  {32, 8, 1, 8, 0},
    {32, 16, 1, 8, 0},  {32, 32, 1, 8, 0},  {64, 16, 2, 4, 0},  {64, 32, 2, 4, 0},
    {128, 16, 4, 2, 0}, {128, 32, 4, 2, 0}, {64, 16, 2, 8, 0},  {64, 32, 2, 8, 0},
    {128, 16, 4, 4, 0}, {128, 32, 4, 4, 0}, {64, 4, 2, 8, 0},   {64, 8, 2, 8, 0},
    {128, 4, 4, 4, 0},
    {64, 4, 1, 8, 1},   {64, 8, 1, 8, 1},   {64, 16, 1, 8, 1},  {128, 4, 2, 4, 1},
    {128, 8, 2, 8, 1},  {128, 16, 2, 4, 1}, {256, 4, 4, 2, 1},  {256, 8, 4, 4, 1}};
constexpr int DN_DEFAULT = 4;   // 256 B segments, 16 rows

template <typename T>
constexpr DnTile dn_tile(int cfg) {
    return sizeof(T) == 4 ? DN_F32[cfg] : DN_F64[cfg];
}

int phub_dn_tile_code(int cfg, bool fp64) {
    if (cfg < 0) return -1;
    const DnTile t = fp64 ? DN_F64[cfg] : DN_F32[cfg];
    return t.cols * (fp64 ? 8 : 4) * 100000 + t.rows * 100 + t.un + (t.v2 ? 50 : 0);
}

// byte-offset entries from the row-major down table: entry k of row s (CSR order)
template <typename T>
__global__ void k_dn_ent(const int32_t *__restrict__ idx, const double *__restrict__ c64,
                         const int32_t *__restrict__ cnt, const int32_t *__restrict__ rowptr,
                         int64_t dim, int maxh, uint64_t rowbytes, DnEnt<T> *__restrict__ ent) {
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= dim) return;
    const int c = cnt[s];
    const int32_t base = rowptr[s];
    for (int k = 0; k < c; ++k) {
        const size_t o = (size_t)s * maxh + k;
        DnEnt<T> e{};
        e.off = (uint32_t)((uint64_t)idx[o] * rowbytes);
        e.c = (T)c64[o];
        ent[base + k] = e;
    }
}

int phub_build_dn_ent(phub_ctx *c, phub_op *op) {
    const HopTable &t = op->dn;
    if (!t.rowptr || t.dim <= 0 || t.maxh <= 0) return PHUB_OK;
    int32_t nnz = 0;
    PHUB_HIP(c, phub_copy(c, &nnz, t.rowptr + t.dim, sizeof(int32_t), hipMemcpyDeviceToHost));
    if (nnz <= 0) return PHUB_OK;
    for (int ti = 0; ti < 2; ++ti) {
        const uint64_t es = ti ? 8 : 4;
        if ((uint64_t)op->dim_up * (uint64_t)op->dim_dn * es >= (1ull << 32)) continue;
        const uint64_t rowbytes = (uint64_t)op->dim_up * es;
        if (ti == 0) {
            DnEnt<float> *e = nullptr;
            PHUB_HIP(c, phub_malloc(&e, (size_t)nnz * sizeof(DnEnt<float>)));
            op->dn_ent[0] = e;
            hipLaunchKernelGGL(k_dn_ent<float>, dim3((unsigned)((t.dim + 255) / 256)), dim3(256), 0, c->stream, t.idx,
                               t.c64, t.cnt, t.rowptr, t.dim, t.maxh, rowbytes, e);
        } else {
            DnEnt<double> *e = nullptr;
            PHUB_HIP(c, phub_malloc(&e, (size_t)nnz * sizeof(DnEnt<double>)));
            op->dn_ent[1] = e;
            hipLaunchKernelGGL(k_dn_ent<double>, dim3((unsigned)((t.dim + 255) / 256)), dim3(256), 0, c->stream, t.idx,
                               t.c64, t.cnt, t.rowptr, t.dim, t.maxh, rowbytes, e);
        }
        PHUB_HIP(c, hipGetLastError());
        op->table_bytes += (double)nnz * (ti ? sizeof(DnEnt<double>) : sizeof(DnEnt<float>));
    }
    PHUB_HIP(c, hipStreamSynchronize(c->stream));
    return PHUB_OK;
}

// the buffer-load kernel needs the byte-offset entries (a state below 4 GiB)
template <typename T>
static bool dn_tile_ok(const phub_op *op, int cfg) {
    return !dn_tile<T>(cfg).v2 || op->dn_ent[sizeof(T) == 8];
}

template <typename T, bool EPI, int CFG>
static void dn_go(phub_ctx *c, const phub_op *op, const HubArgs<T> &a, const T *x, T *out,
                  double scale, const T *prev, double beta, int *nblocks) {
    constexpr DnTile t = dn_tile<T>(CFG);
    const int nchunks = (int)((op->dim_up + t.cols - 1) / t.cols);
    const int nrb = (int)((op->dim_dn + t.rows - 1) / t.rows);
    *nblocks = (int)std::min<int64_t>((int64_t)8 * ((nchunks + 7) / 8) * nrb, INT32_MAX);
    if constexpr (t.v2) {
        static_assert(t.cols == 64 * t.vw, "one row per wave");
        const uint32_t nrec = (uint32_t)std::min<uint64_t>((uint64_t)op->dim * sizeof(T), 0xffffffffull);
        hipLaunchKernelGGL((k_hv_dn2<T, EPI, t.rows, t.vw, t.un>), dim3(*nblocks), dim3(256), 0,
                           c->stream, a, (const DnEnt<T> *)op->dn_ent[sizeof(T) == 8], nrec, x, out,
                           scale, prev, beta, c->d_red, nchunks, nrb);
    } else {
        hipLaunchKernelGGL((k_hv_dn<T, EPI, t.cols, t.rows, t.vw, t.un>), dim3(*nblocks), dim3(256), 0,
                           c->stream, a, x, out, scale, prev, beta, c->d_red, nchunks, nrb);
    }
}

#define PHUB_DN_CASE(k) \
    case k: dn_go<T, EPI, k>(c, op, a, x, out, scale, prev, beta, nb); break;

template <typename T, bool EPI>
static void dn_dispatch(phub_ctx *c, const phub_op *op, const HubArgs<T> &a, const T *x, T *out,
                        double scale, const T *prev, double beta, int cfg, int *nb) {
    static_assert(DN_NCFG == 25, "dispatch cases");
    switch (cfg) {
    PHUB_DN_CASE(0) PHUB_DN_CASE(1) PHUB_DN_CASE(2) PHUB_DN_CASE(3) PHUB_DN_CASE(4)
    PHUB_DN_CASE(5) PHUB_DN_CASE(6) PHUB_DN_CASE(7) PHUB_DN_CASE(8) PHUB_DN_CASE(9)
    PHUB_DN_CASE(10) PHUB_DN_CASE(11) PHUB_DN_CASE(12) PHUB_DN_CASE(13) PHUB_DN_CASE(14)
    PHUB_DN_CASE(15) PHUB_DN_CASE(16) PHUB_DN_CASE(17) PHUB_DN_CASE(18) PHUB_DN_CASE(19)
    PHUB_DN_CASE(20) PHUB_DN_CASE(21) PHUB_DN_CASE(22) PHUB_DN_CASE(23)
    default: dn_go<T, EPI, 24>(c, op, a, x, out, scale, prev, beta, nb); break;
    }
}
#undef PHUB_DN_CASE

template <typename T>
static int launch_dn(phub_ctx *c, const phub_op *op, const HubArgs<T> &a, const T *x, T *out,
                     double scale, const T *prev, double beta, bool epi, int cfg, int *nblocks) {
    if (!dn_tile_ok<T>(op, cfg)) cfg = DN_DEFAULT;
    const DnTile t = dn_tile<T>(cfg);
    const int cols = t.cols, rows = t.rows;
    const int64_t nchunks = (op->dim_up + cols - 1) / cols, nrb = (op->dim_dn + rows - 1) / rows;
    const int64_t g2 = 8 * ((nchunks + 7) / 8) * nrb;
    if (g2 > INT32_MAX) return phub_fail(c, PHUB_ERR_UNSUPPORTED, "grid too large");
    if (epi) {
        int rc = phub_ensure_red(c, (size_t)g2 * 2 + 64);
        if (rc) return rc;
        dn_dispatch<T, true>(c, op, a, x, out, scale, prev, beta, cfg, nblocks);
    } else {
        dn_dispatch<T, false>(c, op, a, x, out, scale, prev, beta, cfg, nblocks);
    }
    PHUB_HIP(c, hipGetLastError());
    return PHUB_OK;
}

// Tuner: on an operator's first H*v in a precision, time every down-hop tile on
// the caller's own vectors (up kernel, then the candidate; the best of two runs
// each) and keep the fastest.  The pick is remembered on the context by operator
// shape (sector dimensions, maximal hop counts, precision), and a later operator
// of the same shape takes it without timing: workloads that build one operator
// per parameter set would otherwise pay about 24 extra H*v per operator.
// PHUB_TUNE=0 keeps the default tile, PHUB_DN_TILE=k forces tile k.  Every run
// rewrites out completely, so the caller's result is recomputed afterwards.
template <typename T, int R>
static int tune_dn(phub_ctx *c, phub_op *op, const HubArgs<T> &a, const T *x, T *out,
                   double scale, const T *prev, double beta, bool epi, unsigned g1, size_t lds,
                   const T *cls) {
    const int ti = sizeof(T) == 8;
    const char *env = std::getenv("PHUB_TUNE");
    if (env && env[0] == '0') {
        op->tune_dn[ti] = DN_DEFAULT;
        return PHUB_OK;
    }
    const char *force = std::getenv("PHUB_DN_TILE");   // a fixed tile index
    if (force && std::atoi(force) >= 0 && std::atoi(force) < DN_NCFG) {
        op->tune_dn[ti] = std::atoi(force);
        return PHUB_OK;
    }
    for (const auto &t : c->tuned)
        if (t.dim_up == op->dim_up && t.dim_dn == op->dim_dn && t.maxh_up == op->up.maxh &&
            t.maxh_dn == op->dn.maxh && t.fp64 == ti) {
            op->tune_dn[ti] = t.cfg;
            return PHUB_OK;
        }
    float best = 1e30f;
    int bestc = DN_DEFAULT;
    for (int cfg = 0; cfg < DN_NCFG; ++cfg) {
        if (!dn_tile_ok<T>(op, cfg)) continue;
        float tmin = 1e30f;
        for (int rep = 0; rep < 2; ++rep) {
            hipLaunchKernelGGL((k_hv_up<T, R>), dim3(g1), dim3(UP_THREADS), lds, c->stream, a,
                               op->up.packed, cls, op->up.ncls, x, out);
            PHUB_HIP(c, hipEventRecord(c->ev0, c->stream));
            int nb = 0;
            int rc = launch_dn<T>(c, op, a, x, out, scale, prev, beta, epi, cfg, &nb);
            if (rc) return rc;
            PHUB_HIP(c, hipEventRecord(c->ev1, c->stream));
            PHUB_HIP(c, hipEventSynchronize(c->ev1));
            float ms = 0.f;
            PHUB_HIP(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
            tmin = std::min(tmin, ms);
        }
        if (std::getenv("PHUB_TUNE_VERBOSE"))
            std::fprintf(stderr, "phub tune: %s down tile %d: %s %d cols x %d rows (%d per lane, %d in flight): %.3f ms\n",
                         sizeof(T) == 4 ? "fp32" : "fp64", cfg,
                         dn_tile<T>(cfg).v2 ? "k_hv_dn2" : "k_hv_dn",
                         dn_tile<T>(cfg).cols, dn_tile<T>(cfg).rows, dn_tile<T>(cfg).vw,
                         dn_tile<T>(cfg).un, tmin);
        if (tmin < best) { best = tmin; bestc = cfg; }
    }
    op->tune_dn[ti] = bestc;
    c->tuned.push_back(phub_ctx::TunedTile{op->dim_up, op->dim_dn, op->up.maxh, op->dn.maxh, ti, bestc});
    return PHUB_OK;
}

template <typename T>
static int hv_launch_main(phub_ctx *c, const phub_op *op, const T *x, T *out, double scale,
                          const T *prev, double beta, bool want_dots, int *nblocks_out);

template <typename T>
int phub_hv_launch(phub_ctx *c, const phub_op *op, const T *x, T *out, double scale,
                   const T *prev, double beta, bool want_dots, int *nblocks_out) {
    // the epilogue kernel applies scale / prev and writes dot partials
    const bool epi = want_dots || prev != nullptr || scale != 1.0;
    if (op->kind == 0 && op->d_ul) {
        // one-body part, then the 4-index term with the epilogue on the sum
        int rc = hv_launch_main<T>(c, op, x, out, 1.0, (const T *)nullptr, 0.0, false, nullptr);
        if (rc) return rc;
        return phub_hv_ul<T>(c, op, x, out, scale, prev, beta, epi, nblocks_out);
    }
    return hv_launch_main<T>(c, op, x, out, scale, prev, beta, want_dots, nblocks_out);
}

// the two-kernel Hubbard H*v with R rows per up-kernel workgroup; the operator's
// first call in a precision runs the down-tile tuner
template <typename T, int R>
static int hv_two_kernel(phub_ctx *c, const phub_op *op, const T *x, T *out, double scale,
                         const T *prev, double beta, bool epi, int *nblocks_out) {
    int nblocks = 0;
    const size_t lds = 0;   // k_hv_up's LDS is static (UP_LDS_MAX)
    HubArgs<T> a = hub_args<T>(op, c);
    const T *cls = (sizeof(T) == 8) ? (const T *)op->up.cls64 : (const T *)op->up.cls32;
    const unsigned g1 = (unsigned)((op->dim_dn + R - 1) / R);
    hipLaunchKernelGGL((k_hv_up<T, R>), dim3(g1), dim3(UP_THREADS), lds, c->stream, a,
                       op->up.packed, cls, op->up.ncls, x, out);
    PHUB_HIP(c, hipGetLastError());
    const int ti = sizeof(T) == 8;
    phub_op *mop = const_cast<phub_op *>(op);
    if (mop->tune_dn[ti] < 0) {
        int rc = tune_dn<T, R>(c, mop, a, x, out, scale, prev, beta, epi, g1, lds, cls);
        if (rc) return rc;
        // the tuner's runs overwrote out: recompute the up part
        hipLaunchKernelGGL((k_hv_up<T, R>), dim3(g1), dim3(UP_THREADS), lds, c->stream, a,
                           op->up.packed, cls, op->up.ncls, x, out);
    }
    int rc = launch_dn<T>(c, op, a, x, out, scale, prev, beta, epi, mop->tune_dn[ti], &nblocks);
    if (rc) return rc;
    PHUB_HIP(c, hipGetLastError());
    if (nblocks_out) *nblocks_out = nblocks;
    return PHUB_OK;
}

// PHUB_UP_ROWS=1 keeps FP64 on the one-row up kernel even where two rows fit
// (the path of the large sectors, exercised on small ones by the tests)
static HeisStgArgs heis_stg_args(const phub_op *op) {
    const HeisStage &s = op->h.stg;
    HeisStgArgs a{};
    a.nh = s.nh;
    a.k = s.k;
    a.cntmax = s.cntmax;
    a.ncx = s.ncx;
    a.nlo = s.nlo;
    a.nll = s.nll;
    a.npair = s.npair;
    a.nphase = s.nphase;
    a.nx = s.nx;
    a.nblk = s.nblk;
    a.GS = s.GS;
    a.cnt = s.cnt;
    a.goff = s.goff;
    a.first = s.first;
    a.blk = s.blk;
    a.hht = s.hht;
    a.xt = s.xt;
    a.cx = s.cx;
    a.ehh32 = s.ehh32;
    a.ehh64 = s.ehh64;
    a.ll = s.ll;
    a.hp = s.hp;
    a.xb = s.xb;
    a.ph = s.ph;
    a.dconst = op->h.diag_const;
    return a;
}

// FP32: 256 threads, three workgroups per CU in LDS; FP64: 512 threads, one per CU
template <typename T>
constexpr int heis_stg_threads() { return 512; }

template <typename T, bool EPI>
static int heis_stg_go(phub_ctx *c, const HeisStgArgs &a, int nblocks, size_t lds, const T *x,
                       T *out, double scale, const T *prev, double beta) {
    constexpr int NT = heis_stg_threads<T>();
    static bool attr_set[PHUB_MAX_DEVICES];   // set once per device and instantiation
    if (c->device < 0 || c->device >= PHUB_MAX_DEVICES)
        return phub_fail(c, PHUB_ERR_UNSUPPORTED, "device index beyond PHUB_MAX_DEVICES");
    if (!attr_set[c->device]) {
        PHUB_HIP(c, hipFuncSetAttribute((const void *)k_hv_heis_stg<T, NT, EPI>,
                                        hipFuncAttributeMaxDynamicSharedMemorySize,
                                        (int)HEIS_STG_LDS_MAX));
        attr_set[c->device] = true;
    }
    hipLaunchKernelGGL((k_hv_heis_stg<T, NT, EPI>), dim3(nblocks), dim3(NT), lds, c->stream, a, x,
                       out, scale, prev, beta, c->d_red);
    PHUB_HIP(c, hipGetLastError());
    return PHUB_OK;
}

template <typename T>
static int hv_heis_stage(phub_ctx *c, const phub_op *op, const T *x, T *out, double scale,
                         const T *prev, double beta, bool epi, int *nblocks_out) {
    const HeisStage &s = op->h.stg;
    // the grid covers whole runs of 64 blocks on each of the 8 XCDs (every map below
    // sends workgroup ids past the list to no block)
    const int64_t nb = 512 * ((s.nblk + 511) / 512);
    if (nb > INT32_MAX) return phub_fail(c, PHUB_ERR_UNSUPPORTED, "grid too large");
    const int nblocks = (int)nb;
    if (epi) {
        int rc = phub_ensure_red(c, (size_t)nblocks * 2 + 64);
        if (rc) return rc;
    }
    HeisStgArgs a = heis_stg_args(op);
    // LDS: the own block, one source block, the cross Ising tables, then the low-low
    // sources and coefficients and the phase sources (FP32: two workgroups per CU)
    const size_t tbytes = (((size_t)2 * HEIS_STG_SLOT + 2 * HEIS_STG_NXMAX + ((size_t)1 << s.nlo) +
                            ((size_t)1 << (s.ncx - s.nlo))) * sizeof(T) + 7) & ~(size_t)7;
    const size_t lds = tbytes + (size_t)(s.nll + s.nphase) * sizeof(int64_t) + (size_t)s.nll * sizeof(T);
    if (lds > HEIS_STG_LDS_MAX) return phub_fail(c, PHUB_ERR_UNSUPPORTED, "Heisenberg blocks exceed LDS");
    // code: bits 0-1 the workgroup -> block map, bit 2 the block list (0 memory
    // order, 1 popcount-major)
    auto go = [&](int code) {
        a.order = code & 3;
        a.blk = (code & 4) ? s.blkm : s.blk;
        return epi ? heis_stg_go<T, true>(c, a, nblocks, lds, x, out, scale, prev, beta)
                   : heis_stg_go<T, false>(c, a, nblocks, lds, x, out, scale, prev, beta);
    };
    const int ti = sizeof(T) == 8;
    phub_op *mop = const_cast<phub_op *>(op);
    if (mop->tune_heis[ti] < 0) {
        // time the eight block orders on the caller's vectors (every launch rewrites
        // out completely); PHUB_HEIS_ORDER=k forces code k, PHUB_TUNE=0 takes code 0
        const char *env = std::getenv("PHUB_TUNE");
        const char *force = std::getenv("PHUB_HEIS_ORDER");
        int best = 0;
        if (force && std::atoi(force) >= 0 && std::atoi(force) < 8) {
            best = std::atoi(force);
        } else if (!(env && env[0] == '0')) {
            float tb = 1e30f;
            for (int order = 0; order < 8; ++order) {
                PHUB_HIP(c, hipEventRecord(c->ev0, c->stream));
                int rc = go(order);
                if (rc) return rc;
                PHUB_HIP(c, hipEventRecord(c->ev1, c->stream));
                PHUB_HIP(c, hipEventSynchronize(c->ev1));
                float ms = 0.f;
                PHUB_HIP(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
                if (std::getenv("PHUB_TUNE_VERBOSE"))
                    std::fprintf(stderr, "phub tune: %s heisenberg block order %d: %.3f ms\n",
                                 sizeof(T) == 4 ? "fp32" : "fp64", order, ms);
                if (ms < tb) { tb = ms; best = order; }
            }
        }
        mop->tune_heis[ti] = best;
    }
    int rc = go(mop->tune_heis[ti]);
    if (rc) return rc;
    if (nblocks_out) *nblocks_out = nblocks;
    return PHUB_OK;
}

static bool phub_force_r1() {
    const char *e = std::getenv("PHUB_UP_ROWS");
    return e && e[0] == '1';
}
// PHUB_UP_ROWS=2 keeps FP32 on two rows per workgroup where three fit
static bool phub_force_r2() {
    const char *e = std::getenv("PHUB_UP_ROWS");
    return e && e[0] == '2';
}

// the one-body (T, per-site U) Hubbard operator and the Heisenberg operator
template <typename T>
static int hv_launch_main(phub_ctx *c, const phub_op *op, const T *x, T *out, double scale,
                          const T *prev, double beta, bool want_dots, int *nblocks_out) {
    int nblocks = 0;
    const bool epi = want_dots || prev != nullptr || scale != 1.0;
    if (op->kind == 0 && op->up.packed && op->dn.csr) {
        // two-kernel path: R rows of the state in LDS for the up hops (FP32: 3 when
        // three rows fit, else 2; FP64: 2 when two rows fit, else 1)
        const size_t row_bytes = (size_t)op->dim_up * sizeof(T);
        if (sizeof(T) == 8 && UP_CLS_BYTES + 2 * row_bytes <= UP_LDS_MAX && !phub_force_r1())
            return hv_two_kernel<T, 2>(c, op, x, out, scale, prev, beta, epi, nblocks_out);
        if constexpr (sizeof(T) == 4)
            if (UP_CLS_BYTES + 3 * row_bytes <= UP_LDS_MAX && !phub_force_r2())
                return hv_two_kernel<T, 3>(c, op, x, out, scale, prev, beta, epi, nblocks_out);
        constexpr int R = sizeof(T) == 4 ? 2 : 1;
        if (UP_CLS_BYTES + R * row_bytes <= UP_LDS_MAX)
            return hv_two_kernel<T, R>(c, op, x, out, scale, prev, beta, epi, nblocks_out);
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
        HubArgs<T> a = hub_args<T>(op, c);
        if (epi)
            hipLaunchKernelGGL((k_hv_hub<T, HUB_R, true>), dim3(nblocks), dim3(256), 0, c->stream,
                               a, x, out, scale, prev, beta, c->d_red, ncb);
        else
            hipLaunchKernelGGL((k_hv_hub<T, HUB_R, false>), dim3(nblocks), dim3(256), 0,
                               c->stream, a, x, out, scale, prev, beta, c->d_red, ncb);
    } else {
        return hv_heis_stage<T>(c, op, x, out, scale, prev, beta, epi, nblocks_out);
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
