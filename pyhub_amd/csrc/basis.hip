// basis.hip -- context, device-built bases and ranking / hop tables.
//
// Basis order (reference: HILBERT, pyhub/libs/src/compute_basis.f90:67-96) is the
// lexicographic order of the increasing set-bit positions; its closed form is
// rank(x) = C(n,k) - 1 - colex(rev_n(x)).  Bases are produced here by unranking
// that closed form on the device, one thread per state.
#include "phub_internal.h"

#include <algorithm>
#include <cmath>
#include <cstring>

std::atomic<unsigned long long> g_phub_alloc_count{0};

extern "C" uint64_t phub_alloc_count(void) {
    return (uint64_t)g_phub_alloc_count.load(std::memory_order_relaxed);
}

int phub_fail(phub_ctx *ctx, int code, const std::string &msg) {
    if (ctx) ctx->err = msg;
    return code;
}

int phub_hip_fail(phub_ctx *ctx, hipError_t e, const char *what) {
    std::string m = std::string(what) + ": " + hipGetErrorString(e);
    if (ctx) ctx->err = m;
    return e == hipErrorOutOfMemory ? PHUB_ERR_OOM : PHUB_ERR_HIP;
}

static void host_binom(uint64_t *tab) {
    for (int n = 0; n <= PHUB_MAXN; ++n)
        for (int k = 0; k <= PHUB_MAXN; ++k) {
            uint64_t v = 0;
            if (k <= n) {
                v = 1;
                for (int i = 1; i <= k; ++i) v = v * (uint64_t)(n - k + i) / (uint64_t)i;
            }
            tab[n * (PHUB_MAXN + 1) + k] = v;
        }
}

extern "C" uint64_t phub_binom(int n, int k) {
    if (n < 0 || n > PHUB_MAXN || k < 0 || k > n) return 0;
    uint64_t v = 1;
    for (int i = 1; i <= k; ++i) v = v * (uint64_t)(n - k + i) / (uint64_t)i;
    return v;
}

extern "C" int phub_ctx_create(int device, void *stream, phub_ctx **out) {
    if (!out) return PHUB_ERR_ARG;
    phub_ctx *c = new phub_ctx();
    c->device = device;
    c->stream = (hipStream_t)stream;
    hipError_t e = hipSetDevice(device);
    if (e != hipSuccess) { int r = phub_hip_fail(c, e, "hipSetDevice"); *out = c; return r; }
    std::vector<uint64_t> tab((PHUB_MAXN + 1) * (PHUB_MAXN + 1));
    host_binom(tab.data());
    e = phub_malloc(&c->d_binom, tab.size() * sizeof(uint64_t));
    if (e == hipSuccess)
        e = phub_copy(c, c->d_binom, tab.data(), tab.size() * sizeof(uint64_t), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipEventCreate(&c->ev0);
    if (e == hipSuccess) e = hipEventCreate(&c->ev1);
    *out = c;
    if (e != hipSuccess) return phub_hip_fail(c, e, "phub_ctx_create");
    return phub_ensure_red(c, 1 << 16);
}

extern "C" void phub_ctx_destroy(phub_ctx *c) {
    if (!c) return;
    hipSetDevice(c->device);
    if (c->d_binom) hipFree(c->d_binom);
    if (c->d_red) hipFree(c->d_red);
    if (c->arena) hipFree(c->arena);
    if (c->capture_stream) hipStreamDestroy(c->capture_stream);
    if (c->h_lcz) hipHostFree(c->h_lcz);
    for (void *p : c->deferred) hipFree(p);
    if (c->h_red) hipHostFree(c->h_red);
    if (c->ev0) hipEventDestroy(c->ev0);
    if (c->ev1) hipEventDestroy(c->ev1);
    delete c;
}

extern "C" const char *phub_last_error(const phub_ctx *c) { return c ? c->err.c_str() : "null context"; }

extern "C" int phub_sync(phub_ctx *c) {
    if (!c) return PHUB_ERR_ARG;
    PHUB_HIP(c, hipStreamSynchronize(c->stream));
    return PHUB_OK;
}

extern "C" double phub_last_kernel_ms(const phub_ctx *c) { return c ? c->last_ms : 0.0; }

int phub_ensure_red(phub_ctx *c, size_t nd) {
    if (c->red_cap < nd) {
        phub_dfree(c, c->d_red);
        c->d_red = nullptr;
        c->red_cap = 0;
        PHUB_HIP(c, phub_malloc(&c->d_red, nd * sizeof(double)));
        c->red_cap = nd;
    }
    if (c->h_red_cap < 4096) {
        PHUB_HIP(c, phub_host_malloc(&c->h_red, 4096 * sizeof(double), hipHostMallocCoherent));
        c->h_red_cap = 4096;
    }
    return PHUB_OK;
}

// The context's scratch arena: small per-call buffers (RDM sums, Krylov
// coefficients, staged inputs) carved from one device block that grows to the
// largest request and is kept until the context is destroyed, so a call repeated
// on the same operator allocates nothing.  The old block is released (or kept,
// for a context that defers frees) before the new one is made; its previous
// users on this stream have been ordered before by hipFree's implicit wait.
int phub_arena(phub_ctx *c, size_t bytes, void **out) {
    bytes = (bytes + 255) & ~(size_t)255;
    if (c->arena_cap < bytes) {
        phub_dfree(c, c->arena);
        c->arena = nullptr;
        c->arena_cap = 0;
        PHUB_HIP(c, phub_malloc(&c->arena, bytes));
        c->arena_cap = bytes;
    }
    *out = c->arena;
    return PHUB_OK;
}

// ---------------------------------------------------------------------------
// reductions
// ---------------------------------------------------------------------------
// partial layout: [block][comp]; one workgroup per component sums all blocks.
// one block of 1024 threads per component; four independent partial sums per
// thread keep four loads in flight (the partial arrays of the H*v grids hold ~1e5
// entries per component)
__global__ __launch_bounds__(1024) void k_reduce_partials(const double *__restrict__ part,
                                                          int nblocks, int ncomp,
                                                          double *__restrict__ out) {
    __shared__ double sh[16];
    const int c = blockIdx.x;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    int b = threadIdx.x;
    for (; b + 3 * 1024 < nblocks; b += 4 * 1024) {
        s0 += part[(size_t)b * ncomp + c];
        s1 += part[(size_t)(b + 1024) * ncomp + c];
        s2 += part[(size_t)(b + 2048) * ncomp + c];
        s3 += part[(size_t)(b + 3072) * ncomp + c];
    }
    for (; b < nblocks; b += 1024) s0 += part[(size_t)b * ncomp + c];
    double s = (s0 + s1) + (s2 + s3);
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < 16; ++w) t += sh[w];
        out[c] = t;
    }
}

// the same sums into device memory, asynchronously (no host wait; graph-capturable)
int phub_reduce_partials_dev(phub_ctx *c, int nblocks, int ncomp, double *dev_out) {
    if ((size_t)nblocks * ncomp + ncomp > c->red_cap)
        return phub_fail(c, PHUB_ERR_ARG, "reduction buffer too small");
    hipLaunchKernelGGL(k_reduce_partials, dim3(ncomp), dim3(1024), 0, c->stream, c->d_red,
                       nblocks, ncomp, dev_out);
    PHUB_HIP(c, hipGetLastError());
    return PHUB_OK;
}

int phub_reduce_partials(phub_ctx *c, int nblocks, int ncomp, double *host_out) {
    if ((size_t)nblocks * ncomp + ncomp > c->red_cap || (size_t)ncomp > c->h_red_cap)
        return phub_fail(c, PHUB_ERR_ARG, "reduction buffer too small");
    // the sums go straight to the pinned host buffer (coherent host memory, mapped
    // into the device address space): no device-to-host copy after the kernel
    hipLaunchKernelGGL(k_reduce_partials, dim3(ncomp), dim3(1024), 0, c->stream, c->d_red,
                       nblocks, ncomp, c->h_red);
    PHUB_HIP(c, hipGetLastError());
    PHUB_HIP(c, hipStreamSynchronize(c->stream));
    std::memcpy(host_out, c->h_red, ncomp * sizeof(double));
    return PHUB_OK;
}

// ---------------------------------------------------------------------------
// basis: unrank (reference order) and rank
// ---------------------------------------------------------------------------
template <typename M>
__global__ void k_unrank(const uint64_t *__restrict__ binom, int n, int k, int64_t count,
                         M *__restrict__ out) {
    int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= count) return;
    uint64_t c = binom_lookup(binom, n, k) - 1 - (uint64_t)s;   // colex rank of rev_n(x)
    uint64_t x = 0;
    int p = n - 1;
    for (int i = k; i >= 1; --i) {
        while (binom_lookup(binom, p, i) > c) --p;             // largest p with C(p,i) <= c
        x |= 1ull << (n - 1 - p);
        c -= binom_lookup(binom, p, i);
        --p;
    }
    out[s] = (M)x;
}

__global__ void k_rank(const uint64_t *__restrict__ binom, int n, int k,
                       const uint64_t *__restrict__ masks, int64_t count,
                       int64_t *__restrict__ ranks) {
    int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= count) return;
    uint64_t x = masks[s];
    ranks[s] = (__popcll(x) == k) ? rank_ref_dev(binom, x, n, k) : -1;
}

static inline unsigned nblk(int64_t count, int bs) { return (unsigned)((count + bs - 1) / bs); }

extern "C" int phub_basis_build(phub_ctx *c, int n, int k, uint64_t *d_masks) {
    if (!c || !d_masks) return phub_fail(c, PHUB_ERR_ARG, "null argument");
    if (n < 0 || n > PHUB_MAXN || k < 0 || k > n) return phub_fail(c, PHUB_ERR_ARG, "bad (n,k)");
    int64_t cnt = (int64_t)phub_binom(n, k);
    if (cnt <= 0) return PHUB_OK;
    hipLaunchKernelGGL(k_unrank<uint64_t>, dim3(nblk(cnt, 256)), dim3(256), 0, c->stream,
                       c->d_binom, n, k, cnt, d_masks);
    PHUB_HIP(c, hipGetLastError());
    return PHUB_OK;
}

extern "C" int phub_basis_rank(phub_ctx *c, int n, int k, const uint64_t *d_masks,
                               int64_t count, int64_t *d_ranks) {
    if (!c || (!d_masks && count > 0) || (!d_ranks && count > 0))
        return phub_fail(c, PHUB_ERR_ARG, "null argument");
    if (n < 0 || n > PHUB_MAXN || k < 0 || k > n) return phub_fail(c, PHUB_ERR_ARG, "bad (n,k)");
    if (count <= 0) return PHUB_OK;
    hipLaunchKernelGGL(k_rank, dim3(nblk(count, 256)), dim3(256), 0, c->stream, c->d_binom, n,
                       k, d_masks, count, d_ranks);
    PHUB_HIP(c, hipGetLastError());
    return PHUB_OK;
}

// ---------------------------------------------------------------------------
// hop tables (Hubbard)
// ---------------------------------------------------------------------------
struct PairList {
    int np;
    const int *bj;        // bit of the annihilated site J  (n-J)
    const int *bk;        // bit of the created site K      (n-K)
    const double *t;      // T(J,K)
    const int16_t *pid;   // (J-1)*n+(K-1)
};

__global__ void k_hop_count(const uint32_t *__restrict__ masks, int64_t dim, PairList pl,
                            int32_t *__restrict__ cnt, int *__restrict__ maxh) {
    int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int c = 0;
    if (s < dim) {
        uint32_t x = masks[s];
        for (int p = 0; p < pl.np; ++p)
            c += (((x >> pl.bj[p]) & 1u) && !((x >> pl.bk[p]) & 1u)) ? 1 : 0;
        cnt[s] = c;
    }
    // wave max then one atomic per wave
    for (int off = 32; off > 0; off >>= 1) c = max(c, __shfl_xor(c, off));
    if ((threadIdx.x & 63) == 0) atomicMax(maxh, c);
}

// One wave per state: lane p of each 64-pair chunk tests pair p, a ballot gives every
// valid pair its slot (the number of valid pairs before it, so the entries keep the
// pair order), and only the valid lanes rank their target.  Launch with
// hop_fill_grid(dim) blocks of 256 threads (four states per block, grid-stride).
__global__ __launch_bounds__(256) void k_hop_fill(const uint64_t *__restrict__ binom, int n, int k,
                           const uint32_t *__restrict__ masks, int64_t dim, PairList pl, int maxh,
                           bool kmajor, int32_t *__restrict__ idx, double *__restrict__ c64,
                           float *__restrict__ c32, int16_t *__restrict__ pair,
                           double *__restrict__ eloc, const double *__restrict__ tdiag) {
    const int lane = threadIdx.x & 63;
    const int64_t nwaves = (int64_t)gridDim.x * 4;
    for (int64_t s = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6); s < dim; s += nwaves) {
        const uint32_t x = masks[s];
        int c = 0;   // entries of this state so far (wave-uniform)
        for (int p0 = 0; p0 < pl.np; p0 += 64) {
            const int p = p0 + lane;
            int bj = 0, bk = 0;
            bool ok = false;
            if (p < pl.np) {
                bj = pl.bj[p];
                bk = pl.bk[p];
                ok = ((x >> bj) & 1u) && !((x >> bk) & 1u);
            }
            const uint64_t m = __ballot(ok);
            if (ok) {
                const int slot = c + __popcll(m & ((1ull << lane) - 1ull));
                const uint32_t xn = x ^ (1u << bj) ^ (1u << bk);
                // JW sign: occupied sites strictly between J and K (ANTICOM(J,x)*ANTICOM(K,x'))
                const int lo = min(bj, bk), hi = max(bj, bk);
                const uint32_t between = (hi - lo > 1) ? (((1u << (hi - lo - 1)) - 1u) << (lo + 1)) : 0u;
                const double sg = (__popc(x & between) & 1) ? -1.0 : 1.0;
                const int64_t pos = rank_ref_dev(binom, xn, n, k);
                const size_t o = kmajor ? ((size_t)slot * dim + s) : ((size_t)s * maxh + slot);
                idx[o] = (int32_t)pos;
                const double cc = sg * pl.t[p];
                c64[o] = cc;
                c32[o] = (float)cc;
                if (pair) pair[o] = pl.pid[p];
            }
            c += __popcll(m);
        }
        if (lane == 0) {
            double e = 0.0;
            for (int b = 0; b < n; ++b)
                if ((x >> b) & 1u) e += tdiag[n - 1 - b];
            eloc[s] = e;
        }
    }
}

static inline unsigned hop_fill_grid(int64_t dim) {
    return (unsigned)std::max<int64_t>(1, std::min<int64_t>((dim + 3) / 4, 1 << 22));
}

void phub_free_hop_table(HopTable *t) {
    if (t->masks) hipFree(t->masks);
    if (t->eloc) hipFree(t->eloc);
    if (t->cnt) hipFree(t->cnt);
    if (t->idx) hipFree(t->idx);
    if (t->c64) hipFree(t->c64);
    if (t->c32) hipFree(t->c32);
    if (t->pair) hipFree(t->pair);
    if (t->packed) hipFree(t->packed);
    if (t->rowptr) hipFree(t->rowptr);
    if (t->csr) hipFree(t->csr);
    if (t->cls64) hipFree(t->cls64);
    if (t->cls32) hipFree(t->cls32);
    *t = HopTable();
}

void phub_free_hop_batch(HopBatch *b) {
    for (void *&p : b->arena)
        if (p) { hipFree(p); p = nullptr; }
    b->tabs.clear();
}

int phub_build_hop_tables(phub_ctx *c, const std::vector<HopSpec> &specs, HopBatch *out) {
    phub_free_hop_batch(out);
    const size_t nt = specs.size();
    out->tabs.assign(nt, HopTable());
    if (nt == 0) return PHUB_OK;
    // host staging of every table's pair list (the rules of phub_build_hop_table)
    std::vector<int> bj, bk;
    std::vector<double> tv, td;
    std::vector<int16_t> pid;
    std::vector<size_t> poff(nt), np(nt), toff(nt);
    std::vector<int64_t> dim(nt), moff(nt);
    int64_t mtot = 0;
    for (size_t t = 0; t < nt; ++t) {
        const HopSpec &h = specs[t];
        const int n = h.n;
        if (n < 1 || n > 32 || h.k < 0 || h.k > n) return phub_fail(c, PHUB_ERR_ARG, "hop table sector out of range");
        poff[t] = bj.size();
        toff[t] = td.size();
        for (int J = 1; J <= n; ++J) {
            const double dj = h.T[(J - 1) * n + (J - 1)];
            td.push_back(std::fabs(dj) >= 1e-14 ? dj : 0.0);
            for (int K = 1; K <= n; ++K) {
                if (J == K) continue;
                const double tjk = h.T[(J - 1) * n + (K - 1)];
                if (!(std::fabs(tjk) >= 1e-14)) continue;
                bj.push_back(n - J);
                bk.push_back(n - K);
                tv.push_back(tjk);
                pid.push_back((int16_t)((J - 1) * n + (K - 1)));
            }
        }
        np[t] = bj.size() - poff[t];
        dim[t] = (int64_t)phub_binom(n, h.k);
        moff[t] = mtot;
        mtot += dim[t];
    }
    // arena 0: pairs and per-table maxima, one copy
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    const size_t npair = std::max<size_t>(bj.size(), 1);
    const size_t o_bk = al(npair * sizeof(int)), o_t = o_bk + al(npair * sizeof(int));
    const size_t o_pid = o_t + al(npair * sizeof(double)), o_td = o_pid + al(npair * sizeof(int16_t));
    const size_t o_max = o_td + al(std::max<size_t>(td.size(), 1) * sizeof(double));
    const size_t bytes0 = o_max + al(nt * sizeof(int));
    std::vector<char> stage(bytes0, 0);
    if (!bj.empty()) {
        std::memcpy(stage.data(), bj.data(), bj.size() * sizeof(int));
        std::memcpy(stage.data() + o_bk, bk.data(), bk.size() * sizeof(int));
        std::memcpy(stage.data() + o_t, tv.data(), tv.size() * sizeof(double));
        std::memcpy(stage.data() + o_pid, pid.data(), pid.size() * sizeof(int16_t));
    }
    std::memcpy(stage.data() + o_td, td.data(), td.size() * sizeof(double));
    PHUB_HIP(c, phub_malloc(&out->arena[0], bytes0));
    char *a0 = (char *)out->arena[0];
    PHUB_HIP(c, hipMemcpyAsync(a0, stage.data(), bytes0, hipMemcpyHostToDevice, c->stream));
    // arena 1: masks (u32), eloc (f64), cnt (i32) of every table
    const size_t o_el = al((size_t)mtot * sizeof(uint32_t)), o_cn = o_el + al((size_t)mtot * sizeof(double));
    PHUB_HIP(c, phub_malloc(&out->arena[1], o_cn + al((size_t)mtot * sizeof(int32_t))));
    char *a1 = (char *)out->arena[1];
    auto pl_of = [&](size_t t) {
        return PairList{(int)np[t], (const int *)(a0) + poff[t], (const int *)(a0 + o_bk) + poff[t],
                        (const double *)(a0 + o_t) + poff[t], (const int16_t *)(a0 + o_pid) + poff[t]};
    };
    for (size_t t = 0; t < nt; ++t) {
        HopTable &h = out->tabs[t];
        h.n = specs[t].n;
        h.k = specs[t].k;
        h.kmajor = specs[t].kmajor;
        h.dim = dim[t];
        h.masks = (uint32_t *)a1 + moff[t];
        h.eloc = (double *)(a1 + o_el) + moff[t];
        h.cnt = (int32_t *)(a1 + o_cn) + moff[t];
        hipLaunchKernelGGL(k_unrank<uint32_t>, dim3(nblk(dim[t], 256)), dim3(256), 0, c->stream,
                           c->d_binom, h.n, h.k, dim[t], h.masks);
        hipLaunchKernelGGL(k_hop_count, dim3(nblk(dim[t], 256)), dim3(256), 0, c->stream, h.masks,
                           dim[t], pl_of(t), h.cnt, (int *)(a0 + o_max) + t);
    }
    PHUB_HIP(c, hipGetLastError());
    std::vector<int> mh(nt, 0);
    PHUB_HIP(c, hipMemcpyAsync(mh.data(), a0 + o_max, nt * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    PHUB_HIP(c, hipStreamSynchronize(c->stream));
    // arena 2: idx (i32), c64, c32 of every table
    std::vector<size_t> eoff(nt);
    size_t etot = 0;
    for (size_t t = 0; t < nt; ++t) {
        eoff[t] = etot;
        etot += (size_t)dim[t] * std::max(mh[t], 1);
    }
    const size_t o_c64 = al(etot * sizeof(int32_t)), o_c32 = o_c64 + al(etot * sizeof(double));
    const size_t bytes2 = o_c32 + al(etot * sizeof(float));
    PHUB_HIP(c, phub_malloc(&out->arena[2], bytes2));
    char *a2 = (char *)out->arena[2];
    PHUB_HIP(c, hipMemsetAsync(a2, 0, bytes2, c->stream));
    for (size_t t = 0; t < nt; ++t) {
        HopTable &h = out->tabs[t];
        h.maxh = mh[t];
        h.idx = (int32_t *)a2 + eoff[t];
        h.c64 = (double *)(a2 + o_c64) + eoff[t];
        h.c32 = (float *)(a2 + o_c32) + eoff[t];
        hipLaunchKernelGGL(k_hop_fill, dim3(hop_fill_grid(dim[t])), dim3(256), 0, c->stream, c->d_binom,
                           h.n, h.k, h.masks, dim[t], pl_of(t), mh[t], h.kmajor, h.idx, h.c64, h.c32,
                           (int16_t *)nullptr, h.eloc, (const double *)(a0 + o_td) + toff[t]);
    }
    PHUB_HIP(c, hipGetLastError());
    return PHUB_OK;
}

// compressed entries: (target index << 8) | class of the signed coefficient
__global__ void k_hop_pack(const int32_t *__restrict__ idx, const double *__restrict__ c64,
                           const int32_t *__restrict__ cnt, int64_t dim, int maxh,
                           const double *__restrict__ cls, int ncls,
                           uint32_t *__restrict__ packed) {
    int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= dim) return;
    const int c = cnt[s];
    for (int k = 0; k < maxh; ++k) {
        const size_t o = (size_t)k * dim + s;
        uint32_t e = 0;
        if (k < c) {
            int q = 0;
            while (q < ncls - 1 && cls[q] != c64[o]) ++q;
            // (idx * 8) << 11 | q * 8: the LDS byte offset of column idx (8 bytes per
            // column) and of the coefficient slot q (8 bytes per slot)
            e = ((uint32_t)idx[o] << 14) | ((uint32_t)q << 3);
        }
        packed[o] = e;
    }
}

// row-major table -> CSR of compressed entries
__global__ void k_hop_pack_csr(const int32_t *__restrict__ idx, const double *__restrict__ c64,
                               const int32_t *__restrict__ cnt, const int32_t *__restrict__ rowptr,
                               int64_t dim, int maxh, const double *__restrict__ cls, int ncls,
                               uint32_t *__restrict__ csr) {
    int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= dim) return;
    const int c = cnt[s];
    const int32_t base = rowptr[s];
    for (int k = 0; k < c; ++k) {
        const size_t o = (size_t)s * maxh + k;
        int q = 0;
        while (q < ncls - 1 && cls[q] != c64[o]) ++q;
        csr[base + k] = ((uint32_t)idx[o] << 8) | (uint32_t)q;
    }
}

int phub_build_hop_table(phub_ctx *c, int n, int k, const double *T, bool kmajor,
                         bool all_pairs, HopTable *out) {
    phub_free_hop_table(out);
    out->n = n;
    out->k = k;
    out->kmajor = kmajor;
    int64_t dim = (int64_t)phub_binom(n, k);
    out->dim = dim;
    if (dim > (int64_t)INT32_MAX) return phub_fail(c, PHUB_ERR_UNSUPPORTED, "species dimension exceeds int32");
    // directed pairs J != K in the reference loop order (hprod_hubbard.f90:26,36-37)
    std::vector<int> bj, bk;
    std::vector<double> tv, tdiag(n);
    std::vector<int16_t> pid;
    for (int J = 1; J <= n; ++J) {
        tdiag[J - 1] = (std::fabs(T[(J - 1) * n + (J - 1)]) >= 1e-14) ? T[(J - 1) * n + (J - 1)] : 0.0;
        for (int K = 1; K <= n; ++K) {
            if (J == K) continue;
            double t = T[(J - 1) * n + (K - 1)];
            if (!all_pairs && !(std::fabs(t) >= 1e-14)) continue;
            bj.push_back(n - J);
            bk.push_back(n - K);
            tv.push_back(all_pairs ? 1.0 : t);
            pid.push_back((int16_t)((J - 1) * n + (K - 1)));
        }
    }
    int np = (int)bj.size();
    int *d_bj = nullptr, *d_bk = nullptr, *d_max = nullptr;
    double *d_t = nullptr, *d_td = nullptr;
    int16_t *d_pid = nullptr;
    size_t npa = std::max(np, 1);
    PHUB_HIP(c, phub_malloc(&d_bj, npa * sizeof(int)));
    PHUB_HIP(c, phub_malloc(&d_bk, npa * sizeof(int)));
    PHUB_HIP(c, phub_malloc(&d_t, npa * sizeof(double)));
    PHUB_HIP(c, phub_malloc(&d_pid, npa * sizeof(int16_t)));
    PHUB_HIP(c, phub_malloc(&d_td, std::max(n, 1) * sizeof(double)));
    PHUB_HIP(c, phub_malloc(&d_max, sizeof(int)));
    if (np) {
        PHUB_HIP(c, phub_copy(c, d_bj, bj.data(), np * sizeof(int), hipMemcpyHostToDevice));
        PHUB_HIP(c, phub_copy(c, d_bk, bk.data(), np * sizeof(int), hipMemcpyHostToDevice));
        PHUB_HIP(c, phub_copy(c, d_t, tv.data(), np * sizeof(double), hipMemcpyHostToDevice));
        PHUB_HIP(c, phub_copy(c, d_pid, pid.data(), np * sizeof(int16_t), hipMemcpyHostToDevice));
    }
    if (n) PHUB_HIP(c, phub_copy(c, d_td, tdiag.data(), n * sizeof(double), hipMemcpyHostToDevice));
    PHUB_HIP(c, phub_clear(c, d_max, sizeof(int)));
    PHUB_HIP(c, phub_malloc(&out->masks, dim * sizeof(uint32_t)));
    PHUB_HIP(c, phub_malloc(&out->eloc, dim * sizeof(double)));
    PHUB_HIP(c, phub_malloc(&out->cnt, dim * sizeof(int32_t)));
    hipLaunchKernelGGL(k_unrank<uint32_t>, dim3(nblk(dim, 256)), dim3(256), 0, c->stream,
                       c->d_binom, n, k, dim, out->masks);
    PairList pl{np, d_bj, d_bk, d_t, d_pid};
    hipLaunchKernelGGL(k_hop_count, dim3(nblk(dim, 256)), dim3(256), 0, c->stream, out->masks,
                       dim, pl, out->cnt, d_max);
    PHUB_HIP(c, hipGetLastError());
    int mh = 0;
    PHUB_HIP(c, hipMemcpyAsync(&mh, d_max, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    PHUB_HIP(c, hipStreamSynchronize(c->stream));
    out->maxh = mh;
    size_t ne = (size_t)dim * std::max(mh, 1);
    PHUB_HIP(c, phub_malloc(&out->idx, ne * sizeof(int32_t)));
    PHUB_HIP(c, phub_malloc(&out->c64, ne * sizeof(double)));
    PHUB_HIP(c, phub_malloc(&out->c32, ne * sizeof(float)));
    PHUB_HIP(c, hipMemsetAsync(out->idx, 0, ne * sizeof(int32_t), c->stream));
    PHUB_HIP(c, hipMemsetAsync(out->c64, 0, ne * sizeof(double), c->stream));
    PHUB_HIP(c, hipMemsetAsync(out->c32, 0, ne * sizeof(float), c->stream));
    if (all_pairs) PHUB_HIP(c, phub_malloc(&out->pair, ne * sizeof(int16_t)));
    hipLaunchKernelGGL(k_hop_fill, dim3(hop_fill_grid(dim)), dim3(256), 0, c->stream, c->d_binom, n,
                       k, out->masks, dim, pl, mh, kmajor, out->idx, out->c64, out->c32,
                       out->pair, out->eloc, d_td);
    PHUB_HIP(c, hipGetLastError());
    PHUB_HIP(c, hipStreamSynchronize(c->stream));
    hipFree(d_bj); hipFree(d_bk); hipFree(d_t); hipFree(d_pid); hipFree(d_td); hipFree(d_max);
    // distinct signed coefficients (+-T(J,K)); compressed table when they fit a byte
    if (!all_pairs && mh > 0 && dim < (1ll << 24)) {
        std::vector<double> cls;
        for (double t : tv)
            for (double sg : {1.0, -1.0})
                if (std::find(cls.begin(), cls.end(), sg * t) == cls.end()) cls.push_back(sg * t);
        if (cls.size() <= 255) {
            out->ncls = (int)cls.size();
            std::vector<float> cls32(cls.begin(), cls.end());
            PHUB_HIP(c, phub_malloc(&out->cls64, cls.size() * sizeof(double)));
            PHUB_HIP(c, phub_malloc(&out->cls32, cls.size() * sizeof(float)));
            PHUB_HIP(c, phub_copy(c, out->cls64, cls.data(), cls.size() * sizeof(double), hipMemcpyHostToDevice));
            PHUB_HIP(c, phub_copy(c, out->cls32, cls32.data(), cls.size() * sizeof(float), hipMemcpyHostToDevice));
            if (kmajor && dim < (1ll << 18)) {   // the byte offset idx*8 must fit 21 bits
                PHUB_HIP(c, phub_malloc(&out->packed, (size_t)dim * mh * sizeof(uint32_t)));
                hipLaunchKernelGGL(k_hop_pack, dim3(nblk(dim, 256)), dim3(256), 0, c->stream,
                                   out->idx, out->c64, out->cnt, dim, mh, out->cls64, out->ncls,
                                   out->packed);
            } else if (!kmajor) {
                std::vector<int32_t> hcnt(dim), rp(dim + 1, 0);
                PHUB_HIP(c, phub_copy(c, hcnt.data(), out->cnt, dim * sizeof(int32_t), hipMemcpyDeviceToHost));
                for (int64_t s = 0; s < dim; ++s) rp[s + 1] = rp[s] + hcnt[s];
                PHUB_HIP(c, phub_malloc(&out->rowptr, (dim + 1) * sizeof(int32_t)));
                PHUB_HIP(c, phub_malloc(&out->csr, std::max<int64_t>(rp[dim], 1) * sizeof(uint32_t)));
                PHUB_HIP(c, phub_copy(c, out->rowptr, rp.data(), (dim + 1) * sizeof(int32_t), hipMemcpyHostToDevice));
                hipLaunchKernelGGL(k_hop_pack_csr, dim3(nblk(dim, 256)), dim3(256), 0, c->stream,
                                   out->idx, out->c64, out->cnt, out->rowptr, dim, mh, out->cls64,
                                   out->ncls, out->csr);
            }
            PHUB_HIP(c, hipGetLastError());
            PHUB_HIP(c, hipStreamSynchronize(c->stream));
        }
    }
    return PHUB_OK;
}

// ---------------------------------------------------------------------------
// Heisenberg tables: masks, bonds and the tables of the staged H*v
// ---------------------------------------------------------------------------
static void free_heis_stage(HeisStage *s);

void phub_free_heis(HeisTables *t) {
    free_heis_stage(&t->stg);
    if (t->masks) hipFree(t->masks);
    if (t->bj) hipFree(t->bj);
    if (t->bk) hipFree(t->bk);
    if (t->bJ) hipFree(t->bJ);
    if (t->bJr) hipFree(t->bJr);
    if (t->bsign) hipFree(t->bsign);
    if (t->bond_ising) hipFree(t->bond_ising);
    *t = HeisTables();
}

static void free_heis_stage(HeisStage *s) {
    hipFree(s->cnt);
    hipFree(s->goff);
    hipFree(s->first);
    hipFree(s->pos);
    hipFree(s->blk);
    hipFree(s->blkm);
    hipFree(s->hht);
    hipFree(s->xt);
    hipFree(s->cx);
    hipFree(s->ehh32);
    hipFree(s->ehh64);
    hipFree(s->ll);
    hipFree(s->hp);
    hipFree(s->xb);
    hipFree(s->ph);
    *s = HeisStage();
}

template <typename V>
static int stage_upload(phub_ctx *c, const std::vector<V> &h, V **d, double *bytes) {
    const size_t nb = std::max<size_t>(h.size(), 1) * sizeof(V);
    PHUB_HIP(c, phub_malloc(d, nb));
    if (!h.empty()) PHUB_HIP(c, phub_copy(c, *d, h.data(), h.size() * sizeof(V), hipMemcpyHostToDevice));
    *bytes += (double)nb;
    return PHUB_OK;
}

// Tables of the staged Heisenberg H*v.  Bonds come as bit positions p > q (bit = n - site)
// with their flip coefficient cf and half Ising weight hw.  The index of a state (xh, xl)
// is C(n, n/2) - 1 - (A(xh) + B(m, xl)): the closed form of the basis order above
// (compute_basis.f90:67-96) split at bit L, A summing the set bits of the high half
// counted from the top and B continuing that count into the low half.  The block of xl
// (all xh of popcount m = n/2 - popc(xl)) is contiguous; a state sits at position
// j = C(nh, m) - 1 - A(xh) from the block's first index.  Per-pattern tables are
// indexed by goff[m] + j, flip targets are 4 * (position in the source block).
static int build_heis_stage(phub_ctx *c, int n, int64_t dim, const std::vector<int> &bp,
                            const std::vector<int> &bq, const std::vector<double> &cf,
                            const std::vector<double> &hw, HeisStage *s) {
    free_heis_stage(s);
    int nh = std::min(n - n / 2, HEIS_STG_NHMAX);
    if (const char *e = std::getenv("PHUB_HEIS_NH")) {   // a smaller high half
        const int v = std::atoi(e);
        if (v >= 1 && v < nh) nh = v;
    }
    const int L = n - nh, k = n / 2;
    if (L > 30) return phub_fail(c, PHUB_ERR_UNSUPPORTED, "Heisenberg H*v needs n - 15 <= 30");
    s->nh = nh;
    s->L = L;
    s->k = k;
    const int64_t GS = (int64_t)1 << nh;
    s->GS = GS;
    s->cntmax = (int)phub_binom(nh, nh / 2);
    std::vector<uint64_t> bn((size_t)(n + 1) * (n + 1), 0);
    for (int a = 0; a <= n; ++a) {
        bn[(size_t)a * (n + 1)] = 1;
        for (int b = 1; b <= a; ++b)
            bn[(size_t)a * (n + 1) + b] = bn[(size_t)(a - 1) * (n + 1) + b - 1] +
                                          (b <= a - 1 ? bn[(size_t)(a - 1) * (n + 1) + b] : 0);
    }
    auto C = [&](int a, int b) -> uint64_t {
        return (b < 0 || b > a) ? 0 : bn[(size_t)a * (n + 1) + b];
    };
    std::vector<int32_t> cnt(nh + 1);
    std::vector<int64_t> goff(nh + 1);
    int64_t acc = 0;
    for (int m = 0; m <= nh; ++m) {
        cnt[m] = (int32_t)C(nh, m);
        goff[m] = acc;
        acc += cnt[m];
    }
    std::vector<uint32_t> pos(GS);   // position j of a high pattern in its block
    std::vector<int64_t> row(GS);
    for (int64_t h = 0; h < GS; ++h) {
        uint64_t a = 0;
        int i = 0;
        for (int b = nh - 1; b >= 0; --b)
            if ((h >> b) & 1) { ++i; a += C(nh - 1 - b, i); }
        const int m = __builtin_popcountll((uint64_t)h);
        pos[h] = (uint32_t)(cnt[m] - 1 - (int64_t)a);
        row[h] = goff[m] + pos[h];
    }
    std::vector<int> hh, ll, xx;
    for (size_t b = 0; b < bp.size(); ++b) {
        if (bq[b] >= L) hh.push_back((int)b);
        else if (bp[b] < L) ll.push_back((int)b);
        else xx.push_back((int)b);
    }
    const uint32_t zero = 4u * (uint32_t)s->cntmax;   // byte offset of the zero slot
    auto anti = [](uint64_t v, int p, int q) { return ((v >> p) ^ (v >> q)) & 1u; };
    // high-high: two targets per word, a missing flip points at the zero slot
    const int npair = ((int)hh.size() + 1) / 2;
    std::vector<uint32_t> hht((size_t)npair * GS, 0);
    std::vector<HeisStgPair> hp(npair);
    std::vector<double> ehh(GS, 0.0);
    for (int pi = 0; pi < npair; ++pi)
        for (int e = 0; e < 2; ++e) {
            const int bi = 2 * pi + e;
            const bool have = bi < (int)hh.size();
            const int b = have ? hh[bi] : 0;
            hp[pi].c64[e] = have ? cf[b] : 0.0;
            hp[pi].c32[e] = (float)hp[pi].c64[e];
            const int p = have ? bp[b] - L : 0, q = have ? bq[b] - L : 0;
            for (int64_t h = 0; h < GS; ++h) {
                uint32_t t = zero;
                if (have && anti((uint64_t)h, p, q)) {
                    t = 4u * pos[h ^ (((int64_t)1 << p) | ((int64_t)1 << q))];
                    ehh[row[h]] += hw[b];
                }
                hht[(size_t)pi * GS + row[h]] |= t << (16 * e);
            }
        }
    // cross: the distinct high sites, and one phase per distinct low site
    std::vector<int> xs;
    for (int b : xx)
        if (std::find(xs.begin(), xs.end(), bp[b]) == xs.end()) xs.push_back(bp[b]);
    std::sort(xs.begin(), xs.end());
    const int ncx = (int)xs.size();
    if (ncx > HEIS_STG_NXMAX) return phub_fail(c, PHUB_ERR_UNSUPPORTED, "too many cross sites");
    // per high site: 4 * the target's position in the source block, bit 15 set when
    // the pattern's high bit is 0.  A target flips only where its high bit differs
    // from its low bit tb, i.e. where bit 15 of (entry ^ (tb << 15)) is clear; the
    // kernel takes min(entry ^ (tb << 15), zero slot), which sends the others to the
    // zero slot (4 * position < 2^15 <= every entry with bit 15 set).
    std::vector<uint16_t> xt((size_t)std::max(ncx, 1) * GS, 0);
    std::vector<uint32_t> cxp(GS, 0);
    for (int jx = 0; jx < ncx; ++jx) {
        const int hj = xs[jx] - L;
        for (int64_t h = 0; h < GS; ++h) {
            const uint32_t hb = (uint32_t)(h >> hj) & 1u;
            cxp[row[h]] |= hb << jx;
            const uint32_t t = 4u * pos[h ^ ((int64_t)1 << hj)];
            xt[(size_t)jx * GS + row[h]] = (uint16_t)(t | ((1u - hb) << 15));
        }
    }
    std::vector<int> ks;
    for (int b : xx)
        if (std::find(ks.begin(), ks.end(), bq[b]) == ks.end()) ks.push_back(bq[b]);
    std::sort(ks.begin(), ks.end());
    std::vector<HeisStgX> xb;
    std::vector<HeisStgPhase> ph;
    for (int kk : ks) {
        HeisStgPhase P{kk, (int)xb.size(), 0, 0};
        for (int b : xx) {
            if (bq[b] != kk) continue;
            const int jx = (int)(std::find(xs.begin(), xs.end(), bp[b]) - xs.begin());
            xb.push_back(HeisStgX{jx, kk, (float)cf[b], (float)hw[b], cf[b], hw[b]});
            ++P.nb;
        }
        ph.push_back(P);
    }
    std::vector<HeisStgLL> lv;
    for (int b : ll) lv.push_back(HeisStgLL{bp[b], bq[b], (float)cf[b], (float)hw[b], cf[b], hw[b]});
    // blocks: every xl whose high popcount m = n/2 - popc(xl) exists, in memory order
    const int64_t NL = (int64_t)1 << L;
    std::vector<int64_t> first(NL, -1);
    std::vector<std::pair<int64_t, uint32_t>> blocks;
    for (int64_t xl = 0; xl < NL; ++xl) {
        const int m = k - __builtin_popcountll((uint64_t)xl);
        if (m < 0 || m > nh) continue;
        uint64_t B = 0;
        int i = m;
        for (int b = L - 1; b >= 0; --b)
            if ((xl >> b) & 1) { ++i; B += C(n - 1 - b, i); }
        first[xl] = dim - 1 - (int64_t)B - (cnt[m] - 1);
        blocks.push_back({first[xl], (uint32_t)xl});
    }
    std::sort(blocks.begin(), blocks.end());
    std::vector<uint32_t> blk(blocks.size()), blkm;
    for (size_t i = 0; i < blocks.size(); ++i) blk[i] = blocks[i].second;
    blkm.reserve(blk.size());
    for (int m = 0; m <= nh; ++m)   // popcount-major: one popcount's tables at a time
        for (uint32_t xl : blk)
            if (k - __builtin_popcount(xl) == m) blkm.push_back(xl);
    std::vector<float> ehh32(ehh.begin(), ehh.end());
    s->ncx = ncx;
    s->nlo = ncx / 2;
    s->nll = (int)lv.size();
    s->npair = npair;
    s->nx = (int)xb.size();
    s->nphase = (int)ph.size();
    s->nblk = (int64_t)blk.size();
    double by = 0.0;
    int rc = 0;
    if ((rc = stage_upload(c, cnt, &s->cnt, &by)) || (rc = stage_upload(c, goff, &s->goff, &by)) ||
        (rc = stage_upload(c, first, &s->first, &by)) || (rc = stage_upload(c, pos, &s->pos, &by)) || (rc = stage_upload(c, blk, &s->blk, &by)) ||
        (rc = stage_upload(c, blkm, &s->blkm, &by)) ||
        (rc = stage_upload(c, hht, &s->hht, &by)) || (rc = stage_upload(c, xt, &s->xt, &by)) ||
        (rc = stage_upload(c, cxp, &s->cx, &by)) || (rc = stage_upload(c, ehh32, &s->ehh32, &by)) ||
        (rc = stage_upload(c, ehh, &s->ehh64, &by)) || (rc = stage_upload(c, lv, &s->ll, &by)) ||
        (rc = stage_upload(c, hp, &s->hp, &by)) || (rc = stage_upload(c, xb, &s->xb, &by)) ||
        (rc = stage_upload(c, ph, &s->ph, &by)))
        return rc;
    s->bytes = by;
    return PHUB_OK;
}

int phub_build_heis(phub_ctx *c, int n, const double *Jm, int conv, HeisTables *out) {
    phub_free_heis(out);
    if (n < 2 || n > 40) return phub_fail(c, PHUB_ERR_UNSUPPORTED, "Heisenberg needs 2 <= n <= 40");
    int64_t dim = (int64_t)phub_binom(n, n / 2);
    out->n = n;
    out->dim = dim;
    // the mask table is only kept when small (reduced quantities use it); the H*v
    // enumerates blocks of equal low-half pattern instead
    if (dim <= (1ll << 28)) {
        PHUB_HIP(c, phub_malloc(&out->masks, dim * sizeof(uint64_t)));
        hipLaunchKernelGGL(k_unrank<uint64_t>, dim3(nblk(dim, 256)), dim3(256), 0, c->stream,
                           c->d_binom, n, n / 2, dim, out->masks);
        PHUB_HIP(c, hipGetLastError());
    }
    // bonds: unordered {J<K} with either J(J,K) or J(K,J) non-zero
    std::vector<int> bj, bk;
    std::vector<double> bJ, bJr, bs, bis;
    double dconst = 0.0;
    for (int J = 1; J <= n; ++J) {
        double jj = Jm[(J - 1) * n + (J - 1)];
        if (std::fabs(jj) > 1e-14) dconst += 0.5 * jj;      // hprod_heis.f90:11-21
        dconst += 0.25 * jj;                               // Ising J=K term (:54-58)
        for (int K = J + 1; K <= n; ++K) {
            double jk = Jm[(J - 1) * n + (K - 1)], kj = Jm[(K - 1) * n + (J - 1)];
            if (jk == 0.0 && kj == 0.0) continue;
            bj.push_back(n - J);
            bk.push_back(n - K);
            // flip element for the source with the spin-up at the first site of the
            // ordered pair: -(-1)^{J+K} J(J,K) (hprod_heis.f90:39-47)
            bJ.push_back(std::fabs(jk) > 1e-14 ? jk : 0.0);
            bJr.push_back(std::fabs(kj) > 1e-14 ? kj : 0.0);
            // reference HPROD: -(-1)^{J+K}; proper spin model in this fermion basis
            // (the sign J_SPIN measures, static_rq_heis.f90:94-102): +(-1)^{J+K}
            {
                const double pr = ((J + K) & 1) ? -1.0 : 1.0;
                bs.push_back(conv == 0 ? -pr : pr);
            }
            bis.push_back(jk + kj);
            dconst += 0.25 * (jk + kj);
        }
    }
    int nb = (int)bj.size();
    out->nb = nb;
    out->diag_const = dconst;
    {
        std::vector<double> cf(nb), hw(nb);
        for (int b = 0; b < nb; ++b) {
            cf[b] = bs[b] * bJ[b];
            hw[b] = 0.5 * bis[b];
        }
        int rc = build_heis_stage(c, n, dim, bj, bk, cf, hw, &out->stg);
        if (rc) return rc;
    }
    size_t nba = std::max(nb, 1);
    PHUB_HIP(c, phub_malloc(&out->bj, nba * sizeof(int)));
    PHUB_HIP(c, phub_malloc(&out->bk, nba * sizeof(int)));
    PHUB_HIP(c, phub_malloc(&out->bJ, nba * sizeof(double)));
    PHUB_HIP(c, phub_malloc(&out->bJr, nba * sizeof(double)));
    PHUB_HIP(c, phub_malloc(&out->bsign, nba * sizeof(double)));
    PHUB_HIP(c, phub_malloc(&out->bond_ising, nba * sizeof(double)));
    if (nb) {
        PHUB_HIP(c, phub_copy(c, out->bj, bj.data(), nb * sizeof(int), hipMemcpyHostToDevice));
        PHUB_HIP(c, phub_copy(c, out->bk, bk.data(), nb * sizeof(int), hipMemcpyHostToDevice));
        PHUB_HIP(c, phub_copy(c, out->bJ, bJ.data(), nb * sizeof(double), hipMemcpyHostToDevice));
        PHUB_HIP(c, phub_copy(c, out->bJr, bJr.data(), nb * sizeof(double), hipMemcpyHostToDevice));
        PHUB_HIP(c, phub_copy(c, out->bsign, bs.data(), nb * sizeof(double), hipMemcpyHostToDevice));
        PHUB_HIP(c, phub_copy(c, out->bond_ising, bis.data(), nb * sizeof(double), hipMemcpyHostToDevice));
    }
    PHUB_HIP(c, hipStreamSynchronize(c->stream));
    return PHUB_OK;
}
