// phub_internal.h -- internal types shared by the HIP sources of libphub.so.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <atomic>
#include <string>
#include <vector>
#include "../../include/phub.h"

#define PHUB_MAXN 64

struct phub_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double last_ms = 0.0;
    uint64_t *d_binom = nullptr;     // (PHUB_MAXN+1)^2 binomials on the device
    double *d_red = nullptr;         // reduction partials
    size_t red_cap = 0;              // doubles
    double *h_red = nullptr;         // pinned host mirror of reduction results
    size_t h_red_cap = 0;
    void *arena = nullptr;           // scratch arena (phub_arena)
    size_t arena_cap = 0;            // bytes
    // set only while a Lanczos step is captured into a graph: the Hubbard H*v
    // epilogue then reads its scale and beta from this device array ([0], [1])
    const double *dev_scalars = nullptr;
    // the stream graph captures are recorded on (lazily created; the graphs are then
    // launched on `stream`, which may be the legacy null stream)
    hipStream_t capture_stream = nullptr;
    // pinned (coherent) host array the graph-captured Lanczos steps write alpha and
    // beta into; grown on demand and kept
    double *h_lcz = nullptr;
    size_t h_lcz_cap = 0;            // doubles
    // hipFree waits for every stream of the device: a context that works beside a
    // long kernel on another stream sets defer_free and its device memory is kept
    // until the context is destroyed
    bool defer_free = false;
    std::vector<void *> deferred;
    // down-hop tiles the tuner picked on this context, by operator shape: operators
    // with the same sector dimensions and maximal hop counts reuse the pick
    struct TunedTile {
        int64_t dim_up, dim_dn;
        int maxh_up, maxh_dn, fp64, cfg;
    };
    std::vector<TunedTile> tuned;
};

// Copies and clears ordered on the context's stream.  The synchronous hipMemcpy /
// hipMemset run on the null stream, which neither waits for nor is waited on by a
// non-blocking stream (torch's side streams are such streams): a table uploaded or
// a counter cleared that way can still be in flight when a kernel queued on
// c->stream reads it.  phub_copy returns when the copy is complete, so the host
// buffer may be reused or freed at once.
inline hipError_t phub_copy(phub_ctx *c, void *dst, const void *src, size_t bytes,
                            hipMemcpyKind kind) {
    hipError_t e = hipMemcpyAsync(dst, src, bytes, kind, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    return e;
}
inline hipError_t phub_clear(phub_ctx *c, void *dst, size_t bytes) {
    return hipMemsetAsync(dst, 0, bytes, c->stream);
}

// Every device and pinned-host allocation of the library goes through these two
// wrappers, which count them (phub_alloc_count): the tests of the allocation
// contract read the count around a call.
extern std::atomic<unsigned long long> g_phub_alloc_count;
template <typename P>
inline hipError_t phub_malloc(P **p, size_t bytes) {
    g_phub_alloc_count.fetch_add(1, std::memory_order_relaxed);
    return hipMalloc(reinterpret_cast<void **>(p), bytes);
}
template <typename P>
inline hipError_t phub_host_malloc(P **p, size_t bytes, unsigned flags) {
    g_phub_alloc_count.fetch_add(1, std::memory_order_relaxed);
    return hipHostMalloc(reinterpret_cast<void **>(p), bytes, flags);
}

// Frees device memory that context c allocated (c may be null: plain hipFree).
inline void phub_dfree(phub_ctx *c, void *p) {
    if (!p) return;
    if (c && c->defer_free) c->deferred.push_back(p);
    else hipFree(p);
}

// Per-species hop table: for every state s the list of (target index, coefficient)
// produced by c^dag_K c_J with J != K and |T(J,K)| >= 1e-14 (hprod_hubbard.f90:37,39-50).
// Layout KMAJOR: entry k of state s at [k*dim + s] (lane-varying consumers);
// ROWMAJOR: at [s*maxh + k] (wave-uniform consumers).
struct HopTable {
    int n = 0, k = 0;
    int64_t dim = 0;
    int maxh = 0;
    bool kmajor = true;
    uint32_t *masks = nullptr;    // device, dim
    double *eloc = nullptr;       // device, dim: sum of T(J,J) over occupied J
    int32_t *cnt = nullptr;       // device, dim
    int32_t *idx = nullptr;       // device, dim*maxh
    double *c64 = nullptr;        // device, dim*maxh
    float *c32 = nullptr;         // device, dim*maxh
    int16_t *pair = nullptr;      // device, dim*maxh, (J-1)*n+(K-1) (all-pairs tables only)
    // compressed entries when the signed hopping values take at most 255 distinct
    // values: k-major ((idx * 8) << 8 | class, the target's byte offset in an LDS
    // array of 8-byte columns), CSR (idx << 8 | class)
    uint32_t *packed = nullptr;   // device, dim*maxh (k-major tables, dim < 2^21)
    int32_t *rowptr = nullptr;    // device, dim+1 (row-major tables: CSR of packed entries)
    uint32_t *csr = nullptr;      // device, rowptr[dim]
    int ncls = 0;
    double *cls64 = nullptr;      // device, ncls signed coefficient values
    float *cls32 = nullptr;
};

// Staged Heisenberg H*v (k_hv_heis_stg).  The high half holds the top nh bits of a
// state (sites 1..nh), nh <= HEIS_STG_NHMAX so that one block of C(nh, nh/2) states
// and a neighbouring block fit in LDS together in FP64, and so that LDS byte offsets
// of FP32 positions (4 * j < 2^16) fit the 16-bit flip targets.
constexpr int HEIS_STG_NHMAX = 15;
constexpr int HEIS_STG_CNTMAX = 6435;    // C(15, 7)
constexpr int HEIS_STG_NXMAX = 16;       // cross high sites (<= nh)

struct HeisStgLL {      // low-low bond: low bit positions p > q
    int p, q;
    float c32, h32;     // flip coefficient, half the Ising weight
    double c64, h64;
};
struct HeisStgPair {    // two high-high bonds sharing one word of the target table
    float c32[2];
    double c64[2];
};
struct HeisStgX {       // cross bond: high site index jx (into the cross-site list), low bit kk
    int jx, kk;
    float c32, h32;
    double c64, h64;
};
struct HeisStgPhase {   // cross bonds [b0, b0 + nb) share the low bit kk
    int kk, b0, nb, pad;
};

struct HeisStage {
    int nh = 0, L = 0, k = 0;
    int cntmax = 0;                // C(nh, nh/2): index of the zero slot
    int ncx = 0, nlo = 0;          // cross high sites; the first nlo index the low Ising table
    int nll = 0, npair = 0, nx = 0, nphase = 0;
    int64_t GS = 0;                // 2^nh: per-pattern table stride
    int64_t nblk = 0;              // valid low patterns
    int32_t *cnt = nullptr;        // device, nh+1: C(nh, m)
    int64_t *goff = nullptr;       // device, nh+1: first table row of popcount m
    int64_t *first = nullptr;      // device, 2^L: first index of the block of xl (-1: none)
    uint32_t *pos = nullptr;       // device, 2^nh: position of a high pattern in its block
    uint32_t *blk = nullptr;       // device, nblk: valid xl in memory order
    uint32_t *blkm = nullptr;      // device, nblk: the same ordered by popcount, then memory
    uint32_t *hht = nullptr;       // device, npair x GS: two 16-bit high-high flip targets
    uint16_t *xt = nullptr;        // device, ncx x GS: cross flip targets and direction bit
    uint32_t *cx = nullptr;        // device, GS: bits of xh at the cross high sites
    float *ehh32 = nullptr;        // device, GS: half Ising weight of antiparallel high-high bonds
    double *ehh64 = nullptr;
    HeisStgLL *ll = nullptr;       // device, nll
    HeisStgPair *hp = nullptr;     // device, npair
    HeisStgX *xb = nullptr;        // device, nx (phase order)
    HeisStgPhase *ph = nullptr;    // device, nphase
    double bytes = 0.0;            // device bytes of the tables above
};

struct HeisTables {
    int n = 0;
    int64_t dim = 0;
    uint64_t *masks = nullptr;     // device, dim (kept when dim <= 2^28)
    int nb = 0;                    // number of undirected bonds
    int *bj = nullptr, *bk = nullptr;   // device, bit positions of the two sites
    double *bJ = nullptr;          // device, J(J,K) used when site J is up in the source
    double *bJr = nullptr;         // device, J(K,J)
    double *bsign = nullptr;       // device, flip sign factor per bond
    double diag_const = 0.0;       // sum of 0.75 J(J,J) + 0.5 * (on-site) terms
    double *bond_ising = nullptr;  // device, Ising weight per bond
    HeisStage stg;                 // tables of the H*v
};

// A down-hop entry of the buffer-load kernel: the byte offset of the source row in the
// state (source index * dim_up * sizeof(T), below 2^32) and the signed coefficient.
template <typename T>
struct DnEnt;
template <>
struct DnEnt<float> {
    uint32_t off;
    float c;
};
template <>
struct DnEnt<double> {
    uint32_t off, pad;
    double c;
};

struct phub_op {
    int kind = 0;                  // 0 Hubbard, 1 Heisenberg
    int n = 0, nup = 0, ndn = 0;
    int64_t dim_up = 0, dim_dn = 0, dim = 0;
    // Hubbard
    HopTable up, dn;               // up KMAJOR, dn ROWMAJOR (hop tables of T)
    HopTable rup, rdn;             // all-pairs tables for reduced quantities (lazy)
    // down-spin pair map (lazy, when it fits 1 GiB): for the ordered pair p = (J-1)*n +
    // (K-1) and down state d, (d' << 1) | sign of c^dag_J c_K d, or -1; [d][p]
    int32_t *dn_pairmap = nullptr;
    double U[PHUB_MAXN];
    int u_uniform = 1;
    double u0 = 0.0;
    // 4-index interaction (is_Uloc = 3): device n^4, entries below the reference's
    // thresholds zeroed; the all-pairs tables rup / rdn are built with it
    double *d_ul = nullptr;
    std::vector<double> T;         // host copy, row-major
    // Heisenberg
    HeisTables h;
    std::vector<double> J;         // host copy
    int sign_convention = 0;
    double table_bytes = 0.0;
    phub_ctx *ctx = nullptr;
    // down-hop entries with the source row's byte offset in the state (CSR order of
    // dn.rowptr), per precision (FP32 DnEnt<float>, FP64 DnEnt<double>); null when the
    // state's bytes do not fit 32 bits or the down table has no CSR
    void *dn_ent[2] = {nullptr, nullptr};
    // down-hop tile chosen by the H*v tuner, per precision (FP32, FP64); -1 = not yet
    int tune_dn[2] = {-1, -1};
    // Heisenberg work-item order chosen by the tuner (0 XCD-contiguous, 1 round robin)
    int tune_heis[2] = {-1, -1};
    // Heisenberg: the all-pairs view for spin correlations (lazy, owned)
    phub_op *heis_all = nullptr;
};

// error helpers ------------------------------------------------------------------
int phub_fail(phub_ctx *ctx, int code, const std::string &msg);
int phub_hip_fail(phub_ctx *ctx, hipError_t e, const char *what);

#define PHUB_HIP(ctx, call)                                                   \
    do {                                                                      \
        hipError_t _e = (call);                                               \
        if (_e != hipSuccess) return phub_hip_fail((ctx), _e, #call);         \
    } while (0)

// shared device helpers ----------------------------------------------------------
__host__ __device__ inline uint64_t binom_lookup(const uint64_t *tab, int n, int k) {
    return (k < 0 || k > n) ? 0 : tab[n * (PHUB_MAXN + 1) + k];
}

// reference-order rank: C(n,k) - 1 - colex(rev_n(x)), closed form of HILBERT's order
__device__ inline int64_t rank_ref_dev(const uint64_t *binom, uint64_t x, int n, int k) {
    uint64_t c = 0;
    int i = 0;
    for (int b = n - 1; b >= 0; --b) {
        if ((x >> b) & 1ull) {
            ++i;
            c += binom_lookup(binom, n - 1 - b, i);
        }
    }
    return (int64_t)(binom_lookup(binom, n, k) - 1 - c);
}

// internal launches (implemented in the .hip files) ------------------------------
int phub_build_hop_table(phub_ctx *ctx, int n, int k, const double *T_host, bool kmajor,
                         bool all_pairs, HopTable *out);
void phub_free_hop_table(HopTable *t);
// Many uncompressed hop tables at once (masks, eloc, cnt, idx, c64, c32 of each),
// in three device arenas: one host-device synchronisation for all of them.  The
// HopTable entries point into the arenas and must not be freed one by one.
struct HopSpec {
    int n, k;
    const double *T;      // host, n*n
    bool kmajor;
};
struct HopBatch {
    std::vector<HopTable> tabs;
    void *arena[3] = {nullptr, nullptr, nullptr};
};
int phub_build_hop_tables(phub_ctx *ctx, const std::vector<HopSpec> &specs, HopBatch *out);
void phub_free_hop_batch(HopBatch *b);
// down-hop entries of the buffer-load down kernel (op->dn_ent), both precisions (hv.hip)
int phub_build_dn_ent(phub_ctx *ctx, phub_op *op);
int phub_build_heis(phub_ctx *ctx, int n, const double *J_host, int sign_convention,
                    HeisTables *out);
void phub_free_heis(HeisTables *t);

// the context's scratch arena: a device block of at least `bytes`, valid until the
// next phub_arena call on this context (one request per call: carve it)
int phub_arena(phub_ctx *ctx, size_t bytes, void **out);

// reductions: sums of `ncomp` interleaved per-block partials -> host (synchronous)
int phub_reduce_partials(phub_ctx *ctx, int nblocks, int ncomp, double *host_out);
// the same sums into device memory, asynchronously (graph-capturable)
int phub_reduce_partials_dev(phub_ctx *ctx, int nblocks, int ncomp, double *dev_out);
int phub_ensure_red(phub_ctx *ctx, size_t ndoubles);

// H*v launchers with a fused Lanczos epilogue:
//   out = scale * (H x) - beta * prev   (prev may be null)
//   partials: <x_scaled, out>, <out, out> per block (when want_dots)
// all-pairs hop tables of an operator (reduced quantities, 4-index interaction)
int phub_ensure_pair_tables(phub_ctx *ctx, phub_op *op);
// the 4-index interaction term, added to out (hv.hip)
template <typename T>
int phub_hv_ul(phub_ctx *ctx, const phub_op *op, const T *x, T *out, double scale,
               const T *prev, double beta, bool epi, int *nblocks_out);

// tuner tile index -> (segment bytes * 1000 + rows), -1 if untuned (hv.hip)
int phub_dn_tile_code(int cfg, bool fp64);

template <typename T>
int phub_hv_launch(phub_ctx *ctx, const phub_op *op, const T *x, T *out, double scale,
                   const T *prev, double beta, bool want_dots, int *nblocks_out);

// minv = 1 / max(diag(H) - e, 1e-3) of a one-body Hubbard operator, dim entries of V
// (float or double; hv.hip); PHUB_ERR_UNSUPPORTED for other operators
template <typename V>
int phub_hub_minv(phub_ctx *ctx, const phub_op *op, double e, V *minv);
// the Jacobi preconditioner of the MINRES solves applies to this operator (one-body
// Hubbard) and is not switched off (PHUB_JD_PRECOND=0)
bool phub_minres_precond(const phub_op *op);
