// gf.hip -- Green's-function building blocks.
//
// Excitation vectors c^dag_i psi / c_i psi into the N+-1 sector: reference
// GF_E_UP / GF_E_DOWN / GF_H_UP / GF_H_DOWN (pyhub/libs/src/spgf_ed.f90:304-429).
// Sign: ANTICOM(i, x_spin) (occupied sites before i), times (-1)^{N_dn} for up
// operators (:321, :386).  Pull form: each target element finds its unique source.
//
// Continued-fraction evaluation of a diagonal G(omega) from Lanczos coefficients:
// the single-vector case of the band-Lanczos poles/weights (spgf_bandlanczos.f90:309-448)
// summed as Pole.retarded does (pyhub/tools/pole.py:81-88).
#include "phub_internal.h"

#include <cmath>
#include <mutex>
#include <string>
#include <vector>

__global__ __launch_bounds__(256) void k_excite(const uint64_t *__restrict__ binom, int n,
                                                int bit, int spin, int create, double parity,
                                                int nup_s, int ndn_s, int64_t nu_s,
                                                const uint32_t *__restrict__ mu_t,
                                                const uint32_t *__restrict__ md_t, int64_t nu_t,
                                                int64_t nd_t, const double *__restrict__ psi,
                                                double *__restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nu_t * nd_t) return;
    const int64_t d = i / nu_t, u = i % nu_t;
    const uint32_t b = 1u << bit;
    const uint32_t above = (bit + 1 < 32) ? ~((b << 1) - 1u) : 0u;   // sites before `site`
    double v = 0.0;
    if (spin == 0) {
        const uint32_t xt = mu_t[u];
        const bool occ = (xt & b) != 0;
        if (create ? occ : !occ) {
            const uint32_t xs = xt ^ b;
            const int64_t us = rank_ref_dev(binom, xs, n, nup_s);
            const double ac = (__popc(xs & above) & 1) ? -1.0 : 1.0;
            v = parity * ac * psi[d * nu_s + us];
        }
    } else {
        const uint32_t xt = md_t[d];
        const bool occ = (xt & b) != 0;
        if (create ? occ : !occ) {
            const uint32_t xs = xt ^ b;
            const int64_t ds = rank_ref_dev(binom, xs, n, ndn_s);
            const double ac = (__popc(xs & above) & 1) ? -1.0 : 1.0;
            v = ac * psi[ds * nu_s + u];
        }
    }
    out[i] = v;
}

extern "C" int phub_excite(phub_ctx *c, const phub_op *src, const phub_op *dst, int site,
                           int spin, int create, const double *psi, double *out) {
    if (!c || !src || !dst || !psi || !out) return phub_fail(c, PHUB_ERR_ARG, "null argument");
    if (src->kind != 0 || dst->kind != 0) return phub_fail(c, PHUB_ERR_ARG, "Hubbard operators required");
    const int n = src->n;
    if (dst->n != n || site < 1 || site > n || (spin != 0 && spin != 1))
        return phub_fail(c, PHUB_ERR_ARG, "bad site / spin / cluster size");
    const int dn = create ? 1 : -1;
    const int eu = src->nup + (spin == 0 ? dn : 0), ed = src->ndn + (spin == 1 ? dn : 0);
    if (dst->nup != eu || dst->ndn != ed)
        return phub_fail(c, PHUB_ERR_ARG, "target sector does not match the excitation");
    if (dst->dim == 0) return PHUB_OK;
    const double parity = (spin == 0 && (src->ndn % 2 == 1)) ? -1.0 : 1.0;
    hipLaunchKernelGGL(k_excite, dim3((unsigned)((dst->dim + 255) / 256)), dim3(256), 0, c->stream,
                       c->d_binom, n, n - site, spin, create ? 1 : 0, parity, src->nup, src->ndn,
                       src->dim_up, dst->up.masks, dst->dn.masks, dst->dim_up, dst->dim_dn, psi,
                       out);
    PHUB_HIP(c, hipGetLastError());
    return PHUB_OK;
}

// G = sign * norm2 / (z - a0 - b0^2 / (z - a1 - b1^2 / ...)), z = sign*(w + i eta) + shift
__global__ void k_cf_eval(int nch, int m, const double *__restrict__ al,
                          const double *__restrict__ be, const double *__restrict__ n2,
                          const double *__restrict__ shift, const double *__restrict__ sgn,
                          int nw, const double *__restrict__ w, double eta,
                          double *__restrict__ out) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)nch * nw) return;
    const int ch = (int)(t / nw), iw = (int)(t % nw);
    const double s = sgn[ch];
    const double zr = s * w[iw] + shift[ch], zi = s * eta;
    const double *a = al + (size_t)ch * m, *b = be + (size_t)ch * m;
    // find the effective chain length (trailing zeros after a breakdown)
    int len = m;
    while (len > 1 && a[len - 1] == 0.0 && b[len - 2] == 0.0) --len;
    double fr = 0.0, fi = 0.0;   // f = 1/(z - a_k - b_k^2 f_{k+1})
    for (int k = len - 1; k >= 0; --k) {
        const double bb = (k < len - 1) ? b[k] * b[k] : 0.0;
        const double dr = zr - a[k] - bb * fr, di = zi - bb * fi;
        const double den = dr * dr + di * di;
        fr = dr / den;
        fi = -di / den;
    }
    out[2 * t] = s * n2[ch] * fr;
    out[2 * t + 1] = s * n2[ch] * fi;
}

// Inputs staged at d (2 nch m + 3 nch + nw doubles), then the output (2 nch nw).
static size_t cf_stage_doubles(int nch, int m, int nw) {
    return 2 * (size_t)nch * m + 3 * (size_t)nch + (size_t)nw + 2 * (size_t)nch * nw;
}

static int cf_eval_staged(phub_ctx *c, double *d, int nch, int m, const double *alpha,
                          const double *beta, const double *norm2, const double *shift,
                          const double *sign, int nw, const double *omega, double eta,
                          double *out) {
    const size_t na = (size_t)nch * m;
    double *da = d, *db = da + na, *dn = db + na, *ds = dn + nch, *dg = ds + nch, *dw = dg + nch,
           *dout = dw + nw;
    const int64_t nt = (int64_t)nch * nw;
    if ((nt + 255) / 256 > (int64_t)INT32_MAX) return phub_fail(c, PHUB_ERR_UNSUPPORTED, "grid too large");
    PHUB_HIP(c, hipMemcpyAsync(da, alpha, na * sizeof(double), hipMemcpyHostToDevice, c->stream));
    PHUB_HIP(c, hipMemcpyAsync(db, beta, na * sizeof(double), hipMemcpyHostToDevice, c->stream));
    PHUB_HIP(c, hipMemcpyAsync(dn, norm2, nch * sizeof(double), hipMemcpyHostToDevice, c->stream));
    PHUB_HIP(c, hipMemcpyAsync(ds, shift, nch * sizeof(double), hipMemcpyHostToDevice, c->stream));
    PHUB_HIP(c, hipMemcpyAsync(dg, sign, nch * sizeof(double), hipMemcpyHostToDevice, c->stream));
    PHUB_HIP(c, hipMemcpyAsync(dw, omega, nw * sizeof(double), hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(k_cf_eval, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, c->stream, nch,
                       m, da, db, dn, ds, dg, nw, dw, eta, dout);
    PHUB_HIP(c, hipGetLastError());
    PHUB_HIP(c, hipMemcpyAsync(out, dout, 2 * (size_t)nt * sizeof(double), hipMemcpyDeviceToHost,
                               c->stream));
    PHUB_HIP(c, hipStreamSynchronize(c->stream));
    return PHUB_OK;
}

extern "C" int phub_cf_eval(phub_ctx *c, int nch, int m, const double *alpha, const double *beta,
                            const double *norm2, const double *shift, const double *sign,
                            int nw, const double *omega, double eta, double *out) {
    if (!c || nch < 0 || m < 1 || nw < 0) return phub_fail(c, PHUB_ERR_ARG, "bad argument");
    if (nch == 0 || nw == 0) return PHUB_OK;
    double *d = nullptr;
    int rc = phub_arena(c, cf_stage_doubles(nch, m, nw) * sizeof(double), (void **)&d);
    if (rc) return rc;
    return cf_eval_staged(c, d, nch, m, alpha, beta, norm2, shift, sign, nw, omega, eta, out);
}

// ---------------------------------------------------------------------------
// phub_cf_eval without any allocation, for callers that cannot allocate: the
// staged inputs and the output live in `scratch`, a file-scope device array of
// PHUB_CF_NOALLOC_DOUBLES doubles (32 MiB) that every GPU the library runs on
// holds until the process ends.  Host code never passes `scratch` itself (its name
// is a host-side stand-in, not a device address): the first call on a device,
// with that device current, looks up the array's address there with
// hipGetSymbolAddress and keeps it per device.  All calls on one GPU share the
// array, so calls that overlap on one GPU (contexts on different streams) race;
// calls on one stream are ordered.  Nothing in it needs clearing: each call
// writes every staged input and the kernel writes every output element.
// ---------------------------------------------------------------------------
static __device__ __align__(256) double scratch[PHUB_CF_NOALLOC_DOUBLES];

static std::mutex g_scratch_mu;
static double *g_scratch_ptr[PHUB_MAX_DEVICES];

static int scratch_for_device(phub_ctx *c, double **out) {
    if (c->device < 0 || c->device >= PHUB_MAX_DEVICES)
        return phub_fail(c, PHUB_ERR_UNSUPPORTED, "device index beyond PHUB_MAX_DEVICES");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[c->device]) {
        PHUB_HIP(c, hipSetDevice(c->device));
        void *p = nullptr;
        PHUB_HIP(c, hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return phub_fail(c, PHUB_ERR_HIP, "hipGetSymbolAddress returned no address");
        g_scratch_ptr[c->device] = (double *)p;
    }
    *out = g_scratch_ptr[c->device];
    return PHUB_OK;
}

extern "C" int phub_cf_eval_noalloc(phub_ctx *c, int nch, int m, const double *alpha,
                                    const double *beta, const double *norm2, const double *shift,
                                    const double *sign, int nw, const double *omega, double eta,
                                    double *out) {
    if (!c || nch < 0 || m < 1 || nw < 0) return phub_fail(c, PHUB_ERR_ARG, "bad argument");
    if (nch == 0 || nw == 0) return PHUB_OK;
    const size_t need = cf_stage_doubles(nch, m, nw);
    if (need > (size_t)PHUB_CF_NOALLOC_DOUBLES)
        return phub_fail(c, PHUB_ERR_UNSUPPORTED,
                         "phub_cf_eval_noalloc: 2*nch*m + 3*nch + nw + 2*nch*nw = " +
                             std::to_string(need) + " doubles exceeds the fixed scratch of " +
                             std::to_string((size_t)PHUB_CF_NOALLOC_DOUBLES) +
                             " doubles; use phub_cf_eval");
    double *d = nullptr;
    int rc = scratch_for_device(c, &d);
    if (rc) return rc;
    return cf_eval_staged(c, d, nch, m, alpha, beta, norm2, shift, sign, nw, omega, eta, out);
}
