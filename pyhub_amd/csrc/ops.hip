// ops.hip -- operator objects: the device-side state of one model + sector.
// Replaces the HUBMOD globals the reference fills from solver.h5 before every
// executable run (pyhub/libs/src/hubmod.f90:1-29, compute_solver.f90:39-179).
#include "phub_internal.h"

#include <cmath>
#include <cstring>

static bool symmetric(const double *a, int n) {
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < i; ++j)
            if (!(std::fabs(a[i * n + j] - a[j * n + i]) <= 1e-8 + 1e-5 * std::fabs(a[j * n + i])))
                return false;
    return true;
}

extern "C" int phub_op_hubbard(phub_ctx *c, const phub_hubbard_desc *d, phub_op **out) {
    if (!c || !d || !out || !d->t || (!d->u && !d->u4)) return phub_fail(c, PHUB_ERR_ARG, "null argument");
    const int n = d->n_sites;
    if (n < 1 || n > 32) return phub_fail(c, PHUB_ERR_UNSUPPORTED, "Hubbard operator needs 1 <= n_sites <= 32");
    if (d->n_up < 0 || d->n_up > n || d->n_dn < 0 || d->n_dn > n)
        return phub_fail(c, PHUB_ERR_ARG, "particle numbers out of range");
    for (int i = 0; i < n * n; ++i)
        if (!std::isfinite(d->t[i])) return phub_fail(c, PHUB_ERR_ARG, "t_matrix has non-finite entries");
    const int n4 = n * n * n * n;
    if (d->u4) {
        // core_solver.py:142-145: np.isclose(U, einsum('ijkl->jilk', U)) everywhere
        for (int i = 0; i < n4; ++i)
            if (!std::isfinite(d->u4[i])) return phub_fail(c, PHUB_ERR_ARG, "U has non-finite entries");
        for (int J = 0; J < n; ++J)
            for (int K = 0; K < n; ++K)
                for (int M = 0; M < n; ++M)
                    for (int N = 0; N < n; ++N) {
                        const double a = d->u4[((J * n + K) * n + M) * n + N];
                        const double b = d->u4[((K * n + J) * n + N) * n + M];
                        if (!(std::fabs(a - b) <= 1e-8 + 1e-5 * std::fabs(b)))
                            return phub_fail(c, PHUB_ERR_ARG, "U tensor must be symmetric");
                    }
    } else {
        for (int i = 0; i < n; ++i)
            if (!std::isfinite(d->u[i])) return phub_fail(c, PHUB_ERR_ARG, "U has non-finite entries");
    }
    // fermi_hubbard.py:26-27
    if (!symmetric(d->t, n)) return phub_fail(c, PHUB_ERR_ARG, "t_matrix must be symmetric");
    const double *tdn = d->t_dn ? d->t_dn : d->t;
    for (int i = 0; i < n * n; ++i)
        if (!std::isfinite(tdn[i])) return phub_fail(c, PHUB_ERR_ARG, "t_dn has non-finite entries");
    if (!symmetric(tdn, n)) return phub_fail(c, PHUB_ERR_ARG, "t_dn must be symmetric");
    PHUB_HIP(c, hipSetDevice(c->device));
    phub_op *op = new phub_op();
    op->kind = 0;
    op->ctx = c;
    op->n = n;
    op->nup = d->n_up;
    op->ndn = d->n_dn;
    op->T.assign(d->t, d->t + n * n);
    op->u_uniform = 1;
    op->u0 = d->u4 ? 0.0 : d->u[0];
    for (int i = 0; i < PHUB_MAXN; ++i) op->U[i] = (i < n && !d->u4) ? d->u[i] : 0.0;
    for (int i = 1; i < n && !d->u4; ++i)
        if (d->u[i] != d->u[0]) op->u_uniform = 0;
    int rc = phub_build_hop_table(c, n, d->n_up, d->t, true, false, &op->up);
    if (!rc) rc = phub_build_hop_table(c, n, d->n_dn, tdn, false, false, &op->dn);
    if (rc) { phub_op_destroy(op); return rc; }
    op->dim_up = op->up.dim;
    op->dim_dn = op->dn.dim;
    op->dim = op->dim_up * op->dim_dn;
    auto tb = [](const HopTable &t) {
        return (double)t.dim * (4.0 + 8.0 + 4.0) + (double)t.dim * t.maxh * (4.0 + 8.0 + 4.0);
    };
    op->table_bytes = tb(op->up) + tb(op->dn);
    rc = phub_build_dn_ent(c, op);
    if (rc) { phub_op_destroy(op); return rc; }
    if (d->u4) {
        // UPARTIAL(J,K) = sum_{MN} UL(J,K,M,N)^2; a term enters when UPARTIAL(J,K) >= 1e-14
        // and |UL(J,K,M,N)| >= 1e-14 (hprod_hubbard.f90:62-65): keep exactly those
        std::vector<double> ul(d->u4, d->u4 + n4);
        for (int jk = 0; jk < n * n; ++jk) {
            double up = 0.0;
            for (int mn = 0; mn < n * n; ++mn) up += ul[(size_t)jk * n * n + mn] * ul[(size_t)jk * n * n + mn];
            for (int mn = 0; mn < n * n; ++mn) {
                double &x = ul[(size_t)jk * n * n + mn];
                if (!(up >= 1e-14) || !(std::fabs(x) >= 1e-14)) x = 0.0;
            }
        }
        hipError_t e = phub_malloc(&op->d_ul, n4 * sizeof(double));
        if (e == hipSuccess) e = phub_copy(c, op->d_ul, ul.data(), n4 * sizeof(double), hipMemcpyHostToDevice);
        if (e != hipSuccess) { phub_op_destroy(op); return phub_hip_fail(c, e, "4-index U upload"); }
        rc = phub_ensure_pair_tables(c, op);
        if (rc) { phub_op_destroy(op); return rc; }
        op->table_bytes += n4 * 8.0;
    }
    *out = op;
    return PHUB_OK;
}

extern "C" int phub_op_heisenberg(phub_ctx *c, const phub_heis_desc *d, phub_op **out) {
    if (!c || !d || !out || !d->j) return phub_fail(c, PHUB_ERR_ARG, "null argument");
    const int n = d->n_sites;
    if (n < 2 || n > 40) return phub_fail(c, PHUB_ERR_UNSUPPORTED, "Heisenberg operator needs 2 <= n_sites <= 40");
    for (int i = 0; i < n * n; ++i)
        if (!std::isfinite(d->j[i])) return phub_fail(c, PHUB_ERR_ARG, "J_matrix has non-finite entries");
    // heisenberg.py:26-27
    if (!symmetric(d->j, n)) return phub_fail(c, PHUB_ERR_ARG, "J_matrix must be symmetric");
    PHUB_HIP(c, hipSetDevice(c->device));
    phub_op *op = new phub_op();
    op->kind = 1;
    op->ctx = c;
    op->n = n;
    op->nup = n / 2;
    op->ndn = 0;
    op->J.assign(d->j, d->j + n * n);
    op->sign_convention = d->sign_convention;
    int rc = phub_build_heis(c, n, d->j, d->sign_convention, &op->h);
    if (rc) { phub_op_destroy(op); return rc; }
    op->dim_up = op->h.dim;
    op->dim_dn = 1;
    op->dim = op->h.dim;
    op->table_bytes = (op->h.masks ? (double)op->h.dim * 8.0 : 0.0) + op->h.stg.bytes;
    *out = op;
    return PHUB_OK;
}

extern "C" void phub_op_destroy(phub_op *op) {
    if (!op) return;
    phub_free_hop_table(&op->up);
    phub_free_hop_table(&op->dn);
    phub_free_hop_table(&op->rup);
    phub_free_hop_table(&op->rdn);
    if (op->dn_pairmap) hipFree(op->dn_pairmap);
    for (void *p : op->dn_ent)
        if (p) hipFree(p);

    phub_free_heis(&op->h);
    if (op->d_ul) hipFree(op->d_ul);
    if (op->heis_all) phub_op_destroy(op->heis_all);
    delete op;
}

extern "C" int phub_op_get_info(const phub_op *op, phub_op_info *info) {
    if (!op || !info) return PHUB_ERR_ARG;
    info->kind = op->kind;
    info->n_sites = op->n;
    info->n_up = op->nup;
    info->n_dn = op->ndn;
    info->dim_up = op->dim_up;
    info->dim_dn = op->dim_dn;
    info->dim = op->dim;
    info->max_hops_up = op->up.maxh;
    info->max_hops_dn = op->dn.maxh;
    info->table_bytes = op->table_bytes;
    for (int t = 0; t < 2; ++t) info->dn_tile[t] = phub_dn_tile_code(op->tune_dn[t], t == 1);
    return PHUB_OK;
}

__global__ void k_widen(const uint32_t *__restrict__ a, int64_t n, uint64_t *__restrict__ b) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = a[i];
}

extern "C" int phub_op_masks(phub_ctx *c, const phub_op *op, uint64_t *d_up, uint64_t *d_dn) {
    if (!c || !op) return phub_fail(c, PHUB_ERR_ARG, "null argument");
    if (op->kind == 0) {
        if (d_up && op->dim_up)
            hipLaunchKernelGGL(k_widen, dim3((op->dim_up + 255) / 256), dim3(256), 0, c->stream,
                               op->up.masks, op->dim_up, d_up);
        if (d_dn && op->dim_dn)
            hipLaunchKernelGGL(k_widen, dim3((op->dim_dn + 255) / 256), dim3(256), 0, c->stream,
                               op->dn.masks, op->dim_dn, d_dn);
    } else {
        if (d_up && op->dim)
            PHUB_HIP(c, hipMemcpyAsync(d_up, op->h.masks, op->dim * sizeof(uint64_t),
                                       hipMemcpyDeviceToDevice, c->stream));
    }
    PHUB_HIP(c, hipGetLastError());
    return PHUB_OK;
}
