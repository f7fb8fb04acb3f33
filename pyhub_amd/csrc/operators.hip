// operators.hip -- generic operator strings on a product basis of fixed-N sectors.
// Replaces operators.x: the driver compute_operators.f90:1-72 and the elementary
// routines of operators.f90 (C_DAGGER_C, C_DAGGER_C_UNRES, NI, NIM1, SZ_OPE, SISJ,
// C_DAGGER, C_), which apply the factors of a product one after the other to the
// whole vector and sum the products with their coefficients.
//
// Here one kernel evaluates every product for a target state at once, in pull
// form: starting from the target's masks it walks the factors backwards through
// their inverse maps (each factor maps a basis state to at most one basis state),
// collects the forward sign of each factor, ranks the source state and adds
// coeff * sign * in[source].  No scatter, no atomics: every output element is
// written by one thread, in a fixed order.
#include "phub_internal.h"

#include <cstring>
#include <vector>

struct phub_fock {
    int n = 0;
    int64_t nsup = 0, nsdn = 0;
    uint64_t *bup = nullptr, *bdn = nullptr;   // device, concatenated sectors
    int64_t off_up[PHUB_MAXN + 1], off_dn[PHUB_MAXN + 1];   // -1: sector absent
    int64_t *d_off = nullptr;                  // device copy: [2][n+1]
};

// One expanded program per evaluation path: the reference evaluates a product on a
// vector with operators.f90 and on a single basis element (to_matrix, the
// comparisons, bit_repr) with operators_basis.f90, and the two differ for SZ and
// S.S (see phub_opprog_create).
struct OpProgDev {
    int nterms = 0;
    double *coef = nullptr;      // device, nterms
    int *start = nullptr;        // device, nterms + 1
    int *elem = nullptr;         // device, 5 per factor: code, site1, spin1, site2, spin2
};

struct phub_opprog {
    OpProgDev vec, bas;
    int max_site = 0;            // checked against the basis at every apply
};

static int fock_fail_free(phub_ctx *c, phub_fock *f, hipError_t e, const char *what) {
    phub_fock_destroy(f);
    return phub_hip_fail(c, e, what);
}

extern "C" int phub_fock_create(phub_ctx *c, int n, const int *n_up, int n_up_count,
                                const int *n_dn, int n_dn_count, phub_fock **out) {
    if (!c || !out || (!n_up && n_up_count) || (!n_dn && n_dn_count))
        return phub_fail(c, PHUB_ERR_ARG, "null argument");
    if (n < 1 || n > 32) return phub_fail(c, PHUB_ERR_UNSUPPORTED, "operators need 1 <= n_sites <= 32");
    PHUB_HIP(c, hipSetDevice(c->device));
    phub_fock *f = new phub_fock();
    f->n = n;
    for (int i = 0; i <= PHUB_MAXN; ++i) f->off_up[i] = f->off_dn[i] = -1;
    // the species list is the concatenation of its sectors in the given order
    // (compute_operators.f90:10-26); a repeated N is kept at its first position
    for (int s = 0; s < 2; ++s) {
        const int *ks = s == 0 ? n_up : n_dn;
        const int cnt = s == 0 ? n_up_count : n_dn_count;
        int64_t *off = s == 0 ? f->off_up : f->off_dn;
        int64_t total = 0;
        for (int i = 0; i < cnt; ++i) {
            const int k = ks[i];
            if (k < 0 || k > n) {
                phub_fock_destroy(f);
                return phub_fail(c, PHUB_ERR_ARG, "particle number out of range");
            }
            if (off[k] >= 0) continue;
            off[k] = total;
            total += (int64_t)phub_binom(n, k);
        }
        (s == 0 ? f->nsup : f->nsdn) = total;
        uint64_t *&dst = s == 0 ? f->bup : f->bdn;
        hipError_t e = hipMalloc(&dst, std::max<int64_t>(total, 1) * sizeof(uint64_t));
        if (e != hipSuccess) return fock_fail_free(c, f, e, "species list");
        for (int k = 0; k <= n; ++k)
            if (off[k] >= 0) {
                int rc = phub_basis_build(c, n, k, dst + off[k]);
                if (rc) { phub_fock_destroy(f); return rc; }
            }
    }
    hipError_t e = hipMalloc(&f->d_off, 2 * (PHUB_MAXN + 1) * sizeof(int64_t));
    if (e == hipSuccess)
        e = hipMemcpy(f->d_off, f->off_up, (PHUB_MAXN + 1) * sizeof(int64_t), hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(f->d_off + PHUB_MAXN + 1, f->off_dn, (PHUB_MAXN + 1) * sizeof(int64_t),
                      hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return fock_fail_free(c, f, e, "sector offsets");
    *out = f;
    return PHUB_OK;
}

extern "C" void phub_fock_destroy(phub_fock *f) {
    if (!f) return;
    if (f->bup) hipFree(f->bup);
    if (f->bdn) hipFree(f->bdn);
    if (f->d_off) hipFree(f->d_off);
    delete f;
}

extern "C" int phub_fock_dims(const phub_fock *f, int64_t *nsup, int64_t *nsdn) {
    if (!f) return PHUB_ERR_ARG;
    if (nsup) *nsup = f->nsup;
    if (nsdn) *nsdn = f->nsdn;
    return PHUB_OK;
}

extern "C" int phub_fock_masks(phub_ctx *c, const phub_fock *f, uint64_t *d_up, uint64_t *d_dn) {
    if (!c || !f) return phub_fail(c, PHUB_ERR_ARG, "null argument");
    if (d_up && f->nsup)
        PHUB_HIP(c, hipMemcpyAsync(d_up, f->bup, f->nsup * sizeof(uint64_t), hipMemcpyDeviceToDevice, c->stream));
    if (d_dn && f->nsdn)
        PHUB_HIP(c, hipMemcpyAsync(d_dn, f->bdn, f->nsdn * sizeof(uint64_t), hipMemcpyDeviceToDevice, c->stream));
    return PHUB_OK;
}

// internal factor codes beyond the reference's OPE codes (compute_operators.f90:36-63):
// the host expands SISJ (8) into two pair moves and one of these diagonals
enum { OP_SZSZ = 12, OP_SISJ_SAME = 13, OP_SISJ_SAME_HALF = 14, OP_SZ_PROPER = 15 };

static void free_prog(OpProgDev &d) {
    if (d.coef) hipFree(d.coef);
    if (d.start) hipFree(d.start);
    if (d.elem) hipFree(d.elem);
    d = OpProgDev();
}

struct Fac { int code, s1, p1, s2, p2; };

// Distribute the products over S.S and map the reference's codes to the factors the
// kernel evaluates.  basis = false: the vector path (operators.f90); true: the
// basis-element path (operators_basis.f90).  They differ in:
//   SZ (5):   SZ_OPE gives (1 - n_up)/2 - (1 - n_dn)/2 = -Sz (operators.f90:224-245),
//             SZ_BASIS the proper Sz (operators_basis.f90:134-143);
//   S.S (8):  SISJ(K, J) = -1/2 [up J->K with down K->J] - 1/2 [down J->K with up K->J]
//             + SZ(J) SZ(K) (operators.f90:303-394); SISJ_BASIS drops the SZ SZ term
//             (operators_basis.f90:147-186).  For J = K the vector path reads the
//             K-th up state instead of the current one (operators.f90:311-314); here
//             both paths use the current state: 1/2 + 1/4 = 3/4 on a singly occupied
//             site (vector), 1/2 (basis element).
static int expand(phub_ctx *c, int nterms, const double *coef, const int *term_len,
                  const int *elems, bool basis, std::vector<double> &cf, std::vector<int> &st,
                  std::vector<int> &el, int &max_site) {
    st.assign(1, 0);
    int pos = 0;
    for (int t = 0; t < nterms; ++t) {
        const int len = term_len[t];
        if (len < 0) return phub_fail(c, PHUB_ERR_ARG, "negative term length");
        std::vector<std::vector<Fac>> alts(1);
        std::vector<double> wts(1, coef[t]);
        for (int q = 0; q < len; ++q) {
            const int *e = elems + 5 * (size_t)(pos + q);
            Fac f{e[0], e[1], e[2], e[3], e[4]};
            if (f.code < 1 || f.code > 11) return phub_fail(c, PHUB_ERR_ARG, "unknown operator code");
            const bool one = f.code == 2 || f.code == 4 || f.code == 5 || f.code == 6 ||
                             f.code == 7 || f.code == 10 || f.code == 11;
            const bool two = f.code == 1 || f.code == 8 || f.code == 9;
            if ((one || two) && (f.s1 < 1 || f.s1 > 32)) return phub_fail(c, PHUB_ERR_ARG, "site out of range");
            if (two && (f.s2 < 1 || f.s2 > 32)) return phub_fail(c, PHUB_ERR_ARG, "site out of range");
            if (one || two) max_site = std::max(max_site, f.s1);
            if (two) max_site = std::max(max_site, f.s2);
            if (f.code == 5 && basis) f.code = OP_SZ_PROPER;
            if (f.code != 8) {
                for (auto &a : alts) a.push_back(f);
                continue;
            }
            const int K = f.s1, J = f.s2;
            std::vector<std::vector<Fac>> na;
            std::vector<double> nw;
            for (size_t a = 0; a < alts.size(); ++a) {
                if (J != K) {
                    auto x = alts[a];
                    x.push_back({1, K, 1, J, 1});   // up: c^dag_K c_J
                    x.push_back({1, J, 2, K, 2});   // down: c^dag_J c_K
                    na.push_back(x);
                    nw.push_back(-0.5 * wts[a]);
                    auto y = alts[a];
                    y.push_back({1, K, 2, J, 2});   // down: c^dag_K c_J
                    y.push_back({1, J, 1, K, 1});   // up: c^dag_J c_K
                    na.push_back(y);
                    nw.push_back(-0.5 * wts[a]);
                    if (!basis) {
                        auto z = alts[a];
                        z.push_back({OP_SZSZ, J, 0, K, 0});
                        na.push_back(z);
                        nw.push_back(wts[a]);
                    }
                } else {
                    auto z = alts[a];
                    z.push_back({basis ? OP_SISJ_SAME_HALF : OP_SISJ_SAME, J, 0, J, 0});
                    na.push_back(z);
                    nw.push_back(wts[a]);
                }
            }
            alts.swap(na);
            wts.swap(nw);
        }
        pos += len;
        // the reference skips |COEFF| <= 1e-14 (compute_operators.f90:33)
        if (!(std::fabs(coef[t]) > 1e-14)) continue;
        for (size_t a = 0; a < alts.size(); ++a) {
            cf.push_back(wts[a]);
            for (const Fac &f : alts[a]) {
                el.push_back(f.code);
                el.push_back(f.s1);
                el.push_back(f.p1);
                el.push_back(f.s2);
                el.push_back(f.p2);
            }
            st.push_back((int)(el.size() / 5));
        }
    }
    return PHUB_OK;
}

static hipError_t upload_prog(OpProgDev &d, const std::vector<double> &cf,
                              const std::vector<int> &st, const std::vector<int> &el) {
    d.nterms = (int)cf.size();
    hipError_t e = hipMalloc(&d.coef, std::max<size_t>(cf.size(), 1) * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&d.start, st.size() * sizeof(int));
    if (e == hipSuccess) e = hipMalloc(&d.elem, std::max<size_t>(el.size(), 1) * sizeof(int));
    if (e == hipSuccess && !cf.empty())
        e = hipMemcpy(d.coef, cf.data(), cf.size() * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d.start, st.data(), st.size() * sizeof(int), hipMemcpyHostToDevice);
    if (e == hipSuccess && !el.empty())
        e = hipMemcpy(d.elem, el.data(), el.size() * sizeof(int), hipMemcpyHostToDevice);
    return e;
}

extern "C" int phub_opprog_create(phub_ctx *c, int nterms, const double *coef,
                                  const int *term_len, const int *elems, phub_opprog **out) {
    if (!c || !out || nterms < 0 || (nterms > 0 && (!coef || !term_len)))
        return phub_fail(c, PHUB_ERR_ARG, "null argument");
    PHUB_HIP(c, hipSetDevice(c->device));
    phub_opprog *p = new phub_opprog();
    for (int b = 0; b < 2; ++b) {
        std::vector<double> cf;
        std::vector<int> st, el;
        int rc = expand(c, nterms, coef, term_len, elems, b == 1, cf, st, el, p->max_site);
        if (rc) { phub_opprog_destroy(p); return rc; }
        hipError_t e = upload_prog(b == 0 ? p->vec : p->bas, cf, st, el);
        if (e != hipSuccess) {
            phub_opprog_destroy(p);
            return phub_hip_fail(c, e, "operator program upload");
        }
    }
    *out = p;
    return PHUB_OK;
}

extern "C" void phub_opprog_destroy(phub_opprog *p) {
    if (!p) return;
    free_prog(p->vec);
    free_prog(p->bas);
    delete p;
}

// products after the expansion of S.S (vector path)
extern "C" int phub_opprog_terms(const phub_opprog *p) { return p ? p->vec.nterms : 0; }

// (-1)^(number of occupied sites with index < J) in m: ANTICOM (funcmod.f90:29-41);
// site J (1-based) is bit n - J, so the sites before J are the bits above it
__device__ __forceinline__ double anticom(int n, int J, uint64_t m) {
    if (J <= 1) return 1.0;
    const uint64_t above = (~0ull << (n - J + 1)) & ((n >= 64) ? ~0ull : ((1ull << n) - 1));
    return (__popcll(m & above) & 1) ? -1.0 : 1.0;
}

// Inverse of one factor: (xu, xd) is the state after it; on success the state
// before it replaces (xu, xd) and the factor's forward sign / weight multiplies w.
__device__ bool undo_factor(const int *e, int n, uint64_t &xu, uint64_t &xd, double &w) {
    const int code = e[0], s1 = e[1], p1 = e[2], s2 = e[3], p2 = e[4];
    const uint64_t b1 = 1ull << (n - s1);
    switch (code) {
    case 3:   // identity
        return true;
    case 1: {   // c^dag_{s1} c_{s2} on spin p1 (C_DAGGER_C, operators.f90:1-55)
        uint64_t &m = p1 == 1 ? xu : xd;
        const uint64_t b2 = 1ull << (n - s2);
        if (s1 == s2) return (m & b1) != 0;   // n_J: KINOP keeps an occupied state
        if (!(m & b1) || (m & b2)) return false;
        const uint64_t after = m;
        m = m - b1 + b2;
        w *= anticom(n, s2, m) * anticom(n, s1, after);
        return true;
    }
    case 2:   // NI: keeps the states with site s1 occupied (operators.f90:148-172)
        return ((p1 == 1 ? xu : xd) & b1) != 0;
    case 4:   // NIM1: keeps the empty ones (operators.f90:196-220)
        return ((p1 == 1 ? xu : xd) & b1) == 0;
    case 5: {   // SZ_OPE (operators.f90:224-245): (1 - n_up)/2 - (1 - n_dn)/2
        const double v = ((xu & b1) ? 0.0 : 0.5) - ((xd & b1) ? 0.0 : 0.5);
        w *= v;
        return v != 0.0;
    }
    case 6:   // S+ : C_DAGGER_C_UNRES(J, J, 2), a down removed and an up added at J
    case 7: {   // S- : C_DAGGER_C_UNRES(J, J, 1), an up removed and a down added at J
        uint64_t &add = code == 6 ? xu : xd, &rem = code == 6 ? xd : xu;
        if (!(add & b1) || (rem & b1)) return false;
        w *= anticom(n, s1, rem) * anticom(n, s1, add);   // ANTICOM(K, NEW) * ANTICOM(J, NEW2)
        add -= b1;
        rem += b1;
        return true;
    }
    case 9: {   // c^dag_{s1, other spin} c_{s2, p2} (C_DAGGER_C_UNRES, operators.f90:58-144)
        const uint64_t b2 = 1ull << (n - s2);
        uint64_t &rem = p2 == 1 ? xu : xd, &add = p2 == 1 ? xd : xu;
        if ((rem & b2) || !(add & b1)) return false;
        w *= anticom(n, s2, rem) * anticom(n, s1, add);
        add -= b1;
        rem += b2;
        return true;
    }
    case 10: {   // C_DAGGER (operators.f90:397-446): sign from the state before
        uint64_t &m = p1 == 1 ? xu : xd;
        if (!(m & b1)) return false;
        m -= b1;
        w *= anticom(n, s1, m);
        return true;
    }
    case 11: {   // C_ (operators.f90:450-499)
        uint64_t &m = p1 == 1 ? xu : xd;
        if (m & b1) return false;
        m += b1;
        w *= anticom(n, s1, m);
        return true;
    }
    case OP_SZSZ: {   // SZ(J) * SZ(K) with funcmod's SZ (+1/2 up, -1/2 down)
        const uint64_t b2 = 1ull << (n - s2);
        const double z1 = ((xu & b1) ? 0.5 : 0.0) - ((xd & b1) ? 0.5 : 0.0);
        const double z2 = ((xu & b2) ? 0.5 : 0.0) - ((xd & b2) ? 0.5 : 0.0);
        w *= z1 * z2;
        return z1 * z2 != 0.0;
    }
    case OP_SISJ_SAME:        // S_J . S_J on the vector path: 3/4 on a singly occupied site
    case OP_SISJ_SAME_HALF: {   // SISJ_BASIS with SI = SJ: 1/2 there
        const bool u = (xu & b1) != 0, d = (xd & b1) != 0;
        const double v = (u != d) ? (code == OP_SISJ_SAME ? 0.75 : 0.5) : 0.0;
        w *= v;
        return v != 0.0;
    }
    case OP_SZ_PROPER: {   // SZ_BASIS: funcmod's SZ, +1/2 up, -1/2 down
        const double v = ((xu & b1) ? 0.5 : 0.0) - ((xd & b1) ? 0.5 : 0.0);
        w *= v;
        return v != 0.0;
    }
    default:
        return false;
    }
}

__device__ __forceinline__ int64_t fock_rank(const uint64_t *binom, const int64_t *off, uint64_t m,
                                             int n) {
    const int k = __popcll(m);
    const int64_t o = off[k];
    if (o < 0) return -1;
    return o + rank_ref_dev(binom, m, n, k);
}

struct OpArgs {
    int n;
    int64_t nsup, nsdn;
    const uint64_t *bup, *bdn;
    const int64_t *off;        // [2][PHUB_MAXN + 1]
    const uint64_t *binom;
    int nterms;
    const double *coef;
    const int *start, *elem;
};

// mode 0: out[v][t] = sum_terms coef * sign * in[v][source(t)]    (operator applied)
// mode 1: out[t]    = sum_terms coef * sign * [source(t) == t]    (diagonal, basis-element mode)
// mode 2: out[source(t) * ns + t] += coef * sign                  (matrix, row i = O e_i)
template <int MODE>
__global__ __launch_bounds__(256) void k_op_apply(OpArgs a, const double *__restrict__ in,
                                                  double *__restrict__ out, int nvec) {
    const int64_t ns = a.nsup * a.nsdn;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ns) return;
    const int64_t td = t / a.nsup, tu = t - td * a.nsup;
    const uint64_t mu0 = a.bup[tu], md0 = a.bdn[td];
    double acc0 = 0.0;
    for (int q = 0; q < a.nterms; ++q) {
        uint64_t xu = mu0, xd = md0;
        double w = a.coef[q];
        bool ok = true;
        // every intermediate state must lie in the basis too: the reference applies
        // the factors one by one to vectors of this basis, so a path through an
        // absent sector is cut (its POS lookup fails)
        for (int f = a.start[q + 1] - 1; f >= a.start[q] && ok; --f) {
            ok = undo_factor(a.elem + 5 * (size_t)f, a.n, xu, xd, w);
            ok = ok && a.off[__popcll(xu)] >= 0 && a.off[PHUB_MAXN + 1 + __popcll(xd)] >= 0;
        }
        if (!ok) continue;
        const int64_t su = fock_rank(a.binom, a.off, xu, a.n);
        const int64_t sd = fock_rank(a.binom, a.off + PHUB_MAXN + 1, xd, a.n);
        if (su < 0 || sd < 0) continue;   // the source lies outside the listed sectors
        const int64_t s = sd * a.nsup + su;
        if (MODE == 0) {
            for (int v = 0; v < nvec; ++v) {
                if (v == 0) acc0 += w * in[s];
                else out[(size_t)v * ns + t] += w * in[(size_t)v * ns + s];
            }
        } else if (MODE == 1) {
            if (s == t) acc0 += w;
        } else {
            out[(size_t)s * ns + t] += w;
        }
    }
    if (MODE == 0 || MODE == 1) out[t] = acc0;
}

static OpArgs op_args(phub_ctx *c, const phub_fock *f, const OpProgDev &p) {
    OpArgs a{};
    a.n = f->n;
    a.nsup = f->nsup;
    a.nsdn = f->nsdn;
    a.bup = f->bup;
    a.bdn = f->bdn;
    a.off = f->d_off;
    a.binom = c->d_binom;
    a.nterms = p.nterms;
    a.coef = p.coef;
    a.start = p.start;
    a.elem = p.elem;
    return a;
}

extern "C" int phub_operator_apply(phub_ctx *c, const phub_fock *f, const phub_opprog *p,
                                   const double *d_in, double *d_out, int nvec) {
    if (!c || !f || !p || !d_in || !d_out) return phub_fail(c, PHUB_ERR_ARG, "null argument");
    if (nvec < 1) return phub_fail(c, PHUB_ERR_ARG, "nvec must be >= 1");
    if (d_in == d_out) return phub_fail(c, PHUB_ERR_ARG, "out must not alias in");
    if (p->max_site > f->n) return phub_fail(c, PHUB_ERR_ARG, "operator site beyond the basis");
    PHUB_HIP(c, hipSetDevice(c->device));
    const int64_t ns = f->nsup * f->nsdn;
    if (ns == 0) return PHUB_OK;
    // vectors past the first accumulate: clear them (the first is written once)
    if (nvec > 1)
        PHUB_HIP(c, hipMemsetAsync(d_out + ns, 0, (size_t)(nvec - 1) * ns * sizeof(double), c->stream));
    hipLaunchKernelGGL(k_op_apply<0>, dim3((unsigned)((ns + 255) / 256)), dim3(256), 0, c->stream,
                       op_args(c, f, p->vec), d_in, d_out, nvec);
    PHUB_HIP(c, hipGetLastError());
    return PHUB_OK;
}

extern "C" int phub_operator_diag(phub_ctx *c, const phub_fock *f, const phub_opprog *p,
                                  double *d_diag) {
    if (!c || !f || !p || !d_diag) return phub_fail(c, PHUB_ERR_ARG, "null argument");
    if (p->max_site > f->n) return phub_fail(c, PHUB_ERR_ARG, "operator site beyond the basis");
    PHUB_HIP(c, hipSetDevice(c->device));
    const int64_t ns = f->nsup * f->nsdn;
    if (ns == 0) return PHUB_OK;
    hipLaunchKernelGGL(k_op_apply<1>, dim3((unsigned)((ns + 255) / 256)), dim3(256), 0, c->stream,
                       op_args(c, f, p->bas), (const double *)nullptr, d_diag, 1);
    PHUB_HIP(c, hipGetLastError());
    return PHUB_OK;
}

extern "C" int phub_operator_matrix(phub_ctx *c, const phub_fock *f, const phub_opprog *p,
                                    double *d_mat) {
    if (!c || !f || !p || !d_mat) return phub_fail(c, PHUB_ERR_ARG, "null argument");
    if (p->max_site > f->n) return phub_fail(c, PHUB_ERR_ARG, "operator site beyond the basis");
    PHUB_HIP(c, hipSetDevice(c->device));
    const int64_t ns = f->nsup * f->nsdn;
    if (ns == 0) return PHUB_OK;
    if ((double)ns * ns * 8.0 > 16e9)
        return phub_fail(c, PHUB_ERR_UNSUPPORTED, "dense operator matrix above 16 GB");
    PHUB_HIP(c, hipMemsetAsync(d_mat, 0, (size_t)ns * ns * sizeof(double), c->stream));
    hipLaunchKernelGGL(k_op_apply<2>, dim3((unsigned)((ns + 255) / 256)), dim3(256), 0, c->stream,
                       op_args(c, f, p->bas), (const double *)nullptr, d_mat, 1);
    PHUB_HIP(c, hipGetLastError());
    return PHUB_OK;
}
