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

#include <atomic>
#include <cstdlib>
#include <cstring>
#include <vector>

// operators act on at most 32 sites: a species mask then fits 32 bits, which the
// ranking below reverses with one instruction
#define OP_MAXN 32
#define OP_NCHUNK_MAX (OP_MAXN / 4)
#define OP_NIB_MAX (OP_NCHUNK_MAX * (OP_MAXN + 1) * 16)

struct phub_fock {
    int n = 0;
    int64_t nsup = 0, nsdn = 0;
    uint64_t *bup = nullptr, *bdn = nullptr;   // device, concatenated sectors
    int64_t off_up[PHUB_MAXN + 1], off_dn[PHUB_MAXN + 1];   // -1: sector absent
    uint64_t pres_up = 0, pres_dn = 0;         // bit k set: sector k listed
    // device [2][OP_MAXN + 1]: off[k] + C(n, k) - 1, so that the position of a mask
    // in the list is base[popcount] minus its colex sum
    int64_t *d_base = nullptr;
    // device [nchunk][n + 1][16]: the colex sum of the sites 4c..4c+3 (nibble v of
    // the site-ordered mask) when s particles sit on the sites before them
    uint32_t *d_nib = nullptr;
    int nchunk = 0;
    uint64_t id = 0;                           // distinct per basis, keys the source maps
};

static std::atomic<uint64_t> g_fock_id{0};

// One expanded program per evaluation path: the reference evaluates a product on a
// vector with operators.f90 and on a single basis element (to_matrix, the
// comparisons, bit_repr) with operators_basis.f90, and the two differ for SZ and
// S.S (see phub_opprog_create).
struct OpProgDev {
    int nterms = 0;
    double *coef = nullptr;      // device, nterms
    int *start = nullptr;        // device, nterms + 1
    int *elem = nullptr;         // device, 5 per factor: code, site1, spin1, site2, spin2
    int *touch = nullptr;        // device, nterms: bit 0 / 1 set if the product may move
                                 // the up / down mask (else the source keeps the target's),
                                 // bit 2 if it changes a particle number, bit 3 if a
                                 // factor reads both species
    // products whose factors each act on one species, and that have factors of the
    // up (down) species, own a row of the up (down) source map: urow / drow per
    // product (-1: none), rowq_up / rowq_dn the product of each row
    int *urow = nullptr, *drow = nullptr, *rowq_up = nullptr, *rowq_dn = nullptr;
    int nrow_up = 0, nrow_dn = 0;
    bool has_mix = false;        // some product has a factor that reads both species
};

struct phub_opprog {
    OpProgDev vec, bas;
    int max_site = 0;            // checked against the basis at every apply
    // source maps of the vector program on the basis last applied to (built on the
    // first apply on a basis, kept until the program is destroyed)
    uint64_t map_fock = 0;       // id of that basis, 0 = none
    int32_t *d_map = nullptr;
    size_t map_cap = 0;          // bytes allocated
};

static int fock_fail_free(phub_ctx *c, phub_fock *f, hipError_t e, const char *what) {
    phub_fock_destroy(f);
    return phub_hip_fail(c, e, what);
}

extern "C" int phub_fock_create(phub_ctx *c, int n, const int *n_up, int n_up_count,
                                const int *n_dn, int n_dn_count, phub_fock **out) {
    if (!c || !out || (!n_up && n_up_count) || (!n_dn && n_dn_count))
        return phub_fail(c, PHUB_ERR_ARG, "null argument");
    if (n < 1 || n > OP_MAXN) return phub_fail(c, PHUB_ERR_UNSUPPORTED, "operators need 1 <= n_sites <= 32");
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
        hipError_t e = phub_malloc(&dst, std::max<int64_t>(total, 1) * sizeof(uint64_t));
        if (e != hipSuccess) return fock_fail_free(c, f, e, "species list");
        for (int k = 0; k <= n; ++k)
            if (off[k] >= 0) {
                int rc = phub_basis_build(c, n, k, dst + off[k]);
                if (rc) { phub_fock_destroy(f); return rc; }
            }
    }
    // ranking tables.  The list position of a k-particle mask m of a listed sector is
    // off[k] + C(n, k) - 1 - sum_i C(p_i, i), p_1 < p_2 < ... the 0-based sites of
    // its particles (site J is bit n - J): the closed form of HILBERT's order.  The
    // sum is taken four sites at a time from a table indexed by the particles
    // before the nibble and the nibble itself.
    auto C = [&](int a, int b) -> uint64_t {
        if (b < 0 || b > a) return 0;
        uint64_t r = 1;
        for (int i = 1; i <= b; ++i) r = r * (uint64_t)(a - b + i) / (uint64_t)i;
        return r;
    };
    f->nchunk = (n + 3) / 4;
    f->id = ++g_fock_id;
    std::vector<int64_t> base(2 * (OP_MAXN + 1), INT64_MIN);
    for (int k = 0; k <= n; ++k) {
        if (f->off_up[k] >= 0) { base[k] = f->off_up[k] + (int64_t)C(n, k) - 1; f->pres_up |= 1ull << k; }
        if (f->off_dn[k] >= 0) {
            base[OP_MAXN + 1 + k] = f->off_dn[k] + (int64_t)C(n, k) - 1;
            f->pres_dn |= 1ull << k;
        }
    }
    std::vector<uint32_t> nib((size_t)f->nchunk * (n + 1) * 16, 0);
    for (int ch = 0; ch < f->nchunk; ++ch)
        for (int s0 = 0; s0 <= n; ++s0)
            for (int v = 0; v < 16; ++v) {
                uint64_t sum = 0;
                int before = s0;
                for (int j = 0; j < 4; ++j)
                    if (v >> j & 1) {
                        ++before;
                        if (4 * ch + j < n) sum += C(4 * ch + j, before);
                    }
                nib[((size_t)ch * (n + 1) + s0) * 16 + v] = (uint32_t)sum;
            }
    hipError_t e = phub_malloc(&f->d_base, base.size() * sizeof(int64_t));
    if (e == hipSuccess) e = phub_malloc(&f->d_nib, nib.size() * sizeof(uint32_t));
    if (e == hipSuccess)
        e = phub_copy(c, f->d_base, base.data(), base.size() * sizeof(int64_t), hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = phub_copy(c, f->d_nib, nib.data(), nib.size() * sizeof(uint32_t), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return fock_fail_free(c, f, e, "ranking tables");
    *out = f;
    return PHUB_OK;
}

extern "C" void phub_fock_destroy(phub_fock *f) {
    if (!f) return;
    if (f->bup) hipFree(f->bup);
    if (f->bdn) hipFree(f->bdn);
    if (f->d_base) hipFree(f->d_base);
    if (f->d_nib) hipFree(f->d_nib);
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
    if (d.touch) hipFree(d.touch);
    for (int *q : {d.urow, d.drow, d.rowq_up, d.rowq_dn})
        if (q) hipFree(q);
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

// which species' masks a factor may change (bits 1 up, 2 down), whether it changes
// a particle number (bit 4) and whether it reads both species at once (bit 8)
static int factor_touch(const int *e) {
    switch (e[0]) {
    case 1: return e[1] == e[3] ? 0 : (e[2] == 1 ? 1 : 2);
    case 2: case 3: case 4: return 0;
    case 6: case 7: case 9: return 3 | 4 | 8;
    case 10: case 11: return (e[2] == 1 ? 1 : 2) | 4;
    default: return 8;   // SZ, SZ.SZ and the S.S diagonals
    }
}

static hipError_t upload_prog(phub_ctx *c, OpProgDev &d, const std::vector<double> &cf,
                              const std::vector<int> &st, const std::vector<int> &el) {
    d.nterms = (int)cf.size();
    std::vector<int> touch(std::max<size_t>(cf.size(), 1), 0);
    for (size_t q = 0; q < cf.size(); ++q)
        for (int f = st[q]; f < st[q + 1]; ++f) touch[q] |= factor_touch(&el[5 * (size_t)f]);
    std::vector<int> urow(touch.size(), -1), drow(touch.size(), -1), rq_up, rq_dn;
    for (size_t q = 0; q < cf.size(); ++q) {
        if (touch[q] & 8) { d.has_mix = true; continue; }
        bool up = false, dn = false;
        for (int f = st[q]; f < st[q + 1]; ++f) {
            const int *e = &el[5 * (size_t)f];
            if (e[0] == 3) continue;
            (e[2] == 1 ? up : dn) = true;
        }
        if (up) { urow[q] = (int)rq_up.size(); rq_up.push_back((int)q); }
        if (dn) { drow[q] = (int)rq_dn.size(); rq_dn.push_back((int)q); }
    }
    d.nrow_up = (int)rq_up.size();
    d.nrow_dn = (int)rq_dn.size();
    rq_up.push_back(0);
    rq_dn.push_back(0);
    hipError_t e = hipSuccess;
    for (auto pr : {std::make_pair(&d.urow, &urow), std::make_pair(&d.drow, &drow),
                    std::make_pair(&d.rowq_up, &rq_up), std::make_pair(&d.rowq_dn, &rq_dn)}) {
        if (e == hipSuccess) e = phub_malloc(pr.first, pr.second->size() * sizeof(int));
        if (e == hipSuccess)
            e = phub_copy(c, *pr.first, pr.second->data(), pr.second->size() * sizeof(int),
                          hipMemcpyHostToDevice);
    }
    if (e == hipSuccess) e = phub_malloc(&d.coef, std::max<size_t>(cf.size(), 1) * sizeof(double));
    if (e == hipSuccess) e = phub_malloc(&d.start, st.size() * sizeof(int));
    if (e == hipSuccess) e = phub_malloc(&d.elem, std::max<size_t>(el.size(), 1) * sizeof(int));
    if (e == hipSuccess) e = phub_malloc(&d.touch, touch.size() * sizeof(int));
    if (e == hipSuccess)
        e = phub_copy(c, d.touch, touch.data(), touch.size() * sizeof(int), hipMemcpyHostToDevice);
    if (e == hipSuccess && !cf.empty())
        e = phub_copy(c, d.coef, cf.data(), cf.size() * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = phub_copy(c, d.start, st.data(), st.size() * sizeof(int), hipMemcpyHostToDevice);
    if (e == hipSuccess && !el.empty())
        e = phub_copy(c, d.elem, el.data(), el.size() * sizeof(int), hipMemcpyHostToDevice);
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
        hipError_t e = upload_prog(c, b == 0 ? p->vec : p->bas, cf, st, el);
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
    if (p->d_map) hipFree(p->d_map);
    delete p;
}

// products after the expansion of S.S (vector path)
extern "C" int phub_opprog_terms(const phub_opprog *p) { return p ? p->vec.nterms : 0; }

// parity of the number of occupied sites with index < J in m: ANTICOM
// (funcmod.f90:29-41) is -1 when it is odd.  Site J (1-based) is bit n - J, so the
// sites before J are the bits above it.  Masks are 32-bit here (n <= 32).
__device__ __forceinline__ uint32_t anticom_par(int n, int J, uint32_t m) {
    if (J <= 1) return 0u;
    return __popc(m & (0xffffffffu << (n - J + 1))) & 1u;
}

// Inverse of one factor: (xu, xd) is the state after it; on success the state
// before it replaces (xu, xd), the factor's forward sign flips par when negative
// and its weight (diagonal factors) multiplies w.
__device__ __forceinline__ bool undo_factor(int code, int s1, int p1, int s2, int p2, int n,
                                            uint32_t &xu, uint32_t &xd, uint32_t &par, double &w) {
    const uint32_t b1 = 1u << ((n - s1) & 31);   // the identity carries site 0
    switch (code) {
    case 3:   // identity
        return true;
    case 1: {   // c^dag_{s1} c_{s2} on spin p1 (C_DAGGER_C, operators.f90:1-55)
        uint32_t &m = p1 == 1 ? xu : xd;
        const uint32_t b2 = 1u << (n - s2);
        if (s1 == s2) return (m & b1) != 0;   // n_J: KINOP keeps an occupied state
        if (!(m & b1) || (m & b2)) return false;
        const uint32_t after = m;
        m = m - b1 + b2;
        par ^= anticom_par(n, s2, m) ^ anticom_par(n, s1, after);
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
        uint32_t &add = code == 6 ? xu : xd, &rem = code == 6 ? xd : xu;
        if (!(add & b1) || (rem & b1)) return false;
        par ^= anticom_par(n, s1, rem) ^ anticom_par(n, s1, add);   // ANTICOM(K, NEW) * ANTICOM(J, NEW2)
        add -= b1;
        rem += b1;
        return true;
    }
    case 9: {   // c^dag_{s1, other spin} c_{s2, p2} (C_DAGGER_C_UNRES, operators.f90:58-144)
        const uint32_t b2 = 1u << (n - s2);
        uint32_t &rem = p2 == 1 ? xu : xd, &add = p2 == 1 ? xd : xu;
        if ((rem & b2) || !(add & b1)) return false;
        par ^= anticom_par(n, s2, rem) ^ anticom_par(n, s1, add);
        add -= b1;
        rem += b2;
        return true;
    }
    case 10: {   // C_DAGGER (operators.f90:397-446): sign from the state before
        uint32_t &m = p1 == 1 ? xu : xd;
        if (!(m & b1)) return false;
        m -= b1;
        par ^= anticom_par(n, s1, m);
        return true;
    }
    case 11: {   // C_ (operators.f90:450-499)
        uint32_t &m = p1 == 1 ? xu : xd;
        if (m & b1) return false;
        m += b1;
        par ^= anticom_par(n, s1, m);
        return true;
    }
    case OP_SZSZ: {   // SZ(J) * SZ(K) with funcmod's SZ (+1/2 up, -1/2 down)
        const uint32_t b2 = 1u << (n - s2);
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

// Position of mask m in its species list: base[popcount] - colex sum, the sum taken
// one nibble of the site-ordered mask (bit p = site p + 1) at a time from LDS.
__device__ __forceinline__ int64_t op_rank(const uint32_t *nib, int64_t base_k, uint32_t m, int n,
                                           int nchunk) {
    uint32_t r = __brev(m) >> (32 - n);
    uint32_t s = 0, sum = 0;
    for (int ch = 0; ch < nchunk; ++ch) {
        const uint32_t v = r & 15u;
        sum += nib[(ch * (n + 1) + (int)s) * 16 + (int)v];
        s += __popc(v);
        r >>= 4;
    }
    return base_k - (int64_t)sum;
}

struct OpArgs {
    int n, nchunk;
    int64_t nsup, nsdn;
    const uint64_t *bup, *bdn;
    const int64_t *base;       // [2][OP_MAXN + 1]
    const uint32_t *nib;       // [nchunk][n + 1][16]
    uint64_t pres_up, pres_dn;
    int nterms;
    const double *coef;
    const int *start, *elem, *touch;
    // source maps (use_map = 1): entry (source << 1) | negative sign, -1 annihilated
    int use_map;
    const int *urow, *drow;
    const int32_t *map_up, *map_dn;
};

// touch bits of a product: 1 / 2 it may move the up / down mask, 4 it changes a
// particle number (then every intermediate state is checked against the lists)
enum { TOUCH_UP = 1, TOUCH_DN = 2, TOUCH_N = 4, TOUCH_MIX = 8 };

// One lane per target state t = d * nsup + u, grid-stride.  The products are walked
// in the same order by every lane, so a product's factors and coefficient are
// wave-uniform and come through scalar loads (the program is read-only here).
// mode 0: out[v][t] = sum_terms coef * sign * in[v][source(t)]    (operator applied)
// mode 1: out[t]    = sum_terms coef * sign * [source(t) == t]    (diagonal, basis-element mode)
// mode 2: out[source(t) * ns + t] += coef * sign                  (matrix, row i = O e_i)
using op_ci32 = const __attribute__((address_space(4))) int;
using op_cf64 = const __attribute__((address_space(4))) double;

// Source state and weight of product q for the target (mu0, md0) = (tu, td): false
// when the product annihilates the target.  The program comes through scalar loads.
__device__ __forceinline__ bool op_source(const OpArgs &a, const uint32_t *s_nib,
                                          const int64_t *s_base, int q, uint32_t mu0,
                                          uint32_t md0, int64_t tu, int64_t td, int64_t &s,
                                          double &w) {
    op_ci32 *st = (op_ci32 *)a.start;
    op_ci32 *el = (op_ci32 *)a.elem;
    const int f0 = st[q], f1 = st[q + 1], touch = ((op_ci32 *)a.touch)[q];
    uint32_t xu = mu0, xd = md0, par = 0u;
    w = ((op_cf64 *)a.coef)[q];
    bool ok = true;
    // every intermediate state must lie in the basis too: the reference applies the
    // factors one by one to vectors of this basis, so a path through an absent
    // sector is cut (its POS lookup fails)
    for (int f = f1 - 1; f >= f0; --f) {
        op_ci32 *e = el + 5 * f;
        const int code = e[0], s1 = e[1], p1 = e[2], s2 = e[3], p2 = e[4];
        if (ok) ok = undo_factor(code, s1, p1, s2, p2, a.n, xu, xd, par, w);
        if (touch & TOUCH_N)
            ok = ok && ((a.pres_up >> __popc(xu)) & 1ull) && ((a.pres_dn >> __popc(xd)) & 1ull);
    }
    if (!ok) return false;
    if (par) w = -w;
    const int64_t su = (touch & TOUCH_UP) ? op_rank(s_nib, s_base[__popc(xu)], xu, a.n, a.nchunk) : tu;
    const int64_t sd = (touch & TOUCH_DN)
        ? op_rank(s_nib, s_base[OP_MAXN + 1 + __popc(xd)], xd, a.n, a.nchunk) : td;
    s = sd * a.nsup + su;
    return true;
}

// One lane per target state t = d * nsup + u, grid-stride, for the basis-element
// path.  The products are walked in the same order by every lane, so a product's
// factors and coefficient are wave-uniform.
// mode 1: out[t]                 = sum_terms coef * sign * [source(t) == t]   (diagonal)
// mode 2: out[source(t) * ns + t] += coef * sign          (matrix, row i = O e_i)
template <int MODE>
__global__ __launch_bounds__(256) void k_op_apply(OpArgs a, double *__restrict__ out) {
    __shared__ uint32_t s_nib[OP_NIB_MAX];
    __shared__ int64_t s_base[2 * (OP_MAXN + 1)];
    const int nn = a.nchunk * (a.n + 1) * 16;
    for (int i = threadIdx.x; i < nn; i += blockDim.x) s_nib[i] = a.nib[i];
    for (int i = threadIdx.x; i < 2 * (OP_MAXN + 1); i += blockDim.x) s_base[i] = a.base[i];
    __syncthreads();
    const int64_t ns = a.nsup * a.nsdn;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t t0 = (int64_t)blockIdx.x * blockDim.x; t0 < ns; t0 += stride) {
        // lanes past the end evaluate the last state and store nothing
        const bool live = t0 + threadIdx.x < ns;
        const int64_t t = live ? t0 + threadIdx.x : ns - 1;
        const int64_t td = t / a.nsup, tu = t - td * a.nsup;
        const uint32_t mu0 = (uint32_t)a.bup[tu], md0 = (uint32_t)a.bdn[td];
        double acc0 = 0.0;
        for (int q = 0; q < a.nterms; ++q) {
            int64_t s;
            double w;
            if (!op_source(a, s_nib, s_base, q, mu0, md0, tu, td, s, w)) continue;
            if (MODE == 1) {
                if (s == t) acc0 += w;
            } else if (live) {
                out[(size_t)s * ns + t] += w;
            }
        }
        if (MODE == 1 && live) out[t] = acc0;
    }
}

// Inverse of one single-species factor (codes 1, 2, 3, 4, 10, 11) on the mask m of
// its species; the same steps as undo_factor.  With uniform arguments the compiler
// keeps it in scalar registers.
__device__ __forceinline__ bool undo_one(int code, int s1, int s2, int n, uint32_t &m, uint32_t &par) {
    const uint32_t b1 = 1u << ((n - s1) & 31);
    switch (code) {
    case 1: {
        const uint32_t b2 = 1u << ((n - s2) & 31);
        if (s1 == s2) return (m & b1) != 0;
        if (!(m & b1) || (m & b2)) return false;
        const uint32_t after = m;
        m = m - b1 + b2;
        par ^= anticom_par(n, s2, m) ^ anticom_par(n, s1, after);
        return true;
    }
    case 2: return (m & b1) != 0;
    case 4: return (m & b1) == 0;
    case 10:
        if (!(m & b1)) return false;
        m -= b1;
        par ^= anticom_par(n, s1, m);
        return true;
    case 11:
        if (m & b1) return false;
        m += b1;
        par ^= anticom_par(n, s1, m);
        return true;
    default:   // identity
        return true;
    }
}

// op_rank with the tables read through scalar loads (a wave-uniform mask)
__device__ __forceinline__ int64_t op_rank_s(const uint32_t *nib_g, const int64_t *base_g, int sp,
                                             uint32_t m, int n, int nchunk) {
    using cu32 = const __attribute__((address_space(4))) uint32_t;
    using ci64 = const __attribute__((address_space(4))) int64_t;
    cu32 *nib = (cu32 *)nib_g;
    uint32_t r = __brev(m) >> (32 - n);
    uint32_t s = 0, sum = 0;
    for (int ch = 0; ch < nchunk; ++ch) {
        const uint32_t v = r & 15u;
        sum += nib[(ch * (n + 1) + (int)s) * 16 + (int)v];
        s += __popc(v);
        r >>= 4;
    }
    return ((ci64 *)base_g)[sp * (OP_MAXN + 1) + __popc(m)] - (int64_t)sum;
}

// Source map of one species: map[row][x] for x over the species list is the source
// of x under the factors of that species in product rowq[row] (walked backwards),
// as (source << 1) | (sign negative), or -1 when they annihilate x or pass through
// an unlisted sector.  Products whose factors each act on one species factorise
// into an up half and a down half (the sign of a factor counts its own species
// only), so O*v needs each half once per product and list element.
__global__ __launch_bounds__(256) void k_op_map(OpArgs a, int sp, const int *__restrict__ rowq,
                                                int nrows, int32_t *__restrict__ map) {
    const int64_t nx = sp == 0 ? a.nsup : a.nsdn;
    const uint64_t *list = sp == 0 ? a.bup : a.bdn;
    const uint64_t pres = sp == 0 ? a.pres_up : a.pres_dn;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < (int64_t)nrows * nx;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int row = (int)(i / nx);
        const int64_t x = i - (int64_t)row * nx;
        const int q = rowq[row];
        uint32_t m = (uint32_t)list[x], par = 0u;
        bool ok = true;
        for (int f = a.start[q + 1] - 1; f >= a.start[q] && ok; --f) {
            const int *e = a.elem + 5 * f;
            if (e[0] == 3 || (e[2] == 1) != (sp == 0)) continue;
            ok = undo_one(e[0], e[1], e[3], a.n, m, par) && ((pres >> __popc(m)) & 1ull);
        }
        map[i] = ok ? (int32_t)((op_rank(a.nib, a.base[sp * (OP_MAXN + 1) + __popc(m)], m, a.n,
                                         a.nchunk) << 1) | par)
                    : -1;
    }
}

// O*v for one vector, a tile of OPR_R rows x 256 columns per workgroup (grid-stride
// over tiles), one column per lane, each wave on the tile's R rows.  For a product
// whose factors each act on one species (hops, densities, c and c^dag: the usual
// Hamiltonian terms) the up half is walked and ranked once per lane for the R rows,
// and the down half, common to the whole wave, once per row in scalar registers.
// Products with a factor that reads both species take the per-state walk.
constexpr int OPR_R = 8;

// MAPONLY: every product takes the map path (no mixed-species factor, maps built):
// the per-state walks are compiled out, which frees registers for occupancy.
template <bool MAPONLY>
__global__ __launch_bounds__(256) void k_op_apply_rows(OpArgs a, const double *__restrict__ in,
                                                       double *__restrict__ out) {
    // the ranking tables serve the per-state walks only
    __shared__ uint32_t s_nib[MAPONLY ? 1 : OP_NIB_MAX];
    __shared__ int64_t s_base[MAPONLY ? 1 : 2 * (OP_MAXN + 1)];
    if constexpr (!MAPONLY) {
        const int nn = a.nchunk * (a.n + 1) * 16;
        for (int i = threadIdx.x; i < nn; i += blockDim.x) s_nib[i] = a.nib[i];
        for (int i = threadIdx.x; i < 2 * (OP_MAXN + 1); i += blockDim.x) s_base[i] = a.base[i];
        __syncthreads();
    }
    op_ci32 *st = (op_ci32 *)a.start;
    op_ci32 *el = (op_ci32 *)a.elem;
    op_ci32 *tc = (op_ci32 *)a.touch;
    op_cf64 *cf = (op_cf64 *)a.coef;
    using cu64 = const __attribute__((address_space(4))) uint64_t;
    const int64_t ncb = (a.nsup + 255) / 256, nrb = (a.nsdn + OPR_R - 1) / OPR_R;
    for (int64_t tile = blockIdx.x; tile < ncb * nrb; tile += gridDim.x) {
        const int64_t rb = tile / ncb, cb = tile - rb * ncb;
        const int64_t d0 = rb * OPR_R;
        const int64_t u = cb * 256 + threadIdx.x;
        const bool live = u < a.nsup;
        const int64_t tu = live ? u : a.nsup - 1;
        const uint32_t mu0 = (uint32_t)a.bup[tu];
        uint32_t md[OPR_R];
#pragma unroll
        for (int r = 0; r < OPR_R; ++r)
            md[r] = (uint32_t)((cu64 *)a.bdn)[d0 + r < a.nsdn ? d0 + r : a.nsdn - 1];
        double acc[OPR_R];
#pragma unroll
        for (int r = 0; r < OPR_R; ++r) acc[r] = 0.0;
        if constexpr (MAPONLY) {
            // every product from the maps: the next product's up entry is loaded
            // before this product's gathers, each gather is issued as soon as its
            // row's source is known, and the row signs stay wave-uniform bits
            using ci32m = const __attribute__((address_space(4))) int32_t;
            auto up_entry = [&](int q) -> int32_t {
                const int ur = ((op_ci32 *)a.urow)[q];
                return ur < 0 ? (int32_t)(tu << 1) : a.map_up[(int64_t)ur * a.nsup + tu];
            };
            int32_t eu_next = a.nterms > 0 ? up_entry(0) : -1;
            for (int q = 0; q < a.nterms; ++q) {
                const int32_t eu = eu_next;
                if (q + 1 < a.nterms) eu_next = up_entry(q + 1);
                if (__ballot(eu >= 0) == 0) continue;
                const double c = cf[q];
                const int dr = ((op_ci32 *)a.drow)[q];
                const int64_t su = eu >= 0 ? (int64_t)(eu >> 1) : tu;
                const double wu = eu < 0 ? 0.0 : ((eu & 1) ? -c : c);
                double g[OPR_R];
                uint32_t neg = 0u;
#pragma unroll
                for (int r = 0; r < OPR_R; ++r) {
                    const int64_t d = d0 + r;
                    const int32_t ed = d >= a.nsdn ? -1
                        : dr < 0 ? (int32_t)(d << 1) : ((ci32m *)a.map_dn)[(int64_t)dr * a.nsdn + d];
                    neg |= (ed >= 0 && (ed & 1)) ? (1u << r) : 0u;
                    g[r] = (ed >= 0 && eu >= 0) ? in[(int64_t)(ed >> 1) * a.nsup + su] : 0.0;
                }
#pragma unroll
                for (int r = 0; r < OPR_R; ++r) acc[r] += ((neg >> r) & 1u ? -wu : wu) * g[r];
            }
        } else
        for (int q = 0; q < a.nterms; ++q) {
            const int f0 = st[q], f1 = st[q + 1], touch = tc[q];
            const double c = cf[q];
            int64_t idx[OPR_R];
            double wr[OPR_R];
            if (touch & TOUCH_MIX) {
#pragma unroll
                for (int r = 0; r < OPR_R; ++r) {
                    int64_t sq;
                    double w;
                    const bool ok = d0 + r < a.nsdn &&
                                    op_source(a, s_nib, s_base, q, mu0, md[r], tu, d0 + r, sq, w);
                    idx[r] = ok ? sq : tu;
                    wr[r] = ok ? w : 0.0;
                }
            } else if (a.use_map) {
                // up half from the map (one coalesced load per lane), down half from
                // the map through scalar loads (wave-uniform rows)
                using ci32m = const __attribute__((address_space(4))) int32_t;
                const int ur = ((op_ci32 *)a.urow)[q], dr = ((op_ci32 *)a.drow)[q];
                const int32_t eu = ur < 0 ? (int32_t)(tu << 1) : a.map_up[(int64_t)ur * a.nsup + tu];
                if (__ballot(eu >= 0) == 0) continue;
                const int64_t su = eu >= 0 ? (int64_t)(eu >> 1) : tu;
                const double wu = eu < 0 ? 0.0 : ((eu & 1) ? -c : c);
#pragma unroll
                for (int r = 0; r < OPR_R; ++r) {
                    const int64_t d = d0 + r;
                    const int32_t ed = d >= a.nsdn ? -1
                        : dr < 0 ? (int32_t)(d << 1) : ((ci32m *)a.map_dn)[(int64_t)dr * a.nsdn + d];
                    idx[r] = ed >= 0 ? (int64_t)(ed >> 1) * a.nsup + su : tu;
                    wr[r] = ed < 0 ? 0.0 : ((ed & 1) ? -wu : wu);
                }
            } else {
                // up half: per lane, once for the R rows
                uint32_t xu = mu0, pu = 0u;
                bool oku = true;
                for (int f = f1 - 1; f >= f0; --f) {
                    op_ci32 *e = el + 5 * f;
                    const int code = e[0];
                    if (code == 3 || e[2] != 1) continue;
                    if (oku) oku = undo_one(code, e[1], e[3], a.n, xu, pu);
                    if (touch & TOUCH_N) oku = oku && ((a.pres_up >> __popc(xu)) & 1ull);
                }
                if (__ballot(oku) == 0) continue;
                const int64_t su = !oku ? tu
                    : (touch & TOUCH_UP) ? op_rank(s_nib, s_base[__popc(xu)], xu, a.n, a.nchunk) : tu;
                const double wu = !oku ? 0.0 : (pu ? -c : c);
                // down half: wave-uniform, once per row
#pragma unroll
                for (int r = 0; r < OPR_R; ++r) {
                    uint32_t xd = md[r], pd = 0u;
                    bool okd = d0 + r < a.nsdn;
                    for (int f = f1 - 1; f >= f0 && okd; --f) {
                        op_ci32 *e = el + 5 * f;
                        const int code = e[0];
                        if (code == 3 || e[2] == 1) continue;
                        okd = undo_one(code, e[1], e[3], a.n, xd, pd);
                        if (touch & TOUCH_N) okd = okd && ((a.pres_dn >> __popc(xd)) & 1ull);
                    }
                    const int64_t sd = !okd ? 0
                        : (touch & TOUCH_DN) ? op_rank_s(a.nib, a.base, 1, xd, a.n, a.nchunk) : d0 + r;
                    idx[r] = okd ? sd * a.nsup + su : tu;
                    wr[r] = okd ? (pd ? -wu : wu) : 0.0;
                }
            }
            // lanes the product annihilates load nothing
            double g[OPR_R];
#pragma unroll
            for (int r = 0; r < OPR_R; ++r) g[r] = wr[r] != 0.0 ? in[idx[r]] : 0.0;
#pragma unroll
            for (int r = 0; r < OPR_R; ++r) acc[r] += wr[r] * g[r];
        }
        if (live)
#pragma unroll
            for (int r = 0; r < OPR_R; ++r)
                if (d0 + r < a.nsdn) out[(d0 + r) * a.nsup + u] = acc[r];
    }
}

static unsigned op_grid(int64_t ns) {
    // enough 256-lane blocks to fill the chip several times over; each loads the
    // ranking tables once
    return (unsigned)std::min<int64_t>((ns + 255) / 256, 8192);
}

static OpArgs op_args(phub_ctx *c, const phub_fock *f, const OpProgDev &p) {
    (void)c;
    OpArgs a{};
    a.n = f->n;
    a.nchunk = f->nchunk;
    a.nsup = f->nsup;
    a.nsdn = f->nsdn;
    a.bup = f->bup;
    a.bdn = f->bdn;
    a.base = f->d_base;
    a.nib = f->d_nib;
    a.pres_up = f->pres_up;
    a.pres_dn = f->pres_dn;
    a.nterms = p.nterms;
    a.coef = p.coef;
    a.start = p.start;
    a.elem = p.elem;
    a.touch = p.touch;
    a.use_map = 0;
    a.urow = p.urow;
    a.drow = p.drow;
    return a;
}

// Build (or reuse) the source maps of the vector program on basis f.  They are used
// when both lists stay below 2^30 elements and the maps within 1 GiB; otherwise (or
// with PHUB_OP_MAPS=0 in the environment) the kernel walks the factors itself.
static int op_maps(phub_ctx *c, const phub_fock *f, const phub_opprog *pc, OpArgs &a) {
    if (const char *env = getenv("PHUB_OP_MAPS"))
        if (env[0] == '0') return PHUB_OK;
    phub_opprog *p = const_cast<phub_opprog *>(pc);
    const OpProgDev &v = p->vec;
    if (f->nsup >= (1ll << 30) || f->nsdn >= (1ll << 30)) return PHUB_OK;
    const size_t nu = (size_t)v.nrow_up * (size_t)f->nsup, nd = (size_t)v.nrow_dn * (size_t)f->nsdn;
    const size_t bytes = (nu + nd) * sizeof(int32_t);
    if (bytes > (size_t(1) << 30)) return PHUB_OK;
    if (p->map_fock != f->id) {
        p->map_fock = 0;
        if (bytes > p->map_cap) {
            if (p->d_map) hipFree(p->d_map);
            p->d_map = nullptr;
            p->map_cap = 0;
            PHUB_HIP(c, phub_malloc(&p->d_map, std::max<size_t>(bytes, 4)));
            p->map_cap = std::max<size_t>(bytes, 4);
        }
        const OpArgs b = a;
        if (nu)
            hipLaunchKernelGGL(k_op_map, dim3((unsigned)std::min<size_t>((nu + 255) / 256, 65536)),
                               dim3(256), 0, c->stream, b, 0, v.rowq_up, v.nrow_up, p->d_map);
        if (nd)
            hipLaunchKernelGGL(k_op_map, dim3((unsigned)std::min<size_t>((nd + 255) / 256, 65536)),
                               dim3(256), 0, c->stream, b, 1, v.rowq_dn, v.nrow_dn, p->d_map + nu);
        PHUB_HIP(c, hipGetLastError());
        p->map_fock = f->id;
    }
    a.use_map = 1;
    a.map_up = p->d_map;
    a.map_dn = p->d_map + nu;
    return PHUB_OK;
}

extern "C" int phub_operator_apply(phub_ctx *c, const phub_fock *f, const phub_opprog *p,
                                   const double *d_in, double *d_out, int nvec) {
    if (!c || !f || !p || !d_in || !d_out) return phub_fail(c, PHUB_ERR_ARG, "null argument");
    if (nvec < 1) return phub_fail(c, PHUB_ERR_ARG, "nvec must be >= 1");
    if (p->max_site > f->n) return phub_fail(c, PHUB_ERR_ARG, "operator site beyond the basis");
    const int64_t ns = f->nsup * f->nsdn;
    if (ns == 0) return PHUB_OK;
    const size_t span = (size_t)nvec * (size_t)ns;
    if (d_in < d_out + span && d_out < d_in + span)
        return phub_fail(c, PHUB_ERR_ARG, "out must not overlap in");
    PHUB_HIP(c, hipSetDevice(c->device));
    OpArgs a = op_args(c, f, p->vec);
    if (int rc = op_maps(c, f, p, a)) return rc;
    const int64_t tiles = ((f->nsup + 255) / 256) * ((f->nsdn + OPR_R - 1) / OPR_R);
    const unsigned grid = (unsigned)std::min<int64_t>(tiles, 16384);
    const bool maponly = a.use_map && !p->vec.has_mix;
    for (int v = 0; v < nvec; ++v) {
        if (maponly)
            hipLaunchKernelGGL(k_op_apply_rows<true>, dim3(grid), dim3(256), 0, c->stream, a,
                               d_in + (size_t)v * ns, d_out + (size_t)v * ns);
        else
            hipLaunchKernelGGL(k_op_apply_rows<false>, dim3(grid), dim3(256), 0, c->stream, a,
                               d_in + (size_t)v * ns, d_out + (size_t)v * ns);
    }
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
    hipLaunchKernelGGL(k_op_apply<1>, dim3(op_grid(ns)), dim3(256), 0, c->stream,
                       op_args(c, f, p->bas), d_diag);
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
    hipLaunchKernelGGL(k_op_apply<2>, dim3(op_grid(ns)), dim3(256), 0, c->stream,
                       op_args(c, f, p->bas), d_mat);
    PHUB_HIP(c, hipGetLastError());
    return PHUB_OK;
}
