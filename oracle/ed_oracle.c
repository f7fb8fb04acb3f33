/*
 * ed_oracle.c -- CPU restatement of PyHub's exact-diagonalization hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the oracle: it is linked and called
 * only by tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg,
 * and only as the checker (or the timed CPU baseline) -- never by the product
 * path in pyhub_amd/, which must run on the HIP extension or fail.
 *
 * Every function restates one routine of the reference (Fortran, serial, FP64)
 * and cites the file:line it follows.  Conventions kept from the reference:
 *   - site J (1-based) is bit NORB-J of a mask (funcmod.f90:22-23);
 *   - state index (d)*NSUP+u, u fastest (hprod_hubbard.f90:31);
 *   - basis order is the one HILBERT/MOVE/BMOVE generate (compute_basis.f90:67-167).
 * Masks are uint64 here (the reference uses INTEGER*4, so NORB<=31 there).
 *
 * Pinning of each function (how a plausible mistake would be caught):
 *   orc_hilbert        brute-force enumeration + the closed-form rank, n<=14
 *                      (tests/test_oracle.py) and GPU unrank parity.
 *   orc_hprod_hubbard  bit-exact against the reference HPROD compiled from
 *                      its own sources (oracle/_ref, tests/test_oracle.py);
 *                      U=0 free-fermion energies, Hubbard dimer closed form,
 *                      Lieb-Wu finite-ring Bethe ansatz.
 *   orc_hprod_heis     bit-exact against the reference Heisenberg HPROD
 *                      (oracle/_ref); 2x standard spectrum on bipartite rings.
 *   orc_hprod_heis_fast bit-identical to orc_hprod_heis (tests/test_oracle.py).
 *   orc_hprod_hubbard_ul bit-exact against the reference HPROD with is_Uloc = 3
 *                      (oracle/_ref); equal to the per-site U form when UL holds
 *                      only UL(J,J,J,J) = U_J.
 *   orc_two_rdm        its (J,J,J,J) entries equal the double occupancy NI;
 *                      E_int = sum UL * DBL_OCC equals <psi|H_U|psi>.
 *   orc_lanczos        converged E0 against dense FP64 eigh.
 *   orc_one_rdm, orc_ni, orc_two_rdm, orc_j_spin
 *                      bit-exact against the reference's own ONE_RDM / TWO_RDM
 *                      / J_SPIN (static_rq.f90:153-395, oracle/_ref); sum rules
 *                      (tr rho = N), Hellmann-Feynman vs finite differences.
 *   orc_j_spin_heis    bit-exact against the reference's Heisenberg J_SPIN
 *                      (static_rq_heis.f90:58-116, oracle/_ref).
 *   orc_excite         bit-exact against the reference's GF_E_UP / GF_E_DOWN /
 *                      GF_H_UP / GF_H_DOWN (spgf_ed.f90:304-429, oracle/_ref);
 *                      sum rule sum_k w_k = 1.
 *   orc_cf_chain       on a sector the chain exhausts, its continued fraction
 *                      equals the dense-ED spectral sum sum_k |<k|x>|^2/(z-E_k);
 *                      alpha_0 = <x|H|x>/<x|x> of the pinned HPROD.
 *   orc_tridiag_eig    through orc_lanczos against dense eigh.
 * The whole file also runs under ASan + UBSan (Makefile target `sanitize`).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>

typedef uint64_t u64;

/* ---------------------------------------------------------------------------
 * bit primitives: funcmod.f90:18-94
 * ------------------------------------------------------------------------- */
static inline u64 sitebit(int n, int pos) { return (u64)1 << (n - pos); }

/* KINOP (funcmod.f90:18-27): c^dag_posf c_posi on x, 0 when forbidden. */
static inline long long kinop(int n, u64 x, int posi, int posf) {
    u64 a = sitebit(n, posi), b = sitebit(n, posf);
    long long r = 0;
    if ((x & a) == a && (x & b) == 0) r = (long long)(x - a + b);
    if ((x & a) == a && posi == posf) r = (long long)x;
    return r;
}

/* ANTICOM (funcmod.f90:29-41): (-1)^{#occupied sites 1..pos-1}. */
static inline double anticom(int n, int pos, u64 x) {
    double s = 1.0;
    for (int i = 1; i < pos; ++i)
        if ((x & sitebit(n, i)) == sitebit(n, i)) s = -s;
    return s;
}

static inline int is_part(int n, int pos, u64 x) { return (x & sitebit(n, pos)) != 0; }

/* SZ (funcmod.f90:76-84) */
static inline double szf(int n, int pos, u64 xu, u64 xd) {
    double s = 0.0;
    if (is_part(n, pos, xu)) s += 0.5;
    if (is_part(n, pos, xd)) s -= 0.5;
    return s;
}

/* recursive BINOM (funcmod.f90:7-16), restated iteratively */
u64 orc_binom(int n, int k) {
    if (k < 0 || k > n) return 0;
    u64 r = 1;
    for (int i = 1; i <= k; ++i) r = r * (u64)(n - k + i) / (u64)i;
    return r;
}

/* ---------------------------------------------------------------------------
 * basis: HILBERT / MOVE / BMOVE (compute_basis.f90:67-96,127-135,139-167)
 * ------------------------------------------------------------------------- */
static u64 move_(u64 s, int posf) {            /* compute_basis.f90:127-135 */
    u64 a = (u64)1 << (posf - 1);
    u64 b = s ^ a;
    a = (u64)1 << posf;
    return b ^ a;
}

static void bmove_(int no, u64 statef, u64 *statem, int *posm) { /* :139-167 */
    int nstatem = 0;
    u64 s = statef;
    u64 a = (u64)1 << (no - nstatem - 1);
    while ((statef & a) == a) {
        nstatem++;
        a = (u64)1 << (no - nstatem - 1);
    }
    for (int i = 1; i <= nstatem; ++i) s -= (u64)1 << (no - i);
    int posle = no;
    a = (u64)1 << (posle - 1);
    while ((s & a) == 0) {
        posle--;
        a = (u64)1 << (posle - 1);
    }
    s -= (u64)1 << (posle - 1);
    u64 m = s;
    for (int i = 1; i <= nstatem + 1; ++i) m += (u64)1 << (posle - 1 + i);
    *statem = m;
    *posm = posle + 1 + nstatem;
}

/* HILBERT (compute_basis.f90:67-96): fills basis[ns] in the reference order. */
void orc_hilbert(int no, int ne, long long ns, u64 *basis) {
    for (long long i = 0; i < ns; ++i) basis[i] = 0;
    for (int i = 1; i <= ne; ++i) basis[0] += (u64)1 << (i - 1);
    long long cpt = 1;                      /* 1-based count of filled entries */
    int pos = ne;
    while (cpt < ns) {
        while (pos < no) {
            basis[cpt] = move_(basis[cpt - 1], pos);
            cpt++;
            pos++;
            if (cpt >= ns) break;
        }
        if (cpt < ns) {
            u64 m;
            bmove_(no, basis[cpt - 1], &m, &pos);
            basis[cpt] = m;
            cpt++;
        }
    }
}

/* linear search for a mask, as the reference does (hprod_hubbard.f90:42-47);
 * returns the 0-based position or -1 */
static inline long long lsearch(const u64 *b, long long ns, u64 x) {
    for (long long i = 0; i < ns; ++i)
        if (b[i] == x) return i;
    return -1;
}

/* ---------------------------------------------------------------------------
 * Hubbard H*v: HPROD (hprod_hubbard.f90:17-118), is_Uloc = 2 (per-site U).
 * T is row-major T[(J-1)*n+(K-1)] = T(J,K).  Same loop order and the same
 * linear search as the reference; the 4-index UL branch (:62-89) is not
 * restated (UPARTIAL = 0 for is_Uloc = 2).
 * ------------------------------------------------------------------------- */
void orc_hprod_hubbard(int n, int nup, int ndn, long long nsup, long long nsdn,
                       const u64 *bup, const u64 *bdn, const double *T,
                       const double *U, const double *v, double *w) {
    long long nst = nsup * nsdn;
    for (long long i = 0; i < nst; ++i) w[i] = 0.0;
    for (int J = 1; J <= n; ++J) {
        /* U diagonal :27-35 */
        for (long long u = 0; u < nsup; ++u) {
            if (is_part(n, J, bup[u])) {
                for (long long d = 0; d < nsdn; ++d)
                    if (is_part(n, J, bdn[d]))
                        w[d * nsup + u] = w[d * nsup + u] + v[d * nsup + u] * U[J - 1];
            }
        }
        for (int K = 1; K <= n; ++K) {
            double t = T[(J - 1) * n + (K - 1)];
            if (!(fabs(t) >= 1.e-14)) continue;
            /* up hops :38-61 */
            for (long long u = 0; u < nsup; ++u) {
                long long nw = kinop(n, bup[u], J, K);
                if (nw > 0) {
                    long long pos = lsearch(bup, nsup, (u64)nw);
                    if (pos >= 0) {
                        double ac = anticom(n, J, bup[u]);
                        ac = ac * anticom(n, K, (u64)nw);
                        if (nup == ndn) {
                            for (long long d = 0; d < nsdn; ++d) {
                                w[d * nsup + pos] = w[d * nsup + pos] + ac * v[d * nsup + u] * t;
                                w[pos * nsup + d] = w[pos * nsup + d] + ac * v[u * nsup + d] * t;
                            }
                        } else {
                            for (long long d = 0; d < nsdn; ++d)
                                w[d * nsup + pos] = w[d * nsup + pos] + ac * v[d * nsup + u] * t;
                        }
                    }
                }
            }
            /* down hops :93-114 */
            if (nup != ndn) {
                for (long long d = 0; d < nsdn; ++d) {
                    long long nw = kinop(n, bdn[d], J, K);
                    if (nw > 0) {
                        long long pos = lsearch(bdn, nsdn, (u64)nw);
                        if (pos >= 0) {
                            double ac = anticom(n, J, bdn[d]);
                            ac = ac * anticom(n, K, (u64)nw);
                            for (long long u = 0; u < nsup; ++u)
                                w[pos * nsup + u] = w[pos * nsup + u] + ac * v[d * nsup + u] * t;
                        }
                    }
                }
            }
        }
    }
}

/* ---------------------------------------------------------------------------
 * Hubbard H*v with a 4-index interaction (is_Uloc = 3): HPROD
 * (hprod_hubbard.f90:17-118) with
 *   sum_{JKMN} UL(J,K,M,N) c^dag_{K up} c_{J up} c^dag_{N dn} c_{M dn}   (:62-89)
 * and no per-site U.  UL row-major UL[(((J-1)*n+(K-1))*n+(M-1))*n+(N-1)];
 * UP[(J-1)*n+(K-1)] = UPARTIAL(J,K) = sum_{MN} UL(J,K,M,N)^2 (core_solver.py:146).
 * Pairs (J,K) enter when |T(J,K)| >= 1e-14 or UPARTIAL(J,K) >= 1e-14 (:37).
 * ------------------------------------------------------------------------- */
void orc_hprod_hubbard_ul(int n, int nup, int ndn, long long nsup, long long nsdn,
                          const u64 *bup, const u64 *bdn, const double *T,
                          const double *UL, const double *UP, const double *v, double *w) {
    long long nst = nsup * nsdn;
    for (long long i = 0; i < nst; ++i) w[i] = 0.0;
    for (int J = 1; J <= n; ++J) {
        for (int K = 1; K <= n; ++K) {
            double t = T[(J - 1) * n + (K - 1)];
            double upj = UP[(J - 1) * n + (K - 1)];
            if (!((fabs(t) >= 1.e-14) || (upj >= 1.e-14))) continue;
            for (long long u = 0; u < nsup; ++u) {
                long long nw = kinop(n, bup[u], J, K);
                if (nw > 0) {
                    long long pos = lsearch(bup, nsup, (u64)nw);
                    if (pos >= 0) {
                        double ac = anticom(n, J, bup[u]);
                        ac = ac * anticom(n, K, (u64)nw);
                        if (nup == ndn) {
                            for (long long d = 0; d < nsdn; ++d) {
                                w[d * nsup + pos] = w[d * nsup + pos] + ac * v[d * nsup + u] * t;
                                w[pos * nsup + d] = w[pos * nsup + d] + ac * v[u * nsup + d] * t;
                            }
                        } else {
                            for (long long d = 0; d < nsdn; ++d)
                                w[d * nsup + pos] = w[d * nsup + pos] + ac * v[d * nsup + u] * t;
                        }
                        if (upj >= 1.e-14) {
                            for (int M = 1; M <= n; ++M)
                                for (int N = 1; N <= n; ++N) {
                                    double ul = UL[(((size_t)(J - 1) * n + (K - 1)) * n + (M - 1)) * n + (N - 1)];
                                    if (!(fabs(ul) >= 1.e-14)) continue;
                                    for (long long d = 0; d < nsdn; ++d) {
                                        long long nw2 = kinop(n, bdn[d], M, N);
                                        if (nw2 > 0) {
                                            long long pos2 = lsearch(bdn, nsdn, (u64)nw2);
                                            if (pos2 >= 0) {
                                                double ac2 = anticom(n, M, bdn[d]);
                                                ac2 = ac2 * anticom(n, N, (u64)nw2);
                                                w[pos2 * nsup + pos] = w[pos2 * nsup + pos] +
                                                                       ac * ac2 * v[d * nsup + u] * ul;
                                            }
                                        }
                                    }
                                }
                        }
                    }
                }
            }
            if (nup != ndn) {
                for (long long d = 0; d < nsdn; ++d) {
                    long long nw = kinop(n, bdn[d], J, K);
                    if (nw > 0) {
                        long long pos = lsearch(bdn, nsdn, (u64)nw);
                        if (pos >= 0) {
                            double ac = anticom(n, J, bdn[d]);
                            ac = ac * anticom(n, K, (u64)nw);
                            for (long long u = 0; u < nsup; ++u)
                                w[pos * nsup + u] = w[pos * nsup + u] + ac * v[d * nsup + u] * t;
                        }
                    }
                }
            }
        }
    }
}

/* DBL_OCC of TWO_RDM (static_rq.f90:247-306) for <bra| . |ket>: computed for
 * K <= J, N <= M and mirrored to (J,K,N,M), (K,J,M,N), (K,J,N,M); out row-major
 * as UL above. */
void orc_two_rdm(int n, long long nsup, long long nsdn, const u64 *bup, const u64 *bdn,
                 const double *bra, const double *ket, double *out) {
    for (int J = 1; J <= n; ++J)
        for (int K = 1; K <= J; ++K)
            for (int M = 1; M <= n; ++M)
                for (int N = 1; N <= M; ++N) {
                    double acc = 0.0;
                    for (long long i = 0; i < nsup; ++i) {
                        long long nw = kinop(n, bup[i], J, K);
                        if (nw <= 0) continue;
                        long long pos = lsearch(bup, nsup, (u64)nw);
                        if (pos < 0) continue;
                        double ac = anticom(n, J, bup[i]);
                        ac = ac * anticom(n, K, (u64)nw);
                        for (long long l1 = 0; l1 < nsdn; ++l1) {
                            long long nw2 = kinop(n, bdn[l1], M, N);
                            if (nw2 <= 0) continue;
                            long long pos2 = lsearch(bdn, nsdn, (u64)nw2);
                            if (pos2 < 0) continue;
                            double ac2 = anticom(n, M, bdn[l1]);
                            ac2 = ac2 * anticom(n, N, (u64)nw2);
                            acc = acc + ac * ac2 * bra[l1 * nsup + i] * ket[pos2 * nsup + pos];
                        }
                    }
#define IX(a, b, c, d) ((((size_t)(a - 1) * n + (b - 1)) * n + (c - 1)) * n + (d - 1))
                    out[IX(J, K, M, N)] = acc;
                    out[IX(J, K, N, M)] = acc;
                    out[IX(K, J, M, N)] = acc;
                    out[IX(K, J, N, M)] = acc;
#undef IX
                }
}

/* ---------------------------------------------------------------------------
 * Heisenberg H*v: HPROD (hprod_heis.f90:1-61), S_z = 0 sector, down = complement.
 * Jm row-major Jm[(J-1)*n+(K-1)] = J_MATRIX(J,K).
 * ------------------------------------------------------------------------- */
void orc_hprod_heis(int n, long long nsup, const u64 *bup, const double *Jm,
                    const double *v, double *w) {
    u64 full = (n >= 64) ? ~(u64)0 : (((u64)1 << n) - 1);
    for (long long i = 0; i < nsup; ++i) w[i] = 0.0;
    for (int J = 1; J <= n; ++J) {
        double jjj = Jm[(J - 1) * n + (J - 1)];
        if (fabs(jjj) > 1.e-14) {                                    /* :11-21 */
            for (long long i = 0; i < nsup; ++i) {
                u64 bd = full - bup[i];
                if (is_part(n, J, bup[i]) && !is_part(n, J, bd))
                    w[i] = w[i] + v[i] * 0.5 * jjj;
                else if (!is_part(n, J, bup[i]) && is_part(n, J, bd))
                    w[i] = w[i] + v[i] * 0.5 * jjj;
            }
        }
        for (int K = 1; K <= n; ++K) {
            double jjk = Jm[(J - 1) * n + (K - 1)];
            if (fabs(jjk) > 1.e-14 && J != K) {                      /* :22-53 */
                for (long long i = 0; i < nsup; ++i) {
                    long long nw = kinop(n, bup[i], J, K);
                    if (nw > 0) {
                        long long pos = lsearch(bup, nsup, (u64)nw);
                        if (pos >= 0) {
                            u64 bd = full - bup[i];
                            double ac = anticom(n, J, bup[i]);
                            ac = ac * anticom(n, K, (u64)nw);
                            long long nw2 = kinop(n, bd, K, J);
                            if ((u64)nw2 == full - bup[pos]) {
                                ac = ac * anticom(n, J, bd);
                                ac = ac * anticom(n, K, (u64)nw2);
                                w[pos] = w[pos] - ac * v[i] * jjk;
                            }
                        }
                    }
                }
            }
            for (long long i = 0; i < nsup; ++i) {                    /* :54-58 */
                u64 bd = full - bup[i];
                w[i] = w[i] + v[i] * szf(n, J, bup[i], bd) * szf(n, K, bup[i], bd) * jjk;
            }
        }
    }
}

/* ---------------------------------------------------------------------------
 * Faster CPU baseline ("best effort"): same operator, rank tables instead of
 * the linear search, OpenMP over down rows (no counterpart in the reference;
 * used only as the multi-core CPU baseline).
 * ------------------------------------------------------------------------- */
static long long colex_rank_ref(int n, int k, u64 x) {
    /* reference-order rank: C(n,k)-1-colex(rev_n(x)) (closed form of HILBERT) */
    u64 c = 0;
    int i = 0;
    for (int b = n - 1; b >= 0; --b)
        if (x >> b & 1) { ++i; c += orc_binom(n - 1 - b, i); }
    return (long long)(orc_binom(n, k) - 1 - c);
}

long long orc_rank_ref(int n, int k, u64 x) { return colex_rank_ref(n, k, x); }

/* orc_hprod_heis with the linear search replaced by the closed-form rank (the
 * basis is complete and in reference order, so both give the same position);
 * identical arithmetic in identical order, hence bit-identical output. */
void orc_hprod_heis_fast(int n, long long nsup, const u64 *bup, const double *Jm,
                         const double *v, double *w) {
    u64 full = (n >= 64) ? ~(u64)0 : (((u64)1 << n) - 1);
    for (long long i = 0; i < nsup; ++i) w[i] = 0.0;
    for (int J = 1; J <= n; ++J) {
        double jjj = Jm[(J - 1) * n + (J - 1)];
        if (fabs(jjj) > 1.e-14) {
            for (long long i = 0; i < nsup; ++i) {
                u64 bd = full - bup[i];
                if (is_part(n, J, bup[i]) && !is_part(n, J, bd))
                    w[i] = w[i] + v[i] * 0.5 * jjj;
                else if (!is_part(n, J, bup[i]) && is_part(n, J, bd))
                    w[i] = w[i] + v[i] * 0.5 * jjj;
            }
        }
        for (int K = 1; K <= n; ++K) {
            double jjk = Jm[(J - 1) * n + (K - 1)];
            if (fabs(jjk) > 1.e-14 && J != K) {
                for (long long i = 0; i < nsup; ++i) {
                    long long nw = kinop(n, bup[i], J, K);
                    if (nw > 0) {
                        long long pos = colex_rank_ref(n, n / 2, (u64)nw);
                        u64 bd = full - bup[i];
                        double ac = anticom(n, J, bup[i]);
                        ac = ac * anticom(n, K, (u64)nw);
                        long long nw2 = kinop(n, bd, K, J);
                        if ((u64)nw2 == full - bup[pos]) {
                            ac = ac * anticom(n, J, bd);
                            ac = ac * anticom(n, K, (u64)nw2);
                            w[pos] = w[pos] - ac * v[i] * jjk;
                        }
                    }
                }
            }
            for (long long i = 0; i < nsup; ++i) {
                u64 bd = full - bup[i];
                w[i] = w[i] + v[i] * szf(n, J, bup[i], bd) * szf(n, K, bup[i], bd) * jjk;
            }
        }
    }
}

typedef struct { long long pos; double c; } hop_t;

static void build_hops(int n, long long ns, int k, const u64 *b, const double *T,
                       hop_t **tab, int **cnt, int *maxh) {
    int mh = 0;
    *cnt = (int *)calloc((size_t)ns, sizeof(int));
    for (long long s = 0; s < ns; ++s) {
        int c = 0;
        for (int J = 1; J <= n; ++J)
            for (int K = 1; K <= n; ++K) {
                if (J == K) continue;
                double t = T[(J - 1) * n + (K - 1)];
                if (!(fabs(t) >= 1.e-14)) continue;
                if (kinop(n, b[s], J, K) > 0) c++;
            }
        (*cnt)[s] = c;
        if (c > mh) mh = c;
    }
    *maxh = mh;
    *tab = (hop_t *)malloc(sizeof(hop_t) * (size_t)(mh > 0 ? mh : 1) * (size_t)ns);
    for (long long s = 0; s < ns; ++s) {
        int c = 0;
        for (int J = 1; J <= n; ++J)
            for (int K = 1; K <= n; ++K) {
                if (J == K) continue;
                double t = T[(J - 1) * n + (K - 1)];
                if (!(fabs(t) >= 1.e-14)) continue;
                long long nw = kinop(n, b[s], J, K);
                if (nw > 0) {
                    double ac = anticom(n, J, b[s]) * anticom(n, K, (u64)nw);
                    (*tab)[s * mh + c].pos = colex_rank_ref(n, k, (u64)nw);
                    (*tab)[s * mh + c].c = ac * t;
                    c++;
                }
            }
    }
}

void orc_hprod_hubbard_fast(int n, int nup, int ndn, long long nsup, long long nsdn,
                            const u64 *bup, const u64 *bdn, const double *T,
                            const double *U, const double *v, double *w, int nthreads) {
    hop_t *tu, *td;
    int *cu, *cd, mu, md;
    build_hops(n, nsup, nup, bup, T, &tu, &cu, &mu);
    build_hops(n, nsdn, ndn, bdn, T, &td, &cd, &md);
    double *eu = (double *)malloc(sizeof(double) * (size_t)nsup);
    double *ed = (double *)malloc(sizeof(double) * (size_t)nsdn);
    for (long long s = 0; s < nsup; ++s) {
        double e = 0;
        for (int J = 1; J <= n; ++J) if (is_part(n, J, bup[s])) e += T[(J - 1) * n + J - 1];
        eu[s] = e;
    }
    for (long long s = 0; s < nsdn; ++s) {
        double e = 0;
        for (int J = 1; J <= n; ++J) if (is_part(n, J, bdn[s])) e += T[(J - 1) * n + J - 1];
        ed[s] = e;
    }
    (void)nthreads;
#pragma omp parallel for schedule(dynamic, 4) num_threads(nthreads)
    for (long long d = 0; d < nsdn; ++d) {
        const double *vd = v + d * nsup;
        double *wd = w + d * nsup;
        for (long long u = 0; u < nsup; ++u) {
            u64 both = bup[u] & bdn[d];
            double dg = eu[u] + ed[d];
            for (int J = 1; J <= n; ++J) if (both & sitebit(n, J)) dg += U[J - 1];
            double acc = dg * vd[u];
            const hop_t *h = tu + u * mu;
            for (int k = 0; k < cu[u]; ++k) acc += h[k].c * vd[h[k].pos];
            const hop_t *g = td + d * md;
            for (int k = 0; k < cd[d]; ++k) acc += g[k].c * v[g[k].pos * nsup + u];
            wd[u] = acc;
        }
    }
    free(tu); free(td); free(cu); free(cd); free(eu); free(ed);
}

/* ---------------------------------------------------------------------------
 * Symmetric tridiagonal eigen-solver (implicit QL with Wilkinson shifts),
 * written from the textbook algorithm.  It stands in for LAPACK DSYEVX in
 * DIAGMAT (diagmat.f90:1-45), which the reference calls on the Lanczos
 * tridiagonal (lanczos.f90:61-71).  d[0..m-1] diagonal, e[0..m-2] off-diagonal.
 * On return d holds eigenvalues ascending, z (m*m, column-major, column j =
 * eigenvector j) if non-NULL.  Returns 0 on success.
 * ------------------------------------------------------------------------- */
int orc_tridiag_eig(int m, double *d, const double *e_in, double *z) {
    double *e = (double *)calloc((size_t)m + 1, sizeof(double));
    for (int i = 0; i + 1 < m; ++i) e[i] = e_in[i];
    if (z) {
        for (int i = 0; i < m * m; ++i) z[i] = 0.0;
        for (int i = 0; i < m; ++i) z[i * m + i] = 1.0;
    }
    for (int l = 0; l < m; ++l) {
        int iter = 0;
        for (;;) {
            int mm;
            for (mm = l; mm < m - 1; ++mm) {
                double dd = fabs(d[mm]) + fabs(d[mm + 1]);
                if (fabs(e[mm]) <= 1e-300 + 2.2e-16 * dd) break;
            }
            if (mm == l) break;
            if (++iter > 200) { free(e); return 1; }
            double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
            double r = hypot(g, 1.0);
            g = d[mm] - d[l] + e[l] / (g + (g >= 0 ? fabs(r) : -fabs(r)));
            double s = 1.0, c = 1.0, p = 0.0;
            int i;
            for (i = mm - 1; i >= l; --i) {
                double f = s * e[i], b = c * e[i];
                r = hypot(f, g);
                e[i + 1] = r;
                if (r == 0.0) { d[i + 1] -= p; e[mm] = 0.0; break; }
                s = f / r; c = g / r;
                g = d[i + 1] - p;
                r = (d[i] - g) * s + 2.0 * c * b;
                p = s * r;
                d[i + 1] = g + p;
                g = c * r - b;
                if (z) {
                    for (int k = 0; k < m; ++k) {
                        double zf = z[(i + 1) * m + k];
                        z[(i + 1) * m + k] = s * z[i * m + k] + c * zf;
                        z[i * m + k] = c * z[i * m + k] - s * zf;
                    }
                }
            }
            if (r == 0.0 && i >= l) continue;
            d[l] -= p; e[l] = g; e[mm] = 0.0;
        }
    }
    /* sort ascending */
    for (int i = 0; i < m - 1; ++i) {
        int kmin = i;
        for (int j = i + 1; j < m; ++j) if (d[j] < d[kmin]) kmin = j;
        if (kmin != i) {
            double t = d[i]; d[i] = d[kmin]; d[kmin] = t;
            if (z) for (int k = 0; k < m; ++k) { t = z[i * m + k]; z[i * m + k] = z[kmin * m + k]; z[kmin * m + k] = t; }
        }
    }
    free(e);
    return 0;
}

/* ---------------------------------------------------------------------------
 * Lanczos: LANCZOS (lanczos.f90:1-163), plain three-term recurrence without
 * reorthogonalisation, NADD = 5 steps between tridiagonal solves, the same
 * convergence / NDEG / norm-lost logic.  The start vector is supplied by the
 * caller (the reference draws RAND()-0.5, :22-25).  All Krylov vectors are kept
 * in RAM (the reference writes them to HDF5, :49) for the ground-vector
 * reconstruction (LANCZOS_NORM :137-161, READ_PSI_HUBBARD compute_solver.f90:282).
 * model: 0 = Hubbard, 1 = Heisenberg; the faithful HPROD of the model is used.
 * Outputs: alpha[maxlcz], beta[maxlcz] (beta[0] = start norm, as BETA(1) :28),
 * eval[*nblcz] Ritz values, psi[nst] ground vector; return value nblcz.
 * ------------------------------------------------------------------------- */
typedef struct {
    int model, n, nup, ndn;
    long long nsup, nsdn;
    const u64 *bup, *bdn;
    const double *T, *U, *Jm;
    int fast;          /* 1: use the rank-table H*v instead of the faithful one */
    int nthreads;
} orc_model;

static void hprod_model(const orc_model *m, const double *v, double *w) {
    if (m->model == 0) {
        if (m->fast)
            orc_hprod_hubbard_fast(m->n, m->nup, m->ndn, m->nsup, m->nsdn, m->bup, m->bdn,
                                   m->T, m->U, v, w, m->nthreads);
        else
            orc_hprod_hubbard(m->n, m->nup, m->ndn, m->nsup, m->nsdn, m->bup, m->bdn,
                              m->T, m->U, v, w);
    } else {
        orc_hprod_heis(m->n, m->nsup, m->bup, m->Jm, v, w);
    }
}

int orc_lanczos(int model, int n, int nup, int ndn, long long nsup, long long nsdn,
                const u64 *bup, const u64 *bdn, const double *T, const double *U,
                const double *Jm, int fast, int nthreads,
                int maxlcz, double acc_lcz, const double *v_start,
                double *alpha, double *beta, double *eval, double *psi,
                int *ndeg_out, int *norm_lost_out) {
    orc_model md = {model, n, nup, ndn, nsup, nsdn, bup, bdn, T, U, Jm, fast, nthreads};
    long long nst = (model == 0) ? nsup * nsdn : nsup;
    const int NADD = 5;
    double *V0 = (double *)calloc((size_t)nst, sizeof(double));
    double *V = (double *)malloc(sizeof(double) * (size_t)nst);
    double *V2 = (double *)malloc(sizeof(double) * (size_t)nst);
    double *basis = (double *)malloc(sizeof(double) * (size_t)nst * (size_t)maxlcz);
    double nr = 0;
    for (long long i = 0; i < nst; ++i) { V[i] = v_start[i]; nr += V[i] * V[i]; }
    nr = sqrt(nr);
    for (long long i = 0; i < nst; ++i) V[i] = (1.0 / nr) * V[i];
    beta[0] = nr;
    int cv = 0, nblcz = 0, ndeg = 0, norm_lost = 0;
    double ev = 0.0;
    double *val = (double *)malloc(sizeof(double) * (size_t)maxlcz);
    double *vec = (double *)malloc(sizeof(double) * (size_t)maxlcz * (size_t)maxlcz);
    double *off = (double *)malloc(sizeof(double) * (size_t)maxlcz);
    while (!cv && nblcz <= maxlcz - 10) {
        for (int it = 0; it < NADD; ++it) {
            memcpy(basis + (size_t)nblcz * nst, V, sizeof(double) * (size_t)nst);
            nblcz++;
            hprod_model(&md, V, V2);
            double a = 0;
            for (long long i = 0; i < nst; ++i) a += V[i] * V2[i];
            alpha[nblcz - 1] = a;
            for (long long i = 0; i < nst; ++i)
                V2[i] = V2[i] - a * V[i] - beta[nblcz - 1] * V0[i];
            nr = 0;
            for (long long i = 0; i < nst; ++i) nr += V2[i] * V2[i];
            nr = sqrt(nr);
            beta[nblcz] = nr;
            memcpy(V0, V, sizeof(double) * (size_t)nst);
            for (long long i = 0; i < nst; ++i) V[i] = (1.0 / nr) * V2[i];
        }
        for (int i = 0; i < nblcz; ++i) val[i] = alpha[i];
        for (int i = 0; i + 1 < nblcz; ++i) off[i] = beta[i + 1];
        orc_tridiag_eig(nblcz, val, off, vec);
        /* LANCZOS_NORM: |V * VEC(:,1)| */
        double nrm = 0;
        for (long long i = 0; i < nst; ++i) {
            double s = 0;
            for (int k = 0; k < nblcz; ++k) s += basis[(size_t)k * nst + i] * vec[k];
            nrm += s * s;
        }
        nrm = sqrt(nrm);
        if (fabs(1 - nrm) > 1.e-5) {
            cv = 1; norm_lost = 1;
        } else if (fabs(val[ndeg] - ev) < acc_lcz) {
            if (fabs(val[ndeg] - val[0]) > acc_lcz && (ndeg + 1) > 1) cv = 1;
            else ndeg = ndeg + 1;
        } else {
            if (nblcz <= maxlcz - 10) ev = val[ndeg];
        }
    }
    for (int i = 0; i < nblcz; ++i) eval[i] = val[i];
    /* psi = V * VEC(:,1) / |.| (compute_solver.f90:282-284) */
    double nrm = 0;
    for (long long i = 0; i < nst; ++i) {
        double s = 0;
        for (int k = 0; k < nblcz; ++k) s += basis[(size_t)k * nst + i] * vec[k];
        psi[i] = s;
        nrm += s * s;
    }
    nrm = sqrt(nrm);
    for (long long i = 0; i < nst; ++i) psi[i] /= nrm;
    if (ndeg_out) *ndeg_out = ndeg;
    if (norm_lost_out) *norm_lost_out = norm_lost;
    free(V0); free(V); free(V2); free(basis); free(val); free(vec); free(off);
    return nblcz;
}

/* ---------------------------------------------------------------------------
 * Static reduced quantities (static_rq.f90:153-395).
 * dens[2*n*n]: dens[s*n*n + (J-1)*n + (K-1)] = DENS_MAT(J,K,s+1)
 * ------------------------------------------------------------------------- */
void orc_one_rdm(int n, int nup, int ndn, long long nsup, long long nsdn,
                 const u64 *bup, const u64 *bdn, const double *bra, const double *ket,
                 double *dens) {
#define DM(J, K, s) dens[(s) * n * n + ((J) - 1) * n + ((K) - 1)]
    for (int J = 1; J <= n; ++J) {
        for (int K = 1; K <= J; ++K) {
            DM(J, K, 0) = 0.0;
            for (long long i = 0; i < nsup; ++i) {
                long long nw = kinop(n, bup[i], J, K);
                if (nw > 0) {
                    long long pos = lsearch(bup, nsup, (u64)nw);
                    if (pos >= 0) {
                        double ac = anticom(n, J, bup[i]);
                        ac = ac * anticom(n, K, (u64)nw);
                        for (long long l1 = 0; l1 < nsdn; ++l1)
                            DM(J, K, 0) = DM(J, K, 0) + ac * bra[l1 * nsup + i] * ket[l1 * nsup + pos];
                    }
                }
            }
            DM(K, J, 0) = DM(J, K, 0);
            if (nup != ndn) {                                   /* :187-209 */
                DM(J, K, 1) = 0.0;
                for (long long l1 = 0; l1 < nsdn; ++l1) {
                    long long nw = kinop(n, bdn[l1], J, K);
                    if (nw > 0) {
                        long long pos = lsearch(bdn, nsdn, (u64)nw);
                        if (pos >= 0) {
                            double ac = anticom(n, J, bdn[l1]);
                            ac = ac * anticom(n, K, (u64)nw);
                            for (long long i = 0; i < nsup; ++i)
                                DM(J, K, 1) = DM(J, K, 1) + ac * bra[l1 * nsup + i] * ket[pos * nsup + i];
                        }
                    }
                }
                DM(K, J, 1) = DM(J, K, 1);
            } else {                                            /* :238-239 */
                DM(J, K, 1) = DM(J, K, 0);
                DM(K, J, 1) = DM(J, K, 0);
            }
        }
    }
#undef DM
}

/* NI = double occupancy (static_rq.f90:307-316) */
void orc_ni(int n, long long nsup, long long nsdn, const u64 *bup, const u64 *bdn,
            const double *bra, const double *ket, double *ni) {
    for (int J = 1; J <= n; ++J) {
        ni[J - 1] = 0.0;
        for (long long u = 0; u < nsup; ++u)
            if (is_part(n, J, bup[u]))
                for (long long d = 0; d < nsdn; ++d)
                    if (is_part(n, J, bdn[d]))
                        ni[J - 1] = ni[J - 1] + bra[d * nsup + u] * ket[d * nsup + u];
    }
}

/* SMATRIX = <S_J . S_K>, Hubbard (static_rq.f90:323-395) */
void orc_j_spin(int n, long long nsup, long long nsdn, const u64 *bup, const u64 *bdn,
                const double *bra, const double *ket, double *smat) {
#define SM(J, K) smat[((J) - 1) * n + ((K) - 1)]
    for (int J = 1; J <= n; ++J) {
        for (int K = 1; K <= J; ++K) {
            SM(J, K) = 0.0;
            if (J == K) {
                for (long long i = 0; i < nsup; ++i)
                    for (long long l1 = 0; l1 < nsdn; ++l1) {
                        if (is_part(n, J, bup[i]) && !is_part(n, J, bdn[l1]))
                            SM(J, J) = SM(J, J) + 0.5 * bra[l1 * nsup + i] * ket[l1 * nsup + i];
                        if (!is_part(n, J, bup[i]) && is_part(n, J, bdn[l1]))
                            SM(J, J) = SM(J, J) + 0.5 * bra[l1 * nsup + i] * ket[l1 * nsup + i];
                    }
            } else {
                for (long long i = 0; i < nsup; ++i) {
                    long long nw = kinop(n, bup[i], J, K);
                    if (nw > 0) {
                        long long pos = lsearch(bup, nsup, (u64)nw);
                        if (pos >= 0) {
                            double ac = anticom(n, J, bup[i]);
                            ac = ac * anticom(n, K, (u64)nw);
                            for (long long l1 = 0; l1 < nsdn; ++l1) {
                                long long nw2 = kinop(n, bdn[l1], K, J);
                                if (nw2 > 0) {
                                    long long pos2 = lsearch(bdn, nsdn, (u64)nw2);
                                    if (pos2 >= 0) {
                                        double ac2 = anticom(n, K, bdn[l1]);
                                        ac2 = ac2 * anticom(n, J, (u64)nw2);
                                        SM(J, K) = SM(J, K) - ac * ac2 * bra[l1 * nsup + i] * ket[pos2 * nsup + pos];
                                    }
                                }
                            }
                        }
                    }
                }
            }
            for (long long i = 0; i < nsup; ++i)
                for (long long l1 = 0; l1 < nsdn; ++l1)
                    SM(J, K) = SM(J, K) + bra[l1 * nsup + i] * ket[l1 * nsup + i] *
                               szf(n, J, bup[i], bdn[l1]) * szf(n, K, bup[i], bdn[l1]);
            SM(K, J) = SM(J, K);
        }
    }
#undef SM
}

/* SMATRIX, Heisenberg (static_rq_heis.f90:58-116) */
void orc_j_spin_heis(int n, long long nsup, const u64 *bup, const double *bra,
                     const double *ket, double *smat) {
    u64 full = (n >= 64) ? ~(u64)0 : (((u64)1 << n) - 1);
#define SM(J, K) smat[((J) - 1) * n + ((K) - 1)]
    for (int J = 1; J <= n; ++J) {
        for (int K = 1; K <= n; ++K) {
            SM(J, K) = 0.0;
            if (J == K) {
                for (long long i = 0; i < nsup; ++i) {
                    u64 bd = full - bup[i];
                    if (is_part(n, J, bup[i]) && !is_part(n, J, bd))
                        SM(J, J) = SM(J, J) + 0.5 * bra[i] * ket[i];
                    else if (!is_part(n, J, bup[i]) && is_part(n, J, bd))
                        SM(J, J) = SM(J, J) + 0.5 * bra[i] * ket[i];
                }
            } else {
                for (long long i = 0; i < nsup; ++i) {
                    long long nw = kinop(n, bup[i], J, K);
                    if (nw > 0) {
                        long long pos = lsearch(bup, nsup, (u64)nw);
                        if (pos >= 0) {
                            u64 bd = full - bup[i];
                            double ac = anticom(n, J, bup[i]);
                            ac = ac * anticom(n, K, (u64)nw);
                            long long nw2 = kinop(n, bd, K, J);
                            if ((u64)nw2 == full - bup[pos]) {
                                ac = ac * anticom(n, K, bd);
                                ac = ac * anticom(n, J, (u64)nw2);
                                SM(J, K) = SM(J, K) - ac * bra[pos] * ket[i];
                            }
                        }
                    }
                }
            }
            for (long long i = 0; i < nsup; ++i) {
                u64 bd = full - bup[i];
                SM(J, K) = SM(J, K) + bra[i] * ket[i] * szf(n, J, bup[i], bd) * szf(n, K, bup[i], bd);
            }
        }
    }
#undef SM
}

/* ---------------------------------------------------------------------------
 * Green's-function excitation vectors (spgf_ed.f90:304-429).
 * create = 1: c^dag_site (GF_E_*), 0: c_site (GF_H_*); spin 0 = up, 1 = down.
 * Source sector (nup_, ndn_) with bases b*_src; target bases b*_dst.
 * out has nsup_dst*nsdn_dst entries (zeroed here).
 * ------------------------------------------------------------------------- */
void orc_excite(int n, int site, int spin, int create, int ndn_src,
                long long nsup_s, long long nsdn_s, const u64 *bup_s, const u64 *bdn_s,
                long long nsup_t, long long nsdn_t, const u64 *bup_t, const u64 *bdn_t,
                const double *psi, double *out) {
    for (long long i = 0; i < nsup_t * nsdn_t; ++i) out[i] = 0.0;
    if (spin == 0) {
        for (long long i = 0; i < nsup_s; ++i) {
            int occ = is_part(n, site, bup_s[i]);
            if (create ? occ : !occ) continue;
            u64 a = create ? bup_s[i] + sitebit(n, site) : bup_s[i] - sitebit(n, site);
            for (long long j = 0; j < nsup_t; ++j) {
                if (bup_t[j] == a) {
                    double ac = anticom(n, site, bup_s[i]);
                    if (ndn_src % 2 == 1) ac = -ac;               /* :321, :386 */
                    for (long long k = 0; k < nsdn_t; ++k)
                        out[k * nsup_t + j] = ac * psi[k * nsup_s + i];
                }
            }
        }
    } else {
        for (long long i = 0; i < nsdn_s; ++i) {
            int occ = is_part(n, site, bdn_s[i]);
            if (create ? occ : !occ) continue;
            u64 a = create ? bdn_s[i] + sitebit(n, site) : bdn_s[i] - sitebit(n, site);
            for (long long j = 0; j < nsdn_t; ++j) {
                if (a == bdn_t[j]) {
                    double ac = anticom(n, site, bdn_s[i]);
                    for (long long k = 0; k < nsup_t; ++k)
                        out[j * nsup_t + k] = ac * psi[i * nsup_s + k];
                }
            }
        }
    }
}

/* ---------------------------------------------------------------------------
 * Continued-fraction chain for a diagonal Green's function: the single-vector
 * case of GF_BANDLANCZOS (spgf_bandlanczos.f90:309-448, NORB2COMP = 1) run for
 * a fixed number of steps from `start`.  norm2 = <start|start>; alpha[nsteps],
 * beta[nsteps] (beta[k] couples k and k+1).  FP64, no reorthogonalisation.
 * ------------------------------------------------------------------------- */
void orc_cf_chain(int model, int n, int nup, int ndn, long long nsup, long long nsdn,
                  const u64 *bup, const u64 *bdn, const double *T, const double *U,
                  const double *Jm, int fast, int nthreads, const double *start, int nsteps,
                  double *alpha, double *beta, double *norm2) {
    orc_model md = {model, n, nup, ndn, nsup, nsdn, bup, bdn, T, U, Jm, fast, nthreads};
    long long nst = (model == 0) ? nsup * nsdn : nsup;
    double *V0 = (double *)calloc((size_t)nst, sizeof(double));
    double *V = (double *)malloc(sizeof(double) * (size_t)nst);
    double *V2 = (double *)malloc(sizeof(double) * (size_t)nst);
    double nr = 0;
    for (long long i = 0; i < nst; ++i) nr += start[i] * start[i];
    *norm2 = nr;
    nr = sqrt(nr);
    for (long long i = 0; i < nst; ++i) V[i] = start[i] / nr;
    double bprev = 0.0;
    for (int k = 0; k < nsteps; ++k) {
        hprod_model(&md, V, V2);
        double a = 0;
        for (long long i = 0; i < nst; ++i) a += V[i] * V2[i];
        alpha[k] = a;
        double b2 = 0;
        for (long long i = 0; i < nst; ++i) {
            V2[i] = V2[i] - a * V[i] - bprev * V0[i];
            b2 += V2[i] * V2[i];
        }
        double b = sqrt(b2);
        beta[k] = b;
        if (b == 0.0) { for (int q = k + 1; q < nsteps; ++q) { alpha[q] = 0; beta[q] = 0; } break; }
        memcpy(V0, V, sizeof(double) * (size_t)nst);
        for (long long i = 0; i < nst; ++i) V[i] = V2[i] / b;
        bprev = b;
    }
    free(V0); free(V); free(V2);
}

/* ---------------------------------------------------------------------------
 * One column of the Heisenberg HPROD matrix for clusters far beyond a dense or
 * even a full-vector restatement (n = 36: 9.1e9 states, indices past 2^32).
 * orc_unrank_ref inverts the closed-form reference rank (the order of
 * compute_basis.f90:67-167): the colex index C(n,k)-1-idx of rev_n(x), taken
 * greedily from the highest bit.  orc_heis_column gives H e_idx as (row, value)
 * pairs: for every ordered pair (J,K) the same KINOP / ANTICOM factors as
 * hprod_heis.f90:22-53 (sign_convention 0: the reference's -ac*J(J,K); 1: +ac*J(J,K),
 * the element of the physical spin model), the J=J term of :11-21 and the Ising
 * products of :54-58 on the diagonal.  Rows are unsorted; a row can repeat when
 * two pairs reach it.  Returns the number of pairs written (at most n*n+1).
 * ------------------------------------------------------------------------- */
u64 orc_unrank_ref(int n, int k, u64 idx) {
    u64 c = orc_binom(n, k) - 1 - idx, y = 0;
    int p = n - 1;
    for (int i = k; i >= 1; --i) {
        while (p >= 0 && orc_binom(p, i) > c) --p;
        y |= (u64)1 << p;
        c -= orc_binom(p, i);
        --p;
    }
    u64 x = 0;                                   /* x = rev_n(y) */
    for (int b = 0; b < n; ++b)
        if (y >> b & 1) x |= (u64)1 << (n - 1 - b);
    return x;
}

long long orc_heis_column(int n, const double *Jm, int sign_convention, u64 idx,
                          u64 *rows, double *vals) {
    const int k = n / 2;
    const u64 full = (n >= 64) ? ~(u64)0 : (((u64)1 << n) - 1);
    const u64 x = orc_unrank_ref(n, k, idx), bd = full - x;
    long long cnt = 0;
    double dg = 0.0;
    for (int J = 1; J <= n; ++J) {
        double jjj = Jm[(J - 1) * n + (J - 1)];
        if (fabs(jjj) > 1.e-14)
            if ((is_part(n, J, x) && !is_part(n, J, bd)) || (!is_part(n, J, x) && is_part(n, J, bd)))
                dg += 0.5 * jjj;
        for (int K = 1; K <= n; ++K) {
            double jjk = Jm[(J - 1) * n + (K - 1)];
            if (fabs(jjk) > 1.e-14 && J != K) {
                long long nw = kinop(n, x, J, K);
                if (nw > 0) {
                    double ac = anticom(n, J, x);
                    ac = ac * anticom(n, K, (u64)nw);
                    long long nw2 = kinop(n, bd, K, J);
                    if ((u64)nw2 == full - (u64)nw) {
                        ac = ac * anticom(n, J, bd);
                        ac = ac * anticom(n, K, (u64)nw2);
                        rows[cnt] = (u64)colex_rank_ref(n, k, (u64)nw);
                        vals[cnt] = sign_convention ? ac * jjk : -ac * jjk;
                        ++cnt;
                    }
                }
            }
            dg += szf(n, J, x, bd) * szf(n, K, x, bd) * jjk;
        }
    }
    rows[cnt] = idx;
    vals[cnt] = dg;
    return cnt + 1;
}
