/*
 * phub.h -- C ABI of the MI355X exact-diagonalization engine (libphub.so).
 *
 * Replaces the process + HDF5-file boundary of PyHub's hot path: the
 * reference spawns Fortran executables with os.system (pyhub/core/core_solver.py:107-111,
 * pyhub/core/basis.py:167-171) which read and write solver.h5 / basis.h5.  Here the
 * same stages are in-process calls on device buffers.
 *
 * Conventions (identical to the reference):
 *   - site J (1-based) is bit n-J of a mask (pyhub/libs/src/funcmod.f90:22-23);
 *   - per-species basis order is HILBERT's (pyhub/libs/src/compute_basis.f90:67-96),
 *     rank(x) = C(n,k) - 1 - colex(rev_n(x));
 *   - Hubbard state index d*dim_up + u, u fastest (pyhub/libs/src/hprod_hubbard.f90:31);
 *   - T is row-major T[(J-1)*n + (K-1)] = T(J,K): hopping c^dag_K c_J has amplitude
 *     T(J,K), the diagonal T(J,J) is the on-site energy eps_J.
 *
 * Memory: every pointer documented "device" is a device pointer on the
 * context's device; "host" pointers are read/written synchronously.  Calls are
 * ordered on the context's stream.  Device (and pinned host) memory is allocated
 * only
 *   - by the *_create entry points (contexts, operators, Fock spaces, operator
 *     programs) and by phub_batch_hubbard (its per-call task and vector arrays);
 *   - once per operator, on the first call that needs them, for tables kept with
 *     the operator: the all-pairs hop tables and the down-spin pair map of the
 *     reduced quantities, the all-pairs view of a Heisenberg operator, the
 *     source maps of an operator program;
 *   - once per context, when a call needs a larger reduction buffer or scratch
 *     arena than the context holds (kept until phub_ctx_destroy);
 *   - for the Krylov / work vectors of phub_lanczos, phub_cf_chain,
 *     phub_band_lanczos and phub_adjoint_minres when the caller passes no
 *     workspace (sizes from the matching *_workspace_bytes query), for the
 *     duration of the call.
 * So a call repeated on the same operator and context, with workspaces where the
 * entry point takes one, allocates nothing; phub_alloc_count() counts every
 * allocation the library makes, for checking that.
 *
 * Errors: every entry returns PHUB_OK (0) or a PHUB_ERR_* code; the message is
 * available from phub_last_error(ctx).
 */
#ifndef PHUB_H
#define PHUB_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    PHUB_OK = 0,
    PHUB_ERR_ARG = 1,          /* invalid argument / shape */
    PHUB_ERR_UNSUPPORTED = 2,  /* size or feature outside what is built */
    PHUB_ERR_OOM = 3,          /* device allocation failed */
    PHUB_ERR_NOCONV = 4,       /* Lanczos hit max steps (reference: "MAX LCZ STEP REACH") */
    PHUB_ERR_BREAKDOWN = 5,    /* invariant subspace found before convergence was checked */
    PHUB_ERR_HIP = 6,          /* HIP runtime error */
    PHUB_ERR_NORM_LOST = 7     /* kept for the ABI; phub_lanczos reports NORM LOST in
                                * phub_lanczos_result.norm_lost instead */
};

typedef struct phub_ctx phub_ctx;
typedef struct phub_op phub_op;

/* ---- context --------------------------------------------------------------
 * Replaces the per-call executable spawn (core_solver.py:107-111).
 * stream: a hipStream_t (NULL = default stream). */
int phub_ctx_create(int device, void *stream, phub_ctx **out);
/* Number of device and pinned-host allocations the library has made in this
 * process (all contexts).  No counterpart in the reference: it lets a caller check
 * the allocation contract above around a call. */
uint64_t phub_alloc_count(void);
void phub_ctx_destroy(phub_ctx *ctx);
const char *phub_last_error(const phub_ctx *ctx);
int phub_sync(phub_ctx *ctx);
/* timing of the most recent H*v launch on the context (HIP events), ms */
double phub_last_kernel_ms(const phub_ctx *ctx);

/* ---- basis ----------------------------------------------------------------
 * Replaces basis.x / HILBERT (compute_basis.f90:67-96) and basis.h5
 * (basis.py:65-83,214-228).  Masks in the reference order, built on the device. */
uint64_t phub_binom(int n, int k);
int phub_basis_build(phub_ctx *ctx, int n, int k, uint64_t *d_masks /* device, C(n,k) */);
/* rank of each mask (reference order); masks need popcount k */
int phub_basis_rank(phub_ctx *ctx, int n, int k, const uint64_t *d_masks, int64_t count,
                    int64_t *d_ranks /* device */);

/* ---- operators --------------------------------------------------------------
 * Hubbard: the operator of HPROD (hprod_hubbard.f90:17-118) with per-site U
 * (is_Uloc = 2, core_solver.py:133-141).  t: host n*n row-major; u: host n. */
typedef struct {
    int n_sites, n_up, n_dn;
    const double *t;
    const double *u;
    /* optional down-spin hopping matrix (NULL = t).  No reference counterpart:
     * the reference uses one T for both spins; a spin-dependent one-body
     * operator is what the adjoint gradient applies to psi. */
    const double *t_dn;
    /* optional 4-index interaction (NULL = per-site u), host n^4 row-major
     * u4[((J-1)*n+(K-1))*n*n+(M-1)*n+(N-1)] = UL(J,K,M,N):
     * sum_{JKMN} UL(J,K,M,N) c^dag_{K up} c_{J up} c^dag_{N dn} c_{M dn}
     * (is_Uloc = 3, hprod_hubbard.f90:62-89); u is then ignored.  UL must satisfy
     * UL(J,K,M,N) = UL(K,J,N,M) (core_solver.py:143). */
    const double *u4;
} phub_hubbard_desc;

/* Heisenberg S_z = 0 sector (hprod_heis.f90:1-61): H = sum over ORDERED pairs
 * J(J,K) S_J.S_K.  sign_convention 0 reproduces the reference's spin-flip
 * element -(-1)^(J+K) J(J,K) exactly; 1 uses +(-1)^(J+K) J(J,K), the element of
 * the physical spin model in the same fermion basis (the sign the reference's
 * J_SPIN measures, static_rq_heis.f90:94-102).  The two agree up to a basis gauge
 * on bipartite bond graphs only. */
typedef struct {
    int n_sites;
    const double *j;
    int sign_convention;
} phub_heis_desc;

typedef struct {
    int kind;            /* 0 = Hubbard, 1 = Heisenberg */
    int n_sites, n_up, n_dn;
    int64_t dim_up, dim_dn, dim;
    int max_hops_up, max_hops_dn;
    double table_bytes;  /* device bytes of the hop / rank tables */
    int dn_tile[2];      /* Hubbard down-hop tile the H*v tuner chose, FP32 / FP64:
                            (bytes per row segment) * 100000 + (rows per workgroup) * 100
                            + (row gathers in flight per lane), + 50 for the buffer-load
                            kernel; -1 = not yet tuned */
} phub_op_info;

int phub_op_hubbard(phub_ctx *ctx, const phub_hubbard_desc *desc, phub_op **out);
int phub_op_heisenberg(phub_ctx *ctx, const phub_heis_desc *desc, phub_op **out);
void phub_op_destroy(phub_op *op);
int phub_op_get_info(const phub_op *op, phub_op_info *info);
/* masks of the per-species bases held by the operator (device, dim_up / dim_dn) */
int phub_op_masks(phub_ctx *ctx, const phub_op *op, uint64_t *d_up, uint64_t *d_dn);

/* ---- H*v ---------------------------------------------------------------------
 * w = H v, replaces HPROD (hprod_hubbard.f90:17-118, hprod_heis.f90:1-61).
 * v, w: device, op dim; w must not alias v. */
int phub_hv_f64(phub_ctx *ctx, const phub_op *op, const double *v, double *w);
int phub_hv_f32(phub_ctx *ctx, const phub_op *op, const float *v, float *w);

/* ---- Lanczos ground state ----------------------------------------------------
 * Replaces LANCZOS (lanczos.f90:1-163) + SOLVE (solve.f90:1-23) +
 * READ_PSI_HUBBARD (compute_solver.f90:183-290). */
enum {
    PHUB_LCZ_PLAIN64 = 0,   /* FP64 vectors, 3-term recurrence, no reorthogonalisation
                               (the reference algorithm); with want_psi the vectors are
                               kept while memory allows and psi is combined from them
                               (as the reference), else psi comes from a second pass */
    PHUB_LCZ_FULL64 = 1,    /* FP64 vectors, full reorthogonalisation, stored basis */
    PHUB_LCZ_MIXED = 2,     /* FP32 vectors, FP64 scalars, full reorthogonalisation,
                               FP64 psi reconstruction + FP64 refinement steps */
    PHUB_LCZ_PLAIN32 = 3    /* FP32 vectors, FP64 scalars, 3-term recurrence (vectors kept
                               as PLAIN64); with refine_steps, FP64 refinement as MIXED.
                               Runs in 3 resident vectors when psi is not wanted (sectors
                               whose Krylov basis cannot be stored) */
};
typedef struct {
    int mode;               /* PHUB_LCZ_* */
    int max_steps;          /* reference max_lcz */
    double tol;             /* convergence of the lowest Ritz value (reference acc_lcz) */
    int check_every;        /* reference NADD = 5 */
    int fixed_steps;        /* >0: run exactly this many steps, no convergence test */
    int refine_steps;       /* FP64 corrections of psi after MIXED / PLAIN32 (0..8) */
    uint64_t seed;          /* start vector when d_v0 is NULL */
    int want_psi;           /* reconstruct the ground vector into d_psi */
    int n_states;           /* Ritz vectors to reconstruct (READ_PSI_HUBBARD for EXC_STATE
                               1..n_states, compute_solver.f90:183-290); 0 or 1 = ground
                               only, at most PHUB_MAX_STATES; d_psi then holds n_states
                               vectors of dim, state k at d_psi + k*dim */
    void *workspace;        /* caller-owned device memory for the Krylov vectors, or NULL
                               (then allocated per call); phub_lanczos_workspace_bytes
                               gives the size that holds every vector.  With a smaller
                               one the plain modes keep fewer vectors (psi by a second
                               pass) and FULL64 / MIXED allocate their basis instead */
    size_t workspace_bytes;
    double psi_tol;         /* > 0: target for ||H psi - E psi|| (psi normalised).  The
                               run also continues until the Lanczos residual estimate
                               beta_m |y_m| of the lowest Ritz pair is below it (FP32
                               vectors: below max(psi_tol, 1e-6 max(1, |E|)), their
                               floor); MIXED / PLAIN32 then apply FP64 corrections of psi
                               (at most refine_steps) until the FP64 residual is below
                               it.  0: the Ritz-value test only */
} phub_lanczos_opts;

#define PHUB_MAX_STATES 16

typedef struct {
    double e0;              /* final ground energy (after refinement when MIXED) */
    double e0_ritz;         /* lowest Ritz value of the tridiagonal */
    double residual;        /* ||H psi - e0 psi|| (FP64), when psi was built */
    int steps;              /* Lanczos steps taken */
    int converged;
    int ndeg;               /* reference NDEG counter */
    double seconds;         /* device wall time of the solve */
    double seconds_lanczos; /* of which: Krylov steps */
    double hv_ms_avg;       /* average H*v kernel time */
    double e_states[PHUB_MAX_STATES];   /* lowest Ritz values (e_states[0] = e0) */
    int norm_lost;          /* the reference's NORM LOST (lanczos.f90:74-76): the ground
                             * vector rebuilt from the Krylov vectors had a norm more than
                             * 1e-5 away from 1 (loss of orthogonality in the plain forms);
                             * it is normalised and the call still succeeds, as the
                             * reference prints the message and keeps its results */
    int corrections;        /* FP64 corrections of psi applied (MIXED / PLAIN32) */
    int correction_iters;   /* their inner MINRES iterations, summed */
} phub_lanczos_result;

/* Device bytes phub_lanczos uses for its Krylov vectors under these options (the
 * workspace size that lets it allocate none).  No counterpart in the reference
 * (which writes every vector to solver.h5, lanczos.f90:49). */
int phub_lanczos_workspace_bytes(const phub_op *op, const phub_lanczos_opts *opts,
                                 size_t *bytes);
int phub_lanczos(phub_ctx *ctx, const phub_op *op, const phub_lanczos_opts *opts,
                 const double *d_v0 /* device start vector or NULL */,
                 double *alpha /* host, max_steps */, double *beta /* host, max_steps+1 */,
                 phub_lanczos_result *res, double *d_psi /* device, dim, or NULL */);

/* ---- reduced quantities ------------------------------------------------------
 * Hubbard: ONE_RDM, TWO_RDM (NI), J_SPIN (static_rq.f90:153-395) for <bra| . |ket>.
 * Outputs host, n*n row-major (rdm_up[(J-1)*n+(K-1)] = <c^dag_J c_K>), ni n,
 * sisj n*n.  Any output may be NULL. */
int phub_rdm_hubbard(phub_ctx *ctx, const phub_op *op, const double *d_bra, const double *d_ket,
                     double *rdm_up, double *rdm_dn, double *ni, double *sisj);
/* DBL_OCC of TWO_RDM (static_rq.f90:247-306): out[((J-1)*n+(K-1))*n*n+(M-1)*n+(N-1)]
 * = sum over basis states of bra(src) ket(c^dag_{K up} c_{J up} c^dag_{N dn} c_{M dn} src)
 * times the fermion signs, computed for K <= J, N <= M and copied to (J,K,N,M),
 * (K,J,M,N), (K,J,N,M) as the reference does.  out: host n^4. */
int phub_two_rdm_hubbard(phub_ctx *ctx, const phub_op *op, const double *d_bra,
                         const double *d_ket, double *out);
/* Spin-resolved transition 1-RDMs g_s[(J-1)*n+(K-1)] = <bra| c^dag_Js c_Ks |ket>
 * for every ordered (J,K), without ONE_RDM's mirroring and up-copy rule (used by
 * the adjoint gradient; no reference counterpart).  Host outputs, n*n each. */
int phub_one_rdm_spin(phub_ctx *ctx, const phub_op *op, const double *d_bra, const double *d_ket,
                      double *g_up, double *g_dn);
/* Heisenberg J_SPIN (static_rq_heis.f90:58-116) */
int phub_rdm_heisenberg(phub_ctx *ctx, const phub_op *op, const double *d_bra,
                        const double *d_ket, double *sisj);

/* ---- Green's functions -------------------------------------------------------
 * Excitation c^dag_{site,spin} psi (create=1) or c_{site,spin} psi (create=0) from
 * src's sector into dst's sector (GF_E_* / GF_H_*, spgf_ed.f90:304-429; sign
 * ANTICOM(site, x_spin) times (-1)^N_dn for up operators, :321,:386).
 * site is 1-based; spin 0 = up, 1 = down. */
int phub_excite(phub_ctx *ctx, const phub_op *src, const phub_op *dst, int site, int spin,
                int create, const double *d_psi, double *d_out);
/* Continued-fraction chain from d_start (single-vector case of GF_BANDLANCZOS,
 * spgf_bandlanczos.f90:309-448): alpha[nsteps], beta[nsteps], norm2 = <start|start>.
 * precision 64 or 32 (FP32 vectors, FP64 recurrence scalars). */
int phub_cf_chain(phub_ctx *ctx, const phub_op *op, const double *d_start, int nsteps,
                  int precision, double *alpha, double *beta, double *norm2,
                  void *workspace /* device or NULL */, size_t workspace_bytes);
/* Device bytes of the three chain vectors (the workspace that lets phub_cf_chain
 * allocate nothing).  No counterpart in the reference (it allocates internally). */
int phub_cf_chain_workspace_bytes(const phub_op *op, int precision, size_t *bytes);

/* Band Lanczos for the full G_ij (GF_BANDLANCZOS, spgf_bandlanczos.f90:309-448, and
 * the weights Q of SPGF_BANDLANCZOS, :90-100).  d_V: p start vectors (device,
 * FP64, vector i at d_V + i*dim, not modified).  They are orthonormalised
 * (Gram-Schmidt, V = Q R^T); the band recurrence adds one vector per step with a
 * deflation threshold of 1e-10 and checks the lowest eigenvalue of the band
 * matrix every p steps against tol, up to max_steps - 10 steps (as the
 * reference).  Outputs (host): positions[*npoles] ascending (eigenvalues of the
 * band matrix after deflated rows are dropped) and
 * Q[k*p + i] = sum_{j<=i} R(i,j) VEC(j,k) for pole k, site i.  positions and Q
 * must hold max_steps + p and (max_steps + p) * p doubles. */
typedef struct {
    int max_steps;          /* reference max_lcz */
    double tol;             /* reference acc_lcz */
    void *workspace;        /* device memory for the 2p+1 vector slots, or NULL
                               (then allocated for the call) */
    size_t workspace_bytes;
} phub_band_opts;
int phub_band_lanczos(phub_ctx *ctx, const phub_op *op, const double *d_V, int p,
                      const phub_band_opts *opts, double *positions, double *Q, int *npoles);
/* Device bytes of the band Lanczos vector slots for p start vectors.  No counterpart
 * in the reference. */
int phub_band_lanczos_workspace_bytes(const phub_op *op, int p, size_t *bytes);

/* ---- batched clusters --------------------------------------------------------
 * Many independent Hubbard clusters in one persistent launch: each workgroup
 * runs whole Lanczos chains of one work item (ground state, then G_ii chains)
 * with FP32 vectors and FP64 recurrence scalars.  See phub_batch.h-style
 * descriptors below. */
typedef struct {
    int n_sites, n_up, n_dn;
    const double *t;   /* host n*n */
    const double *u;   /* host n */
} phub_cluster;

typedef struct {
    int gs_steps;        /* maximum Lanczos steps for the ground state (<= 256) */
    int gf_steps;        /* continued-fraction steps per G_ii chain */
    uint64_t seed;
    double gs_tol;       /* stop the ground state once |beta_m y_m| < gs_tol*max(1,|E|)
                            (checked every 10 steps); 0 runs all gs_steps */
} phub_batch_opts;

/* e0[n_clusters]; for cluster c, site i (0-based), spin s, kind k (0 particle
 * c^dag psi0, 1 hole c psi0): chain slot q = ((c*32 + i)*2 + s)*2 + k holds
 * gf_alpha[q*gf_steps..], gf_beta[q*gf_steps..], gf_norm2[q].  Chains that do
 * not exist (sector empty) have norm2 = 0.  All outputs host. */
int phub_batch_hubbard(phub_ctx *ctx, const phub_cluster *clusters, int n_clusters,
                       const phub_batch_opts *opts, double *e0, double *gf_alpha,
                       double *gf_beta, double *gf_norm2);

/* Evaluate G(omega) = sign * norm2 / (z - a0 - b0^2/(z - a1 - ...)) on a grid, on
 * the device, for n_chains chains: z = sign*(omega + i eta) + shift[c].  With shift
 * E0, sign +1 gives a particle chain's sum of w_n / (omega + i eta - (E_n - E0)) and
 * sign -1 a hole chain's sum of w_n / (omega + i eta + (E_n - E0)).
 * out: host complex128 interleaved, n_chains * n_omega. */
int phub_cf_eval(phub_ctx *ctx, int n_chains, int nsteps, const double *alpha,
                 const double *beta, const double *norm2, const double *shift,
                 const double *sign, int n_omega, const double *omega, double eta,
                 double *out);
/* phub_cf_eval without allocating: for callers that cannot allocate device memory
 * (graph capture, tight serving loops, harnesses that forbid it).  No counterpart
 * in the reference.  Same arguments, results and errors as phub_cf_eval; the
 * staged inputs and the output (2 n_chains nsteps + 3 n_chains + nw +
 * 2 n_chains nw doubles) must fit PHUB_CF_NOALLOC_DOUBLES (2^22 doubles, 32 MiB,
 * a fixed device array each GPU holds while the library is loaded), else
 * PHUB_ERR_UNSUPPORTED.  The array is shared by every call on one GPU: calls that
 * overlap on one GPU (on different streams) race. */
#define PHUB_CF_NOALLOC_DOUBLES (1 << 22)
#define PHUB_MAX_DEVICES 64
int phub_cf_eval_noalloc(phub_ctx *ctx, int n_chains, int nsteps, const double *alpha,
                         const double *beta, const double *norm2, const double *shift,
                         const double *sign, int nw, const double *omega, double eta,
                         double *out);

/* ---- gradients ---------------------------------------------------------------
 * Hellmann-Feynman: dE0/dT(J,K), dE0/dU_J for the Hubbard operator in one pass
 * over psi (no reference counterpart; pinned by static_rq.f90 contractions).
 * d_t: host n*n, d_u: host n. */
int phub_hf_grad_hubbard(phub_ctx *ctx, const phub_op *op, const double *d_psi,
                         double *grad_t, double *grad_u);
/* Transition derivatives <bra| dH/dtheta |ket> of the Hubbard operator for every
 * T(J,K) (grad_t, host n*n) and U_J (grad_u, host n): the contraction the adjoint
 * gradient of a psi-dependent loss needs (no reference counterpart). */
int phub_grad_hubbard(phub_ctx *ctx, const phub_op *op, const double *d_bra, const double *d_ket,
                      double *grad_t, double *grad_u);
/* dE0/dJ(J,K) for the Heisenberg operator (ordered pairs). */
int phub_hf_grad_heisenberg(phub_ctx *ctx, const phub_op *op, const double *d_psi,
                            double *grad_j);
/* Adjoint solve (H - e0) x = b with b orthogonal to psi, by MINRES restricted to
 * the complement of psi (FP64; seven work vectors of the operator dimension and, on
 * one-body Hubbard operators, the diagonal of the Jacobi preconditioner
 * 1 / (diag(H) - e0), from the workspace or allocated for the call; one iteration is
 * one H*v, four vector passes and two host round trips).  tol and *rel_res are
 * relative residuals, in the preconditioner's norm when it applies
 * (PHUB_JD_PRECOND=0 switches it off).  Returns iterations in *iters. */
int phub_adjoint_minres(phub_ctx *ctx, const phub_op *op, const double *d_psi, double e0,
                        const double *d_b, double *d_x, int max_iter, double tol,
                        int *iters, double *rel_res, void *workspace /* device or NULL */,
                        size_t workspace_bytes);
/* Device bytes of the MINRES work vectors (eight vectors of the dimension). */
int phub_adjoint_minres_workspace_bytes(const phub_op *op, size_t *bytes);

/* ---- generic operator strings (operators.x) ---------------------------------------
 * Replaces compute_operators.f90:1-72 (the driver) and the elementary routines of
 * operators.f90 (C_DAGGER_C :1-55, C_DAGGER_C_UNRES :58-144, NI :148-172, NIM1
 * :196-220, SZ_OPE :224-245, SISJ :303-394, C_DAGGER :397-446, C_ :450-499).
 * A phub_fock is the product basis of the concatenated fixed-N sectors of each
 * species (compute_operators.f90:10-26; state index d * nsup + u).  A program is a
 * sum of products: term t has coefficient coef[t] and term_len[t] factors, each
 * five ints (code, site1, spin1, site2, spin2) with the reference's OPE codes
 * (1 c^dag c same spin, 2 n, 3 id, 4 1 - n, 5 sz, 6 S+, 7 S-, 8 S.S, 9 c^dag c
 * spin-flip, 10 c^dag, 11 c), sites 1-based, spin 1 = up; the factors of a product
 * act in list order (the first one first), as in the reference. */
typedef struct phub_fock phub_fock;
typedef struct phub_opprog phub_opprog;
int phub_fock_create(phub_ctx *ctx, int n_sites, const int *n_up, int n_up_count,
                     const int *n_dn, int n_dn_count, phub_fock **out);
void phub_fock_destroy(phub_fock *f);
int phub_fock_dims(const phub_fock *f, int64_t *nsup, int64_t *nsdn);
int phub_fock_masks(phub_ctx *ctx, const phub_fock *f, uint64_t *d_up, uint64_t *d_dn);
int phub_opprog_create(phub_ctx *ctx, int n_terms, const double *coef, const int *term_len,
                       const int *elems, phub_opprog **out);
void phub_opprog_destroy(phub_opprog *p);
int phub_opprog_terms(const phub_opprog *p);
/* d_out[v] = O d_in[v] for nvec vectors of nsup*nsdn (vector v at offset v*dim).  The
 * first call on a basis builds the program's per-species source maps (at most 1 GiB,
 * kept with the program); calls on one program must not overlap. */
int phub_operator_apply(phub_ctx *ctx, const phub_fock *f, const phub_opprog *p,
                        const double *d_in, double *d_out, int nvec);
/* d_diag[i] = <e_i|O|e_i> for every basis element (the basis-element mode,
 * operators_basis.f90, used by the comparisons and bit_repr). */
int phub_operator_diag(phub_ctx *ctx, const phub_fock *f, const phub_opprog *p,
                       double *d_diag);
/* d_mat[i * dim + j] = (O e_i)[j], the reference's to_matrix layout (at most 16 GB). */
int phub_operator_matrix(phub_ctx *ctx, const phub_fock *f, const phub_opprog *p,
                         double *d_mat);

#ifdef __cplusplus
}
#endif
#endif /* PHUB_H */
