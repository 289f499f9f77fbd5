/*
 * alt3 -- MI355X-native three-phonon path of THERMACOND (C ABI).
 *
 * One context (alt3_ctx) holds a crystal, its force constants and a q-mesh in device
 * memory; the entry points below replace the reference subroutines named in each comment
 * (paths relative to the reference repository).  All buffers are host pointers owned by
 * the caller; arrays use the reference's Fortran column-major layouts with 0-based index
 * values unless stated.  A NULL output pointer means "keep on the device, do not copy".
 * Every function returns 0 on success or a negative ALT3_E* code; alt3_last_error()
 * returns the message of the last failure on the calling thread.  One host thread per
 * context; distinct contexts may be used from distinct threads.  No MPI inside.
 */
#ifndef ALT3_H
#define ALT3_H

#ifdef __cplusplus
extern "C" {
#endif

#define ALT3_OK 0
#define ALT3_E_ARG (-1)      /* bad argument / size mismatch */
#define ALT3_E_HIP (-2)      /* HIP runtime error */
#define ALT3_E_MESH (-3)     /* q is not a mesh point (get_k_info inside=2), odd mesh */
#define ALT3_E_STATE (-4)    /* prerequisite stage not run */
#define ALT3_E_DEN0 (-5)     /* zero denominator in V3 (matrix_elt_simplified den=0) */
#define ALT3_E_NOMEM (-6)    /* device allocation failed */
#define ALT3_E_NBE (-7)      /* occupation of a non-positive frequency (nbe w<=0) */

/* V3 arithmetic */
#define ALT3_V3_FP64_REF 0   /* per-element sum over all FC3 terms (matrix_elt_simplified) */
#define ALT3_V3_FP64 1       /* batched phase-GEMM + contraction, FP64 */
#define ALT3_V3_MIXED 2      /* phase sum and contractions with FP32 accumulation, FP64 scale:
                                24 branches run all three products on matrix cores (split FP16,
                                k_v3_chain16, or FP32, k_v3_chain: alt3_set_v3_mma), 6 and 12
                                branches on FP32 VALU (k_v3_rows), other counts the phase GEMM on
                                FP32 MFMA and the contractions on VALU (k_v3_mfma); cells whose
                                16-point MFMA tile exceeds the LDS (nb > ~28) run the FP64 path */

/* matrix-core arithmetic of the mixed V3 on 24-branch cells (alt3_set_v3_mma) */
#define ALT3_MMA_F16X3 0     /* FP16 MFMA on split operands: x = hi + lo in two halves, three
                                products hi.hi + hi.lo + lo.hi per product, FP32 accumulation
                                (k_v3_chain16) */
#define ALT3_MMA_F32 1       /* FP32 MFMA (k_v3_chain) */

/* delta-function integration of the scattering rates */
#define ALT3_DELTA_TETRA 0   /* linear tetrahedra, weight_tet (tetrahedron.f90:188-384) */
#define ALT3_DELTA_GAUSS 1   /* Gaussian delta_g (extratools.f90:239-246) at the mesh points */

typedef struct alt3_ctx alt3_ctx;

/* Crystal, force constants and mesh: the module state THERMACOND holds after
 * read_lattice/read_fc234 (THERMACOND/rest.f90:152-443). */
typedef struct {
    int natoms0;               /* atoms in the primitive cell */
    const double *mass;        /* [natoms0] amu */
    const double *r123;        /* [3][3] rows r1,r2,r3 (Angstrom), lat_fc.dat */
    const double *latparams;   /* [6] structure.params line 1 (a,b,c scaled, angles deg) */
    const double *primlatt;    /* [9] column-major, column j = primitive vector j, conv units */
    int natoms;                /* atoms in the neighbour list */
    const double *atompos;     /* [natoms][3] Cartesian */
    const int *iatomcell0;     /* [natoms] 1-based basis atom of each listed atom */
    int nt2;                   /* FC2 terms (fc2.dat) */
    const int *fc2_atoms;      /* [nt2][2] 1-based atom indices */
    const int *fc2_xyz;        /* [nt2][2] 1..3 */
    const double *fc2;         /* [nt2] eV/A^2 */
    int nt3;                   /* FC3 terms (fc3.dat) */
    const int *fc3_atoms;      /* [nt3][3] */
    const int *fc3_xyz;        /* [nt3][3] */
    const double *fc3;         /* [nt3] eV/A^3 */
    int nc[3];                 /* q-mesh (even), latdyn.params record 1 */
    int classical;             /* 0 quantum, 1 classical occupations */
    int host_only;             /* 1: build the mesh/symmetry tables only (no device needed);
                                  such a context answers alt3_sizes/alt3_get_mesh/alt3_kstars */
    /* dielectric.params (read_born, rest.f90:3580-3600).  born_flag != 0 adds the
     * non-analytic term of nonanal (rest.f90:3517-3578) to every dynamical matrix, as get_freq
     * does (rest.f90:4342-4348): D += (q.Z_a)_i (q.Z_b)_j / (q.eps.q eps0') / V / sqrt(m_a m_b)
     * * exp(-q^2/rho^2), rho = (|g1||g2||g3|)^0.33 / 4, q -> q + 1e-10 at q = 0. */
    int born_flag;
    const double *epsil;       /* [3][3] row i = epsil(i,1:3) (records 2-4); NULL when born_flag == 0 */
    const double *zeu;         /* [natoms0][3][3]: [k][i][j] = atom0(k)%charge(i,j) (record i of atom k) */
} alt3_system;

int alt3_create(const alt3_system *sys, alt3_ctx **out);
int alt3_destroy(alt3_ctx *ctx);
const char *alt3_last_error(void);
int alt3_device_bytes(const alt3_ctx *ctx, long long *bytes);
/* Free and total memory of the current device (hipMemGetInfo), for admission by footprint.
 * No counterpart in the reference, which does not manage device memory. */
int alt3_device_free(long long *free_bytes, long long *total_bytes);

/* Device bytes alt3_solve needs on this context for ntemps temperatures at the given V3
 * precision (an upper bound of the library's allocations; answers on host-only contexts,
 * so that a scheduler can admit work by HBM footprint before creating a device context).
 * No counterpart in the reference, which sizes its arrays per MPI rank. */
int alt3_footprint(const alt3_ctx *ctx, int ntemps, int precision, long long *bytes);
/* Sizes: nk = nc1*nc2*nc3, nibz, nb = 3*natoms0, nops (lattice point group). */
int alt3_sizes(const alt3_ctx *ctx, int *nk, int *nibz, int *nb, int *nops);

/* Mesh and symmetry tables built at creation, copied out:
 *   kpc_eig [nk][3]  make_kp_reg mesh (REAL*4 fractions), modules.f90:1721-1784
 *   kpc     [nk][3]  make_kp_reg_tet mesh (FP64), tetrahedron.f90:2-142
 *   mapibz  [nk], mapinv [nibz], kibz [nibz][3], wibz [nibz]   get_weights, modules.f90:2006-2162
 *   tet     [6nk][4] tetrahedron corners (make_kp_reg_tet)
 *   opk     [nops][9] op_kmatrix column-major (force_constants.f90:154-186)
 *   mappos  [nk]     get_negatives map, rest.f90:3208-3260; *npos its positive count
 * Any pointer may be NULL. */
int alt3_get_mesh(const alt3_ctx *ctx, double *kpc_eig, double *kpc, int *mapibz, int *mapinv,
                  double *kibz, double *wibz, int *tet, double *opk, int *mappos, int *npos);

/* Stars on the mesh of IBZ point i (kstarsinfbz1, sub.f90:5551-5590): returns narms,
 * fills ops[48] (0-based op index) and n0[48] (mesh index). */
int alt3_kstars(const alt3_ctx *ctx, int ibz, int *narms, int *ops, int *n0);

/* Harmonic phonons on the mesh: get_frequencies (rest.f90:2230-2405) with finite-difference
 * velocities (finitedif_vel, :2587-2631), then subst_eivecs (:3262-3293).  FP64 on device.
 *   eigenval [nk][nb]  (= Fortran eigenval(nb,nk)), |omega^2| in eV/A^2/amu
 *   eigenvec [nk][nb][nb] complex interleaved (= Fortran eigenvec(nb,nb,nk))
 *   veloc    [nk][nb][3]  (= Fortran veloc(3,nb,nk)), units of c */
int alt3_eig(alt3_ctx *ctx, double *eigenval, double *eigenvec, double *veloc);

/* Load externally computed eigen data instead of alt3_eig (same layouts). */
int alt3_set_eig(alt3_ctx *ctx, const double *eigenval, const double *eigenvec, const double *veloc);

/* |V3|^2 for q1 = IBZ points ibz_lo..ibz_hi (0-based, inclusive), q2 over the mesh,
 * q3 = -q1-q2: calculate_w3_ibz_split_sq (sub.f90:1050-1233).  Kept on the device for
 * alt3_fgr; v33s8 (optional) receives Fortran v33s8(ksub,nk,nb,nb,nb), FP64. */
int alt3_v3sq_ibz(alt3_ctx *ctx, int ibz_lo, int ibz_hi, int precision, double *v33s8);

/* alt3_v3sq_ibz over the whole IBZ of n contexts (same branch count, same device) in one
 * shared persistent launch: a grid sized to the device pops work items (context, q1, tile of
 * q2) from one counter, so small meshes fill the chip together.  The work item is the
 * reference's IBZ block of calculate_w3_ibz_split_sq (kappa.f90:420-431, sub.f90:1050-1233)
 * cut into q2 tiles; each context then holds its |V3|^2 exactly as after its own
 * alt3_v3sq_ibz(0, nibz-1) call (bit-identical: the tile is computed by the same code) and
 * alt3_solve reuses it.  precision: ALT3_V3_FP64 or ALT3_V3_MIXED.  tile: q2 points per work
 * item (6 branches: 16, 32 or 64; 12 branches: 8, 16 or 32), or 0 for the autotuned choice
 * (the fastest candidate timed once on the first context, cached per precision, branch count
 * and FC3 pair-vector count for the life of the process); tile_used (optional) receives it.
 * Other branch counts run one launch per context (tile_used = 0).  No counterpart in the
 * reference, which runs one IBZ block per MPI rank. */
int alt3_v3_batch(alt3_ctx *const *ctxs, int n, int precision, int tile, int *tile_used);

/* Copy out the |V3|^2 the context holds (from alt3_v3sq_ibz or alt3_v3_batch), in the layout
 * of alt3_v3sq_ibz's v33s8, without recomputing it: the v33s8 array calculate_w3_ibz_split_sq
 * leaves in memory (sub.f90:1050-1233). */
int alt3_get_v3sq(alt3_ctx *ctx, double *v33s8);

/* Tetrahedron-integrated Fermi golden rule: calculate_FGR_mpi2 (sub.f90:7199-7499,
 * weight_tet tetrahedron.f90:188-384).  Uses the |V3|^2 of the last alt3_v3sq_ibz call
 * (same ibz range).  P1sm/P2sm optional, Fortran (ksub,nk,nb,nb,nb). */
int alt3_fgr(alt3_ctx *ctx, double *P1sm, double *P2sm);

/* Gaussian-smeared variant of alt3_fgr: the same processes, corner energies and acoustic
 * masks, with delta(w1 - e) replaced by delta_g(w1 - e, sigma) = exp(-x^2/2 sigma^2) /
 * (sqrt(2 pi) sigma) (extratools.f90:239-246) evaluated at each mesh q2; sigma in cm^-1
 * (0.1 THz = 3.33564 cm^-1).  The reference's job 1 has no caller of this form (its
 * Gaussian rates, cubic_rates rest.f90:1698-1775, are unreachable from kappa.f90).
 * P1sm/P2sm optional, Fortran (ksub,nk,nb,nb,nb). */
int alt3_fgr_gauss(alt3_ctx *ctx, double sigma, double *P1sm, double *P2sm);
/* Load P1sm/P2sm computed elsewhere (the reference's v33sq_delta.NNN.dat split files, read
 * by calculate_tauRTA_iso sub.f90:7598-7613 and job 5's route) for irreducible points
 * ibz_lo..ibz_hi, Fortran (ksub,nk,nb,nb,nb) FP64, in place of alt3_v3sq_ibz + alt3_fgr.
 * Needs the eigen data (alt3_eig or alt3_set_eig); alt3_tau and what follows use them. */
int alt3_set_fgr(alt3_ctx *ctx, int ibz_lo, int ibz_hi, const double *P1sm, const double *P2sm);

/* Delta integration used by alt3_solve: ALT3_DELTA_TETRA (default) or ALT3_DELTA_GAUSS with
 * width sigma (cm^-1). */
int alt3_set_delta(alt3_ctx *ctx, int kind, double sigma);

/* Matrix-core arithmetic of the mixed-precision V3 on 24-branch cells: ALT3_MMA_F16X3 (default)
 * or ALT3_MMA_F32.  No reference counterpart: the reference computes V3 in FP64
 * (matrix_elt_simplified, THERMACOND/sub.f90:5888-6011); both kinds are its batched form. */
int alt3_set_v3_mma(alt3_ctx *ctx, int kind);
/* Occupations and scattering rates at tempk: distribution (sub.f90:1955-1980),
 * calculate_tauRTA_mpi2 (:7953-8194), Qandtau (:8198-8305).  Outputs Fortran (ksub,nb):
 * Q, QN, QU, tauinv_N, tauinv_U, tauinv_tot (1/s).  Any may be NULL. */
int alt3_tau(alt3_ctx *ctx, double tempk, double *Q, double *QN, double *QU,
             double *tauN, double *tauU, double *tautot);

/* RTA conductivity: calculate_RTA_distributionfunction (sub.f90:1238-1289) and mode_kappa
 * (:1292-1401).  Needs alt3_tau over the whole IBZ.  kappa Fortran (nb,3,3) W/m/K;
 * FRTA (nibz,nb,3); taueff (nibz,nb). */
int alt3_kappa_rta(alt3_ctx *ctx, double tempk, double *kappa, double *FRTA, double *taueff);

/* Collision matrix and right-hand side: CollisionMatrix_mpi (sub.f90:3034-3210) for the
 * current ibz range, cal_RHS_IBZ (:2618-2691).  CM Fortran (ksub*nb*3, nibz*nb*3); RHS
 * (nibz*nb*3), index 3*(i*nb+l)+alpha.  Both stay on the device for alt3_cg. */
int alt3_collision(alt3_ctx *ctx, double tempk, double *CM, double *RHS);

/* Conjugate gradients, cg_quadratic (cg.f90:19-69): minimises 1/2 x.A.x - b.x from the
 * initial x (overwritten).  A == NULL and b == NULL use the device collision matrix and
 * RHS of the last alt3_collision (n must then be nibz*nb*3).  Stops when
 * (dx/x)^2 < 10^(-2*ndig) or after max_iter iterations. */
int alt3_cg(alt3_ctx *ctx, int n, const double *A, const double *b, double *x, int max_iter,
            int ndig, double *funcmin, int *iters);

/* alt3_cg with no device allocation, for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc); no counterpart in the
 * reference.  The same solve through the same kernels, with A, b, x, the CG vectors and the
 * reduction partials staged in one file-scope static device array of 6*2^20 doubles (48 MiB)
 * per GPU, held until the process ends: n*n + 6n + 529 doubles must fit with A given
 * (n <= 2505), 5n + 529 with A == NULL (the context's collision matrix); larger n returns
 * ALT3_E_ARG.  Every call on one GPU uses that GPU's array: calls that overlap on one GPU race
 * (they need one host thread, or the caller's own serialisation, per GPU). */
int alt3_cg_static(alt3_ctx *ctx, int n, const double *A, const double *b, double *x, int max_iter,
                   int ndig, double *funcmin, int *iters);

/* Full-BTE conductivity from the CG solution x: mode_kappa_d (sub.f90:1407-1530) with
 * veloc_avergeIBZ_invsyFBZ (:5479-5537).  taueff accumulates onto its input, like the
 * reference module array (pass the alt3_kappa_rta output). */
int alt3_kappa_bte(alt3_ctx *ctx, double tempk, const double *x, double *kappa, double *taueff);

/* One update of the iterative BTE, update_mpi (sub.f90:2295-2449), for the irreducible points
 * of the context:  F2 = F_RTA + A(FFF)/Q  with A summed over the mesh as the reference does
 * (P1 (F(q1+q2,l3) - F(q2,l2)) + P2/2 (F(q2,l2) + F(q1-q2,l3)), F on the mesh rotated from the
 * IBZ by the star operations).  Needs alt3_tau at tempk.  FFF and FRTA Fortran (nibz,nb,3)
 * (the module arrays FFF and F_RTA); F2 Fortran (ksub,nb,3) (F2mpi).  A is held on the device
 * as a matrix built once per temperature. */
int alt3_update(alt3_ctx *ctx, double tempk, const double *FFF, const double *FRTA, double *F2);

/* The iterative BTE of job 3 (kappa.f90:1373-1487) at tempk on the device, after alt3_tau over
 * the whole IBZ: F starts at F_RTA (calculate_RTA_distributionfunction), then per iteration
 * F <- F + update_mix (update_mpi(F) - F) and mode_kappa (sub.f90:1292-1401), until
 * sqrt(sum (kappa - kappa_old)^2 / (9 nb)) < tol or max_iter iterations.  The reference uses
 * max_iter = 150 and update_mix = 0.4 (rest.f90:36,44) and stops at normk < 0.0001
 * (kappa.f90:1476); the two constants are REAL*4 literals: pass (double)0.4f and (double)1e-4f.
 * kappa Fortran (nb,3,3) of the last iteration; kappa_hist [iters][nb*9] or NULL and
 * taueff_hist [iters] or NULL (room for max_iter): after each iteration kappa and
 * tauinv_eff(nibz,nb), the only element of tauinv_eff mode_kappa leaves nonzero (it clears
 * the array for every mode); FFF Fortran (nibz,nb,3) final or NULL; iters = iterations done. */
int alt3_kappa_iter(alt3_ctx *ctx, double tempk, int max_iter, double update_mix, double tol, double *kappa,
                    double *kappa_hist, double *taueff_hist, double *FFF, int *iters);

/* Three-phonon self-energy of jobs 2/6/7 (kappa.f90:900-925) at nq points q off the mesh
 * (Cartesian, the units of the mesh points, [nq][3]): for every branch la, function_self_w3
 * (rest.f90:4012-4124) at omega = omega(q, la), with matrix_elt_full (:3664-3719) and get_freq
 * (:4319-4445) at q and at -q-k for every mesh point k, at tempk (K) and Lorentzian width eta
 * (etaz of latdyn.params, cm^-1).  The reference keeps |V3| >= v3_threshold (1E-7, a REAL*4
 * literal: pass (double)1e-7f).  Needs the mesh eigen data (alt3_eig or alt3_set_eig); cells
 * of 6, 12 or 24 branches.  omega0 (nb,nq) cm^-1; nself, uself the normal and umklapp parts,
 * complex (nb,nq) as (re,im) pairs; any may be NULL.  The linewidth of jobs 2/6/7 is
 * 2 Im(nself + uself).  Points q where -q-k is Gamma for a mesh k give a zero omega3 and the
 * reference's den = 0 stop (ALT3_E_DEN0) or, with round-off, an ill-conditioned sum. */
int alt3_self_w3(alt3_ctx *ctx, int nq, const double *q, double tempk, double eta, double v3_threshold, double *omega0,
                 double *nself, double *uself);

/* The Gaussian form of the same self-energy, function_self_w4 (rest.f90:4126-4240; self_energy
 * calls it for iread = 4, kappa.f90:3311, and function_self_w3 for iread = 3): the
 * imaginary part alone, sum over k, l2, l3 of |V3|^2 [(n2+n3+1) delta_g(omega-omega3-omega2)
 * + (n2-n3) delta_g(omega-omega3+omega2) + (n3-n2) delta_g(omega-omega2+omega3)] pi / nk, with
 * delta_g(x) = exp(-x^2/2/eta^2)/sqrt(2 pi)/eta (extratools.f90:239-246) cut at |x| < 6 eta; the
 * real parts of nself and uself are zero.  Arguments, units and conditions as alt3_self_w3;
 * eta must be positive. */
int alt3_self_w4(alt3_ctx *ctx, int nq, const double *q, double tempk, double eta, double v3_threshold, double *omega0,
                 double *nself, double *uself);

/* Isotope scattering, iso = 1 (kappa.f90:522-543).  giso (natoms0): the mass-variance factor
 * of each atom of the primitive cell (cal_giso, sub.f90:4695-4725; alt3_isotope_g gives it per
 * element).  The isotope term Piso (nibz,nb) of P_iso_split (sub.f90:4825-4961) is formed
 * from the eigen data of the context (again after alt3_eig or alt3_set_eig) and, while set,
 *   alt3_tau adds it to Q, QN and QU (calculate_tauRTA_iso, sub.f90:8311-8549),
 *   alt3_collision builds CollisionMatrix_iso (sub.f90:3387-3570) in place of
 *     CollisionMatrix_mpi: C = -P1(q0) + P1(q1+q0) + P2(q0) + Piso off the diagonal,
 *   alt3_update and alt3_kappa_iter apply update_iso (sub.f90:2456-2610),
 * and alt3_solve follows.  giso == NULL switches the term off.  Piso (optional) receives
 * Piso, Fortran (nibz,nb); it needs the eigen data.  Clears tau and what depends on it. */
int alt3_set_iso(alt3_ctx *ctx, const double *giso, double *Piso);

/* g of an element from its natural isotopes, cal_g (sub.f90:3766-4693):
 * g = sum_i f_i (1 - m_i/m_ave)^2 with abundances f_i and m_ave = sum_i f_i m_i; element is
 * compared as the reference's CHARACTER*2 names are.  An element without isotope data gets
 * g = 0, as in the reference.  Needs no context and no GPU. */
int alt3_isotope_g(const char *element, double *g);

/* The whole job-1 path on the device for each temperature: eig (unless set), V3 over the
 * IBZ and FGR (unless the context still holds them from an earlier call with the same
 * precision and delta integration; alt3_eig/alt3_set_eig clear them), then per T tau (one pass over P1sm/P2sm per group of four temperatures),
 * kappa_RTA, collision, CG (x0 = F_RTA, 100 iterations, 2 digits), kappa_BTE.  kappa_rta/kappa_bte: [ntemps][nb*9] Fortran (nb,3,3) per T.
 * stage_ms (optional, 8 doubles): eig, v3, fgr, tau, kappa_rta, collision, cg, kappa_bte. */
int alt3_solve(alt3_ctx *ctx, int ntemps, const double *temps, int precision,
               double *kappa_rta, double *kappa_bte, int *cg_iters, double *stage_ms);

/* Wall time (ms, HIP events on the context stream) of the kernels of the last call. */
int alt3_last_kernel_ms(const alt3_ctx *ctx, double *ms);
/* HIP-event duration (ms, context stream) of the last launch of the |V3|^2 contraction kernel
 * and of the tetrahedron FGR kernel (-1 if none ran yet), for roofline accounting.  No
 * counterpart in the reference. */
int alt3_kernel_ms(const alt3_ctx *ctx, double *v3_ms, double *fgr_ms);

#ifdef __cplusplus
}
#endif
#endif
