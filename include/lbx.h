/* lbx — MI355X-native SHARK electron-repulsion integrals: the C-ABI boundary.
 *
 * Plain C types only.  Every array argument marked "device" is a pointer into HBM on the
 * calling process's current HIP device (hipMalloc / torch.cuda memory); "host" arrays are
 * ordinary process memory.  Functions return 0 on success and a negative status otherwise;
 * lbx_last_error() then describes the failure.  No exception crosses this boundary.
 *
 * Each entry point names the reference interface (MihkuU/Lible, src/lible/ints/...) whose
 * behaviour it replaces.  The reference's eri4/eri3/eri2 return dense row-major tensors with AO
 * order = atoms in input order, shells in basis order, spherical m order 0,+1,-1,+2,-2,...;
 * the dense entry points here write exactly that layout.
 */
#ifndef LBX_H
#define LBX_H

#ifdef __cplusplus
extern "C" {
#endif

#define LBX_OK 0
#define LBX_ERR_INVALID (-1)   /* bad argument: sizes, null pointers, l out of range */
#define LBX_ERR_HIP (-2)       /* a HIP runtime call failed */
#define LBX_ERR_NOMEM (-3)     /* device allocation failed */
#define LBX_ERR_STATE (-4)     /* call not valid for this context (e.g. no auxiliary basis) */

/* One basis on a set of atoms, flattened.  Shells are ordered by atom (atoms in input order)
 * and in basis order within an atom, as constructShells builds them (shell.cpp:89-135).
 * Coordinates in bohr.  coeffs are the raw contraction coefficients; primitive and contracted
 * normalisation are applied inside (shell.cpp:11-46, utils.cpp:6-10). */
typedef struct lbx_basis {
    int natoms;
    const double *coords;      /* host, natoms*3, bohr */
    int nshells;
    const int *shell_atom;     /* host, nshells */
    const int *shell_l;        /* host, nshells; main basis l <= 4, auxiliary l <= 5 */
    const int *shell_nprim;    /* host, nshells */
    const double *exps;        /* host, sum(nprim) */
    const double *coeffs;      /* host, sum(nprim) */
} lbx_basis;

typedef struct lbx_context lbx_context;

/* Replaces Structure(basis_atoms_t, [aux,] atomic_nrs, coords) (structure.cpp:66-92) and the
 * per-class ShellPairData/ShellData + E-matrix set-up of the kernel functors
 * (shell_pair_data.cpp:73-202, ecoeffs.cpp:253-403, eri_kernels.cpp:1101-1162).  Builds the
 * normalised shells, screened primitive pairs (threshold 1e-15, shell_pair_data.hpp:55-57) and
 * the SHARK E matrices on the current device.  aux may be NULL. */
int lbx_create(lbx_context **ctx, const lbx_basis *main_basis, const lbx_basis *aux_basis);
int lbx_destroy(lbx_context *ctx);
const char *lbx_last_error(void);

/* Structure::getDimAO / getDimAOAux / getNAtoms (structure.hpp:79-91). */
int lbx_dims(const lbx_context *ctx, long *nbf, long *nbf_aux, long *natoms);

/* All work of a context is queued on this stream (default: the null stream).  hipStream_t. */
int lbx_set_stream(lbx_context *ctx, void *stream);
int lbx_synchronize(lbx_context *ctx);

/* Boys function F_0..F_n at count points (device x[count] -> device out[count*(n+1)]).
 * Replaces calcBoysF (boys_function.cpp:71-117); n <= 17. */
int lbx_boys(int n, long count, const double *x, double *out, void *stream);

/* Dense spherical ERIs, device outputs, row-major:
 *   eri4      (nbf,nbf,nbf,nbf)   replaces eri4()         (eri4.cpp:38-95)
 *   eri4_diag (nbf,nbf) (mn|mn)   replaces eri4Diagonal() (eri4.cpp:152-174)
 *   eri3      (nbf,nbf,naux)      replaces eri3()         (eri3.cpp:7-46)
 *   eri2      (naux,naux)         replaces eri2()         (eri2.cpp:27-72)
 *   eri2_diag (naux)              replaces eri2Diagonal() (eri2.cpp:74-96) */
int lbx_eri4(lbx_context *ctx, double *out);
int lbx_eri4_diag(lbx_context *ctx, double *out);
int lbx_eri3(lbx_context *ctx, double *out);
int lbx_eri2(lbx_context *ctx, double *out);
int lbx_eri2_diag(lbx_context *ctx, double *out);

/* Dense first-derivative ERI tensors, device outputs, row-major.  Component 3k + x (x, y, z) is
 * the derivative with respect to coordinate x of the centre of the k-th index's shell, i.e. the
 * blocks of the reference's ERI4D1 / ERI3D1 / ERI2D1 kernels ([Ax..Dz], [Ax..Cz], [Ax..Bz];
 * eri_kernels.hpp:113-170, eri_kernel_funs.hpp:350-732) placed at their AO positions:
 *   eri4d1 (12, nbf, nbf, nbf, nbf)   replaces ERI4D1Kernel::operator()
 *   eri3d1 (9, nbf, nbf, naux)        replaces ERI3D1Kernel::operator()
 *   eri2d1 (6, naux, naux)            replaces ERI2D1Kernel::operator()
 * Small systems only (the sizes above); classes up to (ff|ff), (ff|h), (h|h). */
int lbx_eri4d1(lbx_context *ctx, double *out);
int lbx_eri3d1(lbx_context *ctx, double *out);
int lbx_eri2d1(lbx_context *ctx, double *out);

/* Blocks of listed shell quartets, triplets or doublets, the per-call unit of the reference's
 * kernel functors (ERI4Kernel / ERI3Kernel / ERI2Kernel / ERI4D1Kernel / ERI3D1Kernel /
 * ERI2D1Kernel::operator(), eri_kernels.hpp:60-170, eri_kernel_funs.hpp:104-346, 350-732), for
 * many entries in one launch.  shells: host, n entries of 4 (a, b, c, d: main-basis shell
 * indices), 3 (a, b main, c auxiliary) or 2 (a, b auxiliary) shell indices, shells numbered in
 * AO order from 0 within their basis; every entry of one call has the angular momenta of the
 * first.  out: device, entry t at out + t * ncomp * Sa*Sb*Sc*Sd (S = 2l + 1 of each shell), row
 * major in the entry's shell order (the reference's vec4d / vec3d / vec2d), ncomp = 1 for the
 * integrals and 12 / 9 / 6 for the derivatives (component 3k + x: coordinate x of the centre of
 * the entry's k-th shell, the reference's [Ax..Dz] order).  Entries whose shell pair has no
 * primitive pair above the 1e-15 screen (absent from the reference's ShellPairData) are zero. */
int lbx_eri4_batch(lbx_context *ctx, long n, const int *shells, double *out);
int lbx_eri3_batch(lbx_context *ctx, long n, const int *shells, double *out);
int lbx_eri2_batch(lbx_context *ctx, long n, const int *shells, double *out);
int lbx_eri4d1_batch(lbx_context *ctx, long n, const int *shells, double *out);
int lbx_eri3d1_batch(lbx_context *ctx, long n, const int *shells, double *out);
int lbx_eri2d1_batch(lbx_context *ctx, long n, const int *shells, double *out);

/* The same block layout for the reference's remaining per-block kernels:
 *   eri4soc (ncomp 3)  replaces ERI4SOCKernel::operator() (eri_kernels.hpp:188-204,
 *                      eri_kernel_funs.cpp:846-974): components x, y, z of
 *                      sum E'_{d1} R_{+d2} E - E'_{d2} R_{+d1} E, (d1, d2) = (y,z), (z,x), (x,y),
 *                      E' = dE/d(A - B) of the bra pair; odd under the exchange of a and b
 *   eri3soc (ncomp 3)  replaces ERI3SOCKernel::operator() (eri_kernels.hpp:206-222,
 *                      eri_kernel_funs.cpp:976-1090)
 *   eri2d2  (ncomp 36) replaces ERI2D2Kernel::operator() (eri_kernels.hpp:171-186,
 *                      eri_kernel_funs.cpp:343-433): d2(a|b)/dX_i dX_j, i and j over
 *                      [Ax, Ay, Az, Bx, By, Bz], component 6 i + j */
int lbx_eri4soc_batch(lbx_context *ctx, long n, const int *shells, double *out);
int lbx_eri3soc_batch(lbx_context *ctx, long n, const int *shells, double *out);
int lbx_eri2d2_batch(lbx_context *ctx, long n, const int *shells, double *out);

/* lbx_eri4_batch that allocates nothing.  No counterpart in the reference: for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  Same arguments, layout and results as lbx_eri4_batch; the per-call
 * task list is staged in a device array of the library (one per GPU, held until the process
 * ends) instead of a buffer of the call, so n is at most 1048576 (LBX_ERR_INVALID beyond).
 * Calls on one GPU share that array: calls that overlap on one GPU (from two contexts on it, on
 * different threads) race; calls on one context are serialised by its mutex. */
int lbx_eri4_batch_static(lbx_context *ctx, long n, const int *shells, double *out);

/* Coulomb and exchange matrices, no reference counterpart (the reference stops at the dense
 * tensor).  D, J, K: device (nbf,nbf) row-major, D symmetric.
 *   J_mn = sum_kl (mn|kl) D_kl,   K_mk = sum_nl (mn|kl) D_nl
 * over unique shell quartets whose Schwarz bound Q_ab*Q_cd >= schwarz_thresh
 * (Q_ab = sqrt(max |(ab|ab)|)).  precision: LBX_FP64, or LBX_MIXED (quartets with bound below
 * mixed_thresh evaluated in FP32 with FP64 accumulation). */
#define LBX_FP64 0
#define LBX_MIXED 1
typedef struct lbx_jk_stats {
    double quartets;       /* shell quartets evaluated */
    double flops;          /* algorithmic FLOPs of those quartets */
    double quartets_fp32;  /* of which evaluated in FP32 (mixed precision) */
} lbx_jk_stats;
int lbx_jk(lbx_context *ctx, const double *D, double *J, double *K, double schwarz_thresh,
           int precision, double mixed_thresh, lbx_jk_stats *stats);

/* Density-fitted (RI) Coulomb and exchange, no reference counterpart (the consumer of the
 * reference's eri3() / eri2(), eri3.cpp:7-46, eri2.cpp:27-72): with V = (P|Q) = L L^T,
 *   J_mn = sum_PQ (mn|P) (V^-1)_PQ sum_kl (Q|kl) D_kl
 *   K_mk = sum_P sum_nl W_P(mn) D_nl W_P(kl),   W_P = sum_Q (L^-1)_PQ (Q|..)
 * D, J, K: device (nbf, nbf) row-major, D symmetric; K may be NULL (J only).  Needs an
 * auxiliary basis.  The first call builds the fitted 3-index tensor W (8 nbf^2 naux bytes of
 * device memory, plus a work buffer of the same size while it is built or once K is asked for)
 * and the context keeps both until lbx_destroy, so later calls (an SCF loop) cost only the
 * products with D.  LBX_ERR_NOMEM when they do not fit; LBX_ERR_STATE when (P|Q) is not
 * numerically positive definite. */
int lbx_ri_jk(lbx_context *ctx, const double *D, double *J, double *K);

/* One-electron integrals over the main basis, dense (nbf, nbf) row-major device matrices, each
 * optional (NULL: not computed): the overlap S (replaces overlap(), oneelints.cpp:169-197), the
 * kinetic energy T (kineticEnergy(), oneelints.cpp:348-376) and the attraction of an electron to
 * ncharges point charges V (externalCharges(), oneelints.cpp:1087-1120; nuclearAttraction(),
 * oneelints.cpp:1703-1744, is the case of the nuclei with q = Z).  charges: host, (x, y, z, q)
 * per charge, bohr; ncharges = 0 gives V = 0.  Outputs are overwritten. */
int lbx_one_electron(lbx_context *ctx, double *S, double *T, int ncharges, const double *charges,
                     double *V);

/* Nuclear gradient of the one-electron energy E1 = sum_mn D_mn (T + V)_mn - sum_mn W_mn S_mn
 * (D the density, W the energy-weighted density; both device (nbf, nbf); W may be NULL), the
 * contraction of the reference's overlapD1Kernel, kineticEnergyD1Kernel and
 * externalChargesD1Kernel (orbital part, oneelints.cpp) blocks.  grad: device (natoms, 3),
 * overwritten: derivatives with respect to the basis-function centres.  charge_grad: device
 * (ncharges, 3) or NULL, overwritten: dE1/d(charge position), the operator part
 * (externalChargesOperatorD1Kernel).  For the nuclei as charges the full one-electron gradient
 * of atom A is grad[A] + charge_grad[A]. */
int lbx_one_electron_grad(lbx_context *ctx, const double *D, const double *W, int ncharges,
                          const double *charges, double *grad, double *charge_grad);

/* Batched RI integrals over many independent molecules (no reference counterpart: the
 * reference runs eri3() / eri2() one Structure at a time, eri3.cpp:7-46, eri2.cpp:27-72).
 * lbx_create_batch builds one context over nmol molecules (main_bases[k], aux_bases[k] for
 * molecule k, each with its own atoms); shell pairs only form within a molecule, so the
 * dense-tensor entry points, J/K, the gradient and the tuner reject a batched context.
 * lbx_ri_sums evaluates every molecule's (mn|P) and (P|Q) and reduces them on the fly: out is a
 * device (nmol, 4) array of the sums of |value| and of value over the dense eri3 tensor
 * (nbf, nbf, naux) and over the dense eri2 tensor (naux, naux) of each molecule, i.e. the
 * checksums the reference's tests take of eri3() and eri2() (tests_ints.cpp:694-746).  With
 * lbx_set_graphs on, the whole launch sequence is captured into a HIP graph on the first call
 * (per out pointer) and replayed afterwards.  Also valid on an ordinary context with an auxiliary
 * basis (nmol = 1). */
typedef struct lbx_ri_stats {
    double triplets;        /* (ab|P) shell triplets evaluated */
    double doublets;        /* (P|Q) shell doublets evaluated */
    double flops;           /* algorithmic FLOPs */
    double triplets_fp32;   /* of which in FP32 (LBX_MIXED) */
    double doublets_fp32;
} lbx_ri_stats;
int lbx_create_batch(lbx_context **ctx, int nmol, const lbx_basis *main_bases,
                     const lbx_basis *aux_bases);
/* precision LBX_FP64, or LBX_MIXED: the triplets (doublets) whose Schwarz bound Q_ab Q_P
 * (Q_P Q_Q), Q = sqrt(max |(xx|xx)|), is below mixed_thresh in FP32 arithmetic with FP64 Boys
 * values and sums; nothing is screened out.  With graphs on, one graph per (out, precision
 * split). */
int lbx_ri_sums(lbx_context *ctx, double *out, int precision, double mixed_thresh,
                lbx_ri_stats *stats);

/* Nuclear gradient of E = 1/2 tr(DJ) - 1/4 tr(DK) for a fixed density, no reference counterpart
 * (the reference supplies the derivative integrals, ERI4D1Kernel, eri_kernels.hpp:113-131 /
 * eri_kernel_funs.cpp:582-748; this contracts them with the density inside the kernel):
 *   grad[3 A + x] = sum over unique quartets with Q_ab Q_cd >= schwarz_thresh of
 *                   f (4 D_mn D_kl - D_mk D_nl - D_ml D_nk) d(mn|kl)/dR_{A,x}
 * with f the permutational weight.  D: device (nbf, nbf) symmetric; grad: device 3 natoms
 * (atom-major, x y z), Eh/bohr.  precision LBX_FP64, or LBX_MIXED: quartets with bound below
 * mixed_thresh in FP32 arithmetic (R, E, E', transforms) with FP64 Boys values and sums; stats
 * count the quartets and those in FP32. */
int lbx_jk_grad(lbx_context *ctx, const double *D, double *grad, double schwarz_thresh,
                int precision, double mixed_thresh, lbx_jk_stats *stats);

/* Autotuner (no reference counterpart).  Each class pair (la lb|lc ld) has up to eight kernel
 * variants: 0 one lane per quartet with the bra pairs on the lanes, 1 the same with the ket
 * pairs on the lanes, 2 one wave per quartet on FP64 MFMA, 6 one workgroup per bra pair and
 * block of ket pairs (ket transform on the VALU, bra transform on FP64 MFMA with the ket pairs
 * along N), and 3, 4, 5, 7 the forms 0, 1, 2, 6 with R and both transforms in FP32 (5 and 7 on
 * the FP32 MFMA; the small-bound quartets of LBX_MIXED).  lbx_tune times every variant of every
 * class pair on this density and threshold (on a sample of ~2M quartets of each class pair:
 * every k-th pair of one side with all its partners) and keeps the fastest per quartet: the
 * fastest FP64 form for FP64 work and the fastest of all forms for the FP32 tier of mixed
 * builds; without it lbx_jk picks by a work estimate from the primitive-pair counts.  On a
 * batched context (lbx_create_batch) it tunes lbx_ri_sums instead: every eri3 / eri2 class pair's
 * FP64 sum kernel, one-lane or MFMA form, timed over its whole task table (D and the threshold
 * are not used; D may be NULL). */
int lbx_tune(lbx_context *ctx, const double *D, double schwarz_thresh);
/* Fixes the variant of class pair (la lb|lc ld), la >= lb, lc >= ld, bra class >= ket class in
 * the order (0,0),(1,0),(1,1),(2,0),...: variants 0-2 and 6 for FP64 work, 3-5 and 7 for the
 * FP32 tier;
 * fails if the class pair has no such variant.  variant -1 restores both defaults. */
int lbx_set_variant(lbx_context *ctx, int la, int lb, int lc, int ld, int variant);

/* Per-class timing of lbx_jk (the reference's only timing facility is eri4Benchmark,
 * eri4.cpp:97-150, which prints per-(la,lb,lc,ld) wall time).  While enabled, every lbx_jk
 * records a HIP event pair around each class-pair kernel on the context stream;
 * lbx_profile_read returns the accumulated totals since the last enable. */
typedef struct lbx_class_profile {
    int la, lb, lc, ld;
    int kernel;            /* variant: 0 bra on lanes, 1 ket on lanes, 2 MFMA, 6 block;
                              3 / 4 / 5 / 7 their FP32 forms */
    int calls;
    double ms;             /* summed kernel time */
    double quartets;       /* summed quartets */
    double flops;          /* summed algorithmic FLOPs */
} lbx_class_profile;
int lbx_profile_enable(lbx_context *ctx, int on);
int lbx_profile_read(lbx_context *ctx, lbx_class_profile *out, int capacity, int *count);

/* Restricts the following lbx_jk calls on this context to one class pair (la lb|lc ld)
 * (canonical: la >= lb, lc >= ld, bra class >= ket class); all four -1 lifts the restriction.
 * No reference counterpart: the per-(la,lb,lc,ld) loop of eri4Benchmark (eri4.cpp:97-150) is
 * the model.  J and K then hold that class pair's contribution only (symmetrised), which gives
 * per-class-pair error and timing reports. */
int lbx_set_class_filter(lbx_context *ctx, int la, int lb, int lc, int ld);

/* Restricts the following lbx_jk calls on this context to the quartets whose Schwarz bound
 * Q_ab Q_cd lies in [lo, hi) (besides the call's own threshold; with LBX_MIXED each precision
 * tier keeps its part of the window); lo = 0, hi = INFINITY lifts it.  No reference
 * counterpart: it isolates the error of the FP32 tier of a mixed-precision build (window
 * [threshold, split): the same quartets in FP64 and in FP32, with partial sums small enough
 * that FP64 rounding does not hide the difference). */
int lbx_set_bound_window(lbx_context *ctx, double lo, double hi);

/* HIP graphs for repeated calls, no reference counterpart; off by default.  On: lbx_jk captures
 * the launch sequence of a (D, J, K, thresholds) set into a graph after its first call and
 * replays it on later calls with the same pointers (the kernels read the buffers' current
 * contents, so an SCF loop that reuses its buffers replays); lbx_ri_sums captures its launches
 * per output buffer and precision split.  The capture runs on a private non-blocking stream, but
 * while it is in progress the HIP runtime rejects synchronous calls that other threads of the
 * process make on the legacy (null) stream, so a program that drives the GPU from several
 * threads keeps graphs off.  Turning graphs off drops the captured ones. */
int lbx_set_graphs(lbx_context *ctx, int on);

#ifdef __cplusplus
}
#endif

#endif
