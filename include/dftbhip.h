/*
 * dftbhip: MI355X (gfx950) backend for the SCC-DFTB hot path of DFTB+.
 *
 * C ABI, plain pointers and sizes. All matrices are column-major (Fortran order), all
 * indices passed in neighbour-list arrays are the 1-based values DFTB+ stores. Every function
 * returns 0 on success, a negative code on argument/device errors (message from
 * dftbhip_last_error) and, for the eigensolver, a LAPACK-style positive info.
 */
#ifndef DFTBHIP_H
#define DFTBHIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- device management: replaces dftbp_extlibs_magma_get_gpus_available / magma_init
 *      (src/dftbp/extlibs/magmac.c:1-46, src/dftbp/extlibs/magma.F90:30-60) ---------------- */
int dftbhip_get_gpus_available(int* ngpus);
int dftbhip_init(int device);
int dftbhip_finalize(void);
int dftbhip_last_error(char* buf, int len);

/* ---- generalized symmetric-definite eigensolver: replaces magmaHegvd (dreal)
 *      (src/dftbp/math/eigensolver.F90:1143-1245; contract of hegvd/hegvr :561-920).
 * itype must be 1, uplo 'L'. On exit a holds the S-orthonormal eigenvectors (jobz = 'V'),
 * b the Cholesky factor, w the ascending eigenvalues. info > n: leading minor info-n of b is
 * not positive definite. */
int dftbhip_sygvd(int itype, char jobz, char uplo, int n, double* a, int lda, double* b, int ldb,
                  double* w);

/* ---- complex Hermitian twin: replaces magmaHegvd (dcomplex) / zhegvd
 *      (src/dftbp/math/eigensolver.F90:113-117, 1143-1245). a and b are complex(dp) arrays
 *      (interleaved re, im; lda, ldb in complex elements), lower triangles read. On exit a holds
 *      the S-orthonormal eigenvectors (C^H S C = I, jobz = 'V'), the lower triangle of b the
 *      Cholesky factor L (S = L L^H; its strict upper part is left as given), w the ascending
 *      eigenvalues. info > n: leading minor info-n of b is not positive definite. */
int dftbhip_hegvd(int itype, char jobz, char uplo, int n, double* a, int lda, double* b, int ldb,
                  double* w);

/* ---- BLAS / LAPACK style drop-ins (the dgemm and dstedc DFTB+ reaches through
 *      src/dftbp/math/blasroutines.F90 and the divide-and-conquer solvers of
 *      src/dftbp/math/eigensolver.F90:561-660). dstedc: on exit d holds the ascending
 *      eigenvalues and z (n x n) the eigenvectors of the tridiagonal (d, e). */
int dftbhip_dgemm(char transa, char transb, int m, int n, int k, double alpha, const double* a,
                  int lda, const double* b, int ldb, double beta, double* c, int ldc);
int dftbhip_dstedc(int n, double* d, double* e, double* z, int ldz);
/* LAPACK dsytrd (uplo = 'L'): Householder tridiagonalisation of the symmetric a (lower
 * triangle read); on exit d, e, tau and the reflectors below the subdiagonal of a, as LAPACK
 * (the reduction DFTB+ reaches through the dsygvd / dsyevd routes of
 * src/dftbp/math/eigensolver.F90:561-660). */
int dftbhip_dsytrd(char uplo, int n, double* a, int lda, double* d, double* e, double* tau);
/* dftbhip_dsytrd on device arrays (a, d, e, tau device pointers) without any allocation: no
 * counterpart in the reference; for callers that cannot allocate device memory (code under
 * graph capture, serving loops). The workspace is a fixed 96 MiB device array per GPU
 * (n up to about 17500); larger n is rejected with kErrArg. dftbhip_init must have been
 * called. Calls that overlap on one GPU share the workspace and race. */
int dftbhip_dsytrd_noalloc(int n, double* a, int lda, double* d, double* e, double* tau);
/* Kernel tuning. param 0: row tiles per workgroup of the tridiagonalisation's SYMV (2, 4, 8,
 * 16, or 0 = chosen from the trailing size per column, the default: the autotuned table of the
 * arithmetic type when there is one, else fixed thresholds); param 1, value 0: drop the
 * autotuned tables; param 2: k-points that are all real and of one size are tridiagonalised
 * together, one launch per column kernel for all points (1, the default) or one by one (0);
 * param 3: the neighbour list, pairs and SK blocks of a geometry are built on the device (1,
 * the default) or on the host (0); param 4: the molecule batch's per-molecule Fermi levels,
 * energies, convergence tests and Broyden mixing run on host threads (0), as device kernels
 * (1), or as device kernels with every warm SCC iteration replayed from one captured HIP
 * graph (2, the default); param 5: tile order of full-output GEMMs, 0 the launch grid's or 1
 * XCD-grouped (the default: the tiles an XCD runs together share their panels in its L2);
 * param 6: Ogita-Aishima refinement of the sparse tail, pairs outside the first-order range
 * re-orthonormalised only (0, the default) or rotated by the exact 2x2 angle (1); param 7:
 * the Broyden history restarts at the iteration that seeds the mixed-precision tail (1) or is
 * kept (0, the default); param 8: arithmetic of a step of the tail's refinement, 0: X^T H X,
 * X^T S X and the correction product X E in FP64, 1: the same with X E in FP32, 2 (the
 * default): residual form, X^T S X, the Rayleigh quotients and the residuals
 * H X - S X diag(lam) in FP64, the products with the residuals and with E in FP16 (FP32
 * sums); param 10: residual
 * form, tile rows of X^T S X below its 128 x 128 diagonal tiles formed per step (1, the
 * default; -1 the whole lower triangle); param 11: residual form, iterations whose previous
 * charge error is at most 10^-value use arithmetic 1 (0, the default: never); param 12: the
 * tridiagonalisation's launch sequence is captured in a HIP graph and replayed for repeated
 * calls on the same buffers (1) or launched kernel by kernel (0, the default: the replay saves
 * about 2 % of an FP32 iteration, and rocprofv3 kernel tracing fails on the capture); param 9: vectors per lane of the tail's
 * block-sparse products (1, the default, 2 or 4); param 13: the two-stage tridiagonalisation's
 * bulge chasing runs on one workgroup of an ordinary launch (1) or on the cooperative grid (0,
 * the default; the one-workgroup launch is taken by itself when the device refuses the
 * cooperative one, as it does while another process holds its cooperative-launch resource).
 * No counterpart in the reference. */
int dftbhip_set_tuning(int param, int value);
/* Half neighbour list on the device, as updateNeighbourList (src/dftbp/dftb/periodic.F90:446-696)
 * builds it: atom a lists itself, then the images of atoms b >= a within cutoff by increasing
 * distance (ties by cell translation, then atom); image atoms follow the central ones in
 * (translation, atom) order. latVecs: the three lattice vectors (9 doubles) or NULL for a
 * cluster. Query: with coordsAll NULL only *nAll and *mNeighbour (largest neighbour count,
 * self excluded) are set. Fill: *nAll and *mNeighbour give the capacities of coordsAll
 * (3 x nAll), img2CentCell (nAll, 1-based), iNeighbour and neighDist2 ((mNeighbour + 1) x
 * nAtom, column k = 0 the atom itself, 1-based image indices, 0 past the last neighbour;
 * neighDist2 optional), nNeighbour (nAtom, self excluded); on exit the actual sizes. Errors:
 * atoms closer than 0.01 bohr, capacities too small. */
int dftbhip_neighbour_list(int nAtom, const double* coords, const double* latVecs, double cutoff,
                           int* nAll, int* mNeighbour, double* coordsAll, int* img2CentCell,
                           int* iNeighbour, int* nNeighbour, double* neighDist2);
/* Autotuner of the SYMV tile count (no counterpart in the reference): tridiagonalises a
 * synthetic symmetric n x n matrix (256 <= n <= 65535) once per option 2, 4, 8, 16 with every
 * SYMV launch timed, picks per 256-row bucket of trailing sizes the option of least time
 * (keeping the option of the bucket below unless another is at least 2 % faster) and
 * installs the result for precision 0 (FP64) or 1 (FP32). table: up to maxRows rows
 * (first trailing size, tiles), ascending; *nRows: rows in the installed table. detail
 * (optional, detailLen doubles): summed seconds, option-major: detail[o * nb + b] for option
 * o and bucket b, nb = ceil(n / 256). Allocates device memory for the test matrix. */
int dftbhip_autotune(int n, int precision, int* table, int maxRows, int* nRows, double* detail,
                     int detailLen);
/* LAPACK dlacpy (uplo = 'A'): b = a through the device copy kernel (the copy used for HBM
 * counter calibration; DFTB+ copies square matrices with plain array assignment). */
int dftbhip_dlacpy(char uplo, int m, int n, const double* a, int lda, double* b, int ldb);

/* ---- density matrices: replace makeDensityMtxRealGPU / makeEnergyDensityMtxRealGPU
 *      (src/dftbp/dftb/densitymatrix.F90:1105-1395). Lower triangle of dm/edm is set:
 *      dm = C diag(f) C^T, edm = C diag(f*eig) C^T over the first nlevels columns. */
int dftbhip_dm_real(int n, int nlevels, const double* evec, int lde, const double* filling,
                    double* dm, int lddm);
int dftbhip_edm_real(int n, int nlevels, const double* evec, int lde, const double* filling,
                     const double* eigen, double* edm, int lddm);
/* ---- complex twins: replace makeDensityMtxCmplxGPU / makeEnergyDensityMtxCmplxGPU
 *      (src/dftbp/dftb/densitymatrix.F90:1401-1539, 1545-1691). evec and dm/edm are complex(dp)
 *      arrays, i.e. interleaved (re, im) doubles with leading dimensions lde, lddm counted in
 *      complex elements. The lower triangle of dm = C diag(f) C^H (edm: weights f*eig) over
 *      the first nlevels columns is set, the strict upper triangle is zeroed; evec is not
 *      modified. */
int dftbhip_dm_cmplx(int n, int nlevels, const double* evec, int lde, const double* filling,
                     double* dm, int lddm);
int dftbhip_edm_cmplx(int n, int nlevels, const double* evec, int lde, const double* filling,
                      const double* eigen, double* edm, int lddm);

/* ---- Slater-Koster parameters and repulsives (the content of the .skf files after
 *      getFullTable, src/dftbp/dftbplus/parser.F90:3668-4130). Species have at most two
 *      shells with l <= 1. Tables are row-major (nrows x ninteg[sp1*nspecies+sp2]), row i at
 *      r = (i+1)*dist. ------------------------------------------------------------------- */
#define DFTBHIP_MAX_SPECIES 8
typedef struct {
  int nspecies;
  int nshell[DFTBHIP_MAX_SPECIES];
  int angshell[DFTBHIP_MAX_SPECIES][2];
  double onsite[DFTBHIP_MAX_SPECIES][2];  /* per shell (Ha) */
  double hubbu[DFTBHIP_MAX_SPECIES];      /* first-shell U (non shell-resolved SCC) */
  double occ[DFTBHIP_MAX_SPECIES][2];     /* neutral shell occupations */
  int nrows;
  double dist;
  int ninteg[DFTBHIP_MAX_SPECIES * DFTBHIP_MAX_SPECIES];
  const double* htab[DFTBHIP_MAX_SPECIES * DFTBHIP_MAX_SPECIES];
  const double* stab[DFTBHIP_MAX_SPECIES * DFTBHIP_MAX_SPECIES];
  int rep_nint[DFTBHIP_MAX_SPECIES * DFTBHIP_MAX_SPECIES];
  double rep_cutoff[DFTBHIP_MAX_SPECIES * DFTBHIP_MAX_SPECIES];
  double rep_expc[DFTBHIP_MAX_SPECIES * DFTBHIP_MAX_SPECIES][3];
  const double* rep_xs[DFTBHIP_MAX_SPECIES * DFTBHIP_MAX_SPECIES];   /* rep_nint */
  const double* rep_coef[DFTBHIP_MAX_SPECIES * DFTBHIP_MAX_SPECIES]; /* (rep_nint-1) x 4 */
  const double* rep_last[DFTBHIP_MAX_SPECIES * DFTBHIP_MAX_SPECIES]; /* 6 */
} dftbhip_params;

typedef struct dftbhip_sk dftbhip_sk;
int dftbhip_sk_create(dftbhip_sk** sk, const dftbhip_params* p);
int dftbhip_sk_destroy(dftbhip_sk* sk);

/* ---- sparse kernels with the reference's array conventions ------------------------------
 * coords(3, nAll); species(nAll) 1-based; img2CentCell(nAll) 1-based;
 * iNeighbour(0:mNeighbour, nAtom) 1-based image indices; nNeighbourSK(nAtom);
 * iPair(0:mNeighbour, nAtom) 0-based sparse offsets; iAtomStart(nAtom+1) 1-based. */

/* replaces buildH0 / buildS (src/dftbp/dftb/nonscc.F90:66-181) */
int dftbhip_build_h0_s(const dftbhip_sk* sk, int nAtom, int nAll, const double* coords,
                       const int* species, const int* img2CentCell, int mNeighbour,
                       const int* iNeighbour, const int* nNeighbourSK, const int* iPair,
                       int sparseSize, double* h0, double* over);
/* replaces unpackHS_real (src/dftbp/dftb/sparse2dense.F90:266-321); writes both triangles */
int dftbhip_unpack_hs_real(int nAtom, const int* iAtomStart, const int* img2CentCell,
                           int mNeighbour, const int* iNeighbour, const int* nNeighbourSK,
                           const int* iPair, int sparseSize, const double* orig, double* square,
                           int ldsq);
/* replaces packHS_real (src/dftbp/dftb/sparse2dense.F90:1024-1097); primitive is overwritten */
int dftbhip_pack_hs_real(int nAtom, const int* iAtomStart, const int* img2CentCell,
                         int mNeighbour, const int* iNeighbour, const int* nNeighbourSK,
                         const int* iPair, int sparseSize, const double* square, int ldsq,
                         double* primitive);
/* replaces mullikenPerOrbital (src/dftbp/dftb/populations.F90:120-193); qorb(nOrb) */
int dftbhip_mulliken(int nAtom, const int* iAtomStart, const int* img2CentCell, int mNeighbour,
                     const int* iNeighbour, const int* nNeighbourSK, const int* iPair,
                     int sparseSize, const double* over, const double* rho, double* qorb);

/* ---- resident SCC context: the full hot path on the device -------------------------------
 * (SCC loop of src/dftbp/dftbplus/main.F90:1429-1563, gradients main.F90:7058-7440) */
typedef struct {
  double kT;            /* electron temperature (Ha), floored at 1e-8 */
  double scc_tol;       /* max |q_out - q_in| */
  int max_scc_iter;
  int min_scc_iter;
  double mix_param, omega0, min_weight, max_weight, weight_fac;   /* Broyden */
  double fd_delta;      /* finite-difference step of the SK derivatives */
  double ewald_tol;
  int force_mode;       /* 0: finite differences (reference default), 1: analytic adjoint */
  int precision;        /* 0: FP64, 1: mixed (FP32 MFMA bulk + FP64 refinement) */
  /* mixed precision only: an SCC iteration refines its eigenpairs in FP64 (Ogita-Aishima,
   * refine_steps steps, then FP64 density matrices) when the charge error of the previous
   * iteration is at most refine_below (0: never, +inf: every iteration); with refine_steps = 0
   * such iterations run the plain FP64 path instead */
  double refine_below;
  int refine_steps;
  int tridiag;          /* 0: two-stage (dense -> band -> tridiagonal), 1: one-stage Householder */
  /* mixed precision only (0: off): with tail_below > 0 the refined iterations keep all n
   * refined eigenvectors, and a refined iteration whose previous charge error is at most
   * tail_below skips the eigensolve: one Ogita-Aishima step refines the previous iteration's
   * FP64 eigenpairs against the new H (late in the SCC cycle H changes by much less than the
   * level spacing from one iteration to the next) */
  double tail_below;
  /* Ogita-Aishima steps per tail iteration (0 is taken as 1); more steps keep the eigenpairs
   * closer to the new H, which the SCC convergence of large systems with dense spectra needs */
  int tail_steps;
} dftbhip_options;

typedef struct dftbhip_ctx dftbhip_ctx;
int dftbhip_ctx_create(dftbhip_ctx** ctx, const dftbhip_params* p, const dftbhip_options* opt);
int dftbhip_ctx_destroy(dftbhip_ctx* ctx);
/* k-points (periodic systems): kpts(3, nK) in units of the reciprocal lattice vectors, weights
 * (normalised here; NULL = equal). nK = 0 returns to the Gamma point in real arithmetic. Each
 * k-point's complex Hermitian problem is solved through its real 2n x 2n embedding; FP64 only.
 * Replaces the k-point loop of the reference (kPoint/kWeight, unpackHS_cmplx_kpts,
 * packHS_cmplx_kpts: src/dftbp/dftb/sparse2dense.F90:185-262, 820-918; Efilling with kWeight:
 * src/dftbp/dftb/etemp.F90:59-194). Call before or after set_geometry. */
int dftbhip_ctx_set_kpoints(dftbhip_ctx* ctx, int nK, const double* kpts, const double* weights);
/* coords(3, nAtom) bohr, species 0-based, latvecs(3,3) rows = lattice vectors or NULL */
int dftbhip_ctx_set_geometry(dftbhip_ctx* ctx, int nAtom, const double* coords,
                             const int* species, const double* latvecs);
/* energies: mermin, total, nonscc, scc, repulsive, TS, Eband, Ef; gross charges q0 - q */
int dftbhip_ctx_scc(dftbhip_ctx* ctx, double* energies, double* charges, int* niter,
                    int* converged);
/* the same cycle one iteration at a time: begin (input charges = reference charges), step
 * (one full iteration incl. mixing; converged flag and max |q_out - q_in|), end (energies and
 * charges of the last iteration, as dftbhip_ctx_scc) */
int dftbhip_ctx_scc_begin(dftbhip_ctx* ctx);
/* warm = 1: every SCC cycle after the first starts from the input charges the previous cycle
 * ended with (on the same atoms), as the reference's geometry loop keeps qInput from one
 * geometry step to the next (src/dftbp/dftbplus/main.F90:219-357, getNextInputCharges
 * :4772-4963); warm = 0 (the default): every cycle starts from the neutral reference charges. */
int dftbhip_ctx_set_charge_restart(dftbhip_ctx* ctx, int warm);
/* Input charges of the next SCC cycle, as ReadInitialCharges = Yes reads them from a charge
 * file (initQFromFile, src/dftbp/dftb/sccinit.F90:185-615, called from
 * src/dftbp/dftbplus/initprogram.F90:4604-4630): qin[nAtom] are the atoms' electron
 * populations (the context's charges are atom resolved); with checkCount != 0 their sum must
 * equal the number of electrons to 1e-3, as the reference checks ("External file of charges
 * has a total charge: ..., instead of ..."); checkCount = 0 skips the check, as its
 * SkipChargeTest = Yes does. Used by the next dftbhip_ctx_scc_begin / dftbhip_ctx_scc only;
 * call after dftbhip_ctx_set_geometry. */
int dftbhip_ctx_set_input_charges(dftbhip_ctx* ctx, int nAtom, const double* qin,
                                  int checkCount);
/* The input charges of the next SCC iteration (the mixed charges after the last step), as the
 * reference writes them to its charge file every iteration (writeCharges, mainio.F90:4319-4370
 * -> writeQToFile, src/dftbp/dftb/sccinit.F90:619-905, called from main.F90:1016): electron
 * populations per atom, qin[nAtom]. */
int dftbhip_ctx_get_input_charges(dftbhip_ctx* ctx, int nAtom, double* qin);
int dftbhip_ctx_scc_step(dftbhip_ctx* ctx, int* converged, double* qerr);
int dftbhip_ctx_scc_end(dftbhip_ctx* ctx, double* energies, double* charges, int* niter,
                        int* converged);
/* State of the last SCC iteration for the SCC block of detailed.out (writeDetailedOut1,
 * src/dftbp/dftbplus/mainio.F90:2794-2926): out[0..n-1] of iteration number, charge error
 * (sccErrorQ), electronic energy, electronic energy of the iteration before (NaN after the
 * first; diffElec is their difference), input electrons of the last iteration, the number
 * of eigenvector pairs outside the first-order range of the last Ogita-Aishima refinement
 * step of the mixed-precision tail and (residual form of the refinement) the number of such
 * pairs beyond the diagonal band of X^T S X in the last refinement, for which the whole
 * X^T S X was formed, and 1 if the tail has been switched off for the rest of the cycle (a tail
 * iteration whose charge error grew more than fourfold past tail_below is repeated with an
 * eigensolve) (the last three have no counterpart in the reference). */
int dftbhip_ctx_scc_state(dftbhip_ctx* ctx, double* out, int n);
/* gradient dE/dR (3, nAtom); needs a converged dftbhip_ctx_scc */
int dftbhip_ctx_gradients(dftbhip_ctx* ctx, double* grad);
int dftbhip_ctx_eigenvalues(dftbhip_ctx* ctx, double* eig, double* fill, int n);
/* eigenvalues and fillings of k-point k (0 without k-points) of the last SCC iteration, the
 * eigen(:, iK, 1) / filling(:, iK, 1) arrays of getFillingsAndBandEnergies
 * (src/dftbp/dftbplus/main.F90:4544-4651); fillings are electrons per level (0..2) */
int dftbhip_ctx_eigenvalues_k(dftbhip_ctx* ctx, int k, double* eig, double* fill, int n);
/* output Mulliken populations per orbital of the last SCC iteration, qOutput of
 * getMullikenPopulation (src/dftbp/dftbplus/main.F90:4653-4753), orbitals atom by atom in the
 * order s, p_y, p_z, p_x */
int dftbhip_ctx_orbital_populations(dftbhip_ctx* ctx, double* q, int n);
int dftbhip_ctx_nr_orbitals(dftbhip_ctx* ctx);
/* per-phase wall times (s) of the last SCC/gradient call and the number of calls */
int dftbhip_ctx_timings(dftbhip_ctx* ctx, double* t, int n);

/* ---- molecule batch: many independent non-periodic molecules in one set of launches -------
 * Replaces a loop over single-geometry calculations (one dftbp_* instance per molecule:
 * dftbp_init / dftbp_process_input / dftbp_get_energy, src/dftbp/api/dftbplus.h:215-550,
 * src/dftbp/api/capi.F90:90-135) with one block-diagonal system: per molecule its
 * own electron count, Fermi level, Broyden mixer and convergence. FP64 only; at most 512
 * orbitals per molecule. coords(3, sum nAtomMol) bohr, species 0-based, molecules one after
 * the other. energies(8, nMol) as dftbhip_ctx_scc; charges(sum nAtomMol) = q0 - q. */
typedef struct dftbhip_batch dftbhip_batch;
int dftbhip_batch_create(dftbhip_batch** batch, const dftbhip_params* p,
                         const dftbhip_options* opt);
int dftbhip_batch_destroy(dftbhip_batch* batch);
int dftbhip_batch_set_geometries(dftbhip_batch* batch, int nMol, const int* nAtomMol,
                                 const double* coords, const int* species);
int dftbhip_batch_scc(dftbhip_batch* batch, double* energies, double* charges, int* niter,
                      int* converged);
/* one SCC iteration at a time (as dftbhip_ctx_scc_begin/step/end); nactive = molecules not yet
 * converged after the iteration */
int dftbhip_batch_scc_begin(dftbhip_batch* batch);
int dftbhip_batch_scc_step(dftbhip_batch* batch, int* nactive);
int dftbhip_batch_scc_end(dftbhip_batch* batch, double* energies, double* charges, int* niter,
                          int* converged);
int dftbhip_batch_gradients(dftbhip_batch* batch, double* grad);

/* ---- host-side pieces of the SCC cycle (no device work, callable without a GPU) ---------
 * Fermi filling of one spin channel over nKPoints k-points, as Efilling with the Fermi
 * distribution (src/dftbp/dftb/etemp.F90:59-194: zero-electron and full-band shortcuts, middle
 * gap for integer counts, bisection, final rescale to the exact count). eigenvals(nLevels,
 * nKPoints) ascending per k-point, kWeight(nKPoints); filling(nLevels, nKPoints) in [0, 1];
 * Ebs, TS, E0 k-weighted. The SCC cycle calls it with half the electron count and doubles the
 * results (getFillingsAndBandEnergies, src/dftbp/dftbplus/main.F90:4544-4651). */
int dftbhip_efilling(int nLevels, int nKPoints, const double* eigenvals, const double* kWeight,
                     double nElectrons, double kT, double* filling, double* Ef, double* Ebs,
                     double* TS, double* E0);
/* Modified Broyden mixer, TBroydenMixerReal (src/dftbp/mixer/broydenmixer.F90:112-342):
 * create = TBroydenMixerReal_init (maxSccIter, mixParam, omega0, minWeight, maxWeight,
 * weightFac), reset = reset(nElem), mix = mix(qInpResult, qDiff): qInpResult holds the input
 * charges on entry and the mixed charges on exit. More than maxSccIter mix calls after a reset
 * is an error ("Broyden mixer: Maximal nr. of steps exceeded"), as in the reference. */
typedef struct dftbhip_mixer dftbhip_mixer;
int dftbhip_mixer_create(dftbhip_mixer** mixer, int maxSccIter, double mixParam, double omega0,
                         double minWeight, double maxWeight, double weightFac);
int dftbhip_mixer_reset(dftbhip_mixer* mixer, int nElem);
int dftbhip_mixer_mix(dftbhip_mixer* mixer, double* qInpResult, const double* qDiff);
int dftbhip_mixer_destroy(dftbhip_mixer* mixer);

/* ---- measurement: per-launch HIP-event timing of the tridiagonalisation's symmetric
 *      product (the HBM-bound kernel of the eigensolver). stats: out[0] launches, out[1] summed
 *      duration (s), out[2] algorithmic bytes (lower triangle of each trailing matrix).
 *      (no counterpart in the reference) */
int dftbhip_kernel_profile(int enable);
int dftbhip_kernel_stats(double* out);

/* ---- diagnostics: accumulator lane layouts of the FP64 and FP32 16x16x4 MFMAs (one wave
 *      each), out[512]: FP64 in out[0:256], FP32 in out[256:512] ----------------------------
 * (no counterpart in the reference; used by the tests to pin the MFMA tiling) */
int dftbhip_probe_mfma_f64(double* out);
/* the same for the FP16 16x16x32 MFMA (FP32 accumulator), out[256] */
int dftbhip_probe_mfma_f16(double* out);
/* FP16 product on host arrays, c = alpha a^T b: a (k x m) and b (k x n) FP16 bit patterns, c
 * (m x n) FP32, all column-major; the kernel of the refinement's correction products
 * (no counterpart in the reference; used by the tests) */
int dftbhip_gemm_f16_tn(int m, int n, int k, float alpha, const uint16_t* a, int lda,
                        const uint16_t* b, int ldb, float* c, int ldc);

#ifdef __cplusplus
}
#endif

#endif
