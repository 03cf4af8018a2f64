/*
 * semigpu.h -- C ABI of the MI355X-native MNDO SCF engine (libsemigpu.so).
 *
 * Plain C types only: ints, doubles, opaque handles.  Every entry point
 * returns SG_OK (0) or a negative SG_ERR_* code; sg_last_error() describes the
 * last failure of the calling thread.  Units follow the reference at its
 * object boundary: atom coordinates in Angstrom, matrices and energies in eV,
 * packed symmetric matrices as the lower triangle in row order (index
 * r(r+1)/2 + c, mymemory.cpp:363-371), eigenvectors column-major
 * (scfcalculation.cpp:175-179).
 *
 * A batch holds any number of molecules; a single molecule is a batch of one.
 * Each entry point below names the reference member it replaces
 * (LuisAlfredoNu/semiGPU, src/common/...).  One HIP stream per context;
 * a context is not thread-safe, separate contexts are.
 */
#ifndef SEMIGPU_H
#define SEMIGPU_H

#ifdef __cplusplus
extern "C" {
#endif

#define SG_OK 0
#define SG_ERR_INVALID (-1)           /* bad argument / handle */
#define SG_ERR_UNSUPPORTED (-2)       /* element outside H, C, N, O */
#define SG_ERR_HIP (-3)               /* HIP runtime failure */
#define SG_ERR_ORDER (-4)             /* stage called before its inputs exist */
#define SG_ERR_NO_DEVICE (-5)         /* no usable GPU */

enum sg_matrix { SG_HCORE = 0, SG_FOCK = 1, SG_DENSITY = 2 };

typedef struct sg_context sg_context;
typedef struct sg_batch sg_batch;

typedef struct {
  double e_el;    /* SCFCalculation::finalEnergy (scfcalculation.h:22), eV */
  double e_cc;    /* CoreCoreRepulsion::ComputeRepulsion, eV */
  double e_tot;   /* e_el + e_cc (semiGPU.cc:95) */
  int steps;      /* SCF iterations performed */
  int converged;  /* SCFCalculation::energyConverge (scfcalculation.h:24) */
  int status;     /* 0, or >0 when the eigensolver failed (dsyev info > 0) */
} sg_scf_result;

const char* sg_last_error(void);
int sg_version(void);

/* Context = one GPU + one stream. */
int sg_ctx_create(int device, sg_context** out);
int sg_ctx_destroy(sg_context* ctx);
/* Basis size from which the molecules of batches created afterwards take the
 * multi-workgroup eigensolver (one molecule at a time over the whole GPU)
 * instead of one workgroup per molecule.  Default 1281: exactly the molecules
 * above the one-workgroup limit of 1280.  Values from 128 up are accepted, so
 * that the two solvers can be compared on the same molecules.  The largest
 * basis a batch accepts is 8192.  No counterpart in the reference, whose
 * dsyev_ call (scfcalculation.cpp:169-203) takes any size on the host. */
int sg_ctx_set_large_basis(sg_context* ctx, int n_min);
/* Autotuner of the one-workgroup eigensolver's FP64 reduction (no
 * counterpart in the reference).  The reduction holds each molecule's
 * trailing block (up to 256 x 256) in the registers of one workgroup, launched
 * in SG_TUNE_BUCKETS size buckets (N above 224, (192, 224], (160, 192],
 * (128, 160], (96, 128], (64, 96], (32, 64], up to 32); a bucket can run in up
 * to SG_TUNE_CANDIDATES launch shapes (wavefronts per molecule x register
 * tiles per wavefront), which give the same results bit for bit.
 * sg_ctx_autotune times every shape of every bucket on `molecules` random
 * symmetric matrices of the bucket's largest N (best of two launches after a
 * warmup) and makes the fastest the context's choice for batches and servers
 * created afterwards (and the running servers' next iterations).
 * ms[SG_TUNE_BUCKETS * SG_TUNE_CANDIDATES] gets the times (-1: no such shape)
 * and choice[SG_TUNE_BUCKETS] the choices (either may be null).  Without a
 * call the first shape of each bucket is used.  sg_ctx_set_tri_variant sets one
 * bucket's shape; sg_ctx_tri_shapes reports the choices (variant[]), the
 * number of shapes per bucket (count[]) and their wavefronts and register
 * tiles (nw[], nr[], SG_TUNE_CANDIDATES per bucket). */
#define SG_TUNE_BUCKETS 8
#define SG_TUNE_CANDIDATES 3
int sg_ctx_autotune(sg_context* ctx, int molecules, double* ms, int* choice);
int sg_ctx_set_tri_variant(sg_context* ctx, int bucket, int variant);
int sg_ctx_tri_shapes(const sg_context* ctx, int* variant, int* count, int* nw, int* nr);
/* Whether sg_scf_run / sg_scf_run_policy capture one SCF iteration of a batch
 * as a HIP graph on first use and replay it (on, the default), or launch each
 * iteration's kernels directly (off: no capture and instantiation per batch,
 * the cheaper choice for batches that run once, e.g. a single molecule).  The
 * results are the same either way.  No counterpart in the reference. */
int sg_ctx_set_graphs(sg_context* ctx, int on);
/* Whether sg_scf_run / sg_scf_run_policy iterate a batch whose molecules lie
 * on both sides of 256 basis functions as two groups on two streams (on, the
 * default: the large molecules' HBM-bound reduction overlaps the small
 * molecules' on-chip one), or as one group on the context's stream (off: each
 * kernel runs alone on the device, e.g. to time it).  The results are the
 * same either way.  No counterpart in the reference. */
int sg_ctx_set_overlap(sg_context* ctx, int on);

/* Molecules: atoms of molecule m are atom_offsets[m] .. atom_offsets[m+1]-1 of
 * Z[] / xyz_ang[3*i..3*i+2].  Replaces ReadXYZFile + ListAtomicOrbitals::SetOrbitals
 * (readxyzfile.cpp:27-55, atomicOrbitals.cpp:76-119) and the allocations of
 * TwoCenterIntegral::Alloc4AllTwoCenterIntegral (twocenterintegral.cpp:55-107),
 * BaseMatrix::Alloc4Matrix (basematrix.cpp:23-27) and
 * SCFCalculation::AllocSCFData (scfcalculation.cpp:28-46). */
int sg_batch_create(sg_context* ctx, int n_mol, const int* atom_offsets, const int* Z,
                    const double* xyz_ang, sg_batch** out);
/* Same, with coordinates already in bohr (the AtomicOrbital::coordinates of
 * atomicOrbitals.h:24, as the reference's kernels see them). */
int sg_batch_create_bohr(sg_context* ctx, int n_mol, const int* atom_offsets, const int* Z,
                         const double* xyz_bohr, sg_batch** out);
int sg_batch_destroy(sg_batch* b);
int sg_batch_molecule_info(const sg_batch* b, int mol, int* n_atoms, int* n_ao, int* n_occ);

/* TwoCenterIntegral::ComputeAllTwoCenterIntegral + To_device
 * (twocenterintegral.cpp:23-53, 229-279). */
int sg_two_center(sg_batch* b);
/* Hcore::ComputeMatrix (hcore.cpp:104-124 / basematrix.cpp:39-55). */
int sg_hcore(sg_batch* b);
/* CoreCoreRepulsion::ComputeRepulsion (corecorerepulsion.cpp:12-41); e_cc[n_mol]. */
int sg_core_core(sg_batch* b, double* e_cc);
/* DensityMatrix guess (GuessDensityMatrixSwitch(true) + ComputeMatrix,
 * densitymatrix.cpp:26-36). */
int sg_guess_density(sg_batch* b);
/* FockMatrix::ComputeMatrix (fockmatrix.cpp:32-53) for every molecule. */
int sg_fock(sg_batch* b);
/* SCFCalculation::GetEigenValVecOfFmatrix, i.e. dsyev_('V','L') on F
 * (scfcalculation.cpp:169-203). */
int sg_eigh(sg_batch* b);
/* The full eigendecomposition of sg_eigh for one symmetric n x n matrix in
 * device memory, without allocating: no counterpart in the reference (whose
 * dsyev_ call, scfcalculation.cpp:169-203, allocates its own workspace); for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops).  a_dev: n x n row-major (only the lower triangle is
 * read); eigval_dev[n]: ascending eigenvalues; eigvec_dev[n*n]: eigenvectors
 * column-major.  All three are device pointers of the caller's.  The kernels
 * are sg_eigh's; their workspace is one fixed device array per GPU held by the
 * library for the life of the process (about 10.5 MB), so n is at most 512
 * and larger n is rejected with SG_ERR_INVALID.  Every call on a GPU uses that
 * GPU's array: calls that overlap on one GPU (other streams, other contexts,
 * other host threads) race, and must be ordered by the caller.  The work is
 * queued on `stream` (a hipStream_t; NULL = the context's stream) and the call
 * returns without waiting; it calls no allocation routine. */
int sg_eigh_noalloc(sg_context* ctx, int n, const double* a_dev, double* eigval_dev,
                    double* eigvec_dev, void* stream);
/* DensityMatrix::ComputeMatrix from the eigenvectors (densitymatrix.cpp:38-45). */
int sg_density(sg_batch* b);
/* ElectronicEnergy::ComputeEnergy (electronicenergy.cpp:12-36); e_el[n_mol]. */
int sg_energy(sg_batch* b, double* e_el);
/* SCFCalculation::ComputeSCF (scfcalculation.cpp:66-146) plus the core-core
 * energy, for every molecule of the batch; runs sg_two_center / sg_hcore first
 * when they have not run.  threshold_ev = 5e-3 and max_steps = 100 reproduce
 * the reference.  results[n_mol]. */
int sg_scf_run(sg_batch* b, double threshold_ev, int max_steps, sg_scf_result* results);

/* Same SCF with a precision policy (no counterpart in the reference, which is
 * FP64 throughout).  SG_POLICY_FP64 is sg_scf_run.  The mixed policies run
 * the early steps of each molecule at reduced precision, until its
 * |dE| < switch_ev (1e-3 eV suggested), or |dE| falls below what the reduced
 * arithmetic resolves on its energy (2^-20 |E_el|, 8 FP32 ulps; 2^-14 |E_el|
 * for split BF16), or |dE| stops falling by half per step, or 40 reduced
 * steps; then in FP64 to convergence, which is only tested between two FP64
 * steps.  Integrals, Hcore, the energy sums and the multi-workgroup
 * eigensolver are always FP64.
 *   SG_POLICY_MIXED        Fock contractions in FP32 (over an FP32 copy of the
 *                          integrals) and the density product in FP32 on the
 *                          matrix cores; the eigensolver stays FP64.
 *   SG_POLICY_MIXED_BF16   Fock contractions in FP32; the density product as
 *                          split BF16 on the matrix cores (each FP64 vector
 *                          entry y = hi + lo with hi, lo BF16; hi hi + hi lo +
 *                          lo hi accumulated in FP32); the eigensolver stays
 *                          FP64.
 *   SG_POLICY_MIXED_EIG32  as SG_POLICY_MIXED, and in addition the Householder
 *                          reduction and back-transformation of the
 *                          one-workgroup eigensolver (N <= 1280) run in FP32,
 *                          with bisection to 1e-10 ||T|| and one
 *                          inverse-iteration solve (factors stored in FP32).
 *                          Opt-in: it departs from an FP64 eigensolver.
 *                          Every policy gives a molecule the same result alone
 *                          and in any batch, bit for bit. */
#define SG_POLICY_FP64 0
#define SG_POLICY_MIXED 1
#define SG_POLICY_MIXED_BF16 2
#define SG_POLICY_MIXED_EIG32 3
int sg_scf_run_policy(sg_batch* b, double threshold_ev, int max_steps, int policy,
                      double switch_ev, sg_scf_result* results);
/* Density damping for the batch's later sg_scf_run / sg_scf_run_policy calls
 * (no counterpart in the reference, whose SCF takes each new density as it
 * is; opt-in for oscillating cases): every step's density becomes
 * beta P_old + (1 - beta) P_new before the Fock build, beta in [0, 1), 0 (the
 * default) = off.  The converged energy is the same fixed point; the server
 * does not damp. */
int sg_scf_set_damping(sg_batch* b, double beta);
/* Pulay DIIS for the batch's later sg_scf_run / sg_scf_run_policy calls (no
 * counterpart in the reference, whose SCF diagonalises each new Fock matrix
 * as it is; opt-in for slow or oscillating cases): after each step's
 * convergence test, a molecule that goes on iterating stores F and its error
 * F P - P F (the last `depth` of each), and the next eigensolve takes
 * sum_j c_j F_j, c minimising |sum_j c_j e_j| with sum_j c_j = 1.  depth 0
 * (the default) = off, else 2..8; the histories take 2 depth sum N^2 doubles,
 * allocated here.  Under SG_POLICY_MIXED the FP32 steps extrapolate over
 * FP32 steps only, and the history restarts at the first FP64 step.  A
 * converged molecule keeps the Fock matrix of its final density; for the
 * server, sg_server_set_diis. */
int sg_scf_set_diis(sg_batch* b, int depth);
/* Per-iteration energies (opt-in; no counterpart in the reference, whose SCF
 * keeps only the last two energies, scfcalculation.cpp:82-146).  With cap > 0,
 * the batch's later sg_scf_run / sg_scf_run_policy calls record every
 * molecule's E_el of the guess (entry 0) and of each step k < cap (entry k);
 * cap = 0 frees the record.  It takes cap doubles per molecule. */
int sg_scf_record_energies(sg_batch* b, int cap);
/* The record of molecule mol into out (cap doubles): E_el(0) .. E_el(steps),
 * *count = min(steps + 1, cap) entries; SG_ERR_ORDER without a record. */
int sg_scf_energy_history(const sg_batch* b, int mol, double* out, int* count);
/* FP32 steps of each molecule in the last sg_scf_run_policy (0 under
 * SG_POLICY_FP64); steps32 gets n_mol entries.  No counterpart in the
 * reference: it splits the steps for per-precision rates. */
int sg_scf_fp32_steps(const sg_batch* b, int* steps32);

/* Per-molecule eigensolver status of the last sg_eigh / SCF step
 * (status[n_mol]; > 0 where dsyev would report info > 0). */
int sg_get_status(const sg_batch* b, int* status);

/* Heat of formation (kcal/mol) of a molecule from its SCF total energy (eV):
 * (E_tot - sum E_isol) x 23.060529 + sum dH_f(atom), with the atomic tables
 * the reference carries unused (heat, eisol: MNDO_parameters.h:40-41,
 * MNDO_parameters.cpp:58-121).  Host arithmetic only (no device); no
 * counterpart in the reference, which stops at E_tot (semiGPU.cc:95). */
int sg_heat_of_formation(int n_atoms, const int* Z, double e_tot_ev, double* dhf_kcal);
/* Nuclear gradient of the MNDO total energy by central differences (no
 * counterpart in the reference, whose README lists gradients and a geometry
 * optimiser as to do; analytic gradients are not built): one batch of the
 * molecule and its 6 n_atoms displaced copies (each atom moved by +-step_ang
 * along x, y, z), FP64 SCF to threshold_ev (200 steps at most; DIIS of
 * diis_depth, 0 = off), grad[3 a + k] = (E(+) - E(-)) / (2 step) in eV/A.
 * The truncation error is O(step^2), the SCF's is about threshold / step:
 * step 1e-3 A with threshold 1e-9 eV gives about 1e-6 eV/A.  e_tot (may be
 * null) gets the energy of the undisplaced geometry, all_converged (may be
 * null) whether every SCF converged. */
int sg_gradient_fd(sg_context* ctx, int n_atoms, const int* Z, const double* xyz_ang,
                   double step_ang, double threshold_ev, int diis_depth, double* e_tot,
                   double* grad, int* all_converged);

/* Analytic nuclear gradient of the MNDO total energy (E_el + E_cc) of every
 * molecule of the batch at its current density, i.e. after sg_scf_run /
 * sg_scf_run_policy has converged it: the derivatives of every atom pair's
 * two-center integrals, overlaps and core-core term at fixed density (exact
 * for a converged NDDO SCF, whose basis is orthonormal by assumption; the
 * error is first order in the SCF residual, about 1e-6 eV/A after an energy
 * test of 1e-12 eV, about 1e-8 after further steps with threshold 0).
 * grad[3 i + k] for atom i of the batch (the atom_offsets order), axis k, in
 * eV/Angstrom.  No counterpart in the reference, whose README lists gradients
 * and a geometry optimiser as to do (README.md:109-110). */
int sg_gradient(sg_batch* b, double* grad);

/* Host copies for the reference's host-side views. */
long sg_pair_block_count(const sg_batch* b, int mol);
/* double**** all2CenterIntegral contents in the reference jagged order
 * [i][j<=i][k][l] (twocenterintegral.cpp:55-107); read by GetValueFromArray. */
int sg_get_pair_blocks(const sg_batch* b, int mol, double* out);
/* The reverse copy (TwoCenterIntegral::To_device, twocenterintegral.cpp:23-53):
 * host pair blocks become the device's (FP64 and the FP32 copy). */
int sg_set_pair_blocks(sg_batch* b, int mol, const double* in);
/* BaseMatrix::matrixHold_ of Hcore / FockMatrix / DensityMatrix (packed). */
int sg_get_packed(const sg_batch* b, int mol, int which, double* out);
int sg_set_packed(sg_batch* b, int mol, int which, const double* in);
/* SCFCalculation::eigenVal / eigenVec (eigenvectors column-major, n_ao*n_ao). */
int sg_get_eigen(const sg_batch* b, int mol, double* eigval, double* eigvec);
/* Uploads eigenvalues and/or eigenvectors (either may be NULL), so that
 * sg_density builds P from a host eigenVec as DensityMatrix does
 * (densitymatrix.h:13, the eigenVec pointer it aliases). */
int sg_set_eigen(sg_batch* b, int mol, const double* eigval, const double* eigvec);

/* Continuous-batching SCF server (no counterpart in the reference, whose main()
 * runs one molecule per process, semiGPU.cc:30-110).  Requests (one molecule
 * each) go to the size class of their basis size N and queue there in arrival
 * order.  Each class has its own pools, stream and SCF iterations, so that
 * small molecules are not paced by large ones; their kernels overlap on the GPU.
 * sg_server_step() visits every class whose last iteration has finished: it
 * retires the molecules that converged, reached max_steps or failed, admits
 * queued molecules while their HBM footprint fits the class's pools (first
 * fit) and a slot is free, sets them up (integrals, Hcore, core-core, guess,
 * E_0) and launches ONE SCF iteration for every molecule in flight in the
 * class.  State stays in HBM between steps.  A molecule's result does not
 * depend on what else is in flight.
 *   hbm_bytes  device memory the server may hold (one pool for H, F, P, C,
 *              workspace, pair blocks and vectors, shared by the classes)
 *   max_slots  molecules in flight at most;  max_atoms  atoms in flight at most
 *              (both split evenly between the classes)
 *   policy     SG_POLICY_FP64 / SG_POLICY_MIXED (switch_ev as for sg_scf_run_policy)
 * Times are seconds on the server's clock (sg_server_clock, 0 at creation). */
typedef struct sg_server sg_server;
typedef struct {
  long long request_id;
  double e_el, e_cc, e_tot;      /* eV, as sg_scf_result */
  int steps, converged, status;
  int n_ao;
  double t_submit, t_admit, t_done;
} sg_served;
int sg_server_create(sg_context* ctx, double hbm_bytes, int max_slots, int max_atoms,
                     double threshold_ev, int max_steps, int policy, double switch_ev,
                     sg_server** out);
/* Same with n_classes size classes: class i takes N in [bounds[i-1], bounds[i])
 * (bounds[-1] = 0, the last class up to the largest N the server accepts, 1280);
 * bounds[] has n_classes - 1 increasing entries.  sg_server_create is one class.
 * Each class has its own stream; HIP maps streams onto GPU_MAX_HW_QUEUES
 * hardware queues per process (4 by default) and streams that share one run
 * in sequence, so with more than one or two classes raise it (8 gives each
 * class its own queue). */
int sg_server_create_classes(sg_context* ctx, double hbm_bytes, int max_slots, int max_atoms,
                             double threshold_ev, int max_steps, int policy, double switch_ev,
                             int n_classes, const int* bounds, sg_server** out);
int sg_server_destroy(sg_server* srv);
/* Pulay DIIS for the server's SCFs (no counterpart in the reference; the
 * batch's sg_scf_set_diis for a stream): depth 0 = off (the default), else
 * 2..8.  Allocates 2 depth doubles per element of the server's matrix pool
 * beyond hbm_bytes.  Only while no request is queued or in flight. */
int sg_server_set_diis(sg_server* srv, int depth);
/* Queues one molecule (Angstrom); rejects elements outside H, C, N, O and
 * molecules larger than the server's pools. */
int sg_server_submit(sg_server* srv, long long request_id, int n_atoms, const int* Z,
                     const double* xyz_ang);
/* One retirement + admission + SCF-iteration round for every class that is
 * ready (does not wait for the others; sleeps ~100 us when no class was ready);
 * reports the molecules in flight and still queued afterwards. */
int sg_server_step(sg_server* srv, int* n_inflight, int* n_queued);
/* Moves up to max_out finished results (completion order) into out[];
 * returns how many, or a negative error code. */
int sg_server_poll(sg_server* srv, sg_served* out, int max_out);
double sg_server_clock(const sg_server* srv);

/* Per-stage GPU time (HIP events on the context stream), milliseconds summed
 * since the last reset; stage order: two_center, hcore, guess, fock, tridiag,
 * bisect, invit, backtransform (the eigensolver's kernels), density, energy,
 * control, tridiag_fp32, tridiag_head.  tridiag is the on-chip FP64 reduction
 * of each eigensolve (its size-bucketed launches, plus the multi-workgroup
 * molecules); tridiag_head the first N - 256 columns of the molecules above
 * N = 256, which precede it; tridiag_fp32 the FP32 reduction of the
 * mixed_eig32 policy.  counts[] gets the number of passes timed. */
#define SG_N_STAGES 13
/* One plain device-to-device copy kernel (k_calibration_copy) over freshly
 * allocated buffers: reads and writes exactly `bytes` (a multiple of 16).  No
 * counterpart in the reference: it is the known-size kernel that hardware byte
 * counters (FETCH_SIZE, WRITE_SIZE) are scaled by. */
int sg_calibration_copy(sg_context* ctx, long long bytes);
int sg_timing_enable(sg_context* ctx, int on);
int sg_timing_read(sg_context* ctx, double* ms, int* counts);
int sg_timing_reset(sg_context* ctx);

#ifdef __cplusplus
}
#endif
#endif /* SEMIGPU_H */
