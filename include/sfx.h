/*
 * sfx.h -- C ABI of the MI355X (gfx950) Scheutjens-Fleer hot path.
 *
 * Plain C, no HIP or torch types in any signature. Every entry point returns an
 * int status (SFX_OK == 0, negative on error) and leaves a message for
 * sfx_last_error(). Arrays are caller-owned; device buffers are owned by the
 * handle that allocated them.
 *
 * Site layout (identical to the reference's lattice vectors, 0-based here):
 *   a lattice with g gradients and n_layers[k] interior layers keeps one halo
 *   layer on each side of every used gradient: MX = n_layers[0]+2,
 *   MY = n_layers[1]+2 (2-D/3-D, else 1), MZ = n_layers[2]+2 (3-D, else 1),
 *   M = MX*MY*MZ and site (x,y,z) sits at (x*MY + y)*MZ + z, so the last used
 *   gradient is contiguous (Lat1DFlat.cpp:25, Lat2DFlat.cpp:44-46,
 *   Lat3D.cpp:59-68). Interior coordinates run 1..n_layers[k].
 *
 * Reference interfaces replaced (file:line in HaripriyaGnanasekaran/field):
 *   - Cuda_tools.h:3-26 (flat C GPU API): sfx_lattice_* primitives below.
 *   - Lat1stO.h:19-50 / Lat2ndO.h:73-119 (virtual lattice API): the same
 *     primitives, called per rank by the molecule drivers.
 *   - SF_Copol1stO/SF_Homopol1stO/SF_Monomer1stO ::ComputePhi (Matrix1,
 *     SymMatrix1, Matrix1Long) and SF_MolList1stO.cpp:246-295: sfx_compute_phi.
 *   - SF_Solve::residuals / GetResiduals (SF_Solve.cpp:527-591, 647-656) and
 *     MPS.cpp:187-198 ComputeG: sfx_residual.
 *   - SF_MolList1stO::GetFreeEnergy (SF_MolList1stO.cpp:606-635) and
 *     MPS.cpp:253-261 Helmholtz: sfx_free_energy.
 *   - SFNewton::iterate / PlainLBFGSDIIS (sfnewton.cpp:566-664, 844-1143) and
 *     MPS.cpp:228-284 DIIS/SCF: sfx_solve.
 */
#ifndef SFX_H
#define SFX_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SFX_ABI_VERSION 3

/* ---- status codes ------------------------------------------------------ */
enum {
  SFX_OK = 0,
  SFX_E_INVALID = -1,      /* bad argument or descriptor */
  SFX_E_OOM = -2,          /* device allocation failed */
  SFX_E_UNSUPPORTED = -3,  /* combination not implemented */
  SFX_E_NONFINITE = -4,    /* NaN/Inf produced (e.g. ln of a non-positive G_N) */
  SFX_E_HIP = -5,          /* HIP runtime error */
  SFX_E_NOTCONVERGED = -6  /* solver hit its iteration limit */
};

/* ---- lattice ----------------------------------------------------------- */
enum { SFX_FLAT = 0, SFX_CYLINDER = 1, SFX_SPHERE = 2 };
enum { SFX_LAT_STANDARD = 0, SFX_LAT_FCC = 1, SFX_LAT_HEX = 2 };
/* Boundary choices, numbered as the reference's input choices
 * (Lat1DFlat.cpp:26-31: surface, mirror1, mirror2, periodic). 3-D lattices
 * accept mirror1/mirror2/periodic only (Lat3D.cpp:69-73). */
enum {
  SFX_BOUND_SURFACE = 1,
  SFX_BOUND_MIRROR1 = 2,
  SFX_BOUND_MIRROR2 = 3,
  SFX_BOUND_PERIODIC = 4
};

typedef struct sfx_lattice_desc {
  int gradients;          /* 1, 2 or 3 */
  int geometry;           /* SFX_FLAT / SFX_CYLINDER / SFX_SPHERE (curved: 1-D, 2-D cyl) */
  int lattice_type;       /* SFX_LAT_STANDARD / SFX_LAT_FCC / SFX_LAT_HEX (3-D) */
  int order;              /* 1: flexible chains, 2: semiflexible (direction-resolved) */
  int n_layers[3];        /* interior layers per gradient */
  int lowerbound[3];      /* SFX_BOUND_* per gradient */
  int upperbound[3];
  double lambda;          /* 1-D/2-D step weight l1 (required, Lat1DFlat.cpp:109) */
  double lambda_chi;      /* contact-average weight; <= 0 selects lambda */
  double layer_adjustment;   /* 1-D: default 1 (Lat1DFlat.cpp:139) */
  double offset_first_layer; /* curved geometries: default 0 */
} sfx_lattice_desc;

/* ---- system (one SCF problem on one lattice) ---------------------------- */
enum { SFX_FREE = 0, SFX_RESTRICTED = 1, SFX_SOLVENT = 2, SFX_RANGE_RESTRICTED = 3 };

typedef struct sfx_molecule_desc {
  int length;             /* chain length N */
  const int* sequence;    /* N state indices (0-based); rank s=1..N uses sequence[s-1] */
  int freedom;            /* SFX_FREE (phibulk), SFX_RESTRICTED (theta), SFX_SOLVENT,
                             SFX_RANGE_RESTRICTED (phi_range in range_lo..range_hi) */
  double phibulk;         /* SFX_FREE: bulk volume fraction */
  double theta;           /* SFX_RESTRICTED: amount, sum over interior sites of phi */
  /* SFX_RANGE_RESTRICTED (freedom range_restricted, SF_Molecule.cpp:169-180;
   * SF_Copol1stO.cpp:1246-1280): the molecule's phi, summed over its states
   * and over the interior sites of the box range_lo..range_hi (1-based
   * coordinates per gradient, inclusive; unused axes ignored), equals
   * phi_range; theta and phi_b follow from it. Residuals and solves by DIIS
   * and SFNewton; JVP and VJP return SFX_E_UNSUPPORTED. */
  int range_lo[3];
  int range_hi[3];
  double phi_range;
} sfx_molecule_desc;

/* Residual / bookkeeping conventions:
 *  SFX_FORM_SFBOX: SF_Solve::GetResiduals with solvent closure
 *                  (SF_MolList1stO.cpp:280-292); x is the SF_Solve layout:
 *                  for every iteration site (interior, not frozen) in site
 *                  order, one value per distinct iteration variable.
 *  SFX_FORM_MELT:  MPS.cpp incompressible melt: no solvent, explicit bulk
 *                  fractions in the chi term (melt_phibulk), x is
 *                  [state][M] with halo entries zero (MPS.cpp:309-320). */
enum { SFX_FORM_SFBOX = 0, SFX_FORM_MELT = 1 };

typedef struct sfx_system_desc {
  int n_states;                   /* segment types, one state each */
  const double* chi;              /* n_states*n_states, row-major, symmetric */
  const int* frozen;              /* n_states flags (NULL: none frozen) */
  const double* frozen_profile;   /* n_states*M range values of frozen states (rows of
                                     liquid states ignored); NULL if none frozen */
  int n_molecules;
  const sfx_molecule_desc* molecules;
  double stiffness;               /* 2nd order: straight-bond weight S; L=(1-S)/(2d-2) */
  int form;                       /* SFX_FORM_SFBOX / SFX_FORM_MELT */
  const double* melt_phibulk;     /* SFX_FORM_MELT: n_states bulk fractions */
} sfx_system_desc;

/* Propagator precision of a sweep. FP32 carries a per-rank log-scale. */
enum { SFX_FP64 = 0, SFX_FP32 = 1 };

typedef struct sfx_options {
  int precision;          /* default precision of sfx_residual/sfx_compute_phi */
  size_t memory_budget;   /* bytes for the stored propagator; 0 = unlimited.
                             A smaller budget selects sqrt(N) checkpointing. */
  int sweep_kernel;       /* first-order 3-D simple-cubic sweeps: SFX_SWEEP_AUTO,
                             SFX_SWEEP_PER_RANK (one launch per rank) or
                             SFX_SWEEP_BLOCKED (up to four ranks per launch through
                             LDS, where every side exceeds four layers); on 1-D
                             first-order lattices of at most 4000 layers, AUTO and
                             BLOCKED run each FP64 stretch of a chain's ranks in
                             one workgroup launch (same bits as per rank) */
  int chain_schedule;     /* SFX_SCHEDULE_AUTO: a first-order chain without mirror
                             symmetry, fully stored, on a blocked lattice advances its
                             forward and backward chains in the same launches (each
                             stores half and composes the other half); elsewhere, and
                             with SFX_SCHEDULE_SEQUENTIAL, the forward chain is stored
                             and the backward chain composes with it
                             (SF_Copol1stO.cpp:1186-1281). The two schedules agree to
                             rounding; a checkpointed store repeats the sequential one
                             bit for bit. No counterpart in the reference. */
  int transfer_kernel;    /* second order, FP32 sweeps: how the 6x6 direction transfer
                             (Lat3D2ndO.cpp:64-152, 597-678) runs at every site:
                             SFX_TRANSFER_ROWS, its structured form S in_k +
                             L (sum - in_k - in_5-k) on the vector units;
                             SFX_TRANSFER_MFMA, the dense 6x6 product on the matrix
                             cores (v_mfma_f32_4x4x1f32, each lane keeping its site);
                             SFX_TRANSFER_AUTO, the matrix cores in the tangent sweeps
                             and the rows in the base sweeps (each the faster of the two
                             there on MI355X). The two agree to FP32 rounding; FP64
                             sweeps use the rows. No counterpart in the reference. */
} sfx_options;

enum { SFX_SWEEP_AUTO = 0, SFX_SWEEP_PER_RANK = 1, SFX_SWEEP_BLOCKED = 2 };
enum { SFX_SCHEDULE_AUTO = 0, SFX_SCHEDULE_SEQUENTIAL = 1 };
enum { SFX_TRANSFER_AUTO = 0, SFX_TRANSFER_ROWS = 1, SFX_TRANSFER_MFMA = 2 };

typedef struct sfx_system sfx_system;

/* Pointer-location flag for array arguments. */
enum { SFX_HOST = 0, SFX_DEVICE = 1 };

/* ---- library ----------------------------------------------------------- */
int sfx_abi_version(void);
/* Copies the last error message of the calling thread into buf. */
int sfx_last_error(char* buf, size_t size);
/* Selects the HIP device for handles created afterwards by this thread. */
int sfx_set_device(int device);
int sfx_device_count(int* count);

/* ---- system lifecycle -------------------------------------------------- */
int sfx_system_create(const sfx_lattice_desc* lat, const sfx_system_desc* sys,
                      const sfx_options* opt, sfx_system** out);
int sfx_system_destroy(sfx_system* s);
/* Number of iteration variables (length of x and f) and of lattice sites M. */
int sfx_system_sizes(const sfx_system* s, long* nvar, long* M);

/* ---- hot path ---------------------------------------------------------- */
/* u -> SWF -> propagators -> phi, ln G_N (UpdateSWF + MolQ->ComputePhi). */
int sfx_compute_phi(sfx_system* s, const double* x, int where, int precision);
/* One residual evaluation f(x) (SF_Solve::residuals / MPS ComputeG). */
int sfx_residual(sfx_system* s, const double* x, double* f, int where, int precision);
/* Jacobian-vector product J(x) v of the residual, by tangent propagators. */
int sfx_jvp(sfx_system* s, const double* x, const double* v, double* jv, int where,
            int precision);
/* Transposed product J(x)^T w. The composition of a chain is the derivative
 * of its partition function wherever the backward sweep is the forward sweep's
 * adjoint in the site-volume inner product, so the transposed sweep is the
 * tangent sweep with a rescaled direction. SFX_E_UNSUPPORTED where that does
 * not hold: a curved lattice with a mirror2 side, the second-order 2-D
 * cylinder (the reference's backward sweep does not split its outward step,
 * Lat2DCyl2ndO.cpp:223-227), a 1-D flat layer adjustment other than 1, a
 * mirror side on a HEX lattice (its neighbour set is not reflection
 * symmetric), a second-order 1-D or 2-D lattice with a mirror2 lower bound
 * (copied without reversing the components, as the reference does). */
int sfx_vjp(sfx_system* s, const double* x, const double* w, double* jtw, int where,
            int precision);
/* Results of the last sfx_compute_phi / sfx_residual. */
int sfx_get_phi(sfx_system* s, int state, double* phi_out /* M */, int where);
int sfx_get_lngn(sfx_system* s, int molecule, double* lngn);
int sfx_get_phibulk(sfx_system* s, int molecule, double* phibulk);
int sfx_free_energy(sfx_system* s, double* F);
/* Grand potential (SF_MolList1stO.cpp:637-647 over the profile of :1001-1100)
 * and chemical potential of a molecule (SF_MolList1stO.cpp:389-452), both of
 * the last sfx_compute_phi / sfx_residual, for linear chains without charges,
 * grafting or internal states; the melt form returns SFX_E_UNSUPPORTED. */
int sfx_grand_potential(sfx_system* s, double* omega);
int sfx_chemical_potential(sfx_system* s, int molecule, double* mu);
/* Sets chi(a,b) = chi(b,a) = value for the following calls, keeping the
 * current iteration vector: one step of a `var` sweep over a chi parameter
 * (Vary.cpp:3-80; SF_Box.cpp:145-176 re-solves each step from the previous
 * solution). SFX_E_INVALID if the change would merge or split iteration
 * variables (states with equal chi rows share one, SF_Solve.cpp:53-86). */
int sfx_set_chi(sfx_system* s, int a, int b, double value);

/* Runs the following calls of s on `stream` (a hipStream_t of the current
 * device; NULL: the stream the system created for itself), after finishing
 * what s queued on its previous stream. The caller keeps the stream alive
 * while s uses it. Calls still wait for their own results before returning,
 * as on the system's stream; the stream orders them after work the caller
 * queued there. Replaces the reference GPU API's use of the default stream
 * (Cuda_tools.cu:180-223 and on: every launch and cudaMemcpy on stream 0). */
int sfx_stream_set(sfx_system* s, void* stream);

/* ---- solvers ----------------------------------------------------------- */
/* SFX_SOLVER_NEWTON: the reference's default iteration, SFNewton with a dense
 * pseudo-Hessian (a float L D U factorisation on the host, one rank-one update
 * per step) and the Scheutjens line search and trust region
 * (sfnewton.cpp:844-1143,1433-1530,1916-2056; newttool.cpp), with the reset
 * rules of SF_Solve::inneriteration (SF_Solve.cpp:659-760). Every residual is
 * FP64 on the GPU. It stops as the reference does: accuracy
 * sqrt(|p| |g| / (1 + |x|)) <= tolerance and |g| <= 10 tolerance. At most
 * 16384 iteration variables (SFX_E_UNSUPPORTED beyond). */
enum { SFX_SOLVER_DIIS = 0, SFX_SOLVER_NK = 1, SFX_SOLVER_NEWTON = 2 };

typedef struct sfx_solver_opts {
  int method;             /* SFX_SOLVER_DIIS (MPS.cpp:228-284) or SFX_SOLVER_NK */
  double tolerance;       /* stop when ||f||_2 (FP64 residual) < tolerance */
  int max_iter;           /* outer iterations */
  int diis_m;             /* DIIS memory (MPS.cpp:306: 10) */
  double diis_eta;        /* DIIS step (MPS.cpp:13: 0.1) */
  int gmres_restart;      /* NK: GMRES restart length */
  int gmres_max_iter;     /* NK: Krylov iterations per Newton step */
  double gmres_rtol;      /* NK: relative Krylov tolerance (forcing term) */
  int jvp_precision;      /* NK: precision of the tangent sweeps */
  double fp64_switch;     /* NK: ||f|| below which residual sweeps run in FP64;
                             a trial point whose predicted ||f|| (the Newton
                             step's linear model) is below it runs in FP64 too.
                             Finite and >= 0 (0: FP64 throughout). FP32 residuals
                             stop improving near their rounding floor (about
                             1e-4 at 64^3); with a switch below it, an FP32 line
                             search that finds no descent moves the solve to
                             FP64 instead of ending it. */
  int precondition;       /* NK: 1 = right-precondition GMRES with the Fourier-space Jacobian of the
                             homogeneous melt where it applies (periodic simple-cubic melt form);
                             0 = none. No counterpart in the reference (SFNewton has none). */
  double ptc_dt0;         /* NK: > 0 globalises by pseudo-transient continuation, solving
                             (I/dt + J) dx = -f from dt = ptc_dt0 with dt grown as ||f|| falls;
                             0 = Newton steps with a backtracking line search */
  int verbose;
  int diis_backtrack;     /* batched DIIS: 1 = per problem, a non-finite residual or one 100x
                             above the best so far returns to the best point, clears the DIIS
                             history and halves that problem's step (giving up below 1e-4 of
                             diis_eta, at the best point); 0 = the plain iteration of
                             MPS.cpp:228-284. No counterpart in the reference. */
  double newton_deltamax;       /* Newton: largest step length (SF_Solve.cpp:1888: 0.5) */
  int newton_linesearch_limit;  /* Newton: evaluations per line search (sfnewton.cpp:31: 20) */
  double newton_reset_criterion; /* Newton: reset the pseudo-Hessian when the accuracy
                                    exceeds this multiple of the best so far
                                    (SF_Solve.cpp:2176: 1e3) */
} sfx_solver_opts;

typedef struct sfx_solve_stats {
  int iterations;         /* outer iterations */
  int residual_evals;
  int residual_evals_fp32;
  int jvp_evals;
  double residual_norm;   /* final ||f||_2 of the FP64 residual */
  double seconds;         /* wall time of the solve (HIP events) */
  int converged;
  double accuracy;        /* Newton: the reference's accuracy sqrt(|p| |g| / (1 + |x|))
                             at the end (sfnewton.cpp:2040); 0 for the other methods */
  int graph_replays;      /* residual evaluations replayed from the system's captured
                             HIP graph (the others ran as separate launches: the first
                             per precision, and all of them while kernel timing is on) */
} sfx_solve_stats;

int sfx_solver_defaults(sfx_solver_opts* o);
/* Applies the NK preconditioner (see sfx_solver_opts.precondition) to r,
 * both in the iteration-vector layout; SFX_E_UNSUPPORTED where it does not
 * apply. A diagnostic with no counterpart in the reference. */
int sfx_precondition(sfx_system* s, const double* r, double* out, int where);
/* Solves f(x)=0 from x (in/out, host or device per `where`). */
int sfx_solve(sfx_system* s, const sfx_solver_opts* o, double* x, int where,
              sfx_solve_stats* stats);

/* SFNewton::iterate(x, n) (sfnewton.h:87, sfnewton.cpp:844) on a residual
 * the caller supplies, as a subclass's override of SFNewton::residuals(g, x)
 * (sfnewton.h:84) supplies it: the same iteration as SFX_SOLVER_NEWTON, on
 * host memory, with no GPU work of its own. Reads tolerance, max_iter and the
 * newton_* fields of o; stats gets iterations, residual_evals, residual_norm,
 * accuracy, converged and the wall time. SFX_E_NOTCONVERGED if it stopped at
 * max_iter or on a residual that is not a number (x holds the last point). */
typedef void (*sfx_residual_fn)(void* ctx, const double* x, double* g);
int sfx_newton_iterate(sfx_residual_fn f, void* ctx, double* x, long n, const sfx_solver_opts* o,
                       sfx_solve_stats* stats);

/* ---- batched 1-D adsorption solves --------------------------------------
 * Replaces running a Vary sweep's calculations one after the other
 * (Vary.cpp:3-80, SF_Box.cpp:145-176: one SF_System + SF_Solve per `start`
 * block). Every problem is the calculation
 *   lat : Lat1DFlat1stO, n_layers, lowerbound surface (frozen wall W in
 *         layer 0), upperbound mirror1, lambda;
 *   mol A_N free at phibulk; mol S monomer solvent;
 *   chi A-S = chi_as, A-W = chi_s, S-W = 0.
 * The iteration vector of a problem has two values per layer z = 1..n in
 * layer order (A, then S), as sfx_residual lays out the same system; the
 * problems are concatenated in order (offsets from sfx_batch_sizes). */
typedef struct sfx_adsorption_desc {
  int n_layers;    /* 2..1024 */
  int N;           /* homopolymer length, >= 2 */
  double lambda;   /* step weight, (0, 1/2] */
  double chi_as;   /* polymer-solvent */
  double chi_s;    /* polymer-wall */
  double phibulk;  /* homopolymer bulk volume fraction, (0, 1) */
} sfx_adsorption_desc;

typedef struct sfx_batch sfx_batch;
/* opt may be NULL; only SFX_FP64 is accepted. */
int sfx_batch_create(const sfx_adsorption_desc* problems, int count, const sfx_options* opt,
                     sfx_batch** out);
int sfx_batch_destroy(sfx_batch* b);
/* Total iteration variables; offsets (count+1 entries, may be NULL) of every
 * problem's part of the iteration vector. */
int sfx_batch_sizes(const sfx_batch* b, long long* nvar, long long* offsets);
/* Threads per problem the tuner chose at creation and the residual time (ms)
 * it measured at 64, 128, 256 and 512 threads (ms: 4 floats, may be NULL). */
int sfx_batch_tuning(const sfx_batch* b, int* threads, float* ms);
/* Residuals of every problem at x (host or device per `where`). */
int sfx_batch_residual(sfx_batch* b, const double* x, double* f, int where);
/* DIIS (MPS.cpp:228-284) on every problem at once from x (in/out), in one
 * kernel launch: each problem's workgroup iterates until that problem is done,
 * with no host synchronisation in between. Per
 * problem: evaluations after the first (iterations) and the final residual
 * norm (host arrays of count, may be NULL). A converged problem returns the
 * point whose residual was measured. SFX_E_NOTCONVERGED if any problem did
 * not converge; x and the arrays are filled in either case. */
int sfx_batch_solve(sfx_batch* b, const sfx_solver_opts* o, double* x, int where, int* iterations,
                    double* norms, sfx_solve_stats* stats);
/* ln G_N of every problem's homopolymer at the last evaluation (count values). */
int sfx_batch_get_lngn(const sfx_batch* b, double* out);
/* phi of one problem at the last evaluation: state 0 = A, 1 = S; n_layers+2
 * values, halo layers 0. */
int sfx_batch_get_phi(const sfx_batch* b, int problem, int state, double* out);

/* ---- per-rank lattice primitives (Cuda_tools.h:3-26 / Lat1stO.h) -------- */
/* All arrays are device pointers of M doubles; stream may be NULL. */
typedef struct sfx_lattice sfx_lattice;
int sfx_lattice_create(const sfx_lattice_desc* d, sfx_lattice** out);
int sfx_lattice_destroy(sfx_lattice* l);
int sfx_lattice_sites(const sfx_lattice* l, long* M);
/* Gout = G * stencil(Gin), halos of Gout filled (Lat1stO::PropagateG(Gi,G,Gout)). */
int sfx_lattice_propagate(const sfx_lattice* l, const double* Gin, const double* G,
                          double* Gout, void* stream);
/* out += A * B on every site (Lat1stO::ConnectG(A,B,out)). */
int sfx_lattice_connect(const sfx_lattice* l, const double* A, const double* B,
                        double* out, void* stream);
/* in /= G where G > 0, else 0 (CorrectDoubleCountG). */
int sfx_lattice_correct_double_count(const sfx_lattice* l, double* in, const double* G,
                                     void* stream);
/* ln of the site-weighted sum of G over the interior (ComputeLnGN). On a 3-D
 * second-order lattice G holds six direction components (6*M) and the result
 * is ln(sum_k sum G^k / 6) (Lat3D2ndO::ComputeLnGN, Lat3D2ndO.cpp:996-1025). */
int sfx_lattice_lngn(const sfx_lattice* l, const double* G, double* lngn, void* stream);
/* Fills the halo layers of P (SetBoundaries). */
int sfx_lattice_set_boundaries(const sfx_lattice* l, double* P, void* stream);
/* sfx_lattice_lngn without allocating: no hipMalloc, hipFree or allocating
 * library call. It has no counterpart in the reference; it is for callers that
 * cannot allocate (code under HIP graph capture, tight serving loops, harnesses
 * that forbid it). The partial sums live in a static device array of 512 KiB per
 * GPU, so the lattice may hold at most 16,777,216 interior sites (larger ones:
 * SFX_E_INVALID). where = SFX_HOST: *lngn is a host double, the call waits for
 * the stream and gives the same bits as sfx_lattice_lngn. where = SFX_DEVICE:
 * lngn points to device memory, ln G_N is written there by a kernel and the call
 * returns without waiting. Calls that overlap on one GPU share the array and race. */
int sfx_lattice_lngn_noalloc(const sfx_lattice* l, const double* G, double* lngn, int where,
                             void* stream);
/* Contact average <P> (SideFraction on every site / Lat3D::Sides). */
int sfx_lattice_sides(const sfx_lattice* l, const double* P, double* out, void* stream);
/* Gout = G * stencil_f(Gin) under a force f along the last gradient: a
 * neighbour one layer below weighs e^f, one above e^-f, normalised as the
 * reference does (Lat1stO::PropagateG(Gi,G,Gout,f): Lat1DFlat1stO.cpp:110-130,
 * Lat3D1stO.cpp:94-165, Lat3DFCC1stO.cpp:102-145, Lat3DHEX1stO.cpp:102-145).
 * Flat 1-D and 3-D lattices; 2-D and curved ones: SFX_E_UNSUPPORTED (the
 * reference stops there: Lat2DFlat1stO.cpp:55-57). A surface neighbour
 * contributes 0, as in sfx_lattice_propagate (a surface's propagator halo
 * holds 0 in the reference's sweeps). Halos of Gout filled. */
int sfx_lattice_propagate_force(const sfx_lattice* l, const double* Gin, const double* G, double* Gout,
                                double f, void* stream);
/* Normalisations (Lat1stO.h:19-50, Lat2ndO.h:62-66): phi *= C on every site
 * (NormPhiFree, Lat1DFlat1stO.cpp:318-322), or phi *= C / exp(ComputeLnGN(Gi))
 * (NormPhiRestr, :324-331); on a 3-D second-order lattice Gi holds six
 * components, the divisor is 6 exp(ComputeLnGN(Gi)) and PHI (3*M, the per-axis
 * parts, may be NULL) is scaled alike (Lat3D2ndO.cpp:1027-1047). The restricted
 * form reduces Gi on the device: no host synchronisation. */
int sfx_lattice_norm_phi_free(const sfx_lattice* l, double* phi, double* PHI, double C, void* stream);
int sfx_lattice_norm_phi_restr(sfx_lattice* l, double* phi, double* PHI, const double* Gi, double C,
                               void* stream);
/* ---- second order, 3-D (Lat2ndO.h:19-66, Cuda_tools.h Forward/Backward
 * with S): arrays of six direction components, component k of site i at
 * k*M + i, k the bond direction {+x, +y, +z, -z, -y, -x}; S the straight-bond
 * weight, (1-S)/4 per perpendicular bond, reversal 0. 1-D and 2-D
 * second-order lattices (the reference's three and five lumped components):
 * SFX_E_UNSUPPORTED. */
/* backward = 0: PropagateF, Gout^k(r) = G(r) sum_j T_kj Gin^j(r - e_k)
 * (Lat3D2ndO.cpp:64-152); backward = 1: PropagateB, Gout^k(r) =
 * G(r) sum_j T_jk Gin^j(r + e_j) (:597-678). Input halos are read through the
 * boundary maps (mirror sides reverse their axis' components), Gout's halos
 * are filled. Gin and Gout must not overlap. */
int sfx_lattice_propagate2(const sfx_lattice* l, const double* Gin, const double* G, double* Gout,
                           double S, int backward, void* stream);
/* out += sum_k A^k B^k on every site; OUT (3*M, may be NULL) += A^k B^k into
 * the part of k's axis (ConnectG with PHI, Lat3D2ndO.cpp:895-955). */
int sfx_lattice_connect2(const sfx_lattice* l, const double* A, const double* B, double* out,
                         double* OUT, void* stream);
/* in /= G where G > 0, else 0, and the same for the three parts of IN (3*M,
 * may be NULL) (CorrectDoubleCountG with PHI, Lat3D2ndO.cpp:988-994). */
int sfx_lattice_correct_double_count2(const sfx_lattice* l, double* in, double* IN, const double* G,
                                      void* stream);

#ifdef __cplusplus
}
#endif
#endif /* SFX_H */
