/*
 * pbsam_amd: MI355X (gfx950) implementation of PB-SAM's mutual-polarization
 * solve, with the C ABI that replaces the reference's CMolecule entry points
 * on that path.  Plain C types only; the library owns all device memory.
 *
 * Units follow the reference: lengths in Angstrom, charges in e, energies in
 * the reference's internal units (e^2/A scaled by 1/eps_in); conversions to
 * kT or kcal/mol are left to the caller (reference: src/BDnam.h:10,19-20).
 *
 * Coefficient blocks use the reference layout (src/expansion.cpp:57-63): a block
 * of pole order p holds p*p doubles; order n occupies [n*n, n*n+2n] as
 * [c(n,0), Re c(n,1), Im c(n,1), ..., Re c(n,n), Im c(n,n)].
 */
#ifndef PBSAM_AMD_H
#define PBSAM_AMD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes (the reference exits or asserts instead; these never exit) */
#define PBSAM_OK 0
#define PBSAM_ERR_BADARG 1      /* invalid argument */
#define PBSAM_ERR_OVERLAP 2     /* two spheres of different molecules overlap
                                   (reference: computeForces returns false,
                                   src/molecule.cpp:1090-1097) */
#define PBSAM_ERR_NOCONV 3      /* round limit reached before the tolerance */
#define PBSAM_ERR_HIP 4         /* HIP runtime error */
#define PBSAM_ERR_OOM 5         /* device allocation failed */
#define PBSAM_ERR_UNSUPPORTED 6 /* input needs a feature this build lacks */
#define PBSAM_ERR_CHARGE 7      /* a charge lies outside every sphere
                                   (reference exits, src/molecule.cpp:1019-1020) */
#define PBSAM_ERR_STATE 8       /* call out of order (e.g. no self-polarization yet) */

#define PBSAM_PMAX 32           /* largest pole order */

typedef struct pbsam_ctx pbsam_ctx;
typedef struct pbsam_moltype pbsam_moltype;
typedef struct pbsam_system pbsam_system;

/* Context: one HIP device and stream, the physical constants and recursion tables.
 * Replaces the process-wide statics set by CSystem::initConstants
 * (src/system.cpp:16-25) and CMolecule::initConstants (src/molecule.cpp:42-70). */
int pbsam_ctx_create(double kappa, double eps_out, int device, pbsam_ctx **out);
void pbsam_ctx_destroy(pbsam_ctx *ctx);
const char *pbsam_last_error(const pbsam_ctx *ctx);
/* Memory plan.  The reference recomputes every re-expansion operator per use ("LowMemory",
 * src/molecule.cpp:1286-1293); this library stores two kinds of them when they fit: the
 * far-pair translation matrices of a molecule type (made for its frozen far values and
 * reused by the gradient phase) and its numerical re-expansion maps (gradient phase).
 * Stores that would take the context's total past `bytes` are not made (default: a quarter
 * of the device's memory); their users then make the operators again in every step that
 * needs them, with results equal bit for bit.  Applies to molecule types whose far values
 * or maps are built after the call.  pbsam_ctx_stored reports the bytes stored and the cap. */
int pbsam_ctx_set_store_limit(pbsam_ctx *ctx, long long bytes);
int pbsam_ctx_stored(const pbsam_ctx *ctx, long long *bytes, long long *limit);
/* HIP graphs (no counterpart in the reference, whose sweeps are host loops): the gradient
 * phase's Gauss-Seidel sweep runs in windows of steps (every launch of a window is known
 * before it starts); a window whose launches recur with the same buffers and sizes is
 * captured as a HIP graph the second time and replayed after that.  Results are the same
 * with graphs on (default) or off.  pbsam_system_graph_stats counts a system's graph
 * launches and captures. */
int pbsam_ctx_set_graphs(pbsam_ctx *ctx, int on);
int pbsam_system_graph_stats(const pbsam_system *sys, long long *launches, long long *captures);
/* Debye kappa (1/A) from a salt concentration in mol/L, as src/selfp_makemat.cpp:59-60 */
double pbsam_kappa_from_salt(double salt_molar, double eps_out, double temp_k);

/* Molecule type: spheres, charges, surface points, per-sphere constants.
 * Replaces readpqr + generateMolSPX + generateMolTypeIntraPolLists + the CMolecule
 * self-polarization constructor (src/molecule.cpp:355-415, 501-535, 756-780, 878-1023;
 * src/expcenter.cpp:375-390, 470-552, 1042-1060). */
typedef struct pbsam_moltype_desc {
  int n_sph;
  const double *cen;     /* 3*n_sph sphere centres (input frame) */
  const double *rad;     /* n_sph radii */
  int n_chg;
  const double *chg_xyz; /* 3*n_chg */
  const double *chg_q;   /* n_chg, in e */
  const double *rcen;    /* 3: molecule centre; NULL = origin */
  int p;                 /* pole order, 1..PBSAM_PMAX */
  double eps_in;         /* interior dielectric */
  const int *npts;       /* per-sphere spiral size, NULL = the reference's radius rule */
  int npts_cap;          /* cap of the radius rule (reference: 10000) */
  double intra_rcutoff;  /* near/far split (reference: 30 A) */
  const double *imats;   /* n_sph*p^4 column-major surface-integral matrices, or NULL */
  int imat_npts;         /* with imats NULL: spiral size of the matrix build (on the device,
                            as pbsam_imat_build) for spheres with buried points; the others get
                            the exact identity matrix.  0 with overlapping spheres and no
                            imats: PBSAM_ERR_UNSUPPORTED */
} pbsam_moltype_desc;

int pbsam_moltype_create(pbsam_ctx *ctx, const pbsam_moltype_desc *desc, pbsam_moltype **out);
void pbsam_moltype_destroy(pbsam_moltype *mt);
int pbsam_moltype_info(const pbsam_moltype *mt, int *n_sph, int *p, long long *n_points);
/* exposed surface points of sphere k (relative to its centre) and the spiral size */
int pbsam_moltype_points(const pbsam_moltype *mt, int k, double *xyz, int *n_exposed, int *n_spiral);

/* Self-polarization (CMolecule::polarize_self, src/molecule.cpp:1899-1979) in the
 * reference's Gauss-Seidel order: spheres in ascending index, selective recompute
 * (dev > 0.1*maxdev or every 5th round), until maxdev <= tol or max_rounds.
 * tol <= 0 / max_rounds <= 0 select the reference's values (1e-4, or 1e-2 for one sphere;
 * 1700 rounds).  rounds_out may be NULL.  Stores F/H as the self values.
 * Returns PBSAM_ERR_NOCONV when the round limit ends the loop. */
int pbsam_selfpol(pbsam_moltype *mt, double tol, int max_rounds, int *rounds_out);
/* fixed number of simultaneous (Jacobi) rounds from the current state, every sphere
 * from the previous round's values (does not store self values) */
int pbsam_selfpol_rounds(pbsam_moltype *mt, int rounds, double *maxdev_out);
/* Inject self-polarized F/H (n_sph*p^2 each), as reading .exp files
 * (readExpansionsP, src/readutil.cpp:564-666). */
int pbsam_moltype_set_self(pbsam_moltype *mt, const double *F, const double *H);
/* Exposed self charges + frozen intra-molecular re-expansions:
 * generateMolExposedChargesFH + computeMolTypeValues (src/molecule.cpp:545-810). */
int pbsam_moltype_values(pbsam_moltype *mt);
/* current self-polarization state of sphere k (any pointer may be NULL);
 * tq = {TQF, TQH, dev} */
int pbsam_moltype_get(const pbsam_moltype *mt, int k, double *F, double *H, double *qF, double *qH,
                      double *tq);
int pbsam_moltype_get_intra(const pbsam_moltype *mt, int k, double *LF, double *LH, double *LFfar,
                            double *LHfar);

/* System of rigid molecules; pose = centre position + unit quaternion (w,x,y,z).
 * Replaces the BD CMolecule constructor + setPos/setOrient (src/molecule.cpp:422-488,
 * src/molecule.h:325-329) and CMolecule::initMutualConstants (src/molecule.cpp:78-95). */
int pbsam_system_create(pbsam_ctx *ctx, int n_mol, pbsam_moltype *const *mts, const double *pos,
                        const double *quat, pbsam_system **out);
void pbsam_system_destroy(pbsam_system *sys);
/* A batch of n_sys independent systems (BD snapshots or trajectories) held as one system
 * whose molecules come in groups: system g has n_mol[g] molecules, listed after those of
 * systems 0..g-1 in mts/pos/quat, and interacts only with itself.  Every call on a system
 * handle works on a batch with molecule indices running over all groups; each group
 * follows its own reference schedule exactly (its results equal those of a system made
 * of that group alone up to the rounding of different launch shapes), while the groups'
 * Gauss-Seidel chains share kernel launches.
 * The batched counterpart of calling CMolecule::computeForces (src/moldynamics.cpp:554-587)
 * once per independent system; there is no single reference entry point for it.
 * Systems may use different pole orders (the molecules of one system share one): the
 * handle then holds one batch per order and every call runs them in turn, with results
 * and indices in the caller's order. */
int pbsam_batch_create(pbsam_ctx *ctx, int n_sys, const int *n_mol, pbsam_moltype *const *mts,
                       const double *pos, const double *quat, pbsam_system **out);
/* per-group results of the last solve / gradient (arrays of n_sys, any may be NULL):
 * status (PBSAM_OK, PBSAM_ERR_OVERLAP, PBSAM_ERR_NOCONV), mutual rounds, gradient passes */
int pbsam_batch_status(const pbsam_system *sys, int *status, int *mutual_rounds, int *grad_passes);
/* Changing a pose or the cutoffs discards the last solution: pbsam_gradient, pbsam_forces and
 * pbsam_system_get_grad return PBSAM_ERR_STATE until the next pbsam_solve_mutual. */
int pbsam_system_set_pose(pbsam_system *sys, int i, const double *pos, const double *quat);
int pbsam_system_set_cutoffs(pbsam_system *sys, double inter_rcutoff, double interact_rcutoff);

/* Mutual polarization (CMolecule::polarize_mutual phases 0-2, src/molecule.cpp:1989-2075)
 * preceded by the pair lists (generateInterXFormsForPolarize_LowMemory :1036-1126).
 * Phase 0 rotates the self values to the lab frame, round 1 updates every polarized
 * sphere from the self values, later rounds run in the reference's Gauss-Seidel order
 * (molecules in order, spheres ascending, those with dev > 0.1*MAXDEV) until the mean
 * change <= tol or max_rounds (tol <= 0 / max_rounds <= 0: the reference's 1e-4 and
 * 1000).  flags: PBSAM_HROT_REFRESH refreshes the rotated H used for inter-molecular
 * coupling after each update (the reference keeps round-1 values).
 * Returns PBSAM_ERR_OVERLAP when two molecules' spheres overlap and PBSAM_ERR_NOCONV
 * when the round limit ends the loop (for a batch: the first group's non-OK status; the
 * other groups are solved, an overlapping group gets no pairs; rounds_out is the largest
 * group's count). */
#define PBSAM_HROT_REFRESH 1
/* System settings (pbsam_system_set_flags) beyond the reference's behaviour.  Together,
 * PBSAM_EXACT_DERIVATIVE, they make pbsam_forces the exact derivative of the energy of the
 * converged solution (-d/dr of sum_k <H_k, LS_k> per molecule, orientation applied as the
 * reference does):
 *   PBSAM_HROT_REFRESH  the mutual solve refreshes the rotated H after each update (the
 *                       reference keeps round-1 values, so its fixed point is not
 *                       self-consistent)
 *   PBSAM_FULL_ORDERS   no adaptive truncation (computeOrder, src/xforms.cpp:402-414,
 *                       returns p everywhere)
 *   PBSAM_NO_GRAD_SKIP  gradient surface charges with |g|^2 < 1e-15 are kept (the reference
 *                       skips them, src/triexpan.cpp:136,146)
 *   PBSAM_EXT_SIGN      the geometric derivative of a re-expansion from a sphere of molecule
 *                       j onto a sphere of another molecule enters the gradient of j as -d/dP
 *                       (P = c_i - c_m); the reference adds +d/dP
 *                       (prepareDTA_LowMemory, src/molecule.cpp:1495-1521), which leaves its
 *                       forces 0.5-2 % off the energy gradient
 *   PBSAM_SELF_FRAME    a sphere with interaction partners but none that polarize it takes
 *                       its force from its rotated (lab-frame) self H against the lab-frame
 *                       gradient of its partners' re-expanded self H; the reference pairs
 *                       the unrotated self H with that lab-frame gradient
 *                       (computeForceOn_0, src/expcenter.cpp:140, after
 *                       prepareDTA_iself_LowMemory, src/molecule.cpp:1585-1629), so its
 *                       forces change when the whole system is rotated
 * Interactions of such spheres are not part of computePot's energy (their LS stays zero),
 * so the exact-derivative statement covers the polarized spheres' energy.
 * Setting flags discards the last solution (as pbsam_system_set_pose does). */
#define PBSAM_FULL_ORDERS 2
#define PBSAM_NO_GRAD_SKIP 4
#define PBSAM_EXT_SIGN 8
#define PBSAM_SELF_FRAME 16
#define PBSAM_EXACT_DERIVATIVE 31
/* With PBSAM_EXACT_DERIVATIVE: pbsam_bd_run takes each step's forces and torques from
 * pbsam_solve_exact and pbsam_adjoint_forces (squared relative residual 1e-16) instead of the
 * reference's force evaluation, and moves the molecules by those lab-frame vectors (-dE/dr,
 * -dE/domega).  The reference's CBDnam moves them by its returned force conj(q) F and torque,
 * which pbsam_compute_forces reproduces.  No counterpart in the reference; needs identity
 * surface matrices and the interaction cutoff at the polarization cutoff
 * (PBSAM_ERR_UNSUPPORTED from pbsam_bd_run otherwise). */
#define PBSAM_BD_ADJOINT 32
int pbsam_system_set_flags(pbsam_system *sys, unsigned flags);
int pbsam_solve_mutual(pbsam_system *sys, double tol, int max_rounds, unsigned flags, int *rounds_out);
/* round 1 from self values, then a fixed number of simultaneous (Jacobi) rounds */
int pbsam_mutual_rounds(pbsam_system *sys, int rounds, unsigned flags, double *maxdev_out);
int pbsam_system_get(const pbsam_system *sys, int mol, int k, double *F, double *H, double *LS,
                     double *Hrot);
/* CMolecule::computePot (src/moldynamics.cpp:509-521) per molecule */
int pbsam_energy(pbsam_system *sys, double *e_per_mol);
int pbsam_system_counts(const pbsam_system *sys, int *n_inter_pairs, int *n_polarized);

/* Gradient phase of CMolecule::polarize_mutual (src/molecule.cpp:2077-2228) after a
 * mutual solve of the current poses: forward-mode derivatives of every polarized
 * sphere's H with respect to each molecule's position, in the reference's Gauss-Seidel
 * order (prepareDTA_LowMemory :1483-1571, recomputeGrad_LowMemory :1811-1843), until the
 * mean change <= tol or max_rounds passes (tol <= 0 / max_rounds <= 0: 1e-4 and 1000).
 * passes_out: total passes over the per-molecule lists. */
int pbsam_gradient(pbsam_system *sys, double tol, int max_rounds, int *passes_out);
/* CMolecule::computeMol_Force_and_Torque (src/moldynamics.cpp:525-550) for every
 * molecule after pbsam_gradient: force and torque (3 doubles each per molecule), the
 * torque being sum c_k x f_k over spheres (the surface term is off in the reference), both
 * multiplied by conj(orientation) as the reference does before returning them. */
int pbsam_forces(pbsam_system *sys, double *force, double *torque);
/* CMolecule::computeForces (src/moldynamics.cpp:554-587): pair lists, mutual solve,
 * gradient, forces and torques.  Returns PBSAM_ERR_OVERLAP when spheres of two molecules
 * overlap; PBSAM_ERR_NOCONV when the mutual solve hit its round limit (forces are still
 * computed, as the reference does). */
int pbsam_compute_forces(pbsam_system *sys, double tol, int max_rounds, double *force, double *torque,
                         int *rounds_out);
/* Torques that are the exact derivative of the energy (no counterpart in the reference, whose
 * torque is sum c x f with the surface term disabled, src/moldynamics.cpp:540-541): after a
 * mutual solve with PBSAM_EXACT_DERIVATIVE set, the gradient sweeps of polarize_mutual run
 * once more for a rotation of each molecule about its centre (sources: the sphere
 * displacements omega x (c - r) and the turn of the rotated expansions), until the mean
 * change <= tol or max_rounds passes, and torque (3 per molecule, lab frame) receives
 * -dE_i/domega of each molecule's energy sum_k <H_k, LS_k>.  The gradient blocks then hold
 * the rotational derivative (pbsam_forces returns PBSAM_ERR_STATE until pbsam_gradient runs
 * again).  PBSAM_ERR_STATE without the flags or the solve; PBSAM_ERR_UNSUPPORTED when a
 * sphere has interaction partners but no polarization partner. */
int pbsam_torques_exact(pbsam_system *sys, double tol, int max_rounds, double *torque);
/* The fixed point of the mutual solve under PBSAM_EXACT_DERIVATIVE by restarted GMRES (no
 * counterpart in the reference, which iterates polarize_mutual's Gauss-Seidel rounds,
 * src/molecule.cpp:1989-2075): pair lists and phase 0 as pbsam_solve_mutual, then GMRES on the
 * linear system of the polarized spheres' F and H blocks (one operator application = one
 * Jacobi round of the mutual solve without its constant terms, with the inner solve's limit)
 * until the squared residual <= tol times the squared right-hand side, or max_rounds
 * applications.  Leaves the state pbsam_energy, pbsam_gradient and pbsam_adjoint_forces read.
 * The reference's iteration need not converge (on molecules with nearly touching spheres its
 * Jacobi rounds diverge and its Gauss-Seidel rounds wander); GMRES needs only the linear
 * system to be invertible.  PBSAM_ERR_STATE without the flags; PBSAM_ERR_UNSUPPORTED for
 * polarized spheres with buried points; PBSAM_ERR_NOCONV at the limit. */
int pbsam_solve_exact(pbsam_system *sys, double tol, int max_rounds, int *applications_out);
/* Forces and torques from the adjoint of the converged mutual-polarization fixed point (no
 * counterpart in the reference, which differentiates by forward sweeps,
 * src/molecule.cpp:2077-2228): after a mutual solve with PBSAM_EXACT_DERIVATIVE set, one
 * transposed solve for all two-molecule systems of the handle, and for a system of n > 2
 * molecules one per molecule (solved for every such system of the handle at once, n passes
 * for the largest), each restarted GMRES until the squared residual <= tol times the squared
 * right-hand side, or max_rounds operator applications (each application one pass of the
 * re-expansion kernels over every polarized sphere); then each solve's contraction with the
 * source planes of the positions and orientations.  force and torque
 * (3 doubles per molecule, lab frame) receive -dE_i/dr_i and -dE_i/domega_i of each molecule's energy
 * sum_k <H_k, LS_k>, the derivative of the solved energy (pbsam_forces in this mode returns
 * conj(q_i) times the same force when every polarized pair of a molecule lies within the
 * intra cutoff; pbsam_torques_exact the same torque).  sweeps_out: operator applications,
 * summed over the passes.  PBSAM_ERR_STATE without the flags or the solve;
 * PBSAM_ERR_UNSUPPORTED for spheres with interaction partners beyond the polarization cutoff
 * and polarized spheres with buried points; PBSAM_ERR_NOCONV at the limit (of any pass). */
int pbsam_adjoint_forces(pbsam_system *sys, double tol, int max_rounds, double *force, double *torque,
                         int *sweeps_out);
/* gradient blocks of sphere k of molecule mol: gH (3 p^2, polarized spheres) and gLHN
 * (3 p^2), as left by the last pbsam_gradient */
int pbsam_system_get_grad(const pbsam_system *sys, int mol, int k, double *gH, double *gLHN);

/* Launch-shape autotuning (no counterpart in the reference): times complete force
 * evaluations of this system (pbsam_compute_forces, stock tolerances, reps calls each) while
 * it varies, one at a time, the settings that only change how the work is split over
 * workgroups and launches (the sweep look-ahead window, the multi-plane split threshold,
 * the order-class binning threshold, the point split of the stored-map products, the
 * bundle size from which they run on matrix cores, the point split of the numerical
 * re-expansions and the lanes per point of the surface charges), and keeps the fastest.  The
 * choice is used by this system and by every later system of the context with the same pole
 * order, sphere count and number of systems.  Results stay within rounding of the untuned
 * ones (the point splits change summation orders).  ms_per_eval: the best time per
 * evaluation (summed over the parts of a mixed-order batch); settings (7 ints, may be NULL):
 * window, split threshold, binning threshold, GEMV workgroup target, matrix-core bundle
 * minimum, point-to-expansion workgroup target, surface-charge workgroup target. */
int pbsam_autotune(pbsam_system *sys, int reps, double *ms_per_eval, int *settings);

/* ---- energy and potential queries ----
 * pbsam_potential: CMolecule::computePotInSpace (src/moldynamics.cpp:469-507) at n points
 * (xyz: 3n, lab frame, pot: n): the sum over every molecule and sphere of the sphere's H
 * contracted with the screened local expansion of a unit charge at the point (order
 * computeOrder(rho, TQH, a)), from the current state (self values before any solve).
 * pbsam_total_int_energy: CMolecule::computeTotalIntEnergy (src/moldynamics.cpp:358-386 with
 * interactCenters_LowMemory, src/molecule.cpp:1851-1890), one value per system of the handle
 * (e: n_sys): for every molecule pair, each sphere's H with the re-expansion of the other
 * molecule's H onto it.  The reference forms both with the spheres' unrotated centres and
 * molecule-frame blocks, which is right only at the identity orientation; by default the
 * library uses the rotated centres (and lab-frame blocks), and PBSAM_POT_REFERENCE_FRAME asks
 * for the reference's formula exactly.  The reference's aggregated molecular multipoles (an
 * optional shortcut for far points, off by default there) are not used. */
#define PBSAM_POT_REFERENCE_FRAME 1u
int pbsam_potential(pbsam_system *sys, int n, const double *xyz, double *pot, unsigned flags);
int pbsam_total_int_energy(pbsam_system *sys, unsigned flags, double *e);

/* ---- Brownian dynamics of two-molecule trajectories (CBDnam, src/BDnam.cpp:166-417) ----
 * A batch made by pbsam_batch_create with two molecules in every system runs one trajectory
 * per system.  A step of a running trajectory is one pass of CBDnam::run's loop
 * (src/BDnam.cpp:196-246): the time step of compute_dt (:409-417) or dt_fixed, forces and
 * torques from the batched force evaluation when the centres are closer than f_dist (the
 * system's cutoffs apply: set them as CBDnam does, 10 and 100 A, src/BDnam.cpp:138-140), and
 * makeMove (:262-370) with its collision retries, escape, stored rotations and docking test
 * (:377-401).  Each trajectory draws from its own drand48 stream seeded as srand48(seed), with
 * normRand's polar method (src/util.h:514-537): trajectory g is the reference's run of one
 * trajectory after srand48(seeds[g]).  Status values are CBDnam::STATUS (src/BDnam.h:39);
 * PBSAM_BD_FAILED marks a trajectory whose force evaluation found overlapping spheres (the
 * reference exits there, src/BDnam.cpp:211-213). */
#define PBSAM_BD_ESCAPED 0
#define PBSAM_BD_DOCKED 1
#define PBSAM_BD_RUNNING 2
#define PBSAM_BD_STUCK 3
#define PBSAM_BD_FAILED 4
typedef struct pbsam_bd_params {
  double dtr;      /* translational diffusion of molecule 1 (m_Dtr, A^2/ps) */
  double dr[2];    /* rotational diffusion of molecules 0 and 1 (m_Dr1, m_Dr2, rad^2/ps) */
  double b_dist;   /* start distance of the random placement (b_DIST) */
  double q_dist;   /* escape distance (q_DIST) */
  double f_dist;   /* centre distance below which forces are computed (f_DIST) */
  double fact_tr;  /* force -> displacement factor (FACT2) */
  double ikbt;     /* 1/kT (IKbT) */
  double coul_k;   /* torque unit factor (COUL_K) */
  double dt_fixed; /* > 0: fixed time step instead of compute_dt */
  int max_tries;   /* collision retries (500) */
} pbsam_bd_params;
/* CBDnam's values: src/BDnam.cpp:143, src/BDnam.h:10,18-25, compute_dt, 500 tries */
void pbsam_bd_defaults(pbsam_bd_params *prm);
/* Start one trajectory per system of the batch, seeds: n_sys values.  place != 0: CBDnam::run's
 * random start (molecule 0 at the origin, molecule 1 at b_dist in a random direction, random
 * orientations; src/BDnam.cpp:171-181); place == 0: the systems' current poses. */
int pbsam_bd_start(pbsam_system *batch, const pbsam_bd_params *prm, const long long *seeds, int place);
/* Docking criteria (CBDnam::addMolContact + IsDocked, src/BDnam.cpp:38-41,377-401): n_def
 * definitions; definition d holds n_pair[d] sphere pairs (pairs: 2 ints each, sphere of
 * molecule 0, sphere of molecule 1, all definitions in order) with separation thresholds
 * dist, and docks when n_contact[d] of its pairs are within their thresholds. */
int pbsam_bd_contacts(pbsam_system *batch, int n_def, const int *n_contact, const int *n_pair, const int *pairs,
                      const double *dist);
/* Advance every running trajectory by up to n_steps steps.  status and steps (n_sys each,
 * either may be NULL): each trajectory's status and total steps taken.  The systems' poses
 * follow the trajectories.  Returns PBSAM_ERR_NOCONV when a force evaluation hit its round
 * limit (the reference logs and goes on). */
int pbsam_bd_run(pbsam_system *batch, int n_steps, int *status, long long *steps);
/* trajectory g: pos (6), quat (8, w x y z, as the reference keeps them: products of unit
 * rotations, not renormalized), stored rotations m_rot1/m_rot2 (8), the last time step */
int pbsam_bd_get(const pbsam_system *batch, int g, double *pos, double *quat, double *rot, double *dt);

/* ---- the reference's file formats (src/readutil.cpp; host only, no device needed) ----
 * Errors are PBSAM_ERR_BADARG (unreadable file, a matrix of lower order than asked) where the
 * reference exits. */
/* readpqr (src/readutil.cpp:285-355): "ATOM" records, x y z charge radius from column 31;
 * "CEN" at column 18 marks a dielectric sphere (centre, radius), other records are charges
 * (position, charge).  Call with NULL arrays for the counts, then with arrays of at least
 * that many (*n_sph, *n_chg give their sizes).  err (err_len bytes, may be NULL) gets a message. */
int pbsam_read_pqr(const char *path, int *n_sph, double *cen, double *rad, int *n_chg, double *chg_xyz,
                   double *chg_q, char *err, int err_len);
/* writeMat_Unformatted / readMat_Unformatted (src/readutil.cpp:378-435): an int pole order,
 * then p^4 doubles (the p^2 x p^2 matrix, column-major).  Reading order p from a file of
 * order p_file >= p keeps the leading p^2 rows of the leading p^2 columns. */
int pbsam_write_imat(const char *path, int p, const double *imat);
int pbsam_read_imat(const char *path, int p, double *imat, int *p_file);
/* writeExpansion / readExpansion (src/readutil.cpp:541-595): text, pole order, scale, cutoff,
 * then the p^2 coefficients (precision 8, one line per order); a file of lower order is
 * padded with zeros, one of higher order cut */
int pbsam_write_expansion(const char *path, int p, double scale, double rcut, const double *coef);
int pbsam_read_expansion(const char *path, int p, double *scale, double *rcut, double *coef, int *p_file);
/* readConfigWithMolType (src/readutil.cpp:835-886) and CMolecule::saveConfig
 * (src/molecule.cpp:254-270): lines "x y z w qx qy qz moltype"; read quaternions are
 * normalized.  pbsam_read_config: NULL arrays for the count first. */
int pbsam_read_config(const char *path, int *n, double *pos, double *quat, int *moltype);
int pbsam_write_config(const char *path, int n, const double *pos, const double *quat, const int *moltype);
/* readMats and readExpansionsP (src/readutil.cpp:493-506,603-630): a molecule type's files in
 * the reference's layout, <dir>/imat.sp<k>.out.bin and <dir>/<run>.<k>.{F,H}.exp, k < ns.
 * imats: ns * p^4; F, H: ns * p^2; intra_rcutoff: the cutoff the expansion files carry. */
int pbsam_read_imats(const char *dir, int ns, int p, double *imats);
int pbsam_read_expansions(const char *dir, const char *run, int ns, int p, double *F, double *H,
                          double *intra_rcutoff);

/* ---- kernel-level entry points (one batch of independent jobs; host arrays in/out) ---- */
/* CSolExpCenter::build_IMat (src/expcenter.h:207, src/expcenter.cpp:667-902), the `pbsam mat`
 * stage for one sphere: from its exposed (npe) and buried (npb) surface points (3 doubles
 * each, relative to the centre, on the sphere's spiral), the p^2 x p^2 column-major
 * surface-integral matrix.  pbsam_moltype_create builds these itself when imats is NULL and
 * imat_npts > 0. */
int pbsam_imat_build(pbsam_ctx *ctx, int p, int npe, const double *spe, int npb, const double *spb,
                     double *imat);
/* CXFormA::reset + xformH (src/xforms.cpp:51-77,123-132) for n pairs.
 * P: 3n (target - source), a: 2n (source radius, target radius), order: n (<= p),
 * in/out: n*p*p blocks. */
int pbsam_xform_h_batch(pbsam_ctx *ctx, int n, int p, const double *P, const double *a, const int *order,
                        const double *in, double *out);
/* Allocation-free variant of pbsam_xform_h_batch; it has no counterpart in the reference.
 * For callers that cannot allocate device memory (code under HIP graph capture, serving
 * loops, harnesses that forbid hipMalloc).  P, a, order, in and out are DEVICE pointers
 * with the meanings of pbsam_xform_h_batch (the caller owns them and orders its writes to
 * them before the call); the results equal pbsam_xform_h_batch's.  The job tables and
 * translation matrices live in a fixed per-device scratch array of
 * PBSAM_NOALLOC_SCRATCH_BYTES, held until the process ends: a batch needing more
 * (pbsam_xform_h_batch_noalloc_bytes; 64 MiB hold about 880 pairs at p = 30) returns
 * PBSAM_ERR_BADARG before any launch, as does an order outside [1, p] (checked on the
 * device; the call then waits for the stream).  Calls never allocate and return after the
 * context's stream has finished them.  Calls on one device share the scratch: calls that
 * overlap (other threads, other contexts on the same device) race. */
#define PBSAM_NOALLOC_SCRATCH_BYTES (64u << 20)
int pbsam_xform_h_batch_noalloc(pbsam_ctx *ctx, int n, int p, const double *P, const double *a, const int *order,
                                const double *in, double *out);
/* scratch bytes a call of n pairs at order p needs (0 for invalid sizes) */
size_t pbsam_xform_h_batch_noalloc_bytes(int n, int p);
/* CXFormAIntra::reset + xformF/xformH (src/xforms.cpp:214-270) */
int pbsam_xform_intra_batch(pbsam_ctx *ctx, int n, int p, const double *P, const double *a,
                            const int *orders /* 2n: pH, pF */, const double *H, const double *F,
                            double *outH, double *outF);
/* CXFormNIntra::xformFH (src/xforms.cpp:528-550) for one source point set and n targets */
int pbsam_xform_num_batch(pbsam_ctx *ctx, int n, int p, const double *P, const double *a, int npts,
                          const double *xyz, const double *qF, const double *qH, double tQF, double tQH,
                          double *outF, double *outH);
/* CXFormA::reset + xformH(CExpan, CGradExpan) (src/xforms.cpp:51-77,86-115,165-192):
 * derivative of the re-expansion with respect to P, out3: n * 3 * p*p (x, y, z blocks) */
int pbsam_xform_h_grad_batch(pbsam_ctx *ctx, int n, int p, const double *P, const double *a, const int *order,
                             const double *in, double *out3);
/* CRotCoeff::reset(CQuat) + rotateWithXi (src/rotcoeff.cpp:181-201,250-350) */
int pbsam_rotate_quat_batch(pbsam_ctx *ctx, int n, int p, const double *quat, int fwd, const double *in,
                            double *out);

/* ---- instrumentation (no counterpart in the reference) ----
 * With profiling on, every launch of the kernels below is bracketed by HIP events on the
 * context's stream, and the kernels add their algorithmic work to device counters:
 * flops and HBM bytes by the reference's loop bounds for the units they process
 * (pairs at their truncation orders, points, inner sweeps).  pbsam_profile(ctx, 1)
 * resets and starts; (ctx, 0) stops.  ms is the summed event time of the launches. */
#define PBSAM_K_XFORM 0   /* analytical re-expansion R^H T R, one workgroup per pair */
#define PBSAM_K_P2L 1     /* numerical re-expansion of surface charges */
#define PBSAM_K_SOLVE 2   /* per-sphere inner F/H iteration */
#define PBSAM_K_SURFACE 3 /* exposed surface charges */
#define PBSAM_K_XGRAD 4   /* gradient re-expansion, one workgroup per pair */
#define PBSAM_K_SURFGRAD 5 /* exposed-surface gradient charges */
#define PBSAM_K_COUNT 6
int pbsam_profile(pbsam_ctx *ctx, int on);
int pbsam_profile_read(pbsam_ctx *ctx, int kernel, long long *launches, double *ms, double *flops,
                       double *bytes);
/* wait for the context's stream */
int pbsam_sync(pbsam_ctx *ctx);

#ifdef __cplusplus
}
#endif
#endif
