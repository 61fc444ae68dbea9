/*
 * CPU oracle for the PB-SAM mutual-polarization path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library restates the reference's algorithms
 * on the CPU so that tests, the smoke check and the benchmark's CPU baseline
 * can check the HIP path against it.  Nothing in the product library links,
 * loads or calls it.
 *
 * Coefficient blocks use the reference layout (src/expansion.cpp:57-63): for
 * order n there are 2n+1 reals [c(n,0), Re c(n,1), Im c(n,1), ...] at offset n*n.
 * A block of order p holds p*p doubles.
 *
 * What pins each part (tests/test_oracle_ref.py, test_oracle_math.py,
 * test_oracle_forces.py):
 *   primitives, re-expansions (and their gradient), computeOrder, sphere points,
 *   IMat of a buried sphere, molecule-type set-up and frozen intra values:
 *     known-answer vectors of the reference compiled from its sources (p <= 15),
 *     and at p = 30 closed forms (harmonics, Bessel, addition theorems, rotation
 *     of charges, central differences);
 *   orc_selfpol (reference order): Kirkwood's closed form for one sphere;
 *   orc_mutual / orc_system_pot: Debye-Hueckel interaction of two spheres;
 *   orc_gradient / orc_forces / orc_compute_forces: central differences of the energy;
 *   orc_bd_*: drand48 against the C library's, and whole trajectories against the
 *     reference's own CBDnam::makeMove / compute_dt compiled from its sources.
 *   orc_system_total_int_energy, orc_pot_in_space: the reference's own computeTotalIntEnergy
 *     and computePotInSpace compiled from its sources (tests/golden/pot_ref.npz), and the
 *     rotation-correct potential by invariance under a rotation of the whole system.
 * Parity unpinned: the Jacobi variants (orc_selfpol_round_jacobi, schedule 1 of orc_selfpol / orc_mutual /
 * orc_gradient), which have no reference counterpart and serve only to check the
 * device's Jacobi rounds against the same restatement.
 */
#ifndef PBSAM_ORACLE_H
#define PBSAM_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_ctx orc_ctx;
typedef struct orc_moltype orc_moltype;
typedef struct orc_system orc_system;

/* context: physical constants + recursion tables for orders up to pmax */
orc_ctx *orc_create(double kappa, double eps_out, int pmax);
void orc_destroy(orc_ctx *c);

/* ---- primitives (function-level checks) ---- */
void orc_sh(orc_ctx *c, int P, double theta, double phi, double *out);
void orc_bessel_k(orc_ctx *c, int L, double val, double *out);
void orc_bessel_i(orc_ctx *c, int L, double val, double *out);
void orc_local_expan(orc_ctx *c, int n, const double *q, const double *xyz, int p,
                     int bkappa, double scale, double *out);
void orc_mul_expan(orc_ctx *c, int n, const double *q, const double *xyz, int p,
                   int bkappa, double scale, double *out);
int orc_compute_order(double rho, double tq, double scale, int p);
/* rotation coefficients R(n,s,m) for n<nmax (nmax<=2p-1), all |s|<=n, 0<=m<=n;
 * out: complex (re,im) in layout off(n)+(s+n)*(n+1)+m. returns singular flag */
int orc_rot_coeff(orc_ctx *c, double theta, double phi, double xi, int p, int nmax,
                  double *out, double *dout);
void orc_rotate(orc_ctx *c, double theta, double phi, double xi, int p, int p2, int fwd,
                const double *in, double *out);
void orc_rotate_quat(orc_ctx *c, const double *qwxyz, int p, int fwd, const double *in,
                     double *out);
void orc_rotate_grad(orc_ctx *c, double theta, double phi, double xi, int p, int which,
                     int fwd, const double *in, double *out); /* which 0: dT, 1: dP */
void orc_trans_coeff(orc_ctx *c, double rho, double kappa_seed, int p, double scalei,
                     double scaleo, double *T, double *dT);
void orc_translate(orc_ctx *c, double rho, double kappa_seed, int p, double scalei,
                   double scaleo, int grad, const double *in, double *out);
void orc_xform_h(orc_ctx *c, const double *P, double a_src, double a_tgt, int p,
                 const double *in, double *out);
void orc_xform_intra(orc_ctx *c, const double *P, double a_src, double a_tgt, int pH,
                     int pF, const double *H, const double *F, double *outH, double *outF);
void orc_xform_h_grad(orc_ctx *c, const double *P, double a_src, double a_tgt, int p,
                      const double *in, double *out3);
void orc_xform_num(orc_ctx *c, const double *P, int npts, const double *xyz,
                   const double *qF, const double *qH, double tQF, double tQH,
                   double a_src, double a_tgt, int p, double *outF, double *outH);
void orc_grad_local(orc_ctx *c, int npts, const double *g3, const double *xyz, int p,
                    int bkappa, double scale, double *out3);
void orc_sphere_points(int N, double r, double *xyz);
int orc_npts_rule(double r, int nin);
void orc_imat_build(orc_ctx *c, int p, int npe, const double *spe, int npb,
                    const double *spb, double *imat);

/* ---- molecule types ---- */
typedef struct orc_moltype_desc {
  int n_sph;
  const double *cen;   /* 3*n_sph, frame of the input file */
  const double *rad;   /* n_sph */
  int n_chg;
  const double *chg_xyz; /* 3*n_chg */
  const double *chg_q;   /* n_chg */
  const double *rcen;    /* 3: molecule centre */
  int p;
  double eps_in;
  const int *npts;     /* per-sphere spiral size, or NULL for the radius rule */
  int npts_cap;        /* cap for the radius rule (reference: 10000) */
  double intra_rcutoff;
  const double *imats; /* n_sph*p^4 column-major, or NULL: build (identity fast path) */
  int imat_npts;       /* spiral size for IMat quadrature when built */
} orc_moltype_desc;

orc_moltype *orc_moltype_create(orc_ctx *c, const orc_moltype_desc *d);
void orc_moltype_destroy(orc_moltype *m);
int orc_moltype_nsph(const orc_moltype *m);
int orc_moltype_npts(const orc_moltype *m, int k, int *nspx_total);
void orc_moltype_points(const orc_moltype *m, int k, double *xyz);
int orc_moltype_lists(const orc_moltype *m, int k, int *near, int *far, int *nfar);
void orc_moltype_consts(const orc_moltype *m, int k, double *c12p, double *dfix,
                        double *vfix, double *e0);
/* schedule: 0 = reference order (Gauss-Seidel, selective recompute, stock rule),
 * 1 = Jacobi rounds until max dev <= tol.  returns rounds used */
int orc_selfpol(orc_moltype *m, int schedule, double tol, int max_rounds);
void orc_selfpol_round_jacobi(orc_moltype *m, double *maxdev);
void orc_moltype_get(const orc_moltype *m, int k, double *F, double *H, double *qF,
                     double *qH, double *tq);
void orc_moltype_set_self(orc_moltype *m, const double *F, const double *H);
int orc_moltype_values(orc_moltype *m); /* exposed charges + intra self values; -1: no self values */
void orc_moltype_get_intra(const orc_moltype *m, int k, double *LF, double *LH,
                           double *LFfar, double *LHfar);
double orc_moltype_energy_self(const orc_moltype *m);

/* ---- systems of rigid molecules ---- */
orc_system *orc_system_create(orc_ctx *c, int n_mol, orc_moltype *const *mts,
                              const double *pos, const double *quat);
void orc_system_destroy(orc_system *s);
void orc_system_set_cutoffs(orc_system *s, double inter_rc, double interact_rc);
void orc_system_set_pose(orc_system *s, int i, const double *pos, const double *quat);
/* 0 ok, 1 overlap */
int orc_system_lists(orc_system *s);
/* schedule as for selfpol; hrot_refresh!=0 refreshes Hrot every round */
int orc_mutual(orc_system *s, int schedule, double tol, int max_rounds, int hrot_refresh);
int orc_gradient(orc_system *s, int schedule, double tol, int max_rounds);
void orc_forces(orc_system *s, double *force, double *torque);
int orc_compute_forces(orc_system *s, int schedule, double tol, int max_rounds,
                       double *force, double *torque);
void orc_system_get(const orc_system *s, int i, int k, double *F, double *H, double *LS,
                    double *Hrot);
void orc_system_get_grad(const orc_system *s, int i, int k, double *gH, double *gLHN);
double orc_system_pot(const orc_system *s, int i);
double orc_system_total_int_energy(orc_system *s);
/* computePotInSpace at n points (xyz: 3n); ref_frame != 0: the reference's unrotated formula */
void orc_pot_in_space(orc_system *s, int n, const double *xyz, double *out, int ref_frame);
int orc_system_counts(const orc_system *s, int *n_inter_pairs, int *n_polarized);

/* ---- Brownian dynamics of one two-molecule trajectory (CBDnam) ----
 * prm: dtr dr0 dr1 b_dist q_dist f_dist fact_tr ikbt coul_k dt_fixed(0: compute_dt)
 * max_tries.  place != 0: CBDnam::run's random start from srand48(seed); else pos(6),
 * quat(8).  Status: ESCAPED 0, DOCKED 1, RUNNING 2, STUCK 3 (-1: forces failed). */
typedef struct orc_bd orc_bd;
orc_bd *orc_bd_create(int ns0, const double *cen0, const double *rad0, int ns1,
                      const double *cen1, const double *rad1, const double *prm,
                      long long seed, int place, const double *pos, const double *quat);
void orc_bd_destroy(orc_bd *b);
void orc_bd_set_system(orc_bd *b, orc_system *s);
void orc_bd_add_contact(orc_bd *b, int ncontact, int npair, const int *pairs,
                        const double *dist);
int orc_bd_step(orc_bd *b, const double *ft);
void orc_bd_get(const orc_bd *b, double *pos, double *quat, double *rot, double *dt,
                long long *steps);
void orc_bd_rng(long long seed, int n, double *out);

#ifdef __cplusplus
}
#endif
#endif
