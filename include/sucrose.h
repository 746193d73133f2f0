/*
 * sucrose.h -- C ABI of the MI355X (gfx950) implementation of the SUCROSE
 * per-step force-and-integrate path.
 *
 * The reference keeps one system in Fortran module globals and is driven by a
 * user loop (README.md:58-118) that calls the routines below.  This ABI keeps
 * the same routine names and argument meaning; the state lives on the GPU in
 * a context and moves to/from host arrays only on explicit upload/download.
 *
 * Conventions
 *  - Every entry point returns SX_OK (0) or a negative SX_ERR_* code; the
 *    message is available from sx_last_error().  The reference's `stop`
 *    conditions map to error codes (callers that want the reference's
 *    behaviour stop on a non-zero return).
 *  - Host arrays are caller-owned FP64 (Fortran real64), contiguous, read or
 *    written only during the call.  Particle indices passed in are 1-based
 *    (Fortran), exactly as the reference's I, J arguments.
 *  - Calls are ordered on the context's HIP stream; only calls that return
 *    data to the host synchronise.
 *  - One host thread per context.  Contexts are independent.
 */
#ifndef SUCROSE_H
#define SUCROSE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct sx_ctx sx_ctx;

/* ---- status codes ---- */
#define SX_OK              0
#define SX_ERR_ARG        -1   /* bad argument (index out of range, n <= 0, ...) */
#define SX_ERR_HIP        -2   /* HIP runtime error                                */
#define SX_ERR_STATE      -3   /* call order (e.g. forces before nlist_init)      */
#define SX_ERR_CAPACITY   -4   /* device capacity exceeded (grown for the next call) */
#define SX_ERR_CONSTRAINT -5   /* RATTLE failure: system.f90:1101-1105            */
#define SX_ERR_CUTOFF     -6   /* cutoff > box/2: system.f90:902-906              */
#define SX_ERR_NOTEMP     -7   /* target temperature unset: integrators.f90:311   */

/* ---- per-particle fields: the arrays of `module system` (system.f90:7-22)
 *      and RX_old.. (:22) ---- */
enum sx_field {
    SX_RX, SX_RY, SX_RZ, SX_VX, SX_VY, SX_VZ, SX_FX, SX_FY, SX_FZ,
    SX_QW, SX_QX, SX_QY, SX_QZ, SX_LX, SX_LY, SX_LZ, SX_TX, SX_TY, SX_TZ,
    SX_MASS, SX_INV_MASS, SX_CHARGE, SX_SIG_X, SX_SIG_Y, SX_SIG_Z,
    SX_EPS_X, SX_EPS_Y, SX_EPS_Z, SX_IXX, SX_IYY, SX_IZZ,
    SX_IXX_INV, SX_IYY_INV, SX_IZZ_INV, SX_RX_OLD, SX_RY_OLD, SX_RZ_OLD,
    SX_NFIELDS
};

/* ---- scalars: module system (system.f90:27-30) and nose_hoover
 *      (integrators.f90:229-230) state ---- */
enum sx_scalar {
    SX_POT_ENG, SX_KIN_ENG, SX_KIN_ROT, SX_TEMP, SX_PRESSURE, SX_DENSITY,
    SX_VIRIAL, SX_VIRIAL_BONDED, SX_TARGET_TEMP, SX_DOF,
    SX_HB0, SX_HB1, SX_HB2, SX_ETA0, SX_ETA1, SX_ETA2, SX_PETA0, SX_PETA1, SX_PETA2,
    SX_NSCALARS
};

/* ---- pair potentials: one code per `*_calculate_forces` routine ----
 * params[] per kind (same order as the Fortran dummy arguments):
 *   SX_LJ    sigma, epsilon, cutoff                 force.f90:6
 *   SX_LJ12  sigma, epsilon, cutoff                 force.f90:51
 *   SX_LJW   sigma, epsilon, cutoff, wf             force.f90:95
 *   SX_WCA   sigma, epsilon                         force.f90:155
 *   SX_MORSE dissoc, width, rmin, cutoff            force.f90:199
 *   SX_GB    cutoff [, mu, nu]  (reference mu=1 nu=2, force.f90:240)   :308
 *   SX_GBR   [unused, mu, nu]                       force.f90:430
 *   SX_ECP   cutoff                                 force.f90:1202
 *   SX_ECPR  (none)                                 force.f90:1387
 *   SX_DH    cutoff, db_len, dielectric             force.f90:2589
 *   SX_GBW   cutoff, mu, nu, wf                     force.f90:535 (welded)
 *   SX_GBC   cutoff, mu, nu, chirality              force.f90:665 (chiral)
 *   SX_GBWC  cutoff, mu, nu, wf, chirality          force.f90:845
 *   SX_ECPW  cutoff, wf                             force.f90:1553 (welded)
 *   SX_ECPC  cutoff, chirality                      force.f90:1744 (chiral)
 *   SX_ECPWC cutoff, wf, chirality                  force.f90:1986
 *   SX_DIPOLE cutoff, dielectric                    force.f90:2687 (point dipoles:
 *            the body z axis times the charge; sx_forces also applies the
 *            routine's reaction-field side effects, see sx_init_reaction_field)
 *   SX_COULOMB cutoff, dielectric                   force.f90:2627 (switched over
 *            the last 2 units before the cutoff)
 */
enum sx_kind {
    SX_LJ = 1, SX_LJ12, SX_LJW, SX_WCA, SX_MORSE, SX_GB, SX_GBR, SX_ECP, SX_ECPR, SX_DH,
    SX_GBW, SX_GBC, SX_GBWC, SX_ECPW, SX_ECPC, SX_ECPWC, SX_DIPOLE, SX_COULOMB
};

/* ECP contact-parameter solver (sx_set_ecp_solver) */
enum sx_ecp_solver {
    SX_ECP_BRENT_REFERENCE = 0, /* the reference's Brent, atol 1e-3 rtol 1e-2 (force.f90:1046) */
    SX_ECP_CONVERGED       = 1, /* Newton on dF/dlambda from a cold start to |dlambda| < 1e-10, Brent fallback */
    SX_ECP_LEARNED         = 2  /* as CONVERGED, warm-started by the lambda* MLP */
};

/* neighbour-list rebuild rule (sx_nlist_set_rule) */
enum sx_rebuild_rule {
    SX_REBUILD_REFERENCE = 0,  /* 4 max|dr|^2 > (rc+skin)^2   (system.f90:1033 after :933) */
    SX_REBUILD_HALF_SKIN = 1   /* max|dr| > skin/2: no interacting pair can be missed       */
};

/* ------------------------------------------------------------------ */
/* context                                                             */
/* ------------------------------------------------------------------ */
int  sx_create(int device, sx_ctx **out);
int  sx_destroy(sx_ctx *ctx);
const char *sx_last_error(const sx_ctx *ctx);   /* ctx may be NULL (last create error) */
int  sx_synchronize(sx_ctx *ctx);
int  sx_version(void);

/* ------------------------------------------------------------------ */
/* system state (module system, system.f90)                            */
/* ------------------------------------------------------------------ */
/* set_num_atoms (system.f90:43): allocates n = num_atoms*rigid_atoms
 * particles with the reference's sentinel values (-1 params, q = (1,0,0,0));
 * bonded terms registered for the previous system are dropped. */
int sx_set_num_atoms(sx_ctx *ctx, int num_atoms, int rigid_atoms);
/* set_box (system.f90:185) + set_periodic_boundary (:209) */
int sx_set_box(sx_ctx *ctx, const double box[3], int periodic);
int sx_get_box(sx_ctx *ctx, double box[3]);
/* set_density (system.f90:167-183): a cubic box of edge ((N/Nrigid)/dens)^(1/3) */
int sx_set_density(sx_ctx *ctx, double dens);
/* generate_fcc_lattice (system.f90:438-523): 4 unit_cells^3 sites of the four
 * body-diagonal orientations, the box from set_density, positions shifted by
 * -box/2; sets the number of atoms (rigid_atoms per site when > 0) if none
 * is set yet, as the reference does; the periodic flag is left as it is */
int sx_generate_fcc_lattice(sx_ctx *ctx, int unit_cells, double dens, int rigid_atoms);
/* count values of one field from 1-based atom first1 on: the per-atom setters
 * set_atom_mass/charge/shape/energy/position/velocity/inertia (system.f90:
 * 129-343) are one or three such calls */
int sx_set_atom_values(sx_ctx *ctx, int field, int first1, int count, const double *values);
int sx_get_atom_values(sx_ctx *ctx, int field, int first1, int count, double *values);
/* set_atom_type (system.f90:109) for all n atoms at once; rtype is host
 * metadata (-1 until set; initialize_system turns negatives into 0), read by
 * the trajectory and checkpoint reporters */
int sx_set_atom_types(sx_ctx *ctx, const int *rtype);
int sx_get_atom_types(sx_ctx *ctx, int *rtype);
int sx_upload(sx_ctx *ctx, int field, const double *host, int n);
int sx_download(sx_ctx *ctx, int field, double *host, int n);
/* device address of a field's n FP64 values (valid until the next sx_set_num_atoms) */
int sx_field_ptr(sx_ctx *ctx, int field, double **dev_ptr);
int sx_set_scalar(sx_ctx *ctx, int id, double value);
int sx_get_scalar(sx_ctx *ctx, int id, double *value);
/* set_num_constraints + add_constraint (system.f90:97, :345); 1-based atoms */
int sx_set_constraints(sx_ctx *ctx, int nc, const int *bbi1, const int *bbj1, const double *len);
/* initialize_system defaults (system.f90:525-718): mass 1, shape/energy 1,
 * inertia (b^2+c^2)/20, dof = 6N - NC - 3*remove_com, then initialize_forces */
int sx_initialize_system(sx_ctx *ctx, int remove_com);
/* set_velocity_to_temperature (system.f90:362): Philox normals, COM removal, scaling */
int sx_set_velocity_to_temperature(sx_ctx *ctx, double temperature, int seed, int remove_com);
int sx_initialize_forces(sx_ctx *ctx);          /* system.f90:423 */
int sx_calculate_thermo(sx_ctx *ctx);           /* system.f90:725 */

/* ------------------------------------------------------------------ */
/* neighbour list (module neighbour_list, system.f90:885-1053)         */
/* ------------------------------------------------------------------ */
int sx_nlist_init(sx_ctx *ctx, double skin, double cutoff);   /* init_nlist :895 (builds) */
int sx_nlist_build(sx_ctx *ctx);                              /* build_nlist :966 */
int sx_nlist_needs_rebuild(sx_ctx *ctx, int *out);            /* needs_rebuild :1005 */
int sx_nlist_update(sx_ctx *ctx);   /* needs_rebuild -> build, decided on the device */
int sx_nlist_set_rule(sx_ctx *ctx, int rule);
int sx_nlist_rebuild_count(sx_ctx *ctx, long *count);
/* the reference layout: offset(1..N) 1-based row starts with offset(N) = nnz+1,
 * nlist(1..nnz) 1-based J ascending within each row */
int sx_nlist_export(sx_ctx *ctx, int *offset1, int *nlist1, long cap, long *nnz);

/* ------------------------------------------------------------------ */
/* forces (module lennard_jones .. debye_huckel, force.f90)            */
/* ------------------------------------------------------------------ */
/* One call replaces the README pair loop (README.md:86-103) for one potential
 * over the current neighbour list: F(I) += FXIJ, F(J) -= FXIJ, T(I) += TXI,
 * T(J) += TXJ, pot_eng += PE. */
int sx_forces(sx_ctx *ctx, int kind, const double *params, int nparams);
/* Per-pair evaluation of `*_calculate_forces(I, J, ...)` for explicit 1-based
 * index pairs; host outputs pe[np], f[3*np], ti[3*np], tj[3*np] (ti/tj may be
 * NULL for isotropic kinds); lambda[np] (may be NULL) gets ECP's lambda. */
int sx_pair_eval(sx_ctx *ctx, int kind, const double *params, int nparams, long npairs,
                 const int *i1, const int *j1, double *pe, double *f, double *ti, double *tj,
                 double *lambda);
int sx_set_ecp_solver(sx_ctx *ctx, int solver);
/* Arithmetic of the pair terms, 64 (default) or 32 bits.  No counterpart in
 * the reference, whose force.f90 (6-49, 155-190, 308-428, 2589-2617) is FP64
 * throughout; the workloads quote FP32 pair terms.  At 32 bits the terms of
 * SX_LJ, SX_WCA and SX_DH (sx_forces, the WCA + Debye-Hueckel pass of the
 * step driver and the batched replicas) are evaluated in FP32 with hardware
 * reciprocal, square root and exp2, and SX_GB in an FP64 form with shared
 * divisions instead of the reference's expression order (FP32 is too coarse
 * for gb_ near contact); positions, the separation, the cutoff / neighbour
 * decisions, every sum over pairs, the integrators and all other kinds (ECP
 * included) stay FP64.  Invalidates captured steps. */
int sx_set_pair_precision(sx_ctx *ctx, int bits);
int sx_get_pair_precision(sx_ctx *ctx, int *bits);
/* ECP diagnostics of the last sx_forces(SX_ECP*): pairs screened, pairs solved,
 * total solver iterations, fallbacks from the warm start to Brent. */
int sx_ecp_stats(sx_ctx *ctx, long out[4]);
/* Nose-Hoover diagnostics (cumulative over sx_nht_* calls of 2048 particles or
 * more, which evaluate the chain recurrence in parallel by Newton iteration):
 * Newton iterations, calls, calls whose passes did not converge (evaluated
 * window by window instead). */
int sx_nht_stats(sx_ctx *ctx, long out[3]);

/* bonded terms (module bonded, force.f90:2244-2579); kind 1 bend_, 2 cosine_bend_,
 * 3 dih_, 4 cosine_dih_, 5 angle_bend_ (:2464, harmonic in the angle),
 * 6 angle_dih_ (:2508); idx1 holds 3 (bends: 1, 2, 5) or 4 (dihedrals: 3, 4, 6)
 * 1-based atoms per term.  Each call adds its terms to those registered;
 * sx_bonded_forces evaluates them all (sx_set_num_atoms drops them). */
int sx_set_bonded(sx_ctx *ctx, int kind, int count, const int *idx1, double coeff, double angle);
int sx_bonded_forces(sx_ctx *ctx);

/* ------------------------------------------------------------------ */
/* integrators (integrators.f90)                                       */
/* ------------------------------------------------------------------ */
int sx_vv_initial_step(sx_ctx *ctx, double dt);                       /* :7   */
int sx_vv_final_step(sx_ctx *ctx, double dt);                         /* :36  */
int sx_vvc_initial_step(sx_ctx *ctx, double dt);                      /* :55  */
int sx_vvc_final_step(sx_ctx *ctx, double dt);                        /* :89  */
int sx_qq_initial_step(sx_ctx *ctx, double dt);                       /* :123 */
int sx_qq_final_step(sx_ctx *ctx, double dt);                         /* :185 */
/* coupling: NULL = the reference default (HB = T tau^2, HB(1) = n_dof T tau^2) */
int sx_nht_initial_step(sx_ctx *ctx, double dt, const double *coupling);   /* :303 */
int sx_nht_final_step(sx_ctx *ctx, double dt, const double *coupling);     /* :356 */
int sx_nhtc_initial_step(sx_ctx *ctx, double dt, const double *coupling);  /* :394 */
int sx_nhtc_final_step(sx_ctx *ctx, double dt, const double *coupling);    /* :452 */
int sx_lgv_initial_step(sx_ctx *ctx, double dt, int step, double damping); /* :507 */
int sx_lgv_final_step(sx_ctx *ctx, double dt, double damping);             /* :629 */
int sx_lgvc_initial_step(sx_ctx *ctx, double dt, int step, double damping);/* :678 */
int sx_lgvc_final_step(sx_ctx *ctx, double dt, double damping);            /* :806 */
int sx_apply_constraints_a(sx_ctx *ctx, double dt);                   /* system.f90:1060 */
int sx_apply_constraints_b(sx_ctx *ctx, double dt);                   /* system.f90:1144 */

/* ------------------------------------------------------------------ */
/* whole steps, captured once as a HIP graph and replayed              */
/* ------------------------------------------------------------------ */
enum sx_integrator { SX_INT_VV = 1, SX_INT_NHT = 4, SX_INT_LGV = 6, SX_INT_LGVC = 7 };
typedef struct sx_step_spec {
    int integrator;        /* SX_INT_*; NHT also runs the quaternion steps (qq_) */
    int nforces;           /* number of pair potentials (<= 4)                    */
    int kinds[4];
    double params[4][8];
    int bonded;            /* evaluate registered bonded terms                    */
    double dt, damping;
    int first_step;        /* Langevin counter of the first replayed step; < 0:
                              continue from the device's counter (the step after
                              the last one run, or the one sx_load_checkpoint
                              restored)                                           */
} sx_step_spec;
/* Runs nsteps of: initial step, placement of registered composite-body
 * satellites, nlist update, initialize_forces, pair loops (then
 * eval_reaction_field when a kind is SX_DIPOLE), bonded, transfer of the
 * satellites' forces to their centres, final step, calculate_thermo.  The
 * first call captures the graph (again after a new target temperature or a
 * change of the list's storage).  Before replaying, the neighbour list is
 * given room for half as many entries again as its last build.  A rebuild
 * inside the replays that still does not fit empties the list from then on
 * (no pair forces, no out-of-range access); the call then grows the list,
 * rebuilds it at the current positions and returns SX_ERR_CAPACITY. */
int sx_run(sx_ctx *ctx, const sx_step_spec *spec, int nsteps);

/* ------------------------------------------------------------------ */
/* composite rigid bodies (system.f90:743-881)                          */
/* ------------------------------------------------------------------ */
/* copy_transform (:743-842) for count (from, to) pairs, 1-based, as if called
 * in order: `to` is placed at `from` plus offsets[3t..3t+2] rotated by from's
 * orientation and takes from's quaternion; set_inertia -1 (argument absent),
 * 0 or 1 (the inertia bookkeeping, with the reference's three errors).  The
 * `to` atoms must be distinct and none may be a `from` of the same call. */
int sx_copy_transform(sx_ctx *ctx, int count, const int *from1, const int *to1, const double *offsets,
                      int set_inertia);
/* transfer_forces (:844-881) for count (from, to) pairs, 1-based, as if
 * called in order: from's force and torque and the torque of that force about
 * `to` are added to `to`, from's are zeroed.  The `from` atoms must be
 * distinct and none may be a `to`. */
int sx_transfer_forces(sx_ctx *ctx, int count, const int *from1, const int *to1);
/* Composite bodies held on the device for repeated use: satellite t sits at
 * offsets[3t..] from centre centre1[t].  sx_composite_place runs the
 * copy_transform of every satellite, sx_composite_transfer the
 * transfer_forces of every satellite onto its centre; sx_run calls them
 * after the initial step and after the pair loops and bonded terms. */
int sx_set_composites(sx_ctx *ctx, int count, const int *centre1, const int *sat1, const double *offsets);
int sx_composite_place(sx_ctx *ctx);
int sx_composite_transfer(sx_ctx *ctx);

/* ------------------------------------------------------------------ */
/* reaction field of the point dipoles (module dipole, force.f90:2669-2767) */
/* ------------------------------------------------------------------ */
/* init_reaction_field (:2677): the dipole table DIP (3 x n) and the reaction
 * field RFL set to zero.  sx_forces(SX_DIPOLE) then does what the reference's
 * pair loop over dipole_calculate_forces does besides forces and torques: a
 * pair (i < j) within the cutoff sets DIP(i) to particle i's scaled axis and
 * adds 2(eps-1)/(2 eps+1)/rc^3 times particle j's scaled axis to the field of
 * EVERY particle (the whole-array update, :2745-2747), so the field is one
 * vector shared by all particles, accumulated over calls until the next
 * sx_init_reaction_field.  sx_pair_eval has no such side effects. */
int sx_init_reaction_field(sx_ctx *ctx);
/* eval_reaction_field (:2753-2766): for each charged particle pot_eng +=
 * DIP.RFL / 2 and T -= DIP x RFL */
int sx_eval_reaction_field(sx_ctx *ctx);
/* dip[3*n] (x then y then z rows), rfl[3] */
int sx_get_reaction_field(sx_ctx *ctx, double *dip, double rfl[3]);

/* ------------------------------------------------------------------ */
/* Monte Carlo barostat (module monte_carlo_barostat, integrators.f90:859-947) */
/* ------------------------------------------------------------------ */
int sx_set_target_pressure(sx_ctx *ctx, double pres);   /* set_target_pressure system.f90:224 */
/* mc_barostat (integrators.f90:864-945): an isotropic volume move of at most
 * volume_change * V (volume_change < 0: the reference's default 0.01), the
 * positions scaled with the box, accepted on log(u[1]) < -beta dU - beta P dV
 * + N log(V_new/V_old) with beta = 1/temp.  compute_potential is the
 * neighbour list rebuilt at the current box, initialize_forces and the pair
 * loops (and, with spec->bonded, the bonded terms) of spec (its integrator,
 * dt, damping and first_step are not used).  u[0], u[1] are the two numbers
 * the reference draws with random_number.  An accepted move leaves the
 * forces, pot_eng and the list of the new state, sets density = N / V and,
 * as compute_potential's initialize_forces does in the reference, zero
 * kin_eng, kin_rot, temp, pressure and the virials; a rejected one (or
 * V_new <= 0) restores the box, the positions, the list, the forces, torques
 * and every scalar the move started from.  RX_old.. hold the positions
 * before the move, as in the reference.  accept may be NULL. */
int sx_mc_barostat(sx_ctx *ctx, const sx_step_spec *spec, double volume_change, const double u[2], int *accept);

/* ------------------------------------------------------------------ */
/* reporters (reporters.f90): text files in the reference's formats     */
/* ------------------------------------------------------------------ */
enum sx_report_kind {
    SX_REPORT_STATE_DATA = 1,  /* module state_data_reporter   reporters.f90:1-46    */
    SX_REPORT_TRAJECTORY = 2,  /* module trajectory_reporter   reporters.f90:48-101  */
    SX_REPORT_CONFIG     = 3,  /* module config_reporter       reporters.f90:103-139 */
    SX_REPORT_CHECKPOINT = 4   /* module checkpoint_reporter   reporters.f90:141-231 */
};
typedef struct sx_reporter sx_reporter;
/* open_state_data_file / open_trajectory_file / open_confwriter_file /
 * open_checkpoint_file (reporters.f90:7, :54, :109, :148): creates (truncates)
 * the file; the state-data reporter writes its header line.  ctx may be NULL
 * for a reporter fed only by sx_report_host. */
int sx_reporter_open(sx_ctx *ctx, int kind, const char *filename, sx_reporter **out);
/* report_state_data / report_trajectory / report_confwriter /
 * report_checkpoint (reporters.f90:20, :65, :120, :160) of the context's state
 * as it stands after the work queued before the call: the frame is copied on
 * the device, in stream order, and formatted and written by the reporter's
 * host thread; the call does not wait for the device.  step is printed by the
 * state-data and trajectory frames.  A checkpoint also writes <file>.state
 * (target temperature, dof, Nose-Hoover HB/eta/p_eta, Langevin step counter,
 * periodic flag), which the reference's checkpoint does not hold. */
int sx_report(sx_reporter *r, int step);
/* the same frame from host arrays (fields: SX_NFIELDS pointers indexed by
 * sx_field, only the frame's fields read; scal: SX_NSCALARS values; bonds
 * 1-based), written before the call returns */
int sx_report_host(sx_reporter *r, int step, int n, const double *const *fields, const int *rtype,
                   const double *scal, const double box[3], int nc, const int *bbi1, const int *bbj1,
                   const double *bblen);
/* config_index written at the end of each checkpoint (0 at open, reporters.f90:156) */
int sx_reporter_set_config_index(sx_reporter *r, int config_index);
int sx_reporter_flush(sx_reporter *r);        /* waits until every frame is written */
const char *sx_reporter_error(const sx_reporter *r);
int sx_reporter_close(sx_reporter *r);        /* close_*_file: flush, close, free   */
/* A reporter fed by a context must be closed before that context is
 * destroyed or given a different number of atoms by sx_load_checkpoint or
 * sx_set_num_atoms while frames are in flight (sx_reporter_flush first). */
/* load_checkpoint (reporters.f90:188-221): set_num_atoms, set_num_constraints,
 * set_box, then the atom and bond lines (the mass of a line goes to the atom
 * the line names, the other columns to the line's position; inv_mass and the
 * inverse inertia stay -1 as set_num_atoms leaves them) and config_index
 * (may be NULL); restores <file>.state when present */
int sx_load_checkpoint(sx_ctx *ctx, const char *filename, int *config_index);
/* the parse of sx_load_checkpoint into host arrays (no device): sizes first;
 * SX_ERR_CAPACITY when n > cap_n or nc > cap_nc; fields[k] may be NULL */
int sx_read_checkpoint_host(const char *filename, int *n, int *nc, double box[3], int *config_index, int cap_n,
                            int cap_nc, double *const *fields, int *rtype, int *bbi1, int *bbj1, double *bblen);

/* ------------------------------------------------------------------ */
/* batched replicas: independent small systems advanced together, one       */
/* launch per step (or three, by autotuning), graph-replayed               */
/* ------------------------------------------------------------------ */
/* The context's particles are split into nsys contiguous systems of counts[s]
 * particles, each in its own cubic periodic box of edge box[s], with pair
 * kind kind[s] (SX_LJ: params sigma, epsilon, cutoff; SX_GB: cutoff, mu, nu)
 * from params[4 s ..], and Langevin temperature temp[s].  Each step runs
 * lgv_initial_step (Philox counters: the particle's 1-based index within its
 * system), initialize_forces, the pair loop over all pairs I < J of the
 * system, lgv_final_step and the system's thermo.  Interaction ranges must not
 * exceed half the box (SX_ERR_CUTOFF). */
int sx_batch_set(sx_ctx *ctx, int nsys, const int *counts, const double *box, const int *kind, const double *params,
                 const double *temp);
int sx_batch_run(sx_ctx *ctx, int nsteps, double dt, double damping, int first_step);
/* per system: KE, KErot, PE, T = 2 (KE + KErot) / (6 N - 3); out[4 nsys] */
int sx_batch_thermo(sx_ctx *ctx, double *out);
/* times nrep steps of each of the 10 launch shapes from the same state,
 * restores it, keeps the fastest; ms[10] (may be NULL) gets ms per step.
 * Shapes 0-7 take three launches per step (initial step, forces, final
 * step); as (particles per force workgroup, lanes per particle): 0 (64,1)
 * 1 (16,4) 2 (32,4) 3 (32,8) 4 (64,8) 5 (16,64) 6 (32,64) 7 (64,64); 64 lanes
 * means one wavefront per particle that queues the pairs in range and
 * evaluates them 64 at a time.  Shape 0 adds each particle's pair terms in
 * the README loop's order; the others in a fixed tree order.  Shapes 8 and
 * 9 run the whole step in one launch: workgroups of 256 or 512 threads take
 * whole systems from a queue, most expensive first, one compacting
 * wavefront per particle. */
int sx_batch_autotune(sx_ctx *ctx, double dt, double damping, int first_step, int nrep, int *choice, double ms[10]);
int sx_batch_set_shape(sx_ctx *ctx, int choice);

/* ------------------------------------------------------------------ */
/* measurement and synthetic set-up (no counterpart in the reference:   */
/* benchmark and test harness support)                                 */
/* ------------------------------------------------------------------ */
/* HIP-event timing of the ECP pipeline on the current state, each stage
 * launched nrep times in a captured graph as inside a replayed step: the
 * median time per launch over three timed replays after an untimed one,
 * ms = {per-particle precompute, screen, solve, apply} (apply as screen +
 * solve + apply less the screen and the solve alone);
 * stats = {half pairs screened, pairs solved, solver iterations} per run.
 * Forces, torques and pair energies are restored afterwards. */
int sx_profile_ecp(sx_ctx *ctx, int kind, const double *params, int nparams, int nrep, double ms[4],
                   double stats[3]);
/* random sequential addition of identical ellipsoids (semi-axes a, b, c) with
 * uniform orientations in a periodic cube of edge L; pos[3n], quat[4n]
 * (w x y z); returns the number placed */
int sx_gen_ellipsoids_rsa(int n, double a, double b, double c, double L, uint32_t seed, long max_attempts,
                          double *pos, double *quat);
/* number of overlapping pairs (Perram-Wertheim contact function < 1) */
long sx_count_ellipsoid_overlaps(int n, double a, double b, double c, double L, const double *pos,
                                 const double *quat);
/* n uniformly distributed unit quaternions (w x y z), Philox stream of seed */
void sx_gen_uniform_quaternions(int n, uint32_t seed, double *quat);
/* self-avoiding random-walk chains (bond length `bond`, no two beads closer
 * than min_dist) in a periodic cube of edge L; pos[3 nchains nbeads];
 * returns the number of complete chains */
int sx_gen_chains(int nchains, int nbeads, double L, double bond, double min_dist, uint32_t seed,
                  long max_attempts, double *pos);

/* ------------------------------------------------------------------ */
/* learned lambda* warm start for the ECP contact solve (no counterpart  */
/* in the reference: it accelerates the solve of force.f90:1046-1200)    */
/* ------------------------------------------------------------------ */
/* Network 7 -> 64 -> 64 -> 64 -> 2 (SiLU), BF16 matrix-core products with FP32
 * accumulation.  Inputs: unit separation in particle i's body frame and the
 * relative quaternion conj(q_i) q_j (w >= 0); outputs: lambda* and the contact
 * distance over 2 max(semi-axis).  Parameters: one flat FP32 vector
 * W1 (64x7) b1 W2 (64x64) b2 W3 (64x64) b3 W4 (2x64) b4, W row-major (out, in). */
int sx_mlp_num_params(void);
int sx_mlp_init(sx_ctx *ctx, uint32_t seed);                 /* uniform +-1/sqrt(fan_in), Philox */
int sx_mlp_set_params(sx_ctx *ctx, const float *params);
int sx_mlp_get_params(sx_ctx *ctx, float *params);
/* outputs Y[n][2] for inputs X[n][7] */
int sx_mlp_forward(sx_ctx *ctx, long n, const float *X, float *Y);
/* loss sum_o w_o (y_o - t_o)^2 / n of an explicit batch and its gradient */
int sx_mlp_loss_grad(sx_ctx *ctx, long n, const float *X, const float *T, const float w[2], float *loss,
                     float *grad);
/* appends the ECP work list of the current state (kind SX_ECP/SX_ECPR) as
 * training samples: inputs, converged lambda*, contact distance / (2 max a);
 * the device store holds up to `capacity` samples; *total = samples held */
int sx_mlp_collect(sx_ctx *ctx, int kind, const double *params, int nparams, long capacity, long *total);
int sx_mlp_samples(sx_ctx *ctx, long first, long n, float *X, float *T);
int sx_mlp_clear_samples(sx_ctx *ctx);
/* Adam (lr, beta 0.9/0.999, eps 1e-8) on minibatches of `batch` samples (a
 * multiple of 64) drawn with Philox(seed); loss_trace[steps] may be NULL */
int sx_mlp_train(sx_ctx *ctx, int steps, int batch, double lr, const float w[2], uint32_t seed, float *loss_trace);
/* contact-parameter solvers on the current state's ECP work list, each timed
 * alone (HIP events, nrep runs).  out[19]: [0] live pairs, ms of [1] Brent
 * (reference tolerances) [2] Newton from psi [3] network [4] Newton from the
 * network's warm start [5] full converged solve (lambda + forces) [6] full
 * learned solve incl. network [7] Brent at 1e-10; iterations per pair [8]
 * plain Newton [9] warm Newton; mean per-wave maximum iterations [10] plain
 * [11] warm; |lambda0 - lambda*| [12] mean [13] max [14] rms; [15] max
 * |lambda_warm - lambda_plain|; [16] iterations per pair and [17] mean
 * per-wave maximum of the 1e-10 Brent; [18] pairs the learned solve proved out
 * of range from the warm start alone (F(lambda0) <= F(lambda*)) */
int sx_ecp_compare_solvers(sx_ctx *ctx, int kind, const double *params, int nparams, int nrep, double out[19]);

/* ------------------------------------------------------------------ */
/* Philox4x32-10 (xmath.f90:14-113)                                    */
/* ------------------------------------------------------------------ */
/* host: the reference's fixed key */
void sx_philox4x32_10(const uint32_t ctr[4], uint32_t out[4]);
/* device batch of philox_normal_3seq(i, c2, c3, c4) for i = 1..n -> out[3*n] */
int sx_philox_normals(sx_ctx *ctx, int n, int c2, int c3, int c4, double *out);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): sx_philox_normals with the same results, staged in a
 * fixed per-GPU scratch array (4 MiB, set up by sx_create) instead of a buffer
 * allocated per call.  n <= 174762 (SX_ERR_CAPACITY beyond).  It calls no
 * hipMalloc or hipFree.  Calls on one GPU share its scratch: calls that
 * overlap (other host threads, other contexts on the same GPU) race. */
int sx_philox_normals_noalloc(sx_ctx *ctx, int n, int c2, int c3, int c4, double *out);

#ifdef __cplusplus
}
#endif
#endif
