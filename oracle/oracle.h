/*
 * CPU oracle: an FP64 C restatement of the SUCROSE per-step force-and-integrate
 * path (fortran/xmath.f90, fortran/system.f90, fortran/force.f90,
 * fortran/integrators.f90).  TEST INFRASTRUCTURE ONLY: tests, the smoke check
 * and the benchmark's CPU-baseline leg use it as the checker.  No product code
 * links or calls it.
 *
 * Every routine keeps the reference's operation order (no FMA contraction,
 * ANINT = round-half-away-from-zero) so that, on the same FP64 inputs, its
 * outputs can be compared bit-for-bit with the reference built from source
 * (oracle/_ref).  Indices are 0-based; the Philox counters keep the
 * reference's 1-based particle index.
 */
#ifndef SUCROSE_ORACLE_H
#define SUCROSE_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Mirror of the module globals of `module system` (fortran/system.f90:7-33)
 * plus the thermostat state of `module nose_hoover` (integrators.f90:227-230).
 * All arrays are caller-owned, length n (bonds: length nc). */
typedef struct or_sys {
    int n, nc, nrigid, dof, periodic, _pad;
    double box[3];
    double *rx, *ry, *rz, *vx, *vy, *vz, *fx, *fy, *fz;
    double *qw, *qx, *qy, *qz, *lx, *ly, *lz, *tx, *ty, *tz;
    double *mass, *inv_mass, *charge;
    double *sig_x, *sig_y, *sig_z, *eps_x, *eps_y, *eps_z;
    double *ixx, *iyy, *izz, *ixx_inv, *iyy_inv, *izz_inv;
    double *rx_old, *ry_old, *rz_old;
    int *bbi, *bbj;          /* 0-based bond atoms */
    double *bb_len;
    double pot_eng, kin_eng, kin_rot, temp, pressure, density, virial, virial_bonded;
    double target_temp;
    double hb[3], eta[3], p_eta[3];   /* Nose-Hoover chain, N_th = 3 */
} or_sys;

/* Half Verlet list in the reference CSR layout (system.f90:927-1003) but
 * 0-based: row i holds the j > i with r^2 < (rc+skin)^2 in ascending order,
 * entries offset[i] .. offset[i+1]-1.  offset has n+1 entries. */
typedef struct or_nlist {
    int n;
    long cap;
    long nnz;
    long *offset;
    int *nbr;
    double *x_last, *y_last, *z_last;
    double cutoff, skin, r_skin_sq;
} or_nlist;

/* ---- Philox4x32-10 and Box-Muller (xmath.f90:14-113) ---- */
void or_philox4x32_10(const uint32_t ctr[4], uint32_t out[4]);   /* reference key */
void or_philox4x32_10_key(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
void or_philox_uniform(const int32_t ctr[4], double out[4]);
void or_philox_normal_3seq(int32_t c1, int32_t c2, int32_t c3, int32_t c4,
                           double *n1, double *n2, double *n3);

/* ---- 3x3 helpers used by ECP (xmath.f90:141-304); column-major a[i + 3*j] ---- */
void or_cholesky_solve(const double a[9], const double b[3], double out[3]);
void or_inverse(const double m[9], double out[9]);
double or_determinant(const double m[9]);

/* ---- system set-up (system.f90:362-523, 525-741) ---- */
void or_set_velocity_to_temperature(or_sys *s, double rtemp, int seed, int remove_com);
void or_initialize_defaults(or_sys *s, int remove_com);  /* initialize_system defaults */
void or_initialize_forces(or_sys *s);
void or_calculate_thermo(or_sys *s);
/* generate_fcc_lattice positions/orientations for n = 4*uc^3 sites; sets box. */
void or_generate_fcc_lattice(or_sys *s, int unit_cells, double dens);

/* ---- neighbour list (system.f90:895-1051) ---- */
int  or_nlist_build(or_sys *s, or_nlist *nl);           /* returns 0 or -1 on capacity */
int  or_nlist_needs_rebuild(const or_sys *s, const or_nlist *nl);

/* ---- pair kernels (force.f90); i, j 0-based, reference order I=i, J=j ---- */
void or_lj(const or_sys *s, int i, int j, double sigma, double epsilon, double cutoff,
           double *pe, double f[3]);
void or_lj12(const or_sys *s, int i, int j, double sigma, double epsilon, double cutoff,
             double *pe, double f[3]);
void or_ljw(const or_sys *s, int i, int j, double sigma, double epsilon, double cutoff,
            double wf, double *pe, double f[3]);
void or_wca(const or_sys *s, int i, int j, double sigma, double epsilon,
            double *pe, double f[3]);
void or_morse(const or_sys *s, int i, int j, double dissoc, double width, double rmin,
              double cutoff, double *pe, double f[3]);
void or_dh(const or_sys *s, int i, int j, double cutoff, double db_len, double dielectric,
           double *pe, double f[3]);
void or_coulomb(const or_sys *s, int i, int j, double cutoff, double dielectric, double *pe, double f[3]);
int or_copy_transform(or_sys *s, int from, int to, double xo, double yo, double zo, int set_inertia);
void or_transfer_forces(or_sys *s, int from, int to);
void or_dipole(const or_sys *s, int i, int j, double cutoff, double dielectric, double *pe, double f[3],
               double ti[3], double tj[3], double *dip, double *rfl);
void or_eval_reaction_field(or_sys *s, const double *dip, const double *rfl);
void or_dipole_loop(or_sys *s, const or_nlist *nl, double cutoff, double dielectric, double *dip, double *rfl);
/* Gay-Berne; mu, nu are the exponents the reference fixes to (1, 2)
 * (force.f90:240).  variant: 0 gb (cutoff + smooth switch), 1 gbr (repulsive). */
/* variant 0 gb_, 1 gbr_, 2 gbw_, 3 gbc_, 4 gbwc_ */
void or_gbx(const or_sys *s, int i, int j, int variant, double cutoff, double wf, double chirality, int mu, int nu,
            double *pe, double f[3], double ti[3], double tj[3]);
void or_gb(const or_sys *s, int i, int j, int variant, double cutoff, int mu, int nu,
           double *pe, double f[3], double ti[3], double tj[3]);
/* ECP (Perram-Wertheim).  variant: 0 ecp (cutoff + switch), 1 ecpr (repulsive).
 * lambda_out (may be NULL) receives the Brent lambda; niter_out the iteration
 * count of the Brent loop. */
/* variant 0 ecp_, 1 ecpr_, 2 ecpw_, 3 ecpc_, 4 ecpwc_ */
void or_ecpx(const or_sys *s, int i, int j, int variant, double cutoff, double wf, double chirality,
             double *pe, double f[3], double ti[3], double tj[3], double *lambda_out, int *niter_out);
void or_ecpx_at_lambda(const or_sys *s, int i, int j, int variant, double cutoff, double wf, double chirality,
                       double lambda, double *pe, double f[3], double ti[3], double tj[3]);
void or_ecp(const or_sys *s, int i, int j, int variant, double cutoff,
            double *pe, double f[3], double ti[3], double tj[3],
            double *lambda_out, int *niter_out);
/* brent(optim_dist) on given AR, BR (column-major) and r (force.f90:1046-1200). */
void or_ecp_at_lambda(const or_sys *s, int i, int j, int variant, double cutoff, double lambda,
                      double *pe, double f[3], double ti[3], double tj[3]);
void or_ecp_shape(const or_sys *s, int i, int j, double ar[9], double br[9], double rij[3]);
double or_ecp_lambda_converged(const double ar[9], const double br[9], const double r[3], double l0, double tol,
                               int *iters);
double or_ecp_contact_distance(const double ar[9], const double br[9], const double r[3], double l);
void or_ecp_features(const or_sys *s, int i, int j, double x[7]);
double or_ecp_brent(const double ar[9], const double br[9], const double r[3], int *niter);
double or_ecp_optim_dist(double l, const double ar[9], const double br[9], const double r[3]);

/* bonded (force.f90:2244-2579); f arrays are length 3 each */
void or_bend(const or_sys *s, int i, int j, int k, double coeff,
             double *pe, double fi[3], double fj[3], double fk[3]);
void or_cosine_bend(const or_sys *s, int i, int j, int k, double coeff, double angle,
                    double *pe, double fi[3], double fj[3], double fk[3]);
void or_dih(const or_sys *s, int i, int j, int k, int l, double coeff,
            double *pe, double fi[3], double fj[3], double fk[3], double fl[3]);
void or_cosine_dih(const or_sys *s, int i, int j, int k, int l, double coeff, double angle,
                   double *pe, double fi[3], double fj[3], double fk[3], double fl[3]);
void or_angle_bend(const or_sys *s, int i, int j, int k, double coeff, double angle,
                   double *pe, double fi[3], double fj[3], double fk[3]);
void or_angle_dih(const or_sys *s, int i, int j, int k, int l, double coeff, double angle,
                  double *pe, double fi[3], double fj[3], double fk[3], double fl[3]);

/* ---- integrators (integrators.f90) ---- */
void or_vv_initial_step(or_sys *s, double dt);
void or_vv_final_step(or_sys *s, double dt);
void or_vvc_initial_step(or_sys *s, double dt);
void or_vvc_final_step(or_sys *s, double dt);
void or_qq_initial_step(or_sys *s, double dt);
void or_qq_final_step(or_sys *s, double dt);
/* coupling: NULL = reference default (HB = T tau^2, HB1 = n_dof T tau^2) */
void or_nht_initial_step(or_sys *s, double dt, const double *coupling);
void or_nht_final_step(or_sys *s, double dt, const double *coupling);
void or_nhtc_initial_step(or_sys *s, double dt, const double *coupling);
void or_nhtc_final_step(or_sys *s, double dt, const double *coupling);
void or_lgv_initial_step(or_sys *s, double dt, int step, double damping);
void or_lgv_final_step(or_sys *s, double dt, double damping);
void or_lgvc_initial_step(or_sys *s, double dt, int step, double damping);
void or_lgvc_final_step(or_sys *s, double dt, double damping);
/* RATTLE (system.f90:1060-1209); return 0, or -(1+bond) on constraint failure */
int or_apply_constraints_a(or_sys *s, double dt);
int or_apply_constraints_b(or_sys *s, double dt);
/* exprel of module nose_hoover (integrators.f90:233-272) */
double or_exprel(double x);

/* ---- whole-loop helpers: the README pair loop (README.md:86-103) over a half list ---- */
/* kind codes as in include/sucrose.h: 1 LJ, 2 LJ12, 3 LJW, 4 WCA, 5 MORSE, 6 GB, 7 GBR,
 * 8 ECP, 9 ECPR, 10 DH.  p = potential parameters (see sx_forces_* in sucrose.h). */
void or_pair_loop(or_sys *s, const or_nlist *nl, int kind, const double *p);
/* all-cores variants (CPU baseline): the same results as or_pair_loop and
   or_nlist_build; buf holds 10 doubles per half-list entry, cnt n longs */
void or_pair_loop_threads(or_sys *s, const or_nlist *nl, int kind, const double *p, double *buf);
int or_nlist_build_threads(or_sys *s, or_nlist *nl, long *cnt);
int or_threads(void);

#ifdef __cplusplus
}
#endif
#endif
