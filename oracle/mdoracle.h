/*
 * mdoracle.h -- CPU FP64 oracle for the MDBas nonbonded force path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a from-scratch C restatement of the
 * reference semantics, used as the checker for the HIP library.  Only the
 * test suite, the smoke check and the CPU-baseline leg of the benchmark may
 * load it; the product library never links or calls it.
 *
 * Every function cites the reference file:line it restates.  Build with
 * -O2 -ffp-contract=off (no FMA contraction), which is what the reference's
 * x86-64 SSE2 build does, so that the neighbour-list distance tests round
 * exactly as the reference's do.
 *
 * Pinning (what checks each function beyond retyping it):
 *   oc_init_box, oc_nint, oc_dist        : compiled reference (oracle/_ref),
 *                                          special cases at +-0.5.
 *   oc_exclusions                        : compiled reference exclude_list /
 *                                          link_cell_exclude_list (as sets).
 *   oc_verlet_list, oc_link_cell_list    : compiled reference list builders,
 *                                          element-by-element; brute force.
 *   oc_ewald_real                        : compiled reference ewald_dir /
 *                                          ewald_corr / vdw_switch; finite
 *                                          differences; Newton's third law.
 *   oc_spme                              : classic Ewald (oc_ewald_classic and
 *                                          the compiled reference ewald_rec) at
 *                                          convergence; B-spline identities;
 *                                          FD of the energy.
 *   oc_fft3d                             : direct O(n^2) DFT on small grids.
 *   oc_spme_pot, oc_energy_grad          : central finite differences of the
 *                                          oracle energy (numderiv.c:39, 102,
 *                                          190 stencils); Euler's theorem for
 *                                          the degree-2 homogeneity of E_elec
 *                                          in q and E_vdw in eps.
 *   oc_vv_nve / oc_shake_*               : compiled reference vv_integrate /
 *                                          vv_shake_r / vv_shake_v.
 *   oc_hvp, oc_spme_hvp                  : central differences of the oracle's
 *                                          own forces and reciprocal gradient;
 *                                          the symmetry v.Hu = u.Hv; parameter
 *                                          differences of grad E . u.
 *   oc_traj_adjoint                      : central differences of J over whole
 *                                          oracle trajectories (SHAKE to 1e-12).
 *   oc_vv_ens                            : compiled reference vv_integrate for
 *                                          NVT_B, NPT_B, NVT_H, NPT_H (with and
 *                                          without constraints); the Berendsen
 *                                          kinetic-energy fixed point and the
 *                                          Nose-Hoover conserved quantity.
 */
#ifndef MDORACLE_H
#define MDORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Box: the nine box vectors in, everything else derived by oc_init_box.
 * type follows the reference BOX_TYPE: 0 none, 1 cubic, 2 orthorhombic,
 * 3 triclinic (global.h enum BOX_TYPE). */
typedef struct {
    int type;
    double a1, a2, a3, b1, b2, b3, c1, c2, c3;
    double u1, u2, u3, v1, v2, v3, w1, w2, w3;
    double pa, pb, pc, u, v, w, det, vol;
} oc_box;

/* Run parameters (the subset of the reference PARAM/EWALD/NEIGH used on
 * the nonbonded path) plus the derived constants of io.c:714-723. */
typedef struct {
    int nAtom;
    double cutOff, cutOn, delr, chargeConst;
    double alpha;
    int m1, m2, m3, nbsp;
    double timeStep, tolShake;
    int maxCycle, update, linkRatio;
    /* derived by oc_param_derive */
    double cutOff2, cutOn2, switch2, rcutOff, rcutOff2, rTimeStep;
} oc_param;

/* Energies / virials of one force evaluation (subset of ENERGY). */
typedef struct {
    double eDir, eCorr, eRec, eVdw;
    double virDir, virCorr, virRec, virVdw;
    double stressRec[6];
    int nInCut;      /* listed pairs with r2 <= cutOff2 */
    int nListed;     /* listed pairs */
} oc_energy;

void   oc_param_derive(oc_param *p);
void   oc_init_box(oc_box *b);
int    oc_nint(double x);
double oc_dist(const oc_box *b, double d[3]);
void   oc_image_update(const oc_box *b, int n, double *x, double *y, double *z);

/* Decide the list builder the way makelist does on its first call:
 * 1 = link cells, 0 = O(N^2) Verlet list. */
int    oc_choose_link(const oc_param *p, const oc_box *b);

/* Exclusion sets from bonds, constraints and angles plus the 1-3
 * connectivity walk.  Output CSR (off[n+1], lst).  sym=0 stores each pair
 * once at the smaller index (exclude_list); sym=1 stores it at both atoms
 * (link_cell_exclude_list).  Returns the number of entries, or -110 when a
 * 1-4 path exists (the reference raises DIHE_NONPARAM_ERROR without
 * dihedrals), or -(needed) when cap is too small. */
int oc_exclusions(int n, int nBond, const int *ba, const int *bb,
                  int nConst, const int *ca, const int *cb,
                  int nAngle, const int *aa, const int *ab, const int *ac,
                  int sym, int *off, int *lst, int cap);

/* O(N^2) Verlet list (list.c:896-1123).  Writes neighPair[n],
 * neighOrder[n] and up to cap entries of neighList; returns the pair count
 * (which may exceed cap: call again with a larger buffer). */
long oc_verlet_list(const oc_param *p, const oc_box *b,
                    const double *x, const double *y, const double *z,
                    const int *frozen, const int *exOff, const int *exLst,
                    int *neighList, long cap, int *neighPair, int *neighOrder);

/* Link-cell list (list.c:1793-2619).  build=1 wraps x,y,z first (the
 * first-call builder); build=0 is the update builder, which does not.
 * Returns the pair count, or -412 when there are too few cells. */
long oc_link_cell_list(const oc_param *p, const oc_box *b, int build,
                       double *x, double *y, double *z,
                       const int *frozen, const int *exOff, const int *exLst,
                       int *neighList, long cap, int *neighPair, int *neighOrder);

/* Real-space Ewald + switched LJ over the list plus the excluded-pair erf
 * correction (energy.c:480-627 with ewald.c:356-384, vdw.c:144-184).
 * Forces accumulate (+=) into fx,fy,fz. */
void oc_ewald_real(const oc_param *p, const oc_box *b,
                   const double *x, const double *y, const double *z,
                   double *fx, double *fy, double *fz,
                   const double *q, const double *eps, const double *sig,
                   const int *neighList, const int *neighPair, const int *neighOrder,
                   const int *exOff, const int *exLst, oc_energy *e);

/* Non-Ewald "plain" electrostatics with the SHIFT1 (2), SHIFT2 (3) or
 * SWITCH (4) functor plus switched LJ (energy.c:216-345, elec.c:137-263). */
void oc_nonbond(const oc_param *p, const oc_box *b, int elecType,
                const double *x, const double *y, const double *z,
                double *fx, double *fy, double *fz,
                const double *q, const double *eps, const double *sig,
                const int *neighList, const int *neighPair, const int *neighOrder,
                double *eElec, double *eVdw, double *virElec, double *virVdw);

/* One pair: the functors themselves, for table tests. */
double oc_ewald_dir(double alpha, double qel, double r, double rt, double *dE);
double oc_ewald_corr(double alpha, double qel, double r, double rt, double *dE);
double oc_vdw_switch(const oc_param *p, double veps, double vsig, double r2,
                     double rt, double *dE);

/* SPME reciprocal energy, forces (+=, then the mean of the whole force
 * array is removed as spme.c:648-670 does), stress[6] and virial.  The
 * self and net-charge terms are computed on every call. */
double oc_spme(const oc_param *p, const oc_box *b,
               const double *x, const double *y, const double *z,
               double *fx, double *fy, double *fz, const double *q,
               double stress[6], double *virRec);

/* dE_rec/dq per atom of the SPME energy (self / net-charge terms not
 * included). */
void oc_spme_pot(const oc_param *p, const oc_box *b, const double *x, const double *y,
                 const double *z, const double *q, double *dq);

/* Gradient of the Ewald-path potential (real space over the list,
 * excluded-pair correction, SPME, self and net-charge terms) with respect
 * to the per-atom q, eps, sig.  Overwrites dq, deps, dsig. */
void oc_energy_grad(const oc_param *p, const oc_box *b,
                    const double *x, const double *y, const double *z,
                    const double *q, const double *eps, const double *sig,
                    const int *neighList, const int *neighPair, const int *neighOrder,
                    const int *exOff, const int *exLst,
                    double *dq, double *deps, double *dsig);

/* B-spline values/derivatives for one fractional coordinate s, as bspgen
 * does (spme.c:269-355): bsp[i] = M_p(u+i), bsd[i] = M_p'(u+i). */
void oc_bspline(double s, int p, double *bsp, double *bsd);

/* |b(m)|^2 moduli for one axis (spme.c:170-267). */
void oc_bspmod(int m, int p, double *mod2);

/* Classic Ewald reciprocal sum over |m_i| <= kmax, for pinning SPME. */
double oc_ewald_classic(const oc_param *p, const oc_box *b, int kmax,
                        const double *x, const double *y, const double *z,
                        double *fx, double *fy, double *fz, const double *q);

/* In-place 3-D complex FFT (interleaved re,im doubles, index 3 fastest),
 * unnormalised, exponent sign `sign` (+1 = FFTW_BACKWARD). */
int oc_fft3d(double *data, int n1, int n2, int n3, int sign);

/* Velocity-Verlet NVE stage 1 or 2 (integrate.c:1518-1634) with SHAKE
 * (shake.c:209-457).  Returns 0, or 311 when SHAKE does not converge.
 * ekin/virshake receive the kinetic energy (stage 2) and SHAKE virial
 * (stage 1). */
int oc_vv_nve(const oc_param *p, const oc_box *b, int stage,
              int nConst, const int *ca, const int *cb, const double *rc2,
              double *x, double *y, double *z,
              double *vx, double *vy, double *vz,
              const double *fx, const double *fy, const double *fz,
              const double *mass, const double *rmass, const int *nAtConst,
              double *ekin, double *virshake, int *ncycle);

/* oc_vv_nve, recording each constraint's convergence test at every SHAKE
 * sweep: conv[sweep * nConst + c] (cap sweeps); *ncycle = sweeps run. */
int oc_vv_nve_conv(const oc_param *p, const oc_box *b, int stage,
                   int nConst, const int *ca, const int *cb, const double *rc2,
                   double *x, double *y, double *z, double *vx, double *vy, double *vz,
                   const double *fx, const double *fy, const double *fz,
                   const double *mass, const double *rmass, const int *nAtConst,
                   unsigned char *conv, int cap, int *ncycle);

/* Thermostat / barostat state of one VV call: BATH (global.h:240-245),
 * the PARAM targets, and the ENERGY members the integrators read (virpot)
 * and write (kin, virshake, conint, consv); stress[9] accumulates into the
 * box stresses stress1..stress9 as stage 2 does. */
typedef struct {
    int ens;                       /* ENSEMBLE_TYPE: 1 NVT_B, 2 NPT_B, 3 NVT_H, 4 NPT_H */
    double tauT, tauP, chiT, chiP;
    double kinTemp0, temp0, press0;
    double virpot, conint, consv, kin, virshake;
    double stress[9];
} oc_bath;

/* vv_nvt_b / vv_npt_b / vv_nvt_h / vv_npt_h (integrate.c:1636-2618), one
 * stage.  The NPT integrators rescale *b (scale_box, vv_scale_box,
 * utils.c:291-459).  Returns 0, 311 when SHAKE does not converge, 1 for an
 * ensemble outside 1..4. */
int oc_vv_ens(const oc_param *p, oc_box *b, oc_bath *s, int stage,
              int nConst, const int *ca, const int *cb, const double *rc2,
              double *x, double *y, double *z,
              double *vx, double *vy, double *vz,
              const double *fx, const double *fy, const double *fz,
              const double *mass, const double *rmass, const int *nAtConst, int *ncycle);

/* Whole MD run: energy at step 0, then nsteps of VV stage 1, list update
 * every p->update steps, energy, VV stage 2 (mdbas.c:160-300).
 * etot/epot/ekin (length nsteps+1, may be NULL) receive per-step energies.
 * keyLink: 1 link cells, 0 O(N^2).  Returns 0 or an error code. */
int oc_md_run(const oc_param *p, const oc_box *b, int keyLink, int nsteps,
              double *x, double *y, double *z,
              double *vx, double *vy, double *vz,
              double *fx, double *fy, double *fz,
              const double *q, const double *eps, const double *sig,
              const double *mass, const double *rmass, const int *frozen,
              const int *nAtConst, int nConst, const int *ca, const int *cb,
              const double *rc2, double *etot, double *epot, double *ekin,
              double *seconds);

/* ---- second derivatives and the trajectory adjoint -------------- */

/* H u (hu, 3n component-major) and d/d(q, eps, sig) (grad_x E . u) (mq,
 * me, ms) of the Ewald-path potential at (x, y, z): real space over the
 * list (pairs within cutOff), the excluded-pair correction over the
 * SYMMETRIC exclusion table (sym = 1 of oc_exclusions), SPME (splines,
 * their first and second derivatives, the dipole mesh sum_l q_l grad W_l
 * . u_l through the influence function).  All outputs accumulate. */
void oc_hvp(const oc_param *p, const oc_box *b, const double *x, const double *y,
            const double *z, const double *q, const double *eps, const double *sig,
            const int *neighList, const int *neighPair, const int *neighOrder,
            const int *exOff, const int *exLst, const double *u, double *hu, double *mq,
            double *me, double *ms);

/* The SPME part of oc_hvp; grad (3n, may be NULL) receives the raw
 * reciprocal gradient grad E_rec (before any mean removal); u may be NULL
 * when only grad is wanted. */
void oc_spme_hvp(const oc_param *p, const oc_box *b, const double *x, const double *y,
                 const double *z, const double *q, const double *u, double *hu, double *mq,
                 double *grad);

/* J = (1/T) sum_{t=1..T} E_pot(x_t) of T steps of the NVE MD loop of
 * oc_md_run from (x0, v0) (not modified), and dJ/d(q, eps, sig) by a
 * reverse sweep over the stored steps: SHAKE-V and SHAKE-R through the
 * implicit function theorem on each rigid molecule, the forces through
 * oc_hvp with the mean-force removal of spme.c:648-670.  Rigid three-site
 * molecules (constraints (0,1), (0,2), (1,2) of atoms 3w..3w+2), no frozen
 * atoms.  Returns 0, 311 (SHAKE) or 1 (layout not covered). */
int oc_traj_adjoint(const oc_param *p, const oc_box *b, int keyLink, int T,
                    const double *x0, const double *y0, const double *z0,
                    const double *vx0, const double *vy0, const double *vz0,
                    const double *q, const double *eps, const double *sig,
                    const double *mass, const double *rmass, const int *frozen,
                    const int *nAtConst, int nConst, const int *ca, const int *cb,
                    const double *rc2, double *J, double *dq, double *deps, double *dsig);

/* ---- bonded terms and 1-4 pairs ---------------------------------- */

/* Term records with the reference's layouts (BOND, ANGLE, DIHE of
 * global.h:247-267; enums are int-sized). */
typedef struct { int type; int a, b; double k, r0, beta; } oc_bond;
typedef struct { int type; int a, b, c; double k, theta0; } oc_angle;
typedef struct { int type; int order; int a, b, c, d; double k, phi0, mult; } oc_dihe;

/* Exclusions with dihedrals, impropers and the 1-4 list (list.c:108-894,
 * 1125-1791); d4 / i4 hold the a,b,c,d quadruples.  list14 needs room for
 * 5 nDihe pairs.  Return as oc_exclusions. */
int oc_exclusions14(int n, int nBond, const int *ba, const int *bb,
                    int nConst, const int *ca, const int *cb,
                    int nAngle, const int *aa, const int *ab, const int *ac,
                    int nDihe, const int *d4, int nImpr, const int *i4,
                    int sym, int *off, int *lst, int cap, int *list14, int *n14);

/* bond_energy, ub_energy, angle_energy, dihedral_energy, improper_energy
 * (internal.c:21-604) in energy()'s order (energy.c:163-189); forces
 * accumulate.  out = ebond, eub, eangle, edihe, eimpr, virbond, virub. */
void oc_bonded(const oc_box *b, int nBond, const oc_bond *bond, int nUb, const oc_bond *ub,
               int nAngle, const oc_angle *angle, int nDihe, const oc_dihe *dihe,
               int nImpr, const oc_dihe *impr,
               const double *x, const double *y, const double *z,
               double *fx, double *fy, double *fz, double out[7]);

/* 1-4 pairs: ewald14_energy (ewald = 1, energy.c:654-742) or
 * nonbond14_energy with the coulomb14 functor of elecType (energy.c:
 * 356-439); switched LJ when vdwOn.  out = elec, vdw, virelec, virvdw
 * (the amounts the reference adds to ENERGY). */
void oc_nb14(const oc_param *p, const oc_box *b, int ewald, int elecType, int vdwOn,
             double scal14, int nPair14, const int *list14,
             const double *x, const double *y, const double *z,
             double *fx, double *fy, double *fz,
             const double *q, const double *eps14, const double *sig14, double out[4]);

#ifdef __cplusplus
}
#endif
#endif
