/*
 * mdgpu.h -- C ABI of libmdgpu, the MI355X (gfx950) nonbonded force path
 * for MDBas-style molecular dynamics.
 *
 * The structs below have the same member order and types as the
 * reference's include/global.h (enums are int-sized), so a reference host
 * can pass its own CTRL / PARAM / PBC / EWALD / NEIGH / ENERGY / BATH /
 * CONSTRAINT / BOND / ANGLE / DIHE pointers by cast.  Every drop-in entry
 * point takes the reference function's parameter list with a leading
 * context pointer and returns an int status instead of calling exit():
 * 0 on success, otherwise the reference's ERROR_TYPE number (errors.h)
 * where one applies, or one of the MDG_E* codes below.
 *
 * Threading: one host thread per context; every call is ordered on the
 * context's single HIP stream.  Host arrays stay owned by the caller; the
 * context owns all device memory.
 */
#ifndef MDGPU_H
#define MDGPU_H

#include <stdio.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- reference struct mirrors (global.h:165-316) ---------------------- */

typedef struct {                       /* PBC, global.h:175-184 */
    int type;                          /* 0 none, 1 cubic, 2 orthorhombic, 3 triclinic */
    double a, a1, a2, a3, b, b1, b2, b3, c, c1, c2, c3;
    double u, u1, u2, u3, v, v1, v2, v3, w, w1, w2, w3;
    double pa, pb, pc, det, vol, vol0;
    double stress1, stress2, stress3;
    double stress4, stress5, stress6;
    double stress7, stress8, stress9;
} MDG_PBC;

typedef struct {                       /* CTRL, global.h:195-217 */
    int newjob;
    int keyRand, keyTraj, keyProp, keyForF, keyRest;
    int printOut, printProp, printTraj, printRest;
    int keyMd, mdType;
    int seed;
    int keyMinim;
    int keyLink, noLink;
    int keyConstH;
    int keyNb14, keyNumForce;
    int keyEwald, keyAlpha, keyMmax;
    int elecType;                      /* 0 none, 1 full, 2 shift1, 3 shift2, 4 switch */
    int vdwType;                       /* 0 none, 1 full, 2 switch */
    int integrator;                    /* 0 leap-frog, 1 velocity Verlet */
    int ens;                           /* 0 NVE */
} MDG_CTRL;

typedef struct {                       /* PARAM, global.h:219-238 */
    int step, nSteps;
    double timeStep, rTimeStep;
    int nAtom, nDegFree, nFrozen;
    int nBond, nAngle, nDihedral, nImproper, nUb;
    int nConst, maxCycle;
    double tolShake;
    double chargeConst, delr, scal14;
    double cutOff, rcutOff, cutOff2, rcutOff2;
    double cutOn, cutOn2, switch2;
    double temp0, press0, kinTemp0;
    double tolMinim, maxminsiz, maxminst;
} MDG_PARAM;

typedef struct { double tauT, chiT, tauP, chiP, compress; } MDG_BATH;   /* global.h:240-245 */

typedef struct { int type; int a, b; double k, r0, beta; } MDG_BOND;     /* global.h:247-252 */
typedef struct { int type; int a, b, c; double k, theta0; } MDG_ANGLE;  /* global.h:254-259 */
typedef struct { int type; int order; int a, b, c, d; double k, phi0, mult; } MDG_DIHE;

typedef struct {                       /* NEIGH, global.h:269-276 */
    int update, linkRatio, nCells, sizeList, nPair14;
} MDG_NEIGH;

typedef struct {                       /* ENERGY, global.h:282-290 */
    double tot, pot, kin;
    double elec, vdw;
    double bond, ang, ub, dihe, impr;
    double conint, consv;
    double virelec, virvdw, virbond, virub;
    double virshake, virpot, virtot;
} MDG_ENERGY;

typedef struct { int a, b; double rc2; } MDG_CONSTRAINT;                /* global.h:296-300 */

typedef struct {                       /* EWALD, global.h:311-316 */
    int nbsp;
    int mmax, m1max, m2max, m3max;
    double prec, tol, tol1, alpha;
} MDG_EWALD;

#define MDG_FINAMELEN 512              /* FINAMELEN, global.h:61 */

typedef struct {                       /* ATOM, global.h:165-169 */
    char label[5], segn[5], resn[5];
    int type, resi, ires;
} MDG_ATOM;

typedef struct {                       /* IO, global.h:186-193 */
    FILE *confFile, *forfFile, *propFile, *restFile, *rconFile, *trajFile;
    FILE *simuFile;
    char confName[MDG_FINAMELEN], forfName[MDG_FINAMELEN], propName[MDG_FINAMELEN];
    char restName[MDG_FINAMELEN], rconName[MDG_FINAMELEN], trajName[MDG_FINAMELEN];
    char simuName[MDG_FINAMELEN];
} MDG_IO;

/* ---- library codes ---------------------------------------------------- */

enum {
    MDG_FP64 = 0,                      /* FP64 everywhere */
    MDG_MIXED = 1                      /* FP32 pair/spline math, int64 fixed-point forces */
};

enum {
    MDG_OK = 0,
    MDG_EHIP = 1001,                   /* a HIP / rocFFT call failed */
    MDG_EUNSUPPORTED = 1002,           /* a setting this path does not implement */
    MDG_ELIST = 1003,                  /* device list capacity exceeded */
    MDG_EARG = 1004,                   /* bad argument (NULL, size mismatch) */
    MDG_ESHAKE_ORDER = 1005,           /* a constraint group re-diverged after converging */
    MDG_ESCRATCH = 1006                /* input larger than a fixed scratch array */
};

/* Largest system mdg_nonbond_energy_noalloc takes (its fixed scratch). */
#define MDG_NOALLOC_MAX_ATOMS 262144

typedef struct mdg_ctx mdg_ctx;

/* Per-step energies written by mdg_run_steps (internal units, 10 J/mol). */
typedef struct {
    int step;
    int listRebuilt;
    double pot, kin, tot;
    double elec, vdw;                  /* elec = direct + reciprocal + correction */
    double eDir, eRec, eCorr;
    double virelec, virvdw, virshake;
    double nInCut;                     /* listed pairs with r <= rc this step */
    double ebond, eub, eang, edihe, eimpr;   /* bonded terms (ENERGY bond, ub, ang, dihe, impr);
                                              * elec / vdw include the 1-4 pairs */
    double virbond, virub;             /* bonded virials (ENERGY virbond, virub) */
    double conint, consv, chiT;        /* thermostat of the device step (mdg_set_bath): ENERGY
                                        * conint and consv, BATH chiT; zero for NVE */
} MDG_STEP;

/* ---- context ---------------------------------------------------------- */

/* Creates a context for nAtom = param->nAtom atoms: copies the run
 * settings, picks the list builder the way makelist does on its first call
 * (list.c:38-57), sizes the SPME mesh from ewald->m1max..m3max as given
 * (the reference's init_spme has already rounded them, spme.c:54-73) and
 * builds the rocFFT plans.  precision is MDG_FP64 or MDG_MIXED. */
mdg_ctx *mdg_create(const MDG_CTRL *ctrl, const MDG_PARAM *param, const MDG_PBC *box,
                    const MDG_EWALD *ewald, const MDG_NEIGH *neigh, int precision,
                    int *status);
void mdg_destroy(mdg_ctx *ctx);
const char *mdg_last_error(const mdg_ctx *ctx);
int mdg_precision(const mdg_ctx *ctx);
int mdg_key_link(const mdg_ctx *ctx);  /* 1 link cells, 0 O(N^2) */

/* Per-atom parameters and the constraint list after setup() (init.c:248-616):
 * charges, eps (sqrt(eps_kcal*418.4), io.c:1174), sig, masses, inverse
 * masses, frozen flags, constraint counts per atom.  Builds the exclusion
 * table from the constraints (consth; list.c:176-216 and the connectivity
 * walk 764-862). */
int mdg_set_topology(mdg_ctx *ctx, const double q[], const double eps[], const double sig[],
                     const double mass[], const double rmass[], const int frozen[],
                     const int nAtConst[], const MDG_CONSTRAINT constList[], int nConst);

/* Bonded topology for the device-resident step (mdg_run_steps): the term
 * tables of energy()'s bonded calls (param->nBond / nUb / nAngle /
 * nDihedral / nImproper entries of bond / ub / angle / dihe / impr,
 * internal.c:21-604) and, under ctrl->keyNb14, the 1-4 pairs with eps14 /
 * sig14 and param->scal14 (ewald14_energy, energy.c:654-742).  Rebuilds
 * the exclusion table from every term and the constraints of
 * mdg_set_topology, as makelist does on its first call (list.c:108-894 /
 * 1125-1791), and derives the 1-4 list from it.  Each device step then
 * adds these forces (FP64 terms, int64 fixed-point accumulation) and the
 * step records carry the bonded energies.  Call after mdg_set_topology.
 * No single reference counterpart: the reference hands the same tables to
 * energy() on every step (energy.h:30-37). */
int mdg_set_bonded(mdg_ctx *ctx, const MDG_CTRL *ctrl, const MDG_PARAM *param,
                   const MDG_BOND bond[], const MDG_BOND ub[], const MDG_ANGLE angle[],
                   const MDG_DIHE dihe[], const MDG_DIHE impr[], const double eps14[],
                   const double sig14[]);

/* ---- drop-in mirrors (host arrays in, host arrays out) ------------------ */

/* makelist, list.c:29-106 / list.h:28-31.  First call (ctrl->newjob == 1):
 * exclusions plus the O(N^2) or link-cell list, built on the GPU, returned
 * in the reference format through malloc'd out-parameters; the link-cell
 * build wraps x,y,z in place as list.c:1927 does.  Later calls rebuild when
 * param->step % neigh->update == 0 (realloc'ing *neighList if it grows).
 * Bonds, constraints (under keyConstH), angles, dihedrals and impropers
 * feed the exclusions (list.c:108-894 / 1125-1791); *neighList14 receives
 * the neigh->nPair14 1-4 pairs as (i, j) couples (list.c:864-873). */
int mdg_makelist(mdg_ctx *ctx, MDG_CTRL *ctrl, MDG_PARAM *param, MDG_PBC *box, MDG_NEIGH *neigh,
                 MDG_CONSTRAINT constList[], MDG_BOND bond[], MDG_ANGLE angle[],
                 MDG_DIHE dihe[], MDG_DIHE impr[], double x[], double y[], double z[],
                 int frozen[], int **neighList, int **neighPair, int **neighOrder,
                 int **neighList14, int ***exclList, int **exclPair);

/* ewald_energy, energy.c:450-643 / energy.h:49-53: SPME reciprocal part
 * (with its mean-force removal), real-space Ewald + switched LJ over the
 * given list, excluded-pair correction.  Forces and ENERGY fields
 * accumulate (+=) exactly as the reference's do; fx,fy,fz should hold
 * zeros on entry, as energy() guarantees. */
int mdg_ewald_energy(mdg_ctx *ctx, MDG_CTRL *ctrl, MDG_PARAM *param, MDG_ENERGY *ener,
                     MDG_EWALD *ewald, MDG_PBC *box, const double x[], const double y[],
                     const double z[], double fx[], double fy[], double fz[], const double q[],
                     const double eps[], const double sig[], const int neighList[],
                     const int neighPair[], const int neighOrder[], int **exclList,
                     const int exclPair[]);

/* spme_energy, spme.c:357-699 / spme.h:44-45: adds reciprocal forces to
 * fx,fy,fz, then removes the mean of the whole force arrays; writes
 * stress[6], *virEwaldRec and the energy (self and net-charge terms
 * included, computed on every call). */
int mdg_spme_energy(mdg_ctx *ctx, MDG_PARAM *param, MDG_EWALD *ewald, MDG_PBC *box,
                    const double x[], const double y[], const double z[], double fx[],
                    double fy[], double fz[], const double q[], double stress[6],
                    double *virEwaldRec, double *eEwaldRec);

/* nonbond_energy, energy.c:216-345 / energy.h:39-42: cutoff Coulomb with
 * the SHIFT1 / SHIFT2 / SWITCH functor of ctrl->elecType (elec.c:137-263)
 * plus switched LJ over the given list; accumulates into ener and f.
 * Deviation from the reference's parameter list: a leading ctrl.  The
 * reference picks the functor through the global pointers ptr_coulomb /
 * ptr_vdw that init_energy_ptrs sets from CTRL (energy.c:35-93); a
 * re-entrant library has no such globals, so the caller's CTRL is passed
 * instead. */
int mdg_nonbond_energy(mdg_ctx *ctx, MDG_CTRL *ctrl, MDG_PARAM *param, MDG_ENERGY *ener,
                       MDG_PBC *box, const double x[], const double y[], const double z[],
                       double fx[], double fy[], double fz[], const double q[],
                       const double eps[], const double sig[], const int neighList[],
                       const int neighPair[], const int neighOrder[]);

/* energy, energy.c:103-205 / energy.h:30-37, the reference's full
 * parameter list: zeroes the energy and virial fields, the stress and the
 * forces; runs the Ewald path (ewald_energy, keyEwald != 0) or the plain
 * cutoff path (nonbond_energy), each with its 1-4 pairs when keyNb14 is
 * set (ewald14_energy / nonbond14_energy over neighList14, neigh->nPair14
 * pairs, with eps14 / sig14 and param->scal14); then bond_energy,
 * angle_energy, ub_energy, dihedral_energy and improper_energy
 * (internal.c:21-604) for param->nBond / nAngle / nUb / nDihedral /
 * nImproper terms; sums pot and virpot; zeroes frozen atoms' forces and
 * velocities.  Every term runs on the GPU in FP64. */
int mdg_energy(mdg_ctx *ctx, MDG_CTRL *ctrl, MDG_PARAM *param, MDG_ENERGY *ener,
               MDG_EWALD *ewald, MDG_PBC *box, MDG_NEIGH *neigh, MDG_BOND bond[],
               MDG_BOND ub[], MDG_ANGLE angle[], MDG_DIHE dihe[], MDG_DIHE impr[],
               const double x[], const double y[], const double z[], double vx[], double vy[],
               double vz[], double fx[], double fy[], double fz[], const double q[],
               const double eps[], const double sig[], const double eps14[],
               const double sig14[], const int frozen[], const int neighList[],
               const int neighPair[], const int neighOrder[], const int neighList14[],
               int **exclList, const int exclPair[]);

/* vv_integrate, integrate.c:1479-1634 / integrate.h:72-76.  NVE: stage 1
 * = half kick, drift, SHAKE-R; stage 2 = half kick, SHAKE-V, kinetic
 * energy, kinetic stress, wrap.  SHAKE keeps the reference's global
 * stopping rule (shake.c:243-338) without a host round trip.  ctrl->ens
 * selects the thermostat or barostat as the reference's switch does
 * (integrate.c:1490-1511): NVT_B vv_nvt_b, NPT_B vv_npt_b, NVT_H vv_nvt_h,
 * NPT_H vv_npt_h (integrate.c:1636-2618), reading bath and the targets
 * param->kinTemp0 / temp0 / press0 and ener->virpot, and updating bath
 * (chiT, chiP), ener (kin, virshake, conint, consv) and, for NPT, the box
 * (every derived member, as scale_box / vv_scale_box do); later calls on
 * the context follow that box.  Per-atom work runs on the device, the bath
 * scalars on the host.  A changing box needs an MDG_FP64 context. */
int mdg_vv_integrate(mdg_ctx *ctx, MDG_CTRL *ctrl, MDG_PARAM *param, MDG_ENERGY *ener,
                     MDG_PBC *box, MDG_BATH *bath, MDG_CONSTRAINT constList[], double *x,
                     double *y, double *z, double *vx, double *vy, double *vz, double *fx,
                     double *fy, double *fz, double *mass, double *rmass, int *nAtConst,
                     int stage);

/* No counterpart in the reference: mdg_nonbond_energy's computation
 * (nonbond_energy, energy.c:216-345) for callers that cannot allocate
 * device memory -- code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc.  It allocates nothing and takes no
 * context: every array is the caller's DEVICE memory (x, y, z, q, eps, sig
 * and fx, fy, fz of param->nAtom atoms; neighList, neighPair, neighOrder
 * in the reference list format), forces and ener[0..3] (elec, vdw,
 * virelec, virvdw) accumulate (+=) as in the reference, and the work is
 * enqueued on stream (a hipStream_t; NULL for the default stream) without
 * a host wait.  List offsets and partial sums go to a fixed device array of
 * the library (one per GPU, held until the process ends), so at most
 * MDG_NOALLOC_MAX_ATOMS atoms (else MDG_ESCRATCH), and calls that overlap
 * on one GPU race; ordered calls do not.  ctrl->elecType selects the
 * functor as ptr_coulomb does (energy.c:35-93), ctrl->vdwType the LJ
 * switch. */
int mdg_nonbond_energy_noalloc(const MDG_CTRL *ctrl, const MDG_PARAM *param, const MDG_PBC *box,
                               const double *x, const double *y, const double *z, double *fx,
                               double *fy, double *fz, const double *q, const double *eps,
                               const double *sig, const int *neighList, const int *neighPair,
                               const int *neighOrder, double *ener, void *stream);

/* ---- device-resident MD ------------------------------------------------ */

int mdg_upload_state(mdg_ctx *ctx, const double x[], const double y[], const double z[],
                     const double vx[], const double vy[], const double vz[]);
int mdg_download_state(mdg_ctx *ctx, double x[], double y[], double z[], double vx[],
                       double vy[], double vz[], double fx[], double fy[], double fz[]);

/* The thermostat of the device-resident step (mdg_run_steps): ctrl->ens
 * NVE, NVT_B (vv_nvt_b, integrate.c:1636-1768) or NVT_H (vv_nvt_h,
 * integrate.c:1998-2169) with bath->tauT, bath->chiT, param->kinTemp0 and
 * ener->conint as the starting state.  The bath scalars live on the
 * device and are advanced by one-block kernels between the kinetic-energy
 * partials and a velocity-scaling pass, inside the step (and its graph);
 * each MDG_STEP carries conint, consv and chiT.  The barostats (NPT_B,
 * NPT_H) change the box, which the device step holds fixed: they return
 * MDG_EUNSUPPORTED here and run through mdg_vv_integrate.  A context made
 * with ctrl->ens != NVE runs mdg_run_steps only after this call. */
int mdg_set_bath(mdg_ctx *ctx, const MDG_CTRL *ctrl, const MDG_PARAM *param,
                 const MDG_BATH *bath, const MDG_ENERGY *ener);
/* The device thermostat's state now: bath->chiT, ener->conint, ener->consv. */
int mdg_get_bath(mdg_ctx *ctx, MDG_BATH *bath, MDG_ENERGY *ener);

/* Force and energy of the uploaded state (the step-0 energy() call of
 * mdbas.c:174), building the device list first. */
int mdg_initial_energy(mdg_ctx *ctx, MDG_STEP *rec);

/* nsteps MD steps in the order of mdbas.c:244-301 (VV stage 1, list
 * rebuild every neigh->update steps, energy, VV stage 2), entirely on the
 * device; with use_graph the step sequence is captured once in a hipGraph
 * and replayed.  records (may be NULL) receives nsteps entries, copied
 * after the last step: there is no per-step host round trip. */
int mdg_run_steps(mdg_ctx *ctx, int nsteps, int use_graph, MDG_STEP *records);

/* ---- replica batch ------------------------------------------------------ */

/* nrep independent periodic boxes advanced together in one context: atoms
 * (and constraints, by global index) of replica r are nAtom[r] consecutive
 * entries of the concatenated arrays; box[r] is its cubic or orthorhombic
 * box and mesh[r] its cubic SPME mesh edge; ctrl / param / ewald / neigh
 * hold the settings every replica shares (param->nAtom is ignored).  Every
 * replica must be one the reference lists with verlet_list (makelist's
 * O(N^2) choice, list.c:38-57).  The context then takes mdg_set_topology,
 * mdg_upload_state, mdg_download_state, mdg_initial_energy (nrep records)
 * and mdg_run_steps (records[step * nrep + r]); each step is one captured
 * graph whose kernels cover all replicas.  No reference counterpart (the
 * reference runs one system per process). */
mdg_ctx *mdg_batch_create(const MDG_CTRL *ctrl, const MDG_PARAM *param, const MDG_EWALD *ewald,
                          const MDG_NEIGH *neigh, int nrep, const int nAtom[], const MDG_PBC box[],
                          const int mesh[], int precision, int *status);
int mdg_batch_size(const mdg_ctx *ctx);   /* replicas, 0 for a single system */

/* ---- adjoint ------------------------------------------------------------ */

/* New per-atom charges and LJ parameters (eps = sqrt(eps_kcal*418.4),
 * sig as io.c:1174-1175 store them) for the current topology: the
 * parameters the adjoint differentiates.  No reference counterpart (the
 * reference reads them once, io.c:1150-1187). */
int mdg_set_nonbond_params(mdg_ctx *ctx, const double q[], const double eps[],
                           const double sig[]);

/* dL/dq, dL/deps, dL/dsig = dL_dE * dE_pot/d(q, eps, sig) per atom, FP64,
 * for the Ewald-path potential of the current device state (energy.c:450-643
 * with spme.c:357-699: real-space pairs over the device list, SPME, the
 * excluded-pair correction, the self and net-charge terms).  Call after
 * mdg_initial_energy or mdg_run_steps.  No reference counterpart: it is the
 * analytic form of what numderiv.c:39-190 estimates by finite differences. */
int mdg_backward(mdg_ctx *ctx, double dL_dE, double dL_dq[], double dL_deps[], double dL_dsig[]);

/* Runs nsteps MD steps from the current device state (as mdg_run_steps),
 * writes J = the mean potential energy over steps 1..nsteps, and returns
 * dJ/d(q, eps, sig) per atom through the whole trajectory: checkpoints
 * every ckpt steps, recomputation of each segment, reverse sweep over the
 * VV + SHAKE steps.  FP64 contexts of rigid three-site molecules.  The
 * device state is left at step nsteps.  No reference counterpart. */
int mdg_backward_traj(mdg_ctx *ctx, int nsteps, int ckpt, double *J, double dJ_dq[],
                      double dJ_deps[], double dJ_dsig[]);

/* Autotuner: times nsteps graph-replayed steps of the uploaded system for
 * each candidate launch shape (the three-site SHAKE kernels' molecules per
 * wavefront, the cluster pair kernel's grid, the branch order and wave
 * priority of the SPME kernels beside it, and whether the pruned list pays
 * for its prune kernel when pruning is on) and keeps the fastest; the
 * state is restored and the initial energy recomputed.  out[0] = SHAKE
 * spread (16/spread molecules per wavefront, or 0 for one molecule per
 * lane), out[1] = pair grid cap; the
 * pruning choice reads back through mdg_prune_skin.  No reference
 * counterpart. */
int mdg_autotune(mdg_ctx *ctx, int nsteps, int out[2]);
/* The launch shapes in use, as the autotuner sets them: [0] SHAKE spread,
 * [1] pair grid cap, [2] 1 when steps without a rebuild enqueue the SPME
 * branch before the pair branch, [3] 1 when the binned SPME kernels raise
 * their wave priority above the pair kernel's beside them, 2 when the
 * mixed pair kernel raises its own instead.  mdg_set_launch
 * restores a saved set (per-atom results do not depend on them).  No
 * reference counterpart. */
int mdg_get_launch(const mdg_ctx *ctx, int out[4]);
int mdg_set_launch(mdg_ctx *ctx, const int in[4]);

/* Average kernel durations (ms) of the last mdg_run_steps, measured with
 * HIP events on the context stream: [0] pair kernel, [1] SPME (spline to
 * gather), [2] integrator, [3] list rebuild (per rebuild), [4] whole step.
 * Events are recorded only when profiling is switched on. */
int mdg_set_profiling(mdg_ctx *ctx, int on);
int mdg_phase_times(mdg_ctx *ctx, double out[5]);

/* Dynamic pruning of the cluster-pair list (mixed precision).  After each
 * device step a kernel keeps, per work chunk, the entries with an atom pair
 * within cutOff + skin at that step's positions; the next step's pair
 * kernel runs on that shorter list while no atom has moved more than
 * skin / 2 since, and on the full list otherwise.  Forces, energies and
 * virials are the same bit for bit either way.  skin = 0 switches it off;
 * it is capped at the list skin (delr).  Default for mixed contexts:
 * 0.8 A where every box edge exceeds 2 (cutOff + 0.8 A + 3 A), else off
 * (the prune kernel's test needs that room, and small boxes gain
 * little).  No reference counterpart: the reference evaluates every
 * listed pair (energy.c:480-541), which this reproduces. */
int mdg_set_prune(mdg_ctx *ctx, double skin);
double mdg_prune_skin(const mdg_ctx *ctx);   /* the current pruning distance */
/* [0] pair kernel launches that ran on the pruned list (since creation),
 * [1] entries in the current pruned list. */
int mdg_prune_stats(mdg_ctx *ctx, long out[2]);

/* How the mixed pair kernel of one rectangular box forms its pair
 * displacements.  0 (default): in FP32 from each slot's offset to its
 * cluster's origin plus the entry's origin-to-origin offset, while every
 * offset stays below (shortest edge / 2 - cutOff) / 2 less 0.05 A, with the
 * cutoff test in FP64 from the positions for any entry that has a listed
 * pair whose FP32 r^2 lies within a rounding bound of cutOff^2 (so the
 * in-cutoff set is the reference's exactly); otherwise, and for the
 * cutoff functors, as 1.  1: in FP64 from the slot positions (the
 * reference's dist(), utils.c:38-80), converted to FP32.  2: as 0 with
 * the FP64 cutoff test on every entry (a check of the bound: its results
 * equal 0's bit for bit).  No reference counterpart. */
int mdg_set_pair_path(mdg_ctx *ctx, int path);

/* Counters of the device cluster-pair list: [0] cluster pairs,
 * [1] listed atom pairs (mask bits), [2] clusters. */
int mdg_list_stats(mdg_ctx *ctx, long out[3]);

/* The device cluster-pair list as atom pairs (i < j), sorted, for the
 * bit-exact comparison with the reference list: writes up to cap pairs,
 * returns the count through *npairs. */
int mdg_list_pairs(mdg_ctx *ctx, int *pairs_i, int *pairs_j, long cap, long *npairs);

/* ---- reference input, output and set-up (host side, no device work) ---- */

/* read_SIMU, io.c:29-726: defaults, then every keyword of io->simuName
 * (case-insensitive, up to "end"), then the derived cut-off constants
 * (io.c:714-723).  Returns 60 (no file), 61 (unknown keyword), 62 (bad
 * value), 63 (missing value) where the reference exits with those codes. */
int mdg_read_simu(MDG_IO *io, MDG_CTRL *ctrl, MDG_PARAM *param, MDG_BATH *bath,
                  MDG_NEIGH *neigh, MDG_EWALD *ewald, MDG_PBC *box);
/* read_CONF, io.c:728-783: '*' title lines, the atom count, then
 * "%d %d %4s %4s %lf %lf %lf %4s %d %lf" per atom; *atom, *x, *y, *z are
 * malloc'd. */
int mdg_read_conf(MDG_IO *io, MDG_PARAM *param, MDG_ATOM **atom, double **x, double **y,
                  double **z);
/* read_rest, io.c:785-870: the binary REST file mdg_write_rest writes. */
int mdg_read_rest(MDG_IO *io, MDG_PARAM *param, MDG_ENERGY *ener, MDG_BATH *bath,
                  MDG_ATOM **atom, double **x, double **y, double **z, double **vx, double **vy,
                  double **vz, double **fx, double **fy, double **fz);
/* read_FORF, io.c:872-1199: atoms, bonds, constraints, urey-bradley,
 * angles, dihedrals, impropers, vdw sections (energies in kcal/mol scaled
 * to internal units, angles in degrees to radians, eps = sqrt(eps*418.4));
 * term arrays are malloc'd, per-atom arrays are the caller's. */
int mdg_read_forf(MDG_IO *io, MDG_PARAM *param, MDG_ATOM atom[], MDG_CONSTRAINT **constList,
                  MDG_BOND **bond, MDG_ANGLE **angle, MDG_DIHE **dihe, MDG_DIHE **impr,
                  MDG_BOND **ub, double *eps, double *sig, double *eps14, double *sig14,
                  double *mass, double *q, int *frozen, int *nAtConst);
/* setup, init.c:248-616: masses of frozen atoms, consth constraints from
 * bonds with a hydrogen, terms dropped among frozen or fully constrained
 * atoms; the term arrays are reallocated (freed and set NULL when empty). */
int mdg_setup(MDG_CTRL *ctrl, MDG_PARAM *param, MDG_ATOM atom[], MDG_CONSTRAINT **constList,
              MDG_BOND **bond, MDG_ANGLE **angle, MDG_DIHE **dihe, MDG_DIHE **impr,
              MDG_BOND **ub, double mass[], double rmass[], int frozen[], int nAtConst[]);
/* init_system's Coulomb constant by mdType, init.c:102-109. */
double mdg_charge_const(int mdType);
/* init_box, init.c:840-908. */
void mdg_init_box(MDG_PBC *box);
/* get_kinfromtemp with get_degfree, utils.c:551-582. */
void mdg_get_kinfromtemp(MDG_PARAM *param, const MDG_PBC *box);
/* image_update, utils.c:162-218 (host arrays, in place). */
void mdg_image_update(const MDG_PARAM *param, const MDG_PBC *box, double x[], double y[],
                      double z[]);
/* kinetic, utils.c:502-515. */
double mdg_kinetic(const MDG_PARAM *param, const double vx[], const double vy[],
                   const double vz[], const double mass[]);
/* init_spme's Ewald width, tolerances and mesh, spme.c:32-73: call before
 * mdg_create so the context gets the reference's (power-of-two) mesh. */
int mdg_spme_mesh(const MDG_CTRL *ctrl, const MDG_PARAM *param, MDG_EWALD *ewald,
                  const MDG_PBC *box);
/* init_vel + init_constvel, init.c:627-833, with the 3N standard normal
 * deviates supplied by the caller (gauss: vx of every atom, then vy, then
 * vz) in place of the reference's Box-Muller draws from its dSFMT stream.
 * Returns 310 when the constraint projection does not converge. */
int mdg_init_vel(const MDG_PARAM *param, const MDG_PBC *box, const MDG_CONSTRAINT constList[],
                 const double x[], const double y[], const double z[], double vx[], double vy[],
                 double vz[], const double mass[], const double rmass[], const int frozen[],
                 const int nAtConst[], const double gauss[]);

/* write_CONF, io.c:1201-1228 (to io->rconName). */
int mdg_write_conf(MDG_IO *io, const MDG_PARAM *param, const MDG_ATOM atom[], const double x[],
                   const double y[], const double z[]);
/* write_prop, io.c:1230-1274: appends one 29-double record. */
int mdg_write_prop(MDG_IO *io, const MDG_PARAM *param, const MDG_ENERGY *ener,
                   const MDG_PBC *box);
/* write_rest, io.c:1276-1369. */
int mdg_write_rest(MDG_IO *io, const MDG_PARAM *param, const MDG_ENERGY *ener,
                   const MDG_BATH *bath, const MDG_ATOM atom[], const double x[],
                   const double y[], const double z[], const double vx[], const double vy[],
                   const double vz[], const double fx[], const double fy[], const double fz[]);
/* write_DCD_header, io.c:1371-1460. */
int mdg_write_dcd_header(MDG_IO *io, const MDG_CTRL *ctrl, const MDG_PARAM *param,
                         const MDG_PBC *box, const int frozen[]);
/* write_DCD_traj, io.c:1462-1522: the first frame after the header holds
 * every atom, later frames the free atoms only. */
int mdg_write_dcd_traj(MDG_IO *io, const MDG_PARAM *param, const MDG_PBC *box, const double x[],
                       const double y[], const double z[], const int frozen[]);

#ifdef __cplusplus
}
#endif
#endif
