/*
 * ref_driver.c -- flat-array entry points around the compiled reference
 * (test infrastructure only).
 *
 * oracle/Makefile compiles the reference's own src/list.c, ewald.c, vdw.c,
 * elec.c, utils.c, shake.c, integrate.c, init.c, energy.c, internal.c,
 * memory.c, errors.c and io.c from where they lie under the reference
 * tree, links them with this file into oracle/_ref/libmdbas_ref.so,
 * exports only the ref_* symbols and lets the linker drop every function
 * nothing here reaches.  That drops energy.c's ewald_energy (it calls
 * spme_energy, which lives in spme.c, which needs the absent fftw3.h) and
 * init.c's init_system (it calls init_spme), so no stand-in for any
 * missing header or library is written.
 *
 * This file plays the role of the reference's mdbas.c for the objects it
 * links: it owns outFile and the structs the reference functions take.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "global.h"
#include "list.h"
#include "ewald.h"
#include "vdw.h"
#include "elec.h"
#include "utils.h"
#include "shake.h"
#include "integrate.h"
#include "energy.h"
#include "internal.h"

FILE *outFile;   /* defined by mdbas.c in the reference */

void init_box(PBC *box);   /* init.c:840 */

typedef struct {
    CTRL ctrl;
    PARAM param;
    PBC box;
    NEIGH neigh;
    EWALD ewald;
    ENERGY ener;
    BATH bath;
    CONSTRAINT *cons;
    int *frozen, *nAtConst;
    int *neighList, *neighPair, *neighOrder, *neighList14, **exclList, *exclPair;
    int listed;
    int shakeAlloc;
    BOND *bond, *ub;
    ANGLE *angle;
    DIHE *dihe, *impr;
} ref_sys;

static void ensure_out(void)
{
    if (!outFile) outFile = fopen("/dev/null", "w");
}

/* box9 = a1 a2 a3 b1 b2 b3 c1 c2 c3; out[24] = u1..w3 (9), pa pb pc,
 * u v w, det vol, a b c (norms), 3 spare. */
void ref_init_box(int type, const double *box9, double *out)
{
    PBC b;
    memset(&b, 0, sizeof b);
    b.type = (enum BOX_TYPE)type;
    b.a1 = box9[0]; b.a2 = box9[1]; b.a3 = box9[2];
    b.b1 = box9[3]; b.b2 = box9[4]; b.b3 = box9[5];
    b.c1 = box9[6]; b.c2 = box9[7]; b.c3 = box9[8];
    init_box(&b);
    double o[24] = {b.u1, b.u2, b.u3, b.v1, b.v2, b.v3, b.w1, b.w2, b.w3,
                    b.pa, b.pb, b.pc, b.u, b.v, b.w, b.det, b.vol, b.a, b.b, b.c, 0, 0, 0, 0};
    memcpy(out, o, sizeof o);
}

int ref_nint(double x) { return nint(x); }

ref_sys *ref_create(int n, int boxType, const double *box9,
                    double cutOff, double cutOn, double delr, double alpha,
                    int linkRatio, int update, int noLink,
                    double timeStep, double tolShake, int maxCycle,
                    int nConst, const int *ca, const int *cb, const double *rc2,
                    const int *frozen, const int *nAtConst)
{
    ensure_out();
    ref_sys *s = (ref_sys *)calloc(1, sizeof(ref_sys));
    s->box.type = (enum BOX_TYPE)boxType;
    s->box.a1 = box9[0]; s->box.a2 = box9[1]; s->box.a3 = box9[2];
    s->box.b1 = box9[3]; s->box.b2 = box9[4]; s->box.b3 = box9[5];
    s->box.c1 = box9[6]; s->box.c2 = box9[7]; s->box.c3 = box9[8];
    init_box(&s->box);

    s->ctrl.newjob = 1;
    s->ctrl.noLink = noLink;
    s->ctrl.keyConstH = 1;
    s->ctrl.elecType = SWITCH;
    s->ctrl.vdwType = VSWITCH;
    s->ctrl.integrator = VELOCITY;
    s->ctrl.ens = NVE;
    s->ctrl.keyEwald = 2;

    s->param.nAtom = n;
    s->param.nConst = nConst;
    s->param.cutOff = cutOff;
    s->param.cutOn = cutOn;
    s->param.delr = delr;
    s->param.chargeConst = chgcharmm * kcaltoiu;   /* init.c:102-103 */
    s->param.timeStep = timeStep;
    s->param.tolShake = tolShake;
    s->param.maxCycle = maxCycle;
    /* io.c:714-723 */
    s->param.rcutOff = 1. / s->param.cutOff;
    s->param.cutOff2 = X2(s->param.cutOff);
    s->param.rcutOff2 = 1. / s->param.cutOff2;
    s->param.cutOn2 = X2(s->param.cutOn);
    s->param.switch2 = 1. / X3(s->param.cutOff2 - s->param.cutOn2);
    s->param.rTimeStep = 1. / s->param.timeStep;

    s->neigh.update = update;
    s->neigh.linkRatio = linkRatio;
    s->ewald.alpha = alpha;

    s->cons = (CONSTRAINT *)malloc((size_t)(nConst > 0 ? nConst : 1) * sizeof(CONSTRAINT));
    for (int i = 0; i < nConst; i++) {
        s->cons[i].a = ca[i];
        s->cons[i].b = cb[i];
        s->cons[i].rc2 = rc2[i];
    }
    s->frozen = (int *)malloc((size_t)n * sizeof(int));
    s->nAtConst = (int *)malloc((size_t)n * sizeof(int));
    memcpy(s->frozen, frozen, (size_t)n * sizeof(int));
    memcpy(s->nAtConst, nAtConst, (size_t)n * sizeof(int));
    return s;
}

/* makelist (list.c:29-106): first call builds exclusions and the list
 * (link cells or O(N^2)); later calls rebuild when step % update == 0.
 * x,y,z may be wrapped in place, as the reference does. */
int ref_makelist(ref_sys *s, double *x, double *y, double *z, int step)
{
    s->param.step = step;
    makelist(&s->ctrl, &s->param, &s->box, &s->neigh, s->cons, s->bond, s->angle, s->dihe,
             s->impr, x, y, z, s->frozen, &s->neighList, &s->neighPair, &s->neighOrder,
             &s->neighList14, &s->exclList, &s->exclPair);
    s->listed = 1;
    return s->ctrl.keyLink;
}

long ref_list_count(ref_sys *s)
{
    long t = 0;
    for (int i = 0; i < s->param.nAtom; i++) t += s->neighPair[i];
    return t;
}

void ref_list_copy(ref_sys *s, int *neighList, int *neighPair, int *neighOrder)
{
    long t = ref_list_count(s);
    memcpy(neighList, s->neighList, (size_t)t * sizeof(int));
    memcpy(neighPair, s->neighPair, (size_t)s->param.nAtom * sizeof(int));
    memcpy(neighOrder, s->neighOrder, (size_t)s->param.nAtom * sizeof(int));
}

/* Exclusions as CSR; rows beyond the table (the ordered table has N-1
 * rows) are empty. */
int ref_excl_copy(ref_sys *s, int *off, int *lst, int cap)
{
    int n = s->param.nAtom, o = 0;
    int rows = s->ctrl.keyLink ? n : n - 1;
    for (int i = 0; i < n; i++) {
        off[i] = o;
        if (i < rows)
            for (int k = 0; k < s->exclPair[i]; k++) {
                if (o < cap) lst[o] = s->exclList[i][k];
                o++;
            }
    }
    off[n] = o;
    return o;
}

double ref_ewald_dir(double alpha, double qel, double r, double rt, double *dE)
{
    EWALD e;
    memset(&e, 0, sizeof e);
    e.alpha = alpha;
    return ewald_dir(&e, dE, qel, r, rt);
}

double ref_ewald_corr(double alpha, double qel, double r, double rt, double *dE)
{
    EWALD e;
    memset(&e, 0, sizeof e);
    e.alpha = alpha;
    return ewald_corr(&e, dE, qel, r, rt);
}

double ref_vdw_switch(ref_sys *s, double veps, double vsig, double r2, double rt, double *dE)
{
    return vdw_switch(&s->param, dE, veps, vsig, r2, rt);
}

double ref_coulomb(ref_sys *s, int type, double qel, double r2, double rt, double *dE)
{
    if (type == SHIFT1) return coulomb_shift1(&s->param, dE, qel, r2, rt);
    if (type == SHIFT2) return coulomb_shift2(&s->param, dE, qel, r2, rt);
    if (type == SWITCH) return coulomb_switch(&s->param, dE, qel, r2, rt);
    *dE = 0.;
    return 0.;
}

/* energy.c:216-345 over the current list; out = elec, vdw, virelec, virvdw. */
void ref_nonbond(ref_sys *s, int elecType, const double *x, const double *y, const double *z,
                 double *fx, double *fy, double *fz, const double *q,
                 const double *eps, const double *sig, double *out)
{
    s->ctrl.elecType = (enum ELEC_TYPE)elecType;
    s->ctrl.vdwType = VSWITCH;
    init_energy_ptrs(&s->ctrl);
    memset(&s->ener, 0, sizeof s->ener);
    nonbond_energy(&s->param, &s->ener, &s->box, x, y, z, fx, fy, fz, q, eps, sig,
                   s->neighList, s->neighPair, s->neighOrder);
    out[0] = s->ener.elec;
    out[1] = s->ener.vdw;
    out[2] = s->ener.virelec;
    out[3] = s->ener.virvdw;
}

/* Classic Ewald (ewald.c:10-354) with explicit kmax on each axis. */
double ref_ewald_rec(ref_sys *s, int kmax, const double *x, const double *y, const double *z,
                     double *fx, double *fy, double *fz, const double *q,
                     double *stress, double *vir)
{
    s->ctrl.keyAlpha = 1;
    s->ctrl.keyMmax = 1;
    s->ewald.prec = 1e-6;
    s->ewald.m1max = s->ewald.m2max = s->ewald.m3max = kmax;
    init_ewald(&s->ctrl, &s->param, &s->ewald, &s->box);
    double e = ewald_rec(&s->param, &s->ewald, &s->box, x, y, z, fx, fy, fz, q, stress, vir);
    ewald_free(&s->ewald);
    return e;
}

/* vv_integrate (integrate.c:1479-1634) for NVE.  out = kin, virshake. */
/* integrate.c and shake.c keep their work arrays in file-scope globals
 * sized for one system and (integrators_allocate_arrays,
 * integrate.c:37-109) one ensemble: allocate again whenever the system,
 * its size or the ensemble differs from the last allocation (the old
 * arrays are left allocated). */
static void integ_alloc(ref_sys *s)
{
    static const ref_sys *sys;
    static int ens = -1, n = -1, nc = -1;
    if (sys != s || ens != (int)s->ctrl.ens || n != s->param.nAtom || nc != s->param.nConst) {
        integrators_allocate_arrays(&s->ctrl, &s->param);
        shake_allocate_arrays(&s->param);
        sys = s;
        ens = (int)s->ctrl.ens;
        n = s->param.nAtom;
        nc = s->param.nConst;
    }
    s->shakeAlloc = 1;
}

void ref_vv(ref_sys *s, int stage, double *x, double *y, double *z,
            double *vx, double *vy, double *vz, double *fx, double *fy, double *fz,
            double *mass, double *rmass, double *out)
{
    integ_alloc(s);
    vv_integrate(&s->ctrl, &s->param, &s->ener, &s->box, &s->bath, s->cons, x, y, z,
                 vx, vy, vz, fx, fy, fz, mass, rmass, s->nAtConst, stage);
    out[0] = s->ener.kin;
    out[1] = s->ener.virshake;
}

/* vv_integrate with a thermostat or barostat (integrate.c:1479-1516 ->
 * 1636-2618).  st (in/out): ens, tauT, tauP, chiT, chiP, kinTemp0, temp0,
 * press0, virpot, conint, consv, kin, virshake, box a1..c3 (9), then the
 * nine box stresses; the box is re-derived by init_box on the way in, and
 * the stresses start from st and accumulate. */
void ref_vv_ens(ref_sys *s, int stage, double *st, double *x, double *y, double *z,
                double *vx, double *vy, double *vz, double *fx, double *fy, double *fz,
                double *mass, double *rmass)
{
    s->ctrl.ens = (enum ENSEMBLE_TYPE)(int)st[0];
    integ_alloc(s);
    s->bath.tauT = st[1];
    s->bath.tauP = st[2];
    s->bath.chiT = st[3];
    s->bath.chiP = st[4];
    s->param.kinTemp0 = st[5];
    s->param.temp0 = st[6];
    s->param.press0 = st[7];
    s->ener.virpot = st[8];
    s->ener.conint = st[9];
    s->ener.consv = st[10];
    s->ener.kin = st[11];
    s->ener.virshake = st[12];
    s->box.a1 = st[13]; s->box.a2 = st[14]; s->box.a3 = st[15];
    s->box.b1 = st[16]; s->box.b2 = st[17]; s->box.b3 = st[18];
    s->box.c1 = st[19]; s->box.c2 = st[20]; s->box.c3 = st[21];
    init_box(&s->box);
    s->box.stress1 = st[22]; s->box.stress2 = st[23]; s->box.stress3 = st[24];
    s->box.stress4 = st[25]; s->box.stress5 = st[26]; s->box.stress6 = st[27];
    s->box.stress7 = st[28]; s->box.stress8 = st[29]; s->box.stress9 = st[30];
    vv_integrate(&s->ctrl, &s->param, &s->ener, &s->box, &s->bath, s->cons, x, y, z,
                 vx, vy, vz, fx, fy, fz, mass, rmass, s->nAtConst, stage);
    st[3] = s->bath.chiT;
    st[4] = s->bath.chiP;
    st[9] = s->ener.conint;
    st[10] = s->ener.consv;
    st[11] = s->ener.kin;
    st[12] = s->ener.virshake;
    st[13] = s->box.a1; st[14] = s->box.a2; st[15] = s->box.a3;
    st[16] = s->box.b1; st[17] = s->box.b2; st[18] = s->box.b3;
    st[19] = s->box.c1; st[20] = s->box.c2; st[21] = s->box.c3;
    st[22] = s->box.stress1; st[23] = s->box.stress2; st[24] = s->box.stress3;
    st[25] = s->box.stress4; st[26] = s->box.stress5; st[27] = s->box.stress6;
    st[28] = s->box.stress7; st[29] = s->box.stress8; st[30] = s->box.stress9;
    s->ctrl.ens = NVE;
}

/* init_constvel (init.c:724-833): projects velocities onto the
 * constraint manifold. */
void init_constvel(PARAM *param, PBC *box, CONSTRAINT constList[], double x[],
                   double y[], double z[], double vx[], double vy[], double vz[],
                   double mass[], int nAtConst[]);

void ref_constvel(ref_sys *s, double *x, double *y, double *z,
                  double *vx, double *vy, double *vz, double *mass)
{
    init_constvel(&s->param, &s->box, s->cons, x, y, z, vx, vy, vz, mass, s->nAtConst);
}

/* Bonded topology (global.h:247-267 records) for makelist's exclusions
 * and the bonded energies; scal14 is PARAM.scal14. */
void ref_set_bonded(ref_sys *s, int nBond, const BOND *bond, int nUb, const BOND *ub,
                    int nAngle, const ANGLE *angle, int nDihe, const DIHE *dihe,
                    int nImpr, const DIHE *impr, double scal14)
{
    s->param.nBond = nBond;
    s->param.nUb = nUb;
    s->param.nAngle = nAngle;
    s->param.nDihedral = nDihe;
    s->param.nImproper = nImpr;
    s->param.scal14 = scal14;
    s->bond = (BOND *)malloc((size_t)(nBond + 1) * sizeof(BOND));
    s->ub = (BOND *)malloc((size_t)(nUb + 1) * sizeof(BOND));
    s->angle = (ANGLE *)malloc((size_t)(nAngle + 1) * sizeof(ANGLE));
    s->dihe = (DIHE *)malloc((size_t)(nDihe + 1) * sizeof(DIHE));
    s->impr = (DIHE *)malloc((size_t)(nImpr + 1) * sizeof(DIHE));
    memcpy(s->bond, bond, (size_t)nBond * sizeof(BOND));
    memcpy(s->ub, ub, (size_t)nUb * sizeof(BOND));
    memcpy(s->angle, angle, (size_t)nAngle * sizeof(ANGLE));
    memcpy(s->dihe, dihe, (size_t)nDihe * sizeof(DIHE));
    memcpy(s->impr, impr, (size_t)nImpr * sizeof(DIHE));
}

int ref_n14(ref_sys *s) { return s->neigh.nPair14; }

void ref_list14_copy(ref_sys *s, int *out)
{
    memcpy(out, s->neighList14, (size_t)(2 * s->neigh.nPair14) * sizeof(int));
}

/* energy()'s bonded calls (energy.c:163-189) in its order; forces
 * accumulate.  out = bond, ub, ang, dihe, impr, virbond, virub. */
void ref_bonded(ref_sys *s, const double *x, const double *y, const double *z,
                double *fx, double *fy, double *fz, double *out)
{
    memset(&s->ener, 0, sizeof s->ener);
    if (s->param.nBond > 0) bond_energy(&s->param, &s->ener, &s->box, s->bond, x, y, z, fx, fy, fz);
    if (s->param.nAngle > 0) angle_energy(&s->param, &s->ener, &s->box, s->angle, x, y, z, fx, fy, fz);
    if (s->param.nUb > 0) ub_energy(&s->param, &s->ener, &s->box, s->ub, x, y, z, fx, fy, fz);
    if (s->param.nDihedral > 0)
        dihedral_energy(&s->param, &s->ener, &s->box, s->dihe, x, y, z, fx, fy, fz);
    if (s->param.nImproper > 0)
        improper_energy(&s->param, &s->ener, &s->box, s->impr, x, y, z, fx, fy, fz);
    out[0] = s->ener.bond;
    out[1] = s->ener.ub;
    out[2] = s->ener.ang;
    out[3] = s->ener.dihe;
    out[4] = s->ener.impr;
    out[5] = s->ener.virbond;
    out[6] = s->ener.virub;
}

/* ewald14_energy (ewald = 1) or nonbond14_energy over the 1-4 list of the
 * last makelist; out = elec, vdw, virelec, virvdw added. */
void ref_nb14(ref_sys *s, int ewald, int elecType, int vdwType, const double *x, const double *y,
              const double *z, double *fx, double *fy, double *fz, const double *q,
              const double *eps14, const double *sig14, double *out)
{
    s->ctrl.elecType = (enum ELEC_TYPE)elecType;
    s->ctrl.vdwType = (enum VDW_TYPE)vdwType;
    init_energy_ptrs(&s->ctrl);
    memset(&s->ener, 0, sizeof s->ener);
    if (ewald)
        ewald14_energy(&s->param, &s->ener, &s->ewald, &s->box, &s->neigh, x, y, z, fx, fy, fz,
                       q, eps14, sig14, s->neighList14);
    else
        nonbond14_energy(&s->param, &s->ener, &s->box, &s->neigh, x, y, z, fx, fy, fz, q, eps14,
                         sig14, s->neighList14);
    out[0] = s->ener.elec;
    out[1] = s->ener.vdw;
    out[2] = s->ener.virelec;
    out[3] = s->ener.virvdw;
}

/* ---- reference input and output (io.c, init.c setup) -------------------- */

#include "io.h"

void setup(CTRL *ctrl, PARAM *param, ATOM atom[], CONSTRAINT **constList, BOND **bond,
           ANGLE **angle, DIHE **dihe, DIHE **impr, BOND **ub, double mass[], double rmass[],
           int frozen[], int nAtConst[]);

typedef struct {
    IO io;
    CTRL ctrl;
    PARAM param;
    BATH bath;
    NEIGH neigh;
    EWALD ewald;
    PBC box;
    ENERGY ener;
    ATOM *atom;
    double *x, *y, *z, *vx, *vy, *vz, *fx, *fy, *fz;
    double *q, *eps, *sig, *eps14, *sig14, *mass, *rmass;
    int *frozen, *nAtConst;
    CONSTRAINT *cons;
    BOND *bond, *ub;
    ANGLE *angle;
    DIHE *dihe, *impr;
} ref_deck;

/* init_system's reading chain (init.c:88-172) in the current directory:
 * read_SIMU, the Coulomb constant, read_CONF (or read_rest under
 * "restart"), read_FORF, setup, get_kinfromtemp, init_box, image_update. */
ref_deck *ref_read_deck(const char *simu)
{
    ensure_out();
    ref_deck *d = (ref_deck *)calloc(1, sizeof(ref_deck));
    strncpy(d->io.simuName, simu, FINAMELEN - 1);
    read_SIMU(&d->io, &d->ctrl, &d->param, &d->bath, &d->neigh, &d->ewald, &d->box);
    if (d->ctrl.mdType == 1)
        d->param.chargeConst = chgcharmm * kcaltoiu;
    else if (d->ctrl.mdType == 2)
        d->param.chargeConst = chgnamd * kcaltoiu;
    else if (d->ctrl.mdType == 3)
        d->param.chargeConst = chgdlpolyiu;
    else
        d->param.chargeConst = mu0 * X2(clight) * X2(elemchg) * NA * 0.1 / (angstr);
    if (d->ctrl.keyRest)
        read_rest(&d->io, &d->param, &d->ener, &d->bath, &d->atom, &d->x, &d->y, &d->z, &d->vx,
                  &d->vy, &d->vz, &d->fx, &d->fy, &d->fz);
    else
        read_CONF(&d->io, &d->param, &d->atom, &d->x, &d->y, &d->z);
    int n = d->param.nAtom;
    double **arr[] = {&d->q, &d->eps, &d->sig, &d->eps14, &d->sig14, &d->mass, &d->rmass};
    for (unsigned k = 0; k < sizeof arr / sizeof arr[0]; k++)
        *arr[k] = (double *)calloc((size_t)n, sizeof(double));
    d->frozen = (int *)calloc((size_t)n, sizeof(int));
    d->nAtConst = (int *)calloc((size_t)n, sizeof(int));
    read_FORF(&d->io, &d->param, d->atom, &d->cons, &d->bond, &d->angle, &d->dihe, &d->impr,
              &d->ub, d->eps, d->sig, d->eps14, d->sig14, d->mass, d->q, d->frozen, d->nAtConst);
    setup(&d->ctrl, &d->param, d->atom, &d->cons, &d->bond, &d->angle, &d->dihe, &d->impr,
          &d->ub, d->mass, d->rmass, d->frozen, d->nAtConst);
    get_kinfromtemp(&d->param, &d->box);
    init_box(&d->box);
    image_update(&d->param, &d->box, d->x, d->y, d->z);
    return d;
}

/* The settings structs, byte for byte. */
void ref_deck_structs(const ref_deck *d, void *ctrl, void *param, void *bath, void *neigh,
                      void *ewald, void *box, void *ener, void *io)
{
    memcpy(ctrl, &d->ctrl, sizeof d->ctrl);
    memcpy(param, &d->param, sizeof d->param);
    memcpy(bath, &d->bath, sizeof d->bath);
    memcpy(neigh, &d->neigh, sizeof d->neigh);
    memcpy(ewald, &d->ewald, sizeof d->ewald);
    memcpy(box, &d->box, sizeof d->box);
    memcpy(ener, &d->ener, sizeof d->ener);
    memcpy(io, &d->io, sizeof d->io);
}

/* Per-atom data: atoms, x y z, q eps sig eps14 sig14 mass rmass (double
 * rows of n), frozen and nAtConst. */
void ref_deck_atoms(const ref_deck *d, void *atom, double *xyz, double *par, int *frozen,
                    int *nAtConst)
{
    size_t n = (size_t)d->param.nAtom;
    memcpy(atom, d->atom, n * sizeof(ATOM));
    memcpy(xyz, d->x, n * sizeof(double));
    memcpy(xyz + n, d->y, n * sizeof(double));
    memcpy(xyz + 2 * n, d->z, n * sizeof(double));
    const double *p[] = {d->q, d->eps, d->sig, d->eps14, d->sig14, d->mass, d->rmass};
    for (size_t k = 0; k < 7; k++) memcpy(par + k * n, p[k], n * sizeof(double));
    memcpy(frozen, d->frozen, n * sizeof(int));
    memcpy(nAtConst, d->nAtConst, n * sizeof(int));
}

/* Term tables after setup (counts in PARAM). */
void ref_deck_terms(const ref_deck *d, void *cons, void *bond, void *ub, void *angle, void *dihe,
                    void *impr)
{
    const PARAM *p = &d->param;
    if (p->nConst) memcpy(cons, d->cons, (size_t)p->nConst * sizeof(CONSTRAINT));
    if (p->nBond) memcpy(bond, d->bond, (size_t)p->nBond * sizeof(BOND));
    if (p->nUb) memcpy(ub, d->ub, (size_t)p->nUb * sizeof(BOND));
    if (p->nAngle) memcpy(angle, d->angle, (size_t)p->nAngle * sizeof(ANGLE));
    if (p->nDihedral) memcpy(dihe, d->dihe, (size_t)p->nDihedral * sizeof(DIHE));
    if (p->nImproper) memcpy(impr, d->impr, (size_t)p->nImproper * sizeof(DIHE));
}

/* The reference writers on the deck's atoms and the given state, into the
 * names of io (paths relative to the current directory): RCON, one PROP
 * record, REST, and a DCD header with two frames. */
void ref_deck_write(ref_deck *d, const void *io, const void *ener, const double *xyz,
                    const double *vxyz, const double *fxyz, double chiT, double chiP)
{
    size_t n = (size_t)d->param.nAtom;
    memcpy(&d->io, io, sizeof d->io);
    memcpy(&d->ener, ener, sizeof d->ener);
    d->bath.chiT = chiT;
    d->bath.chiP = chiP;
    double *x = (double *)xyz, *y = x + n, *z = y + n;
    double *vx = (double *)vxyz, *vy = vx + n, *vz = vy + n;
    double *fx = (double *)fxyz, *fy = fx + n, *fz = fy + n;
    write_CONF(&d->io, &d->param, d->atom, x, y, z);
    write_prop(&d->io, &d->param, &d->ener, &d->box);
    write_rest(&d->io, &d->param, &d->ener, &d->bath, d->atom, x, y, z, vx, vy, vz, fx, fy, fz);
    write_DCD_header(&d->io, &d->ctrl, &d->param, &d->box, d->frozen);
    write_DCD_traj(&d->io, &d->param, &d->box, x, y, z, d->frozen);
    write_DCD_traj(&d->io, &d->param, &d->box, vx, vy, vz, d->frozen);
}

void ref_deck_set_step(ref_deck *d, int step) { d->param.step = step; }
