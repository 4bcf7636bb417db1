/*
 * mdbas_gpu -- MDBas's main program over libmdgpu: reads a reference deck
 * (SIMU, CONF or REST, FORF) in the current directory, runs the MD loop
 * on the GPU and writes the reference's OUTPUT, PROP, DCD, RCON and REST
 * files.  It restates the order of the reference's main program
 * (mdbas.c:36-363 with init_system, init.c:30-246) around the C ABI of
 * include/mdgpu.h.
 *
 *   mdbas_gpu [-i SIMU] [-o OUTPUT] [-p fp64|mixed] [-m dropin|device]
 *             [-g 0|1|2] [-s seed]
 *
 * -m dropin : every step calls the drop-in mirrors in the reference's
 *             order: mdg_vv_integrate(stage 1), mdg_makelist, mdg_energy,
 *             mdg_vv_integrate(stage 2) (mdbas.c:244-301), host arrays
 *             copied in and out on each call.
 * -m device : the state stays on the GPU; mdg_run_steps advances it in
 *             chunks that end where an output is due.  PROP records and
 *             the energy table come from the per-step device records;
 *             DCD frames, RCON and REST take a state snapshot
 *             (mdg_download_state) that a writer thread writes while the
 *             GPU runs the next chunk.
 * -p         : precision of the context (device mode; the drop-ins are
 *             FP64 in either case).
 * -g         : graph mode of mdg_run_steps (default 1).
 * -s         : seed of the Gaussian deviates for the initial velocities
 *             (default: SIMU "seed", else the clock).  The deviates come
 *             from this program's generator (splitmix64 + Box-Muller), not
 *             the reference's dSFMT stream; mdg_init_vel applies the
 *             reference's procedure to them.
 *
 * Exit status: 0, or the reference's error number (or an MDG_E* code)
 * of the first call that fails.
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "mdgpu.h"

#define KCAL 418.4
#define RGAS 0.83144621
#define BAR 6.02214129e-3

static FILE *out;

static void fail(int code, const char *what, mdg_ctx *ctx)
{
    fprintf(stderr, "mdbas_gpu: %s failed with %d%s%s\n", what, code, ctx ? ": " : "",
            ctx ? mdg_last_error(ctx) : "");
    if (out && out != stdout) fprintf(out, "\nMDBas error %d in %s\n", code, what);
    exit(code > 255 ? 1 : code);
}

#define CHECK(call, ctx)                                                                           \
    do {                                                                                           \
        int st_ = (call);                                                                          \
        if (st_) fail(st_, #call, ctx);                                                            \
    } while (0)

/* ---- Gaussian deviates for mdg_init_vel ------------------------------- */

static uint64_t splitmix64(uint64_t *s)
{
    uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static double uniform01(uint64_t *s)   /* (0, 1] */
{
    return ((splitmix64(s) >> 11) + 1) * (1.0 / 9007199254740992.0);
}

static void gaussians(double *g, long n, uint64_t seed)
{
    uint64_t s = seed;
    for (long i = 0; i < n; i += 2) {
        const double r = sqrt(-2.0 * log(uniform01(&s))), t = 6.283185307179586 * uniform01(&s);
        g[i] = r * cos(t);
        if (i + 1 < n) g[i + 1] = r * sin(t);
    }
}

/* ---- the energy table of OUTPUT (mdbas.c:186-225, 313-341) ------------ */

static const char *DASHES =
    "--------------------------------------------------------------------------------";

static void print_energies(const MDG_PARAM *p, const MDG_ENERGY *e, const MDG_PBC *b, int *nPrint)
{
    const double temp = 2. * e->kin / ((double)p->nDegFree * RGAS);
    const double press = b->type > 0 ? (2. * e->kin - e->virtot) / (3. * b->vol * BAR) : 0.;
    if (*nPrint % 5 == 0) {
        fprintf(out, "\n%s\n\n", DASHES);
        fprintf(out,
                "%13s\t%13s\t%13s\t%13s\t%13s\n%13s\t%13s\t%13s\t%13s\t%13s\n%13s\t%13s\t%13s\t%13s\t"
                "%13s\n",
                "Step", "Time", "Temp", "Press", "Vol", "Etot", "Ekin", "Epot", "Ecoul", "Evdw",
                "Ebond", "Eangle", "Eub", "Edihe", "Eimpr");
        fprintf(out, "\n%s\n", DASHES);
        *nPrint = 0;
    }
    fprintf(out,
            "%13d\t%#13.5le\t%#13.5le\t%#13.5le\t%#13.5le\n%#13.5le\t%#13.5le\t%#13.5le\t%#13.5le\t"
            "%#13.5le\n%#13.5le\t%#13.5le\t%#13.5le\t%#13.5le\t%#13.5le\n",
            p->step, p->step * p->timeStep, temp, press, b->vol, e->tot / KCAL, e->kin / KCAL,
            e->pot / KCAL, e->elec / KCAL, e->vdw / KCAL, e->bond / KCAL, e->ang / KCAL,
            e->ub / KCAL, e->dihe / KCAL, e->impr / KCAL);
    fprintf(out, "%s\n", DASHES);
    (*nPrint)++;
}

/* ---- the run ----------------------------------------------------------- */

typedef struct {
    MDG_IO io;
    MDG_CTRL ctrl;
    MDG_PARAM param;
    MDG_ENERGY ener;
    MDG_PBC box;
    MDG_BATH bath;
    MDG_EWALD ewald;
    MDG_NEIGH neigh;
    MDG_ATOM *atom;
    MDG_BOND *bond, *ub;
    MDG_ANGLE *angle;
    MDG_DIHE *dihe, *impr;
    MDG_CONSTRAINT *cons;
    double *x, *y, *z, *vx, *vy, *vz, *fx, *fy, *fz;
    double *q, *mass, *rmass, *eps, *sig, *eps14, *sig14;
    int *frozen, *nAtConst;
    int *neighList, *neighPair, *neighOrder, *neighList14, **exclList, *exclPair;
    mdg_ctx *ctx;
    int nPrint;
} Run;

/* ---- asynchronous writer for device snapshots --------------------------- */

typedef struct Job {
    int step, traj, rest;
    MDG_ENERGY ener;
    MDG_BATH bath;
    double *s;           /* 9 N doubles: x y z vx vy vz fx fy fz */
    struct Job *next;
} Job;

static struct {
    pthread_mutex_t mu;
    pthread_cond_t cv;
    Job *head, *tail;
    int done, error;
    Run *run;
} W = {PTHREAD_MUTEX_INITIALIZER, PTHREAD_COND_INITIALIZER, NULL, NULL, 0, 0, NULL};

static void *writer_main(void *arg)
{
    (void)arg;
    Run *R = W.run;
    const long n = R->param.nAtom;
    for (;;) {
        pthread_mutex_lock(&W.mu);
        while (!W.head && !W.done) pthread_cond_wait(&W.cv, &W.mu);
        Job *j = W.head;
        if (j) {
            W.head = j->next;
            if (!W.head) W.tail = NULL;
        }
        pthread_mutex_unlock(&W.mu);
        if (!j) return NULL;
        MDG_PARAM p = R->param;
        p.step = j->step;
        double *s = j->s;
        int st = 0;
        if (j->traj) st = st ? st : mdg_write_dcd_traj(&R->io, &p, &R->box, s, s + n, s + 2 * n, R->frozen);
        if (j->rest) {
            st = st ? st : mdg_write_conf(&R->io, &p, R->atom, s, s + n, s + 2 * n);
            st = st ? st : mdg_write_rest(&R->io, &p, &j->ener, &j->bath, R->atom, s, s + n,
                                          s + 2 * n, s + 3 * n, s + 4 * n, s + 5 * n, s + 6 * n,
                                          s + 7 * n, s + 8 * n);
        }
        if (st) W.error = st;
        free(s);
        free(j);
    }
}

static void enqueue(Run *R, int step, int traj, int rest)
{
    const long n = R->param.nAtom;
    Job *j = (Job *)calloc(1, sizeof(Job));
    j->step = step;
    j->traj = traj;
    j->rest = rest;
    j->ener = R->ener;
    j->bath = R->bath;
    j->s = (double *)malloc(9 * n * sizeof(double));
    double *s = j->s;
    CHECK(mdg_download_state(R->ctx, s, s + n, s + 2 * n, s + 3 * n, s + 4 * n, s + 5 * n,
                             s + 6 * n, s + 7 * n, s + 8 * n),
          R->ctx);
    pthread_mutex_lock(&W.mu);
    if (W.tail) W.tail->next = j;
    else W.head = j;
    W.tail = j;
    pthread_cond_signal(&W.cv);
    pthread_mutex_unlock(&W.mu);
}

/* ENERGY of one device record (energy.c:193-196, mdbas.c:303-305). */
static void energy_of(MDG_ENERGY *e, const MDG_STEP *r)
{
    e->pot = r->pot;
    e->kin = r->kin;
    e->tot = r->kin + r->pot;
    e->elec = r->elec;
    e->vdw = r->vdw;
    e->bond = r->ebond;
    e->ub = r->eub;
    e->ang = r->eang;
    e->dihe = r->edihe;
    e->impr = r->eimpr;
    e->virelec = r->virelec;
    e->virvdw = r->virvdw;
    e->virbond = r->virbond;
    e->virub = r->virub;
    e->virshake = r->virshake;
    e->virpot = e->virbond + e->virub + e->virelec + e->virvdw;
    e->virtot = e->virbond + e->virub + e->virelec + e->virvdw + e->virshake;
}

static int due(int step, int every) { return every > 0 && step % every == 0; }

/* Step-0 output (mdbas.c:178-225). */
static void output_step0(Run *R)
{
    MDG_ENERGY *e = &R->ener;
    e->tot = e->kin + e->pot;
    e->virtot = e->virbond + e->virub + e->virelec + e->virvdw + e->virshake;
    if (R->ctrl.keyProp && due(R->param.step, R->ctrl.printProp))
        CHECK(mdg_write_prop(&R->io, &R->param, e, &R->box), NULL);
    if (due(R->param.step, R->ctrl.printOut)) print_energies(&R->param, e, &R->box, &R->nPrint);
}

static void run_dropin(Run *R)
{
    MDG_PARAM *p = &R->param;
    R->ener.kin = mdg_kinetic(p, R->vx, R->vy, R->vz, R->mass);
    CHECK(mdg_energy(R->ctx, &R->ctrl, p, &R->ener, &R->ewald, &R->box, &R->neigh, R->bond,
                     R->ub, R->angle, R->dihe, R->impr, R->x, R->y, R->z, R->vx, R->vy, R->vz,
                     R->fx, R->fy, R->fz, R->q, R->eps, R->sig, R->eps14, R->sig14, R->frozen,
                     R->neighList, R->neighPair, R->neighOrder, R->neighList14, R->exclList,
                     R->exclPair),
          R->ctx);
    output_step0(R);
    if (!R->ctrl.keyMd) return;
    for (++p->step; p->step <= p->nSteps; p->step++) {
        R->ener.consv = 0.;
        CHECK(mdg_vv_integrate(R->ctx, &R->ctrl, p, &R->ener, &R->box, &R->bath, R->cons, R->x,
                               R->y, R->z, R->vx, R->vy, R->vz, R->fx, R->fy, R->fz, R->mass,
                               R->rmass, R->nAtConst, 1),
              R->ctx);
        CHECK(mdg_makelist(R->ctx, &R->ctrl, p, &R->box, &R->neigh, R->cons, R->bond, R->angle,
                           R->dihe, R->impr, R->x, R->y, R->z, R->frozen, &R->neighList,
                           &R->neighPair, &R->neighOrder, &R->neighList14, &R->exclList,
                           &R->exclPair),
              R->ctx);
        CHECK(mdg_energy(R->ctx, &R->ctrl, p, &R->ener, &R->ewald, &R->box, &R->neigh, R->bond,
                         R->ub, R->angle, R->dihe, R->impr, R->x, R->y, R->z, R->vx, R->vy,
                         R->vz, R->fx, R->fy, R->fz, R->q, R->eps, R->sig, R->eps14, R->sig14,
                         R->frozen, R->neighList, R->neighPair, R->neighOrder, R->neighList14,
                         R->exclList, R->exclPair),
              R->ctx);
        CHECK(mdg_vv_integrate(R->ctx, &R->ctrl, p, &R->ener, &R->box, &R->bath, R->cons, R->x,
                               R->y, R->z, R->vx, R->vy, R->vz, R->fx, R->fy, R->fz, R->mass,
                               R->rmass, R->nAtConst, 2),
              R->ctx);
        R->ener.tot = R->ener.kin + R->ener.pot;
        R->ener.virtot = R->ener.virbond + R->ener.virub + R->ener.virelec + R->ener.virvdw +
                         R->ener.virshake;
        if (R->ctrl.keyProp && due(p->step, R->ctrl.printProp))
            CHECK(mdg_write_prop(&R->io, p, &R->ener, &R->box), NULL);
        if (due(p->step, R->ctrl.printOut)) print_energies(p, &R->ener, &R->box, &R->nPrint);
        if (R->ctrl.keyTraj && due(p->step, R->ctrl.printTraj))
            CHECK(mdg_write_dcd_traj(&R->io, p, &R->box, R->x, R->y, R->z, R->frozen), NULL);
        if (due(p->step, R->ctrl.printRest)) {
            CHECK(mdg_write_conf(&R->io, p, R->atom, R->x, R->y, R->z), NULL);
            CHECK(mdg_write_rest(&R->io, p, &R->ener, &R->bath, R->atom, R->x, R->y, R->z, R->vx,
                                 R->vy, R->vz, R->fx, R->fy, R->fz),
                  NULL);
        }
    }
    p->step--;   /* the loop leaves step at nSteps + 1 */
}

/* Steps from s+1 to the next step at which an output is due (at most
 * nSteps). */
static int next_stop(const Run *R, int s)
{
    const int every[4] = {R->ctrl.keyProp ? R->ctrl.printProp : 0, R->ctrl.printOut,
                          R->ctrl.keyTraj ? R->ctrl.printTraj : 0, R->ctrl.printRest};
    int stop = R->param.nSteps;
    for (int k = 0; k < 4; k++)
        if (every[k] > 0) {
            const int t = (s / every[k] + 1) * every[k];
            if (t < stop) stop = t;
        }
    return stop;
}

static void run_device(Run *R, int graph)
{
    MDG_PARAM *p = &R->param;
    CHECK(mdg_upload_state(R->ctx, R->x, R->y, R->z, R->vx, R->vy, R->vz), R->ctx);
    MDG_STEP r0;
    CHECK(mdg_initial_energy(R->ctx, &r0), R->ctx);
    /* step 0 as energy() leaves it: kinetic energy of the host velocities,
     * virshake and the conserved-quantity terms as they were (zero, or the
     * REST file's) */
    const double kin0 = mdg_kinetic(p, R->vx, R->vy, R->vz, R->mass);
    const double virshake0 = R->ener.virshake;
    energy_of(&R->ener, &r0);
    R->ener.kin = kin0;
    R->ener.virshake = virshake0;
    output_step0(R);
    if (!R->ctrl.keyMd) return;
    pthread_t th;
    W.run = R;
    pthread_create(&th, NULL, writer_main, NULL);
    MDG_STEP *rec = NULL;
    int cap = 0;
    while (p->step < p->nSteps) {
        const int stop = next_stop(R, p->step), k = stop - p->step;
        if (k > cap) {
            free(rec);
            cap = k;
            rec = (MDG_STEP *)malloc((size_t)cap * sizeof(MDG_STEP));
        }
        CHECK(mdg_run_steps(R->ctx, k, graph, rec), R->ctx);
        for (int s = 0; s < k; s++) {
            p->step++;
            energy_of(&R->ener, &rec[s]);
            R->ener.consv = 0.;
            if (R->ctrl.ens) {   /* the device thermostat's state after this step */
                R->ener.conint = rec[s].conint;
                R->ener.consv = rec[s].consv;
                R->bath.chiT = rec[s].chiT;
            }
            if (R->ctrl.keyProp && due(p->step, R->ctrl.printProp))
                CHECK(mdg_write_prop(&R->io, p, &R->ener, &R->box), NULL);
            if (due(p->step, R->ctrl.printOut)) print_energies(p, &R->ener, &R->box, &R->nPrint);
        }
        const int traj = R->ctrl.keyTraj && due(p->step, R->ctrl.printTraj);
        const int rest = due(p->step, R->ctrl.printRest);
        if (traj || rest) enqueue(R, p->step, traj, rest);
    }
    pthread_mutex_lock(&W.mu);
    W.done = 1;
    pthread_cond_broadcast(&W.cv);
    pthread_mutex_unlock(&W.mu);
    pthread_join(th, NULL);
    free(rec);
    if (W.error) fail(W.error, "snapshot writer", NULL);
    CHECK(mdg_download_state(R->ctx, R->x, R->y, R->z, R->vx, R->vy, R->vz, R->fx, R->fy, R->fz),
          R->ctx);
}

int main(int argc, char **argv)
{
    static Run R;
    char outName[MDG_FINAMELEN] = "OUTPUT";
    int precision = MDG_FP64, device = 0, graph = 1, haveSeed = 0;
    uint64_t seed = 0;
    strcpy(R.io.simuName, "SIMU");
    for (int i = 1; i < argc; i++) {
        const char *a = argv[i];
        const char *v = i + 1 < argc ? argv[i + 1] : NULL;
        if (!strcmp(a, "--help")) {
            printf("%s [-i SIMU] [-o OUTPUT] [-p fp64|mixed] [-m dropin|device] [-g 0|1|2] "
                   "[-s seed]\n",
                   argv[0]);
            return 0;
        }
        if (!v) fail(MDG_EARG, "command line", NULL);
        if (!strcmp(a, "-i")) snprintf(R.io.simuName, MDG_FINAMELEN, "%s", v);
        else if (!strcmp(a, "-o")) snprintf(outName, MDG_FINAMELEN, "%s", v);
        else if (!strcmp(a, "-p")) precision = !strcmp(v, "mixed") ? MDG_MIXED : MDG_FP64;
        else if (!strcmp(a, "-m")) device = !strcmp(v, "device");
        else if (!strcmp(a, "-g")) graph = atoi(v);
        else if (!strcmp(a, "-s")) {
            seed = strtoull(v, NULL, 10);
            haveSeed = 1;
        } else fail(MDG_EARG, "command line", NULL);
        i++;
    }
    out = fopen(outName, "w");
    if (!out) out = stdout;

    /* init_system (init.c:30-246) */
    MDG_PARAM *p = &R.param;
    CHECK(mdg_read_simu(&R.io, &R.ctrl, p, &R.bath, &R.neigh, &R.ewald, &R.box), NULL);
    fprintf(out, "%s file read\n", R.io.simuName);
    p->chargeConst = mdg_charge_const(R.ctrl.mdType);
    if (R.ctrl.keyRest) {
        CHECK(mdg_read_rest(&R.io, p, &R.ener, &R.bath, &R.atom, &R.x, &R.y, &R.z, &R.vx, &R.vy,
                            &R.vz, &R.fx, &R.fy, &R.fz),
              NULL);
        fprintf(out, "%s file read\n", R.io.restName);
    } else {
        CHECK(mdg_read_conf(&R.io, p, &R.atom, &R.x, &R.y, &R.z), NULL);
        fprintf(out, "%s file read\n", R.io.confName);
        double **v[] = {&R.vx, &R.vy, &R.vz, &R.fx, &R.fy, &R.fz};
        for (int k = 0; k < 6; k++) *v[k] = (double *)calloc(p->nAtom, sizeof(double));
    }
    const int n = p->nAtom;
    double **pa[] = {&R.q, &R.mass, &R.rmass, &R.eps, &R.sig, &R.eps14, &R.sig14};
    for (int k = 0; k < 7; k++) *pa[k] = (double *)calloc(n, sizeof(double));
    R.frozen = (int *)calloc(n, sizeof(int));
    R.nAtConst = (int *)calloc(n, sizeof(int));
    CHECK(mdg_read_forf(&R.io, p, R.atom, &R.cons, &R.bond, &R.angle, &R.dihe, &R.impr, &R.ub,
                        R.eps, R.sig, R.eps14, R.sig14, R.mass, R.q, R.frozen, R.nAtConst),
          NULL);
    fprintf(out, "%s file read\n", R.io.forfName);
    CHECK(mdg_setup(&R.ctrl, p, R.atom, &R.cons, &R.bond, &R.angle, &R.dihe, &R.impr, &R.ub,
                    R.mass, R.rmass, R.frozen, R.nAtConst),
          NULL);
    fprintf(out, "Setup done\n");
    mdg_get_kinfromtemp(p, &R.box);
    mdg_init_box(&R.box);
    mdg_image_update(p, &R.box, R.x, R.y, R.z);
    if (R.ctrl.keyEwald == 2) {
        CHECK(mdg_spme_mesh(&R.ctrl, p, &R.ewald, &R.box), NULL);
    } else if (R.ctrl.keyEwald == 1) {
        fail(MDG_EUNSUPPORTED, "classic Ewald sum (elec ewald): use elec spme", NULL);
    }
    if (R.ctrl.integrator != 1) fail(MDG_EUNSUPPORTED, "leap-frog integrator", NULL);
    if (device && R.ctrl.ens != 0 && R.ctrl.ens != 1 && R.ctrl.ens != 3)
        fail(MDG_EUNSUPPORTED, "device-resident steps run NVE, NVT_B or NVT_H: run the barostats "
             "with -m dropin", NULL);

    int st = 0;
    R.ctx = mdg_create(&R.ctrl, p, &R.box, &R.ewald, &R.neigh, device ? precision : MDG_FP64, &st);
    if (!R.ctx) fail(st, "mdg_create", NULL);
    CHECK(mdg_set_topology(R.ctx, R.q, R.eps, R.sig, R.mass, R.rmass, R.frozen, R.nAtConst,
                           R.cons, p->nConst),
          R.ctx);
    /* makelist's first call (init.c:172), which also builds the exclusions */
    CHECK(mdg_makelist(R.ctx, &R.ctrl, p, &R.box, &R.neigh, R.cons, R.bond, R.angle, R.dihe,
                       R.impr, R.x, R.y, R.z, R.frozen, &R.neighList, &R.neighPair, &R.neighOrder,
                       &R.neighList14, &R.exclList, &R.exclPair),
          R.ctx);
    if (device && (p->nBond + p->nUb + p->nAngle + p->nDihedral + p->nImproper > 0 ||
                   R.ctrl.keyNb14))
        CHECK(mdg_set_bonded(R.ctx, &R.ctrl, p, R.bond, R.ub, R.angle, R.dihe, R.impr, R.eps14,
                             R.sig14),
              R.ctx);
    if (R.ctrl.keyEwald == 2)
        fprintf(out, "\nSmooth Particle-Mesh Ewald: order %d, alpha %lf, mesh %d %d %d\n\n",
                R.ewald.nbsp, R.ewald.alpha, R.ewald.m1max, R.ewald.m2max, R.ewald.m3max);
    if (!R.ctrl.keyRest) {
        remove(R.io.propName);
        if (!haveSeed) seed = R.ctrl.keyRand ? (uint64_t)R.ctrl.seed : (uint64_t)time(NULL);
        double *g = (double *)malloc(3 * (size_t)n * sizeof(double));
        gaussians(g, 3L * n, seed);
        CHECK(mdg_init_vel(p, &R.box, R.cons, R.x, R.y, R.z, R.vx, R.vy, R.vz, R.mass, R.rmass,
                           R.frozen, R.nAtConst, g),
              NULL);
        free(g);
        R.bath.chiT = R.bath.chiP = 0.;
        R.ener.conint = R.ener.consv = R.ener.virshake = 0.;
    }
    R.box.stress1 = R.box.stress2 = R.box.stress3 = R.box.stress4 = R.box.stress5 = 0.;
    R.box.stress6 = R.box.stress7 = R.box.stress8 = R.box.stress9 = 0.;
    if (R.ctrl.keyTraj) CHECK(mdg_write_dcd_header(&R.io, &R.ctrl, p, &R.box, R.frozen), NULL);
    fprintf(out, "GPU path: %s, %s\n\n", device ? "device-resident steps" : "drop-in calls",
            device && precision == MDG_MIXED ? "mixed precision" : "FP64");

    if (device && R.ctrl.ens) CHECK(mdg_set_bath(R.ctx, &R.ctrl, p, &R.bath, &R.ener), R.ctx);
    if (device) run_device(&R, graph);
    else run_dropin(&R);

    /* mdbas.c:351-360 */
    CHECK(mdg_write_conf(&R.io, p, R.atom, R.x, R.y, R.z), NULL);
    fprintf(out, "\n%s file written\n", R.io.rconName);
    CHECK(mdg_write_rest(&R.io, p, &R.ener, &R.bath, R.atom, R.x, R.y, R.z, R.vx, R.vy, R.vz,
                         R.fx, R.fy, R.fz),
          NULL);
    fprintf(out, "\n%s file written\n", R.io.restName);
    fprintf(out, "\nNormal Termination of MDBas.\n");
    mdg_destroy(R.ctx);
    if (out != stdout) fclose(out);
    return 0;
}
