// mdg_run.hip -- device-resident MD: the step order of mdbas.c:244-301
// (VV stage 1, list rebuild every `update` steps, energy, VV stage 2) with
// every array on the device and no host round trip per step.  A step is
// captured once per variant (plain / with list rebuild) in a hipGraph and
// replayed.
#include "mdg_internal.h"
#include "mdg_shake.h"
#include "mdg_side.h"

#include <cstring>
#include <initializer_list>

#define TWOPI_R 6.283185307179586


// Per atom: F = slot pair force + (reciprocal force - mean reciprocal
// force) + excluded-pair corrections.  The mean is over the reciprocal
// part only (spme.c:648-670 runs on arrays energy() has just zeroed).
// The correction of each excluded pair (energy.c:567-621) is evaluated
// from both of its atoms with the reference's orientation (delta = x_hi -
// x_lo), so every atom sums its own terms without atomics; energy and
// virial count each pair once.  Frozen atoms get zero force and velocity
// (utils.c:131-153).
template <bool MIXED>
__global__ void __launch_bounds__(256) k_finalize(FinArgs A)
{
    __shared__ double mean[3], red[4][3];
    fin_mean(A, mean, red);
    // both SHAKE masks: stage 1 is done with its own, the next stage 2
    // has not started
    if (blockIdx.x == 0 && threadIdx.x < 2 * SHK_MASKW) A.mask[threadIdx.x] = ~0ull;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double acc[2] = {0, 0};
    if (i < A.n) {
        double gx, gy, gz;
        (void)fin_atom<MIXED>(A, i, mean, acc, gx, gy, gz);
    }
    mdg_block_partials<2>(acc, A.part);
}

__global__ void k_kinetic(int n, const double *__restrict__ mass, const double *__restrict__ vx,
                          const double *__restrict__ vy, const double *__restrict__ vz,
                          double *__restrict__ part)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    double acc[7] = {0, 0, 0, 0, 0, 0, 0};
    if (i < n) acc[0] = mass[i] * (vx[i] * vx[i] + vy[i] * vy[i] + vz[i] * vz[i]);
    mdg_block_partials<7>(acc, part);
}

// The step record as a kernel of its own (record_block, mdg_side.h).
__global__ void __launch_bounds__(SIDE_BLOCK) k_record(RecArgs R, int *ctrl, MDG_STEP *rec,
                                                       int rebuilt)
{
    record_block(R, ctrl, rec, rebuilt);
}

static int ensure_slots(mdg_ctx *ctx)
{
    if (ctx->sx) return MDG_OK;
    const int ns = ctx->cl.nclu * MDG_CL + 64;
    HIPCHK(hipMalloc(&ctx->sx, ns * sizeof(double)));
    HIPCHK(hipMalloc(&ctx->sy, ns * sizeof(double)));
    HIPCHK(hipMalloc(&ctx->sz, ns * sizeof(double)));
    HIPCHK(hipMalloc(&ctx->sp32, ns * sizeof(float4)));
    HIPCHK(hipMalloc(&ctx->sq, ns * sizeof(double)));
    HIPCHK(hipMalloc(&ctx->seps, ns * sizeof(double)));
    HIPCHK(hipMalloc(&ctx->ssig, ns * sizeof(double)));
    HIPCHK(hipMalloc(&ctx->fdx, ns * sizeof(double)));
    HIPCHK(hipMalloc(&ctx->fdy, ns * sizeof(double)));
    HIPCHK(hipMalloc(&ctx->fdz, ns * sizeof(double)));
    // x, y, z fixed-point accumulators in one allocation (the mixed pair
    // kernel addresses a component as fix + c * ns)
    HIPCHK(hipMalloc(&ctx->fix, 3 * (size_t)ns * sizeof(long long)));
    ctx->fiy = ctx->fix + ns;
    ctx->fiz = ctx->fix + 2 * (size_t)ns;
    if (ctx->precision == MDG_MIXED) {
        ctx->nDispPart = nblk(ctx->cl.nclu * MDG_CL, 256);
        // room for one partial per stage-1 SHAKE block as well (at most one
        // molecule per wavefront, four wavefronts per block)
        const int cap = std::max(ctx->nDispPart, nblk(std::max(ctx->nGrp, 1), 4) + 1);
        HIPCHK(hipMalloc(&ctx->dispPart, cap * sizeof(double)));
        if (!ctx->bt) {
            HIPCHK(hipMalloc(&ctx->relPart, cap * sizeof(double)));
            HIPCHK(hipMalloc(&ctx->spr, ns * sizeof(float4)));
        }
    }
    return MDG_OK;
}

// Captured step graphs hold kernel arguments by value (record capacity,
// self and net-charge energies, ...): drop them when those change.
void mdg_invalidate_graphs(mdg_ctx *ctx)
{
    for (int k = 0; k < 2; k++)
        if (ctx->graphExec[k]) {
            (void)hipGraphExecDestroy(ctx->graphExec[k]);
            ctx->graphExec[k] = nullptr;
        }
    if (ctx->graphBlock) {
        (void)hipGraphExecDestroy(ctx->graphBlock);
        ctx->graphBlock = nullptr;
    }
}

static int ensure_records(mdg_ctx *ctx, int n)
{
    if (ctx->bt) {
        MdgBatch &T = *ctx->bt;
        if (T.recCap >= n) return MDG_OK;
        if (T.rec) HIPCHK(hipFree(T.rec));
        T.recCap = std::max(n, 64);
        HIPCHK(hipMalloc(&T.rec, (size_t)T.recCap * T.R * sizeof(MDG_STEP)));
        mdg_invalidate_graphs(ctx);
        return MDG_OK;
    }
    if (ctx->recCap >= n) return MDG_OK;
    if (ctx->recDev) HIPCHK(hipFree(ctx->recDev));
    ctx->recCap = std::max(n, 64);
    HIPCHK(hipMalloc(&ctx->recDev, ctx->recCap * sizeof(MDG_STEP)));
    mdg_invalidate_graphs(ctx);
    return MDG_OK;
}

// Forces of the current positions on the device (energy() of
// energy.c:103-205), optionally rebuilding the cluster list first.
static int b_energy(mdg_ctx *ctx, bool rebuild);

int mdg_device_energy(mdg_ctx *ctx, bool rebuild, bool fin)
{
    if (ctx->prunePending) {
        // the previous step's list prune reads the list and the slot positions
        HIPCHK(hipStreamWaitEvent(ctx->stream, ctx->evPruneDone, 0));
        ctx->prunePending = false;
    }
    if (ctx->bt) return b_energy(ctx, rebuild);
    const int n = ctx->n;
    const bool ew = ctx->P.ewald != 0;
    if (ctx->P.ewald == 1) {
        ctx->err = "classic k-space Ewald (keyEwald 1) is not on this path: use SPME";
        return MDG_EUNSUPPORTED;
    }
    // SPME on the second stream, beside the list rebuild and the pair kernel;
    // the branch enqueued first starts at the fork, the other one a few
    // microseconds later: the pair branch unless the tuner found the SPME
    // branch the longer one on steps without a rebuild
    if (ew) RETCHK(mdg_spme_fork(ctx));
    else ctx->spmeForked = false;
    // (a rebuild step always starts the SPME branch first: the second
    // branch starts only after the first one's nodes, and the rebuild has
    // some thirty of them)
    const bool spmeFirst = ctx->tune.spmeFirst || rebuild;
    if (ew && spmeFirst) RETCHK(mdg_spme_on(ctx, mdg_spme_device));
    if (rebuild) {
        if (ctx->profiling) HIPCHK(hipEventRecord(ctx->evCur[6], ctx->stream));
        RETCHK(mdg_build_cluster_list(ctx));
        RETCHK(ensure_slots(ctx));
        RETCHK(mdg_slot_params(ctx));
        if (ctx->profiling) HIPCHK(hipEventRecord(ctx->evCur[7], ctx->stream));
    }
    RETCHK(mdg_pair_cluster(ctx));
    // bonded terms and 1-4 pairs (energy.c:157-189) into their int64
    // per-atom accumulators, added by the force assembly
    if (ctx->bd.on) RETCHK(mdg_bonded_forces(ctx, nullptr, nullptr, nullptr, ctx->bd.fix, false));
    if (ew && !spmeFirst) RETCHK(mdg_spme_on(ctx, mdg_spme_device));
    if (ctx->spmeForked) HIPCHK(hipStreamWaitEvent(ctx->stream, ctx->evJoin, 0));
    if (!fin) return MDG_OK;
    const bool mixed = ctx->precision == MDG_MIXED;
    const FinArgs F = mdg_fin_args(ctx);
    if (mixed)
        hipLaunchKernelGGL(k_finalize<true>, dim3(nblk(n, 256)), dim3(256), 0, ctx->stream, F);
    else
        hipLaunchKernelGGL(k_finalize<false>, dim3(nblk(n, 256)), dim3(256), 0, ctx->stream, F);
    LAUNCHCHK();
    return MDG_OK;
}

FinArgs mdg_fin_args(mdg_ctx *ctx)
{
    const int n = ctx->n;
    FinArgs F;
    F.P = ctx->P;
    F.B = ctx->B;
    F.n = n;
    F.hasFrozen = ctx->hasFrozen ? 1 : 0;
    F.nbGather = mdg_gather_blocks(ctx);
    F.atomSlot = ctx->cl.atomSlot;
    F.fix = ctx->fix;
    F.fiy = ctx->fiy;
    F.fiz = ctx->fiz;
    F.fdx = ctx->fdx;
    F.fdy = ctx->fdy;
    F.fdz = ctx->fdz;
    F.frx = ctx->frx;
    F.fry = ctx->fry;
    F.frz = ctx->frz;
    F.frx64 = ctx->frx64;
    F.fry64 = ctx->fry64;
    F.frz64 = ctx->frz64;
    F.gpart = ctx->epart + PART_GATHER;
    F.exOff = ctx->exSymOff;
    F.exLst = ctx->exSymLst;
    F.x = ctx->x;
    F.y = ctx->y;
    F.z = ctx->z;
    F.q = ctx->q;
    F.frozen = ctx->frozen;
    F.fx = ctx->fx;
    F.fy = ctx->fy;
    F.fz = ctx->fz;
    F.vx = ctx->vx;
    F.vy = ctx->vy;
    F.vz = ctx->vz;
    F.part = ctx->epart + PART_EXCL;
    F.mask = ctx->shkMask;
    F.bfix = ctx->bd.on ? ctx->bd.fix : nullptr;
    return F;
}

bool mdg_fin_fused(const mdg_ctx *ctx) { return !ctx->bt && ctx->shakeW3; }

int mdg_excl_rows(const mdg_ctx *ctx)
{
    return mdg_fin_fused(ctx) ? mdg_vv_blocks(ctx) : nblk(ctx->n, 256);
}

RecArgs mdg_rec_args(mdg_ctx *ctx, int nbKin, int nGrpVir, int nbExcl)
{
    RecArgs R;
    R.fact = (TWOPI_R / ctx->B.vol) * ctx->P.chargeConst;
    const bool ew = ctx->P.ewald != 0;   // else no reciprocal, self or net-charge terms
    R.eSelf = ew ? ctx->eSelf : 0.;
    R.eNet = ew ? ctx->eNet : 0.;
    R.recCap = ctx->recCap;
    R.nbPair = mdg_pair_blocks(ctx);
    R.nbExcl = nbExcl;
    R.nbInfl = ew ? mdg_infl_blocks((long)ctx->m1 * ctx->m2 * ctx->mh) : 0;
    R.nbKin = nbKin;
    R.nGrpVir = nGrpVir;
    R.pPair = ctx->epart + PART_PAIR;
    R.pExcl = ctx->epart + PART_EXCL;
    R.pInfl = ctx->epart + PART_INFL;
    R.pKin = ctx->epart + PART_KIN;
    R.grpVir = ctx->grpVir;
    R.nbBond = ctx->bd.on ? mdg_bonded_blocks(ctx) : 0;
    R.pBond = ctx->epart + PART_BOND;
    R.bath = ctx->ens ? ctx->bathDev : nullptr;
    return R;
}

static int b_record(mdg_ctx *ctx, bool rebuild);

// One MD step (mdbas.c:244-301).
int mdg_device_step(mdg_ctx *ctx, bool rebuild)
{
    if (ctx->profiling) HIPCHK(hipEventRecord(ctx->evCur[8], ctx->stream));
    // without a rebuild the slots keep their atoms: stage 1 writes them
    ctx->slotsInStage1 = !rebuild && ctx->precision == MDG_MIXED && !ctx->bt && ctx->shakeW3 &&
                         ctx->sx && ctx->dispPart && ctx->cl.atomSlot;
    ctx->slotsDone = false;
    if (ctx->ens) RETCHK(mdg_thermo_stage(ctx, 1));
    const int st1 = mdg_vv_stage(ctx, 1, ctx->fx, ctx->fy, ctx->fz, nullptr, nullptr, false);
    ctx->slotsInStage1 = false;
    RETCHK(st1);
    if (ctx->profiling) HIPCHK(hipEventRecord(ctx->evCur[9], ctx->stream));
    const bool fused = mdg_fin_fused(ctx);
    RETCHK(mdg_device_energy(ctx, rebuild, !fused));
    if (ctx->profiling) HIPCHK(hipEventRecord(ctx->evCur[10], ctx->stream));
    if (fused) {
        const FinArgs F = mdg_fin_args(ctx);
        RETCHK(mdg_vv_stage(ctx, 2, ctx->fx, ctx->fy, ctx->fz, nullptr, nullptr, false, &F));
    } else {
        RETCHK(mdg_vv_stage(ctx, 2, ctx->fx, ctx->fy, ctx->fz, nullptr, nullptr, false));
    }
    if (ctx->ens) RETCHK(mdg_thermo_stage(ctx, 2));
    // In a step block of one box with three-site SHAKE, the record and the
    // list prune go with the next step's first SHAKE kernel (its spare
    // blocks); the block's last step runs them after the loop.
    if (ctx->recOverlap && !ctx->profiling && ctx->shakeW3) {
        ctx->sidePending = true;
        ctx->sideRebuilt = rebuild;
        if (ctx->profiling) HIPCHK(hipEventRecord(ctx->evCur[11], ctx->stream));
        return MDG_OK;
    }
    // Else, in a step block (batch) the record goes to stream2, beside the
    // next step's first SHAKE kernel; the next step's second one waits for it.
    const bool fork = ctx->recOverlap && !ctx->profiling;
    hipStream_t main = ctx->stream;
    if (fork) {
        HIPCHK(hipEventRecord(ctx->evRecFork, main));
        HIPCHK(hipStreamWaitEvent(ctx->stream2, ctx->evRecFork, 0));
        ctx->stream = ctx->stream2;
    }
    int st = MDG_OK;
    if (ctx->bt)
        st = b_record(ctx, rebuild);
    else
        hipLaunchKernelGGL(k_record, dim3(1), dim3(SIDE_BLOCK), 0, ctx->stream,
                           mdg_rec_args(ctx, mdg_vv_blocks(ctx), mdg_vir_rows(ctx), mdg_excl_rows(ctx)),
                           ctx->ctrl, ctx->recDev, rebuild ? 1 : 0);
    if (fork && st == MDG_OK) {
        // a failure here must not leave ctx->stream on stream2: record the
        // status and fall through to the restore below
        if (hipEventRecord(ctx->evRecDone, ctx->stream2) == hipSuccess) {
            ctx->recPending = true;
        } else {
            ctx->err = "hipEventRecord failed at the record fork";
            st = MDG_EHIP;
        }
    }
    // the next step's pruned list, from this step's slot positions (in a
    // step block beside the next step's SHAKE kernels, joined before its
    // list rebuild / pair kernel)
    const bool prune = st == MDG_OK && ctx->precision == MDG_MIXED && ctx->pruneSkin > 0.;
    if (prune) st = mdg_prune_list(ctx);
    ctx->stream = main;
    RETCHK(st);
    LAUNCHCHK();
    if (fork && prune) {
        HIPCHK(hipEventRecord(ctx->evPruneDone, ctx->stream2));
        ctx->prunePending = true;
    }
    if (ctx->profiling) HIPCHK(hipEventRecord(ctx->evCur[11], ctx->stream));
    return MDG_OK;
}

int mdg_check_flags(mdg_ctx *ctx)
{
    int c[CTRL_WORDS];
    HIPCHK(hipMemcpyAsync(c, ctx->ctrl, sizeof c, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    if (c[CTRL_SHAKE_FAIL]) {
        ctx->err = "SHAKE did not converge within maxCycle";
        return 311;
    }
    if (c[CTRL_SHAKE_ORDER]) {
        ctx->err = "a constraint group re-diverged after converging";
        return MDG_ESHAKE_ORDER;
    }
    if (c[CTRL_LIST_OVF]) {
        ctx->err = "cluster-pair list capacity exceeded";
        return MDG_ELIST;
    }
    if (c[CTRL_SEG_OVF]) {
        ctx->err = "a cell holds more atoms than one LDS sort segment";
        return MDG_EUNSUPPORTED;
    }
    return MDG_OK;
}

__global__ void k_set_int(int *p, int v)
{
    if (threadIdx.x == 0) *p = v;
}

// every word to all ones (the SHAKE masks' reset value)
__global__ void k_not64(long m, unsigned long long *p)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) p[i] = ~0ull;
}

__global__ void k_reset_mask(unsigned long long *mask)
{
    if (threadIdx.x < 2 * SHK_MASKW) mask[threadIdx.x] = ~0ull;
}

// Record index 0 and a fresh SHAKE mask, before a sequence of steps.
int mdg_reset_step_state(mdg_ctx *ctx)
{
    if (ctx->bt) {
        RETCHK(mdg_fill_i32(ctx, ctx->bt->stepCtr, ctx->bt->R, 0));
        const long nm = (long)ctx->bt->R * 2 * SHK_MASKW;
        hipLaunchKernelGGL(k_not64, dim3(nblk(nm, 256)), dim3(256), 0, ctx->stream, nm,
                           ctx->bt->masks);
        LAUNCHCHK();
        return MDG_OK;
    }
    hipLaunchKernelGGL(k_set_int, dim3(1), dim3(64), 0, ctx->stream, ctx->ctrl + CTRL_STEP, 0);
    LAUNCHCHK();
    hipLaunchKernelGGL(k_reset_mask, dim3(1), dim3(64), 0, ctx->stream, ctx->shkMask);
    LAUNCHCHK();
    return MDG_OK;
}

static int b_group_kinetic(mdg_ctx *ctx);

extern "C" int mdg_upload_state(mdg_ctx *ctx, const double x[], const double y[],
                                const double z[], const double vx[], const double vy[],
                                const double vz[])
{
    if (!ctx || !x || !vx) return MDG_EARG;
    const size_t b = ctx->n * sizeof(double);
    HIPCHK(hipMemcpyAsync(ctx->x, x, b, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(ctx->y, y, b, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(ctx->z, z, b, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(ctx->vx, vx, b, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(ctx->vy, vy, b, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(ctx->vz, vz, b, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    ctx->step = 0;
    return MDG_OK;
}

extern "C" int mdg_download_state(mdg_ctx *ctx, double x[], double y[], double z[], double vx[],
                                  double vy[], double vz[], double fx[], double fy[], double fz[])
{
    if (!ctx) return MDG_EARG;
    const size_t b = ctx->n * sizeof(double);
    double *h[9] = {x, y, z, vx, vy, vz, fx, fy, fz};
    double *d[9] = {ctx->x, ctx->y, ctx->z, ctx->vx, ctx->vy, ctx->vz, ctx->fx, ctx->fy, ctx->fz};
    for (int k = 0; k < 9; k++)
        if (h[k]) HIPCHK(hipMemcpyAsync(h[k], d[k], b, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return MDG_OK;
}

extern "C" int mdg_initial_energy(mdg_ctx *ctx, MDG_STEP *rec)
{
    if (!ctx || !ctx->grpAtomOff || !ctx->exSymOff) return MDG_EARG;
    if (ctx->P.ewald != 0 && !ctx->theta) {
        ctx->err = "no SPME mesh: the context was created without ewald m1max..m3max";
        return MDG_EARG;
    }
    if (ctx->bt) {
        RETCHK(ensure_records(ctx, 64));
        RETCHK(mdg_reset_step_state(ctx));
        RETCHK(b_energy(ctx, true));
        RETCHK(b_group_kinetic(ctx));
        RETCHK(b_record(ctx, true));
        RETCHK(mdg_check_flags(ctx));
        if (rec)
            HIPCHK(hipMemcpy(rec, ctx->bt->rec, ctx->bt->R * sizeof(MDG_STEP),
                             hipMemcpyDeviceToHost));
        ctx->stepsSinceList = 0;
        return MDG_OK;
    }
    if (ctx->B.type == 3) {
        // the cluster list bins atoms in Cartesian columns
        ctx->err = "the device-resident path needs a cubic or orthorhombic box";
        return MDG_EUNSUPPORTED;
    }
    RETCHK(ensure_records(ctx, 64));
    hipLaunchKernelGGL(k_set_int, dim3(1), dim3(64), 0, ctx->stream, ctx->ctrl + CTRL_STEP, 0);
    RETCHK(mdg_device_energy(ctx, true));
    hipLaunchKernelGGL(k_kinetic, dim3(nblk(ctx->n, 256)), dim3(256), 0, ctx->stream, ctx->n,
                       ctx->mass, ctx->vx, ctx->vy, ctx->vz, ctx->epart + PART_KIN);
    LAUNCHCHK();
    hipLaunchKernelGGL(k_record, dim3(1), dim3(SIDE_BLOCK), 0, ctx->stream,
                       mdg_rec_args(ctx, nblk(ctx->n, 256), 0, nblk(ctx->n, 256)), ctx->ctrl, ctx->recDev, 1);
    LAUNCHCHK();
    RETCHK(mdg_check_flags(ctx));
    if (rec) HIPCHK(hipMemcpy(rec, ctx->recDev, sizeof(MDG_STEP), hipMemcpyDeviceToHost));
    ctx->stepsSinceList = 0;
    return MDG_OK;
}

static int capture(mdg_ctx *ctx, int variant)
{
    hipGraph_t g;
    HIPCHK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    int st = mdg_device_step(ctx, variant == 1);
    hipError_t e = hipStreamEndCapture(ctx->stream, &g);
    if (st) return st;
    if (e != hipSuccess) {
        ctx->err = std::string("hipStreamEndCapture: ") + hipGetErrorString(e);
        return MDG_EHIP;
    }
    HIPCHK(hipGraphInstantiate(&ctx->graphExec[variant], g, nullptr, nullptr, 0));
    HIPCHK(hipGraphDestroy(g));
    return MDG_OK;
}

// neigh.update consecutive steps, from one just after a rebuild to the
// next rebuild, captured as one graph: inside it each step's record runs
// beside the next step's first SHAKE kernel, and consecutive steps follow
// each other without a graph boundary.
static int capture_block(mdg_ctx *ctx)
{
    const int U = ctx->P.update;
    hipGraph_t g;
    HIPCHK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    ctx->recOverlap = true;
    int st = MDG_OK;
    for (int k = 0; k < U && st == MDG_OK; k++) st = mdg_device_step(ctx, k == U - 1);
    if (st == MDG_OK && ctx->sidePending) {
        // the last step's record and prune
        if (ctx->bt) {
            st = b_record(ctx, ctx->sideRebuilt);
        } else {
            hipLaunchKernelGGL(k_record, dim3(1), dim3(SIDE_BLOCK), 0, ctx->stream,
                               mdg_rec_args(ctx, mdg_vv_blocks(ctx), mdg_vir_rows(ctx),
                                            mdg_excl_rows(ctx)),
                               ctx->ctrl, ctx->recDev, ctx->sideRebuilt ? 1 : 0);
            if (hipGetLastError() != hipSuccess) st = MDG_EHIP;
        }
        if (st == MDG_OK) st = mdg_prune_list(ctx);
    }
    ctx->sidePending = false;
    if (st == MDG_OK && ctx->recPending) {
        // the last record joins the captured stream
        if (hipStreamWaitEvent(ctx->stream, ctx->evRecDone, 0) != hipSuccess) st = MDG_EHIP;
    }
    if (st == MDG_OK && ctx->prunePending) {
        if (hipStreamWaitEvent(ctx->stream, ctx->evPruneDone, 0) != hipSuccess) st = MDG_EHIP;
    }
    ctx->recOverlap = false;
    ctx->recPending = false;
    ctx->prunePending = false;
    hipError_t e = hipStreamEndCapture(ctx->stream, &g);
    if (st) return st;
    if (e != hipSuccess) {
        ctx->err = std::string("hipStreamEndCapture: ") + hipGetErrorString(e);
        return MDG_EHIP;
    }
    HIPCHK(hipGraphInstantiate(&ctx->graphBlock, g, nullptr, nullptr, 0));
    HIPCHK(hipGraphDestroy(g));
    return MDG_OK;
}

static void accumulate_phase(mdg_ctx *ctx, const hipEvent_t *e, bool rebuild, int &nReb)
{
    float t, t1, t2;
    hipEventElapsedTime(&t, e[0], e[1]);
    ctx->phase[0] += t;
    hipEventElapsedTime(&t, e[2], e[3]);
    ctx->phase[1] += t;
    hipEventElapsedTime(&t1, e[8], e[9]);
    hipEventElapsedTime(&t2, e[10], e[11]);
    ctx->phase[2] += t1 + t2;
    if (rebuild) {
        hipEventElapsedTime(&t, e[6], e[7]);
        ctx->phase[3] += t;
        nReb++;
    }
    hipEventElapsedTime(&t, e[8], e[11]);
    ctx->phase[4] += t;
}

// use_graph: 0 plain launches, 1 one graph per step variant (replayed),
// whole rebuild periods as one block graph where they fit,
// 2 the whole run captured as one graph.  With profiling on, every mode
// runs plain launches with HIP events bracketing the pair kernel, SPME,
// the integrator, the list rebuild and the step on the context stream,
// read after each step.
extern "C" int mdg_run_steps(mdg_ctx *ctx, int nsteps, int use_graph, MDG_STEP *records)
{
    if (!ctx || nsteps < 0) return MDG_EARG;
    if (ctx->ensCreate != ctx->ens) {
        ctx->err = "the context was made for a thermostat: call mdg_set_bath first (NVT), or run "
                   "the barostats through mdg_vv_integrate";
        return MDG_EUNSUPPORTED;
    }
    if (!ctx->cl.cpJ) {
        ctx->err = "call mdg_initial_energy first";
        return MDG_EARG;
    }
    RETCHK(ensure_records(ctx, nsteps));
    // the step sequence expects a fresh SHAKE mask for its first stage 1
    // (each step's record kernel resets it for the next)
    RETCHK(mdg_reset_step_state(ctx));
    const bool prof = ctx->profiling;
    for (int k = 0; k < 5; k++) ctx->phase[k] = 0.;
    const int stepBase = ctx->step + 1;
    int nReb = 0;
    if (use_graph == 2 && !prof) {
        hipGraph_t g;
        HIPCHK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        int st = MDG_OK;
        for (int s = 0; s < nsteps && st == MDG_OK; s++)
            st = mdg_device_step(ctx, ((ctx->step + 1 + s) % ctx->P.update) == 0);
        hipError_t e = hipStreamEndCapture(ctx->stream, &g);
        if (st) return st;
        if (e != hipSuccess) {
            ctx->err = std::string("hipStreamEndCapture: ") + hipGetErrorString(e);
            return MDG_EHIP;
        }
        hipGraphExec_t ge;
        HIPCHK(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
        HIPCHK(hipGraphDestroy(g));
        HIPCHK(hipGraphLaunch(ge, ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
        HIPCHK(hipGraphExecDestroy(ge));
        ctx->step += nsteps;
    } else {
        const int U = ctx->P.update;
        for (int s = 0; s < nsteps; s++) {
            const int step = ctx->step + 1;
            const bool rebuild = (step % ctx->P.update) == 0;
            if (use_graph && !prof && U > 1 && step % U == 1 && nsteps - s >= U) {
                // a whole rebuild period as one graph
                if (!ctx->graphBlock) RETCHK(capture_block(ctx));
                HIPCHK(hipGraphLaunch(ctx->graphBlock, ctx->stream));
                ctx->step += U;
                s += U - 1;
                continue;
            }
            if (use_graph && !prof) {
                const int v = rebuild ? 1 : 0;
                if (!ctx->graphExec[v]) RETCHK(capture(ctx, v));
                HIPCHK(hipGraphLaunch(ctx->graphExec[v], ctx->stream));
            } else {
                ctx->evCur = ctx->ev;
                RETCHK(mdg_device_step(ctx, rebuild));
                if (prof) {
                    HIPCHK(hipEventSynchronize(ctx->ev[11]));
                    accumulate_phase(ctx, ctx->ev, rebuild, nReb);
                }
            }
            ctx->step = step;
        }
    }
    if (prof && nsteps > 0) {
        for (int k = 0; k < 5; k++)
            if (k != 3) ctx->phase[k] /= nsteps;
        ctx->phase[3] = nReb ? ctx->phase[3] / nReb : 0.;
    }
    RETCHK(mdg_check_flags(ctx));
    if (records && nsteps > 0) {
        const int R = ctx->bt ? ctx->bt->R : 1;
        HIPCHK(hipMemcpy(records, ctx->bt ? ctx->bt->rec : ctx->recDev,
                         (size_t)nsteps * R * sizeof(MDG_STEP), hipMemcpyDeviceToHost));
        for (long s = 0; s < (long)nsteps * R; s++) records[s].step = stepBase + (int)(s / R);
    }
    return MDG_OK;
}

extern "C" int mdg_set_prune(mdg_ctx *ctx, double skin)
{
    if (!ctx || !(skin >= 0.)) return MDG_EARG;
    if (skin > 0. && ctx->precision != MDG_MIXED) {
        ctx->err = "list pruning is a mixed-precision option";
        return MDG_EUNSUPPORTED;
    }
    if (skin > ctx->P.delr) skin = ctx->P.delr;
    ctx->pruneSkin = skin;
    mdg_invalidate_graphs(ctx);
    return MDG_OK;
}

extern "C" double mdg_prune_skin(const mdg_ctx *ctx) { return ctx ? ctx->pruneSkin : 0.; }

extern "C" int mdg_set_pair_path(mdg_ctx *ctx, int path)
{
    if (!ctx || path < 0 || path > 2) return MDG_EARG;
    ctx->pairPath = path;
    mdg_invalidate_graphs(ctx);
    return MDG_OK;
}

extern "C" int mdg_get_launch(const mdg_ctx *ctx, int out[4])
{
    if (!ctx || !out) return MDG_EARG;
    out[0] = ctx->tune.shkSpread;
    out[1] = ctx->tune.pairBlocks;
    out[2] = ctx->tune.spmeFirst;
    out[3] = ctx->tune.spmePrio;
    return MDG_OK;
}

extern "C" int mdg_set_launch(mdg_ctx *ctx, const int in[4])
{
    if (!ctx || !in) return MDG_EARG;
    if (!(in[0] == 0 || in[0] == 1 || in[0] == 2 || in[0] == 4 || in[0] == 8 || in[0] == 16) ||
        in[1] < 1 || (in[2] != 0 && in[2] != 1) || in[3] < 0 || in[3] > 2) {
        ctx->err = "launch shapes: SHAKE spread 0, 1, 2, 4, 8 or 16, pair grid >= 1, SPME-first "
                   "0 or 1, priority 0, 1 (SPME kernels) or 2 (pair kernel)";
        return MDG_EARG;
    }
    ctx->tune.shkSpread = in[0];
    ctx->tune.pairBlocks = in[1];
    ctx->tune.spmeFirst = in[2];
    ctx->tune.spmePrio = in[3];
    mdg_invalidate_graphs(ctx);
    return MDG_OK;
}

extern "C" int mdg_prune_stats(mdg_ctx *ctx, long out[2])
{
    if (!ctx || !out) return MDG_EARG;
    int c[CTRL_WORDS];
    HIPCHK(hipStreamSynchronize(ctx->stream));
    HIPCHK(hipMemcpy(c, ctx->ctrl, sizeof(c), hipMemcpyDeviceToHost));
    out[0] = c[CTRL_PRUNED];
    out[1] = 0;
    ClusterList &C = ctx->cl;
    if (C.pCnt && C.chOff && ctx->pruneSkin > 0.) {
        int nch = 0;
        HIPCHK(hipMemcpy(&nch, C.chOff + C.nclu, sizeof(int), hipMemcpyDeviceToHost));
        nch = std::min(nch, C.maxChunks);
        std::vector<int> cnt(std::max(nch, 1));
        if (nch > 0) HIPCHK(hipMemcpy(cnt.data(), C.pCnt, nch * sizeof(int), hipMemcpyDeviceToHost));
        for (int k = 0; k < nch; k++) out[1] += cnt[k];
    }
    return MDG_OK;
}

extern "C" int mdg_set_profiling(mdg_ctx *ctx, int on)
{
    if (!ctx) return MDG_EARG;
    ctx->profiling = on != 0;
    if (ctx->profiling)
        for (int k = 0; k < 12; k++)
            if (!ctx->ev[k]) HIPCHK(hipEventCreate(&ctx->ev[k]));
    return MDG_OK;
}

extern "C" int mdg_phase_times(mdg_ctx *ctx, double out[5])
{
    if (!ctx || !out) return MDG_EARG;
    for (int k = 0; k < 5; k++) out[k] = ctx->phase[k];
    return MDG_OK;
}

extern "C" int mdg_list_stats(mdg_ctx *ctx, long out[3])
{
    if (!ctx || !out || !ctx->cl.cpOff) return MDG_EARG;
    const int nclu = ctx->cl.nclu;
    int ncp = 0;
    HIPCHK(hipMemcpy(&ncp, ctx->cl.cpOff + nclu, sizeof(int), hipMemcpyDeviceToHost));
    std::vector<unsigned long long> m(ncp);
    if (ncp) HIPCHK(hipMemcpy(m.data(), ctx->cl.cpMask, ncp * sizeof(unsigned long long),
                              hipMemcpyDeviceToHost));
    long bits = 0;
    for (auto v : m) bits += __builtin_popcountll(v);
    out[0] = ncp;
    out[1] = bits;
    out[2] = nclu;
    return MDG_OK;
}

extern "C" int mdg_list_pairs(mdg_ctx *ctx, int *pi, int *pj, long cap, long *npairs)
{
    if (!ctx || !npairs || !ctx->cl.cpOff) return MDG_EARG;
    const int nclu = ctx->cl.nclu;
    int ncp = 0;
    HIPCHK(hipMemcpy(&ncp, ctx->cl.cpOff + nclu, sizeof(int), hipMemcpyDeviceToHost));
    std::vector<int> I(ncp), J(ncp), slot(nclu * MDG_CL);
    std::vector<unsigned long long> m(ncp);
    if (ncp) {
        HIPCHK(hipMemcpy(I.data(), ctx->cl.cpI, ncp * sizeof(int), hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(J.data(), ctx->cl.cpJ, ncp * sizeof(int), hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(m.data(), ctx->cl.cpMask, ncp * sizeof(unsigned long long),
                         hipMemcpyDeviceToHost));
    }
    HIPCHK(hipMemcpy(slot.data(), ctx->cl.slotAtom, slot.size() * sizeof(int), hipMemcpyDeviceToHost));
    long k = 0;
    for (int e = 0; e < ncp; e++)
        for (int bit = 0; bit < 64; bit++)
            if ((m[e] >> bit) & 1ull) {
                int a = slot[I[e] * MDG_CL + (bit & 7)], b = slot[J[e] * MDG_CL + (bit >> 3)];
                if (k < cap && pi && pj) {
                    pi[k] = std::min(a, b);
                    pj[k] = std::max(a, b);
                }
                k++;
            }
    *npairs = k;
    return MDG_OK;
}

// ---------------------------------------------------------------- replica batch

struct BFinArgs {
    MdgParams P;
    int n, hasFrozen;
    const MdgBox *boxes;
    const int *arep, *aoff, *atomSlot;
    const long long *fix, *fiy, *fiz;
    const double *fdx, *fdy, *fdz;
    const float *frx, *fry, *frz;
    const double *frx64, *fry64, *frz64;
    const double *repRec;                // per replica sums of the reciprocal forces
    const int *exOff, *exLst;
    const double *x, *y, *z, *q;
    const int *frozen;
    double *fx, *fy, *fz, *vx, *vy, *vz;
    double *atomE;                       // per atom: correction energy, virial
    unsigned long long *masks;           // SHAKE masks, reset for VV stage 2
    int nmask;
};

// k_finalize for a batch: each atom's reciprocal-force mean is its own
// replica's, the correction uses the replica's box, and the correction
// energy is kept per atom for the per-replica sums.
template <bool MIXED>
__global__ void __launch_bounds__(256) k_b_finalize(BFinArgs A)
{
    if (blockIdx.x == 0)
        for (int t = threadIdx.x; t < A.nmask; t += blockDim.x) A.masks[t] = ~0ull;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= A.n) return;
    const int r = A.arep[i];
    const MdgBox &B = A.boxes[r];
    const double nr = (double)(A.aoff[r + 1] - A.aoff[r]);
    const double m0 = A.repRec[3 * r] / nr, m1 = A.repRec[3 * r + 1] / nr,
                 m2 = A.repRec[3 * r + 2] / nr;
    const int s = A.atomSlot[i];
    double gx, gy, gz;
    if (MIXED) {
        gx = (double)A.fix[s] * (1.0 / MDG_FSCALE) + ((double)A.frx[i] - m0);
        gy = (double)A.fiy[s] * (1.0 / MDG_FSCALE) + ((double)A.fry[i] - m1);
        gz = (double)A.fiz[s] * (1.0 / MDG_FSCALE) + ((double)A.frz[i] - m2);
    } else {
        gx = (A.frx64[i] - m0) + A.fdx[s];
        gy = (A.fry64[i] - m1) + A.fdy[s];
        gz = (A.frz64[i] - m2) + A.fdz[s];
    }
    double ec = 0., vc = 0.;
    for (int k = A.exOff[i]; k < A.exOff[i + 1]; k++) {
        const int j = A.exLst[k];
        const int lo = min(i, j), hi = max(i, j);
        double dx = A.x[hi] - A.x[lo], dy = A.y[hi] - A.y[lo], dz = A.z[hi] - A.z[lo];
        const double r2 = mdg_dist(B, dx, dy, dz);
        const double rr = sqrt(r2), rt = 1. / rr;
        double dC;
        const double e = ewald_corr64(A.P.alpha, A.P.chargeConst * A.q[lo] * A.q[hi], rr, rt, dC);
        const double sgn = (i == lo) ? 1.0 : -1.0;
        gx += sgn * (dC * dx * rt);
        gy += sgn * (dC * dy * rt);
        gz += sgn * (dC * dz * rt);
        if (i == lo) {
            ec += e;
            vc += dC * rr;
        }
    }
    if (A.hasFrozen && A.frozen[i]) {
        gx = gy = gz = 0.;
        A.vx[i] = A.vy[i] = A.vz[i] = 0.;
    }
    A.fx[i] = gx;
    A.fy[i] = gy;
    A.fz[i] = gz;
    A.atomE[2 * i] = ec;
    A.atomE[2 * i + 1] = vc;
}

// sum m v^2 per group (the step-0 record, before any VV stage 2)
__global__ void k_b_group_kin(int nGrp, const int *__restrict__ grpAtomOff,
                              const int *__restrict__ grpAtoms, const double *__restrict__ mass,
                              const double *__restrict__ vx, const double *__restrict__ vy,
                              const double *__restrict__ vz, double *__restrict__ grpKin,
                              double *__restrict__ grpVir)
{
    int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= nGrp) return;
    double s = 0.;
    for (int k = grpAtomOff[g]; k < grpAtomOff[g + 1]; k++) {
        const int i = grpAtoms[k];
        s += mass[i] * (vx[i] * vx[i] + vy[i] * vy[i] + vz[i] * vz[i]);
    }
    grpKin[g] = s;
    grpVir[g] = 0.;
}


// The batch records as a kernel of their own (b_record_block, mdg_side.h):
// one block per replica.
__global__ void __launch_bounds__(SIDE_BLOCK) k_b_record(BRecArgs R, int rebuilt)
{
    b_record_block(R, blockIdx.x, rebuilt);
}

static int b_energy(mdg_ctx *ctx, bool rebuild)
{
    MdgBatch &T = *ctx->bt;
    const int n = ctx->n;
    RETCHK(mdg_spme_fork(ctx));
    // (a rebuild step always starts the SPME branch first: the second
    // branch starts only after the first one's nodes, and the rebuild has
    // some thirty of them)
    const bool spmeFirst = ctx->tune.spmeFirst || rebuild;
    if (spmeFirst) RETCHK(mdg_spme_on(ctx, mdg_b_spme));
    if (rebuild) {
        if (ctx->profiling) HIPCHK(hipEventRecord(ctx->evCur[6], ctx->stream));
        RETCHK(mdg_b_build_list(ctx));
        RETCHK(ensure_slots(ctx));
        RETCHK(mdg_slot_params(ctx));
        if (ctx->profiling) HIPCHK(hipEventRecord(ctx->evCur[7], ctx->stream));
    }
    RETCHK(mdg_b_pair(ctx));
    if (!spmeFirst) RETCHK(mdg_spme_on(ctx, mdg_b_spme));
    if (ctx->spmeForked) HIPCHK(hipStreamWaitEvent(ctx->stream, ctx->evJoin, 0));
    BFinArgs F;
    F.P = ctx->P;
    F.n = n;
    F.hasFrozen = ctx->hasFrozen ? 1 : 0;
    F.boxes = T.boxes;
    F.arep = T.arep;
    F.aoff = T.aoff;
    F.atomSlot = ctx->cl.atomSlot;
    F.fix = ctx->fix;
    F.fiy = ctx->fiy;
    F.fiz = ctx->fiz;
    F.fdx = ctx->fdx;
    F.fdy = ctx->fdy;
    F.fdz = ctx->fdz;
    F.frx = ctx->frx;
    F.fry = ctx->fry;
    F.frz = ctx->frz;
    F.frx64 = ctx->frx64;
    F.fry64 = ctx->fry64;
    F.frz64 = ctx->frz64;
    F.repRec = T.repRec;
    F.exOff = ctx->exSymOff;
    F.exLst = ctx->exSymLst;
    F.x = ctx->x;
    F.y = ctx->y;
    F.z = ctx->z;
    F.q = ctx->q;
    F.frozen = ctx->frozen;
    F.fx = ctx->fx;
    F.fy = ctx->fy;
    F.fz = ctx->fz;
    F.vx = ctx->vx;
    F.vy = ctx->vy;
    F.vz = ctx->vz;
    F.atomE = T.atomE;
    F.masks = T.masks;
    F.nmask = T.R * 2 * SHK_MASKW;
    if (ctx->precision == MDG_MIXED)
        hipLaunchKernelGGL(k_b_finalize<true>, dim3(nblk(n, 256)), dim3(256), 0, ctx->stream, F);
    else
        hipLaunchKernelGGL(k_b_finalize<false>, dim3(nblk(n, 256)), dim3(256), 0, ctx->stream, F);
    LAUNCHCHK();
    return MDG_OK;
}

static int b_group_kinetic(mdg_ctx *ctx)
{
    hipLaunchKernelGGL(k_b_group_kin, dim3(nblk(ctx->nGrp, 256)), dim3(256), 0, ctx->stream,
                       ctx->nGrp, ctx->grpAtomOff, ctx->grpAtoms, ctx->mass, ctx->vx, ctx->vy,
                       ctx->vz, ctx->bt->grpKin, ctx->grpVir);
    LAUNCHCHK();
    return MDG_OK;
}

BRecArgs mdg_b_rec_args(mdg_ctx *ctx)
{
    MdgBatch &T = *ctx->bt;
    BRecArgs R;
    R.chargeConst = ctx->P.chargeConst;
    R.R = T.R;
    R.recCap = T.recCap;
    R.boxes = T.boxes;
    R.eSelf = T.eSelf;
    R.eNet = T.eNet;
    R.aoff = T.aoff;
    R.goff = T.goff;
    R.cluStart = T.cluStart;
    R.chOff = ctx->cl.chOff;
    R.maxChunks = ctx->cl.maxChunks;
    R.chE = T.chE;
    R.atomE = T.atomE;
    R.grpKin = T.grpKin;
    R.grpVir = ctx->grpVir;
    R.repInfl = T.repInfl;
    R.stepCtr = T.stepCtr;
    R.rec = T.rec;
    return R;
}

static int b_record(mdg_ctx *ctx, bool rebuild)
{
    hipLaunchKernelGGL(k_b_record, dim3(ctx->bt->R), dim3(SIDE_BLOCK), 0, ctx->stream,
                       mdg_b_rec_args(ctx), rebuild ? 1 : 0);
    LAUNCHCHK();
    return MDG_OK;
}

// ---------------------------------------------------------------- autotuner

// Wall time (ms, HIP events on the context stream) of nsteps steps from a
// fresh start at state (x, v), after a warm-up that captures the graphs.
static int time_steps(mdg_ctx *ctx, const double *state, int nsteps, float &ms)
{
    const size_t n = ctx->n;
    double *d[6] = {ctx->x, ctx->y, ctx->z, ctx->vx, ctx->vy, ctx->vz};
    for (int k = 0; k < 6; k++)
        HIPCHK(hipMemcpyAsync(d[k], state + k * n, n * sizeof(double), hipMemcpyDeviceToDevice,
                              ctx->stream));
    const int step0 = ctx->step;
    RETCHK(mdg_initial_energy(ctx, nullptr));
    RETCHK(mdg_run_steps(ctx, ctx->P.update, 1, nullptr));   // captures both variants
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    // the faster of two timed runs (a candidate is not penalised by a
    // one-off stall)
    int st = MDG_OK;
    ms = 1e30f;
    for (int rep = 0; rep < 2 && st == MDG_OK; rep++) {
        HIPCHK(hipEventRecord(e0, ctx->stream));
        st = mdg_run_steps(ctx, nsteps, 1, nullptr);
        if (st == MDG_OK) {
            float t = 0.f;
            HIPCHK(hipEventRecord(e1, ctx->stream));
            HIPCHK(hipEventSynchronize(e1));
            HIPCHK(hipEventElapsedTime(&t, e0, e1));
            ms = std::min(ms, t);
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    ctx->step = step0;
    return st;
}

// Picks the launch shapes of this system on this GPU by timing whole steps
// (graph-replayed, as mdg_run_steps runs them): the three-site SHAKE
// spread, the pair kernel's grid cap, the branch order at the fork, the
// SPME kernels' wave priority, and (when pruning is on) the pruned list
// against the full one, one after the other with the rest held.  The uploaded state is restored afterwards and
// the initial energy recomputed.  out (may be NULL): the chosen spread and
// pair-grid cap (mdg_get_launch has all of them).
extern "C" int mdg_autotune(mdg_ctx *ctx, int nsteps, int out[2])
{
    if (!ctx || nsteps <= 0) return MDG_EARG;
    if (!ctx->cl.cpJ) {
        ctx->err = "call mdg_initial_energy first";
        return MDG_EARG;
    }
    const size_t n = ctx->n;
    double *state = nullptr;
    HIPCHK(hipMalloc(&state, 6 * n * sizeof(double)));
    double *d[6] = {ctx->x, ctx->y, ctx->z, ctx->vx, ctx->vy, ctx->vz};
    for (int k = 0; k < 6; k++)
        HIPCHK(hipMemcpyAsync(state + k * n, d[k], n * sizeof(double), hipMemcpyDeviceToDevice,
                              ctx->stream));
    int st = MDG_OK;
    auto sweep = [&](int *param, std::initializer_list<int> values) {
        int best = *param;
        float bestMs = 1e30f;
        for (int v : values) {
            if (st) return;
            *param = v;
            mdg_invalidate_graphs(ctx);
            float ms = 0.f;
            st = time_steps(ctx, state, nsteps, ms);
            if (st == MDG_OK && ms < bestMs) {
                bestMs = ms;
                best = v;
            }
        }
        *param = best;
    };
    if (ctx->shakeW3) sweep(&ctx->tune.shkSpread, {0, 1, 2, 4, 8});
    // the pair grid, the fork order and the priorities interact (a full
    // grid leaves the SPME branch no room beside it, which only matters
    // when that branch starts early): two rounds of the three sweeps
    for (int round = 0; round < 2; round++) {
        sweep(&ctx->tune.pairBlocks, {512, 768, 1024, 2048, 4096, 8192, 16384});
        sweep(&ctx->tune.spmeFirst, {0, 1});
        sweep(&ctx->tune.spmePrio, {0, 1, 2});
    }
    if (ctx->pruneSkin > 0. && st == MDG_OK) {
        // the pruned list, or the full list every step (no prune kernel)
        const double skin = ctx->pruneSkin;
        float msOn = 0.f, msOff = 0.f;
        mdg_invalidate_graphs(ctx);
        st = time_steps(ctx, state, nsteps, msOn);
        ctx->pruneSkin = 0.;
        mdg_invalidate_graphs(ctx);
        if (st == MDG_OK) st = time_steps(ctx, state, nsteps, msOff);
        ctx->pruneSkin = (st == MDG_OK && msOff < msOn) ? 0. : skin;
    }
    mdg_invalidate_graphs(ctx);
    for (int k = 0; k < 6 && st == MDG_OK; k++)
        if (hipMemcpyAsync(d[k], state + k * n, n * sizeof(double), hipMemcpyDeviceToDevice,
                           ctx->stream) != hipSuccess)
            st = MDG_EHIP;
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipFree(state);
    if (st == MDG_OK) st = mdg_initial_energy(ctx, nullptr);
    if (out) {
        out[0] = ctx->tune.shkSpread;
        out[1] = ctx->tune.pairBlocks;
    }
    return st;
}
