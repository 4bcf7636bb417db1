// mdg_internal.h -- shared state and device helpers of libmdgpu (gfx950).
#pragma once

#include <hip/hip_runtime.h>
#include <rocfft/rocfft.h>

#include <algorithm>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/mdgpu.h"

#define MDG_WAVE 64
#define MDG_CL 8                 // atoms per cluster (8 x 8 = one wavefront per cluster pair)
#define MDG_GRP_ATOMS 8          // max atoms in one constraint group
#define MDG_GRP_CONS 16          // max constraints in one constraint group
#define MDG_SEG_MAX 4096         // max atoms in one cell segment sorted in LDS
#define MDG_CHUNK 32             // j-clusters per wavefront in the pair kernel
#define MDG_PRUNE_SKIN 0.8       // default pruning distance beyond the cutoff (mixed), A

// Fixed-point scales: forces in internal units (10 J/mol/A) carry 2^32
// fractional steps in an int64 (|F| < 2^31); energies and virials 2^24
// (|E| < 2^39); spread charges 2^40 (|Q| < 2^23).
#define MDG_FSCALE 4294967296.0
#define MDG_ESCALE 16777216.0
#define MDG_QSCALE 1099511627776.0

// Host/device copy of the run settings (PARAM/EWALD/NEIGH subset with the
// derived constants of io.c:714-723) and the box (PBC after init_box).
struct MdgBox {
    int type;
    double a1, a2, a3, b1, b2, b3, c1, c2, c3;
    double u1, u2, u3, v1, v2, v3, w1, w2, w3;
    double pa, pb, pc, u, v, w, det, vol;
};

struct MdgParams {
    int n;
    double cutOff, cutOn, delr, chargeConst, alpha;
    double cutOff2, cutOn2, switch2, rcutOff, rcutOff2, cutnb2;
    double timeStep, rTimeStep, tolShake;
    int maxCycle, update, linkRatio;
    int m1, m2, m3, nbsp;
    int elecType;
    int ewald;                   // keyEwald: 2 Ewald real space + SPME; 0 the cutoff
                                 // Coulomb functor of elecType (energy.c:216-345); 1
                                 // (classic k-space sum) is not on the device path
    int vdwOn;                   // vdwType == VSWITCH (vdw full / none evaluate nothing)
    int spmePrio;                // 1: the binned SPME kernels raise their wave priority
                                 // (s_setprio); 2: the cluster pair kernel does
};

// Reference-format list on the device.
struct RefList {
    int *neighList = nullptr;    // flat j indices, in neighOrder order of i
    long cap = 0;
    long count = 0;
    int *neighPair = nullptr;    // per atom
    int *neighOrder = nullptr;   // position -> atom
    int *cnt = nullptr;          // per position
    long *off = nullptr;         // per position exclusive offsets (+1)
};

// Cluster-pair list: atoms live in "slots" (cluster*8 + k), ordered by
// spatial columns; each entry (I, J, mask) covers the 64 atom pairs of
// clusters I <= J, bit a*8+b set iff (slot I*8+a, slot J*8+b) is a listed
// pair of the reference list.
struct ClusterList {
    int nclu = 0, nslot = 0;
    int *slotAtom = nullptr;     // slot -> atom (-1 padding)
    int *atomSlot = nullptr;     // atom -> slot
    float4 *bb = nullptr;        // per cluster: centre xyz, unused
    float4 *bbh = nullptr;       // per cluster: half extents xyz, unused
    int *cpOff = nullptr;        // per I cluster: first entry (nclu+1)
    int *cpJ = nullptr;          // entry -> J
    unsigned long long *cpMask = nullptr;
    int *cpI = nullptr;          // entry -> I
    long cap = 0;
    int *chOff = nullptr;        // per I: first chunk (nclu+1)
    int2 *chunk = nullptr;       // chunk -> (I, first entry)
    int maxChunks = 0;
    int *tmpCnt = nullptr;       // per I: entries found
    int *tmpJ = nullptr;         // per I scratch [nclu][kmax]
    unsigned long long *tmpMask = nullptr;
    int kmax = 0;
    // spatial sort scratch
    int ncol = 0, nfx = 0, nfy = 0;
    int *colCnt = nullptr, *colStart = nullptr, *colPad = nullptr, *colSlot = nullptr;
    int *cluCol = nullptr;       // column of each cluster (-1 unused)
    // per slot, for the list predicate: the atom's predicate coordinates
    // (scaled for link cells, else positions), link cell and atom
    double4 *pk = nullptr;
    int *sortTmp = nullptr;
    float *sortKey = nullptr;
    // pruned copy of the list (mixed pair kernel): per chunk of the full
    // list, its entries with a pair within cutOff + pruneSkin at the slot
    // positions of the last prune, at the same offsets
    int *pJ = nullptr;
    unsigned long long *pMask = nullptr;
    int *pCnt = nullptr;         // per chunk: entries kept
    // relative coordinates (mixed pair kernel, one box): a fixed FP64
    // origin per cluster for the life of the list (its build-time bounding
    // box centre), and per entry the minimum image of O_J - O_I in FP32
    // (pD: the same for the pruned list's entries)
    double *cO = nullptr;        // [3][maxclu]
    float4 *cpD = nullptr, *pD = nullptr;
};

// Replica batch (mdg_batch.hip): many independent boxes concatenated in
// one context.  Atoms, constraint groups, clusters and chunks of replica r
// are contiguous; every kernel of a step covers all replicas and reads the
// replica's box, mesh and offsets from these tables.
struct MdgFftGroup {
    int m = 0, count = 0;        // cubic mesh edge and replicas in the group
    long g0 = 0, h0 = 0;         // first real-mesh / half-spectrum element
    rocfft_plan fwd = nullptr, bwd = nullptr, fwd32 = nullptr, bwd32 = nullptr;
};

struct MdgBatch {
    int R = 0;
    std::vector<int> hAoff, hGoff, hMesh, hColOff, hNfx, hNfy;
    std::vector<long> hGridOff, hHOff;
    std::vector<MdgBox> hBox;
    std::vector<double> hESelf, hENet, hSystq;
    int *aoff = nullptr, *goff = nullptr, *arep = nullptr, *grep = nullptr;
    int *colOff = nullptr, *nfx = nullptr, *nfy = nullptr, *mesh = nullptr;
    long *gridOff = nullptr, *hOff = nullptr;
    MdgBox *boxes = nullptr;
    double *eSelf = nullptr, *eNet = nullptr;
    int *cluStart = nullptr;     // per replica first cluster (R + 1), per build
    int *crep = nullptr;         // per cluster: replica (-1 empty)
    double *chE = nullptr;       // per chunk: eDir, eVdw, virDir, virVdw, nIn
    double *atomE = nullptr;     // per atom: correction energy, virial
    double *grpKin = nullptr;    // per group: sum m v^2
    double *repInfl = nullptr;   // per replica: sum theta|F|^2, st0..st5
    double *repRec = nullptr;    // per replica: sums of the reciprocal forces
    unsigned long long *masks = nullptr;   // per replica SHAKE masks
    MDG_STEP *rec = nullptr;     // [step][replica]
    int recCap = 0;              // steps
    int *stepCtr = nullptr;      // per replica: next record index
    long Gtot = 0, Htot = 0;
    std::vector<MdgFftGroup> fft;
    // binned spreading (mixed): per-replica bin and tile offsets, tile -> replica
    bool binned = false;
    int nTiles = 0;
    int *binOff = nullptr, *tileOff = nullptr, *tileRep = nullptr;
    hipGraphExec_t graph[2] = {nullptr, nullptr};
};

// Bonded topology on the device (mdg_bonded.hip): the term tables in the
// reference layouts, the 1-4 pair list and its LJ parameters, and how the
// 1-4 pairs are evaluated (ewald14_energy or nonbond14_energy).
struct MdgBonded {
    MDG_BOND *bond = nullptr, *ub = nullptr;
    MDG_ANGLE *angle = nullptr;
    MDG_DIHE *dihe = nullptr, *impr = nullptr;
    int *l14 = nullptr;          // 1-4 pairs (i, j)
    double *eps14 = nullptr, *sig14 = nullptr;
    long long *fix = nullptr;    // int64 per-atom accumulators (x | y | z) of the drop-in
    int nBond = 0, nUb = 0, nAngle = 0, nDihe = 0, nImpr = 0, n14 = 0;
    int capBond = 0, capUb = 0, capAngle = 0, capDihe = 0, capImpr = 0, cap14 = 0;
    double scal14 = 0.;
    int ewald14 = 1, elec14 = 0, vdw14 = 1;
    bool on = false;             // the device-resident step adds these terms
};

// Launch shapes the autotuner chooses (mdg_autotune).
struct MdgTune {
    int shkSpread = 1;           // three-site SHAKE: quads per molecule slot (1, 2, 4, 8, 16),
                                 // or 0: one molecule per lane
    int pairBlocks = 1024;       // cap on the cluster pair kernel's grid (4 blocks per CU leave
                                 // wave slots for the SPME kernels beside it)
    int spmeFirst = 0;           // steps without a rebuild enqueue the SPME branch before the
                                 // pair branch (the first branch after a fork starts first;
                                 // rebuild steps always do)
    int spmePrio = 0;            // 1: the binned SPME kernels' waves issue ahead of the pair
                                 // kernel's when both share a SIMD; 2: the pair kernel's ahead
};

struct mdg_ctx {
    int precision = MDG_FP64;
    int n = 0;
    int keyLink = 0;
    int nlc[3] = {0, 0, 0};
    MdgParams P{};
    MdgBox B{};
    std::string err;
    hipStream_t stream = nullptr;
    // second stream of the step: SPME runs on it beside the list rebuild and
    // the pair kernel (forked and joined with events, so a captured step
    // graph has the two branches)
    hipStream_t stream2 = nullptr;
    hipEvent_t evFork = nullptr, evJoin = nullptr;
    // step blocks: each step's record kernel runs on stream2 beside the next
    // step's first SHAKE kernel (joined before that step's second one)
    hipEvent_t evRecFork = nullptr, evRecDone = nullptr;
    bool recOverlap = false, recPending = false;
    // ... and after the record the list prune (evPruneDone), joined before
    // the next step's list rebuild / pair kernel
    hipEvent_t evPruneDone = nullptr;
    bool prunePending = false;
    // step blocks of one box with three-site SHAKE: a step's record and
    // list prune run in spare blocks of the next step's first SHAKE kernel
    bool sidePending = false, sideRebuilt = false;
    bool spmeForked = false;     // the current step's SPME went to stream2
    int device = 0;

    // state (FP64, reference atom order)
    double *x = nullptr, *y = nullptr, *z = nullptr;
    double *vx = nullptr, *vy = nullptr, *vz = nullptr;
    double *fx = nullptr, *fy = nullptr, *fz = nullptr;
    // topology
    double *q = nullptr, *eps = nullptr, *sig = nullptr, *mass = nullptr, *rmass = nullptr;
    int *frozen = nullptr, *nAtConst = nullptr;
    bool hasFrozen = false;
    int nConst = 0;
    int *ca = nullptr, *cb = nullptr;
    double *rc2 = nullptr;
    double *shkThr = nullptr;    // per constraint: exact threshold of the stage-1 check
    double *dd = nullptr;        // constraint vectors of the current VV stage [3][nConst]
    // constraint groups (connected components; singletons for free atoms)
    int nGrp = 0;
    int *grpAtomOff = nullptr, *grpAtoms = nullptr, *grpConOff = nullptr, *grpCons = nullptr;
    int *grpCycle = nullptr;     // per group: cycles to own convergence
    double *grpVir = nullptr;    // per group: SHAKE virial
    int *ctrl = nullptr;         // device control words (see CTRL_*)
    double *shkScr = nullptr, *shkVir = nullptr;   // SHAKE state window (mdg_integrate.hip)
    int *shkFirst = nullptr, *shkKend = nullptr;
    int *shkRing = nullptr;      // per group: the sweeps in the two ring slots
    unsigned long long *shkMask = nullptr;   // [stage 1 | stage 2][SHK_MASKW]
    bool shakeW3 = false;        // all groups are three-site molecules
    // exclusions (CSR; symmetric for link cells, ordered otherwise)
    int *exOff = nullptr, *exLst = nullptr;
    int nExcl = 0;
    int exSpan = 0;              // largest |i - j| of an excluded pair (the list skips the lookup beyond)
    int *exSymOff = nullptr, *exSymLst = nullptr;   // both orientations of every pair
    int nExSym = 0;
    // parameter adjoints dE/d(q, eps, sig) and the SPME mesh potential (mdg_adjoint.hip)
    double *adjQ = nullptr, *adjE = nullptr, *adjS = nullptr, *adjPot = nullptr;
    // trajectory adjoint: the mesh of theta * Q_u (G) and the second-
    // derivative scratch (slot-order u, H u and mixed terms; atom-order sums)
    double *gridU = nullptr, *hvpBuf = nullptr;
    std::vector<int> hExOff, hExLst;
    std::vector<MDG_CONSTRAINT> hCons;   // the constraint list of mdg_set_topology
    bool exclBuilt = false;

    // link-cell / Verlet scratch
    double *xu = nullptr, *yu = nullptr, *zu = nullptr;
    int *cellOf = nullptr, *cellCnt = nullptr, *cellStart = nullptr;
    int ncell = 0;
    unsigned long long *lcKeys = nullptr;
    double *lcS3 = nullptr;

    RefList rl;
    ClusterList cl;
    MdgBonded bd;

    // forces / energies of the device-resident path (slot order)
    double *sx = nullptr, *sy = nullptr, *sz = nullptr;        // slot positions
    double *dispPart = nullptr;  // per k_slot_pos block: max squared move since the last prune
    int nDispPart = 0;
    // mixed, one box: slot positions relative to their cluster's origin
    // (FP32, minimum image), and per slot-writing block the largest
    // |component| of them (relPart); pairPath 0 automatic, 1 FP64
    // displacements, 2 relative with every entry's cutoff test in FP64
    float4 *spr = nullptr;
    double *relPart = nullptr;
    int pairPath = 0;
    double pruneSkin = 0.;       // > 0: the mixed pair kernel may use the pruned list
    bool pruneFresh = true;      // the list was rebuilt since the last slot update
    // device steps without a rebuild (one box of three-site molecules,
    // mixed): stage 1's replay kernel writes the slot positions
    bool slotsInStage1 = false;  // set by the step for its stage 1
    bool slotsDone = false;      // stage 1 wrote them: the pair step skips k_slot_pos
    float4 *sp32 = nullptr;                                     // slot q, eps, sig
    double *sq = nullptr, *seps = nullptr, *ssig = nullptr;
    double *fdx = nullptr, *fdy = nullptr, *fdz = nullptr;     // FP64 slot forces
    double *scal = nullptr;                                     // reduced scalars (S_*)
    long long *fix = nullptr, *fiy = nullptr, *fiz = nullptr;  // int64 fixed point
    float *frx = nullptr, *fry = nullptr, *frz = nullptr;      // SPME forces (mixed)
    double *frx64 = nullptr, *fry64 = nullptr, *frz64 = nullptr;
    double *epart = nullptr;     // per-block FP64 partials
    int epartCap = 0;
    MDG_STEP *recDev = nullptr;  // per-step records
    int recCap = 0;

    // SPME
    int m1 = 0, m2 = 0, m3 = 0, mh = 0;   // mh = m3/2+1
    double *theta = nullptr;     // [m1][m2][mh]
    float *theta32 = nullptr;
    double *grid64 = nullptr;    // real mesh (FP64)
    double *cgrid64 = nullptr;   // complex half spectrum (FP64, interleaved)
    float *grid32 = nullptr;
    float *cgrid32 = nullptr;
    long long *gridq = nullptr;  // fixed-point spread mesh (mixed)
    // binned spreading (mixed): atoms sorted by mesh bin every step
    int nbx = 0, nby = 0, nbz = 0;   // bins per dimension (0: not used)
    int *binCnt = nullptr;       // atoms per bin; zero between steps
    int *binBucket = nullptr;    // per bin: up to SPME_BCAP atom indices
    int *binOut = nullptr;       // atoms not in a bucket (points not one run, full bucket)
    int *binNOut = nullptr;      // [0] count of binOut; zero between steps
    int4 *splI = nullptr;        // per atom: spline base index per dimension (binned path)
    float4 *splB = nullptr;      // per atom: FP32 spline values [3 * nq][n], nq = (order + 3) / 4
    float4 *splD = nullptr;      // per atom: their derivatives, same layout
    rocfft_plan fwd = nullptr, bwd = nullptr, fwd32 = nullptr, bwd32 = nullptr;
    rocfft_execution_info info = nullptr;
    void *fftWork = nullptr;
    size_t fftWorkSize = 0;
    double eSelf = 0., eNet = 0., systq = 0.;
    // the box vectors (a1..c3) the derived state above (link-cell counts,
    // net-charge energy, influence function) was made for
    double boxKey[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    double *ensSave = nullptr;   // positions and velocities [6][n] across a barostat trial cycle
    // device-resident thermostat (mdg_set_bath): the ensemble the context was
    // made for and the one set; bathDev = chiT, conint, lambda, kin, then the
    // last finished step's conint, consv, chiT (read by its record)
    int ensCreate = 0, ens = 0;
    double tauT = 0., kinTemp0 = 0.;
    double *bathDev = nullptr;

    // timing
    bool profiling = false;
    hipEvent_t ev[12] = {};
    hipEvent_t *evCur = ev;      // the event set the next step records into
    double phase[5] = {0, 0, 0, 0, 0};

    // scan scratch (mdg_util.hip)
    void *scanBuf[8] = {};
    size_t scanSize[8] = {};

    // graph
    hipGraphExec_t graphExec[2] = {nullptr, nullptr};
    hipGraphExec_t graphBlock = nullptr;   // neigh.update consecutive steps ending in a rebuild
    MdgBatch *bt = nullptr;      // replica batch, when created by mdg_batch_create
    MdgTune tune;
    int graphStepsPerLaunch = 0;
    int step = 0;                // last completed MD step
    int stepsSinceList = 0;
};

// control words in ctx->ctrl
enum {
    CTRL_SHAKE_MAX = 0,          // max cycles of any group (stage A -> B)
    CTRL_SHAKE_FAIL = 1,         // a group did not converge within maxCycle
    CTRL_SHAKE_ORDER = 2,        // a group re-diverged after converging
    CTRL_LIST_OVF = 3,           // cluster scratch overflow
    CTRL_STEP = 4,               // device-side step counter
    CTRL_REBUILT = 5,
    CTRL_SEG_OVF = 6,
    CTRL_PRUNE = 7,              // 1: the pruned list belongs to the current slot positions
    CTRL_PRUNED = 8,             // pair launches that ran on the pruned list
    CTRL_WORDS = 12
};

// reduced scalars in ctx->scal
enum {
    S_EDIR = 0, S_EVDW, S_VIRDIR, S_VIRVDW, S_NIN,   // pair kernel (5)
    S_ECORR = 5, S_VIRCORR,                          // exclusion correction (2)
    S_EREC = 7,                                      // sum theta|F|^2, then st0..st5 (7)
    S_FRX = 14, S_FRY, S_FRZ,                        // sums of reciprocal forces (3)
    S_KIN = 17,                                      // sum m v^2, kinetic stress (7)
    S_VIRSHAKE = 24,
    S_REF = 25,                                      // drop-in pair loop (6)
    S_REFCORR = 31,                                  // drop-in correction (2)
    S_BONDED = 33,                                   // bonded and 1-4 sums (11)
    S_ENS = 48,                                      // thermostat sums (7)
    S_COUNT = 56
};

// regions of ctx->epart (per-block partial sums), each PART_R doubles
#define PART_R (1 << 18)
// the cluster pair kernel loops over chunks with at most this many blocks
#define MDG_PAIR_MAX_BLOCKS 16384
enum {
    PART_PAIR = 0, PART_EXCL = PART_R, PART_INFL = 2 * PART_R, PART_GATHER = 3 * PART_R,
    PART_KIN = 4 * PART_R, PART_MISC = 5 * PART_R, PART_REF = 6 * PART_R,
    PART_BOND = 7 * PART_R, PART_ENS = 8 * PART_R, PART_TOTAL = 9 * PART_R
};

// ---------------------------------------------------------------- helpers
#define HIPCHK(call)                                                          \
    do {                                                                      \
        hipError_t e_ = (call);                                               \
        if (e_ != hipSuccess) {                                               \
            ctx->err = std::string(#call) + ": " + hipGetErrorString(e_);     \
            return MDG_EHIP;                                                  \
        }                                                                     \
    } while (0)

#define LAUNCHCHK()                                                           \
    do {                                                                      \
        hipError_t e_ = hipGetLastError();                                    \
        if (e_ != hipSuccess) {                                               \
            ctx->err = std::string("kernel launch: ") + hipGetErrorString(e_);\
            return MDG_EHIP;                                                  \
        }                                                                     \
    } while (0)

#define RETCHK(call)                                                          \
    do {                                                                      \
        int r_ = (call);                                                      \
        if (r_ != 0) return r_;                                               \
    } while (0)

static inline int nblk(long n, int b) { return (int)((n + b - 1) / b); }
static inline int mdg_pair_blocks(const mdg_ctx *ctx)
{
    return std::min(nblk(ctx->cl.maxChunks, 4), std::min(ctx->tune.pairBlocks, MDG_PAIR_MAX_BLOCKS));
}

// utils.c:606-609: round half away from zero through a truncating cast.
__host__ __device__ inline int mdg_nint(double x)
{
    return x >= 0. ? (int)(x + 0.5) : (int)(x - 0.5);
}

// utils.c:38-80 (dist): minimum image of d in place, |d|^2 as
// (dx^2 + dy^2) + dz^2 without contraction.
__host__ __device__ inline double mdg_dist(const MdgBox &b, double &dx, double &dy, double &dz)
{
#pragma clang fp contract(off)
    double xt, yt, zt;
    switch (b.type) {
    case 1:
        dx -= b.a1 * mdg_nint(dx * b.u1);
        dy -= b.a1 * mdg_nint(dy * b.u1);
        dz -= b.a1 * mdg_nint(dz * b.u1);
        break;
    case 2:
        dx -= b.a1 * mdg_nint(dx * b.u1);
        dy -= b.b2 * mdg_nint(dy * b.v2);
        dz -= b.c3 * mdg_nint(dz * b.w3);
        break;
    case 3:
        xt = dx * b.u1 + dy * b.u2 + dz * b.u3;
        yt = dx * b.v1 + dy * b.v2 + dz * b.v3;
        zt = dx * b.w1 + dy * b.w2 + dz * b.w3;
        xt -= mdg_nint(xt);
        yt -= mdg_nint(yt);
        zt -= mdg_nint(zt);
        dx = xt * b.a1 + yt * b.b1 + zt * b.c1;
        dy = xt * b.a2 + yt * b.b2 + zt * b.c2;
        dz = xt * b.a3 + yt * b.b3 + zt * b.c3;
        break;
    default:
        break;
    }
    return dx * dx + dy * dy + dz * dz;
}

#define SQRTPI 1.772453850905516     // global.h:54

// ewald.c:371-384: erf correction of an excluded pair; returns the energy,
// dE receives dE/dr.
__device__ inline double ewald_corr64(double alpha, double qel, double r, double rt, double &dE)
{
    double ar = alpha * r;
    double e = -qel * erf(ar) * rt;
    double d = 2.0 * qel * alpha * exp(-(ar * ar)) / SQRTPI;
    dE = -(e + d) * rt;
    return e;
}

// Fixed-point conversions (round to nearest).
__device__ inline long long mdg_tofix(double v, double scale)
{
    return (long long)__double2ll_rn(v * scale);
}

__device__ inline void mdg_atomic_add_fix(long long *p, double v, double scale)
{
    long long iv = mdg_tofix(v, scale);
    if (iv != 0) atomicAdd((unsigned long long *)p, (unsigned long long)iv);
}

// Wave-uniform load through the scalar data cache (constant address
// space): i must be the same in every lane of the wavefront, and the data
// must not be written by the running kernel.
template <class T>
__device__ inline T mdg_uload(const T *p, long i)
{
    return ((const __attribute__((address_space(4))) T *)p)[i];
}

// Wave-level reductions (64 lanes).
__device__ inline double mdg_wave_sum(double v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ inline float mdg_wave_sumf(float v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Block reduction of NV doubles into out[blockIdx.x*NV + k] (blockDim <= 1024).
template <int NV>
__device__ inline void mdg_block_partials(double (&v)[NV], double *out)
{
    __shared__ double red[16][NV];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < NV; k++) {
        double s = mdg_wave_sum(v[k]);
        if (lane == 0) red[wid][k] = s;
    }
    __syncthreads();
    if (threadIdx.x < NV) {
        double s = 0.;
        const int nw = (blockDim.x + 63) >> 6;
        for (int w = 0; w < nw; w++) s += red[w][threadIdx.x];
        out[(long)blockIdx.x * NV + threadIdx.x] = s;
    }
}

// Relative coordinates of slot s (cluster c) for the mixed pair kernel:
// the minimum image of the position minus the cluster's origin (a
// rectangular box; cO is [3][nclu]), stored as FP32.  Returns the largest
// |component| (+inf for a NaN position).
__device__ inline double mdg_slot_rel(const MdgBox &B, const double *cO, int nclu, int c, double x,
                                      double y, double z, float4 *spr, int s)
{
    const double L0 = B.a1, L1 = B.type == 1 ? B.a1 : B.b2, L2 = B.type == 1 ? B.a1 : B.c3;
    const double i0 = B.u1, i1 = B.type == 1 ? B.u1 : B.v2, i2 = B.type == 1 ? B.u1 : B.w3;
    double u = x - cO[c], v = y - cO[nclu + c], w = z - cO[2 * (long)nclu + c];
    u -= L0 * rint(u * i0);
    v -= L1 * rint(v * i1);
    w -= L2 * rint(w * i2);
    spr[s] = make_float4((float)u, (float)v, (float)w, 0.f);
    const double r = fmax(fabs(u), fmax(fabs(v), fabs(w)));
    return r <= 1e300 ? r : INFINITY;
}

// Largest value over the block into out[blockIdx.x] (every thread calls).
__device__ inline void mdg_block_max(double v, double *out)
{
    __shared__ double wm[16];
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    if ((threadIdx.x & 63) == 0) wm[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = wm[0];
        for (int w = 1; w < (int)((blockDim.x + 63) >> 6); w++) m = fmax(m, wm[w]);
        out[blockIdx.x] = m;
    }
}

// ---------------------------------------------------------------- internal API
void mdg_box_from(const MDG_PBC *in, MdgBox &b);
void mdg_params_from(const MDG_CTRL *ctrl, const MDG_PARAM *p, MdgParams &P);
// the reference-list pair kernel (k_pair_ref64) on a stream: per-block
// partials [eElec, eVdw, virElec, virVdw, nIn, nListed] into part
int mdg_pair_ref64_launch(hipStream_t stream, const MdgParams &P, const MdgBox &B, int elecType,
                          const double *x, const double *y, const double *z, const double *q,
                          const double *eps, const double *sig, const int *neighList,
                          const long *off, const int *order, double *fx, double *fy, double *fz,
                          double *part);
int mdg_reduce_partials_on(hipStream_t stream, const double *part, int nblocks, int nv,
                           double *out);
int mdg_alloc_state(mdg_ctx *ctx);
int mdg_build_exclusions(mdg_ctx *ctx, int nBond, const MDG_BOND *bond, int nConst,
                         const MDG_CONSTRAINT *cons, int nAngle, const MDG_ANGLE *angle,
                         int nDihe = 0, const MDG_DIHE *dihe = nullptr, int nImpr = 0,
                         const MDG_DIHE *impr = nullptr, std::vector<int> *list14 = nullptr);

// bonded terms and 1-4 pairs (mdg_bonded.hip)
int mdg_bonded_upload(mdg_ctx *ctx, int nBond, const MDG_BOND *bond, int nUb, const MDG_BOND *ub,
                      int nAngle, const MDG_ANGLE *angle, int nDihe, const MDG_DIHE *dihe,
                      int nImpr, const MDG_DIHE *impr, int n14, const int *list14,
                      const double *eps14, const double *sig14);
int mdg_bonded_forces(mdg_ctx *ctx, double *fx, double *fy, double *fz, long long *fix,
                      bool reduce = true);
int mdg_bonded_blocks(const mdg_ctx *ctx);
int mdg_fix_to_f64(mdg_ctx *ctx, long long *fix, double *fx, double *fy, double *fz);
void mdg_bonded_free(mdg_ctx *ctx);
int mdg_upload_exclusions(mdg_ctx *ctx);

// scan / sort utilities (mdg_util.hip)
int mdg_scan_int_to_long(mdg_ctx *ctx, const int *in, long *out, int n);  // out[0]=0, out[n]=total
int mdg_scan_int(mdg_ctx *ctx, const int *in, int *out, int n);           // exclusive, out[n]=total
int mdg_fill_i64(mdg_ctx *ctx, long long *p, long n);
int mdg_fill_f64(mdg_ctx *ctx, double *p, long n, double v);
int mdg_fill_i32(mdg_ctx *ctx, int *p, long n, int v);
int mdg_reduce_partials(mdg_ctx *ctx, const double *part, int nblocks, int nv, double *out);

// list (mdg_list.hip)
int mdg_build_ref_list(mdg_ctx *ctx, bool wrap);
int mdg_build_cluster_list(mdg_ctx *ctx);

// pair (mdg_pair.hip)
int mdg_pair_ref64(mdg_ctx *ctx, int elecType, double *fx, double *fy, double *fz,
                   const int *neighList, const long *off, const int *order, double *eout);
int mdg_excl_corr64(mdg_ctx *ctx, double *fx, double *fy, double *fz, double *eout);
int mdg_pair_cluster(mdg_ctx *ctx);
int mdg_slot_params(mdg_ctx *ctx);
int mdg_prune_list(mdg_ctx *ctx);   // mdg_pair.hip: refresh the pruned list (mixed)

// spme (mdg_spme.hip)
int mdg_spme_setup(mdg_ctx *ctx);
int mdg_spme_free(mdg_ctx *ctx);
int mdg_spme64(mdg_ctx *ctx, double *fx, double *fy, double *fz, double eout[8]);
int mdg_spme_device(mdg_ctx *ctx);   // device-resident: frx.. per atom, partials in epart
int mdg_spme_fork(mdg_ctx *ctx);   // marks the point stream2's work starts from
int mdg_spme_on(mdg_ctx *ctx, int (*fn)(mdg_ctx *));   // runs fn's launches on stream2 from it
// blocks (= partial rows) of the SPME influence kernel over H half-spectrum points
inline int mdg_infl_blocks(long H) { return (int)std::min<long>(nblk(H, 256), 256); }
int mdg_gather_blocks(const mdg_ctx *ctx);   // block partials of the reciprocal forces
int mdg_spme_pot(mdg_ctx *ctx, double *pot);   // FP64 dE_rec/dq per atom
int mdg_adj_pair(mdg_ctx *ctx);      // real-space dE/d(q, eps, sig) into fdx/fdy/fdz slots

// device-resident step pieces (mdg_run.hip)
int mdg_device_energy(mdg_ctx *ctx, bool rebuild, bool fin = true);   // fin: run k_finalize
int mdg_device_step(mdg_ctx *ctx, bool rebuild);
int mdg_reset_step_state(mdg_ctx *ctx);
int mdg_check_flags(mdg_ctx *ctx);
void mdg_invalidate_graphs(mdg_ctx *ctx);

// integrate (mdg_integrate.hip)
int mdg_build_groups(mdg_ctx *ctx, const MDG_CONSTRAINT *cons, int nConst);
int mdg_vv_blocks(const mdg_ctx *ctx);
int mdg_vir_rows(const mdg_ctx *ctx);   // rows of grpVir the record sums
struct FinArgs;
// fin (stage 2 of three-site molecules): the first stage-2 kernel also
// assembles the forces that k_finalize would (mdg_side.h)
// Options of the thermostat and barostat integrators (mdg_ensemble.hip): a
// stage on the generic-group kernels with the ShakeArgs fields of the same
// names; needs sync and a single system.
struct VvOpts {
    int ddStored = 0, noKick = 0, npt = 0, noWrap = 0;
    double dscale = 1., com[3] = {0., 0., 0.};
};
int mdg_vv_stage(mdg_ctx *ctx, int stage, double *fx, double *fy, double *fz, double *kin_out,
                 double *virshake_out, bool sync, const FinArgs *fin = nullptr,
                 const VvOpts *opt = nullptr);
// the device-resident thermostat of one step, before stage 1 (stage 1) or
// after stage 2 (stage 2) (mdg_ensemble.hip)
int mdg_thermo_stage(mdg_ctx *ctx, int stage);
// thermostats and barostats behind the drop-in vv_integrate (mdg_ensemble.hip)
int mdg_vv_ensemble(mdg_ctx *ctx, const MDG_PARAM *param, MDG_ENERGY *ener, MDG_PBC *box,
                    MDG_BATH *bath, int ens, int stage);
// second derivatives for the trajectory adjoint (mdg_pair.hip, mdg_spme.hip)
struct HvpSlots {
    const double *ux, *uy, *uz;      // the direction u in slot order
    double *hx, *hy, *hz;            // H u
    double *mq, *me, *ms;            // d/d(q, eps, sig) (grad E . u)
};
int mdg_hvp_pair(mdg_ctx *ctx, const HvpSlots &H);
struct HvpVec {
    double s[3];                     // the direction of every atom when u is null
};
// the SPME part at the context's positions: hu (3n) += H_rec u and mq (n)
// += d/dq (grad E_rec . u); u (3n, component-major) or, when u is null,
// the same vector s for every atom
int mdg_spme_hvp(mdg_ctx *ctx, const double *u, const double s[3], double *hu, double *mq);
// follow a box the caller changed (an NPT step): derived vectors, link-cell
// counts, net-charge energy and the SPME influence function
int mdg_refresh_box(mdg_ctx *ctx, const MDG_PBC *box);
int mdg_spme_rebox(mdg_ctx *ctx);

// replica batch pieces
int mdg_b_build_list(mdg_ctx *ctx);                  // mdg_list.hip
int mdg_b_spme_setup(mdg_ctx *ctx);                  // mdg_spme.hip
int mdg_b_spme(mdg_ctx *ctx);
void mdg_b_spme_free(mdg_ctx *ctx);
int mdg_b_pair(mdg_ctx *ctx);                        // mdg_pair.hip
