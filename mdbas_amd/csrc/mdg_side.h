// mdg_side.h -- two pieces of per-step work that need little of the GPU
// and run either as kernels of their own or in spare blocks appended to
// the next step's first SHAKE kernel (which keeps one wavefront per SIMD
// busy with latency-bound sweeps, so the extra blocks cost it little and
// the step needs no second stream):
//
//  * the step record: the step's per-block partial sums reduced in a fixed
//    order into one MDG_STEP;
//  * the list prune: per chunk of the cluster-pair list, the entries with a
//    pair within cutOff + skin at the current slot positions.
//
// Both run in blocks of 256 threads in either form, so their results do
// not depend on where they run.
#pragma once
#include "mdg_internal.h"

#define SIDE_BLOCK 256

// Adds rows [0, nrow) of a row-major partial table with NC columns into
// acc[0..NC) (this thread's rows only).  Four rows per thread are loaded
// before they are added, so a thread has several loads in flight (the
// partials were just written, on every XCD).
template <int NC>
__device__ inline void rows_add(const double *__restrict__ p, int nrow, double *acc)
{
    constexpr int U = 4;
    const int step = blockDim.x;
    int r = threadIdx.x;
    for (; r + (U - 1) * step < nrow; r += U * step) {
        double v[U][NC];
#pragma unroll
        for (int u = 0; u < U; u++)
#pragma unroll
            for (int c = 0; c < NC; c++) v[u][c] = p[(long)(r + u * step) * NC + c];
#pragma unroll
        for (int u = 0; u < U; u++)
#pragma unroll
            for (int c = 0; c < NC; c++) acc[c] += v[u][c];
    }
    for (; r < nrow; r += step)
#pragma unroll
        for (int c = 0; c < NC; c++) acc[c] += p[(long)r * NC + c];
}

// ---------------------------------------------------------------- forces

struct FinArgs {
    MdgParams P;
    MdgBox B;
    int n, hasFrozen, nbGather;
    const int *atomSlot;
    const long long *fix, *fiy, *fiz;
    const double *fdx, *fdy, *fdz;
    const float *frx, *fry, *frz;
    const double *frx64, *fry64, *frz64;
    const double *gpart;                 // per-block sums of the reciprocal forces
    const int *exOff, *exLst;            // symmetric exclusion table
    const double *x, *y, *z, *q;
    const int *frozen;
    double *fx, *fy, *fz, *vx, *vy, *vz;
    double *part;                        // per block: eCorr, virCorr
    unsigned long long *mask;            // the SHAKE masks (stage 1, then stage 2)
    long long *bfix;                     // bonded / 1-4 forces, int64 per atom (x | y | z), or null
};

// Mean reciprocal force over the atoms, from the gather's block partials
// (thread-strided with batched loads, then across the block: a fixed
// order), into mean[3]; every thread of the block takes part.
__device__ inline void fin_mean(const FinArgs &A, double *mean, double (*red)[3])
{
    if (A.P.ewald == 0) {   // no reciprocal part (cutoff electrostatics)
        if (threadIdx.x < 3) mean[threadIdx.x] = 0.;
        __syncthreads();
        return;
    }
    double s[3] = {0., 0., 0.};
    rows_add<3>(A.gpart, A.nbGather, s);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (int c = 0; c < 3; c++) {
        const double v = mdg_wave_sum(s[c]);
        if (lane == 0) red[wid][c] = v;
    }
    __syncthreads();
    if (threadIdx.x < 3)
        mean[threadIdx.x] = (((red[0][threadIdx.x] + red[1][threadIdx.x]) + red[2][threadIdx.x]) +
                             red[3][threadIdx.x]) / (double)A.n;
    __syncthreads();
}

// Atom i's force: F = slot pair force + (reciprocal force - mean
// reciprocal force) + excluded-pair corrections.  The mean is over the
// reciprocal part only (spme.c:648-670 runs on arrays energy() has just
// zeroed).  The correction of each excluded pair (energy.c:567-621) is
// evaluated from both of its atoms with the reference's orientation
// (delta = x_hi - x_lo), so every atom sums its own terms without atomics;
// energy and virial (acc) count each pair once.  Frozen atoms get zero
// force and velocity (utils.c:131-153); returns whether atom i is frozen.
template <bool MIXED>
__device__ inline bool fin_atom(const FinArgs &A, int i, const double *mean, double *acc,
                                double &gx, double &gy, double &gz)
{
    const int s = A.atomSlot[i];
    const bool ew = A.P.ewald != 0;   // else cutoff electrostatics: pair forces only
    if (MIXED) {
        gx = (double)A.fix[s] * (1.0 / MDG_FSCALE);
        gy = (double)A.fiy[s] * (1.0 / MDG_FSCALE);
        gz = (double)A.fiz[s] * (1.0 / MDG_FSCALE);
        if (ew) {
            gx += (double)A.frx[i] - mean[0];
            gy += (double)A.fry[i] - mean[1];
            gz += (double)A.frz[i] - mean[2];
        }
    } else {
        gx = ew ? (A.frx64[i] - mean[0]) + A.fdx[s] : A.fdx[s];
        gy = ew ? (A.fry64[i] - mean[1]) + A.fdy[s] : A.fdy[s];
        gz = ew ? (A.frz64[i] - mean[2]) + A.fdz[s] : A.fdz[s];
    }
    for (int k = A.exOff[i]; ew && k < A.exOff[i + 1]; k++) {
        const int j = A.exLst[k];
        const int lo = min(i, j), hi = max(i, j);
        double dx = A.x[hi] - A.x[lo], dy = A.y[hi] - A.y[lo], dz = A.z[hi] - A.z[lo];
        const double r2 = mdg_dist(A.B, dx, dy, dz);
        const double r = sqrt(r2), rt = 1. / r;
        double dC;
        const double e = ewald_corr64(A.P.alpha, A.P.chargeConst * A.q[lo] * A.q[hi], r, rt, dC);
        const double sgn = (i == lo) ? 1.0 : -1.0;
        gx += sgn * (dC * dx * rt);
        gy += sgn * (dC * dy * rt);
        gz += sgn * (dC * dz * rt);
        if (i == lo) {
            acc[0] += e;
            acc[1] += dC * r;
        }
    }
    if (A.bfix) {
        // bonded and 1-4 forces (mdg_bonded.hip), cleared for the next step
        gx += (double)A.bfix[i] * (1.0 / MDG_FSCALE);
        gy += (double)A.bfix[A.n + i] * (1.0 / MDG_FSCALE);
        gz += (double)A.bfix[2 * (long)A.n + i] * (1.0 / MDG_FSCALE);
        A.bfix[i] = 0;
        A.bfix[A.n + i] = 0;
        A.bfix[2 * (long)A.n + i] = 0;
    }
    const bool fr = A.hasFrozen && A.frozen[i];
    if (fr) {
        gx = gy = gz = 0.;
        A.vx[i] = A.vy[i] = A.vz[i] = 0.;
    }
    A.fx[i] = gx;
    A.fy[i] = gy;
    A.fz[i] = gz;
    return fr;
}

// ---------------------------------------------------------------- record

struct RecArgs {
    double fact, eSelf, eNet;
    int recCap;
    // per-block partial counts of this step's producers
    int nbPair, nbExcl, nbInfl, nbKin, nGrpVir, nbBond;
    const double *pPair, *pExcl, *pInfl, *pKin, *grpVir, *pBond;
    const double *bath;           // device thermostat: conint, consv, chiT at [4..6] (or null)
};

// One MDG_STEP record by one block of SIDE_BLOCK threads: reduces the
// step's partials (pair, correction, influence function, kinetic energy,
// SHAKE virial) in a fixed order and assembles the energies
// (spme.c:674-697 for the reciprocal energy and virial, energy.c:623-627
// and 193-196 for the sums).
#define REC_NV 27   // 16 nonbonded / kinetic / SHAKE sums + 11 bonded and 1-4 sums
__device__ inline void record_block(const RecArgs &R, int *ctrl, MDG_STEP *rec, int rebuilt)
{
    __shared__ double red[SIDE_BLOCK / 64][REC_NV];
    __shared__ double tot[REC_NV];
    double v[REC_NV];
#pragma unroll
    for (int k = 0; k < REC_NV; k++) v[k] = 0.;
    rows_add<5>(R.pPair, R.nbPair, v);
    rows_add<2>(R.pExcl, R.nbExcl, v + 5);
    rows_add<7>(R.pInfl, R.nbInfl, v + 7);
    for (int r = threadIdx.x; r < R.nbKin; r += blockDim.x) v[14] += R.pKin[(long)r * 7];
    rows_add<1>(R.grpVir, R.nGrpVir, v + 15);
    if (R.nbBond) rows_add<11>(R.pBond, R.nbBond, v + 16);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < REC_NV; k++) {
        const double s = mdg_wave_sum(v[k]);
        if (lane == 0) red[wid][k] = s;
    }
    __syncthreads();
    // thread k adds value k over the wavefronts (in wavefront order)
    if (threadIdx.x < REC_NV) {
        double s = 0.;
        for (int w = 0; w < SIDE_BLOCK / 64; w++) s += red[w][threadIdx.x];
        tot[threadIdx.x] = s;
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    for (int k = 0; k < REC_NV; k++) v[k] = tot[k];
    const int idx = ctrl[CTRL_STEP];
    const double er = v[7];
    const double st0 = R.fact * (v[8] + er) + R.eNet;
    const double st3 = R.fact * (v[11] + er) + R.eNet;
    const double st5 = R.fact * (v[13] + er) + R.eNet;
    const double eRec = R.fact * er + R.eSelf + R.eNet;
    const double virRec = -(st0 + st3 + st5);
    MDG_STEP s;
    s.step = idx;
    s.listRebuilt = rebuilt;
    s.eDir = v[0];
    s.eRec = eRec;
    s.eCorr = v[5];
    // energy.c:623-627 plus the 1-4 pairs (energy.c:431-435 / 726-730)
    s.elec = (s.eDir + eRec + s.eCorr) + v[23];
    s.vdw = v[1] + v[24];
    s.ebond = v[16];
    s.eub = v[17];
    s.eang = v[18];
    s.edihe = v[19];
    s.eimpr = v[20];
    // energy.c:193-194
    s.pot = s.elec + s.vdw + s.ebond + s.eang + s.eub + s.edihe + s.eimpr;
    s.kin = 0.5 * v[14];
    s.tot = s.pot + s.kin;
    s.virelec = (v[2] + virRec + v[6]) + v[25];
    s.virvdw = v[3] + v[26];
    s.virshake = v[15];
    s.virbond = v[21];
    s.virub = v[22];
    s.conint = R.bath ? R.bath[4] : 0.;
    s.consv = R.bath ? R.bath[5] : 0.;
    s.chiT = R.bath ? R.bath[6] : 0.;
    s.nInCut = v[4];
    if (idx < R.recCap) rec[idx] = s;
    ctrl[CTRL_STEP] = idx + 1;
}

// Batch: one MDG_STEP per replica, by one block each (replica r): the
// replica's chunk, atom and group sums in a fixed order, assembled as
// record_block does.
struct BRecArgs {
    double chargeConst;
    int R, recCap;
    const MdgBox *boxes;
    const double *eSelf, *eNet;
    const int *aoff, *goff, *cluStart, *chOff;
    int maxChunks;
    const double *chE, *atomE, *grpKin, *grpVir, *repInfl;
    int *stepCtr;
    MDG_STEP *rec;                       // [step][replica]
};

__device__ inline void b_record_block(const BRecArgs &R, int r, int rebuilt)
{
    __shared__ double red[SIDE_BLOCK / 64][10];
    double v[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    const int c0 = min(R.chOff[R.cluStart[r]], R.maxChunks);
    const int c1 = min(R.chOff[R.cluStart[r + 1]], R.maxChunks);
    for (int c = c0 + threadIdx.x; c < c1; c += blockDim.x)
        for (int k = 0; k < 5; k++) v[k] += R.chE[(long)c * 5 + k];
    for (int i = R.aoff[r] + threadIdx.x; i < R.aoff[r + 1]; i += blockDim.x) {
        v[5] += R.atomE[2 * (long)i];
        v[6] += R.atomE[2 * (long)i + 1];
    }
    for (int g = R.goff[r] + threadIdx.x; g < R.goff[r + 1]; g += blockDim.x) {
        v[7] += R.grpKin[g];
        v[8] += R.grpVir[g];
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (int k = 0; k < 9; k++) {
        const double s = mdg_wave_sum(v[k]);
        if (lane == 0) red[wid][k] = s;
    }

    __syncthreads();
    if (threadIdx.x != 0) return;
    for (int k = 0; k < 9; k++) v[k] = red[0][k] + red[1][k] + red[2][k] + red[3][k];
    const double *inf = R.repInfl + 7 * (long)r;
    const double fact = (6.283185307179586 / R.boxes[r].vol) * R.chargeConst;
    const double er = inf[0];
    const double st0 = fact * (inf[1] + er) + R.eNet[r];
    const double st3 = fact * (inf[4] + er) + R.eNet[r];
    const double st5 = fact * (inf[6] + er) + R.eNet[r];
    const double eRec = fact * er + R.eSelf[r] + R.eNet[r];
    const int idx = R.stepCtr[r];
    MDG_STEP s;
    s.step = idx;
    s.listRebuilt = rebuilt;
    s.eDir = v[0];
    s.eRec = eRec;
    s.eCorr = v[5];
    s.elec = s.eDir + eRec + s.eCorr;
    s.vdw = v[1];
    s.pot = s.elec + s.vdw;
    s.ebond = s.eub = s.eang = s.edihe = s.eimpr = 0.;
    s.virbond = s.virub = 0.;
    s.conint = s.consv = s.chiT = 0.;
    s.kin = 0.5 * v[7];
    s.tot = s.pot + s.kin;
    s.virelec = v[2] + (-(st0 + st3 + st5)) + v[6];
    s.virvdw = v[3];
    s.virshake = v[8];
    s.nInCut = v[4];
    if (idx < R.recCap) R.rec[(long)idx * R.R + r] = s;
    R.stepCtr[r] = idx + 1;
}

// ---------------------------------------------------------------- prune

struct PruneArgs {
    int nclu, maxChunks, cap;
    const int *chOff, *cpOff, *cpJ;
    const unsigned long long *cpMask;
    const int2 *chunk;               // chunk -> (I, first entry)
    const double *sx, *sy, *sz;      // slot positions
    double L0, L1, L2, iL0, iL1, iL2;
    const MdgBox *boxes;             // batch: box of each replica
    const int *crep;                 // batch: replica of each cluster
    int *pJ;
    unsigned int *pMask32;           // the 64-bit masks as two 32-bit halves
    int *pCnt;
    const float4 *cpD;               // entry origin offsets, copied with the
    float4 *pD;                      // kept entries (relative coordinates)
    double rp2, rp;                  // (cutOff + skin)^2, cutOff + skin
    int *ctrl;
};

__device__ inline double readlane_d(double v, int l)
{
    return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), l),
                            __builtin_amdgcn_readlane(__double2loint(v), l));
}

// The pruned list from the current slot positions: per chunk of the full
// list, the entries with a pair within rp, in order, at the chunk's own
// offsets.  Wavefront w0 of nw takes chunks w0, w0 + nw, ...
//
// One wavefront per chunk (at most MDG_CHUNK = 32 entries): lane t + 32 h
// takes entry t and the j atoms 4h .. 4h+3 of its J cluster and tests them
// against the 8 i atoms, which makes the 32-bit half h of the entry's pair
// mask (bit 8 j + i).  Distances go through the first i atom c: with the
// i offsets u_i = image(x_i - x_c) at most e per axis, every pair within
// rp has image(x_j - x_c) - u_i as its minimum-image vector whenever
// rp + e < L / 2 (the vector is the unique representative below L / 2).
// A chunk where that fails keeps its entries whole.  The kept entries are
// packed with a wave prefix count.  The first wavefront marks the pruned
// list current (CTRL_PRUNE).
template <bool BATCH>
__device__ inline void prune_waves(const PruneArgs &A, int w0, int nw)
{
    static_assert(MDG_CHUNK <= 32, "one lane pair per entry");
    const int lane = threadIdx.x & 63;
    const int t = lane & 31, h = lane >> 5, ia = lane & 7;
    const int nch = min(mdg_uload(A.chOff, A.nclu), A.maxChunks);
    if (w0 == 0 && lane == 0) A.ctrl[CTRL_PRUNE] = 1;
    for (int W = w0; W < nch; W += nw) {
        const long long ch = mdg_uload((const long long *)A.chunk, W);   // int2 (I, e0)
        const int I = (int)(ch & 0xffffffffll), e0 = (int)(ch >> 32);
        const int e1 = min(min(e0 + MDG_CHUNK, mdg_uload(A.cpOff, I + 1)), A.cap);
        const int n = max(e1 - e0, 0);
        double L0 = A.L0, L1 = A.L1, L2 = A.L2, iL0 = A.iL0, iL1 = A.iL1, iL2 = A.iL2;
        if (BATCH && n > 0) {
            const MdgBox &Bx = A.boxes[mdg_uload(A.crep, I)];
            L0 = Bx.a1;
            L1 = Bx.type == 1 ? Bx.a1 : Bx.b2;
            L2 = Bx.type == 1 ? Bx.a1 : Bx.c3;
            iL0 = 1.0 / L0;
            iL1 = 1.0 / L1;
            iL2 = 1.0 / L2;
        }
        const bool live = t < n;
        const int J = live ? A.cpJ[e0 + t] : 0;
        const unsigned long long m = live ? A.cpMask[e0 + t] : 0ull;
        double xj[4], yj[4], zj[4];
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int sj = J * MDG_CL + 4 * h + q;
            xj[q] = A.sx[sj];
            yj[q] = A.sy[sj];
            zj[q] = A.sz[sj];
        }
        // i offsets: lane ia holds atom ia's (padding slots, far points
        // never in a mask, are parked far away and left out of e)
        const double xc = mdg_uload(A.sx, (long)I * MDG_CL), yc = mdg_uload(A.sy, (long)I * MDG_CL),
                     zc = mdg_uload(A.sz, (long)I * MDG_CL);
        const long si = (long)I * MDG_CL + ia;
        const double xi = A.sx[si];
        double ux = xi - xc, uy = A.sy[si] - yc, uz = A.sz[si] - zc;
        ux = fma(-L0, rint(ux * iL0), ux);
        uy = fma(-L1, rint(uy * iL1), uy);
        uz = fma(-L2, rint(uz * iL2), uz);
        const bool pad = !(xi < 1e6);
        double ex = pad ? 0. : fabs(ux), ey = pad ? 0. : fabs(uy), ez = pad ? 0. : fabs(uz);
        for (int o = 1; o < 8; o <<= 1) {
            ex = fmax(ex, __shfl_xor(ex, o, 64));
            ey = fmax(ey, __shfl_xor(ey, o, 64));
            ez = fmax(ez, __shfl_xor(ez, o, 64));
        }
        if (pad) ux = uy = uz = 1e7;
        const bool ok = __builtin_amdgcn_readfirstlane(
                            (int)(A.rp + ex < 0.5 * L0 - 1e-6 && A.rp + ey < 0.5 * L1 - 1e-6 &&
                                  A.rp + ez < 0.5 * L2 - 1e-6)) != 0;
        unsigned int bits = 0xffffffffu;
        if (ok) {
            bits = 0;
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const double dx = xj[q] - xc, dy = yj[q] - yc, dz = zj[q] - zc;
                xj[q] = fma(-L0, rint(dx * iL0), dx);
                yj[q] = fma(-L1, rint(dy * iL1), dy);
                zj[q] = fma(-L2, rint(dz * iL2), dz);
            }
#pragma unroll
            for (int i = 0; i < MDG_CL; i++) {
                const double vx = readlane_d(ux, i), vy = readlane_d(uy, i), vz = readlane_d(uz, i);
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    const double dx = xj[q] - vx, dy = yj[q] - vy, dz = zj[q] - vz;
                    const double r2 = dx * dx + dy * dy + dz * dz;
                    bits |= (r2 <= A.rp2 ? 1u : 0u) << (8 * q + i);
                }
            }
        }
        const unsigned int keep = bits & (unsigned int)(m >> (32 * h));
        const unsigned long long nz = __builtin_amdgcn_ballot_w64(keep != 0);
        const unsigned int kept = (unsigned int)nz | (unsigned int)(nz >> 32);   // per entry t
        const int pos = __builtin_popcount(kept & ((1u << t) - 1u));
        if (live && ((kept >> t) & 1u)) {
            if (h == 0) A.pJ[e0 + pos] = J;
            if (h == 0 && A.pD) A.pD[e0 + pos] = A.cpD[e0 + t];
            A.pMask32[2 * (long)(e0 + pos) + h] = keep;
        }
        if (lane == 0) A.pCnt[W] = __builtin_popcount(kept);
    }
}

// ---------------------------------------------------------------- spare blocks

// Work for the blocks a kernel launches beyond its own nbMain: the nbRec
// blocks after them write the record (one block; a batch one per
// replica), the nbPrune blocks after those prune the list.  Blocks have
// SIDE_BLOCK threads.
struct SideWork {
    int nbMain, nbRec, nbPrune, rebuilt;
    RecArgs R;
    int *ctrl;
    MDG_STEP *out;
    BRecArgs BR;
    PruneArgs Pr;
};

template <bool BATCH>
__device__ inline void side_work(const SideWork &S)
{
    int b = blockIdx.x - S.nbMain;
    if (b < S.nbRec) {
        if (BATCH)
            b_record_block(S.BR, b, S.rebuilt);
        else
            record_block(S.R, S.ctrl, S.out, S.rebuilt);
        return;
    }
    b -= S.nbRec;
    if (b < S.nbPrune) prune_waves<BATCH>(S.Pr, b * (SIDE_BLOCK / 64) + (threadIdx.x >> 6),
                                          S.nbPrune * (SIDE_BLOCK / 64));
}

// Host side (mdg_run.hip, mdg_pair.hip).
RecArgs mdg_rec_args(mdg_ctx *ctx, int nbKin, int nGrpVir, int nbExcl);
BRecArgs mdg_b_rec_args(mdg_ctx *ctx);
FinArgs mdg_fin_args(mdg_ctx *ctx);
// Device steps of one box of three-site molecules compute the forces in
// the first stage-2 SHAKE kernel (no k_finalize): its blocks then hold the
// correction partials.
bool mdg_fin_fused(const mdg_ctx *ctx);
int mdg_excl_rows(const mdg_ctx *ctx);
PruneArgs mdg_prune_args(mdg_ctx *ctx);
bool mdg_prune_on(const mdg_ctx *ctx);       // mixed, pruning on, list built
int mdg_prune_blocks(const mdg_ctx *ctx);    // blocks of a prune kernel of its own
